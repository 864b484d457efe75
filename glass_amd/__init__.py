"""glass_amd: MI355X-native field-generation hot path of GLASS.

Drop-in counterparts of the reference's generation path (same names, argument
meaning and error messages):

    glass_amd.generate / generate_lognormal / generate_gaussian / iternorm /
    cls2cov / getcl / nfields_from_nspectra          (glass/fields.py)
    glass_amd.harmonics.multalm                      (glass/harmonics.py)
    glass_amd.healpix.alm2map / map2alm / alm2map_spin / almxfl / ...
                                                     (glass/healpix.py)
    glass_amd.grf.Normal / Lognormal / SquaredNormal (glass/grf)
    glass_amd.positions_from_delta / linear_bias / loglinear_bias (glass/points.py)
    glass_amd.from_convergence / shear_from_convergence / MultiPlaneConvergence /
    multi_plane_matrix / multi_plane_weights         (glass/lensing.py)

All numerics run in hand-written HIP kernels for gfx950 behind the C ABI in
include/glass_amd.h; there is no CPU fallback.
"""

from . import grf, harmonics, healpix, points, rng
from .fields import (
    cls2cov,
    generate,
    generate_gaussian,
    generate_lognormal,
    getcl,
    iternorm,
    nfields_from_nspectra,
)
from .lensing import (
    MultiPlaneConvergence,
    RadialWindow,
    from_convergence,
    multi_plane_matrix,
    multi_plane_weights,
    shear_from_convergence,
)
from .points import linear_bias, loglinear_bias, positions_from_delta
from .rng import DeviceRNG

__all__ = [
    "DeviceRNG", "MultiPlaneConvergence", "RadialWindow", "cls2cov", "from_convergence", "generate", "generate_gaussian",
    "generate_lognormal", "getcl", "grf", "harmonics", "healpix", "iternorm",
    "linear_bias", "loglinear_bias", "multi_plane_matrix", "multi_plane_weights", "points",
    "positions_from_delta",
    "nfields_from_nspectra", "rng", "shear_from_convergence",
]
