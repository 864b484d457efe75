"""glass_amd: MI355X-native field-generation hot path of GLASS.

Drop-in counterparts of the reference's generation path (same names, argument
meaning and error messages):

    glass_amd.generate / generate_lognormal / generate_gaussian / iternorm /
    cls2cov / getcl / nfields_from_nspectra / discretized_cls /
    compute_gaussian_spectra / solve_gaussian_spectra (glass/fields.py)
    glass_amd.grf.compute / solve / corr / icorr / dcorr (glass/grf)
    glass_amd.save_cls / load_cls                    (glass/user.py)
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

from . import grf, harmonics, healpix, points, rng, user
from .fields import (
    cls2cov,
    compute_gaussian_spectra,
    discretized_cls,
    enumerate_spectra,
    solve_gaussian_spectra,
    spectra_indices,
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
from .user import load_cls, save_cls

__all__ = [
    "DeviceRNG", "MultiPlaneConvergence", "RadialWindow", "cls2cov", "compute_gaussian_spectra",
    "discretized_cls", "enumerate_spectra", "load_cls", "save_cls", "solve_gaussian_spectra",
    "spectra_indices", "user", "from_convergence", "generate", "generate_gaussian",
    "generate_lognormal", "getcl", "grf", "harmonics", "healpix", "iternorm",
    "linear_bias", "loglinear_bias", "multi_plane_matrix", "multi_plane_weights", "points",
    "positions_from_delta",
    "nfields_from_nspectra", "rng", "shear_from_convergence",
]
