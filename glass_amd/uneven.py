"""Uneven batch: many independent lognormal maps of different (nside, lmax),
with their shear, submitted to the GPU as ONE graph launch.

Maps are grouped by geometry.  Each group is a batched engine
(``LognormalShells`` with ncorr = 0, so every map has its own spectrum and
its own normals, keyed by (seed, group, map)); the groups' kernel sequences
(iternorm -> alm -> MFMA Legendre -> ring FFT + lognormal -> analysis ->
spin-2 shear) are captured once into a HIP graph, one concurrent branch per
geometry, so a replay submits the whole uneven batch with one host call and
the small geometries fill the device while the large ones run.  The Legendre
tile configuration of every geometry is chosen by the autotuner before the
capture (``glass_amd.autotune``).

The synthetic workload: 128 maps, nside uniform over {256, 512, 1024, 2048},
lmax = 2 nside - 1, C_l = 1e-4 (l+1)^t exp(-(l/1500)^2) with a tilt t uniform
in [-1.5, -1.0] per map, lognormal shift 1.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib
from .batch import LognormalShells


@dataclasses.dataclass
class MapSpec:
    index: int
    nside: int
    lmax: int
    tilt: float


def uneven_specs(n: int = 128, seed: int = 3, nsides=(256, 512, 1024, 2048),
                 tilt=(-1.5, -1.0)) -> list[MapSpec]:
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(nsides), size=n)
    t = rng.uniform(tilt[0], tilt[1], size=n)
    return [MapSpec(i, int(ns[i]), 2 * int(ns[i]) - 1, float(t[i])) for i in range(n)]


def map_spectrum(lmax: int, tilt: float, scale: float = 1500.0, amp: float = 1e-4):
    ell = np.arange(lmax + 1, dtype=np.float64)
    return amp * (ell + 1.0) ** tilt * np.exp(-((ell / scale) ** 2))


def independent_gls(spectra):
    """GLASS-order spectra of uncorrelated fields: autos, empty crosses."""
    gls = []
    for i, c in enumerate(spectra):
        gls.append(np.asarray(c))
        gls.extend(np.zeros(0) for _ in range(i))
    return gls


class UnevenBatch:
    def __init__(self, specs, *, shift: float = 1.0, shear: bool = True, seed: int = 3,
                 tune: bool = True):
        T = _lib.torch()
        _lib.require_gpu()
        self.specs = list(specs)
        groups: dict = {}
        for sp in self.specs:
            groups.setdefault((sp.nside, sp.lmax), []).append(sp)
        # largest geometry first: its branch is captured (and starts) first
        self.keys = sorted(groups, key=lambda k: -k[0])
        self.groups = {k: groups[k] for k in self.keys}
        self.engines = {}
        self.tuning = {}
        for gi, key in enumerate(self.keys):
            nside, lmax = key
            sps = self.groups[key]
            gls = independent_gls([map_spectrum(lmax, sp.tilt) for sp in sps])
            self.engines[key] = LognormalShells(gls, nside, ncorr=0, shift=shift,
                                                seed=seed * 1000 + gi, lmax=lmax, shear=shear)
        if tune:
            from . import autotune

            for key, eng in self.engines.items():
                self.tuning[key] = autotune.tune_legendre(eng)
        self.streams = {key: T.cuda.Stream() for key in self.keys}
        self.graph = None
        self.nmaps = len(self.specs)

    def run_eager(self):
        for key in self.keys:
            self.engines[key].run()

    def _branches(self):
        T = _lib.torch()
        main = T.cuda.current_stream()
        for key in self.keys:
            s = self.streams[key]
            s.wait_stream(main)
            with T.cuda.stream(s):
                self.engines[key].run()
        for key in self.keys:
            main.wait_stream(self.streams[key])

    def capture(self):
        """Record the whole batch into one HIP graph (all buffers are already
        allocated; the library's entry points neither allocate nor set memory
        with memset, so they are capturable)."""
        T = _lib.torch()
        # one eager pass on the branch streams first (lazy tables, allocator state)
        self._branches()
        T.cuda.synchronize()
        g = T.cuda.CUDAGraph()
        with T.cuda.graph(g):
            self._branches()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def check(self):
        for eng in self.engines.values():
            eng.check()

    def maps_of(self, index: int):
        """The lognormal map (and shear, if computed) of spec ``index``."""
        sp = self.specs[index]
        key = (sp.nside, sp.lmax)
        j = [s.index for s in self.groups[key]].index(index)
        eng = self.engines[key]
        return eng.maps[j], (eng.gamma[j] if eng.gamma is not None else None)
