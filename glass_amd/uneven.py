"""Uneven batch: many independent lognormal maps of different (nside, lmax),
with their shear, submitted to the GPU as ONE graph launch.

Maps are grouped by geometry.  Each group is a batched engine
(``LognormalShells`` with ncorr = 0, so every map has its own spectrum and
its own normals, keyed by (seed, group, map)); the groups' kernel sequences
(iternorm -> alm -> MFMA Legendre -> ring FFT + lognormal -> analysis ->
spin-2 shear) are captured once into a HIP graph through the library's graph
entry points (``ga_graph_begin`` / ``ga_graph_end``), so a replay submits the
whole uneven batch with one host call.  The dominant stage, the split-FP16
Legendre contraction, runs for all geometries in ONE grouped launch
(``ga_legendre_contract_grouped_f32``: the largest geometry's blocks first,
the small geometries' blocks filling the device as it drains); the stages
before and after it run as one concurrent graph branch per geometry.  The
tile configuration of every geometry (one or two 32-column tiles per wave)
is chosen by the autotuner before the capture (``glass_amd.autotune``).
``grouped=False`` keeps one contraction launch per geometry inside its
branch.

The synthetic workload: 128 maps, nside uniform over {256, 512, 1024, 2048},
lmax = 2 nside - 1, C_l = 1e-4 (l+1)^t exp(-(l/1500)^2) with a tilt t uniform
in [-1.5, -1.0] per map, lognormal shift 1.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib
from .batch import LognormalShells


GROUP_MAX = 8  # geometries per grouped contraction launch


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
                 tune: bool = True, grouped: bool = True):
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
        self.grouped = bool(grouped)
        if tune:
            from . import autotune

            cands = autotune.GROUPED_CANDIDATES if self.grouped else autotune.CANDIDATES
            for key, eng in self.engines.items():
                self.tuning[key] = autotune.tune_legendre(eng, candidates=cands)
        self.streams = {key: T.cuda.Stream() for key in self.keys}
        self.capture_stream = T.cuda.Stream()
        self.graph = None
        self.nmaps = len(self.specs)

    def run_eager(self):
        for key in self.keys:
            self.engines[key].run()

    def _fork(self, work):
        """work(engine) on every geometry's branch stream, joined back."""
        T = _lib.torch()
        main = T.cuda.current_stream()
        for key in self.keys:
            s = self.streams[key]
            s.wait_stream(main)
            with T.cuda.stream(s):
                work(self.engines[key])
        for key in self.keys:
            main.wait_stream(self.streams[key])

    def _branches(self):
        if not self.grouped:
            self._fork(lambda eng: eng.run())
            return

        def before(eng):
            eng.stage_fields()
            eng.stage_operands()

        def after(eng):
            eng.stage_rings()
            eng.stage_shear()

        self._fork(before)
        for i in range(0, len(self.keys), GROUP_MAX):  # the library takes 8 per launch
            engs = [self.engines[k] for k in self.keys[i:i + GROUP_MAX]]
            _lib.contract_grouped([e.plan for e in engs], [e.S for e in engs],
                                  [e.F for e in engs])
        self._fork(after)

    def capture(self):
        """Record the whole batch into one HIP graph through the library's
        graph entry points (all buffers are already allocated; the library's
        entry points neither allocate nor set memory with memset)."""
        T = _lib.torch()
        cs = self.capture_stream
        cs.wait_stream(T.cuda.current_stream())
        with T.cuda.stream(cs):
            # one eager pass first (code objects, lazily built tables)
            self._branches()
            cs.synchronize()
            self.graph = _lib.Graph().capture(self._branches)
        T.cuda.current_stream().wait_stream(cs)
        return self.graph

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.launch()

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
