"""Uneven batch under one HIP graph, and the Legendre autotuner.

Every graph replay (with eager work and synchronisations between replays) must
reproduce the eager run exactly (same kernels, same inputs); each map is
within 1e-5 of the FP64 path (max |error| / map RMS).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from glass_amd import autotune  # noqa: E402
from glass_amd.batch import LognormalShells, shell_spectra  # noqa: E402
from glass_amd.uneven import UnevenBatch, independent_gls, map_spectrum, uneven_specs  # noqa: E402


@pytest.fixture(autouse=True)
def _tune_cache(tmp_path, monkeypatch):
    # tuning choices live in each engine's plan; only the cache file is shared
    monkeypatch.setenv("GLASS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))


def rel(a, b):
    a, b = a.double(), b.double()
    return float(torch.max(torch.abs(a - b)) / torch.sqrt(torch.mean(b * b)))


def test_uneven_specs_workload():
    sp = uneven_specs()
    assert len(sp) == 128
    ns = np.array([s.nside for s in sp])
    assert set(ns) <= {256, 512, 1024, 2048} and len(set(ns)) == 4
    assert all(s.lmax == 2 * s.nside - 1 for s in sp)
    t = np.array([s.tilt for s in sp])
    assert t.min() >= -1.5 and t.max() <= -1.0
    assert uneven_specs() == sp


@pytest.mark.parametrize("grouped", [True, False])
def test_graph_replays_match_eager_and_fp64(grouped):
    """Graph replays (grouped: one contraction launch for all geometries)
    equal the eager run (one contraction launch per geometry) bit for bit."""
    specs = uneven_specs(n=12, seed=3, nsides=(16, 32, 64))
    ub = UnevenBatch(specs, shear=True, tune=True, grouped=grouped)
    ub.run_eager()
    torch.cuda.synchronize()
    ub.check()
    ref = {s.index: tuple(x.clone() for x in ub.maps_of(s.index)) for s in specs}
    ub.capture()
    for rep in range(3):
        # eager work between replays overwrites the buffers
        for eng in ub.engines.values():
            eng.maps.zero_()
            eng.gamma.zero_()
        torch.cuda.synchronize()
        ub.replay()
        torch.cuda.synchronize()
        for s in specs:
            m, g = ub.maps_of(s.index)
            assert torch.equal(m, ref[s.index][0]), (rep, s.index)
            assert torch.equal(g, ref[s.index][1]), (rep, s.index)
    # accuracy of a few maps against the FP64 path of their engine
    for key, eng in ub.engines.items():
        for j in (0, eng.S - 1):
            assert rel(eng.maps[j], eng.fp64_reference_map(j)) < 1e-5, (key, j)
    # the per-map spectra: engine j of a group uses map_spectrum(lmax, tilt)
    key = ub.keys[0]
    sps = ub.groups[key]
    gls = independent_gls([map_spectrum(key[1], s.tilt) for s in sps])
    assert len(gls) == len(sps) * (len(sps) + 1) // 2


def test_autotuner_picks_and_caches_a_candidate():
    S, nside, lmax = 8, 32, 63
    eng = LognormalShells(shell_spectra(S, lmax), nside, ncorr=3, seed=2, lmax=lmax)
    cfg = autotune.tune_legendre(eng, reps=2)
    assert (cfg["path"], cfg["nt"]) in autotune.CANDIDATES
    assert set(cfg["all_ms"]) == {f"{p}/{n}" for p, n in autotune.CANDIDATES}
    assert autotune.current(eng.plan)["path"] == cfg["path"]
    again = autotune.tune_legendre(eng)
    assert (again["path"], again["nt"]) == (cfg["path"], cfg["nt"])
    # every candidate gives the same maps to FP32 accuracy
    outs = []
    for p, n in autotune.CANDIDATES:
        autotune.apply(eng.plan, {"path": p, "nt": n})
        eng.run()
        torch.cuda.synchronize()
        outs.append(eng.maps.clone())
    for o in outs[1:]:
        for j in range(S):
            assert rel(o[j], outs[0][j]) < 1e-5


def test_c4_geometries_graph_eager_and_fp64():
    """The uneven batch at its real geometries (two maps at each nside in
    {256, 512, 1024, 2048}, lmax = 2 nside - 1, lognormal + shear): graph
    replays equal the eager run bit for bit, and every map is within 1e-5 of
    the FP64 path, every shear component within 1e-5 of the FP64 shear of the
    same map (max |error| / RMS, printed per geometry)."""
    from glass_amd import healpix as hp
    from glass_amd.uneven import MapSpec

    tilts = np.linspace(-1.5, -1.0, 8)
    specs = [MapSpec(i, ns, 2 * ns - 1, float(tilts[i]))
             for i, ns in enumerate((256, 256, 512, 512, 1024, 1024, 2048, 2048))]
    ub = UnevenBatch(specs, shear=True, tune=True)
    ub.run_eager()
    torch.cuda.synchronize()
    ub.check()
    ref = {s.index: tuple(x.clone() for x in ub.maps_of(s.index)) for s in specs}
    ub.capture()
    for rep in range(2):
        for eng in ub.engines.values():
            eng.maps.zero_()
            eng.gamma.zero_()
        torch.cuda.synchronize()
        ub.replay()
        torch.cuda.synchronize()
        for s in specs:
            m, g = ub.maps_of(s.index)
            assert torch.equal(m, ref[s.index][0]) and torch.equal(g, ref[s.index][1]), (rep, s)
    for (nside, lmax), eng in ub.engines.items():
        ell = np.arange(lmax + 1, dtype=np.float64)
        fl = np.zeros(lmax + 1)
        fl[2:] = -np.sqrt((ell[2:] - 1) * (ell[2:] + 2) / (ell[2:] * (ell[2:] + 1)))
        for j in range(eng.S):
            e_map = rel(eng.maps[j], eng.fp64_reference_map(j))
            alm = hp.map2alm(eng.maps[j].double(), lmax=lmax, pol=False, precision="fp64",
                             use_pixel_weights=True)
            alm = hp.almxfl(alm, fl)
            q, u = hp.alm2map_spin([alm, torch.zeros_like(alm)], nside, 2, lmax)
            e_q, e_u = rel(eng.gamma[j, 0], q), rel(eng.gamma[j, 1], u)
            print(f"uneven batch nside={nside} map {j}: {e_map:.2e} gamma {e_q:.2e} {e_u:.2e}", flush=True)
            assert e_map < 1e-5 and e_q < 1e-5 and e_u < 1e-5, (nside, j, e_map, e_q, e_u)
