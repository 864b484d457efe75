"""The grouped contraction (several geometries in one launch) and the graph
entry points of the C ABI.

* ga_legendre_contract_grouped_f32 writes, for every geometry, exactly the
  ring buffer ga_legendre_contract_batch_f32 writes for it alone (same
  kernel body, same operands): bit-identical, with one and two column tiles
  per wave and with one accumulation chain or l-chunks, in any mix.
* ga_graph_begin / ga_graph_end / ga_graph_launch: a captured sequence of
  entry points replays to the eager result, several times, with eager work
  and synchronisations between replays.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from glass_amd import _lib  # noqa: E402
from glass_amd.batch import LognormalShells, shell_spectra  # noqa: E402


def _engine(nside, lmax, S, seed, nt, flush):
    print(f"grouped geometry nside={nside} lmax={lmax} S={S} nt={nt} flush={flush}", flush=True)
    eng = LognormalShells(shell_spectra(S, lmax), nside, ncorr=2, seed=seed, lmax=lmax)
    eng.plan.set_tuning(_lib.TUNE_MFMA_PATH, 2)
    eng.plan.set_tuning(_lib.TUNE_MFMA_TILES, nt)
    eng.plan.set_tuning(_lib.TUNE_F16_FLUSH, flush)
    eng.plan.set_tuning(_lib.TUNE_CHUNK_MIN, 2)  # park even at these small lmax
    # the workspace is sized for the plan's accumulation form
    eng.F = _lib.workspace(eng.plan.workspace_bytes(S, 4))
    eng.stage_fields()
    eng.stage_operands()
    return eng


def _ring_bytes(eng):
    nb = (eng.S + 7) // 8 * 8
    return (4 * eng.nside - 1) * (eng.lmax + 1) * nb * 8


def test_grouped_contraction_equals_single_launches():
    geoms = [(128, 255, 20, 1, 2, -2), (64, 127, 9, 2, 1, -1), (32, 63, 40, 3, 2, -2),
             (16, 40, 3, 4, 1, -2), (8, 30, 33, 5, 2, -1)]
    run_grouped(geoms)
    # the other parking form, in its own launch (one form per grid)
    run_grouped([(128, 255, 20, 1, 2, -3), (32, 63, 40, 3, 1, -3), (16, 40, 3, 4, 2, -1)])


def run_grouped(geoms):
    engs = [_engine(*g) for g in geoms]
    torch.cuda.synchronize()
    refs = []
    for e in engs:
        # bytes of the compared range that no launch writes (past the last
        # column group) must hold the same value in both runs
        e.F[: _ring_bytes(e)].zero_()
        e.stage_contract()
        torch.cuda.synchronize()
        refs.append(e.F[: _ring_bytes(e)].clone())
        e.F[: _ring_bytes(e)].zero_()
    torch.cuda.synchronize()
    _lib.contract_grouped([e.plan for e in engs], [e.S for e in engs], [e.F for e in engs])
    torch.cuda.synchronize()
    for e, r, g in zip(engs, refs, geoms):
        assert torch.equal(e.F[: _ring_bytes(e)], r), g
    # and the maps made from the grouped ring buffers are the maps of the
    # one-geometry path
    for e in engs:
        e.stage_rings()
    torch.cuda.synchronize()
    for e in engs:
        m = e.maps.clone()
        e.run()
        torch.cuda.synchronize()
        assert torch.equal(m, e.maps)


def test_grouped_contraction_rejects():
    e1 = _engine(32, 63, 4, 1, 2, -1)
    e2 = _engine(16, 31, 4, 2, 2, -1)
    e2.plan.set_tuning(_lib.TUNE_MFMA_PATH, 1)
    with pytest.raises(NotImplementedError, match="FP32 MFMA path"):
        _lib.contract_grouped([e1.plan, e2.plan], [4, 4], [e1.F, e2.F])
    e2.plan.set_tuning(_lib.TUNE_MFMA_PATH, 2)
    e2.plan.set_tuning(_lib.TUNE_F16_FLUSH, 8)
    with pytest.raises(NotImplementedError, match="master sums"):
        _lib.contract_grouped([e1.plan, e2.plan], [4, 4], [e1.F, e2.F])
    e2.plan.set_tuning(_lib.TUNE_F16_FLUSH, -3)
    e1.plan.set_tuning(_lib.TUNE_F16_FLUSH, -2)
    e2.F = _lib.workspace(e2.plan.workspace_bytes(4, 4))
    e1.F = _lib.workspace(e1.plan.workspace_bytes(4, 4))
    with pytest.raises(NotImplementedError, match="one two-chunk parking form"):
        _lib.contract_grouped([e1.plan, e2.plan], [4, 4], [e1.F, e2.F])
    e2.plan.set_tuning(_lib.TUNE_F16_FLUSH, -1)
    with pytest.raises(RuntimeError, match="workspace"):
        _lib.contract_grouped([e1.plan, e2.plan], [4, 4], [e1.F, e2.F[:16]])
    with pytest.raises(ValueError):
        _lib.contract_grouped([e1.plan] * 9, [4] * 9, [e1.F] * 9)
    torch.cuda.synchronize()


def test_graph_entry_points_replay_a_batched_synthesis():
    nside, lmax, S = 64, 127, 6
    eng = LognormalShells(shell_spectra(S, lmax), nside, ncorr=3, seed=5, lmax=lmax)
    eng.run()
    torch.cuda.synchronize()
    ref = eng.maps.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        g = _lib.Graph().capture(eng.run)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        eng.maps.fill_(float("nan"))
        eng.F.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.launch()
        s.synchronize()
        assert torch.equal(eng.maps, ref), rep
    # the null stream cannot be captured
    with pytest.raises(ValueError, match="null stream"):
        _lib.call("ga_graph_begin", None)
    del g
    torch.cuda.synchronize()
