"""Throughput of the lognormal-shell generation path on one MI355X.

Workload (the headline configuration): S correlated lognormal shells at
nside = 2048, lmax = 4096, ncorr = 3 (neighbour correlation 0.6 / 0.3 / 0.1,
auto-spectra 1e-4 (l+1)^-1.2 exp(-(l/1500)^2) with amplitude rising x1 -> x4
across shells), normals from the device Philox generator.  One step = one
batch of S shells through the whole path, inputs resident in HBM:

    iternorm -> correlated alm (+ healpy reorder, m = 0 fix)
    -> Legendre stage on MFMA (operand preparation, then the contraction with
       the recurrence in FP64; split-FP16 products by default)
    -> ring FFT + lognormal.

value = maps/s over the timed steps.  Also reported: the plain FP64 GPU path's
maps/s on the same configuration, the FP32 maps' error against the FP64 path
(same normals), the roofline of the Legendre contraction kernel (algorithmic
dense FLOP/s from HIP events on the launch stream), per-step p50/p99 latency,
an uneven batch under one HIP graph, the serving loop's p50/p99 request
latency over a slice of the 600 s serving workload (``--serve-seconds``; the
full run is ``--serve-seconds 600``), and the FP64 CPU oracle timed on one
whole shell on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

MFMA_FP32_PEAK = 157.3e12  # MI355X dense FP32 matrix FLOP/s
MFMA_FP16_PEAK = 2.5e15    # MI355X dense FP16 matrix FLOP/s
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--shells", type=int, default=64)
    ap.add_argument("--fp64-maps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="seconds of Poisson arrivals for the serving-loop section (0 = skip)")
    ap.add_argument("--uneven-maps", type=int, default=128,
                    help="maps in the uneven-batch section (0 = skip)")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="skip the kernels of the rows beside the hot path")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first (a known-size "
                         "kernel that calibrates hardware byte counters)")
    return ap.parse_args()


def read_traffic(nside: int, lmax: int, shells: int):
    """HBM bytes per launch of the Legendre contraction kernel from the
    committed rocprof PMC summary of this workload (its line
    'traffic_bytes_per_launch: <n>': FETCH_SIZE + WRITE_SIZE, calibrated)."""
    p = ROOT / "profiles" / f"pmc_lognormal_nside{nside}_lmax{lmax}_S{shells}.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(nside: int, lmax: int, gls):
    """The FP64 oracle on the host: one whole shell of the workload through the
    full pipeline (cls2cov, iternorm, numpy PCG64 normals, multalm sum, healpy
    reorder, synthesis of every ring, lognormal).  The synthesis is the
    oracle's symmetric form (one recurrence per north ring and m serves its
    south mirror; values below 2^-900 left out), on every host core."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    alm = next(O.generate_grf_alm(gls[:1], ncorr=0, rng=np.random.default_rng(0)))
    t_alm = time.perf_counter() - t0
    t0 = time.perf_counter()
    x = O.alm2map_sym(alm, nside, nthreads=cores)
    t_synth = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.lognormal(x, O.cltovar(gls[0]))
    t_ln = time.perf_counter() - t0
    total = t_alm + t_synth + t_ln
    return {
        "value": 1.0 / total, "unit": "maps/s", "cores": cores, "kind": "port",
        "sample": (f"1 whole lognormal shell at nside {nside}, lmax {lmax} through the FP64 "
                   f"C/numpy oracle on {cores} host threads (synthesis with equatorial "
                   f"symmetry, all {4 * nside - 1} rings, nothing extrapolated): alm "
                   f"{t_alm:.2f} s, synthesis {t_synth:.1f} s, lognormal {t_ln:.2f} s"),
    }


def serving_section(seconds: float):
    """Serving loop: Poisson arrivals at 5 requests/s (seed 4) for ``seconds``,
    nside in {256, 512, 1024, 2048, 4096} (weights .3/.3/.2/.15/.05), 10-80
    shells, lmax = 2 nside, ncorr 3, FP32 MFMA synthesis + lognormal,
    admission against a 240 GB HBM budget, continuous batching per geometry."""
    import torch

    from glass_amd import _lib, serving

    reqs = serving.poisson_requests(rate=5.0, duration=seconds, seed=4)
    ex = serving.DeviceExecutor()
    for ns in sorted({r.nside for r in reqs}):  # plans are built at server start-up
        _lib.get_plan(ns, 2 * ns)
    torch.cuda.synchronize()

    def progress(now, done, total):
        print(f"[bench] serving t={now:.0f}s completed {done}/{total}", file=sys.stderr, flush=True)

    sched = serving.run(reqs, ex, 240e9, progress=progress)
    out = serving.summarize(sched)
    out.update({"arrival_seconds": seconds,
                "slice": ("full 600 s arrival process" if seconds >= 600 else
                          f"first {seconds:.0f} s of the 600 s arrival process"),
                "rate_per_s": 5.0, "seed": 4, "budget_gb": 240,
                "wall_s": sched.wall,
                "latency_definition": "arrival to all maps of the request resident in HBM"})
    return out


def uneven_section(nmaps: int):
    """Uneven batch: nmaps maps (seed 3), nside uniform over {256..2048}, lmax =
    2 nside - 1, per-map tilt, lognormal + shear, captured in one HIP graph."""
    import torch

    from glass_amd.uneven import UnevenBatch, uneven_specs

    ub = UnevenBatch(uneven_specs(n=nmaps, seed=3), shear=True, tune=True)
    ub.capture()
    torch.cuda.synchronize()
    ts = []
    for _ in range(2):
        t0 = time.perf_counter()
        ub.replay()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ub.check()
    errs = {}
    for key, eng in ub.engines.items():
        acc = accuracy_all_shells(eng, eng.S)
        errs[f"nside{key[0]}"] = {k: acc[k] for k in ("shells_checked", "worst_max_rel",
                                                       "max_rms_rel", "median_max_rel")}
    t = min(ts)
    return {"maps": nmaps, "groups": {f"nside{k[0]}": len(v) for k, v in ub.groups.items()},
            "graph_replay_ms": 1e3 * t, "maps_per_s": nmaps / t, "shear": True,
            "tuning": {f"nside{k[0]}": {"path": c["path"], "nt": c["nt"]}
                       for k, c in ub.tuning.items()},
            "accuracy_vs_fp64_all_maps": errs}


def extras_section(nside: int):
    """The kernels of the rows beside the hot path, each timed with HIP events
    on the launch stream against its own bound, on maps of the headline
    geometry (nside 2048):

    * multi-plane convergence update, float32 maps: reads kappa1, kappa2,
      delta2, writes kappa3 (16 bytes per pixel), HBM roofline;
    * positions from a lognormal density map, device Philox counts and
      positions, ~1e8 points: expected counts + Poisson counts (reads the map,
      writes int64 counts), the offset scan, the placement (reads counts,
      writes two float64 per point); points/s;
    * the pixel window of the geometry (spin 0 and 2, up to l = 4 nside);
    * the pixel quadrature weights of the geometry (exact to l = 2 nside);
    * the Gaussian-spectra solver on the 64 lognormal spectra of the headline
      batch (host loop over device Gauss-Legendre transforms)."""
    import torch

    from glass_amd import _lib, healpix, points
    from glass_amd.batch import shell_spectra
    from glass_amd import grf

    out = {}
    npix = 12 * nside * nside
    st = _lib.stream_ptr()

    def timed(fn, reps=5):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            a.record()
            fn()
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    g = torch.Generator(device="cuda").manual_seed(5)
    k1 = torch.randn(npix, device="cuda", generator=g)
    k2 = torch.randn(npix, device="cuda", generator=g)
    d2 = torch.randn(npix, device="cuda", generator=g)
    k3 = torch.empty_like(k1)
    ms = timed(lambda: _lib.call("ga_multi_plane_update", npix, 4, _lib.ptr(k1), _lib.ptr(k2),
                                 _lib.ptr(d2), 0.3, 0.01, _lib.ptr(k3), st))
    gbs = 16.0 * npix / (ms * 1e-3) / 1e9
    out["multi_plane_update"] = {"nside": nside, "ms": ms, "bytes": 16 * npix, "GB_per_s": gbs,
                                 "frac_of_8TBs": gbs / 8000.0}
    del k1, k2, k3

    delta = torch.expm1(0.5 * d2.double() - 0.125)  # lognormal contrast, mean ~0
    nb = _lib.ctypes.c_size_t()
    _lib.call("ga_points_workspace_size", npix, _lib.ctypes.byref(nb))
    ws = _lib.workspace(nb.value)
    counts = torch.empty(npix, dtype=torch.int64, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    ngal = 1e8 / points.ARCMIN2_SPHERE
    scale = points.ARCMIN2_SPHERE / npix * ngal

    def do_counts():
        _lib.call("ga_points_counts", npix, 8, _lib.ptr(delta), 1, 1.0, 0, scale, None, 7, 0,
                  None, _lib.ptr(counts), _lib.ptr(total), _lib.ptr(ws), ws.numel(), st)

    ms_counts = timed(do_counts)
    ntot = int(total.item())
    lon = torch.empty(ntot, dtype=torch.float64, device="cuda")
    lat = torch.empty(ntot, dtype=torch.float64, device="cuda")
    ms_pos = timed(lambda: _lib.call("ga_points_positions", nside, _lib.ptr(counts), 7, 0,
                                     _lib.ptr(lon), _lib.ptr(lat), _lib.ptr(ws), ws.numel(), st))
    out["positions_from_delta"] = {
        "nside": nside, "points": ntot, "counts_ms": ms_counts, "positions_ms": ms_pos,
        "points_per_s": ntot / ((ms_counts + ms_pos) * 1e-3),
        "positions_GB_per_s": (16.0 * ntot + 8.0 * npix) / (ms_pos * 1e-3) / 1e9}
    del lon, lat, counts, delta, d2, ws

    healpix._pixwin_cache.clear()
    t0 = time.perf_counter()
    healpix.pixwin(nside, lmax=4 * nside, pol=True)
    out["pixwin"] = {"nside": nside, "lmax": 4 * nside, "pol": True,
                     "s": time.perf_counter() - t0}

    hp_plan = _lib.get_plan(nside, 2 * nside)
    n = _lib.ctypes.c_size_t()
    _lib.call("ga_pixel_weights_workspace_size", hp_plan.h, _lib.ctypes.byref(n))
    pws = _lib.workspace(n.value)
    pw = torch.empty(npix, dtype=torch.float64, device="cuda")
    it, res = _lib.ctypes.c_int32(), _lib.ctypes.c_double()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.call("ga_pixel_weights", hp_plan.h, 1e-12, 1000, _lib.ptr(pw), _lib.ctypes.byref(it),
              _lib.ctypes.byref(res), _lib.ptr(pws), pws.numel(), st)
    torch.cuda.synchronize()
    out["pixel_weights"] = {"nside": nside, "exact_to_l": 2 * nside, "s": time.perf_counter() - t0,
                            "cg_iterations": it.value, "relative_residual": res.value}
    del pws, pw

    gls = shell_spectra(64, 2 * nside, amp=lambda i: 1.0 + 3.0 * i / 63, rho=(0.6, 0.3, 0.1),
                        scale=1500.0)
    fields = [grf.Lognormal(1.0) for _ in range(64)]
    import glass_amd

    t0 = time.perf_counter()
    sol = glass_amd.solve_gaussian_spectra(fields, gls)
    out["solve_gaussian_spectra"] = {"spectra": len(sol), "lmax": 2 * nside,
                                     "s": time.perf_counter() - t0}
    torch.cuda.empty_cache()
    return out


def accuracy_all_shells(eng, S: int):
    """Max and RMS error of every FP32 map against the FP64 path on the same
    normals, relative to the FP64 map's RMS."""
    import torch

    per = []
    for s in range(S):
        ref = eng.fp64_reference_map(s)
        got = eng.maps[s].double()
        rms = torch.sqrt(torch.mean(ref * ref))
        per.append((float(torch.max(torch.abs(got - ref)) / rms),
                    float(torch.sqrt(torch.mean((got - ref) ** 2)) / rms)))
    worst = int(np.argmax([p[0] for p in per]))
    return {"shells_checked": S, "worst_shell": worst, "worst_max_rel": per[worst][0],
            "max_rms_rel": max(p[1] for p in per),
            "median_max_rel": float(np.median([p[0] for p in per])),
            "per_shell_max_rel": [float(f"{p[0]:.3e}") for p in per]}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(shells and requests are independent, so N GPUs run N independent batches).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import glass_amd
    from glass_amd.batch import LognormalShells, shell_spectra

    nside = args.nside
    lmax = args.lmax if args.lmax is not None else 2 * nside
    S = args.shells
    torch.cuda.set_device(0)
    gls = shell_spectra(S, lmax, amp=lambda i: 1.0 + 3.0 * i / max(S - 1, 1),
                        rho=(0.6, 0.3, 0.1), scale=1500.0)
    print(f"[bench] building batch: nside={nside} lmax={lmax} shells={S}", file=sys.stderr,
          flush=True)
    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    eng = LognormalShells(gls, nside, ncorr=3, shift=1.0, seed=1234, lmax=lmax)
    for _ in range(args.warmup):
        eng.run()
    torch.cuda.synchronize()
    eng.check()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # HIP events on the stream the library launches on (torch's current stream,
    # which every stage passes to the C ABI)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        e = ev[k]
        e[0].record()
        eng.stage_fields()
        e[1].record()
        eng.stage_operands()
        e[2].record()
        eng.stage_contract()
        e[3].record()
        eng.stage_rings()
        e[4].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.check()
    wall = t1 - t0

    def span(a, b):
        return float(np.mean([ev[k][a].elapsed_time(ev[k][b]) for k in range(args.steps)]))

    step_ms = [ev[k][0].elapsed_time(ev[k][4]) for k in range(args.steps)]
    fields_ms, prep_ms, leg_ms, ring_ms = span(0, 1), span(1, 2), span(2, 3), span(3, 4)
    from glass_amd import _lib

    path = 1 if eng.plan.get_tuning(_lib.TUNE_MFMA_PATH) == 1 else 2
    value = args.steps * S / wall
    print(f"[bench] timed {args.steps} steps: {wall:.3f} s", file=sys.stderr, flush=True)

    # accuracy of every FP32 map against the FP64 path on identical normals
    print("[bench] accuracy of all shells against FP64", file=sys.stderr, flush=True)
    errs = accuracy_all_shells(eng, S)

    # plain FP64 GPU path: one map per launch
    fp64 = None
    if args.fp64_maps > 0:
        eng.fp64_reference_map(0)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for i in range(args.fp64_maps):
            eng.fp64_reference_map(i % S)
        torch.cuda.synchronize()
        fp64 = args.fp64_maps / (time.perf_counter() - ta)

    nalm = (lmax + 1) * (lmax + 2) // 2
    flop = 4.0 * nalm * (2 * nside) * S  # dense Legendre contraction, ring pairs
    achieved = flop / (leg_ms * 1e-3)
    traffic = read_traffic(nside, lmax, S)
    if path == 1:
        peak, kname, peak_note = MFMA_FP32_PEAK, "legendre_synth_mfma_kernel", "FP32 MFMA dense"
    else:
        # three FP16 products (hi*hi + hi*lo + lo*hi) per FP32-accurate product
        peak, kname = MFMA_FP16_PEAK / 3.0, "legendre_synth_f16_kernel"
        peak_note = ("FP16 MFMA dense 2.5 PFLOP/s / 3: the split-FP16 contraction issues "
                     "three FP16 products per algorithmic FP32 product")
    out = {
        "metric": "lognormal maps/s at nside 2048 (FP32-MFMA vs FP64 plain, same GPU); p99 latency",
        "value": value,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "batched lognormal shells", "nside": nside, "lmax": lmax,
                   "shells_per_batch": S, "ncorr": 3, "rng": "device Philox4x32-10",
                   "stages": "iternorm + alm + MFMA Legendre (FP64 recurrence) + ring FFT "
                             "+ lognormal",
                   "legendre_path": "FP32 MFMA" if path == 1 else "split-FP16 MFMA"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": kname, "flop_per_launch": flop, "avg_launch_ms": leg_ms,
                     "peak_note": peak_note},
        "stages_ms": {"fields": fields_ms, "legendre_operands": prep_ms, "legendre_mfma": leg_ms,
                      "ring_fft_lognormal": ring_ms},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99)), "unit": "per batch"},
        "fp64_plain_maps_per_s": fp64,
        "speedup_vs_fp64_plain": (value / fp64) if fp64 else None,
        "accuracy_vs_fp64": errs,
        "cpu_baseline": None,
    }
    # the extra sections run after the headline numbers, on freed memory
    del eng
    torch.cuda.empty_cache()
    if args.uneven_maps > 0:
        print("[bench] uneven batch", file=sys.stderr, flush=True)
        try:
            out["uneven_batch"] = uneven_section(args.uneven_maps)
        except Exception as e:  # the headline line is still printed
            out["uneven_batch"] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    if args.serve_seconds > 0:
        print("[bench] serving loop", file=sys.stderr, flush=True)
        try:
            out["serving"] = serving_section(args.serve_seconds)
        except Exception as e:
            out["serving"] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    if args.extras:
        print("[bench] extras", file=sys.stderr, flush=True)
        try:
            out["extras"] = extras_section(nside)
        except Exception as e:
            out["extras"] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(nside, lmax, gls)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
