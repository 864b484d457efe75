"""Autotuner of the batched MFMA Legendre stage.

For one geometry and batch width it times the candidate tile shapes of the
Legendre contraction -- split FP16 with one or two 32-column tiles per wave,
and the FP32 MFMA path with two tiles -- on the engine's own alm (HIP events on the
launch stream, median of ``reps``), keeps the fastest, and caches the choice
per (nside, lmax, shells) in a JSON file (``GLASS_AMD_TUNE_CACHE``, default
``~/.cache/glass_amd/legendre_tune.json``).  The choice is stored in the
engine's plan (``ga_plan_set_tuning``): every launch on that plan reads it,
also during graph capture, and plans of other geometries and the process
default are left as they were.
"""

from __future__ import annotations

import json
import os
import pathlib

import numpy as np

from . import _lib

# (path, tiles): path 2 = split FP16 with one or two 32-column tiles per wave,
# path 1 = FP32 MFMA with two tiles
CANDIDATES = ((2, 1), (2, 2), (1, 2))
# the tile shapes a grouped (several geometries in one grid) launch runs
GROUPED_CANDIDATES = ((2, 1), (2, 2))


def cache_path() -> pathlib.Path:
    p = os.environ.get("GLASS_AMD_TUNE_CACHE")
    if p:
        return pathlib.Path(p)
    return pathlib.Path.home() / ".cache" / "glass_amd" / "legendre_tune.json"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _save(d: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(d, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


_KEYS = (("path", _lib.TUNE_MFMA_PATH), ("nt", _lib.TUNE_MFMA_TILES),
         ("chunk", _lib.TUNE_MFMA_CHUNK))


def apply(plan, cfg: dict) -> None:
    """Store a Legendre configuration in ``plan`` (a ``_lib.Plan``)."""
    plan.set_tuning(_lib.TUNE_MFMA_PATH, int(cfg["path"]))
    plan.set_tuning(_lib.TUNE_MFMA_TILES, int(cfg["nt"]))
    plan.set_tuning(_lib.TUNE_MFMA_CHUNK, int(cfg.get("chunk", 0)))


def current(plan) -> dict:
    return {name: plan.get_tuning(key) for name, key in _KEYS}


def tune_legendre(eng, candidates=CANDIDATES, reps: int = 3, use_cache: bool = True) -> dict:
    """Pick the fastest Legendre configuration for a ``LognormalShells``-like
    engine (attributes nside, lmax, S, plan and methods stage_fields /
    stage_legendre) and store it in the engine's plan."""
    T = _lib.torch()
    key = f"{eng.nside}:{eng.lmax}:{eng.S}"
    if tuple(candidates) != CANDIDATES:
        key += ":" + ",".join(f"{p}/{t}" for p, t in candidates)
    cache = _load() if use_cache else {}
    if key in cache:
        apply(eng.plan, cache[key])
        return cache[key]
    eng.stage_fields()
    times = {}
    for path, nt in candidates:
        cfg = {"path": path, "nt": nt, "chunk": 0}
        apply(eng.plan, cfg)
        eng.stage_legendre()  # warm-up (code load, caches)
        ts = []
        for _ in range(reps):
            a = T.cuda.Event(enable_timing=True)
            b = T.cuda.Event(enable_timing=True)
            a.record()
            eng.stage_legendre()
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        times[f"{path}/{nt}"] = float(np.median(ts))
    best = min(times, key=times.get)
    path, nt = (int(x) for x in best.split("/"))
    cfg = {"path": path, "nt": nt, "chunk": 0, "ms": times[best], "all_ms": times}
    apply(eng.plan, cfg)
    if use_cache:
        cache = _load()
        cache[key] = cfg
        _save(cache)
    return cfg
