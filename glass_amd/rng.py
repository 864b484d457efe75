"""Random-number sources for the generation path (glass/rng.py).

The reference draws its normals from a numpy PCG64 ``Generator`` (seed 42 by
default, glass/rng.py:18, :94-102).  Passing such a generator (or ``None``) to
``glass_amd.generate`` reproduces that stream exactly: the normals are drawn
on the host and uploaded.  ``DeviceRNG`` instead selects the counter-based
Philox4x32-10 generator that runs inside the alm kernel, keyed by
(seed, shell, GLASS alm index): no host work, no upload, and every shell's
normals are independent of how shells are batched.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SEED = 42


def default_rng(*, seed: int = SEED, xp=None) -> np.random.Generator:
    """glass/rng.py:21-55 for the numpy backend."""
    return np.random.default_rng(seed=seed)


@dataclass(frozen=True)
class DeviceRNG:
    """Counter-based device normals (statistically equivalent, not the numpy stream)."""

    seed: int = SEED
