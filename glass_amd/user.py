"""The spectra file format of glass/user.py:42-86 (``.npz`` with the
concatenated values and the split points), the input format of a serving
request's spectra."""

from __future__ import annotations

import numpy as np


def save_cls(filename, cls) -> None:
    """glass/user.py:42-62: one ``values`` array and the ``split`` offsets."""
    split = np.cumsum([np.asarray(cl).shape[0] for cl in cls[:-1]])
    values = np.concatenate([np.asarray(cl) for cl in cls])
    np.savez(filename, values=values, split=split)


def load_cls(filename):
    """glass/user.py:65-86 (no pickled objects are read)."""
    with np.load(filename, allow_pickle=False) as npz:
        values = npz["values"]
        split = npz["split"]
    return np.split(values, split)
