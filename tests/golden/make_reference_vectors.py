"""Writes reference_vectors.json: the literal known-answer vectors that the
reference's own unit tests hold for this path, as data (inputs and expected
outputs), plus a short numpy PCG64 stream that pins the RNG the reference draws
its normals from (glass/fields.py:407, glass/rng.py:18).

Sources (reference files, read as text):
  tests/core/test_fields.py:41-63     iternorm covariances
  tests/core/test_fields.py:166-252   cls2cov inputs and expected columns
  tests/core/test_fields.py:649-655   GLASS -> healpy alm permutation
  tests/core/test_harmonics.py:18-52  multalm inputs and expected outputs
"""

import json
import pathlib

import numpy as np

OUT = pathlib.Path(__file__).with_name("reference_vectors.json")

data = {
    "iternorm_cov_0d": [[1.0, 0.2, 0.1], [0.2, 0.5, 0.2], [0.1, 0.2, 0.3]],
    "iternorm_cov_nd": [
        [[1.0, 0.2, 0.1], [0.2, 0.5, 0.2], [0.1, 0.2, 0.3]],
        [[1.4, 0.4, 0.5], [0.4, 1.5, 0.8], [0.5, 0.8, 1.3]],
    ],
    "cls2cov_single": {
        "cls": [[1.0, 0.5, 0.3], None, [0.7, 0.6, 0.1]],
        "nl": 3, "nf": 2, "nc": 2,
        "col0": [0.5, 0.25, 0.15],
    },
    "cls2cov_negative": {"cls": [[-1.0, 0.5, 0.3], [0.8, 0.4, 0.2], [0.7, 0.6, 0.1]],
                         "nl": 3, "nf": 2, "nc": 2},
    "cls2cov_multi": {
        "cls": [[1.0, 0.5, 0.3], [0.8, 0.4, 0.2], [0.7, 0.6, 0.1],
                [0.9, 0.5, 0.3], [0.6, 0.3, 0.2], [0.8, 0.7, 0.4]],
        "nl": 3, "nf": 3, "nc": 2,
        "first_col0": [0.5, 0.25, 0.15],
        "final_col0": [0.45, 0.25, 0.15],
    },
    "glass_to_healpix_alm": {
        "in": [0, 10, 11, 20, 21, 22, 30, 31, 32, 33],
        "out": [0, 10, 20, 30, 11, 21, 31, 22, 32, 33],
    },
    "multalm": [
        {"alm": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], "bl": [2.0, 0.5, 1.0],
         "out": [2.0, 1.0, 1.5, 4.0, 5.0, 6.0]},
        {"alm": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], "bl": [1.0, 1.0, 1.0],
         "out": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]},
        {"alm": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], "bl": [0.0, 1.0, 0.0],
         "out": [0.0, 2.0, 3.0, 0.0, 0.0, 0.0]},
        {"alm": [], "bl": [], "out": []},
    ],
    "pcg64_seed42_normals_6x2": np.random.default_rng(42).standard_normal((6, 2)).tolist(),
}

OUT.write_text(json.dumps(data, indent=1) + "\n")
print("wrote", OUT)
