"""Multi-plane convergence on device maps (glass/lensing.py:430-673).

The device update restates the reference's three float64 array operations
with the same roundings, so every convergence plane is bit-identical to the
oracle's numpy restatement, for float64 and float32 maps.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import glass_amd  # noqa: E402


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_multi_plane_bit_identical_to_oracle(cosmo, shells, dtype):
    npix = 12 * 32 * 32
    rng = np.random.default_rng(3)
    deltas = (rng.lognormal(0.0, 0.5, (len(shells), npix)) - 1.0).astype(dtype)
    ref = O.MultiPlaneConvergence(cosmo)
    dev = glass_amd.MultiPlaneConvergence(cosmo)
    tdev = glass_amd.MultiPlaneConvergence(cosmo)
    for i, w in enumerate(shells):
        print(f"multi-plane shell {i} dtype={np.dtype(dtype).name}", flush=True)
        ref.add_window(deltas[i].copy(), w)
        dev.add_window(deltas[i], w)
        tdev.add_window(torch.as_tensor(deltas[i]).cuda(), w)
        k = dev.kappa
        assert isinstance(k, np.ndarray) and k.dtype == dtype
        assert np.array_equal(k, ref.kappa), i
        kt = tdev.kappa
        assert isinstance(kt, torch.Tensor) and kt.is_cuda
        assert np.array_equal(kt.cpu().numpy(), ref.kappa), i
        assert dev.zsrc == ref.z3 and dev.wlens == ref.w3
        assert np.array_equal(dev.delta, deltas[i])


def test_multi_plane_matrix_and_weights_on_device(cosmo, shells):
    """tests/core/test_lensing.py:64-120 on the drop-in, and the matrix equal
    to the oracle's."""
    mat = glass_amd.multi_plane_matrix(shells, cosmo)
    assert np.array_equal(mat, O.multi_plane_matrix(shells, cosmo))
    assert np.array_equal(mat, np.tril(mat))
    rng = np.random.default_rng(4)
    deltas = rng.random((len(shells), 10))
    conv = glass_amd.MultiPlaneConvergence(cosmo)
    kappas = []
    for i, w in enumerate(shells):
        conv.add_window(deltas[i], w)
        kappas.append(conv.kappa.copy())
    assert np.allclose(mat @ deltas, np.stack(kappas), rtol=1e-12, atol=1e-15)
    weights = rng.random((len(shells), 3))
    wmat = glass_amd.multi_plane_weights(weights, shells, cosmo)
    assert np.array_equal(wmat, O.multi_plane_weights(weights, shells, cosmo))
    with pytest.raises(ValueError, match="shape mismatch between weights and shells"):
        glass_amd.multi_plane_weights(np.ones(2), shells, cosmo)


def test_multi_plane_errors(cosmo):
    conv = glass_amd.MultiPlaneConvergence(cosmo)
    conv.add_plane(np.zeros(8), 0.5)
    with pytest.raises(ValueError, match="source redshift must be increasing"):
        conv.add_plane(np.zeros(8), 0.5)
    with pytest.raises(ValueError, match="same shape"):
        conv.add_plane(np.zeros(9), 0.7)
