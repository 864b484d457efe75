"""Spectra preparation on the device: Gauss-Legendre transforms (against the
oracle's numpy restatement), the Gaussian-spectra solver (the reference's own
solver tests, tests/core/grf/test_solver.py), solve_gaussian_spectra and
discretized_cls.

The transforms replace transformcl's (not available here): the solver's
solutions agree with the reference's to the discretisation of the transform,
which is not pinned; its convergence properties are.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import glass_amd  # noqa: E402
from glass_amd import grf  # noqa: E402


@pytest.fixture
def cl():
    lmax = 100
    ell = np.arange(lmax + 1)
    return 1e-2 / (2 * ell + 1) ** 2


@pytest.mark.parametrize("n", [1, 2, 5, 64, 301])
def test_nodes_and_transforms_match_oracle(n):
    print(f"gauss-legendre n={n}", flush=True)
    x, w = (t.cpu().numpy() for t in grf._nodes(n))
    rx, rw = O.gl_nodes(n)
    # float64 Newton on the device, extended precision in the oracle: nodes to
    # an ulp, weights to the conditioning of 1 - x^2 next to +-1
    assert np.max(np.abs(x - rx)) <= 2.3e-16 and np.max(np.abs(w - rw) / rw) < 1e-11
    rng = np.random.default_rng(n)
    c = rng.random(n) / (1.0 + np.arange(n)) ** 2
    corr = grf.cltocorr(c)
    assert np.max(np.abs(corr - O.gl_cltocorr(c))) < 1e-13 * max(1.0, np.max(np.abs(corr)))
    back = grf.corrtocl(corr)
    assert np.max(np.abs(back - c)) < 1e-12


def test_compute_normal_is_identity_and_lognormal_relation(cl):
    assert np.allclose(grf.compute(cl, grf.Normal()), cl, rtol=0, atol=1e-15)
    t = grf.Lognormal(0.7)
    gl = grf.compute(cl, t)
    # C(theta) of the result is log1p of the target's over lambda^2
    assert np.allclose(grf.cltocorr(gl), np.log1p(grf.cltocorr(cl) / 0.49), rtol=1e-10,
                       atol=1e-16)


def test_solver_reference_properties(cl):
    """tests/core/grf/test_solver.py:20-116."""
    rng = np.random.default_rng(0)
    lam = rng.random()
    t = grf.Lognormal(lam)
    gl1, _, _ = grf.solve(cl, t)
    gl2, _, _ = grf.solve(cl, t, t)
    assert np.array_equal(gl1, gl2)
    _, cl_out, _ = grf.solve(cl, t, pad=2 * cl.shape[0])
    assert cl_out.shape[0] == 3 * cl.shape[0]
    with pytest.raises(ValueError, match="pad must be a positive integer"):
        grf.solve(cl, t, pad=-1)
    gl = grf.compute(cl, t)
    gl1, _, _ = grf.solve(cl, t)
    gl2, _, _ = grf.solve(cl, t, initial=gl)
    assert np.array_equal(gl1, gl2)
    t0 = grf.Lognormal()
    assert np.array_equal(grf.compute(cl, t0), grf.solve(cl, t0, maxiter=0)[0])
    # lognormal with a fixed monopole converges to cltol
    gl0 = rng.random()
    cltol = 1e-7
    gl, cl_, info = grf.solve(cl, t0, grf.Lognormal(), monopole=gl0, cltol=cltol)
    assert info > 0
    assert np.allclose(cl_[1: cl.shape[0]], cl[1:], atol=0.0, rtol=cltol)
    gl_ = grf.compute(cl_, t0, grf.Lognormal())
    assert gl[0] == gl0
    assert np.allclose(gl_[1: gl.shape[0]], gl[1:])
    # monopole free / fixed
    c = cl.copy()
    c[0] = rng.random()
    gl, cl_out, _ = grf.solve(c, t0, monopole=None, gltol=1e-8)
    assert gl[0] != 0.0 and np.isclose(cl_out[0], c[0])
    gl, cl_out, _ = grf.solve(c, t0, monopole=gl0, gltol=1e-8)
    assert gl[0] == gl0 and not np.isclose(cl_out[0], c[0])


def test_solve_gaussian_spectra_and_discretized_cls():
    """glass/fields.py:239-300, :745-836 on three lognormal fields."""
    lmax, n = 200, 3
    ell = np.arange(lmax + 1)
    auto = [1e-3 * (i + 1) / (ell + 1.0) ** 2 for i in range(n)]
    spectra = []
    for i in range(n):
        for j in range(i, -1, -1):
            spectra.append(auto[i] if i == j else 0.4 * np.sqrt(auto[i] * auto[j]))
    fields = [grf.Lognormal(1.0) for _ in range(n)]
    gls = glass_amd.solve_gaussian_spectra(fields, spectra)
    assert len(gls) == len(spectra)
    for (i, j, cl_t), gl in zip(glass_amd.enumerate_spectra(spectra), gls):
        # the solution's transformed spectrum reproduces the target
        rl = grf.corrtocl(grf.corr(fields[i], fields[j], grf.cltocorr(np.pad(gl, (0, 2 * gl.size)))))
        assert np.allclose(rl[1: lmax + 1], cl_t[1:], rtol=1e-4, atol=1e-12), (i, j)
    bl = glass_amd.compute_gaussian_spectra(fields, spectra)
    assert len(bl) == len(spectra) and bl[0].shape == spectra[0].shape
    with pytest.raises(ValueError, match="mismatch between number of fields and spectra"):
        glass_amd.solve_gaussian_spectra(fields[:2], spectra)
    # discretisation: truncation, correlation cut and the squared pixel window
    d = glass_amd.discretized_cls(spectra, lmax=100, ncorr=1, nside=64)
    pw = glass_amd.healpix.pixwin(64, lmax=100)
    assert d[0].shape == (101,) and np.allclose(d[0], spectra[0][:101] * pw**2, rtol=1e-15)
    assert d[5].shape == (0,)  # (2, 0): beyond ncorr = 1
    assert glass_amd.discretized_cls([]) == []
