"""Pin the oracle's GLL rule and edge maps.

* against the reference's own femlib/polylib.c (zwglj, Dglj) and
  femlib/mapping.c (edgemaps), compiled unmodified into oracle/_ref;
* against what the mathematics fixes: exact integration of polynomials of
  degree <= 2N-1 and exact differentiation of degree <= N (the checks of
  femlib/tests/polylib_test.c:92,125-260, EPS 1e-12).
"""
import ctypes as C

import numpy as np
import pytest

from semtex_amd import mesh as M


# np = 2 is excluded: the reference's Dglj reads past its arrays there
# (jacobd with zero interior points); semtex sessions use N_P >= 3.
@pytest.mark.parametrize("np_", range(3, 17))
def test_gll_matches_reference_polylib(oracle, np_):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    zr, wr = np.zeros(np_), np.zeros(np_)
    Dr, DTr = np.zeros(np_ * np_), np.zeros(np_ * np_)
    R.zwglj(zr, wr, np_, 0.0, 0.0)
    R.Dglj(Dr, DTr, zr, np_, 0.0, 0.0)
    z, w, D, DT = oracle.gll(np_)
    assert np.max(np.abs(z - zr)) < 4e-15
    assert np.max(np.abs(w - wr)) < 4e-15
    scale = max(1.0, np.abs(Dr).max())
    assert np.max(np.abs(D.ravel() - Dr)) / scale < 2e-14
    assert np.max(np.abs(DT.ravel() - DTr)) / scale < 2e-14


@pytest.mark.parametrize("np_", range(2, 17))
def test_edgemaps_match_reference_mapping(oracle, np_):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    em_p, pm_p = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    R.edgemaps(np_, 2, C.byref(em_p), C.byref(pm_p))
    n = np_ * np_
    em_ref = np.array([em_p[i] for i in range(n)])
    pm_ref = np.array([pm_p[i] for i in range(n)])
    em, pm = oracle.edgemaps(np_)
    assert np.array_equal(em, em_ref)
    assert np.array_equal(pm, pm_ref)
    em2, pm2 = M.edgemaps(np_)
    assert np.array_equal(em2, em_ref) and np.array_equal(pm2, pm_ref)


@pytest.mark.parametrize("np_", [3, 5, 8, 10, 12, 16])
def test_gll_polynomial_exactness(oracle, np_):
    z, w, D, _ = oracle.gll(np_)
    N = np_ - 1
    for k in range(0, 2 * N):          # exact for degree <= 2N-1
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-13
    for k in range(0, N + 1):          # D exact for degree <= N
        d = D @ z ** k
        ex = k * z ** (k - 1) if k else np.zeros_like(z)
        assert np.max(np.abs(d - ex)) < 1e-11 * max(1, k * k)


@pytest.mark.parametrize("np_", [2, 4, 8, 10, 12])
def test_host_gll_equals_oracle(oracle, np_):
    """The product's host-side rule (semtex_amd.mesh.gll) is the same rule."""
    z, w, D, _ = oracle.gll(np_)
    z2, w2, D2 = M.gll(np_)
    assert np.max(np.abs(z - z2)) < 4e-15
    assert np.max(np.abs(w - w2)) < 4e-15
    assert np.max(np.abs(D - D2)) < 1e-13 * max(1, np_ * np_)


def test_dft_packing_matches_numpy_rfft(oracle):
    """femlib/fourier.c:40-165: forward scaled 1/N, packing
    [Re0, ReNyq, Re1, Im1, ...], exp(-i) convention (FFTPACK, which the
    reference's femlib/tests/testFFT.C cross-checks against Temperton)."""
    rng = np.random.default_rng(1)
    nz, ntrn = 16, 6
    data = rng.standard_normal((nz, ntrn))
    out = oracle.dft(data, nz, ntrn, +1)
    F = np.fft.rfft(data, axis=0) / nz
    assert np.allclose(out[0], F[0].real, atol=1e-14)
    assert np.allclose(out[1], F[nz // 2].real, atol=1e-14)
    for k in range(1, nz // 2):
        assert np.allclose(out[2 * k], F[k].real, atol=1e-14)
        assert np.allclose(out[2 * k + 1], F[k].imag, atol=1e-14)
    back = oracle.dft(out, nz, ntrn, -1)
    assert np.allclose(back, data, atol=1e-13)
