"""The oracle's z transforms against the reference's own dDFTr
(femlib/fourier.c:40-165 with the Temperton FFT of femlib/temfftd.F,
compiled unmodified into oracle/_ref and linked to the image's BLAS).

Both oracle restatements are pinned: or_dft (plain C, the oracle of the
library's transform tests) and fwd/inv of the DNS oracle (numpy), in both
directions, for every transform length the reference accepts that the
tests and workloads use (products of 2, 3 and 5), on seeded data laid out
as the reference's AuxField planes (ntrn interleaved transforms).
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libfourier_ref.so")

LENGTHS = [2, 4, 6, 8, 10, 12, 16, 30, 32, 48, 64, 128, 256]


@pytest.fixture(scope="module")
def dDFTr():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libfourier_ref.so not built (reference tree, Fortran or BLAS absent)")
    L = C.CDLL(REF_LIB)
    L.dDFTr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.dDFTr.restype = None

    def run(x, sign):
        d = np.ascontiguousarray(x, np.float64).copy()
        L.dDFTr(d.ctypes.data_as(C.c_void_p), d.shape[0], d.shape[1], sign)
        return d
    return run


@pytest.mark.parametrize("nz", LENGTHS)
def test_oracle_transforms_equal_reference_dDFTr(oracle, dDFTr, nz):
    from oracle.dns_oracle import fwd, inv
    ntrn = 24
    rng = np.random.default_rng(1000 + nz)
    x = rng.standard_normal((nz, ntrn))
    ref_f = dDFTr(x, +1)
    scale = np.max(np.abs(ref_f))
    assert np.max(np.abs(oracle.dft(x, nz, ntrn, +1) - ref_f)) <= 4e-16 * nz * scale + 1e-300
    assert np.max(np.abs(fwd(x) - ref_f)) <= 4e-16 * nz * scale
    X = rng.standard_normal((nz, ntrn))
    ref_i = dDFTr(X, -1)
    scale = np.max(np.abs(ref_i))
    assert np.max(np.abs(oracle.dft(X, nz, ntrn, -1) - ref_i)) <= 4e-16 * nz * scale
    assert np.max(np.abs(inv(X) - ref_i)) <= 4e-16 * nz * scale


def test_reference_dDFTr_packing_known_answer(dDFTr):
    """A known answer the packing fixes: cos(2 pi z/L) + 0.5 sin(4 pi z/L) + 3
    on 8 planes transforms to mode 0 = 3, mode 1 = (1/2, 0), mode 2 =
    (0, -1/4), Nyquist 0 (forward scaled by 1/nz, e^{-ikz} convention)."""
    nz = 8
    z = np.arange(nz) / nz
    x = (np.cos(2 * np.pi * z) + 0.5 * np.sin(4 * np.pi * z) + 3.0)[:, None] * np.ones((1, 2))
    got = dDFTr(x, +1)[:, 0]
    want = np.array([3.0, 0.0, 0.5, 0.0, 0.0, -0.25, 0.0, 0.0])
    assert np.max(np.abs(got - want)) < 1e-15
