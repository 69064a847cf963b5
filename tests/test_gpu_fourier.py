"""GPU parity of the z transform (sx_fourier_transform, the drop-in for
AuxField::transform / dDFTr) against the oracle's packed DFT
(femlib/fourier.c:40-165 conventions; the oracle itself is pinned to
numpy.fft.rfft).  Power-of-two nz run the LDS FFT in place and out of
place; other even nz run the matrix form.  Bar: 1e-13 relative to the
field's size (FFT and DFT round differently)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


SIZES = [(8, 484), (16, 1000), (32, 777), (64, 130), (128, 18432), (128, 35), (256, 333),
         (1024, 70), (12, 500), (2, 64)]
# the two-level transform (nz 32..256) also on its 256-thread tile
CASES = [(nz, S, "128") for nz, S in SIZES] + \
        [(nz, S, "256") for nz, S in SIZES if nz in (32, 64, 128, 256)]


@pytest.mark.parametrize("nz,S,tile", CASES)
def test_transform_matches_oracle(torch_cuda, oracle, monkeypatch, nz, S, tile):
    import torch
    from oracle import oracle as O
    from semtex_amd.solver import fourier_transform
    monkeypatch.setenv("SX_FFT_THREADS", tile)
    print("case: fourier nz", nz, "S", S, "tile", tile, flush=True)
    rng = np.random.default_rng(nz * 7 + S)
    x = rng.standard_normal((nz, S))
    want_f = O.dft(x, nz, S, +1)
    g = torch.from_numpy(x).cuda()
    if nz >= 8 and (nz & (nz - 1)) == 0:
        fourier_transform(g, +1)                       # in place
        got_f = g.cpu().numpy()
    else:
        out = torch.empty_like(g)
        fourier_transform(g, +1, out)
        got_f = out.cpu().numpy()
    assert np.max(np.abs(got_f - want_f)) < 1e-13 * np.max(np.abs(want_f)) * np.sqrt(nz)
    # inverse of the oracle's spectrum, out of place
    spec = torch.from_numpy(want_f).cuda()
    back = torch.empty_like(spec)
    fourier_transform(spec, -1, back)
    want_b = O.dft(want_f, nz, S, -1)
    got_b = back.cpu().numpy()
    assert np.max(np.abs(got_b - want_b)) < 1e-13 * np.max(np.abs(want_b)) * np.sqrt(nz)
    assert np.max(np.abs(got_b - x)) < 1e-12 * np.max(np.abs(x))


def test_transform_rejects_bad_arguments(torch_cuda):
    import torch
    from semtex_amd import _lib
    from semtex_amd.solver import fourier_transform
    g = torch.zeros((3, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.SxError):
        fourier_transform(g, +1)                       # odd nz
    h = torch.zeros((12, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.SxError):
        fourier_transform(h, +1)                       # matrix form in place
