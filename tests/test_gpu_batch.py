"""Uneven batch: independent 2D problems of different order, size, aspect,
lambda2, Cartesian and cylindrical (axisymmetric, axis on y = 0), solved by
one captured HIP graph (sx_batch_*), its problems on concurrent branches.

* FP64 graph replays (three, with eager work and synchronisation between
  them) equal the eager per-problem solves exactly (same kernels: the
  graph records no split band sweeps, so the eager references run with
  SX_BAND_SPLIT=0), and the systems keep their split sweeps outside the
  batch;
* the eager FP64 solves match the oracle to 1e-12;
* the mixed-precision graph (FP32 element work, 3 FP64 refinement passes)
  is within 1e-10 of FP64 for every problem.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PICK = [0, 1, 2, 3, 4, 5, 32, 33, 34, 35, 36, 37]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def batch_problems():
    from semtex_amd.problems import uneven_batch
    allp = uneven_batch()
    return [allp[i] for i in PICK]


def _build(torch, probs, precision):
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    out = []
    for p in probs:
        m = p["mesh"]
        gm = GpuMesh(m)
        M = ModalMatrixSys(p["lambda2"], 0.0, 0, 1, gm, precision=precision)
        u = Field(gm, 1)
        src = torch.zeros((1, m.plane_size), dtype=torch.float64, device="cuda")
        src[0, :m.nel * m.npnp] = torch.from_numpy(p["forcing"][0])
        out.append((M, u, src, gm))
    return out


def test_batch_graph_matches_eager_and_oracle(torch_cuda, oracle, batch_problems, monkeypatch):
    import torch
    from semtex_amd.solver import SX_FP64, SX_MIXED, Batch
    print("case: uneven batch of", len(batch_problems), flush=True)
    built = _build(torch, batch_problems, SX_FP64)
    B = Batch([(M, u, src) for (M, u, src, _) in built])
    # the systems keep their own sweep choice outside the batch: a twisted
    # system still gets the split sweeps (two workgroups per chain) ...
    split_seen = False
    for (M, u, src, _) in built:
        u.solve(src.clone(), M)
        split_seen = split_seen or M.layout()["kernel"] == 3
    torch.cuda.synchronize()
    assert split_seen, [M.layout() for (M, _, _, _) in built]
    # ... while the graph's recorded solves use none (its branches run
    # concurrently); eager solves without them are its exact reference
    monkeypatch.setenv("SX_BAND_SPLIT", "0")
    eager = []
    for (M, u, src, _) in built:
        u.solve(src.clone(), M)
        eager.append(u.data.clone())
        assert M.layout()["kernel"] != 3, M.layout()
    monkeypatch.delenv("SX_BAND_SPLIT")
    torch.cuda.synchronize()
    # FP64 oracle parity of the eager solves
    for p, e in zip(batch_problems, eager):
        m = p["mesh"]
        n = m.nel * m.npnp
        uo = oracle.solve_modes(m, p["lambda2"], 0.0, p["forcing"][:, :n])[0]
        got = e[0, :n].cpu().numpy()
        assert np.linalg.norm(got - uo) / np.linalg.norm(uo) < 1e-12, (p["np"], p["cyl"])
    for rep in range(3):
        for (_, u, _, _) in built:
            u.data.zero_()
        junk = torch.randn(1000, device="cuda").sum()      # eager work between replays
        torch.cuda.synchronize()
        B.run()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for (_, u, _, _), e in zip(built, eager):
            assert torch.equal(u.data, e), rep
    # mixed precision graph
    builtm = _build(torch, batch_problems, SX_MIXED)
    Bm = Batch([(M, u, src) for (M, u, src, _) in builtm], refine_passes=3)
    for rep in range(2):
        Bm.run()
        torch.cuda.synchronize()
        for (p, (_, u, _, _), e) in zip(batch_problems, builtm, eager):
            n = p["mesh"].nel * p["mesh"].npnp
            rel = float(torch.linalg.norm(u.data[0, :n] - e[0, :n]) / torch.linalg.norm(e[0, :n]))
            assert rel < 1e-10, (rep, p["np"], p["lambda2"], p["cyl"], rel)


def test_batch_problems_sharing_a_system(torch_cuda, batch_problems, monkeypatch):
    """Problems that share a system (and so its workspace) run in order on
    one branch of the graph: three forcings on one system, two on another,
    each equal to its eager solve."""
    import torch
    from semtex_amd.solver import SX_FP64, Batch, Field
    print("case: batch, 5 problems on 2 systems", flush=True)
    built = _build(torch, batch_problems[:2], SX_FP64)
    entries, refs = [], []
    for j, (M, _, src, gm) in enumerate(built):
        for r in range(3 - j):
            u = Field(gm, 1)
            s = src * (1.0 + r) + 0.25 * r
            entries.append((M, u, s))
    B = Batch(entries)
    monkeypatch.setenv("SX_BAND_SPLIT", "0")          # as the graph's recorded solves
    for (M, u, s) in entries:
        v = Field(M.gmesh, 1)
        v.solve(s.clone(), M)
        refs.append(v.data.clone())
    monkeypatch.delenv("SX_BAND_SPLIT")
    torch.cuda.synchronize()
    for rep in range(2):
        for (_, u, _) in entries:
            u.data.zero_()
        B.run()
        torch.cuda.synchronize()
        for (_, u, _), ref in zip(entries, refs):
            assert torch.equal(u.data, ref), rep


def test_fused_batch_all_64_problems_vs_oracle(torch_cuda, oracle, monkeypatch):
    """The whole uneven batch (64 problems) as one fused graph: FP64 against
    the oracle to 1e-12 for every problem, before and after the autotuner
    (sx_batch_tune), equal to the graph of per-problem launches
    (SX_BATCH_FUSED=0) to round-off; the mixed fused graph (one FP64
    refinement) within 1e-10 of the oracle."""
    import re

    import torch
    from semtex_amd.problems import uneven_batch
    from semtex_amd.solver import SX_FP64, SX_MIXED, Batch
    probs = uneven_batch()
    print("case: fused batch of", len(probs), flush=True)
    refs = []
    for p in probs:
        m = p["mesh"]
        n = m.nel * m.npnp
        refs.append(oracle.solve_modes(m, p["lambda2"], 0.0, p["forcing"][:, :n])[0])

    def check(built, tol, what):
        worst = 0.0
        for p, (_, u, _, _), uo in zip(probs, built, refs):
            n = p["mesh"].nel * p["mesh"].npnp
            got = u.data[0, :n].cpu().numpy()
            worst = max(worst, float(np.linalg.norm(got - uo) / np.linalg.norm(uo)))
        print(f"  {what}: worst rel L2 vs oracle {worst:.2e}", flush=True)
        assert worst < tol, (what, worst)

    built = _build(torch, probs, SX_FP64)
    B = Batch([(M, u, src) for (M, u, src, _) in built])
    info = B.info()
    print("  ", info, flush=True)
    assert info.startswith("fused=1")
    assert int(re.search(r"launches=(\d+)", info).group(1)) < 200
    for (_, u, _, _) in built:
        u.data.zero_()
    B.run()
    torch.cuda.synchronize()
    check(built, 1e-12, "fused FP64")
    fused = [u.data.clone() for (_, u, _, _) in built]
    print("   tuned:", B.tune(reps=3), flush=True)
    for (_, u, _, _) in built:
        u.data.zero_()
    B.run()
    torch.cuda.synchronize()
    check(built, 1e-12, "tuned fused FP64")
    # the graph of per-problem launches
    monkeypatch.setenv("SX_BATCH_FUSED", "0")
    B0 = Batch([(M, u, src) for (M, u, src, _) in built])
    monkeypatch.delenv("SX_BATCH_FUSED")
    assert B0.info().startswith("fused=0")
    B0.run()
    torch.cuda.synchronize()
    for (_, u, _, _), f in zip(built, fused):
        assert torch.allclose(u.data, f, rtol=0, atol=1e-13 * float(f.abs().max())), "fused vs lanes"
    del B, B0
    builtm = _build(torch, probs, SX_MIXED)
    Bm = Batch([(M, u, src) for (M, u, src, _) in builtm], refine_passes=2)
    assert Bm.info().startswith("fused=1")
    Bm.run()
    torch.cuda.synchronize()
    check(builtm, 1e-10, "fused mixed, one refinement")
