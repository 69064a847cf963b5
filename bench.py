#!/usr/bin/env python3
"""Benchmark: batched per-Fourier-mode Helmholtz solve on one MI355X.

Workload (default): 16x16 quadrilateral elements on [0,2pi]x[-1,1], N_P=10,
N_Z=64 planes = 32 Fourier modes (beta=1), lambda2=1 so mode k solves with
1+k^2; u=0 on y=+-1, periodic in x; seeded Gaussian forcing with amplitude
(1+k)^-2 per mode.  One step = one Field::solve of all 63 planes (32
mode-solves: mode 0 real, modes 1..31 Re+Im; the Nyquist plane is not
solved), inputs resident in HBM.  FP64 throughout.

Prints ONE JSON line with the metric, a roofline entry for the dominant
kernel (the banded FP64 triangular sweeps, HBM-bound) measured with HIP
events on the solve stream, and the CPU oracle timed on a bounded sample of
the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mode-solves/s at ≤1e-10 rel err (mixed vs FP64 GPU); DNS steps/s"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="bounded CPU-baseline sample length")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-parity", action="store_true")
    return p.parse_args()


def cpu_baseline(P, seconds):
    """The CPU oracle (a plain-C FP64 port of the reference's solve) on one
    core: Field::solve of modes 0..7 of the same workload, repeated until
    `seconds` of solve time; set-up (factorisation) excluded."""
    from oracle import oracle as O
    m = P["mesh"]
    n = m.nel * m.npnp
    om = O.OracleMesh(m)
    modes = list(range(8))
    systems = {md: O.OracleMsys(om, P["lambda2"], (P["beta"] * md) ** 2) for md in modes}
    planes = [(0, 0)] + [(md, 2 * md + r) for md in modes[1:] for r in (0, 1)]
    F = P["forcing"]
    t0 = time.perf_counter()
    count = 0
    reps = 0
    while True:
        for md, k in planes:
            systems[md].solve_plane(F[k, :n])
        count += len(modes)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": count / el, "unit": "mode-solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle Field::solve, modes 0-7 (15 planes) x {reps} repetitions, "
                      f"{el:.1f} s, same mesh/numbering/forcing, 1 thread"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from semtex_amd import _lib
    from semtex_amd.problems import helmholtz_fourier
    from semtex_amd.solver import SX_FP64, Field, GpuMesh, ModalMatrixSys

    assert torch.cuda.is_available(), "bench needs a GPU"
    L = _lib.load()
    P = helmholtz_fourier()
    m, nz, nmodes = P["mesh"], P["nz"], P["nmodes"]
    print(f"# mesh {m.info['nx']}x{m.info['ny']} np={m.np} nz={nz} nglobal={m.nglobal} "
          f"nsolve={m.nsolve} nband={m.nband} ({m.info['ordering']})", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    gm = GpuMesh(m)
    M = ModalMatrixSys(P["lambda2"], P["beta"], 0, nmodes, gm, SX_FP64)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"# setup (geometry, condensation, band Cholesky of {nmodes} modes): "
          f"{setup_s:.3f} s", file=sys.stderr, flush=True)

    F_src = torch.from_numpy(P["forcing"]).cuda()
    u = Field(gm, nz)
    nbuf = a.warmup + a.steps
    fs = [F_src.clone() for _ in range(nbuf)]
    fs_t = [F_src.clone() for _ in range(a.steps)]
    torch.cuda.synchronize()

    for i in range(a.warmup):
        u.solve(fs[i], M)
    torch.cuda.synchronize()

    rel_err = None
    if not a.no_parity:
        from oracle import oracle as O
        n = m.nel * m.npnp
        ug = u.data.cpu().numpy()
        om = O.OracleMesh(m)
        errs = []
        for md in (0, 1, nmodes - 1):
            S = O.OracleMsys(om, P["lambda2"], (P["beta"] * md) ** 2)
            for k in ([0] if md == 0 else [2 * md, 2 * md + 1]):
                uo = S.solve_plane(P["forcing"][k, :n])
                errs.append(np.linalg.norm(ug[k, :n] - uo) / np.linalg.norm(uo))
        rel_err = float(max(errs))
        print(f"# parity vs oracle (modes 0,1,{nmodes-1}): max rel L2 {rel_err:.3e}",
              file=sys.stderr, flush=True)

    # timed region: K solves, inputs resident in HBM
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record()
    for i in range(a.steps):
        u.solve(fs[a.warmup + i], M)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1)
    ms_per_step = ms / a.steps
    value = nmodes * a.steps / (ms / 1e3)

    # per-kernel timing (HIP events on the solve stream)
    L.sx_timing_enable(1)
    for i in range(a.steps):
        u.solve(fs_t[i], M)
    torch.cuda.synchronize()
    pm = (C.c_double * _lib.SX_NPHASE)()
    pc = (C.c_int * _lib.SX_NPHASE)()
    _lib.check(L.sx_timing_read(pm, pc, _lib.SX_NPHASE), "sx_timing_read")
    L.sx_timing_enable(0)
    phases = {name: pm[i] / max(pc[i], 1) for i, name in enumerate(_lib.PHASES)}

    # algorithmic bytes of one band-solve launch: every mode's factor once
    # (8 nsolve nband) + its right-hand sides read and solutions written
    band_bytes = 0
    for md in range(nmodes):
        ns = M.nsolve(md)
        nrhs = 1 if md == 0 else 2
        band_bytes += 8 * ns * m.nband + 16 * ns * nrhs
    band_ms = phases["band_solve"]
    achieved = band_bytes / (band_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "mode-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched 2.5D Helmholtz, 16x16 elements on [0,2pi]x[-1,1], N_P=10, "
                        "N_Z=64 (32 modes, beta=1, lambda2=1+k^2), Dirichlet y=+-1, periodic x, "
                        "seeded Gaussian forcing ~(1+k)^-2; FP64 static condensation + "
                        "banded Cholesky",
            "precision": "fp64",
            "mode_solves_per_step": nmodes,
            "nglobal": m.nglobal, "nsolve": m.nsolve, "nband": m.nband,
            "ordering": m.info["ordering"],
        },
        "roofline": {
            "kernel": "k_band_solve",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes": band_bytes,
            "avg_ms": round(band_ms, 5),
        },
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "setup_s": round(setup_s, 3),
        "rel_err_vs_oracle": rel_err,
        "wall_s_timed": round(wall, 4),
    }
    if not a.no_cpu_baseline:
        cb = cpu_baseline(P, a.cpu_seconds)
        cb["value"] = round(cb["value"], 2)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
