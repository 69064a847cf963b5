"""Writes tests/golden/dns_channel_12x24_np8_nz128.npz: per-step fixtures of
the DNS oracle on the full channel workload (12x24 elements on
[0,4pi]x[-1,1], N_P=8, N_Z=128, Lz=2pi, Re=2000, dt=1e-3, N_TIME=2,
skew-symmetric advection, CONST_X = 2/Re, Poiseuille + seeded U(-1e-2,1e-2)
perturbation; semtex_amd.problems.dns_channel), 100 steps.

For every step and field (u, v, w, p; Nyquist plane excluded):
  norm      L2 norm of the oracle's Fourier field,
  samples   the oracle's values at 256 seeded (plane, node) positions,
  vsamples  the variant's values there (see floor),
  floor     ||oracle - variant|| / norm, where the variant is the same
            oracle with the pressure systems of modes 0 and 1 solved by
            oracle/dense_alt.py (an independently ordered FP64 solve of the
            same discrete system): how far two valid FP64 implementations
            drift apart, step by step.  Mode 0 of the pressure is singular
            (pure Neumann, last unknown dropped) with a condensed matrix of
            condition ~2.6e5, which sets this floor.
The first three steps' full fields are not stored (75 MB a step); the GPU
test recomputes them with the oracle.

    python tests/golden/make_dns_channel.py        (about 10 minutes)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "dns_channel_12x24_np8_nz128.npz")
STEPS = 100
NSAMP = 256
SAMPLE_SEED = 4242


def sample_positions(nz, n, count=NSAMP, seed=SAMPLE_SEED):
    """(plane, node) pairs, Nyquist plane 1 excluded; plane 0 (mode 0, where
    the pressure floor lives) always gets a share."""
    rng = np.random.default_rng(seed)
    planes = np.concatenate([np.zeros(count // 8, int),
                             rng.choice([k for k in range(nz) if k != 1], count - count // 8)])
    nodes = rng.integers(0, n, count)
    return planes, nodes


def main():
    from oracle.dense_alt import DenseAltSystem
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.problems import dns_channel
    P = dns_channel()
    mu, mp, nz = P["mesh_u"], P["mesh_p"], P["nz"]
    n = mu.nel * mu.npnp
    kw = dict(const_force=P["const_force"], advection="skew")
    A = DnsOracle(mu, mp, nz, P["beta"], P["kinvis"], P["dt"], **kw)
    B = DnsOracle(mu, mp, nz, P["beta"], P["kinvis"], P["dt"], **kw)
    for md in (0, 1):
        ns = A.system("p", 0.0, md).nsolve
        B._sys[("p", 0.0, md)] = DenseAltSystem(mp, 0.0, (P["beta"] * md) ** 2, ns)
    A.set_initial(P["u0"])
    B.set_initial(P["u0"])
    pl = [k for k in range(nz) if k != 1]
    sp, sn = sample_positions(nz, n)
    norms = np.zeros((STEPS, 4))
    floors = np.zeros((STEPS, 4))
    samples = np.zeros((STEPS, 4, NSAMP))
    vsamples = np.zeros((STEPS, 4, NSAMP))
    t0 = time.time()
    for s in range(STEPS):
        A.step()
        B.step()
        for i in range(4):
            a = A.u[i] if i < 3 else A.p
            b = B.u[i] if i < 3 else B.p
            norms[s, i] = np.linalg.norm(a[pl])
            floors[s, i] = np.linalg.norm(a[pl] - b[pl]) / norms[s, i]
            samples[s, i] = a[sp, sn]
            vsamples[s, i] = b[sp, sn]
        print(f"step {s + 1}: floor u v w p " + " ".join(f"{x:.2e}" for x in floors[s]) +
              f"  ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(OUT, norms=norms, floors=floors, samples=samples, vsamples=vsamples,
                        sample_planes=sp, sample_nodes=sn, steps=STEPS)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
