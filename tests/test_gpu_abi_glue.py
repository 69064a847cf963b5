"""The C ABI driven from C++ exactly as the drop-in glue drives it
(tests/abi_glue.cpp): host arrays into sx_mesh_create through a mesh cache
keyed by the numbering system, ModalMatrixSys + Field::solve through
sx_field_solve_host, and integrate() through sx_dns_create with the
velocity and the pressure numbering (dns/drive.cpp:54-55, src/field.h:27,
src/matrix.h:61-66).  The C++ program's results are compared with the
oracle: the Helmholtz solve on the velocity numbering at the FP64 bar, and
the walled channel's fields after 4 DNS steps at the DNS bar.
"""
import os
import subprocess

import numpy as np
import pytest

from glue_build import build_glue

pytestmark = pytest.mark.gpu


def _write_case(d, mu, mp, nz, stride, params, F, lines, u0):
    def w(name, a, dt):
        np.ascontiguousarray(a, dt).tofile(os.path.join(d, name))
    w("x", mu.x, np.float64)
    w("y", mu.y, np.float64)
    for s, m in (("u", mu), ("p", mp)):
        w(f"{s}_btog", m.btog, np.int32)
        w(f"{s}_bmask", m.bmask, np.int32)
        w(f"{s}_emask", m.emask, np.int32)
        w(f"{s}_bnd", np.stack([m.bnd_elmt, m.bnd_side, m.bnd_kind], 1), np.int32)
        params[f"{s}_nglobal"], params[f"{s}_nsolve"] = m.nglobal, m.nsolve
        params[f"{s}_nband"], params[f"{s}_nbound"] = m.nband, m.nbound
    w("helm_f", F, np.float64)
    w("helm_lines", lines, np.float64)
    w("dns_u0", u0, np.float64)
    params.update(np=mu.np, nel=mu.nel, nz=nz, stride=stride)
    with open(os.path.join(d, "params.txt"), "w") as f:
        for k, v in params.items():
            f.write(f"{k} {v!r}\n")


def test_cpp_glue_solve_and_dns_match_oracle(oracle, tmp_path):
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.mesh import edgemaps, rect_mesh
    from semtex_amd.problems import fourier_forcing
    print("case: C++ glue, channel 4x6 np=6 nz=8: Helmholtz host solve + 4 DNS steps",
          flush=True)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(4, 6, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(4, 6, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    nz, n = 8, mu.nel * mu.npnp
    stride = mu.plane_size
    beta, nu, dt, lam2 = 1.0, 1.0 / 100, 2e-3, 3.0
    F = fourier_forcing(mu, nz, seed=11)
    lines = np.zeros((nz, mu.nbound * mu.np))
    for k in (0, 2, 3, 4):
        for b in range(mu.nbound):
            xs, _ = mu.side_xy(b)
            lines[k, b * mu.np:(b + 1) * mu.np] = np.cos((k + 1) * xs)
    rng = np.random.default_rng(5)
    wall = 1.0 - mu.y.ravel() ** 2
    em, _ = edgemaps(mu.np)
    u0 = np.zeros((3, nz, stride))
    for i in range(3):
        u0[i, :, :n] = (wall if i == 0 else 0.0) + 0.1 * rng.uniform(-1, 1, (nz, n)) * wall
        for k in range(nz):
            f = u0[i, k, :n].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, em[:mu.next]]
            f[:, em[:mu.next]] = g[mu.btog]
    steps = 4
    params = dict(lambda2=lam2, beta=beta, kinvis=nu, dt=dt, cf_x=2 * nu, dns_steps=steps)
    _write_case(str(tmp_path), mu, mp, nz, stride, params, F, lines, u0)
    exe = build_glue()
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    ug = np.fromfile(os.path.join(tmp_path, "helm_u"), np.float64).reshape(nz, stride)
    uo = oracle.solve_modes(mu, lam2, beta, np.ascontiguousarray(F[:, :n]), lines)
    for k in [0] + list(range(2, nz)):
        rel = np.linalg.norm(ug[k, :n] - uo[k]) / np.linalg.norm(uo[k])
        assert rel < 1e-12, (k, rel)

    out = np.fromfile(os.path.join(tmp_path, "dns_out"), np.float64).reshape(4, nz, stride)
    O = DnsOracle(mu, mp, nz, beta, nu, dt, const_force=(2 * nu, 0.0, 0.0), advection="skew")
    O.set_initial(u0[:, :, :n])
    for _ in range(steps):
        O.step()
    planes = [0] + list(range(2, nz))
    for i in range(4):
        o = O.u[i] if i < 3 else O.p
        rel = np.linalg.norm(out[i][planes, :n] - o[planes]) / np.linalg.norm(o[planes])
        assert rel < 1e-11, (i, rel)
