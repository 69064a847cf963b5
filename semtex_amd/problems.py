"""Workload builders for the Helmholtz solve: meshes, forcing and boundary
data of the benchmark configurations (synthetic, seeded).

  helmholtz_2d_square   4x4 elements on the unit square, N_P=8, lambda2=1,
                        u = sin(pi x) sin(pi y), f = -(2 pi^2 + 1) u.
  helmholtz_fourier     16x16 elements on [0,2pi]x[-1,1], N_P=10, N_Z=64
                        (32 modes, beta=1), lambda2=1 so mode k solves with
                        1+k^2; u=0 on y=+-1, periodic in x; forcing per mode
                        i.i.d. N(0,1) x (1+k)^-2 (Re and Im; mode 0 real;
                        Nyquist plane zero).
"""
from __future__ import annotations

import numpy as np

from .mesh import rect_mesh

SEED = 20261202


def helmholtz_2d_square(np_=8, nel_side=4, ordering="auto"):
    m = rect_mesh(nel_side, nel_side, np_, ordering=ordering)
    exact = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    f = -(2 * np.pi ** 2 + 1.0) * exact
    return dict(mesh=m, lambda2=1.0, beta=0.0, nz=1, nmodes=1,
                forcing=f.reshape(1, -1), exact=exact.reshape(1, -1))


def fourier_forcing(mesh, nz, seed=SEED, amp_pow=2.0):
    """[nz][plane_size] Fourier planes, amplitude (1+k)^-amp_pow."""
    rng = np.random.default_rng(seed)
    ps = mesh.plane_size
    n = mesh.nel * mesh.npnp
    F = np.zeros((nz, ps))
    for k in range(nz):
        if k == 1 and nz > 1:
            continue                       # Nyquist plane stays zero
        mode = k >> 1
        F[k, :n] = rng.standard_normal(n) * (1.0 + mode) ** (-amp_pow)
    return F


def helmholtz_fourier(nx=16, ny=16, np_=10, nz=64, seed=SEED, ordering="auto"):
    m = rect_mesh(nx, ny, np_, 0.0, 2 * np.pi, -1.0, 1.0,
                  bc=("D", "N", "D", "N"), periodic_x=True, ordering=ordering)
    return dict(mesh=m, lambda2=1.0, beta=1.0, nz=nz, nmodes=nz // 2,
                forcing=fourier_forcing(m, nz, seed))
