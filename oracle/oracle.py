"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
compiled quadrature/edge-map routines (_ref/libsemtex_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_gll.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_edgemaps.argtypes = [C.c_int, ip, ip]
        L.or_mapping.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_mapping.restype = C.c_int
        L.or_helmholtz_row.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double,
                                       C.c_double, C.c_int, C.c_int, dp]
        L.or_helmholtz_sc.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double,
                                      C.c_double, dp, dp, dp]
        L.or_helmholtz_sc.restype = C.c_int
        L.or_helmholtz_op.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double,
                                      C.c_double, dp, dp]
        L.or_grad.argtypes = [C.c_int, dp, C.c_void_p, C.c_void_p]
        L.or_pbtrf.argtypes = [C.c_int, C.c_int, dp, C.c_int]
        L.or_pbtrf.restype = C.c_int
        L.or_pbtrs.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp]
        L.or_msys_build.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        L.or_pcg_diag.argtypes = [C.c_void_p, C.c_void_p, dp]
        L.or_solve_plane_pcg.argtypes = [C.c_void_p, C.c_void_p, dp, dp, dp, C.c_double, C.c_int]
        L.or_solve_plane_pcg.restype = C.c_int
        L.or_msys_build.restype = C.c_int
        L.or_msys_free.argtypes = [C.c_void_p]
        L.or_side_geom.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp, dp, dp]
        L.or_solve_plane.argtypes = [C.c_void_p, C.c_void_p, dp, dp, dp]
        L.or_solve_plane.restype = C.c_int
        L.or_dft.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own polylib/mapping, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libsemtex_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.zwglj.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double]
        R.Dglj.argtypes = [dp, dp, dp, C.c_int, C.c_double, C.c_double]
        R.edgemaps.argtypes = [C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_int)),
                               C.POINTER(C.POINTER(C.c_int))]
        _REF = R
    return _REF


NGEO = 10
Q1, Q2, Q3, Q4, Q8, DRDX, DRDY, DSDX, DSDY, DELTA = range(NGEO)


def gll(np_):
    z, w = np.zeros(np_), np.zeros(np_)
    D, DT = np.zeros(np_ * np_), np.zeros(np_ * np_)
    lib().or_gll(np_, z, w, D, DT)
    return z, w, D.reshape(np_, np_), DT.reshape(np_, np_)


def edgemaps(np_):
    em, pm = np.zeros(np_ * np_, np.int32), np.zeros(np_ * np_, np.int32)
    lib().or_edgemaps(np_, em, pm)
    return em, pm


def mapping(np_, x, y, cyl=0):
    geo = np.zeros(NGEO * np_ * np_)
    rc = lib().or_mapping(np_, cyl, np.ascontiguousarray(x, np.float64),
                          np.ascontiguousarray(y, np.float64), geo)
    return rc, geo.reshape(NGEO, np_ * np_)


def helmholtz_full(np_, geo, y, lambda2, betak2, cyl=0):
    """Full elemental matrix (np^2 x np^2), rows from HelmholtzRow."""
    n = np_ * np_
    H = np.zeros((n, n))
    g = np.ascontiguousarray(geo.ravel())
    yy = np.ascontiguousarray(y, np.float64)
    row = np.zeros(n)
    for i in range(np_):
        for j in range(np_):
            lib().or_helmholtz_row(np_, cyl, g, yy, lambda2, betak2, i, j, row)
            H[i * np_ + j] = row
    return H


def helmholtz_sc(np_, geo, y, lambda2, betak2, cyl=0):
    next_, nint = 4 * (np_ - 1), (np_ - 2) ** 2
    hbb = np.zeros(next_ * next_)
    hbi = np.zeros(max(next_ * nint, 1))
    hii = np.zeros(max(nint * nint, 1))
    rc = lib().or_helmholtz_sc(np_, cyl, np.ascontiguousarray(geo.ravel()),
                               np.ascontiguousarray(y, np.float64), lambda2,
                               betak2, hbb, hbi, hii)
    return rc, hbb.reshape(next_, next_), hbi[: next_ * nint].reshape(next_, nint), \
        hii[: nint * nint].reshape(nint, nint)


def helmholtz_op(np_, geo, y, lambda2, betak2, src, cyl=0):
    out = np.zeros(np_ * np_)
    lib().or_helmholtz_op(np_, cyl, np.ascontiguousarray(geo.ravel()),
                          np.ascontiguousarray(y, np.float64), lambda2, betak2,
                          np.ascontiguousarray(src, np.float64), out)
    return out


def grad(np_, geo, u):
    gx = np.ascontiguousarray(u, np.float64).copy()
    gy = gx.copy()
    lib().or_grad(np_, np.ascontiguousarray(geo.ravel()),
                  gx.ctypes.data_as(C.c_void_p), gy.ctypes.data_as(C.c_void_p))
    return gx, gy


def pbtrf(ab, n, kd):
    ab = np.ascontiguousarray(ab, np.float64).copy()
    rc = lib().or_pbtrf(n, kd, ab.ravel(), kd + 1)
    return rc, ab


def pbtrs(ab, n, kd, b):
    x = np.ascontiguousarray(b, np.float64).copy()
    lib().or_pbtrs(n, kd, np.ascontiguousarray(ab.ravel()), kd + 1, x)
    return x


def dft(data, nz, ntrn, sign):
    d = np.ascontiguousarray(data, np.float64).copy()
    lib().or_dft(d.ravel(), nz, ntrn, sign)
    return d


# --------------------------------------------------------------------------
class _OrMesh(C.Structure):
    _fields_ = [("np", C.c_int), ("nel", C.c_int), ("cyl", C.c_int),
                ("x", C.c_void_p), ("y", C.c_void_p), ("btog", C.c_void_p),
                ("bmask", C.c_void_p), ("emask", C.c_void_p),
                ("nglobal", C.c_int), ("nsolve", C.c_int), ("nband", C.c_int),
                ("nbound", C.c_int), ("bnd_elmt", C.c_void_p),
                ("bnd_side", C.c_void_p), ("bnd_kind", C.c_void_p)]


class _OrMsys(C.Structure):
    _fields_ = [("lambda2", C.c_double), ("betak2", C.c_double),
                ("nsolve", C.c_int), ("nband", C.c_int), ("singular", C.c_int),
                ("H", C.c_void_p), ("hbi", C.c_void_p), ("hii", C.c_void_p),
                ("geo", C.c_void_p), ("hbb_cache", C.c_void_p)]


class OracleMesh:
    """Keeps numpy arrays alive behind the C struct."""

    def __init__(self, m):
        self.m = m
        self._keep = [np.ascontiguousarray(a) for a in (
            m.x.astype(np.float64), m.y.astype(np.float64),
            m.btog.astype(np.int32), m.bmask.astype(np.int32),
            m.emask.astype(np.int32), m.bnd_elmt.astype(np.int32),
            m.bnd_side.astype(np.int32), m.bnd_kind.astype(np.int32))]
        x, y, bt, bm, em, be, bs, bk = self._keep
        p = lambda a: a.ctypes.data if a.size else None  # noqa: E731
        self.s = _OrMesh(m.np, m.nel, m.cyl, p(x), p(y), p(bt), p(bm), p(em),
                         m.nglobal, m.nsolve, m.nband, m.nbound, p(be), p(bs), p(bk))


class OracleMsys:
    """One mode's DIRECT MatrixSys (src/matrix.cpp:132-267)."""

    def __init__(self, om: OracleMesh, lambda2, betak2):
        self.om = om
        self.s = _OrMsys()
        rc = lib().or_msys_build(C.byref(om.s), lambda2, betak2, C.byref(self.s))
        if rc:
            raise RuntimeError(f"or_msys_build failed: {rc}")

    @property
    def nsolve(self):
        return self.s.nsolve

    def band(self):
        n = self.s.nband * max(self.s.nsolve, 1)
        return np.ctypeslib.as_array(C.cast(self.s.H, C.POINTER(C.c_double)), (n,)).copy()

    def solve_plane(self, f, line=None):
        m = self.om.m
        f = np.ascontiguousarray(f, np.float64).reshape(-1).copy()
        line = np.zeros(max(m.nbound * m.np, 1)) if line is None else \
            np.ascontiguousarray(line, np.float64).reshape(-1)
        u = np.zeros(m.nel * m.np * m.np)
        lib().or_solve_plane(C.byref(self.om.s), C.byref(self.s), f, line, u)
        return u

    def solve_plane_pcg(self, f, line=None, u0=None, tol_rel=1e-8, step_max=500):
        """JACPCG solve of one plane (src/field.cpp:711-805) from the initial
        guess u0 (zero if None); returns (u, iterations)."""
        m = self.om.m
        f = np.ascontiguousarray(f, np.float64).reshape(-1).copy()
        line = np.zeros(max(m.nbound * m.np, 1)) if line is None else \
            np.ascontiguousarray(line, np.float64).reshape(-1)
        u = np.zeros(m.nel * m.np * m.np) if u0 is None else \
            np.ascontiguousarray(u0, np.float64).reshape(-1).copy()
        it = lib().or_solve_plane_pcg(C.byref(self.om.s), C.byref(self.s), f, line, u,
                                      float(tol_rel), int(step_max))
        return u, it

    def pcg_diag(self):
        m = self.om.m
        pc = np.zeros(m.nglobal + m.nel * (m.np - 2) ** 2)
        lib().or_pcg_diag(C.byref(self.om.s), C.byref(self.s), pc)
        return pc

    def __del__(self):
        try:
            lib().or_msys_free(C.byref(self.s))
        except Exception:
            pass


def solve_modes(mesh, lambda2, beta, planes, lines=None):
    """Field::solve over Fourier planes [nz][nel*np*np] (src/field.cpp:648):
    plane k belongs to mode k>>1; the Nyquist plane 1 is skipped (left 0)."""
    om = OracleMesh(mesh)
    nz = planes.shape[0]
    out = np.zeros_like(planes)
    systems = {}
    for k in range(nz):
        if nz > 1 and k == 1:
            continue
        mode = k >> 1
        if mode not in systems:
            systems[mode] = OracleMsys(om, lambda2, (beta * mode) ** 2)
        line = None if lines is None else lines[k]
        out[k] = systems[mode].solve_plane(planes[k], line)
    return out


# --------------------------------------------------------------------------
# Cylindrical 3D: mode constants, coupling, per-mode-class numbering
def mode_constant(name, cyl, mode, beta):
    """Field::modeConstant (src/field.cpp:1204-1229): beta*mode, except for
    the coupled radial / azimuthal velocities of a cylindrical problem:
    v~ beta*mode + 1, w~ beta*mode - 1, and 1 for both at mode 0."""
    if not cyl or name not in ("v", "w"):
        return beta * mode
    if mode == 0:
        return 1.0
    return beta * mode + 1.0 if name == "v" else beta * mode - 1.0


def couple(V, W, direction, base_mode=0):
    """AuxField::couple (src/auxfield.cpp:1376-1451) on [nz][n] Fourier
    planes, returning new arrays: FORWARD (+1) v~ = v + i w, w~ = v - i w per
    mode (Re planes 2k, Im 2k+1); INVERSE (-1) v = (v~ + w~)/2,
    w = i (w~ - v~)/2.  Mode 0 of base mode 0 and nz = 1 stay unchanged."""
    V, W = np.array(V, np.float64), np.array(W, np.float64)
    nz = V.shape[0]
    if nz < 2:
        return V, W
    for k in range(1 if base_mode == 0 else 0, nz // 2):
        vr, vi, wr, wi = V[2 * k].copy(), V[2 * k + 1].copy(), W[2 * k].copy(), W[2 * k + 1].copy()
        if direction > 0:                       # src/auxfield.cpp:1411-1424
            V[2 * k], V[2 * k + 1] = vr - wi, vi + wr
            W[2 * k], W[2 * k + 1] = vr + wi, vi - wr
        else:                                   # src/auxfield.cpp:1428-1445
            V[2 * k], V[2 * k + 1] = 0.5 * (vr + wr), 0.5 * (vi + wi)
            W[2 * k], W[2 * k + 1] = 0.5 * (vi - wi), 0.5 * (wr - vr)
    return V, W


def solve_field_cyl(meshes, name, lambda2, beta, planes, lines=None, base_mode=0):
    """Field::solve over a BoundarySys whose numbering depends on the mode
    (src/field.cpp:648-700 with _bsys->Nsys(mode), src/bsys.cpp:42-123):
    plane k of mode base_mode + k>>1 is solved on meshes[min(mode, 2)] with
    betak2 = mode_constant(name, cyl, mode, beta)^2; the Nyquist plane 1 of
    base mode 0 is skipped (left 0)."""
    nz = planes.shape[0]
    out = np.zeros_like(planes)
    oms, systems = {}, {}
    for k in range(nz):
        if base_mode == 0 and nz > 1 and k == 1:
            continue
        mode = base_mode + (k >> 1)
        cls = min(mode, 2)
        m = meshes[cls]
        if cls not in oms:
            oms[cls] = OracleMesh(m)
        if mode not in systems:
            bk = mode_constant(name, m.cyl, mode, beta)
            systems[mode] = OracleMsys(oms[cls], lambda2, bk * bk)
        line = None if lines is None else lines[k]
        out[k] = systems[mode].solve_plane(planes[k], line)
    return out
