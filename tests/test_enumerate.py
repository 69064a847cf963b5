"""The numbering generator (sx_enumerate, `enumerate -O1`) on sessions of
the reference's own mesh directory (tests/golden/sessions: periodic,
walled, mixed, outflow and axis boundaries, 1 to 411 elements).

* The oracle's reverse Cuthill-McKee core (oracle/enumerate_oracle.py)
  equals the reference's own SPARSPAK FNROOT + RCM (femlib/netlib.f,
  compiled from its source into oracle/_ref) on every session's graph.
* sx_enumerate equals the oracle bit for bit: btog, bmask, nglobal,
  nsolve and the band width, for every field with distinct BCs.
* The numbering is a valid one: element-boundary nodes share a global
  number exactly when they sit at the same point (periodic images
  identified), essential nodes are numbered last, and the band width is
  the element spread of the unknowns.
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from semtex_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
SESSIONS = sorted(glob.glob(os.path.join(HERE, "golden", "sessions", "*")))


def _fields(S):
    f = set()
    for kinds in S.bcs.values():
        f.update(kinds)
    return sorted(f) or ["u"]


def product_enumerate(S, field, np_=None, optlev=1):
    np_ = np_ or S.np
    en, nnode, surf = S.topology(field)
    nel, next_ = len(en), 4 * (np_ - 1)
    E = np.ascontiguousarray(en, np.int32)
    Sf = np.ascontiguousarray(surf, np.int32).reshape(-1, 5)
    btog = np.zeros(nel * next_, np.int32)
    bmask = np.zeros(nel * next_, np.int32)
    ng, ns, nb = C.c_int(), C.c_int(), C.c_int()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = _lib.load().sx_enumerate(nel, nnode, p(E), len(surf), p(Sf), np_, optlev, p(btog),
                                  p(bmask), C.byref(ng), C.byref(ns), C.byref(nb))
    _lib.check(rc, "sx_enumerate")
    return btog, bmask, ng.value, ns.value, nb.value


@pytest.mark.parametrize("path", SESSIONS, ids=os.path.basename)
def test_oracle_rcm_equals_reference_sparspak(path):
    from oracle import enumerate_oracle as EO
    from semtex_amd.session import read_session
    if EO.netlib() is None:
        pytest.skip("reference netlib.f not built")
    S = read_session(path)
    for field in _fields(S):
        en, nnode, surf = S.topology(field)
        bmap, bmsk, nglobal, nsolve, _ = EO.enumerate_mesh(en, nnode, surf, S.np, optlev=0)
        if nsolve < 2:
            continue
        xadj, adj = EO.adjacency(bmap, bmsk, nsolve, len(en), 4 * (S.np - 1))
        ref_perm, ref_cc = EO.ref_fnroot_rcm(xadj, adj, nsolve)
        mask = [1] * (nsolve + 2)
        xls, perm = [0] * (nsolve + 2), [0] * (nsolve + 2)
        root, _ = EO.py_fnroot(1, xadj, adj, mask, xls, perm)
        mask = [1] * (nsolve + 2)
        cc = EO.py_rcm(root, list(xadj), adj, mask, perm, xls)
        assert cc == ref_cc == nsolve
        assert perm[1:nsolve + 1] == ref_perm[1:nsolve + 1], field


@pytest.mark.parametrize("path", SESSIONS, ids=os.path.basename)
def test_enumerate_equals_oracle(path):
    from oracle import enumerate_oracle as EO
    from semtex_amd.session import read_session
    S = read_session(path)
    for field in _fields(S):
        for np_ in (S.np, 3):
            got = product_enumerate(S, field, np_)
            want = EO.enumerate_mesh(*S.topology(field), np_)
            assert np.array_equal(got[0], want[0]), (field, np_, "btog")
            assert np.array_equal(got[1], want[1]), (field, np_, "bmask")
            assert got[2:] == want[2:], (field, np_, got[2:], want[2:])


def _node_points(S, np_):
    """Physical position of every element-boundary node (emap order) of the
    straight-sided elements, from the corner nodes (bilinear map)."""
    from semtex_amd.mesh import edgemaps, gll
    z, _, _ = gll(np_)
    em, _ = edgemaps(np_)
    pts = []
    for en in S.elements:
        c = np.array([S.nodes[k] for k in en])
        out = np.zeros((np_ * np_, 2))
        for i in range(np_):
            for j in range(np_):
                r, s = z[j], z[i]
                w = [(1 - r) * (1 - s), (1 + r) * (1 - s), (1 + r) * (1 + s), (1 - r) * (1 + s)]
                out[i * np_ + j] = 0.25 * (np.array(w)[:, None] * c).sum(0)
        pts.append(out[em[:4 * (np_ - 1)]])
    return np.concatenate(pts)


@pytest.mark.parametrize("name", ["kovas5", "laplace5", "cav01", "backfs1", "bfs01"])
def test_enumerate_numbering_is_valid(name):
    from semtex_amd.session import read_session
    S = read_session(os.path.join(HERE, "golden", "sessions", name))
    field = _fields(S)[0]
    btog, bmask, nglobal, nsolve, nband = product_enumerate(S, field)
    P = _node_points(S, S.np)
    key = np.round(P / 1e-9).astype(np.int64)
    # same global number <=> same point (no periodic surfaces in these)
    for g in range(nglobal):
        sel = key[btog == g]
        assert len(sel) and (sel == sel[0]).all(), g
    uniq = {tuple(k) for k in key}
    assert len(uniq) == nglobal
    assert set(np.unique(btog)) == set(range(nglobal))
    # essential nodes last
    assert (btog[bmask == 1] >= nsolve).all() and (btog[bmask == 0] < nsolve).all()
    # band width: the element spread of the unknowns, plus the diagonal
    nx = 4 * (S.np - 1)
    spread = max(int(b[m == 0].max() - b[m == 0].min()) for b, m in
                 zip(btog.reshape(-1, nx), bmask.reshape(-1, nx)) if (m == 0).any())
    assert nband == spread + 1
