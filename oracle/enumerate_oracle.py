"""Restatement of semtex's `enumerate` numbering (utility/enumerate.cpp) for
checking the library's sx_enumerate.

TEST INFRASTRUCTURE ONLY.  Pure Python, following the reference step by
step:
  Mesh::assemble (mates by node identity)      src/mesh.cpp:261-347
  Mesh::surfaces + chooseNode (periodic pairs) src/mesh.cpp:350-536
  Mesh::buildMap + Side::connect               src/mesh.cpp:884-1009
  Mesh::buildMask                              src/mesh.cpp:1226-1326
  Nsys::sortGid                                utility/enumerate.cpp:636-688
  Nsys::renumber, buildAdjncy, fillAdjncy      utility/enumerate.cpp:691-893
  Nsys::globalBandwidth / bandwidthSC          utility/enumerate.cpp:896-940
The reverse Cuthill-McKee core (FNROOT, ROOTLS, DEGREE, RCM of George &
Liu, femlib/netlib.f:1507-1890) is called from the reference's own
Fortran, compiled into oracle/_ref/libnetlib_ref.so by oracle/Makefile,
when that library is present; otherwise from the restatement below, which
tests pin against the compiled routines.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
UNSET = -1
_NETLIB = None


def netlib():
    """The reference's SPARSPAK routines (fnroot_, rcm_) or None."""
    global _NETLIB
    if _NETLIB is None:
        path = os.path.join(HERE, "_ref", "libnetlib_ref.so")
        if not os.path.exists(path):
            return None
        _NETLIB = C.CDLL(path)
    return _NETLIB


# --------------------------------------------------------------------- RCM
def _rootls(root, xadj, adj, mask, xls, ls):
    mask[root] = 0
    ls[1] = root
    nlvl, lvlend, ccsize = 0, 0, 1
    while True:
        lbegin = lvlend + 1
        lvlend = ccsize
        nlvl += 1
        xls[nlvl] = lbegin
        for i in range(lbegin, lvlend + 1):
            node = ls[i]
            for j in range(xadj[node], xadj[node + 1]):
                nbr = adj[j]
                if mask[nbr] == 0:
                    continue
                ccsize += 1
                ls[ccsize] = nbr
                mask[nbr] = 0
        if ccsize - lvlend <= 0:
            break
    xls[nlvl + 1] = lvlend + 1
    for i in range(1, ccsize + 1):
        mask[ls[i]] = 1
    return nlvl


def py_fnroot(root, xadj, adj, mask, xls, ls):
    nlvl = _rootls(root, xadj, adj, mask, xls, ls)
    ccsize = xls[nlvl + 1] - 1
    if nlvl == 1 or nlvl == ccsize:
        return root, nlvl
    while True:
        jstrt = xls[nlvl]
        mindeg = ccsize
        root = ls[jstrt]
        if ccsize != jstrt:
            for j in range(jstrt, ccsize + 1):
                node = ls[j]
                ndeg = sum(1 for k in range(xadj[node], xadj[node + 1]) if mask[adj[k]] > 0)
                if ndeg < mindeg:
                    root, mindeg = node, ndeg
        nunlvl = _rootls(root, xadj, adj, mask, xls, ls)
        if nunlvl <= nlvl:
            return root, nlvl
        nlvl = nunlvl
        if nlvl >= ccsize:
            return root, nlvl


def _degree(root, xadj, adj, mask, deg, ls):
    ls[1] = root
    xadj[root] = -xadj[root]
    lvlend, ccsize = 0, 1
    while True:
        lbegin = lvlend + 1
        lvlend = ccsize
        for i in range(lbegin, lvlend + 1):
            node = ls[i]
            jstrt, jstop = -xadj[node], abs(xadj[node + 1]) - 1
            ideg = 0
            for j in range(jstrt, jstop + 1):
                nbr = adj[j]
                if mask[nbr] == 0:
                    continue
                ideg += 1
                if xadj[nbr] < 0:
                    continue
                xadj[nbr] = -xadj[nbr]
                ccsize += 1
                ls[ccsize] = nbr
            deg[node] = ideg
        if ccsize - lvlend <= 0:
            break
    for i in range(1, ccsize + 1):
        xadj[ls[i]] = -xadj[ls[i]]
    return ccsize


def py_rcm(root, xadj, adj, mask, perm, deg):
    ccsize = _degree(root, xadj, adj, mask, deg, perm)
    mask[root] = 0
    if ccsize <= 1:
        return ccsize
    lvlend, lnbr = 0, 1
    while True:
        lbegin = lvlend + 1
        lvlend = lnbr
        for i in range(lbegin, lvlend + 1):
            node = perm[i]
            fnbr = lnbr + 1
            for j in range(xadj[node], xadj[node + 1]):
                nbr = adj[j]
                if mask[nbr] == 0:
                    continue
                lnbr += 1
                mask[nbr] = 0
                perm[lnbr] = nbr
            if fnbr >= lnbr:
                continue
            k = fnbr
            while True:
                l = k
                k += 1
                nbr = perm[k]
                while l >= fnbr and deg[perm[l]] > deg[nbr]:
                    perm[l + 1] = perm[l]
                    l -= 1
                perm[l + 1] = nbr
                if k >= lnbr:
                    break
        if lnbr <= lvlend:
            break
    l = ccsize
    for i in range(1, ccsize // 2 + 1):
        perm[l], perm[i] = perm[i], perm[l]
        l -= 1
    return ccsize


def ref_fnroot_rcm(xadj, adj, nsolve, root=1):
    """FNROOT then RCM by the reference's compiled Fortran (1-based arrays
    passed from index 1).  Returns (perm[1..n], ccsize)."""
    L = netlib()
    ia = lambda a: np.ascontiguousarray(a[1:], dtype=np.int32)  # noqa: E731
    X, A = ia(xadj), ia(adj)
    mask = np.ones(nsolve + 1, np.int32)
    xls = np.zeros(nsolve + 2, np.int32)
    perm = np.zeros(nsolve + 1, np.int32)
    r, nlvl = C.c_int(root), C.c_int(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    L.fnroot_(C.byref(r), p(X), p(A), p(mask), C.byref(nlvl), p(xls), p(perm))
    mask[:] = 1
    L.rcm_(C.byref(r), p(X), p(A), p(mask), p(perm), C.byref(nlvl), p(xls))
    return [0] + [int(v) for v in perm[:nsolve]], int(nlvl.value)


def adjacency(bmap, bmsk, nsolve, nel, next_):
    """Nsys::buildAdjncy + fillAdjncy (utility/enumerate.cpp:792-866): the
    unknowns' graph in 1-based xadj / adjncy form (index 0 unused), each
    node's neighbours in the order the elements first connect them."""
    adjl = [[] for _ in range(nsolve)]
    for e in range(nel):
        bm, mk = bmap[e * next_:(e + 1) * next_], bmsk[e * next_:(e + 1) * next_]
        for i in range(next_):
            if mk[i]:
                continue
            L = adjl[bm[i]]
            for j in range(next_):
                if i != j and not mk[j] and bm[j] not in L:
                    L.append(int(bm[j]))
    xadj, adj = [0] * (nsolve + 2), [0]
    k = 1
    for i in range(nsolve):
        xadj[i + 1] = k
        for v in adjl[i]:
            adj.append(v + 1)
            k += 1
    xadj[nsolve + 1] = k
    adj.append(0)
    return xadj, adj


# ---------------------------------------------------------------- topology
class _Node:
    __slots__ = ("id", "gid", "periodic")

    def __init__(self, i):
        self.id, self.gid, self.periodic = i, UNSET, None


class _Side:
    __slots__ = ("start", "end", "mate", "group", "essential", "gid")

    def __init__(self, a, b, ni):
        self.start, self.end, self.mate = a, b, None
        self.group, self.essential = False, False
        self.gid = [UNSET] * ni


def _choose(n1, n2):
    """Mesh::chooseNode (src/mesh.cpp:477-536)."""
    if n1.periodic is None:
        if n2.periodic is None:
            pn = n1 if n1.id < n2.id else n2
            n1.periodic = n2.periodic = pn
        elif n2.periodic is not n2:
            pn = n1 if n1.id < n2.periodic.id else n2.periodic
            n2.periodic.periodic = pn
            n2.periodic = pn
            n1.periodic = pn
        else:
            pn = n1 if n1.id < n2.id else n2
            n1.periodic = n2.periodic = pn
    elif n1.periodic is not n1:
        if n2.periodic is None:
            pn = n1.periodic if n1.periodic.id < n2.id else n2
            n1.periodic.periodic = pn
            n1.periodic = pn
            n2.periodic = pn
        elif n2.periodic is not n2:
            pn = n1.periodic if n1.periodic.id < n2.periodic.id else n2.periodic
            a, b = n1.periodic, n2.periodic
            b.periodic = pn
            a.periodic = pn
            n2.periodic = pn
            n1.periodic = pn
        else:
            pn = n1.periodic if n1.periodic.id < n2.id else n2
            n1.periodic.periodic = pn
            n1.periodic = pn
            n2.periodic = pn
    else:
        if n2.periodic is None:
            pn = n1 if n1.id < n2.id else n2
            n1.periodic = n2.periodic = pn
        elif n2.periodic is not n2:
            pn = n1 if n1.id < n2.periodic.id else n2
            n1.periodic.periodic = pn
            n1.periodic = pn
            n2.periodic = pn
        else:
            pn = n1 if n1.id < n2.id else n2
            n1.periodic = n2.periodic = pn


def enumerate_mesh(elem_nodes, nnode, surfaces, np_, optlev=1):
    """elem_nodes [nel][4] 0-based node indices; surfaces: list of
    (elmt, side, kind, a, b) 0-based, kind 0 = boundary with essential flag
    a, kind 1 = periodic with mate (a, b).  Returns (btog, bmask, nglobal,
    nsolve, nband) as enumerate -O<optlev> numbers one field."""
    nel = len(elem_nodes)
    ni, next_ = np_ - 2, 4 * (np_ - 1)
    nodes = [_Node(i) for i in range(nnode)]
    sides = [[_Side(nodes[elem_nodes[e][j]], nodes[elem_nodes[e][(j + 1) % 4]], ni)
              for j in range(4)] for e in range(nel)]
    for e in range(nel):                       # Mesh::assemble
        for j in range(4):
            s = sides[e][j]
            for r in range(nel):
                m = next((m for m in sides[r] if s.start is m.end and m.start is s.end), None)
                if m is not None:
                    s.mate = m
                    break
    for (e, j, kind, a, b) in surfaces:         # Mesh::surfaces
        s = sides[e][j]
        if kind == 0:
            s.group, s.essential = True, bool(a)
        else:
            m = sides[a][b]
            s.mate, m.mate = m, s
            _choose(s.start, m.end)
            _choose(s.end, m.start)
    # Mesh::buildMap
    counter = [0]

    def node_gid(n):
        if n.periodic is not None:
            if n.periodic.gid == UNSET:
                n.periodic.gid = counter[0]
                counter[0] += 1
            if n.gid == UNSET:
                n.gid = n.periodic.gid
        if n.gid == UNSET:
            n.gid = counter[0]
            counter[0] += 1

    for row in sides:
        for s in row:
            s.start.gid = s.end.gid = UNSET
            s.gid = [UNSET] * ni
    for row in sides:
        for s in row:
            node_gid(s.start)
            if ni:
                if s.mate is not None and s.mate.gid[0] != UNSET:
                    s.gid = list(reversed(s.mate.gid))
                else:
                    s.gid = list(range(counter[0], counter[0] + ni))
                    counter[0] += ni
            node_gid(s.end)
    bmap = [v for row in sides for s in row for v in [s.start.gid] + s.gid]
    # Mesh::buildMask
    for row in sides:
        for s in row:
            s.start.gid = s.end.gid = 0
            s.gid = [0] * ni
    for row in sides:
        for s in row:
            if s.mate is None and s.essential:
                s.start.gid = s.end.gid = 1
                s.gid = [1] * ni
                if s.start.periodic is not None:
                    s.start.periodic.gid = 1
                if s.end.periodic is not None:
                    s.end.periodic.gid = 1
    bmsk = [v for row in sides for s in row for v in [s.start.gid] + s.gid]
    bmap, bmsk = np.array(bmap), np.array(bmsk)
    # Nsys::sortGid
    nglobal = int(bmap.max()) + 1
    tmp = np.zeros(nglobal, int)
    for g, k in zip(bmap, bmsk):
        tmp[g] = k
    unknowns = int(nglobal - np.count_nonzero(tmp))
    if unknowns < nglobal:
        order = sorted(range(nglobal), key=lambda g: (tmp[g] != 0, g))
        renum = np.empty(nglobal, int)
        renum[order] = np.arange(nglobal)
        bmap = renum[bmap]
    nsolve = unknowns
    if optlev and nsolve:                      # Nsys::renumber, -O1
        assert optlev == 1, "the oracle restates enumerate -O1"
        xadj, adj = adjacency(bmap, bmsk, nsolve, nel, next_)
        if netlib() is not None:
            perm, cc = ref_fnroot_rcm(xadj, adj, nsolve)
        else:
            mask = [1] * (nsolve + 2)
            xls, perm = [0] * (nsolve + 2), [0] * (nsolve + 2)
            root, _ = py_fnroot(1, xadj, adj, mask, xls, perm)
            mask = [1] * (nsolve + 2)
            cc = py_rcm(root, list(xadj), adj, mask, perm, xls)
        assert cc == nsolve, "unknowns' graph not connected"
        invperm = np.arange(nglobal)
        for i in range(nsolve):
            invperm[perm[i + 1] - 1] = i
        bmap = invperm[bmap]
    nband = 0
    for e in range(nel):
        sel = bmap[e * next_:(e + 1) * next_][bmsk[e * next_:(e + 1) * next_] == 0]
        nband = max(nband, int(sel.max() - sel.min()) if sel.size else 1)
    return bmap.astype(np.int32), bmsk.astype(np.int32), nglobal, nsolve, nband + 1
