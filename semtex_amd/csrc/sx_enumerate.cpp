// sx_enumerate.cpp -- the numbering generator semtex's `enumerate` utility
// provides (utility/enumerate.cpp), host code: from a session's mesh
// topology it builds the element-boundary to global node map of the
// statically condensed system, masks the essential-BC nodes of one field,
// numbers them last and renumbers the unknowns by reverse Cuthill-McKee to
// reduce the band width of the assembled matrix:
//   Mesh::assemble / surfaces / chooseNode   src/mesh.cpp:261-556
//   Mesh::buildMap / Side::connect           src/mesh.cpp:884-1009
//   Mesh::buildMask                          src/mesh.cpp:1226-1326
//   Nsys::sortGid / renumber / bandwidth     utility/enumerate.cpp:636-940
//   FNROOT / ROOTLS / DEGREE / RCM (George & Liu, "Computer Solution of
//   Large Sparse Positive Definite Systems", 1981; femlib/netlib.f)
// The traversal orders are the reference's, because the numbering (and so
// the band layout and rounding of every solve) follows from them.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <string>
#include <vector>

#include "sx_internal.h"

using namespace sx;

namespace {

constexpr int UNSET = -1;

struct Node {
  int id = 0;
  int gid = UNSET;
  Node* periodic = nullptr;
};

struct Side {
  Node *start = nullptr, *end = nullptr;
  int elmt = -1, mate_elmt = -1;
  Side* mate = nullptr;
  int group = 0;            // 1 on a boundary surface (group set)
  int essential = 0;        // that surface's group is essential for the field
  std::vector<int> gid;     // side-internal numbers (np - 2)
};

struct Topology {
  std::vector<Node> nodes;
  std::vector<Side> sides;  // [nel][4]
  int nel = 0;
  Side& side (int e, int j) { return sides[(size_t) e * 4 + j]; }
};

// Mesh::chooseNode: join two nodes into one periodic set, represented by
// the lowest node ID among the alternatives (the 9 cases of whether each is
// already periodic and whether that is self-referential).
void choose_node (Node* n1, Node* n2)
{
  Node* pn = nullptr;
  if (!n1->periodic) {
    if (!n2->periodic) {
      pn = (n1->id < n2->id) ? n1 : n2;
      n1->periodic = n2->periodic = pn;
    } else if (n2->periodic != n2) {
      pn = (n1->id < n2->periodic->id) ? n1 : n2->periodic;
      n1->periodic = n2->periodic = n2->periodic->periodic = pn;
    } else {
      pn = (n1->id < n2->id) ? n1 : n2;
      n1->periodic = n2->periodic = pn;
    }
  } else if (n1->periodic != n1) {
    if (!n2->periodic) {
      pn = (n1->periodic->id < n2->id) ? n1->periodic : n2;
      n2->periodic = n1->periodic = n1->periodic->periodic = pn;
    } else if (n2->periodic != n2) {
      pn = (n1->periodic->id < n2->periodic->id) ? n1->periodic : n2->periodic;
      Node* a = n1->periodic;
      Node* b = n2->periodic;
      n1->periodic = n2->periodic = a->periodic = b->periodic = pn;
    } else {
      pn = (n1->periodic->id < n2->id) ? n1->periodic : n2;
      n2->periodic = n1->periodic = n1->periodic->periodic = pn;
    }
  } else {
    if (!n2->periodic) {
      pn = (n1->id < n2->id) ? n1 : n2;
      n1->periodic = n2->periodic = pn;
    } else if (n2->periodic != n2) {
      pn = (n1->id < n2->periodic->id) ? n1 : n2;
      n2->periodic = n1->periodic = n1->periodic->periodic = pn;
    } else {
      pn = (n1->id < n2->id) ? n1 : n2;
      n1->periodic = n2->periodic = pn;
    }
  }
}

// Side::connect: global numbers of a side's start node, interior knots
// (copied reversed from an already numbered mate) and end node
void connect (Side& s, int ni, int& gid)
{
  auto node_gid = [&] (Node* n) {
    if (n->periodic) {
      if (n->periodic->gid == UNSET) n->periodic->gid = gid++;
      if (n->gid == UNSET) n->gid = n->periodic->gid;
    }
    if (n->gid == UNSET) n->gid = gid++;
  };
  node_gid (s.start);
  if (ni) {
    if (s.mate) {
      if (s.mate->gid[0] == UNSET)
        for (int i = 0; i < ni; i++) s.gid[i] = gid++;
      else
        for (int i = 0, k = ni - 1; i < ni; i++, k--) s.gid[i] = s.mate->gid[k];
    } else {
      for (int i = 0; i < ni; i++) s.gid[i] = gid++;
    }
  }
  node_gid (s.end);
}

// -------------------------------------------------------------------------
// George & Liu's level structures and reverse Cuthill-McKee ordering, on a
// graph in the 1-based xadj / adjncy form.  All arrays 1-based (index 0
// unused) to keep the published algorithm's indexing.
// -------------------------------------------------------------------------
void rootls (int root, const std::vector<int>& xadj, const std::vector<int>& adj,
             std::vector<int>& mask, int& nlvl, std::vector<int>& xls, std::vector<int>& ls)
{
  mask[root] = 0;
  ls[1] = root;
  nlvl = 0;
  int lvlend = 0, ccsize = 1;
  for (;;) {
    const int lbegin = lvlend + 1;
    lvlend = ccsize;
    nlvl++;
    xls[nlvl] = lbegin;
    for (int i = lbegin; i <= lvlend; i++) {
      const int node = ls[i];
      for (int j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ls[++ccsize] = nbr;
        mask[nbr] = 0;
      }
    }
    if (ccsize - lvlend <= 0) break;
  }
  xls[nlvl + 1] = lvlend + 1;
  for (int i = 1; i <= ccsize; i++) mask[ls[i]] = 1;
}

void fnroot (int& root, const std::vector<int>& xadj, const std::vector<int>& adj,
             std::vector<int>& mask, int& nlvl, std::vector<int>& xls, std::vector<int>& ls)
{
  rootls (root, xadj, adj, mask, nlvl, xls, ls);
  const int ccsize = xls[nlvl + 1] - 1;
  if (nlvl == 1 || nlvl == ccsize) return;
  for (;;) {
    const int jstrt = xls[nlvl];
    int mindeg = ccsize;
    root = ls[jstrt];
    if (ccsize != jstrt) {
      for (int j = jstrt; j <= ccsize; j++) {
        const int node = ls[j];
        int ndeg = 0;
        for (int k = xadj[node]; k <= xadj[node + 1] - 1; k++)
          if (mask[adj[k]] > 0) ndeg++;
        if (ndeg >= mindeg) continue;
        root = node;
        mindeg = ndeg;
      }
    }
    int nunlvl = 0;
    rootls (root, xadj, adj, mask, nunlvl, xls, ls);
    if (nunlvl <= nlvl) return;
    nlvl = nunlvl;
    if (nlvl >= ccsize) return;
  }
}

void degree (int root, std::vector<int>& xadj, const std::vector<int>& adj,
             const std::vector<int>& mask, std::vector<int>& deg, int& ccsize,
             std::vector<int>& ls)
{
  ls[1] = root;
  xadj[root] = -xadj[root];
  int lvlend = 0;
  ccsize = 1;
  for (;;) {
    const int lbegin = lvlend + 1;
    lvlend = ccsize;
    for (int i = lbegin; i <= lvlend; i++) {
      const int node = ls[i];
      const int jstrt = -xadj[node], jstop = std::abs (xadj[node + 1]) - 1;
      int ideg = 0;
      for (int j = jstrt; j <= jstop; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ideg++;
        if (xadj[nbr] < 0) continue;
        xadj[nbr] = -xadj[nbr];
        ls[++ccsize] = nbr;
      }
      deg[node] = ideg;
    }
    if (ccsize - lvlend <= 0) break;
  }
  for (int i = 1; i <= ccsize; i++) xadj[ls[i]] = -xadj[ls[i]];
}

void rcm (int root, std::vector<int>& xadj, const std::vector<int>& adj,
          std::vector<int>& mask, std::vector<int>& perm, int& ccsize, std::vector<int>& deg)
{
  degree (root, xadj, adj, mask, deg, ccsize, perm);
  mask[root] = 0;
  if (ccsize <= 1) return;
  int lvlend = 0, lnbr = 1;
  do {
    const int lbegin = lvlend + 1;
    lvlend = lnbr;
    for (int i = lbegin; i <= lvlend; i++) {
      const int node = perm[i];
      const int fnbr = lnbr + 1;
      for (int j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        const int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        lnbr++;
        mask[nbr] = 0;
        perm[lnbr] = nbr;
      }
      if (fnbr >= lnbr) continue;
      // the new neighbours in order of increasing degree (insertion sort)
      int k = fnbr;
      do {
        int l = k;
        k++;
        const int nbr = perm[k];
        while (l >= fnbr && deg[perm[l]] > deg[nbr]) {
          perm[l + 1] = perm[l];
          l--;
        }
        perm[l + 1] = nbr;
      } while (k < lnbr);
    }
  } while (lnbr > lvlend);
  for (int i = 1, l = ccsize; i <= ccsize / 2; i++, l--) std::swap (perm[l], perm[i]);
}

int band_of (const std::vector<int>& bmap, const std::vector<int>& bmsk, int nel, int next)
{
  int nband = 0;
  for (int e = 0; e < nel; e++) {
    int lo = INT_MAX, hi = INT_MIN;
    bool any = false;
    for (int i = 0; i < next; i++)
      if (!bmsk[(size_t) e * next + i]) {
        lo = std::min (lo, bmap[(size_t) e * next + i]);
        hi = std::max (hi, bmap[(size_t) e * next + i]);
        any = true;
      }
    // an element with no unknown node: the reference's INT_MIN - INT_MAX,
    // which wraps to 1 in two's complement
    nband = std::max (nband, any ? hi - lo : 1);
  }
  return nband + 1;
}

}  // namespace

extern "C" int sx_enumerate (int nel, int nnode, const int* elem_nodes, int nsurf,
                             const int* surf, int np, int optlev, int* btog, int* bmask,
                             int* nglobal_out, int* nsolve_out, int* nband_out)
{
  if (nel < 1 || nnode < 4 || !elem_nodes || nsurf < 0 || (nsurf && !surf) || np < 2 ||
      optlev < 0 || optlev > 3 || !btog || !bmask || !nglobal_out || !nsolve_out ||
      !nband_out) {
    set_error ("sx_enumerate: bad arguments");
    return SX_ERR_ARG;
  }
  Topology T;
  T.nel = nel;
  T.nodes.resize (nnode);
  for (int i = 0; i < nnode; i++) T.nodes[i].id = i;
  T.sides.resize ((size_t) nel * 4);
  const int ni = np - 2, next = 4 * (np - 1);
  for (int e = 0; e < nel; e++)
    for (int j = 0; j < 4; j++) {
      const int a = elem_nodes[e * 4 + j], b = elem_nodes[e * 4 + (j + 1) % 4];
      if (a < 0 || a >= nnode || b < 0 || b >= nnode) {
        set_error ("sx_enumerate: element node out of range");
        return SX_ERR_ARG;
      }
      Side& s = T.side (e, j);
      s.start = &T.nodes[a];
      s.end = &T.nodes[b];
      s.elmt = e;
      s.gid.assign (ni, UNSET);
    }
  // Mesh::assemble: the first side, in element order, running the other way
  for (int e = 0; e < nel; e++)
    for (int j = 0; j < 4; j++) {
      Side& s = T.side (e, j);
      bool found = false;
      for (int r = 0; !found && r < nel; r++)
        for (int q = 0; !found && q < 4; q++) {
          Side& m = T.side (r, q);
          if ((found = (s.start == m.end && m.start == s.end))) {
            s.mate = &m;
            s.mate_elmt = r;
          }
        }
    }
  // Mesh::surfaces, in the session's order
  for (int k = 0; k < nsurf; k++) {
    const int* t = surf + (size_t) k * 5;
    const int e = t[0], j = t[1];
    if (e < 0 || e >= nel || j < 0 || j > 3) {
      set_error ("sx_enumerate: surface element or side out of range");
      return SX_ERR_ARG;
    }
    Side& s = T.side (e, j);
    if (s.mate) {
      set_error ("sx_enumerate: surface " + std::to_string (k + 1) + " already mated");
      return SX_ERR_ARG;
    }
    if (t[2] == 0) {
      s.group = 1;
      s.essential = t[3] != 0;
    } else {
      const int me = t[3], ms = t[4];
      if (me < 0 || me >= nel || ms < 0 || ms > 3 || T.side (me, ms).mate ||
          T.side (me, ms).group) {
        set_error ("sx_enumerate: periodic mate of surface " + std::to_string (k + 1) +
                   " out of range or already set");
        return SX_ERR_ARG;
      }
      Side& m = T.side (me, ms);
      s.mate = &m;
      s.mate_elmt = me;
      m.mate = &s;
      m.mate_elmt = e;
      choose_node (s.start, m.end);
      choose_node (s.end, m.start);
    }
  }
  for (int e = 0; e < nel; e++)
    for (int j = 0; j < 4; j++)
      if (!T.side (e, j).mate && !T.side (e, j).group) {
        set_error ("sx_enumerate: element " + std::to_string (e + 1) + " side " +
                   std::to_string (j + 1) + " has neither a mate nor a surface");
        return SX_ERR_ARG;
      }
  // Mesh::buildMap
  const int nb = nel * next;
  std::vector<int> bmap (nb), bmsk (nb);
  int ngid = 0;
  for (auto& s : T.sides) { s.start->gid = s.end->gid = UNSET; std::fill (s.gid.begin (), s.gid.end (), UNSET); }
  for (auto& s : T.sides) connect (s, ni, ngid);
  for (int e = 0, p = 0; e < nel; e++)
    for (int j = 0; j < 4; j++) {
      Side& s = T.side (e, j);
      bmap[p++] = s.start->gid;
      for (int k = 0; k < ni; k++) bmap[p++] = s.gid[k];
    }
  // Mesh::buildMask (node and side flags reuse the gid slots as there)
  for (auto& s : T.sides) { s.start->gid = s.end->gid = 0; std::fill (s.gid.begin (), s.gid.end (), 0); }
  for (auto& s : T.sides)
    if (!s.mate && s.essential) {
      s.start->gid = s.end->gid = 1;
      std::fill (s.gid.begin (), s.gid.end (), 1);
      if (s.start->periodic) s.start->periodic->gid = 1;
      if (s.end->periodic) s.end->periodic->gid = 1;
    }
  for (int e = 0, p = 0; e < nel; e++)
    for (int j = 0; j < 4; j++) {
      Side& s = T.side (e, j);
      bmsk[p++] = s.start->gid;
      for (int k = 0; k < ni; k++) bmsk[p++] = s.gid[k];
    }
  // Nsys::sortGid: essential nodes last, each partition in ascending order
  const int nglobal = *std::max_element (bmap.begin (), bmap.end ()) + 1;
  std::vector<int> tmp (nglobal, 0);
  for (int i = 0; i < nb; i++) tmp[bmap[i]] = bmsk[i];
  const int unknowns = nglobal - (int) std::count_if (tmp.begin (), tmp.end (),
                                                      [] (int v) { return v != 0; });
  if (unknowns < nglobal) {
    std::vector<int> order (nglobal);
    for (int g = 0; g < nglobal; g++) order[g] = g;
    std::stable_sort (order.begin (), order.end (),
                      [&] (int a, int b) { return (tmp[a] != 0) < (tmp[b] != 0); });
    std::vector<int> renum (nglobal);
    for (int k = 0; k < nglobal; k++) renum[order[k]] = k;
    for (int i = 0; i < nb; i++) bmap[i] = renum[bmap[i]];
  }
  const int nsolve = unknowns;
  // Nsys::renumber: RCM of the unknowns' graph (element boundary nodes of
  // one element are all connected; essential nodes left out)
  if (optlev && nsolve) {
    std::vector<std::vector<int>> adjl (nsolve);
    for (int e = 0; e < nel; e++) {
      const int* bm = &bmap[(size_t) e * next];
      const int* mk = &bmsk[(size_t) e * next];
      for (int i = 0; i < next; i++) {
        if (mk[i]) continue;
        std::vector<int>& L = adjl[bm[i]];
        for (int j = 0; j < next; j++)
          if (i != j && !mk[j] && std::find (L.begin (), L.end (), bm[j]) == L.end ())
            L.push_back (bm[j]);
      }
    }
    std::vector<int> xadj (nsolve + 2), adj (1);
    for (int i = 0, k = 1; i < nsolve; i++) {
      xadj[i + 1] = k;
      for (int v : adjl[i]) { adj.push_back (v + 1); k++; }
      xadj[nsolve + 1] = k;
    }
    adj.push_back (0);
    std::vector<int> perm (nsolve + 2), mask (nsolve + 2), xls (nsolve + 2),
        invperm (nglobal);
    const std::vector<int> bsave = bmap;
    for (int i = nsolve; i < nglobal; i++) invperm[i] = i;
    auto apply = [&] () {
      for (int i = 0; i < nsolve; i++) invperm[perm[i + 1] - 1] = i;
      for (int i = 0; i < nb; i++) bmap[i] = invperm[bsave[i]];
    };
    int nlvl = 0, root = 1;
    if (optlev == 1) {
      std::fill (mask.begin (), mask.end (), 1);
      fnroot (root, xadj, adj, mask, nlvl, xls, perm);
      rcm (root, xadj, adj, mask, perm, nlvl, xls);
    } else {
      int best = 1, bwmin = INT_MAX;
      for (int r = 1; r <= nsolve; r += (optlev == 2 ? 10 : 1)) {
        int rt = r;
        std::fill (mask.begin (), mask.end (), 1);
        if (optlev == 2) fnroot (rt, xadj, adj, mask, nlvl, xls, perm);
        rcm (rt, xadj, adj, mask, perm, nlvl, xls);
        if (nlvl != nsolve) continue;     // root's component only: not a full ordering
        apply ();
        const int bw = band_of (bmap, bmsk, nel, next);
        if (bw < bwmin) { bwmin = bw; best = rt; }
      }
      std::fill (mask.begin (), mask.end (), 1);
      rcm (best, xadj, adj, mask, perm, nlvl, xls);
    }
    if (nlvl != nsolve) {
      set_error ("sx_enumerate: the unknowns' graph is not connected (RCM numbered " +
                 std::to_string (nlvl) + " of " + std::to_string (nsolve) + ")");
      return SX_ERR_ARG;
    }
    apply ();
  }
  std::copy (bmap.begin (), bmap.end (), btog);
  std::copy (bmsk.begin (), bmsk.end (), bmask);
  *nglobal_out = nglobal;
  *nsolve_out = nsolve;
  *nband_out = band_of (bmap, bmsk, nel, next);
  return SX_OK;
}
