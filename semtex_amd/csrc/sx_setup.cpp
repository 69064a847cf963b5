// sx_setup.cpp -- host-side set-up of the semtex_amd library: quadrature,
// element geometry, element-class sharing, and the gather/scatter and
// assembly maps that the GPU kernels consume.  None of this is on the timed
// path; it runs once per mesh.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>

#include "sx_internal.h"

namespace sx {

static thread_local std::string g_err;

void set_error (const std::string& msg) { g_err = msg; }

int hip_fail (hipError_t e, const char* what)
{
  set_error (std::string ("HIP error ") + hipGetErrorString (e) + " in " + what);
  return SX_ERR_HIP;
}

// GLL points by Newton iteration on the Legendre recurrence (published
// method: G. von Winckel, "lglnodes"), weights 2/(N(N+1)P_N^2), classical
// collocation derivative (Canuto et al., Spectral Methods, eq. 2.4.34).
static void legendre (int N, double x, double& PN, double& PNm1)
{
  double p0 = 1.0, p1 = x;
  if (N == 0) { PN = 1.0; PNm1 = 0.0; return; }
  for (int k = 2; k <= N; k++) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1; p1 = p2;
  }
  PN = p1; PNm1 = p0;
}

void gll (int np, double* z, double* w, double* D)
{
  const int N = np - 1;
  if (np == 1) { z[0] = 0.0; w[0] = 2.0; if (D) D[0] = 0.0; return; }
  for (int i = 0; i < np; i++) {
    double x = -std::cos (M_PI * i / N);
    if (i == 0) x = -1.0;
    else if (i == N) x = 1.0;
    else
      for (int it = 0; it < 100; it++) {
        double pN, pNm1;
        legendre (N, x, pN, pNm1);
        const double xold = x;
        x = xold - (xold * pN - pNm1) / ((N + 1.0) * pN);
        if (std::fabs (x - xold) < 1e-16) break;
      }
    z[i] = x;
  }
  for (int i = 0; i < np / 2; i++) {
    const double a = 0.5 * (z[N - i] - z[i]);
    z[i] = -a; z[N - i] = a;
  }
  if (np & 1) z[N / 2] = 0.0;
  std::vector<double> PN (np);
  for (int i = 0; i < np; i++) {
    double pNm1;
    legendre (N, z[i], PN[i], pNm1);
    w[i] = 2.0 / (N * (N + 1.0) * PN[i] * PN[i]);
  }
  if (D)
    for (int i = 0; i < np; i++)
      for (int j = 0; j < np; j++) {
        double d = 0.0;
        if (i != j) d = PN[i] / (PN[j] * (z[i] - z[j]));
        else if (i == 0) d = -0.25 * N * (N + 1.0);
        else if (i == N) d = 0.25 * N * (N + 1.0);
        D[i * np + j] = d;
      }
}

void edgemaps (int np, int* emap, int* pmap)
{
  const int nm = np - 1;
  int k = 0, n = 0;
  emap[k++] = 0;
  const int steps[4] = {1, np, -1, -np};
  for (int s = 0; s < 4; s++)
    for (int i = 1; i < np; i++) { n += steps[s]; if (k < 4 * nm) emap[k++] = n; }
  for (int i = 1; i < nm; i++)
    for (int j = 1; j < nm; j++) emap[k++] = i * np + j;
  if (pmap) for (int i = 0; i < np * np; i++) pmap[emap[i]] = i;
}

static void mxm (const double* A, int nra, const double* B, int nca, double* C,
                 int ncb)
{
  for (int i = 0; i < nra; i++)
    for (int j = 0; j < ncb; j++) {
      double s = 0.0;
      for (int k = 0; k < nca; k++) s += A[i * nca + k] * B[k * ncb + j];
      C[i * ncb + j] = s;
    }
}

bool element_geometry (int np, int cyl, const double* x, const double* y,
                       double* geo)
{
  const int npnp = np * np;
  std::vector<double> z (np), w (np), D (npnp), DT (npnp), buf (4 * npnp);
  gll (np, z.data (), w.data (), D.data ());
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++) DT[j * np + i] = D[i * np + j];
  double *dxdr = buf.data (), *dxds = dxdr + npnp, *dydr = dxds + npnp,
         *dyds = dydr + npnp;
  mxm (x, np, DT.data (), np, dxdr, np);
  mxm (D.data (), np, x, np, dxds, np);
  mxm (y, np, DT.data (), np, dydr, np);
  mxm (D.data (), np, y, np, dyds, np);
  bool ok = true;
  const double EPS = 4.0 * 6.0e-14;
  for (int i = 0, ij = 0; i < np; i++)
    for (int j = 0; j < np; j++, ij++) {
      const double jac = dxdr[ij] * dyds[ij] - dxds[ij] * dydr[ij];
      if (jac < EPS) ok = false;
      const double WW = w[i] * w[j];
      const double r = cyl ? y[ij] : 1.0;
      geo[G_Q1 * npnp + ij] = (dyds[ij] * dyds[ij] + dxds[ij] * dxds[ij]) / jac * WW * r;
      geo[G_Q2 * npnp + ij] = (dydr[ij] * dydr[ij] + dxdr[ij] * dxdr[ij]) / jac * WW * r;
      geo[G_Q3 * npnp + ij] = (-dydr[ij] * dyds[ij] - dxdr[ij] * dxds[ij]) / jac * WW * r;
      geo[G_Q4 * npnp + ij] = jac * WW * r;
      geo[G_Q8 * npnp + ij] = jac * WW;
      geo[G_RX * npnp + ij] = dyds[ij] / jac;
      geo[G_RY * npnp + ij] = -dxds[ij] / jac;
      geo[G_SX * npnp + ij] = -dydr[ij] / jac;
      geo[G_SY * npnp + ij] = dxdr[ij] / jac;
      geo[G_Y * npnp + ij] = y[ij];
    }
  // null-mapping optimisation of the reference: arrays that vanish to
  // round-off are exactly zero (src/element.cpp:1402-1407).
  for (int a : {G_Q3, G_RX, G_RY, G_SX, G_SY}) {
    double s = 0.0;
    for (int ij = 0; ij < npnp; ij++) s += geo[a * npnp + ij] * geo[a * npnp + ij];
    if (std::sqrt (s) < EPS)
      std::fill (geo + a * npnp, geo + (a + 1) * npnp, 0.0);
  }
  return ok;
}

void side_area (int np, const double* x, const double* y, int side,
                double* area)
{
  std::vector<double> z (np), w (np), D (np * np);
  gll (np, z.data (), w.data (), D.data ());
  for (int k = 0; k < np; k++) {
    double dx = 0.0, dy = 0.0;
    if (side == 0 || side == 2) {
      const int row = side == 0 ? 0 : np - 1, col = side == 0 ? k : np - 1 - k;
      for (int q = 0; q < np; q++) {
        dx += D[col * np + q] * x[row * np + q];
        dy += D[col * np + q] * y[row * np + q];
      }
      area[k] = std::sqrt (dx * dx + dy * dy) * w[col];
    } else {
      const int col = side == 1 ? np - 1 : 0, row = side == 1 ? k : np - 1 - k;
      for (int q = 0; q < np; q++) {
        dx += D[row * np + q] * x[q * np + col];
        dy += D[row * np + q] * y[q * np + col];
      }
      area[k] = std::sqrt (dx * dx + dy * dy) * w[row];
    }
  }
}

}  // namespace sx

using namespace sx;

extern "C" const char* sx_last_error (void) { return g_err.c_str (); }

extern "C" int sx_device_info (char* buf, int len)
{
  int n = 0;
  if (hipGetDeviceCount (&n) != hipSuccess || n == 0) {
    if (buf && len > 0) std::snprintf (buf, len, "no HIP device");
    set_error ("no HIP device");
    return SX_ERR_HIP;
  }
  hipDeviceProp_t p;
  SX_HIP (hipGetDeviceProperties (&p, 0));
  if (buf && len > 0)
    std::snprintf (buf, len, "%s %s CUs=%d", p.name, p.gcnArchName,
                   p.multiProcessorCount);
  if (std::strncmp (p.gcnArchName, "gfx950", 6) != 0) {
    set_error (std::string ("expected gfx950, found ") + p.gcnArchName);
    return SX_ERR_HIP;
  }
  return SX_OK;
}

extern "C" int sx_gll (int np, double* z, double* w, double* D)
{
  if (np < 1 || !z || !w) { set_error ("sx_gll: bad arguments"); return SX_ERR_ARG; }
  gll (np, z, w, D);
  return SX_OK;
}

template <typename T>
static int upload (T** dst, const std::vector<T>& src)
{
  const size_t n = std::max<size_t> (src.size (), 1);
  SX_HIP (hipMalloc ((void**) dst, n * sizeof (T)));
  if (!src.empty ())
    SX_HIP (hipMemcpy (*dst, src.data (), src.size () * sizeof (T),
                       hipMemcpyHostToDevice));
  return SX_OK;
}

extern "C" int sx_mesh_destroy (sx_mesh_t m)
{
  if (!m) return SX_OK;
  void* ptrs[] = {m->d_geo, m->d_cgeo, m->d_D, m->d_cls, m->d_btog, m->d_bmask,
                  m->d_emask, m->d_emap, m->d_gptr, m->d_gsrc, m->d_ess,
                  m->d_nptr, m->d_nidx, m->d_nw, m->d_ppos, m->d_pmn,
                  m->d_pptr, m->d_psrc, m->d_tile, m->d_elist, m->d_ess_elem,
                  m->d_ttab, m->d_ess_slot, m->d_gpos, m->d_gq4};
  for (void* p : ptrs) if (p) (void) hipFree (p);
  delete m;
  return SX_OK;
}

extern "C" int sx_mesh_nclass (sx_mesh_t m) { return m ? m->nclass : 0; }

extern "C" int sx_mesh_create (int np, int nel, int cyl, const double* x,
                               const double* y, const int* btog,
                               const int* bmask, const int* emask, int nglobal,
                               int nsolve, int nband, int nbound,
                               const int* bnd_elmt, const int* bnd_side,
                               const int* bnd_kind, sx_mesh_t* out)
{
  if (!out || np < 3 || np > 16 || nel < 1 || !x || !y || !btog || !bmask ||
      !emask || nglobal < 1 || nsolve < 0 || nsolve > nglobal || nband < 1 ||
      nbound < 0 || (nbound && (!bnd_elmt || !bnd_side || !bnd_kind))) {
    set_error ("sx_mesh_create: bad arguments (np must be 3..16)");
    return SX_ERR_ARG;
  }
  for (int b = 0; b < nbound; b++)
    if (bnd_kind[b] != SX_ESSENTIAL && bnd_kind[b] != SX_NATURAL && bnd_kind[b] != SX_AXIS) {
      set_error ("sx_mesh_create: boundary kind must be SX_ESSENTIAL, SX_NATURAL or SX_AXIS");
      return SX_ERR_ARG;
    }
  sx_mesh* m = new sx_mesh;
  *out = nullptr;
  m->np = np; m->nel = nel; m->cyl = cyl ? 1 : 0;
  m->npnp = np * np; m->next = 4 * (np - 1); m->nint = m->npnp - m->next;
  m->nglobal = nglobal; m->nsolve = nsolve; m->nband = nband; m->kd = nband - 1;
  m->nbound = nbound;
  const int npnp = m->npnp, next = m->next;

  m->btog.assign (btog, btog + (size_t) nel * next);
  m->hx.assign (x, x + (size_t) nel * npnp);
  m->hy.assign (y, y + (size_t) nel * npnp);
  m->bnd_elmt.assign (bnd_elmt, bnd_elmt + nbound);
  m->bnd_side.assign (bnd_side, bnd_side + nbound);
  m->bnd_kind.assign (bnd_kind, bnd_kind + nbound);
  m->bmask.assign (bmask, bmask + (size_t) nel * next);
  m->emask.assign (emask, emask + nel);
  for (int v : m->btog)
    if (v < 0 || v >= nglobal) {
      sx_mesh_destroy (m);
      set_error ("sx_mesh_create: btog entry out of range");
      return SX_ERR_ARG;
    }
  for (int b = 0; b < nbound; b++)
    if (bnd_elmt[b] < 0 || bnd_elmt[b] >= nel || bnd_side[b] < 0 || bnd_side[b] > 3) {
      sx_mesh_destroy (m);
      set_error ("sx_mesh_create: boundary segment out of range");
      return SX_ERR_ARG;
    }

  // band width consistency (utility/enumerate.cpp:923-965)
  for (int e = 0; e < nel; e++) {
    int lo = INT32_MAX, hi = -1;
    for (int i = 0; i < next; i++) {
      const int g = m->btog[(size_t) e * next + i];
      if (g < nsolve) { lo = std::min (lo, g); hi = std::max (hi, g); }
    }
    if (hi >= 0 && hi - lo + 1 > nband) {
      sx_mesh_destroy (m);
      set_error ("sx_mesh_create: nband smaller than the numbering's band width");
      return SX_ERR_ARG;
    }
  }

  m->emap.resize (npnp);
  std::vector<int> pmap (npnp);
  edgemaps (np, m->emap.data (), pmap.data ());

  // geometry per element, and classes of identical geometry (Family
  // sharing, src/family.cpp; tolerance relative to each array's size)
  std::vector<double> geo ((size_t) nel * NGEO * npnp);
  for (int e = 0; e < nel; e++)
    if (!element_geometry (np, m->cyl, x + (size_t) e * npnp, y + (size_t) e * npnp,
                           geo.data () + (size_t) e * NGEO * npnp)) {
      sx_mesh_destroy (m);
      set_error ("sx_mesh_create: element " + std::to_string (e + 1) +
                 " has a non-positive Jacobian");
      return SX_ERR_GEOM;
    }
  m->cls.assign (nel, -1);
  std::vector<int> rep;
  const int keyarr[] = {G_Q1, G_Q2, G_Q3, G_Q4, G_Y};
  const int nkey = m->cyl ? 5 : 4;
  std::vector<double> scale (nkey, 0.0);
  for (int e = 0; e < nel; e++)
    for (int a = 0; a < nkey; a++)
      for (int ij = 0; ij < npnp; ij++)
        scale[a] = std::max (scale[a], std::fabs (geo[((size_t) e * NGEO + keyarr[a]) * npnp + ij]));
  for (int e = 0; e < nel; e++) {
    const double* ge = geo.data () + (size_t) e * NGEO * npnp;
    for (size_t c = 0; c < rep.size () && m->cls[e] < 0; c++) {
      const double* gr = geo.data () + (size_t) rep[c] * NGEO * npnp;
      bool same = true;
      for (int a = 0; a < nkey && same; a++) {
        const double tol = 1e-13 * std::max (scale[a], 1e-300);
        for (int ij = 0; ij < npnp; ij++)
          if (std::fabs (ge[keyarr[a] * npnp + ij] - gr[keyarr[a] * npnp + ij]) > tol) {
            same = false; break;
          }
      }
      if (same) m->cls[e] = (int) c;
    }
    if (m->cls[e] < 0) { m->cls[e] = (int) rep.size (); rep.push_back (e); }
  }
  m->nclass = (int) rep.size ();
  std::vector<double> cgeo ((size_t) m->nclass * NGEO * npnp);
  for (int c = 0; c < m->nclass; c++)
    std::copy (geo.begin () + (size_t) rep[c] * NGEO * npnp,
               geo.begin () + (size_t) (rep[c] + 1) * NGEO * npnp,
               cgeo.begin () + (size_t) c * NGEO * npnp);

  std::vector<double> z (np), w (np), D (npnp);
  gll (np, z.data (), w.data (), D.data ());

  // global node -> element-boundary slots (deterministic direct-stiffness
  // summation order: increasing element, then CCW position)
  std::vector<int> cnt (nglobal + 1, 0);
  for (int v : m->btog) cnt[v + 1]++;
  for (int g = 0; g < nglobal; g++) cnt[g + 1] += cnt[g];
  std::vector<int> gsrc (cnt[nglobal]), fill (cnt.begin (), cnt.end () - 1);
  for (int e = 0; e < nel; e++)
    for (int i = 0; i < next; i++)
      gsrc[fill[m->btog[(size_t) e * next + i]]++] = e * npnp + i;

  // essential values: last writer in boundary order (Field::getEssential)
  // and natural-BC quadrature contributions (Natural::sum)
  std::vector<int> ess (nglobal, -1);
  std::vector<std::vector<std::pair<int, double>>> nat (nglobal);
  const int nm = np - 1;
  std::vector<double> area (np);
  for (int b = 0; b < nbound; b++) {
    const int e = bnd_elmt[b], s = bnd_side[b];
    const int* bm = m->btog.data () + (size_t) e * next;
    if (bnd_kind[b] == SX_NATURAL)
      side_area (np, x + (size_t) e * npnp, y + (size_t) e * npnp, s, area.data ());
    for (int k = 0; k < np; k++) {
      const int loc = (k < nm) ? s * nm + k : ((s == 3) ? 0 : s * nm + nm);
      const int g = bm[loc];
      if (bnd_kind[b] == SX_ESSENTIAL) ess[g] = b * np + k;
      else if (bnd_kind[b] == SX_NATURAL) nat[g].push_back ({b * np + k, area[k]});
      // SX_AXIS: natural with no flux, no contribution
    }
  }
  std::vector<int> nptr (nglobal + 1, 0), nidx;
  std::vector<double> nw;
  for (int g = 0; g < nglobal; g++) {
    for (auto& p : nat[g]) { nidx.push_back (p.first); nw.push_back (p.second); }
    nptr[g + 1] = (int) nidx.size ();
  }

  // assembled band entries (m <= n < nsolve) and their condensed-matrix
  // sources (src/matrix.cpp:229-234)
  std::map<std::pair<int, int>, std::vector<int>> pairs;
  for (int e = 0; e < nel; e++) {
    const int* bm = m->btog.data () + (size_t) e * next;
    for (int i = 0; i < next; i++) {
      if (bm[i] >= nsolve) continue;
      for (int k = 0; k < next; k++)
        if (bm[k] < nsolve && bm[k] >= bm[i])
          pairs[{bm[i], bm[k]}].push_back (m->cls[e] * next * next + i * next + k);
    }
  }
  std::vector<long long> ppos;
  std::vector<int> pmn, pptr (1, 0), psrc;
  for (auto& kv : pairs) {
    const int r = kv.first.first, c = kv.first.second;
    ppos.push_back ((long long) c * nband + (nband - 1) + r - c);
    pmn.push_back (r); pmn.push_back (c);
    for (int s : kv.second) psrc.push_back (s);
    pptr.push_back ((int) psrc.size ());
  }
  m->npairs = (int) ppos.size ();

  // class-sorted element tiles and the elements with essential nodes
  std::vector<int> elist, tile, ess_elem;
  for (int c = 0; c < m->nclass; c++) {
    const int first = (int) elist.size ();
    for (int e = 0; e < nel; e++)
      if (m->cls[e] == c) elist.push_back (e);
    for (int t = first; t < (int) elist.size (); t += ETILE) {
      tile.push_back (c);
      tile.push_back (t);
      tile.push_back (std::min (ETILE, (int) elist.size () - t));
    }
  }
  for (int e = 0; e < nel; e++)
    if (m->emask[e]) ess_elem.push_back (e);
  m->ntile = (int) tile.size () / 3;
  const int ti = tile_ints (m->next);
  std::vector<int> ttab ((size_t) m->ntile * ti, 0);
  for (int t = 0; t < m->ntile; t++) {
    int* T = ttab.data () + (size_t) t * ti;
    const int t0 = tile[3 * t + 1], cnt = tile[3 * t + 2];
    T[0] = tile[3 * t];
    T[1] = cnt;
    for (int j = 0; j < ETILE; j++) {
      const int e = elist[t0 + std::min (j, cnt - 1)];
      T[TT_EL + j] = e;
      T[TT_EM + j] = m->emask[e];
      for (int a = 0; a < m->next; a++) {
        T[TT_BT + j * m->next + a] = m->btog[(size_t) e * m->next + a];
        T[TT_BT + (ETILE + j) * m->next + a] = m->bmask[(size_t) e * m->next + a];
      }
    }
  }
  m->ness = (int) ess_elem.size ();
  if (ess_elem.empty ()) ess_elem.push_back (0);
  // every slot of the CSR lists as its row-major plane offset and its Q4
  // weight (the smoothing's operands without the index chain)
  std::vector<int> gpos (std::max<size_t> (gsrc.size (), 1), 0);
  std::vector<double> gq4 (gpos.size (), 0.0);
  for (size_t c = 0; c < gsrc.size (); c++) {
    const int slot = gsrc[c], e = slot / npnp, ij = m->emap[slot % npnp];
    gpos[c] = e * npnp + ij;
    gq4[c] = geo[((size_t) e * NGEO + G_Q4) * npnp + ij];
  }
  // the line slot of every boundary slot of the essential elements (-1: none)
  std::vector<int> ess_slot ((size_t) ess_elem.size () * m->next, -1);
  for (int i = 0; i < m->ness; i++)
    for (int a = 0; a < m->next; a++)
      ess_slot[(size_t) i * m->next + a] = ess[m->btog[(size_t) ess_elem[i] * m->next + a]];

  int rc = SX_OK;
  std::vector<int> cls (m->cls);
  if ((rc = upload (&m->d_geo, geo)) || (rc = upload (&m->d_cgeo, cgeo)) ||
      (rc = upload (&m->d_D, D)) || (rc = upload (&m->d_cls, cls)) ||
      (rc = upload (&m->d_btog, m->btog)) || (rc = upload (&m->d_bmask, m->bmask)) ||
      (rc = upload (&m->d_emask, m->emask)) || (rc = upload (&m->d_emap, m->emap)) ||
      (rc = upload (&m->d_gptr, cnt)) || (rc = upload (&m->d_gsrc, gsrc)) ||
      (rc = upload (&m->d_ess, ess)) || (rc = upload (&m->d_nptr, nptr)) ||
      (rc = upload (&m->d_nidx, nidx)) || (rc = upload (&m->d_nw, nw)) ||
      (rc = upload (&m->d_ppos, ppos)) || (rc = upload (&m->d_pmn, pmn)) ||
      (rc = upload (&m->d_pptr, pptr)) || (rc = upload (&m->d_psrc, psrc)) ||
      (rc = upload (&m->d_tile, tile)) || (rc = upload (&m->d_elist, elist)) ||
      (rc = upload (&m->d_ess_elem, ess_elem)) || (rc = upload (&m->d_ttab, ttab)) ||
      (rc = upload (&m->d_ess_slot, ess_slot)) || (rc = upload (&m->d_gpos, gpos)) ||
      (rc = upload (&m->d_gq4, gq4))) {
    sx_mesh_destroy (m);
    return rc;
  }
  *out = m;
  return SX_OK;
}
