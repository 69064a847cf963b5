// sx_pcg.hip -- Field::solve, JACPCG branch (src/field.cpp:711-805), for
// every solved plane of a field at once.  Each plane runs its own Jacobi-
// preconditioned conjugate-gradient iteration on the uncondensed system in
// the reference's global numbering (nglobal element-boundary values, then
// nel blocks of nint interior values in emap order), with its own stopping
// test, so every plane stops at the iteration the reference's loop would.
//
//   set-up  (src/matrix.cpp:269-310)   pc = 1 / assembled diag(A)
//   solve   getEssential, constrain, buildRHS (interiors not condensed),
//           x0 from the field, r = b - A x0 (essential entries masked),
//           while r.r > (TOL_REL |b|)^2 and ++i < STEP_MAX:
//             z = pc r; rho = r.z; p = z + rho/rho_old p; q = A p (masked);
//             alpha = rho / p.q; x += alpha p; r -= alpha q
//           global2local, essential values inserted.
//
// A p is matrix-free: one wavefront per (element, plane) gathers p, applies
// the elemental operator (Element::HelmholtzOp) and writes its local
// result; a second kernel sums them per global node (deterministic CSR
// order, no atomics) and fuses the dot products and vector updates, one
// workgroup per plane.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sx_internal.h"

using namespace sx;

namespace {

constexpr double EPSDP = 6.0e-14;
constexpr int PV_THREADS = 1024;   // one workgroup per plane for vector work

__device__ __forceinline__ int solved_plane (int y, int nz, int base_mode)
{
  return (base_mode == 0 && nz > 1 && y >= 1) ? y + 1 : y;
}

__device__ double block_sum (double v, double* red)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads ();
  if (lane == 0) red[w] = v;
  __syncthreads ();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; i++) s += red[i];
  return s;
}

// elemental operator on one element by one wavefront (Element::HelmholtzOp,
// src/element.cpp:1593-1622); src/out row-major in LDS
__device__ void helm_op_elem (int np, const double* D, const double* geo, double lambda2,
                              double betak2, int cyl, const double* src, double* R,
                              double* S, double* out, int lane)
{
  const int npnp = np * np;
  for (int ij = lane; ij < npnp; ij += 64) {
    const int i = ij / np, j = ij % np;
    double r = 0.0, s = 0.0;
    for (int k = 0; k < np; k++) {
      r += src[i * np + k] * D[j * np + k];
      s += D[i * np + k] * src[k * np + j];
    }
    const double g1 = geo[G_Q1 * npnp + ij], g2 = geo[G_Q2 * npnp + ij],
                 g3 = geo[G_Q3 * npnp + ij], g4 = geo[G_Q4 * npnp + ij];
    double hCon;
    if (cyl) {
      const double y = geo[G_Y * npnp + ij], r2 = y * y;
      hCon = (r2 > EPSDP) ? betak2 / r2 + lambda2 : 0.0;
    } else hCon = betak2 + lambda2;
    R[ij] = g1 * r + g3 * s;
    S[ij] = g2 * s + g3 * r;
    out[ij] = g4 * src[ij] * hCon;
  }
  __syncthreads ();
  for (int ij = lane; ij < npnp; ij += 64) {
    const int i = ij / np, j = ij % np;
    double t = 0.0;
    for (int k = 0; k < np; k++) t += D[k * np + i] * S[k * np + j] + R[i * np + k] * D[k * np + j];
    out[ij] += t;
  }
  __syncthreads ();
}

// Diagonal of the elemental matrix (Element::HelmholtzDiag,
// src/element.cpp:1484-1533) in emap order, per (element, mode).  The cross
// term reads the derivative matrix at linear indices j*j and i*i, as the
// reference does (it is absent on affine elements, where Q3 = 0).
__global__ __launch_bounds__(64) void k_pcg_elem_diag (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap, double* L, size_t lstride)
{
  const int npnp = np * np, e = blockIdx.x, mode = blockIdx.y;
  const double* g = geo + (size_t) e * NGEO * npnp;
  const double bk2 = betak2[mode];
  bool q3 = false;
  for (int ij = 0; ij < npnp; ij++) q3 = q3 || std::fabs (g[G_Q3 * npnp + ij]) > 0.0;
  for (int a = threadIdx.x; a < npnp; a += 64) {
    const int ij = emap[a], i = ij / np, j = ij % np;
    double hcon;
    if (cyl) {
      const double y = g[G_Y * npnp + ij], r2 = y * y;
      hcon = (r2 > EPSDP) ? bk2 / r2 + lambda2 : 0.0;
    } else hcon = lambda2 + bk2;
    double d = 0.0;
    for (int k = 0; k < np; k++) d += g[G_Q1 * npnp + i * np + k] * Dg[k * np + j] * Dg[k * np + j];
    for (int k = 0; k < np; k++) d += g[G_Q2 * npnp + k * np + j] * Dg[k * np + i] * Dg[k * np + i];
    if (q3) d += 2.0 * g[G_Q3 * npnp + ij] * Dg[j * j] * Dg[i * i];
    d += hcon * g[G_Q4 * npnp + ij];
    L[(size_t) mode * lstride + (size_t) e * npnp + a] = d;
  }
}

// pc = 1 / (assembled diagonal), per mode
__global__ void k_pcg_assemble_diag (int nglobal, int nel, int npnp, int next,
                                     const int* __restrict__ gptr,
                                     const int* __restrict__ gsrc,
                                     const double* __restrict__ L, size_t lstride, double* pc,
                                     size_t npts)
{
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int mode = blockIdx.y;
  if (idx >= npts) return;
  const double* Lm = L + (size_t) mode * lstride;
  double d;
  if (idx < (size_t) nglobal) {
    d = 0.0;
    for (int c = gptr[idx]; c < gptr[idx + 1]; c++) d += Lm[gsrc[c]];
  } else {
    const int nint = npnp - next, q = (int) (idx - nglobal), e = q / nint;
    d = Lm[(size_t) e * npnp + next + q % nint];
  }
  pc[(size_t) mode * npts + idx] = 1.0 / d;
}

// constrain: W = -(M f + H g) in emap order into L (per solved plane)
__global__ __launch_bounds__(64) void k_pcg_weak (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ emask, const double* __restrict__ G, int nglobal,
    const double* __restrict__ F, int stride, double* L, size_t lstride, int nz,
    int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1);
  const int e = blockIdx.x, y = blockIdx.y, k = solved_plane (y, nz, base_mode);
  const int lane = threadIdx.x;
  __shared__ double D[256], buf[256], g[256], R[256], S[256], Hg[256];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) { D[i] = Dg[i]; emap[i] = emap_g[i]; }
  const double* ge = geo + (size_t) e * NGEO * npnp;
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) buf[i] = F[off + i] * ge[G_Q8 * npnp + i];
  __syncthreads ();
  if (emask[e]) {
    for (int i = lane; i < npnp; i += 64) g[i] = 0.0;
    __syncthreads ();
    for (int i = lane; i < next; i += 64)
      g[emap[i]] = G[(size_t) k * nglobal + btog[(size_t) e * next + i]];
    __syncthreads ();
    helm_op_elem (np, D, ge, lambda2, betak2[k >> 1], cyl, g, R, S, Hg, lane);
    for (int i = lane; i < npnp; i += 64) buf[i] += Hg[i];
    __syncthreads ();
  }
  double* Le = L + (size_t) y * lstride + (size_t) e * npnp;
  for (int a = lane; a < npnp; a += 64) Le[a] = -buf[emap[a]];
}

// y_p = A p locally: gather the plane's global vector into element e,
// apply the elemental operator, write the local result in emap order
__global__ __launch_bounds__(64) void k_pcg_op (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog, int nglobal,
    const double* __restrict__ V, size_t npts, double* L, size_t lstride,
    const int* __restrict__ active, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, y = blockIdx.y, k = solved_plane (y, nz, base_mode);
  if (!active[y]) return;
  const int lane = threadIdx.x;
  __shared__ double D[256], src[256], R[256], S[256], out[256];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) { D[i] = Dg[i]; emap[i] = emap_g[i]; }
  __syncthreads ();
  const double* v = V + (size_t) y * npts;
  for (int i = lane; i < next; i += 64) src[emap[i]] = v[btog[(size_t) e * next + i]];
  for (int q = lane; q < nint; q += 64) src[emap[next + q]] = v[nglobal + (size_t) e * nint + q];
  __syncthreads ();
  helm_op_elem (np, D, geo + (size_t) e * NGEO * npnp, lambda2, betak2[k >> 1], cyl, src, R,
                S, out, lane);
  double* Le = L + (size_t) y * lstride + (size_t) e * npnp;
  for (int a = lane; a < npnp; a += 64) Le[a] = out[emap[a]];
}

// global value idx of plane y from local results (direct stiffness sum)
__device__ __forceinline__ double assemble_at (size_t idx, int nglobal, int npnp, int next,
                                               const int* gptr, const int* gsrc,
                                               const double* Ly)
{
  if (idx < (size_t) nglobal) {
    double s = 0.0;
    for (int c = gptr[idx]; c < gptr[idx + 1]; c++) s += Ly[gsrc[c]];
    return s;
  }
  const int nint = npnp - next, q = (int) (idx - nglobal);
  return Ly[(size_t) (q / nint) * npnp + next + q % nint];
}

struct PcgState {
  double *x, *r, *p, *q, *z;        // [npl][npts]
  double *rho, *eps, *r2;           // [npl]
  int *it, *active;                 // [npl]
};

// b = assembled W (+ <h, w>), essential tail zero; x0 from the field,
// essential tail zero; eps = (tol |b|)^2.  One workgroup per plane.
__global__ __launch_bounds__(PV_THREADS) void k_pcg_setup (
    int nglobal, int npnp, int next, int nel, const int* __restrict__ gptr,
    const int* __restrict__ gsrc, const int* __restrict__ nptr,
    const int* __restrict__ nidx, const double* __restrict__ nw,
    const double* __restrict__ lines, int lstride_lines, const int* __restrict__ nsolve,
    const int* __restrict__ emap, const double* __restrict__ U, int stride,
    const double* __restrict__ L, size_t lstride, PcgState st, size_t npts, double tol,
    int nz, int base_mode)
{
  __shared__ double red[PV_THREADS / 64];
  const int y = blockIdx.x, k = solved_plane (y, nz, base_mode), ns = nsolve[k >> 1];
  const int nint = npnp - next;
  const double* Ly = L + (size_t) y * lstride;
  double* r = st.r + (size_t) y * npts;
  double* x = st.x + (size_t) y * npts;
  double bb = 0.0;
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS) {
    double b = assemble_at (i, nglobal, npnp, next, gptr, gsrc, Ly);
    if (i < (size_t) nglobal) {
      if (lines)
        for (int c = nptr[i]; c < nptr[i + 1]; c++)
          b += lines[(size_t) k * lstride_lines + nidx[c]] * nw[c];
      if ((int) i >= ns) b = 0.0;
    }
    r[i] = b;
    bb += b * b;
    // initial guess (Field::local2global: the last element writing a node wins)
    double xv;
    if (i < (size_t) nglobal) {
      const int slot = gsrc[gptr[i + 1] - 1], e = slot / npnp, a = slot % npnp;
      xv = ((int) i >= ns) ? 0.0 : U[(size_t) k * stride + (size_t) e * npnp + emap[a]];
    } else {
      const int q = (int) (i - nglobal), e = q / nint;
      xv = U[(size_t) k * stride + (size_t) e * npnp + emap[next + q % nint]];
    }
    x[i] = xv;
  }
  bb = block_sum (bb, red);
  if (threadIdx.x == 0) {
    const double e = tol * std::sqrt (bb);
    st.eps[y] = e * e;
  }
}

// r -= (A x0) (masked); r2; active for iteration 1
__global__ __launch_bounds__(PV_THREADS) void k_pcg_resid0 (
    int nglobal, int npnp, int next, const int* __restrict__ gptr,
    const int* __restrict__ gsrc, const int* __restrict__ nsolve,
    const double* __restrict__ L, size_t lstride, PcgState st, size_t npts, int step_max,
    int nz, int base_mode)
{
  __shared__ double red[PV_THREADS / 64];
  const int y = blockIdx.x, k = solved_plane (y, nz, base_mode), ns = nsolve[k >> 1];
  const double* Ly = L + (size_t) y * lstride;
  double* r = st.r + (size_t) y * npts;
  double rr = 0.0;
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS) {
    double v = r[i];
    if (!(i < (size_t) nglobal && (int) i >= ns)) v -= assemble_at (i, nglobal, npnp, next, gptr, gsrc, Ly);
    else v = 0.0;
    r[i] = v;
    rr += v * v;
  }
  rr = block_sum (rr, red);
  if (threadIdx.x == 0) {
    st.r2[y] = rr;
    st.it[y] = 0;
    st.active[y] = (rr > st.eps[y]) && (1 < step_max);
  }
}

// z = pc r, rho = r.z, p = z (+ rho/rho_old p)
__global__ __launch_bounds__(PV_THREADS) void k_pcg_dir (
    const double* __restrict__ pc, size_t npts, PcgState st, int nz, int base_mode)
{
  __shared__ double red[PV_THREADS / 64];
  const int y = blockIdx.x, k = solved_plane (y, nz, base_mode);
  if (!st.active[y]) return;
  const double* pcm = pc + (size_t) (k >> 1) * npts;
  double* r = st.r + (size_t) y * npts;
  double* z = st.z + (size_t) y * npts;
  double* p = st.p + (size_t) y * npts;
  double rz = 0.0;
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS) {
    const double zi = pcm[i] * r[i];
    z[i] = zi;
    rz += r[i] * zi;
  }
  rz = block_sum (rz, red);
  const bool first = st.it[y] == 0;
  const double beta = first ? 0.0 : rz / st.rho[y];
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS)
    p[i] = first ? z[i] : beta * p[i] + z[i];
  __syncthreads ();
  if (threadIdx.x == 0) st.rho[y] = rz;
}

// q = A p (masked); alpha = rho / p.q; x += alpha p; r -= alpha q; r2;
// the next iteration runs while r2 > eps and ++i < STEP_MAX
__global__ __launch_bounds__(PV_THREADS) void k_pcg_update (
    int nglobal, int npnp, int next, const int* __restrict__ gptr,
    const int* __restrict__ gsrc, const int* __restrict__ nsolve,
    const double* __restrict__ L, size_t lstride, PcgState st, size_t npts, int step_max,
    int nz, int base_mode)
{
  __shared__ double red[PV_THREADS / 64];
  const int y = blockIdx.x, k = solved_plane (y, nz, base_mode), ns = nsolve[k >> 1];
  if (!st.active[y]) return;
  const double* Ly = L + (size_t) y * lstride;
  double* p = st.p + (size_t) y * npts;
  double* q = st.q + (size_t) y * npts;
  double pq = 0.0;
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS) {
    double v = 0.0;
    if (!(i < (size_t) nglobal && (int) i >= ns))
      v = assemble_at (i, nglobal, npnp, next, gptr, gsrc, Ly);
    q[i] = v;
    pq += p[i] * v;
  }
  pq = block_sum (pq, red);
  const double alpha = st.rho[y] / pq;
  double* x = st.x + (size_t) y * npts;
  double* r = st.r + (size_t) y * npts;
  double rr = 0.0;
  for (size_t i = threadIdx.x; i < npts; i += PV_THREADS) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rr += ri * ri;
  }
  rr = block_sum (rr, red);
  if (threadIdx.x == 0) {
    const int it = st.it[y] + 1;
    st.it[y] = it;
    st.r2[y] = rr;
    st.active[y] = (rr > st.eps[y]) && (it + 1 < step_max);
  }
}

// u = x (global2local), then the essential values (setEssential)
__global__ __launch_bounds__(64) void k_pcg_finish (
    int np, const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ G, int nglobal, const double* __restrict__ X, size_t npts,
    double* U, int stride, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, y = blockIdx.y, k = solved_plane (y, nz, base_mode);
  const double* x = X + (size_t) y * npts;
  double* u = U + (size_t) k * stride + (size_t) e * npnp;
  for (int i = threadIdx.x; i < next; i += 64) {
    const int g = btog[(size_t) e * next + i];
    double v = x[g];
    if (emask[e] && bmask[(size_t) e * next + i]) v = G[(size_t) k * nglobal + g];
    u[emap_g[i]] = v;
  }
  for (int q = threadIdx.x; q < nint; q += 64)
    u[emap_g[next + q]] = x[nglobal + (size_t) e * nint + q];
}

__global__ void k_fill_int (int* a, int v, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_pcg_essential (int nglobal, const int* __restrict__ ess,
                                 const double* __restrict__ lines, int lstride,
                                 double* G, int nz, int base_mode)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  if (g >= nglobal) return;
  const int s = ess[g];
  G[(size_t) k * nglobal + g] = (s >= 0 && lines) ? lines[(size_t) k * lstride + s] : 0.0;
}

}  // namespace

namespace sx {

// Set-up of a JACPCG system: the preconditioner of every mode.
int pcg_setup (sx_msys* s, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const size_t npts = (size_t) m->nglobal + (size_t) m->nel * m->nint;
  const size_t lstride = (size_t) m->nel * m->npnp;
  double* L = nullptr;
  SX_HIP (hipMalloc (&s->d_pc, sizeof (double) * s->nmodes * npts));
  SX_HIP (hipMalloc (&L, sizeof (double) * s->nmodes * lstride));
  hipLaunchKernelGGL (k_pcg_elem_diag, dim3 (m->nel, s->nmodes), dim3 (64), 0, st, m->np,
                      m->cyl, s->lambda2, s->d_betak2, m->d_geo, m->d_D, m->d_emap, L,
                      lstride);
  hipLaunchKernelGGL (k_pcg_assemble_diag, dim3 ((unsigned) ((npts + 255) / 256), s->nmodes),
                      dim3 (256), 0, st, m->nglobal, m->nel, m->npnp, m->next, m->d_gptr,
                      m->d_gsrc, L, lstride, s->d_pc, npts);
  SX_HIP (hipGetLastError ());
  SX_HIP (hipStreamSynchronize (st));
  (void) hipFree (L);
  return SX_OK;
}

}  // namespace sx

extern "C" int sx_field_solve_pcg (sx_msys_t s, int nz, int stride, double* d_f, double* d_u,
                                   const double* d_lines, double tol_rel, int step_max,
                                   int* iters, void* stream)
{
  if (!s || !d_f || !d_u || nz < 1 || stride < s->m->nel * s->m->npnp ||
      (nz > 2 * s->nmodes && nz != 1) || step_max < 1 || !(tol_rel >= 0.0)) {
    set_error ("sx_field_solve_pcg: bad arguments");
    return SX_ERR_ARG;
  }
  if (s->precision != SX_JACPCG || !s->d_pc) {
    set_error ("sx_field_solve_pcg: system was not created with SX_JACPCG");
    return SX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t) stream;
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const size_t npts = (size_t) m->nglobal + (size_t) m->nel * m->nint;
  const size_t lstride = (size_t) m->nel * m->npnp;
  const int lstride_lines = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : nullptr;
  // workspace: L [npl][lstride], x r p q z [npl][npts], scalars
  const size_t need = (size_t) npl * (lstride + 5 * npts) + 3 * (size_t) npl + 2 * (size_t) npl;
  if (s->work_doubles < need) {
    if (s->d_work) (void) hipFree (s->d_work);
    s->d_work = nullptr;
    s->work_doubles = 0;
    SX_HIP (hipMalloc (&s->d_work, sizeof (double) * need));
    s->work_doubles = need;
  }
  double* L = s->d_work;
  PcgState ps;
  ps.x = L + (size_t) npl * lstride;
  ps.r = ps.x + (size_t) npl * npts;
  ps.p = ps.r + (size_t) npl * npts;
  ps.q = ps.p + (size_t) npl * npts;
  ps.z = ps.q + (size_t) npl * npts;
  ps.rho = ps.z + (size_t) npl * npts;
  ps.eps = ps.rho + npl;
  ps.r2 = ps.eps + npl;
  ps.it = (int*) (ps.r2 + npl);
  ps.active = ps.it + npl;     // 2*npl ints fit in the npl + npl doubles left

  hipLaunchKernelGGL (k_pcg_essential, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                      st, m->nglobal, m->d_ess, lines, lstride_lines, s->d_G, nz,
                      s->base_mode);
  hipLaunchKernelGGL (k_pcg_weak, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl,
                      s->lambda2, s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog,
                      m->d_emask, s->d_G, m->nglobal, d_f, stride, L, lstride, nz,
                      s->base_mode);
  hipLaunchKernelGGL (k_pcg_setup, dim3 (npl), dim3 (PV_THREADS), 0, st, m->nglobal, m->npnp,
                      m->next, m->nel, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw,
                      lines, lstride_lines, s->d_nsolve, m->d_emap, d_u, stride, L, lstride, ps,
                      npts, tol_rel, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  // r = b - A x0: every plane is active for this product
  hipLaunchKernelGGL (k_fill_int, dim3 ((npl + 255) / 256), dim3 (256), 0, st, ps.active, 1, npl);
  hipLaunchKernelGGL (k_pcg_op, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl,
                      s->lambda2, s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog,
                      m->nglobal, ps.x, npts, L, lstride, ps.active, nz, s->base_mode);
  hipLaunchKernelGGL (k_pcg_resid0, dim3 (npl), dim3 (PV_THREADS), 0, st, m->nglobal,
                      m->npnp, m->next, m->d_gptr, m->d_gsrc, s->d_nsolve, L, lstride, ps,
                      npts, step_max, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  std::vector<int> act (npl);
  for (int iter = 1; iter < step_max; iter++) {
    hipLaunchKernelGGL (k_pcg_dir, dim3 (npl), dim3 (PV_THREADS), 0, st, s->d_pc, npts, ps, nz,
                        s->base_mode);
    hipLaunchKernelGGL (k_pcg_op, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl,
                        s->lambda2, s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog,
                        m->nglobal, ps.p, npts, L, lstride, ps.active, nz, s->base_mode);
    hipLaunchKernelGGL (k_pcg_update, dim3 (npl), dim3 (PV_THREADS), 0, st, m->nglobal,
                        m->npnp, m->next, m->d_gptr, m->d_gsrc, s->d_nsolve, L, lstride, ps,
                        npts, step_max, nz, s->base_mode);
    SX_HIP (hipGetLastError ());
    if (iter % 4 == 0 || iter + 1 >= step_max) {
      SX_HIP (hipMemcpyAsync (act.data (), ps.active, sizeof (int) * npl,
                              hipMemcpyDeviceToHost, st));
      SX_HIP (hipStreamSynchronize (st));
      if (std::none_of (act.begin (), act.end (), [] (int a) { return a != 0; })) break;
    }
  }
  hipLaunchKernelGGL (k_pcg_finish, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->d_emap,
                      m->d_btog, m->d_bmask, m->d_emask, s->d_G, m->nglobal, ps.x, npts, d_u,
                      stride, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  if (iters) {
    std::vector<int> its (npl);
    SX_HIP (hipMemcpyAsync (its.data (), ps.it, sizeof (int) * npl, hipMemcpyDeviceToHost, st));
    SX_HIP (hipStreamSynchronize (st));
    *iters = *std::max_element (its.begin (), its.end ());
  }
  return SX_OK;
}
