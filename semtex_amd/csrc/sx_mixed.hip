// sx_mixed.hip -- mixed-precision Field::solve: FP32 elemental static
// condensation and interior recovery (hbi hii^-1 and hii^-1 rounded to
// FP32, FP32 arithmetic), FP64 banded boundary solve, and FP64 residual
// refinement with the elemental Helmholtz operator until the correction is
// below the requested relative size.
//
//   W   = -(M f + H g)                      weak forcing, FP64   (constrain)
//   v_0 = S32(W)                            inner solve
//   R_k = W - H v_k   (FP64, per element; boundary rows summed in the
//                      inner solve's assembly together with <h, w>)
//   v_{k+1} = v_k + S32(R_k)
//   u   = v with the essential values inserted                (setEssential)
//
// S32(X): per element F = X[emap], F_b -= (hbi hii^-1)_32 F_i (FP32);
// direct-stiffness sum + natural BC integrals (FP64); FP64 band sweeps;
// u_i = (hii^-1)_32 F_i - (hbi hii^-1)_32^T u_b (FP32).
#include <algorithm>
#include <cmath>
#include <vector>

#include "sx_internal.h"

using namespace sx;

namespace {

constexpr double EPSDP = 6.0e-14;

__device__ __forceinline__ int solved_plane (int y, int nz, int base_mode)
{
  return (base_mode == 0 && nz > 1 && y >= 1) ? y + 1 : y;
}

// Elemental operator out = H src on one element by one wavefront
// (Element::HelmholtzOp, src/element.cpp:1593-1622).
__device__ void helm_op_elem (int np, const double* D, const double* geo,
                              double lambda2, double betak2, int cyl,
                              const double* src, double* R, double* S,
                              double* out, int lane)
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

// W = -(M f + H g) (Field::constrain, src/field.cpp:812-852), row-major.
__device__ __forceinline__ void k_weak_body (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ emask, const double* __restrict__ G, int nglobal,
    const double* __restrict__ F, double* W, int stride, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1);
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode), lane = threadIdx.x;
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
  for (int i = lane; i < npnp; i += 64) W[off + i] = -buf[i];
}

__global__ __launch_bounds__(64) void k_weak (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ emask, const double* __restrict__ G, int nglobal,
    const double* __restrict__ F, double* W, int stride, int nz, int base_mode)
{
  k_weak_body (np, cyl, lambda2, betak2, geo, Dg, emap_g, btog, emask, G, nglobal, F, W, stride, nz, base_mode);
}

// R = W - H v (FP64), row-major.
__device__ __forceinline__ void k_resid_body (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const double* __restrict__ W, const double* __restrict__ V, double* R_out,
    int stride, int nz, int base_mode)
{
  const int npnp = np * np;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode), lane = threadIdx.x;
  __shared__ double D[256], src[256], R[256], S[256], out[256];
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) { D[i] = Dg[i]; src[i] = V[off + i]; }
  __syncthreads ();
  helm_op_elem (np, D, geo + (size_t) e * NGEO * npnp, lambda2, betak2[k >> 1], cyl, src,
                R, S, out, lane);
  for (int i = lane; i < npnp; i += 64) R_out[off + i] = W[off + i] - out[i];
}

__global__ __launch_bounds__(64) void k_resid (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const double* __restrict__ W, const double* __restrict__ V, double* R_out,
    int stride, int nz, int base_mode)
{
  k_resid_body (np, cyl, lambda2, betak2, geo, Dg, W, V, R_out, stride, nz, base_mode);
}

// Inner solve, part 1: permute to emap order, condense the boundary
// forcing with the FP32 hbi hii^-1.  Fp gets [F_b condensed | F_i].
__device__ __forceinline__ void k_condense32_body (
    int np, int nclass, const int* __restrict__ emap_g, const int* __restrict__ cls,
    const float* __restrict__ Bm, const double* __restrict__ X, double* Fp,
    int stride, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode), lane = threadIdx.x;
  __shared__ float fi[256];
  __shared__ double fp[256];
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  for (int a = lane; a < npnp; a += 64) fp[a] = X[off + emap_g[a]];
  __syncthreads ();
  for (int a = lane; a < nint; a += 64) fi[a] = (float) fp[next + a];
  __syncthreads ();
  const float* B = Bm + ((size_t) (k >> 1) * nclass + cls[e]) * next * nint;
  for (int a = lane; a < next; a += 64) {
    float s = 0.0f;
    for (int q = 0; q < nint; q++) s += B[a * nint + q] * fi[q];
    Fp[off + a] = fp[a] - (double) s;
  }
  for (int a = next + lane; a < npnp; a += 64) Fp[off + a] = fp[a];
}

__global__ __launch_bounds__(64) void k_condense32 (
    int np, int nclass, const int* __restrict__ emap_g, const int* __restrict__ cls,
    const float* __restrict__ Bm, const double* __restrict__ X, double* Fp,
    int stride, int nz, int base_mode)
{
  k_condense32_body (np, nclass, emap_g, cls, Bm, X, Fp, stride, nz, base_mode);
}

// Inner solve, part 3: FP32 interior recovery into row-major V (no
// essential insertion: essential boundary values are zero here).
__device__ __forceinline__ void k_recover32_body (
    int np, int nclass, const int* __restrict__ emap_g, const int* __restrict__ cls,
    const int* __restrict__ btog, const float* __restrict__ Bm,
    const float* __restrict__ Hinv, const double* __restrict__ Fp, double* V,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs,
    int nmax, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1, rr = k & 1, lane = threadIdx.x;
  __shared__ float ub[64], fi[256];
  __shared__ double ubd[64];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) emap[i] = emap_g[i];
  const int* bt = btog + (size_t) e * next;
  const int ns = nsolve[mode];
  const double* x = rhs + ((size_t) mode * 2 + rr) * nmax;
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < next; i += 64) {
    const int gg = bt[i];
    const double v = gg < ns ? x[gg] : 0.0;
    ubd[i] = v;
    ub[i] = (float) v;
  }
  for (int a = lane; a < nint; a += 64) fi[a] = (float) Fp[off + next + a];
  __syncthreads ();
  const size_t mc = (size_t) mode * nclass + cls[e];
  const float* B = Bm + mc * next * nint;
  const float* Hi = Hinv + mc * nint * nint;
  for (int a = lane; a < nint; a += 64) {
    float s = 0.0f;
    for (int b = 0; b < nint; b++) s += Hi[b * nint + a] * fi[b];
    for (int c = 0; c < next; c++) s -= B[c * nint + a] * ub[c];
    V[off + emap[next + a]] = (double) s;
  }
  for (int i = lane; i < next; i += 64) V[off + emap[i]] = ubd[i];
}

__global__ __launch_bounds__(64) void k_recover32 (
    int np, int nclass, const int* __restrict__ emap_g, const int* __restrict__ cls,
    const int* __restrict__ btog, const float* __restrict__ Bm,
    const float* __restrict__ Hinv, const double* __restrict__ Fp, double* V,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs,
    int nmax, int nz, int base_mode)
{
  k_recover32_body (np, nclass, emap_g, cls, btog, Bm, Hinv, Fp, V, stride, nsolve, rhs, nmax, nz, base_mode);
}

__device__ __forceinline__ void k_axpy_planes_body (double* V, const double* dV, int npl, int nz,
                               int base_mode, int n_plane, int stride)
{
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) npl * n_plane) return;
  const int y = (int) (idx / n_plane), j = (int) (idx % n_plane);
  const int k = solved_plane (y, nz, base_mode);
  V[(size_t) k * stride + j] += dV[(size_t) k * stride + j];
}

__global__ void k_axpy_planes (double* V, const double* dV, int npl, int nz,
                               int base_mode, int n_plane, int stride)
{
  k_axpy_planes_body (V, dV, npl, nz, base_mode, n_plane, stride);
}

// sum of squares of V and of dV over the solved planes (partials per block)
__global__ __launch_bounds__(256) void k_norms (const double* V, const double* dV,
                                                int npl, int nz, int base_mode,
                                                int n_plane, int stride, double* part)
{
  __shared__ double sv[256], sd[256];
  double a = 0.0, b = 0.0;
  const size_t tot = (size_t) npl * n_plane;
  for (size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
       idx += (size_t) gridDim.x * blockDim.x) {
    const int y = (int) (idx / n_plane), j = (int) (idx % n_plane);
    const size_t o = (size_t) solved_plane (y, nz, base_mode) * stride + j;
    a += V[o] * V[o];
    b += dV[o] * dV[o];
  }
  sv[threadIdx.x] = a;
  sd[threadIdx.x] = b;
  __syncthreads ();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int) threadIdx.x < s) { sv[threadIdx.x] += sv[threadIdx.x + s]; sd[threadIdx.x] += sd[threadIdx.x + s]; }
    __syncthreads ();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = sv[0]; part[2 * blockIdx.x + 1] = sd[0]; }
}

// u = V with the essential values of G inserted (Field::setEssential).
__device__ __forceinline__ void k_finish_body (
    int np, const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ G, int nglobal, const double* __restrict__ V,
    double* U, int stride, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1);
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode), lane = threadIdx.x;
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) U[off + i] = V[off + i];
  if (!emask[e]) return;
  __syncthreads ();
  for (int i = lane; i < next; i += 64)
    if (bmask[(size_t) e * next + i])
      U[off + emap_g[i]] = G[(size_t) k * nglobal + btog[(size_t) e * next + i]];
}

__global__ __launch_bounds__(64) void k_finish (
    int np, const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ G, int nglobal, const double* __restrict__ V,
    double* U, int stride, int nz, int base_mode)
{
  k_finish_body (np, emap_g, btog, bmask, emask, G, nglobal, V, U, stride, nz, base_mode);
}

__device__ __forceinline__ void k_essential_m_body (int nglobal, const int* __restrict__ ess,
                               const double* __restrict__ lines, int lstride,
                               double* G, int nz, int base_mode)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  if (g >= nglobal) return;
  const int s = ess[g];
  G[(size_t) k * nglobal + g] = (s >= 0 && lines) ? lines[(size_t) k * lstride + s] : 0.0;
}

__global__ void k_essential_m (int nglobal, const int* __restrict__ ess,
                               const double* __restrict__ lines, int lstride,
                               double* G, int nz, int base_mode)
{
  k_essential_m_body (nglobal, ess, lines, lstride, G, nz, base_mode);
}

__device__ __forceinline__ void k_assemble_m_body (int nglobal, const int* __restrict__ gptr,
                              const int* __restrict__ gsrc, const int* __restrict__ nptr,
                              const int* __restrict__ nidx, const double* __restrict__ nw,
                              const double* __restrict__ F, int stride,
                              const double* __restrict__ lines, int lstride,
                              const int* __restrict__ nsolve, double* rhs, int nmax,
                              int nz, int base_mode)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1, rr = k & 1;
  if (g >= nsolve[mode]) return;
  const double* Fk = F + (size_t) k * stride;
  double s = 0.0;
  for (int c = gptr[g]; c < gptr[g + 1]; c++) s += Fk[gsrc[c]];
  if (lines)
    for (int c = nptr[g]; c < nptr[g + 1]; c++) s += lines[(size_t) k * lstride + nidx[c]] * nw[c];
  rhs[((size_t) mode * 2 + rr) * nmax + g] = s;
}

__global__ void k_assemble_m (int nglobal, const int* __restrict__ gptr,
                              const int* __restrict__ gsrc, const int* __restrict__ nptr,
                              const int* __restrict__ nidx, const double* __restrict__ nw,
                              const double* __restrict__ F, int stride,
                              const double* __restrict__ lines, int lstride,
                              const int* __restrict__ nsolve, double* rhs, int nmax,
                              int nz, int base_mode)
{
  k_assemble_m_body (nglobal, gptr, gsrc, nptr, nidx, nw, F, stride, lines, lstride, nsolve, rhs, nmax, nz, base_mode);
}

int inner_solve (sx_msys* s, int nz, int stride, const double* X, double* Fp,
                 double* V, const double* lines, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const int lstride = m->nbound * m->np;
  if (s->d_Mcon32) {
    if (int rc = condense32_tiled (s, nz, stride, X, Fp, st)) return rc;
  } else {
    hipLaunchKernelGGL (k_condense32, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->nclass,
                        m->d_emap, m->d_cls, s->d_Bm32, X, Fp, stride, nz, s->base_mode);
  }
  hipLaunchKernelGGL (k_assemble_m, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0, st,
                      m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw, Fp,
                      stride, lines, lstride, s->d_nsolve, s->d_rhs, s->nmax, nz,
                      s->base_mode);
  SX_HIP (hipGetLastError ());
  int rc = band_solve (s, 2, st);
  if (rc) return rc;
  if (s->d_Mrec32) {
    if ((rc = recover32_tiled (s, nz, stride, Fp, V, st))) return rc;
  } else {
    hipLaunchKernelGGL (k_recover32, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->nclass,
                        m->d_emap, m->d_cls, m->d_btog, s->d_Bm32, s->d_Hinv32, Fp, V, stride,
                        s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode);
  }
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace

namespace sx {

}  // namespace sx

namespace {
#define SX_BODY_FN(Name, body)                                                   \
  struct Name {                                                                 \
    template <typename... A> __device__ __forceinline__ void operator() (A... a) const \
    { body (a...); }                                                            \
  };
SX_BODY_FN (FnEssentialM, k_essential_m_body)
SX_BODY_FN (FnWeak, k_weak_body)
SX_BODY_FN (FnAssembleM, k_assemble_m_body)
SX_BODY_FN (FnResid, k_resid_body)
SX_BODY_FN (FnAxpyPlanes, k_axpy_planes_body)
SX_BODY_FN (FnFinish, k_finish_body)
SX_BODY_FN (FnCondense32, k_condense32_body)
SX_BODY_FN (FnRecover32, k_recover32_body)
#undef SX_BODY_FN

// the inner solve of the mixed passes, recorded (FP32 MFMA tiles only)
int plan_inner (BatchPlan& P, sx_msys* s, int nz, int stride, const double* X, double* Fp,
                double* V, const double* lines)
{
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const int lstride = m->nbound * m->np;
  const bool tiled = s->d_Mcon32 && s->d_Mrec32;
  if (tiled) {
    if (int rc = plan_condense32_tiled (P, s, nz, stride, X, Fp)) return rc;
  } else {
    batch_rec<64, FnCondense32> (P, dim3 (m->nel, npl), 64, 0, m->np, m->nclass,
                                 (const int*) m->d_emap, (const int*) m->d_cls,
                                 (const float*) s->d_Bm32, X, Fp, stride, nz, s->base_mode);
    P.phase++;
  }
  batch_rec<256, FnAssembleM> (P, dim3 ((m->nglobal + 255) / 256, npl), 256, 0, m->nglobal,
                               (const int*) m->d_gptr, (const int*) m->d_gsrc,
                               (const int*) m->d_nptr, (const int*) m->d_nidx,
                               (const double*) m->d_nw, (const double*) Fp, stride, lines,
                               lstride, (const int*) s->d_nsolve, s->d_rhs, s->nmax, nz,
                               s->base_mode);
  P.phase++;
  if (int rc = plan_band_solve (P, s, 2)) return rc;
  if (tiled) return plan_recover32_tiled (P, s, nz, stride, Fp, V);
  batch_rec<64, FnRecover32> (P, dim3 (m->nel, npl), 64, 0, m->np, m->nclass,
                              (const int*) m->d_emap, (const int*) m->d_cls,
                              (const int*) m->d_btog, (const float*) s->d_Bm32,
                              (const float*) s->d_Hinv32, (const double*) Fp, V, stride,
                              (const int*) s->d_nsolve, (const double*) s->d_rhs, s->nmax, nz,
                              s->base_mode);
  P.phase++;
  return SX_OK;
}
}  // namespace

namespace sx {

// field_solve_mixed_passes, recorded for a fused batch; the system's
// workspace is allocated here (planning is set-up, not the replayed work)
int plan_mixed_passes (BatchPlan& P, sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                       const double* d_lines, int passes)
{
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const size_t nplane = (size_t) nz * stride;
  const size_t need = mixed_work_doubles (s, nz, stride);
  if (s->work_doubles < need) {
    if (s->d_work) (void) hipFree (s->d_work);
    s->d_work = nullptr;
    s->work_doubles = 0;
    SX_HIP (hipMalloc (&s->d_work, sizeof (double) * need));
    s->work_doubles = need;
  }
  double *W = s->d_work, *X = W + nplane, *Fp = X + nplane, *V = Fp + nplane,
         *dV = V + nplane;
  const int lstride = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : (const double*) nullptr;
  batch_rec<256, FnEssentialM> (P, dim3 ((m->nglobal + 255) / 256, npl), 256, 0, m->nglobal,
                                (const int*) m->d_ess, lines, lstride, s->d_G, nz, s->base_mode);
  P.phase++;
  batch_rec<64, FnWeak> (P, dim3 (m->nel, npl), 64, 0, m->np, m->cyl, s->lambda2,
                         (const double*) s->d_betak2, (const double*) m->d_geo,
                         (const double*) m->d_D, (const int*) m->d_emap, (const int*) m->d_btog,
                         (const int*) m->d_emask, (const double*) s->d_G, m->nglobal,
                         (const double*) d_f, W, stride, nz, s->base_mode);
  P.phase++;
  if (int rc = plan_inner (P, s, nz, stride, W, Fp, V, lines)) return rc;
  const int n_plane = m->nel * m->npnp;
  for (int it = 1; it < passes; it++) {
    batch_rec<64, FnResid> (P, dim3 (m->nel, npl), 64, 0, m->np, m->cyl, s->lambda2,
                            (const double*) s->d_betak2, (const double*) m->d_geo,
                            (const double*) m->d_D, (const double*) W, (const double*) V, X,
                            stride, nz, s->base_mode);
    P.phase++;
    if (int rc = plan_inner (P, s, nz, stride, X, Fp, dV, lines)) return rc;
    batch_rec<256, FnAxpyPlanes> (P, dim3 ((unsigned) (((size_t) npl * n_plane + 255) / 256)),
                                  256, 0, V, (const double*) dV, npl, nz, s->base_mode, n_plane,
                                  stride);
    P.phase++;
  }
  batch_rec<64, FnFinish> (P, dim3 (m->nel, npl), 64, 0, m->np, (const int*) m->d_emap,
                           (const int*) m->d_btog, (const int*) m->d_bmask,
                           (const int*) m->d_emask, (const double*) s->d_G, m->nglobal,
                           (const double*) V, d_u, stride, nz, s->base_mode);
  P.phase++;
  return SX_OK;
}

// The mixed solve with a fixed number of refinement passes and no host
// synchronisation (graph capture): the first pass is the FP32-condensed
// solve, each further pass adds the solve of the FP64 residual.
int field_solve_mixed_passes (sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                              const double* d_lines, int passes, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const size_t nplane = (size_t) nz * stride;
  const size_t need = mixed_work_doubles (s, nz, stride);
  if (s->work_doubles < need) {
    if (s->d_work) (void) hipFree (s->d_work);
    s->d_work = nullptr;
    s->work_doubles = 0;
    SX_HIP (hipMalloc (&s->d_work, sizeof (double) * need));
    s->work_doubles = need;
  }
  double *W = s->d_work, *X = W + nplane, *Fp = X + nplane, *V = Fp + nplane,
         *dV = V + nplane;
  const int lstride = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : nullptr;
  hipLaunchKernelGGL (k_essential_m, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0, st,
                      m->nglobal, m->d_ess, lines, lstride, s->d_G, nz, s->base_mode);
  hipLaunchKernelGGL (k_weak, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl, s->lambda2,
                      s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog, m->d_emask, s->d_G,
                      m->nglobal, d_f, W, stride, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  int rc = inner_solve (s, nz, stride, W, Fp, V, lines, st);
  if (rc) return rc;
  const int n_plane = m->nel * m->npnp;
  for (int it = 1; it < passes; it++) {
    hipLaunchKernelGGL (k_resid, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl,
                        s->lambda2, s->d_betak2, m->d_geo, m->d_D, W, V, X, stride, nz,
                        s->base_mode);
    SX_HIP (hipGetLastError ());
    if ((rc = inner_solve (s, nz, stride, X, Fp, dV, lines, st))) return rc;
    hipLaunchKernelGGL (k_axpy_planes, dim3 ((unsigned) (((size_t) npl * n_plane + 255) / 256)),
                        dim3 (256), 0, st, V, dV, npl, nz, s->base_mode, n_plane, stride);
    SX_HIP (hipGetLastError ());
  }
  hipLaunchKernelGGL (k_finish, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->d_emap,
                      m->d_btog, m->d_bmask, m->d_emask, s->d_G, m->nglobal, V, d_u, stride,
                      nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace sx

namespace sx {

size_t mixed_work_doubles (const sx_msys* s, int nz, int stride)
{
  (void) s;
  return 5 * (size_t) nz * stride + 4 * MIXED_NBLK;
}

// The mixed solve of sx_field_solve_mixed on a caller-provided workspace of
// at least mixed_work_doubles() doubles: W, X, Fp, V, dV and the norm
// partials.  Every kernel writes the entries of the solved planes it reads
// later, so the workspace needs no clearing.
int field_solve_mixed_ws (sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                          const double* d_lines, double refine_tol, int max_iter, int* iters,
                          double* work, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const int npl = (s->base_mode == 0 && nz > 1) ? nz - 1 : nz;
  const size_t nplane = (size_t) nz * stride;
  const double tol = refine_tol > 0.0 ? refine_tol : 1e-12;
  double *W = work, *X = W + nplane, *Fp = X + nplane, *V = Fp + nplane,
         *dV = V + nplane, *part = dV + nplane;
  const int lstride = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : nullptr;
  hipLaunchKernelGGL (k_essential_m, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0, st,
                      m->nglobal, m->d_ess, lines, lstride, s->d_G, nz, s->base_mode);
  hipLaunchKernelGGL (k_weak, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl, s->lambda2,
                      s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog, m->d_emask, s->d_G,
                      m->nglobal, d_f, W, stride, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  int rc = inner_solve (s, nz, stride, W, Fp, V, lines, st);
  if (rc) return rc;
  const int n_plane = m->nel * m->npnp;
  double hp[2 * MIXED_NBLK];
  int it = 0;
  for (; it < max_iter; it++) {
    hipLaunchKernelGGL (k_resid, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->cyl,
                        s->lambda2, s->d_betak2, m->d_geo, m->d_D, W, V, X, stride, nz,
                        s->base_mode);
    SX_HIP (hipGetLastError ());
    if ((rc = inner_solve (s, nz, stride, X, Fp, dV, lines, st))) return rc;
    hipLaunchKernelGGL (k_norms, dim3 (MIXED_NBLK), dim3 (256), 0, st, V, dV, npl, nz,
                        s->base_mode, n_plane, stride, part);
    hipLaunchKernelGGL (k_axpy_planes, dim3 ((unsigned) (((size_t) npl * n_plane + 255) / 256)),
                        dim3 (256), 0, st, V, dV, npl, nz, s->base_mode, n_plane, stride);
    SX_HIP (hipGetLastError ());
    SX_HIP (hipMemcpyAsync (hp, part, sizeof (double) * 2 * MIXED_NBLK, hipMemcpyDeviceToHost, st));
    SX_HIP (hipStreamSynchronize (st));
    double nv = 0.0, nd = 0.0;
    for (int b = 0; b < MIXED_NBLK; b++) { nv += hp[2 * b]; nd += hp[2 * b + 1]; }
    if (nd <= tol * tol * std::max (nv, 1e-300)) { it++; break; }
  }
  hipLaunchKernelGGL (k_finish, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->d_emap,
                      m->d_btog, m->d_bmask, m->d_emask, s->d_G, m->nglobal, V, d_u, stride,
                      nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  if (iters) *iters = it;
  return SX_OK;
}

int check_mixed_solve_args (sx_msys* s, int nz, int stride, const void* f, const void* u,
                            int max_iter, const char* who)
{
  if (!s || !f || !u || nz < 1 || stride < s->m->nel * s->m->npnp ||
      (nz > 2 * s->nmodes && nz != 1) || max_iter < 0) {
    set_error (std::string (who) + ": bad arguments");
    return SX_ERR_ARG;
  }
  if (s->precision != SX_MIXED || !s->d_Bm32) {
    set_error (std::string (who) + ": system was not created with SX_MIXED");
    return SX_ERR_ARG;
  }
  return SX_OK;
}

}  // namespace sx

extern "C" int sx_field_solve_mixed (sx_msys_t s, int nz, int stride, double* d_f,
                                     double* d_u, const double* d_lines,
                                     double refine_tol, int max_iter, int* iters,
                                     void* stream)
{
  if (int rc = check_mixed_solve_args (s, nz, stride, d_f, d_u, max_iter, "sx_field_solve_mixed"))
    return rc;
  // workspace: W, X, Fp, V, dV (+ norm partials), kept with the system
  const size_t need = mixed_work_doubles (s, nz, stride);
  if (s->work_doubles < need) {
    if (s->d_work) (void) hipFree (s->d_work);
    s->d_work = nullptr;
    s->work_doubles = 0;
    SX_HIP (hipMalloc (&s->d_work, sizeof (double) * need));
    s->work_doubles = need;
  }
  return field_solve_mixed_ws (s, nz, stride, d_f, d_u, d_lines, refine_tol, max_iter, iters,
                               s->d_work, (hipStream_t) stream);
}
