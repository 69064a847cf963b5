// sx_solver.hip -- gfx950 kernels and C ABI of the per-Fourier-mode
// Helmholtz solver (semtex ModalMatrixSys + Field::solve, DIRECT).
//
// Set-up (once per system, not timed):
//   k_elem_sc       elemental Helmholtz matrix + static condensation per
//                   (mode, element class)  [Element::HelmholtzSC]
//   k_band_assemble deterministic direct-stiffness assembly of the band
//   band_factor     FP64 banded Cholesky + inverted diagonal blocks (sx_band.hip)
// Solve (the timed path, per call):
//   k_essential     essential BC values into global numbering (k_ess_weak:
//                   with the essential elements' -(M f + H g) in one launch)
//   k_condense      -(M f + H g), emap gather, F_b -= hbi hii^-1 F_i
//   k_assemble      direct-stiffness sum of F_b (+ natural BC integrals)
//   band_solve      FP64 U^T U x = b, one workgroup per mode (sx_band.hip)
//   k_recover       u_i = hii^-1 F_i - (hbi hii^-1)^T u_b, essential insert
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sx_internal.h"

using namespace sx;

namespace {

constexpr double EPSDP = 6.0e-14;   // femlib/cfemdef.h
constexpr double EPSSP = 6.0e-7;

__device__ __forceinline__ double wave_sum (double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off, 64);
  return v;
}

__device__ __forceinline__ double readlane_d (double v, int l)
{
  const long long x = __double_as_longlong (v);
  const int lo = __builtin_amdgcn_readlane ((int) (x & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane ((int) (x >> 32), l);
  return __longlong_as_double (((long long) hi << 32) | (unsigned) lo);
}

// Map the y-th solved plane to its plane index: plane 1 (Nyquist of mode
// 0) is skipped when the base mode is 0 and there is more than one plane
// (src/field.cpp:650).
__device__ __forceinline__ int solved_plane (int y, int nz, int base_mode)
{
  return (base_mode == 0 && nz > 1 && y >= 1) ? y + 1 : y;
}

// ---------------------------------------------------------------------------
// Element matrix entry h[r][c] (r, c row-major node indices), following
// Element::HelmholtzRow (src/element.cpp:1411-1481), SVV off.
// ---------------------------------------------------------------------------
__device__ double helm_entry (int np, int r, int c, const double* Q1,
                              const double* Q2, const double* Q3,
                              const double* Q4, const double* Y,
                              const double* D, double lambda2, double betak2,
                              int cyl, bool hasQ3)
{
  const int i = r / np, j = r % np, m = c / np, n = c % np;
  double h = 0.0;
  if (i == m) {
    double s = 0.0;
    for (int q = 0; q < np; q++) s += Q1[i * np + q] * (D[q * np + j] * D[q * np + n]);
    h += s;
  }
  if (j == n) {
    double s = 0.0;
    for (int p = 0; p < np; p++) s += Q2[p * np + j] * (D[p * np + i] * D[p * np + m]);
    h += s;
  }
  if (hasQ3) {
    h += Q3[i * np + n] * D[n * np + j] * D[i * np + m];
    h += Q3[m * np + j] * D[j * np + n] * D[m * np + i];
  }
  if (r == c) {
    const double r2 = Y[r] * Y[r];
    const double hCon = (cyl && r2 > EPSDP) ? betak2 / r2 + lambda2 : betak2 + lambda2;
    h += Q4[r] * hCon;
  }
  return h;
}

// One workgroup per (mode, class).  ws: per-block workspace of
// npnp*npnp + nint*nint doubles.  Outputs: S (condensed hbb, next^2),
// Bm = hbi inv(hii) (next x nint), Hinv = inv(hii) (nint^2).
__global__ __launch_bounds__(256) void k_elem_sc (
    int np, int nclass, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ cgeo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, double* ws, double* S, double* Bm,
    double* Hinv, int* status)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int blk = blockIdx.x, mode = blk / nclass, cls = blk % nclass;
  const int tid = threadIdx.x, nt = blockDim.x;
  __shared__ double D[256], G[5 * 256];
  __shared__ int emap[256];
  __shared__ int s_bad, s_hasQ3;
  __shared__ double s_piv;

  const double* geo = cgeo + (size_t) cls * NGEO * npnp;
  for (int k = tid; k < npnp; k += nt) {
    D[k] = Dg[k];
    emap[k] = emap_g[k];
    G[0 * npnp + k] = geo[G_Q1 * npnp + k];
    G[1 * npnp + k] = geo[G_Q2 * npnp + k];
    G[2 * npnp + k] = geo[G_Q3 * npnp + k];
    G[3 * npnp + k] = geo[G_Q4 * npnp + k];
    G[4 * npnp + k] = geo[G_Y * npnp + k];
  }
  if (tid == 0) { s_bad = 0; s_hasQ3 = 0; }
  __syncthreads ();
  for (int k = tid; k < npnp; k += nt)
    if (G[2 * npnp + k] != 0.0) s_hasQ3 = 1;
  __syncthreads ();
  const bool hasQ3 = s_hasQ3 != 0;
  const double bk2 = betak2[mode];

  double* H = ws + (size_t) blk * (npnp * npnp + nint * nint);  // permuted
  double* Li = H + npnp * npnp;                                  // inv(L)
  for (int idx = tid; idx < npnp * npnp; idx += nt) {
    const int a = idx / npnp, b = idx % npnp;
    H[idx] = helm_entry (np, emap[a], emap[b], G, G + npnp, G + 2 * npnp,
                         G + 3 * npnp, G + 4 * npnp, D, lambda2, bk2, cyl, hasQ3);
  }
  __syncthreads ();

  // Cholesky of the interior block hii (lower, in place)
#define HII(a, b) H[(size_t) (next + (a)) * npnp + next + (b)]
  for (int k = 0; k < nint; k++) {
    if (tid == 0) {
      const double p = HII (k, k);
      if (!(p > 0.0)) s_bad = 1;
      s_piv = std::sqrt (p > 0.0 ? p : 1.0);
      HII (k, k) = s_piv;
    }
    __syncthreads ();
    const double piv = s_piv;
    for (int i = k + 1 + tid; i < nint; i += nt) HII (i, k) /= piv;
    __syncthreads ();
    const int w = nint - k - 1;
    for (int idx = tid; idx < w * w; idx += nt) {
      const int ii = idx / w, jj = idx % w;
      if (jj <= ii) HII (k + 1 + ii, k + 1 + jj) -= HII (k + 1 + ii, k) * HII (k + 1 + jj, k);
    }
    __syncthreads ();
  }
  if (s_bad) { if (tid == 0) atomicExch (status, 1); return; }

  // inverse of L, column by column
  for (int c = tid; c < nint; c += nt) {
    for (int i = 0; i < c; i++) Li[i * nint + c] = 0.0;
    Li[c * nint + c] = 1.0 / HII (c, c);
    for (int i = c + 1; i < nint; i++) {
      double s = 0.0;
      for (int k = c; k < i; k++) s += HII (i, k) * Li[k * nint + c];
      Li[i * nint + c] = -s / HII (i, i);
    }
  }
  __syncthreads ();
  double* Hi = Hinv + ((size_t) mode * nclass + cls) * nint * nint;
  for (int idx = tid; idx < nint * nint; idx += nt) {
    const int a = idx / nint, b = idx % nint;
    double s = 0.0;
    for (int k = max (a, b); k < nint; k++) s += Li[k * nint + a] * Li[k * nint + b];
    Hi[idx] = s;
  }
  __syncthreads ();
  double* B = Bm + ((size_t) mode * nclass + cls) * next * nint;
  for (int idx = tid; idx < next * nint; idx += nt) {
    const int a = idx / nint, k = idx % nint;
    double s = 0.0;
    for (int b = 0; b < nint; b++) s += H[(size_t) a * npnp + next + b] * Hi[b * nint + k];
    B[idx] = s;
  }
  __syncthreads ();
  double* So = S + ((size_t) mode * nclass + cls) * next * next;
  for (int idx = tid; idx < next * next; idx += nt) {
    const int a = idx / next, b = idx % next;
    double s = 0.0;
    for (int k = 0; k < nint; k++) s += B[a * nint + k] * H[(size_t) b * npnp + next + k];
    So[idx] = H[(size_t) a * npnp + b] - s;
  }
#undef HII
}

__global__ void k_band_assemble (int npairs, const long long* __restrict__ ppos,
                                 const int* __restrict__ pmn,
                                 const int* __restrict__ pptr,
                                 const int* __restrict__ psrc,
                                 const double* __restrict__ S, size_t s_stride,
                                 double* band, size_t band_stride,
                                 const int* __restrict__ nsolve)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x, mode = blockIdx.y;
  if (p >= npairs) return;
  if (pmn[2 * p + 1] >= nsolve[mode]) return;
  const double* Sm = S + (size_t) mode * s_stride;
  double s = 0.0;
  for (int c = pptr[p]; c < pptr[p + 1]; c++) s += Sm[psrc[c]];
  band[(size_t) mode * band_stride + ppos[p]] = s;
}

// ---------------------------------------------------------------------------
// Solve path
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_essential_body (int nglobal, const int* __restrict__ ess,
                             const double* __restrict__ lines, int lstride,
                             double* G, int nz, int base_mode)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  if (g >= nglobal) return;
  const int s = ess[g];
  G[(size_t) k * nglobal + g] = (s >= 0 && lines) ? lines[(size_t) k * lstride + s] : 0.0;
}

__global__ void k_essential (int nglobal, const int* __restrict__ ess,
                             const double* __restrict__ lines, int lstride,
                             double* G, int nz, int base_mode)
{
  k_essential_body (nglobal, ess, lines, lstride, G, nz, base_mode);
}

// Elemental Helmholtz operator out = H src on one element by one
// wavefront (Element::HelmholtzOp/HelmholtzKern, src/element.cpp:1536-1622).
// src, R, S, out are LDS arrays of npnp; geo points to the element's
// NGEO arrays in global memory.
__device__ void helm_op_wave (int np, const double* D, const double* geo,
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

// One wavefront (= one block of 64) per (element, solved plane).
__global__ __launch_bounds__(64) void k_condense (
    int np, int nel, int nclass, int cyl, double lambda2,
    const double* __restrict__ betak2, const double* __restrict__ geo,
    const double* __restrict__ Dg, const int* __restrict__ emap_g,
    const int* __restrict__ cls, const int* __restrict__ btog,
    const int* __restrict__ emask, const double* __restrict__ Bm,
    const double* __restrict__ G, int nglobal, double* F, int stride, int nz,
    int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1, lane = threadIdx.x;
  __shared__ double D[256], buf[256], g[256], R[256], S[256], Hg[256];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) { D[i] = Dg[i]; emap[i] = emap_g[i]; }
  const double* ge = geo + (size_t) e * NGEO * npnp;
  double* f = F + (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) buf[i] = f[i] * ge[G_Q8 * npnp + i];
  __syncthreads ();
  if (emask[e]) {
    for (int i = lane; i < npnp; i += 64) g[i] = 0.0;
    __syncthreads ();
    const int* bt = btog + (size_t) e * next;
    for (int i = lane; i < next; i += 64) g[emap[i]] = G[(size_t) k * nglobal + bt[i]];
    __syncthreads ();
    helm_op_wave (np, D, ge, lambda2, betak2[mode], cyl, g, R, S, Hg, lane);
    for (int i = lane; i < npnp; i += 64) buf[i] += Hg[i];
    __syncthreads ();
  }
  // permuted, negated forcing: F = -(M f + H g) in emap order
  for (int a = lane; a < npnp; a += 64) g[a] = -buf[emap[a]];
  __syncthreads ();
  const double* B = Bm + ((size_t) mode * nclass + cls[e]) * next * nint;
  for (int a = lane; a < next; a += 64) {
    double s = 0.0;
    for (int q = 0; q < nint; q++) s += B[a * nint + q] * g[next + q];
    f[a] = g[a] - s;
  }
  for (int a = next + lane; a < npnp; a += 64) f[a] = g[a];
}

// ---------------------------------------------------------------------------
// Batched condensation / recovery: one workgroup per (tile of same-class
// elements, Fourier mode).  The mode's condensed matrices of that class
// are staged in LDS once and applied to all element-planes of the tile
// (both planes of the mode) as one product: the element-planes are the
// columns of an LDS operand [k][CG_EP], the matrix rows [k][a] are padded
// to 4, and each thread accumulates a 4 (rows) x 2 (element-planes) tile
// from 16-byte LDS reads -- 8 FMAs per 3 LDS instructions.
// ---------------------------------------------------------------------------
constexpr int CT_THREADS = 256;
constexpr int CG_EP = 34;          // element-plane stride in LDS (>= 2 ETILE, bank-shifted)

__host__ __device__ constexpr int pad4 (int x) { return (x + 3) & ~3; }
__host__ __device__ constexpr int pad16 (int x) { return (x + 15) & ~15; }

// Operand layout of the tiled kernels.  MF (FP64 MFMA 16x16x4 tiles): rows
// of the matrix operand padded to 16 columns and its k range to a multiple
// of 4 (zeros); vector path: columns padded to 4.
//   recovery      M[k][a], k < KR(npnp), a < NA(nint)
//   condensation  M[q][a], q < KR(nint), a < NA(next)
struct TileShape {
  int NA, KR;
};
__host__ __device__ inline TileShape rec_shape (int np, bool mf)
{
  const int npnp = np * np, nint = npnp - 4 * (np - 1);
  return mf ? TileShape {pad16 (nint), pad4 (npnp)} : TileShape {pad4 (nint), npnp};
}
__host__ __device__ inline TileShape con_shape (int np, bool mf)
{
  const int next = 4 * (np - 1), nint = np * np - next;
  return mf ? TileShape {pad16 (next), pad4 (nint)} : TileShape {pad4 (next), nint};
}
typedef double v4d __attribute__ ((ext_vector_type (4)));

// v_mfma_f64_16x16x4f64, D += A B for one 16x16 output tile: lane l
// supplies A[l%16][k + l/16] and B[k + l/16][l%16]; D[4r + l/16][l%16] is
// component r of lane l.  The products are formed in chunks of MF_U k-steps
// whose operands are read before their products, alternating between two
// accumulators (mfma_strip), so the operand latency and the MFMA latency
// overlap.
constexpr int MF_U = 8;

// dst(i) <- src(i) for i < n with SU loads in flight per thread (the loads
// of a batch are issued before any of them is waited for)
constexpr int SU = 8;
template <typename Load, typename Store>
__device__ __forceinline__ void stage (int n, Load load, Store store)
{
  for (int i0 = threadIdx.x; i0 < n; i0 += SU * CT_THREADS) {
    double v[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int i = i0 + u * CT_THREADS;
      v[u] = i < n ? load (i) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int i = i0 + u * CT_THREADS;
      if (i < n) store (i, v[u]);
    }
  }
}

// dst[i] = src[i] for i < n (n even, both 16-byte aligned): 16-byte loads,
// sixteen in flight per thread
__device__ __forceinline__ void copy16 (double* dst, const double* __restrict__ src, int n)
{
  constexpr int CU = 16;
  const double2* s2 = (const double2*) src;
  double2* d2 = (double2*) dst;
  const int n2 = n / 2;
  for (int i0 = threadIdx.x; i0 < n2; i0 += CU * CT_THREADS) {
    double2 v[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const int i = i0 + u * CT_THREADS;
      v[u] = i < n2 ? s2[i] : make_double2 (0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const int i = i0 + u * CT_THREADS;
      if (i < n2) d2[i] = v[u];
    }
  }
}

// The same product in FP32 (v_mfma_f32_16x16x4f32, FP32 accumulation):
// operands as for mfma_tile, but D[4 (l/16) + r][l%16] is component r of
// lane l
typedef float v4f __attribute__ ((ext_vector_type (4)));
__device__ __forceinline__ v4f mfma_tile32 (const float* __restrict__ M, int NA,
                                            const float* __restrict__ X, int K, int lane)
{
  v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
  const float* A = M + (lane >> 4) * NA + (lane & 15);
  const float* B = X + (lane >> 4) * CG_EP + (lane & 15);
  int k = 0;
  for (; k + 4 * MF_U <= K; k += 4 * MF_U) {
    float a[MF_U], b[MF_U];
#pragma unroll
    for (int u = 0; u < MF_U; u++) {
      a[u] = A[(size_t) (k + 4 * u) * NA];
      b[u] = B[(size_t) (k + 4 * u) * CG_EP];
    }
#pragma unroll
    for (int u = 0; u < MF_U; u += 2) {
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32 (a[u], b[u], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32 (a[u + 1], b[u + 1], acc1, 0, 0, 0);
    }
  }
  for (; k < K; k += 4)
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32 (A[(size_t) k * NA], B[(size_t) k * CG_EP], acc0,
                                                 0, 0, 0);
  return acc0 + acc1;
}

// dst[i] = src[i] for i < n floats (n a multiple of 4, 16-byte aligned)
__device__ __forceinline__ void copy16f (float* dst, const float* __restrict__ src, int n)
{
  copy16 ((double*) dst, (const double*) src, n / 2);
}

// The tiled kernels' operands per (mode, class), in their LDS layout
// (TileShape; padding zero): Mrec[k][a] = -Bm[k][a] (k < next),
// Hinv[k-next][a] (next <= k < npnp); Mcon[q][a] = Bm[a][q]
__global__ void k_tiled_ops (int np, int nmc, int mf, const double* __restrict__ Bm,
                             const double* __restrict__ Hinv, double* Mrec, double* Mcon)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape R = rec_shape (np, mf), Cs = con_shape (np, mf);
  const size_t nr = (size_t) R.KR * R.NA, nc = (size_t) Cs.KR * Cs.NA;
  for (size_t mc = blockIdx.y; mc < (size_t) nmc; mc += gridDim.y)
  for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < nr + nc;
       i += (size_t) gridDim.x * blockDim.x) {
    if (i < nr) {
      const int k = (int) (i / R.NA), a = (int) (i % R.NA);
      double v = 0.0;
      if (a < nint && k < npnp)
        v = k < next ? -Bm[(mc * next + k) * nint + a] : Hinv[(mc * nint + k - next) * nint + a];
      Mrec[mc * nr + i] = v;
    } else {
      const size_t j = i - nr;
      const int q = (int) (j / Cs.NA), a = (int) (j % Cs.NA);
      Mcon[mc * nc + j] = (a < next && q < nint) ? Bm[(mc * next + a) * nint + q] : 0.0;
    }
  }
}

// acc[r][j] += sum_k A[k][a0 + r] X[k][e0 + j], k < K (A rows of stride NA)
__device__ __forceinline__ void tile_4x2 (const double* __restrict__ A, int NA,
                                          const double* __restrict__ X, int K,
                                          double (&acc)[4][2])
{
  for (int k = 0; k < K; k++) {
    const double2 a01 = *(const double2*) (A + (size_t) k * NA);
    const double2 a23 = *(const double2*) (A + (size_t) k * NA + 2);
    const double2 x = *(const double2*) (X + (size_t) k * CG_EP);
    acc[0][0] += a01.x * x.x; acc[0][1] += a01.x * x.y;
    acc[1][0] += a01.y * x.x; acc[1][1] += a01.y * x.y;
    acc[2][0] += a23.x * x.x; acc[2][1] += a23.x * x.y;
    acc[3][0] += a23.y * x.x; acc[3][1] += a23.y * x.y;
  }
}

// the solved planes of mode m: k0 and their count (1 or 2)
__device__ __forceinline__ int mode_planes (int m, int nz, int base_mode, int* k0)
{
  if (nz == 1 || (base_mode == 0 && m == 0)) { *k0 = 0; return 1; }
  *k0 = 2 * m;
  return 2;
}

// Constrain on the elements with essential nodes (Field::constrain,
// src/field.cpp:812-852): f <- -(M f + H g), row-major, in place.
__device__ __forceinline__ void k_weak_ess_body (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ ess_elem, const double* __restrict__ G, int nglobal,
    double* F, int stride, int nz, int base_mode)
{
  const int npnp = np * np, next = 4 * (np - 1);
  const int e = ess_elem[blockIdx.x], k = solved_plane (blockIdx.y, nz, base_mode);
  const int lane = threadIdx.x;
  __shared__ double D[256], g[256], R[256], S[256], Hg[256];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) { D[i] = Dg[i]; emap[i] = emap_g[i]; g[i] = 0.0; }
  __syncthreads ();
  const int* bt = btog + (size_t) e * next;
  for (int i = lane; i < next; i += 64) g[emap[i]] = G[(size_t) k * nglobal + bt[i]];
  __syncthreads ();
  const double* ge = geo + (size_t) e * NGEO * npnp;
  helm_op_wave (np, D, ge, lambda2, betak2[k >> 1], cyl, g, R, S, Hg, lane);
  double* f = F + (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) f[i] = -(f[i] * ge[G_Q8 * npnp + i] + Hg[i]);
}

__global__ __launch_bounds__(64) void k_weak_ess (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ ess_elem, const double* __restrict__ G, int nglobal,
    double* F, int stride, int nz, int base_mode)
{
  k_weak_ess_body (np, cyl, lambda2, betak2, geo, Dg, emap_g, btog, ess_elem, G, nglobal, F, stride, nz, base_mode);
}

// boundary lines, essential values and field of up to three fields
// (blockIdx.z selects)
struct FieldSet {
  const double* lines[3];
  double* G[3];
  double* F[3];
};

// k_essential and k_weak_ess in one launch, the weak form reading the
// essential values from the boundary lines itself (the value k_essential
// writes to G: lines[ess[g]], 0 when the node takes none; ess_slot holds
// ess[g] for the essential elements' boundary slots): blocks
// [0, ness) form -(M f + H g) of one essential element, the rest fill G,
// 256 global nodes each.  blockIdx.y: plane, blockIdx.z: field.
__global__ __launch_bounds__(64) void k_ess_weak (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const int* __restrict__ emap_g, const int* __restrict__ ess_elem,
    const int* __restrict__ ess_slot, int ness, const int* __restrict__ ess, FieldSet fs,
    int lstride, int nglobal, int stride, int nz, int base_mode)
{
  const int k = solved_plane (blockIdx.y, nz, base_mode), lane = threadIdx.x;
  const double* lines = fs.lines[blockIdx.z];
  if ((int) blockIdx.x >= ness) {
    double* G = fs.G[blockIdx.z] + (size_t) k * nglobal;
    const int g0 = ((int) blockIdx.x - ness) * 256;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int g = g0 + 64 * q + lane;
      if (g < nglobal) {
        const int sl = ess[g];
        G[g] = (sl >= 0 && lines) ? lines[(size_t) k * lstride + sl] : 0.0;
      }
    }
    return;
  }
  // -(M f + H g) as helm_op_wave forms H g (the same expressions), with
  // every load issued up front: the operator, the boundary slot's line
  // value (next <= 44 < 64 on this path, np <= 12: one slot a lane), and
  // the metric factors, mass weights and field of the lane's nodes (at
  // most 3 a lane)
  const int npnp = np * np, next = 4 * (np - 1);
  const int e = ess_elem[blockIdx.x];
  __shared__ double D[256], g[256], R[256], S[256];
  const double* ge = geo + (size_t) e * NGEO * npnp;
  double* f = fs.F[blockIdx.z] + (size_t) k * stride + (size_t) e * npnp;
  const double bk2 = betak2[k >> 1];
  int ei = 0;
  double gv = 0.0;
  if (lane < next) {
    const int sl = ess_slot[(size_t) blockIdx.x * next + lane];
    ei = emap_g[lane];
    gv = (sl >= 0 && lines) ? lines[(size_t) k * lstride + sl] : 0.0;
  }
  double q1[4], q2[4], q3[4], q4[4], q8[4], fv[4], hc[4], o[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij < npnp) {
      D[ij] = Dg[ij];
      g[ij] = 0.0;
      q1[t] = ge[G_Q1 * npnp + ij];
      q2[t] = ge[G_Q2 * npnp + ij];
      q3[t] = ge[G_Q3 * npnp + ij];
      q4[t] = ge[G_Q4 * npnp + ij];
      q8[t] = ge[G_Q8 * npnp + ij];
      fv[t] = f[ij];
      if (cyl) {
        const double y = ge[G_Y * npnp + ij], r2 = y * y;
        hc[t] = (r2 > EPSDP) ? bk2 / r2 + lambda2 : 0.0;
      } else hc[t] = bk2 + lambda2;
    }
  }
  __syncthreads ();
  if (lane < next) g[ei] = gv;
  __syncthreads ();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij < npnp) {
      const int i = ij / np, j = ij % np;
      double r = 0.0, s = 0.0;
      for (int q = 0; q < np; q++) {
        r += g[i * np + q] * D[j * np + q];
        s += D[i * np + q] * g[q * np + j];
      }
      R[ij] = q1[t] * r + q3[t] * s;
      S[ij] = q2[t] * s + q3[t] * r;
      o[t] = q4[t] * g[ij] * hc[t];
    }
  }
  __syncthreads ();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij < npnp) {
      const int i = ij / np, j = ij % np;
      double tt = 0.0;
      for (int q = 0; q < np; q++) tt += D[q * np + i] * S[q * np + j] + R[i * np + q] * D[q * np + j];
      o[t] += tt;
      f[ij] = -(fv[t] * q8[t] + o[t]);
    }
  }
}

// f (row-major; already weak on essential elements, else -M f is formed
// here) -> [F_b | F_i] in emap order with F_b = w_b - Bm w_i.
__device__ __forceinline__ void k_condense_t_body (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ emask,
    const double* __restrict__ geo, const double* __restrict__ Mcon, double* F, int stride,
    int nz, int base_mode, int weak)
{
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape Cs = con_shape (np, false);
  const int NE = Cs.NA, WR = next + Cs.KR;           // Ws rows (the k range padded)
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int c = tiles[3 * blockIdx.x], t0 = tiles[3 * blockIdx.x + 1],
            cnt = tiles[3 * blockIdx.x + 2], nep = cnt * npm;
  double* As = sm;                                   // [KR][NE]: As[q][a] = Bm[a][q]
  double* Ws = As + (size_t) Cs.KR * NE;             // [WR][CG_EP]: w in emap order
  int* emap = (int*) (Ws + (size_t) WR * CG_EP);
  const int tid = threadIdx.x;
  copy16 (As, Mcon + ((size_t) mode * nclass + c) * Cs.KR * NE, Cs.KR * NE);
  for (int i = tid; i < npnp; i += CT_THREADS) emap[i] = emap_g[i];
  for (int i = tid; i < npnp * (CG_EP - nep); i += CT_THREADS)
    Ws[(i / (CG_EP - nep)) * CG_EP + nep + i % (CG_EP - nep)] = 0.0;
  for (int i = npnp * CG_EP + tid; i < WR * CG_EP; i += CT_THREADS) Ws[i] = 0.0;
  __syncthreads ();
  // w = -Q8 f in emap order (f as given on essential elements / weak input)
  stage (nep * npnp,
         [&] (int i) {
           const int ep = i / npnp, a = i % npnp, e = elist[t0 + ep / npm], k = k0 + ep % npm;
           const int r = emap[a];
           const double v = F[(size_t) k * stride + (size_t) e * npnp + r];
           return (weak || emask[e]) ? v : -v * geo[((size_t) e * NGEO + G_Q8) * npnp + r];
         },
         [&] (int i, double v) { Ws[(i % npnp) * CG_EP + i / npnp] = v; });
  __syncthreads ();
  const int nag = NE / 4, neg = (nep + 1) / 2;
  for (int task = tid; task < nag * neg; task += CT_THREADS) {
    const int a0 = 4 * (task % nag), e0 = 2 * (task / nag);
    double acc[4][2] = {};
    tile_4x2 (As + a0, NE, Ws + (size_t) next * CG_EP + e0, nint, acc);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int ep = e0 + j;
      if (ep >= nep) continue;
      const int e = elist[t0 + ep / npm], k = k0 + ep % npm;
      double* out = F + (size_t) k * stride + (size_t) e * npnp;
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (a0 + r < next) out[a0 + r] = Ws[(a0 + r) * CG_EP + ep] - acc[r][j];
    }
  }
  for (int i = tid; i < nep * nint; i += CT_THREADS) {
    const int ep = i / nint, q = i % nint, e = elist[t0 + ep / npm], k = k0 + ep % npm;
    F[(size_t) k * stride + (size_t) e * npnp + next + q] = Ws[(next + q) * CG_EP + ep];
  }
}

__global__ __launch_bounds__(CT_THREADS) void k_condense_t (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ emask,
    const double* __restrict__ geo, const double* __restrict__ Mcon, double* F, int stride,
    int nz, int base_mode, int weak)
{
  k_condense_t_body (np, nclass, tiles, elist, emap_g, emask, geo, Mcon, F, stride, nz, base_mode, weak);
}

// u_i = Hinv^T F_i - Bm^T u_b with u_b from the band solution, then the
// essential values (Element::global2localSC src/element.cpp:368-407,
// Field::setEssential src/field.cpp:1105-1124).  One operand
// M[k][a] = (k < next ? -Bm[k][a] : Hinv[k-next][a]) against [u_b | F_i].
__device__ __forceinline__ void k_recover_t_body (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ Mrec,
    const double* __restrict__ G, int nglobal, const double* __restrict__ F, double* U,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax,
    int nz, int base_mode, int rnr, int roff)
{
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape R = rec_shape (np, false);
  const int NI = R.NA;
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int c = tiles[3 * blockIdx.x], t0 = tiles[3 * blockIdx.x + 1],
            cnt = tiles[3 * blockIdx.x + 2], nep = cnt * npm;
  double* Ms = sm;                                   // [KR][NI]
  double* Xs = Ms + (size_t) R.KR * NI;              // [KR][CG_EP]: u_b | F_i, then u
  int* emap = (int*) (Xs + (size_t) R.KR * CG_EP);
  int* pm = emap + npnp;                             // row-major position -> emap index
  const int tid = threadIdx.x, ns = nsolve[mode];
  const size_t mc = (size_t) mode * nclass + c;
  copy16 (Ms, Mrec + mc * R.KR * NI, R.KR * NI);
  for (int i = npnp * CG_EP + tid; i < R.KR * CG_EP; i += CT_THREADS) Xs[i] = 0.0;
  for (int i = tid; i < npnp; i += CT_THREADS) {
    const int r = emap_g[i];
    emap[i] = r;
    pm[r] = i;
  }
  for (int i = tid; i < npnp * (CG_EP - nep); i += CT_THREADS)
    Xs[(i / (CG_EP - nep)) * CG_EP + nep + i % (CG_EP - nep)] = 0.0;
  // u_b from the band solution (zero on essential nodes, whose values
  // entered the right-hand side); the values the output takes there
  // (essential: G) are kept in registers, BO per thread
  constexpr int BO = 6;                              // >= 32 * 44 / CT_THREADS
  double bout[BO];
#pragma unroll
  for (int u = 0; u < BO; u++) {
    const int i = tid + u * CT_THREADS;
    bout[u] = 0.0;
    if (i < nep * next) {
      const int ep = i / next, a = i % next, e = elist[t0 + ep / npm], k = k0 + ep % npm;
      const int gg = btog[(size_t) e * next + a];
      const double v = gg < ns ? rhs[((size_t) mode * rnr + roff + (k & 1)) * nmax + gg] : 0.0;
      bout[u] = (emask[e] && bmask[(size_t) e * next + a]) ? G[(size_t) k * nglobal + gg] : v;
      Xs[a * CG_EP + ep] = v;
    }
  }
  stage (nep * nint,
         [&] (int i) {
           const int ep = i / nint, q = i % nint, e = elist[t0 + ep / npm], k = k0 + ep % npm;
           return F[(size_t) k * stride + (size_t) e * npnp + next + q];
         },
         [&] (int i, double v) { Xs[(next + i % nint) * CG_EP + i / nint] = v; });
  __syncthreads ();
  const int nag = NI / 4, neg = (nep + 1) / 2, ntask = nag * neg;
  // at most two tasks per thread (nint <= 100, nep <= 32): results are
  // held in registers until every thread has finished reading Xs
  double acc[2][4][2] = {};
  for (int s = 0; s < 2; s++) {
    const int task = tid + s * CT_THREADS;
    if (task < ntask)
      tile_4x2 (Ms + 4 * (task % nag), NI, Xs + 2 * (task / nag), npnp, acc[s]);
  }
  __syncthreads ();
  for (int s = 0; s < 2; s++) {
    const int task = tid + s * CT_THREADS;
    if (task >= ntask) continue;
    const int a0 = 4 * (task % nag), e0 = 2 * (task / nag);
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        if (a0 + r < nint && e0 + j < nep) Xs[(next + a0 + r) * CG_EP + e0 + j] = acc[s][r][j];
  }
#pragma unroll
  for (int u = 0; u < BO; u++) {
    const int i = tid + u * CT_THREADS;
    if (i < nep * next) Xs[(i % next) * CG_EP + i / next] = bout[u];
  }
  __syncthreads ();
  // whole element-planes out, row-major
  for (int i = tid; i < nep * npnp; i += CT_THREADS) {
    const int ep = i / npnp, r = i % npnp, e = elist[t0 + ep / npm], k = k0 + ep % npm;
    U[(size_t) k * stride + (size_t) e * npnp + r] = Xs[pm[r] * CG_EP + ep];
  }
}

__global__ __launch_bounds__(CT_THREADS) void k_recover_t (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ Mrec,
    const double* __restrict__ G, int nglobal, const double* __restrict__ F, double* U,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax,
    int nz, int base_mode, int rnr, int roff)
{
  k_recover_t_body (np, nclass, tiles, elist, emap_g, btog, bmask, emask, Mrec, G, nglobal, F, U, stride, nsolve, rhs, nmax, nz, base_mode, rnr, roff);
}

// ---------------------------------------------------------------------------
// Register-operand MFMA condensation and recovery.  Each wavefront reads the
// (mode, class) operator columns of its output tiles straight into
// registers -- a 16-column strip, one double per lane and k-step (the A
// operand of v_mfma_f64_16x16x4f64: lane l holds M[4s + l/16][a0 + l%16] at
// step s) -- from L2, where every tile of the mode finds it.  The LDS holds
// only the element-plane operand, and a workgroup's global loads form two
// dependent rounds: (1) the tile's element numbers and tile tables
// (class-sorted btog, bmask, emask rows) with the operator strips, (2) the
// gathers that need them (field values, band solution, essential values).
// NPM bounds np: KS k-steps of 4 per strip, NV gathered values per thread.
// ---------------------------------------------------------------------------
template <int NPM> struct RegShape {
  static constexpr int NPNP = NPM * NPM, NEXT = 4 * (NPM - 1), NINT = NPNP - NEXT;
  static constexpr int KS_REC = pad4 (NPNP) / 4, KS_CON = pad4 (NINT) / 4;
  static constexpr int NV_REC = (2 * ETILE * NINT + CT_THREADS - 1) / CT_THREADS;
  static constexpr int NV_CON = (2 * ETILE * NPNP + CT_THREADS - 1) / CT_THREADS;
  static constexpr int NB_REC = (2 * ETILE * NEXT + CT_THREADS - 1) / CT_THREADS;
};
// the smallest instantiated bound for np
inline int reg_npm (int np) { return np <= 8 ? 8 : (np <= 10 ? 10 : 12); }

// wave w's output tiles: a-tiles w and w + 4 with every ep-tile when there
// are at least 3 a-tiles; otherwise a-tile w % nat with ep-tiles w / nat
// and w / nat + 4 / nat
struct WaveTiles {
  int a0, astep, e0, estep;
};
__device__ __forceinline__ WaveTiles wave_tiles (int w, int nat)
{
  if (nat >= 3) return {w, 4, 0, 1};
  return {w % nat, 1 << 20, w / nat, 4 / nat};
}

template <int KS>
__device__ __forceinline__ void load_strip (double (&a)[KS], const double* __restrict__ M, int NA,
                                            int ks, int a0, int lane)
{
  const double* p = M + (size_t) (lane >> 4) * NA + a0 + (lane & 15);
#pragma unroll
  for (int s = 0; s < KS; s++) a[s] = s < ks ? p[(size_t) 4 * s * NA] : 0.0;
}

// D = sum over k-steps s < ks of A_s B_s with A from the strip and B from
// the LDS operand X [k][CG_EP] (column block at X): the same chunks of MF_U
// steps on two accumulators and the same tail as mfma_tile, so the sums
// are formed in the same order
template <int KS>
__device__ __forceinline__ v4d mfma_strip (const double (&a)[KS], int ks,
                                           const double* __restrict__ X, int lane)
{
  v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  const double* B = X + (lane >> 4) * CG_EP + (lane & 15);
  const int kc = ks / MF_U * MF_U;
#pragma unroll
  for (int c = 0; c < KS / MF_U; c++) {
    if (c * MF_U < kc) {
      double b[MF_U];
#pragma unroll
      for (int u = 0; u < MF_U; u++) b[u] = B[(size_t) 4 * (c * MF_U + u) * CG_EP];
#pragma unroll
      for (int u = 0; u < MF_U; u += 2) {
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64 (a[c * MF_U + u], b[u], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64 (a[c * MF_U + u + 1], b[u + 1], acc1, 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < KS; s++)
    if (s >= kc && s < ks)
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64 (a[s], B[(size_t) 4 * s * CG_EP], acc0, 0, 0, 0);
  return acc0 + acc1;
}

// f (row-major; already weak on essential elements, else -M f is formed
// here) -> [F_b | F_i] in emap order with F_b = w_b - Bm w_i, as
// k_condense_t.  Round 1 reads the tile table (class, count, element
// numbers, emask) and the emap entries each thread needs; round 2 the
// field values, the mass weights and the operator strip.
template <int NPM>
__device__ __forceinline__ void k_condense_r_body (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ geo, const double* __restrict__ Mcon, double* F, int stride,
    int nz, int base_mode, int weak)
{
  typedef RegShape<NPM> RS;
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1);
  const TileShape Cs = con_shape (np, true);
  const int NE = Cs.NA, ks = Cs.KR / 4, WR = next + Cs.KR;
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int* T = ttab + (size_t) blockIdx.x * tile_ints (next);
  double* Ws = sm;                                   // [WR][CG_EP]: w in emap order
  int* el = (int*) (Ws + (size_t) WR * CG_EP);       // [ETILE] element numbers
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int nat = NE / 16;
  const WaveTiles wt = wave_tiles (wave, nat);
  // round 1
  const int c = T[0], cnt = T[1], nep = cnt * npm, nept = (nep + 15) / 16;
  if (tid < ETILE) el[tid] = T[TT_EL + tid];
  int ev[RS::NV_CON], rv[RS::NV_CON];
  bool mv[RS::NV_CON];
#pragma unroll
  for (int u = 0; u < RS::NV_CON; u++) {
    const int i = min (tid + u * CT_THREADS, 2 * ETILE * npnp - 1);
    const int j = min (i / npnp / npm, ETILE - 1);
    ev[u] = T[TT_EL + j];
    mv[u] = T[TT_EM + j] != 0;
    rv[u] = emap_g[i % npnp];
  }
  for (int i = npnp * CG_EP + tid; i < WR * CG_EP; i += CT_THREADS) Ws[i] = 0.0;
  // round 2: w = -Q8 f in emap order (f as given on essential elements /
  // weak input) and the operator strip
  const double* Mg = Mcon + ((size_t) mode * nclass + c) * Cs.KR * NE;
  double am[RS::KS_CON];
  if (wt.a0 < nat) load_strip<RS::KS_CON> (am, Mg, NE, ks, 16 * wt.a0, lane);
  {
    double fv[RS::NV_CON], sc[RS::NV_CON];
#pragma unroll
    for (int u = 0; u < RS::NV_CON; u++) {
      const int i = tid + u * CT_THREADS;
      fv[u] = 0.0;
      sc[u] = 1.0;
      if (i < nep * npnp) {
        const int ep = i / npnp, e = ev[u], r = rv[u];
        fv[u] = F[(size_t) (k0 + ep % npm) * stride + (size_t) e * npnp + r];
        if (!weak && !mv[u]) sc[u] = -geo[((size_t) e * NGEO + G_Q8) * npnp + r];
      }
    }
#pragma unroll
    for (int u = 0; u < RS::NV_CON; u++) {
      const int i = tid + u * CT_THREADS;
      if (i < nep * npnp) Ws[(i % npnp) * CG_EP + i / npnp] = fv[u] * sc[u];
    }
  }
  __syncthreads ();
  // F_b = w_b - Bm w_i over this wave's 16x16 tiles, written over w_b (each
  // entry by the lane that reads it; the products read rows next.. only)
#pragma unroll
  for (int si = 0; si < 2; si++) {
    const int ta = wt.a0 + si * wt.astep;
    if (ta >= nat) break;
    if (si) load_strip<RS::KS_CON> (am, Mg, NE, ks, 16 * ta, lane);
#pragma unroll
    for (int ei = 0; ei < 2; ei++) {
      const int te = wt.e0 + ei * wt.estep;
      if (te >= nept) continue;
      const v4d acc = mfma_strip<RS::KS_CON> (am, ks, Ws + (size_t) next * CG_EP + 16 * te, lane);
      const int ep = 16 * te + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int a = 16 * ta + 4 * r + (lane >> 4);
        if (a < next && ep < nep) Ws[a * CG_EP + ep] -= acc[r];
      }
    }
  }
  __syncthreads ();
  for (int i = tid; i < nep * npnp; i += CT_THREADS) {
    const int ep = i / npnp, a = i % npnp, e = el[ep / npm], k = k0 + ep % npm;
    F[(size_t) k * stride + (size_t) e * npnp + a] = Ws[a * CG_EP + ep];
  }
}

template <int NPM>
__global__ __launch_bounds__(CT_THREADS) void k_condense_r (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ geo, const double* __restrict__ Mcon, double* F, int stride,
    int nz, int base_mode, int weak)
{
  k_condense_r_body<NPM> (np, nclass, ttab, emap_g, geo, Mcon, F, stride, nz, base_mode, weak);
}

// u_i = Hinv^T F_i - Bm^T u_b and the essential values, as k_recover_t.
// Round 1 reads the tile table (class, count, element numbers, the btog and
// mask entries of each thread's boundary values) and the emap; round 2 the
// band solution, the essential values, F_i and the operator strip.
template <int NPM>
__device__ __forceinline__ void k_recover_r_body (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ Mrec,
    const double* __restrict__ G, int nglobal, const double* __restrict__ F, double* U,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax,
    int nz, int base_mode, int rnr, int roff, const int* __restrict__ tw,
    const double* __restrict__ rhsb, int nmaxb, int tS)
{
  typedef RegShape<NPM> RS;
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape R = rec_shape (np, true);
  const int NI = R.NA, ks = R.KR / 4;
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int* T = ttab + (size_t) blockIdx.x * tile_ints (next);
  double* Xs = sm;                                   // [KR][CG_EP]: u_b | F_i, then u
  int* el = (int*) (Xs + (size_t) R.KR * CG_EP);     // [ETILE] element numbers
  int* pm = el + ETILE;                              // row-major position -> emap index
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int nat = NI / 16;
  const WaveTiles wt = wave_tiles (wave, nat);
  // round 1
  const int c = T[0], cnt = T[1], nep = cnt * npm, nept = (nep + 15) / 16, ns = nsolve[mode];
  if (tid < ETILE) el[tid] = T[TT_EL + tid];
  for (int i = tid; i < npnp; i += CT_THREADS) pm[emap_g[i]] = i;
  int gg[RS::NB_REC], ev[RS::NV_REC];
  bool ess[RS::NB_REC];
#pragma unroll
  for (int u = 0; u < RS::NB_REC; u++) {
    const int i = min (tid + u * CT_THREADS, 2 * ETILE * next - 1);
    const int j = min (i / next / npm, ETILE - 1), a = i % next;
    gg[u] = T[TT_BT + j * next + a];
    ess[u] = (T[TT_EM + j] != 0) & (T[TT_BT + (ETILE + j) * next + a] != 0);
  }
#pragma unroll
  for (int u = 0; u < RS::NV_REC; u++) {
    const int i = min (tid + u * CT_THREADS, 2 * ETILE * nint - 1);
    ev[u] = T[TT_EL + min (i / nint / npm, ETILE - 1)];
  }
  for (int i = npnp * CG_EP + tid; i < R.KR * CG_EP; i += CT_THREADS) Xs[i] = 0.0;
  // twisted system whose band solve left B's solution in the bottom
  // piece (tw given): rows g >= m+S come from bottom row p + n-1-g
  int tbase = 1 << 30, tp = 0, tn = 0;
  if (tw) {
    const int* t = tw + TW * mode;
    tbase = t[TW_M] + tS;
    tp = t[TW_PAD];
    tn = t[TW_N];
  }
  // round 2: u_b from the band solution (zero on essential nodes, whose
  // values entered the right-hand side; the output takes G there), F_i and
  // the operator strip
  const double* Mg = Mrec + ((size_t) mode * nclass + c) * R.KR * NI;
  double am[RS::KS_REC];
  if (wt.a0 < nat) load_strip<RS::KS_REC> (am, Mg, NI, ks, 16 * wt.a0, lane);
  double ub[RS::NB_REC], bout[RS::NB_REC], fv[RS::NV_REC];
#pragma unroll
  for (int u = 0; u < RS::NB_REC; u++) {
    const int i = tid + u * CT_THREADS;
    ub[u] = 0.0;
    bout[u] = 0.0;
    if (i < nep * next) {
      const int k = k0 + (i / next) % npm;
      const size_t col = (size_t) mode * rnr + roff + (k & 1);
      if (gg[u] < ns) {
        const bool bot = gg[u] >= tbase;
        const double* src = bot ? rhsb : rhs;
        const size_t at = bot ? col * nmaxb + (size_t) (tp + tn - 1 - gg[u]) : col * nmax + gg[u];
        ub[u] = src[at];
      }
      if (ess[u]) bout[u] = G[(size_t) k * nglobal + gg[u]];
    }
  }
#pragma unroll
  for (int u = 0; u < RS::NV_REC; u++) {
    const int i = tid + u * CT_THREADS;
    fv[u] = 0.0;
    if (i < nep * nint) {
      const int ep = i / nint, q = i % nint;
      fv[u] = F[(size_t) (k0 + ep % npm) * stride + (size_t) ev[u] * npnp + next + q];
    }
  }
#pragma unroll
  for (int u = 0; u < RS::NB_REC; u++) {
    const int i = tid + u * CT_THREADS;
    if (i < nep * next) {
      Xs[(i % next) * CG_EP + i / next] = ub[u];
      if (!ess[u]) bout[u] = ub[u];
    }
  }
#pragma unroll
  for (int u = 0; u < RS::NV_REC; u++) {
    const int i = tid + u * CT_THREADS;
    if (i < nep * nint) Xs[(next + i % nint) * CG_EP + i / nint] = fv[u];
  }
  __syncthreads ();
  // u_i over this wave's 16x16 tiles, held until every wave has read Xs
  v4d acc[4];
#pragma unroll
  for (int si = 0; si < 2; si++) {
    const int ta = wt.a0 + si * wt.astep;
#pragma unroll
    for (int ei = 0; ei < 2; ei++) acc[2 * si + ei] = v4d {0.0, 0.0, 0.0, 0.0};
    if (ta >= nat) continue;
    if (si) load_strip<RS::KS_REC> (am, Mg, NI, ks, 16 * ta, lane);
#pragma unroll
    for (int ei = 0; ei < 2; ei++) {
      const int te = wt.e0 + ei * wt.estep;
      if (te < nept) acc[2 * si + ei] = mfma_strip<RS::KS_REC> (am, ks, Xs + 16 * te, lane);
    }
  }
  __syncthreads ();
#pragma unroll
  for (int si = 0; si < 2; si++) {
    const int ta = wt.a0 + si * wt.astep;
    if (ta >= nat) continue;
#pragma unroll
    for (int ei = 0; ei < 2; ei++) {
      const int te = wt.e0 + ei * wt.estep, ep = 16 * te + (lane & 15);
      if (te >= nept || ep >= nep) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int a = 16 * ta + 4 * r + (lane >> 4);
        if (a < nint) Xs[(next + a) * CG_EP + ep] = acc[2 * si + ei][r];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < RS::NB_REC; u++) {
    const int i = tid + u * CT_THREADS;
    if (i < nep * next) Xs[(i % next) * CG_EP + i / next] = bout[u];
  }
  __syncthreads ();
  // whole element-planes out, row-major
  for (int i = tid; i < nep * npnp; i += CT_THREADS) {
    const int ep = i / npnp, r = i % npnp, e = el[ep / npm], k = k0 + ep % npm;
    U[(size_t) k * stride + (size_t) e * npnp + r] = Xs[pm[r] * CG_EP + ep];
  }
}

template <int NPM>
__global__ __launch_bounds__(CT_THREADS) void k_recover_r (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ Mrec,
    const double* __restrict__ G, int nglobal, const double* __restrict__ F, double* U,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax,
    int nz, int base_mode, int rnr, int roff, const int* __restrict__ tw,
    const double* __restrict__ rhsb, int nmaxb, int tS)
{
  k_recover_r_body<NPM> (np, nclass, ttab, emap_g, Mrec, G, nglobal, F, U, stride, nsolve, rhs,
                         nmax, nz, base_mode, rnr, roff, tw, rhsb, nmaxb, tS);
}

// Up to three fields of one system in one launch (blockIdx.z selects the
// field): the grouped solve's condensations and recoveries.
struct FieldPtrs {
  double* p[3];
};
struct FieldCPtrs {
  const double* p[3];
};
template <int NPM>
__global__ __launch_bounds__(CT_THREADS) void k_condense_rb (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ geo, const double* __restrict__ Mcon, FieldPtrs F, int stride,
    int nz, int base_mode, int weak)
{
  k_condense_r_body<NPM> (np, nclass, ttab, emap_g, geo, Mcon, F.p[blockIdx.z], stride, nz,
                          base_mode, weak);
}
template <int NPM>
__global__ __launch_bounds__(CT_THREADS) void k_recover_rb (
    int np, int nclass, const int* __restrict__ ttab, const int* __restrict__ emap_g,
    const double* __restrict__ Mrec, FieldCPtrs G, int nglobal, FieldCPtrs F, FieldPtrs U,
    int stride, const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax,
    int nz, int base_mode, int rnr)
{
  k_recover_r_body<NPM> (np, nclass, ttab, emap_g, Mrec, G.p[blockIdx.z], nglobal,
                         F.p[blockIdx.z], U.p[blockIdx.z], stride, nsolve, rhs, nmax, nz,
                         base_mode, rnr, 2 * blockIdx.z, (const int*) nullptr,
                         (const double*) nullptr, 0, 0);
}

// FP32 forms of the tiled condensation and recovery for the mixed solve's
// inner passes (FP32 operands and MFMA tiles, FP32 accumulation; the
// boundary values stay FP64).
// Condensation: X (row-major, already weak) -> Fp = [w_b - Bm w_i | w_i] in
// emap order.
__device__ __forceinline__ void k_condense32_t_body (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const float* __restrict__ Mcon,
    const double* __restrict__ X, double* Fp, int stride, int nz, int base_mode)
{
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape Cs = con_shape (np, true);
  const int NE = Cs.NA;
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int c = tiles[3 * blockIdx.x], t0 = tiles[3 * blockIdx.x + 1],
            cnt = tiles[3 * blockIdx.x + 2], nep = cnt * npm;
  double* Wb = sm;                                   // [next][CG_EP] w_b (FP64)
  float* As = (float*) (Wb + (size_t) next * CG_EP); // [KR][NE]
  float* Wi = As + (size_t) Cs.KR * NE;              // [KR][CG_EP] w_i (FP32)
  int* emap = (int*) (Wi + (size_t) Cs.KR * CG_EP);
  const int tid = threadIdx.x;
  copy16f (As, Mcon + ((size_t) mode * nclass + c) * Cs.KR * NE, Cs.KR * NE);
  for (int i = tid; i < npnp; i += CT_THREADS) emap[i] = emap_g[i];
  for (int i = tid; i < Cs.KR * CG_EP; i += CT_THREADS) Wi[i] = 0.0f;
  __syncthreads ();
  stage (nep * npnp,
         [&] (int i) {
           const int ep = i / npnp, a = i % npnp, e = elist[t0 + ep / npm], k = k0 + ep % npm;
           return X[(size_t) k * stride + (size_t) e * npnp + emap[a]];
         },
         [&] (int i, double v) {
           const int ep = i / npnp, a = i % npnp;
           if (a < next) {
             Wb[a * CG_EP + ep] = v;
           } else {
             Wi[(a - next) * CG_EP + ep] = (float) v;
             const int e = elist[t0 + ep / npm], k = k0 + ep % npm;
             Fp[(size_t) k * stride + (size_t) e * npnp + a] = v;
           }
         });
  __syncthreads ();
  const int wave = tid >> 6, lane = tid & 63, nat = NE / 16, ntile = nat * ((nep + 15) / 16);
  for (int t = wave; t < ntile; t += CT_THREADS / 64) {
    const int a0 = 16 * (t % nat), ep = 16 * (t / nat) + (lane & 15);
    const v4f acc = mfma_tile32 (As + a0, NE, Wi + 16 * (t / nat), Cs.KR, lane);
    if (ep >= nep) continue;
    const int e = elist[t0 + ep / npm], k = k0 + ep % npm;
    double* out = Fp + (size_t) k * stride + (size_t) e * npnp;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int a = a0 + 4 * (lane >> 4) + r;
      if (a < next) out[a] = Wb[a * CG_EP + ep] - (double) acc[r];
    }
  }
}

__global__ __launch_bounds__(CT_THREADS) void k_condense32_t (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const float* __restrict__ Mcon,
    const double* __restrict__ X, double* Fp, int stride, int nz, int base_mode)
{
  k_condense32_t_body (np, nclass, tiles, elist, emap_g, Mcon, X, Fp, stride, nz, base_mode);
}

// Recovery: V (row-major) = [u_b | Hinv^T F_i - Bm^T u_b] with u_b from the
// band solution (no essential insertion: the correction's essential values
// are zero).
__device__ __forceinline__ void k_recover32_t_body (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const float* __restrict__ Mrec, const double* __restrict__ Fp, double* V, int stride,
    const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax, int nz,
    int base_mode)
{
  extern __shared__ double sm[];
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const TileShape R = rec_shape (np, true);
  const int NI = R.NA;
  const int mode = blockIdx.y;
  int k0;
  const int npm = mode_planes (mode, nz, base_mode, &k0);
  const int c = tiles[3 * blockIdx.x], t0 = tiles[3 * blockIdx.x + 1],
            cnt = tiles[3 * blockIdx.x + 2], nep = cnt * npm;
  double* Ub = sm;                                   // [next][CG_EP] u_b (FP64)
  float* Ms = (float*) (Ub + (size_t) next * CG_EP); // [KR][NI]
  float* Xs = Ms + (size_t) R.KR * NI;               // [KR][CG_EP] u_b | F_i, then u_i
  int* pm = (int*) (Xs + (size_t) R.KR * CG_EP);     // row-major position -> emap index
  const int tid = threadIdx.x, ns = nsolve[mode];
  copy16f (Ms, Mrec + ((size_t) mode * nclass + c) * R.KR * NI, R.KR * NI);
  for (int i = tid; i < npnp; i += CT_THREADS) pm[emap_g[i]] = i;
  for (int i = tid; i < R.KR * CG_EP; i += CT_THREADS) Xs[i] = 0.0f;
  __syncthreads ();
  stage (nep * npnp,
         [&] (int i) {
           const int ep = i / npnp, a = i % npnp, e = elist[t0 + ep / npm], k = k0 + ep % npm;
           if (a < next) {
             const int gg = btog[(size_t) e * next + a];
             return gg < ns ? rhs[((size_t) mode * 2 + (k & 1)) * nmax + gg] : 0.0;
           }
           return Fp[(size_t) k * stride + (size_t) e * npnp + a];
         },
         [&] (int i, double v) {
           const int ep = i / npnp, a = i % npnp;
           Xs[a * CG_EP + ep] = (float) v;
           if (a < next) Ub[a * CG_EP + ep] = v;
         });
  __syncthreads ();
  const int wave = tid >> 6, lane = tid & 63, nat = NI / 16, ntile = nat * ((nep + 15) / 16);
  v4f acc[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int t = wave + 4 * s;
    acc[s] = v4f {0.0f, 0.0f, 0.0f, 0.0f};
    if (t < ntile) acc[s] = mfma_tile32 (Ms + 16 * (t % nat), NI, Xs + 16 * (t / nat), R.KR, lane);
  }
  __syncthreads ();
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int t = wave + 4 * s;
    if (t >= ntile) continue;
    const int a0 = 16 * (t % nat), ep = 16 * (t / nat) + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int a = a0 + 4 * (lane >> 4) + r;
      if (a < nint && ep < nep) Xs[(next + a) * CG_EP + ep] = acc[s][r];
    }
  }
  __syncthreads ();
  for (int i = tid; i < nep * npnp; i += CT_THREADS) {
    const int ep = i / npnp, r = i % npnp, e = elist[t0 + ep / npm], k = k0 + ep % npm;
    const int a = pm[r];
    V[(size_t) k * stride + (size_t) e * npnp + r] =
        a < next ? Ub[a * CG_EP + ep] : (double) Xs[a * CG_EP + ep];
  }
}

__global__ __launch_bounds__(CT_THREADS) void k_recover32_t (
    int np, int nclass, const int* __restrict__ tiles, const int* __restrict__ elist,
    const int* __restrict__ emap_g, const int* __restrict__ btog,
    const float* __restrict__ Mrec, const double* __restrict__ Fp, double* V, int stride,
    const int* __restrict__ nsolve, const double* __restrict__ rhs, int nmax, int nz,
    int base_mode)
{
  k_recover32_t_body (np, nclass, tiles, elist, emap_g, btog, Mrec, Fp, V, stride, nsolve, rhs, nmax, nz, base_mode);
}

// FP32 copies of the MFMA-layout operands (mixed systems)
__global__ void k_ops32 (const double* __restrict__ a, float* b, size_t n)
{
  for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t) gridDim.x * blockDim.x)
    b[i] = (float) a[i];
}

__device__ __forceinline__ void k_assemble_body (int nglobal, const int* __restrict__ gptr,
                            const int* __restrict__ gsrc,
                            const int* __restrict__ nptr,
                            const int* __restrict__ nidx,
                            const double* __restrict__ nw,
                            const double* __restrict__ F, int stride,
                            const double* __restrict__ lines, int lstride,
                            const int* __restrict__ nsolve, double* rhs,
                            int nmax, int nz, int base_mode, int rnr, int roff)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1, rr = roff + (k & 1);
  if (g >= nsolve[mode]) return;
  const double* Fk = F + (size_t) k * stride;
  double s = 0.0;
  for (int c = gptr[g]; c < gptr[g + 1]; c++) s += Fk[gsrc[c]];
  if (lines)
    for (int c = nptr[g]; c < nptr[g + 1]; c++)
      s += lines[(size_t) k * lstride + nidx[c]] * nw[c];
  rhs[((size_t) mode * rnr + rr) * nmax + g] = s;
}

__global__ void k_assemble (int nglobal, const int* __restrict__ gptr,
                            const int* __restrict__ gsrc,
                            const int* __restrict__ nptr,
                            const int* __restrict__ nidx,
                            const double* __restrict__ nw,
                            const double* __restrict__ F, int stride,
                            const double* __restrict__ lines, int lstride,
                            const int* __restrict__ nsolve, double* rhs,
                            int nmax, int nz, int base_mode, int rnr, int roff)
{
  k_assemble_body (nglobal, gptr, gsrc, nptr, nidx, nw, F, stride, lines, lstride, nsolve, rhs, nmax, nz, base_mode, rnr, roff);
}

// Assembly for a twisted system that also writes the bottom piece's
// right-hand side (the gather the band solve would otherwise do): bottom row
// q = p + n-1-g holds row g of B = [m+S, n); its separator rows
// [m, m+S) and its p identity rows start at zero.
__global__ void k_assemble_tw (int nglobal, const int* __restrict__ gptr,
                               const int* __restrict__ gsrc, const int* __restrict__ nptr,
                               const int* __restrict__ nidx, const double* __restrict__ nw,
                               const double* __restrict__ F, int stride,
                               const double* __restrict__ lines, int lstride,
                               const int* __restrict__ nsolve, double* rhs, int nmax, int nz,
                               int base_mode, int rnr, const int* __restrict__ tw,
                               double* rhsb, int nmaxb, int S)
{
  k_assemble_body (nglobal, gptr, gsrc, nptr, nidx, nw, F, stride, lines, lstride, nsolve, rhs,
                   nmax, nz, base_mode, rnr, 0);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1;
  const int* t = tw + TW * mode;
  const int p = t[TW_PAD], m = t[TW_M], n = t[TW_N];
  if (g >= n) return;
  double* rb = rhsb + ((size_t) mode * rnr + (k & 1)) * nmaxb;
  if (g < p) rb[g] = 0.0;
  if (g >= m) rb[p + n - 1 - g] = g >= m + S ? rhs[((size_t) mode * rnr + (k & 1)) * nmax + g] : 0.0;
}

// The grouped solve's per-field kernels for up to three fields in one
// launch (blockIdx.z selects the field; its right-hand sides are 2z, 2z+1)
__global__ void k_assemble3 (int nglobal, const int* __restrict__ gptr,
                             const int* __restrict__ gsrc, const int* __restrict__ nptr,
                             const int* __restrict__ nidx, const double* __restrict__ nw,
                             FieldSet fs, int stride, int lstride,
                             const int* __restrict__ nsolve, double* rhs, int nmax, int nz,
                             int base_mode, int rnr)
{
  k_assemble_body (nglobal, gptr, gsrc, nptr, nidx, nw, fs.F[blockIdx.z], stride,
                   fs.lines[blockIdx.z], lstride, nsolve, rhs, nmax, nz, base_mode, rnr,
                   2 * blockIdx.z);
}

// One wavefront per (element, solved plane): interior recovery and
// essential-BC insertion (Element::global2localSC src/element.cpp:368-407,
// Field::setEssential src/field.cpp:1105-1124).
__global__ __launch_bounds__(64) void k_recover (
    int np, int nclass, const int* __restrict__ emap_g,
    const int* __restrict__ cls, const int* __restrict__ btog,
    const int* __restrict__ bmask, const int* __restrict__ emask,
    const double* __restrict__ Bm, const double* __restrict__ Hinv,
    const double* __restrict__ G, int nglobal, const double* __restrict__ F,
    double* U, int stride, const int* __restrict__ nsolve,
    const double* __restrict__ rhs, int nmax, int nz, int base_mode, int rnr, int roff)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode);
  const int mode = k >> 1, rr = roff + (k & 1), lane = threadIdx.x;
  __shared__ double ub[64], fi[256];
  __shared__ int emap[256];
  for (int i = lane; i < npnp; i += 64) emap[i] = emap_g[i];
  const int* bt = btog + (size_t) e * next;
  const int ns = nsolve[mode];
  const double* x = rhs + ((size_t) mode * rnr + rr) * nmax;
  const double* f = F + (size_t) k * stride + (size_t) e * npnp;
  double* u = U + (size_t) k * stride + (size_t) e * npnp;
  for (int i = lane; i < next; i += 64) { const int gg = bt[i]; ub[i] = gg < ns ? x[gg] : 0.0; }
  for (int a = lane; a < nint; a += 64) fi[a] = f[next + a];
  __syncthreads ();
  const size_t mc = (size_t) mode * nclass + cls[e];
  const double* B = Bm + mc * next * nint;
  const double* Hi = Hinv + mc * nint * nint;
  for (int a = lane; a < nint; a += 64) {
    double s = 0.0;
    for (int b = 0; b < nint; b++) s += Hi[b * nint + a] * fi[b];
    for (int c = 0; c < next; c++) s -= B[c * nint + a] * ub[c];
    u[emap[next + a]] = s;
  }
  for (int i = lane; i < next; i += 64) {
    double v = ub[i];
    if (emask[e] && bmask[(size_t) e * next + i]) v = G[(size_t) k * nglobal + bt[i]];
    u[emap[i]] = v;
  }
}

__global__ __launch_bounds__(64) void k_helm_op (
    int np, int cyl, double lambda2, const double* __restrict__ betak2,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const double* __restrict__ Uin, double* Uout, int stride, int nz,
    int base_mode)
{
  const int npnp = np * np;
  const int e = blockIdx.x, k = solved_plane (blockIdx.y, nz, base_mode);
  const int lane = threadIdx.x;
  __shared__ double D[256], src[256], R[256], S[256], out[256];
  for (int i = lane; i < npnp; i += 64) {
    D[i] = Dg[i];
    src[i] = Uin[(size_t) k * stride + (size_t) e * npnp + i];
  }
  __syncthreads ();
  helm_op_wave (np, D, geo + (size_t) e * NGEO * npnp, lambda2, betak2[k >> 1],
                cyl, src, R, S, out, lane);
  for (int i = lane; i < npnp; i += 64)
    Uout[(size_t) k * stride + (size_t) e * npnp + i] = out[i];
}

// ---------------------------------------------------------------------------
int n_solved_planes (int nz, int base_mode) { return (base_mode == 0 && nz > 1) ? nz - 1 : nz; }

// Optional per-phase event timing of the solve (sx_timing_enable).
struct PhaseRec { int phase; hipEvent_t a, b; };
struct Timing {
  bool on = false;
  std::vector<PhaseRec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t get ()
  {
    if (!pool.empty ()) { hipEvent_t e = pool.back (); pool.pop_back (); return e; }
    hipEvent_t e;
    (void) hipEventCreate (&e);
    return e;
  }
} g_timing;

struct PhaseScope {
  PhaseRec r;
  hipStream_t st;
  bool on;
  PhaseScope (int phase, hipStream_t s, bool allowed = true)
    : st (s), on (allowed && g_timing.on)
  {
    if (!on) return;
    r.phase = phase; r.a = g_timing.get (); r.b = g_timing.get ();
    (void) hipEventRecord (r.a, st);
  }
  ~PhaseScope ()
  {
    if (!on) return;
    (void) hipEventRecord (r.b, st);
    g_timing.recs.push_back (r);
  }
};

}  // namespace

extern "C" int sx_timing_enable (int on)
{
  g_timing.on = on != 0;
  return SX_OK;
}

extern "C" int sx_timing_read (double* ms, int* count, int nphase)
{
  if (!ms || !count || nphase < 0) { set_error ("sx_timing_read: bad arguments"); return SX_ERR_ARG; }
  for (int p = 0; p < nphase; p++) { ms[p] = 0.0; count[p] = 0; }
  for (auto& r : g_timing.recs) {
    SX_HIP (hipEventSynchronize (r.b));
    float t = 0.0f;
    SX_HIP (hipEventElapsedTime (&t, r.a, r.b));
    if (r.phase < nphase) { ms[r.phase] += t; count[r.phase]++; }
    g_timing.pool.push_back (r.a);
    g_timing.pool.push_back (r.b);
  }
  g_timing.recs.clear ();
  return SX_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" int sx_msys_destroy (sx_msys_t s)
{
  if (!s) return SX_OK;
  void* ptrs[] = {s->d_betak2, s->d_nsolve, s->d_Bm, s->d_Hinv, s->d_Mrec, s->d_Mcon,
                  s->d_Mrec32, s->d_Mcon32, s->d_Bm32,
                  s->d_Hinv32, s->d_bandU, s->d_swF, s->d_swR, s->d_rhs, s->d_G,
                  s->d_work, s->d_pc, s->d_tw, s->d_rhsb, s->d_xch, s->d_xep};
  for (void* p : ptrs) if (p) (void) hipFree (p);
  delete s;
  return SX_OK;
}

extern "C" int sx_msys_nsolve (sx_msys_t s, int mode)
{
  if (!s || mode < 0 || mode >= s->nmodes) return -1;
  return s->nsolve[mode];
}

extern "C" int sx_msys_layout (sx_msys_t s, int* info, int n)
{
  if (!s || !info || n < 0) {
    sx::set_error ("sx_msys_layout: null system or output");
    return SX_ERR_ARG;
  }
  const int v[7] = {s->sw_panel, s->sw_D, s->tw, s->m->nband, s->sw_unit, s->last_kernel,
                    s->last_nr};
  for (int i = 0; i < n && i < 7; i++) info[i] = v[i];
  return SX_OK;
}

__global__ void k_to_f32 (const double* a, float* b, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float) a[i];
}

namespace { bool tiled_ok (const sx_mesh* m); bool tiled_mfma (const sx_mesh* m); }

// The Fourier constant of a field's mode (Field::modeConstant,
// src/field.cpp:1204-1229): beta k, except for the radial and azimuthal
// velocity of a cylindrical mesh, whose coupled components v~ = v + i w and
// w~ = v - i w solve with beta k + 1 and beta k - 1 (1 for mode 0, where
// v and w are uncoupled).
static double mode_constant (int cyl, char name, int mode, double beta)
{
  if (!cyl || (name != 'v' && name != 'w')) return beta * mode;
  if (mode == 0) return 1.0;
  return name == 'v' ? beta * mode + 1.0 : beta * mode - 1.0;
}

extern "C" int sx_msys_create (sx_mesh_t m, double lambda2, double beta,
                               int base_mode, int num_modes, int precision,
                               void* stream, sx_msys_t* out)
{
  return sx_msys_create_named (m, lambda2, beta, base_mode, num_modes, 'c', precision, stream,
                               out);
}

extern "C" int sx_msys_create_named (sx_mesh_t m, double lambda2, double beta,
                                     int base_mode, int num_modes, char name, int precision,
                                     void* stream, sx_msys_t* out)
{
  if (!m || !out || num_modes < 1 || base_mode < 0 ||
      (precision != SX_FP64 && precision != SX_MIXED && precision != SX_JACPCG) ||
      !(name == 'u' || name == 'v' || name == 'w' || name == 'p' || name == 'c')) {
    set_error ("sx_msys_create: bad arguments");
    return SX_ERR_ARG;
  }
  *out = nullptr;
  hipStream_t st = (hipStream_t) stream;
  sx_msys* s = new sx_msys;
  s->m = m; s->nmodes = num_modes; s->base_mode = base_mode;
  s->precision = precision; s->lambda2 = lambda2; s->beta = beta;
  const int next = m->next, nint = m->nint, npnp = m->npnp, ld = m->nband;
  const bool fixed = m->nglobal > m->nsolve;
  for (int md = 0; md < num_modes; md++) {
    const double bk = mode_constant (m->cyl, name, base_mode + md, beta);
    const double bk2 = bk * bk;
    s->betak2.push_back (bk2);
    const bool singular = (lambda2 + bk2) < EPSSP && !fixed;
    s->nsolve.push_back (singular ? m->nsolve - 1 : m->nsolve);
  }
  s->nmax = *std::max_element (s->nsolve.begin (), s->nsolve.end ());
  s->band_stride = (size_t) std::max (s->nmax, 1) * ld;

  int rc = SX_OK;
  auto fail = [&] (int r) { sx_msys_destroy (s); return r; };
#define TRY(call) do { hipError_t _e = (call); if (_e != hipSuccess) return fail (hip_fail (_e, #call)); } while (0)
  const size_t nmc = (size_t) num_modes * m->nclass;
  TRY (hipMalloc (&s->d_betak2, sizeof (double) * num_modes));
  TRY (hipMalloc (&s->d_nsolve, sizeof (int) * num_modes));
  TRY (hipMemcpy (s->d_betak2, s->betak2.data (), sizeof (double) * num_modes, hipMemcpyHostToDevice));
  TRY (hipMemcpy (s->d_nsolve, s->nsolve.data (), sizeof (int) * num_modes, hipMemcpyHostToDevice));
  if (precision == SX_JACPCG) {      // MatrixSys JACPCG: the preconditioner only
    TRY (hipMalloc (&s->d_G, sizeof (double) * 2 * num_modes * m->nglobal));
    if ((rc = pcg_setup (s, st))) return fail (rc);
    *out = s;
    return SX_OK;
  }
  TRY (hipMalloc (&s->d_Bm, sizeof (double) * std::max<size_t> (nmc * next * nint, 1)));
  TRY (hipMalloc (&s->d_Hinv, sizeof (double) * std::max<size_t> (nmc * nint * nint, 1)));
  TRY (hipMalloc (&s->d_bandU, sizeof (double) * num_modes * s->band_stride));
  TRY (hipMalloc (&s->d_rhs, sizeof (double) * num_modes * 2 * std::max (s->nmax, 1)));
  TRY (hipMalloc (&s->d_G, sizeof (double) * 2 * num_modes * m->nglobal));
  TRY (hipMemsetAsync (s->d_rhs, 0, sizeof (double) * num_modes * 2 * std::max (s->nmax, 1), st));
  TRY (hipMemsetAsync (s->d_bandU, 0, sizeof (double) * num_modes * s->band_stride, st));

  double *d_ws = nullptr, *d_S = nullptr;
  int* d_status = nullptr;
  const size_t ws_per = (size_t) npnp * npnp + (size_t) nint * nint;
  TRY (hipMalloc (&d_ws, sizeof (double) * nmc * ws_per));
  TRY (hipMalloc (&d_S, sizeof (double) * nmc * next * next));
  TRY (hipMalloc (&d_status, sizeof (int) * (num_modes + 1)));
  TRY (hipMemsetAsync (d_status, 0, sizeof (int) * (num_modes + 1), st));

  hipLaunchKernelGGL (k_elem_sc, dim3 ((unsigned) nmc), dim3 (256), 0, st, m->np,
                      m->nclass, m->cyl, lambda2, s->d_betak2, m->d_cgeo, m->d_D,
                      m->d_emap, d_ws, d_S, s->d_Bm, s->d_Hinv, d_status + num_modes);
  TRY (hipGetLastError ());
  if (tiled_ok (m)) {
    s->tile_mf = tiled_mfma (m) ? 1 : 0;
    if (const char* e = std::getenv ("SX_TILE_MFMA")) s->tile_mf = s->tile_mf && std::atoi (e);
    const TileShape R = rec_shape (m->np, s->tile_mf), Cs = con_shape (m->np, s->tile_mf);
    const size_t nr = (size_t) R.KR * R.NA, nc = (size_t) Cs.KR * Cs.NA;
    TRY (hipMalloc (&s->d_Mrec, sizeof (double) * nmc * nr));
    TRY (hipMalloc (&s->d_Mcon, sizeof (double) * nmc * nc));
    hipLaunchKernelGGL (k_tiled_ops, dim3 ((unsigned) ((nr + nc + 255) / 256),
                                          (unsigned) std::min<size_t> (nmc, 65535)),
                        dim3 (256), 0, st, m->np, (int) nmc, s->tile_mf, s->d_Bm, s->d_Hinv,
                        s->d_Mrec, s->d_Mcon);
    TRY (hipGetLastError ());
  }
  if (m->npairs)
    hipLaunchKernelGGL (k_band_assemble, dim3 ((m->npairs + 255) / 256, num_modes),
                        dim3 (256), 0, st, m->npairs, m->d_ppos, m->d_pmn, m->d_pptr,
                        m->d_psrc, d_S, (size_t) m->nclass * next * next, s->d_bandU,
                        s->band_stride, s->d_nsolve);
  TRY (hipGetLastError ());
  if ((rc = band_factor (s, st, d_status))) return fail (rc);
  std::vector<int> status (num_modes + 1);
  TRY (hipMemcpyAsync (status.data (), d_status, sizeof (int) * (num_modes + 1),
                       hipMemcpyDeviceToHost, st));
  TRY (hipStreamSynchronize (st));
  (void) hipFree (d_ws); (void) hipFree (d_S); (void) hipFree (d_status);
  (void) hipFree (s->d_bandU);
  s->d_bandU = nullptr;
  if (status[num_modes]) {
    set_error ("sx_msys_create: matrix hii is not positive definite");
    return fail (SX_ERR_FACTOR);
  }
  for (int md = 0; md < num_modes; md++)
    if (status[md]) {
      set_error ("sx_msys_create: failed to factor Helmholtz matrix of mode " +
                 std::to_string (base_mode + md) + " (column " +
                 std::to_string (status[md]) + ")");
      return fail (SX_ERR_FACTOR);
    }
  if (precision == SX_MIXED && s->tile_mf) {
    const TileShape R = rec_shape (m->np, true), Cs = con_shape (m->np, true);
    const size_t nr = nmc * R.KR * R.NA, nc = nmc * Cs.KR * Cs.NA;
    TRY (hipMalloc (&s->d_Mrec32, sizeof (float) * nr));
    TRY (hipMalloc (&s->d_Mcon32, sizeof (float) * nc));
    hipLaunchKernelGGL (k_ops32, dim3 ((unsigned) std::min<size_t> ((nr + 255) / 256, 65535)),
                        dim3 (256), 0, st, s->d_Mrec, s->d_Mrec32, nr);
    hipLaunchKernelGGL (k_ops32, dim3 ((unsigned) std::min<size_t> ((nc + 255) / 256, 65535)),
                        dim3 (256), 0, st, s->d_Mcon, s->d_Mcon32, nc);
    TRY (hipGetLastError ());
  }
  if (precision == SX_MIXED) {
    TRY (hipMalloc (&s->d_Bm32, sizeof (float) * std::max<size_t> (nmc * next * nint, 1)));
    TRY (hipMalloc (&s->d_Hinv32, sizeof (float) * std::max<size_t> (nmc * nint * nint, 1)));
    const size_t n1 = nmc * next * nint, n2 = nmc * nint * nint;
    if (n1) hipLaunchKernelGGL (k_to_f32, dim3 ((unsigned) ((n1 + 255) / 256)), dim3 (256), 0, st, s->d_Bm, s->d_Bm32, n1);
    if (n2) hipLaunchKernelGGL (k_to_f32, dim3 ((unsigned) ((n2 + 255) / 256)), dim3 (256), 0, st, s->d_Hinv, s->d_Hinv32, n2);
    TRY (hipGetLastError ());
    TRY (hipStreamSynchronize (st));
  }
#undef TRY
  (void) rc;
  *out = s;
  return SX_OK;
}

// direct: the entry point runs the DIRECT solve, which a JACPCG system
// (preconditioner only, no condensed matrices or band factor) cannot do
static int check_solve_args (sx_msys_t s, int nz, int stride, const void* f,
                             const void* u, bool direct = true)
{
  if (!s || !f || !u || nz < 1 || stride < s->m->nel * s->m->npnp) {
    set_error ("sx_field_solve: bad arguments");
    return SX_ERR_ARG;
  }
  if (nz > 2 * s->nmodes && !(nz == 1)) {
    set_error ("sx_field_solve: more planes than modes in the system");
    return SX_ERR_ARG;
  }
  if (direct && s->precision == SX_JACPCG) {
    set_error ("sx_field_solve: a JACPCG system is solved with sx_field_solve_pcg");
    return SX_ERR_ARG;
  }
  return SX_OK;
}

namespace sx {
int check_direct_solve_args (sx_msys* s, int nz, int stride, const void* f, const void* u)
{
  return check_solve_args (s, nz, stride, f, u, true);
}
}  // namespace sx

namespace {

size_t tiled_lds (const sx_mesh* m, bool recover, bool mf)
{
  const size_t next = m->next, npnp = m->npnp;
  if (recover) {
    const TileShape R = rec_shape (m->np, mf);
    return sizeof (double) * ((size_t) R.KR * R.NA + (size_t) R.KR * CG_EP) +
           sizeof (int) * 2 * npnp;
  }
  const TileShape Cs = con_shape (m->np, mf);
  return sizeof (double) * ((size_t) Cs.KR * Cs.NA + (next + Cs.KR) * CG_EP) +
         sizeof (int) * npnp;
}

// the tiled kernels hold at most 2 product tasks and 6 boundary values per
// thread: nint <= 100 and next <= 48 (np <= 12, which the LDS bound implies)
bool tiled_ok (const sx_mesh* m)
{
  return tiled_lds (m, true, false) <= 160 * 1024 && m->nint <= 100 &&
         2 * ETILE * m->next <= 6 * 256;
}

// the MFMA form where its (16-padded) operands fit the LDS as well
bool tiled_mfma (const sx_mesh* m)
{
  return tiled_ok (m) && tiled_lds (m, true, true) <= 160 * 1024 &&
         tiled_lds (m, false, true) <= 160 * 1024;
}

int n_mode_blocks (int nz) { return nz == 1 ? 1 : nz / 2; }

// constrain + condense of one field (F: forcing in, condensed out)
// LDS of the register-operand tile kernels
size_t reg_lds (const sx_mesh* m, bool recover)
{
  if (recover) {
    const TileShape R = rec_shape (m->np, true);
    return sizeof (double) * (size_t) R.KR * CG_EP + sizeof (int) * (ETILE + m->npnp);
  }
  const TileShape Cs = con_shape (m->np, true);
  return sizeof (double) * (size_t) (m->next + Cs.KR) * CG_EP + sizeof (int) * ETILE;
}

template <int NPM>
int launch_condense_r (const sx_msys* s, int nz, int stride, double* F, int weak, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const size_t lds = reg_lds (m, false);
  auto kern = k_condense_r<NPM>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS), lds, st,
                      m->np, m->nclass, m->d_ttab, m->d_emap, m->d_geo, s->d_Mcon, F, stride, nz,
                      s->base_mode, weak);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

template <int NPM>
int launch_recover_r (const sx_msys* s, int nz, int stride, const double* F, double* U,
                      const double* G, int rnr, int roff, hipStream_t st, bool bottom = false)
{
  const sx_mesh* m = s->m;
  const size_t lds = reg_lds (m, true);
  auto kern = k_recover_r<NPM>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS), lds, st,
                      m->np, m->nclass, m->d_ttab, m->d_emap, s->d_Mrec, G, m->nglobal, F, U,
                      stride, s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, rnr, roff,
                      bottom ? (const int*) s->d_tw : nullptr, (const double*) s->d_rhsb,
                      s->nmaxb, s->tw_S);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// the register-operand condensation / recovery of nf <= 3 fields of one
// system in one launch each (the grouped solve)
template <int NPM>
int launch_condense_group (const sx_msys* s, int nf, int nz, int stride, const FieldPtrs& F,
                           hipStream_t st)
{
  const sx_mesh* m = s->m;
  const size_t lds = reg_lds (m, false);
  auto kern = k_condense_rb<NPM>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz), nf), dim3 (CT_THREADS), lds, st,
                      m->np, m->nclass, m->d_ttab, m->d_emap, m->d_geo, s->d_Mcon, F, stride, nz,
                      s->base_mode, 0);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}
template <int NPM>
int launch_recover_group (const sx_msys* s, int nf, int nz, int stride, const FieldCPtrs& F,
                          const FieldPtrs& U, const FieldCPtrs& G, int rnr, hipStream_t st)
{
  const sx_mesh* m = s->m;
  const size_t lds = reg_lds (m, true);
  auto kern = k_recover_rb<NPM>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz), nf), dim3 (CT_THREADS), lds, st,
                      m->np, m->nclass, m->d_ttab, m->d_emap, s->d_Mrec, G, m->nglobal, F, U,
                      stride, s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, rnr);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// ess_done: k_ess_weak has formed the essential elements' weak form
int condense (const sx_msys* s, int nz, int stride, double* F, const double* G, hipStream_t st,
              int weak = 0, bool ess_done = false)
{
  const sx_mesh* m = s->m;
  const int npl = n_solved_planes (nz, s->base_mode);
  if (weak && !tiled_ok (m)) {
    set_error ("weak-form solve: element matrices too large for the tiled kernels (np > 12)");
    return SX_ERR_LIMIT;
  }
  if (!tiled_ok (m)) {
    hipLaunchKernelGGL (k_condense, dim3 (m->nel, npl), dim3 (64), 0, st, m->np, m->nel,
                        m->nclass, m->cyl, s->lambda2, s->d_betak2, m->d_geo, m->d_D,
                        m->d_emap, m->d_cls, m->d_btog, m->d_emask, s->d_Bm, G,
                        m->nglobal, F, stride, nz, s->base_mode);
    SX_HIP (hipGetLastError ());
    return SX_OK;
  }
  if (m->ness && !weak && !ess_done)
    hipLaunchKernelGGL (k_weak_ess, dim3 (m->ness, npl), dim3 (64), 0, st, m->np, m->cyl,
                        s->lambda2, s->d_betak2, m->d_geo, m->d_D, m->d_emap, m->d_btog,
                        m->d_ess_elem, G, m->nglobal, F, stride, nz, s->base_mode);
  const bool mf = s->tile_mf;
  if (mf) {
    switch (reg_npm (m->np)) {
    case 8: return launch_condense_r<8> (s, nz, stride, F, weak, st);
    case 10: return launch_condense_r<10> (s, nz, stride, F, weak, st);
    default: return launch_condense_r<12> (s, nz, stride, F, weak, st);
    }
  }
  const size_t lds = tiled_lds (m, false, false);
  auto kern = k_condense_t;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS),
                      lds, st, m->np, m->nclass, m->d_tile, m->d_elist, m->d_emap, m->d_emask,
                      m->d_geo, s->d_Mcon, F, stride, nz, s->base_mode, weak);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// interior recovery + essential insertion of one field
int recover (const sx_msys* s, int nz, int stride, const double* F, double* U, const double* G,
             int rnr, int roff, hipStream_t st)
{
  const sx_mesh* m = s->m;
  if (!tiled_ok (m)) {
    hipLaunchKernelGGL (k_recover, dim3 (m->nel, n_solved_planes (nz, s->base_mode)), dim3 (64),
                        0, st, m->np, m->nclass, m->d_emap, m->d_cls, m->d_btog, m->d_bmask,
                        m->d_emask, s->d_Bm, s->d_Hinv, G, m->nglobal, F, U, stride,
                        s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, rnr, roff);
    SX_HIP (hipGetLastError ());
    return SX_OK;
  }
  const bool mf = s->tile_mf;
  if (mf) {
    switch (reg_npm (m->np)) {
    case 8: return launch_recover_r<8> (s, nz, stride, F, U, G, rnr, roff, st);
    case 10: return launch_recover_r<10> (s, nz, stride, F, U, G, rnr, roff, st);
    default: return launch_recover_r<12> (s, nz, stride, F, U, G, rnr, roff, st);
    }
  }
  const size_t lds = tiled_lds (m, true, false);
  auto kern = k_recover_t;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS), lds,
                      st, m->np, m->nclass, m->d_tile, m->d_elist, m->d_emap, m->d_btog,
                      m->d_bmask, m->d_emask, s->d_Mrec, G, m->nglobal, F, U, stride,
                      s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, rnr, roff);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace

namespace sx {

// The DIRECT solve of Field::solve over nz planes with the system's
// workspace (d_G, d_rhs, d_rhsb, and the split sweeps' rings when
// no_split is 0) as `s` names it: sx_field_solve passes the system itself,
// sx_field_solve_noalloc a copy whose workspace pointers lead into its
// static scratch.  timed: record the per-phase HIP events when timing is on.
int field_solve_direct (const sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                        const double* d_lines, hipStream_t st, bool timed)
{
  int rc;
  const sx_mesh* m = s->m;
  const int npl = n_solved_planes (nz, s->base_mode);
  const int lstride = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : nullptr;
  // the essential values and the essential elements' weak form in one
  // launch where the tiled condensation follows
  const bool ew = tiled_ok (m) && m->ness;
  {
    PhaseScope ps (0, st, timed);
    if (ew) {
      FieldSet fs {};
      fs.lines[0] = lines;
      fs.G[0] = s->d_G;
      fs.F[0] = d_f;
      hipLaunchKernelGGL (k_ess_weak, dim3 (m->ness + (m->nglobal + 255) / 256, npl), dim3 (64),
                          0, st, m->np, m->cyl, s->lambda2, s->d_betak2, m->d_geo, m->d_D,
                          m->d_emap, m->d_ess_elem, m->d_ess_slot, m->ness, m->d_ess, fs,
                          lstride, m->nglobal, stride, nz, s->base_mode);
    } else if (m->ness) {
      // (with no element holding an essential node no kernel reads G)
      hipLaunchKernelGGL (k_essential, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                          st, m->nglobal, m->d_ess, lines, lstride, s->d_G, nz, s->base_mode);
    }
  }
  {
    PhaseScope ps (1, st, timed);
    if ((rc = condense (s, nz, stride, d_f, s->d_G, st, 0, ew))) return rc;
  }
  // a twisted system's bottom piece addressed by the assembly and the
  // recovery themselves (no gather and scatter launches around the sweeps)
  const bool fold = s->tw && s->sw_panel && tiled_ok (m) && s->tile_mf;
  {
    PhaseScope ps (2, st, timed);
    if (fold)
      hipLaunchKernelGGL (k_assemble_tw, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                          st, m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw,
                          d_f, stride, lines, lstride, s->d_nsolve, s->d_rhs, s->nmax, nz,
                          s->base_mode, 2, s->d_tw, s->d_rhsb, s->nmaxb, s->tw_S);
    else
      hipLaunchKernelGGL (k_assemble, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                          st, m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw,
                          d_f, stride, lines, lstride, s->d_nsolve, s->d_rhs, s->nmax, nz,
                          s->base_mode, 2, 0);
  }
  SX_HIP (hipGetLastError ());
  {
    PhaseScope ps (3, st, timed);
    if ((rc = band_solve (s, 2, st, fold ? 0 : 3))) return rc;
  }
  {
    PhaseScope ps (4, st, timed);
    if (fold) {
      switch (reg_npm (m->np)) {
      case 8: rc = launch_recover_r<8> (s, nz, stride, d_f, d_u, s->d_G, 2, 0, st, true); break;
      case 10: rc = launch_recover_r<10> (s, nz, stride, d_f, d_u, s->d_G, 2, 0, st, true); break;
      default: rc = launch_recover_r<12> (s, nz, stride, d_f, d_u, s->d_G, 2, 0, st, true); break;
      }
      if (rc) return rc;
    } else if ((rc = recover (s, nz, stride, d_f, d_u, s->d_G, 2, 0, st))) {
      return rc;
    }
  }
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace sx

extern "C" int sx_field_solve (sx_msys_t s, int nz, int stride, double* d_f,
                               double* d_u, const double* d_lines, void* stream)
{
  const int rc = check_solve_args (s, nz, stride, d_f, d_u);
  if (rc) return rc;
  return field_solve_direct (s, nz, stride, d_f, d_u, d_lines, (hipStream_t) stream, true);
}

namespace sx {

int msys_reserve_rhs (sx_msys* s, int nr)
{
  if (nr != 2 && nr != 4 && nr != 6) {
    set_error ("msys_reserve_rhs: 2, 4 or 6 right-hand sides per mode");
    return SX_ERR_ARG;
  }
  if (nr <= s->rhs_cap) return SX_OK;
  if (!s->sw_panel || s->sw_D > 6) {
    set_error ("msys_reserve_rhs: more than 2 right-hand sides per mode need the panel "
               "sweeps with at most 6 off-diagonal blocks (band offsets kd <= 447)");
    return SX_ERR_LIMIT;
  }
  double *r = nullptr, *g = nullptr;
  if (s->tw) {
    double* rb = nullptr;
    const size_t nb = sizeof (double) * s->nmodes * nr * s->nmaxb;
    SX_HIP (hipMalloc (&rb, nb));
    SX_HIP (hipMemset (rb, 0, nb));
    (void) hipFree (s->d_rhsb);
    s->d_rhsb = rb;
  }
  SX_HIP (hipMalloc (&r, sizeof (double) * s->nmodes * nr * std::max (s->nmax, 1)));
  SX_HIP (hipMalloc (&g, sizeof (double) * (nr / 2) * 2 * s->nmodes * s->m->nglobal));
  SX_HIP (hipMemset (r, 0, sizeof (double) * s->nmodes * nr * std::max (s->nmax, 1)));
  SX_HIP (hipMemset (g, 0, sizeof (double) * (nr / 2) * 2 * s->nmodes * s->m->nglobal));
  (void) hipFree (s->d_rhs);
  (void) hipFree (s->d_G);
  s->d_rhs = r;
  s->d_G = g;
  s->rhs_cap = nr;
  return SX_OK;
}

int condense32_tiled (const sx_msys* s, int nz, int stride, const double* X, double* Fp,
                      hipStream_t st)
{
  const sx_mesh* m = s->m;
  const TileShape Cs = con_shape (m->np, true);
  const size_t lds = sizeof (double) * m->next * CG_EP +
                     sizeof (float) * ((size_t) Cs.KR * Cs.NA + (size_t) Cs.KR * CG_EP) +
                     sizeof (int) * m->npnp;
  SX_HIP (hipFuncSetAttribute ((const void*) k_condense32_t,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (k_condense32_t, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS),
                      lds, st, m->np, m->nclass, m->d_tile, m->d_elist, m->d_emap, s->d_Mcon32,
                      X, Fp, stride, nz, s->base_mode);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

int recover32_tiled (const sx_msys* s, int nz, int stride, const double* Fp, double* V,
                     hipStream_t st)
{
  const sx_mesh* m = s->m;
  const TileShape R = rec_shape (m->np, true);
  const size_t lds = sizeof (double) * m->next * CG_EP +
                     sizeof (float) * ((size_t) R.KR * R.NA + (size_t) R.KR * CG_EP) +
                     sizeof (int) * m->npnp;
  SX_HIP (hipFuncSetAttribute ((const void*) k_recover32_t,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (k_recover32_t, dim3 (m->ntile, n_mode_blocks (nz)), dim3 (CT_THREADS),
                      lds, st, m->np, m->nclass, m->d_tile, m->d_elist, m->d_emap, m->d_btog,
                      s->d_Mrec32, Fp, V, stride, s->d_nsolve, s->d_rhs, s->nmax, nz,
                      s->base_mode);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// ---------------------------------------------------------------------------
// Fused-batch plans of the element phases (the launches of
// field_solve_direct, condense32_tiled and recover32_tiled, recorded)
// ---------------------------------------------------------------------------
}  // namespace sx

namespace {
struct FnEssential {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_essential_body (a...); }
};
struct FnWeakEss {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_weak_ess_body (a...); }
};
struct FnAssemble {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_assemble_body (a...); }
};
struct FnCondenseT {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_condense_t_body (a...); }
};
struct FnRecoverT {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_recover_t_body (a...); }
};
template <int NPM> struct FnCondenseR {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_condense_r_body<NPM> (a...); }
};
template <int NPM> struct FnRecoverR {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_recover_r_body<NPM> (a...); }
};
struct FnCondense32T {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_condense32_t_body (a...); }
};
struct FnRecover32T {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_recover32_t_body (a...); }
};
}  // namespace

namespace sx {

int plan_field_solve_direct (BatchPlan& P, const sx_msys* s, int nz, int stride, double* d_f,
                             double* d_u, const double* d_lines)
{
  const sx_mesh* m = s->m;
  if (!tiled_ok (m) || !s->d_Mcon || !s->d_Mrec || s->precision != SX_FP64) return SX_ERR_LIMIT;
  const int npl = n_solved_planes (nz, s->base_mode);
  const int lstride = m->nbound * m->np;
  const double* lines = (d_lines && m->nbound) ? d_lines : (const double*) nullptr;
  const unsigned gb = (unsigned) ((m->nglobal + 255) / 256);
  batch_rec<256, FnEssential> (P, dim3 (gb, npl), 256, 0, m->nglobal, (const int*) m->d_ess,
                               lines, lstride, s->d_G, nz, s->base_mode);
  P.phase++;
  if (m->ness) {
    batch_rec<64, FnWeakEss> (P, dim3 (m->ness, npl), 64, 0, m->np, m->cyl, s->lambda2,
                              (const double*) s->d_betak2, (const double*) m->d_geo,
                              (const double*) m->d_D, (const int*) m->d_emap,
                              (const int*) m->d_btog, (const int*) m->d_ess_elem,
                              (const double*) s->d_G, m->nglobal, d_f, stride, nz,
                              s->base_mode);
    P.phase++;
  }
  const dim3 tg (m->ntile, n_mode_blocks (nz));
  const int weak = 0;
  const int npmr = reg_npm (m->np);
  auto con_r = [&] (auto fn) {
    batch_rec<CT_THREADS, decltype (fn)> (
        P, tg, CT_THREADS, reg_lds (m, false), m->np, m->nclass, (const int*) m->d_ttab,
        (const int*) m->d_emap, (const double*) m->d_geo, (const double*) s->d_Mcon, d_f, stride,
        nz, s->base_mode, weak);
  };
  if (s->tile_mf) {
    if (npmr == 8) con_r (FnCondenseR<8> {});
    else if (npmr == 10) con_r (FnCondenseR<10> {});
    else con_r (FnCondenseR<12> {});
  } else {
    batch_rec<CT_THREADS, FnCondenseT> (
        P, tg, CT_THREADS, tiled_lds (m, false, false), m->np, m->nclass,
        (const int*) m->d_tile, (const int*) m->d_elist, (const int*) m->d_emap,
        (const int*) m->d_emask, (const double*) m->d_geo, (const double*) s->d_Mcon, d_f,
        stride, nz, s->base_mode, weak);
  }
  P.phase++;
  const int rnr = 2, roff = 0;
  batch_rec<256, FnAssemble> (P, dim3 (gb, npl), 256, 0, m->nglobal, (const int*) m->d_gptr,
                              (const int*) m->d_gsrc, (const int*) m->d_nptr,
                              (const int*) m->d_nidx, (const double*) m->d_nw,
                              (const double*) d_f, stride, lines, lstride,
                              (const int*) s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode,
                              rnr, roff);
  P.phase++;
  if (int rc = plan_band_solve (P, s, 2)) return rc;
  auto rec_r = [&] (auto fn) {
    batch_rec<CT_THREADS, decltype (fn)> (
        P, tg, CT_THREADS, reg_lds (m, true), m->np, m->nclass, (const int*) m->d_ttab,
        (const int*) m->d_emap, (const double*) s->d_Mrec,
        (const double*) s->d_G, m->nglobal, (const double*) d_f, d_u, stride,
        (const int*) s->d_nsolve, (const double*) s->d_rhs, s->nmax, nz, s->base_mode, rnr,
        roff, (const int*) nullptr, (const double*) nullptr, 0, 0);
  };
  if (s->tile_mf) {
    if (npmr == 8) rec_r (FnRecoverR<8> {});
    else if (npmr == 10) rec_r (FnRecoverR<10> {});
    else rec_r (FnRecoverR<12> {});
  } else {
    batch_rec<CT_THREADS, FnRecoverT> (
        P, tg, CT_THREADS, tiled_lds (m, true, false), m->np, m->nclass,
        (const int*) m->d_tile, (const int*) m->d_elist, (const int*) m->d_emap,
        (const int*) m->d_btog, (const int*) m->d_bmask, (const int*) m->d_emask,
        (const double*) s->d_Mrec, (const double*) s->d_G, m->nglobal, (const double*) d_f,
        d_u, stride, (const int*) s->d_nsolve, (const double*) s->d_rhs, s->nmax, nz,
        s->base_mode, rnr, roff);
  }
  P.phase++;
  return SX_OK;
}

int plan_condense32_tiled (BatchPlan& P, const sx_msys* s, int nz, int stride, const double* X,
                           double* Fp)
{
  const sx_mesh* m = s->m;
  const TileShape Cs = con_shape (m->np, true);
  const size_t lds = sizeof (double) * m->next * CG_EP +
                     sizeof (float) * ((size_t) Cs.KR * Cs.NA + (size_t) Cs.KR * CG_EP) +
                     sizeof (int) * m->npnp;
  batch_rec<CT_THREADS, FnCondense32T> (P, dim3 (m->ntile, n_mode_blocks (nz)), CT_THREADS, lds,
                                        m->np, m->nclass, (const int*) m->d_tile,
                                        (const int*) m->d_elist, (const int*) m->d_emap,
                                        (const float*) s->d_Mcon32, X, Fp, stride, nz,
                                        s->base_mode);
  P.phase++;
  return SX_OK;
}

int plan_recover32_tiled (BatchPlan& P, const sx_msys* s, int nz, int stride, const double* Fp,
                          double* V)
{
  const sx_mesh* m = s->m;
  const TileShape R = rec_shape (m->np, true);
  const size_t lds = sizeof (double) * m->next * CG_EP +
                     sizeof (float) * ((size_t) R.KR * R.NA + (size_t) R.KR * CG_EP) +
                     sizeof (int) * m->npnp;
  batch_rec<CT_THREADS, FnRecover32T> (P, dim3 (m->ntile, n_mode_blocks (nz)), CT_THREADS, lds,
                                       m->np, m->nclass, (const int*) m->d_tile,
                                       (const int*) m->d_elist, (const int*) m->d_emap,
                                       (const int*) m->d_btog, (const float*) s->d_Mrec32, Fp,
                                       V, stride, (const int*) s->d_nsolve,
                                       (const double*) s->d_rhs, s->nmax, nz, s->base_mode);
  P.phase++;
  return SX_OK;
}

int field_solve_group (sx_msys* s, int nf, int nz, int stride, double* const* d_f,
                       double* const* d_u, const double* const* d_lines, hipStream_t st)
{
  if (nf < 1 || 2 * nf > s->rhs_cap) {
    set_error ("field_solve_group: more fields than the reserved right-hand sides");
    return SX_ERR_ARG;
  }
  for (int f = 0; f < nf; f++)
    if (int rc = check_solve_args (s, nz, stride, d_f[f], d_u[f])) return rc;
  const sx_mesh* m = s->m;
  const int npl = n_solved_planes (nz, s->base_mode), nr = 2 * nf;
  const int lstride = m->nbound * m->np;
  const size_t gstride = (size_t) 2 * s->nmodes * m->nglobal;
  // the register-operand element kernels take up to three fields per launch
  const bool grouped = nf <= 3 && tiled_ok (m) && s->tile_mf;
  const int npmr = reg_npm (m->np);
  if (grouped) {
    FieldSet fs {};
    FieldPtrs F {};
    for (int f = 0; f < nf; f++) {
      fs.lines[f] = (d_lines && d_lines[f] && m->nbound) ? d_lines[f] : nullptr;
      fs.G[f] = s->d_G + f * gstride;
      fs.F[f] = d_f[f];
      F.p[f] = d_f[f];
    }
    if (m->ness)
      hipLaunchKernelGGL (k_ess_weak, dim3 (m->ness + (m->nglobal + 255) / 256, npl, nf),
                          dim3 (64), 0, st, m->np, m->cyl, s->lambda2, s->d_betak2, m->d_geo,
                          m->d_D, m->d_emap, m->d_ess_elem, m->d_ess_slot, m->ness, m->d_ess,
                          fs, lstride, m->nglobal, stride, nz, s->base_mode);
    // (with no element holding an essential node no kernel reads G)
    const int rc = npmr == 8 ? launch_condense_group<8> (s, nf, nz, stride, F, st)
                 : npmr == 10 ? launch_condense_group<10> (s, nf, nz, stride, F, st)
                              : launch_condense_group<12> (s, nf, nz, stride, F, st);
    if (rc) return rc;
    hipLaunchKernelGGL (k_assemble3, dim3 ((m->nglobal + 255) / 256, npl, nf), dim3 (256), 0,
                        st, m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw, fs,
                        stride, lstride, s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, nr);
  }
  for (int f = 0; f < nf && !grouped; f++) {
    const double* lines = (d_lines && d_lines[f] && m->nbound) ? d_lines[f] : nullptr;
    double* G = s->d_G + f * gstride;
    hipLaunchKernelGGL (k_essential, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                        st, m->nglobal, m->d_ess, lines, lstride, G, nz, s->base_mode);
    if (int rc = condense (s, nz, stride, d_f[f], G, st)) return rc;
    hipLaunchKernelGGL (k_assemble, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0,
                        st, m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw,
                        d_f[f], stride, lines, lstride, s->d_nsolve, s->d_rhs, s->nmax, nz,
                        s->base_mode, nr, 2 * f);
  }
  SX_HIP (hipGetLastError ());
  if (int rc = band_solve (s, nr, st)) return rc;
  if (grouped) {
    FieldCPtrs F {}, G {};
    FieldPtrs U {};
    for (int f = 0; f < nf; f++) {
      F.p[f] = d_f[f];
      U.p[f] = d_u[f];
      G.p[f] = s->d_G + f * gstride;
    }
    return npmr == 8 ? launch_recover_group<8> (s, nf, nz, stride, F, U, G, nr, st)
         : npmr == 10 ? launch_recover_group<10> (s, nf, nz, stride, F, U, G, nr, st)
                      : launch_recover_group<12> (s, nf, nz, stride, F, U, G, nr, st);
  }
  for (int f = 0; f < nf; f++)
    if (int rc = recover (s, nz, stride, d_f[f], d_u[f], s->d_G + f * gstride, nr, 2 * f, st))
      return rc;
  return SX_OK;
}

}  // namespace sx

extern "C" int sx_field_solve_host (sx_msys_t s, int nz, int stride, const double* f,
                                    double* u, const double* lines)
{
  int rc = check_solve_args (s, nz, stride, f, u);
  if (rc) return rc;
  const size_t nb = sizeof (double) * (size_t) nz * stride;
  const size_t lb = sizeof (double) * (size_t) nz * s->m->nbound * s->m->np;
  double *df = nullptr, *du = nullptr, *dl = nullptr;
  SX_HIP (hipMalloc (&df, nb));
  SX_HIP (hipMalloc (&du, nb));
  SX_HIP (hipMemcpy (df, f, nb, hipMemcpyHostToDevice));
  SX_HIP (hipMemcpy (du, u, nb, hipMemcpyHostToDevice));
  if (lines && lb) {
    SX_HIP (hipMalloc (&dl, lb));
    SX_HIP (hipMemcpy (dl, lines, lb, hipMemcpyHostToDevice));
  }
  rc = sx_field_solve (s, nz, stride, df, du, dl, nullptr);
  if (rc == SX_OK) rc = split_status (s, nullptr);
  if (rc == SX_OK) {
    SX_HIP (hipMemcpy (u, du, nb, hipMemcpyDeviceToHost));
  }
  (void) hipFree (df); (void) hipFree (du); if (dl) (void) hipFree (dl);
  return rc;
}

extern "C" int sx_msys_status (sx_msys_t s, void* stream)
{
  if (!s) { set_error ("sx_msys_status: null system"); return SX_ERR_ARG; }
  return split_status (s, (hipStream_t) stream);
}

extern "C" int sx_helmholtz_op (sx_msys_t s, int nz, int stride, const double* d_u,
                                double* d_hu, void* stream)
{
  int rc = check_solve_args (s, nz, stride, d_u, d_hu, false);
  if (rc) return rc;
  const sx_mesh* m = s->m;
  hipLaunchKernelGGL (k_helm_op, dim3 (m->nel, n_solved_planes (nz, s->base_mode)),
                      dim3 (64), 0, (hipStream_t) stream, m->np, m->cyl, s->lambda2,
                      s->d_betak2, m->d_geo, m->d_D, d_u, d_hu, stride, nz,
                      s->base_mode);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// ---------------------------------------------------------------------------
// Adjoint building blocks
// ---------------------------------------------------------------------------
namespace {

// out[k] = sum over elements and nodes of Q4 a b (the mass-weighted inner
// product of plane k, element quadrature as AuxField::integral does);
// one workgroup per plane
__global__ __launch_bounds__(256) void k_mass_inner (int npnp, int nel,
                                                     const double* __restrict__ geo,
                                                     const double* __restrict__ A,
                                                     const double* __restrict__ B, int stride,
                                                     double* out)
{
  __shared__ double red[256];
  const int k = blockIdx.x;
  const size_t n = (size_t) nel * npnp;
  double s = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) {
    const size_t e = i / npnp, ij = i % npnp;
    s += geo[(e * NGEO + G_Q4) * npnp + ij] * A[(size_t) k * stride + i] * B[(size_t) k * stride + i];
  }
  red[threadIdx.x] = s;
  __syncthreads ();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int) threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads ();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// out = scale Q8 in, element by element (Element::weight)
__global__ void k_mass_weight (int npnp, int nel, const double* __restrict__ geo,
                               const double* __restrict__ in, double* out, int stride,
                               double scale, int nz)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t) nel * npnp;
  if (i >= n * nz) return;
  const size_t k = i / n, r = i % n, e = r / npnp, ij = r % npnp;
  out[k * stride + r] = scale * geo[(e * NGEO + G_Q8) * npnp + ij] * in[k * stride + r];
}

}  // namespace

extern "C" int sx_field_solve_weak (sx_msys_t s, int nz, int stride, double* d_w, double* d_u,
                                    void* stream)
{
  int rc = check_solve_args (s, nz, stride, d_w, d_u);
  if (rc) return rc;
  if (s->precision == SX_JACPCG) {
    set_error ("sx_field_solve_weak: needs a DIRECT system");
    return SX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t) stream;
  const sx_mesh* m = s->m;
  const int npl = n_solved_planes (nz, s->base_mode);
  hipLaunchKernelGGL (k_essential, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0, st,
                      m->nglobal, m->d_ess, nullptr, 0, s->d_G, nz, s->base_mode);
  if ((rc = condense (s, nz, stride, d_w, s->d_G, st, 1))) return rc;
  hipLaunchKernelGGL (k_assemble, dim3 ((m->nglobal + 255) / 256, npl), dim3 (256), 0, st,
                      m->nglobal, m->d_gptr, m->d_gsrc, m->d_nptr, m->d_nidx, m->d_nw, d_w,
                      stride, nullptr, 0, s->d_nsolve, s->d_rhs, s->nmax, nz, s->base_mode, 2, 0);
  SX_HIP (hipGetLastError ());
  if ((rc = band_solve (s, 2, st))) return rc;
  return recover (s, nz, stride, d_w, d_u, s->d_G, 2, 0, st);
}

extern "C" int sx_mass_inner (sx_mesh_t m, int nz, int stride, const double* d_a,
                              const double* d_b, double* d_out, void* stream)
{
  if (!m || nz < 1 || stride < m->nel * m->npnp || !d_a || !d_b || !d_out) {
    set_error ("sx_mass_inner: bad arguments");
    return SX_ERR_ARG;
  }
  hipLaunchKernelGGL (k_mass_inner, dim3 (nz), dim3 (256), 0, (hipStream_t) stream, m->npnp,
                      m->nel, m->d_geo, d_a, d_b, stride, d_out);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

extern "C" int sx_mass_weight (sx_mesh_t m, int nz, int stride, const double* d_in,
                               double* d_out, double scale, void* stream)
{
  if (!m || nz < 1 || stride < m->nel * m->npnp || !d_in || !d_out) {
    set_error ("sx_mass_weight: bad arguments");
    return SX_ERR_ARG;
  }
  const size_t tot = (size_t) nz * m->nel * m->npnp;
  hipLaunchKernelGGL (k_mass_weight, dim3 ((unsigned) ((tot + 255) / 256)), dim3 (256), 0,
                      (hipStream_t) stream, m->npnp, m->nel, m->d_geo, d_in, d_out, stride,
                      scale, nz);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}
