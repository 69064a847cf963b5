// sx_band.hip -- FP64 banded Cholesky factorisation and solve on gfx950,
// one workgroup per Fourier mode (LAPACK dpbtrf/dpbtrs('U') semantics,
// called by the reference at src/matrix.cpp:252 and src/field.cpp:693).
//
// The factor is assembled and factorised in LAPACK 'U' band storage
// (column j holds U(j-kd..j, j) at offsets 0..kd, ld = kd+1,
// veclib/lapack.h:107).  The solve U^T U x = b is then two identical
// forward sweeps:
//   forward  U^T y = b   on the factor itself,
//   backward U x = y     as a forward sweep on the reversed factor V,
//                        V(i', j') = U(n-1-j', n-1-i'), with reversed vectors.
// Each sweep processes blocks of B columns (B <= kd+1).  In block J
//   r_J = b_J - sum_{I<J} W_IJ^T y_I     (the band rows above the block)
//   y_J = W_JJ^-T r_J                    (W_JJ^-1 precomputed in set-up)
// so no sequential substitution remains inside a block.  For the sweep
// each block of the band is stored offset-major, X[J][q][c] = column j0+c
// at band offset q (q < kd - c: rows above the block; q >= kd - c: the
// inverted triangle), so a wavefront reads one coalesced row q for all B
// columns and lane c accumulates column c's dot product by itself -- no
// cross-lane reductions.  Eight wavefronts split the offsets and meet
// through LDS.  The solved values live in a circular LDS window of W >=
// kd+B rows, so LDS does not bound n or the number of right-hand sides.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <cstdlib>

#include "sx_internal.h"

namespace sx {
namespace {

// ---------------------------------------------------------------------------
// Factorisation (set-up): right-looking, the pivot row staged in LDS, the
// trailing window updated one column per wavefront.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_band_factor (
    double* band, size_t band_stride, int ld, const int* __restrict__ nsolve,
    int* status, const int* __restrict__ nelim)
{
  extern __shared__ double row[];
  __shared__ double s_dinv;
  __shared__ int s_bad;
  // nelim (optional): eliminate only the first nelim[mode] columns, leaving
  // the Schur complement of the rest in the trailing band
  const int mode = blockIdx.x, n = nsolve[mode], kd = ld - 1;
  const int ne = nelim ? nelim[mode] : n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6,
            nw = blockDim.x >> 6;
  double* A = band + (size_t) mode * band_stride;
  if (tid == 0) s_bad = 0;
  for (int j = 0; j < ne; j++) {
    if (tid == 0) {
      const double d = A[(size_t) j * ld + kd];
      if (!(d > 0.0)) { s_bad = j + 1; s_dinv = 0.0; }
      else { const double r = std::sqrt (d); A[(size_t) j * ld + kd] = r; s_dinv = 1.0 / r; }
    }
    __syncthreads ();
    if (s_bad) break;
    const double dinv = s_dinv;
    const int last = min (n - 1, j + kd), w = last - j;
    for (int t = tid; t < w; t += blockDim.x) {
      const int i = j + 1 + t;
      const size_t a = (size_t) i * ld + kd + j - i;
      const double v = A[a] * dinv;
      A[a] = v;
      row[t] = v;
    }
    __syncthreads ();
    for (int cl = wave; cl < w; cl += nw) {
      const int l = j + 1 + cl;
      const double rl = row[cl];
      double* col = A + (size_t) l * ld + kd - l;
      for (int ii = lane; ii <= cl; ii += 64) col[j + 1 + ii] -= row[ii] * rl;
    }
    __syncthreads ();
  }
  if (tid == 0 && s_bad) status[mode] = s_bad;
}

// Invert the BxB upper-triangular diagonal blocks in place; one wavefront
// per (block, mode), lane c computes column c by back substitution.
__global__ __launch_bounds__(64) void k_band_tinv (
    double* band, size_t band_stride, int ld, int B,
    const int* __restrict__ nsolve)
{
  __shared__ double T[64 * 64];
  const int mode = blockIdx.y, n = nsolve[mode], kd = ld - 1;
  const int j0 = blockIdx.x * B;
  if (j0 >= n) return;
  const int bs = min (B, n - j0), c = threadIdx.x;
  double* A = band + (size_t) mode * band_stride;
  if (c < bs)
    for (int t = 0; t <= c; t++) T[t * 64 + c] = A[(size_t) (j0 + c) * ld + kd - c + t];
  __syncthreads ();
  if (c < bs) {
    // column c of inv(T) overwrites column c of the band (T stays in LDS);
    // the lane re-reads only its own writes
    double* X = A + (size_t) (j0 + c) * ld + kd - c;   // X[t] = inv(T)[t][c]
    X[c] = 1.0 / T[c * 64 + c];
    for (int t = c - 1; t >= 0; t--) {
      double s = 0.0;
      for (int k = t + 1; k <= c; k++) s += T[t * 64 + k] * X[k];
      X[t] = -s / T[t * 64 + t];
    }
  }
}

// ---------------------------------------------------------------------------
// Set-up: reversed band and offset-major sweep blocks
// ---------------------------------------------------------------------------
// V[i'][o] = U(i, i+kd-o), i = n-1-i'  (the backward sweep U x = y is a
// forward sweep on V with reversed vectors).
__global__ void k_band_reverse (const double* __restrict__ U, double* V,
                                size_t band_stride, int ld,
                                const int* __restrict__ nsolve)
{
  const int mode = blockIdx.y, n = nsolve[mode], kd = ld - 1;
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) n * ld) return;
  const int ip = (int) (idx / ld), o = (int) (idx % ld);
  const int i = n - 1 - ip, jj = i + kd - o;
  const double* Um = U + (size_t) mode * band_stride;
  V[(size_t) mode * band_stride + idx] = (jj < n) ? Um[(size_t) jj * ld + o] : 0.0;
}

// X[J][q][c] = band column J*B+c at offset q (zero past the last column).
__global__ void k_band_blockify (const double* __restrict__ U, double* X,
                                 size_t band_stride, size_t sw_stride, int ld,
                                 int B, const int* __restrict__ nsolve)
{
  const int mode = blockIdx.y, n = nsolve[mode];
  const int nblk = (n + B - 1) / B;
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) nblk * ld * B) return;
  const int J = (int) (idx / ((size_t) ld * B));
  const int rem = (int) (idx % ((size_t) ld * B)), q = rem / B, c = rem % B;
  const int j = J * B + c;
  X[(size_t) mode * sw_stride + idx] =
      (j < n) ? U[(size_t) mode * band_stride + (size_t) j * ld + q] : 0.0;
}

// ---------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------
constexpr int SW_NW = 8;   // wavefronts per sweep workgroup

template <int B, int QS, int NR>
struct Sweep {
  static constexpr int NT = SW_NW * 64;
  double v[QS];            // this wave's rows q = w*QS + k of the block, lane = column
  double bin[NR];          // right-hand side of column lane
};

template <int B, int QS, int NR>
__device__ __forceinline__ void sweep_load (Sweep<B, QS, NR>& S,
                                            const double* __restrict__ X,
                                            const double* __restrict__ in,
                                            int nmax, int ld, int n, int J,
                                            int reversed, int w, int lane)
{
  const int c = lane & (B - 1);
  const double* blk = X + (size_t) J * ld * B + (size_t) w * QS * B + c;
  if (w * QS + QS <= ld) {          // wave-uniform: all rows inside the band
#pragma unroll
    for (int k = 0; k < QS; k++) S.v[k] = blk[k * B];
  } else {                          // last wave: rows past the band re-read row ld-1
#pragma unroll
    for (int k = 0; k < QS; k++) S.v[k] = blk[min (k, ld - 1 - w * QS) * B];
  }
  const int j = min (J * B + c, n - 1);
  const int gi = reversed ? n - 1 - j : j;
#pragma unroll
  for (int r = 0; r < NR; r++) S.bin[r] = in[(size_t) r * nmax + gi];
}

template <int B, int QS, int NR>
__device__ __forceinline__ void sweep_step (const Sweep<B, QS, NR>& S, double* win,
                                            int W, double* part, double* rpriv,
                                            double* part2, double* out, int nmax,
                                            int ld, int n, int J, int reversed,
                                            int w, int lane)
{
  const int kd = ld - 1, c = lane, j0 = J * B, Wm = W - 1;
  // far part: rows above the block, q < kd - c.  The window holds every
  // row twice (slots s and s+W), so the QS consecutive rows of this wave
  // are read at consecutive addresses from slot (first row mod W).
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0.0;
  const int s0 = (j0 + c - kd + w * QS) & Wm;         // row j0+c-kd+q0 (may be < 0)
  const double* wr = win + s0;
  if (w * QS + QS - 1 < kd - (B - 1)) {               // wave-uniform: all rows far
#pragma unroll
    for (int k = 0; k < QS; k++)
#pragma unroll
      for (int r = 0; r < NR; r++) acc[r] += S.v[k] * wr[r * 2 * W + k];
  } else {
#pragma unroll
    for (int k = 0; k < QS; k++) {
      const int q = w * QS + k;
      const double u = (q < kd - c) ? S.v[k] : 0.0;
#pragma unroll
      for (int r = 0; r < NR; r++) acc[r] += u * wr[r * 2 * W + k];
    }
  }
  if (c < B)
#pragma unroll
    for (int r = 0; r < NR; r++) part[(w * B + c) * NR + r] = acc[r];
  __syncthreads ();
  // triangle rows q in [kd-B+1, kd]: the waves holding them form the full
  // r vector privately, then apply their rows of inv(W_JJ)^T
  const bool tri = (w * QS + QS - 1 >= kd - (B - 1)) && (w * QS <= kd);
  double acc2[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc2[r] = 0.0;
  if (tri) {
    if (c < B)
#pragma unroll
      for (int r = 0; r < NR; r++) {
        double rr = S.bin[r];
#pragma unroll
        for (int ww = 0; ww < SW_NW; ww++) rr -= part[(ww * B + c) * NR + r];
        rpriv[(w * B + c) * NR + r] = rr;
      }
#pragma unroll
    for (int k = 0; k < QS; k++) {
      const int q = w * QS + k, t = c - kd + q;
      const bool ok = t >= 0 && q <= kd && c < B;
      const double u = ok ? S.v[k] : 0.0;
      const int tc = min (max (t, 0), B - 1);
#pragma unroll
      for (int r = 0; r < NR; r++) acc2[r] += u * rpriv[(w * B + tc) * NR + r];
    }
  }
  if (c < B)
#pragma unroll
    for (int r = 0; r < NR; r++) part2[(w * B + c) * NR + r] = acc2[r];
  __syncthreads ();
  if (w == 0 && c < B && j0 + c < n) {
    const int gi = reversed ? n - 1 - (j0 + c) : j0 + c;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      double y = 0.0;
#pragma unroll
      for (int ww = 0; ww < SW_NW; ww++) y += part2[(ww * B + c) * NR + r];
      const int slot = (j0 + c) & Wm;
      win[r * 2 * W + slot] = y;
      win[r * 2 * W + slot + W] = y;
      out[(size_t) r * nmax + gi] = y;
    }
  }
  __syncthreads ();
}

template <int B, int QS, int NR>
__global__ __launch_bounds__(SW_NW * 64) void k_band_sweeps (
    const double* __restrict__ XF, const double* __restrict__ XR,
    size_t sw_stride, int ld, int W, const int* __restrict__ nsolve,
    double* rhs, int nmax)
{
  extern __shared__ double sm[];
  const int mode = blockIdx.x, n = nsolve[mode];
  if (n == 0) return;
  double* win = sm;                          // [NR][2W] (each row stored twice)
  double* part = win + 2 * NR * W;           // [NW][B][NR]
  double* rpriv = part + SW_NW * B * NR;     // [NW][B][NR]
  double* part2 = rpriv + SW_NW * B * NR;    // [NW][B][NR]
  double* bg = rhs + (size_t) mode * NR * nmax;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane (tid >> 6);
  const int nblk = (n + B - 1) / B;

  for (int pass = 0; pass < 2; pass++) {
    const double* X = (pass == 0 ? XF : XR) + (size_t) mode * sw_stride;
    for (int i = tid; i < 2 * NR * W; i += SW_NW * 64) win[i] = 0.0;
    __syncthreads ();
    // the next block's rows are loaded while the current one is processed;
    // prefetches are unconditional (block index clamped) so waiting for the
    // current block never drains them
    Sweep<B, QS, NR> sa, sb;
    sweep_load (sa, X, bg, nmax, ld, n, 0, pass, w, lane);
    for (int J = 0; J < nblk; J += 2) {
      sweep_load (sb, X, bg, nmax, ld, n, min (J + 1, nblk - 1), pass, w, lane);
      sweep_step (sa, win, W, part, rpriv, part2, bg, nmax, ld, n, J, pass, w, lane);
      sweep_load (sa, X, bg, nmax, ld, n, min (J + 2, nblk - 1), pass, w, lane);
      if (J + 1 < nblk)
        sweep_step (sb, win, W, part, rpriv, part2, bg, nmax, ld, n, J + 1, pass, w, lane);
    }
  }
}

template <int B>
size_t sweep_lds (int W, int NR)
{
  return sizeof (double) * (2 * (size_t) NR * W + 3 * (size_t) SW_NW * B * NR);
}

int window (int nband, int B)
{
  int W = 1;
  while (W < nband - 1 + B) W <<= 1;
  return W;
}

template <int B, int QS>
int launch_sweeps (const sx_msys* s, hipStream_t st)
{
  const int W = window (s->m->nband, B);
  const size_t lds = sweep_lds<B> (W, 2);
  auto kern = k_band_sweeps<B, QS, 2>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (s->nmodes), dim3 (SW_NW * 64), lds, st, s->d_swF,
                      s->d_swR, s->sw_stride, s->m->nband, W, s->d_nsolve, s->d_rhs,
                      s->nmax);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// ---------------------------------------------------------------------------
// Right-looking panel sweeps (default path, band offsets kd <= 511)
// ---------------------------------------------------------------------------
// The factor is cut into 64x64 blocks: row block J couples to column blocks
// J..J+D, D = floor((kd+63)/64).  Panel J holds the inverse of the diagonal
// block, T_J = U_JJ^-1, and the D off-diagonal blocks
// U_{J,J+d} = U(64J+t, 64(J+d)+c) (zero outside the band).  A block is
// stored as row pairs [t/2][c][2] so that lane c reads rows t and t+1 of
// its column with one 16-byte load and a wavefront reads 1 KB contiguously.
// A forward sweep U^T y = b is
//   y_J = T_J^T r_J,    r_K -= U_{J,K}^T y_J   for K = J+1..J+D,
// with r_K initialised to b_K.  Wave w (w < D) owns the pending column
// block J+w at step J and holds the next step's 64x64 block in registers,
// each pair reloaded right after its use, so a whole step's loads are in
// flight behind the compute.  Every wave issues the same loads in the same
// order (waves without a block read a zero panel), which keeps the memory
// wait counts exact; the triangle is shared by all waves, TRW rows each.
// One extra zero panel follows the nb panels of each mode.
constexpr int PB = 64;             // panel block
constexpr int PMAXD = 8;           // off-diagonal blocks per panel supported

__device__ __forceinline__ void lds_sync ()
{
  // LDS-only workgroup barrier: global loads stay in flight across it
  asm volatile ("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// the value of lane l (wave-uniform l) in every lane
__device__ __forceinline__ double lane_bcast (double v, int l)
{
  const long long x = __double_as_longlong (v);
  const int lo = __builtin_amdgcn_readlane ((int) (x & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane ((int) (x >> 32), l);
  return __longlong_as_double (((long long) hi << 32) | (unsigned) lo);
}

// the sums over the 64 lanes of p[0..7]: lane l returns the sum of p[l >> 3]
// (a halving exchange: across lane halves, quarters and eighths each lane
// keeps half of its values and adds its partner's, then the eight lanes
// holding one value sum it)
__device__ __forceinline__ double lane_reduce8 (const double (&p)[8], int lane)
{
  const bool h32 = lane & 32, h16 = lane & 16, h8 = lane & 8;
  double a[4], b[2];
#pragma unroll
  for (int j = 0; j < 4; j++)
    a[j] = (h32 ? p[j + 4] : p[j]) + __shfl_xor (h32 ? p[j] : p[j + 4], 32);
#pragma unroll
  for (int j = 0; j < 2; j++)
    b[j] = (h16 ? a[j + 2] : a[j]) + __shfl_xor (h16 ? a[j] : a[j + 2], 16);
  double c = (h8 ? b[1] : b[0]) + __shfl_xor (h8 ? b[0] : b[1], 8);
  c += __shfl_xor (c, 4);
  c += __shfl_xor (c, 2);
  c += __shfl_xor (c, 1);
  return c;
}

// the same for p[0..3]: lane l returns the sum of p[l >> 4]
__device__ __forceinline__ double lane_reduce4 (const double (&p)[4], int lane)
{
  const bool h32 = lane & 32, h16 = lane & 16;
  double a[2];
#pragma unroll
  for (int j = 0; j < 2; j++)
    a[j] = (h32 ? p[j + 2] : p[j]) + __shfl_xor (h32 ? p[j] : p[j + 2], 32);
  double c = (h16 ? a[1] : a[0]) + __shfl_xor (h16 ? a[0] : a[1], 16);
  c += __shfl_xor (c, 8);
  c += __shfl_xor (c, 4);
  c += __shfl_xor (c, 2);
  c += __shfl_xor (c, 1);
  return c;
}

__device__ __forceinline__ size_t pidx (int t, int c)
{
  return ((size_t) (t >> 1) * PB + c) * 2 + (t & 1);
}

// T_J = inverse of the diagonal block (zero outside the band, identity
// past row n-1), one wavefront per (J, mode); lane c back-substitutes
// column c.  J = nb writes the zero panel.
__global__ __launch_bounds__(64) void k_panel_tinv (
    const double* __restrict__ band, size_t band_stride, int ld,
    const int* __restrict__ nsolve, double* P, size_t pstride, int D)
{
  __shared__ double T[PB * (PB + 1)], X[PB * (PB + 1)];
  const int mode = blockIdx.y, n = nsolve[mode], kd = ld - 1;
  const int J = blockIdx.x, j0 = J * PB, c = threadIdx.x;
  double* out = P + (size_t) mode * pstride + (size_t) J * (D + 1) * PB * PB;
  const int bs = max (0, min (PB, n - j0));
  const double* A = band + (size_t) mode * band_stride;
  for (int t = 0; t < PB; t++) {
    double v = 0.0;
    if (c < bs && t <= c && c - t <= kd) v = A[(size_t) (j0 + c) * ld + kd - c + t];
    T[t * (PB + 1) + c] = v;
  }
  if (c >= bs) T[c * (PB + 1) + c] = 1.0;
  __syncthreads ();
  X[c * (PB + 1) + c] = 1.0 / T[c * (PB + 1) + c];
  for (int t = c - 1; t >= 0; t--) {
    double sum = 0.0;
    for (int k = t + 1; k <= c; k++) sum += T[t * (PB + 1) + k] * X[k * (PB + 1) + c];
    X[t * (PB + 1) + c] = -sum / T[t * (PB + 1) + t];
  }
  for (int t = 0; t < PB; t++)
    out[pidx (t, c)] = (c < bs && t <= c) ? X[t * (PB + 1) + c] : 0.0;
}

// Off-diagonal panel blocks U(64J+t, 64(J+d)+c), d = 1..D, J = 0..nb
// (J = nb: the zero panel).
__global__ void k_panel_offdiag (const double* __restrict__ band, size_t band_stride,
                                 int ld, const int* __restrict__ nsolve, double* P,
                                 size_t pstride, int D, int nbp)
{
  const int mode = blockIdx.y, n = nsolve[mode], kd = ld - 1;
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) nbp * D * PB * PB) return;
  const int J = (int) (idx / ((size_t) D * PB * PB));
  const int rem = (int) (idx % ((size_t) D * PB * PB));
  const int d = 1 + rem / (PB * PB), t = (rem % (PB * PB)) / PB, c = rem % PB;
  const int i = J * PB + t, j = (J + d) * PB + c;
  double v = 0.0;
  if (j < n && j - i <= kd) v = band[(size_t) mode * band_stride + (size_t) j * ld + kd - (j - i)];
  P[(size_t) mode * pstride + ((size_t) J * (D + 1) + d) * PB * PB + pidx (t, c)] = v;
}

// wavefronts per pending column block: two (rows 0-31 and 32-63) while
// 2D waves fit the register file at two waves per SIMD, else one
constexpr int panel_h (int D) { return D > 0 ? 2 : 2; }
constexpr int panel_nw (int D) { return panel_h (D) * D < 4 ? 4 : panel_h (D) * D; }

template <int NR, int D>
struct PanelSweep {
  static constexpr int H = panel_h (D);
  static constexpr int NW = panel_nw (D);
  static constexpr int TRW = ((PB + NW - 1) / NW + 1) & ~1;  // triangle rows per wave
  static constexpr int HB = PB / H;                          // block rows per wave
};

// Twisted solve.  The rows are split into T = [0, m), S = [m, m+64D) and
// B = [m+64D, n) (|S| >= kd, so T and B do not couple).  B is eliminated
// from its far end (the bottom piece: rows [m, n) reversed and padded in
// front by p identity rows to whole blocks, factorised through B only),
// and S's Schur complement A_SS - U_BS^T U_BS closes the top piece, rows
// [0, m+64D), which is factorised completely.  Then
//   forward   y_B = U_BB^-T b_B  and  top U^T y = (b_T, b_S - U_BS^T y_B),
//   backward  top U x = y        and  x_B = U_BB^-1 (y_B - U_BS x_S),
// so each sweep runs the two pieces side by side and meets at S: a chain
// of about nb + 2D block steps instead of 2 nb.  A launch runs a segment
// of blocks of each pass; a segment that stops early writes its pending
// window r_{J1..J1+D-1} back into the right-hand side, where the next
// segment reads it as its b.  Per-mode parameters, TW ints per mode:
// (TW, TW_* in sx_internal.h)

// One sweep launch's work for this workgroup: its mode and right-hand sides,
// the piece's size, panels and block ranges per pass ([f0, f1) forward,
// [b0, b1) backward; wf/wb: hand the window on).  G workgroups share a mode
// (and piece), NR of its G*NR right-hand sides each; they get block indices
// congruent mod 8, i.e. the same XCD and L2.  False: nothing to do.
struct SweepSeg {
  int mode, n, f0, f1, b0, b1, rs, role, chain, grp;
  int l0, l1;                        // left-looking backward: blocks [l0, l1), descending
  bool wf, wb;
  size_t poff;
  double* bg;
};

// roles: workgroups per chain (a chain: one piece of one mode, NR right-hand
// sides); they also get block indices congruent mod 8
__device__ __forceinline__ bool sweep_seg (SweepSeg& g, int NR, int D,
                                           const int* __restrict__ nsolve, double* rhs,
                                           int nmax, int nmodes, int G,
                                           const int* __restrict__ tw, int stage, size_t boff,
                                           double* rhsb, int nmaxb, int roles = 1)
{
  const int per = (nmodes + 7) / 8 * 8 * G * roles;
  const int piece = blockIdx.x / per, b = blockIdx.x % per;
  const int mode = (b / (8 * G * roles)) * 8 + b % 8, gr = (b / 8) % (G * roles);
  const int grp = gr / roles;
  if (mode >= nmodes) return false;
  g.role = gr % roles;
  g.grp = grp;
  g.chain = (piece * nmodes + mode) * G + grp;
  g.mode = mode;
  g.f0 = g.f1 = g.b0 = g.b1 = g.l0 = g.l1 = 0;
  g.wf = g.wb = false;
  g.poff = 0;
  g.bg = rhs + ((size_t) mode * G + grp) * NR * nmax;
  g.rs = nmax;
  if (stage == 0) {
    g.n = nsolve[mode];
    g.f1 = g.b1 = g.l1 = (g.n + PB - 1) / PB;
  } else {
    const int* t = tw + TW * mode;
    const int nbT = t[TW_NBT], nbB = t[TW_NBB];
    if (piece == 0) {
      g.n = t[TW_NTOP];
      if (stage == 1) { g.f1 = nbT; g.wf = true; }
      if (stage == 2) { g.f0 = nbT; g.f1 = nbT + D; g.b1 = D; g.wb = true; g.l0 = nbT; g.l1 = nbT + D; }
      if (stage == 3) { g.b0 = D; g.b1 = nbT + D; g.l1 = nbT; }
    } else {
      g.n = t[TW_NBOT];
      g.poff = boff;
      g.bg = rhsb + ((size_t) mode * G + grp) * NR * nmaxb;
      g.rs = nmaxb;
      if (stage == 1) { g.f1 = nbB; g.wf = true; }
      if (stage == 3) { g.b1 = nbB + D; g.l1 = nbB; }
    }
  }
  return g.n > 0;
}

template <int NR, int D>
__device__ __forceinline__ void k_panel_sweeps_body (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd)
{
  constexpr int NW = PanelSweep<NR, D>::NW, TRW = PanelSweep<NR, D>::TRW;
  constexpr int HB = PanelSweep<NR, D>::HB, H = PanelSweep<NR, D>::H;
  extern __shared__ double sm[];
  double* rbuf = sm;                                // [2][D][64][NR] pending r, by half
  double* part = rbuf + 2 * D * PB * NR;            // [NW][64][NR] triangle partials
  SweepSeg g;
  if (!sweep_seg (g, NR, D, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb, nmaxb)) return;
  const int mode = g.mode, n = g.n, f0 = g.f0, f1 = g.f1, b0 = g.b0, b1 = g.b1, rs = g.rs;
  const bool wf = g.wf, wb = g.wb;
  const size_t poff = g.poff;
  double* bg = g.bg;
  const int nb = (n + PB - 1) / PB;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
  const int slot = w / H, half = w % H;             // pending block J+slot, rows half*HB..
  const bool owner = slot < D;
  const size_t PS = (size_t) (D + 1) * PB * PB;     // doubles per panel
  double u[HB], tr[TRW], bp[NR];

  for (int pass = 0; pass < 2; pass++) {
    const int J0 = pass ? b0 : f0, J1 = pass ? b1 : f1;
    const bool wout = pass ? wb : wf;
    if (J0 >= J1) continue;
    const double* P = (pass == 0 ? PF : PR) + (size_t) mode * pstride + poff;
    auto gidx = [&] (int j) __attribute__ ((always_inline)) {
      j = min (j, n - 1);
      return pass ? n - 1 - j : j;
    };
    // Panel loads go through a buffer descriptor over the piece's nb+1
    // panels: a lane whose row pair has no entry inside the band (the lower
    // half of T_J, the tail of the last off-diagonal block, waves without a
    // block) is given an offset past the range and reads 0 with no memory
    // access -- the stored values there are 0 as well.  Every wave still
    // issues the same loads in the same order.
    const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc (
        (void*) P, (short) 0, (int) (sizeof (double) * (size_t) (nb + 1) * PS), 0x00020000);
    constexpr int OOB = 0x7ffffff0;
    auto pld = [&] (int boff) __attribute__ ((always_inline)) {
      typedef int v4i __attribute__ ((ext_vector_type (4)));
      const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (prs, boff, 0, 0);
      return __builtin_bit_cast (double2, t);
    };
    // byte offset of this wave's rows of block U_{J, J+1+slot}, lane's column;
    // row pair t of it reaches into the band for this lane iff lane <= ulim + t
    auto uoff = [&] (int J) __attribute__ ((always_inline)) {
      return (int) (sizeof (double) * ((size_t) J * PS + (size_t) (slot + 1) * PB * PB +
                                       (size_t) half * HB * PB + 2 * lane));
    };
    const int ulim = owner ? kd - PB * (slot + 1) + half * HB + 1 : -2 * PB;
    // triangle rows [w*TRW, w*TRW+TRW) of T_J (upper triangular: row pair t
    // is needed by lanes >= t; rows past 63 by none)
    auto toff = [&] (int J, int k) __attribute__ ((always_inline)) {
      const int t = min (w * TRW + k, PB - 2);
      return (int) (sizeof (double) * ((size_t) J * PS + pidx (t, lane)));
    };
    auto tneed = [&] (int k) __attribute__ ((always_inline)) {
      const int t = w * TRW + k;
      return t < PB && lane >= t;
    };
    auto load_b = [&] (int K) __attribute__ ((always_inline)) {
#pragma unroll
      for (int r = 0; r < NR; r++) bp[r] = bg[(size_t) r * rs + gidx (K * PB + lane)];
    };
    // triangle partials of block J, then prefetch of T_{J+1}
    auto tri = [&] (int J) __attribute__ ((always_inline)) {
      const double* r0 = rbuf + (size_t) (J % D) * PB * NR;
      const double* r1 = r0 + (size_t) D * PB * NR;
      const int Jn = min (J + 1, nb - 1);
      double pp[NR];
#pragma unroll
      for (int r = 0; r < NR; r++) pp[r] = 0.0;
#pragma unroll
      for (int k = 0; k < TRW; k += 2) {
        const double t0 = tr[k], t1 = tr[k + 1];
        const int ta = min (w * TRW + k, PB - 2);
        const double wt = (w * TRW + k < PB) ? 1.0 : 0.0;   // rows past 63 add nothing
#pragma unroll
        for (int r = 0; r < NR; r++)
          pp[r] += wt * (t0 * (r0[ta * NR + r] + r1[ta * NR + r]) +
                         t1 * (r0[(ta + 1) * NR + r] + r1[(ta + 1) * NR + r]));
        // reload behind the use (offset tied to the sums, as for the blocks)
        int off = toff (Jn, k);
#pragma unroll
        for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(pp[r]));
        const double2 v = pld (tneed (k) ? off : OOB);
        tr[k] = v.x;
        tr[k + 1] = v.y;
      }
#pragma unroll
      for (int r = 0; r < NR; r++) part[(w * PB + lane) * NR + r] = pp[r];
    };
    // every wave sums the triangle partials itself: lane c holds y_J[c],
    // which the update reads back by lane broadcast (v_readlane), so no
    // barrier and no LDS round trip is needed; wave 0 writes y_J out
    double yreg[NR];
    auto finish = [&] (int J) __attribute__ ((always_inline)) {
      const int j = J * PB + lane;
#pragma unroll
      for (int r = 0; r < NR; r++) {
        double y = 0.0;
#pragma unroll
        for (int g = 0; g < NW; g++) y += part[(g * PB + lane) * NR + r];
        yreg[r] = y;
        if (w == 0 && j < n) bg[(size_t) r * rs + gidx (j)] = y;
      }
    };

    // r of blocks J0..J0+D-1 from b (pass 1 of a whole solve: the forward
    // result, reversed); the second half's partial sums start at zero
    if (owner) {
      const int K = J0 + slot;
      const bool in = K * PB + lane < n;
#pragma unroll
      for (int r = 0; r < NR; r++)
        rbuf[((size_t) (half * D + K % D) * PB + lane) * NR + r] =
            (in && half == 0) ? bg[(size_t) r * rs + gidx (K * PB + lane)] : 0.0;
    }
    // prefetch in the steady-state order: T_J0 rows, U_{J0,J0+1+slot}, b of
    // block J0+D
#pragma unroll
    for (int k = 0; k < TRW; k += 2) {
      const double2 v = pld (tneed (k) ? toff (J0, k) : OOB);
      tr[k] = v.x;
      tr[k + 1] = v.y;
    }
    {
      const int ub = uoff (J0);
#pragma unroll
      for (int t = 0; t < HB; t += 2) {
        const double2 v = pld (lane <= ulim + t ? ub + t * PB * (int) sizeof (double) : OOB);
        u[t] = v.x;
        u[t + 1] = v.y;
      }
    }
    load_b (J0 + D);
    // drain once per pass: the loop's wait counts then follow its own,
    // fixed load order
    __builtin_amdgcn_s_waitcnt (0);
    __syncthreads ();

    // step J: y_J from the triangle, then r_{J+1+slot} -= U_{J,J+1+slot}^T y_J
    for (int J = J0; J < J1; J++) {
      tri (J);
      lds_sync ();
      finish (J);
      // U_{J+1,J+2+slot} is loaded pair by pair right behind the use: the
      // load offset is tied to the running sums, so the loads cannot be
      // hoisted above the FMAs (which would need a second register buffer)
      {
        const int K = J + 1 + slot;
        const int row0 = half * HB;                  // this wave's rows of y_J
        double acc0[NR], acc1[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) acc0[r] = acc1[r] = 0.0;
        const int un = uoff (J + 1);
#pragma unroll
        for (int t = 0; t < HB; t += 2) {
          const double u0 = u[t], u1 = u[t + 1];
#pragma unroll
          for (int r = 0; r < NR; r++) {
            acc0[r] += u0 * lane_bcast (yreg[r], row0 + t);
            acc1[r] += u1 * lane_bcast (yreg[r], row0 + t + 1);
          }
          int off = un + t * PB * (int) sizeof (double);
#pragma unroll
          for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(acc0[r]), "v"(acc1[r]));
          const double2 v = pld (lane <= ulim + t ? off : OOB);
          u[t] = v.x;
          u[t + 1] = v.y;
        }
        double acc[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
          acc[r] = acc0[r] + acc1[r];
          // bp is read only after the sums (its wait must not move up)
          asm ("" : "+v"(bp[r]) : "v"(acc[r]));
        }
        if (owner) {
          double* rp = rbuf + ((size_t) (half * D + K % D) * PB + lane) * NR;
          if (slot == D - 1) {                // block K enters the window now
            const bool in = K * PB + lane < n && half == 0;
#pragma unroll
            for (int r = 0; r < NR; r++) rp[r] = (in ? bp[r] : 0.0) - acc[r];
          } else {
#pragma unroll
            for (int r = 0; r < NR; r++) rp[r] -= acc[r];
          }
        }
        // b of the block entering next step, loaded after the use of bp
        {
          int off = 0;
#pragma unroll
          for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(acc[r]), "v"(bp[r]));
          const int gi = gidx ((K + 1) * PB + lane) + off;
#pragma unroll
          for (int r = 0; r < NR; r++) bp[r] = bg[(size_t) r * rs + gi];
        }
      }
      lds_sync ();
    }
    // a segment that stops early hands its pending window on through b
    if (wout && owner && half == 0) {
      const int K = J1 + slot;
      if (K * PB + lane < n) {
        const double* r0 = rbuf + ((size_t) (K % D) * PB + lane) * NR;
        const double* r1 = r0 + (size_t) D * PB * NR;
#pragma unroll
        for (int r = 0; r < NR; r++) bg[(size_t) r * rs + gidx (K * PB + lane)] = r0[r] + r1[r];
      }
    }
    // the next pass reads what this one wrote to global memory: written and
    // read by waves of this workgroup only (one CU, one L1), so a
    // workgroup-scope fence orders it; a device-scope fence would write back
    // and invalidate the XCD's L2 under the other workgroups' panel streams
    __threadfence_block ();
    __syncthreads ();
  }
}

template <int NR, int D>
__global__ __launch_bounds__(panel_nw (D) * 64) void k_panel_sweeps (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd)
{
  k_panel_sweeps_body<NR, D> (PF, PR, pstride, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb, nmaxb, kd);
}

// Moves between the right-hand side (rows in global order) and the bottom
// piece's vectors (rows [m, n) reversed, p zero rows in front), per
// (row, column, mode):
//   0  gather   rb[q] = b[n-1-(q-p)] for B rows, zero for pad and S rows
//   1  combine  b[m+a] += rb[nbot-1-a]    (S: the two pieces' residuals)
//   2  x_S      rb[nbot-1-a] = b[m+a]
//   3  scatter  b[n-1-(q-p)] = rb[q]      for B rows
__device__ __forceinline__ void k_twist_move_body (double* rhs, int nmax, double* rhsb,
                                                   int nmaxb, int nr, const int* __restrict__ tw,
                                                   int S, int what, int mode, int col)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int* t = tw + TW * mode;
  const int p = t[TW_PAD], m = t[TW_M], nbot = t[TW_NBOT], n = t[TW_N];
  const int nbr = nbot - p - S;                     // B rows
  double* b = rhs + ((size_t) mode * nr + col) * nmax;
  double* rb = rhsb + ((size_t) mode * nr + col) * nmaxb;
  switch (what) {
  case 0:
    if (q < nbot) rb[q] = (q >= p && q < p + nbr) ? b[n - 1 - (q - p)] : 0.0;
    break;
  case 1:
    if (q < S) b[m + q] += rb[nbot - 1 - q];
    break;
  case 2:
    if (q < S) rb[nbot - 1 - q] = b[m + q];
    break;
  default:
    if (q >= p && q < p + nbr) b[n - 1 - (q - p)] = rb[q];
    break;
  }
}

__global__ void k_twist_move (double* rhs, int nmax, double* rhsb, int nmaxb, int nr,
                              const int* __restrict__ tw, int S, int what)
{
  k_twist_move_body (rhs, nmax, rhsb, nmaxb, nr, tw, S, what, blockIdx.z, blockIdx.y);
}

// The bottom piece of a twisted solve: rows [m, n) of the band reversed,
// p identity rows in front, R(q, q') = A(n-1-(q-p), n-1-(q'-p)).
__global__ void k_twist_rbuild (const double* __restrict__ A, size_t astride, double* R,
                                size_t rstride, int ld, const int* __restrict__ tw)
{
  const int mode = blockIdx.y, kd = ld - 1;
  const int* t = tw + TW * mode;
  const int p = t[TW_PAD], nbot = t[TW_NBOT], n = t[TW_N];
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) nbot * ld) return;
  const int q = (int) (idx / ld), o = (int) (idx % ld);
  double v = 0.0;
  if (q < p) v = (o == kd) ? 1.0 : 0.0;
  else {
    const int r = n - 1 - (q - p) + kd - o;        // global row of R(q-kd+o, q)
    if (r < n && q - kd + o >= p) v = A[(size_t) mode * astride + (size_t) r * ld + o];
  }
  R[(size_t) mode * rstride + idx] = v;
}

// S's Schur complement, left in the bottom piece's trailing block by its
// partial factorisation, into the top piece's band rows [m, m+S); the
// bottom piece's S block becomes the identity (its backward sweep takes
// x_S as given).  Thread (be, o) owns entry (m+al, m+be), al = be-kd+o.
__global__ void k_twist_sback (double* A, size_t astride, double* R, size_t rstride,
                               int ld, const int* __restrict__ tw, int S)
{
  const int mode = blockIdx.y, kd = ld - 1;
  const int* t = tw + TW * mode;
  const int m = t[TW_M], nbot = t[TW_NBOT];
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) S * ld) return;
  const int be = (int) (idx / ld), o = (int) (idx % ld), al = be - kd + o;
  if (al < 0) return;
  double* rv = R + (size_t) mode * rstride + (size_t) (nbot - 1 - al) * ld + o;
  A[(size_t) mode * astride + (size_t) (m + be) * ld + o] = *rv;
  *rv = (o == kd) ? 1.0 : 0.0;
}

int panel_D (int nband)
{
  const int kd = nband - 1;
  return std::max (1, (kd + PB - 1) / PB);
}

// NR right-hand sides per workgroup, G workgroups per mode
template <int NR, int D>
int launch_panel_d (const sx_msys* s, int G, int stage, hipStream_t st)
{
  constexpr int NW = PanelSweep<NR, D>::NW;
  const size_t lds = sizeof (double) * ((size_t) PB * NR * (2 * D + NW));
  auto kern = k_panel_sweeps<NR, D>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  const int pieces = (stage == 1 || stage == 3) ? 2 : 1;
  const int nblocks = (s->nmodes + 7) / 8 * 8 * G * pieces;
  hipLaunchKernelGGL (kern, dim3 (nblocks), dim3 (NW * 64), lds, st, s->d_swF, s->d_swR,
                      s->sw_stride, s->d_nsolve, s->d_rhs, s->nmax, s->nmodes, G,
                      s->d_tw, stage, s->tw_boff, s->d_rhsb, s->nmaxb, s->m->nband - 1);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

template <int NR>
int launch_panel (const sx_msys* s, int G, int stage, hipStream_t st)
{
  switch (s->sw_D) {
  case 1: return launch_panel_d<NR, 1> (s, G, stage, st);
  case 2: return launch_panel_d<NR, 2> (s, G, stage, st);
  case 3: return launch_panel_d<NR, 3> (s, G, stage, st);
  case 4: return launch_panel_d<NR, 4> (s, G, stage, st);
  case 5: return launch_panel_d<NR, 5> (s, G, stage, st);
  case 6: return launch_panel_d<NR, 6> (s, G, stage, st);
  default: break;
  }
  // more than 2 right-hand sides per workgroup need D <= 6 (msys_reserve_rhs)
  if constexpr (NR <= 2) {
    if (s->sw_D == 7) return launch_panel_d<NR, 7> (s, G, stage, st);
    if (s->sw_D == 8) return launch_panel_d<NR, 8> (s, G, stage, st);
  }
  set_error ("panel sweep: unsupported block count");
  return SX_ERR_LIMIT;
}

// ---------------------------------------------------------------------------
// Unit-diagonal panel sweeps (D = 3..5): two steps of panel loads in flight
// ---------------------------------------------------------------------------
// With U = diag(U_JJ) V, the off-diagonal blocks of V are W_{J,K} = T_J U_{J,K}
// and a forward sweep U^T y = b becomes
//   z_J = r_J,   r_K -= W_{J,K}^T z_J  (K = J+1..J+D),   y_J = T_J^T z_J,
// so the chain of dependent block steps carries no triangle: a step is the
// update alone and one barrier, and y_J is formed beside it.  The backward
// sweep is the same on the reversed factor (its own blocking, its own T_J).
// W_{J,J+D} fills its upper triangle (columns c <= kd - 64D + 63 in every
// row), the other blocks are full.
//
// Eight wavefronts (two per SIMD, up to 256 VGPRs each) share a panel by
// rows: wave w takes row pairs 4w..4w+3 (rows 8w..8w+7) of each of its D+1
// blocks, so per step it broadcasts only its eight entries of z_J into
// SGPRs and every pair costs two FMAs.  Each wave holds its pairs of panels
// J+1 and J+2 in registers and reloads a pair right behind its use, two
// steps ahead.  Pending right-hand sides live in LDS in a ring of D+1 entries
// (block K at K mod (D+1), so the block entering the window never shares an
// entry with the one being read), one slot per wave; the triangle partials
// are double-buffered and summed into y one step later.
constexpr int UW = 8;                               // wavefronts per workgroup
constexpr int UPW = 4;                              // row pairs per wave and block

// Left-looking backward sweep on the forward panels.  With the forward
// sweep's y_J = T_J^T z_J in b, V x = (T_J y_J)_J is solved block by block
// from the last:
//   x_J = T_J y_J - sum_{d=1..D} W_{J,J+d} x_{J+d},   J = l1-1 .. l0,
// reading the same panels as the forward sweep (block 0 of panel J is T_J,
// block d is W_{J,J+d}), in reverse order, so a backward sweep that follows
// its forward sweep finds the recently read panels in the memory-side cache.
// Wave w, lane c forms for its rows t = 8w..8w+7 the terms of column c,
// T_J(t,c) y_J[c] - sum_d W_{J,J+d}(t,c) x_{J+d}[c], and the lanes sum them
// (lane_reduce8); x_J goes to b and to a ring of D+1 blocks in LDS, from
// which every lane reads x_K[c] of the blocks it multiplies.  Blocks at or
// past l1 enter from b (x already known there: the separator of a twisted
// solve), and past the piece's end they are zero.  Panels two steps ahead
// stay in registers as in the forward sweep.
template <int NR, int D>
__device__ __forceinline__ void unit_backward_ll (
    const double* P, int n, int nb, double* bg, int rs, int l0, int l1, int lane, int w, int kd,
    double* hist, double2 (&ua)[UPW * (D + 1)], double2 (&ub)[UPW * (D + 1)], double (&ya)[NR],
    double (&yb)[NR])
{
  if (l0 >= l1) return;
  const size_t PS = (size_t) (D + 1) * PB * PB;
  const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc (
      (void*) P, (short) 0, (int) (sizeof (double) * (size_t) (nb + 1) * PS), 0x00020000);
  constexpr int OOB = 0x7f000000;
  auto pld = [&] (int voff, int soff) __attribute__ ((always_inline)) {
    typedef int v4i __attribute__ ((ext_vector_type (4)));
    const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (prs, voff, soff, 0);
    return __builtin_bit_cast (double2, t);
  };
  auto vbase = [&] (int Jn) __attribute__ ((always_inline)) {
    return Jn >= l0 && Jn <= nb
               ? (int) (sizeof (double) * (size_t) Jn * PS) + 1024 * UPW * w + 16 * lane
               : OOB;
  };
  const bool dneed = lane <= kd - PB * D + PB - 1;
  auto pvoff = [&] (int vb, int d, int k) __attribute__ ((always_inline)) {
    if (d == 0) return lane >= 8 * w + 2 * k ? vb : OOB;
    if (d == D) return dneed ? vb : OOB;
    return vb;
  };
  auto hist_at = [&] (int K, int row) __attribute__ ((always_inline)) {
    return hist + ((size_t) (K % (D + 1)) * PB + row) * NR;
  };
  auto yload = [&] (int J, double (&yq)[NR]) __attribute__ ((always_inline)) {
    const int j = J * PB + lane;
#pragma unroll
    for (int r = 0; r < NR; r++) yq[r] = J >= l0 && j < n ? bg[(size_t) r * rs + j] : 0.0;
  };
  // history: blocks l1 .. l1+D-1 (zero past n), one wave per block
  if (w < D) {
    const int K = l1 + w, j = K * PB + lane;
#pragma unroll
    for (int r = 0; r < NR; r++) hist_at (K, lane)[r] = j < n ? bg[(size_t) r * rs + j] : 0.0;
  }
  {
    const int v0 = vbase (l1 - 1), v1 = vbase (l1 - 2);
#pragma unroll
    for (int k = 0; k < UPW; k++)
#pragma unroll
      for (int d = 0; d <= D; d++) ua[k * (D + 1) + d] = pld (pvoff (v0, d, k), 32768 * d + 1024 * k);
    yload (l1 - 1, ya);
#pragma unroll
    for (int k = 0; k < UPW; k++)
#pragma unroll
      for (int d = 0; d <= D; d++) ub[k * (D + 1) + d] = pld (pvoff (v1, d, k), 32768 * d + 1024 * k);
    yload (l1 - 2, yb);
  }
  __builtin_amdgcn_s_waitcnt (0);
  __syncthreads ();

  auto step = [&] (int J, double2 (&u)[UPW * (D + 1)], double (&yq)[NR]) __attribute__ ((always_inline)) {
    // two halves of the wave's rows (pairs 2h, 2h+1), block by block: the
    // terms of T_J (d = 0) and of W_{J,J+d} x_{J+d}, then the lane sums
    const int vb = vbase (J - 2);
    double xo[2][NR];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      double p[NR][4];
#pragma unroll
      for (int d = 0; d <= D; d++) {
        double xd[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) xd[r] = d == 0 ? yq[r] : -hist_at (J + d, lane)[r];
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
          const int k = 2 * h + kk;
          const double2 v = u[k * (D + 1) + d];
#pragma unroll
          for (int r = 0; r < NR; r++) {
            if (d == 0) {
              p[r][2 * kk] = v.x * xd[r];
              p[r][2 * kk + 1] = v.y * xd[r];
            } else {
              p[r][2 * kk] = fma (v.x, xd[r], p[r][2 * kk]);
              p[r][2 * kk + 1] = fma (v.y, xd[r], p[r][2 * kk + 1]);
            }
          }
          int so = 32768 * d + 1024 * k;
#pragma unroll
          for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(p[r][2 * kk + 1]));
          u[k * (D + 1) + d] = pld (pvoff (vb, d, k), so);
        }
      }
#pragma unroll
      for (int r = 0; r < NR; r++) xo[h][r] = lane_reduce4 (p[r], lane);
    }
    // y of the block two steps on, loaded after this one's use
    {
      int off = 0;
#pragma unroll
      for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(xo[1][r]), "v"(yq[r]));
      const int j = (J - 2) * PB + lane + off;
#pragma unroll
      for (int r = 0; r < NR; r++) yq[r] = J - 2 >= l0 && j < n ? bg[(size_t) r * rs + j] : 0.0;
    }
    // row 8w + 4h + lane/16 of x_J, from the first lane of its 16
    if ((lane & 15) == 0) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = 8 * w + 4 * h + (lane >> 4), j = J * PB + row;
#pragma unroll
        for (int r = 0; r < NR; r++) {
          hist_at (J, row)[r] = xo[h][r];
          if (j < n) bg[(size_t) r * rs + j] = xo[h][r];
        }
      }
    }
    lds_sync ();
  };

  int J = l1 - 1;
  for (; J - 1 >= l0; J -= 2) {
    step (J, ua, ya);
    step (J - 1, ub, yb);
  }
  if (J >= l0) step (J, ua, ya);
}

template <int NR, int D, bool LL = false>
__device__ __forceinline__ void k_unit_sweeps_body (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd)
{
  constexpr int NP = UPW * (D + 1);                 // pairs per wave and panel
  static_assert (UW * UPW * 2 == PB, "rows per wave");
  extern __shared__ double sm[];
  double* ring = sm;                                // [D+1][UW][64][NR] pending r
  double* part = ring + (D + 1) * UW * PB * NR;     // [2][UW][64][NR] triangle partials
  SweepSeg g;
  if (!sweep_seg (g, NR, D, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb, nmaxb)) return;
  const int n = g.n, rs = g.rs;
  double* bg = g.bg;
  const int nb = (n + PB - 1) / PB;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
  const size_t PS = (size_t) (D + 1) * PB * PB;     // doubles per panel
  double2 ua[NP], ub[NP];
  double ba[NR], bb[NR];
  // the separator stage of a twisted solve (top piece rows [m, m+S)) adds
  // the bottom piece's residual of S, which its forward sweep left in its
  // last S rows (reversed), before the sweeps, and hands x_S back to those
  // rows after them: this chain's columns only, so the workgroup's own
  // barriers order it
  double* sb = nullptr;
  int sm0 = 0, snb = 0;
  if (stage == 2) {
    const int* t = tw + TW * g.mode;
    sm0 = t[TW_M];
    snb = t[TW_NBOT];
    sb = rhsb + ((size_t) g.mode * G + g.grp) * NR * nmaxb;
    for (int i = threadIdx.x; i < NR * PB * D; i += UW * 64) {
      const int r = i / (PB * D), q = i % (PB * D);
      bg[(size_t) r * rs + sm0 + q] += sb[(size_t) r * nmaxb + snb - 1 - q];
    }
    __threadfence_block ();
    __syncthreads ();
  }

  for (int pass = 0; pass < 2; pass++) {
    if (LL && pass == 1) {
      unit_backward_ll<NR, D> (PF + (size_t) g.mode * pstride + g.poff, n, nb, bg, rs, g.l0,
                               g.l1, lane, w, kd, ring, ua, ub, ba, bb);
      continue;
    }
    const int J0 = pass ? g.b0 : g.f0, J1 = pass ? g.b1 : g.f1;
    const bool wout = pass ? g.wb : g.wf;
    if (J0 >= J1) continue;
    const double* P = (pass == 0 ? PF : PR) + (size_t) g.mode * pstride + g.poff;
    auto gidx = [&] (int j) __attribute__ ((always_inline)) {
      j = min (j, n - 1);
      return pass ? n - 1 - j : j;
    };
    // Panel loads through a buffer descriptor over the piece's nb+1 panels.
    // Pair (d, k) of this wave sits at byte 32768 d + 1024 (4w + k) of a
    // panel, lane c at 16 c in it: the lane's offset (one register per
    // step) plus a constant.  A lane whose pair holds no entry in its column
    // (T_J: c < 2p; W_{J,J+D}: c > kd - 64D + 63), and every lane past panel
    // nb, gets a base past the range and reads 0 with no memory access.
    const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc (
        (void*) P, (short) 0, (int) (sizeof (double) * (size_t) (nb + 1) * PS), 0x00020000);
    constexpr int OOB = 0x7f000000;
    auto pld = [&] (int voff, int soff) __attribute__ ((always_inline)) {
      typedef int v4i __attribute__ ((ext_vector_type (4)));
      const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (prs, voff, soff, 0);
      return __builtin_bit_cast (double2, t);
    };
    auto vbase = [&] (int Jn) __attribute__ ((always_inline)) {
      return Jn <= nb ? (int) (sizeof (double) * (size_t) Jn * PS) + 1024 * UPW * w + 16 * lane
                      : OOB;
    };
    const bool dneed = lane <= kd - PB * D + PB - 1;
    auto ring_at = [&] (int K, int s) __attribute__ ((always_inline)) {
      return ring + ((size_t) ((K % (D + 1)) * UW + s) * PB + lane) * NR;
    };
    auto part_at = [&] (int J, int s) __attribute__ ((always_inline)) {
      return part + ((size_t) ((J & 1) * UW + s) * PB + lane) * NR;
    };
    // the voffset of pair (d, k) for a panel whose base is vb
    auto pvoff = [&] (int vb, int d, int k) __attribute__ ((always_inline)) {
      if (d == 0) return lane >= 8 * w + 2 * k ? vb : OOB;
      if (d == D) return dneed ? vb : OOB;
      return vb;
    };
    // y_J = the sum of the triangle partials, out by wave 0
    auto put_y = [&] (int J) __attribute__ ((always_inline)) {
      const int j = J * PB + lane;
#pragma unroll
      for (int r = 0; r < NR; r++) {
        double y = part_at (J, 0)[r];
#pragma unroll
        for (int s = 1; s < UW; s++) y += part_at (J, s)[r];
        if (w == 0 && j < n) bg[(size_t) r * rs + gidx (j)] = y;
      }
    };

    // window: blocks J0..J0+D-1 from b, one wave per block
    if (w < D) {
      const int K = J0 + w;
      const bool in = K * PB + lane < n;
#pragma unroll
      for (int s = 0; s < UW; s++)
#pragma unroll
        for (int r = 0; r < NR; r++)
          ring_at (K, s)[r] = (in && s == 0) ? bg[(size_t) r * rs + gidx (K * PB + lane)] : 0.0;
    }
    // panels J0 and J0+1, b of the blocks entering at steps J0 and J0+1
    {
      const int v0 = vbase (J0), v1 = vbase (J0 + 1);
#pragma unroll
      for (int k = 0; k < UPW; k++)
#pragma unroll
        for (int d = 0; d <= D; d++) ua[k * (D + 1) + d] = pld (pvoff (v0, d, k), 32768 * d + 1024 * k);
#pragma unroll
      for (int r = 0; r < NR; r++) ba[r] = bg[(size_t) r * rs + gidx ((J0 + D) * PB + lane)];
#pragma unroll
      for (int k = 0; k < UPW; k++)
#pragma unroll
        for (int d = 0; d <= D; d++) ub[k * (D + 1) + d] = pld (pvoff (v1, d, k), 32768 * d + 1024 * k);
#pragma unroll
      for (int r = 0; r < NR; r++) bb[r] = bg[(size_t) r * rs + gidx ((J0 + D + 1) * PB + lane)];
    }
    __builtin_amdgcn_s_waitcnt (0);
    __syncthreads ();

    auto step = [&] (int J, double2 (&u)[NP], double (&bq)[NR]) __attribute__ ((always_inline)) {
      // z_J = r_J (the sum of its slots); this wave's eight rows of it
      double zs[NR][2 * UPW];
#pragma unroll
      for (int r = 0; r < NR; r++) {
        double z = ring_at (J, 0)[r];
#pragma unroll
        for (int s = 1; s < UW; s++) z += ring_at (J, s)[r];
#pragma unroll
        for (int j = 0; j < 2 * UPW; j++) zs[r][j] = lane_bcast (z, 8 * w + j);
      }
      if (J > J0) put_y (J - 1);
      // products with this wave's pairs of panel J (all D+1 targets); each
      // pair reloaded from panel J+2 right behind its use, its offset tied
      // to the sums so the loads stay behind the FMAs
      double acc[D + 1][NR];
#pragma unroll
      for (int d = 0; d <= D; d++)
#pragma unroll
        for (int r = 0; r < NR; r++) acc[d][r] = 0.0;
      const int vb = vbase (J + 2);
#pragma unroll
      for (int k = 0; k < UPW; k++) {
#pragma unroll
        for (int d = 0; d <= D; d++) {
          const double2 v = u[k * (D + 1) + d];
#pragma unroll
          for (int r = 0; r < NR; r++)
            acc[d][r] = fma (v.y, zs[r][2 * k + 1], fma (v.x, zs[r][2 * k], acc[d][r]));
          int so = 32768 * d + 1024 * k;
#pragma unroll
          for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(acc[d][r]));
          u[k * (D + 1) + d] = pld (pvoff (vb, d, k), so);
        }
      }
      // into the targets: the triangle partial, pending blocks J+1..J+D
      // (block J+D enters the window: every wave sets its slot, wave 0 adds b)
#pragma unroll
      for (int r = 0; r < NR; r++) part_at (J, w)[r] = acc[0][r];
#pragma unroll
      for (int d = 1; d < D; d++)
#pragma unroll
        for (int r = 0; r < NR; r++) ring_at (J + d, w)[r] -= acc[d][r];
      {
        const bool in = w == 0 && (J + D) * PB + lane < n;
#pragma unroll
        for (int r = 0; r < NR; r++) ring_at (J + D, w)[r] = (in ? bq[r] : 0.0) - acc[D][r];
      }
      // b of the block entering two steps on, loaded after this one's use
      {
        int off = 0;
#pragma unroll
        for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(acc[D][r]), "v"(bq[r]));
        const int gi = gidx ((J + D + 2) * PB + lane) + off;
#pragma unroll
        for (int r = 0; r < NR; r++) bq[r] = bg[(size_t) r * rs + gi];
      }
      lds_sync ();
    };

    // both steps on every path to the back edge, so that the compiler's wait
    // counts see buffer a's loads two steps old and b's one step old
    int J = J0;
    for (; J + 1 < J1; J += 2) {
      step (J, ua, ba);
      step (J + 1, ub, bb);
    }
    if (J < J1) step (J, ua, ba);
    put_y (J1 - 1);
    // a segment that stops early hands its pending window on through b
    if (wout && w < D) {
      const int K = J1 + w;
      if (K * PB + lane < n) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
          double v = ring_at (K, 0)[r];
#pragma unroll
          for (int s = 1; s < UW; s++) v += ring_at (K, s)[r];
          bg[(size_t) r * rs + gidx (K * PB + lane)] = v;
        }
      }
    }
    // the next pass reads what this one wrote to global memory: written and
    // read by waves of this workgroup only (one CU, one L1), so a
    // workgroup-scope fence orders it; a device-scope fence would write back
    // and invalidate the XCD's L2 under the other workgroups' panel streams
    __threadfence_block ();
    __syncthreads ();
  }
  if (stage == 2) {
    __threadfence_block ();
    __syncthreads ();
    for (int i = threadIdx.x; i < NR * PB * D; i += UW * 64) {
      const int r = i / (PB * D), q = i % (PB * D);
      sb[(size_t) r * nmaxb + snb - 1 - q] = bg[(size_t) r * rs + sm0 + q];
    }
  }
}

template <int NR, int D, bool LL>
__global__ __launch_bounds__(UW * 64) void k_unit_sweeps (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd)
{
  k_unit_sweeps_body<NR, D, LL> (PF, PR, pstride, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb, nmaxb, kd);
}

// ---------------------------------------------------------------------------
// Split unit sweeps: one chain of block steps on two CUs
// ---------------------------------------------------------------------------
// A step moves the whole panel (up to 192 KB) through one CU's vector
// memory path, which bounds it.  Here the panel of each step is shared by
// two workgroups on one XCD: the near one (role 0) takes T_J and
// W_{J,J+1..J+H} (H = D/2), the far one (role 1) W_{J,J+H+1..J+D}.
//   near, step J:  z_J = (own pending r_J) + p_J;  publish z_J;  y_J; near updates
//   far,  step J:  wait z_J;  far updates;  block J+H complete -> publish p_{J+H}
// p_K (the far contributions to block K) is published H steps before the
// near side needs it, so neither side waits in the steady state.  Messages
// go through L2 as 16-byte records {value, epoch, seq} in rings of XR
// entries per direction: a record is accepted when its epoch and sequence
// number match, so no flag has to be ordered against the data.  The epoch
// of a chain lives in device memory and advances once per launch (graph
// replays included).  Both workgroups of a chain must be resident at once:
// the host launches this kernel only when every workgroup of the grid fits
// on the device at once.  A wait that exceeds its bound marks the side dead:
// it skips further waits and its values become NaN, so a failure shows in
// the solution rather than as a hang.
constexpr int XR = 8;                               // ring entries per direction
constexpr int XSPIN = 1 << 19;                      // polls before a wait gives up
                                                    // (SX_SPLIT_SPIN overrides)

__device__ __forceinline__ void msg_put (__amdgpu_buffer_rsrc_t rs, int off, double v,
                                         unsigned ep, unsigned seq)
{
  typedef int v4i __attribute__ ((ext_vector_type (4)));
  const long long x = __double_as_longlong (v);
  v4i t;
  t.x = (int) (x & 0xffffffffll);
  t.y = (int) (x >> 32);
  t.z = (int) ep;
  t.w = (int) seq;
  __builtin_amdgcn_raw_buffer_store_b128 (t, rs, off, 0, 17);   // sc0 sc1: to L2
}

// every lane's record of (ep, seq), or NaN once the wait gives up
template <int NR>
__device__ __forceinline__ void msg_wait (__amdgpu_buffer_rsrc_t rs, int off0, int rstride,
                                          unsigned ep, unsigned seq, double (&v)[NR],
                                          bool& dead, int spin)
{
  typedef int v4i __attribute__ ((ext_vector_type (4)));
  // spin < 0: every wait gives up at once (the host forces the failure path)
  if (dead || spin < 0) {
    dead = true;
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = __builtin_nan ("");
    return;
  }
  for (int it = 0;; it++) {
    bool ok = true;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      asm volatile ("" ::: "memory");
      const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (rs, off0 + r * rstride, 0, 17);
      v[r] = __longlong_as_double (((long long) t.y << 32) | (unsigned) t.x);
      ok = ok && (unsigned) t.z == ep && (unsigned) t.w == seq;
    }
    if (__builtin_amdgcn_ballot_w64 (!ok) == 0) return;
    if (it >= spin) break;
    __builtin_amdgcn_s_sleep (2);
  }
  dead = true;
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = __builtin_nan ("");
}

// Roles of a split chain (R workgroups): the near one (role 0) takes T_J
// and W_{J,J+1..J+H}, H = ceil((D+1)/R) - 1 (at least 1); far role
// r = 1..R-1 takes a contiguous range [a_r, b_r] of the remaining blocks
// H+1..D, as even as they divide (the first ones one block more).
template <int D, int R> __host__ __device__ constexpr int sp_h ()
{
  return (D + R) / R - 1 < 1 ? 1 : (D + R) / R - 1;
}
template <int D, int R> __host__ __device__ constexpr int sp_fbmax ()
{
  return (D - sp_h<D, R> () + R - 2) / (R - 1);
}
// first block and block count of far role r (1..R-1)
template <int D, int R> __host__ __device__ constexpr int sp_a (int r)
{
  const int F = D - sp_h<D, R> (), base = F / (R - 1), extra = F % (R - 1);
  int a = sp_h<D, R> () + 1;
  for (int q = 1; q < r; q++) a += base + (q - 1 < extra ? 1 : 0);
  return a;
}
template <int D, int R> __host__ __device__ constexpr int sp_n (int r)
{
  const int F = D - sp_h<D, R> (), base = F / (R - 1), extra = F % (R - 1);
  return base + (r - 1 < extra ? 1 : 0);
}

template <int NR, int D, int R, bool LL = false, int PD = 2>
__device__ __forceinline__ void split_body (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd,
    char* xch, unsigned ep, int spin, bool& dead)
{
  constexpr int H = sp_h<D, R> (), FBM = sp_fbmax<D, R> ();
  constexpr int NA = UPW * (H + 1), NB = UPW * FBM;  // pairs per wave
  constexpr int NP = NA > NB ? NA : NB;
  constexpr int NF = R - 1;                         // far roles
  constexpr int MS = PB * 16;                       // bytes per record row (64 lanes)
  extern __shared__ double sm[];
  SweepSeg g;
  if (!sweep_seg (g, NR, D, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb, nmaxb, R))
    return;
  const int n = g.n, rs = g.rs, role = g.role;
  double* bg = g.bg;
  // Stage 2 of a twisted solve (the separator S = rows [m, m+S) of the top
  // piece) folds in the two moves around it: S's right-hand side is read
  // with the bottom piece's residual added (row m+a of b plus row nbot-1-a
  // of the bottom piece's vector), and x_S is written to both.
  double* rbS = nullptr;
  int tm = 0, tnb = 0;
  if (stage == 2) {
    tm = tw[TW * g.mode + TW_M];
    tnb = tw[TW * g.mode + TW_NBOT];
    rbS = rhsb + ((size_t) g.mode * G + g.grp) * NR * nmaxb;
  }
  auto bS = [&] (int r, int row) __attribute__ ((always_inline)) {
    return rbS ? rbS[(size_t) r * nmaxb + (tnb - 1 - (row - tm))] : 0.0;
  };
  const int nb = (n + PB - 1) / PB;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
  const size_t PS = (size_t) (D + 1) * PB * PB;
  // this chain's rings: [dir][XR][NR] rows of 64 records; dir 0: z (near ->
  // far roles), dir r: p of far role r (-> near)
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc (
      xch + (size_t) g.chain * R * XR * NR * MS, (short) 0, R * XR * NR * MS, 0x00020000);
  auto xoff = [&] (int dir, int K) __attribute__ ((always_inline)) {
    return ((dir * XR + (K & (XR - 1))) * NR) * MS + 16 * lane;
  };
  // the top bit keeps a never-written (zero) record from matching
  auto seq = [&] (int pass, int K) __attribute__ ((always_inline)) {
    return 0x80000000u | ((unsigned) pass << 24) | (unsigned) K;
  };
  // panels of the next PD steps and b of the blocks entering with them
  double2 ubuf[PD][NP];
  double bbuf[PD][NR];
  bool ran = false;                                  // a previous pass ran
  // The next step's records, read at the start of a step before its panel
  // reloads: a wait for them then leaves those reloads in flight (vector
  // loads retire in order, so a wait for a load issued after them would
  // not).  Taken when their tags match, else the step polls.
  typedef int v4i __attribute__ ((ext_vector_type (4)));
  v4i spec[NF][NR];
  int spec_k = -1;                                   // step whose records spec holds
  auto spec_read = [&] (int dir0, int K) __attribute__ ((always_inline)) {
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int r = 0; r < NR; r++)
        spec[f][r] = __builtin_amdgcn_raw_buffer_load_b128 (xrs, xoff (dir0 + f, K) + r * MS, 0, 17);
    spec_k = K;
  };
  // this step's records of direction dir0 + f: the speculative ones when
  // all lanes carry the tag
  auto take = [&] (int dir0, int f, int K, unsigned sq, double (&v)[NR]) __attribute__ ((always_inline)) {
    bool ok = spec_k == K && !dead;
    if (ok) {
      bool lane_ok = true;
#pragma unroll
      for (int r = 0; r < NR; r++) {
        v[r] = __longlong_as_double (((long long) spec[f][r].y << 32) | (unsigned) spec[f][r].x);
        lane_ok = lane_ok && (unsigned) spec[f][r].z == ep && (unsigned) spec[f][r].w == sq;
      }
      ok = __builtin_amdgcn_ballot_w64 (!lane_ok) == 0;
    }
    if (!ok) msg_wait<NR> (xrs, xoff (dir0 + f, K), MS, ep, sq, v, dead, spin);
  };

  for (int pass = 0; pass < 2; pass++) {
    if (LL && pass == 1) {
      // Left-looking backward (unit_backward_ll) on two workgroups: the
      // near one forms x_J = T_J y_J - sum_{d<=H} W_{J,J+d} x_{J+d} - f_J and
      // publishes x_J (one record per lane, direction 0); the far one forms
      // f_K = sum_{d>H} W_{K,K+d} x_{K+d} as soon as x_{K+H+1} is known and
      // publishes it by row (direction 1), H steps before the near one needs
      // it.  Blocks at or past l1 are known (b); the separator's solution of
      // a twisted solve also goes to the bottom piece's vector.
      const int l0 = g.l0, l1 = g.l1;
      if (l0 >= l1) break;
      const double* P = PF + (size_t) g.mode * pstride + g.poff;
      const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc (
          (void*) P, (short) 0, (int) (sizeof (double) * (size_t) (nb + 1) * PS), 0x00020000);
      constexpr int OOB = 0x7f000000;
      auto pld = [&] (int voff, int soff) __attribute__ ((always_inline)) {
        typedef int v4i __attribute__ ((ext_vector_type (4)));
        const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (prs, voff, soff, 0);
        return __builtin_bit_cast (double2, t);
      };
      auto vbase = [&] (int Jn) __attribute__ ((always_inline)) {
        return Jn >= l0 && Jn <= nb
                   ? (int) (sizeof (double) * (size_t) Jn * PS) + 1024 * UPW * w + 16 * lane
                   : OOB;
      };
      const bool dneed = lane <= kd - PB * D + PB - 1;
      auto pvoff = [&] (int vb, int d, int k) __attribute__ ((always_inline)) {
        if (d == 0) return lane >= 8 * w + 2 * k ? vb : OOB;
        if (d == D) return dneed ? vb : OOB;
        return vb;
      };
      // x of block K >= l1 (known), lane c
      auto xknown = [&] (int K, int r) __attribute__ ((always_inline)) {
        const int j = K * PB + lane;
        return j < n ? bg[(size_t) r * rs + j] : 0.0;
      };
      if (role == 0) {
        constexpr int RE = H + 1;                    // x of blocks J..J+H
        double* hist = sm;                           // [RE][64][NR]
        auto hist_at = [&] (int K, int row) __attribute__ ((always_inline)) {
          return hist + ((size_t) (K % RE) * PB + row) * NR;
        };
        auto yload = [&] (int J, double (&yq)[NR]) __attribute__ ((always_inline)) {
          const int j = J * PB + lane;
#pragma unroll
          for (int r = 0; r < NR; r++) yq[r] = J >= l0 && j < n ? bg[(size_t) r * rs + j] : 0.0;
        };
        if (w < H) {
          const int K = l1 + w;
#pragma unroll
          for (int r = 0; r < NR; r++) hist_at (K, lane)[r] = xknown (K, r);
        }
        {
#pragma unroll
          for (int i = 0; i < PD; i++) {
            const int v0 = vbase (l1 - 1 - i);
#pragma unroll
            for (int k = 0; k < UPW; k++)
#pragma unroll
              for (int d = 0; d <= H; d++)
                ubuf[i][k * (H + 1) + d] = pld (pvoff (v0, d, k), 32768 * d + 1024 * k);
            yload (l1 - 1 - i, bbuf[i]);
          }
        }
        __builtin_amdgcn_s_waitcnt (0);
        __syncthreads ();
        // f_J of row 8w + 4 ((lane >> 3) & 1) + (lane >> 4): the row whose sum
        // lane_reduce4 leaves in lane (lane & 15) == 0 (first half) or 8
        // (second half) of the lane's sixteen
        auto foff = [&] (int J) __attribute__ ((always_inline)) {
          return xoff (1, J) - 16 * lane + 16 * (8 * w + 4 * ((lane >> 3) & 1) + (lane >> 4));
        };
        auto step = [&] (int J, double2 (&u)[NP], double (&yq)[NR]) __attribute__ ((always_inline)) {
          // f_J's records read before this step's panel reloads: waiting for
          // them below waits for no reload
          typedef int v4i __attribute__ ((ext_vector_type (4)));
          v4i fr[NR];
#pragma unroll
          for (int r = 0; r < NR; r++)
            fr[r] = __builtin_amdgcn_raw_buffer_load_b128 (xrs, foff (J) + r * MS, 0, 17);
          const int vb = vbase (J - PD);
          double xo[2][NR];
#pragma unroll
          for (int h = 0; h < 2; h++) {
            double pp[NR][4];
#pragma unroll
            for (int d = 0; d <= H; d++) {
              double xd[NR];
#pragma unroll
              for (int r = 0; r < NR; r++) xd[r] = d == 0 ? yq[r] : -hist_at (J + d, lane)[r];
#pragma unroll
              for (int kk = 0; kk < 2; kk++) {
                const int k = 2 * h + kk;
                const double2 v = u[k * (H + 1) + d];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                  if (d == 0) {
                    pp[r][2 * kk] = v.x * xd[r];
                    pp[r][2 * kk + 1] = v.y * xd[r];
                  } else {
                    pp[r][2 * kk] = fma (v.x, xd[r], pp[r][2 * kk]);
                    pp[r][2 * kk + 1] = fma (v.y, xd[r], pp[r][2 * kk + 1]);
                  }
                }
                int so = 32768 * d + 1024 * k;
#pragma unroll
                for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(pp[r][2 * kk + 1]));
                u[k * (H + 1) + d] = pld (pvoff (vb, d, k), so);
              }
            }
#pragma unroll
            for (int r = 0; r < NR; r++) xo[h][r] = lane_reduce4 (pp[r], lane);
          }
          {
            int off = 0;
#pragma unroll
            for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(xo[1][r]), "v"(yq[r]));
            const int j = (J - PD) * PB + lane + off;
#pragma unroll
            for (int r = 0; r < NR; r++)
              yq[r] = J - PD >= l0 && j < n ? bg[(size_t) r * rs + j] : 0.0;
          }
          // f_J: the early records when every lane's tag matches, else wait
          double fv[NR];
          {
            bool ok = !dead;
            if (ok) {
              bool lane_ok = true;
#pragma unroll
              for (int r = 0; r < NR; r++) {
                fv[r] = __longlong_as_double (((long long) fr[r].y << 32) | (unsigned) fr[r].x);
                lane_ok = lane_ok && (unsigned) fr[r].z == ep && (unsigned) fr[r].w == seq (1, J);
              }
              ok = __builtin_amdgcn_ballot_w64 (!lane_ok) == 0;
            }
            if (!ok) msg_wait<NR> (xrs, foff (J), MS, ep, seq (1, J), fv, dead, spin);
          }
          const int hw = (lane & 15) == 0 ? 0 : ((lane & 15) == 8 ? 1 : -1);
          if (hw >= 0) {
            const int row = 8 * w + 4 * hw + (lane >> 4), j = J * PB + row;
#pragma unroll
            for (int r = 0; r < NR; r++) {
              const double x = (hw ? xo[1][r] : xo[0][r]) - fv[r];
              hist_at (J, row)[r] = x;
              if (j < n) {
                bg[(size_t) r * rs + j] = x;
                if (rbS) rbS[(size_t) r * nmaxb + (tnb - 1 - (j - tm))] = x;
              }
            }
          }
          lds_sync ();
          if (w == 0)
#pragma unroll
            for (int r = 0; r < NR; r++)
              msg_put (xrs, xoff (0, J) + r * MS, hist_at (J, lane)[r], ep, seq (1, J));
        };
        int J = l1 - 1;
        for (; J - (PD - 1) >= l0; J -= PD) {
#pragma unroll
          for (int i = 0; i < PD; i++) step (J - i, ubuf[i], bbuf[i]);
        }
#pragma unroll
        for (int i = 0; i < PD - 1; i++)
          if (J - i >= l0) step (J - i, ubuf[i], bbuf[i]);
      } else {
        constexpr int FB = D - H;                    // far blocks H+1..D
        constexpr int RE = FB + 1;                   // x of blocks K+H+1..K+D
        double* hist = sm;                           // [RE][64][NR]
        auto hist_at = [&] (int K) __attribute__ ((always_inline)) {
          return hist + ((size_t) (K % RE) * PB + lane) * NR;
        };
        auto fvoff = [&] (int vb, int dd, int k) __attribute__ ((always_inline)) {
          return dd < FB ? pvoff (vb, H + 1 + dd, k) : OOB;
        };
        // the near side must be done with the previous pass's records
        if (ran) {
          double dv[NR];
          const int Jp = g.f1;
          msg_wait<NR> (xrs, xoff (0, Jp), MS, ep, seq (0, Jp), dv, dead, spin);
        }
        {
#pragma unroll
          for (int i = 0; i < PD; i++) {
            const int v0 = vbase (l1 - 1 - i);
#pragma unroll
            for (int k = 0; k < UPW; k++)
#pragma unroll
              for (int dd = 0; dd < FBM; dd++)
                ubuf[i][k * FBM + dd] = pld (fvoff (v0, dd, k), 32768 * (H + 1 + dd) + 1024 * k);
          }
        }
        __builtin_amdgcn_s_waitcnt (0);
        __syncthreads ();
        auto step = [&] (int K, double2 (&u)[NP]) __attribute__ ((always_inline)) {
          const int Kx = K + H + 1;
          // x_{K+H+1}'s records read before this step's panel reloads; the
          // older blocks' products run first
          typedef int v4i __attribute__ ((ext_vector_type (4)));
          v4i xr[NR];
          if (Kx < l1)
#pragma unroll
            for (int r = 0; r < NR; r++)
              xr[r] = __builtin_amdgcn_raw_buffer_load_b128 (xrs, xoff (0, Kx) + r * MS, 0, 17);
          const int vb = vbase (K - PD);
          double pp[NR][2 * UPW];
#pragma unroll
          for (int dd = FB - 1; dd >= 0; dd--) {
            double xd[NR];
            if (dd > 0) {
#pragma unroll
              for (int r = 0; r < NR; r++)
                xd[r] = Kx + dd < l1 ? hist_at (Kx + dd)[r] : xknown (Kx + dd, r);
            } else if (Kx < l1) {
              bool ok = !dead;
              if (ok) {
                bool lane_ok = true;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                  xd[r] = __longlong_as_double (((long long) xr[r].y << 32) | (unsigned) xr[r].x);
                  lane_ok = lane_ok && (unsigned) xr[r].z == ep && (unsigned) xr[r].w == seq (1, Kx);
                }
                ok = __builtin_amdgcn_ballot_w64 (!lane_ok) == 0;
              }
              if (!ok) msg_wait<NR> (xrs, xoff (0, Kx), MS, ep, seq (1, Kx), xd, dead, spin);
              if (w == 0)
#pragma unroll
                for (int r = 0; r < NR; r++) hist_at (Kx)[r] = xd[r];
            } else {
#pragma unroll
              for (int r = 0; r < NR; r++) xd[r] = xknown (Kx, r);
            }
#pragma unroll
            for (int k = 0; k < UPW; k++) {
              const double2 v = u[k * FBM + dd];
#pragma unroll
              for (int r = 0; r < NR; r++) {
                if (dd == FB - 1) {
                  pp[r][2 * k] = v.x * xd[r];
                  pp[r][2 * k + 1] = v.y * xd[r];
                } else {
                  pp[r][2 * k] = fma (v.x, xd[r], pp[r][2 * k]);
                  pp[r][2 * k + 1] = fma (v.y, xd[r], pp[r][2 * k + 1]);
                }
              }
              int so = 32768 * (H + 1 + dd) + 1024 * k;
#pragma unroll
              for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(pp[r][2 * k + 1]));
              u[k * FBM + dd] = pld (fvoff (vb, dd, k), so);
            }
          }
          // f_K row t to record slot t
#pragma unroll
          for (int r = 0; r < NR; r++) {
            const double f = lane_reduce8 (pp[r], lane);
            if ((lane & 7) == 0)
              msg_put (xrs, xoff (1, K) - 16 * lane + 16 * (8 * w + (lane >> 3)) + r * MS, f, ep,
                       seq (1, K));
          }
          lds_sync ();
        };
        int K = l1 - 1;
        for (; K - (PD - 1) >= l0; K -= PD) {
#pragma unroll
          for (int i = 0; i < PD; i++) step (K - i, ubuf[i]);
        }
#pragma unroll
        for (int i = 0; i < PD - 1; i++)
          if (K - i >= l0) step (K - i, ubuf[i]);
      }
      break;
    }
    const int J0 = pass ? g.b0 : g.f0, J1 = pass ? g.b1 : g.f1;
    const bool wout = pass ? g.wb : g.wf;
    if (J0 >= J1) continue;
    const double* P = (pass == 0 ? PF : PR) + (size_t) g.mode * pstride + g.poff;
    auto gidx = [&] (int j) __attribute__ ((always_inline)) {
      j = min (j, n - 1);
      return pass ? n - 1 - j : j;
    };
    const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc (
        (void*) P, (short) 0, (int) (sizeof (double) * (size_t) (nb + 1) * PS), 0x00020000);
    constexpr int OOB = 0x7f000000;
    auto pld = [&] (int voff, int soff) __attribute__ ((always_inline)) {
      typedef int v4i __attribute__ ((ext_vector_type (4)));
      const v4i t = __builtin_amdgcn_raw_buffer_load_b128 (prs, voff, soff, 0);
      return __builtin_bit_cast (double2, t);
    };
    auto vbase = [&] (int Jn) __attribute__ ((always_inline)) {
      return Jn <= nb ? (int) (sizeof (double) * (size_t) Jn * PS) + 1024 * UPW * w + 16 * lane
                      : OOB;
    };
    const bool dneed = lane <= kd - PB * D + PB - 1;
    auto pvoff = [&] (int vb, int d, int k) __attribute__ ((always_inline)) {
      if (d == 0) return lane >= 8 * w + 2 * k ? vb : OOB;
      if (d == D) return dneed ? vb : OOB;
      return vb;
    };

    if (role == 0) {
      // ---------------- near side: T_J, W_{J,J+1..J+H} ----------------
      constexpr int RE = H + 1;                      // ring entries: blocks J..J+H
      double* ring = sm;                             // [RE][UW][64][NR]
      double* part = ring + RE * UW * PB * NR;       // [2][UW][64][NR]
      auto ring_at = [&] (int K, int s) __attribute__ ((always_inline)) {
        return ring + ((size_t) ((K % RE) * UW + s) * PB + lane) * NR;
      };
      auto part_at = [&] (int J, int s) __attribute__ ((always_inline)) {
        return part + ((size_t) ((J & 1) * UW + s) * PB + lane) * NR;
      };
      auto put_y = [&] (int J) __attribute__ ((always_inline)) {
        const int j = J * PB + lane;
#pragma unroll
        for (int r = 0; r < NR; r++) {
          double y = part_at (J, 0)[r];
#pragma unroll
          for (int s = 1; s < UW; s++) y += part_at (J, s)[r];
          if (w == 0 && j < n) {
            bg[(size_t) r * rs + gidx (j)] = y;
            if (rbS && pass) rbS[(size_t) r * nmaxb + (tnb - 1 - (gidx (j) - tm))] = y;
          }
        }
      };
      // window: blocks J0..J0+H-1 from b, one wave per block
      if (w < H) {
        const int K = J0 + w;
        const bool in = K * PB + lane < n;
#pragma unroll
        for (int s = 0; s < UW; s++)
#pragma unroll
          for (int r = 0; r < NR; r++)
            ring_at (K, s)[r] = (in && s == 0) ? bg[(size_t) r * rs + gidx (K * PB + lane)] +
                                                   (pass ? 0.0 : bS (r, gidx (K * PB + lane)))
                                               : 0.0;
      }
      {
#pragma unroll
        for (int i = 0; i < PD; i++) {
          const int v0 = vbase (J0 + i);
#pragma unroll
          for (int k = 0; k < UPW; k++)
#pragma unroll
            for (int d = 0; d <= H; d++)
              ubuf[i][k * (H + 1) + d] = pld (pvoff (v0, d, k), 32768 * d + 1024 * k);
#pragma unroll
          for (int r = 0; r < NR; r++)
            bbuf[i][r] = bg[(size_t) r * rs + gidx ((J0 + H + i) * PB + lane)] +
                         (pass ? 0.0 : bS (r, gidx ((J0 + H + i) * PB + lane)));
        }
      }
      __builtin_amdgcn_s_waitcnt (0);
      __syncthreads ();

      auto step = [&] (int J, double2 (&u)[NP], double (&bq)[NR]) __attribute__ ((always_inline)) {
        double pv[NR], zs[NR][2 * UPW];
#pragma unroll
        for (int r = 0; r < NR; r++) pv[r] = 0.0;
#pragma unroll
        for (int f = 0; f < NF; f++) {
          double pf[NR];
          take (1, f, J, seq (pass, J), pf);
#pragma unroll
          for (int r = 0; r < NR; r++) pv[r] += pf[r];
        }
        if (J + 1 < J1) spec_read (1, J + 1);
#pragma unroll
        for (int r = 0; r < NR; r++) {
          double z = pv[r];
#pragma unroll
          for (int s = 0; s < UW; s++) z += ring_at (J, s)[r];
          if (w == 0) msg_put (xrs, xoff (0, J) + r * MS, z, ep, seq (pass, J));
#pragma unroll
          for (int j = 0; j < 2 * UPW; j++) zs[r][j] = lane_bcast (z, 8 * w + j);
        }
        if (J > J0) put_y (J - 1);
        double acc[H + 1][NR];
#pragma unroll
        for (int d = 0; d <= H; d++)
#pragma unroll
          for (int r = 0; r < NR; r++) acc[d][r] = 0.0;
        const int vb = vbase (J + PD);
#pragma unroll
        for (int k = 0; k < UPW; k++) {
#pragma unroll
          for (int d = 0; d <= H; d++) {
            const double2 v = u[k * (H + 1) + d];
#pragma unroll
            for (int r = 0; r < NR; r++)
              acc[d][r] = fma (v.y, zs[r][2 * k + 1], fma (v.x, zs[r][2 * k], acc[d][r]));
            int so = 32768 * d + 1024 * k;
#pragma unroll
            for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(acc[d][r]));
            u[k * (H + 1) + d] = pld (pvoff (vb, d, k), so);
          }
        }
#pragma unroll
        for (int r = 0; r < NR; r++) part_at (J, w)[r] = acc[0][r];
#pragma unroll
        for (int d = 1; d < H; d++)
#pragma unroll
          for (int r = 0; r < NR; r++) ring_at (J + d, w)[r] -= acc[d][r];
        {
          const bool in = w == 0 && (J + H) * PB + lane < n;
#pragma unroll
          for (int r = 0; r < NR; r++) ring_at (J + H, w)[r] = (in ? bq[r] : 0.0) - acc[H][r];
        }
        {
          int off = 0;
#pragma unroll
          for (int r = 0; r < NR; r++) asm ("" : "+v"(off) : "v"(acc[H][r]), "v"(bq[r]));
          const int gi = gidx ((J + H + PD) * PB + lane) + off;
#pragma unroll
          for (int r = 0; r < NR; r++)
            bq[r] = bg[(size_t) r * rs + gi] + (pass ? 0.0 : bS (r, gi));
        }
        lds_sync ();
      };
      int J = J0;
      for (; J + PD - 1 < J1; J += PD) {
#pragma unroll
        for (int i = 0; i < PD; i++) step (J + i, ubuf[i], bbuf[i]);
      }
#pragma unroll
      for (int i = 0; i < PD - 1; i++)
        if (J + i < J1) step (J + i, ubuf[i], bbuf[i]);
      put_y (J1 - 1);
      // the far roles' last records: p_{J1..J1+D-1}; with them, the window
      // handed on (blocks that never entered the near ring keep b)
      for (int K = J1; K < J1 + D; K++) {
        double pv[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) pv[r] = 0.0;
#pragma unroll
        for (int f = 0; f < NF; f++) {
          double pf[NR];
          msg_wait<NR> (xrs, xoff (1 + f, K), MS, ep, seq (pass, K), pf, dead, spin);
#pragma unroll
          for (int r = 0; r < NR; r++) pv[r] += pf[r];
        }
        if (wout && w == K - J1 && K * PB + lane < n) {
#pragma unroll
          for (int r = 0; r < NR; r++) {
            double v = pv[r];
            if (K < J1 + H) {
#pragma unroll
              for (int s = 0; s < UW; s++) v += ring_at (K, s)[r];
            } else {
              v += bg[(size_t) r * rs + gidx (K * PB + lane)];
            }
            bg[(size_t) r * rs + gidx (K * PB + lane)] = v;
          }
        }
      }
      // the far roles have read every z of this pass: tell them the pass is over
      if (w == 0)
#pragma unroll
        for (int r = 0; r < NR; r++) msg_put (xrs, xoff (0, J1) + r * MS, 0.0, ep, seq (pass, J1));
    } else {
      // ---------------- far role: W_{J,J+a..J+a+FB-1} ----------------
      // a, FB from the role; register arrays and loads for FBM blocks (the
      // loads past FB read nothing: out-of-range offsets), so every role
      // issues the same instruction stream
      const int a = R == 2 || role == 1 ? sp_a<D, R> (1) : role == 2 ? sp_a<D, R> (2)
                                                                   : sp_a<D, R> (3);
      const int FB = R == 2 || role == 1 ? sp_n<D, R> (1) : role == 2 ? sp_n<D, R> (2)
                                                                    : sp_n<D, R> (3);
      const int bl = a + FB - 1;                     // last block
      const int RE = FB + 1;                         // ring entries: blocks J+a-1..J+bl
      double* ring = sm;                             // [RE][UW][64][NR]
      auto ring_at = [&] (int K, int s) __attribute__ ((always_inline)) {
        return ring + ((size_t) ((K % RE) * UW + s) * PB + lane) * NR;
      };
      // p_K of this role, published by wave 0
      auto put_p = [&] (int K) __attribute__ ((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
          double v = ring_at (K, 0)[r];
#pragma unroll
          for (int s = 1; s < UW; s++) v += ring_at (K, s)[r];
          msg_put (xrs, xoff (role, K) + r * MS, v, ep, seq (pass, K));
        }
      };
      // the near side must be done with the previous pass's records
      if (ran) {
        double dv[NR];
        const int Jp = pass ? g.f1 : 0;
        msg_wait<NR> (xrs, xoff (0, Jp), MS, ep, seq (pass - 1, Jp), dv, dead, spin);
      }
      // blocks J0+a..J0+bl-1 start empty here (their earlier terms are in
      // the window the near side reads from b); p_K = 0 for K <= J0+a-1
      if (w < FB - 1) {
        const int K = J0 + a + w;
#pragma unroll
        for (int s = 0; s < UW; s++)
#pragma unroll
          for (int r = 0; r < NR; r++) ring_at (K, s)[r] = 0.0;
      }
      if (w == 0)
        for (int K = J0; K <= J0 + a - 1; K++)
#pragma unroll
          for (int r = 0; r < NR; r++) msg_put (xrs, xoff (role, K) + r * MS, 0.0, ep, seq (pass, K));
      auto fvoff = [&] (int vb, int d, int k) __attribute__ ((always_inline)) {
        return d < FB ? pvoff (vb, a + d, k) : OOB;
      };
      {
#pragma unroll
        for (int i = 0; i < PD; i++) {
          const int v0 = vbase (J0 + i);
#pragma unroll
          for (int k = 0; k < UPW; k++)
#pragma unroll
            for (int d = 0; d < FBM; d++)
              ubuf[i][k * FBM + d] = pld (fvoff (v0, d, k), 32768 * (a + d) + 1024 * k);
        }
      }
      __builtin_amdgcn_s_waitcnt (0);
      __syncthreads ();

      auto step = [&] (int J, double2 (&u)[NP]) __attribute__ ((always_inline)) {
        if (J > J0 && w == 0) put_p (J + a - 1);
        double zv[NR], zs[NR][2 * UPW];
        take (0, 0, J, seq (pass, J), zv);
        if (J + 1 < J1) spec_read (0, J + 1);
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
          for (int j = 0; j < 2 * UPW; j++) zs[r][j] = lane_bcast (zv[r], 8 * w + j);
        double acc[FBM][NR];
#pragma unroll
        for (int d = 0; d < FBM; d++)
#pragma unroll
          for (int r = 0; r < NR; r++) acc[d][r] = 0.0;
        const int vb = vbase (J + PD);
#pragma unroll
        for (int k = 0; k < UPW; k++) {
#pragma unroll
          for (int d = 0; d < FBM; d++) {
            const double2 v = u[k * FBM + d];
#pragma unroll
            for (int r = 0; r < NR; r++)
              acc[d][r] = fma (v.y, zs[r][2 * k + 1], fma (v.x, zs[r][2 * k], acc[d][r]));
            int so = 32768 * (a + d) + 1024 * k;
#pragma unroll
            for (int r = 0; r < NR; r++) asm ("" : "+s"(so) : "v"(acc[d][r]));
            u[k * FBM + d] = pld (fvoff (vb, d, k), so);
          }
        }
#pragma unroll
        for (int d = 0; d < FBM; d++)
          if (d < FB - 1) {
#pragma unroll
            for (int r = 0; r < NR; r++) ring_at (J + a + d, w)[r] -= acc[d][r];
          } else if (d == FB - 1) {
#pragma unroll
            for (int r = 0; r < NR; r++) ring_at (J + bl, w)[r] = -acc[d][r];
          }
        lds_sync ();
      };
      int J = J0;
      for (; J + PD - 1 < J1; J += PD) {
#pragma unroll
        for (int i = 0; i < PD; i++) step (J + i, ubuf[i]);
      }
#pragma unroll
      for (int i = 0; i < PD - 1; i++)
        if (J + i < J1) step (J + i, ubuf[i]);
      // the block completed by the last step and the partial sums still
      // pending: p_{J1+a-1..J1+bl-1}; past its last block this role adds
      // nothing: zero records up to J1+D-1
      if (w == 0) {
        for (int K = J1 + a - 1; K < J1 + bl; K++) put_p (K);
        for (int K = J1 + bl; K < J1 + D; K++)
#pragma unroll
          for (int r = 0; r < NR; r++) msg_put (xrs, xoff (role, K) + r * MS, 0.0, ep, seq (pass, K));
      }
      ran = true;
    }
    // the next pass reads what this workgroup wrote (workgroup scope, as above)
    __threadfence_block ();
    __syncthreads ();
  }
}

// sync[0]: the epoch of the system's current launch, sync[1]: workgroups of
// that launch that have finished, sync[2]: status (bit 0: some wait gave up
// since the host last read it).  Every workgroup reads the epoch when it
// starts; the last one to finish advances it, so both workgroups of a chain
// always use one epoch (also when one of them starts only after its partner
// has given up and finished), and a record written under an earlier launch
// of any layout never carries the current epoch.
template <int NR, int D, int R, bool LL, int PD>
__global__ __launch_bounds__(UW * 64) void k_split_sweeps (
    const double* __restrict__ PF, const double* __restrict__ PR, size_t pstride,
    const int* __restrict__ nsolve, double* rhs, int nmax, int nmodes, int G,
    const int* __restrict__ tw, int stage, size_t boff, double* rhsb, int nmaxb, int kd,
    char* xch, unsigned* sync, int spin)
{
  // relaxed throughout: successive launches are ordered by the kernel
  // boundary, and within a launch every workgroup has used its epoch long
  // before it counts itself finished (an agent-scope acquire or release here
  // would make every workgroup invalidate or write back its L2 at its end)
  const unsigned ep = __hip_atomic_load (sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool dead = false;
  split_body<NR, D, R, LL, PD> (PF, PR, pstride, nsolve, rhs, nmax, nmodes, G, tw, stage, boff, rhsb,
                     nmaxb, kd, xch, ep, spin, dead);
  const int any_dead = __syncthreads_or (dead ? 1 : 0);
  if (threadIdx.x == 0) {
    if (any_dead)
      __hip_atomic_fetch_or (sync + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned done = __hip_atomic_fetch_add (sync + 1, 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
    if (done == gridDim.x - 1) {
      __hip_atomic_store (sync + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store (sync, ep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// W_{J,J+d} = T_J U_{J,J+d} in place of U_{J,J+d}, d = 1..D: one workgroup
// per (J, d, mode), T_J and the block staged in LDS, sums in a fixed order
__global__ __launch_bounds__(256) void k_panel_unit (double* P, size_t pstride, int D, int nbp)
{
  extern __shared__ double sm[];
  double* T = sm;                                   // [64][65]
  double* B = sm + PB * (PB + 1);                   // [64][65]
  const int mode = blockIdx.y, J = blockIdx.x / D, d = 1 + blockIdx.x % D;
  if (J >= nbp) return;
  const double* base = P + (size_t) mode * pstride + (size_t) J * (D + 1) * PB * PB;
  double* blk = P + (size_t) mode * pstride + ((size_t) J * (D + 1) + d) * PB * PB;
  for (int k = threadIdx.x; k < PB * PB; k += blockDim.x) {
    const int t = k / PB, c = k % PB;
    T[t * (PB + 1) + c] = base[pidx (t, c)];
    B[t * (PB + 1) + c] = blk[pidx (t, c)];
  }
  __syncthreads ();
  for (int k = threadIdx.x; k < PB * PB; k += blockDim.x) {
    const int t = k / PB, c = k % PB;
    double acc = 0.0;
    for (int s = t; s < PB; s++) acc = fma (T[t * (PB + 1) + s], B[s * (PB + 1) + c], acc);
    blk[pidx (t, c)] = acc;
  }
}

bool unit_ok (int D) { return D >= 3 && D <= 5; }

bool band_ll ()
{
  static const bool ll = [] { const char* e = std::getenv ("SX_BAND_LL"); return !e || std::atoi (e); } ();
  return ll;
}

template <int NR, int D>
int launch_unit_d (const sx_msys* s, int G, int stage, hipStream_t st)
{
  const size_t lds = sizeof (double) * ((size_t) (D + 3) * UW * PB * NR);
  auto kern = band_ll () ? k_unit_sweeps<NR, D, true> : k_unit_sweeps<NR, D, false>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  const int pieces = (stage == 1 || stage == 3) ? 2 : 1;
  const int nblocks = (s->nmodes + 7) / 8 * 8 * G * pieces;
  hipLaunchKernelGGL (kern, dim3 (nblocks), dim3 (UW * 64), lds, st, s->d_swF, s->d_swR,
                      s->sw_stride, s->d_nsolve, s->d_rhs, s->nmax, s->nmodes, G,
                      s->d_tw, stage, s->tw_boff, s->d_rhsb, s->nmaxb, s->m->nband - 1);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

template <int NR>
int launch_unit (const sx_msys* s, int G, int stage, hipStream_t st)
{
  switch (s->sw_D) {
  case 3: return launch_unit_d<NR, 3> (s, G, stage, st);
  case 4: return launch_unit_d<NR, 4> (s, G, stage, st);
  case 5: return launch_unit_d<NR, 5> (s, G, stage, st);
  default: break;
  }
  set_error ("unit panel sweep: unsupported block count");
  return SX_ERR_LIMIT;
}

// polls before a wait of the split sweeps gives up: XSPIN, or SX_SPLIT_SPIN
// in the environment (-1: every wait gives up at once, which tests use to
// force the failure path)
int split_spin ()
{
  const char* e = std::getenv ("SX_SPLIT_SPIN");
  return e ? std::max (-1, std::atoi (e)) : XSPIN;
}

template <int NR, int D, int R>
size_t split_lds ()
{
  constexpr int H = sp_h<D, R> (), FBM = sp_fbmax<D, R> ();
  return sizeof (double) * (size_t) std::max (H + 3, FBM + 1) * UW * PB * NR;
}

// workgroups of k_split_sweeps<NR, D> the device holds at once; 0 when
// the kernel uses scratch (private memory): the occupancy query does not
// count the runtime's limit on waves holding scratch at once, so a grid it
// reports as fitting could leave a partner workgroup waiting for scratch
template <int NR, int D, int R, bool LL, int PD>
int split_capacity ()
{
  static int cap = -1;
  if (cap < 0) {
    int dev = 0, cus = 0, per = 0;
    auto kern = k_split_sweeps<NR, D, R, LL, PD>;
    const int lds = (int) split_lds<NR, D, R> ();
    hipFuncAttributes fa;
    if (hipGetDevice (&dev) != hipSuccess ||
        hipDeviceGetAttribute (&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipFuncSetAttribute ((const void*) kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                             lds) != hipSuccess ||
        hipFuncGetAttributes (&fa, (const void*) kern) != hipSuccess || fa.localSizeBytes != 0 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor (&per, (const void*) kern, UW * 64,
                                                      (size_t) lds) != hipSuccess)
      cap = 0;
    else
      cap = cus * per;
  }
  return cap;
}

template <int NR, int D, int R>
int launch_split_d (const sx_msys* s, int G, int stage, hipStream_t st, bool probe)
{
  const int pieces = (stage == 1 || stage == 3) ? 2 : 1;
  const int nblocks = (s->nmodes + 7) / 8 * 8 * G * R * pieces;
  // panels two steps ahead: three (PD = 3) need more than the 256 VGPRs
  // of two waves per SIMD for D > 3 and spill
  const bool ll = band_ll ();
  if (probe)
    return nblocks <= (ll ? split_capacity<NR, D, R, true, 2> () : split_capacity<NR, D, R, false, 2> ())
               ? SX_OK : SX_ERR_LIMIT;
  const size_t lds = split_lds<NR, D, R> ();
  auto kern = ll ? k_split_sweeps<NR, D, R, true, 2> : k_split_sweeps<NR, D, R, false, 2>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (nblocks), dim3 (UW * 64), lds, st, s->d_swF, s->d_swR,
                      s->sw_stride, s->d_nsolve, s->d_rhs, s->nmax, s->nmodes, G,
                      s->d_tw, stage, s->tw_boff, s->d_rhsb, s->nmaxb, s->m->nband - 1,
                      s->d_xch, s->d_xep, split_spin ());
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// workgroups per chain of the split sweeps (near and far).  split_body
// takes the role count as a parameter; three roles (the far blocks on two
// workgroups) let waits run out on the 32-mode Helmholtz system and are not
// instantiated.
constexpr int SPLIT_ROLES = 2;

// probe: SX_OK iff every workgroup of the launch fits on the device at once
template <int NR>
int launch_split (const sx_msys* s, int G, int stage, hipStream_t st, bool probe = false)
{
  switch (s->sw_D) {
  case 3: return launch_split_d<NR, 3, SPLIT_ROLES> (s, G, stage, st, probe);
  case 4: return launch_split_d<NR, 4, SPLIT_ROLES> (s, G, stage, st, probe);
  case 5: return launch_split_d<NR, 5, SPLIT_ROLES> (s, G, stage, st, probe);
  default: break;
  }
  set_error ("split panel sweep: unsupported block count");
  return SX_ERR_LIMIT;
}

// message rings of the split sweeps: up to 2 pieces x nmodes x 6
// right-hand sides (one ring pair per right-hand side and chain), and the
// system's epoch, arrival count and status words (k_split_sweeps)
constexpr int XMAXR = 6;
int split_alloc (sx_msys* s, hipStream_t st)
{
  const size_t rows = (size_t) 2 * s->nmodes * XMAXR;
  s->xch_bytes = rows * SPLIT_ROLES * XR * PB * 16;  // one direction per role
  SX_HIP (hipMalloc (&s->d_xch, s->xch_bytes));
  SX_HIP (hipMalloc (&s->d_xep, sizeof (unsigned) * 4));
  SX_HIP (hipMemsetAsync (s->d_xch, 0, s->xch_bytes, st));
  SX_HIP (hipMemsetAsync (s->d_xep, 0, sizeof (unsigned) * 4, st));
  return SX_OK;
}

// scale the off-diagonal blocks of nbp panels per mode to W = T_J U
int unit_panels (const sx_msys* s, double* P, int nbp, hipStream_t st)
{
  const size_t lds = sizeof (double) * 2 * PB * (PB + 1);
  SX_HIP (hipFuncSetAttribute ((const void*) k_panel_unit,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (k_panel_unit, dim3 ((unsigned) (nbp * s->sw_D), s->nmodes), dim3 (256),
                      lds, st, P, s->sw_stride, s->sw_D, nbp);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// rows per wavefront, rounded up to a multiple of 4
#define SX_QS_CASES(B) \
  switch (qs) { \
  case 4: return launch_sweeps<B, 4> (s, st);   case 8: return launch_sweeps<B, 8> (s, st); \
  case 12: return launch_sweeps<B, 12> (s, st); case 16: return launch_sweeps<B, 16> (s, st); \
  case 20: return launch_sweeps<B, 20> (s, st); case 24: return launch_sweeps<B, 24> (s, st); \
  case 28: return launch_sweeps<B, 28> (s, st); case 32: return launch_sweeps<B, 32> (s, st); \
  case 36: return launch_sweeps<B, 36> (s, st); case 40: return launch_sweeps<B, 40> (s, st); \
  case 44: return launch_sweeps<B, 44> (s, st); case 48: return launch_sweeps<B, 48> (s, st); \
  case 52: return launch_sweeps<B, 52> (s, st); case 56: return launch_sweeps<B, 56> (s, st); \
  case 60: return launch_sweeps<B, 60> (s, st); case 64: return launch_sweeps<B, 64> (s, st); \
  case 72: return launch_sweeps<B, 72> (s, st); case 80: return launch_sweeps<B, 80> (s, st); \
  case 88: return launch_sweeps<B, 88> (s, st); case 96: return launch_sweeps<B, 96> (s, st); \
  default: break; }

}  // namespace

// Block size: the largest power of two <= 64 that keeps the inverted
// diagonal triangles inside the band (B <= kd+1).  SX_BAND_BLOCK in the
// environment caps it (used by the tuner).
int band_block (int nband)
{
  int cap = 64;
  if (const char* e = std::getenv ("SX_BAND_BLOCK")) cap = std::max (1, std::atoi (e));
  int B = 64;
  while (B > 1 && (B > nband || B > cap)) B >>= 1;
  return B;
}

static int rows_per_wave (int nband)
{
  int qs = (nband + SW_NW - 1) / SW_NW;
  qs = (qs + 3) / 4 * 4;
  if (qs > 64) qs = (qs + 7) / 8 * 8;
  return qs;
}

// Unit-diagonal panels and their sweeps for D = 3..5 (SX_BAND_UNIT=0 in the
// environment keeps the triangle in the chain, k_panel_sweeps)
static bool use_unit (int D)
{
  if (const char* e = std::getenv ("SX_BAND_UNIT"))
    if (std::atoi (e) == 0) return false;
  return unit_ok (D);
}

// The panel kernels address a piece's panels with 32-bit byte offsets (and
// a buffer descriptor of 32-bit size): systems whose panels per mode would
// pass 2^31 bytes (about 700,000 rows at D = 5) use the offset-major sweeps.
static bool use_panels (int nband, int nmax)
{
  if (const char* e = std::getenv ("SX_BAND_SWEEP"))
    if (std::string (e) == "offset") return false;
  const size_t bytes = sizeof (double) * ((size_t) (nmax + PB - 1) / PB + 3) *
                       (panel_D (nband) + 1) * PB * PB;
  return panel_D (nband) <= PMAXD && bytes < ((size_t) 1 << 31);
}

size_t band_sweep_bytes_per_mode (int nmax, int nband)
{
  if (use_panels (nband, nmax))
    return (size_t) ((nmax + PB - 1) / PB + 1) * (panel_D (nband) + 1) * PB * PB;
  const int B = band_block (nband);
  return (size_t) ((nmax + B - 1) / B) * nband * B;
}

// Split of every mode for the twisted solve (the TW ints per mode), or
// false when some mode is too small for it to shorten the chain of block
// steps by a quarter.  A sweep step is bound by how many loads one CU
// keeps in flight, so the second piece pays off while the modes leave CUs
// idle: systems of at most 32 modes.  (At 64 modes of one field, the DNS
// pressure, the twisted sweeps run 176 us against 245 us single-piece, but
// over 100 channel steps the pressure, whose mode 0 is pure Neumann with
// condition ~2.6e5, then drifted from the oracle to 1.06 times its parity
// bar, against 0.95 single-piece, whose elimination order is the oracle's;
// the grouped velocity, 6 right-hand sides per mode, runs 301 us twisted
// against 271 us.)  SX_BAND_TWIST=0/1 in the environment forces it off/on.
static bool twist_split (const sx_msys* s, std::vector<int>& tw)
{
  if (const char* e = std::getenv ("SX_BAND_TWIST")) {
    if (std::atoi (e) == 0) return false;
  } else if (s->nmodes > 32) {
    return false;
  }
  const int D = panel_D (s->m->nband), S = PB * D;
  const bool ll = band_ll () && use_unit (D);
  tw.assign ((size_t) TW * s->nmodes, 0);
  for (int md = 0; md < s->nmodes; md++) {
    const int n = s->nsolve[md], nbtot = (n + PB - 1) / PB;
    int best = -1, bestT = 0;
    for (int nbT = 1; nbT < nbtot; nbT++) {
      const int nbr = n - nbT * PB - S;
      if (nbr < 1) break;
      const int nbB = (nbr + PB - 1) / PB;
      // stage 1 (both forward), stage 2 (S both ways), stage 3 (both
      // backward; the right-looking sweep of the bottom piece also passes
      // its D separator blocks, the left-looking one starts below them)
      const int cost = std::max (nbT, nbB) + 2 * D + std::max (nbT, ll ? nbB : nbB + D);
      if (best < 0 || cost < best) { best = cost; bestT = nbT; }
    }
    if (best < 0 || 4 * best > 3 * 2 * nbtot) return false;
    const int m = bestT * PB, nbr = n - m - S, nbB = (nbr + PB - 1) / PB;
    int* t = &tw[(size_t) TW * md];
    t[TW_NBT] = bestT;
    t[TW_NBB] = nbB;
    t[TW_PAD] = nbB * PB - nbr;
    t[TW_M] = m;
    t[TW_NTOP] = m + S;
    t[TW_NBOT] = (nbB + D) * PB;
    t[TW_N] = n;
  }
  return true;
}

// Panels of one piece: T_J and the off-diagonal blocks of the factor U
// (band `src`, per-mode sizes `d_n`) and of its reversal, into P + off.
// the backward sweeps read a reversed copy of the panels unless they are
// the left-looking sweeps on the forward panels (unit-diagonal systems)
static bool reversed_panels (const sx_msys* s) { return !(s->sw_unit && band_ll ()); }

static int build_panels (sx_msys* s, const double* src, double* V, size_t stride,
                         const int* d_n, int nbp, size_t off, hipStream_t st)
{
  const int ld = s->m->nband, D = s->sw_D;
  const bool rev = reversed_panels (s);
  if (rev)
    hipLaunchKernelGGL (k_band_reverse, dim3 ((unsigned) ((stride + 255) / 256), s->nmodes),
                        dim3 (256), 0, st, src, V, stride, ld, d_n);
  const double* from[2] = {src, V};
  double* to[2] = {s->d_swF + off, rev ? s->d_swR + off : nullptr};
  for (int k = 0; k < (rev ? 2 : 1); k++) {
    hipLaunchKernelGGL (k_panel_tinv, dim3 (nbp, s->nmodes), dim3 (64), 0, st, from[k],
                        stride, ld, d_n, to[k], s->sw_stride, D);
    const size_t tot = (size_t) nbp * D * PB * PB;
    hipLaunchKernelGGL (k_panel_offdiag, dim3 ((unsigned) ((tot + 255) / 256), s->nmodes),
                        dim3 (256), 0, st, from[k], stride, ld, d_n, to[k], s->sw_stride, D,
                        nbp);
    if (s->sw_unit)
      if (int rc = unit_panels (s, to[k], nbp, st)) return rc;
  }
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// Twisted factorisation: the bottom piece built and factorised through B,
// S's Schur complement moved into the top piece, the top piece factorised,
// then the panels of both pieces (top at offset 0, bottom at tw_boff).
static int band_factor_twisted (sx_msys* s, const std::vector<int>& tw, hipStream_t st,
                                int* d_status)
{
  const int ld = s->m->nband, D = panel_D (ld), S = PB * D, nm = s->nmodes;
  std::vector<int> aux (3 * nm);
  int nbtm = 0, nbbm = 0;
  for (int md = 0; md < nm; md++) {
    const int* t = &tw[(size_t) TW * md];
    aux[md] = t[TW_NTOP];
    aux[nm + md] = t[TW_NBOT];
    aux[2 * nm + md] = t[TW_NBOT] - S;
    nbtm = std::max (nbtm, t[TW_NTOP] / PB);
    nbbm = std::max (nbbm, t[TW_NBOT] / PB);
  }
  SX_HIP (hipMalloc (&s->d_tw, sizeof (int) * tw.size ()));
  SX_HIP (hipMemcpy (s->d_tw, tw.data (), sizeof (int) * tw.size (), hipMemcpyHostToDevice));
  int* d_aux = nullptr;
  SX_HIP (hipMalloc (&d_aux, sizeof (int) * aux.size ()));
  SX_HIP (hipMemcpy (d_aux, aux.data (), sizeof (int) * aux.size (), hipMemcpyHostToDevice));
  const int *d_ntop = d_aux, *d_nbot = d_aux + nm, *d_nelim = d_aux + 2 * nm;
  const size_t rstride = (size_t) nbbm * PB * ld;
  double *R = nullptr, *V = nullptr;
  const size_t vstride = std::max (rstride, s->band_stride);
  SX_HIP (hipMalloc (&R, sizeof (double) * nm * rstride));
  SX_HIP (hipMalloc (&V, sizeof (double) * nm * vstride));
  SX_HIP (hipMemsetAsync (V, 0, sizeof (double) * nm * vstride, st));
  const size_t lds = sizeof (double) * std::max (ld, 1);
  hipLaunchKernelGGL (k_twist_rbuild, dim3 ((unsigned) ((rstride + 255) / 256), nm),
                      dim3 (256), 0, st, s->d_bandU, s->band_stride, R, rstride, ld, s->d_tw);
  hipLaunchKernelGGL (k_band_factor, dim3 (nm), dim3 (1024), lds, st, R, rstride, ld, d_nbot,
                      d_status, d_nelim);
  hipLaunchKernelGGL (k_twist_sback, dim3 ((unsigned) (((size_t) S * ld + 255) / 256), nm),
                      dim3 (256), 0, st, s->d_bandU, s->band_stride, R, rstride, ld, s->d_tw, S);
  hipLaunchKernelGGL (k_band_factor, dim3 (nm), dim3 (1024), lds, st, s->d_bandU,
                      s->band_stride, ld, d_ntop, d_status, nullptr);
  SX_HIP (hipGetLastError ());
  s->sw_panel = 1;
  s->sw_D = D;
  s->sw_unit = use_unit (D);
  if (s->sw_unit)
    if (int rc = split_alloc (s, st)) return rc;
  s->sw_block = PB;
  s->tw = 1;
  s->tw_S = S;
  s->tw_boff = (size_t) (nbtm + 1) * (D + 1) * PB * PB;
  s->sw_stride = s->tw_boff + (size_t) (nbbm + 1) * (D + 1) * PB * PB;
  SX_HIP (hipMalloc (&s->d_swF, sizeof (double) * nm * s->sw_stride));
  if (reversed_panels (s)) SX_HIP (hipMalloc (&s->d_swR, sizeof (double) * nm * s->sw_stride));
  if (int rc = build_panels (s, s->d_bandU, V, s->band_stride, d_ntop, nbtm + 1, 0, st))
    return rc;
  SX_HIP (hipMemsetAsync (V, 0, sizeof (double) * nm * vstride, st));
  if (int rc = build_panels (s, R, V, rstride, d_nbot, nbbm + 1, s->tw_boff, st)) return rc;
  s->nmaxb = nbbm * PB;
  const size_t nb_rhs = (size_t) nm * s->rhs_cap * s->nmaxb;
  SX_HIP (hipMalloc (&s->d_rhsb, sizeof (double) * nb_rhs));
  SX_HIP (hipMemsetAsync (s->d_rhsb, 0, sizeof (double) * nb_rhs, st));
  SX_HIP (hipStreamSynchronize (st));
  (void) hipFree (R);
  (void) hipFree (V);
  (void) hipFree (d_aux);
  return SX_OK;
}

int band_factor (sx_msys* s, hipStream_t st, int* d_status)
{
  const int ld = s->m->nband;
  const bool panels = use_panels (ld, s->nmax);
  std::vector<int> tw;
  if (panels && twist_split (s, tw)) return band_factor_twisted (s, tw, st, d_status);
  hipLaunchKernelGGL (k_band_factor, dim3 (s->nmodes), dim3 (1024),
                      sizeof (double) * std::max (ld, 1), st, s->d_bandU,
                      s->band_stride, ld, s->d_nsolve, d_status, nullptr);
  SX_HIP (hipGetLastError ());
  const int B = panels ? PB : band_block (ld);
  const int nblk = (s->nmax + B - 1) / B;
  if (panels) {
    s->sw_unit = use_unit (panel_D (ld));
  }
  const bool rev = !panels || reversed_panels (s);
  double* V = nullptr;
  if (rev) {
    SX_HIP (hipMalloc (&V, sizeof (double) * s->nmodes * s->band_stride));
    SX_HIP (hipMemsetAsync (V, 0, sizeof (double) * s->nmodes * s->band_stride, st));
    hipLaunchKernelGGL (k_band_reverse,
                        dim3 ((unsigned) ((s->band_stride + 255) / 256), s->nmodes),
                        dim3 (256), 0, st, s->d_bandU, V, s->band_stride, ld, s->d_nsolve);
  }
  if (panels) {
    const int D = panel_D (ld);
    s->sw_panel = 1;
    s->sw_D = D;
    if (s->sw_unit)
      if (int rc = split_alloc (s, st)) return rc;
    s->sw_block = PB;
    s->sw_stride = band_sweep_bytes_per_mode (s->nmax, ld);
    SX_HIP (hipMalloc (&s->d_swF, sizeof (double) * s->nmodes * s->sw_stride));
    if (rev) SX_HIP (hipMalloc (&s->d_swR, sizeof (double) * s->nmodes * s->sw_stride));
    const double* src[2] = {s->d_bandU, V};
    double* dst[2] = {s->d_swF, s->d_swR};
    // panels 0..nblk (per-mode nb <= nblk; everything past a mode's own
    // blocks is zero, including its zero panel at index nb)
    for (int k = 0; k < (rev ? 2 : 1); k++) {
      hipLaunchKernelGGL (k_panel_tinv, dim3 (nblk + 1, s->nmodes), dim3 (64), 0, st, src[k],
                          s->band_stride, ld, s->d_nsolve, dst[k], s->sw_stride, D);
      const size_t tot = (size_t) (nblk + 1) * D * PB * PB;
      hipLaunchKernelGGL (k_panel_offdiag, dim3 ((unsigned) ((tot + 255) / 256), s->nmodes),
                          dim3 (256), 0, st, src[k], s->band_stride, ld, s->d_nsolve,
                          dst[k], s->sw_stride, D, nblk + 1);
      if (s->sw_unit)
        if (int rc = unit_panels (s, dst[k], nblk + 1, st)) return rc;
    }
    SX_HIP (hipGetLastError ());
    SX_HIP (hipStreamSynchronize (st));
    (void) hipFree (V);
    return SX_OK;
  }
  s->sw_panel = 0;
  if (nblk > 0) {
    hipLaunchKernelGGL (k_band_tinv, dim3 (nblk, s->nmodes), dim3 (64), 0, st,
                        s->d_bandU, s->band_stride, ld, B, s->d_nsolve);
    hipLaunchKernelGGL (k_band_tinv, dim3 (nblk, s->nmodes), dim3 (64), 0, st, V,
                        s->band_stride, ld, B, s->d_nsolve);
  }
  SX_HIP (hipGetLastError ());
  s->sw_block = B;
  s->sw_stride = band_sweep_bytes_per_mode (s->nmax, ld);
  SX_HIP (hipMalloc (&s->d_swF, sizeof (double) * s->nmodes * s->sw_stride));
  SX_HIP (hipMalloc (&s->d_swR, sizeof (double) * s->nmodes * s->sw_stride));
  const unsigned gx = (unsigned) ((s->sw_stride + 255) / 256);
  hipLaunchKernelGGL (k_band_blockify, dim3 (gx, s->nmodes), dim3 (256), 0, st,
                      s->d_bandU, s->d_swF, s->band_stride, s->sw_stride, ld, B,
                      s->d_nsolve);
  hipLaunchKernelGGL (k_band_blockify, dim3 (gx, s->nmodes), dim3 (256), 0, st, V,
                      s->d_swR, s->band_stride, s->sw_stride, ld, B, s->d_nsolve);
  SX_HIP (hipGetLastError ());
  SX_HIP (hipStreamSynchronize (st));
  (void) hipFree (V);
  return SX_OK;
}

int band_solve (const sx_msys* s, int nr, hipStream_t st, int moves)
{
  if (s->sw_panel) {
    // right-hand sides per workgroup: SX_PANEL_NR in the environment, else
    // the fewest that keep one workgroup per CU (256 CUs), counting both
    // pieces of a twisted solve
    const int pieces = s->tw ? 2 : 1, slots = (s->nmodes + 7) / 8 * 8 * pieces;
    int per = nr;
    if (const char* e = std::getenv ("SX_PANEL_NR")) per = std::atoi (e);
    else
      for (int c : {1, 2, 3, 6})
        if (nr % c == 0 && slots * (nr / c) <= 256) { per = c; break; }
    if (s->sw_unit && per > 2) per = nr % 2 == 0 ? 2 : 1;
    if (per < 1 || nr % per) per = nr;
    // Split sweeps (two workgroups per chain) for twisted systems, where the
    // pieces leave the chain short: the right-hand sides per workgroup
    // (1, 2 or 3) bring the grid closest to one workgroup per CU (256),
    // among the launches that fit on the device at once (with the
    // left-looking backward sweeps 256 workgroups measured faster than 128:
    // 0.208 against 0.231 ms at 32 modes).  SX_BAND_SPLIT=0/1 in the
    // environment forces them off/on (on: also for single-piece systems).
    int sper = 0;
    if (s->sw_unit && s->d_xch && nr <= XMAXR && !s->no_split) {
      const char* e = std::getenv ("SX_BAND_SPLIT");
      const bool want = e ? std::atoi (e) != 0 : s->tw != 0;
      if (want) {
        const char* ep = std::getenv ("SX_PANEL_NR");
        const int stg = s->tw ? 1 : 0, slots1 = (s->nmodes + 7) / 8 * 8 * (s->tw ? 2 : 1);
        int best = 1 << 30;
        for (int c : {1, 2, 3}) {
          if (nr % c || (ep && std::atoi (ep) != c)) continue;
          const int fits = c == 1 ? launch_split<1> (s, nr, stg, st, true)
                         : c == 2 ? launch_split<2> (s, nr / 2, stg, st, true)
                                  : launch_split<3> (s, nr / 3, stg, st, true);
          const int dist = std::abs (slots1 * (nr / c) * SPLIT_ROLES - 256);
          if (fits == SX_OK && dist <= best) { best = dist; sper = c; }
        }
      }
    }
    if (sper) per = sper;
    const int G = nr / per;
    s->last_kernel = sper ? 3 : s->sw_unit ? 2 : 1;
    s->last_nr = per;
    auto sweep = [&] (int stage) {
      if (sper) {
        if (per == 1) return launch_split<1> (s, G, stage, st);
        if (per == 2) return launch_split<2> (s, G, stage, st);
        return launch_split<3> (s, G, stage, st);
      }
      if (s->sw_unit) {
        if (per == 1) return launch_unit<1> (s, G, stage, st);
        if (per == 2) return launch_unit<2> (s, G, stage, st);
        set_error ("band solve: unit panel sweeps take 1 or 2 right-hand sides per workgroup");
        return (int) SX_ERR_ARG;
      }
      switch (per) {
      case 1: return launch_panel<1> (s, G, stage, st);
      case 2: return launch_panel<2> (s, G, stage, st);
      case 3: return launch_panel<3> (s, G, stage, st);
      case 4: return launch_panel<4> (s, G, stage, st);
      case 6: return launch_panel<6> (s, G, stage, st);
      default: break;
      }
      set_error ("band solve: 1, 2, 3, 4 or 6 right-hand sides per workgroup");
      return (int) SX_ERR_ARG;
    };
    if (!s->tw) return sweep (0);
    // twisted: gather B, both forward, S, both backward, scatter B
    auto move = [&] (int what) {
      const int rows = (what == 1 || what == 2) ? s->tw_S : s->nmaxb;
      hipLaunchKernelGGL (k_twist_move, dim3 ((rows + 255) / 256, nr, s->nmodes), dim3 (256),
                          0, st, s->d_rhs, s->nmax, s->d_rhsb, s->nmaxb, nr, s->d_tw,
                          s->tw_S, what);
      const hipError_t e = hipGetLastError ();
      return e == hipSuccess ? (int) SX_OK : hip_fail (e, "k_twist_move");
    };
    // moves & 1: gather B here (else the caller's assembly wrote the bottom
    // piece's right-hand side); moves & 2: scatter B (else the caller's
    // recovery reads the bottom piece's solution)
    int rc;
    if ((moves & 1) && (rc = move (0))) return rc;
    if (sper || s->sw_unit) {
      // the split and the unit-diagonal sweeps fold the separator's combine
      // and x_S copy into stage 2
      if ((rc = sweep (1)) || (rc = sweep (2)) || (rc = sweep (3))) return rc;
    } else if ((rc = sweep (1)) || (rc = move (1)) || (rc = sweep (2)) || (rc = move (2)) ||
               (rc = sweep (3))) {
      return rc;
    }
    if ((moves & 2) && (rc = move (3))) return rc;
    return SX_OK;
  }
  if (nr != 2) {
    set_error ("band solve: the offset-major sweeps take 2 right-hand sides per mode");
    return SX_ERR_ARG;
  }
  const int qs = rows_per_wave (s->m->nband);
  switch (s->sw_block) {
  case 64: SX_QS_CASES (64) break;
  case 32:
    if (qs == 4) return launch_sweeps<32, 4> (s, st);
    if (qs == 8) return launch_sweeps<32, 8> (s, st);
    break;
  case 16: return launch_sweeps<16, 4> (s, st);
  case 8: return launch_sweeps<8, 4> (s, st);
  case 4: return launch_sweeps<4, 4> (s, st);
  case 2: return launch_sweeps<2, 4> (s, st);
  case 1: return launch_sweeps<1, 4> (s, st);
  default: break;
  }
  set_error ("band solver: band width " + std::to_string (s->m->nband) +
             " outside the instantiated range (nband <= 768)");
  return SX_ERR_LIMIT;
}

// ---------------------------------------------------------------------------
// Fused-batch plan of band_solve without the split sweeps (the batch's
// solves run side by side, so no launch may need its whole grid resident)
// ---------------------------------------------------------------------------
namespace {
template <int NR, int D, bool LL> struct FnUnit {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_unit_sweeps_body<NR, D, LL> (a...); }
};
template <int NR, int D> struct FnPanel {
  template <typename... A> __device__ __forceinline__ void operator() (A... a) const
  { k_panel_sweeps_body<NR, D> (a...); }
};
// blockIdx.y = mode * nr + column (k_twist_move has them in z and y)
struct FnTwistMove {
  __device__ __forceinline__ void operator() (double* rhs, int nmax, double* rhsb, int nmaxb,
                                              int nr, const int* tw, int S, int what) const
  { k_twist_move_body (rhs, nmax, rhsb, nmaxb, nr, tw, S, what, blockIdx.y / nr,
                       blockIdx.y % nr); }
};

template <int NR, int D>
void plan_unit_d (BatchPlan& P, const sx_msys* s, int G, int stage)
{
  const size_t lds = sizeof (double) * ((size_t) (D + 3) * UW * PB * NR);
  const int pieces = (stage == 1 || stage == 3) ? 2 : 1;
  const int nblocks = (s->nmodes + 7) / 8 * 8 * G * pieces;
  auto rec = [&] (auto fn) {
    batch_rec<UW * 64, decltype (fn)> (P, dim3 (nblocks), UW * 64, lds,
                                       (const double*) s->d_swF, (const double*) s->d_swR,
                                       s->sw_stride, (const int*) s->d_nsolve, s->d_rhs, s->nmax,
                                       s->nmodes, G, (const int*) s->d_tw, stage, s->tw_boff,
                                       s->d_rhsb, s->nmaxb, s->m->nband - 1);
  };
  if (band_ll ()) rec (FnUnit<NR, D, true> {});
  else rec (FnUnit<NR, D, false> {});
}

template <int NR, int D>
void plan_panel_d (BatchPlan& P, const sx_msys* s, int G, int stage)
{
  constexpr int NW = PanelSweep<NR, D>::NW;
  const size_t lds = sizeof (double) * ((size_t) PB * NR * (2 * D + NW));
  const int pieces = (stage == 1 || stage == 3) ? 2 : 1;
  const int nblocks = (s->nmodes + 7) / 8 * 8 * G * pieces;
  batch_rec<panel_nw (D) * 64, FnPanel<NR, D>> (P, dim3 (nblocks), NW * 64, lds,
                                               (const double*) s->d_swF,
                                               (const double*) s->d_swR, s->sw_stride,
                                               (const int*) s->d_nsolve, s->d_rhs, s->nmax,
                                               s->nmodes, G, (const int*) s->d_tw, stage,
                                               s->tw_boff, s->d_rhsb, s->nmaxb,
                                               s->m->nband - 1);
}

template <int NR>
int plan_sweep (BatchPlan& P, const sx_msys* s, int G, int stage)
{
  if (s->sw_unit) {
    switch (s->sw_D) {
    case 3: plan_unit_d<NR, 3> (P, s, G, stage); return SX_OK;
    case 4: plan_unit_d<NR, 4> (P, s, G, stage); return SX_OK;
    case 5: plan_unit_d<NR, 5> (P, s, G, stage); return SX_OK;
    default: return SX_ERR_LIMIT;
    }
  }
  switch (s->sw_D) {
  case 1: plan_panel_d<NR, 1> (P, s, G, stage); return SX_OK;
  case 2: plan_panel_d<NR, 2> (P, s, G, stage); return SX_OK;
  case 3: plan_panel_d<NR, 3> (P, s, G, stage); return SX_OK;
  case 4: plan_panel_d<NR, 4> (P, s, G, stage); return SX_OK;
  case 5: plan_panel_d<NR, 5> (P, s, G, stage); return SX_OK;
  case 6: plan_panel_d<NR, 6> (P, s, G, stage); return SX_OK;
  case 7: plan_panel_d<NR, 7> (P, s, G, stage); return SX_OK;
  case 8: plan_panel_d<NR, 8> (P, s, G, stage); return SX_OK;
  default: return SX_ERR_LIMIT;
  }
}
}  // namespace

// The right-hand sides per sweep workgroup: P.tune when it divides nr and
// the kernel takes it, else band_solve's default (the fewest that keep one
// workgroup per CU, at most 2 for the unit-diagonal sweeps).
int plan_band_solve (BatchPlan& P, const sx_msys* s, int nr)
{
  if (!s->sw_panel || nr != 2) return SX_ERR_LIMIT;
  const int pieces = s->tw ? 2 : 1, slots = (s->nmodes + 7) / 8 * 8 * pieces;
  int per = nr;
  for (int c : {1, 2})
    if (nr % c == 0 && slots * (nr / c) <= 256) { per = c; break; }
  if (P.tune == 1 || P.tune == 2) per = P.tune;
  const int G = nr / per;
  auto sweep = [&] (int stage) {
    return per == 1 ? plan_sweep<1> (P, s, G, stage) : plan_sweep<2> (P, s, G, stage);
  };
  auto move = [&] (int what) {
    const int rows = (what == 1 || what == 2) ? s->tw_S : s->nmaxb;
    batch_rec<256, FnTwistMove> (P, dim3 ((rows + 255) / 256, nr * s->nmodes), 256, 0,
                                 s->d_rhs, s->nmax, s->d_rhsb, s->nmaxb, nr,
                                 (const int*) s->d_tw, s->tw_S, what);
    P.phase++;
  };
  int rc;
  if (!s->tw) {
    if ((rc = sweep (0))) return rc;
    P.phase++;
    return SX_OK;
  }
  move (0);
  for (int stage = 1; stage <= 3; stage++) {
    if ((rc = sweep (stage))) return rc;
    P.phase++;
    // the unit-diagonal sweeps fold the separator's combine and x_S copy
    // into stage 2; their phases stay aligned with the other systems'
    // (the fused launches of one phase serve every system of the batch)
    if (stage < 3) {
      if (s->sw_unit) P.phase++;
      else move (stage);
    }
  }
  move (3);
  return SX_OK;
}

// status word of the split sweeps: SX_OK, or SX_ERR_SYNC (and the word
// cleared) when a wait of some launch since the last query gave up
int split_status (sx_msys* s, hipStream_t st)
{
  if (!s->d_xep) return SX_OK;
  unsigned w[4] = {0, 0, 0, 0};
  SX_HIP (hipMemcpyAsync (w, s->d_xep, sizeof w, hipMemcpyDeviceToHost, st));
  SX_HIP (hipStreamSynchronize (st));
  if (!(w[2] & 1u)) return SX_OK;
  const unsigned zero = 0;
  SX_HIP (hipMemcpyAsync (s->d_xep + 2, &zero, sizeof zero, hipMemcpyHostToDevice, st));
  SX_HIP (hipStreamSynchronize (st));
  set_error ("band solve: a wait of the split sweeps gave up (a partner workgroup was not "
             "resident); the affected solutions hold NaN");
  return SX_ERR_SYNC;
}


}  // namespace sx
