// sx_fourier.hip -- real transforms along z of every (element, node) line
// of a field of nz planes, in the reference's packing and scaling
// (AuxField::transform, src/auxfield.cpp:1157-1194 -> dDFTr,
// femlib/fourier.c:40-165):
//
//   forward (sign +1):  plane 0 = Re X_0, plane 1 = Re X_{nz/2},
//                       planes 2k, 2k+1 = Re X_k, Im X_k   (k = 1..nz/2-1),
//                       X_k = (1/nz) sum_n x_n exp(-2 pi i k n / nz);
//   inverse (sign -1):  x_n = X_0 + (-1)^n X_{nz/2}
//                             + 2 sum_k Re(X_k exp(+2 pi i k n / nz)).
//
// For nz a power of two (8..1024) one workgroup transforms a tile of
// columns in LDS: columns 2q and 2q+1 form the complex sequence
// x_2q + i x_2q+1, which a radix-2 Stockham FFT transforms in natural order;
// the two real spectra are separated (forward) or combined (inverse) from
// Z_k and conj(Z_{nz-k}).  Each tile is read once and written once, so the
// transform is HBM-bound (16 bytes per point per pass) and may run in place.
// Other nz use the packed DFT matrix, one thread per column.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "sx_internal.h"

namespace {

constexpr int FT_THREADS = 256;

template <int N>
struct FftTile {
  static constexpr int TS = (4096 / N) < 4 ? 4 : 4096 / N;   // columns per tile
  static constexpr int NQ = TS / 2;                          // complex sequences
  static size_t lds () { return sizeof (double2) * (2 * (size_t) N * NQ + N / 2); }
};

__device__ __forceinline__ double2 cmul (double2 a, double2 b)
{
  return make_double2 (a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int N>
__global__ __launch_bounds__(FT_THREADS) void k_fft_z (const double* in, double* out, int S,
                                                       int sign, const double2* __restrict__ tw)
{
  constexpr int TS = FftTile<N>::TS, NQ = FftTile<N>::NQ;
  extern __shared__ double2 lds[];
  double2* A = lds;                   // [N][NQ]
  double2* B = A + N * NQ;            // [N][NQ], also the raw tile [N][TS]
  double2* W = B + N * NQ;            // exp(-2 pi i j / N), j < N/2
  double* raw = (double*) B;
  const int s0 = blockIdx.x * TS, tid = threadIdx.x;
  for (int i = tid; i < N / 2; i += FT_THREADS) W[i] = tw[i];
  {
    // the whole tile's loads in flight before any LDS write
    constexpr int PER = N * TS / FT_THREADS;
    static_assert (N * TS % FT_THREADS == 0, "tile must split evenly over the threads");
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int i = tid + u * FT_THREADS, n = i / TS, s = s0 + i % TS;
      v[u] = s < S ? in[(size_t) n * S + s] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PER; u++) raw[tid + u * FT_THREADS] = v[u];
  }
  __syncthreads ();
  if (sign > 0) {
    // Z_q[n] = x_2q[n] + i x_2q+1[n]
    for (int i = tid; i < N * NQ; i += FT_THREADS) {
      const int n = i / NQ, q = i % NQ;
      A[i] = make_double2 (raw[n * TS + 2 * q], raw[n * TS + 2 * q + 1]);
    }
  } else {
    // Z_q[k] = Xa[k] + i Xb[k], spectra extended by X[N-k] = conj X[k]
    for (int i = tid; i < N * NQ; i += FT_THREADS) {
      const int k = i / NQ, q = i % NQ;
      const int kk = k <= N / 2 ? k : N - k;
      double ar, ai, br, bi;
      if (kk == 0) {
        ar = raw[2 * q]; ai = 0.0; br = raw[2 * q + 1]; bi = 0.0;
      } else if (kk == N / 2) {
        ar = raw[TS + 2 * q]; ai = 0.0; br = raw[TS + 2 * q + 1]; bi = 0.0;
      } else {
        ar = raw[2 * kk * TS + 2 * q];       ai = raw[(2 * kk + 1) * TS + 2 * q];
        br = raw[2 * kk * TS + 2 * q + 1];   bi = raw[(2 * kk + 1) * TS + 2 * q + 1];
      }
      if (k > N / 2) { ai = -ai; bi = -bi; }
      A[i] = make_double2 (ar - bi, ai + br);
    }
  }
  __syncthreads ();
  // radix-2 Stockham: y[2j-k] = a + w b, y[2j-k+m] = a - w b, w = W^(k N/2m)
  double2 *x = A, *y = B;
#pragma unroll 1
  for (int m = 1; m < N; m <<= 1) {
    const int ws = N / (2 * m);
    for (int b = tid; b < (N / 2) * NQ; b += FT_THREADS) {
      const int j = b / NQ, q = b % NQ, k = j & (m - 1);
      double2 w = W[k * ws];
      if (sign < 0) w.y = -w.y;
      const double2 a = x[j * NQ + q];
      const double2 t = cmul (x[(j + N / 2) * NQ + q], w);
      y[(2 * j - k) * NQ + q] = make_double2 (a.x + t.x, a.y + t.y);
      y[(2 * j - k + m) * NQ + q] = make_double2 (a.x - t.x, a.y - t.y);
    }
    __syncthreads ();
    double2* sw = x; x = y; y = sw;
  }
  if (sign > 0) {
    const double scale = 1.0 / N;
    for (int i = tid; i < N * TS; i += FT_THREADS) {
      const int p = i / TS, c = i % TS, s = s0 + c;
      if (s >= S) continue;
      const int q = c >> 1, k = p == 0 ? 0 : (p == 1 ? N / 2 : p >> 1);
      const double2 zk = x[k * NQ + q], zr = x[((N - k) & (N - 1)) * NQ + q];
      // Xa = (Z_k + conj Z_{N-k}) / 2,  Xb = (Z_k - conj Z_{N-k}) / 2i
      double re, im;
      if ((c & 1) == 0) { re = 0.5 * (zk.x + zr.x); im = 0.5 * (zk.y - zr.y); }
      else              { re = 0.5 * (zk.y + zr.y); im = -0.5 * (zk.x - zr.x); }
      out[(size_t) p * S + s] = ((p < 2 || (p & 1) == 0) ? re : im) * scale;
    }
  } else {
    for (int i = tid; i < N * TS; i += FT_THREADS) {
      const int n = i / TS, c = i % TS, s = s0 + c;
      if (s >= S) continue;
      const double2 z = x[n * NQ + (c >> 1)];
      out[(size_t) n * S + s] = (c & 1) ? z.y : z.x;
    }
  }
}

// ---------------------------------------------------------------------------
// Two-level FFT for nz = N = R1 R2 (32..256): with n = R2 n1 + n2 and
// k = k1 + R1 k2,
//   X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_N^(n2 k1) sum_n1 W_R1^(n1 k1) x[R2 n1 + n2].
// Thread (q, n2) loads its R1 points straight from HBM and transforms them
// in registers; one LDS exchange; thread (q, k1) transforms its R2 points
// in registers.  The inverse writes its columns from there; the forward
// separates the two real spectra through one more exchange.
// ---------------------------------------------------------------------------
template <int N, int R1, int TH>
struct Fft2 {
  static constexpr int R2 = N / R1;
  static constexpr int NQ = TH / R2;              // complex sequences per tile
  static constexpr int TS = 2 * NQ;               // real columns per tile
  static size_t lds () { return sizeof (double2) * ((size_t) N * NQ + N / 2); }
};

// in-register radix-2 DFT of R points (natural order in and out);
// W_R^j = W[j wstep] with W the N/2 twiddles of the full length
template <int R>
__device__ __forceinline__ void dft_regs (double2 (&v)[R], const double2* W, int wstep, bool inv)
{
#pragma unroll
  for (int i = 0; i < R; i++) {
    int j = 0;
#pragma unroll
    for (int b = 1, r = R >> 1; b < R; b <<= 1, r >>= 1)
      if (i & b) j |= r;
    if (j > i) { const double2 t = v[i]; v[i] = v[j]; v[j] = t; }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int j = 0; j < len / 2; j++) {
      double2 w = W[j * (R / len) * wstep];
      if (inv) w.y = -w.y;
#pragma unroll
      for (int i = 0; i < R; i += len) {
        const double2 u = v[i + j], t = cmul (v[i + j + len / 2], w);
        v[i + j] = make_double2 (u.x + t.x, u.y + t.y);
        v[i + j + len / 2] = make_double2 (u.x - t.x, u.y - t.y);
      }
    }
  }
}

// Fused forms (Fused): forward input in .* mul, forward output
// scale * (X - d_z fsub) (+ add0 on plane 0; d_z multiplies by i fdz k,
// planes 0 and 1 drop out); inverse input multiplied by i beta k (the
// z derivative, AuxField::gradient(2): Nyquist and k = 0 dropped), inverse
// output acc - amul .* x.  Null pointers / zero values switch a part off.
struct Fused {
  const double* mul = nullptr;
  double dz = 0.0, scale = 1.0, add0 = 0.0;
  const double* acc = nullptr;
  const double* amul = nullptr;
  // forward: the z derivative (times fdz k) of the packed spectra fsub is
  // subtracted before scaling
  const double* fsub = nullptr;
  double fdz = 0.0;
  const double* in = nullptr;      // the transform's input and output
  double* out = nullptr;
};
// up to FB_MAX transforms of one shape in one launch (blockIdx.y selects)
constexpr int FB_MAX = 3;
struct FusedBatch {
  Fused f[FB_MAX];
};

// V16: column pairs moved as 16-byte accesses (S even, rows 16-byte
// aligned).
template <int N, int R1, int TH, bool V16>
__global__ __launch_bounds__(TH) void k_fft2_z (FusedBatch bt, int S, int sign,
                                                        const double2* __restrict__ tw)
{
  const Fused fu = bt.f[blockIdx.y];
  const double* in = fu.in;
  double* out = fu.out;
  constexpr int R2 = Fft2<N, R1, TH>::R2, NQ = Fft2<N, R1, TH>::NQ, TS = Fft2<N, R1, TH>::TS;
  static_assert (NQ * R2 == TH && NQ * R1 <= TH, "tile shape");
  extern __shared__ double2 lds[];
  double2* T = lds;                   // [N][NQ] exchange, then the spectrum
  double2* W = T + N * NQ;            // exp(-2 pi i j / N), j < N/2
  const int tid = threadIdx.x, ntiles = (S + TS - 1) / TS;
  const bool inv = sign < 0;
  for (int i = tid; i < N / 2; i += TH) W[i] = tw[i];
  // step A is thread (qa, n2): columns c0 and c0 + 1 as one complex
  // sequence, points k = R2 n1 + n2.  Raw operands: forward (x, mul) column
  // pairs, inverse the (Re, Im) rows of the packed spectra.
  const int qa = tid % NQ, n2 = tid / NQ;
  double2 p0[R1], p1[R1];
  auto pair = [&] (const double* row, int c0) __attribute__ ((always_inline)) {
    if (V16) return c0 < S ? *(const double2*) (row + c0) : make_double2 (0.0, 0.0);
    return make_double2 (c0 < S ? row[c0] : 0.0, c0 + 1 < S ? row[c0 + 1] : 0.0);
  };
  auto fetch = [&] (int t) __attribute__ ((always_inline)) {
    const int c0 = t * TS + 2 * qa;
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
      const int k = R2 * n1 + n2;
      if (!inv) {
        p0[n1] = pair (in + (size_t) k * S, c0);
        if (fu.mul) p1[n1] = pair (fu.mul + (size_t) k * S, c0);
      } else {
        const int kk = k <= N / 2 ? k : N - k;
        const int pr = kk == 0 ? 0 : (kk == N / 2 ? 1 : 2 * kk);
        p0[n1] = pair (in + (size_t) pr * S, c0);
        p1[n1] = kk != 0 && kk != N / 2 ? pair (in + (size_t) (2 * kk + 1) * S, c0)
                                        : make_double2 (0.0, 0.0);
      }
    }
  };
  const int t = blockIdx.x, s0 = t * TS;
  if (t < ntiles) fetch (t);
  __syncthreads ();                   // twiddles in LDS
  if (t >= ntiles) return;
  {
    double2 v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
      const int k = R2 * n1 + n2;
      if (!inv) {
        v[n1] = p0[n1];
        if (fu.mul) { v[n1].x *= p1[n1].x; v[n1].y *= p1[n1].y; }
      } else {
        // Z_k = Xa_k + i Xb_k from the packed spectra (X_{N-k} = conj X_k)
        double ai = p1[n1].x, bi = p1[n1].y;
        if (k > N / 2) { ai = -ai; bi = -bi; }
        v[n1] = make_double2 (p0[n1].x - bi, ai + p0[n1].y);
        if (fu.dz != 0.0) {         // times i beta k_s, k_s = k - N above N/2
          const double a = fu.dz * (k < N / 2 ? k : (k == N / 2 ? 0 : k - N));
          v[n1] = make_double2 (-a * v[n1].y, a * v[n1].x);
        }
      }
    }
    dft_regs<R1> (v, W, N / R1, inv);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      const int m = n2 * k1;        // < N
      double2 w = m < N / 2 ? W[m] : make_double2 (-W[m - N / 2].x, -W[m - N / 2].y);
      if (inv) w.y = -w.y;
      T[(k1 * R2 + n2) * NQ + qa] = cmul (v[k1], w);
    }
  }
  __syncthreads ();
  // step B: thread (q, k1)
  const int q = tid % NQ, k1 = tid / NQ, c0 = s0 + 2 * q;
  const bool act = tid < NQ * R1;
  double2 y[R2];
  if (act) {
#pragma unroll
    for (int j = 0; j < R2; j++) y[j] = T[(k1 * R2 + j) * NQ + q];
    dft_regs<R2> (y, W, N / R2, inv);
  }
  if (inv) {
    if (act)
#pragma unroll
      for (int k2 = 0; k2 < R2; k2++) {
        const size_t ro = (size_t) (k1 + R1 * k2) * S;
        double a = y[k2].x, b = y[k2].y;
        if (V16) {
          if (c0 >= S) continue;
          if (fu.acc) {
            if (fu.amul) {
              const double2 m = *(const double2*) (fu.amul + ro + c0);
              a *= m.x;
              b *= m.y;
            }
            const double2 q2 = *(const double2*) (fu.acc + ro + c0);
            a = q2.x - a;
            b = q2.y - b;
          }
          *(double2*) (out + ro + c0) = make_double2 (a, b);
          continue;
        }
        if (fu.acc) {
          if (fu.amul) {
            a *= c0 < S ? fu.amul[ro + c0] : 0.0;
            b *= c0 + 1 < S ? fu.amul[ro + c0 + 1] : 0.0;
          }
          a = (c0 < S ? fu.acc[ro + c0] : 0.0) - a;
          b = (c0 + 1 < S ? fu.acc[ro + c0 + 1] : 0.0) - b;
        }
        if (c0 < S) out[ro + c0] = a;
        if (c0 + 1 < S) out[ro + c0 + 1] = b;
      }
    return;
  }
  __syncthreads ();                 // every step-B read of T is done
  if (act)
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) T[(k1 + R1 * k2) * NQ + q] = y[k2];
  __syncthreads ();
  const double scale = fu.scale / N;
  if (V16) {
    // one thread per plane and column pair: both columns from Z_k, Z_{N-k}
    for (int i = tid; i < N * NQ; i += TH) {
      const int pp = i / NQ, qq = i % NQ, s = s0 + 2 * qq;
      if (s >= S) continue;
      const int k = pp == 0 ? 0 : (pp == 1 ? N / 2 : pp >> 1);
      const double2 zk = T[k * NQ + qq], zr = T[((N - k) & (N - 1)) * NQ + qq];
      const bool re = pp < 2 || (pp & 1) == 0;
      const double ve = re ? 0.5 * (zk.x + zr.x) : 0.5 * (zk.y - zr.y);
      const double vo = re ? 0.5 * (zk.y + zr.y) : -0.5 * (zk.x - zr.x);
      double ae = pp == 0 ? fu.add0 : 0.0, ao = ae;
      if (fu.fsub && pp >= 2) {
        // Re_k -= -fdz k Im_k (fsub), Im_k -= fdz k Re_k (fsub); fsub is
        // normalised, X (ve, vo) is not yet
        const double bk = fu.scale * fu.fdz * (pp >> 1);
        const double2 t = *(const double2*) (fu.fsub + (size_t) (pp ^ 1) * S + s);
        ae = (pp & 1) ? -bk * t.x : bk * t.x;
        ao = (pp & 1) ? -bk * t.y : bk * t.y;
      }
      *(double2*) (out + (size_t) pp * S + s) = make_double2 (ve * scale + ae, vo * scale + ao);
    }
  } else {
    for (int i = tid; i < N * TS; i += TH) {
      const int pp = i / TS, c = i % TS, s = s0 + c;
      if (s >= S) continue;
      const int qq = c >> 1, k = pp == 0 ? 0 : (pp == 1 ? N / 2 : pp >> 1);
      const double2 zk = T[k * NQ + qq], zr = T[((N - k) & (N - 1)) * NQ + qq];
      // Xa = (Z_k + conj Z_{N-k}) / 2,  Xb = (Z_k - conj Z_{N-k}) / 2i
      double re, im;
      if ((c & 1) == 0) { re = 0.5 * (zk.x + zr.x); im = 0.5 * (zk.y - zr.y); }
      else              { re = 0.5 * (zk.y + zr.y); im = -0.5 * (zk.x - zr.x); }
      double a = pp == 0 ? fu.add0 : 0.0;
      if (fu.fsub && pp >= 2) {
        const double bk = fu.scale * fu.fdz * (pp >> 1), t = fu.fsub[(size_t) (pp ^ 1) * S + s];
        a = (pp & 1) ? -bk * t : bk * t;
      }
      out[(size_t) pp * S + s] = ((pp < 2 || (pp & 1) == 0) ? re : im) * scale + a;
    }
  }
}



// Packed DFT matrix product out[k][s] = sum_n M[k][n] in[n][s] (any nz);
// one thread per column, KG output planes per pass (M rows in LDS).
constexpr int KG = 16;
__global__ __launch_bounds__(256) void k_dft_z (const double* __restrict__ M,
                                                const double* __restrict__ in, double* out,
                                                int nz, int S)
{
  extern __shared__ double Ms[];   // [KG][nz]
  const int k0 = blockIdx.y * KG, kg = min (KG, nz - k0);
  for (int i = threadIdx.x; i < kg * nz; i += blockDim.x) Ms[i] = M[(size_t) k0 * nz + i];
  __syncthreads ();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double acc[KG];
#pragma unroll
  for (int k = 0; k < KG; k++) acc[k] = 0.0;
  for (int n = 0; n < nz; n++) {
    const double xv = in[(size_t) n * S + s];
#pragma unroll
    for (int k = 0; k < KG; k++)
      if (k < kg) acc[k] += Ms[k * nz + n] * xv;
  }
#pragma unroll
  for (int k = 0; k < KG; k++)
    if (k < kg) out[(size_t) (k0 + k) * S + s] = acc[k];
}

// Per (device, nz) tables: FFT twiddles, or forward/inverse DFT matrices.
struct Tables {
  double2* tw = nullptr;
  double *Cf = nullptr, *Ci = nullptr;
};
std::mutex g_mu;
std::map<std::pair<int, int>, Tables> g_tables;

bool pow2_fft (int nz) { return sx::fourier_in_place (nz); }

int tables (int nz, Tables** out)
{
  int dev = 0;
  SX_HIP (hipGetDevice (&dev));
  std::lock_guard<std::mutex> lk (g_mu);
  Tables& t = g_tables[{dev, nz}];
  const long double twopi = 2.0L * 3.14159265358979323846264338327950288L;
  if (pow2_fft (nz) && !t.tw) {
    std::vector<double2> h (nz / 2);
    for (int j = 0; j < nz / 2; j++) {
      const long double a = twopi * j / nz;
      h[j] = make_double2 ((double) cosl (a), (double) -sinl (a));
    }
    SX_HIP (hipMalloc (&t.tw, sizeof (double2) * h.size ()));
    SX_HIP (hipMemcpy (t.tw, h.data (), sizeof (double2) * h.size (), hipMemcpyHostToDevice));
  }
  if (!pow2_fft (nz) && !t.Cf) {
    std::vector<double> Cf ((size_t) nz * nz, 0.0), Ci ((size_t) nz * nz, 0.0);
    for (int n = 0; n < nz; n++) {
      Cf[n] = 1.0 / nz;
      Ci[(size_t) n * nz] = 1.0;
      if (nz > 1) {
        Cf[(size_t) nz + n] = ((n & 1) ? -1.0 : 1.0) / nz;
        Ci[(size_t) n * nz + 1] = (n & 1) ? -1.0 : 1.0;
      }
      for (int k = 1; k < nz / 2; k++) {
        const long double a = twopi * (long double) ((long) k * n % nz) / nz;
        const double c = (double) cosl (a), s = (double) sinl (a);
        Cf[(size_t) (2 * k) * nz + n] = c / nz;
        Cf[(size_t) (2 * k + 1) * nz + n] = -s / nz;
        Ci[(size_t) n * nz + 2 * k] = 2.0 * c;
        Ci[(size_t) n * nz + 2 * k + 1] = -2.0 * s;
      }
    }
    SX_HIP (hipMalloc (&t.Cf, sizeof (double) * Cf.size ()));
    SX_HIP (hipMalloc (&t.Ci, sizeof (double) * Ci.size ()));
    SX_HIP (hipMemcpy (t.Cf, Cf.data (), sizeof (double) * Cf.size (), hipMemcpyHostToDevice));
    SX_HIP (hipMemcpy (t.Ci, Ci.data (), sizeof (double) * Ci.size (), hipMemcpyHostToDevice));
  }
  *out = &t;
  return SX_OK;
}

template <int N, int R1, int TH>
int launch_fft2_t (int S, int nb, const FusedBatch& bt, int sign, const double2* tw,
                   hipStream_t st)
{
  const size_t lds = Fft2<N, R1, TH>::lds ();
  const int TS = Fft2<N, R1, TH>::TS, ntiles = (S + TS - 1) / TS;
  // column pairs as 16-byte accesses when every row starts 16-byte aligned
  auto al = [] (const double* p) { return ((uintptr_t) p & 15) == 0; };
  bool v16 = S % 2 == 0;
  for (int b = 0; b < nb; b++) {
    const Fused& fu = bt.f[b];
    v16 = v16 && al (fu.in) && al (fu.out) && al (fu.mul) && al (fu.acc) && al (fu.amul) &&
          al (fu.fsub);
  }
  auto kern = v16 ? k_fft2_z<N, R1, TH, true> : k_fft2_z<N, R1, TH, false>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  hipLaunchKernelGGL (kern, dim3 (ntiles, nb), dim3 (TH), lds, st, bt, S, sign, tw);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// Threads per transform tile: 128 (16 real columns of 128-byte rows per
// tile, more and smaller workgroups than 256: 9.9 against 10.2 us for 128
// planes of 18,432 columns); SX_FFT_THREADS=256 selects the larger tile.
int fft_threads ()
{
  const char* e = getenv ("SX_FFT_THREADS");
  return e && atoi (e) == 256 ? 256 : 128;
}

template <int N, int R1>
int launch_fft2b (int S, int nb, const FusedBatch& bt, int sign, const double2* tw,
                  hipStream_t st)
{
  if (fft_threads () == 128) return launch_fft2_t<N, R1, 128> (S, nb, bt, sign, tw, st);
  return launch_fft2_t<N, R1, 256> (S, nb, bt, sign, tw, st);
}

template <int N, int R1>
int launch_fft2 (int S, const double* in, double* out, int sign, const double2* tw,
                 hipStream_t st, Fused fu = Fused ())
{
  FusedBatch bt;
  fu.in = in;
  fu.out = out;
  bt.f[0] = fu;
  return launch_fft2b<N, R1> (S, 1, bt, sign, tw, st);
}

template <int N>
int launch_fft (int S, const double* in, double* out, int sign, const double2* tw, hipStream_t st)
{
  const size_t lds = FftTile<N>::lds ();
  auto kern = k_fft_z<N>;
  SX_HIP (hipFuncSetAttribute ((const void*) kern,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
  const int TS = FftTile<N>::TS;
  hipLaunchKernelGGL (kern, dim3 ((S + TS - 1) / TS), dim3 (FT_THREADS), lds, st, in, out, S,
                      sign, tw);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace

namespace sx {

int fourier_prepare (int nz)
{
  Tables* t = nullptr;
  return nz > 1 ? tables (nz, &t) : SX_OK;
}

bool fourier_fused_ok (int nz) { return nz == 32 || nz == 64 || nz == 128 || nz == 256; }

int fourier_z_fused (int nz, int S, const double* in, const double* mul, double* out, int sign,
                     double dz, const double* acc, const double* amul, double scale, double add0,
                     hipStream_t st)
{
  if (!fourier_fused_ok (nz) || S < 1 || (sign != 1 && sign != -1)) {
    set_error ("fourier_z_fused: nz must be 32, 64, 128 or 256");
    return SX_ERR_ARG;
  }
  Tables* t = nullptr;
  if (int rc = tables (nz, &t)) return rc;
  Fused fu;
  fu.mul = sign > 0 ? mul : nullptr;
  fu.scale = sign > 0 ? scale : 1.0;
  fu.add0 = sign > 0 ? add0 : 0.0;
  fu.dz = sign < 0 ? dz : 0.0;
  fu.acc = sign < 0 ? acc : nullptr;
  fu.fsub = sign > 0 ? acc : nullptr;
  fu.fdz = sign > 0 ? dz : 0.0;
  fu.amul = sign < 0 ? amul : nullptr;
  switch (nz) {
  case 32: return launch_fft2<32, 4> (S, in, out, sign, t->tw, st, fu);
  case 64: return launch_fft2<64, 8> (S, in, out, sign, t->tw, st, fu);
  case 128: return launch_fft2<128, 8> (S, in, out, sign, t->tw, st, fu);
  default: return launch_fft2<256, 16> (S, in, out, sign, t->tw, st, fu);
  }
}

int fourier_z_fused_batch (int nz, int S, int nb, const double* const* in,
                           const double* const* mul, double* const* out, int sign, double dz,
                           const double* const* acc, const double* const* amul,
                           const double* scale, const double* add0, hipStream_t st)
{
  if (!fourier_fused_ok (nz) || S < 1 || (sign != 1 && sign != -1) || nb < 1 || nb > FB_MAX) {
    set_error ("fourier_z_fused_batch: nz must be 32, 64, 128 or 256, 1..3 transforms");
    return SX_ERR_ARG;
  }
  Tables* t = nullptr;
  if (int rc = tables (nz, &t)) return rc;
  FusedBatch bt;
  for (int b = 0; b < nb; b++) {
    Fused& fu = bt.f[b];
    fu.in = in[b];
    fu.out = out[b];
    fu.mul = sign > 0 && mul ? mul[b] : nullptr;
    fu.scale = sign > 0 && scale ? scale[b] : 1.0;
    fu.add0 = sign > 0 && add0 ? add0[b] : 0.0;
    fu.dz = sign < 0 ? dz : 0.0;
    fu.acc = sign < 0 && acc ? acc[b] : nullptr;
    fu.fsub = sign > 0 && acc ? acc[b] : nullptr;
    fu.fdz = sign > 0 ? dz : 0.0;
    fu.amul = sign < 0 && amul ? amul[b] : nullptr;
  }
  switch (nz) {
  case 32: return launch_fft2b<32, 4> (S, nb, bt, sign, t->tw, st);
  case 64: return launch_fft2b<64, 8> (S, nb, bt, sign, t->tw, st);
  case 128: return launch_fft2b<128, 8> (S, nb, bt, sign, t->tw, st);
  default: return launch_fft2b<256, 16> (S, nb, bt, sign, t->tw, st);
  }
}

int fourier_z (int nz, int S, const double* in, double* out, int sign, hipStream_t st)
{
  if (nz < 2 || S < 1) {
    if (in != out && nz >= 1 && S >= 1)
      SX_HIP (hipMemcpyAsync (out, in, sizeof (double) * nz * S, hipMemcpyDeviceToDevice, st));
    return SX_OK;
  }
  Tables* t = nullptr;
  int rc = tables (nz, &t);
  if (rc) return rc;
  if (pow2_fft (nz)) {
    switch (nz) {
    case 8: return launch_fft<8> (S, in, out, sign, t->tw, st);
    case 16: return launch_fft<16> (S, in, out, sign, t->tw, st);
    case 32: return launch_fft2<32, 4> (S, in, out, sign, t->tw, st);
    case 64: return launch_fft2<64, 8> (S, in, out, sign, t->tw, st);
    case 128: return launch_fft2<128, 8> (S, in, out, sign, t->tw, st);
    case 256: return launch_fft2<256, 16> (S, in, out, sign, t->tw, st);
    case 512: return launch_fft<512> (S, in, out, sign, t->tw, st);
    case 1024: return launch_fft<1024> (S, in, out, sign, t->tw, st);
    default: break;
    }
  }
  if (in == out) {
    set_error ("fourier transform: the matrix form (nz not a power of two in 8..1024) "
               "needs distinct input and output");
    return SX_ERR_ARG;
  }
  dim3 grid ((S + 255) / 256, (nz + KG - 1) / KG);
  hipLaunchKernelGGL (k_dft_z, grid, dim3 (256), sizeof (double) * KG * nz, st,
                      sign > 0 ? t->Cf : t->Ci, in, out, nz, S);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

}  // namespace sx

extern "C" int sx_fourier_transform (int nz, int stride, const double* d_in, double* d_out,
                                     int sign, void* stream)
{
  if (nz < 1 || stride < 1 || !d_in || !d_out || (sign != 1 && sign != -1) ||
      (nz > 1 && (nz & 1))) {
    sx::set_error ("sx_fourier_transform: bad arguments (nz must be 1 or even, sign +-1)");
    return SX_ERR_ARG;
  }
  return sx::fourier_z (nz, stride, d_in, d_out, sign, (hipStream_t) stream);
}
