// sx_dns.hip -- the DNS time step of semtex's dns driver on gfx950:
// integrate() (dns/integrate.cpp:66-201) with skewSymmetric or
// altSkewSymmetric advection (dns/nonlinear.cpp:38-381, Cartesian), the
// computed pressure boundary conditions (BCmgr::maintainFourier /
// evaluateCNBCp, src/bcmgr.cpp:1037-1335, Edge::curlCurl src/edge.cpp:98),
// Field::smooth (src/field.cpp:282-327) and ConstForce
// (dns/fieldforce.cpp:268).  All fields stay on the GPU; the pressure and
// the grouped velocity solves go through the library's Field::solve, the
// z transforms through sx_fourier.hip (with the z terms of the advection
// folded into them where nz allows).
//
// Kernels:
//   k_adv_xy     x and y terms of the nonlinear term of one component per
//                element-plane (products, Element::grad contractions)
//   k_wave_div   waveProp and setPForce: H of every component and div H / dt
//                per element and plane pair
//   k_project3   project: F_i = H_i alpha + beta d_i p, all components
//   k_grad       x / y derivative per element (Element::grad)
//   k_gradz      Fourier z derivative (AuxField::gradient(2))
//   k_ew_*       element-wise field algebra (the unfused z terms)
//   k_smooth_*   mass-averaged direct stiffness sum on element boundaries
//   k_hopbc      -d(u.n)/dt + n.N - nu n.curlcurl u on boundary segments
//   k_plines     extrapolated dp/dn line data for the pressure solve
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocblas/rocblas.h>

#include "sx_internal.h"

using namespace sx;

namespace {

constexpr double EPSDP = 6.0e-14;

// x or y derivative of every element of every plane (Element::grad,
// src/element.cpp:521-565): d/dx = (u DT) drdx + (D u) dsdx.
__global__ __launch_bounds__(64) void k_grad (int np, const double* __restrict__ geo,
                                              const double* __restrict__ Dg,
                                              const double* __restrict__ in,
                                              double* out, int dir, int stride)
{
  const int npnp = np * np, e = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
  __shared__ double D[256], u[256];
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  const double* ge = geo + (size_t) e * NGEO * npnp;
  const double* gr = ge + (dir == 0 ? G_RX : G_RY) * npnp;
  const double* gs = ge + (dir == 0 ? G_SX : G_SY) * npnp;
  // every load (operator, field, metric factors) in one round trip;
  // np <= 16, so a lane holds at most 4 nodes
  double r_[4], s_[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int i = lane + 64 * t;
    if (i < npnp) { D[i] = Dg[i]; u[i] = in[off + i]; r_[t] = gr[i]; s_[t] = gs[i]; }
  }
  __syncthreads ();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij >= npnp) break;
    const int i = ij / np, j = ij % np;
    double ur = 0.0, us = 0.0;
    for (int q = 0; q < np; q++) {
      ur += u[i * np + q] * D[j * np + q];
      us += D[i * np + q] * u[q * np + j];
    }
    out[off + ij] = ur * r_[t] + us * s_[t];
  }
}

// The x and y terms of the nonlinear term of component u, per element-plane
// in physical space (skewSymmetric / altSkewSymmetric, dns/nonlinear.cpp,
// Cartesian branches :162-193, :342-380, with Element::grad for every
// derivative):
//   N = -[conv (ux d_x u + uy d_y u) + cons (d_x (u ux) + d_y (u uy))]
// written, not accumulated: the z terms are subtracted afterwards.
// per-component arguments of a launch over up to three components
// (blockIdx.z selects one)
struct Comp3 {
  const double* in[3];
  double* out[3];
};
__global__ __launch_bounds__(64) void k_adv_xy (int np, const double* __restrict__ geo,
                                                const double* __restrict__ Dg,
                                                const double* __restrict__ ux,
                                                const double* __restrict__ uy, Comp3 cmp,
                                                int stride, int conv, int cons)
{
  const double* __restrict__ ui = cmp.in[blockIdx.z];
  double* N = cmp.out[blockIdx.z];
  const int npnp = np * np, e = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
  __shared__ double D[256], fu[256], fx[256], fy[256];     // u, u ux, u uy
  const size_t off = (size_t) k * stride + (size_t) e * npnp;
  const double* ge = geo + (size_t) e * NGEO * npnp;
  double ax[4], ay[4], rx[4], ry[4], sx[4], sy[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {       // np <= 16: at most 4 nodes per lane
    const int i = lane + 64 * t;
    if (i < npnp) {
      D[i] = Dg[i];
      const double a = ui[off + i];
      ax[t] = ux[off + i];
      ay[t] = uy[off + i];
      fu[i] = a;
      fx[i] = a * ax[t];
      fy[i] = a * ay[t];
      rx[t] = ge[G_RX * npnp + i];
      ry[t] = ge[G_RY * npnp + i];
      sx[t] = ge[G_SX * npnp + i];
      sy[t] = ge[G_SY * npnp + i];
    }
  }
  __syncthreads ();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij >= npnp) break;
    const int i = ij / np, j = ij % np;
    double ur = 0.0, us = 0.0, xr = 0.0, xs = 0.0, yr = 0.0, ys = 0.0;
    for (int q = 0; q < np; q++) {
      const double dj = D[j * np + q], di = D[i * np + q];
      ur += fu[i * np + q] * dj;
      us += di * fu[q * np + j];
      xr += fx[i * np + q] * dj;
      xs += di * fx[q * np + j];
      yr += fy[i * np + q] * dj;
      ys += di * fy[q * np + j];
    }
    double v = 0.0;
    if (conv) v += ax[t] * (ur * rx[t] + us * sx[t]) + ay[t] * (ur * ry[t] + us * sy[t]);
    if (cons) v += (xr * rx[t] + xs * sx[t]) + (yr * ry[t] + ys * sy[t]);
    N[off + ij] = -v;
  }
}

// the z derivative of a Fourier field at one node of plane k (k's partner
// plane holds the other half of the mode): planes 0, 1 -> 0,
// Re_k <- -beta m Im_k, Im_k <- beta m Re_k, m = k/2
__device__ __forceinline__ double dz_at (const double* f, int k, int stride, size_t node,
                                         double beta)
{
  const int m = k >> 1;
  if (m < 1) return 0.0;
  const double bk = beta * m;
  return (k & 1) ? bk * f[(size_t) (k - 1) * stride + node]
                 : -bk * f[(size_t) (k + 1) * stride + node];
}


// project (dns/integrate.cpp:260-289) for every component at once:
// F_i = H_i alpha + beta d_i p, per element-plane (x, y by Element::grad,
// z as AuxField::gradient(2)); nc = 2 or 3
__global__ __launch_bounds__(64) void k_project3 (int np, int nc,
                                                  const double* __restrict__ geo,
                                                  const double* __restrict__ Dg,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ H0,
                                                  const double* __restrict__ H1,
                                                  const double* __restrict__ H2, double* F0,
                                                  double* F1, double* F2, int stride,
                                                  double beta, double alpha, double betap)
{
  const int npnp = np * np, e = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
  __shared__ double D[256], P[256];
  const size_t eo = (size_t) e * npnp, off = (size_t) k * stride + eo;
  const double* ge = geo + (size_t) e * NGEO * npnp;
  double rx[4], sx[4], ry[4], sy[4], gz[4], h0[4], h1[4], h2[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int i = lane + 64 * t;
    if (i < npnp) {
      D[i] = Dg[i];
      P[i] = p[off + i];
      rx[t] = ge[G_RX * npnp + i];
      sx[t] = ge[G_SX * npnp + i];
      ry[t] = ge[G_RY * npnp + i];
      sy[t] = ge[G_SY * npnp + i];
      h0[t] = H0[off + i];
      h1[t] = H1[off + i];
      h2[t] = nc == 3 ? H2[off + i] : 0.0;
      gz[t] = nc == 3 ? dz_at (p, k, stride, eo + i, beta) : 0.0;
    }
  }
  __syncthreads ();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij >= npnp) break;
    const int i = ij / np, j = ij % np;
    double pr = 0.0, ps = 0.0;
    for (int q = 0; q < np; q++) {
      pr += P[i * np + q] * D[j * np + q];
      ps += D[i * np + q] * P[q * np + j];
    }
    const double gx = pr * rx[t] + ps * sx[t], gy = pr * ry[t] + ps * sy[t];
    double f = h0[t] * alpha;
    f += betap * gx;
    F0[off + ij] = f;
    f = h1[t] * alpha;
    f += betap * gy;
    F1[off + ij] = f;
    if (nc == 3) {
      f = h2[t] * alpha;
      f += betap * gz[t];
      F2[off + ij] = f;
    }
  }
}

// AuxField::gradient(2) on Fourier planes: planes 0 and 1 zero,
// Re_k <- -beta k Im_k, Im_k <- beta k Re_k.
__global__ void k_gradz (const double* __restrict__ in, double* out, int nz,
                         int S, double beta)
{
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t) nz * S) return;
  const int k = (int) (idx / S);
  const size_t s = idx % S;
  const int m = k >> 1;
  double v = 0.0;
  if (m >= 1) {
    const double bk = beta * m;
    v = (k & 1) ? bk * in[(size_t) (k - 1) * S + s] : -bk * in[(size_t) (k + 1) * S + s];
  }
  out[idx] = v;
}

// Element-wise kernels over n = nz*S entries.
__global__ void k_ew_submul (double* N, const double* a, const double* b, size_t n)
{   // N -= a*b  (AuxField::timesMinus)
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) N[i] -= a[i] * b[i];
}
__global__ void k_ew_mul (double* out, const double* a, const double* b, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i];
}
__global__ void k_ew_sub (double* N, const double* t, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) N[i] -= t[i];
}
__global__ void k_ew_scale (double* x, double a, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}
__global__ void k_ew_addc (double* x, double c, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += c;
}
// N -= d_z F in Fourier space (planes 0 and 1 carry no z derivative):
// Re_k -= -beta k Im_k(F), Im_k -= beta k Re_k(F)
__global__ void k_sub_dz (double* N, const double* F, int nz, int S, double beta)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t half = (size_t) (nz / 2 - 1) * S;
  if (i >= half) return;
  const int k = 1 + (int) (i / S);
  const size_t c = i % S, re = (size_t) 2 * k * S + c, im = re + S;
  const double bk = beta * k, fr = F[re], fi = F[im];
  N[re] += bk * fi;
  N[im] -= bk * fr;
}

// Time levels q = 0..Je-1 of one field (Je <= 3) and their coefficients.
struct Levels {
  const double* f[3];
  double c[3];
};
struct Coef3 { double c[3]; };
// waveProp (dns/integrate.cpp:204-239): H = sum_q -alpha[q+1] Us[q] + beta[q] dt Uf[q],
// summed in the reference's order (q ascending, the Us term first)
struct Wave3 {
  double* H[3];
  Levels us[3], uf[3];
  double* unyq[3];
};
// waveProp and setPForce in one pass (dns/integrate.cpp:204-257): one
// workgroup per element and plane pair (2y, 2y+1), a wave per plane
// (threadIdx.y; np <= 16, so a lane holds at most 4 nodes).  Each node's
// H_c = sum_q us.c[q] Us[q] + uf.c[q] Uf[q] (summed in the reference's
// order) is formed and written; the divergence out = (d_x H0 + d_y H1 +
// d_z H2) / dt (x, y by Element::grad, z as AuxField::gradient(2)) is then
// formed from the values in LDS, the z derivative from the pair's other
// plane.  The unsolved Nyquist plane 1 of Us[0] is carried into unyq, and
// the element-plane padding of H (S > nel np^2) is written too.
__global__ __launch_bounds__(128) void k_wave_div (Wave3 wv, int Je, int nc, int nz, int np,
                                                   int nel, const double* __restrict__ geo,
                                                   const double* __restrict__ Dg, double* out,
                                                   int S, double beta, double dt)
{
  const int npnp = np * np, e = blockIdx.x, pl = threadIdx.y, lane = threadIdx.x;
  const int k = 2 * blockIdx.y + pl;
  const bool kin = k < nz;
  __shared__ double D[256], hs[3][2][256];           // H0, H1, H2 of both planes
  const size_t eo = (size_t) e * npnp, off = (size_t) k * S + eo;
  const double* ge = geo + (size_t) e * NGEO * npnp;
  auto wave_at = [&] (int c, size_t i) __attribute__ ((always_inline)) {
    const Levels& us = wv.us[c];
    const Levels& uf = wv.uf[c];
    double h = 0.0;
    for (int q = 0; q < Je; q++) {
      h += us.c[q] * us.f[q][i];
      h += uf.c[q] * uf.f[q][i];
    }
    return h;
  };
  double rx[4], sx[4], ry[4], sy[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int i = lane + 64 * t;
    if (i < npnp) {
      if (pl == 0) D[i] = Dg[i];
      rx[t] = ge[G_RX * npnp + i];
      sx[t] = ge[G_SX * npnp + i];
      ry[t] = ge[G_RY * npnp + i];
      sy[t] = ge[G_SY * npnp + i];
      for (int c = 0; c < nc; c++) {
        double h = 0.0;
        if (kin) {
          h = wave_at (c, off + i);
          wv.H[c][off + i] = h;
          if (wv.unyq[c] && k == 1) wv.unyq[c][off + i] = wv.us[c].f[0][off + i];
        }
        hs[c][pl][i] = h;
      }
    }
  }
  if (kin && e == 0 && lane == 0)
    for (size_t i = (size_t) k * S + (size_t) nel * npnp; i < (size_t) (k + 1) * S; i++)
      for (int c = 0; c < nc; c++) {
        wv.H[c][i] = wave_at (c, i);
        if (wv.unyq[c] && k == 1) wv.unyq[c][i] = wv.us[c].f[0][i];
      }
  __syncthreads ();
  if (!kin) return;
  const double* a = hs[0][pl];
  const double* b = hs[1][pl];
  const int m = k >> 1;
  const double bk = beta * m;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ij = lane + 64 * t;
    if (ij >= npnp) break;
    const int i = ij / np, j = ij % np;
    double ar = 0.0, as = 0.0, br = 0.0, bs = 0.0;
    for (int q = 0; q < np; q++) {
      const double dj = D[j * np + q], di = D[i * np + q];
      ar += a[i * np + q] * dj;
      as += di * a[q * np + j];
      br += b[i * np + q] * dj;
      bs += di * b[q * np + j];
    }
    double sum = (ar * rx[t] + as * sx[t]) + (br * ry[t] + bs * sy[t]);
    if (nc == 3)
      sum += m < 1 ? 0.0 : ((k & 1) ? bk * hs[2][0][ij] : -bk * hs[2][1][ij]);
    out[off + ij] = sum / dt;
  }
}

// setPForce (dns/integrate.cpp:242-257): f = (g0 + g1 [+ g2]) / dt
// Field::smooth: the average imass[g] sum Q4 f over the element-boundary
// slots of global node g, written back to every one of those slots.  The
// slot lists of the global nodes are disjoint, so each thread reads and
// writes only its own slots (one pass, no intermediate global array).
__global__ void k_smooth (int nglobal, const int* __restrict__ gptr,
                          const int* __restrict__ gpos, const double* __restrict__ gq4,
                          const double* __restrict__ imass, Comp3 cmp, int S)
{
  double* f = cmp.out[blockIdx.z];
  const int g = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (g >= nglobal) return;
  double* fk = f + (size_t) k * S;
  const int c0 = gptr[g], c1 = gptr[g + 1];
  // the first four slots' offsets and weights, then their values, each in
  // one round (a node shared by more elements takes the rest one by one);
  // summed in slot order
  int p[4];
  double w[4], v[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    p[u] = c0 + u < c1 ? gpos[c0 + u] : 0;
    w[u] = c0 + u < c1 ? gq4[c0 + u] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 4; u++) v[u] = c0 + u < c1 ? fk[p[u]] : 0.0;
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (c0 + u < c1) s += w[u] * v[u];
  for (int c = c0 + 4; c < c1; c++) s += gq4[c] * fk[gpos[c]];
  s *= imass[g];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (c0 + u < c1) fk[p[u]] = s;
  for (int c = c0 + 4; c < c1; c++) fk[gpos[c]] = s;
}

// Computed pressure BC terms on boundary segment b, Fourier mode m
// (BCmgr::maintainFourier, src/bcmgr.cpp:1037-1296; Edge::curlCurl,
// src/edge.cpp:98-220), Cartesian.  hop_new is level 0 of the rolled
// storage, un_new receives this step's u.n and un1 holds the previous
// step's.  The time-derivative term is formed as the reference forms it
// (src/bcmgr.cpp:1124-1130): alpha_0 u.n + sum_{q<Je} alpha_{q+1} times
// the *previous* level for every q, i.e. a0 u.n + asum un1 with asum =
// alpha_1 + .. + alpha_Je.
__global__ __launch_bounds__(64) void k_hopbc (
    int np, const int* __restrict__ seg_elmt, const int* __restrict__ seg_idx,
    const double* __restrict__ seg_nx, const double* __restrict__ seg_ny,
    const double* __restrict__ geo, const double* __restrict__ Dg,
    const double* __restrict__ U0, const double* __restrict__ U1,
    const double* __restrict__ U2, const double* __restrict__ N0,
    const double* __restrict__ N1, double* hop_new, double* un_new,
    const double* __restrict__ un1, int Je, double a0, Coef3 acoef, double inv_dt, double nu,
    double beta, int nz, int nline, int S)
{
  const int npnp = np * np, b = blockIdx.x, m = blockIdx.y, lane = threadIdx.x;
  const int e = seg_elmt[b];
  __shared__ double D[256], ur[256], ui[256], vr[256], vi[256], wr[256], wi[256];
  __shared__ double gwr[2][16], gwi[2][16], wzr[256], wzi[256];
  __shared__ int idx[16];
  __shared__ double nx[16], ny[16];
  for (int i = lane; i < npnp; i += 64) D[i] = Dg[i];
  if (lane < np) {
    idx[lane] = seg_idx[b * np + lane];
    nx[lane] = seg_nx[b * np + lane];
    ny[lane] = seg_ny[b * np + lane];
  }
  const int kr = (m == 0) ? 0 : 2 * m, ki = kr + 1;
  const bool three = (m > 0);
  const size_t eoff = (size_t) e * npnp;
  for (int i = lane; i < npnp; i += 64) {
    ur[i] = U0[(size_t) kr * S + eoff + i];
    vr[i] = U1[(size_t) kr * S + eoff + i];
    if (three) {
      ui[i] = U0[(size_t) ki * S + eoff + i];
      vi[i] = U1[(size_t) ki * S + eoff + i];
      wr[i] = U2[(size_t) kr * S + eoff + i];
      wi[i] = U2[(size_t) ki * S + eoff + i];
    }
  }
  __syncthreads ();
  const double* ge = geo + (size_t) e * NGEO * npnp;
  const double *rx = ge + G_RX * npnp, *ry = ge + G_RY * npnp,
               *sx = ge + G_SX * npnp, *sy = ge + G_SY * npnp;
  // vorticity w_z = dv/dx - du/dy over the element (Re, and Im for m > 0)
  for (int ij = lane; ij < npnp; ij += 64) {
    const int i = ij / np, j = ij % np;
    double vr_r = 0, vr_s = 0, ur_r = 0, ur_s = 0, vi_r = 0, vi_s = 0, ui_r = 0, ui_s = 0;
    for (int q = 0; q < np; q++) {
      const double dj = D[j * np + q], di = D[i * np + q];
      vr_r += vr[i * np + q] * dj; vr_s += di * vr[q * np + j];
      ur_r += ur[i * np + q] * dj; ur_s += di * ur[q * np + j];
      if (three) {
        vi_r += vi[i * np + q] * dj; vi_s += di * vi[q * np + j];
        ui_r += ui[i * np + q] * dj; ui_s += di * ui[q * np + j];
      }
    }
    wzr[ij] = (vr_r * rx[ij] + vr_s * sx[ij]) - (ur_r * ry[ij] + ur_s * sy[ij]);
    if (three) wzi[ij] = (vi_r * rx[ij] + vi_s * sx[ij]) - (ui_r * ry[ij] + ui_s * sy[ij]);
  }
  __syncthreads ();
  // gradients of w_z at the side nodes
  if (lane < np) {
    const int ij = idx[lane], i = ij / np, j = ij % np;
    double a_r = 0, a_s = 0, b_r = 0, b_s = 0;
    for (int q = 0; q < np; q++) {
      const double dj = D[j * np + q], di = D[i * np + q];
      a_r += wzr[i * np + q] * dj; a_s += di * wzr[q * np + j];
      if (three) { b_r += wzi[i * np + q] * dj; b_s += di * wzi[q * np + j]; }
    }
    gwr[0][lane] = a_r * rx[ij] + a_s * sx[ij];     // d(w_r)/dx
    gwr[1][lane] = a_r * ry[ij] + a_s * sy[ij];     // d(w_r)/dy
    if (three) {
      gwi[0][lane] = b_r * rx[ij] + b_s * sx[ij];
      gwi[1][lane] = b_r * ry[ij] + b_s * sy[ij];
    }
  }
  __syncthreads ();
  if (lane < np) {
    const int ij = idx[lane], i = ij / np, j = ij % np, q0 = lane;
    const int jl = b * np + q0;
    // -- plane(s) of this mode: u.n, time derivative, n.N
    for (int pl = 0; pl < (three ? 2 : 1); pl++) {
      const int k = pl ? ki : kr;
      const double un = U0[(size_t) k * S + eoff + ij] * nx[q0] + U1[(size_t) k * S + eoff + ij] * ny[q0];
      double h = 0.0;
      if (Je) {
        double w = a0 * un;
        for (int q = 0; q < Je; q++) w += acoef.c[q] * un1[(size_t) k * nline + jl];
        h = -inv_dt * w;
      }
      double nn = 0.0;
      nn += N0[(size_t) k * S + eoff + ij] * nx[q0];
      nn += N1[(size_t) k * S + eoff + ij] * ny[q0];
      h += nn;
      un_new[(size_t) k * nline + jl] = un;
      hop_new[(size_t) k * nline + jl] = h;
    }
    // -- -nu n.curl curl u
    double xr = gwr[1][q0], yr = -gwr[0][q0];
    if (!three) {
      const size_t o = (size_t) kr * nline + jl;
      hop_new[o] += -nu * xr * nx[q0];
      hop_new[o] += -nu * yr * ny[q0];
    } else {
      const double bk = beta * m, bk2 = bk * bk;
      double xi = gwi[1][q0], yi = -gwi[0][q0];
      // semi-Fourier terms: grads of Wr, Wi at the node
      double wr_x = 0, wr_y = 0, wi_x = 0, wi_y = 0;
      {
        double c_r = 0, c_s = 0, d_r = 0, d_s = 0;
        for (int q = 0; q < np; q++) {
          const double dj = D[j * np + q], di = D[i * np + q];
          c_r += wr[i * np + q] * dj; c_s += di * wr[q * np + j];
          d_r += wi[i * np + q] * dj; d_s += di * wi[q * np + j];
        }
        wr_x = c_r * rx[ij] + c_s * sx[ij]; wr_y = c_r * ry[ij] + c_s * sy[ij];
        wi_x = d_r * rx[ij] + d_s * sx[ij]; wi_y = d_r * ry[ij] + d_s * sy[ij];
      }
      xi += bk * wr_x;
      yi += bk * wr_y;
      xr += -bk * wi_x;
      yr += -bk * wi_y;
      xr += bk2 * ur[ij];
      xi += bk2 * ui[ij];
      yr += bk2 * vr[ij];
      yi += bk2 * vi[ij];
      const size_t orr = (size_t) kr * nline + jl, oi = (size_t) ki * nline + jl;
      hop_new[orr] += -nu * xr * nx[q0];
      hop_new[oi] += -nu * xi * nx[q0];
      hop_new[orr] += -nu * yr * ny[q0];
      hop_new[oi] += -nu * yi * ny[q0];
    }
  }
}

// evaluateCNBCp: lines[k] = sum_q beta_q hop[q][k]; Nyquist plane and
// essential pressure segments zero.
__global__ void k_plines (double* lines, Levels hop, int Je, int nz, int nline, int np,
                          const int* __restrict__ kind)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t) nz * nline) return;
  const int k = (int) (i / nline), j = (int) (i % nline);
  double v = 0.0;
  if (!(nz > 1 && k == 1) && kind[j / np] == SX_NATURAL)
    for (int q = 0; q < Je; q++) v += hop.c[q] * hop.f[q][i];
  lines[i] = v;
}

inline unsigned nb256 (size_t n) { return (unsigned) ((n + 255) / 256); }

}  // namespace

// ===========================================================================
struct sx_dns {
  sx_mesh *mu = nullptr, *mp = nullptr;
  int nz = 1, ncom = 2, S = 0, advection = SX_SKEW, nord = 2, toggle = 1, step = 0;
  double beta = 0, nu = 0, dt = 0, cf[3] = {0, 0, 0};
  size_t n = 0;                        // nz * S
  sx_msys *Mv[3] = {nullptr, nullptr, nullptr}, *Mp = nullptr;   // Mv[Je-1]: order Je
  bool group_solve = false;            // velocity solves share one band launch
  int startup = SX_FP64;               // start-up velocity solves: direct or SX_JACPCG
  double tol_rel = 1e-8;
  int step_max = 500;
  sx_msys* Mpcg[2] = {nullptr, nullptr};   // start-up JACPCG systems, order 1, 2
  std::vector<double*> bufs;           // every device allocation
  double *u[3] = {}, *p = nullptr;
  double *Us[3][3] = {}, *Uf[3][3] = {};   // [level][component], nord levels
  double *phys[3] = {}, *tmp = nullptr, *tmp2 = nullptr, *H[3] = {}, *F[3] = {};
  double* imass = nullptr;
  int nb = 0, nline = 0, hop_head = 0, un_head = 0;
  int *seg_elmt = nullptr, *seg_idx = nullptr, *seg_kind = nullptr;
  double *seg_nx = nullptr, *seg_ny = nullptr;
  double *hop[3] = {}, *un[2] = {}, *plines = nullptr;
  double* ulines[3] = {};              // velocity boundary values, Fourier [nz][nline]
  // 3/2-rule dealiasing: M = 3 nz / 2 physical planes, the transforms as
  // FP64 GEMMs with the packed DFT matrices (column-major)
  int dealias = 0, M = 0;
  size_t nM = 0;                       // M * S
  double *physM[3] = {}, *tmpM = nullptr, *tmpM2 = nullptr;
  double *Wi = nullptr, *Wdz = nullptr, *Wf = nullptr;   // [N][M], [N][M], [M][N]
  rocblas_handle blas = nullptr;

  double* alloc (size_t cnt)
  {
    double* q = nullptr;
    if (hipMalloc (&q, sizeof (double) * std::max<size_t> (cnt, 1)) != hipSuccess) return nullptr;
    bufs.push_back (q);
    return q;
  }
};

namespace {

// z transform of a whole field, in place (tmp2 is the scratch when the
// transform cannot run in place)
int transform (sx_dns* d, double* f, int sign, hipStream_t st)
{
  if (d->nz < 2) return SX_OK;
  if (fourier_in_place (d->nz)) return fourier_z (d->nz, d->S, f, f, sign, st);
  int rc = fourier_z (d->nz, d->S, f, d->tmp2, sign, st);
  if (rc) return rc;
  SX_HIP (hipMemcpyAsync (f, d->tmp2, sizeof (double) * d->n, hipMemcpyDeviceToDevice, st));
  return SX_OK;
}

// out-of-place z transform (out may not alias in)
int transform_to (sx_dns* d, const double* in, double* out, int sign, hipStream_t st)
{
  if (d->nz < 2) {
    SX_HIP (hipMemcpyAsync (out, in, sizeof (double) * d->n, hipMemcpyDeviceToDevice, st));
    return SX_OK;
  }
  return fourier_z (d->nz, d->S, in, out, sign, st);
}

int grad (sx_dns* d, const double* in, double* out, int dir, hipStream_t st)
{
  const sx_mesh* m = d->mu;
  if (dir == 2) {
    hipLaunchKernelGGL (k_gradz, dim3 (nb256 (d->n)), dim3 (256), 0, st, in, out, d->nz,
                        d->S, d->beta);
  } else {
    hipLaunchKernelGGL (k_grad, dim3 (m->nel, d->nz), dim3 (64), 0, st, m->np, m->d_geo,
                        m->d_D, in, out, dir, d->S);
  }
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// smoothing of nf <= 3 fields in one launch
int smooth_n (sx_dns* d, double* const* f, int nf, hipStream_t st)
{
  const sx_mesh* m = d->mu;
  Comp3 c {};
  for (int i = 0; i < nf; i++) c.out[i] = f[i];
  hipLaunchKernelGGL (k_smooth, dim3 ((m->nglobal + 255) / 256, d->nz, nf), dim3 (256), 0, st,
                      m->nglobal, m->d_gptr, m->d_gpos, m->d_gq4, d->imass, c, d->S);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}
int smooth (sx_dns* d, double* f, hipStream_t st) { return smooth_n (d, &f, 1, st); }

// the x/y advection terms of nf <= 3 components in one launch
void adv_xy (sx_dns* d, double* const* ui, double* const* N, int nf, bool conv, bool cons,
             hipStream_t st)
{
  const sx_mesh* m = d->mu;
  Comp3 c {};
  for (int i = 0; i < nf; i++) {
    c.in[i] = ui[i];
    c.out[i] = N[i];
  }
  hipLaunchKernelGGL (k_adv_xy, dim3 (m->nel, d->nz, nf), dim3 (64), 0, st, m->np, m->d_geo,
                      m->d_D, d->phys[0], d->phys[1], c, d->S, (int) conv, (int) cons);
}

#define EW(kern, ...) \
  do { hipLaunchKernelGGL (kern, dim3 (nb256 (d->n)), dim3 (256), 0, st, __VA_ARGS__); } while (0)

// skewSymmetric / altSkewSymmetric (dns/nonlinear.cpp): U holds the
// Fourier velocity (the reference's Us[0] copy of D->u), N -> Uf0; the
// physical-space velocity goes to phys.
// out (S x cols, ld S) = alpha X (S x inner, ld S) W (inner x cols, ld inner): a
// z transform of all S columns of a field as one FP64 GEMM
int zgemm (sx_dns* d, const double* X, const double* W, double* out, int inner, int cols,
           double alpha)
{
  const double beta0 = 0.0;
  const rocblas_status r = rocblas_dgemm (d->blas, rocblas_operation_none,
                                          rocblas_operation_none, d->S, cols, inner, &alpha,
                                          X, d->S, W, inner, &beta0, out, d->S);
  if (r != rocblas_status_success) {
    set_error (std::string ("rocblas_dgemm: ") + rocblas_status_to_string (r));
    return SX_ERR_HIP;
  }
  return SX_OK;
}

// skewSymmetric (dns/nonlinear.cpp:162-193) with every product in z formed
// on M = 3N/2 physical planes (AuxField::transform32,
// src/auxfield.cpp:1197-1258, with the 3/2 rule): the velocity and its z
// derivative are zero-padded and inverse transformed onto M planes, the x-y
// terms and the convective z term summed there, then forward transformed
// and truncated to modes 0..N/2-1; d_z(u_i w) is taken in Fourier space
// from the truncated product.
int advection32 (sx_dns* d, double** U, double** N, hipStream_t st)
{
  const int nc = d->ncom, nz = d->nz, M = d->M;
  const sx_mesh* m = d->mu;
  int rc;
  if (rocblas_set_stream (d->blas, st) != rocblas_status_success) {
    set_error ("rocblas_set_stream failed");
    return SX_ERR_HIP;
  }
  for (int i = 0; i < nc; i++)
    if ((rc = zgemm (d, U[i], d->Wi, d->physM[i], nz, M, 1.0))) return rc;
  const unsigned gM = nb256 (d->nM);
  for (int i = 0; i < nc; i++) {
    Comp3 c {};
    c.in[0] = d->physM[i];
    c.out[0] = d->tmpM;
    hipLaunchKernelGGL (k_adv_xy, dim3 (m->nel, M), dim3 (64), 0, st, m->np, m->d_geo, m->d_D,
                        d->physM[0], d->physM[1], c, d->S, 1, 1);
    if ((rc = zgemm (d, U[i], d->Wdz, d->tmpM2, nz, M, 1.0))) return rc;
    hipLaunchKernelGGL (k_ew_submul, dim3 (gM), dim3 (256), 0, st, d->tmpM, d->physM[2],
                        d->tmpM2, d->nM);
    hipLaunchKernelGGL (k_ew_mul, dim3 (gM), dim3 (256), 0, st, d->tmpM2, d->physM[i],
                        d->physM[2], d->nM);
    SX_HIP (hipGetLastError ());
    if ((rc = zgemm (d, d->tmpM, d->Wf, N[i], M, nz, 0.5))) return rc;
    if ((rc = zgemm (d, d->tmpM2, d->Wf, d->tmp, M, nz, 0.5))) return rc;
    hipLaunchKernelGGL (k_sub_dz, dim3 (nb256 ((size_t) (nz / 2 - 1) * d->S)), dim3 (256), 0,
                        st, N[i], d->tmp, nz, d->S, d->beta);
    if (std::fabs (d->cf[i]) > EPSDP)
      hipLaunchKernelGGL (k_ew_addc, dim3 (nb256 (d->S)), dim3 (256), 0, st, N[i], d->cf[i],
                          (size_t) d->S);
    SX_HIP (hipGetLastError ());
    if ((rc = smooth (d, N[i], st))) return rc;
  }
  return SX_OK;
}

int advection (sx_dns* d, double** U, double** N, hipStream_t st)
{
  if (d->dealias) return advection32 (d, U, N, st);
  const int nc = d->ncom;
  int rc;
  const bool conv = d->advection == SX_SKEW || d->toggle == 1;
  const bool cons = d->advection == SX_SKEW || d->toggle == 0;
  const sx_mesh* m = d->mu;
  if (nc == 3 && fourier_fused_ok (d->nz)) {
    // The three components' transforms of each kind in one launch (the
    // launch has three times the workgroups, so the loads of later ones
    // overlap the stores of earlier ones); the same operations per
    // component as one at a time.  N needs no clearing: k_adv_xy writes
    // every element entry of it (the plane padding stays zero from
    // sx_dns_set_state); H is free until the time-level combination.
    const double* Uc[3] = {U[0], U[1], U[2]};
    const double* Nc[3] = {N[0], N[1], N[2]};
    const double* P3[3] = {d->phys[2], d->phys[2], d->phys[2]};
    const double* Pc[3] = {d->phys[0], d->phys[1], d->phys[2]};
    const double* Hc[3] = {d->H[0], d->H[1], d->H[2]};
    double* Ho[3] = {d->H[0], d->H[1], d->H[2]};
    if ((rc = fourier_z_fused_batch (d->nz, d->S, 3, Uc, nullptr, d->phys, -1, 0.0, nullptr,
                                     nullptr, nullptr, nullptr, st)))
      return rc;
    adv_xy (d, d->phys, N, 3, conv, cons, st);
    // N -= w .* INV(d_z U) in physical space
    if (conv && (rc = fourier_z_fused_batch (d->nz, d->S, 3, Uc, nullptr, N, -1, d->beta, Nc,
                                             P3, nullptr, nullptr, st)))
      return rc;
    // the conservative term d_z (u w) = d_z FWD(u .* w) subtracted in
    // Fourier space, in the transform that takes N back
    // (FWD (N - INV (d_z F)) = FWD (N) - d_z F), with the skew-symmetric
    // 1/2 and ConstForce
    if (cons && (rc = fourier_z_fused_batch (d->nz, d->S, 3, Pc, P3, Ho, +1, 0.0, nullptr,
                                             nullptr, nullptr, nullptr, st)))
      return rc;
    double sc[3], ad[3];
    for (int i = 0; i < 3; i++) {
      sc[i] = d->advection == SX_SKEW ? 0.5 : 1.0;
      ad[i] = std::fabs (d->cf[i]) > EPSDP ? d->cf[i] : 0.0;
    }
    if ((rc = fourier_z_fused_batch (d->nz, d->S, 3, Nc, nullptr, N, +1, cons ? d->beta : 0.0,
                                     cons ? Hc : nullptr, nullptr, sc, ad, st)))
      return rc;
    if ((rc = smooth_n (d, N, 3, st))) return rc;
    if (d->advection == SX_ALTSKEW) d->toggle = 1 - d->toggle;
    SX_HIP (hipGetLastError ());
    return SX_OK;
  }
  for (int i = 0; i < nc; i++)
    if ((rc = transform_to (d, U[i], d->phys[i], -1, st))) return rc;
  for (int i = 0; i < nc; i++) {
    // x and y terms in one pass per element-plane, then the z terms
    adv_xy (d, &d->phys[i], &N[i], 1, conv, cons, st);
    if (nc == 3) {
      // z terms through Fourier space: w d_z u and d_z (u w)
      if (conv) {
        if ((rc = grad (d, U[i], d->tmp, 2, st))) return rc;
        if ((rc = transform (d, d->tmp, -1, st))) return rc;
        EW (k_ew_submul, N[i], d->phys[2], d->tmp, d->n);
      }
      if (cons) {
        EW (k_ew_mul, d->tmp, d->phys[i], d->phys[2], d->n);
        if ((rc = transform (d, d->tmp, +1, st))) return rc;
        if ((rc = grad (d, d->tmp, d->tmp2, 2, st))) return rc;
        if ((rc = transform_to (d, d->tmp2, d->tmp, -1, st))) return rc;
        EW (k_ew_sub, N[i], d->tmp, d->n);
      }
    }
    if (d->advection == SX_SKEW) EW (k_ew_scale, N[i], 0.5, d->n);
    if ((rc = transform (d, N[i], +1, st))) return rc;
    if (std::fabs (d->cf[i]) > EPSDP)
      hipLaunchKernelGGL (k_ew_addc, dim3 (nb256 (d->S)), dim3 (256), 0, st, N[i], d->cf[i],
                          (size_t) d->S);
    if ((rc = smooth (d, N[i], st))) return rc;
  }
  if (d->advection == SX_ALTSKEW) d->toggle = 1 - d->toggle;
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// Integration::StifflyStable / Extrapolation (src/integration.cpp:99-164)
const double SSc[4][4] = {{0, 0, 0, 0}, {1.0, -1.0, 0, 0}, {1.5, -2.0, 0.5, 0},
                          {11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0}};
const double EXc[4][3] = {{0, 0, 0}, {1.0, 0, 0}, {2.0, -1.0, 0}, {3.0, -3.0, 1.0}};

// rollm (dns/integrate.cpp): level q moves to q+1, the oldest to level 0
void roll (double* (&X)[3][3], int nord, int nc)
{
  for (int i = 0; i < nc; i++) {
    double* last = X[nord - 1][i];
    for (int q = nord - 1; q > 0; q--) X[q][i] = X[q - 1][i];
    X[0][i] = last;
  }
}

int startup_system (sx_dns* d, int Je, sx_msys** M, hipStream_t)
{
  *M = (Je >= 1 && Je <= 2) ? d->Mpcg[Je - 1] : nullptr;
  if (!*M) { set_error ("sx_dns_step: start-up JACPCG system missing"); return SX_ERR_ARG; }
  return SX_OK;
}

int dns_step (sx_dns* d, hipStream_t st)
{
  const int nc = d->ncom, nord = d->nord;
  int rc;
  // Us[0] <- u: Us[0] holds the oldest level, which nothing needs any more,
  // so the buffers are exchanged; u is rewritten by the viscous solves (its
  // Nyquist plane, which they skip, is copied back from U below)
  for (int i = 0; i < nc; i++) std::swap (d->u[i], d->Us[0][i]);
  double** U = d->Us[0];
  double** N = d->Uf[0];
  if ((rc = advection (d, U, N, st))) return rc;
  d->step += 1;
  // -- HOPBC: maintainFourier + evaluateCNBCp
  double* plines = nullptr;
  if (d->nb) {
    const int Je = std::min (d->step - 1, nord);
    d->hop_head = (d->hop_head + nord - 1) % nord;     // roll: new level 0
    double* hnew = d->hop[d->hop_head];
    double* un1 = d->un[d->un_head];                   // the previous step's u.n
    double* unnew = d->un[1 - d->un_head];
    // k_hopbc writes every entry of the segments of hnew and unnew outright
    // (the plane padding is never written and stays zero)
    const int nmodes = std::max (1, d->nz / 2);
    Coef3 ac = {{SSc[Je][1], SSc[Je][2], SSc[Je][3]}};
    hipLaunchKernelGGL (k_hopbc, dim3 (d->nb, nmodes), dim3 (64), 0, st, d->mu->np,
                        d->seg_elmt, d->seg_idx, d->seg_nx, d->seg_ny, d->mu->d_geo,
                        d->mu->d_D, U[0], U[1], nc == 3 ? U[2] : U[0], N[0], N[1], hnew,
                        unnew, un1, Je, SSc[Je][0], ac, 1.0 / d->dt, d->nu, d->beta, d->nz,
                        d->nline, d->S);
    SX_HIP (hipGetLastError ());
    d->un_head = 1 - d->un_head;
    const int Jb = std::min (d->step, nord);
    Levels hl = {};
    for (int q = 0; q < Jb; q++) {
      hl.f[q] = d->hop[(d->hop_head + q) % nord];
      hl.c[q] = EXc[Jb][q];
    }
    hipLaunchKernelGGL (k_plines, dim3 (nb256 ((size_t) d->nz * d->nline)), dim3 (256), 0,
                        st, d->plines, hl, Jb, d->nz, d->nline, d->mu->np, d->seg_kind);
    SX_HIP (hipGetLastError ());
    plines = d->plines;
  }
  // -- waveProp
  const int Je = std::min (d->step, nord);
  {
    // all components in one launch (blockIdx.y selects).  D->u after
    // swapData holds u^n = Us[0]: only the Nyquist plane, which the solves
    // leave alone, has to be carried over (by the same kernel)
    Wave3 wv {};
    for (int i = 0; i < nc; i++) {
      wv.H[i] = d->H[i];
      wv.unyq[i] = d->nz > 1 ? d->u[i] : nullptr;
      for (int q = 0; q < Je; q++) {
        wv.us[i].f[q] = d->Us[q][i];
        wv.us[i].c[q] = -SSc[Je][q + 1];
        wv.uf[i].f[q] = d->Uf[q][i];
        wv.uf[i].c[q] = EXc[Je][q] * d->dt;
      }
    }
    // -- and setPForce: div u^ / dt, in the same pass
    hipLaunchKernelGGL (k_wave_div, dim3 (d->mu->nel, std::max (1, d->nz / 2)), dim3 (64, 2), 0,
                        st, wv, Je, nc, d->nz, d->mu->np, d->mu->nel, d->mu->d_geo, d->mu->d_D,
                        d->tmp, d->S, d->beta, d->dt);
  }
  roll (d->Uf, nord, nc);
  const sx_mesh* m = d->mu;
  SX_HIP (hipGetLastError ());
  // -- pressure
  if ((rc = sx_field_solve (d->Mp, d->nz, d->S, d->tmp, d->p, plines, st))) return rc;
  // -- project: F_i = -u^_i/(dt nu) + grad_i p / nu
  const double alpha = -1.0 / (d->dt * d->nu), betap = 1.0 / d->nu;
  hipLaunchKernelGGL (k_project3, dim3 (m->nel, d->nz), dim3 (64), 0, st, m->np, nc, m->d_geo,
                      m->d_D, d->p, d->H[0], d->H[1], nc == 3 ? d->H[2] : nullptr, d->F[0],
                      d->F[1], nc == 3 ? d->F[2] : nullptr, d->S, d->beta, alpha, betap);
  SX_HIP (hipGetLastError ());
  // -- Us: u^n is already in Us[0]; roll
  roll (d->Us, nord, nc);
  // -- viscous substep, with the velocity boundary values (bTransform'd
  //    once by sx_dns_set_velocity_bc; steady)
  const double* lines[3] = {d->ulines[0], d->ulines[1], d->ulines[2]};
  if (d->startup == SX_JACPCG && d->step < nord) {
    // Solve, dns/integrate.cpp:336-350: a JACPCG system of this step's
    // order for the start-up steps, iterating from the current velocity
    // u^n (now in Us[1])
    sx_msys* Mpcg = nullptr;
    if ((rc = startup_system (d, Je, &Mpcg, st))) return rc;
    for (int i = 0; i < nc; i++) {
      SX_HIP (hipMemcpyAsync (d->u[i], d->Us[1][i], sizeof (double) * d->n,
                              hipMemcpyDeviceToDevice, st));
      if ((rc = sx_field_solve_pcg (Mpcg, d->nz, d->S, d->F[i], d->u[i], lines[i],
                                    d->tol_rel, d->step_max, nullptr, st)))
        return rc;
    }
    return SX_OK;
  }
  sx_msys* Mv = d->Mv[Je - 1];
  if (d->group_solve) return field_solve_group (Mv, nc, d->nz, d->S, d->F, d->u, lines, st);
  for (int i = 0; i < nc; i++)
    if ((rc = sx_field_solve (Mv, d->nz, d->S, d->F[i], d->u[i], lines[i], st))) return rc;
  return SX_OK;
}

}  // namespace

extern "C" int sx_dns_destroy (sx_dns_t d)
{
  if (!d) return SX_OK;
  for (double* q : d->bufs) (void) hipFree (q);
  if (d->blas) (void) rocblas_destroy_handle (d->blas);
  void* ip[] = {d->seg_elmt, d->seg_idx, d->seg_kind};
  for (void* q : ip) if (q) (void) hipFree (q);
  for (sx_msys* M : {d->Mv[0], d->Mv[1], d->Mv[2], d->Mp, d->Mpcg[0], d->Mpcg[1]})
    if (M) sx_msys_destroy (M);
  delete d;
  return SX_OK;
}

extern "C" int sx_dns_plane_stride (sx_dns_t d) { return d ? d->S : 0; }

extern "C" int sx_dns_create (sx_mesh_t mu, sx_mesh_t mp, int nz, double beta,
                              double kinvis, double dt, const double* cf, int advection,
                              int nord, void* stream, sx_dns_t* out)
{
  if (!mu || !mp || !out || nz < 1 || (nz > 1 && (nz & 1)) || kinvis <= 0 || dt <= 0 ||
      nord < 1 || nord > 3 || (advection != SX_SKEW && advection != SX_ALTSKEW) ||
      mu->nel != mp->nel || mu->np != mp->np || mu->nbound != mp->nbound || mu->cyl ||
      mp->cyl || mu->np > 16) {
    set_error ("sx_dns_create: bad arguments (Cartesian, nz 1 or even, N_TIME 1..3, "
               "same mesh and boundary segments for velocity and pressure)");
    return SX_ERR_ARG;
  }
  for (int b = 0; b < mu->nbound; b++)
    if (mu->bnd_elmt[b] != mp->bnd_elmt[b] || mu->bnd_side[b] != mp->bnd_side[b]) {
      set_error ("sx_dns_create: velocity and pressure boundary segments differ");
      return SX_ERR_ARG;
    }
  *out = nullptr;
  hipStream_t st = (hipStream_t) stream;
  sx_dns* d = new sx_dns;
  d->mu = mu; d->mp = mp; d->nz = nz; d->ncom = (nz == 1) ? 2 : 3;
  d->beta = beta; d->nu = kinvis; d->dt = dt; d->nord = nord; d->advection = advection;
  if (cf) for (int i = 0; i < 3; i++) d->cf[i] = cf[i];
  const int npnp = mu->npnp;
  d->S = mu->nel * npnp + ((mu->nel * npnp) & 1);
  d->n = (size_t) nz * d->S;
  int rc;
  auto fail = [&] (int r) { sx_dns_destroy (d); return r; };
  const int nmodes = std::max (1, nz / 2);
  // Helmholtz systems: velocity for every order 1..nord (the start-up
  // steps and the full-order ones), pressure (preSolve,
  // dns/integrate.cpp:292-325)
  for (int q = 0; q < nord; q++)
    if ((rc = sx_msys_create (mu, SSc[q + 1][0] / (dt * kinvis), beta, 0, nmodes, SX_FP64,
                              stream, &d->Mv[q])))
      return fail (rc);
  if ((rc = sx_msys_create (mp, 0.0, beta, 0, nmodes, SX_FP64, stream, &d->Mp)))
    return fail (rc);
  // fields
  const int nc = d->ncom;
  // the velocity components share their systems: one band launch solves
  // all of them (falls back to one solve per component when the band is
  // too wide for the panel sweeps)
  d->group_solve = true;
  for (int q = 0; q < nord; q++) {
    rc = msys_reserve_rhs (d->Mv[q], 2 * nc);
    if (rc == SX_ERR_LIMIT) d->group_solve = false;
    else if (rc) return fail (rc);
  }
  for (int i = 0; i < nc; i++) {
    if (!(d->u[i] = d->alloc (d->n)) || !(d->phys[i] = d->alloc (d->n)) ||
        !(d->H[i] = d->alloc (d->n)) || !(d->F[i] = d->alloc (d->n)))
      return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    for (int q = 0; q < nord; q++)
      if (!(d->Us[q][i] = d->alloc (d->n)) || !(d->Uf[q][i] = d->alloc (d->n)))
        return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
  }
  if (!(d->p = d->alloc (d->n)) || !(d->tmp = d->alloc (d->n)) || !(d->tmp2 = d->alloc (d->n)))
    return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
  for (double* q : d->bufs)
    if (hipMemsetAsync (q, 0, sizeof (double) * d->n, st) != hipSuccess)
      return fail (hip_fail (hipErrorUnknown, "sx_dns_create memset"));
  if (int rc = fourier_prepare (nz)) return fail (rc);
  // smoothing: inverse global boundary mass (BCmgr::buildnum imass)
  {
    std::vector<double> mass (mu->nglobal, 0.0), q4 ((size_t) mu->nel * npnp);
    std::vector<double> geo (NGEO * npnp);
    for (int e = 0; e < mu->nel; e++) {
      element_geometry (mu->np, 0, &mu->hx[(size_t) e * npnp], &mu->hy[(size_t) e * npnp], geo.data ());
      for (int i = 0; i < mu->next; i++)
        mass[mu->btog[(size_t) e * mu->next + i]] += geo[G_Q4 * npnp + mu->emap[i]];
    }
    for (double& v : mass) v = 1.0 / v;
    if (!(d->imass = d->alloc (mass.size ())))
      return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    if (hipMemcpy (d->imass, mass.data (), sizeof (double) * mass.size (), hipMemcpyHostToDevice))
      return fail (hip_fail (hipErrorUnknown, "sx_dns_create copy"));
  }
  // boundary segments for the computed pressure conditions
  d->nb = mu->nbound;
  d->nline = d->nb * mu->np;
  if (d->nb) {
    const int np = mu->np;
    std::vector<int> el (d->nb), ix ((size_t) d->nb * np), kind (d->nb);
    std::vector<double> nxv ((size_t) d->nb * np), nyv ((size_t) d->nb * np), geo (NGEO * npnp);
    for (int b = 0; b < d->nb; b++) {
      const int e = mu->bnd_elmt[b], s = mu->bnd_side[b];
      el[b] = e;
      kind[b] = mp->bnd_kind[b];
      element_geometry (np, 0, &mu->hx[(size_t) e * npnp], &mu->hy[(size_t) e * npnp], geo.data ());
      for (int k = 0; k < np; k++) {
        int node;
        double nx, ny;
        switch (s) {      // CCW traverse and outward normal (Element::sideGeom)
        case 0: node = k; nx = -geo[G_SX * npnp + node]; ny = -geo[G_SY * npnp + node]; break;
        case 1: node = k * np + np - 1; nx = geo[G_RX * npnp + node]; ny = geo[G_RY * npnp + node]; break;
        case 2: node = (np - 1) * np + (np - 1 - k); nx = geo[G_SX * npnp + node]; ny = geo[G_SY * npnp + node]; break;
        default: node = (np - 1 - k) * np; nx = -geo[G_RX * npnp + node]; ny = -geo[G_RY * npnp + node]; break;
        }
        const double len = std::hypot (nx, ny);
        ix[(size_t) b * np + k] = node;
        nxv[(size_t) b * np + k] = nx / len;
        nyv[(size_t) b * np + k] = ny / len;
      }
    }
    if (hipMalloc (&d->seg_elmt, sizeof (int) * d->nb) || hipMalloc (&d->seg_idx, sizeof (int) * ix.size ()) ||
        hipMalloc (&d->seg_kind, sizeof (int) * d->nb))
      return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    if (!(d->seg_nx = d->alloc (nxv.size ())) || !(d->seg_ny = d->alloc (nyv.size ())))
      return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    for (int q = 0; q < nord; q++)
      if (!(d->hop[q] = d->alloc ((size_t) nz * d->nline)))
        return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    for (int q = 0; q < 2; q++)
      if (!(d->un[q] = d->alloc ((size_t) nz * d->nline)))
        return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    if (!(d->plines = d->alloc ((size_t) nz * d->nline)))
      return fail (hip_fail (hipErrorOutOfMemory, "sx_dns_create"));
    if (hipMemcpy (d->seg_elmt, el.data (), sizeof (int) * el.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->seg_idx, ix.data (), sizeof (int) * ix.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->seg_kind, kind.data (), sizeof (int) * kind.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->seg_nx, nxv.data (), sizeof (double) * nxv.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->seg_ny, nyv.data (), sizeof (double) * nyv.size (), hipMemcpyHostToDevice))
      return fail (hip_fail (hipErrorUnknown, "sx_dns_create copy"));
    for (double* q : {d->hop[0], d->hop[1], d->hop[2], d->un[0], d->un[1]})
      if (q && hipMemsetAsync (q, 0, sizeof (double) * nz * d->nline, st))
        return fail (hip_fail (hipErrorUnknown, "sx_dns_create memset"));
  }
  if (hipStreamSynchronize (st) != hipSuccess) return fail (hip_fail (hipErrorUnknown, "sync"));
  *out = d;
  return SX_OK;
}

extern "C" int sx_dns_set_state (sx_dns_t d, const double* u_phys, void* stream)
{
  if (!d || !u_phys) { set_error ("sx_dns_set_state: bad arguments"); return SX_ERR_ARG; }
  hipStream_t st = (hipStream_t) stream;
  int rc;
  for (int i = 0; i < d->ncom; i++) {
    SX_HIP (hipMemcpyAsync (d->u[i], u_phys + (size_t) i * d->n, sizeof (double) * d->n,
                            hipMemcpyDeviceToDevice, st));
    if ((rc = transform (d, d->u[i], +1, st))) return rc;
    if (d->nz > 1) SX_HIP (hipMemsetAsync (d->u[i] + d->S, 0, sizeof (double) * d->S, st));
    for (int q = 0; q < d->nord; q++) {
      SX_HIP (hipMemsetAsync (d->Us[q][i], 0, sizeof (double) * d->n, st));
      SX_HIP (hipMemsetAsync (d->Uf[q][i], 0, sizeof (double) * d->n, st));
    }
  }
  SX_HIP (hipMemsetAsync (d->p, 0, sizeof (double) * d->n, st));
  for (double* q : {d->hop[0], d->hop[1], d->hop[2], d->un[0], d->un[1]})
    if (q) SX_HIP (hipMemsetAsync (q, 0, sizeof (double) * d->nz * d->nline, st));
  d->step = 0; d->toggle = 1; d->hop_head = 0; d->un_head = 0;
  return SX_OK;
}

extern "C" int sx_dns_set_velocity_bc (sx_dns_t d, const double* d_lines_phys, void* stream)
{
  if (!d) { set_error ("sx_dns_set_velocity_bc: bad arguments"); return SX_ERR_ARG; }
  hipStream_t st = (hipStream_t) stream;
  const size_t len = (size_t) d->nz * d->nline;
  if (!d_lines_phys || !len) {
    for (int i = 0; i < d->ncom; i++)
      if (d->ulines[i]) SX_HIP (hipMemsetAsync (d->ulines[i], 0, sizeof (double) * len, st));
    return SX_OK;
  }
  for (int i = 0; i < d->ncom; i++) {
    if (!d->ulines[i] && !(d->ulines[i] = d->alloc (len)))
      return hip_fail (hipErrorOutOfMemory, "sx_dns_set_velocity_bc");
    const double* src = d_lines_phys + (size_t) i * len;
    if (d->nz == 2) {
      // Field::bTransform for nz == 2: the first plane keeps its values,
      // the second is zeroed
      SX_HIP (hipMemcpyAsync (d->ulines[i], src, sizeof (double) * d->nline,
                              hipMemcpyDeviceToDevice, st));
      SX_HIP (hipMemsetAsync (d->ulines[i] + d->nline, 0, sizeof (double) * d->nline, st));
    } else if (int rc = fourier_z (d->nz, d->nline, src, d->ulines[i], +1, st)) {
      return rc;
    }
  }
  return SX_OK;
}

extern "C" int sx_dns_set_dealias (sx_dns_t d, int on, void* stream)
{
  if (!d || (on && (d->ncom != 3 || d->nz < 4 || d->nz % 4 || d->advection != SX_SKEW))) {
    set_error ("sx_dns_set_dealias: needs a 3D field (nz a multiple of 4) and the "
               "skew-symmetric advection");
    return SX_ERR_ARG;
  }
  (void) stream;
  if (on && !d->blas) {
    const int N = d->nz, M = 3 * N / 2;
    d->M = M;
    d->nM = (size_t) M * d->S;
    if (rocblas_create_handle (&d->blas) != rocblas_status_success) {
      set_error ("rocblas_create_handle failed");
      return SX_ERR_HIP;
    }
    for (int i = 0; i < 3; i++)
      if (!(d->physM[i] = d->alloc (d->nM))) return hip_fail (hipErrorOutOfMemory, "dealias");
    if (!(d->tmpM = d->alloc (d->nM)) || !(d->tmpM2 = d->alloc (d->nM)) ||
        !(d->Wi = d->alloc ((size_t) N * M)) || !(d->Wdz = d->alloc ((size_t) N * M)) ||
        !(d->Wf = d->alloc ((size_t) N * M)))
      return hip_fail (hipErrorOutOfMemory, "dealias");
    // packed real DFT matrices (femlib/fourier.c packing, the Nyquist plane
    // of the N planes dropped): Wi (N x M, ld N) evaluates modes 0..N/2-1
    // at the M points, Wdz the same after d/dz (i beta k), Wf (M x N, ld M)
    // the forward transform scaled 1/M truncated to those modes
    std::vector<double> wi ((size_t) N * M, 0.0), wdz ((size_t) N * M, 0.0),
        wf ((size_t) N * M, 0.0);
    const double tp = 2.0 * M_PI;
    for (int j = 0; j < M; j++) {
      wi[(size_t) j * N] = 1.0;
      wf[j] = 1.0 / M;
      for (int k = 1; k < N / 2; k++) {
        const double a = tp * (double) (((long long) j * k) % M) / M;
        const double c = std::cos (a), sn = std::sin (a), bk = d->beta * k;
        wi[(size_t) j * N + 2 * k] = 2.0 * c;
        wi[(size_t) j * N + 2 * k + 1] = -2.0 * sn;
        wdz[(size_t) j * N + 2 * k] = -2.0 * bk * sn;
        wdz[(size_t) j * N + 2 * k + 1] = -2.0 * bk * c;
        wf[(size_t) (2 * k) * M + j] = c / M;
        wf[(size_t) (2 * k + 1) * M + j] = -sn / M;
      }
    }
    if (hipMemcpy (d->Wi, wi.data (), sizeof (double) * wi.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->Wdz, wdz.data (), sizeof (double) * wdz.size (), hipMemcpyHostToDevice) ||
        hipMemcpy (d->Wf, wf.data (), sizeof (double) * wf.size (), hipMemcpyHostToDevice))
      return hip_fail (hipErrorUnknown, "dealias matrices");
  }
  d->dealias = on ? 1 : 0;
  return SX_OK;
}

extern "C" int sx_dns_set_startup (sx_dns_t d, int method, double tol_rel, int step_max,
                                   void* stream)
{
  if (!d || (method != SX_FP64 && method != SX_JACPCG) || !(tol_rel >= 0.0) || step_max < 1) {
    set_error ("sx_dns_set_startup: bad arguments");
    return SX_ERR_ARG;
  }
  // the start-up steps 1..nord-1 run at orders Je = step, constant
  // alpha_0(Je) / (dt nu) (dns/integrate.cpp:336-350)
  for (int q = 0; method == SX_JACPCG && q + 1 < d->nord; q++)
    if (!d->Mpcg[q]) {
      const int rc = sx_msys_create (d->mu, SSc[q + 1][0] / (d->dt * d->nu), d->beta, 0,
                                     std::max (1, d->nz / 2), SX_JACPCG, stream, &d->Mpcg[q]);
      if (rc) return rc;
    }
  d->startup = method;
  d->tol_rel = tol_rel;
  d->step_max = step_max;
  return SX_OK;
}

extern "C" int sx_dns_step (sx_dns_t d, int nsteps, void* stream)
{
  if (!d || nsteps < 0) { set_error ("sx_dns_step: bad arguments"); return SX_ERR_ARG; }
  for (int s = 0; s < nsteps; s++) {
    const int rc = dns_step (d, (hipStream_t) stream);
    if (rc) return rc;
  }
  return SX_OK;
}

extern "C" int sx_dns_get (sx_dns_t d, int which, double* d_out, int* step, void* stream)
{
  if (!d || !d_out || which < 0 || which > d->ncom) {
    set_error ("sx_dns_get: bad arguments");
    return SX_ERR_ARG;
  }
  const double* src = (which == d->ncom) ? d->p : d->u[which];
  SX_HIP (hipMemcpyAsync (d_out, src, sizeof (double) * d->n, hipMemcpyDeviceToDevice,
                          (hipStream_t) stream));
  if (step) *step = d->step;
  return SX_OK;
}
