// sx_tape.hip -- the element-local operators of the DNS step as plane
// kernels with their transposes, for the hand-written adjoint of the step
// (no reference counterpart: semtex has no adjoint).  Planes are
// [nz][stride] doubles, element-major, row-major np x np per element, as
// the reference's AuxField.
//
//   grad   out = d/dx (dir 0) or d/dy (dir 1) per element
//          (Element::grad, src/element.cpp:521-565):
//            out[i][j] = rx[i][j] sum_k u[i][k] D[j][k] + sx[i][j] sum_k D[i][k] u[k][j]
//          transpose: u'[a][b] = sum_j g[a][j] rx[a][j] D[j][b]
//                                + sum_i g[i][b] sx[i][b] D[i][a]
//   smooth Field::smooth (src/field.cpp:282-327): on element-boundary nodes
//          the mass-weighted average over the elements sharing the global
//          node, m_g^-1 sum_s Q4_s f_s; interior nodes unchanged.
//          transpose: boundary slot s of node g gets Q4_s m_g^-1 sum_s' g_s'.
//   dz     d/dz of Fourier planes (AuxField::gradient, direction 2,
//          src/auxfield.cpp:640-668): Re_k <- -beta k Im_k,
//          Im_k <- beta k Re_k, planes 0 and 1 zero; its transpose is -dz.
#include "sx_internal.h"

using namespace sx;

namespace {

constexpr int TB = 256;

__global__ __launch_bounds__(TB) void k_tape_grad (int np, int nel, const double* __restrict__ geo,
                                                   const double* __restrict__ Dg,
                                                   const double* __restrict__ in, double* out,
                                                   int stride, int dir, int transpose)
{
  const int npnp = np * np;
  const size_t n = (size_t) nel * npnp;
  const size_t idx = (size_t) blockIdx.x * TB + threadIdx.x;
  __shared__ double D[256];
  for (int q = threadIdx.x; q < npnp; q += TB) D[q] = Dg[q];
  __syncthreads ();
  if (idx >= n) return;
  const int e = (int) (idx / npnp), ij = (int) (idx % npnp), i = ij / np, j = ij % np;
  const double* ge = geo + (size_t) e * NGEO * npnp;
  const double* r = ge + (dir == 0 ? G_RX : G_RY) * npnp;
  const double* s = ge + (dir == 0 ? G_SX : G_SY) * npnp;
  const double* u = in + (size_t) blockIdx.y * stride + (size_t) e * npnp;
  double acc = 0.0;
  if (!transpose) {
    double ur = 0.0, us = 0.0;
    for (int k = 0; k < np; k++) {
      ur += u[i * np + k] * D[j * np + k];
      us += D[i * np + k] * u[k * np + j];
    }
    acc = r[ij] * ur + s[ij] * us;
  } else {
    // (a, b) = (i, j): sum_q g[i][q] r[i][q] D[q][j] + sum_q g[q][j] s[q][j] D[q][i]
    for (int q = 0; q < np; q++) {
      acc += u[i * np + q] * r[i * np + q] * D[q * np + j];
      acc += u[q * np + j] * s[q * np + j] * D[q * np + i];
    }
  }
  out[(size_t) blockIdx.y * stride + idx] = acc;
}

__global__ __launch_bounds__(TB) void k_tape_smooth (int np, int nel, int next,
                                                     const int* __restrict__ emap_g,
                                                     const int* __restrict__ btog,
                                                     const int* __restrict__ gptr,
                                                     const int* __restrict__ gsrc,
                                                     const double* __restrict__ geo,
                                                     const double* __restrict__ in, double* out,
                                                     int stride, int transpose)
{
  const int npnp = np * np;
  const size_t n = (size_t) nel * npnp;
  const size_t idx = (size_t) blockIdx.x * TB + threadIdx.x;
  __shared__ int emap[256], pmap[256];
  for (int q = threadIdx.x; q < npnp; q += TB) emap[q] = emap_g[q];
  __syncthreads ();
  for (int q = threadIdx.x; q < npnp; q += TB) pmap[emap[q]] = q;
  __syncthreads ();
  if (idx >= n) return;
  const int e = (int) (idx / npnp), ij = (int) (idx % npnp), p = pmap[ij];
  const double* f = in + (size_t) blockIdx.y * stride;
  double* o = out + (size_t) blockIdx.y * stride;
  if (p >= next) { o[idx] = f[idx]; return; }      // interior node
  const int g = btog[(size_t) e * next + p];
  double mass = 0.0, sum = 0.0;
  for (int c = gptr[g]; c < gptr[g + 1]; c++) {
    const int slot = gsrc[c], es = slot / npnp, ijs = emap[slot % npnp];
    const double q4 = geo[((size_t) es * NGEO + G_Q4) * npnp + ijs];
    const double v = f[(size_t) es * npnp + ijs];
    mass += q4;
    sum += transpose ? v : q4 * v;
  }
  o[idx] = transpose ? geo[((size_t) e * NGEO + G_Q4) * npnp + ij] * sum / mass : sum / mass;
}

__global__ __launch_bounds__(TB) void k_tape_dz (int nz, int n, int stride, double beta,
                                                 const double* __restrict__ in, double* out,
                                                 int transpose)
{
  const int j = blockIdx.x * TB + threadIdx.x, k = blockIdx.y;
  if (j >= n) return;
  double v = 0.0;
  if (k >= 2) {
    const double bk = beta * (k >> 1) * (transpose ? -1.0 : 1.0);
    v = (k & 1) ? bk * in[(size_t) (k - 1) * stride + j] : -bk * in[(size_t) (k + 1) * stride + j];
  }
  out[(size_t) k * stride + j] = v;
}

}  // namespace

extern "C" int sx_grad_planes (sx_mesh_t m, int nz, int stride, const double* d_in, double* d_out,
                               int dir, int transpose, void* stream)
{
  if (!m || nz < 1 || stride < m->nel * m->npnp || !d_in || !d_out || d_in == d_out ||
      (dir != 0 && dir != 1) || m->npnp > 256) {
    set_error ("sx_grad_planes: bad arguments (out of place, dir 0 or 1)");
    return SX_ERR_ARG;
  }
  const size_t n = (size_t) m->nel * m->npnp;
  hipLaunchKernelGGL (k_tape_grad, dim3 ((unsigned) ((n + TB - 1) / TB), nz), dim3 (TB), 0,
                      (hipStream_t) stream, m->np, m->nel, m->d_geo, m->d_D, d_in, d_out, stride,
                      dir, transpose ? 1 : 0);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

extern "C" int sx_smooth_planes (sx_mesh_t m, int nz, int stride, const double* d_in,
                                 double* d_out, int transpose, void* stream)
{
  if (!m || nz < 1 || stride < m->nel * m->npnp || !d_in || !d_out || d_in == d_out ||
      m->npnp > 256) {
    set_error ("sx_smooth_planes: bad arguments (out of place)");
    return SX_ERR_ARG;
  }
  const size_t n = (size_t) m->nel * m->npnp;
  hipLaunchKernelGGL (k_tape_smooth, dim3 ((unsigned) ((n + TB - 1) / TB), nz), dim3 (TB), 0,
                      (hipStream_t) stream, m->np, m->nel, m->next, m->d_emap, m->d_btog,
                      m->d_gptr, m->d_gsrc, m->d_geo, d_in, d_out, stride, transpose ? 1 : 0);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

extern "C" int sx_dz_planes (int nz, int stride, int n, double beta, const double* d_in,
                             double* d_out, int transpose, void* stream)
{
  if (nz < 1 || n < 0 || stride < n || !d_in || !d_out || d_in == d_out || (nz > 1 && (nz & 1))) {
    set_error ("sx_dz_planes: bad arguments (out of place, nz 1 or even)");
    return SX_ERR_ARG;
  }
  if (n == 0) return SX_OK;
  hipLaunchKernelGGL (k_tape_dz, dim3 ((unsigned) ((n + TB - 1) / TB), nz), dim3 (TB), 0,
                      (hipStream_t) stream, nz, n, stride, beta, d_in, d_out, transpose ? 1 : 0);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}
