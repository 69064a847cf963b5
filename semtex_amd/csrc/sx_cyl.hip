// sx_cyl.hip -- coupling of the radial and azimuthal velocity fields of a
// cylindrical 3D problem (AuxField::couple, src/auxfield.cpp:1376-1451;
// Field::coupleBCs FORWARD, src/field.cpp:1127-1201).  Per Fourier mode,
// with v = Vr + i Vi and w = Wr + i Wi:
//   FORWARD  v~ = v + i w = (Vr - Wi) + i (Vi + Wr)
//            w~ = v - i w = (Vr + Wi) + i (Vi - Wr)
//   INVERSE  v  = (v~ + w~)/2,  w = i (w~ - v~)/2
// One thread per (mode, node); all four values of a node are read before
// any is written, so the update is in place.
#include "sx_internal.h"

using namespace sx;

namespace {

__global__ __launch_bounds__(256) void k_couple (int n, int stride, int k0, int nmode,
                                                 double* __restrict__ v, double* __restrict__ w,
                                                 int dir)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int k = k0 + (int) blockIdx.y;
  if (i >= (size_t) n || k >= nmode) return;
  const size_t re = (size_t) (2 * k) * stride + i, im = re + stride;
  const double vr = v[re], vi = v[im], wr = w[re], wi = w[im];
  if (dir > 0) {
    v[re] = vr - wi;
    v[im] = vi + wr;
    w[re] = vr + wi;
    w[im] = vi - wr;
  } else {
    v[re] = 0.5 * (vr + wr);
    v[im] = 0.5 * (vi + wi);
    w[re] = 0.5 * (vi - wi);
    w[im] = 0.5 * (wr - vr);
  }
}

}  // namespace

extern "C" int sx_couple (int nz, int stride, int n, double* d_v, double* d_w, int dir,
                          int base_mode, void* stream)
{
  if (nz < 1 || n < 0 || stride < n || !d_v || !d_w || d_v == d_w ||
      (dir != SX_FORWARD && dir != SX_INVERSE) || base_mode < 0 || (nz > 1 && (nz & 1))) {
    set_error ("sx_couple: bad arguments (nz 1 or even, stride >= n, dir +-1, distinct fields)");
    return SX_ERR_ARG;
  }
  if (nz == 1 || n == 0) return SX_OK;              // 2D: no coupling
  const int nmode = nz / 2, k0 = base_mode == 0 ? 1 : 0;
  if (k0 >= nmode) return SX_OK;
  hipLaunchKernelGGL (k_couple, dim3 ((unsigned) ((n + 255) / 256), (unsigned) (nmode - k0)),
                      dim3 (256), 0, (hipStream_t) stream, n, stride, k0, nmode, d_v, d_w, dir);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}
