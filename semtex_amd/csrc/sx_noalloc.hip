// sx_noalloc.hip -- the mixed-precision Field::solve without device
// allocation, for callers that cannot allocate (tight serving loops,
// harnesses that forbid hipMalloc).
//
// sx_field_solve_mixed keeps its workspace (five planes-sets of the field
// plus the correction-norm partials) with the system and grows it with
// hipMalloc on first use.  sx_field_solve_mixed_noalloc runs the same
// kernels on a file-scope device array instead.  Each GPU the library runs
// on holds its own copy of the array until the process ends; its device
// address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels and copies.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two host threads) race.  The kernels write every
// workspace entry they read later, so the array needs no clearing between
// calls.
#include <mutex>

#include "sx_internal.h"

using namespace sx;

static __device__ __align__(256) double scratch[SX_NOALLOC_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES];          // per device, filled on first use

// The current device's address of `scratch` (nullptr with the error set).
double* device_scratch ()
{
  int dev = -1;
  if (hipGetDevice (&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    set_error ("sx_field_solve_mixed_noalloc: no usable current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock (scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));
    if (e != hipSuccess || !p) {
      hip_fail (e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    scratch_ptr[dev] = static_cast<double*> (p);
  }
  return scratch_ptr[dev];
}

}  // namespace

extern "C" int sx_field_solve_mixed_noalloc (sx_msys_t s, int nz, int stride, double* d_f,
                                             double* d_u, const double* d_lines,
                                             double refine_tol, int max_iter, int* iters,
                                             void* stream)
{
  if (int rc = check_mixed_solve_args (s, nz, stride, d_f, d_u, max_iter,
                                       "sx_field_solve_mixed_noalloc"))
    return rc;
  const size_t need = mixed_work_doubles (s, nz, stride);
  if (need > (size_t) SX_NOALLOC_DOUBLES) {
    set_error ("sx_field_solve_mixed_noalloc: the workspace of nz = " + std::to_string (nz) +
               " planes of stride " + std::to_string (stride) + " needs " +
               std::to_string (need) + " doubles, above the fixed " +
               std::to_string ((size_t) SX_NOALLOC_DOUBLES));
    return SX_ERR_LIMIT;
  }
  double* work = device_scratch ();
  if (!work) return SX_ERR_HIP;
  return field_solve_mixed_ws (s, nz, stride, d_f, d_u, d_lines, refine_tol, max_iter, iters,
                               work, (hipStream_t) stream);
}
