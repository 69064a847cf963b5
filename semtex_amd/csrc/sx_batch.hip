// sx_batch.hip -- a batch of independent Helmholtz problems (each its own
// mesh, order, lambda2, geometry) solved by one captured HIP graph: for
// every problem the forcing is copied from its source buffer (Field::solve
// destroys it) and the problem's solve kernels follow, FP64 or mixed
// precision with a fixed number of FP64 refinement passes.  The problems
// are spread over lane streams forked from and joined back into the
// capturing stream, so the graph holds independent branches that run
// concurrently.  Replaying the graph costs one launch for the whole batch.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "sx_internal.h"

using namespace sx;

constexpr int BL = 16;             // lane streams at most

struct sx_batch {
  hipStream_t cap = nullptr;
  hipStream_t lane[BL] = {};
  hipEvent_t fork = nullptr, join[BL] = {};
  int nl = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int n = 0;
};

extern "C" int sx_batch_destroy (sx_batch_t b)
{
  if (!b) return SX_OK;
  if (b->exec) (void) hipGraphExecDestroy (b->exec);
  if (b->graph) (void) hipGraphDestroy (b->graph);
  for (int l = 0; l < BL; l++) {
    if (b->lane[l]) (void) hipStreamDestroy (b->lane[l]);
    if (b->join[l]) (void) hipEventDestroy (b->join[l]);
  }
  if (b->fork) (void) hipEventDestroy (b->fork);
  if (b->cap) (void) hipStreamDestroy (b->cap);
  delete b;
  return SX_OK;
}

// problem i on the lane of its system (systems taken round-robin), the
// lanes forked from st and joined into it
static int record (const sx_batch* b, int n, const sx_msys_t* sys, const int* nz,
                   const int* stride, const double* const* f_src, double* const* f_work,
                   double* const* u, const double* const* lines, int passes, hipStream_t st)
{
  SX_HIP (hipEventRecord (b->fork, st));
  for (int l = 0; l < b->nl; l++) SX_HIP (hipStreamWaitEvent (b->lane[l], b->fork, 0));
  // problems that share a system (its workspace) keep to one lane, in order
  std::vector<sx_msys_t> seen;
  for (int i = 0; i < n; i++) {
    int sl = 0;
    while (sl < (int) seen.size () && seen[sl] != sys[i]) sl++;
    if (sl == (int) seen.size ()) seen.push_back (sys[i]);
    const hipStream_t ls = b->lane[sl % b->nl];
    const size_t bytes = sizeof (double) * (size_t) nz[i] * stride[i];
    SX_HIP (hipMemcpyAsync (f_work[i], f_src[i], bytes, hipMemcpyDeviceToDevice, ls));
    const double* li = lines ? lines[i] : nullptr;
    int rc;
    if (sys[i]->precision == SX_MIXED)
      rc = field_solve_mixed_passes (sys[i], nz[i], stride[i], f_work[i], u[i], li, passes, ls);
    else
      rc = sx_field_solve (sys[i], nz[i], stride[i], f_work[i], u[i], li, ls);
    if (rc) return rc;
  }
  for (int l = 0; l < b->nl; l++) {
    SX_HIP (hipEventRecord (b->join[l], b->lane[l]));
    SX_HIP (hipStreamWaitEvent (st, b->join[l], 0));
  }
  return SX_OK;
}

extern "C" int sx_batch_create (int n, const sx_msys_t* systems, const int* nz,
                                const int* plane_stride, const double* const* d_f_src,
                                double* const* d_f_work, double* const* d_u,
                                const double* const* d_lines, int refine_passes,
                                sx_batch_t* out)
{
  if (n < 1 || !systems || !nz || !plane_stride || !d_f_src || !d_f_work || !d_u || !out ||
      refine_passes < 1) {
    set_error ("sx_batch_create: bad arguments");
    return SX_ERR_ARG;
  }
  for (int i = 0; i < n; i++)
    if (!systems[i] || systems[i]->precision == SX_JACPCG) {
      set_error ("sx_batch_create: DIRECT systems (SX_FP64 or SX_MIXED) only");
      return SX_ERR_ARG;
    }
  *out = nullptr;
  sx_batch* b = new sx_batch;
  b->n = n;
  auto fail = [&] (int r) { sx_batch_destroy (b); return r; };
  if (hipStreamCreate (&b->cap) != hipSuccess) return fail (hip_fail (hipErrorUnknown, "hipStreamCreate"));
  // lanes: 4, as many as the hardware queues a process gets by default
  // (measured: 4 lanes 2.9 ms FP64 / 7.4 ms mixed per 64-problem batch, 8
  // lanes 3.0 / 9.5, 16 lanes 2.9 / 8.3, 1 lane 8.8 / 22.3);
  // SX_BATCH_LANES in the environment (1..16) overrides
  int nl = 4;
  if (const char* e = std::getenv ("SX_BATCH_LANES")) nl = std::max (1, std::min (BL, std::atoi (e)));
  b->nl = n < nl ? n : nl;
  if (hipEventCreateWithFlags (&b->fork, hipEventDisableTiming) != hipSuccess)
    return fail (hip_fail (hipErrorUnknown, "hipEventCreate"));
  for (int l = 0; l < b->nl; l++)
    if (hipStreamCreate (&b->lane[l]) != hipSuccess ||
        hipEventCreateWithFlags (&b->join[l], hipEventDisableTiming) != hipSuccess)
      return fail (hip_fail (hipErrorUnknown, "lane stream"));
  // the branches run concurrently: the recorded launches use no split
  // sweeps (they need their whole grid resident at once); the systems'
  // own setting is restored once the graph is recorded
  struct NoSplit {
    std::vector<std::pair<sx_msys*, int>> saved;
    ~NoSplit () { for (auto it = saved.rbegin (); it != saved.rend (); ++it) it->first->no_split = it->second; }
  } guard;
  for (int i = 0; i < n; i++) {
    guard.saved.emplace_back (systems[i], systems[i]->no_split);
    systems[i]->no_split = 1;
  }
  // one eager pass first: workspaces and kernel attributes are set up
  // outside the capture
  int rc = record (b, n, systems, nz, plane_stride, d_f_src, d_f_work, d_u, d_lines,
                   refine_passes, b->cap);
  if (rc) return fail (rc);
  if (hipStreamSynchronize (b->cap) != hipSuccess) return fail (hip_fail (hipErrorUnknown, "sync"));
  if (hipStreamBeginCapture (b->cap, hipStreamCaptureModeRelaxed) != hipSuccess)
    return fail (hip_fail (hipErrorUnknown, "hipStreamBeginCapture"));
  rc = record (b, n, systems, nz, plane_stride, d_f_src, d_f_work, d_u, d_lines, refine_passes,
               b->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture (b->cap, &g);
  b->graph = g;
  if (rc) return fail (rc);
  if (ec != hipSuccess) return fail (hip_fail (ec, "hipStreamEndCapture"));
  hipError_t ei = hipGraphInstantiate (&b->exec, b->graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) return fail (hip_fail (ei, "hipGraphInstantiate"));
  *out = b;
  return SX_OK;
}

extern "C" int sx_batch_run (sx_batch_t b, void* stream)
{
  if (!b || !b->exec) { set_error ("sx_batch_run: bad arguments"); return SX_ERR_ARG; }
  SX_HIP (hipGraphLaunch (b->exec, (hipStream_t) stream));
  return SX_OK;
}
