// sx_batch.hip -- a batch of independent Helmholtz problems (each its own
// mesh, order, lambda2, geometry) solved by one captured HIP graph: for
// every problem the forcing is copied from its source buffer (Field::solve
// destroys it) and the problem's solve follows, FP64 or mixed precision
// with a fixed number of FP64 refinement passes.
//
// Fused form (the default): the problems' solves are planned, not
// launched (plan_* in the solver sources), and every kernel instance of a
// phase runs once for all the problems that use it (k_batch: blockIdx.z
// picks the problem), so the graph holds a few dozen launches instead of
// ~10 per problem.  Problems that share a system (and so its workspace)
// are planned one after the other.  The right-hand sides per band-sweep
// workgroup are chosen per class of band system (D, unit-diagonal,
// twisted) by sx_batch_tune, which times the alternatives.
//
// Graph-of-launches form (SX_BATCH_FUSED=0, or a system the plans do not
// cover: offset-major band sweeps, untiled element kernels): the
// problems' ordinary launches, spread over lane streams forked from and
// joined back into the capturing stream, so the graph holds independent
// branches that run concurrently.
#include <algorithm>
#include <array>
#include <cstdio>
#include <string>
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
  // fused form
  int fused = 0, passes = 1, launches = 0;
  std::vector<sx_msys_t> sys;
  std::vector<int> nz, stride;
  std::vector<const double*> fsrc, lines;
  std::vector<double*> fwork, u;
  std::vector<int> cls;                // band-system class of every problem
  std::vector<std::array<int, 3>> keys;  // class keys: sw_D, sw_unit, tw
  std::vector<int> per;                // right-hand sides per sweep workgroup, per class (0: default)
  void* d_plan = nullptr;              // the merged launches' arguments and grids
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
  if (b->d_plan) (void) hipFree (b->d_plan);
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

namespace {

__device__ __forceinline__ void copy_body (double* dst, const double* src, size_t n)
{
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
struct FnCopy {
  __device__ __forceinline__ void operator() (double* d, const double* s, size_t n) const
  { copy_body (d, s, n); }
};

// the band-system class of a system: (off-diagonal blocks, unit-diagonal
// panels, twisted), which decides the sweep kernel instance
int class_of (sx_batch* b, const sx_msys* s)
{
  const std::array<int, 3> k = {s->sw_D, s->sw_unit, s->tw};
  for (size_t c = 0; c < b->keys.size (); c++)
    if (b->keys[c] == k) return (int) c;
  b->keys.push_back (k);
  b->per.push_back (0);
  return (int) b->keys.size () - 1;
}

// every problem's copy and solve, recorded; problems that share a system
// start where the previous one on it ended
int fused_plan (const sx_batch* b, BatchPlan& P)
{
  std::vector<std::pair<const sx_msys*, int>> end;
  for (int i = 0; i < b->n; i++) {
    const sx_msys* s = b->sys[i];
    int start = 0;
    for (auto& e : end)
      if (e.first == s) start = std::max (start, e.second);
    P.phase = start;
    const size_t cnt = (size_t) b->nz[i] * b->stride[i];
    batch_rec<256, FnCopy> (P, dim3 ((unsigned) ((cnt + 255) / 256)), 256, 0, b->fwork[i],
                            b->fsrc[i], cnt);
    P.phase++;
    P.tune = b->per[b->cls[i]];
    const int rc = s->precision == SX_MIXED
        ? plan_mixed_passes (P, b->sys[i], b->nz[i], b->stride[i], b->fwork[i], b->u[i],
                             b->lines[i], b->passes)
        : plan_field_solve_direct (P, s, b->nz[i], b->stride[i], b->fwork[i], b->u[i],
                                   b->lines[i]);
    if (rc) return rc;
    bool found = false;
    for (auto& e : end)
      if (e.first == s) { e.second = P.phase; found = true; }
    if (!found) end.push_back ({s, P.phase});
  }
  return SX_OK;
}

// the merged launches in phase order, their arguments uploaded, captured
// into the batch's graph (replacing any previous one)
int fused_record (sx_batch* b)
{
  BatchPlan P;
  if (int rc = fused_plan (b, P)) return rc;
  std::stable_sort (P.groups.begin (), P.groups.end (),
                    [] (const BatchPlan::Group& x, const BatchPlan::Group& y) {
                      return x.phase < y.phase;
                    });
  auto al = [] (size_t v) { return (v + 255) / 256 * 256; };
  size_t bytes = 0;
  for (auto& g : P.groups) bytes += al (g.blob.size ()) + al (sizeof (uint2) * g.grid.size ());
  if (b->d_plan) (void) hipFree (b->d_plan);
  b->d_plan = nullptr;
  SX_HIP (hipMalloc (&b->d_plan, std::max<size_t> (bytes, 256)));
  char* base = (char*) b->d_plan;
  size_t off = 0;
  for (auto& g : P.groups) {
    g.d_args = base + off;
    SX_HIP (hipMemcpy (g.d_args, g.blob.data (), g.blob.size (), hipMemcpyHostToDevice));
    off += al (g.blob.size ());
    g.d_grid = (uint2*) (base + off);
    SX_HIP (hipMemcpy (g.d_grid, g.grid.data (), sizeof (uint2) * g.grid.size (),
                       hipMemcpyHostToDevice));
    off += al (sizeof (uint2) * g.grid.size ());
  }
  if (b->exec) { (void) hipGraphExecDestroy (b->exec); b->exec = nullptr; }
  if (b->graph) { (void) hipGraphDestroy (b->graph); b->graph = nullptr; }
  // the kernel instances of one phase are independent: they run on lane
  // streams forked from the capturing stream and joined back before the
  // next phase
  size_t most = 1;
  for (size_t i = 0, j; i < P.groups.size (); i = j) {
    for (j = i; j < P.groups.size () && P.groups[j].phase == P.groups[i].phase; j++) {}
    most = std::max (most, j - i);
  }
  const int want = (int) std::min<size_t> (most, BL);
  for (int l = b->nl; l < want; l++) {
    SX_HIP (hipStreamCreate (&b->lane[l]));
    SX_HIP (hipEventCreateWithFlags (&b->join[l], hipEventDisableTiming));
  }
  b->nl = std::max (b->nl, want);
  if (!b->fork) SX_HIP (hipEventCreateWithFlags (&b->fork, hipEventDisableTiming));
  auto run_groups = [&] () -> int {
    for (size_t i = 0, j; i < P.groups.size (); i = j) {
      for (j = i; j < P.groups.size () && P.groups[j].phase == P.groups[i].phase; j++) {}
      if (j - i == 1) {
        auto& g = P.groups[i];
        if (int rc = g.launch (g.d_args, g.d_grid, (int) g.grid.size (), g.gx, g.gy, g.block,
                               g.lds, b->cap))
          return rc;
        continue;
      }
      const int nl = (int) std::min<size_t> (j - i, (size_t) b->nl);
      SX_HIP (hipEventRecord (b->fork, b->cap));
      for (int l = 0; l < nl; l++) SX_HIP (hipStreamWaitEvent (b->lane[l], b->fork, 0));
      for (size_t k = i; k < j; k++) {
        auto& g = P.groups[k];
        if (int rc = g.launch (g.d_args, g.d_grid, (int) g.grid.size (), g.gx, g.gy, g.block,
                               g.lds, b->lane[(k - i) % nl]))
          return rc;
      }
      for (int l = 0; l < nl; l++) {
        SX_HIP (hipEventRecord (b->join[l], b->lane[l]));
        SX_HIP (hipStreamWaitEvent (b->cap, b->join[l], 0));
      }
    }
    return SX_OK;
  };
  // one eager pass (kernel attributes set outside the capture), then the capture
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1)
      SX_HIP (hipStreamBeginCapture (b->cap, hipStreamCaptureModeRelaxed));
    const int rc = run_groups ();
    if (pass == 1) {
      hipGraph_t gr = nullptr;
      const hipError_t ec = hipStreamEndCapture (b->cap, &gr);
      b->graph = gr;
      if (rc) return rc;
      if (ec != hipSuccess) return hip_fail (ec, "hipStreamEndCapture");
    } else {
      if (rc) return rc;
      SX_HIP (hipStreamSynchronize (b->cap));
    }
  }
  SX_HIP (hipGraphInstantiate (&b->exec, b->graph, nullptr, nullptr, 0));
  b->launches = (int) P.groups.size ();
  return SX_OK;
}

}  // namespace

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
  // fused form first
  b->passes = refine_passes;
  const char* ef = std::getenv ("SX_BATCH_FUSED");
  if (!ef || std::atoi (ef) != 0) {
    for (int i = 0; i < n; i++) {
      b->sys.push_back (systems[i]);
      b->nz.push_back (nz[i]);
      b->stride.push_back (plane_stride[i]);
      b->fsrc.push_back (d_f_src[i]);
      b->fwork.push_back (d_f_work[i]);
      b->u.push_back (d_u[i]);
      b->lines.push_back (d_lines ? d_lines[i] : nullptr);
      b->cls.push_back (class_of (b, systems[i]));
    }
    const int rc = fused_record (b);
    if (rc == SX_OK) {
      b->fused = 1;
      *out = b;
      return SX_OK;
    }
    if (rc != SX_ERR_LIMIT) return fail (rc);
    if (b->d_plan) { (void) hipFree (b->d_plan); b->d_plan = nullptr; }
  }
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

extern "C" int sx_batch_tune (sx_batch_t b, int reps, void* stream)
{
  if (!b || !b->exec || reps < 1) { set_error ("sx_batch_tune: bad arguments"); return SX_ERR_ARG; }
  if (!b->fused) return SX_OK;                       // nothing to choose
  hipStream_t st = (hipStream_t) stream;
  hipEvent_t e0, e1;
  SX_HIP (hipEventCreate (&e0));
  SX_HIP (hipEventCreate (&e1));
  auto time_ms = [&] (double& ms) -> int {
    SX_HIP (hipGraphLaunch (b->exec, st));           // warm-up
    SX_HIP (hipEventRecord (e0, st));
    for (int r = 0; r < reps; r++) SX_HIP (hipGraphLaunch (b->exec, st));
    SX_HIP (hipEventRecord (e1, st));
    SX_HIP (hipEventSynchronize (e1));
    float f = 0.0f;
    SX_HIP (hipEventElapsedTime (&f, e0, e1));
    ms = f / reps;
    return SX_OK;
  };
  double best = 0.0;
  int rc = time_ms (best);
  // coordinate search: each class's right-hand sides per sweep workgroup
  for (size_t c = 0; !rc && c < b->keys.size (); c++) {
    const int keep = b->per[c];
    for (int cand : {1, 2}) {
      b->per[c] = cand;
      double t = 0.0;
      if ((rc = fused_record (b)) || (rc = time_ms (t))) break;
      if (t < 0.98 * best) { best = t; b->per[c] = cand; goto kept; }
      b->per[c] = keep;
    }
    if (!rc) rc = fused_record (b);
  kept:;
  }
  if (!rc) rc = fused_record (b);
  (void) hipEventDestroy (e0);
  (void) hipEventDestroy (e1);
  return rc;
}

extern "C" int sx_batch_info (sx_batch_t b, char* buf, int len)
{
  if (!b || !buf || len < 1) { set_error ("sx_batch_info: bad arguments"); return SX_ERR_ARG; }
  std::string t = "fused=" + std::to_string (b->fused) + " launches=" +
                  std::to_string (b->fused ? b->launches : -1);
  for (size_t c = 0; c < b->keys.size (); c++) {
    int cnt = 0;
    for (int x : b->cls) cnt += x == (int) c;
    t += "; D=" + std::to_string (b->keys[c][0]) + " unit=" + std::to_string (b->keys[c][1]) +
         " twisted=" + std::to_string (b->keys[c][2]) + " problems=" + std::to_string (cnt) +
         " nr=" + std::to_string (b->per[c]);
  }
  std::snprintf (buf, (size_t) len, "%s", t.c_str ());
  return SX_OK;
}
