// sx_internal.h -- shared declarations of the semtex_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/semtex_amd.h"

namespace sx {

// Per-node geometric arrays kept per element on the device, in this order
// (Element::mapping, src/element.cpp:1253-1408).
enum { G_Q1 = 0, G_Q2, G_Q3, G_Q4, G_Q8, G_RX, G_RY, G_SX, G_SY, G_Y, NGEO };

void set_error (const std::string& msg);
int  hip_fail (hipError_t e, const char* what);

#define SX_HIP(call)                                                  \
  do {                                                                \
    hipError_t _e = (call);                                           \
    if (_e != hipSuccess) return sx::hip_fail (_e, #call);            \
  } while (0)

// GLL rule (femlib/polylib.c:312 zwglj, :513 Dglj), host.
void gll (int np, double* z, double* w, double* D);
// edge-major permutation (femlib/mapping.c:46-132), host.
void edgemaps (int np, int* emap, int* pmap);
// geometric factors of one element (src/element.cpp:1253-1408), host.
// geo: NGEO*np*np.  Returns false if the Jacobian is not positive.
bool element_geometry (int np, int cyl, const double* x, const double* y,
                       double* geo);
// side quadrature weights |dx/dr| w of side `side` in CCW order
// (src/element.cpp:1625-1742 sideGeom).
void side_area (int np, const double* x, const double* y, int side,
                double* area);

}  // namespace sx

struct sx_msys;
namespace sx {
// Banded FP64 Cholesky of every mode's band, the offset-major sweep copies
// of the factor and of its reversal (with inverted diagonal blocks), and
// the two-sweep solve of the right-hand sides in d_rhs ([mode][2][nmax])
// -- sx_band.hip.
int    band_block (int nband);
size_t band_sweep_bytes_per_mode (int nmax, int nband);
int    band_factor (sx_msys* s, hipStream_t st, int* d_status);
// nr RHS per mode; moves (twisted systems): 1 gather the bottom piece's
// right-hand side, 2 scatter its solution (3: both; 0 when the caller's
// assembly and recovery address the bottom piece themselves)
int    band_solve (const sx_msys* s, int nr, hipStream_t st, int moves = 3);
// per-mode split of a twisted system (d_tw): TW ints per mode
constexpr int TW = 8;
enum { TW_NBT = 0, TW_NBB, TW_PAD, TW_M, TW_NTOP, TW_NBOT, TW_N };
// SX_OK, or SX_ERR_SYNC when a split-sweep wait gave up since the last call
// (synchronises st; clears the status word)
int    split_status (sx_msys* s, hipStream_t st);
// Grow the per-mode right-hand-side workspace to nr (2, 4 or 6) columns.
int    msys_reserve_rhs (sx_msys* s, int nr);
// Field::solve for nf fields sharing one system (the DNS velocity
// components): one band launch for all 2*nf right-hand sides per mode.
int    field_solve_group (sx_msys* s, int nf, int nz, int stride, double* const* d_f,
                          double* const* d_u, const double* const* d_lines, hipStream_t st);
// Real z transforms of nz planes of stride S in the reference's packing
// (sx_fourier.hip); in place when fourier_in_place (nz).
int    fourier_z (int nz, int S, const double* in, double* out, int sign, hipStream_t st);
int    fourier_prepare (int nz);   // builds the per-device tables (synchronous)
// Fused z transforms (nz 32, 64, 128, 256: fourier_fused_ok):
//   sign +1: out = scale (FWD(in .* mul) - dz-derivative of acc) + add0 on
//            plane 0 (mul, acc may be null; acc packed spectra, dz = beta)
//   sign -1: out = acc - amul .* INV(dz-derivative of in)   (dz = beta: multiply
//            by i beta k first, 0: none; acc null: out = INV(..); amul null: 1)
bool   fourier_fused_ok (int nz);
int    fourier_z_fused (int nz, int S, const double* in, const double* mul, double* out,
                        int sign, double dz, const double* acc, const double* amul,
                        double scale, double add0, hipStream_t st);
// nb (1..3) fused transforms of one shape in one launch: array arguments
// per transform (null arrays: no mul / acc / amul; scale, add0 null: 1, 0)
int    fourier_z_fused_batch (int nz, int S, int nb, const double* const* in,
                              const double* const* mul, double* const* out, int sign, double dz,
                              const double* const* acc, const double* const* amul,
                              const double* scale, const double* add0, hipStream_t st);
// Mixed-precision solve with a fixed number of refinement passes, no host
// synchronisation (sx_mixed.hip).
int    field_solve_mixed_passes (sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                                 const double* d_lines, int passes, hipStream_t st);
// The tolerance-driven mixed solve (sx_field_solve_mixed) on a workspace of
// at least mixed_work_doubles() doubles that the caller provides; the
// argument check shared by its entry points.
constexpr int MIXED_NBLK = 256;     // workgroups of the correction-norm reduction
size_t mixed_work_doubles (const sx_msys* s, int nz, int stride);
int    field_solve_mixed_ws (sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                             const double* d_lines, double refine_tol, int max_iter,
                             int* iters, double* work, hipStream_t st);
int    check_mixed_solve_args (sx_msys* s, int nz, int stride, const void* f, const void* u,
                               int max_iter, const char* who);
// The mixed solve's FP32 element passes on MFMA tiles (systems with
// d_Mrec32): X (weak, row-major) -> Fp [condensed F_b | F_i] in emap order;
// Fp + band solution (d_rhs, 2 per mode) -> V row-major.
int    condense32_tiled (const sx_msys* s, int nz, int stride, const double* X, double* Fp,
                         hipStream_t st);
int    recover32_tiled (const sx_msys* s, int nz, int stride, const double* Fp, double* V,
                        hipStream_t st);
// The DIRECT solve of sx_field_solve with the workspace `s` names (d_G,
// d_rhs, d_rhsb; split sweeps unless no_split) -- sx_solver.hip; and the
// argument check of the DIRECT entry points.
int    field_solve_direct (const sx_msys* s, int nz, int stride, double* d_f, double* d_u,
                           const double* d_lines, hipStream_t st, bool timed);
int    check_direct_solve_args (sx_msys* s, int nz, int stride, const void* f, const void* u);
// JACPCG set-up: the Jacobi preconditioner of every mode (sx_pcg.hip).
int    pcg_setup (sx_msys* s, hipStream_t st);
inline bool fourier_in_place (int nz) { return nz >= 8 && nz <= 1024 && (nz & (nz - 1)) == 0; }
}  // namespace sx

struct sx_mesh {
  int np = 0, nel = 0, cyl = 0, npnp = 0, next = 0, nint = 0;
  int nglobal = 0, nsolve = 0, nband = 0, kd = 0, nbound = 0, nclass = 0;
  int npairs = 0;
  std::vector<int> btog, bmask, emask, cls, emap;
  std::vector<double> hx, hy;  // host node coordinates [nel][npnp]
  std::vector<int> bnd_elmt, bnd_side, bnd_kind;

  // device arrays
  double* d_geo = nullptr;     // [nel][NGEO][npnp]
  double* d_cgeo = nullptr;    // [nclass][NGEO][npnp] class representatives
  double* d_D = nullptr;       // [np][np]
  int* d_cls = nullptr;        // [nel]
  int* d_btog = nullptr;       // [nel][next]
  int* d_bmask = nullptr;      // [nel][next]
  int* d_emask = nullptr;      // [nel]
  int* d_emap = nullptr;       // [npnp]
  int* d_gptr = nullptr;       // [nglobal+1] CSR: global node -> plane slots
  int* d_gsrc = nullptr;       //   slot = e*npnp + i (emap-ordered storage)
  int* d_gpos = nullptr;       //   the slot's row-major plane offset e*npnp + emap[i]
  double* d_gq4 = nullptr;     //   and its Q4 weight
  int* d_ess = nullptr;        // [nglobal] line index of essential value / -1
  int* d_nptr = nullptr;       // [nglobal+1] CSR natural-BC contributions
  int* d_nidx = nullptr;       //   line index
  double* d_nw = nullptr;      //   quadrature weight
  long long* d_ppos = nullptr; // [npairs] band slot of assembled entry
  int* d_pmn = nullptr;        // [npairs][2] (m, n) global rows
  int* d_pptr = nullptr;       // [npairs+1] CSR: entry -> condensed-matrix slots
  int* d_psrc = nullptr;       //   slot = cls*next*next + i*next + k
  // element tiles for the batched condensation / recovery: elements sorted
  // by class, tiles of at most ETILE elements of one class
  int ntile = 0, ness = 0;
  int* d_tile = nullptr;       // [ntile][3] class, first index into d_elist, count
  int* d_elist = nullptr;      // [nel] element numbers, class-sorted
  int* d_ess_elem = nullptr;   // [ness] elements with essential nodes (emask)
  int* d_ess_slot = nullptr;   // [ness][next] their slots' line index (d_ess) / -1
  // per-tile tables of the register-operand tile kernels, read by block
  // index with no prior load: [ntile][tile_ints (next)], per tile the class,
  // the element count, 2 unused, the ETILE element numbers (the last one
  // repeated past the count), their emask, then their btog and bmask rows
  int* d_ttab = nullptr;
};
constexpr int ETILE = 16;
constexpr int TT_EL = 4, TT_EM = TT_EL + ETILE, TT_BT = TT_EM + ETILE;
__host__ __device__ constexpr int tile_ints (int next) { return TT_BT + 2 * ETILE * next; }

struct sx_msys {
  sx_mesh* m = nullptr;
  int nmodes = 0, base_mode = 0, precision = SX_FP64;
  double lambda2 = 0.0, beta = 0.0;
  std::vector<double> betak2;
  std::vector<int> nsolve;
  int nmax = 0;                // max nsolve over modes
  size_t band_stride = 0;      // doubles per mode band
  double* d_betak2 = nullptr;  // [nmodes]
  int* d_nsolve = nullptr;     // [nmodes]
  double* d_Bm = nullptr;      // [nmodes][nclass][next][nint]  hbi*inv(hii)
  double* d_Hinv = nullptr;    // [nmodes][nclass][nint][nint]  inv(hii)
  double* d_Mrec = nullptr;    // [nmodes][nclass][npnp][pad4 nint] [-Bm; Hinv] (recovery)
  double* d_Mcon = nullptr;    // [nmodes][nclass][nint][pad4 next] Bm^T (condensation)
  int tile_mf = 0;             // 1: tiled kernels on FP64 MFMA (operands padded to 16)
  float* d_Mrec32 = nullptr;   // SX_MIXED with tile_mf: FP32 copies of d_Mrec, d_Mcon
  float* d_Mcon32 = nullptr;
  float* d_Bm32 = nullptr;     // FP32 copies for SX_MIXED
  float* d_Hinv32 = nullptr;
  double* d_bandU = nullptr;   // [nmodes][band_stride] 'U' band, ld = nband
                               // (assembly + factorisation; freed after set-up)
  int sw_block = 0;            // sweep block size B
  int sw_panel = 0;            // 1: right-looking panel sweeps (sw_D off-diagonal blocks)
  int sw_D = 0;
  int sw_unit = 0;             // 1: off-diagonal panels scaled to W = T_J U (unit diagonal blocks)
  char* d_xch = nullptr;       // split sweeps: per-chain message rings
  unsigned* d_xep = nullptr;   //   [4] epoch, arrivals, status (bit 0: a wait gave up)
  size_t xch_bytes = 0;
  // the panel sweep kernel of the last band solve (diagnostic): 0 none,
  // 1 k_panel_sweeps, 2 k_unit_sweeps, 3 k_split_sweeps; right-hand sides
  // per workgroup
  mutable int last_kernel = 0, last_nr = 0;
  // 1: solves may run concurrently with other kernels (a batch graph's
  // branches), so the split sweeps, which need their whole grid resident,
  // are not used
  int no_split = 0;
  size_t sw_stride = 0;        // doubles per mode of each sweep copy
  double* d_swF = nullptr;     // [nmodes][nblk][nband][B] forward sweep blocks
  double* d_swR = nullptr;     // [nmodes][nblk][nband][B] reversed (backward) blocks
  int tw = 0;                  // 1: twisted two-piece panel solve (sx_band.hip)
  int tw_S = 0;                //   separator rows, 64 sw_D
  size_t tw_boff = 0;          //   doubles from a mode's top-piece panels to its bottom ones
  int* d_tw = nullptr;         //   [nmodes][8] split of every mode
  double* d_rhsb = nullptr;    //   [nmodes][nr][nmaxb] bottom-piece right-hand sides
  int nmaxb = 0;
  double* d_rhs = nullptr;     // [nmodes][nr][nmax] solve workspace (nr <= rhs_cap)
  double* d_G = nullptr;       // [fields][2*nmodes][nglobal] essential values
  int rhs_cap = 2;             // right-hand sides per mode d_rhs / d_G can hold
  double* d_work = nullptr;    // mixed-precision / PCG workspace
  double* d_pc = nullptr;      // SX_JACPCG: [nmodes][npts] inverse assembled diagonal
  size_t work_doubles = 0;
};

// ---------------------------------------------------------------------------
// Fused batches (sx_batch.hip).  A batch of independent solves is planned
// instead of launched: every launch of every solve is recorded with its
// phase (its place in that solve's dependency order) and merged with the
// launches of the other solves that run the same kernel in the same phase.
// A merged launch runs k_batch: blockIdx.z picks the solve, whose recorded
// arguments (a Pack) and grid it reads; workgroups past that solve's own
// grid return at once, and the kernel body sees its usual blockIdx.x/.y.
// ---------------------------------------------------------------------------
#include <cstring>
#include <utility>

namespace sx {

template <typename... A> struct Pack;
template <> struct Pack<> {};
template <typename T, typename... R> struct Pack<T, R...> {
  T h;
  Pack<R...> t;
};

// The host writes Packs that the device reads: both compilation passes
// must lay them out alike (checked here for the member kinds used).
static_assert (sizeof (Pack<int, const int*, double>) == 32 &&
               sizeof (Pack<double*, size_t, int, int>) == 32 &&
               sizeof (Pack<int, double, const double*>) == 32 &&
               sizeof (Pack<int, int>) == 12 && alignof (Pack<int, double*>) == 8,
               "argument packs must have one layout on host and device");

inline Pack<> make_pack () { return {}; }
template <typename T, typename... R> Pack<T, R...> make_pack (T h, R... r)
{
  return Pack<T, R...> {h, make_pack (r...)};
}

template <int I, typename T, typename... R>
__host__ __device__ __forceinline__ const auto& pack_get (const Pack<T, R...>& p)
{
  if constexpr (I == 0) return p.h;
  else return pack_get<I - 1> (p.t);
}

template <typename Fn, typename... A, int... I>
__device__ __forceinline__ void pack_call (const Pack<A...>& p, std::integer_sequence<int, I...>)
{
  Fn {} (pack_get<I> (p)...);
}

template <int LB, typename Fn, typename... A>
__global__ __launch_bounds__(LB) void k_batch (const Pack<A...>* __restrict__ args,
                                               const uint2* __restrict__ grid)
{
  const uint2 g = grid[blockIdx.z];
  if (blockIdx.x >= g.x || blockIdx.y >= g.y) return;
  pack_call<Fn> (args[blockIdx.z], std::make_integer_sequence<int, sizeof... (A)> {});
}

typedef int (*BatchLauncher) (const void* d_args, const uint2* d_grid, int n, unsigned gx,
                              unsigned gy, int block, size_t lds, hipStream_t st);

template <int LB, typename Fn, typename... A>
int batch_launcher (const void* d_args, const uint2* d_grid, int n, unsigned gx, unsigned gy,
                    int block, size_t lds, hipStream_t st)
{
  auto k = k_batch<LB, Fn, A...>;
  if (lds > 0)
    SX_HIP (hipFuncSetAttribute ((const void*) k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) lds));
  hipLaunchKernelGGL (k, dim3 (gx, gy, (unsigned) n), dim3 (block), lds, st,
                      (const Pack<A...>*) d_args, d_grid);
  SX_HIP (hipGetLastError ());
  return SX_OK;
}

// The launches of a batch, merged by (phase, kernel instance).
struct BatchPlan {
  struct Group {
    int phase = 0;
    BatchLauncher launch = nullptr;
    int block = 0;
    size_t lds = 0, psize = 0;
    std::vector<char> blob;          // the solves' argument Packs, psize bytes each
    std::vector<uint2> grid;         // the solves' own grids
    unsigned gx = 0, gy = 0;         // their maximum
    void* d_args = nullptr;
    uint2* d_grid = nullptr;
  };
  std::vector<Group> groups;
  int phase = 0;                     // the phase the next recorded launch belongs to
  int tune = 0;                      // right-hand sides per sweep workgroup (0: default)
};

template <int LB, typename Fn, typename... A>
void batch_rec (BatchPlan& P, dim3 grid, int block, size_t lds, A... a)
{
  const BatchLauncher L = &batch_launcher<LB, Fn, A...>;
  const Pack<A...> pk = make_pack (a...);
  BatchPlan::Group* g = nullptr;
  for (auto& x : P.groups)
    if (x.phase == P.phase && x.launch == L && x.block == block) { g = &x; break; }
  if (!g) {
    P.groups.emplace_back ();
    g = &P.groups.back ();
    g->phase = P.phase;
    g->launch = L;
    g->block = block;
    g->psize = sizeof (pk);
  }
  const size_t o = g->blob.size ();
  g->blob.resize (o + sizeof (pk));
  std::memcpy (g->blob.data () + o, &pk, sizeof (pk));
  g->grid.push_back (make_uint2 (grid.x, grid.y));
  g->gx = std::max (g->gx, grid.x);
  g->gy = std::max (g->gy, grid.y);
  g->lds = std::max (g->lds, lds);
}

// Plans of the solve paths (each starts at P.phase and leaves P.phase at
// the phase after its last launch); SX_ERR_LIMIT when a system needs a
// kernel the fused batch does not merge (the caller then records launches).
int plan_field_solve_direct (BatchPlan& P, const sx_msys* s, int nz, int stride, double* d_f,
                             double* d_u, const double* d_lines);
int plan_band_solve (BatchPlan& P, const sx_msys* s, int nr);
int plan_mixed_passes (BatchPlan& P, sx_msys* s, int nz, int stride, double* d_f,
                       double* d_u, const double* d_lines, int passes);
int plan_copy (BatchPlan& P, double* dst, const double* src, size_t n);
int plan_condense32_tiled (BatchPlan& P, const sx_msys* s, int nz, int stride, const double* X,
                           double* Fp);
int plan_recover32_tiled (BatchPlan& P, const sx_msys* s, int nz, int stride, const double* Fp,
                          double* V);

}  // namespace sx
