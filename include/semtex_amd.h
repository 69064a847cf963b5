/*
 * semtex_amd.h -- C ABI of the MI355X (gfx950) per-Fourier-mode Helmholtz
 * solver that stands in for semtex's DIRECT solve path.
 *
 * All entry points are extern "C", take plain pointers and sizes, and return
 * an int status (SX_OK = 0, negative on error; sx_last_error() gives the
 * message).  Device pointers are HIP device addresses in HBM; `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Storage conventions are the reference's:
 *   - a Fourier field is [nz][plane_stride] doubles, plane k holding mode
 *     k>>1 (real part for even k, imaginary for odd k; plane 1 of mode 0 is
 *     the Nyquist plane and is never solved) -- src/field.cpp:648-655;
 *   - within a plane, element e occupies [e*np*np, (e+1)*np*np), row-major
 *     np x np (src/element.h, src/auxfield.cpp:1062);
 *   - btog/bmask are [nel][4(np-1)] in emap order, essential nodes numbered
 *     last (src/numbersys.h);
 *   - boundary line data are [nz][nbound*np]: segment b's np values in CCW
 *     traverse order (src/field.cpp:1078-1102, src/condition.cpp:80).
 */
#ifndef SEMTEX_AMD_H
#define SEMTEX_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SX_OK            0
#define SX_ERR_ARG      -1   /* bad argument / size                        */
#define SX_ERR_HIP      -2   /* HIP runtime error                          */
#define SX_ERR_GEOM     -3   /* non-positive element Jacobian              */
#define SX_ERR_FACTOR   -4   /* Cholesky breakdown (matrix not SPD)        */
#define SX_ERR_LIMIT    -5   /* problem exceeds an implemented size limit  */
#define SX_ERR_SYNC     -6   /* a split band sweep's bounded wait gave up:  */
                             /* the affected solutions hold NaN            */

#define SX_ESSENTIAL 0       /* boundary kinds (src/condition.h)           */
#define SX_NATURAL   1
#define SX_AXIS      2       /* cylindrical axis, natural with no flux      */
                             /* (class Axis, src/condition.cpp): keeps its  */
                             /* place in the boundary-line layout           */

#define SX_FP64  0           /* solver precisions                          */
#define SX_MIXED 1           /* FP32 elemental applies + FP64 refinement   */
#define SX_JACPCG 2          /* Jacobi-preconditioned CG (SolverKind JACPCG) */

typedef struct sx_mesh*  sx_mesh_t;   /* geometry + numbering, on the GPU  */
typedef struct sx_msys*  sx_msys_t;   /* factorised per-mode systems       */

/* Last error message of this thread ("" if none). */
const char* sx_last_error (void);

/* Fills `buf` with the device name / arch; returns SX_OK when a gfx950
 * device is present. */
int sx_device_info (char* buf, int len);

/* GLL points, weights, derivative matrix D[i*np+j] = h_j'(z_i) as the
 * library computes them (femlib/polylib.c:312 zwglj, :513 Dglj).  Host
 * arrays, no GPU needed. */
int sx_gll (int np, double* z, double* w, double* D);

/* Geometry, numbering and boundary system of one 2D mesh, replacing the
 * Element vector + NumberSys + BoundarySys that ModalMatrixSys receives
 * (src/matrix.h:61-66; Element::mapping src/element.cpp:1253-1408;
 * NumberSys src/numbersys.h:4-80; BoundarySys src/bsys.cpp:42-157).
 * x, y: [nel][np*np] node coordinates (Element::_xmesh/_ymesh).
 * cyl: 1 for cylindrical (y = r) coordinates.  All arrays are host. */
int sx_mesh_create (int np, int nel, int cyl,
                    const double* x, const double* y,
                    const int* btog, const int* bmask, const int* emask,
                    int nglobal, int nsolve, int nband,
                    int nbound, const int* bnd_elmt, const int* bnd_side,
                    const int* bnd_kind, sx_mesh_t* out);
int sx_mesh_destroy (sx_mesh_t m);
/* Number of distinct element geometries found (Family sharing,
 * src/family.cpp). */
int sx_mesh_nclass (sx_mesh_t m);

/* ModalMatrixSys::ModalMatrixSys (src/matrix.cpp:33-105) with method
 * DIRECT: builds, statically condenses (Element::HelmholtzSC
 * src/element.cpp:131-214) and Cholesky-factors (MatrixSys::MatrixSys
 * src/matrix.cpp:185-267) the system of every mode base_mode ..
 * base_mode+num_modes-1, with betak2 = (beta*mode)^2
 * (Field::modeConstant, Cartesian), all on the GPU.  A mode whose
 * lambda2+betak2 vanishes with no essential nodes is singular and drops its
 * last unknown (src/matrix.cpp:154-162).  precision: SX_FP64 or SX_MIXED
 * (DIRECT), or SX_JACPCG: MatrixSys with method JACPCG
 * (src/matrix.cpp:269-310), which holds only the Jacobi preconditioner
 * (inverse of the assembled diagonal, Element::HelmholtzDiag
 * src/element.cpp:1484-1533) and is solved with sx_field_solve_pcg. */
int sx_msys_create (sx_mesh_t m, double lambda2, double beta, int base_mode,
                    int num_modes, int precision, void* stream,
                    sx_msys_t* out);
/* The same for the field `name` ('u', 'v', 'w', 'p' or 'c'), whose Fourier
 * constant is Field::modeConstant (src/field.cpp:1204-1229): beta*mode,
 * except on a cylindrical mesh (cyl = 1) for the coupled radial and
 * azimuthal velocities v~ = v + i w and w~ = v - i w, which solve with
 * beta*mode + 1 ('v') and beta*mode - 1 ('w'), and with 1 for mode 0.
 * sx_msys_create is this with name 'c'.  As in ModalMatrixSys over a
 * BoundarySys (src/bsys.cpp:42-123), a cylindrical 3D field whose axis
 * condition changes with the mode (BCmgr::getCondition,
 * src/bcmgr.cpp:478-530: u, p, c natural on the axis for mode 0 and
 * essential above; v~ essential; w~ natural for mode 1 only) takes one
 * mesh (numbering) per mode class and one system per class: base_mode 0,
 * 1 and 2 with 1, 1 and num_modes-2 modes, solved on the matching planes
 * ([0,2), [2,4), [4,nz)) of the field. */
int sx_msys_create_named (sx_mesh_t m, double lambda2, double beta, int base_mode,
                          int num_modes, char name, int precision, void* stream,
                          sx_msys_t* out);
int sx_msys_destroy (sx_msys_t s);

/* AuxField::couple (src/auxfield.cpp:1376-1451) and the FORWARD branch of
 * Field::coupleBCs (src/field.cpp:1127-1201) on nz planes of n values at
 * stride plane_stride (fields: n = nel*np*np; boundary lines: n =
 * plane_stride = nbound*np), modes base_mode+k for planes 2k, 2k+1:
 *   dir SX_FORWARD (+1):  v~ = v + i w,  w~ = v - i w
 *   dir SX_INVERSE (-1):  v = (v~ + w~)/2,  w = i (w~ - v~)/2
 * in place in d_v and d_w; mode 0 (base_mode 0) and nz = 1 are left
 * unchanged.  (The reference's Field::coupleBCs INVERSE branch, never
 * called, is not the inverse of its FORWARD branch and is not provided.) */
#define SX_FORWARD  1
#define SX_INVERSE -1
int sx_couple (int nz, int plane_stride, int n, double* d_v, double* d_w, int dir,
               int base_mode, void* stream);

/* Status of the solves issued on this system since the last query:
 * synchronises `stream`, returns SX_OK, or SX_ERR_SYNC (and clears the
 * status) when a wait of the split band sweeps gave up.  The split sweeps
 * run each chain of block steps on two workgroups that exchange partial
 * sums and need both resident at once; the library launches them only when
 * the whole grid fits on the device, and if another process's kernels keep
 * a partner workgroup off the device past the wait bound, that solve's
 * results are NaN instead of the kernel hanging.  sx_field_solve and
 * sx_field_solve_mixed run asynchronously on their stream, so a caller that
 * keeps the reference's message(ERROR) behaviour (veclib/util.c:24-46) calls
 * this after the stream work it depends on; sx_field_solve_host checks it
 * itself.  No reference counterpart (the CPU band solve cannot fail this
 * way). */
int sx_msys_status (sx_msys_t s, void* stream);
/* Unknowns solved for mode index `mode` (0-based within the system). */
int sx_msys_nsolve (sx_msys_t s, int mode);
/* Layout of the factor's solve copies (diagnostic; no reference
 * counterpart): info[0] 1 = 64x64 panel sweeps, 0 = offset-major blocks;
 * info[1] off-diagonal panel blocks D; info[2] 1 = twisted two-piece solve;
 * info[3] band width nband; info[4] 1 = unit-diagonal panels (W = T_J U);
 * info[5] kernel of the last band solve: 0 none yet, 1 panel sweeps with
 * the triangle in the chain, 2 unit-diagonal sweeps, 3 split sweeps (two
 * workgroups per chain); info[6] its right-hand sides per workgroup.
 * Fills min(n, 7) entries. */
int sx_msys_layout (sx_msys_t s, int* info, int n);

/* Field::solve (src/field.cpp:598-709, DIRECT branch) over nz planes.
 * d_f:     forcing planes [nz][plane_stride] (device; DESTROYED, as in the
 *          reference, src/field.cpp:611).
 * d_u:     output planes [nz][plane_stride] (device).
 * d_lines: boundary values [nz][nbound*np] (device) or NULL for all zero.
 * Planes k >= 2*num_modes are rejected; the Nyquist plane 1 of base mode 0
 * is left untouched.  For SX_MIXED systems, refine_tol is the relative
 * residual target of the FP64 refinement (<= 0: 1e-12) and max_iter its
 * iteration cap; *iters (may be NULL) returns the iterations used. */
int sx_field_solve (sx_msys_t s, int nz, int plane_stride, double* d_f,
                    double* d_u, const double* d_lines, void* stream);
int sx_field_solve_mixed (sx_msys_t s, int nz, int plane_stride, double* d_f,
                          double* d_u, const double* d_lines,
                          double refine_tol, int max_iter, int* iters,
                          void* stream);

/* sx_field_solve_mixed without device allocation.  No reference
 * counterpart: for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, same
 * kernels and the same results; the workspace (5*nz*plane_stride + 1024
 * doubles) lives in a fixed device array of SX_NOALLOC_DOUBLES doubles
 * (80 MiB) that each GPU holds until the process ends, instead of one the
 * system allocates.  Larger nz*plane_stride is rejected with SX_ERR_LIMIT.
 * All calls on one GPU share that array: calls that overlap on one GPU
 * (two streams or two host threads) race. */
#define SX_NOALLOC_DOUBLES (10 * 1024 * 1024)
int sx_field_solve_mixed_noalloc (sx_msys_t s, int nz, int plane_stride, double* d_f,
                                  double* d_u, const double* d_lines,
                                  double refine_tol, int max_iter, int* iters,
                                  void* stream);

/* Field::solve, JACPCG branch (src/field.cpp:711-805), every solved plane
 * at once, each with its own stopping test: d_u holds the initial guess on
 * entry (the field's current values, as in the reference) and the solution
 * on exit; iterations stop when |r|^2 <= (tol_rel |b|)^2 (TOL_REL) or the
 * iteration index reaches step_max (STEP_MAX).  *iters (may be NULL)
 * returns the largest iteration count over the planes. */
int sx_field_solve_pcg (sx_msys_t s, int nz, int plane_stride, double* d_f, double* d_u,
                        const double* d_lines, double tol_rel, int step_max, int* iters,
                        void* stream);

/* A batch of independent problems (each system with its own mesh, order,
 * lambda2, Cartesian or cylindrical geometry) solved by one captured HIP
 * graph.  By default the graph is fused: every kernel instance runs once
 * per phase for all the problems that use it (a few dozen launches for
 * the whole batch); SX_BATCH_FUSED=0 in the environment, or a system the
 * fused plans do not cover (offset-major band sweeps for band offsets past
 * 511, untiled element kernels for N_P > 12), records each problem's own
 * launches instead.  For problem i the graph copies d_f_src[i] into d_f_work[i] (the
 * solve destroys its forcing) and runs the solve into d_u[i], with
 * d_lines[i] (d_lines or an entry may be NULL).  SX_MIXED systems run
 * refine_passes passes (the first solve plus refine_passes-1 FP64
 * residual corrections) without host synchronisation; SX_FP64 systems run
 * sx_field_solve.  The buffers are baked into the graph: refill
 * d_f_src / d_lines in place between runs.  sx_batch_run launches the
 * whole batch on `stream`.  Problems of different systems run concurrently
 * (branches of the graph over 4 streams), so their d_f_work and d_u
 * buffers must not overlap; problems that share a system run in order.
 * The recorded solves use no split band sweeps (they need their whole grid
 * resident at once); sx_field_solve on the same systems outside the batch
 * keeps them.  No reference counterpart: for callers that
 * solve many independent small problems (uneven batches). */
typedef struct sx_batch* sx_batch_t;
int sx_batch_create (int n, const sx_msys_t* systems, const int* nz, const int* plane_stride,
                     const double* const* d_f_src, double* const* d_f_work, double* const* d_u,
                     const double* const* d_lines, int refine_passes, sx_batch_t* out);
int sx_batch_run (sx_batch_t b, void* stream);
int sx_batch_destroy (sx_batch_t b);
/* Autotuner of a fused batch: for every class of band system in it (off-
 * diagonal blocks D, unit-diagonal panels, twisted layout) tries 1 and 2
 * right-hand sides per band-sweep workgroup, times `reps` replays of the
 * whole batch on `stream` for each, keeps what is fastest and re-records
 * the graph with it.  No-op for the graph-of-launches form. */
int sx_batch_tune (sx_batch_t b, int reps, void* stream);
/* The batch's form and choices as text: "fused=1 launches=<merged launches
 * in the graph>; D=.. unit=.. twisted=.. problems=.. nr=..; ..." (nr 0:
 * the default choice). */
int sx_batch_info (sx_batch_t b, char* buf, int len);

/* Adjoint building blocks.  The DIRECT system is symmetric, so the adjoint
 * of Field::solve is a solve with a weak-form right-hand side:
 * sx_field_solve_weak solves A u = L^T w for element-local w (the direct
 * stiffness sum of w, no mass weighting, zero essential values; d_w is
 * destroyed).  sx_mass_inner returns per plane k (device d_out[nz]) the
 * element-quadrature inner product sum_e sum_ij Q4 a b (AuxField::integral
 * weights); sx_mass_weight writes d_out = scale * Q8 * d_in element by
 * element (Element::weight).  With J(u), lam = solve_weak(dJ/du):
 * dJ/df = -Q8 lam, dJ/dlambda2 = -sum_k inner(lam, u).
 * No reference counterpart: semtex has no adjoint of its solve. */
int sx_field_solve_weak (sx_msys_t s, int nz, int plane_stride, double* d_w, double* d_u,
                         void* stream);
int sx_mass_inner (sx_mesh_t m, int nz, int plane_stride, const double* d_a,
                   const double* d_b, double* d_out, void* stream);
int sx_mass_weight (sx_mesh_t m, int nz, int plane_stride, const double* d_in, double* d_out,
                    double scale, void* stream);

/* Element-local operators of the DNS step and their transposes, for the
 * hand-written adjoint of the step (no reference counterpart).  Out of
 * place on nz planes [nz][plane_stride]; transpose = 1 applies the
 * transpose (adjoint) of the operator.
 *   sx_grad_planes:   d/dx (dir 0) or d/dy (dir 1) per element
 *                     (Element::grad, src/element.cpp:521-565);
 *   sx_smooth_planes: Field::smooth (src/field.cpp:282-327), the
 *                     mass-weighted average on element-boundary nodes;
 *   sx_dz_planes:     d/dz of Fourier planes (i beta k; planes 0, 1 zero;
 *                     AuxField::gradient direction 2), n values per plane. */
int sx_grad_planes (sx_mesh_t m, int nz, int plane_stride, const double* d_in, double* d_out,
                    int dir, int transpose, void* stream);
int sx_smooth_planes (sx_mesh_t m, int nz, int plane_stride, const double* d_in, double* d_out,
                      int transpose, void* stream);
int sx_dz_planes (int nz, int plane_stride, int n, double beta, const double* d_in,
                  double* d_out, int transpose, void* stream);

/* Host-pointer convenience wrapper for drop-in callers whose Field data
 * lives in host memory (Domain::u, src/domain.cpp:75-79): copies in,
 * solves, copies u back (f is not copied back). */
int sx_field_solve_host (sx_msys_t s, int nz, int plane_stride,
                         const double* f, double* u, const double* lines);

/* Elemental Helmholtz operator applied plane by plane, u -> H u for every
 * element (Element::HelmholtzOp src/element.cpp:1593-1622), mode-dependent
 * betak2 as in sx_msys_create.  Device planes [nz][plane_stride]. */
int sx_helmholtz_op (sx_msys_t s, int nz, int plane_stride, const double* d_u,
                     double* d_hu, void* stream);

/* Real transforms along z of every line of nz planes of stride
 * plane_stride (AuxField::transform src/auxfield.cpp:1157-1194 -> dDFTr
 * femlib/fourier.c:40-165): sign +1 physical -> Fourier, packed
 * [Re X0, Re X_nz/2, Re X1, Im X1, ...] and scaled 1/nz; sign -1 the
 * inverse, unscaled.  nz == 1 or even.  In place (d_in == d_out) when nz is
 * a power of two in 8..1024; other nz need distinct buffers. */
int sx_fourier_transform (int nz, int plane_stride, const double* d_in, double* d_out,
                          int sign, void* stream);

/* ------------------------------------------------------------------------
 * DNS time step: integrate() (dns/integrate.cpp:66-201) for Cartesian
 * 2D (nz == 1: u, v, p) or 2.5D (nz even >= 2: u, v, w, p) flows.
 * mesh_u: velocity numbering (its essential segments carry the velocity
 * boundary values of sx_dns_set_velocity_bc, zero by default: no-slip
 * walls), mesh_p: pressure numbering, same boundary segments in the same
 * order; natural segments of mesh_p carry the computed high-order
 * pressure condition (BCmgr::maintainFourier/evaluateCNBCp,
 * src/bcmgr.cpp:1037-1335).  advection: SX_SKEW (skewSymmetric, "-S") or
 * SX_ALTSKEW (altSkewSymmetric, the driver default), dns/nonlinear.cpp.
 * const_force: ConstForce CONST_X/Y/Z added to mode 0 (may be NULL).
 * nord: N_TIME, 1..3 (Integration::StifflyStable/Extrapolation,
 * src/integration.cpp:99-164).  Step s < nord runs at order s: by default
 * with the direct solve of that order's system (the limit of the
 * reference's start-up Jacobi-PCG, dns/integrate.cpp:340-353), or with
 * JACPCG after sx_dns_set_startup.
 * ---------------------------------------------------------------------- */
#define SX_SKEW    0
#define SX_ALTSKEW 1
typedef struct sx_dns* sx_dns_t;
int sx_dns_create (sx_mesh_t mesh_u, sx_mesh_t mesh_p, int nz, double beta,
                   double kinvis, double dt, const double* const_force,
                   int advection, int nord, void* stream, sx_dns_t* out);
int sx_dns_destroy (sx_dns_t d);
/* Domain::restart (src/domain.cpp:123-157): physical velocity
 * [ncom][nz][plane_stride] (device) -> forward transform, Nyquist plane
 * zeroed; pressure and history cleared, step = 0. */
int sx_dns_set_state (sx_dns_t d, const double* d_u_phys, void* stream);
/* Steady velocity boundary values: d_lines_phys [ncom][nz][nbound*np]
 * (device) in physical space, segment b's np values in CCW traverse order,
 * as Field::evaluateBoundaries fills _line (src/field.cpp:206-240); they
 * are forward transformed once as Field::bTransform does
 * (src/field.cpp:174-203) and used by every velocity solve (only the
 * essential segments' values matter).  NULL resets them to zero. */
int sx_dns_set_velocity_bc (sx_dns_t d, const double* d_lines_phys, void* stream);
/* 3/2-rule dealiasing of the z products of the skew-symmetric nonlinear
 * term (on = 1), the treatment AuxField::transform32 provides
 * (src/auxfield.cpp:1197-1258; the reference's nonlinear terms no longer
 * call it, dns/nonlinear.cpp:72-76, so it is off by default and the
 * parity mode is the reference's): the velocity and its z derivative are
 * zero-padded onto M = 3 nz / 2 physical planes, the products formed there
 * and the result truncated to modes 0..nz/2-1 (the Nyquist plane of the
 * nonlinear term is zero).  The z transforms run as FP64 GEMMs with the
 * packed DFT matrices.  3D fields with nz a multiple of 4, SX_SKEW. */
int sx_dns_set_dealias (sx_dns_t d, int on, void* stream);
/* Advance nsteps steps of the integrate() loop body. */
int sx_dns_step (sx_dns_t d, int nsteps, void* stream);
/* Velocity solves of the start-up steps (step < nord): SX_FP64 (direct,
 * the default) or SX_JACPCG as the reference's Solve does
 * (dns/integrate.cpp:328-356): a JACPCG system iterated from the current
 * velocity to tol_rel (TOL_REL) within step_max (STEP_MAX). */
int sx_dns_set_startup (sx_dns_t d, int method, double tol_rel, int step_max, void* stream);
/* Copy field `which` (0..ncom-1 velocity, ncom pressure) in Fourier space
 * [nz][plane_stride] to device memory d_out.  Returns the current step
 * count through *step (may be NULL). */
int sx_dns_get (sx_dns_t d, int which, double* d_out, int* step, void* stream);
/* Plane stride of the DNS fields (Geometry::planeSize()). */
int sx_dns_plane_stride (sx_dns_t d);

/* ------------------------------------------------------------------------
 * Field files (binary nekton format, host memory): the 10-line header of
 * writeField (src/misc.cpp:106-180) / Header (src/data2df.cpp:437-516)
 * followed, for each field named in `fields` (one character each, in
 * order), by nz planes of nel*nr*ns doubles in the machine's byte order
 * (AuxField operator<<, src/auxfield.cpp:1021-1070).  Dumps may be
 * appended one after another (the reference's .fld with CHKPOINT = 0).
 * ---------------------------------------------------------------------- */
typedef struct {
  char   session[64];     /* "Session" (the reader keeps 24 characters)    */
  char   created[64];     /* "Created"; empty on write: the current time   */
  int    nr, ns, nz, nel; /* "Nr, Ns, Nz, Elements"                        */
  int    step;            /* "Step"                                        */
  double time, dt;        /* "Time", "Time step"                           */
  double kinvis, beta;    /* "Kinvis", "Beta"                              */
  char   fields[32];      /* "Fields written", e.g. "uvwp"                 */
  char   format[64];      /* "Format": "binary IEEE little-endian"; set on
                             write from the machine                        */
} sx_field_header;
/* writeField: data[i] ([nz][plane_stride] host doubles) for fields[i];
 * append = 1 adds a dump at the end of the file. */
int sx_field_write (const char* path, const sx_field_header* hdr,
                    const double* const* data, int plane_stride, int append);
/* Header of dump number `dump` (0 = first) of a field file. */
int sx_field_read_header (const char* path, int dump, sx_field_header* hdr);
/* readField for one field: field `name` of dump `dump` into data
 * [nz][plane_stride] (plane padding zeroed), byte-swapped when the file's
 * format names the other endianness (Header::swab, src/data2df.cpp:527). */
int sx_field_read (const char* path, int dump, char name, double* data, int plane_stride);

/* The numbering `enumerate -O<optlev>` generates for one field
 * (utility/enumerate.cpp:110-199 with Mesh::buildMap/buildMask,
 * src/mesh.cpp:884-1326): the element-boundary to global map of the
 * condensed system, essential nodes numbered last, the unknowns renumbered
 * by reverse Cuthill-McKee from a pseudo-peripheral root (optlev 1; 2 and
 * 3 search roots for the least band width; 0 keeps the traversal order).
 * elem_nodes: [nel][4] 0-based node indices (the session's NODES order) of
 * each quadrilateral, counter-clockwise.  surf: [nsurf][5] the SURFACES in
 * session order, 0-based: {elmt, side, 0, essential, 0} for a boundary
 * side whose group's BC for this field is essential (<D>, or <A>/<O> as
 * Mesh::buildMask decides), {elmt, side, 1, mate elmt, mate side} for a
 * periodic pair.  Out: btog, bmask [nel][4(np-1)] (emap order), nglobal,
 * nsolve, nband (band width with the diagonal, Nsys::globalBandwidth). */
int sx_enumerate (int nel, int nnode, const int* elem_nodes, int nsurf, const int* surf,
                  int np, int optlev, int* btog, int* bmask, int* nglobal, int* nsolve,
                  int* nband);

/* session.num numbering files.  sx_num_write prints the enumerate summary
 * and table (utility/enumerate.cpp:441-552): `fields` all field names,
 * nset numbering systems, set_fields[j] the fields sharing system j,
 * btog[j]/bmask[j] [nel][4(np-1)] per element side in emap order,
 * nglobal/nsolve/optlev/nband per system.  sx_num_read parses it as
 * BCmgr::buildnum does (src/bcmgr.cpp:586-780), checking nel and np:
 * set_fields receives nset strings of field_cap chars, btog/bmask
 * [nset][nel*4(np-1)], the rest [nset]. */
int sx_num_write (const char* path, const char* fields, int nset,
                  const char* const* set_fields, int nel, int np,
                  const int* const* btog, const int* const* bmask,
                  const int* nglobal, const int* nsolve, const int* optlev,
                  const int* nband);
int sx_num_read (const char* path, int nel, int np, int max_sets, int* nset,
                 char* set_fields, int field_cap, int* btog, int* bmask,
                 int* nglobal, int* nsolve, int* optlev, int* nband);

/* Per-kernel timing of sx_field_solve with HIP events recorded on the
 * solve's stream (no counterpart in the reference, whose only timer is the CPU
 * clock dclock(), veclib/util.c:68).  Phases: 0 essential BCs,
 * 1 condense, 2 assemble, 3 band solve, 4 recover.  sx_timing_read
 * synchronises the recorded events, fills ms[p] (total milliseconds) and
 * count[p] (launches) for p < nphase, and clears the record. */
#define SX_NPHASE 5
int sx_timing_enable (int on);
int sx_timing_read (double* ms, int* count, int nphase);

#ifdef __cplusplus
}
#endif
#endif
