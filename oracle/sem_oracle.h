/*
 * sem_oracle.h -- CPU restatement (FP64, plain C) of semtex's per-Fourier-mode
 * Helmholtz solve, used ONLY as the checker for the GPU library.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product (semtex_amd/, include/) may
 * link, load or call this code: only tests/, the smoke check and the CPU
 * baseline leg of the benchmark do.
 *
 * Every function names the reference routine it restates (file:line in
 * jyalim/semtex).  Storage conventions follow the reference:
 *   - element data are row-major np x np, index i*np+j with i the s
 *     (second coordinate) index and j the r index (src/element.h);
 *   - the banded matrix is LAPACK 'U' symmetric band storage with leading
 *     dimension nband, column j holding rows j-kd..j, kd = nband-1
 *     (veclib/lapack.h:107 band_addr);
 *   - global numbering puts essential (Dirichlet) nodes last, so unknowns
 *     are 0..nsolve-1 (utility/enumerate.cpp:635-692).
 */
#ifndef SEM_ORACLE_H
#define SEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Number of per-node geometric arrays returned by or_mapping. */
#define OR_NGEO 10
enum { OR_Q1 = 0, OR_Q2, OR_Q3, OR_Q4, OR_Q8, OR_DRDX, OR_DRDY, OR_DSDX,
       OR_DSDY, OR_DELTA };

/* Boundary-segment kinds (src/condition.h: Essential, Natural). */
enum { OR_ESSENTIAL = 0, OR_NATURAL = 1 };

/* GLL points z, weights w and derivative matrix D[i*np+j] = h_j'(z_i),
 * DT its transpose (femlib/polylib.c:312 zwglj, :513 Dglj). */
void or_gll (int np, double* z, double* w, double* D, double* DT);

/* Edge-major permutation emap and its inverse pmap (femlib/mapping.c:46). */
void or_edgemaps (int np, int* emap, int* pmap);

/* Geometric factors of one element from its nodal coordinates
 * (src/element.cpp:1253-1408 Element::mapping).  geo has OR_NGEO*np*np
 * entries.  Returns 0, or -1 if the Jacobian is not positive. */
int or_mapping (int np, int cyl, const double* x, const double* y, double* geo);

/* Row (i,j) of the elemental Helmholtz matrix, hij is np*np
 * (src/element.cpp:1411-1481 Element::HelmholtzRow). */
void or_helmholtz_row (int np, int cyl, const double* geo, const double* y,
                       double lambda2, double betak2, int i, int j, double* hij);

/* Statically condensed elemental matrices (src/element.cpp:131-214
 * Element::HelmholtzSC): hbb (next^2) condensed, hbi (next*nint) holds
 * hbi*inv(hii), hii (nint^2) holds inv(hii).  Returns 0 or -1 (singular). */
int or_helmholtz_sc (int np, int cyl, const double* geo, const double* y,
                     double lambda2, double betak2,
                     double* hbb, double* hbi, double* hii);

/* Apply the elemental Helmholtz operator tgt = H src (np*np row-major)
 * (src/element.cpp:1536-1622 HelmholtzKern/HelmholtzOp). */
void or_helmholtz_op (int np, int cyl, const double* geo, const double* y,
                      double lambda2, double betak2,
                      const double* src, double* tgt);

/* Gradient of one element's data (src/element.cpp:521-565 Element::grad).
 * Either tgt may be NULL; on input each holds the data to differentiate. */
void or_grad (int np, const double* geo, double* tgtX, double* tgtY);

/* Banded Cholesky factorisation, LAPACK dpbtrf('U') semantics.
 * Returns 0 or the 1-based column where the matrix is not positive. */
int or_pbtrf (int n, int kd, double* ab, int ldab);
/* Banded solve U^T U x = b in place, dpbtrs('U') with nrhs = 1. */
void or_pbtrs (int n, int kd, const double* ab, int ldab, double* b);

/* --------------------------------------------------------------------------
 * A whole mesh + numbering system + boundary segments, as semtex's Domain,
 * NumberSys (src/numbersys.h) and BoundarySys hold them.
 * ------------------------------------------------------------------------ */
typedef struct {
  int     np, nel, cyl;
  const double* x;        /* [nel][np*np] node x                            */
  const double* y;        /* [nel][np*np] node y                            */
  const int*    btog;     /* [nel][next]  element-boundary -> global         */
  const int*    bmask;    /* [nel][next]  1 where essential                  */
  const int*    emask;    /* [nel]        1 if element has essential nodes   */
  int     nglobal, nsolve, nband;
  int     nbound;         /* number of boundary segments                    */
  const int*    bnd_elmt; /* [nbound] element of segment                     */
  const int*    bnd_side; /* [nbound] side 0..3                              */
  const int*    bnd_kind; /* [nbound] OR_ESSENTIAL / OR_NATURAL              */
} or_mesh;

/* Per-mode DIRECT system (src/matrix.cpp:132-267 MatrixSys::MatrixSys). */
typedef struct {
  double  lambda2, betak2;
  int     nsolve, nband, singular;
  double* H;              /* nband*nsolve factored band                     */
  double* hbi;            /* [nel][next*nint]                               */
  double* hii;            /* [nel][nint*nint]                               */
  double* geo;            /* [nel][OR_NGEO*np*np]                           */
  double* hbb_cache;      /* build-time scratch, NULL afterwards            */
} or_msys;

/* Build (and Cholesky-factor) one mode's system.  Returns 0 or an error. */
int  or_msys_build (const or_mesh* M, double lambda2, double betak2,
                    or_msys* S);
void or_msys_free  (or_msys* S);

/* Side quadrature weights (|dx/dr| w) and unit outward normal of side
 * `side` of one element (src/element.cpp:1625-1742 Element::sideGeom). */
void or_side_geom (int np, const double* geo, const double* x,
                   const double* y, int side,
                   double* area, double* nx, double* ny);

/* One plane of Field::solve DIRECT (src/field.cpp:664-709): forcing f
 * (np*np per element, destroyed), line = boundary values (nbound*np),
 * u = output plane.  Returns 0. */
int or_solve_plane (const or_mesh* M, const or_msys* S,
                    double* f, const double* line, double* u);

/* Real DFT of nlines interleaved lines of length nz along z with stride
 * `ntrn` (femlib/fourier.c:40-165 dDFTr): sign +1 forward (scaled 1/nz),
 * -1 inverse.  Packing [Re0, ReNyq, Re1, Im1, ...]. */
/* JACPCG (src/field.cpp:711-805, src/matrix.cpp:269-310): inverse of the
 * assembled diagonal (globally numbered, nglobal + nel*nint), and the
 * preconditioned CG solve of one plane from the initial guess in u.
 * Returns the iteration count.                                            */
void or_pcg_diag (const or_mesh* M, const or_msys* S, double* pc);
int  or_solve_plane_pcg (const or_mesh* M, const or_msys* S, double* f, const double* line,
                         double* u, double tol_rel, int step_max);

void or_dft (double* data, int nz, int ntrn, int sign);

#ifdef __cplusplus
}
#endif
#endif
