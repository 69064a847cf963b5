/*
 * sem_oracle.c -- CPU restatement (FP64) of semtex's Helmholtz solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  The product
 * never links or calls it.  Each routine cites the reference routine it
 * restates; the arithmetic order follows the reference where that is cheap
 * so that agreement is at round-off level.
 *
 * Pinning (tests/test_oracle_*.py): GLL rule and edge maps against the
 * reference's own polylib.c / mapping.c compiled into oracle/_ref; element
 * matrices against a dense Kronecker-product assembly and the constant-mode
 * identity H*1 = lambda2*M*1; condensation against the dense Schur
 * complement; banded Cholesky against dense numpy solves; the whole solve
 * against the reference's analytic elliptic sessions (mesh/helmholtz1 ...)
 * and manufactured solutions.
 */
#include "sem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPSDP 6.0e-14   /* femlib/cfemdef.h: EPSDP (double tolerance)       */
#define EPSSP 6.0e-7    /* femlib/cfemdef.h: EPSSP                          */

static double sqr (double x) { return x * x; }

/* ------------------------------------------------------------------------ */
/* GLL quadrature.  Published method: Newton iteration on the Legendre
 * three-term recurrence (G. von Winckel, "lglnodes"), weights
 * 2/(N(N+1)P_N(z)^2), and the classical collocation derivative matrix
 * (Canuto, Hussaini, Quarteroni & Zang, Spectral Methods, eq. 2.4.34).
 * Same rule as femlib/polylib.c:312 zwglj(alpha=beta=0), :513 Dglj.      */
/* ------------------------------------------------------------------------ */
static void legendre (int N, double x, double* PN, double* PNm1)
{
  double p0 = 1.0, p1 = x, p2;
  int k;
  if (N == 0) { *PN = 1.0; *PNm1 = 0.0; return; }
  for (k = 2; k <= N; k++) {
    p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1; p1 = p2;
  }
  *PN = p1; *PNm1 = p0;
}

void or_gll (int np, double* z, double* w, double* D, double* DT)
{
  const int N = np - 1;
  int i, j, it;
  double* PN = (double*) malloc (sizeof (double) * np);

  if (np == 1) { z[0] = 0.0; w[0] = 2.0; if (D) D[0] = 0.0;
                 if (DT) DT[0] = 0.0; free (PN); return; }

  for (i = 0; i < np; i++) {
    double x = -cos (M_PI * i / N), xold, pN, pNm1, pNp1;
    if (i == 0)      x = -1.0;
    else if (i == N) x =  1.0;
    else {
      for (it = 0; it < 100; it++) {
        /* x_new = x - (x P_N - P_{N-1}) / ((N+1) P_N)  (Newton on the
           polynomial (1-x^2) P_N'(x) written through P_{N+1}).          */
        legendre (N, x, &pN, &pNm1);
        pNp1 = ((2.0 * N + 1.0) * x * pN - N * pNm1) / (N + 1.0);
        (void) pNp1;
        xold = x;
        x = xold - (xold * pN - pNm1) / ((N + 1.0) * pN);
        if (fabs (x - xold) < 1.0e-16) break;
      }
    }
    z[i] = x;
  }
  /* enforce exact antisymmetry of the rule */
  for (i = 0; i < np / 2; i++) { double a = 0.5 * (z[N - i] - z[i]);
                                 z[i] = -a; z[N - i] = a; }
  if (np & 1) z[N / 2] = 0.0;

  for (i = 0; i < np; i++) {
    double pNm1;
    legendre (N, z[i], &PN[i], &pNm1);
    w[i] = 2.0 / (N * (N + 1.0) * PN[i] * PN[i]);
  }
  if (D) {
    for (i = 0; i < np; i++)
      for (j = 0; j < np; j++) {
        double d;
        if (i != j)              d = PN[i] / (PN[j] * (z[i] - z[j]));
        else if (i == 0)         d = -0.25 * N * (N + 1.0);
        else if (i == N)         d =  0.25 * N * (N + 1.0);
        else                     d = 0.0;
        D[i * np + j] = d;
        if (DT) DT[j * np + i] = d;
      }
  }
  free (PN);
}

/* femlib/mapping.c:46-132 edgemaps(dim=2): CCW around the edges starting at
 * node (0,0), then the interior in row-major order.                       */
void or_edgemaps (int np, int* emap, int* pmap)
{
  const int nm = np - 1;
  int i, j, k = 0, n = 0;
  emap[k++] = 0;
  for (i = 1; i < np; i++) { n += 1;  if (k < 4 * nm) emap[k++] = n; }
  for (i = 1; i < np; i++) { n += np; if (k < 4 * nm) emap[k++] = n; }
  for (i = 1; i < np; i++) { n -= 1;  if (k < 4 * nm) emap[k++] = n; }
  for (i = 1; i < np; i++) { n -= np; if (k < 4 * nm) emap[k++] = n; }
  for (i = 1; i < nm; i++)
    for (j = 1; j < nm; j++) emap[k++] = i * np + j;
  if (pmap) for (i = 0; i < np * np; i++) pmap[emap[i]] = i;
}

/* row-major C = A (nra x nca) * B (nca x ncb): veclib/blas.h mxm          */
static void mxm (const double* A, int nra, const double* B, int nca,
                 double* C, int ncb)
{
  int i, j, k;
  for (i = 0; i < nra; i++)
    for (j = 0; j < ncb; j++) {
      double s = 0.0;
      for (k = 0; k < nca; k++) s += A[i * nca + k] * B[k * ncb + j];
      C[i * ncb + j] = s;
    }
}

/* src/element.cpp:1253-1408 Element::mapping.  Null-mapping optimisation
 * (arrays with norm < 4 EPSDP dropped) is represented by zero arrays. */
int or_mapping (int np, int cyl, const double* x, const double* y, double* geo)
{
  const int npnp = np * np;
  const double EPS = 4.0 * EPSDP;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp), *DT = malloc (sizeof (double) * npnp),
         *dxdr = malloc (sizeof (double) * 5 * npnp);
  double *dxds = dxdr + npnp, *dydr = dxds + npnp, *dyds = dydr + npnp,
         *jac = dyds + npnp;
  double *Q1 = geo + OR_Q1 * npnp, *Q2 = geo + OR_Q2 * npnp,
         *Q3 = geo + OR_Q3 * npnp, *Q4 = geo + OR_Q4 * npnp,
         *Q8 = geo + OR_Q8 * npnp, *rx = geo + OR_DRDX * npnp,
         *ry = geo + OR_DRDY * npnp, *sx = geo + OR_DSDX * npnp,
         *sy = geo + OR_DSDY * npnp, *dl = geo + OR_DELTA * npnp;
  int ij, i, j, bad = 0, a;

  or_gll (np, z, w, D, DT);
  mxm (x, np, DT, np, dxdr, np);
  mxm (D, np, x,  np, dxds, np);
  mxm (y, np, DT, np, dydr, np);
  mxm (D, np, y,  np, dyds, np);

  for (ij = 0; ij < npnp; ij++) {
    jac[ij] = dxdr[ij] * dyds[ij] - dxds[ij] * dydr[ij];
    if (jac[ij] < EPS) bad = 1;
  }
  for (i = 0, ij = 0; i < np; i++)
    for (j = 0; j < np; j++, ij++) {
      const double WW = w[i] * w[j];
      Q1[ij] = (dyds[ij] * dyds[ij] + dxds[ij] * dxds[ij]) / jac[ij] * WW;
      Q2[ij] = (dydr[ij] * dydr[ij] + dxdr[ij] * dxdr[ij]) / jac[ij] * WW;
      Q3[ij] = (-dydr[ij] * dyds[ij] - dxdr[ij] * dxds[ij]) / jac[ij] * WW;
      Q4[ij] = jac[ij] * WW;
      rx[ij] =  dyds[ij] / jac[ij];
      ry[ij] = -dxds[ij] / jac[ij];
      sx[ij] = -dydr[ij] / jac[ij];
      sy[ij] =  dxdr[ij] / jac[ij];
      dl[ij] = sqrt (2.0 * Q4[ij] / 2.0);
      Q8[ij] = Q4[ij];
    }
  if (cyl)
    for (ij = 0; ij < npnp; ij++) {
      Q1[ij] *= y[ij]; Q2[ij] *= y[ij]; Q3[ij] *= y[ij]; Q4[ij] *= y[ij];
    }
  /* null-mapping optimisations */
  for (a = OR_Q3; a <= OR_DSDY; a++) {
    double s = 0.0;
    if (a == OR_Q4 || a == OR_Q8) continue;
    if (a == OR_Q3 || a >= OR_DRDX) {
      double* v = geo + a * npnp;
      for (ij = 0; ij < npnp; ij++) s += v[ij] * v[ij];
      if (sqrt (s) < EPS) memset (v, 0, sizeof (double) * npnp);
    }
  }
  free (z); free (w); free (D); free (DT); free (dxdr);
  return bad ? -1 : 0;
}

static int is_zero (const double* v, int n)
{ int i; for (i = 0; i < n; i++) if (v[i] != 0.0) return 0; return 1; }

/* src/element.cpp:1411-1481 Element::HelmholtzRow (SVV off: SDV == DV). */
void or_helmholtz_row (int np, int cyl, const double* geo, const double* y,
                       double lambda2, double betak2, int i, int j, double* h)
{
  const int npnp = np * np;
  const double *Q1 = geo + OR_Q1 * npnp, *Q2 = geo + OR_Q2 * npnp,
               *Q3 = geo + OR_Q3 * npnp, *Q4 = geo + OR_Q4 * npnp;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp);
  const double r2 = sqr (y[i * np + j]);
  const double hCon = (cyl && r2 > EPSDP) ? betak2 / r2 + lambda2
                                          : betak2 + lambda2;
  const int hasQ3 = !is_zero (Q3, npnp);
  int m, n, q;

  or_gll (np, z, w, D, 0);
  memset (h, 0, sizeof (double) * npnp);
  /* DT[a][b] = D[b][a] */
  for (n = 0; n < np; n++) {
    double s = 0.0;
    for (q = 0; q < np; q++) s += Q1[i * np + q] * (D[q * np + j] * D[q * np + n]);
    h[i * np + n] = s;
  }
  for (m = 0; m < np; m++) {
    double s = 0.0;
    for (q = 0; q < np; q++) s += Q2[q * np + j] * (D[q * np + i] * D[q * np + m]);
    h[m * np + j] += s;
  }
  if (hasQ3)
    for (m = 0; m < np; m++)
      for (n = 0; n < np; n++) {
        h[m * np + n] += Q3[i * np + n] * D[n * np + j] * D[i * np + m];
        h[m * np + n] += Q3[m * np + j] * D[j * np + n] * D[m * np + i];
      }
  h[i * np + j] += Q4[i * np + j] * hCon;
  free (z); free (w); free (D);
}

/* In-place inverse of a dense n x n matrix by LU with partial pivoting
 * (the job of LAPACK getrf + getri in src/element.cpp:199,207). */
static int invert (int n, double* A)
{
  int* piv = malloc (sizeof (int) * n);
  double* B = malloc (sizeof (double) * n * n);
  int i, j, k, p;
  memcpy (B, A, sizeof (double) * n * n);
  for (k = 0; k < n; k++) {
    double amax = 0.0;
    p = k;
    for (i = k; i < n; i++) if (fabs (B[i * n + k]) > amax) { amax = fabs (B[i * n + k]); p = i; }
    if (amax == 0.0) { free (piv); free (B); return -1; }
    piv[k] = p;
    if (p != k) for (j = 0; j < n; j++) { double t = B[k*n+j]; B[k*n+j] = B[p*n+j]; B[p*n+j] = t; }
    for (i = k + 1; i < n; i++) {
      double l = B[i * n + k] /= B[k * n + k];
      for (j = k + 1; j < n; j++) B[i * n + j] -= l * B[k * n + j];
    }
  }
  /* solve for each column of the identity */
  for (j = 0; j < n; j++) {
    double* x = A + 0;  (void) x;
  }
  {
    double* col = malloc (sizeof (double) * n);
    for (j = 0; j < n; j++) {
      for (i = 0; i < n; i++) col[i] = (i == j) ? 1.0 : 0.0;
      for (k = 0; k < n; k++) { if (piv[k] != k) { double t = col[k]; col[k] = col[piv[k]]; col[piv[k]] = t; } }
      for (i = 0; i < n; i++) { double s = col[i]; for (k = 0; k < i; k++) s -= B[i*n+k] * col[k]; col[i] = s; }
      for (i = n - 1; i >= 0; i--) { double s = col[i]; for (k = i + 1; k < n; k++) s -= B[i*n+k] * col[k]; col[i] = s / B[i*n+i]; }
      for (i = 0; i < n; i++) A[i * n + j] = col[i];
    }
    free (col);
  }
  free (piv); free (B);
  return 0;
}

/* src/element.cpp:131-214 Element::HelmholtzSC. */
int or_helmholtz_sc (int np, int cyl, const double* geo, const double* y,
                     double lambda2, double betak2,
                     double* hbb, double* hbi, double* hii)
{
  const int npnp = np * np, next = 4 * (np - 1), nint = npnp - next;
  int *emap = malloc (sizeof (int) * npnp), *pmap = malloc (sizeof (int) * npnp);
  double *rmat = malloc (sizeof (double) * npnp), *rwrk = malloc (sizeof (double) * npnp);
  double *tmp;
  int i, j, ij = 0, eq, a, b, k;

  or_edgemaps (np, emap, pmap);
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++, ij++) {
      or_helmholtz_row (np, cyl, geo, y, lambda2, betak2, i, j, rmat);
      for (k = 0; k < npnp; k++) rwrk[k] = rmat[emap[k]];
      if ((eq = pmap[ij]) < next) {
        memcpy (hbb + eq * next, rwrk,        sizeof (double) * next);
        memcpy (hbi + eq * nint, rwrk + next, sizeof (double) * nint);
      } else
        memcpy (hii + (eq - next) * nint, rwrk + next, sizeof (double) * nint);
    }
  if (nint) {
    if (invert (nint, hii)) { free (emap); free (pmap); free (rmat); free (rwrk); return -1; }
    /* rwrk = hbi * inv(hii)  (next x nint); hbb -= rwrk * hbi^T */
    tmp = malloc (sizeof (double) * next * nint);
    mxm (hbi, next, hii, nint, tmp, nint);
    for (a = 0; a < next; a++)
      for (b = 0; b < next; b++) {
        double s = 0.0;
        for (k = 0; k < nint; k++) s += tmp[a * nint + k] * hbi[b * nint + k];
        hbb[a * next + b] -= s;
      }
    memcpy (hbi, tmp, sizeof (double) * next * nint);
    free (tmp);
  }
  free (emap); free (pmap); free (rmat); free (rwrk);
  return 0;
}

/* src/element.cpp:1536-1622 HelmholtzKern + HelmholtzOp (SVV off). */
void or_helmholtz_op (int np, int cyl, const double* geo, const double* y,
                      double lambda2, double betak2,
                      const double* src, double* tgt)
{
  const int npnp = np * np;
  const double *g1 = geo + OR_Q1 * npnp, *g2 = geo + OR_Q2 * npnp,
               *g3 = geo + OR_Q3 * npnp, *g4 = geo + OR_Q4 * npnp;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp), *DT = malloc (sizeof (double) * npnp),
         *R = malloc (sizeof (double) * 4 * npnp);
  double *S = R + npnp, *T = S + npnp, *U = T + npnp;
  const int hasQ3 = !is_zero (g3, npnp);
  int ij;

  or_gll (np, z, w, D, DT);
  mxm (src, np, DT, np, R, np);      /* R = src * DT  (d/dr)  */
  mxm (D,   np, src, np, S, np);     /* S = D * src   (d/ds)  */
  for (ij = 0; ij < npnp; ij++) {
    double hCon, r2, t = R[ij];
    if (cyl) { r2 = y[ij] * y[ij]; hCon = (r2 > EPSDP) ? betak2 / r2 + lambda2 : 0.0; }
    else hCon = betak2 + lambda2;
    if (hasQ3) {
      R[ij] = g1[ij] * R[ij] + g3[ij] * S[ij];
      S[ij] = g2[ij] * S[ij] + g3[ij] * t;
    } else { R[ij] *= g1[ij]; S[ij] *= g2[ij]; }
    U[ij] = g4[ij] * src[ij] * hCon;
  }
  /* tgt = U + DT*S + R*D  (Blas::mxma accumulations) */
  mxm (DT, np, S, np, T, np);
  for (ij = 0; ij < npnp; ij++) U[ij] += T[ij];
  mxm (R, np, D, np, T, np);
  for (ij = 0; ij < npnp; ij++) tgt[ij] = U[ij] + T[ij];
  free (z); free (w); free (D); free (DT); free (R);
}

/* src/element.cpp:521-565 Element::grad. */
void or_grad (int np, const double* geo, double* tgtX, double* tgtY)
{
  const int npnp = np * np;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp), *DT = malloc (sizeof (double) * npnp),
         *A = malloc (sizeof (double) * 2 * npnp);
  double *B = A + npnp;
  double* tg[2] = { tgtX, tgtY };
  int c, ij;

  or_gll (np, z, w, D, DT);
  for (c = 0; c < 2; c++) {
    const double *dr = geo + (c ? OR_DRDY : OR_DRDX) * npnp,
                 *ds = geo + (c ? OR_DSDY : OR_DSDX) * npnp;
    double* t = tg[c];
    if (!t) continue;
    mxm (t, np, DT, np, A, np);
    mxm (D, np, t,  np, B, np);
    for (ij = 0; ij < npnp; ij++) t[ij] = A[ij] * dr[ij] + B[ij] * ds[ij];
  }
  free (z); free (w); free (D); free (DT); free (A);
}

/* ------------------------------------------------------------------------ */
/* Banded Cholesky.  Storage: ab[j*ldab + kd + i - j] = A(i,j), i <= j.     */
/* LAPACK dpbtrf/dpbtrs('U') (call sites src/matrix.cpp:252,
 * src/field.cpp:693).  Right-looking, column by column.                    */
/* ------------------------------------------------------------------------ */
int or_pbtrf (int n, int kd, double* ab, int ldab)
{
  int j, i, l;
  double* row = malloc (sizeof (double) * (kd + 1));
#define AB(r, c) ab[(size_t) (c) * ldab + kd + (r) - (c)]
  for (j = 0; j < n; j++) {
    const int last = (j + kd < n - 1) ? j + kd : n - 1;
    double d = AB (j, j);
    if (d <= 0.0) { free (row); return j + 1; }
    d = sqrt (d);
    AB (j, j) = d;
    for (i = j + 1; i <= last; i++) row[i - j] = (AB (j, i) /= d);
    /* U(i,l) -= U(j,i) U(j,l) for j < i <= l <= last, column by column
     * (each column of the band is contiguous in i)                      */
    for (l = j + 1; l <= last; l++) {
      const double ujl = row[l - j];
      double* col = &AB (0, l);
      for (i = j + 1; i <= l; i++) col[i] -= row[i - j] * ujl;
    }
  }
  free (row);
  return 0;
}

void or_pbtrs (int n, int kd, const double* ab, int ldab, double* b)
{
  int j, i;
  /* U^T y = b (forward) */
  for (j = 0; j < n; j++) {
    const int first = (j - kd > 0) ? j - kd : 0;
    double s = b[j];
    for (i = first; i < j; i++) s -= AB (i, j) * b[i];
    b[j] = s / AB (j, j);
  }
  /* U x = y (backward, column oriented) */
  for (j = n - 1; j >= 0; j--) {
    const int first = (j - kd > 0) ? j - kd : 0;
    const double xj = b[j] / AB (j, j);
    b[j] = xj;
    for (i = first; i < j; i++) b[i] -= AB (i, j) * xj;
  }
#undef AB
}

/* ------------------------------------------------------------------------ */
/* Mode system and whole-plane solve.                                       */
/* ------------------------------------------------------------------------ */
int or_msys_build (const or_mesh* M, double lambda2, double betak2, or_msys* S)
{
  const int np = M->np, npnp = np * np, next = 4 * (np - 1),
            nint = npnp - next, nel = M->nel;
  double* hbb = malloc (sizeof (double) * next * next);
  int e, i, k, rc = 0, fmask = M->nglobal - M->nsolve;

  S->lambda2 = lambda2; S->betak2 = betak2;
  /* src/matrix.cpp:154-156: singular when lambda2+betak2 ~ 0, no fixed
   * nodes and no mixed BCs -- the last unknown is then dropped.          */
  S->singular = (lambda2 + betak2) < EPSSP && !fmask;
  S->nsolve = S->singular ? M->nsolve - 1 : M->nsolve;
  S->nband  = M->nband;
  S->H   = calloc ((size_t) S->nband * (S->nsolve > 0 ? S->nsolve : 1), sizeof (double));
  S->hbi = calloc ((size_t) nel * next * (nint ? nint : 1), sizeof (double));
  S->hii = calloc ((size_t) nel * (nint ? nint * nint : 1), sizeof (double));
  S->geo = calloc ((size_t) nel * OR_NGEO * npnp, sizeof (double));
  S->hbb_cache = calloc ((size_t) nel * next * next, sizeof (double));

  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    double* geo = S->geo + (size_t) e * OR_NGEO * npnp;
    int twin = -1, f;
    if (or_mapping (np, M->cyl, M->x + e * npnp, M->y + e * npnp, geo)) rc = -2;
    /* Family storage (femlib/family.c:171-192, veclib/lxsame.c:37-49):
     * the reference keeps one copy of each geometric-factor array, and an
     * element whose array matches an earlier element's to EPSDP (absolute,
     * entry by entry) adopts the earlier copy (src/element.cpp:82-93).
     * The elemental matrices then depend only on adopted arrays; when
     * Q1..Q4 (and, cylindrical, the radii) are the same storage as an
     * earlier element's, its matrix packs are the same too
     * (src/matrix.cpp:236-237), and are copied instead of recomputed.   */
    for (k = 0; k < OR_NGEO; k++) {
      double* ga = geo + (size_t) k * npnp;
      for (f = 0; f < e; f++) {
        const double* gb = S->geo + ((size_t) f * OR_NGEO + k) * npnp;
        int same = 1;
        for (i = 0; i < npnp && same; i++) same = fabs (ga[i] - gb[i]) <= EPSDP;
        if (same) { memcpy (ga, gb, sizeof (double) * npnp); break; }
      }
    }
    for (f = 0; f < e && twin < 0; f++)
      if (!memcmp (geo, S->geo + (size_t) f * OR_NGEO * npnp, sizeof (double) * 4 * npnp) &&
          (!M->cyl || !memcmp (M->y + e * npnp, M->y + f * npnp, sizeof (double) * npnp)))
        twin = f;
    if (twin >= 0) {
      memcpy (hbb, S->hbb_cache + (size_t) twin * next * next, sizeof (double) * next * next);
      memcpy (S->hbi + (size_t) e * next * nint, S->hbi + (size_t) twin * next * nint,
              sizeof (double) * next * nint);
      memcpy (S->hii + (size_t) e * nint * nint, S->hii + (size_t) twin * nint * nint,
              sizeof (double) * nint * nint);
    } else if (or_helmholtz_sc (np, M->cyl, geo, M->y + e * npnp, lambda2, betak2, hbb,
                                S->hbi + (size_t) e * next * nint,
                                S->hii + (size_t) e * nint * nint)) rc = -3;
    memcpy (S->hbb_cache + (size_t) e * next * next, hbb, sizeof (double) * next * next);
    /* src/matrix.cpp:229-234 band assembly */
    for (i = 0; i < next; i++) {
      const int m = bmap[i];
      if (m >= S->nsolve) continue;
      for (k = 0; k < next; k++) {
        const int n = bmap[k];
        if (n < S->nsolve && n >= m)
          S->H[(size_t) n * S->nband + (S->nband - 1) + m - n] += hbb[i * next + k];
      }
    }
  }
  free (hbb);
  free (S->hbb_cache);
  S->hbb_cache = NULL;
  if (rc) return rc;
  if (S->nsolve && or_pbtrf (S->nsolve, S->nband - 1, S->H, S->nband)) return -4;
  return 0;
}

void or_msys_free (or_msys* S)
{ free (S->H); free (S->hbi); free (S->hii); free (S->geo); memset (S, 0, sizeof (*S)); }

/* src/element.cpp:1625-1742 Element::sideGeom. */
void or_side_geom (int np, const double* geo, const double* x, const double* y,
                   int side, double* area, double* nx, double* ny)
{
  const int npnp = np * np;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp);
  const double *rx = geo + OR_DRDX * npnp, *ry = geo + OR_DRDY * npnp,
               *sx = geo + OR_DSDX * npnp, *sy = geo + OR_DSDY * npnp;
  int k, q;
  or_gll (np, z, w, D, 0);
  for (k = 0; k < np; k++) {
    /* node index along the CCW traverse, and the tangent derivative */
    int node; double dx = 0.0, dy = 0.0;
    switch (side) {
    case 0: node = k;                      break;
    case 1: node = k * np + (np - 1);      break;
    case 2: node = (np - 1) * np + (np - 1 - k); break;
    default: node = (np - 1 - k) * np;     break;
    }
    if (side == 0 || side == 2) {
      const int row = (side == 0) ? 0 : np - 1, col = (side == 0) ? k : np - 1 - k;
      for (q = 0; q < np; q++) { dx += D[col * np + q] * x[row * np + q];
                                 dy += D[col * np + q] * y[row * np + q]; }
      area[k] = sqrt (dx * dx + dy * dy) * w[col];
      nx[k] = (side == 0) ? -sx[node] : sx[node];
      ny[k] = (side == 0) ? -sy[node] : sy[node];
    } else {
      const int col = (side == 1) ? np - 1 : 0, row = (side == 1) ? k : np - 1 - k;
      for (q = 0; q < np; q++) { dx += D[row * np + q] * x[q * np + col];
                                 dy += D[row * np + q] * y[q * np + col]; }
      area[k] = sqrt (dx * dx + dy * dy) * w[row];
      nx[k] = (side == 1) ? rx[node] : -rx[node];
      ny[k] = (side == 1) ? ry[node] : -ry[node];
    }
    { const double len = hypot (nx[k], ny[k]); nx[k] /= len; ny[k] /= len; }
  }
  free (z); free (w); free (D);
}

/* Offset of side `side`'s first node in an element's btog block, and the
 * node that closes the side (src/condition.cpp:80-99 Essential::set).    */
static void side_nodes (int np, int side, int* idx)
{
  const int nm = np - 1;
  int k;
  for (k = 0; k < nm; k++) idx[k] = side * nm + k;
  idx[nm] = (side == 3) ? 0 : side * nm + nm;
}

/* src/field.cpp:1078-1102 Field::getEssential (+ Essential::set). */
static void get_essential (const or_mesh* M, const double* line, double* tgt)
{
  const int np = M->np, next = 4 * (np - 1);
  int b, k, idx[64];
  for (b = 0; b < M->nbound; b++) {
    const int* bmap = M->btog + M->bnd_elmt[b] * next;
    if (M->bnd_kind[b] != OR_ESSENTIAL) continue;
    side_nodes (np, M->bnd_side[b], idx);
    for (k = 0; k < np; k++) tgt[bmap[idx[k]]] = line[b * np + k];
  }
}

int or_solve_plane (const or_mesh* M, const or_msys* S, double* f,
                    const double* line, double* u)
{
  const int np = M->np, npnp = np * np, next = 4 * (np - 1),
            nint = npnp - next, nel = M->nel, nglobal = M->nglobal,
            nsolve = S->nsolve;
  double *RHS = calloc (nglobal, sizeof (double)),
         *g = malloc (sizeof (double) * npnp), *work = malloc (sizeof (double) * npnp),
         *area = malloc (sizeof (double) * np), *nx = malloc (sizeof (double) * np),
         *ny = malloc (sizeof (double) * np);
  int *emap = malloc (sizeof (int) * npnp), *pmap = malloc (sizeof (int) * npnp);
  int e, i, k, b, idx[64];

  or_edgemaps (np, emap, pmap);

  /* getEssential, then constrain: f <- -(M f + H g)  (src/field.cpp:812) */
  get_essential (M, line, RHS);
  for (e = 0; e < nel; e++) {
    const double* geo = S->geo + (size_t) e * OR_NGEO * npnp;
    const double* Q8 = geo + OR_Q8 * npnp;
    double* fe = f + (size_t) e * npnp;
    for (i = 0; i < npnp; i++) fe[i] *= Q8[i];
    if (M->emask[e]) {
      const int* bmap = M->btog + e * next;
      memset (g, 0, sizeof (double) * npnp);
      for (i = 0; i < next; i++) g[emap[i]] = RHS[bmap[i]];
      or_helmholtz_op (np, M->cyl, geo, M->y + e * npnp, S->lambda2, S->betak2, g, work);
      for (i = 0; i < npnp; i++) fe[i] += work[i];
    }
    for (i = 0; i < npnp; i++) fe[i] = -fe[i];
  }

  /* buildRHS (src/field.cpp:943-1009) with local2globalSumSC */
  memset (RHS, 0, sizeof (double) * nglobal);
  for (e = 0; e < nel; e++) {
    double* fe = f + (size_t) e * npnp;
    const int* bmap = M->btog + e * next;
    const double* hbi = S->hbi + (size_t) e * next * nint;
    for (i = 0; i < npnp; i++) work[i] = fe[emap[i]];
    memcpy (fe, work, sizeof (double) * npnp);
    for (i = 0; i < next; i++) {
      double s = 0.0;
      for (k = 0; k < nint; k++) s += hbi[i * nint + k] * fe[next + k];
      fe[i] -= s;
    }
    for (i = 0; i < next; i++) RHS[bmap[i]] += fe[i];
  }
  /* natural BCs <h, w> (Natural::sum, src/condition.cpp:192) */
  for (b = 0; b < M->nbound; b++) {
    const int el = M->bnd_elmt[b];
    const int* bmap = M->btog + el * next;
    if (M->bnd_kind[b] != OR_NATURAL) continue;
    or_side_geom (np, S->geo + (size_t) el * OR_NGEO * npnp, M->x + el * npnp,
                  M->y + el * npnp, M->bnd_side[b], area, nx, ny);
    side_nodes (np, M->bnd_side[b], idx);
    for (k = 0; k < np; k++) RHS[bmap[idx[k]]] += line[b * np + k] * area[k];
  }
  for (i = nsolve; i < nglobal; i++) RHS[i] = 0.0;

  if (nsolve) or_pbtrs (nsolve, S->nband - 1, S->H, S->nband, RHS);

  /* global2localSC (src/element.cpp:368-407) */
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    const double* hbi = S->hbi + (size_t) e * next * nint;
    const double* hii = S->hii + (size_t) e * nint * nint;
    double* fe = f + (size_t) e * npnp;
    double* ue = u + (size_t) e * npnp;
    for (i = 0; i < next; i++) { work[i] = RHS[bmap[i]]; ue[emap[i]] = work[i]; }
    for (i = 0; i < nint; i++) {
      double s = 0.0;
      for (k = 0; k < nint; k++) s += hii[k * nint + i] * fe[next + k];
      for (k = 0; k < next; k++) s -= hbi[k * nint + i] * work[k];
      ue[emap[next + i]] = s;
    }
  }
  /* setEssential (src/field.cpp:1105-1124, Element::bndryMask) */
  memset (RHS, 0, sizeof (double) * nglobal);
  get_essential (M, line, RHS);
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    const int* bm = M->bmask + e * next;
    double* ue = u + (size_t) e * npnp;
    if (!M->emask[e]) continue;
    for (i = 0; i < next; i++) if (bm[i]) ue[emap[i]] = RHS[bmap[i]];
  }
  free (RHS); free (g); free (work); free (area); free (nx); free (ny);
  free (emap); free (pmap);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* JACPCG path.                                                              */
/* ------------------------------------------------------------------------ */
/* src/element.cpp:1484-1533 Element::HelmholtzDiag, row-major (the emap
 * gather is done by the caller).  The cross term reads the derivative
 * matrix at linear indices j*j and i*i, as the reference does; it only
 * contributes where Q3 is nonzero (non-affine elements).                  */
static void helmholtz_diag (int np, int cyl, const double* geo, const double* y,
                            double lambda2, double betak2, double* dg)
{
  const int npnp = np * np;
  double *z = malloc (sizeof (double) * np), *w = malloc (sizeof (double) * np),
         *D = malloc (sizeof (double) * npnp);
  const double *Q1 = geo + OR_Q1 * npnp, *Q2 = geo + OR_Q2 * npnp,
               *Q3 = geo + OR_Q3 * npnp, *Q4 = geo + OR_Q4 * npnp;
  const int q3 = !is_zero (Q3, npnp);
  int i, j, k;
  or_gll (np, z, w, D, 0);
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) {
      const int ij = i * np + j;
      double hcon, d = 0.0;
      if (cyl) { const double r2 = y[ij] * y[ij]; hcon = (r2 > EPSDP) ? betak2 / r2 + lambda2 : 0.0; }
      else hcon = lambda2 + betak2;
      for (k = 0; k < np; k++) d += Q1[i * np + k] * sqr (D[k * np + j]);
      for (k = 0; k < np; k++) d += Q2[k * np + j] * sqr (D[k * np + i]);
      if (q3) d += 2.0 * Q3[ij] * D[j * j] * D[i * i];
      d += hcon * Q4[ij];
      dg[ij] = d;
    }
  free (z); free (w); free (D);
}

/* src/matrix.cpp:269-310: inverse of the assembled diagonal, globally
 * numbered (nglobal boundary values, then nel blocks of nint interior
 * values in emap order).                                                   */
void or_pcg_diag (const or_mesh* M, const or_msys* S, double* pc)
{
  const int np = M->np, npnp = np * np, next = 4 * (np - 1), nint = npnp - next,
            nglobal = M->nglobal, npts = nglobal + M->nel * nint;
  double* dg = malloc (sizeof (double) * npnp);
  int *emap = malloc (sizeof (int) * npnp), *pmap = malloc (sizeof (int) * npnp);
  int e, i;
  or_edgemaps (np, emap, pmap);
  memset (pc, 0, sizeof (double) * npts);
  for (e = 0; e < M->nel; e++) {
    const int* bmap = M->btog + e * next;
    helmholtz_diag (np, M->cyl, S->geo + (size_t) e * OR_NGEO * npnp, M->y + e * npnp,
                    S->lambda2, S->betak2, dg);
    for (i = 0; i < next; i++) pc[bmap[i]] += dg[emap[i]];
    for (i = 0; i < nint; i++) pc[nglobal + e * nint + i] = dg[emap[next + i]];
  }
  for (i = 0; i < npts; i++) pc[i] = 1.0 / pc[i];
  free (dg); free (emap); free (pmap);
}

/* src/field.cpp:855-940 Field::HelmholtzOperator (no mixed BCs): y = A x
 * on globally numbered vectors, element by element.                       */
static void global_helmholtz (const or_mesh* M, const or_msys* S, const int* emap,
                              const double* x, double* y, double* P, double* work)
{
  const int np = M->np, npnp = np * np, next = 4 * (np - 1), nint = npnp - next,
            nglobal = M->nglobal, npts = nglobal + M->nel * nint;
  int e, i;
  memset (y, 0, sizeof (double) * npts);
  for (e = 0; e < M->nel; e++) {
    const int* bmap = M->btog + e * next;
    for (i = 0; i < next; i++) P[emap[i]] = x[bmap[i]];
    for (i = 0; i < nint; i++) P[emap[next + i]] = x[nglobal + e * nint + i];
    or_helmholtz_op (np, M->cyl, S->geo + (size_t) e * OR_NGEO * npnp, M->y + e * npnp,
                     S->lambda2, S->betak2, P, work);
    for (i = 0; i < next; i++) y[bmap[i]] += work[emap[i]];
    for (i = 0; i < nint; i++) y[nglobal + e * nint + i] += work[emap[next + i]];
  }
}

static double dot (int n, const double* a, const double* b)
{
  double s = 0.0;
  int i;
  for (i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* src/field.cpp:711-805 Field::solve, JACPCG branch, one plane.  f is
 * destroyed; u holds the initial guess on entry and the solution on exit.
 * Returns the iteration count i of the reference's loop.                  */
int or_solve_plane_pcg (const or_mesh* M, const or_msys* S, double* f, const double* line,
                        double* u, double tol_rel, int step_max)
{
  const int np = M->np, npnp = np * np, next = 4 * (np - 1), nint = npnp - next,
            nel = M->nel, nglobal = M->nglobal, nsolve = S->nsolve,
            npts = nglobal + nel * nint;
  double *r = calloc (npts, sizeof (double)), *p = calloc (npts, sizeof (double)),
         *q = calloc (npts, sizeof (double)), *x = calloc (npts, sizeof (double)),
         *z = calloc (npts, sizeof (double)), *pc = malloc (sizeof (double) * npts),
         *g = malloc (sizeof (double) * npnp), *work = malloc (sizeof (double) * npnp),
         *P = malloc (sizeof (double) * npnp), *area = malloc (sizeof (double) * np),
         *nx = malloc (sizeof (double) * np), *ny = malloc (sizeof (double) * np);
  int *emap = malloc (sizeof (int) * npnp), *pmap = malloc (sizeof (int) * npnp);
  int e, i, b, k, it, idx[64];
  double epsb2, r2, rho1, rho2 = 1.0, alpha, beta, dotp;

  or_edgemaps (np, emap, pmap);
  or_pcg_diag (M, S, pc);

  /* getEssential into x, constrain f <- -(M f + H g) */
  get_essential (M, line, x);
  for (e = 0; e < nel; e++) {
    const double* geo = S->geo + (size_t) e * OR_NGEO * npnp;
    double* fe = f + (size_t) e * npnp;
    for (i = 0; i < npnp; i++) fe[i] *= geo[OR_Q8 * npnp + i];
    if (M->emask[e]) {
      const int* bmap = M->btog + e * next;
      memset (g, 0, sizeof (double) * npnp);
      for (i = 0; i < next; i++) g[emap[i]] = x[bmap[i]];
      or_helmholtz_op (np, M->cyl, geo, M->y + e * npnp, S->lambda2, S->betak2, g, work);
      for (i = 0; i < npnp; i++) fe[i] += work[i];
    }
    for (i = 0; i < npnp; i++) fe[i] = -fe[i];
  }
  /* buildRHS with element interiors (local2globalSum) + <h, w>, essential
   * tail zeroed (src/field.cpp:943-1009)                                   */
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    const double* fe = f + (size_t) e * npnp;
    for (i = 0; i < next; i++) r[bmap[i]] += fe[emap[i]];
    for (i = 0; i < nint; i++) r[nglobal + e * nint + i] += fe[emap[next + i]];
  }
  for (b = 0; b < M->nbound; b++) {
    const int el = M->bnd_elmt[b];
    const int* bmap = M->btog + el * next;
    if (M->bnd_kind[b] != OR_NATURAL) continue;
    or_side_geom (np, S->geo + (size_t) el * OR_NGEO * npnp, M->x + el * npnp,
                  M->y + el * npnp, M->bnd_side[b], area, nx, ny);
    side_nodes (np, M->bnd_side[b], idx);
    for (k = 0; k < np; k++) r[bmap[idx[k]]] += line[b * np + k] * area[k];
  }
  for (i = nsolve; i < nglobal; i++) r[i] = 0.0;

  epsb2 = tol_rel * sqrt (dot (npts, r, r));
  epsb2 *= epsb2;

  /* initial guess from the field (local2global, last element writes) */
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    const double* ue = u + (size_t) e * npnp;
    for (i = 0; i < next; i++) x[bmap[i]] = ue[emap[i]];
    for (i = 0; i < nint; i++) x[nglobal + e * nint + i] = ue[emap[next + i]];
  }
  for (i = nsolve; i < nglobal; i++) x[i] = 0.0;
  global_helmholtz (M, S, emap, x, p, P, work);
  for (i = nsolve; i < nglobal; i++) { p[i] = 0.0; r[i] = 0.0; }
  for (i = 0; i < npts; i++) r[i] -= p[i];
  r2 = dot (npts, r, r);

  it = 0;
  while (r2 > epsb2 && ++it < step_max) {
    for (i = 0; i < npts; i++) z[i] = pc[i] * r[i];
    rho1 = dot (npts, r, z);
    if (it == 1) memcpy (p, z, sizeof (double) * npts);
    else {
      beta = rho1 / rho2;
      for (i = 0; i < npts; i++) p[i] = beta * p[i] + z[i];
    }
    global_helmholtz (M, S, emap, p, q, P, work);
    for (i = nsolve; i < nglobal; i++) q[i] = 0.0;
    dotp = dot (npts, p, q);
    alpha = rho1 / dotp;
    for (i = 0; i < npts; i++) x[i] += alpha * p[i];
    for (i = 0; i < npts; i++) r[i] -= alpha * q[i];
    rho2 = rho1;
    r2 = dot (npts, r, r);
  }

  /* global2local, then the essential values (getEssential, setEssential) */
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    double* ue = u + (size_t) e * npnp;
    for (i = 0; i < next; i++) ue[emap[i]] = x[bmap[i]];
    for (i = 0; i < nint; i++) ue[emap[next + i]] = x[nglobal + e * nint + i];
  }
  get_essential (M, line, x);
  for (e = 0; e < nel; e++) {
    const int* bmap = M->btog + e * next;
    const int* bm = M->bmask + e * next;
    double* ue = u + (size_t) e * npnp;
    if (!M->emask[e]) continue;
    for (i = 0; i < next; i++) if (bm[i]) ue[emap[i]] = x[bmap[i]];
  }
  free (r); free (p); free (q); free (x); free (z); free (pc); free (g); free (work);
  free (P); free (area); free (nx); free (ny); free (emap); free (pmap);
  return it;
}

/* ------------------------------------------------------------------------ */
/* Real DFT along z (femlib/fourier.c:40-165).  Direct O(N^2) evaluation,
 * exp(-i) convention for the forward transform, scaled by 1/N.            */
/* ------------------------------------------------------------------------ */
void or_dft (double* data, int nz, int ntrn, int sign)
{
  double* line = malloc (sizeof (double) * nz);
  double* out  = malloc (sizeof (double) * nz);
  int t, n, k;
  if (nz < 2) { free (line); free (out); return; }
  for (t = 0; t < ntrn; t++) {
    for (n = 0; n < nz; n++) line[n] = data[(size_t) n * ntrn + t];
    if (sign > 0) {
      for (k = 0; k <= nz / 2; k++) {
        double re = 0.0, im = 0.0;
        for (n = 0; n < nz; n++) {
          const double a = 2.0 * M_PI * (double) ((long) k * n % nz) / nz;
          re += line[n] * cos (a); im -= line[n] * sin (a);
        }
        re /= nz; im /= nz;
        if (k == 0) out[0] = re;
        else if (k == nz / 2) out[1] = re;
        else { out[2 * k] = re; out[2 * k + 1] = im; }
      }
    } else {
      for (n = 0; n < nz; n++) {
        double s = line[0] + line[1] * ((n & 1) ? -1.0 : 1.0);
        for (k = 1; k < nz / 2; k++) {
          const double a = 2.0 * M_PI * (double) ((long) k * n % nz) / nz;
          s += 2.0 * (line[2 * k] * cos (a) - line[2 * k + 1] * sin (a));
        }
        out[n] = s;
      }
    }
    for (n = 0; n < nz; n++) data[(size_t) n * ntrn + t] = out[n];
  }
  free (line); free (out);
}
