/*
 * p3d_oracle.c -- CPU restatement of POT3D's preconditioned conjugate-
 * gradient solve of the finite-difference spherical Laplace equation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library (libpot3d_hip.so) never links, loads or calls it.
 *
 * Every routine names the reference routine it restates (file:line in
 * predsci/POT3D).  Layout follows the reference exactly:
 *   padded 3-D arrays s(nr,nt,np), Fortran order, 1-based:
 *       s(i,j,k) = s[(i-1) + nr*((j-1) + nt*(k-1))]
 *   packed interior vectors x(N), N=(nr-2)(nt-2)(np-2), r fastest
 *   (src/pot3d.F90:3907-3917, 5323):
 *       x(i,j,k) = x[(i-2) + a*((j-2) + b*(k-2))], a=nr-2, b=nt-2
 *   2-D boundary arrays br0(nt,np): br0[(j-1) + nt*(k-1)]
 *   1-D mesh arrays are passed 0-based: R(i) = r[i-1].
 * Built with -ffp-contract=off so products and sums round in the order
 * they are written, which is the order of the reference's Fortran source.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   coefficients, ax, rhs  -- explicit scipy sparse assembly, symmetry,
 *                             row sums, closed-form RHS, tiny brute force
 *   getM/diacsr/ilu0/lusol -- dense textbook ILU(0) and dense triangular
 *                             solves on tiny grids
 *   dilu_*                 -- equality with ilu0 (to FP32 rounding) and
 *                             with dense ILU(0) in FP64
 *   cgsolve                -- dense direct solve on tiny grids; the three
 *                             testsuite decks' energies and iteration counts
 *   getb/energy            -- closed-form fields on uniform meshes; the
 *                             testsuite energies
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nr, nt, np;
  const double *r, *dr;    /* nr-1 entries: main r mesh r(1:nrm1), dr(1:nrm1) */
  const double *rh, *drh;  /* nr entries: half r mesh */
  const double *st, *dt;   /* nt-1 entries: sin(t(1:ntm1)), dt(1:ntm1) */
  const double *sth, *dth; /* nt entries: sin(th(1:nt)), dth(1:nt) */
  const double *dp, *dph;  /* np entries each (periodic phi) */
  double pm_r1;            /* +1 'potential', -1 'open'/'ss' */
} orc_grid;

#define PI_F90 3.1415926535897932
/* pl_i = one/pl, pl = two*pi - 0 (src/pot3d.F90:192-197) */
static double pl_i(void) { return 1.0 / (2.0 * PI_F90 - 0.0); }

#define IDX3(i, j, k) ((size_t)((i)-1) + (size_t)nr * ((size_t)((j)-1) + (size_t)nt * (size_t)((k)-1)))
#define PIDX(i, j, k) ((size_t)((i)-2) + (size_t)A * ((size_t)((j)-2) + (size_t)B * (size_t)((k)-2)))

/* ---------------------------------------------------------------------- */
/* Deterministic FP64 dot product: fixed 4096-element chunks summed in
 * order, so the result does not depend on the thread count.            */
static double det_dot(const double *x, const double *y, size_t n) {
  const size_t CH = 4096;
  size_t nch = (n + CH - 1) / CH;
  double *part = (double *)malloc(sizeof(double) * (nch ? nch : 1));
  long c;
#pragma omp parallel for schedule(static)
  for (c = 0; c < (long)nch; c++) {
    size_t lo = (size_t)c * CH, hi = lo + CH < n ? lo + CH : n;
    double s = 0.0;
    for (size_t i = lo; i < hi; i++) s = s + x[i] * y[i];
    part[c] = s;
  }
  double s = 0.0;
  for (size_t c2 = 0; c2 < nch; c2++) s = s + part[c2];
  free(part);
  return s;
}

double orc_dot(const double *x, const double *y, int64_t n) { return det_dot(x, y, (size_t)n); }

/* ---------------------------------------------------------------------- */
/* load_matrix_pot3d_solve: src/pot3d.F90:4565-4613.
 * a[m*N + n], m=0..6 holds a(i,j,k,m+1).                                */
void orc_coefs(const orc_grid *g, double *a) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2, C = np - 2;
  const size_t N = (size_t)A * B * C;
  (void)C;
  long k;
#pragma omp parallel for schedule(static)
  for (k = 2; k <= np - 1; k++) {
    for (int j = 2; j <= nt - 1; j++) {
      for (int i = 2; i <= nr - 1; i++) {
        const size_t n = PIDX(i, j, k);
        const double drh = g->drh[i - 1], dth = g->dth[j - 1], sth = g->sth[j - 1];
        const double sth_i = 1.0 / sth;
        const double dph = g->dph[k - 1];
        const double dpi_km1 = 1.0 / g->dp[k - 2], dpi_k = 1.0 / g->dp[k - 1];
        const double st_jm1 = g->st[j - 2], st_j = g->st[j - 1];
        const double dti_jm1 = 1.0 / g->dt[j - 2], dti_j = 1.0 / g->dt[j - 1];
        const double r2_im1 = g->r[i - 2] * g->r[i - 2], r2_i = g->r[i - 1] * g->r[i - 1];
        const double dri_im1 = 1.0 / g->dr[i - 2], dri_i = 1.0 / g->dr[i - 1];
        const double a1 = -(drh * dth * sth_i * dpi_km1);
        const double a2 = -(drh * dph * st_jm1 * dti_jm1);
        const double a3 = -(sth * dth * dph * r2_im1 * dri_im1);
        const double a5 = -(sth * dth * dph * r2_i * dri_i);
        const double a6 = -(drh * dph * st_j * dti_j);
        const double a7 = -(drh * dth * sth_i * dpi_k);
        const double a4 = -(a1 + a2 + a3 + a5 + a6 + a7);
        a[0 * N + n] = a1;
        a[1 * N + n] = a2;
        a[2 * N + n] = a3;
        a[3 * N + n] = a4;
        a[4 * N + n] = a5;
        a[5 * N + n] = a6;
        a[6 * N + n] = a7;
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* set_boundary_points: src/pot3d.F90:5376-5487 (one rank: rb0, rb1, tb0,
 * tb1 all true; sum_over_phi is then the identity).  br1 may be NULL
 * (the validation-run Neumann branch at r1, :5420-5423, uses it).       */
void orc_set_boundary_points(const orc_grid *g, const double *br0, const double *br1,
                             double vmask, double *x) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const double dr1 = g->dr[0], drn = g->dr[nr - 2];
  const double pli = pl_i();
  for (int k = 2; k <= np - 1; k++)
    for (int j = 2; j <= nt - 1; j++) {
      const double b0 = br0 ? br0[(j - 1) + (size_t)nt * (k - 1)] : 0.0;
      x[IDX3(1, j, k)] = x[IDX3(2, j, k)] - vmask * b0 * dr1;
    }
  for (int k = 2; k <= np - 1; k++)
    for (int j = 2; j <= nt - 1; j++) {
      if (br1)
        x[IDX3(nr, j, k)] = x[IDX3(nr - 1, j, k)] + vmask * br1[(j - 1) + (size_t)nt * (k - 1)] * drn;
      else
        x[IDX3(nr, j, k)] = g->pm_r1 * x[IDX3(nr - 1, j, k)];
    }
  double *sum0 = (double *)malloc(sizeof(double) * nr);
  double *sum1 = (double *)malloc(sizeof(double) * nr);
  for (int i = 1; i <= nr; i++) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 2; k <= np - 1; k++) {
      s0 = s0 + x[IDX3(i, 2, k)] * g->dph[k - 1] * pli;
      s1 = s1 + x[IDX3(i, nt - 1, k)] * g->dph[k - 1] * pli;
    }
    sum0[i - 1] = s0;
    sum1[i - 1] = s1;
  }
  for (int k = 2; k <= np - 1; k++)
    for (int i = 1; i <= nr; i++) {
      x[IDX3(i, 1, k)] = 2.0 * sum0[i - 1] - x[IDX3(i, 2, k)];
      x[IDX3(i, nt, k)] = 2.0 * sum1[i - 1] - x[IDX3(i, nt - 1, k)];
    }
  free(sum0);
  free(sum1);
}

/* seam_hhh, periodic phi at one rank: src/pot3d.F90:5738-5758. */
void orc_seam(const orc_grid *g, double *x) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const size_t plane = (size_t)nr * nt;
  memcpy(x + 0 * plane, x + (size_t)(np - 2) * plane, plane * sizeof(double));
  memcpy(x + (size_t)(np - 1) * plane, x + 1 * plane, plane * sizeof(double));
}

/* unpack_scalar: src/pot3d.F90:5304-5335. */
void orc_unpack(const orc_grid *g, const double *xp, double *s) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2;
  long k;
#pragma omp parallel for schedule(static)
  for (k = 2; k <= np - 1; k++)
    for (int j = 2; j <= nt - 1; j++)
      for (int i = 2; i <= nr - 1; i++) s[IDX3(i, j, k)] = xp[PIDX(i, j, k)];
}

/* delsq: src/pot3d.F90:5337-5374. */
void orc_delsq(const orc_grid *g, const double *a, const double *x, double *y) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2, C = np - 2;
  const size_t N = (size_t)A * B * C;
  long k;
#pragma omp parallel for schedule(static)
  for (k = 2; k <= np - 1; k++)
    for (int j = 2; j <= nt - 1; j++)
      for (int i = 2; i <= nr - 1; i++) {
        const size_t n = PIDX(i, j, k);
        y[n] = a[0 * N + n] * x[IDX3(i, j, k - 1)] + a[1 * N + n] * x[IDX3(i, j - 1, k)] +
               a[2 * N + n] * x[IDX3(i - 1, j, k)] + a[3 * N + n] * x[IDX3(i, j, k)] +
               a[4 * N + n] * x[IDX3(i + 1, j, k)] + a[5 * N + n] * x[IDX3(i, j + 1, k)] +
               a[6 * N + n] * x[IDX3(i, j, k + 1)];
      }
}

/* ax: src/pot3d.F90:5102-5150.  xa is the padded scratch x_ax(nr,nt,np). */
void orc_ax(const orc_grid *g, const double *a, const double *x, double *y, double *xa) {
  orc_unpack(g, x, xa);
  orc_set_boundary_points(g, NULL, NULL, 0.0, xa);
  orc_seam(g, xa);
  orc_delsq(g, a, xa, y);
}

/* RHS "boundary trick": src/pot3d.F90:3941-3950, starting from x_ax = 0. */
void orc_rhs(const orc_grid *g, const double *a, const double *br0, const double *br1,
             double *rhs, double *xa) {
  const size_t tot = (size_t)g->nr * g->nt * g->np;
  const size_t N = (size_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  memset(xa, 0, tot * sizeof(double));
  orc_set_boundary_points(g, br0, br1, 1.0, xa);
  orc_seam(g, xa);
  orc_delsq(g, a, xa, rhs);
  for (size_t n = 0; n < N; n++) rhs[n] = -rhs[n];
}

/* Post-solve phi: unpack, inhomogeneous BCs, seam (src/pot3d.F90:3966-3969). */
void orc_fill_phi(const orc_grid *g, const double *x, const double *br0, const double *br1,
                  double *phi) {
  orc_unpack(g, x, phi);
  orc_set_boundary_points(g, br0, br1, 1.0, phi);
  orc_seam(g, phi);
}

/* ---------------------------------------------------------------------- */
/* Diagonal preconditioner: a_i = real(one/a4, REAL32) (src/pot3d.F90:
 * 4649-4661); apply x = a_i*x (:5189-5191).                             */
void orc_diag_prec(const orc_grid *g, const double *a, float *a_i) {
  const size_t N = (size_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  for (size_t n = 0; n < N; n++) a_i[n] = (float)(1.0 / a[3 * N + n]);
}

/* ---------------------------------------------------------------------- */
/* getM: src/pot3d.F90:5020-5100 (nproc_p = 1).  IA is 1-based, N+1 long;
 * returns M = IA(N+1) = nnz+1, the reference's allocated length.        */
int64_t orc_getM(const orc_grid *g, int *IA) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  size_t i = 0;
  IA[0] = 1;
  for (int mk = 2; mk <= np - 1; mk++)
    for (int mj = 2; mj <= nt - 1; mj++)
      for (int mi = 2; mi <= nr - 1; mi++) {
        int ok[7] = {1, 1, 1, 1, 1, 1, 1};
        if (mi == 2) ok[2] = 0;
        if (mi == nr - 1) ok[4] = 0;
        if (mj == 2) ok[1] = 0;
        if (mj == nt - 1) ok[5] = 0;
        int cnt = 0;
        for (int jj = 0; jj < 7; jj++) cnt += ok[jj];
        IA[i + 1] = IA[i] + cnt;
        i++;
      }
  return IA[i];
}

/* diacsr: src/pot3d.F90:4870-5018 (nproc_p = 1: periodic wrap kept).
 * Acsr FP32, JA and Adptr 1-based.                                       */
void orc_diacsr(const orc_grid *g, const double *a, const int *IA, float *Acsr, int *JA,
                int *Adptr) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2, C = np - 2;
  const int64_t N = (int64_t)A * B * C;
  const int64_t ioff[7] = {-(int64_t)A * B, -A, -1, 0, 1, A, (int64_t)A * B};
  for (int mk = 2; mk <= np - 1; mk++)
    for (int mj = 2; mj <= nt - 1; mj++)
      for (int mi = 2; mi <= nr - 1; mi++) {
        const int64_t i = (int64_t)(mk - 2) * B * A + (int64_t)(mj - 2) * A + (mi - 1);
        int64_t ko = IA[i - 1];
        int ok[7] = {1, 1, 1, 1, 1, 1, 1};
        if (mi == 2) ok[2] = 0;
        if (mi == nr - 1) ok[4] = 0;
        if (mj == 2) ok[1] = 0;
        if (mj == nt - 1) ok[5] = 0;
        for (int jj = 0; jj < 7; jj++)
          if (ok[jj]) {
            int64_t j = i + ioff[jj];
            if (j > N) {
              j -= N;
              Acsr[ko - 1] = (float)a[jj * N + (i - 1)];
              JA[ko - 1] = (int)j;
              ko++;
            }
          }
        for (int jj = 0; jj < 7; jj++)
          if (ok[jj]) {
            int64_t j = i + ioff[jj];
            if (j >= 1 && j <= N) {
              if (jj == 3) Adptr[i - 1] = (int)ko;
              Acsr[ko - 1] = (float)a[jj * N + (i - 1)];
              JA[ko - 1] = (int)j;
              ko++;
            }
          }
        for (int jj = 0; jj < 7; jj++)
          if (ok[jj]) {
            int64_t j = i + ioff[jj];
            if (j < 1) {
              j += N;
              Acsr[ko - 1] = (float)a[jj * N + (i - 1)];
              JA[ko - 1] = (int)j;
              ko++;
            }
          }
        (void)C;
      }
}

/* ilu0 (Saad IKJ, in place, FP32): src/pot3d.F90:4717-4813.
 * Returns icode (0, or the 1-based row of a zero pivot).                 */
int orc_ilu0(int64_t N, float *Av, const int *JA, const int *IA, const int *A_da) {
  int64_t *iw = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
  int icode = 0;
  for (int64_t i = 2; i <= N; i++) {
    const int64_t IA_i = IA[i - 1], IA_ip1m1 = IA[i] - 1;
    for (int64_t ij = IA_i; ij <= IA_ip1m1; ij++) iw[JA[ij - 1]] = ij;
    int64_t ik;
    for (ik = IA_i; ik <= A_da[i - 1] - 1; ik++) {
      const int64_t k = JA[ik - 1];
      const float Aik = Av[ik - 1] / Av[A_da[k - 1] - 1];
      Av[ik - 1] = Aik;
      for (int64_t kj = A_da[k - 1] + 1; kj <= IA[k] - 1; kj++) {
        const int64_t ij = iw[JA[kj - 1]];
        if (ij != 0) Av[ij - 1] = Av[ij - 1] - Aik * Av[kj - 1];
      }
    }
    /* after the loop ik == A_da(i): the diagonal (Fortran DO-variable rule) */
    if (Av[ik - 1] == 0.0f) {
      icode = (int)i;
      break;
    }
    for (int64_t ij = IA_i; ij <= IA_ip1m1; ij++) iw[JA[ij - 1]] = 0;
  }
  free(iw);
  return icode;
}

/* lu2luopt: src/pot3d.F90:4815-4868. */
void orc_lu2luopt(int64_t N, float *LU, const float *Av, const int *IA, const int *JA, int *LUJA,
                  const int *A_da, int *N1, int *N2) {
  int64_t ii = 0;
  for (int64_t i = 1; i <= N; i++) {
    for (int64_t k = IA[i - 1]; k <= A_da[i - 1] - 1; k++) {
      ii++;
      LU[ii - 1] = Av[k - 1];
      LUJA[ii - 1] = JA[k - 1];
    }
    N1[i - 1] = A_da[i - 1] - 1 - IA[i - 1];
    N2[i - 1] = IA[i] - 2 - A_da[i - 1];
  }
  for (int64_t i = N; i >= 1; i--)
    for (int64_t k = A_da[i - 1] + 1; k <= IA[i] - 1; k++) {
      ii++;
      LU[ii - 1] = Av[k - 1];
      LUJA[ii - 1] = JA[k - 1];
    }
}

/* lusol (forward unit-L, backward U with inverse diagonal), FP32:
 * src/pot3d.F90:5227-5302.                                               */
void orc_lusol(int64_t N, float *x, const float *LU, const int *LUJA, const int *N1,
               const int *N2, const float *LUd_i) {
  int64_t k1, k2 = 0;
  for (int64_t i = 1; i <= N; i++) {
    k1 = k2 + 1;
    k2 = k1 + N1[i - 1];
    for (int64_t k = k1; k <= k2; k++) x[i - 1] = x[i - 1] - LU[k - 1] * x[LUJA[k - 1] - 1];
  }
  for (int64_t i = N; i >= 1; i--) {
    k1 = k2 + 1;
    k2 = k1 + N2[i - 1];
    for (int64_t k = k1; k <= k2; k++) x[i - 1] = x[i - 1] - LU[k - 1] * x[LUJA[k - 1] - 1];
    x[i - 1] = x[i - 1] * LUd_i[i - 1];
  }
}

/* ---------------------------------------------------------------------- */
/* D-ILU: on the reference's ILU pattern (7-point stencil, r/theta
 * boundary couplings dropped, phi wrap kept; np >= 6) ILU(0) fill lands
 * only on the diagonal, so L and U keep A's off-diagonals and only the
 * pivots change:  d_p = a4_p - sum_{q in lower(p)} (a_pq/d_q)*a_qp,
 * summed in the column order ilu0 uses (wrap, k-1, j-1, i-1).
 * FP64 here; the product path stores dinv = real(1/d, REAL32).          */
/* Block-Jacobi variant: the grid is cut into blocks of bi x bj x bk points
 * (block starts at multiples of bi, bj, bk) and every coupling between two
 * blocks is dropped from the ILU matrix -- the reference's per-rank ILU0 on
 * a decomposed grid (diacsr drops the couplings at each rank's local r and
 * theta boundaries, and the phi ones when nproc_p > 1, :4940-4967).  The
 * phi wrap is kept only when there is a single phi block (bk >= C).
 * bi >= A, bj >= B, bk >= C is the unblocked D-ILU above.                */
static int same_lo(int x, int b) { return x > 0 && (x % b) != 0; }
static int same_hi(int x, int n, int b) { return x < n - 1 && ((x + 1) % b) != 0; }

void orc_dilu_pivots_blk(const orc_grid *g, const double *a, int bi, int bj, int bk,
                         double *d) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2, C = np - 2;
  const size_t N = (size_t)A * B * C;
  const size_t sj = (size_t)A, sk = (size_t)A * B;
  const int wrap = bk >= C;
  for (int k = 0; k < C; k++)
    for (int j = 0; j < B; j++)
      for (int i = 0; i < A; i++) {
        const size_t p = i + sj * j + sk * k;
        double dp_ = a[3 * N + p];
        if (wrap && k == C - 1) { /* wrap: q=(i,j,0); a_pq = a7(p), a_qp = a1(q) */
          const size_t q = i + sj * j;
          dp_ = dp_ - a[6 * N + p] / d[q] * a[0 * N + q];
        }
        if (same_lo(k, bk)) { /* q=(i,j,k-1): a_pq = a1(p), a_qp = a7(q) */
          const size_t q = p - sk;
          dp_ = dp_ - a[0 * N + p] / d[q] * a[6 * N + q];
        }
        if (same_lo(j, bj)) { /* q=(i,j-1,k): a2(p), a6(q) */
          const size_t q = p - sj;
          dp_ = dp_ - a[1 * N + p] / d[q] * a[5 * N + q];
        }
        if (same_lo(i, bi)) { /* q=(i-1,j,k): a3(p), a5(q) */
          const size_t q = p - 1;
          dp_ = dp_ - a[2 * N + p] / d[q] * a[4 * N + q];
        }
        d[p] = dp_;
      }
}

void orc_dilu_pivots(const orc_grid *g, const double *a, double *d) {
  orc_dilu_pivots_blk(g, a, g->nr, g->nt, g->np, d);
}

/* D-ILU apply in FP32, the arithmetic of the product's trisolve kernels:
 *   forward  y_p = r_p - sum_{q<p} a_pq * (dinv_q * y_q)   (wrap,k-1,j-1,i-1)
 *   backward z_p = (y_p - sum_{q>p} a_pq * z_q) * dinv_p    (i+1,j+1,k+1,wrap)
 * with a_pq = real(a, REAL32); couplings between blocks dropped as in
 * orc_dilu_pivots_blk.  x is overwritten (in: r, out: z).                */
void orc_dilu_apply_blk(const orc_grid *g, const double *a, const float *dinv, int bi, int bj,
                        int bk, float *x) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int A = nr - 2, B = nt - 2, C = np - 2;
  const size_t N = (size_t)A * B * C;
  const size_t sj = (size_t)A, sk = (size_t)A * B;
  const int wrap = bk >= C;
  for (int k = 0; k < C; k++)
    for (int j = 0; j < B; j++)
      for (int i = 0; i < A; i++) {
        const size_t p = i + sj * j + sk * k;
        float y = x[p];
        if (wrap && k == C - 1) y = y - (float)a[6 * N + p] * (dinv[i + sj * j] * x[i + sj * j]);
        if (same_lo(k, bk)) y = y - (float)a[0 * N + p] * (dinv[p - sk] * x[p - sk]);
        if (same_lo(j, bj)) y = y - (float)a[1 * N + p] * (dinv[p - sj] * x[p - sj]);
        if (same_lo(i, bi)) y = y - (float)a[2 * N + p] * (dinv[p - 1] * x[p - 1]);
        x[p] = y;
      }
  for (int k = C - 1; k >= 0; k--)
    for (int j = B - 1; j >= 0; j--)
      for (int i = A - 1; i >= 0; i--) {
        const size_t p = i + sj * j + sk * k;
        float s = x[p];
        if (same_hi(i, A, bi)) s = s - (float)a[4 * N + p] * x[p + 1];
        if (same_hi(j, B, bj)) s = s - (float)a[5 * N + p] * x[p + sj];
        if (same_hi(k, C, bk)) s = s - (float)a[6 * N + p] * x[p + sk];
        if (wrap && k == 0) s = s - (float)a[0 * N + p] * x[i + sj * j + sk * (C - 1)];
        x[p] = s * dinv[p];
      }
}

void orc_dilu_apply(const orc_grid *g, const double *a, const float *dinv, float *x) {
  orc_dilu_apply_blk(g, a, dinv, g->nr, g->nt, g->np, x);
}

/* ---------------------------------------------------------------------- */
/* cgsolve + ernorm: src/pot3d.F90:4231-4485.
 * ifprec: 0 none, 1 diagonal, 2 ILU0 via diacsr+ilu0+lusol (the
 * reference's CPU path, :4684-4709, :5202-5220), 3 D-ILU (FP64 pivots,
 * FP32 apply; the product's ILU0 arithmetic).
 * x: initial guess in, solution out.  r: rhs in, residual out.
 * hist (optional, length nhist): epsn recorded at every iteration.
 * Returns ierr (0 converged, 1 ncgmax reached); *ncg_out, *epsn_out.     */
typedef struct {
  int ifprec;
  const orc_grid *g;
  const double *a;
  float *a_i;                    /* ifprec 1 */
  float *lu, *lud_i;             /* ifprec 2 */
  int *luja, *n1, *n2;
  float *dinv;                   /* ifprec 3 */
  float *x32;
} orc_prec;

static void prec_apply(orc_prec *P, double *x, size_t N) {
  if (P->ifprec == 1) {
    for (size_t n = 0; n < N; n++) x[n] = (double)P->a_i[n] * x[n];
  } else if (P->ifprec == 2) {
    for (size_t n = 0; n < N; n++) P->x32[n] = (float)x[n];
    orc_lusol((int64_t)N, P->x32, P->lu, P->luja, P->n1, P->n2, P->lud_i);
    for (size_t n = 0; n < N; n++) x[n] = (double)P->x32[n];
  } else if (P->ifprec == 3) {
    for (size_t n = 0; n < N; n++) P->x32[n] = (float)x[n];
    orc_dilu_apply(P->g, P->a, P->dinv, P->x32);
    for (size_t n = 0; n < N; n++) x[n] = (double)P->x32[n];
  }
}

static int prec_setup(orc_prec *P, const orc_grid *g, const double *a, int ifprec) {
  const size_t N = (size_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  memset(P, 0, sizeof(*P));
  P->ifprec = ifprec;
  P->g = g;
  P->a = a;
  if (ifprec == 1) {
    P->a_i = (float *)malloc(N * sizeof(float));
    orc_diag_prec(g, a, P->a_i);
  } else if (ifprec == 2) {
    int *IA = (int *)malloc((N + 1) * sizeof(int));
    const int64_t M = orc_getM(g, IA);
    float *Av = (float *)malloc((size_t)M * sizeof(float));
    int *JA = (int *)malloc((size_t)M * sizeof(int));
    int *dptr = (int *)malloc(N * sizeof(int));
    orc_diacsr(g, a, IA, Av, JA, dptr);
    int icode = orc_ilu0((int64_t)N, Av, JA, IA, dptr);
    P->lu = (float *)malloc((size_t)M * sizeof(float));
    P->luja = (int *)malloc((size_t)M * sizeof(int));
    P->n1 = (int *)malloc(N * sizeof(int));
    P->n2 = (int *)malloc(N * sizeof(int));
    P->lud_i = (float *)malloc(N * sizeof(float));
    orc_lu2luopt((int64_t)N, P->lu, Av, IA, JA, P->luja, dptr, P->n1, P->n2);
    for (size_t n = 0; n < N; n++) P->lud_i[n] = (float)(1.0 / (double)Av[dptr[n] - 1]);
    free(IA);
    free(Av);
    free(JA);
    free(dptr);
    P->x32 = (float *)malloc(N * sizeof(float));
    if (icode != 0) return icode; /* reference prints and continues */
  } else if (ifprec == 3) {
    double *d = (double *)malloc(N * sizeof(double));
    orc_dilu_pivots(g, a, d);
    P->dinv = (float *)malloc(N * sizeof(float));
    for (size_t n = 0; n < N; n++) P->dinv[n] = (float)(1.0 / d[n]);
    free(d);
    P->x32 = (float *)malloc(N * sizeof(float));
  }
  return 0;
}

static void prec_free(orc_prec *P) {
  free(P->a_i);
  free(P->lu);
  free(P->lud_i);
  free(P->luja);
  free(P->n1);
  free(P->n2);
  free(P->dinv);
  free(P->x32);
}

/* Apply the chosen preconditioner once (for per-routine parity tests). */
int orc_prec_apply(const orc_grid *g, const double *a, int ifprec, double *x) {
  orc_prec P;
  const size_t N = (size_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  int rc = prec_setup(&P, g, a, ifprec);
  prec_apply(&P, x, N);
  prec_free(&P);
  return rc;
}

static double ernorm(double bdotb, double rdotr) {
  const double epssq = rdotr / bdotb;
  return copysign(1.0, epssq) * sqrt(fabs(epssq));
}

int orc_cgsolve(const orc_grid *g, const double *a, int ifprec, double epscg, int ncgmax,
                double *x, double *r, int *ncg_out, double *epsn_out, double *hist, int nhist) {
  const size_t N = (size_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  const size_t tot = (size_t)g->nr * g->nt * g->np;
  double *p = (double *)malloc(N * sizeof(double));
  double *ap = (double *)malloc(N * sizeof(double));
  double *xa = (double *)calloc(tot, sizeof(double));
  orc_prec P;
  prec_setup(&P, g, a, ifprec);
  int ncg = 0, ierr = -1;
  double epsn = 0.0;
  memcpy(p, r, N * sizeof(double));
  prec_apply(&P, p, N);
  const double bdotb = det_dot(r, p, N);
  if (bdotb == 0.0) {
    memset(x, 0, N * sizeof(double));
    ierr = 0;
    epsn = 0.0;
    goto done;
  }
  orc_ax(g, a, x, ap, xa);
  for (size_t n = 0; n < N; n++) {
    r[n] = r[n] - ap[n];
    p[n] = r[n];
  }
  prec_apply(&P, p, N);
  double rdotr = det_dot(r, p, N);
  epsn = ernorm(bdotb, rdotr);
  if (hist && nhist > 0) hist[0] = epsn;
  if (epsn < epscg) {
    ierr = 0;
    goto done;
  } else if (ncg >= ncgmax) {
    ierr = 1;
    goto done;
  }
  for (;;) {
    ncg++;
    orc_ax(g, a, p, ap, xa);
    const double pdotap = det_dot(p, ap, N);
    const double alphai = rdotr / pdotap;
    long n;
#pragma omp parallel for schedule(static)
    for (n = 0; n < (long)N; n++) {
      x[n] = x[n] + alphai * p[n];
      r[n] = r[n] - alphai * ap[n];
      ap[n] = r[n];
    }
    prec_apply(&P, ap, N);
    const double rdotr_old = rdotr;
    rdotr = det_dot(r, ap, N);
    epsn = ernorm(bdotb, rdotr);
    if (hist && ncg < nhist) hist[ncg] = epsn;
    if (epsn < epscg) {
      ierr = 0;
      break;
    } else if (ncg >= ncgmax) {
      ierr = 1;
      break;
    }
    const double betai = rdotr / rdotr_old;
#pragma omp parallel for schedule(static)
    for (n = 0; n < (long)N; n++) p[n] = betai * p[n] + ap[n];
  }
done:
  prec_free(&P);
  free(p);
  free(ap);
  free(xa);
  *ncg_out = ncg;
  *epsn_out = epsn;
  return ierr;
}

/* ---------------------------------------------------------------------- */
/* getb: src/pot3d.F90:6187-6233.  br(nr-1,nt,np), bt(nr,nt-1,np),
 * bp(nr,nt,np-1), Fortran order.                                         */
void orc_getb(const orc_grid *g, const double *phi, double *br, double *bt, double *bp) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int nrm1 = nr - 1, ntm1 = nt - 1, npm1 = np - 1;
  for (int k = 1; k <= np; k++)
    for (int j = 1; j <= nt; j++)
      for (int i = 1; i <= nrm1; i++)
        br[(i - 1) + (size_t)nrm1 * ((j - 1) + (size_t)nt * (k - 1))] =
            (phi[IDX3(i + 1, j, k)] - phi[IDX3(i, j, k)]) / g->dr[i - 1];
  for (int k = 1; k <= np; k++)
    for (int j = 1; j <= ntm1; j++)
      for (int i = 1; i <= nr; i++)
        bt[(i - 1) + (size_t)nr * ((j - 1) + (size_t)ntm1 * (k - 1))] =
            (phi[IDX3(i, j + 1, k)] - phi[IDX3(i, j, k)]) / (g->rh[i - 1] * g->dt[j - 1]);
  for (int k = 1; k <= npm1; k++)
    for (int j = 1; j <= nt; j++)
      for (int i = 1; i <= nr; i++)
        bp[(i - 1) + (size_t)nr * ((j - 1) + (size_t)nt * (k - 1))] =
            (phi[IDX3(i, j, k + 1)] - phi[IDX3(i, j, k)]) /
            (g->rh[i - 1] * g->sth[j - 1] * g->dp[k - 1]);
}

/* magnetic_energy: src/pot3d.F90:6235-6321.  w[0..2] = Br^2, Bt^2, Bp^2. */
void orc_energy(const orc_grid *g, const double *br, const double *bt, const double *bp,
                double *w) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  const int nrm1 = nr - 1, ntm1 = nt - 1;
  double wr = 0.0, wt = 0.0, wp = 0.0;
#define BR(i, j, k) br[((i)-1) + (size_t)nrm1 * (((j)-1) + (size_t)nt * ((k)-1))]
#define BT(i, j, k) bt[((i)-1) + (size_t)nr * (((j)-1) + (size_t)ntm1 * ((k)-1))]
#define BP(i, j, k) bp[((i)-1) + (size_t)nr * (((j)-1) + (size_t)nt * ((k)-1))]
  for (int k = 2; k <= np - 1; k++)
    for (int j = 2; j <= nt - 1; j++)
      for (int i = 2; i <= nr - 1; i++) {
        const double rh = g->rh[i - 1];
        const double dv = rh * rh * g->drh[i - 1] * g->dth[j - 1] * g->sth[j - 1] * g->dph[k - 1];
        const double brav = 0.5 * (BR(i, j, k) + BR(i - 1, j, k));
        const double btav = 0.5 * (BT(i, j, k) + BT(i, j - 1, k));
        const double bpav = 0.5 * (BP(i, j, k) + BP(i, j, k - 1));
        wr = wr + 0.5 * (brav * brav) * dv;
        wt = wt + 0.5 * (btav * btav) * dv;
        wp = wp + 0.5 * (bpav * bpav) * dv;
      }
#undef BR
#undef BT
#undef BP
  w[0] = wr;
  w[1] = wt;
  w[2] = wp;
}

/* One call: phi -> energies (getb + magnetic_energy). */
void orc_phi_energy(const orc_grid *g, const double *phi, double *w) {
  const int nr = g->nr, nt = g->nt, np = g->np;
  double *br = (double *)malloc(sizeof(double) * (size_t)(nr - 1) * nt * np);
  double *bt = (double *)malloc(sizeof(double) * (size_t)nr * (nt - 1) * np);
  double *bp = (double *)malloc(sizeof(double) * (size_t)nr * nt * (np - 1));
  orc_getb(g, phi, br, bt, bp);
  orc_energy(g, br, bt, bp, w);
  free(br);
  free(bt);
  free(bp);
}

/* ---------------------------------------------------------------------- */
/* The reference's CPU solve with ifprec=2 on P MPI ranks, restated with
 * one OpenMP thread per rank (shared memory in place of messages):
 *  - rank topology nb[0] x nb[1] x nb[2] and each rank's interior index
 *    ranges: set_proc_topology (src/pot3d.F90:1386-1530) and
 *    decompose_dimension (:2116-2167), computed by the caller; lo[d] holds
 *    nb[d]+1 block starts (0-based interior indices, lo[d][nb[d]] = extent);
 *  - ax on the whole grid with the stored coefficients (orc_ax), the CG
 *    vector updates and the dot products over all ranks (:4231-4369);
 *  - prec_inv with ifprec=2 per rank (:5202-5220): each rank's diacsr
 *    drops the couplings at its local r and theta boundaries, and its phi
 *    ones when nproc_p > 1 (:4940-4967), then ilu0 + lu2luopt
 *    (:4684-4709) and lusol on an FP32 copy of the rank's part of x.
 * stop_mode 0: epsn < epscg (ernorm, :4432-4433); 1: ||r||2/||b||2 of the
 * recurrence residual < epscg.  x: guess in, solution out; r: rhs in,
 * residual out.  t[0] = preconditioner set-up seconds, t[1] = iteration
 * seconds (both wall clock).  Returns ierr (0 converged, 1 ncgmax, 2 a
 * rank's ilu0 found a zero pivot).                                       */
#include <omp.h>

typedef struct {
  int i0, ni, j0, nj, k0, nk, drop_k;
  int64_t nloc;
  float *lu, *lud_i, *x32;
  int *luja, *n1, *n2;
  int icode;
} orc_rank;

/* diacsr of one rank's block (:4870-5018) in its local numbering. */
static int rank_setup(const orc_grid *g, const double *a, orc_rank *R) {
  const int A = g->nr - 2, B = g->nt - 2;
  const int64_t N = (int64_t)A * B * (g->np - 2);
  const int ni = R->ni, nj = R->nj, nk = R->nk;
  const int64_t nl = (int64_t)ni * nj * nk, ab = (int64_t)ni * nj;
  const int64_t ioff[7] = {-ab, -ni, -1, 0, 1, ni, ab};
  int *IA = (int *)malloc((size_t)(nl + 1) * sizeof(int));
  IA[0] = 1;
  int64_t p = 0;
  for (int mk = 0; mk < nk; mk++)
    for (int mj = 0; mj < nj; mj++)
      for (int mi = 0; mi < ni; mi++, p++) {
        int cnt = 7;
        if (mi == 0) cnt--;
        if (mi == ni - 1) cnt--;
        if (mj == 0) cnt--;
        if (mj == nj - 1) cnt--;
        if (R->drop_k && mk == 0) cnt--;
        if (R->drop_k && mk == nk - 1) cnt--;
        IA[p + 1] = IA[p] + cnt;
      }
  const int64_t M = IA[nl];
  float *Av = (float *)malloc((size_t)M * sizeof(float));
  int *JA = (int *)malloc((size_t)M * sizeof(int));
  int *dptr = (int *)malloc((size_t)nl * sizeof(int));
  p = 0;
  for (int mk = 0; mk < nk; mk++)
    for (int mj = 0; mj < nj; mj++)
      for (int mi = 0; mi < ni; mi++, p++) {
        const int64_t i = p + 1; /* 1-based local row */
        const int64_t ng = (int64_t)(R->i0 + mi) + (int64_t)A * ((R->j0 + mj) + (int64_t)B * (R->k0 + mk));
        int ok[7] = {1, 1, 1, 1, 1, 1, 1};
        if (mi == 0) ok[2] = 0;
        if (mi == ni - 1) ok[4] = 0;
        if (mj == 0) ok[1] = 0;
        if (mj == nj - 1) ok[5] = 0;
        if (R->drop_k && mk == 0) ok[0] = 0;
        if (R->drop_k && mk == nk - 1) ok[6] = 0;
        int64_t ko = IA[i - 1];
        for (int pass = 0; pass < 3; pass++) /* right wrap, in range, left wrap */
          for (int jj = 0; jj < 7; jj++) {
            if (!ok[jj]) continue;
            int64_t j = i + ioff[jj];
            const int where = j > nl ? 0 : (j >= 1 ? 1 : 2);
            if (where != pass) continue;
            if (pass == 0) j -= nl;
            if (pass == 2) j += nl;
            if (pass == 1 && jj == 3) dptr[i - 1] = (int)ko;
            Av[ko - 1] = (float)a[jj * N + ng];
            JA[ko - 1] = (int)j;
            ko++;
          }
      }
  R->nloc = nl;
  R->icode = orc_ilu0(nl, Av, JA, IA, dptr);
  R->lu = (float *)malloc((size_t)M * sizeof(float));
  R->luja = (int *)malloc((size_t)M * sizeof(int));
  R->n1 = (int *)malloc((size_t)nl * sizeof(int));
  R->n2 = (int *)malloc((size_t)nl * sizeof(int));
  R->lud_i = (float *)malloc((size_t)nl * sizeof(float));
  R->x32 = (float *)malloc((size_t)nl * sizeof(float));
  orc_lu2luopt(nl, R->lu, Av, IA, JA, R->luja, dptr, R->n1, R->n2);
  for (int64_t n = 0; n < nl; n++) R->lud_i[n] = (float)(1.0 / (double)Av[dptr[n] - 1]);
  free(IA);
  free(Av);
  free(JA);
  free(dptr);
  return R->icode;
}

/* prec_inv of one rank: x_32 = real(x), lusol, x = x_32 on its points. */
static void rank_apply(const orc_grid *g, orc_rank *R, double *x) {
  const int A = g->nr - 2, B = g->nt - 2;
  int64_t l = 0;
  for (int mk = 0; mk < R->nk; mk++)
    for (int mj = 0; mj < R->nj; mj++) {
      const double *xs = x + (R->i0 + (int64_t)A * ((R->j0 + mj) + (int64_t)B * (R->k0 + mk)));
      for (int mi = 0; mi < R->ni; mi++) R->x32[l++] = (float)xs[mi];
    }
  orc_lusol(R->nloc, R->x32, R->lu, R->luja, R->n1, R->n2, R->lud_i);
  l = 0;
  for (int mk = 0; mk < R->nk; mk++)
    for (int mj = 0; mj < R->nj; mj++) {
      double *xs = x + (R->i0 + (int64_t)A * ((R->j0 + mj) + (int64_t)B * (R->k0 + mk)));
      for (int mi = 0; mi < R->ni; mi++) xs[mi] = (double)R->x32[l++];
    }
}

static double par_dot(const double *x, const double *y, int64_t n) {
  double s = 0.0;
  int64_t i;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

int orc_cgsolve_ranks(const orc_grid *g, const double *a, const int *nb, const int *lo_r,
                      const int *lo_t, const int *lo_p, int stop_mode, double epscg, int ncgmax,
                      double *x, double *r, int *ncg_out, double *epsn_out, double *t) {
  const int64_t N = (int64_t)(g->nr - 2) * (g->nt - 2) * (g->np - 2);
  const size_t tot = (size_t)g->nr * g->nt * g->np;
  const int P = nb[0] * nb[1] * nb[2];
  omp_set_num_threads(P);
  double t0 = omp_get_wtime();
  orc_rank *R = (orc_rank *)calloc((size_t)P, sizeof(orc_rank));
  int bad = 0, q;
#pragma omp parallel for schedule(static, 1) reduction(| : bad)
  for (q = 0; q < P; q++) {
    const int br = q % nb[0], bt = (q / nb[0]) % nb[1], bp = q / (nb[0] * nb[1]);
    R[q].i0 = lo_r[br];
    R[q].ni = lo_r[br + 1] - lo_r[br];
    R[q].j0 = lo_t[bt];
    R[q].nj = lo_t[bt + 1] - lo_t[bt];
    R[q].k0 = lo_p[bp];
    R[q].nk = lo_p[bp + 1] - lo_p[bp];
    R[q].drop_k = nb[2] > 1;
    if (rank_setup(g, a, &R[q])) bad = 1;
  }
  double *p = (double *)malloc((size_t)N * sizeof(double));
  double *ap = (double *)malloc((size_t)N * sizeof(double));
  double *xa = (double *)calloc(tot, sizeof(double));
  double t1 = omp_get_wtime();
  t[0] = t1 - t0;
  int ncg = 0, ierr = 1;
  double epsn = 0.0, rdotr, bdotb, bb = 0.0, rr;
  int64_t n;
#define PREC(v)                                             \
  do {                                                      \
    _Pragma("omp parallel for schedule(static, 1)")         \
    for (q = 0; q < P; q++) rank_apply(g, &R[q], (v));      \
  } while (0)
  if (bad) {
    ierr = 2;
    goto done;
  }
  memcpy(p, r, (size_t)N * sizeof(double));
  PREC(p);
  bdotb = par_dot(r, p, N);
  bb = par_dot(r, r, N);
  if (bdotb == 0.0) {
    memset(x, 0, (size_t)N * sizeof(double));
    ierr = 0;
    goto done;
  }
  orc_ax(g, a, x, ap, xa);
#pragma omp parallel for schedule(static)
  for (n = 0; n < N; n++) {
    r[n] = r[n] - ap[n];
    p[n] = r[n];
  }
  PREC(p);
  rdotr = par_dot(r, p, N);
  rr = par_dot(r, r, N);
  for (;;) {
    epsn = ernorm(bdotb, rdotr);
    const double crit = stop_mode ? sqrt(rr / bb) : epsn;
    if (crit < epscg) {
      ierr = 0;
      break;
    }
    if (ncg >= ncgmax) {
      ierr = 1;
      break;
    }
    ncg++;
    orc_ax(g, a, p, ap, xa);
    const double alphai = rdotr / par_dot(p, ap, N);
#pragma omp parallel for schedule(static)
    for (n = 0; n < N; n++) {
      x[n] = x[n] + alphai * p[n];
      r[n] = r[n] - alphai * ap[n];
      ap[n] = r[n];
    }
    PREC(ap);
    const double rdotr_old = rdotr;
    rdotr = par_dot(r, ap, N);
    if (stop_mode) rr = par_dot(r, r, N);
    const double betai = rdotr / rdotr_old;
#pragma omp parallel for schedule(static)
    for (n = 0; n < N; n++) p[n] = betai * p[n] + ap[n];
  }
#undef PREC
done:
  t[1] = omp_get_wtime() - t1;
  for (q = 0; q < P; q++) {
    free(R[q].lu);
    free(R[q].lud_i);
    free(R[q].x32);
    free(R[q].luja);
    free(R[q].n1);
    free(R[q].n2);
  }
  free(R);
  free(p);
  free(ap);
  free(xa);
  *ncg_out = ncg;
  *epsn_out = epsn;
  return ierr;
}
