/*
 * CPU restatement of the ASTR per-timestep right-hand-side path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP solver.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each routine follows the reference routine cited in its comment, including
 * the evaluation order of its floating-point expressions (Fortran evaluates
 * a+b+c as (a+b)+c and a*b*c as (a*b)*c).  Paths are relative to the ASTR
 * source tree.
 *
 * Parity status of each part (how it is pinned): the line operators are
 * pinned bit-for-bit against the reference
 * modules compiled from source (oracle/_ref), the mini-app variant against
 * the reference-held series miniapps/tgv_solver_3d/state.ref_128, and the
 * main-path stage order (qswap averaging, stale primitives) by invariants
 * only ("parity unpinned" against a run of the full reference, which needs
 * MPI and cannot be built here).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HM ORC_HM

int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================= banded solver ======================= */

/* tridiagonal_thomas_proprocess, src/commfunc.F90:752-774.
 * Row 0 keeps its unit diagonal; ac(2,1), ac(3,1) are never read by the
 * reference solver, here they are set to 1 and 0 so that the same table can
 * drive a solver that treats row 0 like the others. */
void orc_thomas_prep(const double *a, const double *c, int n, double *ac)
{
    ac[0] = c[0];
    ac[1] = 1.0;
    ac[2] = 0.0;
    for (int i = 1; i < n; ++i) {
        ac[3 * i + 0] = c[i] / (1.0 - a[i] * ac[3 * (i - 1) + 0]);
        ac[3 * i + 1] = 1.0 / (1.0 - a[i] * ac[3 * (i - 1) + 0]);
        ac[3 * i + 2] = a[i] / (1.0 - a[i] * ac[3 * (i - 1) + 0]);
    }
}

/* tridiagonal_thomas_solver, src/commfunc.F90:790-813 */
void orc_thomas_solve(const double *ac, double *d, int n, double *x)
{
    for (int i = 1; i < n; ++i)
        d[i] = d[i] * ac[3 * i + 1] - d[i - 1] * ac[3 * i + 2];
    x[n - 1] = d[n - 1];
    for (int i = n - 2; i >= 0; --i)
        x[i] = d[i] - ac[3 * i + 0] * x[i + 1];
}

/* ======================= compact derivative ======================= */

static const double num1d3 = 1.0 / 3.0, num2d3 = 2.0 / 3.0,
                    num7d9 = 7.0 / 9.0, num1d36 = 1.0 / 36.0,
                    num1d60 = 1.0 / 60.0, num1d12 = 1.0 / 12.0;

static void node_range(int ntype, int dim, int lo_if, int hi_if,
                       int *first, int *last)
{
    /* ntype 1: boundary at 0; 2: boundary at dim; 3: inner; 4: both. */
    *first = (ntype == 1 || ntype == 4) ? 0 : -lo_if;
    *last = (ntype == 2 || ntype == 4) ? dim : dim + hi_if;
}

/* fd_scheme_initiate for '643c', src/derivative.F90:63-158 */
int orc_fd_init(int ntype, int dim, int *first, int *last,
                double *a, double *c, double *ac)
{
    int i0, im_;
    node_range(ntype, dim, 1, 1, &i0, &im_);
    int n = im_ - i0 + 1;
    for (int r = 0; r < n; ++r) {
        a[r] = num1d3;
        c[r] = num1d3;
    }
    a[0] = 0.0; c[0] = 0.0;          /* explicit central at interface */
    a[n - 1] = 0.0; c[n - 1] = 0.0;
    if (ntype == 1 || ntype == 4) {
        a[0] = 2.0; c[0] = 2.0;      /* 3rd-order biased compact */
        a[1] = 0.25; c[1] = 0.25;    /* 4th-order compact central */
    }
    if (ntype == 2 || ntype == 4) {
        a[n - 1] = 2.0; c[n - 1] = 2.0;
        a[n - 2] = 0.25; c[n - 2] = 0.25;
    }
    orc_thomas_prep(a, c, n, ac);
    *first = i0;
    *last = im_;
    return n;
}

/* compact_fd_rhs, src/derivative.F90:210-306.  fp(node) = f[node+HM];
 * d is indexed by row (node - i0). */
static void fd_rhs(int ntype, int i0, int im_, const double *f, double *d)
{
#define F(nd) f[(nd) + HM]
#define D(nd) d[(nd) - i0]
    int is, ie, j;
    if (ntype == 1 || ntype == 4) {
        is = i0 + 2;
        j = i0 + 1;
        D(j) = 0.75 * (F(j + 1) - F(j - 1));
        j = i0;
        D(j) = -2.5 * F(j) + 2.0 * F(j + 1) + 0.5 * F(j + 2);
    } else {
        is = i0 + 1;
        j = i0;
        double v1 = F(j + 1) - F(j - 1), v2 = F(j + 2) - F(j - 2),
               v3 = F(j + 3) - F(j - 3);
        D(j) = 0.75 * v1 - 0.15 * v2 + num1d60 * v3;
    }
    if (ntype == 2 || ntype == 4) {
        ie = im_ - 2;
        j = im_ - 1;
        D(j) = 0.75 * (F(j + 1) - F(j - 1));
        j = im_;
        D(j) = 2.5 * F(j) - 2.0 * F(j - 1) - 0.5 * F(j - 2);
    } else {
        ie = im_ - 1;
        j = im_;
        double v1 = F(j + 1) - F(j - 1), v2 = F(j + 2) - F(j - 2),
               v3 = F(j + 3) - F(j - 3);
        D(j) = 0.75 * v1 - 0.15 * v2 + num1d60 * v3;
    }
    for (j = is; j <= ie; ++j) {
        double v1 = F(j + 1) - F(j - 1), v2 = F(j + 2) - F(j - 2);
        D(j) = num7d9 * v1 + num1d36 * v2;
    }
#undef F
#undef D
}

typedef struct {
    int ntype, dim, first, last, n;
    double *ac;
} scheme_t;

static void scheme_fd(scheme_t *s, int ntype, int dim)
{
    int nmax = dim + 3;
    double *a = malloc(sizeof(double) * nmax), *c = malloc(sizeof(double) * nmax);
    s->ac = malloc(sizeof(double) * 3 * nmax);
    s->n = orc_fd_init(ntype, dim, &s->first, &s->last, a, c, s->ac);
    s->ntype = ntype;
    s->dim = dim;
    free(a);
    free(c);
}

/* df_compact, src/derivative.F90:171-198; work holds 2*n doubles */
static void line_df(const scheme_t *s, const double *f, double *df, double *work)
{
    double *d = work, *x = work + s->n;
    fd_rhs(s->ntype, s->first, s->last, f, d);
    orc_thomas_solve(s->ac, d, s->n, x);
    for (int i = 0; i <= s->dim; ++i) df[i] = x[i - s->first];
}

void orc_df_compact(int ntype, int dim, const double *f, double *df)
{
    scheme_t s;
    scheme_fd(&s, ntype, dim);
    double *work = malloc(sizeof(double) * 2 * s.n);
    line_df(&s, f, df, work);
    free(work);
    free(s.ac);
}

/* diff6ec, src/derivative.F90:350-413 */
void orc_diff6ec(int ntype, int dim, const double *vin, double *vout)
{
#define V(nd) vin[(nd) + HM]
    int i;
#define E6(i) (0.75 * (V(i + 1) - V(i - 1)) - 0.15 * (V(i + 2) - V(i - 2)) + \
               num1d60 * (V(i + 3) - V(i - 3)))
    switch (ntype) {
    case 1:
        vout[0] = -0.5 * V(2) + 2.0 * V(1) - 1.5 * V(0);
        vout[1] = 0.5 * (V(2) - V(0));
        vout[2] = num2d3 * (V(3) - V(1)) - num1d12 * (V(4) - V(0));
        for (i = 3; i <= dim; ++i) vout[i] = E6(i);
        break;
    case 2:
        for (i = 0; i <= dim - 3; ++i) vout[i] = E6(i);
        vout[dim - 2] = num2d3 * (V(dim - 1) - V(dim - 3)) -
                        num1d12 * (V(dim) - V(dim - 4));
        vout[dim - 1] = 0.5 * (V(dim) - V(dim - 2));
        vout[dim] = 0.5 * V(dim - 2) - 2.0 * V(dim - 1) + 1.5 * V(dim);
        break;
    case 3:
        for (i = 0; i <= dim; ++i) vout[i] = E6(i);
        break;
    case 4:
        vout[0] = -0.5 * V(2) + 2.0 * V(1) - 1.5 * V(0);
        vout[1] = 0.5 * (V(2) - V(0));
        vout[2] = num2d3 * (V(3) - V(1)) - num1d12 * (V(4) - V(0));
        for (i = 3; i <= dim - 3; ++i) vout[i] = E6(i);
        vout[dim - 2] = num2d3 * (V(dim - 1) - V(dim - 3)) -
                        num1d12 * (V(dim) - V(dim - 4));
        vout[dim - 1] = 0.5 * (V(dim) - V(dim - 2));
        vout[dim] = 0.5 * V(dim - 2) - 2.0 * V(dim - 1) + 1.5 * V(dim);
        break;
    }
#undef E6
#undef V
}

/* ======================= compact filter ======================= */

#define C10(b) ((b) + 0)
#define CB(b, k, j) ((b)[6 + (k) * 9 + (j)])
#define CH(b, k, j) ((b)[51 + (k) * 11 + (j)])
#define C6(b) ((b) + 106)
#define C8(b) ((b) + 110)

/* filter_coefficient_cal, src/filter.F90:299-432 (coef2i/coef4i unused on
 * the path and omitted). */
void orc_filter_coef(double alfa, double bh, double bb, double *cf)
{
    memset(cf, 0, sizeof(double) * ORC_NCOEF);
    double *c6 = C6(cf), *c8 = C8(cf), *c10 = C10(cf);
    c6[0] = (11.0 + 10.0 * alfa) / 32.0;
    c6[1] = (15.0 + 34.0 * alfa) / 64.0;
    c6[2] = (-3.0 + 6.0 * alfa) / 32.0;
    c6[3] = (1.0 - 2.0 * alfa) / 64.0;

    c8[0] = (93.0 + 70.0 * alfa) / 256.0;
    c8[1] = (7.0 + 18.0 * alfa) / 32.0;
    c8[2] = (-7.0 + 14.0 * alfa) / 64.0;
    c8[3] = (1.0 - 2.0 * alfa) / 32.0;
    c8[4] = (-1.0 + 2.0 * alfa) / 256.0;

    c10[0] = (193.0 + 126.0 * alfa) / 512.0;
    c10[1] = (105.0 + 302.0 * alfa) / 512.0;
    c10[2] = (-15.0 + 30.0 * alfa) / 128.0;
    c10[3] = (45.0 - 90.0 * alfa) / 1024.0;
    c10[4] = (-5.0 + 10.0 * alfa) / 512.0;
    c10[5] = (1.0 - 2.0 * alfa) / 1024.0;

    CB(cf, 3, 0) = (1.0 - 2.0 * alfa) / 256.0;
    CB(cf, 3, 1) = (-1.0 + 2.0 * alfa) / 32.0;
    CB(cf, 3, 2) = (7.0 + 50.0 * alfa) / 64.0;
    CB(cf, 3, 3) = (25.0 + 14.0 * alfa) / 32.0;
    CB(cf, 3, 4) = (35.0 + 58.0 * alfa) / 128.0;
    CB(cf, 3, 5) = (-7.0 + 14.0 * alfa) / 32.0;
    CB(cf, 3, 6) = (7.0 - 14.0 * alfa) / 64.0;
    CB(cf, 3, 7) = (-1.0 + 2.0 * alfa) / 32.0;
    CB(cf, 3, 8) = (1.0 - 2.0 * alfa) / 256.0;

    CB(cf, 2, 0) = (-1.0 + 2.0 * alfa) / 64.0;
    CB(cf, 2, 1) = (3.0 + 26.0 * alfa) / 32.0;
    CB(cf, 2, 2) = (49.0 + 30.0 * alfa) / 64.0;
    CB(cf, 2, 3) = (5.0 + 6.0 * alfa) / 16.0;
    CB(cf, 2, 4) = (-15.0 + 30.0 * alfa) / 64.0;
    CB(cf, 2, 5) = (3.0 - 6.0 * alfa) / 32.0;
    CB(cf, 2, 6) = (-1.0 + 2.0 * alfa) / 64.0;

    CB(cf, 1, 0) = (1.0 + 62.0 * alfa) / 64.0;
    CB(cf, 1, 1) = (29.0 + 6.0 * alfa) / 32.0;
    CB(cf, 1, 2) = (15.0 + 34.0 * alfa) / 64.0;
    CB(cf, 1, 3) = (-5.0 + 10.0 * alfa) / 16.0;
    CB(cf, 1, 4) = (15.0 - 30.0 * alfa) / 64.0;
    CB(cf, 1, 5) = (-3.0 + 6.0 * alfa) / 32.0;
    CB(cf, 1, 6) = (1.0 - 2.0 * alfa) / 64.0;

    CB(cf, 0, 0) = (63.0 + 1.0 * bb) / 64.0;
    CB(cf, 0, 1) = (3.0 + 29.0 * bb) / 32.0;
    CB(cf, 0, 2) = (-15.0 + 15.0 * bb) / 64.0;
    CB(cf, 0, 3) = (5.0 - 5.0 * bb) / 16.0;
    CB(cf, 0, 4) = (-15.0 + 15.0 * bb) / 64.0;
    CB(cf, 0, 5) = (3.0 - 3.0 * bb) / 32.0;
    CB(cf, 0, 6) = (-1.0 + 1.0 * bb) / 64.0;

    CH(cf, 0, 0) = (-1.0 + 1.0 * bh) / 1024.0;
    CH(cf, 0, 1) = (5.0 - 5.0 * bh) / 512.0;
    CH(cf, 0, 2) = (979.0 + 45.0 * bh) / 1024.0;
    CH(cf, 0, 3) = (15.0 + 113.0 * bh) / 128.0;
    CH(cf, 0, 4) = (-105.0 + 105.0 * bh) / 512.0;
    CH(cf, 0, 5) = (63.0 - 63.0 * bh) / 256.0;
    CH(cf, 0, 6) = (-105.0 + 105.0 * bh) / 512.0;
    CH(cf, 0, 7) = (15.0 - 15.0 * bh) / 128.0;
    CH(cf, 0, 8) = (-45.0 + 45.0 * bh) / 1024.0;
    CH(cf, 0, 9) = (5.0 - 5.0 * bh) / 512.0;
    CH(cf, 0, 10) = (-1.0 + 1.0 * bh) / 1024.0;

    /* The reference writes 1024., 512., 256. as default-real literals in
     * four of these entries; they are powers of two, exact in any kind. */
    CH(cf, 1, 0) = (1.0 - 2.0 * alfa) / 1024.0;
    CH(cf, 1, 1) = (-5.0 + 10.0 * alfa) / 512.0;
    CH(cf, 1, 2) = (45.0 + 934.0 * alfa) / 1024.0;
    CH(cf, 1, 3) = (113.0 + 30.0 * alfa) / 128.0;
    CH(cf, 1, 4) = (105.0 + 302.0 * alfa) / 512.0;
    CH(cf, 1, 5) = (-63.0 + 126.0 * alfa) / 256.0;
    CH(cf, 1, 6) = (105.0 - 210.0 * alfa) / 512.0;
    CH(cf, 1, 7) = (-15.0 + 30.0 * alfa) / 128.0;
    CH(cf, 1, 8) = (45.0 - 90.0 * alfa) / 1024.0;
    CH(cf, 1, 9) = (-5.0 + 10.0 * alfa) / 512.0;
    CH(cf, 1, 10) = (1.0 - 2.0 * alfa) / 1024.0;

    CH(cf, 2, 0) = (-1.0 + 2.0 * alfa) / 1024.0;
    CH(cf, 2, 1) = (5.0 - 10.0 * alfa) / 512.0;
    CH(cf, 2, 2) = (-45.0 + 90.0 * alfa) / 1024.0;
    CH(cf, 2, 3) = (15.0 + 98.0 * alfa) / 128.0;
    CH(cf, 2, 4) = (407.0 + 210.0 * alfa) / 512.0;
    CH(cf, 2, 5) = (63.0 + 130.0 * alfa) / 256.0;
    CH(cf, 2, 6) = (-105.0 + 210.0 * alfa) / 512.0;
    CH(cf, 2, 7) = (15.0 - 30.0 * alfa) / 128.0;
    CH(cf, 2, 8) = (-45.0 + 90.0 * alfa) / 1024.0;
    CH(cf, 2, 9) = (5.0 - 10.0 * alfa) / 512.0;
    CH(cf, 2, 10) = (-1.0 + 2.0 * alfa) / 1024.0;
}

/* compact_filter_initiate, src/filter.F90:31-100 (note absent).  The LHS
 * end-row betas are the hard-coded 1.11 (interface) / 0.98 (boundary) of
 * the reference, independent of filter_coefficient_cal's arguments. */
int orc_filter_init(int ntype, int dim, double alfa, int *first, int *last,
                    double *a, double *c, double *ac)
{
    int i0, im_;
    node_range(ntype, dim, 3, 3, &i0, &im_);
    double b0 = (ntype == 1 || ntype == 4) ? 0.98 : 1.11;
    double bm = (ntype == 2 || ntype == 4) ? 0.98 : 1.11;
    int n = im_ - i0 + 1;
    for (int r = 0; r < n; ++r) {
        a[r] = alfa;
        c[r] = alfa;
    }
    a[0] = b0; c[0] = b0;
    a[n - 1] = bm; c[n - 1] = bm;
    orc_thomas_prep(a, c, n, ac);
    *first = i0;
    *last = im_;
    return n;
}

static void scheme_filter(scheme_t *s, int ntype, int dim, double alfa)
{
    int nmax = dim + 7;
    double *a = malloc(sizeof(double) * nmax), *c = malloc(sizeof(double) * nmax);
    s->ac = malloc(sizeof(double) * 3 * nmax);
    s->n = orc_filter_init(ntype, dim, alfa, &s->first, &s->last, a, c, s->ac);
    s->ntype = ntype;
    s->dim = dim;
    free(a);
    free(c);
}

/* compact_filter_rhs, src/filter.F90:156-285 */
static void filter_rhs(int ntype, int i0, int im_, const double *cf,
                       const double *f, double *d)
{
#define F(nd) f[(nd) + HM]
#define D(nd) d[(nd) - i0]
    const double *c6 = C6(cf), *c8 = C8(cf), *c10 = C10(cf);
    int is, ie, j, k;
    if (ntype == 1 || ntype == 4) {
        is = i0 + 5;
        for (k = i0; k <= i0 + 2; ++k) {
            double v0 = 0.0;
            for (j = 0; j <= 6; ++j) v0 = v0 + CB(cf, k - i0, j) * F(i0 + j);
            D(k) = v0;
        }
        j = i0 + 3;
        D(j) = c6[0] * (F(j) + F(j)) + c6[1] * (F(j + 1) + F(j - 1)) +
               c6[2] * (F(j + 2) + F(j - 2)) + c6[3] * (F(j + 3) + F(j - 3));
        j = i0 + 4;
        D(j) = c8[0] * (F(j) + F(j)) + c8[1] * (F(j + 1) + F(j - 1)) +
               c8[2] * (F(j + 2) + F(j - 2)) + c8[3] * (F(j + 3) + F(j - 3)) +
               c8[4] * (F(j + 4) + F(j - 4));
    } else {
        is = i0 + 3;
        for (k = i0; k <= i0 + 2; ++k) {
            double v0 = 0.0;
            for (j = 0; j <= 10; ++j) v0 = v0 + CH(cf, k - i0, j) * F(i0 - 2 + j);
            D(k) = v0;
        }
    }
    if (ntype == 2 || ntype == 4) {
        ie = im_ - 5;
        for (k = im_ - 2; k <= im_; ++k) {
            double v0 = 0.0;
            for (j = 0; j <= 6; ++j) v0 = v0 + CB(cf, im_ - k, j) * F(im_ - j);
            D(k) = v0;
        }
        j = im_ - 3;
        D(j) = c6[0] * (F(j) + F(j)) + c6[1] * (F(j + 1) + F(j - 1)) +
               c6[2] * (F(j + 2) + F(j - 2)) + c6[3] * (F(j + 3) + F(j - 3));
        j = im_ - 4;
        D(j) = c8[0] * (F(j) + F(j)) + c8[1] * (F(j + 1) + F(j - 1)) +
               c8[2] * (F(j + 2) + F(j - 2)) + c8[3] * (F(j + 3) + F(j - 3)) +
               c8[4] * (F(j + 4) + F(j - 4));
    } else {
        ie = im_ - 3;
        for (k = im_ - 2; k <= im_; ++k) {
            double v0 = 0.0;
            for (j = 0; j <= 10; ++j) v0 = v0 + CH(cf, im_ - k, j) * F(im_ + 2 - j);
            D(k) = v0;
        }
    }
    for (j = is; j <= ie; ++j) {
        double v0 = F(j) + F(j), v1 = F(j + 1) + F(j - 1),
               v2 = F(j + 2) + F(j - 2), v3 = F(j + 3) + F(j - 3),
               v4 = F(j + 4) + F(j - 4), v5 = F(j + 5) + F(j - 5);
        D(j) = c10[0] * v0 + c10[1] * v1 + c10[2] * v2 + c10[3] * v3 +
               c10[4] * v4 + c10[5] * v5;
    }
#undef F
#undef D
}

/* compact_filter, src/filter.F90:112-144 */
static void line_filter(const scheme_t *s, const double *cf, const double *f,
                        double *ff, double *work)
{
    double *d = work, *x = work + s->n;
    filter_rhs(s->ntype, s->first, s->last, cf, f, d);
    orc_thomas_solve(s->ac, d, s->n, x);
    for (int i = 0; i <= s->dim; ++i) ff[i] = x[i - s->first];
    if (s->ntype == 1 || s->ntype == 4) ff[0] = f[HM];
    if (s->ntype == 2 || s->ntype == 4) ff[s->dim] = f[HM + s->dim];
}

void orc_compact_filter(int ntype, int dim, double alfa, double beta_halo,
                        double beta_bound, const double *f, double *ff)
{
    double cf[ORC_NCOEF];
    orc_filter_coef(alfa, beta_halo, beta_bound, cf);
    scheme_t s;
    scheme_filter(&s, ntype, dim, alfa);
    double *work = malloc(sizeof(double) * 2 * s.n);
    line_filter(&s, cf, f, ff, work);
    free(work);
    free(s.ac);
}

/* ======================= box (periodic, or bounded on chosen axes) ======================= */

struct orc_box {
    int im, jm, km, variant;
    int nt[3]; /* npdc{i,j,k}: 3 periodic, 1/2/4 bounded (single rank) */
    int ni, nj, nk;
    size_t n3;
    double *q[5], *rho, *vel[3], *prs, *tmp;
    double *qrhs[5], *qsave[5], *dvel[9], *dtmp[3];
    double *sigma[6], *qflux[3];
    double *jacob, *dxi[9]; /* dxi[(d-1)*3+(n-1)] = dxi(:,:,:,d,n) */
    scheme_t fd[3], fl[3];
    double coef[ORC_NCOEF];
    double reynolds, mach, prandtl, gamma, ref_tem, alfa;
    double const1, const2, const3, const4, const5, const6, const7;
    double tempconst, tempconst1, roinf, uinf, pinf, tinf;
    double lx, ly, lz, dx, dy, dz;
    int bc[6]; /* bctype of faces 1..6: 0 nothing done here, else boucon's */
    double twall[6];
    double xslip;
};

#define IX(b, i, j, k) ((size_t)((i) + HM) + (size_t)(b)->ni * ((size_t)((j) + HM) + (size_t)(b)->nj * (size_t)((k) + HM)))

static double *field(const orc_box *b)
{
    return calloc(b->n3, sizeof(double));
}

void orc_box_set_sutherland(orc_box *b, double s)
{
    b->tempconst = s / b->ref_tem;
    b->tempconst1 = 1.0 + b->tempconst;
}

orc_box *orc_box_new(int im, int jm, int km, int variant, double reynolds,
                     double mach, double prandtl, double gamma, double ref_tem,
                     double alfa, double lx, double ly, double lz)
{
    orc_box *b = calloc(1, sizeof(orc_box));
    for (int f = 0; f < 6; ++f) b->twall[f] = 1.0;
    b->im = im; b->jm = jm; b->km = km; b->variant = variant;
    b->ni = im + 1 + 2 * HM; b->nj = jm + 1 + 2 * HM; b->nk = km + 1 + 2 * HM;
    b->n3 = (size_t)b->ni * b->nj * b->nk;
    for (int m = 0; m < 5; ++m) {
        b->q[m] = field(b); b->qrhs[m] = field(b); b->qsave[m] = field(b);
    }
    b->rho = field(b); b->prs = field(b); b->tmp = field(b);
    for (int m = 0; m < 3; ++m) { b->vel[m] = field(b); b->dtmp[m] = field(b); b->qflux[m] = field(b); }
    for (int m = 0; m < 9; ++m) { b->dvel[m] = field(b); b->dxi[m] = field(b); }
    for (int m = 0; m < 6; ++m) b->sigma[m] = field(b);
    b->jacob = field(b);

    b->reynolds = reynolds; b->mach = mach; b->prandtl = prandtl;
    b->gamma = gamma; b->ref_tem = ref_tem; b->alfa = alfa;
    /* refcal, src/solver.F90:101-126 (nondimensional branch) */
    b->const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach));
    b->const2 = gamma * (mach * mach);
    b->const3 = (gamma - 1.0) / 3.0 * prandtl * (mach * mach);
    b->const4 = (gamma - 1.0) * (mach * mach) * reynolds * prandtl;
    b->const5 = (gamma - 1.0) * (mach * mach);
    b->const6 = 1.0 / (gamma - 1.0);
    b->const7 = (gamma - 1.0) * (mach * mach) * reynolds * prandtl;
    b->roinf = 1.0; b->uinf = 1.0; b->tinf = 1.0;
    b->pinf = b->roinf * b->tinf / b->const2;
    orc_box_set_sutherland(b, variant == 0 ? 110.3 : 110.4);  /* 2: the mini-app's */
    b->lx = lx; b->ly = ly; b->lz = lz;
    b->dx = lx / (double)im; b->dy = ly / (double)jm; b->dz = lz / (double)km;

    int dims[3] = {im, jm, km};
    for (int d = 0; d < 3; ++d) {
        b->nt[d] = 3;
        scheme_fd(&b->fd[d], 3, dims[d]);
        scheme_filter(&b->fl[d], 3, dims[d], alfa);
    }
    orc_filter_coef(alfa, 1.11, variant == 0 ? 0.98 : 1.09, b->coef);

    /* Metrics of the uniform periodic box: d(xi)/dx = 1/dx on the diagonal,
     * zero off it, J = dx*dy*dz (what gridgeom, src/geom.F90:99-917,
     * computes for gridcube up to round-off). */
    for (size_t p = 0; p < b->n3; ++p) {
        b->jacob[p] = b->dx * b->dy * b->dz;
        b->dxi[0][p] = 1.0 / b->dx;
        b->dxi[4][p] = 1.0 / b->dy;
        b->dxi[8][p] = 1.0 / b->dz;
    }
    return b;
}

static void free_scheme(scheme_t *s) { free(s->ac); }

/* The closure of each axis (npdc{i,j,k}, src/parallel.F90:1044-1249 on a
 * single rank): 3 periodic, 1/2/4 a boundary at 0 / at n / at both.  On a
 * bounded axis qswap and dataswap copy nothing (their single-rank branches
 * run only when the axis is periodic, lihomo/ljhomo/lkhomo:
 * src/parallel.F90:4877-4915, 4170-4175), and the line operators take the
 * axis's boundary closures.  Main variant only; returns -1 otherwise. */
int orc_box_set_ntype(orc_box *b, int n0, int n1, int n2)
{
    if (b->variant != 0) return -1;
    int nt[3] = {n0, n1, n2}, dims[3] = {b->im, b->jm, b->km};
    for (int d = 0; d < 3; ++d)
        if (nt[d] < 1 || nt[d] > 4) return -1;
    for (int d = 0; d < 3; ++d) {
        free_scheme(&b->fd[d]);
        free_scheme(&b->fl[d]);
        b->nt[d] = nt[d];
        scheme_fd(&b->fd[d], nt[d], dims[d]);
        scheme_filter(&b->fl[d], nt[d], dims[d], b->alfa);
    }
    return 0;
}

int orc_box_set_bc(orc_box *b, const int *bctype)
{
    if (b->variant != 0) return -1;
    for (int f = 0; f < 6; ++f)
        if (bctype[f] != 0 && bctype[f] != 23 && bctype[f] != 50 && bctype[f] != 60 &&
            !((bctype[f] == 41 || bctype[f] == 42) && f < 4) &&
            !(bctype[f] == 21 && (f == 1 || f == 3)) && bctype[f] != 31 &&
            !((bctype[f] == 421 || bctype[f] == 411) && (f == 2 || f == 3)))
            return -1;
    for (int f = 0; f < 6; ++f) b->bc[f] = bctype[f];
    return 0;
}

/* boucon (src/bc.F90:327-411) face by face in the order 1..6: gcnscbc
 * (src/bc.F90:7907-8140, faces 2 and 4), zeroextrap (src/bc.F90:2331-2627)
 * symmetry (src/bc.F90:2639-2971, the box's rectilinear metrics) and the
 * walls (src/bc.F90:6281-6919), the restatements of upwind.c run on this
 * box's fields (the same haloed
 * layout); gcnscbc's coordinates are the gridcube's x(i) = lx/im*i,
 * y(j) = ly/jm*j */
void orc_box_set_twall(orc_box *b, const double *twall)
{
    for (int f = 0; f < 6; ++f) b->twall[f] = twall[f];
}

void orc_box_set_xslip(orc_box *b, double xslip) { b->xslip = xslip; }

static void box_boucon(orc_box *b, double dt)
{
    int any = 0;
    for (int f = 0; f < 6; ++f) any |= b->bc[f] != 0;
    if (!any) return;
    orc_uwd u;
    memset(&u, 0, sizeof u);
    u.im = b->im; u.jm = b->jm; u.km = b->km;
    u.gamma = b->gamma; u.mach = b->mach;
    u.lx = b->lx; u.ly = b->ly; u.lz = b->lz;
    for (int f = 0; f < 6; ++f) u.twall[f] = b->twall[f];
    u.deltat = dt;
    u.xslip = b->xslip;
    const size_t n = b->n3;
    double *q = malloc(5 * n * sizeof(double)), *vel = malloc(3 * n * sizeof(double));
    for (int m = 0; m < 5; ++m) memcpy(q + m * n, b->q[m], n * sizeof(double));
    for (int c = 0; c < 3; ++c) memcpy(vel + c * n, b->vel[c], n * sizeof(double));
    for (int f = 0; f < 6; ++f) {
        if (b->bc[f] == 23) orc_gcnscbc(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 50) orc_zeroextrap(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 60) orc_symmetry(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 41 || b->bc[f] == 42)
            orc_wall(&u, f + 1, b->bc[f] == 41, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 21) orc_outflow(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 31) orc_fixbc(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 421) orc_slipadib(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
        else if (b->bc[f] == 411) orc_slipisot(&u, f + 1, q, b->rho, vel, b->prs, b->tmp);
    }
    for (int m = 0; m < 5; ++m) memcpy(b->q[m], q + m * n, n * sizeof(double));
    for (int c = 0; c < 3; ++c) memcpy(b->vel[c], vel + c * n, n * sizeof(double));
    free(q);
    free(vel);
}

void orc_box_free(orc_box *b)
{
    if (!b) return;
    for (int m = 0; m < 5; ++m) { free(b->q[m]); free(b->qrhs[m]); free(b->qsave[m]); }
    free(b->rho); free(b->prs); free(b->tmp);
    for (int m = 0; m < 3; ++m) { free(b->vel[m]); free(b->dtmp[m]); free(b->qflux[m]); }
    for (int m = 0; m < 9; ++m) { free(b->dvel[m]); free(b->dxi[m]); }
    for (int m = 0; m < 6; ++m) free(b->sigma[m]);
    free(b->jacob);
    for (int d = 0; d < 3; ++d) { free_scheme(&b->fd[d]); free_scheme(&b->fl[d]); }
    free(b);
}

/* ---------- pointwise thermodynamics (src/fludyna.F90) ---------- */

/* q2fvar_3da, src/fludyna.F90:545-634 (nondimensional, no species) */
static void q2fvar_pt(const orc_box *b, size_t p)
{
    double rho = b->q[0][p];
    b->rho[p] = rho;
    double u = b->q[1][p] / rho, v = b->q[2][p] / rho, w = b->q[3][p] / rho;
    b->vel[0][p] = u; b->vel[1][p] = v; b->vel[2][p] = w;
    double pr = (b->q[4][p] - 0.5 * rho * (u * u + v * v + w * w)) / b->const6;
    b->prs[p] = pr;
    b->tmp[p] = pr / rho * b->const2; /* thermal(pressure,density) */
}

/* miucal, src/fludyna.F90:791-820 (nondimensional) */
static double miucal(const orc_box *b, double t)
{
    return t * sqrt(t) * b->tempconst1 / (t + b->tempconst);
}

/* ---------- line gather / scatter helpers ---------- */

static void gather(const orc_box *b, const double *a, int dir, int i, int j,
                   int k, int dim, double *f)
{
    for (int n = -HM; n <= dim + HM; ++n) {
        size_t p = dir == 0 ? IX(b, n, j, k) : dir == 1 ? IX(b, i, n, k) : IX(b, i, j, n);
        f[n + HM] = a[p];
    }
}

static size_t lidx(const orc_box *b, int dir, int i, int j, int k, int n)
{
    return dir == 0 ? IX(b, n, j, k) : dir == 1 ? IX(b, i, n, k) : IX(b, i, j, n);
}

/* number of lines in a sweep along dir, and their (i,j,k) */
static void line_ijk(const orc_box *b, int dir, long l, int *i, int *j, int *k)
{
    if (dir == 0) { *i = 0; *j = (int)(l % (b->jm + 1)); *k = (int)(l / (b->jm + 1)); }
    else if (dir == 1) { *i = (int)(l % (b->im + 1)); *j = 0; *k = (int)(l / (b->im + 1)); }
    else { *i = (int)(l % (b->im + 1)); *j = (int)(l / (b->im + 1)); *k = 0; }
}

static long nlines(const orc_box *b, int dir)
{
    if (dir == 0) return (long)(b->jm + 1) * (b->km + 1);
    if (dir == 1) return (long)(b->im + 1) * (b->km + 1);
    return (long)(b->im + 1) * (b->jm + 1);
}

static int dimof(const orc_box *b, int dir)
{
    return dir == 0 ? b->im : dir == 1 ? b->jm : b->km;
}

/* ---------- halo copies (src/parallel.F90 single-rank branches) ---------- */

/* array4d_sendrecv periodic branch, src/parallel.F90:4170-4175 (i),
 * :4230-4240 (j), :4295-4305 (k): plain copy of hm planes, over the
 * 0:jm,0:km (resp. 0:im,0:km / 0:im,0:jm) face only. */
static void halo_copy(const orc_box *b, double *a, int dir)
{
    int im = b->im, jm = b->jm, km = b->km;
    if (dir == 0) {
        for (int k = 0; k <= km; ++k)
            for (int j = 0; j <= jm; ++j)
                for (int m = 1; m <= HM; ++m) {
                    a[IX(b, -m, j, k)] = a[IX(b, im - m, j, k)];
                    a[IX(b, im + m, j, k)] = a[IX(b, m, j, k)];
                }
    } else if (dir == 1) {
        for (int k = 0; k <= km; ++k)
            for (int i = 0; i <= im; ++i)
                for (int m = 1; m <= HM; ++m) {
                    a[IX(b, i, -m, k)] = a[IX(b, i, jm - m, k)];
                    a[IX(b, i, jm + m, k)] = a[IX(b, i, m, k)];
                }
    } else {
        for (int j = 0; j <= jm; ++j)
            for (int i = 0; i <= im; ++i)
                for (int m = 1; m <= HM; ++m) {
                    a[IX(b, i, j, -m)] = a[IX(b, i, j, km - m)];
                    a[IX(b, i, j, km + m)] = a[IX(b, i, j, m)];
                }
    }
}

/* ---------- main variant ---------- */

/* filterq, src/comsolver.F90:514-632: dataswap then filter, per direction,
 * in place. */
static void filterq_dir(orc_box *b, int dir)
{
    if (b->nt[dir] == 3)
        for (int m = 0; m < 5; ++m) halo_copy(b, b->q[m], dir);
    int dim = dimof(b, dir);
    long nl = nlines(b, dir);
    const scheme_t *s = &b->fl[dir];
#pragma omp parallel
    {
        double *f = malloc(sizeof(double) * (dim + 1 + 2 * HM));
        double *ff = malloc(sizeof(double) * (dim + 1));
        double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
        for (long l = 0; l < nl; ++l) {
            int i, j, k;
            line_ijk(b, dir, l, &i, &j, &k);
            for (int m = 0; m < 5; ++m) {
                gather(b, b->q[m], dir, i, j, k, dim, f);
                line_filter(s, b->coef, f, ff, work);
                for (int n = 0; n <= dim; ++n) b->q[m][lidx(b, dir, i, j, k, n)] = ff[n];
            }
        }
        free(f); free(ff); free(work);
    }
}

void orc_box_filterq(orc_box *b)
{
    if (b->variant == 0) {
        for (int d = 0; d < 3; ++d) filterq_dir(b, d);
    } else {
        /* tgvsolver.F90 filterq uses bchomovec, the same plain copy. */
        for (int d = 0; d < 3; ++d) filterq_dir(b, d);
    }
}

/* qswap single-rank periodic branch, src/parallel.F90:4849-5319:
 * per direction: halo copy, duplicate-node average, q2fvar on the
 * strips -hm:0 and dim:dim+hm over the face 0:jm,0:km (resp. ...). */
void orc_box_qswap(orc_box *b)
{
    int im = b->im, jm = b->jm, km = b->km;
    /* i direction */
    if (b->nt[0] == 3) {
    for (int m = 0; m < 5; ++m) {
        double *a = b->q[m];
        halo_copy(b, a, 0);
        for (int k = 0; k <= km; ++k)
            for (int j = 0; j <= jm; ++j) {
                a[IX(b, 0, j, k)] = 0.5 * (a[IX(b, 0, j, k)] + a[IX(b, im, j, k)]);
                a[IX(b, im, j, k)] = a[IX(b, 0, j, k)];
            }
    }
    for (int k = 0; k <= km; ++k)
        for (int j = 0; j <= jm; ++j) {
            for (int i = im; i <= im + HM; ++i) q2fvar_pt(b, IX(b, i, j, k));
            for (int i = -HM; i <= 0; ++i) q2fvar_pt(b, IX(b, i, j, k));
        }
    }
    /* j direction */
    if (b->nt[1] == 3) {
    for (int m = 0; m < 5; ++m) {
        double *a = b->q[m];
        halo_copy(b, a, 1);
        for (int k = 0; k <= km; ++k)
            for (int i = 0; i <= im; ++i) {
                a[IX(b, i, 0, k)] = 0.5 * (a[IX(b, i, 0, k)] + a[IX(b, i, jm, k)]);
                a[IX(b, i, jm, k)] = a[IX(b, i, 0, k)];
            }
    }
    for (int k = 0; k <= km; ++k)
        for (int i = 0; i <= im; ++i) {
            for (int j = -HM; j <= 0; ++j) q2fvar_pt(b, IX(b, i, j, k));
            for (int j = jm; j <= jm + HM; ++j) q2fvar_pt(b, IX(b, i, j, k));
        }
    }
    /* k direction */
    if (b->nt[2] == 3) {
    for (int m = 0; m < 5; ++m) {
        double *a = b->q[m];
        halo_copy(b, a, 2);
        for (int j = 0; j <= jm; ++j)
            for (int i = 0; i <= im; ++i) {
                a[IX(b, i, j, 0)] = 0.5 * (a[IX(b, i, j, 0)] + a[IX(b, i, j, km)]);
                a[IX(b, i, j, km)] = a[IX(b, i, j, 0)];
            }
    }
    for (int j = 0; j <= jm; ++j)
        for (int i = 0; i <= im; ++i) {
            for (int k = -HM; k <= 0; ++k) q2fvar_pt(b, IX(b, i, j, k));
            for (int k = km; k <= km + HM; ++k) q2fvar_pt(b, IX(b, i, j, k));
        }
    }
}

/* gradcal, src/comsolver.F90:244-497 (no species, no k-omega) */
static void gradcal_main(orc_box *b)
{
    for (int m = 0; m < 9; ++m) memset(b->dvel[m], 0, sizeof(double) * b->n3);
    for (int m = 0; m < 3; ++m) memset(b->dtmp[m], 0, sizeof(double) * b->n3);
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
#pragma omp parallel
        {
            double *f = malloc(sizeof(double) * (dim + 1 + 2 * HM));
            double *df = malloc(sizeof(double) * 4 * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                const double *src[4] = {b->vel[0], b->vel[1], b->vel[2], b->tmp};
                for (int c = 0; c < 4; ++c) {
                    gather(b, src[c], dir, i, j, k, dim, f);
                    line_df(s, f, df + c * (dim + 1), work);
                }
                for (int n = 0; n <= dim; ++n) {
                    size_t p = lidx(b, dir, i, j, k, n);
                    for (int m = 0; m < 3; ++m)
                        for (int nn = 0; nn < 3; ++nn)
                            b->dvel[m * 3 + nn][p] = b->dvel[m * 3 + nn][p] +
                                df[m * (dim + 1) + n] * b->dxi[dir * 3 + nn][p];
                    for (int nn = 0; nn < 3; ++nn)
                        b->dtmp[nn][p] = b->dtmp[nn][p] +
                            df[3 * (dim + 1) + n] * b->dxi[dir * 3 + nn][p];
                }
            }
            free(f); free(df); free(work);
        }
    }
}

/* convrsdcal6, src/solver.F90:2174-2342 (numq=5) */
static void convrsdcal6(orc_box *b)
{
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
        const double *x1 = b->dxi[dir * 3 + 0], *x2 = b->dxi[dir * 3 + 1],
                     *x3 = b->dxi[dir * 3 + 2];
#pragma omp parallel
        {
            int nn = dim + 1 + 2 * HM;
            double *fcs = malloc(sizeof(double) * 5 * nn);
            double *df = malloc(sizeof(double) * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                for (int n = -HM; n <= dim + HM; ++n) {
                    size_t p = lidx(b, dir, i, j, k, n);
                    double uu = x1[p] * b->vel[0][p] + x2[p] * b->vel[1][p] +
                                x3[p] * b->vel[2][p];
                    double jc = b->jacob[p], pr = b->prs[p];
                    fcs[0 * nn + n + HM] = jc * b->q[0][p] * uu;
                    fcs[1 * nn + n + HM] = jc * (b->q[1][p] * uu + x1[p] * pr);
                    fcs[2 * nn + n + HM] = jc * (b->q[2][p] * uu + x2[p] * pr);
                    fcs[3 * nn + n + HM] = jc * (b->q[3][p] * uu + x3[p] * pr);
                    fcs[4 * nn + n + HM] = jc * (b->q[4][p] + pr) * uu;
                }
                for (int m = 0; m < 5; ++m) {
                    line_df(s, fcs + m * nn, df, work);
                    for (int n = 0; n <= dim; ++n) {
                        size_t p = lidx(b, dir, i, j, k, n);
                        b->qrhs[m][p] = b->qrhs[m][p] + df[n];
                    }
                }
            }
            free(fcs); free(df); free(work);
        }
    }
}

/* diffrsdcal6, src/solver.F90:2355-2874 (turbmode none, no species) */
static void diffrsdcal6(orc_box *b)
{
    int im = b->im, jm = b->jm, km = b->km;
    for (int m = 0; m < 6; ++m) memset(b->sigma[m], 0, sizeof(double) * b->n3);
    for (int m = 0; m < 3; ++m) memset(b->qflux[m], 0, sizeof(double) * b->n3);
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= km; ++k)
        for (int j = 0; j <= jm; ++j)
            for (int i = 0; i <= im; ++i) {
                size_t p = IX(b, i, j, k);
                double miu = miucal(b, b->tmp[p]) / b->reynolds;
                double s11 = b->dvel[0][p];
                double s12 = 0.5 * (b->dvel[1][p] + b->dvel[3][p]);
                double s13 = 0.5 * (b->dvel[2][p] + b->dvel[6][p]);
                double s22 = b->dvel[4][p];
                double s23 = 0.5 * (b->dvel[5][p] + b->dvel[7][p]);
                double s33 = b->dvel[8][p];
                double skk = num1d3 * (s11 + s22 + s33);
                double miu2 = 2.0 * miu;
                double hcc = (miu / b->prandtl) / b->const5;
                double detk = 0.0;
                double s1 = miu2 * (s11 - skk) - detk + 0.0;
                double s2 = miu2 * s12 + 0.0;
                double s3 = miu2 * s13 + 0.0;
                double s4 = miu2 * (s22 - skk) - detk + 0.0;
                double s5 = miu2 * s23 + 0.0;
                double s6 = miu2 * (s33 - skk) - detk + 0.0;
                b->sigma[0][p] = s1; b->sigma[1][p] = s2; b->sigma[2][p] = s3;
                b->sigma[3][p] = s4; b->sigma[4][p] = s5; b->sigma[5][p] = s6;
                double u = b->vel[0][p], v = b->vel[1][p], w = b->vel[2][p];
                b->qflux[0][p] = hcc * b->dtmp[0][p] + s1 * u + s2 * v + s3 * w;
                b->qflux[1][p] = hcc * b->dtmp[1][p] + s2 * u + s4 * v + s5 * w;
                b->qflux[2][p] = hcc * b->dtmp[2][p] + s3 * u + s5 * v + s6 * w;
            }
    for (int d = 0; d < 3; ++d) {
        if (b->nt[d] != 3) continue;
        for (int m = 0; m < 6; ++m) halo_copy(b, b->sigma[m], d);
    }
    for (int d = 0; d < 3; ++d) {
        if (b->nt[d] != 3) continue;
        for (int m = 0; m < 3; ++m) halo_copy(b, b->qflux[m], d);
    }
    /* flux components per direction: rows of sigma used with dxi(d,1..3) */
    static const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
        const double *x1 = b->dxi[dir * 3 + 0], *x2 = b->dxi[dir * 3 + 1],
                     *x3 = b->dxi[dir * 3 + 2];
#pragma omp parallel
        {
            int nn = dim + 1 + 2 * HM;
            double *ff = malloc(sizeof(double) * 4 * nn);
            double *df = malloc(sizeof(double) * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                for (int n = -HM; n <= dim + HM; ++n) {
                    size_t p = lidx(b, dir, i, j, k, n);
                    double jc = b->jacob[p];
                    for (int c = 0; c < 3; ++c)
                        ff[c * nn + n + HM] =
                            (b->sigma[sidx[c][0]][p] * x1[p] +
                             b->sigma[sidx[c][1]][p] * x2[p] +
                             b->sigma[sidx[c][2]][p] * x3[p]) * jc;
                    ff[3 * nn + n + HM] = (b->qflux[0][p] * x1[p] +
                                           b->qflux[1][p] * x2[p] +
                                           b->qflux[2][p] * x3[p]) * jc;
                }
                for (int c = 0; c < 4; ++c) {
                    line_df(s, ff + c * nn, df, work);
                    for (int n = 0; n <= dim; ++n) {
                        size_t p = lidx(b, dir, i, j, k, n);
                        b->qrhs[c + 1][p] = b->qrhs[c + 1][p] + df[n];
                    }
                }
            }
            free(ff); free(df); free(work);
        }
    }
}

static void zero_qrhs(orc_box *b)
{
    for (int m = 0; m < 5; ++m) memset(b->qrhs[m], 0, sizeof(double) * b->n3);
}

static void negate_qrhs(orc_box *b)
{
    for (int m = 0; m < 5; ++m)
        for (int k = 0; k <= b->km; ++k)
            for (int j = 0; j <= b->jm; ++j)
                for (int i = 0; i <= b->im; ++i) {
                    size_t p = IX(b, i, j, k);
                    b->qrhs[m][p] = -b->qrhs[m][p];
                }
}

/* rhscal, src/solver.F90:184-283: gradcal, convrsdcal6, qrhs=-qrhs,
 * diffrsdcal6.  qrhs is cleared by the caller (src/mainloop.F90:404). */
void orc_box_rhscal(orc_box *b)
{
    gradcal_main(b);
    convrsdcal6(b);
    negate_qrhs(b);
    diffrsdcal6(b);
}

/* updatefvar, src/fludyna.F90:191-242 */
void orc_box_updatefvar(orc_box *b)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= b->km; ++k)
        for (int j = 0; j <= b->jm; ++j)
            for (int i = 0; i <= b->im; ++i) q2fvar_pt(b, IX(b, i, j, k));
}

/* kenergycal / enstophycal / diss_rate_cal, src/statistic.F90:870-1042
 * (3-D branch, sums over 1:im,1:jm,1:km in loop order k,j,i). */
void orc_box_stats(orc_box *b, double *st)
{
    int im = b->im, jm = b->jm, km = b->km;
    double ke = 0.0, en = 0.0, di = 0.0;
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j)
            for (int i = 1; i <= im; ++i) {
                size_t p = IX(b, i, j, k);
                double u = b->vel[0][p], v = b->vel[1][p], w = b->vel[2][p];
                double var1 = u * u + v * v + w * w;
                ke = ke + b->rho[p] * var1;
                double o1 = b->dvel[7][p] - b->dvel[5][p];
                double o2 = b->dvel[2][p] - b->dvel[6][p];
                double o3 = b->dvel[3][p] - b->dvel[1][p];
                double om = o1 * o1 + o2 * o2 + o3 * o3;
                en = en + b->rho[p] * om;
                double miu = miucal(b, b->tmp[p]) / b->reynolds;
                double s11 = b->dvel[0][p], s22 = b->dvel[4][p], s33 = b->dvel[8][p];
                double s12 = 0.5 * (b->dvel[1][p] + b->dvel[3][p]);
                double s13 = 0.5 * (b->dvel[2][p] + b->dvel[6][p]);
                double s23 = 0.5 * (b->dvel[5][p] + b->dvel[7][p]);
                double div = s11 + s22 + s33;
                double va = 2.0 * miu * (s11 * s11 + s22 * s22 + s33 * s33 +
                                         2.0 * (s12 * s12 + s13 * s13 + s23 * s23) -
                                         num1d3 * (div * div));
                di = di + va;
            }
    double n3 = (double)im * (double)jm * (double)km;
    if (b->variant == 0) {
        ke = 0.5 * ke / n3;
        ke = ke / (b->roinf * b->uinf * b->uinf);
        en = 0.5 * en / n3;
        double l0 = b->lx / (2.0 * (4.0 * atan(1.0)));
        en = en / (b->roinf * ((b->uinf / l0) * (b->uinf / l0)));
        di = di / n3;
    } else {
        /* stacal, tgvsolver.F90:893-942 */
        ke = 0.5 * ke / n3;
        en = 0.5 * en / n3;
        di = di / n3;
    }
    st[0] = ke; st[1] = en; st[2] = di;
}

/* ---------- mini-app variant (miniapps/tgv_solver_3d/tgvsolver.F90) ---------- */

/* fvar2q with pressure, tgvsolver.F90:45-66 */
static void fvar2q_p_pt(orc_box *b, size_t p)
{
    double rho = b->rho[p], u = b->vel[0][p], v = b->vel[1][p], w = b->vel[2][p];
    b->q[0][p] = rho;
    b->q[1][p] = rho * u;
    b->q[2][p] = rho * v;
    b->q[3][p] = rho * w;
    double var1 = 0.5 * (u * u + v * v + w * w);
    b->q[4][p] = b->prs[p] * b->const6 + rho * var1;
}

/* bchomo, tgvsolver.F90:271-369: copy primitives to the halos and rebuild q
 * there from them. */
static void bchomo(orc_box *b)
{
    int im = b->im, jm = b->jm, km = b->km;
    double *pr[6] = {b->rho, b->vel[0], b->vel[1], b->vel[2], b->tmp, b->prs};
    for (int k = 0; k <= km; ++k)
        for (int j = 0; j <= jm; ++j)
            for (int m = 1; m <= HM; ++m) {
                size_t d1 = IX(b, -m, j, k), s1 = IX(b, im - m, j, k);
                size_t d2 = IX(b, im + m, j, k), s2 = IX(b, m, j, k);
                for (int c = 0; c < 6; ++c) { pr[c][d1] = pr[c][s1]; pr[c][d2] = pr[c][s2]; }
                fvar2q_p_pt(b, d1);
                fvar2q_p_pt(b, d2);
            }
    for (int k = 0; k <= km; ++k)
        for (int i = 0; i <= im; ++i)
            for (int m = 1; m <= HM; ++m) {
                size_t d1 = IX(b, i, -m, k), s1 = IX(b, i, jm - m, k);
                size_t d2 = IX(b, i, jm + m, k), s2 = IX(b, i, m, k);
                for (int c = 0; c < 6; ++c) { pr[c][d1] = pr[c][s1]; pr[c][d2] = pr[c][s2]; }
                fvar2q_p_pt(b, d1);
                fvar2q_p_pt(b, d2);
            }
    for (int j = 0; j <= jm; ++j)
        for (int i = 0; i <= im; ++i)
            for (int m = 1; m <= HM; ++m) {
                size_t d1 = IX(b, i, j, -m), s1 = IX(b, i, j, km - m);
                size_t d2 = IX(b, i, j, km + m), s2 = IX(b, i, j, m);
                for (int c = 0; c < 6; ++c) { pr[c][d1] = pr[c][s1]; pr[c][d2] = pr[c][s2]; }
                fvar2q_p_pt(b, d1);
                fvar2q_p_pt(b, d2);
            }
}

/* gradcal, convection, diffusion of tgvsolver.F90:374-843 (d/dx by
 * division, no metrics). */
static void rhscal_mini(orc_box *b)
{
    double h[3] = {b->dx, b->dy, b->dz};
    /* gradcal: dvel(.,m,d) = df/h */
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
#pragma omp parallel
        {
            double *f = malloc(sizeof(double) * (dim + 1 + 2 * HM));
            double *df = malloc(sizeof(double) * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                const double *src[4] = {b->vel[0], b->vel[1], b->vel[2], b->tmp};
                for (int c = 0; c < 4; ++c) {
                    gather(b, src[c], dir, i, j, k, dim, f);
                    line_df(s, f, df, work);
                    for (int n = 0; n <= dim; ++n) {
                        size_t p = lidx(b, dir, i, j, k, n);
                        if (c < 3) b->dvel[c * 3 + dir][p] = df[n] / h[dir];
                        else b->dtmp[dir][p] = df[n] / h[dir];
                    }
                }
            }
            free(f); free(df); free(work);
        }
    }
    /* convection */
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
#pragma omp parallel
        {
            int nn = dim + 1 + 2 * HM;
            double *fcs = malloc(sizeof(double) * 5 * nn);
            double *df = malloc(sizeof(double) * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                for (int n = -HM; n <= dim + HM; ++n) {
                    size_t p = lidx(b, dir, i, j, k, n);
                    double uu = b->vel[dir][p], pr = b->prs[p];
                    /* the mini-app adds the pressure only to the momentum
                     * component of the sweep direction (tgvsolver.F90:578-631) */
                    for (int m = 0; m < 4; ++m)
                        fcs[m * nn + n + HM] = m == dir + 1 ? b->q[m][p] * uu + pr
                                                            : b->q[m][p] * uu;
                    fcs[4 * nn + n + HM] = (b->q[4][p] + pr) * uu;
                }
                for (int m = 0; m < 5; ++m) {
                    line_df(s, fcs + m * nn, df, work);
                    for (int n = 0; n <= dim; ++n) {
                        size_t p = lidx(b, dir, i, j, k, n);
                        b->qrhs[m][p] = b->qrhs[m][p] + df[n] / h[dir];
                    }
                }
            }
            free(fcs); free(df); free(work);
        }
    }
    negate_qrhs(b);
    /* diffusion */
    int im = b->im, jm = b->jm, km = b->km;
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= km; ++k)
        for (int j = 0; j <= jm; ++j)
            for (int i = 0; i <= im; ++i) {
                size_t p = IX(b, i, j, k);
                double s11 = b->dvel[0][p];
                double s12 = 0.5 * (b->dvel[1][p] + b->dvel[3][p]);
                double s13 = 0.5 * (b->dvel[2][p] + b->dvel[6][p]);
                double s22 = b->dvel[4][p];
                double s23 = 0.5 * (b->dvel[5][p] + b->dvel[7][p]);
                double s33 = b->dvel[8][p];
                double skk = num1d3 * (s11 + s22 + s33);
                double miu = miucal(b, b->tmp[p]) / b->reynolds;
                double miu2 = 2.0 * miu;
                double hcc = (miu / b->prandtl) / b->const5;
                double s1 = miu2 * (s11 - skk), s2 = miu2 * s12, s3 = miu2 * s13;
                double s4 = miu2 * (s22 - skk), s5 = miu2 * s23, s6 = miu2 * (s33 - skk);
                b->sigma[0][p] = s1; b->sigma[1][p] = s2; b->sigma[2][p] = s3;
                b->sigma[3][p] = s4; b->sigma[4][p] = s5; b->sigma[5][p] = s6;
                double u = b->vel[0][p], v = b->vel[1][p], w = b->vel[2][p];
                b->qflux[0][p] = hcc * b->dtmp[0][p] + s1 * u + s2 * v + s3 * w;
                b->qflux[1][p] = hcc * b->dtmp[1][p] + s2 * u + s4 * v + s5 * w;
                b->qflux[2][p] = hcc * b->dtmp[2][p] + s3 * u + s5 * v + s6 * w;
            }
    for (int d = 0; d < 3; ++d) {
        for (int m = 0; m < 6; ++m) halo_copy(b, b->sigma[m], d);
        for (int m = 0; m < 3; ++m) halo_copy(b, b->qflux[m], d);
    }
    static const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int dir = 0; dir < 3; ++dir) {
        int dim = dimof(b, dir);
        long nl = nlines(b, dir);
        const scheme_t *s = &b->fd[dir];
#pragma omp parallel
        {
            double *f = malloc(sizeof(double) * (dim + 1 + 2 * HM));
            double *df = malloc(sizeof(double) * (dim + 1));
            double *work = malloc(sizeof(double) * 2 * s->n);
#pragma omp for schedule(static)
            for (long l = 0; l < nl; ++l) {
                int i, j, k;
                line_ijk(b, dir, l, &i, &j, &k);
                for (int c = 0; c < 4; ++c) {
                    const double *src = c < 3 ? b->sigma[sidx[dir][c]] : b->qflux[dir];
                    gather(b, src, dir, i, j, k, dim, f);
                    line_df(s, f, df, work);
                    for (int n = 0; n <= dim; ++n) {
                        size_t p = lidx(b, dir, i, j, k, n);
                        b->qrhs[c + 1][p] = b->qrhs[c + 1][p] + df[n] / h[dir];
                    }
                }
            }
            free(f); free(df); free(work);
        }
    }
}

/* ---------- initial field ---------- */

void orc_box_init_tgv(orc_box *b)
{
    int im = b->im, jm = b->jm, km = b->km;
    double pi = 4.0 * atan(1.0);
    for (int k = 0; k <= km; ++k)
        for (int j = 0; j <= jm; ++j)
            for (int i = 0; i <= im; ++i) {
                size_t p = IX(b, i, j, k);
                if (b->variant == 0) {
                    /* gridcube, src/gridgeneration.F90:233-267; tgvini,
                     * src/initialisation.F90:1079-1149 (l_0 = 1) */
                    double x = b->lx / (double)im * (double)i;
                    double y = b->ly / (double)jm * (double)j;
                    double z = b->lz / (double)km * (double)k;
                    double rho = b->roinf;
                    b->rho[p] = rho;
                    b->vel[0][p] = b->uinf * sin(x) * cos(y) * cos(z);
                    b->vel[1][p] = -(b->uinf * cos(x) * sin(y) * cos(z));
                    b->vel[2][p] = 0.0;
                    b->prs[p] = b->pinf + rho / 16.0 * (b->uinf * b->uinf) *
                                              (cos(2.0 * x) + cos(2.0 * y)) *
                                              (cos(2.0 * z) + 2.0);
                    b->tmp[p] = b->prs[p] / rho * b->const2;
                    /* updateq -> fvar2q_3da with temperature,
                     * src/fludyna.F90:449-533 */
                    double u = b->vel[0][p], v = b->vel[1][p], w = b->vel[2][p];
                    b->q[0][p] = rho;
                    b->q[1][p] = rho * u;
                    b->q[2][p] = rho * v;
                    b->q[3][p] = rho * w;
                    b->q[4][p] = rho * (b->tmp[p] * b->const1 +
                                        0.5 * (u * u + v * v + w * w));
                } else {
                    /* flowfield_init, tgvsolver.F90:176-234 */
                    double x1 = 2.0 * pi / (double)im * (double)i;
                    double x2 = 2.0 * pi / (double)jm * (double)j;
                    double x3 = 2.0 * pi / (double)km * (double)k;
                    b->rho[p] = 1.0;
                    b->vel[0][p] = sin(x1) * cos(x2) * cos(x3);
                    b->vel[1][p] = -cos(x1) * sin(x2) * cos(x3);
                    b->vel[2][p] = 0.0;
                    b->prs[p] = b->pinf + 1.0 / 16.0 * (cos(2.0 * x1) + cos(2.0 * x2)) *
                                              (cos(2.0 * x3) + 2.0);
                    b->tmp[p] = b->prs[p] / b->rho[p] * b->const2;
                    fvar2q_p_pt(b, p);
                }
            }
}

/* ---------- state access ---------- */

static double *which_field(const orc_box *b, int which)
{
    if (which >= 0 && which < 5) return b->q[which];
    switch (which) {
    case 5: return b->rho;
    case 6: case 7: case 8: return b->vel[which - 6];
    case 9: return b->prs;
    case 10: return b->tmp;
    }
    return NULL;
}

static void copy_out(const orc_box *b, const double *a, double *out)
{
    size_t c = 0;
    for (int k = 0; k <= b->km; ++k)
        for (int j = 0; j <= b->jm; ++j)
            for (int i = 0; i <= b->im; ++i) out[c++] = a[IX(b, i, j, k)];
}

void orc_box_get(const orc_box *b, int which, double *out)
{
    copy_out(b, which_field(b, which), out);
}

void orc_box_set(orc_box *b, int which, const double *in)
{
    double *a = which_field(b, which);
    size_t c = 0;
    for (int k = 0; k <= b->km; ++k)
        for (int j = 0; j <= b->jm; ++j)
            for (int i = 0; i <= b->im; ++i) a[IX(b, i, j, k)] = in[c++];
}

/* The metrics gridgeom would compute for the box's grid (src/geom.F90:
 * 99-917), haloed, Fortran order: jacob(-hm:im+hm,...) and dxi(...,1:3,1:3)
 * with dxi(:,:,:,d,e) = d(xi_d)/d(x_e) at offset (d-1 + 3(e-1)) n3.  The
 * box's kernels read every component (convrsdcal6, diffrsdcal6, gradcal). */
void orc_box_set_metrics(orc_box *b, const double *jacob, const double *dxi)
{
    memcpy(b->jacob, jacob, sizeof(double) * b->n3);
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e)
            memcpy(b->dxi[d * 3 + e], dxi + (size_t)(d + 3 * e) * b->n3, sizeof(double) * b->n3);
}

void orc_box_get_haloed(const orc_box *b, int which, double *out)
{
    memcpy(out, which_field(b, which), sizeof(double) * b->n3);
}

void orc_box_get_deriv(const orc_box *b, int which, double *out)
{
    const double *a = which < 9 ? b->dvel[which] : which < 12 ? b->dtmp[which - 9]
                                                              : b->qrhs[which - 12];
    copy_out(b, a, out);
}

/* ---------- one RK3 step ---------- */

static const double rkc[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25},
                                 {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

void orc_box_step(orc_box *b, double dt, double *stats)
{
    int im = b->im, jm = b->jm, km = b->km;
    if (b->variant == 0) {
        /* time_integration_rk, src/mainloop.F90:298-595 (rk3, lfilter) */
        for (int s = 0; s < 3; ++s) {
            orc_box_filterq(b);
            zero_qrhs(b);
            box_boucon(b, dt);  /* boucon before qswap, src/mainloop.F90:413-415 */
            orc_box_qswap(b);
            orc_box_rhscal(b);
            if (s == 0) {
                for (int m = 0; m < 5; ++m)
                    for (int k = 0; k <= km; ++k)
                        for (int j = 0; j <= jm; ++j)
                            for (int i = 0; i <= im; ++i) {
                                size_t p = IX(b, i, j, k);
                                b->qsave[m][p] = b->q[m][p] * b->jacob[p];
                            }
                if (stats) orc_box_stats(b, stats);
            }
#pragma omp parallel for schedule(static)
            for (int k = 0; k <= km; ++k)
                for (int m = 0; m < 5; ++m)
                    for (int j = 0; j <= jm; ++j)
                        for (int i = 0; i <= im; ++i) {
                            size_t p = IX(b, i, j, k);
                            double v = rkc[s][0] * b->qsave[m][p] +
                                       rkc[s][1] * b->q[m][p] * b->jacob[p] +
                                       rkc[s][2] * b->qrhs[m][p] * dt;
                            b->q[m][p] = v / b->jacob[p];
                        }
            orc_box_updatefvar(b);
        }
    } else {
        /* rk3, tgvsolver.F90:994-1082.  Variant 2 keeps this stage order
         * (bchomo halos, right-hand side, qsave, RK3, filter after the
         * update) but computes the right-hand side and the RK3 combine as
         * the main path does: gradcal's metric products, convrsdcal6's and
         * diffrsdcal6's J-weighted fluxes in metric form, qsave = q J and
         * the combine divided by J (src/comsolver.F90:244-497,
         * src/solver.F90:2174-2874, src/mainloop.F90:423-445).  On a
         * uniform box this is the mini-app's arithmetic up to rounding. */
        const int mainform = b->variant == 2;
        for (int s = 0; s < 3; ++s) {
            bchomo(b);
            zero_qrhs(b);
            if (mainform) orc_box_rhscal(b);
            else rhscal_mini(b);
            if (s == 0) {
                if (stats) orc_box_stats(b, stats);
                for (int m = 0; m < 5; ++m)
                    for (int k = 0; k <= km; ++k)
                        for (int j = 0; j <= jm; ++j)
                            for (int i = 0; i <= im; ++i) {
                                size_t p = IX(b, i, j, k);
                                b->qsave[m][p] = mainform ? b->q[m][p] * b->jacob[p] : b->q[m][p];
                            }
            }
#pragma omp parallel for schedule(static)
            for (int k = 0; k <= km; ++k)
                for (int m = 0; m < 5; ++m)
                    for (int j = 0; j <= jm; ++j)
                        for (int i = 0; i <= im; ++i) {
                            size_t p = IX(b, i, j, k);
                            if (mainform) {
                                double v = rkc[s][0] * b->qsave[m][p] +
                                           rkc[s][1] * b->q[m][p] * b->jacob[p] +
                                           rkc[s][2] * b->qrhs[m][p] * dt;
                                b->q[m][p] = v / b->jacob[p];
                            } else {
                                b->q[m][p] = rkc[s][0] * b->qsave[m][p] +
                                             rkc[s][1] * b->q[m][p] +
                                             rkc[s][2] * b->qrhs[m][p] * dt;
                            }
                        }
            orc_box_filterq(b);
            orc_box_updatefvar(b);
        }
    }
}
