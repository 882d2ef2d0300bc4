/* Checker for the explicit upwind convection path of ASTR: a plain C
 * restatement of the reference routines, cited file:line, used only by the
 * tests and the CPU baseline.  Arithmetic follows the Fortran expression
 * order (left to right, parentheses as written), compiled without FMA
 * contraction, so the HIP kernels (built with contraction off for this
 * path) can be compared bit for bit.
 *
 * Layouts (host, the reference's own):
 *   haloed field f(-hm:im+hm,-hm:jm+hm,-hm:km+hm), i fastest;
 *   q(...,1:5), vel(...,1:3): 5 / 3 haloed fields back to back;
 *   dxi(...,1:3,1:3): 9 haloed fields, field (d-1)+3*(n-1) = dxi(.,d,n);
 *   qrhs(0:im,0:jm,0:km,1:5), dvel(0:im,..,1:3,1:3), dtmp(..,1:3);
 *   lshock(0:im,0:jm,0:km) as bytes.
 * km = 0 is a two-dimensional box (ndims = 2): no k sweeps, k halos of
 * hm planes on each side as the reference allocates them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define HMU 5

/* a bounded low / high end: ntype 1 / 2, and both ends of ntype 4 */
#define LO_END(nt) ((nt) == 1 || (nt) == 4)
#define HI_END(nt) ((nt) == 2 || (nt) == 4)

/* The closure type recons_exp sees at interface inode+1/2 of an axis of
 * ntype nt: ntype 4 acts as ntype 1 on the low half and ntype 2 on the
 * high half, which is what a run on several ranks along the axis computes
 * (the explicit upwind flux has no closure at rank interfaces: its stencils
 * read the neighbour's nodes through the halos). */
static int end_type(int nt, int inode, int dim)
{
    if (nt != 4) return nt;
    return inode < dim / 2 ? 1 : 2;
}

/* ---------------------------------------------------------------- scalars */

/* minmod2, minmod4: src/commfunc.F90:699-738 */
double orc_minmod2(double a, double b)
{
    if (a > 0.0 && b > 0.0) return fmin(fabs(a), fabs(b));
    if (a < 0.0 && b < 0.0) return -1.0 * fmin(fabs(a), fabs(b));
    return 0.0;
}

static double min2(double a, double b) { return a < b ? a : b; }
static double max2(double a, double b) { return a > b ? a : b; }

double orc_minmod4(double a, double b, double c, double d)
{
    if (a > 0.0 && b > 0.0 && c > 0.0 && d > 0.0)
        return min2(min2(min2(fabs(a), fabs(b)), fabs(c)), fabs(d));
    if (a < 0.0 && b < 0.0 && c < 0.0 && d < 0.0)
        return -1.0 * min2(min2(min2(fabs(a), fabs(b)), fabs(c)), fabs(d));
    return 0.0;
}

/* suw3/suw5/suw7, src/flux.F90:388-422 (constdef num1d6, num5d6, num1d3) */
double orc_suw3(const double *u)
{
    return -(1.0 / 6.0) * u[0] + (5.0 / 6.0) * u[1] + (1.0 / 3.0) * u[2];
}

double orc_suw5(const double *u)
{
    return (2.0 * u[0] - 13.0 * u[1] + 47.0 * u[2] + 27.0 * u[3] - 3.0 * u[4]) / 60.0;
}

double orc_suw7(const double *u)
{
    return (-3.0 * u[0] + 25.0 * u[1] - 101.0 * u[2] + 319.0 * u[3] + 214.0 * u[4] -
            38.0 * u[5] + 4.0 * u[6]) / 420.0;
}

/* The monotonicity-preserving limiter shared by MP5/MP7/MP5LD/MP7LD
 * (Suresh & Huynh 1997), src/flux.F90:456-490: u[0..4] are the points
 * u(1..5) of MP5 (u(i-2..i+2)), ul the linear value. */
static double mp_limit(const double *u, double ulinear)
{
    const double dm1 = u[0] - 2.0 * u[1] + u[2];
    const double d0 = u[1] - 2.0 * u[2] + u[3];
    const double d1 = u[2] - 2.0 * u[3] + u[4];
    const double dhm1 = orc_minmod4(4.0 * dm1 - d0, 4.0 * d0 - dm1, dm1, d0);
    const double dh0 = orc_minmod4(4.0 * d0 - d1, 4.0 * d1 - d0, d0, d1);
    const double uUL = u[2] + 4.0 * (u[2] - u[1]);
    const double uAV = 0.5 * (u[2] + u[3]);
    const double uMD = uAV - 0.5 * dh0;
    const double uLC = u[2] + 0.5 * (u[2] - u[1]) + 1.333333333333333 * dhm1;
    double v1 = min2(min2(u[2], u[3]), uMD);
    double v2 = min2(min2(u[2], uUL), uLC);
    const double uMIN = max2(v1, v2);
    v1 = max2(max2(u[2], u[3]), uMD);
    v2 = max2(max2(u[2], uUL), uLC);
    const double uMAX = min2(v1, v2);
    return ulinear + orc_minmod2(uMIN - ulinear, uMAX - ulinear);
}

/* MP5, src/flux.F90:434-496.  ul < 0 pointer: linear value from suw5;
 * discont < 0: argument absent (limiter allowed). */
double orc_mp5(const double *u, const double *ul, int discont)
{
    const double ulinear = ul ? *ul : orc_suw5(u);
    const int lskt = discont < 0 ? 1 : discont;
    const double uMP = u[2] + orc_minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
    const double v = (ulinear - u[2]) * (ulinear - uMP);
    if (lskt && v >= 1.e-10) return mp_limit(u, ulinear);
    return ulinear;
}

/* MP7, src/flux.F90:498-551: u(0:6), the limiter on u(1:5) */
double orc_mp7(const double *u)
{
    const double ulinear = (-3.0 * u[0] + 25.0 * u[1] - 101.0 * u[2] + 319.0 * u[3] +
                            214.0 * u[4] - 38.0 * u[5] + 4.0 * u[6]) / 420.0;
    const double uMP = u[3] + orc_minmod2(u[4] - u[3], 4.0 * (u[3] - u[2]));
    const double v = (ulinear - u[3]) * (ulinear - uMP);
    if (v >= 1.e-10) return mp_limit(u + 1, ulinear);
    return ulinear;
}

/* MP5LD, src/flux.F90:556-633: u(1:6) */
double orc_mp5ld(const double *u, double weightbw)
{
    const double vadp = 1.666666666666667e-2 * weightbw;
    const double b1 = -vadp + 3.333333333333333e-2;
    const double b2 = 5.0 * vadp - 2.166666666666667e-1;
    const double b3 = -10.0 * vadp + 7.833333333333333e-1;
    const double b4 = 10.0 * vadp + 0.45;
    const double b5 = -5.0 * vadp - 5.e-2;
    const double b6 = vadp;
    const double ulinear = b1 * u[0] + b2 * u[1] + b3 * u[2] + b4 * u[3] + b5 * u[4] + b6 * u[5];
    const double uMP = u[2] + orc_minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
    const double v = (ulinear - u[2]) * (ulinear - uMP);
    if (v >= 1.e-10) return mp_limit(u, ulinear);
    return ulinear;
}

/* MP7LD, src/flux.F90:635-706: u(0:7); the limiter is switched by the
 * shock flag alone */
double orc_mp7ld(const double *u, double weightbw, int lskt)
{
    const double vadp = 3.571428571428571e-3 * weightbw;
    double b[8];
    b[0] = 1.0 * vadp - 7.142857142857143e-3;
    b[1] = -7.0 * vadp + 5.952380952380952e-2;
    b[2] = 21.0 * vadp - 0.240476190476190;
    b[3] = -35.0 * vadp + 0.759523809523809;
    b[4] = 35.0 * vadp + 0.509523809523809;
    b[5] = -21.0 * vadp - 9.047619047619047e-2;
    b[6] = 7.0 * vadp + 9.523809523809525e-3;
    b[7] = -1.0 * vadp;
    const double ulinear = b[0] * u[0] + b[1] * u[1] + b[2] * u[2] + b[3] * u[3] +
                           b[4] * u[4] + b[5] * u[5] + b[6] * u[6] + b[7] * u[7];
    if (lskt) return mp_limit(u + 1, ulinear);
    return ulinear;
}

/* WENO5, src/flux.F90:761-821 */
static void weno5_parts(const double *u, double *uh, double *bt)
{
    uh[0] = 0.333333333333333 * u[0] - 1.16666666666667 * u[1] + 1.83333333333333 * u[2];
    uh[1] = -0.166666666666667 * u[1] + 0.833333333333333 * u[2] + 0.333333333333333 * u[3];
    uh[2] = 0.333333333333333 * u[2] + 0.833333333333333 * u[3] - 0.166666666666667 * u[4];
    double a = u[0] - 2.0 * u[1] + u[2], b = u[0] - 4.0 * u[1] + 3.0 * u[2];
    bt[0] = 1.08333333333333 * (a * a) + 0.25 * (b * b);
    a = u[1] - 2.0 * u[2] + u[3];
    b = u[1] - u[3];
    bt[1] = 1.08333333333333 * (a * a) + 0.25 * (b * b);
    a = u[2] - 2.0 * u[3] + u[4];
    b = 3.0 * u[2] - 4.0 * u[3] + u[4];
    bt[2] = 1.08333333333333 * (a * a) + 0.25 * (b * b);
}

double orc_weno5(const double *u)
{
    double uh[3], bt[3];
    weno5_parts(u, uh, bt);
    const double eps = 1.e-6;
    double e;
    e = bt[0] + eps; const double a1 = 0.1 / (e * e);
    e = bt[1] + eps; const double a2 = 0.6 / (e * e);
    e = bt[2] + eps; const double a3 = 0.3 / (e * e);
    const double as = a1 + a2 + a3;
    return a1 / as * uh[0] + a2 / as * uh[1] + a3 / as * uh[2];
}

/* WENO5Z, src/flux.F90:936-994 */
double orc_weno5z(const double *u)
{
    double uh[3], bt[3];
    weno5_parts(u, uh, bt);
    const double eps = 1.e-6;
    const double tau5 = fabs(bt[2] - bt[0]);
    double e;
    e = bt[0] + eps; const double a1 = 0.1 + 0.1 * (tau5 / (e * e));
    e = bt[1] + eps; const double a2 = 0.6 + 0.6 * (tau5 / (e * e));
    e = bt[2] + eps; const double a3 = 0.3 + 0.3 * (tau5 / (e * e));
    const double as = a1 + a2 + a3;
    return a1 / as * uh[0] + a2 / as * uh[1] + a3 / as * uh[2];
}

/* WENO7, src/flux.F90:826-925 */
static void weno7_parts(const double *u, double *uh, double *bt)
{
    uh[0] = -0.25 * u[0] + 1.083333333333333 * u[1] - 1.916666666666667 * u[2] +
            2.083333333333333 * u[3];
    uh[1] = 8.333333333333333e-2 * u[1] - 4.166666666666667e-1 * u[2] +
            1.083333333333333 * u[3] + 0.25 * u[4];
    uh[2] = -8.333333333333333e-2 * u[2] + 5.833333333333333e-1 * u[3] +
            5.833333333333333e-1 * u[4] - 8.333333333333333e-2 * u[5];
    uh[3] = 0.25 * u[3] + 1.083333333333333 * u[4] - 4.166666666666667e-1 * u[5] +
            8.333333333333333e-2 * u[6];
    double s1, s2, s3;
    s1 = -2.0 * u[0] + 9.0 * u[1] - 18.0 * u[2] + 11.0 * u[3];
    s2 = -1.0 * u[0] + 4.0 * u[1] - 5.0 * u[2] + 2.0 * u[3];
    s3 = -1.0 * u[0] + 3.0 * u[1] - 3.0 * u[2] + 1.0 * u[3];
    bt[0] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) +
            1.084722222222222 * (s3 * s3);
    s1 = u[1] - 6.0 * u[2] + 3.0 * u[3] + 2.0 * u[4];
    s2 = u[2] - 2.0 * u[3] + u[4];
    s3 = -1.0 * u[1] + 3.0 * u[2] - 3.0 * u[3] + 1.0 * u[4];
    bt[1] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) +
            1.084722222222222 * (s3 * s3);
    s1 = -2.0 * u[2] - 3.0 * u[3] + 6.0 * u[4] - 1.0 * u[5];
    s2 = u[2] - 2.0 * u[3] + u[4];
    s3 = -1.0 * u[2] + 3.0 * u[3] - 3.0 * u[4] + 1.0 * u[5];
    bt[2] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) +
            1.084722222222222 * (s3 * s3);
    s1 = -11.0 * u[3] + 18.0 * u[4] - 9.0 * u[5] + 2.0 * u[6];
    s2 = 2.0 * u[3] - 5.0 * u[4] + 4.0 * u[5] - u[6];
    s3 = -1.0 * u[3] + 3.0 * u[4] - 3.0 * u[5] + 1.0 * u[6];
    bt[3] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) +
            1.084722222222222 * (s3 * s3);
}

static const double C7[4] = {2.857142857142857e-2, 3.428571428571429e-1, 5.142857142857143e-1,
                             1.142857142857143e-1};

double orc_weno7(const double *u)
{
    double uh[4], bt[4], a[4];
    weno7_parts(u, uh, bt);
    for (int k = 0; k < 4; ++k) {
        const double e = bt[k] + 1.e-6;
        a[k] = C7[k] / (e * e);
    }
    const double as = a[0] + a[1] + a[2] + a[3];
    return a[0] / as * uh[0] + a[1] / as * uh[1] + a[2] / as * uh[2] + a[3] / as * uh[3];
}

/* WENO7Z, src/flux.F90:999-1050 */
double orc_weno7z(const double *u)
{
    double uh[4], bt[4], a[4];
    weno7_parts(u, uh, bt);
    const double tau7 = fabs(bt[3] - bt[0]);
    for (int k = 0; k < 4; ++k) {
        const double e = bt[k] + 1.e-6;
        a[k] = C7[k] + C7[k] * (tau7 / (e * e));
    }
    const double as = a[0] + a[1] + a[2] + a[3];
    return a[0] / as * uh[0] + a[1] / as * uh[1] + a[2] / as * uh[2] + a[3] / as * uh[3];
}

/* ROUND, src/flux.F90:717-752 */
double orc_round(const double *u)
{
    const double eps = 1.e-16;
    const double z0c = (u[1] - u[0] + eps) / (u[2] - u[0] + eps);
    const double a1c = 1.0 + 12.0 * z0c * z0c;
    const double a2c = 1.0 + 5.0 * (z0c - 1.0) * (z0c - 1.0);
    const double zl = z0c - 0.05, zr = 0.47 - z0c;
    const double plc = 1100.0 * zl * zl * zl * zr * zr * zr;
    const double zp = z0c - 0.55, zq = 0.97 - z0c;
    const double prc = 18000.0 * zp * zp * zp * zq * zq * zq * zq * zq;
    const double p1c = 0.833333333333333 * z0c + 0.333333333333333 + max2(plc, 0.0) + max2(prc, 0.0);
    const double p2c = 1.5 * z0c;
    const double p3c = 0.5 * z0c + 0.5;
    const double wc1c = 1.0 / a1c / a1c / a1c / a1c;
    const double wc2c = 1.0 / a2c / a2c / a2c / a2c / a2c / a2c / a2c / a2c;
    const double gc = (p1c * (1.0 - wc1c) + p2c * wc1c) * (1.0 - wc2c) + p3c * wc2c;
    return gc * (u[2] - u[0]) + u[0];
}

/* recons_exp, src/flux.F90:269-350: f(1:8) of the interface inode+1/2,
 * degraded near ntype 1/2 boundaries.  Returns NAN for an undefined
 * scheme (the reference stops). */
double orc_recons_exp(const double *f, int inode, int dim, int ntype, int reschem, int shock,
                      double bfacmpld)
{
    if ((ntype == 1 && inode == 0) || (ntype == 2 && inode == dim - 1))
        return reschem == -1 ? f[3] : 0.5 * (f[3] + f[4]);
    if ((ntype == 1 && inode == 1) || (ntype == 2 && inode == dim - 2))
        return reschem == -1 ? f[3] : orc_suw3(f + 2);
    if ((ntype == 1 && inode == 2) || (ntype == 2 && inode == dim - 3)) {
        switch (reschem) {
        case -1: return f[3];
        case 0: return orc_suw5(f + 1);
        case 1: return orc_weno5(f + 1);
        case 2: return orc_weno5z(f + 1);
        case 3: return orc_mp5(f + 1, NULL, -1);
        case 5: return orc_mp5ld(f + 1, bfacmpld);
        case 6: return orc_round(f + 2);
        }
        return NAN;
    }
    switch (reschem) {
    case -1: return f[3];
    case 0: return orc_suw7(f);
    case 1: return orc_weno7(f);
    case 2: return orc_weno7z(f);
    case 3: return orc_mp7(f);
    case 5: return orc_mp7ld(f, bfacmpld, shock);
    case 6: return orc_round(f + 2);
    }
    return NAN;
}

/* ------------------------------------------------------- flux splitting */

/* flux_steger_warming at one node, src/riemann.F90:23-161 (no species),
 * sound speed sos = sqrt(T)/Ma (src/fludyna.F90:832-859, nondimen). */
void orc_sw_split(double gamma, double mach, double rho, const double *vel, double prs,
                  double tmp, const double *q, const double *dxi, double jacob, double *fp,
                  double *fm)
{
    const double eps = 0.04;
    const double uu = dxi[0] * vel[0] + dxi[1] * vel[1] + dxi[2] * vel[2];
    const double var0 = 1.0 / sqrt(dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2]);
    const double gpd[3] = {dxi[0] * var0, dxi[1] * var0, dxi[2] * var0};
    const double gm2 = 0.5 / gamma;
    const double css = sqrt(tmp) / mach;
    const double csa = css / var0;
    const double lmach = uu / csa;
    const double lmda[5] = {uu, uu, uu, uu + csa, uu - csa};
    double lp[5], lm[5];
    for (int k = 0; k < 5; ++k) {
        lp[k] = 0.5 * (lmda[k] + sqrt(lmda[k] * lmda[k] + eps * eps));
        lm[k] = lmda[k] - lp[k];
    }
    if (lmach >= 1.0 || lmach <= -1.0) {
        double *f = lmach >= 1.0 ? fp : fm, *z = lmach >= 1.0 ? fm : fp;
        f[0] = jacob * q[0] * uu;
        f[1] = jacob * (q[1] * uu + dxi[0] * prs);
        f[2] = jacob * (q[2] * uu + dxi[1] * prs);
        f[3] = jacob * (q[3] * uu + dxi[2] * prs);
        f[4] = jacob * (q[4] + prs) * uu;
        for (int k = 0; k < 5; ++k) z[k] = 0.0;
        return;
    }
    const double fhi = 0.5 * (gamma - 1.0) * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double jro = jacob * rho;
    for (int s = 0; s < 2; ++s) {
        const double *l = s == 0 ? lp : lm;
        double *f = s == 0 ? fp : fm;
        const double var1 = l[0];
        const double var2 = l[3] - l[4];
        const double var3 = 2.0 * l[0] - l[3] - l[4];
        const double var4 = var1 - var3 * gm2;
        f[0] = jro * var4;
        f[1] = jro * (var4 * vel[0] + var2 * css * gpd[0] * gm2);
        f[2] = jro * (var4 * vel[1] + var2 * css * gpd[1] * gm2);
        f[3] = jro * (var4 * vel[2] + var2 * css * gpd[2] * gm2);
        f[4] = jacob * (var1 * q[4] + rho * (var2 * uu * var0 * css * gm2 -
                                             var3 * (fhi + css * css) * gm2 / (gamma - 1.0)));
    }
}

/* chardecomp, src/solver.F90:1964-2163 (no species): Roe-averaged left
 * and right eigenvectors along the mean metric normal.  LEV, REV are
 * row-major 5x5, L[m][n] = LEV(m+1,n+1).  Returns 1 when the normal is
 * zero (the reference stops). */
int orc_chardecomp(double gamma, double tmp_l, double ro_l, double p_l, double e_l,
                   const double *vel_l, const double *ddi_l, double tmp_r, double ro_r,
                   double p_r, double e_r, const double *vel_r, const double *ddi_r, double *R,
                   double *Lv)
{
    (void)tmp_l;
    (void)tmp_r;
    const double rero = 1.e-12;
#define LEV(a, b) Lv[((a)-1) * 5 + (b)-1]
#define REV(a, b) R[((a)-1) * 5 + (b)-1]
    const double sl = sqrt(ro_l), sr = sqrt(ro_r);
    const double wroe = sl / (sl + sr);
    const double wroe1 = 1.0 - wroe;
    const double u1 = wroe * vel_l[0] + wroe1 * vel_r[0];
    const double u2 = wroe * vel_l[1] + wroe1 * vel_r[1];
    const double u3 = wroe * vel_l[2] + wroe1 * vel_r[2];
    const double kroe = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
    const double hl = (e_l + p_l) / ro_l;
    const double hr = (e_r + p_r) / ro_r;
    const double hroe = wroe * hl + wroe1 * hr;
    const double cs = sqrt((gamma - 1.0) * (hroe - kroe));
    const double rcs = 1.0 / cs;
    const double var1 = 0.5 * (ddi_l[0] + ddi_r[0]);
    const double var2 = 0.5 * (ddi_l[1] + ddi_r[1]);
    const double var3 = 0.5 * (ddi_l[2] + ddi_r[2]);
    const double var4 = 1.0 / sqrt(var1 * var1 + var2 * var2 + var3 * var3);
    const double g1 = var1 * var4, g2 = var2 * var4, g3 = var3 * var4;
    const double ugp = u1 * g1 + u2 * g2 + u3 * g3;
    const double b1 = (gamma - 1.0) / (cs * cs);
    const double b2 = 1.0 + 2.0 * b1 * kroe - b1 * hroe;

    LEV(1, 1) = 0.5 * (b2 + ugp * rcs);
    LEV(1, 2) = -0.5 * (b1 * u1 + g1 * rcs);
    LEV(1, 3) = -0.5 * (b1 * u2 + g2 * rcs);
    LEV(1, 4) = -0.5 * (b1 * u3 + g3 * rcs);
    LEV(1, 5) = 0.5 * b1;
    REV(1, 1) = 1.0;
    REV(2, 1) = u1 - cs * g1;
    REV(3, 1) = u2 - cs * g2;
    REV(4, 1) = u3 - cs * g3;
    REV(5, 1) = hroe - ugp * cs;

    LEV(2, 1) = 1.0 - b2;
    LEV(2, 2) = b1 * u1;
    LEV(2, 3) = b1 * u2;
    LEV(2, 4) = b1 * u3;
    LEV(2, 5) = -b1;
    REV(1, 2) = 1.0;
    REV(2, 2) = u1;
    REV(3, 2) = u2;
    REV(4, 2) = u3;
    REV(5, 2) = hroe - 1.0 / b1;

    if (fabs(var1) > rero) {
        const double rgp = 1.0 / g1;
        LEV(3, 1) = (ugp * g2 - u2) * rgp;
        LEV(3, 2) = -g2;
        LEV(3, 3) = (1.0 - g2 * g2) * rgp;
        LEV(3, 4) = -g2 * g3 * rgp;
        LEV(3, 5) = 0.0;
        LEV(4, 1) = (ugp * g3 - u3) * rgp;
        LEV(4, 2) = -g3;
        LEV(4, 3) = -g2 * g3 * rgp;
        LEV(4, 4) = (1.0 - g3 * g3) * rgp;
        LEV(4, 5) = 0.0;
        REV(1, 3) = 0.0;
        REV(2, 3) = -g2;
        REV(3, 3) = g1;
        REV(4, 3) = 0.0;
        REV(5, 3) = u2 * g1 - u1 * g2;
        REV(1, 4) = 0.0;
        REV(2, 4) = -g3;
        REV(3, 4) = 0.0;
        REV(4, 4) = g1;
        REV(5, 4) = u3 * g1 - u1 * g3;
    } else if (fabs(var2) > rero) {
        const double rgp = 1.0 / g2;
        LEV(3, 1) = (ugp * g1 - u1) * rgp;
        LEV(3, 2) = (1.0 - g1 * g1) * rgp;
        LEV(3, 3) = -g1;
        LEV(3, 4) = -g1 * g3 * rgp;
        LEV(3, 5) = 0.0;
        LEV(4, 1) = (ugp * g3 - u3) * rgp;
        LEV(4, 2) = -g1 * g3 * rgp;
        LEV(4, 3) = -g3;
        LEV(4, 4) = (1.0 - g3 * g3) * rgp;
        LEV(4, 5) = 0.0;
        REV(1, 3) = 0.0;
        REV(2, 3) = g2;
        REV(3, 3) = -g1;
        REV(4, 3) = 0.0;
        REV(5, 3) = u1 * g2 - u2 * g1;
        REV(1, 4) = 0.0;
        REV(2, 4) = 0.0;
        REV(3, 4) = -g3;
        REV(4, 4) = g2;
        REV(5, 4) = u3 * g2 - u2 * g3;
    } else if (fabs(var3) > rero) {
        const double rgp = 1.0 / g3;
        LEV(3, 1) = (ugp * g1 - u1) * rgp;
        LEV(3, 2) = (1.0 - g1 * g1) * rgp;
        LEV(3, 3) = -g1 * g2 * rgp;
        LEV(3, 4) = -g1;
        LEV(3, 5) = 0.0;
        LEV(4, 1) = (ugp * g2 - u2) * rgp;
        LEV(4, 2) = -g1 * g2 * rgp;
        LEV(4, 3) = (1.0 - g2 * g2) * rgp;
        LEV(4, 4) = -g2;
        LEV(4, 5) = 0.0;
        REV(1, 3) = 0.0;
        REV(2, 3) = g3;
        REV(3, 3) = 0.0;
        REV(4, 3) = -g1;
        REV(5, 3) = u1 * g3 - u3 * g1;
        REV(1, 4) = 0.0;
        REV(2, 4) = 0.0;
        REV(3, 4) = g3;
        REV(4, 4) = -g2;
        REV(5, 4) = u2 * g3 - u3 * g2;
    } else {
        return 1;
    }

    LEV(5, 1) = 0.5 * (b2 - ugp * rcs);
    LEV(5, 2) = -0.5 * (b1 * u1 - g1 * rcs);
    LEV(5, 3) = -0.5 * (b1 * u2 - g2 * rcs);
    LEV(5, 4) = -0.5 * (b1 * u3 - g3 * rcs);
    LEV(5, 5) = 0.5 * b1;
    REV(1, 5) = 1.0;
    REV(2, 5) = u1 + cs * g1;
    REV(3, 5) = u2 + cs * g2;
    REV(4, 5) = u3 + cs * g3;
    REV(5, 5) = hroe + ugp * cs;
#undef LEV
#undef REV
    return 0;
}

/* iwind8, src/solver.F90:1214-1235: stencil node n (1..8) of interface
 * i+1/2 for the + or - flux, clamped to [imin, imax] */
int orc_iwind8(int i, int n, int imin, int imax, int plus)
{
    int w = plus ? i + n - 4 : i + 5 - n;
    if (w < imin) w = imin;
    if (w > imax) w = imax;
    return w;
}

/* ------------------------------------------------------------- box level */

typedef struct {
    long ni, nj, nk, nf;  /* haloed extents, elements per haloed field */
    long ii, ij, ik, nint;
} shape_t;

static shape_t shape_of(const orc_uwd *u)
{
    shape_t s;
    s.ni = u->im + 1 + 2 * HMU;
    s.nj = u->jm + 1 + 2 * HMU;
    s.nk = u->km + 1 + 2 * HMU;
    s.nf = s.ni * s.nj * s.nk;
    s.ii = u->im + 1;
    s.ij = u->jm + 1;
    s.ik = u->km + 1;
    s.nint = s.ii * s.ij * s.ik;
    return s;
}

static long hidx(const shape_t *s, int i, int j, int k)
{
    return (i + HMU) + s->ni * ((j + HMU) + s->nj * (long)(k + HMU));
}

static long iidx(const shape_t *s, int i, int j, int k)
{
    return i + s->ii * (j + s->ij * (long)k);
}

/* is..ie of an axis: single-rank ranges of src/parallel.F90:1043-1160
 * for ntype 1 (is=1), 2 (ie=n-1), 3 (0..n); ntype 4 (both ends bounded,
 * one rank) leaves is/ie unset in the reference and is taken as 1..n-1. */
void orc_uwd_range(int ntype, int n, int *lo, int *hi)
{
    *lo = (ntype == 1 || ntype == 4) ? 1 : 0;
    *hi = (ntype == 2 || ntype == 4) ? n - 1 : n;
}

int orc_ndims(const orc_uwd *u) { return u->km == 0 ? 2 : 3; }

/* gradcal, src/comsolver.F90:244-497 (no species, turbmode none):
 * dvel(m,n) = sum_d d/dxi_d(u_m) * dxi(d,n), accumulated from zero in the
 * order d = 1, 2, 3; the derivative is df_compact (explicit_fd = 0) or
 * diff6ec (explicit_fd = 1) with the axis' ntype. */
void orc_gradcal(const orc_uwd *u, int explicit_fd, const double *vel, const double *tmp,
                 const double *dxi, double *dvel, double *dtmp)
{
    const shape_t s = shape_of(u);
    const int nd = orc_ndims(u);
    const int dims[3] = {u->im, u->jm, u->km};
    memset(dvel, 0, sizeof(double) * 9 * s.nint);
    memset(dtmp, 0, sizeof(double) * 3 * s.nint);
    for (int d = 0; d < nd; ++d) {
        const int n = dims[d];
        double *f = malloc(sizeof(double) * (n + 1 + 2 * HMU));
        double *df = malloc(sizeof(double) * (n + 1) * 4);
        const int na = d == 0 ? u->jm : u->im, nb = d == 2 ? u->jm : u->km;
        for (int b = 0; b <= nb; ++b)
            for (int a = 0; a <= na; ++a) {
                for (int c = 0; c < 4; ++c) {
                    const double *src = c < 3 ? vel + c * s.nf : tmp;
                    for (int t = -HMU; t <= n + HMU; ++t) {
                        const int i = d == 0 ? t : a, j = d == 1 ? t : (d == 0 ? a : b);
                        const int k = d == 2 ? t : b;
                        f[t + HMU] = src[hidx(&s, i, j, k)];
                    }
                    if (explicit_fd) orc_diff6ec(u->npdc[d], n, f, df + c * (n + 1));
                    else orc_df_compact(u->npdc[d], n, f, df + c * (n + 1));
                }
                for (int t = 0; t <= n; ++t) {
                    const int i = d == 0 ? t : a, j = d == 1 ? t : (d == 0 ? a : b);
                    const int k = d == 2 ? t : b;
                    const long p = iidx(&s, i, j, k), h = hidx(&s, i, j, k);
                    for (int c = 0; c < 4; ++c) {
                        const double g = df[c * (n + 1) + t];
                        for (int nn = 0; nn < 3; ++nn) {
                            const double x = dxi[(d + 3 * nn) * s.nf + h];
                            if (c < 3) dvel[(c + 3 * nn) * s.nint + p] += g * x;
                            else dtmp[nn * s.nint + p] += g * x;
                        }
                    }
                }
            }
        free(f);
        free(df);
    }
}

/* ducrossensor, src/commcal.F90:196-353: ssf on 0..n, its halo swap
 * (periodic axes: plain copy, src/parallel.F90:4170-4175), then lshock =
 * (max of ssf over -hm+1..hm along each axis) > shkcrt.  The pressure
 * stencil and the window are clamped at a bounded end: the low end of
 * ntype 1 and the high end of ntype 2 (as in the reference), and both ends
 * of ntype 4.  ntype 4 (one domain bounded at both ends) is the union of
 * the two: the reference assigns it only to a single rank along a
 * non-periodic axis, where it leaves is/ie unset (src/parallel.F90:
 * 1070-1078), so no reference run defines these reads; a run on several
 * ranks, which the reference can complete, has ntype 1 and 2 at the
 * physical ends and clamps there.  Elsewhere the stencils read halos: the
 * pressure halos as given, ssf halos of non-periodic ends zero. */
void orc_ducros(const orc_uwd *u, const double *dvel, const double *prs, double *ssf,
                unsigned char *lshock)
{
    const shape_t s = shape_of(u);
    const int n[3] = {u->im, u->jm, u->km};
    const int nd = orc_ndims(u);
    for (long e = 0; e < s.nf; ++e) ssf[e] = 0.0;
    for (int k = 0; k <= u->km; ++k)
        for (int j = 0; j <= u->jm; ++j)
            for (int i = 0; i <= u->im; ++i) {
                const long p = iidx(&s, i, j, k);
#define DV(m, nn) dvel[((m)-1 + 3 * ((nn)-1)) * s.nint + p]
                double div = DV(1, 1) + DV(2, 2) + DV(3, 3);
                const double div2 = div * div;
                const double vx = DV(3, 2) - DV(2, 3);
                const double vy = DV(1, 3) - DV(3, 1);
                const double vz = DV(2, 1) - DV(1, 2);
#undef DV
                const double vort = vx * vx + vy * vy + vz * vz;
                int ip1 = i + 1, im1 = i - 1, jp1 = j + 1, jm1 = j - 1, kp1 = k + 1, km1 = k - 1;
                if (LO_END(u->npdc[0]) && im1 < 0) im1 = 0;
                if (LO_END(u->npdc[1]) && jm1 < 0) jm1 = 0;
                if (LO_END(u->npdc[2]) && km1 < 0) km1 = 0;
                if (HI_END(u->npdc[0]) && ip1 > u->im) ip1 = u->im;
                if (HI_END(u->npdc[1]) && jp1 > u->jm) jp1 = u->jm;
                if (HI_END(u->npdc[2]) && kp1 > u->km) kp1 = u->km;
                const double pc = prs[hidx(&s, i, j, k)];
#define P(a, b, c) prs[hidx(&s, a, b, c)]
                const double dpdi = fabs(P(ip1, j, k) - 2.0 * pc + P(im1, j, k)) /
                                    (P(ip1, j, k) + 2.0 * pc + P(im1, j, k));
                const double dpdj = fabs(P(i, jp1, k) - 2.0 * pc + P(i, jm1, k)) /
                                    (P(i, jp1, k) + 2.0 * pc + P(i, jm1, k));
                const double dpdk = fabs(P(i, j, kp1) - 2.0 * pc + P(i, j, km1)) /
                                    (P(i, j, kp1) + 2.0 * pc + P(i, j, km1));
#undef P
                ssf[hidx(&s, i, j, k)] =
                    div2 / (div2 + vort + 1.e-30) * max2(max2(dpdi, dpdj), dpdk);
            }
    /* halo swap of ssf along the periodic axes, in the order i, j, k, each
     * over the full extent of the other axes */
    for (int d = 0; d < nd; ++d) {
        if (u->npdc[d] != 3) continue;
        for (long e = 0; e < s.nf; ++e) {
            int c[3];
            c[0] = (int)(e % s.ni) - HMU;
            c[1] = (int)((e / s.ni) % s.nj) - HMU;
            c[2] = (int)(e / (s.ni * s.nj)) - HMU;
            int t = c[d];
            if (t >= 0 && t <= n[d]) continue;
            c[d] = t < 0 ? t + n[d] : t - n[d];
            ssf[e] = ssf[hidx(&s, c[0], c[1], c[2])];
        }
    }
    for (int k = 0; k <= u->km; ++k)
        for (int j = 0; j <= u->jm; ++j)
            for (int i = 0; i <= u->im; ++i) {
                double m = 0.0;
                const int c0[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    for (int o = -HMU + 1; o <= HMU; ++o) {
                        int c[3] = {c0[0], c0[1], c0[2]};
                        int t = c[d] + o;
                        if (LO_END(u->npdc[d]) && t < 0) t = 0;
                        if (HI_END(u->npdc[d]) && t > n[d]) t = n[d];
                        c[d] = t;
                        m = max2(m, ssf[hidx(&s, c[0], c[1], c[2])]);
                    }
                }
                lshock[iidx(&s, i, j, k)] = m > u->shkcrt;
            }
}

/* convrsduwd, src/solver.F90:549-1202 (no species, crinod false, no
 * solids): per axis, Steger-Warming split fluxes on iss..iee, the
 * interface flux Fh(i+1/2) for i = is-1..ie (characteristic projection
 * when lchardecomp and the shock flag), qrhs += Fh(i) - Fh(i-1) on is..ie.
 * lshock NULL: the sensor is off (sson false, lsh true).  Returns nonzero
 * on an unsupported ntype (the reference stops: ntype 4 on j and k). */
int orc_convrsduwd(const orc_uwd *u, const double *q, const double *rho, const double *vel,
                   const double *prs, const double *tmp, const double *dxi, const double *jacob,
                   const unsigned char *lshock, double *qrhs)
{
    const shape_t s = shape_of(u);
    const int nd = orc_ndims(u);
    const int n[3] = {u->im, u->jm, u->km};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) orc_uwd_range(u->npdc[d], n[d], &lo[d], &hi[d]);
    if (nd == 2) lo[2] = hi[2] = 0;
    for (int d = 0; d < nd; ++d) {
        const int nt = u->npdc[d];
        if (nt < 1 || nt > 4) return 1;
        const int iss = (nt == 1 || nt == 4) ? 0 : -HMU;
        const int iee = (nt == 2 || nt == 4) ? n[d] : n[d] + HMU;
        const int len = n[d] + 1 + 2 * HMU;
        double *fp = malloc(sizeof(double) * 5 * len), *fm = malloc(sizeof(double) * 5 * len);
        double *fh = malloc(sizeof(double) * 5 * (n[d] + 2));  /* Fh(-1:n) */
        /* lines: d=0 over (j,k) in js..je x ks..ke, d=1 over (i,k), d=2 (i,j) */
        const int a0 = d == 0 ? lo[1] : lo[0], a1 = d == 0 ? hi[1] : hi[0];
        const int b0 = d == 2 ? lo[1] : lo[2], b1 = d == 2 ? hi[1] : hi[2];
        for (int b = b0; b <= b1; ++b)
            for (int a = a0; a <= a1; ++a) {
#define NODE(t, I, J, K)                                   \
    const int I = d == 0 ? (t) : a;                        \
    const int J = d == 1 ? (t) : (d == 0 ? a : b);         \
    const int K = d == 2 ? (t) : b;
                for (int t = iss; t <= iee; ++t) {
                    NODE(t, i, j, k)
                    const long h = hidx(&s, i, j, k);
                    double v[3], qq[5], dd[3];
                    for (int c = 0; c < 3; ++c) {
                        v[c] = vel[c * s.nf + h];
                        dd[c] = dxi[(d + 3 * c) * s.nf + h];
                    }
                    for (int m = 0; m < 5; ++m) qq[m] = q[m * s.nf + h];
                    orc_sw_split(u->gamma, u->mach, rho[h], v, prs[h], tmp[h], qq, dd, jacob[h],
                                 fp + 5 * (t + HMU), fm + 5 * (t + HMU));
                }
                for (int i = lo[d] - 1; i <= hi[d]; ++i) {
                    int lsh = 1;
                    if (lshock) {
                        NODE(i < 0 ? i + 1 : i, i1, j1, k1)
                        NODE(i + 1 > n[d] ? i : i + 1, i2, j2, k2)
                        lsh = lshock[iidx(&s, i1, j1, k1)] || lshock[iidx(&s, i2, j2, k2)];
                    }
                    double Lv[25], Rv[25], fcp[5][8], fcm[5][8], fhc[5];
                    const int chr = u->lchardecomp && lsh;
                    if (chr) {
                        NODE(i, il, jl, kl)
                        NODE(i + 1, ir, jr, kr)
                        const long hl = hidx(&s, il, jl, kl), hr = hidx(&s, ir, jr, kr);
                        double vl[3], vr[3], dl[3], dr[3];
                        for (int c = 0; c < 3; ++c) {
                            vl[c] = vel[c * s.nf + hl];
                            vr[c] = vel[c * s.nf + hr];
                            dl[c] = dxi[(d + 3 * c) * s.nf + hl];
                            dr[c] = dxi[(d + 3 * c) * s.nf + hr];
                        }
                        if (orc_chardecomp(u->gamma, tmp[hl], rho[hl], prs[hl], q[4 * s.nf + hl],
                                           vl, dl, tmp[hr], rho[hr], prs[hr], q[4 * s.nf + hr],
                                           vr, dr, Rv, Lv))
                            return 2;
                        for (int m = 0; m < 5; ++m)
                            for (int w = 1; w <= 8; ++w) {
                                const double *P = fp + 5 * (orc_iwind8(i, w, iss, iee, 1) + HMU);
                                const double *M = fm + 5 * (orc_iwind8(i, w, iss, iee, 0) + HMU);
                                const double *l = Lv + 5 * m;
                                fcp[m][w - 1] = l[0] * P[0] + l[1] * P[1] + l[2] * P[2] +
                                                l[3] * P[3] + l[4] * P[4];
                                fcm[m][w - 1] = l[0] * M[0] + l[1] * M[1] + l[2] * M[2] +
                                                l[3] * M[3] + l[4] * M[4];
                            }
                    } else {
                        for (int w = 1; w <= 8; ++w) {
                            const double *P = fp + 5 * (orc_iwind8(i, w, iss, iee, 1) + HMU);
                            const double *M = fm + 5 * (orc_iwind8(i, w, iss, iee, 0) + HMU);
                            for (int m = 0; m < 5; ++m) {
                                fcp[m][w - 1] = P[m];
                                fcm[m][w - 1] = M[m];
                            }
                        }
                    }
                    const int nte = end_type(nt, i, n[d]);
                    for (int m = 0; m < 5; ++m) {
                        const double v1 = orc_recons_exp(fcp[m], i, n[d], nte, u->recon_schem, lsh,
                                                         u->bfacmpld);
                        const double v2 = orc_recons_exp(fcm[m], i, n[d], nte, u->recon_schem, lsh,
                                                         u->bfacmpld);
                        fhc[m] = v1 + v2;
                    }
                    double *F = fh + 5 * (i + 1);
                    if (chr) {
                        for (int m = 0; m < 5; ++m) {
                            const double *r = Rv + 5 * m;
                            F[m] = r[0] * fhc[0] + r[1] * fhc[1] + r[2] * fhc[2] + r[3] * fhc[3] +
                                   r[4] * fhc[4];
                        }
                    } else {
                        for (int m = 0; m < 5; ++m) F[m] = fhc[m];
                    }
                }
                for (int i = lo[d]; i <= hi[d]; ++i) {
                    NODE(i, ii, jj, kk)
                    const long p = iidx(&s, ii, jj, kk);
                    for (int m = 0; m < 5; ++m)
                        qrhs[m * s.nint + p] = qrhs[m * s.nint + p] + fh[5 * (i + 1) + m] -
                                               fh[5 * i + m];
                }
#undef NODE
            }
        free(fp);
        free(fm);
        free(fh);
    }
    return 0;
}

/* rhscal for an odd-order explicit convection scheme (conschm(4:4)='e'),
 * src/solver.F90:184-283 without viscous terms: qrhs = 0, gradcal,
 * ducrossensor when recon_schem = 5 or lchardecomp, convrsduwd, then
 * qrhs = -qrhs.  ssf/lshock may be NULL when the sensor is not run. */
int orc_uwd_rhscal(const orc_uwd *u, int explicit_fd, const double *q, const double *rho,
                   const double *vel, const double *prs, const double *tmp, const double *dxi,
                   const double *jacob, double *qrhs, double *dvel, double *dtmp, double *ssf,
                   unsigned char *lshock)
{
    const shape_t s = shape_of(u);
    memset(qrhs, 0, sizeof(double) * 5 * s.nint);
    orc_gradcal(u, explicit_fd, vel, tmp, dxi, dvel, dtmp);
    const int sensor = u->recon_schem == 5 || u->lchardecomp;
    if (sensor) orc_ducros(u, dvel, prs, ssf, lshock);
    const int rc = orc_convrsduwd(u, q, rho, vel, prs, tmp, dxi, jacob, sensor ? lshock : NULL,
                                  qrhs);
    if (rc) return rc;
    for (long e = 0; e < 5 * s.nint; ++e) qrhs[e] = -qrhs[e];
    return 0;
}

/* ------------------------------------------------ compact upwind ('543c') */

static const double n1d6 = 1.0 / 6.0, n37d60 = 37.0 / 60.0, n2d15 = 2.0 / 15.0,
                    n1d60 = 1.0 / 60.0, n1d18 = 1.0 / 18.0, n1d36 = 1.0 / 36.0,
                    n19d18 = 19.0 / 18.0, n9d36 = 9.0 / 36.0, n5d9 = 5.0 / 9.0;

/* compact_flux_initiate, src/flux.F90:32-108 (scheme 543): rows first..last
 * (ntype 1: -1..dim+1, 2: -2..dim, 3: -2..dim+1, 4: -1..dim), the LHS of
 * the 5th-order compact upwind interface flux for the + or - wind, and its
 * Thomas factors ac[3*row].  Returns the row count. */
int orc_flux_compact_init(int ntype, int dim, int plus, double bfac, int *first, int *last,
                          double *ac)
{
    const int i0 = (ntype == 1 || ntype == 4) ? -1 : -2;
    const int im_ = (ntype == 2 || ntype == 4) ? dim : dim + 1;
    const int n = im_ - i0 + 1;
    double *a = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
    for (int r = 0; r < n; ++r) {
        a[r] = plus ? 0.5 - n1d6 * bfac : n1d6 + n1d6 * bfac;
        c[r] = plus ? n1d6 + n1d6 * bfac : 0.5 - n1d6 * bfac;
    }
    a[0] = c[0] = 0.0;
    a[n - 1] = c[n - 1] = 0.0;
    if (ntype == 1 || ntype == 4) {
        a[0] = c[0] = 2.0;
        a[1] = c[1] = 0.25;
    }
    if (ntype == 2 || ntype == 4) {
        a[n - 1] = c[n - 1] = 2.0;
        a[n - 2] = c[n - 2] = 0.25;
    }
    if (ac) orc_thomas_prep(a, c, n, ac);
    free(a);
    free(c);
    *first = i0;
    *last = im_;
    return n;
}

/* compact_flux_rhs, src/flux.F90:158-256, rows first..last; f(node) =
 * f[node + HMU], d indexed by row - first */
static void flux_rhs(int ntype, int i0, int im_, int plus, double bfac, const double *f,
                     double *d)
{
#define F(nd) f[(nd) + HMU]
#define D(nd) d[(nd) - i0]
    int is = i0 + 1, ie = im_ - 1, j;
    if (ntype == 1 || ntype == 4) {
        j = i0;
        D(j) = 2.5 * F(j + 1) + 0.5 * F(j + 2);
        j = i0 + 1;
        D(j) = 0.75 * F(j) + 0.75 * F(j + 1);
        is = i0 + 2;
    } else {
        j = i0;
        const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
        D(j) = n37d60 * v1 - n2d15 * v2 + n1d60 * v3;
    }
    if (ntype == 2 || ntype == 4) {
        j = im_ - 1;
        D(j) = 0.75 * F(j + 1) + 0.75 * F(j);
        j = im_;
        D(j) = 2.5 * F(j) + 0.5 * F(j - 1);
        ie = im_ - 2;
    } else {
        j = im_;
        const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
        D(j) = n37d60 * v1 - n2d15 * v2 + n1d60 * v3;
    }
    for (j = is; j <= ie; ++j) {
        if (plus)
            D(j) = (n1d18 - n1d36 * bfac) * F(j - 1) + (n19d18 - n9d36 * bfac) * F(j) +
                   (n5d9 + n9d36 * bfac) * F(j + 1) + n1d36 * bfac * F(j + 2);
        else
            D(j) = (n1d18 - n1d36 * bfac) * F(j + 2) + (n19d18 - n9d36 * bfac) * F(j + 1) +
                   (n5d9 + n9d36 * bfac) * F(j) + n1d36 * bfac * F(j - 1);
    }
#undef F
#undef D
}

/* flux_compact, src/flux.F90:125-147: fh(-1:dim) of f(-hm:dim+hm) */
void orc_flux_compact(int ntype, int dim, int plus, double bfac, const double *f, double *fh)
{
    int i0, im_;
    const int n = orc_flux_compact_init(ntype, dim, plus, bfac, &i0, &im_, NULL);
    double *ac = malloc(sizeof(double) * 3 * n), *d = malloc(sizeof(double) * n),
           *x = malloc(sizeof(double) * n);
    orc_flux_compact_init(ntype, dim, plus, bfac, &i0, &im_, ac);
    flux_rhs(ntype, i0, im_, plus, bfac, f, d);
    orc_thomas_solve(ac, d, n, x);
    for (int i = -1; i <= dim; ++i) fh[i + 1] = x[i - i0];
    free(ac);
    free(d);
    free(x);
}

/* mplimiter, src/flux.F90:363-379 */
double orc_mplimiter(const double *f, double fl, int shock, int inode, int dim, int ntype)
{
    if ((ntype == 1 && inode == 0) || (ntype == 2 && inode == dim - 1) ||
        (ntype == 1 && inode == 1) || (ntype == 2 && inode == dim - 2))
        return fl;
    return orc_mp5(f, &fl, shock);
}

/* iwind6, src/solver.F90:1237-1258 */
static int iwind6(int i, int n, int imin, int imax, int plus)
{
    int w = plus ? i + n - 3 : i + 4 - n;
    if (w < imin) w = imin;
    if (w > imax) w = imax;
    return w;
}

/* convrsdcmp, src/solver.F90:1272-1938 (no species, crinod false): per
 * axis, Steger-Warming split, the compact interface fluxes of F+ and F-,
 * the MP limiter about them on 5-point stencils (characteristic variables
 * when lchardecomp, for every interface), qrhs += Fh(i) - Fh(i-1). */
int orc_convrsdcmp(const orc_uwd *u, const double *q, const double *rho, const double *vel,
                   const double *prs, const double *tmp, const double *dxi, const double *jacob,
                   const unsigned char *lshock, double *qrhs)
{
    const shape_t s = shape_of(u);
    const int nd = orc_ndims(u);
    const int n[3] = {u->im, u->jm, u->km};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) orc_uwd_range(u->npdc[d], n[d], &lo[d], &hi[d]);
    if (nd == 2) lo[2] = hi[2] = 0;
    for (int d = 0; d < nd; ++d) {
        const int nt = u->npdc[d];
        if (nt < 1 || nt > 4) return 1;
        const int iss = (nt == 1 || nt == 4) ? 0 : -HMU;
        const int iee = (nt == 2 || nt == 4) ? n[d] : n[d] + HMU;
        const int len = n[d] + 1 + 2 * HMU;
        double *fp = calloc(5 * len, sizeof(double)), *fm = calloc(5 * len, sizeof(double));
        double *line = malloc(sizeof(double) * len);
        double *hp = malloc(sizeof(double) * 5 * (n[d] + 2)), *hmn = malloc(sizeof(double) * 5 * (n[d] + 2));
        double *fh = malloc(sizeof(double) * 5 * (n[d] + 2));
        const int a0 = d == 0 ? lo[1] : lo[0], a1 = d == 0 ? hi[1] : hi[0];
        const int b0 = d == 2 ? lo[1] : lo[2], b1 = d == 2 ? hi[1] : hi[2];
        for (int b = b0; b <= b1; ++b)
            for (int a = a0; a <= a1; ++a) {
#define NODE(t, I, J, K)                                   \
    const int I = d == 0 ? (t) : a;                        \
    const int J = d == 1 ? (t) : (d == 0 ? a : b);         \
    const int K = d == 2 ? (t) : b;
                for (int t = iss; t <= iee; ++t) {
                    NODE(t, i, j, k)
                    const long h = hidx(&s, i, j, k);
                    double v[3], qq[5], dd[3];
                    for (int c = 0; c < 3; ++c) {
                        v[c] = vel[c * s.nf + h];
                        dd[c] = dxi[(d + 3 * c) * s.nf + h];
                    }
                    for (int m = 0; m < 5; ++m) qq[m] = q[m * s.nf + h];
                    double P[5], M[5];
                    orc_sw_split(u->gamma, u->mach, rho[h], v, prs[h], tmp[h], qq, dd, jacob[h], P,
                                 M);
                    for (int m = 0; m < 5; ++m) {
                        fp[m * len + t + HMU] = P[m];
                        fm[m * len + t + HMU] = M[m];
                    }
                }
                for (int m = 0; m < 5; ++m) {
                    orc_flux_compact(nt, n[d], 1, u->bfacmpld, fp + m * len, hp + m * (n[d] + 2));
                    orc_flux_compact(nt, n[d], 0, u->bfacmpld, fm + m * len, hmn + m * (n[d] + 2));
                }
                (void)line;
                for (int i = lo[d] - 1; i <= hi[d]; ++i) {
                    double Lv[25], Rv[25], fcp[5][5], fcm[5][5], cp[5], cm[5], fhc[5];
                    if (u->lchardecomp) {
                        NODE(i, il, jl, kl)
                        NODE(i + 1, ir, jr, kr)
                        const long hl = hidx(&s, il, jl, kl), hr = hidx(&s, ir, jr, kr);
                        double vl[3], vr[3], dl[3], dr[3];
                        for (int c = 0; c < 3; ++c) {
                            vl[c] = vel[c * s.nf + hl];
                            vr[c] = vel[c * s.nf + hr];
                            dl[c] = dxi[(d + 3 * c) * s.nf + hl];
                            dr[c] = dxi[(d + 3 * c) * s.nf + hr];
                        }
                        if (orc_chardecomp(u->gamma, tmp[hl], rho[hl], prs[hl], q[4 * s.nf + hl],
                                           vl, dl, tmp[hr], rho[hr], prs[hr], q[4 * s.nf + hr],
                                           vr, dr, Rv, Lv))
                            return 2;
                        for (int m = 0; m < 5; ++m) {
                            const double *l = Lv + 5 * m;
                            for (int w = 1; w <= 5; ++w) {
                                const int tp = iwind6(i, w, iss, iee, 1) + HMU;
                                const int tm = iwind6(i, w, iss, iee, 0) + HMU;
                                fcp[m][w - 1] = l[0] * fp[tp] + l[1] * fp[len + tp] +
                                                l[2] * fp[2 * len + tp] + l[3] * fp[3 * len + tp] +
                                                l[4] * fp[4 * len + tp];
                                fcm[m][w - 1] = l[0] * fm[tm] + l[1] * fm[len + tm] +
                                                l[2] * fm[2 * len + tm] + l[3] * fm[3 * len + tm] +
                                                l[4] * fm[4 * len + tm];
                            }
                            const int ix = i + 1;
                            const int st = n[d] + 2;
                            cp[m] = l[0] * hp[ix] + l[1] * hp[st + ix] + l[2] * hp[2 * st + ix] +
                                    l[3] * hp[3 * st + ix] + l[4] * hp[4 * st + ix];
                            cm[m] = l[0] * hmn[ix] + l[1] * hmn[st + ix] + l[2] * hmn[2 * st + ix] +
                                    l[3] * hmn[3 * st + ix] + l[4] * hmn[4 * st + ix];
                        }
                    } else {
                        for (int m = 0; m < 5; ++m) {
                            for (int w = 1; w <= 5; ++w) {
                                fcp[m][w - 1] = fp[m * len + iwind6(i, w, iss, iee, 1) + HMU];
                                fcm[m][w - 1] = fm[m * len + iwind6(i, w, iss, iee, 0) + HMU];
                            }
                            cp[m] = hp[m * (n[d] + 2) + i + 1];
                            cm[m] = hmn[m * (n[d] + 2) + i + 1];
                        }
                    }
                    int lsh = 1;
                    if (lshock) {
                        NODE(i < 0 ? i + 1 : i, i1, j1, k1)
                        NODE(i + 1 > n[d] ? i : i + 1, i2, j2, k2)
                        lsh = lshock[iidx(&s, i1, j1, k1)] || lshock[iidx(&s, i2, j2, k2)];
                    }
                    for (int m = 0; m < 5; ++m) {
                        const double v1 = orc_mplimiter(fcp[m], cp[m], lsh, i, n[d], nt);
                        const double v2 = orc_mplimiter(fcm[m], cm[m], lsh, i, n[d], nt);
                        fhc[m] = v1 + v2;
                    }
                    double *F = fh + 5 * (i + 1);
                    if (u->lchardecomp) {
                        for (int m = 0; m < 5; ++m) {
                            const double *r = Rv + 5 * m;
                            F[m] = r[0] * fhc[0] + r[1] * fhc[1] + r[2] * fhc[2] + r[3] * fhc[3] +
                                   r[4] * fhc[4];
                        }
                    } else {
                        for (int m = 0; m < 5; ++m) F[m] = fhc[m];
                    }
                }
                for (int i = lo[d]; i <= hi[d]; ++i) {
                    NODE(i, ii, jj, kk)
                    const long p = iidx(&s, ii, jj, kk);
                    for (int m = 0; m < 5; ++m)
                        qrhs[m * s.nint + p] = qrhs[m * s.nint + p] + fh[5 * (i + 1) + m] -
                                               fh[5 * i + m];
                }
#undef NODE
            }
        free(fp);
        free(fm);
        free(line);
        free(hp);
        free(hmn);
        free(fh);
    }
    return 0;
}

/* rhscal for the compact upwind scheme (conschm(4:4)='c', odd order),
 * src/solver.F90:226-236: qrhs = 0, gradcal, ducrossensor when
 * lchardecomp, convrsdcmp, qrhs = -qrhs. */
int orc_cmp_rhscal(const orc_uwd *u, int explicit_fd, const double *q, const double *rho,
                   const double *vel, const double *prs, const double *tmp, const double *dxi,
                   const double *jacob, double *qrhs, double *dvel, double *dtmp, double *ssf,
                   unsigned char *lshock)
{
    const shape_t s = shape_of(u);
    memset(qrhs, 0, sizeof(double) * 5 * s.nint);
    orc_gradcal(u, explicit_fd, vel, tmp, dxi, dvel, dtmp);
    if (u->lchardecomp) orc_ducros(u, dvel, prs, ssf, lshock);
    const int rc = orc_convrsdcmp(u, q, rho, vel, prs, tmp, dxi, jacob,
                                  u->lchardecomp ? lshock : NULL, qrhs);
    if (rc) return rc;
    for (long e = 0; e < 5 * s.nint; ++e) qrhs[e] = -qrhs[e];
    return 0;
}

/* ------------------------------------ boundary conditions and RK3 steps */

/* q2fvar_3da at one node, src/fludyna.F90:545-634 (nondimensional, no
 * species): rho, u, v, w from q; p = (q5 - rho/2 |u|^2)/const6;
 * T = thermal(pressure, density) = p/rho*const2. */
static void q2fvar_node(const shape_t *s, long h, double const2, double const6, const double *q,
                        double *rho, double *vel, double *prs, double *tmp)
{
    const double r = q[h];
    const double u = q[s->nf + h] / r, v = q[2 * s->nf + h] / r, w = q[3 * s->nf + h] / r;
    const double p = (q[4 * s->nf + h] - 0.5 * r * (u * u + v * v + w * w)) / const6;
    rho[h] = r;
    vel[h] = u;
    vel[s->nf + h] = v;
    vel[2 * s->nf + h] = w;
    prs[h] = p;
    tmp[h] = p / r * const2;
}

/* The primitives of one boundary node as udf_bc sets them, then
 * tmp = thermal(pressure, density) and q = fvar2q_sca's pressure form,
 * q5 = p*const6 + rho*(0.5*sum(u*u)) (src/fludyna.F90:45-70, 310-362). */
static void bc_node(const shape_t *s, long h, double const2, double const6, double r, double u,
                    double v, double w, double p, double *q, double *rho, double *vel, double *prs,
                    double *tmp)
{
    rho[h] = r;
    vel[h] = u;
    vel[s->nf + h] = v;
    vel[2 * s->nf + h] = w;
    prs[h] = p;
    tmp[h] = p / r * const2;
    q[h] = r;
    q[s->nf + h] = r * u;
    q[2 * s->nf + h] = r * v;
    q[3 * s->nf + h] = r * w;
    const double var1 = 0.5 * (u * u + v * v + w * w);
    q[4 * s->nf + h] = p * const6 + r * var1;
}

/* shockspeed, examples/Riemann2D/user_define_module/userdefine_riemann2d.F90:
 * 670-686: the Rankine-Hugoniot speed of the shock that raises the pressure
 * p1 -> p2 in a gas at temperature t1, sqrt(1 + (p2/p1-1)(g+1)/(2g)) times
 * the sound speed sos(t1) = sqrt(t1)/Ma (src/fludyna.F90:832-859). */
double orc_riem2d_shockspeed(double gamma, double mach, double p1, double t1, double p2)
{
    const double mx2 = (p2 / p1 - 1.0) * (gamma + 1.0) / (2.0 * gamma) + 1.0;
    const double ss = sqrt(t1) / mach;
    return sqrt(mx2) * ss;
}

/* udf_bc of the 2-D Riemann example (userdefine_riemann2d.F90:421-668), as
 * boucon calls it (src/bc.F90:327-411: faces n = 1..6 in order, udf_bc
 * where bctype(n) = 0), on the grid of gridcube(1,1,0) (x = i/im,
 * y = j/jm; src/gridgeneration.F90:233-267, single rank):
 *   face 1 (i=0):  the quadrant states behind/ahead of the moving shock
 *                  y > 0.8 - vss*time, u = 1.206;
 *   face 2 (i=im): rho, p, v extrapolated, (4f(im-1) - f(im-2))/3, u = 0;
 *                  then the corner (im,jm) fixed at the NE state;
 *   face 3 (j=0):  the states by x > 0.8 - vss*time, v = 1.206;
 *   face 4 (j=jm): rho, p, u extrapolated along j, v = 0; corner again.
 * The NW pressure literal is single precision (0.30) as in the source.
 * Faces 5 and 6 have no user condition. */
void orc_riem2d_bc(const orc_uwd *u, const int *bctype, double time, double *q, double *rho,
                   double *vel, double *prs, double *tmp)
{
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double r01 = 0.5323, r11 = 1.5, u01 = 1.206, v01 = 0.0, v11 = 0.0;
    const double p01 = (double)0.30f, p11 = 1.5;
    const double r00 = 0.138, r10 = 0.5323, u00 = 1.206, v00 = 1.206, u10 = 0.0, v10 = 1.206;
    const double p00 = 0.029, p10 = 0.3;
    const double n1d3 = 1.0 / 3.0;
    const int im = u->im, jm = u->jm, km = u->km;
    const double dx = 1.0 / (double)im, dy = 1.0 / (double)jm;
#define H(i, j, k) hidx(&s, i, j, k)
#define V(c, h) vel[(c) * s.nf + (h)]
    if (bctype[0] == 0) {
        const double t1 = p01 / r01 * const2;
        const double vss = orc_riem2d_shockspeed(gamma, mach, p01, t1, p00);
        for (int k = 0; k <= km; ++k)
            for (int j = 0; j <= jm; ++j) {
                const double ys = 0.8 - vss * time;
                const double y = dy * (double)j;
                const int up = y > ys;
                bc_node(&s, H(0, j, k), const2, const6, up ? r01 : r00, u01, up ? v01 : v00, 0.0,
                        up ? p01 : p00, q, rho, vel, prs, tmp);
            }
    }
    if (bctype[1] == 0) {
        for (int k = 0; k <= km; ++k)
            for (int j = 0; j <= jm; ++j) {
                const long a = H(im - 1, j, k), b = H(im - 2, j, k);
                const double r = n1d3 * (4.0 * rho[a] - rho[b]);
                const double p = n1d3 * (4.0 * prs[a] - prs[b]);
                const double v = n1d3 * (4.0 * V(1, a) - V(1, b));
                bc_node(&s, H(im, j, k), const2, const6, r, 0.0, v, 0.0, p, q, rho, vel, prs,
                        tmp);
            }
        for (int k = 0; k <= km; ++k)
            bc_node(&s, H(im, jm, k), const2, const6, r11, 0.0, v11, 0.0, p11, q, rho, vel, prs,
                    tmp);
    }
    if (bctype[2] == 0) {
        const double t1 = p10 / r10 * const2;
        const double vss = orc_riem2d_shockspeed(gamma, mach, p10, t1, p00);
        for (int k = 0; k <= km; ++k)
            for (int i = 0; i <= im; ++i) {
                const double xs = 0.8 - vss * time;
                const double x = dx * (double)i;
                const int up = x > xs;
                bc_node(&s, H(i, 0, k), const2, const6, up ? r10 : r00, up ? u10 : u00, v10, 0.0,
                        up ? p10 : p00, q, rho, vel, prs, tmp);
            }
    }
    if (bctype[3] == 0) {
        for (int k = 0; k <= km; ++k)
            for (int i = 0; i <= im; ++i) {
                const long a = H(i, jm - 1, k), b = H(i, jm - 2, k);
                const double r = n1d3 * (4.0 * rho[a] - rho[b]);
                const double p = n1d3 * (4.0 * prs[a] - prs[b]);
                const double uu = n1d3 * (4.0 * V(0, a) - V(0, b));
                bc_node(&s, H(i, jm, k), const2, const6, r, uu, 0.0, 0.0, p, q, rho, vel, prs,
                        tmp);
            }
        for (int k = 0; k <= km; ++k)
            bc_node(&s, H(im, jm, k), const2, const6, r11, 0.0, 0.0, 0.0, p11, q, rho, vel, prs,
                    tmp);
    }
#undef V
#undef H
}

/* gcnscbc, src/bc.F90:7907-8137 (Motheau, Almgren & Bell's ghost-cell
 * NSCBC), as boucon calls it for bctype 23 with prs_t = pinf
 * (src/bc.F90:389-391), nondimensional, no species.  Face 2 (i = im) and
 * face 4 (j = jm) only; the routine has no branch for the other faces.  At
 * each node of the face, from the two nodes inside it (one-sided first
 * differences, deriv_1o, src/commfunc.F90:213-222, over the gridcube
 * spacing x(n-1) - x(n-2)):
 *   LODI amplitudes l_k = lambda_k * (characteristic derivative), with the
 *   normal velocity vnorm, the sound speed css = sos(T) = sqrt(T)/Ma and
 *   css2 = css*css held in INTEGER variables as the reference declares
 *   them (truncated toward zero), mavg = vnorm/css an integer quotient,
 *   lambda = (vnorm-css, vnorm, vnorm, vnorm, vnorm) + 1e-10 (the
 *   reference's lambda(5:numq) = vnorm overwrites vnorm+css when numq = 5);
 *   the incoming wave l_1 replaced by the soft outflow relaxation
 *   0.25 css (1 - mavg^2)/(xmax-xmin) (p - pinf), or, on face 4 where
 *   vnorm < 0, by the inflow relaxations of v, T, u, w towards vinf, tinf,
 *   uinf, winf (eta = 0.25);
 *   l / lambda projected back to d(rho, T, p, u) (Poinsot & Lele 1992,
 *   Eqs. 33-36), the boundary node extrapolated f(n) = f(n-1) + 2 dx df,
 *   rho = thermal(p, T) = p/T*const2, q = fvar2q(rho, u, p).
 * Nondimensional freestream: uinf = 1, vinf = winf = 0, tinf = 1,
 * pinf = 1/const2 (refcal, src/solver.F90:100-120).  The reference's
 * qrhs = 0 on the face is followed by rhscal's own sums on every step, so
 * it is not restated. */
void orc_gcnscbc(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                 double *tmp)
{
    if (face != 2 && face != 4) return;
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double pinf = 1.0 / const2, uinf = 1.0, vinf = 0.0, winf = 0.0, tinf = 1.0;
    const double eps = 1.e-10;
    const int im = u->im, jm = u->jm, km = u->km;
    const int ax = face == 2 ? 0 : 1;       /* normal axis */
    const int n = ax == 0 ? im : jm;
    const double len = ax == 0 ? u->lx : u->ly;
    const double hx = len / (double)n;      /* gridcube: x(i) = len/n*i */
    const double dx = hx * (double)(n - 1) - hx * (double)(n - 2);
    const double span = hx * (double)n - 0.0;  /* xmax - xmin */
    const int na = ax == 0 ? jm : im;
    for (int k = 0; k <= km; ++k)
        for (int a = 0; a <= na; ++a) {
            const int i = ax == 0 ? im : a, j = ax == 0 ? a : jm;
            const long h = hidx(&s, i, j, k);
            const long h1 = ax == 0 ? hidx(&s, i - 1, j, k) : hidx(&s, i, j - 1, k);
            const long h2 = ax == 0 ? hidx(&s, i - 2, j, k) : hidx(&s, i, j - 2, k);
            const double t1 = tmp[h1], r1 = rho[h1], p1 = prs[h1];
            const double v1[3] = {vel[h1], vel[s.nf + h1], vel[2 * s.nf + h1]};
            const double v2[3] = {vel[h2], vel[s.nf + h2], vel[2 * s.nf + h2]};
            const int css = (int)(sqrt(t1) / mach);
            if (css == 0) {
                /* the reference divides by the zero integer sound speed here
                 * (a trap in Fortran); the node is set to NaN instead */
                const double nan = NAN;
                rho[h] = prs[h] = tmp[h] = nan;
                for (int c = 0; c < 3; ++c) vel[c * s.nf + h] = nan;
                for (int m = 0; m < 5; ++m) q[m * s.nf + h] = nan;
                continue;
            }
            const int css2 = css * css;
            const int vnorm = (int)v1[ax];
            const double mavg = (double)(vnorm / css);
            double lambda[5];
            lambda[0] = (double)(vnorm - css);
            lambda[1] = lambda[2] = lambda[3] = (double)vnorm;
            lambda[4] = (double)vnorm;
            for (int m = 0; m < 5; ++m) lambda[m] = lambda[m] + eps;
            double dvel[3];
            for (int c = 0; c < 3; ++c) dvel[c] = -(-v1[c] + v2[c]) / dx;
            double dprs = -(-p1 + prs[h2]) / dx;
            double drho = -(-r1 + rho[h2]) / dx;
            double dtmp = -(-t1 + tmp[h2]) / dx;
            (void)dtmp;
            const int tg = ax == 0 ? 1 : 0;  /* the tangential components */
            double lodi[5];
            lodi[0] = lambda[0] * (dprs - r1 * (double)css * dvel[ax]);
            lodi[1] = lambda[1] * ((double)css2 * drho - dprs);
            lodi[2] = lambda[2] * dvel[tg];
            lodi[3] = lambda[3] * dvel[2];
            lodi[4] = lambda[4] * (dprs + r1 * (double)css * dvel[ax]);
            if (face == 2 || vnorm >= 0) {
                const double sigma = 0.25;
                const double kamma = sigma * (double)css * (1.0 - mavg * mavg) / span;
                lodi[0] = kamma * (p1 - pinf);
            } else {
                const double eta = 0.25;
                lodi[0] = eta * r1 * (double)css2 * (1.0 - mavg * mavg) / span * (v1[1] - vinf);
                lodi[1] = eta * r1 * (double)css2 / gamma / span * (t1 - tinf);
                lodi[2] = eta * (double)css / span * (v1[0] - uinf);
                lodi[3] = eta * (double)css / span * (v1[2] - winf);
            }
            for (int m = 0; m < 5; ++m) lodi[m] = lodi[m] / lambda[m];
            drho = 1.0 / (double)css2 * (lodi[1] + 0.5 * (lodi[0] + lodi[4]));
            dtmp = t1 / r1 / (double)css * (-lodi[1] + 0.5 * (gamma - 1.0) * (lodi[0] + lodi[4]));
            dprs = 0.5 * (lodi[0] + lodi[4]);
            dvel[ax] = 1.0 / (2.0 * r1 * (double)css) * (lodi[4] - lodi[0]);
            dvel[tg] = lodi[2];
            dvel[2] = lodi[3];
            const double tn = t1 + 2.0 * dx * dtmp;
            const double pn = p1 + 2.0 * dx * dprs;
            double vn[3];
            for (int c = 0; c < 3; ++c) vn[c] = v1[c] + 2.0 * dx * dvel[c];
            const double rn = pn / tn * const2;
            rho[h] = rn;
            vel[h] = vn[0];
            vel[s.nf + h] = vn[1];
            vel[2 * s.nf + h] = vn[2];
            prs[h] = pn;
            tmp[h] = tn;
            q[h] = rn;
            q[s.nf + h] = rn * vn[0];
            q[2 * s.nf + h] = rn * vn[1];
            q[3 * s.nf + h] = rn * vn[2];
            const double var1 = 0.5 * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
            q[4 * s.nf + h] = pn * const6 + rn * var1;
        }
}

/* zeroextrap (src/bc.F90:2331-2627), bctype 50, on face 1..6 (i = 0, i = im,
 * j = 0, j = jm, k = 0, k = km) over 0..n of the other two axes: the
 * boundary node's velocity and pressure from the two inner nodes by
 * extrapolate_2o with dv = 0 (src/commfunc.F90:277-285: v0 = (4 v1 - v2 -
 * 2 dv)/3, the value at the face of the parabola with zero slope there), and
 * on faces 1-2 the temperature too, the density from thermal(p, T) =
 * p/T*const2; on faces 3-6 the density, the temperature from thermal(p, rho)
 * = p/rho*const2; q = fvar2q.  The reference's qrhs = 0 on the face is
 * followed by rhscal's own sums, so it is not restated. */
void orc_zeroextrap(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                    double *tmp)
{
    if (face < 1 || face > 6) return;
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double n1d3 = 1.0 / 3.0;
    const int ax = (face - 1) / 2, hi = (face - 1) % 2;
    const int n[3] = {u->im, u->jm, u->km};
    const int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
    for (int b = 0; b <= n[ob]; ++b)
        for (int a = 0; a <= n[oa]; ++a) {
            int c[3], c1[3], c2[3];
            c[oa] = c1[oa] = c2[oa] = a;
            c[ob] = c1[ob] = c2[ob] = b;
            c[ax] = hi ? n[ax] : 0;
            c1[ax] = hi ? n[ax] - 1 : 1;
            c2[ax] = hi ? n[ax] - 2 : 2;
            const long h = hidx(&s, c[0], c[1], c[2]);
            const long h1 = hidx(&s, c1[0], c1[1], c1[2]), h2 = hidx(&s, c2[0], c2[1], c2[2]);
            double ve[3];
            for (int m = 0; m < 3; ++m)
                ve[m] = n1d3 * (4.0 * vel[m * s.nf + h1] - vel[m * s.nf + h2] - 2.0 * 0.0);
            const double pe = n1d3 * (4.0 * prs[h1] - prs[h2] - 2.0 * 0.0);
            double re, te;
            if (ax == 0) {
                te = n1d3 * (4.0 * tmp[h1] - tmp[h2] - 2.0 * 0.0);
                re = pe / te * const2;
            } else {
                re = n1d3 * (4.0 * rho[h1] - rho[h2] - 2.0 * 0.0);
                te = pe / re * const2;
            }
            prs[h] = pe;
            tmp[h] = te;
            rho[h] = re;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = ve[m];
            q[h] = re;
            q[s.nf + h] = re * ve[0];
            q[2 * s.nf + h] = re * ve[1];
            q[3 * s.nf + h] = re * ve[2];
            const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
            q[4 * s.nf + h] = pe * const6 + re * var1;
        }
}

/* symmetry (src/bc.F90:2639-2971), bctype 60, on face 1..6 of a rectilinear
 * grid: the boundary node from its two inner neighbours by extrapolate_2o
 * with dv = 0 like zeroextrap, its velocity normal to the face zero (faces
 * 1, 2 and 6 project the extrapolated velocity on the face through the
 * boundary normal bnorm = dxi(d,:)/|dxi(d,:)| of src/geom.F90:700-790,
 * which is +-e_d on such a grid, so the normal component is exactly zero
 * and the others are kept; faces 3, 4 and 5 set it to 0); the temperature
 * extrapolated and the density from thermal(p, T) = p/T*const2, except on
 * face 6, which extrapolates the density and takes T = p/rho*const2; q by
 * fvar2q's temperature form q5 = rho*(T*const1 + |u|^2/2) on faces 1, 2, 4
 * and 6 and its pressure form q5 = p*const6 + rho*|u|^2/2 on faces 3 and 5,
 * as the reference calls it (src/fludyna.F90:312-376). */
void orc_symmetry(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp)
{
    if (face < 1 || face > 6) return;
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach));
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double n1d3 = 1.0 / 3.0;
    const int ax = (face - 1) / 2, hi = (face - 1) % 2;
    const int rho_ext = face == 6, tform = face != 3 && face != 5;
    const int n[3] = {u->im, u->jm, u->km};
    const int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
    for (int b = 0; b <= n[ob]; ++b)
        for (int a = 0; a <= n[oa]; ++a) {
            int c[3], c1[3], c2[3];
            c[oa] = c1[oa] = c2[oa] = a;
            c[ob] = c1[ob] = c2[ob] = b;
            c[ax] = hi ? n[ax] : 0;
            c1[ax] = hi ? n[ax] - 1 : 1;
            c2[ax] = hi ? n[ax] - 2 : 2;
            const long h = hidx(&s, c[0], c[1], c[2]);
            const long h1 = hidx(&s, c1[0], c1[1], c1[2]), h2 = hidx(&s, c2[0], c2[1], c2[2]);
            double ve[3];
            for (int m = 0; m < 3; ++m)
                ve[m] = n1d3 * (4.0 * vel[m * s.nf + h1] - vel[m * s.nf + h2] - 2.0 * 0.0);
            ve[ax] = 0.0;
            const double pe = n1d3 * (4.0 * prs[h1] - prs[h2] - 2.0 * 0.0);
            double re, te;
            if (rho_ext) {
                re = n1d3 * (4.0 * rho[h1] - rho[h2] - 2.0 * 0.0);
                te = pe / re * const2;
            } else {
                te = n1d3 * (4.0 * tmp[h1] - tmp[h2] - 2.0 * 0.0);
                re = pe / te * const2;
            }
            prs[h] = pe;
            tmp[h] = te;
            rho[h] = re;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = ve[m];
            q[h] = re;
            q[s.nf + h] = re * ve[0];
            q[2 * s.nf + h] = re * ve[1];
            q[3 * s.nf + h] = re * ve[2];
            const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
            q[4 * s.nf + h] = tform ? re * (te * const1 + var1) : pe * const6 + re * var1;
        }
}

/* outflow (src/bc.F90:3379-3592), bctype 21, on face 2 or 4 (the faces it
 * acts on).  Face 2 copies the velocity, pressure and temperature of node
 * n-1 (the reference's zero-order branch, taken always), the density from
 * thermal(p, T) = p/T*const2 and q by fvar2q's temperature form.  Face 4
 * extrapolates velocity, pressure, temperature and density by
 * extrapolate_2o with dv = 0; where the node's own normal velocity v is at
 * least its sound speed css = sos(T) = sqrt(T)/Ma it takes them as they
 * are, else its pressure relaxes towards pinf, p = (p + alpha dt pinf +
 * rho css (ve - v))/(1 + alpha dt) with alpha = 0.25, the node's own p, rho
 * and v and the step's dt, and its density comes from (p, Te); on either
 * branch T = thermal(p, rho) = p/rho*const2 afterwards and q by the pressure
 * form.  pinf = 1/const2 (the nondimensional freestream). */
void orc_outflow(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                 double *tmp)
{
    if (face != 2 && face != 4) return;
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach));
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double pinf = 1.0 / const2, alpha = 0.25, dt = u->deltat;
    const double n1d3 = 1.0 / 3.0;
    const int ax = face == 2 ? 0 : 1;
    const int n[3] = {u->im, u->jm, u->km};
    const int oa = ax == 0 ? 1 : 0, ob = 2;
    for (int b = 0; b <= n[ob]; ++b)
        for (int a = 0; a <= n[oa]; ++a) {
            int c[3], c1[3], c2[3];
            c[oa] = c1[oa] = c2[oa] = a;
            c[ob] = c1[ob] = c2[ob] = b;
            c[ax] = n[ax];
            c1[ax] = n[ax] - 1;
            c2[ax] = n[ax] - 2;
            const long h = hidx(&s, c[0], c[1], c[2]);
            const long h1 = hidx(&s, c1[0], c1[1], c1[2]), h2 = hidx(&s, c2[0], c2[1], c2[2]);
            double ve[3], pn, rn, tn;
            if (face == 2) {
                for (int m = 0; m < 3; ++m) ve[m] = vel[m * s.nf + h1];
                pn = prs[h1];
                tn = tmp[h1];
                rn = pn / tn * const2;
            } else {
                const double css = sqrt(tmp[h]) / mach;
                const double ub = vel[s.nf + h];
                for (int m = 0; m < 3; ++m)
                    ve[m] = n1d3 * (4.0 * vel[m * s.nf + h1] - vel[m * s.nf + h2] - 2.0 * 0.0);
                const double pe = n1d3 * (4.0 * prs[h1] - prs[h2] - 2.0 * 0.0);
                const double te = n1d3 * (4.0 * tmp[h1] - tmp[h2] - 2.0 * 0.0);
                const double roe = n1d3 * (4.0 * rho[h1] - rho[h2] - 2.0 * 0.0);
                if (ub >= css) {
                    pn = pe;
                    rn = roe;
                } else {
                    pn = (prs[h] + alpha * dt * pinf + rho[h] * css * (ve[1] - vel[s.nf + h])) /
                         (1.0 + alpha * dt);
                    rn = pn / te * const2;
                }
                tn = pn / rn * const2;
            }
            prs[h] = pn;
            tmp[h] = tn;
            rho[h] = rn;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = ve[m];
            q[h] = rn;
            q[s.nf + h] = rn * ve[0];
            q[2 * s.nf + h] = rn * ve[1];
            q[3 * s.nf + h] = rn * ve[2];
            const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
            q[4 * s.nf + h] = face == 2 ? rn * (tn * const1 + var1) : pn * const6 + rn * var1;
        }
}

/* slipadibwall (src/bc.F90:7114-7313), bctype 421, on face 3 or 4 (the
 * faces it acts on): pressure, temperature and u by num1d3*(4 f1 - f2), w
 * zero, v zero on face 3 and extrapolated alike on face 4; the density from
 * thermal(p, T) = p/T*const2 and q by fvar2q's pressure form. */
void orc_slipadib(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp)
{
    if (face != 3 && face != 4) return;
    const shape_t s = shape_of(u);
    const double const2 = u->gamma * (u->mach * u->mach), const6 = 1.0 / (u->gamma - 1.0);
    const double n1d3 = 1.0 / 3.0;
    const int hi = face == 4;
    for (int k = 0; k <= u->km; ++k)
        for (int i = 0; i <= u->im; ++i) {
            const long h = hidx(&s, i, hi ? u->jm : 0, k);
            const long h1 = hidx(&s, i, hi ? u->jm - 1 : 1, k), h2 = hidx(&s, i, hi ? u->jm - 2 : 2, k);
            const double pe = n1d3 * (4.0 * prs[h1] - prs[h2]);
            const double te = n1d3 * (4.0 * tmp[h1] - tmp[h2]);
            double ve[3];
            ve[0] = n1d3 * (4.0 * vel[h1] - vel[h2]);
            ve[1] = hi ? n1d3 * (4.0 * vel[s.nf + h1] - vel[s.nf + h2]) : 0.0;
            ve[2] = 0.0;
            const double re = pe / te * const2;
            prs[h] = pe;
            tmp[h] = te;
            rho[h] = re;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = ve[m];
            q[h] = re;
            q[s.nf + h] = re * ve[0];
            q[2 * s.nf + h] = re * ve[1];
            q[3 * s.nf + h] = re * ve[2];
            const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
            q[4 * s.nf + h] = pe * const6 + re * var1;
        }
}

/* slipisotwall (src/bc.F90:6921-7112), bctype 411, on face 3 or 4: on face
 * 3 the pressure by num1d3*(4 p1 - p2) and, where the node's gridcube
 * x = lx/im*i is at most xslip, u and T extrapolated alike with v = 0, else
 * the node at rest (the wall blowing of datin/wallbs.dat, unread, zero) at
 * the face's wall temperature; w = 0.  Face 4 is the isothermal noslip wall.
 * The density from p/T*const2, q by fvar2q's pressure form. */
void orc_slipisot(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp)
{
    if (face == 4) {
        orc_wall(u, 4, 1, q, rho, vel, prs, tmp);
        return;
    }
    if (face != 3) return;
    const shape_t s = shape_of(u);
    const double const2 = u->gamma * (u->mach * u->mach), const6 = 1.0 / (u->gamma - 1.0);
    const double n1d3 = 1.0 / 3.0;
    for (int k = 0; k <= u->km; ++k)
        for (int i = 0; i <= u->im; ++i) {
            const long h = hidx(&s, i, 0, k), h1 = hidx(&s, i, 1, k), h2 = hidx(&s, i, 2, k);
            const double pe = n1d3 * (4.0 * prs[h1] - prs[h2]);
            const double x = u->lx / (double)u->im * (double)i;
            double ue, te;
            if (x <= u->xslip) {
                ue = n1d3 * (4.0 * vel[h1] - vel[h2]);
                te = n1d3 * (4.0 * tmp[h1] - tmp[h2]);
            } else {
                ue = 0.0;
                te = u->twall[2];
            }
            const double re = pe / te * const2;
            prs[h] = pe;
            tmp[h] = te;
            rho[h] = re;
            vel[h] = ue;
            vel[s.nf + h] = 0.0;
            vel[2 * s.nf + h] = 0.0;
            q[h] = re;
            q[s.nf + h] = re * ue;
            q[2 * s.nf + h] = re * 0.0;
            q[3 * s.nf + h] = re * 0.0;
            q[4 * s.nf + h] = pe * const6 + re * (0.5 * (ue * ue + 0.0 * 0.0 + 0.0 * 0.0));
        }
}

/* fixbc (src/bc.F90:7562-7648), bctype 31: face 3 held at rho 2, p 1 and
 * face 4 at rho 1, p 2.5, both at rest, T = thermal(rho, p) = p/rho*const2,
 * q by fvar2q's pressure form; the other faces are left alone. */
void orc_fixbc(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
               double *tmp)
{
    if (face != 3 && face != 4) return;
    const shape_t s = shape_of(u);
    const double const2 = u->gamma * (u->mach * u->mach), const6 = 1.0 / (u->gamma - 1.0);
    const int j = face == 3 ? 0 : u->jm;
    const double rf = face == 3 ? 2.0 : 1.0, pf = face == 3 ? 1.0 : 2.5;
    for (int k = 0; k <= u->km; ++k)
        for (int i = 0; i <= u->im; ++i) {
            const long h = hidx(&s, i, j, k);
            rho[h] = rf;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = 0.0;
            prs[h] = pf;
            tmp[h] = pf / rf * const2;
            q[h] = rf;
            q[s.nf + h] = rf * 0.0;
            q[2 * s.nf + h] = rf * 0.0;
            q[3 * s.nf + h] = rf * 0.0;
            const double var1 = 0.5 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
            q[4 * s.nf + h] = pf * const6 + rf * var1;
        }
}

/* noslip (src/bc.F90:6281-6595), bctype 41, and noslip_adibatic
 * (src/bc.F90:6611-6919), bctype 42, on face 1..4 (the reference stops on
 * faces 5 and 6): the boundary node at rest, its pressure num1d3*(4 p1 -
 * p2) from the two inner neighbours, its temperature the face's wall
 * temperature (41) or extrapolated alike (42), the density from
 * thermal(p, T) = p/T*const2 and q by fvar2q's pressure form.  The wall
 * blowing and suction of face 3 (datin/wallbs.dat) is not read: without
 * that file the reference's wall velocity is zero too. */
void orc_wall(const orc_uwd *u, int face, int isothermal, double *q, double *rho, double *vel,
              double *prs, double *tmp)
{
    if (face < 1 || face > 4) return;
    const shape_t s = shape_of(u);
    const double gamma = u->gamma, mach = u->mach;
    const double const2 = gamma * (mach * mach), const6 = 1.0 / (gamma - 1.0);
    const double n1d3 = 1.0 / 3.0;
    const int ax = (face - 1) / 2, hi = (face - 1) % 2;
    const int n[3] = {u->im, u->jm, u->km};
    const int oa = ax == 0 ? 1 : 0, ob = 2;
    for (int b = 0; b <= n[ob]; ++b)
        for (int a = 0; a <= n[oa]; ++a) {
            int c[3], c1[3], c2[3];
            c[oa] = c1[oa] = c2[oa] = a;
            c[ob] = c1[ob] = c2[ob] = b;
            c[ax] = hi ? n[ax] : 0;
            c1[ax] = hi ? n[ax] - 1 : 1;
            c2[ax] = hi ? n[ax] - 2 : 2;
            const long h = hidx(&s, c[0], c[1], c[2]);
            const long h1 = hidx(&s, c1[0], c1[1], c1[2]), h2 = hidx(&s, c2[0], c2[1], c2[2]);
            const double pe = n1d3 * (4.0 * prs[h1] - prs[h2]);
            const double te = isothermal ? u->twall[face - 1] : n1d3 * (4.0 * tmp[h1] - tmp[h2]);
            const double re = pe / te * const2;
            prs[h] = pe;
            tmp[h] = te;
            rho[h] = re;
            for (int m = 0; m < 3; ++m) vel[m * s.nf + h] = 0.0;
            q[h] = re;
            q[s.nf + h] = re * 0.0;
            q[2 * s.nf + h] = re * 0.0;
            q[3 * s.nf + h] = re * 0.0;
            const double var1 = 0.5 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
            q[4 * s.nf + h] = pe * const6 + re * var1;
        }
}

void orc_boucon(const orc_uwd *u, const int *bctype, int udf, double time, double *q,
                double *rho, double *vel, double *prs, double *tmp)
{
    for (int f = 0; f < 6; ++f) {
        if (bctype[f] == 0 && udf == 1) {
            int one[6] = {1, 1, 1, 1, 1, 1};
            one[f] = 0;
            orc_riem2d_bc(u, one, time, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 23) {
            orc_gcnscbc(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 50) {
            orc_zeroextrap(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 60) {
            orc_symmetry(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 41 || bctype[f] == 42) {
            orc_wall(u, f + 1, bctype[f] == 41, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 21) {
            orc_outflow(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 31) {
            orc_fixbc(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 421) {
            orc_slipadib(u, f + 1, q, rho, vel, prs, tmp);
        } else if (bctype[f] == 411) {
            orc_slipisot(u, f + 1, q, rho, vel, prs, tmp);
        }
    }
}

/* qswap, single rank (src/parallel.F90:4849-5319, turbmode none): along
 * each periodic axis in the order i, j, k the halo copy, the duplicate
 * average q(0) = q(n) = (q(0)+q(n))/2 and q2fvar on the strips -hm..0 and
 * n..n+hm, over 0..n of the other axes; a two-dimensional box (km = 0)
 * copies the plane k = 0 into every k plane -hm..hm and refreshes the
 * primitives of all of them (the ka = 0 branch).  Non-periodic axes are
 * left alone (no neighbour on one rank). */
void orc_uwd_qswap(const orc_uwd *u, double *q, double *rho, double *vel, double *prs,
                   double *tmp)
{
    const shape_t s = shape_of(u);
    const double const2 = u->gamma * (u->mach * u->mach), const6 = 1.0 / (u->gamma - 1.0);
    const int n[3] = {u->im, u->jm, u->km};
    for (int d = 0; d < 3; ++d) {
        const int two_d = d == 2 && n[2] == 0;
        if (!two_d && u->npdc[d] != 3) continue;
        const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
        for (int b = 0; b <= n[ob]; ++b)
            for (int a = 0; a <= n[oa]; ++a) {
                int c[3];
                c[oa] = a;
                c[ob] = b;
                c[d] = 0;
                const long h0 = hidx(&s, c[0], c[1], c[2]);
                const long sd = d == 0 ? 1 : d == 1 ? s.ni : s.ni * s.nj;
#define AT(t) (h0 + (long)(t) * sd)
                for (int m = 0; m < 5; ++m) {
                    double *f = q + m * s.nf;
                    if (two_d) {
                        const double v0 = f[AT(0)];
                        for (int t = -HMU; t <= HMU; ++t)
                            if (t) f[AT(t)] = v0;
                        continue;
                    }
                    for (int t = 1; t <= HMU; ++t) {
                        const double lo = f[AT(n[d] - t)];
                        f[AT(-t)] = lo;
                    }
                    for (int t = 1; t <= HMU; ++t) {
                        const double hi = f[AT(t)];
                        f[AT(n[d] + t)] = hi;
                    }
                    const double avg = 0.5 * (f[AT(0)] + f[AT(n[d])]);
                    f[AT(0)] = avg;
                    f[AT(n[d])] = avg;
                }
                for (int t = n[d]; t <= n[d] + HMU; ++t)
                    q2fvar_node(&s, AT(t), const2, const6, q, rho, vel, prs, tmp);
                for (int t = -HMU; t <= 0; ++t)
                    q2fvar_node(&s, AT(t), const2, const6, q, rho, vel, prs, tmp);
#undef AT
            }
    }
}

/* One RK3 step of time_integration_rk (src/mainloop.F90:392-476, rk3, no
 * filter) on the upwind path: per stage qrhs = 0, boucon (the 2-D Riemann
 * udf_bc on faces with bctype 0 when udf = 1) at the step's start time,
 * qswap, rhscal (convrsduwd, or convrsdcmp with compact; explicit gradient
 * with explicit_fd), qsave = q*J at stage 1, the combine
 * q = (a qsave + b q J + c qrhs dt)/J on 0..n with (a,b,c) = (1,0,1),
 * (3/4,1/4,1/4), (1/3,2/3,2/3), and updatefvar on 0..n.  Returns nonzero
 * on an unsupported configuration. */
int orc_uwd_rk3_step(const orc_uwd *u, int compact, int explicit_fd, const int *bctype, int udf,
                     double time, double dt, const double *dxi, const double *jacob, double *q,
                     double *rho, double *vel, double *prs, double *tmp)
{
    const shape_t s = shape_of(u);
    const double const2 = u->gamma * (u->mach * u->mach), const6 = 1.0 / (u->gamma - 1.0);
    const double rk[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25},
                             {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    int anyudf = 0, anybc = 0;
    for (int f = 0; f < 6; ++f) {
        if (bctype[f] != 0 && bctype[f] != 1 && bctype[f] != 23 && bctype[f] != 50 &&
            bctype[f] != 60 && bctype[f] != 41 && bctype[f] != 42 && bctype[f] != 21 &&
            bctype[f] != 31 && bctype[f] != 421 && bctype[f] != 411)
            return 3;
        anyudf |= bctype[f] == 0;
        anybc |= bctype[f] != 1;
    }
    if (anyudf && udf != 1) return 3;
    double *qrhs = malloc(sizeof(double) * 5 * s.nint), *qsave = malloc(sizeof(double) * 5 * s.nint);
    double *dvel = malloc(sizeof(double) * 9 * s.nint), *dtmp = malloc(sizeof(double) * 3 * s.nint);
    double *ssf = malloc(sizeof(double) * s.nf);
    unsigned char *lsh = malloc(s.nint);
    int rc = 0;
    for (int st = 0; st < 3 && !rc; ++st) {
        if (anybc) {
            orc_uwd ud = *u;  /* boucon's outflow reads the step's deltat */
            ud.deltat = dt;
            orc_boucon(&ud, bctype, udf, time, q, rho, vel, prs, tmp);
        }
        orc_uwd_qswap(u, q, rho, vel, prs, tmp);
        rc = (compact ? orc_cmp_rhscal : orc_uwd_rhscal)(u, explicit_fd, q, rho, vel, prs, tmp,
                                                         dxi, jacob, qrhs, dvel, dtmp, ssf, lsh);
        if (rc) break;
        for (int k = 0; k <= u->km; ++k)
            for (int j = 0; j <= u->jm; ++j)
                for (int i = 0; i <= u->im; ++i) {
                    const long h = hidx(&s, i, j, k), p = iidx(&s, i, j, k);
                    const double jc = jacob[h];
                    for (int m = 0; m < 5; ++m) {
                        const double qo = q[m * s.nf + h];
                        if (st == 0) qsave[m * s.nint + p] = qo * jc;
                        const double v = rk[st][0] * qsave[m * s.nint + p] + rk[st][1] * qo * jc +
                                         rk[st][2] * qrhs[m * s.nint + p] * dt;
                        q[m * s.nf + h] = v / jc;
                    }
                    q2fvar_node(&s, h, const2, const6, q, rho, vel, prs, tmp);
                }
    }
    free(qrhs);
    free(qsave);
    free(dvel);
    free(dtmp);
    free(ssf);
    free(lsh);
    return rc;
}
