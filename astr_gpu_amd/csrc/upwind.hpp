// Explicit upwind convection path on the device: Steger-Warming flux
// splitting, Roe characteristic projection, the upwind reconstructions of
// recons_exp, the interface-flux difference of convrsduwd, the Ducros shock
// sensor and the explicit 6th-order gradient.
//
// FMA contraction is off in this file: every expression is evaluated in the
// order the reference writes it (left to right, parentheses as written), as
// the checker does, so the limiter and sensor branches take the same side
// and the results agree bit for bit.
#pragma once

#pragma clang fp contract(off)

namespace astr {
namespace uwd {

// ------------------------------------------------------------- scalars

__device__ __forceinline__ double min2(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double max2(double a, double b) { return a > b ? a : b; }

// minmod2 / minmod4, src/commfunc.F90:699-738
__device__ __forceinline__ double minmod2(double a, double b)
{
    if (a > 0.0 && b > 0.0) return min2(fabs(a), fabs(b));
    if (a < 0.0 && b < 0.0) return -1.0 * min2(fabs(a), fabs(b));
    return 0.0;
}

__device__ __forceinline__ double minmod4(double a, double b, double c, double d)
{
    if (a > 0.0 && b > 0.0 && c > 0.0 && d > 0.0)
        return min2(min2(min2(fabs(a), fabs(b)), fabs(c)), fabs(d));
    if (a < 0.0 && b < 0.0 && c < 0.0 && d < 0.0)
        return -1.0 * min2(min2(min2(fabs(a), fabs(b)), fabs(c)), fabs(d));
    return 0.0;
}

// suw3/5/7, src/flux.F90:388-422
__device__ __forceinline__ double suw3(const double* u)
{
    return -(1.0 / 6.0) * u[0] + (5.0 / 6.0) * u[1] + (1.0 / 3.0) * u[2];
}
__device__ __forceinline__ double suw5(const double* u)
{
    return (2.0 * u[0] - 13.0 * u[1] + 47.0 * u[2] + 27.0 * u[3] - 3.0 * u[4]) / 60.0;
}
__device__ __forceinline__ double suw7(const double* u)
{
    return (-3.0 * u[0] + 25.0 * u[1] - 101.0 * u[2] + 319.0 * u[3] + 214.0 * u[4] -
            38.0 * u[5] + 4.0 * u[6]) / 420.0;
}

// Suresh-Huynh monotonicity-preserving bounds on u(i-2..i+2) = u[0..4]
// around the linear value, src/flux.F90:465-490
__device__ double mp_limit(const double* u, double ul)
{
    const double dm1 = u[0] - 2.0 * u[1] + u[2];
    const double d0 = u[1] - 2.0 * u[2] + u[3];
    const double d1 = u[2] - 2.0 * u[3] + u[4];
    const double dhm1 = minmod4(4.0 * dm1 - d0, 4.0 * d0 - dm1, dm1, d0);
    const double dh0 = minmod4(4.0 * d0 - d1, 4.0 * d1 - d0, d0, d1);
    const double uUL = u[2] + 4.0 * (u[2] - u[1]);
    const double uAV = 0.5 * (u[2] + u[3]);
    const double uMD = uAV - 0.5 * dh0;
    const double uLC = u[2] + 0.5 * (u[2] - u[1]) + 1.333333333333333 * dhm1;
    const double lo = max2(min2(min2(u[2], u[3]), uMD), min2(min2(u[2], uUL), uLC));
    const double hi = min2(max2(max2(u[2], u[3]), uMD), max2(max2(u[2], uUL), uLC));
    return ul + minmod2(lo - ul, hi - ul);
}

// MP5 without the optional arguments, src/flux.F90:434-496
__device__ double mp5(const double* u)
{
    const double ul = suw5(u);
    const double uMP = u[2] + minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
    if ((ul - u[2]) * (ul - uMP) >= 1.e-10) return mp_limit(u, ul);
    return ul;
}

// MP5 with a supplied linear value and shock switch, and mplimiter,
// src/flux.F90:363-379,434-496
__device__ double mplimiter(const double* u, double ul, bool lskt, int inode, int dim, int ntype)
{
    if ((ntype == 1 && (inode == 0 || inode == 1)) ||
        (ntype == 2 && (inode == dim - 1 || inode == dim - 2)))
        return ul;
    const double uMP = u[2] + minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
    if (lskt && (ul - u[2]) * (ul - uMP) >= 1.e-10) return mp_limit(u, ul);
    return ul;
}

// MP7, src/flux.F90:498-551
__device__ double mp7(const double* u)
{
    const double ul = suw7(u);
    const double uMP = u[3] + minmod2(u[4] - u[3], 4.0 * (u[3] - u[2]));
    if ((ul - u[3]) * (ul - uMP) >= 1.e-10) return mp_limit(u + 1, ul);
    return ul;
}

// MP5LD, src/flux.F90:556-633
__device__ double mp5ld(const double* u, double w)
{
    const double vadp = 1.666666666666667e-2 * w;
    const double ul = (-vadp + 3.333333333333333e-2) * u[0] + (5.0 * vadp - 2.166666666666667e-1) * u[1] +
                      (-10.0 * vadp + 7.833333333333333e-1) * u[2] + (10.0 * vadp + 0.45) * u[3] +
                      (-5.0 * vadp - 5.e-2) * u[4] + vadp * u[5];
    const double uMP = u[2] + minmod2(u[3] - u[2], 4.0 * (u[2] - u[1]));
    if ((ul - u[2]) * (ul - uMP) >= 1.e-10) return mp_limit(u, ul);
    return ul;
}

// MP7LD, src/flux.F90:635-706: the limiter is switched by the shock flag
__device__ double mp7ld(const double* u, double w, bool lskt)
{
    const double vadp = 3.571428571428571e-3 * w;
    const double ul = (1.0 * vadp - 7.142857142857143e-3) * u[0] +
                      (-7.0 * vadp + 5.952380952380952e-2) * u[1] +
                      (21.0 * vadp - 0.240476190476190) * u[2] +
                      (-35.0 * vadp + 0.759523809523809) * u[3] +
                      (35.0 * vadp + 0.509523809523809) * u[4] +
                      (-21.0 * vadp - 9.047619047619047e-2) * u[5] +
                      (7.0 * vadp + 9.523809523809525e-3) * u[6] + (-1.0 * vadp) * u[7];
    if (lskt) return mp_limit(u + 1, ul);
    return ul;
}

// WENO5 / WENO5Z, src/flux.F90:761-821,936-994
__device__ double weno5(const double* u, bool z)
{
    const double uh0 = 0.333333333333333 * u[0] - 1.16666666666667 * u[1] + 1.83333333333333 * u[2];
    const double uh1 = -0.166666666666667 * u[1] + 0.833333333333333 * u[2] + 0.333333333333333 * u[3];
    const double uh2 = 0.333333333333333 * u[2] + 0.833333333333333 * u[3] - 0.166666666666667 * u[4];
    double a = u[0] - 2.0 * u[1] + u[2], b = u[0] - 4.0 * u[1] + 3.0 * u[2];
    const double bt0 = 1.08333333333333 * (a * a) + 0.25 * (b * b);
    a = u[1] - 2.0 * u[2] + u[3];
    b = u[1] - u[3];
    const double bt1 = 1.08333333333333 * (a * a) + 0.25 * (b * b);
    a = u[2] - 2.0 * u[3] + u[4];
    b = 3.0 * u[2] - 4.0 * u[3] + u[4];
    const double bt2 = 1.08333333333333 * (a * a) + 0.25 * (b * b);
    const double e0 = bt0 + 1.e-6, e1 = bt1 + 1.e-6, e2 = bt2 + 1.e-6;
    double a1, a2, a3;
    if (z) {
        const double tau5 = fabs(bt2 - bt0);
        a1 = 0.1 + 0.1 * (tau5 / (e0 * e0));
        a2 = 0.6 + 0.6 * (tau5 / (e1 * e1));
        a3 = 0.3 + 0.3 * (tau5 / (e2 * e2));
    } else {
        a1 = 0.1 / (e0 * e0);
        a2 = 0.6 / (e1 * e1);
        a3 = 0.3 / (e2 * e2);
    }
    const double as = a1 + a2 + a3;
    return a1 / as * uh0 + a2 / as * uh1 + a3 / as * uh2;
}

// WENO7 / WENO7Z, src/flux.F90:826-925,999-1050
__device__ double weno7(const double* u, bool z)
{
    double uh[4], bt[4];
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
    bt[0] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) + 1.084722222222222 * (s3 * s3);
    s1 = u[1] - 6.0 * u[2] + 3.0 * u[3] + 2.0 * u[4];
    s2 = u[2] - 2.0 * u[3] + u[4];
    s3 = -1.0 * u[1] + 3.0 * u[2] - 3.0 * u[3] + 1.0 * u[4];
    bt[1] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) + 1.084722222222222 * (s3 * s3);
    s1 = -2.0 * u[2] - 3.0 * u[3] + 6.0 * u[4] - 1.0 * u[5];
    s2 = u[2] - 2.0 * u[3] + u[4];
    s3 = -1.0 * u[2] + 3.0 * u[3] - 3.0 * u[4] + 1.0 * u[5];
    bt[2] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) + 1.084722222222222 * (s3 * s3);
    s1 = -11.0 * u[3] + 18.0 * u[4] - 9.0 * u[5] + 2.0 * u[6];
    s2 = 2.0 * u[3] - 5.0 * u[4] + 4.0 * u[5] - u[6];
    s3 = -1.0 * u[3] + 3.0 * u[4] - 3.0 * u[5] + 1.0 * u[6];
    bt[3] = 2.777777777777778e-2 * (s1 * s1) + 1.083333333333333 * (s2 * s2) + 1.084722222222222 * (s3 * s3);
    const double C[4] = {2.857142857142857e-2, 3.428571428571429e-1, 5.142857142857143e-1,
                         1.142857142857143e-1};
    const double tau7 = fabs(bt[3] - bt[0]);
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double e = bt[k] + 1.e-6;
        a[k] = z ? C[k] + C[k] * (tau7 / (e * e)) : C[k] / (e * e);
    }
    const double as = a[0] + a[1] + a[2] + a[3];
    return a[0] / as * uh[0] + a[1] / as * uh[1] + a[2] / as * uh[2] + a[3] / as * uh[3];
}

// ROUND, src/flux.F90:717-752
__device__ double round3(const double* u)
{
    const double z0c = (u[1] - u[0] + 1.e-16) / (u[2] - u[0] + 1.e-16);
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

// recons_exp, src/flux.F90:269-350 (schemes checked at set-up).  LD false
// leaves out the interior MP7-LD branch (reschem 5), whose registers would
// otherwise set every scheme's occupancy: a caller with reschem 5 takes
// LD true.
template <bool LD = true>
__device__ double recons_exp(const double* f, int inode, int dim, int ntype, int reschem,
                             bool shock, double bfac)
{
    if ((ntype == 1 && inode == 0) || (ntype == 2 && inode == dim - 1))
        return reschem == -1 ? f[3] : 0.5 * (f[3] + f[4]);
    if ((ntype == 1 && inode == 1) || (ntype == 2 && inode == dim - 2))
        return reschem == -1 ? f[3] : suw3(f + 2);
    if ((ntype == 1 && inode == 2) || (ntype == 2 && inode == dim - 3)) {
        switch (reschem) {
        case -1: return f[3];
        case 0: return suw5(f + 1);
        case 1: return weno5(f + 1, false);
        case 2: return weno5(f + 1, true);
        case 3: return mp5(f + 1);
        case 5: return mp5ld(f + 1, bfac);
        default: return round3(f + 2);
        }
    }
    switch (reschem) {
    case -1: return f[3];
    case 0: return suw7(f);
    case 1: return weno7(f, false);
    case 2: return weno7(f, true);
    case 3: return mp7(f);
    case 5:
        if constexpr (LD) return mp7ld(f, bfac, shock);
        else return __builtin_nan("");  // not reached: reschem 5 runs with LD true
    default: return round3(f + 2);
    }
}

// flux_steger_warming at one node, src/riemann.F90:23-161 (no species);
// sos = sqrt(T)/Ma (src/fludyna.F90:832-859)
__device__ void sw_split(double gamma, double mach, double rho, const double* vel, double prs,
                         double tmp, const double* q, const double* dxi, double jacob, double* fp,
                         double* fm)
{
    const double eps = 0.04;
    const double uu = dxi[0] * vel[0] + dxi[1] * vel[1] + dxi[2] * vel[2];
    const double var0 = 1.0 / sqrt(dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2]);
    const double gpd0 = dxi[0] * var0, gpd1 = dxi[1] * var0, gpd2 = dxi[2] * var0;
    const double gm2 = 0.5 / gamma;
    const double css = sqrt(tmp) / mach;
    const double csa = css / var0;
    const double lmach = uu / csa;
    if (lmach >= 1.0 || lmach <= -1.0) {
        double* f = lmach >= 1.0 ? fp : fm;
        double* z = lmach >= 1.0 ? fm : fp;
        f[0] = jacob * q[0] * uu;
        f[1] = jacob * (q[1] * uu + dxi[0] * prs);
        f[2] = jacob * (q[2] * uu + dxi[1] * prs);
        f[3] = jacob * (q[3] * uu + dxi[2] * prs);
        f[4] = jacob * (q[4] + prs) * uu;
#pragma unroll
        for (int k = 0; k < 5; ++k) z[k] = 0.0;
        return;
    }
    const double l0 = uu, l3 = uu + csa, l4 = uu - csa;
    const double p0 = 0.5 * (l0 + sqrt(l0 * l0 + eps * eps));
    const double p3 = 0.5 * (l3 + sqrt(l3 * l3 + eps * eps));
    const double p4 = 0.5 * (l4 + sqrt(l4 * l4 + eps * eps));
    const double lp[3] = {p0, p3, p4};
    const double lm[3] = {l0 - p0, l3 - p3, l4 - p4};
    const double fhi = 0.5 * (gamma - 1.0) * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double jro = jacob * rho;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const double* l = s == 0 ? lp : lm;
        double* f = s == 0 ? fp : fm;
        const double var1 = l[0];
        const double var2 = l[1] - l[2];
        const double var3 = 2.0 * l[0] - l[1] - l[2];
        const double var4 = var1 - var3 * gm2;
        f[0] = jro * var4;
        f[1] = jro * (var4 * vel[0] + var2 * css * gpd0 * gm2);
        f[2] = jro * (var4 * vel[1] + var2 * css * gpd1 * gm2);
        f[3] = jro * (var4 * vel[2] + var2 * css * gpd2 * gm2);
        f[4] = jacob * (var1 * q[4] + rho * (var2 * uu * var0 * css * gm2 -
                                             var3 * (fhi + css * css) * gm2 / (gamma - 1.0)));
    }
}

// Roe average and eigenvector scalars of chardecomp, src/solver.F90:
// 1964-2163 (no species); rows of LEV and REV are formed on demand.
struct Roe {
    double u1, u2, u3, hroe, cs, rcs, g1, g2, g3, ugp, b1, b2, rgp;
    int br;  // which normal component carries the 3rd/4th eigenvectors
};

__device__ Roe roe_state(double gamma, double ro_l, double p_l, double e_l, const double* vel_l,
                         const double* ddi_l, double ro_r, double p_r, double e_r,
                         const double* vel_r, const double* ddi_r)
{
    Roe r;
    const double sl = sqrt(ro_l), sr = sqrt(ro_r);
    const double w = sl / (sl + sr);
    const double w1 = 1.0 - w;
    r.u1 = w * vel_l[0] + w1 * vel_r[0];
    r.u2 = w * vel_l[1] + w1 * vel_r[1];
    r.u3 = w * vel_l[2] + w1 * vel_r[2];
    const double kroe = 0.5 * (r.u1 * r.u1 + r.u2 * r.u2 + r.u3 * r.u3);
    const double hl = (e_l + p_l) / ro_l;
    const double hr = (e_r + p_r) / ro_r;
    r.hroe = w * hl + w1 * hr;
    r.cs = sqrt((gamma - 1.0) * (r.hroe - kroe));
    r.rcs = 1.0 / r.cs;
    const double v1 = 0.5 * (ddi_l[0] + ddi_r[0]);
    const double v2 = 0.5 * (ddi_l[1] + ddi_r[1]);
    const double v3 = 0.5 * (ddi_l[2] + ddi_r[2]);
    const double v4 = 1.0 / sqrt(v1 * v1 + v2 * v2 + v3 * v3);
    r.g1 = v1 * v4;
    r.g2 = v2 * v4;
    r.g3 = v3 * v4;
    r.ugp = r.u1 * r.g1 + r.u2 * r.g2 + r.u3 * r.g3;
    r.b1 = (gamma - 1.0) / (r.cs * r.cs);
    r.b2 = 1.0 + 2.0 * r.b1 * kroe - r.b1 * r.hroe;
    const double rero = 1.e-12;
    r.br = fabs(v1) > rero ? 0 : (fabs(v2) > rero ? 1 : 2);
    r.rgp = 1.0 / (r.br == 0 ? r.g1 : (r.br == 1 ? r.g2 : r.g3));
    return r;
}

// row m (0..4) of LEV
__device__ __forceinline__ void lev_row(const Roe& r, int m, double* l)
{
    switch (m) {
    case 0:
        l[0] = 0.5 * (r.b2 + r.ugp * r.rcs);
        l[1] = -0.5 * (r.b1 * r.u1 + r.g1 * r.rcs);
        l[2] = -0.5 * (r.b1 * r.u2 + r.g2 * r.rcs);
        l[3] = -0.5 * (r.b1 * r.u3 + r.g3 * r.rcs);
        l[4] = 0.5 * r.b1;
        return;
    case 1:
        l[0] = 1.0 - r.b2;
        l[1] = r.b1 * r.u1;
        l[2] = r.b1 * r.u2;
        l[3] = r.b1 * r.u3;
        l[4] = -r.b1;
        return;
    case 4:
        l[0] = 0.5 * (r.b2 - r.ugp * r.rcs);
        l[1] = -0.5 * (r.b1 * r.u1 - r.g1 * r.rcs);
        l[2] = -0.5 * (r.b1 * r.u2 - r.g2 * r.rcs);
        l[3] = -0.5 * (r.b1 * r.u3 - r.g3 * r.rcs);
        l[4] = 0.5 * r.b1;
        return;
    }
    const double rgp = r.rgp;
    l[4] = 0.0;
    if (r.br == 0) {
        if (m == 2) {
            l[0] = (r.ugp * r.g2 - r.u2) * rgp;
            l[1] = -r.g2;
            l[2] = (1.0 - r.g2 * r.g2) * rgp;
            l[3] = -r.g2 * r.g3 * rgp;
        } else {
            l[0] = (r.ugp * r.g3 - r.u3) * rgp;
            l[1] = -r.g3;
            l[2] = -r.g2 * r.g3 * rgp;
            l[3] = (1.0 - r.g3 * r.g3) * rgp;
        }
    } else if (r.br == 1) {
        if (m == 2) {
            l[0] = (r.ugp * r.g1 - r.u1) * rgp;
            l[1] = (1.0 - r.g1 * r.g1) * rgp;
            l[2] = -r.g1;
            l[3] = -r.g1 * r.g3 * rgp;
        } else {
            l[0] = (r.ugp * r.g3 - r.u3) * rgp;
            l[1] = -r.g1 * r.g3 * rgp;
            l[2] = -r.g3;
            l[3] = (1.0 - r.g3 * r.g3) * rgp;
        }
    } else {
        if (m == 2) {
            l[0] = (r.ugp * r.g1 - r.u1) * rgp;
            l[1] = (1.0 - r.g1 * r.g1) * rgp;
            l[2] = -r.g1 * r.g2 * rgp;
            l[3] = -r.g1;
        } else {
            l[0] = (r.ugp * r.g2 - r.u2) * rgp;
            l[1] = -r.g1 * r.g2 * rgp;
            l[2] = (1.0 - r.g2 * r.g2) * rgp;
            l[3] = -r.g2;
        }
    }
}

// row m (0..4) of REV
__device__ __forceinline__ void rev_row(const Roe& r, int m, double* v)
{
    const double c = r.cs;
    switch (m) {
    case 0:
        v[0] = 1.0; v[1] = 1.0; v[2] = 0.0; v[3] = 0.0; v[4] = 1.0;
        return;
    case 4:
        v[0] = r.hroe - r.ugp * c;
        v[1] = r.hroe - 1.0 / r.b1;
        v[4] = r.hroe + r.ugp * c;
        if (r.br == 0) {
            v[2] = r.u2 * r.g1 - r.u1 * r.g2;
            v[3] = r.u3 * r.g1 - r.u1 * r.g3;
        } else if (r.br == 1) {
            v[2] = r.u1 * r.g2 - r.u2 * r.g1;
            v[3] = r.u3 * r.g2 - r.u2 * r.g3;
        } else {
            v[2] = r.u1 * r.g3 - r.u3 * r.g1;
            v[3] = r.u2 * r.g3 - r.u3 * r.g2;
        }
        return;
    }
    const double um = m == 1 ? r.u1 : (m == 2 ? r.u2 : r.u3);
    const double gm = m == 1 ? r.g1 : (m == 2 ? r.g2 : r.g3);
    v[0] = um - c * gm;
    v[1] = um;
    v[4] = um + c * gm;
    // columns 3 and 4 (src/solver.F90:2095-2146)
    if (r.br == 0) {
        v[2] = m == 1 ? -r.g2 : (m == 2 ? r.g1 : 0.0);
        v[3] = m == 1 ? -r.g3 : (m == 2 ? 0.0 : r.g1);
    } else if (r.br == 1) {
        v[2] = m == 1 ? r.g2 : (m == 2 ? -r.g1 : 0.0);
        v[3] = m == 1 ? 0.0 : (m == 2 ? -r.g3 : r.g2);
    } else {
        v[2] = m == 1 ? r.g3 : (m == 2 ? 0.0 : -r.g1);
        v[3] = m == 1 ? 0.0 : (m == 2 ? r.g3 : -r.g2);
    }
}

}  // namespace uwd
}  // namespace astr

// back to the default contraction of the rest of the library
#pragma clang fp contract(fast)
