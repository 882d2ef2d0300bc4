// Device-side line engine: one 64-lane wavefront solves one compact line.
//
// A line of R tridiagonal rows (the pre-factored Thomas system of
// tridiagonal_thomas_solver, src/commfunc.F90:790-813) is spread over the 64
// lanes in chunks of L consecutive rows.  Both first-order recurrences of the
// Thomas solve,
//     forward   y_r = rhs_r * ac2_r - ac3_r * y_{r-1}
//     backward  x_r = y_r - ac1_r * x_{r+1},
// are evaluated as a chunked scan: each lane runs its chunk sequentially from
// a zero carry, the chunk-end values are combined across lanes with a
// Kogge-Stone scan of affine maps (6 shuffle steps), and every row is then
// corrected by (product of the chunk's multipliers) x (carry).  The products
// g_r = prod(-ac3) and h_r = prod(-ac1) over the chunk depend only on the row,
// so they are precomputed on the host with the factors.  The result equals
// the sequential solve up to rounding (a few ulp), with no per-line scratch.
//
// The right-hand sides restate compact_fd_rhs (src/derivative.F90:210-306)
// and compact_filter_rhs (src/filter.F90:156-285); they read the line from
// LDS, where node n of the line sits at ln[n + HM].
#pragma once
#include <hip/hip_runtime.h>

namespace astr {

constexpr int HM = 5;  // halo width, src/commvar.F90:190

enum LineOp : int { OP_DERIV = 0, OP_FILTER = 1 };

// One pre-factored line system on the device.  t points at 5 arrays of
// 64*L doubles: ac1, ac2, ac3, g, h (rows >= R are zero).
struct LineTab {
    int R;      // rows
    int first;  // node index of row 0
    int last;   // node index of row R-1
    int ntype;  // 1 boundary at 0, 2 boundary at dim, 3 inner, 4 both
    int dim;    // line nodes 0..dim
    int op;     // LineOp
    int L;      // rows per lane
    int pad;
    const double* __restrict__ t;
};

// filter_coefficient_cal tables (src/filter.F90:299-432) used on the path.
struct FiltCoef {
    double c10[6];
    double c6[4];
    double c8[5];
    double cb[3][7];
    double ch[3][11];
};

__device__ __forceinline__ bool lo_boundary(int nt) { return nt == 1 || nt == 4; }
__device__ __forceinline__ bool hi_boundary(int nt) { return nt == 2 || nt == 4; }

// compact_fd_rhs for row r (node tb.first + r), 643c scheme.
__device__ __forceinline__ double rhs_deriv(const double* ln, int r, const LineTab& tb)
{
    const double* f = ln + HM + tb.first + r;  // f[s] = node (first+r+s)
    const bool lob = lo_boundary(tb.ntype), hib = hi_boundary(tb.ntype);
    if (r == 0) {
        if (lob) return -2.5 * f[0] + 2.0 * f[1] + 0.5 * f[2];
        return 0.75 * (f[1] - f[-1]) - 0.15 * (f[2] - f[-2]) + (1.0 / 60.0) * (f[3] - f[-3]);
    }
    if (r == tb.R - 1) {
        if (hib) return 2.5 * f[0] - 2.0 * f[-1] - 0.5 * f[-2];
        return 0.75 * (f[1] - f[-1]) - 0.15 * (f[2] - f[-2]) + (1.0 / 60.0) * (f[3] - f[-3]);
    }
    if ((r == 1 && lob) || (r == tb.R - 2 && hib)) return 0.75 * (f[1] - f[-1]);
    return (7.0 / 9.0) * (f[1] - f[-1]) + (1.0 / 36.0) * (f[2] - f[-2]);
}

// compact_filter_rhs for row r.
__device__ __forceinline__ double rhs_filter(const double* ln, int r, const LineTab& tb,
                                             const FiltCoef* __restrict__ fc)
{
    const double* f = ln + HM;  // f[n] = node n
    const int n = tb.first + r;
    const int R = tb.R;
    const bool lob = lo_boundary(tb.ntype), hib = hi_boundary(tb.ntype);
    if (r >= R - 3) {
        const int k = R - 1 - r;
        double v = 0.0;
        if (hib) {
            for (int j = 0; j <= 6; ++j) v = v + fc->cb[k][j] * f[tb.last - j];
        } else {
            for (int j = 0; j <= 10; ++j) v = v + fc->ch[k][j] * f[tb.last + 2 - j];
        }
        return v;
    }
    if (r <= 2) {
        double v = 0.0;
        if (lob) {
            for (int j = 0; j <= 6; ++j) v = v + fc->cb[r][j] * f[tb.first + j];
        } else {
            for (int j = 0; j <= 10; ++j) v = v + fc->ch[r][j] * f[tb.first - 2 + j];
        }
        return v;
    }
    const double* g = f + n;
    if ((lob && r == 3) || (hib && r == R - 4)) {
        return fc->c6[0] * (g[0] + g[0]) + fc->c6[1] * (g[1] + g[-1]) +
               fc->c6[2] * (g[2] + g[-2]) + fc->c6[3] * (g[3] + g[-3]);
    }
    if ((lob && r == 4) || (hib && r == R - 5)) {
        return fc->c8[0] * (g[0] + g[0]) + fc->c8[1] * (g[1] + g[-1]) +
               fc->c8[2] * (g[2] + g[-2]) + fc->c8[3] * (g[3] + g[-3]) +
               fc->c8[4] * (g[4] + g[-4]);
    }
    return fc->c10[0] * (g[0] + g[0]) + fc->c10[1] * (g[1] + g[-1]) +
           fc->c10[2] * (g[2] + g[-2]) + fc->c10[3] * (g[3] + g[-3]) +
           fc->c10[4] * (g[4] + g[-4]) + fc->c10[5] * (g[5] + g[-5]);
}

// Solve the line held in LDS at ln (nodes -HM..dim+HM); on return lane l
// holds the solution of rows l*L .. l*L+L-1 in x[].  Must be called by all
// 64 lanes of a wavefront together.
template <int L>
__device__ __forceinline__ void solve_line(const double* ln, const LineTab& tb,
                                           const FiltCoef* __restrict__ fc, double (&x)[L])
{
    const int lane = threadIdx.x & 63;
    const int stride = 64 * L;
    const double* __restrict__ ac1 = tb.t;
    const double* __restrict__ ac2 = tb.t + stride;
    const double* __restrict__ ac3 = tb.t + 2 * stride;
    const double* __restrict__ gp = tb.t + 3 * stride;
    const double* __restrict__ hp = tb.t + 4 * stride;
    const int r0 = lane * L;

    double y[L];
    double prev = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
        const int r = r0 + t;
        double b = 0.0;
        if (r < tb.R) b = (tb.op == OP_DERIV) ? rhs_deriv(ln, r, tb) : rhs_filter(ln, r, tb, fc);
        const double v = b * ac2[r] - ac3[r] * prev;
        y[t] = v;
        prev = v;
    }
    // inclusive scan over lanes of Y_l = A_l * Y_{l-1} + B_l
    double A = gp[r0 + L - 1], B = prev;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double Ap = __shfl_up(A, off, 64);
        const double Bp = __shfl_up(B, off, 64);
        if (lane >= off) {
            B = fma(A, Bp, B);
            A = A * Ap;
        }
    }
    double cin = __shfl_up(B, 1, 64);
    if (lane == 0) cin = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) y[t] = fma(gp[r0 + t], cin, y[t]);

    double nxt = 0.0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
        const double v = y[t] - ac1[r0 + t] * nxt;
        x[t] = v;
        nxt = v;
    }
    // reverse inclusive scan of X_l = H_l * X_{l+1} + x_loc(first row of l)
    A = hp[r0];
    B = nxt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double Ap = __shfl_down(A, off, 64);
        const double Bp = __shfl_down(B, off, 64);
        if (lane + off < 64) {
            B = fma(A, Bp, B);
            A = A * Ap;
        }
    }
    cin = __shfl_down(B, 1, 64);
    if (lane == 63) cin = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) x[t] = fma(hp[r0 + t], cin, x[t]);
}

}  // namespace astr
