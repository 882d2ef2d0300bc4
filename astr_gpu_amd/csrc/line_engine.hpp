// Device-side line engine: 32 lanes of a wavefront solve one compact line.
//
// A line of R tridiagonal rows (the pre-factored Thomas system of
// tridiagonal_thomas_solver, src/commfunc.F90:790-813) is spread over the 32
// lanes of a half-wavefront in chunks of L consecutive rows, so one
// wavefront solves two lines at once.  Both first-order recurrences of the
// Thomas solve,
//     forward   y_r = rhs_r * ac2_r - ac3_r * y_{r-1}
//     backward  x_r = y_r - ac1_r * x_{r+1},
// are evaluated as a chunked scan: each lane runs its chunk sequentially from
// a zero carry, the chunk-end values are combined across the 32 lanes with a
// Kogge-Stone scan of affine maps (four DPP row shifts inside each 16-lane
// row, then one step across the two rows), and every row is corrected by
// (product of the chunk's multipliers) x (carry).  The result equals the
// sequential solve up to rounding (a few ulp), with no per-line scratch.
//
// The right-hand sides restate compact_fd_rhs (src/derivative.F90:210-306)
// and compact_filter_rhs (src/filter.F90:156-285).  A lane loads its rows'
// stencil window from the LDS line (node n at ln[n + HM]) into registers
// once.  The filter's closure rows at the low end of a line without a
// boundary there (the 11-point halo rows) are evaluated by lane 0 from its
// window; the other closure rows are evaluated wave-uniformly (scalar
// coefficient loads) and handed to the lane that owns them.
#pragma once
#include <hip/hip_runtime.h>

namespace astr {

constexpr int HM = 5;    // halo width, src/commvar.F90:190
constexpr int LPL = 32;  // lanes per line

enum LineOp : int { OP_DERIV = 0, OP_FILTER = 1 };

// One pre-factored line system on the device.  t points at 3 arrays of
// LPL*L doubles: ac1, ac2, ac3 (rows >= R are zero).
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

// ---------------------------------------------------------------- DPP helpers
// update_dpp on the two halves of a double; lanes whose source is outside
// the pattern keep `ident` (bound_ctrl off).
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp(double v, double ident)
{
    const long long b = __double_as_longlong(v), ib = __double_as_longlong(ident);
    const int lo = __builtin_amdgcn_update_dpp((int)ib, (int)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(ib >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int DPP_ROW_SR1 = 0x111, DPP_ROW_SR2 = 0x112, DPP_ROW_SR4 = 0x114, DPP_ROW_SR8 = 0x118;
constexpr int DPP_ROW_SL1 = 0x101, DPP_ROW_SL2 = 0x102, DPP_ROW_SL4 = 0x104, DPP_ROW_SL8 = 0x108;
constexpr int DPP_WAVE_SR1 = 0x138, DPP_WAVE_SL1 = 0x130, DPP_ROW_BCAST15 = 0x142;

// one Kogge-Stone step: (A,B) <- (A,B) o (Ap,Bp), i.e. B = A*Bp + B, A = A*Ap
__device__ __forceinline__ void ks(double& A, double& B, double Ap, double Bp)
{
    B = fma(A, Bp, B);
    A = A * Ap;
}

// Inclusive scan over each 32-lane group of Y_l = A_l Y_{l-1} + B_l
// (forward: lane order); returns the exclusive carry Y_{l-1} (0 at lane 0).
__device__ __forceinline__ double scan_fwd(double A, double B, int lane)
{
    ks(A, B, dpp<DPP_ROW_SR1>(A, 1.0), dpp<DPP_ROW_SR1>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SR2>(A, 1.0), dpp<DPP_ROW_SR2>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SR4>(A, 1.0), dpp<DPP_ROW_SR4>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SR8>(A, 1.0), dpp<DPP_ROW_SR8>(B, 0.0));
    // rows 1 and 3 take the totals of rows 0 and 2 (lanes 15, 47)
    ks(A, B, dpp<DPP_ROW_BCAST15, 0xA>(A, 1.0), dpp<DPP_ROW_BCAST15, 0xA>(B, 0.0));
    double c = dpp<DPP_WAVE_SR1>(B, 0.0);
    return (lane & 31) == 0 ? 0.0 : c;
}

// Same in reverse lane order: X_l = A_l X_{l+1} + B_l; returns X_{l+1}.
__device__ __forceinline__ double scan_bwd(double A, double B, int lane)
{
    ks(A, B, dpp<DPP_ROW_SL1>(A, 1.0), dpp<DPP_ROW_SL1>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SL2>(A, 1.0), dpp<DPP_ROW_SL2>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SL4>(A, 1.0), dpp<DPP_ROW_SL4>(B, 0.0));
    ks(A, B, dpp<DPP_ROW_SL8>(A, 1.0), dpp<DPP_ROW_SL8>(B, 0.0));
    // rows 0 and 2 take the totals of rows 1 and 3 (lanes 16, 48)
    const double a16 = readlane(A, 16), b16 = readlane(B, 16);
    const double a48 = readlane(A, 48), b48 = readlane(B, 48);
    if ((lane & 16) == 0) {
        const bool hi = lane >= 32;
        ks(A, B, hi ? a48 : a16, hi ? b48 : b16);
    }
    double c = dpp<DPP_WAVE_SL1>(B, 0.0);
    return (lane & 31) == 31 ? 0.0 : c;
}

// ------------------------------------------------------------ right-hand sides

// Special rows of compact_fd_rhs (src/derivative.F90:230-294) for row r
// (node n = first + r); f points at node n in the LDS line.
__device__ __forceinline__ double rhs_deriv_edge(const double* f, int r, const LineTab& tb)
{
    const bool lob = lo_boundary(tb.ntype), hib = hi_boundary(tb.ntype);
    if (r == 0 && lob) return -2.5 * f[0] + 2.0 * f[1] + 0.5 * f[2];
    if (r == tb.R - 1 && hib) return 2.5 * f[0] - 2.0 * f[-1] - 0.5 * f[-2];
    if ((r == 1 && lob) || (r == tb.R - 2 && hib)) return 0.75 * (f[1] - f[-1]);
    // explicit 6th-order interface row
    return 0.75 * (f[1] - f[-1]) - 0.15 * (f[2] - f[-2]) + (1.0 / 60.0) * (f[3] - f[-3]);
}

// Special rows of compact_filter_rhs (src/filter.F90:176-264) for row r;
// ln is the LDS line (node n at ln[n + HM]).  Called with r uniform across
// the wavefront, so the coefficient reads are scalar.
__device__ __forceinline__ double rhs_filter_edge(const double* ln, int r, const LineTab& tb,
                                                  const FiltCoef* __restrict__ fc)
{
    const double* f = ln + HM;  // f[n] = node n
    const int R = tb.R;
    const bool lob = lo_boundary(tb.ntype), hib = hi_boundary(tb.ntype);
    double v = 0.0;
    if (r >= R - 3) {
        const int k = R - 1 - r;
        if (hib) {
            for (int j = 0; j <= 6; ++j) v = v + fc->cb[k][j] * f[tb.last - j];
        } else {
            for (int j = 0; j <= 10; ++j) v = v + fc->ch[k][j] * f[tb.last + 2 - j];
        }
        return v;
    }
    if (r <= 2) {
        if (lob) {
            for (int j = 0; j <= 6; ++j) v = v + fc->cb[r][j] * f[tb.first + j];
        } else {
            for (int j = 0; j <= 10; ++j) v = v + fc->ch[r][j] * f[tb.first - 2 + j];
        }
        return v;
    }
    const double* g = f + tb.first + r;
    if (r == 3 || r == R - 4)  // 6th-order row next to a boundary
        return fc->c6[0] * (g[0] + g[0]) + fc->c6[1] * (g[1] + g[-1]) +
               fc->c6[2] * (g[2] + g[-2]) + fc->c6[3] * (g[3] + g[-3]);
    // 8th-order row
    return fc->c8[0] * (g[0] + g[0]) + fc->c8[1] * (g[1] + g[-1]) +
           fc->c8[2] * (g[2] + g[-2]) + fc->c8[3] * (g[3] + g[-3]) + fc->c8[4] * (g[4] + g[-4]);
}

// Interior rows: 6th-order compact derivative / 10th-order filter.
__device__ __forceinline__ double rhs_deriv_in(const double* f)
{
    return (7.0 / 9.0) * (f[1] - f[-1]) + (1.0 / 36.0) * (f[2] - f[-2]);
}

__device__ __forceinline__ double rhs_filter_in(const double* g, const FiltCoef* __restrict__ fc)
{
    return fc->c10[0] * (g[0] + g[0]) + fc->c10[1] * (g[1] + g[-1]) +
           fc->c10[2] * (g[2] + g[-2]) + fc->c10[3] * (g[3] + g[-3]) +
           fc->c10[4] * (g[4] + g[-4]) + fc->c10[5] * (g[5] + g[-5]);
}

// Put v into b[r - sub*L] on the lane that owns row r (r uniform).  The
// register index is selected by a branch on the uniform t, so the write is
// one move on one lane rather than a select over the whole array.
template <int L>
__device__ __forceinline__ void set_row(double (&b)[L], int r, int sub, double v)
{
    const int owner = r / L, t = __builtin_amdgcn_readfirstlane(r - owner * L);
    if (sub == owner) {
        switch (t) {
#define ASTR_SET_ROW(k) \
    case k:             \
        if constexpr (k < L) b[k] = v; \
        break;
            ASTR_SET_ROW(0) ASTR_SET_ROW(1) ASTR_SET_ROW(2) ASTR_SET_ROW(3)
            ASTR_SET_ROW(4) ASTR_SET_ROW(5) ASTR_SET_ROW(6) ASTR_SET_ROW(7)
            ASTR_SET_ROW(8) ASTR_SET_ROW(9) ASTR_SET_ROW(10) ASTR_SET_ROW(11)
            ASTR_SET_ROW(12) ASTR_SET_ROW(13) ASTR_SET_ROW(14) ASTR_SET_ROW(15)
#undef ASTR_SET_ROW
        default: break;
        }
    }
}

template <int L>
__device__ __forceinline__ void thomas_rows(const LineTab& tb, double (&x)[L]);

// Solve the line held in LDS at ln (nodes -HM..dim+HM).  Called by a whole
// wavefront; lanes 0-31 and 32-63 each solve their own line (same system),
// lane sub of a half holds rows sub*L .. sub*L+L-1 in x[] on return.
template <int L, int OP>
__device__ __forceinline__ void solve_line(const double* ln, const LineTab& tb,
                                           const FiltCoef* __restrict__ fc, double (&x)[L])
{
    constexpr int HW = OP == OP_DERIV ? 2 : 5;  // interior stencil half-width
    constexpr int W = L + 2 * HW;
    const int lane = threadIdx.x & 63;
    const int sub = lane & 31;
    const int r0 = sub * L;
    const int R = tb.R;

    // interior right-hand side of every row from a register window
    {
        double w[W];
        // the window of the first and last lanes reaches into guard and
        // padding slots of the LDS row; only closure rows, which are
        // replaced below, or rows >= R read them
        const double* wp = ln + HM + tb.first + r0 - HW;  // node of w[0]
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = wp[k];
#pragma unroll
        for (int t = 0; t < L; ++t) {
            if constexpr (OP == OP_DERIV) x[t] = rhs_deriv_in(w + t + HW);
            else x[t] = rhs_filter_in(w + t + HW, fc);
        }
        if constexpr (OP == OP_FILTER && L >= 4) {
            // The closure rows at the low end of a line without a boundary
            // there (the 11-point ch rows of compact_filter_rhs) from lane
            // 0's window, in the reference's summation order: rows 0..2 are
            // its x[0..2], and the window (nodes first-5..first+L+4, so
            // L >= 4) holds their nodes first-2..first+8 at w[3..13].  (The
            // high end's rows sit at a run-time place in their owner's
            // window and are taken from LDS below.)
            if (!lo_boundary(tb.ntype) && sub == 0) {
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j <= 10; ++j) v = v + fc->ch[r][j] * w[j + 3];
                    x[r] = v;
                }
            }
        }
    }
    // the closure rows at both ends, computed wave-uniformly
    if constexpr (OP == OP_DERIV) {
        const int ne = (lo_boundary(tb.ntype) ? 2 : 1);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (j < ne) set_row(x, j, sub, rhs_deriv_edge(ln + HM + tb.first + j, j, tb));
        const int me = (hi_boundary(tb.ntype) ? 2 : 1);
        for (int j = 0; j < me; ++j) {
            const int r = R - 1 - j;
            set_row(x, r, sub, rhs_deriv_edge(ln + HM + tb.first + r, r, tb));
        }
    } else {
        // boundary closures, and the halo closure rows the windows above
        // did not give
        const bool lo_done = L >= 4 && !lo_boundary(tb.ntype);
        const int ne = lo_boundary(tb.ntype) ? 5 : (lo_done ? 0 : 3);
#pragma unroll
        for (int r = 0; r < 5; ++r)
            if (r < ne) set_row(x, r, sub, rhs_filter_edge(ln, r, tb, fc));
        const int me = hi_boundary(tb.ntype) ? 5 : 3;
        for (int j = 0; j < me; ++j) {
            const int r = R - 1 - j;
            set_row(x, r, sub, rhs_filter_edge(ln, r, tb, fc));
        }
    }

    thomas_rows<L>(tb, x);
}

// The pre-factored Thomas recurrences of tb on the rows a half-wavefront
// holds in x (lane sub: rows sub*L .. sub*L+L-1), as a chunked scan.
template <int L>
__device__ __forceinline__ void thomas_rows(const LineTab& tb, double (&x)[L])
{
    const int lane = threadIdx.x & 63;
    const int sub = lane & 31;
    const int S = LPL * L;
    const double* __restrict__ ac1 = tb.t;
    const double* __restrict__ ac2 = tb.t + S;
    const double* __restrict__ ac3 = tb.t + 2 * S;
    const int r0 = sub * L;
    const int R = tb.R;
    // forward elimination, chunk-local, with the chunk product g
    double prev = 0.0, g = 1.0;
    double gr[L];
#pragma unroll
    for (int t = 0; t < L; ++t) {
        const int r = r0 + t;
        const double b = r < R ? x[t] : 0.0;
        const double c3 = ac3[r];
        const double v = b * ac2[r] - c3 * prev;
        g = -c3 * g;
        gr[t] = g;
        x[t] = v;
        prev = v;
    }
    double cin = scan_fwd(g, prev, lane);
#pragma unroll
    for (int t = 0; t < L; ++t) x[t] = fma(gr[t], cin, x[t]);

    // back substitution, chunk-local, with the chunk product h
    double nxt = 0.0, h = 1.0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
        const double c1 = ac1[r0 + t];
        const double v = x[t] - c1 * nxt;
        h = -c1 * h;
        gr[t] = h;
        x[t] = v;
        nxt = v;
    }
    cin = scan_bwd(h, nxt, lane);
#pragma unroll
    for (int t = 0; t < L; ++t) x[t] = fma(gr[t], cin, x[t]);
}

}  // namespace astr
