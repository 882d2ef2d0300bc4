// MI355X (gfx950) FP64 solver for ASTR's per-timestep right-hand side.
//
// Device layout.  Every field is one padded 3-D array of doubles holding
// nodes -HM..n+HM on each axis (the reference's halo width, so the qswap
// halo semantics can be reproduced exactly):
//     index(i,j,k) = org + i + j*s1 + k*s2,
// with i fastest; node i=0 sits at a 128-byte boundary and the row pitch s1
// is a multiple of 16 doubles, so 16 consecutive i form one 128-byte line.
//
// Line sweeps.  A workgroup of 256 threads (4 wavefronts) owns a tile of
// lines and one quantity: 8 x-lines (one per half-wavefront) or 16 y/z
// lines (16 consecutive i, one 128-byte segment per node).  It stages the
// lines (nodes -HM..n+HM) in LDS with coalesced loads, each half-wavefront
// solves one line with the chunked-scan Thomas solver of line_engine.hpp,
// and the results overwrite the staged line and leave with coalesced stores.
//
// Per RK3 stage of the central path (src/mainloop.F90:392-476) the device
// runs:
//   filter x,y,z (in place)           filterq      src/comsolver.F90:514
//   qswap i, j, k parts                qswap        src/parallel.F90:4849
//   gradient sweeps x,y,z              gradcal      src/comsolver.F90:244
//   k_gflux (pointwise)                sigma, qflux and the axis fluxes
//                                      G = Fv - F of convrsdcal6/diffrsdcal6
//   flux sweeps x,y,z                  one derivative of G per axis into qrhs
//                                      (qrhs=-qrhs folded in); the z sweep
//                                      applies the RK3 combine
//   k_prim                             updatefvar   src/mainloop.F90:476
// The upwind paths (convrsduwd, convrsdcmp) run k_uwd per axis with the
// Ducros sensor and the explicit gradient; their RK3 combine is k_rk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/astr_hip.h"
#include "line_engine.hpp"
#include "upwind.hpp"

using namespace astr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ASTR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int TXY = 16;    // lines per y/z tile: 16 consecutive i = 128 bytes
constexpr int TXX = 8;     // lines per x tile: one per half-wavefront
constexpr int TX = TXY;    // largest tile
constexpr int WGT = 256;   // threads per sweep workgroup
constexpr int TG = 3;      // LDS guard slots before each staged line
constexpr int OI = 16;     // element offset of node i=0 in a row
static_assert(OI % 16 == 0, "node i = 0 starts a 128-byte line");
constexpr int LMAX = 16;   // rows per lane supported (line length <= 32*LMAX)
constexpr int STATS_SLOTS = 1024;  // steps whose statistics stay on the device

struct Geo {
    int n[3];    // cells per axis (nodes 0..n)
    long s[3];   // element stride per axis (s[0] = 1)
    long org;    // element offset of node (0,0,0)
};

__host__ __device__ __forceinline__ long gidx(const Geo& g, int i, int j, int k)
{
    return g.org + i + j * g.s[1] + k * g.s[2];
}

// Node i = 0 of every row starts a 128-byte line (OI = 16 doubles, rows a
// multiple of 16).  A kernel that writes nodes 0..n of a row also writes,
// where the field's cells past n are dead or rewritten with their own
// values, the nodes up to the end of the line holding node n: HBM takes a
// partly written line at a far higher cost than a whole one.
__host__ __device__ constexpr int line_end_i(int n) { return (n / 16 + 1) * 16; }

// The interior i-rows (j, k in 0..n) cut at line_end_i(n), laid end to end
// and taken one node per thread: a row is a whole number of 128-byte lines,
// so a wavefront covers whole lines only, also where it crosses from one
// row to the next.  False past the last row.
__device__ __forceinline__ bool row_node(const Geo& g, unsigned bid, int& i, int& j, int& k)
{
    const unsigned le = (unsigned)line_end_i(g.n[0]);
    const unsigned t = bid * blockDim.x + threadIdx.x;
    const unsigned row = t / le;
    const unsigned nj = (unsigned)g.n[1] + 1;
    if (row >= nj * ((unsigned)g.n[2] + 1)) return false;
    i = (int)(t - row * le);
    k = (int)(row / nj);
    j = (int)(row - (unsigned)k * nj);
    return true;
}
__device__ __forceinline__ bool row_node(const Geo& g, int& i, int& j, int& k)
{
    return row_node(g, blockIdx.x, i, j, k);
}

// store v back to f[gi] (a cell past node n that keeps its value, written so
// that its line is written whole)
__device__ __forceinline__ void keep_cell(double* f, long gi)
{
    double v = f[gi];
    asm volatile("" : "+v"(v));  // the store is not dropped as redundant
    f[gi] = v;
}

// ---------------------------------------------------------------- sweeps

enum Pass : int {
    PASS_FILTER = 0,       // filterq along one axis, q(1:5) in place
    PASS_GRAD = 1,         // gradcal along one axis: u,v,w,T
    PASS_RHS = 2,          // convrsdcal6 + diffrsdcal6 along one axis: the
                           // derivative of (viscous - inviscid) flux of q(1:5)
    PASS_DERIV_FIELD = 4,  // fds%central of one field (halos as given)
    PASS_FILTER_FIELD = 5  // compact_filter of one field (halos as given)
};

struct SweepArgs {
    Geo g;
    int dir;          // 0,1,2
    int pass;
    LineTab tab;      // line system of this axis (derivative or filter)
    const FiltCoef* fc;
    int na, nb;       // line index ranges: a in [0,na), b in [0,nb)
    int ntile_a;      // tiles along a
    int rs_in;
    // fields
    double* q[5];
    const double* prim[6];  // rho,u,v,w,p,T
    double* dvel[9];
    double* dtmp[3];
    double* qrhs[5];
    const double* gf[5];    // PASS_RHS: flux of this axis, G = Fv - F, q(1:5)
    const double* jac;
    const double* xi;       // metric along dir, xi[n+HM]
    const double* fin;      // field in (PASS_*_FIELD)
    double* fout;           // field out (PASS_*_FIELD)
    // PASS_RHS along z only: RK3 stage (1..3) applied to q in the epilogue
    // instead of storing qrhs (0: store qrhs)
    int rk_stage;
    double rka, rkb, rkc, dt;
    double* qsave[5];
    // where PASS_FILTER and the RK3 combine write q (q itself unless a
    // caller sends the result to another buffer)
    double* qo[5];
    int visc;  // PASS_RHS: include the viscous fluxes (diffterm)
    // mini-app variant (tgvsolver.F90): derivatives divided by the uniform
    // spacing h instead of multiplied by the metric; RK3 without J
    int mini;
    // PASS_GRAD on a grid with general metrics: the index-space derivative
    // is stored as it is, and k_metric_grad combines the three axes
    int raw;
    double h;
    // jac holds one value everywhere (a uniform grid): j0, read from no
    // array
    int juni;
    double j0;
    // the RK3 combine with qsave holding the stage-1 state q itself rather
    // than q J: stage 1 stores no qsave (its state stays in the buffer it
    // was read from while the result goes to qo), stages 2-3 form q J from
    // it (the same product, rounded alike)
    int qsraw;
};

// The RK3 combine in the reference's operation order, without FMA
// contraction (as k_rk and the checker evaluate it): q = (a qsave + b q J +
// c qrhs dt)/J (src/mainloop.F90:435-445), and the mini-app's form without
// J (tgvsolver.F90:1040-1060).
#pragma clang fp contract(off)
__device__ __forceinline__ double rk_combine(double ra, double rb, double rc, double dt, double qs,
                                             double qo, double jc, double r)
{
    return (ra * qs + rb * qo * jc + rc * r * dt) / jc;
}
__device__ __forceinline__ double rk_combine_mini(double ra, double rb, double rc, double dt,
                                                  double qs, double qo, double r)
{
    return ra * qs + rb * qo + rc * r * dt;
}
#pragma clang fp contract(fast)

// node -> element offset along the line, with periodic wrap (dataswap:
// plain copy of hm planes, src/parallel.F90:4170-4175) or direct.
__device__ __forceinline__ int wrap_node(int n, int dim)
{
    return n < 0 ? n + dim : (n > dim ? n - dim : n);
}

// Every workgroup solves one quantity of one tile.  Block ids are laid out
// so that the NQ quantity-workgroups of a tile are dispatched within a few
// dozen blocks of each other onto the same XCD (workgroups go round-robin
// over the 8 XCDs): the inputs they share (u_d, p, J of the flux pass) are
// then re-read from that XCD's L2, not from HBM.
template <int PASS>
struct PassCfg {
    static constexpr int NQ = PASS == PASS_FILTER ? 5 : PASS == PASS_GRAD ? 4 : PASS == PASS_RHS ? 5 : 1;
};

// block id -> (tile, quantity)
__device__ __forceinline__ void block_tile(int bid, int nq, int ntiles, int& tile, int& q)
{
    const int grp = 8 * nq;
    const int g = bid / grp, r = bid - g * grp;
    q = r >> 3;
    tile = g * 8 + (r & 7);
    (void)ntiles;
}

// load / store through a 32-bit byte offset from a wave-uniform base (every
// field is below 4 GB), so the address is one VGPR added to an SGPR pair
// (every field is global memory: the accesses say so, so that pointers read
// from an argument table rather than the kernel arguments also give global
// rather than flat instructions).  Both are streaming (nontemporal): a
// sweep reads each staged node once and its outputs are read back from HBM
// by a later launch, so neither is worth a place in the caches.  Streaming
// sweep accesses, k_prim's and k_gflux's streaming stores took a 256^3 step
// from 15.0 to 14.2 ms; streaming loads in k_gflux, whose halo strips
// re-read their periodic images, cost it 60 us per launch and are not used.
typedef __attribute__((address_space(1))) const char* gcptr;
typedef __attribute__((address_space(1))) char* gptr;
__device__ __forceinline__ double ldb(const double* base, unsigned int boff)
{
    return __builtin_nontemporal_load((__attribute__((address_space(1))) const double*)((gcptr)base + boff));
}
__device__ __forceinline__ void stb(double* base, unsigned int boff, double v)
{
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) double*)((gptr)base + boff));
}
// streaming store of one double of a pointwise kernel
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
// the same through the caches: the transposed sweeps (streaming accesses
// there measured 4 % slower over a whole reversal)
__device__ __forceinline__ double ldc(const double* base, unsigned int boff)
{
    return *(__attribute__((address_space(1))) const double*)((gcptr)base + boff);
}
__device__ __forceinline__ void stc(double* base, unsigned int boff, double v)
{
    *(__attribute__((address_space(1))) double*)((gptr)base + boff) = v;
}

// Write-back of the y/z flux sweeps for the nodes n0 + nstep*e (e <
// EMAX) of one lane, given the derivative o of G in the LDS row trow:
//   MODE 0  qrhs += o (y axis, or z when no RK3 combine follows)
//   MODE 1  z axis, RK3 stage 1: qsave = q J, then the combine
//   MODE 2  z axis, RK3 stages 2-3: the combine with the saved qsave
//   (qsraw: stage 1 stores nothing, stages 2-3 take qsave J)
//   MODE 3/4 the mini-app's combine without J (stage 1 / stages 2-3)
// The operands of a batch of RB nodes are loaded before the stores of the
// previous batch (a store to q or qrhs could alias a later load, so the
// compiler keeps program order); nodes past dim load node dim and store
// nothing.
// keep: a lane of a line past the tile's valid ones (i past n, up to the
// end of the 128-byte line) stores q unchanged, so the line of q is
// written whole and its halo cells keep their values.
// A sweep's arguments as one tile of quantity m sees them: the
// quantity's pointers picked out and the axis' extents read, so that only
// plain pointers and scalars remain (nothing indexed at run time).  The
// sweep body takes this view whether the arguments were kernel arguments
// (one box) or an entry of a member table (the line-queue engine): either
// way the values sit in scalar registers.
struct TileView {
    int dir, dim, sd, n0, na, ntile_a, rs_in, org, s1, s2;
    LineTab tab;
    const FiltCoef* fc;
    const double* src;   // the staged field of this quantity
    double* gdst;        // gradient: dvel(m, dir) or dtmp(dir)
    double *q, *qo, *qrhs, *qsave, *fout;
    const double *jac, *xi;
    int rk_stage, mini, raw, juni, qsraw;
    double j0;
    double rka, rkb, rkc, dt, h;
};

template <int PASS, bool XD, class A>
__device__ __forceinline__ TileView tile_view(const A& a, int m)
{
    TileView v;
    v.dir = XD ? 0 : a.dir;
    v.dim = a.g.n[v.dir];
    v.sd = (int)a.g.s[v.dir];
    v.n0 = a.g.n[0];
    v.na = a.na;
    v.ntile_a = a.ntile_a;
    v.rs_in = a.rs_in;
    v.org = (int)a.g.org;
    v.s1 = (int)a.g.s[1];
    v.s2 = (int)a.g.s[2];
    v.tab = a.tab;
    v.fc = a.fc;
    if constexpr (PASS == PASS_FILTER) v.src = a.q[m];
    else if constexpr (PASS == PASS_GRAD) v.src = a.prim[m < 3 ? 1 + m : 5];
    else if constexpr (PASS == PASS_RHS) v.src = a.gf[m];
    else v.src = a.fin;
    v.gdst = nullptr;
    if constexpr (PASS == PASS_GRAD) v.gdst = m < 3 ? a.dvel[m * 3 + v.dir] : a.dtmp[v.dir];
    const bool qm = PASS == PASS_FILTER || PASS == PASS_RHS;
    v.q = qm ? a.q[m] : nullptr;
    v.qo = qm ? a.qo[m] : nullptr;
    v.qrhs = PASS == PASS_RHS ? a.qrhs[m] : nullptr;
    v.qsave = PASS == PASS_RHS ? a.qsave[m] : nullptr;
    v.fout = a.fout;
    v.jac = a.jac;
    v.xi = a.xi;
    v.rk_stage = a.rk_stage;
    v.mini = a.mini;
    v.raw = a.raw;
    v.juni = a.juni;
    v.qsraw = a.qsraw;
    v.j0 = a.j0;
    v.rka = a.rka;
    v.rkb = a.rkb;
    v.rkc = a.rkc;
    v.dt = a.dt;
    v.h = a.h;
    return v;
}

template <int MODE, int EMAX>
__device__ __forceinline__ void rhs_epilogue(const TileView& a, const double* trow,
                                             int n0, int nstep, int dim, unsigned int lob,
                                             unsigned int sdb, bool keep)
{
    constexpr int RB = MODE == 0 ? 8 : 4;
    constexpr int NB = (EMAX + RB - 1) / RB;
    double cr[RB] = {}, cq[RB] = {}, cs[RB] = {}, cj[RB] = {};  // batch being combined
    double nr[RB] = {}, nq[RB] = {}, ns[RB] = {}, nj[RB] = {};  // next batch, loading
    auto load = [&](int b, double* r, double* q, double* sv, double* jv) {
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            const int e = b * RB + g;
            if (e < EMAX) {
                const int n = min(n0 + nstep * e, dim);
                const unsigned int od = lob + (unsigned int)n * sdb;
                r[g] = ldb(a.qrhs, od);
                if constexpr (MODE != 0) q[g] = ldb(a.q, od);
                if constexpr (MODE == 1 || MODE == 2) jv[g] = a.juni ? a.j0 : ldb(a.jac, od);
                if constexpr (MODE == 2 || MODE == 4) sv[g] = ldb(a.qsave, od);
            }
        }
    };
    load(0, cr, cq, cs, cj);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b + 1 < NB) load(b + 1, nr, nq, ns, nj);
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            const int e = b * RB + g;
            const int n = n0 + nstep * e;
            if (e < EMAX && n <= dim) {
                const double o = trow[n + HM];
                const unsigned int od = lob + (unsigned int)n * sdb;
                if constexpr (MODE == 0) {
                    stb(a.qrhs, od, cr[g] + (a.mini ? o / a.h : o));
                } else if constexpr (MODE == 1 || MODE == 2) {
                    // the RK3 combine of src/mainloop.F90:423-445 for this
                    // quantity; the workgroup owns these nodes of q_m
                    const double r = cr[g] + o;
                    double qs;
                    if constexpr (MODE == 1) {
                        qs = cq[g] * cj[g];
                        if (!a.qsraw) stb(a.qsave, od, qs);
                    } else {
                        qs = a.qsraw ? cs[g] * cj[g] : cs[g];
                    }
                    stb(a.qo, od,
                        keep ? cq[g] : rk_combine(a.rka, a.rkb, a.rkc, a.dt, qs, cq[g], cj[g], r));
                } else {
                    // rk3 of tgvsolver.F90:1040-1060: q = a qsave + b q + c dt qrhs
                    const double r = cr[g] + o / a.h;
                    double qs;
                    if constexpr (MODE == 3) {
                        qs = cq[g];
                        stb(a.qsave, od, qs);
                    } else {
                        qs = cs[g];
                    }
                    stb(a.qo, od, keep ? cq[g] : rk_combine_mini(a.rka, a.rkb, a.rkc, a.dt, qs, cq[g], r));
                }
            }
        }
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            cr[g] = nr[g];
            cq[g] = nq[g];
            cs[g] = ns[g];
            cj[g] = nj[g];
        }
    }
}

// One tile of lines and quantity m: stage, solve, write back.  Called by
// every thread of the workgroup with the same tile and quantity.  The
// staging and the solve read the view of args, the write-back that of
// args_w, formed after the solve (the same arguments; the queue kernel
// hands them over through a second pointer so that the write-back's
// pointers are read again rather than kept in registers through the solve).
template <int L, int PASS, bool XD, class A>
__device__ __forceinline__ void sweep_tile(const A& args, const A& args_w, int tile_id, int m)
{
    const TileView a = tile_view<PASS, XD>(args, m);
    const LineTab& tab = a.tab;
    // One LDS tile; line l stages node n-HM at tile[l*rs_in + TG + n].  The TG
    // guard slots and the tail padding keep every window read and every
    // (clamped) staging write inside the tile, so neither needs a branch.
    // The solution of a line overwrites that line's slots in place, which is
    // safe because only the half-wavefront solving a line reads it.
    extern __shared__ double tile[];
    constexpr int OP = (PASS == PASS_FILTER || PASS == PASS_FILTER_FIELD) ? OP_FILTER : OP_DERIV;
    // XD: x-lines (dir 0), else y- or z-lines
    const int dir = a.dir;
    constexpr bool xdir = XD;
    constexpr int tx = XD ? TXX : TXY;
    // staged nodes per thread: x 32 per step (L+1 steps), y/z 16 per step
    // (2L+1 steps); dim+11 <= 32*L+4
    constexpr int EMAX = 2 * L + 1;

    const int ta = tile_id % a.ntile_a;
    const int b = tile_id / a.ntile_a;
    const int a0 = ta * tx;
    const int dim = a.dim;
    const int sd = a.sd;
    const int nvalid = min(tx, a.na - a0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int half = tid >> 5;  // half-wavefront: solves lines half (+8 for y/z)

    // Staging map.  x-lines are contiguous: the half-wavefront that solves a
    // line also stages and stores it (32 consecutive nodes per access).  y/z
    // tiles are 16 consecutive i: a wavefront moves 4 nodes x 16 lines.
    const int ll = xdir ? half : (tid & 15);
    const int n0 = xdir ? (tid & 31) : (tid >> 4);
    constexpr int nstep = XD ? 32 : 16;
    constexpr int esteps = XD ? L + 1 : EMAX;
    const bool lvalid = ll < nvalid;
    const int lc = min(ll, nvalid - 1);  // staged line (a valid one for padding lanes)
    int lo;  // element offset of node 0 of line lc
    {
        const int av = a0 + lc;
        if (XD) lo = a.org + av * a.s1 + b * a.s2;
        else if (dir == 1) lo = a.org + av + b * a.s2;
        else lo = a.org + av + b * a.s1;
    }
    const unsigned int lob = (unsigned int)lo * 8u, sdb = (unsigned int)sd * 8u;
    double* const trow = tile + ll * a.rs_in + TG;
    // Fields whose cells past node n are dead (dvel, dtmp, qrhs) are written
    // to the end of the 128-byte line holding node n (line_end_i): x lines
    // run on past dim, and in y/z tiles the lanes of lines a0 + ll past the
    // last valid one store too, at their own line (lob_w), whatever their
    // LDS row holds.
    const bool dead = PASS == PASS_GRAD || (PASS == PASS_RHS && (XD || a.rk_stage == 0));
    // the RK3 combine of the z sweep stores q unchanged on those lines; the
    // y/z filter stages those lines themselves, leaves them unsolved and
    // stores them back as staged
    const bool keepq = PASS == PASS_RHS && !XD && a.rk_stage != 0;
    const bool keepf = PASS == PASS_FILTER && !XD;
    const bool lw = lvalid || (!XD && (dead || keepq || keepf) && a0 + ll < line_end_i(a.n0));
    unsigned int lob_w = lob;
    if (!XD && !lvalid) {
        const int av = a0 + ll;
        const int low = dir == 1 ? a.org + av + b * a.s2 : a.org + av + b * a.s1;
        lob_w = (unsigned int)low * 8u;
    }
    const int nlast = XD && dead ? line_end_i(dim) - 1 : dim;  // last node stored
    const unsigned int lob_s = keepf && lw ? lob_w : lob;      // staged line
    auto sync = [&]() {
        if constexpr (XD) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            __syncthreads();
        }
    };

    // the metric factors of the rows this lane solves (gradient pass)
    double xr[PASS == PASS_GRAD ? L : 1];
    if constexpr (PASS == PASS_GRAD) {
        const int rb = tab.first + (lane & 31) * L;
#pragma unroll
        for (int t = 0; t < L; ++t) xr[t] = a.xi[min(max(rb + t, -HM), dim + HM) + HM];
    }

    // ---------------- stage the lines of this quantity in LDS
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        if (e < esteps) {  // compile-time
            const int n = n0 + nstep * e;
            const int node = min(n - HM, dim + HM);  // clamp the tail padding
            const unsigned int od = lob + (unsigned int)node * sdb;
            double v;
            if constexpr (PASS == PASS_FILTER) {
                // filterq's dataswap: a periodic line reads its halos as
                // copies of its nodes; a bounded line's closures read none
                const int nd = tab.ntype == 3 ? wrap_node(node, dim) : node;
                v = ldb(a.src, lob_s + (unsigned int)nd * sdb);
            } else if constexpr (PASS == PASS_GRAD) {
                v = ldb(a.src, od);
            } else if constexpr (PASS == PASS_RHS) {
                // the flux of this axis, viscous minus inviscid (k_gflux)
                v = ldb(a.src, od);
            } else {
                v = ldb(a.src, od);
            }
            trow[n] = v;
        }
    }
    sync();

    // ---------------- solve: each half-wavefront takes its lines
    for (int l = half; l < tx; l += 8) {
        if (l - (lane >> 5) >= nvalid) break;  // wave-uniform: both halves invalid
        double x[L];
        double* ln = tile + l * a.rs_in + TG;
        const double keep0 = ln[HM], keepd = ln[HM + dim];
        solve_line<L, OP>(ln, tab, a.fc, x);
        if constexpr (PASS == PASS_GRAD) {
            // gradcal's metric factor dxi(d,d) (src/comsolver.F90:304-318)
            // on the lane's rows, loaded with the staging
            if (!a.mini && !a.raw) {
#pragma unroll
                for (int t = 0; t < L; ++t) x[t] *= xr[t];
            }
        }
        // every row goes back to its node's slot: rows outside 0..dim (and
        // the zero padding rows past R) land in slots of the staged input,
        // which the half-wavefront has finished reading and nothing reads
        // again
        const int sub = lane & 31;
        if (!keepf || l < nvalid) {
#pragma unroll
            for (int t = 0; t < L; ++t) ln[HM + tab.first + sub * L + t] = x[t];
        }
        if constexpr (OP == OP_FILTER) {
            // compact_filter keeps the boundary node (src/filter.F90:141-142);
            // the lane that wrote the node overwrites it
            if (lo_boundary(tab.ntype) && sub == (0 - tab.first) / L) ln[HM] = keep0;
            if (hi_boundary(tab.ntype) && sub == (dim - tab.first) / L) ln[HM + dim] = keepd;
        }
    }
    sync();

    // ---------------- write back nodes 0..dim
    const TileView w = tile_view<PASS, XD>(args_w, m);
    if constexpr (PASS == PASS_RHS && !XD) {
        // read-modify-write of qrhs (y) or the RK3 combine (z), specialised
        // per mode so that each is straight-line code
        if (lw) {
            const int mode = w.rk_stage == 0 ? 0 : (w.mini ? 3 : 1) + (w.rk_stage > 1);
            switch (mode) {
            case 0: rhs_epilogue<0, EMAX>(w, trow, n0, nstep, dim, lob_w, sdb, false); break;
            case 1: rhs_epilogue<1, EMAX>(w, trow, n0, nstep, dim, lob_w, sdb, !lvalid); break;
            case 2: rhs_epilogue<2, EMAX>(w, trow, n0, nstep, dim, lob_w, sdb, !lvalid); break;
            case 3: rhs_epilogue<3, EMAX>(w, trow, n0, nstep, dim, lob_w, sdb, !lvalid); break;
            default: rhs_epilogue<4, EMAX>(w, trow, n0, nstep, dim, lob_w, sdb, !lvalid); break;
            }
        }
        return;
    }
    if (lw) {
#pragma unroll
        for (int e = 0; e < EMAX; ++e) {
            const int n = n0 + nstep * e;
            if (e < esteps && n <= nlast) {
                const double o = trow[n + HM];
                const unsigned int od = lob_w + (unsigned int)n * sdb;
                if constexpr (PASS == PASS_FILTER) {
                    stb(w.qo, od, o);
                } else if constexpr (PASS == PASS_GRAD) {
                    // gradcal: d/dxi times dxi(d,d) (src/comsolver.F90:304-318),
                    // or df/dx (tgvsolver.F90:420-470)
                    const double gv = w.mini ? o / w.h : o;  // scaled by xi in the solve
                    stb(w.gdst, od, gv);
                } else if constexpr (PASS == PASS_RHS) {
                    // x: the first axis sets qrhs
                    stb(w.qrhs, od, w.mini ? o / w.h : o);
                } else {
                    stb(w.fout, od, o);
                }
            }
        }
    }
}

template <int L, int PASS, bool XD>
__global__ __launch_bounds__(WGT, 4) void k_sweep(SweepArgs a)
{
    int tile_id, m;
    block_tile(blockIdx.x, PassCfg<PASS>::NQ, a.ntile_a * a.nb, tile_id, m);
    if (tile_id >= a.ntile_a * a.nb) return;  // padding blocks of the last group
    sweep_tile<L, PASS, XD>(a, a, tile_id, m);
}

// ------------------------------------------------------ transposed sweeps
//
// The adjoint of a compact line operator.  The forward maps the staged line
// f (nodes -HM..dim+HM) to x = A^-1 B f on the rows first..last and keeps
// x on the nodes 0..dim; its transpose maps the adjoint g of those nodes to
//     f = B^T A^-T (g on the rows 0..dim, zero on the others).
// A^T is tridiagonal too: its Thomas factors are built on the host and
// solved by the same chunked scan.  B^T is the transposed interior stencil
// (antisymmetric for the derivative, symmetric for the filter) plus, for
// the closure rows at each end, the difference between their own stencil
// and the interior one, held as dense blocks of TC columns.

constexpr int TC = 16;  // columns of a closure-correction block

struct TransTab {
    int nlo, nhi;       // closure rows: 0..nlo-1 and R-nhi..R-1
    int lo0, hi0;       // first column node of each block (first-5, last-10)
    const double* dlo;  // nlo x TC: B - B_interior of row r at column lo0 + c
    const double* dhi;  // nhi x TC: the same for row R-1-k at column hi0 + c
};

struct TSweepArgs {
    Geo g;
    int dir;
    LineTab tab;        // the transposed system A^T of this axis and operator
    TransTab tt;
    const FiltCoef* fc;
    int na, nb, ntile_a, rs_in;
    int nq;              // quantities of this launch
    const double* fin[5];  // adjoint of the forward output, nodes 0..dim of a line
    double* fout[5];       // adjoint of the forward input
    const double* xi;   // when set, fin is scaled by xi[n+HM] first (gradcal's metric)
    int wrap;           // 1: the forward read its halos as periodic copies of the
                        // nodes (dataswap): fold their adjoint into 0..dim;
                        // 0: the halos are outputs too (nodes -HM..dim+HM)
    int accumulate;     // add into fout instead of overwriting it
};

// the interior filter coefficients c10 (filter_coefficient_cal), held in
// registers for the whole kernel
struct C10 {
    double c[6];
};

// LDS guard slots before each line of a transposed sweep, zero: the
// interior stencil of B^T then reads zeros beyond node -HM (the tail slots
// past dim+HM hold zeros from the staging)
constexpr int TGT = 5;

// (B^T w) at column node j of one line with the interior stencil of every
// row, closure rows included (their difference is added separately); w
// holds the solved rows at ln[n + HM] and zero elsewhere
template <int OP>
__device__ __forceinline__ double bt_interior(const double* ln, int j, const C10& cf)
{
    const double* w = ln + j + HM;
    if constexpr (OP == OP_DERIV) {
        return (7.0 / 9.0) * (w[-1] - w[1]) + (1.0 / 36.0) * (w[-2] - w[2]);
    } else {
        double v = cf.c[0] * (w[0] + w[0]);
#pragma unroll
        for (int k = 1; k <= 5; ++k) v += cf.c[k] * (w[-k] + w[k]);
        return v;
    }
}

// One workgroup per tile of lines and quantity (the forward sweep's tiles,
// staging map and XCD-grouped block ids): stage g, solve A^T per
// half-wavefront, form B^T w in place, write the adjoint of the line's
// input.
// 5 waves per SIMD where the registers allow it without spilling (L <= 6:
// the 128^3 adjoint's reversal 221 -> 212 ms), else 4
template <int L, int OP, bool XD>
__global__ __launch_bounds__(WGT, L <= 6 ? 5 : 4) void k_sweep_t(TSweepArgs a)
{
    extern __shared__ double tile[];
    const int dir = XD ? 0 : a.dir;
    constexpr int tx = XD ? TXX : TXY;
    constexpr int EMAX = 2 * L + 1;
    int tile_id, m;
    block_tile(blockIdx.x, a.nq, a.ntile_a * a.nb, tile_id, m);
    if (tile_id >= a.ntile_a * a.nb) return;  // padding blocks of the last group
    const double* const fin = a.fin[m];
    double* const fout = a.fout[m];
    const int ta = tile_id % a.ntile_a;
    const int b = tile_id / a.ntile_a;
    const int a0 = ta * tx;
    const int dim = a.g.n[dir];
    const int sd = (int)a.g.s[dir];
    const int nvalid = min(tx, a.na - a0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int half = tid >> 5;
    const int ll = XD ? half : (tid & 15);
    const int n0 = XD ? (tid & 31) : (tid >> 4);
    constexpr int nstep = XD ? 32 : 16;
    constexpr int esteps = XD ? L + 1 : EMAX;
    const bool lvalid = ll < nvalid;
    const int lc = min(ll, nvalid - 1);
    int lo;
    {
        const int av = a0 + lc;
        if (XD) lo = (int)a.g.org + av * (int)a.g.s[1] + b * (int)a.g.s[2];
        else if (dir == 1) lo = (int)a.g.org + av + b * (int)a.g.s[2];
        else lo = (int)a.g.org + av + b * (int)a.g.s[1];
    }
    const unsigned int lob = (unsigned int)lo * 8u, sdb = (unsigned int)sd * 8u;
    double* const trow = tile + ll * a.rs_in + TGT;
    C10 cf;  // wave-uniform loads before any divergent branch
#pragma unroll
    for (int k = 0; k < 6; ++k) cf.c[k] = OP == OP_FILTER ? a.fc->c10[k] : 0.0;
    // the closure rows' corrections this lane applies on every line: lanes
    // 0-15 column lo0 + sub of the low block, lanes 16-31 column hi0 + sub -
    // 16 of the high block (rows past nlo / nhi contribute zero)
    const int csub = lane & 31;
    double dcf[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        if (csub < TC) dcf[r] = r < a.tt.nlo ? a.tt.dlo[r * TC + csub] : 0.0;
        else dcf[r] = r < a.tt.nhi ? a.tt.dhi[r * TC + csub - TC] : 0.0;
    }
    auto wsync = []() {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };
    auto sync = [&]() {
        if constexpr (XD) wsync();
        else __syncthreads();
    };

    // ---------------- stage g on the nodes 0..dim, zero on the rest and on
    // the guard slots (x: each half-wavefront its own line)
    if (XD) {
        if (csub < TGT) tile[half * a.rs_in + csub] = 0.0;
    } else if (tid < tx * TGT) {
        tile[(tid / TGT) * a.rs_in + tid % TGT] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        if (e < esteps) {
            const int n = n0 + nstep * e;
            const int node = n - HM;
            double v = 0.0;
            if (node >= 0 && node <= dim) {
                v = ldc(fin, lob + (unsigned int)node * sdb);
                if (a.xi) v *= a.xi[node + HM];
            }
            trow[n] = v;
        }
    }
    sync();

    // ---------------- A^T w = g, then B^T w, one line per half-wavefront
    for (int l = half; l < tx; l += 8) {
        if (l - (lane >> 5) >= nvalid) break;  // wave-uniform: both halves invalid
        double* ln = tile + l * a.rs_in + TGT;
        const int sub = lane & 31;
        const int r0 = sub * L;
        const int R = a.tab.R;
        // compact_filter keeps the input at a bounded end node (src/filter.F90:
        // 141-142): the adjoint there bypasses the solve
        const bool k0 = OP == OP_FILTER && lo_boundary(a.tab.ntype);
        const bool kd = OP == OP_FILTER && hi_boundary(a.tab.ntype);
        const double g0 = ln[HM], gd = ln[HM + dim];
        double x[L];
#pragma unroll
        for (int t = 0; t < L; ++t) {
            const int node = a.tab.first + r0 + t;
            const bool kept = (k0 && node == 0) || (kd && node == dim);
            x[t] = r0 + t < R && !kept ? ln[HM + node] : 0.0;
        }
        thomas_rows<L>(a.tab, x);
        wsync();
        if (l < nvalid) {
#pragma unroll
            for (int t = 0; t < L; ++t)
                if (r0 + t < R) ln[HM + a.tab.first + r0 + t] = x[t];
        }
        wsync();
        // the closure rows' difference from the interior stencil, on this
        // lane's column of the low or high block (the rows' w are
        // wave-uniform reads)
        double corr = 0.0;
#pragma unroll
        for (int r = 0; r < 5; ++r)
            corr += dcf[r] * ln[HM + (sub < TC ? a.tab.first + r : a.tab.last - r)];
        // B^T w with the interior stencil, columns j = -HM + sub + 32 t (a
        // half-wavefront reads 32 consecutive LDS slots at a time)
        double f[L + 1];
#pragma unroll
        for (int t = 0; t <= L; ++t) {
            const int j = -HM + sub + 32 * t;
            f[t] = j <= dim + HM ? bt_interior<OP>(ln, j, cf) : 0.0;
            if (k0 && j == 0) f[t] += g0;
            if (kd && j == dim) f[t] += gd;
        }
        wsync();
        if (l < nvalid) {
#pragma unroll
            for (int t = 0; t <= L; ++t) {
                const int j = -HM + sub + 32 * t;
                if (j <= dim + HM) ln[HM + j] = f[t];
            }
        }
        wsync();
        // the corrections, low block first (the blocks overlap on short lines)
        const int col = sub < TC ? a.tt.lo0 + sub : a.tt.hi0 + sub - TC;
        const bool cin = l < nvalid && col >= -HM && col <= dim + HM;
        if (cin && sub < TC) ln[HM + col] += corr;
        wsync();
        if (cin && sub >= TC) ln[HM + col] += corr;
    }
    sync();

    // ---------------- write the adjoint of the line's input
    if (lvalid) {
        // accumulate: the old values of every node of the lane are loaded
        // before the first store (a store could alias a later load)
        double old[EMAX];
#pragma unroll
        for (int e = 0; e < EMAX; ++e) {
            old[e] = 0.0;
            const int node = min(n0 + nstep * e - HM, dim + HM);
            if (a.accumulate && e < esteps) old[e] = ldc(fout, lob + (unsigned int)node * sdb);
        }
#pragma unroll
        for (int e = 0; e < EMAX; ++e) {
            if (e >= esteps) continue;
            const int n = n0 + nstep * e;
            const int node = n - HM;
            double v;
            if (a.wrap) {
                if (node < 0 || node > dim) continue;
                // the forward's halos were f(-k) = f(dim-k), f(dim+k) = f(k)
                v = trow[n];
                if (node >= dim - HM && node < dim) v += trow[node - dim + HM];
                if (node >= 1 && node <= HM) v += trow[dim + node + HM];
            } else {
                if (node > dim + HM) continue;
                v = trow[n];
            }
            stc(fout, lob + (unsigned int)node * sdb, v + old[e]);
        }
    }
}

// ---------------------------------------------------------------- pointwise

struct PtArgs {
    Geo g;
    double* q[5];
    double* prim[6];
    double* qsave[5];
    const double* qrhs[5];
    const double* dvel[9];
    const double* dtmp[3];
    double* gfl[15];        // k_gflux: G = Fv - F per axis d and quantity m, 5d+m
    const double* xi[3];    // metric along each axis, xi[n+HM]
    int visc;               // include the viscous fluxes (diffterm)
    int mini;               // mini-app variant (tgvsolver.F90) flux forms
    int per[3];             // axis d periodic (npdc 3): k_gflux writes its halo G
    // general metrics (curv 1): dxf[3d+e] = d(xi_d)/d(x_e) at every node;
    // else the rectilinear xi lines
    int curv;
    const double* dxf[9];
    int juni;  // jac is j0 everywhere
    double j0;
    // k_prim leaves the density out: within a step of the central path
    // nothing reads it before the last stage's k_prim (the statistics read
    // it at the next step's first stage)
    int skip_rho;
    const double* jac;
    double const1, const2, const5, const6;
    double reynolds, prandtl, tempconst, tempconst1;
    double rka, rkb, rkc, dt;
    int stage;
    int axis;               // k_qswap2d_line: the periodic in-plane axis
    double* part;
};

// (i, j, k) of flat index p over an ni x nj x nk box, i fastest, in 32-bit
// arithmetic (every box here has fewer than 2^32 nodes; a 64-bit division
// is a long instruction sequence on the GPU)
__device__ __forceinline__ void split3(long p, long ni, long nj, int& i, int& j, int& k)
{
    const unsigned up = (unsigned)p, ui = (unsigned)ni, uj = (unsigned)nj;
    const unsigned r = up / ui;
    i = (int)(up - r * ui);
    k = (int)(r / uj);
    j = (int)(r - (unsigned)k * uj);
}

__device__ __forceinline__ void interior_ijk(const Geo& g, long p, int& i, int& j, int& k)
{
    split3(p, g.n[0] + 1, g.n[1] + 1, i, j, k);
}

// q2fvar_3da, src/fludyna.F90:545-634 (nondimensional, numq = 5), evaluated
// without FMA contraction (in the reference's operation order, as the
// upwind checker does)
#pragma clang fp contract(off)
__device__ __forceinline__ void q2fvar(const PtArgs& a, const double qv[5], long gi,
                                       bool write_rho = true)
{
    const double rho = qv[0];
    const double u = qv[1] / rho, v = qv[2] / rho, w = qv[3] / rho;
    const double p = (qv[4] - 0.5 * rho * (u * u + v * v + w * w)) / a.const6;
    if (write_rho) st_nt(&a.prim[0][gi], rho);
    st_nt(&a.prim[1][gi], u);
    st_nt(&a.prim[2][gi], v);
    st_nt(&a.prim[3][gi], w);
    st_nt(&a.prim[4][gi], p);
    st_nt(&a.prim[5][gi], p / rho * a.const2);
}
#pragma clang fp contract(fast)

// miucal, src/fludyna.F90:791-820
__device__ __forceinline__ double miucal(const PtArgs& a, double t)
{
    return t * sqrt(t) * a.tempconst1 / (t + a.tempconst);
}

// qswap of a periodic three-dimensional box, single rank
// (src/parallel.F90:4877-5200), one kernel per axis in the reference's
// order:
//   i  duplicate nodes q(0) = q(n) = (q(0) + q(n))/2 on every interior row,
//      then the i-halos copied from the interior before any other average
//      (src/parallel.F90:4884-4885, 4877-4915);
//   j  the same along j, over i and k in 0..n, from the i-averaged state;
//   k  the same along k, over i and j in 0..n, from the i- and j-averaged
//      state.
// Every halo cell and duplicate node written gets its primitives refreshed
// (q2fvar); a node that is a duplicate along several axes ends with the
// primitives of its last average.  Within one kernel the cells read (the
// interior sources and the duplicate pairs) and the cells written are
// disjoint or owned by the same thread, so no thread reads what another
// writes.
constexpr int QSX_LANES = 64;  // threads per row of the i part

__device__ __forceinline__ void store_q(const PtArgs& a, long gi, const double (&qv)[5])
{
#pragma unroll
    for (int m = 0; m < 5; ++m) a.q[m][gi] = qv[m];
    q2fvar(a, qv, gi);
}

// i part: one wavefront per interior row (j, k).  Lanes 0-15 write the
// 128-byte line of the nodes -16..-1 (the halo -HM..-1 from n-HM..n-1, the
// row padding before it as zeros, so the line is written whole); lane 16
// node 0 (alone on its line: rewriting the nodes 1..15 with their own
// values to make that line whole measured slower, 166 against 102 us per
// launch at 256^3); lanes 32.. node n and the cells after it to the end of
// the line of node n+HM (the halo n+1..n+HM from 1..HM, then zeros).
__device__ __forceinline__ void qswap_x_body(const PtArgs& a, unsigned bid)
{
    const Geo& g = a.g;
    const int n0 = g.n[0];
    const unsigned t = bid * blockDim.x + threadIdx.x;
    const unsigned row = t / QSX_LANES, lane = t % QSX_LANES;
    const unsigned nj = (unsigned)g.n[1] + 1;
    if (row >= nj * ((unsigned)g.n[2] + 1)) return;
    const int k = (int)(row / nj), j = (int)(row - (unsigned)k * nj);
    int i;
    if (lane < 16) i = (int)lane - 16;
    else if (lane == 16) i = 0;
    else if (lane >= 32 && (int)lane - 32 < line_end_i(n0 + HM) - n0) i = n0 + (int)lane - 32;
    else return;
    const long gi = gidx(g, i, j, k);
    double qv[5];
    if (i == 0 || i == n0) {
        const long g0 = gidx(g, 0, j, k), gn = gidx(g, n0, j, k);
#pragma unroll
        for (int m = 0; m < 5; ++m) qv[m] = 0.5 * (a.q[m][g0] + a.q[m][gn]);
    } else if (i >= -HM && i <= n0 + HM) {
        const long gs = gidx(g, i < 0 ? i + n0 : i - n0, j, k);
#pragma unroll
        for (int m = 0; m < 5; ++m) qv[m] = a.q[m][gs];
    } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) a.q[m][gi] = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) a.prim[m][gi] = 0.0;
        return;
    }
    store_q(a, gi, qv);
}

// j part (AX = 1) or k part (AX = 2): threads over the nodes i = 0..n of the
// rows of each outer index o (k for the j part, j for the k part) and 2*HM+1
// slots: slot 0 the duplicate pair (the rows at 0 and n of axis AX), slots
// 1..HM the halo rows -1..-HM from n-1..n-HM, slots HM+1..2*HM the halo rows
// n+1..n+HM from 1..HM.
template <int AX>
__device__ __forceinline__ void qswap_jk_body(const PtArgs& a, unsigned bid)
{
    const Geo& g = a.g;
    const unsigned ni = (unsigned)g.n[0] + 1;
    const unsigned t = bid * blockDim.x + threadIdx.x;
    const unsigned r = t / ni;
    const int i = (int)(t - r * ni);
    const unsigned o = r / (2 * HM + 1);
    const int slot = (int)(r - o * (2 * HM + 1));
    const int na = g.n[AX], no = AX == 1 ? g.n[2] : g.n[1];
    if ((int)o > no) return;
    auto at = [&](int c) { return AX == 1 ? gidx(g, i, c, (int)o) : gidx(g, i, (int)o, c); };
    double qv[5];
    if (slot == 0) {
        const long g0 = at(0), gn = at(na);
#pragma unroll
        for (int m = 0; m < 5; ++m) qv[m] = 0.5 * (a.q[m][g0] + a.q[m][gn]);
        store_q(a, g0, qv);
        store_q(a, gn, qv);
        return;
    }
    const int h = slot <= HM ? -slot : na + slot - HM;  // halo row
    const long gs = at(h < 0 ? h + na : h - na);
#pragma unroll
    for (int m = 0; m < 5; ++m) qv[m] = a.q[m][gs];
    store_q(a, at(h), qv);
}

__global__ __launch_bounds__(256) void k_qswap_x(PtArgs a) { qswap_x_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void k_qswap_y(PtArgs a) { qswap_jk_body<1>(a, blockIdx.x); }
__global__ __launch_bounds__(256) void k_qswap_z(PtArgs a) { qswap_jk_body<2>(a, blockIdx.x); }

// threads of the i, j and k parts of a box
__host__ __device__ inline long qswap_threads(const Geo& g, int axis)
{
    const long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    if (axis == 0) return (long)QSX_LANES * (n1 + 1) * (n2 + 1);
    return (n0 + 1) * (2 * HM + 1) * (axis == 1 ? n2 + 1 : n1 + 1);
}

// qswap of a two-dimensional box (km = 0), single rank: along a periodic
// in-plane axis (src/parallel.F90:4877-4915 for i, 5019-5065 for j) the
// halo copy, the duplicate average q(0) = q(n) = (q(0)+q(n))/2 and q2fvar
// on the strips -hm..0 and n..n+hm, one thread per line of the plane k = 0.
__global__ void k_qswap2d_line(PtArgs a)
{
    const int d = a.axis, o = 1 - d;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > a.g.n[o]) return;
    const int n = a.g.n[d];
    const long h0 = gidx(a.g, d == 0 ? 0 : t, d == 0 ? t : 0, 0);
    const long sd = a.g.s[d];
    for (int m = 0; m < 5; ++m) {
        double* f = a.q[m] + h0;
        for (int r = 1; r <= HM; ++r) f[-r * sd] = f[(n - r) * sd];
        for (int r = 1; r <= HM; ++r) f[(n + r) * sd] = f[r * sd];
        const double avg = 0.5 * (f[0] + f[n * sd]);
        f[0] = avg;
        f[n * sd] = avg;
    }
    // the strips n..n+hm and -hm..0
    for (int r = 0; r <= HM; ++r)
        for (int e = 0; e < 2; ++e) {
            const long gi = h0 + (e ? -r : n + r) * sd;
            double qv[5];
            for (int m = 0; m < 5; ++m) qv[m] = a.q[m][gi];
            q2fvar(a, qv, gi);
        }
}

// The k part of qswap on a two-dimensional box, the ka = 0 branch
// (src/parallel.F90:5161-5200): q of the plane k = 0 is copied into every k
// plane -hm..hm and q2fvar refreshes all of them.  Of the copies only the
// pressure of the planes k = -1 and k = +1 is written: it is the one k-halo
// value a two-dimensional kernel reads (the pressure stencil of the Ducros
// sensor); the plane k = 0 gets q2fvar in full.
__global__ void k_qswap2d_plane(PtArgs a)
{
    const long ni = a.g.n[0] + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * (a.g.n[1] + 1)) return;
    const long gi = gidx(a.g, (int)(t % ni), (int)(t / ni), 0);
    double qv[5];
    for (int m = 0; m < 5; ++m) qv[m] = a.q[m][gi];
    q2fvar(a, qv, gi);
    const double p = a.prim[4][gi];
    a.prim[4][gi - a.g.s[2]] = p;
    a.prim[4][gi + a.g.s[2]] = p;
}

// s + e = a + b exactly (Knuth's TwoSum; plain IEEE additions)
__host__ __device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (ah, al) += (bh, bl) in double-double arithmetic
__host__ __device__ __forceinline__ void dd_add(double& ah, double& al, double bh, double bl)
{
    double s, e;
    two_sum(ah, bh, s, e);
    e += al + bl;
    ah = s + e;
    al = e - (ah - s);
}

// The per-node terms of kenergycal / enstophycal / diss_rate_cal
// (src/statistic.F90:870-1042): rho |u|^2, rho |omega|^2 and
// 2 mu (S:S - (div u)^2 / 3)
__device__ __forceinline__ void stats_terms(const PtArgs& a, long gi, double& ke, double& en,
                                            double& di)
{
    const double rho = a.prim[0][gi];
    const double u = a.prim[1][gi], v = a.prim[2][gi], w = a.prim[3][gi];
    ke = rho * (u * u + v * v + w * w);
    const double d11 = a.dvel[0][gi], d12 = a.dvel[1][gi], d13 = a.dvel[2][gi];
    const double d21 = a.dvel[3][gi], d22 = a.dvel[4][gi], d23 = a.dvel[5][gi];
    const double d31 = a.dvel[6][gi], d32 = a.dvel[7][gi], d33 = a.dvel[8][gi];
    const double o1 = d32 - d23, o2 = d13 - d31, o3 = d21 - d12;
    en = rho * (o1 * o1 + o2 * o2 + o3 * o3);
    const double miu = miucal(a, a.prim[5][gi]) / a.reynolds;
    const double s12 = 0.5 * (d12 + d21), s13 = 0.5 * (d13 + d31), s23 = 0.5 * (d23 + d32);
    const double div = d11 + d22 + d33;
    di = 2.0 * miu * (d11 * d11 + d22 * d22 + d33 * d33 +
                      2.0 * (s12 * s12 + s13 * s13 + s23 * s23) - (1.0 / 3.0) * (div * div));
}

// The block's sums of (ke, en, di) into part[6 * blockIdx.x ...] as
// double-double pairs: a butterfly of plain additions over each
// wavefront's 64 lanes (its rounding, ~64 ulp of a wavefront's sum, is far
// below what the per-step statistics are compared to), then the
// wavefronts' sums in double-double in LDS.  Every thread of the block
// takes part.
__device__ __forceinline__ void stats_block_sum(double ke, double en, double di, double* part)
{
    double h[3] = {ke, en, di}, l[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) h[c] += __shfl_xor(h[c], off, 64);
    __shared__ double ws[4][6];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) {
            ws[wv][2 * c] = h[c];
            ws[wv][2 * c + 1] = l[c];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x + 63) >> 6;
        for (int c = 0; c < 3; ++c) {
            double hh = ws[0][2 * c], ll = ws[0][2 * c + 1];
            for (int v = 1; v < nw; ++v) dd_add(hh, ll, ws[v][2 * c], ws[v][2 * c + 1]);
            part[6 * blockIdx.x + 2 * c] = hh;
            part[6 * blockIdx.x + 2 * c + 1] = ll;
        }
    }
}

// The inputs of G at one node, loaded once: every output store could alias
// an input array, so loads interleaved with the stores would be repeated.
struct NodeIn {
    double q[5], u[3], p, jc;
};

__device__ __forceinline__ NodeIn load_node(const PtArgs& a, long gi)
{
    NodeIn v;
#pragma unroll
    for (int m = 0; m < 5; ++m) v.q[m] = a.q[m][gi];
#pragma unroll
    for (int c = 0; c < 3; ++c) v.u[c] = a.prim[1 + c][gi];
    v.p = a.prim[4][gi];
    v.jc = a.mini ? 1.0 : (a.juni ? a.j0 : a.jac[gi]);
    return v;
}

// sigma and qflux of diffrsdcal6 (src/solver.F90:2434-2546, turbmode none)
// at one node from its gradients, temperature and velocity:
// sg = (s11, s12, s13, s22, s23, s33), qf(1:3)
__device__ __forceinline__ void sigma_of(const PtArgs& a, long gi, const double* uvw, double* sg,
                                         double* qf)
{
    double dv[9], dt[3];
#pragma unroll
    for (int c = 0; c < 9; ++c) dv[c] = a.dvel[c][gi];
#pragma unroll
    for (int c = 0; c < 3; ++c) dt[c] = a.dtmp[c][gi];
    const double miu = miucal(a, a.prim[5][gi]) / a.reynolds;
    const double d11 = dv[0], d12 = dv[1], d13 = dv[2];
    const double d21 = dv[3], d22 = dv[4], d23 = dv[5];
    const double d31 = dv[6], d32 = dv[7], d33 = dv[8];
    const double s11 = d11, s12 = 0.5 * (d12 + d21), s13 = 0.5 * (d13 + d31);
    const double s22 = d22, s23 = 0.5 * (d23 + d32), s33 = d33;
    const double skk = (1.0 / 3.0) * (s11 + s22 + s33);
    const double miu2 = 2.0 * miu;
    const double hcc = (miu / a.prandtl) / a.const5;
    const double g1 = miu2 * (s11 - skk), g2 = miu2 * s12, g3 = miu2 * s13;
    const double g4 = miu2 * (s22 - skk), g5 = miu2 * s23, g6 = miu2 * (s33 - skk);
    sg[0] = g1; sg[1] = g2; sg[2] = g3; sg[3] = g4; sg[4] = g5; sg[5] = g6;
    const double u = uvw[0], v = uvw[1], w = uvw[2];
    qf[0] = hcc * dt[0] + g1 * u + g2 * v + g3 * w;
    qf[1] = hcc * dt[1] + g2 * u + g4 * v + g5 * w;
    qf[2] = hcc * dt[2] + g3 * u + g5 * v + g6 * w;
}

// The flux of axis d that the flux sweep differentiates, G = Fv - F, at one
// node: inviscid fcs(:,1..5) of convrsdcal6 (src/solver.F90:2201-2207) with
// the off-diagonal metrics zero, F = J*(q_m*U + [m = d+1] dxi(d,d)*p),
// F5 = J*(q5+p)*U, and viscous ff(:,2..5) of diffrsdcal6 (src/solver.F90:
// 2628-2639), where the diagonal metric picks sigma(d,c) and qflux(d).
// The derivative is linear, so the reference's two passes (qrhs = -sum D F,
// then qrhs += sum D Fv) become one derivative of Fv - F per axis; only the
// summation order differs.
__device__ __forceinline__ void gflux_of(const PtArgs& a, int d, double xn, const NodeIn& n,
                                         const double* sg, const double* qf, double* out)
{
    static constexpr int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const double p = n.p;
    if (a.mini) {
        // tgvsolver.F90:578-631,760-843: fcs = q_m u_d (+ p on the momentum
        // of the sweep axis), (q5 + p) u_d; ff = sigma(d,c), qflux(d); no
        // metrics (the derivative is divided by the spacing)
        const double uu = n.u[d];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double qm = n.q[m];
            double fc;
            if (m == 4) fc = (qm + p) * uu;
            else if (m == d + 1) fc = qm * uu + p;
            else fc = qm * uu;
            if (m == 0 || !a.visc) out[m] = -fc;
            else out[m] = (m < 4 ? sg[sidx[d][m - 1]] : qf[d]) - fc;
        }
        return;
    }
    const double jc = n.jc;
    const double uu = xn * n.u[d];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double qm = n.q[m];
        double fi;
        if (m == 0) fi = jc * qm * uu;
        else if (m == 4) fi = jc * (qm + p) * uu;
        else if (m == d + 1) fi = jc * (qm * uu + xn * p);
        else fi = jc * (qm * uu);
        if (m == 0 || !a.visc) {
            out[m] = -fi;
        } else {
            const double sv = m < 4 ? sg[sidx[d][m - 1]] : qf[d];
            out[m] = (sv * xn) * jc - fi;
        }
    }
}

// gflux_of with the three metric components of axis d at the node (a grid
// with general metrics): U = dxi(d,:).u, the pressure on every momentum
// component, and sigma and the heat flux projected on dxi(d,:)
// (src/solver.F90:2239-2253, 2628-2639)
__device__ __forceinline__ void gflux_of_gen(const PtArgs& a, int d, long gm, const NodeIn& n,
                                             const double* sg, const double* qf, double* out)
{
    static constexpr int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const double x1 = a.dxf[3 * d][gm], x2 = a.dxf[3 * d + 1][gm], x3 = a.dxf[3 * d + 2][gm];
    const double p = n.p, jc = n.jc;
    const double uu = x1 * n.u[0] + x2 * n.u[1] + x3 * n.u[2];
    const double xe[3] = {x1, x2, x3};
    double fi[5];
    fi[0] = jc * n.q[0] * uu;
#pragma unroll
    for (int c = 0; c < 3; ++c) fi[1 + c] = jc * (n.q[1 + c] * uu + xe[c] * p);
    fi[4] = jc * (n.q[4] + p) * uu;
    out[0] = -fi[0];
    if (!a.visc) {
#pragma unroll
        for (int m = 1; m < 5; ++m) out[m] = -fi[m];
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[1 + c] = (sg[sidx[c][0]] * x1 + sg[sidx[c][1]] * x2 + sg[sidx[c][2]] * x3) * jc - fi[1 + c];
    out[4] = (qf[0] * x1 + qf[1] * x2 + qf[2] * x3) * jc - fi[4];
}

// G_d at a node of the halo strip of axis d (the nodes -hm..-1 and
// n+1..n+hm of a d-line, where the flux sweep also reads G_d): the inviscid
// part at the node itself, the viscous part (sigma, heat flux) that of its
// periodic image, as the halos of sigma and qflux are plain copies
// (src/solver.F90:2605-2607)
__device__ __forceinline__ void gflux_strip(const PtArgs& a, const int (&c)[3], int d, long gi)
{
    const NodeIn nd = load_node(a, gi);
    double sg[6] = {0, 0, 0, 0, 0, 0}, qf[3] = {0, 0, 0};
    if (a.visc) {
        int w[3] = {c[0], c[1], c[2]};
        w[d] = wrap_node(c[d], a.g.n[d]);
        const long gw = gidx(a.g, w[0], w[1], w[2]);
        const double uw[3] = {a.prim[1][gw], a.prim[2][gw], a.prim[3][gw]};
        sigma_of(a, gw, uw, sg, qf);
    }
    // a compile-time axis keeps sg and qf in registers
    double out[5];
    const double xn = a.xi[d][c[d] + HM];
    if (a.curv) gflux_of_gen(a, d, gi, nd, sg, qf, out);
    else if (d == 0) gflux_of(a, 0, xn, nd, sg, qf, out);
    else if (d == 1) gflux_of(a, 1, xn, nd, sg, qf, out);
    else gflux_of(a, 2, xn, nd, sg, qf, out);
#pragma unroll
    for (int m = 0; m < 5; ++m) st_nt(&a.gfl[5 * d + m][gi], out[m]);
}

__device__ __forceinline__ void gflux_zero(const PtArgs& a, int d, long gi)
{
#pragma unroll
    for (int m = 0; m < 5; ++m) st_nt(&a.gfl[5 * d + m][gi], 0.0);
}

// The halo node along an axis of extent n whose periodic image is node c
// (wrap_node's inverse on the strips): c + n for c in 1..HM, c - n for c in
// n-HM..n-1, else no halo node (NO_HALO).
constexpr int NO_HALO = -(1 << 20);
__device__ __forceinline__ int halo_of_image(int c, int n)
{
    return (c >= 1 && c <= HM) ? c + n : ((c >= n - HM && c <= n - 1) ? c - n : NO_HALO);
}

// k_gflux at cell i of the interior row (j, k); hj, hk are the y and z
// halo rows of which this row is the periodic image (or NO_HALO).
template <bool STATS>
__device__ __forceinline__ void gflux_cell(const PtArgs& a, int i, int j, int k, int hj, int hk,
                                           double (&st)[3])
{
    const int n0 = a.g.n[0];
    const int c[3] = {i, j, k};
    const long gi = gidx(a.g, c[0], c[1], c[2]);
    if (i >= 0 && i <= n0) {
        const NodeIn nd = load_node(a, gi);
        double sg[6] = {0, 0, 0, 0, 0, 0}, qf[3] = {0, 0, 0};
        if (a.visc) sigma_of(a, gi, nd.u, sg, qf);
        // rkfirst's statistics of this state (nodes 1..n), from the inputs
        // already loaded (the stored primitive rho aside), before any store
        if (STATS && i >= 1 && j >= 1 && k >= 1 && j <= a.g.n[1] && k <= a.g.n[2])
            stats_terms(a, gi, st[0], st[1], st[2]);
        double out[15];
        if (a.curv) {
#pragma unroll
            for (int e = 0; e < 3; ++e) gflux_of_gen(a, e, gi, nd, sg, qf, out + 5 * e);
        } else {
#pragma unroll
            for (int e = 0; e < 3; ++e) gflux_of(a, e, a.xi[e][c[e] + HM], nd, sg, qf, out + 5 * e);
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) st_nt(&a.gfl[k][gi], out[k]);
        if (hj != NO_HALO) {
            const long gh = gidx(a.g, i, hj, c[2]);
            const NodeIn nh = load_node(a, gh);
            if (a.curv) gflux_of_gen(a, 1, gh, nh, sg, qf, out);
            else gflux_of(a, 1, a.xi[1][hj + HM], nh, sg, qf, out);
#pragma unroll
            for (int m = 0; m < 5; ++m) st_nt(&a.gfl[5 + m][gh], out[m]);
        }
        if (hk != NO_HALO) {
            const long gh = gidx(a.g, i, c[1], hk);
            const NodeIn nh = load_node(a, gh);
            if (a.curv) gflux_of_gen(a, 2, gh, nh, sg, qf, out);
            else gflux_of(a, 2, a.xi[2][hk + HM], nh, sg, qf, out);
#pragma unroll
            for (int m = 0; m < 5; ++m) st_nt(&a.gfl[10 + m][gh], out[m]);
        }
        return;
    }
    if (a.per[0] && i >= -HM && i <= n0 + HM) gflux_strip(a, c, 0, gi);
    else gflux_zero(a, 0, gi);
    if (i > n0) {
        gflux_zero(a, 1, gi);
        gflux_zero(a, 2, gi);
        if (hj != NO_HALO) gflux_zero(a, 1, gidx(a.g, i, hj, c[2]));
        if (hk != NO_HALO) gflux_zero(a, 2, gidx(a.g, i, c[1], hk));
    }
}

// One thread per cell of the interior i-rows (j, k in 0..n), i from -16
// (the start of the line of node -HM) to line_end_i(n + HM) (the end of the
// line of node n + HM, which is the row pitch), the rows laid end to end (a
// whole number of 128-byte lines each, so a wavefront covers whole lines
// only).  A row gives G of all axes on its interior nodes and G_x on its x
// halo strips.  A row that is the periodic image of a y (z) halo row also
// gives that row's G_y (G_z): the viscous part is its own sigma and heat
// flux, already in registers, the inviscid part that of the halo node.  The
// threads of cells that no flux sweep reads (row padding past the strips,
// i past n in halo rows) store zeros, so each field's lines are written
// whole.
//
// STATS (stage 1 of a step): the workgroup also sums the statistics terms
// of its nodes in 1..n into a.part (double-double, 6 per workgroup), which
// k_stats_fin reduces; the inputs are those of G, so the statistics cost
// one more load per node (the primitive rho) instead of a pass of their own.
template <bool STATS>
__device__ __forceinline__ void gflux_body(const PtArgs& a, unsigned bid)
{
    const int iend = line_end_i(a.g.n[0] + HM);
    const unsigned rl = (unsigned)(iend + 16);  // cells per row
    const unsigned t = bid * blockDim.x + threadIdx.x;
    const unsigned row = t / rl;
    const unsigned nj = (unsigned)a.g.n[1] + 1;
    double st[3] = {0.0, 0.0, 0.0};
    if (row < nj * (unsigned)(a.g.n[2] + 1)) {
        const unsigned kk = row / nj;
        const int j = (int)(row - kk * nj), k = (int)kk;
        gflux_cell<STATS>(a, (int)(t - row * rl) - 16, j, k,
                          a.per[1] ? halo_of_image(j, a.g.n[1]) : NO_HALO,
                          a.per[2] ? halo_of_image(k, a.g.n[2]) : NO_HALO, st);
    }
    if (STATS) stats_block_sum(st[0], st[1], st[2], a.part);
}

template <bool STATS>
__global__ __launch_bounds__(256) void k_gflux(PtArgs a) { gflux_body<STATS>(a, blockIdx.x); }

// bchomo of the mini-app (tgvsolver.F90:271-369): the primitives of the
// halo faces copied from their periodic images (x, y and z faces, each over
// the interior extent of the other axes, so every source is interior) and q
// rebuilt there with fvar2q from the pressure (tgvsolver.F90:45-66).
__global__ void k_bchomo(PtArgs a)
{
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1, nk = a.g.n[2] + 1;
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long s0 = 2L * HM * nj * nk, s1 = 2L * HM * ni * nk, s2 = 2L * HM * ni * nj;
    int d;
    if (r < s0) d = 0;
    else if ((r -= s0) < s1) d = 1;
    else if ((r -= s1) < s2) d = 2;
    else return;
    const long ext[3] = {ni, nj, nk};
    const int h = (int)(r % (2 * HM));
    r /= 2 * HM;
    const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
    int c[3], w[3];
    c[oa] = w[oa] = (int)(r % ext[oa]);
    c[ob] = w[ob] = (int)(r / ext[oa]);
    const int n = a.g.n[d];
    c[d] = h < HM ? -(h + 1) : n + (h - HM + 1);
    w[d] = h < HM ? n - (h + 1) : (h - HM + 1);
    const long gd = gidx(a.g, c[0], c[1], c[2]), gs = gidx(a.g, w[0], w[1], w[2]);
    double pr[6];
    for (int k = 0; k < 6; ++k) {
        pr[k] = a.prim[k][gs];
        a.prim[k][gd] = pr[k];
    }
    const double rho = pr[0], u = pr[1], v = pr[2], ww = pr[3];
    a.q[0][gd] = rho;
    a.q[1][gd] = rho * u;
    a.q[2][gd] = rho * v;
    a.q[3][gd] = rho * ww;
    const double var1 = 0.5 * (u * u + v * v + ww * ww);
    a.q[4][gd] = pr[4] * a.const6 + rho * var1;
}

// RK3 combine (src/mainloop.F90:423-445) and updatefvar (src/fludyna.F90:
// 191-242) on the interior, in the reference's operation order (no FMA
// contraction).
#pragma clang fp contract(off)
__global__ void k_rk(PtArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ntot = (long)(a.g.n[0] + 1) * (a.g.n[1] + 1) * (a.g.n[2] + 1);
    if (p >= ntot) return;
    int i, j, k;
    interior_ijk(a.g, p, i, j, k);
    const long gi = gidx(a.g, i, j, k);
    const double jc = a.jac[gi];
    // every operand is read before the first store (which could alias it)
    double qo[5], qs[5], qr[5], qv[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        qo[m] = a.q[m][gi];
        qr[m] = a.qrhs[m][gi];
        qs[m] = a.stage == 1 ? qo[m] * jc : a.qsave[m][gi];
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        if (a.stage == 1) st_nt(&a.qsave[m][gi], qs[m]);
        const double vn = a.rka * qs[m] + a.rkb * qo[m] * jc + a.rkc * qr[m] * a.dt;
        qv[m] = vn / jc;
        st_nt(&a.q[m][gi], qv[m]);
    }
    q2fvar(a, qv, gi);
}
#pragma clang fp contract(fast)

// updatefvar on the nodes 0..n (src/fludyna.F90:191-242), over row_node's
// rows; the threads past node n up to the end of its 128-byte line store
// the primitives already there, so every line is written whole.
__device__ __forceinline__ void prim_body(const PtArgs& a, unsigned bid)
{
    int i, j, k;
    if (!row_node(a.g, bid, i, j, k)) return;
    const long gi = gidx(a.g, i, j, k);
    if (i <= a.g.n[0]) {
        double qv[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) qv[m] = a.q[m][gi];
        q2fvar(a, qv, gi, !a.skip_rho);
    } else {
#pragma unroll
        for (int m = 0; m < 6; ++m)
            if (m > 0 || !a.skip_rho) keep_cell(a.prim[m], gi);
    }
}

__global__ __launch_bounds__(256) void k_prim(PtArgs a) { prim_body(a, blockIdx.x); }

// gradcal's metric combination on a grid with general metrics
// (src/comsolver.F90:304-318, 344-358, 384-398): the sweeps left the
// index-space derivative of u_m (and T) along axis d in dvel(m, d) (dtmp(d));
// dvel(m, e) = sum over d, in the order 1, 2, 3, of that derivative times
// dxi(d, e).  Over row_node's rows (the cells past node n are dead).
struct MetricArgs {
    Geo g;
    double* dv[12];  // dvel(0..8), dtmp(0..2)
    const double* dxf[9];
};

__global__ __launch_bounds__(256) void k_metric_grad(MetricArgs a)
{
    int i, j, k;
    if (!row_node(a.g, i, j, k)) return;
    const long gi = gidx(a.g, i, j, k);
    double x[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) x[m] = a.dxf[m][gi];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = a.dv[3 * q + d][gi];
#pragma unroll
        for (int e = 0; e < 3; ++e)
            a.dv[3 * q + e][gi] = 0.0 + r[0] * x[e] + r[1] * x[3 + e] + r[2] * x[6 + e];
    }
}

// Partial sums of kenergycal / enstophycal / diss_rate_cal over 1..n
// (src/statistic.F90:870-1042), one double-double (hi, lo) pair per
// statistic and workgroup: the sums are then exact to about 1e-30 relative
// of each term, so two statistics of nearby states differ by what the
// states do, not by summation order (central differences of the enstrophy
// at eps 1e-6 rely on that).
__global__ __launch_bounds__(256) void k_stats(PtArgs a)
{
    __shared__ double sh[3][256], sl[3][256];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ntot = (long)a.g.n[0] * a.g.n[1] * a.g.n[2];
    double ke = 0.0, en = 0.0, di = 0.0;
    if (p < ntot) {
        const long n0 = a.g.n[0], n1 = a.g.n[1];
        const int i = 1 + (int)(p % n0);
        const long r = p / n0;
        const int j = 1 + (int)(r % n1);
        const int k = 1 + (int)(r / n1);
        stats_terms(a, gidx(a.g, i, j, k), ke, en, di);
    }
    sh[0][threadIdx.x] = ke;
    sh[1][threadIdx.x] = en;
    sh[2][threadIdx.x] = di;
    for (int c = 0; c < 3; ++c) sl[c][threadIdx.x] = 0.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) {
                double h = sh[c][threadIdx.x], l = sl[c][threadIdx.x];
                dd_add(h, l, sh[c][threadIdx.x + s], sl[c][threadIdx.x + s]);
                sh[c][threadIdx.x] = h;
                sl[c][threadIdx.x] = l;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            a.part[6 * blockIdx.x + 2 * c] = sh[c][0];
            a.part[6 * blockIdx.x + 2 * c + 1] = sl[c][0];
        }
}

// Workgroup partials of the statistics (double-double pairs, 6 per
// partial) summed in double-double: workgroup w of the launch takes the
// partials [w*chunk, (w+1)*chunk) of nb, each thread a strided share, then
// a tree over the threads.  With out3 null the workgroup's sum goes to
// part2 (6 doubles per workgroup: the first level over every k_stats or
// k_gflux workgroup); otherwise (one workgroup) the means kenergycal /
// enstophycal / diss_rate_cal report go to out3 on the device: ke and
// enstrophy 1/2 of the mean, dissipation the mean (roinf = uinf = l_0 = 1),
// so a step's statistics need no host round trip.
__global__ __launch_bounds__(256) void k_stats_fin(const double* __restrict__ part, long nb,
                                                  long chunk, double* __restrict__ part2,
                                                  double n3, double* __restrict__ out3)
{
    __shared__ double sh[3][256], sl[3][256];
    double h[3] = {0.0, 0.0, 0.0}, l[3] = {0.0, 0.0, 0.0};
    const long b0 = (long)blockIdx.x * chunk, b1 = min(nb, b0 + chunk);
    for (long b = b0 + threadIdx.x; b < b1; b += 256)
        for (int c = 0; c < 3; ++c) dd_add(h[c], l[c], part[6 * b + 2 * c], part[6 * b + 2 * c + 1]);
    for (int c = 0; c < 3; ++c) {
        sh[c][threadIdx.x] = h[c];
        sl[c][threadIdx.x] = l[c];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) {
                double hh = sh[c][threadIdx.x], ll = sl[c][threadIdx.x];
                dd_add(hh, ll, sh[c][threadIdx.x + s], sl[c][threadIdx.x + s]);
                sh[c][threadIdx.x] = hh;
                sl[c][threadIdx.x] = ll;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (out3) {
            out3[0] = 0.5 * sh[0][0] / n3;
            out3[1] = 0.5 * sh[1][0] / n3;
            out3[2] = sh[2][0] / n3;
        } else {
            for (int c = 0; c < 3; ++c) {
                part2[6 * blockIdx.x + 2 * c] = sh[c][0];
                part2[6 * blockIdx.x + 2 * c + 1] = sl[c][0];
            }
        }
    }
}

// Statistics with a fixed number of workgroups (the allocation-free entry
// point, astr_stats_device): workgroup b sums, in double-double, the nodes
// 1..n taken grid-stride from its own start, then its threads in an LDS
// tree; part[6 b ...] receives the (hi, lo) pairs.  The node index is
// 32-bit (the entry point refuses boxes of 2^31 nodes or more).
__global__ __launch_bounds__(256) void k_stats_gs(PtArgs a, double* __restrict__ part)
{
    __shared__ double sh[3][256], sl[3][256];
    const unsigned n0 = (unsigned)a.g.n[0], n1 = (unsigned)a.g.n[1];
    const unsigned ntot = n0 * n1 * (unsigned)a.g.n[2];
    double h[3] = {0.0, 0.0, 0.0}, l[3] = {0.0, 0.0, 0.0};
    for (unsigned p = blockIdx.x * blockDim.x + threadIdx.x; p < ntot; p += gridDim.x * blockDim.x) {
        const unsigned r = p / n0;
        const int i = 1 + (int)(p - r * n0);
        const int j = 1 + (int)(r % n1);
        const int k = 1 + (int)(r / n1);
        double t[3];
        stats_terms(a, gidx(a.g, i, j, k), t[0], t[1], t[2]);
        for (int c = 0; c < 3; ++c) dd_add(h[c], l[c], t[c], 0.0);
    }
    for (int c = 0; c < 3; ++c) {
        sh[c][threadIdx.x] = h[c];
        sl[c][threadIdx.x] = l[c];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) {
                double hh = sh[c][threadIdx.x], ll = sl[c][threadIdx.x];
                dd_add(hh, ll, sh[c][threadIdx.x + s], sl[c][threadIdx.x + s]);
                sh[c][threadIdx.x] = hh;
                sl[c][threadIdx.x] = ll;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            part[6 * blockIdx.x + 2 * c] = sh[c][0];
            part[6 * blockIdx.x + 2 * c + 1] = sl[c][0];
        }
}

// maxmincal's max_q (src/statistic.F90:737-763): the largest value of each
// q(m) over the nodes 0..n; each workgroup strides over the nodes and
// leaves 5 partial maxima (max is exact, so the order does not matter)
__global__ __launch_bounds__(256) void k_qmax(PtArgs a)
{
    __shared__ double sh[5][256];
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    const long ntot = ni * nj * (a.g.n[2] + 1);
    double mx[5] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ntot;
         p += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(p, ni, nj, i, j, k);
        const long gi = gidx(a.g, i, j, k);
#pragma unroll
        for (int m = 0; m < 5; ++m) mx[m] = fmax(mx[m], a.q[m][gi]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) sh[m][threadIdx.x] = mx[m];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < 5; ++m)
                sh[m][threadIdx.x] = fmax(sh[m][threadIdx.x], sh[m][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < 5; ++m) a.part[5 * blockIdx.x + m] = sh[m][0];
}

// ------------------------------------------------------ discrete adjoint
//
// Reverse-mode kernels of one stage of the central path (filterq, qswap,
// gradcal, the axis fluxes G, their derivatives, the RK3 combine and
// updatefvar).  x-bar is the derivative of the objective with respect to
// x.  The forward values each kernel needs (the state after qswap, the RK3
// output, the gradients) come from the stage's stored or recomputed fields.

struct AdjArgs {
    Geo g;
    const double* q[5];     // forward q after qswap (or of the state)
    const double* prim[6];  // forward primitives after qswap
    const double* qn[5];    // the stage's RK3 output
    const double* dvel[9];
    const double* dtmp[3];
    const double* jac;
    const double* xi[3];
    double* qb[5];          // q-bar
    double* pb[6];          // primitive-bar (rho, u, v, w, p, T)
    double* db[9];          // dvel-bar
    double* tb[3];          // dtmp-bar
    double* gb[15];         // G-bar, 5d+m
    double* rb[5];          // qrhs-bar
    double* sb[5];          // qsave-bar
    double const2, const5, const6, reynolds, prandtl, tempconst, tempconst1;
    double rka, rkb, rkc, dt;
    int stage, axis, visc;
    double scale;           // 1/(n0 n1 n2) of the statistic
};

// q-bar += (d q2fvar / d q)^T primitive-bar at one node, on registers: q is
// the state the primitives were computed from, pbv their adjoint
__device__ __forceinline__ void q2fvar_adj(const AdjArgs& a, const double* q, const double* pbv,
                                           double* qbv)
{
    const double rho = q[0];
    const double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
    const double uu = u * u + v * v + w * w;
    const double p = (q[4] - 0.5 * rho * uu) / a.const6;
    double rb = pbv[0], ub = pbv[1], vb = pbv[2], wb = pbv[3];
    double pbar = pbv[4];
    const double tbv = pbv[5];
    // T = p/rho*const2
    pbar += tbv * a.const2 / rho;
    rb -= tbv * p * a.const2 / (rho * rho);
    // p = (q5 - rho |u|^2/2)/const6
    const double q4b = pbar / a.const6;
    rb -= pbar * 0.5 * uu / a.const6;
    ub -= pbar * rho * u / a.const6;
    vb -= pbar * rho * v / a.const6;
    wb -= pbar * rho * w / a.const6;
    // u = q2/rho, ...
    qbv[1] += ub / rho;
    qbv[2] += vb / rho;
    qbv[3] += wb / rho;
    rb -= (ub * u + vb * v + wb * w) / rho;
    qbv[0] += rb;
    qbv[4] += q4b;
}

// qswap's refresh of the primitives reversed at the halo cells (one axis
// outside 0..n) and the seam nodes (interior, some coordinate 0 or n), at
// q after qswap: q-bar collects the primitive-bar there, which is zeroed
// (the forward overwrote those primitives).  One thread per node of the
// padded box (x over the (i,j) plane, y over k).
__global__ __launch_bounds__(256) void k_q2f_adj(AdjArgs a)
{
    const unsigned pw = (unsigned)(a.g.n[0] + 1 + 2 * HM);
    const unsigned ph = (unsigned)(a.g.n[1] + 1 + 2 * HM);
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pw * ph) return;
    const unsigned jj = t / pw;
    const int c[3] = {(int)(t - jj * pw) - HM, (int)jj - HM, (int)blockIdx.y - HM};
    int nout = 0;
    bool seam = false;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (c[e] < 0 || c[e] > a.g.n[e]) ++nout;
        else if (c[e] == 0 || c[e] == a.g.n[e]) seam = true;
    }
    if (nout > 1 || (nout == 0 && !seam)) return;
    const long gi = gidx(a.g, c[0], c[1], c[2]);
    // every operand before the first store (a store could alias a later load)
    double q[5], pbv[6], qbv[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        q[m] = a.q[m][gi];
        qbv[m] = a.qb[m][gi];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) pbv[k] = a.pb[k][gi];
    q2fvar_adj(a, q, pbv, qbv);
#pragma unroll
    for (int m = 0; m < 5; ++m) a.qb[m][gi] = qbv[m];
#pragma unroll
    for (int k = 0; k < 6; ++k) a.pb[k][gi] = 0.0;
}

// updatefvar and the RK3 combine reversed on the interior, one node per
// thread.  The primitive-bar of the stage output qn is spent into q-bar
// (it is not zeroed: k_gflux_adj_int overwrites the interior primitive-bar
// next); q-bar, then the adjoint of q' = (a qsave + b q J + c qrhs dt)/J,
// becomes the adjoint of q after qswap and qrhs-bar is set.  qsave-bar
// collects a/J q'-bar over stages 3 and 2 and is spent at stage 1, where
// qsave = q J; stage 3 (the first reversed stage of a step) neither reads
// it nor needs it cleared.
__global__ __launch_bounds__(256) void k_update_adj(AdjArgs a)
{
    int ci, cj, ck;
    if (!row_node(a.g, ci, cj, ck)) return;
    const long gi = gidx(a.g, ci, cj, ck);
    if (ci > a.g.n[0]) {
        // the rest of the line holding node n: q-bar keeps its values; the
        // qrhs-bar and qsave-bar cells there are dead
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            keep_cell(a.qb[m], gi);
            st_nt(&a.rb[m][gi], 0.0);
            if (a.stage != 1) st_nt(&a.sb[m][gi], 0.0);
        }
        return;
    }
    const double jc = a.jac[gi];
    double qn[5], pbv[6], ob[5], sv[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        qn[m] = a.qn[m][gi];
        ob[m] = a.qb[m][gi];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) pbv[k] = a.pb[k][gi];
    if (a.stage != 3) {
#pragma unroll
        for (int m = 0; m < 5; ++m) sv[m] = a.sb[m][gi];
    }
    q2fvar_adj(a, qn, pbv, ob);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        st_nt(&a.rb[m][gi], a.rkc * a.dt * ob[m] / jc);
        double qv = a.rkb * ob[m];
        if (a.stage == 1) qv += a.rka * ob[m] + jc * sv[m];
        else st_nt(&a.sb[m][gi], sv[m] + a.rka * ob[m] / jc);
        st_nt(&a.qb[m][gi], qv);
    }
}

// The inviscid part of G_d (convrsdcal6's fcs) reversed at one node, on
// register copies: qbv and pbv (u_d, p) collect -dF/d(.) G-bar; q, p, u_d
// are the forward values at the node
__device__ __forceinline__ void inviscid_adj(int d, double xn, double jc, const double* qv,
                                             double p, double ud, const double* gbv, double* qbv,
                                             double* pbv)
{
    const double uu = xn * ud;
    double ubar = 0.0, pbar = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double g = -gbv[m];  // G = -F + (viscous)
        const double qm = qv[m];
        qbv[m] += g * jc * uu;
        if (m == 4) {
            ubar += g * jc * (qm + p);
            pbar += g * jc * uu;
        } else {
            ubar += g * jc * qm;
            if (m == d + 1) pbar += g * jc * xn;
        }
    }
    pbv[1 + d] += ubar * xn;
    pbv[4] += pbar;
}

// G on the halo strips of every axis d: the inviscid part at the halo node
// itself (its viscous part, taken at the periodic image, is reversed by
// k_gflux_adj_int).  One thread per halo node, the x strips first, then y,
// then z (disjoint node sets), i fastest within a row.
__host__ __device__ inline long gflux_adj_halo_count(const Geo& g, int d)
{
    const long e0 = g.n[0] + 1L, e1 = g.n[1] + 1L, e2 = g.n[2] + 1L;
    return 2L * HM * (d == 0 ? e1 * e2 : d == 1 ? e0 * e2 : e0 * e1);
}

// the halo node r of axis D's strips (a compile-time axis keeps every
// index and array in registers)
template <int D>
__device__ __forceinline__ void gflux_adj_halo_node(const AdjArgs& a, long r)
{
    constexpr int OA = D == 0 ? 1 : 0, OB = D == 2 ? 1 : 2;
    const long ea = a.g.n[OA] + 1L, eb = a.g.n[OB] + 1L;
    int h, ca, cb;
    if (D == 0) {
        // x strips: the 2 HM halo nodes of a row are the contiguous ones
        h = (int)(r % (2 * HM));
        r /= 2 * HM;
        ca = (int)(r % ea);
        cb = (int)(r / ea);
    } else {
        // y/z strips: consecutive i
        ca = (int)(r % ea);
        r /= ea;
        cb = (int)(r % eb);
        h = (int)(r / eb);
    }
    const int cd = h < HM ? -(h + 1) : a.g.n[D] + (h - HM + 1);
    const int ci = D == 0 ? cd : ca, cj = D == 1 ? cd : (D == 0 ? ca : cb), ck = D == 2 ? cd : cb;
    const long gi = gidx(a.g, ci, cj, ck);
    double gbv[5], qv[5], qbv[5], pbv[6];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        gbv[m] = a.gb[5 * D + m][gi];
        qv[m] = a.q[m][gi];
        qbv[m] = a.qb[m][gi];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) pbv[k] = a.pb[k][gi];
    const double xn = a.xi[D][cd + HM], jc = a.jac[gi], pf = a.prim[4][gi];
    const double ud = a.prim[1 + D][gi];
    inviscid_adj(D, xn, jc, qv, pf, ud, gbv, qbv, pbv);
#pragma unroll
    for (int m = 0; m < 5; ++m) a.qb[m][gi] = qbv[m];
    a.pb[1 + D][gi] = pbv[1 + D];
    a.pb[4][gi] = pbv[4];
}

__global__ void k_gflux_adj_halo(AdjArgs a)
{
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n0 = gflux_adj_halo_count(a.g, 0), n1 = gflux_adj_halo_count(a.g, 1);
    if (r < n0) gflux_adj_halo_node<0>(a, r);
    else if ((r -= n0) < n1) gflux_adj_halo_node<1>(a, r);
    else if ((r -= n1) < gflux_adj_halo_count(a.g, 2)) gflux_adj_halo_node<2>(a, r);
}

// G on the interior reversed at one node: the inviscid parts of its three
// axis fluxes, and sigma and the heat flux, whose adjoint gathers the
// node's own viscous terms and those of the halo nodes that took it as
// their periodic image; then sigma -> dvel, T (through miucal), u and the
// heat flux -> dtmp.
__global__ __launch_bounds__(256) void k_gflux_adj_int(AdjArgs a)
{
    static constexpr int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    int c[3];
    if (!row_node(a.g, c[0], c[1], c[2])) return;
    const long gi = gidx(a.g, c[0], c[1], c[2]);
    if (c[0] > a.g.n[0]) {
        // the rest of the line holding node n: q-bar and primitive-bar keep
        // their values; the dvel-bar and dtmp-bar cells there are dead
#pragma unroll
        for (int m = 0; m < 5; ++m) keep_cell(a.qb[m], gi);
#pragma unroll
        for (int m = 0; m < 6; ++m) keep_cell(a.pb[m], gi);
        if (a.visc) {
#pragma unroll
            for (int m = 0; m < 9; ++m) st_nt(&a.db[m][gi], 0.0);
#pragma unroll
            for (int m = 0; m < 3; ++m) st_nt(&a.tb[m][gi], 0.0);
        }
        return;
    }
    const double jc = a.jac[gi];
    // every operand first (a store could alias a later load); primitive-bar
    // is zero on entry (spent by updatefvar's reversal)
    double qv[5], qbv[5], pbv[6] = {0, 0, 0, 0, 0, 0}, gball[15];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        qv[m] = a.q[m][gi];
        qbv[m] = a.qb[m][gi];
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) gball[k] = a.gb[k][gi];
    const double pf = a.prim[4][gi];
    const double uvw[3] = {a.prim[1][gi], a.prim[2][gi], a.prim[3][gi]};
    double sgb[6] = {0, 0, 0, 0, 0, 0}, qfb[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double xn = a.xi[e][c[e] + HM];
        const double* gbv = gball + 5 * e;
        inviscid_adj(e, xn, jc, qv, pf, uvw[e], gbv, qbv, pbv);
        if (a.visc) {
#pragma unroll
            for (int k = 0; k < 3; ++k) sgb[sidx[e][k]] += gbv[1 + k] * xn * jc;
            qfb[e] += gbv[4] * xn * jc;
            // the halo node of this axis whose image is this node
            const int n = a.g.n[e];
            int hc = 1 << 30;
            if (c[e] >= 1 && c[e] <= HM) hc = n + c[e];
            else if (c[e] >= n - HM && c[e] <= n - 1) hc = c[e] - n;
            if (hc != (1 << 30)) {
                int w[3] = {c[0], c[1], c[2]};
                w[e] = hc;
                const long gh = gidx(a.g, w[0], w[1], w[2]);
                const double xh = a.xi[e][hc + HM], jh = a.jac[gh];
#pragma unroll
                for (int k = 0; k < 3; ++k) sgb[sidx[e][k]] += a.gb[5 * e + 1 + k][gh] * xh * jh;
                qfb[e] += a.gb[5 * e + 4][gh] * xh * jh;
            }
        }
    }
    if (!a.visc) {
#pragma unroll
        for (int m = 0; m < 5; ++m) st_nt(&a.qb[m][gi], qbv[m]);
#pragma unroll
        for (int k = 0; k < 6; ++k) st_nt(&a.pb[k][gi], pbv[k]);
        return;
    }
    // forward sigma and heat flux at this node
    double dv[9], dt[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) dv[k] = a.dvel[k][gi];
#pragma unroll
    for (int k = 0; k < 3; ++k) dt[k] = a.dtmp[k][gi];
    const double T = a.prim[5][gi];
    const double st = sqrt(T);
    const double mu0 = T * st * a.tempconst1 / (T + a.tempconst);
    const double miu = mu0 / a.reynolds;
    const double s11 = dv[0], s22 = dv[4], s33 = dv[8];
    const double s12 = 0.5 * (dv[1] + dv[3]), s13 = 0.5 * (dv[2] + dv[6]), s23 = 0.5 * (dv[5] + dv[7]);
    const double skk = (1.0 / 3.0) * (s11 + s22 + s33);
    const double miu2 = 2.0 * miu;
    const double hcc = (miu / a.prandtl) / a.const5;
    const double g[6] = {miu2 * (s11 - skk), miu2 * s12, miu2 * s13, miu2 * (s22 - skk), miu2 * s23,
                         miu2 * (s33 - skk)};
    const double* uv = uvw;
    // qf_d = hcc dtmp_d + sum_c sigma(d,c) u_c
    double hccb = 0.0, ubv[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        st_nt(&a.tb[d][gi], hcc * qfb[d]);
        hccb += qfb[d] * dt[d];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sgb[sidx[d][k]] += qfb[d] * uv[k];
            ubv[k] += qfb[d] * g[sidx[d][k]];
        }
    }
    for (int k = 0; k < 3; ++k) pbv[1 + k] += ubv[k];
    // sigma = 2 mu (S - skk I)
    const double miu2b = sgb[0] * (s11 - skk) + sgb[1] * s12 + sgb[2] * s13 +
                         sgb[3] * (s22 - skk) + sgb[4] * s23 + sgb[5] * (s33 - skk);
    const double skkb = -miu2 * (sgb[0] + sgb[3] + sgb[5]);
    const double s11b = miu2 * sgb[0] + skkb / 3.0;
    const double s22b = miu2 * sgb[3] + skkb / 3.0;
    const double s33b = miu2 * sgb[5] + skkb / 3.0;
    const double s12b = miu2 * sgb[1], s13b = miu2 * sgb[2], s23b = miu2 * sgb[4];
    // dvel-bar and dtmp-bar are zero on entry (spent by the previous
    // gradient reversal): stored, not accumulated
    st_nt(&a.db[0][gi], s11b);
    st_nt(&a.db[4][gi], s22b);
    st_nt(&a.db[8][gi], s33b);
    st_nt(&a.db[1][gi], 0.5 * s12b);
    st_nt(&a.db[3][gi], 0.5 * s12b);
    st_nt(&a.db[2][gi], 0.5 * s13b);
    st_nt(&a.db[6][gi], 0.5 * s13b);
    st_nt(&a.db[5][gi], 0.5 * s23b);
    st_nt(&a.db[7][gi], 0.5 * s23b);
    // miu = miucal(T)/Re, hcc = miu/Pr/const5
    const double miub = 2.0 * miu2b + hccb / a.prandtl / a.const5;
    const double dmu = a.tempconst1 * st * (0.5 * T + 1.5 * a.tempconst) /
                       ((T + a.tempconst) * (T + a.tempconst));
    pbv[5] += miub * dmu / a.reynolds;
#pragma unroll
    for (int m = 0; m < 5; ++m) st_nt(&a.qb[m][gi], qbv[m]);
#pragma unroll
    for (int k = 0; k < 6; ++k) st_nt(&a.pb[k][gi], pbv[k]);
}

// qswap's copies and averages along a.axis reversed, for the lines of that
// axis over 0..n of the other two: the average q(0) = q(n) = (q(0)+q(n))/2,
// and the halo copies q(-k) = q(n-k), q(n+k) = q(k).  The eleven pieces
// touch different nodes (the copies' sources exclude 0 and n), so each is
// its own thread: op 0 the average, op 1..HM the low halos, HM+1..2HM the
// high ones.  Run for the axes in the order z, y, x.
__global__ void k_qswap_adj_line(AdjArgs a)
{
    const int d = a.axis;
    const long ext[3] = {a.g.n[0] + 1, a.g.n[1] + 1, a.g.n[2] + 1};
    const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
    const long nl = ext[oa] * ext[ob];
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (2 * HM + 1) * 5 * nl) return;
    int op, m;
    if (d == 0) {
        // x lines: the pieces of one row are neighbours (the same lines)
        op = (int)(r % (2 * HM + 1));
        r /= 2 * HM + 1;
        m = (int)(r % 5);
        r /= 5;
    } else {
        // y and z lines: neighbouring threads take neighbouring i
        op = (int)(r / (5 * nl));
        r -= op * 5 * nl;
        m = (int)(r / nl);
        r -= m * nl;
    }
    int c[3];
    c[oa] = (int)(r % ext[oa]);
    c[ob] = (int)(r / ext[oa]);
    c[d] = 0;
    const long sd = a.g.s[d];
    const int n = a.g.n[d];
    double* f = a.qb[m] + gidx(a.g, c[0], c[1], c[2]);
    if (op == 0) {
        const double t = f[0] + f[n * sd];
        f[0] = 0.5 * t;
        f[n * sd] = 0.5 * t;
    } else if (op <= HM) {
        const int k = op;
        f[(n - k) * sd] += f[-k * sd];
        f[-k * sd] = 0.0;
    } else {
        const int k = op - HM;
        f[k * sd] += f[(n + k) * sd];
        f[(n + k) * sd] = 0.0;
    }
}

// The enstrophy statistic (k_stats, 1/2 sum over nodes 1..n of rho |omega|^2
// times scale) reversed: rho-bar and dvel-bar at the nodes 1..n
__global__ void k_ens_adj(AdjArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n0 = a.g.n[0], n1 = a.g.n[1];
    if (p >= n0 * n1 * a.g.n[2]) return;
    const int i = 1 + (int)(p % n0);
    const long r = p / n0;
    const int j = 1 + (int)(r % n1), k = 1 + (int)(r / n1);
    const long gi = gidx(a.g, i, j, k);
    const double rho = a.prim[0][gi];
    const double d12 = a.dvel[1][gi], d13 = a.dvel[2][gi], d21 = a.dvel[3][gi];
    const double d23 = a.dvel[5][gi], d31 = a.dvel[6][gi], d32 = a.dvel[7][gi];
    const double o1 = d32 - d23, o2 = d13 - d31, o3 = d21 - d12;
    a.pb[0][gi] += 0.5 * a.scale * (o1 * o1 + o2 * o2 + o3 * o3);
    const double c = rho * a.scale;
    a.db[7][gi] += c * o1;
    a.db[5][gi] -= c * o1;
    a.db[2][gi] += c * o2;
    a.db[6][gi] -= c * o2;
    a.db[3][gi] += c * o3;
    a.db[1][gi] -= c * o3;
}

// ------------------------------------------------------- upwind convection
//
// Contraction is off here as in upwind.hpp: the explicit gradient, the
// sensor and the interface fluxes keep the reference's expression order.
#pragma clang fp contract(off)

struct UwdArgs {
    Geo g;
    int dir, ntype, ndims;
    int lo, hi;          // qrhs nodes along dir (is..ie of the axis)
    int iss, iee;        // nodes the split fluxes are taken on
    int a0, a1, b0, b1;  // line ranges over the two other axes
    int ntile_line, ntile_a;
    int recon, lchar, sensor, last;
    double bfac, gamma, mach;
    const double* q[5];
    const double* prim[6];  // rho,u,v,w,p,T
    const double* jac;
    const double* xi;       // metric along dir, xi[n+HM]
    const unsigned char* lshock;
    double* qrhs[5];
    // compact upwind: interface fluxes of F+ / F- at node index i (i+1/2)
    const double* fhp[5];
    const double* fhm[5];
};

struct SensorArgs {
    Geo g;
    int ntype[3], ndims;
    double shkcrt;
    const double* dvel[9];
    const double* prs;
    double* ssf;
    unsigned char* lshock;
};

// A bounded low / high end of an axis: ntype 1 / 2, and both ends of ntype
// 4.  The reference assigns ntype 4 only to a single rank along a
// non-periodic axis, where it leaves is/ie unset (src/parallel.F90:
// 1070-1078); on several ranks, the decomposition it runs bounded axes on,
// the physical ends have ntype 1 and 2.  ntype 4 here is their union.
__host__ __device__ __forceinline__ bool lo_end(int nt) { return nt == 1 || nt == 4; }
__host__ __device__ __forceinline__ bool hi_end(int nt) { return nt == 2 || nt == 4; }

// ducrossensor, part 1 (src/commcal.F90:236-272): ssf on 0..n; the
// pressure stencil is clamped at a bounded end (ntype 1 and 2 in the
// reference, both ends of ntype 4)
__global__ void k_ducros_ssf(SensorArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    if (p >= ni * nj * (a.g.n[2] + 1)) return;
    int i, j, k;
    split3(p, ni, nj, i, j, k);
    const long gi = gidx(a.g, i, j, k);
    const double div = a.dvel[0][gi] + a.dvel[4][gi] + a.dvel[8][gi];
    const double div2 = div * div;
    // dvel(m,n) is field 3(m-1)+(n-1)
    const double vx = a.dvel[7][gi] - a.dvel[5][gi];  // dvel(3,2) - dvel(2,3)
    const double vy = a.dvel[2][gi] - a.dvel[6][gi];  // dvel(1,3) - dvel(3,1)
    const double vz = a.dvel[3][gi] - a.dvel[1][gi];  // dvel(2,1) - dvel(1,2)
    const double vort = vx * vx + vy * vy + vz * vz;
    int ip1 = i + 1, im1 = i - 1, jp1 = j + 1, jm1 = j - 1, kp1 = k + 1, km1 = k - 1;
    if (lo_end(a.ntype[0]) && im1 < 0) im1 = 0;
    if (lo_end(a.ntype[1]) && jm1 < 0) jm1 = 0;
    if (lo_end(a.ntype[2]) && km1 < 0) km1 = 0;
    if (hi_end(a.ntype[0]) && ip1 > a.g.n[0]) ip1 = a.g.n[0];
    if (hi_end(a.ntype[1]) && jp1 > a.g.n[1]) jp1 = a.g.n[1];
    if (hi_end(a.ntype[2]) && kp1 > a.g.n[2]) kp1 = a.g.n[2];
    const double pc = a.prs[gi];
    const double* P = a.prs;
    const double pi1 = P[gidx(a.g, ip1, j, k)], pi0 = P[gidx(a.g, im1, j, k)];
    const double pj1 = P[gidx(a.g, i, jp1, k)], pj0 = P[gidx(a.g, i, jm1, k)];
    const double pk1 = P[gidx(a.g, i, j, kp1)], pk0 = P[gidx(a.g, i, j, km1)];
    const double dpdi = fabs(pi1 - 2.0 * pc + pi0) / (pi1 + 2.0 * pc + pi0);
    const double dpdj = fabs(pj1 - 2.0 * pc + pj0) / (pj1 + 2.0 * pc + pj0);
    const double dpdk = fabs(pk1 - 2.0 * pc + pk0) / (pk1 + 2.0 * pc + pk0);
    a.ssf[gi] = div2 / (div2 + vort + 1.e-30) * uwd::max2(uwd::max2(dpdi, dpdj), dpdk);
}

// dataswap(ssf) (src/commcal.F90:276; src/parallel.F90:4170-4175): a halo
// cell takes the interior value at its periodically wrapped coordinates;
// halo cells of non-periodic axes are zero (no neighbour on one domain).
// Reads the interior, writes halo cells only.  A two-dimensional box needs
// the plane k = 0 only: the window max skips the k axis there (the k halo
// planes would repeat the plane itself, which the window already holds).
__global__ void k_ssf_halo(SensorArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1 + 2 * HM, nj = a.g.n[1] + 1 + 2 * HM;
    const long nk = a.ndims == 2 ? 1 : a.g.n[2] + 1 + 2 * HM;
    if (p >= ni * nj * nk) return;
    int c[3];
    split3(p, ni, nj, c[0], c[1], c[2]);
    c[0] -= HM;
    c[1] -= HM;
    c[2] = a.ndims == 2 ? 0 : c[2] - HM;
    bool halo = false, keep = true;
    int s[3];
    for (int d = 0; d < 3; ++d) {
        s[d] = c[d];
        if (c[d] >= 0 && c[d] <= a.g.n[d]) continue;
        halo = true;
        if (d < a.ndims && a.ntype[d] == 3) s[d] = c[d] < 0 ? c[d] + a.g.n[d] : c[d] - a.g.n[d];
        else keep = false;
    }
    if (!halo) return;
    a.ssf[gidx(a.g, c[0], c[1], c[2])] = keep ? a.ssf[gidx(a.g, s[0], s[1], s[2])] : 0.0;
}

// ducrossensor, part 2 (src/commcal.F90:278-316): lshock = max of ssf over
// offsets -hm+1..hm along each axis > shkcrt, clamped at bounded ends
__global__ void k_ducros_flag(SensorArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    if (p >= ni * nj * (a.g.n[2] + 1)) return;
    int c0[3];
    split3(p, ni, nj, c0[0], c0[1], c0[2]);
    double m = 0.0;
    for (int d = 0; d < a.ndims; ++d)
        for (int o = -HM + 1; o <= HM; ++o) {
            int c[3] = {c0[0], c0[1], c0[2]};
            int t = c[d] + o;
            if (lo_end(a.ntype[d]) && t < 0) t = 0;
            if (hi_end(a.ntype[d]) && t > a.g.n[d]) t = a.g.n[d];
            c[d] = t;
            m = uwd::max2(m, a.ssf[gidx(a.g, c[0], c[1], c[2])]);
        }
    a.lshock[gidx(a.g, c0[0], c0[1], c0[2])] = m > a.shkcrt ? 1 : 0;
}

struct GradXArgs {
    Geo g;
    int ntype[3], ndims;
    const double* prim[6];
    const double* xi[3];
    double* dvel[9];
    double* dtmp[3];
};

// diff6ec at node t of a line v(-hm..dim+hm), src/derivative.F90:350-413
__device__ __forceinline__ double diff6ec_at(const double* base, long stride, int t, int dim,
                                             int ntype)
{
    auto V = [&](int n) { return base[(long)n * stride]; };
    const bool lob = ntype == 1 || ntype == 4, hib = ntype == 2 || ntype == 4;
    if (lob && t == 0) return -0.5 * V(2) + 2.0 * V(1) - 1.5 * V(0);
    if (lob && t == 1) return 0.5 * (V(2) - V(0));
    if (lob && t == 2) return (2.0 / 3.0) * (V(3) - V(1)) - (1.0 / 12.0) * (V(4) - V(0));
    if (hib && t == dim) return 0.5 * V(dim - 2) - 2.0 * V(dim - 1) + 1.5 * V(dim);
    if (hib && t == dim - 1) return 0.5 * (V(dim) - V(dim - 2));
    if (hib && t == dim - 2)
        return (2.0 / 3.0) * (V(dim - 1) - V(dim - 3)) - (1.0 / 12.0) * (V(dim) - V(dim - 4));
    return 0.75 * (V(t + 1) - V(t - 1)) - 0.15 * (V(t + 2) - V(t - 2)) +
           (1.0 / 60.0) * (V(t + 3) - V(t - 3));
}

// gradcal with the explicit 6th-order derivative (difschm '643e'),
// src/comsolver.F90:244-497: dvel(m,d) = d/dxi_d(u_m) * dxi(d,d)
// (streaming stores, as in k_rk: 2048^2 Riemann step 4.01 -> 3.89 ms; the
// qrhs that k_zero_qrhs clears and k_uwd accumulates stays cached, which
// measured 40 us per step faster than streaming it)
__global__ void k_grad_explicit(GradXArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    if (p >= ni * nj * (a.g.n[2] + 1)) return;
    int c[3];
    split3(p, ni, nj, c[0], c[1], c[2]);
    const long gi = gidx(a.g, c[0], c[1], c[2]);
    for (int d = 0; d < 3; ++d) {
        if (d >= a.ndims) {
            for (int m = 0; m < 3; ++m) st_nt(&a.dvel[3 * m + d][gi], 0.0);
            st_nt(&a.dtmp[d][gi], 0.0);
            continue;
        }
        const long st = a.g.s[d];
        const long line0 = gi - (long)c[d] * st;  // node 0 of this line
        const double x = a.xi[d][c[d] + HM];
        for (int m = 0; m < 4; ++m) {
            const double* f = a.prim[m < 3 ? 1 + m : 5] + line0;
            const double df = diff6ec_at(f, st, c[d], a.g.n[d], a.ntype[d]);
            if (m < 3) st_nt(&a.dvel[3 * m + d][gi], df * x);
            else st_nt(&a.dtmp[d][gi], df * x);
        }
    }
}

struct FieldXArgs {
    Geo g;
    int dir, ntype;
    const double* fin;
    double* fout;
};

// df_explicit (fds%central with difschm '643e') of every line of a field,
// src/derivative.F90:318-337
__global__ void k_deriv_explicit_field(FieldXArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    if (p >= ni * nj * (a.g.n[2] + 1)) return;
    int c[3];
    split3(p, ni, nj, c[0], c[1], c[2]);
    const long gi = gidx(a.g, c[0], c[1], c[2]);
    const long st = a.g.s[a.dir];
    a.fout[gi] = diff6ec_at(a.fin + (gi - (long)c[a.dir] * st), st, c[a.dir], a.g.n[a.dir], a.ntype);
}

struct CmpArgs {
    Geo g;
    int dir, ntype, ndims;
    int iss, iee;
    int a0, a1, b0, b1, na;
    int first, last;           // rows of the compact flux system
    double bfac, gamma, mach;
    const double* q[5];
    const double* prim[6];
    const double* jac;
    const double* xi;          // metric along dir, xi[n+HM]
    double* sw[10];            // F+(1:5), F-(1:5) on iss..iee
    double* fh[10];            // interface fluxes of F+ / F-, rows first..last
    const double* ac[2];       // Thomas factors (3 per row) of the + / - system
    int comp0, ncomp;          // k_flux_compact: components comp0..comp0+ncomp-1
};

// flux_steger_warming on the nodes iss..iee of every line of one axis,
// src/solver.F90:1326-1336, into the 10 split-flux fields
__global__ void k_sw_field(CmpArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nt = a.iee - a.iss + 1;
    const long nl = (long)a.na * (a.b1 - a.b0 + 1);
    if (p >= nt * nl) return;
    // consecutive threads walk the fastest-varying memory axis
    int t, av, b;
    if (a.dir == 0) {
        t = a.iss + (int)(p % nt);
        const long r = p / nt;
        av = a.a0 + (int)(r % a.na);
        b = a.b0 + (int)(r / a.na);
    } else {
        av = a.a0 + (int)(p % a.na);
        const long r = p / a.na;
        t = a.iss + (int)(r % nt);
        b = a.b0 + (int)(r / nt);
    }
    // node t along the axis, av and b along the others (in axis order)
    const int c0 = a.dir == 0 ? t : av;
    const int c1 = a.dir == 1 ? t : (a.dir == 0 ? av : b);
    const int c2 = a.dir == 2 ? t : b;
    const long gi = gidx(a.g, c0, c1, c2);
    const double vel[3] = {a.prim[1][gi], a.prim[2][gi], a.prim[3][gi]};
    double q[5];
    for (int m = 0; m < 5; ++m) q[m] = a.q[m][gi];
    const double xv = a.xi[t + HM];
    const double dxi[3] = {a.dir == 0 ? xv : 0.0, a.dir == 1 ? xv : 0.0, a.dir == 2 ? xv : 0.0};
    double fp[5], fm[5];
    uwd::sw_split(a.gamma, a.mach, a.prim[0][gi], vel, a.prim[4][gi], a.prim[5][gi], q, dxi,
                  a.jac[gi], fp, fm);
    for (int m = 0; m < 5; ++m) {
        a.sw[m][gi] = fp[m];
        a.sw[5 + m][gi] = fm[m];
    }
}

// flux_compact (src/flux.F90:125-256) of one split-flux component along one
// line, per thread: the right-hand side of compact_flux_rhs, then the
// sequential tridiagonal_thomas_solver (src/commfunc.F90:790-813) in place
// in the output field (row r at node index r).
__global__ void k_flux_compact(CmpArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nl = (long)a.na * (a.b1 - a.b0 + 1);
    if (p >= nl * a.ncomp) return;
    const int av = a.a0 + (int)(p % a.na);
    const long r = p / a.na;
    const int b = a.b0 + (int)(r % (a.b1 - a.b0 + 1));
    const int comp = a.comp0 + (int)(r / (a.b1 - a.b0 + 1));  // 0..4 F+, 5..9 F-
    const bool plus = comp < 5;
    int c[3];
    c[a.dir] = 0;
    c[a.dir == 0 ? 1 : 0] = av;
    c[a.dir == 2 ? 1 : 2] = b;
    const long st = a.g.s[a.dir];
    const double* f = a.sw[comp] + gidx(a.g, c[0], c[1], c[2]);
    double* d = a.fh[comp] + gidx(a.g, c[0], c[1], c[2]);
    const double* ac = a.ac[plus ? 0 : 1];
    auto F = [&](int n) { return f[(long)n * st]; };
    const int i0 = a.first, im_ = a.last, nt = a.ntype;
    const double bf = a.bfac;
    int is = i0 + 1, ie = im_ - 1;
    // compact_flux_rhs, src/flux.F90:187-254
    if (nt == 1 || nt == 4) {
        d[(long)i0 * st] = 2.5 * F(i0 + 1) + 0.5 * F(i0 + 2);
        d[(long)(i0 + 1) * st] = 0.75 * F(i0 + 1) + 0.75 * F(i0 + 2);
        is = i0 + 2;
    } else {
        const int j = i0;
        const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
        d[(long)j * st] = (37.0 / 60.0) * v1 - (2.0 / 15.0) * v2 + (1.0 / 60.0) * v3;
    }
    if (nt == 2 || nt == 4) {
        d[(long)(im_ - 1) * st] = 0.75 * F(im_) + 0.75 * F(im_ - 1);
        d[(long)im_ * st] = 2.5 * F(im_) + 0.5 * F(im_ - 1);
        ie = im_ - 2;
    } else {
        const int j = im_;
        const double v1 = F(j) + F(j + 1), v2 = F(j - 1) + F(j + 2), v3 = F(j - 2) + F(j + 3);
        d[(long)j * st] = (37.0 / 60.0) * v1 - (2.0 / 15.0) * v2 + (1.0 / 60.0) * v3;
    }
    const double c0 = (1.0 / 18.0) - (1.0 / 36.0) * bf, c1 = (19.0 / 18.0) - (9.0 / 36.0) * bf;
    const double c2 = (5.0 / 9.0) + (9.0 / 36.0) * bf, c3 = (1.0 / 36.0) * bf;
    for (int j = is; j <= ie; ++j) {
        double v;
        if (plus) v = c0 * F(j - 1) + c1 * F(j) + c2 * F(j + 1) + c3 * F(j + 2);
        else v = c0 * F(j + 2) + c1 * F(j + 1) + c2 * F(j) + c3 * F(j - 1);
        d[(long)j * st] = v;
    }
    // forward elimination and back substitution (rows 0..n-1 = i0..im_)
    const int n = im_ - i0 + 1;
    double prev = d[(long)i0 * st];
    for (int k = 1; k < n; ++k) {
        double* dk = d + (long)(i0 + k) * st;
        prev = *dk * ac[3 * k + 1] - prev * ac[3 * k + 2];
        *dk = prev;
    }
    double nxt = prev;
    for (int k = n - 2; k >= 0; --k) {
        double* dk = d + (long)(i0 + k) * st;
        nxt = *dk - ac[3 * k + 0] * nxt;
        *dk = nxt;
    }
}

struct ReconsArgs {
    Geo g;
    int dir, ntype, lo, hi;  // interfaces lo..hi (node i: the interface i+1/2)
    int plus, recon, shock;
    double bfac;
    const double* f;
    double* fh;
};

// recons_exp (src/flux.F90:269-350) of one field at the interfaces lo..hi
// of every line of an axis, one thread per interface: the stencil f(i-3:i+4)
// for the + wind, f(i+4:i-3) for the - wind (convrsduwd's iwind8 order,
// src/solver.F90:620-650), halos as given.  ntype 4 is taken as ntype 1 on
// the low half of the line and 2 on the high half, as in k_uwd.
__global__ void k_recons_field(ReconsArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int d = a.dir, dim = a.g.n[d];
    const long ni = a.hi - a.lo + 1;
    const long na = a.g.n[d == 0 ? 1 : 0] + 1, nb = a.g.n[d == 2 ? 1 : 2] + 1;
    if (p >= ni * na * nb) return;
    const int i = a.lo + (int)(p % ni);
    const long r = p / ni;
    int c[3];
    c[d] = 0;
    c[d == 0 ? 1 : 0] = (int)(r % na);
    c[d == 2 ? 1 : 2] = (int)(r / na);
    const long st = a.g.s[d], g0 = gidx(a.g, c[0], c[1], c[2]);
    double f8[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) f8[w] = a.f[g0 + (long)(a.plus ? i - 3 + w : i + 4 - w) * st];
    const int nte = a.ntype == 4 ? (i < dim / 2 ? 1 : 2) : a.ntype;
    a.fh[g0 + (long)i * st] = uwd::recons_exp(f8, i, dim, nte, a.recon, a.shock != 0, a.bfac);
}

// Zero qrhs on the interior (qrhs=0 before rhscal, src/mainloop.F90:404)
__global__ void k_zero_qrhs(UwdArgs a)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ni = a.g.n[0] + 1, nj = a.g.n[1] + 1;
    if (p >= ni * nj * (a.g.n[2] + 1)) return;
    int ci, cj, ck;
    split3(p, ni, nj, ci, cj, ck);
    const long gi = gidx(a.g, ci, cj, ck);
    for (int m = 0; m < 5; ++m) a.qrhs[m][gi] = 0.0;
}

// convrsduwd along one axis, src/solver.F90:549-1202 (no species, crinod
// false, no solids).  A tile is TL lines x S interfaces i0-1 .. i0+S-2:
//  1. the Steger-Warming split fluxes of nodes i0-4 .. i0+S+2 (their index
//     clamped to iss..iee, which is iwind8's clamp) go to LDS;
//  2. each thread forms the interface flux Fh(i) from the 8-point stencils
//     of F+ and F-, projected on the Roe eigenvectors where the scheme and
//     the shock flag ask for it;
//  3. qrhs(i) += Fh(i) - Fh(i-1) for the nodes i0 .. i0+S-2 within is..ie
//     (negated on the last axis: qrhs=-qrhs, src/solver.F90:243).
// x tiles are 1 line x 256 interfaces; y/z tiles 16 lines (consecutive i,
// 128-byte rows) x 16 interfaces.
// LD: the explicit path with MP7-LD (reschem 5), whose reconstruction needs
// 236 registers (2 waves per SIMD).  The other explicit schemes run with 4
// waves per SIMD (128 registers; the dozen the compiler then keeps in
// scratch cost less than the waves gained: MP7 at 2048^2 from 445 to 314 us
// per launch).
template <bool XD, bool CMP, bool LD = false>
__global__ __launch_bounds__(256, CMP || LD ? 1 : 4) void k_uwd(UwdArgs a)
{
    constexpr int TL = XD ? 1 : 16;
    constexpr int S = XD ? 256 : 16;
    constexpr int NN = S + 7;
    __shared__ double sw[TL * NN * 10];
    __shared__ double fh[TL * S * 5];
    const int tid = threadIdx.x;
    int bid = blockIdx.x;
    const int seg = bid % a.ntile_line;
    bid /= a.ntile_line;
    const int ta = bid % a.ntile_a;
    const int b = a.b0 + bid / a.ntile_a;
    const int i0 = a.lo + seg * (S - 1);
    const int dir = a.dir;
    const int dim = a.g.n[dir];
    auto line_node = [&](int l, int t, int& i, int& j, int& k) {
        const int av = min(a.a0 + ta * TL + l, a.a1);
        if (dir == 0) { i = t; j = av; k = b; }
        else if (dir == 1) { i = av; j = t; k = b; }
        else { i = av; j = b; k = t; }
    };
    // 1. split fluxes
    for (int e = tid; e < TL * NN; e += 256) {
        const int l = e % TL, t = e / TL;
        const int node = min(max(i0 - 4 + t, a.iss), a.iee);
        int i, j, k;
        line_node(l, node, i, j, k);
        const long gi = gidx(a.g, i, j, k);
        double vel[3] = {a.prim[1][gi], a.prim[2][gi], a.prim[3][gi]};
        double q[5];
        for (int m = 0; m < 5; ++m) q[m] = a.q[m][gi];
        // dxi(dir, dir) in its slot (selects: an array indexed by the
        // run-time axis would live in scratch memory)
        const double xv = a.xi[node + HM];
        const double dxi[3] = {dir == 0 ? xv : 0.0, dir == 1 ? xv : 0.0, dir == 2 ? xv : 0.0};
        double* o = sw + (l * NN + t) * 10;
        uwd::sw_split(a.gamma, a.mach, a.prim[0][gi], vel, a.prim[4][gi], a.prim[5][gi], q, dxi,
                      a.jac[gi], o, o + 5);
    }
    __syncthreads();
    // 2. interface fluxes
    const int l = XD ? 0 : (tid & 15);
    const int s = XD ? tid : (tid >> 4);
    const int ii = i0 - 1 + s;  // interface ii+1/2
    if (ii <= a.hi) {
        bool lsh = true;
        if (a.sensor) {
            int i, j, k;
            line_node(l, ii < 0 ? ii + 1 : ii, i, j, k);
            const bool s1 = a.lshock[gidx(a.g, i, j, k)] != 0;
            line_node(l, ii + 1 > dim ? ii : ii + 1, i, j, k);
            lsh = s1 || a.lshock[gidx(a.g, i, j, k)] != 0;
        }
        // convrsdcmp projects every interface, convrsduwd only flagged ones
        const bool chr = a.lchar && (CMP || lsh);
        uwd::Roe r;
        if (chr) {
            int i, j, k;
            line_node(l, ii, i, j, k);
            const long gl = gidx(a.g, i, j, k);
            line_node(l, ii + 1, i, j, k);
            const long gr = gidx(a.g, i, j, k);
            const double vl[3] = {a.prim[1][gl], a.prim[2][gl], a.prim[3][gl]};
            const double vr[3] = {a.prim[1][gr], a.prim[2][gr], a.prim[3][gr]};
            const double xl = a.xi[ii + HM], xr = a.xi[ii + 1 + HM];
            const double dl[3] = {dir == 0 ? xl : 0.0, dir == 1 ? xl : 0.0, dir == 2 ? xl : 0.0};
            const double dr[3] = {dir == 0 ? xr : 0.0, dir == 1 ? xr : 0.0, dir == 2 ? xr : 0.0};
            r = uwd::roe_state(a.gamma, a.prim[0][gl], a.prim[4][gl], a.q[4][gl], vl, dl,
                               a.prim[0][gr], a.prim[4][gr], a.q[4][gr], vr, dr);
        }
        const double* base = sw + l * NN * 10;
        double fhc[5];
        if constexpr (CMP) {
            // convrsdcmp (src/solver.F90:1340-1460): 5-point stencils of
            // iwind6 about the compact interface values, then mplimiter
            int i, j, k;
            line_node(l, ii, i, j, k);
            const long gf = gidx(a.g, i, j, k);  // interface ii+1/2 sits at node ii
            double cp5[5], cm5[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                cp5[v] = a.fhp[v][gf];
                cm5[v] = a.fhm[v][gf];
            }
            for (int m = 0; m < 5; ++m) {
                double fp[5], fm[5], cpl, cml;
                if (chr) {
                    double lv[5];
                    uwd::lev_row(r, m, lv);
#pragma unroll
                    for (int w = 0; w < 5; ++w) {
                        const double* P = base + (s + 1 + w) * 10;      // node ii+w-2
                        const double* M = base + (s + 6 - w) * 10 + 5;  // node ii+3-w
                        fp[w] = lv[0] * P[0] + lv[1] * P[1] + lv[2] * P[2] + lv[3] * P[3] + lv[4] * P[4];
                        fm[w] = lv[0] * M[0] + lv[1] * M[1] + lv[2] * M[2] + lv[3] * M[3] + lv[4] * M[4];
                    }
                    cpl = lv[0] * cp5[0] + lv[1] * cp5[1] + lv[2] * cp5[2] + lv[3] * cp5[3] + lv[4] * cp5[4];
                    cml = lv[0] * cm5[0] + lv[1] * cm5[1] + lv[2] * cm5[2] + lv[3] * cm5[3] + lv[4] * cm5[4];
                } else {
#pragma unroll
                    for (int w = 0; w < 5; ++w) {
                        fp[w] = base[(s + 1 + w) * 10 + m];
                        fm[w] = base[(s + 6 - w) * 10 + 5 + m];
                    }
                    cpl = cp5[m];
                    cml = cm5[m];
                }
                const double v1 = uwd::mplimiter(fp, cpl, lsh, ii, dim, a.ntype);
                const double v2 = uwd::mplimiter(fm, cml, lsh, ii, dim, a.ntype);
                fhc[m] = v1 + v2;
            }
        } else {
        // recons_exp sees ntype 4 as ntype 1 on the low half of the line
        // and ntype 2 on the high half: what the reference computes with
        // the axis split over several ranks (the explicit flux has no
        // closure at a rank interface)
        const int nte = a.ntype == 4 ? (ii < dim / 2 ? 1 : 2) : a.ntype;
        for (int m = 0; m < 5; ++m) {
            // the F+ stencil and its reconstruction, then F-'s: the empty
            // asm keeps the second's LDS reads after the first's result, so
            // that the two stencils are not live at once
            double lv[5];
            if (chr) uwd::lev_row(r, m, lv);
            double v1, v2;
            {
                double fp[8];
                if (chr) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        const double* P = base + (s + w) * 10;  // node ii+w-3
                        fp[w] = lv[0] * P[0] + lv[1] * P[1] + lv[2] * P[2] + lv[3] * P[3] + lv[4] * P[4];
                    }
                } else {
#pragma unroll
                    for (int w = 0; w < 8; ++w) fp[w] = base[(s + w) * 10 + m];
                }
                v1 = uwd::recons_exp<LD>(fp, ii, dim, nte, a.recon, lsh, a.bfac);
            }
            asm volatile("" : "+v"(v1) : : "memory");
            {
                double fm[8];
                if (chr) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        const double* M = base + (s + 7 - w) * 10 + 5;  // node ii+4-w
                        fm[w] = lv[0] * M[0] + lv[1] * M[1] + lv[2] * M[2] + lv[3] * M[3] + lv[4] * M[4];
                    }
                } else {
#pragma unroll
                    for (int w = 0; w < 8; ++w) fm[w] = base[(s + 7 - w) * 10 + 5 + m];
                }
                v2 = uwd::recons_exp<LD>(fm, ii, dim, nte, a.recon, lsh, a.bfac);
            }
            fhc[m] = v1 + v2;
        }
        }
        double* F = fh + (l * S + s) * 5;
        if (chr) {
            for (int m = 0; m < 5; ++m) {
                double rv[5];
                uwd::rev_row(r, m, rv);
                F[m] = rv[0] * fhc[0] + rv[1] * fhc[1] + rv[2] * fhc[2] + rv[3] * fhc[3] +
                       rv[4] * fhc[4];
            }
        } else {
            for (int m = 0; m < 5; ++m) F[m] = fhc[m];
        }
    }
    __syncthreads();
    // 3. flux difference
    const int node = i0 - 1 + s;
    const bool lvalid = XD || (a.a0 + ta * TL + l <= a.a1);
    if (s >= 1 && node <= a.hi && lvalid) {
        int i, j, k;
        line_node(l, node, i, j, k);
        const long gi = gidx(a.g, i, j, k);
        const double* F1 = fh + (l * S + s) * 5;
        const double* F0 = F1 - 5;
        double qr[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) qr[m] = a.qrhs[m][gi];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double v = qr[m] + F1[m] - F0[m];
            a.qrhs[m][gi] = a.last ? -v : v;
        }
    }
}

// ---- user boundary condition of the 2-D Riemann example

struct BcArgs {
    Geo g;
    double* q[5];
    double* prim[6];
    int face[4];    // faces 1..4 (i=0, i=im, j=0, j=jm) under the user condition
    double time;    // the step's start time
    double gamma, mach, const1, const2, const6;
    double len[2];  // gridcube lengths of x and y (x(i) = len/n*i), for gcnscbc
    double tw[6];   // wall temperature of each face (noslip)
    double dt;      // the step's time step (outflow)
    double xslip;   // slip-nonslip point of the isothermal slip wall
};

// the boundary conditions of k_bc_extrap
enum {
    BCX_ZEROEXTRAP = 0,
    BCX_SYMMETRY = 1,
    BCX_NOSLIP = 2,
    BCX_ADIABATIC = 3,
    BCX_FIX = 4,
    BCX_SLIPADIAB = 5,
    BCX_SLIPISOT = 6
};

// One boundary node as udf_bc sets it: the primitives, tmp = thermal(
// pressure, density) = p/rho*const2 and q from fvar2q_sca's pressure form,
// q5 = p*const6 + rho*(0.5*sum(u*u)) (src/fludyna.F90:45-70, 310-362).
__device__ __forceinline__ void bc_node(const BcArgs& a, long gi, double r, double u, double v,
                                        double w, double p)
{
    a.prim[0][gi] = r;
    a.prim[1][gi] = u;
    a.prim[2][gi] = v;
    a.prim[3][gi] = w;
    a.prim[4][gi] = p;
    a.prim[5][gi] = p / r * a.const2;
    a.q[0][gi] = r;
    a.q[1][gi] = r * u;
    a.q[2][gi] = r * v;
    a.q[3][gi] = r * w;
    const double var1 = 0.5 * (u * u + v * v + w * w);
    a.q[4][gi] = p * a.const6 + r * var1;
}

// shockspeed, examples/Riemann2D/user_define_module/userdefine_riemann2d.F90:
// 670-686: the Rankine-Hugoniot speed of the shock p1 -> p2 into a gas at
// temperature t1, sqrt(1 + (p2/p1-1)(g+1)/(2g)) sos(t1), sos = sqrt(T)/Ma.
__device__ __forceinline__ double shockspeed(const BcArgs& a, double p1, double t1, double p2)
{
    const double mx2 = (p2 / p1 - 1.0) * (a.gamma + 1.0) / (2.0 * a.gamma) + 1.0;
    return sqrt(mx2) * (sqrt(t1) / a.mach);
}

// the quadrant states of the example (rho, u, v, p): NE 11, NW 01, SW 00,
// SE 10; the NW pressure literal is single precision (0.30) in the source
#define R2D_STATES                                                    \
    const double r01 = 0.5323, r11 = 1.5, u01 = 1.206, v01 = 0.0;     \
    const double v11 = 0.0, p01 = (double)0.30f, p11 = 1.5;           \
    const double r00 = 0.138, r10 = 0.5323, u00 = 1.206, v00 = 1.206; \
    const double u10 = 0.0, v10 = 1.206, p00 = 0.029, p10 = 0.3;

// udf_bc faces 1 and 2 (userdefine_riemann2d.F90:446-541), one thread per
// (j,k): face 1 (i=0) takes the states behind and ahead of the shock that
// moves down from y = 0.8 at vss, with u = 1.206; face 2 (i=im) extrapolates
// rho, p and v as (4f(im-1) - f(im-2))/3 with u = w = 0, and the corner
// (im,jm) is the north-east state at rest.  The grid is gridcube(1,1,0)'s:
// x = i/im, y = j/jm (src/gridgeneration.F90:233-267).
__global__ void k_bc_riem2d_i(BcArgs a)
{
    const long nj = a.g.n[1] + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nj * (a.g.n[2] + 1)) return;
    const int j = (int)(t % nj), k = (int)(t / nj);
    const int im = a.g.n[0], jm = a.g.n[1];
    R2D_STATES
    (void)u10; (void)v10; (void)p10; (void)r10;
    if (a.face[0]) {
        const double t1 = p01 / r01 * a.const2;
        const double vss = shockspeed(a, p01, t1, p00);
        const double ys = 0.8 - vss * a.time;
        const bool up = (1.0 / (double)jm) * (double)j > ys;
        bc_node(a, gidx(a.g, 0, j, k), up ? r01 : r00, u01, up ? v01 : v00, 0.0, up ? p01 : p00);
    }
    if (a.face[1]) {
        const long gd = gidx(a.g, im, j, k);
        if (j == jm) {
            bc_node(a, gd, r11, 0.0, v11, 0.0, p11);
        } else {
            const long g1 = gidx(a.g, im - 1, j, k), g2 = gidx(a.g, im - 2, j, k);
            const double n1d3 = 1.0 / 3.0;
            const double r = n1d3 * (4.0 * a.prim[0][g1] - a.prim[0][g2]);
            const double p = n1d3 * (4.0 * a.prim[4][g1] - a.prim[4][g2]);
            const double v = n1d3 * (4.0 * a.prim[2][g1] - a.prim[2][g2]);
            bc_node(a, gd, r, 0.0, v, 0.0, p);
        }
    }
}

// udf_bc faces 3 and 4 (userdefine_riemann2d.F90:543-666), one thread per
// (i,k), after faces 1 and 2: face 3 (j=0) takes the states either side of
// the shock that moves left from x = 0.8, with v = 1.206; face 4 (j=jm)
// extrapolates rho, p and u along j with v = w = 0 (its nodes i = 0 and
// i = im read what faces 1 and 2 wrote), and the corner again.
__global__ void k_bc_riem2d_j(BcArgs a)
{
    const long ni = a.g.n[0] + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * (a.g.n[2] + 1)) return;
    const int i = (int)(t % ni), k = (int)(t / ni);
    const int im = a.g.n[0], jm = a.g.n[1];
    R2D_STATES
    (void)u01; (void)v01; (void)p01; (void)r01;
    if (a.face[2]) {
        const double t1 = p10 / r10 * a.const2;
        const double vss = shockspeed(a, p10, t1, p00);
        const double xs = 0.8 - vss * a.time;
        const bool up = (1.0 / (double)im) * (double)i > xs;
        bc_node(a, gidx(a.g, i, 0, k), up ? r10 : r00, up ? u10 : u00, v10, 0.0, up ? p10 : p00);
    }
    if (a.face[3]) {
        const long gd = gidx(a.g, i, jm, k);
        if (i == im) {
            bc_node(a, gd, r11, 0.0, 0.0, 0.0, p11);
        } else {
            const long g1 = gidx(a.g, i, jm - 1, k), g2 = gidx(a.g, i, jm - 2, k);
            const double n1d3 = 1.0 / 3.0;
            const double r = n1d3 * (4.0 * a.prim[0][g1] - a.prim[0][g2]);
            const double p = n1d3 * (4.0 * a.prim[4][g1] - a.prim[4][g2]);
            const double u = n1d3 * (4.0 * a.prim[1][g1] - a.prim[1][g2]);
            bc_node(a, gd, r, u, 0.0, 0.0, p);
        }
    }
}
#undef R2D_STATES

// gcnscbc (bctype 23), src/bc.F90:7907-8137: the ghost-cell NSCBC of
// Motheau, Almgren & Bell on face 2 (i = im, AX 0) or face 4 (j = jm,
// AX 1), with prs_t = pinf, nondimensional, no species.  One thread per
// node of the face: the LODI amplitudes from one-sided first differences
// of the two nodes inside (deriv_1o, src/commfunc.F90:213-222, over the
// gridcube spacing x(n-1) - x(n-2)); the incoming wave replaced by the soft
// outflow relaxation towards pinf (or, on face 4 where the normal velocity
// is negative, the inflow relaxations towards the freestream); projected
// back (Poinsot & Lele 1992, Eqs. 33-36) and extrapolated to the face node,
// whose density is thermal(p, T) and whose q is fvar2q(rho, u, p).  The
// normal velocity, the sound speed and its square are INTEGER variables in
// the reference (truncated toward zero), the Mach number their integer
// quotient, and lambda(5) takes the normal velocity (the reference's
// lambda(5:numq) = vnorm with numq = 5): all kept.  Freestream of the
// nondimensional run: uinf = 1, vinf = winf = 0, tinf = 1, pinf = 1/const2.
template <int AX>
__global__ void k_bc_gcnscbc(BcArgs a)
{
    const int na = a.g.n[AX == 0 ? 1 : 0];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)(na + 1) * (a.g.n[2] + 1)) return;
    const int ia = (int)(t % (na + 1)), k = (int)(t / (na + 1));
    const int n = a.g.n[AX];
    const int i = AX == 0 ? n : ia, j = AX == 0 ? ia : n;
    const long gi = gidx(a.g, i, j, k);
    const long g1 = AX == 0 ? gidx(a.g, i - 1, j, k) : gidx(a.g, i, j - 1, k);
    const long g2 = AX == 0 ? gidx(a.g, i - 2, j, k) : gidx(a.g, i, j - 2, k);
    const double gamma = a.gamma, pinf = 1.0 / a.const2;
    const double uinf = 1.0, vinf = 0.0, winf = 0.0, tinf = 1.0, eps = 1.e-10;
    const double hx = a.len[AX] / (double)n;
    const double dx = hx * (double)(n - 1) - hx * (double)(n - 2);
    const double span = hx * (double)n - 0.0;
    const double t1 = a.prim[5][g1], r1 = a.prim[0][g1], p1 = a.prim[4][g1];
    const double v1[3] = {a.prim[1][g1], a.prim[2][g1], a.prim[3][g1]};
    const double v2[3] = {a.prim[1][g2], a.prim[2][g2], a.prim[3][g2]};
    const int css = (int)(sqrt(t1) / a.mach);
    if (css == 0) {
        // the reference divides by the zero integer sound speed here (a
        // trap in Fortran); the node is set to NaN instead
        const double nan = __builtin_nan("");
        for (int m = 0; m < 6; ++m) a.prim[m][gi] = nan;
        for (int m = 0; m < 5; ++m) a.q[m][gi] = nan;
        return;
    }
    const int css2 = css * css;
    const int vnorm = (int)v1[AX];
    const double mavg = (double)(vnorm / css);
    double lambda[5];
    lambda[0] = (double)(vnorm - css) + eps;
    lambda[1] = (double)vnorm + eps;
    lambda[2] = (double)vnorm + eps;
    lambda[3] = (double)vnorm + eps;
    lambda[4] = (double)vnorm + eps;
    double dvel[3];
    for (int c = 0; c < 3; ++c) dvel[c] = -(-v1[c] + v2[c]) / dx;
    double dprs = -(-p1 + a.prim[4][g2]) / dx;
    double drho = -(-r1 + a.prim[0][g2]) / dx;
    constexpr int TG = AX == 0 ? 1 : 0;  // the in-plane tangential component
    double lodi[5];
    lodi[0] = lambda[0] * (dprs - r1 * (double)css * dvel[AX]);
    lodi[1] = lambda[1] * ((double)css2 * drho - dprs);
    lodi[2] = lambda[2] * dvel[TG];
    lodi[3] = lambda[3] * dvel[2];
    lodi[4] = lambda[4] * (dprs + r1 * (double)css * dvel[AX]);
    if (AX == 0 || vnorm >= 0) {
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
    const double dtmp = t1 / r1 / (double)css * (-lodi[1] + 0.5 * (gamma - 1.0) * (lodi[0] + lodi[4]));
    dprs = 0.5 * (lodi[0] + lodi[4]);
    dvel[AX] = 1.0 / (2.0 * r1 * (double)css) * (lodi[4] - lodi[0]);
    dvel[TG] = lodi[2];
    dvel[2] = lodi[3];
    const double tn = t1 + 2.0 * dx * dtmp;
    const double pn = p1 + 2.0 * dx * dprs;
    double vn[3];
    for (int c = 0; c < 3; ++c) vn[c] = v1[c] + 2.0 * dx * dvel[c];
    const double rn = pn / tn * a.const2;
    a.prim[0][gi] = rn;
    a.prim[1][gi] = vn[0];
    a.prim[2][gi] = vn[1];
    a.prim[3][gi] = vn[2];
    a.prim[4][gi] = pn;
    a.prim[5][gi] = tn;
    a.q[0][gi] = rn;
    a.q[1][gi] = rn * vn[0];
    a.q[2][gi] = rn * vn[1];
    a.q[3][gi] = rn * vn[2];
    const double var1 = 0.5 * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    a.q[4][gi] = pn * a.const6 + rn * var1;
}

// zeroextrap (src/bc.F90:2331-2627, bctype 50; MODE BCX_ZEROEXTRAP), symmetry
// (src/bc.F90:2639-2971, bctype 60; BCX_SYMMETRY) and the walls (bctypes 41,
// 42; BCX_NOSLIP, BCX_ADIABATIC) on the face of axis AX at its
// low (HI false) or high end, over 0..n of the other two axes.  Velocity and
// pressure of the boundary node by extrapolate_2o with dv = 0 from its two
// inner neighbours (src/commfunc.F90:277-285: v0 = (4 v1 - v2 - 2 dv)/3).
//   zeroextrap: on the x faces the temperature too and the density from
//   thermal(p, T) = p/T*const2, on the y and z faces the density and the
//   temperature from thermal(p, rho) = p/rho*const2; q by fvar2q's pressure
//   form.
//   symmetry (rectilinear grids): the velocity normal to the face zero
//   (faces 1, 2 and 6 project the extrapolated velocity through the boundary
//   normal dxi(d,:)/|dxi(d,:)| = +-e_d, which leaves exactly that; faces
//   3-5 set it); the temperature extrapolated and the density from p/T,
//   except on face 6 (the density extrapolated, T from p/rho); q by fvar2q's
//   temperature form rho*(T*const1 + |u|^2/2) on faces 1, 2, 4 and 6, its
//   pressure form on faces 3 and 5 (src/fludyna.F90:312-376).
// In the reference's operation order, like the rest of this region.
//   slipadibwall (MODE 5, src/bc.F90:7114-7313, faces 3 and 4): pressure,
//   temperature and u extrapolated, w = 0, v = 0 on face 3 and extrapolated
//   on face 4, the density from p/T; q by the pressure form.
//   slipisotwall (MODE 6, src/bc.F90:6921-7112) on face 3: the slip wall's
//   node where x <= xslip, else at rest at the face's wall temperature; its
//   face 4 is the noslip wall's (MODE 2).
//   fixbc (MODE 4, src/bc.F90:7562-7648): the fixed states of faces 3
//   (rho 2, p 1) and 4 (rho 1, p 2.5) at rest, T = p/rho*const2.
//   walls (MODE 2 noslip, src/bc.F90:6281-6595; MODE 3 noslip_adibatic,
//   :6611-6919; faces 1-4): the node at rest, the pressure num1d3*(4 p1 -
//   p2), the temperature the face's wall temperature (2) or extrapolated
//   alike (3), the density from p/T; q by fvar2q's pressure form.
template <int AX, bool HI, int MODE>
__global__ void k_bc_extrap(BcArgs a)
{
    constexpr bool SYM = MODE == BCX_SYMMETRY, WALL = MODE == BCX_NOSLIP || MODE == BCX_ADIABATIC;
    constexpr bool SLIP = MODE == BCX_SLIPADIAB, SLIPT = MODE == BCX_SLIPISOT;
    constexpr int OA = AX == 0 ? 1 : 0, OB = AX == 2 ? 1 : 2;
    constexpr bool RHO_EXT = (WALL || SLIP || SLIPT) ? false : (SYM ? (AX == 2 && HI) : AX != 0);
    constexpr bool TFORM = SYM && (AX == 0 || HI);
    const long ea = a.g.n[OA] + 1L;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ea * (a.g.n[OB] + 1L)) return;
    const int ca = (int)(t % ea), cb = (int)(t / ea);
    const int n = a.g.n[AX];
    const int c0 = HI ? n : 0, c1 = HI ? n - 1 : 1, c2 = HI ? n - 2 : 2;
    auto at = [&](int cd) {
        const int i = AX == 0 ? cd : ca;
        const int j = AX == 1 ? cd : (AX == 0 ? ca : cb);
        const int k = AX == 2 ? cd : cb;
        return gidx(a.g, i, j, k);
    };
    const long gi = at(c0), g1 = at(c1), g2 = at(c2);
    const double n1d3 = 1.0 / 3.0;
    double ve[3];
    for (int m = 0; m < 3; ++m)
        ve[m] = WALL ? 0.0 : n1d3 * (4.0 * a.prim[1 + m][g1] - a.prim[1 + m][g2] - 2.0 * 0.0);
    if (SYM) ve[AX] = 0.0;
    if (SLIP || SLIPT) {
        // slipadibwall: u extrapolated, w = 0; v = 0 on face 3, extrapolated
        // on face 4.  slipisotwall (face 3 only): the same where the node's
        // gridcube x = lx/im*i is at most xslip, else at rest (the unread wall
        // blowing zero) at the face's wall temperature
        if (!HI) ve[1] = 0.0;
        ve[2] = 0.0;
    }
    bool stick = false;
    if (SLIPT) {
        const double xn = a.len[0] / (double)a.g.n[0] * (double)ca;
        stick = !(xn <= a.xslip);
        if (stick) ve[0] = 0.0;
    }
    const double pe = n1d3 * (4.0 * a.prim[4][g1] - a.prim[4][g2] - 2.0 * 0.0);
    double re, te;
    if (MODE == BCX_FIX) {
        // fixbc's states of faces 3 and 4 (the others are left alone)
        re = HI ? 1.0 : 2.0;
        for (int m = 0; m < 3; ++m) ve[m] = 0.0;
        const double pf = HI ? 2.5 : 1.0;
        te = pf / re * a.const2;
        a.prim[0][gi] = re;
        for (int m = 0; m < 3; ++m) a.prim[1 + m][gi] = 0.0;
        a.prim[4][gi] = pf;
        a.prim[5][gi] = te;
        a.q[0][gi] = re;
        a.q[1][gi] = re * 0.0;
        a.q[2][gi] = re * 0.0;
        a.q[3][gi] = re * 0.0;
        const double var0 = 0.5 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
        a.q[4][gi] = pf * a.const6 + re * var0;
        return;
    } else if (MODE == BCX_NOSLIP || stick) {
        te = a.tw[2 * AX + (HI ? 1 : 0)];
        re = pe / te * a.const2;
    } else if (RHO_EXT) {
        re = n1d3 * (4.0 * a.prim[0][g1] - a.prim[0][g2] - 2.0 * 0.0);
        te = pe / re * a.const2;
    } else {
        te = n1d3 * (4.0 * a.prim[5][g1] - a.prim[5][g2] - 2.0 * 0.0);
        re = pe / te * a.const2;
    }
    a.prim[0][gi] = re;
    for (int m = 0; m < 3; ++m) a.prim[1 + m][gi] = ve[m];
    a.prim[4][gi] = pe;
    a.prim[5][gi] = te;
    a.q[0][gi] = re;
    a.q[1][gi] = re * ve[0];
    a.q[2][gi] = re * ve[1];
    a.q[3][gi] = re * ve[2];
    const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
    a.q[4][gi] = TFORM ? re * (te * a.const1 + var1) : pe * a.const6 + re * var1;
}

// outflow (src/bc.F90:3379-3592), bctype 21, on face 2 (AX 0) or face 4
// (AX 1), the faces it acts on.  Face 2 copies velocity, pressure and
// temperature of node n-1 (the reference's zero-order branch, always
// taken), the density from thermal(p, T) = p/T*const2, q by fvar2q's
// temperature form.  Face 4 extrapolates velocity, pressure, temperature
// and density by extrapolate_2o with dv = 0; where the node's own normal
// velocity is at least its sound speed css = sqrt(T)/Ma it takes them as
// they are, else its pressure relaxes towards pinf = 1/const2,
// p = (p + alpha dt pinf + rho css (ve - v))/(1 + alpha dt), alpha = 0.25,
// with the node's own p, rho and v and the step's dt, and the density from
// (p, Te); then T = p/rho*const2 and q by the pressure form.
template <int AX>
__global__ void k_bc_outflow(BcArgs a)
{
    const long ea = (AX == 0 ? a.g.n[1] : a.g.n[0]) + 1L;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ea * (a.g.n[2] + 1L)) return;
    const int ca = (int)(t % ea), k = (int)(t / ea);
    const int n = a.g.n[AX];
    auto at = [&](int cd) { return AX == 0 ? gidx(a.g, cd, ca, k) : gidx(a.g, ca, cd, k); };
    const long gi = at(n), g1 = at(n - 1), g2 = at(n - 2);
    const double n1d3 = 1.0 / 3.0;
    double ve[3], pn, rn, tn;
    if (AX == 0) {
        for (int m = 0; m < 3; ++m) ve[m] = a.prim[1 + m][g1];
        pn = a.prim[4][g1];
        tn = a.prim[5][g1];
        rn = pn / tn * a.const2;
    } else {
        const double pinf = 1.0 / a.const2, alpha = 0.25;
        const double css = sqrt(a.prim[5][gi]) / a.mach;
        const double ub = a.prim[2][gi];
        for (int m = 0; m < 3; ++m)
            ve[m] = n1d3 * (4.0 * a.prim[1 + m][g1] - a.prim[1 + m][g2] - 2.0 * 0.0);
        const double pe = n1d3 * (4.0 * a.prim[4][g1] - a.prim[4][g2] - 2.0 * 0.0);
        const double te = n1d3 * (4.0 * a.prim[5][g1] - a.prim[5][g2] - 2.0 * 0.0);
        const double roe = n1d3 * (4.0 * a.prim[0][g1] - a.prim[0][g2] - 2.0 * 0.0);
        if (ub >= css) {
            pn = pe;
            rn = roe;
        } else {
            pn = (a.prim[4][gi] + alpha * a.dt * pinf + a.prim[0][gi] * css * (ve[1] - ub)) /
                 (1.0 + alpha * a.dt);
            rn = pn / te * a.const2;
        }
        tn = pn / rn * a.const2;
    }
    a.prim[0][gi] = rn;
    for (int m = 0; m < 3; ++m) a.prim[1 + m][gi] = ve[m];
    a.prim[4][gi] = pn;
    a.prim[5][gi] = tn;
    a.q[0][gi] = rn;
    a.q[1][gi] = rn * ve[0];
    a.q[2][gi] = rn * ve[1];
    a.q[3][gi] = rn * ve[2];
    const double var1 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
    a.q[4][gi] = AX == 0 ? rn * (tn * a.const1 + var1) : pn * a.const6 + rn * var1;
}

// back to the default contraction of the rest of the library
#pragma clang fp contract(fast)

// Plain copy of n doubles (16-byte accesses): the bandwidth reference for the
// roofline and the calibration of the HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, long n2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------- batched line-queue engine
//
// An ensemble of boxes advances through launches that each cover every
// member.  A sweep launch takes the members whose lines along its axis
// share one rows-per-lane count L (one template instance); its work items
// are tiles of lines of those members, listed in a device table in the
// order the autotuner chose.  In the persistent form a grid sized to the
// resident workgroups pulls items from a device counter until the table is
// exhausted; in the one-shot form workgroup b takes item b.  An item is a
// tile with all its quantities (the workgroup solves them in turn, so the
// inputs they share stay in its XCD's L2) or a tile and one quantity.

struct QItem {
    int member;  // index into the launch's argument array
    int tile;    // tile of that member (the plain sweep's tile numbering)
    int q;       // quantity, or -1: every quantity of the tile in turn
    int pad;
};

struct QueueLaunch {
    const SweepArgs* args;  // per member of the launch
    const QItem* items;
    int nitems;
    int* counter;           // persistent form: the launch's own counter (zero at start)
};

// A member's entry of a sweep argument table as the queue kernel reads it:
// in the constant address space (address space 4) in the device pass.  The
// host pass of this single-source file only checks the kernel's body, for
// which the plain type stands in.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const SweepArgs CArgs;
#else
typedef const SweepArgs CArgs;
#endif

// The argument table and the items are kernel parameters of their own,
// const and __restrict__: the compiler may then load a member's arguments
// once into scalar registers, knowing that the kernel's stores to the fields
// do not change them (through a pointer inside a struct it would reload
// them after every store).
template <int L, int PASS, bool XD>
__global__ __launch_bounds__(WGT, 4) void k_sweep_q(const SweepArgs* __restrict__ margs,
                                                    const QItem* __restrict__ items, int nitems,
                                                    int* counter)
{
    __shared__ int s_item;
    constexpr int NQ = PassCfg<PASS>::NQ;
    int next = blockIdx.x;
    for (;;) {
        if (counter) {
            if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
            __syncthreads();
            next = s_item;
            // every thread has read the item before thread 0 fetches the next
            __syncthreads();
        }
        const int item = __builtin_amdgcn_readfirstlane(next);
        if (item >= nitems) return;
        const QItem it = items[item];
        // the table is read through the constant address space: the loads
        // are then invariant like those of kernel arguments
        CArgs& a = ((CArgs*)margs)[__builtin_amdgcn_readfirstlane(it.member)];
        // one quantity, or every quantity of the tile in turn (one call
        // site, not unrolled: a copy of the body per quantity would not fit
        // the registers)
        const int mlo = it.q >= 0 ? it.q : 0, mhi = it.q >= 0 ? it.q + 1 : NQ;
#pragma unroll 1
        for (int m = mlo; m < mhi; ++m) {
            // the write-back's view is read through a pointer the compiler
            // cannot tie to the first (the table is invariant either way)
            CArgs* pw = &a;
            asm volatile("" : "+s"(pw));
            sweep_tile<L, PASS, XD>(a, *pw, it.tile, m);
            // the next quantity stages into the same LDS rows
            __syncthreads();
        }
        if (!counter) return;
        __syncthreads();
    }
}

// The pointwise kernels over every member in one launch: the members' blocks
// are numbered consecutively (bstart[k] is member k's first block), and a
// block runs the kernel body of its member with its own block number.
enum PtKind : int { PT_QSWAP_X = 0, PT_QSWAP_Y = 1, PT_QSWAP_Z = 2, PT_GFLUX = 3, PT_PRIM = 4,
                    PT_KINDS = 5 };

struct PtBatch {
    const PtArgs* args;
    const unsigned* bstart;  // nmem + 1 entries
    int nmem;
};

// (the tables are const __restrict__ kernel parameters, as in k_sweep_q)
template <int K>
__global__ __launch_bounds__(256) void k_pt_batch(const PtArgs* __restrict__ margs,
                                                  const unsigned* __restrict__ bstart, int nmem)
{
    const unsigned bid = blockIdx.x;
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {  // the last member whose first block is <= bid
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    const int mi = __builtin_amdgcn_readfirstlane(lo);
    const PtArgs& a = margs[mi];
    const unsigned lb = bid - bstart[mi];
    if constexpr (K == PT_QSWAP_X) qswap_x_body(a, lb);
    else if constexpr (K == PT_QSWAP_Y) qswap_jk_body<1>(a, lb);
    else if constexpr (K == PT_QSWAP_Z) qswap_jk_body<2>(a, lb);
    else if constexpr (K == PT_GFLUX) gflux_body<false>(a, lb);
    else prim_body(a, lb);
}

// zero n ints (the queue counters of a step; a kernel rather than a memset,
// so that the step can be captured into a graph)
__global__ void k_zero_ints(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

// ====================================================================== host

struct astr_ctx {
    astr_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    Geo g{};
    long pi = 0, pj = 0, pk = 0, nfield = 0;  // padded dims, elements per field
    double* base = nullptr;
    double *q[5], *prim[6], *qsave[5], *qrhs[5], *dvel[9], *dtmp[3], *gfl[15], *jac;
    double* swf[10] = {};           // compact upwind: split fluxes F+, F-
    double* fhf[10] = {};           // compact upwind: their interface fluxes
    double* cmp_ac[3][2] = {};      // compact upwind Thomas factors per axis, wind
    double len[3] = {1.0, 1.0, 0.0};  // gridcube lengths (astr_set_lengths)
    double twall[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};  // wall temperatures (astr_set_twall)
    double bc_dt = 0.0;  // time step of the latest whole step, which outflow reads
    double xslip = 0.0;  // slip-nonslip point (astr_set_xslip)
    bool curv = false;                // general metrics uploaded (dxf)
    bool juni = false;                // jacob uniform (j0): read from no array
    double j0 = 0.0;
    double* dxf[9] = {};              // dxi(d, e) per node, 3d+e
    int cmp_first[3] = {}, cmp_last[3] = {};
    bool lines_ok[3] = {};          // axis fits the compact line sweeps
    double h[3] = {0.0, 0.0, 0.0};  // uniform spacing of the mini-app variant
    bool have_h = false;
    double* ssf = nullptr;          // Ducros sensor (upwind path), haloed
    unsigned char* lshock = nullptr; // shock flags, one byte per padded cell
    double* scratch = nullptr;  // field in of the field-level operators
    double* scratch2 = nullptr; // field out of the field-level operators
    double* xi = nullptr;       // 3 metric lines, each LPL*LMAX+2*HM+1
    long xi_stride = 0;
    double* tabs = nullptr;
    LineTab tab[3][2];
    double* adj_base = nullptr; // discrete adjoint workspace (fields)
    long adj_nf = 0;            // fields it holds
    // the trajectory astr_adjoint_record kept there (rec_nsteps < 0: none)
    long rec_nsteps = -1, rec_ckpt = 1;
    double rec_dt = 0.0, rec_t_end = 0.0;
    double* tabs_t = nullptr;   // transposed systems and closure corrections
    LineTab tab_t[3][2];
    TransTab tt[3][2];
    int L[3];
    FiltCoef* fc = nullptr;
    double* part = nullptr;
    long npart = 0;
    std::vector<double> host_part;
    // per-step statistics of astr_rk3_steps on the device (3 per step,
    // STATS_SLOTS steps), copied to the host after the steps
    double* stats_dev = nullptr;
    double const1, const2, const3, const4, const5, const6, const7, tempconst, tempconst1;
    bool have_geom = false, have_state = false;
    int bctype[6] = {1, 1, 1, 1, 1, 1};  // boucon per face: 1 nothing done, 0 user
    int udf = 0;                          // the user condition (ASTR_UDF_*), 0 none
    double time = 0.0;                    // simulation time (commvar time)
    int timing = 0;  // 0 off, 1 per kernel family, 2 whole calls of astr_rk3_steps
    // bumped by the setters whose values kernels take by value (spacing,
    // boundary conditions, metrics): a batch graph captured before is stale
    unsigned long gen = 0;
    hipEvent_t ev[16] = {};
    double tms[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

// ---- line systems (host): fd_scheme_initiate, compact_filter_initiate and
// tridiagonal_thomas_proprocess restated, then the chunk products.
struct HostTab {
    int R, first, last;
    std::vector<double> t;  // 3 * LPL * L
};

void node_range(int ntype, int dim, int ext, int& first, int& last)
{
    first = (ntype == 1 || ntype == 4) ? 0 : -ext;
    last = (ntype == 2 || ntype == 4) ? dim : dim + ext;
}

// The LHS rows of a line system, row r: a[r] x(r-1) + x(r) + c[r] x(r+1)
// (a[0] and c[R-1] are outside the matrix)
void lhs_rows(int op, int ntype, int R, double alfa, std::vector<double>& a, std::vector<double>& c)
{
    a.assign(R, 0.0);
    c.assign(R, 0.0);
    if (op == OP_DERIV) {
        // src/derivative.F90:104-139 (643c)
        for (int r = 0; r < R; ++r) a[r] = c[r] = 1.0 / 3.0;
        a[0] = c[0] = 0.0;
        a[R - 1] = c[R - 1] = 0.0;
        if (ntype == 1 || ntype == 4) { a[0] = c[0] = 2.0; a[1] = c[1] = 0.25; }
        if (ntype == 2 || ntype == 4) { a[R - 1] = c[R - 1] = 2.0; a[R - 2] = c[R - 2] = 0.25; }
    } else {
        // src/filter.F90:44-96
        const double b0 = (ntype == 1 || ntype == 4) ? 0.98 : 1.11;
        const double bm = (ntype == 2 || ntype == 4) ? 0.98 : 1.11;
        for (int r = 0; r < R; ++r) a[r] = c[r] = alfa;
        a[0] = c[0] = b0;
        a[R - 1] = c[R - 1] = bm;
    }
}

// tridiagonal_thomas_proprocess, src/commfunc.F90:760-770, in the device
// table layout (ac1, ac2, ac3, LPL*L each)
void factor_rows(const std::vector<double>& a, const std::vector<double>& c, int L, HostTab& h)
{
    const int R = h.R;
    const int S = LPL * L;
    h.t.assign(3 * S, 0.0);
    double* ac1 = h.t.data();
    double* ac2 = ac1 + S;
    double* ac3 = ac1 + 2 * S;
    ac1[0] = c[0];
    ac2[0] = 1.0;
    ac3[0] = 0.0;
    for (int r = 1; r < R; ++r) {
        const double den = 1.0 - a[r] * ac1[r - 1];
        ac1[r] = c[r] / den;
        ac2[r] = 1.0 / den;
        ac3[r] = a[r] / den;
    }
    ac1[R - 1] = 0.0;  // the solver starts the back substitution at x(n)=d(n)
}

HostTab build_tab(int op, int ntype, int dim, double alfa, int L)
{
    HostTab h;
    node_range(ntype, dim, op == OP_DERIV ? 1 : 3, h.first, h.last);
    h.R = h.last - h.first + 1;
    std::vector<double> a, c;
    lhs_rows(op, ntype, h.R, alfa, a, c);
    factor_rows(a, c, L, h);
    return h;
}

// The transposed system A^T: sub-diagonal c[r-1], super-diagonal a[r+1]
HostTab build_tab_t(int op, int ntype, int dim, double alfa, int L)
{
    HostTab h;
    node_range(ntype, dim, op == OP_DERIV ? 1 : 3, h.first, h.last);
    const int R = h.last - h.first + 1;
    h.R = R;
    std::vector<double> a, c, at(R, 0.0), ct(R, 0.0);
    lhs_rows(op, ntype, R, alfa, a, c);
    for (int r = 1; r < R; ++r) at[r] = c[r - 1];
    for (int r = 0; r + 1 < R; ++r) ct[r] = a[r + 1];
    factor_rows(at, ct, L, h);
    return h;
}

// Coefficient of the interior right-hand-side stencil at offset off from
// the row's node: 643c (derivative) or the 10th-order filter
double b_interior(int op, const FiltCoef& f, int off)
{
    if (op == OP_DERIV) {
        switch (off) {
        case 1: return 7.0 / 9.0;
        case -1: return -7.0 / 9.0;
        case 2: return 1.0 / 36.0;
        case -2: return -1.0 / 36.0;
        default: return 0.0;
        }
    }
    const int k = off < 0 ? -off : off;
    if (k > 5) return 0.0;
    return k == 0 ? 2.0 * f.c10[0] : f.c10[k];
}

// The (node, coefficient) pairs of closure row r of the right-hand side, as
// rhs_deriv_edge and rhs_filter_edge evaluate it
std::vector<std::pair<int, double>> b_closure_row(int op, int ntype, int first, int last, int R,
                                                  int r, const FiltCoef& f)
{
    std::vector<std::pair<int, double>> v;
    const bool lob = ntype == 1 || ntype == 4, hib = ntype == 2 || ntype == 4;
    const int m = first + r;
    if (op == OP_DERIV) {
        if (r == 0 && lob) return {{m, -2.5}, {m + 1, 2.0}, {m + 2, 0.5}};
        if (r == R - 1 && hib) return {{m, 2.5}, {m - 1, -2.0}, {m - 2, -0.5}};
        if ((r == 1 && lob) || (r == R - 2 && hib)) return {{m + 1, 0.75}, {m - 1, -0.75}};
        return {{m + 1, 0.75}, {m - 1, -0.75}, {m + 2, -0.15}, {m - 2, 0.15},
                {m + 3, 1.0 / 60.0}, {m - 3, -1.0 / 60.0}};
    }
    if (r >= R - 3) {
        const int k = R - 1 - r;
        if (hib)
            for (int j = 0; j <= 6; ++j) v.push_back({last - j, f.cb[k][j]});
        else
            for (int j = 0; j <= 10; ++j) v.push_back({last + 2 - j, f.ch[k][j]});
        return v;
    }
    if (r <= 2) {
        if (lob)
            for (int j = 0; j <= 6; ++j) v.push_back({first + j, f.cb[r][j]});
        else
            for (int j = 0; j <= 10; ++j) v.push_back({first - 2 + j, f.ch[r][j]});
        return v;
    }
    if (r == 3 || r == R - 4) {
        v.push_back({m, 2.0 * f.c6[0]});
        for (int k = 1; k <= 3; ++k) { v.push_back({m + k, f.c6[k]}); v.push_back({m - k, f.c6[k]}); }
        return v;
    }
    v.push_back({m, 2.0 * f.c8[0]});
    for (int k = 1; k <= 4; ++k) { v.push_back({m + k, f.c8[k]}); v.push_back({m - k, f.c8[k]}); }
    return v;
}

// The closure-correction blocks of B^T for one line system: nlo rows at the
// low end (columns first-5 ..), nhi at the high end (columns last-10 ..)
void build_trans_blocks(int op, int ntype, const HostTab& h, const FiltCoef& f, int& nlo, int& nhi,
                        std::vector<double>& dlo, std::vector<double>& dhi)
{
    const bool lob = ntype == 1 || ntype == 4, hib = ntype == 2 || ntype == 4;
    if (op == OP_DERIV) { nlo = lob ? 2 : 1; nhi = hib ? 2 : 1; }
    else { nlo = lob ? 5 : 3; nhi = hib ? 5 : 3; }
    const int lo0 = h.first - 5, hi0 = h.last - 10;
    dlo.assign(nlo * TC, 0.0);
    dhi.assign(nhi * TC, 0.0);
    auto fill = [&](int r, int c0, double* row) {
        const int m = h.first + r;
        for (int c = 0; c < TC; ++c) row[c] = -b_interior(op, f, c0 + c - m);
        for (const auto& e : b_closure_row(op, ntype, h.first, h.last, h.R, r, f)) {
            const int c = e.first - c0;
            if (c >= 0 && c < TC) row[c] += e.second;
        }
    };
    for (int r = 0; r < nlo; ++r) fill(r, lo0, dlo.data() + r * TC);
    for (int k = 0; k < nhi; ++k) fill(h.R - 1 - k, hi0, dhi.data() + k * TC);
}

// compact_flux_initiate (scheme 543), src/flux.F90:32-108, and its Thomas
// factors (tridiagonal_thomas_proprocess, src/commfunc.F90:752-774), three
// per row as the sequential solver reads them
std::vector<double> build_flux_tab(int ntype, int dim, bool plus, double bf, int& first, int& last)
{
    first = (ntype == 1 || ntype == 4) ? -1 : -2;
    last = (ntype == 2 || ntype == 4) ? dim : dim + 1;
    const int n = last - first + 1;
    std::vector<double> a(n), c(n), ac(3 * n);
    for (int r = 0; r < n; ++r) {
        a[r] = plus ? 0.5 - (1.0 / 6.0) * bf : (1.0 / 6.0) + (1.0 / 6.0) * bf;
        c[r] = plus ? (1.0 / 6.0) + (1.0 / 6.0) * bf : 0.5 - (1.0 / 6.0) * bf;
    }
    a[0] = c[0] = 0.0;
    a[n - 1] = c[n - 1] = 0.0;
    if (ntype == 1 || ntype == 4) { a[0] = c[0] = 2.0; a[1] = c[1] = 0.25; }
    if (ntype == 2 || ntype == 4) { a[n - 1] = c[n - 1] = 2.0; a[n - 2] = c[n - 2] = 0.25; }
    ac[0] = c[0];
    ac[1] = 1.0;
    ac[2] = 0.0;
    for (int r = 1; r < n; ++r) {
        ac[3 * r + 0] = c[r] / (1.0 - a[r] * ac[3 * (r - 1)]);
        ac[3 * r + 1] = 1.0 / (1.0 - a[r] * ac[3 * (r - 1)]);
        ac[3 * r + 2] = a[r] / (1.0 - a[r] * ac[3 * (r - 1)]);
    }
    return ac;
}

// filter_coefficient_cal closed forms, src/filter.F90:303-431
FiltCoef build_coef(double al, double bh, double bb)
{
    FiltCoef f{};
    f.c6[0] = (11.0 + 10.0 * al) / 32.0;
    f.c6[1] = (15.0 + 34.0 * al) / 64.0;
    f.c6[2] = (-3.0 + 6.0 * al) / 32.0;
    f.c6[3] = (1.0 - 2.0 * al) / 64.0;
    f.c8[0] = (93.0 + 70.0 * al) / 256.0;
    f.c8[1] = (7.0 + 18.0 * al) / 32.0;
    f.c8[2] = (-7.0 + 14.0 * al) / 64.0;
    f.c8[3] = (1.0 - 2.0 * al) / 32.0;
    f.c8[4] = (-1.0 + 2.0 * al) / 256.0;
    f.c10[0] = (193.0 + 126.0 * al) / 512.0;
    f.c10[1] = (105.0 + 302.0 * al) / 512.0;
    f.c10[2] = (-15.0 + 30.0 * al) / 128.0;
    f.c10[3] = (45.0 - 90.0 * al) / 1024.0;
    f.c10[4] = (-5.0 + 10.0 * al) / 512.0;
    f.c10[5] = (1.0 - 2.0 * al) / 1024.0;
    const double cb[3][7] = {
        {(63.0 + 1.0 * bb) / 64.0, (3.0 + 29.0 * bb) / 32.0, (-15.0 + 15.0 * bb) / 64.0,
         (5.0 - 5.0 * bb) / 16.0, (-15.0 + 15.0 * bb) / 64.0, (3.0 - 3.0 * bb) / 32.0,
         (-1.0 + 1.0 * bb) / 64.0},
        {(1.0 + 62.0 * al) / 64.0, (29.0 + 6.0 * al) / 32.0, (15.0 + 34.0 * al) / 64.0,
         (-5.0 + 10.0 * al) / 16.0, (15.0 - 30.0 * al) / 64.0, (-3.0 + 6.0 * al) / 32.0,
         (1.0 - 2.0 * al) / 64.0},
        {(-1.0 + 2.0 * al) / 64.0, (3.0 + 26.0 * al) / 32.0, (49.0 + 30.0 * al) / 64.0,
         (5.0 + 6.0 * al) / 16.0, (-15.0 + 30.0 * al) / 64.0, (3.0 - 6.0 * al) / 32.0,
         (-1.0 + 2.0 * al) / 64.0}};
    const double ch[3][11] = {
        {(-1.0 + 1.0 * bh) / 1024.0, (5.0 - 5.0 * bh) / 512.0, (979.0 + 45.0 * bh) / 1024.0,
         (15.0 + 113.0 * bh) / 128.0, (-105.0 + 105.0 * bh) / 512.0, (63.0 - 63.0 * bh) / 256.0,
         (-105.0 + 105.0 * bh) / 512.0, (15.0 - 15.0 * bh) / 128.0, (-45.0 + 45.0 * bh) / 1024.0,
         (5.0 - 5.0 * bh) / 512.0, (-1.0 + 1.0 * bh) / 1024.0},
        {(1.0 - 2.0 * al) / 1024.0, (-5.0 + 10.0 * al) / 512.0, (45.0 + 934.0 * al) / 1024.0,
         (113.0 + 30.0 * al) / 128.0, (105.0 + 302.0 * al) / 512.0, (-63.0 + 126.0 * al) / 256.0,
         (105.0 - 210.0 * al) / 512.0, (-15.0 + 30.0 * al) / 128.0, (45.0 - 90.0 * al) / 1024.0,
         (-5.0 + 10.0 * al) / 512.0, (1.0 - 2.0 * al) / 1024.0},
        {(-1.0 + 2.0 * al) / 1024.0, (5.0 - 10.0 * al) / 512.0, (-45.0 + 90.0 * al) / 1024.0,
         (15.0 + 98.0 * al) / 128.0, (407.0 + 210.0 * al) / 512.0, (63.0 + 130.0 * al) / 256.0,
         (-105.0 + 210.0 * al) / 512.0, (15.0 - 30.0 * al) / 128.0, (-45.0 + 90.0 * al) / 1024.0,
         (5.0 - 10.0 * al) / 512.0, (-1.0 + 2.0 * al) / 1024.0}};
    memcpy(f.cb, cb, sizeof cb);
    memcpy(f.ch, ch, sizeof ch);
    return f;
}

PtArgs pt_args(astr_ctx* c)
{
    PtArgs a{};
    a.g = c->g;
    for (int m = 0; m < 5; ++m) { a.q[m] = c->q[m]; a.qsave[m] = c->qsave[m]; a.qrhs[m] = c->qrhs[m]; }
    for (int m = 0; m < 6; ++m) a.prim[m] = c->prim[m];
    for (int m = 0; m < 15; ++m) a.gfl[m] = c->gfl[m];
    for (int d = 0; d < 3; ++d) a.xi[d] = c->xi + d * c->xi_stride;
    a.visc = c->p.diffterm;
    a.mini = c->p.variant == 1;
    a.curv = c->curv;
    for (int m = 0; m < 9; ++m) a.dxf[m] = c->dxf[m];
    a.juni = c->juni;
    a.j0 = c->j0;
    a.per[0] = c->p.npdci == 3;
    a.per[1] = c->p.npdcj == 3;
    a.per[2] = c->p.npdck == 3;
    for (int m = 0; m < 9; ++m) a.dvel[m] = c->dvel[m];
    for (int m = 0; m < 3; ++m) a.dtmp[m] = c->dtmp[m];
    a.jac = c->jac;
    a.const1 = c->const1; a.const2 = c->const2; a.const5 = c->const5; a.const6 = c->const6;
    a.reynolds = c->p.reynolds; a.prandtl = c->p.prandtl;
    a.tempconst = c->tempconst; a.tempconst1 = c->tempconst1;
    a.part = c->part;
    return a;
}

// LDS row of a staged line: guard + staged nodes including the clamped tail
// (x: 32(L+1) slots, y/z: 16(2L+1)), which also covers the solver's window
// reads up to slot 32L+15; odd stride against bank conflicts
int lds_row(int dir, int L)
{
    return (TG + (dir == 0 ? 32 * (L + 1) : 16 * (2 * L + 1))) | 1;
}

size_t lds_bytes(int dir, int L)
{
    return sizeof(double) * (dir == 0 ? TXX : TXY) * (size_t)lds_row(dir, L);
}

SweepArgs sweep_args(astr_ctx* c, int dir, int pass)
{
    SweepArgs a{};
    a.g = c->g;
    a.dir = dir;
    a.pass = pass;
    const bool filt = pass == PASS_FILTER || pass == PASS_FILTER_FIELD;
    a.tab = c->tab[dir][filt ? 1 : 0];
    a.fc = c->fc;
    const int n0 = c->g.n[0], n1 = c->g.n[1], n2 = c->g.n[2];
    if (dir == 0) { a.na = n1 + 1; a.nb = n2 + 1; }
    else if (dir == 1) { a.na = n0 + 1; a.nb = n2 + 1; }
    else { a.na = n0 + 1; a.nb = n1 + 1; }
    const int tx = dir == 0 ? TXX : TXY;
    a.ntile_a = (a.na + tx - 1) / tx;
    const int dim = c->g.n[dir];
    a.rs_in = lds_row(dir, c->L[dir]);
    for (int m = 0; m < 5; ++m) { a.q[m] = c->q[m]; a.qo[m] = c->q[m]; a.qrhs[m] = c->qrhs[m]; }
    for (int m = 0; m < 6; ++m) a.prim[m] = c->prim[m];
    for (int m = 0; m < 9; ++m) a.dvel[m] = c->dvel[m];
    for (int m = 0; m < 3; ++m) a.dtmp[m] = c->dtmp[m];
    for (int m = 0; m < 5; ++m) a.gf[m] = c->gfl[5 * dir + m];
    a.jac = c->jac;
    a.xi = c->xi + dir * c->xi_stride;
    for (int m = 0; m < 5; ++m) a.qsave[m] = c->qsave[m];
    a.visc = c->p.diffterm;
    a.mini = c->p.variant == 1;
    a.raw = pass == PASS_GRAD && c->curv;
    a.juni = c->juni;
    a.j0 = c->j0;
    a.h = c->h[dir];
    return a;
}

template <int PASS>
int launch_sweep_p(astr_ctx* c, const SweepArgs& a, hipStream_t st)
{
    const int L = c->L[a.dir];
    const size_t lds = lds_bytes(a.dir, L);
    const int nq = PassCfg<PASS>::NQ, ntiles = a.ntile_a * a.nb;
    const dim3 grid((ntiles + 7) / 8 * 8 * nq);
    switch (L) {
#define CASE(LL)                                                                        \
    case LL:                                                                            \
        if (a.dir == 0)                                                                 \
            hipLaunchKernelGGL((k_sweep<LL, PASS, true>), grid, dim3(WGT), lds, st, a); \
        else                                                                            \
            hipLaunchKernelGGL((k_sweep<LL, PASS, false>), grid, dim3(WGT), lds, st, a); \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default:
        return fail(ASTR_E_UNSUPPORTED, "line too long");
    }
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// st: the stream of the launch (default: the context's)
int launch_sweep(astr_ctx* c, const SweepArgs& a, hipStream_t st = nullptr)
{
    if (!c->lines_ok[a.dir])
        return fail(ASTR_E_UNSUPPORTED, "axis too long (or absent) for the compact line operators");
    if (!st) st = c->stream;
    switch (a.pass) {
    case PASS_FILTER: return launch_sweep_p<PASS_FILTER>(c, a, st);
    case PASS_GRAD: return launch_sweep_p<PASS_GRAD>(c, a, st);
    case PASS_RHS: return launch_sweep_p<PASS_RHS>(c, a, st);
    case PASS_DERIV_FIELD: return launch_sweep_p<PASS_DERIV_FIELD>(c, a, st);
    case PASS_FILTER_FIELD: return launch_sweep_p<PASS_FILTER_FIELD>(c, a, st);
    }
    return fail(ASTR_E_ARG, "unknown pass");
}


// ---- transposed sweeps

// LDS row of a transposed sweep: TGT zero guard slots + the staged nodes
int lds_row_t(int dir, int L)
{
    return (TGT + (dir == 0 ? 32 * (L + 1) : 16 * (2 * L + 1))) | 1;
}

TSweepArgs tsweep_args(astr_ctx* c, int dir, int op)
{
    TSweepArgs a{};
    a.g = c->g;
    a.dir = dir;
    a.tab = c->tab_t[dir][op];
    a.tt = c->tt[dir][op];
    a.fc = c->fc;
    const int n0 = c->g.n[0], n1 = c->g.n[1], n2 = c->g.n[2];
    if (dir == 0) { a.na = n1 + 1; a.nb = n2 + 1; }
    else if (dir == 1) { a.na = n0 + 1; a.nb = n2 + 1; }
    else { a.na = n0 + 1; a.nb = n1 + 1; }
    const int tx = dir == 0 ? TXX : TXY;
    a.ntile_a = (a.na + tx - 1) / tx;
    a.rs_in = lds_row_t(dir, c->L[dir]);
    return a;
}

template <int OP>
int launch_sweep_t_op(astr_ctx* c, const TSweepArgs& a)
{
    const int L = c->L[a.dir];
    const size_t lds = sizeof(double) * (a.dir == 0 ? TXX : TXY) * (size_t)a.rs_in;
    const int ntiles = a.ntile_a * a.nb;
    const dim3 grid((ntiles + 7) / 8 * 8 * a.nq);
    switch (L) {
#define CASE(LL)                                                                              \
    case LL:                                                                                  \
        if (a.dir == 0)                                                                       \
            hipLaunchKernelGGL((k_sweep_t<LL, OP, true>), grid, dim3(WGT), lds, c->stream, a); \
        else                                                                                  \
            hipLaunchKernelGGL((k_sweep_t<LL, OP, false>), grid, dim3(WGT), lds, c->stream, a); \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default:
        return fail(ASTR_E_UNSUPPORTED, "line too long");
    }
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// fout[m] (+)= the transpose of the compact derivative (op OP_DERIV) or
// filter (OP_FILTER) along dir applied to fin[m], m < nq, in one launch;
// see TSweepArgs for wrap and xi
int launch_sweep_t(astr_ctx* c, int dir, int op, int nq, const double* const* fin,
                   double* const* fout, const double* xi, bool wrap, bool accumulate)
{
    if (!c->lines_ok[dir])
        return fail(ASTR_E_UNSUPPORTED, "axis too long (or absent) for the compact line operators");
    TSweepArgs a = tsweep_args(c, dir, op);
    a.nq = nq;
    for (int m = 0; m < nq; ++m) {
        a.fin[m] = fin[m];
        a.fout[m] = fout[m];
    }
    a.xi = xi;
    a.wrap = wrap;
    a.accumulate = accumulate;
    return op == OP_DERIV ? launch_sweep_t_op<OP_DERIV>(c, a) : launch_sweep_t_op<OP_FILTER>(c, a);
}

long n_interior(const astr_ctx* c)
{
    return (long)(c->g.n[0] + 1) * (c->g.n[1] + 1) * (c->g.n[2] + 1);
}

int launch_pt(astr_ctx* c, void (*k)(PtArgs), long nthreads, const PtArgs& a, int block = 256)
{
    if (nthreads <= 0) return ASTR_OK;
    const long nb = (nthreads + block - 1) / block;
    hipLaunchKernelGGL(k, dim3((unsigned)nb), dim3(block), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// k_prim over the interior rows cut at the end of the line holding node n
// (skip_rho: the density left as it is)
int launch_prim(astr_ctx* c, bool skip_rho = false)
{
    const long n = (long)(c->g.n[1] + 1) * (c->g.n[2] + 1) * line_end_i(c->g.n[0]);
    const dim3 grid((unsigned)((n + 255) / 256));
    PtArgs a = pt_args(c);
    a.skip_rho = skip_rho;
    hipLaunchKernelGGL(k_prim, grid, dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

int set_dev(astr_ctx* c)
{
    HIPCHK(hipSetDevice(c->device));
    return ASTR_OK;
}

#define TRY(x)                      \
    do {                            \
        int rc_ = (x);              \
        if (rc_ != ASTR_OK) return rc_; \
    } while (0)

int do_filterq(astr_ctx* c)
{
    for (int d = 0; d < 3; ++d) TRY(launch_sweep(c, sweep_args(c, d, PASS_FILTER)));
    return ASTR_OK;
}

int do_qswap(astr_ctx* c)
{
    PtArgs a = pt_args(c);
    const int n0 = c->g.n[0], n1 = c->g.n[1], n2 = c->g.n[2];
    if (n2 == 0) {
        // two-dimensional box: periodic in-plane axes, then the k part
        const int nts[2] = {c->p.npdci, c->p.npdcj};
        for (int d = 0; d < 2; ++d) {
            if (nts[d] != 3) continue;
            a.axis = d;
            TRY(launch_pt(c, k_qswap2d_line, (long)c->g.n[1 - d] + 1, a));
        }
        return launch_pt(c, k_qswap2d_plane, (long)(n0 + 1) * (n1 + 1), a);
    }
    // each part runs on a periodic axis only (the single-rank branches of
    // qswap copy nothing on a bounded one, src/parallel.F90:4877-4915)
    if (c->p.npdci == 3) TRY(launch_pt(c, k_qswap_x, qswap_threads(c->g, 0), a));
    if (c->p.npdcj == 3) TRY(launch_pt(c, k_qswap_y, qswap_threads(c->g, 1), a));
    if (c->p.npdck == 3) TRY(launch_pt(c, k_qswap_z, qswap_threads(c->g, 2), a));
    return ASTR_OK;
}

// boucon (src/bc.F90:327-411) for the faces under the user condition of
// the 2-D Riemann example, at the context's time
// k_bc_extrap's condition MODE on face f (0..5), one thread per node of the
// face
template <int MODE>
static int launch_extrap(astr_ctx* c, const BcArgs& a, int f)
{
    const int ax = f / 2, oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
    const long n = (c->g.n[oa] + 1L) * (c->g.n[ob] + 1L);
    const dim3 grid((unsigned)((n + 255) / 256));
    switch (f) {
    case 0: hipLaunchKernelGGL((k_bc_extrap<0, false, MODE>), grid, dim3(256), 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL((k_bc_extrap<0, true, MODE>), grid, dim3(256), 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_bc_extrap<1, false, MODE>), grid, dim3(256), 0, c->stream, a); break;
    case 3: hipLaunchKernelGGL((k_bc_extrap<1, true, MODE>), grid, dim3(256), 0, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_bc_extrap<2, false, MODE>), grid, dim3(256), 0, c->stream, a); break;
    default: hipLaunchKernelGGL((k_bc_extrap<2, true, MODE>), grid, dim3(256), 0, c->stream, a); break;
    }
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

int do_boucon(astr_ctx* c)
{
    bool any = c->udf == ASTR_UDF_RIEMANN2D;
    for (int f = 0; f < 6; ++f)
        any = any || c->bctype[f] == 50 || c->bctype[f] == 60 || c->bctype[f] == 23 ||
              c->bctype[f] == 41 || c->bctype[f] == 42 || c->bctype[f] == 21 ||
              c->bctype[f] == 31 || c->bctype[f] == 421 || c->bctype[f] == 411;
    if (!any) return ASTR_OK;
    BcArgs a{};
    a.g = c->g;
    for (int m = 0; m < 5; ++m) a.q[m] = c->q[m];
    for (int m = 0; m < 6; ++m) a.prim[m] = c->prim[m];
    for (int f = 0; f < 4; ++f) a.face[f] = c->udf == ASTR_UDF_RIEMANN2D && c->bctype[f] == 0;
    a.time = c->time;
    a.gamma = c->p.gamma;
    a.mach = c->p.mach;
    a.const1 = c->const1;
    a.const2 = c->const2;
    a.const6 = c->const6;
    a.len[0] = c->len[0];
    a.len[1] = c->len[1];
    a.xslip = c->xslip;
    for (int f = 0; f < 6; ++f) a.tw[f] = c->twall[f];
    a.dt = c->bc_dt;
    // boucon's face order (src/bc.F90:338-404), face by face 1..6: the user
    // condition of faces 1 and 2 (and of 3 and 4) in one kernel (it is never
    // combined with zeroextrap or symmetry), gcnscbc on faces 2 and 4,
    // zeroextrap and symmetry on any
    const long nk = c->g.n[2] + 1;
    const long ni = (long)(c->g.n[0] + 1) * nk, nj = (long)(c->g.n[1] + 1) * nk;
    for (int f = 0; f < 6; ++f) {
        if (f == 0 && (a.face[0] || a.face[1])) {
            hipLaunchKernelGGL(k_bc_riem2d_i, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (f == 2 && (a.face[2] || a.face[3])) {
            hipLaunchKernelGGL(k_bc_riem2d_j, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (c->bctype[f] == 50) TRY(launch_extrap<BCX_ZEROEXTRAP>(c, a, f));
        if (c->bctype[f] == 60) TRY(launch_extrap<BCX_SYMMETRY>(c, a, f));
        if (c->bctype[f] == 41) TRY(launch_extrap<BCX_NOSLIP>(c, a, f));
        if (c->bctype[f] == 31 && (f == 2 || f == 3)) TRY(launch_extrap<BCX_FIX>(c, a, f));
        if (c->bctype[f] == 421) TRY(launch_extrap<BCX_SLIPADIAB>(c, a, f));
        if (c->bctype[f] == 411 && f == 2) TRY(launch_extrap<BCX_SLIPISOT>(c, a, f));
        if (c->bctype[f] == 411 && f == 3) TRY(launch_extrap<BCX_NOSLIP>(c, a, f));
        if (c->bctype[f] == 21 && f == 1) {
            hipLaunchKernelGGL(k_bc_outflow<0>, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (c->bctype[f] == 21 && f == 3) {
            hipLaunchKernelGGL(k_bc_outflow<1>, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (c->bctype[f] == 42) TRY(launch_extrap<BCX_ADIABATIC>(c, a, f));
        if (c->bctype[f] == 23 && f == 1) {
            hipLaunchKernelGGL(k_bc_gcnscbc<0>, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
        if (c->bctype[f] == 23 && f == 3) {
            hipLaunchKernelGGL(k_bc_gcnscbc<1>, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0,
                               c->stream, a);
            HIPCHK(hipGetLastError());
        }
    }
    return ASTR_OK;
}

// gradcal, then sigma/qflux, then the flux derivatives along x, y, z; the
// z sweep either finishes qrhs (rk_stage 0) or applies the RK3 combine of
// that stage to q (rk_stage 1..3), after which k_prim refreshes the
// primitives.
const double RKC[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

int do_flux_sweeps(astr_ctx* c, int rk_stage, double dt, double* const* qo = nullptr,
                   bool qsraw = false)
{
    for (int d = 0; d < 3; ++d) {
        SweepArgs a = sweep_args(c, d, PASS_RHS);
        if (qo)
            for (int m = 0; m < 5; ++m) a.qo[m] = qo[m];
        if (d == 2 && rk_stage > 0) {
            a.qsraw = qsraw;
            a.rk_stage = rk_stage;
            a.rka = RKC[rk_stage - 1][0];
            a.rkb = RKC[rk_stage - 1][1];
            a.rkc = RKC[rk_stage - 1][2];
            a.dt = dt;
        }
        TRY(launch_sweep(c, a));
    }
    return ASTR_OK;
}

// sigma/qflux and the flux G = Fv - F of each axis (interior and the halo
// strips the flux sweeps read)
// workgroups of k_gflux: one thread per cell of the interior (j, k) rows, i
// from -16 to line_end_i(n + HM)
long gflux_blocks(const Geo& g)
{
    const long rows = (long)(g.n[1] + 1) * (g.n[2] + 1);
    return (rows * (line_end_i(g.n[0] + HM) + 16) + 255) / 256;
}

int stats_finish(astr_ctx* c, long nb, double* dst);

// stats_slot (nullable, device memory): also the statistics of the state
// (rkfirst), summed by the same workgroups and reduced into the slot
int do_gflux(astr_ctx* c, double* stats_slot = nullptr)
{
    const long nb = gflux_blocks(c->g);
    if (stats_slot) {
        hipLaunchKernelGGL(k_gflux<true>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                           pt_args(c));
        HIPCHK(hipGetLastError());
        return stats_finish(c, nb, stats_slot);
    } else {
        hipLaunchKernelGGL(k_gflux<false>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                           pt_args(c));
    }
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

int ndims_of(const astr_ctx* c) { return c->g.n[2] == 0 ? 2 : 3; }

long n_haloed(const astr_ctx* c)
{
    return (long)(c->g.n[0] + 1 + 2 * HM) * (c->g.n[1] + 1 + 2 * HM) * (c->g.n[2] + 1 + 2 * HM);
}

// ducrossensor runs for recon_schem 5 or lchardecomp with the explicit
// upwind scheme (src/solver.F90:227), for lchardecomp with the compact one
// (src/solver.F90:233)
bool sensor_on(const astr_ctx* c)
{
    if (c->p.convection == 1) return c->p.recon_schem == 5 || c->p.lchardecomp;
    if (c->p.convection == 2) return c->p.lchardecomp != 0;
    return false;
}

// gradcal with the explicit derivative (difschm '643e')
int do_grad_explicit(astr_ctx* c)
{
    GradXArgs a{};
    a.g = c->g;
    a.ntype[0] = c->p.npdci;
    a.ntype[1] = c->p.npdcj;
    a.ntype[2] = c->p.npdck;
    a.ndims = ndims_of(c);
    for (int m = 0; m < 6; ++m) a.prim[m] = c->prim[m];
    for (int d = 0; d < 3; ++d) a.xi[d] = c->xi + d * c->xi_stride;
    for (int m = 0; m < 9; ++m) a.dvel[m] = c->dvel[m];
    for (int d = 0; d < 3; ++d) a.dtmp[d] = c->dtmp[d];
    const long n = n_interior(c);
    hipLaunchKernelGGL(k_grad_explicit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// ducrossensor: ssf, its halo swap, the windowed flag
int do_sensor(astr_ctx* c)
{
    SensorArgs a{};
    a.g = c->g;
    a.ntype[0] = c->p.npdci;
    a.ntype[1] = c->p.npdcj;
    a.ntype[2] = c->p.npdck;
    a.ndims = ndims_of(c);
    a.shkcrt = c->p.shkcrt;
    for (int m = 0; m < 9; ++m) a.dvel[m] = c->dvel[m];
    a.prs = c->prim[4];
    a.ssf = c->ssf;
    a.lshock = c->lshock;
    const long n = n_interior(c);
    const long nh = a.ndims == 2 ? (long)(c->g.n[0] + 1 + 2 * HM) * (c->g.n[1] + 1 + 2 * HM)
                                 : n_haloed(c);
    hipLaunchKernelGGL(k_ducros_ssf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_ssf_halo, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_ducros_flag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// is..ie of an axis on one domain (src/parallel.F90:1043-1160): ntype 1
// starts at 1, ntype 2 ends at n-1; ntype 4 (bounded at both ends) is left
// unset by the reference on one rank and taken as 1..n-1
void uwd_range(int ntype, int n, int& lo, int& hi)
{
    lo = (ntype == 1 || ntype == 4) ? 1 : 0;
    hi = (ntype == 2 || ntype == 4) ? n - 1 : n;
}

// qrhs=0, then convrsduwd (convection 1) or convrsdcmp (convection 2)
// along each axis, the last one negating
int do_upwind(astr_ctx* c)
{
    const bool cmp = c->p.convection == 2;
    const int nd = ndims_of(c);
    const int nts[3] = {c->p.npdci, c->p.npdcj, c->p.npdck};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) uwd_range(nts[d], c->g.n[d], lo[d], hi[d]);
    if (nd == 2) lo[2] = hi[2] = 0;
    UwdArgs a{};
    a.g = c->g;
    a.ndims = nd;
    a.recon = c->p.recon_schem;
    a.lchar = c->p.lchardecomp;
    a.sensor = sensor_on(c) ? 1 : 0;
    a.bfac = c->p.bfacmpld;
    a.gamma = c->p.gamma;
    a.mach = c->p.mach;
    for (int m = 0; m < 5; ++m) { a.q[m] = c->q[m]; a.qrhs[m] = c->qrhs[m]; }
    for (int m = 0; m < 6; ++m) a.prim[m] = c->prim[m];
    a.jac = c->jac;
    a.lshock = c->lshock;
    const long n = n_interior(c);
    hipLaunchKernelGGL(k_zero_qrhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    for (int d = 0; d < nd; ++d) {
        a.dir = d;
        a.ntype = nts[d];
        a.lo = lo[d];
        a.hi = hi[d];
        a.iss = (nts[d] == 1 || nts[d] == 4) ? 0 : -HM;
        a.iee = (nts[d] == 2 || nts[d] == 4) ? c->g.n[d] : c->g.n[d] + HM;
        const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
        a.a0 = lo[oa];
        a.a1 = hi[oa];
        a.b0 = lo[ob];
        a.b1 = hi[ob];
        a.xi = c->xi + d * c->xi_stride;
        a.last = d == nd - 1;
        const int S = d == 0 ? 256 : 16, TL = d == 0 ? 1 : 16;
        a.ntile_line = (hi[d] - lo[d] + 1 + S - 2) / (S - 1);
        a.ntile_a = (a.a1 - a.a0 + 1 + TL - 1) / TL;
        const long nblk = (long)a.ntile_line * a.ntile_a * (a.b1 - a.b0 + 1);
        if (nblk <= 0) continue;
        if (cmp) {
            // split fluxes and their compact interface fluxes first
            CmpArgs ca{};
            ca.g = c->g;
            ca.dir = d;
            ca.ntype = nts[d];
            ca.ndims = nd;
            ca.iss = a.iss;
            ca.iee = a.iee;
            ca.a0 = a.a0;
            ca.a1 = a.a1;
            ca.b0 = a.b0;
            ca.b1 = a.b1;
            ca.na = a.a1 - a.a0 + 1;
            ca.first = c->cmp_first[d];
            ca.last = c->cmp_last[d];
            ca.bfac = c->p.bfacmpld;
            ca.gamma = c->p.gamma;
            ca.mach = c->p.mach;
            for (int m = 0; m < 5; ++m) ca.q[m] = c->q[m];
            for (int m = 0; m < 6; ++m) ca.prim[m] = c->prim[m];
            ca.jac = c->jac;
            ca.xi = a.xi;
            for (int m = 0; m < 10; ++m) {
                ca.sw[m] = c->swf[m];
                ca.fh[m] = c->fhf[m];
            }
            ca.ac[0] = c->cmp_ac[d][0];
            ca.ac[1] = c->cmp_ac[d][1];
            ca.comp0 = 0;
            ca.ncomp = 10;
            const long nl = (long)ca.na * (ca.b1 - ca.b0 + 1);
            const long nsw = nl * (ca.iee - ca.iss + 1);
            hipLaunchKernelGGL(k_sw_field, dim3((unsigned)((nsw + 255) / 256)), dim3(256), 0,
                               c->stream, ca);
            hipLaunchKernelGGL(k_flux_compact, dim3((unsigned)((10 * nl + 127) / 128)), dim3(128),
                               0, c->stream, ca);
            HIPCHK(hipGetLastError());
            for (int m = 0; m < 5; ++m) {
                a.fhp[m] = c->fhf[m];
                a.fhm[m] = c->fhf[5 + m];
            }
            if (d == 0)
                hipLaunchKernelGGL((k_uwd<true, true>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_uwd<false, true>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
        } else {
            const bool ld = c->p.recon_schem == 5;
            if (d == 0 && ld)
                hipLaunchKernelGGL((k_uwd<true, false, true>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
            else if (d == 0)
                hipLaunchKernelGGL((k_uwd<true, false>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
            else if (ld)
                hipLaunchKernelGGL((k_uwd<false, false, true>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_uwd<false, false>), dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
        }
        HIPCHK(hipGetLastError());
    }
    return ASTR_OK;
}

int do_grad(astr_ctx* c)
{
    if (c->p.deriv_explicit) return do_grad_explicit(c);
    for (int d = 0; d < 3; ++d) TRY(launch_sweep(c, sweep_args(c, d, PASS_GRAD)));
    if (c->curv) {
        MetricArgs m{};
        m.g = c->g;
        for (int q = 0; q < 9; ++q) m.dv[q] = c->dvel[q];
        for (int q = 0; q < 3; ++q) m.dv[9 + q] = c->dtmp[q];
        for (int q = 0; q < 9; ++q) m.dxf[q] = c->dxf[q];
        const long n = (long)(c->g.n[1] + 1) * (c->g.n[2] + 1) * line_end_i(c->g.n[0]);
        hipLaunchKernelGGL(k_metric_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, m);
        HIPCHK(hipGetLastError());
    }
    return ASTR_OK;
}

int time_mark(astr_ctx* c, int slot);

int do_rhscal(astr_ctx* c, bool timed = false)
{
    if (c->p.convection != 0) {
        // timing families of an rhscal call: 2 gradient, 8 shock sensor,
        // 5 interface-flux sweeps, 7 total
        if (timed) TRY(time_mark(c, 0));
        TRY(do_grad(c));
        if (timed) TRY(time_mark(c, 1));
        if (sensor_on(c)) TRY(do_sensor(c));
        if (timed) TRY(time_mark(c, 2));
        TRY(do_upwind(c));
        if (timed) {
            TRY(time_mark(c, 3));
            HIPCHK(hipEventSynchronize(c->ev[3]));
            const int slot[3] = {2, 8, 5};
            for (int f = 0; f < 3; ++f) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, c->ev[f], c->ev[f + 1]));
                c->tms[slot[f]] += ms;
                c->tms[7] += ms;
            }
        }
        return ASTR_OK;
    }
    TRY(do_grad(c));
    TRY(do_gflux(c));
    return do_flux_sweeps(c, 0, 0.0);
}

int do_rk(astr_ctx* c, int stage, double dt)
{
    PtArgs a = pt_args(c);
    a.stage = stage;
    a.rka = RKC[stage - 1][0];
    a.rkb = RKC[stage - 1][1];
    a.rkc = RKC[stage - 1][2];
    a.dt = dt;
    return launch_pt(c, k_rk, n_interior(c), a);
}

// the nb partials in c->part reduced to the three statistics in dst (3
// doubles of device memory): STATS_FIN workgroups sum slices of them into
// the second-level partials after the first npart, one workgroup sums those
constexpr int STATS_FIN = 256;
int stats_finish(astr_ctx* c, long nb, double* dst)
{
    const double n3 = (double)c->g.n[0] * (double)c->g.n[1] * (double)c->g.n[2];
    double* part2 = c->part + 6 * c->npart;
    const long chunk = (nb + STATS_FIN - 1) / STATS_FIN;
    hipLaunchKernelGGL(k_stats_fin, dim3(STATS_FIN), dim3(256), 0, c->stream,
                       (const double*)c->part, nb, chunk, part2, n3, (double*)nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_stats_fin, dim3(1), dim3(256), 0, c->stream, (const double*)part2,
                       (long)STATS_FIN, (long)STATS_FIN, (double*)nullptr, n3, dst);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// k_stats and the finish: the three statistics of the current state into
// dst (3 doubles of device memory), queued on the solver's stream
int launch_stats(astr_ctx* c, double* dst)
{
    PtArgs a = pt_args(c);
    const long n = (long)c->g.n[0] * c->g.n[1] * c->g.n[2];
    TRY(launch_pt(c, k_stats, n, a, 256));
    return stats_finish(c, (n + 255) / 256, dst);
}

int do_stats(astr_ctx* c, double* out)
{
    TRY(launch_stats(c, c->stats_dev));
    HIPCHK(hipMemcpyAsync(out, c->stats_dev, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ASTR_OK;
}

// host haloed Fortran array <-> device padded field
int upload_field(astr_ctx* c, double* dev, const double* host, std::vector<double>& stage)
{
    // work queued on the solver's stream may still read the field
    HIPCHK(hipStreamSynchronize(c->stream));
    const int ni = c->g.n[0] + 1 + 2 * HM, nj = c->g.n[1] + 1 + 2 * HM, nk = c->g.n[2] + 1 + 2 * HM;
    stage.assign(c->nfield, 0.0);
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            memcpy(&stage[(OI - HM) + (long)j * c->pi + (long)k * c->pi * c->pj],
                   host + (long)ni * (j + (long)nj * k), sizeof(double) * ni);
    HIPCHK(hipMemcpy(dev, stage.data(), sizeof(double) * c->nfield, hipMemcpyHostToDevice));
    return ASTR_OK;
}

int download_field(astr_ctx* c, const double* dev, double* host, std::vector<double>& stage,
                   bool interior)
{
    stage.resize(c->nfield);
    HIPCHK(hipMemcpy(stage.data(), dev, sizeof(double) * c->nfield, hipMemcpyDeviceToHost));
    if (interior) {
        const int ni = c->g.n[0] + 1, nj = c->g.n[1] + 1, nk = c->g.n[2] + 1;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                memcpy(host + (long)ni * (j + (long)nj * k),
                       &stage[c->g.org + (long)j * c->pi + (long)k * c->pi * c->pj],
                       sizeof(double) * ni);
    } else {
        const int ni = c->g.n[0] + 1 + 2 * HM, nj = c->g.n[1] + 1 + 2 * HM, nk = c->g.n[2] + 1 + 2 * HM;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                memcpy(host + (long)ni * (j + (long)nj * k),
                       &stage[(OI - HM) + (long)j * c->pi + (long)k * c->pi * c->pj],
                       sizeof(double) * ni);
    }
    return ASTR_OK;
}

int time_mark(astr_ctx* c, int slot)
{
    if (c->timing == 1) HIPCHK(hipEventRecord(c->ev[slot], c->stream));
    return ASTR_OK;
}

// alt (whole steps of one box, central path): stage 1's RK3 combine writes
// its result into the qsave buffer and leaves the stage's state where it
// was read, and the two buffers swap roles (q <-> qsave) for the rest of the
// step, so no qsave = q J is stored (5 fields per step fewer).  The
// buffers then alternate from step to step; a generation bump makes a
// batch that holds this context rebuild its launches.
int do_step(astr_ctx* c, double dt, double* stats_slot, bool timed, bool alt = false)
{
    c->bc_dt = dt;  // boucon's outflow reads the step's deltat
    // timing families: 0 filter, 1 boucon and qswap, 2 grad, 3 central:
    // sigma and the axis fluxes (k_gflux), upwind: the shock sensor,
    // 4 stats, 5 flux sweeps (central: with the RK3 combine; upwind: the
    // interface fluxes), 6 update (central: primitives; upwind: RK3 combine
    // and primitives)
    const bool upwind = c->p.convection != 0;
    bool bc_reads_prim = c->udf == ASTR_UDF_RIEMANN2D;
    for (int f = 0; f < 6; ++f)
        bc_reads_prim = bc_reads_prim || c->bctype[f] == 23 || c->bctype[f] == 50 ||
                        c->bctype[f] == 60 || c->bctype[f] == 41 || c->bctype[f] == 42 ||
                        c->bctype[f] == 21 || c->bctype[f] == 31 || c->bctype[f] == 421 ||
                        c->bctype[f] == 411;
    if (c->p.variant == 1) {
        // rk3 of tgvsolver.F90:994-1082: bchomo, rhscal (gradcal, convection,
        // diffusion by df/dx), stage-1 statistics and qsave, the RK3 combine
        // (in the last flux sweep), then the filter and q2fvar
        for (int s = 1; s <= 3; ++s) {
            if (timed) TRY(time_mark(c, 1));
            {
                const long ni = c->g.n[0] + 1, nj = c->g.n[1] + 1, nk = c->g.n[2] + 1;
                TRY(launch_pt(c, k_bchomo, 2L * HM * (nj * nk + ni * nk + ni * nj), pt_args(c)));
            }
            if (timed) TRY(time_mark(c, 2));
            TRY(do_grad(c));
            if (timed) TRY(time_mark(c, 3));
            // rkfirst's statistics of the stage-1 state, summed by k_gflux
            TRY(do_gflux(c, s == 1 ? stats_slot : nullptr));
            if (timed) TRY(time_mark(c, 4));
            if (timed) TRY(time_mark(c, 5));
            TRY(do_flux_sweeps(c, s, dt));
            if (timed) TRY(time_mark(c, 6));
            if (c->p.lfilter) TRY(do_filterq(c));
            TRY(launch_prim(c));
            if (timed) {
                TRY(time_mark(c, 7));
                HIPCHK(hipEventSynchronize(c->ev[7]));
                float ms = 0.f;
                for (int f = 1; f < 7; ++f) {
                    HIPCHK(hipEventElapsedTime(&ms, c->ev[f], c->ev[f + 1]));
                    c->tms[f] += ms;
                    c->tms[7] += ms;
                }
            }
        }
        c->time = c->time + dt;
        return ASTR_OK;
    }
    for (int s = 1; s <= 3; ++s) {
        if (timed) TRY(time_mark(c, 0));
        if (c->p.lfilter) TRY(do_filterq(c));
        if (timed) TRY(time_mark(c, 1));
        // boucon before qswap (src/mainloop.F90:413-415), at the step's
        // start time in every stage
        TRY(do_boucon(c));
        TRY(do_qswap(c));
        if (timed) TRY(time_mark(c, 2));
        TRY(do_grad(c));
        if (timed) TRY(time_mark(c, 3));
        // rkfirst's statcal sees the stage-1 state and its gradients; the
        // central path sums it inside k_gflux, which loads the same inputs
        if (!upwind) TRY(do_gflux(c, s == 1 ? stats_slot : nullptr));
        else if (sensor_on(c)) TRY(do_sensor(c));
        if (timed) TRY(time_mark(c, 4));
        if (upwind && s == 1 && stats_slot) TRY(launch_stats(c, stats_slot));
        if (timed) TRY(time_mark(c, 5));
        if (upwind) {
            TRY(do_upwind(c));
        } else if (alt) {
            TRY(do_flux_sweeps(c, s, dt, s == 1 ? c->qsave : nullptr, true));
            if (s == 1) {
                for (int m = 0; m < 5; ++m) std::swap(c->q[m], c->qsave[m]);
                ++c->gen;
            }
        } else {
            TRY(do_flux_sweeps(c, s, dt));
        }
        if (timed) TRY(time_mark(c, 6));
        if (upwind) TRY(do_rk(c, s, dt));
        // without a boundary kernel, the density of stages 1 and 2 is read by
        // nothing before stage 3's k_prim writes it
        else TRY(launch_prim(c, s < 3 && !bc_reads_prim));
        if (timed) {
            TRY(time_mark(c, 7));
            HIPCHK(hipEventSynchronize(c->ev[7]));
            float ms = 0.f;
            for (int f = 0; f < 7; ++f) {
                HIPCHK(hipEventElapsedTime(&ms, c->ev[f], c->ev[f + 1]));
                c->tms[f] += ms;
                c->tms[7] += ms;
            }
        }
    }
    c->time = c->time + dt;  // time=time+deltat, src/mainloop.F90:200
    return ASTR_OK;
}

// ---- discrete adjoint of the central path (host)

// adjoint fields: q 5, primitives 6, dvel 9, dtmp 3, G 15, qrhs 5, qsave 5
constexpr int ADJ_FIELDS = 48;
// what a stage stores for its reversal: q and the primitives after qswap,
// their gradients (dvel, dtmp), the RK3 output
constexpr int STAGE_FIELDS = 28;
constexpr int STATE_FIELDS = 11;  // q and the primitives

struct AdjFields {
    double *qb[5], *pb[6], *db[9], *tb[3], *gb[15], *rb[5], *sb[5];
};
struct StageStore {
    double *q[5], *prim[6], *qn[5], *dvel[9], *dtmp[3];
};

double* adj_field(astr_ctx* c, long k) { return c->adj_base + k * c->nfield; }

AdjFields adj_fields(astr_ctx* c)
{
    AdjFields f;
    long k = 0;
    for (int m = 0; m < 5; ++m) f.qb[m] = adj_field(c, k++);
    for (int m = 0; m < 6; ++m) f.pb[m] = adj_field(c, k++);
    for (int m = 0; m < 9; ++m) f.db[m] = adj_field(c, k++);
    for (int m = 0; m < 3; ++m) f.tb[m] = adj_field(c, k++);
    for (int m = 0; m < 15; ++m) f.gb[m] = adj_field(c, k++);
    for (int m = 0; m < 5; ++m) f.rb[m] = adj_field(c, k++);
    for (int m = 0; m < 5; ++m) f.sb[m] = adj_field(c, k++);
    return f;
}

// a saved state (q, primitives) k of the checkpoint area
void state_slot(astr_ctx* c, long k, double** q, double** prim)
{
    const long b = ADJ_FIELDS + k * STATE_FIELDS;
    for (int m = 0; m < 5; ++m) q[m] = adj_field(c, b + m);
    for (int m = 0; m < 6; ++m) prim[m] = adj_field(c, b + 5 + m);
}

StageStore stage_slot(astr_ctx* c, long nstates, long k)
{
    StageStore st;
    const long b = ADJ_FIELDS + nstates * STATE_FIELDS + k * STAGE_FIELDS;
    for (int m = 0; m < 5; ++m) st.q[m] = adj_field(c, b + m);
    for (int m = 0; m < 6; ++m) st.prim[m] = adj_field(c, b + 5 + m);
    for (int m = 0; m < 5; ++m) st.qn[m] = adj_field(c, b + 11 + m);
    for (int m = 0; m < 9; ++m) st.dvel[m] = adj_field(c, b + 16 + m);
    for (int m = 0; m < 3; ++m) st.dtmp[m] = adj_field(c, b + 25 + m);
    return st;
}

int copy_fields(astr_ctx* c, double* const* dst, double* const* src, int n)
{
    for (int m = 0; m < n; ++m)
        HIPCHK(hipMemcpyAsync(dst[m], src[m], sizeof(double) * c->nfield, hipMemcpyDeviceToDevice,
                              c->stream));
    return ASTR_OK;
}

int save_state(astr_ctx* c, long k)
{
    double *q[5], *pr[6];
    state_slot(c, k, q, pr);
    TRY(copy_fields(c, q, c->q, 5));
    return copy_fields(c, pr, c->prim, 6);
}

int load_state(astr_ctx* c, long k)
{
    double *q[5], *pr[6];
    state_slot(c, k, q, pr);
    TRY(copy_fields(c, c->q, q, 5));
    return copy_fields(c, c->prim, pr, 6);
}

// the adjoint kernels' operands at the state (q, prim) with its gradients
// (dvel, dtmp) and the stage output qn
AdjArgs adj_args(astr_ctx* c, double* const* q, double* const* prim, double* const* qn,
                 double* const* dvel, double* const* dtmp)
{
    AdjArgs a{};
    a.g = c->g;
    const AdjFields f = adj_fields(c);
    for (int m = 0; m < 5; ++m) {
        a.q[m] = q[m];
        a.qn[m] = qn ? qn[m] : nullptr;
        a.qb[m] = f.qb[m];
        a.rb[m] = f.rb[m];
        a.sb[m] = f.sb[m];
    }
    for (int m = 0; m < 6; ++m) { a.prim[m] = prim[m]; a.pb[m] = f.pb[m]; }
    for (int m = 0; m < 9; ++m) { a.dvel[m] = dvel[m]; a.db[m] = f.db[m]; }
    for (int m = 0; m < 3; ++m) { a.dtmp[m] = dtmp[m]; a.tb[m] = f.tb[m]; }
    for (int m = 0; m < 15; ++m) a.gb[m] = f.gb[m];
    a.jac = c->jac;
    for (int d = 0; d < 3; ++d) a.xi[d] = c->xi + d * c->xi_stride;
    a.const2 = c->const2; a.const5 = c->const5; a.const6 = c->const6;
    a.reynolds = c->p.reynolds; a.prandtl = c->p.prandtl;
    a.tempconst = c->tempconst; a.tempconst1 = c->tempconst1;
    a.visc = c->p.diffterm;
    a.scale = 1.0 / ((double)c->g.n[0] * (double)c->g.n[1] * (double)c->g.n[2]);
    return a;
}

dim3 padded_grid(const astr_ctx* c)
{
    const long plane = (long)(c->g.n[0] + 1 + 2 * HM) * (c->g.n[1] + 1 + 2 * HM);
    return dim3((unsigned)((plane + 255) / 256), (unsigned)(c->g.n[2] + 1 + 2 * HM));
}

// a row_node kernel over the interior rows
template <class K>
int launch_rows(astr_ctx* c, K kernel, const AdjArgs& a)
{
    const long n = (long)(c->g.n[1] + 1) * (c->g.n[2] + 1) * line_end_i(c->g.n[0]);
    hipLaunchKernelGGL(kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

template <class K>
int launch_adj(astr_ctx* c, K kernel, long n, const AdjArgs& a)
{
    if (n <= 0) return ASTR_OK;
    hipLaunchKernelGGL(kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

// Stages k0..k0+nst-1 (stage s = k % 3 + 1) as do_step runs them, from the
// state in c->q and c->prim, each writing what its reversal reads straight
// into its store slot instead of copying it there: the filter's x sweep
// reads the incoming state and writes the slot's q (the y and z sweeps then
// work in place), qswap completes q and the primitives in place, gradcal
// writes the slot's dvel and dtmp, the RK3 combine writes the slot's qn,
// and the primitive refresh writes the next slot's primitives.  The last
// stage's output is copied back into c->q and its primitives go to c->prim;
// the context's own buffers are restored.
int segment_fwd(astr_ctx* c, long nstates, long nst, double dt)
{
    if (nst <= 0) return ASTR_OK;
    double *own_q[5], *own_prim[6], *own_dvel[9], *own_dtmp[3];
    for (int m = 0; m < 5; ++m) own_q[m] = c->q[m];
    for (int m = 0; m < 6; ++m) own_prim[m] = c->prim[m];
    for (int m = 0; m < 9; ++m) own_dvel[m] = c->dvel[m];
    for (int m = 0; m < 3; ++m) own_dtmp[m] = c->dtmp[m];
    auto restore = [&] {
        for (int m = 0; m < 5; ++m) c->q[m] = own_q[m];
        for (int m = 0; m < 6; ++m) c->prim[m] = own_prim[m];
        for (int m = 0; m < 9; ++m) c->dvel[m] = own_dvel[m];
        for (int m = 0; m < 3; ++m) c->dtmp[m] = own_dtmp[m];
    };
    auto run = [&]() -> int {
        TRY(copy_fields(c, stage_slot(c, nstates, 0).prim, own_prim, 6));
        double* qin[5];
        for (int m = 0; m < 5; ++m) qin[m] = own_q[m];
        for (long k = 0; k < nst; ++k) {
            const int s = (int)(k % 3) + 1;
            const StageStore st = stage_slot(c, nstates, k);
            if (c->p.lfilter) {
                for (int d = 0; d < 3; ++d) {
                    SweepArgs a = sweep_args(c, d, PASS_FILTER);
                    for (int m = 0; m < 5; ++m) {
                        a.q[m] = d == 0 ? qin[m] : st.q[m];
                        a.qo[m] = st.q[m];
                    }
                    TRY(launch_sweep(c, a));
                }
            } else {
                TRY(copy_fields(c, st.q, qin, 5));
            }
            for (int m = 0; m < 5; ++m) c->q[m] = st.q[m];
            for (int m = 0; m < 6; ++m) c->prim[m] = st.prim[m];
            for (int m = 0; m < 9; ++m) c->dvel[m] = st.dvel[m];
            for (int m = 0; m < 3; ++m) c->dtmp[m] = st.dtmp[m];
            TRY(do_qswap(c));
            TRY(do_grad(c));
            TRY(do_gflux(c));
            TRY(do_flux_sweeps(c, s, dt, st.qn));
            for (int m = 0; m < 5; ++m) c->q[m] = st.qn[m];
            if (k + 1 < nst) {
                const StageStore nx = stage_slot(c, nstates, k + 1);
                for (int m = 0; m < 6; ++m) c->prim[m] = nx.prim[m];
            } else {
                for (int m = 0; m < 6; ++m) c->prim[m] = own_prim[m];
            }
            TRY(launch_prim(c));
            for (int m = 0; m < 5; ++m) qin[m] = st.qn[m];
            if (s == 3) c->time = c->time + dt;
        }
        return copy_fields(c, own_q, qin, 5);
    };
    const int rc = run();
    restore();
    return rc;
}

// gradcal reversed: primitive-bar (u, v, w, T; halos too) += D_d^T (xi_d
// dvel-bar(.,d)); dvel-bar and dtmp-bar are spent
int grad_adj(astr_ctx* c)
{
    const AdjFields f = adj_fields(c);
    for (int d = 0; d < 3; ++d) {
        const double* xi = c->xi + d * c->xi_stride;
        const double* fin[4] = {f.db[d], f.db[3 + d], f.db[6 + d], f.tb[d]};
        double* fout[4] = {f.pb[1], f.pb[2], f.pb[3], f.pb[5]};
        TRY(launch_sweep_t(c, d, OP_DERIV, 4, fin, fout, xi, false, true));
    }
    // with viscous terms k_gflux_adj_int overwrites them in the next stage
    if (!c->p.diffterm)
        HIPCHK(hipMemsetAsync(f.db[0], 0, sizeof(double) * c->nfield * 12, c->stream));
    return ASTR_OK;
}

// qswap reversed at the state q after it: the primitive refresh of the
// halo and seam nodes, then the copies and averages in the order z, y, x
int qswap_adj(astr_ctx* c, AdjArgs a)
{
    hipLaunchKernelGGL(k_q2f_adj, padded_grid(c), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    const long ext[3] = {c->g.n[0] + 1L, c->g.n[1] + 1L, c->g.n[2] + 1L};
    for (int d = 2; d >= 0; --d) {
        a.axis = d;
        const long nl = d == 0 ? ext[1] * ext[2] : d == 1 ? ext[0] * ext[2] : ext[0] * ext[1];
        TRY(launch_adj(c, k_qswap_adj_line, (2 * HM + 1) * 5 * nl, a));
    }
    return ASTR_OK;
}

// one stage reversed: q-bar and primitive-bar hold the adjoint of the
// stage's outputs on entry and of its inputs on return
int stage_adj(astr_ctx* c, const StageStore& st, int s, double dt)
{
    AdjArgs a = adj_args(c, st.q, st.prim, st.qn, st.dvel, st.dtmp);
    a.stage = s;
    a.rka = RKC[s - 1][0];
    a.rkb = RKC[s - 1][1];
    a.rkc = RKC[s - 1][2];
    a.dt = dt;
    const AdjFields f = adj_fields(c);
    // updatefvar and the RK3 combine
    TRY(launch_rows(c, k_update_adj, a));
    // the flux derivatives: G-bar_d = D_d^T qrhs-bar
    for (int d = 0; d < 3; ++d)
        TRY(launch_sweep_t(c, d, OP_DERIV, 5, f.rb, f.gb + 5 * d, nullptr, false, false));
    // the axis fluxes, at the stored gradients
    TRY(launch_rows(c, k_gflux_adj_int, a));
    TRY(launch_adj(c, k_gflux_adj_halo,
                   gflux_adj_halo_count(c->g, 0) + gflux_adj_halo_count(c->g, 1) +
                       gflux_adj_halo_count(c->g, 2),
                   a));
    TRY(grad_adj(c));
    TRY(qswap_adj(c, a));
    // filterq: the filters of z, y, x transposed, in place
    if (c->p.lfilter)
        for (int d = 2; d >= 0; --d)
            TRY(launch_sweep_t(c, d, OP_FILTER, 5, f.qb, f.qb, nullptr, true, false));
    return ASTR_OK;
}

// the enstrophy of the state after a qswap, and its adjoint in q-bar and
// primitive-bar (the state is left after the qswap)
int objective_adj(astr_ctx* c, double* J)
{
    TRY(do_qswap(c));
    TRY(do_grad(c));
    double st[3];
    TRY(do_stats(c, st));
    *J = st[1];
    const AdjFields f = adj_fields(c);
    HIPCHK(hipMemsetAsync(f.qb[0], 0, sizeof(double) * c->nfield * ADJ_FIELDS, c->stream));
    AdjArgs a = adj_args(c, c->q, c->prim, nullptr, c->dvel, c->dtmp);
    TRY(launch_adj(c, k_ens_adj, (long)c->g.n[0] * c->g.n[1] * c->g.n[2], a));
    TRY(grad_adj(c));
    return qswap_adj(c, a);
}


// ---- batched line-queue engine (host)

// Tuning choices of a queue launch, a bit set: QC_PERSISTENT a grid of the
// resident workgroups pulling items from the counter (else one workgroup
// per item); QC_INTERLEAVE items alternate between the members (else each
// member's tiles in a run, the largest member first); QC_PER_QUANTITY an
// item is one quantity of a tile (else all quantities of the tile in turn)
enum : int { QC_PERSISTENT = 1, QC_INTERLEAVE = 2, QC_PER_QUANTITY = 4, QC_COUNT = 8 };

struct QLaunchHost {
    int pass, dir, stage;      // stage: the RK3 stage applied by a z flux launch, else 0
    int L;
    std::vector<int> members;  // batch members (contexts) of the launch
    size_t lds;
    int args_off = 0;          // first SweepArgs of the launch in the device array
    int items_off = 0, nitems = 0;
    int cfg = QC_PERSISTENT;
    int grid = 0;
};

struct BatchQueue {
    std::vector<QLaunchHost> launches;
    SweepArgs* d_args = nullptr;
    QItem* d_items = nullptr;
    long items_cap = 0;
    int* d_counters = nullptr;
    PtArgs* d_pt = nullptr;
    unsigned* d_bstart = nullptr;  // 4 kinds x (nmem + 1)
    unsigned nblocks[PT_KINDS] = {};
    int cus = 0;
    int tuned[3][LMAX + 1];        // per pass (filter, grad, flux) and L; -1 untuned
    double tune_us[3][LMAX + 1][QC_COUNT];
    ~BatchQueue()
    {
        (void)hipFree(d_args);
        (void)hipFree(d_items);
        (void)hipFree(d_counters);
        (void)hipFree(d_pt);
        (void)hipFree(d_bstart);
    }
};

int pass_slot(int pass) { return pass == PASS_FILTER ? 0 : pass == PASS_GRAD ? 1 : 2; }

template <int PASS>
int queue_launch_p(const QLaunchHost& h, const QueueLaunch& ql, hipStream_t st, bool occupancy,
                   int* occ)
{
    const dim3 grid((unsigned)std::max(h.grid, 1));
    switch (h.L) {
#define QCASE(LL)                                                                            \
    case LL:                                                                                 \
        if (occupancy) {                                                                     \
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(                             \
                occ, h.dir == 0 ? (const void*)k_sweep_q<LL, PASS, true>                     \
                                : (const void*)k_sweep_q<LL, PASS, false>, WGT, h.lds));     \
        } else if (h.dir == 0) {                                                             \
            hipLaunchKernelGGL((k_sweep_q<LL, PASS, true>), grid, dim3(WGT), h.lds, st,     \
                               ql.args, ql.items, ql.nitems, ql.counter);                \
        } else {                                                                             \
            hipLaunchKernelGGL((k_sweep_q<LL, PASS, false>), grid, dim3(WGT), h.lds, st,    \
                               ql.args, ql.items, ql.nitems, ql.counter);                \
        }                                                                                    \
        break;
        QCASE(1) QCASE(2) QCASE(3) QCASE(4) QCASE(5) QCASE(6) QCASE(7) QCASE(8)
        QCASE(9) QCASE(10) QCASE(11) QCASE(12) QCASE(13) QCASE(14) QCASE(15) QCASE(16)
#undef QCASE
    default:
        return fail(ASTR_E_UNSUPPORTED, "line too long");
    }
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

int queue_launch(const QLaunchHost& h, const QueueLaunch& ql, hipStream_t st, bool occupancy = false,
                 int* occ = nullptr)
{
    switch (h.pass) {
    case PASS_FILTER: return queue_launch_p<PASS_FILTER>(h, ql, st, occupancy, occ);
    case PASS_GRAD: return queue_launch_p<PASS_GRAD>(h, ql, st, occupancy, occ);
    default: return queue_launch_p<PASS_RHS>(h, ql, st, occupancy, occ);
    }
}

int nq_of(int pass) { return pass == PASS_GRAD ? 4 : 5; }

// the items of one launch in the order its configuration asks for
void queue_items(const QLaunchHost& h, const std::vector<astr_ctx*>& ctx,
                 const std::vector<SweepArgs>& args, std::vector<QItem>& out)
{
    const int nm = (int)h.members.size();
    const int nq = nq_of(h.pass);
    const bool perq = h.cfg & QC_PER_QUANTITY;
    std::vector<int> ntiles(nm), order(nm);
    for (int k = 0; k < nm; ++k) {
        const SweepArgs& a = args[h.args_off + k];
        ntiles[k] = a.ntile_a * a.nb;
        order[k] = k;
    }
    // largest member first (its lines are the longest work of the launch)
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const Geo& gx = ctx[h.members[x]]->g;
        const Geo& gy = ctx[h.members[y]]->g;
        return (long)gx.n[0] * gx.n[1] * gx.n[2] > (long)gy.n[0] * gy.n[1] * gy.n[2];
    });
    auto push = [&](int k, int t) {
        if (perq) {
            for (int m = 0; m < nq; ++m) out.push_back(QItem{k, t, m, 0});
        } else {
            out.push_back(QItem{k, t, -1, 0});
        }
    };
    if (h.cfg & QC_INTERLEAVE) {
        int tmax = 0;
        for (int k = 0; k < nm; ++k) tmax = std::max(tmax, ntiles[k]);
        for (int t = 0; t < tmax; ++t)
            for (int k : order)
                if (t < ntiles[k]) push(k, t);
    } else {
        for (int k : order)
            for (int t = 0; t < ntiles[k]; ++t) push(k, t);
    }
}

// (re)build the item table and grids of every launch from their configurations
int queue_configure(astr_batch* b);
int queue_build(astr_batch* b);
int queue_step(astr_batch* b, double dt, std::vector<hipEvent_t>* ev);
}  // namespace

// ====================================================================== C ABI

struct astr_batch {
    std::vector<astr_ctx*> ctx;
    int device = 0;
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double dt = 0.0;
    bool built = false;
    // 0: one HIP graph per step over the members' streams; 1: direct
    // launches on the members' streams; 2: the line-queue engine (launches
    // over every member), one HIP graph per step
    int mode = 0;
    std::vector<unsigned long> gens;  // members' set-up generations the graph holds
    BatchQueue* qe = nullptr;
    std::vector<int> seq;             // the queue engine's launch sequence of a step
    int nqueue = 0;                   // queue launches per step (counters)
};

namespace {

// The launch templates of the queue engine: per sweep (pass, axis) one per
// rows-per-lane class L of the members along that axis; the z flux sweep
// once per RK3 stage (its arguments carry the stage).  seq lists a step:
// template indices, and -1 - kind for the pointwise launches.
int queue_build(astr_batch* b)
{
    delete b->qe;
    b->qe = new BatchQueue();
    BatchQueue* qe = b->qe;
    const int nm = (int)b->ctx.size();
    for (int p = 0; p < 3; ++p)
        for (int l = 0; l <= LMAX; ++l) {
            qe->tuned[p][l] = -1;
            for (int k = 0; k < QC_COUNT; ++k) qe->tune_us[p][l][k] = 0.0;
        }
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, b->device));
    qe->cus = prop.multiProcessorCount;
    std::vector<SweepArgs> args;
    auto add_group = [&](int pass, int dir, int stage, std::vector<int>& ids) {
        // one template per L class among the members of this sweep
        std::vector<int> Ls;
        for (int i = 0; i < nm; ++i) {
            astr_ctx* c = b->ctx[i];
            if (pass == PASS_FILTER && !c->p.lfilter) continue;
            if (std::find(Ls.begin(), Ls.end(), c->L[dir]) == Ls.end()) Ls.push_back(c->L[dir]);
        }
        std::sort(Ls.begin(), Ls.end(), std::greater<int>());
        for (int L : Ls) {
            QLaunchHost h;
            h.pass = pass;
            h.dir = dir;
            h.stage = stage;
            h.L = L;
            h.lds = lds_bytes(dir, L);
            h.args_off = (int)args.size();
            for (int i = 0; i < nm; ++i) {
                astr_ctx* c = b->ctx[i];
                if (c->L[dir] != L || (pass == PASS_FILTER && !c->p.lfilter)) continue;
                h.members.push_back(i);
                SweepArgs a = sweep_args(c, dir, pass);
                if (stage > 0) {
                    a.rk_stage = stage;
                    a.rka = RKC[stage - 1][0];
                    a.rkb = RKC[stage - 1][1];
                    a.rkc = RKC[stage - 1][2];
                }
                args.push_back(a);
            }
            ids.push_back((int)qe->launches.size());
            qe->launches.push_back(h);
        }
    };
    std::vector<int> filt[3], grad[3], flux[2], fluxz[3];
    for (int d = 0; d < 3; ++d) add_group(PASS_FILTER, d, 0, filt[d]);
    for (int d = 0; d < 3; ++d) add_group(PASS_GRAD, d, 0, grad[d]);
    for (int d = 0; d < 2; ++d) add_group(PASS_RHS, d, 0, flux[d]);
    for (int s = 1; s <= 3; ++s) add_group(PASS_RHS, 2, s, fluxz[s - 1]);
    b->seq.clear();
    for (int s = 1; s <= 3; ++s) {
        for (int d = 0; d < 3; ++d) b->seq.insert(b->seq.end(), filt[d].begin(), filt[d].end());
        for (int k = PT_QSWAP_X; k <= PT_QSWAP_Z; ++k) b->seq.push_back(-1 - k);
        for (int d = 0; d < 3; ++d) b->seq.insert(b->seq.end(), grad[d].begin(), grad[d].end());
        b->seq.push_back(-1 - PT_GFLUX);
        for (int d = 0; d < 2; ++d) b->seq.insert(b->seq.end(), flux[d].begin(), flux[d].end());
        b->seq.insert(b->seq.end(), fluxz[s - 1].begin(), fluxz[s - 1].end());
        b->seq.push_back(-1 - PT_PRIM);
    }
    b->nqueue = 0;
    for (int k : b->seq)
        if (k >= 0) ++b->nqueue;
    // the time step enters the z flux arguments at capture (queue_set_dt)
    HIPCHK(hipMalloc(&qe->d_args, sizeof(SweepArgs) * std::max<size_t>(args.size(), 1)));
    HIPCHK(hipMemcpy(qe->d_args, args.data(), sizeof(SweepArgs) * args.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&qe->d_counters, sizeof(int) * std::max(b->nqueue, 1)));
    // pointwise launches
    std::vector<PtArgs> pts(nm);
    std::vector<unsigned> bs(PT_KINDS * (nm + 1), 0);
    for (int i = 0; i < nm; ++i) {
        astr_ctx* c = b->ctx[i];
        pts[i] = pt_args(c);
        const long np = (long)(c->g.n[1] + 1) * (c->g.n[2] + 1) * line_end_i(c->g.n[0]);
        const long nb[PT_KINDS] = {(qswap_threads(c->g, 0) + 255) / 256,
                                   (qswap_threads(c->g, 1) + 255) / 256,
                                   (qswap_threads(c->g, 2) + 255) / 256, gflux_blocks(c->g),
                                   (np + 255) / 256};
        for (int k = 0; k < PT_KINDS; ++k)
            bs[k * (nm + 1) + i + 1] = bs[k * (nm + 1) + i] + (unsigned)nb[k];
    }
    for (int k = 0; k < PT_KINDS; ++k) qe->nblocks[k] = bs[k * (nm + 1) + nm];
    HIPCHK(hipMalloc(&qe->d_pt, sizeof(PtArgs) * nm));
    HIPCHK(hipMemcpy(qe->d_pt, pts.data(), sizeof(PtArgs) * nm, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&qe->d_bstart, sizeof(unsigned) * bs.size()));
    HIPCHK(hipMemcpy(qe->d_bstart, bs.data(), sizeof(unsigned) * bs.size(), hipMemcpyHostToDevice));
    return queue_configure(b);
}

// the z flux arguments carry dt
int queue_set_dt(astr_batch* b, double dt)
{
    BatchQueue* qe = b->qe;
    for (const QLaunchHost& h : qe->launches) {
        if (h.stage == 0) continue;
        const int nmem = (int)h.members.size();
        std::vector<SweepArgs> a(nmem);
        HIPCHK(hipMemcpy(a.data(), qe->d_args + h.args_off, sizeof(SweepArgs) * nmem,
                         hipMemcpyDeviceToHost));
        for (SweepArgs& x : a) x.dt = dt;
        HIPCHK(hipMemcpy(qe->d_args + h.args_off, a.data(), sizeof(SweepArgs) * nmem,
                         hipMemcpyHostToDevice));
    }
    return ASTR_OK;
}

int queue_configure(astr_batch* b)
{
    BatchQueue* qe = b->qe;
    std::vector<SweepArgs> args;
    std::vector<QItem> items;
    for (QLaunchHost& h : qe->launches) {
        const int t = qe->tuned[pass_slot(h.pass)][h.L];
        h.cfg = t >= 0 ? t : QC_PERSISTENT;
        // the launch's own arguments (host copies, for the tile counts)
        const int nmem = (int)h.members.size();
        args.resize((size_t)h.args_off + nmem);
        for (int k = 0; k < nmem; ++k)
            args[h.args_off + k] = sweep_args(b->ctx[h.members[k]], h.dir, h.pass);
        h.items_off = (int)items.size();
        queue_items(h, b->ctx, args, items);
        h.nitems = (int)items.size() - h.items_off;
        if (h.cfg & QC_PERSISTENT) {
            int occ = 0;
            TRY(queue_launch(h, QueueLaunch{}, nullptr, true, &occ));
            h.grid = std::min(h.nitems, std::max(occ, 1) * qe->cus);
        } else {
            h.grid = h.nitems;
        }
    }
    if ((long)items.size() > qe->items_cap) {
        (void)hipFree(qe->d_items);
        qe->d_items = nullptr;
        qe->items_cap = 0;
        HIPCHK(hipMalloc(&qe->d_items, sizeof(QItem) * items.size()));
        qe->items_cap = (long)items.size();
    }
    HIPCHK(hipMemcpy(qe->d_items, items.data(), sizeof(QItem) * items.size(), hipMemcpyHostToDevice));
    b->built = false;
    return ASTR_OK;
}

// One RK3 step of every member through the queue engine on the batch's
// root stream; ev (tuning) receives an event before and after each queue
// launch of the sequence.
int queue_step(astr_batch* b, std::vector<hipEvent_t>* ev)
{
    BatchQueue* qe = b->qe;
    const int nm = (int)b->ctx.size();
    hipLaunchKernelGGL(k_zero_ints, dim3((b->nqueue + 255) / 256), dim3(256), 0, b->root,
                       qe->d_counters, b->nqueue);
    HIPCHK(hipGetLastError());
    int qi = 0;
    for (int k : b->seq) {
        if (k < 0) {
            const int kind = -1 - k;
            const PtBatch pb{qe->d_pt, qe->d_bstart + kind * (nm + 1), nm};
            const dim3 grid(qe->nblocks[kind]);
            switch (kind) {
            case PT_QSWAP_X:
                hipLaunchKernelGGL(k_pt_batch<PT_QSWAP_X>, grid, dim3(256), 0, b->root, pb.args,
                                   pb.bstart, pb.nmem);
                break;
            case PT_QSWAP_Y:
                hipLaunchKernelGGL(k_pt_batch<PT_QSWAP_Y>, grid, dim3(256), 0, b->root, pb.args,
                                   pb.bstart, pb.nmem);
                break;
            case PT_QSWAP_Z:
                hipLaunchKernelGGL(k_pt_batch<PT_QSWAP_Z>, grid, dim3(256), 0, b->root, pb.args,
                                   pb.bstart, pb.nmem);
                break;
            case PT_GFLUX:
                hipLaunchKernelGGL(k_pt_batch<PT_GFLUX>, grid, dim3(256), 0, b->root, pb.args,
                                   pb.bstart, pb.nmem);
                break;
            default:
                hipLaunchKernelGGL(k_pt_batch<PT_PRIM>, grid, dim3(256), 0, b->root, pb.args,
                                   pb.bstart, pb.nmem);
                break;
            }
            HIPCHK(hipGetLastError());
            continue;
        }
        const QLaunchHost& h = qe->launches[k];
        QueueLaunch ql{qe->d_args + h.args_off, qe->d_items + h.items_off, h.nitems,
                       (h.cfg & QC_PERSISTENT) ? qe->d_counters + qi : nullptr};
        if (ev) HIPCHK(hipEventRecord((*ev)[2 * qi], b->root));
        TRY(queue_launch(h, ql, b->root));
        if (ev) HIPCHK(hipEventRecord((*ev)[2 * qi + 1], b->root));
        ++qi;
    }
    return ASTR_OK;
}

}  // namespace

extern "C" {

const char* astr_last_error(void) { return g_err.c_str(); }

void astr_default_params(astr_params* p, int32_t n)
{
    memset(p, 0, sizeof *p);
    p->im = p->jm = p->km = n;
    p->hm = HM;
    p->numq = 5;
    p->npdci = p->npdcj = p->npdck = 3;
    p->variant = 0;
    p->alfa_filter = 0.49;
    p->reynolds = 1600.0;
    p->mach = 0.1;
    p->prandtl = 0.72;
    p->gamma = 1.4;
    p->ref_tem = 273.15;
    p->sutherland = 110.3;
    p->beta_halo = 1.11;
    p->beta_bound = 0.98;
    p->convection = 0;
    p->deriv_explicit = 0;
    p->recon_schem = 3;
    p->lchardecomp = 1;
    p->lfilter = 1;
    p->diffterm = 1;
    p->shkcrt = 0.01;
    p->bfacmpld = 0.7;
}

int astr_create(const astr_params* p, int32_t device, astr_ctx** out)
{
    if (!p || !out) return fail(ASTR_E_ARG, "null argument");
    *out = nullptr;
    if (p->hm != HM) return fail(ASTR_E_UNSUPPORTED, "hm must be 5");
    if (p->numq != 5) return fail(ASTR_E_UNSUPPORTED, "numq must be 5");
    if (p->variant != 0 && p->variant != 1)
        return fail(ASTR_E_ARG, "variant must be 0 (src/mainloop.F90) or 1 (tgvsolver.F90)");
    if (p->variant == 1 && (p->convection != 0 || p->km == 0))
        return fail(ASTR_E_UNSUPPORTED, "the mini-app variant is the central scheme on a 3-D box");
    const int dims[3] = {p->im, p->jm, p->km};
    const int nts[3] = {p->npdci, p->npdcj, p->npdck};
    if (p->convection < 0 || p->convection > 2)
        return fail(ASTR_E_ARG, "convection must be 0 (central), 1 (upwind) or 2 (compact upwind)");
    if (p->deriv_explicit != 0 && p->deriv_explicit != 1)
        return fail(ASTR_E_ARG, "deriv_explicit must be 0 or 1");
    const bool two_d = p->km == 0;
    if (p->convection == 2 && p->diffterm)
        return fail(ASTR_E_UNSUPPORTED, "the upwind path is built for inviscid runs (diffterm 0)");
    if (p->convection == 1) {
        const int r = p->recon_schem;
        if (!(r == -1 || r == 0 || r == 1 || r == 2 || r == 3 || r == 5 || r == 6))
            return fail(ASTR_E_ARG, "recon_schem must be -1, 0, 1, 2, 3, 5 or 6");
        if (p->diffterm)
            return fail(ASTR_E_UNSUPPORTED, "the upwind path is built for inviscid runs (diffterm 0)");
    }
    if (two_d && (p->convection == 0 || p->deriv_explicit != 1 || nts[2] != 3))
        return fail(ASTR_E_UNSUPPORTED,
                    "two-dimensional boxes need an upwind convection, deriv_explicit 1 and npdck 3");
    // the compact line sweeps (filter, compact derivatives, central fluxes)
    // take axes of up to 32*LMAX-7 = 505 cells, with an LDS tile that fits
    // the device's per-workgroup limit (160 KB on gfx950, where the y/z
    // tiles of the longest lines take 69 KB); an upwind context without the
    // filter and with the explicit gradient never runs them
    const bool need_lines = p->convection == 0 || p->lfilter || !p->deriv_explicit;
    bool lines_ok[3] = {false, false, false};
    for (int d = 0; d < 3; ++d) {
        if (d == 2 && two_d) continue;
        if (dims[d] < 16) return fail(ASTR_E_UNSUPPORTED, "each axis needs at least 16 cells");
        if (nts[d] < 1 || nts[d] > 4) return fail(ASTR_E_ARG, "npdc must be 1..4");
        if (dims[d] > 1000000) return fail(ASTR_E_UNSUPPORTED, "axis longer than 10^6 cells");
        const int L = (dims[d] + 7 + LPL - 1) / LPL;
        lines_ok[d] = L <= LMAX;
        if (!lines_ok[d] && need_lines)
            return fail(ASTR_E_UNSUPPORTED, "axis longer than the compact line operators take "
                                            "(505 cells)");
    }
    // the argument checks above touch no GPU; the LDS tiles of the line
    // sweeps must also fit the device's limit per workgroup
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        lds_max <= 0)
        return fail(ASTR_E_HIP, "cannot query the device's LDS per workgroup");
    for (int d = 0; d < 3; ++d) {
        if (!lines_ok[d]) continue;
        const int L = (dims[d] + 7 + LPL - 1) / LPL;
        lines_ok[d] = lds_bytes(0, L) <= (size_t)lds_max && lds_bytes(1, L) <= (size_t)lds_max;
        if (!lines_ok[d] && need_lines)
            return fail(ASTR_E_UNSUPPORTED,
                        "the line sweeps' LDS tiles exceed the device's limit per workgroup");
    }
    auto* c = new astr_ctx();
    c->p = *p;
    c->device = device;
    for (int d = 0; d < 3; ++d) c->lines_ok[d] = lines_ok[d];
    auto cleanup = [&](int rc) {
        astr_destroy(c);
        return rc;
    };
    if (set_dev(c) != ASTR_OK) return cleanup(ASTR_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(ASTR_E_HIP, "hipStreamCreate failed"));

    for (int d = 0; d < 3; ++d) c->g.n[d] = dims[d];
    c->pi = ((OI + dims[0] + 1 + HM) + 15) / 16 * 16;
    c->pj = dims[1] + 1 + 2 * HM;
    c->pk = dims[2] + 1 + 2 * HM;
    c->nfield = c->pi * c->pj * c->pk;
    c->g.s[0] = 1;
    c->g.s[1] = c->pi;
    c->g.s[2] = c->pi * c->pj;
    c->g.org = OI + HM * c->g.s[1] + HM * c->g.s[2];
    const int nfields = 5 + 6 + 5 + 5 + 9 + 3 + 15 + 1 + 2 + (p->convection != 0 ? 1 : 0) +
                        (p->convection == 2 ? 20 : 0);
    if (hipMalloc(&c->base, sizeof(double) * c->nfield * nfields) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "device allocation of the fields failed"));
    double* f = c->base;
    auto take = [&]() { double* r = f; f += c->nfield; return r; };
    for (int m = 0; m < 5; ++m) c->q[m] = take();
    for (int m = 0; m < 6; ++m) c->prim[m] = take();
    for (int m = 0; m < 5; ++m) c->qsave[m] = take();
    for (int m = 0; m < 5; ++m) c->qrhs[m] = take();
    for (int m = 0; m < 9; ++m) c->dvel[m] = take();
    for (int m = 0; m < 3; ++m) c->dtmp[m] = take();
    for (int m = 0; m < 15; ++m) c->gfl[m] = take();
    c->jac = take();
    c->scratch = take();
    c->scratch2 = take();
    if (p->convection != 0) c->ssf = take();
    if (p->convection == 2)
        for (int m = 0; m < 10; ++m) {
            c->swf[m] = take();
            c->fhf[m] = take();
        }
    if (hipMemset(c->base, 0, sizeof(double) * c->nfield * nfields) != hipSuccess)
        return cleanup(fail(ASTR_E_HIP, "hipMemset failed"));
    if (hipMalloc(&c->lshock, c->nfield) != hipSuccess ||
        hipMemset(c->lshock, 0, c->nfield) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "device allocation of the shock flags failed"));

    // constants of refcal, src/solver.F90:105-122
    const double g = p->gamma, ma = p->mach;
    c->const1 = 1.0 / (g * (g - 1.0) * (ma * ma));
    c->const2 = g * (ma * ma);
    c->const3 = (g - 1.0) / 3.0 * p->prandtl * (ma * ma);
    c->const4 = (g - 1.0) * (ma * ma) * p->reynolds * p->prandtl;
    c->const5 = (g - 1.0) * (ma * ma);
    c->const6 = 1.0 / (g - 1.0);
    c->const7 = c->const4;
    c->tempconst = p->sutherland / p->ref_tem;
    c->tempconst1 = 1.0 + c->tempconst;

    // line systems
    std::vector<double> tabs;
    std::vector<long> offs;
    for (int d = 0; d < 3; ++d) {
        // an axis the compact sweeps cannot take (the k axis of a 2-D box, a
        // long axis of an upwind context) keeps the line systems of a
        // 16-cell axis, which no entry point then uses
        const int nd = lines_ok[d] ? dims[d] : 16;
        c->L[d] = (nd + 7 + LPL - 1) / LPL;
        for (int op = 0; op < 2; ++op) {
            HostTab h = build_tab(op, nts[d], nd, p->alfa_filter, c->L[d]);
            offs.push_back((long)tabs.size());
            tabs.insert(tabs.end(), h.t.begin(), h.t.end());
            LineTab& t = c->tab[d][op];
            t.R = h.R;
            t.first = h.first;
            t.last = h.last;
            t.ntype = nts[d];
            t.dim = dims[d];
            t.op = op;
            t.L = c->L[d];
        }
    }
    if (hipMalloc(&c->tabs, sizeof(double) * tabs.size()) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "table allocation failed"));
    if (hipMemcpy(c->tabs, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(ASTR_E_HIP, "table copy failed"));
    for (int d = 0, s = 0; d < 3; ++d)
        for (int op = 0; op < 2; ++op, ++s) c->tab[d][op].t = c->tabs + offs[s];

    if (p->convection == 2)
        for (int d = 0; d < (two_d ? 2 : 3); ++d)
            for (int w = 0; w < 2; ++w) {
                std::vector<double> ac = build_flux_tab(nts[d], dims[d], w == 0, p->bfacmpld,
                                                        c->cmp_first[d], c->cmp_last[d]);
                if (hipMalloc(&c->cmp_ac[d][w], sizeof(double) * ac.size()) != hipSuccess ||
                    hipMemcpy(c->cmp_ac[d][w], ac.data(), sizeof(double) * ac.size(),
                              hipMemcpyHostToDevice) != hipSuccess)
                    return cleanup(fail(ASTR_E_HIP, "compact flux table upload failed"));
            }
    FiltCoef fc = build_coef(p->alfa_filter, p->beta_halo, p->beta_bound);
    if (hipMalloc(&c->fc, sizeof(FiltCoef)) != hipSuccess ||
        hipMemcpy(c->fc, &fc, sizeof fc, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(ASTR_E_HIP, "coefficient upload failed"));

    // transposed line systems and the closure blocks of B^T (the adjoint
    // sweeps)
    {
        std::vector<double> tt;
        std::vector<long> toff;
        int nl[3][2], nh[3][2];
        for (int d = 0; d < 3; ++d) {
            const int nd = lines_ok[d] ? dims[d] : 16;
            for (int op = 0; op < 2; ++op) {
                HostTab h = build_tab_t(op, nts[d], nd, p->alfa_filter, c->L[d]);
                std::vector<double> dlo, dhi;
                build_trans_blocks(op, nts[d], h, fc, nl[d][op], nh[d][op], dlo, dhi);
                toff.push_back((long)tt.size());
                tt.insert(tt.end(), h.t.begin(), h.t.end());
                toff.push_back((long)tt.size());
                tt.insert(tt.end(), dlo.begin(), dlo.end());
                toff.push_back((long)tt.size());
                tt.insert(tt.end(), dhi.begin(), dhi.end());
                LineTab& t = c->tab_t[d][op];
                t = c->tab[d][op];
                TransTab& b = c->tt[d][op];
                b.nlo = nl[d][op];
                b.nhi = nh[d][op];
                b.lo0 = h.first - 5;
                b.hi0 = h.last - 10;
            }
        }
        if (hipMalloc(&c->tabs_t, sizeof(double) * tt.size()) != hipSuccess)
            return cleanup(fail(ASTR_E_NOMEM, "table allocation failed"));
        if (hipMemcpy(c->tabs_t, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return cleanup(fail(ASTR_E_HIP, "table copy failed"));
        for (int d = 0, s = 0; d < 3; ++d)
            for (int op = 0; op < 2; ++op, s += 3) {
                c->tab_t[d][op].t = c->tabs_t + toff[s];
                c->tt[d][op].dlo = c->tabs_t + toff[s + 1];
                c->tt[d][op].dhi = c->tabs_t + toff[s + 2];
            }
    }

    c->xi_stride = std::max(LPL * LMAX, std::max(dims[0], std::max(dims[1], dims[2])) + 1 + LPL) +
                   2 * HM + 1;
    if (hipMalloc(&c->xi, sizeof(double) * 3 * c->xi_stride) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "metric allocation failed"));
    (void)hipMemset(c->xi, 0, sizeof(double) * 3 * c->xi_stride);

    // partial sums of k_stats (one workgroup per 256 nodes) or of k_gflux's
    // statistics (one per workgroup)
    c->npart = ((long)dims[0] * dims[1] * std::max(dims[2], 1) + 255) / 256;
    if (dims[2] > 0) c->npart = std::max(c->npart, gflux_blocks(c->g));
    c->host_part.resize(6 * c->npart);
    // the partials, then the second-level partials of the statistics
    if (hipMalloc(&c->part, sizeof(double) * 6 * (c->npart + STATS_FIN)) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "partials allocation failed"));
    if (hipMalloc(&c->stats_dev, sizeof(double) * 3 * STATS_SLOTS) != hipSuccess)
        return cleanup(fail(ASTR_E_NOMEM, "statistics allocation failed"));
    for (int e = 0; e < 16; ++e)
        if (hipEventCreate(&c->ev[e]) != hipSuccess) return cleanup(fail(ASTR_E_HIP, "event"));
    *out = c;
    return ASTR_OK;
}

int astr_destroy(astr_ctx* c)
{
    if (!c) return ASTR_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->base);
    (void)hipFree(c->tabs);
    (void)hipFree(c->fc);
    (void)hipFree(c->tabs_t);
    (void)hipFree(c->adj_base);
    (void)hipFree(c->xi);
    (void)hipFree(c->part);
    (void)hipFree(c->stats_dev);
    (void)hipFree(c->lshock);
    for (int d = 0; d < 3; ++d)
        for (int w = 0; w < 2; ++w) (void)hipFree(c->cmp_ac[d][w]);
    if (c->dxf[0]) (void)hipFree(c->dxf[0]);
    for (int e = 0; e < 16; ++e)
        if (c->ev[e]) (void)hipEventDestroy(c->ev[e]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ASTR_OK;
}

int astr_upload_geom(astr_ctx* c, const double* jacob, const double* dxi)
{
    if (!c || !jacob || !dxi) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    ++c->gen;
    const int n[3] = {c->g.n[0], c->g.n[1], c->g.n[2]};
    const long ni = n[0] + 1 + 2 * HM, nj = n[1] + 1 + 2 * HM, nk = n[2] + 1 + 2 * HM;
    const long nt = ni * nj * nk;
    // rectilinear (dxi diagonal, dxi(d,d) a function of the d-th index
    // alone, over the whole haloed array): the metric lines xi; otherwise
    // the nine components per node (general metrics)
    std::vector<double> xi(3 * c->xi_stride, 0.0);
    bool rect = true;
    for (int d = 0; d < 3 && rect; ++d)
        for (int e = 0; e < 3 && rect; ++e) {
            const double* m = dxi + (long)(d + 3 * e) * nt;  // dxi(:,:,:,d+1,e+1)
            for (long k = 0; k < nk && rect; ++k)
                for (long j = 0; j < nj && rect; ++j)
                    for (long i = 0; i < ni; ++i) {
                        const double v = m[i + ni * (j + nj * k)];
                        if (d != e) {
                            if (v != 0.0) { rect = false; break; }
                            continue;
                        }
                        const long idx = d == 0 ? i : d == 1 ? j : k;
                        double& slot = xi[d * c->xi_stride + idx];
                        const bool first = (d == 0 ? (j == 0 && k == 0)
                                                   : d == 1 ? (i == 0 && k == 0) : (i == 0 && j == 0));
                        if (first) slot = v;
                        else if (slot != v) { rect = false; break; }
                    }
        }
    std::vector<double> stage;
    if (rect) {
        HIPCHK(hipMemcpy(c->xi, xi.data(), sizeof(double) * xi.size(), hipMemcpyHostToDevice));
        c->curv = false;
    } else {
        // general metrics: the central path of the main variant on a
        // three-dimensional box reads all nine components
        if (c->p.convection != 0 || c->p.variant != 0 || c->g.n[2] == 0)
            return fail(ASTR_E_UNSUPPORTED, "curvilinear metrics (off-diagonal or varying dxi) "
                                            "are built for the central path of the main variant "
                                            "on three-dimensional boxes");
        if (!c->dxf[0]) {
            double* base = nullptr;
            if (hipMalloc(&base, sizeof(double) * c->nfield * 9) != hipSuccess)
                return fail(ASTR_E_NOMEM, "device allocation of the metric fields failed");
            HIPCHK(hipMemset(base, 0, sizeof(double) * c->nfield * 9));
            for (int m = 0; m < 9; ++m) c->dxf[m] = base + (long)m * c->nfield;
        }
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e)
                TRY(upload_field(c, c->dxf[3 * d + e], dxi + (long)(d + 3 * e) * nt, stage));
        c->curv = true;
    }
    TRY(upload_field(c, c->jac, jacob, stage));
    // a uniform grid's jacob is one value: the kernels of the central path
    // then take it as a scalar (the array stays uploaded for the others)
    c->juni = true;
    c->j0 = jacob[0];
    for (long t = 1; t < nt && c->juni; ++t) c->juni = jacob[t] == c->j0;
    c->have_geom = true;
    return ASTR_OK;
}

int astr_upload_state(astr_ctx* c, const double* q, const double* rho, const double* vel,
                      const double* prs, const double* tmp)
{
    if (!c || !q || !rho || !vel || !prs || !tmp) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    const long nt = (long)(c->g.n[0] + 1 + 2 * HM) * (c->g.n[1] + 1 + 2 * HM) * (c->g.n[2] + 1 + 2 * HM);
    std::vector<double> stage;
    for (int m = 0; m < 5; ++m) TRY(upload_field(c, c->q[m], q + m * nt, stage));
    TRY(upload_field(c, c->prim[0], rho, stage));
    for (int m = 0; m < 3; ++m) TRY(upload_field(c, c->prim[1 + m], vel + m * nt, stage));
    TRY(upload_field(c, c->prim[4], prs, stage));
    TRY(upload_field(c, c->prim[5], tmp, stage));
    c->have_state = true;
    return ASTR_OK;
}

int astr_download_state(astr_ctx* c, double* q, double* rho, double* vel, double* prs, double* tmp)
{
    if (!c || !q || !rho || !vel || !prs || !tmp) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    const long nt = (long)(c->g.n[0] + 1 + 2 * HM) * (c->g.n[1] + 1 + 2 * HM) * (c->g.n[2] + 1 + 2 * HM);
    std::vector<double> stage;
    for (int m = 0; m < 5; ++m) TRY(download_field(c, c->q[m], q + m * nt, stage, false));
    TRY(download_field(c, c->prim[0], rho, stage, false));
    for (int m = 0; m < 3; ++m) TRY(download_field(c, c->prim[1 + m], vel + m * nt, stage, false));
    TRY(download_field(c, c->prim[4], prs, stage, false));
    TRY(download_field(c, c->prim[5], tmp, stage, false));
    return ASTR_OK;
}

int astr_download_rhs(astr_ctx* c, double* qrhs, double* dvel, double* dtmp)
{
    if (!c) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    const long ni = n_interior(c);
    std::vector<double> stage;
    if (qrhs)
        for (int m = 0; m < 5; ++m) TRY(download_field(c, c->qrhs[m], qrhs + m * ni, stage, true));
    if (dvel)
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d)
                // dvel(:,:,:,m,d) is column (m-1) + 3*(d-1) of the Fortran array
                TRY(download_field(c, c->dvel[m * 3 + d], dvel + (m + 3 * d) * ni, stage, true));
    if (dtmp)
        for (int d = 0; d < 3; ++d) TRY(download_field(c, c->dtmp[d], dtmp + d * ni, stage, true));
    return ASTR_OK;
}

static int ready_state(astr_ctx* c)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    if (!c->have_geom || !c->have_state) return fail(ASTR_E_STATE, "upload geometry and state first");
    return set_dev(c);
}

bool all_periodic(const astr_ctx* c)
{
    return c->p.npdci == 3 && c->p.npdcj == 3 && c->p.npdck == 3;
}

// stage pieces and steps: three-dimensional boxes, bounded axes on the
// central path of the main variant
// a boundary condition of bc.F90 on some face: gcnscbc (it acts on faces 2
// and 4 only), zeroextrap, symmetry or a noslip wall
static bool has_face_bc(const astr_ctx* c)
{
    for (int f = 0; f < 6; ++f)
        if ((c->bctype[f] == 23 && (f == 1 || f == 3)) || c->bctype[f] == 50 ||
            c->bctype[f] == 60 || c->bctype[f] == 41 || c->bctype[f] == 42 || c->bctype[f] == 21 ||
            c->bctype[f] == 31 || c->bctype[f] == 421 || c->bctype[f] == 411)
            return true;
    return false;
}

static int ready(astr_ctx* c)
{
    TRY(ready_state(c));
    if (c->curv)
        for (int f = 0; f < 6; ++f)
            if (c->bctype[f] == 60)
                return fail(ASTR_E_UNSUPPORTED, "symmetry (bctype 60) needs rectilinear metrics");
    if (c->p.variant == 1 && !c->have_h)
        return fail(ASTR_E_STATE, "the mini-app variant needs astr_set_spacing first");
    if (c->g.n[2] == 0) {
        // two-dimensional boxes (upwind convection, explicit gradient):
        // any closure on i and j, boundary conditions by astr_set_bc
        if (c->p.lfilter)
            return fail(ASTR_E_UNSUPPORTED, "two-dimensional boxes step without the filter (lfilter 0)");
        return ASTR_OK;
    }
    if (!all_periodic(c) && (c->p.convection != 0 || c->p.variant != 0))
        return fail(ASTR_E_UNSUPPORTED, "bounded axes of a three-dimensional box step on the "
                                        "central path of the main variant only");
    // a single box's axis is periodic (3) or bounded at both ends (4); npdc
    // 1 and 2 are a rank's share of an axis split over several ranks, whose
    // interface halos a neighbour fills (src/parallel.F90:1044-1110)
    for (int d = 0; d < 3; ++d) {
        const int nt = d == 0 ? c->p.npdci : d == 1 ? c->p.npdcj : c->p.npdck;
        if (nt != 3 && nt != 4)
            return fail(ASTR_E_UNSUPPORTED, "a single box steps with npdc 3 (periodic) or 4 "
                                            "(bounded at both ends) on each axis");
    }
    return ASTR_OK;
}

// batch members: three-dimensional periodic boxes without a time-dependent
// boundary condition (a captured step replays the same launches)
static int ready_batch(astr_ctx* c)
{
    TRY(ready(c));
    if (c->g.n[2] == 0) return fail(ASTR_E_UNSUPPORTED, "batched members are three-dimensional boxes");
    if (!all_periodic(c)) return fail(ASTR_E_UNSUPPORTED, "batched members are periodic boxes");
    if (c->curv) return fail(ASTR_E_UNSUPPORTED, "batched members have rectilinear metrics");
    if (c->udf || has_face_bc(c))
        return fail(ASTR_E_UNSUPPORTED, "batched members take no boundary condition");
    return ASTR_OK;
}

int astr_filterq(astr_ctx* c)
{
    TRY(ready(c));
    if (c->g.n[2] == 0) return fail(ASTR_E_UNSUPPORTED, "the filter runs on three-dimensional boxes");
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant runs whole steps only");
    return do_filterq(c);
}

int astr_qswap(astr_ctx* c)
{
    TRY(ready(c));
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant runs whole steps only");
    return do_qswap(c);
}

int astr_rhscal(astr_ctx* c)
{
    // the upwind right-hand side takes any closure type with halos as given
    if (c && c->p.convection != 0) TRY(ready_state(c));
    else TRY(ready(c));
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant runs whole steps only");
    return do_rhscal(c, c->timing == 1);
}

int astr_download_shock(astr_ctx* c, uint8_t* lshock)
{
    if (!c || !lshock) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<unsigned char> all(c->nfield);
    HIPCHK(hipMemcpy(all.data(), c->lshock, c->nfield, hipMemcpyDeviceToHost));
    const int ni = c->g.n[0] + 1, nj = c->g.n[1] + 1, nk = c->g.n[2] + 1;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i)
                lshock[i + (long)ni * (j + (long)nj * k)] = all[gidx(c->g, i, j, k)];
    return ASTR_OK;
}

int astr_rk3_update(astr_ctx* c, int32_t stage, double dt)
{
    TRY(ready(c));
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant runs whole steps only");
    if (stage < 1 || stage > 3) return fail(ASTR_E_ARG, "stage must be 1..3");
    c->bc_dt = dt;  // the next boucon's outflow reads it
    return do_rk(c, stage, dt);
}

int astr_stats(astr_ctx* c, double* out3)
{
    TRY(ready(c));
    if (!out3) return fail(ASTR_E_ARG, "null argument");
    if (c->g.n[2] == 0) return fail(ASTR_E_UNSUPPORTED, "statistics of three-dimensional boxes only");
    return do_stats(c, out3);
}

// ---- statistics into device memory, allocating nothing
//
// The partial sums live in `scratch`, a file-scope device array of a fixed
// size: NA_BLOCKS first-level partials and STATS_FIN second-level ones, 6
// doubles each (110.6 KB per GPU, held until the process ends).  Its
// device address is looked up once per device with hipGetSymbolAddress
// (the array's name is only a host-side stand-in in host code).  Every
// partial a call reads is written by that call first, so nothing needs
// clearing; two calls that overlap on one GPU (two streams) race on it.
constexpr int NA_BLOCKS = 2048;
static __device__ __align__(256) double scratch[6 * (NA_BLOCKS + STATS_FIN)];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[64] = {};

static int scratch_ptr(int device, double** ptr)
{
    if (device < 0 || device >= 64) return fail(ASTR_E_ARG, "device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        // the context's device is current (set_dev)
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = p;
    }
    *ptr = static_cast<double*>(g_scratch_dev[device]);
    return ASTR_OK;
}

int astr_stats_device(astr_ctx* c, double* out3_dev)
{
    TRY(ready(c));
    if (!out3_dev) return fail(ASTR_E_ARG, "null argument");
    if (c->g.n[2] == 0) return fail(ASTR_E_UNSUPPORTED, "statistics of three-dimensional boxes only");
    const double nodes = (double)c->g.n[0] * (double)c->g.n[1] * (double)c->g.n[2];
    if (nodes >= 2147483648.0)
        return fail(ASTR_E_UNSUPPORTED,
                    "astr_stats_device takes boxes of fewer than 2^31 nodes (32-bit node index)");
    double* part = nullptr;
    TRY(scratch_ptr(c->device, &part));
    double* part2 = part + 6 * NA_BLOCKS;
    hipLaunchKernelGGL(k_stats_gs, dim3(NA_BLOCKS), dim3(256), 0, c->stream, pt_args(c), part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_stats_fin, dim3(STATS_FIN), dim3(256), 0, c->stream, (const double*)part,
                       (long)NA_BLOCKS, (long)((NA_BLOCKS + STATS_FIN - 1) / STATS_FIN), part2,
                       nodes, (double*)nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_stats_fin, dim3(1), dim3(256), 0, c->stream, (const double*)part2,
                       (long)STATS_FIN, (long)STATS_FIN, (double*)nullptr, nodes, out3_dev);
    HIPCHK(hipGetLastError());
    return ASTR_OK;
}

int astr_qmax(astr_ctx* c, double* qmax5)
{
    TRY(ready(c));
    if (!qmax5) return fail(ASTR_E_ARG, "null argument");
    // the partial-sum buffer holds 6 doubles per block of npart; 5 per
    // workgroup of at most npart workgroups fit
    const long nb = std::min<long>(c->npart, 1024);
    hipLaunchKernelGGL(k_qmax, dim3((unsigned)nb), dim3(256), 0, c->stream, pt_args(c));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->host_part.data(), c->part, sizeof(double) * 5 * nb,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int m = 0; m < 5; ++m) {
        double v = -HUGE_VAL;
        for (long b = 0; b < nb; ++b) v = std::max(v, c->host_part[5 * b + m]);
        qmax5[m] = v;
    }
    return ASTR_OK;
}

// nsteps whole steps; restart: the per-family kernel times start again at
// zero (astr_rk3_steps) rather than adding up over calls (astr_rk3_step)
static int rk3_steps(astr_ctx* c, int32_t nsteps, double dt, double* stats, bool restart)
{
    TRY(ready(c));
    if (nsteps < 0) return fail(ASTR_E_ARG, "nsteps < 0");
    if (stats && c->g.n[2] == 0)
        return fail(ASTR_E_UNSUPPORTED, "statistics of three-dimensional boxes only");
    if (c->timing == 1 && restart)
        for (int f = 0; f < 9; ++f) c->tms[f] = 0.0;
    if (c->timing == 2) HIPCHK(hipEventRecord(c->ev[14], c->stream));
    // each step's statistics go to a device slot; the slots are copied to
    // the host when they are all used and after the last step
    int done = 0;
    for (int s = 0; s < nsteps; ++s) {
        const int slot = s - done;
        TRY(do_step(c, dt, stats ? c->stats_dev + 3 * slot : nullptr, c->timing == 1, true));
        if (stats && (slot + 1 == STATS_SLOTS || s + 1 == nsteps)) {
            HIPCHK(hipMemcpyAsync(stats + 3 * (long)done, c->stats_dev,
                                  sizeof(double) * 3 * (slot + 1), hipMemcpyDeviceToHost,
                                  c->stream));
            done = s + 1;
        }
    }
    if (c->timing == 2) HIPCHK(hipEventRecord(c->ev[15], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->timing == 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[14], c->ev[15]));
        c->tms[7] += ms;
    }
    return ASTR_OK;
}

int astr_rk3_step(astr_ctx* c, double dt, double* stats)
{
    return rk3_steps(c, 1, dt, stats, false);
}

int astr_rk3_steps(astr_ctx* c, int32_t nsteps, double dt, double* stats)
{
    return rk3_steps(c, nsteps, dt, stats, true);
}

int astr_bench_copy(astr_ctx* c, int32_t reps, double* ms_per_copy, int64_t* bytes_moved)
{
    if (!c || !ms_per_copy || !bytes_moved || reps < 1) return fail(ASTR_E_ARG, "bad argument");
    TRY(set_dev(c));
    const long n2 = c->nfield / 2;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream,
                       (const double2*)c->scratch, (double2*)c->scratch2, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[14], c->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream,
                           (const double2*)c->scratch, (double2*)c->scratch2, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[15], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[15]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[14], c->ev[15]));
    *ms_per_copy = ms / reps;
    *bytes_moved = (int64_t)(2 * n2 * 16);
    return ASTR_OK;
}



int astr_batch_destroy(astr_batch* b)
{
    if (!b) return ASTR_OK;
    (void)hipSetDevice(b->device);
    if (b->root) (void)hipStreamSynchronize(b->root);
    delete b->qe;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->root) (void)hipStreamDestroy(b->root);
    delete b;
    return ASTR_OK;
}

int astr_batch_create(astr_ctx* const* ctxs, int32_t n, astr_batch** out)
{
    if (!ctxs || !out || n < 1) return fail(ASTR_E_ARG, "bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        TRY(ready_batch(ctxs[i]));
        if (ctxs[i]->device != ctxs[0]->device)
            return fail(ASTR_E_ARG, "batch members must share one device");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(ASTR_E_ARG, "a context appears twice in the batch");
    }
    auto* b = new astr_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->device = ctxs[0]->device;
    auto cleanup = [&](int rc) {
        astr_batch_destroy(b);
        return rc;
    };
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->root, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(ASTR_E_HIP, "batch stream set-up failed"));
    b->join.assign(n, nullptr);
    for (int i = 0; i < n; ++i)
        if (hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(ASTR_E_HIP, "batch event set-up failed"));
    *out = b;
    return ASTR_OK;
}

// one RK3 step of every member, each on its own stream, forked from and
// joined back into the root stream under capture
static int batch_capture(astr_batch* b, double dt)
{
    if (b->exec) {
        HIPCHK(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        HIPCHK(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    // members' own queued work must be done before the graph reads their fields
    for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamBeginCapture(b->root, hipStreamCaptureModeThreadLocal));
    int rc = ASTR_OK;
    if (hipEventRecord(b->fork, b->root) != hipSuccess) rc = fail(ASTR_E_HIP, "fork record failed");
    for (size_t i = 0; rc == ASTR_OK && i < b->ctx.size(); ++i) {
        astr_ctx* c = b->ctx[i];
        if (hipStreamWaitEvent(c->stream, b->fork, 0) != hipSuccess) {
            rc = fail(ASTR_E_HIP, "fork wait failed");
            break;
        }
        const double t0 = c->time;  // the replays advance it, not the capture
        rc = do_step(c, dt, nullptr, false);
        c->time = t0;
        if (rc == ASTR_OK && (hipEventRecord(b->join[i], c->stream) != hipSuccess ||
                              hipStreamWaitEvent(b->root, b->join[i], 0) != hipSuccess))
            rc = fail(ASTR_E_HIP, "join failed");
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->root, &g);
    if (rc != ASTR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(ASTR_E_HIP, "graph capture failed");
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dt = dt;
    b->built = true;
    return ASTR_OK;
}

// the queue engine runs the central scheme of the main path
static int queue_ok(astr_batch* b)
{
    for (astr_ctx* c : b->ctx) {
        if (c->p.convection != 0 || c->p.variant != 0)
            return fail(ASTR_E_UNSUPPORTED,
                        "the line-queue engine runs the central scheme of the main path");
        for (int d = 0; d < 3; ++d)
            if (!c->lines_ok[d]) return fail(ASTR_E_UNSUPPORTED, "axis too long for the line sweeps");
    }
    return ASTR_OK;
}

static void drop_graph(astr_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->built = false;
}

// one step of the queue engine captured into a graph
static int queue_capture(astr_batch* b, double dt)
{
    drop_graph(b);
    for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    if (!b->qe) TRY(queue_build(b));
    TRY(queue_set_dt(b, dt));
    HIPCHK(hipStreamBeginCapture(b->root, hipStreamCaptureModeThreadLocal));
    const int rc = queue_step(b, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->root, &g);
    if (rc != ASTR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(ASTR_E_HIP, "graph capture failed");
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dt = dt;
    b->built = true;
    return ASTR_OK;
}

// members changed by a setter since the graph was captured: every member
// is checked again, and the graph (and the queue engine's arguments) are
// rebuilt
static int batch_refresh(astr_batch* b)
{
    std::vector<unsigned long> g;
    for (astr_ctx* c : b->ctx) {
        TRY(ready_batch(c));
        g.push_back(c->gen);
    }
    if (g != b->gens) {
        drop_graph(b);
        if (b->qe) {
            // keep the tuning across the rebuild
            int tuned[3][LMAX + 1];
            memcpy(tuned, b->qe->tuned, sizeof tuned);
            TRY(queue_build(b));
            memcpy(b->qe->tuned, tuned, sizeof tuned);
            TRY(queue_configure(b));
        }
        b->gens = g;
    }
    return ASTR_OK;
}

int astr_batch_set_mode(astr_batch* b, int32_t mode)
{
    if (!b || mode < 0 || mode > 2)
        return fail(ASTR_E_ARG, "mode must be 0 (graph), 1 (streams) or 2 (line queues)");
    if (mode == 2) TRY(queue_ok(b));
    if (mode != b->mode) drop_graph(b);
    b->mode = mode;
    return ASTR_OK;
}

int astr_batch_autotune(astr_batch* b, double dt)
{
    if (!b) return fail(ASTR_E_ARG, "null batch");
    HIPCHK(hipSetDevice(b->device));
    TRY(queue_ok(b));
    TRY(batch_refresh(b));
    for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    if (!b->qe) TRY(queue_build(b));
    TRY(queue_set_dt(b, dt));
    BatchQueue* qe = b->qe;
    const int nm = (int)b->ctx.size();
    // the members' q and primitives, restored after every trial step
    std::vector<double*> bk(nm, nullptr);
    std::vector<hipEvent_t> ev(2 * (size_t)b->nqueue, nullptr);
    int rc = ASTR_OK;
    auto copy_state = [&](bool save) -> int {
        for (int i = 0; i < nm && rc == ASTR_OK; ++i) {
            astr_ctx* c = b->ctx[i];
            const size_t fb = sizeof(double) * c->nfield;
            for (int f = 0; f < 11; ++f) {
                double* dev = f < 5 ? c->q[f] : c->prim[f - 5];
                double* bak = bk[i] + (size_t)f * c->nfield;
                if (hipMemcpyAsync(save ? bak : dev, save ? dev : bak, fb, hipMemcpyDeviceToDevice,
                                   b->root) != hipSuccess)
                    return fail(ASTR_E_HIP, "state copy failed");
            }
        }
        return ASTR_OK;
    };
    for (int i = 0; i < nm && rc == ASTR_OK; ++i)
        if (hipMalloc(&bk[i], sizeof(double) * 11 * b->ctx[i]->nfield) != hipSuccess)
            rc = fail(ASTR_E_NOMEM, "tuning backup allocation failed");
    for (size_t e = 0; e < ev.size() && rc == ASTR_OK; ++e)
        if (hipEventCreate(&ev[e]) != hipSuccess) rc = fail(ASTR_E_HIP, "event");
    if (rc == ASTR_OK) rc = copy_state(true);
    double t[3][LMAX + 1][QC_COUNT] = {};
    bool seen[3][LMAX + 1] = {};
    // a warm-up step (every kernel loaded), then one step per configuration
    for (int cfg = -1; cfg < QC_COUNT && rc == ASTR_OK; ++cfg) {
        for (int p = 0; p < 3; ++p)
            for (int l = 0; l <= LMAX; ++l) qe->tuned[p][l] = cfg < 0 ? QC_PERSISTENT : cfg;
        rc = queue_configure(b);
        if (rc == ASTR_OK) rc = queue_step(b, &ev);
        if (rc == ASTR_OK && hipStreamSynchronize(b->root) != hipSuccess)
            rc = fail(ASTR_E_HIP, "tuning step failed");
        if (rc == ASTR_OK && cfg >= 0) {
            int qi = 0;
            for (int k : b->seq) {
                if (k < 0) continue;
                const QLaunchHost& h = qe->launches[k];
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[2 * qi], ev[2 * qi + 1]) != hipSuccess) {
                    rc = fail(ASTR_E_HIP, "event time");
                    break;
                }
                t[pass_slot(h.pass)][h.L][cfg] += 1e3 * ms;
                seen[pass_slot(h.pass)][h.L] = true;
                ++qi;
            }
        }
        if (rc == ASTR_OK) rc = copy_state(false);
    }
    if (rc == ASTR_OK && hipStreamSynchronize(b->root) != hipSuccess)
        rc = fail(ASTR_E_HIP, "tuning restore failed");
    for (int p = 0; p < 3; ++p)
        for (int l = 0; l <= LMAX; ++l) {
            qe->tuned[p][l] = -1;
            if (!seen[p][l] || rc != ASTR_OK) continue;
            int best = 0;
            for (int k = 0; k < QC_COUNT; ++k) {
                qe->tune_us[p][l][k] = t[p][l][k];
                if (t[p][l][k] < t[p][l][best]) best = k;
            }
            qe->tuned[p][l] = best;
        }
    if (rc == ASTR_OK) rc = queue_configure(b);
    for (double* x : bk) (void)hipFree(x);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    drop_graph(b);
    return rc;
}

int astr_batch_get_tuning(astr_batch* b, double* rows, int32_t cap, int32_t* nrows)
{
    if (!b || !nrows || (cap > 0 && !rows)) return fail(ASTR_E_ARG, "null argument");
    int n = 0;
    if (b->qe)
        for (int p = 0; p < 3; ++p)
            for (int l = 0; l <= LMAX; ++l) {
                if (b->qe->tuned[p][l] < 0) continue;
                if (n < cap) {
                    double* r = rows + (size_t)n * (3 + QC_COUNT);
                    r[0] = p;
                    r[1] = l;
                    r[2] = b->qe->tuned[p][l];
                    for (int k = 0; k < QC_COUNT; ++k) r[3 + k] = b->qe->tune_us[p][l][k];
                }
                ++n;
            }
    *nrows = n;
    return ASTR_OK;
}

int astr_batch_set_tuning(astr_batch* b, const int32_t* choice)
{
    if (!b || !choice) return fail(ASTR_E_ARG, "null argument");
    TRY(queue_ok(b));
    HIPCHK(hipSetDevice(b->device));
    for (int p = 0; p < 3; ++p)
        for (int l = 0; l <= LMAX; ++l)
            if (choice[p * (LMAX + 1) + l] >= QC_COUNT)
                return fail(ASTR_E_ARG, "configuration out of range");
    TRY(batch_refresh(b));
    if (!b->qe) TRY(queue_build(b));
    for (int p = 0; p < 3; ++p)
        for (int l = 0; l <= LMAX; ++l) b->qe->tuned[p][l] = choice[p * (LMAX + 1) + l];
    TRY(queue_configure(b));
    drop_graph(b);
    return ASTR_OK;
}

int astr_batch_rk3_steps(astr_batch* b, int32_t nsteps, double dt)
{
    if (!b || nsteps < 0) return fail(ASTR_E_ARG, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    TRY(batch_refresh(b));
    if (b->mode == 2) {
        if (!b->built || b->dt != dt) TRY(queue_capture(b, dt));
        else
            for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
        for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(b->exec, b->root));
        HIPCHK(hipStreamSynchronize(b->root));
        for (astr_ctx* c : b->ctx)
            for (int s = 0; s < nsteps; ++s) c->time = c->time + dt;
        return ASTR_OK;
    }
    if (b->mode == 1) {
        // every member on its own stream, one step of all members at a time;
        // the runtime spreads the streams over its hardware queues
        for (int s = 0; s < nsteps; ++s)
            for (astr_ctx* c : b->ctx) TRY(do_step(c, dt, nullptr, false));
        for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
        return ASTR_OK;
    }
    if (!b->built || b->dt != dt) TRY(batch_capture(b, dt));
    else
        for (astr_ctx* c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(b->exec, b->root));
    HIPCHK(hipStreamSynchronize(b->root));
    for (astr_ctx* c : b->ctx)
        for (int s = 0; s < nsteps; ++s) c->time = c->time + dt;
    return ASTR_OK;
}

int astr_set_spacing(astr_ctx* c, double dx, double dy, double dz)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    if (!(dx > 0.0 && dy > 0.0 && dz > 0.0)) return fail(ASTR_E_ARG, "spacings must be positive");
    ++c->gen;
    c->h[0] = dx;
    c->h[1] = dy;
    c->h[2] = dz;
    c->have_h = true;
    return ASTR_OK;
}

int astr_set_bc(astr_ctx* c, const int32_t* bctype, int32_t udf)
{
    if (!c || !bctype) return fail(ASTR_E_ARG, "null argument");
    bool any = false, ext = false;
    for (int f = 0; f < 6; ++f) {
        const bool wall = bctype[f] == 41 || bctype[f] == 42;
        if (bctype[f] != 0 && bctype[f] != 1 && bctype[f] != 23 && bctype[f] != 50 &&
            bctype[f] != 60 && bctype[f] != 21 && bctype[f] != 31 && bctype[f] != 421 &&
            bctype[f] != 411 && !wall)
            return fail(ASTR_E_UNSUPPORTED,
                        "bctype 1 (nothing done), 0 (user condition), 23 (gcnscbc), 50 "
                        "(zeroextrap), 60 (symmetry), 41 and 42 (noslip walls), 411 and 421 "
                        "(slip walls), 21 (outflow), 31 (fixbc) are built; the inflow, far-field "
                        "and other NSCBC conditions of src/bc.F90 are not");
        if (wall && f >= 4)
            return fail(ASTR_E_UNSUPPORTED, "noslip walls act on faces 1-4 (src/bc.F90:6281-6919)");
        if (bctype[f] == 21 && f != 1 && f != 3)
            return fail(ASTR_E_UNSUPPORTED, "outflow acts on faces 2 and 4 (src/bc.F90:3379-3592)");
        if ((bctype[f] == 421 || bctype[f] == 411) && f != 2 && f != 3)
            return fail(ASTR_E_UNSUPPORTED,
                        "the slip walls act on faces 3 and 4 (src/bc.F90:6921-7313)");
        const bool ex = bctype[f] == 50 || bctype[f] == 60 || wall || bctype[f] == 21 ||
                        bctype[f] == 31 || bctype[f] == 421 || bctype[f] == 411;
        if (ex && f >= 4 && c->g.n[2] == 0)
            return fail(ASTR_E_UNSUPPORTED,
                        "zeroextrap and symmetry on the k faces need a three-dimensional box");
        any = any || bctype[f] == 0;
        ext = ext || ex;
    }
    if (any && ext)
        return fail(ASTR_E_UNSUPPORTED,
                    "zeroextrap, symmetry and walls are not combined with a user condition");
    if (udf != 0 && udf != ASTR_UDF_RIEMANN2D)
        return fail(ASTR_E_ARG, "udf must be 0 (none) or ASTR_UDF_RIEMANN2D");
    if (any && udf == 0) return fail(ASTR_E_ARG, "bctype 0 needs a user condition (udf)");
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant has its own halos");
    for (int f = 0; f < 6; ++f) c->bctype[f] = bctype[f];
    c->udf = any ? udf : 0;
    ++c->gen;
    return ASTR_OK;
}

int astr_set_twall(astr_ctx* c, const double* twall)
{
    if (!c || !twall) return fail(ASTR_E_ARG, "null argument");
    for (int f = 0; f < 6; ++f)
        if (!(twall[f] > 0.0)) return fail(ASTR_E_ARG, "wall temperatures must be positive");
    for (int f = 0; f < 6; ++f) c->twall[f] = twall[f];
    ++c->gen;
    return ASTR_OK;
}

int astr_set_xslip(astr_ctx* c, double xslip)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    if (!std::isfinite(xslip)) return fail(ASTR_E_ARG, "xslip must be finite");
    c->xslip = xslip;
    return ASTR_OK;
}

int astr_set_lengths(astr_ctx* c, double lx, double ly, double lz)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    if (!(lx > 0.0) || !(ly > 0.0) || !(lz >= 0.0)) return fail(ASTR_E_ARG, "lengths must be positive");
    c->len[0] = lx;
    c->len[1] = ly;
    c->len[2] = lz;
    ++c->gen;
    return ASTR_OK;
}

int astr_set_time(astr_ctx* c, double time)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    c->time = time;
    return ASTR_OK;
}

int astr_get_time(astr_ctx* c, double* time)
{
    if (!c || !time) return fail(ASTR_E_ARG, "null argument");
    *time = c->time;
    return ASTR_OK;
}

int astr_boucon(astr_ctx* c)
{
    TRY(ready(c));
    if (c->p.variant == 1) return fail(ASTR_E_UNSUPPORTED, "the mini-app variant runs whole steps only");
    return do_boucon(c);
}

int astr_synchronize(astr_ctx* c)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    TRY(set_dev(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ASTR_OK;
}

int astr_set_timing(astr_ctx* c, int32_t on)
{
    if (!c) return fail(ASTR_E_ARG, "null context");
    if (on < 0 || on > 2) return fail(ASTR_E_ARG, "timing mode must be 0, 1 or 2");
    c->timing = on;
    for (int f = 0; f < 9; ++f) c->tms[f] = 0.0;
    return ASTR_OK;
}

int astr_get_timing(astr_ctx* c, double* ms9)
{
    if (!c || !ms9) return fail(ASTR_E_ARG, "null argument");
    for (int f = 0; f < 9; ++f) ms9[f] = c->tms[f];
    return ASTR_OK;
}

static int field_op(astr_ctx* c, int dir, const double* f, double* df, int pass)
{
    if (!c || !f || !df) return fail(ASTR_E_ARG, "null argument");
    if (dir < 1 || dir > ndims_of(c)) return fail(ASTR_E_ARG, "dir must be 1..3 (1..2 in 2-D)");
    TRY(set_dev(c));
    std::vector<double> stage;
    TRY(upload_field(c, c->scratch, f, stage));
    if (pass == PASS_DERIV_FIELD && c->p.deriv_explicit) {
        FieldXArgs a{};
        a.g = c->g;
        a.dir = dir - 1;
        a.ntype = dir == 1 ? c->p.npdci : dir == 2 ? c->p.npdcj : c->p.npdck;
        a.fin = c->scratch;
        a.fout = c->scratch2;
        const long n = n_interior(c);
        hipLaunchKernelGGL(k_deriv_explicit_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, a);
        HIPCHK(hipGetLastError());
    } else {
        SweepArgs a = sweep_args(c, dir - 1, pass);
        a.fin = c->scratch;
        a.fout = c->scratch2;
        TRY(launch_sweep(c, a));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    TRY(download_field(c, c->scratch2, df, stage, true));
    return ASTR_OK;
}

int astr_deriv_field(astr_ctx* c, int32_t dir, const double* f, double* df)
{
    return field_op(c, dir, f, df, PASS_DERIV_FIELD);
}

int astr_filter_field(astr_ctx* c, int32_t dir, const double* f, double* ff)
{
    return field_op(c, dir, f, ff, PASS_FILTER_FIELD);
}

int astr_flux_field(astr_ctx* c, int32_t dir, int32_t wind, int32_t shock, const double* f,
                    double* fh)
{
    if (!c || !f || !fh) return fail(ASTR_E_ARG, "null argument");
    if (dir < 1 || dir > ndims_of(c)) return fail(ASTR_E_ARG, "dir must be 1..3 (1..2 in 2-D)");
    if (wind != 1 && wind != -1) return fail(ASTR_E_ARG, "wind must be +1 or -1");
    if (c->p.convection == 0)
        return fail(ASTR_E_UNSUPPORTED, "the interface fluxes need an upwind context (convection 1 or 2)");
    TRY(set_dev(c));
    const int d = dir - 1;
    const int nts[3] = {c->p.npdci, c->p.npdcj, c->p.npdck};
    std::vector<double> stage;
    TRY(upload_field(c, c->scratch, f, stage));
    HIPCHK(hipMemsetAsync(c->scratch2, 0, sizeof(double) * c->nfield, c->stream));
    const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
    if (c->p.convection == 2) {
        CmpArgs ca{};
        ca.g = c->g;
        ca.dir = d;
        ca.ntype = nts[d];
        ca.ndims = ndims_of(c);
        ca.a0 = 0;
        ca.a1 = c->g.n[da];
        ca.b0 = 0;
        ca.b1 = c->g.n[db];
        ca.na = ca.a1 + 1;
        ca.first = c->cmp_first[d];
        ca.last = c->cmp_last[d];
        ca.bfac = c->p.bfacmpld;
        ca.comp0 = wind > 0 ? 0 : 5;
        ca.ncomp = 1;
        ca.sw[ca.comp0] = c->scratch;
        ca.fh[ca.comp0] = c->scratch2;
        ca.ac[0] = c->cmp_ac[d][0];
        ca.ac[1] = c->cmp_ac[d][1];
        const long nl = (long)ca.na * (ca.b1 + 1);
        hipLaunchKernelGGL(k_flux_compact, dim3((unsigned)((nl + 127) / 128)), dim3(128), 0,
                           c->stream, ca);
    } else {
        ReconsArgs ra{};
        ra.g = c->g;
        ra.dir = d;
        ra.ntype = nts[d];
        // convrsduwd's interfaces is-1..ie, with is = 1 at a bounded low end
        // and ie = n-1 at a bounded high end (src/parallel.F90:1043-1160)
        ra.lo = (nts[d] == 1 || nts[d] == 4) ? 0 : -1;
        ra.hi = (nts[d] == 2 || nts[d] == 4) ? c->g.n[d] - 1 : c->g.n[d];
        ra.plus = wind > 0;
        ra.recon = c->p.recon_schem;
        ra.shock = shock != 0;
        ra.bfac = c->p.bfacmpld;
        ra.f = c->scratch;
        ra.fh = c->scratch2;
        const long n = (long)(ra.hi - ra.lo + 1) * (c->g.n[da] + 1) * (c->g.n[db] + 1);
        hipLaunchKernelGGL(k_recons_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, ra);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    TRY(download_field(c, c->scratch2, fh, stage, false));
    return ASTR_OK;
}

int astr_adjoint_record(astr_ctx* c, int32_t nsteps, double dt, int32_t ckpt, double* J)
{
    TRY(ready(c));
    if (!J) return fail(ASTR_E_ARG, "null argument");
    if (nsteps < 0 || ckpt < 1) return fail(ASTR_E_ARG, "nsteps >= 0 and ckpt >= 1 required");
    if (c->p.convection != 0 || c->p.variant != 0 || c->g.n[2] == 0 || !all_periodic(c) ||
        c->curv)
        return fail(ASTR_E_UNSUPPORTED, "the adjoint is built for the central path on periodic 3-D "
                                        "boxes with rectilinear metrics");
    if (c->udf || has_face_bc(c))
        return fail(ASTR_E_UNSUPPORTED, "the adjoint takes no boundary condition");
    c->rec_nsteps = -1;
    const long nseg = nsteps > 0 ? (nsteps + ckpt - 1) / ckpt : 0;
    const long nstates = nseg + 1;  // the segment starts and the final state
    const long seglen = std::min<long>(ckpt, std::max(nsteps, 1));
    const long need = ADJ_FIELDS + nstates * STATE_FIELDS + 3 * seglen * STAGE_FIELDS;
    if (need > c->adj_nf) {
        (void)hipFree(c->adj_base);
        c->adj_base = nullptr;
        c->adj_nf = 0;
        if (hipMalloc(&c->adj_base, sizeof(double) * c->nfield * need) != hipSuccess)
            return fail(ASTR_E_NOMEM, "device allocation of the adjoint workspace failed");
        c->adj_nf = need;
        // the stage slots' cells that no stage writes (edges and corners of
        // the halo) hold zeros rather than whatever the allocation held
        HIPCHK(hipMemsetAsync(c->adj_base, 0, sizeof(double) * c->nfield * need, c->stream));
    }
    // forward, keeping the start of every segment; the last segment's
    // stages are stored as they run
    for (long seg = 0; seg < nseg; ++seg) {
        const long t0 = seg * ckpt, t1 = std::min<long>(nsteps, t0 + ckpt);
        TRY(save_state(c, seg));
        if (seg == nseg - 1) {
            TRY(segment_fwd(c, nstates, 3 * (t1 - t0), dt));
        } else {
            for (long t = t0; t < t1; ++t) TRY(do_step(c, dt, nullptr, false, true));
        }
    }
    TRY(save_state(c, nseg));
    // J at the final state after a qswap (the kept state is the one before)
    TRY(do_qswap(c));
    TRY(do_grad(c));
    double st[3];
    TRY(do_stats(c, st));
    *J = st[1];
    c->rec_nsteps = nsteps;
    c->rec_ckpt = ckpt;
    c->rec_dt = dt;
    c->rec_t_end = c->time;
    return ASTR_OK;
}

int astr_adjoint_reverse(astr_ctx* c, double* qbar, double* pbar, double* ms)
{
    TRY(ready(c));
    if (!qbar || !pbar) return fail(ASTR_E_ARG, "null argument");
    if (c->rec_nsteps < 0) return fail(ASTR_E_ARG, "no trajectory recorded (astr_adjoint_record)");
    const long nsteps = c->rec_nsteps, ckpt = c->rec_ckpt;
    const double dt = c->rec_dt;
    const long nseg = nsteps > 0 ? (nsteps + ckpt - 1) / ckpt : 0;
    const long nstates = nseg + 1;
    HIPCHK(hipEventRecord(c->ev[13], c->stream));
    // the objective at the kept final state, then segment by segment:
    // recompute the segment's stages from its kept state (the last one is
    // stored already), then reverse them
    TRY(load_state(c, nseg));
    double J = 0.0;
    TRY(objective_adj(c, &J));
    for (long seg = nseg - 1; seg >= 0; --seg) {
        const long t0 = seg * ckpt, t1 = std::min<long>(nsteps, t0 + ckpt);
        if (seg != nseg - 1) {
            TRY(load_state(c, seg));
            TRY(segment_fwd(c, nstates, 3 * (t1 - t0), dt));
        }
        for (long t = t1 - 1; t >= t0; --t)
            for (int s = 3; s >= 1; --s)
                TRY(stage_adj(c, stage_slot(c, nstates, 3 * (t - t0) + (s - 1)), s, dt));
    }
    HIPCHK(hipEventRecord(c->ev[14], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[14]));
    if (ms) {
        float b = 0.f;
        HIPCHK(hipEventElapsedTime(&b, c->ev[13], c->ev[14]));
        *ms = b;
    }
    // the adjoint of the initial state, haloed host layout
    const AdjFields f = adj_fields(c);
    std::vector<double> stage, one((size_t)n_haloed(c));
    const long nh = n_haloed(c);
    for (int m = 0; m < 5; ++m) {
        TRY(download_field(c, f.qb[m], one.data(), stage, false));
        memcpy(qbar + m * nh, one.data(), sizeof(double) * nh);
    }
    for (int m = 0; m < 6; ++m) {
        TRY(download_field(c, f.pb[m], one.data(), stage, false));
        memcpy(pbar + m * nh, one.data(), sizeof(double) * nh);
    }
    // leave the device state at the end of the forward steps, as the
    // record left it (after the objective's qswap and gradients)
    TRY(load_state(c, nseg));
    TRY(do_qswap(c));
    TRY(do_grad(c));
    c->time = c->rec_t_end;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ASTR_OK;
}

int astr_adjoint_enstrophy(astr_ctx* c, int32_t nsteps, double dt, int32_t ckpt, double* J,
                           double* qbar, double* pbar, double* ms2)
{
    if (!c || !J || !qbar || !pbar) return fail(ASTR_E_ARG, "null argument");
    TRY(set_dev(c));
    HIPCHK(hipEventRecord(c->ev[12], c->stream));
    TRY(astr_adjoint_record(c, nsteps, dt, ckpt, J));
    HIPCHK(hipEventRecord(c->ev[15], c->stream));
    double rev = 0.0;
    TRY(astr_adjoint_reverse(c, qbar, pbar, &rev));
    if (ms2) {
        float a = 0.f;
        HIPCHK(hipEventElapsedTime(&a, c->ev[12], c->ev[15]));
        ms2[0] = a;
        ms2[1] = rev;
    }
    return ASTR_OK;
}

int astr_line_adjoint(astr_ctx* c, int32_t op, int32_t dir, int32_t wrap, const double* gbar,
                      double* fbar)
{
    if (!c || !gbar || !fbar) return fail(ASTR_E_ARG, "null argument");
    if (op != 0 && op != 1) return fail(ASTR_E_ARG, "op must be 0 (derivative) or 1 (filter)");
    if (dir < 1 || dir > ndims_of(c)) return fail(ASTR_E_ARG, "dir must be 1..3 (1..2 in 2-D)");
    if (op == 0 && c->p.deriv_explicit)
        return fail(ASTR_E_UNSUPPORTED, "the transposed derivative is the compact one");
    TRY(set_dev(c));
    std::vector<double> stage;
    TRY(upload_field(c, c->scratch, gbar, stage));
    HIPCHK(hipMemsetAsync(c->scratch2, 0, sizeof(double) * c->nfield, c->stream));
    const double* fin[1] = {c->scratch};
    double* fout[1] = {c->scratch2};
    TRY(launch_sweep_t(c, dir - 1, op == 0 ? OP_DERIV : OP_FILTER, 1, fin, fout, nullptr, wrap != 0,
                       false));
    HIPCHK(hipStreamSynchronize(c->stream));
    return download_field(c, c->scratch2, fbar, stage, false);
}

}  // extern "C"
