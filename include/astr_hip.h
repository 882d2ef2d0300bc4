/*
 * astr_hip.h - C ABI of the MI355X (gfx950) FP64 solver for ASTR's
 * per-timestep right-hand-side path.
 *
 * The Fortran host keeps its module state (src/commarray.F90) and calls
 * these entry points through iso_c_binding (astr_gpu_amd/fortran/
 * astr_hip_mod.f90); Python binds the same symbols with ctypes.  Host
 * arrays are passed in the reference's own layout: Fortran column-major,
 * i fastest, lower bound -hm (hm = 5) on every haloed axis, variable index
 * last, e.g. q(-hm:im+hm,-hm:jm+hm,-hm:km+hm,1:5).  Arrays without halo
 * (qrhs, dvel, dtmp) are dimensioned (0:im,0:jm,0:km,...).
 *
 * Every entry point returns 0 on success or an ASTR_E_* code; the Fortran
 * shim turns a nonzero code into mpistop, which is what the reference does
 * on any error (e.g. src/mainloop.F90:381).
 *
 * Reference citations are relative to the ASTR source tree.
 */
#ifndef ASTR_HIP_H
#define ASTR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ASTR_OK = 0,
    ASTR_E_ARG = 1,         /* bad argument (null pointer, bad direction) */
    ASTR_E_UNSUPPORTED = 2, /* configuration outside what the solver does */
    ASTR_E_HIP = 3,         /* a HIP runtime call failed */
    ASTR_E_NOMEM = 4,       /* device allocation failed */
    ASTR_E_STATE = 5        /* call out of order (e.g. step before upload) */
};

/* Solver configuration: the values refcal/readinput/solvrinit put in
 * module commvar (src/solver.F90:28-172, src/comsolver.F90:47-146).
 * Layout is fixed so that a Fortran bind(c) derived type can mirror it. */
typedef struct astr_params {
    int32_t im, jm, km;          /* nodes 0..im etc. (commvar im,jm,km)   */
    int32_t hm;                  /* halo width of host arrays, must be 5  */
    int32_t numq;                /* must be 5 (no species, no k-omega)    */
    int32_t npdci, npdcj, npdck; /* closure type per axis; the solver
                                    steps periodic boxes (3 on all axes)  */
    int32_t variant;             /* 0: src/mainloop.F90 stage order; 1: the
                                    mini-app tgvsolver.F90 (bchomo halos,
                                    df/dx, filter after the update)       */
    int32_t reserved;
    double alfa_filter;          /* 0.49 in every deck                    */
    double reynolds, mach, prandtl, gamma;
    double ref_tem;              /* tempconst = sutherland / ref_tem      */
    double sutherland;           /* 110.3 (src/solver.F90:121)            */
    double beta_halo;            /* 1.11 (src/comsolver.F90:121)          */
    double beta_bound;           /* 0.98 (src/comsolver.F90:121)          */
    /* scheme switches of the input deck (src/readwrite.F90:466-771)       */
    int32_t convection;          /* 0: central compact, conschm '643c'
                                    (convrsdcal6); 1: explicit upwind,
                                    conschm odd order with 'e' (convrsduwd) */
    int32_t deriv_explicit;      /* difschm(4:4): 0 'c' compact (df_compact),
                                    1 'e' explicit (diff6ec)               */
    int32_t recon_schem;         /* -1, 0 SUW, 1 WENO, 2 WENO-Z, 3 MP,
                                    5 MP-LD, 6 ROUND (src/flux.F90:269)    */
    int32_t lchardecomp;         /* characteristic projection (1/0)       */
    int32_t lfilter;             /* filter q at every stage (1/0)         */
    int32_t diffterm;            /* viscous terms (1/0)                   */
    double shkcrt;               /* Ducros sensor threshold               */
    double bfacmpld;             /* MP-LD dissipation weight              */
} astr_params;

typedef struct astr_ctx astr_ctx;

/* Fill *p with the Taylor-Green deck values (examples/Taylor_Green_Vortex/
 * datin/input.tgv, Re 1600, Ma 0.1, '643c' central scheme, filter and
 * viscous terms on) for an n x n x n periodic box.  The upwind fields take
 * the Riemann2D deck's values (examples/Riemann2D/datin/input.riemann2d:
 * recon_schem 3, lchardecomp on, bfacmpld 0.7, shkcrt 0.01). */
void astr_default_params(astr_params *p, int32_t n);

/* Replaces solvrinit (src/comsolver.F90:47-146) and allocommarray: builds
 * the pre-factored line systems (fd_scheme_initiate, src/derivative.F90:
 * 63-158; compact_filter_initiate + filter_coefficient_cal, src/filter.F90:
 * 31-100,299-432) and allocates the device fields on HIP device `device`.
 * km = 0 is a two-dimensional box (an upwind convection with
 * deriv_explicit 1 and npdck 3; right-hand sides and whole steps without
 * the filter).  npdc 4 (bounded at both ends on one domain) is the union of
 * ntype 1 at the low end and ntype 2 at the high end in the upwind path:
 * recons_exp (convrsduwd), the Ducros sensor's clamps, diff6ec and the
 * ranges is..ie = 1..n-1.  The reference gives ntype 4 only to a single rank
 * along a non-periodic axis, where convrsduwd stops on j and k
 * (src/solver.F90:808-822) and is/ie stay unset (src/parallel.F90:
 * 1070-1078); it runs bounded axes on several ranks, whose result the
 * union reproduces (the explicit upwind flux has no closure at a rank
 * interface). */
int astr_create(const astr_params *p, int32_t device, astr_ctx **out);
int astr_destroy(astr_ctx *ctx);

/* Metrics from gridgeom (src/geom.F90:99-917): jacob(-hm:im+hm,...) and
 * dxi(-hm:im+hm,...,1:3,1:3).  Rectilinear metrics (off-diagonal dxi zero,
 * dxi(d,d) a function of the d-th index alone) are kept as one line per
 * axis; any other (curvilinear) metrics as the nine components per node,
 * which the central path of the main variant on three-dimensional boxes
 * reads (gradcal's combination of the axes, the fluxes' projections on
 * dxi(d,:)); other contexts return ASTR_E_UNSUPPORTED for them, as do the
 * adjoint and the batch. */
int astr_upload_geom(astr_ctx *ctx, const double *jacob, const double *dxi);

/* Module state of src/commarray.F90 in and out: q(.,.,.,1:5), rho, vel
 * (.,.,.,1:3), prs, tmp, all haloed.  Download writes the full haloed
 * arrays as the device holds them. */
int astr_upload_state(astr_ctx *ctx, const double *q, const double *rho,
                      const double *vel, const double *prs, const double *tmp);
int astr_download_state(astr_ctx *ctx, double *q, double *rho, double *vel,
                        double *prs, double *tmp);

/* --- stage pieces: one per call site of time_integration_rk ---------- */

/* call filterq, src/mainloop.F90:395 -> src/comsolver.F90:514-632 */
int astr_filterq(astr_ctx *ctx);
/* call qswap, src/mainloop.F90:415 -> src/parallel.F90:4849-5319 */
int astr_qswap(astr_ctx *ctx);
/* qrhs=0 (src/mainloop.F90:404) then call rhscal (src/solver.F90:184-283:
 * gradcal, convrsdcal6, qrhs=-qrhs, diffrsdcal6).  The inviscid and viscous
 * flux derivatives of one axis are taken as one derivative of their
 * difference (the operator is linear; only the summation order differs).
 * With convection 1: gradcal, ducrossensor when recon_schem is 5 or
 * lchardecomp is set, convrsduwd, qrhs=-qrhs; any closure type, halos used
 * as given; qrhs is set on is..ie of each axis (ntype 1: 1..n, 2: 0..n-1,
 * 3: 0..n, 4: 1..n-1) and zero elsewhere. */
int astr_rhscal(astr_ctx *ctx);
/* RK3 combine of stage s (1..3) and updatefvar, src/mainloop.F90:423-476;
 * stage 1 also saves qsave = q*jacob. */
int astr_rk3_update(astr_ctx *ctx, int32_t stage, double dt);
/* kenergycal, enstophycal, diss_rate_cal (src/statistic.F90:870-1042) of
 * the current device state -> out[0..2]. */
int astr_stats(astr_ctx *ctx, double *out3);
/* The three statistics of astr_stats, written to out3_dev (3 doubles of
 * device memory on the context's device) by kernels queued on the
 * context's stream; the call returns without waiting (astr_synchronize, or
 * an event on that stream, orders a later read).  No counterpart in the
 * reference: it is for callers that cannot allocate or wait, such as work
 * captured into a HIP graph or a tight serving loop.  It allocates
 * nothing: its partial sums live in a fixed device array of the library
 * (110.6 KB per GPU, kept until the process ends) that every call on that
 * GPU shares, so calls that overlap on one GPU (on two streams) race.
 * Boxes of 2^31 nodes or more are refused (ASTR_E_UNSUPPORTED). */
int astr_stats_device(astr_ctx *ctx, double *out3_dev);
/* maxmincal's max_q (src/statistic.F90:737-763): the largest value of each
 * q(m), m = 1..5, over the nodes 0..n of the current device state ->
 * qmax[0..4]; the columns statout writes to flowstate.dat for flow types
 * without their own statistics (src/statistic.F90:835-838).  Two- and
 * three-dimensional boxes. */
int astr_qmax(astr_ctx *ctx, double *qmax5);

/* --- whole steps: time_integration_rk, src/mainloop.F90:298-595 ------- */

/* One RK3 step: three-dimensional periodic boxes (any scheme) and
 * two-dimensional upwind boxes (any closures, boundary conditions by
 * astr_set_bc, no filter).  If stats is non-null (three-dimensional boxes)
 * it receives (ke, enstrophy, dissipation) of the stage-1 state and
 * velocity gradients, which rkfirst (src/mainloop.F90:608-688) evaluates
 * after rhscal; they are reduced on the device and copied to the host
 * after the steps, so computing them does not stall the stream.  The central path applies the RK3 combine inside the last
 * flux sweep of each stage, so qrhs is left holding the x+y part;
 * astr_rhscal gives the whole qrhs.  The step advances the time by dt. */
int astr_rk3_step(astr_ctx *ctx, double dt, double *stats);
/* nsteps RK3 steps back to back; stats (nullable) receives 3 per step. */
int astr_rk3_steps(astr_ctx *ctx, int32_t nsteps, double dt, double *stats);

/* --- results of the last rhscal, for parity checks --------------------- */
/* lshock(0:im,0:jm,0:km) of the last ducrossensor (src/commcal.F90:196-353)
 * as bytes 0/1; all zero when the upwind path has not run the sensor. */
int astr_download_shock(astr_ctx *ctx, uint8_t *lshock);
/* qrhs(0:im,0:jm,0:km,1:5), dvel(0:im,0:jm,0:km,1:3,1:3),
 * dtmp(0:im,0:jm,0:km,1:3) */
int astr_download_rhs(astr_ctx *ctx, double *qrhs, double *dvel, double *dtmp);

/* --- field-level line operators --------------------------------------- */

/* fds%central along axis dir (1,2,3) for every line of a haloed field:
 * f(-hm:im+hm,-hm:jm+hm,-hm:km+hm) -> df(0:im,0:jm,0:km), index-space
 * derivative (src/derivative.F90:16-46,171-306; diff6ec, :350-413, when
 * deriv_explicit is set), closure npdc{i,j,k} of the context.  Halos are
 * used as given. */
int astr_deriv_field(astr_ctx *ctx, int32_t dir, const double *f, double *df);
/* compact_filter along axis dir for every line (src/filter.F90:112-285):
 * f(-hm:...) haloed in, ff(0:im,0:jm,0:km) out.  Halos used as given. */
int astr_filter_field(astr_ctx *ctx, int32_t dir, const double *f, double *ff);
/* The upwind interface values of one split-flux field along axis dir for
 * every line: f(-hm:...) haloed in (halos used as given), fh(-hm:...)
 * haloed out, fh(i) the value at the interface i+1/2, zero elsewhere.
 * wind +1 reconstructs F+ (stencil f(i-3:i+4)), -1 F- (f(i+4:i-3)).
 *   convection 2: flux_compact (src/flux.F90:125-256), the 543c compact
 *     upwind system of the context's closure, rows -1..dim (0.. at a
 *     bounded low end, ..dim-1 at a bounded high end);
 *   convection 1: recons_exp (src/flux.F90:269-350) with the context's
 *     recon_schem at the interfaces is-1..ie of convrsduwd
 *     (src/solver.F90:549-700); shock (0/1) is the lshock flag MP-LD reads.
 * ASTR_E_UNSUPPORTED for a central context (convection 0). */
int astr_flux_field(astr_ctx *ctx, int32_t dir, int32_t wind, int32_t shock, const double *f,
                    double *fh);
/* The transpose of one of the two operators above (op 0: the compact
 * derivative, 1: the filter) along dir: gbar(-hm:..) holds the adjoint of
 * the output on 0..im etc. (halo entries are ignored), fbar(-hm:..)
 * receives the adjoint of the input: f = B^T A^-T g per line.  wrap 1 folds
 * the adjoint of the halos into the nodes they copy periodically (f(-k) =
 * f(n-k), f(n+k) = f(k), as filterq's dataswap fills them) and leaves the
 * halo entries zero.  No counterpart in the reference: the building block
 * of the discrete adjoint. */
int astr_line_adjoint(astr_ctx *ctx, int32_t op, int32_t dir, int32_t wrap, const double *gbar,
                      double *fbar);

/* --- discrete adjoint ---------------------------------------------------- */

/* The forward half of the gradient: nsteps plain RK3 steps from the
 * current (uploaded) state, keeping the state at the start of every ckpt
 * steps and every stage of the last segment in the adjoint workspace (device
 * memory the context owns), then J = the enstrophy of the final state after
 * a qswap (as astr_adjoint_enstrophy).  The device state is the one the
 * plain steps reach, bit for bit, followed by that qswap.  A later
 * astr_adjoint_reverse uses what is kept, whatever runs in between;
 * another record replaces it.  No counterpart in the reference. */
int astr_adjoint_record(astr_ctx *ctx, int32_t nsteps, double dt, int32_t ckpt, double *J);

/* The reverse half: qbar and pbar (as astr_adjoint_enstrophy) of the
 * trajectory astr_adjoint_record kept; the segments before the last are
 * recomputed from their kept states.  ms (nullable) receives the
 * milliseconds of the device work.  The device state is left at the
 * recorded final state.  ASTR_E_ARG if nothing is recorded.  No
 * counterpart in the reference. */
int astr_adjoint_reverse(astr_ctx *ctx, double *qbar, double *pbar, double *ms);

/* Gradient of the enstrophy after nsteps RK3 steps with respect to the
 * current (uploaded) state: astr_adjoint_record, then
 * astr_adjoint_reverse.  The steps are the plain forward (the central
 * path on a periodic box, filter and viscous terms as configured).  J is the
 * enstrophy statistic of the final state after a qswap (enstophycal,
 * src/statistic.F90:870-925: 1/2 mean over the nodes 1..n of rho |omega|^2).
 * qbar(-hm:..,1:5) and pbar(-hm:..,1:6) receive dJ/dq and dJ/d(rho, u, v,
 * w, p, T) of the initial state.  The interior primitives are their own
 * inputs, because the filter leaves them stale.  Hand-written reverse
 * kernels of every stage piece:
 *   - the transposed compact solves of the filter, gradient and flux
 *     sweeps;
 *   - the Jacobian-transposes of the axis fluxes and of q2fvar;
 *   - qswap's copies and averages reversed;
 *   - the RK3 combine reversed.
 * The steps are reversed segment by segment: the state is kept every ckpt
 * steps, and a segment's stages are recomputed from it before they are
 * reversed.  ms2 (nullable) receives the milliseconds of the recorded
 * forward (with the objective) and of the reversal.  The device state is left at
 * the end of the forward steps, with the time advanced.  No counterpart in
 * the reference. */
int astr_adjoint_enstrophy(astr_ctx *ctx, int32_t nsteps, double dt, int32_t ckpt, double *J,
                           double *qbar, double *pbar, double *ms2);

/* Kernel time since timing was (re)enabled with astr_set_timing(ctx, 1),
 * in milliseconds, per kernel family (HIP events on the solver's stream):
 * [0] filter sweeps, [1] boucon and qswap, [2] gradient, [3] central:
 * sigma/qflux and the flux of each axis, upwind steps: the shock sensor,
 * [4] statistics, [5] flux sweeps (central: inviscid and viscous flux
 * derivatives, the z sweep applying the RK3 combine; upwind: interface
 * fluxes), [6] update (central: primitives; upwind: RK3 combine and
 * primitives), [7] total, [8] shock sensor of astr_rhscal.
 * astr_rk3_step(s) fill 0-7; astr_rhscal of an upwind context fills 2, 5,
 * 7 and 8.  astr_rk3_steps restarts the count at each call.  Mode 1 puts
 * events between the families and waits for them at every stage.
 * astr_set_timing(ctx, 2) instead brackets each astr_rk3_step(s) call with
 * two events and nothing in between: [7] adds up the device time of those
 * calls (the other entries stay zero).  0 turns timing off. */
int astr_set_timing(astr_ctx *ctx, int32_t on);
int astr_get_timing(astr_ctx *ctx, double *ms9);

/* --- batched ensemble -------------------------------------------------- */

/* A group of contexts (same device; each created, with geometry and state
 * uploaded; three-dimensional periodic boxes without a user boundary
 * condition) that advance together.
 * astr_batch_rk3_steps captures one RK3 step of every member into one HIP
 * graph, each member's kernels on its own stream branching from the batch's
 * root stream so that boxes of any size run concurrently, and replays the
 * graph once per step (it is rebuilt when dt changes).  No counterpart in
 * the reference, which advances one box per run. */
typedef struct astr_batch astr_batch;
int astr_batch_create(astr_ctx *const *ctxs, int32_t n, astr_batch **out);
int astr_batch_destroy(astr_batch *b);
int astr_batch_rk3_steps(astr_batch *b, int32_t nsteps, double dt);
/* mode 0 (default): the graph above; mode 1: no graph, every member's
 * kernels launched directly on its own stream, all members one step at a
 * time (the runtime spreads the streams over its hardware queues);
 * mode 2: the line-queue engine (members on the central scheme of the main
 * path only, else ASTR_E_UNSUPPORTED): each line sweep is one launch per
 * rows-per-lane class over every member, whose workgroups pull tiles of
 * lines of any member from a device work queue (persistent form) or take
 * one tile each, and each pointwise kernel (qswap, the fluxes G, the
 * primitives) is one launch over every member; one such step is captured
 * into a HIP graph and replayed.  Every mode gives each member the steps
 * astr_rk3_step gives it, bit for bit.  A setter called on a member after
 * a graph was captured (spacing, boundary conditions, metrics) makes the
 * next call capture again. */
int astr_batch_set_mode(astr_batch *b, int32_t mode);

/* Autotuner of the line-queue engine: runs one trial step per
 * configuration of the sweep launches, timing every launch with HIP events,
 * and keeps for each sweep kind (filter, gradient, flux) and rows-per-lane
 * class L the fastest: persistent grid pulling from the queue or one
 * workgroup per item (bit 1), the members' tiles interleaved or in runs,
 * largest member first (bit 2), an item of one quantity or of all the
 * quantities of a tile (bit 4).  The members' state is restored after each
 * trial, so the call leaves every member where it was.  dt is the step the
 * trials take. */
int astr_batch_autotune(astr_batch *b, double dt);
/* The tuning kept: *nrows rows (at most cap written) of 11 doubles: sweep
 * kind (0 filter, 1 gradient, 2 flux), L, the chosen configuration, and the
 * trial time in microseconds of configurations 0..7 (summed over the axes
 * and stages of that kind and class). */
int astr_batch_get_tuning(astr_batch *b, double *rows, int32_t cap, int32_t *nrows);
/* Apply recorded choices without trials: choice[kind * 17 + L] (kind 0..2,
 * L 0..16), each a configuration 0..7 or -1 for the default (persistent,
 * runs, whole tiles). */
int astr_batch_set_tuning(astr_batch *b, const int32_t *choice);

/* Uniform spacing of the mini-app variant (variant 1,
 * miniapps/tgv_solver_3d/tgvsolver.F90), which divides every derivative
 * by dx, dy, dz (tgvsolver.F90:420-470) instead of using metrics; its
 * steps need it.  Variant 1 runs whole steps (astr_rk3_step(s)) only. */
int astr_set_spacing(astr_ctx *ctx, double dx, double dy, double dz);

/* Boundary conditions, boucon (src/bc.F90:327-411), applied in every stage
 * before qswap (src/mainloop.F90:413-415) at the step's start time.
 * bctype[0..5] per face (i=0, i=im, j=0, j=jm, k=0, k=km): 1 nothing done
 * (periodic, or a bounded face left as it is), 0 the user condition udf.
 * udf ASTR_UDF_RIEMANN2D is udf_bc of the 2-D Riemann example
 * (examples/Riemann2D/user_define_module/userdefine_riemann2d.F90:421-668)
 * on the grid of gridcube(1,1,0): shock-moving quadrant states on faces
 * 1 and 3, zero-slope extrapolation on faces 2 and 4, the north-east state
 * in the corner (im,jm).  23 is gcnscbc (src/bc.F90:7907-8137, boucon's
 * bctype 23 with prs_t = pinf): the ghost-cell NSCBC on face 2 (i=im) and
 * face 4 (j=jm), as the reference has it (it acts on no other face),
 * over every k: two-dimensional boxes on the upwind path and
 * three-dimensional boxes on the central path (a bounded x or y axis);
 * batched members and the adjoint take no boundary condition
 * (ASTR_E_UNSUPPORTED there).  It reads node coordinates of gridcube's
 * grid (astr_set_lengths).  50 is zeroextrap (src/bc.F90:2331-2627): the
 * boundary node from its two inner neighbours by extrapolate_2o with
 * dv = 0; 60 is symmetry (src/bc.F90:2639-2971): the same with the velocity
 * normal to the face zero (rectilinear metrics only).  Both on any face
 * (the k faces of three-dimensional boxes only).  41 and 42 are the noslip
 * walls (src/bc.F90:6281-6919): at rest, the pressure extrapolated, the
 * temperature the face's wall temperature (41, astr_set_twall) or
 * extrapolated (42), on faces 1-4 as the reference has them.  21 is outflow
 * (src/bc.F90:3379-3592) on faces 2 and 4: face 2 copies node n-1, face 4
 * extrapolates and, below the sound speed, relaxes the pressure towards
 * pinf at the rate 0.25 dt (dt of the latest step or astr_rk3_update).
 * 31 is fixbc (src/bc.F90:7562-7648): faces 3 and 4 held at its two fixed
 * states, the other faces left alone.  421 is slipadibwall
 * (src/bc.F90:7114-7313) on faces 3 and 4: pressure, temperature and u
 * extrapolated, w = 0, v = 0 on face 3 and extrapolated on face 4, the
 * density from p/T (its unused wall blowing and k-omega branches aside).
 * 411 is slipisotwall (src/bc.F90:6921-7112): on face 3 the same where the
 * node's gridcube x = lx/im*i is at most xslip (astr_set_xslip), else at
 * rest at the face's wall temperature; on face 4 the isothermal noslip wall.
 * None of these is combined with a user condition; batched members and the
 * adjoint take none.  Faces act in the reference's order 1..6.  The other bctypes of
 * bc.F90 (inflow, far field, the other NSCBCs) return
 * ASTR_E_UNSUPPORTED.  Default: all faces 1. */
#define ASTR_UDF_RIEMANN2D 1
int astr_set_bc(astr_ctx *ctx, const int32_t *bctype, int32_t udf);
/* The wall temperature of each face for bctype 41 (twall of src/bc.F90:29,
 * read with the face's bctype by readinput, src/readwrite.F90:565).
 * Default 1 on every face.  No counterpart call in the reference: the deck
 * holds it. */
int astr_set_twall(astr_ctx *ctx, const double *twall6);
/* The slip-nonslip point xslip of bctype 411 (src/bc.F90:29, read with the
 * face's bctype by readinput, src/readwrite.F90:569).  Default 0.  No
 * counterpart call in the reference: the deck holds it. */
int astr_set_xslip(astr_ctx *ctx, double xslip);
/* The box lengths of gridcube (src/gridgeneration.F90:233-267: x(i) =
 * lx/im*i, y(j) = ly/jm*j, z(k) = lz/km*k), which replace the reference's
 * node coordinates x(:,:,:,1:3) where a boundary condition reads them
 * (gcnscbc's spacing x(n-1)-x(n-2) and xmax-xmin).  Default 1, 1, 0 (the
 * Riemann example's gridcube(1,1,0)).  No counterpart call in the
 * reference, which holds the coordinates themselves. */
int astr_set_lengths(astr_ctx *ctx, double lx, double ly, double lz);
/* call boucon alone (src/mainloop.F90:413), at the context's time */
int astr_boucon(astr_ctx *ctx);
/* The simulation time (commvar time) that boucon sees; each whole step
 * adds dt after its three stages (time=time+deltat, src/mainloop.F90:200),
 * batched steps included.  Starts at 0. */
int astr_set_time(astr_ctx *ctx, double time);
int astr_get_time(astr_ctx *ctx, double *time);

/* Wait for the work queued on the context's stream (entry points that
 * only queue kernels, such as astr_rhscal, return before it ends).  No
 * counterpart in the reference, whose calls are synchronous. */
int astr_synchronize(astr_ctx *ctx);

/* Measurement utility with no counterpart in the reference: copies one
 * field-sized device buffer (16-byte accesses) `reps` times and reports the
 * time per copy and the bytes read plus written per copy.  It is the stream
 * bandwidth reference of the roofline and calibrates the HBM traffic
 * counters. */
int astr_bench_copy(astr_ctx *ctx, int32_t reps, double *ms_per_copy, int64_t *bytes_moved);

/* Human-readable text of the last error on this thread. */
const char *astr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
