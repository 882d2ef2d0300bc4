/*
 * CPU restatement of the ASTR per-timestep right-hand-side path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * it may be loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else.  The product path never calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the ASTR source tree).  Arrays that carry halos follow the Fortran layout of
 * the reference: f(-hm:n+hm) for a line, and a(-hm:im+hm,-hm:jm+hm,-hm:km+hm)
 * with i fastest for fields, hm = 5 (src/commvar.F90:190).
 */
#ifndef ASTR_ORACLE_H
#define ASTR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HM 5

/* ---------------- banded solver (src/commfunc.F90) ---------------- */
/* tridiagonal_thomas_proprocess, src/commfunc.F90:752-774 */
void orc_thomas_prep(const double *a, const double *c, int n, double *ac);
/* tridiagonal_thomas_solver, src/commfunc.F90:790-813; d is modified */
void orc_thomas_solve(const double *ac, double *d, int n, double *x);

/* ---------------- line operators ---------------- */
/* fd_scheme_initiate (643c), src/derivative.F90:63-158.
 * Returns the row count; first/last receive the node range.  a,c,ac have
 * room for dim+3 rows (ac is 3 per row, row-major [row][3]). */
int orc_fd_init(int ntype, int dim, int *first, int *last,
                double *a, double *c, double *ac);
/* df_compact + compact_fd_rhs, src/derivative.F90:171-306.
 * f points at node -hm of f(-hm:dim+hm); df receives df(0:dim). */
void orc_df_compact(int ntype, int dim, const double *f, double *df);
/* diff6ec, src/derivative.F90:350-413 */
void orc_diff6ec(int ntype, int dim, const double *f, double *df);

/* filter_coefficient_cal, src/filter.F90:299-432.
 * Layout of coef (out, 6+45+55 = 106 doubles):
 *   [0..5]    coef10i(0:5)
 *   [6..50]   coefb(0:4,0:8)  as coefb[k*9+j]
 *   [51..105] coefh(0:4,0:10) as coefh[k*11+j]
 * plus coef6i/coef8i in the trailing 9 (total 115 doubles):
 *   [106..109] coef6i(0:3), [110..114] coef8i(0:4) */
#define ORC_NCOEF 115
void orc_filter_coef(double alfa, double beta_halo, double beta_bound,
                     double *coef);
/* compact_filter_initiate, src/filter.F90:31-100 (note absent). */
int orc_filter_init(int ntype, int dim, double alfa, int *first, int *last,
                    double *a, double *c, double *ac);
/* compact_filter + compact_filter_rhs, src/filter.F90:112-285 */
void orc_compact_filter(int ntype, int dim, double alfa, double beta_halo,
                        double beta_bound, const double *f, double *ff);

/* ---------------- periodic-box solver ---------------- */
typedef struct orc_box orc_box;

/* variant 0: main path (src/mainloop.F90 time_integration_rk with rk3,
 *            src/solver.F90 rhscal with convrsdcal6/diffrsdcal6,
 *            src/comsolver.F90 filterq, src/parallel.F90 qswap).
 * variant 1: mini-app (miniapps/tgv_solver_3d/tgvsolver.F90 rk3).
 * variant 2: the mini-app's stage order and constants with the main path's
 *            right-hand side (metric products, J-weighted fluxes) and RK3
 *            combine (qsave = q J, divided by J): the bridge that carries
 *            the reference-held mini-app series over to the main path's
 *            arithmetic. */
orc_box *orc_box_new(int im, int jm, int km, int variant,
                     double reynolds, double mach, double prandtl,
                     double gamma, double ref_tem, double alfa_filter,
                     double lx, double ly, double lz);
void orc_box_free(orc_box *b);
/* Sutherland constant numerator (110.3 in src/solver.F90:121, 110.4 in
 * tgvsolver.F90:133) and the filter's boundary beta (0.98 main, 1.09 mini)
 * are set by the variant; they can be overridden. */
void orc_box_set_sutherland(orc_box *b, double s_over_tref_numerator);
/* metrics of the grid (haloed jacob and dxi(..,d,e) at (d-1+3(e-1))*n3);
 * the box starts with those of its uniform grid */
void orc_box_set_metrics(orc_box *b, const double *jacob, const double *dxi);
/* closure per axis (3 periodic, 1/2/4 bounded); main variant; 0 or -1 */
int orc_box_set_ntype(orc_box *b, int n0, int n1, int n2);
/* bctype of faces 1..6 for the main variant's steps: 0 (nothing done), 23
 * (gcnscbc, src/bc.F90:7907-8140, acting on faces 2 and 4), 50
 * (zeroextrap, src/bc.F90:2331-2627), 60 (symmetry, src/bc.F90:
 * 2639-2971, rectilinear grids), 41 or 42 (noslip, noslip_adibatic,
 * src/bc.F90:6281-6919, faces 1-4) or 21 (outflow, src/bc.F90:3379-3592,
 * faces 2 and 4); applied in face order; 0 or -1 */
int orc_box_set_bc(orc_box *b, const int *bctype);
/* wall temperature of each face for bctype 41 (twall of src/bc.F90:29) */
void orc_box_set_twall(orc_box *b, const double *twall);
/* slip-nonslip point of bctype 411 (xslip of src/bc.F90:29) */
void orc_box_set_xslip(orc_box *b, double xslip);

/* Set the Taylor-Green initial field: tgvini (src/initialisation.F90:1079-1149)
 * followed by updateq (src/initialisation.F90:30-162 / src/fludyna.F90:254)
 * for variant 0, flowfield_init (tgvsolver.F90:176-234) for variant 1. */
void orc_box_init_tgv(orc_box *b);
/* Load / store the interior (0:im,0:jm,0:km) of the state arrays.
 * Each buffer holds (im+1)*(jm+1)*(km+1) values, i fastest.
 * which: 0..4 q(1..5), 5 rho, 6..8 vel(1..3), 9 prs, 10 tmp.          */
void orc_box_get(const orc_box *b, int which, double *out);
void orc_box_set(orc_box *b, int which, const double *in);
/* The whole haloed array (im+11)*(jm+11)*(km+11) of the same fields. */
void orc_box_get_haloed(const orc_box *b, int which, double *out);
/* dvel(m,n) for which = 3*(m-1)+(n-1), dtmp(n) for which = 9+(n-1),
 * qrhs(m) for which = 12+(m-1). */
void orc_box_get_deriv(const orc_box *b, int which, double *out);

/* One RK3 step; stats (ke, enstrophy, dissipation) are computed at
 * stage 1 after the right-hand side, where rkfirst (src/mainloop.F90:608)
 * or stacal (tgvsolver.F90:893) evaluates them.  stats may be NULL. */
void orc_box_step(orc_box *b, double dt, double *stats);
/* The individual stage pieces of the main variant, for operator tests. */
void orc_box_filterq(orc_box *b);
void orc_box_qswap(orc_box *b);
void orc_box_rhscal(orc_box *b);
void orc_box_updatefvar(orc_box *b);
void orc_box_stats(orc_box *b, double *stats);

int orc_omp_threads(void);

/* ---------------- explicit upwind convection path (upwind.c) ---------- */

typedef struct {
    int im, jm, km;      /* km = 0: two-dimensional box */
    int npdc[3];         /* closure type of each axis, 1..4 */
    int recon_schem;     /* -1, 0 SUW, 1 WENO, 2 WENO-Z, 3 MP, 5 MP-LD, 6 ROUND */
    int lchardecomp;     /* characteristic projection where the shock flag is set */
    double bfacmpld;     /* MP-LD dissipation weight */
    double shkcrt;       /* Ducros sensor threshold */
    double gamma, mach;
    double lx, ly, lz;   /* gridcube lengths (x(i) = lx/im*i): gcnscbc's coordinates */
    double twall[6];     /* wall temperature of each face (bctype 41) */
    double deltat;       /* the step's time step, which outflow (bctype 21) reads */
    double xslip;        /* slip-nonslip point of bctype 411 */
} orc_uwd;

double orc_minmod2(double a, double b);
double orc_minmod4(double a, double b, double c, double d);
double orc_suw3(const double *u);
double orc_suw5(const double *u);
double orc_suw7(const double *u);
double orc_mp5(const double *u, const double *ul, int discont);
double orc_mp7(const double *u);
double orc_mp5ld(const double *u, double weightbw);
double orc_mp7ld(const double *u, double weightbw, int lskt);
double orc_weno5(const double *u);
double orc_weno5z(const double *u);
double orc_weno7(const double *u);
double orc_weno7z(const double *u);
double orc_round(const double *u);
double orc_recons_exp(const double *f, int inode, int dim, int ntype, int reschem, int shock,
                      double bfacmpld);
void orc_sw_split(double gamma, double mach, double rho, const double *vel, double prs,
                  double tmp, const double *q, const double *dxi, double jacob, double *fp,
                  double *fm);
int orc_chardecomp(double gamma, double tmp_l, double ro_l, double p_l, double e_l,
                   const double *vel_l, const double *ddi_l, double tmp_r, double ro_r,
                   double p_r, double e_r, const double *vel_r, const double *ddi_r, double *R,
                   double *Lv);
int orc_iwind8(int i, int n, int imin, int imax, int plus);
void orc_uwd_range(int ntype, int n, int *lo, int *hi);
int orc_ndims(const orc_uwd *u);
void orc_gradcal(const orc_uwd *u, int explicit_fd, const double *vel, const double *tmp,
                 const double *dxi, double *dvel, double *dtmp);
void orc_ducros(const orc_uwd *u, const double *dvel, const double *prs, double *ssf,
                unsigned char *lshock);
int orc_convrsduwd(const orc_uwd *u, const double *q, const double *rho, const double *vel,
                   const double *prs, const double *tmp, const double *dxi, const double *jacob,
                   const unsigned char *lshock, double *qrhs);
int orc_flux_compact_init(int ntype, int dim, int plus, double bfac, int *first, int *last,
                          double *ac);
void orc_flux_compact(int ntype, int dim, int plus, double bfac, const double *f, double *fh);
double orc_mplimiter(const double *f, double fl, int shock, int inode, int dim, int ntype);
int orc_convrsdcmp(const orc_uwd *u, const double *q, const double *rho, const double *vel,
                   const double *prs, const double *tmp, const double *dxi, const double *jacob,
                   const unsigned char *lshock, double *qrhs);
int orc_cmp_rhscal(const orc_uwd *u, int explicit_fd, const double *q, const double *rho,
                   const double *vel, const double *prs, const double *tmp, const double *dxi,
                   const double *jacob, double *qrhs, double *dvel, double *dtmp, double *ssf,
                   unsigned char *lshock);
int orc_uwd_rhscal(const orc_uwd *u, int explicit_fd, const double *q, const double *rho,
                   const double *vel, const double *prs, const double *tmp, const double *dxi,
                   const double *jacob, double *qrhs, double *dvel, double *dtmp, double *ssf,
                   unsigned char *lshock);
double orc_riem2d_shockspeed(double gamma, double mach, double p1, double t1, double p2);
void orc_riem2d_bc(const orc_uwd *u, const int *bctype, double time, double *q, double *rho,
                   double *vel, double *prs, double *tmp);
/* gcnscbc (bctype 23) on face n = 2 (i = im) or 4 (j = jm); other faces
 * are left alone, as in the reference */
void orc_gcnscbc(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                 double *tmp);
/* zeroextrap (bctype 50) on face 1..6: the boundary node from its two inner
 * neighbours by extrapolate_2o with dv = 0 (src/bc.F90:2331-2627) */
void orc_zeroextrap(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                    double *tmp);
/* symmetry (bctype 60) on face 1..6 of a rectilinear grid: zeroextrap's
 * extrapolation with the normal velocity zero (src/bc.F90:2639-2971) */
void orc_symmetry(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp);
/* outflow (bctype 21) on face 2 or 4 (src/bc.F90:3379-3592; u->deltat the
 * step's time step) */
void orc_outflow(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                 double *tmp);
/* slipisotwall (bctype 411) on face 3 or 4 (src/bc.F90:6921-7112) */
void orc_slipisot(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp);
/* slipadibwall (bctype 421) on face 3 or 4 (src/bc.F90:7114-7313) */
void orc_slipadib(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
                  double *tmp);
/* fixbc (bctype 31): faces 3 and 4 held at fixed states (src/bc.F90:7562-7648) */
void orc_fixbc(const orc_uwd *u, int face, double *q, double *rho, double *vel, double *prs,
               double *tmp);
/* noslip (bctype 41, wall temperature u->twall[face-1]) and noslip_adibatic
 * (42) on face 1..4 (src/bc.F90:6281-6919) */
void orc_wall(const orc_uwd *u, int face, int isothermal, double *q, double *rho, double *vel,
              double *prs, double *tmp);
/* boucon: faces 1..6 in order, udf_bc (bctype 0, udf 1) or gcnscbc (23) */
void orc_boucon(const orc_uwd *u, const int *bctype, int udf, double time, double *q,
                double *rho, double *vel, double *prs, double *tmp);
void orc_uwd_qswap(const orc_uwd *u, double *q, double *rho, double *vel, double *prs,
                   double *tmp);
int orc_uwd_rk3_step(const orc_uwd *u, int compact, int explicit_fd, const int *bctype, int udf,
                     double time, double dt, const double *dxi, const double *jacob, double *q,
                     double *rho, double *vel, double *prs, double *tmp);

#ifdef __cplusplus
}
#endif
#endif
