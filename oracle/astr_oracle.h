/*
 * CPU restatement of the ASTR per-step hot path (test infrastructure only).
 *
 * This library is the oracle that the HIP path is checked against.  It is
 * never linked into, loaded by or called from the product library
 * (astr_amd/csrc); only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as a checker.
 *
 * Every routine follows a named piece of the reference and cites it
 * (file:line relative to the astr-code/astr tree).  Pinning of each routine
 * (what checks it beyond itself) is listed in the header comment of
 * astr_oracle.c.
 */
#ifndef ASTR_ORACLE_H
#define ASTR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_HM 5 /* halo depth, src/commvar.F90:184 */

typedef struct or_cfg {
  int im, jm, km;          /* computed lattice 0..im etc.; km==0 means 2-D */
  double alfa_filter;      /* 0.49 */
  double beta_halo;        /* 1.11 */
  double beta_bound;       /* 1.09 */
  double gamma, mach, reynolds, prandtl, ref_tem;
  double suth;             /* Sutherland constant in ref_tem units: 110.4/ref_tem */
  double dx, dy, dz;
  double dt;
} or_cfg;

/* ---- line operators (one line of f(-hm:n+hm), stored f[t+hm]) ---- */
void or_thomas_prep(int n, const double *a, const double *c, double *ac);
void or_thomas_solve(int n, const double *ac, double *d, double *x);
int  or_deriv_nrows(int n);
void or_deriv_lhs(int n, double *a, double *c);
void or_deriv_rhs(int n, const double *f, double *d);
void or_df_compact(int n, const double *f, double *df);
/* ntype 1: physical boundary at 0, 2: at n, 3: interfaces, 4: both physical;
 * return -1 for another ntype */
int or_df_compact_nt(int n, int ntype, const double *f, double *df);
void or_filter_coefs_b(double alfa, double beta_bound, double *coefb, double *coef6i,
                       double *coef8i);
int or_compact_filter_nt(int n, int ntype, double alfa, double beta_halo, double beta_bound,
                         const double *f, double *ff);
int  or_filter_nrows(int n);
void or_filter_coefs(double alfa, double beta_halo, double beta_bound,
                     double *coef10i, double *coefh);
void or_filter_lhs(int n, double alfa, double beta_halo, double *a, double *c);
void or_filter_rhs(int n, const double *coef10i, const double *coefh,
                   const double *f, double *d);
void or_compact_filter(int n, double alfa, double beta_halo, double beta_bound,
                       const double *f, double *ff);

/* ---- field state: q[5][N], prims rho,u,v,w,p,T each [N]; N=(im+1)(jm+1)(km+1),
 *      index i + (im+1)*(j + (jm+1)*k) ---- */
typedef struct or_state or_state;
or_state *or_create(const or_cfg *cfg);
void or_destroy(or_state *s);
long or_npoints(const or_state *s);
double *or_q(or_state *s);       /* 5*N */
double *or_prims(or_state *s);   /* 6*N: rho,u,v,w,p,T */
double *or_qrhs(or_state *s);    /* 5*N */
double *or_dvel(or_state *s);    /* 9*N, component m, direction d at (3*d+m)... see .c */
double *or_dtmp(or_state *s);    /* 3*N */

void or_tgv_init(or_state *s, double noise_amp, uint64_t seed);
void or_set_q(or_state *s, const double *q);      /* q -> prims by q2fvar */
void or_q2fvar_all(or_state *s);
void or_rhscal(or_state *s);
void or_filterq(or_state *s);
void or_stacal(or_state *s, double *ke, double *ens);
void or_rk3_steps(or_state *s, int nsteps, double *ke, double *ens);
void or_diss_rate(or_state *s, double *diss);
void or_rk3_steps3(or_state *s, int nsteps, double *ke, double *ens, double *diss);

/* ---- main solver's stage (astr_oracle_main.c): arrays with halos of depth
 *      OR_HM, as the reference keeps them ---- */
typedef struct or_main or_main;
or_main *or_main_create(const or_cfg *cfg);
void or_main_destroy(or_main *s);
long or_main_npoints(const or_main *s);
long or_main_nhalo(const or_main *s);
double *or_main_qh(or_main *s);      /* 5 x halo box, i fastest */
double *or_main_primsh(or_main *s);  /* 6 x halo box: rho, u, v, w, p, T */
double *or_main_qrhs(or_main *s);    /* 5 N */
double *or_main_dvel(or_main *s);    /* 9 N, component m, direction n at m + 3n */
double *or_main_dtmp(or_main *s);    /* 3 N */
void or_main_set_metrics(or_main *s, const double *jac, const double *dxi);
void or_main_set_npdc(or_main *s, const int *npdc);
void or_main_set_q(or_main *s, const double *q);
void or_main_set_q_only(or_main *s, const double *q);
void or_main_get(or_main *s, double *q, double *prims);
void or_main_updatefvar(or_main *s);
void or_main_filterq(or_main *s);
void or_main_qswap(or_main *s);
void or_main_gradcal(or_main *s);
void or_main_rhscal(or_main *s);
void or_main_stats(or_main *s, double *ke, double *ens, double *diss);
void or_main_rk3_steps(or_main *s, int nsteps, double *ke, double *ens, double *diss);
int or_main_set_conv(or_main *s, int upwind, int recon, int chardec, double bfac, double shkcrt);
double *or_main_ssf(or_main *s);              /* N */
unsigned char *or_main_lshock(or_main *s);    /* N */

/* pointwise pieces of the upwind-biased convection (astr_oracle_upwind.c) */
void or_sw_split(double gamma, double mach, double jac, const double *dxi, const double *q,
                 const double *pv, double *fp, double *fm);
double or_recons(int scheme, const double *f, double bfac, int shock);
int or_chardecomp(double gamma, const double *l, const double *r, double *rev, double *lev);
double or_mp5(const double *u, double ul, int discont);
void or_cuw_lu(int n, double bfac, int plus, double *ac);
void or_cuw_flux(int n, double bfac, int plus, const double *ac, const double *f, double *fh);

/* counter-based noise shared with the device path (not a reference routine) */
double or_gauss(uint64_t seed, uint32_t i, uint32_t j, uint32_t k, uint32_t c);

#ifdef __cplusplus
}
#endif
#endif
