/*
 * astr_hip.h -- C ABI of the MI355X (gfx950) implementation of ASTR's
 * per-step hot path.
 *
 * The reference has no FFI: its hot path is a set of Fortran module
 * procedures working on module-global arrays (src/commarray.F90:273-285,
 * miniapps/tgv_solver_3d/tgvsolver.F90:1-40).  Each entry point below
 * replaces one of those procedures; a Fortran host binds them with
 * ISO_C_BINDING (astr_amd/fortran/astr_hip_mod.F90) under the reference's
 * own subroutine names.
 *
 * Conventions
 *   - Return value 0 = success, nonzero = error; astr_hip_last_error()
 *     returns the message (per host thread).  The reference stops the
 *     program on error (`stop '<msg>'`); the Fortran binding does the same.
 *   - Host arrays use the reference's layout: column-major, i fastest, the
 *     variable index slowest, each spatial index running -hm..n+hm, e.g.
 *     q(-hm:im+hm,-hm:jm+hm,-hm:km+hm,1:5).  hm = 0 passes the computed
 *     lattice 0..im only (halo cells are never read: on one device the
 *     periodic halo is the index wrap the reference's halo copy produces).
 *   - The library owns device memory; every field stays resident in HBM
 *     between calls.  Calls on one context are stream-ordered and not
 *     thread-safe (one host thread per context, as in the reference).
 */
#ifndef ASTR_HIP_H
#define ASTR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* semantics mode: ASTR_SEM_MINIAPP = the stage order, halo refresh and
 * constants of miniapps/tgv_solver_3d/tgvsolver.F90 (rk3 :994-1082).
 * ASTR_SEM_MAIN = the main solver's (astr run) single-rank periodic stage,
 * time_integration_rk (src/mainloop.F90:396-482): filterq at the start of
 * every stage, qswap (duplicate-plane averages and the primitives of the
 * boundary planes, src/parallel.F90:4876-4920, 5018-5060, 5159-5200),
 * gradients and fluxes from the primitives the last updatefvar stored
 * (src/comsolver.F90:244-497, src/solver.F90:2173-2873), fluxes carrying
 * the metrics J and dxi, the RK3 update (q J-weighted), then updatefvar.
 * The metrics are any grid's J and dxi (astr_hip_set_metrics); periodic or
 * closed directions per astr_hip_set_boundary.
 * Batches take ASTR_SEM_MINIAPP cases only. */
#define ASTR_SEM_MINIAPP 0
#define ASTR_SEM_MAIN 1

/* Solver set-up.  Replaces program_init / solver_init / refcal constants
 * (tgvsolver.F90:152-261; src/solver.F90:104-128). */
typedef struct astr_hip_config {
  int im, jm, km;          /* computed lattice 0..im, 0..jm, 0..km; km = 0: 2-D */
  int semantics;           /* ASTR_SEM_MINIAPP or ASTR_SEM_MAIN */
  double alfa_filter;      /* compact filter alpha (0.49); filter.F90:79-80 */
  double beta_halo;        /* halo-row beta (1.11); filter.F90:57-62 */
  double beta_bound;       /* boundary-row beta (1.09; ntype 1,2,4 only) */
  double gamma, mach, reynolds, prandtl;
  double ref_tem;          /* reference temperature (K) */
  double sutherland;       /* Sutherland constant (K): 110.4 in tgvsolver.F90:133,
                              110.3 in the main solver (src/solver.F90:122) */
  double lx, ly, lz;       /* domain lengths; dx = lx/im (tgvsolver.F90:199-201) */
  double deltat;           /* time step */
} astr_hip_config;

typedef struct astr_hip_ctx astr_hip_ctx;

/* device layout of the resident fields (for callers that share device memory) */
typedef struct astr_hip_layout {
  int pitch;               /* doubles between consecutive j rows */
  long long plane;         /* doubles between consecutive k planes */
  long long field_stride;  /* doubles between variables of q / qrhs */
  double *q;               /* device pointer, 5 fields */
  double *qrhs;            /* device pointer, 5 fields */
} astr_hip_layout;

const char *astr_hip_last_error(void);

/* alloccommarray + solver_init: tgvsolver.F90:13-38, :236-261 */
int astr_hip_create(const astr_hip_config *cfg, astr_hip_ctx **out);
int astr_hip_destroy(astr_hip_ctx *ctx);
int astr_hip_set_stream(astr_hip_ctx *ctx, void *hip_stream);
int astr_hip_set_deltat(astr_hip_ctx *ctx, double deltat);
int astr_hip_field_info(astr_hip_ctx *ctx, astr_hip_layout *layout);

/* flowfield_init (tgvsolver.F90:172-234), on the device; noise_amp > 0 adds
 * seeded Gaussian velocity noise (Philox4x32-10 + Box-Muller, not part of
 * the reference) on i,j,k < n, copied to the duplicate planes. */
int astr_hip_init_tgv(astr_hip_ctx *ctx, double noise_amp, unsigned long long seed);

/* Two-dimensional isentropic vortex (km = 0; no reference counterpart, a
 * build-defined initial field for the batched and served workloads): free
 * stream rho = 1, T = 1, velocity (u0, v0); a vortex of strength beta and unit
 * radius at (xc, yc):  du = -beta/(2 pi) (y-yc) e,  dv = beta/(2 pi) (x-xc) e,
 * e = exp((1-r^2)/2),  T = 1 - (gamma-1) Ma^2 beta^2/(8 pi^2) e^2,
 * rho = T^(1/(gamma-1)),  p = rho T/(gamma Ma^2). */
int astr_hip_init_vortex(astr_hip_ctx *ctx, double beta, double xc, double yc, double u0, double v0);

/* set / read the conservative state q(-hm:..,1:5); download also returns the
 * primitives rho,u,v,w,p,T (q2fvar, tgvsolver.F90:92-121) when prims != 0. */
int astr_hip_upload(astr_hip_ctx *ctx, const double *q, int hm);
int astr_hip_download(astr_hip_ctx *ctx, double *q, double *prims, int hm);

/* updatefvar: fludyna.F90:191-242 / q2fvar tgvsolver.F90:92-121 -- the
 * primitives of the current q into the reference's separate arrays
 * rho(-hm:im+hm,-hm:jm+hm,-hm:km+hm), vel(...,1:3), prs, tmp (each
 * nullable), periodic halos of depth hm. */
int astr_hip_updatefvar(astr_hip_ctx *ctx, double *rho, double *vel, double *prs, double *tmp, int hm);

/* ASTR_SEM_MAIN: set the stored primitives (the reference's module arrays
 * rho, vel(...,1:3), prs, tmp, src/commarray.F90:273-285) from host arrays
 * in the layout of astr_hip_updatefvar, e.g. after a flow initialisation
 * that sets them and q = fvar2q of them (flowinit + updateq,
 * src/initialisation.F90:60-140).  Halo cells are not read. */
int astr_hip_set_fvar(astr_hip_ctx *ctx, const double *rho, const double *vel, const double *prs,
                      const double *tmp, int hm);

/* writechkpt / readcheckpoint (src/readwrite.F90:2390-2459, 1381-1531): the
 * reference writes ro, u1, u2, u3, p, t, nstep and time to HDF5 and rebuilds
 * q from them at restart; here one raw little-endian FP64 file with a fixed
 * 64-byte header ("ASTRCHK1", int32 header bytes, im, jm, km, semantics,
 * nfields = 11, int64 nstep, double time) holds ro, u1, u2, u3, p, t (the
 * stored primitives under ASTR_SEM_MAIN) and q1..q5 on the computed points,
 * i fastest, so a resumed run continues bit for bit.  read checks the grid
 * size and the semantics mode; nstep and time (nullable) receive the
 * header's values. */
int astr_hip_write_checkpoint(astr_hip_ctx *ctx, const char *path, long long nstep, double time);
int astr_hip_read_checkpoint(astr_hip_ctx *ctx, const char *path, long long *nstep, double *time);

/* filterq: tgvsolver.F90:791-883 (x, y, z sweeps of the compact filter) */
int astr_hip_filterq(astr_hip_ctx *ctx);

/* gradcal: tgvsolver.F90:450-537; dvel(0:im,0:jm,0:km,1:3,1:3), dtmp(...,1:3) */
int astr_hip_gradcal(astr_hip_ctx *ctx, double *dvel, double *dtmp);

/* bchomo + rhscal: tgvsolver.F90:1039-1041, :962-992; qrhs(0:im,..,1:5) */
int astr_hip_rhscal(astr_hip_ctx *ctx, double *qrhs);

/* steploop/rk3: tgvsolver.F90:944-960, :994-1082.  ke/ens (nullable) receive
 * stacal's kinetic energy and enstrophy (:893-942) at the start of each step. */
int astr_hip_rk3_steps(astr_hip_ctx *ctx, int nsteps, double *ke, double *ens);
/* rk3_steps with the main path's statcal statistics of each step's first
 * stage (statistic.F90:593-595): ke and ens as above, and diss[it] =
 * sum 2 mu (s:s - div^2/3) / (im*jm*km) over i,j,k = 1..im,jm,km
 * (diss_rate_cal, statistic.F90:994-1043).  Any of the three may be NULL. */
int astr_hip_rk3_steps_stats(astr_hip_ctx *ctx, int nsteps, double *ke, double *ens, double *diss);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): astr_hip_rk3_steps_stats with the same results, whose
 * statistics buffers (per-tile partial sums, dissipation block partials, the
 * per-step values) live in a fixed 8 MiB device array, one copy per GPU, held
 * until the process ends.  It calls no hipMalloc/hipFree.  Limit:
 * 2 x (tiles of the last gradient pass) + 1024 + 3 x nsteps doubles must fit
 * in 2^20 (e.g. nsteps <= 170000 at 512^3); larger requests return an error
 * before any work is queued.  Every call on one GPU shares that GPU's array,
 * so calls that overlap on one GPU (other threads, other contexts on other
 * streams) race; serialise them. */
int astr_hip_rk3_steps_stats_noalloc(astr_hip_ctx *ctx, int nsteps, double *ke, double *ens,
                                     double *diss);
int astr_hip_sync(astr_hip_ctx *ctx);

/* qswap, single rank, periodic (parallel.F90:4876-4884, 5018-5030,
 * 5159-5170): the main path's duplicate-plane average q(0) = q(n) =
 * 0.5 (q(0) + q(n)) along i, then j, then k, on the device; the halo copies
 * it also makes are the passes' own index wrap.  The miniapp's stage order
 * (astr_hip_rk3_steps) does not call it. */
int astr_hip_qswap(astr_hip_ctx *ctx);

/* gridgeom: geom.F90:99-692 (single rank, periodic).  x(0:im,0:jm,0:km,1:3)
 * are the grid coordinates; on the device: dx(m,d) = d x_m / d xi_d by the
 * compact derivative with gridsendrecv's single-rank halo
 * (parallel.F90:2859-2864), jacob = det(dx) (:350-381), the symmetric
 * conservative form of jacob * dxi (:395-540), divided by jacob (:664-668).
 * jacob(0:im,0:jm,0:km) and dxi(0:im,0:jm,0:km,1:3,1:3) are nullable. */
int astr_hip_gridgeom(astr_hip_ctx *ctx, const double *x, double *jacob, double *dxi);
/* gridgeom with a boundary type per direction, npdc[d] = npdci/npdcj/npdck
 * (commvar.F90; fds_compact_i/j/k of geom.F90:128-161): 3 periodic, 4
 * physical at both ends, the two a single rank has.  A physical direction
 * takes the compact derivative's boundary rows (derivative.F90:128-139,
 * 229-289) and reads no halo. */
int astr_hip_gridgeom_nt(astr_hip_ctx *ctx, const int *npdc, const double *x, double *jacob,
                         double *dxi);
/* Metrics for the passes, in gridgeom's layout.  The metrics of a uniform
 * Cartesian grid (dxi diagonal and constant, jacob constant, to 1e-12) run
 * on constants, the means of dxi(d,d) and jacob.  Any other grid
 * (ASTR_SEM_MAIN only: stretched, curvilinear) keeps J and all nine dxi(d,n)
 * per point, which the main path's fluxes, gradients and RK update read as
 * the reference does (src/solver.F90:2196-2207, 2623-2635,
 * src/comsolver.F90:280-321, src/mainloop.F90:441-449); the miniapp
 * semantics refuse it. */
int astr_hip_set_metrics(astr_hip_ctx *ctx, const double *jacob, const double *dxi);
/* ASTR_SEM_MAIN: the boundary type of each direction, npdc[d] = npdci,
 * npdcj, npdck of one rank (src/parallel.F90:1042-1106): 3 periodic (the
 * default), 4 physical at both ends.  A physical direction takes the
 * derivative's and the filter's boundary closures (derivative.F90:128-139,
 * 229-289; filter.F90:31-100, 141-142, 175-262, with beta_bound of the
 * config, 0.98 in the main solver) in every pass, reads no halo, and qswap
 * neither averages nor refreshes it.  Boundary conditions (boucon) are not
 * applied. */
int astr_hip_set_boundary(astr_hip_ctx *ctx, const int *npdc);

/* Main path: the convective scheme (rhscal, src/solver.F90:208-237).
 * upwind = 0: the central compact scheme (convrsdcal6, conschm 643c);
 * upwind = 1: the explicit upwind-biased scheme (conschm with an odd first
 * digit and suffix 'e'; convrsduwd, src/solver.F90:548-1201): Steger-Warming
 * splitting (src/riemann.F90:23-166), reconstruction recon_schem = -1, 0
 * SUW7, 1 WENO7, 2 WENO7-Z, 3 MP7, 5 MP7-LD (bandwidth factor bfacmpld),
 * 6 ROUND (recons_exp, src/flux.F90:269-383), characteristic decomposition
 * where the shock sensor flags an interface when lchardecomp
 * (chardecomp, :1958-2162), and the Ducros sensor with threshold shkcrt
 * when recon_schem = 5 or lchardecomp (ducrossensor,
 * src/commcal.F90:196-360) -- the input line "recon_schem, lchardecomp,
 * bfacmpld, shkcrt" of the reference's input file (readwrite.F90:546).
 * upwind = 2: the compact upwind-biased scheme (conschm 543c; convrsdcmp,
 * src/solver.F90:1271-1937): the same splitting, the interface fluxes of the
 * 5th-order compact upwind scheme with bandwidth factor bfacmpld
 * (compact_flux_initiate / flux_compact, src/flux.F90:33-254), limited by
 * MP5 on the 5-point windows (mplimiter, :363-495), characteristic
 * decomposition at every interface when lchardecomp, and the Ducros sensor
 * when lchardecomp (recon_schem is not used).
 * Every direction must be periodic (npdc 3); the scheme is kept until
 * changed. */
int astr_hip_set_convection(astr_hip_ctx *ctx, int upwind, int recon_schem, int lchardecomp,
                            double bfacmpld, double shkcrt);

/* Per-pass timing with HIP events on the context's stream (no reference
 * counterpart; the reference's timerept accumulators, src/utility.F90:27-297,
 * are no-ops).  Slot = pass*3 + direction, pass 0..8 = convective+gradient
 * first/middle/last, viscous middle/last(rhs)/last(RK3+filter), filter, and
 * the two halves of the split last convective+gradient pass (gradients and
 * stress, convective); 9..11 = the upwind-biased convection's split,
 * interface-flux and shock-sensor kernels (the sensor in slot 33); 12 = the
 * first convective+gradient pass with the previous stage's x filter sweep
 * (program 9, slot 36). */
#define ASTR_PROF_SLOTS 39
int astr_hip_profile_begin(astr_hip_ctx *ctx, int max_launches);
int astr_hip_profile_end(astr_hip_ctx *ctx, double *ms, int *count);

/* Launch selection of the line passes (no reference counterpart: the
 * reference's loops have no launch parameters).  A pass is identified by its
 * program (0..6 as in the profile slots) and direction; its selection is a
 * launch shape and a register budget minb (1 or 2 workgroups per CU).  The
 * shapes (256 threads each) are 0: 4 lines x 64 chunks of 5 unknowns,
 * 1: 4 x 64 x 9, 2: 16 x 16 x 5, 3: 8 x 32 x 9, 4: 8 x 32 x 5; a shape fits a
 * line of n intervals when its chunks hold n+7 unknowns.  astr_hip_autotune times
 * every fitting selection of every pass the case launches (reps launches
 * each, on the context's stream, state restored afterwards) and keeps the
 * fastest; get/set read and override single selections. */
int astr_hip_autotune(astr_hip_ctx *ctx, int reps);
/* The last direction's gradient/convective work as one pass (split = 0) or as
 * two lighter passes (split = 1, programs 7 and 8); astr_hip_autotune sets
 * it from the timings. */
int astr_hip_set_last_split(astr_hip_ctx *ctx, int split);
/* Filter-sweep fusion of astr_hip_rk3_steps* (no reference counterpart).
 * filterq's sweeps act on different indices and are linear, so they commute.
 * - Compact-scheme solver's stage: it ends with z (in the RK pass), then y,
 *   then x.  With on != 0 the x sweep of every stage but a call's last runs
 *   inside the next stage's first convective+gradient pass (program 9),
 *   which loads the same x-lines.  Off by default; astr_hip_autotune turns
 *   it on when its timing of program 9 is below those of programs 0 and 6
 *   along x together.
 * - Main path (ASTR_SEM_MAIN), whose stage starts with filterq: with
 *   on != 0 (the default) every RK pass but a call's last also applies the
 *   next stage's sweep along the last direction to the new q, after
 *   updatefvar took its primitives, and that stage's filterq skips it.
 * The state between calls is the same up to rounding either way. */
int astr_hip_set_filter_fusion(astr_hip_ctx *ctx, int on);
int astr_hip_get_last_split(astr_hip_ctx *ctx, int *split);
int astr_hip_get_pass_config(astr_hip_ctx *ctx, int prog, int dir, int *shape, int *minb, float *ms);
int astr_hip_set_pass_config(astr_hip_ctx *ctx, int prog, int dir, int shape, int minb);

/* ------------------------------------------------------------------------
 * Batching (no reference counterpart: the reference runs one case per job).
 * A batch advances many independent contexts together:
 *   - each pass of a stage is one persistent launch per group of cases that
 *     share program, launch shape and register budget;
 *   - with use_graph != 0, a whole RK3 step is captured once as a hipGraph
 *     and replayed.
 * Each case's result equals its own astr_hip_rk3_steps under the same launch
 * selection.
 * Rules:
 *   - the cases stay owned by the caller and must outlive the batch;
 *   - astr_hip_batch_autotune picks one selection per class of cases (same
 *     program and set of fitting shapes) per pass, by timing, and writes it
 *     into the cases' own selections;
 *   - ke/ens (nullable) receive ncases x nsteps values, case-major;
 *   - rk3_steps is rk3_begin (enqueue the steps on the batch's stream and
 *     return) followed by rk3_end (wait, then copy the histories of those
 *     steps); rk3_query sets *done to 1 once the enqueued steps finished, so
 *     a host loop can tend other work meanwhile.
 * ------------------------------------------------------------------------ */
typedef struct astr_hip_batch astr_hip_batch;
int astr_hip_batch_create(astr_hip_ctx *const *cases, int ncases, astr_hip_batch **out);
int astr_hip_batch_destroy(astr_hip_batch *batch);
int astr_hip_batch_autotune(astr_hip_batch *batch, int reps);
int astr_hip_batch_rk3_steps(astr_hip_batch *batch, int nsteps, int use_graph, double *ke, double *ens);
int astr_hip_batch_rk3_begin(astr_hip_batch *batch, int nsteps, int use_graph);
int astr_hip_batch_query(astr_hip_batch *batch, int *done);
int astr_hip_batch_rk3_end(astr_hip_batch *batch, double *ke, double *ens);
int astr_hip_batch_info(astr_hip_batch *batch, int *ngroups, int *launches_per_step);

/* fds%central / compact_filter on a batch of lines: derivative.F90:171-198
 * (op 0, df_compact, scheme 643c) and filter.F90:112-144 (op 1), ntype=3.
 * f holds nlines lines of f(-5:n+5) (n+11 values each), out nlines x (n+1). */
int astr_hip_line_batch(int op, int n, long long nlines, const double *f, double *out,
                        double alfa_filter, double beta_halo);
/* The same with the boundary type ntype of a block's ends, as
 * fd_scheme_initiate / compact_filter_initiate take it
 * (derivative.F90:63-158, filter.F90:31-100): 1 physical boundary at i=0,
 * 2 at i=n, 3 interfaces at both ends (astr_hip_line_batch), 4 physical at
 * both ends.  A physical end takes the derivative's 3rd-order biased and
 * 4th-order compact rows (derivative.F90:229-289) and the filter's
 * one-sided coefb rows, 6th- and 8th-order rows and reset of the boundary
 * value (filter.F90:141-142, 175-250); its halo values are not read.
 * beta_bound is filter_coefficient_cal's beter_bouond (comsolver.F90:121
 * passes 0.98); the filter's LHS takes beta_halo at an interface and 0.98
 * at a physical end. */
int astr_hip_line_batch_nt(int op, int ntype, int n, long long nlines, const double *f,
                           double *out, double alfa_filter, double beta_halo, double beta_bound);

/* ------------------------------------------------------------------------
 * Chemical source terms (per cell), replacing the reference's Cantera calls
 * chemrate (src/thermchem.F90:1553-1575: wi = net production rate * W_k,
 * kg/m^3/s), cpeval (:1342-1428: cp_mass, J/kg/K) and enthpy (:1498-1541:
 * h_k = (H/RT)_k * (8314.2/W_k) * T, J/kg) for a CTML mechanism file
 * (chemread, :59-98).  Arrays are structure-of-arrays: species k of cell i
 * at [k*n + i].
 * ------------------------------------------------------------------------ */
typedef struct astr_chem astr_chem;

const char *astr_hip_chem_last_error(void);
int astr_hip_chem_create(const char *ctml_path, const char *phase_id, astr_chem **out);
int astr_hip_chem_destroy(astr_chem *chem);
/* species count, reaction count, molecular weights (kg/kmol), names (name_len
 * bytes each, NUL-terminated); any output may be null */
int astr_hip_chem_info(astr_chem *chem, int *nsp, int *nrx, double *mw, char *names, int name_len);
/* host buffers: T (K), rho (kg/m^3) or p (Pa) per input_is_p, Y (nsp x n);
 * out wdot (nsp x n), h_mix, cp_mix (n each, nullable), h_k (nsp x n, nullable) */
int astr_hip_chem_rates(astr_chem *chem, long long n, const double *T, const double *rho_or_p,
                        int input_is_p, const double *Y, double *wdot, double *h_mix,
                        double *cp_mix, double *h_k);
/* the same on device buffers already resident in HBM, stream-ordered on stream */
int astr_hip_chem_rates_device(astr_chem *chem, long long n, const double *T,
                               const double *rho_or_p, int input_is_p, const double *Y,
                               double *wdot, double *h_mix, double *cp_mix, double *h_k,
                               void *stream);

/* srccomb (solver.F90:500-537): the chemical source of every computed point
 * added to the species equations, qrhs(i,j,k,6:numq) += wirate * jacob(i,j,k),
 * wirate = chemrate(rho, tmp, spc) = wdot_k W_k (kg/m^3/s), for i, j, k in
 * 0:im, 0:jm, 0:km.  The arrays are the reference's (commarray.F90),
 * column-major: rho, tmp, jacob (-hm:im+hm, -hm:jm+hm, -hm:km+hm); spc the
 * same with a last index 1:nsp; qrhs (0:im, 0:jm, 0:km, 1:numq).  numq must
 * be 5 + nsp.  The first form takes host buffers (qrhs updated in place); the
 * _device form takes buffers in HBM and is stream-ordered on stream. */
int astr_hip_srccomb(astr_chem *chem, int im, int jm, int km, int hm, int numq, const double *rho,
                     const double *tmp, const double *spc, const double *jacob, double *qrhs);
int astr_hip_srccomb_device(astr_chem *chem, int im, int jm, int km, int hm, int numq,
                            const double *rho, const double *tmp, const double *spc,
                            const double *jacob, double *qrhs, void *stream);

/* temperature_calc (thermchem.F90:1775-1786; Cantera setState_UV): T from
 * the mass-specific internal energy e (J/kg) and Y (nsp x n, SoA); T holds
 * the initial guesses on entry and the temperatures on return.  Newton on
 * e(T) = sum_k Y_k (h_k(T) - R T/W_k) to |dT| <= 1e-13 T. */
int astr_hip_chem_temperature(astr_chem *chem, long long n, const double *Y, const double *e, double *T);
int astr_hip_chem_temperature_device(astr_chem *chem, long long n, const double *Y, const double *e,
                                     double *T, void *stream);

/* ------------------------------------------------------------------------
 * 0-D constant-pressure adiabatic reactors (no reference counterpart: the
 * reference's 0dreactor, src/solver.F90:208,256-258, is constant-volume).
 *
 * The model: state (T, Y_k), dY_k/dt = wdot_k W_k / rho,
 * dT/dt = -sum_k h_k dY_k/dt / cp.
 *
 * The integrator: RODAS4, the stiffly accurate L-stable Rosenbrock method of
 * order 4 with an embedded order-3 solution (Hairer and Wanner 1996,
 * Sect. VI.4), with a finite-difference Jacobian and error control at rtol,
 * atol (mass fractions; temperature uses atol 1e-6 K).  reactor_method
 * writes its table: gamma, a21, a31, a32, a41..a43, a51..a54, c21, c31,
 * c32, c41..c43, c51..c54, c61..c65 (26 values; n must be at least 26).
 *
 * Lanes per reactor (set_lanes / get_lanes): 1, one thread per reactor and 64
 * reactors per wavefront, for throughput over many thousands of reactors; or
 * 16, a team per reactor that evaluates the Jacobian's columns on separate
 * lanes and eliminates one matrix row per lane, for latency when reactors are
 * few (the default for 9-species mechanisms; other sizes use 1).
 *
 * A pool holds `capacity` slots:
 *   - set starts a reactor in a slot;
 *   - advance launches up to max_tries step attempts for every running slot
 *     (asynchronous, on the pool's own stream); query sets *idle to 1 once
 *     the pool's launches finished, sync waits for them;
 *   - status reads all slots: 0 free, 1 running, 2 done, 3 failed;
 *   - get reads one slot's t, (T, Y) and step counts;
 *   - release frees a slot.
 * A reactor's result does not depend on the other slots or on how its run is
 * split into launches.
 * ------------------------------------------------------------------------ */
typedef struct astr_reactor_pool astr_reactor_pool;
int astr_hip_reactor_create(astr_chem *chem, int capacity, double rtol, double atol,
                            astr_reactor_pool **out);
int astr_hip_reactor_destroy(astr_reactor_pool *pool);
int astr_hip_reactor_set(astr_reactor_pool *pool, int slot, double T0, double p, const double *Y0,
                         double t_end, double h0);
int astr_hip_reactor_advance(astr_reactor_pool *pool, int max_tries);
int astr_hip_reactor_set_lanes(astr_reactor_pool *pool, int lanes);
int astr_hip_reactor_get_lanes(astr_reactor_pool *pool, int *lanes);
int astr_hip_reactor_query(astr_reactor_pool *pool, int *idle);
int astr_hip_reactor_sync(astr_reactor_pool *pool);
int astr_hip_reactor_status(astr_reactor_pool *pool, int *status);
int astr_hip_reactor_get(astr_reactor_pool *pool, int slot, double *t, double *TY, int *nstep, int *nrej);
int astr_hip_reactor_release(astr_reactor_pool *pool, int slot);
int astr_hip_reactor_method(double *coef, int n);
/* The reactor's right-hand side d(T, Y)/dt at n states (p[n], y and f n x
 * (ns+1) row-major) and, if jac is not null, the analytic Jacobian the
 * integrator uses (n x (ns+1) x (ns+1), row k = component, column j =
 * variable); 9-species mechanisms. */
int astr_hip_reactor_rhs(astr_chem *chem, int n, const double *p, const double *y, double *f, double *jac);

/* ------------------------------------------------------------------------
 * HBM streaming probe (no reference counterpart): dst[0:n] = src[0:n] in
 * FP64 with width doubles per lane per access (1, or 2 and 4 as 16-byte
 * loads; even n and 16-byte aligned buffers), nontemporal, stream-ordered.
 * It reads and writes exactly 8n bytes each: the achievable-bandwidth
 * reference of the GPU and the calibration of the FETCH_SIZE / WRITE_SIZE
 * counters.  Errors: astr_hip_probe_last_error().
 * ------------------------------------------------------------------------ */
const char *astr_hip_probe_last_error(void);
int astr_hip_stream_probe(const double *src, double *dst, long long n, int width, void *stream);

#ifdef __cplusplus
}
#endif
#endif
