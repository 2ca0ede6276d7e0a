/*
 * iwgpu.h -- C-ABI of the MI355X (gfx950) implementation of one time step of
 * the stratified-Boussinesq solver of jmsegret/IWrepo (mode1_bbl): Fourier
 * pseudo-spectral in x,y, Legendre Gauss-Lobatto multidomain penalty in z.
 *
 * Plain C types only.  Every array argument is a host array in the reference's
 * own layout a(nxpp,ny,nz) (Fortran column order: x fastest), nxpp = nx+2,
 * nz = nsubd*nzloc; in spectral space entries 2ii-1/2ii (Fortran, 1-based) are
 * the real/imaginary parts of mode (kx=(ii-1)*2pi/xlen, ky=yw(j)).
 *
 * All entry points return 0 on success and a negative code on failure;
 * iw_last_error() then describes the failure.  A solver handle is not
 * thread-safe (the reference is single-threaded, comscratch.h).
 */
#ifndef IWGPU_H
#define IWGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define IW_MAXSUB 64

/* Run parameters: dim.h:12-28 sizes plus the &inputs/&params/&mflags
 * namelists of main.f:185-194 that the step reads. */
typedef struct iw_params {
    int nx, ny, nzloc, nsubd;      /* dim.h:12 (nzloc = GLL order + 1)            */
    int bc_uv_bottom;              /* 0 free slip (calcuvwt.f:291-295),
                                      1 no slip (calcuvwt_noslip.f:294-299)       */
    int ispecfilternl;             /* main.f mflags: filter after NL step          */
    int ispecfilterexp;            /* filter penalty solutions ...                 */
    int nspecfilterexp;            /* ... every nspecfilterexp steps               */
    int navgintrfc;                /* interface averaging period (main.f:920)      */
    int ifilter, nstepfilter;      /* estimation-model filter every nstepfilter
                                      steps (main.f:880-895)                      */
    int npostp;                    /* ... with the SGS tensors every npostp steps  */
    int imaxcalc;                  /* 1: min/max/CFL reductions of maxcalc every
                                      step (main.f:47, various.f:749-851)          */
    int dtadapt;                   /* adaptive dt of various.f:911-985: 0 off
                                      (fixed dt), 1 with Tso kept from the last
                                      change, 2 with Tso read as 0 at every call  */
    int n3doutput;                 /* every n3doutput steps the 3-D output routine
                                      plotf takes u,v,w,temp through a horfft
                                      round trip (main.f:996-1039); 0: never      */
    int ncalclegspec;              /* every ncalclegspec steps speclegencalc
                                      takes the Legendre spectra of u,v,w,temp
                                      at j = ny/2 (main.f:905-911), a round
                                      trip of each; 0: never                      */
    double xlen, ylen, zlen;
    double xnu, xkappa, brunt, grav, rho0, omz;
    double dt;
    double amp, p, pf, fackc;      /* exponential filters (filtering.f:1225-1268)  */
    double facdir, facneumn, facrobin; /* penalty factors (penalty.f:57-157)      */
    double zcen, xbcrit;           /* mean-u damping (calcuvwt.f:343-361)          */
    double xacrit, umax, fwave;    /* virtual paddle (forcing.f:325-447)           */
    double fackcbv;                /* Boyd-Vandeven cut-off fraction (filtering.f:1040) */
    double fcshear, uo;            /* shear current uo*fcshear*fshear(z) (main.f:120,
                                      352; forcing.f:190-214): 0 and 0.07 there   */
    double z0[IW_MAXSUB], zh[IW_MAXSUB]; /* subdomains (subd.dat, setup.f:316-325) */
} iw_params;

typedef struct iw_solver iw_solver;

int iw_params_size(void);
const char *iw_last_error(void);

/* Set-up of one solver instance on the current HIP device: GLL nodes,
 * weights, Costa-Don matrices, wavenumbers and grid (replaces setup,
 * setup.f:6-267, and fftfax/cftfax, main.f:210-218) and the device mirror
 * of the COMMON blocks /flow/ and /matrix/ (comflow.h, commatrix.h).
 * nx and ny (ny > 1) even, of the form 2^a 3^b 5^c (the lengths FFTFAX and
 * CFTFAX factor, fft.f:407-500, 1503-1622), nx in [4, 1024], ny <= 1024;
 * ny = 1 is the 2-D case. */
int iw_create(const iw_params *p, iw_solver **out);
int iw_destroy(iw_solver *s);

/* Device mirror of the 24 state arrays: 0..3 u,v,w,temp; 4..7 un..tn;
 * 8..11 unm..tnm; 12..15 su..st; 16..19 sun..stn; 20..23 sunm..stnm. */
int iw_set_state(iw_solver *s, int idx, const double *host);
int iw_get_state(iw_solver *s, int idx, double *host);
/* /pressbc/ wnl, wviscrot (comrotpressbc.h), each (nxpp,ny,2) */
int iw_get_pressbc(iw_solver *s, double *wnl, double *wviscrot);
int iw_set_pressbc(iw_solver *s, const double *wnl, const double *wviscrot);
/* dt of /timevars/ (comtime.h:1); the u,v,w,rho operators are refactored
 * at the next calcuvwt when dt or bdfv has changed */
int iw_set_dt(iw_solver *s, double dt);
/* Step clock and the SS/BDF coefficients of /ab/, /bdf/ (comtimeadvance.h) */
int iw_set_coeffs(iw_solver *s, double t, const double fac[3], const double bdf[3], double bdfv);
int iw_get_coeffs(iw_solver *s, double *t, double fac[3], double bdf[3], double *bdfv);

/* Start from rest-frame physical fields (istart=0): forward transform of the
 * four physical fields and the start coefficients of main.f:543-552. */
int iw_set_initial(iw_solver *s, const double *u, const double *v,
                   const double *w, const double *temp);

/* Device-resident step loop replaying main.f:653-1066 for nsteps steps:
 * maxcalc (its round trip; with imaxcalc its records, see iw_maxcalc; with
 * dtadapt the adaptive step), convec, SS/BDF coefficient selection with the
 * variable-step branches after a restart or a change of dt, calcuvwt, the
 * estimation filter, interface_avg3d.  counts (may be NULL) receives 4
 * interface-average counts per step (icu,icv,icw,ict). */
int iw_run(iw_solver *s, int nsteps, int *counts);
double iw_time(iw_solver *s);
int iw_sync(iw_solver *s);

/* various.f:665  subroutine maxcalc(istep), called by main.f:657 every step:
 * on the device state, the physical-space extrema and CFL maxima in the
 * reference's loop order (a later point wins a tie), the horfft round trip
 * of u,v,w,temp, and, with dtadapt, the adaptive step of various.f:887-985
 * (new dt, cdt, variable-step coefficients, ivardt = 1, idtstep = 0).
 * rec (may be NULL) receives IW_MAXCALC_NREC doubles, indices 1-based as the
 * reference prints them (0 = none): 0 istep, 1 t, 2 umin, 3 umaxt, 4 kumin,
 * 5 kumax, 6 vmin, 7 vmax, 8 kvmin, 9 kvmax, 10 wmin, 11 wmax, 12 kwmin,
 * 13 kwmax, 14 tmin, 15 tmax, 16 ktmin, 17 ktmax, 18 itmin, 19 jtmin,
 * 20 itmax, 21 jtmax, 22 cflxmax, 23-25 its i,j,k, 26 cflymax, 27-29 i,j,k,
 * 30 cflzmax, 31-33 i,j,k, 34 dz at the CFLZ maximum, 35 dt after the
 * decision, 36 cdt, 37 1 if dt changed, 38 ivardt, 39 idtstep,
 * 40 u(32,1,10), 41 u(32,1,50) (physical, the uveltest.dat probe; NaN when
 * that position lies outside u). */
#define IW_MAXCALC_NREC 42
int iw_maxcalc(iw_solver *s, int istep, double *rec);
/* The maxcalc records of the steps of the last iw_run (imaxcalc or dtadapt
 * set): n records of IW_MAXCALC_NREC doubles. */
int iw_maxcalc_records(iw_solver *s, double *out, int n);
/* dt, cdt of /timevars/ and ivardt, idtstep of /vardtflags/ (comtime.h:1-3),
 * which main.f:660-664, 838-854 change between maxcalc calls */
int iw_get_dtstate(iw_solver *s, double *dt, double *cdt, int *ivardt, int *idtstep);
int iw_set_dtstate(iw_solver *s, double cdt, int ivardt, int idtstep);

/* /sol2dvar/ (comsoliton.h:9-13): the nine (nxpp,nz) mean-wave arrays in
 * COMMON order, umwv, wmwv, dumdx, dumdz, dwmdx, dwmdz, rhomwv, drhomdx,
 * drhomdz (9*nxpp*nz doubles).  They enter rhsfm (solver.f:848-931), the
 * interface upwind penalty (solver.f:723-724) and maxcalc (various.f:757-808);
 * zero unless set here or read from a restart file (restart.f:65-69). */
int iw_set_meanwave(iw_solver *s, const double *mw);
int iw_get_meanwave(iw_solver *s, double *mw);

/* Diagnostics.  Replaces postp(initenergy,t,brunt,tstart) (postp.f:9-76):
 * the TKE budget of energycalc_postp (postp.f:415-732) of the device state,
 * out = the sixteen values of the energy.dat line (postp.f:53-58) in its
 * order: tketot, u1tke, u2tke, u3tke, prody, prodz, prod, bflux, epsilon,
 * epssgs, epstot, chi, chisgs, chitot, tkemean, tempvar (out may be NULL).
 * The state changes as postp's horfft round trips change it.  iw_run calls
 * it every npostp steps (main.f:935-936) and iw_load_restart once
 * (main.f:635). */
int iw_postp(iw_solver *s, double *out);
/* Replaces speclegencalc(filein,t,fin) (filtering.f:1110-1221) for field idx
 * (0 u, 1 v, 2 w, 3 temp): out[kloc*nsubd + ks] = the x-average of the
 * |Legendre coefficient kloc| of subdomain ks at j = ny/2, the numbers of
 * the speclegen<c><t>.dat file; the field takes the horfft round trip.
 * iw_run calls it for u,v,w,temp every ncalclegspec steps (main.f:905-911). */
int iw_speclegen(iw_solver *s, int idx, double *out);
/* Kinetic and available potential energy of the state, from the offline
 * tool CalcE (CalcE/integration.f:14-226, KEh/KEv/PEh over 2-D dumps) taken
 * to 3-D: out = KE_u, KE_v, KE_w, PE with KE_q = int rho0/2 q^2 dV and
 * PE = int g^2 rho'^2/(2 rho0 N^2) dV; horizontal mean times xlen (times
 * ylen when ny > 1; per unit span in 2-D as the tool), the z integral by
 * integrcol (derv.f:479-512) per subdomain, PE without the tool's double
 * count of i = 1 (integration.f:221).  The state is not changed. */
int iw_kepe(iw_solver *s, double *out);
/* iw_kepe without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The per-level partial
 * sums live in a static device array of 8192 doubles per GPU (nz up to
 * 2047); larger nz is rejected.  Concurrent calls on one GPU share that
 * array and race.  Same result as iw_kepe. */
int iw_kepe_noalloc(iw_solver *s, double *out);
/* The records of the last iw_run (or of iw_load_restart): postp records of
 * 18 doubles (istep, t, the sixteen values) and speclegencalc records of
 * 2 + 4*nz doubles (istep, t, the u, v, w, temp spectra).  Copies n records
 * and returns how many there are (n = 0: just the count), or -1. */
int iw_energy_records(iw_solver *s, double *out, int n);
int iw_speclegen_records(iw_solver *s, double *out, int n);

/* Threads per block of the horizontal transforms (64, 128 or 256, the
 * default): every block holds 8 values per thread, i.e. 8 ft / (nx/2) rows
 * of the x transform or 8 ft / ny columns of the y transform; smaller blocks
 * give small grids more blocks.  The results do not depend on it (each
 * sequence takes the same passes).  A tuning knob with no counterpart in the
 * reference (iw_ensemble_tune sets it per member). */
int iw_set_fft_threads(iw_solver *s, int ft);

/* Mixed precision (no counterpart in the reference, whose arithmetic is FP64
 * throughout).  on = 1: the physical-space part of convec (solver.f:6-514,
 * 523-1206) -- the inverse and forward horizontal transforms, the z
 * derivatives of the fluxes and of u, v, w (on FP32 matrix cores) and the
 * nonlinear products -- keeps its working set in FP32; the state, the SS/BDF
 * assembly, the filters, the penalty solves and the projection stay FP64.
 * maxcalc's extrema come from the FP64 values of the inverse transform, its
 * umaxt chain from FP32 u.  3-D runs without mean-wave terms; GLL orders
 * 8 to 24.  on = 0: the plain FP64 path (the default). */
int iw_set_mixed(iw_solver *s, int on);

/* ---- learned sub-grid closure (no counterpart in the reference, whose
 * sub-grid model is the estimation filter with sgscalc's tensors,
 * filtering.f:8-93, sgscalc.f:125-219) ----
 * A per-point network 12 -> 64 -> 64 -> 64 -> 9 (tanh; BF16 matrix-core
 * forward with FP32 accumulation, FP32 backward, Adam) predicts the SGS
 * stresses t11 t12 t13 t22 t23 t33 and buoyancy fluxes th1 th2 th3 of a
 * coarse grid from its filtered velocity gradients (dudx dudy dudz dvdx dvdy
 * dvdz dwdx dwdy dwdz), d rho/dz, |grad_h rho| and z.  Training samples come
 * from fine runs coarse-grained onto the coarse grid (sharp spectral cut-off
 * in x and y, Legendre L2 projection in z onto coarse subdomains made of
 * whole fine subdomains): tau_ij = f(u_i u_j) - f(u_i) f(u_j), theta_i =
 * f(u_i rho) - f(u_i) f(rho).  In a coarse run (iw_set_closure) the
 * predicted tensors enter rhsfm's flux form (solver.f:975-1063, 1155-1161). */
typedef struct iw_mlp iw_mlp;
/* parameters + normalisation floats: W1[64][16] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[16][64] b4[16] (unused pads zero), then the mean and
 * deviation of the 12 inputs and of the 9 outputs */
int iw_mlp_nparams(void);
int iw_mlp_create(unsigned long long seed, iw_mlp **out);
int iw_mlp_destroy(iw_mlp *m);
int iw_mlp_get_params(iw_mlp *m, float *out);
int iw_mlp_set_params(iw_mlp *m, const float *in);
/* host arrays: raw inputs x (n x 12) -> predicted tensors y (n x 9) */
int iw_mlp_forward(iw_mlp *m, const float *x, int n, float *y);
/* mean squared error of the normalised outputs for targets t (n x 9) and its
 * gradient with respect to the first nparams - 42 parameters */
int iw_mlp_gradient(iw_mlp *m, const float *x, const float *t, int n, float *grad, float *loss);
/* coarse-grain the fine instance's u,v,w,rho into the coarse instance's
 * state and (m not null) append the coarse grid's nx ny nz samples to m's
 * training set */
int iw_closure_sample(iw_mlp *m, iw_solver *fine, iw_solver *coarse);
/* size of the training set; copies the first n samples when n > 0 */
long long iw_closure_dataset(iw_mlp *m, float *x, float *t, long long n);
/* normalisation from the training set (unless set), then `steps` Adam steps
 * on minibatches of nb samples; loss_hist (may be null) per step */
int iw_mlp_train(iw_mlp *m, int steps, int nb, float lr, unsigned long long seed, float *loss_hist);
/* the network's tensors at the coarse instance's sample points (nx ny nz x 9) */
int iw_closure_predict(iw_mlp *m, iw_solver *coarse, float *tau);
/* use the network in the coarse instance's steps (m null: off; plain FP64
 * path, 3-D, no mean-wave terms) */
int iw_set_closure(iw_solver *coarse, iw_mlp *m);

/* ---- ensembles (no counterpart in the reference, which sizes one run at
 * compile time, dim.h:12-28) ----
 * Independent instances of uneven sizes (each from iw_create and
 * iw_set_initial, on one device) advanced together: after their start-up
 * steps, three steps of every member are captured once into one HIP graph
 * (members dealt to independent chains) and replayed, one launch per three
 * steps.  Members run without maxcalc records, adaptive dt, the estimation
 * filter, the output routines and the paddle; their state after
 * iw_ensemble_run equals that of iw_run on each member.  While a solver is a
 * member, iw_run keeps all of its launches on the solver's own stream (no
 * second stream for the rho solve); iw_ensemble_destroy restores it, so the
 * members must outlive the ensemble. */
typedef struct iw_ensemble iw_ensemble;
int iw_ensemble_create(iw_solver **members, int n, iw_ensemble **out);
int iw_ensemble_destroy(iw_ensemble *e);
int iw_ensemble_run(iw_ensemble *e, int nsteps);
/* Autotuner: per member the transform block size (iw_set_fft_threads), then
 * the number of graph chains (or no graph), each timed over `steps` steps on
 * the members' own state, which is restored.  out (may be null): chosen
 * chains (0: no graph), ms per ensemble step chosen, ms with all chains and
 * 256 threads. */
int iw_ensemble_tune(iw_ensemble *e, int steps, double *out);
int iw_ensemble_config(iw_ensemble *e, int *lanes, int *ft);
int iw_ensemble_set_lanes(iw_ensemble *e, int lanes);

/* ---- reference routines on host arrays (parity / Fortran-shim mode) ---- */
/* horfft.f:2    subroutine horfft(b,is,nk): is=-1 physical->spectral */
int iw_horfft(iw_solver *s, double *b, int is, int nk);
/* derv.f:212    subroutine dpdz(a,b) */
int iw_dpdz(iw_solver *s, const double *a, double *b);
/* penalty.f:9, 162-235  subroutine penalty_setup(epsilon) and the boundary
 * part of penalty_main: the /penaltyparams/ alpha, beta, gamma, delta, tau1,
 * tau2 of every subdomain (nsubd values each) for (epsilon, ifbcbot,
 * ifbctop); boundary types 0 Dirichlet, 1 Neumann, 2 Robin (parambctype.h) */
int iw_penalty_params(iw_solver *s, double eps, int ifbcbot, int ifbctop, double *alpha, double *beta,
                      double *gamma, double *delta, double *tau1, double *tau2);
/* penalty.f:57   subroutine param_set(ifbc,alpha0,beta0,tau0,epsilon,ybar,ks) */
int iw_param_set(iw_solver *s, int ifbc, double eps, double ybar, double *alpha0, double *beta0, double *tau0);
/* penalty.f:255  subroutine helmv_penalty(fvr,fvi,epsilon,rhsinr,rhsini,
 * bcrtop,bcitop,bcrbot,bcibot,ifbctop,ifbcbot,ifuncreal) for one column pair
 * (with the boundary-row terms of matrix_elmsetup, penalty.f:583-604, and the
 * banded LU of bandlu_solve, band_lu.f:15-52, on the device); rhsr, rhsi as
 * penalty_main hands them over (already times -epsilon), the penalty
 * parameters those of iw_penalty_params; ifbcbot, ifbctop 0 or 1 */
int iw_penalty_solve(iw_solver *s, double eps, int ifbcbot, int ifbctop, double bcrbot, double bcibot,
                     double bcrtop, double bcitop, const double *rhsr, const double *rhsi, double *fr, double *fi);
/* filtering.f:811  subroutine filter_3dlegenfourier(fin3d,ifourier,ilegen,iout) */
int iw_filter_3dlegenfourier(iw_solver *s, double *f, int ifourier, int ilegen);
/* penalty.f:714  subroutine interface_avg3d(fin,ic) */
int iw_interface_avg3d(iw_solver *s, double *f, int *ic);
/* solver.f:6    subroutine convec(istep,istart,ispecfilternl), on the device
 * mirror of /flow/,/matrix/,/pressbc/ (upload with iw_set_state first) */
int iw_convec(iw_solver *s, int istep, int ispecfilternl);
/* calcuvwt.f:2  subroutine calcuvwt(izmode,ispecfilter,irelax,cpenspecfilter) */
int iw_calcuvwt(iw_solver *s, int cpenspecfilter);

/* filtering.f:8  subroutine filtering(isgscalc): the estimation-model
 * filter of u,v,w (modecut with the Boyd-Vandeven vertical filter) on the
 * device mirror; isgscalc = 1 also forms the SGS tensors (sgscalc.f:4-223) */
int iw_filtering(iw_solver *s, int isgscalc);
/* /comsgs/ (comsgs.h) tensor idx in physical space: 0 t11, 1 t12, 2 t13,
 * 3 t22, 4 t23, 5 t33, 6 th1, 7 th2, 8 th3 */
int iw_get_sgs(iw_solver *s, int idx, double *host);
/* The same tensors from the host, for callers whose /comsgs/ was written
 * elsewhere (postp reads them, postp.f:622-665). */
int iw_set_sgs(iw_solver *s, int idx, const double *host);

/* ---- restart files in the record layout of output.f:27-104 and
 * restart.f:26-115 (Fortran sequential unformatted, REAL*8, INTEGER*4) ---- */
/* Device state to a file the reference can restart from (main.f:986-990). */
int iw_save_restart(iw_solver *s, const char *path);
/* Restart from a reference start.dat (main.f:563-634): the 20 saved fields,
 * t and umax, then the variable-step first-step coefficients with cdtin =
 * cdt; the next two steps of iw_run use the restart branches of
 * main.f:703-734 and 818-825. */
int iw_load_restart(iw_solver *s, const char *path, double cdt);
/* The same file format on host arrays, no device needed: state holds the 24
 * arrays in iw_set_state order (su..st are not stored and read as zero). */
int iw_write_restart_file(const char *path, const iw_params *p, double t, const double *state);
int iw_read_restart_file(const char *path, const iw_params *p, double *t, double *umax, double *state);

/* multistep.f:7-89 variable-step AB3/BDF3 coefficient functions */
double iw_abvar1(double dt1, double dt2, double dt3);
double iw_abvar2(double dt1, double dt2, double dt3);
double iw_abvar3(double dt1, double dt2, double dt3);
double iw_bdfvarv(double dt1, double dt2, double dt3);
double iw_bdfvar1(double dt1, double dt2, double dt3);
double iw_bdfvar2(double dt1, double dt2, double dt3);
double iw_bdfvar3(double dt1, double dt2, double dt3);

/* ---- measurement ---- */
/* Per kernel class recorded while profiling is on: average device time (ms,
 * HIP events on the solver's stream) and average algorithmic bytes per
 * launch; names are written as a ';'-separated list into names (capacity
 * cap bytes).  Returns the number of classes. */
int iw_set_profiling(iw_solver *s, int on);
int iw_kernel_stats(iw_solver *s, char *names, int cap, double *avg_ms, int *launches,
                    double *avg_bytes, int maxk);
/* Bytes of device memory held by the instance. */
long long iw_device_bytes(iw_solver *s);

#ifdef __cplusplus
}
#endif
#endif
