/*
 * iw_oracle.h -- CPU FP64 restatement of one time step of the reference
 * stratified-Boussinesq solver (jmsegret/IWrepo, mode1_bbl).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product never links it.
 *
 * Every function cites the reference file:line it restates.  The parameter
 * block has the same layout as iw_params in include/iwgpu.h so one ctypes
 * structure feeds both libraries (checked by the tests through
 * iwo_params_size / iw_params_size).
 */
#ifndef IW_ORACLE_H
#define IW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define IWO_MAXSUB 64

typedef struct iwo_params {
    int nx, ny, nzloc, nsubd;
    int bc_uv_bottom;   /* 0 Neumann (calcuvwt.f:291-295), 1 Dirichlet (calcuvwt_noslip.f:294-299) */
    int ispecfilternl, ispecfilterexp, nspecfilterexp, navgintrfc;
    int ifilter, nstepfilter, npostp; /* estimation filter, main.f:880-895 */
    int imaxcalc, dtadapt;            /* maxcalc reductions, adaptive dt (various.f:665-987) */
    int n3doutput;                    /* plotf round trip every n3doutput steps (main.f:996-1039) */
    int ncalclegspec;                 /* speclegencalc every ncalclegspec steps (main.f:905-911) */
    double xlen, ylen, zlen;
    double xnu, xkappa, brunt, grav, rho0, omz;
    double dt;
    double amp, p, pf, fackc;
    double facdir, facneumn, facrobin;
    double zcen, xbcrit;
    double xacrit, umax, fwave;
    double fackcbv;
    double fcshear, uo; /* shear current, main.f:120, 352 */
    double z0[IWO_MAXSUB], zh[IWO_MAXSUB];
} iwo_params;

typedef struct iwo_state iwo_state;

int iwo_params_size(void);
iwo_state *iwo_create(const iwo_params *p);
void iwo_destroy(iwo_state *s);

/* 24 state arrays (nxpp,ny,nz): 0..3 u,v,w,temp; 4..7 un..tn; 8..11 unm..tnm;
   12..15 su..st; 16..19 sun..stn; 20..23 sunm..stnm. */
double *iwo_field(iwo_state *s, int idx);
/* setup tables: 0 zpts, 1 wg, 2 d (row-major d[k1*nzloc+k2]), 3 d2, 4 alpol
   (alpol[j*nzloc+k] = P_j(x_k)), 5 xw, 6 yw, 7 z, 8 x */
double *iwo_table(iwo_state *s, int idx);
double iwo_time(iwo_state *s);

void iwo_horfft(iwo_state *s, double *b, int is, int nk);
void iwo_dpdz(iwo_state *s, const double *a, double *b);
void iwo_filter_3dlegenfourier(iwo_state *s, double *f, int ifourier, int ilegen);
int  iwo_interface_avg3d(iwo_state *s, double *f);
/* one column penalty Helmholtz solve (penalty.f:162-251 with penalty_setup);
   rhsr/rhsi are overwritten (scaled by -epsilon) as in the reference. */
void iwo_penalty_solve(iwo_state *s, double epsilon, int ifbcbot, int ifbctop,
                       double bcrbot, double bcibot, double bcrtop, double bcitop,
                       double *rhsr, double *rhsi, double *fr, double *fi);
/* load physical u,v,w,temp (nxpp,ny,nz each) and transform them (istart=0) */
void iwo_set_initial(iwo_state *s, const double *u, const double *v,
                     const double *w, const double *t);
/* advance nsteps; counts (4 per step, may be NULL) receive interface counts */
void iwo_run(iwo_state *s, int nsteps, int *counts);
/* restart from 24 state arrays at time t with cdtin = cdt (main.f:563-634) */
void iwo_set_restart(iwo_state *s, const double *state, double t, double cdt);
int  iwo_nthreads(void);
void iwo_set_threads(int n);
/* estimation-model filter of u,v,w (filtering.f:8-93) on the spectral
   state; isgscalc also computes the SGS tensors (sgscalc.f:4-223) */
void iwo_filtering(iwo_state *s, int isgscalc);
/* SGS tensors in physical space: 0 t11, 1 t12, 2 t13, 3 t22, 4 t23,
   5 t33, 6 th1, 7 th2, 8 th3 (comsgs.h) */
double *iwo_sgs(iwo_state *s, int idx);

/* postp (postp.f:9-76) on the spectral state: the TKE budget of
   energycalc_postp in the energy.dat order of postp.f:53-58 (out, 16
   values: tketot, u1tke, u2tke, u3tke, prody, prodz, prod, bflux, epsilon,
   epssgs, epstot, chi, chisgs, chitot, tkemean, tempvar) and the state
   change of its horfft round trips */
void iwo_postp(iwo_state *s, double *out);
/* postp calls of the last iwo_run (every npostp steps, main.f:935), or
   the one of iwo_set_restart (main.f:635, istep 0): n records of IWO_NEREC
   doubles, istep, t and the 16 values above */
#define IWO_NEREC 18
double *iwo_energy_records(iwo_state *s, int *n);
/* Kinetic and available potential energy of the state (CalcE/
   integration.f:14-226 over the x-z dumps, generalised to 3-D): out =
   KE_u, KE_v, KE_w, PE with KE_q = int rho0/2 q^2 dV and PE = int g^2
   rho'^2 / (2 rho0 N^2) dV.  Horizontal: mean over i < nx (and j) times xlen
   (times ylen when ny > 1; per unit span in 2-D, as the 2-D tool does);
   vertical: GLL quadrature per subdomain, integrcol (derv.f:479-512).  PE
   without the tool's double count of i = 1 (integration.f:221).  The state
   is not changed. */
void iwo_kepe(iwo_state *s, double *out);
/* speclegencalc (filtering.f:1110-1221) of spectral field f (in place: its
   horfft round trip): the x-averaged |Legendre coefficients| at j = ny/2,
   out[kloc * nsubd + ks] (the rows of speclegen<c><t>.dat) */
void iwo_speclegen(iwo_state *s, double *f, double *out);
/* speclegencalc calls of the last iwo_run (every ncalclegspec steps, for
   u,v,w,temp, main.f:905-911): n records of 2 + 4 nzloc nsubd doubles,
   istep, t and the four spectra */
double *iwo_speclegen_records(iwo_state *s, int *n);
/* the nine (nxpp,nz) arrays of /sol2dvar/ (comsoliton.h:9-13) in COMMON
   order: umwv, wmwv, dumdx, dumdz, dwmdx, dwmdz, rhomwv, drhomdx, drhomdz */
void iwo_set_meanwave(iwo_state *s, const double *mw);
/* maxcalc (various.f:665-987) on the spectral state: the round trip, the
   extrema/CFL record and, with dtadapt, the adaptive step.  Record layout
   (IWO_NREC doubles; indices 1-based as the reference prints them, 0 =
   none): 0 istep, 1 t, 2 umin, 3 umaxt, 4 kumin, 5 kumax, 6 vmin, 7 vmax,
   8 kvmin, 9 kvmax, 10 wmin, 11 wmax, 12 kwmin, 13 kwmax, 14 tmin, 15 tmax,
   16 ktmin, 17 ktmax, 18 itmin, 19 jtmin, 20 itmax, 21 jtmax, 22 cflxmax,
   23-25 i,j,k, 26 cflymax, 27-29 i,j,k, 30 cflzmax, 31-33 i,j,k, 34 dzmax,
   35 dt (after the decision), 36 cdt, 37 dt changed, 38 ivardt, 39 idtstep,
   40 u(32,1,10), 41 u(32,1,50) (physical, uveltest.dat). */
#define IWO_NREC 42
void iwo_maxcalc(iwo_state *s, int istep, double *rec);
/* records of the steps of the last iwo_run (n of them) */
double *iwo_maxcalc_records(iwo_state *s, int *n);
double iwo_abvar1(double, double, double);
double iwo_abvar2(double, double, double);
double iwo_abvar3(double, double, double);
double iwo_bdfvarv(double, double, double);
double iwo_bdfvar1(double, double, double);
double iwo_bdfvar2(double, double, double);
double iwo_bdfvar3(double, double, double);

#ifdef __cplusplus
}
#endif
#endif
