/*
 * iw_oracle.c -- CPU FP64 restatement of the reference time step.
 *
 * TEST INFRASTRUCTURE ONLY (see iw_oracle.h).  Reference: jmsegret/IWrepo,
 * directory mode1_bbl; file:line citations below are relative to it.  The
 * reference is Fortran compiled with ifort -autodouble (makefile:24-31), so
 * every REAL is FP64 here as well.
 *
 * Array convention: a Fortran array a(nxpp,ny,nz) is a[i + nxpp*(j + ny*k)]
 * with 0-based i,j,k.  In spectral space entries 2ii/2ii+1 of a row are the
 * real/imaginary parts of the Fourier mode (kx = ii*alpha, ky = yw[j])
 * (comflow.h:6-11, complexflow.h:1-9).
 */
#include "iw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { IDIR = 0, INEU = 1, IROB = 2 }; /* parambctype.h */

struct iwo_state {
    iwo_params P;
    int nx, ny, nz, nzloc, nzm, nsubd, nxpp, nxh, nxhp, nyh, ndims;
    long nxppy, ntot;
    double alpha, beta, pi2, zdim, xkcrit, xmcrit, xomega, twave;
    double *zpts, *wg, *d, *d2, *alpol;
    double *xw, *xsq, *yw, *ysq, *x, *z;
    double *F[24];
    double *g[12]; /* dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz dtdx dtdy dtdz */
    double *ss1, *ss2, *ox, *work[8];
    double *wnl, *wviscrot;
    double *ub, *vb, *wb, *phi, *dphidz;
    double fac1, fac2, fac3, bdf1, bdf2, bdf3, bdfv, t, dt, facvel, fact;
    int istep;
    int istart;  /* 1 after iwo_set_restart (main.f:563-634) */
    double cdt;
    /* FFT twiddle tables: cos/sin(2*pi*k/n) for n = nx and n = ny */
    double *twx_c, *twx_s, *twy_c, *twy_s;
    /* estimation filter: /comsgs/ t11,t12,t13,t22,t23,t33,th1,th2,th3
       (comsgs.h) and the /comfiltertmp/ work arrays (comfiltertmp.h) */
    double *sgs[9], *tmps[6];
    /* maxcalc (various.f:665-987): Tso, the /vardtflags/ ivardt, idtstep
       (comtime.h:3), itmin/jtmin and dz, which the reference keeps from
       the previous call, and the per-step records of the last iwo_run */
    /* /sol2dvar/ (comsoliton.h:9-13), each (nxpp,nz): 0 umwv, 1 wmwv,
       2 dumdx, 3 dumdz, 4 dwmdx, 5 dwmdz, 6 rhomwv, 7 drhomdx, 8 drhomdz;
       zero unless set (wavesetup is not called, setup.f:191) */
    double *mw[9];
    double Tso, dzlast;
    int ivardt, idtstep, itmin, jtmin;
    double *mcrec;
    int mcn, mccap;
    double *erec; /* postp records of the last run */
    int en;
    double *lrec; /* speclegencalc records of the last run */
    int ln;
};

#define U(s) ((s)->F[0])
#define V(s) ((s)->F[1])
#define W(s) ((s)->F[2])
#define T(s) ((s)->F[3])

int iwo_params_size(void) { return (int)sizeof(iwo_params); }
int iwo_nthreads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
/* threads of the following calls (the 1-core baseline timing) */
void iwo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* multistep.f:7-89 -- variable-step AB3/BDF3 coefficients             */
/* ------------------------------------------------------------------ */
double iwo_abvar1(double dt1, double dt2, double dt3)
{ /* multistep.f:10-15 */
    double dt123 = 2. * dt1 + 6. * dt2 + 3. * dt3, dt23 = dt2 + dt3;
    return 1. + dt1 * dt123 / (6. * dt2 * dt23);
}
double iwo_abvar2(double dt1, double dt2, double dt3)
{ /* multistep.f:19-25 */
    double dt123 = 2. * dt1 + 3. * dt2 + 3. * dt3;
    return -dt1 * dt123 / (6. * dt2 * dt3);
}
double iwo_abvar3(double dt1, double dt2, double dt3)
{ /* multistep.f:29-36 */
    double dt12 = 2. * dt1 + 3. * dt2, dt23 = dt2 + dt3;
    return dt1 * dt12 / (6. * dt3 * dt23);
}
double iwo_bdfvarv(double dt1, double dt2, double dt3)
{ /* multistep.f:39-48 */
    double dt12 = dt1 + dt2, dt123 = dt1 + dt2 + dt3;
    return 1. + dt1 / dt12 + dt1 / dt123;
}
double iwo_bdfvar1(double dt1, double dt2, double dt3)
{ /* multistep.f:50-62 */
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    double b = -1. - dt1 / dt2 - dt1 / dt123 - dt1 * dt1 / (dt123 * dt2)
               - dt1 * dt1 * dt12 / (dt123 * dt23 * dt2);
    return -b;
}
double iwo_bdfvar2(double dt1, double dt2, double dt3)
{ /* multistep.f:64-77 */
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    double b = dt1 * dt1 / (dt12 * dt2) + dt1 * dt1 / (dt2 * dt123)
               + dt1 * dt1 * dt12 / (dt2 * dt23 * dt123)
               + dt1 * dt1 * dt12 / (dt3 * dt23 * dt123);
    return -b;
}
double iwo_bdfvar3(double dt1, double dt2, double dt3)
{ /* multistep.f:79-89 */
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    return dt1 * dt1 * dt12 / (dt3 * dt23 * dt123);
}

/* ------------------------------------------------------------------ */
/* GLL setup: setup.f:334-361 (legen), 395-518 (jacobl/jacobf),        */
/* 527-600 (quad), 636-798 (derv)                                       */
/* ------------------------------------------------------------------ */
static void jacobf(int n, double alp, double bet, double rv, double x, double *poly,
                   double *pder, double *polym1, double *pderm1, double *polym2,
                   double *pderm2)
{ /* setup.f:486-518, Jacobi recurrence */
    double apb = alp + bet;
    double p = 1.0, pd = 0.0, plst = 0, pdlst = 0, psave = 0, pdsave = 0;
    *poly = 1.0;
    *pder = 0.0;
    if (n == 0) return;
    plst = p;
    pdlst = pd;
    p = rv * x;
    pd = rv;
    if (n == 1) {
        *poly = p; *pder = pd;
        return;
    }
    for (int k = 2; k <= n; k++) {
        double a1 = 2. * k * (k + apb) * (2. * k + apb - 2.);
        double a2 = (2. * k + apb - 1.) * (alp * alp - bet * bet);
        double b3 = (2. * k + apb - 2.);
        double a3 = b3 * (b3 + 1.) * (b3 + 2.);
        double a4 = 2. * (k + alp - 1.) * (k + bet - 1.) * (2. * k + apb);
        double pn = ((a2 + a3 * x) * p - a4 * plst) / a1;
        double pdn = ((a2 + a3 * x) * pd - a4 * pdlst + a3 * p) / a1;
        psave = plst;
        pdsave = pdlst;
        plst = p;
        p = pn;
        pdlst = pd;
        pd = pdn;
    }
    *poly = p; *pder = pd;
    *polym1 = plst; *pderm1 = pdlst;
    *polym2 = psave; *pderm2 = pdsave;
}

static void jacobl(int n, double *xcol)
{ /* setup.f:395-479: Gauss-Lobatto points (alpha=beta=0), Newton with
     deflation, tolerance 1e-12, at most 10 iterations per root */
    double alp = 0., bet = 0., rv = 1. + alp;
    int np = n + 1, kstop = 10;
    double eps = 1.0e-12;
    double *xjac = calloc((size_t)np + 2, sizeof(double));
    double pnp1p, pdnp1p, pnp, pdnp, pnm1p, pdnm1, pnp1m, pdnp1m, pnm, pdnm, pnm1m;
    jacobf(np, alp, bet, rv, 1.0, &pnp1p, &pdnp1p, &pnp, &pdnp, &pnm1p, &pdnm1);
    jacobf(np, alp, bet, rv, -1.0, &pnp1m, &pdnp1m, &pnm, &pdnm, &pnm1m, &pdnm1);
    double det = pnp * pnm1m - pnm * pnm1p;
    double rp = -pnp1p, rm = -pnp1m;
    double a = (rp * pnm1m - rm * pnm1p) / det;
    double b = (rm * pnp - rp * pnm) / det;
    xjac[1] = 1.0;
    int nh = (n + 1) / 2;
    double dth = 3.14159265 / (2 * n + 1); /* setup.f:429 keeps this pi */
    double cd = cos(2. * dth), sd = sin(2. * dth), cs = cos(dth), ss = sin(dth);
    for (int j = 2; j <= nh; j++) {
        double x = cs;
        for (int k = 1; k <= kstop; k++) {
            double pnp1, pdnp1, pn, pdn, pnm1, pdnm1b;
            jacobf(np, alp, bet, rv, x, &pnp1, &pdnp1, &pn, &pdn, &pnm1, &pdnm1b);
            double poly = pnp1 + a * pn + b * pnm1;
            double pder = pdnp1 + a * pdn + b * pdnm1b;
            double recsum = 0.0;
            for (int i = 1; i <= j - 1; i++) recsum += 1.0 / (x - xjac[i]);
            double delx = -poly / (pder - recsum * poly);
            x = x + delx;
            if (fabs(delx) < eps) break;
        }
        xjac[j] = x;
        double cssave = cs * cd - ss * sd;
        ss = cs * sd + ss * cd;
        cs = cssave;
    }
    xjac[np] = -1.0;
    int npp = n + 2;
    for (int i = 2; i <= nh; i++) xjac[npp - i] = -xjac[i];
    if (n % 2 == 0) xjac[nh + 1] = 0.0;
    for (int k = 0; k <= n; k++) xcol[k] = xjac[n - k + 1];
    free(xjac);
}

static void legen(double *al, double *alp, int n, double xc)
{ /* setup.f:334-361 */
    al[0] = 1.;
    al[1] = xc;
    alp[0] = 0;
    alp[1] = 1.;
    for (int k = 1; k <= n; k++) {
        int kp = k + 1, km = k - 1;
        al[kp] = (2. * k + 1.) * xc * al[k] / kp - k * al[km] / kp;
    }
    for (int k = 1; k <= n; k++) {
        int kp = k + 1, km = k - 1;
        alp[kp] = (2. * k + 1.) * (al[k] + xc * alp[k]) / kp - k * alp[km] / kp;
    }
}

static void setup_gll(iwo_state *s)
{
    int N = s->nzm, nl = s->nzloc;
    jacobl(N, s->zpts);
    double *al = calloc((size_t)N + 3, sizeof(double)), *alp = calloc((size_t)N + 3, sizeof(double));
    for (int k = 0; k <= N; k++) { /* quad, setup.f:541-556 */
        legen(al, alp, N, s->zpts[k]);
        s->wg[k] = 2. / (N * (N + 1) * al[N] * al[N] + 1.0e-30);
        for (int j = 0; j <= N; j++) s->alpol[j * nl + k] = al[j];
    }
    free(al);
    free(alp);
    /* Costa-Don differentiation matrices, setup.f:666-760 */
    double *c = calloc((size_t)nl, sizeof(double));
    for (int k = 0; k <= N; k++) {
        double prod = 1.;
        for (int l = 0; l <= N; l++)
            if (l != k) prod = prod * (s->zpts[k] - s->zpts[l]);
        c[k] = prod;
    }
    double *d = s->d, *d2 = s->d2;
    for (int k = 0; k <= N; k++)
        for (int j = 0; j <= N; j++)
            d[k * nl + j] = (k != j) ? c[k] / (c[j] * (s->zpts[k] - s->zpts[j])) : 0.;
    for (int k = 0; k <= N; k++) {
        double sum = 0.;
        for (int j = 0; j <= N; j++)
            if (k != j) sum += d[k * nl + j];
        d[k * nl + k] = -sum;
    }
    for (int k = 0; k <= N; k++)
        for (int j = 0; j <= N; j++)
            d2[k * nl + j] = (k != j) ? 2.0 * (d[k * nl + k] * d[k * nl + j]
                                               - d[k * nl + j] / (s->zpts[k] - s->zpts[j]))
                                      : 0.;
    for (int k = 0; k <= N; k++) {
        double sum = 0.;
        for (int j = 0; j <= N; j++)
            if (k != j) sum += d2[k * nl + j];
        d2[k * nl + k] = -sum;
    }
    free(c);
}

/* ------------------------------------------------------------------ */
/* creation: main.f:210-472 derived parameters, setup.f:77-166 grid     */
/* ------------------------------------------------------------------ */
static double *dalloc(long n) { return calloc((size_t)n, sizeof(double)); }

static void twiddles(int n, double pi2, double **c, double **sn)
{
    *c = dalloc(n);
    *sn = dalloc(n);
    for (int k = 0; k < n; k++) {
        (*c)[k] = cos(pi2 * k / n);
        (*sn)[k] = sin(pi2 * k / n);
    }
}

iwo_state *iwo_create(const iwo_params *p)
{
    iwo_state *s = calloc(1, sizeof(iwo_state));
    s->P = *p;
    s->nx = p->nx;
    s->ny = p->ny;
    s->nzloc = p->nzloc;
    s->nzm = p->nzloc - 1;
    s->nsubd = p->nsubd;
    s->nz = p->nzloc * p->nsubd;
    s->nxpp = p->nx + 2;
    s->nxh = p->nx / 2;
    s->nxhp = s->nxh + 1;
    s->nyh = p->ny / 2;
    s->ndims = (p->ny > 1) ? 3 : 2;
    s->nxppy = (long)s->nxpp * s->ny;
    s->ntot = s->nxppy * s->nz;
    s->pi2 = 8. * atan(1.0);
    s->alpha = s->pi2 / p->xlen;
    s->beta = s->pi2 / p->ylen;
    s->zdim = p->xlen; /* main.f:339 */
    s->xkcrit = s->pi2 / p->xlen;
    s->xmcrit = s->pi2 / (2. * p->zlen);
    s->xomega = (p->brunt * s->xkcrit) / sqrt(s->xkcrit * s->xkcrit + s->xmcrit * s->xmcrit);
    s->twave = s->pi2 / s->xomega;
    s->dt = p->dt;
    s->cdt = 1.; /* cdtin (input.dat) */
    for (int q = 0; q < 9; q++) s->mw[q] = (double *)calloc((size_t)s->nxpp * s->nz, sizeof(double));
    s->fact = 1. / (p->xkappa * s->dt);
    s->facvel = 1. / (p->xnu * s->dt);
    int nl = s->nzloc;
    s->zpts = dalloc(nl);
    s->wg = dalloc(nl);
    s->d = dalloc((long)nl * nl);
    s->d2 = dalloc((long)nl * nl);
    s->alpol = dalloc((long)nl * nl);
    setup_gll(s);
    s->xw = dalloc(s->nxhp);
    s->xsq = dalloc(s->nxhp);
    s->yw = dalloc(s->ny);
    s->ysq = dalloc(s->ny);
    for (int i = 0; i < s->nxhp; i++) { /* setup.f:77-80 */
        s->xw[i] = i * s->alpha;
        s->xsq[i] = s->xw[i] * s->xw[i];
    }
    for (int j = 1; j <= s->nyh; j++) { /* setup.f:83-89 */
        s->yw[j - 1] = (j - 1) * s->beta;
        s->yw[s->ny - j] = -j * s->beta;
    }
    for (int j = 0; j < s->ny; j++) s->ysq[j] = s->yw[j] * s->yw[j];
    s->x = dalloc(s->nx);
    double xl = s->pi2 / s->alpha;
    for (int i = 0; i < s->nx; i++) s->x[i] = i * xl / s->nx; /* setup.f:126-129 */
    s->z = dalloc(s->nz);
    for (int ks = 0; ks < s->nsubd; ks++) /* setup.f:151-160 */
        for (int kl = 0; kl < nl; kl++)
            s->z[ks * nl + kl] = p->z0[ks] + 0.5 * p->zh[ks] * (s->zpts[kl] + 1.0);
    for (int f = 0; f < 24; f++) s->F[f] = dalloc(s->ntot);
    for (int f = 0; f < 12; f++) s->g[f] = dalloc(s->ntot);
    s->ss1 = dalloc(s->ntot);
    s->ss2 = dalloc(s->ntot);
    s->ox = dalloc(s->ntot);
    for (int q = 0; q < 9; q++) s->sgs[q] = dalloc(s->ntot);
    for (int q = 0; q < 6; q++) s->tmps[q] = dalloc(s->ntot);
    for (int f = 0; f < 8; f++) s->work[f] = dalloc(s->ntot);
    s->wnl = dalloc(s->nxppy * 2);
    s->wviscrot = dalloc(s->nxppy * 2);
    s->ub = dalloc(s->ntot);
    s->vb = dalloc(s->ntot);
    s->wb = dalloc(s->ntot);
    s->phi = dalloc(s->ntot);
    s->dphidz = dalloc(s->ntot);
    twiddles(s->nx, s->pi2, &s->twx_c, &s->twx_s);
    twiddles(s->ny, s->pi2, &s->twy_c, &s->twy_s);
    /* main.f:543-552 start values */
    s->fac1 = 1.0;
    s->bdf1 = 1.;
    s->bdfv = 1.;
    return s;
}

void iwo_destroy(iwo_state *s)
{
    if (!s) return;
    free(s->zpts); free(s->wg); free(s->d); free(s->d2); free(s->alpol);
    free(s->xw); free(s->xsq); free(s->yw); free(s->ysq); free(s->x); free(s->z);
    for (int f = 0; f < 24; f++) free(s->F[f]);
    for (int f = 0; f < 12; f++) free(s->g[f]);
    free(s->ss1); free(s->ss2); free(s->ox);
    for (int q = 0; q < 9; q++) free(s->sgs[q]);
    for (int q = 0; q < 6; q++) free(s->tmps[q]);
    for (int f = 0; f < 8; f++) free(s->work[f]);
    free(s->wnl); free(s->wviscrot);
    free(s->mcrec);
    free(s->erec);
    free(s->lrec);
    for (int q = 0; q < 9; q++) free(s->mw[q]);
    free(s->ub); free(s->vb); free(s->wb); free(s->phi); free(s->dphidz);
    free(s->twx_c); free(s->twx_s); free(s->twy_c); free(s->twy_s);
    free(s);
}

double *iwo_field(iwo_state *s, int idx) { return (idx >= 0 && idx < 24) ? s->F[idx] : NULL; }
double iwo_time(iwo_state *s) { return s->t; }
double *iwo_table(iwo_state *s, int idx)
{
    switch (idx) {
    case 0: return s->zpts;
    case 1: return s->wg;
    case 2: return s->d;
    case 3: return s->d2;
    case 4: return s->alpol;
    case 5: return s->xw;
    case 6: return s->yw;
    case 7: return s->z;
    case 8: return s->x;
    default: return NULL;
    }
}

/* ------------------------------------------------------------------ */
/* Complex FFT (any correct DFT may stand in for Temperton's passes,    */
/* fft.f:501-837): iterative radix-2 for powers of two, plain DFT else. */
/* sign = -1: sum g_j exp(-2 pi i jk/n); +1: exp(+...). Unnormalised.   */
/* ------------------------------------------------------------------ */
static void cfft(double *re, double *im, int n, int sign, const double *tc,
                 const double *ts, double *wr, double *wi)
{
    if (n == 1) return;
    if ((n & (n - 1)) == 0) {
        /* bit reversal */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < half; k++) {
                    double c = tc[k * step], sn = sign * ts[k * step];
                    double xr = re[i + k + half], xi = im[i + k + half];
                    double tr = xr * c - xi * sn, ti = xr * sn + xi * c;
                    re[i + k + half] = re[i + k] - tr;
                    im[i + k + half] = im[i + k] - ti;
                    re[i + k] += tr;
                    im[i + k] += ti;
                }
        }
    } else {
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                int m = (int)(((long)j * k) % n);
                double c = tc[m], sn = sign * ts[m];
                sr += re[j] * c - im[j] * sn;
                si += re[j] * sn + im[j] * c;
            }
            wr[k] = sr;
            wi[k] = si;
        }
        memcpy(re, wr, sizeof(double) * n);
        memcpy(im, wi, sizeof(double) * n);
    }
}

/* horfft.f:2-69 with FFT991 (fft.f:1361-1502) and CFFT99 (fft.f:1-406)  */
void iwo_horfft(iwo_state *s, double *b, int is, int nk)
{
    int nx = s->nx, ny = s->ny, nxpp = s->nxpp, nxh = s->nxh;
    long nxppy = s->nxppy;
    double scaley = 1. / (double)ny;
#pragma omp parallel
    {
        int nmax = nx > ny ? nx : ny;
        double *re = dalloc(nmax), *im = dalloc(nmax), *wr = dalloc(nmax), *wi = dalloc(nmax);
#pragma omp for schedule(static)
        for (int kz = 0; kz < nk; kz++) {
            double *a = b + (long)kz * nxppy;
            if (is == -1) {
                /* FFT991 forward per row: a_k = (1/N) sum x cos, b_k = -(1/N) sum x sin */
                for (int j = 0; j < ny; j++) {
                    double *row = a + (long)j * nxpp;
                    for (int i = 0; i < nx; i++) { re[i] = row[i]; im[i] = 0.; }
                    cfft(re, im, nx, -1, s->twx_c, s->twx_s, wr, wi);
                    double sc = 1.0 / (double)nx;
                    for (int k = 0; k <= nxh; k++) {
                        row[2 * k] = sc * re[k];
                        row[2 * k + 1] = sc * im[k];
                    }
                    row[1] = 0.0;              /* b0 = 0, fft.f:1304-1305 */
                    row[2 * nxh + 1] = 0.0;    /* b_{N/2} = 0 */
                }
                if (s->ndims == 3) { /* CFFT99 over nxh columns, Nyquist skipped */
                    for (int ii = 0; ii < nxh; ii++) {
                        for (int j = 0; j < ny; j++) {
                            re[j] = a[(long)j * nxpp + 2 * ii];
                            im[j] = a[(long)j * nxpp + 2 * ii + 1];
                        }
                        cfft(re, im, ny, -1, s->twy_c, s->twy_s, wr, wi);
                        for (int j = 0; j < ny; j++) {
                            a[(long)j * nxpp + 2 * ii] = re[j];
                            a[(long)j * nxpp + 2 * ii + 1] = im[j];
                        }
                    }
                }
                for (long i = 0; i < nxppy; i++) a[i] = scaley * a[i]; /* horfft.f:28-30 */
            } else {
                if (s->ndims == 3) {
                    for (int ii = 0; ii < nxh; ii++) {
                        for (int j = 0; j < ny; j++) {
                            re[j] = a[(long)j * nxpp + 2 * ii];
                            im[j] = a[(long)j * nxpp + 2 * ii + 1];
                        }
                        cfft(re, im, ny, +1, s->twy_c, s->twy_s, wr, wi);
                        for (int j = 0; j < ny; j++) {
                            a[(long)j * nxpp + 2 * ii] = re[j];
                            a[(long)j * nxpp + 2 * ii + 1] = im[j];
                        }
                    }
                }
                /* FFT991 inverse: x_j = sum_k C_k e^{+ijk}, C_0=A_0, C_{N/2}=A_{N/2}
                   (imaginary parts of both ignored, fft.f:1218-1230), pads zeroed */
                for (int j = 0; j < ny; j++) {
                    double *row = a + (long)j * nxpp;
                    re[0] = row[0];
                    im[0] = 0.;
                    re[nxh] = row[2 * nxh];
                    im[nxh] = 0.;
                    for (int k = 1; k < nxh; k++) {
                        re[k] = row[2 * k];
                        im[k] = row[2 * k + 1];
                        re[nx - k] = row[2 * k];
                        im[nx - k] = -row[2 * k + 1];
                    }
                    cfft(re, im, nx, +1, s->twx_c, s->twx_s, wr, wi);
                    for (int i = 0; i < nx; i++) row[i] = re[i];
                    row[nx] = 0.0;
                    row[nx + 1] = 0.0; /* fft.f:1483-1490 */
                }
            }
        }
        free(re); free(im); free(wr); free(wi);
    }
}

/* derv.f:212-275: b = (2/zh) * D * a within each subdomain */
void iwo_dpdz(iwo_state *s, const double *a, double *b)
{
    int nl = s->nzloc;
    long nxppy = s->nxppy;
#pragma omp parallel for schedule(static)
    for (int kk1 = 0; kk1 < s->nz; kk1++) {
        int ks = kk1 / nl, k1 = kk1 % nl;
        double fac = 2. / s->P.zh[ks];
        double *out = b + (long)kk1 * nxppy;
        for (long ij = 0; ij < nxppy; ij++) {
            double sum = 0.0;
            for (int k2 = 0; k2 < nl; k2++)
                sum = sum + s->d[k1 * nl + k2] * a[(long)(ks * nl + k2) * nxppy + ij];
            out[ij] = sum * fac;
        }
    }
}

/* derv.f:279-329 restricted to one output point k0 (0-based) */
static double dpdzcol_pt(const iwo_state *s, const double *a, double ybar, int k0)
{
    int nl = s->nzloc;
    double sum = 0.0;
    for (int k2 = 0; k2 < nl; k2++) sum = sum + s->d[k0 * nl + k2] * a[k2];
    return sum * (2. / ybar); /* zderivmap, setup.f:802-814 */
}

/* ------------------------------------------------------------------ */
/* filters: filtering.f:1225-1268 (specfilter), 729-808, 811-983        */
/* ------------------------------------------------------------------ */
static double specfilter(int k, int kc, int n, double p, double amp)
{ /* filtering.f:1225-1243 */
    if (k >= kc) {
        double theta = (double)(k - kc) / (double)(n - kc);
        return exp(-amp * pow(theta, 2. * p));
    }
    return 1.;
}

static double specfilter_fourier2d(double ak, double akmax, int kc, double p, double amp)
{ /* filtering.f:1248-1268; the ak < akc branch never occurs for kc = 0 */
    double akc = akmax * (double)kc;
    double theta = (ak - akc) / (akmax - akc);
    return exp(-amp * pow(theta, 2. * p));
}

/* filtering.f:729-808 for one subdomain column of nzloc values */
static void filter_legenpen(const iwo_state *s, const double *fin, double *fout)
{
    int nl = s->nzloc, nzm = s->nzm;
    int kc = (int)(s->P.fackc * nzm);
    double a[IWO_MAXSUB + 64], af[IWO_MAXSUB + 64];
    for (int k = 0; k <= nzm; k++) {
        double sum = 0.;
        for (int k1 = 0; k1 < nl; k1++) sum = sum + s->wg[k1] * fin[k1] * s->alpol[k * nl + k1];
        a[k] = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
    }
    for (int k = 0; k <= nzm; k++) af[k] = a[k] * specfilter(k, kc, nzm, s->P.p, s->P.amp);
    for (int k1 = 0; k1 < nl; k1++) {
        double sum = 0.;
        for (int k = 0; k <= nzm; k++) sum = sum + af[k] * s->alpol[k * nl + k1];
        fout[k1] = sum;
    }
}

void iwo_filter_3dlegenfourier(iwo_state *s, double *f, int ifourier, int ilegen)
{ /* filtering.f:811-983 */
    int nl = s->nzloc, nzm = s->nzm, nxpp = s->nxpp;
    int kc = (int)(s->P.fackc * nzm);
    double akmax = sqrt(s->xsq[s->nxhp - 1] + s->ysq[s->nyh]);
    long nxppy = s->nxppy;
    double sig[128];
    for (int k = 0; k <= nzm; k++) sig[k] = specfilter(k, kc, nzm, s->P.p, s->P.amp);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < s->ny; j++) {
        double ar[128], ai[128];
        for (int ii = 0; ii < s->nxhp; ii++) {
            double ak = sqrt(s->xsq[ii] + s->ysq[j]);
            double cf = specfilter_fourier2d(ak, akmax, 0, s->P.pf, s->P.amp);
            for (int ks = 0; ks < s->nsubd; ks++) {
                if (ifourier) {
                    for (int kl = 0; kl < nl; kl++) {
                        long o = (long)(ks * nl + kl) * nxppy + (long)j * nxpp + 2 * ii;
                        double fr = f[o], fi = f[o + 1];
                        double fmag = sqrt(fr * fr + fi * fi);
                        double cosc = fr / fmag, sinc = fi / fmag;
                        if (fmag != 0.) {
                            double phase = (sinc >= 0.) ? acos(cosc) : -acos(cosc);
                            double fmagf = fmag * cf;
                            fr = fmagf * cos(phase);
                            fi = fmagf * sin(phase);
                        } else {
                            fr = 0.;
                            fi = 0.;
                        }
                        f[o] = fr;
                        f[o + 1] = fi;
                    }
                }
                if (ilegen) {
                    for (int k = 0; k <= nzm; k++) {
                        double sumr = 0., sumi = 0.;
                        for (int k1 = 0; k1 < nl; k1++) {
                            long o = (long)(ks * nl + k1) * nxppy + (long)j * nxpp + 2 * ii;
                            sumr = sumr + s->wg[k1] * f[o] * s->alpol[k * nl + k1];
                            sumi = sumi + s->wg[k1] * f[o + 1] * s->alpol[k * nl + k1];
                        }
                        double fk = (k == nzm) ? 0.5 * nzm : (k + 0.5);
                        ar[k] = fk * sumr * sig[k];
                        ai[k] = fk * sumi * sig[k];
                    }
                    for (int k1 = 0; k1 < nl; k1++) {
                        double sumr = 0., sumi = 0.;
                        for (int k = 0; k <= nzm; k++) {
                            sumr = sumr + ar[k] * s->alpol[k * nl + k1];
                            sumi = sumi + ai[k] * s->alpol[k * nl + k1];
                        }
                        long o = (long)(ks * nl + k1) * nxppy + (long)j * nxpp + 2 * ii;
                        f[o] = sumr;
                        f[o + 1] = sumi;
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* penalty method: penalty.f:9-251 (parameters), 255-634 (assembly),    */
/* band_lu.f:15-52 + SGBTF2/SGBTRS (bandludcmp.f:238-440, 883-1069)      */
/* ------------------------------------------------------------------ */
typedef struct {
    double alpha[IWO_MAXSUB], beta[IWO_MAXSUB], gamma[IWO_MAXSUB], delta[IWO_MAXSUB];
    double tau1[IWO_MAXSUB], tau2[IWO_MAXSUB];
    double omega;
} penparams;

static void param_set(const iwo_state *s, const penparams *pp, int ifbc, double *alpha0,
                      double *beta0, double *tau0, double eps, double ybar)
{ /* penalty.f:57-157 */
    double omega = pp->omega;
    if (ifbc == IDIR) {
        *alpha0 = 1.;
        *beta0 = 0.;
        double fac = s->P.facdir * ((2. / ybar) * (2. / ybar));
        *tau0 = fac * eps / (*alpha0 * (omega * omega));
    }
    if (ifbc == INEU) {
        *alpha0 = 0.;
        *beta0 = 1.;
        double fac = s->P.facneumn * 2. / ybar;
        *tau0 = fac / (*beta0 * omega);
    }
    if (ifbc == IROB) {
        *alpha0 = 1.;
        *beta0 = 1.;
        double ck = *alpha0 * omega / *beta0;
        double dn = 1. / (omega * eps * *beta0);
        double t1 = sqrt(ck * ck + eps * ck);
        double t2 = 2. * ck - 2. * t1;
        double facm = ((1. / eps) >= 1.e5) ? s->P.facrobin : 1.;
        double t3 = facm * dn * (eps + t2) * 2. / ybar;
        *tau0 = t3;
    }
}

static void penalty_params(const iwo_state *s, penparams *pp, double eps, int ifbcbot, int ifbctop)
{
    int ns = s->nsubd;
    double a0, b0, t0;
    pp->omega = 2. / ((s->nzloc + 1) * (s->nzloc + 2)); /* penalty.f:25 */
    for (int ks = 1; ks < ns - 1; ks++) { /* penalty_setup, penalty.f:32-48 */
        double ybar = s->P.zh[ks];
        param_set(s, pp, IROB, &a0, &b0, &t0, eps, ybar);
        pp->alpha[ks] = a0; pp->beta[ks] = b0; pp->tau1[ks] = t0;
        pp->gamma[ks] = a0; pp->delta[ks] = b0; pp->tau2[ks] = t0;
    }
    /* penalty_main, penalty.f:200-238 */
    param_set(s, pp, ifbcbot, &a0, &b0, &t0, eps, s->P.zh[0]);
    pp->alpha[0] = a0; pp->beta[0] = b0; pp->tau1[0] = t0;
    if (ns > 1) {
        param_set(s, pp, IROB, &a0, &b0, &t0, eps, s->P.zh[0]);
        pp->gamma[0] = a0; pp->delta[0] = b0; pp->tau2[0] = t0;
    }
    param_set(s, pp, ifbctop, &a0, &b0, &t0, eps, s->P.zh[ns - 1]);
    pp->gamma[ns - 1] = a0; pp->delta[ns - 1] = b0; pp->tau2[ns - 1] = t0;
    if (ns > 1) {
        param_set(s, pp, IROB, &a0, &b0, &t0, eps, s->P.zh[ns - 1]);
        pp->alpha[ns - 1] = a0; pp->beta[ns - 1] = b0; pp->tau1[ns - 1] = t0;
    }
}

/* matrix_elmsetup, penalty.f:403-634: dense nz x nz operator, row-major
   a[r*nz + c] for Fortran a(r+1,c+1) */
static void matrix_elmsetup(const iwo_state *s, const penparams *pp, double eps, double *a)
{
    int nz = s->nz, nl = s->nzloc, ns = s->nsubd, N = nl - 1;
    const double *d = s->d, *d2 = s->d2;
    memset(a, 0, sizeof(double) * (size_t)nz * nz); /* helmv_penalty, penalty.f:340-344 */
#define A_(r, c) a[(long)(r) * nz + (c)]
    for (int ks = 0; ks < ns; ks++) {
        double ybar = s->P.zh[ks];
        double ybarl = (ks > 0) ? s->P.zh[ks - 1] : 0., ybarr = (ks < ns - 1) ? s->P.zh[ks + 1] : 0.;
        int km = ks * nl;
        double c2 = (2. / ybar) * (2. / ybar);
        for (int k1 = 0; k1 < nl; k1++)
            for (int k2 = 0; k2 < nl; k2++) A_(km + k1, km + k2) = -eps * c2 * d2[k1 * nl + k2];
        for (int k1 = 0; k1 < nl; k1++) A_(km + k1, km + k1) = A_(km + k1, km + k1) + 1.;
        A_(km, km) = A_(km, km) + pp->tau1[ks] * pp->alpha[ks];
        double cf = -pp->tau1[ks] * pp->beta[ks] * eps * 2. / ybar;
        for (int k = 0; k < nl; k++) A_(km, km + k) = A_(km, km + k) + cf * d[0 * nl + k];
        A_(km + N, km + N) = A_(km + N, km + N) + pp->tau2[ks] * pp->gamma[ks];
        cf = +pp->tau2[ks] * pp->delta[ks] * eps * 2. / ybar;
        for (int k = 0; k < nl; k++) A_(km + N, km + k) = A_(km + N, km + k) + cf * d[N * nl + k];
        if (ns > 1 && ks != 0) { /* penalty.f:517-547 */
            A_(km, km) = A_(km, km) - 0.5 * pp->tau1[ks] * pp->alpha[ks];
            cf = +0.5 * pp->tau1[ks] * pp->beta[ks] * eps * 2. / ybar;
            for (int k = 0; k < nl; k++) A_(km, km + k) = A_(km, km + k) + cf * d[0 * nl + k];
            for (int k = 0; k < nl; k++) A_(km, km - nl + k) = 0.;
            A_(km, km - 1) = -0.5 * pp->tau1[ks] * pp->alpha[ks];
            cf = +0.5 * pp->tau1[ks] * pp->beta[ks] * eps * 2. / ybarl;
            for (int k = 0; k < nl; k++)
                A_(km, km - nl + k) = A_(km, km - nl + k) + cf * d[N * nl + k];
        }
        if (ns > 1 && ks != ns - 1) { /* penalty.f:555-578 */
            A_(km + N, km + N) = A_(km + N, km + N) - 0.5 * pp->tau2[ks] * pp->gamma[ks];
            cf = -0.5 * pp->tau2[ks] * pp->delta[ks] * eps * 2. / ybar;
            for (int k = 0; k < nl; k++) A_(km + N, km + k) = A_(km + N, km + k) + cf * d[N * nl + k];
            for (int k = 0; k < nl; k++) A_(km, km + nl + k) = 0.; /* row km as in penalty.f:565 */
            A_(km + N, km + nl) = A_(km + N, km + nl) - 0.5 * pp->tau2[ks] * pp->gamma[ks];
            cf = -0.5 * pp->tau2[ks] * pp->delta[ks] * eps * 2. / ybarr;
            for (int k = 0; k < nl; k++)
                A_(km + N, km + nl + k) = A_(km + N, km + nl + k) + cf * d[0 * nl + k];
        }
    }
#undef A_
}

/* SGBTF2 (bandludcmp.f:238-440): ab is column-major ab[(r-1) + ldab*(c-1)]
   for Fortran AB(r,c); ipiv 1-based row indices */
static void sgbtf2(int m, int n, int kl, int ku, double *ab, int ldab, int *ipiv)
{
#define AB(r, c) ab[(long)((r) - 1) + (long)ldab * ((c) - 1)]
    int kv = ku + kl;
    for (int j = ku + 2; j <= (kv < n ? kv : n); j++)
        for (int i = kv - j + 2; i <= kl; i++) AB(i, j) = 0.;
    int ju = 1;
    int mn = m < n ? m : n;
    for (int j = 1; j <= mn; j++) {
        if (j + kv <= n)
            for (int i = 1; i <= kl; i++) AB(i, j + kv) = 0.;
        int km = kl < m - j ? kl : m - j;
        /* ISAMAX (bandludcmp.f:149-187): first index of max |x| */
        int jp = 1;
        double dmax = fabs(AB(kv + 1, j));
        for (int i = 2; i <= km + 1; i++) {
            double v = fabs(AB(kv + i, j));
            if (v > dmax) { dmax = v; jp = i; }
        }
        ipiv[j - 1] = jp + j - 1;
        if (AB(kv + jp, j) != 0.) {
            int jt = j + ku + jp - 1;
            if (jt > n) jt = n;
            if (jt > ju) ju = jt;
            if (jp != 1) /* SSWAP over columns j..ju, stride ldab-1 */
                for (int c = 0; c <= ju - j; c++) {
                    double t = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv + 1 - c, j + c);
                    AB(kv + 1 - c, j + c) = t;
                }
            if (km > 0) {
                double rcp = 1. / AB(kv + 1, j);
                for (int i = 1; i <= km; i++) AB(kv + 1 + i, j) = rcp * AB(kv + 1 + i, j);
                if (ju > j) /* SGER: A(kv+1+i-c, j+c) -= x(i)*y(c) */
                    for (int c = 1; c <= ju - j; c++) {
                        double yc = AB(kv + 1 - c, j + c);
                        if (yc != 0.)
                            for (int i = 1; i <= km; i++)
                                AB(kv + 1 + i - c, j + c) = AB(kv + 1 + i - c, j + c)
                                                            + (-yc) * AB(kv + 1 + i, j);
                    }
            }
        }
    }
#undef AB
}

/* SGBTRS 'N' (bandludcmp.f:1020-1036) with STBSV upper (ludcmp.f:4698-5046) */
static void sgbtrs(int n, int kl, int ku, const double *ab, int ldab, const int *ipiv, double *b)
{
#define AB(r, c) ab[(long)((r) - 1) + (long)ldab * ((c) - 1)]
    int kd = ku + kl + 1;
    if (kl > 0)
        for (int j = 1; j <= n - 1; j++) {
            int lm = kl < n - j ? kl : n - j;
            int l = ipiv[j - 1];
            if (l != j) {
                double t = b[l - 1]; b[l - 1] = b[j - 1]; b[j - 1] = t;
            }
            double bj = b[j - 1];
            if (bj != 0.) /* SGER with one column */
                for (int i = 1; i <= lm; i++) b[j - 1 + i] = b[j - 1 + i] + (-bj) * AB(kd + i, j);
        }
    int k = kl + ku, kplus1 = k + 1;
    for (int j = n; j >= 1; j--) {
        if (b[j - 1] != 0.) {
            b[j - 1] = b[j - 1] / AB(kplus1, j);
            double temp = b[j - 1];
            int lo = 1 > j - k ? 1 : j - k;
            for (int i = j - 1; i >= lo; i--) b[i - 1] = b[i - 1] - temp * AB(kplus1 + i - j, j);
        }
    }
#undef AB
}

typedef struct {
    double *a, *ab, *rr, *ri;
    int *ipiv;
} solvework;

static void solvework_init(const iwo_state *s, solvework *w)
{
    int nz = s->nz, kl = s->nzloc;
    w->a = dalloc((long)nz * nz);
    w->ab = dalloc((long)(3 * kl + 1) * nz);
    w->rr = dalloc(nz);
    w->ri = dalloc(nz);
    w->ipiv = calloc((size_t)nz, sizeof(int));
}
static void solvework_free(solvework *w)
{
    free(w->a); free(w->ab); free(w->rr); free(w->ri); free(w->ipiv);
}

/* penalty_main + helmv_penalty + bandlu_solve for one column (Re and Im) */
static void penalty_solve(const iwo_state *s, solvework *w, double eps, int ifbcbot, int ifbctop,
                          double bcrbot, double bcibot, double bcrtop, double bcitop,
                          double *rhsr, double *rhsi, double *fr, double *fi)
{
    int nz = s->nz, kl = s->nzloc, ku = s->nzloc, ldab = 2 * kl + ku + 1;
    penparams pp;
    for (int k = 0; k < nz; k++) { /* penalty.f:191-194 */
        rhsr[k] = -rhsr[k] * eps;
        rhsi[k] = -rhsi[k] * eps;
    }
    penalty_params(s, &pp, eps, ifbcbot, ifbctop);
    matrix_elmsetup(s, &pp, eps, w->a);
    for (int k = 0; k < nz; k++) {
        fr[k] = rhsr[k];
        fi[k] = rhsi[k];
    }
    if (ifbcbot == IDIR) { /* penalty.f:583-598 */
        fr[0] = rhsr[0] + pp.tau1[0] * bcrbot;
        fi[0] = rhsi[0] + pp.tau1[0] * bcibot;
    } else {
        fr[0] = rhsr[0] - pp.tau1[0] * bcrbot * eps;
        fi[0] = rhsi[0] - pp.tau1[0] * bcibot * eps;
    }
    if (ifbctop == IDIR) {
        fr[nz - 1] = rhsr[nz - 1] + pp.tau2[s->nsubd - 1] * bcrtop;
        fi[nz - 1] = rhsi[nz - 1] + pp.tau2[s->nsubd - 1] * bcitop;
    } else {
        fr[nz - 1] = rhsr[nz - 1] + pp.tau2[s->nsubd - 1] * bcrtop * eps;
        fi[nz - 1] = rhsi[nz - 1] + pp.tau2[s->nsubd - 1] * bcitop * eps;
    }
    /* band_lu.f:28-34 compaction */
    for (int j = 1; j <= nz; j++) {
        int lo = 1 > j - ku ? 1 : j - ku, hi = nz < j + kl ? nz : j + kl;
        for (int i = lo; i <= hi; i++)
            w->ab[(long)(kl + ku + 1 + i - j - 1) + (long)ldab * (j - 1)] = w->a[(long)(i - 1) * nz + (j - 1)];
    }
    sgbtf2(nz, nz, kl, ku, w->ab, ldab, w->ipiv);
    sgbtrs(nz, kl, ku, w->ab, ldab, w->ipiv, fr);
    sgbtrs(nz, kl, ku, w->ab, ldab, w->ipiv, fi);
}

void iwo_penalty_solve(iwo_state *s, double epsilon, int ifbcbot, int ifbctop, double bcrbot,
                       double bcibot, double bcrtop, double bcitop, double *rhsr, double *rhsi,
                       double *fr, double *fi)
{
    solvework w;
    solvework_init(s, &w);
    penalty_solve(s, &w, epsilon, ifbcbot, ifbctop, bcrbot, bcibot, bcrtop, bcitop, rhsr, rhsi, fr, fi);
    solvework_free(&w);
}

/* Legendre filter of every subdomain of a solved column (calcuvwt.f:164-194) */
static void filter_column(const iwo_state *s, double *fr)
{
    int nl = s->nzloc;
    double fin[128], fout[128];
    for (int ks = 0; ks < s->nsubd; ks++) {
        for (int k = 0; k < nl; k++) fin[k] = fr[ks * nl + k];
        filter_legenpen(s, fin, fout);
        for (int k = 0; k < nl; k++) fr[ks * nl + k] = fout[k];
    }
}

/* ------------------------------------------------------------------ */
/* interface_avg3d, penalty.f:714-775                                    */
/* ------------------------------------------------------------------ */
int iwo_interface_avg3d(iwo_state *s, double *f)
{
    const double ct = 0.005, c1 = 0.5, c2 = 0.;
    int nl = s->nzloc;
    long nxppy = s->nxppy;
    iwo_horfft(s, f, 1, s->nz);
    int ic = 0;
    for (long ij = 0; ij < nxppy; ij++)
        for (int ks = 1; ks < s->nsubd; ks++) {
            long k = (long)ks * nl, km = (long)(ks - 1) * nl + nl - 1;
            double fsp = f[k * nxppy + ij], fsm = f[km * nxppy + ij];
            double favg2 = fsp + fsm;
            double c0 = fabs(fsp - fsm) / fabs(favg2);
            if (c0 > ct) {
                double ffilter = f[(km - 1) * nxppy + ij] * c1 + f[km * nxppy + ij] * c2
                                 + f[k * nxppy + ij] * c2 + f[(k + 1) * nxppy + ij] * c1;
                f[km * nxppy + ij] = ffilter;
                f[k * nxppy + ij] = ffilter;
                ic++;
            }
        }
    iwo_horfft(s, f, -1, s->nz);
    return ic;
}

/* ------------------------------------------------------------------ */
/* forcing.f: background and virtual paddle                              */
/* ------------------------------------------------------------------ */
static double rhograd(const iwo_state *s)
{ /* forcing.f:115-120 with bruntz = brunt (forcing.f:156-162) */
    return -s->P.rho0 * (s->P.brunt * s->P.brunt) / s->P.grav;
}
static double phit(const iwo_state *s, double t)
{ /* forcing.f:424-447 */
    return (t <= s->P.fwave * s->twave) ? 1. : 0.;
}
/* shear-current profile and its derivative, forcing.f:190-214 (their
   zdim, zsh, xacu arguments are not used there) */
static double fshear(double z)
{
    const double dt = 0.13, z0 = 0.2625, zdom = 0.45;
    const double gam = 4.0 / (dt * zdom);
    return 0.5 * (tanh(gam * (z - z0)) + 1);
}
static double dfsheardz(double z)
{
    const double dt = 0.13, z0 = 0.2625, zdom = 0.45;
    const double gam = 4.0 / (dt * zdom);
    const double c = 1.0 / cosh(gam * (z - z0));
    return 0.5 * gam * (c * c);
}

void iwo_set_meanwave(iwo_state *s, const double *mw)
{
    const size_t n = (size_t)s->nxpp * s->nz;
    for (int q = 0; q < 9; q++) memcpy(s->mw[q], mw + q * n, sizeof(double) * n);
}

/* ------------------------------------------------------------------ */
/* rhsfm, solver.f:523-1206 (all arrays except u,v,w,temp physical)      */
/* ------------------------------------------------------------------ */
static void rhsfm(iwo_state *s)
{
    long ntot = s->ntot, nxppy = s->nxppy;
    int nxpp = s->nxpp, nl = s->nzloc, nz = s->nz;
    double *u = U(s), *v = V(s), *w = W(s), *tp = T(s);
    double *su = s->F[12], *sv = s->F[13], *sw = s->F[14], *st = s->F[15];
    double **g = s->g, *ss1 = s->ss1, *ss2 = s->ss2;
    double omz = s->P.omz;
    /* solver.f:627-668 vertical flux derivatives */
    for (long n = 0; n < ntot; n++) ss1[n] = -0.5 * u[n] * w[n];
    iwo_dpdz(s, ss1, su);
    for (long n = 0; n < ntot; n++) ss1[n] = -0.5 * v[n] * w[n];
    iwo_dpdz(s, ss1, sv);
    for (long n = 0; n < ntot; n++) ss1[n] = -0.5 * w[n] * w[n];
    iwo_dpdz(s, ss1, sw);
    double bcoef = (s->P.brunt * s->P.brunt) / rhograd(s); /* solver.f:693-696 */
    for (long n = 0; n < ntot; n++) { /* solver.f:671-702 */
        su[n] = su[n] - 0.5 * (u[n] * g[0][n] + v[n] * (g[1][n] - 4. * omz) + w[n] * g[2][n]);
        sv[n] = sv[n] - 0.5 * (u[n] * (g[3][n] + 4. * omz) + v[n] * g[4][n] + w[n] * g[5][n]);
        sw[n] = sw[n] - 0.5 * (u[n] * g[6][n] + v[n] * g[7][n] + w[n] * g[8][n]);
        sw[n] = sw[n] + bcoef * tp[n];
    }
    /* interface upwind penalty, solver.f:708-839 */
    double omega = 2. / ((nl + 1) * (nl + 2));
    double *fld[4] = {u, v, w, tp}, *sfl[4] = {su, sv, sw, st};
    for (long ij = 0; ij < nxppy; ij++)
        for (int ks = 0; ks < s->nsubd; ks++) {
            double ybar = s->P.zh[ks];
            long kb = (long)ks * nl, kt = (long)ks * nl + nl - 1;
            /* total vertical velocity, perturbation + mean wave */
            const int i = (int)(ij % nxpp);
            double wb = w[kb * nxppy + ij] + s->mw[1][kb * nxpp + i];
            double wt = w[kt * nxppy + ij] + s->mw[1][kt * nxpp + i];
            if (kb == 0) wb = 0.;
            if (kt == nz - 1) wt = 0.;
            double alpha0, tau1, gamma0, tau2;
            if (wb > 0.) { alpha0 = wb; tau1 = 0.5 / omega * (2. / ybar); }
            else { alpha0 = 0.; tau1 = 0.; }
            if (wt < 0.) { gamma0 = fabs(wt); tau2 = 0.5 / omega * (2. / ybar); }
            else { gamma0 = 0.; tau2 = 0.; }
            for (int q = 0; q < 4; q++) {
                double *f = fld[q], *sf = sfl[q];
                double pcb = (kb == 0) ? 0. : alpha0 * f[(kb - 1) * nxppy + ij];
                sf[kb * nxppy + ij] = sf[kb * nxppy + ij] - tau1 * (alpha0 * f[kb * nxppy + ij] - pcb);
                double pct = (kt == nz - 1) ? 0. : gamma0 * f[(kt + 1) * nxppy + ij];
                sf[kt * nxppy + ij] = sf[kt * nxppy + ij] - tau2 * (gamma0 * f[kt * nxppy + ij] - pct);
            }
        }
    /* mean-wave and shear-current terms, solver.f:848-931: advection of the
       perturbation by the mean flow and of the mean flow by the
       perturbation; d2rhomeandz2 = 0 (forcing.f:167-183) */
    {
        const double uofc = s->P.uo * s->P.fcshear;
        double **mw = s->mw;
        for (int k = 0; k < nz; k++) {
            const double zz = s->z[k];
            const double fsh = fshear(zz), dfsh = dfsheardz(zz);
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < nxpp; i++) {
                    const long n = (long)k * nxppy + (long)j * nxpp + i;
                    const long m = (long)k * nxpp + i;
                    const double advmean = mw[0][m] + uofc * fsh;
                    su[n] = su[n] - advmean * g[0][n] - mw[1][m] * g[2][n];
                    sv[n] = sv[n] - advmean * g[3][n] - mw[1][m] * g[5][n];
                    sw[n] = sw[n] - advmean * g[6][n] - mw[1][m] * g[8][n];
                    st[n] = st[n] - advmean * g[9][n] - mw[1][m] * g[11][n];
                    const double dusheardz = uofc * dfsh;
                    su[n] = su[n] - u[n] * mw[2][m] - w[n] * mw[3][m] - w[n] * dusheardz;
                    sw[n] = sw[n] - u[n] * mw[4][m] - w[n] * mw[5][m];
                    st[n] = st[n] - u[n] * mw[7][m] - w[n] * mw[8][m];
                    st[n] = st[n] + s->P.xkappa * 0.0;
                }
        }
    }
    /* virtual paddle solver.f:941-956 */
    if (s->P.xacrit != 0.) {
        double ph = phit(s, s->t), A = s->P.xacrit * s->P.umax, um = s->P.umax;
        for (int k = 0; k < nz; k++) {
            double zz = s->z[k];
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx; i++) {
                    double xx = s->x[i];
                    double arg = s->xkcrit * xx - s->xomega * s->t;
                    double wzp = s->xmcrit * cos(s->xmcrit * zz), wz = sin(s->xmcrit * zz);
                    double fu = -(1 / s->xkcrit) * A * wzp * sin(arg) * um / s->zdim * ph;
                    double fw = A * wz * cos(arg) * (um / s->zdim) * ph;
                    double fr = A * wz * sin(arg) * -(1 / s->xomega) * s->P.brunt * s->P.brunt
                                * (s->P.rho0 / s->P.grav) * (um / s->zdim) * ph;
                    long n = (long)k * nxppy + (long)j * nxpp + i;
                    su[n] = su[n] + fu;
                    sw[n] = sw[n] + fw;
                    st[n] = st[n] + fr;
                }
        }
    }
    /* horizontal flux derivatives, solver.f:975-1063 */
    double *sarr[3] = {su, sv, sw};
    for (int q = 0; q < 3; q++) {
        for (long n = 0; n < ntot; n++) {
            if (q == 0) { ss1[n] = -0.5 * u[n] * u[n]; ss2[n] = -0.5 * u[n] * v[n]; }
            if (q == 1) { ss1[n] = -0.5 * u[n] * v[n]; ss2[n] = -0.5 * v[n] * v[n]; }
            if (q == 2) { ss1[n] = -0.5 * u[n] * w[n]; ss2[n] = -0.5 * v[n] * w[n]; }
        }
        iwo_horfft(s, sarr[q], -1, nz);
        iwo_horfft(s, ss1, -1, nz);
        iwo_horfft(s, ss2, -1, nz);
        double *sc = sarr[q];
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < s->ny; j++)
                for (int ii = 0; ii < s->nxhp; ii++) {
                    long o = (long)k * nxppy + (long)j * nxpp + 2 * ii;
                    double xr = s->xw[ii] * ss1[o] + s->yw[j] * ss2[o];
                    double xi = s->xw[ii] * ss1[o + 1] + s->yw[j] * ss2[o + 1];
                    sc[o] = sc[o] + (-xi);
                    sc[o + 1] = sc[o + 1] + xr;
                }
    }
    /* density equation, solver.f:1085-1169 */
    double rg = rhograd(s);
    for (long n = 0; n < ntot; n++) ss1[n] = -w[n] * tp[n];
    iwo_dpdz(s, ss1, s->ox);
    for (long n = 0; n < ntot; n++) st[n] = st[n] + s->ox[n] - w[n] * rg;
    for (long n = 0; n < ntot; n++) {
        ss1[n] = -u[n] * tp[n];
        ss2[n] = -v[n] * tp[n];
    }
    iwo_horfft(s, st, -1, nz);
    iwo_horfft(s, ss1, -1, nz);
    iwo_horfft(s, ss2, -1, nz);
    iwo_horfft(s, tp, -1, nz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int ii = 0; ii < s->nxhp; ii++) {
                long o = (long)k * nxppy + (long)j * nxpp + 2 * ii;
                double xr = s->xw[ii] * ss1[o] + s->yw[j] * ss2[o];
                double xi = s->xw[ii] * ss1[o + 1] + s->yw[j] * ss2[o + 1];
                st[o] = st[o] + (-xi);
                st[o + 1] = st[o + 1] + xr;
            }
    iwo_horfft(s, u, -1, nz); /* solver.f:1201-1203 */
    iwo_horfft(s, v, -1, nz);
    iwo_horfft(s, w, -1, nz);
}

/* rotvisc (solver.f:1210-1253): vertical component of curl curl u at point
   k0 of subdomain ks for the mode whose real part sits at plane offset o */
static void rot_at(const iwo_state *s, const double *u, const double *v, const double *w, int ks,
                   long o, double xw, double yw, int k0, double *rr, double *ri)
{ /* solver.f:1226-1249 */
    int nl = s->nzloc;
    long nxppy = s->nxppy;
    double fur[128], fui[128], fvr[128], fvi[128];
    for (int kl = 0; kl < nl; kl++) {
        long n = (long)(ks * nl + kl) * nxppy + o;
        fur[kl] = u[n];
        fui[kl] = u[n + 1];
        fvr[kl] = v[n];
        fvi[kl] = v[n + 1];
    }
    double ybar = s->P.zh[ks];
    double dfur = dpdzcol_pt(s, fur, ybar, k0), dfui = dpdzcol_pt(s, fui, ybar, k0);
    double dfvr = dpdzcol_pt(s, fvr, ybar, k0), dfvi = dpdzcol_pt(s, fvi, ybar, k0);
    long nk = (long)(ks * nl + k0) * nxppy + o;
    double wr = w[nk], wi = w[nk + 1];
    /* (xw^2,0)*ws + (yw^2,0)*ws + (0,xw)*(dfur,dfui) + (0,yw)*(dfvr,dfvi) */
    *rr = xw * xw * wr + yw * yw * wr + (-xw * dfui) + (-yw * dfvi);
    *ri = xw * xw * wi + yw * yw * wi + xw * dfur + yw * dfvr;
}

/* ------------------------------------------------------------------ */
/* convec, solver.f:6-514                                               */
/* ------------------------------------------------------------------ */
static void convec(iwo_state *s)
{
    long ntot = s->ntot, nxppy = s->nxppy;
    int nxpp = s->nxpp, nz = s->nz, nl = s->nzloc;
    double *u = U(s), *v = V(s), *w = W(s), *tp = T(s);
    double **g = s->g;
    for (int k = 0; k < nz; k++) { /* solver.f:66-69: mean w = 0 */
        w[(long)k * nxppy] = 0.;
        w[(long)k * nxppy + 1] = 0.;
    }
    iwo_dpdz(s, u, g[2]); /* solver.f:74-77 */
    iwo_dpdz(s, v, g[5]);
    iwo_dpdz(s, w, g[8]);
    iwo_dpdz(s, tp, g[11]);
    for (int k = 0; k < nz; k++) /* solver.f:86-103: i*kx*f, i*ky*f */
        for (int j = 0; j < s->ny; j++)
            for (int ii = 0; ii < s->nxhp; ii++) {
                long o = (long)k * nxppy + (long)j * nxpp + 2 * ii;
                double xw = s->xw[ii], yw = s->yw[j];
                double *src[4] = {u, v, w, tp};
                int dx[4] = {0, 3, 6, 9}, dy[4] = {1, 4, 7, 10};
                for (int q = 0; q < 4; q++) {
                    double fr = src[q][o], fi = src[q][o + 1];
                    g[dx[q]][o] = -xw * fi;
                    g[dx[q]][o + 1] = xw * fr;
                    g[dy[q]][o] = -yw * fi;
                    g[dy[q]][o + 1] = yw * fr;
                }
            }
    /* spectral-to-physical, solver.f:107-123 (the copy of divg in ox,
       solver.f:73, 119, is never read and is skipped) */
    iwo_horfft(s, u, 1, nz);
    iwo_horfft(s, v, 1, nz);
    iwo_horfft(s, w, 1, nz);
    for (int q = 0; q < 9; q++) iwo_horfft(s, g[q], 1, nz);
    iwo_horfft(s, tp, 1, nz);
    iwo_horfft(s, g[9], 1, nz);
    iwo_horfft(s, g[10], 1, nz);
    iwo_horfft(s, g[11], 1, nz);
    for (int q = 12; q < 16; q++) memset(s->F[q], 0, sizeof(double) * (size_t)ntot);
    rhsfm(s);
    double *su = s->F[12], *sv = s->F[13], *sw = s->F[14], *st = s->F[15];
    double *sn[4] = {s->F[16], s->F[17], s->F[18], s->F[19]};
    double *snm[4] = {s->F[20], s->F[21], s->F[22], s->F[23]};
    double *scur[4] = {su, sv, sw, st};
    double *fn[4] = {s->F[4], s->F[5], s->F[6], s->F[7]};
    double *fnm[4] = {s->F[8], s->F[9], s->F[10], s->F[11]};
    double *fcur[4] = {u, v, w, tp};
    /* save NL(n) and state(n), solver.f:248-265 */
    for (int q = 0; q < 4; q++) {
        memcpy(s->work[q], scur[q], sizeof(double) * (size_t)ntot);
        memcpy(s->work[4 + q], fcur[q], sizeof(double) * (size_t)ntot);
    }
    /* wnl, solver.f:270-284 */
    for (long ij = 0; ij < nxppy; ij++) {
        long kt = (long)(nz - 1) * nxppy + ij;
        s->wnl[ij] = s->fac1 * sw[ij] + s->fac2 * sn[2][ij] + s->fac3 * snm[2][ij];
        s->wnl[nxppy + ij] = s->fac1 * sw[kt] + s->fac2 * sn[2][kt] + s->fac3 * snm[2][kt];
    }
    /* SS extrapolation, solver.f:294-316 */
    for (int q = 0; q < 4; q++)
        for (long n = 0; n < ntot; n++)
            scur[q][n] = s->fac1 * scur[q][n] + s->fac2 * sn[q][n] + s->fac3 * snm[q][n];
    /* rotational viscous pressure-BC term, solver.f:320-394 */
    for (int j = 0; j < s->ny; j++)
        for (int ii = 0; ii < s->nxhp; ii++) {
            long o = (long)j * nxpp + 2 * ii;
            double xw = s->xw[ii], yw = s->yw[j];
            for (int side = 0; side < 2; side++) {
                int ks = side == 0 ? 0 : s->nsubd - 1, k0 = side == 0 ? 0 : nl - 1;
                double r0r, r0i, r1r, r1i, r2r, r2i;
                rot_at(s, u, v, w, ks, o, xw, yw, k0, &r0r, &r0i);
                rot_at(s, fn[0], fn[1], fn[2], ks, o, xw, yw, k0, &r1r, &r1i);
                rot_at(s, fnm[0], fnm[1], fnm[2], ks, o, xw, yw, k0, &r2r, &r2i);
                double rfr = s->fac1 * r0r + s->fac2 * r1r + s->fac3 * r2r;
                double rfi = s->fac1 * r0i + s->fac2 * r1i + s->fac3 * r2i;
                s->wviscrot[side * nxppy + o] = -s->P.xnu * rfr;
                s->wviscrot[side * nxppy + o + 1] = -s->P.xnu * rfi;
            }
        }
    /* BDF assembly, solver.f:402-430 */
    for (int q = 0; q < 4; q++)
        for (long n = 0; n < ntot; n++)
            fcur[q][n] = s->dt * scur[q][n] + s->bdf1 * fcur[q][n] + s->bdf2 * fn[q][n]
                         + s->bdf3 * fnm[q][n];
    /* history rotation, solver.f:447-478 */
    for (int q = 0; q < 4; q++) {
        memcpy(snm[q], sn[q], sizeof(double) * (size_t)ntot);
        memcpy(sn[q], s->work[q], sizeof(double) * (size_t)ntot);
        memcpy(fnm[q], fn[q], sizeof(double) * (size_t)ntot);
        memcpy(fn[q], s->work[4 + q], sizeof(double) * (size_t)ntot);
    }
    if (s->P.ispecfilternl == 1) /* solver.f:483-490 */
        for (int q = 0; q < 4; q++) iwo_filter_3dlegenfourier(s, fcur[q], 1, 1);
}

/* ------------------------------------------------------------------ */
/* calcuvwt, calcuvwt.f:2-746 (+ calcuvwt_noslip.f), update 826-901      */
/* ------------------------------------------------------------------ */
static void calcuvwt(iwo_state *s, int cpenspecfilter)
{
    int nz = s->nz, nl = s->nzloc, nxpp = s->nxpp;
    long nxppy = s->nxppy;
    double *u = U(s), *v = V(s), *w = W(s), *tp = T(s);
    double dt = s->dt, xnu = s->P.xnu;
    int nmode = s->ny * s->nxhp;
    /* pressure, calcuvwt.f:87-222 */
#pragma omp parallel
    {
        solvework sw;
        solvework_init(s, &sw);
        double *rr = dalloc(nz), *ri = dalloc(nz), *fr = dalloc(nz), *fi = dalloc(nz);
        double *col = dalloc(nl * 2);
#pragma omp for schedule(dynamic)
        for (int m = 0; m < nmode; m++) {
            int j = m / s->nxhp, ii = m % s->nxhp;
            long o = (long)j * nxpp + 2 * ii;
            if (ii == 0 && j == 0) {
                for (int k = 0; k < nz; k++) {
                    s->phi[(long)k * nxppy + o] = 0.;
                    s->phi[(long)k * nxppy + o + 1] = 0.;
                }
                continue;
            }
            double ak = s->xsq[ii] + s->ysq[j];
            double eps = 1. / ak;
            double xw = s->xw[ii], yw = s->yw[j];
            for (int ks = 0; ks < s->nsubd; ks++) { /* divcolc, derv.f:35-85 */
                double wr[128], wi[128];
                for (int kl = 0; kl < nl; kl++) {
                    long n = (long)(ks * nl + kl) * nxppy + o;
                    wr[kl] = w[n];
                    wi[kl] = w[n + 1];
                }
                for (int kl = 0; kl < nl; kl++) {
                    long n = (long)(ks * nl + kl) * nxppy + o;
                    double dfr = dpdzcol_pt(s, wr, s->P.zh[ks], kl);
                    double dfi = dpdzcol_pt(s, wi, s->P.zh[ks], kl);
                    double dr = (0. + (-xw * u[n + 1])) + (-yw * v[n + 1]);
                    double di = (0. + xw * u[n]) + yw * v[n];
                    dr = dr + dfr;
                    di = di + dfi;
                    rr[ks * nl + kl] = dr / dt;
                    ri[ks * nl + kl] = di / dt;
                }
            }
            double bcrtop = s->wnl[nxppy + o] + s->wviscrot[nxppy + o];
            double bcrbot = s->wnl[o] + s->wviscrot[o];
            double bcitop = s->wnl[nxppy + o + 1] + s->wviscrot[nxppy + o + 1];
            double bcibot = s->wnl[o + 1] + s->wviscrot[o + 1];
            penalty_solve(s, &sw, eps, INEU, INEU, bcrbot, bcibot, bcrtop, bcitop, rr, ri, fr, fi);
            if (cpenspecfilter) {
                filter_column(s, fr);
                filter_column(s, fi);
            }
            for (int k = 0; k < nz; k++) {
                s->phi[(long)k * nxppy + o] = fr[k];
                s->phi[(long)k * nxppy + o + 1] = fi[k];
            }
        }
        (void)col;
        free(rr); free(ri); free(fr); free(fi); free(col);
        solvework_free(&sw);
    }
    /* update, calcuvwt.f:826-901 (no-slip file: k = 2..nz, calcuvwt_noslip.f:851) */
    iwo_dpdz(s, s->phi, s->dphidz);
    int k0 = (s->P.bc_uv_bottom == 1) ? 1 : 0;
    for (int k = k0; k < nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int ii = 0; ii < s->nxhp; ii++) {
                long o = (long)k * nxppy + (long)j * nxpp + 2 * ii;
                double tr = s->phi[o], ti = s->phi[o + 1];
                double cx = dt * s->xw[ii], cy = dt * s->yw[j];
                s->ub[o] = u[o] - (0. * tr - cx * ti);
                s->ub[o + 1] = u[o + 1] - (0. * ti + cx * tr);
                s->vb[o] = v[o] - (0. * tr - cy * ti);
                s->vb[o + 1] = v[o + 1] - (0. * ti + cy * tr);
                if (ii == 0 && j == 0) {
                    s->wb[o] = 0.;
                    s->wb[o + 1] = 0.;
                } else {
                    s->wb[o] = w[o] - dt * s->dphidz[o];
                    s->wb[o + 1] = w[o + 1] - dt * s->dphidz[o + 1];
                }
            }
    /* viscous steps, calcuvwt.f:247-741 */
    int ubot = (s->P.bc_uv_bottom == 1) ? IDIR : INEU;
#pragma omp parallel
    {
        solvework sw;
        solvework_init(s, &sw);
        double *rr = dalloc(nz), *ri = dalloc(nz), *fr = dalloc(nz), *fi = dalloc(nz);
#pragma omp for schedule(dynamic)
        for (int m = 0; m < nmode; m++) {
            int j = m / s->nxhp, ii = m % s->nxhp;
            long o = (long)j * nxpp + 2 * ii;
            int zero = (ii == 0 && j == 0);
            double ak = s->xsq[ii] + s->ysq[j];
            double epsv = 1. / (ak + s->bdfv * s->facvel);
            double epst = 1. / (ak + s->bdfv * s->fact);
            double *src[4] = {s->ub, s->vb, s->wb, tp};
            double *dst[4] = {u, v, w, tp};
            for (int q = 0; q < 4; q++) {
                if (q == 2 && zero) { /* calcuvwt.f:496-503 */
                    for (int k = 0; k < nz; k++) {
                        w[(long)k * nxppy + o] = 0.;
                        w[(long)k * nxppy + o + 1] = 0.;
                    }
                    continue;
                }
                for (int k = 0; k < nz; k++) {
                    long n = (long)k * nxppy + o;
                    if (q < 3) {
                        rr[k] = -src[q][n] / (xnu * dt);
                        ri[k] = -src[q][n + 1] / (xnu * dt);
                    } else {
                        rr[k] = -src[q][n] / (dt * s->P.xkappa);
                        ri[k] = -src[q][n + 1] / (dt * s->P.xkappa);
                    }
                }
                int bb = (q < 2) ? ubot : (q == 2 ? IDIR : INEU);
                int bt = (q == 2) ? IDIR : INEU;
                penalty_solve(s, &sw, q < 3 ? epsv : epst, bb, bt, 0., 0., 0., 0., rr, ri, fr, fi);
                if (cpenspecfilter) {
                    filter_column(s, fr);
                    filter_column(s, fi);
                }
                if (q == 0 && zero) /* mean-u damping, calcuvwt.f:343-361 */
                    for (int k = 0; k < nz; k++) {
                        double zz = (s->z[k] - s->P.zcen) / s->zdim;
                        double g = exp(-s->P.xbcrit * (zz * zz));
                        fr[k] = (1.0 - g) * fr[k];
                        fi[k] = (1.0 - g) * fi[k];
                    }
                for (int k = 0; k < nz; k++) {
                    dst[q][(long)k * nxppy + o] = fr[k];
                    dst[q][(long)k * nxppy + o + 1] = fi[k];
                }
            }
        }
        free(rr); free(ri); free(fr); free(fi);
        solvework_free(&sw);
    }
}

/* ------------------------------------------------------------------ */
/* driver: main.f:653-1066 (istart=0, fixed dt)                           */
/* ------------------------------------------------------------------ */
/* plotf (output.f:125-260): u,v,w,temp to physical space for the dump and
   back, the only effect on the state */
static void plotf_trip(iwo_state *s)
{
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, s->nz);
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, s->nz);
}

void iwo_set_initial(iwo_state *s, const double *u, const double *v, const double *w, const double *t)
{
    long nb = sizeof(double) * (size_t)s->ntot;
    memcpy(U(s), u, nb);
    memcpy(V(s), v, nb);
    memcpy(W(s), w, nb);
    memcpy(T(s), t, nb);
    for (int q = 4; q < 24; q++) memset(s->F[q], 0, nb);
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, s->nz); /* init.f:196-205 */
    plotf_trip(s); /* output of the initial field, main.f:537-544 */
    s->t = 0.;
    s->istep = 0;
    s->istart = 0;
    s->fac1 = 1.0; s->fac2 = 0.; s->fac3 = 0.;
    s->bdf1 = 1.; s->bdf2 = 0.; s->bdf3 = 0.;
    s->bdfv = 1.;
}

/* ------------------------------------------------------------------ */
/* Estimation-model filter and SGS tensors: filtering.f:8-270,         */
/* 989-1105, 1272-1310, sgscalc.f:4-223                                */
/* ------------------------------------------------------------------ */

/* Boyd-Vandeven filter (filtering.f:1272-1310).  Its erfc is the Fortran
 * 2008 intrinsic (no EXTERNAL declaration in filtering.f), i.e. the C
 * library's erfc: the compiled reference references libm's erfc there and
 * leaves the Numerical Recipes version of erfc.f unused. */
static double specfilter_bv(int k, int kc, int n, double p)
{
    if (k > kc && k < n) {
        double theta = (double)(k - kc) / (double)(n - kc);
        double omega = fabs(theta) - 0.5, chi;
        if (theta == 0.5) {
            chi = 1.;
        } else {
            double c1 = 4. * pow(omega, 2.);
            double c2 = -log(1. - c1) / c1;
            chi = sqrt(c2);
        }
        double arg = 2. * sqrt(p) * chi * omega;
        return 0.5 * erfc(arg);
    }
    return (k == n) ? 0. : 1.;
}

/* 3-point periodic top-hat filters in x and y (filtering.f:162-270):
 * uf = u everywhere, then uf = u(i-1)/4 + u(i)/2 + u(i+1)/4 for i <= nx */
static void filterx3pv(const iwo_state *s, const double *u, double *uf)
{
    const double a = 1. / 4., b = 2. / 4., c = 1. / 4.;
    const int nx = s->nx, nxpp = s->nxpp;
    memcpy(uf, u, sizeof(double) * (size_t)s->ntot);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)s->ny * s->nz; r++) {
        const double *row = u + r * nxpp;
        double *o = uf + r * nxpp;
        for (int i = 0; i < nx; i++) {
            double um = row[i == 0 ? nx - 1 : i - 1], up = row[i == nx - 1 ? 0 : i + 1];
            o[i] = a * um + b * row[i] + c * up;
        }
    }
}

static void filtery3pv(const iwo_state *s, const double *u, double *uf)
{
    const double a = 1. / 4., b = 2. / 4., c = 1. / 4.;
    const int nx = s->nx, ny = s->ny, nxpp = s->nxpp;
    memcpy(uf, u, sizeof(double) * (size_t)s->ntot);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < ny; j++) {
            const double *rm = u + ((long)k * ny + (j == 0 ? ny - 1 : j - 1)) * nxpp;
            const double *r0 = u + ((long)k * ny + j) * nxpp;
            const double *rp = u + ((long)k * ny + (j == ny - 1 ? 0 : j + 1)) * nxpp;
            double *o = uf + ((long)k * ny + j) * nxpp;
            for (int i = 0; i < nx; i++) o[i] = a * rm[i] + b * r0[i] + c * rp[i];
        }
}

/* filter_legensgsz (filtering.f:989-1105): Legendre transform of every
 * column i <= nx of each subdomain, Boyd-Vandeven filter of order 2 with
 * cut-off kc = int(fackcbv*nzm), back-transform */
static void filter_legensgsz(const iwo_state *s, double *uin)
{
    const int nl = s->nzloc, nzm = s->nzm, nx = s->nx, nxpp = s->nxpp;
    const int kc = (int)(s->P.fackcbv * nzm);
    const long nxppy = s->nxppy;
    double sig[128];
    for (int k = 0; k <= nzm; k++) sig[k] = specfilter_bv(k, kc, nzm, 2.);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < s->ny; j++) {
        double fin[128], a[128];
        for (int i = 0; i < nx; i++)
            for (int ks = 0; ks < s->nsubd; ks++) {
                const long o = (long)ks * nl * nxppy + (long)j * nxpp + i;
                for (int kl = 0; kl < nl; kl++) fin[kl] = uin[o + kl * nxppy];
                for (int k = 0; k <= nzm; k++) {
                    double sum = 0.;
                    for (int k1 = 0; k1 < nl; k1++) sum = sum + s->wg[k1] * fin[k1] * s->alpol[k * nl + k1];
                    a[k] = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
                }
                for (int k = 0; k <= nzm; k++) a[k] = a[k] * sig[k];
                for (int k1 = 0; k1 < nl; k1++) {
                    double sum = 0.;
                    for (int k = 0; k <= nzm; k++) sum = sum + a[k] * s->alpol[k * nl + k1];
                    uin[o + k1 * nxppy] = sum;
                }
            }
    }
}

/* modecut (filtering.f:100-156): u0 = 3 f(u) - 3 f(f(u)) + f(f(f(u))) with
 * f = y-filter after x-filter, for i <= nx, then the vertical SGS filter */
static void modecut(iwo_state *s, const double *u, double *u0)
{
    const long n = s->ntot;
    double *tmp2 = dalloc(n), *g3 = dalloc(n), *g3s = dalloc(n), *g3c = dalloc(n);
    filterx3pv(s, u, tmp2);
    filtery3pv(s, tmp2, g3);
    filterx3pv(s, g3, tmp2);
    filtery3pv(s, tmp2, g3s);
    filterx3pv(s, g3s, tmp2);
    filtery3pv(s, tmp2, g3c);
    for (long r = 0; r < (long)s->ny * s->nz; r++)
        for (int i = 0; i < s->nx; i++) {
            long o = r * s->nxpp + i;
            u0[o] = 3. * g3[o] - 3. * g3s[o] + g3c[o];
        }
    free(tmp2);
    free(g3);
    free(g3s);
    free(g3c);
    filter_legensgsz(s, u0);
}

/* mean of temp per level through its horfft round trip (sgscalc.f:21-28,
 * 129-136); the round trip itself rescales temp's kx-Nyquist content */
static void temp_mean(iwo_state *s, double *tm)
{
    iwo_horfft(s, T(s), -1, s->nz);
    for (int k = 0; k < s->nz; k++) tm[k] = T(s)[(long)k * s->nxppy];
    iwo_horfft(s, T(s), 1, s->nz);
}

static void sgscalc_pre(iwo_state *s)
{ /* sgscalc.f:4-57 */
    double **t = s->sgs, tm[4096];
    const long nxppy = s->nxppy;
    temp_mean(s, tm);
    for (int j = 0; j < s->ny; j++)
        for (int i = 0; i < s->nx; i++) {
            long o1 = (long)j * s->nxpp + i;
            for (int q = 0; q < 9; q++) t[q][o1] = 0.;
            for (int k = 1; k < s->nz; k++) {
                long o = (long)k * nxppy + o1;
                double u = U(s)[o], v = V(s)[o], w = W(s)[o];
                t[0][o] = u * u;
                t[3][o] = v * v;
                t[5][o] = w * w;
                t[1][o] = u * v;
                t[2][o] = u * w;
                t[4][o] = v * w;
                s->ox[o] = T(s)[o] - tm[k];
                t[6][o] = u * s->ox[o];
                t[7][o] = v * s->ox[o];
                t[8][o] = w * s->ox[o];
            }
        }
}

static void sgscalc_post(iwo_state *s)
{ /* sgscalc.f:62-223 */
    double **t = s->sgs, **tp = s->tmps, tm[4096];
    const long nxppy = s->nxppy;
    const int nz = s->nz;
    modecut(s, t[0], tp[0]); /* t11 */
    modecut(s, t[3], tp[1]); /* t22 */
    modecut(s, t[5], tp[2]); /* t33 */
    modecut(s, t[1], tp[3]); /* t12 */
    modecut(s, t[2], tp[4]); /* t13 */
    modecut(s, t[4], tp[5]); /* t23 */
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                long o = (long)k * nxppy + (long)j * s->nxpp + i;
                double u = U(s)[o], v = V(s)[o], w = W(s)[o];
                t[0][o] = tp[0][o] - u * u;
                t[3][o] = tp[1][o] - v * v;
                t[1][o] = tp[3][o] - u * v;
                if (k == nz - 1) {
                    t[5][o] = 0.;
                    t[2][o] = 0.;
                    t[4][o] = 0.;
                } else {
                    t[5][o] = tp[2][o] - w * w;
                    t[2][o] = tp[4][o] - u * w;
                    t[4][o] = tp[5][o] - v * w;
                }
            }
    temp_mean(s, tm);
    modecut(s, t[6], tp[0]);
    modecut(s, t[7], tp[1]);
    modecut(s, t[8], tp[2]);
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                long o = (long)k * nxppy + (long)j * s->nxpp + i;
                s->ox[o] = T(s)[o] - tm[k];
                if (k == nz - 1) {
                    t[6][o] = 0.;
                    t[7][o] = 0.;
                    t[8][o] = 0.;
                } else {
                    t[6][o] = tp[0][o] - U(s)[o] * s->ox[o];
                    t[7][o] = tp[1][o] - V(s)[o] * s->ox[o];
                    t[8][o] = tp[2][o] - W(s)[o] * s->ox[o];
                }
            }
}

void iwo_filtering(iwo_state *s, int isgscalc)
{ /* filtering.f:8-93 */
    double **tp = s->tmps;
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, s->nz);
    if (isgscalc) sgscalc_pre(s);
    modecut(s, U(s), tp[0]);
    modecut(s, V(s), tp[1]);
    modecut(s, W(s), tp[2]);
    modecut(s, T(s), tp[3]); /* computed, not used (filtering.f:48, 67) */
    const int nz = s->nz;
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                long o = (long)k * s->nxppy + (long)j * s->nxpp + i;
                U(s)[o] = tp[0][o];
                V(s)[o] = tp[1][o];
                if (k != nz - 1) W(s)[o] = tp[2][o];
            }
    if (isgscalc) sgscalc_post(s);
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, s->nz);
}

double *iwo_sgs(iwo_state *s, int idx) { return (idx >= 0 && idx < 9) ? s->sgs[idx] : 0; }

void iwo_set_restart(iwo_state *s, const double *state, double t, double cdt)
{ /* restart.f:26-115 then main.f:619-634 */
    for (int q = 0; q < 24; q++)
        memcpy(s->F[q], state + (size_t)q * s->ntot, sizeof(double) * (size_t)s->ntot);
    s->t = t;
    s->cdt = cdt;
    double d1 = s->dt, d2 = cdt * s->dt;
    s->fac1 = iwo_abvar1(d1, d2, d2); s->fac2 = iwo_abvar2(d1, d2, d2); s->fac3 = iwo_abvar3(d1, d2, d2);
    s->bdf1 = iwo_bdfvar1(d1, d2, d2); s->bdf2 = iwo_bdfvar2(d1, d2, d2); s->bdf3 = iwo_bdfvar3(d1, d2, d2);
    s->bdfv = iwo_bdfvarv(d1, d2, d2);
    s->istep = 0;
    s->istart = 1;
    for (int q = 0; q < 9; q++) /* a restarted run has no SGS tensors yet */
        memset(s->sgs[q], 0, sizeof(double) * (size_t)s->ntot);
    plotf_trip(s); /* output of the restart field, main.f:599-608 */
    if (s->mccap < 1) { /* main.f:634-635: its record is the one energy record until the next run */
        free(s->mcrec);
        free(s->erec);
        free(s->lrec);
        s->mcrec = (double *)calloc(IWO_NREC, sizeof(double));
        s->erec = (double *)calloc(IWO_NEREC, sizeof(double));
        s->lrec = (double *)calloc((size_t)(2 + 4 * s->nz), sizeof(double));
        s->mccap = 1;
    }
    s->erec[0] = 0.;
    s->erec[1] = s->t;
    iwo_postp(s, s->erec + 2);
    s->en = 1;
    s->ln = 0;
}

/* variable-step coefficients for the step sizes (dt1, dt2, dt3), the
   blocks of main.f:619-634 / various.f:945-958 */
static void var_coeffs(iwo_state *s, double d1, double d2, double d3)
{
    s->fac1 = iwo_abvar1(d1, d2, d3); s->fac2 = iwo_abvar2(d1, d2, d3); s->fac3 = iwo_abvar3(d1, d2, d3);
    s->bdf1 = iwo_bdfvar1(d1, d2, d3); s->bdf2 = iwo_bdfvar2(d1, d2, d3); s->bdf3 = iwo_bdfvar3(d1, d2, d3);
}

/* maxcalc (various.f:665-987): the physical u,v,w,temp (horfft +1), the
   running extrema and CFL numbers in the reference's loop order (i, j, then
   z; a later point wins a tie, various.f:755-845), the return to spectral
   space (the state-changing round trip) and, with dtadapt, the adaptive
   step of various.f:887-985.  rec (IWO_NREC doubles) receives the outputs. */
void iwo_maxcalc(iwo_state *s, int istep, double *rec)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz, nzl = s->nzloc, ns = s->nsubd;
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, nz);
    const double pi2 = 8. * atan(1.0); /* various.f:696-700 */
    const double xl = pi2 / s->alpha, yl = pi2 / s->beta;
    const double dy = yl / ny, dx = xl / nx;
    /* zf (comgridptxyz.h) as setup.f:150-161 fills it (the upper subdomain
       writes an interface point last); zf(0) is the word before zf in
       /comgridptxyz/, z(nz) */
    const int nz2 = ns * (nzl - 1) + 1;
    double *zf = (double *)malloc(sizeof(double) * (size_t)(nz2 + 1));
    zf[0] = s->z[nz - 1];
    for (int ks = 0; ks < ns; ks++)
        for (int kl = 0; kl < nzl; kl++) zf[ks * (nzl - 1) + kl + 1] = s->z[ks * nzl + kl];
    const double b = s->P.brunt;
    const double rhograd = -(s->P.rho0 * (b * b)) / s->P.grav; /* forcing.f:115-120 */
    double umin = 0., umaxt = 0., vmin = 0., vmax = 0., wmin = 0., wmax = 0., tmin = 1., tmax = 0.;
    double cflxmax = 0., cflymax = 0., cflzmax = 0., dzmax = s->dzlast, dz = s->dzlast;
    int kumin = 0, kumax = 0, kvmin = 0, kvmax = 0, kwmin = 0, kwmax = 0, ktmin = 0, ktmax = 0;
    int itmax = 0, jtmax = 0, icflx = 0, icfly = 0, icflz = 0, jcflx = 0, jcfly = 0, jcflz = 0;
    int kcflx = 0, kcfly = 0, kcflz = 0;
    const double *u = U(s), *v = V(s), *w = W(s), *tp = T(s);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int ks = 0; ks < ns; ks++)
                for (int kl = 0; kl < nzl; kl++) {
                    const int k = ks * nzl + kl, kg = ks * (nzl - 1) + kl + 1;
                    const long o = (long)k * s->nxppy + (long)j * s->nxpp + i;
                    const long m = (long)k * s->nxpp + i;
                    const double uu = u[o] + s->mw[0][m], vv = v[o], ww = w[o] + s->mw[1][m];
                    if (uu >= umaxt) { umaxt = uu + s->P.umax; kumax = k + 1; }
                    if (uu <= umin) { umin = uu; kumin = k + 1; }
                    if (vv >= vmax) { vmax = vv; kvmax = k + 1; }
                    if (vv <= vmin) { vmin = vv; kvmin = k + 1; }
                    if (ww >= wmax) { wmax = ww; kwmax = k + 1; }
                    if (ww <= wmin) { wmin = ww; kwmin = k + 1; }
                    const double rm = 1. + rhograd * s->z[k]; /* rhomean, forcing.f:134-141 */
                    const double tt = (tp[o] + s->mw[6][m]) + rm;
                    if (tt >= tmax) { tmax = tt; ktmax = k + 1; itmax = i + 1; jtmax = j + 1; }
                    if (tt <= tmin) { tmin = tt; ktmin = k + 1; s->itmin = i + 1; s->jtmin = j + 1; }
                    const double cflx = fabs(((s->P.uo * s->P.fcshear * fshear(s->z[k]) + u[o]) * s->dt) / dx);
                    const double cfly = fabs((v[o] * s->dt) / dy);
                    double cflz;
                    if (k + 1 == nz) {
                        cflz = 0.;
                    } else {
                        dz = zf[kg] - zf[kg - 1];
                        cflz = fabs((w[o] * s->dt) / dz);
                    }
                    if (cflx >= cflxmax) { cflxmax = cflx; icflx = i + 1; jcflx = j + 1; kcflx = k + 1; }
                    if (cfly >= cflymax) { cflymax = cfly; icfly = i + 1; jcfly = j + 1; kcfly = k + 1; }
                    if (cflz >= cflzmax) {
                        cflzmax = cflz; icflz = i + 1; jcflz = j + 1; kcflz = k + 1;
                        dzmax = dz;
                    }
                }
    free(zf);
    s->dzlast = dz;
    /* uveltest.dat probe (various.f:744-746): u(32,1,10), u(32,1,50) by
       their linear position in u */
    const long p1 = 31 + (long)s->nxpp * ((long)ny * 9), p2 = 31 + (long)s->nxpp * ((long)ny * 49);
    const double pr1 = p1 < s->ntot ? u[p1] : NAN, pr2 = p2 < s->ntot ? u[p2] : NAN;
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, nz);
    double r[IWO_NREC] = {0};
    r[0] = istep; r[1] = s->t;
    r[2] = umin; r[3] = umaxt; r[4] = kumin; r[5] = kumax;
    r[6] = vmin; r[7] = vmax; r[8] = kvmin; r[9] = kvmax;
    r[10] = wmin; r[11] = wmax; r[12] = kwmin; r[13] = kwmax;
    r[14] = tmin; r[15] = tmax; r[16] = ktmin; r[17] = ktmax;
    r[18] = s->itmin; r[19] = s->jtmin; r[20] = itmax; r[21] = jtmax;
    r[22] = cflxmax; r[23] = icflx; r[24] = jcflx; r[25] = kcflx;
    r[26] = cflymax; r[27] = icfly; r[28] = jcfly; r[29] = kcfly;
    r[30] = cflzmax; r[31] = icflz; r[32] = jcflz; r[33] = kcflz;
    r[34] = dzmax;
    r[40] = pr1; r[41] = pr2;
    int changed = 0;
    if (s->P.dtadapt) { /* various.f:887-985 */
        const int change_dt = (cflzmax > 0.7) || (cflzmax < 0.1) || (cflxmax > 0.16);
        const double tlim = pi2 / (s->xomega * 50.);
        const double tso = s->P.dtadapt == 2 ? 0. : s->Tso;
        const int nst = (int)((s->t - tso) / s->dt);
        if (nst > 20 && change_dt) {
            s->Tso = s->t;
            if (cflzmax < 0.1) s->cdt = fmax(0.8, s->dt / tlim);
            if ((cflzmax > 0.7) || (cflxmax > 0.16)) s->cdt = 1.25;
            s->dt = s->dt / s->cdt;
            var_coeffs(s, s->dt, s->cdt * s->dt, s->cdt * s->dt);
            s->bdfv = iwo_bdfvarv(s->dt, s->cdt * s->dt, s->cdt * s->dt);
            s->fact = 1. / (s->P.xkappa * s->dt); /* pr = xnu */
            s->facvel = 1. / (s->P.xnu * s->dt);
            s->ivardt = 1;
            s->idtstep = 0;
            changed = 1;
        }
    }
    r[35] = s->dt; r[36] = s->cdt; r[37] = changed; r[38] = s->ivardt; r[39] = s->idtstep;
    if (rec) memcpy(rec, r, sizeof r);
}

double *iwo_speclegen_records(iwo_state *s, int *n)
{
    if (n) *n = s->ln;
    return s->lrec;
}

/* speclegencalc (filtering.f:1110-1221) */
void iwo_speclegen(iwo_state *s, double *f, double *out)
{
    const int nl = s->nzloc, nzm = nl - 1, nx = s->nx;
    /* j = nyh (1-based); ny = 1 has no such row (the reference reads
       outside the array there), row 1 is taken instead */
    const long j = s->ny >= 2 ? s->ny / 2 - 1 : 0;
    iwo_horfft(s, f, 1, s->nz);
    double *acc = (double *)malloc(sizeof(double) * (size_t)nl);
    for (int ks = 0; ks < s->nsubd; ks++) {
        for (int kl = 0; kl < nl; kl++) acc[kl] = 0.;
        for (int i = 0; i < nx; i++) {
            for (int k = 0; k <= nzm; k++) {
                double sum = 0.;
                for (int k1 = 0; k1 < nl; k1++)
                    sum = sum + s->wg[k1] * f[(long)(ks * nl + k1) * s->nxppy + j * s->nxpp + i]
                                    * s->alpol[k * nl + k1];
                const double a = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
                acc[k] = acc[k] + fabs(a);
            }
        }
        for (int kl = 0; kl < nl; kl++) out[kl * s->nsubd + ks] = acc[kl] / nx;
    }
    free(acc);
    iwo_horfft(s, f, -1, s->nz);
}

double *iwo_energy_records(iwo_state *s, int *n)
{
    if (n) *n = s->en;
    return s->erec;
}

double *iwo_maxcalc_records(iwo_state *s, int *n)
{
    if (n) *n = s->mcn;
    return s->mcrec;
}

void iwo_run(iwo_state *s, int nsteps, int *counts)
{
    s->mcn = 0;
    s->en = 0;
    s->ln = 0;
    if (nsteps > s->mccap) {
        free(s->mcrec);
        free(s->erec);
        free(s->lrec);
        s->mcrec = (double *)calloc((size_t)nsteps * IWO_NREC, sizeof(double));
        s->erec = (double *)calloc((size_t)nsteps * IWO_NEREC, sizeof(double));
        s->lrec = (double *)calloc((size_t)nsteps * (2 + 4 * s->nz), sizeof(double));
        s->mccap = nsteps;
    }
    for (int n = 0; n < nsteps; n++) {
        int istep = ++s->istep;
        if (s->P.imaxcalc || s->P.dtadapt) { /* main.f:657 */
            iwo_maxcalc(s, istep, s->mcrec + (size_t)n * IWO_NREC);
            s->mcn = n + 1;
        } else { /* only the state-changing part, various.f:739-742, 854-857 */
            for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, s->nz);
            for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, s->nz);
        }
        if (s->ivardt == 1) s->idtstep++; /* main.f:660-664 */
        convec(s);
        if (istep < 2) { /* main.f:700-734 */
            if (!s->istart) {
                s->fac1 = 2.; s->fac2 = -1.; s->fac3 = 0.;
                s->bdf1 = 2.; s->bdf2 = -0.5; s->bdf3 = 0.;
            } else {
                var_coeffs(s, s->dt, s->dt, s->cdt * s->dt);
            }
        } else if (!(s->ivardt == 1 && s->idtstep < 2)) { /* main.f:741-756 */
            s->fac1 = 3.; s->fac2 = -3.; s->fac3 = 1.;
            s->bdf1 = 3.; s->bdf2 = -1.5; s->bdf3 = 1. / 3.;
        } else { /* first step after a change of dt, main.f:757-776 */
            var_coeffs(s, s->dt, s->dt, s->cdt * s->dt);
        }
        for (int q = 0; q < 4; q++) /* su = u ... (main.f:785-795) */
            memcpy(s->F[12 + q], s->F[q], sizeof(double) * (size_t)s->ntot);
        int cpen = (s->P.ispecfilterexp == 1) && (istep % s->P.nspecfilterexp == 0);
        calcuvwt(s, cpen);
        if (istep < 2) { /* main.f:813-826 */
            s->bdfv = s->istart ? iwo_bdfvarv(s->dt, s->dt, s->cdt * s->dt) : 1.5;
        } else if (!(s->ivardt == 1 && s->idtstep < 2)) { /* main.f:832-837 */
            s->bdfv = 11. / 6.;
        } else { /* main.f:838-854 */
            s->bdfv = iwo_bdfvarv(s->dt, s->dt, s->cdt * s->dt);
            s->cdt = 1.;
            s->ivardt = 0;
        }
        if (s->P.ifilter == 1 && s->P.nstepfilter > 0 && istep % s->P.nstepfilter == 0) /* main.f:880-895 */
            iwo_filtering(s, s->P.npostp > 0 && istep % s->P.npostp == 0);
        s->t = s->t + s->dt;
        if (s->P.ncalclegspec > 0 && istep % s->P.ncalclegspec == 0) { /* main.f:905-911 */
            double *e = s->lrec + (size_t)s->ln * (2 + 4 * s->nz);
            e[0] = istep;
            e[1] = s->t;
            for (int q = 0; q < 4; q++) iwo_speclegen(s, s->F[q], e + 2 + (size_t)q * s->nz);
            s->ln++;
        }
        int ic[4] = {0, 0, 0, 0};
        if (s->P.navgintrfc > 0 && istep % s->P.navgintrfc == 0) /* main.f:920-932 */
            for (int q = 0; q < 4; q++) ic[q] = iwo_interface_avg3d(s, s->F[q]);
        if (counts)
            for (int q = 0; q < 4; q++) counts[4 * n + q] = ic[q];
        if (s->P.npostp > 0 && istep % s->P.npostp == 0) { /* main.f:935-936 */
            double *e = s->erec + (size_t)s->en * IWO_NEREC;
            e[0] = istep;
            e[1] = s->t;
            iwo_postp(s, e + 2);
            s->en++;
        }
        memset(s->F[12], 0, sizeof(double) * (size_t)s->ntot); /* main.f:938-944 */
        if (s->P.n3doutput > 0 && istep % s->P.n3doutput == 0) plotf_trip(s); /* main.f:996-1039 */
    }
}

/* ------------------------------------------------------------------ */
/* postp (postp.f:9-76): umeancalc, dudzmeancalc, dudymeancalc,          */
/* perturbcalc(0) and energycalc_postp (postp.f:129-363, 415-732)        */
/* ------------------------------------------------------------------ */
void iwo_postp(iwo_state *s, double *out)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz, nxpp = s->nxpp, nl = s->nzloc;
    const long nxppy = s->nxppy;
    double *u = U(s), *v = V(s), *w = W(s), *tp = T(s);
    double **g = s->g;
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, nz); /* postp.f:23-26 */
    /* umeancalc: x means at every (j,k), k >= 2 (k = 1 stays 0) */
    double *um = (double *)calloc((size_t)ny * nz, sizeof(double));
    double *tm = (double *)calloc((size_t)6 * ny * nz, sizeof(double));
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            const long o = (long)k * nxppy + (long)j * nxpp;
            double a = 0., t6[6] = {0., 0., 0., 0., 0., 0.};
            for (int i = 0; i < nx; i++) {
                a = a + u[o + i];
                for (int q = 0; q < 6; q++)
                    if (s->sgs[q]) t6[q] = t6[q] + s->sgs[q][o + i];
            }
            um[(long)k * ny + j] = a / nx;
            for (int q = 0; q < 6; q++) tm[((long)q * nz + k) * ny + j] = t6[q] / nx;
        }
    /* tempmean through a round trip of temp (postp.f:189-196) */
    iwo_horfft(s, tp, -1, nz);
    iwo_horfft(s, tp, 1, nz);
    /* dudzmeancalc: dpdzyzplane (postp.f:984-1030) */
    double *umz = (double *)calloc((size_t)ny * nz, sizeof(double));
    for (int j = 0; j < ny; j++)
        for (int ks = 0; ks < s->nsubd; ks++)
            for (int k1 = 0; k1 < nl; k1++) {
                double sum = 0.0;
                for (int k2 = 0; k2 < nl; k2++) sum = sum + s->d[k1 * nl + k2] * um[(long)(ks * nl + k2) * ny + j];
                umz[(long)(ks * nl + k1) * ny + j] = sum * (2. / s->P.zh[ks]);
            }
    /* dudymeancalc (postp.f:237-296): spectral y derivative of the real
       sequence umean(.,k) (FFT991: forward normalised by 1/ny, coefficient
       m times i m beta, the Nyquist coefficient dropping out) */
    double *umy = (double *)calloc((size_t)ny * nz, sizeof(double));
    if (ny > 1) {
        double *cr = (double *)calloc((size_t)ny, sizeof(double)), *ci = (double *)calloc((size_t)ny, sizeof(double));
        for (int k = 0; k < nz; k++) {
            const double *a = um + (long)k * ny;
            for (int m = 1; m < ny / 2; m++) {
                double sr = 0., si = 0.;
                for (int q = 0; q < ny; q++) {
                    const double ang = s->pi2 * (double)((long)m * q % ny) / ny;
                    sr = sr + a[q] * cos(ang);
                    si = si - a[q] * sin(ang);
                }
                const double kb = m * s->beta;
                cr[m] = -kb * (si / ny); /* i m beta (a_m + i b_m) */
                ci[m] = kb * (sr / ny);
            }
            for (int j = 0; j < ny; j++) {
                double sum = 0.;
                for (int m = 1; m < ny / 2; m++) {
                    const double ang = s->pi2 * (double)((long)m * j % ny) / ny;
                    sum = sum + 2. * (cr[m] * cos(ang) - ci[m] * sin(ang));
                }
                umy[(long)k * ny + j] = sum;
            }
        }
        free(cr);
        free(ci);
    }
    /* perturbcalc(0): u' = u - umean (postp.f:333-345) */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) u[(long)k * nxppy + (long)j * nxpp + i] -= um[(long)k * ny + j];
    /* energycalc_postp: gradients of the perturbation fields (postp.f:471-514) */
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, nz);
    iwo_dpdz(s, u, g[2]);
    iwo_dpdz(s, v, g[5]);
    iwo_dpdz(s, w, g[8]);
    iwo_dpdz(s, tp, g[11]);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int ii = 0; ii < s->nxhp; ii++) {
                const long o = (long)k * nxppy + (long)j * nxpp + 2 * ii;
                const double xw = s->xw[ii], yw = s->yw[j];
                double *src[4] = {u, v, w, tp};
                const int dx[4] = {0, 3, 6, 9}, dy[4] = {1, 4, 7, 10};
                for (int q = 0; q < 4; q++) {
                    const double fr = src[q][o], fi = src[q][o + 1];
                    g[dx[q]][o] = -xw * fi;
                    g[dx[q]][o + 1] = xw * fr;
                    g[dy[q]][o] = -yw * fi;
                    g[dy[q]][o + 1] = yw * fr;
                }
            }
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], 1, nz);
    for (int q = 0; q < 12; q++) iwo_horfft(s, g[q], 1, nz);
    /* TKE budget sums over k >= 2, all j, i <= nx (postp.f:553-694); the
       wake-interior variants are empty: wakelengthcalc is not called
       (postp.f:39-41) and the zero wake sizes exclude every point */
    double tketot = 0., u1 = 0., u2 = 0., u3 = 0., prody = 0., prodz = 0., bflux = 0., tempvar = 0.;
    double eps = 0., epssgs = 0., chi = 0., chisgs = 0., tkemean = 0.;
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) {
                const long n = (long)k * nxppy + (long)j * nxpp + i;
                const double uu = u[n], vv = v[n], ww = w[n], tt = tp[n];
                tketot = tketot + (uu * uu + vv * vv + ww * ww);
                u1 = u1 + uu * uu;
                u2 = u2 + vv * vv;
                u3 = u3 + ww * ww;
                prody = prody - umy[(long)k * ny + j] * (uu * vv);
                prodz = prodz - umz[(long)k * ny + j] * (uu * ww);
                bflux = bflux + ww * tt;
                tempvar = tempvar + tt * tt;
                eps = eps + g[0][n] * g[0][n] + g[1][n] * g[1][n] + g[2][n] * g[2][n] + g[3][n] * g[3][n]
                      + g[4][n] * g[4][n] + g[5][n] * g[5][n] + g[6][n] * g[6][n] + g[7][n] * g[7][n]
                      + g[8][n] * g[8][n];
                const double s11 = g[0][n], s22 = g[4][n], s33 = g[8][n];
                const double s12 = 0.5 * (g[1][n] + g[3][n]), s13 = 0.5 * (g[2][n] + g[6][n]);
                const double s23 = 0.5 * (g[5][n] + g[7][n]);
                double f[6];
                for (int q = 0; q < 6; q++)
                    f[q] = (s->sgs[q] ? s->sgs[q][n] : 0.) - tm[((long)q * nz + k) * ny + j];
                /* t11 t12 t13 t22 t23 t33 */
                const double sg = f[0] * s11 + f[3] * s22 + f[5] * s33 + 2. * f[1] * s12 + 2. * f[2] * s13
                                  + 2. * f[4] * s23;
                epssgs = epssgs - sg;
                chi = chi + g[9][n] * g[9][n] + g[10][n] * g[10][n] + g[11][n] * g[11][n];
                const double th1 = s->sgs[6] ? s->sgs[6][n] : 0., th2 = s->sgs[7] ? s->sgs[7][n] : 0.;
                const double th3 = s->sgs[8] ? s->sgs[8][n] : 0.;
                chisgs = chisgs - (th1 * g[9][n] + th2 * g[10][n] + th3 * g[11][n]);
            }
            tkemean = tkemean + um[(long)k * ny + j] * um[(long)k * ny + j];
        }
    const double pr = s->P.xnu, ra = 0.; /* main.f:428; ra is 0 (output:73) */
    tketot = tketot * 0.5;
    tkemean = tkemean * 0.5 * (double)nx;
    bflux = bflux * ra * pr;
    eps = eps * pr;
    chi = chi * s->P.xkappa;
    const double r[16] = {tketot, u1, u2, u3, prody, prodz, prody + prodz, bflux, eps, epssgs, eps + epssgs,
                          chi, chisgs, chisgs + chi, tkemean, tempvar};
    if (out) memcpy(out, r, sizeof r);
    /* the mean back into u (postp.f:712-719), then to spectral space */
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) u[(long)k * nxppy + (long)j * nxpp + i] += um[(long)k * ny + j];
    for (int q = 0; q < 4; q++) iwo_horfft(s, s->F[q], -1, nz);
    free(um); free(tm); free(umz); free(umy);
}

/* ------------------------------------------------------------------ */
/* KE / PE (CalcE/integration.f:14-226; quadrature derv.f:479-512)    */
/* ------------------------------------------------------------------ */
void iwo_kepe(iwo_state *s, double *out)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz, nl = s->nzloc;
    const double span = s->P.xlen * (ny > 1 ? s->P.ylen : 1.0);
    const double rho0 = s->P.rho0, g = s->P.grav, bn = s->P.brunt;
    const double cke = rho0 / 2., cpe = 1. / 2. / rho0 * g * g / (bn * bn);
    double *f = (double *)malloc(sizeof(double) * (size_t)s->ntot);
    double *lvl = (double *)malloc(sizeof(double) * (size_t)nz);
    for (int q = 0; q < 4; q++) {
        memcpy(f, s->F[q], sizeof(double) * (size_t)s->ntot);
        iwo_horfft(s, f, 1, nz);
        const double c = q < 3 ? cke : cpe;
        for (int k = 0; k < nz; k++) { /* KEh / PEh: the horizontal mean times the span */
            double sum = 0.;
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const double v = f[(long)k * s->nxppy + (long)j * s->nxpp + i];
                    sum = sum + c * (v * v);
                }
            lvl[k] = span * sum / ((double)nx * ny);
        }
        double tot = 0.;
        for (int ks = 0; ks < s->nsubd; ks++) { /* integrcol per subdomain */
            double sum = 0.;
            for (int k = 0; k < nl; k++) sum = sum + s->wg[k] * lvl[ks * nl + k];
            tot = tot + sum * s->P.zh[ks] / 2.;
        }
        out[q] = tot;
    }
    free(f);
    free(lvl);
}
