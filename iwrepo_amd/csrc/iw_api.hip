// Host side of libiwgpu: set-up tables, factor cache, the device-resident step
// driver and the C-ABI of include/iwgpu.h.
//
// Reference: jmsegret/IWrepo, mode1_bbl.  The step order replays
// main.f:653-1066 for istart=0 and fixed dt.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "iw_internal.h"

using namespace iw;

static thread_local std::string g_err;

static int fail(const std::string &m)
{
    g_err = m;
    return -1;
}

int iw_fail(const std::string &m) { return fail(m); }

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(std::string(#x) + ": " + hipGetErrorString(e_));                   \
    } while (0)

extern "C" int iw_params_size(void) { return (int)sizeof(iw_params); }
extern "C" const char *iw_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// multistep.f:7-89 variable-step coefficients
// ---------------------------------------------------------------------------
extern "C" double iw_abvar1(double dt1, double dt2, double dt3)
{
    double dt123 = 2. * dt1 + 6. * dt2 + 3. * dt3, dt23 = dt2 + dt3;
    return 1. + dt1 * dt123 / (6. * dt2 * dt23);
}
extern "C" double iw_abvar2(double dt1, double dt2, double dt3)
{
    double dt123 = 2. * dt1 + 3. * dt2 + 3. * dt3;
    return -dt1 * dt123 / (6. * dt2 * dt3);
}
extern "C" double iw_abvar3(double dt1, double dt2, double dt3)
{
    double dt12 = 2. * dt1 + 3. * dt2, dt23 = dt2 + dt3;
    return dt1 * dt12 / (6. * dt3 * dt23);
}
extern "C" double iw_bdfvarv(double dt1, double dt2, double dt3)
{
    return 1. + dt1 / (dt1 + dt2) + dt1 / (dt1 + dt2 + dt3);
}
extern "C" double iw_bdfvar1(double dt1, double dt2, double dt3)
{
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    return -(-1. - dt1 / dt2 - dt1 / dt123 - dt1 * dt1 / (dt123 * dt2)
             - dt1 * dt1 * dt12 / (dt123 * dt23 * dt2));
}
extern "C" double iw_bdfvar2(double dt1, double dt2, double dt3)
{
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    return -(dt1 * dt1 / (dt12 * dt2) + dt1 * dt1 / (dt2 * dt123)
             + dt1 * dt1 * dt12 / (dt2 * dt23 * dt123) + dt1 * dt1 * dt12 / (dt3 * dt23 * dt123));
}
extern "C" double iw_bdfvar3(double dt1, double dt2, double dt3)
{
    double dt12 = dt1 + dt2, dt23 = dt2 + dt3, dt123 = dt1 + dt2 + dt3;
    return dt1 * dt1 * dt12 / (dt3 * dt23 * dt123);
}

// ---------------------------------------------------------------------------
// GLL tables (setup.f:37-42): Gauss-Lobatto points by the Jacobi-polynomial
// Newton iteration with deflation of jacobl (setup.f:395-479), weights and
// Legendre table of quad (setup.f:527-600), Costa-Don D and D2 of derv
// (setup.f:636-798).
// ---------------------------------------------------------------------------
struct Jac {
    double poly, pder, polym1, pderm1, polym2, pderm2;
};

// Jacobi recurrence for alpha = beta = 0 (jacobf, setup.f:486-518)
static Jac jacobf(int n, double x)
{
    const double alp = 0., bet = 0., apb = 0., rv = 1.;
    Jac r = {1.0, 0.0, 0., 0., 0., 0.};
    if (n == 0) return r;
    double plst = r.poly, pdlst = r.pder, psave = 0., pdsave = 0.;
    double p = rv * x, pd = rv;
    if (n == 1) {
        r.poly = p;
        r.pder = pd;
        return r;
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
    r.poly = p; r.pder = pd;
    r.polym1 = plst; r.pderm1 = pdlst;
    r.polym2 = psave; r.pderm2 = pdsave;
    return r;
}

// Gauss-Lobatto points from the largest down by Newton with deflation,
// tolerance 1e-12 and at most 10 steps (jacobl, setup.f:395-479)
static std::vector<double> gll_points(int n)
{
    const int np = n + 1;
    std::vector<double> xj(np + 2, 0.0), out(n + 1);
    Jac jp = jacobf(np, 1.0), jm = jacobf(np, -1.0);
    double det = jp.polym1 * jm.polym2 - jm.polym1 * jp.polym2;
    double rp = -jp.poly, rm = -jm.poly;
    double a = (rp * jm.polym2 - rm * jp.polym2) / det;
    double b = (rm * jp.polym1 - rp * jm.polym1) / det;
    xj[1] = 1.0;
    int nh = (n + 1) / 2;
    double dth = 3.14159265 / (2 * n + 1); // the reference's truncated pi
    double cd = cos(2. * dth), sd = sin(2. * dth), cs = cos(dth), ss = sin(dth);
    for (int j = 2; j <= nh; j++) {
        double x = cs;
        for (int k = 1; k <= 10; k++) {
            Jac q = jacobf(np, x);
            double poly = q.poly + a * q.polym1 + b * q.polym2;
            double pder = q.pder + a * q.pderm1 + b * q.pderm2;
            double recsum = 0.0;
            for (int i = 1; i <= j - 1; i++) recsum += 1.0 / (x - xj[i]);
            double delx = -poly / (pder - recsum * poly);
            x = x + delx;
            if (fabs(delx) < 1.0e-12) break;
        }
        xj[j] = x;
        double cssave = cs * cd - ss * sd;
        ss = cs * sd + ss * cd;
        cs = cssave;
    }
    xj[np] = -1.0;
    for (int i = 2; i <= nh; i++) xj[n + 2 - i] = -xj[i];
    if (n % 2 == 0) xj[nh + 1] = 0.0;
    for (int k = 0; k <= n; k++) out[k] = xj[n - k + 1];
    return out;
}

static void setup_tables(iw_solver *s)
{
    const int nl = s->P.nzloc, N = nl - 1;
    s->zpts = gll_points(N);
    s->wg.assign(nl, 0.);
    s->alpol.assign((size_t)nl * nl, 0.);
    std::vector<double> al(nl + 2), alp(nl + 2);
    for (int k = 0; k <= N; k++) {
        double xc = s->zpts[k];
        al[0] = 1.;
        al[1] = xc;
        alp[0] = 0.;
        alp[1] = 1.;
        for (int m = 1; m <= N; m++) {
            al[m + 1] = (2. * m + 1.) * xc * al[m] / (m + 1) - m * al[m - 1] / (m + 1);
            alp[m + 1] = (2. * m + 1.) * (al[m] + xc * alp[m]) / (m + 1) - m * alp[m - 1] / (m + 1);
        }
        s->wg[k] = 2. / (N * (N + 1) * al[N] * al[N] + 1.0e-30);
        for (int j = 0; j <= N; j++) s->alpol[j * nl + k] = al[j];
    }
    std::vector<double> c(nl);
    for (int k = 0; k <= N; k++) {
        double prod = 1.;
        for (int l = 0; l <= N; l++)
            if (l != k) prod = prod * (s->zpts[k] - s->zpts[l]);
        c[k] = prod;
    }
    s->d.assign((size_t)nl * nl, 0.);
    s->d2.assign((size_t)nl * nl, 0.);
    for (int k = 0; k <= N; k++)
        for (int j = 0; j <= N; j++)
            if (k != j) s->d[k * nl + j] = c[k] / (c[j] * (s->zpts[k] - s->zpts[j]));
    for (int k = 0; k <= N; k++) {
        double sum = 0.;
        for (int j = 0; j <= N; j++)
            if (k != j) sum += s->d[k * nl + j];
        s->d[k * nl + k] = -sum;
    }
    for (int k = 0; k <= N; k++)
        for (int j = 0; j <= N; j++)
            if (k != j)
                s->d2[k * nl + j] = 2.0 * (s->d[k * nl + k] * s->d[k * nl + j]
                                           - s->d[k * nl + j] / (s->zpts[k] - s->zpts[j]));
    for (int k = 0; k <= N; k++) {
        double sum = 0.;
        for (int j = 0; j <= N; j++)
            if (k != j) sum += s->d2[k * nl + j];
        s->d2[k * nl + k] = -sum;
    }
    // wavenumbers (setup.f:77-89) and grid (setup.f:126-160)
    const double pi2 = 8. * atan(1.0);
    const double alpha = pi2 / s->P.xlen, beta = pi2 / s->P.ylen;
    const int nxhp = s->P.nx / 2 + 1, ny = s->P.ny, nyh = ny / 2;
    s->xw.assign(nxhp, 0.);
    s->xsq.assign(nxhp, 0.);
    s->yw.assign(ny, 0.);
    s->ysq.assign(ny, 0.);
    for (int i = 0; i < nxhp; i++) {
        s->xw[i] = i * alpha;
        s->xsq[i] = s->xw[i] * s->xw[i];
    }
    for (int j = 1; j <= nyh; j++) {
        s->yw[j - 1] = (j - 1) * beta;
        s->yw[ny - j] = -j * beta;
    }
    for (int j = 0; j < ny; j++) s->ysq[j] = s->yw[j] * s->yw[j];
    s->z.assign((size_t)nl * s->P.nsubd, 0.);
    for (int ks = 0; ks < s->P.nsubd; ks++)
        for (int k = 0; k < nl; k++) s->z[ks * nl + k] = s->P.z0[ks] + 0.5 * s->P.zh[ks] * (s->zpts[k] + 1.0);
}

// twiddle e^{-2 pi i k/n} in extended precision, rounded once
static double2 twiddle(long k, long n)
{
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    long double cr = cosl(a), sr = sinl(a);
    if (4 * k == n) { cr = 0; sr = -1; }
    if (2 * k == n) { cr = -1; sr = 0; }
    if (4 * k == 3 * n) { cr = 0; sr = 1; }
    return make_double2((double)cr, (double)sr);
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src, long long &acc)
{
    size_t b = sizeof(T) * (src.size() ? src.size() : 1);
    HIPCHK(hipMalloc((void **)dst, b));
    acc += (long long)b;
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return 0;
}

static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return (1 << l) == v ? l : -1;
}

// The passes of a transform of length n = 2^a 3^b 5^c (the factors FFTFAX /
// CFTFAX accept, fft.f:407-500, 1503-1622): radix 8 while 8 divides what
// remains, then 4 or 2, then the fives and threes; 4 bits per pass, first
// pass lowest.  0 when n has another prime factor (or n = 1).
static unsigned fft_plan(int n)
{
    unsigned plan = 0;
    int shift = 0;
    auto take = [&](int r) {
        plan |= (unsigned)r << shift;
        shift += 4;
        n /= r;
    };
    if (n < 2) return 0;
    while (n % 8 == 0) take(8);
    if (n % 4 == 0) take(4);
    if (n % 2 == 0) take(2);
    while (n % 5 == 0) take(5);
    while (n % 3 == 0) take(3);
    return (n == 1 && shift <= 32) ? plan : 0;
}

// ---------------------------------------------------------------------------
// penalty parameters (penalty.f:9-157) and per-subdomain assembly
// coefficients of matrix_elmsetup (penalty.f:403-634)
// ---------------------------------------------------------------------------
enum { BC_DIR = 0, BC_NEU = 1, BC_ROB = 2 };

struct Pen {
    std::vector<double> alpha, beta, gamma, delta, tau1, tau2;
};

static void param_set(const iw_params &P, double omega, int ifbc, double eps, double ybar, double &a0,
                      double &b0, double &t0)
{
    if (ifbc == BC_DIR) {
        a0 = 1.;
        b0 = 0.;
        double fac = P.facdir * ((2. / ybar) * (2. / ybar));
        t0 = fac * eps / (a0 * (omega * omega));
    } else if (ifbc == BC_NEU) {
        a0 = 0.;
        b0 = 1.;
        double fac = P.facneumn * 2. / ybar;
        t0 = fac / (b0 * omega);
    } else {
        a0 = 1.;
        b0 = 1.;
        double ck = a0 * omega / b0;
        double dn = 1. / (omega * eps * b0);
        double t1 = sqrt(ck * ck + eps * ck);
        double t2 = 2. * ck - 2. * t1;
        double facm = ((1. / eps) >= 1.e5) ? P.facrobin : 1.;
        t0 = facm * dn * (eps + t2) * 2. / ybar;
    }
}

static Pen penalty_params(const iw_params &P, double eps, int bcbot, int bctop)
{
    const int ns = P.nsubd;
    const double omega = 2. / ((P.nzloc + 1) * (P.nzloc + 2));
    Pen p;
    p.alpha.assign(ns, 0.);
    p.beta = p.gamma = p.delta = p.tau1 = p.tau2 = p.alpha;
    double a0, b0, t0;
    for (int ks = 1; ks < ns - 1; ks++) {
        param_set(P, omega, BC_ROB, eps, P.zh[ks], a0, b0, t0);
        p.alpha[ks] = a0; p.beta[ks] = b0; p.tau1[ks] = t0;
        p.gamma[ks] = a0; p.delta[ks] = b0; p.tau2[ks] = t0;
    }
    param_set(P, omega, bcbot, eps, P.zh[0], a0, b0, t0);
    p.alpha[0] = a0; p.beta[0] = b0; p.tau1[0] = t0;
    if (ns > 1) {
        param_set(P, omega, BC_ROB, eps, P.zh[0], a0, b0, t0);
        p.gamma[0] = a0; p.delta[0] = b0; p.tau2[0] = t0;
    }
    param_set(P, omega, bctop, eps, P.zh[ns - 1], a0, b0, t0);
    p.gamma[ns - 1] = a0; p.delta[ns - 1] = b0; p.tau2[ns - 1] = t0;
    if (ns > 1) {
        param_set(P, omega, BC_ROB, eps, P.zh[ns - 1], a0, b0, t0);
        p.alpha[ns - 1] = a0; p.beta[ns - 1] = b0; p.tau1[ns - 1] = t0;
    }
    return p;
}

static void assembly_coef(const iw_params &P, const Pen &p, double eps, double *q /*12*ns*/)
{
    const int ns = P.nsubd;
    for (int ks = 0; ks < ns; ks++) {
        double ybar = P.zh[ks];
        double ybarl = ks > 0 ? P.zh[ks - 1] : 1., ybarr = ks < ns - 1 ? P.zh[ks + 1] : 1.;
        double *c = q + 12 * ks;
        double c2 = (2. / ybar) * (2. / ybar);
        c[0] = eps * c2;
        c[1] = p.tau1[ks] * p.alpha[ks];
        c[2] = -p.tau1[ks] * p.beta[ks] * eps * 2. / ybar;
        c[3] = p.tau2[ks] * p.gamma[ks];
        c[4] = +p.tau2[ks] * p.delta[ks] * eps * 2. / ybar;
        c[5] = 0.5 * p.tau1[ks] * p.alpha[ks];
        c[6] = +0.5 * p.tau1[ks] * p.beta[ks] * eps * 2. / ybar;
        c[7] = 0.5 * p.tau2[ks] * p.gamma[ks];
        c[8] = -0.5 * p.tau2[ks] * p.delta[ks] * eps * 2. / ybar;
        c[9] = +0.5 * p.tau1[ks] * p.beta[ks] * eps * 2. / ybarl;
        c[10] = -0.5 * p.tau2[ks] * p.delta[ks] * eps * 2. / ybarr;
        c[11] = 0.;
    }
}

// Build the system table for the current (bdfv, dt): every mode needs a
// pressure system (eps = 1/k^2, Neumann/Neumann, calcuvwt.f:114-152), a u/v
// system (eps = 1/(k^2 + bdfv/(nu dt)), calcuvwt.f:276-295), a w system
// (Dirichlet/Dirichlet, calcuvwt.f:516-532) and a density system
// (eps = 1/(k^2 + bdfv/(kappa dt)), calcuvwt.f:634-653).  The matrices depend
// only on (eps, bottom BC, top BC), so identical triples (same k^2 bits)
// share one factor.  Returns the first system index that needs factoring.
static int build_systems(iw_solver *s, bool refactor_p, int &first, int &count)
{
    const iw_params &P = s->P;
    const int nxhp = s->g.nxhp, ny = s->g.ny, nmode = nxhp * ny, ns = P.nsubd;
    const double facvel = 1. / (P.xnu * P.dt), fact = 1. / (P.xkappa * P.dt);
    // the tables below are rewritten by synchronous copies on the null
    // stream, which does not wait for the solver's non-blocking stream: let
    // the sweeps of the previous step (readers of sysinfo, modesys and the
    // factors) finish first
    HIPCHK(hipStreamSynchronize(s->stream));
    const int ubot = P.bc_uv_bottom == 1 ? BC_DIR : BC_NEU;
    typedef std::tuple<unsigned long long, int, int> Key;
    std::map<Key, int> idx;
    std::vector<double> eps_of;
    std::vector<int> bb, bt;
    std::vector<int> modesys(4 * (size_t)nmode, -1);
    auto key_of = [](double e, int b, int t) {
        unsigned long long u;
        memcpy(&u, &e, 8);
        return Key(u, b, t);
    };
    auto get = [&](double e, int b, int t) {
        Key k = key_of(e, b, t);
        auto it = idx.find(k);
        if (it != idx.end()) return it->second;
        int id = (int)eps_of.size();
        idx[k] = id;
        eps_of.push_back(e);
        bb.push_back(b);
        bt.push_back(t);
        return id;
    };
    // pressure systems first so that they keep their slots across refactors
    for (int m = 0; m < nmode; m++) {
        int ii = m % nxhp, j = m / nxhp;
        if (m == 0) continue;
        double ak = s->xsq[ii] + s->ysq[j];
        modesys[4 * m] = get(1. / ak, BC_NEU, BC_NEU);
    }
    int np = (int)eps_of.size();
    for (int m = 0; m < nmode; m++) {
        int ii = m % nxhp, j = m / nxhp;
        double ak = s->xsq[ii] + s->ysq[j];
        double ev = 1. / (ak + s->bdfv * facvel);
        double et = 1. / (ak + s->bdfv * fact);
        modesys[4 * m + 1] = get(ev, ubot, BC_NEU);
        if (m != 0) modesys[4 * m + 2] = get(ev, BC_DIR, BC_DIR);
        modesys[4 * m + 3] = get(et, BC_NEU, BC_NEU);
    }
    const int nsys = (int)eps_of.size();
    SysTables &S = s->sys;
    S.nsys_w = 0; // Dirichlet at both ends: the w systems
    for (int q = np; q < nsys; q++) S.nsys_w += (bb[q] == BC_DIR && bt[q] == BC_DIR) ? 1 : 0;
    if ((size_t)nsys > S.cap_sys) {
        if (S.AB) hipFree(S.AB);
        if (S.FB) hipFree(S.FB);
        if (S.SPK) hipFree(S.SPK);
        if (S.SPB) hipFree(S.SPB);
        if (S.ipiv) hipFree(S.ipiv);
        if (S.coef) hipFree(S.coef);
        if (S.sysinfo) hipFree(S.sysinfo);
        size_t cap = (size_t)nsys;
        HIPCHK(hipMalloc((void **)&S.AB, sizeof(double) * cap * S.ldab * s->g.nz));
        // band positions outside the matrix are never written by the
        // factorisation and must read as zero in the lane solvers
        HIPCHK(hipMemsetAsync(S.AB, 0, sizeof(double) * cap * S.ldab * s->g.nz, s->stream));
        const size_t fbn = (size_t)(tri_sf(P.nzloc) + tri_sb(P.nzloc, false)) * ns; // double2 per system
        HIPCHK(hipMalloc((void **)&S.FB, sizeof(double) * 2 * cap * fbn));
        s->dev_bytes += (long long)(sizeof(double) * 2 * cap * fbn);
        const size_t spn = (size_t)spk_stride(P.nzloc) * ns; // doubles per system
        HIPCHK(hipMalloc((void **)&S.SPK, sizeof(double) * cap * spn));
        HIPCHK(hipMalloc((void **)&S.SPB, sizeof(double) * cap * 4 * ns));
        s->dev_bytes += (long long)(sizeof(double) * cap * (spn + 4 * ns));
        HIPCHK(hipMalloc((void **)&S.ipiv, sizeof(int) * cap * s->g.nz));
        HIPCHK(hipMalloc((void **)&S.coef, sizeof(double) * cap * 12 * ns));
        HIPCHK(hipMalloc((void **)&S.sysinfo, sizeof(double) * cap * 4));
        s->dev_bytes += (long long)(sizeof(double) * cap * S.ldab * s->g.nz + sizeof(int) * cap * s->g.nz
                                    + sizeof(double) * cap * 12 * ns + sizeof(double) * cap * 4);
        S.cap_sys = cap;
        refactor_p = true;
    }
    {   // packed pressure factors
        const int npc = std::max(np, 1);
        if ((size_t)npc > S.cap_p) {
            if (S.PB) hipFree(S.PB);
            const size_t pbn = (size_t)(tri_sf(P.nzloc) + tri_sb(P.nzloc, true)) * ns; // double2 per system
            HIPCHK(hipMalloc((void **)&S.PB, sizeof(double) * 2 * npc * pbn));
            s->dev_bytes += (long long)(sizeof(double) * 2 * npc * pbn);
            S.cap_p = npc;
            refactor_p = true;
        }
    }
    std::vector<double> coef((size_t)nsys * 12 * ns), info((size_t)nsys * 4);
    for (int q = 0; q < nsys; q++) {
        Pen p = penalty_params(P, eps_of[q], bb[q], bt[q]);
        assembly_coef(P, p, eps_of[q], &coef[(size_t)q * 12 * ns]);
        info[4 * q] = eps_of[q];
        info[4 * q + 1] = p.tau1[0];
        info[4 * q + 2] = p.tau2[ns - 1];
        info[4 * q + 3] = bb[q] * 4 + bt[q];
    }
    HIPCHK(hipMemcpy(S.coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(S.sysinfo, info.data(), sizeof(double) * info.size(), hipMemcpyHostToDevice));
    if (!S.modesys) {
        HIPCHK(hipMalloc((void **)&S.modesys, sizeof(int) * 4 * (size_t)nmode));
        s->dev_bytes += (long long)(sizeof(int) * 4 * (size_t)nmode);
    }
    HIPCHK(hipMemcpy(S.modesys, modesys.data(), sizeof(int) * modesys.size(), hipMemcpyHostToDevice));
    // XCD-aware task order: group tasks by the system they solve, deal the
    // groups round-robin to the 8 XCDs, and interleave the per-XCD queues so
    // that block b runs on XCD b % 8 (the hardware's round-robin dispatch)
    auto make_order = [&](const std::vector<std::pair<int, int>> &tasks, std::vector<int> &order) {
        std::vector<std::pair<int, int>> t = tasks; // (system, task)
        std::stable_sort(t.begin(), t.end());
        const int nx = 8;
        std::vector<std::vector<int>> q(nx);
        int g = -1, lastsys = -2;
        for (auto &e : t) {
            if (e.first != lastsys) {
                lastsys = e.first;
                g++;
            }
            q[g % nx].push_back(e.second);
        }
        size_t len = 0;
        for (auto &v : q) len = std::max(len, v.size());
        order.assign(len * nx, -1);
        for (int x = 0; x < nx; x++)
            for (size_t i = 0; i < q[x].size(); i++) order[i * nx + x] = q[x][i];
    };
    std::vector<std::pair<int, int>> tp, tv;
    for (int m = 0; m < nmode; m++) {
        tp.push_back({m == 0 ? -1 : modesys[4 * m], m});
        tv.push_back({modesys[4 * m + 1], 4 * m});
        tv.push_back({modesys[4 * m + 1], 4 * m + 1});
        if (m != 0) tv.push_back({modesys[4 * m + 2], 4 * m + 2});
        tv.push_back({modesys[4 * m + 3], 4 * m + 3});
    }
    std::vector<int> op, ov;
    make_order(tp, op);
    make_order(tv, ov);
    if (S.order_p) hipFree(S.order_p);
    if (S.order_v) hipFree(S.order_v);
    HIPCHK(hipMalloc((void **)&S.order_p, sizeof(int) * op.size()));
    HIPCHK(hipMalloc((void **)&S.order_v, sizeof(int) * ov.size()));
    HIPCHK(hipMemcpy(S.order_p, op.data(), sizeof(int) * op.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(S.order_v, ov.data(), sizeof(int) * ov.size(), hipMemcpyHostToDevice));
    S.nblk_p = (int)op.size();
    S.nblk_v = (int)ov.size();
    S.nsys = nsys;
    S.np = np;
    first = refactor_p ? 0 : np;
    count = nsys - first;
    return 0;
}

// Place the 64-lane blocks of a lane table on the XCDs: XCD x gets the
// x-th eighth of the table, in order (block b of the grid runs on XCD b mod 8,
// the hardware's round-robin dispatch).  Consecutive blocks hold consecutive
// kx of the same ky rows, so the cache lines of a state row that two blocks
// share are fetched into one XCD's L2 by blocks that run at the same time;
// dealing blocks to XCDs by the systems they share instead (modes with equal
// k^2) splits those lines between XCDs and measured slower.  Padding blocks
// are -1.
static std::vector<int> xcd_place(const std::vector<int> &tasks)
{
    const int nx = 8, nblk = (int)(tasks.size() / 64);
    const int cap = (nblk + nx - 1) / nx;
    std::vector<int> out((size_t)cap * nx * 64, -1);
    for (int b = 0; b < nblk; b++) {
        const int x = b / cap, i = b % cap;
        std::copy(tasks.begin() + (size_t)b * 64, tasks.begin() + (size_t)b * 64 + 64,
                  out.begin() + ((size_t)i * nx + x) * 64);
    }
    return out;
}

static int ensure_factors(iw_solver *s)
{
    if (s->factored_bdfv == s->bdfv && s->factored_dt == s->P.dt && s->p_factored) return 0;
    int first, count;
    if (build_systems(s, !s->p_factored, first, count)) return -1;
    // [0], [1]: pattern flags of the pressure and the other factors; [2]:
    // the u/v and w factor blocks below the top subdomain differ
    if (!s->sys.tri_bad) HIPCHK(hipMalloc((void **)&s->sys.tri_bad, 3 * sizeof(int)));
    // pattern flags of this packing: [0] pressure factors (only when they are
    // packed again), [1] the u/v, w and density factors
    launch_zero_int(first == 0 ? s->sys.tri_bad : s->sys.tri_bad + 1, first == 0 ? 2 : 1, s->stream);
    launch_factor(s->g, s->sys, first, count, s->stream);
    launch_compact(s->g, s->sys, first, count, s->stream);
    launch_spike_pre(s->g, s->sys, first, count, s->stream);
    // u, v and w below the top subdomain over one factor (no-slip bottom:
    // the u/v and w systems differ only in the top boundary condition)
    const bool uvw = s->P.bc_uv_bottom == 1 && s->P.nsubd > 1;
    if (uvw) {
        launch_zero_int(s->sys.tri_bad + 2, 1, s->stream);
        launch_share_check(s->g, s->sys, s->sys.tri_bad + 2, s->stream);
    }
    HIPCHK(hipGetLastError());
    // the u/v, w and density systems go to the packed-factor sweeps (lane =
    // one column); if any of them pivoted, or a factor breaks the packed
    // pattern, all of them take the wavefront path
    SysTables &S = s->sys;
    const int nmode = s->g.nxhp * s->g.ny;
    std::vector<double> info((size_t)S.nsys * 4);
    HIPCHK(hipMemcpyAsync(info.data(), S.sysinfo, sizeof(double) * info.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    int tri_bad[3] = {0, 0, 1};
    HIPCHK(hipMemcpy(tri_bad, S.tri_bad, (uvw ? 3 : 2) * sizeof(int), hipMemcpyDeviceToHost));
    // the packed sweeps need every factor of their kind in the tri_*
    // pattern; the band solvers on AB take over otherwise, for the pressure
    // and the other systems separately
    S.tri_ok_p = tri_bad[0] == 0;
    S.tri_ok_v = tri_bad[1] == 0;
    bool ok = solve_np_supported(s->P.nzloc, s->P.nsubd) && S.tri_ok_v;
    for (int q = S.np; q < S.nsys && ok; q++)
        if (info[4 * q + 3] != 0.) ok = false;
    S.np_ok = ok;
    if (solve_np_supported(s->P.nzloc, s->P.nsubd) && !S.gwt_tasks) {
        const Geo &g = s->g;
        // rows ky = 0 and ky = -ny/2 pair with themselves
        std::vector<int> single = {0};
        if (g.ny > 1) single.push_back(g.ny / 2);
        // grouped layouts: G lanes per system, groups of one kx on the rows
        // +ky and -ky (u and v together for G = 8); rows without a partner
        // fill their group with idle lanes
        auto groups = [&](int G, bool both, std::vector<int> &out) -> int {
            auto ent = [&](int jy, int ii, std::vector<int> &v) {
                const int m = jy * g.nxhp + ii;
                for (int f = 0; f < (both ? 2 : 1); f++)
                    for (int c = 0; c < 2; c++) v.push_back((2 * m + c) | (f << 28));
            };
            auto add = [&](std::vector<int> &v) {
                const int first = v[0];
                while ((int)v.size() < G) v.push_back(first | (1 << 29));
                out.insert(out.end(), v.begin(), v.end());
            };
            for (int p = 1; p < (g.ny + 1) / 2; p++) {
                const int q = g.ny - p;
                if (q == p) continue;
                for (int ii = 0; ii < g.nxh; ii++) {
                    std::vector<int> v;
                    ent(p, ii, v);
                    ent(q, ii, v);
                    add(v);
                }
            }
            for (int jy : single)
                for (int ii = 0; ii < g.nxh; ii++) {
                    std::vector<int> v;
                    ent(jy, ii, v);
                    add(v);
                }
            for (int jy : single) {
                std::vector<int> v;
                ent(jy, g.nxh, v);
                add(v);
            }
            for (int p = 1; p < (g.ny + 1) / 2; p++) {
                const int q = g.ny - p;
                if (q == p) continue;
                std::vector<int> v;
                ent(p, g.nxh, v);
                ent(q, g.nxh, v);
                add(v);
            }
            long live = 0;
            for (int e : out) live += ((e >> 29) & 1) ? 0 : 1;
            if (live != (both ? 2 : 1) * g.nxppy) return -1;
            while (out.size() % 64) out.push_back(-1);
            return 0;
        };
        std::vector<int> guv, gwt;
        if (groups(8, true, guv) || groups(4, false, gwt))
            return fail("internal: grouped lane tables do not cover the plane");
        if (uvw) {
            // groups of 12 lanes: u, v of +ky and -ky, then w of +ky and
            // -ky (re, im each), five groups per 64 lanes; field in bits
            // 28-29, idle in bit 30; w(0,0) is not solved (idle)
            std::vector<int> out, blk;
            auto entf = [&](int jy, int ii, int f, std::vector<int> &v) {
                const int m = jy * g.nxhp + ii;
                for (int c = 0; c < 2; c++)
                    v.push_back((f == 2 && m == 0) ? ((2 * m + c) | (1 << 30)) : ((2 * m + c) | (f << 28)));
            };
            auto add = [&](std::vector<int> &v) {
                const int first = v[0] & ((1 << 28) - 1);
                while ((int)v.size() < 12) v.push_back(first | (1 << 30));
                if (blk.size() + 12 > 64) {
                    while (blk.size() < 64) blk.push_back(-1);
                    out.insert(out.end(), blk.begin(), blk.end());
                    blk.clear();
                }
                blk.insert(blk.end(), v.begin(), v.end());
            };
            auto pair = [&](int p, int q, int ii) {
                std::vector<int> v;
                entf(p, ii, 0, v);
                entf(p, ii, 1, v);
                if (q >= 0) {
                    entf(q, ii, 0, v);
                    entf(q, ii, 1, v);
                }
                entf(p, ii, 2, v);
                if (q >= 0) entf(q, ii, 2, v);
                add(v);
            };
            for (int p = 1; p < (g.ny + 1) / 2; p++)
                if (g.ny - p != p)
                    for (int ii = 0; ii < g.nxh; ii++) pair(p, g.ny - p, ii);
            for (int jy : single)
                for (int ii = 0; ii < g.nxh; ii++) pair(jy, -1, ii);
            for (int jy : single) pair(jy, -1, g.nxh);
            for (int p = 1; p < (g.ny + 1) / 2; p++)
                if (g.ny - p != p) pair(p, g.ny - p, g.nxh);
            if (!blk.empty()) {
                while (blk.size() < 64) blk.push_back(-1);
                out.insert(out.end(), blk.begin(), blk.end());
            }
            long live = 0;
            for (int e : out) live += (e >= 0 && !((e >> 30) & 1)) ? 1 : 0;
            if (live != 3 * g.nxppy - 2) return fail("internal: u/v/w lane table does not cover the plane");
            out = xcd_place(out);
            HIPCHK(hipMalloc((void **)&S.guvw_tasks, sizeof(int) * out.size()));
            HIPCHK(hipMemcpy(S.guvw_tasks, out.data(), sizeof(int) * out.size(), hipMemcpyHostToDevice));
            S.guvw_ntasks = (int)out.size();
        }
        guv = xcd_place(guv);
        gwt = xcd_place(gwt);
        HIPCHK(hipMalloc((void **)&S.guv_tasks, sizeof(int) * guv.size()));
        HIPCHK(hipMemcpy(S.guv_tasks, guv.data(), sizeof(int) * guv.size(), hipMemcpyHostToDevice));
        S.guv_ntasks = (int)guv.size();
        HIPCHK(hipMalloc((void **)&S.gwt_tasks, sizeof(int) * gwt.size()));
        HIPCHK(hipMemcpy(S.gwt_tasks, gwt.data(), sizeof(int) * gwt.size(), hipMemcpyHostToDevice));
        S.gwt_ntasks = (int)gwt.size();
        // coupling scalars of the subdomain-parallel solve: 4 fields x nsubd x columns, twice
        const size_t nal = (size_t)4 * g.nsubd * g.nxppy;
        HIPCHK(hipMalloc((void **)&S.spk_al, sizeof(double) * nal));
        HIPCHK(hipMalloc((void **)&S.spk_au, sizeof(double) * nal));
        s->dev_bytes += (long long)(2 * sizeof(double) * nal);
    }
    S.share_uvw = uvw && tri_bad[2] == 0 && S.np_ok && S.guvw_tasks;
    s->factored_bdfv = s->bdfv;
    s->factored_dt = s->P.dt;
    s->p_factored = true;
    return 0;
}

// ---------------------------------------------------------------------------
// profiling: event pairs per kernel class, resolved after the run
// ---------------------------------------------------------------------------
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
    double bytes;
};
static thread_local std::vector<ProfRec> g_prec;
static thread_local std::vector<hipEvent_t> g_pool;
static thread_local size_t g_pool_next = 0;

static hipEvent_t pool_event()
{
    if (g_pool_next == g_pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        g_pool.push_back(e);
    }
    return g_pool[g_pool_next++];
}

struct Scope {
    iw_solver *s;
    const char *name;
    hipEvent_t a;
    double bytes;
    Scope(iw_solver *s_, const char *n, double b = 0.) : s(s_), name(n), a(nullptr), bytes(b)
    {
        if (s->profiling) {
            a = pool_event();
            hipEventRecord(a, s->stream);
        }
    }
    ~Scope()
    {
        if (s->profiling) {
            hipEvent_t b = pool_event();
            hipEventRecord(b, s->stream);
            g_prec.push_back({name, a, b, bytes});
        }
    }
};

static void prof_resolve(iw_solver *s)
{
    if (g_prec.empty()) return;
    hipStreamSynchronize(s->stream);
    for (auto &r : g_prec) {
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        size_t k = 0;
        for (; k < s->prof.names.size(); k++)
            if (s->prof.names[k] == r.name) break;
        if (k == s->prof.names.size()) {
            s->prof.names.push_back(r.name);
            s->prof.ms.push_back(0.);
            s->prof.n.push_back(0);
            s->prof.bytes.push_back(0.);
        }
        s->prof.ms[k] += ms;
        s->prof.n[k] += 1;
        s->prof.bytes[k] += r.bytes;
    }
    g_prec.clear();
    g_pool_next = 0;
}

// ---------------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------------
static int scratch_init(int dev);

extern "C" int iw_create(const iw_params *p, iw_solver **out)
{
    if (!p || !out) return fail("null argument");
    const iw_params &P = *p;
    if (P.nsubd < 1 || P.nsubd > IW_MAXSUB) return fail("nsubd out of range");
    if (P.nzloc < 3 || P.nzloc > kMaxNl - 1) return fail("nzloc must lie in [3, 31]");
    // FFT991 / CFFT99 lengths: nx and ny even and free of prime factors
    // above 5 (fft.f:532 rejects the others); ny = 1 is the 2-D case
    const unsigned planM = (P.nx % 2 == 0) ? fft_plan(P.nx / 2) : 0;
    const unsigned planY = (P.ny % 2 == 0) ? fft_plan(P.ny) : 0;
    if (P.nx < 4 || P.nx > 1024 || !planM) return fail("nx must be even, 2^a 3^b 5^c, in [4, 1024]");
    if (P.ny < 1 || P.ny > 1024 || (P.ny > 1 && !planY)) return fail("ny must be 1 or even, 2^a 3^b 5^c, <= 1024");
    const int lx = ilog2(P.nx), ly = ilog2(P.ny);
    if (P.dt <= 0) return fail("dt must be positive");
    if (P.ncalclegspec > 0 && P.ny < 2)
        return fail("ncalclegspec needs ny >= 2 (speclegencalc reads the row j = ny/2)");
    // the estimation filter runs inside iw_run every nstepfilter steps: an
    // order it lacks must fail here, not after steps of the run have advanced
    if (P.ifilter == 1 && !solve_np_supported(P.nzloc, P.nsubd))
        return fail("ifilter = 1: estimation filter GLL order must lie in 8..24");
    iw_solver *s = new iw_solver();
    s->P = P;
    {
        const char *e = getenv("IW_SWEEPS");
        s->seq_sweeps = e && !strcmp(e, "seq");
        const char *m = getenv("IW_MIXED_MFMA");
        s->mixed_mfma = m && !strcmp(m, "1");
    }
    Geo &g = s->g;
    g.nx = P.nx;
    g.ny = P.ny;
    g.nzloc = P.nzloc;
    g.nsubd = P.nsubd;
    g.nz = P.nzloc * P.nsubd;
    g.nxpp = P.nx + 2;
    g.nxh = P.nx / 2;
    g.nxhp = g.nxh + 1;
    g.nyh = P.ny / 2;
    g.ndims = P.ny > 1 ? 3 : 2;
    g.nxppy = (long)g.nxpp * g.ny;
    g.ntot = g.nxppy * g.nz;
    g.M = P.nx / 2;
    g.p2 = lx > 0 && ly >= 0;
    g.logM = g.p2 ? lx - 1 : 0;
    g.logNy = g.p2 ? ly : 0;
    g.planM = planM;
    g.planY = planY;
    g.scaley = 1. / (double)P.ny;
    HIPCHK(hipGetDevice(&s->device));
    if (scratch_init(s->device)) {
        delete s;
        return -1;
    }
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    setup_tables(s);
    long long &acc = s->dev_bytes;
    // filter tables: Legendre sigma (filtering.f:1225-1243), Fourier factor
    // per mode (filtering.f:849, 868-901), mean-u damping (calcuvwt.f:343-361)
    const int nzm = P.nzloc - 1;
    const int kc = (int)(P.fackc * nzm);
    std::vector<double> sig(P.nzloc);
    for (int k = 0; k <= nzm; k++)
        sig[k] = (k >= kc) ? exp(-P.amp * pow((double)(k - kc) / (double)(nzm - kc), 2. * P.p)) : 1.;
    // the same filter as one matrix: out[q] = sum_t fl[t nl + q] x[t] with
    // fl = w_t sum_k P_k(x_t) c_k sig_k P_k(x_q), c_k = k + 1/2 (nzm/2 for
    // k = nzm), the two sums of filtering.f:912-955 taken together
    const int nl = P.nzloc;
    std::vector<double> fl((size_t)nl * nl, 0.);
    for (int t = 0; t < nl; t++)
        for (int q = 0; q < nl; q++) {
            double sum = 0.;
            for (int k = 0; k < nl; k++) {
                const double c = (k == nzm) ? 0.5 * nzm : (k + 0.5);
                sum += s->alpol[k * nl + t] * c * sig[k] * s->alpol[k * nl + q];
            }
            fl[(size_t)t * nl + q] = s->wg[t] * sum;
        }
    std::vector<double> fcf((size_t)g.nxhp * g.ny);
    double akmax = sqrt(s->xsq[g.nxhp - 1] + s->ysq[g.nyh]);
    for (int j = 0; j < g.ny; j++)
        for (int ii = 0; ii < g.nxhp; ii++) {
            double ak = sqrt(s->xsq[ii] + s->ysq[j]);
            double theta = (ak - akmax * 0.) / (akmax - akmax * 0.);
            fcf[(size_t)j * g.nxhp + ii] = exp(-P.amp * pow(theta, 2. * P.pf));
        }
    // Boyd-Vandeven filter of order 2 with cut-off int(fackcbv*nzm)
    // (filtering.f:1040, 1085, 1272-1310); erfc is the C library's, as the
    // Fortran intrinsic the reference resolves to
    std::vector<double> sigbv(P.nzloc);
    {
        const int kcb = (int)(P.fackcbv * nzm);
        for (int k = 0; k <= nzm; k++) {
            double v = 1.;
            if (k > kcb && k < nzm) {
                const double theta = (double)(k - kcb) / (double)(nzm - kcb);
                const double omega = fabs(theta) - 0.5;
                double chi = 1.;
                if (theta != 0.5) {
                    const double c1 = 4. * pow(omega, 2.);
                    chi = sqrt(-log(1. - c1) / c1);
                }
                v = 0.5 * erfc(2. * sqrt(2.) * chi * omega);
            } else if (k == nzm) {
                v = 0.;
            }
            sigbv[k] = v;
        }
    }
    std::vector<double> damp(g.nz);
    for (int k = 0; k < g.nz; k++) {
        double q = (s->z[k] - P.zcen) / P.xlen;
        damp[k] = 1.0 - exp(-P.xbcrit * (q * q));
    }
    std::vector<double> zh(P.zh, P.zh + P.nsubd);
    std::vector<double> xg(P.nx);
    const double pi2 = 8. * atan(1.0);
    for (int i = 0; i < P.nx; i++) xg[i] = i * (pi2 / (pi2 / P.xlen)) / P.nx;
    std::vector<double2> twM(g.M), twN(g.M + 1), twY(g.ny);
    for (int t = 0; t < g.M; t++) twM[t] = twiddle(t, g.M);
    for (int k = 0; k <= g.M; k++) twN[k] = twiddle(k, 2 * g.M);
    for (int t = 0; t < g.ny; t++) twY[t] = twiddle(t, g.ny);
    if (upload(&g.D, s->d, acc) || upload(&g.D2, s->d2, acc) || upload(&g.AL, s->alpol, acc)
        || upload(&g.wg, s->wg, acc) || upload(&g.sig, sig, acc) || upload(&g.xw, s->xw, acc)
        || upload(&g.yw, s->yw, acc) || upload(&g.zh, zh, acc) || upload(&g.z, s->z, acc)
        || upload(&g.damp, damp, acc) || upload(&g.fcf, fcf, acc) || upload(&g.x, xg, acc)
        || upload(&g.sigbv, sigbv, acc) || upload(&g.FL, fl, acc)
        || upload(&g.twM, twM, acc) || upload(&g.twN, twN, acc) || upload(&g.twY, twY, acc)) {
        delete s;
        return -1;
    }
    size_t fb = sizeof(double) * (size_t)g.ntot;
    for (int q = 0; q < 24; q++) {
        HIPCHK(hipMalloc((void **)&s->buf[q], fb));
        HIPCHK(hipMemset(s->buf[q], 0, fb));
    }
    for (int q = 0; q < 9; q++) HIPCHK(hipMalloc((void **)&s->grad[q], fb));
    for (int q = 0; q < 7; q++) HIPCHK(hipMalloc((void **)&s->prod[q], fb));
    HIPCHK(hipMalloc((void **)&s->ox, fb));
    HIPCHK(hipMalloc((void **)&s->hbuf, fb));
    for (int q = 0; q < 4; q++) HIPCHK(hipMalloc((void **)&s->phys[q], fb));
    {
        const int nint = P.nsubd > 1 ? P.nsubd - 1 : 1, nl = P.nzloc;
        HIPCHK(hipMalloc((void **)&s->icomp, 4 * sizeof(double) * 4 * (size_t)nint * g.nxppy));
        std::vector<int> src, msrc, mdst;
        for (int i = 0; i < P.nsubd - 1; i++) {
            const int k = (i + 1) * nl; // first plane of subdomain i+1
            src.push_back(k - 2); src.push_back(k - 1); src.push_back(k); src.push_back(k + 1);
            msrc.push_back(4 * i + 1); msrc.push_back(4 * i + 2);
            mdst.push_back(k - 1); mdst.push_back(k);
        }
        if (src.empty()) { src.push_back(0); msrc.push_back(0); mdst.push_back(0); }
        if (upload(&s->iface_src, src, acc) || upload(&s->iface_mid_src, msrc, acc)
            || upload(&s->iface_mid_dst, mdst, acc))
            return -1;
        acc += (long long)fb * 4 + 4LL * sizeof(double) * 4 * nint * g.nxppy;
    }
    HIPCHK(hipMalloc((void **)&s->wnl, 2 * sizeof(double) * g.nxppy));
    HIPCHK(hipMalloc((void **)&s->wviscrot, 2 * sizeof(double) * g.nxppy));
    HIPCHK(hipMemset(s->wnl, 0, 2 * sizeof(double) * g.nxppy));
    HIPCHK(hipMemset(s->wviscrot, 0, 2 * sizeof(double) * g.nxppy));
    HIPCHK(hipMalloc((void **)&s->d_count, 4 * sizeof(int)));
    HIPCHK(hipHostMalloc((void **)&s->h_count, 4 * sizeof(int)));
    {   // maxcalc tables (various.f:696-700, 790-822; forcing.f:115-141)
        if ((double)P.nx * P.ny * g.nz >= 2147483647.) return fail("grid too large for maxcalc positions");
        const double pi2 = 8. * atan(1.0);
        s->mc_dx = (pi2 / (pi2 / P.xlen)) / P.nx;
        s->mc_dy = (pi2 / (pi2 / P.ylen)) / P.ny;
        const double xkc = pi2 / P.xlen, xmc = pi2 / (2. * P.zlen);
        s->xomega = (P.brunt * xkc) / sqrt(xkc * xkc + xmc * xmc); // main.f:341-346
        const double rhograd = -(P.rho0 * (P.brunt * P.brunt)) / P.grav;
        std::vector<double> rm(g.nz), dz(g.nz);
        // zf as setup.f:150-161 fills it (the upper subdomain writes an
        // interface point last); zf(0) is the word before zf in
        // /comgridptxyz/, z(nz)
        const int nl = P.nzloc, ns = P.nsubd;
        std::vector<double> zf(ns * (nl - 1) + 2);
        zf[0] = s->z[g.nz - 1];
        for (int ks = 0; ks < ns; ks++)
            for (int kl = 0; kl < nl; kl++) zf[ks * (nl - 1) + kl + 1] = s->z[ks * nl + kl];
        for (int ks = 0; ks < ns; ks++)
            for (int kl = 0; kl < nl; kl++) {
                const int k = ks * nl + kl, kg = ks * (nl - 1) + kl + 1;
                rm[k] = 1. + rhograd * s->z[k];
                dz[k] = zf[kg] - zf[kg - 1];
            }
        // the top point has no CFLZ and keeps the dz of the point below
        // it (various.f:817-822)
        if (g.nz > 1) dz[g.nz - 1] = dz[g.nz - 2];
        std::vector<double> rdz(g.nz);
        for (int k = 0; k < g.nz; k++) rdz[k] = 1. / dz[k];
        if (upload(&s->mc_rm, rm, acc) || upload(&s->mc_dz, dz, acc) || upload(&s->mc_rdz, rdz, acc)) return -1;
        const int nxy = P.nx * P.ny;
        // sized for the smallest transform block (the most blocks)
        // + kMcPre slots after the largest count (block size 64) for the two-level reduction
        HIPCHK(hipMalloc((void **)&s->mc_part, sizeof(double) * 2 * kMcQ * ((size_t)mc_blocks(g, 64) + kMcPre)));
        HIPCHK(hipMalloc((void **)&s->mc_rec, sizeof(double) * kMcRec));
        HIPCHK(hipMemset(s->mc_rec, 0, sizeof(double) * kMcRec));
        HIPCHK(hipMalloc((void **)&s->mc_w00, sizeof(double) * g.nz));
        HIPCHK(hipMalloc((void **)&s->mc_colmax, sizeof(double) * (size_t)(kMcSlices + 1) * nxy));
        HIPCHK(hipMalloc((void **)&s->mc_gmax, sizeof(double) * (size_t)((nxy + 255) / 256)));
        HIPCHK(hipHostMalloc((void **)&s->h_mc, sizeof(double) * kMcRec));
        // mean-wave arrays (zero) and the shear-current profiles
        // (forcing.f:190-214; their zdim, zsh, xacu arguments are unused)
        const size_t nmw = 9 * (size_t)g.nxpp * g.nz;
        HIPCHK(hipMalloc((void **)&s->mw, sizeof(double) * nmw));
        HIPCHK(hipMemset(s->mw, 0, sizeof(double) * nmw));
        const double uofc = P.uo * P.fcshear;
        const double sdt = 0.13, sz0 = 0.2625, szdom = 0.45, gam = 4.0 / (sdt * szdom);
        std::vector<double> ush(g.nz), dush(g.nz);
        for (int k = 0; k < g.nz; k++) {
            const double zz = s->z[k];
            const double c = 1.0 / cosh(gam * (zz - sz0));
            ush[k] = uofc * (0.5 * (tanh(gam * (zz - sz0)) + 1));
            dush[k] = uofc * (0.5 * gam * (c * c));
        }
        if (upload(&s->ush, ush, acc) || upload(&s->dush, dush, acc)) return -1;
    }
    acc += (long long)fb * (24 + 9 + 7 + 2) + 4 * sizeof(double) * g.nxppy;
    s->sys.kl = P.nzloc;
    s->sys.kv = 2 * P.nzloc;
    s->sys.ldab = 3 * P.nzloc + 1;
    HIPCHK(hipEventCreate(&s->ev0));
    HIPCHK(hipEventCreate(&s->ev1));
    s->t = 0.;
    s->fac[0] = 1.; s->fac[1] = 0.; s->fac[2] = 0.;
    s->bdf[0] = 1.; s->bdf[1] = 0.; s->bdf[2] = 0.;
    s->bdfv = 1.;
    s->istep = 0;
    *out = s;
    return 0;
}

extern "C" int iw_destroy(iw_solver *s)
{
    if (!s) return 0;
    hipStreamSynchronize(s->stream);
    Geo &g = s->g;
    double *tabs[] = {g.D, g.D2, g.AL, g.wg, g.sig, g.xw, g.yw, g.zh, g.z, g.damp, g.fcf, g.x, g.sigbv, g.FL};
    for (int q = 0; q < 9; q++) hipFree(s->sgs[q]);
    hipFree(s->tmean);
    for (double *t : tabs) hipFree(t);
    hipFree(g.twM);
    hipFree(g.twN);
    hipFree(g.twY);
    for (int q = 0; q < 24; q++) hipFree(s->buf[q]);
    for (int q = 0; q < 9; q++) hipFree(s->grad[q]);
    for (int q = 0; q < 7; q++) hipFree(s->prod[q]);
    hipFree(s->ox);
    hipFree(s->hbuf);
    for (int q = 0; q < 4; q++) hipFree(s->phys[q]);
    for (int q = 0; q < 12; q++) hipFree(s->s32[q]);
    hipFree(s->icomp);
    hipFree(s->iface_src);
    hipFree(s->iface_mid_src);
    hipFree(s->iface_mid_dst);
    hipFree(s->wnl);
    hipFree(s->wviscrot);
    hipFree(s->d_count);
    hipHostFree(s->h_count);
    hipFree(s->mc_rm);
    hipFree(s->mc_dz);
    hipFree(s->mc_rdz);
    hipFree(s->mc_part);
    hipFree(s->mc_rec);
    hipFree(s->mc_colmax);
    hipFree(s->mc_w00);
    hipFree(s->mw);
    hipFree(s->ush);
    hipFree(s->dush);
    hipFree(s->gT[0]);
    hipFree(s->gT[1]);
    hipFree(s->mc_gmax);
    hipFree(s->mc_ring);
    hipHostFree(s->h_mc);
    double *pp[9] = {s->pp_um, s->pp_tm, s->pp_dy, s->pp_dz, s->pp_part, s->pp_rec, s->e_ring, s->l_ring, s->kp};
    for (double *q : pp) hipFree(q);
    hipFree(s->sys.AB);
    hipFree(s->sys.FB);
    hipFree(s->sys.PB);
    hipFree(s->pen_buf);
    hipFree(s->pen_piv);
    hipFree(s->sys.SPK);
    hipFree(s->sys.SPB);
    hipFree(s->sys.spk_al);
    hipFree(s->sys.spk_au);
    hipFree(s->sys.tri_bad);
    for (double *q : s->cl_buf)
        if (q) hipFree(q);
    hipFree(s->sys.ipiv);
    hipFree(s->sys.coef);
    hipFree(s->sys.sysinfo);
    hipFree(s->sys.modesys);
    hipFree(s->sys.order_p);
    hipFree(s->sys.order_v);
    hipFree(s->sys.guv_tasks);
    hipFree(s->sys.gwt_tasks);
    hipFree(s->sys.guvw_tasks);
    hipEventDestroy(s->ev0);
    hipEventDestroy(s->ev1);
    hipStreamDestroy(s->stream);
    if (s->stream2) hipStreamDestroy(s->stream2);
    if (s->ev_fork) hipEventDestroy(s->ev_fork);
    if (s->ev_join) hipEventDestroy(s->ev_join);
    delete s;
    return 0;
}

extern "C" long long iw_device_bytes(iw_solver *s) { return s ? s->dev_bytes : 0; }

// ---------------------------------------------------------------------------
// state transfer
// ---------------------------------------------------------------------------
extern "C" int iw_set_state(iw_solver *s, int idx, const double *host)
{
    if (!s || idx < 0 || idx >= 24 || !host) return fail("iw_set_state: bad argument");
    HIPCHK(hipMemcpyAsync(s->buf[idx], host, sizeof(double) * s->g.ntot, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_get_state(iw_solver *s, int idx, double *host)
{
    if (!s || idx < 0 || idx >= 24 || !host) return fail("iw_get_state: bad argument");
    HIPCHK(hipMemcpyAsync(host, s->buf[idx], sizeof(double) * s->g.ntot, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_get_pressbc(iw_solver *s, double *wnl, double *wviscrot)
{
    if (!s) return fail("null solver");
    size_t b = 2 * sizeof(double) * s->g.nxppy;
    if (wnl) HIPCHK(hipMemcpyAsync(wnl, s->wnl, b, hipMemcpyDeviceToHost, s->stream));
    if (wviscrot) HIPCHK(hipMemcpyAsync(wviscrot, s->wviscrot, b, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_set_pressbc(iw_solver *s, const double *wnl, const double *wviscrot)
{
    if (!s) return fail("null solver");
    size_t b = 2 * sizeof(double) * s->g.nxppy;
    if (wnl) HIPCHK(hipMemcpyAsync(s->wnl, wnl, b, hipMemcpyHostToDevice, s->stream));
    if (wviscrot) HIPCHK(hipMemcpyAsync(s->wviscrot, wviscrot, b, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// dt of /timevars/ (comtime.h); the implicit operators are refactored on
// the next calcuvwt when it changes
extern "C" int iw_set_dt(iw_solver *s, double dt)
{
    if (!s) return fail("null solver");
    if (!(dt > 0.)) return fail("iw_set_dt: dt must be positive");
    s->P.dt = dt;
    return 0;
}

extern "C" int iw_set_coeffs(iw_solver *s, double t, const double fac[3], const double bdf[3], double bdfv)
{
    if (!s) return fail("null solver");
    s->t = t;
    for (int q = 0; q < 3; q++) {
        s->fac[q] = fac[q];
        s->bdf[q] = bdf[q];
    }
    s->bdfv = bdfv;
    return 0;
}

extern "C" int iw_get_coeffs(iw_solver *s, double *t, double fac[3], double bdf[3], double *bdfv)
{
    if (!s) return fail("null solver");
    if (t) *t = s->t;
    for (int q = 0; q < 3; q++) {
        if (fac) fac[q] = s->fac[q];
        if (bdf) bdf[q] = s->bdf[q];
    }
    if (bdfv) *bdfv = s->bdfv;
    return 0;
}

extern "C" double iw_time(iw_solver *s) { return s ? s->t : 0.; }

extern "C" int iw_sync(iw_solver *s)
{
    if (!s) return fail("null solver");
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_set_profiling(iw_solver *s, int on)
{
    if (!s) return fail("null solver");
    s->profiling = on != 0;
    s->prof = Prof();
    return 0;
}

extern "C" int iw_kernel_stats(iw_solver *s, char *names, int cap, double *avg_ms, int *launches,
                               double *avg_bytes, int maxk)
{
    if (!s) return fail("null solver");
    prof_resolve(s);
    std::string all;
    int k = 0;
    for (; k < (int)s->prof.names.size() && k < maxk; k++) {
        if (k) all += ";";
        all += s->prof.names[k];
        if (avg_ms) avg_ms[k] = s->prof.n[k] ? s->prof.ms[k] / s->prof.n[k] : 0.;
        if (launches) launches[k] = s->prof.n[k];
        if (avg_bytes) avg_bytes[k] = s->prof.n[k] ? s->prof.bytes[k] / s->prof.n[k] : 0.;
    }
    if (names && cap > 0) {
        strncpy(names, all.c_str(), (size_t)cap - 1);
        names[cap - 1] = 0;
    }
    return k;
}

// ---------------------------------------------------------------------------
// step pieces
// ---------------------------------------------------------------------------
static FieldBatch batch(std::initializer_list<double *> l)
{
    FieldBatch b;
    b.n = 0;
    b.smap = b.dmap = nullptr;
    for (double *p : l) {
        b.f[b.n] = p;
        b.g[b.n] = p;
        b.n++;
    }
    return b;
}

// horfft (horfft.f:2-69) over nk logical planes of a batch: the first pass
// reads b.f (through smap) and writes b.g (through dmap), the second pass
// works in place on b.g.  is = -1: physical -> spectral.
enum FftUse { kFftMain = 0, kFftIface = 1, kFftMaxcalc = 2 };

static void fft_dev(iw_solver *s, FieldBatch b, int is, int nk, int use = kFftMain, const McFuse *mf = nullptr)
{
    if (b.n == 0 || nk == 0) return;
    // timing scopes: the interface-plane and maxcalc transforms are kept
    // apart from the others
    static const char *names[3][4] = {{"fft_x_fwd", "fft_y_fwd", "fft_x_inv", "fft_y_inv"},
                                      {"iface_fft_x_fwd", "iface_fft_y_fwd", "iface_fft_x_inv", "iface_fft_y_inv"},
                                      {"mc_fft_x_fwd", "mc_fft_y_fwd", "mc_fft_x_inv", "mc_fft_y_inv"}};
    const char *nxf = names[use][0], *nyf = names[use][1], *nxi = names[use][2], *nyi = names[use][3];
    const double bx = 16. * (double)s->g.nxppy * nk * b.n;         // each double read + written
    const double by = 32. * (double)s->g.nxh * s->g.ny * nk * b.n; // columns ii < nxh
    FieldBatch b2 = b;
    for (int q = 0; q < b.n; q++) b2.f[q] = b.g[q];
    b2.smap = b.dmap;
    const bool y = s->g.ndims == 3;
    if (is == -1) {
        {
            Scope sc(s, nxf, bx);
            launch_fft_x(s->g, b, nk, -1, s->stream);
        }
        if (y) {
            Scope sc(s, nyf, by);
            launch_fft_y(s->g, b2, nk, -1, s->stream);
        }
    } else {
        if (y) {
            {
                Scope sc(s, nyi, by);
                launch_fft_y(s->g, b, nk, 1, s->stream);
            }
            Scope sc(s, nxi, bx);
            launch_fft_x(s->g, b2, nk, 1, s->stream, mf);
        } else {
            Scope sc(s, nxi, bx);
            launch_fft_x(s->g, b, nk, 1, s->stream, mf);
        }
    }
}

static void horfft_dev(iw_solver *s, FieldBatch b, int is, int nk) { fft_dev(s, b, is, nk); }

void iw_fft_batch(iw_solver *s, const std::vector<double *> &f, const std::vector<double *> &g, int is)
{
    for (size_t i = 0; i < f.size(); i += kMaxFields) {
        FieldBatch b;
        b.n = 0;
        b.smap = b.dmap = nullptr;
        for (size_t q = i; q < f.size() && q < i + kMaxFields; q++) {
            b.f[b.n] = f[q];
            b.g[b.n] = g[q];
            b.n++;
        }
        fft_dev(s, b, is, s->g.nz);
    }
}

static void batch_fft(iw_solver *s, std::vector<double *> v, int is)
{
    for (size_t i = 0; i < v.size(); i += kMaxFields) {
        FieldBatch b;
        b.n = 0;
        b.smap = b.dmap = nullptr;
        for (size_t q = i; q < v.size() && q < i + kMaxFields; q++) {
            b.f[b.n] = v[q];
            b.g[b.n] = v[q];
            b.n++;
        }
        fft_dev(s, b, is, s->g.nz);
    }
}

// spectral effect of an inverse+forward horfft round trip on u,v,w,temp
static void project_dev(iw_solver *s, double *const *f, int nf)
{
    FieldBatch b;
    b.n = nf;
    b.smap = b.dmap = nullptr;
    for (int q = 0; q < nf; q++) b.f[q] = b.g[q] = f[q];
    Scope sc(s, "project", 16. * (double)(s->g.ny + s->g.ny) * s->g.nz * nf);
    launch_project(s->g, b, s->stream);
}

// are the mean-wave / shear-current terms of solver.f:848-931 active?
static bool mean_active(const iw_solver *s) { return s->mw_on || s->P.uo * s->P.fcshear != 0.0; }

// maxcalc reductions in the epilogue of an inverse x transform of u,v,w,
// temp: cny = ny-1 and the removed mean of w when the transform is convec's
static McFuse mc_fuse(iw_solver *s, double cny, const double *w00)
{
    McFuse f;
    if (s->mw_on) {
        const size_t nm = (size_t)s->g.nxpp * s->g.nz;
        f.umwv = s->mw;
        f.wmwv = s->mw + nm;
        f.rhomwv = s->mw + 6 * nm;
    }
    if (s->P.uo * s->P.fcshear != 0.0) f.ush = s->ush;
    f.part = s->mc_part;
    f.rm = s->mc_rm;
    f.rdz = s->mc_rdz;
    f.w00 = w00;
    f.dt = s->P.dt;
    f.cx = s->P.dt / s->mc_dx;
    f.cy = s->P.dt / s->mc_dy;
    f.cny = cny;
    return f;
}

// the maxcalc record (s->mc_rec) from the partials of that transform
static void mc_record(iw_solver *s, int istep, double cny, const float *u32 = nullptr)
{
    const Geo &g = s->g;
    McArgs A;
    A.u = s->phys[0];
    A.u32 = u32;
    A.spec = s->buf[0];
    A.cny = cny;
    A.dz = s->mc_dz;
    A.nx = g.nx; A.ny = g.ny; A.nz = g.nz; A.nxpp = g.nxpp; A.nxppy = g.nxppy;
    A.part = s->mc_part;
    A.nblk = mc_blocks(g);
    A.part2 = s->mc_part + (size_t)2 * kMcQ * mc_blocks(g, 64);
    A.umwv = s->mw_on ? s->mw : nullptr;
    {
        Scope sc(s, "maxcalc");
        launch_maxcalc_final(g, A, istep, s->t, s->P.umax, s->mc_rec, s->stream);
    }
    if (s->P.umax != 0.0) {
        Scope sc(s, "mc_umaxt", 8. * (double)g.ntot);
        launch_maxcalc_umaxt(g, A, s->P.umax, s->mc_colmax, s->mc_gmax, s->mc_rec, s->stream);
    }
}

static int mc_after(iw_solver *s, int slot);
static int convec_tail(iw_solver *s, int ispecfilternl);

// convec (solver.f:6-514) on the device state.  mcstep > 0: maxcalc of that
// step (its record from convec's inverse transforms, then the adaptive step
// and main.f:660-664) before anything that reads dt or the coefficients;
// its record goes to slot mcslot of the run.
static int convec_dev(iw_solver *s, int ispecfilternl, int mcstep = 0, int mcslot = 0)
{
    const Geo &g = s->g;
    double **B = s->buf;
    double *uvwt[4] = {B[0], B[1], B[2], B[3]};
    double *S[4] = {B[12], B[13], B[14], B[15]};
    double **gr = s->grad;
    const double nt = (double)g.ntot;
    {
        Scope sc(s, "zero_meanw");
        launch_zero_meanw(g, B[2], s->mc_w00, s->stream);
    }
    const McFuse mf = mc_fuse(s, g.ndims == 3 ? (double)(g.ny - 1) : 0.0, s->mc_w00);
    // the mean-wave / shear-current terms also need dtdx, dtdz
    // (solver.f:74-103)
    const bool mean = mean_active(s);
    if (mean && !s->gT[0]) {
        for (int q = 0; q < 2; q++) HIPCHK(hipMalloc((void **)&s->gT[q], sizeof(double) * (size_t)g.ntot));
        s->dev_bytes += 2 * (long long)sizeof(double) * g.ntot;
    }
    if (g.ndims == 3) {
        // spectral-to-physical (solver.f:107-123) with the x/y gradients of
        // u,v,w (solver.f:86-103) formed inside the transforms: the y pass
        // turns u,v,w,temp into the physical buffers and emits i ky f of
        // u,v,w; the x pass turns all seven and emits i kx f of u,v,w (and
        // temp).  The z gradients are taken after the transforms, from the
        // physical fields, together with the vertical fluxes below.
        FieldBatch by = batch({B[0], B[1], B[2], B[3]});
        for (int q = 0; q < 4; q++) by.g[q] = s->phys[q];
        by.dy[0] = gr[1];
        by.dy[1] = gr[4];
        by.dy[2] = gr[7];
        if (s->closure) by.dy[3] = s->cl_buf[1]; // d rho/dy for the closure's features
        by.kw = g.yw;
        {
            Scope sc(s, "fft_y_inv", 32. * g.nxh * g.ny * g.nz * 7);
            launch_fft_y(g, by, g.nz, 1, s->stream);
        }
        FieldBatch bx = batch({s->phys[0], s->phys[1], s->phys[2], s->phys[3], gr[1], gr[4], gr[7]});
        bx.dx[0] = gr[0];
        bx.dx[1] = gr[3];
        bx.dx[2] = gr[6];
        if (mean) bx.dx[3] = s->gT[0];
        if (s->closure) { // d rho/dy to physical space, d rho/dx
            bx.f[bx.n] = bx.g[bx.n] = s->cl_buf[1];
            bx.n++;
            bx.dx[3] = s->cl_buf[0];
        }
        bx.kw = g.xw;
        {
            Scope sc(s, "fft_x_inv", 16. * g.nxppy * g.nz * (mean ? 11 : 10));
            launch_fft_x(g, bx, g.nz, 1, s->stream, mcstep > 0 ? &mf : nullptr);
        }
    } else {
        {
            Scope sc(s, "dpdz", 48. * nt);
            double *o3[3] = {gr[2], gr[5], gr[8]};
            launch_dpdz3(g, uvwt, o3, s->stream);
        }
        if (mean) {
            Scope sc(s, "dpdz", 16. * nt);
            launch_dpdz(g, B[3], s->gT[1], s->stream);
        }
        {
            Scope sc(s, "grad", 72. * nt);
            launch_grad(g, uvwt, gr, s->stream);
        }
        FieldBatch b = batch({B[0], B[1], B[2], B[3], gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7],
                              gr[8]});
        if (mean) {
            launch_dx(g, B[3], s->gT[0], s->stream);
            b.f[b.n] = b.g[b.n] = s->gT[0];
            b.n++;
            b.f[b.n] = b.g[b.n] = s->gT[1];
            b.n++;
        }
        for (int q = 0; q < 4; q++) b.g[q] = s->phys[q];
        fft_dev(s, b, 1, g.nz, kFftMain, mcstep > 0 ? &mf : nullptr);
    }
    if (mcstep > 0) {
        mc_record(s, mcstep, mf.cny);
        if (mc_after(s, mcslot)) return -1;
    }
    double *ph[4] = {s->phys[0], s->phys[1], s->phys[2], s->phys[3]};
    // z flux derivatives, z gradients and the physical right-hand sides in
    // one pass when nothing else needs the z gradients (no mean-wave terms,
    // no closure): 10 fields in and 11 out instead of 4 in and 7 out, then 18
    // in and 11 out
    bool fused = false;
    if (g.ndims == 3 && !mean && !s->closure) {
        Scope sc(s, "conv_rhs", 168. * nt);
        fused = launch_conv_rhs(g, s->P, s->t, ph, gr, S, s->prod, s->stream);
    }
    if (!fused) {
        if (g.ndims == 3) { // fluxes and z gradients: 4 fields in, 7 (or 8) out
            Scope sc(s, "dpdz_flux", (mean ? 96. : 88.) * nt);
            double *o4[4] = {S[0], S[1], S[2], s->ox};
            double *dz4[4] = {gr[2], gr[5], gr[8], mean ? s->gT[1] : (s->closure ? s->cl_buf[2] : nullptr)};
            launch_dpdz_conv_grad(g, ph, o4, dz4, s->stream);
        } else {
            Scope sc(s, "dpdz_flux", 64. * nt);
            double *o4[4] = {S[0], S[1], S[2], s->ox};
            launch_dpdz_conv(g, ph, o4, s->stream);
        }
        {
            Scope sc(s, "rhs_phys", 224. * nt);
            MeanTerms mt;
            if (mean) {
                mt.mw = s->mw;
                mt.ush = s->ush;
                mt.dush = s->dush;
                mt.dtdx = s->gT[0];
                mt.dtdz = s->gT[1];
            }
            launch_rhs_phys(g, s->P, s->t, ph, gr, S, s->ox, s->prod, s->stream, &mt);
        }
    }
    if (s->closure) {
        Scope sc(s, "closure");
        if (g.ndims != 3 || iw_closure_apply(s, S, s->prod)) return -1;
    }
    double **pr = s->prod;
    if (g.ndims == 3) {
        // physical-to-spectral of the four right-hand sides and the seven
        // flux products (solver.f:1167-1180) with the flux divergence
        // i kx P_a + i ky P_b (solver.f:985-1063, 1163-1196) added inside the
        // transforms: the x pass turns S_u,S_v,S_w,S_T adding i kx of
        // P0,P1,P3,P5 and writes the x transforms of P1,P2,P4,P6 to free
        // gradient buffers; the y pass turns S_u.. adding i ky of those
        // P1 is both S_v's x flux (ax) and S_u's y flux: its transform in
        // S_v's blocks is stored for the y pass as well (axo)
        FieldBatch bx = batch({S[0], S[1], S[2], S[3], pr[2], pr[4], pr[6]});
        const int ia[4] = {0, 1, 3, 5};
        for (int q = 0; q < 4; q++) bx.ax[q] = pr[ia[q]];
        bx.axo[1] = gr[0];
        for (int q = 1; q < 4; q++) bx.g[3 + q] = gr[q];
        bx.kw = g.xw;
        {
            Scope sc(s, "fft_x_fwd", 16. * g.nxppy * g.nz * 9.5);
            launch_fft_x(g, bx, g.nz, -1, s->stream);
        }
        FieldBatch by = batch({S[0], S[1], S[2], S[3]});
        for (int q = 0; q < 4; q++) by.ay[q] = gr[q];
        by.kw = g.yw;
        {
            Scope sc(s, "fft_y_fwd", 32. * g.nxh * g.ny * g.nz * 6);
            launch_fft_y(g, by, g.nz, -1, s->stream);
        }
    } else {
        batch_fft(s, {S[0], S[1], S[2], S[3], pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], pr[6]}, -1);
        Scope sc(s, "rhs_spec", 120. * nt);
        launch_rhs_spec(g, S, pr, s->stream);
    }
    return convec_tail(s, ispecfilternl);
}

// The spectral part of convec after the right-hand sides S are formed: the
// round trip of u,v,w,temp, the pressure BC terms, SS/BDF assembly with the
// filters, history rotation.
static int convec_tail(iw_solver *s, int ispecfilternl)
{
    const Geo &g = s->g;
    double **B = s->buf;
    double *uvwt[4] = {B[0], B[1], B[2], B[3]};
    const double nt = (double)g.ntot;
    // u,v,w,temp back to spectral space (solver.f:1180, 1201-1203): their
    // physical values were not changed, so the round trip reduces to its
    // spectral effect on the kx=0 and kx-Nyquist columns
    project_dev(s, uvwt, 4);
    {
        Scope sc(s, "wnl", 64. * (double)g.nxppy);
        if (s->mixed)
            launch_wnl(g, s->s32[2], s->s32[6], s->s32[10], s->fac, s->wnl, s->stream);
        else
            launch_wnl(g, B[14], B[18], B[22], s->fac, s->wnl, s->stream);
    }
    {
        Scope sc(s, "rotvisc");
        double *lev[9] = {B[0], B[1], B[2], B[4], B[5], B[6], B[8], B[9], B[10]};
        launch_rotvisc(g, s->P, lev, s->fac, s->wviscrot, s->stream);
    }
    // with the packed pressure path set up (from the first step on), the
    // same pass forms the pressure right-hand side for this step's calcuvwt
    const bool prhs = s->fuse_prhs && s->sys.gwt_tasks && s->sys.tri_ok_p && s->p_factored;
    s->phi_ready = false;
    if (s->mixed) { // FP32 right-hand-side histories
        Scope sc(s, "bdf_filter", (prhs ? 4 * 44. + 8. : 4 * 44.) * nt);
        float *S12[12];
        double *F12[12], *out4[4];
        for (int q = 0; q < 4; q++) {
            S12[3 * q] = s->s32[q]; S12[3 * q + 1] = s->s32[4 + q]; S12[3 * q + 2] = s->s32[8 + q];
            F12[3 * q] = B[q]; F12[3 * q + 1] = B[4 + q]; F12[3 * q + 2] = B[8 + q];
            out4[q] = B[8 + q];
        }
        if (prhs)
            s->phi_ready = launch_bdf_filter_prhs(g, S12, F12, out4, s->fac, s->bdf, s->P.dt, ispecfilternl == 1,
                                                  s->sys, s->grad[0], s->wnl, s->wviscrot, s->stream);
        if (!s->phi_ready)
            launch_bdf_filter(g, S12, F12, out4, s->fac, s->bdf, s->P.dt, ispecfilternl == 1, s->stream);
    } else {
        Scope sc(s, "bdf_filter", (prhs ? 4 * 56. + 8. : 4 * 56.) * nt);
        double *S12[12], *F12[12], *out4[4];
        for (int q = 0; q < 4; q++) {
            S12[3 * q] = B[12 + q]; S12[3 * q + 1] = B[16 + q]; S12[3 * q + 2] = B[20 + q];
            F12[3 * q] = B[q]; F12[3 * q + 1] = B[4 + q]; F12[3 * q + 2] = B[8 + q];
            out4[q] = B[8 + q];
        }
        if (prhs)
            s->phi_ready = launch_bdf_filter_prhs(g, S12, F12, out4, s->fac, s->bdf, s->P.dt, ispecfilternl == 1,
                                                  s->sys, s->grad[0], s->wnl, s->wviscrot, s->stream);
        if (!s->phi_ready)
            launch_bdf_filter(g, S12, F12, out4, s->fac, s->bdf, s->P.dt, ispecfilternl == 1, s->stream);
    }
    for (int q = 0; q < 4; q++) { // history rotation (solver.f:447-478) by pointer swap
        double *nw = B[8 + q];
        B[8 + q] = B[4 + q];
        B[4 + q] = B[q];
        B[q] = nw;
        if (s->mixed) { // the FP64 S, Sn, Snm are written when the mixed run ends
            float *fr = s->s32[8 + q];
            s->s32[8 + q] = s->s32[4 + q];
            s->s32[4 + q] = s->s32[q];
            s->s32[q] = fr;
            continue;
        }
        double *fr = B[20 + q];
        B[20 + q] = B[16 + q];
        B[16 + q] = B[12 + q];
        B[12 + q] = fr;
    }
    return hipGetLastError() == hipSuccess ? 0 : fail("convec launch failed");
}

// convec on the mixed path: the same sequence as convec_dev (3-D, no
// maxcalc reductions, no mean-wave terms) with the physical-space working set
// in FP32.  The inverse y pass reads the FP64 spectral state and writes FP32;
// the x passes, the z contraction (FP32 matrix cores) and the products run on
// FP32 arrays; the forward y pass writes the spectral right-hand sides in
// FP32 (s32: the products they come from carry FP32 rounding already), which
// the SS/BDF assembly widens.  From there on (round-trip projection, pressure
// BC terms, SS/BDF assembly, filters and the implicit step) everything is FP64
// as in the plain path.
// The FP32 arrays live in the FP64 scratch of the plain path: phys[q] holds
// the physical field (first half) and its right-hand side (second half),
// grad[q], prod[q] and ox one FP32 array each.
static int convec_mixed(iw_solver *s, int ispecfilternl, int mcstep = 0, int mcslot = 0)
{
    const Geo &g = s->g;
    double **B = s->buf;
    float *ph[4], *S32[4], *g32[9], *pr32[7];
    for (int q = 0; q < 4; q++) {
        ph[q] = reinterpret_cast<float *>(s->phys[q]);
        S32[q] = ph[q] + g.ntot;
    }
    for (int q = 0; q < 9; q++) g32[q] = reinterpret_cast<float *>(s->grad[q]);
    for (int q = 0; q < 7; q++) pr32[q] = reinterpret_cast<float *>(s->prod[q]);
    float *ox32 = reinterpret_cast<float *>(s->ox);
    auto D = [](float *p) { return reinterpret_cast<double *>(p); };
    const double nt = (double)g.ntot, nc = (double)g.nxh * g.ny * g.nz;
    {
        Scope sc(s, "zero_meanw");
        launch_zero_meanw(g, B[2], s->mc_w00, s->stream);
    }
    // maxcalc's reductions in the inverse x pass read the FP64 values of the
    // transform before they are stored in FP32; the umaxt chain reads FP32 u
    const McFuse mf = mc_fuse(s, (double)(g.ny - 1), s->mc_w00);
    {   // u,v,w,temp (FP64 spectral) -> FP32, with i ky f of u,v,w
        FieldBatch by = batch({B[0], B[1], B[2], B[3]});
        for (int q = 0; q < 4; q++) by.g[q] = D(ph[q]);
        by.dy[0] = D(g32[1]);
        by.dy[1] = D(g32[4]);
        by.dy[2] = D(g32[7]);
        by.kw = g.yw;
        Scope sc(s, "m_fft_y_inv", (4 * 16. + 7 * 8.) * nc);
        launch_fft_y(g, by, g.nz, 1, s->stream, kPrecF32);
    }
    {   // x pass in place on the seven FP32 arrays, with i kx f of u,v,w
        FieldBatch bx = batch({D(ph[0]), D(ph[1]), D(ph[2]), D(ph[3]), D(g32[1]), D(g32[4]), D(g32[7])});
        bx.dx[0] = D(g32[0]);
        bx.dx[1] = D(g32[3]);
        bx.dx[2] = D(g32[6]);
        bx.kw = g.xw;
        Scope sc(s, "m_fft_x_inv", 17 * 4. * (double)g.nxppy * g.nz);
        launch_fft_x(g, bx, g.nz, 1, s->stream, mcstep > 0 ? &mf : nullptr, kPrecF32);
    }
    if (mcstep > 0) {
        mc_record(s, mcstep, mf.cny, ph[0]);
        if (mc_after(s, mcslot)) return -1;
    }
    if (s->mixed_mfma) { // the z contraction on the matrix cores, then the products
        {
            float *o4[4] = {S32[0], S32[1], S32[2], ox32};
            float *dz3[3] = {g32[2], g32[5], g32[8]};
            Scope sc(s, "m_dpdz_mfma", 11 * 4. * nt);
            if (!launch_dpdz_conv_grad_mfma(g, ph, o4, dz3, s->stream))
                return fail("mixed path: no matrix-core z contraction for this GLL order");
        }
        {
            Scope sc(s, "m_rhs_phys", 28 * 4. * nt);
            launch_rhs_phys_f32(g, s->P, s->t, ph, g32, S32, ox32, pr32, s->stream);
        }
    } else { // vertical fluxes, z gradients and the products in one pass (the z sums on the VALU from LDS)
        Scope sc(s, "m_conv_rhs", 21 * 4. * nt);
        if (!launch_conv_rhs(g, s->P, s->t, ph, g32, S32, pr32, s->stream))
            return fail("mixed path: no fused z contraction for this GLL order");
    }
    {   // forward x pass: as convec_dev, on FP32 arrays
        FieldBatch bx = batch({D(S32[0]), D(S32[1]), D(S32[2]), D(S32[3]), D(pr32[2]), D(pr32[4]), D(pr32[6])});
        const int ia[4] = {0, 1, 3, 5};
        for (int q = 0; q < 4; q++) bx.ax[q] = D(pr32[ia[q]]);
        bx.axo[1] = D(g32[0]);
        for (int q = 1; q < 4; q++) bx.g[3 + q] = D(g32[q]);
        bx.kw = g.xw;
        Scope sc(s, "m_fft_x_fwd", 19 * 4. * (double)g.nxppy * g.nz);
        launch_fft_x(g, bx, g.nz, -1, s->stream, nullptr, kPrecF32);
    }
    {   // forward y pass: FP32 in, FP32 spectral right-hand sides out
        FieldBatch by = batch({D(S32[0]), D(S32[1]), D(S32[2]), D(S32[3])});
        for (int q = 0; q < 4; q++) {
            by.g[q] = D(s->s32[q]);
            by.ay[q] = D(g32[q]);
        }
        by.kw = g.yw;
        Scope sc(s, "m_fft_y_fwd", 12 * 8. * nc);
        launch_fft_y(g, by, g.nz, -1, s->stream, kPrecF32S);
    }
    return convec_tail(s, ispecfilternl);
}

static int calcuvwt_dev(iw_solver *s, int cpen)
{
    if (ensure_factors(s)) return -1;
    double *uvwt[4] = {s->buf[0], s->buf[1], s->buf[2], s->buf[3]};
    // the rho solve needs nothing the pressure step changes: with the u/v/w
    // groups it runs with its coupling pass and filter on the second stream
    // beside the pressure step and the u/v/w solves, joined after the u/v/w
    // coupling pass (not when profiling, so that the per-class times stay
    // those of one stream)
    const bool rho_side = !s->seq_sweeps && s->sys.np_ok && s->sys.share_uvw && s->stream2 && !s->profiling &&
                          !s->one_stream;
    if (rho_side) {
        HIPCHK(hipEventRecord(s->ev_fork, s->stream));
        HIPCHK(hipStreamWaitEvent(s->stream2, s->ev_fork, 0));
        launch_solve_spike(s->g, s->P, s->sys, uvwt, s->P.dt, s->stream2, 2);
        launch_spike_finish(s->g, s->sys, uvwt, cpen, s->stream2, 2);
        HIPCHK(hipEventRecord(s->ev_join, s->stream2));
    }
    // algorithmic bytes: state columns in and out + every distinct cached factor once
    const double fcol = (double)s->sys.ldab * s->g.nz * 8. + s->g.nz * 4.;
    const double nt = (double)s->g.ntot;
    if (s->sys.gwt_tasks && s->sys.tri_ok_p) {
        const char *name[3] = {"p_rhs", "p_solve", "p_project"};
        // u,v,w in, phi out | phi in and out + each P factor once |
        // phi, u,v,w in, u,v,w out
        // the pressure factor's structural nonzeros: per row nl - r below the
        // diagonal, r + nl + 1 on and above it (2 nl + 1 in all), and the pivot
        const double fpt = (double)(2 * s->P.nzloc + 1) * s->g.nz * 8. + s->g.nz * 4.;
        const double bytes[3] = {32. * nt, 16. * nt + s->sys.np * fpt, 56. * nt};
        for (int part = 0; part < 3; part++) {
            if (part == 0 && s->phi_ready) continue; // formed by convec's SS/BDF pass
            Scope sc(s, name[part], bytes[part]);
            launch_pressure_nat(s->g, s->P, s->sys, uvwt, s->grad[0], s->wnl, s->wviscrot, s->P.dt, cpen, part,
                                s->stream);
        }
    } else {
        Scope sc(s, "pressure", 56. * nt + s->sys.np * fcol); // u,v,w in; u,v,w out; w re-read
        launch_pressure(s->g, s->P, s->sys, uvwt, s->wnl, s->wviscrot, s->P.dt, cpen, s->stream);
    }
    s->phi_ready = false;
    if (s->sys.np_ok) {
        {
            // right-hand sides in, solutions out (16 B per point and field)
            // + each distinct no-interchange factor's structural nonzeros
            // once (per row nl - r below the diagonal, r + 2 on and above it)
            const double fnp = (double)(s->P.nzloc + 2) * s->g.nz * 8.;
            // + the coupling scalars written (4 fields x nsubd x columns x 2);
            // with the u/v/w groups the w factors below the top subdomain are
            // the u/v ones, read once
            const double cpl = 2. * 4. * s->g.nsubd * s->g.nxppy * 8.;
            const double nfac = (s->sys.nsys - s->sys.np) -
                                ((!s->seq_sweeps && s->sys.share_uvw)
                                     ? s->sys.nsys_w * (s->P.nsubd - 1.) / s->P.nsubd : 0.);
            Scope sc(s, "solve_nat", 4 * 16. * nt + nfac * fnp + (s->seq_sweeps ? 0. : cpl));
            if (s->seq_sweeps)
                launch_solve_grp(s->g, s->P, s->sys, uvwt, s->P.dt, s->stream);
            else
                launch_solve_spike(s->g, s->P, s->sys, uvwt, s->P.dt, s->stream, rho_side ? 1 : 0);
        }
        if (s->seq_sweeps) {
            Scope sc(s, "lfilter", 4 * 16. * nt);
            launch_lfilter(s->g, uvwt, cpen, s->stream);
        } else {
            // + the coupling scalars (the recurrences read both and write
            // both, the correction reads both) and each distinct system's
            // k, h and beta, gamma, delta once
            const double fix = 6. * 4. * s->g.nsubd * s->g.nxppy * 8. +
                               (s->sys.nsys - s->sys.np) * (2. * s->P.nzloc + 4.) * s->P.nsubd * 8.;
            Scope sc(s, "lfilter", 4 * 16. * nt + fix);
            launch_spike_finish(s->g, s->sys, uvwt, cpen, s->stream, rho_side ? 1 : 0);
        }
        if (rho_side) HIPCHK(hipStreamWaitEvent(s->stream, s->ev_join, 0));
    } else {
        Scope sc(s, "viscous", 64. * nt + (s->sys.nsys - s->sys.np) * fcol);
        launch_viscous(s->g, s->P, s->sys, uvwt, s->P.dt, cpen, s->stream);
    }
    return hipGetLastError() == hipSuccess ? 0 : fail("calcuvwt launch failed");
}

// interface_avg3d (penalty.f:714-775): only the four planes km-1, km, k, k+1
// around each interface are taken to physical space; after the averaging the
// two modified planes are transformed back, and every other plane receives
// the spectral effect of the round trip
static int iface_dev(iw_solver *s, double *const *f, int nf)
{
    const int nint = s->g.nsubd - 1;
    launch_zero_int(s->d_count, 4, s->stream);
    if (nint <= 0) {
        project_dev(s, f, nf);
        return 0;
    }
    FieldBatch bi;
    bi.n = nf;
    bi.smap = s->iface_src;
    bi.dmap = nullptr;
    for (int q = 0; q < nf; q++) {
        bi.f[q] = f[q];
        bi.g[q] = s->icomp + (size_t)q * 4 * nint * s->g.nxppy;
    }
    fft_dev(s, bi, 1, 4 * nint, kFftIface);
    {
        Scope sc(s, "iface", 48. * (double)s->g.nxppy * nint * nf);
        FieldBatch b;
        b.n = nf;
        b.smap = b.dmap = nullptr;
        for (int q = 0; q < nf; q++) b.f[q] = b.g[q] = bi.g[q];
        launch_iface(s->g, b, s->d_count, s->stream);
    }
    project_dev(s, f, nf);
    FieldBatch bf;
    bf.n = nf;
    bf.smap = s->iface_mid_src;
    bf.dmap = s->iface_mid_dst;
    for (int q = 0; q < nf; q++) {
        bf.f[q] = bi.g[q];
        bf.g[q] = f[q];
    }
    fft_dev(s, bf, -1, 2 * nint, kFftIface);
    return 0;
}

// maxcalc (various.f:665-987) on the device state: physical u,v,w,temp
// (horfft +1) into the physical scratch fields, the extrema and CFL record
// (s->mc_rec), and the return to spectral space, whose only effect on the
// spectrum is on the kx=0 and kx-Nyquist columns
static void maxcalc_dev(iw_solver *s, int istep)
{
    const Geo &g = s->g;
    FieldBatch b = batch({s->buf[0], s->buf[1], s->buf[2], s->buf[3]});
    for (int q = 0; q < 4; q++) b.g[q] = s->phys[q];
    const McFuse mf = mc_fuse(s, 0.0, nullptr);
    fft_dev(s, b, 1, g.nz, kFftMaxcalc, &mf);
    mc_record(s, istep, 0.0);
    project_dev(s, s->buf, 4);
}

// variable-step AB3/BDF3 coefficients for step sizes (d1, d2, d3)
static void var_coeffs(iw_solver *s, double d1, double d2, double d3)
{
    s->fac[0] = iw_abvar1(d1, d2, d3); s->fac[1] = iw_abvar2(d1, d2, d3); s->fac[2] = iw_abvar3(d1, d2, d3);
    s->bdf[0] = iw_bdfvar1(d1, d2, d3); s->bdf[1] = iw_bdfvar2(d1, d2, d3); s->bdf[2] = iw_bdfvar3(d1, d2, d3);
}

// adaptive step (various.f:887-985) from the record on the host; fills
// the record's dt, cdt, change flag, ivardt, idtstep
static void adapt_dt(iw_solver *s, double *rec)
{
    int changed = 0;
    if (s->P.dtadapt) {
        const double cflx = rec[22], cflz = rec[30];
        const bool change_dt = (cflz > 0.7) || (cflz < 0.1) || (cflx > 0.16);
        const double pi2 = 8. * atan(1.0);
        const double tlim = pi2 / (s->xomega * 50.);
        const double tso = s->P.dtadapt == 2 ? 0. : s->Tso;
        const int nst = (int)((s->t - tso) / s->P.dt);
        if (nst > 20 && change_dt) {
            s->Tso = s->t;
            if (cflz < 0.1) s->cdt = fmax(0.8, s->P.dt / tlim);
            if ((cflz > 0.7) || (cflx > 0.16)) s->cdt = 1.25;
            s->P.dt = s->P.dt / s->cdt;
            const double dt = s->P.dt, d2 = s->cdt * dt;
            var_coeffs(s, dt, d2, d2);
            s->bdfv = iw_bdfvarv(dt, d2, d2);
            s->ivardt = 1;
            s->idtstep = 0;
            changed = 1;
        }
    }
    rec[35] = s->P.dt;
    rec[36] = s->cdt;
    rec[37] = changed;
    rec[38] = s->ivardt;
    rec[39] = s->idtstep;
}

// maxcalc record to the host (waits for the stream)
static int mc_fetch(iw_solver *s)
{
    HIPCHK(hipMemcpyAsync(s->h_mc, s->mc_rec, sizeof(double) * kMcRec, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// after the record of a step of iw_run: the adaptive step (the record to
// the host) or the record into the run's ring; then main.f:660-664
static int mc_after(iw_solver *s, int slot)
{
    if (s->P.dtadapt) {
        if (mc_fetch(s)) return -1;
        adapt_dt(s, s->h_mc);
        s->mc_hist.insert(s->mc_hist.end(), s->h_mc, s->h_mc + kMcRec);
    } else {
        HIPCHK(hipMemcpyAsync(s->mc_ring + (size_t)slot * kMcRec, s->mc_rec, sizeof(double) * kMcRec,
                              hipMemcpyDeviceToDevice, s->stream));
    }
    if (s->ivardt == 1) s->idtstep++;
    return 0;
}

// ---------------------------------------------------------------------------
// diagnostics: postp (postp.f:9-76) and speclegencalc (filtering.f:1110-1221)
// ---------------------------------------------------------------------------
static int pp_alloc(iw_solver *s)
{
    if (s->pp_um) return 0;
    const size_t r = (size_t)s->g.ny * s->g.nz;
    const size_t n[6] = {r, 6 * r, r, r, (size_t)pp_blocks(s->g) * 12, 18 + 2 + 4 * (size_t)s->g.nz};
    double **p[6] = {&s->pp_um, &s->pp_tm, &s->pp_dy, &s->pp_dz, &s->pp_part, &s->pp_rec};
    for (int q = 0; q < 6; q++) {
        HIPCHK(hipMalloc((void **)p[q], sizeof(double) * n[q]));
        s->dev_bytes += (long long)(sizeof(double) * n[q]);
    }
    return 0;
}

// grow a device record ring to hold n doubles (contents not kept)
static int ring_reserve(iw_solver *s, double **ring, int *cap, size_t n)
{
    if ((size_t)*cap >= n) return 0;
    hipFree(*ring);
    *ring = nullptr;
    *cap = 0;
    HIPCHK(hipMalloc((void **)ring, sizeof(double) * n));
    *cap = (int)n;
    return 0;
}

// postp on the device state, out (device): istep, t and the sixteen values
// of the energy.dat line.  The state changes as the reference's does: u,v,w
// take two horfft round trips (postp.f:23-25, 47-49 and energycalc_postp's
// own, postp.f:466-495), temp three (umeancalc's, postp.f:189-196, besides).
// Scratch: phys, grad, prod[0..3] (free between steps).
static int postp_dev(iw_solver *s, int istep, double *out)
{
    if (pp_alloc(s)) return -1;
    const Geo &g = s->g;
    double **B = s->buf, **gr = s->grad, **pr = s->prod, **ph = s->phys;
    const int nz = g.nz;
    { // u in physical space (postp.f:23); its x-means, taken out of it
        FieldBatch b = batch({B[0]});
        b.g[0] = ph[0];
        fft_dev(s, b, 1, nz);
    }
    double *const *sgs = s->sgs[0] ? s->sgs : nullptr;
    {
        Scope sc(s, "postp_means");
        launch_pp_means(g, ph[0], sgs, s->pp_um, s->pp_tm, s->pp_dy, s->pp_dz, 8. * atan(1.0) / s->P.ylen,
                        s->stream);
    }
    { // the perturbation u' back to spectral space (postp.f:466)
        FieldBatch b = batch({ph[0]});
        b.g[0] = pr[0];
        fft_dev(s, b, -1, nz);
    }
    // v, w, temp as energycalc_postp transforms them: after one round trip,
    // temp after two
    project_dev(s, B + 1, 3);
    project_dev(s, B + 3, 1);
    double *spec[4] = {pr[0], B[1], B[2], B[3]};
    double *gx[4] = {gr[0], gr[3], gr[6], pr[1]}, *gy[4] = {gr[1], gr[4], gr[7], pr[2]};
    double *gz[4] = {gr[2], gr[5], gr[8], pr[3]};
    {
        Scope sc(s, "postp_grad");
        for (int q = 0; q < 4; q++) { // postp.f:471-489
            launch_grad_xy(g, spec[q], gx[q], gy[q], s->stream);
            launch_dpdz(g, spec[q], gz[q], s->stream);
        }
    }
    FieldBatch b = batch({spec[0], spec[1], spec[2], spec[3], gx[0], gy[0], gz[0], gx[1], gy[1], gz[1], gx[2],
                          gy[2], gz[2], gx[3], gy[3], gz[3]});
    for (int q = 0; q < 4; q++) b.g[q] = ph[q];
    fft_dev(s, b, 1, nz); // postp.f:493-508
    double *g12[12] = {gx[0], gy[0], gz[0], gx[1], gy[1], gz[1], gx[2], gy[2], gz[2], gx[3], gy[3], gz[3]};
    {
        Scope sc(s, "postp_sums");
        launch_pp_sums(g, ph, g12, sgs, s->pp_um, s->pp_dy, s->pp_dz, s->pp_tm, s->pp_part, s->P.xnu,
                       s->P.xkappa, (double)istep, s->t, out, s->stream);
    }
    // the state: u,v,w after two round trips, temp after three
    project_dev(s, B, 4);
    project_dev(s, B, 1);
    HIPCHK(hipGetLastError());
    return 0;
}

// speclegencalc of spectral field f; out (device) nzloc*nsubd, hdr (device,
// may be null) istep and t.  f takes the round trip of filtering.f:1152, 1206.
static int speclegen_dev(iw_solver *s, double *f, int istep, double *hdr, double *out)
{
    FieldBatch b = batch({f});
    b.g[0] = s->phys[0];
    fft_dev(s, b, 1, s->g.nz);
    {
        Scope sc(s, "speclegen");
        launch_speclegen(s->g, s->phys[0], (double)istep, s->t, hdr, out, s->stream);
    }
    project_dev(s, &f, 1);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int iw_postp(iw_solver *s, double *out)
{
    if (!s) return fail("null solver");
    if (pp_alloc(s) || postp_dev(s, s->istep, s->pp_rec)) return -1;
    double r[18];
    HIPCHK(hipMemcpyAsync(r, s->pp_rec, sizeof r, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (out) memcpy(out, r + 2, 16 * sizeof(double));
    return 0;
}

extern "C" int iw_speclegen(iw_solver *s, int idx, double *out)
{
    if (!s) return fail("null solver");
    if (idx < 0 || idx > 3) return fail("iw_speclegen: field index must be 0..3 (u, v, w, temp)");
    if (s->g.ny < 2) return fail("iw_speclegen: needs ny >= 2 (speclegencalc reads the row j = ny/2)");
    if (pp_alloc(s)) return -1;
    const int n = s->g.nz;
    if (speclegen_dev(s, s->buf[idx], s->istep, nullptr, s->pp_rec + 20)) return -1;
    std::vector<double> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), s->pp_rec + 20, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (out) memcpy(out, h.data(), sizeof(double) * n);
    return 0;
}

// kinetic and potential energy of the device state (CalcE/integration.f:
// 14-226 in 3-D, integrcol quadrature); the state is not changed
extern "C" int iw_kepe(iw_solver *s, double *out)
{
    if (!s || !out) return fail("iw_kepe: null argument");
    const Geo &g = s->g;
    if (!s->kp) {
        HIPCHK(hipMalloc((void **)&s->kp, sizeof(double) * (4 * (size_t)g.nz + 4)));
        s->dev_bytes += (long long)(sizeof(double) * (4 * (size_t)g.nz + 4));
    }
    FieldBatch b = batch({s->buf[0], s->buf[1], s->buf[2], s->buf[3]});
    for (int q = 0; q < 4; q++) b.g[q] = s->phys[q];
    fft_dev(s, b, 1, g.nz);
    const iw_params &P = s->P;
    const double span = P.xlen * (g.ny > 1 ? P.ylen : 1.0);
    const double cpe = 1. / 2. / P.rho0 * P.grav * P.grav / (P.brunt * P.brunt);
    {
        Scope sc(s, "kepe");
        launch_kepe(g, s->phys, P.rho0 / 2., cpe, span, s->kp, s->kp + 4 * (size_t)g.nz, s->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->kp + 4 * (size_t)g.nz, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// KE/PE without allocation.  The per-level partial sums and the four results
// live in a file-scope device array (one copy per GPU, held until the process
// ends); its address on each device is looked up once, at the first
// iw_create on that device, and only that address reaches kernels and
// copies.  Every slot the kernels read is written by them first in the same
// call, so nothing needs clearing.  Calls that overlap on one GPU share the
// array and race.
// ---------------------------------------------------------------------------
constexpr int kKepeScratch = 8192; // doubles: 4 nz + 4 for nz <= 2047
static __device__ __align__(256) double scratch[kKepeScratch];
constexpr int kMaxDev = 64;
static double *g_scratch_ptr[kMaxDev] = {};
static std::mutex g_scratch_mu;

static int scratch_init(int dev)
{
    if (dev < 0 || dev >= kMaxDev) return fail("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
    return 0;
}

extern "C" int iw_kepe_noalloc(iw_solver *s, double *out)
{
    if (!s || !out) return fail("iw_kepe_noalloc: null argument");
    const Geo &g = s->g;
    if (4 * (long)g.nz + 4 > kKepeScratch)
        return fail("iw_kepe_noalloc: nz = " + std::to_string(g.nz) + " needs " + std::to_string(4 * g.nz + 4) +
                    " doubles of scratch, beyond the static " + std::to_string(kKepeScratch));
    double *ptr = (s->device >= 0 && s->device < kMaxDev) ? g_scratch_ptr[s->device] : nullptr;
    if (!ptr) return fail("iw_kepe_noalloc: no scratch address for this device");
    FieldBatch b = batch({s->buf[0], s->buf[1], s->buf[2], s->buf[3]});
    for (int q = 0; q < 4; q++) b.g[q] = s->phys[q];
    fft_dev(s, b, 1, g.nz);
    const iw_params &P = s->P;
    const double span = P.xlen * (g.ny > 1 ? P.ylen : 1.0);
    const double cpe = 1. / 2. / P.rho0 * P.grav * P.grav / (P.brunt * P.brunt);
    {
        Scope sc(s, "kepe");
        launch_kepe(g, s->phys, P.rho0 / 2., cpe, span, ptr, ptr + 4 * (size_t)g.nz, s->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ptr + 4 * (size_t)g.nz, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// The thread count of the horizontal transforms' blocks (each holds 8
// values per thread, so it bounds the rows and columns per block).
bool iw::fft_threads_ok(const Geo &g, int ft)
{
    if (ft != 64 && ft != 128 && ft != 256) return false;
    if (2 * g.M > 8 * ft) return false;                      // two rows per block (forward with flux products)
    if (g.ndims == 3 && g.ny > 8 * ft) return false;         // one column per block
    return true;
}

extern "C" int iw_set_fft_threads(iw_solver *s, int ft)
{
    if (!s) return fail("iw_set_fft_threads: null solver");
    if (!fft_threads_ok(s->g, ft))
        return fail("iw_set_fft_threads: " + std::to_string(ft) + " threads do not fit this grid (64, 128 or 256 with "
                    "8 ft >= nx and 8 ft >= ny)");
    HIPCHK(hipStreamSynchronize(s->stream));
    s->g.ft = ft;
    return 0;
}

extern "C" int iw_set_mixed(iw_solver *s, int on)
{
    if (!s) return fail("iw_set_mixed: null solver");
    if (on) {
        const Geo &g = s->g;
        if (g.ndims != 3) return fail("iw_set_mixed: the mixed path is 3-D only (ny >= 2)");
        if (!solve_np_supported(g.nzloc, g.nsubd))
            return fail("iw_set_mixed: GLL order must lie in 8..24");
        if (mean_active(s)) return fail("iw_set_mixed: mean-wave and shear-current terms are FP64 only");
        if (s->closure) return fail("iw_set_mixed: the learned closure runs on the plain FP64 path");
    }
    if (on && !s->s32[0]) {
        for (int q = 0; q < 12; q++) HIPCHK(hipMalloc((void **)&s->s32[q], sizeof(float) * (size_t)s->g.ntot));
        s->dev_bytes += 12 * (long long)sizeof(float) * s->g.ntot;
    }
    s->mixed = on != 0;
    return 0;
}

// Mixed runs keep S, Sn, Snm in FP32 (s32); outside a run the FP64 buffers
// hold them.  At the start of a mixed run Sn, Snm are rounded to FP32 (S is
// formed anew in the step); at its end, on every return, all three are
// widened back into buf[12..23].
struct MixedHistories {
    iw_solver *s;
    explicit MixedHistories(iw_solver *s_) : s(s_)
    {
        if (!s->mixed) return;
        for (int i = 0; i < 8; i++) launch_cvt(s->buf[16 + i], s->s32[4 + i], s->g.ntot, s->stream);
    }
    ~MixedHistories()
    {
        if (!s->mixed) return;
        for (int i = 0; i < 12; i++) launch_cvt(s->s32[i], s->buf[12 + i], s->g.ntot, s->stream);
    }
};

extern "C" int iw_energy_records(iw_solver *s, double *out, int n)
{
    if (!s) return fail("null solver");
    const int have = (int)(s->e_hist.size() / 18);
    if (n < 0 || n > have) return fail("iw_energy_records: more records asked for than there are");
    if (n > 0 && !out) return fail("iw_energy_records: null argument");
    memcpy(out, s->e_hist.data(), sizeof(double) * 18 * (size_t)n);
    return have;
}

extern "C" int iw_speclegen_records(iw_solver *s, double *out, int n)
{
    if (!s) return fail("null solver");
    const size_t m = 2 + 4 * (size_t)s->g.nz;
    const int have = (int)(s->l_hist.size() / m);
    if (n < 0 || n > have) return fail("iw_speclegen_records: more records asked for than there are");
    if (n > 0 && !out) return fail("iw_speclegen_records: null argument");
    memcpy(out, s->l_hist.data(), sizeof(double) * m * (size_t)n);
    return have;
}

extern "C" int iw_set_meanwave(iw_solver *s, const double *mw)
{
    if (!s || !mw) return fail("iw_set_meanwave: null argument");
    const size_t n = 9 * (size_t)s->g.nxpp * s->g.nz;
    HIPCHK(hipMemcpyAsync(s->mw, mw, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->mw_on = false;
    for (size_t i = 0; i < n && !s->mw_on; i++) s->mw_on = mw[i] != 0.0;
    return 0;
}

extern "C" int iw_get_meanwave(iw_solver *s, double *mw)
{
    if (!s || !mw) return fail("iw_get_meanwave: null argument");
    const size_t n = 9 * (size_t)s->g.nxpp * s->g.nz;
    HIPCHK(hipMemcpyAsync(mw, s->mw, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_maxcalc(iw_solver *s, int istep, double *rec)
{
    if (!s) return fail("null solver");
    maxcalc_dev(s, istep);
    HIPCHK(hipGetLastError());
    if (mc_fetch(s)) return -1;
    adapt_dt(s, s->h_mc);
    if (rec) memcpy(rec, s->h_mc, sizeof(double) * kMcRec);
    return 0;
}

extern "C" int iw_maxcalc_records(iw_solver *s, double *out, int n)
{
    if (!s || (n > 0 && !out)) return fail("iw_maxcalc_records: null argument");
    if (n < 0 || (size_t)n * kMcRec > s->mc_hist.size())
        return fail("iw_maxcalc_records: more records asked for than the last iw_run produced");
    memcpy(out, s->mc_hist.data(), sizeof(double) * kMcRec * (size_t)n);
    return 0;
}

extern "C" int iw_get_dtstate(iw_solver *s, double *dt, double *cdt, int *ivardt, int *idtstep)
{
    if (!s) return fail("null solver");
    if (dt) *dt = s->P.dt;
    if (cdt) *cdt = s->cdt;
    if (ivardt) *ivardt = s->ivardt;
    if (idtstep) *idtstep = s->idtstep;
    return 0;
}

extern "C" int iw_set_dtstate(iw_solver *s, double cdt, int ivardt, int idtstep)
{
    if (!s) return fail("null solver");
    s->cdt = cdt;
    s->ivardt = ivardt;
    s->idtstep = idtstep;
    return 0;
}

// ---------------------------------------------------------------------------
// device-resident step loop (main.f:653-1066)
// ---------------------------------------------------------------------------
extern "C" int iw_set_initial(iw_solver *s, const double *u, const double *v, const double *w,
                              const double *temp)
{
    if (!s || !u || !v || !w || !temp) return fail("iw_set_initial: null argument");
    size_t fb = sizeof(double) * s->g.ntot;
    const double *src[4] = {u, v, w, temp};
    for (int q = 0; q < 24; q++) HIPCHK(hipMemsetAsync(s->buf[q], 0, fb, s->stream));
    for (int q = 0; q < 4; q++)
        HIPCHK(hipMemcpyAsync(s->buf[q], src[q], fb, hipMemcpyHostToDevice, s->stream));
    batch_fft(s, {s->buf[0], s->buf[1], s->buf[2], s->buf[3]}, -1);
    project_dev(s, s->buf, 4); // plotf of the initial field (main.f:537-544)
    HIPCHK(hipStreamSynchronize(s->stream));
    s->t = 0.;
    s->fac[0] = 1.; s->fac[1] = 0.; s->fac[2] = 0.; // main.f:543-552
    s->bdf[0] = 1.; s->bdf[1] = 0.; s->bdf[2] = 0.;
    s->bdfv = 1.;
    s->istep = 0;
    s->istart = 0;
    return 0;
}

// ---------------------------------------------------------------------------
// Restart files in the record layout of output.f:27-104 / restart.f:26-115:
// Fortran sequential unformatted records (4-byte length, payload, 4-byte
// length), REAL = 8 bytes, INTEGER = 4 bytes.  29 records of set-up data
// (which restart.f reads and main.f:591 then recomputes with setup), the
// time level, then 20 fields: u,v,w,temp, the n-1 and n-2 levels and the NL
// terms at n-1 and n-2.
// ---------------------------------------------------------------------------
static const int kRestartMeta = 29;

struct RecFile {
    FILE *f = nullptr;
    bool ok = true;
    ~RecFile() { if (f) fclose(f); }
    void put(const void *p, size_t n)
    {
        const int m = (int)n;
        ok = ok && fwrite(&m, 4, 1, f) == 1 && (n == 0 || fwrite(p, 1, n, f) == n) && fwrite(&m, 4, 1, f) == 1;
    }
    void putd(const std::vector<double> &v) { put(v.data(), sizeof(double) * v.size()); }
    // one record into out (its length must be known in advance unless out is null: skipped)
    long get(std::vector<char> &buf)
    {
        int m = 0, m2 = 0;
        if (fread(&m, 4, 1, f) != 1 || m < 0) return -1;
        buf.resize((size_t)m);
        if (m && fread(buf.data(), 1, (size_t)m, f) != (size_t)m) return -1;
        if (fread(&m2, 4, 1, f) != 1 || m2 != m) return -1;
        return m;
    }
};

// host copy of the set-up tables of a parameter block (no device needed)
static int host_tables(const iw_params *p, iw_solver &h)
{
    if (p->nzloc < 3 || p->nzloc > kMaxNl - 1 || p->nx < 4 || p->ny < 1 || p->nsubd < 1 || p->nsubd > IW_MAXSUB)
        return fail("restart file: parameter block out of range");
    h.P = *p;
    setup_tables(&h);
    return 0;
}

// the mean-wave records (output.f:55-59 / restart.f:65-69): pairs of
// /sol2dvar/ arrays interleaved per (i,k), or one array
static const int kMwRec[5][2] = {{0, 1}, {2, 3}, {4, 5}, {6, -1}, {7, 8}};

static int write_restart(const char *path, const iw_params *p, iw_solver &h, double t, double umax,
                         const double *const *fields, const double *mw = nullptr)
{
    const int nx = p->nx, ny = p->ny, nl = p->nzloc, ns = p->nsubd, nz = nl * ns, nzm = nl - 1;
    const int nxpp = nx + 2, nxhp = nx / 2 + 1;
    const long nxppy = (long)nxpp * ny, ntot = nxppy * nz;
    const int nz2 = ns * (nl - 1) + 1, nsdtype = 1;
    const double pi2 = 8. * atan(1.0);
    RecFile r;
    r.f = fopen(path, "wb");
    if (!r.f) return fail(std::string("cannot write restart file ") + path);
    r.put(&t, 8);
    const double prra[2] = {p->xnu, 0.};
    r.put(prra, 16);
    const int dims[9] = {nx, ny, nz, nzm, (int)nxppy, (int)ntot, ns, nl, nsdtype};
    r.put(dims, sizeof(dims));
    const double ab[2] = {pi2 / p->xlen, pi2 / p->ylen};
    r.put(ab, 16);
    // (xw(i),xsq(i),i=1,nx): entries past nxhp lie outside xw in the
    // reference's /wave/ block and are recomputed by setup; written as zero
    std::vector<double> v(2 * (size_t)nx, 0.);
    for (int i = 0; i < nxhp && i < nx; i++) { v[2 * i] = h.xw[i]; v[2 * i + 1] = h.xsq[i]; }
    r.putd(v);
    v.assign(2 * (size_t)ny, 0.);
    for (int j = 0; j < ny; j++) { v[2 * j] = h.yw[j]; v[2 * j + 1] = h.ysq[j]; }
    r.putd(v);
    v.assign(nx, 0.);
    for (int i = 0; i < nx; i++) v[i] = i * p->xlen / nx;
    r.putd(v);
    v.assign(ny, 0.);
    for (int j = 0; j < ny; j++) v[j] = j * p->ylen / ny;
    r.putd(v);
    r.putd(h.z);
    v.assign(nz2, 0.);
    for (int ks = 0; ks < ns; ks++)
        for (int k = 0; k < nl; k++) v[ks * (nl - 1) + k] = h.z[ks * nl + k];
    r.putd(v);
    std::vector<char> sub((size_t)ns * 20);
    for (int ks = 0; ks < ns; ks++) {
        const int it = 1;
        memcpy(&sub[ks * 20], &p->z0[ks], 8);
        memcpy(&sub[ks * 20 + 8], &p->zh[ks], 8);
        memcpy(&sub[ks * 20 + 16], &it, 4);
    }
    r.put(sub.data(), sub.size());
    v.assign(nsdtype, p->zlen / ns);
    r.putd(v);
    std::vector<double> al((size_t)nl * nl);
    for (int k = 0; k < nl; k++)
        for (int j = 0; j < nl; j++) al[(size_t)k * nl + j] = h.alpol[(size_t)j * nl + k]; // alpol(j,k), j fastest
    r.putd(al);
    r.putd(std::vector<double>((size_t)nl * nl, 0.)); // dalpol, recomputed by setup
    r.putd(h.zpts);
    r.putd(h.wg);
    std::vector<double> dd((size_t)nl * nl);
    for (int j = 0; j < nl; j++)
        for (int i = 0; i < nl; i++) dd[(size_t)j * nl + i] = h.d[(size_t)i * nl + j]; // d(i,j), i fastest
    r.putd(dd);
    for (int j = 0; j < nl; j++)
        for (int i = 0; i < nl; i++) dd[(size_t)j * nl + i] = h.d2[(size_t)i * nl + j];
    r.putd(dd);
    r.putd(std::vector<double>((size_t)nl * nl, 0.));          // d3
    r.putd(std::vector<double>(3 * (size_t)nz, 0.));           // atpz, btpz, ctpz
    r.putd(std::vector<double>(2 * (size_t)nxppy, 0.));        // tubc, tbbc (homogeneous walls)
    r.putd(std::vector<double>(12 * (size_t)nxppy, 0.));       // uw, vw, ww
    r.put(&umax, 8);
    const double solks = 0.;
    r.put(&solks, 8);
    { // mean-wave fields (zero when none were given)
        const size_t nm = (size_t)nxpp * nz;
        for (int rr = 0; rr < 5; rr++) {
            const int a = kMwRec[rr][0], b = kMwRec[rr][1];
            std::vector<double> v(b < 0 ? nm : 2 * nm, 0.);
            if (mw)
                for (size_t m = 0; m < nm; m++) {
                    if (b < 0) {
                        v[m] = mw[a * nm + m];
                    } else {
                        v[2 * m] = mw[a * nm + m];
                        v[2 * m + 1] = mw[b * nm + m];
                    }
                }
            r.putd(v);
        }
    }
    for (int q = 0; q < 20; q++) r.put(fields[q], sizeof(double) * (size_t)ntot);
    if (!r.ok) return fail(std::string("write error on restart file ") + path);
    return 0;
}

// state index of restart field q (the file holds u..tnm, then sun..stnm)
static int restart_state_index(int q) { return q < 12 ? q : q + 4; }

static int read_restart(const char *path, const iw_params *p, double *t, double *umax, double *const *fields,
                        double *mw = nullptr)
{
    const int nx = p->nx, ny = p->ny, nl = p->nzloc, nz = nl * p->nsubd;
    const long ntot = (long)(nx + 2) * ny * nz;
    RecFile r;
    r.f = fopen(path, "rb");
    if (!r.f) return fail(std::string("cannot open restart file ") + path);
    std::vector<char> b;
    for (int q = 0; q < kRestartMeta; q++) {
        const long n = r.get(b);
        if (n < 0) return fail(std::string("restart file ") + path + ": truncated set-up records");
        if (q == 0 && n == 8) memcpy(t, b.data(), 8);
        if (q == 2) {
            int d[9];
            if (n != 36) return fail("restart file: bad dimension record");
            memcpy(d, b.data(), 36);
            if (d[0] != nx || d[1] != ny || d[2] != nz || d[7] != nl)
                return fail("restart file: dimensions differ from the parameter block");
        }
        if (q == 22 && n == 8) memcpy(umax, b.data(), 8);
        if (q >= 24 && q < 29 && mw) { // mean-wave fields
            const size_t nm = (size_t)(nx + 2) * nz;
            const int a = kMwRec[q - 24][0], c = kMwRec[q - 24][1];
            if ((size_t)n != sizeof(double) * (c < 0 ? nm : 2 * nm))
                return fail("restart file: bad mean-wave record");
            const double *v = reinterpret_cast<const double *>(b.data());
            for (size_t m = 0; m < nm; m++) {
                if (c < 0) {
                    mw[a * nm + m] = v[m];
                } else {
                    mw[a * nm + m] = v[2 * m];
                    mw[c * nm + m] = v[2 * m + 1];
                }
            }
        }
    }
    for (int q = 0; q < 20; q++) {
        const long n = r.get(b);
        if (n != (long)sizeof(double) * ntot) return fail(std::string("restart file ") + path + ": bad field record");
        memcpy(fields[q], b.data(), (size_t)n);
    }
    return 0;
}

extern "C" int iw_write_restart_file(const char *path, const iw_params *p, double t, const double *state)
{
    if (!path || !p || !state) return fail("iw_write_restart_file: null argument");
    iw_solver h;
    if (host_tables(p, h)) return -1;
    const long ntot = (long)(p->nx + 2) * p->ny * p->nzloc * p->nsubd;
    const double *f[20];
    for (int q = 0; q < 20; q++) f[q] = state + (size_t)restart_state_index(q) * ntot;
    return write_restart(path, p, h, t, p->umax, f);
}

extern "C" int iw_read_restart_file(const char *path, const iw_params *p, double *t, double *umax, double *state)
{
    if (!path || !p || !t || !umax || !state) return fail("iw_read_restart_file: null argument");
    const long ntot = (long)(p->nx + 2) * p->ny * p->nzloc * p->nsubd;
    double *f[20];
    for (int q = 0; q < 20; q++) f[q] = state + (size_t)restart_state_index(q) * ntot;
    for (int q = 12; q < 16; q++) memset(state + (size_t)q * ntot, 0, sizeof(double) * ntot);
    return read_restart(path, p, t, umax, f);
}

extern "C" int iw_save_restart(iw_solver *s, const char *path)
{
    if (!s || !path) return fail("iw_save_restart: null argument");
    const size_t fb = sizeof(double) * (size_t)s->g.ntot;
    std::vector<double> host(20 * (size_t)s->g.ntot);
    const double *f[20];
    for (int q = 0; q < 20; q++) {
        HIPCHK(hipMemcpyAsync(host.data() + (size_t)q * s->g.ntot, s->buf[restart_state_index(q)], fb,
                              hipMemcpyDeviceToHost, s->stream));
        f[q] = host.data() + (size_t)q * s->g.ntot;
    }
    std::vector<double> mw(9 * (size_t)s->g.nxpp * s->g.nz);
    HIPCHK(hipMemcpyAsync(mw.data(), s->mw, sizeof(double) * mw.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return write_restart(path, &s->P, *s, s->t, s->P.umax, f, mw.data());
}

extern "C" int iw_load_restart(iw_solver *s, const char *path, double cdt)
{
    if (!s || !path) return fail("iw_load_restart: null argument");
    if (!(cdt > 0.)) return fail("iw_load_restart: cdt must be positive");
    const size_t fb = sizeof(double) * (size_t)s->g.ntot;
    std::vector<double> host(20 * (size_t)s->g.ntot);
    double *f[20];
    for (int q = 0; q < 20; q++) f[q] = host.data() + (size_t)q * s->g.ntot;
    double t = 0., umax = s->P.umax;
    std::vector<double> mw(9 * (size_t)s->g.nxpp * s->g.nz, 0.);
    if (read_restart(path, &s->P, &t, &umax, f, mw.data())) return -1;
    if (iw_set_meanwave(s, mw.data())) return -1; // restart.f:65-69
    for (int q = 0; q < 24; q++) HIPCHK(hipMemsetAsync(s->buf[q], 0, fb, s->stream));
    for (int q = 0; q < 20; q++)
        HIPCHK(hipMemcpyAsync(s->buf[restart_state_index(q)], f[q], fb, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // restart.f:69 overwrites umax; main.f:619-634 first-step coefficients
    s->P.umax = umax;
    s->t = t;
    s->cdt = cdt;
    const double dt = s->P.dt, d2 = cdt * dt;
    s->fac[0] = iw_abvar1(dt, d2, d2); s->fac[1] = iw_abvar2(dt, d2, d2); s->fac[2] = iw_abvar3(dt, d2, d2);
    s->bdf[0] = iw_bdfvar1(dt, d2, d2); s->bdf[1] = iw_bdfvar2(dt, d2, d2); s->bdf[2] = iw_bdfvar3(dt, d2, d2);
    s->bdfv = iw_bdfvarv(dt, d2, d2);
    s->istep = 0;
    s->istart = 1;
    // a restarted reference run has no SGS tensors yet (/comsgs/ is zero)
    for (int q = 0; q < 9; q++)
        if (s->sgs[q]) HIPCHK(hipMemsetAsync(s->sgs[q], 0, fb, s->stream));
    project_dev(s, s->buf, 4); // plotf of the restart field (main.f:599-608)
    // postp (main.f:634-635): its record is the one energy record until the
    // next iw_run
    if (ring_reserve(s, &s->e_ring, &s->e_cap, 18) || postp_dev(s, 0, s->e_ring)) return -1;
    s->e_hist.assign(18, 0.);
    HIPCHK(hipMemcpyAsync(s->e_hist.data(), s->e_ring, 18 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->l_hist.clear();
    return 0;
}

// filtering(isgscalc) (filtering.f:8-93) on the device state: the
// estimation model applied to u,v,w in physical space, and with isgscalc the
// SGS tensors of sgscalc.f:4-223
static int rho_mean(iw_solver *s)
{ // sgscalc.f:21-28: rho's level means through its horfft round trip
    batch_fft(s, {s->buf[3]}, -1);
    launch_plane_mean(s->g, s->buf[3], s->tmean, s->stream);
    batch_fft(s, {s->buf[3]}, 1);
    return 0;
}

// the nine /comsgs/ tensors (zero) and the level means of rho, on first use
static int sgs_alloc(iw_solver *s)
{
    if (s->sgs[0]) return 0;
    const size_t fb = sizeof(double) * (size_t)s->g.ntot;
    for (int q = 0; q < 9; q++) {
        HIPCHK(hipMalloc((void **)&s->sgs[q], fb));
        HIPCHK(hipMemset(s->sgs[q], 0, fb));
    }
    HIPCHK(hipMalloc((void **)&s->tmean, sizeof(double) * s->g.nz));
    s->dev_bytes += 9 * (long long)fb + sizeof(double) * s->g.nz;
    return 0;
}

static int filtering_dev(iw_solver *s, int isgscalc)
{
    const Geo &g = s->g;
    if (!solve_np_supported(s->P.nzloc, s->P.nsubd))
        return fail("estimation filter: GLL order must lie in 8..24");
    double **B = s->buf;
    double *uvwt[4] = {B[0], B[1], B[2], B[3]};
    double *u0 = s->grad[0];
    if (isgscalc && sgs_alloc(s)) return -1;
    Scope sc(s, "estimation_filter");
    batch_fft(s, {B[0], B[1], B[2], B[3]}, 1);
    if (isgscalc) {
        rho_mean(s);
        launch_sgs_pre(g, uvwt, s->tmean, s->sgs, s->ox, s->stream);
    }
    // modecut of u, v, w (filtering.f:43-47) written on levels 2..nz
    // (w: 2..nz-1) of the 1-based reference index (filtering.f:59-77)
    launch_modecut(g, B[0], u0, B[0], 1, g.nz - 1, s->stream);
    launch_modecut(g, B[1], u0, B[1], 1, g.nz - 1, s->stream);
    launch_modecut(g, B[2], u0, B[2], 1, g.nz - 2, s->stream);
    if (isgscalc) {
        double **tp = s->grad + 1;
        const int srcs[6] = {0, 3, 5, 1, 2, 4}; // t11 t22 t33 t12 t13 t23 (sgscalc.f:71-82)
        for (int q = 0; q < 6; q++) launch_modecut(g, s->sgs[srcs[q]], u0, tp[q], 0, g.nz - 1, s->stream);
        launch_sgs_post(g, uvwt, s->tmean, s->sgs, tp, s->ox, 0, s->stream);
        rho_mean(s);
        for (int q = 0; q < 3; q++) launch_modecut(g, s->sgs[6 + q], u0, tp[q], 0, g.nz - 1, s->stream);
        launch_sgs_post(g, uvwt, s->tmean, s->sgs, tp, s->ox, 1, s->stream);
    }
    batch_fft(s, {B[0], B[1], B[2], B[3]}, -1);
    return hipGetLastError() == hipSuccess ? 0 : fail("filtering launch failed");
}

extern "C" int iw_filtering(iw_solver *s, int isgscalc)
{
    if (!s) return fail("null solver");
    if (filtering_dev(s, isgscalc)) return -1;
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_get_sgs(iw_solver *s, int idx, double *host)
{
    if (!s || !host) return fail("iw_get_sgs: null argument");
    if (idx < 0 || idx > 8) return fail("iw_get_sgs: index must lie in [0, 8]");
    if (!s->sgs[idx]) return fail("iw_get_sgs: no SGS tensors computed yet (filtering with isgscalc = 1)");
    HIPCHK(hipMemcpyAsync(host, s->sgs[idx], sizeof(double) * (size_t)s->g.ntot, hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_set_sgs(iw_solver *s, int idx, const double *host)
{
    if (!s || !host) return fail("iw_set_sgs: null argument");
    if (idx < 0 || idx > 8) return fail("iw_set_sgs: index must lie in [0, 8]");
    if (sgs_alloc(s)) return -1;
    HIPCHK(hipMemcpyAsync(s->sgs[idx], host, sizeof(double) * (size_t)s->g.ntot, hipMemcpyHostToDevice,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_run(iw_solver *s, int nsteps, int *counts)
{
    if (!s || nsteps < 0) return fail("iw_run: bad argument");
    const bool mc = s->P.imaxcalc || s->P.dtadapt;
    if (s->mixed && mean_active(s)) return fail("iw_run: the mixed path runs without mean-wave terms");
    const MixedHistories mixed_hist(s);
    struct FuseScope { // convec may form the pressure right-hand side for the calcuvwt of its step
        iw_solver *s;
        explicit FuseScope(iw_solver *s_) : s(s_) { s->fuse_prhs = true; }
        ~FuseScope() { s->fuse_prhs = s->phi_ready = false; }
    } fuse_scope(s);
    const bool adapt = s->P.dtadapt != 0;
    s->mc_hist.clear();
    if (mc && !adapt && nsteps > s->mc_ring_cap) { // device ring of the step records
        hipFree(s->mc_ring);
        s->mc_ring = nullptr;
        s->mc_ring_cap = 0;
        HIPCHK(hipMalloc((void **)&s->mc_ring, sizeof(double) * kMcRec * (size_t)nsteps));
        s->mc_ring_cap = nsteps;
    }
    // output routines of the step loop: speclegencalc (main.f:905-911),
    // postp (main.f:935-936), plotf (main.f:996-1039); their records go to
    // device rings, read back after the loop
    const int npp = s->P.npostp, nleg = s->P.ncalclegspec, n3d = s->P.n3doutput;
    const int ne = npp > 0 ? (s->istep + nsteps) / npp - s->istep / npp : 0;
    const int nlg = nleg > 0 ? (s->istep + nsteps) / nleg - s->istep / nleg : 0;
    const size_t lrec = 2 + 4 * (size_t)s->g.nz;
    if (ring_reserve(s, &s->e_ring, &s->e_cap, 18 * (size_t)ne) ||
        ring_reserve(s, &s->l_ring, &s->l_cap, lrec * (size_t)nlg))
        return -1;
    if ((ne > 0 || nlg > 0) && pp_alloc(s)) return -1;
    int en = 0, ln = 0;
    for (int n = 0; n < nsteps; n++) {
        int istep = ++s->istep;
        // maxcalc (main.f:657): its horfft round trip (various.f:739-742,
        // 854-857) changes the spectrum only through the kx=0 and kx-Nyquist
        // columns; its reductions come from convec's own inverse transforms
        // of the same fields, and the adaptive step and main.f:660-664 are
        // taken there before dt or the coefficients are read
        project_dev(s, s->buf, 4);
        if (!mc && s->ivardt == 1) s->idtstep++; // main.f:660-664
        if (s->mixed ? convec_mixed(s, s->P.ispecfilternl, mc ? istep : 0, n)
                     : convec_dev(s, s->P.ispecfilternl, mc ? istep : 0, n))
            return -1;
        const bool ldtadapt = s->ivardt == 1 && s->idtstep < 2;
        if (istep < 2 && s->istart == 0) { // main.f:700-777
            s->fac[0] = 2.; s->fac[1] = -1.; s->fac[2] = 0.;
            s->bdf[0] = 2.; s->bdf[1] = -0.5; s->bdf[2] = 0.;
        } else if (istep < 2 || ldtadapt) {
            // variable-step AB3/BDF3 after a restart (main.f:719-734) or on
            // the step after a change of dt (main.f:757-776)
            const double dt = s->P.dt;
            var_coeffs(s, dt, dt, s->cdt * dt);
        } else {
            s->fac[0] = 3.; s->fac[1] = -3.; s->fac[2] = 1.;
            s->bdf[0] = 3.; s->bdf[1] = -1.5; s->bdf[2] = 1. / 3.;
        }
        int cpen = (s->P.ispecfilterexp == 1) && (istep % s->P.nspecfilterexp == 0);
        if (calcuvwt_dev(s, cpen)) return -1;
        if (istep < 2) { // main.f:813-856
            s->bdfv = (s->istart == 0) ? 1.5 : iw_bdfvarv(s->P.dt, s->P.dt, s->cdt * s->P.dt);
        } else if (!(s->ivardt == 1 && s->idtstep < 2)) {
            s->bdfv = 11. / 6.;
        } else {
            s->bdfv = iw_bdfvarv(s->P.dt, s->P.dt, s->cdt * s->P.dt);
            s->cdt = 1.;
            s->ivardt = 0;
        }
        if (s->P.ifilter == 1 && s->P.nstepfilter > 0 && istep % s->P.nstepfilter == 0) // main.f:880-895
            if (filtering_dev(s, s->P.npostp > 0 && istep % s->P.npostp == 0)) return -1;
        s->t = s->t + s->P.dt;
        if (nleg > 0 && istep % nleg == 0) {
            double *r = s->l_ring + lrec * (size_t)ln++;
            for (int q = 0; q < 4; q++)
                if (speclegen_dev(s, s->buf[q], istep, q == 0 ? r : nullptr, r + 2 + (size_t)q * s->g.nz))
                    return -1;
        }
        bool avg = s->P.navgintrfc > 0 && istep % s->P.navgintrfc == 0;
        if (avg) {
            if (iface_dev(s, s->buf, 4)) return -1;
        }
        if (counts) {
            if (avg) {
                HIPCHK(hipMemcpyAsync(s->h_count, s->d_count, 4 * sizeof(int), hipMemcpyDeviceToHost,
                                      s->stream));
                HIPCHK(hipStreamSynchronize(s->stream));
                for (int q = 0; q < 4; q++) counts[4 * n + q] = s->h_count[q];
            } else {
                for (int q = 0; q < 4; q++) counts[4 * n + q] = 0;
            }
        }
        if (npp > 0 && istep % npp == 0)
            if (postp_dev(s, istep, s->e_ring + 18 * (size_t)en++)) return -1;
        if (n3d > 0 && istep % n3d == 0) project_dev(s, s->buf, 4);
    }
    HIPCHK(hipGetLastError());
    s->e_hist.assign(18 * (size_t)en, 0.);
    s->l_hist.assign(lrec * (size_t)ln, 0.);
    if (en > 0)
        HIPCHK(hipMemcpyAsync(s->e_hist.data(), s->e_ring, sizeof(double) * s->e_hist.size(), hipMemcpyDeviceToHost,
                              s->stream));
    if (ln > 0)
        HIPCHK(hipMemcpyAsync(s->l_hist.data(), s->l_ring, sizeof(double) * s->l_hist.size(), hipMemcpyDeviceToHost,
                              s->stream));
    if (en > 0 || ln > 0) HIPCHK(hipStreamSynchronize(s->stream));
    if (mc && !adapt && nsteps > 0) { // records of a fixed-dt run
        s->mc_hist.resize((size_t)nsteps * kMcRec);
        HIPCHK(hipMemcpyAsync(s->mc_hist.data(), s->mc_ring, sizeof(double) * kMcRec * (size_t)nsteps,
                              hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int n = 0; n < nsteps; n++) adapt_dt(s, s->mc_hist.data() + (size_t)n * kMcRec);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// reference routines on host arrays
// ---------------------------------------------------------------------------
extern "C" int iw_horfft(iw_solver *s, double *b, int is, int nk)
{
    if (!s || !b) return fail("iw_horfft: null argument");
    if (nk < 1 || nk > s->g.nz) return fail("iw_horfft: nk out of range");
    if (is != 1 && is != -1) return fail("iw_horfft: is must be +1 or -1");
    size_t bytes = sizeof(double) * s->g.nxppy * nk;
    HIPCHK(hipMemcpyAsync(s->hbuf, b, bytes, hipMemcpyHostToDevice, s->stream));
    horfft_dev(s, batch({s->hbuf}), is, nk);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b, s->hbuf, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// penalty_setup + the boundary part of penalty_main (penalty.f:9-55,
// 183-235): alpha, beta, gamma, delta, tau1, tau2 of every subdomain
extern "C" int iw_penalty_params(iw_solver *s, double eps, int ifbcbot, int ifbctop, double *alpha, double *beta,
                                 double *gamma, double *delta, double *tau1, double *tau2)
{
    if (!s || !alpha || !beta || !gamma || !delta || !tau1 || !tau2) return fail("iw_penalty_params: null argument");
    if (ifbcbot < 0 || ifbcbot > 2 || ifbctop < 0 || ifbctop > 2)
        return fail("iw_penalty_params: boundary types must be 0 (Dirichlet), 1 (Neumann) or 2 (Robin)");
    const Pen p = penalty_params(s->P, eps, ifbcbot, ifbctop);
    for (int k = 0; k < s->P.nsubd; k++) {
        alpha[k] = p.alpha[k];
        beta[k] = p.beta[k];
        gamma[k] = p.gamma[k];
        delta[k] = p.delta[k];
        tau1[k] = p.tau1[k];
        tau2[k] = p.tau2[k];
    }
    return 0;
}

// param_set (penalty.f:57-157): the parameters of one boundary of subdomain
// height ybar
extern "C" int iw_param_set(iw_solver *s, int ifbc, double eps, double ybar, double *alpha0, double *beta0,
                            double *tau0)
{
    if (!s || !alpha0 || !beta0 || !tau0) return fail("iw_param_set: null argument");
    if (ifbc < 0 || ifbc > 2) return fail("iw_param_set: ifbc must be 0, 1 or 2");
    const double omega = 2. / ((s->P.nzloc + 1) * (s->P.nzloc + 2));
    param_set(s->P, omega, ifbc, eps, ybar, *alpha0, *beta0, *tau0);
    return 0;
}

// helmv_penalty (penalty.f:255-400) on one column pair: the boundary rows of
// the right-hand side take tau (bc) (Dirichlet) or -tau (bc) eps / +tau (bc)
// eps (Neumann, bottom / top) as in penalty.f:583-604, the matrix of
// matrix_elmsetup is assembled and factored with row interchanges
// (SGBTF2 order, k_factor) and both columns solved (SGBTRS) on the device.
// rhsr, rhsi are the right-hand sides as helmv_penalty receives them
// (penalty_main has multiplied them by -eps); the penalty parameters are
// those of penalty_setup + penalty_main for (eps, ifbcbot, ifbctop).
extern "C" int iw_penalty_solve(iw_solver *s, double eps, int ifbcbot, int ifbctop, double bcrbot, double bcibot,
                                double bcrtop, double bcitop, const double *rhsr, const double *rhsi, double *fr,
                                double *fi)
{
    if (!s || !rhsr || !rhsi || !fr || !fi) return fail("iw_penalty_solve: null argument");
    if (ifbcbot < 0 || ifbcbot > 1 || ifbctop < 0 || ifbctop > 1)
        return fail("iw_penalty_solve: boundary types must be 0 (Dirichlet) or 1 (Neumann)");
    if (!(eps > 0.)) return fail("iw_penalty_solve: epsilon must be positive");
    const iw_params &P = s->P;
    const int nz = s->g.nz, ns = P.nsubd, ldab = s->sys.ldab;
    const Pen p = penalty_params(P, eps, ifbcbot, ifbctop);
    std::vector<double> coef((size_t)12 * ns), info(4), x((size_t)2 * nz);
    assembly_coef(P, p, eps, coef.data());
    info[0] = eps;
    info[1] = p.tau1[0];
    info[2] = p.tau2[ns - 1];
    info[3] = ifbcbot * 4 + ifbctop;
    for (int k = 0; k < nz; k++) {
        x[k] = rhsr[k];
        x[nz + k] = rhsi[k];
    }
    if (ifbcbot == BC_DIR) {
        x[0] = rhsr[0] + p.tau1[0] * bcrbot;
        x[nz] = rhsi[0] + p.tau1[0] * bcibot;
    } else {
        x[0] = rhsr[0] - p.tau1[0] * bcrbot * eps;
        x[nz] = rhsi[0] - p.tau1[0] * bcibot * eps;
    }
    if (ifbctop == BC_DIR) {
        x[nz - 1] = rhsr[nz - 1] + p.tau2[ns - 1] * bcrtop;
        x[2 * nz - 1] = rhsi[nz - 1] + p.tau2[ns - 1] * bcitop;
    } else {
        x[nz - 1] = rhsr[nz - 1] + p.tau2[ns - 1] * bcrtop * eps;
        x[2 * nz - 1] = rhsi[nz - 1] + p.tau2[ns - 1] * bcitop * eps;
    }
    if (!s->pen_buf) {
        // coef, sysinfo, band, solution | pivots
        const size_t nd = (size_t)12 * ns + 4 + (size_t)ldab * nz + 2 * (size_t)nz;
        HIPCHK(hipMalloc((void **)&s->pen_buf, sizeof(double) * nd));
        HIPCHK(hipMalloc((void **)&s->pen_piv, sizeof(int) * (size_t)nz));
        s->dev_bytes += (long long)(sizeof(double) * nd + sizeof(int) * nz);
    }
    double *dcoef = s->pen_buf, *dinfo = dcoef + 12 * ns, *dab = dinfo + 4, *dx = dab + (size_t)ldab * nz;
    HIPCHK(hipMemcpyAsync(dcoef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(dinfo, info.data(), sizeof(double) * 4, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(dx, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, s->stream));
    // band positions the factorisation does not write read as zero
    HIPCHK(hipMemsetAsync(dab, 0, sizeof(double) * (size_t)ldab * nz, s->stream));
    launch_factor_one(s->g, ldab, dcoef, dinfo, dab, s->pen_piv, s->stream);
    launch_band_solve1(s->g, ldab, dab, s->pen_piv, dx, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(x.data(), dx, sizeof(double) * x.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int k = 0; k < nz; k++) {
        fr[k] = x[k];
        fi[k] = x[nz + k];
    }
    return 0;
}

extern "C" int iw_dpdz(iw_solver *s, const double *a, double *b)
{
    if (!s || !a || !b) return fail("iw_dpdz: null argument");
    size_t bytes = sizeof(double) * s->g.ntot;
    HIPCHK(hipMemcpyAsync(s->hbuf, a, bytes, hipMemcpyHostToDevice, s->stream));
    launch_dpdz(s->g, s->hbuf, s->ox, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b, s->ox, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_filter_3dlegenfourier(iw_solver *s, double *f, int ifourier, int ilegen)
{
    if (!s || !f) return fail("iw_filter_3dlegenfourier: null argument");
    size_t bytes = sizeof(double) * s->g.ntot;
    HIPCHK(hipMemcpyAsync(s->hbuf, f, bytes, hipMemcpyHostToDevice, s->stream));
    launch_filter3d(s->g, s->hbuf, ifourier == 1, ilegen == 1, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, s->hbuf, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_interface_avg3d(iw_solver *s, double *f, int *ic)
{
    if (!s || !f) return fail("iw_interface_avg3d: null argument");
    size_t bytes = sizeof(double) * s->g.ntot;
    HIPCHK(hipMemcpyAsync(s->hbuf, f, bytes, hipMemcpyHostToDevice, s->stream));
    double *one[1] = {s->hbuf};
    if (iface_dev(s, one, 1)) return -1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, s->hbuf, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_count, s->d_count, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (ic) *ic = s->h_count[0];
    return 0;
}

extern "C" int iw_convec(iw_solver *s, int istep, int ispecfilternl)
{
    if (!s) return fail("null solver");
    (void)istep;
    if (convec_dev(s, ispecfilternl)) return -1;
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_calcuvwt(iw_solver *s, int cpenspecfilter)
{
    if (!s) return fail("null solver");
    if (calcuvwt_dev(s, cpenspecfilter != 0)) return -1;
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}
