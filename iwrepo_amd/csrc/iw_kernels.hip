// Device kernels of the stratified-Boussinesq step for MI355X (gfx950).
//
// Reference: jmsegret/IWrepo, mode1_bbl (Fortran, FP64).  Each kernel cites the
// reference loop it replaces.  All arithmetic is FP64.  Layout of every field
// is the reference's a(nxpp,ny,nz): a[i + nxpp*(j + ny*k)], x fastest; the
// spectral plane is nxhp*ny contiguous complex values.
#include <hip/hip_runtime.h>

#include "iw_internal.h"

namespace iw {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Radix-2 Stockham autosort FFT on nrows rows of length M held in LDS.
// sign -1: sum g_j e^{-2 pi i jk/M}; +1: e^{+...}.  tw[t] = e^{-2 pi i t/M}.
// Returns the buffer holding the result (natural order).
__device__ double2 *stockham(double2 *in, double2 *out, int M, int logM, int nrows,
                             const double2 *__restrict__ tw, int sign)
{
    const int half = M >> 1;
    const int nb = nrows * half;
    for (int s = 0; s < logM; s++) {
        const int Ns = 1 << s;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            int r = b / half;
            int j = b - r * half;
            int k = j & (Ns - 1);
            double2 a = in[r * M + j];
            double2 c = in[r * M + j + half];
            double2 w = tw[k << (logM - 1 - s)];
            if (sign > 0) w.y = -w.y;
            double2 cw = cmul(c, w);
            int o = r * M + ((j - k) << 1) + k;
            out[o] = make_double2(a.x + cw.x, a.y + cw.y);
            out[o + Ns] = make_double2(a.x - cw.x, a.y - cw.y);
        }
        __syncthreads();
        double2 *t = in;
        in = out;
        out = t;
    }
    return in;
}

// ---------------------------------------------------------------------------
// x transform (FFT991, fft.f:1361-1502, as used by horfft.f:21 / 58):
// real length-N rows <-> N/2+1 coefficients A_k + i B_k, forward normalised
// by 1/N with B_0 = B_{N/2} = 0 (fft.f:1293-1305), inverse ignoring Im of the
// k=0 and k=N/2 coefficients (fft.f:1218-1230) and zeroing the 2 pad words
// (fft.f:1483-1490).  Implemented as one complex FFT of length M=N/2 on the
// even/odd-packed row with LDS-staged radix-2 passes.
// ---------------------------------------------------------------------------
template <int DIR>
__global__ __launch_bounds__(256) void k_fft_x(FieldBatch fb, int nrows, int nxpp, int M,
                                               int logM, int R, const double2 *__restrict__ twM,
                                               const double2 *__restrict__ twN, double scaley)
{
    extern __shared__ double2 smx[];
    double2 *b0 = smx, *b1 = smx + R * M;
    double *base = fb.f[blockIdx.y];
    const int row0 = blockIdx.x * R;
    const int N = 2 * M;
    if (DIR < 0) {
        for (int idx = threadIdx.x; idx < R * M; idx += blockDim.x) {
            int r = idx / M, m = idx - r * M, row = row0 + r;
            double2 z = make_double2(0., 0.);
            if (row < nrows) {
                const double *p = base + (long)row * nxpp + 2 * m;
                z = make_double2(p[0], p[1]);
            }
            b0[idx] = z;
        }
    } else {
        for (int idx = threadIdx.x; idx < R * M; idx += blockDim.x) {
            int r = idx / M, k = idx - r * M, row = row0 + r;
            double2 Z = make_double2(0., 0.);
            if (row < nrows) {
                const double *p = base + (long)row * nxpp;
                double ar, ai, br, bi;
                if (k == 0) {
                    ar = p[0]; ai = 0.; br = p[2 * M]; bi = 0.;
                } else {
                    ar = p[2 * k]; ai = p[2 * k + 1];
                    br = p[2 * (M - k)]; bi = -p[2 * (M - k) + 1];
                }
                double er = ar + br, ei = ai + bi;
                double dr = ar - br, di = ai - bi;
                double2 w = twN[k];
                double c = w.x, s = -w.y; // e^{+2 pi i k/N}
                double orr = dr * c - di * s, oi = dr * s + di * c;
                Z = make_double2(er - oi, ei + orr);
            }
            b0[idx] = Z;
        }
    }
    __syncthreads();
    double2 *res = stockham(b0, b1, M, logM, R, twM, DIR);
    if (DIR < 0) {
        const double sc = 1.0 / (double)N;
        for (int idx = threadIdx.x; idx < R * (M + 1); idx += blockDim.x) {
            int r = idx / (M + 1), k = idx - r * (M + 1), row = row0 + r;
            if (row >= nrows) continue;
            double2 Zk = res[r * M + (k == M ? 0 : k)];
            double2 Zm = res[r * M + (k == 0 ? 0 : M - k)];
            double er = 0.5 * (Zk.x + Zm.x), ei = 0.5 * (Zk.y - Zm.y);
            double dr = Zk.x - Zm.x, di = Zk.y + Zm.y;
            double orr = 0.5 * di, oi = -0.5 * dr;
            double2 w = twN[k];
            double xr = er + (orr * w.x - oi * w.y);
            double xi = ei + (orr * w.y + oi * w.x);
            double A = sc * xr, B = sc * xi;
            if (k == 0 || k == M) B = 0.0;
            if (k == M) A = A * scaley; // horfft.f:28-30 on the untransformed Nyquist column
            double *p = base + (long)row * nxpp;
            p[2 * k] = A;
            p[2 * k + 1] = B;
        }
    } else {
        for (int idx = threadIdx.x; idx < R * M; idx += blockDim.x) {
            int r = idx / M, m = idx - r * M, row = row0 + r;
            if (row >= nrows) continue;
            double *p = base + (long)row * nxpp;
            p[2 * m] = res[idx].x;
            p[2 * m + 1] = res[idx].y;
        }
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            int row = row0 + r;
            if (row < nrows) {
                double *p = base + (long)row * nxpp;
                p[N] = 0.;
                p[N + 1] = 0.;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// y transform (CFFT99, fft.f:1-406, horfft.f:25 / 55): complex length-ny
// transform of columns ii < nxh (the kx-Nyquist column is skipped); forward
// scaled by 1/ny (horfft.f:28-30).  CB columns per block staged in LDS.
// ---------------------------------------------------------------------------
template <int DIR>
__global__ __launch_bounds__(256) void k_fft_y(FieldBatch fb, int nxpp, int nxh, int ny, int logNy,
                                               int CB, const double2 *__restrict__ twY, double scaley)
{
    extern __shared__ double2 smy[];
    double2 *b0 = smy, *b1 = smy + CB * ny;
    double *plane = fb.f[blockIdx.z] + (long)blockIdx.y * nxpp * ny;
    const int ii0 = blockIdx.x * CB;
    for (int idx = threadIdx.x; idx < CB * ny; idx += blockDim.x) {
        int j = idx / CB, c = idx - j * CB, ii = ii0 + c;
        double2 z = make_double2(0., 0.);
        if (ii < nxh) {
            const double *p = plane + (long)j * nxpp + 2 * ii;
            z = make_double2(p[0], p[1]);
        }
        b0[c * ny + j] = z;
    }
    __syncthreads();
    double2 *res = stockham(b0, b1, ny, logNy, CB, twY, DIR);
    for (int idx = threadIdx.x; idx < CB * ny; idx += blockDim.x) {
        int j = idx / CB, c = idx - j * CB, ii = ii0 + c;
        if (ii >= nxh) continue;
        double2 z = res[c * ny + j];
        if (DIR < 0) {
            z.x = scaley * z.x;
            z.y = scaley * z.y;
        }
        double *p = plane + (long)j * nxpp + 2 * ii;
        p[0] = z.x;
        p[1] = z.y;
    }
}

void launch_fft_x(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st)
{
    int R = 512 / g.M;
    if (R < 1) R = 1;
    int nrows = g.ny * nk;
    dim3 grid((nrows + R - 1) / R, b.n);
    size_t sh = 2 * (size_t)R * g.M * sizeof(double2);
    if (dir < 0)
        hipLaunchKernelGGL(k_fft_x<-1>, grid, dim3(256), sh, st, b, nrows, g.nxpp, g.M, g.logM, R,
                           g.twM, g.twN, g.scaley);
    else
        hipLaunchKernelGGL(k_fft_x<1>, grid, dim3(256), sh, st, b, nrows, g.nxpp, g.M, g.logM, R,
                           g.twM, g.twN, g.scaley);
}

void launch_fft_y(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st)
{
    if (g.ndims != 3) return;
    int CB = 1024 / g.ny;
    if (CB < 1) CB = 1;
    if (CB > g.nxh) CB = g.nxh;
    dim3 grid((g.nxh + CB - 1) / CB, nk, b.n);
    size_t sh = 2 * (size_t)CB * g.ny * sizeof(double2);
    if (dir < 0)
        hipLaunchKernelGGL(k_fft_y<-1>, grid, dim3(256), sh, st, b, g.nxpp, g.nxh, g.ny, g.logNy, CB,
                           g.twY, g.scaley);
    else
        hipLaunchKernelGGL(k_fft_y<1>, grid, dim3(256), sh, st, b, g.nxpp, g.nxh, g.ny, g.logNy, CB,
                           g.twY, g.scaley);
}

// ---------------------------------------------------------------------------
// z derivative dpdz (derv.f:212-275): out = (2/zh) D in per subdomain, with
// the input optionally a product c*a*b formed on load (solver.f:627-668,
// 1085-1136).  One thread per (i,j) column, column held in registers.
// ---------------------------------------------------------------------------
struct DpdzArgs {
    const double *a[4];
    const double *b[4];
    double c[4];
    double *out[4];
    int nq;
};

__global__ __launch_bounds__(256) void k_dpdz(DpdzArgs A, long nxppy, int nl,
                                              const double *__restrict__ D,
                                              const double *__restrict__ zh)
{
    long ij = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxppy) return;
    const int ks = blockIdx.y;
    const double fac = 2. / zh[ks];
    for (int q = 0; q < A.nq; q++) {
        double col[kMaxNl];
#pragma unroll
        for (int t = 0; t < kMaxNl; t++) {
            if (t < nl) {
                long o = (long)(ks * nl + t) * nxppy + ij;
                double v = A.a[q][o];
                if (A.b[q]) v = (A.c[q] * v) * A.b[q][o];
                col[t] = v;
            }
        }
        for (int k1 = 0; k1 < nl; k1++) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < kMaxNl; t++)
                if (t < nl) s = s + D[k1 * nl + t] * col[t];
            A.out[q][(long)(ks * nl + k1) * nxppy + ij] = s * fac;
        }
    }
}

static void run_dpdz(const Geo &g, const DpdzArgs &a, hipStream_t st)
{
    dim3 grid((unsigned)((g.nxppy + 255) / 256), g.nsubd);
    hipLaunchKernelGGL(k_dpdz, grid, dim3(256), 0, st, a, g.nxppy, g.nzloc, g.D, g.zh);
}

void launch_dpdz(const Geo &g, const double *a, double *b, hipStream_t st)
{
    DpdzArgs A = {};
    A.a[0] = a;
    A.b[0] = nullptr;
    A.out[0] = b;
    A.nq = 1;
    run_dpdz(g, A, st);
}

void launch_dpdz3(const Geo &g, double *const *in3, double *const *out3, hipStream_t st)
{
    DpdzArgs A = {};
    for (int q = 0; q < 3; q++) {
        A.a[q] = in3[q];
        A.b[q] = nullptr;
        A.out[q] = out3[q];
    }
    A.nq = 3;
    run_dpdz(g, A, st);
}

// solver.f:627-668 (d(-uw/2)/dz, d(-vw/2)/dz, d(-ww/2)/dz) and 1085-1093 (d(-w rho)/dz)
void launch_dpdz_conv(const Geo &g, double *const *uvwt, double *const *out4, hipStream_t st)
{
    DpdzArgs A = {};
    A.a[0] = uvwt[0]; A.b[0] = uvwt[2]; A.c[0] = -0.5;
    A.a[1] = uvwt[1]; A.b[1] = uvwt[2]; A.c[1] = -0.5;
    A.a[2] = uvwt[2]; A.b[2] = uvwt[2]; A.c[2] = -0.5;
    A.a[3] = uvwt[2]; A.b[3] = uvwt[3]; A.c[3] = -1.0;
    for (int q = 0; q < 4; q++) A.out[q] = out4[q];
    A.nq = 4;
    run_dpdz(g, A, st);
}

// mean w = 0 (solver.f:66-69)
__global__ void k_zero_meanw(double *w, long nxppy, int nz)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nz) {
        w[(long)k * nxppy] = 0.;
        w[(long)k * nxppy + 1] = 0.;
    }
}

__global__ void k_zero_int(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero_int(int *p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_int, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_zero_meanw(const Geo &g, double *w, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_meanw, dim3((g.nz + 63) / 64), dim3(64), 0, st, w, g.nxppy, g.nz);
}

// spectral x/y gradients i kx f, i ky f of u,v,w (solver.f:86-103)
__global__ __launch_bounds__(256) void k_grad(const double *__restrict__ u, const double *__restrict__ v,
                                              const double *__restrict__ w, double *dudx, double *dudy,
                                              double *dvdx, double *dvdy, double *dwdx, double *dwdy,
                                              long ncplx, int nxhp, int ny,
                                              const double *__restrict__ xw,
                                              const double *__restrict__ yw)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncplx) return;
    int ii = (int)(c % nxhp);
    int j = (int)((c / nxhp) % ny);
    double kx = xw[ii], ky = yw[j];
    long o = 2 * c;
    double2 fu = make_double2(u[o], u[o + 1]);
    double2 fv = make_double2(v[o], v[o + 1]);
    double2 fw = make_double2(w[o], w[o + 1]);
    dudx[o] = -kx * fu.y; dudx[o + 1] = kx * fu.x;
    dudy[o] = -ky * fu.y; dudy[o + 1] = ky * fu.x;
    dvdx[o] = -kx * fv.y; dvdx[o + 1] = kx * fv.x;
    dvdy[o] = -ky * fv.y; dvdy[o + 1] = ky * fv.x;
    dwdx[o] = -kx * fw.y; dwdx[o + 1] = kx * fw.x;
    dwdy[o] = -ky * fw.y; dwdy[o + 1] = ky * fw.x;
}

void launch_grad(const Geo &g, double *const *uvw, double *const *gr, hipStream_t st)
{
    long ncplx = g.ntot / 2;
    hipLaunchKernelGGL(k_grad, dim3((unsigned)((ncplx + 255) / 256)), dim3(256), 0, st, uvw[0], uvw[1],
                       uvw[2], gr[0], gr[1], gr[3], gr[4], gr[6], gr[7], ncplx, g.nxhp, g.ny, g.xw, g.yw);
}

// ---------------------------------------------------------------------------
// rhsfm physical-space part (solver.f:671-702 advection and buoyancy,
// 708-839 interface upwind penalty, 941-956 virtual paddle, 1136-1148
// density flux) plus the seven horizontal flux products (solver.f:975-1063,
// 1155-1161).  Elementwise over all points.
// ---------------------------------------------------------------------------
struct RhsArgs {
    const double *u, *v, *w, *t;
    const double *g[9];
    double *S[4]; // in: su,sv,sw = vertical flux derivatives; out: su..st
    const double *ox;
    double *pr[7];
    double omz, bcoef, rg, omega;
    double t_now, xacrit_umax, umax_zdim, xkcrit, xomega, xmcrit, phit, rfac, inv_xkcrit;
    int paddle;
};

__global__ __launch_bounds__(256) void k_rhs_phys(RhsArgs A, long ntot, long nxppy, int nxpp, int nx,
                                                  int nz, int nl, const double *__restrict__ zh,
                                                  const double *__restrict__ zc,
                                                  const double *__restrict__ xg)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ntot) return;
    const int k = (int)(n / nxppy);
    const int i = (int)(n % nxpp);
    double u = A.u[n], v = A.v[n], w = A.w[n], tp = A.t[n];
    double su = A.S[0][n], sv = A.S[1][n], sw = A.S[2][n], st = 0.0;
    su = su - 0.5 * (u * A.g[0][n] + v * (A.g[1][n] - 4. * A.omz) + w * A.g[2][n]);
    sv = sv - 0.5 * (u * (A.g[3][n] + 4. * A.omz) + v * A.g[4][n] + w * A.g[5][n]);
    sw = sw - 0.5 * (u * A.g[6][n] + v * A.g[7][n] + w * A.g[8][n]);
    sw = sw + A.bcoef * tp;
    const int k1 = k % nl, ks = k / nl;
    if (k1 == 0) {
        double wb = (k == 0) ? 0. : w;
        if (wb > 0.) {
            double tau1 = 0.5 / A.omega * (2. / zh[ks]);
            long m = n - nxppy;
            su = su - tau1 * (wb * u - wb * A.u[m]);
            sv = sv - tau1 * (wb * v - wb * A.v[m]);
            sw = sw - tau1 * (wb * w - wb * A.w[m]);
            st = st - tau1 * (wb * tp - wb * A.t[m]);
        }
    }
    if (k1 == nl - 1) {
        double wt = (k == nz - 1) ? 0. : w;
        if (wt < 0.) {
            double g0 = fabs(wt);
            double tau2 = 0.5 / A.omega * (2. / zh[ks]);
            long m = n + nxppy;
            su = su - tau2 * (g0 * u - g0 * A.u[m]);
            sv = sv - tau2 * (g0 * v - g0 * A.v[m]);
            sw = sw - tau2 * (g0 * w - g0 * A.w[m]);
            st = st - tau2 * (g0 * tp - g0 * A.t[m]);
        }
    }
    if (A.paddle && i < nx) {
        double zz = zc[k], xx = xg[i];
        double arg = A.xkcrit * xx - A.xomega * A.t_now;
        double wzp = A.xmcrit * cos(A.xmcrit * zz), wz = sin(A.xmcrit * zz);
        su = su + (-A.inv_xkcrit) * A.xacrit_umax * wzp * sin(arg) * A.umax_zdim * A.phit;
        sw = sw + A.xacrit_umax * wz * cos(arg) * A.umax_zdim * A.phit;
        st = st + A.xacrit_umax * wz * sin(arg) * A.rfac * A.umax_zdim * A.phit;
    }
    st = st + A.ox[n] - w * A.rg;
    A.S[0][n] = su;
    A.S[1][n] = sv;
    A.S[2][n] = sw;
    A.S[3][n] = st;
    A.pr[0][n] = (-0.5 * u) * u;
    A.pr[1][n] = (-0.5 * u) * v;
    A.pr[2][n] = (-0.5 * v) * v;
    A.pr[3][n] = (-0.5 * u) * w;
    A.pr[4][n] = (-0.5 * v) * w;
    A.pr[5][n] = -(u * tp);
    A.pr[6][n] = -(v * tp);
}

void launch_rhs_phys(const Geo &g, const iw_params &P, double t, double *const *uvwt,
                     double *const *gr, double *const *S, const double *ox, double *const *pr,
                     hipStream_t st)
{
    RhsArgs A;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3];
    for (int q = 0; q < 9; q++) A.g[q] = gr[q];
    for (int q = 0; q < 4; q++) A.S[q] = S[q];
    A.ox = ox;
    for (int q = 0; q < 7; q++) A.pr[q] = pr[q];
    double rg = -P.rho0 * (P.brunt * P.brunt) / P.grav; // forcing.f:115-120
    A.omz = P.omz;
    A.bcoef = (P.brunt * P.brunt) / rg;                 // solver.f:693-696
    A.rg = rg;
    A.omega = 2. / ((g.nzloc + 1) * (g.nzloc + 2));     // solver.f:708
    double pi2 = 8. * atan(1.0);
    double xkcrit = pi2 / P.xlen, xmcrit = pi2 / (2. * P.zlen);
    double xomega = (P.brunt * xkcrit) / sqrt(xkcrit * xkcrit + xmcrit * xmcrit);
    A.t_now = t;
    A.xacrit_umax = P.xacrit * P.umax;
    A.umax_zdim = P.umax / P.xlen;
    A.xkcrit = xkcrit;
    A.xomega = xomega;
    A.xmcrit = xmcrit;
    A.inv_xkcrit = 1. / xkcrit;
    A.rfac = -(1. / xomega) * P.brunt * P.brunt * (P.rho0 / P.grav);
    A.phit = (t <= P.fwave * (pi2 / xomega)) ? 1. : 0.; // forcing.f:424-447
    A.paddle = (P.xacrit != 0.);
    hipLaunchKernelGGL(k_rhs_phys, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, A, g.ntot,
                       g.nxppy, g.nxpp, g.nx, g.nz, g.nzloc, g.zh, g.z, g.x);
}

// S += i (kx P_a + ky P_b) for the four equations (solver.f:985-1063, 1163-1196)
__global__ __launch_bounds__(256) void k_rhs_spec(double *su, double *sv, double *sw, double *st,
                                                  const double *__restrict__ p0, const double *__restrict__ p1,
                                                  const double *__restrict__ p2, const double *__restrict__ p3,
                                                  const double *__restrict__ p4, const double *__restrict__ p5,
                                                  const double *__restrict__ p6, long ncplx, int nxhp, int ny,
                                                  const double *__restrict__ xw,
                                                  const double *__restrict__ yw)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncplx) return;
    int ii = (int)(c % nxhp);
    int j = (int)((c / nxhp) % ny);
    double kx = xw[ii], ky = yw[j];
    long o = 2 * c;
#define IWADD(S, PA, PB)                                   \
    {                                                      \
        double xr = kx * PA[o] + ky * PB[o];               \
        double xi = kx * PA[o + 1] + ky * PB[o + 1];       \
        S[o] = S[o] + (-xi);                               \
        S[o + 1] = S[o + 1] + xr;                          \
    }
    IWADD(su, p0, p1)
    IWADD(sv, p1, p2)
    IWADD(sw, p3, p4)
    IWADD(st, p5, p6)
#undef IWADD
}

void launch_rhs_spec(const Geo &g, double *const *S, double *const *pr, hipStream_t st)
{
    long ncplx = g.ntot / 2;
    hipLaunchKernelGGL(k_rhs_spec, dim3((unsigned)((ncplx + 255) / 256)), dim3(256), 0, st, S[0], S[1],
                       S[2], S[3], pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], pr[6], ncplx, g.nxhp, g.ny,
                       g.xw, g.yw);
}

// wnl at bottom/top (solver.f:270-284)
__global__ void k_wnl(const double *sw, const double *swn, const double *swnm, double f1, double f2,
                      double f3, double *wnl, long nxppy, int nz)
{
    long ij = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxppy) return;
    long kt = (long)(nz - 1) * nxppy + ij;
    wnl[ij] = f1 * sw[ij] + f2 * swn[ij] + f3 * swnm[ij];
    wnl[nxppy + ij] = f1 * sw[kt] + f2 * swn[kt] + f3 * swnm[kt];
}

void launch_wnl(const Geo &g, const double *sw, const double *swn, const double *swnm, const double *fac,
                double *wnl, hipStream_t st)
{
    hipLaunchKernelGGL(k_wnl, dim3((unsigned)((g.nxppy + 255) / 256)), dim3(256), 0, st, sw, swn, swnm,
                       fac[0], fac[1], fac[2], wnl, g.nxppy, g.nz);
}

// rotvisc (solver.f:1210-1253) at the bottom point of subdomain 1 and the top
// point of subdomain nsubd for levels n, n-1, n-2 (solver.f:320-394)
struct RotArgs {
    const double *lev[9]; // u v w | un vn wn | unm vnm wnm
};

__global__ void k_rotvisc(RotArgs A, double f1, double f2, double f3, double xnu, double *wviscrot,
                          int nmode, int nxhp, int ny, int nl, int nsubd, long nxppy,
                          const double *__restrict__ D, const double *__restrict__ zh,
                          const double *__restrict__ xw, const double *__restrict__ yw)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 2 * nmode) return;
    int side = id / nmode, m = id - side * nmode;
    int ii = m % nxhp, j = m / nxhp;
    int ks = side == 0 ? 0 : nsubd - 1;
    int k0 = side == 0 ? 0 : nl - 1;
    long o = 2L * m;
    double kx = xw[ii], ky = yw[j];
    double fac2z = 2. / zh[ks];
    double rr[3], ri[3];
    for (int L = 0; L < 3; L++) {
        const double *u = A.lev[3 * L], *v = A.lev[3 * L + 1], *w = A.lev[3 * L + 2];
        double sur = 0., sui = 0., svr = 0., svi = 0.;
        for (int t = 0; t < nl; t++) {
            long n = (long)(ks * nl + t) * nxppy + o;
            double dd = D[k0 * nl + t];
            sur = sur + dd * u[n];
            sui = sui + dd * u[n + 1];
            svr = svr + dd * v[n];
            svi = svi + dd * v[n + 1];
        }
        double dfur = sur * fac2z, dfui = sui * fac2z, dfvr = svr * fac2z, dfvi = svi * fac2z;
        long nk = (long)(ks * nl + k0) * nxppy + o;
        double wr = w[nk], wi = w[nk + 1];
        rr[L] = kx * kx * wr + ky * ky * wr + (-kx * dfui) + (-ky * dfvi);
        ri[L] = kx * kx * wi + ky * ky * wi + kx * dfur + ky * dfvr;
    }
    double rfr = f1 * rr[0] + f2 * rr[1] + f3 * rr[2];
    double rfi = f1 * ri[0] + f2 * ri[1] + f3 * ri[2];
    wviscrot[side * nxppy + o] = -xnu * rfr;
    wviscrot[side * nxppy + o + 1] = -xnu * rfi;
}

void launch_rotvisc(const Geo &g, const iw_params &P, double *const *lev, const double *fac,
                    double *wviscrot, hipStream_t st)
{
    RotArgs A;
    for (int q = 0; q < 9; q++) A.lev[q] = lev[q];
    int nmode = g.nxhp * g.ny;
    hipLaunchKernelGGL(k_rotvisc, dim3((2 * nmode + 127) / 128), dim3(128), 0, st, A, fac[0], fac[1],
                       fac[2], P.xnu, wviscrot, nmode, g.nxhp, g.ny, g.nzloc, g.nsubd, g.nxppy, g.D,
                       g.zh, g.xw, g.yw);
}

// ---------------------------------------------------------------------------
// Column tile kernels: a block owns MB consecutive complex modes of one
// subdomain (nl planes x MB complex, coalesced rows of MB*16 bytes).
//   mode 0: filter only (filter_3dlegenfourier, filtering.f:811-983)
//   mode 1: SS extrapolation + BDF assembly (solver.f:294-316, 402-430)
//           followed by the filter (solver.f:483-490)
// ---------------------------------------------------------------------------
constexpr int kMB = 32;

struct TileArgs {
    const double *S[3];  // S, Sn, Snm
    const double *F[3];  // f, fn, fnm
    double *out;
    double f1, f2, f3, b1, b2, b3, dt;
    int mode, ifourier, ilegen;
};

__global__ __launch_bounds__(256) void k_tile(TileArgs A, long ncplx_plane, long nxppy, int nl,
                                              const double *__restrict__ fcf,
                                              const double *__restrict__ wg,
                                              const double *__restrict__ AL,
                                              const double *__restrict__ sig)
{
    __shared__ double tile[kMaxNl * 2 * kMB];
    __shared__ double coef[kMaxNl * 2 * kMB];
    const int ks = blockIdx.y;
    const long c0 = (long)blockIdx.x * kMB;
    const int nmb = (int)((ncplx_plane - c0) < kMB ? (ncplx_plane - c0) : kMB);
    const int W2 = 2 * kMB;
    for (int idx = threadIdx.x; idx < nl * W2; idx += blockDim.x) {
        int t = idx / W2, e = idx - t * W2;
        double val = 0.;
        if (e < 2 * nmb) {
            long o = (long)(ks * nl + t) * nxppy + 2 * c0 + e;
            if (A.mode == 1) {
                double s = A.f1 * A.S[0][o] + A.f2 * A.S[1][o] + A.f3 * A.S[2][o];
                val = A.dt * s + A.b1 * A.F[0][o] + A.b2 * A.F[1][o] + A.b3 * A.F[2][o];
            } else {
                val = A.out[o];
            }
        }
        tile[idx] = val;
    }
    __syncthreads();
    if (A.ifourier) { // Fourier magnitude/phase filter, filtering.f:852-901
        for (int idx = threadIdx.x; idx < nl * kMB; idx += blockDim.x) {
            int t = idx / kMB, cc = idx - t * kMB;
            if (cc >= nmb) continue;
            double fr = tile[t * W2 + 2 * cc], fi = tile[t * W2 + 2 * cc + 1];
            double fmag = sqrt(fr * fr + fi * fi);
            double cosc = fr / fmag, sinc = fi / fmag;
            if (fmag != 0.) {
                double phase = (sinc >= 0.) ? acos(cosc) : -acos(cosc);
                double fmagf = fmag * fcf[c0 + cc];
                fr = fmagf * cos(phase);
                fi = fmagf * sin(phase);
            } else {
                fr = 0.;
                fi = 0.;
            }
            tile[t * W2 + 2 * cc] = fr;
            tile[t * W2 + 2 * cc + 1] = fi;
        }
        __syncthreads();
    }
    if (A.ilegen) { // Legendre exponential filter, filtering.f:905-956
        const int nzm = nl - 1;
        for (int idx = threadIdx.x; idx < nl * W2; idx += blockDim.x) {
            int k = idx / W2, e = idx - k * W2;
            double sum = 0.;
            for (int t = 0; t < nl; t++) sum = sum + wg[t] * tile[t * W2 + e] * AL[k * nl + t];
            double a = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
            coef[idx] = a * sig[k];
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < nl * W2; idx += blockDim.x) {
            int t = idx / W2, e = idx - t * W2;
            double sum = 0.;
            for (int k = 0; k < nl; k++) sum = sum + coef[k * W2 + e] * AL[k * nl + t];
            tile[idx] = sum;
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nl * W2; idx += blockDim.x) {
        int t = idx / W2, e = idx - t * W2;
        if (e < 2 * nmb) A.out[(long)(ks * nl + t) * nxppy + 2 * c0 + e] = tile[idx];
    }
}

static void run_tile(const Geo &g, const TileArgs &A, hipStream_t st)
{
    long ncp = (long)g.nxhp * g.ny;
    dim3 grid((unsigned)((ncp + kMB - 1) / kMB), g.nsubd);
    hipLaunchKernelGGL(k_tile, grid, dim3(256), 0, st, A, ncp, g.nxppy, g.nzloc, g.fcf, g.wg, g.AL, g.sig);
}

void launch_bdf_filter(const Geo &g, const iw_params &P, double *const *S3, double *const *F3, double *out,
                       const double *fac, const double *bdf, double dt, int filt, hipStream_t st)
{
    TileArgs A = {};
    for (int q = 0; q < 3; q++) {
        A.S[q] = S3[q];
        A.F[q] = F3[q];
    }
    A.out = out;
    A.f1 = fac[0]; A.f2 = fac[1]; A.f3 = fac[2];
    A.b1 = bdf[0]; A.b2 = bdf[1]; A.b3 = bdf[2];
    A.dt = dt;
    A.mode = 1;
    A.ifourier = filt;
    A.ilegen = filt;
    (void)P;
    run_tile(g, A, st);
}

void launch_filter3d(const Geo &g, double *f, int ifourier, int ilegen, hipStream_t st)
{
    TileArgs A = {};
    A.out = f;
    A.mode = 0;
    A.ifourier = ifourier;
    A.ilegen = ilegen;
    run_tile(g, A, st);
}

// ---------------------------------------------------------------------------
// interface_avg3d physical-space pass (penalty.f:729-772): one thread per
// (i,j) point and interface; counts by wavefront ballot + one atomic.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_iface(FieldBatch fb, int *count, long nxppy, int nl, int nint)
{
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double *f = fb.f[blockIdx.y];
    bool hit = false;
    if (id < nxppy * nint) {
        int ks = 1 + (int)(id / nxppy);
        long ij = id % nxppy;
        long k = (long)ks * nl, km = (long)(ks - 1) * nl + nl - 1;
        double fsp = f[k * nxppy + ij], fsm = f[km * nxppy + ij];
        double favg2 = fsp + fsm;
        double c0 = fabs(fsp - fsm) / fabs(favg2);
        if (c0 > 0.005) {
            double ff = f[(km - 1) * nxppy + ij] * 0.5 + f[km * nxppy + ij] * 0. + f[k * nxppy + ij] * 0.
                        + f[(k + 1) * nxppy + ij] * 0.5;
            f[km * nxppy + ij] = ff;
            f[k * nxppy + ij] = ff;
            hit = true;
        }
    }
    unsigned long long b = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&count[blockIdx.y], __popcll(b));
}

void launch_iface(const Geo &g, FieldBatch b, int *count, hipStream_t st)
{
    int nint = g.nsubd - 1;
    if (nint <= 0) return;
    long n = g.nxppy * nint;
    dim3 grid((unsigned)((n + 255) / 256), b.n);
    hipLaunchKernelGGL(k_iface, grid, dim3(256), 0, st, b, count, g.nxppy, g.nzloc, nint);
}

// ---------------------------------------------------------------------------
// Banded LU with partial pivoting (SGBTF2, bandludcmp.f:238-440, reached via
// SGBTRF with NB=1, ludcmp.f:7478-7491) of the penalty operator assembled on
// the fly (matrix_elmsetup, penalty.f:403-634).  One wavefront per system;
// the active (kl+1) x (kv+1) window lives in LDS and slides down the band.
// Output: LAPACK band storage columns AB[row + ldab*col] and 0-based ipiv.
// Coefficients per subdomain (12): e2, t1a, cfb1, t2g, cft1, ht1a, cfb2,
// ht2g, cft2, cfl, cfr, unused.
// ---------------------------------------------------------------------------
constexpr int kWR = 32, kWC = 64, kWS = 65; // window rows/cols (power of 2), row stride

__device__ __forceinline__ double assemble(int r, int c, int nl, int ns, const double *__restrict__ cf,
                                           const double *__restrict__ Dl, const double *__restrict__ D2l)
{
    const int ks = r / nl, k1 = r - ks * nl, km = ks * nl, N = nl - 1;
    const double *q = cf + 12 * ks;
    if (c >= km && c < km + nl) {
        int k2 = c - km;
        double v = -(q[0] * D2l[k1 * nl + k2]);
        if (k1 == k2) v = v + 1.;
        if (k1 == 0) {
            if (k2 == 0) v = v + q[1];
            v = v + q[2] * Dl[k2];
        }
        if (k1 == N) {
            if (k2 == N) v = v + q[3];
            v = v + q[4] * Dl[N * nl + k2];
        }
        if (k1 == 0 && ks > 0) {
            if (k2 == 0) v = v - q[5];
            v = v + q[6] * Dl[k2];
        }
        if (k1 == N && ks < ns - 1) {
            if (k2 == N) v = v - q[7];
            v = v + q[8] * Dl[N * nl + k2];
        }
        return v;
    }
    if (k1 == 0 && ks > 0 && c >= km - nl && c < km) {
        int k2 = c - (km - nl);
        double v = (k2 == N) ? -q[5] : 0.;
        return v + q[9] * Dl[N * nl + k2];
    }
    if (k1 == N && ks < ns - 1 && c >= km + nl && c < km + 2 * nl) {
        int k2 = c - (km + nl);
        double v = (k2 == 0) ? (0. - q[7]) : 0.;
        return v + q[10] * Dl[k2];
    }
    return 0.;
}

__global__ __launch_bounds__(64) void k_factor(int first, int n, int nl, int ns, int ldab,
                                               const double *__restrict__ coef, const double *__restrict__ D,
                                               const double *__restrict__ D2, double *AB, int *ipiv)
{
    extern __shared__ double shf[];
    double *Wm = shf;                    // kWR * kWS
    double *Dl = Wm + kWR * kWS;         // nl*nl
    double *D2l = Dl + nl * nl;          // nl*nl
    double *cf = D2l + nl * nl;          // 12*ns
    const int sys = first + blockIdx.x;
    const int lane = threadIdx.x;
    const int L = nl, kv = 2 * nl;
    for (int e = lane; e < nl * nl; e += 64) {
        Dl[e] = D[e];
        D2l[e] = D2[e];
    }
    for (int e = lane; e < 12 * ns; e += 64) cf[e] = coef[(long)sys * 12 * ns + e];
    __syncthreads();
    double *ab = AB + (long)sys * ldab * n;
    int *ip = ipiv + (long)sys * n;
#define WM(r, c) Wm[((r) & (kWR - 1)) * kWS + ((c) & (kWC - 1))]
    for (int r = 0; r <= L && r < n; r++)
        if (lane <= kv) WM(r, lane) = (lane < n) ? assemble(r, lane, nl, ns, cf, Dl, D2l) : 0.;
    __syncthreads();
    int ju = 0;
    for (int j = 0; j < n; j++) {
        const int km = (L < n - 1 - j) ? L : n - 1 - j;
        double av = (lane <= km) ? fabs(WM(j + lane, j)) : -1.0;
        int ai = lane;
        for (int off = 32; off >= 1; off >>= 1) {
            double ov = __shfl_xor(av, off);
            int oi = __shfl_xor(ai, off);
            if (ov > av || (ov == av && oi < ai)) {
                av = ov;
                ai = oi;
            }
        }
        const int p = ai;
        if (lane == 0) ip[j] = j + p;
        const double piv = WM(j + p, j);
        __syncthreads();
        if (piv != 0.) {
            int jt = j + L + p;
            if (jt > n - 1) jt = n - 1;
            if (jt > ju) ju = jt;
            if (p != 0 && lane <= kv) {
                double a = WM(j, j + lane), b = WM(j + p, j + lane);
                WM(j, j + lane) = b;
                WM(j + p, j + lane) = a;
            }
            __syncthreads();
            if (km > 0) {
                const double rcp = 1. / WM(j, j);
                __syncthreads();
                if (lane >= 1 && lane <= km) WM(j + lane, j) = rcp * WM(j + lane, j);
                __syncthreads();
                if (lane >= 1 && lane <= ju - j) {
                    double yc = WM(j, j + lane);
                    if (yc != 0.)
                        for (int t = 1; t <= km; t++)
                            WM(j + t, j + lane) = WM(j + t, j + lane) + (-yc) * WM(j + t, j);
                }
                __syncthreads();
            }
        }
        if (lane <= kv && j + lane < n) ab[(kv - lane) + (long)ldab * (j + lane)] = WM(j, j + lane);
        if (lane >= 1 && lane <= km) ab[(kv + lane) + (long)ldab * j] = WM(j + lane, j);
        __syncthreads();
        const int rn = j + L + 1;
        if (lane >= 1 && lane <= L) WM(j + lane, j + 1 + kv) = 0.;
        if (rn < n && lane <= kv) {
            int c = j + 1 + lane;
            WM(rn, c) = (c < n) ? assemble(rn, c, nl, ns, cf, Dl, D2l) : 0.;
        }
        __syncthreads();
    }
#undef WM
}

void launch_factor(const Geo &g, const SysTables &S, int first, int count, hipStream_t st)
{
    if (count <= 0) return;
    size_t sh = sizeof(double) * ((size_t)kWR * kWS + 2 * g.nzloc * g.nzloc + 12 * g.nsubd);
    hipLaunchKernelGGL(k_factor, dim3(count), dim3(64), sh, st, first, g.nz, g.nzloc, g.nsubd, S.ldab,
                       S.coef, g.D, g.D2, S.AB, S.ipiv);
}

// ---------------------------------------------------------------------------
// Implicit step for one Fourier mode (calcuvwt.f:87-741 with update,
// calcuvwt.f:826-901): pressure Poisson, projection, u/v (shared factor),
// w and density Helmholtz solves, each through SGBTRS (bandludcmp.f:883-1069)
// on the cached factor, Legendre filter (filter_legenpen, filtering.f:729-808)
// and mean-u damping.  One wavefront per mode; columns staged in LDS; the
// band window of the forward/backward sweeps rides in the lanes.
// ---------------------------------------------------------------------------
constexpr int kCH = 16; // factor columns staged per chunk

struct ModeArgs {
    double *u, *v, *w, *t;
    const double *wnl, *wviscrot;
    const double *AB;
    const int *ipiv;
    const double *sysinfo; // eps, tau1 bottom, tau2 top, unused
    const int *modesys;
    const double *D, *AL, *wg, *sig, *zh, *xw, *yw, *damp;
    double dt, xnu, xkappa;
    long nxppy;
    int nz, nl, ns, nxhp, ldab, kl, kv, cpen, noslip;
};

template <int NR>
__device__ void band_solve(const ModeArgs &A, int sys, double *const *b, double *fch, int *ich)
{
    const int lane = threadIdx.x;
    const int n = A.nz, kl = A.kl, kv = A.kv, ldab = A.ldab;
    const double *ab = A.AB + (long)sys * ldab * n;
    const int *ip = A.ipiv + (long)sys * n;
    double v[NR];
    // forward: L with row interchanges (bandludcmp.f:1020-1030)
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = (lane <= kl && lane < n) ? b[r][lane] : 0.;
    for (int j = 0; j < n; j++) {
        const int c = j % kCH;
        if (c == 0) {
            __syncthreads();
            for (int e = lane; e < kCH * kl; e += 64) {
                int cc = e / kl, t = e - cc * kl;
                if (j + cc < n) fch[cc * kl + t] = ab[(kv + 1 + t) + (long)ldab * (j + cc)];
            }
            for (int e = lane; e < kCH; e += 64)
                if (j + e < n) ich[e] = ip[j + e];
            __syncthreads();
        }
        if (j < n - 1) {
            const int p = ich[c] - j;
            const int lm = (kl < n - 1 - j) ? kl : n - 1 - j;
#pragma unroll
            for (int r = 0; r < NR; r++) {
                if (p != 0) {
                    double x0 = __shfl(v[r], 0), xp = __shfl(v[r], p);
                    if (lane == 0) v[r] = xp;
                    else if (lane == p) v[r] = x0;
                }
                double bj = __shfl(v[r], 0);
                if (lane >= 1 && lane <= lm) v[r] = v[r] + fch[c * kl + lane - 1] * (-bj);
            }
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (lane == 0) b[r][j] = v[r];
            double nv = __shfl_down(v[r], 1);
            if (lane == kl) nv = (j + kl + 1 < n) ? b[r][j + kl + 1] : 0.;
            v[r] = nv;
        }
    }
    __syncthreads();
    // backward: upper band triangle with kv superdiagonals (STBSV, ludcmp.f:4698-5046)
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = (lane <= kv && n - 1 - lane >= 0) ? b[r][n - 1 - lane] : 0.;
    const int W = kv + 1;
    for (int j = n - 1; j >= 0; j--) {
        const int c = (n - 1 - j) % kCH;
        if (c == 0) {
            __syncthreads();
            for (int e = lane; e < kCH * W; e += 64) {
                int cc = e / W, t = e - cc * W;
                if (j - cc >= 0) fch[cc * W + t] = ab[t + (long)ldab * (j - cc)];
            }
            __syncthreads();
        }
        const double diag = fch[c * W + kv];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double bj = __shfl(v[r], 0);
            double x = (bj != 0.) ? bj / diag : bj;
            if (lane == 0) {
                v[r] = x;
                b[r][j] = x;
            } else if (lane <= kv && lane <= j) {
                v[r] = v[r] - x * fch[c * W + kv - lane];
            }
            double nv = __shfl_down(v[r], 1);
            if (lane == kv) nv = (j - 1 - kv >= 0) ? b[r][j - 1 - kv] : 0.;
            v[r] = nv;
        }
    }
    __syncthreads();
}

// Legendre filter of NC columns of nz values held in LDS, subdomain by subdomain
__device__ void filter_cols(const ModeArgs &A, double *const *col, int NC, double *tmp)
{
    const int lane = threadIdx.x, nl = A.nl, nzm = nl - 1;
    for (int q = 0; q < NC; q += 2) {
        const int nq = (NC - q) < 2 ? (NC - q) : 2;
        for (int ks = 0; ks < A.ns; ks++) {
            const int comp = lane >> 5, k = lane & 31;
            if (comp < nq && k < nl) {
                const double *f = col[q + comp] + ks * nl;
                double sum = 0.;
                for (int t = 0; t < nl; t++) sum = sum + A.wg[t] * f[t] * A.AL[k * nl + t];
                double a = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
                tmp[comp * 32 + k] = a * A.sig[k];
            }
            __syncthreads();
            if (comp < nq && k < nl) {
                double sum = 0.;
                for (int kk = 0; kk < nl; kk++) sum = sum + tmp[comp * 32 + kk] * A.AL[kk * nl + k];
                col[q + comp][ks * nl + k] = sum;
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(64) void k_calc_mode(ModeArgs A)
{
    extern __shared__ double shm[];
    const int nz = A.nz, nl = A.nl;
    double *wr = shm, *wi = wr + nz, *rr = wi + nz, *ri = rr + nz, *ar = ri + nz, *ai = ar + nz;
    double *ubr = ar, *ubi = ai, *vbr = ai + nz, *vbi = vbr + nz; // ub reuses the dphidz columns
    double *tmp = vbi + nz;              // 64
    double *fch = tmp + 64;              // kCH * (kv+1)
    int *ich = (int *)(fch + kCH * (A.kv + 1));
    const int m = blockIdx.x;
    const int lane = threadIdx.x;
    const int ii = m % A.nxhp, j = m / A.nxhp;
    const long o = 2L * m;
    const bool zero = (m == 0);
    const double xw = A.xw[ii], yw = A.yw[j];
    const int sP = A.modesys[4 * m], sUV = A.modesys[4 * m + 1];
    const int sW = A.modesys[4 * m + 2], sT = A.modesys[4 * m + 3];
    for (int k = lane; k < nz; k += 64) {
        wr[k] = A.w[(long)k * A.nxppy + o];
        wi[k] = A.w[(long)k * A.nxppy + o + 1];
    }
    __syncthreads();
    // ---- pressure (calcuvwt.f:87-222) ----
    if (!zero) {
        const double eps = A.sysinfo[4 * sP];
        for (int k = lane; k < nz; k += 64) {
            int ks = k / nl, k1 = k - ks * nl;
            double fac = 2. / A.zh[ks];
            double sr = 0., si = 0.;
            for (int t = 0; t < nl; t++) {
                double dd = A.D[k1 * nl + t];
                sr = sr + dd * wr[ks * nl + t];
                si = si + dd * wi[ks * nl + t];
            }
            double dfr = sr * fac, dfi = si * fac;
            long n = (long)k * A.nxppy + o;
            double dr = (0. + (-xw * A.u[n + 1])) + (-yw * A.v[n + 1]);
            double di = (0. + xw * A.u[n]) + yw * A.v[n];
            dr = dr + dfr;
            di = di + dfi;
            rr[k] = -(dr / A.dt) * eps; // penalty.f:191-194
            ri[k] = -(di / A.dt) * eps;
        }
        __syncthreads();
        if (lane == 0) { // Neumann boundary data (penalty.f:583-598)
            double t1 = A.sysinfo[4 * sP + 1], t2 = A.sysinfo[4 * sP + 2];
            long nb = A.nxppy;
            double bcrtop = A.wnl[nb + o] + A.wviscrot[nb + o];
            double bcrbot = A.wnl[o] + A.wviscrot[o];
            double bcitop = A.wnl[nb + o + 1] + A.wviscrot[nb + o + 1];
            double bcibot = A.wnl[o + 1] + A.wviscrot[o + 1];
            rr[0] = rr[0] - t1 * bcrbot * eps;
            ri[0] = ri[0] - t1 * bcibot * eps;
            rr[nz - 1] = rr[nz - 1] + t2 * bcrtop * eps;
            ri[nz - 1] = ri[nz - 1] + t2 * bcitop * eps;
        }
        __syncthreads();
        double *bp[2] = {rr, ri};
        band_solve<2>(A, sP, bp, fch, ich);
        if (A.cpen) filter_cols(A, bp, 2, tmp);
        for (int k = lane; k < nz; k += 64) { // dphidz (calcuvwt.f:853)
            int ks = k / nl, k1 = k - ks * nl;
            double fac = 2. / A.zh[ks];
            double sr = 0., si = 0.;
            for (int t = 0; t < nl; t++) {
                double dd = A.D[k1 * nl + t];
                sr = sr + dd * rr[ks * nl + t];
                si = si + dd * ri[ks * nl + t];
            }
            ar[k] = sr * fac;
            ai[k] = si * fac;
        }
    } else {
        for (int k = lane; k < nz; k += 64) {
            rr[k] = 0.; ri[k] = 0.; ar[k] = 0.; ai[k] = 0.;
        }
    }
    __syncthreads();
    // ---- projection (update, calcuvwt.f:826-901) ----
    {
        const double cx = A.dt * xw, cy = A.dt * yw;
        const int k0 = A.noslip ? 1 : 0;
        for (int k = lane; k < nz; k += 64) {
            long n = (long)k * A.nxppy + o;
            if (k < k0) {
                ubr[k] = 0.; ubi[k] = 0.; vbr[k] = 0.; vbi[k] = 0.; wr[k] = 0.; wi[k] = 0.;
                continue;
            }
            double tr = rr[k], ti = ri[k];
            double dzr = ar[k], dzi = ai[k];
            ubr[k] = A.u[n] - (0. * tr - cx * ti);
            ubi[k] = A.u[n + 1] - (0. * ti + cx * tr);
            vbr[k] = A.v[n] - (0. * tr - cy * ti);
            vbi[k] = A.v[n + 1] - (0. * ti + cy * tr);
            if (zero) {
                wr[k] = 0.; wi[k] = 0.;
            } else {
                wr[k] = wr[k] - A.dt * dzr;
                wi[k] = wi[k] - A.dt * dzi;
            }
        }
    }
    __syncthreads();
    const double xnudt = A.xnu * A.dt;
    // ---- w (calcuvwt.f:496-585) ----
    if (zero) {
        for (int k = lane; k < nz; k += 64) {
            A.w[(long)k * A.nxppy + o] = 0.;
            A.w[(long)k * A.nxppy + o + 1] = 0.;
        }
    } else {
        const double eps = A.sysinfo[4 * sW];
        for (int k = lane; k < nz; k += 64) {
            rr[k] = -(-wr[k] / xnudt) * eps;
            ri[k] = -(-wi[k] / xnudt) * eps;
        }
        __syncthreads();
        double *bp[2] = {rr, ri};
        band_solve<2>(A, sW, bp, fch, ich);
        if (A.cpen) filter_cols(A, bp, 2, tmp);
        for (int k = lane; k < nz; k += 64) {
            A.w[(long)k * A.nxppy + o] = rr[k];
            A.w[(long)k * A.nxppy + o + 1] = ri[k];
        }
    }
    // ---- u, v (calcuvwt.f:247-470): one factor, four right-hand sides ----
    {
        const double eps = A.sysinfo[4 * sUV];
        for (int k = lane; k < nz; k += 64) {
            ubr[k] = -(-ubr[k] / xnudt) * eps;
            ubi[k] = -(-ubi[k] / xnudt) * eps;
            vbr[k] = -(-vbr[k] / xnudt) * eps;
            vbi[k] = -(-vbi[k] / xnudt) * eps;
        }
        __syncthreads();
        double *bp[4] = {ubr, ubi, vbr, vbi};
        band_solve<4>(A, sUV, bp, fch, ich);
        if (A.cpen) filter_cols(A, bp, 4, tmp);
        for (int k = lane; k < nz; k += 64) {
            double a = ubr[k], b = ubi[k];
            if (zero) { // calcuvwt.f:343-361
                a = A.damp[k] * a;
                b = A.damp[k] * b;
            }
            long n = (long)k * A.nxppy + o;
            A.u[n] = a;
            A.u[n + 1] = b;
            A.v[n] = vbr[k];
            A.v[n + 1] = vbi[k];
        }
    }
    __syncthreads();
    // ---- density (calcuvwt.f:609-741) ----
    {
        const double eps = A.sysinfo[4 * sT];
        const double dk = A.dt * A.xkappa;
        for (int k = lane; k < nz; k += 64) {
            long n = (long)k * A.nxppy + o;
            rr[k] = -(-A.t[n] / dk) * eps;
            ri[k] = -(-A.t[n + 1] / dk) * eps;
        }
        __syncthreads();
        double *bp[2] = {rr, ri};
        band_solve<2>(A, sT, bp, fch, ich);
        if (A.cpen) filter_cols(A, bp, 2, tmp);
        for (int k = lane; k < nz; k += 64) {
            long n = (long)k * A.nxppy + o;
            A.t[n] = rr[k];
            A.t[n + 1] = ri[k];
        }
    }
}

void launch_calc_mode(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt,
                      const double *wnl, const double *wviscrot, double dt, int cpen, hipStream_t st)
{
    ModeArgs A;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3];
    A.wnl = wnl; A.wviscrot = wviscrot;
    A.AB = S.AB; A.ipiv = S.ipiv; A.sysinfo = S.sysinfo; A.modesys = S.modesys;
    A.D = g.D; A.AL = g.AL; A.wg = g.wg; A.sig = g.sig; A.zh = g.zh; A.xw = g.xw; A.yw = g.yw;
    A.damp = g.damp;
    A.dt = dt; A.xnu = P.xnu; A.xkappa = P.xkappa;
    A.nxppy = g.nxppy;
    A.nz = g.nz; A.nl = g.nzloc; A.ns = g.nsubd; A.nxhp = g.nxhp;
    A.ldab = S.ldab; A.kl = S.kl; A.kv = S.kv;
    A.cpen = cpen;
    A.noslip = P.bc_uv_bottom == 1;
    int nmode = g.nxhp * g.ny;
    size_t sh = sizeof(double) * (8 * (size_t)g.nz + 64 + (size_t)kCH * (S.kv + 1)) + sizeof(int) * kCH;
    hipLaunchKernelGGL(k_calc_mode, dim3(nmode), dim3(64), sh, st, A);
}

} // namespace iw
