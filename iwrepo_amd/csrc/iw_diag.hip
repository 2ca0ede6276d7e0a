// Diagnostics of jmsegret/IWrepo mode1_bbl on the device.
//
// postp (postp.f:9-76): the x-means of u and of the SGS tensors at every
// (y,z) line (umeancalc, postp.f:129-210), d umean/dz (dpdzyzplane,
// postp.f:984-1024), d umean/dy (dudymeancalc, postp.f:237-296: FFT991 of
// each z level, times i m beta, back), and the TKE budget sums of
// energycalc_postp (postp.f:553-694) over the points k >= 2, i <= nx of the
// perturbation fields and their gradients.  The budget is summed per block
// and the block partials in a fixed order, so a run is reproducible.
//
// speclegencalc (filtering.f:1110-1221): at the row j = ny/2 the Legendre
// coefficients of every z column of every subdomain, their magnitudes
// averaged over x.
//
// Kinetic and available potential energy (CalcE/integration.f:14-226, the
// offline tool, in 3-D): per z level the horizontal mean of rho0/2 q^2 and
// g^2 rho'^2/(2 rho0 N^2) times the span, then the GLL quadrature of
// integrcol (derv.f:479-512) per subdomain.
#include "iw_internal.h"

namespace iw {

namespace {

constexpr int kPpThreads = 256;
constexpr int kPpSums = 12; // tketot u1 u2 u3 prody prodz bflux tempvar epsilon epssgs chi chisgs

// sum of NV values over the block; the result is valid in thread 0
template <int NV>
__device__ void block_sum(double (&v)[NV], double *lds /* NV * 16 */)
{
    for (int q = 0; q < NV; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0)
        for (int q = 0; q < NV; q++) lds[q * 16 + wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; q++) {
            double t = 0.;
            for (int w = 0; w < nw; w++) t += lds[q * 16 + w];
            v[q] = t;
        }
}

struct RowMeanArgs {
    double *u;            // physical u: the mean is taken out of it in place
    const double *sgs[6]; // t11 t12 t13 t22 t23 t33 (all null: none)
    double *um;           // ny * nz, row = j + ny k
    double *tm;           // 6 * ny * nz
    int nx, ny, nz, nxpp;
    long nxppy;
};

// one block per (j, k) line; k = 0 keeps mean 0 (postp.f:166-188 start at
// k = 2)
__global__ __launch_bounds__(kPpThreads) void k_pp_rowmean(RowMeanArgs A)
{
    __shared__ double lds[7 * 16];
    __shared__ double mean;
    const int row = blockIdx.x, j = row % A.ny, k = row / A.ny;
    const long rows = (long)A.ny * A.nz;
    if (k == 0) {
        if (threadIdx.x == 0) {
            A.um[row] = 0.;
            for (int q = 0; q < 6; q++) A.tm[q * rows + row] = 0.;
        }
        return;
    }
    const long o = (long)k * A.nxppy + (long)j * A.nxpp;
    const bool sg = A.sgs[0] != nullptr;
    double v[7] = {0., 0., 0., 0., 0., 0., 0.};
    for (int i = threadIdx.x; i < A.nx; i += blockDim.x) {
        v[0] += A.u[o + i];
        if (sg)
            for (int q = 0; q < 6; q++) v[1 + q] += A.sgs[q][o + i];
    }
    block_sum<7>(v, lds);
    if (threadIdx.x == 0) {
        mean = v[0] / A.nx;
        A.um[row] = mean;
        for (int q = 0; q < 6; q++) A.tm[q * rows + row] = v[1 + q] / A.nx;
    }
    __syncthreads();
    const double m = mean;
    for (int i = threadIdx.x; i < A.nx; i += blockDim.x) A.u[o + i] -= m; // perturbcalc, postp.f:333-345
}

// d umean/dz per subdomain (dpdzyzplane with dpdzcol, derv.f:279-325)
__global__ void k_pp_dumdz(const double *__restrict__ um, double *dz, const double *__restrict__ D,
                           const double *__restrict__ zh, int ny, int nz, int nl)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ny * nz) return;
    const int j = id % ny, k = id / ny, ks = k / nl, k1 = k - ks * nl;
    double sum = 0.0;
    for (int k2 = 0; k2 < nl; k2++) sum = sum + D[k1 * nl + k2] * um[(ks * nl + k2) * ny + j];
    dz[id] = sum * (2. / zh[ks]);
}

// d umean/dy (postp.f:237-296): one block per z level; FFT991's forward
// coefficients c_m = (1/ny) sum_q a_q e^{-2 pi i m q / ny}, times i m beta,
// synthesised back for 0 < m < ny/2 (the m = 0 term is zero and the
// Nyquist coefficient's real part, -beta ny/2 Im c, vanishes)
__global__ __launch_bounds__(kPpThreads) void k_pp_dumdy(const double *__restrict__ um, double *dy, int ny,
                                                         double beta, double pi2)
{
    extern __shared__ double sm[]; // a[ny], cr[ny/2], ci[ny/2]
    const int k = blockIdx.x, nh = ny / 2;
    double *a = sm, *cr = sm + ny, *ci = sm + ny + nh;
    for (int q = threadIdx.x; q < ny; q += blockDim.x) a[q] = um[(long)k * ny + q];
    __syncthreads();
    for (int m = 1 + threadIdx.x; m < nh; m += blockDim.x) {
        double sr = 0., si = 0.;
        for (int q = 0; q < ny; q++) {
            double sn, cs;
            sincos(pi2 * (double)((long)m * q % ny) / ny, &sn, &cs);
            sr = sr + a[q] * cs;
            si = si - a[q] * sn;
        }
        const double kb = m * beta;
        cr[m] = -kb * (si / ny);
        ci[m] = kb * (sr / ny);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        double sum = 0.;
        for (int m = 1; m < nh; m++) {
            double sn, cs;
            sincos(pi2 * (double)((long)m * j % ny) / ny, &sn, &cs);
            sum = sum + 2. * (cr[m] * cs - ci[m] * sn);
        }
        dy[(long)k * ny + j] = sum;
    }
}

struct SumArgs {
    const double *u, *v, *w, *t; // physical perturbation u and v, w, temp
    const double *g[12];         // dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz dtdx dtdy dtdz
    const double *sgs[9];        // t11 t12 t13 t22 t23 t33 th1 th2 th3 (null: zero)
    const double *um, *dumdy, *dumdz, *tm;
    double *part; // gridDim.x * kPpSums
    int nx, ny, nz, nxpp;
    long nxppy;
};

// the TKE budget terms (postp.f:553-694) at the points k >= 2, i <= nx
__global__ __launch_bounds__(kPpThreads) void k_pp_sums(SumArgs A)
{
    __shared__ double lds[kPpSums * 16];
    double v[kPpSums];
    for (int q = 0; q < kPpSums; q++) v[q] = 0.;
    const long plane = (long)A.ny * A.nx, npts = (long)(A.nz - 1) * plane;
    const long rows = (long)A.ny * A.nz;
    const bool sg = A.sgs[0] != nullptr;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npts; p += (long)gridDim.x * blockDim.x) {
        const int k = 1 + (int)(p / plane);
        const long r = p - (long)(k - 1) * plane;
        const int j = (int)(r / A.nx), i = (int)(r - (long)j * A.nx);
        const long n = (long)k * A.nxppy + (long)j * A.nxpp + i;
        const long row = (long)k * A.ny + j;
        const double uu = A.u[n], vv = A.v[n], ww = A.w[n], tt = A.t[n];
        double g[12];
        for (int q = 0; q < 12; q++) g[q] = A.g[q][n];
        v[0] += uu * uu + vv * vv + ww * ww;
        v[1] += uu * uu;
        v[2] += vv * vv;
        v[3] += ww * ww;
        v[4] -= A.dumdy[row] * (uu * vv);
        v[5] -= A.dumdz[row] * (uu * ww);
        v[6] += ww * tt;
        v[7] += tt * tt;
        v[8] += g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3] + g[4] * g[4] + g[5] * g[5] + g[6] * g[6]
                + g[7] * g[7] + g[8] * g[8];
        if (sg) {
            const double s11 = g[0], s22 = g[4], s33 = g[8];
            const double s12 = 0.5 * (g[1] + g[3]), s13 = 0.5 * (g[2] + g[6]), s23 = 0.5 * (g[5] + g[7]);
            double f[6];
            for (int q = 0; q < 6; q++) f[q] = A.sgs[q][n] - A.tm[q * rows + row];
            v[9] -= f[0] * s11 + f[3] * s22 + f[5] * s33 + 2. * f[1] * s12 + 2. * f[2] * s13 + 2. * f[4] * s23;
            v[11] -= A.sgs[6][n] * g[9] + A.sgs[7][n] * g[10] + A.sgs[8][n] * g[11];
        }
        v[10] += g[9] * g[9] + g[10] * g[10] + g[11] * g[11];
    }
    block_sum<kPpSums>(v, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < kPpSums; q++) A.part[(long)blockIdx.x * kPpSums + q] = v[q];
}

// block partials in order, tkemean from the means, the scalings of
// postp.f:698-710 (bflux * ra * pr with the Rayleigh number ra = 0 of
// /param/, which nothing sets); out: istep, t and the sixteen values of the
// energy.dat line (postp.f:53-58)
__global__ __launch_bounds__(kPpThreads) void k_pp_final(const double *__restrict__ part, int nblk,
                                                         const double *__restrict__ um, int ny, int nz, int nx,
                                                         double pr, double xkappa, double istep, double t,
                                                         double *out)
{
    __shared__ double lds[(kPpSums + 1) * 16];
    double v[kPpSums + 1];
    for (int q = 0; q <= kPpSums; q++) v[q] = 0.;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
        for (int q = 0; q < kPpSums; q++) v[q] += part[(long)b * kPpSums + q];
    for (int r = ny + threadIdx.x; r < ny * nz; r += blockDim.x) v[kPpSums] += um[r] * um[r];
    block_sum<kPpSums + 1>(v, lds);
    if (threadIdx.x != 0) return;
    const double ra = 0.;
    const double tketot = v[0] * 0.5, prody = v[4], prodz = v[5];
    const double bflux = v[6] * ra * pr, eps = v[8] * pr, epssgs = v[9];
    const double chi = v[10] * xkappa, chisgs = v[11];
    const double r[18] = {istep, t, tketot, v[1], v[2], v[3], prody, prodz, prody + prodz, bflux, eps, epssgs,
                          eps + epssgs, chi, chisgs, chisgs + chi, v[12] * 0.5 * (double)nx, v[7]};
    for (int q = 0; q < 18; q++) out[q] = r[q];
}

// speclegencalc: block (ks, kloc) averages |a_kloc| over i at j = jrow
__global__ __launch_bounds__(kPpThreads) void k_speclegen(const double *__restrict__ f,
                                                          const double *__restrict__ wg,
                                                          const double *__restrict__ AL, int nx, int nl,
                                                          int nsubd, int jrow, int nxpp, long nxppy,
                                                          double istep, double t, double *hdr, double *out)
{
    __shared__ double lds[16];
    const int ks = blockIdx.x % nsubd, kl = blockIdx.x / nsubd, nzm = nl - 1;
    const double c = (kl == nzm) ? 0.5 * nzm : (kl + 0.5);
    double v[1] = {0.};
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        const double *col = f + (long)ks * nl * nxppy + (long)jrow * nxpp + i;
        double sum = 0.;
        for (int k1 = 0; k1 < nl; k1++) sum = sum + wg[k1] * col[(long)k1 * nxppy] * AL[kl * nl + k1];
        v[0] += fabs(c * sum);
    }
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) {
        out[kl * nsubd + ks] = v[0] / nx;
        if (hdr && blockIdx.x == 0) {
            hdr[0] = istep;
            hdr[1] = t;
        }
    }
}

// per z level: span * mean over (j, i < nx) of c q^2, for u, v, w (c = cke)
// and rho' (c = cpe); lvl[q * nz + k]
__global__ __launch_bounds__(kPpThreads) void k_kepe_level(const double *__restrict__ u,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ w,
                                                           const double *__restrict__ t, int nx, int ny, int nz,
                                                           int nxpp, long nxppy, double cke, double cpe,
                                                           double span, double *lvl)
{
    __shared__ double lds[4 * 16];
    const int k = blockIdx.x;
    double s[4] = {0., 0., 0., 0.};
    const long n = (long)nx * ny;
    for (long p = threadIdx.x; p < n; p += blockDim.x) {
        const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
        const long o = (long)k * nxppy + (long)j * nxpp + i;
        const double a = u[o], b = v[o], c = w[o], d = t[o];
        s[0] += cke * (a * a);
        s[1] += cke * (b * b);
        s[2] += cke * (c * c);
        s[3] += cpe * (d * d);
    }
    block_sum<4>(s, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) lvl[(long)q * nz + k] = span * s[q] / ((double)nx * ny);
}

// integrcol per subdomain, summed: out[q] for the four fields
__global__ void k_kepe_final(const double *__restrict__ lvl, const double *__restrict__ wg,
                             const double *__restrict__ zh, int nl, int nsubd, int nz, double *out)
{
    const int q = threadIdx.x;
    if (q >= 4) return;
    double tot = 0.;
    for (int ks = 0; ks < nsubd; ks++) {
        double sum = 0.;
        for (int k = 0; k < nl; k++) sum = sum + wg[k] * lvl[(long)q * nz + ks * nl + k];
        tot = tot + sum * zh[ks] / 2.;
    }
    out[q] = tot;
}

} // namespace

void launch_kepe(const Geo &g, double *const *phys4, double cke, double cpe, double span, double *lvl, double *out,
                 hipStream_t st)
{
    hipLaunchKernelGGL(k_kepe_level, dim3((unsigned)g.nz), dim3(kPpThreads), 0, st, phys4[0], phys4[1], phys4[2],
                       phys4[3], g.nx, g.ny, g.nz, g.nxpp, g.nxppy, cke, cpe, span, lvl);
    hipLaunchKernelGGL(k_kepe_final, dim3(1), dim3(64), 0, st, lvl, g.wg, g.zh, g.nzloc, g.nsubd, g.nz, out);
}

int pp_blocks(const Geo &g)
{
    const long npts = (long)(g.nz - 1) * g.ny * g.nx;
    long b = (npts + kPpThreads - 1) / kPpThreads;
    if (b > 2048) b = 2048;
    return b < 1 ? 1 : (int)b;
}

void launch_pp_means(const Geo &g, double *u, double *const *sgs6, double *um, double *tm, double *dumdy,
                     double *dumdz, double beta, hipStream_t st)
{
    RowMeanArgs A;
    A.u = u;
    for (int q = 0; q < 6; q++) A.sgs[q] = sgs6 ? sgs6[q] : nullptr;
    A.um = um;
    A.tm = tm;
    A.nx = g.nx;
    A.ny = g.ny;
    A.nz = g.nz;
    A.nxpp = g.nxpp;
    A.nxppy = g.nxppy;
    hipLaunchKernelGGL(k_pp_rowmean, dim3((unsigned)(g.ny * g.nz)), dim3(kPpThreads), 0, st, A);
    const int n = g.ny * g.nz;
    hipLaunchKernelGGL(k_pp_dumdz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, um, dumdz, g.D, g.zh,
                       g.ny, g.nz, g.nzloc);
    const size_t lds = sizeof(double) * (size_t)(g.ny + 2 * (g.ny / 2));
    hipLaunchKernelGGL(k_pp_dumdy, dim3((unsigned)g.nz), dim3(kPpThreads), lds, st, um, dumdy, g.ny, beta,
                       8. * atan(1.0));
}

void launch_pp_sums(const Geo &g, double *const *uvwt, double *const *grad12, double *const *sgs9,
                    const double *um, const double *dumdy, const double *dumdz, const double *tm, double *part,
                    double pr, double xkappa, double istep, double t, double *out, hipStream_t st)
{
    SumArgs A;
    A.u = uvwt[0];
    A.v = uvwt[1];
    A.w = uvwt[2];
    A.t = uvwt[3];
    for (int q = 0; q < 12; q++) A.g[q] = grad12[q];
    for (int q = 0; q < 9; q++) A.sgs[q] = sgs9 ? sgs9[q] : nullptr;
    A.um = um;
    A.dumdy = dumdy;
    A.dumdz = dumdz;
    A.tm = tm;
    A.part = part;
    A.nx = g.nx;
    A.ny = g.ny;
    A.nz = g.nz;
    A.nxpp = g.nxpp;
    A.nxppy = g.nxppy;
    const int nb = pp_blocks(g);
    hipLaunchKernelGGL(k_pp_sums, dim3((unsigned)nb), dim3(kPpThreads), 0, st, A);
    hipLaunchKernelGGL(k_pp_final, dim3(1), dim3(kPpThreads), 0, st, part, nb, um, g.ny, g.nz, g.nx, pr, xkappa,
                       istep, t, out);
}

void launch_speclegen(const Geo &g, const double *phys, double istep, double t, double *hdr, double *out,
                      hipStream_t st)
{
    const int jrow = g.ny >= 2 ? g.ny / 2 - 1 : 0; // j = ny/2 (1-based)
    hipLaunchKernelGGL(k_speclegen, dim3((unsigned)(g.nsubd * g.nzloc)), dim3(kPpThreads), 0, st, phys, g.wg,
                       g.AL, g.nx, g.nzloc, g.nsubd, jrow, g.nxpp, g.nxppy, istep, t, hdr, out);
}

// spectral i kx f and i ky f of one field (energycalc_postp, postp.f:476-489)
__global__ void k_grad_xy(const double *__restrict__ a, double *dx, double *dy, long ncplx, int nxhp, int ny,
                          const double *__restrict__ xw, const double *__restrict__ yw)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncplx) return;
    const double kx = xw[(int)(c % nxhp)], ky = yw[(int)((c / nxhp) % ny)];
    const double re = a[2 * c], im = a[2 * c + 1];
    dx[2 * c] = -kx * im;
    dx[2 * c + 1] = kx * re;
    dy[2 * c] = -ky * im;
    dy[2 * c + 1] = ky * re;
}

void launch_grad_xy(const Geo &g, const double *a, double *dx, double *dy, hipStream_t st)
{
    const long ncplx = g.ntot / 2;
    hipLaunchKernelGGL(k_grad_xy, dim3((unsigned)((ncplx + 255) / 256)), dim3(256), 0, st, a, dx, dy, ncplx,
                       g.nxhp, g.ny, g.xw, g.yw);
}

} // namespace iw
