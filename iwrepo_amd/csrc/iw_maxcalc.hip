// maxcalc (jmsegret/IWrepo mode1_bbl, various.f:665-987) on the device: the
// running extrema and CFL maxima of the physical u, v, w, temp.  The
// per-block partial results come from the epilogue of the inverse x
// transform (k_fft_x); the kernels here combine them into the record and
// follow the umaxt chain.
//
// The reference scans i (outer), j, then z (inner) and keeps a later point
// on a tie (.ge./.le. comparisons, various.f:755-845).  With L = (i ny + j)
// nz + k that is: the extreme value, at the largest L among the points that
// reach it, provided it passes the start value (0, or 1 for tmin).  That
// makes every quantity but one an ordinary (value, L) reduction.  The
// exception is umaxt (various.f:755-758): a new maximum stores u + umax, so
// later points are compared against the last record plus umax.  With umax = 0
// this is again a plain maximum; otherwise the chain of records is followed
// exactly by one work-group walking the columns (i,j) in L order, skipping
// columns and groups of 256 columns whose maximum cannot start a record.
#include "iw_internal.h"

#include <cfloat>

namespace iw {

namespace {

constexpr int kMcThreads = 256;

// quantities: 0 umin 1 umax 2 vmin 3 vmax 4 wmin 5 wmax 6 tmin 7 tmax
// 8 cflx 9 cfly 10 cflz; even indices below 8 are minima
__device__ __forceinline__ bool mc_is_min(int q) { return q < 8 && (q & 1) == 0; }

// does (x, lx) replace (b, lb)?  NaN never replaces (Fortran comparisons
// with NaN are false)
__device__ __forceinline__ bool mc_better(bool mn, double x, int lx, double b, int lb)
{
    return mn ? (x < b || (x == b && lx > lb)) : (x > b || (x == b && lx > lb));
}

struct McAcc {
    double v[kMcQ];
    int l[kMcQ];
    __device__ void init()
    {
#pragma unroll
        for (int q = 0; q < kMcQ; q++) {
            v[q] = mc_is_min(q) ? DBL_MAX : -DBL_MAX;
            l[q] = -1;
        }
    }
    __device__ __forceinline__ void add(int q, double x, int lx)
    {
        if (mc_better(mc_is_min(q), x, lx, v[q], l[q])) {
            v[q] = x;
            l[q] = lx;
        }
    }
};

// combine the accumulators of a work-group of kMcThreads threads; the result
// is valid in thread 0
__device__ void mc_block_reduce(McAcc &a)
{
    __shared__ double sv[kMcThreads / 64][kMcQ];
    __shared__ int sl[kMcThreads / 64][kMcQ];
#pragma unroll
    for (int q = 0; q < kMcQ; q++) {
        const bool mn = mc_is_min(q);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            double x = __shfl_down(a.v[q], off, 64);
            int lx = __shfl_down(a.l[q], off, 64);
            if (mc_better(mn, x, lx, a.v[q], a.l[q])) {
                a.v[q] = x;
                a.l[q] = lx;
            }
        }
    }
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < kMcQ; q++) {
            sv[wv][q] = a.v[q];
            sl[wv][q] = a.l[q];
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kMcThreads / 64; w++)
#pragma unroll
            for (int q = 0; q < kMcQ; q++)
                if (mc_better(mc_is_min(q), sv[w][q], sl[w][q], a.v[q], a.l[q])) {
                    a.v[q] = sv[w][q];
                    a.l[q] = sl[w][q];
                }
}

// physical u at (i,j,k) as maxcalc sees it: the transformed value plus,
// after convec's transform, the kx-Nyquist term that the round trip scaled
// by 1/ny (the epilogue of k_fft_x forms the same sum)
__device__ __forceinline__ double mc_u(const McArgs &A, int i, int j, int k)
{
    const long o = (long)k * A.nxppy + (long)j * A.nxpp;
    double x = A.u32 ? (double)A.u32[o + i] : A.u[o + i];
    if (A.cny != 0.0) {
        const double c = A.cny * A.spec[o + A.nx];
        x = (i & 1) ? x - c : x + c;
    }
    return x;
}

// u + umwv as the umaxt chain compares it (various.f:757-758)
__device__ __forceinline__ double mc_uu(const McArgs &A, int i, int j, int k)
{
    const double x = mc_u(A, i, j, k);
    return A.umwv ? x + A.umwv[(long)k * A.nxpp + i] : x + 0.0;
}

// partials -> the record (layout of iw_maxcalc in iwgpu.h); entries the
// reference keeps from the previous call (itmin, jtmin, dzmax) stay when
// there is no new record
__global__ __launch_bounds__(kMcThreads) void k_mc_final(McArgs A, int istep, double t, double umax, double *rec)
{
    McAcc acc;
    acc.init();
    for (int b = threadIdx.x; b < A.nblk; b += kMcThreads) {
        const double *p = A.part + (size_t)b * (2 * kMcQ);
        for (int q = 0; q < kMcQ; q++) acc.add(q, p[2 * q], (int)p[2 * q + 1]);
    }
    mc_block_reduce(acc);
    if (threadIdx.x != 0) return;
    const int nz = A.nz, ny = A.ny;
    auto kk = [&](int L) { return L % nz; };
    auto jj = [&](int L) { return (L / nz) % ny; };
    auto ii = [&](int L) { return L / (nz * ny); };
    // did the running extremum with start value s move at all?
    auto moved = [&](int q, double s) {
        return acc.l[q] >= 0 && (mc_is_min(q) ? acc.v[q] <= s : acc.v[q] >= s);
    };
    rec[0] = istep;
    rec[1] = t;
    const int pairs[6][3] = {{0, 2, 4}, {1, 3, 5}, {2, 6, 8}, {3, 7, 9}, {4, 10, 12}, {5, 11, 13}};
    for (int m = 0; m < 6; m++) { // u, v, w extrema with their k
        const int q = pairs[m][0];
        const bool r = moved(q, 0.0);
        rec[pairs[m][1]] = r ? acc.v[q] : 0.0;
        rec[pairs[m][2]] = r ? kk(acc.l[q]) + 1 : 0;
    }
    if (umax == 0.0 && moved(1, 0.0)) rec[3] = acc.v[1] + umax; // umaxt; the walk sets it otherwise
    if (moved(6, 1.0)) {
        rec[14] = acc.v[6];
        rec[16] = kk(acc.l[6]) + 1;
        rec[18] = ii(acc.l[6]) + 1;
        rec[19] = jj(acc.l[6]) + 1;
    } else {
        rec[14] = 1.0;
        rec[16] = 0;
    }
    if (moved(7, 0.0)) {
        rec[15] = acc.v[7];
        rec[17] = kk(acc.l[7]) + 1;
        rec[20] = ii(acc.l[7]) + 1;
        rec[21] = jj(acc.l[7]) + 1;
    } else {
        rec[15] = 0.0;
        rec[17] = rec[20] = rec[21] = 0;
    }
    for (int m = 0; m < 3; m++) {
        const int q = 8 + m, o = 22 + 4 * m;
        if (moved(q, 0.0)) {
            rec[o] = acc.v[q];
            rec[o + 1] = ii(acc.l[q]) + 1;
            rec[o + 2] = jj(acc.l[q]) + 1;
            rec[o + 3] = kk(acc.l[q]) + 1;
        } else {
            rec[o] = 0.0;
            rec[o + 1] = rec[o + 2] = rec[o + 3] = 0;
        }
    }
    if (moved(10, 0.0)) rec[34] = A.dz[kk(acc.l[10])]; // dz at the CFLZ maximum
    // uveltest.dat probe u(32,1,10), u(32,1,50) by linear position
    const long ntot = (long)A.nxppy * nz;
    const long p1 = 31 + (long)A.nxpp * ny * 9, p2 = 31 + (long)A.nxpp * ny * 49;
    auto probe = [&](long p) {
        if (p >= ntot) return __builtin_nan("");
        const int i = (int)(p % A.nxpp), j = (int)((p / A.nxpp) % ny), k = (int)(p / A.nxppy);
        return i < A.nx ? mc_u(A, i, j, k) : (A.u32 ? (double)A.u32[p] : A.u[p]); // the pads hold zeros
    };
    rec[40] = probe(p1);
    rec[41] = probe(p2);
}

// mc_uu with its three cases fixed at compile time (FP32 u, the Nyquist
// term, the mean-wave term): the same operations in the same order
template <bool F32, bool CNY, bool MW>
__device__ __forceinline__ double mc_uu_t(const McArgs &A, int i, int j, int k)
{
    const long o = (long)k * A.nxppy + (long)j * A.nxpp;
    double x = F32 ? (double)A.u32[o + i] : A.u[o + i];
    if (CNY) {
        const double c = A.cny * A.spec[o + A.nx];
        x = (i & 1) ? x - c : x + c;
    }
    return MW ? x + A.umwv[(long)k * A.nxpp + i] : x + 0.0;
}

// partial column maxima of u over kMcSlices slices of z:
// cp[kc][i ny + j] = max over the slice of u(i,j,k).  Four loads in flight
// per step; the maximum does not depend on the order (fmax skips NaN, and
// -0/+0 compare equal wherever the maxima are used)
template <bool F32, bool CNY, bool MW>
__global__ __launch_bounds__(256) void k_mc_colmax(McArgs A, double *cp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxy = A.nx * A.ny;
    if (e >= nxy) return;
    const int j = e / A.nx, i = e - j * A.nx;
    const int kc = blockIdx.y, k0 = kc * A.nz / kMcSlices, k1 = (kc + 1) * A.nz / kMcSlices;
    double m = -DBL_MAX;
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
        const double a = mc_uu_t<F32, CNY, MW>(A, i, j, k), b = mc_uu_t<F32, CNY, MW>(A, i, j, k + 1);
        const double c = mc_uu_t<F32, CNY, MW>(A, i, j, k + 2), d = mc_uu_t<F32, CNY, MW>(A, i, j, k + 3);
        m = fmax(m, fmax(fmax(a, b), fmax(c, d)));
    }
    for (; k < k1; k++) m = fmax(m, mc_uu_t<F32, CNY, MW>(A, i, j, k));
    cp[(size_t)kc * nxy + i * A.ny + j] = m;
}

// column maxima cm (from the slices) and the maxima of groups of 256
// consecutive columns
__global__ __launch_bounds__(256) void k_mc_groupmax(const double *__restrict__ cp, int ncol, double *cm,
                                                     double *gm)
{
    __shared__ double s[4];
    const int c = blockIdx.x * 256 + threadIdx.x;
    double m = -DBL_MAX;
    if (c < ncol) {
        for (int kc = 0; kc < kMcSlices; kc++) m = fmax(m, cp[(size_t)kc * ncol + c]);
        cm[c] = m;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) gm[blockIdx.x] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
}

// smallest index in [0, 256) whose flag is set, or 256; all threads get it
__device__ int mc_block_first(bool flag)
{
    __shared__ int sw[4];
    const unsigned long long b = __ballot(flag);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = b ? (threadIdx.x & ~63) + __ffsll((long long)b) - 1 : 256;
    __syncthreads();
    const int r = min(min(sw[0], sw[1]), min(sw[2], sw[3]));
    __syncthreads();
    return r;
}

// the umaxt records (various.f:755-758): start value 0, a point with
// u >= umaxt becomes the record and umaxt = u + umax.  One work-group.
// Every loop advances (c, k0) or ends, so the walk finishes after at most
// ncol column steps plus one step per record.
__global__ __launch_bounds__(256) void k_mc_walk(McArgs A, const double *__restrict__ cm,
                                                 const double *__restrict__ gm, double umax, double *rec)
{
    const int ny = A.ny, nz = A.nz;
    const int ncol = A.nx * ny, ngrp = (ncol + 255) / 256;
    double thr = 0.0;
    int kmax = 0;
    int c = 0, k0 = 0;
    // first k >= k0 in column c with u >= thr, or nz
    auto scan_column = [&](int cc, int kb) -> int {
        const int i = cc / ny, j = cc - i * ny;
        for (int k = kb; k < nz; k += 256) {
            const int kt = k + (int)threadIdx.x;
            const bool f = kt < nz && mc_uu(A, i, j, kt) >= thr;
            const int r = mc_block_first(f);
            if (r < 256) return k + r;
        }
        return nz;
    };
    while (c < ncol) {
        if (k0 == 0) { // next column that can hold a record: groups first
            int g = c >> 8;
            int found = -1;
            // rest of the current group
            {
                const int cc = c + (int)threadIdx.x;
                const bool f = cc < min(ncol, (g + 1) * 256) && cm[cc] >= thr;
                const int r = mc_block_first(f);
                if (r < 256) found = c + r;
            }
            for (g = g + 1; found < 0 && g < ngrp; g += 256) {
                const int gg = g + (int)threadIdx.x;
                const int r = mc_block_first(gg < ngrp && gm[gg] >= thr);
                if (r == 256) continue;
                const int g0 = g + r; // first group with a candidate column
                const int cc = g0 * 256 + (int)threadIdx.x;
                const int r2 = mc_block_first(cc < ncol && cm[cc] >= thr);
                if (r2 < 256) {
                    found = g0 * 256 + r2;
                } else { // only when a group maximum is NaN-free but no column passes
                    g = g0 + 1 - 256;
                }
            }
            if (found < 0) break;
            c = found;
        }
        const int k = scan_column(c, k0);
        if (k < nz) {
            const int i = c / ny, j = c - i * ny;
            thr = mc_uu(A, i, j, k) + umax;
            kmax = k + 1;
            k0 = k + 1;
            if (k0 == nz) {
                c++;
                k0 = 0;
            }
        } else {
            c++;
            k0 = 0;
        }
    }
    if (threadIdx.x == 0) {
        rec[3] = thr;
        rec[5] = kmax;
    }
}

} // namespace

// the partials of kMcPre groups of blocks, each reduced by one work-group
// (the reduction is order-free: extreme value, largest L on a tie)
__global__ __launch_bounds__(kMcThreads) void k_mc_pre(const double *__restrict__ part, int nblk, double *part2)
{
    McAcc acc;
    acc.init();
    const int per = (nblk + (int)gridDim.x - 1) / (int)gridDim.x;
    const int b0 = (int)blockIdx.x * per, b1 = b0 + per < nblk ? b0 + per : nblk;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += kMcThreads) {
        const double *p = part + (size_t)b * (2 * kMcQ);
        for (int q = 0; q < kMcQ; q++) acc.add(q, p[2 * q], (int)p[2 * q + 1]);
    }
    mc_block_reduce(acc);
    if (threadIdx.x != 0) return;
    double *o = part2 + (size_t)blockIdx.x * (2 * kMcQ);
    for (int q = 0; q < kMcQ; q++) {
        o[2 * q] = acc.v[q];
        o[2 * q + 1] = (double)acc.l[q];
    }
}

void launch_maxcalc_final(const Geo &g, const McArgs &A, int istep, double t, double umax, double *rec,
                          hipStream_t st)
{
    if (A.part2 && A.nblk > kMcPre) {
        hipLaunchKernelGGL(k_mc_pre, dim3(kMcPre), dim3(kMcThreads), 0, st, A.part, A.nblk, A.part2);
        McArgs B = A;
        B.part = A.part2;
        B.nblk = kMcPre;
        hipLaunchKernelGGL(k_mc_final, dim3(1), dim3(kMcThreads), 0, st, B, istep, t, umax, rec);
        return;
    }
    hipLaunchKernelGGL(k_mc_final, dim3(1), dim3(kMcThreads), 0, st, A, istep, t, umax, rec);
}

// umaxt chain (umax != 0); colmax holds (kMcSlices + 1) nx ny doubles
void launch_maxcalc_umaxt(const Geo &g, const McArgs &A, double umax, double *colmax, double *groupmax,
                          double *rec, hipStream_t st)
{
    const int nxy = g.nx * g.ny;
    double *cm = colmax + (size_t)kMcSlices * nxy;
    const dim3 grid((nxy + 255) / 256, kMcSlices);
#define CM(F, C, M) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mc_colmax<F, C, M>), grid, dim3(256), 0, st, A, colmax)
    const bool c = A.cny != 0.0, w = A.umwv != nullptr;
    if (A.u32) {
        if (c) { if (w) CM(true, true, true); else CM(true, true, false); }
        else { if (w) CM(true, false, true); else CM(true, false, false); }
    } else {
        if (c) { if (w) CM(false, true, true); else CM(false, true, false); }
        else { if (w) CM(false, false, true); else CM(false, false, false); }
    }
#undef CM
    hipLaunchKernelGGL(k_mc_groupmax, dim3((nxy + 255) / 256), dim3(256), 0, st, colmax, nxy, cm, groupmax);
    hipLaunchKernelGGL(k_mc_walk, dim3(1), dim3(256), 0, st, A, cm, groupmax, umax, rec);
}

} // namespace iw
