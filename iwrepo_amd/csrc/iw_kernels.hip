// Device kernels of the stratified-Boussinesq step for MI355X (gfx950).
//
// Reference: jmsegret/IWrepo, mode1_bbl (Fortran, FP64).  Each kernel cites the
// reference loop it replaces.  All arithmetic is FP64.  Layout of every field
// is the reference's a(nxpp,ny,nz): a[i + nxpp*(j + ny*k)], x fastest; the
// spectral plane is nxhp*ny contiguous complex values.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "iw_internal.h"

namespace iw {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// the LDS tile of the transforms holds FP64 values, or FP32 values on the
// mixed path (C = float2); the pre- and post-processing stays FP64
__device__ __forceinline__ double2 tod(double2 v) { return v; }
__device__ __forceinline__ double2 tod(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename C>
__device__ __forceinline__ C toc(double2 v);
template <>
__device__ __forceinline__ double2 toc<double2>(double2 v)
{
    return v;
}
template <>
__device__ __forceinline__ float2 toc<float2>(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}
// tile type of a transform with storage types TI (in) and TO (out)
template <typename TI, typename TO>
using tile_t = typename std::conditional<std::is_same<TI, float>::value || std::is_same<TO, float>::value, float2,
                                         double2>::type;

// ---------------------------------------------------------------------------
// In-place multi-radix Stockham FFT on LDS.  A block of T threads holds
// 8 T complex values (nrows sequences of length N, N a power of two),
// so every pass of radix R has exactly 8/R butterflies per thread.  (Other
// lengths N = 2^a 3^b 5^c: fft_lds_g below, radix 2, 3, 4, 5 and 8.)
// Passes are radix 8 while at least three factors of two remain, then one
// radix-4 or radix-2 pass.  Each pass reads its butterflies into registers,
// synchronises, and writes the outputs to their autosort positions.  Row
// layouts store value i at slot lp(i) = i ^ ((i >> 3) & 15), a permutation
// within each 16-value chunk under which the radix-8 stride-8 writes and
// aligned runs of 8 or 16 values hit distinct 16-byte bank groups.
// sign -1: sum g_j e^{-2 pi i jk/N}; +1: e^{+...}.  tw[t] = e^{-2 pi i t/N}.
// ---------------------------------------------------------------------------
// The block size is a run-time choice (Geo::ft: 64, 128 or 256 threads); a
// block always holds 8 values per thread.
constexpr int kFT = 256; // largest block

__host__ __device__ __forceinline__ int lp(int i) { return i ^ ((i >> 3) & 15); }

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by sign*i
__device__ __forceinline__ double2 cmuli(double2 a, int sign)
{
    return sign < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}
__device__ __forceinline__ float2 cmuli(float2 a, int sign)
{
    return sign < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
template <typename C>
__device__ __forceinline__ C cmake(double x, double y);
template <>
__device__ __forceinline__ double2 cmake<double2>(double x, double y)
{
    return make_double2(x, y);
}
template <>
__device__ __forceinline__ float2 cmake<float2>(double x, double y)
{
    return make_float2((float)x, (float)y);
}

template <int R, typename C>
__device__ __forceinline__ void dft(C *a, int sign)
{
    using S = decltype(a[0].x);
    if (R == 2) {
        const C t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if (R == 4) {
        const C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        const C t2 = cadd(a[1], a[3]), t3 = cmuli(csub(a[1], a[3]), sign);
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    } else if (R == 3) {
        // y_q = sum_m a_m e^{sign 2 pi i m q / 3}
        constexpr S h = (S)0.86602540378443864676372317075293618; // sqrt(3)/2
        const C t1 = cadd(a[1], a[2]);
        const C t2 = cmake<C>(a[0].x - (S)0.5 * t1.x, a[0].y - (S)0.5 * t1.y);
        const C d = cmuli(csub(a[1], a[2]), sign); // sign i (a1 - a2)
        a[0] = cadd(a[0], t1);
        a[1] = cmake<C>(t2.x + h * d.x, t2.y + h * d.y);
        a[2] = cmake<C>(t2.x - h * d.x, t2.y - h * d.y);
    } else if (R == 5) {
        // cos and sin of 2 pi/5 and 4 pi/5: (sqrt5 - 1)/4, sqrt((5 + sqrt5)/8),
        // -(sqrt5 + 1)/4, sqrt((5 - sqrt5)/8)
        constexpr S c1 = (S)0.30901699437494742410229341718281906, s1 = (S)0.95105651629515357211643933337938214;
        constexpr S c2 = (S)-0.80901699437494742410229341718281906, s2 = (S)0.58778525229247312916870595463907277;
        const C b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
        const C d1 = cmuli(csub(a[1], a[4]), sign), d2 = cmuli(csub(a[2], a[3]), sign);
        const C p1 = cmake<C>(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
        const C p2 = cmake<C>(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
        const C q1 = cmake<C>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        const C q2 = cmake<C>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        a[0] = cadd(a[0], cadd(b1, b2));
        a[1] = cadd(p1, q1);
        a[4] = csub(p1, q1);
        a[2] = cadd(p2, q2);
        a[3] = csub(p2, q2);
    } else { // R == 8: two 4-point DFTs of the even and odd inputs
        constexpr S c = (S)0.70710678118654752440084436210484904; // sqrt(2)/2
        C e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
        dft<4>(e, sign);
        dft<4>(o, sign);
        // o_q *= e^{sign 2 pi i q/8}
        const C o1 = cmake<C>(c * (o[1].x - (S)sign * o[1].y), c * (o[1].y + (S)sign * o[1].x));
        const C o2 = cmuli(o[2], sign);
        const C o3 = cmake<C>(c * (-o[3].x - (S)sign * o[3].y), c * (-o[3].y + (S)sign * o[3].x));
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], o1);
        a[5] = csub(e[1], o1);
        a[2] = cadd(e[2], o2);
        a[6] = csub(e[2], o2);
        a[3] = cadd(e[3], o3);
        a[7] = csub(e[3], o3);
    }
}

// COL = false: sequence r occupies slots lp(r*N + i) and consecutive threads take
// consecutive butterflies of one sequence.  COL = true: element i of
// sequence r sits at i*nrows + r and consecutive threads take the same
// butterfly of consecutive sequences (nrows a power of two), so every
// 16-lane access is 16 consecutive slots.
template <int R, int LR, bool COL, typename C>
__device__ __forceinline__ void lds_pass(C *buf, int logN, int logNs, int nrows, int lrows, const C *tw, int sign)
{
    constexpr int BPT = 8 / R;
    const int lNR = logN - LR, NR = 1 << lNR, Ns = 1 << logNs, nb = nrows << lNR;
    const int lstep = logN - logNs - LR;
    C a[BPT][R];
    int base[BPT], rr[BPT];
#pragma unroll
    for (int i = 0; i < BPT; i++) {
        const int b = threadIdx.x + i * (int)blockDim.x;
        base[i] = -1;
        rr[i] = 0;
        if (b < nb) {
            int r, j;
            if (COL) {
                r = b & (nrows - 1);
                j = b >> lrows;
            } else {
                r = b >> lNR;
                j = b & (NR - 1);
            }
            const int k = j & (Ns - 1);
#pragma unroll
            for (int m = 0; m < R; m++)
                a[i][m] = COL ? buf[((j + m * NR) << lrows) + r] : buf[lp((r << logN) + j + m * NR)];
#pragma unroll
            for (int m = 1; m < R; m++) {
                C w = tw[(m * k) << lstep];
                if (sign > 0) w.y = -w.y;
                a[i][m] = cmul(a[i][m], w);
            }
            dft<R>(a[i], sign);
            base[i] = (j - k) * R + k;
            rr[i] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; i++)
        if (base[i] >= 0)
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int e = base[i] + (q << logNs);
                if (COL)
                    buf[(e << lrows) + rr[i]] = a[i][q];
                else
                    buf[lp((rr[i] << logN) + e)] = a[i][q];
            }
    __syncthreads();
}

// nrows sequences of length N = 2^logN in buf, in place
template <bool COL, typename C>
__device__ __forceinline__ void fft_lds(C *buf, int logN, int nrows, const C *tw, int sign)
{
    const int lrows = COL ? __builtin_ctz(nrows) : 0;
    int lNs = 0, left = logN;
    while (left >= 3) {
        lds_pass<8, 3, COL>(buf, logN, lNs, nrows, lrows, tw, sign);
        lNs += 3;
        left -= 3;
    }
    if (left == 2) lds_pass<4, 2, COL>(buf, logN, lNs, nrows, lrows, tw, sign);
    if (left == 1) lds_pass<2, 1, COL>(buf, logN, lNs, nrows, lrows, tw, sign);
}

// Sequences of any length N = 2^a 3^b 5^c (the factors of fft.f:501-837):
// the same self-sorting passes with the radices of `plan` (4 bits each,
// first pass lowest) and division in place of shifts.  A block's tile holds
// nrows whole sequences (nrows N <= 8 blockDim values).
template <int R, bool COL, typename C>
__device__ __forceinline__ void lds_pass_g(C *buf, int N, int Ns, int nrows, const C *tw, int sign)
{
    constexpr int BPT = (8 + R - 1) / R;
    const int NR = N / R, nb = nrows * NR, step = N / (Ns * R);
    C a[BPT][R];
    int base[BPT], rr[BPT];
#pragma unroll
    for (int i = 0; i < BPT; i++) {
        const int b = threadIdx.x + i * (int)blockDim.x;
        base[i] = -1;
        rr[i] = 0;
        if (b < nb) {
            int r, j;
            if (COL) {
                r = b % nrows;
                j = b / nrows;
            } else {
                r = b / NR;
                j = b - r * NR;
            }
            const int k = j % Ns;
#pragma unroll
            for (int m = 0; m < R; m++)
                a[i][m] = COL ? buf[(j + m * NR) * nrows + r] : buf[lp(r * N + j + m * NR)];
#pragma unroll
            for (int m = 1; m < R; m++) {
                C w = tw[m * k * step];
                if (sign > 0) w.y = -w.y;
                a[i][m] = cmul(a[i][m], w);
            }
            dft<R>(a[i], sign);
            base[i] = (j - k) * R + k;
            rr[i] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; i++)
        if (base[i] >= 0)
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int e = base[i] + q * Ns;
                if (COL)
                    buf[e * nrows + rr[i]] = a[i][q];
                else
                    buf[lp(rr[i] * N + e)] = a[i][q];
            }
    __syncthreads();
}

template <bool COL, typename C>
__device__ __forceinline__ void fft_lds_g(C *buf, int N, unsigned plan, int nrows, const C *tw, int sign)
{
    int Ns = 1;
    for (; plan; plan >>= 4) {
        switch (plan & 15) {
        case 8: lds_pass_g<8, COL>(buf, N, Ns, nrows, tw, sign); Ns *= 8; break;
        case 5: lds_pass_g<5, COL>(buf, N, Ns, nrows, tw, sign); Ns *= 5; break;
        case 4: lds_pass_g<4, COL>(buf, N, Ns, nrows, tw, sign); Ns *= 4; break;
        case 3: lds_pass_g<3, COL>(buf, N, Ns, nrows, tw, sign); Ns *= 3; break;
        default: lds_pass_g<2, COL>(buf, N, Ns, nrows, tw, sign); Ns *= 2; break;
        }
    }
}

// filter_legenpen (filtering.f:729-808) of the NL values x of one column in
// place: a_k = c_k sig_k sum_q w_q x_q P_k(x_q), x_q = sum_k a_k P_k(x_q);
// the second sum is accumulated as each a_k becomes available (same order
// of terms), so AL (LDS) is read once per k.
template <int NL>
__device__ __forceinline__ void legfilter(double (&x)[NL], const double *wgs, const double *ALs,
                                          const double *sgs)
{
    constexpr int nzm = NL - 1;
    double y[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) {
        y[q] = wgs[q] * x[q];
        x[q] = 0.;
    }
#pragma unroll 1
    for (int k = 0; k < NL; k++) {
        double sum = 0.;
#pragma unroll
        for (int q = 0; q < NL; q++) sum = sum + y[q] * ALs[k * NL + q];
        const double c = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
        const double a = c * sgs[k];
#pragma unroll
        for (int q = 0; q < NL; q++) x[q] = x[q] + a * ALs[k * NL + q];
    }
}

template <int NL>
__device__ __forceinline__ void stage_tables(double *wgs, double *ALs, double *sgs, const double *wg,
                                             const double *AL, const double *sig)
{
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) ALs[e] = AL[e];
    for (int e = threadIdx.x; e < NL; e += blockDim.x) {
        wgs[e] = wg[e];
        sgs[e] = sig[e];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// x transform (FFT991, fft.f:1361-1502, as used by horfft.f:21 / 58):
// real length-N rows <-> N/2+1 coefficients A_k + i B_k, forward normalised
// by 1/N with B_0 = B_{N/2} = 0 (fft.f:1293-1305), inverse ignoring Im of the
// k=0 and k=N/2 coefficients (fft.f:1218-1230) and zeroing the 2 pad words
// (fft.f:1483-1490).  Implemented as one complex FFT of length M=N/2 on the
// even/odd-packed row with LDS-staged radix-2 passes.
// ---------------------------------------------------------------------------
// maxcalc extrema of one field (various.f:749-851), ties to the later point
// complex loads and stores of a field stored in FP64 or FP32 (the mixed
// path keeps the physical-space working set in FP32); arithmetic stays FP64
__device__ __forceinline__ double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ __forceinline__ double2 ld2(const float *p)
{
    const float2 v = *reinterpret_cast<const float2 *>(p);
    return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ void st2(double *p, double2 v) { *reinterpret_cast<double2 *>(p) = v; }
__device__ __forceinline__ void st2(float *p, double2 v)
{
    *reinterpret_cast<float2 *>(p) = make_float2((float)v.x, (float)v.y);
}

// (x, lx) replaces (b, lb) when it is more extreme, or equal and later;
// selects, not a branch (no divergent region per comparison)
__device__ __forceinline__ void mc_take(bool mn, double x, int lx, double &b, int &lb)
{
    const bool take = (mn ? x < b : x > b) | ((x == b) & (lx > lb));
    b = take ? x : b;
    lb = take ? lx : lb;
}

// AX: forward, the batch's fields have flux products; inverse (P2 only), the
// batch's fields have x gradients, and the tile holds R/2 rows of the field
// and the same rows of i kx f, transformed by the same passes from one load
// of the coefficients.  MAPS: the batch
// has plane maps (else rows are addressed arithmetically, so no row address
// waits on a table load and a thread's loads are all in flight together).
// TI, TO: storage type of the inputs (f, ax) and of the outputs (g, dx, axo).
// P2: M and ny are powers of two (shifts and masks, radix-8/4/2 passes);
// else any M = 2^a 3^b 5^c with the radices of `plan`, and a tile of
// R = floor(8 blockDim / M) rows that may leave slots of the block unused.
template <int DIR, bool AX = false, bool MAPS = false, typename TI = double, typename TO = double, bool P2 = true>
__global__ __launch_bounds__(kFT) void k_fft_x(FieldBatch fb, int nrows, int ny, int nxpp, int M, int logM,
                                               unsigned plan, int R, const double2 *__restrict__ twM,
                                               const double2 *__restrict__ twNg, double scaley, McFuse mf)
{
    using C = tile_t<TI, TO>;
    extern __shared__ double2 smx[];
    // the tile: 8 blockDim slots (R M of them used; lp permutes within
    // aligned runs of 128, so a partial tile's slots stay inside the block)
    C *buf = reinterpret_cast<C *>(smx);
    C *tw = buf + 8 * (int)blockDim.x;
    const int nval = R * M;
    // R entries: the kx-Nyquist coefficient of each row (inverse, for
    // maxcalc), 16-byte aligned after the M twiddles
    double2 *nyq = reinterpret_cast<double2 *>(tw + (sizeof(C) == 8 ? (M + 1) & ~1 : M));
    // the length-N twiddles and the wavenumbers kx of k = 0..M, staged in LDS
    // with tw so that no per-element global table load sits between a
    // thread's data loads or stores
    double2 *twN = nyq + R;
    double *kws = reinterpret_cast<double *>(twN + M + 1);
    // maxcalc's per-level tables for the block's rows (w00, rm, rdz, ush),
    // staged with the others so the epilogue after the stores reads LDS
    double *mcv = kws + M + 1;
    const TI *base = reinterpret_cast<const TI *>(fb.f[blockIdx.y]);
    TO *obase = reinterpret_cast<TO *>(fb.g[blockIdx.y]);
    const int row0 = blockIdx.x * (AX ? R >> 1 : R); // first row of the block's field rows
    const int N = 2 * M;
    const int lny = P2 ? __builtin_ctz(ny) : 0;
    // value idx of the tile -> (row r, index k); row -> (plane, row j)
    auto rowof = [&](int idx) -> int { return P2 ? idx >> logM : idx / M; };
    auto colof = [&](int idx, int r) -> int { return P2 ? idx & (M - 1) : idx - r * M; };
    auto planeof = [&](int row) -> int { return P2 ? row >> lny : row / ny; };
    auto jof = [&](int row, int pl) -> int { return P2 ? row & (ny - 1) : row - pl * ny; };
    auto fft_rows = [&](int nr) {
        if (P2)
            fft_lds<false>(buf, logM, nr, tw, DIR);
        else
            fft_lds_g<false>(buf, M, plan, nr, tw, DIR);
    };
    // logical row -> stored row of the source / destination (plane maps;
    // with maps, the offsets of the block's rows are staged in LDS below)
    auto srow_g = [&](int row) -> long {
        int lpl = planeof(row), j = jof(row, lpl);
        return (long)((fb.smap ? fb.smap[lpl] : lpl) * ny + j) * nxpp;
    };
    auto drow_g = [&](int row) -> long {
        int lpl = planeof(row), j = jof(row, lpl);
        return (long)((fb.dmap ? fb.dmap[lpl] : lpl) * ny + j) * nxpp;
    };
    long *rmap = reinterpret_cast<long *>(kws + M + 1 + 4 * R); // 2 R: source, destination offsets
    const bool staged = MAPS && !AX;
    auto srow = [&](int row) -> long {
        if (!MAPS) return (long)row * nxpp;
        return staged ? rmap[row - row0] : srow_g(row);
    };
    auto drow = [&](int row) -> long {
        if (!MAPS) return (long)row * nxpp;
        return staged ? rmap[R + row - row0] : drow_g(row);
    };
    for (int t = threadIdx.x; t < M; t += blockDim.x) tw[t] = toc<C>(twM[t]);
    for (int t = threadIdx.x; t <= M; t += blockDim.x) {
        twN[t] = twNg[t];
        kws[t] = fb.kw ? fb.kw[t] : 0.;
    }
    if (staged)
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            const int row = row0 + r < nrows ? row0 + r : nrows - 1;
            rmap[r] = srow_g(row);
            rmap[R + r] = drow_g(row);
        }
    if (DIR > 0 && mf.part)
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            const int row = row0 + r, k = planeof(row < nrows ? row : nrows - 1);
            mcv[r] = mf.w00 ? mf.w00[k] : 0.;
            mcv[R + r] = mf.rm ? mf.rm[k] : 0.;
            mcv[2 * R + r] = mf.rdz ? mf.rdz[k] : 0.;
            mcv[3 * R + r] = mf.ush ? mf.ush[k] : 0.;
        }
    __syncthreads();
    // inverse pre-processing (FFT991 inverse, fft.f:1208-1281): the N/2+1
    // coefficients A_k + i B_k of a row -> the packed complex sequence; with
    // deriv the coefficients are first multiplied by i kx (the x gradient of
    // solver.f:86-103)
    // coefficients a = A_k + i B_k, b = A_{M-k} + i B_{M-k} of row `row`
    // (tile row r) -> element k of the packed sequence
    auto pre_inverse = [&](double2 a, double2 b, int k, int r, int row, bool deriv) -> double2 {
        if (!deriv && k == 0) nyq[r] = b; // kx-Nyquist coefficient of the row
        if (deriv) {
            const double ka = kws[k], kb = kws[M - k];
            a = make_double2(-ka * a.y, ka * a.x);
            b = make_double2(-kb * b.y, kb * b.x);
        }
        double ar, ai, br, bi;
        if (k == 0) {
            ar = a.x; ai = 0.; br = b.x; bi = 0.;
        } else {
            ar = a.x; ai = a.y;
            br = b.x; bi = -b.y;
        }
        double er = ar + br, ei = ai + bi;
        double dr = ar - br, di = ai - bi;
        double2 w = twN[k];
        double c = w.x, s = -w.y; // e^{+2 pi i k/N}
        double orr = dr * c - di * s, oi = dr * s + di * c;
        double2 Z = make_double2(er - oi, ei + orr);
        if (row >= nrows) Z = make_double2(0., 0.);
        return Z;
    };
    auto load_inverse = [&](bool deriv) {
        // all of a thread's loads first (rows past the end read the last row
        // and are zeroed below), so they are in flight together
        double2 va[8], vb[8];
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            const int r = rowof(idx), k = colof(idx, r), row = row0 + r;
            const TI *p = base + srow(row < nrows ? row : nrows - 1);
            va[it] = ld2(p + 2 * k);
            vb[it] = ld2(p + 2 * (M - k));
        }
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (!P2 && idx >= nval) continue;
            const int r = rowof(idx), k = colof(idx, r), row = row0 + r;
            buf[lp(idx)] = toc<C>(pre_inverse(va[it], vb[it], k, r, row, deriv));
        }
    };
    auto store_inverse = [&](TO *ob) {
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (!P2 && idx >= nval) continue;
            int r = rowof(idx), m = colof(idx, r), row = row0 + r;
            if (row >= nrows) continue;
            st2(ob + drow(row) + 2 * m, tod(buf[lp(idx)]));
        }
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            int row = row0 + r;
            if (row < nrows) st2(ob + drow(row) + N, make_double2(0., 0.));
        }
    };
    if (DIR > 0 && !AX && fb.dx[blockIdx.y]) {
        // i kx f first: a transform in place overwrites its input rows
        load_inverse(true);
        __syncthreads();
        fft_rows(R);
        store_inverse(reinterpret_cast<TO *>(fb.dx[blockIdx.y]));
        __syncthreads();
    }
    if (DIR < 0) {
        // forward: FFT991 normalised by 1/N with B_0 = B_M = 0 (fft.f:1293-1305)
        const double sc = 1.0 / (double)N;
        auto load_fwd = [&](const TI *src) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                if (!P2 && idx >= nval) continue;
                int r = rowof(idx), m = colof(idx, r), row = row0 + r;
                double2 z = ld2(src + srow(row < nrows ? row : nrows - 1) + 2 * m);
                if (row >= nrows) z = make_double2(0., 0.);
                buf[lp(idx)] = toc<C>(z);
            }
        };
        // coefficient k from Z_k and Z_{M-k}; the thread of k = 0 also
        // forms k = M (both from Z_0)
        auto coef = [&](int r, int k) -> double2 {
            const int r0 = P2 ? r << logM : r * M;
            double2 Zk = tod(buf[lp(r0 + (k == M ? 0 : k))]);
            double2 Zm = tod(buf[lp(r0 + (k == 0 ? 0 : M - k))]);
            double er = 0.5 * (Zk.x + Zm.x), ei = 0.5 * (Zk.y - Zm.y);
            double dr = Zk.x - Zm.x, di = Zk.y + Zm.y;
            double orr = 0.5 * di, oi = -0.5 * dr;
            double2 w = twN[k];
            double xr = er + (orr * w.x - oi * w.y);
            double xi = ei + (orr * w.y + oi * w.x);
            double A = sc * xr, B = sc * xi;
            if (k == 0 || k == M) B = 0.0;
            if (k == M) A = A * scaley; // horfft.f:28-30 on the untransformed Nyquist column
            return make_double2(A, B);
        };
        // + i kx P (the x part of the flux divergence, solver.f:985-1063, 1163-1196)
        auto addk = [&](double2 c, double2 q, int k) {
            const double kx = kws[k];
            return make_double2(c.x + (-kx * q.y), c.y + kx * q.x);
        };
        if (AX) {
            // the tile holds H rows of the field and the same H rows of its
            // flux product P: both are transformed by the same passes and
            // combined from LDS (with R odd the last row of the tile is unused)
            const int H = R >> 1, rowA = blockIdx.x * H;
            const TI *acc = reinterpret_cast<const TI *>(fb.ax[blockIdx.y]);
            TO *axo = reinterpret_cast<TO *>(fb.axo[blockIdx.y]);
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                if (!P2 && idx >= nval) continue;
                const int r = rowof(idx), m = colof(idx, r);
                const bool isp = r >= H;
                const int row = rowA + (isp ? r - H : r);
                double2 z = ld2((isp ? acc : base) + srow(row < nrows ? row : nrows - 1) + 2 * m);
                if (row >= nrows || (!P2 && r >= 2 * H)) z = make_double2(0., 0.);
                buf[lp(idx)] = toc<C>(z);
            }
            __syncthreads();
            fft_rows(R);
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                const int r = rowof(idx), k = colof(idx, r), row = rowA + r;
                if (r >= H || row >= nrows) continue;
                TO *orow = obase + drow(row);
                const double2 q = coef(r + H, k);
                st2(orow + 2 * k, addk(coef(r, k), q, k));
                if (axo) st2(axo + drow(row) + 2 * k, q); // P's own transform
                if (k == 0) {
                    const double2 qm = coef(r + H, M);
                    st2(orow + 2 * M, addk(coef(r, M), qm, M));
                    if (axo) st2(axo + drow(row) + 2 * M, qm);
                }
            }
        } else {
            load_fwd(base);
            __syncthreads();
            fft_rows(R);
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                if (!P2 && idx >= nval) continue;
                int r = rowof(idx), k = colof(idx, r), row = row0 + r;
                if (row >= nrows) continue;
                TO *orow = obase + drow(row);
                st2(orow + 2 * k, coef(r, k));
                if (k == 0) st2(orow + 2 * M, coef(r, M));
            }
        }
    } else {
        // field rows of the tile: the first 8 value slots of each thread, or
        // the first 4 when the other 4 hold the rows' i kx f
        constexpr int NIT = AX ? 4 : 8;
        if (AX) {
            TO *dxo = reinterpret_cast<TO *>(fb.dx[blockIdx.y]);
            const int H = R >> 1;
            const int half = 4 * (int)blockDim.x; // value slot idx + half: tile row r + H, same k
            double2 va[4], vb[4];
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                const int r = rowof(idx), k = colof(idx, r), row = row0 + r;
                const TI *p = base + srow(row < nrows ? row : nrows - 1);
                va[it] = ld2(p + 2 * k);
                vb[it] = ld2(p + 2 * (M - k));
            }
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                const int r = rowof(idx), k = colof(idx, r), row = row0 + r;
                buf[lp(idx)] = toc<C>(pre_inverse(va[it], vb[it], k, r, row, false));
                buf[lp(idx + half)] = toc<C>(pre_inverse(va[it], vb[it], k, r + H, row, true));
            }
            __syncthreads();
            fft_rows(R);
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                const int rr = rowof(idx), m = colof(idx, rr);
                const bool d = rr >= H;
                const int row = row0 + (d ? rr - H : rr);
                if (row >= nrows) continue;
                st2((d ? dxo : obase) + drow(row) + 2 * m, tod(buf[lp(idx)]));
            }
            for (int r = threadIdx.x; r < R; r += blockDim.x) {
                const bool d = r >= H;
                const int row = row0 + (d ? r - H : r);
                if (row < nrows) st2((d ? dxo : obase) + drow(row) + N, make_double2(0., 0.));
            }
        } else {
            load_inverse(false);
            __syncthreads();
            fft_rows(R);
            store_inverse(obase);
        }
        if (mf.part && mf.fid0 + (int)blockIdx.y < 4) { // maxcalc reductions of u, v, w, temp
            const int f = mf.fid0 + blockIdx.y, nz = planeof(nrows);
            const double big = 1.79769313486231570e308;
            double b0 = big, b1 = -big, b2 = -big; // min, max of the value; max of the CFL number
            int l0 = -1, l1 = -1, l2 = -1;
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int idx = threadIdx.x + it * (int)blockDim.x;
                if (!P2 && idx >= nval) continue;
                const int r = rowof(idx), m = colof(idx, r), row = row0 + r;
                if (row >= nrows) continue;
                const int k = planeof(row), j = jof(row, k);
                const double2 z = tod(buf[lp(idx)]);
                const double c = mf.cny * nyq[r].x;
                double xe = z.x + c, xo = z.y - c; // points i = 2m, 2m+1
                if (f == 2 && mf.w00) {
                    xe = xe + mcv[r];
                    xo = xo + mcv[r];
                }
                const int L = ((2 * m) * ny + j) * nz + k, L1 = L + ny * nz;
                // + umwv, wmwv, rhomwv (zero unless set), + rhomean for rho
                const double *mwf = f == 0 ? mf.umwv : f == 2 ? mf.wmwv : f == 3 ? mf.rhomwv : nullptr;
                double ae = 0.0, ao = 0.0;
                if (mwf) {
                    const long mi = (long)k * nxpp + 2 * m;
                    ae = mwf[mi];
                    ao = mwf[mi + 1];
                }
                double ve = xe + ae, vo = xo + ao;
                if (f == 3) {
                    ve = ve + mcv[R + r];
                    vo = vo + mcv[R + r];
                }
                mc_take(true, ve, L, b0, l0);
                mc_take(true, vo, L1, b0, l0);
                mc_take(false, ve, L, b1, l1);
                mc_take(false, vo, L1, b1, l1);
                if (f < 3) { // CFL numbers; the shear current advects in x; no CFLZ at the top point
                    const double cf = f == 0 ? mf.cx : f == 1 ? mf.cy : (k == nz - 1 ? 0.0 : mf.dt * mcv[2 * R + r]);
                    const double sh = (f == 0 && mf.ush) ? mcv[3 * R + r] : 0.0;
                    mc_take(false, fabs((sh + xe) * cf), L, b2, l2);
                    mc_take(false, fabs((sh + xo) * cf), L1, b2, l2);
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                double x;
                int lx;
                x = __shfl_down(b0, off, 64); lx = __shfl_down(l0, off, 64); mc_take(true, x, lx, b0, l0);
                x = __shfl_down(b1, off, 64); lx = __shfl_down(l1, off, 64); mc_take(false, x, lx, b1, l1);
                x = __shfl_down(b2, off, 64); lx = __shfl_down(l2, off, 64); mc_take(false, x, lx, b2, l2);
            }
            __shared__ double rv[kFT / 64][3];
            __shared__ int rl[kFT / 64][3];
            const int wv = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) {
                rv[wv][0] = b0; rv[wv][1] = b1; rv[wv][2] = b2;
                rl[wv][0] = l0; rl[wv][1] = l1; rl[wv][2] = l2;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int w = 1; w < (int)blockDim.x / 64; w++) {
                    mc_take(true, rv[w][0], rl[w][0], b0, l0);
                    mc_take(false, rv[w][1], rl[w][1], b1, l1);
                    mc_take(false, rv[w][2], rl[w][2], b2, l2);
                }
                // quantities: 0 umin 1 umax 2 vmin 3 vmax 4 wmin 5 wmax 6 tmin
                // 7 tmax 8 cflx 9 cfly 10 cflz
                // partial slots per R/2 rows with P2 (the paired tiles' rows):
                // a block of R rows fills two, the second with no candidate
                const int slot = (!AX && P2) ? 2 * (int)blockIdx.x : (int)blockIdx.x;
                for (int q = 0; q < ((!AX && P2) ? 2 : 1); q++) {
                    double *p = mf.part + (size_t)(slot + q) * (2 * kMcQ);
                    p[4 * f] = q ? big : b0;
                    p[4 * f + 1] = q ? -1.0 : (double)l0;
                    p[4 * f + 2] = q ? -big : b1;
                    p[4 * f + 3] = q ? -1.0 : (double)l1;
                    if (f < 3) {
                        p[2 * (8 + f)] = q ? -big : b2;
                        p[2 * (8 + f) + 1] = q ? -1.0 : (double)l2;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// y transform (CFFT99, fft.f:1-406, horfft.f:25 / 55): complex length-ny
// transform of columns ii < nxh (the kx-Nyquist column is skipped); forward
// scaled by 1/ny (horfft.f:28-30).  CB columns per block staged in LDS.
// ---------------------------------------------------------------------------
// TI, TO: storage type of the inputs (f, ay) and of the outputs (g, dy).
// P2: ny is a power of two; else ny = 2^a 3^b 5^c with the radices of
// `plan` and CB = floor(8 blockDim / ny) columns per block.
// One-dimensional grid: ncb column blocks per plane, npl planes per field,
// the (plane, field) units padded to a multiple of 8.
template <int DIR, typename TI = double, typename TO = double, bool P2 = true>
__global__ __launch_bounds__(kFT) void k_fft_y(FieldBatch fb, int nxpp, int nxh, int ny, int logNy, unsigned plan,
                                               int CB, const double2 *__restrict__ twY, double scaley, int ncb,
                                               int npl)
{
    // workgroup -> (column block, plane, field) such that the column blocks
    // of one plane are consecutive workgroups of one XCD (workgroup w runs on
    // XCD w mod 8): the cache lines that neighbouring column blocks share in
    // every row are then read into one L2, by workgroups that run together
    const int wg = blockIdx.x, xcd = wg & 7, slot = wg >> 3;
    const int cbi = slot % ncb, unit = (slot / ncb) * 8 + xcd;
    if (unit >= npl * fb.n) return;
    const int lpl = unit % npl, fz = unit / npl;
    using C = tile_t<TI, TO>;
    extern __shared__ double2 smy[];
    C *buf = reinterpret_cast<C *>(smy);
    C *tw = buf + CB * ny;
    double *kys = reinterpret_cast<double *>(tw + ny); // ky per row, staged like tw
    const TI *plane = reinterpret_cast<const TI *>(fb.f[fz]) + (long)(fb.smap ? fb.smap[lpl] : lpl) * nxpp * ny;
    const long opl = (long)(fb.dmap ? fb.dmap[lpl] : lpl) * nxpp * ny;
    TO *oplane = reinterpret_cast<TO *>(fb.g[fz]) + opl;
    const int ii0 = cbi * CB;
    const int nval = CB * ny; // values of the tile: value idx is column idx % CB of row idx / CB
    const int lcb = P2 ? __builtin_ctz(CB) : 0;
    auto rowof = [&](int idx) -> int { return P2 ? idx >> lcb : idx / CB; };
    auto colof = [&](int idx, int j) -> int { return P2 ? idx & (CB - 1) : idx - j * CB; };
    auto fft_cols = [&]() {
        if (P2)
            fft_lds<true>(buf, logNy, CB, tw, DIR);
        else
            fft_lds_g<true>(buf, ny, plan, CB, tw, DIR);
    };
    for (int t = threadIdx.x; t < ny; t += blockDim.x) {
        tw[t] = toc<C>(twY[t]);
        kys[t] = fb.kw ? fb.kw[t] : 0.;
    }
    __syncthreads();
    // columns ii < nxh of the plane into LDS; with deriv multiplied by i ky
    // (the y gradient of solver.f:86-103)
    auto load = [&](const TI *src, bool deriv) {
        // all loads first (columns past nxh are clamped and their values
        // dropped), so they are in flight together; with P2 the tile is
        // CB x ny = 8 blockDim values, else values idx >= CB ny are unused
        double2 v[8];
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            int j = rowof(idx);
            const int ii = ii0 + colof(idx, j);
            if (!P2 && j >= ny) j = ny - 1;
            v[it] = ld2(src + (long)j * nxpp + 2 * (ii < nxh ? ii : nxh - 1));
        }
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (!P2 && idx >= nval) continue;
            int j = rowof(idx), c = colof(idx, j), ii = ii0 + c;
            double2 z = v[it];
            if (ii >= nxh) z = make_double2(0., 0.);
            if (deriv) {
                const double ky = kys[j];
                z = make_double2(-ky * z.y, ky * z.x);
            }
            buf[idx] = toc<C>(z);
        }
    };
    // forward with a second input P: + i ky P-hat (the y part of the flux
    // divergence, solver.f:985-1063, 1163-1196), P-hat kept in registers
    const TI *accp = (DIR < 0 && fb.ay[fz])
                         ? reinterpret_cast<const TI *>(fb.ay[fz]) + (long)(fb.smap ? fb.smap[lpl] : lpl) * nxpp * ny
                         : nullptr;
    double2 pb[8];
    auto store = [&](TO *op, bool add) {
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (!P2 && idx >= nval) continue;
            int j = rowof(idx), c = colof(idx, j), ii = ii0 + c;
            if (ii >= nxh) continue;
            double2 z = tod(buf[idx]);
            if (DIR < 0) {
                z.x = scaley * z.x;
                z.y = scaley * z.y;
                if (add) {
                    const double ky = kys[j];
                    z.x = z.x + (-ky * pb[it].y);
                    z.y = z.y + ky * pb[it].x;
                }
            }
            st2(op + (long)j * nxpp + 2 * ii, z);
        }
    };
    TO *dyp = (DIR > 0 && fb.dy[fz]) ? reinterpret_cast<TO *>(fb.dy[fz]) + opl : nullptr;
    if (dyp) { // i ky f first: a transform in place overwrites its input
        load(plane, true);
        __syncthreads();
        fft_cols();
        store(dyp, false);
        __syncthreads();
    }
    if (accp) {
        load(accp, false);
        __syncthreads();
        fft_cols();
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            double2 z = make_double2(0., 0.);
            if (idx < nval) {
                z = tod(buf[idx]);
                z.x = scaley * z.x;
                z.y = scaley * z.y;
            }
            pb[it] = z;
        }
        __syncthreads();
    }
    load(plane, false);
    __syncthreads();
    fft_cols();
    store(oplane, accp != nullptr);
    // the untransformed kx-Nyquist column: copied when out of place, plus
    // i ky P when accumulating, and times i ky into the derivative output
    const bool inplace = (const void *)plane == (const void *)oplane;
    if (cbi == 0 && (!inplace || dyp || accp))
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            const TI *p = plane + (long)j * nxpp + 2 * nxh;
            double pr = p[0], pi = p[1];
            if (accp) {
                const double ky = kys[j];
                const TI *a = accp + (long)j * nxpp + 2 * nxh;
                pr = pr + (-ky * (double)a[1]);
                pi = pi + ky * (double)a[0];
            }
            if (!inplace || accp) st2(oplane + (long)j * nxpp + 2 * nxh, make_double2(pr, pi));
            if (dyp) {
                const double ky = kys[j];
                st2(dyp + (long)j * nxpp + 2 * nxh, make_double2(-ky * pi, ky * pr));
            }
        }
}

template <bool MAPS, typename TI, typename TO, bool P2>
static void fft_x_launch(const Geo &g, const FieldBatch &b, int nk, int dir, hipStream_t st, const McFuse *mf)
{
    const int R = 8 * g.ft / g.M; // rows per block (fft_threads_ok: R >= 2)
    int nrows = g.ny * nk;
    dim3 grid((nrows + R - 1) / R, b.n);
    using C = std::conditional_t<std::is_same<TI, float>::value || std::is_same<TO, float>::value, float2, double2>;
    const size_t mt = sizeof(C) == 8 ? (size_t)((g.M + 1) & ~1) : (size_t)g.M; // twiddles, padded to 16 bytes
    size_t sh = sizeof(C) * ((size_t)8 * g.ft + mt) + sizeof(double2) * ((size_t)R + g.M + 1) +
                sizeof(double) * (g.M + 1 + 4 * R) + sizeof(long) * 2 * R;
    McFuse none;
    if (dir < 0) {
        // fields with a flux product share the tile with it (H rows of
        // each): they get their own launch
        FieldBatch ba = b, bp = b;
        ba.n = bp.n = 0;
        for (int q = 0; q < b.n; q++) {
            FieldBatch &t = b.ax[q] ? ba : bp;
            t.f[t.n] = b.f[q];
            t.g[t.n] = b.g[q];
            t.ax[t.n] = b.ax[q];
            t.axo[t.n] = b.axo[q];
            t.ay[t.n] = b.ay[q];
            t.dx[t.n] = b.dx[q];
            t.dy[t.n] = b.dy[q];
            t.n++;
        }
        const int H = R / 2;
        if (ba.n)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_x<-1, true, MAPS, TI, TO, P2>), dim3((nrows + H - 1) / H, ba.n),
                               dim3(g.ft), sh, st, ba, nrows, g.ny, g.nxpp, g.M, g.logM, g.planM, R, g.twM, g.twN,
                               g.scaley, none);
        if (bp.n)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_x<-1, false, MAPS, TI, TO, P2>), dim3(grid.x, bp.n), dim3(g.ft),
                               sh, st, bp, nrows, g.ny, g.nxpp, g.M, g.logM, g.planM, R, g.twM, g.twN, g.scaley, none);
    } else {
        McFuse m = mf ? *mf : none;
        // fields with an x gradient (a leading run of the batch) share their
        // tiles with it (R/2 rows of each), the rest take whole tiles; with
        // P2 the maxcalc partials come per R/2 rows either way (mc_blocks)
        int nd = 0;
        while (nd < b.n && b.dx[nd]) nd++;
        bool lead = P2 && nd > 0;
        for (int q = nd; q < b.n; q++) lead = lead && !b.dx[q];
        if (lead) {
            FieldBatch bd = b, br = b;
            bd.n = nd;
            br.n = b.n - nd;
            for (int q = 0; q < br.n; q++) {
                br.f[q] = b.f[nd + q];
                br.g[q] = b.g[nd + q];
                br.dx[q] = nullptr;
            }
            McFuse mr = m;
            mr.fid0 = m.fid0 + nd;
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_x<1, true, MAPS, TI, TO, P2>), dim3(2 * grid.x, bd.n),
                               dim3(g.ft), sh, st, bd, nrows, g.ny, g.nxpp, g.M, g.logM, g.planM, R, g.twM, g.twN,
                               g.scaley, m);
            if (br.n)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_x<1, false, MAPS, TI, TO, P2>), dim3(grid.x, br.n),
                                   dim3(g.ft), sh, st, br, nrows, g.ny, g.nxpp, g.M, g.logM, g.planM, R, g.twM,
                                   g.twN, g.scaley, mr);
        } else {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_x<1, false, MAPS, TI, TO, P2>), grid, dim3(g.ft), sh, st, b,
                               nrows, g.ny, g.nxpp, g.M, g.logM, g.planM, R, g.twM, g.twN, g.scaley, m);
        }
    }
}

template <bool P2>
static void fft_x_dispatch(const Geo &g, const FieldBatch &b, int nk, int dir, hipStream_t st, const McFuse *mf,
                           int prec)
{
    if (prec == kPrecF32) // FP32 in and out (mixed path; no plane maps)
        fft_x_launch<false, float, float, P2>(g, b, nk, dir, st, mf);
    else if (b.smap || b.dmap)
        fft_x_launch<true, double, double, P2>(g, b, nk, dir, st, mf);
    else
        fft_x_launch<false, double, double, P2>(g, b, nk, dir, st, mf);
}

void launch_fft_x(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st, const McFuse *mf, int prec)
{
    if (g.p2)
        fft_x_dispatch<true>(g, b, nk, dir, st, mf, prec);
    else
        fft_x_dispatch<false>(g, b, nk, dir, st, mf, prec);
}

// maxcalc partial slots of the inverse x transform: one per R rows, two with
// P2 (the tiles that pair a field with its x gradient hold R/2 of its rows)
int mc_blocks(const Geo &g, int ft)
{
    const int R = 8 * (ft > 0 ? ft : g.ft) / g.M;
    return (g.p2 ? 2 : 1) * ((g.ny * g.nz + R - 1) / R);
}

template <bool P2>
static void fft_y_dispatch(const Geo &g, const FieldBatch &b, int nk, int dir, hipStream_t st, int prec)
{
    const int CB = 8 * g.ft / g.ny; // columns per block (fft_threads_ok: CB >= 1)
    const int ncb = (g.nxh + CB - 1) / CB;
    const long units = ((long)nk * b.n + 7) / 8 * 8;
    dim3 grid((unsigned)(ncb * units));
    // the mixed path's tile and twiddles are FP32
    const size_t cs = prec != kPrecF64 ? sizeof(float2) : sizeof(double2);
    size_t sh = cs * ((size_t)CB * g.ny + g.ny) + sizeof(double) * g.ny;
    // mixed path: spectral FP64 <-> FP32 intermediate, or FP32 right-hand sides
    if (dir < 0 && prec == kPrecF32S)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_y<-1, float, float, P2>), grid, dim3(g.ft), sh, st, b, g.nxpp,
                           g.nxh, g.ny, g.logNy, g.planY, CB, g.twY, g.scaley, ncb, nk);
    else if (dir < 0 && prec == kPrecF32)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_y<-1, float, double, P2>), grid, dim3(g.ft), sh, st, b, g.nxpp,
                           g.nxh, g.ny, g.logNy, g.planY, CB, g.twY, g.scaley, ncb, nk);
    else if (prec == kPrecF32)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_y<1, double, float, P2>), grid, dim3(g.ft), sh, st, b, g.nxpp,
                           g.nxh, g.ny, g.logNy, g.planY, CB, g.twY, g.scaley, ncb, nk);
    else if (dir < 0)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_y<-1, double, double, P2>), grid, dim3(g.ft), sh, st, b, g.nxpp,
                           g.nxh, g.ny, g.logNy, g.planY, CB, g.twY, g.scaley, ncb, nk);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_y<1, double, double, P2>), grid, dim3(g.ft), sh, st, b, g.nxpp,
                           g.nxh, g.ny, g.logNy, g.planY, CB, g.twY, g.scaley, ncb, nk);
}

void launch_fft_y(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st, int prec)
{
    if (g.ndims != 3) return;
    if (g.p2)
        fft_y_dispatch<true>(g, b, nk, dir, st, prec);
    else
        fft_y_dispatch<false>(g, b, nk, dir, st, prec);
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

// The same with the GLL order fixed at compile time: the column is loaded in
// one burst and D is read from LDS row by row.
template <int NL>
__global__ __launch_bounds__(256) void k_dpdz_t(DpdzArgs A, long nxppy, const double *__restrict__ D,
                                                const double *__restrict__ zh)
{
    __shared__ double Ds[NL * NL];
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) Ds[e] = D[e];
    __syncthreads();
    const long ij = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxppy) return;
    const int ks = blockIdx.y;
    const double fac = 2. / zh[ks];
    const long b0 = (long)ks * NL * nxppy + ij;
    for (int q = 0; q < A.nq; q++) {
        double col[NL];
        const double *a = A.a[q], *b = A.b[q];
        if (b) {
            const double c = A.c[q];
#pragma unroll
            for (int t = 0; t < NL; t++) col[t] = (c * a[b0 + t * nxppy]) * b[b0 + t * nxppy];
        } else {
#pragma unroll
            for (int t = 0; t < NL; t++) col[t] = a[b0 + t * nxppy];
        }
        double *o = A.out[q] + b0;
#pragma unroll 1
        for (int k1 = 0; k1 < NL; k1++) {
            double sm = 0.0;
#pragma unroll
            for (int t = 0; t < NL; t++) sm = sm + Ds[k1 * NL + t] * col[t];
            o[k1 * nxppy] = sm * fac;
        }
    }
}

static void run_dpdz(const Geo &g, const DpdzArgs &a, hipStream_t st)
{
    dim3 grid((unsigned)((g.nxppy + 255) / 256), g.nsubd);
#define DZ(NL) hipLaunchKernelGGL(k_dpdz_t<NL>, grid, dim3(256), 0, st, a, g.nxppy, g.D, g.zh)
    switch (g.nzloc) {
        IW_NL_SWITCH(DZ)
    default: hipLaunchKernelGGL(k_dpdz, grid, dim3(256), 0, st, a, g.nxppy, g.nzloc, g.D, g.zh); break;
    }
#undef DZ
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

// The vertical flux derivatives of rhsfm above and, from the same columns,
// the z-gradients of u, v, w (and optionally rho) that solver.f:74-77 forms
// in spectral space before the inverse transforms: dpdz acts along z and
// horfft along x and y, so the two commute and the gradients are taken here
// from the physical fields instead of being transformed.  One pass: every
// input column is read once; the products are formed as in k_dpdz_t.
struct ConvArgs {
    const double *u, *v, *w, *t; // physical u, v, w, rho
    double *fu, *fv, *fw, *ft;   // d(-uw/2)/dz, d(-vw/2)/dz, d(-ww/2)/dz, d(-w rho)/dz
    double *du, *dv, *dw, *dt;   // du/dz, dv/dz, dw/dz, drho/dz (dt may be null)
};

// (2/zh) D of a column x (into ox, may be null) and of its product with w,
// (cf x) w (into op), both from each row of D read once
template <int NL>
__device__ __forceinline__ void dz_two(const double *Ds, const double (&x)[NL], const double (&w)[NL], double cf,
                                       double *ox, double *op, long nxppy, double fac)
{
#pragma unroll 1
    for (int k1 = 0; k1 < NL; k1++) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int t = 0; t < NL; t++) {
            const double d = Ds[k1 * NL + t];
            s1 = s1 + d * x[t];
            s2 = s2 + d * ((cf * x[t]) * w[t]);
        }
        if (ox) ox[k1 * nxppy] = s1 * fac;
        op[k1 * nxppy] = s2 * fac;
    }
}

template <int NL>
__global__ __launch_bounds__(256) void k_dpdz_conv_t(ConvArgs A, long nxppy, const double *__restrict__ D,
                                                     const double *__restrict__ zh)
{
    __shared__ double Ds[NL * NL];
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) Ds[e] = D[e];
    __syncthreads();
    const long ij = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxppy) return;
    const int ks = blockIdx.y;
    const double fac = 2. / zh[ks];
    const long b0 = (long)ks * NL * nxppy + ij;
    // each column: its own gradient and that of its product with w
    // ((c x) w, which for rho equals k_dpdz_t's (-w) rho exactly)
    double w[NL], c[NL];
#pragma unroll
    for (int t = 0; t < NL; t++) w[t] = A.w[b0 + t * nxppy];
    dz_two<NL>(Ds, w, w, -0.5, A.dw + b0, A.fw + b0, nxppy, fac);
    const double *src[3] = {A.u, A.v, A.t};
    double *dsts[3] = {A.du, A.dv, A.dt}, *fl[3] = {A.fu, A.fv, A.ft};
#pragma unroll 1
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int t = 0; t < NL; t++) c[t] = src[q][b0 + t * nxppy];
        dz_two<NL>(Ds, c, w, q < 2 ? -0.5 : -1.0, dsts[q] ? dsts[q] + b0 : nullptr, fl[q] + b0, nxppy, fac);
    }
}

void launch_dpdz_conv_grad(const Geo &g, double *const *uvwt, double *const *out4, double *const *dz4,
                           hipStream_t st)
{
    ConvArgs A;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3];
    A.fu = out4[0]; A.fv = out4[1]; A.fw = out4[2]; A.ft = out4[3];
    A.du = dz4[0]; A.dv = dz4[1]; A.dw = dz4[2]; A.dt = dz4[3];
    dim3 grid((unsigned)((g.nxppy + 255) / 256), g.nsubd);
#define DC(NL) hipLaunchKernelGGL(k_dpdz_conv_t<NL>, grid, dim3(256), 0, st, A, g.nxppy, g.D, g.zh)
    switch (g.nzloc) {
        IW_NL_SWITCH(DC)
    default: { // other orders: the products, then the plain gradients
        launch_dpdz_conv(g, uvwt, out4, st);
        launch_dpdz3(g, uvwt, dz4, st);
        if (dz4[3]) launch_dpdz(g, uvwt[3], dz4[3], st);
    }
    }
#undef DC
}

// mean w = 0 (solver.f:66-69)
// w(kx=0,ky=0) = 0 (solver.f:66-69); the removed real part is kept in w00
// (maxcalc's values are recovered from convec's transform with it)
__global__ void k_zero_meanw(double *w, double *w00, long nxppy, int nz)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nz) {
        w00[k] = w[(long)k * nxppy];
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

void launch_zero_meanw(const Geo &g, double *w, double *w00, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_meanw, dim3((g.nz + 63) / 64), dim3(64), 0, st, w, w00, g.nxppy, g.nz);
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
    MeanTerms mt;
    double xkappa;
    int mean;
};

// rhsfm's physical-space part at one point n (plane k, x index i): u, v, w,
// tp and the nine gradients G (du/dx, du/dy, du/dz, dv/.., dw/..) at the
// point, su, sv, sw the vertical flux derivatives, oxv d(-w rho)/dz; writes
// S_u..S_T and the seven flux products.  T: storage and arithmetic type
// (double for the plain path, float for the mixed path, which has no
// mean-wave terms); the pointers of A hold T arrays.
template <typename T>
__device__ __forceinline__ void rhs_point(const RhsArgs &A, long n, int k, int i, long nxppy, int nxpp, int nx,
                                          int nz, int nl, const double *__restrict__ zh,
                                          const double *__restrict__ zc, const double *__restrict__ xg, T u, T v,
                                          T w, T tp, const T (&G)[9], T su, T sv, T sw, T oxv)
{
    const T *Au = reinterpret_cast<const T *>(A.u), *Av = reinterpret_cast<const T *>(A.v);
    const T *Aw = reinterpret_cast<const T *>(A.w), *At = reinterpret_cast<const T *>(A.t);
    T *S0 = reinterpret_cast<T *>(A.S[0]), *S1 = reinterpret_cast<T *>(A.S[1]);
    T *S2 = reinterpret_cast<T *>(A.S[2]), *S3 = reinterpret_cast<T *>(A.S[3]);
    const T half = T(0.5), four = T(4.), omz = T(A.omz);
    T st = T(0.);
    su = su - half * (u * G[0] + v * (G[1] - four * omz) + w * G[2]);
    sv = sv - half * (u * (G[3] + four * omz) + v * G[4] + w * G[5]);
    sw = sw - half * (u * G[6] + v * G[7] + w * G[8]);
    sw = sw + T(A.bcoef) * tp;
    const int k1 = k % nl, ks = k / nl;
    // total vertical velocity at the interfaces, perturbation + mean wave
    // (solver.f:723-724)
    const long mi = (long)k * nxpp + i, nm = (long)nxpp * nz;
    const T wmw = A.mean ? T(A.mt.mw[nm + mi]) : T(0.);
    if (k1 == 0) {
        T wb = (k == 0) ? T(0.) : w + wmw;
        if (wb > T(0.)) {
            T tau1 = T(0.5 / A.omega * (2. / zh[ks]));
            long m = n - nxppy;
            su = su - tau1 * (wb * u - wb * Au[m]);
            sv = sv - tau1 * (wb * v - wb * Av[m]);
            sw = sw - tau1 * (wb * w - wb * Aw[m]);
            st = st - tau1 * (wb * tp - wb * At[m]);
        }
    }
    if (k1 == nl - 1) {
        T wt = (k == nz - 1) ? T(0.) : w + wmw;
        if (wt < T(0.)) {
            T g0 = fabs(wt);
            T tau2 = T(0.5 / A.omega * (2. / zh[ks]));
            long m = n + nxppy;
            su = su - tau2 * (g0 * u - g0 * Au[m]);
            sv = sv - tau2 * (g0 * v - g0 * Av[m]);
            sw = sw - tau2 * (g0 * w - g0 * Aw[m]);
            st = st - tau2 * (g0 * tp - g0 * At[m]);
        }
    }
    if constexpr (std::is_same<T, double>::value) {
        if (A.mean) { // mean-wave and shear-current terms, solver.f:848-931
            const double *mw = A.mt.mw;
            const double advmean = mw[mi] + A.mt.ush[k];
            su = su - advmean * G[0] - wmw * G[2];
            sv = sv - advmean * G[3] - wmw * G[5];
            sw = sw - advmean * G[6] - wmw * G[8];
            st = st - advmean * A.mt.dtdx[n] - wmw * A.mt.dtdz[n];
            const double dusheardz = A.mt.dush[k];
            su = su - u * mw[2 * nm + mi] - w * mw[3 * nm + mi] - w * dusheardz;
            sw = sw - u * mw[4 * nm + mi] - w * mw[5 * nm + mi];
            st = st - u * mw[7 * nm + mi] - w * mw[8 * nm + mi];
            st = st + A.xkappa * 0.0; // d2rhomeandz2 = 0 (forcing.f:167-183)
        }
    }
    if (A.paddle && i < nx) {
        double zz = zc[k], xx = xg[i];
        double arg = A.xkcrit * xx - A.xomega * A.t_now;
        double wzp = A.xmcrit * cos(A.xmcrit * zz), wz = sin(A.xmcrit * zz);
        su = su + T((-A.inv_xkcrit) * A.xacrit_umax * wzp * sin(arg) * A.umax_zdim * A.phit);
        sw = sw + T(A.xacrit_umax * wz * cos(arg) * A.umax_zdim * A.phit);
        st = st + T(A.xacrit_umax * wz * sin(arg) * A.rfac * A.umax_zdim * A.phit);
    }
    st = st + oxv - w * T(A.rg);
    S0[n] = su;
    S1[n] = sv;
    S2[n] = sw;
    S3[n] = st;
    T *const *pr = reinterpret_cast<T *const *>(A.pr);
    pr[0][n] = (-half * u) * u;
    pr[1][n] = (-half * u) * v;
    pr[2][n] = (-half * v) * v;
    pr[3][n] = (-half * u) * w;
    pr[4][n] = (-half * v) * w;
    pr[5][n] = -(u * tp);
    pr[6][n] = -(v * tp);
}

template <typename T>
__global__ __launch_bounds__(256) void k_rhs_phys(RhsArgs A, long ntot, long nxppy, int nxpp, int nx,
                                                  int nz, int nl, const double *__restrict__ zh,
                                                  const double *__restrict__ zc,
                                                  const double *__restrict__ xg)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ntot) return;
    const T *Au = reinterpret_cast<const T *>(A.u), *Av = reinterpret_cast<const T *>(A.v);
    const T *Aw = reinterpret_cast<const T *>(A.w), *At = reinterpret_cast<const T *>(A.t);
    T G[9];
#pragma unroll
    for (int q = 0; q < 9; q++) G[q] = reinterpret_cast<const T *>(A.g[q])[n];
    const T *S0 = reinterpret_cast<const T *>(A.S[0]), *S1 = reinterpret_cast<const T *>(A.S[1]);
    const T *S2 = reinterpret_cast<const T *>(A.S[2]);
    const int k = (int)(n / nxppy);
    const int i = (int)(n % nxpp);
    rhs_point<T>(A, n, k, i, nxppy, nxpp, nx, nz, nl, zh, zc, xg, Au[n], Av[n], Aw[n], At[n], G, S0[n], S1[n],
                 S2[n], reinterpret_cast<const T *>(A.ox)[n]);
}

// The z flux derivatives and z gradients of k_dpdz_conv_t and rhsfm's
// physical part of k_rhs_phys in one pass (plain path, no mean-wave terms, no
// closure).  A work-group is 64 columns of one subdomain: their u, v, w, rho go
// to LDS, then wavefront q takes rows q, q + 4, ... of every column, forms the
// seven sums of dz_two for the row (same terms, same order) and hands them to
// rhs_point, so neither the sums nor the gradients go through HBM.  T: the
// storage and arithmetic type as in k_rhs_phys (float: the mixed path, whose
// separate z contraction runs on the matrix cores with other rounding).
template <int NL, typename T>
__global__ __launch_bounds__(256) void k_conv_rhs(RhsArgs A, long nxppy, int nxpp, int nx, int nz,
                                                  const double *__restrict__ D, const double *__restrict__ zh,
                                                  const double *__restrict__ zc, const double *__restrict__ xg)
{
    __shared__ T cu[NL][64], cv[NL][64], cw[NL][64], ct[NL][64];
    const T *Au = reinterpret_cast<const T *>(A.u), *Av = reinterpret_cast<const T *>(A.v);
    const T *Aw = reinterpret_cast<const T *>(A.w), *At = reinterpret_cast<const T *>(A.t);
    // the row a wavefront works on is the same in all its lanes, so D's
    // entries come through scalar loads (no LDS for D: three work-groups fit
    // a CU)
    const int lane = threadIdx.x & 63, wq = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long ij = (long)blockIdx.x * 64 + lane;
    const int ks = blockIdx.y;
    const bool live = ij < nxppy;
    const long b0 = (long)ks * NL * nxppy + (live ? ij : 0);
    for (int t = wq; t < NL; t += 4) {
        const long o = b0 + (long)t * nxppy;
        cu[t][lane] = Au[o];
        cv[t][lane] = Av[o];
        cw[t][lane] = Aw[o];
        ct[t][lane] = At[o];
    }
    __syncthreads();
    if (!live) return;
    const T fac = T(2. / zh[ks]);
    const T h = T(-0.5), one = T(-1.0);
    const int i = (int)(ij % nxpp);
    for (int k1 = wq; k1 < NL; k1 += 4) {
        // dz_two of w (-0.5), u (-0.5), v (-0.5) and rho (-1.0, gradient unused)
        T gw = T(0.), fw = T(0.), gu = T(0.), fu = T(0.), gv = T(0.), fv = T(0.), ft = T(0.);
        // a few rows of the columns in registers at a time
#pragma unroll 5
        for (int t = 0; t < NL; t++) {
            const T d = T(D[k1 * NL + t]);
            const T wt = cw[t][lane], ut = cu[t][lane], vt = cv[t][lane], tt = ct[t][lane];
            gw = gw + d * wt;
            fw = fw + d * ((h * wt) * wt);
            gu = gu + d * ut;
            fu = fu + d * ((h * ut) * wt);
            gv = gv + d * vt;
            fv = fv + d * ((h * vt) * wt);
            ft = ft + d * ((one * tt) * wt);
        }
        const long n = b0 + (long)k1 * nxppy;
        T G[9];
        G[0] = reinterpret_cast<const T *>(A.g[0])[n];
        G[1] = reinterpret_cast<const T *>(A.g[1])[n];
        G[2] = gu * fac;
        G[3] = reinterpret_cast<const T *>(A.g[3])[n];
        G[4] = reinterpret_cast<const T *>(A.g[4])[n];
        G[5] = gv * fac;
        G[6] = reinterpret_cast<const T *>(A.g[6])[n];
        G[7] = reinterpret_cast<const T *>(A.g[7])[n];
        G[8] = gw * fac;
        rhs_point<T>(A, n, ks * NL + k1, i, nxppy, nxpp, nx, nz, NL, zh, zc, xg, cu[k1][lane], cv[k1][lane],
                     cw[k1][lane], ct[k1][lane], G, fu * fac, fv * fac, fw * fac, ft * fac);
    }
}

// the scalar parameters of rhsfm's physical-space part
static void rhs_args(const Geo &g, const iw_params &P, double t, RhsArgs &A)
{
    A.xkappa = P.xkappa;
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
}

void launch_rhs_phys(const Geo &g, const iw_params &P, double t, double *const *uvwt,
                     double *const *gr, double *const *S, const double *ox, double *const *pr,
                     hipStream_t st, const MeanTerms *mt)
{
    RhsArgs A;
    rhs_args(g, P, t, A);
    A.mean = mt && mt->mw;
    if (A.mean) A.mt = *mt;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3];
    for (int q = 0; q < 9; q++) A.g[q] = gr[q];
    for (int q = 0; q < 4; q++) A.S[q] = S[q];
    A.ox = ox;
    for (int q = 0; q < 7; q++) A.pr[q] = pr[q];
    hipLaunchKernelGGL(k_rhs_phys<double>, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, A, g.ntot,
                       g.nxppy, g.nxpp, g.nx, g.nz, g.nzloc, g.zh, g.z, g.x);
}

template <typename T>
static bool conv_rhs_launch(const Geo &g, const iw_params &P, double t, T *const *uvwt, T *const *gr, T *const *S,
                            T *const *pr, hipStream_t st)
{
    if (g.ndims != 3 || !nl_specialised(g.nzloc)) return false;
    RhsArgs A;
    rhs_args(g, P, t, A);
    A.mean = 0;
    A.u = reinterpret_cast<const double *>(uvwt[0]); A.v = reinterpret_cast<const double *>(uvwt[1]);
    A.w = reinterpret_cast<const double *>(uvwt[2]); A.t = reinterpret_cast<const double *>(uvwt[3]);
    for (int q = 0; q < 9; q++) A.g[q] = reinterpret_cast<const double *>(gr[q]);
    for (int q = 0; q < 4; q++) A.S[q] = reinterpret_cast<double *>(S[q]);
    A.ox = nullptr;
    for (int q = 0; q < 7; q++) A.pr[q] = reinterpret_cast<double *>(pr[q]);
    dim3 grid((unsigned)((g.nxppy + 63) / 64), g.nsubd);
#define CR(NL)                                                                                                   \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv_rhs<NL, T>), grid, dim3(256), 0, st, A, g.nxppy, g.nxpp, g.nx, g.nz, \
                       g.D, g.zh, g.z, g.x)
    switch (g.nzloc) {
        IW_NL_SWITCH(CR)
    default: return false;
    }
#undef CR
    return true;
}

bool launch_conv_rhs(const Geo &g, const iw_params &P, double t, double *const *uvwt, double *const *gr,
                     double *const *S, double *const *pr, hipStream_t st)
{
    return conv_rhs_launch<double>(g, P, t, uvwt, gr, S, pr, st);
}

bool launch_conv_rhs(const Geo &g, const iw_params &P, double t, float *const *uvwt, float *const *gr,
                     float *const *S, float *const *pr, hipStream_t st)
{
    return conv_rhs_launch<float>(g, P, t, uvwt, gr, S, pr, st);
}

void launch_rhs_phys_f32(const Geo &g, const iw_params &P, double t, float *const *uvwt, float *const *gr,
                         float *const *S, const float *ox, float *const *pr, hipStream_t st)
{
    RhsArgs A;
    rhs_args(g, P, t, A);
    A.mean = 0;
    A.u = reinterpret_cast<const double *>(uvwt[0]);
    A.v = reinterpret_cast<const double *>(uvwt[1]);
    A.w = reinterpret_cast<const double *>(uvwt[2]);
    A.t = reinterpret_cast<const double *>(uvwt[3]);
    for (int q = 0; q < 9; q++) A.g[q] = reinterpret_cast<const double *>(gr[q]);
    for (int q = 0; q < 4; q++) A.S[q] = reinterpret_cast<double *>(S[q]);
    A.ox = reinterpret_cast<const double *>(ox);
    for (int q = 0; q < 7; q++) A.pr[q] = reinterpret_cast<double *>(pr[q]);
    hipLaunchKernelGGL(k_rhs_phys<float>, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, A, g.ntot,
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

// b = i kx a (spectral x derivative, solver.f:86-103)
__global__ void k_dx(const double *__restrict__ a, double *b, long ncplx, int nxhp,
                     const double *__restrict__ xw)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncplx) return;
    const double kx = xw[(int)(c % nxhp)];
    const double re = a[2 * c], im = a[2 * c + 1];
    b[2 * c] = -kx * im;
    b[2 * c + 1] = kx * re;
}

void launch_dx(const Geo &g, const double *a, double *b, hipStream_t st)
{
    long ncplx = g.ntot / 2;
    hipLaunchKernelGGL(k_dx, dim3((unsigned)((ncplx + 255) / 256)), dim3(256), 0, st, a, b, ncplx, g.nxhp, g.xw);
}

// wnl at bottom/top (solver.f:270-284)
template <typename T>
__global__ void k_wnl(const T *sw, const T *swn, const T *swnm, double f1, double f2, double f3, double *wnl,
                      long nxppy, int nz)
{
    long ij = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nxppy) return;
    long kt = (long)(nz - 1) * nxppy + ij;
    wnl[ij] = f1 * (double)sw[ij] + f2 * (double)swn[ij] + f3 * (double)swnm[ij];
    wnl[nxppy + ij] = f1 * (double)sw[kt] + f2 * (double)swn[kt] + f3 * (double)swnm[kt];
}

void launch_wnl(const Geo &g, const double *sw, const double *swn, const double *swnm, const double *fac,
                double *wnl, hipStream_t st)
{
    hipLaunchKernelGGL(k_wnl<double>, dim3((unsigned)((g.nxppy + 255) / 256)), dim3(256), 0, st, sw, swn, swnm,
                       fac[0], fac[1], fac[2], wnl, g.nxppy, g.nz);
}

void launch_wnl(const Geo &g, const float *sw, const float *swn, const float *swnm, const double *fac,
                double *wnl, hipStream_t st)
{
    hipLaunchKernelGGL(k_wnl<float>, dim3((unsigned)((g.nxppy + 255) / 256)), dim3(256), 0, st, sw, swn, swnm,
                       fac[0], fac[1], fac[2], wnl, g.nxppy, g.nz);
}

// rotvisc (solver.f:1210-1253) at the bottom point of subdomain 1 and the top
// point of subdomain nsubd for levels n, n-1, n-2 (solver.f:320-394)
struct RotArgs {
    const double *lev[9]; // u v w | un vn wn | unm vnm wnm
};

// NL > 0: the GLL order at compile time, the column loops unrolled so that a
// thread's 300 loads are in flight together (the same sums in the same order)
template <int NL>
__global__ void k_rotvisc(RotArgs A, double f1, double f2, double f3, double xnu, double *wviscrot,
                          int nmode, int nxhp, int ny, int nl_, int nsubd, long nxppy,
                          const double *__restrict__ D, const double *__restrict__ zh,
                          const double *__restrict__ xw, const double *__restrict__ yw)
{
    const int nl = NL > 0 ? NL : nl_;
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
#pragma unroll
    for (int L = 0; L < 3; L++) {
        const double *u = A.lev[3 * L], *v = A.lev[3 * L + 1], *w = A.lev[3 * L + 2];
        double sur = 0., sui = 0., svr = 0., svi = 0.;
#pragma unroll
        for (int t = 0; t < (NL > 0 ? NL : nl); t++) {
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
#define RV(NL)                                                                                                     \
    hipLaunchKernelGGL(k_rotvisc<NL>, dim3((2 * nmode + 127) / 128), dim3(128), 0, st, A, fac[0], fac[1], fac[2], \
                       P.xnu, wviscrot, nmode, g.nxhp, g.ny, g.nzloc, g.nsubd, g.nxppy, g.D, g.zh, g.xw, g.yw)
    switch (g.nzloc) {
        IW_NL_SWITCH(RV)
    default: RV(0);
    }
#undef RV
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
    const double *S[4][3]; // per field (blockIdx.z): S, Sn, Snm
    const float *S32[4][3]; // the same in FP32 (k_tile_t<.., 1, true>)
    const double *F[4][3]; // f, fn, fnm
    double *out[4];
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
    const int ks = blockIdx.y, fz = blockIdx.z;
    const double *const *Sf = A.S[fz], *const *Ff = A.F[fz];
    double *outf = A.out[fz];
    const long c0 = (long)blockIdx.x * kMB;
    const int nmb = (int)((ncplx_plane - c0) < kMB ? (ncplx_plane - c0) : kMB);
    const int W2 = 2 * kMB;
    for (int idx = threadIdx.x; idx < nl * W2; idx += blockDim.x) {
        int t = idx / W2, e = idx - t * W2;
        double val = 0.;
        if (e < 2 * nmb) {
            long o = (long)(ks * nl + t) * nxppy + 2 * c0 + e;
            if (A.mode == 1) {
                double s = A.f1 * Sf[0][o] + A.f2 * Sf[1][o] + A.f3 * Sf[2][o];
                val = A.dt * s + A.b1 * Ff[0][o] + A.b2 * Ff[1][o] + A.b3 * Ff[2][o];
            } else {
                val = outf[o];
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
            fourier_filter_value(fr, fi, fcf[c0 + cc]);
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
        if (e < 2 * nmb) outf[(long)(ks * nl + t) * nxppy + 2 * c0 + e] = tile[idx];
    }
}

// The same per (mode, subdomain) with the GLL order fixed at compile time:
// lane = one mode, its re/im columns in registers (16-byte loads of each
// row), the Fourier filter with one sincos, the Legendre tables in LDS.
// One column (the re or the im part of a mode) per lane, consecutive lanes
// consecutive columns, so that a wavefront reads and writes whole rows; the
// Fourier filter needs both parts of a value and takes the partner lane's
// with a lane exchange (lanes 2i and 2i+1 hold re and im of one mode).  The
// Legendre filter is applied one output row at a time from the inputs in
// registers (the sums in the order of legfilter_mat).
// S32: the right-hand-side histories are FP32 (widened exactly before use)
template <int NL, int MODE, bool S32 = false>
__global__ __launch_bounds__(256) void k_tile_t(TileArgs A, long nxppy, const double *__restrict__ fcf,
                                                const double *__restrict__ FL)
{
    __shared__ double FLs[NL * NL];
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) FLs[e] = FL[e];
    __syncthreads();
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ks = blockIdx.y, fz = blockIdx.z;
    if (col >= nxppy) return; // nxppy is even: both lanes of a mode stay or leave together
    const bool im = col & 1;
    const double *const *Sf = A.S[fz], *const *Ff = A.F[fz];
    double *outf = A.out[fz];
    const long b0 = (long)ks * NL * nxppy + col;
    double x[NL];
    if (MODE == 1) {
#pragma unroll
        for (int t = 0; t < NL; t++) {
            const long o = b0 + (long)t * nxppy;
            double s0, s1, s2;
            if (S32) {
                s0 = (double)A.S32[fz][0][o];
                s1 = (double)A.S32[fz][1][o];
                s2 = (double)A.S32[fz][2][o];
            } else {
                s0 = Sf[0][o];
                s1 = Sf[1][o];
                s2 = Sf[2][o];
            }
            const double f0 = Ff[0][o], f1 = Ff[1][o], f2 = Ff[2][o];
            const double sr = A.f1 * s0 + A.f2 * s1 + A.f3 * s2;
            x[t] = A.dt * sr + A.b1 * f0 + A.b2 * f1 + A.b3 * f2;
        }
    } else {
#pragma unroll
        for (int t = 0; t < NL; t++) x[t] = outf[b0 + (long)t * nxppy];
    }
    if (A.ifourier) { // Fourier magnitude/phase filter, filtering.f:852-901
        const double sigf = fcf[col >> 1];
#pragma unroll
        for (int t = 0; t < NL; t++) {
            const double other = __shfl_xor(x[t], 1, 64);
            double fr = im ? other : x[t], fi = im ? x[t] : other;
            fourier_filter_value(fr, fi, sigf);
            x[t] = im ? fi : fr;
        }
    }
#pragma unroll
    for (int q = 0; q < NL; q++) {
        double o = x[q];
        if (A.ilegen) { // filtering.f:905-956, the two sums as one matrix
            o = 0.;
#pragma unroll
            for (int t = 0; t < NL; t++) o = o + FLs[t * NL + q] * x[t];
        }
        outf[b0 + (long)q * nxppy] = o;
    }
}

static void run_tile(const Geo &g, const TileArgs &A, int nf, hipStream_t st)
{
    long ncp = (long)g.nxhp * g.ny;
    dim3 gt((unsigned)((g.nxppy + 255) / 256), g.nsubd, nf);
#define TT(NL)                                                                                                 \
    if (A.mode == 1 && A.S32[0][0])                                                                            \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_t<NL, 1, true>), gt, dim3(256), 0, st, A, g.nxppy, g.fcf,  \
                           g.FL);                                                                              \
    else if (A.mode == 1)                                                                                      \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_t<NL, 1>), gt, dim3(256), 0, st, A, g.nxppy, g.fcf, g.FL); \
    else                                                                                                       \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_t<NL, 0>), gt, dim3(256), 0, st, A, g.nxppy, g.fcf, g.FL)
    if (nl_specialised(g.nzloc)) {
        switch (g.nzloc) {
            IW_NL_SWITCH(TT)
        default: break;
        }
        return;
    }
#undef TT
    dim3 grid((unsigned)((ncp + kMB - 1) / kMB), g.nsubd, nf);
    hipLaunchKernelGGL(k_tile, grid, dim3(256), 0, st, A, ncp, g.nxppy, g.nzloc, g.fcf, g.wg, g.AL, g.sig);
}

void launch_bdf_filter(const Geo &g, double *const *S12, double *const *F12, double *const *out4, const double *fac,
                       const double *bdf, double dt, int filt, hipStream_t st)
{
    TileArgs A = {};
    for (int f = 0; f < 4; f++) {
        for (int q = 0; q < 3; q++) {
            A.S[f][q] = S12[3 * f + q];
            A.F[f][q] = F12[3 * f + q];
        }
        A.out[f] = out4[f];
    }
    A.f1 = fac[0]; A.f2 = fac[1]; A.f3 = fac[2];
    A.b1 = bdf[0]; A.b2 = bdf[1]; A.b3 = bdf[2];
    A.dt = dt;
    A.mode = 1;
    A.ifourier = filt;
    A.ilegen = filt;
    run_tile(g, A, 4, st);
}

// only the specialised orders (those of the mixed path) read FP32 histories
void launch_bdf_filter(const Geo &g, float *const *S12, double *const *F12, double *const *out4, const double *fac,
                       const double *bdf, double dt, int filt, hipStream_t st)
{
    TileArgs A = {};
    for (int f = 0; f < 4; f++) {
        for (int q = 0; q < 3; q++) {
            A.S32[f][q] = S12[3 * f + q];
            A.F[f][q] = F12[3 * f + q];
        }
        A.out[f] = out4[f];
    }
    A.f1 = fac[0]; A.f2 = fac[1]; A.f3 = fac[2];
    A.b1 = bdf[0]; A.b2 = bdf[1]; A.b3 = bdf[2];
    A.dt = dt;
    A.mode = 1;
    A.ifourier = filt;
    A.ilegen = filt;
    if (nl_specialised(g.nzloc)) run_tile(g, A, 4, st);
}

__global__ __launch_bounds__(256) void k_cvt_df(const double *__restrict__ a, float *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ __launch_bounds__(256) void k_cvt_fd(const float *__restrict__ a, double *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

void launch_cvt(const double *a, float *b, long n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_cvt_df, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n);
}

void launch_cvt(const float *a, double *b, long n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_cvt_fd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n);
}

void launch_filter3d(const Geo &g, double *f, int ifourier, int ilegen, hipStream_t st)
{
    TileArgs A = {};
    A.out[0] = f;
    A.mode = 0;
    A.ifourier = ifourier;
    A.ilegen = ilegen;
    run_tile(g, A, 1, st);
}

// ---------------------------------------------------------------------------
// interface_avg3d physical-space pass (penalty.f:729-772): grid-stride
// over the (i,j) points of every interface; the counts are summed per block
// and added with one atomic per block (per-wavefront atomics on the four
// counters serialised).
// ---------------------------------------------------------------------------
// The field holds, per interface, the four physical planes km-1, km, k, k+1
// (compact: interface i at planes 4i..4i+3).
__global__ __launch_bounds__(256) void k_iface(FieldBatch fb, int *count, long nxppy, int nl, int nint)
{
    __shared__ int ws[4];
    double *f = fb.f[blockIdx.y];
    int hits = 0;
    (void)nl;
    const long n = nxppy * nint;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        int it = (int)(id / nxppy);
        long ij = id % nxppy;
        long k = 4L * it + 2, km = 4L * it + 1;
        double fsp = f[k * nxppy + ij], fsm = f[km * nxppy + ij];
        double favg2 = fsp + fsm;
        double c0 = fabs(fsp - fsm) / fabs(favg2);
        if (c0 > 0.005) {
            double ff = f[(km - 1) * nxppy + ij] * 0.5 + f[km * nxppy + ij] * 0. + f[k * nxppy + ij] * 0.
                        + f[(k + 1) * nxppy + ij] * 0.5;
            f[km * nxppy + ij] = ff;
            f[k * nxppy + ij] = ff;
            hits++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_down(hits, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = hits;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = ws[0] + ws[1] + ws[2] + ws[3];
        if (t) atomicAdd(&count[blockIdx.y], t);
    }
}

void launch_iface(const Geo &g, FieldBatch b, int *count, hipStream_t st)
{
    int nint = g.nsubd - 1;
    if (nint <= 0) return;
    long n = g.nxppy * nint;
    long nb = (n + 255) / 256;
    if (nb > 512) nb = 512;
    dim3 grid((unsigned)nb, b.n);
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
                                               const double *__restrict__ D2, double *AB, int *ipiv,
                                               double *sysinfo)
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
    bool swapped = false;
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
        swapped = swapped || (p != 0);
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
    if (lane == 0) sysinfo[4L * sys + 3] = swapped ? 1. : 0.;
#undef WM
}

void launch_factor(const Geo &g, const SysTables &S, int first, int count, hipStream_t st)
{
    if (count <= 0) return;
    size_t sh = sizeof(double) * ((size_t)kWR * kWS + 2 * g.nzloc * g.nzloc + 12 * g.nsubd);
    hipLaunchKernelGGL(k_factor, dim3(count), dim3(64), sh, st, first, g.nz, g.nzloc, g.nsubd, S.ldab,
                       S.coef, g.D, g.D2, S.AB, S.ipiv, S.sysinfo);
}

// One system, for the single-column solve of iw_penalty_solve: the factor
// of k_factor from its own coefficient and system tables (index 0), then
// SGBTRS with the interchanges (bandludcmp.f:1020-1036, STBSV
// ludcmp.f:4698-5046) of the re and im columns x[0..n), x[n..2n) on one
// lane (a test-size call; the step's solves are the batched sweeps).
void launch_factor_one(const Geo &g, int ldab, const double *coef, double *sysinfo, double *AB, int *ipiv,
                       hipStream_t st)
{
    size_t sh = sizeof(double) * ((size_t)kWR * kWS + 2 * g.nzloc * g.nzloc + 12 * g.nsubd);
    hipLaunchKernelGGL(k_factor, dim3(1), dim3(64), sh, st, 0, g.nz, g.nzloc, g.nsubd, ldab, coef, g.D, g.D2, AB,
                       ipiv, sysinfo);
}

__global__ void k_band_solve1(const double *__restrict__ AB, const int *__restrict__ ipiv, double *x, int n, int ldab,
                              int kl)
{
    if (threadIdx.x >= 2) return;
    double *b = x + (long)threadIdx.x * n;
    const int kv = 2 * kl;
    for (int j = 0; j < n - 1; j++) {
        const int lm = kl < n - 1 - j ? kl : n - 1 - j;
        const int jp = ipiv[j];
        if (jp != j) {
            const double t = b[j];
            b[j] = b[jp];
            b[jp] = t;
        }
        const double y = b[j];
        for (int t = 1; t <= lm; t++) b[j + t] = b[j + t] + AB[(long)j * ldab + kv + t] * (-y);
    }
    for (int j = n - 1; j >= 0; j--) {
        if (b[j] != 0.) b[j] = b[j] / AB[(long)j * ldab + kv];
        const double xj = b[j];
        const int lm = kv < j ? kv : j;
        for (int t = 1; t <= lm; t++) b[j - t] = b[j - t] - xj * AB[(long)j * ldab + kv - t];
    }
}

void launch_band_solve1(const Geo &g, int ldab, const double *AB, const int *ipiv, double *x, hipStream_t st)
{
    hipLaunchKernelGGL(k_band_solve1, dim3(1), dim3(64), 0, st, AB, ipiv, x, g.nz, ldab, g.nzloc);
}

// Packed copies of the structurally nonzero factor entries (tri_* layout in
// SysTables::FB / PB): one block per system; an LDS map from a packed
// position of a subdomain block to its (local row, t) lets consecutive
// threads write consecutive entries.  Entries of rows outside the matrix are
// stored as zero.  A nonzero outside the pattern, or a pressure interchange
// beyond the L column's reach, sets bad[0] (pressure factors) or bad[1] (the
// others): the host then keeps the band solvers that read AB for that kind.
__global__ __launch_bounds__(256) void k_compact(const double *__restrict__ AB, const int *__restrict__ ipiv,
                                                 double *FB, double *PB, int first, int np, int n, int nl,
                                                 int ldab, int kv, int kl, int *bad)
{
    // packed double -> (local row, t, part) of one subdomain block, t < 0 a pad
    __shared__ short map_r[2 * 1024], map_t[2 * 1024];
    const long sys = first + (long)blockIdx.x;
    const double *ab = AB + sys * n * ldab;
    const int *ip = ipiv + sys * n;
    const bool piv = sys < np;
    const int nsub = n / nl;
    const int sf = tri_sf(nl), sb = tri_sb(nl, piv), blk = 2 * (sf + sb);
    double *base = piv ? PB + sys * nsub * 2L * (sf + sb) : FB + sys * nsub * 2L * (sf + sb);
    if (blk > 2 * 1024) { // orders beyond the map (nl > 32): band solvers only
        if (threadIdx.x == 0) atomicOr(bad + (piv ? 0 : 1), 1);
        return;
    }
    for (int r = threadIdx.x; r < nl; r += blockDim.x) {
        const int ef = tri_ef(nl, r), eb = tri_eb(nl, r, piv);
        const int fo = 2 * tri_off_f(nl, r), bo = 2 * (sf + tri_off_b(nl, r, piv));
        for (int q = 0; q < 2 * tri_npf(nl, r); q++) {
            map_r[fo + q] = (short)r;
            map_t[fo + q] = (short)(q < ef ? q + 1 : -1); // L[t] at t - 1
        }
        for (int q = 0; q < 2 * tri_npb(nl, r, piv); q++) {
            map_r[bo + q] = (short)r;
            map_t[bo + q] = (short)(q <= eb ? 1000 + q : -1); // U[t] at t
        }
    }
    __syncthreads();
    // consecutive threads write consecutive packed entries: block layout is
    // nsub forward blocks then nsub backward blocks
    for (long e = threadIdx.x; e < (long)nsub * blk; e += blockDim.x) {
        const bool fwd = e < 2L * nsub * sf;
        const long o = fwd ? e : e - 2L * nsub * sf;
        const int per = fwd ? 2 * sf : 2 * sb;
        const int ks = (int)(o / per), q = (int)(o - (long)ks * per) + (fwd ? 0 : 2 * sf);
        const int r = map_r[q], tt = map_t[q], c = ks * nl + r;
        double v = 0.;
        if (tt > 0 && tt < 1000) {
            if (c + tt < n) v = ab[(long)c * ldab + kv + tt];
        } else if (tt >= 1000) {
            const int t = tt - 1000;
            if (c - t >= 0) v = ab[(long)c * ldab + kv - t];
        }
        base[e] = v;
    }
    // pattern check over the band (coalesced): zero outside it, and for the
    // pressure every interchange within the L column's reach
    int inter = 0;
    if (!piv)
        for (int c = threadIdx.x; c < n; c += blockDim.x) inter |= ip[c] != c;
    if (__syncthreads_or(inter)) return; // such a system sends the solves to the band path anyway
    int err = 0;
    for (long e = threadIdx.x; e < (long)n * ldab; e += blockDim.x) {
        const int c = (int)(e / ldab), row = (int)(e - (long)c * ldab), r = c % nl;
        const double v = ab[e];
        if (row > kv) { // L[t], t = row - kv
            const int t = row - kv;
            err |= (t > tri_ef(nl, r) && c + t < n && v != 0.);
        } else { // U[t], t = kv - row
            const int t = kv - row;
            err |= (t > tri_eb(nl, r, piv) && c - t >= 0 && v != 0.);
        }
    }
    if (piv)
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int d = ip[c] - c;
            err |= (d < 0) || (d > tri_ef(nl, c % nl));
        }
    (void)kl;
    if (err) atomicOr(bad + (piv ? 0 : 1), 1);
}

void launch_compact(const Geo &g, const SysTables &S, int first, int count, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_compact, dim3(count), dim3(256), 0, st, S.AB, S.ipiv, S.FB, S.PB, first, S.np, g.nz,
                       g.nzloc, S.ldab, S.kv, S.kl, S.tri_bad);
}

// ---------------------------------------------------------------------------
// Implicit step (calcuvwt.f:87-741 with update, calcuvwt.f:826-901) in two
// kernels over the cached factors:
//   k_pressure: one wavefront per mode -- divergence (divcolc, derv.f:35-85),
//     pressure Poisson solve with the wnl+wviscrot Neumann data, Legendre
//     filter, dphi/dz and the projection u* - dt grad(phi), written in place;
//   k_viscous: one wavefront per (mode, field) -- u, v, w and density
//     Helmholtz solves, filter, mean-u damping.
// Each solve is SGBTRS (bandludcmp.f:883-1069): the band window of the
// forward and backward sweeps rides in the lanes, the right-hand side column
// sits in LDS, and factor columns are streamed from HBM through a register
// ring kPF columns ahead of use.
// ---------------------------------------------------------------------------
constexpr int kPF = 8;

__device__ __forceinline__ double rdl(double v, int l)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct ModeArgs {
    double *u, *v, *w, *t;
    const double *wnl, *wviscrot;
    const double *AB;
    const int *ipiv;
    const double *sysinfo; // eps, tau1 bottom, tau2 top, bc code
    const int *modesys;
    const int *order; // block -> task (mode, or 4*mode+field); -1 = idle
    const double *D, *AL, *wg, *sig, *zh, *xw, *yw, *damp;
    double dt, xnu, xkappa;
    long nxppy;
    int nz, nl, ns, nxhp, ldab, kl, kv, cpen, noslip;
};

// Solve A x = b for the two columns b0, b1 (LDS) of one system; x overwrites b.
// Factor columns are streamed from global memory through a register ring kPF
// columns ahead of use (each lane holds its own row of the window).
__device__ void band_solve2(const ModeArgs &A, int sys, double *b0, double *b1)
{
    const int lane = threadIdx.x & 63;
    const int n = A.nz, kl = A.kl, kv = A.kv, ldab = A.ldab;
    const double *__restrict__ ab = A.AB + (long)sys * ldab * n;
    const int *__restrict__ ip = A.ipiv + (long)sys * n;
    // forward: L with row interchanges (bandludcmp.f:1020-1030)
    {
        const bool lact = lane >= 1 && lane <= kl;
        double v0 = (lane <= kl && lane < n) ? b0[lane] : 0.;
        double v1 = (lane <= kl && lane < n) ? b1[lane] : 0.;
        double ring[kPF];
#pragma unroll
        for (int q = 0; q < kPF; q++) ring[q] = (lact && q < n) ? ab[(kv + lane) + (long)ldab * q] : 0.;
        int ipv = (lane < n) ? ip[lane] : 0;
        for (int jb = 0; jb < n; jb += kPF) {
#pragma unroll
            for (int q = 0; q < kPF; q++) {
                const int j = jb + q;
                if (j >= n) break;
                if (q == 0 && jb > 0 && (jb & 63) == 0) ipv = (j + lane < n) ? ip[j + lane] : 0;
                const int lm = (kl < n - 1 - j) ? kl : n - 1 - j;
                const double lt = (lane <= lm) ? ring[q] : 0.;
                if (lact && j + kPF < n) ring[q] = ab[(kv + lane) + (long)ldab * (j + kPF)];
                const int nj = (j + kl + 1 < n) ? j + kl + 1 : n - 1;
                const double nb0 = b0[nj], nb1 = b1[nj];
                const int p = __builtin_amdgcn_readlane(ipv, j & 63) - j;
                const double x00 = rdl(v0, 0), x01 = rdl(v1, 0);
                const double bj0 = rdl(v0, p), bj1 = rdl(v1, p);
                if (lane == p) {
                    v0 = x00;
                    v1 = x01;
                }
                v0 = v0 + lt * (-bj0);
                v1 = v1 + lt * (-bj1);
                if (lane == 0) {
                    b0[j] = bj0;
                    b1[j] = bj1;
                }
                v0 = __shfl_down(v0, 1);
                v1 = __shfl_down(v1, 1);
                if (lane == kl) {
                    v0 = (j + kl + 1 < n) ? nb0 : 0.;
                    v1 = (j + kl + 1 < n) ? nb1 : 0.;
                }
            }
        }
    }
    __syncthreads();
    // backward: upper triangle with kv superdiagonals (STBSV, ludcmp.f:4698-5046)
    {
        const bool uact = lane <= kv;
        double v0 = (uact && n - 1 - lane >= 0) ? b0[n - 1 - lane] : 0.;
        double v1 = (uact && n - 1 - lane >= 0) ? b1[n - 1 - lane] : 0.;
        double ring[kPF];
#pragma unroll
        for (int q = 0; q < kPF; q++)
            ring[q] = (uact && n - 1 - q >= 0) ? ab[(kv - lane) + (long)ldab * (n - 1 - q)] : 0.;
        for (int jb = 0; jb < n; jb += kPF) {
#pragma unroll
            for (int q = 0; q < kPF; q++) {
                const int j = n - 1 - (jb + q);
                if (j < 0) break;
                const double uc0 = ring[q];
                if (uact && j - kPF >= 0) ring[q] = ab[(kv - lane) + (long)ldab * (j - kPF)];
                const int lu = (kv < j) ? kv : j;
                const double uc = (lane >= 1 && lane <= lu) ? uc0 : 0.;
                const int nj = (j - kv - 1 >= 0) ? j - kv - 1 : 0;
                const double nb0 = b0[nj], nb1 = b1[nj];
                const double diag = rdl(uc0, 0);
                const double bj0 = rdl(v0, 0), bj1 = rdl(v1, 0);
                const double x0 = (bj0 != 0.) ? bj0 / diag : bj0;
                const double x1 = (bj1 != 0.) ? bj1 / diag : bj1;
                v0 = v0 - x0 * uc;
                v1 = v1 - x1 * uc;
                if (lane == 0) {
                    b0[j] = x0;
                    b1[j] = x1;
                }
                v0 = __shfl_down(v0, 1);
                v1 = __shfl_down(v1, 1);
                if (lane == kv) {
                    v0 = (j - kv - 1 >= 0) ? nb0 : 0.;
                    v1 = (j - kv - 1 >= 0) ? nb1 : 0.;
                }
            }
        }
    }
    __syncthreads();
}

// Legendre filter (filter_legenpen, filtering.f:729-808) of two columns of
// nz values in LDS, subdomain by subdomain (lanes 0-31 real, 32-63 imaginary)
__device__ void filter_pair(const ModeArgs &A, double *c0, double *c1, double *tmp)
{
    const int lane = threadIdx.x & 63, nl = A.nl, nzm = nl - 1;
    const int comp = lane >> 5, k = lane & 31;
    double *col = comp ? c1 : c0;
    for (int ks = 0; ks < A.ns; ks++) {
        if (k < nl) {
            const double *f = col + ks * nl;
            double sum = 0.;
            for (int t = 0; t < nl; t++) sum = sum + A.wg[t] * f[t] * A.AL[k * nl + t];
            double a = (k == nzm) ? 0.5 * nzm * sum : (k + 0.5) * sum;
            tmp[comp * 32 + k] = a * A.sig[k];
        }
        __syncthreads();
        if (k < nl) {
            double sum = 0.;
            for (int kk = 0; kk < nl; kk++) sum = sum + tmp[comp * 32 + kk] * A.AL[kk * nl + k];
            col[ks * nl + k] = sum;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void k_pressure(ModeArgs A)
{
    extern __shared__ double shm[];
    const int nz = A.nz, nl = A.nl;
    double *rr = shm, *ri = rr + nz, *tmp = ri + nz;
    const int task = A.order[blockIdx.x];
    if (task < 0) return;
    const int m = task, lane = threadIdx.x;
    const int ii = m % A.nxhp, j = m / A.nxhp;
    const long o = 2L * m;
    const bool zero = (m == 0);
    const double xw = A.xw[ii], yw = A.yw[j];
    if (!zero) {
        const int sP = A.modesys[4 * m];
        const double eps = A.sysinfo[4 * sP];
        for (int k = lane; k < nz; k += 64) { // divergence, calcuvwt.f:112-136
            const int ks = k / nl, k1 = k - ks * nl;
            const double fac = 2. / A.zh[ks];
            double sr = 0., si = 0.;
            for (int t = 0; t < nl; t++) {
                const double dd = A.D[k1 * nl + t];
                const long nw = (long)(ks * nl + t) * A.nxppy + o;
                sr = sr + dd * A.w[nw];
                si = si + dd * A.w[nw + 1];
            }
            const long n = (long)k * A.nxppy + o;
            double dr = (0. + (-xw * A.u[n + 1])) + (-yw * A.v[n + 1]);
            double di = (0. + xw * A.u[n]) + yw * A.v[n];
            dr = dr + sr * fac;
            di = di + si * fac;
            rr[k] = -(dr / A.dt) * eps; // penalty.f:191-194
            ri[k] = -(di / A.dt) * eps;
        }
        __syncthreads();
        if (lane == 0) { // Neumann data (calcuvwt.f:154-159, penalty.f:583-598)
            const double t1 = A.sysinfo[4 * sP + 1], t2 = A.sysinfo[4 * sP + 2];
            const long nb = A.nxppy;
            const double bcrtop = A.wnl[nb + o] + A.wviscrot[nb + o];
            const double bcrbot = A.wnl[o] + A.wviscrot[o];
            const double bcitop = A.wnl[nb + o + 1] + A.wviscrot[nb + o + 1];
            const double bcibot = A.wnl[o + 1] + A.wviscrot[o + 1];
            rr[0] = rr[0] - t1 * bcrbot * eps;
            ri[0] = ri[0] - t1 * bcibot * eps;
            rr[nz - 1] = rr[nz - 1] + t2 * bcrtop * eps;
            ri[nz - 1] = ri[nz - 1] + t2 * bcitop * eps;
        }
        __syncthreads();
        band_solve2(A, sP, rr, ri);
        if (A.cpen) filter_pair(A, rr, ri, tmp);
    }
    // projection (update, calcuvwt.f:826-901), in place
    const double cx = A.dt * xw, cy = A.dt * yw;
    const int k0 = A.noslip ? 1 : 0;
    for (int k = lane; k < nz; k += 64) {
        const long n = (long)k * A.nxppy + o;
        if (k < k0) { // rows the no-slip variant never assigns (calcuvwt_noslip.f:851)
            A.u[n] = 0.; A.u[n + 1] = 0.;
            A.v[n] = 0.; A.v[n + 1] = 0.;
            A.w[n] = 0.; A.w[n + 1] = 0.;
            continue;
        }
        double tr = 0., ti = 0., dzr = 0., dzi = 0.;
        if (!zero) {
            tr = rr[k];
            ti = ri[k];
            const int ks = k / nl, k1 = k - ks * nl;
            const double fac = 2. / A.zh[ks];
            double sr = 0., si = 0.;
            for (int t = 0; t < nl; t++) {
                const double dd = A.D[k1 * nl + t];
                sr = sr + dd * rr[ks * nl + t];
                si = si + dd * ri[ks * nl + t];
            }
            dzr = sr * fac;
            dzi = si * fac;
        }
        const double ur = A.u[n], ui = A.u[n + 1], vr = A.v[n], vi = A.v[n + 1];
        A.u[n] = ur - (0. * tr - cx * ti);
        A.u[n + 1] = ui - (0. * ti + cx * tr);
        A.v[n] = vr - (0. * tr - cy * ti);
        A.v[n + 1] = vi - (0. * ti + cy * tr);
        if (zero) {
            A.w[n] = 0.;
            A.w[n + 1] = 0.;
        } else {
            A.w[n] = A.w[n] - A.dt * dzr;
            A.w[n + 1] = A.w[n + 1] - A.dt * dzi;
        }
    }
}

__global__ __launch_bounds__(64) void k_viscous(ModeArgs A)
{
    extern __shared__ double shm[];
    const int nz = A.nz;
    double *rr = shm, *ri = rr + nz, *tmp = ri + nz;
    const int task = A.order[blockIdx.x];
    if (task < 0) return;
    const int m = task >> 2, q = task & 3;
    const int lane = threadIdx.x;
    const long o = 2L * m;
    const bool zero = (m == 0);
    if (q == 2 && zero) return; // w(0,0) = 0 already (calcuvwt.f:496-503)
    double *f = q == 0 ? A.u : q == 1 ? A.v : q == 2 ? A.w : A.t;
    const int sys = A.modesys[4 * m + (q == 0 ? 1 : q == 1 ? 1 : q == 2 ? 2 : 3)];
    const double eps = A.sysinfo[4 * sys];
    const double den = (q < 3) ? A.xnu * A.dt : A.dt * A.xkappa;
    for (int k = lane; k < nz; k += 64) { // rhs = -f/(nu dt), scaled by -eps (penalty.f:191-194)
        const long n = (long)k * A.nxppy + o;
        rr[k] = -(-f[n] / den) * eps;
        ri[k] = -(-f[n + 1] / den) * eps;
    }
    __syncthreads();
    band_solve2(A, sys, rr, ri);
    if (A.cpen) filter_pair(A, rr, ri, tmp);
    for (int k = lane; k < nz; k += 64) {
        const long n = (long)k * A.nxppy + o;
        double a = rr[k], b = ri[k];
        if (q == 0 && zero) { // mean-u damping, calcuvwt.f:343-361
            a = A.damp[k] * a;
            b = A.damp[k] * b;
        }
        f[n] = a;
        f[n + 1] = b;
    }
}

static ModeArgs mode_args(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt,
                          const double *wnl, const double *wviscrot, double dt, int cpen)
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
    return A;
}

static size_t mode_lds(const Geo &g)
{
    return sizeof(double) * (2 * (size_t)g.nz + 64);
}

void launch_pressure(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt,
                     const double *wnl, const double *wviscrot, double dt, int cpen, hipStream_t st)
{
    ModeArgs A = mode_args(g, P, S, uvwt, wnl, wviscrot, dt, cpen);
    A.order = S.order_p;
    hipLaunchKernelGGL(k_pressure, dim3(S.nblk_p), dim3(64), mode_lds(g), st, A);
}

void launch_viscous(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                    int cpen, hipStream_t st)
{
    ModeArgs A = mode_args(g, P, S, uvwt, nullptr, nullptr, dt, cpen);
    A.order = S.order_v;
    hipLaunchKernelGGL(k_viscous, dim3(S.nblk_v), dim3(64), mode_lds(g), st, A);
}

// ---------------------------------------------------------------------------
// Band sweeps over the packed factors (tri_* layout), grouped lanes.  Lane =
// one right-hand-side column on the native layout; the G lanes of a group
// share one system (host layout: the +ky and -ky rows of one kx, re and im,
// and for G = 8 both u and v).  Each lane fetches 1/G of a factor segment
// DP rows ahead into a register ring; one row before use the pieces are put
// together in LDS and every lane of the group reads the whole segment from
// there.  The rows of one subdomain are unrolled, so every segment length,
// ring slot and window update is fixed at compile time and only the
// structurally nonzero entries are loaded and applied (an entry outside the
// pattern is an exact zero, and w + 0 * y is w).
//   forward (bandludcmp.f:1020-1030; SGBTRS with interchanges for the
//     pressure): window w[t] = row j + t; y = w[0] (after the swap with
//     w[ip[j] - j]), w[t-1] = w[t] - L[t] y;
//   backward (STBSV, ludcmp.f:4698-5046): window w[t] = row j - t;
//     x = w[0] / U[0], w[t-1] = w[t] - x U[t].
// Task entry: col | field << 28 | idle << 29 (an idle lane helps load its
// group's factors but owns no column); -1 marks an empty group.
// ---------------------------------------------------------------------------
struct SweepArgs {
    double *f[4];
    const int *tasks;
    const int *modesys;
    const double *PF; // FB, or PB for the pressure
    const int *ipiv;
    const double *sysinfo;
    double den[4];
    long nxppy;
    int n, fld_base;
};

template <int NL, int G, int DP, bool PIV>
__global__ __launch_bounds__(64) void k_sweep(SweepArgs A)
{
    static_assert(NL % DP == 0, "ring slots follow the local row");
    static_assert(DP >= 2, "a row is staged one step before use, from a slot filled at least one step earlier");
    constexpr int NG = 64 / G;
    constexpr int SF = tri_sf(NL), SB = tri_sb(NL, PIV);
    constexpr int NPFM = tri_npf(NL, 0), NPBM = tri_npb(NL, NL - 1, PIV);
    constexpr int NPM = NPFM > NPBM ? NPFM : NPBM;
    constexpr int PPF = (NPFM + G - 1) / G, PPB = (NPBM + G - 1) / G; // pieces per lane
    constexpr int EBM = tri_eb(NL, NL - 1, PIV);                     // widest U column
    __shared__ double2 seg[2][NG][NPM + 1];
    const int lane = threadIdx.x, grp = lane / G, rk = lane % G;
    const int e = A.tasks[blockIdx.x * 64 + lane];
    if (e < 0) return; // the whole group is empty
    const bool idle = (e >> 29) & 1;
    const int col = e & ((1 << 28) - 1);
    const int m = col >> 1;
    int fld, sys;
    if constexpr (PIV) {
        fld = 0;
        if (m == 0) return; // no pressure for kx = ky = 0; its group holds only this mode
        sys = A.modesys[4 * m];
    } else {
        fld = A.fld_base + blockIdx.y + ((e >> 28) & 1);
        if (fld == 2 && m == 0) return; // w(0,0) = 0 (calcuvwt.f:496-503); its group holds only this mode
        sys = A.modesys[4 * m + (fld < 2 ? 1 : fld == 2 ? 2 : 3)];
    }
    const double eps = PIV ? 1. : A.sysinfo[4 * sys], den = PIV ? 1. : A.den[fld];
    // right-hand side of the viscous systems: -(-f / den) eps (calcuvwt.f:390-402)
    auto tr = [&](double v) __attribute__((always_inline)) { return PIV ? v : -(-v / den) * eps; };
    const int n = A.n, nsub = n / NL;
    const long st = A.nxppy;
    const double2 *__restrict__ fw = reinterpret_cast<const double2 *>(A.PF) + (long)sys * nsub * (SF + SB);
    const double2 *__restrict__ bw = fw + (long)nsub * SF;
    const int *__restrict__ ip = A.ipiv + (PIV ? (long)sys * n : 0);
    double *__restrict__ b = A.f[fld] + col;
    // ---- forward
    {
        double w[NL + 1];
        double rx[DP][PPF], ry[DP][PPF];
        double bring[DP];
        int pring[DP];
        auto ldp = [&](auto Rc, int ks, double(&dx)[PPF], double(&dy)[PPF]) __attribute__((always_inline)) {
            constexpr int R = decltype(Rc)::value, NP = tri_npf(NL, R);
            const double2 *src = fw + (long)(ks < nsub ? ks : nsub - 1) * SF + tri_off_f(NL, R);
#pragma unroll
            for (int q = 0; q < PPF; q++) {
                const int p = rk + q * G;
                if (q * G < NP) {
                    const double2 v = src[p < NP ? p : 0];
                    dx[q] = v.x;
                    dy[q] = v.y;
                }
            }
        };
        auto stp = [&](auto Rc, int buf, const double(&sx)[PPF], const double(&sy)[PPF]) __attribute__((always_inline)) {
            constexpr int R = decltype(Rc)::value, NP = tri_npf(NL, R);
#pragma unroll
            for (int q = 0; q < PPF; q++) {
                const int p = rk + q * G;
                if (q * G < NP && p < NP) seg[buf][grp][p] = make_double2(sx[q], sy[q]);
            }
        };
#pragma unroll
        for (int t = 0; t <= NL; t++) {
            const double v = b[(long)(t < n ? t : n - 1) * st];
            w[t] = (t < n) ? tr(v) : 0.;
        }
        static_for<0, DP>([&](auto dc) __attribute__((always_inline)) {
            constexpr int d = decltype(dc)::value;
            ldp(IC<d>{}, 0, rx[d], ry[d]);
            const int r = NL + 1 + d;
            bring[d] = b[(long)(r < n ? r : n - 1) * st];
            if constexpr (PIV) pring[d] = ip[d < n ? d : n - 1];
        });
        stp(IC<0>{}, 0, rx[0], ry[0]);
        for (int ks = 0; ks < nsub; ks++) {
            static_for<0, NL>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value, d = r % DP;
                constexpr int EF = tri_ef(NL, r), NP = tri_npf(NL, r);
                const int j = ks * NL + r;
                if (j >= n) return; // never: one block per row step
                stp(IC<(r + 1) % NL>{}, (j + 1) & 1, rx[(d + 1) % DP], ry[(d + 1) % DP]);
                ldp(IC<(r + DP) % NL>{}, ks + (r + DP) / NL, rx[d], ry[d]);
                double fc[2 * NP]; // fc[t - 1] = L[t]
                // the other lanes' pieces were stored one step earlier by this
                // wavefront: order those stores before these reads
                wave_sync();
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const double2 v = seg[j & 1][grp][p];
                    fc[2 * p] = v.x;
                    fc[2 * p + 1] = v.y;
                }
                const double rc_ = bring[d];
                {
                    const int rr = j + NL + 1 + DP;
                    bring[d] = b[(long)(rr < n ? rr : n - 1) * st];
                }
                double y = w[0];
                if constexpr (PIV) {
                    const int dd = pring[d] - j; // 0 <= dd <= EF (checked when packed)
                    {
                        const int jj = j + DP;
                        pring[d] = ip[jj < n ? jj : n - 1];
                    }
#pragma unroll
                    for (int t = 1; t <= EF; t++) y = (dd == t) ? w[t] : y;
                    const double w0 = w[0];
#pragma unroll
                    for (int t = 1; t <= NL; t++) {
                        if (t <= EF) {
                            const double wt = (dd == t) ? w0 : w[t];
                            w[t - 1] = wt + fc[t - 1] * (-y);
                            pin(w[t - 1]);
                        } else {
                            w[t - 1] = w[t];
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 1; t <= NL; t++) {
                        w[t - 1] = (t <= EF) ? w[t] + fc[t - 1] * (-y) : w[t];
                        if (t <= EF) pin(w[t - 1]);
                    }
                }
                w[NL] = (j + NL + 1 < n) ? tr(rc_) : 0.;
                if (!idle) b[(long)j * st] = y;
            });
        }
    }
    // ---- backward
    {
        double w[EBM + 1];
        double rx[DP][PPB], ry[DP][PPB];
        double bring[DP];
        auto ldp = [&](auto Rc, int ks, double(&dx)[PPB], double(&dy)[PPB]) __attribute__((always_inline)) {
            constexpr int R = decltype(Rc)::value, NP = tri_npb(NL, R, PIV);
            const double2 *src = bw + (long)(ks >= 0 ? ks : 0) * SB + tri_off_b(NL, R, PIV);
#pragma unroll
            for (int q = 0; q < PPB; q++) {
                const int p = rk + q * G;
                if (q * G < NP) {
                    const double2 v = src[p < NP ? p : 0];
                    dx[q] = v.x;
                    dy[q] = v.y;
                }
            }
        };
        auto stp = [&](auto Rc, int buf, const double(&sx)[PPB], const double(&sy)[PPB]) __attribute__((always_inline)) {
            constexpr int R = decltype(Rc)::value, NP = tri_npb(NL, R, PIV);
#pragma unroll
            for (int q = 0; q < PPB; q++) {
                const int p = rk + q * G;
                if (q * G < NP && p < NP) seg[buf][grp][p] = make_double2(sx[q], sy[q]);
            }
        };
#pragma unroll
        for (int t = 0; t <= EBM; t++) {
            const int r = n - 1 - t;
            const double v = b[(long)(r >= 0 ? r : 0) * st];
            w[t] = (r >= 0) ? v : 0.;
        }
        static_for<0, DP>([&](auto dc) __attribute__((always_inline)) {
            constexpr int d = decltype(dc)::value;
            ldp(IC<NL - 1 - d>{}, nsub - 1, rx[d], ry[d]);
            const int r = n - 2 - EBM - d;
            bring[d] = b[(long)(r >= 0 ? r : 0) * st];
        });
        stp(IC<NL - 1>{}, 0, rx[0], ry[0]);
        for (int ks = nsub - 1; ks >= 0; ks--) {
            static_for<0, NL>([&](auto qc) __attribute__((always_inline)) {
                constexpr int q = decltype(qc)::value, r = NL - 1 - q, d = q % DP;
                constexpr int EB = tri_eb(NL, r, PIV), NP = tri_npb(NL, r, PIV);
                const int j = ks * NL + r, sidx = n - 1 - j;
                if (j < 0) return; // always false; one block per row step
                stp(IC<NL - 1 - (q + 1) % NL>{}, (sidx + 1) & 1, rx[(d + 1) % DP], ry[(d + 1) % DP]);
                ldp(IC<NL - 1 - (q + DP) % NL>{}, ks - (q + DP) / NL, rx[d], ry[d]);
                double gc[2 * NP]; // gc[t] = U[t], t = 0 the diagonal
                wave_sync();
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const double2 v = seg[sidx & 1][grp][p];
                    gc[2 * p] = v.x;
                    gc[2 * p + 1] = v.y;
                }
                const double rc_ = bring[d];
                {
                    const int rr = j - EBM - 1 - DP;
                    bring[d] = b[(long)(rr >= 0 ? rr : 0) * st];
                }
                const double x = (w[0] != 0.) ? w[0] / gc[0] : w[0];
#pragma unroll
                for (int t = 1; t <= EBM; t++) {
                    w[t - 1] = (t <= EB) ? w[t] - x * gc[t] : w[t];
                    if (t <= EB) pin(w[t - 1]);
                }
                w[EBM] = (j - EBM - 1 >= 0) ? rc_ : 0.;
                if (!idle) b[(long)j * st] = x;
            });
        }
    }
}

// ring depth (rows fetched ahead) per subdomain order: a divisor of nl, as
// deep as the registers allow
template <int NL, bool PIV>
constexpr int sweep_dp()
{
    // nl = 25: 5 rows ahead measured faster than 25 (which halves occupancy);
    // 9, 13, 17: the whole subdomain; otherwise the largest divisor of nl in
    // 4..11, or nl itself when there is none
    if (NL == 25) return 5;
    if (NL == 21) return 7;
    if (NL == 9 || NL == 13 || NL == 17) return NL;
    for (int d = 11; d >= 4; d--)
        if (NL % d == 0) return d;
    return NL;
}

void launch_solve_grp(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                      hipStream_t st)
{
    SweepArgs A;
    for (int q = 0; q < 4; q++) A.f[q] = uvwt[q];
    A.modesys = S.modesys;
    A.PF = S.FB;
    A.ipiv = S.ipiv;
    A.sysinfo = S.sysinfo;
    A.den[0] = A.den[1] = A.den[2] = P.xnu * dt;
    A.den[3] = dt * P.xkappa;
    A.nxppy = g.nxppy;
    A.n = g.nz;
    SweepArgs B = A;
    A.tasks = S.guv_tasks;
    A.fld_base = 0;
    B.tasks = S.gwt_tasks;
    B.fld_base = 2;
    dim3 ga(S.guv_ntasks / 64, 1), gb(S.gwt_ntasks / 64, 2);
#define GS(NL)                                                                                                  \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<NL, 8, sweep_dp<NL, false>(), false>), ga, dim3(64), 0, st, A); \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<NL, 4, sweep_dp<NL, false>(), false>), gb, dim3(64), 0, st, B)
    switch (g.nzloc) {
        IW_NL_SWITCH(GS)
    default: break;
    }
#undef GS
}

// ---------------------------------------------------------------------------
// Pressure step on the native layout (calcuvwt.f:87-200 with update,
// calcuvwt.f:826-901), three kernels:
//   k_prhs:     divergence (divcolc, derv.f:35-85) of u*, scaled right-hand
//               side -(div/dt) eps and the Neumann data wnl + wviscrot
//               (calcuvwt.f:112-159, penalty.f:191-194, 583-598) -> phi;
//   k_sweep<.., true>: SGBTRS with row interchanges (bandludcmp.f:1020-1036),
//               lane = one column, groups of lanes sharing a system;
//   k_pproject: Legendre filter of phi, dphi/dz and u* - dt grad(phi).
// ---------------------------------------------------------------------------
struct PArgs {
    const double *u, *v, *w;
    double *phi;
    const double *wnl, *wviscrot;
    const int *modesys;
    const double *sysinfo;
    const double *D, *zh, *xw, *yw;
    double dt;
    long nxppy;
    int nxhp, nz;
};

template <int NL>
__global__ __launch_bounds__(256) void k_prhs(PArgs A)
{
    __shared__ double Ds[NL * NL];
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) Ds[e] = A.D[e];
    __syncthreads();
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ks = blockIdx.y;
    if (col >= A.nxppy) return;
    const int m = (int)(col >> 1), comp = (int)(col & 1);
    const long st = A.nxppy, base = (long)ks * NL * st;
    double *phi = A.phi + base + col;
    if (m == 0) { // no pressure for kx = ky = 0 (calcuvwt.f:105-109)
#pragma unroll
        for (int k = 0; k < NL; k++) phi[(long)k * st] = 0.;
        return;
    }
    const int sP = A.modesys[4 * m];
    const double eps = A.sysinfo[4 * sP];
    const double xw = A.xw[m % A.nxhp], yw = A.yw[m / A.nxhp];
    const double fac = 2. / A.zh[ks];
    double wc[NL];
#pragma unroll
    for (int t = 0; t < NL; t++) wc[t] = A.w[base + (long)t * st + col];
    const long pc = col ^ 1;
    // every load before the arithmetic (no load behind a branch), so they are
    // in flight together
    const double pw0 = A.sysinfo[4 * sP + 1] * (A.wnl[col] + A.wviscrot[col]) * eps;
    const double pw1 = A.sysinfo[4 * sP + 2] * (A.wnl[st + col] + A.wviscrot[st + col]) * eps;
    double uu[NL], vv[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const long o = base + (long)k * st + pc;
        uu[k] = A.u[o];
        vv[k] = A.v[o];
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double sd = 0.;
#pragma unroll
        for (int t = 0; t < NL; t++) sd = sd + Ds[k * NL + t] * wc[t];
        double d = (comp == 0) ? (0. + (-xw * uu[k])) + (-yw * vv[k]) : (0. + xw * uu[k]) + yw * vv[k];
        d = d + sd * fac;
        double r = -(d / A.dt) * eps;
        const int kg = ks * NL + k;
        if (kg == 0) r = r - pw0;
        if (kg == A.nz - 1) r = r + pw1;
        phi[(long)k * st] = r;
    }
}

template <int NL>
__global__ __launch_bounds__(256) void k_pproject(PArgs A, double *u, double *v, double *w, int cpen,
                                                 int noslip, const double *__restrict__ FL)
{
    __shared__ double FLs[NL * NL], Ds[NL * NL];
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) Ds[e] = A.D[e];
    stage_fl<NL>(FLs, FL);
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ks = blockIdx.y;
    const bool live = col < A.nxppy;
    const long c = live ? col : 0;
    const int m = (int)(c >> 1), comp = (int)(c & 1);
    const long st = A.nxppy, base = (long)ks * NL * st;
    double x[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) x[k] = (m == 0) ? 0. : A.phi[base + (long)k * st + c];
    if (cpen && m != 0) legfilter_mat<NL>(x, FLs); // filter_legenpen of phi
    const double xw = A.xw[m % A.nxhp], yw = A.yw[m / A.nxhp];
    const double cx = A.dt * xw, cy = A.dt * yw;
    const double fac = 2. / A.zh[ks];
    // a few levels at a time: their u, v, w loads first (no load behind a
    // branch, so they are in flight together), then the updates
    constexpr int KB = 5;
#pragma unroll
    for (int k0 = 0; k0 < NL; k0 += KB) {
        double uo[KB], vo[KB], wo[KB];
#pragma unroll
        for (int kk = 0; kk < KB; kk++) {
            const int k = k0 + kk;
            if (k < NL) {
                const long o = base + (long)k * st + c;
                uo[kk] = u[o];
                vo[kk] = v[o];
                wo[kk] = w[o];
            }
        }
#pragma unroll
        for (int kk = 0; kk < KB; kk++) {
            const int k = k0 + kk;
            if (k >= NL) break;
            const double other = __shfl_xor(x[k], 1); // phi of the partner component
            double sd = 0.;
#pragma unroll
            for (int t = 0; t < NL; t++) sd = sd + Ds[k * NL + t] * x[t];
            const double tr = comp == 0 ? x[k] : other, ti = comp == 0 ? other : x[k];
            double nu, nv;
            if (comp == 0) {
                nu = uo[kk] - (0. * tr - cx * ti);
                nv = vo[kk] - (0. * tr - cy * ti);
            } else {
                nu = uo[kk] - (0. * ti + cx * tr);
                nv = vo[kk] - (0. * ti + cy * tr);
            }
            double nw = (m == 0) ? 0. : wo[kk] - A.dt * (sd * fac);
            if (noslip && ks * NL + k == 0) { // rows the no-slip variant never assigns (calcuvwt_noslip.f:851)
                nu = 0.;
                nv = 0.;
                nw = 0.;
            }
            if (live) {
                const long o = base + (long)k * st + c;
                u[o] = nu;
                v[o] = nv;
                w[o] = nw;
            }
        }
    }
}

void launch_pressure_nat(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double *phi,
                         const double *wnl, const double *wviscrot, double dt, int cpen, int part,
                         hipStream_t st)
{
    PArgs A;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2];
    A.phi = phi;
    A.wnl = wnl; A.wviscrot = wviscrot;
    A.modesys = S.modesys; A.sysinfo = S.sysinfo;
    A.D = g.D; A.zh = g.zh; A.xw = g.xw; A.yw = g.yw;
    A.dt = dt;
    A.nxppy = g.nxppy;
    A.nxhp = g.nxhp;
    A.nz = g.nz;
    dim3 gc((unsigned)((g.nxppy + 255) / 256), g.nsubd);
    SweepArgs C;
    C.f[0] = C.f[1] = C.f[2] = C.f[3] = phi;
    C.tasks = S.gwt_tasks; C.modesys = S.modesys; C.PF = S.PB; C.ipiv = S.ipiv; C.sysinfo = S.sysinfo;
    C.den[0] = C.den[1] = C.den[2] = C.den[3] = 1.;
    C.nxppy = g.nxppy; C.n = g.nz; C.fld_base = 0;
    dim3 gg(S.gwt_ntasks / 64);
    const int ns = P.bc_uv_bottom == 1;
#define PSTEP(NL)                                                                                         \
    if (part == 0) hipLaunchKernelGGL(k_prhs<NL>, gc, dim3(256), 0, st, A);                               \
    if (part == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<NL, 4, sweep_dp<NL, true>(), true>), gg, dim3(64), 0, st, C);  \
    if (part == 2)                                                                                        \
        hipLaunchKernelGGL(k_pproject<NL>, gc, dim3(256), 0, st, A, uvwt[0], uvwt[1], uvwt[2], cpen, ns, g.FL)
    switch (g.nzloc) {
        IW_NL_SWITCH(PSTEP)
    default: break;
    }
#undef PSTEP
}

// bdf_filter and the pressure right-hand side in one pass.  A work-group is
// 64 columns (re and im of 32 modes) of one subdomain; wavefront f takes
// field f (u, v, w, rho) through the SS/BDF assembly and the filters of
// k_tile_t<NL,1,S32> and writes it.  The filtered u, v, w also go to LDS,
// and after a barrier the four wavefronts form the pressure right-hand side
// of k_prhs from them (the same sums in the same order), so the pressure
// pass reads no u, v, w back.
template <int NL, bool S32>
__global__ __launch_bounds__(256) void k_tile_prhs(TileArgs A, PArgs P, long nxppy, const double *__restrict__ fcf,
                                                   const double *__restrict__ FL)
{
    __shared__ double fv[3][NL][64]; // filtered u, v, w of the work-group's columns
    // the field is the same in all lanes of a wavefront, so the field's
    // pointers and the filter and D entries (their indices fixed at compile
    // time) come through scalar loads
    const int lane = threadIdx.x & 63, fz = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *__restrict__ FLs = FL;
    const long col = (long)blockIdx.x * 64 + lane;
    const int ks = blockIdx.y;
    const bool live = col < nxppy; // nxppy is even: both lanes of a mode are live or not
    const bool im = col & 1;
    const long b0 = (long)ks * NL * nxppy + (live ? col : 0);
    if (live) {
        const double *const *Ff = A.F[fz];
        double *outf = A.out[fz];
        const double sigf = fcf[col >> 1];
        double x[NL];
#pragma unroll
        for (int t = 0; t < NL; t++) {
            const long o = b0 + (long)t * nxppy;
            double s0, s1, s2;
            if (S32) {
                s0 = (double)A.S32[fz][0][o];
                s1 = (double)A.S32[fz][1][o];
                s2 = (double)A.S32[fz][2][o];
            } else {
                s0 = A.S[fz][0][o];
                s1 = A.S[fz][1][o];
                s2 = A.S[fz][2][o];
            }
            const double f0 = Ff[0][o], f1 = Ff[1][o], f2 = Ff[2][o];
            const double sr = A.f1 * s0 + A.f2 * s1 + A.f3 * s2;
            x[t] = A.dt * sr + A.b1 * f0 + A.b2 * f1 + A.b3 * f2;
        }
        if (A.ifourier) {
#pragma unroll
            for (int t = 0; t < NL; t++) {
                const double other = __shfl_xor(x[t], 1, 64);
                double fr = im ? other : x[t], fi = im ? x[t] : other;
                fourier_filter_value(fr, fi, sigf);
                x[t] = im ? fi : fr;
            }
        }
#pragma unroll
        for (int q = 0; q < NL; q++) {
            double o = x[q];
            if (A.ilegen) {
                o = 0.;
#pragma unroll
                for (int t = 0; t < NL; t++) o = o + FLs[t * NL + q] * x[t];
            }
            outf[b0 + (long)q * nxppy] = o;
            if (fz < 3) fv[fz][q][lane] = o;
        }
    }
    __syncthreads();
    if (!live) return;
    // k_prhs for this column, rows fz, fz + 4, ... (kx = ky = 0: phi = 0)
    const int m = (int)(col >> 1), pl = lane ^ 1;
    double eps = 0., pw0 = 0., pw1 = 0., xw = 0., yw = 0.;
    if (m != 0) {
        const int sP = P.modesys[4 * m];
        eps = P.sysinfo[4 * sP];
        xw = P.xw[m % P.nxhp];
        yw = P.yw[m / P.nxhp];
        pw0 = P.sysinfo[4 * sP + 1] * (P.wnl[col] + P.wviscrot[col]) * eps;
        pw1 = P.sysinfo[4 * sP + 2] * (P.wnl[nxppy + col] + P.wviscrot[nxppy + col]) * eps;
    }
    const double fac = 2. / P.zh[ks];
    for (int k = fz; k < NL; k += 4) {
        double sd = 0.;
#pragma unroll
        for (int t = 0; t < NL; t++) sd = sd + P.D[k * NL + t] * fv[2][t][lane];
        const double uu = fv[0][k][pl], vv = fv[1][k][pl];
        double d = im ? (0. + xw * uu) + yw * vv : (0. + (-xw * uu)) + (-yw * vv);
        d = d + sd * fac;
        double r = -(d / P.dt) * eps;
        const int kg = ks * NL + k;
        if (kg == 0) r = r - pw0;
        if (kg == P.nz - 1) r = r + pw1;
        P.phi[b0 + (long)k * nxppy] = (m == 0) ? 0. : r;
    }
}

template <typename TS>
static bool bdf_prhs_launch(const Geo &g, TS *const *S12, double *const *F12, double *const *out4, const double *fac,
                            const double *bdf, double dt, int filt, const SysTables &S, double *phi,
                            const double *wnl, const double *wviscrot, hipStream_t st)
{
    if (!nl_specialised(g.nzloc) || g.ndims != 3) return false;
    TileArgs A = {};
    for (int f = 0; f < 4; f++) {
        for (int q = 0; q < 3; q++) {
            if constexpr (std::is_same<TS, float>::value)
                A.S32[f][q] = S12[3 * f + q];
            else
                A.S[f][q] = S12[3 * f + q];
            A.F[f][q] = F12[3 * f + q];
        }
        A.out[f] = out4[f];
    }
    A.f1 = fac[0]; A.f2 = fac[1]; A.f3 = fac[2];
    A.b1 = bdf[0]; A.b2 = bdf[1]; A.b3 = bdf[2];
    A.dt = dt;
    A.mode = 1;
    A.ifourier = filt;
    A.ilegen = filt;
    PArgs P;
    P.u = P.v = P.w = nullptr;
    P.phi = phi;
    P.wnl = wnl; P.wviscrot = wviscrot;
    P.modesys = S.modesys; P.sysinfo = S.sysinfo;
    P.D = g.D; P.zh = g.zh; P.xw = g.xw; P.yw = g.yw;
    P.dt = dt;
    P.nxppy = g.nxppy;
    P.nxhp = g.nxhp;
    P.nz = g.nz;
    constexpr bool F32 = std::is_same<TS, float>::value;
    dim3 gt((unsigned)((g.nxppy + 63) / 64), g.nsubd);
#define TP(NL) \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile_prhs<NL, F32>), gt, dim3(256), 0, st, A, P, g.nxppy, g.fcf, g.FL)
    switch (g.nzloc) {
        IW_NL_SWITCH(TP)
    default: return false;
    }
#undef TP
    return true;
}

bool launch_bdf_filter_prhs(const Geo &g, double *const *S12, double *const *F12, double *const *out4,
                            const double *fac, const double *bdf, double dt, int filt, const SysTables &S,
                            double *phi, const double *wnl, const double *wviscrot, hipStream_t st)
{
    return bdf_prhs_launch<double>(g, S12, F12, out4, fac, bdf, dt, filt, S, phi, wnl, wviscrot, st);
}

bool launch_bdf_filter_prhs(const Geo &g, float *const *S12, double *const *F12, double *const *out4,
                            const double *fac, const double *bdf, double dt, int filt, const SysTables &S,
                            double *phi, const double *wnl, const double *wviscrot, hipStream_t st)
{
    return bdf_prhs_launch<float>(g, S12, F12, out4, fac, bdf, dt, filt, S, phi, wnl, wviscrot, st);
}

bool solve_np_supported(int nl, int nsubd)
{
    (void)nsubd;
    return nl_specialised(nl);
}

// ---------------------------------------------------------------------------
// Estimation-model filter (filtering.f:8-270, 989-1105) and SGS tensors
// (sgscalc.f:4-223), physical space.
//   k_sgs_xy:  u0 = 3 f(u) - 3 f(f(u)) + f(f(f(u))), f = periodic 3-point
//              top-hat in y after x (modecut, filtering.f:100-156;
//              filterx3pv/filtery3pv, 162-270), one LDS tile of TX x TY
//              points plus a halo of 3 per plane;
//   k_sgs_z:   Boyd-Vandeven Legendre filter of every column i < nx of a
//              subdomain (filter_legensgsz, filtering.f:989-1105), written to
//              the target on levels kmin..kmax;
//   k_sgs_pre / k_sgs_post: products and tensors of sgscalc.f.
// ---------------------------------------------------------------------------
constexpr int kSX = 32, kSY = 8, kSH = 3;

__global__ __launch_bounds__(256) void k_sgs_xy(const double *__restrict__ f, double *__restrict__ u0, int nx,
                                               int ny, int nxpp)
{
    constexpr int WX = kSX + 2 * kSH, WY = kSY + 2 * kSH;
    __shared__ double A[WY][WX], B[WY][WX], G1[WY][WX], G2[WY][WX];
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * kSX - kSH, j0 = blockIdx.y * kSY - kSH;
    const double *plane = f + (long)k * nxpp * ny;
    const double a = 1. / 4., b = 2. / 4., c = 1. / 4.;
    for (int e = threadIdx.x; e < WX * WY; e += blockDim.x) {
        const int y = e / WX, x = e - y * WX;
        const int i = ((i0 + x) % nx + nx) % nx, j = ((j0 + y) % ny + ny) % ny;
        A[y][x] = plane[(long)j * nxpp + i];
    }
    __syncthreads();
    // three passes: x filter into B, then y filter into the next level
    double(*src)[WX] = A;
    double(*dst[3])[WX] = {G1, G2, A};
    for (int pass = 0; pass < 3; pass++) {
        const int lo = pass + 1, hix = WX - 1 - pass, hiy = WY - 1 - pass;
        for (int e = threadIdx.x; e < WX * WY; e += blockDim.x) {
            const int y = e / WX, x = e - y * WX;
            if (x >= lo && x < hix) B[y][x] = a * src[y][x - 1] + b * src[y][x] + c * src[y][x + 1];
        }
        __syncthreads();
        double(*d)[WX] = dst[pass];
        // the third level overwrites A, whose values are no longer needed
        for (int e = threadIdx.x; e < WX * WY; e += blockDim.x) {
            const int y = e / WX, x = e - y * WX;
            if (x >= lo && x < hix && y >= lo && y < hiy)
                d[y][x] = a * B[y - 1][x] + b * B[y][x] + c * B[y + 1][x];
        }
        __syncthreads();
        src = d;
    }
    double *out = u0 + (long)k * nxpp * ny;
    for (int e = threadIdx.x; e < kSX * kSY; e += blockDim.x) {
        const int y = e / kSX, x = e - y * kSX;
        const int i = blockIdx.x * kSX + x, j = blockIdx.y * kSY + y;
        if (i < nx && j < ny) {
            const int X = x + kSH, Y = y + kSH;
            out[(long)j * nxpp + i] = 3. * G1[Y][X] - 3. * G2[Y][X] + A[Y][X];
        }
    }
}

template <int NL>
__global__ __launch_bounds__(256) void k_sgs_z(const double *__restrict__ u0, double *__restrict__ out, int nx,
                                              int ny, int nxpp, int kmin, int kmax,
                                              const double *__restrict__ wg, const double *__restrict__ AL,
                                              const double *__restrict__ sig)
{
    __shared__ double ALs[NL * NL], wgs[NL], sgs[NL];
    stage_tables<NL>(wgs, ALs, sgs, wg, AL, sig);
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ks = blockIdx.y;
    const long nxppy = (long)nxpp * ny;
    if (col >= nxppy) return;
    const int i = (int)(col % nxpp);
    if (i >= nx) return;
    const long b0 = (long)ks * NL * nxppy + col;
    double x[NL];
#pragma unroll
    for (int t = 0; t < NL; t++) x[t] = u0[b0 + t * nxppy];
    legfilter<NL>(x, wgs, ALs, sgs);
#pragma unroll
    for (int t = 0; t < NL; t++) {
        const int kg = ks * NL + t;
        if (kg >= kmin && kg <= kmax) out[b0 + t * nxppy] = x[t];
    }
}

// mean of rho per level: its kx = ky = 0 coefficient after the forward transform
__global__ void k_plane_mean(const double *__restrict__ tspec, double *__restrict__ tm, int nz, long nxppy)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nz) tm[k] = tspec[(long)k * nxppy];
}

struct SgsArgs {
    const double *u, *v, *w, *t, *tm;
    double *ten[9]; // t11 t12 t13 t22 t23 t33 th1 th2 th3
    const double *tmp[6];
    double *ox;
    long ntot, nxppy;
    int nxpp, nx, nz;
};

// sgscalc.f:29-56: products on levels k >= 1 (0-based), zero on level 0
__global__ void k_sgs_pre(SgsArgs A)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= A.ntot) return;
    const int i = (int)(o % A.nxpp);
    if (i >= A.nx) return;
    const int k = (int)(o / A.nxppy);
    if (k == 0) {
        for (int q = 0; q < 9; q++) A.ten[q][o] = 0.;
        return;
    }
    const double u = A.u[o], v = A.v[o], w = A.w[o];
    const double ox = A.t[o] - A.tm[k];
    A.ox[o] = ox;
    A.ten[0][o] = u * u;
    A.ten[3][o] = v * v;
    A.ten[5][o] = w * w;
    A.ten[1][o] = u * v;
    A.ten[2][o] = u * w;
    A.ten[4][o] = v * w;
    A.ten[6][o] = u * ox;
    A.ten[7][o] = v * ox;
    A.ten[8][o] = w * ox;
}

// sgscalc.f:84-126 (part 0: stresses from tmp 0..5 = filtered t11 t22 t33
// t12 t13 t23) and 137-165 (part 1: fluxes from tmp 0..2 = filtered th1..3)
__global__ void k_sgs_post(SgsArgs A, int part)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= A.ntot) return;
    const int i = (int)(o % A.nxpp);
    if (i >= A.nx) return;
    const int k = (int)(o / A.nxppy);
    if (k == 0) return;
    const double u = A.u[o], v = A.v[o], w = A.w[o];
    const bool top = (k == A.nz - 1);
    if (part == 0) {
        A.ten[0][o] = A.tmp[0][o] - u * u;
        A.ten[3][o] = A.tmp[1][o] - v * v;
        A.ten[1][o] = A.tmp[3][o] - u * v;
        A.ten[5][o] = top ? 0. : A.tmp[2][o] - w * w;
        A.ten[2][o] = top ? 0. : A.tmp[4][o] - u * w;
        A.ten[4][o] = top ? 0. : A.tmp[5][o] - v * w;
    } else {
        const double ox = A.t[o] - A.tm[k];
        A.ox[o] = ox;
        A.ten[6][o] = top ? 0. : A.tmp[0][o] - u * ox;
        A.ten[7][o] = top ? 0. : A.tmp[1][o] - v * ox;
        A.ten[8][o] = top ? 0. : A.tmp[2][o] - w * ox;
    }
}

void launch_modecut(const Geo &g, const double *f, double *u0, double *out, int kmin, int kmax, hipStream_t st)
{
    dim3 gx((g.nx + kSX - 1) / kSX, (g.ny + kSY - 1) / kSY, g.nz);
    hipLaunchKernelGGL(k_sgs_xy, gx, dim3(256), 0, st, f, u0, g.nx, g.ny, g.nxpp);
    dim3 gz((unsigned)((g.nxppy + 255) / 256), g.nsubd);
#define SZ(NL)                                                                                         \
    hipLaunchKernelGGL(k_sgs_z<NL>, gz, dim3(256), 0, st, u0, out, g.nx, g.ny, g.nxpp, kmin, kmax, g.wg, \
                       g.AL, g.sigbv)
    switch (g.nzloc) {
        IW_NL_SWITCH(SZ)
    default: break;
    }
#undef SZ
}

void launch_plane_mean(const Geo &g, const double *tspec, double *tm, hipStream_t st)
{
    hipLaunchKernelGGL(k_plane_mean, dim3((g.nz + 255) / 256), dim3(256), 0, st, tspec, tm, g.nz, g.nxppy);
}

void launch_sgs_pre(const Geo &g, double *const *uvwt, const double *tm, double *const *ten, double *ox,
                    hipStream_t st)
{
    SgsArgs A = {};
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3]; A.tm = tm;
    for (int q = 0; q < 9; q++) A.ten[q] = ten[q];
    A.ox = ox;
    A.ntot = g.ntot; A.nxppy = g.nxppy; A.nxpp = g.nxpp; A.nx = g.nx; A.nz = g.nz;
    hipLaunchKernelGGL(k_sgs_pre, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, A);
}

void launch_sgs_post(const Geo &g, double *const *uvwt, const double *tm, double *const *ten,
                     double *const *tmp, double *ox, int part, hipStream_t st)
{
    SgsArgs A = {};
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3]; A.tm = tm;
    for (int q = 0; q < 9; q++) A.ten[q] = ten[q];
    for (int q = 0; q < 6; q++) A.tmp[q] = tmp[q];
    A.ox = ox;
    A.ntot = g.ntot; A.nxppy = g.nxppy; A.nxpp = g.nxpp; A.nx = g.nx; A.nz = g.nz;
    hipLaunchKernelGGL(k_sgs_post, dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, st, A, part);
}

// ---------------------------------------------------------------------------
// Spectral effect of a horfft inverse+forward round trip (horfft.f:14-64):
// the inverse x transform keeps only Re of the kx=0 and kx-Nyquist entries of
// every row (fft.f:1218-1230) and the y transform skips the Nyquist column
// while 1/ny scales every entry (horfft.f:24-30).  Hence fwd(inv(X)) = X for
// 0 < kx < Nyquist, the kx=0 column becomes its ky-Hermitian part
// (X(ky) + conj X(-ky))/2, and the kx-Nyquist column becomes Re(X)/ny.
// ---------------------------------------------------------------------------
__global__ void k_project(FieldBatch fb, int nk, int ny, int nxpp, int nxh, double scaley)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    double *f = fb.f[blockIdx.y];
    const int nh = ny / 2;
    const int per = nh + 1 + ny; // Hermitian pairs j=0..ny/2, then Nyquist rows
    if (id >= nk * per) return;
    const int k = id / per, r = id - k * per;
    double *plane = f + (long)k * nxpp * ny;
    if (r <= nh) {
        const int j = r, jm = (ny - j) % ny;
        double *a = plane + (long)j * nxpp, *b = plane + (long)jm * nxpp;
        const double ar = a[0], ai = a[1], br = b[0], bi = b[1];
        const double hr = 0.5 * (ar + br), hi = 0.5 * (ai - bi);
        a[0] = hr;
        a[1] = hi;
        if (jm != j) {
            b[0] = hr;
            b[1] = -hi;
        }
    } else {
        const int j = r - nh - 1;
        double *a = plane + (long)j * nxpp + 2 * nxh;
        a[0] = scaley * a[0];
        a[1] = 0.;
    }
}

void launch_project(const Geo &g, FieldBatch b, hipStream_t st)
{
    const int per = g.ny / 2 + 1 + g.ny;
    const int n = g.nz * per;
    hipLaunchKernelGGL(k_project, dim3((n + 255) / 256, b.n), dim3(256), 0, st, b, g.nz, g.ny, g.nxpp, g.nxh,
                       g.scaley);
}

} // namespace iw
