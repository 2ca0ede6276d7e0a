// Learned sub-grid closure of the stratified-Boussinesq step (gfx950).
//
// No counterpart in the reference, whose sub-grid model is the estimation
// filter with the SGS tensors of sgscalc.f:125-219 (tau_ij = f(u_i u_j) -
// f(u_i) f(u_j), theta_i = f(u_i rho) - f(u_i) f(rho)); the closure here
// keeps those definitions with the filter f taken as the coarse-graining of a
// fine run onto a coarse grid:
//   * x, y: a sharp spectral cut-off (the coarse grid's modes |kx|, |ky| below
//     its Nyquist wavenumbers, copied from the fine spectrum; the forward
//     transform is normalised by 1/(nx ny), so coefficients are grid
//     independent);
//   * z: per coarse subdomain (r whole fine subdomains), the L2 projection of
//     the fine piecewise polynomial onto the Legendre polynomials of the
//     coarse order, evaluated at the coarse GLL nodes (exact by the fine GLL
//     quadrature).
// A per-point network 12 -> 64 -> 64 -> 64 -> 9 (tanh) maps the filtered
// velocity gradients (9), d rho/dz, |grad_h rho| and z to the six tau_ij and
// the three theta_i (normalised by the training set's mean and deviation).
// Forward: BF16 matrix cores (v_mfma_f32_16x16x16bf16_1k), FP32 accumulation;
// backward: FP32 (per point deltas on VALU, weight gradients on FP32 matrix
// cores); Adam on FP32 master weights.  In a coarse run the predicted
// tensors enter rhsfm's flux form: -tau_ij is added to the horizontal flux
// products of solver.f:975-1063, 1155-1161, and -d/dz of tau_i3, theta_3 to
// the right-hand sides.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "iw_internal.h"

using namespace iw;

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kIn = 12, kInP = 16, kHid = 64, kOut = 9, kOutP = 16;
// FP32 master parameters: W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[16][64] b4[16]
constexpr int kOW1 = 0, kOB1 = kOW1 + kHid * kInP, kOW2 = kOB1 + kHid, kOB2 = kOW2 + kHid * kHid,
              kOW3 = kOB2 + kHid, kOB3 = kOW3 + kHid * kHid, kOW4 = kOB3 + kHid, kOB4 = kOW4 + kOutP * kHid,
              kNParam = kOB4 + kOutP;
constexpr int kNNorm = 2 * kIn + 2 * kOut; // mean and deviation of inputs and targets

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even (finite values)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float bfr(float f) { return bf2f(f2bf(f)); }

struct MlpDev {
    const unsigned short *wb; // BF16 copy of the parameter vector
    const float *p;           // FP32 parameters (biases are read from here)
    const float *norm;        // mean x[12], dev x[12], mean y[9], dev y[9]
};

// weights of one wave in A-operand layout: A[i][k] = W[16 mt + l % 16][16 s + 4 (l / 16) + 0..3]
struct MlpRegs {
    v4s w1[4], w2[4][4], w3[4][4], w4[4];
    float b1[4][4], b2[4][4], b3[4][4], b4[4];
};

__device__ __forceinline__ v4s ldw(const unsigned short *w, int row, int ncol, int col)
{
    return *reinterpret_cast<const v4s *>(w + (long)row * ncol + col);
}

__device__ void mlp_load(const MlpDev &M, MlpRegs &R)
{
    const int l = threadIdx.x & 63, i = l & 15, kq = 4 * (l >> 4);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        R.w1[mt] = ldw(M.wb + kOW1, 16 * mt + i, kInP, kq);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            R.w2[mt][s] = ldw(M.wb + kOW2, 16 * mt + i, kHid, 16 * s + kq);
            R.w3[mt][s] = ldw(M.wb + kOW3, 16 * mt + i, kHid, 16 * s + kq);
        }
        R.w4[mt] = ldw(M.wb + kOW4, i, kHid, 16 * mt + kq);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int f = 16 * mt + kq + r; // feature of result element r
            R.b1[mt][r] = M.p[kOB1 + f];
            R.b2[mt][r] = M.p[kOB2 + f];
            R.b3[mt][r] = M.p[kOB3 + f];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) R.b4[r] = M.p[kOB4 + kq + r];
}

__device__ __forceinline__ v4s pack_tanh(v4f acc, const float (&b)[4], float (&hout)[4])
{
    v4s a;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float h = tanhf(acc[r] + b[r]);
        hout[r] = h;
        a[r] = (short)f2bf(h);
    }
    return a;
}

// One tile of 16 points.  a0: this lane's B operand of layer 1 (normalised
// inputs 4 (l/16)..+3 of point l % 16, BF16); the result holds outputs
// 4 (l/16) + r of point l % 16 (normalised).  h1..h3 (may be null): this
// lane's post-activation values, 4 per 16-feature block (unused values are
// dropped by the compiler).
__device__ __forceinline__ v4f mlp_tile(const MlpRegs &R, v4s a0, float (&h1)[4][4], float (&h2)[4][4],
                                        float (&h3)[4][4])
{
    const v4f z = {0.f, 0.f, 0.f, 0.f};
    v4s a1[4], a2[4], a3[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const v4f acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w1[mt], a0, z, 0, 0, 0);
        a1[mt] = pack_tanh(acc, R.b1[mt], h1[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        v4f acc = z;
#pragma unroll
        for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w2[mt][s], a1[s], acc, 0, 0, 0);
        a2[mt] = pack_tanh(acc, R.b2[mt], h2[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        v4f acc = z;
#pragma unroll
        for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w3[mt][s], a2[s], acc, 0, 0, 0);
        a3[mt] = pack_tanh(acc, R.b3[mt], h3[mt]);
    }
    v4f y = z;
#pragma unroll
    for (int s = 0; s < 4; s++) y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w4[s], a3[s], y, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; r++) y[r] += R.b4[r];
    return y;
}

// normalised input features 4 (l/16)..+3 of one point as the B operand
__device__ __forceinline__ v4s pack_in(const float (&x)[4], const float *norm, int kq)
{
    v4s a;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = kq + i;
        a[i] = k < kIn ? (short)f2bf((x[i] - norm[k]) / norm[kIn + k]) : (short)0;
    }
    return a;
}

// training / evaluation forward over n points of X (n x 12, raw): Y (n x 9,
// normalised) and, when H is not null, the post-activation values (3 x n x 64)
__global__ __launch_bounds__(256) void k_mlp_fwd(MlpDev M, const float *__restrict__ X, int n, float *Y, float *H)
{
    MlpRegs R;
    mlp_load(M, R);
    const int l = threadIdx.x & 63, kq = 4 * (l >> 4);
    const long base = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {
        const long p = base + 16 * nt + (l & 15);
        const long pc = p < n ? p : n - 1;
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        if (kq < kIn) {
            const float4 v = *reinterpret_cast<const float4 *>(X + pc * kIn + kq);
            x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
        }
        float h1[4][4], h2[4][4], h3[4][4];
        const v4f y = mlp_tile(R, pack_in(x, M.norm, kq), h1, h2, h3);
        if (p >= n) continue;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (kq + r < kOut) Y[p * kOut + kq + r] = y[r];
        if (H) {
            const long nn = (long)n * kHid;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const long o = p * kHid + 16 * mt + kq;
                *reinterpret_cast<float4 *>(H + o) = make_float4(h1[mt][0], h1[mt][1], h1[mt][2], h1[mt][3]);
                *reinterpret_cast<float4 *>(H + nn + o) = make_float4(h2[mt][0], h2[mt][1], h2[mt][2], h2[mt][3]);
                *reinterpret_cast<float4 *>(H + 2 * nn + o) = make_float4(h3[mt][0], h3[mt][1], h3[mt][2], h3[mt][3]);
            }
        }
    }
}

// Backward deltas, one thread per point, FP32.  Loss: mean over the n
// points and 9 outputs of (y - t_norm)^2.  DZ (4 x n x 64): dL/dz of layers
// 1..3 and of the output (first 16 of its row), with the straight-through
// rule for the BF16 rounding of weights and activations: the transposed
// products use the BF16 weights as FP32 values.
__global__ __launch_bounds__(256) void k_mlp_delta(MlpDev M, const float *__restrict__ Y, const float *__restrict__ T,
                                                   const float *__restrict__ H, int n, float *DZ, float *loss)
{
    __shared__ float wt[3][kHid * kHid]; // WT[k][j] = W[j][k] of layers 4 (16 x 64 -> [64][16]), 3, 2
    for (int e = threadIdx.x; e < kHid * kOutP; e += blockDim.x) {
        const int j = e / kHid, k = e % kHid;
        wt[0][k * kOutP + j] = bf2f(M.wb[kOW4 + e]);
    }
    for (int e = threadIdx.x; e < kHid * kHid; e += blockDim.x) {
        const int j = e / kHid, k = e % kHid;
        wt[1][k * kHid + j] = bf2f(M.wb[kOW3 + e]);
        wt[2][k * kHid + j] = bf2f(M.wb[kOW2 + e]);
    }
    __syncthreads();
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.f;
    if (p < n) {
        const long nn = (long)n * kHid;
        const float c = 2.f / ((float)n * kOut);
        float d[kHid];
#pragma unroll
        for (int j = 0; j < kOutP; j++) {
            float v = 0.f;
            if (j < kOut) {
                const float tn = (T[p * kOut + j] - M.norm[2 * kIn + j]) / M.norm[2 * kIn + kOut + j];
                const float e = Y[p * kOut + j] - tn;
                lsum += e * e;
                v = c * e;
            }
            d[j] = v;
            DZ[3 * nn + p * kHid + j] = v;
        }
        // layers 3, 2, 1: dz_in[k] = (sum_j W[j][k] dz_out[j]) (1 - h[k]^2)
#pragma unroll 1
        for (int L = 0; L < 3; L++) {
            const int J = L == 0 ? kOutP : kHid;
            const float *h = H + (long)(2 - L) * nn + p * kHid;
            float *dzo = DZ + (long)(2 - L) * nn + p * kHid;
#pragma unroll 1
            for (int k = 0; k < kHid; k++) {
                const float *w = &wt[L][k * J];
                float s = 0.f;
                if (J == kOutP) {
#pragma unroll
                    for (int j = 0; j < kOutP; j++) s += w[j] * d[j];
                } else {
#pragma unroll
                    for (int j = 0; j < kHid; j++) s += w[j] * d[j];
                }
                const float hk = h[k];
                dzo[k] = s * (1.f - hk * hk);
            }
#pragma unroll
            for (int k = 0; k < kHid; k++) d[k] = dzo[k];
        }
    }
    // loss: block sum, one atomic per block
    __shared__ float red[256];
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / ((float)n * kOut));
}

// Weight gradients on FP32 matrix cores (v_mfma_f32_16x16x4f32): tile t of
// the 40 16x16 output tiles of W1..W4, points [c0, c0 + chunk):
// dW[j][k] += sum_p DZ[p][j] A[p][k], A the BF16-rounded input of the layer
// (normalised x for layer 1).  A[i][k] operand: DZ[p0 + l/16][j0 + l%16];
// B[k][j]: A[p0 + l/16][k0 + l%16]; result rows j0 + 4 (l/16) + r, column
// k0 + l % 16.
__global__ __launch_bounds__(64) void k_mlp_wgrad(MlpDev M, const float *__restrict__ X, const float *__restrict__ H,
                                                  const float *__restrict__ DZ, int n, int chunk, float *G)
{
    const int l = threadIdx.x, t = blockIdx.y;
    int layer, jt, kt, ncol, off;
    if (t < 4) { layer = 0; jt = t; kt = 0; ncol = kInP; off = kOW1; }
    else if (t < 20) { layer = 1; jt = (t - 4) >> 2; kt = (t - 4) & 3; ncol = kHid; off = kOW2; }
    else if (t < 36) { layer = 2; jt = (t - 20) >> 2; kt = (t - 20) & 3; ncol = kHid; off = kOW3; }
    else { layer = 3; jt = 0; kt = t - 36; ncol = kHid; off = kOW4; }
    const long nn = (long)n * kHid;
    const float *dz = DZ + (long)layer * nn;
    const float *in = layer == 0 ? nullptr : H + (long)(layer - 1) * nn;
    const int j = 16 * jt + (l & 15), k = 16 * kt + (l & 15);
    const long c0 = (long)blockIdx.x * chunk;
    const long c1 = c0 + chunk < n ? c0 + chunk : n;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (long p0 = c0; p0 < c1; p0 += 4) {
        const long p = p0 + (l >> 4);
        float a = 0.f, b = 0.f;
        if (p < c1) {
            a = dz[p * kHid + j];
            if (layer == 0)
                b = k < kIn ? bfr((X[p * kIn + k] - M.norm[k]) / M.norm[kIn + k]) : 0.f;
            else
                b = bfr(in[p * kHid + k]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int jj = 16 * jt + 4 * (l >> 4) + r, kk = 16 * kt + (l & 15);
        atomicAdd(G + off + jj * ncol + kk, acc[r]);
    }
}

// bias gradients: db[j] = sum_p DZ[p][j], one block per (layer, chunk)
__global__ __launch_bounds__(256) void k_mlp_bgrad(const float *__restrict__ DZ, int n, int chunk, float *G)
{
    const int layer = blockIdx.y;
    const int off = layer == 0 ? kOB1 : layer == 1 ? kOB2 : layer == 2 ? kOB3 : kOB4;
    const int J = layer == 3 ? kOutP : kHid;
    const long nn = (long)n * kHid;
    const int j = threadIdx.x & 63, pr = threadIdx.x >> 6;
    const long c0 = (long)blockIdx.x * chunk;
    const long c1 = c0 + chunk < n ? c0 + chunk : n;
    float s = 0.f;
    if (j < J)
        for (long p = c0 + pr; p < c1; p += 4) s += DZ[(long)layer * nn + p * kHid + j];
    __shared__ float red[4][64];
    red[pr][j] = s;
    __syncthreads();
    if (pr == 0 && j < J) atomicAdd(G + off + j, red[0][j] + red[1][j] + red[2][j] + red[3][j]);
}

__global__ void k_fill(float *p, float v, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Adam (FP32 master, bias-corrected), then the BF16 copy; padded entries
// have zero gradient and stay zero
__global__ void k_adam(float *p, float *m, float *v, const float *__restrict__ g, unsigned short *wb, int n,
                       float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = w;
    wb[i] = f2bf(w);
}

__global__ void k_tobf(const float *__restrict__ p, unsigned short *wb, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(p[i]);
}

// minibatch gather with a counter-based (SplitMix64) index stream
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ T, long ntot, int nb,
                         unsigned long long seed, float *Xb, float *Tb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nb) return;
    unsigned long long z = seed + (unsigned long long)(p + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const long q = (long)(z % (unsigned long long)ntot);
    for (int k = 0; k < kIn; k++) Xb[(long)p * kIn + k] = X[q * kIn + k];
    for (int k = 0; k < kOut; k++) Tb[(long)p * kOut + k] = T[q * kOut + k];
}

// sums and sums of squares of the columns of A (n x nc) in FP64, one atomic
// per block and column
__global__ __launch_bounds__(256) void k_colstats(const float *__restrict__ A, long n, int nc, double *acc)
{
    __shared__ double s1[256], s2[256];
    for (int c = 0; c < nc; c++) {
        double a = 0., b = 0.;
        for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
            const double x = A[p * nc + c];
            a += x;
            b += x * x;
        }
        s1[threadIdx.x] = a;
        s2[threadIdx.x] = b;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                s1[threadIdx.x] += s1[threadIdx.x + o];
                s2[threadIdx.x] += s2[threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            atomicAdd(acc + 2 * c, s1[0]);
            atomicAdd(acc + 2 * c + 1, s2[0]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// coarse-graining and the training samples
// ---------------------------------------------------------------------------
// coarse spectral column (mode ii_c, j_c; coarse subdomain m) from the fine
// one: sharp cut-off in x, y and the z projection PI (nlc x r nlf) over the
// r fine subdomains of m
struct CoarseArgs {
    const double *f[kMaxFields];
    double *c[kMaxFields];
    int nf;
    int nxpp_f, ny_f, nl_f, nxpp_c, ny_c, nxh_c, nl_c, r;
};

__global__ __launch_bounds__(256) void k_coarsen(CoarseArgs A, const double *__restrict__ PI)
{
    const int nmode = (A.nxh_c + 1) * A.ny_c;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nmode) return;
    const int m = blockIdx.y, q = blockIdx.z;
    const int ii = e % (A.nxh_c + 1), jc = e / (A.nxh_c + 1);
    const int hy = A.ny_c / 2;
    const bool keep = ii < A.nxh_c && (A.ny_c == 1 || jc != hy);
    const int jf = (A.ny_c == 1 || jc < hy) ? jc : A.ny_f - (A.ny_c - jc);
    const long stf = (long)A.nxpp_f * A.ny_f, stc = (long)A.nxpp_c * A.ny_c;
    const double *fc = A.f[q] + (long)jf * A.nxpp_f + 2 * ii + (long)m * A.r * A.nl_f * stf;
    double *cc = A.c[q] + (long)jc * A.nxpp_c + 2 * ii + (long)m * A.nl_c * stc;
    const int K = A.r * A.nl_f;
    for (int c = 0; c < A.nl_c; c++) {
        double sr = 0., si = 0.;
        if (keep)
            for (int k = 0; k < K; k++) {
                const double w = PI[c * K + k];
                sr += w * fc[(long)k * stf];
                si += w * fc[(long)k * stf + 1];
            }
        cc[(long)c * stc] = sr;
        cc[(long)c * stc + 1] = si;
    }
}

// the nine products u u, u v, u w, v v, v w, w w, u rho, v rho, w rho
struct Ptr9 {
    double *p[9];
};
__global__ __launch_bounds__(256) void k_prod9(const double *__restrict__ u, const double *__restrict__ v,
                                               const double *__restrict__ w, const double *__restrict__ t,
                                               Ptr9 out, long ntot)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ntot) return;
    const double a = u[n], b = v[n], c = w[n], d = t[n];
    out.p[0][n] = a * a; out.p[1][n] = a * b; out.p[2][n] = a * c;
    out.p[3][n] = b * b; out.p[4][n] = b * c; out.p[5][n] = c * c;
    out.p[6][n] = a * d; out.p[7][n] = b * d; out.p[8][n] = c * d;
}

struct SampleArgs {
    const double *u, *v, *w, *t; // filtered fields (physical, coarse)
    const double *P[9];          // filtered products (physical, coarse)
    const double *g[9];          // dudx dudy dudz dvdx .. dwdz (physical, coarse)
    const double *rx, *ry, *rz;  // rho gradients
    const double *z;             // height of each level
    int nx, ny, nxpp;
    long nxppy;
};

// the point (k, j, i < nx) of sample p
__device__ __forceinline__ long sample_off(const SampleArgs &A, long p)
{
    const long per = (long)A.nx * A.ny;
    const int k = (int)(p / per);
    const long rem = p - k * per;
    const int j = (int)(rem / A.nx), i = (int)(rem - (long)j * A.nx);
    return (long)k * A.nxppy + (long)j * A.nxpp + i;
}

__global__ __launch_bounds__(256) void k_targets(SampleArgs A, long n, float *T)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long o = sample_off(A, p);
    const double u = A.u[o], v = A.v[o], w = A.w[o], t = A.t[o];
    const double fp[9] = {u * u, u * v, u * w, v * v, v * w, w * w, u * t, v * t, w * t};
    for (int q = 0; q < 9; q++) T[p * kOut + q] = (float)(A.P[q][o] - fp[q]);
}

__device__ __forceinline__ void features(const SampleArgs &A, long o, int k, float (&x)[kIn])
{
    for (int q = 0; q < 9; q++) x[q] = (float)A.g[q][o];
    const double rx = A.rx[o], ry = A.ry[o];
    x[9] = (float)A.rz[o];
    x[10] = (float)sqrt(rx * rx + ry * ry);
    x[11] = (float)A.z[k];
}

__global__ __launch_bounds__(256) void k_inputs(SampleArgs A, long n, float *X)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long o = sample_off(A, p);
    float x[kIn];
    features(A, o, (int)(o / A.nxppy), x);
    for (int q = 0; q < kIn; q++) X[p * kIn + q] = x[q];
}

// The closure in a coarse run (FP64 fields of convec): for every stored
// point, the network's tau from the features; -tau_ij into the horizontal
// flux products pr (the x and y fluxes of S_u .. S_rho, solver.f:975-1063,
// 1155-1161) and tau_13, tau_23, tau_33, theta_3 into V for their z
// derivatives.  Pads (i >= nx) get nothing.
struct ApplyArgs {
    SampleArgs S;
    double *pr[7];
    double *V[4];
    long ntot;
};

__global__ __launch_bounds__(256) void k_closure_apply(MlpDev M, ApplyArgs A)
{
    MlpRegs R;
    mlp_load(M, R);
    const int l = threadIdx.x & 63, kq = 4 * (l >> 4);
    const long base = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    const SampleArgs &S = A.S;
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {
        const long o = base + 16 * nt + (l & 15);
        const long oc = o < A.ntot ? o : A.ntot - 1;
        const int k = (int)(oc / S.nxppy), i = (int)(oc % S.nxpp);
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        if (kq < 8) {
#pragma unroll
            for (int q = 0; q < 4; q++) x[q] = (float)S.g[kq + q][oc];
        } else if (kq == 8) {
            const double rx = S.rx[oc], ry = S.ry[oc];
            x[0] = (float)S.g[8][oc];
            x[1] = (float)S.rz[oc];
            x[2] = (float)sqrt(rx * rx + ry * ry);
            x[3] = (float)S.z[k];
        }
        float h1[4][4], h2[4][4], h3[4][4];
        const v4f y = mlp_tile(R, pack_in(x, M.norm, kq), h1, h2, h3);
        if (o >= A.ntot) continue;
        const bool live = i < S.nx;
        // tau component c = kq + r (t11 t12 t13 t22 t23 t33 th1 th2 th3)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = kq + r;
            if (c >= kOut) continue;
            const double tau = live ? (double)(M.norm[2 * kIn + c] + M.norm[2 * kIn + kOut + c] * y[r]) : 0.0;
            // horizontal fluxes: pr0 x(S_u) <- t11, pr1 y(S_u) and x(S_v) <- t12,
            // pr2 y(S_v) <- t22, pr3 x(S_w) <- t13, pr4 y(S_w) <- t23,
            // pr5 x(S_rho) <- th1, pr6 y(S_rho) <- th2
            const int ph[9] = {0, 1, 3, 2, 4, -1, 5, 6, -1};
            if (ph[c] >= 0) A.pr[ph[c]][o] -= tau;
            // vertical: V0 t13 (S_u), V1 t23 (S_v), V2 t33 (S_w), V3 th3 (S_rho)
            const int pv[9] = {-1, -1, 0, -1, 1, 2, -1, -1, 3};
            if (pv[c] >= 0) A.V[pv[c]][o] = tau;
        }
    }
}

// S_q -= d/dz V_q (the vertical SGS divergence, dz by dpdz)
struct Sub4 {
    double *S[4];
    const double *D[4];
};
__global__ __launch_bounds__(256) void k_sub4(Sub4 A, long ntot)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ntot) return;
    for (int q = 0; q < 4; q++) A.S[q][n] -= A.D[q][n];
}

} // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct iw_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *norm = nullptr;
    unsigned short *wb = nullptr;
    int adam_t = 0;
    bool norm_set = false; // normalisation given by iw_mlp_set_params or taken from the training set
    // training set (device): inputs n x 12, targets n x 9
    float *X = nullptr, *T = nullptr;
    long n = 0, cap = 0;
    // minibatch scratch
    float *Xb = nullptr, *Tb = nullptr, *Yb = nullptr, *Hb = nullptr, *DZb = nullptr, *lossb = nullptr;
    int nb_cap = 0;
    MlpDev dev() const { return MlpDev{wb, p, norm}; }
};

static std::string g_cerr;

#define CK(x)                                                                                                  \
    do {                                                                                                       \
        hipError_t e_ = (x);                                                                                   \
        if (e_ != hipSuccess) return iw_fail(std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

static int mlp_batch_alloc(iw_mlp *m, int nb)
{
    if (nb <= m->nb_cap) return 0;
    for (void *q : {(void *)m->Xb, (void *)m->Tb, (void *)m->Yb, (void *)m->Hb, (void *)m->DZb})
        if (q) (void)hipFree(q);
    CK(hipMalloc((void **)&m->Xb, sizeof(float) * (size_t)nb * kIn));
    CK(hipMalloc((void **)&m->Tb, sizeof(float) * (size_t)nb * kOut));
    CK(hipMalloc((void **)&m->Yb, sizeof(float) * (size_t)nb * kOut));
    CK(hipMalloc((void **)&m->Hb, sizeof(float) * (size_t)nb * kHid * 3));
    CK(hipMalloc((void **)&m->DZb, sizeof(float) * (size_t)nb * kHid * 4));
    m->nb_cap = nb;
    return 0;
}

extern "C" int iw_mlp_nparams(void) { return kNParam + kNNorm; }

extern "C" int iw_mlp_create(unsigned long long seed, iw_mlp **out)
{
    if (!out) return iw_fail("iw_mlp_create: null argument");
    iw_mlp *m = new iw_mlp();
    CK(hipGetDevice(&m->device));
    CK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    CK(hipMalloc((void **)&m->p, sizeof(float) * kNParam));
    CK(hipMalloc((void **)&m->m, sizeof(float) * kNParam));
    CK(hipMalloc((void **)&m->v, sizeof(float) * kNParam));
    CK(hipMalloc((void **)&m->g, sizeof(float) * kNParam));
    CK(hipMalloc((void **)&m->norm, sizeof(float) * kNNorm));
    CK(hipMalloc((void **)&m->wb, sizeof(unsigned short) * kNParam));
    CK(hipMalloc((void **)&m->lossb, sizeof(float)));
    // Glorot-uniform weights from a SplitMix64 stream, zero biases, zero pads
    std::vector<float> h(kNParam, 0.f), nrm(kNNorm, 0.f);
    unsigned long long z = seed;
    auto uni = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        unsigned long long x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int rows, int cols, int nr, int nc) {
        const double a = std::sqrt(6.0 / (nr + nc));
        for (int j = 0; j < rows; j++)
            for (int k = 0; k < cols; k++)
                if (j < nr && k < nc) h[off + j * cols + k] = (float)(a * (2. * uni() - 1.));
    };
    fill(kOW1, kHid, kInP, kHid, kIn);
    fill(kOW2, kHid, kHid, kHid, kHid);
    fill(kOW3, kHid, kHid, kHid, kHid);
    fill(kOW4, kOutP, kHid, kOut, kHid);
    for (int k = 0; k < kIn; k++) nrm[kIn + k] = 1.f;
    for (int k = 0; k < kOut; k++) nrm[2 * kIn + kOut + k] = 1.f;
    CK(hipMemcpy(m->p, h.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice));
    CK(hipMemcpy(m->norm, nrm.data(), sizeof(float) * kNNorm, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fill, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->m, 0.f, (long)kNParam);
    hipLaunchKernelGGL(k_fill, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->v, 0.f, (long)kNParam);
    hipLaunchKernelGGL(k_tobf, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->p, m->wb, kNParam);
    CK(hipStreamSynchronize(m->stream));
    *out = m;
    return 0;
}

extern "C" int iw_mlp_destroy(iw_mlp *m)
{
    if (!m) return 0;
    (void)hipStreamSynchronize(m->stream);
    for (void *q : {(void *)m->p, (void *)m->m, (void *)m->v, (void *)m->g, (void *)m->norm, (void *)m->wb,
                    (void *)m->X, (void *)m->T, (void *)m->Xb, (void *)m->Tb, (void *)m->Yb, (void *)m->Hb,
                    (void *)m->DZb, (void *)m->lossb})
        if (q) (void)hipFree(q);
    (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

// parameters followed by the normalisation (iw_mlp_nparams floats)
extern "C" int iw_mlp_get_params(iw_mlp *m, float *out)
{
    if (!m || !out) return iw_fail("iw_mlp_get_params: null argument");
    CK(hipStreamSynchronize(m->stream));
    CK(hipMemcpy(out, m->p, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    CK(hipMemcpy(out + kNParam, m->norm, sizeof(float) * kNNorm, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int iw_mlp_set_params(iw_mlp *m, const float *in)
{
    if (!m || !in) return iw_fail("iw_mlp_set_params: null argument");
    for (int k = 0; k < kNNorm; k++)
        if (k >= kIn && k < 2 * kIn && !(in[kNParam + k] > 0.f))
            return iw_fail("iw_mlp_set_params: input deviations must be positive");
    CK(hipStreamSynchronize(m->stream));
    CK(hipMemcpy(m->p, in, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    CK(hipMemcpy(m->norm, in + kNParam, sizeof(float) * kNNorm, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tobf, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->p, m->wb, kNParam);
    m->adam_t = 0;
    m->norm_set = true;
    hipLaunchKernelGGL(k_fill, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->m, 0.f, (long)kNParam);
    hipLaunchKernelGGL(k_fill, dim3((kNParam + 255) / 256), dim3(256), 0, m->stream, m->v, 0.f, (long)kNParam);
    CK(hipStreamSynchronize(m->stream));
    return 0;
}

// forward + backward of a device minibatch (Xb, Tb: nb points): loss into
// lossb, gradients into g (zeroed first)
static void mlp_grad_dev(iw_mlp *m, int nb)
{
    hipStream_t st = m->stream;
    const MlpDev M = m->dev();
    hipLaunchKernelGGL(k_fill, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->g, 0.f, (long)kNParam);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, m->lossb, 0.f, 1L);
    hipLaunchKernelGGL(k_mlp_fwd, dim3((nb + 255) / 256), dim3(256), 0, st, M, m->Xb, nb, m->Yb, m->Hb);
    hipLaunchKernelGGL(k_mlp_delta, dim3((nb + 255) / 256), dim3(256), 0, st, M, m->Yb, m->Tb, m->Hb, nb, m->DZb,
                       m->lossb);
    const int chunk = 1024;
    const int nch = (nb + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(nch, 40), dim3(64), 0, st, M, m->Xb, m->Hb, m->DZb, nb, chunk, m->g);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(nch, 4), dim3(256), 0, st, m->DZb, nb, chunk, m->g);
}

// Forward of host inputs x (n x 12, raw): y (n x 9) the predicted tau
// (denormalised)
extern "C" int iw_mlp_forward(iw_mlp *m, const float *x, int n, float *y)
{
    if (!m || !x || !y || n <= 0) return iw_fail("iw_mlp_forward: bad argument");
    if (mlp_batch_alloc(m, n)) return -1;
    CK(hipMemcpyAsync(m->Xb, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_fwd, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->dev(), m->Xb, n, m->Yb,
                       (float *)nullptr);
    std::vector<float> yn((size_t)n * kOut), nrm(kNNorm);
    CK(hipMemcpyAsync(yn.data(), m->Yb, sizeof(float) * yn.size(), hipMemcpyDeviceToHost, m->stream));
    CK(hipMemcpyAsync(nrm.data(), m->norm, sizeof(float) * kNNorm, hipMemcpyDeviceToHost, m->stream));
    CK(hipStreamSynchronize(m->stream));
    for (long p = 0; p < n; p++)
        for (int c = 0; c < kOut; c++)
            y[p * kOut + c] = nrm[2 * kIn + c] + nrm[2 * kIn + kOut + c] * yn[p * kOut + c];
    return 0;
}

// Loss (mean squared error of the normalised outputs) and its gradient with
// respect to the parameters (kNParam floats) for host inputs and targets
extern "C" int iw_mlp_gradient(iw_mlp *m, const float *x, const float *t, int n, float *grad, float *loss)
{
    if (!m || !x || !t || !grad || n <= 0) return iw_fail("iw_mlp_gradient: bad argument");
    if (mlp_batch_alloc(m, n)) return -1;
    CK(hipMemcpyAsync(m->Xb, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice, m->stream));
    CK(hipMemcpyAsync(m->Tb, t, sizeof(float) * (size_t)n * kOut, hipMemcpyHostToDevice, m->stream));
    mlp_grad_dev(m, n);
    CK(hipMemcpyAsync(grad, m->g, sizeof(float) * kNParam, hipMemcpyDeviceToHost, m->stream));
    float l = 0.f;
    CK(hipMemcpyAsync(&l, m->lossb, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    CK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
    return 0;
}

// z projection from the fine subdomains onto the coarse ones (see the file
// header): PI[c][q nl_f + k] for coarse node c, fine subdomain q of the
// coarse one, fine node k
static std::vector<double> zprojection(const iw_solver *f, const iw_solver *c, int r)
{
    const int nlf = f->P.nzloc, nlc = c->P.nzloc, Nc = nlc - 1;
    std::vector<double> PI((size_t)nlc * r * nlf, 0.);
    auto legendre = [](int N, double x, std::vector<double> &P) {
        P.assign(N + 1, 0.);
        P[0] = 1.;
        if (N >= 1) P[1] = x;
        for (int n = 1; n < N; n++) P[n + 1] = ((2. * n + 1.) * x * P[n] - n * P[n - 1]) / (n + 1.);
    };
    std::vector<double> Pc, Pf;
    // coarse subdomain 0 (all coarse subdomains have the same shape when the
    // fine heights are uniform; checked by the caller)
    const double a = c->P.z0[0], b = a + c->P.zh[0];
    for (int cn = 0; cn < nlc; cn++) {
        legendre(Nc, c->zpts[cn], Pc);
        for (int q = 0; q < r; q++) {
            const double zf0 = f->P.z0[q], hq = f->P.zh[q];
            for (int k = 0; k < nlf; k++) {
                const double zz = zf0 + 0.5 * hq * (f->zpts[k] + 1.);
                const double xi = 2. * (zz - a) / (b - a) - 1.;
                const double wq = f->wg[k] * hq / (b - a);
                legendre(Nc, xi, Pf);
                double s = 0.;
                for (int n = 0; n <= Nc; n++) s += 0.5 * (2. * n + 1.) * Pf[n] * Pc[n];
                PI[(size_t)cn * r * nlf + q * nlf + k] = wq * s;
            }
        }
    }
    return PI;
}

static int coarsen_fields(iw_solver *f, iw_solver *c, const std::vector<double *> &src,
                          const std::vector<double *> &dst, const double *PI, int r)
{
    const Geo &gf = f->g, &gc = c->g;
    for (size_t i = 0; i < src.size(); i += kMaxFields) {
        CoarseArgs A;
        A.nf = 0;
        for (size_t q = i; q < src.size() && q < i + kMaxFields; q++) {
            A.f[A.nf] = src[q];
            A.c[A.nf] = dst[q];
            A.nf++;
        }
        A.nxpp_f = gf.nxpp; A.ny_f = gf.ny; A.nl_f = gf.nzloc;
        A.nxpp_c = gc.nxpp; A.ny_c = gc.ny; A.nxh_c = gc.nxh; A.nl_c = gc.nzloc; A.r = r;
        const int nmode = (gc.nxh + 1) * gc.ny;
        hipLaunchKernelGGL(k_coarsen, dim3((nmode + 255) / 256, gc.nsubd, A.nf), dim3(256), 0, c->stream, A, PI);
    }
    return hipGetLastError() == hipSuccess ? 0 : iw_fail("coarsen launch failed");
}

static int check_pair(const iw_solver *f, const iw_solver *c, int &r)
{
    const iw_params &F = f->P, &C = c->P;
    if (f->device != c->device) return iw_fail("closure: fine and coarse instances on different devices");
    if (F.xlen != C.xlen || F.ylen != C.ylen) return iw_fail("closure: fine and coarse boxes differ");
    if (C.nx > F.nx || C.ny > F.ny || C.ny < 2) return iw_fail("closure: the coarse grid must be 3-D and no finer");
    if (F.nsubd % C.nsubd) return iw_fail("closure: fine subdomains must split the coarse ones evenly");
    r = F.nsubd / C.nsubd;
    for (int q = 0; q < F.nsubd; q++)
        if (std::fabs(F.zh[q] - F.zh[0]) > 1e-14 * F.zh[0]) return iw_fail("closure: fine subdomains must be equal");
    for (int m = 0; m < C.nsubd; m++)
        if (std::fabs(C.z0[m] - F.z0[m * r]) > 1e-12 || std::fabs(C.zh[m] - r * F.zh[0]) > 1e-12)
            return iw_fail("closure: coarse subdomains must be unions of fine ones");
    return 0;
}

// physical fields, gradients and rho gradients of the coarse state into the
// coarse instance's scratch: phys u,v,w,rho; grad[0..8]; rx, ry, rz
static void coarse_fields(iw_solver *c, double *rx, double *ry, double *rz)
{
    const Geo &g = c->g;
    double **B = c->buf;
    FieldBatch by;
    by.n = 4;
    by.smap = by.dmap = nullptr;
    for (int q = 0; q < 4; q++) {
        by.f[q] = B[q];
        by.g[q] = c->phys[q];
    }
    by.dy[0] = c->grad[1];
    by.dy[1] = c->grad[4];
    by.dy[2] = c->grad[7];
    by.dy[3] = ry;
    by.kw = g.yw;
    launch_fft_y(g, by, g.nz, 1, c->stream);
    FieldBatch bx;
    bx.n = 8;
    bx.smap = bx.dmap = nullptr;
    double *xf[8] = {c->phys[0], c->phys[1], c->phys[2], c->phys[3], c->grad[1], c->grad[4], c->grad[7], ry};
    for (int q = 0; q < 8; q++) bx.f[q] = bx.g[q] = xf[q];
    bx.dx[0] = c->grad[0];
    bx.dx[1] = c->grad[3];
    bx.dx[2] = c->grad[6];
    bx.dx[3] = rx;
    bx.kw = g.xw;
    launch_fft_x(g, bx, g.nz, 1, c->stream);
    double *in3[3] = {c->phys[0], c->phys[1], c->phys[2]}, *out3[3] = {c->grad[2], c->grad[5], c->grad[8]};
    launch_dpdz3(g, in3, out3, c->stream);
    launch_dpdz(g, c->phys[3], rz, c->stream);
}

static int closure_scratch(iw_solver *c)
{
    if (c->cl_buf[0]) return 0;
    for (int q = 0; q < 8; q++) {
        CK(hipMalloc((void **)&c->cl_buf[q], sizeof(double) * (size_t)c->g.ntot));
        c->dev_bytes += (long long)(sizeof(double) * c->g.ntot);
    }
    return 0;
}

static SampleArgs sample_args(iw_solver *c)
{
    SampleArgs A;
    A.u = c->phys[0]; A.v = c->phys[1]; A.w = c->phys[2]; A.t = c->phys[3];
    for (int q = 0; q < 9; q++) {
        A.P[q] = c->grad[q];
        A.g[q] = c->grad[q];
    }
    A.rx = c->cl_buf[0]; A.ry = c->cl_buf[1]; A.rz = c->cl_buf[2];
    A.z = c->g.z;
    A.nx = c->g.nx; A.ny = c->g.ny; A.nxpp = c->g.nxpp; A.nxppy = c->g.nxppy;
    return A;
}

// Coarse-grain the fine instance's state (u,v,w,rho) into the coarse
// instance's state (its other arrays are left as they are) and, when m is
// not null, append the coarse grid's samples (nx ny nz points: the 12 inputs
// and the 9 SGS targets) to m's training set.  Both instances must be on the
// current device; the fine instance's scratch arrays are overwritten.
extern "C" int iw_closure_sample(iw_mlp *m, iw_solver *f, iw_solver *c)
{
    if (!f || !c) return iw_fail("iw_closure_sample: null solver");
    int r;
    if (check_pair(f, c, r)) return -1;
    if (closure_scratch(c)) return -1;
    CK(hipStreamSynchronize(f->stream));
    CK(hipStreamSynchronize(c->stream));
    const Geo &gf = f->g, &gc = c->g;
    std::vector<double> PI = zprojection(f, c, r);
    double *dPI;
    CK(hipMalloc((void **)&dPI, sizeof(double) * PI.size()));
    CK(hipMemcpy(dPI, PI.data(), sizeof(double) * PI.size(), hipMemcpyHostToDevice));
    // fine: physical u,v,w,rho and the nine products, back to spectral
    iw_fft_batch(f, {f->buf[0], f->buf[1], f->buf[2], f->buf[3]}, {f->phys[0], f->phys[1], f->phys[2], f->phys[3]}, 1);
    Ptr9 d9;
    for (int q = 0; q < 9; q++) d9.p[q] = f->grad[q];
    hipLaunchKernelGGL(k_prod9, dim3((unsigned)((gf.ntot + 255) / 256)), dim3(256), 0, f->stream, f->phys[0],
                       f->phys[1], f->phys[2], f->phys[3], d9, gf.ntot);
    std::vector<double *> g9(f->grad, f->grad + 9);
    iw_fft_batch(f, g9, g9, -1);
    CK(hipStreamSynchronize(f->stream));
    // coarse state and filtered products (into the coarse grad scratch)
    if (coarsen_fields(f, c, {f->buf[0], f->buf[1], f->buf[2], f->buf[3]}, {c->buf[0], c->buf[1], c->buf[2], c->buf[3]},
                       dPI, r))
        return -1;
    std::vector<double *> c9(c->grad, c->grad + 9);
    if (coarsen_fields(f, c, g9, c9, dPI, r)) return -1;
    if (m) {
        const long ns = (long)gc.nx * gc.ny * gc.nz;
        if (m->n + ns > m->cap) {
            const long cap = std::max(2 * m->cap, m->n + ns);
            float *X2, *T2;
            CK(hipMalloc((void **)&X2, sizeof(float) * (size_t)cap * kIn));
            CK(hipMalloc((void **)&T2, sizeof(float) * (size_t)cap * kOut));
            if (m->n) {
                CK(hipMemcpy(X2, m->X, sizeof(float) * (size_t)m->n * kIn, hipMemcpyDeviceToDevice));
                CK(hipMemcpy(T2, m->T, sizeof(float) * (size_t)m->n * kOut, hipMemcpyDeviceToDevice));
            }
            if (m->X) (void)hipFree(m->X);
            if (m->T) (void)hipFree(m->T);
            m->X = X2;
            m->T = T2;
            m->cap = cap;
        }
        // physical filtered fields and products: targets
        iw_fft_batch(c, {c->buf[0], c->buf[1], c->buf[2], c->buf[3]}, {c->phys[0], c->phys[1], c->phys[2], c->phys[3]},
                     1);
        iw_fft_batch(c, c9, c9, 1);
        SampleArgs A = sample_args(c);
        hipLaunchKernelGGL(k_targets, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, A, ns,
                           m->T + (size_t)m->n * kOut);
        // gradients: inputs
        coarse_fields(c, c->cl_buf[0], c->cl_buf[1], c->cl_buf[2]);
        hipLaunchKernelGGL(k_inputs, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, A, ns,
                           m->X + (size_t)m->n * kIn);
        CK(hipStreamSynchronize(c->stream));
        m->n += ns;
    }
    CK(hipStreamSynchronize(c->stream));
    (void)hipFree(dPI);
    return hipGetLastError() == hipSuccess ? 0 : iw_fail("closure sample launch failed");
}

// Number of samples in the training set; with n > 0 copies the first n
// (inputs n x 12, targets n x 9) to the host.
extern "C" long long iw_closure_dataset(iw_mlp *m, float *x, float *t, long long n)
{
    if (!m) return iw_fail("iw_closure_dataset: null argument");
    if (n > m->n || n < 0) return iw_fail("iw_closure_dataset: more samples asked for than there are");
    if (n > 0) {
        if (!x || !t) return iw_fail("iw_closure_dataset: null argument");
        if (hipMemcpy(x, m->X, sizeof(float) * (size_t)n * kIn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(t, m->T, sizeof(float) * (size_t)n * kOut, hipMemcpyDeviceToHost) != hipSuccess)
            return iw_fail("iw_closure_dataset: copy failed");
    }
    return m->n;
}

// The normalisation (mean and deviation of every input and target) from the
// training set, then `steps` Adam steps (lr, betas 0.9 / 0.999, eps 1e-8) on
// minibatches of nb samples drawn with replacement; loss_hist (may be null)
// gets the loss of every step.
extern "C" int iw_mlp_train(iw_mlp *m, int steps, int nb, float lr, unsigned long long seed, float *loss_hist)
{
    if (!m || steps < 0 || nb <= 0) return iw_fail("iw_mlp_train: bad argument");
    if (m->n < 1) return iw_fail("iw_mlp_train: empty training set");
    if (mlp_batch_alloc(m, nb)) return -1;
    hipStream_t st = m->stream;
    if (!m->norm_set) { // normalisation from the training set
        double *acc;
        CK(hipMalloc((void **)&acc, sizeof(double) * 2 * (kIn + kOut)));
        CK(hipMemsetAsync(acc, 0, sizeof(double) * 2 * (kIn + kOut), st));
        hipLaunchKernelGGL(k_colstats, dim3(512), dim3(256), 0, st, m->X, m->n, kIn, acc);
        hipLaunchKernelGGL(k_colstats, dim3(512), dim3(256), 0, st, m->T, m->n, kOut, acc + 2 * kIn);
        std::vector<double> h(2 * (kIn + kOut));
        CK(hipMemcpyAsync(h.data(), acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        (void)hipFree(acc);
        std::vector<float> nrm(kNNorm);
        for (int c = 0; c < kIn + kOut; c++) {
            const double mu = h[2 * c] / m->n, var = std::max(h[2 * c + 1] / m->n - mu * mu, 0.);
            const double sd = var > 0. ? std::sqrt(var) : 1.;
            if (c < kIn) {
                nrm[c] = (float)mu;
                nrm[kIn + c] = (float)sd;
            } else {
                nrm[2 * kIn + (c - kIn)] = (float)mu;
                nrm[2 * kIn + kOut + (c - kIn)] = (float)sd;
            }
        }
        CK(hipMemcpyAsync(m->norm, nrm.data(), sizeof(float) * kNNorm, hipMemcpyHostToDevice, st));
        CK(hipStreamSynchronize(st));
        m->norm_set = true;
    }
    for (int s = 0; s < steps; s++) {
        hipLaunchKernelGGL(k_gather, dim3((nb + 255) / 256), dim3(256), 0, st, m->X, m->T, m->n, nb,
                           seed + 0x632BE59BD9B4E019ull * (unsigned long long)(m->adam_t + 1), m->Xb, m->Tb);
        mlp_grad_dev(m, nb);
        m->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, st, m->p, m->m, m->v, m->g, m->wb,
                           kNParam, lr, b1, b2, 1e-8f, c1, c2);
        if (loss_hist) CK(hipMemcpyAsync(loss_hist + s, m->lossb, sizeof(float), hipMemcpyDeviceToHost, st));
    }
    CK(hipStreamSynchronize(st));
    return hipGetLastError() == hipSuccess ? 0 : iw_fail("iw_mlp_train launch failed");
}

// The network's tau for the coarse instance's current state at its sample
// points (nx ny nz x 9, the order of iw_closure_sample), for a priori tests.
extern "C" int iw_closure_predict(iw_mlp *m, iw_solver *c, float *tau)
{
    if (!m || !c || !tau) return iw_fail("iw_closure_predict: null argument");
    if (c->g.ndims != 3) return iw_fail("iw_closure_predict: 3-D only");
    if (closure_scratch(c)) return -1;
    const long ns = (long)c->g.nx * c->g.ny * c->g.nz;
    if (ns > INT32_MAX) return iw_fail("iw_closure_predict: too many points");
    if (mlp_batch_alloc(m, (int)ns)) return -1;
    CK(hipStreamSynchronize(m->stream));
    coarse_fields(c, c->cl_buf[0], c->cl_buf[1], c->cl_buf[2]);
    SampleArgs A = sample_args(c);
    hipLaunchKernelGGL(k_inputs, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, A, ns, m->Xb);
    CK(hipStreamSynchronize(c->stream));
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, m->stream, m->dev(), m->Xb,
                       (int)ns, m->Yb, (float *)nullptr);
    std::vector<float> yn((size_t)ns * kOut), nrm(kNNorm);
    CK(hipMemcpyAsync(yn.data(), m->Yb, sizeof(float) * yn.size(), hipMemcpyDeviceToHost, m->stream));
    CK(hipMemcpyAsync(nrm.data(), m->norm, sizeof(float) * kNNorm, hipMemcpyDeviceToHost, m->stream));
    CK(hipStreamSynchronize(m->stream));
    for (long p = 0; p < ns; p++)
        for (int q = 0; q < kOut; q++) tau[p * kOut + q] = nrm[2 * kIn + q] + nrm[2 * kIn + kOut + q] * yn[p * kOut + q];
    return 0;
}

// The closure in the coarse instance's steps (m null: off).  m must outlive
// its use by c.
extern "C" int iw_set_closure(iw_solver *c, iw_mlp *m)
{
    if (!c) return iw_fail("iw_set_closure: null solver");
    if (m) {
        if (c->g.ndims != 3) return iw_fail("iw_set_closure: 3-D only");
        if (c->mixed) return iw_fail("iw_set_closure: the closure runs on the plain FP64 path");
        if (c->mw_on || c->P.uo * c->P.fcshear != 0.0)
            return iw_fail("iw_set_closure: not with the mean-wave and shear-current terms");
        if (m->device != c->device) return iw_fail("iw_set_closure: network on another device");
        if (closure_scratch(c)) return -1;
    }
    c->closure = m;
    return 0;
}

// convec's hook (iw_api.hip): after the physical-space right-hand sides and
// flux products of the coarse step are formed, with the rho gradients in
// cl_buf[0..2] (x, y, z)
int iw_closure_apply(iw_solver *c, double *const *S, double *const *pr)
{
    iw_mlp *m = c->closure;
    const Geo &g = c->g;
    CK(hipStreamSynchronize(m->stream));
    ApplyArgs A;
    A.S = sample_args(c);
    for (int q = 0; q < 7; q++) A.pr[q] = pr[q];
    for (int q = 0; q < 4; q++) A.V[q] = c->cl_buf[3 + q];
    A.ntot = g.ntot;
    const long nblk = (g.ntot + 255) / 256;
    hipLaunchKernelGGL(k_closure_apply, dim3((unsigned)nblk), dim3(256), 0, c->stream, m->dev(), A);
    // d/dz of the vertical SGS fluxes: V -> (rx, ry, rz and cl_buf[7]) then S -= it
    double *out[4] = {c->cl_buf[0], c->cl_buf[1], c->cl_buf[2], c->cl_buf[7]};
    double *in3[3] = {A.V[0], A.V[1], A.V[2]};
    launch_dpdz3(g, in3, out, c->stream);
    launch_dpdz(g, A.V[3], out[3], c->stream);
    Sub4 sb;
    for (int q = 0; q < 4; q++) {
        sb.S[q] = S[q];
        sb.D[q] = out[q];
    }
    hipLaunchKernelGGL(k_sub4, dim3((unsigned)nblk), dim3(256), 0, c->stream, sb, g.ntot);
    return hipGetLastError() == hipSuccess ? 0 : iw_fail("closure launch failed");
}
