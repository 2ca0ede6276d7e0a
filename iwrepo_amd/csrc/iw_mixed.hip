// Mixed-precision kernels of the stratified-Boussinesq step (gfx950).
//
// The z contraction of the physical-space pass of rhsfm (solver.f:627-668,
// 1085-1093: the vertical flux derivatives, and the z gradients of u, v, w
// that solver.f:74-77 takes in spectral space) on FP32 matrix cores.  Per
// subdomain the derivative is the dense product (2/zh) D X of the GLL
// differentiation matrix D (nl x nl, derv.f:212-275) with the nl x ncols
// block X of columns; X is taken for w, (-w/2) w, u, (-u/2) w, v, (-v/2) w
// and (-rho) w in turn.
//
// v_mfma_f32_16x16x4f32: a 16x4 block of D times a 4x16 block of X added to
// a 16x16 FP32 accumulator.  Operand layout (lane l of 64): A[i][k] with
// i = l % 16, k = l / 16; B[k][j] with k = l / 16, j = l % 16; the result
// holds rows 4 (l / 16) + r, r = 0..3, of column l % 16.
#include <hip/hip_runtime.h>

#include "iw_internal.h"

namespace iw {

typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kMW = 64;        // columns per wavefront
constexpr int kMWaves = 4;     // wavefronts per block
constexpr int kMStride = 68;   // LDS row stride (floats): rows 4 apart fall on other banks

template <int NL>
struct MfmaShape {
    static constexpr int KS = (NL + 3) / 4;   // k steps of 4
    static constexpr int MB = (NL + 15) / 16; // 16-row blocks of the result
    static constexpr int ROWS = (4 * KS > 16 * MB) ? 4 * KS : 16 * MB;
};

struct ConvArgs32 {
    const float *u, *v, *w, *t;     // physical u, v, w, rho
    float *fu, *fv, *fw, *ft;       // d(-uw/2)/dz, d(-vw/2)/dz, d(-ww/2)/dz, d(-w rho)/dz
    float *du, *dv, *dw;            // du/dz, dv/dz, dw/dz
};

template <int NL>
__global__ __launch_bounds__(kMW *kMWaves) void k_dpdz_conv_mfma(ConvArgs32 A, long nxppy,
                                                                const double *__restrict__ D,
                                                                const double *__restrict__ zh)
{
    using S = MfmaShape<NL>;
    __shared__ float tile[kMWaves][S::ROWS * kMStride];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ks = blockIdx.y;
    const long col0 = ((long)blockIdx.x * kMWaves + wv) * kMW;
    const long col = col0 + lane;
    const bool live = col < nxppy;
    const long b0 = (long)ks * NL * nxppy + (live ? col : nxppy - 1);
    const float fac = (float)(2. / zh[ks]);
    float *T = tile[wv];
    // D in A-operand layout: a[mb][s] = D[16 mb + l % 16][4 s + l / 16]
    float a[S::MB][S::KS];
#pragma unroll
    for (int mb = 0; mb < S::MB; mb++)
#pragma unroll
        for (int s = 0; s < S::KS; s++) {
            const int i = 16 * mb + (lane & 15), k = 4 * s + (lane >> 4);
            a[mb][s] = (i < NL && k < NL) ? (float)D[i * NL + k] : 0.f;
        }
    // rows NL..ROWS-1 of the operand tile stay zero
    for (int e = lane; e < (S::ROWS - NL) * kMW; e += kMW) T[(NL + e / kMW) * kMStride + e % kMW] = 0.f;
    float w[NL], x[NL];
#pragma unroll
    for (int t = 0; t < NL; t++) w[t] = A.w[b0 + t * nxppy];
    // out = fac * D * col for the column each lane holds in x
    auto contract = [&](const float(&c)[NL], float *out) __attribute__((always_inline)) {
#pragma unroll
        for (int t = 0; t < NL; t++) T[t * kMStride + lane] = c[t];
        __syncthreads();
        float b[4][S::KS];
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int s = 0; s < S::KS; s++) b[cb][s] = T[(4 * s + (lane >> 4)) * kMStride + 16 * cb + (lane & 15)];
        __syncthreads();
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int mb = 0; mb < S::MB; mb++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < S::KS; s++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mb][s], b[cb][s], acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = 16 * mb + 4 * (lane >> 4) + r;
                    if (row < NL) T[row * kMStride + 16 * cb + (lane & 15)] = acc[r];
                }
            }
        __syncthreads();
        if (live)
#pragma unroll
            for (int t = 0; t < NL; t++) out[b0 + t * nxppy] = T[t * kMStride + lane] * fac;
        __syncthreads();
    };
    contract(w, A.dw);
#pragma unroll
    for (int t = 0; t < NL; t++) x[t] = (-0.5f * w[t]) * w[t];
    contract(x, A.fw);
    const float *src[3] = {A.u, A.v, A.t};
    float *dz[3] = {A.du, A.dv, nullptr}, *fl[3] = {A.fu, A.fv, A.ft};
#pragma unroll 1
    for (int q = 0; q < 3; q++) {
        const float cf = q < 2 ? -0.5f : -1.f;
#pragma unroll
        for (int t = 0; t < NL; t++) x[t] = src[q][b0 + t * nxppy];
        if (dz[q]) contract(x, dz[q]);
#pragma unroll
        for (int t = 0; t < NL; t++) x[t] = (cf * x[t]) * w[t];
        contract(x, fl[q]);
    }
}

} // namespace

bool launch_dpdz_conv_grad_mfma(const Geo &g, float *const *uvwt, float *const *out4, float *const *dz3,
                                hipStream_t st)
{
    ConvArgs32 A;
    A.u = uvwt[0]; A.v = uvwt[1]; A.w = uvwt[2]; A.t = uvwt[3];
    A.fu = out4[0]; A.fv = out4[1]; A.fw = out4[2]; A.ft = out4[3];
    A.du = dz3[0]; A.dv = dz3[1]; A.dw = dz3[2];
    const long per = (long)kMW * kMWaves;
    dim3 grid((unsigned)((g.nxppy + per - 1) / per), g.nsubd);
    dim3 blk(kMW * kMWaves);
#define MC(NL) hipLaunchKernelGGL(k_dpdz_conv_mfma<NL>, grid, blk, 0, st, A, g.nxppy, g.D, g.zh)
    switch (g.nzloc) {
        IW_NL_SWITCH(MC)
    default: return false;
    }
    return true;
#undef MC
}

} // namespace iw
