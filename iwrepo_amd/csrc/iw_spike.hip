// Subdomain-parallel solve of the u/v, w and rho systems (the systems factored
// without row interchanges) and the Legendre filter of the solutions that
// completes it, for MI355X (gfx950).  Layouts as in iw_kernels.hip.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "iw_internal.h"

namespace iw {

// ---------------------------------------------------------------------------
// Subdomain-parallel solve of the systems without interchanges (the same
// L U x = b as bandludcmp.f:1020-1030 and STBSV, ludcmp.f:4698-5046, split
// along the penalty coupling).  Partition L and U into the nl x nl diagonal
// blocks of the subdomains.  L couples block s to the next one only through
// row 0 of block s+1 (its multipliers l_s = L[(s+1,0), block s]); U couples
// block s to the next one only through its last row (u_s = U[(s,nl-1),
// block s+1]).  With y0_s = L_ss^-1 b_s and x0_s = U_ss^-1 y0_s:
//   y_s = y0_s - c_{s-1} g_s,           g_s = L_ss^-1 e_0,
//   x_s = x0_s - c_{s-1} k_s - d_s h_s, k_s = U_ss^-1 g_s, h_s = U_ss^-1 e_{nl-1},
//   c_s = alpha_s - beta_s c_{s-1},     alpha_s = l_s . y0_s, beta_s = l_s . g_s,
//   d_s = a_s - gamma_s c_s - delta_s d_{s+1},
//                                       a_s = u_s . x0_{s+1}, gamma_s = u_s . k_{s+1},
//                                       delta_s = u_s . h_{s+1},
// with c_{-1} = d_{nsub-1} = 0.  Two passes:
//   k_spike:        one wavefront per (lane block, subdomain): the local
//                   forward and backward sweeps of that block over the packed
//                   factor rows (grouped lanes as in k_sweep), x0_s written in
//                   place, alpha_s and a_{s-1} to per-column scratch;
//   k_spike_couple: one thread per column: the two scalar recurrences, c_{s-1}
//                   and d_s written over alpha_s and a_{s-1};
//   k_spike_finish: one lane per (column, subdomain):
//                   x_s = x0_s - c_{s-1} k_s - d_s h_s, then the Legendre
//                   filter and the mean-u damping that calcuvwt applies to
//                   every solution (filter_legenpen).
// k_spike_pre makes g, k, h, beta, gamma, delta per (system, subdomain) when
// the factors change.  Every block sweep runs the reference's operations
// (w + L (-y), w - x U, the division by the diagonal); the split changes only
// where the coupling terms enter, a rounding-level difference.
// SPK: per system and subdomain s (spk_stride doubles at (system nsub + s)
// spk_stride): (k[i], h[i]) pairs for i < nl, then beta, gamma, delta, 0.
// ---------------------------------------------------------------------------

template <int NL>
__global__ __launch_bounds__(64) void k_spike_pre(const double *__restrict__ FB, double *SPK, double *SPB, int first,
                                                  int count, int np, int nsub)
{
    constexpr int SF = tri_sf(NL), SB = tri_sb(NL, false), ST = spk_stride(NL);
    const long id = (long)blockIdx.x * 64 + threadIdx.x;
    if (id >= (long)count * nsub) return;
    const int sys = first + (int)(id / nsub), s = (int)(id % nsub);
    if (sys < np) return; // pressure systems: interchanges, solved by k_sweep
    const double *fw = FB + 2 * ((long)sys * nsub * (SF + SB) + (long)s * SF);
    const double *bw = FB + 2 * ((long)sys * nsub * (SF + SB) + (long)nsub * SF + (long)s * SB);
    double g[NL], k[NL], h[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) g[i] = i == 0 ? 1. : 0.;
    double beta = 0.;
#pragma unroll
    for (int r = 0; r < NL; r++) { // L[t] of row r at fw[2 off_f(r) + t - 1]
        const double *lr = fw + 2 * tri_off_f(NL, r);
        const double y = g[r];
#pragma unroll
        for (int t = 1; t < NL - r; t++) g[r + t] = g[r + t] + lr[t - 1] * (-y);
        beta = beta + lr[NL - r - 1] * y;
    }
#pragma unroll
    for (int i = 0; i < NL; i++) {
        k[i] = g[i];
        h[i] = i == NL - 1 ? 1. : 0.;
    }
    double gam = 0., del = 0.;
#pragma unroll
    for (int r = NL - 1; r >= 0; r--) { // U[t] of row r at bw[2 off_b(r) + t]
        const double *ur = bw + 2 * tri_off_b(NL, r, false);
        const double d = ur[0];
        k[r] = k[r] / d;
        h[r] = h[r] / d;
#pragma unroll
        for (int t = 1; t <= r; t++) {
            k[r - t] = k[r - t] - k[r] * ur[t];
            h[r - t] = h[r - t] - h[r] * ur[t];
        }
        gam = gam + k[r] * ur[r + 1];
        del = del + h[r] * ur[r + 1];
    }
    double *o = SPK + ((long)sys * nsub + s) * ST;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        o[2 * i] = k[i];
        o[2 * i + 1] = h[i];
    }
    o[2 * NL] = beta;
    o[2 * NL + 3] = 0.;
    if (s == nsub - 1) o[2 * NL + 1] = o[2 * NL + 2] = 0.;
    if (s > 0) { // gamma_{s-1}, delta_{s-1}: u_{s-1} is this block's coupling row
        o[2 * NL + 1 - ST] = gam;
        o[2 * NL + 2 - ST] = del;
    }
    // the same three per subdomain, 4 doubles apart, for the recurrences
    double *p = SPB + ((long)sys * nsub + s) * 4;
    p[0] = beta;
    p[3] = 0.;
    if (s == nsub - 1) p[1] = p[2] = 0.;
    if (s > 0) {
        p[1 - 4] = gam;
        p[2 - 4] = del;
    }
}

void launch_spike_pre(const Geo &g, const SysTables &S, int first, int count, hipStream_t st)
{
    if (count <= 0 || !S.SPK) return;
    const long nt = (long)count * g.nsubd;
    const dim3 grid((unsigned)((nt + 63) / 64));
#define SP(NL) \
    hipLaunchKernelGGL(k_spike_pre<NL>, grid, dim3(64), 0, st, S.FB, S.SPK, S.SPB, first, count, S.np, g.nsubd)
    switch (g.nzloc) {
        IW_NL_SWITCH(SP)
    default: break;
    }
#undef SP
}

struct SpikeArgs {
    double *f[4];
    const int *tasks;
    const int *modesys;
    const double *FB;
    const double *sysinfo;
    double *al, *au; // [field][subdomain][column]: alpha_s, a_{s-1}
    double den[4];
    long nxppy;
    int nsub, fld_base;
    int ks0; // subdomain of grid row 0
};

// rows of the 2 nl sweep steps of one block: forward rows 0..nl-1, then
// backward rows nl-1..0
template <int NL>
constexpr int spike_np(int i)
{
    return i < NL ? tri_npf(NL, i) : tri_npb(NL, 2 * NL - 1 - i, false);
}

// UVW: the u, v and w columns of a (kx, +-ky) pair in one group of G = 12
// lanes (lane table entry: column | field << 28 | idle << 30), all solved
// with the u/v system's factor blocks; used for the subdomains below the
// top one when those blocks of the u/v and w systems are bit for bit the same
// (a no-slip bottom: the two differ only in the top boundary condition)
template <int NL, int G, int DP, bool UVW = false>
__global__ __launch_bounds__(64) void k_spike(SpikeArgs A)
{
    static_assert(DP >= 2 && DP <= 2 * NL, "a row is staged one step before use, from a slot filled earlier");
    constexpr int NG = (64 + G - 1) / G, NS = 2 * NL;
    constexpr int SF = tri_sf(NL), SB = tri_sb(NL, false);
    constexpr int NPM = tri_npf(NL, 0) > tri_npb(NL, NL - 1, false) ? tri_npf(NL, 0) : tri_npb(NL, NL - 1, false);
    constexpr int PP = (NPM + G - 1) / G; // pieces per lane
    __shared__ double2 seg[2][NG][NPM + 1];
    const int lane = threadIdx.x, grp = lane / G, rk = lane % G;
    const int e = A.tasks[blockIdx.x * 64 + lane];
    if (e < 0) return; // the whole group is empty
    const bool idle = UVW ? (e >> 30) & 1 : (e >> 29) & 1;
    const int col = e & ((1 << 28) - 1);
    const int m = col >> 1;
    const int fld = UVW ? (e >> 28) & 3 : A.fld_base + blockIdx.z + ((e >> 28) & 1);
    // w(0,0) = 0 (calcuvwt.f:496-503): its group holds only this mode (in
    // the u/v/w groups its lanes are idle)
    if (!UVW && fld == 2 && m == 0) return;
    const int sys = A.modesys[4 * m + (UVW || fld < 2 ? 1 : fld == 2 ? 2 : 3)];
    const double eps = A.sysinfo[4 * sys], den = A.den[fld];
    const int ks = A.ks0 + blockIdx.y, nsub = A.nsub;
    const long st = A.nxppy;
    const double2 *fw = reinterpret_cast<const double2 *>(A.FB) + (long)sys * nsub * (SF + SB) + (long)ks * SF;
    const double2 *bw = reinterpret_cast<const double2 *>(A.FB) + (long)sys * nsub * (SF + SB) + (long)nsub * SF +
                        (long)ks * SB;
    double *b = A.f[fld] + col + (long)ks * NL * st;
    auto ldp = [&](auto Ic, double(&dx)[PP], double(&dy)[PP]) __attribute__((always_inline)) {
        constexpr int I = decltype(Ic)::value, NP = spike_np<NL>(I);
        const double2 *src = I < NL ? fw + tri_off_f(NL, I) : bw + tri_off_b(NL, NS - 1 - I, false);
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const int p = rk + q * G;
            if (q * G < NP) {
                const double2 v = src[p < NP ? p : 0];
                dx[q] = v.x;
                dy[q] = v.y;
            }
        }
    };
    auto stp = [&](auto Ic, int buf, const double(&sx)[PP], const double(&sy)[PP]) __attribute__((always_inline)) {
        constexpr int I = decltype(Ic)::value, NP = spike_np<NL>(I);
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const int p = rk + q * G;
            if (q * G < NP && p < NP) seg[buf][grp][p] = make_double2(sx[q], sy[q]);
        }
    };
    // right-hand side of the viscous systems: -(-f / den) eps (calcuvwt.f:390-402)
    double w[NL + 1];
#pragma unroll
    for (int t = 0; t < NL; t++) w[t] = -(-b[(long)t * st] / den) * eps;
    w[NL] = 0.; // row 0 of the next block: accumulates -alpha_s
    double rx[DP][PP], ry[DP][PP];
    static_for<0, DP>([&](auto dc) __attribute__((always_inline)) {
        constexpr int d = decltype(dc)::value;
        ldp(IC<d>{}, rx[d], ry[d]);
    });
    stp(IC<0>{}, 0, rx[0], ry[0]);
    double acc = 0.; // last row of the previous block: accumulates -a_{s-1}
    static_for<0, NS>([&](auto ic) __attribute__((always_inline)) {
        constexpr int i = decltype(ic)::value, d = i % DP, NP = spike_np<NL>(i);
        // one row step at a time: without this the scheduler spreads the
        // loads of later steps over the whole unrolled block and doubles the
        // registers
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (i + 1 < NS) stp(IC<i + 1>{}, (i + 1) & 1, rx[(d + 1) % DP], ry[(d + 1) % DP]);
        if constexpr (i + DP < NS) ldp(IC<i + DP>{}, rx[d], ry[d]);
        double fc[2 * NP];
        // the other lanes' pieces were stored one step earlier by this
        // wavefront: order those stores before these reads
        wave_sync();
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const double2 v = seg[i & 1][grp][p];
            fc[2 * p] = v.x;
            fc[2 * p + 1] = v.y;
        }
        if constexpr (i < NL) { // forward row r: fc[t - 1] = L[t], t = 1..nl - r
            constexpr int r = i, EF = tri_ef(NL, r);
            const double y = w[r];
#pragma unroll
            for (int t = 1; t <= EF; t++) {
                w[r + t] = w[r + t] + fc[t - 1] * (-y);
                pin(w[r + t]);
            }
        } else { // backward row r: fc[t] = U[t], t = 0..r + 1
            constexpr int r = NS - 1 - i;
            const double x = (w[r] != 0.) ? w[r] / fc[0] : w[r];
#pragma unroll
            for (int t = 1; t <= r; t++) {
                w[r - t] = w[r - t] - x * fc[t];
                pin(w[r - t]);
            }
            acc = acc - x * fc[r + 1];
            pin(acc);
            if (!idle) b[(long)r * st] = x;
        }
    });
    if (!idle) {
        const long o = ((long)fld * nsub + ks) * st + col;
        A.al[o] = -w[NL];
        A.au[o] = -acc;
    }
}

// ring depth of k_spike (rows fetched ahead): measured at nl = 25, 8 for the
// u/v blocks (G = 8, two pieces per lane) and 5 for w and rho (G = 4, four
// pieces per lane); deeper rings or a register cap that allows a third
// wavefront per SIMD spill and run slower.  The u/v/w groups (G = 12, two
// pieces per lane) take the depth of the u/v ones.
template <int NL, int G>
constexpr int spike_dp()
{
    return (G >= 8 ? 8 : 5) < 2 * NL ? (G >= 8 ? 8 : 5) : 2 * NL;
}

void launch_solve_spike(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                        hipStream_t st, int part)
{
    SpikeArgs A;
    for (int q = 0; q < 4; q++) A.f[q] = uvwt[q];
    A.modesys = S.modesys;
    A.FB = S.FB;
    A.sysinfo = S.sysinfo;
    A.al = S.spk_al;
    A.au = S.spk_au;
    A.den[0] = A.den[1] = A.den[2] = P.xnu * dt;
    A.den[3] = dt * P.xkappa;
    A.nxppy = g.nxppy;
    A.nsub = g.nsubd;
    A.ks0 = 0;
    SpikeArgs B = A;
    A.tasks = S.guv_tasks;
    A.fld_base = 0;
    B.tasks = S.gwt_tasks;
    B.fld_base = 2;
    // subdomain-major grid: the blocks of one subdomain run together, and
    // block x keeps the XCD (x mod 8) the task placement chose for it
    if (S.share_uvw) {
        // below the top subdomain u, v and w in the u/v/w groups over the
        // u/v factor blocks; the top subdomain of u/v and of w and every
        // subdomain of rho in their own groups
        SpikeArgs C = A, W = B, T = B;
        C.tasks = S.guvw_tasks;
        A.ks0 = W.ks0 = g.nsubd - 1;
        T.fld_base = 3;
        dim3 gc(S.guvw_ntasks / 64, g.nsubd - 1, 1), ga(S.guv_ntasks / 64, 1, 1), gw(S.gwt_ntasks / 64, 1, 1),
            gt(S.gwt_ntasks / 64, g.nsubd, 1);
#define GS3(NL)                                                                                                  \
    if (part != 2) {                                                                                             \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 12, spike_dp<NL, 12>(), true>), gc, dim3(64), 0, st, C); \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 8, spike_dp<NL, 8>()>), ga, dim3(64), 0, st, A);         \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 4, spike_dp<NL, 4>()>), gw, dim3(64), 0, st, W);         \
    }                                                                                                            \
    if (part != 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 4, spike_dp<NL, 4>()>), gt, dim3(64), 0, st, T)
        switch (g.nzloc) {
            IW_NL_SWITCH(GS3)
        default: break;
        }
#undef GS3
        return;
    }
    if (part == 2) return; // without the u/v/w groups rho is solved with w
    dim3 ga(S.guv_ntasks / 64, g.nsubd, 1), gb(S.gwt_ntasks / 64, g.nsubd, 2);
#define GS(NL)                                                                                          \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 8, spike_dp<NL, 8>()>), ga, dim3(64), 0, st, A); \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spike<NL, 4, spike_dp<NL, 4>()>), gb, dim3(64), 0, st, B)
    switch (g.nzloc) {
        IW_NL_SWITCH(GS)
    default: break;
    }
#undef GS
}

// Bit-for-bit comparison of the u/v and w factor blocks below the top
// subdomain, forward and backward, for every mode (one thread per packed
// double2); any difference sets *flag.  Run after each packing when the
// bottom condition is the same for both systems.
__global__ __launch_bounds__(256) void k_share_check(const double *__restrict__ FB, const int *__restrict__ modesys,
                                                     long nmode, int nsub, int sf, int sb, int *flag)
{
    const long per = (long)(nsub - 1) * (sf + sb);
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (nmode - 1) * per) return;
    const long m = 1 + id / per, q = id - (m - 1) * per;
    const long off = q < (long)(nsub - 1) * sf ? q : (long)nsub * sf + (q - (long)(nsub - 1) * sf);
    const long su = modesys[4 * m + 1], sw = modesys[4 * m + 2];
    const double2 a = reinterpret_cast<const double2 *>(FB)[su * nsub * (sf + sb) + off];
    const double2 b = reinterpret_cast<const double2 *>(FB)[sw * nsub * (sf + sb) + off];
    if (__double_as_longlong(a.x) != __double_as_longlong(b.x) || __double_as_longlong(a.y) != __double_as_longlong(b.y))
        atomicOr(flag, 1);
}

void launch_share_check(const Geo &g, const SysTables &S, int *flag, hipStream_t st)
{
    const long nmode = (long)g.nxhp * g.ny;
    const int sf = tri_sf(g.nzloc), sb = tri_sb(g.nzloc, false);
    const long n = (nmode - 1) * (long)(g.nsubd - 1) * (sf + sb);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_share_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S.FB, S.modesys, nmode,
                       g.nsubd, sf, sb, flag);
}

// the two scalar recurrences per column (one thread per column and field);
// SPB holds beta, gamma, delta of (system, subdomain) at (system nsub + s) 4
__global__ __launch_bounds__(256) void k_spike_couple(double *al, double *au, const int *__restrict__ modesys,
                                                      const double *__restrict__ SPB, long nxppy, int nsub,
                                                      int fld0)
{
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int fld = fld0 + (int)blockIdx.y;
    if (col >= nxppy) return;
    const int m = (int)(col >> 1);
    if (fld == 2 && m == 0) return;
    const int sys = modesys[4 * m + (fld < 2 ? 1 : fld == 2 ? 2 : 3)];
    const double4 *bgd = reinterpret_cast<const double4 *>(SPB) + (long)sys * nsub;
    double *a1 = al + (long)fld * nsub * nxppy + col, *a2 = au + (long)fld * nsub * nxppy + col;
    // c_s = alpha_s - beta_s c_{s-1}; slot s receives c_{s-1}
    double c = 0.;
#pragma unroll 4
    for (int s = 0; s < nsub; s++) {
        const double alpha = a1[(long)s * nxppy];
        a1[(long)s * nxppy] = c;
        c = alpha - bgd[s].x * c;
    }
    // d_s = a_s - gamma_s c_s - delta_s d_{s+1}, a_s = a_{(s+1)-1} from block
    // s+1, c_s in slot s+1; slot s receives d_s
    double d = 0., anext = 0.;
#pragma unroll 4
    for (int s = nsub - 1; s >= 0; s--) {
        const double ap = a2[(long)s * nxppy];
        if (s < nsub - 1) {
            const double4 q = bgd[s];
            d = anext - q.y * a1[(long)(s + 1) * nxppy] - q.z * d;
        }
        a2[(long)s * nxppy] = d;
        anext = ap;
    }
}

// The coupling correction and the Legendre filter of every subdomain of the
// four solutions (filter_legenpen, filtering.f:729-808) with the mean-u
// damping (calcuvwt.f:343-361), in place on the native layout: lane = one
// column (re or im of a mode) in one subdomain, consecutive lanes consecutive
// columns, the filter matrix in LDS; the filter is applied one output row at
// a time from the inputs in registers (the sums in the order of
// legfilter_mat).
struct FinishArgs {
    double *f[4];
    const int *modesys;
    const double *SPK;
    const double *al, *au; // c_{s-1}, d_s
    const double *FL, *damp;
    long nxppy;
    int nsub, cpen;
    int share_uvw; // w's blocks below the top subdomain are the u/v ones: read their k, h
    int fld0;      // field of the first wavefront
    int nf;        // fields of the launch; with one, a work-group covers blockDim.x columns of it
};

template <int NL>
__global__ __launch_bounds__(256) void k_spike_finish(FinishArgs A)
{
    __shared__ double FLs[NL * NL];
    stage_fl<NL>(FLs, A.FL);
    // wavefront f = field fld0 + f of the work-group's 64 columns: u, v (and
    // w below the top subdomain with the u/v/w groups) read the same k, h,
    // once from HBM for the work-group; a one-field launch covers blockDim.x
    // columns of that field per work-group
    const bool one = A.nf == 1;
    const int fld = one ? A.fld0 : A.fld0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long col = one ? (long)blockIdx.x * blockDim.x + threadIdx.x : (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const int s = blockIdx.y, nsub = A.nsub;
    const long st = A.nxppy;
    if (col >= st) return;
    const long m = col >> 1;
    if (fld == 2 && m == 0) return;
    const bool uvk = fld < 2 || (fld == 2 && A.share_uvw && s < nsub - 1);
    const int sys = A.modesys[4 * m + (uvk ? 1 : fld == 2 ? 2 : 3)];
    const double2 *kh = reinterpret_cast<const double2 *>(A.SPK + ((long)sys * nsub + s) * spk_stride(NL));
    const long o = ((long)fld * nsub + s) * st + col;
    const double c = A.al[o], d = A.au[o];
    double *fs = A.f[fld] + col + (long)s * NL * st;
    double x[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const double2 v = kh[k];
        x[k] = fs[(long)k * st] - c * v.x - d * v.y;
    }
    const bool damp = fld == 0 && m == 0;
#pragma unroll
    for (int q = 0; q < NL; q++) {
        double r = x[q];
        if (A.cpen) {
            r = 0.;
#pragma unroll
            for (int t = 0; t < NL; t++) r = r + FLs[t * NL + q] * x[t];
        }
        if (damp) r = A.damp[s * NL + q] * r;
        fs[(long)q * st] = r;
    }
}

// part 0: the four fields; 1: u, v, w; 2: rho (beside the pressure step on
// a second stream, after the rho solve of launch_solve_spike part 2)
void launch_spike_finish(const Geo &g, const SysTables &S, double *const *uvwt, int cpen, hipStream_t st, int part)
{
    const int f0 = part == 2 ? 3 : 0, nf = part == 0 ? 4 : part == 1 ? 3 : 1;
    hipLaunchKernelGGL(k_spike_couple, dim3((unsigned)((g.nxppy + 255) / 256), nf), dim3(256), 0, st, S.spk_al,
                       S.spk_au, S.modesys, S.SPB, g.nxppy, g.nsubd, f0);
    FinishArgs A;
    for (int q = 0; q < 4; q++) A.f[q] = uvwt[q];
    A.modesys = S.modesys;
    A.SPK = S.SPK;
    A.al = S.spk_al;
    A.au = S.spk_au;
    A.FL = g.FL;
    A.damp = g.damp;
    A.nxppy = g.nxppy;
    A.nsub = g.nsubd;
    A.cpen = cpen;
    A.share_uvw = S.share_uvw ? 1 : 0;
    A.fld0 = f0;
    A.nf = nf;
    const int cpw = nf == 1 ? 256 : 64; // columns per work-group
    dim3 grid((unsigned)((g.nxppy + cpw - 1) / cpw), g.nsubd);
#define FN(NL) hipLaunchKernelGGL(k_spike_finish<NL>, grid, dim3(nf == 1 ? 256 : 64 * nf), 0, st, A)
    switch (g.nzloc) {
        IW_NL_SWITCH(FN)
    default: break;
    }
#undef FN
}

// Legendre filter of every subdomain of the four solutions (filter_legenpen,
// filtering.f:729-808) and the mean-u damping (calcuvwt.f:343-361) after the
// sequential sweeps (k_sweep), in place on the native layout: lane = the re
// and im columns of one mode, the nl values of its subdomain in registers,
// the filter matrix in LDS.
template <int NL>
__global__ __launch_bounds__(256) void k_lfilter(double *u, double *v, double *w, double *t, long nxppy,
                                                int cpen, const double *__restrict__ FL,
                                                const double *__restrict__ damp)
{
    __shared__ double FLs[NL * NL];
    stage_fl<NL>(FLs, FL);
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x; // the re/im column pair of one mode
    const int ks = blockIdx.y, fld = blockIdx.z;
    if (m >= nxppy / 2) return;
    if (fld == 2 && m == 0) return;
    double *f = (fld == 0 ? u : fld == 1 ? v : fld == 2 ? w : t) + (long)ks * NL * nxppy + 2 * m;
    double xr[NL], xi[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const double2 q = *reinterpret_cast<const double2 *>(f + (long)k * nxppy);
        xr[k] = q.x;
        xi[k] = q.y;
    }
    if (cpen) legfilter_mat2<NL>(xr, xi, FLs);
    if (fld == 0 && m == 0)
#pragma unroll
        for (int k = 0; k < NL; k++) {
            xr[k] = damp[ks * NL + k] * xr[k];
            xi[k] = damp[ks * NL + k] * xi[k];
        }
#pragma unroll
    for (int k = 0; k < NL; k++) *reinterpret_cast<double2 *>(f + (long)k * nxppy) = make_double2(xr[k], xi[k]);
}

void launch_lfilter(const Geo &g, double *const *uvwt, int cpen, hipStream_t st)
{
    dim3 grid((unsigned)((g.nxppy / 2 + 255) / 256), g.nsubd, 4);
#define LF(NL)                                                                                             \
    hipLaunchKernelGGL(k_lfilter<NL>, grid, dim3(256), 0, st, uvwt[0], uvwt[1], uvwt[2], uvwt[3], g.nxppy, \
                       cpen, g.FL, g.damp)
    switch (g.nzloc) {
        IW_NL_SWITCH(LF)
    default: break;
    }
#undef LF
}

} // namespace iw
