// Internal declarations shared by the kernels and the host driver of libiwgpu.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <vector>

#include "iwgpu.h"

namespace iw {

constexpr int kMaxNl = 32;     // GLL points per subdomain supported by the kernels
// GLL points per subdomain (orders 8..24) with kernels specialised at
// compile time; other orders take the generic kernels
#define IW_NL_SWITCH(M)                                                                                       \
    case 9: M(9); break;                                                                                       \
    case 10: M(10); break;                                                                                     \
    case 11: M(11); break;                                                                                     \
    case 12: M(12); break;                                                                                     \
    case 13: M(13); break;                                                                                     \
    case 14: M(14); break;                                                                                     \
    case 15: M(15); break;                                                                                     \
    case 16: M(16); break;                                                                                     \
    case 17: M(17); break;                                                                                     \
    case 18: M(18); break;                                                                                     \
    case 19: M(19); break;                                                                                     \
    case 20: M(20); break;                                                                                     \
    case 21: M(21); break;                                                                                     \
    case 22: M(22); break;                                                                                     \
    case 23: M(23); break;                                                                                     \
    case 24: M(24); break;                                                                                     \
    case 25: M(25); break;
constexpr bool nl_specialised(int nl) { return nl >= 9 && nl <= 25; }
constexpr int kMaxFields = 16; // fields per batched transform launch

// Field pointer batch passed by value to batched kernels: source and
// destination of each field (equal for an in-place transform) and optional
// plane maps (logical plane -> stored plane index; null = identity).
struct FieldBatch {
    double *f[kMaxFields];
    double *g[kMaxFields];
    // inverse transforms only: optional second outputs, the transform of
    // i ky f (y pass) or i kx f (x pass), written with the plane map of g
    double *dy[kMaxFields] = {};
    double *dx[kMaxFields] = {};
    // forward transforms only: optional second inputs P whose transform is
    // added as i kx P (x pass) or i ky P (y pass) to the field's own (the
    // horizontal flux terms of rhsfm); read with the plane map of f
    const double *ax[kMaxFields] = {};
    const double *ay[kMaxFields] = {};
    // forward x pass only: where to store the plain transform of ax[q] as
    // well (the y pass needs it for another field), or null
    double *axo[kMaxFields] = {};
    const double *kw = nullptr; // ky (y pass) or kx (x pass) per row / column
    int n = 0;
    const int *smap = nullptr, *dmap = nullptr;
};

// Problem geometry and device-resident tables.
struct Geo {
    int nx, ny, nz, nzloc, nsubd, nxpp, nxh, nxhp, nyh, ndims;
    long nxppy, ntot;
    int M, logM, logNy; // x complex length nx/2 and logs (powers of two)
    bool p2;            // M and ny powers of two: shift-and-mask transforms
    unsigned planM, planY; // radices of the passes for M and ny, 4 bits each, first pass lowest
    double scaley;
    int ft = 256; // threads per transform block (64, 128 or 256), 8 values each
    // device tables
    double *D, *D2, *AL, *wg, *sig, *xw, *yw, *zh, *z, *damp, *fcf, *x;
    double *sigbv; // Boyd-Vandeven filter of the estimation model (filtering.f:1272-1310)
    double *FL;    // the Legendre filter as one nl x nl matrix, out[q] = sum_t FL[t nl + q] x[t]
    double2 *twM;   // e^{-2 pi i t/M}, t < M
    double2 *twN;   // e^{-2 pi i k/nx}, k <= M
    double2 *twY;   // e^{-2 pi i t/ny}, t < ny
};

struct SysTables {
    int nsys = 0;
    int np = 0; // pressure systems occupy ids 0..np-1
    int ldab = 0, kl = 0, kv = 0;
    double *AB = nullptr;   // nsys * ldab * nz, LAPACK band columns
    // the structurally nonzero parts of each factor column, packed per
    // system (tri_* layout): nsubd forward blocks, then nsubd backward
    // blocks, so both sweeps stream contiguous memory; FB for the systems
    // without interchanges, PB for the pressure systems (ids < np)
    double *FB = nullptr;
    double *PB = nullptr;
    // subdomain-parallel solve of the FB systems: per (system, subdomain)
    // k[nl], h[nl], beta, gamma, delta, 0 (spk_stride doubles); per (field,
    // subdomain, column) the coupling scalars alpha_s (then c_{s-1}) and
    // a_{s-1}; SPB: beta, gamma, delta, 0 per (system, subdomain) again,
    // packed for the recurrences
    double *SPK = nullptr, *SPB = nullptr;
    double *spk_al = nullptr, *spk_au = nullptr;
    // set by the packing when a factor breaks the tri_* pattern: [0] the
    // pressure factors, [1] the others
    int *tri_bad = nullptr;
    bool tri_ok_p = false, tri_ok_v = false;
    size_t cap_p = 0;
    int *ipiv = nullptr;    // nsys * nz (0-based)
    double *coef = nullptr; // nsys * nsubd * 12 assembly coefficients
    double *sysinfo = nullptr; // nsys * 4: eps, tau1(bottom), tau2(top), bc flags
    int *modesys = nullptr; // nmode * 4: P, UV, W, T system ids (-1: none)
    // block -> task tables: tasks that share a factor are placed on one XCD
    // (blocks b, b+8, b+16, ...) next to each other so the factor streams
    // from that XCD's L2 once for the whole group
    int *order_p = nullptr, *order_v = nullptr;
    int nblk_p = 0, nblk_v = 0;
    // lane tables of the packed-factor sweeps (lane = one column): 8 lanes
    // per u/v system, 4 per w/rho (and pressure) system
    int *guv_tasks = nullptr;
    int guv_ntasks = 0;
    int *gwt_tasks = nullptr;
    int gwt_ntasks = 0;
    // u, v and w of a (kx, +-ky) pair in one group of 12 lanes (no-slip
    // bottom only), and whether their factor blocks below the top subdomain
    // are the same bits for the u/v and the w system (checked at each packing)
    int *guvw_tasks = nullptr;
    int guvw_ntasks = 0;
    bool share_uvw = false;
    int nsys_w = 0; // systems with Dirichlet ends (w), among the nsys
    bool np_ok = false;
    size_t cap_sys = 0;
};

// maxcalc reductions (various.f:749-851): 11 (value, position) quantities,
// the record layout of iw_maxcalc (iwgpu.h)
constexpr int kMcQ = 11;
constexpr int kMcRec = 42;
constexpr int kMcSlices = 8; // z slices of the column maxima of the umaxt walk
// The reductions run in the epilogue of the inverse x transform of u,v,w,
// temp (fields 0..3 of its batch).  When that transform is convec's
// (solver.f:107-123) rather than maxcalc's own, its input differs from
// maxcalc's by the round trip (kx-Nyquist column scaled by 1/ny) and the
// zeroed mean w (solver.f:66-69); maxcalc's values are then recovered as
// f + (ny-1) A_M(j,k) (-1)^i (+ the removed mean for w), A_M the kx-Nyquist
// coefficient of the row.
struct McFuse {
    double *part = nullptr;      // [blocks][kMcQ] (value, position) pairs; nullptr: off
    const double *rm = nullptr;  // nz: rhomean(z(k)) (forcing.f:134-141)
    const double *rdz = nullptr; // nz: 1/dz of the CFLZ (various.f:818-822)
    const double *w00 = nullptr; // nz: mean w removed before the transform, or nullptr
    double dt = 0, cx = 0, cy = 0; // dt/dx, dt/dy: the CFL numbers are |u| dt/dx ... to
                                   // within an ulp of the reference's (u dt)/dx
    double cny = 0;              // ny-1 for the Nyquist correction, 0: none
    int fid0 = 0;                // field index (u, v, w, temp = 0..3) of the batch's first field
    // mean-wave u, w, rho (nxpp,nz) and the shear current (nz), or nullptr
    const double *umwv = nullptr, *wmwv = nullptr, *rhomwv = nullptr, *ush = nullptr;
};
// what the record kernels read after the transform
struct McArgs {
    const double *u;             // physical u as transformed (convec's or maxcalc's)
    const float *u32 = nullptr;  // the same in FP32 (mixed path), read instead of u when set
    const double *spec;          // spectral u (kx-Nyquist coefficients), used when cny != 0
    double cny;
    const double *dz;
    int nx, ny, nz, nxpp;
    long nxppy;
    const double *part;
    int nblk;
    const double *umwv;          // mean-wave u added for the umaxt chain, or nullptr
    double *part2 = nullptr;     // kMcPre partial slots: the partials reduced in kMcPre groups first
};
constexpr int kMcPre = 64;

struct Prof {
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<int> n;
    std::vector<double> bytes;
};

} // namespace iw

struct iw_solver {
    iw_params P;
    iw::Geo g;
    hipStream_t stream = nullptr;
    int device = 0;
    // state: rotating index tables into buf
    double *buf[24];   // u v w t | un.. | unm.. | S(4) | Sn(4) | Snm(4)
    double *grad[9];   // dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz
    double *prod[7];   // -uu/2 -uv/2 -vv/2 -uw/2 -vw/2 -uT -vT
    double *ox;
    double *wnl, *wviscrot;
    double *hbuf;      // device scratch for host-array entry points
    double *phys[4];      // physical u,v,w,rho during convec
    double *sgs[9] = {};  // SGS tensors t11 t12 t13 t22 t23 t33 th1 th2 th3 (allocated on first use)
    double *tmean = nullptr; // mean of rho per level
    double *icomp;        // 4 planes per interface x 4 fields (interface averaging)
    int *iface_src, *iface_mid_src, *iface_mid_dst; // plane maps
    int *d_count;      // 4 interface counters
    int *h_count;      // pinned
    // time stepping
    double t, fac[3], bdf[3], bdfv;
    int istep;
    int istart = 0;   // 1 after iw_load_restart: first-step coefficients of a restart
    double cdt = 1.;  // cdtin of the restart (main.f:619-634)
    double factored_bdfv = -1, factored_dt = -1;
    bool p_factored = false;
    iw::SysTables sys;
    // host copies of tables
    std::vector<double> zpts, wg, d, d2, alpol, xw, xsq, yw, ysq, z;
    // maxcalc: tables, partials, the record (device and pinned host), the
    // records of the last iw_run, Tso and /vardtflags/ (comtime.h:3)
    double *mc_rm = nullptr, *mc_dz = nullptr, *mc_rdz = nullptr, *mc_part = nullptr, *mc_rec = nullptr;
    double *mc_colmax = nullptr, *mc_gmax = nullptr, *h_mc = nullptr, *mc_w00 = nullptr;
    double *mc_ring = nullptr;
    int mc_ring_cap = 0;
    double mc_dx = 0, mc_dy = 0, xomega = 0;
    std::vector<double> mc_hist;
    double Tso = 0.;
    int ivardt = 0, idtstep = 0;
    // /sol2dvar/ mean-wave arrays (comsoliton.h:9-13), 9 x (nxpp,nz) in
    // COMMON order, and the shear-current profiles uo fcshear fshear(z),
    // uo fcshear dfsheardz(z) (forcing.f:190-214); with either active the
    // terms of solver.f:848-931 need the physical dtdx, dtdz (gT)
    double *mw = nullptr, *ush = nullptr, *dush = nullptr;
    bool mw_on = false;
    // mixed path (iw_set_mixed): FP32 physical-space working set in convec
    bool mixed = false;
    // mixed path: the z contraction on the matrix cores (k_dpdz_conv_mfma)
    // followed by the products, instead of the fused pass: IW_MIXED_MFMA=1
    // at iw_create
    bool mixed_mfma = false;
    // the mixed path's right-hand-side histories S, Sn, Snm of u, v, w, temp
    // in FP32 (the products they come from are FP32): filled from buf[16..23]
    // when a mixed run starts and written back to buf[12..23] when it ends,
    // so that between runs the state is the FP64 one
    float *s32[12] = {};
    // inside iw_run's step loop: convec's SS/BDF pass also forms the pressure
    // right-hand side (phi_ready) for the calcuvwt of the same step
    bool fuse_prhs = false, phi_ready = false;
    // the u/v, w and rho solves by the sequential band sweeps (k_sweep)
    // instead of the subdomain-parallel solve: IW_SWEEPS=seq at iw_create
    bool seq_sweeps = false;
    // learned closure (iw_set_closure): the network and 8 scratch fields
    // (rho gradients x, y, z; the four vertical SGS fluxes; one more)
    struct iw_mlp *closure = nullptr;
    double *cl_buf[8] = {};
    double *gT[2] = {nullptr, nullptr};
    // postp (postp.f:9-76) and speclegencalc (filtering.f:1110-1221): the
    // x-means of u and of t11..t33 at every (y,z) line, d umean/dy and /dz,
    // budget block partials and one record; the records of the last iw_run
    // (device rings, host copies: istep, t, 16 values / istep, t, 4 spectra)
    double *pp_um = nullptr, *pp_tm = nullptr, *pp_dy = nullptr, *pp_dz = nullptr, *pp_part = nullptr;
    double *pp_rec = nullptr;
    double *kp = nullptr; // KE/PE: 4 nz level integrals and the 4 totals
    double *e_ring = nullptr, *l_ring = nullptr;
    int e_cap = 0, l_cap = 0;
    std::vector<double> e_hist, l_hist;
    // the single-column penalty solve (iw_penalty_solve): coefficients,
    // system table, band and solution; pivots (allocated on first use)
    double *pen_buf = nullptr;
    int *pen_piv = nullptr;
    // profiling
    bool profiling = false;
    iw::Prof prof;
    hipEvent_t ev0, ev1;
    // second stream (default priority) for the rho solve beside the pressure
    // step, and the events that fork it from and join it to the solver stream
    hipStream_t stream2 = nullptr;
    // set while the solver is a member of an ensemble: every launch stays on
    // the solver stream (the ensemble's chains are the concurrency, and the
    // extra branches of a second stream per member slow its graph down)
    bool one_stream = false;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    long long dev_bytes = 0;
};

// host helpers of iw_api.hip shared with iw_closure.hip: record a failure
// for iw_last_error (returns -1); transform the fields f (spectral, is = 1,
// or physical, is = -1) into g (may equal f) with horfft's conventions
int iw_fail(const std::string &m);
void iw_fft_batch(iw_solver *s, const std::vector<double *> &f, const std::vector<double *> &g, int is);
// the closure's term in convec (iw_closure.hip)
int iw_closure_apply(iw_solver *c, double *const *S, double *const *pr);

namespace iw {
// kernels launchers (iw_kernels.hip)
// blocks of the inverse x transform (maxcalc partials); ft > 0: for that
// block size instead of g.ft
int mc_blocks(const Geo &g, int ft = 0);
void launch_maxcalc_final(const Geo &g, const McArgs &A, int istep, double t, double umax, double *rec,
                          hipStream_t st);
void launch_maxcalc_umaxt(const Geo &g, const McArgs &A, double umax, double *colmax, double *groupmax,
                          double *rec, hipStream_t st);
// storage of the physical-space side of a transform: kPrecF64 everything
// FP64; kPrecF32 (mixed path) x pass FP32 in and out, y pass FP64 spectral
// in and FP32 out (inverse) or FP32 in and FP64 spectral out (forward)
constexpr int kPrecF64 = 0, kPrecF32 = 1;
// forward y pass only: FP32 in and FP32 spectral out (the mixed path's
// right-hand sides)
constexpr int kPrecF32S = 2;
void launch_fft_x(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st, const McFuse *mf = nullptr,
                  int prec = kPrecF64);
void launch_fft_y(const Geo &g, FieldBatch b, int nk, int dir, hipStream_t st, int prec = kPrecF64);
void launch_dpdz(const Geo &g, const double *a, double *b, hipStream_t st);
void launch_dpdz_conv(const Geo &g, double *const *uvwt, double *const *out4, hipStream_t st);
void launch_dpdz3(const Geo &g, double *const *in3, double *const *out3, hipStream_t st);
// launch_dpdz_conv's four flux derivatives plus du/dz, dv/dz, dw/dz (and
// drho/dz when dz4[3] is not null) of the same physical fields in one pass
void launch_dpdz_conv_grad(const Geo &g, double *const *uvwt, double *const *out4, double *const *dz4,
                           hipStream_t st);
void launch_grad(const Geo &g, double *const *uvw, double *const *gr, hipStream_t st);
// mean-wave and shear-current terms of rhsfm (solver.f:848-931); mw
// nullptr: none
struct MeanTerms {
    const double *mw = nullptr;          // 9 x (nxpp,nz), COMMON order of /sol2dvar/
    const double *ush = nullptr, *dush = nullptr; // nz
    const double *dtdx = nullptr, *dtdz = nullptr; // physical
};
void launch_rhs_phys(const Geo &g, const iw_params &P, double t, double *const *uvwt,
                     double *const *gr, double *const *S, const double *ox, double *const *pr,
                     hipStream_t st, const MeanTerms *mt = nullptr);
// the same on FP32 arrays, FP32 arithmetic (mixed path; no mean-wave terms)
void launch_rhs_phys_f32(const Geo &g, const iw_params &P, double t, float *const *uvwt, float *const *gr,
                         float *const *S, const float *ox, float *const *pr, hipStream_t st);
// mixed path (iw_mixed.hip): launch_dpdz_conv_grad on FP32 fields with the
// contraction by D on FP32 matrix cores (MFMA); false if the order has no
// instantiation
bool launch_dpdz_conv_grad_mfma(const Geo &g, float *const *uvwt, float *const *out4, float *const *dz3,
                                hipStream_t st);
void launch_dx(const Geo &g, const double *a, double *b, hipStream_t st);
// diagnostics (iw_diag.hip): postp's means and d/dy, d/dz of umean, its
// TKE-budget sums (out: istep, t, the 16 energy.dat values), speclegencalc at
// j = ny/2 (out[kloc nsubd + ks]; hdr, if not null, gets istep and t)
int pp_blocks(const Geo &g);
void launch_pp_means(const Geo &g, double *u, double *const *sgs6, double *um, double *tm, double *dumdy,
                     double *dumdz, double beta, hipStream_t st);
void launch_pp_sums(const Geo &g, double *const *uvwt, double *const *grad12, double *const *sgs9,
                    const double *um, const double *dumdy, const double *dumdz, const double *tm, double *part,
                    double pr, double xkappa, double istep, double t, double *out, hipStream_t st);
void launch_speclegen(const Geo &g, const double *phys, double istep, double t, double *hdr, double *out,
                      hipStream_t st);
void launch_grad_xy(const Geo &g, const double *a, double *dx, double *dy, hipStream_t st);
// kinetic energy of u, v, w and potential energy of rho' (physical fields;
// lvl: 4 nz scratch, out: 4 values)
void launch_kepe(const Geo &g, double *const *phys4, double cke, double cpe, double span, double *lvl, double *out,
                 hipStream_t st);
void launch_rhs_spec(const Geo &g, double *const *S, double *const *pr, hipStream_t st);
// the z flux derivatives and z gradients (launch_dpdz_conv_grad) and rhsfm's
// physical part (launch_rhs_phys) in one pass, 3-D plain path without
// mean-wave terms: S and pr as launch_rhs_phys writes them; the z gradients
// (gr[2], gr[5], gr[8]) and d(-w rho)/dz are not stored.  false (nothing
// launched) for orders without specialised kernels or 2-D
bool launch_conv_rhs(const Geo &g, const iw_params &P, double t, double *const *uvwt, double *const *gr,
                     double *const *S, double *const *pr, hipStream_t st);
// the same on the mixed path's FP32 arrays (VALU sums in FP32)
bool launch_conv_rhs(const Geo &g, const iw_params &P, double t, float *const *uvwt, float *const *gr,
                     float *const *S, float *const *pr, hipStream_t st);
void launch_wnl(const Geo &g, const double *sw, const double *swn, const double *swnm,
                const double *fac, double *wnl, hipStream_t st);
void launch_wnl(const Geo &g, const float *sw, const float *swn, const float *swnm, const double *fac,
                double *wnl, hipStream_t st);
// n values FP64 -> FP32 (round to nearest) or FP32 -> FP64
void launch_cvt(const double *a, float *b, long n, hipStream_t st);
void launch_cvt(const float *a, double *b, long n, hipStream_t st);
void launch_rotvisc(const Geo &g, const iw_params &P, double *const *lev /*9*/, const double *fac,
                    double *wviscrot, hipStream_t st);
// SS extrapolation + BDF assembly + filters of u,v,w,rho in one launch:
// S12/F12 = (S, Sn, Snm) and (f, fn, fnm) per field, out4 the results
void launch_bdf_filter(const Geo &g, double *const *S12, double *const *F12, double *const *out4, const double *fac,
                       const double *bdf, double dt, int filt, hipStream_t st);
// bdf_filter fused with the pressure right-hand side (k_prhs) of the packed
// pressure path: phi receives what k_prhs would write from the filtered u, v,
// w; false (nothing launched) for orders without specialised kernels
bool launch_bdf_filter_prhs(const Geo &g, double *const *S12, double *const *F12, double *const *out4,
                            const double *fac, const double *bdf, double dt, int filt, const SysTables &S,
                            double *phi, const double *wnl, const double *wviscrot, hipStream_t st);
bool launch_bdf_filter_prhs(const Geo &g, float *const *S12, double *const *F12, double *const *out4,
                            const double *fac, const double *bdf, double dt, int filt, const SysTables &S,
                            double *phi, const double *wnl, const double *wviscrot, hipStream_t st);
// the same with the right-hand-side histories in FP32 (mixed path)
void launch_bdf_filter(const Geo &g, float *const *S12, double *const *F12, double *const *out4, const double *fac,
                       const double *bdf, double dt, int filt, hipStream_t st);
void launch_filter3d(const Geo &g, double *f, int ifourier, int ilegen, hipStream_t st);
void launch_iface(const Geo &g, FieldBatch b, int *count, hipStream_t st);
void launch_project(const Geo &g, FieldBatch b, hipStream_t st);
void launch_factor(const Geo &g, const SysTables &S, int first, int count, hipStream_t st);
void launch_compact(const Geo &g, const SysTables &S, int first, int count, hipStream_t st);
// one penalty system (coef: 12 nsubd, sysinfo: 4) factored into AB/ipiv,
// and the re and im columns x[0..nz), x[nz..2nz) solved with it
void launch_factor_one(const Geo &g, int ldab, const double *coef, double *sysinfo, double *AB, int *ipiv,
                       hipStream_t st);
void launch_band_solve1(const Geo &g, int ldab, const double *AB, const int *ipiv, double *x, hipStream_t st);
// Packed factors of the z systems, one block of nl rows per subdomain.  The
// penalty matrix couples a subdomain only to the interface rows of its
// neighbours, so below the diagonal the L column of local row r holds
// nl - r nonzeros (the rest of the subdomain and the next one's first row),
// with or without row interchanges; above it the U column holds r + 1
// (without interchanges: the rows above in the subdomain and the previous
// one's last row) or r + nl (with interchanges: the previous subdomain's
// rows as well).  Each segment is padded to whole double2.
constexpr int tri_ef(int nl, int r) { return nl - r; }                    // L[1..ef]
constexpr int tri_eb(int nl, int r, bool piv) { return r + (piv ? nl : 1); } // U[0..eb]
constexpr int tri_npf(int nl, int r) { return (tri_ef(nl, r) + 1) / 2; }
constexpr int tri_npb(int nl, int r, bool piv) { return (tri_eb(nl, r, piv) + 2) / 2; }
constexpr int tri_off_f(int nl, int r)
{
    int s = 0;
    for (int q = 0; q < r; q++) s += tri_npf(nl, q);
    return s;
}
constexpr int tri_off_b(int nl, int r, bool piv)
{
    int s = 0;
    for (int q = 0; q < r; q++) s += tri_npb(nl, q, piv);
    return s;
}
// double2 per subdomain block: forward, backward
constexpr int tri_sf(int nl) { return tri_off_f(nl, nl); }
constexpr int tri_sb(int nl, bool piv) { return tri_off_b(nl, nl, piv); }
// The Fourier filter of one spectral value (filtering.f:852-901): magnitude
// |f|, phase +-acos(f_r/|f|) (the sign of f_i), then |f| sigma (cos, sin) of
// the phase.  cos(acos c) = c and sin(acos c) = sqrt((1 - c)(1 + c)) >= 0
// give the same two values without the transcendental calls; c is the same
// rounded quotient, so the loss of the reference's form for nearly real
// modes (sin of the phase from 1 - c, c rounded) is kept, and the results
// agree with the acos/sin/cos evaluation to a few ulps of |f| sigma.
__device__ __forceinline__ void fourier_filter_value(double &fr, double &fi, double sigma)
{
    const double fmag = sqrt(fr * fr + fi * fi);
    if (fmag != 0.) {
        const double cosc = fr / fmag, sinc = fi / fmag;
        const double fmagf = fmag * sigma;
        const double sn = sqrt((1. - cosc) * (1. + cosc));
        fr = fmagf * cosc;
        fi = (sinc >= 0.) ? fmagf * sn : -(fmagf * sn);
    } else {
        fr = 0.;
        fi = 0.;
    }
}

// doubles per (system, subdomain) of SysTables::SPK
constexpr int spk_stride(int nl) { return 2 * nl + 4; }
void launch_pressure(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt,
                     const double *wnl, const double *wviscrot, double dt, int cpen, hipStream_t st);
void launch_viscous(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                    int cpen, hipStream_t st);
void launch_zero_meanw(const Geo &g, double *w, double *w00, hipStream_t st);
void launch_zero_int(int *p, int n, hipStream_t st);
bool solve_np_supported(int nl, int nsubd);
// whether ft threads per transform block fit the grid (iw_set_fft_threads)
bool fft_threads_ok(const Geo &g, int ft);
void launch_lfilter(const Geo &g, double *const *uvwt, int cpen, hipStream_t st);
// estimation-model filter: modecut of f into out on levels kmin..kmax (u0 scratch)
void launch_modecut(const Geo &g, const double *f, double *u0, double *out, int kmin, int kmax, hipStream_t st);
void launch_plane_mean(const Geo &g, const double *tspec, double *tm, hipStream_t st);
void launch_sgs_pre(const Geo &g, double *const *uvwt, const double *tm, double *const *ten, double *ox,
                    hipStream_t st);
void launch_sgs_post(const Geo &g, double *const *uvwt, const double *tm, double *const *ten,
                     double *const *tmp, double *ox, int part, hipStream_t st);
void launch_solve_grp(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                      hipStream_t st);
// subdomain-parallel solve of the u/v, w and rho systems (k_spike), and
// what completes it in place of launch_lfilter: the coupling recurrences and
// correction, the Legendre filter and the mean-u damping (k_spike_finish)
// part 0: every system; with the u/v/w groups (S.share_uvw) part 1 launches
// all but the rho solve and part 2 the rho solve alone (it needs nothing the
// pressure step changes, so it can run beside it); without them part 1 is
// everything and part 2 nothing
void launch_solve_spike(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double dt,
                        hipStream_t st, int part = 0);
void launch_spike_finish(const Geo &g, const SysTables &S, double *const *uvwt, int cpen, hipStream_t st,
                         int part = 0);
// its per-(system, subdomain) vectors, after launch_compact
void launch_spike_pre(const Geo &g, const SysTables &S, int first, int count, hipStream_t st);
// sets *flag when the u/v and w factor blocks below the top subdomain differ
void launch_share_check(const Geo &g, const SysTables &S, int *flag, hipStream_t st);
// part 0: right-hand side, 1: solve, 2: filter + projection
void launch_pressure_nat(const Geo &g, const iw_params &P, const SysTables &S, double *const *uvwt, double *phi,
                         const double *wnl, const double *wviscrot, double dt, int cpen, int part,
                         hipStream_t st);

// ---- device helpers shared by the kernel files
// compile-time loop: f(std::integral_constant<int, I>) for I = B..E-1
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}
template <int I>
using IC = std::integral_constant<int, I>;
// the row updates of a sweep stay in the step that makes them: without this
// the compiler sinks each update into the step that first reads the value,
// keeping every pending multiplier of a subdomain in registers
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }

// Lanes of one wavefront exchange data through LDS: a release/acquire fence
// at wavefront scope with a wave barrier between them orders this lane's
// earlier LDS stores before the other lanes' later loads.  A single
// wavefront's LDS instructions execute in order, so the fences need no wait
// counter and only fix the compiler's order.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the same filter as one matrix FL (nl x nl, in LDS): x[q] <- sum_t
// FL[t nl + q] x[t], the two sums above taken together (half the
// multiply-adds; equal to them up to rounding)
template <int NL>
__device__ __forceinline__ void legfilter_mat(double (&x)[NL], const double *FLs)
{
    double o[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) o[q] = 0.;
#pragma unroll
    for (int t = 0; t < NL; t++) {
        const double xt = x[t];
#pragma unroll
        for (int q = 0; q < NL; q++) o[q] = o[q] + FLs[t * NL + q] * xt;
    }
#pragma unroll
    for (int q = 0; q < NL; q++) x[q] = o[q];
}

// two columns through the same matrix, each row of FL read once for both
template <int NL>
__device__ __forceinline__ void legfilter_mat2(double (&x)[NL], double (&y)[NL], const double *FLs)
{
    double o[NL], p[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) {
        o[q] = 0.;
        p[q] = 0.;
    }
#pragma unroll
    for (int t = 0; t < NL; t++) {
        const double xt = x[t], yt = y[t];
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const double c = FLs[t * NL + q];
            o[q] = o[q] + c * xt;
            p[q] = p[q] + c * yt;
        }
    }
#pragma unroll
    for (int q = 0; q < NL; q++) {
        x[q] = o[q];
        y[q] = p[q];
    }
}

template <int NL>
__device__ __forceinline__ void stage_fl(double *FLs, const double *FL)
{
    for (int e = threadIdx.x; e < NL * NL; e += blockDim.x) FLs[e] = FL[e];
    __syncthreads();
}

} // namespace iw
