// Ensembles of independent runs of uneven sizes advanced together (gfx950).
//
// No counterpart in the reference, which sizes one run at compile time
// (dim.h:12-28).  Each member is an ordinary instance (iw_create) with its own
// stream; the ensemble advances all of them with one HIP graph launch per
// three steps: the steady step sequence of every member (iw_run after the
// start-up factorisations) is captured once, three steps deep because the
// state histories rotate by pointer with period three, with the members
// dealt to `lanes` independent chains of the graph (members of one chain run
// one after another, chains run side by side).  An autotuner times, on the
// members' own state (saved and restored around the trials), the transform
// block size of every member and the number of chains, and keeps the
// fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "iw_internal.h"

using namespace iw;

struct iw_ensemble {
    std::vector<iw_solver *> m;
    int device = 0;
    hipStream_t cs = nullptr; // fork / join and graph launch stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> phase;   // istep mod 3 of each member when captured
    int lanes = 1;
    bool use_graph = true;
    std::vector<int> lane_of; // member -> chain
    std::vector<double> cost; // measured ms per step of each member alone
    std::vector<hipEvent_t> ev; // fork (0) and one per member, reused
};

#define EK(x)                                                                                                  \
    do {                                                                                                       \
        hipError_t e_ = (x);                                                                                   \
        if (e_ != hipSuccess) return iw_fail(std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

static void drop_graph(iw_ensemble *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

static int check_member(const iw_solver *s)
{
    const iw_params &P = s->P;
    if (P.imaxcalc || P.dtadapt) return iw_fail("ensemble: members run without maxcalc records and adaptive dt");
    if (P.ifilter) return iw_fail("ensemble: members run without the estimation filter");
    if (P.navgintrfc > 1 || P.nspecfilterexp > 1)
        return iw_fail("ensemble: interface averaging and the solution filter must act every step");
    if (P.n3doutput > 0 || P.ncalclegspec > 0 || (P.npostp > 0 && P.npostp < 100000000))
        return iw_fail("ensemble: members run without the output routines");
    // the captured steps keep the kernel arguments of the capture; the
    // paddle forcing is the one term that depends on the time
    if (P.xacrit != 0.) return iw_fail("ensemble: members run without the virtual paddle");
    return 0;
}

extern "C" int iw_ensemble_destroy(iw_ensemble *e);

extern "C" int iw_ensemble_create(iw_solver **members, int n, iw_ensemble **out)
{
    if (!members || n < 1 || !out) return iw_fail("iw_ensemble_create: bad argument");
    iw_ensemble *e = new iw_ensemble();
    e->device = members[0]->device;
    for (int i = 0; i < n; i++) {
        if (!members[i] || members[i]->device != e->device) {
            delete e;
            return iw_fail("iw_ensemble_create: members must be instances on one device");
        }
        if (check_member(members[i])) {
            delete e;
            return -1;
        }
        e->m.push_back(members[i]);
    }
    if (hipStreamCreateWithFlags(&e->cs, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return iw_fail("iw_ensemble_create: stream");
    }
    e->lane_of.assign(n, 0);
    e->cost.assign(n, 1.0);
    e->ev.assign(n + 1, nullptr);
    for (auto &v : e->ev)
        if (hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess) {
            iw_ensemble_destroy(e);
            return iw_fail("iw_ensemble_create: events");
        }
    for (iw_solver *s : e->m) s->one_stream = true;
    *out = e;
    return 0;
}

extern "C" int iw_ensemble_destroy(iw_ensemble *e)
{
    if (!e) return 0;
    drop_graph(e);
    for (iw_solver *s : e->m) s->one_stream = false;
    for (hipEvent_t v : e->ev)
        if (v) (void)hipEventDestroy(v);
    if (e->cs) (void)hipStreamDestroy(e->cs);
    delete e;
    return 0;
}

// members past their start-up steps (the factorisations of steps 1-3)
static int warm(iw_ensemble *e)
{
    for (iw_solver *s : e->m)
        if (s->istep < 3 && iw_run(s, 3 - s->istep, nullptr)) return -1;
    for (iw_solver *s : e->m) EK(hipStreamSynchronize(s->stream));
    return 0;
}

// longest-processing-time deal of the members to `lanes` chains
static void deal(iw_ensemble *e, int lanes)
{
    const int n = (int)e->m.size();
    e->lanes = std::max(1, std::min(lanes, n));
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return e->cost[a] > e->cost[b]; });
    std::vector<double> load(e->lanes, 0.);
    for (int i : idx) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        e->lane_of[i] = l;
        load[l] += e->cost[i];
    }
}

// three steps of every member, chains forked from and joined to cs; under
// capture nothing runs and the host-side step clock is put back afterwards
static int three_steps(iw_ensemble *e, bool capture)
{
    const int n = (int)e->m.size();
    hipEvent_t fork = e->ev[0];
    hipEvent_t *done = e->ev.data() + 1;
    EK(hipEventRecord(fork, e->cs));
    std::vector<int> last(e->lanes, -1);
    // members in a chain wait for the previous member of the chain
    for (int i = 0; i < n; i++) {
        iw_solver *s = e->m[i];
        const int l = e->lane_of[i];
        EK(hipStreamWaitEvent(s->stream, last[l] < 0 ? fork : done[last[l]], 0));
        const double t0 = s->t;
        const int i0 = s->istep;
        if (iw_run(s, 3, nullptr)) return -1;
        if (capture) {
            s->t = t0;
            s->istep = i0;
        }
        EK(hipEventRecord(done[i], s->stream));
        last[l] = i;
    }
    for (int l = 0; l < e->lanes; l++)
        if (last[l] >= 0) EK(hipStreamWaitEvent(e->cs, done[last[l]], 0));
    return 0;
}

static int capture(iw_ensemble *e)
{
    drop_graph(e);
    EK(hipStreamSynchronize(e->cs));
    EK(hipStreamBeginCapture(e->cs, hipStreamCaptureModeThreadLocal));
    const int rc = three_steps(e, true);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->cs, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return -1;
    }
    if (ec != hipSuccess) return iw_fail(std::string("ensemble capture: ") + hipGetErrorString(ec));
    e->graph = g;
    EK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->phase.clear();
    for (iw_solver *s : e->m) e->phase.push_back(s->istep % 3);
    return 0;
}

static bool graph_current(const iw_ensemble *e)
{
    if (!e->exec || e->phase.size() != e->m.size()) return false;
    for (size_t i = 0; i < e->m.size(); i++)
        if (e->m[i]->istep % 3 != e->phase[i]) return false;
    return true;
}

static int advance(iw_ensemble *e, int nsteps)
{
    // start-up steps (each member's first three, with the factorisations)
    // eagerly; then graph replays for the steps every member still has
    const int n = (int)e->m.size();
    std::vector<int> left(n, nsteps);
    for (int i = 0; i < n; i++) {
        iw_solver *s = e->m[i];
        const int w = std::min(left[i], std::max(0, 3 - s->istep));
        if (w > 0 && iw_run(s, w, nullptr)) return -1;
        left[i] -= w;
    }
    for (iw_solver *s : e->m) EK(hipStreamSynchronize(s->stream));
    const int common = *std::min_element(left.begin(), left.end());
    if (e->use_graph && common >= 3) {
        if (!graph_current(e) && capture(e)) return -1;
        for (int k = 0; k + 3 <= common; k += 3) {
            EK(hipGraphLaunch(e->exec, e->cs));
            for (int i = 0; i < n; i++) { // the host-side step clock of the replayed steps
                iw_solver *s = e->m[i];
                s->istep += 3;
                s->t = s->t + s->P.dt;
                s->t = s->t + s->P.dt;
                s->t = s->t + s->P.dt;
                left[i] -= 3;
            }
        }
    }
    EK(hipStreamSynchronize(e->cs));
    // the rest (or everything without the graph) on the members' streams
    for (int i = 0; i < n; i++)
        if (left[i] > 0 && iw_run(e->m[i], left[i], nullptr)) return -1;
    for (iw_solver *s : e->m) EK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int iw_ensemble_run(iw_ensemble *e, int nsteps)
{
    if (!e || nsteps < 0) return iw_fail("iw_ensemble_run: bad argument");
    return advance(e, nsteps);
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
struct Saved {
    std::vector<double *> a;
    std::vector<double> t;
    std::vector<int> istep;
};

static int save(iw_ensemble *e, Saved &sv)
{
    for (iw_solver *s : e->m) {
        const size_t fb = sizeof(double) * (size_t)s->g.ntot, pb = sizeof(double) * 2 * (size_t)s->g.nxppy;
        for (int q = 0; q < 26; q++) {
            double *d = nullptr;
            EK(hipMalloc((void **)&d, q < 24 ? fb : pb));
            EK(hipMemcpyAsync(d, q < 24 ? s->buf[q] : (q == 24 ? s->wnl : s->wviscrot), q < 24 ? fb : pb,
                              hipMemcpyDeviceToDevice, s->stream));
            sv.a.push_back(d);
        }
        sv.t.push_back(s->t);
        sv.istep.push_back(s->istep);
        EK(hipStreamSynchronize(s->stream));
    }
    return 0;
}

static int restore(iw_ensemble *e, Saved &sv)
{
    for (size_t i = 0; i < e->m.size(); i++) {
        iw_solver *s = e->m[i];
        EK(hipStreamSynchronize(s->stream));
        const size_t fb = sizeof(double) * (size_t)s->g.ntot, pb = sizeof(double) * 2 * (size_t)s->g.nxppy;
        for (int q = 0; q < 26; q++)
            EK(hipMemcpyAsync(q < 24 ? s->buf[q] : (q == 24 ? s->wnl : s->wviscrot), sv.a[26 * i + q],
                              q < 24 ? fb : pb, hipMemcpyDeviceToDevice, s->stream));
        s->t = sv.t[i];
        s->istep = sv.istep[i];
        EK(hipStreamSynchronize(s->stream));
    }
    for (double *d : sv.a) (void)hipFree(d);
    sv.a.clear();
    return 0;
}

static double wall_ms(std::chrono::steady_clock::time_point a)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

// Tune the ensemble on `steps` steps per trial (a multiple of 3): first the
// transform block size of each member alone (64, 128, 256 threads where they
// fit), then the number of graph chains (1, 2, 4, 8, 16, all) and the eager
// alternative (members on their own streams, no graph).  The members' state
// is the same before and after.  out (may be null): [chosen chains (0 for no
// graph), ms per ensemble step of the choice, ms of the default (all chains,
// 256 threads)].
extern "C" int iw_ensemble_tune(iw_ensemble *e, int steps, double *out)
{
    if (!e || steps < 3) return iw_fail("iw_ensemble_tune: bad argument");
    steps -= steps % 3;
    if (warm(e)) return -1;
    Saved sv;
    if (save(e, sv)) return -1;
    const int n = (int)e->m.size();
    // per member: block size
    for (int i = 0; i < n; i++) {
        iw_solver *s = e->m[i];
        double best = 1e300;
        int bft = s->g.ft;
        for (int ft : {256, 128, 64}) {
            if (!fft_threads_ok(s->g, ft)) continue;
            s->g.ft = ft;
            if (iw_run(s, 3, nullptr)) return -1; // warm the variant
            EK(hipStreamSynchronize(s->stream));
            const auto t0 = std::chrono::steady_clock::now();
            if (iw_run(s, steps, nullptr)) return -1;
            EK(hipStreamSynchronize(s->stream));
            const double ms = wall_ms(t0) / steps;
            if (ms < best) {
                best = ms;
                bft = ft;
            }
        }
        s->g.ft = bft;
        e->cost[i] = best;
    }
    // graph chains and the eager alternative, timed on the whole ensemble
    std::vector<int> cand;
    for (int l : {1, 2, 4, 8, 16, n})
        if (l <= n && std::find(cand.begin(), cand.end(), l) == cand.end()) cand.push_back(l);
    double best = 1e300, dflt = 0.;
    int bl = n;
    bool bg = true;
    for (int l : cand) {
        deal(e, l);
        e->use_graph = true;
        drop_graph(e);
        if (advance(e, 3)) return -1; // capture and one replay
        const auto t0 = std::chrono::steady_clock::now();
        if (advance(e, steps)) return -1;
        const double ms = wall_ms(t0) / steps;
        if (l == n) dflt = ms;
        if (ms < best) {
            best = ms;
            bl = l;
        }
    }
    e->use_graph = false;
    {
        if (advance(e, 3)) return -1;
        const auto t0 = std::chrono::steady_clock::now();
        if (advance(e, steps)) return -1;
        const double ms = wall_ms(t0) / steps;
        if (ms < best) {
            best = ms;
            bg = false;
        }
    }
    e->use_graph = bg;
    deal(e, bl);
    drop_graph(e);
    if (restore(e, sv)) return -1;
    if (out) {
        out[0] = bg ? bl : 0;
        out[1] = best;
        out[2] = dflt;
    }
    return 0;
}

// The configuration in use: lanes (0 = no graph) and every member's transform
// block size.
extern "C" int iw_ensemble_config(iw_ensemble *e, int *lanes, int *ft)
{
    if (!e) return iw_fail("iw_ensemble_config: null argument");
    if (lanes) *lanes = e->use_graph ? e->lanes : 0;
    if (ft)
        for (size_t i = 0; i < e->m.size(); i++) ft[i] = e->m[i]->g.ft;
    return 0;
}

// Use (lanes > 0) or skip (lanes = 0) the graph with `lanes` chains.
extern "C" int iw_ensemble_set_lanes(iw_ensemble *e, int lanes)
{
    if (!e || lanes < 0) return iw_fail("iw_ensemble_set_lanes: bad argument");
    e->use_graph = lanes > 0;
    if (lanes > 0) deal(e, lanes);
    drop_graph(e);
    return 0;
}
