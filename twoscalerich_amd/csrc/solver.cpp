// solver.cpp - device-resident EXPREM driver and the reference-signature entry points.
//
// One accepted exponential Rosenbrock-Euler step per call (the reference's exprem
// contract, include/main.h:29-32, main.cpp:311-341), with the exprb32 embedded error
// estimate (Hochbruck, Ostermann & Schweitzer, SINUM 47 (2009)):
//     U2 = u + tau phi1(tau J) G(u),  D = G(U2) - G(u) - J (U2 - u),
//     u+ = U2 + 2 tau phi3(tau J) D,  err = WRMS(2 tau phi3(tau J) D).
// phi-vector products by Arnoldi with J v = (G(u + eps v) - G(u)) / eps,
// eps = sqrt(DBL_EPSILON)(1 + ||u||_2), CGS2 orthogonalisation; J (U2 - u) from the
// Arnoldi relation; Krylov stop when tau (tau beta h_{m+1,m} |e_m^T phi1(tau H_m) e1|)
// WRMS(v_{m+1}) <= 0.1; step law tau * min(MaxStepIncFac, max(MinStepDecFac,
// SafetyFac err^(-1/3))); a rejected step reuses the Krylov basis.
//
// Everything of length N lives in HBM; the host holds only the (m+1) x m Hessenberg
// matrices and per-scenario scalars.  All scenarios of a batch advance together: each
// Arnoldi iteration, update and error estimate is one launch over the whole batch
// with a per-scenario activity mask, so converged or accepted scenarios cost nothing.
#include "device.h"
#include "hostpool.h"
#include "tsr_internal.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            tsr::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);        \
            return fail_code;                                                                     \
        }                                                                                         \
    } while (0)

namespace tsr {
constexpr double kKrylovTol = 0.1;
constexpr int kMaxKrylovCap = 120;   // largest MaxKrylov accepted
constexpr int kLearnedCap = 60;       // largest Krylov cap the learned controller may propose
struct Mlp;   // learned controller (mlp.hip)
Mlp*  mlp_create();
void  mlp_free(Mlp*);
int   mlp_set(Mlp*, const float* params, const float* mean, const float* stdv);
int   mlp_infer(Mlp*, const float* feats, int n, float* out, hipStream_t st);
int   mlp_infer_device(Mlp*, const float* feats, int n, const double* stats, hipStream_t st);
const float* mlp_output(const Mlp*);
bool  mlp_ready(const Mlp*);
}  // namespace tsr

using namespace tsr;

// Sampled per-kernel device timing (tsr_solver_ktimer): an event pair around every
// `every`-th launch of each timed class, harvested whenever the solver has synchronised
// its stream (every recorded pair is then complete).  While an Arnoldi window is being
// captured into a HIP graph the pairs are recorded as external event nodes of the graph
// (their own events, sampled by a capture-time count): every replay records them again,
// and the replay hands them, with the window's launch counts, to the pending list.
struct KTimer {
    int every = 0;
    long long seen[KT_COUNT] = {}, timed[KT_COUNT] = {};
    double ms[KT_COUNT] = {}, bytes[KT_COUNT] = {};
    struct Pend { hipEvent_t a, b; int cls; double bytes; bool graph; };
    std::vector<Pend> pend;
    std::vector<hipEvent_t> pool, all;
    int open = -1;
    double open_bytes = 0.0;
    hipEvent_t open_a = nullptr;
    bool capturing = false;
    bool cap_timed = false;                // this capture carries timer pairs
    std::vector<Pend> cap;                 // pairs of the window being captured
    long long cap_seen[KT_COUNT] = {};     // its launches per class
    long long cap_n[KT_COUNT] = {};        // sampling count across captures
};
// what a captured window's replay contributes to the timer
struct GraphTiming {
    std::vector<KTimer::Pend> pairs;
    long long seen[KT_COUNT] = {};
};
// One Arnoldi window shape: its graph, and (while the kernel timer is on) a second
// capture that carries the timer's event pairs, replayed every kGraphTimedEvery-th time
// so the timed launches are an even sample of the replays.
struct GraphEntry {
    hipGraphExec_t exec = nullptr, exec_t = nullptr;
    GraphTiming gt, gt_t;
    long long replays = 0;
};

namespace tsr {
thread_local KTimer* g_kt = nullptr;   // the timer of the solver stepping on this thread
// share of the unknowns in scenarios the current activity mask keeps (masked launches
// move only those scenarios' bytes)
thread_local double g_kt_frac = 1.0;
double kt_active_fraction() { return g_kt_frac; }

static hipEvent_t kt_event(KTimer& k)
{
    if (!k.pool.empty()) { hipEvent_t e = k.pool.back(); k.pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    k.all.push_back(e);
    return e;
}

// an event of its own for a graph (destroyed with the solver, never pooled)
static hipEvent_t kt_graph_event(KTimer& k)
{
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    k.all.push_back(e);
    return e;
}

bool kt_begin(int cls, double bytes, hipStream_t st)
{
    KTimer* k = g_kt;
    if (!k || k->every <= 0 || cls < 0 || cls >= KT_COUNT || k->open >= 0) return false;
    if (k->capturing) {
        k->cap_seen[cls]++;
        if (!k->cap_timed || k->cap_n[cls]++ % k->every) return false;
        hipEvent_t a = kt_graph_event(*k);
        if (!a || hipEventRecordWithFlags(a, st, hipEventRecordExternal) != hipSuccess) return false;
        k->open = cls; k->open_a = a; k->open_bytes = bytes;
        return true;
    }
    if (k->seen[cls]++ % k->every) return false;
    hipEvent_t a = kt_event(*k);
    if (!a || hipEventRecord(a, st) != hipSuccess) return false;
    k->open = cls; k->open_a = a; k->open_bytes = bytes;
    return true;
}

void kt_end(hipStream_t st)
{
    KTimer* k = g_kt;
    if (!k || k->open < 0) return;
    if (k->capturing) {
        hipEvent_t b = kt_graph_event(*k);
        if (b && hipEventRecordWithFlags(b, st, hipEventRecordExternal) == hipSuccess)
            k->cap.push_back({k->open_a, b, k->open, k->open_bytes, true});
        k->open = -1;
        return;
    }
    hipEvent_t b = kt_event(*k);
    if (b && hipEventRecord(b, st) == hipSuccess) k->pend.push_back({k->open_a, b, k->open, k->open_bytes, false});
    else k->pool.push_back(k->open_a);
    k->open = -1;
}

// after a stream synchronisation: every pending pair is complete
static void kt_harvest(KTimer& k)
{
    for (const auto& p : k.pend) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            k.ms[p.cls] += ms; k.bytes[p.cls] += p.bytes; k.timed[p.cls]++;
        }
        if (!p.graph) { k.pool.push_back(p.a); k.pool.push_back(p.b); }
    }
    k.pend.clear();
}
}  // namespace tsr

struct ScState {
    double t = 0.0, proposal = 1e-3, tau = 0.0, tstop = 0.0, tau_prop = 0.0;
    bool landing = false;
    int m = 0, mD = 0, rejects = 0, fe = 0, m_cap = 0, ctrl = 0;
    double beta = 0.0, hn = 0.0, hnD = 0.0, betaD = 0.0, err = 0.0, eps = 0.0, wnext = 0.0, wnextD = 0.0;
    bool bd = false, bdD = false, conv = false, failed = false, accepted = false;
    std::vector<double> H, HD, phi, y, z, cs, csW;
    double phi_tau = -1.0;    // (tau, m) of phi_1(tau H_m) e1, the first column of phi
    int phi_m = -1;
    // phi_k(tau HD) e1, k = 1..3, of the error basis at (ph3_tau, ph3_m) (the stopping
    // rule's last evaluation, reused by the error estimate)
    std::vector<double> ph3;
    double ph3_tau = -1.0;
    int ph3_m = -1;
    exprem_stats_struct stats{};
    // recent history and state summary for the step controller
    double hist_tau[4] = {0, 0, 0, 0};
    double hist_m[4] = {0, 0, 0, 0};
    int    n_hist = 0;            // accepted steps seen (saturates at 4)
    double last_hnorm = 0.0, last_resid = 0.0;
    double mstat[6] = {0, 0, 0, 0, 0, 0};
    float  feat[16] = {};
    double tau_first = 0.0, err_first = 0.0;
    bool   first_done = false;
    int    pred_m = 0, pred_mD = 0;   // predicted Krylov dimensions (deferred checks)
    double last_err = 0.0;            // error estimate of the last accepted step
    unsigned long long rng = 0;       // exploration controller's generator
};

struct tsr_solver {
    DevProblem* P = nullptr;
    hipStream_t st = nullptr;
    long long N = 0;
    int S = 0, mcap = 0;
    exprem_options_struct opt{};
    // vectors
    double *u = nullptr, *g0 = nullptr, *u2 = nullptr, *g2 = nullptr, *D = nullptr, *unew = nullptr;
    double* scratch[2] = {nullptr, nullptr};
    std::vector<double*> V, W;
    const double** Vd = nullptr;
    const double** Wd = nullptr;
    // TSR_PATH_F32BASIS: FP32 copies of the basis columns (allocated by set_path)
    std::vector<float*> V32, W32;
    const float** V32d = nullptr;
    const float** W32d = nullptr;
    // blocks
    int nblk = 0;
    BlkMap M{};
    int* blk_first = nullptr;
    long long* scen_rows = nullptr;
    std::vector<long long> rows_h;
    double* part = nullptr;
    int ldp = 0;
    KrylovState KV{}, KW{};
    int* mask = nullptr;      // device activity mask
    int* nk = nullptr;        // per-scenario column counts for k_lincomb
    // the step's second combination (the Arnoldi-relation term of D), formed in the same
    // pass over the basis as U2: its coefficients and counts, the vector, and the
    // one-column list {p2} with unit coefficients that turns it into D
    double* coef2 = nullptr;
    int* nk2 = nullptr;
    double* p2 = nullptr;
    const double** p2col = nullptr;
    double* one_coef = nullptr;
    int* one_nk = nullptr;
    int* bad = nullptr;       // alias of DBatch::bad
    std::vector<int> mask_h, nk_h, bad_h;
    std::vector<int> mask_dev_h;   // what the device mask holds (uploads only on change)
    std::vector<double> coef_h, res_h, veps_h, cs0_h, coef2_h;
    std::vector<int> nk2_h;
    std::vector<ScState> sc;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_g_ms = 0.0;
    // host RHS mode (exprem_c with a host Gfunc)
    tsr_gfunc_t host_g = nullptr;
    user_data_struct host_ud{};
    std::vector<double> hx, hg, hg0;
    // controller
    int ctrl_mode = 0;
    Mlp* mlp = nullptr;
    // timing of the last run
    double g_ms_acc = 0.0;
    int    g_calls = 0;
    double* mstats = nullptr;   // device [S*6] macro head / saturation summary
    double* mb_h = nullptr;     // mapped pinned readback buffer (host view) ...
    double* mb_d = nullptr;     // ... and its device address
    int mb_stride = 0;          // doubles per scenario
    std::vector<float> ctl_h;   // learned controller's output per scenario (2), read with the step start
    int path = TSR_PATH_PLAIN;
    double* linrec = nullptr;   // TSR_PATH_JVLIN: expansion point of the FD products (kLinBytesPerNode x N bytes)
    double t_dense = 0.0;     // host seconds in the small dense phi routines
    double t_sync = 0.0;      // host seconds waiting for the stream
    KTimer kt;                // sampled per-kernel timing (off unless tsr_solver_ktimer)
    // TSR_PATH_GRAPH: instantiated graphs of Arnoldi windows, keyed by what their kernel
    // arguments depend on (basis, first and last iteration, path bits, state buffer)
    std::map<std::tuple<int, int, int, int, const double*>, GraphEntry> graphs;
    long long graph_launches = 0, graph_captures = 0;
};

namespace {
struct Tic {
    double& acc;
    std::chrono::steady_clock::time_point t0;
    explicit Tic(double& a) : acc(a), t0(std::chrono::steady_clock::now()) {}
    ~Tic() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};
tsr_solver* g_timed = nullptr;   // solver whose dense/sync time is being accumulated
hipError_t timed_sync(tsr_solver* s)
{
    Tic tic(s->t_sync);
    const hipError_t e = hipStreamSynchronize(s->st);
    if (e == hipSuccess && !s->kt.pend.empty()) kt_harvest(s->kt);
    return e;
}

// Stream capture of one Arnoldi window (TSR_PATH_GRAPH): with timed, the kernel timer
// records its sampled pairs into the graph (KTimer::capturing); without, it only counts
// the window's launches.  A capture left open by an early error return is ended and
// discarded.
constexpr long long kGraphTimedEvery = 8;
struct CaptureGuard {
    hipStream_t st;
    bool active = false;
    explicit CaptureGuard(hipStream_t s) : st(s) {}
    hipError_t begin(bool timed)
    {
        if (g_kt) {
            g_kt->capturing = true;
            g_kt->cap_timed = timed;
            g_kt->cap.clear();
            for (int c = 0; c < KT_COUNT; ++c) g_kt->cap_seen[c] = 0;
        }
        const hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        active = e == hipSuccess;
        if (!active && g_kt) g_kt->capturing = false;
        return e;
    }
    // ends the capture; the window's timer pairs and launch counts go to *gt
    hipError_t end(hipGraph_t* g, GraphTiming* gt)
    {
        active = false;
        if (g_kt) {
            g_kt->capturing = false;
            gt->pairs.swap(g_kt->cap);
            for (int c = 0; c < KT_COUNT; ++c) gt->seen[c] = g_kt->cap_seen[c];
        }
        return hipStreamEndCapture(st, g);
    }
    ~CaptureGuard()
    {
        if (!active) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) hipGraphDestroy(g);
        if (g_kt) { g_kt->capturing = false; g_kt->cap.clear(); g_kt->open = -1; }
    }
};

// the solver's kernel timer is active on this thread while it steps
struct KtScope {
    KTimer* keep;
    explicit KtScope(tsr_solver* s) : keep(g_kt) { g_kt = s->kt.every > 0 ? &s->kt : nullptr; }
    ~KtScope() { g_kt = keep; }
};
double dense_dummy = 0.0;
// a host-pool thread evaluating one scenario's dense functions times them here instead
// (the pool's wall time is what the stepping thread waits for, and is counted once)
thread_local double* t_dense_tls = nullptr;
double& dense_acc() { return t_dense_tls ? *t_dense_tls : (g_timed ? g_timed->t_dense : dense_dummy); }

// f(q) for every scenario q with want[q] != 0.  With several such scenarios and more
// than one host thread they are spread over the host pool (f(q) must touch only scenario
// q's state and its own entries of shared per-scenario arrays, call no HIP function and
// leave errors to the caller), and the pool's wall time counts as dense time.  Otherwise
// they run in order on this thread.
void for_scenarios(const std::vector<int>& want, const std::function<void(int)>& f)
{
    std::vector<int> qs;
    for (int q = 0; q < (int)want.size(); ++q)
        if (want[q]) qs.push_back(q);
    if (qs.size() < 2 || HostPool::get().threads() == 1) {
        for (const int q : qs) f(q);
        return;
    }
    Tic tic(dense_acc());
    HostPool::get().run((int)qs.size(), [&](int i) {
        double own = 0.0;
        double* keep = t_dense_tls;
        t_dense_tls = &own;
        f(qs[i]);
        t_dense_tls = keep;
    });
}
}

namespace {

std::mutex g_cache_mu;
std::map<const tsr_batch_view*, tsr_solver*> g_solver_cache;
// solvers replaced when a caller changed options, with their problem's view; freed
// with the problem (solver_cache_drop)
std::vector<std::pair<const tsr_batch_view*, tsr_solver*>> g_retired;

template <class T> int dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1)) == hipSuccess ? 0 : 1;
}

// ---- device helpers with error propagation ----
int G_call(tsr_solver* s, const double* v, const double* veps, const double* g0, const double* ieps, double* out,
           const int* active)
{
    const int fail_code = 2;
    if (s->host_g) {
        // host RHS: state crosses PCIe, Krylov arithmetic stays on the device
        const size_t N = (size_t)s->N;
        HIPCHK(hipMemcpyAsync(s->hx.data(), s->u, N * sizeof(double), hipMemcpyDeviceToHost, s->st));
        std::vector<double> vh;
        double ve = 0.0, ie = 0.0;
        if (v) {
            vh.resize(N);
            HIPCHK(hipMemcpyAsync(vh.data(), v, N * sizeof(double), hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipMemcpyAsync(&ve, veps, sizeof(double), hipMemcpyDeviceToHost, s->st));
        }
        if (g0) HIPCHK(hipMemcpyAsync(&ie, ieps, sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(timed_sync(s));
        if (v) for (size_t i = 0; i < N; ++i) s->hx[i] += ve * vh[i];
        const int flag = s->host_g(s->hg.data(), s->hx.data(), s->host_ud);
        if (g0) for (size_t i = 0; i < N; ++i) s->hg[i] = (s->hg[i] - s->hg0[i]) * ie;
        else if (out == s->g0) s->hg0 = s->hg;
        bool finite = true;
        for (size_t i = 0; i < N && finite; ++i) finite = std::isfinite(s->hg[i]);
        s->bad_h[0] = (flag != 0 || !finite) ? 1 : 0;
        HIPCHK(hipMemcpyAsync(s->bad, s->bad_h.data(), sizeof(int), hipMemcpyHostToDevice, s->st));
        HIPCHK(hipMemcpyAsync(out, s->hg.data(), N * sizeof(double), hipMemcpyHostToDevice, s->st));
        s->g_calls++;
        return 0;
    }
    GArgs a;
    a.u = s->u; a.v = v; a.veps = veps; a.g0 = g0; a.ieps = ieps; a.active = active; a.out = out;
    if (s->path & TSR_PATH_JVLIN) {
        if (!v && out == s->g0) a.rec = s->linrec;          // G(u) at the step start
        else if (v && g0 == s->g0) a.lin = s->linrec;      // FD product around that u
    }
    HIPCHK(launch_gfunc(s->P, a, s->st));
    s->g_calls++;
    return 0;
}

// G(x) for an explicit state x (not s->u): temporarily point u at x.
int G_at(tsr_solver* s, const double* x, double* out, const int* active)
{
    double* keep = s->u;
    s->u = const_cast<double*>(x);
    const int rc = G_call(s, nullptr, nullptr, nullptr, nullptr, out, active);
    s->u = keep;
    return rc;
}

int upload_mask(tsr_solver* s)
{
    const int fail_code = 2;
    {
        double on = 0.0, all = 0.0;
        for (int q = 0; q < s->S; ++q) { all += (double)s->rows_h[q]; if (s->mask_h[q]) on += (double)s->rows_h[q]; }
        g_kt_frac = all > 0.0 ? on / all : 1.0;
    }
    if (s->mask_dev_h == s->mask_h) return 0;   // the device copy is current
    HIPCHK(hipMemcpyAsync(s->mask, s->mask_h.data(), sizeof(int) * s->S, hipMemcpyHostToDevice, s->st));
    s->mask_dev_h = s->mask_h;
    return 0;
}

int clear_bad(tsr_solver* s)
{
    const int fail_code = 2;
    HIPCHK(hipMemsetAsync(s->bad, 0, sizeof(int) * s->S, s->st));
    return 0;
}

// Read back the per-scenario results (res, 8 slots) and bad flags, plus the FD step
// and the macro statistics when eps / stats are given: one pack kernel into the mapped
// buffer, one sync.
// ctl: the controller's device output ([S][2] floats) to read back with the rest (into ctl_h)
int read_iteration(tsr_solver* s, KrylovState& K, double* eps_out = nullptr, bool stats = false,
                   const float* ctl = nullptr)
{
    const int fail_code = 2;
    HIPCHK(launch_pack_small(K.res, s->bad, eps_out ? K.eps : nullptr, stats ? s->mstats : nullptr, ctl, s->S, s->mb_d,
                             s->st));
    HIPCHK(timed_sync(s));
    for (int q = 0; q < s->S; ++q) {
        const double* o = s->mb_h + (size_t)q * kPackSmall;
        std::memcpy(&s->res_h[(size_t)q * 8], o, sizeof(double) * 8);
        s->bad_h[q] = o[8] != 0.0;
        if (eps_out) eps_out[q] = o[9];
        if (stats) std::memcpy(s->sc[q].mstat, o + 10, sizeof(double) * 6);
        if (ctl) { s->ctl_h[(size_t)q * 2] = (float)o[16]; s->ctl_h[(size_t)q * 2 + 1] = (float)o[17]; }
    }
    return 0;
}

// The estimate at (m, tau); phi_1(tau H_m) is evaluated unless c.phi already holds it
// (the caches are dropped whenever H changes).
double krylov_est_main(ScState& c, int m, double tau, int ldh, int ldphi)
{
    if (!(c.phi_m == m && c.phi_tau == tau)) {
        Tic tic(dense_acc());
        c.phi_m = -1;
        if (phi1_pade(c.phi.data(), c.H.data(), tau, m, ldh, ldphi, 1, true)) return INFINITY;
        c.phi_tau = tau;
        c.phi_m = m;
    }
    return tau * (tau * c.beta * c.hn * std::fabs(c.phi[m - 1])) * c.wnext;
}

// Arnoldi over the active scenarios in s->mask_h.  basis 0: main (V, KV, H), 1: error (W, KW, HD).
//
// The host checks the stopping rule at every Krylov dimension m = 1, 2, ... in order and
// stops at the first m that passes (or at the scenario's cap), as the reference's rule
// does.  On the plain path it syncs after every iteration.  With TSR_PATH_DEFERRED the
// iterations up to a predicted dimension (learned, or the last step's, minus two) are
// enqueued back to back and the rule is then evaluated for every m of that window in
// order; iterations beyond the first passing m are wasted work, not a different result.
// Returns 0, or 2 on device error (an RHS failure sets the scenario's failed flag).
// Parts of scenario q's window in the mapped buffer (layout of k_pack_window):
// 0 packed H columns, 1 cs[0..jt], 2 nwj / wrj (jt each), 3 bad.
const double* window_part(const tsr_solver* s, int q, int c0, int jt, int part)
{
    const double* o = s->mb_h + (size_t)q * s->mb_stride;
    const long long nh = (long long)(jt - c0) * (c0 + jt + 3) / 2;   // sum over col of (col + 2)
    if (part == 0) return o;
    if (part == 1) return o + nh;
    if (part == 2) return o + nh + jt + 1;
    return o + nh + jt + 1 + 2 * jt;
}

int read_window(tsr_solver* s, KrylovState& K, int c0, int jt)
{
    const int fail_code = 2;
    HIPCHK(launch_pack_window(K, c0, jt, s->bad, s->mask, s->S, s->mb_d, s->mb_stride, s->st));
    HIPCHK(timed_sync(s));
    for (int q = 0; q < s->S; ++q)
        if (s->mask_h[q]) s->bad_h[q] = window_part(s, q, c0, jt, 3)[0] != 0.0;
    return 0;
}

int arnoldi(tsr_solver* s, int which)
{
    const int fail_code = 2;
    const int S = s->S;
    std::vector<double*>& B = which == 0 ? s->V : s->W;
    const double** Bd = which == 0 ? s->Vd : s->Wd;
    std::vector<float*>& B32 = which == 0 ? s->V32 : s->W32;
    const float** B32d = which == 0 ? s->V32d : s->W32d;
    const bool f32 = (s->path & TSR_PATH_F32BASIS) != 0;
    KrylovState& K = which == 0 ? s->KV : s->KW;
    const int ldh = s->mcap + 1;
    const bool dc = (s->path & TSR_PATH_DCGS2) != 0;
    const bool defer = (s->path & TSR_PATH_DEFERRED) != 0;
    const bool monotone = (s->path & TSR_PATH_MONOTONE) != 0;
    const double at = s->opt.AbsTol, rt = s->opt.RelTol;
    std::vector<int> cap(S, 0), next(S, 0);
    for (int q = 0; q < S; ++q) {
        ScState& c = s->sc[q];
        if (which == 0) { c.m = 0; c.bd = false; c.conv = false; c.phi_m = -1; } else { c.mD = 0; c.bdD = false; c.ph3_m = -1; }
        if (!s->mask_h[q]) continue;
        cap[q] = std::min(which == 0 ? c.m_cap : s->opt.MaxKrylov, s->mcap);
        const int pred = which == 0 ? c.pred_m : c.pred_mD;
        next[q] = defer ? std::min(std::max(pred - 2, 1), cap[q]) : 1;
        if (cap[q] < 1) s->mask_h[q] = 0;
    }
    int j = 0;   // iterations enqueued so far: raw columns 1..j exist
    while (true) {
        int jt = INT_MAX;
        for (int q = 0; q < S; ++q)
            if (s->mask_h[q]) jt = std::min(jt, next[q]);
        if (jt == INT_MAX) break;
        // (the bad flags were cleared by the previous readback)
        if (upload_mask(s)) return 2;
        const int j0 = j;
        // TSR_PATH_GRAPH: the window's launches (every iteration and the readback pack)
        // are captured once per key and replayed; the stream then holds one graph launch
        // instead of several kernels per iteration.  The micro-cell kernels must be tuned
        // already (tuning times candidates with events), the RHS must be the device's, and
        // the kernel timer is off while capturing.
        const bool graph = (s->path & TSR_PATH_GRAPH) && !s->host_g && gfunc_tuned(s->P, s->path & TSR_PATH_JVLIN);
        const std::tuple<int, int, int, int, const double*> gkey{which, j0, jt, s->path, s->u};
        hipGraphExec_t gexec = nullptr;
        GraphEntry* gent = nullptr;
        bool gtimed = false;
        CaptureGuard capture(s->st);
        if (graph) {
            gent = &s->graphs[gkey];
            gtimed = g_kt && gent->replays++ % kGraphTimedEvery == 0;
            gexec = gtimed ? gent->exec_t : gent->exec;
            if (!gexec) HIPCHK(capture.begin(gtimed));
        }
        if (gexec) j = jt;   // replayed below
        for (; j < jt; ++j) {
            if (G_call(s, B[j], K.veps, s->g0, K.ieps, B[j + 1], s->mask)) return 2;
            if (dc && f32) {
                HIPCHK(launch_dcgs2_f32(j + 1, B32d, B[j], B32[j], B[j + 1], s->u, K, s->M, s->nblk, s->part, s->ldp,
                                        s->blk_first, S, s->scen_rows, s->mask, at, rt, s->st));
            } else if (dc) {
                HIPCHK(launch_dcgs2(j + 1, Bd, B[j], B[j + 1], s->u, K, s->M, s->nblk, s->part, s->ldp, s->blk_first, S,
                                    s->scen_rows, s->mask, at, rt, s->st));
            } else {
                HIPCHK(launch_pass(PASS_DOTS, Bd, j + 1, B[j + 1], s->u, K.coef, K.ld, s->M, s->nblk, s->part, s->ldp,
                                   s->mask, at, rt, s->st));
                HIPCHK(launch_reduce(RED_H1, s->part, s->ldp, j + 1, s->blk_first, S, s->scen_rows, K, j, s->mask, s->st));
                HIPCHK(launch_pass(PASS_AXPY_DOTS, Bd, j + 1, B[j + 1], s->u, K.coef, K.ld, s->M, s->nblk, s->part,
                                   s->ldp, s->mask, at, rt, s->st));
                HIPCHK(launch_reduce(RED_H2, s->part, s->ldp, j + 1, s->blk_first, S, s->scen_rows, K, j, s->mask, s->st));
                HIPCHK(launch_pass(PASS_AXPY_NORM, Bd, j + 1, B[j + 1], s->u, K.coef, K.ld, s->M, s->nblk, s->part,
                                   s->ldp, s->mask, at, rt, s->st));
            }
            if (!dc)
                HIPCHK(launch_reduce(RED_NORM, s->part, s->ldp, 2, s->blk_first, S, s->scen_rows, K, j, s->mask, s->st));
        }
        // FP32 copy of the window's last column as it stands (the next window's first
        // iteration corrects it and rewrites the copy; a scenario that stops here uses it)
        if (dc && f32 && !gexec) HIPCHK(launch_to_f32(B32[jt], B[jt], s->M, s->nblk, s->mask, s->st));
        // H columns c0..jt-1 changed in this window (DCGS2 also corrects column j0-1)
        const int c0 = dc ? std::max(j0 - 1, 0) : j0;
        if (graph) {
            GraphTiming& gtime = gtimed ? gent->gt_t : gent->gt;
            if (!gexec) {
                HIPCHK(launch_pack_window(K, c0, jt, s->bad, s->mask, s->S, s->mb_d, s->mb_stride, s->st));
                hipGraph_t gr = nullptr;
                GraphTiming gt;
                HIPCHK(capture.end(&gr, &gt));
                const hipError_t ie = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                HIPCHK(ie);
                (gtimed ? gent->exec_t : gent->exec) = gexec;
                gtime = std::move(gt);
                s->graph_captures++;
            }
            HIPCHK(hipGraphLaunch(gexec, s->st));
            s->graph_launches++;
            if (g_kt) {   // this replay's timer pairs (complete at the sync below) and launches
                for (const auto& p : gtime.pairs) g_kt->pend.push_back(p);
                for (int c = 0; c < KT_COUNT; ++c) g_kt->seen[c] += gtime.seen[c];
            }
            HIPCHK(timed_sync(s));
            for (int q = 0; q < s->S; ++q)
                if (s->mask_h[q]) s->bad_h[q] = window_part(s, q, c0, jt, 3)[0] != 0.0;
        } else if (read_window(s, K, c0, jt)) {
            return 2;
        }
        // the stopping rule per scenario (its phi evaluations on the host pool for a batch)
        const std::vector<int> chk = s->mask_h;
        for_scenarios(chk, [&](int q) {
            ScState& c = s->sc[q];
            c.fe += jt - j0;
            if (s->bad_h[q]) { c.failed = true; s->mask_h[q] = 0; return; }
            std::vector<double>& Hq = which == 0 ? c.H : c.HD;
            std::vector<double>& csq = which == 0 ? c.cs : c.csW;
            if (which == 0) c.phi_m = -1; else c.ph3_m = -1;   // H changes below
            // column col holds rows 0..col+1; the rows below are zero (Hessenberg)
            const double* hw = window_part(s, q, c0, jt, 0);
            for (int col = c0; col < jt; ++col) {
                double* hc = &Hq[(size_t)col * ldh];
                for (int i = 0; i <= col + 1; ++i) hc[i] = hw[i];
                hw += col + 2;
                std::fill(hc + col + 2, hc + ldh, 0.0);
            }
            std::memcpy(csq.data(), window_part(s, q, c0, jt, 1), sizeof(double) * (jt + 1));
            const double* nw = window_part(s, q, c0, jt, 2);
            const double* wr = nw + jt;
            // the stopping rule at dimension m (state for m set in c as a side effect)
            auto passes = [&](int m, bool& bd) {
                const double hn = Hq[(size_t)(m - 1) * ldh + m];
                const double wnorm0 = std::sqrt(nw[m - 1]);
                const double wn = wr[m - 1] / hn;
                bd = !(hn > 1e-12 * wnorm0) || hn == 0.0;
                if (which == 0) {
                    c.m = m; c.hn = hn; c.wnext = wn;
                    if (bd) return true;
                    return m >= s->opt.MinKrylov && krylov_est_main(c, m, c.tau, ldh, s->mcap) <= kKrylovTol;
                }
                c.mD = m; c.hnD = hn; c.wnextD = wn;
                if (bd) return true;
                if (m < s->opt.MinKrylov) return false;
                if (!(c.ph3_m == m && c.ph3_tau == c.tau)) {
                    c.ph3.assign(3 * (size_t)m, 0.0);
                    c.ph3_m = -1;
                    Tic tic(dense_acc());
                    if (phik_e1(c.ph3.data(), c.HD.data(), ldh, m, c.tau, 3, true)) return false;
                    c.ph3_m = m;
                    c.ph3_tau = c.tau;
                }
                return 2.0 * c.tau * c.betaD * (c.tau * hn * std::fabs(c.ph3[2 * m + m - 1])) * wn <= kKrylovTol;
            };
            auto finish = [&](int m, bool bd) {   // stop at m
                bool b2;
                passes(m, b2);
                if (which == 0) { c.bd = bd; c.conv = true; } else { c.bdD = bd; }
            };
            auto c_m_of = [](const ScState& cc, int w) { return w == 0 ? cc.m : cc.mD; };
            auto finish_flags = [](ScState& cc, int w) { if (w == 0) { cc.bd = false; cc.conv = true; } else { cc.bdD = false; } };
            bool stop = false;
            if (monotone && jt - j0 > 1) {
                // breakdowns first (cheap), then the rule at the window's end; only if it
                // passes, scan back for the smallest m from which it passes up to jt
                int mb = 0;
                for (int m = j0 + 1; m <= jt && !mb; ++m) {
                    const double hn = Hq[(size_t)(m - 1) * ldh + m];
                    if (!(hn > 1e-12 * std::sqrt(nw[m - 1])) || hn == 0.0) mb = m;
                }
                bool bd;
                if (mb) {
                    finish(mb, true);
                    stop = true;
                } else if (passes(jt, bd)) {
                    int m0 = jt;
                    while (m0 - 1 > j0 && passes(m0 - 1, bd)) --m0;
                    if (m0 != c_m_of(c, which)) finish(m0, false);   // the scan left c at m0 - 1
                    else finish_flags(c, which);
                    stop = true;
                }   // else: c holds the window's end (not converged)
            } else {
                for (int m = j0 + 1; m <= jt && !stop; ++m) {
                    bool bd;
                    if (passes(m, bd)) {
                        if (which == 0) { c.bd = bd; c.conv = true; } else { c.bdD = bd; }
                        stop = true;
                    }
                }
            }
            if (stop || jt >= cap[q]) s->mask_h[q] = 0;
            else next[q] = jt + 1;
        });
    }
    return 0;
}

int features_and_proposal(tsr_solver* s, const std::vector<int>& act);
void build_features(const ScState& c, float* f);

}  // namespace

// ------------------------------------------------------------------------------------
// One accepted step for every scenario with t < tend.
// ------------------------------------------------------------------------------------
static void fail_scen(const char* what, int q)
{
    set_error(std::string(what) + " (scenario " + std::to_string(q) + ")");
}

static int solver_step(tsr_solver* s, double tend, tsr_step_record* rec)
{
    const int fail_code = -1;
    const exprem_options_struct& o = s->opt;
    const int S = s->S, ldh = s->mcap + 1;
    std::vector<int> act(S, 0);
    int nact = 0;
    if (rec) std::memset(rec, 0, sizeof(tsr_step_record) * S);   // finished scenarios: all-zero records
    for (int q = 0; q < S; ++q) {
        ScState& c = s->sc[q];
        c.failed = false; c.accepted = false; c.rejects = 0; c.fe = 0; c.m = 0; c.mD = 0; c.err = 0.0; c.ctrl = 0;
        if (c.t < tend) { act[q] = 1; ++nact; }
        if (c.stats.NumSteps >= o.MaxNumSteps) { set_error("MaxNumSteps reached"); return -1; }
    }
    if (!nact) return 0;
    s->mask_h = act;
    if (upload_mask(s) || clear_bad(s)) return -1;
    if (G_call(s, nullptr, nullptr, nullptr, nullptr, s->g0, s->mask)) return -1;
    if (launch_norms(s->g0, s->u, s->M, s->nblk, s->part, s->ldp, s->mask, o.AbsTol, o.RelTol, s->st) != hipSuccess)
        return -1;
    if (launch_reduce(RED_START, s->part, s->ldp, 3, s->blk_first, S, s->scen_rows, s->KV, 0, s->mask, s->st) != hipSuccess)
        return -1;
    HIPCHK(launch_macro_stats(batch_of(s->P).scen, S, s->u, s->mstats, s->st));
    // learned controller: its inference is enqueued here (the macro-state features come
    // from the summary just launched, the rest from the host's step history) and its
    // output comes back with the step-start readback
    const bool learned = s->ctrl_mode == 1 && s->mlp && mlp_ready(s->mlp);
    if (learned) {
        std::vector<float> feat((size_t)S * 16, 0.0f);
        for (int q = 0; q < S; ++q)
            if (act[q]) build_features(s->sc[q], &feat[(size_t)q * 16]);
        if (mlp_infer_device(s->mlp, feat.data(), S, s->mstats, s->st)) { set_error("controller inference failed"); return -1; }
    }
    std::vector<double> eps_h(S);
    if (read_iteration(s, s->KV, eps_h.data(), true, learned ? mlp_output(s->mlp) : nullptr)) return -1;
    // proposal: the classical law's (kept in c.proposal), replaced by the learned
    // controller's when it is active and its output is usable
    if (features_and_proposal(s, act)) return -1;
    for (int q = 0; q < S; ++q) {
        if (!act[q]) continue;
        ScState& c = s->sc[q];
        if (c.ctrl != 1 && c.ctrl != 3) { c.m_cap = o.MaxKrylov; }
        c.proposal = std::fmin(std::fmax(c.proposal, o.MinStep), o.MaxStep);
        c.tstop = tend;
        c.tau = c.proposal;
        c.tau_prop = c.proposal;
        c.landing = false;
        c.first_done = false;
        if (c.t + c.tau >= c.tstop - 1e-13 * std::fabs(c.tstop)) { c.tau = c.tstop - c.t; c.landing = true; }
    }
    for (int q = 0; q < S; ++q) {
        if (!act[q]) continue;
        ScState& c = s->sc[q];
        c.fe++;
        if (s->bad_h[q]) { fail_scen("RHS not finite at the step start", q); return -1; }
        c.beta = s->res_h[(size_t)q * 8 + 0];
        c.eps = eps_h[q];
        c.cs[0] = 1.0 / c.beta;
    }
    // main Arnoldi over scenarios with beta > 0
    for (int q = 0; q < S; ++q) s->mask_h[q] = act[q] && s->sc[q].beta > 0.0;
    if (arnoldi(s, 0)) return -1;
    {
        // the Krylov budget per scenario (host pool for a batch); the first failing
        // scenario in order reports
        std::vector<int> why(S, 0);
        for_scenarios(act, [&](int q) {
            ScState& c = s->sc[q];
            if (c.failed) { why[q] = 1; return; }
            if (c.beta > 0.0 && !c.bd) {
                while (krylov_est_main(c, c.m, c.tau, ldh, s->mcap) > kKrylovTol) {
                    c.tau *= 0.5;
                    c.landing = false;
                    if (c.tau < o.MinStep) { why[q] = 2; return; }
                }
            }
        });
        for (int q = 0; q < S; ++q) {
            if (why[q] == 1) { fail_scen("RHS not finite in the Krylov basis", q); return -1; }
            if (why[q] == 2) { fail_scen("Krylov budget: step below MinStep", q); return -1; }
        }
    }
    // attempts
    std::vector<int> pend = act;
    for (int attempt = 0; attempt < 200; ++attempt) {
        int np = 0;
        for (int q = 0; q < S; ++q) np += pend[q];
        if (!np) break;
        // phi_1(tau H) of every pending scenario whose (m, tau) changed, on the host pool
        // for a batch (a failed evaluation is repeated below, which reports it)
        {
            std::vector<int> need(S, 0);
            for (int q = 0; q < S; ++q) {
                const ScState& c = s->sc[q];
                need[q] = pend[q] && c.beta > 0.0 && !(c.phi_m == c.m && c.phi_tau == c.tau);
            }
            for_scenarios(need, [&](int q) {
                ScState& c = s->sc[q];
                Tic tic(dense_acc());
                c.phi_m = -1;
                if (phi1_pade(c.phi.data(), c.H.data(), c.tau, c.m, ldh, s->mcap, 1, true)) return;
                c.phi_tau = c.tau;
                c.phi_m = c.m;
            });
        }
        // u2 = u + V y
        for (int q = 0; q < S; ++q) {
            ScState& c = s->sc[q];
            double* cf = &s->coef_h[(size_t)q * s->KV.ld];
            std::fill(cf, cf + s->KV.ld, 0.0);
            s->nk_h[q] = 0;
            if (!pend[q] || !(c.beta > 0.0)) continue;
            if (!(c.phi_m == c.m && c.phi_tau == c.tau)) {
                Tic tic(dense_acc());
                c.phi_m = -1;
                if (phi1_pade(c.phi.data(), c.H.data(), c.tau, c.m, ldh, s->mcap, 1, true)) { set_error("phipade: singular Pade denominator"); return -1; }
                c.phi_tau = c.tau;
                c.phi_m = c.m;
            }
            for (int k = 0; k < c.m; ++k) { c.y[k] = c.tau * c.beta * c.phi[k]; cf[k] = c.cs[k] * c.y[k]; }
            s->nk_h[q] = c.m;
        }
        // the coefficients of V_{m+1} Hbar y (for D = g2 - g0 - V_{m+1} Hbar y): known now,
        // so that combination is formed in the same pass over the basis as u2
        for (int q = 0; q < S; ++q) {
            ScState& c = s->sc[q];
            double* cf = &s->coef2_h[(size_t)q * s->KV.ld];
            std::fill(cf, cf + s->KV.ld, 0.0);
            s->nk2_h[q] = 0;
            if (!pend[q] || !(c.beta > 0.0)) continue;
            const int mv = c.bd ? c.m : c.m + 1;
            for (int i = 0; i < mv; ++i) {
                double acc = 0.0;
                for (int k = 0; k < c.m; ++k) acc += c.H[i + (size_t)k * ldh] * c.y[k];
                cf[i] = c.cs[i] * acc;
            }
            s->nk2_h[q] = mv;
        }
        s->mask_h = pend;
        if (upload_mask(s)) return -1;
        HIPCHK(hipMemcpyAsync(s->KV.coef, s->coef_h.data(), sizeof(double) * S * s->KV.ld, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipMemcpyAsync(s->nk, s->nk_h.data(), sizeof(int) * S, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipMemcpyAsync(s->coef2, s->coef2_h.data(), sizeof(double) * S * s->KV.ld, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipMemcpyAsync(s->nk2, s->nk2_h.data(), sizeof(int) * S, hipMemcpyHostToDevice, s->st));
        // u2 = u + V y and p2 = V_{m+1} Hbar y
        if (s->path & TSR_PATH_F32BASIS)
            HIPCHK(launch_lincomb2(s->u2, s->u, s->p2, s->V32d, s->nk, s->KV.coef, s->nk2, s->coef2, s->KV.ld, s->M,
                                   s->nblk, s->mask, s->st));
        else
            HIPCHK(launch_lincomb2(s->u2, s->u, s->p2, s->Vd, s->nk, s->KV.coef, s->nk2, s->coef2, s->KV.ld, s->M,
                                   s->nblk, s->mask, s->st));
        if (G_at(s, s->u2, s->g2, s->mask)) return -1;   // bad flags cleared by the last readback
        // D = g2 - p2 - g0 (the one-column combination {p2} with coefficient 1)
        HIPCHK(launch_lincomb(s->D, s->g2, s->p2col, s->one_nk, s->one_coef, 1, -1.0, s->M, s->nblk, s->part, s->ldp, 1,
                              s->u, o.AbsTol, o.RelTol, s->g0, s->mask, s->st));
        HIPCHK(launch_reduce(RED_SUM, s->part, s->ldp, 1, s->blk_first, S, s->scen_rows, s->KW, 0, s->mask, s->st));
        if (read_iteration(s, s->KW)) return -1;
        std::vector<int> okq(S, 0), dq(S, 0);
        for (int q = 0; q < S; ++q) {
            if (!pend[q]) continue;
            ScState& c = s->sc[q];
            c.fe++;
            okq[q] = !s->bad_h[q];
            c.betaD = std::sqrt(s->res_h[(size_t)q * 8 + 0]);
            if (okq[q] && c.betaD > 0.0 && std::isfinite(c.betaD)) dq[q] = 1;
            else if (okq[q] && !(c.betaD == 0.0)) okq[q] = 0;
            c.mD = 0;
        }
        // error-estimate Arnoldi on D
        int anyD = 0;
        for (int q = 0; q < S; ++q) anyD |= dq[q];
        if (anyD) {
            for (int q = 0; q < S; ++q) {
                ScState& c = s->sc[q];
                if (!dq[q]) continue;
                c.csW[0] = 1.0 / c.betaD;
            }
            // csW[0] and veps for the D basis (column 0 of every scenario's cs row;
            // scenarios without a D basis get 0 and take no part)
            std::vector<double> cs0(S, 0.0), vep(S, 0.0);
            for (int q = 0; q < S; ++q) {
                if (!dq[q]) continue;
                cs0[q] = s->sc[q].csW[0];
                vep[q] = s->sc[q].eps * s->sc[q].csW[0];
            }
            HIPCHK(hipMemcpy2DAsync(s->KW.cs, sizeof(double) * s->KW.ld, cs0.data(), sizeof(double), sizeof(double), S,
                                    hipMemcpyHostToDevice, s->st));
            HIPCHK(hipMemcpyAsync(s->KW.veps, vep.data(), sizeof(double) * S, hipMemcpyHostToDevice, s->st));
            s->mask_h = dq;
            if (arnoldi(s, 1)) return -1;
            for (int q = 0; q < S; ++q)
                if (dq[q] && s->sc[q].failed) { fail_scen("RHS not finite in the error-estimate basis", q); return -1; }
        }
        // e = 2 tau betaD W phi3(tau HD) e1; unew = u2 + e; err = WRMS(e).  The phi
        // evaluations first (host pool for a batch), then the coefficients in order.
        // (the stopping rule's evaluation at (mD, tau) is reused when it is the last one)
        std::vector<int> want3(S, 0), prc3(S, 0);
        for (int q = 0; q < S; ++q) {
            const ScState& c = s->sc[q];
            want3[q] = pend[q] && dq[q] && !(c.ph3_m == c.mD && c.ph3_tau == c.tau);
        }
        for_scenarios(want3, [&](int q) {
            ScState& c = s->sc[q];
            c.ph3.assign(3 * (size_t)c.mD, 0.0);
            c.ph3_m = -1;
            Tic tic(dense_acc());
            prc3[q] = phik_e1(c.ph3.data(), c.HD.data(), ldh, c.mD, c.tau, 3, true);
            if (!prc3[q]) { c.ph3_m = c.mD; c.ph3_tau = c.tau; }
        });
        for (int q = 0; q < S; ++q) {
            ScState& c = s->sc[q];
            double* cf = &s->coef_h[(size_t)q * s->KW.ld];
            std::fill(cf, cf + s->KW.ld, 0.0);
            s->nk_h[q] = 0;
            if (!pend[q] || !dq[q]) continue;
            if (prc3[q]) { set_error("phi_3 evaluation failed"); return -1; }
            const std::vector<double>& ph = c.ph3;
            for (int k = 0; k < c.mD; ++k) { c.z[k] = 2.0 * c.tau * c.betaD * ph[2 * c.mD + k]; cf[k] = c.csW[k] * c.z[k]; }
            s->nk_h[q] = c.mD;
        }
        s->mask_h = pend;
        if (upload_mask(s)) return -1;
        HIPCHK(hipMemcpyAsync(s->KW.coef, s->coef_h.data(), sizeof(double) * S * s->KW.ld, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipMemcpyAsync(s->nk, s->nk_h.data(), sizeof(int) * S, hipMemcpyHostToDevice, s->st));
        if (s->path & TSR_PATH_F32BASIS)
            HIPCHK(launch_lincomb(s->unew, s->u2, s->W32d, s->nk, s->KW.coef, s->KW.ld, 1.0, s->M, s->nblk, s->part,
                                  s->ldp, 2, s->u, o.AbsTol, o.RelTol, nullptr, s->mask, s->st));
        else
            HIPCHK(launch_lincomb(s->unew, s->u2, s->Wd, s->nk, s->KW.coef, s->KW.ld, 1.0, s->M, s->nblk, s->part,
                                  s->ldp, 2, s->u, o.AbsTol, o.RelTol, nullptr, s->mask, s->st));
        HIPCHK(launch_reduce(RED_SUM, s->part, s->ldp, 1, s->blk_first, S, s->scen_rows, s->KW, 0, s->mask, s->st));
        if (read_iteration(s, s->KW)) return -1;
        std::vector<int> acc(S, 0);
        int nacc = 0;
        for (int q = 0; q < S; ++q) {
            if (!pend[q]) continue;
            ScState& c = s->sc[q];
            double err = dq[q] ? std::sqrt(s->res_h[(size_t)q * 8 + 0] / (double)s->rows_h[q]) : 0.0;
            c.err = err;
            if (!c.first_done) {
                c.first_done = true;
                c.tau_first = c.tau;
                c.err_first = okq[q] ? err : INFINITY;
            }
            if (okq[q] && std::isfinite(err) && err <= 1.0) { acc[q] = 1; ++nacc; continue; }
            c.rejects++;
            const double fac = (okq[q] && std::isfinite(err)) ? std::fmax(o.MinStepDecFac, o.SafetyFac * std::pow(err, -1.0 / 3.0))
                                                              : 0.5;
            c.tau *= fac;
            c.landing = false;
            if (c.tau < o.MinStep) { fail_scen("error control: step below MinStep", q); return -1; }
        }
        if (nacc) {
            if (S == 1) {
                std::swap(s->u, s->unew);
            } else {
                s->mask_h = acc;
                if (upload_mask(s)) return -1;
                HIPCHK(launch_masked_copy(s->u, s->unew, s->M, s->nblk, s->mask, s->st));
            }
            for (int q = 0; q < S; ++q) if (acc[q]) pend[q] = 0;
        }
    }
    for (int q = 0; q < S; ++q)
        if (pend[q]) { fail_scen("no acceptable step after 200 attempts", q); return -1; }
    // bookkeeping
    for (int q = 0; q < S; ++q) {
        if (!act[q]) continue;
        ScState& c = s->sc[q];
        const double taken = c.tau;
        c.t = c.landing ? c.tstop : c.t + taken;
        const double fac = (c.err > 0.0) ? std::fmin(o.MaxStepIncFac, std::fmax(o.MinStepDecFac, o.SafetyFac * std::pow(c.err, -1.0 / 3.0)))
                                         : o.MaxStepIncFac;
        double next = taken * fac;
        if (c.landing && taken < c.proposal) next = std::fmax(next, c.proposal);
        c.proposal = std::fmin(std::fmax(next, o.MinStep), o.MaxStep);
        c.stats.CurrentStep = c.proposal;
        c.stats.CurrentTime = c.t;
        c.stats.CurrentKryDim = c.m;
        c.stats.NumSteps += 1;
        c.stats.FuncEvals += c.fe;
        double hnorm = 0.0;
        for (int i = 0; i < c.m; ++i) {
            double rs = 0.0;
            for (int k = 0; k < c.m; ++k) rs += std::fabs(c.H[i + (size_t)k * ldh]);
            hnorm = std::max(hnorm, rs);
        }
        for (int k = 3; k > 0; --k) { c.hist_tau[k] = c.hist_tau[k - 1]; c.hist_m[k] = c.hist_m[k - 1]; }
        c.hist_tau[0] = taken;
        c.hist_m[0] = c.m;
        c.pred_m = c.m;
        c.pred_mD = c.mD;
        c.n_hist = std::min(c.n_hist + 1, 4);
        c.last_err = c.err;
        c.last_hnorm = hnorm;
        c.last_resid = c.hn;
        if (rec) {
            tsr_step_record& r = rec[q];
            std::memset(&r, 0, sizeof(r));
            r.t = c.t; r.tau = taken; r.tau_proposed = c.tau_prop; r.err = c.err; r.beta = c.beta;
            r.hnorm = hnorm; r.resid = c.hn; r.m = c.m; r.mD = c.mD; r.rejects = c.rejects; r.fevals = c.fe;
            r.controller = c.ctrl; r.m_cap = c.m_cap;
            r.head_mean = c.mstat[0]; r.head_min = c.mstat[1]; r.head_max = c.mstat[2];
            r.sat_mean = c.mstat[3]; r.sat_min = c.mstat[4]; r.sat_max = c.mstat[5];
            r.tau_first = c.tau_first; r.err_first = c.err_first;
            r.tau_next = c.proposal;
            std::memcpy(r.feat, c.feat, sizeof(r.feat));
        }
    }
    s->mask_h.assign(S, 1);
    return 0;
}

// frees a solver's device and host resources (it must not be in the solver cache)
static void release_solver(tsr_solver* s);

namespace {

// Controller inputs (16 per scenario, order of TSR_FEAT_COUNT in tsr.h) and, with the
// learned controller active, its proposal: step = 10^y0, at most the classical proposal
// (which already carries the error-based growth limit) and at least 1/kLearnedTrust of
// it; Krylov cap kLearnedCap; y1 (the dimension the step is expected to need) as the
// window of the deferred stopping-rule checks.  A scenario whose network output is not
// finite keeps the classical proposal (ctrl = 2).
constexpr double kLearnedTrust = 8.0;

void build_features(const ScState& c, float* f)
{
    const double lp = std::log10(std::max(c.proposal, 1e-300));
    f[0] = (float)lp;
    double last = lp;
    for (int k = 0; k < 3; ++k) {
        if (k < c.n_hist) last = std::log10(std::max(c.hist_tau[k], 1e-300));
        f[1 + k] = (float)last;
    }
    double lm = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (k < c.n_hist) lm = c.hist_m[k];
        f[4 + k] = (float)lm;
    }
    f[8] = c.n_hist ? (float)std::log10(std::max(c.last_hnorm, 1e-300)) : 0.0f;
    f[9] = c.n_hist ? (float)std::log10(std::max(c.last_resid, 1e-300)) : 0.0f;
    for (int k = 0; k < 6; ++k) f[10 + k] = (float)c.mstat[k];
}

// Exploration controller (training-data generation): a random walk of the step size
// that keeps the Krylov dimension of the last step inside [kExploreMlo, kExploreMhi]
// (x0.7 above, x1.3 below, a log-uniform factor in [e^-0.3, e^0.3] inside), never
// above the classical error-based step, with the learned controller's Krylov cap.
constexpr int kExploreMlo = 3, kExploreMhi = 24;

double explore_uniform(unsigned long long& st)
{
    unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int features_and_proposal(tsr_solver* s, const std::vector<int>& act)
{
    const int S = s->S;
    for (int q = 0; q < S; ++q) {
        s->sc[q].ctrl = 0;
        if (act[q]) build_features(s->sc[q], s->sc[q].feat);
    }
    if (s->ctrl_mode == 2) {
        const exprem_options_struct& o = s->opt;
        for (int q = 0; q < S; ++q) {
            if (!act[q]) continue;
            ScState& c = s->sc[q];
            if (!c.n_hist) continue;   // first step: classical proposal
            const double tp = c.hist_tau[0];
            const double mp = c.hist_m[0];
            double f;
            if (mp > kExploreMhi) f = 0.7;
            else if (mp < kExploreMlo) f = 1.3;
            else f = std::exp(0.6 * explore_uniform(c.rng) - 0.3);
            double tl = tp * f;
            if (c.last_err > 0.0) tl = std::fmin(tl, tp * o.SafetyFac * std::pow(c.last_err, -1.0 / 3.0));
            c.proposal = tl;
            c.m_cap = std::min(kLearnedCap, s->mcap);
            c.ctrl = 3;
        }
        return 0;
    }
    if (!(s->ctrl_mode == 1 && s->mlp && mlp_ready(s->mlp))) return 0;
    // the network's output for these features was computed on the device and read back
    // with the step-start results (solver_step)
    const std::vector<float>& out = s->ctl_h;
    const int hi_cap = std::min(kLearnedCap, s->mcap);
    for (int q = 0; q < S; ++q) {
        if (!act[q]) continue;
        ScState& c = s->sc[q];
        const double y0 = out[(size_t)q * 2], y1 = out[(size_t)q * 2 + 1];
        if (std::isfinite(y0) && std::isfinite(y1) && std::fabs(y0) < 300.0) {
            const double tl = std::pow(10.0, y0);
            c.proposal = std::fmin(std::fmax(tl, c.proposal / kLearnedTrust), c.proposal);
            c.m_cap = hi_cap;
            c.pred_m = std::min(std::max((int)std::lround(std::fmin(std::fmax(y1, 0.0), 1000.0)), 5), hi_cap);
            c.ctrl = 1;
        } else {
            c.ctrl = 2;
        }
    }
    return 0;
}

// The step driver syncs on its stream a few times per Arnoldi window and reacts to the
// result at once: ask for spin-waiting syncs (lower wake-up latency than yielding).
// This only takes effect before the process's first HIP call creates the device
// context; a caller that created it already keeps its own choice.
void prefer_spin_sync()
{
    static std::once_flag once;
    std::call_once(once, [] { (void)hipSetDeviceFlags(hipDeviceScheduleSpin); (void)hipGetLastError(); });
}

// a solver whose set-up fails part way is released like a finished one
struct SolverRelease {
    void operator()(tsr_solver* s) const { release_solver(s); }
};

tsr_solver* create(const tsr_batch_view* view, const exprem_options_struct* opt)
{
    prefer_spin_sync();
    std::string err;
    std::unique_ptr<tsr_solver, SolverRelease> s(new tsr_solver());
    s->opt = *opt;
    if (opt->MaxKrylov < 1 || opt->MaxKrylov > kMaxKrylovCap) { set_error("MaxKrylov must be in [1, 120]"); return nullptr; }
    s->mcap = kMaxKrylovCap;
    s->P = upload_problem(*view, err);
    if (!s->P) { set_error(err); return nullptr; }
    if (hipStreamCreate(&s->st) != hipSuccess) { set_error("hipStreamCreate failed"); return nullptr; }
    s->N = view->N;
    s->S = view->n_scen;
    const size_t N = (size_t)s->N;
    const int S = s->S;
    int rc = 0;
    rc |= dalloc(&s->u, N); rc |= dalloc(&s->g0, N); rc |= dalloc(&s->u2, N); rc |= dalloc(&s->g2, N);
    rc |= dalloc(&s->D, N); rc |= dalloc(&s->unew, N); rc |= dalloc(&s->scratch[0], N); rc |= dalloc(&s->scratch[1], N);
    if (rc) { set_error("hipMalloc failed for the state vectors"); return nullptr; }
    // Krylov bases: column 0 aliases the start vector.  Columns are allocated for the
    // options' MaxKrylov; the learned controller may raise a scenario's cap up to 60,
    // so allocate lazily up to mcap.
    const int ncol = std::max(opt->MaxKrylov, kLearnedCap) + 1;
    s->V.assign(s->mcap + 1, nullptr);
    s->W.assign(s->mcap + 1, nullptr);
    s->V[0] = s->g0;
    s->W[0] = s->D;
    for (int k = 1; k < ncol; ++k) {
        if (dalloc(&s->V[k], N) || dalloc(&s->W[k], N)) { set_error("hipMalloc failed for the Krylov basis"); return nullptr; }
    }
    s->mcap = ncol - 1;
    if (dalloc(&s->Vd, s->mcap + 1) || dalloc(&s->Wd, s->mcap + 1)) return nullptr;
    hipMemcpy(s->Vd, s->V.data(), sizeof(double*) * (s->mcap + 1), hipMemcpyHostToDevice);
    hipMemcpy(s->Wd, s->W.data(), sizeof(double*) * (s->mcap + 1), hipMemcpyHostToDevice);
    // row blocks per scenario
    const long long rpb = (long long)kVecBS * kVecR;
    std::vector<int> bs, bf(S + 1, 0);
    std::vector<long long> b0, b1;
    s->rows_h.assign(S, 0);
    for (int q = 0; q < S; ++q) {
        const long long off = view->scen[q].u_offset, n = view->scen[q].ud.N;
        s->rows_h[q] = n;
        bf[q] = (int)bs.size();
        for (long long r = 0; r < n; r += rpb) { bs.push_back(q); b0.push_back(off + r); b1.push_back(off + std::min(n, r + rpb)); }
    }
    bf[S] = (int)bs.size();
    s->nblk = (int)bs.size();
    int *d_bs = nullptr; long long *d_b0 = nullptr, *d_b1 = nullptr;
    rc = dalloc(&d_bs, bs.size()) | dalloc(&d_b0, b0.size()) | dalloc(&d_b1, b1.size()) | dalloc(&s->blk_first, S + 1) |
         dalloc(&s->scen_rows, S);
    s->M.scen = d_bs; s->M.r0 = d_b0; s->M.r1 = d_b1;
    if (rc) return nullptr;
    hipMemcpy(d_bs, bs.data(), sizeof(int) * bs.size(), hipMemcpyHostToDevice);
    hipMemcpy(d_b0, b0.data(), sizeof(long long) * b0.size(), hipMemcpyHostToDevice);
    hipMemcpy(d_b1, b1.data(), sizeof(long long) * b1.size(), hipMemcpyHostToDevice);
    hipMemcpy(s->blk_first, bf.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice);
    hipMemcpy(s->scen_rows, s->rows_h.data(), sizeof(long long) * S, hipMemcpyHostToDevice);
    s->ldp = s->nblk;   // partials are column-major: part[k * nblk + b], k < mcap + 2
    if (dalloc(&s->part, (size_t)s->nblk * (2 * s->mcap + 4))) return nullptr;   // DCGS2 needs 2 (j+1) + 1 columns
    auto kstate = [&](KrylovState& K, double* shared_eps, double* shared_ieps) {
        K.ld = s->mcap + 2;
        K.ldh = s->mcap + 1;
        K.hstride = (s->mcap + 1) * s->mcap;
        K.ld2 = 2 * K.ld + 2;
        int r = dalloc(&K.cs, (size_t)S * K.ld) | dalloc(&K.coef, (size_t)S * K.ld) | dalloc(&K.H, (size_t)S * K.hstride) |
                dalloc(&K.veps, S) | dalloc(&K.res, (size_t)S * 8) | dalloc(&K.coefr, (size_t)S * K.ld) |
                dalloc(&K.dots, (size_t)S * K.ld2) | dalloc(&K.nwj, (size_t)S * K.ld) | dalloc(&K.wrj, (size_t)S * K.ld);
        if (shared_eps) { K.eps = shared_eps; K.ieps = shared_ieps; }
        else r |= dalloc(&K.eps, S) | dalloc(&K.ieps, S);
        return r;
    };
    if (kstate(s->KV, nullptr, nullptr) || kstate(s->KW, s->KV.eps, s->KV.ieps)) { set_error("hipMalloc failed"); return nullptr; }
    hipMemset(s->KV.cs, 0, sizeof(double) * S * s->KV.ld);
    hipMemset(s->KW.cs, 0, sizeof(double) * S * s->KW.ld);
    if (dalloc(&s->mask, S) || dalloc(&s->nk, S)) return nullptr;
    if (dalloc(&s->coef2, (size_t)S * s->KV.ld) || dalloc(&s->nk2, S) || dalloc(&s->p2, N) || dalloc(&s->p2col, 1) ||
        dalloc(&s->one_coef, S) || dalloc(&s->one_nk, S)) {
        set_error("hipMalloc failed for the step combinations");
        return nullptr;
    }
    {
        const std::vector<double> ones(S, 1.0);
        const std::vector<int> ones_i(S, 1);
        const double* pc = s->p2;
        if (hipMemcpy(s->p2col, &pc, sizeof(pc), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->one_coef, ones.data(), sizeof(double) * S, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->one_nk, ones_i.data(), sizeof(int) * S, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("hipMemcpy failed for the step combinations");
            return nullptr;
        }
    }
    s->bad = batch_of(s->P).bad;
    s->mask_h.assign(S, 1); s->nk_h.assign(S, 0); s->bad_h.assign(S, 0);
    s->coef_h.assign((size_t)S * s->KV.ld, 0.0);
    s->coef2_h.assign((size_t)S * s->KV.ld, 0.0);
    s->nk2_h.assign(S, 0);
    s->res_h.assign((size_t)S * 8, 0.0);
    s->sc.resize(S);
    const int ldh = s->mcap + 1;
    for (auto& c : s->sc) {
        c.H.assign((size_t)ldh * s->mcap, 0.0);
        c.HD.assign((size_t)ldh * s->mcap, 0.0);
        c.phi.assign((size_t)s->mcap * s->mcap, 0.0);
        c.y.assign(s->mcap + 1, 0.0);
        c.z.assign(s->mcap + 1, 0.0);
        c.cs.assign(s->mcap + 2, 0.0);
        c.csW.assign(s->mcap + 2, 0.0);
    }
    // mapped readback buffer: a full window (mcap columns) per scenario, at least 16 doubles
    s->mb_stride = std::max(kPackSmall, s->mcap * (s->mcap + 3) / 2 + (s->mcap + 1) + 2 * s->mcap + 1);
    s->ctl_h.assign((size_t)S * 2, 0.0f);
    s->mb_stride = (s->mb_stride + 7) & ~7;
    if (hipHostMalloc((void**)&s->mb_h, sizeof(double) * (size_t)S * s->mb_stride,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&s->mb_d, s->mb_h, 0) != hipSuccess ||
        dalloc(&s->mstats, (size_t)S * 6)) {
        set_error("hipHostMalloc failed");
        return nullptr;
    }
    hipEventCreate(&s->ev0);
    hipEventCreate(&s->ev1);
    tsr_solver_reset(s.get());
    if (hipDeviceSynchronize() != hipSuccess) { set_error("device error during solver set-up"); return nullptr; }
    return s.release();
}

}  // namespace

namespace tsr {
// A problem is being freed: free the solvers the reference-signature entry points made
// for it, so that a later problem allocated at the same address starts afresh.
void solver_cache_drop(const tsr_batch_view* v)
{
    std::vector<tsr_solver*> drop;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_solver_cache.find(v);
        if (it != g_solver_cache.end()) {
            drop.push_back(it->second);
            g_solver_cache.erase(it);
        }
        for (auto r = g_retired.begin(); r != g_retired.end();) {
            if (r->first == v) {
                drop.push_back(r->second);
                r = g_retired.erase(r);
            } else {
                ++r;
            }
        }
    }
    for (tsr_solver* s : drop) tsr_solver_free(s);
}
}  // namespace tsr

extern "C" {

tsr_solver* tsr_solver_create(const tsr_batch_view* view, const exprem_options_struct* opt)
{
    return create(view, opt);
}

void tsr_solver_free(tsr_solver* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (auto it = g_solver_cache.begin(); it != g_solver_cache.end();)
            it = (it->second == s) ? g_solver_cache.erase(it) : std::next(it);
    }
    release_solver(s);
}

}  // extern "C"

static void release_solver(tsr_solver* s)
{
    if (!s) return;
    if (s->P) hipDeviceSynchronize();   // no kernel ran before the problem was uploaded
    for (double* p : {s->u, s->g0, s->u2, s->g2, s->D, s->unew, s->scratch[0], s->scratch[1]}) if (p) hipFree(p);
    for (size_t k = 1; k < s->V.size(); ++k) { if (s->V[k]) hipFree(s->V[k]); if (s->W[k]) hipFree(s->W[k]); }
    hipFree(s->Vd); hipFree(s->Wd);
    for (auto* B32 : {&s->V32, &s->W32})
        for (float* c : *B32) if (c) hipFree(c);
    hipFree(s->V32d); hipFree(s->W32d);
    hipFree((void*)s->M.scen); hipFree((void*)s->M.r0); hipFree((void*)s->M.r1);
    hipFree(s->blk_first); hipFree(s->scen_rows); hipFree(s->part);
    for (KrylovState* K : {&s->KV, &s->KW}) {
        hipFree(K->cs); hipFree(K->coef); hipFree(K->H); hipFree(K->veps); hipFree(K->res);
        hipFree(K->coefr); hipFree(K->dots); hipFree(K->nwj); hipFree(K->wrj);
    }
    hipFree(s->KV.eps); hipFree(s->KV.ieps);
    hipFree(s->mask); hipFree(s->nk);
    hipFree(s->coef2); hipFree(s->nk2); hipFree(s->p2); hipFree((void*)s->p2col); hipFree(s->one_coef); hipFree(s->one_nk);
    if (s->mb_h) hipHostFree(s->mb_h);
    if (s->linrec) hipFree(s->linrec);
    if (s->mstats) hipFree(s->mstats);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->mlp) mlp_free(s->mlp);
    for (auto& g : s->graphs) {
        if (g.second.exec) hipGraphExecDestroy(g.second.exec);
        if (g.second.exec_t) hipGraphExecDestroy(g.second.exec_t);
    }
    for (hipEvent_t e : s->kt.all) hipEventDestroy(e);
    free_problem(s->P);
    if (s->st) hipStreamDestroy(s->st);
    delete s;
}

extern "C" {

void tsr_solver_reset(tsr_solver* s)
{
    for (auto& c : s->sc) {
        c.t = 0.0;
        c.proposal = s->opt.InitStep;
        exprem_stats_init(&c.stats, &s->opt, nullptr);
        for (int k = 0; k < 4; ++k) { c.hist_tau[k] = 0.0; c.hist_m[k] = 0.0; }
        c.n_hist = 0; c.last_hnorm = 0.0; c.last_resid = 0.0; c.pred_m = 0; c.pred_mD = 0; c.last_err = 0.0;
    }
}

int tsr_solver_set_state(tsr_solver* s, const double* u_host, double t)
{
    const int fail_code = 2;
    HIPCHK(hipMemcpy(s->u, u_host, sizeof(double) * s->N, hipMemcpyHostToDevice));
    for (auto& c : s->sc) c.t = t;
    return 0;
}

int tsr_solver_get_state(tsr_solver* s, double* u_host, double* t)
{
    const int fail_code = 2;
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(u_host, s->u, sizeof(double) * s->N, hipMemcpyDeviceToHost));
    if (t) *t = s->sc[0].t;
    return 0;
}

void tsr_solver_stats(const tsr_solver* s, exprem_stats_struct* st) { *st = s->sc[0].stats; }

int tsr_solver_gfunc_host(tsr_solver* s, double* g, const double* u)
{
    const int fail_code = 2;
    HIPCHK(hipMemcpyAsync(s->scratch[0], u, sizeof(double) * s->N, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(s->bad, 0, sizeof(int) * s->S, s->st));
    if (G_at(s, s->scratch[0], s->scratch[1], nullptr)) return 2;
    HIPCHK(hipMemcpyAsync(g, s->scratch[1], sizeof(double) * s->N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipMemcpyAsync(s->bad_h.data(), s->bad, sizeof(int) * s->S, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (int q = 0; q < s->S; ++q) if (s->bad_h[q]) return 1;
    return 0;
}

int tsr_solver_gfunc_dev(tsr_solver* s, double* g_dev, const double* u_dev, int reps)
{
    const int fail_code = 2;
    if (reps < 1) reps = 1;
    HIPCHK(hipEventRecord(s->ev0, s->st));
    for (int r = 0; r < reps; ++r)
        if (G_at(s, u_dev, g_dev, nullptr)) return 2;
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_g_ms = ms / reps;
    return 0;
}

double tsr_solver_last_gfunc_ms(const tsr_solver* s) { return s->last_g_ms; }

void tsr_solver_micro_shape(const tsr_solver* s, int* bs, int* wpe) { micro_shape(s->P, bs, wpe); }

int tsr_solver_time_gfunc_parts(tsr_solver* s, double* g_dev, const double* u_dev, int reps, double* ms3)
{
    return tsr_solver_time_gfunc_mode(s, g_dev, u_dev, 0, reps, ms3);
}

int tsr_solver_time_gfunc_mode(tsr_solver* s, double* g_dev, const double* u_dev, int mode, int reps, double* ms3)
{
    const int fail_code = 2;
    if (reps < 1) reps = 1;
    if (mode < 0 || mode > 2) return 1;
    GArgs a;
    a.u = u_dev; a.v = nullptr; a.veps = nullptr; a.g0 = nullptr; a.ieps = nullptr; a.active = nullptr; a.out = g_dev;
    double* de = nullptr;
    if (mode > 0) {
        if (!s->linrec && tsr_solver_set_path(s, s->path | TSR_PATH_JVLIN)) return 2;
        a.rec = s->linrec;
        HIPCHK(launch_gfunc(s->P, a, s->st));            // the expansion point (and g0 = G(u))
        HIPCHK(hipMemcpyAsync(s->g0, g_dev, sizeof(double) * (size_t)s->N, hipMemcpyDeviceToDevice, s->st));
        if (mode == 2) {
            // an FD product around u_dev along scratch[0] (zero-filled: the timing does
            // not depend on the direction), eps = 1
            std::vector<double> e(2 * (size_t)s->S, 1.0);
            if (hipMalloc((void**)&de, sizeof(double) * e.size()) != hipSuccess) { set_error("hipMalloc failed"); return 2; }
            HIPCHK(hipMemcpyAsync(de, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, s->st));
            HIPCHK(hipMemsetAsync(s->scratch[0], 0, sizeof(double) * (size_t)s->N, s->st));
            a.rec = nullptr;
            a.lin = s->linrec;
            a.v = s->scratch[0]; a.veps = de; a.g0 = s->g0; a.ieps = de + s->S;
        }
    }
    hipEvent_t ev[4];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double acc[3] = {0, 0, 0};
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(ev[0], s->st));
        for (int part = 0; part < 3; ++part) {
            HIPCHK(launch_gfunc_part(s->P, a, part, s->st));
            HIPCHK(hipEventRecord(ev[part + 1], s->st));
        }
        HIPCHK(hipEventSynchronize(ev[3]));
        for (int part = 0; part < 3; ++part) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev[part], ev[part + 1]));
            acc[part] += ms;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    for (int part = 0; part < 3; ++part) ms3[part] = acc[part] / reps;
    if (de) hipFree(de);
    return 0;
}

int tsr_solver_time_copy(tsr_solver* s, int reps, double* ms)
{
    const int fail_code = 2;
    if (reps < 1) reps = 1;
    HIPCHK(launch_masked_copy(s->scratch[1], s->scratch[0], s->M, s->nblk, nullptr, s->st));
    HIPCHK(hipEventRecord(s->ev0, s->st));
    for (int r = 0; r < reps; ++r) HIPCHK(launch_masked_copy(s->scratch[1], s->scratch[0], s->M, s->nblk, nullptr, s->st));
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms = t / reps;
    return 0;
}

int tsr_solver_ktimer(tsr_solver* s, int every)
{
    if (every < 0) return 1;
    if (hipStreamSynchronize(s->st) == hipSuccess && !s->kt.pend.empty()) kt_harvest(s->kt);
    KTimer& k = s->kt;
    k.every = every;
    for (int c = 0; c < KT_COUNT; ++c) { k.seen[c] = 0; k.timed[c] = 0; k.ms[c] = 0.0; k.bytes[c] = 0.0; k.cap_n[c] = 0; }
    // captured windows carry the timer setting they were captured with: recapture
    for (auto& g : s->graphs) {
        if (g.second.exec) hipGraphExecDestroy(g.second.exec);
        if (g.second.exec_t) hipGraphExecDestroy(g.second.exec_t);
    }
    s->graphs.clear();
    return 0;
}

int tsr_solver_ktimer_read(tsr_solver* s, long long* seen, long long* timed, double* ms, double* bytes)
{
    if (hipStreamSynchronize(s->st) == hipSuccess && !s->kt.pend.empty()) kt_harvest(s->kt);
    for (int c = 0; c < KT_COUNT; ++c) {
        if (seen) seen[c] = s->kt.seen[c];
        if (timed) timed[c] = s->kt.timed[c];
        if (ms) ms[c] = s->kt.ms[c];
        if (bytes) bytes[c] = s->kt.bytes[c];
    }
    return 0;
}

int tsr_host_threads(int n)
{
    if (n > 0) tsr::HostPool::get().set_limit(n);
    else if (n == 0) tsr::HostPool::get().set_limit(0);
    return tsr::HostPool::get().threads();
}

void tsr_solver_graph_stats(const tsr_solver* s, long long* launches, long long* captures)
{
    if (launches) *launches = s->graph_launches;
    if (captures) *captures = s->graph_captures;
}

int tsr_solver_set_step(tsr_solver* s, double tau)
{
    if (!(tau > 0.0) || !std::isfinite(tau)) return 1;
    for (auto& c : s->sc) { c.proposal = tau; c.stats.CurrentStep = tau; }
    return 0;
}

int tsr_solver_step(tsr_solver* s, double tend, tsr_step_record* rec)
{
    KtScope kts(s);
    return solver_step(s, tend, rec);
}

int tsr_solver_run(tsr_solver* s, double tend, int max_steps, tsr_run_stats* out, tsr_step_record* records,
                   int max_records)
{
    KtScope kts(s);
    const auto t0 = std::chrono::steady_clock::now();
    int steps = 0, rej = 0, fe = 0, rc = 0;
    s->t_dense = 0.0;
    s->t_sync = 0.0;
    g_timed = s;
    std::vector<tsr_step_record> rec(s->S);
    for (; steps < max_steps; ++steps) {
        bool done = true;
        for (auto& c : s->sc) done = done && !(c.t < tend);
        if (done) break;
        const auto ts = std::chrono::steady_clock::now();
        rc = solver_step(s, tend, rec.data());
        if (rc) break;
        const double wms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
        for (int q = 0; q < s->S; ++q) rec[q].wall_ms = wms;
        for (int q = 0; q < s->S; ++q) { rej += rec[q].rejects; fe += rec[q].fevals; }
        if (records && steps * s->S < max_records)
            for (int q = 0; q < s->S && steps * s->S + q < max_records; ++q) records[steps * s->S + q] = rec[q];
    }
    if (hipStreamSynchronize(s->st) == hipSuccess && !s->kt.pend.empty()) kt_harvest(s->kt);
    if (out) {
        out->t = s->sc[0].t;
        for (auto& c : s->sc) out->t = std::min(out->t, c.t);
        out->steps = steps; out->rejects = rej; out->fevals = fe; out->failed = rc != 0;
        out->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        out->gfunc_ms = 0.0; out->arnoldi_ms = 0.0;
        out->host_dense_s = s->t_dense;
        out->sync_wait_s = s->t_sync;
    }
    g_timed = nullptr;
    return rc;
}

int tsr_solver_jv(tsr_solver* s, const double* v_host, double eps, int lin, double* jv_host)
{
    const int fail_code = 2;
    if (lin && !s->linrec && tsr_solver_set_path(s, s->path | TSR_PATH_JVLIN)) return 2;
    const size_t N = (size_t)s->N;
    std::vector<double> e(2 * (size_t)s->S);
    for (int q = 0; q < s->S; ++q) { e[q] = eps; e[s->S + q] = 1.0 / eps; }
    double* de = nullptr;
    if (hipMalloc((void**)&de, sizeof(double) * e.size()) != hipSuccess) { set_error("hipMalloc failed"); return 2; }
    HIPCHK(hipMemcpyAsync(de, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->scratch[0], v_host, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
    GArgs a;
    a.u = s->u; a.v = nullptr; a.veps = nullptr; a.g0 = nullptr; a.ieps = nullptr; a.active = nullptr; a.out = s->g0;
    a.rec = lin ? s->linrec : nullptr;
    HIPCHK(launch_gfunc(s->P, a, s->st));
    a.v = s->scratch[0]; a.veps = de; a.g0 = s->g0; a.ieps = de + s->S; a.out = s->scratch[1];
    a.rec = nullptr;
    a.lin = lin ? s->linrec : nullptr;
    HIPCHK(launch_gfunc(s->P, a, s->st));
    HIPCHK(hipMemcpyAsync(jv_host, s->scratch[1], sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    hipFree(de);
    return 0;
}

int tsr_solver_jv_noalloc(tsr_solver* s, const double* v_host, double eps, int lin, double* jv_host)
{
    const int fail_code = 2;
    if (s->S > kNoAllocMaxScen) {
        set_error("tsr_solver_jv_noalloc: at most " + std::to_string(kNoAllocMaxScen) + " scenarios");
        return 1;
    }
    if (lin && !s->linrec) {
        set_error("tsr_solver_jv_noalloc: the linearisation record is not allocated (set TSR_PATH_JVLIN first)");
        return 1;
    }
    if (!(eps != 0.0) || !std::isfinite(eps)) { set_error("tsr_solver_jv_noalloc: eps must be finite and nonzero"); return 1; }
    double* de = noalloc_scratch();
    if (!de) { set_error("tsr_solver_jv_noalloc: device scratch address unavailable"); return 2; }
    const size_t N = (size_t)s->N;
    HIPCHK(launch_fill_eps(de, s->S, eps, s->st));
    HIPCHK(hipMemcpyAsync(s->scratch[0], v_host, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
    GArgs a;
    a.u = s->u; a.v = nullptr; a.veps = nullptr; a.g0 = nullptr; a.ieps = nullptr; a.active = nullptr; a.out = s->g0;
    a.rec = lin ? s->linrec : nullptr;
    HIPCHK(launch_gfunc(s->P, a, s->st, false));
    a.v = s->scratch[0]; a.veps = de; a.g0 = s->g0; a.ieps = de + s->S; a.out = s->scratch[1];
    a.rec = nullptr;
    a.lin = lin ? s->linrec : nullptr;
    HIPCHK(launch_gfunc(s->P, a, s->st, false));
    HIPCHK(hipMemcpyAsync(jv_host, s->scratch[1], sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

int tsr_solver_set_path(tsr_solver* s, int path)
{
    if (path & ~(TSR_PATH_BATCHED | TSR_PATH_F32BASIS | TSR_PATH_GRAPH)) return 1;
    if ((path & TSR_PATH_GRAPH) && !(path & TSR_PATH_DEFERRED)) {
        set_error("TSR_PATH_GRAPH needs TSR_PATH_DEFERRED (a window is what is captured)");
        return 1;
    }
    if ((path & TSR_PATH_F32BASIS) && !(path & TSR_PATH_DCGS2)) {
        set_error("TSR_PATH_F32BASIS needs TSR_PATH_DCGS2");
        return 1;
    }
    if ((path & TSR_PATH_F32BASIS) && !s->V32d) {
        // columns 0..mcap of both bases
        const size_t N = (size_t)s->N;
        for (auto* B32 : {&s->V32, &s->W32}) {
            B32->assign(s->mcap + 1, nullptr);
            for (auto& c : *B32)
                if (hipMalloc((void**)&c, sizeof(float) * (N ? N : 1)) != hipSuccess) {
                    set_error("hipMalloc failed for the FP32 basis");
                    return 2;
                }
        }
        const float** d[2] = {nullptr, nullptr};
        for (int w = 0; w < 2; ++w) {
            const std::vector<float*>& B32 = w == 0 ? s->V32 : s->W32;
            if (hipMalloc((void**)&d[w], sizeof(float*) * (s->mcap + 1)) != hipSuccess ||
                hipMemcpy(d[w], B32.data(), sizeof(float*) * (s->mcap + 1), hipMemcpyHostToDevice) != hipSuccess) {
                if (d[w]) hipFree(d[w]);
                if (w == 1) hipFree(d[0]);
                set_error("hipMalloc failed for the FP32 basis");
                return 2;
            }
        }
        s->V32d = d[0];
        s->W32d = d[1];
    }
    if ((path & TSR_PATH_JVLIN) && !lin_mode_ok(s->P)) {
        set_error("linearised FD products need micro templates of at most 1024 nodes");
        return 1;
    }
    if ((path & TSR_PATH_JVLIN) && !s->linrec) {
        if (hipMalloc((void**)&s->linrec, kLinBytesPerNode * (size_t)s->N) != hipSuccess) {
            s->linrec = nullptr;
            set_error("hipMalloc failed for the linearisation record");
            return 2;
        }
    }
    s->path = path;
    return 0;
}

int tsr_solver_set_controller(tsr_solver* s, int mode)
{
    if (mode < 0 || mode > 2) return 1;
    s->ctrl_mode = mode;
    return 0;
}

void tsr_solver_set_seed(tsr_solver* s, unsigned long long seed)
{
    for (int q = 0; q < s->S; ++q) s->sc[q].rng = seed * 0x100000001B3ull + (unsigned long long)q;
}

int tsr_solver_set_mlp(tsr_solver* s, const float* params, const float* mean, const float* stdv)
{
    if (!s->mlp) s->mlp = mlp_create();
    if (!s->mlp) return 2;
    return mlp_set(s->mlp, params, mean, stdv);
}

double* tsr_solver_dev_u(tsr_solver* s) { return s->u; }
double* tsr_solver_dev_scratch(tsr_solver* s, int which) { return s->scratch[which & 1]; }
long long tsr_solver_N(const tsr_solver* s) { return s->N; }
int tsr_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 2; }
int tsr_h2d(double* dst, const double* src, long long n)
{
    return hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess ? 0 : 2;
}
int tsr_d2h(double* dst, const double* src, long long n)
{
    return hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- reference signatures
}  // extern "C"

// The reference's RHS names (include/main.h:21-22), C++ linkage (refabi.cpp): both are
// the device RHS, so exprem keeps the step on the device for them as for Gfunc_hip.
int Gfunc_serial(double* g, double* u, user_data_struct user_data);
int Gfunc_parallel(double* g, double* u, user_data_struct user_data);

static bool device_rhs(tsr_gfunc_t G) { return G == Gfunc_hip || G == Gfunc_parallel || G == Gfunc_serial; }

extern "C" {
// field by field: the struct has padding, whose bytes a caller's copy need not repeat
static bool same_options(const exprem_options_struct& a, const exprem_options_struct& b)
{
    return a.AbsTol == b.AbsTol && a.RelTol == b.RelTol && a.MinKrylov == b.MinKrylov && a.MaxKrylov == b.MaxKrylov &&
           a.MinStep == b.MinStep && a.MaxStep == b.MaxStep && a.MaxNumSteps == b.MaxNumSteps &&
           a.SafetyFac == b.SafetyFac && a.MinStepDecFac == b.MinStepDecFac && a.MaxStepIncFac == b.MaxStepIncFac &&
           a.InitStep == b.InitStep;
}

static tsr_solver* solver_for(const user_data_struct& ud, const exprem_options_struct* opt)
{
    std::string err;
    const tsr_batch_view* v = registry_find(ud, err);
    if (!v) { set_error(err); return nullptr; }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_solver_cache.find(v);
    if (it != g_solver_cache.end() && (!opt || same_options(it->second->opt, *opt))) return it->second;
    exprem_options_struct o;
    if (opt) o = *opt;
    else if (it != g_solver_cache.end()) o = it->second->opt;
    else { set_exprem_options(o); o.MaxKrylov = 30; }
    tsr_solver* s = create(v, &o);
    if (!s) return nullptr;
    if (it != g_solver_cache.end()) {
        // a caller changed options: keep the old solver alive (its device state may
        // still be referenced by that caller's loop) and cache the new one
        g_retired.push_back({v, it->second});
        g_solver_cache.erase(it);
    }
    g_solver_cache[v] = s;
    return s;
}

int Gfunc_hip(double* g, double* u, user_data_struct user_data)
{
    tsr_solver* s = solver_for(user_data, nullptr);
    if (!s) return 2;
    return tsr_solver_gfunc_host(s, g, u);
}

int exprem_c(tsr_gfunc_t Gfunc, int N, double* t, double tend, double* tspan, double* tspan_solns, double* u0,
             user_data_struct ud, exprem_options_struct options, exprem_stats_struct* stats)
{
    tsr_solver* s = solver_for(ud, &options);
    if (!s) return -1;
    if (s->N != N || s->S != 1) {
        set_error("exprem: N = " + std::to_string(N) + " does not match the problem of user_data (N = " +
                  std::to_string(s->N) + ", " + std::to_string(s->S) + " scenarios)");
        return -1;
    }
    if (*t == 0.0) exprem_stats_init(stats, &options, u0);
    ScState& c = s->sc[0];
    c.stats = *stats;
    c.proposal = stats->CurrentStep;
    c.t = *t;
    s->host_g = device_rhs(Gfunc) ? nullptr : Gfunc;
    if (s->host_g) {
        s->host_ud = ud;
        s->hx.resize(N); s->hg.resize(N); s->hg0.resize(N);
    }
    if (hipMemcpy(s->u, u0, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("exprem: copying u0 to the device failed");
        return -1;
    }
    // land exactly on the next output time
    double tstop = tend;
    int out_k = -1;
    const int nout = tspan ? (int)tspan[0] : 0;
    for (int k = 1; k <= nout; ++k)
        if (tspan[k] > *t * (1.0 + 1e-14) + 1e-300 && tspan[k] <= tstop) { tstop = tspan[k]; out_k = k; }
    const int rc = solver_step(s, tstop, nullptr);
    s->host_g = nullptr;
    if (rc) return -1;
    if (hipMemcpy(u0, s->u, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("exprem: copying the solution from the device failed");
        return -1;
    }
    *t = c.t;
    if (out_k > 0 && tspan_solns && c.t == tstop) std::memcpy(tspan_solns + (size_t)out_k * N, u0, sizeof(double) * N);
    *stats = c.stats;
    stats->CurrentSoln = u0;
    return 0;
}

double weighted_norm(double* e, double* u, double AbsTol, double RelTol, int N)
{
    double acc = 0.0;
    for (int i = 0; i < N; ++i) {
        const double r = e[i] / (AbsTol + RelTol * std::fabs(u[i]));
        acc += r * r;
    }
    return std::sqrt(acc / (double)N);
}

}  // extern "C"

int exprem(int (*Gfunc)(double* g, double* u, user_data_struct user_data), int N, double& t, double tend,
           double* tspan, double* tspan_solns, double* u0, user_data_struct user_data, exprem_options_struct options,
           exprem_stats_struct& stats)
{
    // the reference's driver stops on -1 without a message (main.cpp:319-322): say why
    const int rc = exprem_c(Gfunc, N, &t, tend, tspan, tspan_solns, u0, user_data, options, &stats);
    if (rc) std::fprintf(stderr, "exprem: step failed at t = %.6e: %s\n", t, tsr_last_error());
    return rc;
}
