// hvc_engine.cpp -- host side of the HVC engine: the C ABI of include/hvc_engine.h.
//
// A context owns one HIP stream and every device buffer of one network.  The
// trial clock (NGS:8906-8925), the synchronisation steps (NGS:9291, 9563), the
// training-kick steps (NGS:8933-8940) and the stage-time offsets of the
// steppers are evaluated here in IEEE binary64 exactly as the reference
// evaluates them; the kernels in hvc_kernels.hip consume them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hvc_internal.h"
#include "hvc_neuron.hip.h"

namespace hvc {
void launch_decay_table(hipStream_t st, const double* clock, const double* offsets, int num_steps, void* rows);
void launch_trace_ra(hipStream_t st, const NetDev& nd, int n, int n_steps, uint32_t trial, const double* EL,
                     const double* GCa, const double* kick, const double* exc_w, const double* inh_g, double* out,
                     uint8_t* fired);
void launch_trace_i(hipStream_t st, const NetDev& nd, int n, int n_steps, int first_trial, uint32_t trial,
                    const double* exc_g, const double* inh_g, double* out, uint8_t* fired);
void launch_reset(hipStream_t st, const NetDev& nd, int E);
void launch_integrate(hipStream_t st, const NetDev& nd, int t0, int t1, const Tuning& tu, const StdpArgs& sa,
                      int stdp);
void launch_trace_ra_f32(hipStream_t st, const NetDev& nd, int n, int n_steps, uint32_t trial, const double* EL,
                         const double* GCa, const double* kick, const double* exc_w, const double* inh_g, double* out,
                         uint8_t* fired);
void launch_trace_i_f32(hipStream_t st, const NetDev& nd, int n, int n_steps, int first_trial, uint32_t trial,
                        const double* exc_g, const double* inh_g, double* out, uint8_t* fired);
void launch_integrate_f32(hipStream_t st, const NetDev& nd, int t0, int t1, const Tuning& tu, const StdpArgs& sa,
                          int stdp);
void launch_deliver(hipStream_t st, const NetDev& nd, int t0, int t1, const Tuning& tu);
void launch_sync(hipStream_t st, const NetDev& nd, const Tuning& tu);
void launch_stdp(hipStream_t st, const NetDev& nd, const StdpArgs& sa);
void launch_noise_events(hipStream_t st, const NetDev& nd, int tb, int te, int gen_ra, int gen_i);
void launch_add_i32(hipStream_t st, int* a, long long n, int delta);
void launch_rr_pos(hipStream_t st, const NetDev& nd, RrPos* pos, int E);
void launch_decay(hipStream_t st, const NetDev& nd);
void launch_chain_edges(hipStream_t st, const NetDev& nd);
void launch_eval_math(hipStream_t st, int fn, int n, const double* a, const double* b, double* out);
hipError_t noalloc_setup(int device);
void* noalloc_scratch(int device);
void launch_clock_table(hipStream_t st, double* clock, int n_rows, double dt);
}  // namespace hvc

using namespace hvc;

namespace {

enum TimerClass { kTInt = 0, kTDeliver, kTSync, kTDecay, kNumTimers };
const char* kTimerNames[kNumTimers] = {"integrate", "deliver", "sync", "decay"};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// device buffers of one call, freed on every return path
struct CallBufs {
    std::vector<void*> p;
    template <class T>
    hipError_t alloc(T** out, size_t bytes)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes > 0 ? bytes : 1);
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
    ~CallBufs()
    {
        for (void* q : p) hipFree(q);
    }
};

// HIP events of one call, destroyed on every return path
struct CallEvents {
    std::vector<hipEvent_t> ev;
    hipError_t create(size_t n)
    {
        ev.reserve(n);
        for (size_t k = 0; k < n; ++k) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        return hipSuccess;
    }
    ~CallEvents()
    {
        for (auto e : ev) hipEventDestroy(e);
    }
};

}  // namespace

// one resident network of a serving pool, in its own ids
struct PoolNet {
    int handle = 0;
    int n_ra = 0, n_i = 0;
    int E = 0, A = 0, Bn = 0;   // candidates, RA->I and I->RA connections
    uint32_t key[2] = {0, 0};
    std::vector<int> train_ids;
    std::vector<double> spread;
};

// Asynchronous serving on a pool (hvc_submit / hvc_poll): one trial of one
// resident network per ticket, run by the context's serving thread in rounds.
struct ServeTicket {
    int handle = 0;
    hvc_trial trial{};
    long long round = -1;          // -1 while queued, else the round that runs it
    bool done = false;
    hvc_status status = HVC_OK;
    std::string err;
    int bad_values = 0, warnings[3] = {0, 0, 0};
    std::vector<int> counts[3];    // soma, dendrite, HVC-I: per neuron of the network
    std::vector<double> times[3];
};

struct Serving {
    std::mutex q_mu;               // guards everything below (never held while waiting on api_mu)
    std::condition_variable cv;    // the serving thread waits on it for work or stop
    std::thread worker;
    bool stop = false;
    uint64_t next_ticket = 1;
    long long rounds = 0;
    std::deque<uint64_t> queue;    // queued tickets in submission order
    std::unordered_map<uint64_t, ServeTicket> tickets;
};

struct hvc_ctx {
    hvc_params prm;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<DevBuf> bufs;

    int num_steps = 0;
    std::vector<double> clock;
    std::vector<std::pair<int, int>> windows;  // (t0, t1) inclusive, t1 = sync step
    double offsets[kNumOffsets];
    double* d_clock = nullptr;
    void* d_decay = nullptr;
    double* d_offsets = nullptr;

    bool uploaded = false;
    NetDev nd{};
    int n_ra = 0, n_i = 0, E = 0;
    std::vector<int> rr_ptr, rr_tgt;
    std::vector<int> train_ids, train_net;
    std::vector<double> spread;
    int* d_kick = nullptr;
    std::vector<int> h_kick;

    // networks resident in the context (one for hvc_upload_network)
    int n_nets = 0;
    std::vector<int> ra_off, i_off;
    std::vector<uint32_t> keys;
    uint32_t* d_net_trial = nullptr;
    int* d_net_first = nullptr;
    int* d_net_growth = nullptr;
    int* d_ra_net = nullptr;
    int* d_i_net = nullptr;
    int* d_ra_off = nullptr;
    int* d_i_off = nullptr;
    uint32_t* d_key = nullptr;
    double* d_EL = nullptr;
    double* d_GCa = nullptr;
    int* d_rr_ptr = nullptr;
    int* d_rr_tgt = nullptr;
    double* d_rr_d = nullptr;
    int* d_perm = nullptr;
    RrPos* d_rr_pos = nullptr;     // rebuilt from d_perm before the next trial when rr_pos_stale
    bool rr_pos_stale = true;
    int *d_a_ptr = nullptr, *d_a_tgt = nullptr, *d_b_ptr = nullptr, *d_b_tgt = nullptr;
    double *d_a_w = nullptr, *d_a_d = nullptr, *d_b_w = nullptr, *d_b_d = nullptr;
    Tuning tuning;

    // serving pool
    bool pool = false;
    std::mutex api_mu;               // held by the pool calls and by the serving thread through a round
    std::atomic<int> api_waiters{0}; // pool calls waiting for api_mu (the serving thread lets them in first)
    std::unique_ptr<Serving> serve;  // hvc_submit / hvc_poll
    int* d_net_live = nullptr;       // per network: 1 = runs this round (serving rounds only)
    std::vector<int> h_live;
    std::vector<PoolNet> pnets;
    int next_handle = 1;
    long long a_tot = 0, b_tot = 0;
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    // capacities of the device arrays
    int caps_ra = 0, caps_i = 0, caps_nets = 0, caps_indeg = 0;
    long long caps_rr = 0, caps_a2i = 0, caps_i2ra = 0;
    std::vector<uint32_t> h_net_trial;
    std::vector<int> h_net_first, h_net_growth;

    int* d_hits = nullptr;         // branch counters (hvc_set_branch_counting)

    bool trial_done = false;
    int log_counts[3] = {0, 0, 0};
    std::vector<int2> h_log[3];
    std::vector<int2> h_log_raw;     // a spike log as copied from the device
    std::vector<int> h_log_start;    // counting-sort offsets per neuron
    bool h_log_valid[3] = {false, false, false};

    // timing: HIP events recorded on the engine stream around sampled launches,
    // read after the trial's final synchronisation
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_pending;  // (class, index of start event)
    double t_sum[kNumTimers] = {};
    int t_cnt[kNumTimers] = {};
};

// api_mu for a pool call: announced first, so that the serving thread lets the
// call in at the next round boundary instead of starting another round
struct ApiLock {
    hvc_ctx* c;
    explicit ApiLock(hvc_ctx* ctx) : c(ctx)
    {
        c->api_waiters.fetch_add(1);
        c->api_mu.lock();
        c->api_waiters.fetch_sub(1);
    }
    ~ApiLock() { c->api_mu.unlock(); }
    ApiLock(const ApiLock&) = delete;
    ApiLock& operator=(const ApiLock&) = delete;
};

#define HVC_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
            return HVC_E_HIP;                                                                  \
        }                                                                                      \
    } while (0)

static hvc_status fail(hvc_ctx* ctx, hvc_status s, const std::string& msg)
{
    ctx->err = msg;
    return s;
}

template <class T>
static T* dalloc(hvc_ctx* ctx, size_t count, hvc_status* st)
{
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        *st = HVC_E_OOM;
        ctx->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed";
        return nullptr;
    }
    hipMemsetAsync(p, 0, bytes, ctx->stream);
    ctx->bufs.push_back(DevBuf{p, bytes});
    return (T*)p;
}

// free the device buffers allocated since ctx->bufs held `first` of them
static void free_bufs(hvc_ctx* ctx, size_t first = 0)
{
    for (size_t k = first; k < ctx->bufs.size(); ++k) hipFree(ctx->bufs[k].p);
    ctx->bufs.resize(std::min(first, ctx->bufs.size()));
}

extern "C" const char* hvc_last_error(const hvc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" hvc_status hvc_create(const hvc_params* prm, int device, hvc_ctx** out)
{
    if (!prm || !out) return HVC_E_ARG;
    *out = nullptr;
    hvc_ctx* ctx = new hvc_ctx();
    ctx->prm = *prm;
    ctx->device = device;
    auto bad = [&](const char* m) {
        ctx->err = m;
        return HVC_E_ARG;
    };
    if (!(prm->dt > 0) || !(prm->trial_ms > 0) || !(prm->sync_ms > 0)) { hvc_status s = bad("dt, trial_ms, sync_ms must be > 0"); *out = ctx; return s; }
    if (prm->ra_model < 0 || prm->ra_model > 1 || prm->i_model < 0 || prm->i_model > 1) { hvc_status s = bad("model must be 0 or 1"); *out = ctx; return s; }
    if (prm->precision != 0 && prm->precision != 1) { hvc_status s = bad("precision must be 0 (FP64) or 1 (FP32)"); *out = ctx; return s; }
    if (prm->ra_poisson && prm->ra_model != 1) { hvc_status s = bad("RA Poisson noise is part of the RK4 (HH2_buffer) form only"); *out = ctx; return s; }
    if (!(prm->syn.ACTIVATION_THRESHOLD > 0) || !(prm->syn.SUPERSYNAPSE_THRESHOLD > 0) || prm->syn.Nss < 0) {
        hvc_status s = bad("synaptic thresholds must be > 0 and Nss >= 0");
        *out = ctx;
        return s;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        ctx->err = "no HIP device";
        *out = ctx;
        return HVC_E_HIP;
    }
    if (device < 0 || device >= ndev) { hvc_status s = bad("device index out of range"); *out = ctx; return s; }
    *out = ctx;
    HVC_TRY(hipSetDevice(device));
    HVC_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // this device's fixed buffer of hvc_trace_ra_noalloc (looked up once per device)
    HVC_TRY(noalloc_setup(device));
    // default launch shape for one network: eight lanes per HVC-I neuron (one
    // gate rate and one stage combination per lane and DP8 stage: the shortest
    // dependent chain, so the fastest single-network layout)
    ctx->tuning.i_lanes = 3;
    ctx->tuning.ra_lanes = 0;

    // trial clock and synchronisation points (NGS:8906-8925, 9291, 9563)
    const double dt = prm->dt;
    ctx->num_steps = (int)(prm->trial_ms / dt);
    if (ctx->num_steps < 2) return fail(ctx, HVC_E_ARG, "trial shorter than two steps");
    ctx->clock.assign(ctx->num_steps, 0.0);
    double internal_time = 0, network_time = prm->sync_ms;
    int t0 = 1;
    for (int t = 1; t < ctx->num_steps; ++t) {
        internal_time += dt;
        ctx->clock[t] = internal_time;
        if ((internal_time > network_time) || (t == ctx->num_steps - 1)) {
            if (t - t0 + 1 > kMaxWindowSteps)
                return fail(ctx, HVC_E_ARG, "more than 64 steps between synchronisation points");
            ctx->windows.emplace_back(t0, t);
            t0 = t + 1;
            network_time += prm->sync_ms;
        }
    }
    // stage-time offsets of the steppers, written as the reference writes them
    double* o = ctx->offsets;
    o[kOffDt] = dt;
    o[kOffHalf] = 0.5 * dt;
    o[kOffThird] = dt / 3;
    o[kOffTwoThird] = 2 * dt / 3;
    o[kOffDp2] = dt / 18;
    o[kOffDp3] = dt / 12;
    o[kOffDp4] = dt / 8;
    o[kOffDp5] = 5 * dt / 16;
    o[kOffDp6] = 3 * dt / 8;
    o[kOffDp7] = 59 * dt / 400;
    o[kOffDp8] = 93 * dt / 200;
    o[kOffDp9] = 5490023248 * dt / 9719169821;
    o[kOffDp10] = 13 * dt / 20;
    o[kOffDp11] = 1201146811 * dt / 12990119798;

    hvc_status st = HVC_OK;
    ctx->d_clock = dalloc<double>(ctx, ctx->num_steps, &st);
    ctx->d_decay = dalloc<DecayRow>(ctx, ctx->num_steps, &st);
    ctx->d_offsets = dalloc<double>(ctx, kNumOffsets, &st);
    if (st != HVC_OK) return st;
    HVC_TRY(hipMemcpyAsync(ctx->d_clock, ctx->clock.data(), sizeof(double) * ctx->num_steps, hipMemcpyHostToDevice,
                           ctx->stream));
    HVC_TRY(hipMemcpyAsync(ctx->d_offsets, ctx->offsets, sizeof(ctx->offsets), hipMemcpyHostToDevice, ctx->stream));
    launch_decay_table(ctx->stream, ctx->d_clock, ctx->d_offsets, ctx->num_steps, ctx->d_decay);
    HVC_TRY(hipGetLastError());
    HVC_TRY(hipStreamSynchronize(ctx->stream));

    NetDev& nd = ctx->nd;
    nd.ra_model = prm->ra_model;
    nd.i_model = prm->i_model;
    nd.dt = dt;
    nd.num_steps = ctx->num_steps;
    nd.clock = ctx->d_clock;
    nd.decay = ctx->d_decay;
    nd.sq_sig_s = std::sqrt(dt) * prm->sigma_s;
    nd.sq_sig_d = std::sqrt(dt) * prm->sigma_d;
    nd.mu_term_s = 100000 * (0.0 + prm->mu_s) / 5000.0;
    nd.mu_term_d = 100000 * (0.0 + prm->mu_d) / 10000.0;
    nd.ra_poisson = prm->ra_poisson;
    nd.i_poisson = prm->i_poisson;
    nd.seed_lo = prm->seed_lo;
    nd.seed_hi = prm->seed_hi;
    nd.syn = prm->syn;
    return HVC_OK;
}

extern "C" void hvc_destroy(hvc_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->serve) {
        {
            std::lock_guard<std::mutex> lq(ctx->serve->q_mu);
            ctx->serve->stop = true;
        }
        ctx->serve->cv.notify_all();
        if (ctx->serve->worker.joinable()) ctx->serve->worker.join();
    }
    if (ctx->stream) {
        hipStreamSynchronize(ctx->stream);
        free_bufs(ctx);
        for (auto e : ctx->ev_pool) hipEventDestroy(e);
        if (ctx->d_stage) hipFree(ctx->d_stage);
        hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

// ---------------------------------------------------------------------------
// network upload
// ---------------------------------------------------------------------------
static bool check_csr(const int* ptr, const int* tgt, const double* d, int n_pre, int n_post, bool ascending,
                      std::string& why)
{
    if (!ptr || ptr[0] != 0) { why = "CSR ptr[0] must be 0"; return false; }
    for (int i = 0; i < n_pre; ++i) {
        if (ptr[i + 1] < ptr[i]) { why = "CSR ptr not monotone"; return false; }
        std::vector<int> seen;
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
            if (tgt[e] < 0 || tgt[e] >= n_post) { why = "CSR target out of range"; return false; }
            if (!(d[e] >= 0) || !std::isfinite(d[e])) { why = "axonal delays must be finite and >= 0"; return false; }
            if (ascending && e > ptr[i] && tgt[e] <= tgt[e - 1]) { why = "RA->RA targets must be strictly ascending"; return false; }
            seen.push_back(tgt[e]);
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) { why = "duplicate connection"; return false; }
    }
    return true;
}

// edges of each row sorted by delay (stable), for the delay-queue walk
static void sort_fixed(const int* ptr, const int* tgt, const double* w, const double* d, int n, std::vector<int>& otgt,
                       std::vector<double>& ow, std::vector<double>& od)
{
    int E = ptr[n];
    otgt.resize(E); ow.resize(E); od.resize(E);
    std::vector<int> idx;
    for (int i = 0; i < n; ++i) {
        idx.resize(ptr[i + 1] - ptr[i]);
        for (size_t k = 0; k < idx.size(); ++k) idx[k] = ptr[i] + (int)k;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
        for (size_t k = 0; k < idx.size(); ++k) {
            otgt[ptr[i] + k] = tgt[idx[k]];
            ow[ptr[i] + k] = w[idx[k]];
            od[ptr[i] + k] = d[idx[k]];
        }
    }
}

// synapse state of the rows [r0, r1) (edges [E0, E1)) from lists of local ids
// (network-local ids are offset by r0)
static hvc_status set_synapse_state(hvc_ctx* ctx, int r0, int r1, const double* w, const int* act_cnt,
                                    const int* act_ids, const int* sup_cnt, const int* sup_ids, const int* remodeled)
{
    const int n = r1 - r0, Nss = ctx->prm.syn.Nss, NssA = std::max(Nss, 1);
    const int E0 = ctx->rr_ptr[r0], E1 = ctx->rr_ptr[r1], En = E1 - E0;
    std::vector<uint8_t> act(En + 1, 0), sup(En + 1, 0);
    std::vector<int> alist(En + 1, 0), slist((size_t)n * NssA + 1, 0);
    std::vector<int> acnt(n + 1), scnt(n + 1), rem(n + 1);
    auto edge_of = [&](int i, int j) {
        auto b = ctx->rr_tgt.begin() + ctx->rr_ptr[i], e = ctx->rr_tgt.begin() + ctx->rr_ptr[i + 1];
        auto it = std::lower_bound(b, e, j);
        return (it != e && *it == j) ? (int)(it - ctx->rr_tgt.begin()) : -1;
    };
    int oa = 0, os = 0;
    for (int r = 0; r < n; ++r) {
        const int i = r0 + r;
        int len = ctx->rr_ptr[i + 1] - ctx->rr_ptr[i];
        int a = act_cnt ? act_cnt[r] : 0, s = sup_cnt ? sup_cnt[r] : 0;
        if (a < 0 || a > len || s < 0 || s > len || s > Nss) return fail(ctx, HVC_E_ARG, "active/super list size out of range");
        acnt[r] = a; scnt[r] = s;
        for (int k = 0; k < a; ++k) {
            int e = edge_of(i, act_ids[oa++] + r0);
            if (e < 0 || act[e - E0]) return fail(ctx, HVC_E_ARG, "active synapse is not a candidate or repeated");
            act[e - E0] = 1;
            alist[ctx->rr_ptr[i] - E0 + k] = ctx->rr_tgt[e];
        }
        for (int k = 0; k < s; ++k) {
            int e = edge_of(i, sup_ids[os++] + r0);
            if (e < 0 || sup[e - E0]) return fail(ctx, HVC_E_ARG, "supersynapse is not a candidate or repeated");
            sup[e - E0] = 1;
            slist[(size_t)r * NssA + k] = ctx->rr_tgt[e];
        }
        rem[r] = remodeled ? remodeled[r] : 0;
    }
    NetDev& nd = ctx->nd;
    hipStream_t st = ctx->stream;
    if (En > 0) {
        if (w) {
            HVC_TRY(hipMemcpyAsync(nd.w + E0, w, sizeof(double) * En, hipMemcpyHostToDevice, st));
        } else {
            std::vector<double> z(En, 0.0);
            HVC_TRY(hipMemcpy(nd.w + E0, z.data(), sizeof(double) * En, hipMemcpyHostToDevice));
        }
        HVC_TRY(hipMemcpyAsync(nd.act + E0, act.data(), En, hipMemcpyHostToDevice, st));
        HVC_TRY(hipMemcpyAsync(nd.sup + E0, sup.data(), En, hipMemcpyHostToDevice, st));
        HVC_TRY(hipMemcpyAsync(nd.act_list + E0, alist.data(), sizeof(int) * En, hipMemcpyHostToDevice, st));
    }
    HVC_TRY(hipMemcpyAsync(nd.sup_list + (size_t)r0 * NssA, slist.data(), sizeof(int) * (size_t)n * NssA,
                           hipMemcpyHostToDevice, st));
    HVC_TRY(hipMemcpyAsync(nd.act_cnt + r0, acnt.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HVC_TRY(hipMemcpyAsync(nd.sup_cnt + r0, scnt.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HVC_TRY(hipMemcpyAsync(nd.remodeled + r0, rem.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HVC_TRY(hipStreamSynchronize(st));
    return HVC_OK;
}

// Networks laid side by side: network b's HVC-RA ids are offset by ra_off[b],
// its HVC-I ids by i_off[b], its candidate edges by e_off[b].  The merged
// network is block-diagonal, so one launch per kernel advances all of them.
struct Merged {
    std::vector<int> ra2i_ptr, ra2i_tgt, i2ra_ptr, i2ra_tgt, rr_ptr, rr_tgt, act_cnt, act_ids, sup_cnt, sup_ids;
    std::vector<int> remodeled, train_ids, train_net;
    std::vector<double> ra2i_w, ra2i_d, i2ra_w, i2ra_d, rr_d, rr_w, erest, gca, spread;
    std::vector<int> ra_off, i_off;
    hvc_network view{};
};

static void merge_networks(int n, const hvc_network* nets, Merged& m)
{
    m.ra_off.assign(1, 0);
    m.i_off.assign(1, 0);
    m.ra2i_ptr.assign(1, 0);
    m.i2ra_ptr.assign(1, 0);
    m.rr_ptr.assign(1, 0);
    for (int b = 0; b < n; ++b) {
        const hvc_network& x = nets[b];
        const int ro = m.ra_off.back(), io = m.i_off.back();
        for (int i = 0; i < x.n_ra; ++i) {
            for (int e = x.ra2i_ptr[i]; e < x.ra2i_ptr[i + 1]; ++e) {
                m.ra2i_tgt.push_back(x.ra2i_tgt[e] + io);
                m.ra2i_w.push_back(x.ra2i_w[e]);
                m.ra2i_d.push_back(x.ra2i_d[e]);
            }
            m.ra2i_ptr.push_back((int)m.ra2i_tgt.size());
            for (int e = x.rr_ptr[i]; e < x.rr_ptr[i + 1]; ++e) {
                m.rr_tgt.push_back(x.rr_tgt[e] + ro);
                m.rr_d.push_back(x.rr_d[e]);
                m.rr_w.push_back(x.rr_w ? x.rr_w[e] : 0.0);
            }
            m.rr_ptr.push_back((int)m.rr_tgt.size());
            m.act_cnt.push_back(x.act_cnt ? x.act_cnt[i] : 0);
            m.sup_cnt.push_back(x.sup_cnt ? x.sup_cnt[i] : 0);
            m.remodeled.push_back(x.remodeled ? x.remodeled[i] : 0);
            m.erest.push_back(x.erest[i]);
            m.gca.push_back(x.gca[i]);
        }
        int na = 0, ns = 0;
        for (int i = 0; i < x.n_ra; ++i) {
            na += x.act_cnt ? x.act_cnt[i] : 0;
            ns += x.sup_cnt ? x.sup_cnt[i] : 0;
        }
        for (int k = 0; k < na; ++k) m.act_ids.push_back(x.act_ids[k] + ro);
        for (int k = 0; k < ns; ++k) m.sup_ids.push_back(x.sup_ids[k] + ro);
        for (int i = 0; i < x.n_i; ++i) {
            for (int e = x.i2ra_ptr[i]; e < x.i2ra_ptr[i + 1]; ++e) {
                m.i2ra_tgt.push_back(x.i2ra_tgt[e] + ro);
                m.i2ra_w.push_back(x.i2ra_w[e]);
                m.i2ra_d.push_back(x.i2ra_d[e]);
            }
            m.i2ra_ptr.push_back((int)m.i2ra_tgt.size());
        }
        for (int k = 0; k < x.n_train; ++k) {
            m.train_ids.push_back(x.train_ids[k] + ro);
            m.train_net.push_back(b);
            m.spread.push_back(x.spread[k]);
        }
        m.ra_off.push_back(ro + x.n_ra);
        m.i_off.push_back(io + x.n_i);
    }
    hvc_network& v = m.view;
    v.n_ra = m.ra_off.back();
    v.n_i = m.i_off.back();
    auto ip = [](std::vector<int>& a) { return a.empty() ? nullptr : a.data(); };
    auto dp = [](std::vector<double>& a) { return a.empty() ? nullptr : a.data(); };
    v.ra2i_ptr = m.ra2i_ptr.data(); v.ra2i_tgt = ip(m.ra2i_tgt); v.ra2i_w = dp(m.ra2i_w); v.ra2i_d = dp(m.ra2i_d);
    v.i2ra_ptr = m.i2ra_ptr.data(); v.i2ra_tgt = ip(m.i2ra_tgt); v.i2ra_w = dp(m.i2ra_w); v.i2ra_d = dp(m.i2ra_d);
    v.rr_ptr = m.rr_ptr.data(); v.rr_tgt = ip(m.rr_tgt); v.rr_d = dp(m.rr_d); v.rr_w = dp(m.rr_w);
    v.act_cnt = ip(m.act_cnt); v.act_ids = ip(m.act_ids); v.sup_cnt = ip(m.sup_cnt); v.sup_ids = ip(m.sup_ids);
    v.remodeled = ip(m.remodeled); v.erest = dp(m.erest); v.gca = dp(m.gca);
    v.n_train = (int)m.train_ids.size(); v.train_ids = ip(m.train_ids); v.spread = dp(m.spread);
}

static hvc_status check_network(hvc_ctx* ctx, const hvc_network* net, const char* what)
{
    const int nra = net->n_ra, ni = net->n_i;
    if (nra <= 0 || ni < 0) return fail(ctx, HVC_E_ARG, std::string(what) + ": n_ra must be > 0 and n_i >= 0");
    std::string why;
    if (!check_csr(net->ra2i_ptr, net->ra2i_tgt, net->ra2i_d, nra, ni, false, why)) return fail(ctx, HVC_E_ARG, std::string(what) + " RA->I: " + why);
    if (!check_csr(net->i2ra_ptr, net->i2ra_tgt, net->i2ra_d, ni, nra, false, why)) return fail(ctx, HVC_E_ARG, std::string(what) + " I->RA: " + why);
    if (!check_csr(net->rr_ptr, net->rr_tgt, net->rr_d, nra, nra, true, why)) return fail(ctx, HVC_E_ARG, std::string(what) + " RA->RA: " + why);
    if (!net->erest || !net->gca) return fail(ctx, HVC_E_ARG, std::string(what) + ": erest and gca are required");
    for (int k = 0; k < net->n_train; ++k)
        if (net->train_ids[k] < 0 || net->train_ids[k] >= nra) return fail(ctx, HVC_E_ARG, std::string(what) + ": training id out of range");
    return HVC_OK;
}

// per row: candidates sorted by (delay, target), as edge indices + e_base.  Two
// different delays whose sums with a spike time could round to the same
// delivery time would make that order differ from the reference's (time,
// target) order; reject them (|d1 - d0| >= 4 ulp of the largest clock value
// keeps the sums distinct).
static hvc_status rr_order(hvc_ctx* ctx, const int* ptr, const int* tgt, const double* d, int n, int e_base,
                           std::vector<int>& perm)
{
    perm.resize(ptr[n]);
    for (int i = 0; i < n; ++i) {
        int b = ptr[i], e = ptr[i + 1];
        for (int k = b; k < e; ++k) perm[k] = k;
        std::sort(perm.begin() + b, perm.begin() + e, [&](int x, int y) {
            if (d[x] != d[y]) return d[x] < d[y];
            return tgt[x] < tgt[y];
        });
        for (int k = b + 1; k < e; ++k) {
            double d0 = d[perm[k - 1]], d1 = d[perm[k]];
            double tmax = ctx->prm.trial_ms + d1;
            if (d0 != d1 && d1 - d0 < 4 * DBL_EPSILON * tmax)
                return fail(ctx, HVC_E_ARG, "two RA->RA delays of one row differ by less than 4 ulp of the trial clock");
        }
    }
    for (auto& x : perm) x += e_base;
    return HVC_OK;
}

// largest in-degree of a CSR connection set (self-connections skipped if asked)
static int max_indeg(const int* ptr, const int* tgt, int n_pre, int n_post, int skip_self)
{
    std::vector<int> deg(std::max(n_post, 1), 0);
    for (int i = 0; i < n_pre; ++i)
        for (int e = ptr[i]; e < ptr[i + 1]; ++e)
            if (!(skip_self && tgt[e] == i)) deg[tgt[e]]++;
    return *std::max_element(deg.begin(), deg.end());
}

// device capacity of a context: neurons, connections, networks, inbox slots
struct Caps {
    int nets = 1, ra = 0, i = 0;
    long long rr = 0, a2i = 0, i2ra = 0;
    int cap_gi_ra = 8, cap_ge_ra = 8, cap_ge_i = 8;
    long long log_cap = 0;
};

// entries per neuron of a chunk's noise-event list: the chunk spans at most
// kNoiseChunkWindows windows of (sync_ms / dt) + 2 steps; HVC-RA checks fire at
// 4 x 100 Hz and HVC-I checks at 2 x 250 Hz, so a list holds at most
// 0.5 events per ms on average -- sized for mean + 12 standard deviations plus
// slack and the two end marks (an overflow is reported as a capacity error)
static int noise_cap(const hvc_params* prm)
{
    const int steps = std::min(64, (int)(prm->sync_ms / prm->dt) + 2);
    const double mean = 0.5 * kNoiseChunkWindows * steps * prm->dt;
    return (int)std::ceil(mean + 12.0 * std::sqrt(mean)) + 18;
}

// allocates every array of ctx->nd at capacity c (contents undefined)
static hvc_status alloc_device(hvc_ctx* ctx, const Caps& c)
{
    hvc_status st = HVC_OK;
    NetDev& nd = ctx->nd;
    const size_t first_buf = ctx->bufs.size();
    const int nra = std::max(c.ra, 1), ni = std::max(c.i, 1), n_nets = std::max(c.nets, 1);
    const size_t E = (size_t)std::max<long long>(c.rr, 1);
    const size_t A = (size_t)std::max<long long>(c.a2i, 1), Bn = (size_t)std::max<long long>(c.i2ra, 1);
    const int NssA = std::max(ctx->prm.syn.Nss, 1);
    ctx->caps_ra = nra; ctx->caps_i = ni; ctx->caps_nets = n_nets;
    ctx->caps_rr = (long long)E; ctx->caps_a2i = (long long)A; ctx->caps_i2ra = (long long)Bn;
    ctx->caps_indeg = std::min(c.cap_gi_ra, std::min(c.cap_ge_ra, c.cap_ge_i)) - 8;
    ctx->d_ra_net = dalloc<int>(ctx, nra, &st);
    ctx->d_i_net = dalloc<int>(ctx, ni, &st);
    ctx->d_ra_off = dalloc<int>(ctx, n_nets + 1, &st);
    ctx->d_i_off = dalloc<int>(ctx, n_nets + 1, &st);
    ctx->d_key = dalloc<uint32_t>(ctx, 2 * (size_t)n_nets, &st);
    ctx->d_net_trial = dalloc<uint32_t>(ctx, n_nets, &st);
    ctx->d_net_first = dalloc<int>(ctx, n_nets, &st);
    ctx->d_net_growth = dalloc<int>(ctx, n_nets, &st);
    nd.ra_net = ctx->d_ra_net; nd.i_net = ctx->d_i_net; nd.net_ra_off = ctx->d_ra_off; nd.net_i_off = ctx->d_i_off;
    nd.net_key = ctx->d_key; nd.net_trial = ctx->d_net_trial; nd.net_first = ctx->d_net_first;
    nd.net_growth = ctx->d_net_growth;
    for (int k = 0; k < 12; ++k) nd.ra[k] = dalloc<double>(ctx, nra, &st);
    nd.ra_flags = dalloc<int>(ctx, nra, &st);
    nd.ra_noise = dalloc<double>(ctx, 4 * (size_t)nra, &st);
    nd.ra_draw = dalloc<uint32_t>(ctx, nra, &st);
    ctx->d_EL = dalloc<double>(ctx, nra, &st);
    ctx->d_GCa = dalloc<double>(ctx, nra, &st);
    nd.EL = ctx->d_EL;
    nd.GCa = ctx->d_GCa;
    ctx->d_kick = dalloc<int>(ctx, nra, &st);
    nd.kick_step = ctx->d_kick;
    for (int k = 0; k < 7; ++k) nd.in[k] = dalloc<double>(ctx, ni, &st);
    nd.i_flag = dalloc<int>(ctx, ni, &st);
    nd.i_noise = dalloc<double>(ctx, 2 * (size_t)ni, &st);
    nd.i_draw = dalloc<uint32_t>(ctx, ni, &st);
    nd.ev_cap = noise_cap(&ctx->prm);
    nd.ra_ev = dalloc<NoiseEvent>(ctx, (size_t)nra * nd.ev_cap, &st);
    nd.i_ev = dalloc<NoiseEvent>(ctx, (size_t)ni * nd.ev_cap, &st);
    nd.ra_ev_pos = dalloc<int>(ctx, nra, &st);
    nd.i_ev_pos = dalloc<int>(ctx, ni, &st);
    nd.soma_mask = dalloc<uint64_t>(ctx, nra, &st);
    nd.dend_mask = dalloc<uint64_t>(ctx, nra, &st);
    nd.i_mask = dalloc<uint64_t>(ctx, ni, &st);
    long long cap = ctx->prm.spike_capacity > 0 ? ctx->prm.spike_capacity
                                                : std::max<long long>(1 << 16, 256 * ((long long)nra + ni));
    if (c.log_cap > 0) cap = c.log_cap;
    if (cap > (1ll << 30)) cap = 1ll << 30;
    nd.log_cap = (int)cap;
    nd.soma_log = dalloc<int2>(ctx, cap, &st);
    nd.dend_log = dalloc<int2>(ctx, cap, &st);
    nd.i_log = dalloc<int2>(ctx, cap, &st);
    nd.log_cnt = dalloc<int>(ctx, 4, &st);
    nd.post_ring = dalloc<double>(ctx, (size_t)nra * kPostRing, &st);
    nd.post_cnt = dalloc<int>(ctx, nra, &st);
    nd.post_pend = dalloc<int>(ctx, nra, &st);
    // two buffers, alternating by window (window w fills and reads buffer w % 2)
    nd.post_list = dalloc<uint64_t>(ctx, 2 * (size_t)kPostListCap, &st);
    nd.post_list_cnt = dalloc<int>(ctx, 2, &st);
    nd.post_list_next_cnt = nd.post_list_cnt ? nd.post_list_cnt + 1 : nullptr;
    nd.last_pre = dalloc<double>(ctx, nra, &st);
    nd.st_next = dalloc<double>(ctx, nra, &st);
    ctx->d_rr_ptr = dalloc<int>(ctx, nra + 1, &st);
    ctx->d_rr_tgt = dalloc<int>(ctx, E, &st);
    ctx->d_rr_d = dalloc<double>(ctx, E, &st);
    ctx->d_perm = dalloc<int>(ctx, E, &st);
    ctx->d_rr_pos = dalloc<RrPos>(ctx, E, &st);
    nd.rr_ptr = ctx->d_rr_ptr; nd.rr_tgt = ctx->d_rr_tgt; nd.rr_d = ctx->d_rr_d; nd.rr_perm = ctx->d_perm;
    nd.rr_pos = ctx->d_rr_pos;
    nd.w = dalloc<double>(ctx, E, &st);
    nd.act = dalloc<uint8_t>(ctx, E, &st);
    nd.sup = dalloc<uint8_t>(ctx, E, &st);
    nd.act_list = dalloc<int>(ctx, E, &st);
    nd.act_cnt = dalloc<int>(ctx, nra, &st);
    nd.sup_list = dalloc<int>(ctx, (size_t)nra * NssA, &st);
    nd.sup_cnt = dalloc<int>(ctx, nra, &st);
    nd.remodeled = dalloc<int>(ctx, nra, &st);
    nd.pre_ring = dalloc<double>(ctx, E * kPreRing, &st);
    nd.pre_cnt = dalloc<int>(ctx, E, &st);
    nd.st_step = dalloc<int>(ctx, (size_t)nra * kMaxStreams, &st);
    nd.st_mode = dalloc<int>(ctx, (size_t)nra * kMaxStreams, &st);
    nd.st_cursor = dalloc<int>(ctx, (size_t)nra * kMaxStreams, &st);
    nd.st_n = dalloc<int>(ctx, (size_t)nra * kMaxStreams, &st);
    nd.st_count = dalloc<int>(ctx, nra, &st);
    nd.st_items = dalloc<int>(ctx, (size_t)nra * kMaxStreams * NssA, &st);
    ctx->d_a_ptr = dalloc<int>(ctx, nra + 1, &st);
    ctx->d_a_tgt = dalloc<int>(ctx, A, &st);
    ctx->d_a_w = dalloc<double>(ctx, A, &st);
    ctx->d_a_d = dalloc<double>(ctx, A, &st);
    ctx->d_b_ptr = dalloc<int>(ctx, ni + 1, &st);
    ctx->d_b_tgt = dalloc<int>(ctx, Bn, &st);
    ctx->d_b_w = dalloc<double>(ctx, Bn, &st);
    ctx->d_b_d = dalloc<double>(ctx, Bn, &st);
    nd.ra2i = FixedEdges{ctx->d_a_ptr, ctx->d_a_tgt, ctx->d_a_w, ctx->d_a_d};
    nd.i2ra = FixedEdges{ctx->d_b_ptr, ctx->d_b_tgt, ctx->d_b_w, ctx->d_b_d};
    auto mk_inflight = [&](int n) {
        InFlight f;
        f.head = dalloc<FlHead>(ctx, (size_t)n, &st);
        f.more = dalloc<FlSlot>(ctx, (size_t)n * (kMaxInflight - kHeadSlots), &st);
        f.held = dalloc<int>(ctx, (size_t)n, &st);
        return f;
    };
    nd.fl_ra2i = mk_inflight(nra);
    nd.fl_i2ra = mk_inflight(ni);
    nd.fl_rr = mk_inflight(nra);
    nd.chain_e = dalloc<int>(ctx, E, &st);
    auto mk_inbox = [&](int n, int cp) {
        Inbox b;
        b.cap = cp;
        b.rec = dalloc<Record>(ctx, (size_t)n * cp, &st);
        b.cnt = dalloc<int>(ctx, n, &st);
        return b;
    };
    nd.in_gi_ra = mk_inbox(nra, c.cap_gi_ra);
    nd.in_ge_ra = mk_inbox(nra, c.cap_ge_ra);
    nd.in_ge_i = mk_inbox(ni, c.cap_ge_i);
    const int wl_n[3] = {nra, nra, ni};
    for (int q = 0; q < 3; ++q) {
        nd.wl_cap[q] = wl_n[q];
        nd.wl_list[q] = dalloc<int>(ctx, 2 * (size_t)wl_n[q], &st);
        nd.wl_mark[q] = dalloc<int>(ctx, wl_n[q], &st);
    }
    nd.wl_cnt = dalloc<int>(ctx, kWlCounters, &st);
    nd.err = dalloc<unsigned int>(ctx, 1, &st);
    nd.warn = dalloc<int>(ctx, 3 * (size_t)n_nets + 1, &st);
    if (st != HVC_OK) free_bufs(ctx, first_buf);
    return st;
}

// the per-network tables and the network of every neuron, from ctx->ra_off / i_off / keys
static hvc_status upload_tables(hvc_ctx* ctx)
{
    const int B = ctx->n_nets;
    std::vector<int> ra_net(std::max(ctx->n_ra, 1), 0), i_net(std::max(ctx->n_i, 1), 0);
    for (int b = 0; b < B; ++b) {
        for (int i = ctx->ra_off[b]; i < ctx->ra_off[b + 1]; ++i) ra_net[i] = b;
        for (int i = ctx->i_off[b]; i < ctx->i_off[b + 1]; ++i) i_net[i] = b;
    }
    hipStream_t s = ctx->stream;
    if (ctx->n_ra > 0) HVC_TRY(hipMemcpyAsync(ctx->d_ra_net, ra_net.data(), sizeof(int) * ctx->n_ra, hipMemcpyHostToDevice, s));
    if (ctx->n_i > 0) HVC_TRY(hipMemcpyAsync(ctx->d_i_net, i_net.data(), sizeof(int) * ctx->n_i, hipMemcpyHostToDevice, s));
    if (B > 0) {
        HVC_TRY(hipMemcpyAsync(ctx->d_ra_off, ctx->ra_off.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_i_off, ctx->i_off.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_key, ctx->keys.data(), sizeof(uint32_t) * 2 * B, hipMemcpyHostToDevice, s));
    }
    HVC_TRY(hipStreamSynchronize(s));
    ctx->nd.n_ra = ctx->n_ra;
    ctx->nd.n_i = ctx->n_i;
    ctx->nd.n_nets = B;
    return HVC_OK;
}

// upload of the (merged) network; ra_off / i_off / keys describe the networks in it
static hvc_status upload_impl(hvc_ctx* ctx, const hvc_network* net, int n_nets, const std::vector<int>& ra_off,
                              const std::vector<int>& i_off, const uint32_t* keys, const std::vector<int>& train_net)
{
    const int nra = net->n_ra, ni = net->n_i;
    if (nra >= kMaxKeyNeurons) return fail(ctx, HVC_E_ARG, "too many HVC-RA neurons in one context");
    if (ctx->num_steps >= kMaxKeySteps) return fail(ctx, HVC_E_ARG, "trial has too many steps");
    ctx->n_ra = nra;
    ctx->n_i = ni;
    ctx->E = net->rr_ptr[nra];
    ctx->n_nets = n_nets;
    ctx->ra_off = ra_off;
    ctx->i_off = i_off;
    ctx->keys.assign(keys, keys + 2 * n_nets);
    ctx->rr_ptr.assign(net->rr_ptr, net->rr_ptr + nra + 1);
    ctx->rr_tgt.assign(net->rr_tgt, net->rr_tgt + ctx->E);
    ctx->train_ids.assign(net->train_ids, net->train_ids + net->n_train);
    ctx->train_net = train_net;
    ctx->spread.assign(net->spread, net->spread + net->n_train);

    std::vector<int> perm;
    hvc_status r = rr_order(ctx, net->rr_ptr, net->rr_tgt, net->rr_d, nra, 0, perm);
    if (r != HVC_OK) return r;
    // a single network whose rows list every target: edge index = row start + target
    int dense = n_nets == 1 ? 1 : 0;
    for (int i = 0; i < nra && dense; ++i) {
        if (net->rr_ptr[i + 1] - net->rr_ptr[i] != nra) { dense = 0; break; }
        for (int j = 0; j < nra; ++j)
            if (net->rr_tgt[net->rr_ptr[i] + j] != j) { dense = 0; break; }
    }

    std::vector<int> a_tgt, b_tgt;
    std::vector<double> a_w, a_d, b_w, b_d;
    sort_fixed(net->ra2i_ptr, net->ra2i_tgt, net->ra2i_w, net->ra2i_d, nra, a_tgt, a_w, a_d);
    sort_fixed(net->i2ra_ptr, net->i2ra_tgt, net->i2ra_w, net->i2ra_d, ni, b_tgt, b_w, b_d);

    // inbox capacities: largest in-degree per connection type, plus slack
    Caps c;
    c.nets = n_nets; c.ra = nra; c.i = ni;
    c.rr = ctx->E; c.a2i = (long long)a_tgt.size(); c.i2ra = (long long)b_tgt.size();
    c.cap_gi_ra = max_indeg(net->i2ra_ptr, net->i2ra_tgt, ni, nra, 0) + 8;
    c.cap_ge_ra = max_indeg(net->rr_ptr, net->rr_tgt, nra, nra, 1) + 8;
    c.cap_ge_i = max_indeg(net->ra2i_ptr, net->ra2i_tgt, nra, ni, 0) + 8;
    if ((r = alloc_device(ctx, c)) != HVC_OK) return r;
    ctx->nd.dense = dense;
    if ((r = upload_tables(ctx)) != HVC_OK) return r;

    hipStream_t s = ctx->stream;
    const int E = ctx->E;
    HVC_TRY(hipMemcpyAsync(ctx->d_EL, net->erest, sizeof(double) * nra, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_GCa, net->gca, sizeof(double) * nra, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_rr_ptr, net->rr_ptr, sizeof(int) * (nra + 1), hipMemcpyHostToDevice, s));
    if (E > 0) {
        HVC_TRY(hipMemcpyAsync(ctx->d_rr_tgt, net->rr_tgt, sizeof(int) * E, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_rr_d, net->rr_d, sizeof(double) * E, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_perm, perm.data(), sizeof(int) * E, hipMemcpyHostToDevice, s));
        ctx->rr_pos_stale = true;
    }
    HVC_TRY(hipMemcpyAsync(ctx->d_a_ptr, net->ra2i_ptr, sizeof(int) * (nra + 1), hipMemcpyHostToDevice, s));
    if (!a_tgt.empty()) {
        HVC_TRY(hipMemcpyAsync(ctx->d_a_tgt, a_tgt.data(), sizeof(int) * a_tgt.size(), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_a_w, a_w.data(), sizeof(double) * a_w.size(), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_a_d, a_d.data(), sizeof(double) * a_d.size(), hipMemcpyHostToDevice, s));
    }
    HVC_TRY(hipMemcpyAsync(ctx->d_b_ptr, net->i2ra_ptr, sizeof(int) * (ni + 1), hipMemcpyHostToDevice, s));
    if (!b_tgt.empty()) {
        HVC_TRY(hipMemcpyAsync(ctx->d_b_tgt, b_tgt.data(), sizeof(int) * b_tgt.size(), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_b_w, b_w.data(), sizeof(double) * b_w.size(), hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_b_d, b_d.data(), sizeof(double) * b_d.size(), hipMemcpyHostToDevice, s));
    }
    HVC_TRY(hipStreamSynchronize(s));
    ctx->uploaded = true;
    return set_synapse_state(ctx, 0, nra, net->rr_w, net->act_cnt, net->act_ids, net->sup_cnt, net->sup_ids,
                             net->remodeled);
}

extern "C" hvc_status hvc_upload_network(hvc_ctx* ctx, const hvc_network* net)
{
    if (!ctx || !net) return HVC_E_ARG;
    if (ctx->uploaded) return fail(ctx, HVC_E_STATE, "network already uploaded to this context");
    hvc_status r = check_network(ctx, net, "network");
    if (r != HVC_OK) return r;
    uint32_t key[2] = {ctx->prm.seed_lo, ctx->prm.seed_hi};
    return upload_impl(ctx, net, 1, {0, net->n_ra}, {0, net->n_i}, key, std::vector<int>(net->n_train, 0));
}

extern "C" hvc_status hvc_upload_batch(hvc_ctx* ctx, int n_nets, const hvc_network* nets, const uint32_t* keys)
{
    if (!ctx || !nets || !keys || n_nets <= 0) return HVC_E_ARG;
    if (ctx->uploaded) return fail(ctx, HVC_E_STATE, "networks already uploaded to this context");
    if (n_nets >= kMaxNets) return fail(ctx, HVC_E_ARG, "too many networks in one context");
    for (int b = 0; b < n_nets; ++b) {
        char what[32];
        std::snprintf(what, sizeof what, "network %d", b);
        hvc_status r = check_network(ctx, &nets[b], what);
        if (r != HVC_OK) return r;
    }
    Merged m;
    merge_networks(n_nets, nets, m);
    return upload_impl(ctx, &m.view, n_nets, m.ra_off, m.i_off, keys, m.train_net);
}

static hvc_status check_net_index(hvc_ctx* ctx, int net)
{
    if (!ctx->uploaded) return fail(ctx, HVC_E_STATE, "upload a network first");
    if (net < 0 || net >= ctx->n_nets) return fail(ctx, HVC_E_ARG, "network index out of range");
    return HVC_OK;
}

extern "C" hvc_status hvc_batch_upload_synapses(hvc_ctx* ctx, int net, const hvc_synapses* syn)
{
    if (!ctx || !syn) return HVC_E_ARG;
    hvc_status r = check_net_index(ctx, net);
    if (r != HVC_OK) return r;
    return set_synapse_state(ctx, ctx->ra_off[net], ctx->ra_off[net + 1], syn->w, syn->act_cnt, syn->act_ids,
                             syn->sup_cnt, syn->sup_ids, syn->remodeled);
}

extern "C" hvc_status hvc_upload_synapses(hvc_ctx* ctx, const hvc_synapses* syn)
{
    if (!ctx || !syn) return HVC_E_ARG;
    if (ctx->uploaded && ctx->n_nets != 1) return fail(ctx, HVC_E_STATE, "several networks: use hvc_batch_upload_synapses");
    return hvc_batch_upload_synapses(ctx, 0, syn);
}

extern "C" hvc_status hvc_batch_download_synapses(hvc_ctx* ctx, int net, hvc_synapses* syn)
{
    if (!ctx || !syn) return HVC_E_ARG;
    hvc_status r = check_net_index(ctx, net);
    if (r != HVC_OK) return r;
    const int r0 = ctx->ra_off[net], r1 = ctx->ra_off[net + 1], n = r1 - r0;
    const int E0 = ctx->rr_ptr[r0], E1 = ctx->rr_ptr[r1], En = E1 - E0, Nss = ctx->prm.syn.Nss;
    const int NssA = std::max(Nss, 1);
    NetDev& nd = ctx->nd;
    std::vector<int> alist(En + 1), slist((size_t)n * NssA + 1), acnt(n + 1), scnt(n + 1), rem(n + 1);
    hipStream_t s = ctx->stream;
    if (En > 0) {
        if (syn->w) HVC_TRY(hipMemcpyAsync(syn->w, nd.w + E0, sizeof(double) * En, hipMemcpyDeviceToHost, s));
        HVC_TRY(hipMemcpyAsync(alist.data(), nd.act_list + E0, sizeof(int) * En, hipMemcpyDeviceToHost, s));
    }
    HVC_TRY(hipMemcpyAsync(slist.data(), nd.sup_list + (size_t)r0 * NssA, sizeof(int) * (size_t)n * NssA,
                           hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(acnt.data(), nd.act_cnt + r0, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(scnt.data(), nd.sup_cnt + r0, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(rem.data(), nd.remodeled + r0, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipStreamSynchronize(s));
    int oa = 0, os = 0;
    for (int r = 0; r < n; ++r) {
        const int i = r0 + r;
        if (syn->act_cnt) syn->act_cnt[r] = acnt[r];
        if (syn->sup_cnt) syn->sup_cnt[r] = scnt[r];
        if (syn->remodeled) syn->remodeled[r] = rem[r];
        if (syn->act_ids) for (int k = 0; k < acnt[r]; ++k) syn->act_ids[oa++] = alist[ctx->rr_ptr[i] - E0 + k] - r0;
        if (syn->sup_ids) for (int k = 0; k < scnt[r]; ++k) syn->sup_ids[os++] = slist[(size_t)r * NssA + k] - r0;
    }
    return HVC_OK;
}

extern "C" hvc_status hvc_download_synapses(hvc_ctx* ctx, hvc_synapses* syn)
{
    if (!ctx || !syn) return HVC_E_ARG;
    if (ctx->uploaded && ctx->n_nets != 1) return fail(ctx, HVC_E_STATE, "several networks: use hvc_batch_download_synapses");
    return hvc_batch_download_synapses(ctx, 0, syn);
}

extern "C" hvc_status hvc_patch_neurons(hvc_ctx* ctx, const int* ids, int n, const double* erest, const double* gca)
{
    if (!ctx || (n > 0 && (!ids || !erest || !gca))) return HVC_E_ARG;
    if (!ctx->uploaded) return fail(ctx, HVC_E_STATE, "upload a network first");
    bool range = n > 0;   // consecutive ids: one copy per array
    for (int k = 0; k < n && range; ++k) range = ids[k] == ids[0] + k;
    if (range) {
        if (ids[0] < 0 || ids[n - 1] >= ctx->n_ra) return fail(ctx, HVC_E_ARG, "neuron id out of range");
        HVC_TRY(hipMemcpyAsync((double*)ctx->nd.EL + ids[0], erest, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HVC_TRY(hipMemcpyAsync((double*)ctx->nd.GCa + ids[0], gca, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        n = 0;
    }
    for (int k = 0; k < n; ++k) {
        if (ids[k] < 0 || ids[k] >= ctx->n_ra) return fail(ctx, HVC_E_ARG, "neuron id out of range");
        HVC_TRY(hipMemcpyAsync((double*)ctx->nd.EL + ids[k], erest + k, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HVC_TRY(hipMemcpyAsync((double*)ctx->nd.GCa + ids[k], gca + k, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    HVC_TRY(hipStreamSynchronize(ctx->stream));
    return HVC_OK;
}

// ---------------------------------------------------------------------------
// trials
// ---------------------------------------------------------------------------
static hipEvent_t pool_event(hvc_ctx* ctx)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}
static void tstart(hvc_ctx* ctx, int k, bool on)
{
    if (!on) return;
    size_t idx = ctx->ev_used;
    hipEvent_t a = pool_event(ctx);
    hipEvent_t b = pool_event(ctx);
    if (!a || !b) return;
    hipEventRecord(a, ctx->stream);
    ctx->ev_pending.emplace_back(k, idx);
}
static void tstop(hvc_ctx* ctx, int k, bool on)
{
    if (!on || ctx->ev_pending.empty() || ctx->ev_pending.back().first != k) return;
    hipEventRecord(ctx->ev_pool[ctx->ev_pending.back().second + 1], ctx->stream);
}
static void tcollect(hvc_ctx* ctx)
{
    for (auto& p : ctx->ev_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ctx->ev_pool[p.second], ctx->ev_pool[p.second + 1]) == hipSuccess) {
            ctx->t_sum[p.first] += ms;
            ctx->t_cnt[p.first] += 1;
        }
    }
    ctx->ev_pending.clear();
    ctx->ev_used = 0;
}

// the last trial's spike log of one kind, sorted by (neuron, step), fetched once
static hvc_status sorted_log(hvc_ctx* ctx, int kind, const std::vector<int2>** out)
{
    std::vector<int2>& log = ctx->h_log[kind];
    if (!ctx->h_log_valid[kind]) {
        const int2* d_log = kind == 0 ? ctx->nd.soma_log : (kind == 1 ? ctx->nd.dend_log : ctx->nd.i_log);
        const int count = std::min(ctx->log_counts[kind], ctx->nd.log_cap);
        std::vector<int2>& raw = ctx->h_log_raw;
        raw.resize(count);
        if (count > 0) HVC_TRY(hipMemcpy(raw.data(), d_log, sizeof(int2) * count, hipMemcpyDeviceToHost));
        // counting sort by neuron (stable), then by step within each neuron; the
        // log is appended window by window, so a neuron's entries usually come
        // in step order already and the second pass finds nothing to move
        const int n = kind == 2 ? ctx->n_i : ctx->n_ra;
        std::vector<int>& start = ctx->h_log_start;
        start.assign((size_t)n + 1, 0);
        const int nsteps = (int)ctx->clock.size();
        for (const int2& v : raw) {
            if (v.x < 0 || v.x >= n) return fail(ctx, HVC_E_HIP, "spike log holds a neuron id out of range");
            if (v.y < 0 || v.y >= nsteps) return fail(ctx, HVC_E_HIP, "spike log holds a step out of range");
            ++start[v.x + 1];
        }
        for (int i = 0; i < n; ++i) start[i + 1] += start[i];
        log.resize(count);
        for (const int2& v : raw) log[start[v.x]++] = v;
        for (int i = 0, b = 0; i < n; b = start[i], ++i) {
            auto lo = log.begin() + b, hi = log.begin() + start[i];
            if (hi - lo > 1 && !std::is_sorted(lo, hi, [](const int2& a, const int2& c) { return a.y < c.y; }))
                std::sort(lo, hi, [](const int2& a, const int2& c) { return a.y < c.y; });
        }
        ctx->h_log_valid[kind] = true;
    }
    *out = &log;
    return HVC_OK;
}

// spikes of neurons [n0, n1) of one kind: counts per neuron (local index),
// times neuron by neuron
static hvc_status spikes_range(hvc_ctx* ctx, int kind, int n0, int n1, int* counts, double* times, int cap, int* total)
{
    const std::vector<int2>* log = nullptr;
    hvc_status r = sorted_log(ctx, kind, &log);
    if (r != HVC_OK) return r;
    auto lo = std::lower_bound(log->begin(), log->end(), n0, [](const int2& a, int v) { return a.x < v; });
    auto hi = std::lower_bound(log->begin(), log->end(), n1, [](const int2& a, int v) { return a.x < v; });
    const int count = (int)(hi - lo);
    *total = count;
    if (times && count > cap) return fail(ctx, HVC_E_ARG, "spike-time buffer too small");
    if (counts) std::fill(counts, counts + (n1 - n0), 0);
    int k = 0;
    for (auto it = lo; it != hi; ++it, ++k) {
        if (counts) counts[it->x - n0]++;
        if (times) times[k] = ctx->clock[it->y];
    }
    return HVC_OK;
}

extern "C" hvc_status hvc_set_timing(hvc_ctx* ctx, int enabled)
{
    if (!ctx) return HVC_E_ARG;
    ctx->timing = enabled != 0;
    for (int k = 0; k < kNumTimers; ++k) {
        ctx->t_sum[k] = 0;
        ctx->t_cnt[k] = 0;
    }
    return HVC_OK;
}

extern "C" hvc_status hvc_set_branch_counting(hvc_ctx* ctx, int enabled)
{
    if (!ctx) return HVC_E_ARG;
    HVC_TRY(hipSetDevice(ctx->device));
    if (!enabled) {
        ctx->nd.hits = nullptr;
        return HVC_OK;
    }
    if (!ctx->d_hits) {
        hvc_status st = HVC_OK;
        ctx->d_hits = dalloc<int>(ctx, kNumHits, &st);
        if (st != HVC_OK) return st;
    }
    HVC_TRY(hipMemsetAsync(ctx->d_hits, 0, sizeof(int) * kNumHits, ctx->stream));
    HVC_TRY(hipStreamSynchronize(ctx->stream));
    ctx->nd.hits = ctx->d_hits;
    return HVC_OK;
}

extern "C" hvc_status hvc_branch_counts(hvc_ctx* ctx, int* counts)
{
    if (!ctx || !counts) return HVC_E_ARG;
    if (!ctx->d_hits) return fail(ctx, HVC_E_STATE, "branch counting was never enabled");
    HVC_TRY(hipSetDevice(ctx->device));
    HVC_TRY(hipMemcpyAsync(counts, ctx->d_hits, sizeof(int) * kNumHits, hipMemcpyDeviceToHost, ctx->stream));
    HVC_TRY(hipStreamSynchronize(ctx->stream));
    return HVC_OK;
}

extern "C" hvc_status hvc_kernel_timing(hvc_ctx* ctx, const char* name, double* avg_ms, int* launches)
{
    if (!ctx || !name) return HVC_E_ARG;
    for (int k = 0; k < kNumTimers; ++k) {
        if (std::strcmp(name, kTimerNames[k]) == 0) {
            if (avg_ms) *avg_ms = ctx->t_cnt[k] ? ctx->t_sum[k] / ctx->t_cnt[k] : 0.0;
            if (launches) *launches = ctx->t_cnt[k];
            return HVC_OK;
        }
    }
    return fail(ctx, HVC_E_ARG, std::string("unknown kernel class ") + name);
}

extern "C" hvc_status hvc_batch_spikes(hvc_ctx* ctx, int net, int kind, int* counts, double* times, int cap, int* total)
{
    if (!ctx || kind < 0 || kind > 2 || !total) return HVC_E_ARG;
    hvc_status r = check_net_index(ctx, net);
    if (r != HVC_OK) return r;
    if (!ctx->trial_done) return fail(ctx, HVC_E_STATE, "no trial has run");
    const std::vector<int>& off = kind == 2 ? ctx->i_off : ctx->ra_off;
    return spikes_range(ctx, kind, off[net], off[net + 1], counts, times, cap, total);
}

extern "C" hvc_status hvc_trial_spikes(hvc_ctx* ctx, int kind, int* counts, double* times, int cap, int* total)
{
    if (!ctx || kind < 0 || kind > 2 || !total) return HVC_E_ARG;
    if (!ctx->uploaded) return fail(ctx, HVC_E_STATE, "upload a network first");
    if (!ctx->trial_done) return fail(ctx, HVC_E_STATE, "no trial has run");
    return spikes_range(ctx, kind, 0, kind == 2 ? ctx->n_i : ctx->n_ra, counts, times, cap, total);
}

// One trial of every network in the context (trials[b] for network b); the
// networks advance together, one launch per kernel and coupling window.
// noise events of the chunk that starts at window w (k_noise_events)
// the candidate edges by row position (RrPos) after the order tables changed
static void refresh_rr_pos(hvc_ctx* ctx, const NetDev& nd)
{
    if (!ctx->rr_pos_stale) return;
    launch_rr_pos(ctx->stream, nd, ctx->d_rr_pos, ctx->E);
    ctx->rr_pos_stale = false;
}

static void noise_chunk(hvc_ctx* ctx, const NetDev& nd, int w)
{
    const int gen_ra = ctx->prm.ra_model == 1 && ctx->prm.ra_poisson, gen_i = ctx->prm.i_poisson != 0;
    if (w % kNoiseChunkWindows != 0 || !(gen_ra || gen_i)) return;
    const int last = std::min(w + kNoiseChunkWindows, (int)ctx->windows.size()) - 1;
    launch_noise_events(ctx->stream, nd, ctx->windows[w].first, ctx->windows[last].second, gen_ra, gen_i);
}

// One trial of every resident network.  live (host, one flag per network, or
// nullptr: all) marks the networks that run; the others sit the trial out
// (their neurons are integrated but nothing of them is stored, so they neither
// spike nor change).
static hvc_status run_impl(hvc_ctx* ctx, const hvc_trial* trials, hvc_trial_out* outs, const int* live = nullptr)
{
    const int B = ctx->n_nets;
    int any_growth = 0, any_chain = 0;
    for (int b = 0; b < B; ++b) {
        if (trials[b].mode != 0 && trials[b].mode != 1) return fail(ctx, HVC_E_ARG, "mode must be 0 (growth) or 1 (chain test)");
        any_growth |= trials[b].mode == 0;
        any_chain |= trials[b].mode == 1;
    }
    HVC_TRY(hipSetDevice(ctx->device));
    NetDev& nd = ctx->nd;
    hipStream_t s = ctx->stream;

    // training kick: first step with internal_time > spread + mean_injection_time (NGS:8933-8940);
    // a growth trial kicks only when training, a chain test always (NGS:12785-12801)
    ctx->h_kick.assign(ctx->n_ra, -1);
    for (size_t k = 0; k < ctx->train_ids.size(); ++k) {
        const hvc_trial& tr = trials[ctx->train_net[k]];
        if (tr.mode == 0 && !tr.training) continue;
        double thr = ctx->spread[k] + tr.mean_injection_time;
        for (int t = 1; t < ctx->num_steps; ++t)
            if (ctx->clock[t] > thr) {
                ctx->h_kick[ctx->train_ids[k]] = t;
                break;
            }
    }
    ctx->h_net_trial.resize(B);
    ctx->h_net_first.resize(B);
    ctx->h_net_growth.resize(B);
    for (int b = 0; b < B; ++b) {
        ctx->h_net_trial[b] = trials[b].trial_index;
        ctx->h_net_first[b] = trials[b].first_trial ? 1 : 0;
        ctx->h_net_growth[b] = trials[b].mode == 0 ? 1 : 0;
    }
    HVC_TRY(hipMemcpyAsync(ctx->d_kick, ctx->h_kick.data(), sizeof(int) * ctx->n_ra, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_trial, ctx->h_net_trial.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_first, ctx->h_net_first.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_growth, ctx->h_net_growth.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    nd.net_live = nullptr;
    if (live) {
        ctx->h_live.assign(live, live + B);
        HVC_TRY(hipMemcpyAsync(ctx->d_net_live, ctx->h_live.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
        nd.net_live = ctx->d_net_live;
    }
    refresh_rr_pos(ctx, nd);
    launch_reset(s, nd, ctx->E);
    if (any_chain) launch_chain_edges(s, nd);
    const int nw = (int)ctx->windows.size();
    // the window's view: post-event buffer w % 2.  The sync of window w (the
    // inbox sums into the conductances) runs at the start of window w + 1's
    // integration; after the last window it runs on its own.
    nd.growth_mode = any_growth && any_chain ? 2 : (any_growth ? 1 : 0);
    nd.wl_on = ctx->tuning.deliver_lists;
    NetDev ndw[2] = {nd, nd};
    ndw[1].post_list = nd.post_list + kPostListCap;
    ndw[1].post_list_cnt = nd.post_list_cnt + 1;
    ndw[1].post_list_next_cnt = nd.post_list_cnt;
    for (int w = 0; w < nw; ++w) {
        int t0 = ctx->windows[w].first, t1 = ctx->windows[w].second;
        NetDev nv = ndw[w & 1];
        nv.window = w;   // the delivery lists' index
        bool tm = ctx->timing && (w % 16 == 7);
        noise_chunk(ctx, nd, w);
        tstart(ctx, kTInt, tm);
        // the previous window's sync-point STDP runs in blocks of this window's integration
        const NetDev& pv = ndw[(w + 1) & 1];
        const StdpArgs sa{pv.post_list, pv.post_list_cnt, w > 0 ? ctx->windows[w - 1].second : 0};
        const int stdp = any_growth && w > 0;
        if (ctx->prm.precision == 1)
            launch_integrate_f32(s, nv, t0, t1, ctx->tuning, sa, stdp);
        else
            launch_integrate(s, nv, t0, t1, ctx->tuning, sa, stdp);
        tstop(ctx, kTInt, tm);
        tstart(ctx, kTDeliver, tm);
        launch_deliver(s, nv, t0, t1, ctx->tuning);
        tstop(ctx, kTDeliver, tm);
    }
    if (any_growth && nw > 0) {
        const NetDev& lv = ndw[(nw - 1) & 1];
        launch_stdp(s, nd, StdpArgs{lv.post_list, lv.post_list_cnt, ctx->windows[nw - 1].second});
    }
    tstart(ctx, kTSync, ctx->timing);
    launch_sync(s, nd, ctx->tuning);
    tstop(ctx, kTSync, ctx->timing);
    if (any_growth) {
        tstart(ctx, kTDecay, ctx->timing);
        launch_decay(s, nd);
        tstop(ctx, kTDecay, ctx->timing);
    }
    HVC_TRY(hipGetLastError());
    unsigned int err = 0;
    int logc[4] = {0, 0, 0, 0};
    std::vector<int> warn(3 * (size_t)B);
    HVC_TRY(hipMemcpyAsync(&err, nd.err, sizeof(err), hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(logc, nd.log_cnt, sizeof(logc), hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(warn.data(), nd.warn, sizeof(int) * 3 * B, hipMemcpyDeviceToHost, s));
    std::vector<double> vs(ctx->n_ra), vd(ctx->n_ra);
    HVC_TRY(hipMemcpyAsync(vs.data(), nd.ra[0], sizeof(double) * ctx->n_ra, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(vd.data(), nd.ra[1], sizeof(double) * ctx->n_ra, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipStreamSynchronize(s));
    tcollect(ctx);
    for (int k = 0; k < 3; ++k) ctx->h_log_valid[k] = false;
    if (err) {
        // no spike query reads this trial's logs, which may have overflowed
        ctx->trial_done = false;
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "engine capacity exceeded (flags 0x%x); trial result discarded, synapse state undefined", err);
        return fail(ctx, HVC_E_CAPACITY, buf);
    }
    ctx->trial_done = true;
    for (int k = 0; k < 3; ++k) ctx->log_counts[k] = std::min(logc[k], ctx->nd.log_cap);
    // check_bad_values (HH2_test.cpp:244-260), per network
    int any_bad = 0;
    for (int b = 0; b < B; ++b) {
        int bad = 0;
        for (int i = ctx->ra_off[b]; i < ctx->ra_off[b + 1]; ++i) {
            if ((vs[i] > 100) || (vs[i] < -200) || std::isnan(vs[i]) || (vd[i] > 100) || (vd[i] < -200) ||
                std::isnan(vd[i])) {
                bad = -1;
                break;
            }
        }
        any_bad |= bad;
        if (!outs) continue;
        hvc_trial_out* out = &outs[b];
        const int r0 = ctx->ra_off[b], r1 = ctx->ra_off[b + 1], i0 = ctx->i_off[b], i1 = ctx->i_off[b + 1];
        hvc_status r;
        if ((r = spikes_range(ctx, 0, r0, r1, out->soma_counts, out->soma_times, out->soma_cap, &out->soma_total)) != HVC_OK)
            return r;
        if ((r = spikes_range(ctx, 1, r0, r1, out->dend_counts, out->dend_times, out->dend_cap, &out->dend_total)) != HVC_OK)
            return r;
        if ((r = spikes_range(ctx, 2, i0, i1, out->i_counts, out->i_times, out->i_cap, &out->i_total)) != HVC_OK)
            return r;
        out->bad_values = bad;
        for (int k = 0; k < 3; ++k) out->warnings[k] = warn[3 * b + k];
        if (out->ra_state && r1 > r0) {
            std::vector<double> col(r1 - r0);
            for (int k = 0; k < 12; ++k) {
                HVC_TRY(hipMemcpy(col.data(), nd.ra[k] + r0, sizeof(double) * (r1 - r0), hipMemcpyDeviceToHost));
                for (int i = 0; i < r1 - r0; ++i) out->ra_state[(size_t)i * 12 + k] = col[i];
            }
        }
        if (out->i_state && i1 > i0) {
            std::vector<double> col(i1 - i0);
            for (int k = 0; k < 7; ++k) {
                HVC_TRY(hipMemcpy(col.data(), nd.in[k] + i0, sizeof(double) * (i1 - i0), hipMemcpyDeviceToHost));
                for (int i = 0; i < i1 - i0; ++i) out->i_state[(size_t)i * 7 + k] = col[i];
            }
        }
    }
    if (any_bad) return fail(ctx, HVC_E_NUMERIC, "check_bad_values: HVC-RA voltage out of [-200, 100] or NaN");
    return HVC_OK;
}

extern "C" hvc_status hvc_run_trial(hvc_ctx* ctx, const hvc_trial* tr, hvc_trial_out* out)
{
    if (!ctx || !tr) return HVC_E_ARG;
    if (!ctx->uploaded) return fail(ctx, HVC_E_STATE, "upload a network first");
    if (ctx->n_nets != 1) return fail(ctx, HVC_E_STATE, "several networks: use hvc_run_batch");
    return run_impl(ctx, tr, out);
}

extern "C" hvc_status hvc_run_batch(hvc_ctx* ctx, const hvc_trial* trials, hvc_trial_out* outs)
{
    if (!ctx || !trials) return HVC_E_ARG;
    if (!ctx->uploaded) return fail(ctx, HVC_E_STATE, "upload networks first");
    return run_impl(ctx, trials, outs);
}

// ---------------------------------------------------------------------------
// autotuner: launch shapes of the per-window kernels for the resident networks
// ---------------------------------------------------------------------------
static hvc_status check_tuning(hvc_ctx* ctx, const hvc_tuning* t)
{
    auto ok_block = [](int b, int hi) { return b >= 64 && b <= hi && (b % 64) == 0; };
    if (!ok_block(t->integrate_block, 256) || !ok_block(t->deliver_block, 256) || !ok_block(t->sync_block, 512) ||
        t->i_lanes < 0 || t->i_lanes > 3 || (t->ra_lanes != 0 && t->ra_lanes != 1) ||
        (t->deliver_lists != 0 && t->deliver_lists != 1))
        return fail(ctx, HVC_E_ARG, "tuning out of range (blocks: multiples of 64 up to 256, sync up to 512)");
    return HVC_OK;
}

extern "C" hvc_status hvc_get_tuning(hvc_ctx* ctx, hvc_tuning* t)
{
    if (!ctx || !t) return HVC_E_ARG;
    t->integrate_block = ctx->tuning.integrate_block;
    t->i_lanes = ctx->tuning.i_lanes;
    t->ra_lanes = ctx->tuning.ra_lanes;
    t->deliver_block = ctx->tuning.deliver_block;
    t->sync_block = ctx->tuning.sync_block;
    t->deliver_lists = ctx->tuning.deliver_lists;
    return HVC_OK;
}

extern "C" hvc_status hvc_set_tuning(hvc_ctx* ctx, const hvc_tuning* t)
{
    if (!ctx || !t) return HVC_E_ARG;
    ApiLock lk(ctx);   // a pool's serving thread reads the shape at every round
    hvc_status r = check_tuning(ctx, t);
    if (r != HVC_OK) return r;
    ctx->tuning.integrate_block = t->integrate_block;
    ctx->tuning.i_lanes = t->i_lanes;
    ctx->tuning.ra_lanes = t->ra_lanes;
    ctx->tuning.deliver_block = t->deliver_block;
    ctx->tuning.sync_block = t->sync_block;
    ctx->tuning.deliver_lists = t->deliver_lists;
    return HVC_OK;
}

// A scratch chain test of every resident network (synapse state untouched, the
// training kick 5 ms after the start so that chains are active) through the
// first `windows` coupling windows; average device time per window of each
// kernel class under tuning tu.
static hvc_status timed_windows(hvc_ctx* ctx, int windows, const Tuning& tu, double ms[3])
{
    NetDev& nd = ctx->nd;
    hipStream_t s = ctx->stream;
    const int B = ctx->n_nets;
    ctx->h_kick.assign(ctx->n_ra, -1);
    for (size_t k = 0; k < ctx->train_ids.size(); ++k) {
        double thr = ctx->spread[k] + 5.0;
        for (int t = 1; t < ctx->num_steps; ++t)
            if (ctx->clock[t] > thr) {
                ctx->h_kick[ctx->train_ids[k]] = t;
                break;
            }
    }
    std::vector<int> zeros(B, 0);
    std::vector<uint32_t> trial(B, 0);
    HVC_TRY(hipMemcpyAsync(ctx->d_kick, ctx->h_kick.data(), sizeof(int) * ctx->n_ra, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_growth, zeros.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_first, zeros.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_net_trial, trial.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice, s));
    refresh_rr_pos(ctx, nd);
    nd.growth_mode = 0;   // chain tests only
    nd.net_live = nullptr;
    launch_reset(s, nd, ctx->E);
    launch_chain_edges(s, nd);
    const int nw = std::min(windows, (int)ctx->windows.size());
    CallEvents events;
    HVC_TRY(events.create(6 * (size_t)nw));
    std::vector<hipEvent_t>& ev = events.ev;
    hvc_status r = HVC_OK;
    for (int w = 0; w < nw; ++w) {
        int t0 = ctx->windows[w].first, t1 = ctx->windows[w].second;
        noise_chunk(ctx, nd, w);
        hipEventRecord(ev[6 * w + 0], s);
        const StdpArgs none{nd.post_list, nd.post_list_cnt, 0};
        NetDev nv = nd;
        nv.window = w;
        nv.wl_on = tu.deliver_lists;
        if (ctx->prm.precision == 1)
            launch_integrate_f32(s, nv, t0, t1, tu, none, 0);
        else
            launch_integrate(s, nv, t0, t1, tu, none, 0);
        hipEventRecord(ev[6 * w + 1], s);
        launch_deliver(s, nv, t0, t1, tu);
        hipEventRecord(ev[6 * w + 2], s);
        hipEventRecord(ev[6 * w + 3], s);
        launch_sync(s, nd, tu);
        hipEventRecord(ev[6 * w + 4], s);
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) r = fail(ctx, HVC_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
    ms[0] = ms[1] = ms[2] = 0;
    for (int w = 0; w < nw && r == HVC_OK; ++w) {
        float a = 0, b = 0, c = 0;
        hipEventElapsedTime(&a, ev[6 * w + 0], ev[6 * w + 1]);
        hipEventElapsedTime(&b, ev[6 * w + 1], ev[6 * w + 2]);
        hipEventElapsedTime(&c, ev[6 * w + 3], ev[6 * w + 4]);
        ms[0] += a; ms[1] += b; ms[2] += c;
    }
    for (int k = 0; k < 3; ++k) ms[k] /= std::max(nw, 1);
    unsigned int err = 0;
    HVC_TRY(hipMemcpy(&err, nd.err, sizeof(err), hipMemcpyDeviceToHost));
    if (err && r == HVC_OK) r = fail(ctx, HVC_E_CAPACITY, "autotune: engine capacity exceeded in the scratch run");
    ctx->trial_done = false;
    return r;
}

extern "C" hvc_status hvc_autotune(hvc_ctx* ctx, int windows, hvc_tuning* best, double* ms_out)
{
    if (!ctx || windows <= 0) return HVC_E_ARG;
    ApiLock lk(ctx);   // between a pool's serving rounds
    if (!ctx->uploaded || ctx->n_nets == 0) return fail(ctx, HVC_E_STATE, "upload networks first");
    HVC_TRY(hipSetDevice(ctx->device));
    Tuning tu = ctx->tuning;
    double ms[3], best_ms[3];
    hvc_status r;
    // integrate: block size and the HVC-I and HVC-RA layouts (four, five and
    // eight lanes per HVC-I neuron exist for DP8 only, four per HVC-RA neuron
    // for RK4 only)
    double bi = 1e30;
    Tuning ti = tu;
    for (int ra_lanes = (ctx->prm.ra_model == 1 ? 1 : 0); ra_lanes >= 0; --ra_lanes)
        for (int lanes = (ctx->prm.i_model == 0 ? 3 : 0); lanes >= 0; --lanes) {
            for (int bs : {64, 128, 256}) {
                Tuning c = tu;
                c.integrate_block = bs;
                c.i_lanes = lanes;
                c.ra_lanes = ra_lanes;
                if ((r = timed_windows(ctx, windows, c, ms)) != HVC_OK) return r;
                if (ms[0] < bi) { bi = ms[0]; ti = c; }
            }
        }
    tu = ti;
    double bd = 1e30;
    Tuning td = tu;
    for (int lists = 1; lists >= 0; --lists)
        for (int bs : {64, 128, 256}) {
            Tuning c = tu;
            c.deliver_block = bs;
            c.deliver_lists = lists;
            if ((r = timed_windows(ctx, windows, c, ms)) != HVC_OK) return r;
            if (ms[1] < bd) { bd = ms[1]; td = c; }
        }
    tu = td;
    double bsy = 1e30;
    Tuning ts = tu;
    for (int bs : {128, 256, 512}) {
        Tuning c = tu;
        c.sync_block = bs;
        if ((r = timed_windows(ctx, windows, c, ms)) != HVC_OK) return r;
        if (ms[2] < bsy) { bsy = ms[2]; ts = c; }
    }
    tu = ts;
    if ((r = timed_windows(ctx, windows, tu, best_ms)) != HVC_OK) return r;
    ctx->tuning = tu;
    if (best) hvc_get_tuning(ctx, best);
    if (ms_out) for (int k = 0; k < 3; ++k) ms_out[k] = best_ms[k];
    return HVC_OK;
}

// ---------------------------------------------------------------------------
// serving pool: networks admitted and released between trials
// ---------------------------------------------------------------------------
// Device bytes per element of the pool's arrays (the sizes alloc_device uses).
struct PoolCost {
    double ra, i, rr, a2i, i2ra, net;
};

static PoolCost pool_cost(const hvc_params* prm, int max_indeg, long long spikes_per_neuron)
{
    const double NssA = std::max(prm->syn.Nss, 1);
    const double inbox = (double)(max_indeg + 8) * sizeof(Record) + sizeof(int);
    const double logs = 3.0 * spikes_per_neuron * sizeof(int2);
    const double inflight = sizeof(FlHead) + (kMaxInflight - kHeadSlots) * sizeof(FlSlot) + sizeof(int);
    PoolCost c;
    const double noise = (double)noise_cap(prm) * sizeof(NoiseEvent) + sizeof(int);
    c.ra = 12 * 8 + 4 + 4 * 8 + 4 + 2 * 8 + 4 + 2 * 8 + kPostRing * 8 + 16 + 2 * 4 + 4 + 3 * 4 + 4 * NssA +
           4.0 * kMaxStreams * 4 + 4 + kMaxStreams * NssA * 4 + 4 + 2 * inflight + 2 * inbox + 4 + logs + noise +
           2 * 3 * 4;
    c.i = 7 * 8 + 4 + 2 * 8 + 4 + 8 + 4 + inflight + inbox + 4 + logs + noise + 3 * 4;
    c.rr = 4 + 8 + 4 + sizeof(RrPos) + 8 + 1 + 1 + 4 + kPreRing * 8 + 4 + 4;
    c.a2i = 4 + 8 + 8;
    c.i2ra = 4 + 8 + 8;
    c.net = 2 * 4 + 2 * 4 + 4 * 4 + 3 * 4;
    return c;
}

extern "C" hvc_status hvc_pool_footprint(const hvc_params* prm, const hvc_pool_caps* caps, int n_ra, int n_i,
                                         long long n_rr, long long n_ra2i, long long n_i2ra, size_t* network_bytes,
                                         size_t* pool_bytes)
{
    if (!prm || !caps) return HVC_E_ARG;
    const long long spn = caps->spikes_per_neuron > 0 ? caps->spikes_per_neuron : 256;
    PoolCost c = pool_cost(prm, caps->max_in_degree, spn);
    if (network_bytes)
        *network_bytes = (size_t)(n_ra * c.ra + n_i * c.i + n_rr * c.rr + n_ra2i * c.a2i + n_i2ra * c.i2ra + c.net);
    if (pool_bytes)
        *pool_bytes = (size_t)((double)caps->max_ra * c.ra + (double)caps->max_i * c.i + (double)caps->max_rr * c.rr +
                               (double)caps->max_ra2i * c.a2i + (double)caps->max_i2ra * c.i2ra +
                               (double)caps->max_nets * c.net);
    return HVC_OK;
}

extern "C" hvc_status hvc_pool_create(const hvc_params* prm, int device, const hvc_pool_caps* caps, hvc_ctx** out)
{
    if (!caps || !out) return HVC_E_ARG;
    hvc_status r = hvc_create(prm, device, out);
    if (r != HVC_OK) return r;
    hvc_ctx* ctx = *out;
    if (caps->max_nets <= 0 || caps->max_nets >= kMaxNets || caps->max_ra <= 0 || caps->max_i < 0 ||
        caps->max_rr < 0 || caps->max_ra2i < 0 || caps->max_i2ra < 0 || caps->max_in_degree < 0 ||
        caps->max_ra >= kMaxKeyNeurons)
        return fail(ctx, HVC_E_ARG, "pool capacities out of range");
    if (ctx->num_steps >= kMaxKeySteps) return fail(ctx, HVC_E_ARG, "trial has too many steps");
    Caps c;
    c.nets = caps->max_nets; c.ra = caps->max_ra; c.i = caps->max_i;
    c.rr = caps->max_rr; c.a2i = caps->max_ra2i; c.i2ra = caps->max_i2ra;
    c.cap_gi_ra = c.cap_ge_ra = c.cap_ge_i = caps->max_in_degree + 8;
    const long long spn = caps->spikes_per_neuron > 0 ? caps->spikes_per_neuron : 256;
    c.log_cap = spn * ((long long)caps->max_ra + caps->max_i);
    if ((r = alloc_device(ctx, c)) != HVC_OK) return r;
    hvc_status st = HVC_OK;
    ctx->d_net_live = dalloc<int>(ctx, (size_t)std::max(caps->max_nets, 1), &st);
    if (st != HVC_OK) return st;
    ctx->serve.reset(new Serving());
    ctx->pool = true;
    ctx->n_ra = ctx->n_i = ctx->E = 0;
    ctx->n_nets = 0;
    ctx->ra_off.assign(1, 0);
    ctx->i_off.assign(1, 0);
    ctx->keys.clear();
    ctx->rr_ptr.assign(1, 0);
    ctx->rr_tgt.clear();
    ctx->nd.dense = 0;
    int zero = 0;
    HVC_TRY(hipMemcpy(ctx->d_rr_ptr, &zero, sizeof(int), hipMemcpyHostToDevice));
    HVC_TRY(hipMemcpy(ctx->d_a_ptr, &zero, sizeof(int), hipMemcpyHostToDevice));
    HVC_TRY(hipMemcpy(ctx->d_b_ptr, &zero, sizeof(int), hipMemcpyHostToDevice));
    if ((r = upload_tables(ctx)) != HVC_OK) return r;
    ctx->uploaded = true;
    return HVC_OK;
}

// host bookkeeping of the pool from its resident networks (positions = order)
static void pool_rebuild_host(hvc_ctx* ctx)
{
    ctx->n_nets = (int)ctx->pnets.size();
    ctx->ra_off.assign(1, 0);
    ctx->i_off.assign(1, 0);
    ctx->keys.clear();
    ctx->train_ids.clear();
    ctx->train_net.clear();
    ctx->spread.clear();
    ctx->E = 0;
    ctx->a_tot = ctx->b_tot = 0;
    for (size_t p = 0; p < ctx->pnets.size(); ++p) {
        const PoolNet& pn = ctx->pnets[p];
        const int ro = ctx->ra_off.back();
        for (size_t k = 0; k < pn.train_ids.size(); ++k) {
            ctx->train_ids.push_back(pn.train_ids[k] + ro);
            ctx->train_net.push_back((int)p);
            ctx->spread.push_back(pn.spread[k]);
        }
        ctx->ra_off.push_back(ro + pn.n_ra);
        ctx->i_off.push_back(ctx->i_off.back() + pn.n_i);
        ctx->keys.push_back(pn.key[0]);
        ctx->keys.push_back(pn.key[1]);
        ctx->E += pn.E;
        ctx->a_tot += pn.A;
        ctx->b_tot += pn.Bn;
    }
    ctx->n_ra = ctx->ra_off.back();
    ctx->n_i = ctx->i_off.back();
}

extern "C" hvc_status hvc_pool_admit(hvc_ctx* ctx, const hvc_network* net, uint32_t key_lo, uint32_t key_hi,
                                     int* handle)
{
    if (!ctx || !net || !handle) return HVC_E_ARG;
    if (!ctx->pool) return fail(ctx, HVC_E_STATE, "not a pool context (hvc_pool_create)");
    ApiLock lk(ctx);
    hvc_status r = check_network(ctx, net, "network");
    if (r != HVC_OK) return r;
    const int n = net->n_ra, ni = net->n_i;
    const int E = net->rr_ptr[n], A = net->ra2i_ptr[n], Bn = net->i2ra_ptr[ni];
    const int indeg = std::max(max_indeg(net->i2ra_ptr, net->i2ra_tgt, ni, n, 0),
                               std::max(max_indeg(net->rr_ptr, net->rr_tgt, n, n, 1),
                                        max_indeg(net->ra2i_ptr, net->ra2i_tgt, n, ni, 0)));
    if (indeg > ctx->caps_indeg) return fail(ctx, HVC_E_ARG, "network in-degree exceeds the pool's max_in_degree");
    if (ctx->n_nets + 1 > ctx->caps_nets || ctx->n_ra + n > ctx->caps_ra || ctx->n_i + ni > ctx->caps_i ||
        ctx->E + E > ctx->caps_rr || ctx->a_tot + A > ctx->caps_a2i || ctx->b_tot + Bn > ctx->caps_i2ra)
        return fail(ctx, HVC_E_CAPACITY, "the pool has no room for this network now");
    const int ro = ctx->n_ra, io = ctx->n_i, eo = ctx->E;
    const long long ao = ctx->a_tot, bo = ctx->b_tot;
    std::vector<int> perm;
    if ((r = rr_order(ctx, net->rr_ptr, net->rr_tgt, net->rr_d, n, eo, perm)) != HVC_OK) return r;
    std::vector<int> a_tgt, b_tgt;
    std::vector<double> a_w, a_d, b_w, b_d;
    sort_fixed(net->ra2i_ptr, net->ra2i_tgt, net->ra2i_w, net->ra2i_d, n, a_tgt, a_w, a_d);
    sort_fixed(net->i2ra_ptr, net->i2ra_tgt, net->i2ra_w, net->i2ra_d, ni, b_tgt, b_w, b_d);
    std::vector<int> rp(n + 1), rt(E), ap(n + 1), bp(ni + 1);
    for (int k = 0; k <= n; ++k) rp[k] = net->rr_ptr[k] + eo;
    for (int k = 0; k < E; ++k) rt[k] = net->rr_tgt[k] + ro;
    for (int k = 0; k <= n; ++k) ap[k] = net->ra2i_ptr[k] + (int)ao;
    for (int k = 0; k <= ni; ++k) bp[k] = net->i2ra_ptr[k] + (int)bo;
    for (auto& x : a_tgt) x += io;
    for (auto& x : b_tgt) x += ro;
    hipStream_t s = ctx->stream;
    HVC_TRY(hipMemcpyAsync(ctx->d_rr_ptr + ro, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_a_ptr + ro, ap.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_b_ptr + io, bp.data(), sizeof(int) * (ni + 1), hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_EL + ro, net->erest, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_GCa + ro, net->gca, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (E > 0) {
        HVC_TRY(hipMemcpyAsync(ctx->d_rr_tgt + eo, rt.data(), sizeof(int) * E, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_rr_d + eo, net->rr_d, sizeof(double) * E, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_perm + eo, perm.data(), sizeof(int) * E, hipMemcpyHostToDevice, s));
        ctx->rr_pos_stale = true;
    }
    if (A > 0) {
        HVC_TRY(hipMemcpyAsync(ctx->d_a_tgt + ao, a_tgt.data(), sizeof(int) * A, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_a_w + ao, a_w.data(), sizeof(double) * A, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_a_d + ao, a_d.data(), sizeof(double) * A, hipMemcpyHostToDevice, s));
    }
    if (Bn > 0) {
        HVC_TRY(hipMemcpyAsync(ctx->d_b_tgt + bo, b_tgt.data(), sizeof(int) * Bn, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_b_w + bo, b_w.data(), sizeof(double) * Bn, hipMemcpyHostToDevice, s));
        HVC_TRY(hipMemcpyAsync(ctx->d_b_d + bo, b_d.data(), sizeof(double) * Bn, hipMemcpyHostToDevice, s));
    }
    HVC_TRY(hipStreamSynchronize(s));
    PoolNet pn;
    pn.handle = ctx->next_handle++;
    pn.n_ra = n; pn.n_i = ni; pn.E = E; pn.A = A; pn.Bn = Bn;
    pn.key[0] = key_lo; pn.key[1] = key_hi;
    pn.train_ids.assign(net->train_ids, net->train_ids + net->n_train);
    pn.spread.assign(net->spread, net->spread + net->n_train);
    ctx->pnets.push_back(pn);
    // host mirror of the candidate rows (edge lookups of the synapse-state calls)
    ctx->rr_ptr.resize(ro + n + 1);
    for (int k = 0; k <= n; ++k) ctx->rr_ptr[ro + k] = rp[k];
    ctx->rr_tgt.insert(ctx->rr_tgt.end(), rt.begin(), rt.end());
    pool_rebuild_host(ctx);
    if ((r = upload_tables(ctx)) != HVC_OK) return r;
    ctx->trial_done = false;
    *handle = pn.handle;
    return set_synapse_state(ctx, ro, ro + n, net->rr_w, net->act_cnt, net->act_ids, net->sup_cnt, net->sup_ids,
                             net->remodeled);
}

static int pool_position(const hvc_ctx* ctx, int handle)
{
    for (size_t p = 0; p < ctx->pnets.size(); ++p)
        if (ctx->pnets[p].handle == handle) return (int)p;
    return -1;
}

// moves count elements of a device array from index `from` to index `to` (to <
// from, ranges may overlap) through a staging buffer; reloc != 0 adds reloc to
// every moved int32 element (ids and edge indices that shift with the move)
static hvc_status move_down(hvc_ctx* ctx, void* base, size_t esz, long long from, long long to, long long count,
                            int reloc)
{
    if (count <= 0) return HVC_OK;
    const size_t bytes = (size_t)count * esz;
    if (bytes > ctx->stage_bytes) {
        if (ctx->d_stage) HVC_TRY(hipFree(ctx->d_stage));
        ctx->d_stage = nullptr;
        ctx->stage_bytes = 0;
        HVC_TRY(hipMalloc(&ctx->d_stage, bytes));
        ctx->stage_bytes = bytes;
    }
    char* b = (char*)base;
    hipStream_t s = ctx->stream;
    HVC_TRY(hipMemcpyAsync(ctx->d_stage, b + from * esz, bytes, hipMemcpyDeviceToDevice, s));
    HVC_TRY(hipMemcpyAsync(b + to * esz, ctx->d_stage, bytes, hipMemcpyDeviceToDevice, s));
    if (reloc != 0) launch_add_i32(s, (int*)(b + to * esz), count, reloc);
    return HVC_OK;
}

extern "C" hvc_status hvc_pool_release(hvc_ctx* ctx, int handle)
{
    if (!ctx) return HVC_E_ARG;
    if (!ctx->pool) return fail(ctx, HVC_E_STATE, "not a pool context (hvc_pool_create)");
    ApiLock lk(ctx);
    const int p = pool_position(ctx, handle);
    if (p < 0) return fail(ctx, HVC_E_ARG, "unknown network handle");
    {
        std::lock_guard<std::mutex> lq(ctx->serve->q_mu);
        for (const auto& kv : ctx->serve->tickets)
            if (kv.second.handle == handle && !kv.second.done)
                return fail(ctx, HVC_E_STATE, "the network has submitted trials that have not finished");
    }
    const PoolNet gone = ctx->pnets[p];
    const int dn = gone.n_ra, di = gone.n_i, dE = gone.E, dA = gone.A, dB = gone.Bn;
    const int NssA = std::max(ctx->prm.syn.Nss, 1);
    NetDev& nd = ctx->nd;
    hvc_status r;
    // networks after the released one move down, with their ids and edge indices
    for (size_t q = p + 1; q < ctx->pnets.size(); ++q) {
        const PoolNet& pn = ctx->pnets[q];
        const long long ro = ctx->ra_off[q], io = ctx->i_off[q];
        long long eo = 0, ao = 0, bo = 0;
        for (size_t k = 0; k < q; ++k) {
            eo += ctx->pnets[k].E;
            ao += ctx->pnets[k].A;
            bo += ctx->pnets[k].Bn;
        }
        struct Mv { void* base; size_t esz; long long from, to, count; int reloc; };
        const Mv mv[] = {
            {ctx->d_EL, 8, ro, ro - dn, pn.n_ra, 0},
            {ctx->d_GCa, 8, ro, ro - dn, pn.n_ra, 0},
            {nd.act_cnt, 4, ro, ro - dn, pn.n_ra, 0},
            {nd.sup_cnt, 4, ro, ro - dn, pn.n_ra, 0},
            {nd.remodeled, 4, ro, ro - dn, pn.n_ra, 0},
            {nd.sup_list, 4, ro * NssA, (ro - dn) * NssA, (long long)pn.n_ra * NssA, -dn},
            {ctx->d_rr_ptr, 4, ro, ro - dn, pn.n_ra, -dE},
            {ctx->d_a_ptr, 4, ro, ro - dn, pn.n_ra, -dA},
            {ctx->d_b_ptr, 4, io, io - di, pn.n_i, -dB},
            {ctx->d_rr_tgt, 4, eo, eo - dE, pn.E, -dn},
            {ctx->d_rr_d, 8, eo, eo - dE, pn.E, 0},
            {ctx->d_perm, 4, eo, eo - dE, pn.E, -dE},
            {nd.w, 8, eo, eo - dE, pn.E, 0},
            {nd.act, 1, eo, eo - dE, pn.E, 0},
            {nd.sup, 1, eo, eo - dE, pn.E, 0},
            {nd.act_list, 4, eo, eo - dE, pn.E, -dn},
            {ctx->d_a_tgt, 4, ao, ao - dA, pn.A, -di},
            {ctx->d_a_w, 8, ao, ao - dA, pn.A, 0},
            {ctx->d_a_d, 8, ao, ao - dA, pn.A, 0},
            {ctx->d_b_tgt, 4, bo, bo - dB, pn.Bn, -dn},
            {ctx->d_b_w, 8, bo, bo - dB, pn.Bn, 0},
            {ctx->d_b_d, 8, bo, bo - dB, pn.Bn, 0},
        };
        for (const Mv& m : mv)
            if ((r = move_down(ctx, m.base, m.esz, m.from, m.to, m.count, m.reloc)) != HVC_OK) return r;
        ctx->rr_pos_stale = true;
    }
    // host mirror of the candidate rows
    const int ro_g = ctx->ra_off[p];
    const int eo_g = ctx->rr_ptr[ro_g];
    ctx->rr_ptr.erase(ctx->rr_ptr.begin() + ro_g, ctx->rr_ptr.begin() + ro_g + dn);
    for (size_t k = ro_g; k < ctx->rr_ptr.size(); ++k) ctx->rr_ptr[k] -= dE;
    ctx->rr_tgt.erase(ctx->rr_tgt.begin() + eo_g, ctx->rr_tgt.begin() + eo_g + dE);
    for (size_t k = eo_g; k < ctx->rr_tgt.size(); ++k) ctx->rr_tgt[k] -= dn;
    ctx->pnets.erase(ctx->pnets.begin() + p);
    pool_rebuild_host(ctx);
    hipStream_t s = ctx->stream;
    const int te = ctx->E, ta = (int)ctx->a_tot, tb = (int)ctx->b_tot;
    HVC_TRY(hipMemcpyAsync(ctx->d_rr_ptr + ctx->n_ra, &te, sizeof(int), hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_a_ptr + ctx->n_ra, &ta, sizeof(int), hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(ctx->d_b_ptr + ctx->n_i, &tb, sizeof(int), hipMemcpyHostToDevice, s));
    HVC_TRY(hipStreamSynchronize(s));
    ctx->trial_done = false;
    return upload_tables(ctx);
}

extern "C" hvc_status hvc_pool_run(hvc_ctx* ctx, int n, const int* handles, const hvc_trial* trials,
                                   hvc_trial_out* outs)
{
    if (!ctx || (n > 0 && (!handles || !trials))) return HVC_E_ARG;
    if (!ctx->pool) return fail(ctx, HVC_E_STATE, "not a pool context (hvc_pool_create)");
    ApiLock lk(ctx);
    if (n != ctx->n_nets) return fail(ctx, HVC_E_ARG, "hvc_pool_run needs one trial for every resident network");
    if (n == 0) return HVC_OK;
    std::vector<hvc_trial> tp(n);
    std::vector<hvc_trial_out> op(n);
    std::vector<int> where(n, -1);
    for (int k = 0; k < n; ++k) {
        int p = pool_position(ctx, handles[k]);
        if (p < 0 || where[p] >= 0) return fail(ctx, HVC_E_ARG, "unknown or repeated network handle");
        where[p] = k;
        tp[p] = trials[k];
        if (outs) op[p] = outs[k];
    }
    hvc_status r = run_impl(ctx, tp.data(), outs ? op.data() : nullptr);
    if (outs)
        for (int p = 0; p < n; ++p) outs[where[p]] = op[p];
    return r;
}

extern "C" hvc_status hvc_pool_spikes(hvc_ctx* ctx, int handle, int kind, int* counts, double* times, int cap,
                                      int* total)
{
    if (!ctx) return HVC_E_ARG;
    ApiLock lk(ctx);
    const int p = pool_position(ctx, handle);
    if (p < 0) return fail(ctx, HVC_E_ARG, "unknown network handle");
    return hvc_batch_spikes(ctx, p, kind, counts, times, cap, total);
}

extern "C" hvc_status hvc_pool_download_synapses(hvc_ctx* ctx, int handle, hvc_synapses* syn)
{
    if (!ctx) return HVC_E_ARG;
    ApiLock lk(ctx);
    const int p = pool_position(ctx, handle);
    if (p < 0) return fail(ctx, HVC_E_ARG, "unknown network handle");
    return hvc_batch_download_synapses(ctx, p, syn);
}

// ---------------------------------------------------------------------------
// asynchronous serving: hvc_submit / hvc_poll
// ---------------------------------------------------------------------------
// One round of the serving thread (api_mu held): the picked tickets' trials
// run together, the other resident networks sit the round out; results go
// into the tickets.
static void serve_round(hvc_ctx* ctx, const std::vector<uint64_t>& pick, std::vector<ServeTicket>& res)
{
    const int B = ctx->n_nets;
    std::vector<hvc_trial> tp(B, hvc_trial{1, 0, 0u, 0, 100.0});
    std::vector<int> live(B, 0), at(pick.size(), -1);
    for (size_t k = 0; k < pick.size(); ++k) {
        const int p = pool_position(ctx, res[k].handle);
        if (p < 0) {
            res[k].status = HVC_E_STATE;
            res[k].err = "the network is not resident";
            continue;
        }
        tp[p] = res[k].trial;
        live[p] = 1;
        at[k] = p;
    }
    if (std::find(live.begin(), live.end(), 1) == live.end()) return;
    std::vector<hvc_trial_out> op(B, hvc_trial_out{});
    hvc_status r = run_impl(ctx, tp.data(), op.data(), live.data());
    for (size_t k = 0; k < pick.size(); ++k) {
        const int p = at[k];
        if (p < 0) continue;
        ServeTicket& t = res[k];
        if (r != HVC_OK && r != HVC_E_NUMERIC) {
            t.status = r;
            t.err = ctx->err;
            continue;
        }
        const int r0 = ctx->ra_off[p], r1 = ctx->ra_off[p + 1], i0 = ctx->i_off[p], i1 = ctx->i_off[p + 1];
        const int lo[3] = {r0, r0, i0}, hi[3] = {r1, r1, i1};
        hvc_status rs = HVC_OK;
        for (int kind = 0; kind < 3 && rs == HVC_OK; ++kind) {
            int total = 0;
            t.counts[kind].assign(hi[kind] - lo[kind], 0);
            rs = spikes_range(ctx, kind, lo[kind], hi[kind], nullptr, nullptr, 0, &total);
            if (rs != HVC_OK) break;
            t.times[kind].assign(total, 0.0);
            rs = spikes_range(ctx, kind, lo[kind], hi[kind], t.counts[kind].data(), t.times[kind].data(), total,
                              &total);
        }
        if (rs != HVC_OK) {
            t.status = rs;
            t.err = ctx->err;
            continue;
        }
        t.bad_values = op[p].bad_values;
        for (int q = 0; q < 3; ++q) t.warnings[q] = op[p].warnings[q];
        t.status = t.bad_values ? HVC_E_NUMERIC : HVC_OK;
        if (t.bad_values) t.err = "check_bad_values: HVC-RA voltage out of [-200, 100] or NaN";
    }
}

static void serve_loop(hvc_ctx* ctx)
{
    Serving& sv = *ctx->serve;
    const hipError_t dev = hipSetDevice(ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lq(sv.q_mu);
            sv.cv.wait(lq, [&] { return sv.stop || !sv.queue.empty(); });
            if (sv.stop) return;
        }
        // pool calls waiting at the round boundary go first (admissions, releases)
        while (ctx->api_waiters.load() > 0) std::this_thread::yield();
        std::vector<uint64_t> pick;
        std::vector<ServeTicket> res;
        std::unique_lock<std::mutex> la(ctx->api_mu);
        {
            // the oldest queued trial of every network that has one
            std::lock_guard<std::mutex> lq(sv.q_mu);
            std::vector<int> nets;
            for (auto it = sv.queue.begin(); it != sv.queue.end();) {
                ServeTicket& t = sv.tickets[*it];
                if (std::find(nets.begin(), nets.end(), t.handle) == nets.end()) {
                    nets.push_back(t.handle);
                    t.round = sv.rounds;
                    pick.push_back(*it);
                    res.push_back(ServeTicket{});
                    res.back().handle = t.handle;
                    res.back().trial = t.trial;
                    it = sv.queue.erase(it);
                } else {
                    ++it;
                }
            }
            if (!pick.empty()) ++sv.rounds;
        }
        if (pick.empty()) continue;
        if (dev != hipSuccess) {
            for (auto& t : res) {
                t.status = HVC_E_HIP;
                t.err = std::string("hipSetDevice: ") + hipGetErrorString(dev);
            }
        } else {
            serve_round(ctx, pick, res);
        }
        la.unlock();
        std::lock_guard<std::mutex> lq(sv.q_mu);
        for (size_t k = 0; k < pick.size(); ++k) {
            auto it = sv.tickets.find(pick[k]);
            if (it == sv.tickets.end()) continue;
            ServeTicket& t = it->second;
            t.status = res[k].status;
            t.err = std::move(res[k].err);
            t.bad_values = res[k].bad_values;
            for (int q = 0; q < 3; ++q) {
                t.warnings[q] = res[k].warnings[q];
                t.counts[q] = std::move(res[k].counts[q]);
                t.times[q] = std::move(res[k].times[q]);
            }
            t.done = true;
        }
    }
}

extern "C" hvc_status hvc_submit(hvc_ctx* ctx, int handle, const hvc_trial* trial, uint64_t* ticket)
{
    if (!ctx || !trial || !ticket) return HVC_E_ARG;
    if (!ctx->pool || !ctx->serve) return HVC_E_STATE;
    if (trial->mode != 0 && trial->mode != 1) return HVC_E_ARG;
    Serving& sv = *ctx->serve;
    {
        std::lock_guard<std::mutex> lq(sv.q_mu);
        if (sv.stop) return HVC_E_STATE;
        if (!sv.worker.joinable()) {
            try {
                sv.worker = std::thread(serve_loop, ctx);
            } catch (...) {
                return HVC_E_OOM;
            }
        }
        const uint64_t id = sv.next_ticket++;
        ServeTicket& t = sv.tickets[id];
        t.handle = handle;
        t.trial = *trial;
        sv.queue.push_back(id);
        *ticket = id;
    }
    sv.cv.notify_all();
    return HVC_OK;
}

extern "C" hvc_status hvc_poll(hvc_ctx* ctx, uint64_t ticket, hvc_trial_out* out, long long* round)
{
    if (!ctx) return HVC_E_ARG;
    if (!ctx->pool || !ctx->serve) return HVC_E_STATE;
    Serving& sv = *ctx->serve;
    std::lock_guard<std::mutex> lq(sv.q_mu);
    auto it = sv.tickets.find(ticket);
    if (it == sv.tickets.end()) return HVC_E_ARG;
    ServeTicket& t = it->second;
    if (round) *round = t.round;
    if (!t.done) return HVC_PENDING;
    if (!out) return t.status;   // status only: the ticket stays
    if (t.status != HVC_OK && t.status != HVC_E_NUMERIC) {
        const hvc_status st = t.status;
        sv.tickets.erase(it);
        return st;
    }
    {
        int* cnt[3] = {out->soma_counts, out->dend_counts, out->i_counts};
        double* tm[3] = {out->soma_times, out->dend_times, out->i_times};
        const int cap[3] = {out->soma_cap, out->dend_cap, out->i_cap};
        int* tot[3] = {&out->soma_total, &out->dend_total, &out->i_total};
        bool small = false;
        for (int k = 0; k < 3; ++k) {
            *tot[k] = (int)t.times[k].size();
            if (tm[k] && (int)t.times[k].size() > cap[k]) small = true;
        }
        if (small) return HVC_E_ARG;   // the ticket stays; totals give the sizes needed
        for (int k = 0; k < 3; ++k) {
            if (cnt[k] && !t.counts[k].empty()) std::memcpy(cnt[k], t.counts[k].data(), sizeof(int) * t.counts[k].size());
            if (tm[k] && !t.times[k].empty()) std::memcpy(tm[k], t.times[k].data(), sizeof(double) * t.times[k].size());
        }
        out->bad_values = t.bad_values;
        for (int k = 0; k < 3; ++k) out->warnings[k] = t.warnings[k];
    }
    const hvc_status st = t.status;
    sv.tickets.erase(it);
    return st;
}

// ---------------------------------------------------------------------------
// single-neuron traces
// ---------------------------------------------------------------------------
// the trace clock clk[s] = clk[s - 1] + dt (clk[0] = 0) and its decay rows,
// in buffers owned by `bufs`
static hvc_status trace_tables(hvc_ctx* ctx, CallBufs& bufs, int n_steps, double** d_clock, void** d_decay)
{
    std::vector<double> clk(n_steps + 1, 0.0);
    double t = 0;
    for (int s = 1; s <= n_steps; ++s) {
        t += ctx->prm.dt;
        clk[s] = t;
    }
    HVC_TRY(bufs.alloc(d_clock, sizeof(double) * (n_steps + 1)));
    HVC_TRY(bufs.alloc(d_decay, sizeof(DecayRow) * (n_steps + 1)));
    HVC_TRY(hipMemcpyAsync(*d_clock, clk.data(), sizeof(double) * (n_steps + 1), hipMemcpyHostToDevice, ctx->stream));
    launch_decay_table(ctx->stream, *d_clock, ctx->d_offsets, n_steps + 1, *d_decay);
    HVC_TRY(hipGetLastError());
    return HVC_OK;
}

// HVC-RA traces of n neurons over n_steps on device arrays (layouts of
// hvc_trace_ra), with the trace tables at d_clock / d_decay
static hvc_status trace_ra_device(hvc_ctx* ctx, int n, int n_steps, const double* d_el, const double* d_gca,
                                  uint32_t trial, const double* d_kick, const double* d_exc, const double* d_inh,
                                  double* d_out, uint8_t* d_fired, const double* d_clock, const void* d_decay)
{
    NetDev nd = ctx->nd;
    nd.clock = d_clock;
    nd.decay = d_decay;
    if (ctx->prm.precision == 1)
        launch_trace_ra_f32(ctx->stream, nd, n, n_steps, trial, d_el, d_gca, d_kick, d_exc, d_inh, d_out, d_fired);
    else
        launch_trace_ra(ctx->stream, nd, n, n_steps, trial, d_el, d_gca, d_kick, d_exc, d_inh, d_out, d_fired);
    HVC_TRY(hipGetLastError());
    return HVC_OK;
}

extern "C" hvc_status hvc_trace_ra(hvc_ctx* ctx, int n, int n_steps, const double* erest, const double* gca,
                                   uint32_t trial, const double* kick, const double* exc_w, const double* inh_g,
                                   double* out_state, uint8_t* out_fired)
{
    if (!ctx || n <= 0 || n_steps <= 0 || !erest || !gca || !kick || !exc_w || !inh_g || !out_state || !out_fired)
        return HVC_E_ARG;
    HVC_TRY(hipSetDevice(ctx->device));
    CallBufs bufs;
    double* d_clock = nullptr;
    void* d_decay = nullptr;
    hvc_status r = trace_tables(ctx, bufs, n_steps, &d_clock, &d_decay);
    if (r != HVC_OK) return r;
    size_t in_sz = (size_t)n_steps * n;
    double *d_el, *d_gca, *d_kick, *d_exc, *d_inh, *d_out;
    uint8_t* d_fired;
    HVC_TRY(bufs.alloc(&d_el, sizeof(double) * n));
    HVC_TRY(bufs.alloc(&d_gca, sizeof(double) * n));
    HVC_TRY(bufs.alloc(&d_kick, sizeof(double) * in_sz));
    HVC_TRY(bufs.alloc(&d_exc, sizeof(double) * in_sz));
    HVC_TRY(bufs.alloc(&d_inh, sizeof(double) * in_sz));
    HVC_TRY(bufs.alloc(&d_out, sizeof(double) * (in_sz + n) * 12));
    HVC_TRY(bufs.alloc(&d_fired, in_sz));
    hipStream_t s = ctx->stream;
    HVC_TRY(hipMemcpyAsync(d_el, erest, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(d_gca, gca, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(d_kick, kick, sizeof(double) * in_sz, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(d_exc, exc_w, sizeof(double) * in_sz, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(d_inh, inh_g, sizeof(double) * in_sz, hipMemcpyHostToDevice, s));
    r = trace_ra_device(ctx, n, n_steps, d_el, d_gca, trial, d_kick, d_exc, d_inh, d_out, d_fired, d_clock, d_decay);
    if (r != HVC_OK) return r;
    HVC_TRY(hipMemcpyAsync(out_state, d_out, sizeof(double) * (in_sz + n) * 12, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(out_fired, d_fired, in_sz, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipStreamSynchronize(s));
    return HVC_OK;
}

// true when p is device memory of this context's device (or managed memory)
static bool on_device(const hvc_ctx* ctx, const void* p)
{
    hipPointerAttribute_t a;
    std::memset(&a, 0, sizeof a);
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (a.type == hipMemoryTypeDevice) return a.device == ctx->device;
    return a.type == hipMemoryTypeManaged;
}

extern "C" hvc_status hvc_trace_ra_noalloc(hvc_ctx* ctx, int n, int n_steps, const double* d_erest,
                                           const double* d_gca, uint32_t trial, const double* d_kick,
                                           const double* d_exc_w, const double* d_inh_g, double* d_out_state,
                                           uint8_t* d_out_fired)
{
    if (!ctx) return HVC_E_ARG;
    if (n <= 0 || n_steps <= 0) return fail(ctx, HVC_E_ARG, "n and n_steps must be > 0");
    if (n_steps > kNoAllocMaxSteps)
        return fail(ctx, HVC_E_ARG, "n_steps " + std::to_string(n_steps) + " exceeds the fixed trace buffer's " +
                                        std::to_string(kNoAllocMaxSteps) + " steps");
    const void* arrays[] = {d_erest, d_gca, d_kick, d_exc_w, d_inh_g, d_out_state, d_out_fired};
    for (const void* q : arrays)
        if (!q || !on_device(ctx, q))
            return fail(ctx, HVC_E_ARG, "hvc_trace_ra_noalloc takes device arrays of the context's device");
    HVC_TRY(hipSetDevice(ctx->device));
    unsigned char* base = (unsigned char*)noalloc_scratch(ctx->device);
    if (!base) return fail(ctx, HVC_E_STATE, "fixed trace buffer not set up on this device");
    double* d_clock = (double*)base;
    void* d_decay = base + kNoAllocClockBytes;
    launch_clock_table(ctx->stream, d_clock, n_steps + 1, ctx->prm.dt);
    launch_decay_table(ctx->stream, d_clock, ctx->d_offsets, n_steps + 1, d_decay);
    HVC_TRY(hipGetLastError());
    hvc_status r = trace_ra_device(ctx, n, n_steps, d_erest, d_gca, trial, d_kick, d_exc_w, d_inh_g, d_out_state,
                                   d_out_fired, d_clock, d_decay);
    if (r != HVC_OK) return r;
    HVC_TRY(hipStreamSynchronize(ctx->stream));
    return HVC_OK;
}

extern "C" hvc_status hvc_trace_i(hvc_ctx* ctx, int n, int n_steps, int first_trial, uint32_t trial,
                                  const double* exc_g, const double* inh_g, double* out_state, uint8_t* out_fired)
{
    if (!ctx || n <= 0 || n_steps <= 0 || !exc_g || !inh_g || !out_state || !out_fired) return HVC_E_ARG;
    HVC_TRY(hipSetDevice(ctx->device));
    CallBufs bufs;
    double* d_clock = nullptr;
    void* d_decay = nullptr;
    hvc_status r = trace_tables(ctx, bufs, n_steps, &d_clock, &d_decay);
    if (r != HVC_OK) return r;
    size_t in_sz = (size_t)n_steps * n;
    double *d_exc, *d_inh, *d_out;
    uint8_t* d_fired;
    HVC_TRY(bufs.alloc(&d_exc, sizeof(double) * in_sz));
    HVC_TRY(bufs.alloc(&d_inh, sizeof(double) * in_sz));
    HVC_TRY(bufs.alloc(&d_out, sizeof(double) * (in_sz + n) * 7));
    HVC_TRY(bufs.alloc(&d_fired, in_sz));
    hipStream_t s = ctx->stream;
    HVC_TRY(hipMemcpyAsync(d_exc, exc_g, sizeof(double) * in_sz, hipMemcpyHostToDevice, s));
    HVC_TRY(hipMemcpyAsync(d_inh, inh_g, sizeof(double) * in_sz, hipMemcpyHostToDevice, s));
    NetDev nd = ctx->nd;
    nd.clock = d_clock;
    nd.decay = d_decay;
    if (ctx->prm.precision == 1)
        launch_trace_i_f32(s, nd, n, n_steps, first_trial, trial, d_exc, d_inh, d_out, d_fired);
    else
        launch_trace_i(s, nd, n, n_steps, first_trial, trial, d_exc, d_inh, d_out, d_fired);
    HVC_TRY(hipGetLastError());
    HVC_TRY(hipMemcpyAsync(out_state, d_out, sizeof(double) * (in_sz + n) * 7, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipMemcpyAsync(out_fired, d_fired, in_sz, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipStreamSynchronize(s));
    return HVC_OK;
}

extern "C" hvc_status hvc_eval_math(hvc_ctx* ctx, int fn, int n, const double* a, const double* b, double* out)
{
    if (!ctx || n < 0 || fn < 0 || fn > 2 || (n > 0 && (!a || !out || (fn == 2 && !b)))) return HVC_E_ARG;
    if (n == 0) return HVC_OK;
    HVC_TRY(hipSetDevice(ctx->device));
    CallBufs bufs;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    HVC_TRY(bufs.alloc(&da, sizeof(double) * n));
    HVC_TRY(bufs.alloc(&db, sizeof(double) * n));
    HVC_TRY(bufs.alloc(&dout, sizeof(double) * n));
    hipStream_t s = ctx->stream;
    HVC_TRY(hipMemcpyAsync(da, a, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (fn == 2) HVC_TRY(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    launch_eval_math(s, fn, n, da, fn == 2 ? db : da, dout);
    HVC_TRY(hipGetLastError());
    HVC_TRY(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HVC_TRY(hipStreamSynchronize(s));
    return HVC_OK;
}
