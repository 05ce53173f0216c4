// hvc_kernels.hip -- gfx950 kernels of the HVC growth-trial engine (FP64 form).
//
// One trial (NGS:8839-9571 growth, NGS:12747-13118 chain test) runs as a
// sequence of coupling windows: the steps between two synchronisation points
// (NGS:9291).  Neurons only interact at those points, so each window is
//   integrate_ra / integrate_i : every neuron advances through the window in
//                                registers; spikes become per-window bit masks
//                                and a spike log (wave ballot + one atomic).
//   deliver                    : per presynaptic neuron, delay-queue pops whose
//                                arrival falls in the window become records in
//                                the target's inbox; RA->RA rows also run the
//                                order-dependent LTD / update_synapse scan.
//   stdp                       : at the window's sync point, LTP / LTD against
//                                the window's post events and axon remodeling
//                                (NGS:9375-9554), in blocks of the next
//                                window's k_integrate (hvc_stdp.hip.h).
//   sync                       : per target, records are ordered by (arrival
//                                step, presynaptic id) -- the reference's
//                                accumulation order -- summed and applied
//                                (raiseI / raiseE / raiseExcWeight, NGS:9314-9366).
// No floating-point atomics: every sum is evaluated in the reference's order,
// so the FP64 form reproduces the CPU results bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "hvc_internal.h"
#include "hvc_integrate.hip.h"
#include "hvc_stdp.hip.h"

namespace hvc {

// ---------------------------------------------------------------------------
// decay table: row t holds exp(-((c + off) - c)/tau) for c = clock[t-1]
// (HH2_test.cpp:861-864, HHI_buffer.cpp:739-740)
// ---------------------------------------------------------------------------
__global__ void k_decay_table(const double* clock, const double* offsets, int num_steps, DecayRow* rows)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_steps) return;
    double c = (t == 0) ? 0.0 : clock[t - 1];
    DecayRow r;
    for (int o = 0; o < kNumOffsets; ++o) {
        double T = c + offsets[o];
        r.e5[o] = dexp(-(T - c) / 5.0);
        r.e2[o] = dexp(-(T - c) / 2.0);
    }
    rows[t] = r;
}

// ---------------------------------------------------------------------------
// trial start: setToRest_afterEpoch (NGS:4591-4621), HH2_test::set_to_rest
// (HH2_test.cpp:191-242), HHI_buffer::set_to_rest (HHI_buffer.cpp:178-200)
// ---------------------------------------------------------------------------
// no spike in flight: count 0, every slot free
__device__ __forceinline__ void fl_clear(const InFlight& fl, int i)
{
    fl.head[i].count = 0;
    fl.held[i] = 0;
    for (int k = 0; k < kHeadSlots; ++k) fl.head[i].s[k].step = -1;
    FlSlot* m = fl.more + (size_t)i * (kMaxInflight - kHeadSlots);
    for (int k = 0; k < kMaxInflight - kHeadSlots; ++k) m[k].step = -1;
}

__global__ void k_reset(NetDev nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd.n_ra) {
        nd.ra[0][i] = -79.97619025; nd.ra[1][i] = -79.97268759; nd.ra[2][i] = 0.01101284; nd.ra[3][i] = 0.0;
        nd.ra[4][i] = 0.9932845; nd.ra[5][i] = 0.00055429; nd.ra[6][i] = 0.00000261762353; nd.ra[7][i] = 0.01689572;
        nd.ra[8][i] = 0.0; nd.ra[9][i] = 0.0; nd.ra[10][i] = 0.0; nd.ra[11][i] = 0.0;
        nd.ra_flags[i] = 0;
        nd.ra_draw[i] = 0;
        if (nd.ra_poisson) {
            const NoiseId sid = ra_noise_id(nd, i);
            PoissonStream ps{sid.k0, sid.k1, sid.trial, kStreamRAPoisson, sid.local, 0u};
            for (int q = 0; q < 4; ++q) nd.ra_noise[(size_t)q * nd.n_ra + i] = 1000.0 * ps.spike_time(100.0);
            nd.ra_draw[i] = ps.draw;
        }
        nd.post_cnt[i] = 0;
        nd.post_pend[i] = 0;
        nd.last_pre[i] = -1.0e300;
        nd.st_count[i] = 0;
        fl_clear(nd.fl_ra2i, i);
        fl_clear(nd.fl_rr, i);
        nd.in_gi_ra.cnt[i] = 0;
        nd.in_ge_ra.cnt[i] = 0;
        nd.wl_mark[0][i] = -1;
        nd.wl_mark[1][i] = -1;
        nd.soma_mask[i] = 0;   // (a network sitting the trial out never writes its masks)
        nd.dend_mask[i] = 0;
    }
    if (i < nd.n_i) {
        nd.in[0][i] = -66.0; nd.in[1][i] = 0.125; nd.in[2][i] = 0.0; nd.in[3][i] = 0.99; nd.in[4][i] = 0.0;
        nd.in[5][i] = 0.0; nd.in[6][i] = 0.0;
        nd.i_flag[i] = 0;
        const NoiseId sid = i_noise_id(nd, i);
        PoissonStream ps{sid.k0, sid.k1, sid.trial, kStreamIPoisson, sid.local, 0u};
        if (nd.i_poisson) {
            double init_dt = nd.net_first[nd.i_net[i]] ? 0.0 : nd.dt;
            nd.i_noise[i] = i_init_noise(ps, init_dt);
            nd.i_noise[(size_t)nd.n_i + i] = i_init_noise(ps, init_dt);
        }
        nd.i_draw[i] = ps.draw;
        fl_clear(nd.fl_i2ra, i);
        nd.in_ge_i.cnt[i] = 0;
        nd.wl_mark[2][i] = -1;
        nd.i_mask[i] = 0;
    }
    if (i == 0) {
        nd.log_cnt[0] = 0; nd.log_cnt[1] = 0; nd.log_cnt[2] = 0;
        nd.post_list_cnt[0] = 0;
        nd.post_list_cnt[1] = 0;
        *nd.err = 0;
    }
    if (i < 3 * nd.n_nets) nd.warn[i] = 0;
    if (i < kWlCounters) nd.wl_cnt[i] = 0;
}

// ---------------------------------------------------------------------------
// Poisson noise events of a chunk of steps [tb, te].  The noise checks of
// HH2_buffer (HH2_buffer.cpp:592-607) and HHI_buffer (HHI_buffer.cpp:278-294)
// fire at the first step whose neuron time clock[t - 1] exceeds the channel's
// noise time, draw from one stream per neuron in the order of the checks, and
// never look at the neuron's state; so one lane per neuron replays them here,
// channel by channel in step order, with the same draws, and the integration
// only adds the listed increments.  The stream position and noise times carry
// over to the next chunk.
// ---------------------------------------------------------------------------

// first step t in [tmin, te] with clock[t - 1] > nt; te + 1 if there is none
__device__ int noise_step(const double* clock, double dt, int tmin, int te, double nt)
{
    if (tmin > te) return te + 1;
    const double c0 = clock[tmin - 1];
    if (c0 > nt) return tmin;
    const double steps = (nt - c0) / dt;        // the clock advances by dt per step
    if (steps > (double)(te - tmin + 2)) return te + 1;
    int t = tmin + (int)steps;
    if (t > te + 1) t = te + 1;
    while (t > tmin && clock[t - 2] > nt) --t;  // now t - 1 does not fire (or t = tmin)
    while (t <= te && !(clock[t - 1] > nt)) ++t;
    return t;
}

struct NoiseOut {
    NoiseEvent* ev;
    int n, cap;
    unsigned* err;
    __device__ __forceinline__ void put(int step, int target, double inc)
    {
        if (n < cap - 2)
            ev[n++] = NoiseEvent{step, target, inc};
        else
            atomicOr(err, (unsigned)kErrNoise);
    }
    __device__ __forceinline__ void close()
    {
        ev[n] = NoiseEvent{kNoNoiseStep, 0, 0.0};
        ev[n + 1] = NoiseEvent{kNoNoiseStep, 0, 0.0};
    }
};

// one noise check that fired at step t: G += random(G_noise), then the next
// interval, redrawn (with another increment each time) while shorter than dt
__device__ __forceinline__ double noise_fire(PoissonStream& ps, NoiseOut& o, int t, int target, double G_noise,
                                             double lambda, double dt, double noise_time)
{
    o.put(t, target, ps.random(G_noise));
    double random = 1000.0 * ps.spike_time(lambda);
    while (random < dt) {
        random = 1000.0 * ps.spike_time(lambda);
        o.put(t, target, ps.random(G_noise));
    }
    return noise_time + random;
}

// blocks [0, blocks_ra): HVC-RA (RK4 form, channels Gi_s, Gi_d, Ge_s, Ge_d with
// G_noise 0.035, 0.045, 0.035, 0.045 at 100 Hz); the rest HVC-I (Ge, Gi with
// 0.45 at 250 Hz)
__global__ __launch_bounds__(64) void k_noise_events(NetDev nd, int tb, int te, int blocks_ra, int gen_ra, int gen_i)
{
    const double dt = nd.dt;
    if ((int)blockIdx.x < blocks_ra) {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= nd.n_ra) return;
        NoiseOut o{nd.ra_ev + (size_t)i * nd.ev_cap, 0, nd.ev_cap, nd.err};
        nd.ra_ev_pos[i] = 0;
        if (gen_ra) {
            const NoiseId sid = ra_noise_id(nd, i);
            PoissonStream ps{sid.k0, sid.k1, sid.trial, kStreamRAPoisson, sid.local, nd.ra_draw[i]};
            const size_t n = (size_t)nd.n_ra;
            double nt0 = nd.ra_noise[n + i], nt1 = nd.ra_noise[3 * n + i];    // inh soma, inh dend
            double nt2 = nd.ra_noise[i], nt3 = nd.ra_noise[2 * n + i];        // exc soma, exc dend
            int s0 = noise_step(nd.clock, dt, tb, te, nt0), s1 = noise_step(nd.clock, dt, tb, te, nt1);
            int s2 = noise_step(nd.clock, dt, tb, te, nt2), s3 = noise_step(nd.clock, dt, tb, te, nt3);
            for (;;) {
                // earliest check; at one step the channels fire in the order of the checks
                int c = 0, t = s0;
                if (s1 < t) { c = 1; t = s1; }
                if (s2 < t) { c = 2; t = s2; }
                if (s3 < t) { c = 3; t = s3; }
                if (t > te) break;
                const double nt = c == 0 ? nt0 : (c == 1 ? nt1 : (c == 2 ? nt2 : nt3));
                const double nn = noise_fire(ps, o, t, c, (c & 1) ? 0.045 : 0.035, 100.0, dt, nt);
                const int sn = noise_step(nd.clock, dt, t + 1, te, nn);
                if (c == 0) { nt0 = nn; s0 = sn; }
                else if (c == 1) { nt1 = nn; s1 = sn; }
                else if (c == 2) { nt2 = nn; s2 = sn; }
                else { nt3 = nn; s3 = sn; }
            }
            nd.ra_noise[n + i] = nt0; nd.ra_noise[3 * n + i] = nt1;
            nd.ra_noise[i] = nt2; nd.ra_noise[2 * n + i] = nt3;
            nd.ra_draw[i] = ps.draw;
        }
        o.close();
    } else {
        const int i = (blockIdx.x - blocks_ra) * blockDim.x + threadIdx.x;
        if (i >= nd.n_i) return;
        NoiseOut o{nd.i_ev + (size_t)i * nd.ev_cap, 0, nd.ev_cap, nd.err};
        nd.i_ev_pos[i] = 0;
        if (gen_i) {
            const NoiseId sid = i_noise_id(nd, i);
            PoissonStream ps{sid.k0, sid.k1, sid.trial, kStreamIPoisson, sid.local, nd.i_draw[i]};
            const size_t n = (size_t)nd.n_i;
            double nt0 = nd.i_noise[i], nt1 = nd.i_noise[n + i];               // exc, inh
            int s0 = noise_step(nd.clock, dt, tb, te, nt0), s1 = noise_step(nd.clock, dt, tb, te, nt1);
            for (;;) {
                int c = 0, t = s0;
                if (s1 < t) { c = 1; t = s1; }
                if (t > te) break;
                const double nn = noise_fire(ps, o, t, c, 0.45, 250.0, dt, c == 0 ? nt0 : nt1);
                const int sn = noise_step(nd.clock, dt, t + 1, te, nn);
                if (c == 0) { nt0 = nn; s0 = sn; }
                else { nt1 = nn; s1 = sn; }
            }
            nd.i_noise[i] = nt0; nd.i_noise[n + i] = nt1;
            nd.i_draw[i] = ps.draw;
        }
        o.close();
    }
}

__global__ void k_clear_pre(int* pre_cnt, int E)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) pre_cnt[e] = 0;
}

// ---------------------------------------------------------------------------
// delivery helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void inbox_put(const Inbox& ib, int tgt, int step, int pre, double w, unsigned int* err)
{
    int pos = atomicAdd(&ib.cnt[tgt], 1);
    if (pos < ib.cap) {
        Record r;
        r.key = ((uint64_t)(uint32_t)step << 32) | (uint32_t)pre;
        r.w = w;
        ib.rec[(size_t)tgt * ib.cap + pos] = r;
    } else {
        atomicOr(err, (unsigned)kErrInbox);
    }
}

// The window's clock in LDS: wclk[t - t0 + 1] = clock[t] for t0 - 1 <= t <= t1.
// First step a >= lo (lo >= t0) with clock[a] >= T (the drain test
// internal_time >= delivery_time); T <= clock[t1], so a <= t1.
__device__ __forceinline__ int arrival_step(const double* wclk, int t0, int lo, double T)
{
    int a = lo;
    while (wclk[a - t0 + 1] < T) ++a;
    return a;
}

// value of lane K of the quad, in every lane of the quad (DPP quad_perm)
template <int K>
__device__ __forceinline__ int quad_lane(int x)
{
    return __builtin_amdgcn_mov_dpp(x, K * 0x55, 0xF, 0xF, true);
}
template <int G>
__device__ __forceinline__ int group_min(int x)
{
    if constexpr (G == 1) {
        return x;
    } else {
        const int a = min(quad_lane<0>(x), quad_lane<1>(x)), b = min(quad_lane<2>(x), quad_lane<3>(x));
        return min(a, b);
    }
}
template <int G>
__device__ __forceinline__ double group_min(double x)
{
    if constexpr (G == 1) {
        return x;
    } else {
        auto lane = [](double v, auto k) {
            const long long b = __double_as_longlong(v);
            constexpr int K = decltype(k)::value;
            const int lo = quad_lane<K>((int)b), hi = quad_lane<K>((int)(b >> 32));
            return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
        };
        const double a = fmin(lane(x, std::integral_constant<int, 0>{}), lane(x, std::integral_constant<int, 1>{}));
        const double c = fmin(lane(x, std::integral_constant<int, 2>{}), lane(x, std::integral_constant<int, 3>{}));
        return fmin(a, c);
    }
}

// Fixed-weight delay queues of one presynaptic neuron (RA->I NGS:8971-8986 +
// 9039-9062, I->RA NGS:9235-9284, chain-test RA->RA NGS:13222-13250), G lanes
// per neuron (G = 1 or a quad).  Each in-flight spike holds one of the
// neuron's kMaxInflight slots (FlSlot: spike step, edge cursor, delivery time
// of the cursor's edge, clock at the spike step; free slots have step -1).
// The held count and the first slots share one line, read in one go, so a
// window's work starts one memory round trip after the kernel does.  The
// row's edges are sorted by delay, so the edges due by t1 are a prefix from
// the cursor; their pops become inbox records.  The sync sums a window's
// records in the reference's order whatever order they are written in, so
// lane r of the group takes edges cursor + r, cursor + r + G, ... (the records'
// atomics of a spike overlap) and the group agrees on the new cursor, the
// first edge not due.  New spikes take the free slots in slot order.  Returns
// the number of spikes still in flight after the window (in every lane).
template <int G>
__device__ __forceinline__ int push_fixed(const NetDev& nd, const double* wclk, int pre, int r, int row_begin,
                                           int row_end, const int* edge_ids, const int* tgt_arr, const double* w_arr,
                                           const double* d_arr, const InFlight& fl, FlHead h,
                                           uint64_t new_mask, int t0, int t1, const Inbox& ib)
{
    const int deg = row_end - row_begin;
    FlHead* hp = fl.head + pre;
    const int held = h.count;
    if (deg == 0 || (held == 0 && new_mask == 0)) return held;
    FlSlot* more = fl.more + (size_t)pre * (kMaxInflight - kHeadSlots);
    auto slot = [&](int k) -> FlSlot* { return k < kHeadSlots ? &hp->s[k] : more + (k - kHeadSlots); };
    const double tend = wclk[t1 - t0 + 1];
    // the held slots and which of them have an edge due in this window
    unsigned held_mask = 0, due_mask = 0;
#pragma unroll
    for (int k = 0; k < kHeadSlots; ++k) {
        if (h.s[k].step >= 0) {
            held_mask |= 1u << k;
            if (h.s[k].tnext <= tend) due_mask |= 1u << k;
        }
    }
    if (held > __popc(held_mask)) {
        for (int k = kHeadSlots; k < kMaxInflight; ++k) {
            const FlSlot sl = more[k - kHeadSlots];
            if (sl.step >= 0) {
                held_mask |= 1u << k;
                if (sl.tnext <= tend) due_mask |= 1u << k;
            }
        }
    }
    // the first nnew free slots take the window's new spikes (in step order)
    const int nnew = __popcll(new_mask);
    unsigned free_mask = ~held_mask & ((1u << kMaxInflight) - 1u), assigned = 0;
    for (int j = 0; j < nnew && free_mask; ++j) {
        const unsigned low = free_mask & (0u - free_mask);
        assigned |= low;
        free_mask &= ~low;
    }
    if (r == 0 && nnew > __popc(assigned)) atomicOr(nd.err, (unsigned)kErrInflight);
    int now_held = __popc(held_mask & ~due_mask);   // held slots with nothing due stay as they are
    for (unsigned m = due_mask | assigned; m; m &= m - 1) {
        const int k = __ffs(m) - 1;
        const bool fresh = (assigned >> k) & 1u;
        int s, c;
        double cs;   // clock at the spike step (a fresh spike is in this window: from the LDS clock)
        if (fresh) {
            uint64_t nm = new_mask;
            for (int j = __popc(assigned & ((1u << k) - 1u)); j > 0; --j) nm &= nm - 1;
            s = t0 + __ffsll((unsigned long long)nm) - 1;
            c = 0;
            cs = wclk[s - t0 + 1];
        } else {
            FlSlot sl;
            if (k < kHeadSlots) {   // from the registers (no second round trip)
                sl = h.s[0];
#pragma unroll
                for (int j = 1; j < kHeadSlots; ++j)
                    if (k == j) sl = h.s[j];
            } else {
                sl = more[k - kHeadSlots];
            }
            s = sl.step;
            c = sl.cur;
            cs = sl.tspike;
        }
        int lo = s > t0 ? s : t0;
        int cc = c + r, stop = deg;
        double tstop = HUGE_VAL;   // delivery time of this lane's first edge not due
        // an edge's target, weight and delay are loaded together, the next
        // edge's while this edge's record is written
        int e = cc < deg ? (edge_ids ? edge_ids[row_begin + cc] : row_begin + cc) : 0;
        double d = 0.0, w = 0.0;
        int tg = 0;
        if (cc < deg) {
            d = d_arr[e];
            tg = tgt_arr[e];
            w = w_arr[e];
        }
        while (cc < deg) {
            const double T = cs + d;
            if (T > tend) {
                stop = cc;
                tstop = T;
                break;
            }
            const int cn = cc + G;
            const int en = cn < deg ? (edge_ids ? edge_ids[row_begin + cn] : row_begin + cn) : e;
            const double dn = d_arr[en];
            const int tgn = tgt_arr[en];
            const double wn = w_arr[en];
            const int a = arrival_step(wclk, t0, lo, T);
            lo = a;
            inbox_put(ib, tg, a, pre, w, nd.err);
            cc = cn;
            e = en;
            d = dn;
            tg = tgn;
            w = wn;
        }
        // the first edge not due over the group (deg: the spike is done) and its
        // delivery time (delays are sorted, so it is the group's smallest)
        const int cnew = group_min<G>(stop);
        const double tnew = group_min<G>(tstop);
        if (cnew < deg) ++now_held;
        if (r == 0) {
            FlSlot* sp = slot(k);
            if (cnew >= deg) {
                if (!fresh) sp->step = -1;   // (a fresh slot is still free)
            } else {
                *sp = FlSlot{s, cnew, tnew, cs};
            }
        }
    }
    if (r == 0) {
        hp->count = now_held;
        fl.held[pre] = now_held;
    }
    return now_held;
}

// next item of stream sidx of row i from the row's stream state in memory:
// (T, target, edge); returns false when exhausted
__device__ __forceinline__ bool stream_peek(const NetDev& nd, int i, int sidx, double& T, int& tgt, int& e)
{
    int* cur = nd.st_cursor + (size_t)i * kMaxStreams + sidx;
    int n = nd.st_n[(size_t)i * kMaxStreams + sidx];
    int mode = nd.st_mode[(size_t)i * kMaxStreams + sidx];
    int s = nd.st_step[(size_t)i * kMaxStreams + sidx];
    const double cs = nd.clock[s];
    int c = *cur;
    if (mode == 0) {
        const int base = nd.rr_ptr[i];
        while (c < n) {
            const RrPos p = nd.rr_pos[base + c];
            if (p.tgt != i) {
                e = p.e;
                tgt = p.tgt;
                T = cs + p.d;
                *cur = c;
                return true;
            }
            ++c;
        }
        *cur = c;
        return false;
    }
    if (c >= n) return false;
    e = nd.st_items[((size_t)i * kMaxStreams + sidx) * nd.syn.Nss + c];
    tgt = nd.rr_tgt[e];
    T = cs + nd.rr_d[e];
    return true;
}

constexpr int kQuadPushNeurons = 65536;   // neurons per context up to which the pushes take a quad per neuron
constexpr int kDeliverRoleWaves = 256;    // waves per role of a delivery launch, at most

// A row's first kRowHeads streams while a window drains them, one column of the
// block's LDS per thread: the head item (T, target, edge) and the stream's
// cursor, length, mode and spike clock.  The cursors go back to memory when
// the row leaves the cached path.
constexpr int kRowHeads = 8;
constexpr int kHeadDoubles = 2;   // T, spike clock
constexpr int kHeadInts = 5;      // target, edge, cursor, length, mode
struct StreamHeads {
    double* t_;
    int* i_;
    int stride;
    __device__ __forceinline__ double T(int q) const { return t_[q * stride]; }
    __device__ __forceinline__ double cs(int q) const { return t_[(kRowHeads + q) * stride]; }
    __device__ __forceinline__ int tg(int q) const { return i_[q * stride]; }
    __device__ __forceinline__ int e(int q) const { return i_[(kRowHeads + q) * stride]; }
    __device__ __forceinline__ int cur(int q) const { return i_[(2 * kRowHeads + q) * stride]; }
    __device__ __forceinline__ int n(int q) const { return i_[(3 * kRowHeads + q) * stride]; }
    __device__ __forceinline__ int mode(int q) const { return i_[(4 * kRowHeads + q) * stride]; }
    __device__ __forceinline__ void set(int q, double T, int tg, int e)
    {
        t_[q * stride] = T;
        i_[q * stride] = tg;
        i_[(kRowHeads + q) * stride] = e;
    }
    __device__ __forceinline__ void set_cur(int q, int c) const { i_[(2 * kRowHeads + q) * stride] = c; }
    __device__ __forceinline__ void open(int q, double cs, int c, int n, int mode) const
    {
        t_[(kRowHeads + q) * stride] = cs;
        i_[(2 * kRowHeads + q) * stride] = c;
        i_[(3 * kRowHeads + q) * stride] = n;
        i_[(4 * kRowHeads + q) * stride] = mode;
    }
};

// Head of cached stream q at cursor c (c >= the stream's cursor): the first
// item at or after c, self-edges skipped (mode 0).  The edge record p of
// position c may already be loaded (have_p).
__device__ __forceinline__ bool head_find(const NetDev& nd, int i, int base, int q, const StreamHeads& hd, int c,
                                          RrPos p, bool have_p, double& T, int& tgt, int& e)
{
    const int n = hd.n(q);
    const double cs = hd.cs(q);
    if (hd.mode(q) == 0) {
        while (c < n) {
            if (!have_p) p = nd.rr_pos[base + c];
            have_p = false;
            if (p.tgt != i) {
                e = p.e;
                tgt = p.tgt;
                T = cs + p.d;
                hd.set_cur(q, c);
                return true;
            }
            ++c;
        }
        hd.set_cur(q, c);
        return false;
    }
    hd.set_cur(q, c);
    if (c >= n) return false;
    e = have_p ? p.e : nd.st_items[((size_t)i * kMaxStreams + q) * nd.syn.Nss + c];
    tgt = nd.rr_tgt[e];
    T = cs + nd.rr_d[e];
    return true;
}

// the cached cursors back to the row's stream state in memory
__device__ __forceinline__ void heads_store(const NetDev& nd, int i, int sc, const StreamHeads& hd)
{
    int* st_cur = nd.st_cursor + (size_t)i * kMaxStreams;
    for (int q = 0; q < sc; ++q) st_cur[q] = hd.cur(q);
}

// What one drained item reads of its edge, loaded before any of its writes:
// the edge's state and pre-event count, and the target's post-event count (the
// delivery writes no post event).
struct ItemIn {
    int pc, post_n;
    uint8_t act, sup;
    double w;
};
__device__ __forceinline__ ItemIn item_load(const NetDev& nd, int e, int tgt)
{
    ItemIn in;
    in.act = nd.act[e];
    in.sup = nd.sup[e];
    in.w = nd.w[e];
    in.pc = nd.pre_cnt[e];
    in.post_n = nd.post_cnt[tgt];
    return in;
}

// the row's synapse counts and remodeled flag, kept in registers while the row
// drains (only update_synapse changes them here)
struct RowState {
    int scnt, acnt, rem;
    __device__ __forceinline__ void load(const NetDev& nd, int i)
    {
        scnt = nd.sup_cnt[i];
        acnt = nd.act_cnt[i];
        rem = nd.remodeled[i];
    }
};

// update_synapse (NGS:16546-16592) of edge e to target j with the edge's and
// the row's state in registers; the lists change in memory, in the same order
struct SynCache {
    int act, sup, rem, scnt, acnt;
};
__device__ __forceinline__ void update_synapse_c(const NetDev& nd, int i, int j, double w, SynCache& c)
{
    const hvc_synaptic_params& p = nd.syn;
    int* alist = nd.act_list + nd.rr_ptr[i];
    int* slist = nd.sup_list + (size_t)i * p.Nss;
    if ((w >= p.ACTIVATION_THRESHOLD) && (c.act == 0) && (c.rem == 0)) {
        c.act = 1;
        alist[c.acnt++] = j;
        count_hit(nd, kHitActivate);
    }
    if ((w >= p.SUPERSYNAPSE_THRESHOLD) && (c.sup == 0) && (c.scnt < p.Nss)) {
        c.sup = 1;
        slist[c.scnt++] = j;
        count_hit(nd, kHitSuper);
    }
    if ((w < p.SUPERSYNAPSE_THRESHOLD) && (c.sup == 1)) {
        c.sup = 0;
        c.rem = 0;
        if (!list_erase(slist, &c.scnt, j)) atomicAdd(&row_warn(nd, i)[2], 1);
        count_hit(nd, kHitUnsuper);
    }
    if ((w < p.ACTIVATION_THRESHOLD) && (c.act == 1)) {
        c.act = 0;
        if (!list_erase(alist, &c.acnt, j)) atomicAdd(&row_warn(nd, i)[2], 1);
        count_hit(nd, kHitDeactivate);
    }
}

// LTD of edge e against the target's post events not older than STDP_WINDOW
// (NGS:9111-9180), each step followed by update_synapse.  The ring's next
// entry is loaded a step ahead; the edge's and the row's synapse state stay in
// registers through the steps and are written back once (rs follows them).
__device__ __forceinline__ void ltd_on_delivery(const NetDev& nd, int i, int e, int tgt, double T, double now,
                                                const ItemIn& in, RowState& rs, int hit)
{
    const int pc = in.post_n;
    if (pc == 0) return;
    const int first = pc > kPostRing ? pc - kPostRing : 0;
    const double* ring = nd.post_ring + (size_t)tgt * kPostRing;
    const double lim = now - 100.0;
    SynCache c{in.act, in.sup, rs.rem, rs.scnt, rs.acnt};
    double w = in.w;
    bool on = false;   // from the first event not older than the window on
    double vn = ring[first % kPostRing];   // the next event, loaded one step ahead
#pragma unroll 1
    for (int j = 0; j < pc - first; ++j) {
        const double v = vn;
        if (j + 1 < pc - first) vn = ring[(first + j + 1) % kPostRing];
        on = on || !(v < lim);
        if (on) {
            count_hit(nd, hit);
            w = ltd(nd.syn, w, v - T, row_warn(nd, i));
            update_synapse_c(nd, i, tgt, w, c);
        }
    }
    if (!on) return;
    nd.w[e] = w;
    if (c.act != in.act) nd.act[e] = (uint8_t)c.act;
    if (c.sup != in.sup) nd.sup[e] = (uint8_t)c.sup;
    if (c.rem != rs.rem) nd.remodeled[i] = c.rem;
    if (c.scnt != rs.scnt) nd.sup_cnt[i] = c.scnt;
    if (c.acnt != rs.acnt) nd.act_cnt[i] = c.acnt;
    rs.rem = c.rem;
    rs.scnt = c.scnt;
    rs.acnt = c.acnt;
}

// one drained item of row i: edge e to tgt, delivery time T, at step t
// (NGS:9080-9191); lastp is the row's latest pre-event time
__device__ __forceinline__ void rr_item(const NetDev& nd, int i, int e, int tgt, double T, int t, double now, int Nss,
                                        const ItemIn& in, RowState& rs, double& lastp)
{
    double* ring = nd.pre_ring + (size_t)e * kPreRing;
    const double last = in.pc > 0 ? ring[(in.pc - 1) % kPreRing] : 0.0;
    // new presynaptic event (NGS:9091-9102)
    if (in.pc == 0 || T > last + 30.0) {
        ring[in.pc % kPreRing] = T;
        nd.pre_cnt[e] = in.pc + 1;
        if (T > lastp) lastp = T;
        // a saturated row takes LTD on supersynapses only, and reverts its
        // remodeling when one of them falls below the threshold
        const bool saturated = rs.scnt == Nss;
        if (!saturated || in.sup == 1)
            ltd_on_delivery(nd, i, e, tgt, T, now, in, rs, saturated ? kHitLtdDeliverySaturated : kHitLtdDelivery);
        if (saturated && rs.scnt < Nss) {
            count_hit(nd, kHitDesaturation);
            update_row(nd, i);
            rs.load(nd, i);
        }
    }
    // conductance of active synapses with the weight before this item's LTD
    // (NGS:9080-9084; the inbox is read only at the sync)
    if (in.act == 1) inbox_put(nd.in_ge_ra, tgt, t, i, in.w, nd.err);
}

// Growth-mode RA->RA row of presynaptic neuron i for the window t0..t1
// (NGS:8991-9030 enqueue, 9069-9191 drain with LTD, 9375-9554 at the sync).
// Returns the number of the row's spikes still in flight after the window.
__device__ __forceinline__ int rr_row(const NetDev& nd, const double* wclk, int i, int t0, int t1, StreamHeads hd)
{
    const int Nss = nd.syn.Nss;
    uint64_t mask = nd.soma_mask[i];
    int sc = nd.st_count[i];
    // earliest delivery time among the in-flight spikes (their stream heads):
    // at a step before it, with no spike to enqueue, the drain and the
    // exhausted-stream pass change nothing, so the step is skipped
    double next_T = sc > 0 ? nd.st_next[i] : HUGE_VAL;
    if (mask == 0 && !(wclk[t1 - t0 + 1] >= next_T)) return sc;
    int* st_step = nd.st_step + (size_t)i * kMaxStreams;
    int* st_mode = nd.st_mode + (size_t)i * kMaxStreams;
    int* st_cur = nd.st_cursor + (size_t)i * kMaxStreams;
    int* st_n = nd.st_n + (size_t)i * kMaxStreams;
    const int base = nd.rr_ptr[i], deg = nd.rr_ptr[i + 1] - base;
    double lastp = nd.last_pre[i];
    const double lastp0 = lastp;
    RowState rs;
    rs.load(nd, i);
    // the first kRowHeads streams are drained from LDS (hd; T = +inf:
    // exhausted): a drain step then loads only the consumed stream's next item.
    // Rows with more streams take the path through memory.
    bool cached = false;
    for (int t = t0; t <= t1; ++t) {
        const bool spike = (mask >> (t - t0)) & 1ull;
        const double now = wclk[t - t0 + 1];
        if (!spike && !(now >= next_T)) continue;
        bool exhausted = false;   // a stream ran out in this step (the pass below then drops it)
        if (spike) {
            if (sc == kMaxStreams) {
                atomicOr(nd.err, (unsigned)kErrStreams);
            } else {
                st_step[sc] = t;
                st_cur[sc] = 0;
                int mode, n;
                if (rs.scnt == Nss) {
                    // saturated: deliver to supersynaptic targets only, queue sorted by (T, target)
                    mode = 1;
                    n = rs.scnt;
                    int* items = nd.st_items + ((size_t)i * kMaxStreams + sc) * Nss;
                    const double cs = wclk[t - t0 + 1];
                    const int* slist = nd.sup_list + (size_t)i * Nss;
                    for (int k = 0; k < n; ++k) {
                        int e = edge_of(nd, i, slist[k]);
                        double T = cs + nd.rr_d[e];
                        int tg = slist[k];
                        int q = k;
                        while (q > 0) {
                            int pe = items[q - 1];
                            double pT = cs + nd.rr_d[pe];
                            int ptg = nd.rr_tgt[pe];
                            if (pT > T || (pT == T && ptg > tg)) {
                                items[q] = pe;
                                --q;
                            } else
                                break;
                        }
                        items[q] = e;
                    }
                } else {
                    mode = 0;
                    n = deg;
                }
                st_mode[sc] = mode;
                st_n[sc] = n;
                ++sc;
                if (cached) {
                    if (sc > kRowHeads) {
                        heads_store(nd, i, sc - 1, hd);
                        cached = false;
                    } else {
                        hd.open(sc - 1, wclk[t - t0 + 1], 0, n, mode);
                        double T;
                        int tg, e;
                        if (!head_find(nd, i, base, sc - 1, hd, 0, RrPos{}, false, T, tg, e)) {
                            T = HUGE_VAL;
                            exhausted = true;
                        }
                        hd.set(sc - 1, T, tg, e);
                    }
                }
            }
        }
        nd.st_count[i] = sc;
        if (!cached && sc <= kRowHeads) {
            for (int q = 0; q < sc; ++q) hd.open(q, nd.clock[st_step[q]], st_cur[q], st_n[q], st_mode[q]);
            for (int q = 0; q < sc; ++q) {
                double T;
                int tg, e;
                if (!head_find(nd, i, base, q, hd, hd.cur(q), RrPos{}, false, T, tg, e)) {
                    T = HUGE_VAL;
                    exhausted = true;
                }
                hd.set(q, T, tg, e);
            }
            cached = true;
        }
        // drain in (delivery_time, target) order across all in-flight spikes;
        // the pass that finds nothing due leaves the least head time (next_T)
        double heads_min = HUGE_VAL;
        for (;;) {
            if (cached) {
                int best = -1;
                double bT = 0, mT = HUGE_VAL;
                int btg = 0;
                auto consider = [&](int q, double T, int tg) {
                    mT = T < mT ? T : mT;
                    if (now >= T && (best < 0 || T < bT || (T == bT && tg < btg))) {
                        best = q; bT = T; btg = tg;
                    }
                };
                // heads read two at a time (the table holds kRowHeads columns, so
                // a read beyond sc stays inside it and is unused)
                static_assert(kRowHeads % 2 == 0, "heads read in pairs");
                for (int q = 0; q < sc; q += 2) {
                    const double Ta = hd.T(q), Tb = hd.T(q + 1);
                    const int ga = hd.tg(q), gb = hd.tg(q + 1);
                    consider(q, Ta, ga);
                    if (q + 1 < sc) consider(q + 1, Tb, gb);
                }
                if (best < 0) {
                    heads_min = mT;
                    break;
                }
                const int be = hd.e(best);
                // the item's reads and the consumed stream's next item, loaded together
                const ItemIn in = item_load(nd, be, btg);
                const int c = hd.cur(best) + 1;
                const bool more = c < hd.n(best);
                RrPos p{};
                if (more) {
                    if (hd.mode(best) == 0)
                        p = nd.rr_pos[base + c];
                    else
                        p.e = nd.st_items[((size_t)i * kMaxStreams + best) * Nss + c];
                }
                double nT;
                int ntg, ne;
                if (!head_find(nd, i, base, best, hd, c, p, more, nT, ntg, ne)) {
                    nT = HUGE_VAL;
                    exhausted = true;
                }
                hd.set(best, nT, ntg, ne);
                rr_item(nd, i, be, btg, bT, t, now, Nss, in, rs, lastp);
                continue;
            }
            int best = -1;
            double bT = 0;
            int btg = 0, be = 0;
            for (int q = 0; q < sc; ++q) {
                double T;
                int tg, e;
                if (!stream_peek(nd, i, q, T, tg, e)) continue;
                if (!(now >= T)) continue;
                if (best < 0 || T < bT || (T == bT && tg < btg)) {
                    best = q; bT = T; btg = tg; be = e;
                }
            }
            if (best < 0) break;
            st_cur[best] += 1;
            rr_item(nd, i, be, btg, bT, t, now, Nss, item_load(nd, be, btg), rs, lastp);
        }
        if (cached) {
            // no stream ran out: the table is unchanged, next_T from the heads
            if (!exhausted) {
                next_T = heads_min;
                continue;
            }
            heads_store(nd, i, sc, hd);
            cached = false;   // the pass below renumbers the streams
        }
        // drop exhausted streams, keeping order
        int w = 0;
        next_T = HUGE_VAL;
        for (int q = 0; q < sc; ++q) {
            double T;
            int tg, e;
            if (stream_peek(nd, i, q, T, tg, e)) {
                if (T < next_T) next_T = T;
                if (w != q) {
                    st_step[w] = st_step[q]; st_mode[w] = st_mode[q]; st_cur[w] = st_cur[q]; st_n[w] = st_n[q];
                    if (st_mode[q] == 1) {
                        int* dst = nd.st_items + ((size_t)i * kMaxStreams + w) * Nss;
                        int* src = nd.st_items + ((size_t)i * kMaxStreams + q) * Nss;
                        for (int k = 0; k < st_n[q]; ++k) dst[k] = src[k];
                    }
                }
                ++w;
            }
        }
        sc = w;
        nd.st_count[i] = sc;
    }
    if (cached) heads_store(nd, i, sc, hd);
    if (lastp != lastp0) nd.last_pre[i] = lastp;
    nd.st_next[i] = next_T;
    return sc;
}

// ---------------------------------------------------------------------------
// deliver, sweep form: one lane (group) per presynaptic neuron and role
//   role 0: RA->RA rows (growth) or chain-test RA->RA push
//   role 1: RA->I push        role 2: I->RA push
// Every neuron's row state or slot-table head is read, so a push starts one
// memory round trip after the kernel does; the bytes grow with the network
// (128 per neuron and fixed-weight role), which the list form below avoids.
// ---------------------------------------------------------------------------
// two waves per SIMD: the register budget of the fixed-weight pushes and the
// row drain (its LTD steps) is held to 256 so that a launch's ~2 waves per SIMD
// are resident at once
// GA / GB lanes per presynaptic neuron of the RA->I / I->RA pushes
template <int GA, int GB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_deliver_sweep(NetDev nd, int t0,
                                                                                             int t1, int b0, int b1)
{
    const int b = blockIdx.x, bs = blockDim.x, tid = threadIdx.x;
    __shared__ double wclk[kMaxWindowSteps + 1];
    extern __shared__ double heads_lds[];   // StreamHeads of the block's rows
    // the window's clock, and for a fixed-weight push its row, the window's
    // spikes and the slot-table head: all issued before the block waits on the
    // clock
    const int nclk = t1 - t0 + 2;
    const bool ra = b < b0 + b1;
    const int q = (ra ? b - b0 : b - b0 - b1) * bs + tid;
    const int G = ra ? GA : GB;
    const int pre = q / G, r = q % G;
    const bool push = b >= b0 && pre < (ra ? nd.n_ra : nd.n_i);
    const FixedEdges fe = ra ? nd.ra2i : nd.i2ra;
    const InFlight fl = ra ? nd.fl_ra2i : nd.fl_i2ra;
    int rb = 0, re = 0, held = 0;
    uint64_t spikes = 0;
    if (push) {
        held = fl.held[pre];
        rb = fe.ptr[pre];
        re = fe.ptr[pre + 1];
        spikes = ra ? nd.soma_mask[pre] : nd.i_mask[pre];
    }
    const double c0 = tid < nclk ? nd.clock[t0 - 1 + tid] : 0.0;
    const double c1 = tid + bs < nclk ? nd.clock[t0 - 1 + tid + bs] : 0.0;
    // the slot-table head line only of a neuron with edges and spikes held or
    // new (push_fixed does nothing for the others): in most windows few
    // neurons have any, so a window reads ~1.3 MB at 25 500 neurons instead of
    // every neuron's 128-byte line (4.3 MB)
    const bool work = push && re > rb && (held > 0 || spikes != 0);
    FlHead h;
    if (work) h = fl.head[pre];
    if (tid < nclk) wclk[tid] = c0;
    if (tid + bs < nclk) wclk[tid + bs] = c1;
    for (int k = tid + 2 * bs; k < nclk; k += bs) wclk[k] = nd.clock[t0 - 1 + k];
    __syncthreads();
    // the other post-event buffer was read by the previous window's STDP (in
    // this window's integration); the next window's integration fills it
    if (b == 0 && threadIdx.x == 0) *nd.post_list_next_cnt = 0;
    if (b < b0) {
        // RA->RA: one lane per row
        const int i = b * blockDim.x + threadIdx.x;
        if (i >= nd.n_ra) return;
        const StreamHeads hd{heads_lds + tid, (int*)(heads_lds + kHeadDoubles * kRowHeads * bs) + tid, bs};
        const int growth = nd.growth_mode == 2 ? nd.net_growth[nd.ra_net[i]] : nd.growth_mode;
        if (growth)
            rr_row(nd, wclk, i, t0, t1, hd);
        else
            push_fixed<1>(nd, wclk, i, 0, nd.rr_ptr[i], nd.rr_ptr[i] + nd.act_cnt[i], nd.chain_e, nd.rr_tgt, nd.w,
                          nd.rr_d, nd.fl_rr, nd.fl_rr.head[i], nd.soma_mask[i], t0, t1, nd.in_ge_ra);
        return;
    }
    // fixed-weight pushes: G lanes per presynaptic neuron
    if (!work) return;
    if (ra)
        push_fixed<GA>(nd, wclk, pre, r, rb, re, nullptr, fe.tgt, fe.w, fe.d, fl, h, spikes, t0, t1, nd.in_ge_i);
    else
        push_fixed<GB>(nd, wclk, pre, r, rb, re, nullptr, fe.tgt, fe.w, fe.d, fl, h, spikes, t0, t1, nd.in_gi_ra);
}

// ---------------------------------------------------------------------------
// deliver, list form: the presynaptic neurons with work in the window, per
// role (as above).  The window visits the neurons on the role's list only: those that spiked in
// it (listed by the integration) and those with spikes still in flight after
// the previous window (listed here), so it reads the rows, slot tables and
// edges of its active neurons and nothing else.  Item k of a role goes to wave
// k mod W (W = the role's waves): up to W items run one per wave, side by side.
// ---------------------------------------------------------------------------
template <int GA, int GB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_deliver_list(NetDev nd, int t0,
                                                                                            int t1, int b0, int b1)
{
    const int b = blockIdx.x, bs = blockDim.x, tid = threadIdx.x;
    const int w = nd.window;
    __shared__ double wclk[kMaxWindowSteps + 1];
    extern __shared__ double heads_lds[];   // StreamHeads of the block's rows
    const int role = b < b0 ? 0 : (b < b0 + b1 ? 1 : 2);
    const int G = role == 0 ? 1 : (role == 1 ? GA : GB);
    const int wpb = bs >> 6;
    const int nwaves = (role == 0 ? b0 : (role == 1 ? b1 : (int)gridDim.x - b0 - b1)) * wpb;
    const int wave = (b - (role == 0 ? 0 : (role == 1 ? b0 : b0 + b1))) * wpb + (tid >> 6);
    const int lane = tid & 63, r = lane % G;
    const int cap = nd.wl_cap[role];
    const int* list_now = nd.wl_list[role] + (size_t)(w & 1) * cap;
    // the window's clock, the list's count and the lane's first item: all
    // issued before the block waits on the clock
    const int nclk = t1 - t0 + 2;
    const double c0 = tid < nclk ? nd.clock[t0 - 1 + tid] : 0.0;
    const double c1 = tid + bs < nclk ? nd.clock[t0 - 1 + tid + bs] : 0.0;
    const int k0 = (lane / G) * nwaves + wave;
    const int n_items = nd.wl_cnt[3 * role + w % 3];
    const int first = k0 < cap ? list_now[k0] : 0;
    if (tid < nclk) wclk[tid] = c0;
    if (tid + bs < nclk) wclk[tid + bs] = c1;
    for (int k = tid + 2 * bs; k < nclk; k += bs) wclk[k] = nd.clock[t0 - 1 + k];
    if (b == 0 && tid == 0) {
        // the other post-event buffer was read by the previous window's STDP (in
        // this window's integration); the next window's integration fills it
        *nd.post_list_next_cnt = 0;
        // the lists of window w + 2 start empty (window w - 1's, done with)
        for (int q = 0; q < 3; ++q) nd.wl_cnt[3 * q + (w + 2) % 3] = 0;
    }
    __syncthreads();
    int* list_next = nd.wl_list[role] + (size_t)((w + 1) & 1) * cap;
    int* cnt_next = nd.wl_cnt + 3 * role + (w + 1) % 3;
    const StreamHeads hd{heads_lds + tid, (int*)(heads_lds + kHeadDoubles * kRowHeads * bs) + tid, bs};
    for (int k = k0; k < n_items; k += nwaves * (64 / G)) {
        const int i = k == k0 ? first : list_now[k];
        int left;
        if (role == 0) {
            const int growth = nd.growth_mode == 2 ? nd.net_growth[nd.ra_net[i]] : nd.growth_mode;
            if (growth)
                left = rr_row(nd, wclk, i, t0, t1, hd);
            else
                left = push_fixed<1>(nd, wclk, i, 0, nd.rr_ptr[i], nd.rr_ptr[i] + nd.act_cnt[i], nd.chain_e,
                                     nd.rr_tgt, nd.w, nd.rr_d, nd.fl_rr, nd.fl_rr.head[i], nd.soma_mask[i], t0, t1,
                                     nd.in_ge_ra);
        } else if (role == 1) {
            left = push_fixed<GA>(nd, wclk, i, r, nd.ra2i.ptr[i], nd.ra2i.ptr[i + 1], nullptr, nd.ra2i.tgt,
                                  nd.ra2i.w, nd.ra2i.d, nd.fl_ra2i, nd.fl_ra2i.head[i], nd.soma_mask[i], t0, t1,
                                  nd.in_ge_i);
        } else {
            left = push_fixed<GB>(nd, wclk, i, r, nd.i2ra.ptr[i], nd.i2ra.ptr[i + 1], nullptr, nd.i2ra.tgt,
                                  nd.i2ra.w, nd.i2ra.d, nd.fl_i2ra, nd.fl_i2ra.head[i], nd.i_mask[i], t0, t1,
                                  nd.in_gi_ra);
        }
        // still in flight: listed for the next window
        if (left > 0 && r == 0) {
            const int pos = atomicAdd(cnt_next, 1);
            if (pos < cap) {
                list_next[pos] = i;
                nd.wl_mark[role][i] = w + 1;
            } else {
                atomicOr(nd.err, (unsigned)kErrWorklist);
            }
        }
    }
}

// sync: apply the window's conductance updates (NGS:9309-9366).  Within a
// trial k_integrate applies each window's updates as it loads the neurons for
// the next window; this kernel applies the last window's.
__global__ __launch_bounds__(512) void k_sync(NetDev nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd.n_ra) {
        int n = nd.in_gi_ra.cnt[i];
        if (n > 0) {
            if (n > nd.in_gi_ra.cap) { atomicOr(nd.err, (unsigned)kErrInbox); n = nd.in_gi_ra.cap; }
            nd.ra[11][i] += inbox_sum(nd.in_gi_ra, i, n);          // raiseI
            nd.in_gi_ra.cnt[i] = 0;
        }
        n = nd.in_ge_ra.cnt[i];
        if (n > 0) {
            if (n > nd.in_ge_ra.cap) { atomicOr(nd.err, (unsigned)kErrInbox); n = nd.in_ge_ra.cap; }
            nd.ra[10][i] += HVC_CDIV(inbox_sum(nd.in_ge_ra, i, n), 10000);  // raiseExcWeight: w / Ad
            nd.in_ge_ra.cnt[i] = 0;
        }
    }
    if (i < nd.n_i) {
        int n = nd.in_ge_i.cnt[i];
        if (n > 0) {
            if (n > nd.in_ge_i.cap) { atomicOr(nd.err, (unsigned)kErrInbox); n = nd.in_ge_i.cap; }
            nd.in[5][i] += inbox_sum(nd.in_ge_i, i, n);              // HHI raiseE
            nd.in_ge_i.cnt[i] = 0;
        }
    }
}

// end of a growth trial: potentiation_decay + update_all_synapses (NGS:9567-9570)
__global__ __launch_bounds__(256) void k_decay(NetDev nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd.n_ra || !nd.net_growth[nd.ra_net[i]]) return;
    const hvc_synaptic_params& p = nd.syn;
    for (int e = nd.rr_ptr[i]; e < nd.rr_ptr[i + 1]; ++e) {
        double w = nd.w[e];
        if (w >= p.SUPERSYNAPSE_THRESHOLD)
            w -= p.BETA_SUPERSYNAPSE;
        else
            w -= p.BETA;
        if (w < 0) w = 0;
        nd.w[e] = w;
    }
    update_row(nd, i);
}

// chain test: each row's active edges sorted by (delay, target)
__global__ void k_chain_edges(NetDev nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd.n_ra || nd.net_growth[nd.ra_net[i]]) return;
    int base = nd.rr_ptr[i];
    int n = nd.act_cnt[i];
    int* out = nd.chain_e + base;
    for (int k = 0; k < n; ++k) {
        int e = edge_of(nd, i, nd.act_list[nd.rr_ptr[i] + k]);
        double d = nd.rr_d[e];
        int tg = nd.rr_tgt[e];
        int q = k;
        while (q > 0) {
            int pe = out[q - 1];
            if (nd.rr_d[pe] > d || (nd.rr_d[pe] == d && nd.rr_tgt[pe] > tg)) {
                out[q] = pe;
                --q;
            } else
                break;
        }
        out[q] = e;
    }
}

// device evaluation of the engine's math for the CPU cross-check:
// fn 0 exp(a), 1 log(a), 2 a / b through the constant-divisor path (Y = RN(1/b))
__global__ void k_eval_math(int fn, int n, const double* a, const double* b, double* out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double x = a[k];
    double r;
    if (fn == 0)
        r = dexp(x);
    else if (fn == 1)
        r = dlog(x);
    else
        r = cdiv_checked(x, b[k]);
    out[k] = r;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_decay_table(hipStream_t st, const double* clock, const double* offsets, int num_steps, void* rows)
{
    k_decay_table<<<blocks_for(num_steps, 128), 128, 0, st>>>(clock, offsets, num_steps, (DecayRow*)rows);
}

void launch_reset(hipStream_t st, const NetDev& nd, int E)
{
    int n = nd.n_ra > nd.n_i ? nd.n_ra : nd.n_i;
    if (3 * nd.n_nets > n) n = 3 * nd.n_nets;
    if (kWlCounters > n) n = kWlCounters;
    k_reset<<<blocks_for(n, 256), 256, 0, st>>>(nd);
    if (E > 0) k_clear_pre<<<blocks_for(E, 256), 256, 0, st>>>(nd.pre_cnt, E);
}

void launch_deliver(hipStream_t st, const NetDev& nd, int t0, int t1, const Tuning& tu)
{
    const int bs = tu.deliver_block;
    const size_t lds = (size_t)kRowHeads * bs * (kHeadDoubles * sizeof(double) + kHeadInts * sizeof(int));
    // blocks of a role: enough waves for one lane (group) per neuron, at most
    // kDeliverRoleWaves (the items then loop over the role's waves)
    auto role_blocks = [&](long long lanes) {
        const long long waves = tu.deliver_lists ? std::min<long long>((lanes + 63) / 64, kDeliverRoleWaves)
                                                 : (lanes + 63) / 64;
        return (int)((waves * 64 + bs - 1) / bs);
    };
    // a quad per HVC-I neuron shortens the chain of inbox atomics of its ~30
    // I->RA edges; an HVC-RA neuron's few RA->I edges take one lane.  On the
    // lists (at most kDeliverRoleWaves waves per role) quads serve at every
    // size; the sweep's grid grows with the neurons, and above
    // kQuadPushNeurons its extra waves cost more than they save, so every
    // push takes one lane there.
    if (tu.deliver_lists || nd.n_ra + nd.n_i <= kQuadPushNeurons) {
        const int b0 = role_blocks(nd.n_ra), b1 = role_blocks(nd.n_ra), b2 = role_blocks(4ll * nd.n_i);
        if (b0 + b1 + b2 == 0) return;
        if (tu.deliver_lists)
            k_deliver_list<1, 4><<<b0 + b1 + b2, bs, lds, st>>>(nd, t0, t1, b0, b1);
        else
            k_deliver_sweep<1, 4><<<b0 + b1 + b2, bs, lds, st>>>(nd, t0, t1, b0, b1);
    } else {
        const int b0 = role_blocks(nd.n_ra), b1 = role_blocks(nd.n_ra), b2 = role_blocks(nd.n_i);
        if (b0 + b1 + b2 == 0) return;
        k_deliver_sweep<1, 1><<<b0 + b1 + b2, bs, lds, st>>>(nd, t0, t1, b0, b1);
    }
}

__global__ void k_add_i32(int* a, long long n, int delta)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] += delta;
}

void launch_add_i32(hipStream_t st, int* a, long long n, int delta)
{
    if (n > 0) k_add_i32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, n, delta);
}

// pos[k] = (rr_perm[k], its target, its delay) for the positions [0, E)
__global__ void k_rr_pos(NetDev nd, RrPos* pos, int E)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < E) {
        const int e = nd.rr_perm[k];
        pos[k] = RrPos{e, nd.rr_tgt[e], nd.rr_d[e]};
    }
}

void launch_rr_pos(hipStream_t st, const NetDev& nd, RrPos* pos, int E)
{
    if (E > 0) k_rr_pos<<<blocks_for(E, 256), 256, 0, st>>>(nd, pos, E);
}

void launch_noise_events(hipStream_t st, const NetDev& nd, int tb, int te, int gen_ra, int gen_i)
{
    const int bra = blocks_for(nd.n_ra, 64), bi = blocks_for(nd.n_i, 64);
    k_noise_events<<<bra + bi, 64, 0, st>>>(nd, tb, te, bra, gen_ra, gen_i);
}

// the last window's sync-point STDP (earlier windows run it inside the next
// window's k_integrate)
__global__ __launch_bounds__(256) void k_stdp(NetDev nd, StdpArgs sa)
{
    __shared__ uint64_t slice[kStdpSliceCap];
    __shared__ int scount;
    stdp_block(nd, sa, blockIdx.x * blockDim.x, slice, &scount);
}

void launch_stdp(hipStream_t st, const NetDev& nd, const StdpArgs& sa)
{
    k_stdp<<<blocks_for(nd.n_ra, 256), 256, 0, st>>>(nd, sa);
}

void launch_sync(hipStream_t st, const NetDev& nd, const Tuning& tu)
{
    int n = nd.n_ra > nd.n_i ? nd.n_ra : nd.n_i;
    k_sync<<<blocks_for(n, tu.sync_block), tu.sync_block, 0, st>>>(nd);
}

void launch_decay(hipStream_t st, const NetDev& nd)
{
    k_decay<<<blocks_for(nd.n_ra, 64), 64, 0, st>>>(nd);
}

void launch_eval_math(hipStream_t st, int fn, int n, const double* a, const double* b, double* out)
{
    k_eval_math<<<blocks_for(n, 256), 256, 0, st>>>(fn, n, a, b, out);
}

void launch_chain_edges(hipStream_t st, const NetDev& nd)
{
    k_chain_edges<<<blocks_for(nd.n_ra, 64), 64, 0, st>>>(nd);
}

}  // namespace hvc
