// hvc_integrate.hip.h -- neuron integration kernels of one precision form.
//
// Included once by hvc_kernels.hip (FP64 form, bit-exact) and once by
// hvc_kernels_f32.hip (FP32 form, HVC_FP32 defined); the kernels land in
// hvc::p64 or hvc::p32, the launchers are named launch_* and launch_*_f32.
// Neuron state is stored in HBM as FP64 structure-of-arrays in both forms; the
// FP32 form rounds it to float when a window starts and widens it when the
// window ends.
#pragma once
#include <hip/hip_runtime.h>

#include "hvc_internal.h"
#include "hvc_neuron.hip.h"
#include "hvc_stdp.hip.h"

namespace hvc {

static inline int blocks_for(int n, int bs) { return n > 0 ? (n + bs - 1) / bs : 1; }

#ifndef HVC_INBOX_DEFINED
#define HVC_INBOX_DEFINED
// ordered sum of one target's inbox: records sorted by (arrival step, presynaptic id)
__device__ __forceinline__ double inbox_sum(const Inbox& ib, int tgt, int n)
{
    Record* r = ib.rec + (size_t)tgt * ib.cap;
    for (int a = 1; a < n; ++a) {
        Record x = r[a];
        int b = a - 1;
        while (b >= 0 && r[b].key > x.key) {
            r[b + 1] = r[b];
            --b;
        }
        r[b + 1] = x;
    }
    double sum = 0.0;
    for (int a = 0; a < n; ++a) sum += r[a].w;
    return sum;
}

// the previous window's sync for one target (NGS:9309-9366): the ordered sum
// of its inbox, if any, and the inbox emptied
__device__ __forceinline__ bool inbox_take(const NetDev& nd, const Inbox& ib, int tgt, double& sum)
{
    int n = ib.cnt[tgt];
    if (n <= 0) return false;
    if (n > ib.cap) {
        atomicOr(nd.err, (unsigned)kErrInbox);
        n = ib.cap;
    }
    sum = inbox_sum(ib, tgt, n);
    ib.cnt[tgt] = 0;
    return true;
}
#endif

HVC_PREC_BEGIN

// One neuron step.  FP64: with the branch-free constant divisions; if any lane
// of the wave met a quotient outside the proven range, the wave repeats the
// step from the saved state with IEEE division (EXACT), which gives the same
// bits.  FP32: the step as is.
#define HVC_GUARDED_STEP(State, call_fast, call_exact) \
    Guard bad;                                         \
    if constexpr (Guard::kActive) {                    \
        State s0 = s;                                  \
        call_fast;                                     \
        if (__any(bad.failed())) {                     \
            s = s0;                                    \
            call_exact;                                \
        }                                              \
    } else {                                           \
        call_fast;                                     \
    }

__device__ __forceinline__ void ra_step_debraband(RAState& s, real el, real gca, const RAUniform& u, real J1s,
                                                  real J2s, real J1d, real J2d, real E_dt, real E_half)
{
    HVC_GUARDED_STEP(RAState, (ra_debraband<false>(bad, s, el, gca, u, J1s, J2s, J1d, J2d, E_dt, E_half)),
                     (ra_debraband<true>(bad, s, el, gca, u, J1s, J2s, J1d, J2d, E_dt, E_half)))
}

template <bool LANES = false>
__device__ __forceinline__ void ra_step_runge4(RAState& s, real el, real gca, const RAUniform& u, real E_third,
                                               real E_twothird, real E_dt, const RAGateSlot* sl = nullptr)
{
    HVC_GUARDED_STEP(RAState, (ra_runge4<false, LANES>(bad, s, el, gca, u, sl, E_third, E_twothird, E_dt)),
                     (ra_runge4<true, LANES>(bad, s, el, gca, u, sl, E_third, E_twothird, E_dt)))
}

template <int MODEL, class Row>
__device__ __forceinline__ void i_step(IState& s, real dt, const Row& row)
{
    if (MODEL == 0) {
        HVC_GUARDED_STEP(IState, (i_dp8<false>(bad, s, dt, row)), (i_dp8<true>(bad, s, dt, row)))
    } else {
        HVC_GUARDED_STEP(IState, (i_runge4<false>(bad, s, dt, row)), (i_runge4<true>(bad, s, dt, row)))
    }
}

template <class Row>
__device__ __forceinline__ void i_step_lanes(IState& s, real& xlo, int role, int base, real dt, const Row& row)
{
    HVC_GUARDED_STEP(IState, (i_dp8_lanes<false>(bad, s, xlo, role, base, dt, row)),
                     (i_dp8_lanes<true>(bad, s, xlo, role, base, dt, row)))
}

template <class Row>
__device__ __forceinline__ void i_step_quad(IState& s, real& xlo, real& wlo, int role, const GateSlot& ga,
                                            const GateSlot& gb, real dt, const Row& row)
{
    // the low parts are state too: a repeated step starts from their saved values
    real xlo0 = xlo, wlo0 = wlo;
    HVC_GUARDED_STEP(IState, (i_dp8_quad<false>(bad, s, xlo, wlo, role, ga, gb, dt, row)),
                     (xlo = xlo0, wlo = wlo0, i_dp8_quad<true>(bad, s, xlo, wlo, role, ga, gb, dt, row)))
}
template <class Row>
__device__ __forceinline__ void i_step_oct(IState& s, real& xo, real& xlo, int q, const GateSlot& g, real dt,
                                           const Row& row)
{
    // the lane's variable and its low part are state too
    real xo0 = xo, xlo0 = xlo;
    HVC_GUARDED_STEP(IState, (i_dp8_oct<false>(bad, s, xo, xlo, q, g, dt, row)),
                     (xo = xo0, xlo = xlo0, i_dp8_oct<true>(bad, s, xo, xlo, q, g, dt, row)))
}
#undef HVC_GUARDED_STEP

// state variables between HBM (FP64) and registers: the FP32 form keeps a
// variable as float hi + float lo (compensated accumulation), which the FP64
// copy in HBM holds exactly
#ifdef HVC_FP32
#define HVC_LOAD(st, f, d)                            \
    {                                                 \
        double d_ = (d);                              \
        st.f = (real)d_;                              \
        st.f##_lo = (real)(d_ - (double)st.f);        \
    }
#define HVC_WIDE(st, f) ((double)st.f + (double)st.f##_lo)
#else
#define HVC_LOAD(st, f, d) st.f = (d);
#define HVC_WIDE(st, f) (st.f)
#endif

// per-network stepper constants in the form's type
__device__ __forceinline__ RAUniform ra_uniform(const NetDev& nd)
{
    return RAUniform{(real)nd.sq_sig_s, (real)nd.sq_sig_d, (real)nd.mu_term_s, (real)nd.mu_term_d, (real)nd.dt};
}

// The coupling window's per-step tables, staged in LDS by k_integrate: decay
// rows of the steps t0..t1 and the clock at t0 - 1 .. t1 (every step of every
// neuron reads them; from HBM / L2 each read would stall the step).
struct Win {
    const WinRow* rows;     // rows[t - t0]
    const double* clock;    // clock[t - t0 + 1] = clock at step t
    int t0;
    double* lds;            // the block's copy: nsteps rows, then the clock
    int nsteps;
    __device__ __forceinline__ const WinRow& row(int t) const { return rows[t - t0]; }
    __device__ __forceinline__ double clk(int t) const { return clock[t - t0 + 1]; }
    // copy the tables into LDS (every thread of the block, after it has issued
    // its state loads, so their latencies overlap)
    __device__ __forceinline__ void stage(const NetDev& nd) const
    {
        const DecayRow* grow = (const DecayRow*)nd.decay + t0;
        WinRow* rl = (WinRow*)lds;
#ifdef HVC_FP32
        for (int k = threadIdx.x; k < nsteps * kNumOffsets; k += blockDim.x) {
            const int r = k / kNumOffsets, o = k - r * kNumOffsets;
            rl[r].e5[o] = (float)grow[r].e5[o];
            rl[r].e2[o] = (float)grow[r].e2[o];
            if (o == kOffDt) {
                rl[r].e5dt = grow[r].e5[o];
                rl[r].e2dt = grow[r].e2[o];
            }
        }
#else
        constexpr int kRowWords = (int)(sizeof(DecayRow) / sizeof(double));
        for (int k = threadIdx.x; k < nsteps * kRowWords; k += blockDim.x) lds[k] = ((const double*)grow)[k];
#endif
        double* wc = (double*)(rl + nsteps);
        for (int k = threadIdx.x; k <= nsteps; k += blockDim.x) wc[k] = nd.clock[t0 - 1 + k];
        __syncthreads();
    }
};

// The neuron's next two noise events of the chunk (lists are generated by
// k_noise_events).  take() applies every event of step t, in list order; the
// following entry is loaded a step or more before it is looked at, except
// when two events fall on one step.
struct NoiseCursor {
    const NoiseEvent* ev;
    int pos;
    NoiseEvent cur, nxt;
    __device__ __forceinline__ void open(const NoiseEvent* list, int p)
    {
        ev = list;
        pos = p;
        cur = ev[p];
        nxt = ev[p + 1];
    }
    __device__ __forceinline__ void none()
    {
        ev = nullptr;
        pos = 0;
        cur.step = kNoNoiseStep;
        nxt.step = kNoNoiseStep;
    }
    template <class Apply>
    __device__ __forceinline__ void take(int t, Apply apply)
    {
        while (cur.step == t) {
            apply(cur.target, cur.inc);
            cur = nxt;
            ++pos;
            nxt = ev[pos + 1];
        }
    }
};

// ---------------------------------------------------------------------------
// single-neuron traces
// ---------------------------------------------------------------------------
__global__ void k_trace_ra(NetDev nd, int n, int n_steps, uint32_t trial, const double* EL, const double* GCa,
                           const double* kick, const double* exc_w, const double* inh_g, double* out, uint8_t* fired_out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    const RAUniform u = ra_uniform(nd);
    // HH2_test::set_to_rest (HH2_test.cpp:196-205)
    RAState s;
    HVC_LOAD(s, Vs, -79.97619025) HVC_LOAD(s, Vd, -79.97268759) HVC_LOAD(s, n, 0.01101284)
    HVC_LOAD(s, l, 0.0) HVC_LOAD(s, h, 0.9932845) HVC_LOAD(s, r, 0.00055429)
    HVC_LOAD(s, c, 0.00000261762353) HVC_LOAD(s, Ca, 0.01689572)
    s.Ge_s = 0; s.Gi_s = 0; s.Ge_d = 0; s.Gi_d = 0;
    int fs = 0, fd = 0;
    const real el = (real)EL[k], gca = (real)GCa[k];
    PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamRAPoisson, (uint32_t)k, 0u};
    double nes = 0, nis = 0, ned = 0, nid = 0;
    if (nd.ra_poisson) {
        nes = 1000.0 * ps.spike_time(100.0);
        nis = 1000.0 * ps.spike_time(100.0);
        ned = 1000.0 * ps.spike_time(100.0);
        nid = 1000.0 * ps.spike_time(100.0);
    }
    auto dump = [&](int row) {
        double* o = out + ((size_t)row * n + k) * 12;
        o[0] = HVC_WIDE(s, Vs); o[1] = HVC_WIDE(s, Vd); o[2] = HVC_WIDE(s, n); o[3] = HVC_WIDE(s, l);
        o[4] = HVC_WIDE(s, h); o[5] = HVC_WIDE(s, r); o[6] = HVC_WIDE(s, c); o[7] = HVC_WIDE(s, Ca);
        o[8] = s.Ge_s; o[9] = s.Gi_s; o[10] = s.Ge_d; o[11] = s.Gi_d;
    };
    dump(0);
    for (int t = 1; t <= n_steps; ++t) {
        size_t idx = (size_t)(t - 1) * n + k;
        double kk = kick[idx];
        if (kk != 0.0) s.Ge_d += kk;
        int f = ra_state_check(s.Vs, s.Vd, fs, fd);
        const DecayRow& row = rows[t];
        if (nd.ra_model == 0) {
            U4 o = philox4x32((uint32_t)t, (uint32_t)k, trial, kStreamRAWhite, nd.seed_lo, nd.seed_hi);
            ra_step_debraband(s, el, gca, u, kPointDistribution[point_index(o.x)], kPointDistribution[point_index(o.y)],
                         kPointDistribution[point_index(o.z)], kPointDistribution[point_index(o.w)],
                         row.e5[kOffDt], row.e5[kOffHalf]);
        } else {
            if (nd.ra_poisson) {
                double time = nd.clock[t - 1];
                ra_noise_check(ps, time, nd.dt, s.Gi_s, 0.035, nis);
                ra_noise_check(ps, time, nd.dt, s.Gi_d, 0.045, nid);
                ra_noise_check(ps, time, nd.dt, s.Ge_s, 0.035, nes);
                ra_noise_check(ps, time, nd.dt, s.Ge_d, 0.045, ned);
            }
            ra_step_runge4(s, el, gca, u, row.e5[kOffThird], row.e5[kOffTwoThird], row.e5[kOffDt]);
        }
        fired_out[idx] = (uint8_t)f;
        double ig = inh_g[idx];
        if (ig != 0.0) s.Gi_d += ig;
        double ew = exc_w[idx];
        if (ew != 0.0) s.Ge_d += ew / 10000.0;
        dump(t);
    }
}

__global__ void k_trace_i(NetDev nd, int n, int n_steps, int first_trial, uint32_t trial, const double* exc_g,
                          const double* inh_g, double* out, uint8_t* fired_out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    // HHI_buffer::set_to_rest (HHI_buffer.cpp:181-185)
    IState s;
    HVC_LOAD(s, V, -66.0) HVC_LOAD(s, n, 0.125) HVC_LOAD(s, m, 0.0) HVC_LOAD(s, h, 0.99) HVC_LOAD(s, w, 0.0)
    s.Ge = 0; s.Gi = 0;
    int flag = 0;
    PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamIPoisson, (uint32_t)k, 0u};
    double ne = 0, ni = 0;
    if (nd.i_poisson) {
        double init_dt = first_trial ? 0.0 : nd.dt;
        ne = i_init_noise(ps, init_dt);
        ni = i_init_noise(ps, init_dt);
    }
    auto dump = [&](int row) {
        double* o = out + ((size_t)row * n + k) * 7;
        o[0] = HVC_WIDE(s, V); o[1] = HVC_WIDE(s, n); o[2] = HVC_WIDE(s, m); o[3] = HVC_WIDE(s, h);
        o[4] = HVC_WIDE(s, w); o[5] = s.Ge; o[6] = s.Gi;
    };
    dump(0);
    for (int t = 1; t <= n_steps; ++t) {
        size_t idx = (size_t)(t - 1) * n + k;
        int f = i_state_check(s.V, flag);
        if (nd.i_poisson) {
            double time = nd.clock[t - 1];
            i_noise_check(ps, time, nd.dt, s.Ge, ne);
            i_noise_check(ps, time, nd.dt, s.Gi, ni);
        }
        if (nd.i_model == 0)
            i_step<0>(s, nd.dt, rows[t]);
        else
            i_step<1>(s, (real)nd.dt, rows[t]);
        fired_out[idx] = (uint8_t)f;
        double ig = inh_g[idx];
        if (ig != 0.0) s.Gi += ig;
        double eg = exc_g[idx];
        if (eg != 0.0) s.Ge += eg;
        dump(t);
    }
}

// append (neuron, step) to a log with one atomic per wave
__device__ __forceinline__ void wave_append(int2* log, int* cnt, int cap, bool pred, int neuron, int step,
                                            unsigned int* err)
{
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __popcll(m));
    base = __shfl(base, leader);
    if (pred) {
        int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < cap)
            log[pos] = make_int2(neuron, step);
        else
            atomicOr(err, (unsigned)kErrSpikeLog);
    }
}

// the window's spikes of the lanes (bit k of mask: step t0 + k) appended after
// the steps, one atomic per wave and round (a lane spikes at most once per
// window in practice, so one round); the host sorts the log
__device__ __forceinline__ void window_append(int2* log, int* cnt, int cap, uint64_t mask, int neuron, int t0,
                                              unsigned int* err)
{
    for (;;) {
        const bool pred = mask != 0;
        if (__ballot(pred) == 0) return;
        const int step = pred ? t0 + __ffsll((unsigned long long)mask) - 1 : 0;
        wave_append(log, cnt, cap, pred, neuron, step, err);
        mask &= mask - 1;
    }
}

// A neuron that spiked in the window goes on the role's delivery list for the
// window (the delivery kernel visits only listed neurons), unless the previous
// window's delivery listed it already (spikes still in flight); `listed` is
// the neuron's list mark, loaded when the window began.  One atomic per wave.
// A list holds each neuron at most once, so it never outgrows its neurons.
__device__ __forceinline__ void wl_enlist(const NetDev& nd, int role, bool spiked, int listed, int neuron)
{
    if (!nd.wl_on) return;
    const int w = nd.window;
    const bool pred = spiked && listed != w;
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&nd.wl_cnt[3 * role + w % 3], __popcll(m));
    base = __shfl(base, leader);
    if (pred) {
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < nd.wl_cap[role]) {
            nd.wl_list[role][(size_t)(w & 1) * nd.wl_cap[role] + pos] = neuron;
            nd.wl_mark[role][neuron] = w;
        } else {
            atomicOr(nd.err, (unsigned)kErrWorklist);
        }
    }
}

// ---------------------------------------------------------------------------
// integrate: HVC-RA neurons (NGS:8931-9216 minus delivery) and HVC-I neurons
// (NGS:9221-9286 minus delivery) through the steps t0..t1 of one window, in one
// launch.  HVC-I blocks come first: with five lanes per neuron they carry the
// longer per-step work, and the RA blocks run beside them.
// ---------------------------------------------------------------------------
constexpr int kRALanes = 4;                      // lanes per HVC-RA neuron (RK4 form)
constexpr int kRANeuronsPerWave = 64 / kRALanes;  // 16 neurons

// HVC-RA neurons: one thread per neuron (LANES false, `slot` = neuron index) or,
// RK4 form, four lanes per neuron (LANES true, `slot` = wave index; the quad's
// first lane does the neuron's logging and stores)
template <int MODEL, bool LANES>
__device__ void integrate_ra(const NetDev& nd, const Win& win, int slot, int t0, int t1)
{
    static_assert(!LANES || MODEL == 1, "four lanes per neuron: RK4 form only");
    const int lane = LANES ? __lane_id() : 0;
    const int i = LANES ? slot * kRANeuronsPerWave + (lane >> 2) : slot;
    const bool leader = !LANES || (lane & 3) == 0;
    RAGateSlot gs[2];
    if (LANES) {
        gs[0] = kRAGates[2 * (lane & 3)];
        gs[1] = kRAGates[2 * (lane & 3) + 1];
    }
    const bool live = i < nd.n_ra && (!nd.net_live || nd.net_live[nd.ra_net[i]]);
    const int ii = live ? i : 0;
    const int growth = nd.net_growth[nd.ra_net[ii]];
    const int listed0 = nd.wl_on ? nd.wl_mark[0][ii] : 0, listed1 = nd.wl_on ? nd.wl_mark[1][ii] : 0;
    const RAUniform u = ra_uniform(nd);
    RAState s;
    HVC_LOAD(s, Vs, nd.ra[0][ii]) HVC_LOAD(s, Vd, nd.ra[1][ii]) HVC_LOAD(s, n, nd.ra[2][ii])
    HVC_LOAD(s, l, nd.ra[3][ii]) HVC_LOAD(s, h, nd.ra[4][ii]) HVC_LOAD(s, r, nd.ra[5][ii])
    HVC_LOAD(s, c, nd.ra[6][ii]) HVC_LOAD(s, Ca, nd.ra[7][ii])
    win.stage(nd);
    // the previous window's sync: raiseI (Gi_d += G), raiseExcWeight (Ge_d += w / Ad)
    double ge_d = nd.ra[10][ii], gi_d = nd.ra[11][ii];
    if (live && leader) {
        double sum;
        if (inbox_take(nd, nd.in_gi_ra, i, sum)) gi_d += sum;
        if (inbox_take(nd, nd.in_ge_ra, i, sum)) ge_d += HVC_CDIV(sum, 10000);
    }
    if constexpr (LANES) {
        ge_d = quad_bcast<0>(ge_d);
        gi_d = quad_bcast<0>(gi_d);
    }
    s.Ge_s = nd.ra[8][ii]; s.Gi_s = nd.ra[9][ii]; s.Ge_d = ge_d; s.Gi_d = gi_d;
    int flags = nd.ra_flags[ii];
    int fs = flags & 1, fd = (flags >> 1) & 1;
    const real el = (real)nd.EL[ii], gca = (real)nd.GCa[ii];
    const int kick = nd.kick_step[ii];

    // post events detected in the previous window become visible now (NGS:9385-9386)
    int pcnt = nd.post_cnt[ii];
    double last_post = 0.0;
    if (growth && live && leader) {
        int pend = nd.post_pend[ii];
        if (pend > 0) {
            double cs = win.clk(t0 - 1);
            for (int q = 0; q < pend; ++q) {
                nd.post_ring[(size_t)ii * kPostRing + (pcnt % kPostRing)] = cs;
                ++pcnt;
            }
            if (pend > 2) atomicOr(nd.err, (unsigned)kErrPostRing);
            nd.post_cnt[ii] = pcnt;
            nd.post_pend[ii] = 0;
        }
        if (pcnt > 0) last_post = nd.post_ring[(size_t)ii * kPostRing + ((pcnt - 1) % kPostRing)];
    }
    int new_posts = 0;

    const NoiseId sid = ra_noise_id(nd, ii);
    NoiseCursor nc;
    if (MODEL == 1 && nd.ra_poisson)
        nc.open(nd.ra_ev + (size_t)ii * nd.ev_cap, nd.ra_ev_pos[ii]);
    else
        nc.none();

    uint64_t smask = 0, dmask = 0;
    for (int t = t0; t <= t1; ++t) {
        if (kick == t) s.Ge_d += 3.0;  // raiseE(G_TRAINING_KICK), NGS:8933-8940
        int f = ra_state_check(s.Vs, s.Vd, fs, fd);
        const WinRow& row = win.row(t);
        if (MODEL == 0) {
            U4 o = philox4x32((uint32_t)t, sid.local, sid.trial, kStreamRAWhite, sid.k0, sid.k1);
            ra_step_debraband(s, el, gca, u, kPointDistribution[point_index(o.x)], kPointDistribution[point_index(o.y)],
                         kPointDistribution[point_index(o.z)], kPointDistribution[point_index(o.w)],
                         row.e5[kOffDt], row.e5[kOffHalf]);
        } else {
            // HH2_buffer noise checks (Gi_s, Gi_d, Ge_s, Ge_d), as listed
            // (selects, not stores through an index: the state stays in registers)
            nc.take(t, [&](int g, double inc) {
                s.Gi_s = g == 0 ? (real)(s.Gi_s + inc) : s.Gi_s;
                s.Gi_d = g == 1 ? (real)(s.Gi_d + inc) : s.Gi_d;
                s.Ge_s = g == 2 ? (real)(s.Ge_s + inc) : s.Ge_s;
                s.Ge_d = g == 3 ? (real)(s.Ge_d + inc) : s.Ge_d;
            });
            ra_step_runge4<LANES>(s, el, gca, u, row.e5[kOffThird], row.e5[kOffTwoThird], row.e5[kOffDt], gs);
        }
        bool soma = live && leader && (f & 1);
        bool dend = live && leader && (f & 2);
        if (soma) smask |= 1ull << (t - t0);
        if (dend) dmask |= 1ull << (t - t0);
    }
    // after the steps: the spike logs, and the new postsynaptic events (NGS:8956-8967;
    // a window's events become visible in the next window, so each spike is
    // tested against the events of earlier windows only, as in the step loop)
    window_append(nd.soma_log, &nd.log_cnt[0], nd.log_cap, smask, i, t0, nd.err);
    window_append(nd.dend_log, &nd.log_cnt[1], nd.log_cap, dmask, i, t0, nd.err);
    wl_enlist(nd, 0, smask != 0, listed0, i);
    wl_enlist(nd, 1, smask != 0, listed1, i);
    uint64_t pmask = 0;
    if (growth) {
        for (uint64_t m = smask; m; m &= m - 1) {
            const int t = t0 + __ffsll((unsigned long long)m) - 1;
            if (pcnt == 0 || win.clk(t) > last_post + 30.0) pmask |= 1ull << (t - t0);
        }
    }
    for (;;) {
        const bool pred = pmask != 0;
        const uint64_t b = __ballot(pred);
        if (b == 0) break;
        const int lane = __lane_id(), lead = __ffsll((unsigned long long)b) - 1;
        int base = 0;
        if (lane == lead) base = atomicAdd(nd.post_list_cnt, __popcll(b));
        base = __shfl(base, lead);
        if (pred) {
            const int t = t0 + __ffsll((unsigned long long)pmask) - 1;
            const int pos = base + __popcll(b & ((1ull << lane) - 1ull));
            if (pos < kPostListCap)
                nd.post_list[pos] = post_key(nd.ra_net[i], t, i);
            else
                atomicOr(nd.err, (unsigned)kErrPostList);
            ++new_posts;
        }
        pmask &= pmask - 1;
    }
    if (live && leader) {
        nd.ra[0][i] = HVC_WIDE(s, Vs); nd.ra[1][i] = HVC_WIDE(s, Vd); nd.ra[2][i] = HVC_WIDE(s, n);
        nd.ra[3][i] = HVC_WIDE(s, l); nd.ra[4][i] = HVC_WIDE(s, h); nd.ra[5][i] = HVC_WIDE(s, r);
        nd.ra[6][i] = HVC_WIDE(s, c); nd.ra[7][i] = HVC_WIDE(s, Ca);
        nd.ra[8][i] = s.Ge_s; nd.ra[9][i] = s.Gi_s; nd.ra[10][i] = s.Ge_d; nd.ra[11][i] = s.Gi_d;
        nd.ra_flags[i] = fs | (fd << 1);
        nd.soma_mask[i] = smask;
        nd.dend_mask[i] = dmask;
        if (growth) nd.post_pend[i] = new_posts;
        if (MODEL == 1 && nd.ra_poisson) nd.ra_ev_pos[i] = nc.pos;
    }
}

// HVC-I, one lane per neuron (RK4 form)
template <int MODEL>
__device__ void integrate_i_thread(const NetDev& nd, const Win& win, int i, int t0, int t1)
{
    const bool live = i < nd.n_i && (!nd.net_live || nd.net_live[nd.i_net[i]]);
    const int ii = live ? i : 0;
    IState s;
    HVC_LOAD(s, V, nd.in[0][ii]) HVC_LOAD(s, n, nd.in[1][ii]) HVC_LOAD(s, m, nd.in[2][ii])
    HVC_LOAD(s, h, nd.in[3][ii]) HVC_LOAD(s, w, nd.in[4][ii])
    const int listed = nd.wl_on ? nd.wl_mark[2][ii] : 0;
    win.stage(nd);
    // the previous window's sync: HHI raiseE (Ge += G)
    double ge = nd.in[5][ii];
    if (live) {
        double sum;
        if (inbox_take(nd, nd.in_ge_i, i, sum)) ge += sum;
    }
    s.Ge = ge; s.Gi = nd.in[6][ii];
    int flag = nd.i_flag[ii];
    NoiseCursor nc;
    if (nd.i_poisson)
        nc.open(nd.i_ev + (size_t)ii * nd.ev_cap, nd.i_ev_pos[ii]);
    else
        nc.none();
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        // HHI_buffer noise checks (Ge, Gi), as listed
        nc.take(t, [&](int g, double inc) {
            s.Ge = g == 0 ? (real)(s.Ge + inc) : s.Ge;
            s.Gi = g == 1 ? (real)(s.Gi + inc) : s.Gi;
        });
        i_step<MODEL>(s, (real)nd.dt, win.row(t));
        bool fired = live && f;
        if (fired) mask |= 1ull << (t - t0);
    }
    window_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, mask, i, t0, nd.err);
    wl_enlist(nd, 2, mask != 0, listed, i);
    if (live) {
        nd.in[0][i] = HVC_WIDE(s, V); nd.in[1][i] = HVC_WIDE(s, n); nd.in[2][i] = HVC_WIDE(s, m);
        nd.in[3][i] = HVC_WIDE(s, h); nd.in[4][i] = HVC_WIDE(s, w);
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        if (nd.i_poisson) nd.i_ev_pos[i] = nc.pos;
        nd.i_mask[i] = mask;
    }
}

constexpr int kILanes = 5;                    // lanes per HVC-I neuron (DP8)
constexpr int kINeuronsPerWave = 64 / kILanes; // 12 neurons, 4 idle lanes

// HVC-I, five lanes per neuron (DP8 form); one wave holds 12 neurons
__device__ void integrate_i_lanes(const NetDev& nd, const Win& win, int wave, int t0, int t1)
{
    // the four spare lanes 60..63 shadow lanes 55..58 of the last group: same
    // neuron, same roles, results discarded -- their quotients then stay in the
    // range of the real ones and never force the wave to repeat a step
    const int lane = __lane_id();
    const bool spare = lane >= kINeuronsPerWave * kILanes;
    const int role = spare ? lane - kINeuronsPerWave * kILanes : lane % kILanes;
    const int group = spare ? kINeuronsPerWave - 1 : lane / kILanes;
    const int base = group * kILanes;
    const int i = wave * kINeuronsPerWave + group;
    const bool live = (i < nd.n_i) && !spare && (!nd.net_live || nd.net_live[nd.i_net[i]]);
    const int ii = (i < nd.n_i) ? i : 0;
    IState s;
    HVC_LOAD(s, V, nd.in[0][ii]) HVC_LOAD(s, n, nd.in[1][ii]) HVC_LOAD(s, m, nd.in[2][ii])
    HVC_LOAD(s, h, nd.in[3][ii]) HVC_LOAD(s, w, nd.in[4][ii])
    const int listed = nd.wl_on ? nd.wl_mark[2][ii] : 0;
    win.stage(nd);
    // the previous window's sync: HHI raiseE (Ge += G), by the group's first lane
    double ge = nd.in[5][ii];
    if (live && role == 0) {
        double sum;
        if (inbox_take(nd, nd.in_ge_i, i, sum)) ge += sum;
    }
    ge = __shfl(ge, base);
    s.Ge = ge; s.Gi = nd.in[6][ii];
    int flag = nd.i_flag[ii];
    NoiseCursor nc;
    if (nd.i_poisson)
        nc.open(nd.i_ev + (size_t)ii * nd.ev_cap, nd.i_ev_pos[ii]);
    else
        nc.none();
    // low part of this lane's own variable (FP32 form; zero in FP64)
    real xlo = 0;
#ifdef HVC_FP32
    xlo = role == 0 ? s.V_lo : (role == 1 ? s.n_lo : (role == 2 ? s.m_lo : (role == 3 ? s.h_lo : s.w_lo)));
#endif
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        // HHI_buffer noise checks (Ge, Gi), as listed
        nc.take(t, [&](int g, double inc) {
            s.Ge = g == 0 ? (real)(s.Ge + inc) : s.Ge;
            s.Gi = g == 1 ? (real)(s.Gi + inc) : s.Gi;
        });
        i_step_lanes(s, xlo, role, base, (real)nd.dt, win.row(t));
        bool fired = live && f && role == 0;
        if (fired) mask |= 1ull << (t - t0);
    }
    window_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, mask, i, t0, nd.err);
    wl_enlist(nd, 2, mask != 0, listed, i);
    if (live) {
        // each lane writes its own variable, widened with its low part
        const real xr = role == 0 ? s.V : (role == 1 ? s.n : (role == 2 ? s.m : (role == 3 ? s.h : s.w)));
#ifdef HVC_FP32
        nd.in[role][i] = (double)xr + (double)xlo;
#else
        nd.in[role][i] = xr;
#endif
    }
    if (live && role == 0) {
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        if (nd.i_poisson) nd.i_ev_pos[i] = nc.pos;
        nd.i_mask[i] = mask;
    }
}

constexpr int kINeuronsPerQuadWave = 16;       // four lanes per HVC-I neuron

// HVC-I, four lanes per neuron (DP8 form); one wave holds 16 neurons
__device__ void integrate_i_quad(const NetDev& nd, const Win& win, int wave, int t0, int t1)
{
    const int lane = __lane_id();
    const int role = lane & 3;
    const int i = wave * kINeuronsPerQuadWave + (lane >> 2);
    const bool live = i < nd.n_i && (!nd.net_live || nd.net_live[nd.i_net[i]]);
    const int ii = live ? i : 0;
    const GateSlot ga = kIQuadGates[role][0], gb = kIQuadGates[role][1];
    IState s;
    HVC_LOAD(s, V, nd.in[0][ii]) HVC_LOAD(s, n, nd.in[1][ii]) HVC_LOAD(s, m, nd.in[2][ii])
    HVC_LOAD(s, h, nd.in[3][ii]) HVC_LOAD(s, w, nd.in[4][ii])
    const int listed = nd.wl_on ? nd.wl_mark[2][ii] : 0;
    win.stage(nd);
    // the previous window's sync: HHI raiseE (Ge += G), by the quad's first lane
    double ge = nd.in[5][ii];
    if (live && role == 0) {
        double sum;
        if (inbox_take(nd, nd.in_ge_i, i, sum)) ge += sum;
    }
    ge = quad_bcast<0>(ge);
    s.Ge = ge; s.Gi = nd.in[6][ii];
    int flag = nd.i_flag[ii];
    NoiseCursor nc;
    if (nd.i_poisson)
        nc.open(nd.i_ev + (size_t)ii * nd.ev_cap, nd.i_ev_pos[ii]);
    else
        nc.none();
    // low parts of this lane's variable and of w (FP32 form; zero in FP64)
    real xlo = 0, wlo = 0;
#ifdef HVC_FP32
    xlo = role == 0 ? s.V_lo : (role == 1 ? s.n_lo : (role == 2 ? s.m_lo : s.h_lo));
    wlo = s.w_lo;
#endif
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        // HHI_buffer noise checks (Ge, Gi), as listed
        nc.take(t, [&](int g, double inc) {
            s.Ge = g == 0 ? (real)(s.Ge + inc) : s.Ge;
            s.Gi = g == 1 ? (real)(s.Gi + inc) : s.Gi;
        });
        i_step_quad(s, xlo, wlo, role, ga, gb, (real)nd.dt, win.row(t));
        bool fired = live && f && role == 0;
        if (fired) mask |= 1ull << (t - t0);
    }
    window_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, mask, i, t0, nd.err);
    wl_enlist(nd, 2, mask != 0, listed, i);
    if (live) {
        // each lane writes its own variable, widened with its low part; lane 0 also w
        const real xr = role == 0 ? s.V : (role == 1 ? s.n : (role == 2 ? s.m : s.h));
#ifdef HVC_FP32
        nd.in[role][i] = (double)xr + (double)xlo;
        if (role == 0) nd.in[4][i] = (double)s.w + (double)wlo;
#else
        nd.in[role][i] = xr;
        if (role == 0) nd.in[4][i] = s.w;
#endif
    }
    if (live && role == 0) {
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        if (nd.i_poisson) nd.i_ev_pos[i] = nc.pos;
        nd.i_mask[i] = mask;
    }
}

constexpr int kINeuronsPerOctWave = 8;       // eight lanes per HVC-I neuron

// HVC-I, eight lanes per neuron (DP8 form); one wave holds 8 neurons
__device__ void integrate_i_oct(const NetDev& nd, const Win& win, int wave, int t0, int t1)
{
    const int lane = __lane_id();
    const int q = lane & 7;
    const int i = wave * kINeuronsPerOctWave + (lane >> 3);
    const bool live = i < nd.n_i && (!nd.net_live || nd.net_live[nd.i_net[i]]);
    const int ii = live ? i : 0;
    const GateSlot g = kIOctGates[q];
    // the lane's own variable: V (lane 0), n (1, 6), m (2, 5), h (3, 4), w (7)
    const int var = q == 0 ? 0 : (q == 1 || q == 6 ? 1 : (q == 2 || q == 5 ? 2 : (q == 7 ? 4 : 3)));
    IState s;
    HVC_LOAD(s, V, nd.in[0][ii]) HVC_LOAD(s, n, nd.in[1][ii]) HVC_LOAD(s, m, nd.in[2][ii])
    HVC_LOAD(s, h, nd.in[3][ii]) HVC_LOAD(s, w, nd.in[4][ii])
    const int listed = nd.wl_on ? nd.wl_mark[2][ii] : 0;
    win.stage(nd);
    // the previous window's sync: HHI raiseE (Ge += G), by the octet's first lane
    double ge = nd.in[5][ii];
    if (live && q == 0) {
        double sum;
        if (inbox_take(nd, nd.in_ge_i, i, sum)) ge += sum;
    }
    const double ge_lo = quad_bcast<0>(ge), ge_hi = quad_bcast<3>(oct_mirror(ge));
    ge = q < 4 ? ge_lo : ge_hi;
    s.Ge = ge; s.Gi = nd.in[6][ii];
    int flag = nd.i_flag[ii];
    NoiseCursor nc;
    if (nd.i_poisson)
        nc.open(nd.i_ev + (size_t)ii * nd.ev_cap, nd.i_ev_pos[ii]);
    else
        nc.none();
    real xo = var == 0 ? s.V : (var == 1 ? s.n : (var == 2 ? s.m : (var == 3 ? s.h : s.w)));
    real xlo = 0;
#ifdef HVC_FP32
    xlo = var == 0 ? s.V_lo : (var == 1 ? s.n_lo : (var == 2 ? s.m_lo : (var == 3 ? s.h_lo : s.w_lo)));
#endif
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        // HHI_buffer noise checks (Ge, Gi), as listed
        nc.take(t, [&](int gi, double inc) {
            s.Ge = gi == 0 ? (real)(s.Ge + inc) : s.Ge;
            s.Gi = gi == 1 ? (real)(s.Gi + inc) : s.Gi;
        });
        i_step_oct(s, xo, xlo, q, g, (real)nd.dt, win.row(t));
        bool fired = live && f && q == 0;
        if (fired) mask |= 1ull << (t - t0);
    }
    window_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, mask, i, t0, nd.err);
    wl_enlist(nd, 2, mask != 0, listed, i);
    if (live && (q < 4 || q == 7)) {
        // lanes 0-3 write V, n, m, h and lane 7 w, widened with their low parts
#ifdef HVC_FP32
        nd.in[var][i] = (double)xo + (double)xlo;
#else
        nd.in[var][i] = xo;
#endif
    }
    if (live && q == 0) {
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        if (nd.i_poisson) nd.i_ev_pos[i] = nc.pos;
        nd.i_mask[i] = mask;
    }
}

// blocks [0, iblocks) integrate HVC-I (five lanes per neuron, 12 neurons per
// wave; four lanes, 16 per wave; or one thread per neuron), the next rblocks
// HVC-RA (four lanes per neuron, 16 neurons per wave, or one thread per
// neuron); any blocks after them run the previous window's sync-point STDP
// (one row per thread, sa).
// LANES: bit 0 = HVC-I five lanes, bit 1 = HVC-RA four lanes, bit 2 = HVC-I
// four lanes, bit 3 = HVC-I eight lanes (the layouts are template parameters: a
// kernel holds the register allocation of one layout only)
template <int RA_MODEL, int I_MODEL, int LANES>
__global__ __launch_bounds__(256) void k_integrate(NetDev nd, int t0, int t1, int iblocks, int rblocks, StdpArgs sa)
{
    if ((int)blockIdx.x >= iblocks + rblocks) {
        __shared__ uint64_t stdp_slice[kStdpSliceCap];
        __shared__ int stdp_count;
        stdp_block(nd, sa, ((int)blockIdx.x - iblocks - rblocks) * blockDim.x, stdp_slice, &stdp_count);
        return;
    }
    // the window's decay rows and clock, staged in LDS (dynamic shared memory of
    // win_bytes(t1 - t0 + 1) bytes) by Win::stage in the neuron functions
    extern __shared__ double win_lds[];
    const int nsteps = t1 - t0 + 1;
    const Win win{(const WinRow*)win_lds, (const double*)((const WinRow*)win_lds + nsteps), t0, win_lds, nsteps};
    if ((int)blockIdx.x < iblocks) {
        if constexpr ((LANES & 8) != 0)
            integrate_i_oct(nd, win, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), t0, t1);
        else if constexpr ((LANES & 4) != 0)
            integrate_i_quad(nd, win, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), t0, t1);
        else if constexpr ((LANES & 1) != 0)
            integrate_i_lanes(nd, win, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), t0, t1);
        else
            integrate_i_thread<I_MODEL>(nd, win, blockIdx.x * blockDim.x + threadIdx.x, t0, t1);
    } else {
        if constexpr ((LANES & 2) != 0)
            integrate_ra<RA_MODEL, true>(nd, win, (blockIdx.x - iblocks) * (blockDim.x >> 6) + (threadIdx.x >> 6),
                                         t0, t1);
        else
            integrate_ra<RA_MODEL, false>(nd, win, (blockIdx.x - iblocks) * blockDim.x + threadIdx.x, t0, t1);
    }
}

// LDS bytes of k_integrate for a window of n steps
static inline size_t win_bytes(int n) { return (size_t)n * sizeof(WinRow) + (size_t)(n + 1) * sizeof(double); }

}  // namespace p32 / p64

#ifdef HVC_FP32
#define HVC_LAUNCHER(name) name##_f32
#else
#define HVC_LAUNCHER(name) name
#endif

void HVC_LAUNCHER(launch_trace_ra)(hipStream_t st, const NetDev& nd, int n, int n_steps, uint32_t trial, const double* EL,
                     const double* GCa, const double* kick, const double* exc_w, const double* inh_g, double* out,
                     uint8_t* fired)
{
    k_trace_ra<<<blocks_for(n, 64), 64, 0, st>>>(nd, n, n_steps, trial, EL, GCa, kick, exc_w, inh_g, out, fired);
}

void HVC_LAUNCHER(launch_trace_i)(hipStream_t st, const NetDev& nd, int n, int n_steps, int first_trial, uint32_t trial,
                    const double* exc_g, const double* inh_g, double* out, uint8_t* fired)
{
    k_trace_i<<<blocks_for(n, 64), 64, 0, st>>>(nd, n, n_steps, first_trial, trial, exc_g, inh_g, out, fired);
}

void HVC_LAUNCHER(launch_integrate)(hipStream_t st, const NetDev& nd, int t0, int t1, const Tuning& tu,
                                     const StdpArgs& sa, int stdp)
{
    const int bs = tu.integrate_block;
    // HVC-I layout (DP8 only): 0 one thread, 1 five lanes, 2 four lanes, 3 eight lanes per neuron
    const int i_layout = nd.i_model == 0 ? tu.i_lanes : 0;
    const int ra_lanes = (nd.ra_model == 1 && tu.ra_lanes) ? 1 : 0;
    const int i_per_block = i_layout == 1 ? (bs / 64) * kINeuronsPerWave
                            : (i_layout == 2 ? (bs / 64) * kINeuronsPerQuadWave
                                             : (i_layout == 3 ? (bs / 64) * kINeuronsPerOctWave : bs));
    const int ra_per_block = ra_lanes ? (bs / 64) * kRANeuronsPerWave : bs;
    int iblocks = nd.n_i == 0 ? 0 : (nd.n_i + i_per_block - 1) / i_per_block;
    int rblocks = blocks_for(nd.n_ra, ra_per_block);
    const int sblocks = stdp ? blocks_for(nd.n_ra, bs) : 0;
    dim3 grid(iblocks + rblocks + sblocks);
    const size_t lds = win_bytes(t1 - t0 + 1);
#define HVC_K(R, I, L) k_integrate<R, I, L><<<grid, bs, lds, st>>>(nd, t0, t1, iblocks, rblocks, sa)
    if (nd.ra_model == 0) {
        if (nd.i_model == 1) HVC_K(0, 1, 0);
        else if (i_layout == 3) HVC_K(0, 0, 8);
        else if (i_layout == 2) HVC_K(0, 0, 4);
        else if (i_layout == 1) HVC_K(0, 0, 1);
        else HVC_K(0, 0, 0);
    } else if (ra_lanes) {
        if (nd.i_model == 1) HVC_K(1, 1, 2);
        else if (i_layout == 3) HVC_K(1, 0, 10);
        else if (i_layout == 2) HVC_K(1, 0, 6);
        else if (i_layout == 1) HVC_K(1, 0, 3);
        else HVC_K(1, 0, 2);
    } else {
        if (nd.i_model == 1) HVC_K(1, 1, 0);
        else if (i_layout == 3) HVC_K(1, 0, 8);
        else if (i_layout == 2) HVC_K(1, 0, 4);
        else if (i_layout == 1) HVC_K(1, 0, 1);
        else HVC_K(1, 0, 0);
    }
#undef HVC_K
}

#undef HVC_LAUNCHER

}  // namespace hvc
