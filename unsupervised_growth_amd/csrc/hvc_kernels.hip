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
//                                order-dependent LTD / update_synapse scan and,
//                                at the window's sync point, LTP and axon
//                                remodeling (NGS:9375-9554).
//   sync                       : per target, records are ordered by (arrival
//                                step, presynaptic id) -- the reference's
//                                accumulation order -- summed and applied
//                                (raiseI / raiseE / raiseExcWeight, NGS:9314-9366).
// No floating-point atomics: every sum is evaluated in the reference's order,
// so the FP64 form reproduces the CPU results bit for bit.
#include <hip/hip_runtime.h>

#include "hvc_internal.h"
#include "hvc_neuron.hip.h"

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

// One neuron step with the branch-free constant divisions; if any lane of the
// wave met a quotient outside the proven range, the wave repeats the step from
// the saved state with IEEE division (EXACT), which gives the same bits.
__device__ __forceinline__ void ra_step_debraband(RAState& s, double el, double gca, const RAUniform& u, double J1s,
                                                  double J2s, double J1d, double J2d, double E_dt, double E_half)
{
    RAState s0 = s;
    DivGuard bad;
    ra_debraband<false>(bad, s, el, gca, u, J1s, J2s, J1d, J2d, E_dt, E_half);
    if (__any(bad.failed())) {
        s = s0;
        ra_debraband<true>(bad, s, el, gca, u, J1s, J2s, J1d, J2d, E_dt, E_half);
    }
}

__device__ __forceinline__ void ra_step_runge4(RAState& s, double el, double gca, const RAUniform& u, double E_third,
                                               double E_twothird, double E_dt)
{
    RAState s0 = s;
    DivGuard bad;
    ra_runge4<false>(bad, s, el, gca, u, E_third, E_twothird, E_dt);
    if (__any(bad.failed())) {
        s = s0;
        ra_runge4<true>(bad, s, el, gca, u, E_third, E_twothird, E_dt);
    }
}

template <int MODEL>
__device__ __forceinline__ void i_step(IState& s, double dt, const DecayRow& row)
{
    IState s0 = s;
    DivGuard bad;
    if (MODEL == 0) i_dp8<false>(bad, s, dt, row); else i_runge4<false>(bad, s, dt, row);
    if (__any(bad.failed())) {
        s = s0;
        if (MODEL == 0) i_dp8<true>(bad, s, dt, row); else i_runge4<true>(bad, s, dt, row);
    }
}

__device__ __forceinline__ void i_step_lanes(IState& s, int role, int base, double dt, const DecayRow& row)
{
    IState s0 = s;
    DivGuard bad;
    i_dp8_lanes<false>(bad, s, role, base, dt, row);
    if (__any(bad.failed())) {
        s = s0;
        i_dp8_lanes<true>(bad, s, role, base, dt, row);
    }
}

// ---------------------------------------------------------------------------
// single-neuron traces
// ---------------------------------------------------------------------------
__global__ void k_trace_ra(NetDev nd, int n, int n_steps, uint32_t trial, const double* EL, const double* GCa,
                           const double* kick, const double* exc_w, const double* inh_g, double* out, uint8_t* fired_out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    RAUniform u{nd.sq_sig_s, nd.sq_sig_d, nd.mu_term_s, nd.mu_term_d, nd.dt};
    RAState s{-79.97619025, -79.97268759, 0.01101284, 0.0, 0.9932845, 0.00055429, 0.00000261762353, 0.01689572,
              0.0, 0.0, 0.0, 0.0};
    int fs = 0, fd = 0;
    double el = EL[k], gca = GCa[k];
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
        o[0] = s.Vs; o[1] = s.Vd; o[2] = s.n; o[3] = s.l; o[4] = s.h; o[5] = s.r;
        o[6] = s.c; o[7] = s.Ca; o[8] = s.Ge_s; o[9] = s.Gi_s; o[10] = s.Ge_d; o[11] = s.Gi_d;
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
    IState s{-66.0, 0.125, 0.0, 0.99, 0.0, 0.0, 0.0};
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
        o[0] = s.V; o[1] = s.n; o[2] = s.m; o[3] = s.h; o[4] = s.w; o[5] = s.Ge; o[6] = s.Gi;
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
            i_step<1>(s, nd.dt, rows[t]);
        fired_out[idx] = (uint8_t)f;
        double ig = inh_g[idx];
        if (ig != 0.0) s.Gi += ig;
        double eg = exc_g[idx];
        if (eg != 0.0) s.Ge += eg;
        dump(t);
    }
}

// ---------------------------------------------------------------------------
// trial start: setToRest_afterEpoch (NGS:4591-4621), HH2_test::set_to_rest
// (HH2_test.cpp:191-242), HHI_buffer::set_to_rest (HHI_buffer.cpp:178-200)
// ---------------------------------------------------------------------------
__global__ void k_reset(NetDev nd, uint32_t trial, int first_trial)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd.n_ra) {
        nd.ra[0][i] = -79.97619025; nd.ra[1][i] = -79.97268759; nd.ra[2][i] = 0.01101284; nd.ra[3][i] = 0.0;
        nd.ra[4][i] = 0.9932845; nd.ra[5][i] = 0.00055429; nd.ra[6][i] = 0.00000261762353; nd.ra[7][i] = 0.01689572;
        nd.ra[8][i] = 0.0; nd.ra[9][i] = 0.0; nd.ra[10][i] = 0.0; nd.ra[11][i] = 0.0;
        nd.ra_flags[i] = 0;
        nd.ra_draw[i] = 0;
        if (nd.ra_poisson) {
            PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamRAPoisson, (uint32_t)i, 0u};
            for (int q = 0; q < 4; ++q) nd.ra_noise[(size_t)q * nd.n_ra + i] = 1000.0 * ps.spike_time(100.0);
            nd.ra_draw[i] = ps.draw;
        }
        nd.post_cnt[i] = 0;
        nd.post_pend[i] = 0;
        nd.st_count[i] = 0;
        nd.fl_ra2i.head[i] = 0; nd.fl_ra2i.count[i] = 0;
        nd.fl_rr.head[i] = 0; nd.fl_rr.count[i] = 0;
        nd.in_gi_ra.cnt[i] = 0;
        nd.in_ge_ra.cnt[i] = 0;
    }
    if (i < nd.n_i) {
        nd.in[0][i] = -66.0; nd.in[1][i] = 0.125; nd.in[2][i] = 0.0; nd.in[3][i] = 0.99; nd.in[4][i] = 0.0;
        nd.in[5][i] = 0.0; nd.in[6][i] = 0.0;
        nd.i_flag[i] = 0;
        PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamIPoisson, (uint32_t)i, 0u};
        if (nd.i_poisson) {
            double init_dt = first_trial ? 0.0 : nd.dt;
            nd.i_noise[i] = i_init_noise(ps, init_dt);
            nd.i_noise[(size_t)nd.n_i + i] = i_init_noise(ps, init_dt);
        }
        nd.i_draw[i] = ps.draw;
        nd.fl_i2ra.head[i] = 0; nd.fl_i2ra.count[i] = 0;
        nd.in_ge_i.cnt[i] = 0;
    }
    if (i == 0) {
        nd.log_cnt[0] = 0; nd.log_cnt[1] = 0; nd.log_cnt[2] = 0;
        *nd.post_list_cnt = 0;
        *nd.blocks_done = 0;
        *nd.err = 0;
        nd.warn[0] = 0; nd.warn[1] = 0; nd.warn[2] = 0;
    }
}

__global__ void k_clear_pre(int* pre_cnt, int E)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) pre_cnt[e] = 0;
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

// ---------------------------------------------------------------------------
// integrate: HVC-RA neurons (NGS:8931-9216 minus delivery) and HVC-I neurons
// (NGS:9221-9286 minus delivery) through the steps t0..t1 of one window, in one
// launch.  HVC-I blocks come first: with five lanes per neuron they carry the
// longer per-step work, and the RA blocks run beside them.
// ---------------------------------------------------------------------------
template <int MODEL>
__device__ void integrate_ra_thread(const NetDev& nd, int i, int t0, int t1, uint32_t trial, int growth)
{
    const bool live = i < nd.n_ra;
    const int ii = live ? i : 0;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    RAUniform u{nd.sq_sig_s, nd.sq_sig_d, nd.mu_term_s, nd.mu_term_d, nd.dt};
    RAState s;
    s.Vs = nd.ra[0][ii]; s.Vd = nd.ra[1][ii]; s.n = nd.ra[2][ii]; s.l = nd.ra[3][ii];
    s.h = nd.ra[4][ii]; s.r = nd.ra[5][ii]; s.c = nd.ra[6][ii]; s.Ca = nd.ra[7][ii];
    s.Ge_s = nd.ra[8][ii]; s.Gi_s = nd.ra[9][ii]; s.Ge_d = nd.ra[10][ii]; s.Gi_d = nd.ra[11][ii];
    int flags = nd.ra_flags[ii];
    int fs = flags & 1, fd = (flags >> 1) & 1;
    const double el = nd.EL[ii], gca = nd.GCa[ii];
    const int kick = nd.kick_step[ii];

    // post events detected in the previous window become visible now (NGS:9385-9386)
    int pcnt = nd.post_cnt[ii];
    double last_post = 0.0;
    if (growth && live) {
        int pend = nd.post_pend[ii];
        if (pend > 0) {
            double cs = nd.clock[t0 - 1];
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

    PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamRAPoisson, (uint32_t)ii, nd.ra_draw[ii]};
    double nes = 0, nis = 0, ned = 0, nid = 0;
    if (MODEL == 1 && nd.ra_poisson) {
        nes = nd.ra_noise[ii]; nis = nd.ra_noise[(size_t)nd.n_ra + ii];
        ned = nd.ra_noise[2 * (size_t)nd.n_ra + ii]; nid = nd.ra_noise[3 * (size_t)nd.n_ra + ii];
    }

    uint64_t smask = 0, dmask = 0;
    for (int t = t0; t <= t1; ++t) {
        if (kick == t) s.Ge_d += 3.0;  // raiseE(G_TRAINING_KICK), NGS:8933-8940
        int f = ra_state_check(s.Vs, s.Vd, fs, fd);
        const DecayRow& row = rows[t];
        if (MODEL == 0) {
            U4 o = philox4x32((uint32_t)t, (uint32_t)ii, trial, kStreamRAWhite, nd.seed_lo, nd.seed_hi);
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
        bool soma = live && (f & 1);
        bool dend = live && (f & 2);
        if (soma) smask |= 1ull << (t - t0);
        if (dend) dmask |= 1ull << (t - t0);
        wave_append(nd.soma_log, &nd.log_cnt[0], nd.log_cap, soma, i, t, nd.err);
        wave_append(nd.dend_log, &nd.log_cnt[1], nd.log_cap, dend, i, t, nd.err);
        if (growth && soma) {
            // new postsynaptic event (NGS:8956-8967)
            double it = nd.clock[t];
            if (pcnt == 0 || it > last_post + 30.0) {
                int pos = atomicAdd(nd.post_list_cnt, 1);
                if (pos < kPostListCap)
                    nd.post_list[pos] = ((uint64_t)(uint32_t)t << 32) | (uint32_t)i;
                else
                    atomicOr(nd.err, (unsigned)kErrPostList);
                ++new_posts;
            }
        }
    }
    if (live) {
        nd.ra[0][i] = s.Vs; nd.ra[1][i] = s.Vd; nd.ra[2][i] = s.n; nd.ra[3][i] = s.l;
        nd.ra[4][i] = s.h; nd.ra[5][i] = s.r; nd.ra[6][i] = s.c; nd.ra[7][i] = s.Ca;
        nd.ra[8][i] = s.Ge_s; nd.ra[9][i] = s.Gi_s; nd.ra[10][i] = s.Ge_d; nd.ra[11][i] = s.Gi_d;
        nd.ra_flags[i] = fs | (fd << 1);
        nd.soma_mask[i] = smask;
        nd.dend_mask[i] = dmask;
        if (growth) nd.post_pend[i] = new_posts;
        if (MODEL == 1 && nd.ra_poisson) {
            nd.ra_noise[i] = nes; nd.ra_noise[(size_t)nd.n_ra + i] = nis;
            nd.ra_noise[2 * (size_t)nd.n_ra + i] = ned; nd.ra_noise[3 * (size_t)nd.n_ra + i] = nid;
            nd.ra_draw[i] = ps.draw;
        }
    }
}

// HVC-I, one lane per neuron (RK4 form)
__device__ void integrate_i_thread(const NetDev& nd, int i, int t0, int t1, uint32_t trial)
{
    const bool live = i < nd.n_i;
    const int ii = live ? i : 0;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    IState s{nd.in[0][ii], nd.in[1][ii], nd.in[2][ii], nd.in[3][ii], nd.in[4][ii], nd.in[5][ii], nd.in[6][ii]};
    int flag = nd.i_flag[ii];
    PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamIPoisson, (uint32_t)ii, nd.i_draw[ii]};
    double ne = nd.i_noise[ii], ni = nd.i_noise[(size_t)nd.n_i + ii];
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        if (nd.i_poisson) {
            double time = nd.clock[t - 1];
            i_noise_check(ps, time, nd.dt, s.Ge, ne);
            i_noise_check(ps, time, nd.dt, s.Gi, ni);
        }
        i_step<1>(s, nd.dt, rows[t]);
        bool fired = live && f;
        if (fired) mask |= 1ull << (t - t0);
        wave_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, fired, i, t, nd.err);
    }
    if (live) {
        nd.in[0][i] = s.V; nd.in[1][i] = s.n; nd.in[2][i] = s.m; nd.in[3][i] = s.h; nd.in[4][i] = s.w;
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        nd.i_noise[i] = ne; nd.i_noise[(size_t)nd.n_i + i] = ni;
        nd.i_draw[i] = ps.draw;
        nd.i_mask[i] = mask;
    }
}

constexpr int kILanes = 5;                    // lanes per HVC-I neuron (DP8)
constexpr int kINeuronsPerWave = 64 / kILanes; // 12 neurons, 4 idle lanes

// HVC-I, five lanes per neuron (DP8 form); one wave holds 12 neurons
__device__ void integrate_i_lanes(const NetDev& nd, int wave, int t0, int t1, uint32_t trial)
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
    const bool live = (i < nd.n_i) && !spare;
    const int ii = (i < nd.n_i) ? i : 0;
    const DecayRow* rows = (const DecayRow*)nd.decay;
    IState s{nd.in[0][ii], nd.in[1][ii], nd.in[2][ii], nd.in[3][ii], nd.in[4][ii], nd.in[5][ii], nd.in[6][ii]};
    int flag = nd.i_flag[ii];
    PoissonStream ps{nd.seed_lo, nd.seed_hi, trial, kStreamIPoisson, (uint32_t)ii, nd.i_draw[ii]};
    double ne = nd.i_noise[ii], ni = nd.i_noise[(size_t)nd.n_i + ii];
    uint64_t mask = 0;
    for (int t = t0; t <= t1; ++t) {
        int f = i_state_check(s.V, flag);
        if (nd.i_poisson) {
            double time = nd.clock[t - 1];
            i_noise_check(ps, time, nd.dt, s.Ge, ne);
            i_noise_check(ps, time, nd.dt, s.Gi, ni);
        }
        i_step_lanes(s, role, base, nd.dt, rows[t]);
        bool fired = live && f && role == 0;
        if (fired) mask |= 1ull << (t - t0);
        wave_append(nd.i_log, &nd.log_cnt[2], nd.log_cap, fired, i, t, nd.err);
    }
    if (live && role == 0) {
        nd.in[0][i] = s.V; nd.in[1][i] = s.n; nd.in[2][i] = s.m; nd.in[3][i] = s.h; nd.in[4][i] = s.w;
        nd.in[5][i] = s.Ge; nd.in[6][i] = s.Gi;
        nd.i_flag[i] = flag;
        nd.i_noise[i] = ne; nd.i_noise[(size_t)nd.n_i + i] = ni;
        nd.i_draw[i] = ps.draw;
        nd.i_mask[i] = mask;
    }
}

// the last block to finish sorts the window's post-event list by (step, id)
__device__ void sort_post_list_if_last(const NetDev& nd)
{
    __shared__ bool last;
    __shared__ uint64_t keys[kPostListCap];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(nd.blocks_done, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!last) return;
    __threadfence();
    int n = atomicAdd(nd.post_list_cnt, 0);
    if (n > kPostListCap) n = kPostListCap;
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = threadIdx.x; k < np2; k += blockDim.x) keys[k] = (k < n) ? nd.post_list[k] : ~0ull;
    __syncthreads();
    for (int size = 2; size <= np2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < np2; k += blockDim.x) {
                int p = k ^ stride;
                if (p > k) {
                    bool up = ((k & size) == 0);
                    uint64_t a = keys[k], b = keys[p];
                    if ((a > b) == up) { keys[k] = b; keys[p] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) nd.post_list[k] = keys[k];
    if (threadIdx.x == 0) *nd.blocks_done = 0;
}

// one 64-lane wave per block; blocks [0, iblocks) integrate HVC-I, the rest HVC-RA
template <int RA_MODEL, int I_MODEL>
__global__ __launch_bounds__(64) void k_integrate(NetDev nd, int t0, int t1, uint32_t trial, int growth, int iblocks)
{
    if ((int)blockIdx.x < iblocks) {
        if (I_MODEL == 0)
            integrate_i_lanes(nd, blockIdx.x, t0, t1, trial);
        else
            integrate_i_thread(nd, blockIdx.x * 64 + threadIdx.x, t0, t1, trial);
    } else {
        integrate_ra_thread<RA_MODEL>(nd, (blockIdx.x - iblocks) * 64 + threadIdx.x, t0, t1, trial, growth);
    }
    if (growth) sort_post_list_if_last(nd);
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

// first step a >= lo with clock[a] >= T (the drain test internal_time >= delivery_time)
__device__ __forceinline__ int arrival_step(const double* clock, int lo, double T)
{
    int a = lo;
    while (clock[a] < T) ++a;
    return a;
}

// Fixed-weight delay queue of one presynaptic neuron (RA->I NGS:8971-8986 +
// 9039-9062, I->RA NGS:9235-9284, chain-test RA->RA NGS:13222-13250): each
// in-flight spike walks its delay-sorted edges; pops arriving by t1 go to inboxes.
__device__ void push_fixed(const NetDev& nd, int pre, int row_begin, int row_end, const int* edge_ids,
                           const int* tgt_arr, const double* w_arr, const double* d_arr, const InFlight& fl,
                           uint64_t new_mask, int t0, int t1, const Inbox& ib)
{
    int head = fl.head[pre], count = fl.count[pre];
    int* fstep = fl.step + (size_t)pre * kMaxInflight;
    int* fcur = fl.cursor + (size_t)pre * kMaxInflight;
    int deg = row_end - row_begin;
    // enqueue this window's spikes (in step order)
    while (new_mask) {
        int b = __ffsll((unsigned long long)new_mask) - 1;
        new_mask &= new_mask - 1;
        if (deg == 0) continue;
        if (count == kMaxInflight) {
            atomicOr(nd.err, (unsigned)kErrInflight);
            break;
        }
        int slot = (head + count) % kMaxInflight;
        fstep[slot] = t0 + b;
        fcur[slot] = 0;
        ++count;
    }
    const double tend = nd.clock[t1];
    for (int q = 0; q < count; ++q) {
        int slot = (head + q) % kMaxInflight;
        int s = fstep[slot];
        int c = fcur[slot];
        const double cs = nd.clock[s];
        int lo = s > t0 ? s : t0;
        while (c < deg) {
            int e = edge_ids ? edge_ids[row_begin + c] : row_begin + c;
            double T = cs + d_arr[e];
            if (T > tend) break;
            int a = arrival_step(nd.clock, lo, T);
            lo = a;
            inbox_put(ib, tgt_arr[e], a, pre, w_arr[e], nd.err);
            ++c;
        }
        fcur[slot] = c;
    }
    // retire finished spikes from the head (all share the same edges, so FIFO)
    while (count > 0 && fcur[head] >= deg) {
        head = (head + 1) % kMaxInflight;
        --count;
    }
    fl.head[pre] = head;
    fl.count[pre] = count;
}

// ---------------------------------------------------------------------------
// STDP and synapse-state machine (NGS:16546-16652), row-owned by one thread
// ---------------------------------------------------------------------------
__device__ __forceinline__ int edge_of(const NetDev& nd, int i, int j)
{
    int lo = nd.rr_ptr[i], hi = nd.rr_ptr[i + 1] - 1;
    if (nd.dense) return lo + j;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int v = nd.rr_tgt[mid];
        if (v == j) return mid;
        if (v < j) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

__device__ __forceinline__ double ltp(const hvc_synaptic_params& p, double w, double dt, int* warn)
{
    if (dt <= p.T_0) { atomicAdd(&warn[0], 1); return w; }
    if (dt <= p.T_0 + p.T_P)
        w = w + p.A_P * (dt - p.T_0) / p.T_P;
    else
        w = w + p.A_P * dexp(-(dt - p.T_0 - p.T_P) / p.TAU_P);
    if (w > p.WEIGHT_MAX) w = p.WEIGHT_MAX;
    if (w < 0) w = 0;
    return w;
}

__device__ __forceinline__ double ltd(const hvc_synaptic_params& p, double w, double dt, int* warn)
{
    if (dt > p.T_0) { atomicAdd(&warn[1], 1); return w; }
    if (dt >= p.T_0 - p.T_D)
        w = w + w * p.A_D * (dt - p.T_0) / p.T_D;
    else
        w = w - w * p.A_D * dexp((dt - p.T_0 + p.T_D) / p.TAU_D);
    if (w < 0) w = 0;
    return w;
}

__device__ __forceinline__ bool list_erase(int* list, int* cnt, int j)
{
    int n = *cnt;
    for (int k = 0; k < n; ++k) {
        if (list[k] == j) {
            for (int q = k; q < n - 1; ++q) list[q] = list[q + 1];
            *cnt = n - 1;
            return true;
        }
    }
    return false;
}

// NGS:16546-16592
__device__ void update_synapse(const NetDev& nd, int i, int e)
{
    const hvc_synaptic_params& p = nd.syn;
    int j = nd.rr_tgt[e];
    double w = nd.w[e];
    int* alist = nd.act_list + nd.rr_ptr[i];
    int* slist = nd.sup_list + (size_t)i * p.Nss;
    if ((w >= p.ACTIVATION_THRESHOLD) && (nd.act[e] == 0) && (nd.remodeled[i] == 0)) {
        nd.act[e] = 1;
        alist[nd.act_cnt[i]++] = j;
    }
    if ((w >= p.SUPERSYNAPSE_THRESHOLD) && (nd.sup[e] == 0) && (nd.sup_cnt[i] < p.Nss)) {
        nd.sup[e] = 1;
        slist[nd.sup_cnt[i]++] = j;
    }
    if ((w < p.SUPERSYNAPSE_THRESHOLD) && (nd.sup[e] == 1)) {
        nd.sup[e] = 0;
        nd.remodeled[i] = 0;
        if (!list_erase(slist, &nd.sup_cnt[i], j)) atomicAdd(&nd.warn[2], 1);
    }
    if ((w < p.ACTIVATION_THRESHOLD) && (nd.act[e] == 1)) {
        nd.act[e] = 0;
        if (!list_erase(alist, &nd.act_cnt[i], j)) atomicAdd(&nd.warn[2], 1);
    }
}

__device__ void update_row(const NetDev& nd, int i)
{
    for (int e = nd.rr_ptr[i]; e < nd.rr_ptr[i + 1]; ++e) update_synapse(nd, i, e);
}

// NGS:16363-16415
__device__ void axon_remodeling(const NetDev& nd, int i)
{
    int base = nd.rr_ptr[i];
    int* slist = nd.sup_list + (size_t)i * nd.syn.Nss;
    for (int k = 0; k < nd.act_cnt[i]; ++k) nd.act[edge_of(nd, i, nd.act_list[base + k])] = 0;
    for (int k = 0; k < nd.sup_cnt[i]; ++k) nd.act[edge_of(nd, i, slist[k])] = 1;
    for (int k = 0; k < nd.sup_cnt[i]; ++k) nd.act_list[base + k] = slist[k];
    nd.act_cnt[i] = nd.sup_cnt[i];
    nd.remodeled[i] = 1;
}

// LTD against the target's post events not older than STDP_WINDOW (NGS:9111-9180)
__device__ void ltd_on_delivery(const NetDev& nd, int i, int e, int tgt, double T, double now)
{
    int pc = nd.post_cnt[tgt];
    if (pc == 0) return;
    int first = pc > kPostRing ? pc - kPostRing : 0;
    const double* ring = nd.post_ring + (size_t)tgt * kPostRing;
    double lim = now - 100.0;
    int k = first;
    while (k < pc && ring[k % kPostRing] < lim) ++k;
    for (; k < pc; ++k) {
        double dt = ring[k % kPostRing] - T;
        nd.w[e] = ltd(nd.syn, nd.w[e], dt, nd.warn);
        update_synapse(nd, i, e);
    }
}

// next item of a stream: (T, target, edge); returns false when exhausted
__device__ __forceinline__ bool stream_peek(const NetDev& nd, int i, int sidx, double& T, int& tgt, int& e)
{
    int* cur = nd.st_cursor + (size_t)i * kMaxStreams + sidx;
    int n = nd.st_n[(size_t)i * kMaxStreams + sidx];
    int mode = nd.st_mode[(size_t)i * kMaxStreams + sidx];
    int s = nd.st_step[(size_t)i * kMaxStreams + sidx];
    const double cs = nd.clock[s];
    int c = *cur;
    if (mode == 0) {
        int base = nd.rr_ptr[i];
        while (c < n) {
            int ee = nd.rr_perm[base + c];
            if (nd.rr_tgt[ee] != i) {
                e = ee;
                tgt = nd.rr_tgt[ee];
                T = cs + nd.rr_d[ee];
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

// Growth-mode RA->RA row of presynaptic neuron i for the window t0..t1
// (NGS:8991-9030 enqueue, 9069-9191 drain with LTD, 9375-9554 at the sync).
__device__ void rr_row(const NetDev& nd, int i, int t0, int t1)
{
    const int Nss = nd.syn.Nss;
    uint64_t mask = nd.soma_mask[i];
    int sc = nd.st_count[i];
    int* st_step = nd.st_step + (size_t)i * kMaxStreams;
    int* st_mode = nd.st_mode + (size_t)i * kMaxStreams;
    int* st_cur = nd.st_cursor + (size_t)i * kMaxStreams;
    int* st_n = nd.st_n + (size_t)i * kMaxStreams;
    for (int t = t0; t <= t1; ++t) {
        if ((mask >> (t - t0)) & 1ull) {
            if (sc == kMaxStreams) {
                atomicOr(nd.err, (unsigned)kErrStreams);
            } else {
                st_step[sc] = t;
                st_cur[sc] = 0;
                if (nd.sup_cnt[i] == Nss) {
                    // saturated: deliver to supersynaptic targets only, queue sorted by (T, target)
                    st_mode[sc] = 1;
                    int n = nd.sup_cnt[i];
                    int* items = nd.st_items + ((size_t)i * kMaxStreams + sc) * Nss;
                    const double cs = nd.clock[t];
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
                    st_n[sc] = n;
                } else {
                    st_mode[sc] = 0;
                    st_n[sc] = nd.rr_ptr[i + 1] - nd.rr_ptr[i];
                }
                ++sc;
            }
        }
        nd.st_count[i] = sc;
        const double now = nd.clock[t];
        // drain in (delivery_time, target) order across all in-flight spikes
        for (;;) {
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
            // conductance of active synapses (NGS:9080-9084)
            if (nd.act[be] == 1) inbox_put(nd.in_ge_ra, btg, t, i, nd.w[be], nd.err);
            // new presynaptic event (NGS:9091-9102)
            int pc = nd.pre_cnt[be];
            double* ring = nd.pre_ring + (size_t)be * kPreRing;
            bool new_pre = false;
            if (pc == 0 || bT > ring[(pc - 1) % kPreRing] + 30.0) {
                ring[pc % kPreRing] = bT;
                nd.pre_cnt[be] = pc + 1;
                new_pre = true;
            }
            if (new_pre) {
                if (nd.sup_cnt[i] == Nss) {
                    if (nd.sup[be] == 1) ltd_on_delivery(nd, i, be, btg, bT, now);
                    if (nd.sup_cnt[i] < Nss) update_row(nd, i);
                } else {
                    ltd_on_delivery(nd, i, be, btg, bT, now);
                }
            }
        }
        // drop exhausted streams, keeping order
        int w = 0;
        for (int q = 0; q < sc; ++q) {
            double T;
            int tg, e;
            if (stream_peek(nd, i, q, T, tg, e)) {
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

    // sync at step t1: LTP / LTD against this window's new post events (NGS:9375-9542)
    int np = *nd.post_list_cnt;
    if (np > kPostListCap) np = kPostListCap;
    if (np == 0) return;
    const double now = nd.clock[t1];
    const bool saturated = (nd.sup_cnt[i] == Nss);
    const int* slist = nd.sup_list + (size_t)i * Nss;
    for (int k = 0; k < np; ++k) {
        int post = (int)(uint32_t)(nd.post_list[k] & 0xffffffffull);
        if (post == i) continue;
        if (saturated) {
            bool found = false;
            for (int q = 0; q < nd.sup_cnt[i]; ++q)
                if (slist[q] == post) { found = true; break; }
            if (!found) continue;
        }
        int e = edge_of(nd, i, post);
        if (e < 0) continue;
        int pc = nd.pre_cnt[e];
        if (pc == 0) continue;
        const double* ring = nd.pre_ring + (size_t)e * kPreRing;
        int first = pc > kPreRing ? pc - kPreRing : 0;
        double lim = now - 100.0;
        int q = first;
        while (q < pc && ring[q % kPreRing] < lim) ++q;
        for (; q < pc; ++q) {
            double dt = now - ring[q % kPreRing];
            if (dt <= nd.syn.T_0)
                nd.w[e] = ltd(nd.syn, nd.w[e], dt, nd.warn);
            else
                nd.w[e] = ltp(nd.syn, nd.w[e], dt, nd.warn);
            update_synapse(nd, i, e);
        }
    }
    // NGS:9548-9554
    if ((nd.sup_cnt[i] == Nss) && (nd.remodeled[i] == 0)) axon_remodeling(nd, i);
}

// ---------------------------------------------------------------------------
// deliver: one thread per presynaptic neuron and role
//   role 0: RA->RA rows (growth) or chain-test RA->RA push
//   role 1: RA->I push        role 2: I->RA push
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_deliver(NetDev nd, int t0, int t1, int growth, int blocks_ra)
{
    int b = blockIdx.x;
    int role = b / blocks_ra;
    int local = (b % blocks_ra) * blockDim.x + threadIdx.x;
    if (role == 0) {
        if (local >= nd.n_ra) return;
        if (growth)
            rr_row(nd, local, t0, t1);
        else
            push_fixed(nd, local, nd.rr_ptr[local], nd.rr_ptr[local] + nd.act_cnt[local], nd.chain_e, nd.rr_tgt,
                       nd.w, nd.rr_d, nd.fl_rr, nd.soma_mask[local], t0, t1, nd.in_ge_ra);
    } else if (role == 1) {
        if (local >= nd.n_ra) return;
        push_fixed(nd, local, nd.ra2i.ptr[local], nd.ra2i.ptr[local + 1], nullptr, nd.ra2i.tgt, nd.ra2i.w,
                   nd.ra2i.d, nd.fl_ra2i, nd.soma_mask[local], t0, t1, nd.in_ge_i);
    } else {
        if (local >= nd.n_i) return;
        push_fixed(nd, local, nd.i2ra.ptr[local], nd.i2ra.ptr[local + 1], nullptr, nd.i2ra.tgt, nd.i2ra.w,
                   nd.i2ra.d, nd.fl_i2ra, nd.i_mask[local], t0, t1, nd.in_gi_ra);
    }
}

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

// sync: apply the window's conductance updates (NGS:9309-9366)
__global__ __launch_bounds__(256) void k_sync(NetDev nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *nd.post_list_cnt = 0;  // the window's post-event list has been consumed
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
    if (i >= nd.n_ra) return;
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
    if (i >= nd.n_ra) return;
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
static inline int blocks_for(int n, int bs) { return n > 0 ? (n + bs - 1) / bs : 1; }

void launch_decay_table(hipStream_t st, const double* clock, const double* offsets, int num_steps, void* rows)
{
    k_decay_table<<<blocks_for(num_steps, 128), 128, 0, st>>>(clock, offsets, num_steps, (DecayRow*)rows);
}

void launch_trace_ra(hipStream_t st, const NetDev& nd, int n, int n_steps, uint32_t trial, const double* EL,
                     const double* GCa, const double* kick, const double* exc_w, const double* inh_g, double* out,
                     uint8_t* fired)
{
    k_trace_ra<<<blocks_for(n, 64), 64, 0, st>>>(nd, n, n_steps, trial, EL, GCa, kick, exc_w, inh_g, out, fired);
}

void launch_trace_i(hipStream_t st, const NetDev& nd, int n, int n_steps, int first_trial, uint32_t trial,
                    const double* exc_g, const double* inh_g, double* out, uint8_t* fired)
{
    k_trace_i<<<blocks_for(n, 64), 64, 0, st>>>(nd, n, n_steps, first_trial, trial, exc_g, inh_g, out, fired);
}

void launch_reset(hipStream_t st, const NetDev& nd, uint32_t trial, int first_trial, int E)
{
    int n = nd.n_ra > nd.n_i ? nd.n_ra : nd.n_i;
    k_reset<<<blocks_for(n, 256), 256, 0, st>>>(nd, trial, first_trial);
    if (E > 0) k_clear_pre<<<blocks_for(E, 256), 256, 0, st>>>(nd.pre_cnt, E);
}

void launch_integrate(hipStream_t st, const NetDev& nd, int t0, int t1, uint32_t trial, int growth)
{
    int iblocks = nd.n_i == 0 ? 0 : (nd.i_model == 0 ? (nd.n_i + kINeuronsPerWave - 1) / kINeuronsPerWave
                                                     : blocks_for(nd.n_i, 64));
    int rblocks = blocks_for(nd.n_ra, 64);
    dim3 grid(iblocks + rblocks);
    if (nd.ra_model == 0 && nd.i_model == 0)
        k_integrate<0, 0><<<grid, 64, 0, st>>>(nd, t0, t1, trial, growth, iblocks);
    else if (nd.ra_model == 0)
        k_integrate<0, 1><<<grid, 64, 0, st>>>(nd, t0, t1, trial, growth, iblocks);
    else if (nd.i_model == 0)
        k_integrate<1, 0><<<grid, 64, 0, st>>>(nd, t0, t1, trial, growth, iblocks);
    else
        k_integrate<1, 1><<<grid, 64, 0, st>>>(nd, t0, t1, trial, growth, iblocks);
}

void launch_deliver(hipStream_t st, const NetDev& nd, int t0, int t1, int growth)
{
    int nmax = nd.n_ra > nd.n_i ? nd.n_ra : nd.n_i;
    int blocks_ra = blocks_for(nmax, 64);
    k_deliver<<<3 * blocks_ra, 64, 0, st>>>(nd, t0, t1, growth, blocks_ra);
}

void launch_sync(hipStream_t st, const NetDev& nd)
{
    int n = nd.n_ra > nd.n_i ? nd.n_ra : nd.n_i;
    k_sync<<<blocks_for(n, 256), 256, 0, st>>>(nd);
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
