// hvc_stdp.hip.h -- the RA->RA synapse-state machine and the sync-point STDP
// of a growth trial, shared by both precision forms (it is FP64 in both).
//
// A row (presynaptic HVC-RA neuron i and its candidate edges) is owned by one
// thread: update_synapse keeps the reference's list order, so a row's updates
// run in the reference's order.  The sync-point STDP of window w (rr_sync)
// reads the window's post events, the rows' pre-event rings and weights; the
// neuron integration of window w + 1 touches none of these, so k_integrate of
// window w + 1 runs it in blocks of its own (stdp_block), and the delivery of
// window w + 1, which needs its results, is launched after that kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "hvc_internal.h"
#include "hvc_math.hip.h"

namespace hvc {

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

// one pass through branch k of the synapse-state machine (when counting)
__device__ __forceinline__ void count_hit(const NetDev& nd, int k)
{
    if (nd.hits) atomicAdd(nd.hits + k, 1);
}

// warning counters of the network that owns HVC-RA row i
__device__ __forceinline__ int* row_warn(const NetDev& nd, int i) { return nd.warn + 3 * nd.ra_net[i]; }

// first index of the sorted post-event keys that is >= key
__device__ __forceinline__ int post_lower_bound(const uint64_t* keys, int n, uint64_t key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// NGS:16546-16592
static __device__ void update_synapse(const NetDev& nd, int i, int e)
{
    const hvc_synaptic_params& p = nd.syn;
    int j = nd.rr_tgt[e];
    double w = nd.w[e];
    int* alist = nd.act_list + nd.rr_ptr[i];
    int* slist = nd.sup_list + (size_t)i * p.Nss;
    if ((w >= p.ACTIVATION_THRESHOLD) && (nd.act[e] == 0) && (nd.remodeled[i] == 0)) {
        nd.act[e] = 1;
        alist[nd.act_cnt[i]++] = j;
        count_hit(nd, kHitActivate);
    }
    if ((w >= p.SUPERSYNAPSE_THRESHOLD) && (nd.sup[e] == 0) && (nd.sup_cnt[i] < p.Nss)) {
        nd.sup[e] = 1;
        slist[nd.sup_cnt[i]++] = j;
        count_hit(nd, kHitSuper);
    }
    if ((w < p.SUPERSYNAPSE_THRESHOLD) && (nd.sup[e] == 1)) {
        nd.sup[e] = 0;
        nd.remodeled[i] = 0;
        if (!list_erase(slist, &nd.sup_cnt[i], j)) atomicAdd(&row_warn(nd, i)[2], 1);
        count_hit(nd, kHitUnsuper);
    }
    if ((w < p.ACTIVATION_THRESHOLD) && (nd.act[e] == 1)) {
        nd.act[e] = 0;
        if (!list_erase(alist, &nd.act_cnt[i], j)) atomicAdd(&row_warn(nd, i)[2], 1);
        count_hit(nd, kHitDeactivate);
    }
}

static __device__ void update_row(const NetDev& nd, int i)
{
    for (int e = nd.rr_ptr[i]; e < nd.rr_ptr[i + 1]; ++e) update_synapse(nd, i, e);
}

// NGS:16363-16415
static __device__ void axon_remodeling(const NetDev& nd, int i)
{
    int base = nd.rr_ptr[i];
    int* slist = nd.sup_list + (size_t)i * nd.syn.Nss;
    for (int k = 0; k < nd.act_cnt[i]; ++k) nd.act[edge_of(nd, i, nd.act_list[base + k])] = 0;
    for (int k = 0; k < nd.sup_cnt[i]; ++k) nd.act[edge_of(nd, i, slist[k])] = 1;
    for (int k = 0; k < nd.sup_cnt[i]; ++k) nd.act_list[base + k] = slist[k];
    nd.act_cnt[i] = nd.sup_cnt[i];
    nd.remodeled[i] = 1;
    count_hit(nd, kHitRemodel);
}

// LTP / LTD of row i at the sync step t1 against the window's post events of
// its own network, keys sorted by (network, step, neuron) (NGS:9375-9554).
// last_pre[i] is the latest pre-event time of the row: a row whose pre events
// are all older than STDP_WINDOW (100 ms) changes no weight here, so its scan
// is skipped (axon remodeling still runs as in the reference).
static __device__ void rr_sync(const NetDev& nd, int i, int t1, const uint64_t* keys, int np)
{
    if (np == 0) return;
    const int Nss = nd.syn.Nss;
    const int net = nd.ra_net[i];
    const int k_lo = post_lower_bound(keys, np, post_key(net, 0, 0));
    const int k_hi = post_lower_bound(keys, np, post_key(net + 1, 0, 0));
    if (k_lo == k_hi) return;
    const double now = nd.clock[t1];
    const double lim = now - 100.0;
    if (nd.last_pre[i] >= lim) {
        const bool saturated = (nd.sup_cnt[i] == Nss);
        const int* slist = nd.sup_list + (size_t)i * Nss;
        for (int k = k_lo; k < k_hi; ++k) {
            int post = post_key_id(keys[k]);
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
            int q = first;
            while (q < pc && ring[q % kPreRing] < lim) ++q;
            for (; q < pc; ++q) {
                double dt = now - ring[q % kPreRing];
                if (dt <= nd.syn.T_0) {
                    count_hit(nd, saturated ? kHitSyncSaturated : kHitSyncLtd);
                    nd.w[e] = ltd(nd.syn, nd.w[e], dt, row_warn(nd, i));
                } else {
                    count_hit(nd, saturated ? kHitSyncSaturated : kHitSyncLtp);
                    nd.w[e] = ltp(nd.syn, nd.w[e], dt, row_warn(nd, i));
                }
                update_synapse(nd, i, e);
            }
        }
    }
    // NGS:9548-9554
    if ((nd.sup_cnt[i] == Nss) && (nd.remodeled[i] == 0)) axon_remodeling(nd, i);
}

// The sync-point STDP of one window, for the rows [base, base + blockDim.x):
// the block gathers the window's post events of its rows' networks (the list
// is appended in any order), sorts them in LDS (bitonic) and runs rr_sync.

static __device__ void stdp_block(const NetDev& nd, const StdpArgs& sa, int base, uint64_t* slice, int* scount)
{
    const int tid = threadIdx.x;
    if (base >= nd.n_ra) return;
    int n = *sa.count;
    if (n > kPostListCap) n = kPostListCap;
    if (n == 0) return;
    const int last = min(base + (int)blockDim.x, nd.n_ra) - 1;
    const int net_lo = nd.ra_net[base], net_hi = nd.ra_net[last];
    if (tid == 0) *scount = 0;
    __syncthreads();
    for (int k = tid; k < n; k += blockDim.x) {
        const uint64_t key = sa.posts[k];
        const int net = (int)(key >> 52);
        if (net >= net_lo && net <= net_hi) {
            const int pos = atomicAdd(scount, 1);
            if (pos < kStdpSliceCap) slice[pos] = key;
        }
    }
    __syncthreads();
    int m = *scount;
    if (m > kStdpSliceCap) {
        if (tid == 0) atomicOr(nd.err, (unsigned)kErrPostList);
        m = kStdpSliceCap;
    }
    if (m > 1) {
        int np2 = 1;
        while (np2 < m) np2 <<= 1;
        for (int k = m + tid; k < np2; k += blockDim.x) slice[k] = ~0ull;
        __syncthreads();
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = tid; k < np2; k += blockDim.x) {
                    const int p = k ^ stride;
                    if (p > k) {
                        const bool up = ((k & size) == 0);
                        const uint64_t a = slice[k], b = slice[p];
                        if ((a > b) == up) { slice[k] = b; slice[p] = a; }
                    }
                }
                __syncthreads();
            }
        }
    }
    const int i = base + tid;
    if (i < nd.n_ra && nd.net_growth[nd.ra_net[i]]) rr_sync(nd, i, sa.t1, slice, m);
}

}  // namespace hvc
