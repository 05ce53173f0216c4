// hvc_internal.h -- device data layout shared by the engine's kernels and host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hvc_engine.h"

namespace hvc {

constexpr int kPostRing = 8;     // post-event times kept per HVC-RA neuron
constexpr int kPreRing = 4;      // delivered pre-event times kept per candidate edge
constexpr int kMaxStreams = 16;  // in-flight RA->RA spikes per presynaptic row
constexpr int kMaxInflight = 16; // in-flight spikes per neuron for fixed-weight pushes
constexpr int kMaxWindowSteps = 64;  // steps between two synchronisation points
constexpr int kPostListCap = 8192;  // new post events per window, all networks together
constexpr int kMaxNets = 4096;      // networks per context (12 bits of a post-event key)
constexpr int kWlCounters = 9;      // NetDev::wl_cnt: 3 roles x 3 list counts

// hvc_trace_ra_noalloc: its clock and decay tables live in a fixed per-device
// array (hvc_noalloc.hip) sized for this many trace steps
constexpr int kNoAllocMaxSteps = HVC_NOALLOC_MAX_STEPS;
constexpr int kNoAllocMaxDevices = 64;
constexpr size_t kNoAllocDecayRowBytes = 224;   // sizeof(DecayRow)
constexpr size_t kNoAllocClockBytes = ((size_t)(kNoAllocMaxSteps + 1) * sizeof(double) + 255) / 256 * 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocClockBytes + (size_t)(kNoAllocMaxSteps + 1) * kNoAllocDecayRowBytes;

// error flags raised by kernels (bit mask)
enum ErrFlag : uint32_t {
    kErrSpikeLog = 1u,
    kErrPostList = 2u,
    kErrStreams = 4u,
    kErrInflight = 8u,
    kErrInbox = 16u,
    kErrPostRing = 32u,
    kErrNoise = 64u,
    kErrWorklist = 128u,
};

// counters of the synapse-state machine's branches (hvc_set_branch_counting):
// update_synapse transitions, axon remodeling, LTD on delivery (unsaturated /
// saturated row), desaturation, and the sync-point LTP / LTD (unsaturated
// row) and applications in saturated rows
enum BranchHit {
    kHitActivate = 0,
    kHitSuper,
    kHitUnsuper,
    kHitDeactivate,
    kHitRemodel,
    kHitLtdDelivery,
    kHitLtdDeliverySaturated,
    kHitDesaturation,
    kHitSyncLtp,
    kHitSyncLtd,
    kHitSyncSaturated,
    kNumHits
};
static_assert(kNumHits == HVC_NUM_BRANCH_COUNTERS, "branch counters of the C ABI");

// Poisson noise events of one neuron, generated ahead of the integration for a
// chunk of coupling windows (the draws depend on the stream and the clock
// only): at step `step` add `inc` to conductance `target` (HVC-RA 0 Gi_s,
// 1 Gi_d, 2 Ge_s, 3 Ge_d; HVC-I 0 Ge, 1 Gi), in list order.  A list ends with
// two entries of step kNoNoiseStep.
struct NoiseEvent {
    int step;
    int target;
    double inc;
};
constexpr int kNoNoiseStep = 0x7fffffff;

// the sync-point STDP of one window (hvc_stdp.hip.h)
struct StdpArgs {
    const uint64_t* posts;   // the window's post events (unsorted)
    const int* count;
    int t1;                  // the window's sync step
};
constexpr int kStdpSliceCap = 1024;        // post events per window of one block's networks
constexpr int kNoiseChunkWindows = 64;   // windows per generated chunk

// post-event key: network (12 bits), step (24 bits), neuron (28 bits); the
// sorted keys list each network's events by (step, neuron)
__host__ __device__ __forceinline__ uint64_t post_key(int net, int step, int id)
{
    return ((uint64_t)(uint32_t)net << 52) | ((uint64_t)(uint32_t)step << 28) | (uint64_t)(uint32_t)id;
}
__host__ __device__ __forceinline__ int post_key_id(uint64_t k) { return (int)(k & 0xfffffffull); }
constexpr int kMaxKeySteps = 1 << 24;
constexpr int kMaxKeyNeurons = 1 << 28;

// identity of one neuron's noise streams: key of its network, trial index of
// its network, neuron id within its network
struct NoiseId {
    uint32_t k0, k1, trial, local;
};

// launch shapes of the per-window kernels (chosen by the autotuner)
struct Tuning {
    int integrate_block = 64;   // threads per block of k_integrate (64, 128, 256)
    int i_lanes = 3;            // HVC-I DP8: 3 = eight lanes per neuron, 2 = four, 1 = five, 0 = one thread
    int ra_lanes = 0;           // HVC-RA RK4: 1 = four lanes per neuron, 0 = one thread per neuron
    int deliver_block = 64;     // threads per block of k_deliver
    int deliver_lists = 1;      // 1: the delivery visits listed neurons only, 0: every neuron's slot table
    int sync_block = 256;       // threads per block of k_sync
};

struct Record {
    uint64_t key;   // (arrival step << 32) | presynaptic id
    double w;
};

// a candidate edge at one position of its row's (delay, target) order: the
// edge, its target and its delay, read in one load
struct RrPos {
    int e, tgt;
    double d;
};

// Fixed-weight delivery edges sorted by delay within each presynaptic row.
struct FixedEdges {
    const int* ptr;
    const int* tgt;
    const double* w;
    const double* d;
};

struct Inbox {
    Record* rec;    // [n_targets][cap]
    int* cnt;       // [n_targets]
    int cap;
};

// one in-flight spike of a fixed-weight push: spike step (-1: free slot), edge
// cursor, delivery time of the cursor's edge, clock at the spike step
struct FlSlot {
    int step, cur;
    double tnext, tspike;
};
constexpr int kHeadSlots = 5;
// a neuron's count of held slots and its first slots in one 128-byte line, read
// in one go when a window starts; the other slots beside it
struct alignas(128) FlHead {
    int count, pad;
    FlSlot s[kHeadSlots];
};
static_assert(sizeof(FlHead) == 128, "one line per neuron");
struct InFlight {
    FlHead* head;   // [n]
    FlSlot* more;   // [n][kMaxInflight - kHeadSlots]
    int* held;      // [n] the head's count, apart: a window reads the head line
                    // only of a neuron with spikes held or new
};

struct NetDev {
    int n_ra, n_i;
    int ra_model, i_model;
    double dt;
    int num_steps;
    const double* clock;         // [num_steps], clock[t] = internal_time after step t
    const void* decay;           // DecayRow [num_steps] (row t used by step t)
    double sq_sig_s, sq_sig_d, mu_term_s, mu_term_d;
    int ra_poisson, i_poisson;
    uint32_t seed_lo, seed_hi;   // noise key of the single-neuron traces
    hvc_synaptic_params syn;

    // networks resident in this context: neurons of network b are the RA ids
    // [net_ra_off[b], net_ra_off[b+1]) and the I ids [net_i_off[b], net_i_off[b+1])
    int n_nets;
    const int* ra_net;           // [n_ra] network of each HVC-RA
    const int* i_net;            // [n_i]
    const int* net_ra_off;       // [n_nets + 1]
    const int* net_i_off;        // [n_nets + 1]
    const uint32_t* net_key;     // [2 n_nets] noise-stream key
    const uint32_t* net_trial;   // [n_nets] trial index of the running trial
    const int* net_first;        // [n_nets] first trial after initialisation (NGS:849-850)
    const int* net_growth;       // [n_nets] 1: growth trial (STDP), 0: chain-test trial
    const int* net_live;         // [n_nets] 0: the network sits the trial out (nullptr: every network runs)

    // RA state (SoA)
    double* ra[12];              // Vs Vd n l h r c Ca Ge_s Gi_s Ge_d Gi_d
    int* ra_flags;               // bit0 flag_soma, bit1 flag_dend
    double* ra_noise;            // [4][n_ra]: exc_soma inh_soma exc_dend inh_dend
    uint32_t* ra_draw;
    const double* EL;
    const double* GCa;
    const int* kick_step;        // step of the training kick, -1 none
    // I state
    double* in[7];               // V n m h w Ge Gi
    int* i_flag;
    double* i_noise;             // [2][n_i]: exc inh
    uint32_t* i_draw;
    // noise events of the current chunk: [n][ev_cap] per neuron, and the index
    // of each neuron's next unapplied event
    NoiseEvent* ra_ev;
    NoiseEvent* i_ev;
    int* ra_ev_pos;
    int* i_ev_pos;
    int ev_cap;

    // per-window outputs
    uint64_t* soma_mask;
    uint64_t* dend_mask;
    uint64_t* i_mask;
    int2* soma_log; int2* dend_log; int2* i_log;
    int* log_cnt;                // [3]
    int log_cap;

    // post events (growth)
    double* post_ring;           // [n_ra][kPostRing]
    int* post_cnt;
    int* post_pend;
    uint64_t* post_list;         // post_key(network, step, neuron) in any order; the window's buffer
    int* post_list_cnt;          // its count
    int* post_list_next_cnt;     // count of the other buffer (the next window's), reset by k_deliver
    double* last_pre;            // [n_ra] latest pre-event time of the row in this trial
    double* st_next;             // [n_ra] earliest delivery time of the row's in-flight spikes

    // RA->RA candidates
    const int* rr_ptr;
    const int* rr_tgt;
    const double* rr_d;
    const int* rr_perm;          // per row: candidate edges sorted by (delay, target)
    const RrPos* rr_pos;         // [E] rr_perm's edges with their targets and delays
    double* w;
    uint8_t* act;
    uint8_t* sup;
    int* act_list;               // [E] row-segmented
    int* act_cnt;
    int* sup_list;               // [n_ra][Nss]
    int* sup_cnt;
    int* remodeled;
    double* pre_ring;            // [E][kPreRing]
    int* pre_cnt;                // [E]
    int dense;

    // RA->RA in-flight streams (growth)
    int* st_step;                // [n_ra][kMaxStreams]
    int* st_mode;
    int* st_cursor;
    int* st_n;
    int* st_count;               // [n_ra]
    int* st_items;               // [n_ra][kMaxStreams][Nss] edge indices, sorted by (T, target)

    // fixed-weight pushes
    FixedEdges ra2i, i2ra;
    InFlight fl_ra2i, fl_i2ra, fl_rr;
    int* chain_e;                // chain test: row i's active edges sorted by (delay, target) at rr_ptr[i]

    Inbox in_gi_ra, in_ge_ra, in_ge_i;

    // delivery worklists, per role (0 RA->RA rows, 1 RA->I pushes, 2 I->RA
    // pushes): the presynaptic neurons the window's delivery visits.  The
    // delivery of window w - 1 lists those with spikes still in flight, the
    // integration of window w adds those that spike in it.
    int* wl_list[3];             // [2][wl_cap[r]], list of window w in half w % 2
    int* wl_mark[3];             // [wl_cap[r]] window the neuron is listed for (-1: none)
    int* wl_cnt;                 // [3 roles][3] list counts, by window % 3
    int wl_cap[3];
    int window;                  // index of the launch's window in the trial
    int wl_on;                   // the delivery runs on the lists (the integration fills them)
    int growth_mode;             // 1: every network runs a growth trial, 0: every one a chain test, 2: mixed

    unsigned int* err;
    int* warn;                   // [3 n_nets]
    int* hits;                   // [kNumHits] branch counters, nullptr when not counting
};

__device__ __forceinline__ NoiseId ra_noise_id(const NetDev& nd, int i)
{
    int b = nd.ra_net[i];
    return NoiseId{nd.net_key[2 * b], nd.net_key[2 * b + 1], nd.net_trial[b], (uint32_t)(i - nd.net_ra_off[b])};
}

__device__ __forceinline__ NoiseId i_noise_id(const NetDev& nd, int i)
{
    int b = nd.i_net[i];
    return NoiseId{nd.net_key[2 * b], nd.net_key[2 * b + 1], nd.net_trial[b], (uint32_t)(i - nd.net_i_off[b])};
}

}  // namespace hvc
