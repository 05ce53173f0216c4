/*
 * hvc_engine.h -- C ABI of the MI355X (gfx950) HVC growth-trial engine.
 *
 * The engine replaces the per-trial inner loop of the reference growth
 * simulator (paths relative to the reference checkout, "NGS" =
 * c++/src/NetworkGrowthSimulator.cpp, "NGS.h" = c++/include/NetworkGrowthSimulator.h):
 * neuron integration, spike delivery through axonal-delay queues into synaptic
 * conductances, and the STDP weight / synapse-state machine.  The host keeps the
 * growth driver, file formats and replacement bookkeeping and calls these
 * entry points.  Plain C types only; all pointers are borrowed for the call.
 *
 * Errors: every call returns hvc_status; hvc_last_error() gives a message.
 * Threading: a context is not thread-safe; separate contexts are independent.
 */
#ifndef HVC_ENGINE_H
#define HVC_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hvc_ctx hvc_ctx;

typedef enum {
    HVC_OK = 0,
    HVC_E_ARG = 1,      /* invalid argument or network description */
    HVC_E_OOM = 2,      /* device or host allocation failed */
    HVC_E_HIP = 3,      /* HIP runtime error */
    HVC_E_NUMERIC = 4,  /* check_bad_values flagged an HVC-RA neuron (HH2_test.cpp:244-260) */
    HVC_E_STATE = 5,    /* call out of order (e.g. run before upload) */
    HVC_E_CAPACITY = 6, /* an engine buffer overflowed; the trial's spikes are discarded (no spike query
                           answers until the next trial) and its synapse state is undefined */
    HVC_PENDING = 7     /* hvc_poll: the submitted trial has not finished yet (not an error) */
} hvc_status;

/* SynapticParameters, c++/include/ConfigurationNetworkGrowth.h:11-34 */
typedef struct {
    int Nss;
    double A_P, A_D, T_0, T_P, T_D, TAU_P, TAU_D;
    double BETA, BETA_SUPERSYNAPSE;
    double ACTIVATION_THRESHOLD, SUPERSYNAPSE_THRESHOLD, WEIGHT_MAX;
} hvc_synaptic_params;

typedef struct {
    int ra_model;       /* 0: HH2_test + Debraband_step (NGS:8943), 1: HH2_buffer + Runge4_step */
    int i_model;        /* 0: HHI_buffer + DP8_step (NGS:9223),    1: HHI_buffer + Runge4_step  */
    int precision;      /* 0: FP64 state (bit-exact form); 1: FP32 neuron state */
    double dt;          /* TIMESTEP (NGS:15) */
    double trial_ms;    /* TRIAL_DURATION (NGS:17) */
    double sync_ms;     /* NETWORK_UPDATE_FREQUENCY (NGS:16) */
    double mu_s, sigma_s, mu_d, sigma_d;  /* NoiseParameters, ConfigurationNetworkGrowth.h:77-89 */
    int ra_poisson;     /* Poisson conductance noise on HVC-RA (RK4 form only) */
    int i_poisson;      /* Poisson conductance noise on HVC-I (NGS:2061-2075 enables it) */
    uint32_t seed_lo, seed_hi;   /* noise-stream key */
    hvc_synaptic_params syn;
    int spike_capacity; /* max spikes per trial per kind (0 = default) */
} hvc_params;

/* Network.  RA->I and I->RA are CSR by presynaptic neuron (syn_ID_*, weights_*,
 * axonal_delays_* of NGS.h); RA->RA is a candidate set, CSR by presynaptic
 * neuron with ascending targets (the reference's dense N_RA x N_RA arrays,
 * NGS.h:142-172, are the candidate set of all targets).  Synapse state: weight
 * per candidate, active / super lists in insertion order, remodeled flags. */
typedef struct {
    int n_ra, n_i;
    const int* ra2i_ptr; const int* ra2i_tgt; const double* ra2i_w; const double* ra2i_d;
    const int* i2ra_ptr; const int* i2ra_tgt; const double* i2ra_w; const double* i2ra_d;
    const int* rr_ptr; const int* rr_tgt; const double* rr_d;
    const double* rr_w;
    const int* act_cnt; const int* act_ids;
    const int* sup_cnt; const int* sup_ids;
    const int* remodeled;
    const double* erest;   /* per HVC-RA: set_Erest (NGS:2141) */
    const double* gca;     /* per HVC-RA: set_GCa  (NGS:2142) */
    int n_train; const int* train_ids; const double* spread;
} hvc_network;

typedef struct {
    int mode;                   /* 0: growth trial, 1: chain-test trial */
    int training;               /* growth: apply the training kick (NGS:8933) */
    uint32_t trial_index;       /* selects the noise streams of this trial */
    int first_trial;            /* first trial after initialize_network (NGS:849-850) */
    double mean_injection_time; /* NGS:8898 (growth) / 12785 (chain test: 100) */
} hvc_trial;

/* Outputs in caller-owned buffers.  *_counts has one entry per neuron; *_times
 * (may be NULL) holds the spike times neuron by neuron (capacity *_cap); *_total
 * is the spike count.  With *_times NULL only counts are written; the times stay
 * available through hvc_trial_spikes until the next trial. */
typedef struct {
    int* soma_counts; double* soma_times; int soma_cap; int soma_total;
    int* dend_counts; double* dend_times; int dend_cap; int dend_total;
    int* i_counts; double* i_times; int i_cap; int i_total;
    int bad_values;             /* -1 when check_bad_values fails, else 0 */
    int warnings[3];            /* LTP dt<=T_0, LTD dt>T_0, list entry missing (stderr in the reference) */
    double* ra_state;           /* optional: 12 doubles per HVC-RA (Vs Vd n l h r c Ca Ge_s Gi_s Ge_d Gi_d) */
    double* i_state;            /* optional: 7 doubles per HVC-I (V n m h w Ge Gi) */
} hvc_trial_out;

typedef struct {
    double* w;                  /* per candidate edge */
    int* act_cnt; int* act_ids; /* act_ids: concatenated lists, capacity = candidate count */
    int* sup_cnt; int* sup_ids;
    int* remodeled;
} hvc_synapses;

/* Create a context on a device.  Replaces NetworkGrowthSimulator's constructor
 * and set_dynamics / set_noise (NGS:48-82, 2061-2095). */
hvc_status hvc_create(const hvc_params* params, int device, hvc_ctx** out);

/* Upload the network and synapse state.  Replaces initialize_network and the
 * send_* scatters (NGS:745-890, 3536-4232). */
hvc_status hvc_upload_network(hvc_ctx* ctx, const hvc_network* net);

/* Run one trial.  mode 0 replaces trial_1stSoma_pre_1stSoma_post_delays_fixedSpread
 * (NGS:8839-9571, called at NGS:5792) including potentiation_decay and
 * update_all_synapses; mode 1 replaces trial_no_stdp_fixedSpread (NGS:12747-13118,
 * called at NGS:1406).  Every trial starts from rest (setToRest_afterEpoch,
 * NGS:4591-4621).  Returns HVC_E_NUMERIC when check_bad_values fails (NGS:4538-4551). */
hvc_status hvc_run_trial(hvc_ctx* ctx, const hvc_trial* trial, hvc_trial_out* out);

/* Spike times of the last trial: kind 0 somatic, 1 dendritic (HVC-RA), 2 HVC-I.
 * counts gets one entry per neuron; times (capacity cap, may be NULL) the spike
 * times neuron by neuron; *total the number of spikes.  The data of
 * spikes_in_trial_{soma,dend,interneuron} (NGS.h:202-204). */
hvc_status hvc_trial_spikes(hvc_ctx* ctx, int kind, int* counts, double* times, int cap, int* total);

/* Read back weights, active / super lists and remodeled flags (the data of
 * gather_full_state_data / write_full_network_state, NGS:18552-18593). */
hvc_status hvc_download_synapses(hvc_ctx* ctx, hvc_synapses* syn);

/* Replace the synapse state (after replace_neurons, NGS:15757-15881). */
hvc_status hvc_upload_synapses(hvc_ctx* ctx, const hvc_synapses* syn);

/* Set Erest / GCa of n neurons (set_neuron_properties and
 * update_neuron_properties_sameDendrite_diffMaturationRate, NGS:2136-2183). */
hvc_status hvc_patch_neurons(hvc_ctx* ctx, const int* ids, int n, const double* erest, const double* gca);

/* Batched form: n_nets independent networks resident in one context, laid
 * side by side in device memory and advanced together by one launch per kernel
 * and coupling window.  keys holds 2 words per network (its noise-stream key;
 * hvc_params.seed_* is not used).  Network b's results equal what a context
 * holding network b alone (seed_lo/hi = keys[2b], keys[2b+1]) produces.  All
 * networks share hvc_params (models, precision, dt, trial length, synaptic
 * parameters).  Neuron ids in every per-network call are local to the network.
 * No counterpart in the reference, which runs one network per simulator
 * process (NetworkGrowthSimulator, NGS.h:30-40). */
hvc_status hvc_upload_batch(hvc_ctx* ctx, int n_nets, const hvc_network* nets, const uint32_t* keys);

/* One trial of every network: trials[b] and outs[b] (outs may be NULL) belong to
 * network b; each network runs its own mode (growth or chain test).  Returns HVC_E_NUMERIC when
 * any network fails check_bad_values (outs[b].bad_values tells which). */
hvc_status hvc_run_batch(hvc_ctx* ctx, const hvc_trial* trials, hvc_trial_out* outs);

/* Per-network views of the last trial and of the synapse state (as
 * hvc_trial_spikes, hvc_download_synapses, hvc_upload_synapses). */
hvc_status hvc_batch_spikes(hvc_ctx* ctx, int net, int kind, int* counts, double* times, int cap, int* total);
hvc_status hvc_batch_download_synapses(hvc_ctx* ctx, int net, hvc_synapses* syn);
hvc_status hvc_batch_upload_synapses(hvc_ctx* ctx, int net, const hvc_synapses* syn);

/* Launch shapes of the per-window kernels.  hvc_autotune times the candidates
 * on the resident networks (a scratch chain test through the first `windows`
 * coupling windows, kicked early so that activity propagates; synapse state is
 * not touched and the next trial starts from rest as always), keeps the fastest
 * and returns it with its per-window device times (ms) of integrate, deliver
 * and sync.  Results do not depend on the launch shapes.  No counterpart in the
 * reference. */
typedef struct {
    int integrate_block;   /* threads per block of the integration kernel: 64, 128, 256 */
    int i_lanes;           /* HVC-I DP8: 3 = eight lanes per neuron, 2 = four lanes, 1 = five lanes,
                              0 = one thread per neuron */
    int deliver_block;     /* 64, 128, 256 */
    int sync_block;        /* 64 .. 512 */
    int ra_lanes;          /* HVC-RA RK4: 1 = four lanes per neuron, 0 = one thread per neuron */
    int deliver_lists;     /* 1 = the delivery visits only the neurons that spiked in the window or
                              still have spikes in flight (lists kept on the device), 0 = it reads
                              every neuron's queue state each window */
} hvc_tuning;
hvc_status hvc_get_tuning(hvc_ctx* ctx, hvc_tuning* tuning);
hvc_status hvc_set_tuning(hvc_ctx* ctx, const hvc_tuning* tuning);
hvc_status hvc_autotune(hvc_ctx* ctx, int windows, hvc_tuning* best, double* ms);

/* Serving pool: a context whose networks join and leave between trials.  The
 * resident networks are packed side by side in device memory (a release moves
 * the networks behind it down, on the device, with their ids and edge
 * indices); every hvc_pool_run advances all of them by one trial in one launch
 * per kernel and coupling window, each with its own trial (growth or chain
 * test, trial index, kick time).  Network state stays on the device from
 * admission to release.  Capacities fix the pool's device footprint; a network
 * that does not fit returns HVC_E_CAPACITY from hvc_pool_admit.  Handles are
 * stable across admissions and releases.  No counterpart in the reference. */
typedef struct {
    int max_nets;                   /* resident networks */
    int max_ra, max_i;              /* neurons of all resident networks */
    long long max_rr;               /* RA->RA candidates */
    long long max_ra2i, max_i2ra;   /* fixed-weight connections */
    int max_in_degree;              /* per target and connection type */
    int spikes_per_neuron;          /* spike-log capacity per neuron and trial (0: 256) */
} hvc_pool_caps;

/* Device bytes one network of the given sizes occupies in a pool of these
 * capacities (its HBM footprint), and the bytes of the whole pool. */
hvc_status hvc_pool_footprint(const hvc_params* params, const hvc_pool_caps* caps, int n_ra, int n_i,
                              long long n_rr, long long n_ra2i, long long n_i2ra, size_t* network_bytes,
                              size_t* pool_bytes);
hvc_status hvc_pool_create(const hvc_params* params, int device, const hvc_pool_caps* caps, hvc_ctx** out);
hvc_status hvc_pool_admit(hvc_ctx* ctx, const hvc_network* net, uint32_t key_lo, uint32_t key_hi, int* handle);
hvc_status hvc_pool_release(hvc_ctx* ctx, int handle);
/* One trial of every resident network: n must equal the number of resident
 * networks; trials[k] and outs[k] (outs may be NULL) belong to handles[k]. */
hvc_status hvc_pool_run(hvc_ctx* ctx, int n, const int* handles, const hvc_trial* trials, hvc_trial_out* outs);
hvc_status hvc_pool_spikes(hvc_ctx* ctx, int handle, int kind, int* counts, double* times, int cap, int* total);
hvc_status hvc_pool_download_synapses(hvc_ctx* ctx, int handle, hvc_synapses* syn);

/* Asynchronous serving on a pool: the serving entry points of the build's
 * boundary, for the chain-test requests the reference serves one at a time
 * (test_chain, NGS:1313-1482, trial at NGS:1406) and for growth trials.
 * hvc_submit queues one trial of resident network `handle` and returns at
 * once with a ticket.  A host thread of the context (started by the first
 * submission, stopped by hvc_destroy) runs rounds: a round takes the oldest
 * queued trial of every network that has one and runs them together, as
 * hvc_pool_run would; networks with nothing queued sit the round out
 * unchanged.  A network's trials run in submission order, one per round; a
 * trial submitted while a round runs joins a later round.  Results equal the
 * same trials run alone (hvc_run_trial on the network with the same key).
 * hvc_poll returns HVC_PENDING until the ticket's trial is done, with *round
 * (may be NULL) -1 while queued, else the index of the round that runs it;
 * then the trial's status (HVC_OK, HVC_E_NUMERIC as hvc_run_trial, or the
 * round's error) with out filled as by hvc_run_trial (counts per neuron of the
 * network; times, when given, need *_cap >= the totals, else HVC_E_ARG with
 * the totals set and the ticket kept), and the ticket is released.  With out
 * NULL the call reports the status only and keeps the ticket.  hvc_submit and hvc_poll may be called from any thread while a round
 * runs; hvc_pool_admit, _release, _run, _spikes and _download_synapses wait
 * for the round boundary; hvc_pool_release refuses a network whose submitted
 * trials have not finished.  A handle that is not resident when its round
 * comes fails that ticket with HVC_E_STATE.  No counterpart in the reference. */
hvc_status hvc_submit(hvc_ctx* ctx, int handle, const hvc_trial* trial, uint64_t* ticket);
hvc_status hvc_poll(hvc_ctx* ctx, uint64_t ticket, hvc_trial_out* out, long long* round);

/* Single-neuron traces on the device: the neuron seam of the reference
 * (HH2_test::Debraband_step_no_target_update HH2_test.cpp:535-545,
 * HH2_buffer::R4_step_no_target_update HH2_buffer.cpp:521-528,
 * HHI_buffer::DP8_step_no_target_update / R4_step_no_target_update
 * HHI_buffer.cpp:323-339, with raiseE / raiseI / raiseExcWeight between steps).
 * Arrays are host memory laid out [step][neuron]; out_state [n_steps+1][n][12|7]. */
hvc_status hvc_trace_ra(hvc_ctx* ctx, int n, int n_steps, const double* erest, const double* gca,
                        uint32_t trial, const double* kick, const double* exc_w, const double* inh_g,
                        double* out_state, uint8_t* out_fired);
hvc_status hvc_trace_i(hvc_ctx* ctx, int n, int n_steps, int first_trial, uint32_t trial,
                       const double* exc_g, const double* inh_g, double* out_state, uint8_t* out_fired);

/* hvc_trace_ra on device arrays, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every array is device memory of the context's device, in hvc_trace_ra's
 * layouts; results equal hvc_trace_ra's bit for bit.  The trace clock and decay
 * tables live in a fixed buffer that the library holds on each GPU until the
 * process ends (about 3.8 MB), so n_steps is at most HVC_NOALLOC_MAX_STEPS
 * (HVC_E_ARG above).  Every call on one GPU shares that buffer: calls that
 * overlap on one GPU (from several threads or contexts) race.  The call returns
 * after the context's stream has finished the trace. */
#define HVC_NOALLOC_MAX_STEPS 16384
hvc_status hvc_trace_ra_noalloc(hvc_ctx* ctx, int n, int n_steps, const double* d_erest, const double* d_gca,
                                uint32_t trial, const double* d_kick, const double* d_exc_w, const double* d_inh_g,
                                double* d_out_state, uint8_t* d_out_fired);

/* Kernel timing: with timing enabled, HIP events on the engine's stream bracket
 * every 16th window's launches (and the end-of-trial decay); hvc_kernel_timing
 * returns the average device duration (ms) per launch of a kernel class over
 * all trials since hvc_set_timing(ctx, 1) (which clears the averages).
 * names: "integrate" (HVC-RA and HVC-I neurons with the previous window's
 * conductance updates, one launch per window), "deliver" (post-event sort and
 * delivery, per window), "sync" (the last window's updates, once per trial),
 * "decay" (once per growth trial).
 * No counterpart in the reference (it has no profiler). */
hvc_status hvc_kernel_timing(hvc_ctx* ctx, const char* name, double* avg_ms, int* launches);
hvc_status hvc_set_timing(hvc_ctx* ctx, int enabled);

/* Branch counters of the synapse-state machine, summed over the trials (and
 * networks) run since hvc_set_branch_counting(ctx, 1), which clears them
 * (0 stops counting).  counts gets HVC_NUM_BRANCH_COUNTERS entries, in order:
 *   0 activate, 1 super, 2 unsuper, 3 deactivate   (update_synapse transitions, NGS:16546-16592)
 *   4 remodel                                      (axon_remodeling, NGS:16363-16415)
 *   5 ltd_delivery, 6 ltd_delivery_saturated       (LTD on delivery, NGS:9104-9183, per application)
 *   7 desaturation                                 (a saturated row lost a supersynapse, NGS:9144-9146)
 *   8 sync_ltp, 9 sync_ltd, 10 sync_saturated      (sync-point STDP, NGS:9375-9542, per application)
 * No counterpart in the reference: it shows which branches a workload drove. */
#define HVC_NUM_BRANCH_COUNTERS 11
hvc_status hvc_set_branch_counting(hvc_ctx* ctx, int enabled);
hvc_status hvc_branch_counts(hvc_ctx* ctx, int* counts);

/* Device evaluation of the engine's own math on n host values: fn 0 exp(a),
 * 1 log(a), 2 a / b through the constant-divisor path.  No counterpart in the
 * reference: it lets the CPU check the device's exp / log / division bit for
 * bit (the reference calls libm exp / log, HH2_test.h:265-276). */
hvc_status hvc_eval_math(hvc_ctx* ctx, int fn, int n, const double* a, const double* b, double* out);

void hvc_destroy(hvc_ctx* ctx);
const char* hvc_last_error(const hvc_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
