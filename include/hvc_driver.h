/*
 * hvc_driver.h -- C ABI of the host growth driver around the HVC engine.
 *
 * The driver keeps what NetworkGrowthSimulator keeps around the trial (paths
 * relative to the reference checkout, "NGS" = c++/src/NetworkGrowthSimulator.cpp):
 * the growth loop chain_growth (NGS:5731-5900) with its bookkeeping after every
 * trial -- activity history and check_neuron_activity (NGS:9573-9627),
 * maturation update_neuron_properties_sameDendrite_diffMaturationRate
 * (NGS:2152-2183), replacement of silent neurons replace_neurons with
 * remove_neurons_from_network, resampling and update_replaced_neurons
 * (NGS:15757-15881, 157-220, 528-618, 2109-2133) -- and the on-disk state files
 * of write_full_network_state / write_graph_network_state (NGS:18552-18640).
 * Trials run on the engine (include/hvc_engine.h).  Host code only; every
 * decision is exact given the spikes of the trial.
 *
 * Random numbers (injection times, resampling) come from the reference's
 * generator, std::mt19937 with random(G) = G * u / max (c++/src/poisson_noise.cpp),
 * seeded by the caller.  The reference draws its neuron noise from the same
 * generator, the engine from counter-based streams, so the two runs draw
 * different injection times after the first trial; given the same draws the
 * driver's decisions are the reference's.
 */
#ifndef HVC_DRIVER_H
#define HVC_DRIVER_H

#include <stdint.h>

#include "hvc_engine.h"
#include "hvc_topology.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MaturationParameters, c++/include/ConfigurationNetworkGrowth.h:37-75 (the fields
 * the growth loop reads) */
typedef struct {
    double E_REST_MATURE, E_REST_IMMATURE;
    double GCA_MATURE, GCA_IMMATURE;
    double DEATH_RATE_THRESHOLD;
    int RATE_WINDOW_LONG;       /* capacity of the activity history (>= the activity window) */
    int ACTIVITY_WINDOW;        /* trials the activity check reads, and the trials after replacement
                                   before a neuron may be replaced again: 1000 in the reference
                                   (NGS:9586, 9607); 0 = 1000 */
} hvc_maturation_params;

/* Positions of the neurons and the distance unit: what the reference keeps in
 * xx/yy/zz_RA, xx/yy/zz_I and MODEL_INTERNEURON_DISTANCE.  Connection lengths
 * (syn_lengths_*) follow from them (NGS:537, 560). */
typedef struct {
    const double* ra_xyz;            /* 3 per HVC-RA (z ignored on the square) */
    const double* i_xyz;             /* 3 per HVC-I */
    double model_interneuron_distance;
} hvc_geometry;

/* Saved bookkeeping to resume from (all per HVC-RA; NULL = fresh start). */
typedef struct {
    const int* mature;                    /* fresh: 1 for training neurons, else 0 (NGS:1893-1899) */
    const int* maturation_scale;          /* fresh: MATURATION_SCALE_SPONTANEOUS = 100000 (NGS:16137) */
    const int* trials_after_replacement;  /* fresh: 0 */
    const int* history;                   /* [n_ra][RATE_WINDOW_LONG], newest trial first; fresh: 0 */
    int trial_number;                     /* number of the next trial */
} hvc_driver_resume;

typedef struct hvc_driver hvc_driver;

typedef struct {
    int trial;                  /* number of the trial just run */
    double mean_injection_time; /* WAITING_TIME + random(TRIAL_DURATION - 2 WAITING_TIME), NGS:8898 */
    int soma_spikes, dend_spikes, i_spikes;
    int bad_values;             /* -1: check_bad_values failed, the loop stops (NGS:5800-5805) */
    int became_driven;          /* neurons whose maturation scale switched to driven */
    int matured;                /* neurons that became mature */
    int replaced;               /* neurons replaced after this trial */
    int saved;                  /* 1: graph state written, 2: full state, 3: both */
} hvc_driver_report;

/* Create a driver for one network (the engine context is created and the
 * network uploaded).  net->erest / net->gca are the neurons' current
 * properties.  resume may be NULL.  seed seeds the driver's generator. */
hvc_status hvc_driver_create(const hvc_params* params, int device, const hvc_network* net, const hvc_geometry* geometry,
                             const hvc_topology_params* topology, const hvc_maturation_params* maturation,
                             const hvc_driver_resume* resume, uint32_t seed, hvc_driver** out);

/* Periodic state files as chain_growth writes them (NGS:5807-5825): after the
 * trial, before the bookkeeping, graph state every save_freq_short trials and
 * full state (file ending "_<trial>") every save_freq_long trials (0 = never).
 * directory must exist. */
hvc_status hvc_driver_set_output(hvc_driver* d, const char* directory, int save_freq_short, int save_freq_long);

/* One pass of the growth loop (NGS:5765-5898): a growth trial on the engine
 * (training kick if training), check_bad_values, periodic saves, then the
 * bookkeeping.  Returns HVC_E_NUMERIC (report->bad_values = -1) when the trial
 * fails check_bad_values; the driver then refuses further trials. */
hvc_status hvc_driver_trial(hvc_driver* d, int training, hvc_driver_report* report);

/* Write state files now: what = 0 write_graph_network_state (NGS:18620-18640),
 * 1 write_full_network_state with file_ending (NGS:18552-18593), 2 the initial
 * files new_chain_growth writes (I_xy, training_neurons, training_spread,
 * num_neurons and the alterable network with ending "_initial", NGS:1903-1919).
 * Spike times are those of the last trial. */
hvc_status hvc_driver_write_state(hvc_driver* d, const char* directory, const char* file_ending, int what);

/* State views (caller-owned buffers of n_ra entries, history n_ra * RATE_WINDOW_LONG
 * newest first). */
hvc_status hvc_driver_maturation(hvc_driver* d, int* mature, int* maturation_scale, double* erest, double* gca,
                                 int* trials_after_replacement);
hvc_status hvc_driver_history(hvc_driver* d, int* history);
hvc_status hvc_driver_coordinates(hvc_driver* d, double* ra_xyz, double* i_xyz);
/* Every replacement so far: ids[k] replaced after trial trials[k]; *n total (may exceed cap). */
hvc_status hvc_driver_replaced(hvc_driver* d, int* ids, int* trials, int cap, int* n);
/* Fixed-weight connections as the driver holds them (reference order): sizes
 * first (*n_ra2i, *n_i2ra edges), then CSR arrays of those sizes (any may be NULL). */
hvc_status hvc_driver_connections(hvc_driver* d, long long* n_ra2i, long long* n_i2ra, int* ra2i_ptr, int* ra2i_tgt,
                                  double* ra2i_w, double* ra2i_d, int* i2ra_ptr, int* i2ra_tgt, double* i2ra_w,
                                  double* i2ra_d);
/* RA->RA candidate delays (per candidate, the order of net->rr_tgt). */
hvc_status hvc_driver_rr_delays(hvc_driver* d, double* rr_d);
int hvc_driver_trial_number(const hvc_driver* d);
/* The engine context of the current network.  Replacement rebuilds it (launch
 * shapes carry over): the pointer is valid until the next hvc_driver_trial. */
hvc_ctx* hvc_driver_engine(hvc_driver* d);
const char* hvc_driver_last_error(const hvc_driver* d);
void hvc_driver_destroy(hvc_driver* d);

/* Chain test of the current network, test_chain (NGS:1313-1482): num_trials
 * trials of trial_no_stdp_fixedSpread (kick at 100 ms, no STDP, synapse state
 * untouched), noise streams of trial index 0x80000000 + k.  With a directory,
 * each trial's spikes go to test_spike_times_{soma,dend,interneuron}_<k>.bin
 * (write_chain_test, NGS:18595-18604) and the statistics to jitter.bin.  stats
 * (n_ra x 8, may be NULL) receives the jitter statistics in the file's order. */
hvc_status hvc_driver_chain_test(hvc_driver* d, int num_trials, const char* directory, double* stats);

/* Statistics of a chain test, calculate_and_write_jitter (NGS:1641-1782): per
 * HVC-RA neuron over the trials added, probability of a somatic spike, mean
 * number of somatic spikes per trial, mean and standard deviation of the first
 * somatic spike time (-1 when undefined), and the same four for dendritic
 * spikes.  Trials are added in order; spike times neuron by neuron. */
typedef struct hvc_chain_stats hvc_chain_stats;
hvc_status hvc_chain_stats_create(int n_ra, hvc_chain_stats** out);
hvc_status hvc_chain_stats_add(hvc_chain_stats* s, const int* soma_counts, const double* soma_times,
                               const int* dend_counts, const double* dend_times);
/* *num_trials = trials added; out: n_ra x 8 doubles (jitter.bin order). */
hvc_status hvc_chain_stats_result(const hvc_chain_stats* s, int* num_trials, double* out);
/* write_jitter (NGS:17818-17857): N_RA, num_trials, 8 doubles per neuron. */
hvc_status hvc_chain_stats_write(const hvc_chain_stats* s, const char* path);
void hvc_chain_stats_destroy(hvc_chain_stats* s);

/* The bookkeeping steps on host arrays, as the driver runs them.
 * history: [n][window] ring shared by all neurons, logical entry 0 = newest at
 * index *head; a push writes at (*head - 1) mod window.  check_neuron_activity
 * (NGS:9573-9627) with an activity window A (1 <= A <= window; the reference's
 * is 1000): push this trial's somatic spike count; robustness = fraction of the
 * last A trials with a spike; >= 0.5 switches the scale to driven (1000);
 * < death_rate_threshold after more than A trials since replacement marks the
 * neuron for replacement (to_replace, *n_replace, ascending).
 * trials_after_replacement is read as already advanced for this trial
 * (NGS:5831-5832).  recent (may be NULL): per neuron, the number of the newest
 * A history entries that are > 0, kept up to date across calls so that a call
 * costs O(n) instead of O(A n); NULL counts from the history. */
hvc_status hvc_activity_update(int n, int window, int activity_window, int* history, int* head, int* recent,
                               const int* soma_counts,
                               const int* trials_after_replacement, int* maturation_scale,
                               double death_rate_threshold, int* to_replace, int* n_replace, int* became_driven);

/* update_neuron_properties_sameDendrite_diffMaturationRate (NGS:2152-2183) for
 * every immature neuron; *matured counts neurons that became mature. */
hvc_status hvc_maturation_update(int n, const hvc_maturation_params* p, const int* maturation_scale, int* mature,
                                 double* erest, double* gca, int* matured);

#ifdef __cplusplus
}
#endif
#endif
