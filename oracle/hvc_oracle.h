/*
 * hvc_oracle.h -- CPU restatement of the HVC growth-trial hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and bench.py's cpu_baseline leg may load it; the product
 * (unsupervised_growth_amd) never links, loads or calls it.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the reference checkout, "NGS" = c++/src/NetworkGrowthSimulator.cpp).
 * Arithmetic follows the reference expression trees operation by operation,
 * compiled with -ffp-contract=off, so that a GPU path evaluating the same trees
 * produces the same IEEE-754 binary64 results.
 *
 * Parity pins (what fixes each part, see tests/):
 *   ora_exp, ora_log  : checked against glibc exp/log (<= 1 ulp) and against
 *                       60-digit decimal values; special cases exact.
 *   ora_philox4x32    : Random123 known-answer vectors for Philox4x32-10.
 *   RA / I steppers   : bit-exact against the reference neuron classes
 *                       (HH2_test.cpp, HH2_buffer.cpp, HHI_buffer.cpp) compiled
 *                       from the reference sources into oracle/_ref with
 *                       ora_exp/ora_log linked in place of libm's, fed the same
 *                       noise streams; within spike tolerance with glibc exp.
 *   Poisson_noise formulas: against the reference poisson_noise.cpp driven by
 *                       the same mt19937 stream.
 *   trial / STDP      : the reference trial (NGS:8839-9571, 12747-13118) needs
 *                       Boost and libconfig++, which this image lacks, so it is
 *                       not built; the restatement is pinned by known answers
 *                       derived by hand from the NGS lines on networks of one to
 *                       six neurons (kick step and size, arrival step and sync of
 *                       every connection type, zero delay, LTP/LTD in every
 *                       branch with exact pre-event and sync-clock post-event
 *                       times, the update_synapse truth table, axon_remodeling,
 *                       potentiation_decay; tests/test_oracle_pins.py), by an
 *                       independent Python restatement on small networks, and by
 *                       invariants (tests/test_oracle_trial.py).
 */
#ifndef HVC_ORACLE_H
#define HVC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- numerics ---------------- */
double ora_exp(double x);
double ora_log(double x);
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Noise streams (counter-based; one Philox call per (neuron, step) for the
 * four Debraband indices, one u32 per Poisson draw). */
#define ORA_STREAM_RA_WHITE 0x52410000u
#define ORA_STREAM_I_POISSON 0x49000000u
#define ORA_STREAM_RA_POISSON 0x52500000u
void ora_white_indices(uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                       uint32_t neuron, uint32_t step, int idx[4]);
uint32_t ora_poisson_u32(uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                         uint32_t stream, uint32_t neuron, uint32_t draw);

/* ---------------- single-neuron traces ----------------
 * model RA: 0 = HH2_test + Debraband_step (drop-in), 1 = HH2_buffer + Runge4_step
 * model I : 0 = HHI_buffer + DP8_step (drop-in),    1 = HHI_buffer + Runge4_step
 * Step s (1-based) of the trace: kick[s-1] (raiseE) is applied before the step,
 * then <model>_step_no_target_update, then exc/inh inputs of row s-1 are
 * applied (raiseExcWeight / raiseI for RA, raiseE / raiseI for I).
 * out_state row 0 is the state before step 1.
 * first_trial != 0 reproduces the initialisation order of
 * NetworkGrowthSimulator::initialize_network (NGS:849-850: set_noise runs
 * before set_dynamics, so the HVC-I noise initialisation sees timestep 0). */
#define ORA_RA_NSTATE 12 /* Vs Vd n l h r c Ca Ge_s Gi_s Ge_d Gi_d */
#define ORA_I_NSTATE 7   /* V n m h w Ge Gi */
int ora_ra_trace(int model, int n, int n_steps, const double* erest, const double* gca,
                 double mu_s, double sigma_s, double mu_d, double sigma_d, int poisson,
                 uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                 const double* kick, const double* exc_w, const double* inh_g,
                 double* out_state, uint8_t* out_fired);
int ora_i_trace(int model, int n, int n_steps, int poisson, int first_trial,
                uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                const double* exc_g, const double* inh_g,
                double* out_state, uint8_t* out_fired);

/* ---------------- network trial ---------------- */
typedef struct {
    int Nss;
    double A_P, A_D, T_0, T_P, T_D, TAU_P, TAU_D;
    double BETA, BETA_SUPERSYNAPSE;
    double ACTIVATION_THRESHOLD, SUPERSYNAPSE_THRESHOLD, WEIGHT_MAX;
} ora_synaptic_params; /* ConfigurationNetworkGrowth.h:11-34 */

typedef struct {
    int n_ra, n_i;
    int ra_model, i_model;        /* see traces above */
    double dt;                    /* TIMESTEP, NGS:15 */
    double trial_ms;              /* TRIAL_DURATION, NGS:17 */
    double sync_ms;               /* NETWORK_UPDATE_FREQUENCY, NGS:16 */
    double mu_s, sigma_s, mu_d, sigma_d;   /* noiseParameters */
    int ra_poisson;               /* RA Poisson conductance noise (RK4 form only) */
    int i_poisson;                /* HVC-I Poisson noise (NGS:2061-2075: on) */
    uint32_t seed_lo, seed_hi;
    /* RA->I, I->RA: CSR by presynaptic neuron (syn_ID_*, weights_*, axonal_delays_*) */
    const int* ra2i_ptr; const int* ra2i_tgt; const double* ra2i_w; const double* ra2i_d;
    const int* i2ra_ptr; const int* i2ra_tgt; const double* i2ra_w; const double* i2ra_d;
    /* RA->RA candidate set: CSR by presynaptic neuron, targets ascending.
     * The dense reference (NGS.h:142-143) is the candidate set of all j. */
    const int* rr_ptr; const int* rr_tgt; const double* rr_d;
    const double* erest;          /* per RA: EsL = EdL (set_Erest) */
    const double* gca;            /* per RA: GdCa (set_GCa) */
    int n_train; const int* train_ids; const double* spread; /* per training neuron */
    ora_synaptic_params syn;
} ora_net_desc;

typedef struct ora_sim ora_sim;

ora_sim* ora_create(const ora_net_desc* d);
void ora_destroy(ora_sim* s);
/* Synapse state: weights per candidate edge, active/super lists (insertion order), remodeled. */
int ora_set_synapses(ora_sim* s, const double* w, const int* act_cnt, const int* act_ids,
                     const int* sup_cnt, const int* sup_ids, const int* remodeled);
int ora_get_synapses(const ora_sim* s, double* w, int* act_cnt, int* act_ids,
                     int* sup_cnt, int* sup_ids, int* remodeled);
int ora_set_neuron_params(ora_sim* s, const double* erest, const double* gca);
/* mode 0: growth trial (NGS:8839-9571), mode 1: chain-test trial (NGS:12747-13118).
 * Returns 0, or -1 when check_bad_values flags an RA neuron after the trial. */
int ora_run_trial(ora_sim* s, int mode, int training, uint32_t trial_index, int first_trial,
                  double mean_injection_time);
/* Spikes of the last trial: per-neuron counts then times (CSR). */
int ora_threads(void);   /* threads the neuron steps of a large network run on (OpenMP) */
void ora_set_threads(int n);
int ora_spike_counts(const ora_sim* s, int kind, int* counts); /* kind 0 soma,1 dend,2 I; returns total */
int ora_spike_times(const ora_sim* s, int kind, double* times);
/* Counters of reference diagnostics (LTP/LTD rejected dt, missing list entries). */
int ora_warnings(const ora_sim* s, int* out3);
/* Final neuron states of the last trial (RA: 12 per neuron, I: 7 per neuron). */
int ora_final_state(const ora_sim* s, double* ra, double* in);

/* Probe for tests: with out != NULL, the next trials write state variable var
 * of one neuron (kind 0 RA, var 0..11 as ORA_RA_NSTATE; kind 1 HVC-I, var 0..6)
 * into out[t] after every step t (after that step's synchronisation, if any);
 * out[0] is the value at the start of the trial.  out holds num_steps values. */
int ora_set_probe(ora_sim* s, int kind, int neuron, int var, double* out);
/* The synapse-state machine on the current state, for known-answer tests:
 * update_synapse(i, j) (NGS:16546-16592) and axon_remodeling(i) (NGS:16363-16415).
 * Return -1 when (i, j) is not a candidate edge. */
int ora_update_synapse(ora_sim* s, int i, int j);
int ora_axon_remodeling(ora_sim* s, int i);

/* STDP primitives, exposed for known-answer tests (NGS:16615-16652). */
double ora_ltp(const ora_synaptic_params* p, double w, double dt, int* rejected);
double ora_ltd(const ora_synaptic_params* p, double w, double dt, int* rejected);

/* Clock helpers: internal_time after each step (NGS:8925) and the sync steps
 * (NGS:9291, 9563).  Returns num_steps; fills clock[0..num_steps-1]. */
int ora_num_steps(double trial_ms, double dt);
int ora_clock(double trial_ms, double dt, double sync_ms, double* clock, int* is_sync);

#ifdef __cplusplus
}
#endif
#endif
