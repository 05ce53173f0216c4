/*
 * ora_trace.c -- single-neuron traces and the trial clock (test infrastructure,
 * see hvc_oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "hvc_oracle.h"
#include "ora_neuron.h"

int ora_ra_trace(int model, int n, int n_steps, const double* erest, const double* gca,
                 double mu_s, double sigma_s, double mu_d, double sigma_d, int poisson,
                 uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                 const double* kick, const double* exc_w, const double* inh_g,
                 double* out_state, uint8_t* out_fired)
{
    if (n <= 0 || n_steps < 0 || (model != 0 && model != 1)) return -1;
    ora_ra_neuron* a = (ora_ra_neuron*)calloc((size_t)n, sizeof(ora_ra_neuron));
    if (!a) return -1;
    for (int k = 0; k < n; ++k) {
        ora_ra_init(&a[k], model, dt, erest[k], gca[k], mu_s, sigma_s, mu_d, sigma_d, poisson);
        a[k].st.seed_lo = seed_lo; a[k].st.seed_hi = seed_hi;
        a[k].st.trial = trial; a[k].st.neuron = (uint32_t)k; a[k].st.draw = 0;
        ora_ra_set_to_rest(&a[k]);
        ora_ra_get_state(&a[k], out_state + (size_t)k * ORA_RA_NSTATE);
    }
    for (int s = 1; s <= n_steps; ++s) {
        size_t row = (size_t)(s - 1) * (size_t)n;
        for (int k = 0; k < n; ++k) {
            ora_ra_neuron* x = &a[k];
            if (kick && kick[row + k] != 0.0) x->Ge_d += kick[row + k];           /* raiseE */
            ora_ra_step(x);
            out_fired[row + k] = (uint8_t)((x->fired_soma ? 1 : 0) | (x->fired_dend ? 2 : 0));
            if (inh_g && inh_g[row + k] != 0.0) x->Gi_d += inh_g[row + k];        /* raiseI */
            if (exc_w && exc_w[row + k] != 0.0) x->Ge_d += exc_w[row + k] / 10000.0; /* raiseExcWeight */
            ora_ra_get_state(x, out_state + ((size_t)s * n + k) * ORA_RA_NSTATE);
        }
    }
    free(a);
    return 0;
}

int ora_i_trace(int model, int n, int n_steps, int poisson, int first_trial,
                uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                const double* exc_g, const double* inh_g,
                double* out_state, uint8_t* out_fired)
{
    if (n <= 0 || n_steps < 0 || (model != 0 && model != 1)) return -1;
    ora_i_neuron* b = (ora_i_neuron*)calloc((size_t)n, sizeof(ora_i_neuron));
    if (!b) return -1;
    for (int k = 0; k < n; ++k) {
        ora_i_init(&b[k], model, dt, poisson);
        b[k].st.seed_lo = seed_lo; b[k].st.seed_hi = seed_hi;
        b[k].st.trial = trial; b[k].st.neuron = (uint32_t)k; b[k].st.draw = 0;
        ora_i_set_to_rest(&b[k], first_trial ? 0.0 : dt);
        ora_i_get_state(&b[k], out_state + (size_t)k * ORA_I_NSTATE);
    }
    for (int s = 1; s <= n_steps; ++s) {
        size_t row = (size_t)(s - 1) * (size_t)n;
        for (int k = 0; k < n; ++k) {
            ora_i_neuron* x = &b[k];
            ora_i_step(x);
            out_fired[row + k] = (uint8_t)(x->fired ? 1 : 0);
            if (inh_g && inh_g[row + k] != 0.0) x->Gi += inh_g[row + k];          /* raiseI */
            if (exc_g && exc_g[row + k] != 0.0) x->Ge += exc_g[row + k];          /* raiseE */
            ora_i_get_state(x, out_state + ((size_t)s * n + k) * ORA_I_NSTATE);
        }
    }
    free(b);
    return 0;
}

/* num_steps = static_cast<int>(TRIAL_DURATION / TIMESTEP), NGS:8920 */
int ora_num_steps(double trial_ms, double dt) { return (int)(trial_ms / dt); }

/* NGS:8906-8925, 9291, 9563: internal_time starts at 0 and gains TIMESTEP before
 * each step t = 1..num_steps-1; a sync follows step t when internal_time >
 * network_time (which starts at NETWORK_UPDATE_FREQUENCY and grows by it after
 * each sync) or t == num_steps-1.  clock[0] = 0. */
int ora_clock(double trial_ms, double dt, double sync_ms, double* clock, int* is_sync)
{
    int num_steps = ora_num_steps(trial_ms, dt);
    double internal_time = 0;
    double network_time = sync_ms;
    if (clock) clock[0] = 0.0;
    if (is_sync) is_sync[0] = 0;
    for (int t = 1; t < num_steps; ++t) {
        internal_time += dt;
        if (clock) clock[t] = internal_time;
        int s = (internal_time > network_time) || (t == num_steps - 1);
        if (is_sync) is_sync[t] = s;
        if (s) network_time += sync_ms;
    }
    return num_steps;
}
