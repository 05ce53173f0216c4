/*
 * ora_seam.c -- the neuron seam of the oracle, one call per neuron per step
 * (test infrastructure, see hvc_oracle.h).  Lets an independent restatement of
 * the trial logic (tests/pyref_trial.py) drive the oracle's neuron models the
 * way NetworkGrowthSimulator drives HH2_test / HHI_buffer objects
 * (Debraband_step_no_target_update, DP8_step_no_target_update, get_fired_*,
 * raiseE / raiseI / raiseExcWeight, set_to_rest).
 */
#include <stdlib.h>

#include "hvc_oracle.h"
#include "ora_neuron.h"

typedef struct {
    int n_ra, n_i;
    ora_ra_neuron* ra;
    ora_i_neuron* in;
    double dt;
    int ra_model, i_model, ra_poisson, i_poisson;
    double mu_s, sigma_s, mu_d, sigma_d;
    uint32_t seed_lo, seed_hi;
    double *erest, *gca;
} ora_neurons;

void* ora_neurons_create(int n_ra, int n_i, int ra_model, int i_model, double dt, double mu_s, double sigma_s,
                         double mu_d, double sigma_d, int ra_poisson, int i_poisson, uint32_t seed_lo,
                         uint32_t seed_hi, const double* erest, const double* gca)
{
    ora_neurons* h = (ora_neurons*)calloc(1, sizeof(ora_neurons));
    h->n_ra = n_ra; h->n_i = n_i; h->dt = dt;
    h->ra_model = ra_model; h->i_model = i_model; h->ra_poisson = ra_poisson; h->i_poisson = i_poisson;
    h->mu_s = mu_s; h->sigma_s = sigma_s; h->mu_d = mu_d; h->sigma_d = sigma_d;
    h->seed_lo = seed_lo; h->seed_hi = seed_hi;
    h->ra = (ora_ra_neuron*)calloc((size_t)(n_ra ? n_ra : 1), sizeof(ora_ra_neuron));
    h->in = (ora_i_neuron*)calloc((size_t)(n_i ? n_i : 1), sizeof(ora_i_neuron));
    h->erest = (double*)malloc(sizeof(double) * (size_t)(n_ra ? n_ra : 1));
    h->gca = (double*)malloc(sizeof(double) * (size_t)(n_ra ? n_ra : 1));
    for (int i = 0; i < n_ra; ++i) { h->erest[i] = erest[i]; h->gca[i] = gca[i]; }
    return h;
}

void ora_neurons_destroy(void* p)
{
    ora_neurons* h = (ora_neurons*)p;
    if (!h) return;
    free(h->ra); free(h->in); free(h->erest); free(h->gca); free(h);
}

/* set_to_rest of every neuron, streams of trial `trial` */
void ora_neurons_reset(void* p, uint32_t trial, int first_trial)
{
    ora_neurons* h = (ora_neurons*)p;
    for (int i = 0; i < h->n_ra; ++i) {
        ora_ra_neuron* a = &h->ra[i];
        ora_ra_init(a, h->ra_model, h->dt, h->erest[i], h->gca[i], h->mu_s, h->sigma_s, h->mu_d, h->sigma_d,
                    h->ra_poisson);
        a->st.seed_lo = h->seed_lo; a->st.seed_hi = h->seed_hi; a->st.trial = trial;
        a->st.neuron = (uint32_t)i; a->st.draw = 0;
        ora_ra_set_to_rest(a);
    }
    for (int i = 0; i < h->n_i; ++i) {
        ora_i_neuron* b = &h->in[i];
        ora_i_init(b, h->i_model, h->dt, h->i_poisson);
        b->st.seed_lo = h->seed_lo; b->st.seed_hi = h->seed_hi; b->st.trial = trial;
        b->st.neuron = (uint32_t)i; b->st.draw = 0;
        ora_i_set_to_rest(b, first_trial ? 0.0 : h->dt);
    }
}

/* one step; returns fired bits (RA: 1 soma, 2 dend; I: 1) */
int ora_neurons_step(void* p, int kind, int i)
{
    ora_neurons* h = (ora_neurons*)p;
    if (kind == 0) {
        ora_ra_step(&h->ra[i]);
        return (h->ra[i].fired_soma ? 1 : 0) | (h->ra[i].fired_dend ? 2 : 0);
    }
    ora_i_step(&h->in[i]);
    return h->in[i].fired ? 1 : 0;
}

/* which: 0 raiseE, 1 raiseI, 2 raiseExcWeight (RA only) */
void ora_neurons_raise(void* p, int kind, int i, int which, double g)
{
    ora_neurons* h = (ora_neurons*)p;
    if (kind == 0) {
        if (which == 0) h->ra[i].Ge_d += g;
        else if (which == 1) h->ra[i].Gi_d += g;
        else h->ra[i].Ge_d += g / 10000.0;
    } else {
        if (which == 0) h->in[i].Ge += g;
        else h->in[i].Gi += g;
    }
}

void ora_neurons_get(void* p, int kind, int i, double* out)
{
    ora_neurons* h = (ora_neurons*)p;
    if (kind == 0) ora_ra_get_state(&h->ra[i], out);
    else ora_i_get_state(&h->in[i], out);
}

int ora_neurons_bad(void* p, int i)
{
    ora_neurons* h = (ora_neurons*)p;
    return ora_ra_check_bad_values(&h->ra[i]);
}
