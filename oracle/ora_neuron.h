/* ora_neuron.h -- oracle neuron models (test infrastructure, see hvc_oracle.h). */
#ifndef ORA_NEURON_H
#define ORA_NEURON_H
#include <stdint.h>

typedef struct {
    uint32_t seed_lo, seed_hi, trial, neuron;
    uint32_t draw;      /* next Poisson draw index */
} ora_stream;

/* HVC-RA neuron: HH2_test (model 0) or HH2_buffer (model 1) */
typedef struct {
    int model;
    double Vs, Vd, n, l, h, r, c, Ca;
    double Ge_s, Gi_s, Ge_d, Gi_d;
    double time, dt;
    int flag_soma, flag_dend, fired_soma, fired_dend;
    double EL, GCa;                       /* EsL = EdL, GdCa */
    double mu_s, sigma_s, mu_d, sigma_d;
    int poisson;                          /* HH2_buffer Poisson conductance noise */
    double noise_exc_soma, noise_inh_soma, noise_exc_dend, noise_inh_dend;
    ora_stream st;
    uint32_t step;                        /* trial step of the next Debraband step */
} ora_ra_neuron;

/* HVC-I neuron: HHI_buffer, DP8 (model 0) or RK4 (model 1) */
typedef struct {
    int model;
    double V, n, m, h, w, Ge, Gi;
    double time, dt;
    int flag, fired;
    int poisson;
    double noise_exc, noise_inh;
    ora_stream st;
} ora_i_neuron;

void ora_ra_init(ora_ra_neuron* a, int model, double dt, double EL, double GCa,
                 double mu_s, double sigma_s, double mu_d, double sigma_d, int poisson);
void ora_ra_set_to_rest(ora_ra_neuron* a);           /* HH2_test.cpp:191-242 */
void ora_ra_step(ora_ra_neuron* a);                  /* *_step_no_target_update */
int ora_ra_check_bad_values(const ora_ra_neuron* a);  /* HH2_test.cpp:244-260 */
void ora_ra_get_state(const ora_ra_neuron* a, double* s12);

void ora_i_init(ora_i_neuron* b, int model, double dt, int poisson);
void ora_i_set_to_rest(ora_i_neuron* b, double init_dt);  /* HHI_buffer.cpp:178-200 */
void ora_i_step(ora_i_neuron* b);                    /* *_step_no_target_update */
void ora_i_get_state(const ora_i_neuron* b, double* s7);

#endif
