// ref_api.cpp -- test harness only: drives the reference neuron classes
// (HH2_test, HH2_buffer, HHI_buffer, compiled unmodified from the reference
// sources) through the same single-neuron trace protocol as ora_ra_trace /
// ora_i_trace in oracle/ora_trace.c, for bit-level comparison.
#include <cstdint>
#include <cstring>
#include <vector>

#include "HH2_buffer.h"
#include "HH2_test.h"
#include "HHI_buffer.h"
#include "poisson_noise.h"
#include "../hvc_oracle.h"

void ref_bind_stream(const Poisson_noise* p, uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                     uint32_t neuron, uint32_t poisson_stream);
void ref_unbind_stream(const Poisson_noise* p);

namespace {
// Derived probes read the protected state; they add no state or behaviour.
struct RAProbeTest : public HH2_test {
    void dump(double* s) const
    {
        s[0] = _Vs; s[1] = _Vd; s[2] = _n; s[3] = _l; s[4] = _h; s[5] = _r;
        s[6] = _c; s[7] = _Ca; s[8] = _Gexc_s; s[9] = _Ginh_s; s[10] = _Gexc_d; s[11] = _Ginh_d;
    }
};
struct RAProbeBuffer : public HH2_buffer {
    void dump(double* s) const
    {
        s[0] = _Vs; s[1] = _Vd; s[2] = _n; s[3] = 0.0; s[4] = _h; s[5] = _r;
        s[6] = _c; s[7] = _Ca; s[8] = _Gexc_s; s[9] = _Ginh_s; s[10] = _Gexc_d; s[11] = _Ginh_d;
    }
};
struct IProbe : public HHI_buffer {
    void dump(double* s) const
    {
        s[0] = _voltage; s[1] = _n; s[2] = _m; s[3] = _h; s[4] = _w; s[5] = _Gexc; s[6] = _Ginh;
    }
};

template <class N>
int run_ra(int n, int n_steps, int model, const double* erest, const double* gca, double mu_s,
           double sigma_s, double mu_d, double sigma_d, int poisson, uint32_t seed_lo,
           uint32_t seed_hi, uint32_t trial, double dt, const double* kick, const double* exc_w,
           const double* inh_g, double* out_state, uint8_t* out_fired)
{
    std::vector<N> ns(n);
    std::vector<Poisson_noise> gens(n);
    for (int k = 0; k < n; ++k) {
        ref_bind_stream(&gens[k], seed_lo, seed_hi, trial, (uint32_t)k,
                        ORA_STREAM_RA_POISSON);
        ns[k].set_noise_generator(&gens[k]);
        ns[k].set_dynamics(dt);
        ns[k].set_Erest(erest[k]);
        ns[k].set_GCa(gca[k]);
        ns[k].set_white_noise(mu_s, sigma_s, mu_d, sigma_d);
        if (poisson) ns[k].set_poisson_noise();
        ns[k].dump(out_state + (size_t)k * ORA_RA_NSTATE);
    }
    for (int s = 1; s <= n_steps; ++s) {
        size_t row = (size_t)(s - 1) * (size_t)n;
        for (int k = 0; k < n; ++k) {
            N& x = ns[k];
            if (kick && kick[row + k] != 0.0) x.raiseE(kick[row + k]);
            if (model == 0)
                x.Debraband_step_no_target_update();
            else
                x.R4_step_no_target_update();
            out_fired[row + k] = (uint8_t)((x.get_fired_soma() ? 1 : 0) | (x.get_fired_dend() ? 2 : 0));
            if (inh_g && inh_g[row + k] != 0.0) x.raiseI(inh_g[row + k]);
            if (exc_w && exc_w[row + k] != 0.0) x.raiseExcWeight(exc_w[row + k]);
            x.dump(out_state + ((size_t)s * n + k) * ORA_RA_NSTATE);
        }
    }
    for (int k = 0; k < n; ++k) ref_unbind_stream(&gens[k]);
    return 0;
}
}  // namespace

extern "C" int ref_ra_trace(int model, int n, int n_steps, const double* erest, const double* gca,
                            double mu_s, double sigma_s, double mu_d, double sigma_d, int poisson,
                            uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                            const double* kick, const double* exc_w, const double* inh_g,
                            double* out_state, uint8_t* out_fired)
{
    if (n <= 0 || n_steps < 0) return -1;
    if (model == 0)
        return run_ra<RAProbeTest>(n, n_steps, 0, erest, gca, mu_s, sigma_s, mu_d, sigma_d, poisson,
                                   seed_lo, seed_hi, trial, dt, kick, exc_w, inh_g, out_state, out_fired);
    if (model == 1)
        return run_ra<RAProbeBuffer>(n, n_steps, 1, erest, gca, mu_s, sigma_s, mu_d, sigma_d, poisson,
                                     seed_lo, seed_hi, trial, dt, kick, exc_w, inh_g, out_state, out_fired);
    return -1;
}

extern "C" int ref_i_trace(int model, int n, int n_steps, int poisson, int first_trial,
                           uint32_t seed_lo, uint32_t seed_hi, uint32_t trial, double dt,
                           const double* exc_g, const double* inh_g, double* out_state,
                           uint8_t* out_fired)
{
    if (n <= 0 || n_steps < 0 || (model != 0 && model != 1)) return -1;
    std::vector<IProbe> ns(n);
    std::vector<Poisson_noise> gens(n);
    for (int k = 0; k < n; ++k) {
        ref_bind_stream(&gens[k], seed_lo, seed_hi, trial, (uint32_t)k, ORA_STREAM_I_POISSON);
        ns[k].set_noise_generator(&gens[k]);
        // NGS:849-850: set_noise (which enables Poisson noise) precedes set_dynamics
        if (first_trial) {
            if (poisson) ns[k].set_poisson_noise();
            ns[k].set_dynamics(dt);
        } else {
            ns[k].set_dynamics(dt);
            if (poisson) ns[k].set_poisson_noise();
        }
        ns[k].dump(out_state + (size_t)k * ORA_I_NSTATE);
    }
    for (int s = 1; s <= n_steps; ++s) {
        size_t row = (size_t)(s - 1) * (size_t)n;
        for (int k = 0; k < n; ++k) {
            IProbe& x = ns[k];
            if (model == 0)
                x.DP8_step_no_target_update();
            else
                x.R4_step_no_target_update();
            out_fired[row + k] = (uint8_t)(x.get_fired() ? 1 : 0);
            if (inh_g && inh_g[row + k] != 0.0) x.raiseI(inh_g[row + k]);
            if (exc_g && exc_g[row + k] != 0.0) x.raiseE(exc_g[row + k]);
            x.dump(out_state + ((size_t)s * n + k) * ORA_I_NSTATE);
        }
    }
    for (int k = 0; k < n; ++k) ref_unbind_stream(&gens[k]);
    return 0;
}
