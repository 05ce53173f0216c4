// ref_noise_api.cpp -- test harness only: exposes the reference's real
// Poisson_noise (c++/src/poisson_noise.cpp, linked unmodified) and the raw
// std::mt19937 stream with the same seed, so tests can check the replay
// formulas for get_spike_time / random against the reference implementation.
#include <cstdint>
#include <random>

#include "poisson_noise.h"

extern "C" int refnoise_draws(unsigned seed, int kind, double arg, int n, double* out)
{
    Poisson_noise p;
    p.set_seed(seed);
    for (int k = 0; k < n; ++k) {
        if (kind == 0)
            out[k] = p.random(arg);
        else if (kind == 1)
            out[k] = p.get_spike_time(arg);
        else
            out[k] = (double)p.sample_index_for_point_distribution();
    }
    return 0;
}

extern "C" int refnoise_mt_raw(unsigned seed, int n, uint32_t* out)
{
    std::mt19937 g;
    g.seed(seed);
    for (int k = 0; k < n; ++k) out[k] = (uint32_t)g();
    return 0;
}

// The reference's Poisson_noise::random(G[k]) for a sequence of G, one draw each.
extern "C" int refnoise_random_seq(unsigned seed, int n, const double* G, double* out)
{
    Poisson_noise p;
    p.set_seed(seed);
    for (int k = 0; k < n; ++k) out[k] = p.random(G[k]);
    return 0;
}

// The reference's Poisson_noise::sample_integer(0, hi[k]) for a sequence of hi.
extern "C" int refnoise_sample_integer(unsigned seed, int n, const int* hi, int* out)
{
    Poisson_noise p;
    p.set_seed(seed);
    for (int k = 0; k < n; ++k) out[k] = p.sample_integer(0, hi[k]);
    return 0;
}
