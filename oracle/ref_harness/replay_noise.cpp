// replay_noise.cpp -- test harness only: a replay definition of the reference's
// Poisson_noise member functions (declared in c++/include/poisson_noise.h),
// linked into oracle/_ref in place of c++/src/poisson_noise.cpp so that the
// reference neuron classes consume the same counter-based noise streams as the
// oracle and the GPU engine.  Each instance is bound to one neuron's stream by
// ref_bind_stream(); the formulas of get_spike_time and random restate
// poisson_noise.cpp:9-21 and are checked against the real poisson_noise.cpp
// driven by std::mt19937 (see ref_noise_api.cpp).
#include "poisson_noise.h"

#include <cmath>
#include <cstdint>
#include <unordered_map>

#include "../hvc_oracle.h"

namespace {
struct Stream {
    uint32_t seed_lo, seed_hi, trial, neuron, poisson_stream;
    uint32_t index_calls;   // calls of sample_index_for_point_distribution
    uint32_t draws;         // u32 draws for random / get_spike_time
};
std::unordered_map<const Poisson_noise*, Stream>& registry()
{
    static std::unordered_map<const Poisson_noise*, Stream> r;
    return r;
}
uint32_t next_u32(const Poisson_noise* p)
{
    Stream& s = registry().at(p);
    return ora_poisson_u32(s.seed_lo, s.seed_hi, s.trial, s.poisson_stream, s.neuron, s.draws++);
}
}  // namespace

void ref_bind_stream(const Poisson_noise* p, uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                     uint32_t neuron, uint32_t poisson_stream)
{
    registry()[p] = Stream{seed_lo, seed_hi, trial, neuron, poisson_stream, 0u, 0u};
}

void ref_unbind_stream(const Poisson_noise* p) { registry().erase(p); }

void Poisson_noise::set_seed(unsigned s) { seed = s; }

double Poisson_noise::get_spike_time(double lambda)
{
    double random = (double)next_u32(this) / 4294967295.0;
    return (-log(1 - random) / lambda);
}

double Poisson_noise::random(double G) { return (double)G * (double)next_u32(this) / 4294967295.0; }

int Poisson_noise::sample_index_for_point_distribution()
{
    Stream& s = registry().at(this);
    uint32_t c = s.index_calls++;
    int idx[4];
    ora_white_indices(s.seed_lo, s.seed_hi, s.trial, s.neuron, c / 4 + 1, idx);
    return idx[c % 4];
}

void Poisson_noise::set_normal_distribution(double mu, double sigma)
{
    mean = mu;
    sd = sigma;
}

double Poisson_noise::normal_distribution() { return mean + 0.0 * sd; }
