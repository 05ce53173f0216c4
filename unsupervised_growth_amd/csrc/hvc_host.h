// hvc_host.h -- host-side helpers shared by the topology generator and the
// growth driver: the reference's random-number calls and distance functions.
// Host code only; compiled with -ffp-contract=off so every expression rounds
// operation by operation as written.
#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <string>

namespace hvc_host {

// Poisson_noise::random (c++/src/poisson_noise.cpp:17-20): (double)G * x / max,
// evaluated left to right, x one std::mt19937 word.
struct RefRng {
    std::mt19937 g;
    explicit RefRng(uint32_t seed) : g(seed) {}
    double random(double G) { return G * static_cast<double>(g()) / static_cast<double>(std::mt19937::max()); }
    // Poisson_noise::sample_integer (c++/include/poisson_noise.h:26)
    int sample_integer(int lo, int hi) { return std::uniform_int_distribution<int>{lo, hi}(g); }
};

// distance_on_sphere (NGS:36-44): great-circle distance on the unit sphere
inline double distance_on_sphere(double x1, double y1, double z1, double x2, double y2, double z2)
{
    double c = x1 * x2 + y1 * y2 + z1 * z2;
    return c > 1.0 ? 0.0 : std::acos(c);
}

// distance2d (NGS:31-34)
inline double distance2d(double x1, double y1, double x2, double y2)
{
    return std::sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
}

// Constants of NetworkGrowthSimulator (NGS:14-25).
constexpr double kWaitingTime = 100.0;
constexpr double kExperimentalInterneuronDistance = 40.0;   // microns
constexpr double kMinIntersomaticDistance = 10.0;           // microns
constexpr double kAxonalDelayPerMicron = 0.01;              // ms
constexpr int kMaturationScaleSpontaneous = 100000;
constexpr int kMaturationScaleDriven = 1000;
constexpr double kPi = 3.14159265358979323846;

}  // namespace hvc_host
