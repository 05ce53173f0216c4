/*
 * hvc_topology.h -- C ABI of the network topology generator (host code).
 *
 * Restates the reference's topology sampling (paths relative to the reference
 * checkout, "NGS" = c++/src/NetworkGrowthSimulator.cpp):
 *   generate_network_topology (NGS:110-155): sample_coordinates (NGS:222-403),
 *   sample_connections (NGS:620-705) and the two training sets of
 *   write_different_training_arrangements (NGS:406-455);
 *   the preparation new_chain_growth applies before growth (NGS:1882-1885):
 *   sample_axonal_delays (NGS:466-526) and resample_weights (NGS:721-737).
 * Random numbers come from the generator the reference uses, std::mt19937
 * (Poisson_noise, c++/src/poisson_noise.cpp:3-21): random(G) = G * u / max;
 * the random training set draws std::uniform_int_distribution (libstdc++).
 * The reference seeds rank 0 with 15000 (NGS:2443); with that seed the
 * generator gives the reference's own topology for the same sizes and
 * parameters.
 *
 * Units: coordinates on the unit sphere (or in the unit square); connection
 * lengths in microns (distance * EXPERIMENTAL_INTERNEURON_DISTANCE /
 * MODEL_INTERNEURON_DISTANCE, NGS:537); delays in ms.
 */
#ifndef HVC_TOPOLOGY_H
#define HVC_TOPOLOGY_H

#include <stdint.h>

#include "hvc_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* TopologyParameters (c++/include/ConfigurationNetworkTopology.h:11-22) and the
 * ConnectionParameters (c++/include/ConfigurationNetworkGrowth.h:97-105). */
typedef struct {
    int arrangement;                 /* 0: "sphere" (NGS:294-401), 1: "square" (NGS:228-292) */
    double A_RA2I, SIGMA_RA2I;       /* p_RA2I(d) = A exp(-d^2 / (2 SIGMA^2)), NGS:707-712 */
    double B_I2RA, SIGMA_I2RA;       /* p_I2RA(d), NGS:714-719 */
    double Gei_max, Gie_max;         /* weights U(0,1) * G_max (sample_G, NGS:739-742) */
    double delay_constant;           /* delay = length * 0.01 ms/micron * delay_constant (NGS:25, 476) */
} hvc_topology_params;

/* A generated network; arrays are owned by the library (hvc_topology_free).
 * Connections are lists per presynaptic neuron in the reference's order,
 * stored as CSR (ptr has n_pre + 1 entries). */
typedef struct {
    int n_ra, n_i, n_tr;
    int dim;                             /* 3 (sphere) or 2 (square) */
    double model_interneuron_distance;   /* MODEL_INTERNEURON_DISTANCE (NGS:231, 312) */
    double* ra_xyz; double* i_xyz;       /* 3 per neuron (z = 0 for the square) */
    int* ra2i_ptr; int* ra2i_tgt; double* ra2i_w; double* ra2i_len; double* ra2i_d;
    int* i2ra_ptr; int* i2ra_tgt; double* i2ra_w; double* i2ra_len; double* i2ra_d;
    int* training_clustered;             /* training_neurons_clustered.bin (NGS:412-437) */
    int* training_random;                /* training_neurons_random.bin (NGS:439-454) */
} hvc_topology;

/* generate_network_topology with a generator seeded `seed`: positions,
 * connections with weights 0 and delays 0, and both training sets. */
hvc_status hvc_topology_generate(int n_ra, int n_i, int n_tr, const hvc_topology_params* p, uint32_t seed,
                                 hvc_topology** out);

/* What new_chain_growth does to a read topology, with a fresh generator seeded
 * `seed`: sample_axonal_delays (RA->I and I->RA delays from the lengths; RA->RA
 * delays of all pairs into rr_delay[n_ra * n_ra] unless NULL) and then
 * resample_weights (I->RA weights first, then RA->I).  Sphere arrangement only
 * for rr_delay (the reference measures RA->RA lengths on the sphere). */
hvc_status hvc_topology_prepare(hvc_topology* t, const hvc_topology_params* p, uint32_t seed, double* rr_delay);

void hvc_topology_free(hvc_topology* t);

/* The generator itself: n draws of std::mt19937 seeded `seed` (after `skip`
 * draws), as raw 32-bit words.  Lets a caller check the stream. */
hvc_status hvc_mt19937(uint32_t seed, long long skip, int n, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
