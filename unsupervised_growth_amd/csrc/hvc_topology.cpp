// hvc_topology.cpp -- network topology generator (host code).
//
// Restates generate_network_topology (NGS:110-155) and the preparation of
// new_chain_growth (NGS:1882-1885) with the reference's generator, so that a
// seed gives the reference's own network: same draw order, same expression
// trees.  "NGS" = c++/src/NetworkGrowthSimulator.cpp of the reference.
#include "../../include/hvc_topology.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "hvc_host.h"

using namespace hvc_host;

namespace {

struct Lists {
    std::vector<std::vector<int>> tgt;
    std::vector<std::vector<double>> w, len, d;
    explicit Lists(int n) : tgt(n), w(n), len(n), d(n) {}
};

template <class T>
T* dup(const std::vector<T>& v)
{
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

void to_csr(const Lists& L, int** ptr, int** tgt, double** w, double** len, double** d)
{
    std::vector<int> p(L.tgt.size() + 1, 0), t;
    std::vector<double> ww, ll, dd;
    for (size_t i = 0; i < L.tgt.size(); ++i) {
        p[i + 1] = p[i] + static_cast<int>(L.tgt[i].size());
        t.insert(t.end(), L.tgt[i].begin(), L.tgt[i].end());
        ww.insert(ww.end(), L.w[i].begin(), L.w[i].end());
        ll.insert(ll.end(), L.len[i].begin(), L.len[i].end());
        dd.insert(dd.end(), L.d[i].begin(), L.d[i].end());
    }
    *ptr = dup(p);
    *tgt = dup(t);
    *w = dup(ww);
    *len = dup(ll);
    *d = dup(dd);
}

// sample_coordinates, "sphere" (NGS:294-401): HVC-I on a jittered Fibonacci
// lattice, HVC-RA uniform on the sphere, none closer than too_close to another.
double sample_sphere(RefRng& rng, int n_ra, int n_i, std::vector<double>& ra, std::vector<double>& in)
{
    const double noise_shift = 0.15;  // NGS:226
    const double dphi = kPi * (3.0 - std::sqrt(5.0));
    const double dz = 2 / static_cast<double>(n_i);
    double phi = 0;
    double z = 1 - dz / 2.0;
    const double x1 = std::cos(phi) * std::sqrt(1.0 - z * z);
    const double y1 = std::sin(phi) * std::sqrt(1.0 - z * z);
    const double x2 = std::cos(phi + dphi) * std::sqrt(1.0 - (z - dz) * (z - dz));
    const double y2 = std::sin(phi + dphi) * std::sqrt(1.0 - (z - dz) * (z - dz));
    const double mid = distance_on_sphere(x1, y1, z, x2, y2, z - dz);
    const double too_close = mid * kMinIntersomaticDistance / kExperimentalInterneuronDistance;

    for (int i = 0; i < n_i; ++i) {
        double tx, ty, tz;
        bool close = true;
        while (close) {
            close = false;
            double theta = std::acos(z);
            double theta_noise = -noise_shift * mid + rng.random(2 * noise_shift * mid);
            double phi_noise = (-noise_shift * mid + rng.random(2 * noise_shift * mid)) / std::sin(theta);
            tx = std::cos(phi + phi_noise) * std::sin(theta + theta_noise);
            ty = std::sin(phi + phi_noise) * std::sin(theta + theta_noise);
            tz = std::cos(theta + theta_noise);
            for (int j = 0; j < i; ++j)
                if (distance_on_sphere(tx, ty, tz, in[3 * j], in[3 * j + 1], in[3 * j + 2]) < too_close) {
                    close = true;
                    break;
                }
        }
        in[3 * i] = tx;
        in[3 * i + 1] = ty;
        in[3 * i + 2] = tz;
        z = z - dz;
        phi = phi + dphi;
    }
    for (int i = 0; i < n_ra; ++i) {
        double tx, ty, tz;
        bool close = true;
        while (close) {
            close = false;
            z = 2 * rng.random(1.0) - 1;
            phi = rng.random(2.0 * kPi);
            tx = std::sin(std::acos(z)) * std::cos(phi);
            ty = std::sin(std::acos(z)) * std::sin(phi);
            tz = z;
            for (int j = 0; j < n_i; ++j)
                if (distance_on_sphere(tx, ty, tz, in[3 * j], in[3 * j + 1], in[3 * j + 2]) < too_close) {
                    close = true;
                    break;
                }
            if (!close)
                for (int j = 0; j < i; ++j)
                    if (distance_on_sphere(tx, ty, tz, ra[3 * j], ra[3 * j + 1], ra[3 * j + 2]) < too_close) {
                        close = true;
                        break;
                    }
        }
        ra[3 * i] = tx;
        ra[3 * i + 1] = ty;
        ra[3 * i + 2] = tz;
    }
    return mid;
}

// sample_coordinates, "square" (NGS:228-292).  HVC-I indices beyond the
// (int)sqrt(N_I)^2 lattice keep (0, 0), as in the reference.
double sample_square(RefRng& rng, int n_ra, int n_i, std::vector<double>& ra, std::vector<double>& in)
{
    const double noise_shift = 0.15;
    const double mid = 1 / (std::sqrt(n_i) + 1);
    const double too_close = mid * kMinIntersomaticDistance / kExperimentalInterneuronDistance;
    const int side = static_cast<int>(std::sqrt(n_i));
    for (int i = 0; i < side; ++i)
        for (int k = 0; k < side; ++k) {
            int idx = i * side + k;
            in[3 * idx] = static_cast<double>(i + 1) / (std::sqrt(n_i) + 1) - noise_shift * mid +
                          rng.random(2 * noise_shift * mid);
            in[3 * idx + 1] = static_cast<double>(k + 1) / (std::sqrt(n_i) + 1) - noise_shift * mid +
                              rng.random(2 * noise_shift * mid);
        }
    for (int i = 0; i < n_ra; ++i) {
        double tx, ty;
        bool close = true;
        while (close) {
            close = false;
            tx = 0.5 / (std::sqrt(n_i) + 1) + rng.random(1.0 - 1.0 / (std::sqrt(n_i) + 1));
            ty = 0.5 / (std::sqrt(n_i) + 1) + rng.random(1.0 - 1.0 / (std::sqrt(n_i) + 1));
            for (int j = 0; j < n_i; ++j)
                if (distance2d(tx, ty, in[3 * j], in[3 * j + 1]) < too_close) {
                    close = true;
                    break;
                }
            if (!close)
                for (int j = 0; j < i; ++j)
                    if (distance2d(tx, ty, ra[3 * j], ra[3 * j + 1]) < too_close) {
                        close = true;
                        break;
                    }
        }
        ra[3 * i] = tx;
        ra[3 * i + 1] = ty;
    }
    return mid;
}

inline double point_distance(int arrangement, const double* a, const double* b)
{
    return arrangement == 0 ? distance_on_sphere(a[0], a[1], a[2], b[0], b[1], b[2]) : distance2d(a[0], a[1], b[0], b[1]);
}

// p_RA2I / p_I2RA (NGS:707-719)
inline double gaussian_p(double peak, double sigma, double d)
{
    double sigma_square = sigma * sigma;
    return peak * std::exp(-d * d / (2 * sigma_square));
}

}  // namespace

extern "C" hvc_status hvc_topology_generate(int n_ra, int n_i, int n_tr, const hvc_topology_params* p, uint32_t seed,
                                            hvc_topology** out)
{
    if (!p || !out || n_ra <= 0 || n_i <= 0 || n_tr < 0 || n_tr > n_ra) return HVC_E_ARG;
    if (p->arrangement != 0 && p->arrangement != 1) return HVC_E_ARG;
    *out = nullptr;
    RefRng rng(seed);
    std::vector<double> ra(3 * n_ra, 0.0), in(3 * n_i, 0.0);
    double mid = p->arrangement == 0 ? sample_sphere(rng, n_ra, n_i, ra, in) : sample_square(rng, n_ra, n_i, ra, in);

    // sample_connections (NGS:620-705): weight sample_G(0.0) (a draw whose value
    // is 0), delay 0, length in microns
    const double scale_num = kExperimentalInterneuronDistance;
    Lists a(n_ra), b(n_i);
    for (int i = 0; i < n_ra; ++i)
        for (int j = 0; j < n_i; ++j) {
            double d = point_distance(p->arrangement, &ra[3 * i], &in[3 * j]) * scale_num / mid;
            if (rng.random(1) < gaussian_p(p->A_RA2I, p->SIGMA_RA2I, d)) {
                double G = rng.random(0.0);
                a.w[i].push_back(G);
                a.tgt[i].push_back(j);
                a.len[i].push_back(d);
                a.d[i].push_back(0.0);
            }
        }
    for (int i = 0; i < n_i; ++i)
        for (int j = 0; j < n_ra; ++j) {
            double d = point_distance(p->arrangement, &in[3 * i], &ra[3 * j]) * scale_num / mid;
            if (rng.random(1) < gaussian_p(p->B_I2RA, p->SIGMA_I2RA, d)) {
                double G = rng.random(0.0);
                b.w[i].push_back(G);
                b.tgt[i].push_back(j);
                b.len[i].push_back(d);
                b.d[i].push_back(0.0);
            }
        }

    // write_different_training_arrangements (NGS:406-455).  Clustered: neuron 0
    // and its n_tr - 1 nearest neurons (the reference measures on the sphere; the
    // square uses the plane distance).  Random: partial Fisher-Yates with
    // swap-with-last removal.
    std::vector<int> clustered(n_tr), random_set(n_tr);
    if (n_tr > 0) {
        std::vector<double> dist(n_ra - 1);
        for (int i = 1; i < n_ra; ++i) dist[i - 1] = point_distance(p->arrangement, &ra[0], &ra[3 * i]);
        std::vector<size_t> idx(dist.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&dist](size_t i1, size_t i2) { return dist[i1] < dist[i2]; });
        clustered[0] = 0;
        for (int i = 1; i < n_tr; ++i) clustered[i] = static_cast<int>(idx[i - 1] + 1);
        std::vector<int> unsampled(n_ra);
        std::iota(unsampled.begin(), unsampled.end(), 0);
        for (int i = 0; i < n_tr; ++i) {
            int r = rng.sample_integer(0, n_ra - i - 1);
            random_set[i] = unsampled[r];
            unsampled[r] = unsampled.back();
            unsampled.pop_back();
        }
    }

    hvc_topology* t = static_cast<hvc_topology*>(std::calloc(1, sizeof(hvc_topology)));
    if (!t) return HVC_E_OOM;
    t->n_ra = n_ra;
    t->n_i = n_i;
    t->n_tr = n_tr;
    t->dim = p->arrangement == 0 ? 3 : 2;
    t->model_interneuron_distance = mid;
    t->ra_xyz = dup(ra);
    t->i_xyz = dup(in);
    to_csr(a, &t->ra2i_ptr, &t->ra2i_tgt, &t->ra2i_w, &t->ra2i_len, &t->ra2i_d);
    to_csr(b, &t->i2ra_ptr, &t->i2ra_tgt, &t->i2ra_w, &t->i2ra_len, &t->i2ra_d);
    t->training_clustered = dup(clustered);
    t->training_random = dup(random_set);
    if (!t->ra_xyz || !t->i_xyz || !t->ra2i_ptr || !t->ra2i_tgt || !t->ra2i_w || !t->ra2i_len || !t->ra2i_d ||
        !t->i2ra_ptr || !t->i2ra_tgt || !t->i2ra_w || !t->i2ra_len || !t->i2ra_d || !t->training_clustered ||
        !t->training_random) {
        hvc_topology_free(t);
        return HVC_E_OOM;
    }
    *out = t;
    return HVC_OK;
}

extern "C" hvc_status hvc_topology_prepare(hvc_topology* t, const hvc_topology_params* p, uint32_t seed,
                                           double* rr_delay)
{
    if (!t || !p) return HVC_E_ARG;
    if (rr_delay && t->dim != 3) return HVC_E_ARG;
    // sample_axonal_delays (NGS:466-526)
    for (int e = 0; e < t->ra2i_ptr[t->n_ra]; ++e)
        t->ra2i_d[e] = t->ra2i_len[e] * kAxonalDelayPerMicron * p->delay_constant;
    if (rr_delay) {
        const double* x = t->ra_xyz;
        for (int i = 0; i < t->n_ra; ++i)
            for (int j = 0; j < t->n_ra; ++j) {
                double len = j != i ? distance_on_sphere(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * j], x[3 * j + 1],
                                                         x[3 * j + 2]) *
                                          kExperimentalInterneuronDistance / t->model_interneuron_distance
                                    : 0.0;
                rr_delay[static_cast<size_t>(i) * t->n_ra + j] = len * kAxonalDelayPerMicron * p->delay_constant;
            }
    }
    for (int e = 0; e < t->i2ra_ptr[t->n_i]; ++e)
        t->i2ra_d[e] = t->i2ra_len[e] * kAxonalDelayPerMicron * p->delay_constant;
    // resample_weights (NGS:721-737): I->RA first, then RA->I
    RefRng rng(seed);
    for (int e = 0; e < t->i2ra_ptr[t->n_i]; ++e) t->i2ra_w[e] = rng.random(p->Gie_max);
    for (int e = 0; e < t->ra2i_ptr[t->n_ra]; ++e) t->ra2i_w[e] = rng.random(p->Gei_max);
    return HVC_OK;
}

extern "C" void hvc_topology_free(hvc_topology* t)
{
    if (!t) return;
    void* arrays[] = {t->ra_xyz,   t->i_xyz,    t->ra2i_ptr, t->ra2i_tgt, t->ra2i_w,
                      t->ra2i_len, t->ra2i_d,   t->i2ra_ptr, t->i2ra_tgt, t->i2ra_w,
                      t->i2ra_len, t->i2ra_d,   t->training_clustered, t->training_random};
    for (void* a : arrays) std::free(a);
    std::free(t);
}

extern "C" hvc_status hvc_mt19937(uint32_t seed, long long skip, int n, uint32_t* out)
{
    if (skip < 0 || n < 0 || (n > 0 && !out)) return HVC_E_ARG;
    std::mt19937 g(seed);
    g.discard(static_cast<unsigned long long>(skip));
    for (int k = 0; k < n; ++k) out[k] = static_cast<uint32_t>(g());
    return HVC_OK;
}
