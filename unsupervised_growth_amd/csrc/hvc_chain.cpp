// hvc_chain.cpp -- chain-test statistics: calculate_and_write_jitter
// (NGS:1641-1782) and write_jitter (NGS:17818-17857), "NGS" =
// c++/src/NetworkGrowthSimulator.cpp of the reference.  Host code; the
// accumulations run in the reference's order (trial by trial, from 0.0).
#include <cmath>
#include <cstdio>
#include <new>
#include <vector>

#include "../../include/hvc_driver.h"

struct hvc_chain_stats {
    int n_ra = 0;
    int trials = 0;
    std::vector<std::vector<double>> first_soma, first_dend;  // first spike time per trial with a spike
    std::vector<std::vector<int>> n_soma, n_dend;             // spikes per trial
};

extern "C" hvc_status hvc_chain_stats_create(int n_ra, hvc_chain_stats** out)
{
    if (n_ra <= 0 || !out) return HVC_E_ARG;
    hvc_chain_stats* s = new (std::nothrow) hvc_chain_stats();
    if (!s) return HVC_E_OOM;
    s->n_ra = n_ra;
    s->first_soma.resize(n_ra);
    s->first_dend.resize(n_ra);
    s->n_soma.resize(n_ra);
    s->n_dend.resize(n_ra);
    *out = s;
    return HVC_OK;
}

// test_chain's per-trial collection (NGS:1430-1447)
extern "C" hvc_status hvc_chain_stats_add(hvc_chain_stats* s, const int* soma_counts, const double* soma_times,
                                          const int* dend_counts, const double* dend_times)
{
    if (!s || !soma_counts || !dend_counts) return HVC_E_ARG;
    size_t ps = 0, pd = 0;
    for (int i = 0; i < s->n_ra; ++i) {
        if (soma_counts[i] < 0 || dend_counts[i] < 0) return HVC_E_ARG;
        if ((soma_counts[i] && !soma_times) || (dend_counts[i] && !dend_times)) return HVC_E_ARG;
    }
    for (int i = 0; i < s->n_ra; ++i) {
        if (soma_counts[i] >= 1) s->first_soma[i].push_back(soma_times[ps]);
        if (dend_counts[i] >= 1) s->first_dend[i].push_back(dend_times[pd]);
        s->n_soma[i].push_back(soma_counts[i]);
        s->n_dend[i].push_back(dend_counts[i]);
        ps += soma_counts[i];
        pd += dend_counts[i];
    }
    s->trials++;
    return HVC_OK;
}

namespace {

// mean and standard deviation of the first spike times (NGS:1702-1726): -1 when undefined
void jitter(const std::vector<double>& t, double* mean_out, double* std_out)
{
    double mean = -1;
    if (!t.empty()) {
        double sum = 0.0;
        for (double x : t) sum += x;
        mean = sum / static_cast<double>(t.size());
    }
    double accum = 0;
    for (double x : t) accum += (x - mean) * (x - mean);
    *mean_out = mean;
    *std_out = t.size() > 1 ? std::sqrt(accum / (static_cast<double>(t.size()) - 1)) : -1;
}

}  // namespace

extern "C" hvc_status hvc_chain_stats_result(const hvc_chain_stats* s, int* num_trials, double* out)
{
    if (!s || !out) return HVC_E_ARG;
    if (s->trials == 0) return HVC_E_STATE;
    const double T = static_cast<double>(s->trials);
    for (int i = 0; i < s->n_ra; ++i) {
        double* o = out + 8 * static_cast<size_t>(i);
        double sum_s = 0.0, sum_d = 0.0;
        for (int c : s->n_soma[i]) sum_s += c;
        for (int c : s->n_dend[i]) sum_d += c;
        o[0] = static_cast<double>(s->first_soma[i].size()) / T;
        o[1] = sum_s / T;
        jitter(s->first_soma[i], &o[2], &o[3]);
        o[4] = static_cast<double>(s->first_dend[i].size()) / T;
        o[5] = sum_d / T;
        jitter(s->first_dend[i], &o[6], &o[7]);
    }
    if (num_trials) *num_trials = s->trials;
    return HVC_OK;
}

extern "C" hvc_status hvc_chain_stats_write(const hvc_chain_stats* s, const char* path)
{
    if (!s || !path) return HVC_E_ARG;
    std::vector<double> v(8 * static_cast<size_t>(s->n_ra));
    int trials = 0;
    hvc_status r = hvc_chain_stats_result(s, &trials, v.data());
    if (r != HVC_OK) return r;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return HVC_E_ARG;
    bool ok = std::fwrite(&s->n_ra, sizeof(int), 1, f) == 1 && std::fwrite(&trials, sizeof(int), 1, f) == 1 &&
              std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HVC_OK : HVC_E_ARG;
}

extern "C" void hvc_chain_stats_destroy(hvc_chain_stats* s) { delete s; }
