// hvc_driver.cpp -- host growth driver: the loop around the engine's trial.
//
// Restates chain_growth (NGS:5731-5900) and the bookkeeping it runs after every
// trial, with the reference's state files.  "NGS" =
// c++/src/NetworkGrowthSimulator.cpp of the reference.  The engine holds the
// synapse state on the device; the driver reads it back only to replace
// neurons or write files.
#include "../../include/hvc_driver.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "hvc_host.h"

using namespace hvc_host;

namespace {

// Synapse state row by row (the reference's active_synapses / supersynapses
// vectors, NGS.h:177-178), weights per candidate edge.
struct SynRows {
    std::vector<double> w;
    std::vector<std::vector<int>> act, sup;
    std::vector<int> remodeled;
};

}  // namespace

struct hvc_driver {
    hvc_params prm{};
    int device = 0;
    hvc_topology_params top{};
    hvc_maturation_params mat{};
    int n_ra = 0, n_i = 0;
    // fixed-weight connections in the reference's per-neuron order
    std::vector<std::vector<int>> a_tgt, b_tgt;        // RA->I per RA, I->RA per I
    std::vector<std::vector<double>> a_w, a_d, b_w, b_d;
    // RA->RA candidates (CSR, ascending targets), delays per candidate
    std::vector<int> rr_ptr, rr_tgt;
    std::vector<double> rr_d;
    std::vector<int> train_ids;
    std::vector<double> spread;
    // neuron properties and bookkeeping (per HVC-RA)
    std::vector<double> erest, gca;
    std::vector<int> mature, scale, trials_after;
    std::vector<int> history;   // [n_ra][W] ring, logical entry 0 at head
    int head = 0;
    std::vector<int> recent;    // per neuron: trials with a spike among the newest activity-window entries
    int activity_window = 1000; // NGS:9586
    std::vector<double> ra_xyz, i_xyz;
    double mid = 0;
    RefRng rng{0};
    int trial_number = 0;
    bool first = true;
    bool dead = false;
    hvc_ctx* ctx = nullptr;
    bool have_tuning = false;
    hvc_tuning tuning{};
    // last trial's spikes: counts and times per kind (soma, dend, interneuron)
    std::vector<int> cnt[3];
    std::vector<double> times[3];
    bool have_spikes = false;
    std::vector<int> rep_ids, rep_trials;
    std::string out_dir;
    int save_short = 0, save_long = 0;
    std::string err;
};

namespace {

hvc_status fail(hvc_driver* d, hvc_status s, const std::string& m)
{
    d->err = m;
    return s;
}

hvc_status engine_fail(hvc_driver* d, hvc_status s, const char* what)
{
    d->err = std::string(what) + ": " + (d->ctx ? hvc_last_error(d->ctx) : "engine error");
    return s;
}

// syn_lengths_* (NGS:537, 560, 640, 683): distance on the sphere (or in the
// plane for the square arrangement) in microns
double length_microns(const hvc_driver* d, const double* a, const double* b)
{
    double dist = d->top.arrangement == 0 ? distance_on_sphere(a[0], a[1], a[2], b[0], b[1], b[2])
                                          : distance2d(a[0], a[1], b[0], b[1]);
    return dist * kExperimentalInterneuronDistance / d->mid;
}

hvc_status download(hvc_driver* d, SynRows& s)
{
    const int E = d->rr_ptr[d->n_ra];
    std::vector<int> act_cnt(d->n_ra), sup_cnt(d->n_ra), act_ids(std::max(E, 1)), sup_ids(std::max(E, 1));
    s.w.assign(std::max(E, 1), 0.0);
    s.remodeled.assign(d->n_ra, 0);
    hvc_synapses syn{};
    syn.w = s.w.data();
    syn.act_cnt = act_cnt.data();
    syn.act_ids = act_ids.data();
    syn.sup_cnt = sup_cnt.data();
    syn.sup_ids = sup_ids.data();
    syn.remodeled = s.remodeled.data();
    hvc_status r = hvc_download_synapses(d->ctx, &syn);
    if (r != HVC_OK) return engine_fail(d, r, "download synapses");
    s.w.resize(E);
    s.act.assign(d->n_ra, {});
    s.sup.assign(d->n_ra, {});
    size_t pa = 0, ps = 0;
    for (int i = 0; i < d->n_ra; ++i) {
        s.act[i].assign(act_ids.begin() + pa, act_ids.begin() + pa + act_cnt[i]);
        s.sup[i].assign(sup_ids.begin() + ps, sup_ids.begin() + ps + sup_cnt[i]);
        pa += act_cnt[i];
        ps += sup_cnt[i];
    }
    return HVC_OK;
}

// Create the engine context and upload the network with synapse state s.
hvc_status rebuild_engine(hvc_driver* d, const SynRows& s)
{
    if (d->ctx) {
        if (hvc_get_tuning(d->ctx, &d->tuning) == HVC_OK) d->have_tuning = true;
        hvc_destroy(d->ctx);
        d->ctx = nullptr;
    }
    std::vector<int> ap(d->n_ra + 1, 0), at, bp(d->n_i + 1, 0), bt, ac(d->n_ra), sc(d->n_ra), ai, si;
    std::vector<double> aw, ad, bw, bd;
    for (int i = 0; i < d->n_ra; ++i) {
        ap[i + 1] = ap[i] + static_cast<int>(d->a_tgt[i].size());
        at.insert(at.end(), d->a_tgt[i].begin(), d->a_tgt[i].end());
        aw.insert(aw.end(), d->a_w[i].begin(), d->a_w[i].end());
        ad.insert(ad.end(), d->a_d[i].begin(), d->a_d[i].end());
        ac[i] = static_cast<int>(s.act[i].size());
        sc[i] = static_cast<int>(s.sup[i].size());
        ai.insert(ai.end(), s.act[i].begin(), s.act[i].end());
        si.insert(si.end(), s.sup[i].begin(), s.sup[i].end());
    }
    for (int i = 0; i < d->n_i; ++i) {
        bp[i + 1] = bp[i] + static_cast<int>(d->b_tgt[i].size());
        bt.insert(bt.end(), d->b_tgt[i].begin(), d->b_tgt[i].end());
        bw.insert(bw.end(), d->b_w[i].begin(), d->b_w[i].end());
        bd.insert(bd.end(), d->b_d[i].begin(), d->b_d[i].end());
    }
    hvc_network net{};
    net.n_ra = d->n_ra;
    net.n_i = d->n_i;
    net.ra2i_ptr = ap.data(); net.ra2i_tgt = at.data(); net.ra2i_w = aw.data(); net.ra2i_d = ad.data();
    net.i2ra_ptr = bp.data(); net.i2ra_tgt = bt.data(); net.i2ra_w = bw.data(); net.i2ra_d = bd.data();
    net.rr_ptr = d->rr_ptr.data(); net.rr_tgt = d->rr_tgt.data(); net.rr_d = d->rr_d.data(); net.rr_w = s.w.data();
    net.act_cnt = ac.data(); net.act_ids = ai.data();
    net.sup_cnt = sc.data(); net.sup_ids = si.data();
    net.remodeled = s.remodeled.data();
    net.erest = d->erest.data();
    net.gca = d->gca.data();
    net.n_train = static_cast<int>(d->train_ids.size());
    net.train_ids = d->train_ids.data();
    net.spread = d->spread.data();
    hvc_status r = hvc_create(&d->prm, d->device, &d->ctx);
    if (r != HVC_OK) return engine_fail(d, r, "create engine");
    if (d->have_tuning && (r = hvc_set_tuning(d->ctx, &d->tuning)) != HVC_OK) return engine_fail(d, r, "set tuning");
    r = hvc_upload_network(d->ctx, &net);
    if (r != HVC_OK) return engine_fail(d, r, "upload network");
    return HVC_OK;
}

hvc_status fetch_spikes(hvc_driver* d)
{
    for (int k = 0; k < 3; ++k) {
        int total = 0;
        d->cnt[k].assign(k == 2 ? d->n_i : d->n_ra, 0);
        hvc_status r = hvc_trial_spikes(d->ctx, k, d->cnt[k].data(), nullptr, 0, &total);
        if (r != HVC_OK) return engine_fail(d, r, "trial spikes");
        d->times[k].assign(std::max(total, 1), 0.0);
        r = hvc_trial_spikes(d->ctx, k, d->cnt[k].data(), d->times[k].data(), total, &total);
        if (r != HVC_OK) return engine_fail(d, r, "trial spikes");
        d->times[k].resize(total);
    }
    d->have_spikes = true;
    return HVC_OK;
}

// ---------------------------------------------------------------------------
// replacement (replace_neurons, NGS:15858-15870)
// ---------------------------------------------------------------------------
int candidate(const hvc_driver* d, int i, int j)
{
    auto b = d->rr_tgt.begin() + d->rr_ptr[i], e = d->rr_tgt.begin() + d->rr_ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    return it != e && *it == j ? static_cast<int>(it - d->rr_tgt.begin()) : -1;
}

void erase_first(std::vector<int>& v, int x)
{
    auto it = std::find(v.begin(), v.end(), x);
    if (it != v.end()) v.erase(it);
}

// remove_neurons_from_network (NGS:15757-15847)
void remove_neurons(hvc_driver* d, SynRows& s, const std::vector<int>& R)
{
    for (int i = 0; i < d->n_ra; ++i)
        for (int r : R) {
            erase_first(s.act[i], r);
            auto it = std::find(s.sup[i].begin(), s.sup[i].end(), r);
            if (it != s.sup[i].end()) {
                s.sup[i].erase(it);
                s.remodeled[i] = 0;
            }
            int e = candidate(d, i, r);
            if (e >= 0) s.w[e] = 0.0;
        }
    for (int r : R) {
        s.act[r].clear();
        s.sup[r].clear();
        for (int e = d->rr_ptr[r]; e < d->rr_ptr[r + 1]; ++e) s.w[e] = 0.0;
    }
    for (int i = 0; i < d->n_i; ++i)
        for (int r : R) {
            auto it = std::find(d->b_tgt[i].begin(), d->b_tgt[i].end(), r);
            if (it != d->b_tgt[i].end()) {
                size_t k = it - d->b_tgt[i].begin();
                d->b_tgt[i].erase(it);
                d->b_w[i].erase(d->b_w[i].begin() + k);
                d->b_d[i].erase(d->b_d[i].begin() + k);
            }
        }
    for (int r : R) {
        d->a_tgt[r].clear();
        d->a_w[r].clear();
        d->a_d[r].clear();
    }
}

// sample_coordinates_for_replaced (NGS:157-220)
void resample_coordinates(hvc_driver* d, const std::vector<int>& R)
{
    std::vector<char> replaced(d->n_ra, 0);
    for (int r : R) replaced[r] = 1;
    std::vector<int> fixed;
    for (int i = 0; i < d->n_ra; ++i)
        if (!replaced[i]) fixed.push_back(i);
    const double too_close = d->mid * kMinIntersomaticDistance / kExperimentalInterneuronDistance;
    for (int r : R) {
        double tx, ty, tz;
        bool close = true;
        while (close) {
            close = false;
            double z = 2 * d->rng.random(1.0) - 1;
            double phi = d->rng.random(2.0 * kPi);
            tx = std::sin(std::acos(z)) * std::cos(phi);
            ty = std::sin(std::acos(z)) * std::sin(phi);
            tz = z;
            for (int j = 0; j < d->n_i; ++j)
                if (distance_on_sphere(tx, ty, tz, d->i_xyz[3 * j], d->i_xyz[3 * j + 1], d->i_xyz[3 * j + 2]) <
                    too_close) {
                    close = true;
                    break;
                }
            if (!close)
                for (int j : fixed)
                    if (distance_on_sphere(tx, ty, tz, d->ra_xyz[3 * j], d->ra_xyz[3 * j + 1], d->ra_xyz[3 * j + 2]) <
                        too_close) {
                        close = true;
                        break;
                    }
        }
        d->ra_xyz[3 * r] = tx;
        d->ra_xyz[3 * r + 1] = ty;
        d->ra_xyz[3 * r + 2] = tz;
        fixed.push_back(r);
    }
}

// sample_connectionsAndDelays_for_replaced (NGS:528-618)
void resample_connections(hvc_driver* d, const std::vector<int>& R)
{
    const hvc_topology_params& p = d->top;
    auto prob = [](double peak, double sigma, double dd) {
        double sigma_square = sigma * sigma;
        return peak * std::exp(-dd * dd / (2 * sigma_square));
    };
    for (int r : R)
        for (int j = 0; j < d->n_i; ++j) {
            double dd = length_microns(d, &d->ra_xyz[3 * r], &d->i_xyz[3 * j]);
            if (d->rng.random(1) < prob(p.A_RA2I, p.SIGMA_RA2I, dd)) {
                double G = d->rng.random(p.Gei_max);
                d->a_w[r].push_back(G);
                d->a_tgt[r].push_back(j);
                d->a_d[r].push_back(dd * kAxonalDelayPerMicron * p.delay_constant);
            }
        }
    for (int i = 0; i < d->n_i; ++i)
        for (int r : R) {
            double dd = length_microns(d, &d->i_xyz[3 * i], &d->ra_xyz[3 * r]);
            if (d->rng.random(1) < prob(p.B_I2RA, p.SIGMA_I2RA, dd)) {
                double G = d->rng.random(p.Gie_max);
                d->b_w[i].push_back(G);
                d->b_tgt[i].push_back(r);
                d->b_d[i].push_back(dd * kAxonalDelayPerMicron * p.delay_constant);
            }
        }
    // RA->RA delays to and from the replaced neurons (NGS:578-617) on the
    // candidate edges
    std::vector<char> replaced(d->n_ra, 0);
    for (int r : R) replaced[r] = 1;
    for (int i = 0; i < d->n_ra; ++i)
        for (int e = d->rr_ptr[i]; e < d->rr_ptr[i + 1]; ++e) {
            int j = d->rr_tgt[e];
            if (!replaced[i] && !replaced[j]) continue;
            double len = j != i ? length_microns(d, &d->ra_xyz[3 * i], &d->ra_xyz[3 * j]) : 0.0;
            d->rr_d[e] = len * kAxonalDelayPerMicron * p.delay_constant;
        }
}

hvc_status replace_neurons(hvc_driver* d, const std::vector<int>& R)
{
    if (d->top.arrangement != 0) return fail(d, HVC_E_STATE, "replacing neurons needs the sphere arrangement (NGS:157-220)");
    SynRows s;
    hvc_status r = download(d, s);
    if (r != HVC_OK) return r;
    remove_neurons(d, s, R);
    resample_coordinates(d, R);
    resample_connections(d, R);
    // update_replaced_neurons (NGS:2109-2133)
    const int W = d->mat.RATE_WINDOW_LONG;
    for (int n : R) {
        d->erest[n] = d->mat.E_REST_IMMATURE;
        d->gca[n] = d->mat.GCA_IMMATURE;
        d->mature[n] = 0;
        d->scale[n] = kMaturationScaleSpontaneous;
        s.remodeled[n] = 0;
        d->trials_after[n] = 0;
        std::fill(d->history.begin() + static_cast<size_t>(n) * W, d->history.begin() + static_cast<size_t>(n + 1) * W, 0);
        d->recent[n] = 0;
        d->rep_ids.push_back(n);
        d->rep_trials.push_back(d->trial_number);
    }
    return rebuild_engine(d, s);
}

// ---------------------------------------------------------------------------
// state files (NGS:17531-18550)
// ---------------------------------------------------------------------------
struct Out {
    std::FILE* f = nullptr;
    bool ok = true;
    Out(const std::string& path, bool append) { f = std::fopen(path.c_str(), append ? "ab" : "wb"); ok = f != nullptr; }
    ~Out() { close(); }
    bool close()
    {
        if (f) {
            ok = (std::fclose(f) == 0) && ok;
            f = nullptr;
        }
        return ok;
    }
    void i32(int v) { raw(&v, sizeof v, 1); }
    void f64(double v) { raw(&v, sizeof v, 1); }
    void raw(const void* p, size_t sz, size_t n)
    {
        if (ok && n && std::fwrite(p, sz, n, f) != n) ok = false;
    }
};

std::string join(const std::string& dir, const std::string& name)
{
    if (dir.empty() || dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

// write_active_synapses / write_super_synapses (NGS:17531-17578)
bool write_lists(const hvc_driver* d, const std::string& path, const std::vector<std::vector<int>>& rows)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->trial_number);
    for (const auto& r : rows) {
        o.i32(static_cast<int>(r.size()));
        o.raw(r.data(), sizeof(int), r.size());
    }
    return o.close();
}

// write_*_spike_times (NGS:17916-18033): spike times of the last trial
bool write_spikes(const hvc_driver* d, const std::string& path, int kind)
{
    Out o(path, false);
    const int n = kind == 2 ? d->n_i : d->n_ra;
    o.i32(d->trial_number);
    o.f64(d->prm.trial_ms);
    o.i32(n);
    size_t pos = 0;
    for (int i = 0; i < n; ++i) {
        int c = d->have_spikes ? d->cnt[kind][i] : 0;
        o.i32(c);
        if (c) o.raw(d->times[kind].data() + pos, sizeof(double), c);
        pos += c;
    }
    return o.close();
}

// write_num_synapses (NGS:17885-17914): appended, new file at trial 0
bool write_num_synapses(const hvc_driver* d, const std::string& path, const SynRows& s)
{
    Out o(path, d->trial_number != 0);
    int act = 0, sup = 0;
    for (int i = 0; i < d->n_ra; ++i) {
        act += static_cast<int>(s.act[i].size());
        sup += static_cast<int>(s.sup[i].size());
    }
    o.i32(d->trial_number);
    o.i32(act);
    o.i32(sup);
    return o.close();
}

// The dense N_RA x N_RA view of a per-candidate array, row by row; pairs
// outside the candidate set take `other(i, j)`.
template <class F>
void dense_rows(const hvc_driver* d, const std::vector<double>& per_edge, F&& other, std::vector<double>& row, int i)
{
    row.resize(d->n_ra);
    int e = d->rr_ptr[i];
    for (int j = 0; j < d->n_ra; ++j) {
        if (e < d->rr_ptr[i + 1] && d->rr_tgt[e] == j)
            row[j] = per_edge[e++];
        else
            row[j] = other(i, j);
    }
}

// write_weights (NGS:17760-17779): N, trial, N x N doubles (0 outside the candidates)
bool write_weights(const hvc_driver* d, const std::string& path, const SynRows& s)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->trial_number);
    std::vector<double> row;
    for (int i = 0; i < d->n_ra; ++i) {
        dense_rows(d, s.w, [](int, int) { return 0.0; }, row, i);
        o.raw(row.data(), sizeof(double), row.size());
    }
    return o.close();
}

// write_weight_statistics (NGS:17580-17630): trial, mean and standard deviation
// of all N_RA^2 weights, appended (new file at trial 0).  Accumulated in the
// reference's order: row sums added in turn, then squared deviations pair by pair.
bool write_weight_statistics(const hvc_driver* d, const std::string& path, const SynRows& s)
{
    Out o(path, d->trial_number != 0);
    o.i32(d->trial_number);
    const double N = static_cast<double>(d->n_ra);
    double mean = 0.0;
    for (int i = 0; i < d->n_ra; ++i) {
        double row = 0.0;
        for (int e = d->rr_ptr[i]; e < d->rr_ptr[i + 1]; ++e) row += s.w[e];  // adding the zeros between changes nothing
        mean += row;
    }
    mean /= (N * N);
    double sq = 0.0;
    for (int i = 0; i < d->n_ra; ++i) {
        int e = d->rr_ptr[i];
        for (int j = 0; j < d->n_ra; ++j) {
            double w = (e < d->rr_ptr[i + 1] && d->rr_tgt[e] == j) ? s.w[e++] : 0.0;
            sq += (w - mean) * (w - mean);
        }
    }
    if (d->n_ra * d->n_ra > 1) sq = std::sqrt(sq / (N * N - 1));
    o.f64(mean);
    o.f64(sq);
    return o.close();
}

// write_axonal_delays (NGS:17733-17758) of the RA->RA pairs: dense rows; a pair
// outside the candidate set takes the delay its distance gives (NGS:490-499)
bool write_rr_delays(const hvc_driver* d, const std::string& path)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->trial_number);
    const double dc = d->top.delay_constant;
    auto dist_delay = [d, dc](int i, int j) {
        double len = j != i ? length_microns(d, &d->ra_xyz[3 * i], &d->ra_xyz[3 * j]) : 0.0;
        return len * kAxonalDelayPerMicron * dc;
    };
    std::vector<double> row;
    for (int i = 0; i < d->n_ra; ++i) {
        o.i32(d->n_ra);
        dense_rows(d, d->rr_d, dist_delay, row, i);
        o.raw(row.data(), sizeof(double), row.size());
    }
    return o.close();
}

bool write_delay_rows(const hvc_driver* d, const std::string& path, const std::vector<std::vector<double>>& rows)
{
    Out o(path, false);
    o.i32(static_cast<int>(rows.size()));
    o.i32(d->trial_number);
    for (const auto& r : rows) {
        o.i32(static_cast<int>(r.size()));
        o.raw(r.data(), sizeof(double), r.size());
    }
    return o.close();
}

// write_connections_RAandI (NGS:18312-18365): N, per neuron {id, n, {tgt, w, length, delay}}
bool write_connections(const hvc_driver* d, const std::string& path_ra_i, const std::string& path_i_ra)
{
    bool ok = true;
    for (int side = 0; side < 2; ++side) {
        Out o(side == 0 ? path_ra_i : path_i_ra, false);
        const int n = side == 0 ? d->n_ra : d->n_i;
        const auto& T = side == 0 ? d->a_tgt : d->b_tgt;
        const auto& Wt = side == 0 ? d->a_w : d->b_w;
        const auto& D = side == 0 ? d->a_d : d->b_d;
        o.i32(n);
        for (int i = 0; i < n; ++i) {
            o.i32(i);
            o.i32(static_cast<int>(T[i].size()));
            const double* src = side == 0 ? &d->ra_xyz[3 * i] : &d->i_xyz[3 * i];
            for (size_t k = 0; k < T[i].size(); ++k) {
                const double* dst = side == 0 ? &d->i_xyz[3 * T[i][k]] : &d->ra_xyz[3 * T[i][k]];
                o.i32(T[i][k]);
                o.f64(Wt[i][k]);
                o.f64(length_microns(d, src, dst));
                o.f64(D[i][k]);
            }
        }
        ok = o.close() && ok;
    }
    return ok;
}

// write_coordinates (NGS:18431-18500): N, dim, MODEL_INTERNEURON_DISTANCE, coordinates
bool write_coordinates(const hvc_driver* d, const std::string& path, const std::vector<double>& xyz)
{
    Out o(path, false);
    const int n = static_cast<int>(xyz.size() / 3);
    const int dim = d->top.arrangement == 0 ? 3 : 2;
    o.i32(n);
    o.i32(dim);
    o.f64(d->mid);
    for (int i = 0; i < n; ++i) o.raw(&xyz[3 * i], sizeof(double), dim);
    return o.close();
}

bool write_ints(const hvc_driver* d, const std::string& path, const std::vector<int>& v)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->trial_number);
    o.raw(v.data(), sizeof(int), v.size());
    return o.close();
}

// write_maturation_properties (NGS:17650-17668)
bool write_maturation(const hvc_driver* d, const std::string& path)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->trial_number);
    for (int i = 0; i < d->n_ra; ++i) {
        o.i32(d->mature[i]);
        o.i32(d->scale[i]);
        o.f64(d->erest[i]);
        o.f64(d->gca[i]);
    }
    return o.close();
}

void history_logical(const hvc_driver* d, std::vector<int>& out)
{
    const int W = d->mat.RATE_WINDOW_LONG;
    out.resize(static_cast<size_t>(d->n_ra) * W);
    for (int i = 0; i < d->n_ra; ++i)
        for (int j = 0; j < W; ++j)
            out[static_cast<size_t>(i) * W + j] = d->history[static_cast<size_t>(i) * W + (d->head + j) % W];
}

// write_activity_history (NGS:17711-17731): N, RATE_WINDOW_LONG, trial, ints newest first
bool write_history(const hvc_driver* d, const std::string& path)
{
    Out o(path, false);
    o.i32(d->n_ra);
    o.i32(d->mat.RATE_WINDOW_LONG);
    o.i32(d->trial_number);
    std::vector<int> h;
    history_logical(d, h);
    o.raw(h.data(), sizeof(int), h.size());
    return o.close();
}

hvc_status write_graph(hvc_driver* d, const std::string& dir, const SynRows& s)
{
    bool ok = write_num_synapses(d, join(dir, "num_synapses.bin"), s);
    ok = write_spikes(d, join(dir, "spike_times_soma.bin"), 0) && ok;
    ok = write_spikes(d, join(dir, "spike_times_dend.bin"), 1) && ok;
    ok = write_spikes(d, join(dir, "spike_times_interneuron.bin"), 2) && ok;
    ok = write_lists(d, join(dir, "RA_RA_super_connections.bin"), s.sup) && ok;
    ok = write_lists(d, join(dir, "RA_RA_active_connections.bin"), s.act) && ok;
    return ok ? HVC_OK : fail(d, HVC_E_ARG, "cannot write the graph state into " + dir);
}

hvc_status write_full(hvc_driver* d, const std::string& dir, const std::string& e, const SynRows& s)
{
    auto f = [&](const char* name) { return join(dir, std::string(name) + e + ".bin"); };
    bool ok = write_coordinates(d, f("RA_xy"), d->ra_xyz);
    ok = write_connections(d, f("RA_I_connections"), f("I_RA_connections")) && ok;
    ok = write_weights(d, f("weights"), s) && ok;
    ok = write_weight_statistics(d, join(dir, "weight_statistics.bin"), s) && ok;
    ok = write_rr_delays(d, f("axonal_delays_RA2RA")) && ok;
    ok = write_delay_rows(d, f("axonal_delays_RA2I"), d->a_d) && ok;
    ok = write_delay_rows(d, f("axonal_delays_I2RA"), d->b_d) && ok;
    ok = write_lists(d, f("RA_RA_super_connections"), s.sup) && ok;
    ok = write_lists(d, f("RA_RA_active_connections"), s.act) && ok;
    ok = write_ints(d, f("mature"), d->mature) && ok;
    ok = write_maturation(d, f("maturation_properties")) && ok;
    ok = write_history(d, f("activity_history")) && ok;
    ok = write_ints(d, f("replacement_history"), d->trials_after) && ok;
    ok = write_ints(d, f("remodeled_indicators"), s.remodeled) && ok;
    ok = write_spikes(d, f("spike_times_soma"), 0) && ok;
    ok = write_spikes(d, f("spike_times_dend"), 1) && ok;
    ok = write_spikes(d, f("spike_times_interneuron"), 2) && ok;
    return ok ? HVC_OK : fail(d, HVC_E_ARG, "cannot write the full state into " + dir);
}

// the files new_chain_growth writes before growth (NGS:1903-1919)
hvc_status write_initial(hvc_driver* d, const std::string& dir)
{
    bool ok = write_coordinates(d, join(dir, "I_xy.bin"), d->i_xyz);
    {
        Out o(join(dir, "training_neurons.bin"), false);   // write_training_neurons (NGS:18521-18538)
        o.i32(static_cast<int>(d->train_ids.size()));
        o.raw(d->train_ids.data(), sizeof(int), d->train_ids.size());
        ok = o.close() && ok;
    }
    {
        Out o(join(dir, "training_spread.bin"), false);    // write_training_spread (NGS:18502-18519)
        o.i32(static_cast<int>(d->spread.size()));
        o.raw(d->spread.data(), sizeof(double), d->spread.size());
        ok = o.close() && ok;
    }
    {
        Out o(join(dir, "num_neurons.bin"), false);        // write_number_of_neurons (NGS:18540-18550)
        o.i32(d->n_ra);
        o.i32(d->n_i);
        ok = o.close() && ok;
    }
    ok = write_coordinates(d, join(dir, "RA_xy_initial.bin"), d->ra_xyz) && ok;
    ok = write_connections(d, join(dir, "RA_I_connections_initial.bin"), join(dir, "I_RA_connections_initial.bin")) && ok;
    return ok ? HVC_OK : fail(d, HVC_E_ARG, "cannot write the initial files into " + dir);
}

}  // namespace

// ---------------------------------------------------------------------------
// bookkeeping on host arrays
// ---------------------------------------------------------------------------
extern "C" hvc_status hvc_activity_update(int n, int window, int activity_window, int* history, int* head,
                                          int* recent, const int* soma_counts, const int* trials_after_replacement,
                                          int* maturation_scale, double death_rate_threshold, int* to_replace,
                                          int* n_replace, int* became_driven)
{
    const int kWindow = activity_window;   // 1000 in the reference (NGS:9586)
    const double kDriven = 0.50;           // NGS:9588
    if (n < 0 || kWindow < 1 || window < kWindow || !history || !head || *head < 0 || *head >= window ||
        (n > 0 && (!soma_counts || !trials_after_replacement || !maturation_scale || !to_replace)) || !n_replace)
        return HVC_E_ARG;
    const int h0 = *head;
    const int h = (h0 + window - 1) % window;
    *head = h;
    int nr = 0, nd = 0;
    for (int i = 0; i < n; ++i) {
        int* row = history + static_cast<size_t>(i) * window;
        int fired = 0;
        if (recent) {
            // the entry kWindow - 1 trials back leaves the window, the new count enters
            fired = recent[i] - (row[(h0 + kWindow - 1) % window] > 0 ? 1 : 0) + (soma_counts[i] > 0 ? 1 : 0);
            row[h] = soma_counts[i];
            recent[i] = fired;
        } else {
            row[h] = soma_counts[i];
            for (int j = 0; j < kWindow; ++j) fired += row[(h + j) % window] > 0 ? 1 : 0;
        }
        double robustness = static_cast<double>(fired) / static_cast<double>(kWindow);
        if (robustness >= kDriven && maturation_scale[i] != hvc_host::kMaturationScaleDriven) {
            maturation_scale[i] = hvc_host::kMaturationScaleDriven;
            ++nd;
        }
        if (robustness < death_rate_threshold && trials_after_replacement[i] > kWindow) to_replace[nr++] = i;
    }
    *n_replace = nr;
    if (became_driven) *became_driven = nd;
    return HVC_OK;
}

extern "C" hvc_status hvc_maturation_update(int n, const hvc_maturation_params* p, const int* maturation_scale,
                                            int* mature, double* erest, double* gca, int* matured)
{
    if (n < 0 || !p || (n > 0 && (!maturation_scale || !mature || !erest || !gca))) return HVC_E_ARG;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (mature[i] == 1) continue;
        if (maturation_scale[i] == 0) return HVC_E_ARG;
        erest[i] = erest[i] + (erest[i] - p->E_REST_MATURE) * (std::exp(-1.0 / maturation_scale[i]) - 1);
        gca[i] = gca[i] + (gca[i] - p->GCA_MATURE) * (std::exp(-1.0 / maturation_scale[i]) - 1);
        const double potential_threshold = 0.5, conductance_threshold = 0.5;   // NGS:2167-2168
        if (gca[i] >= p->GCA_MATURE - conductance_threshold || erest[i] <= p->E_REST_MATURE - potential_threshold) {
            gca[i] = p->GCA_MATURE;
            erest[i] = p->E_REST_MATURE;
            mature[i] = 1;
            ++m;
        }
    }
    if (matured) *matured = m;
    return HVC_OK;
}

// ---------------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------------
extern "C" hvc_status hvc_driver_create(const hvc_params* params, int device, const hvc_network* net,
                                        const hvc_geometry* geometry, const hvc_topology_params* topology,
                                        const hvc_maturation_params* maturation, const hvc_driver_resume* resume,
                                        uint32_t seed, hvc_driver** out)
{
    if (!params || !net || !geometry || !topology || !maturation || !out) return HVC_E_ARG;
    *out = nullptr;
    hvc_driver* d = new (std::nothrow) hvc_driver();
    if (!d) return HVC_E_OOM;
    *out = d;
    if (net->n_ra <= 0 || net->n_i < 0) return fail(d, HVC_E_ARG, "network needs HVC-RA neurons");
    const int act_window = maturation->ACTIVITY_WINDOW > 0 ? maturation->ACTIVITY_WINDOW : 1000;
    if (maturation->RATE_WINDOW_LONG < act_window)
        return fail(d, HVC_E_ARG, "RATE_WINDOW_LONG must be >= the activity window (1000 in the reference, NGS:9586-9599)");
    if (!geometry->ra_xyz || (net->n_i > 0 && !geometry->i_xyz) || !(geometry->model_interneuron_distance > 0))
        return fail(d, HVC_E_ARG, "geometry needs positions and a model interneuron distance > 0");
    if (!(params->trial_ms > 2 * kWaitingTime)) return fail(d, HVC_E_ARG, "trial must be longer than 2 x WAITING_TIME (NGS:8898)");
    const int nra = net->n_ra, ni = net->n_i, W = maturation->RATE_WINDOW_LONG;
    d->prm = *params;
    d->device = device;
    d->top = *topology;
    d->mat = *maturation;
    d->n_ra = nra;
    d->n_i = ni;
    d->mid = geometry->model_interneuron_distance;
    d->ra_xyz.assign(geometry->ra_xyz, geometry->ra_xyz + 3 * nra);
    if (ni) d->i_xyz.assign(geometry->i_xyz, geometry->i_xyz + 3 * ni);
    d->a_tgt.resize(nra); d->a_w.resize(nra); d->a_d.resize(nra);
    for (int i = 0; i < nra; ++i)
        for (int e = net->ra2i_ptr[i]; e < net->ra2i_ptr[i + 1]; ++e) {
            d->a_tgt[i].push_back(net->ra2i_tgt[e]);
            d->a_w[i].push_back(net->ra2i_w[e]);
            d->a_d[i].push_back(net->ra2i_d[e]);
        }
    d->b_tgt.resize(ni); d->b_w.resize(ni); d->b_d.resize(ni);
    for (int i = 0; i < ni; ++i)
        for (int e = net->i2ra_ptr[i]; e < net->i2ra_ptr[i + 1]; ++e) {
            d->b_tgt[i].push_back(net->i2ra_tgt[e]);
            d->b_w[i].push_back(net->i2ra_w[e]);
            d->b_d[i].push_back(net->i2ra_d[e]);
        }
    const int E = net->rr_ptr[nra];
    d->rr_ptr.assign(net->rr_ptr, net->rr_ptr + nra + 1);
    d->rr_tgt.assign(net->rr_tgt, net->rr_tgt + E);
    d->rr_d.assign(net->rr_d, net->rr_d + E);
    d->train_ids.assign(net->train_ids, net->train_ids + net->n_train);
    d->spread.assign(net->spread, net->spread + net->n_train);
    d->erest.assign(net->erest, net->erest + nra);
    d->gca.assign(net->gca, net->gca + nra);
    d->mature.assign(nra, 0);
    d->scale.assign(nra, kMaturationScaleSpontaneous);
    d->trials_after.assign(nra, 0);
    d->history.assign(static_cast<size_t>(nra) * W, 0);
    if (resume && resume->mature) d->mature.assign(resume->mature, resume->mature + nra);
    else
        for (int id : d->train_ids)
            if (id >= 0 && id < nra) d->mature[id] = 1;
    if (resume && resume->maturation_scale) d->scale.assign(resume->maturation_scale, resume->maturation_scale + nra);
    if (resume && resume->trials_after_replacement)
        d->trials_after.assign(resume->trials_after_replacement, resume->trials_after_replacement + nra);
    if (resume && resume->history) d->history.assign(resume->history, resume->history + static_cast<size_t>(nra) * W);
    d->trial_number = resume ? resume->trial_number : 0;
    d->activity_window = act_window;
    d->recent.assign(nra, 0);
    for (int i = 0; i < nra; ++i)
        for (int j = 0; j < act_window; ++j) d->recent[i] += d->history[static_cast<size_t>(i) * W + j] > 0 ? 1 : 0;
    for (int i = 0; i < nra; ++i)
        if (d->scale[i] == 0) return fail(d, HVC_E_ARG, "maturation scale 0");
    d->rng = RefRng(seed);

    SynRows s;
    s.w.assign(net->rr_w, net->rr_w + E);
    s.act.resize(nra);
    s.sup.resize(nra);
    size_t pa = 0, ps = 0;
    for (int i = 0; i < nra; ++i) {
        s.act[i].assign(net->act_ids + pa, net->act_ids + pa + net->act_cnt[i]);
        s.sup[i].assign(net->sup_ids + ps, net->sup_ids + ps + net->sup_cnt[i]);
        pa += net->act_cnt[i];
        ps += net->sup_cnt[i];
    }
    s.remodeled.assign(net->remodeled, net->remodeled + nra);
    return rebuild_engine(d, s);
}

extern "C" hvc_status hvc_driver_set_output(hvc_driver* d, const char* directory, int save_freq_short, int save_freq_long)
{
    if (!d || save_freq_short < 0 || save_freq_long < 0) return HVC_E_ARG;
    if ((save_freq_short || save_freq_long) && !directory) return fail(d, HVC_E_ARG, "no output directory");
    d->out_dir = directory ? directory : "";
    d->save_short = save_freq_short;
    d->save_long = save_freq_long;
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_trial(hvc_driver* d, int training, hvc_driver_report* report)
{
    if (!d) return HVC_E_ARG;
    if (d->dead) return fail(d, HVC_E_STATE, "a trial failed check_bad_values: the growth loop has stopped");
    if (!d->ctx) return fail(d, HVC_E_STATE, "no engine context");
    hvc_driver_report rep{};
    rep.trial = d->trial_number;
    // trial_1stSoma_pre_1stSoma_post_delays_fixedSpread: injection time (NGS:8898)
    rep.mean_injection_time = kWaitingTime + d->rng.random(d->prm.trial_ms - 2 * kWaitingTime);
    hvc_trial t{};
    t.mode = 0;
    t.training = training ? 1 : 0;
    t.trial_index = static_cast<uint32_t>(d->trial_number);
    t.first_trial = d->first ? 1 : 0;
    t.mean_injection_time = rep.mean_injection_time;
    hvc_trial_out o{};
    hvc_status r = hvc_run_trial(d->ctx, &t, &o);
    d->first = false;
    if (r == HVC_E_NUMERIC) {
        d->dead = true;
        rep.bad_values = -1;
        if (report) *report = rep;
        return fail(d, HVC_E_NUMERIC, "check_bad_values failed: bad values in neuron dynamics (NGS:5800-5805)");
    }
    if (r != HVC_OK) return engine_fail(d, r, "trial");
    if ((r = fetch_spikes(d)) != HVC_OK) return r;
    rep.soma_spikes = static_cast<int>(d->times[0].size());
    rep.dend_spikes = static_cast<int>(d->times[1].size());
    rep.i_spikes = static_cast<int>(d->times[2].size());

    // periodic state files (NGS:5807-5825)
    bool g = d->save_short > 0 && d->trial_number % d->save_short == 0;
    bool f = d->save_long > 0 && d->trial_number % d->save_long == 0;
    if (g || f) {
        SynRows s;
        if ((r = download(d, s)) != HVC_OK) return r;
        if (g && (r = write_graph(d, d->out_dir, s)) != HVC_OK) return r;
        if (f && (r = write_full(d, d->out_dir, "_" + std::to_string(d->trial_number), s)) != HVC_OK) return r;
        rep.saved = (g ? 1 : 0) | (f ? 2 : 0);
    }

    // bookkeeping (NGS:5831-5870)
    for (int& a : d->trials_after) a += 1;
    std::vector<int> to_replace(d->n_ra);
    int n_replace = 0;
    r = hvc_activity_update(d->n_ra, d->mat.RATE_WINDOW_LONG, d->activity_window, d->history.data(), &d->head,
                            d->recent.data(),
                            d->cnt[0].data(), d->trials_after.data(), d->scale.data(), d->mat.DEATH_RATE_THRESHOLD, to_replace.data(),
                            &n_replace, &rep.became_driven);
    if (r != HVC_OK) return fail(d, r, "activity update");
    to_replace.resize(n_replace);
    int lo = d->n_ra, hi = -1;
    for (int i = 0; i < d->n_ra; ++i)
        if (d->mature[i] != 1) {
            lo = std::min(lo, i);
            hi = i;
        }
    r = hvc_maturation_update(d->n_ra, &d->mat, d->scale.data(), d->mature.data(), d->erest.data(), d->gca.data(),
                              &rep.matured);
    if (r != HVC_OK) return fail(d, r, "maturation update");
    if (!to_replace.empty()) {
        if ((r = replace_neurons(d, to_replace)) != HVC_OK) return r;
        rep.replaced = n_replace;
    } else if (hi >= lo) {
        // set_Erest / set_GCa of the neurons that were immature (NGS:2180-2181)
        std::vector<int> ids(hi - lo + 1);
        std::iota(ids.begin(), ids.end(), lo);
        r = hvc_patch_neurons(d->ctx, ids.data(), hi - lo + 1, d->erest.data() + lo, d->gca.data() + lo);
        if (r != HVC_OK) return engine_fail(d, r, "patch neurons");
    }
    d->trial_number++;
    if (report) *report = rep;
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_chain_test(hvc_driver* d, int num_trials, const char* directory, double* stats)
{
    if (!d || num_trials <= 0) return HVC_E_ARG;
    if (d->dead) return fail(d, HVC_E_STATE, "a trial failed check_bad_values: the network is not usable");
    if (!d->ctx) return fail(d, HVC_E_STATE, "no engine context");
    hvc_chain_stats* cs = nullptr;
    hvc_status r = hvc_chain_stats_create(d->n_ra, &cs);
    if (r != HVC_OK) return fail(d, r, "chain statistics");
    for (int k = 0; k < num_trials && r == HVC_OK; ++k) {
        // trial_no_stdp_fixedSpread: kick at mean_injection_time = 100 ms (NGS:12785)
        hvc_trial t{};
        t.mode = 1;
        t.training = 1;
        t.trial_index = 0x80000000u + static_cast<uint32_t>(k);
        t.first_trial = d->first ? 1 : 0;
        t.mean_injection_time = 100.0;
        hvc_trial_out o{};
        r = hvc_run_trial(d->ctx, &t, &o);
        d->first = false;
        if (r == HVC_E_NUMERIC) {
            r = fail(d, r, "check_bad_values failed in the chain test (NGS:1410-1417)");
            break;
        }
        if (r != HVC_OK) {
            r = engine_fail(d, r, "chain-test trial");
            break;
        }
        if ((r = fetch_spikes(d)) != HVC_OK) break;
        r = hvc_chain_stats_add(cs, d->cnt[0].data(), d->times[0].data(), d->cnt[1].data(), d->times[1].data());
        if (r != HVC_OK) {
            r = fail(d, r, "chain statistics");
            break;
        }
        if (directory) {   // write_chain_test (NGS:18595-18604)
            const std::string e = "_" + std::to_string(k) + ".bin";
            bool ok = write_spikes(d, join(directory, "test_spike_times_soma" + e), 0);
            ok = write_spikes(d, join(directory, "test_spike_times_dend" + e), 1) && ok;
            ok = write_spikes(d, join(directory, "test_spike_times_interneuron" + e), 2) && ok;
            if (!ok) r = fail(d, HVC_E_ARG, std::string("cannot write chain-test spikes into ") + directory);
        }
    }
    if (r == HVC_OK && stats) r = hvc_chain_stats_result(cs, nullptr, stats);
    if (r == HVC_OK && directory && (r = hvc_chain_stats_write(cs, join(directory, "jitter.bin").c_str())) != HVC_OK)
        fail(d, r, std::string("cannot write jitter.bin into ") + directory);
    hvc_chain_stats_destroy(cs);
    return r;
}

extern "C" hvc_status hvc_driver_write_state(hvc_driver* d, const char* directory, const char* file_ending, int what)
{
    if (!d || !directory || what < 0 || what > 2) return HVC_E_ARG;
    if (!d->ctx) return fail(d, HVC_E_STATE, "no engine context");
    if (what == 2) return write_initial(d, directory);
    SynRows s;
    hvc_status r = download(d, s);
    if (r != HVC_OK) return r;
    if (what == 0) return write_graph(d, directory, s);
    return write_full(d, directory, file_ending ? file_ending : "", s);
}

extern "C" hvc_status hvc_driver_maturation(hvc_driver* d, int* mature, int* maturation_scale, double* erest,
                                            double* gca, int* trials_after_replacement)
{
    if (!d) return HVC_E_ARG;
    if (mature) std::copy(d->mature.begin(), d->mature.end(), mature);
    if (maturation_scale) std::copy(d->scale.begin(), d->scale.end(), maturation_scale);
    if (erest) std::copy(d->erest.begin(), d->erest.end(), erest);
    if (gca) std::copy(d->gca.begin(), d->gca.end(), gca);
    if (trials_after_replacement) std::copy(d->trials_after.begin(), d->trials_after.end(), trials_after_replacement);
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_history(hvc_driver* d, int* history)
{
    if (!d || !history) return HVC_E_ARG;
    std::vector<int> h;
    history_logical(d, h);
    std::copy(h.begin(), h.end(), history);
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_coordinates(hvc_driver* d, double* ra_xyz, double* i_xyz)
{
    if (!d) return HVC_E_ARG;
    if (ra_xyz) std::copy(d->ra_xyz.begin(), d->ra_xyz.end(), ra_xyz);
    if (i_xyz) std::copy(d->i_xyz.begin(), d->i_xyz.end(), i_xyz);
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_replaced(hvc_driver* d, int* ids, int* trials, int cap, int* n)
{
    if (!d || !n || cap < 0) return HVC_E_ARG;
    const int total = static_cast<int>(d->rep_ids.size());
    for (int k = 0; k < std::min(cap, total); ++k) {
        if (ids) ids[k] = d->rep_ids[k];
        if (trials) trials[k] = d->rep_trials[k];
    }
    *n = total;
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_connections(hvc_driver* d, long long* n_ra2i, long long* n_i2ra, int* ra2i_ptr,
                                             int* ra2i_tgt, double* ra2i_w, double* ra2i_d, int* i2ra_ptr,
                                             int* i2ra_tgt, double* i2ra_w, double* i2ra_d)
{
    if (!d) return HVC_E_ARG;
    long long na = 0, nb = 0;
    for (int i = 0; i < d->n_ra; ++i) {
        if (ra2i_ptr) ra2i_ptr[i] = static_cast<int>(na);
        for (size_t k = 0; k < d->a_tgt[i].size(); ++k, ++na) {
            if (ra2i_tgt) ra2i_tgt[na] = d->a_tgt[i][k];
            if (ra2i_w) ra2i_w[na] = d->a_w[i][k];
            if (ra2i_d) ra2i_d[na] = d->a_d[i][k];
        }
    }
    if (ra2i_ptr) ra2i_ptr[d->n_ra] = static_cast<int>(na);
    for (int i = 0; i < d->n_i; ++i) {
        if (i2ra_ptr) i2ra_ptr[i] = static_cast<int>(nb);
        for (size_t k = 0; k < d->b_tgt[i].size(); ++k, ++nb) {
            if (i2ra_tgt) i2ra_tgt[nb] = d->b_tgt[i][k];
            if (i2ra_w) i2ra_w[nb] = d->b_w[i][k];
            if (i2ra_d) i2ra_d[nb] = d->b_d[i][k];
        }
    }
    if (i2ra_ptr) i2ra_ptr[d->n_i] = static_cast<int>(nb);
    if (n_ra2i) *n_ra2i = na;
    if (n_i2ra) *n_i2ra = nb;
    return HVC_OK;
}

extern "C" hvc_status hvc_driver_rr_delays(hvc_driver* d, double* rr_d)
{
    if (!d || !rr_d) return HVC_E_ARG;
    std::copy(d->rr_d.begin(), d->rr_d.end(), rr_d);
    return HVC_OK;
}

extern "C" int hvc_driver_trial_number(const hvc_driver* d) { return d ? d->trial_number : -1; }

extern "C" hvc_ctx* hvc_driver_engine(hvc_driver* d) { return d ? d->ctx : nullptr; }

extern "C" const char* hvc_driver_last_error(const hvc_driver* d) { return d ? d->err.c_str() : "null driver"; }

extern "C" void hvc_driver_destroy(hvc_driver* d)
{
    if (!d) return;
    if (d->ctx) hvc_destroy(d->ctx);
    delete d;
}
