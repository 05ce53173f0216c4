/*
 * ora_trial.c -- oracle restatement of the growth trial, the chain-test trial and
 * the STDP / synapse-state machine (test infrastructure, see hvc_oracle.h).
 *
 *   growth trial      : NGS:8839-9571 trial_1stSoma_pre_1stSoma_post_delays_fixedSpread
 *   chain-test trial  : NGS:12747-13118 trial_no_stdp_fixedSpread
 *   STDP              : NGS:16346-16361 potentiation_decay, 16363-16415 axon_remodeling,
 *                       16432-16443 update_all_synapses, 16546-16592 update_synapse,
 *                       16615-16634 LTP, 16636-16652 LTD
 *
 * One process holds every neuron, so the MPI reductions of the reference
 * (NGS:9298-9382) reduce to the single-rank case: the local and global
 * accumulators coincide and the gathered post-event list keeps detection order.
 *
 * The RA->RA pairs are a candidate set (CSR by presynaptic neuron, targets
 * ascending); the reference's dense N_RA x N_RA arrays are the candidate set of
 * all j.  Non-candidate pairs keep w = 0 and never change, which requires
 * ACTIVATION_THRESHOLD > 0 and SUPERSYNAPSE_THRESHOLD > 0 (checked at create).
 *
 * Delivery queues are sorted vectors of (delivery_time, id) pairs ordered as
 * std::pair (NGS:9005-9006).  Inserting one spike's items one at a time with
 * upper_bound gives the same sequence as a stable merge of the sorted new
 * items after the existing ones, which is what queue_insert_batch does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hvc_oracle.h"
#include "ora_neuron.h"

#ifdef _OPENMP
#include <omp.h>
#endif
/* the neuron steps of a step run on OpenMP threads from this many neurons on */
#define ORA_OMP_MIN_NEURONS 512

#define STDP_WINDOW 100.0      /* NGS:13 */
#define G_TRAINING_KICK 3.0    /* NGS:22 */
#define EVENT_WINDOW 30.0      /* NGS:8841 */

typedef struct { double t; int id; } qitem;
typedef struct { qitem* a; int n, cap; } queue;
typedef struct { double* a; int n, cap; } dvec;
typedef struct { int* a; int n, cap; } ivec;

static int dvec_push(dvec* v, double x)
{
    if (v->n == v->cap) {
        int nc = v->cap ? 2 * v->cap : 4;
        double* na = (double*)realloc(v->a, (size_t)nc * sizeof(double));
        if (!na) return -1;
        v->a = na; v->cap = nc;
    }
    v->a[v->n++] = x;
    return 0;
}
static int ivec_push(ivec* v, int x)
{
    if (v->n == v->cap) {
        int nc = v->cap ? 2 * v->cap : 8;
        int* na = (int*)realloc(v->a, (size_t)nc * sizeof(int));
        if (!na) return -1;
        v->a = na; v->cap = nc;
    }
    v->a[v->n++] = x;
    return 0;
}

/* std::pair<double,int> operator< */
static int qless(const qitem* x, const qitem* y)
{
    return (x->t < y->t) || (!(y->t < x->t) && x->id < y->id);
}
static int qcmp(const void* p, const void* q)
{
    const qitem* x = (const qitem*)p;
    const qitem* y = (const qitem*)q;
    if (qless(x, y)) return -1;
    if (qless(y, x)) return 1;
    return 0;
}

/* insert items (any order) as repeated upper_bound + insert would */
static int queue_insert_batch(queue* q, qitem* items, int k)
{
    if (k == 0) return 0;
    qsort(items, (size_t)k, sizeof(qitem), qcmp);
    if (q->n + k > q->cap) {
        int nc = q->cap ? q->cap : 16;
        while (nc < q->n + k) nc *= 2;
        qitem* na = (qitem*)realloc(q->a, (size_t)nc * sizeof(qitem));
        if (!na) return -1;
        q->a = na; q->cap = nc;
    }
    /* merge from the back: existing items precede equal new items */
    int i = q->n - 1, j = k - 1, o = q->n + k - 1;
    while (j >= 0) {
        if (i >= 0 && qless(&items[j], &q->a[i]))
            q->a[o--] = q->a[i--];
        else
            q->a[o--] = items[j--];
    }
    q->n += k;
    return 0;
}
static void queue_pop_front(queue* q, int k)
{
    if (k <= 0) return;
    memmove(q->a, q->a + k, (size_t)(q->n - k) * sizeof(qitem));
    q->n -= k;
}

struct ora_sim {
    ora_net_desc d;
    int num_steps;
    double* clock;
    int* is_sync;
    /* connectivity (copied) */
    int *ra2i_ptr, *ra2i_tgt, *i2ra_ptr, *i2ra_tgt, *rr_ptr, *rr_tgt;
    double *ra2i_w, *ra2i_d, *i2ra_w, *i2ra_d, *rr_d;
    int n_edges;
    int dense;                    /* every row holds all N_RA targets */
    /* synapse state */
    double* w;
    uint8_t *act_ind, *sup_ind;
    int *act_list, *act_cnt, *sup_list, *sup_cnt, *remodeled;
    /* neurons */
    double *erest, *gca;
    ora_ra_neuron* ra;
    ora_i_neuron* in;
    uint8_t* is_training;
    double* spread;               /* per RA id */
    /* outputs of the last trial */
    dvec *sp_soma, *sp_dend, *sp_i;
    long warn_ltp, warn_ltd, warn_missing;
    /* per-trial scratch */
    queue *q_ra_i, *q_rr, *q_i_ra;
    dvec* pre_ev;                 /* per candidate edge */
    dvec* post_ev;                /* per RA */
    qitem* scratch;
    /* probe: one state variable of one neuron after every step (tests) */
    int probe_kind, probe_neuron, probe_var;
    double* probe_out;
};

static double probe_value(const ora_sim* s)
{
    double st[ORA_RA_NSTATE];
    if (s->probe_kind == 0)
        ora_ra_get_state(&s->ra[s->probe_neuron], st);
    else
        ora_i_get_state(&s->in[s->probe_neuron], st);
    return st[s->probe_var];
}

int ora_set_probe(ora_sim* s, int kind, int neuron, int var, double* out)
{
    if (out && ((kind == 0 && (neuron < 0 || neuron >= s->d.n_ra || var < 0 || var >= ORA_RA_NSTATE)) ||
                (kind == 1 && (neuron < 0 || neuron >= s->d.n_i || var < 0 || var >= ORA_I_NSTATE)) ||
                (kind != 0 && kind != 1)))
        return -1;
    s->probe_kind = kind;
    s->probe_neuron = neuron;
    s->probe_var = var;
    s->probe_out = out;
    return 0;
}

static int edge_of(const ora_sim* s, int i, int j)
{
    if (s->dense) return s->rr_ptr[i] + j;
    int lo = s->rr_ptr[i], hi = s->rr_ptr[i + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int v = s->rr_tgt[mid];
        if (v == j) return mid;
        if (v < j) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static void* dupmem(const void* p, size_t bytes)
{
    void* q = malloc(bytes ? bytes : 1);
    if (q && p && bytes) memcpy(q, p, bytes);
    return q;
}

ora_sim* ora_create(const ora_net_desc* d)
{
    if (d->n_ra <= 0 || d->n_i < 0) return NULL;
    if (!(d->syn.ACTIVATION_THRESHOLD > 0) || !(d->syn.SUPERSYNAPSE_THRESHOLD > 0)) return NULL;
    ora_sim* s = (ora_sim*)calloc(1, sizeof(ora_sim));
    if (!s) return NULL;
    s->d = *d;
    int nra = d->n_ra, ni = d->n_i;
    s->num_steps = ora_num_steps(d->trial_ms, d->dt);
    s->clock = (double*)malloc(sizeof(double) * (size_t)s->num_steps);
    s->is_sync = (int*)malloc(sizeof(int) * (size_t)s->num_steps);
    ora_clock(d->trial_ms, d->dt, d->sync_ms, s->clock, s->is_sync);

    s->ra2i_ptr = (int*)dupmem(d->ra2i_ptr, sizeof(int) * (size_t)(nra + 1));
    int e1 = s->ra2i_ptr[nra];
    s->ra2i_tgt = (int*)dupmem(d->ra2i_tgt, sizeof(int) * (size_t)e1);
    s->ra2i_w = (double*)dupmem(d->ra2i_w, sizeof(double) * (size_t)e1);
    s->ra2i_d = (double*)dupmem(d->ra2i_d, sizeof(double) * (size_t)e1);
    s->i2ra_ptr = (int*)dupmem(d->i2ra_ptr, sizeof(int) * (size_t)(ni + 1));
    int e2 = s->i2ra_ptr[ni];
    s->i2ra_tgt = (int*)dupmem(d->i2ra_tgt, sizeof(int) * (size_t)e2);
    s->i2ra_w = (double*)dupmem(d->i2ra_w, sizeof(double) * (size_t)e2);
    s->i2ra_d = (double*)dupmem(d->i2ra_d, sizeof(double) * (size_t)e2);
    s->rr_ptr = (int*)dupmem(d->rr_ptr, sizeof(int) * (size_t)(nra + 1));
    s->n_edges = s->rr_ptr[nra];
    s->rr_tgt = (int*)dupmem(d->rr_tgt, sizeof(int) * (size_t)s->n_edges);
    s->rr_d = (double*)dupmem(d->rr_d, sizeof(double) * (size_t)s->n_edges);
    s->dense = 1;
    for (int i = 0; i < nra && s->dense; ++i) {
        if (s->rr_ptr[i + 1] - s->rr_ptr[i] != nra) { s->dense = 0; break; }
        for (int j = 0; j < nra; ++j)
            if (s->rr_tgt[s->rr_ptr[i] + j] != j) { s->dense = 0; break; }
    }
    int E = s->n_edges;
    s->w = (double*)calloc((size_t)E + 1, sizeof(double));
    s->act_ind = (uint8_t*)calloc((size_t)E + 1, 1);
    s->sup_ind = (uint8_t*)calloc((size_t)E + 1, 1);
    s->act_list = (int*)calloc((size_t)E + 1, sizeof(int));
    s->sup_list = (int*)calloc((size_t)E + 1, sizeof(int));
    s->act_cnt = (int*)calloc((size_t)nra, sizeof(int));
    s->sup_cnt = (int*)calloc((size_t)nra, sizeof(int));
    s->remodeled = (int*)calloc((size_t)nra, sizeof(int));
    s->erest = (double*)dupmem(d->erest, sizeof(double) * (size_t)nra);
    s->gca = (double*)dupmem(d->gca, sizeof(double) * (size_t)nra);
    s->ra = (ora_ra_neuron*)calloc((size_t)nra, sizeof(ora_ra_neuron));
    s->in = (ora_i_neuron*)calloc((size_t)(ni ? ni : 1), sizeof(ora_i_neuron));
    s->is_training = (uint8_t*)calloc((size_t)nra, 1);
    s->spread = (double*)calloc((size_t)nra, sizeof(double));
    for (int k = 0; k < d->n_train; ++k) {
        s->is_training[d->train_ids[k]] = 1;
        s->spread[d->train_ids[k]] = d->spread[k];
    }
    s->sp_soma = (dvec*)calloc((size_t)nra, sizeof(dvec));
    s->sp_dend = (dvec*)calloc((size_t)nra, sizeof(dvec));
    s->sp_i = (dvec*)calloc((size_t)(ni ? ni : 1), sizeof(dvec));
    s->q_ra_i = (queue*)calloc((size_t)nra, sizeof(queue));
    s->q_rr = (queue*)calloc((size_t)nra, sizeof(queue));
    s->q_i_ra = (queue*)calloc((size_t)(ni ? ni : 1), sizeof(queue));
    s->pre_ev = (dvec*)calloc((size_t)E + 1, sizeof(dvec));
    s->post_ev = (dvec*)calloc((size_t)nra, sizeof(dvec));
    int maxrow = 1;
    for (int i = 0; i < nra; ++i) {
        int a = s->rr_ptr[i + 1] - s->rr_ptr[i];
        int b = s->ra2i_ptr[i + 1] - s->ra2i_ptr[i];
        if (a > maxrow) maxrow = a;
        if (b > maxrow) maxrow = b;
    }
    for (int i = 0; i < ni; ++i) {
        int b = s->i2ra_ptr[i + 1] - s->i2ra_ptr[i];
        if (b > maxrow) maxrow = b;
    }
    s->scratch = (qitem*)malloc(sizeof(qitem) * (size_t)maxrow);
    return s;
}

static void free_dvecs(dvec* v, int n) { if (v) { for (int i = 0; i < n; ++i) free(v[i].a); free(v); } }
static void free_queues(queue* q, int n) { if (q) { for (int i = 0; i < n; ++i) free(q[i].a); free(q); } }

void ora_destroy(ora_sim* s)
{
    if (!s) return;
    int nra = s->d.n_ra, ni = s->d.n_i ? s->d.n_i : 1;
    free(s->clock); free(s->is_sync);
    free(s->ra2i_ptr); free(s->ra2i_tgt); free(s->ra2i_w); free(s->ra2i_d);
    free(s->i2ra_ptr); free(s->i2ra_tgt); free(s->i2ra_w); free(s->i2ra_d);
    free(s->rr_ptr); free(s->rr_tgt); free(s->rr_d);
    free(s->w); free(s->act_ind); free(s->sup_ind); free(s->act_list); free(s->sup_list);
    free(s->act_cnt); free(s->sup_cnt); free(s->remodeled);
    free(s->erest); free(s->gca); free(s->ra); free(s->in); free(s->is_training); free(s->spread);
    free_dvecs(s->sp_soma, nra); free_dvecs(s->sp_dend, nra); free_dvecs(s->sp_i, ni);
    free_queues(s->q_ra_i, nra); free_queues(s->q_rr, nra); free_queues(s->q_i_ra, ni);
    free_dvecs(s->pre_ev, s->n_edges + 1); free_dvecs(s->post_ev, nra);
    free(s->scratch);
    free(s);
}

int ora_set_synapses(ora_sim* s, const double* w, const int* act_cnt, const int* act_ids,
                     const int* sup_cnt, const int* sup_ids, const int* remodeled)
{
    int nra = s->d.n_ra;
    memcpy(s->w, w, sizeof(double) * (size_t)s->n_edges);
    memset(s->act_ind, 0, (size_t)s->n_edges);
    memset(s->sup_ind, 0, (size_t)s->n_edges);
    int oa = 0, os = 0;
    for (int i = 0; i < nra; ++i) {
        int len = s->rr_ptr[i + 1] - s->rr_ptr[i];
        if (act_cnt[i] < 0 || act_cnt[i] > len || sup_cnt[i] < 0 || sup_cnt[i] > len) return -1;
        if (sup_cnt[i] > s->d.syn.Nss) return -1;
        s->act_cnt[i] = act_cnt[i];
        for (int k = 0; k < act_cnt[i]; ++k) {
            int j = act_ids[oa++];
            int e = edge_of(s, i, j);
            if (e < 0 || s->act_ind[e]) return -1;
            s->act_list[s->rr_ptr[i] + k] = j;
            s->act_ind[e] = 1;
        }
        s->sup_cnt[i] = sup_cnt[i];
        for (int k = 0; k < sup_cnt[i]; ++k) {
            int j = sup_ids[os++];
            int e = edge_of(s, i, j);
            if (e < 0 || s->sup_ind[e]) return -1;
            s->sup_list[s->rr_ptr[i] + k] = j;
            s->sup_ind[e] = 1;
        }
        s->remodeled[i] = remodeled[i];
    }
    return 0;
}

int ora_get_synapses(const ora_sim* s, double* w, int* act_cnt, int* act_ids,
                     int* sup_cnt, int* sup_ids, int* remodeled)
{
    int nra = s->d.n_ra;
    if (w) memcpy(w, s->w, sizeof(double) * (size_t)s->n_edges);
    int oa = 0, os = 0;
    for (int i = 0; i < nra; ++i) {
        if (act_cnt) act_cnt[i] = s->act_cnt[i];
        if (sup_cnt) sup_cnt[i] = s->sup_cnt[i];
        if (act_ids) for (int k = 0; k < s->act_cnt[i]; ++k) act_ids[oa++] = s->act_list[s->rr_ptr[i] + k];
        if (sup_ids) for (int k = 0; k < s->sup_cnt[i]; ++k) sup_ids[os++] = s->sup_list[s->rr_ptr[i] + k];
        if (remodeled) remodeled[i] = s->remodeled[i];
    }
    return 0;
}

int ora_set_neuron_params(ora_sim* s, const double* erest, const double* gca)
{
    memcpy(s->erest, erest, sizeof(double) * (size_t)s->d.n_ra);
    memcpy(s->gca, gca, sizeof(double) * (size_t)s->d.n_ra);
    return 0;
}

/* ---------------- STDP (NGS:16615-16652) ---------------- */
double ora_ltp(const ora_synaptic_params* p, double w, double dt, int* rejected)
{
    if (dt <= p->T_0) { if (rejected) *rejected = 1; return w; }
    if (rejected) *rejected = 0;
    if (dt <= p->T_0 + p->T_P)
        w = w + p->A_P * (dt - p->T_0) / p->T_P;
    else
        w = w + p->A_P * ora_exp(-(dt - p->T_0 - p->T_P) / p->TAU_P);
    if (w > p->WEIGHT_MAX) w = p->WEIGHT_MAX;
    if (w < 0) w = 0;
    return w;
}

double ora_ltd(const ora_synaptic_params* p, double w, double dt, int* rejected)
{
    if (dt > p->T_0) { if (rejected) *rejected = 1; return w; }
    if (rejected) *rejected = 0;
    if (dt >= p->T_0 - p->T_D)
        w = w + w * p->A_D * (dt - p->T_0) / p->T_D;
    else
        w = w - w * p->A_D * ora_exp((dt - p->T_0 + p->T_D) / p->TAU_D);
    if (w < 0) w = 0;
    return w;
}

static void do_ltp(ora_sim* s, int e, double dt)
{
    int rej;
    s->w[e] = ora_ltp(&s->d.syn, s->w[e], dt, &rej);
    s->warn_ltp += rej;
}
static void do_ltd(ora_sim* s, int e, double dt)
{
    int rej;
    s->w[e] = ora_ltd(&s->d.syn, s->w[e], dt, &rej);
    s->warn_ltd += rej;
}

/* erase first occurrence of j from list (std::find + erase) */
static int list_erase(int* list, int* cnt, int j)
{
    for (int k = 0; k < *cnt; ++k) {
        if (list[k] == j) {
            memmove(list + k, list + k + 1, (size_t)(*cnt - k - 1) * sizeof(int));
            *cnt -= 1;
            return 1;
        }
    }
    return 0;
}

/* NGS:16546-16592 */
static void update_synapse(ora_sim* s, int i, int e)
{
    const ora_synaptic_params* p = &s->d.syn;
    int j = s->rr_tgt[e];
    double w = s->w[e];
    int* alist = s->act_list + s->rr_ptr[i];
    int* slist = s->sup_list + s->rr_ptr[i];
    if ((w >= p->ACTIVATION_THRESHOLD) && (s->act_ind[e] == 0) && (s->remodeled[i] == 0)) {
        s->act_ind[e] = 1;
        alist[s->act_cnt[i]++] = j;
    }
    if ((w >= p->SUPERSYNAPSE_THRESHOLD) && (s->sup_ind[e] == 0) && (s->sup_cnt[i] < p->Nss)) {
        s->sup_ind[e] = 1;
        slist[s->sup_cnt[i]++] = j;
    }
    if ((w < p->SUPERSYNAPSE_THRESHOLD) && (s->sup_ind[e] == 1)) {
        s->sup_ind[e] = 0;
        s->remodeled[i] = 0;
        if (!list_erase(slist, &s->sup_cnt[i], j)) s->warn_missing++;
    }
    if ((w < p->ACTIVATION_THRESHOLD) && (s->act_ind[e] == 1)) {
        s->act_ind[e] = 0;
        if (!list_erase(alist, &s->act_cnt[i], j)) s->warn_missing++;
    }
}

/* update_synapse(i, j) for j = 0..N_RA-1 (NGS:9144-9146, 16432-16443) */
static void update_row(ora_sim* s, int i)
{
    for (int e = s->rr_ptr[i]; e < s->rr_ptr[i + 1]; ++e) update_synapse(s, i, e);
}

/* NGS:16363-16415 */
static void axon_remodeling(ora_sim* s, int i)
{
    int base = s->rr_ptr[i];
    for (int k = 0; k < s->act_cnt[i]; ++k) s->act_ind[edge_of(s, i, s->act_list[base + k])] = 0;
    for (int k = 0; k < s->sup_cnt[i]; ++k) s->act_ind[edge_of(s, i, s->sup_list[base + k])] = 1;
    memcpy(s->act_list + base, s->sup_list + base, sizeof(int) * (size_t)s->sup_cnt[i]);
    s->act_cnt[i] = s->sup_cnt[i];
    s->remodeled[i] = 1;
}

/* NGS:16346-16361 */
static void potentiation_decay(ora_sim* s)
{
    const ora_synaptic_params* p = &s->d.syn;
    for (int e = 0; e < s->n_edges; ++e) {
        if (s->w[e] >= p->SUPERSYNAPSE_THRESHOLD)
            s->w[e] -= p->BETA_SUPERSYNAPSE;
        else
            s->w[e] -= p->BETA;
        if (s->w[e] < 0) s->w[e] = 0;
    }
}

/* first index k with v[k] >= x (std::lower_bound) */
static int lower_bound_d(const double* v, int n, double x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void reset_trial_state(ora_sim* s, uint32_t trial_index, int first_trial)
{
    const ora_net_desc* d = &s->d;
    for (int i = 0; i < d->n_ra; ++i) {
        ora_ra_neuron* a = &s->ra[i];
        ora_ra_init(a, d->ra_model, d->dt, s->erest[i], s->gca[i], d->mu_s, d->sigma_s, d->mu_d,
                    d->sigma_d, d->ra_poisson);
        a->st.seed_lo = d->seed_lo; a->st.seed_hi = d->seed_hi;
        a->st.trial = trial_index; a->st.neuron = (uint32_t)i; a->st.draw = 0;
        ora_ra_set_to_rest(a);
        s->sp_soma[i].n = 0; s->sp_dend[i].n = 0;
        s->q_ra_i[i].n = 0; s->q_rr[i].n = 0;
        s->post_ev[i].n = 0;
    }
    for (int i = 0; i < d->n_i; ++i) {
        ora_i_neuron* b = &s->in[i];
        ora_i_init(b, d->i_model, d->dt, d->i_poisson);
        b->st.seed_lo = d->seed_lo; b->st.seed_hi = d->seed_hi;
        b->st.trial = trial_index; b->st.neuron = (uint32_t)i; b->st.draw = 0;
        ora_i_set_to_rest(b, first_trial ? 0.0 : d->dt);
        s->sp_i[i].n = 0;
        s->q_i_ra[i].n = 0;
    }
    for (int e = 0; e < s->n_edges; ++e) s->pre_ev[e].n = 0;
    s->warn_ltp = s->warn_ltd = s->warn_missing = 0;
}

/* LTD over the not-too-old post events of target (NGS:9111-9137 / 9150-9180) */
static void ltd_on_delivery(ora_sim* s, int i, int e, int tgt, double delivery_time, double internal_time)
{
    dvec* pe = &s->post_ev[tgt];
    if (pe->n == 0) return;
    int k = lower_bound_d(pe->a, pe->n, internal_time - STDP_WINDOW);
    for (; k < pe->n; ++k) {
        double dt = pe->a[k] - delivery_time;
        do_ltd(s, e, dt);
        update_synapse(s, i, e);
    }
}

int ora_run_trial(ora_sim* s, int mode, int training, uint32_t trial_index, int first_trial,
                  double mean_injection_time)
{
    const ora_net_desc* d = &s->d;
    const int nra = d->n_ra, ni = d->n_i;
    const int Nss = d->syn.Nss;
    const int growth = (mode == 0);
    if (mode != 0 && mode != 1) return -2;
    if (!growth) training = 1;  /* the chain test always kicks the training neurons (NGS:12807) */
    reset_trial_state(s, trial_index, first_trial);

    double* update_Ge_RA = (double*)calloc((size_t)nra, sizeof(double));
    double* update_Gi_RA = (double*)calloc((size_t)nra, sizeof(double));
    double* update_Ge_I = (double*)calloc((size_t)(ni ? ni : 1), sizeof(double));
    uint8_t* injected = (uint8_t*)calloc((size_t)nra, 1);
    ivec new_post = {0, 0, 0};
    int flag_ra_exc = 0, flag_ra_inh = 0, flag_i_exc = 0, flag_post = 0;

    double internal_time = 0;
    double network_time = d->sync_ms;
    const int num_steps = s->num_steps;
    if (s->probe_out) s->probe_out[0] = probe_value(s);
    for (int t = 1; t < num_steps; ++t) {
        internal_time += d->dt;

        /* ---------------- HVC-RA neurons ----------------
         * The steps first: a neuron's step reads only its own state (the
         * deliveries of this step go to the update_* sums, applied at the
         * sync), so stepping every neuron before any delivery gives the
         * reference's values; the threads share nothing. */
#pragma omp parallel for schedule(static) if (nra >= ORA_OMP_MIN_NEURONS)
        for (int i = 0; i < nra; ++i) {
            ora_ra_neuron* a = &s->ra[i];
            if (training && s->is_training[i] && !injected[i]) {
                if (internal_time > s->spread[i] + mean_injection_time) {
                    a->Ge_d += G_TRAINING_KICK;      /* raiseE */
                    injected[i] = 1;
                }
            }
            ora_ra_step(a);
        }
        /* spikes and delivery queues, neuron by neuron in the reference's order */
        for (int i = 0; i < nra; ++i) {
            ora_ra_neuron* a = &s->ra[i];
            if (a->fired_soma) {
                dvec_push(&s->sp_soma[i], internal_time);
                if (growth) {
                    dvec* pe = &s->post_ev[i];
                    if (pe->n == 0 || internal_time > pe->a[pe->n - 1] + EVENT_WINDOW) {
                        ivec_push(&new_post, i);
                        flag_post = 1;
                    }
                }
                /* RA -> I */
                int k = 0;
                for (int j = s->ra2i_ptr[i]; j < s->ra2i_ptr[i + 1]; ++j) {
                    s->scratch[k].t = internal_time + s->ra2i_d[j];
                    s->scratch[k].id = j - s->ra2i_ptr[i];
                    ++k;
                }
                queue_insert_batch(&s->q_ra_i[i], s->scratch, k);
                /* RA -> RA */
                k = 0;
                int base = s->rr_ptr[i];
                if (growth) {
                    if (s->sup_cnt[i] == Nss) {
                        for (int q = 0; q < s->sup_cnt[i]; ++q) {
                            int tgt = s->sup_list[base + q];
                            s->scratch[k].t = internal_time + s->rr_d[edge_of(s, i, tgt)];
                            s->scratch[k].id = tgt;
                            ++k;
                        }
                    } else {
                        for (int e = base; e < s->rr_ptr[i + 1]; ++e) {
                            if (s->rr_tgt[e] == i) continue;
                            s->scratch[k].t = internal_time + s->rr_d[e];
                            s->scratch[k].id = s->rr_tgt[e];
                            ++k;
                        }
                    }
                } else {
                    for (int q = 0; q < s->act_cnt[i]; ++q) {
                        int tgt = s->act_list[base + q];
                        s->scratch[k].t = internal_time + s->rr_d[edge_of(s, i, tgt)];
                        s->scratch[k].id = tgt;
                        ++k;
                    }
                }
                queue_insert_batch(&s->q_rr[i], s->scratch, k);
            }

            /* RA -> I delivery queue (NGS:9039-9062) */
            {
                queue* q = &s->q_ra_i[i];
                int k = 0;
                for (; k < q->n; ++k) {
                    if (internal_time >= q->a[k].t) {
                        flag_i_exc = 1;
                        int slot = s->ra2i_ptr[i] + q->a[k].id;
                        update_Ge_I[s->ra2i_tgt[slot]] += s->ra2i_w[slot];
                    } else
                        break;
                }
                queue_pop_front(q, k);
            }
            /* RA -> RA delivery queue (NGS:9069-9191 growth, 13234-13250 chain) */
            {
                queue* q = &s->q_rr[i];
                int k = 0;
                for (; k < q->n; ++k) {
                    double delivery_time = q->a[k].t;
                    if (internal_time >= delivery_time) {
                        int tgt = q->a[k].id;
                        int e = edge_of(s, i, tgt);
                        if (!growth) {
                            flag_ra_exc = 1;
                            update_Ge_RA[tgt] += s->w[e];
                            continue;
                        }
                        if (s->act_ind[e] == 1) {
                            flag_ra_exc = 1;
                            update_Ge_RA[tgt] += s->w[e];
                        }
                        int new_pre = 0;
                        dvec* pv = &s->pre_ev[e];
                        if (pv->n > 0) {
                            if (delivery_time > pv->a[pv->n - 1] + EVENT_WINDOW) {
                                dvec_push(pv, delivery_time);
                                new_pre = 1;
                            }
                        } else {
                            dvec_push(pv, delivery_time);
                            new_pre = 1;
                        }
                        if (new_pre) {
                            if (s->sup_cnt[i] == Nss) {
                                if (s->sup_ind[e] == 1) ltd_on_delivery(s, i, e, tgt, delivery_time, internal_time);
                                if (s->sup_cnt[i] < Nss) update_row(s, i);
                            } else {
                                ltd_on_delivery(s, i, e, tgt, delivery_time, internal_time);
                            }
                        }
                    } else
                        break;
                }
                queue_pop_front(q, k);
            }
            if (a->fired_dend) dvec_push(&s->sp_dend[i], internal_time);
        }

        /* ---------------- HVC-I neurons (steps first, as above) ---------------- */
#pragma omp parallel for schedule(static) if (ni >= ORA_OMP_MIN_NEURONS)
        for (int i = 0; i < ni; ++i) ora_i_step(&s->in[i]);
        for (int i = 0; i < ni; ++i) {
            ora_i_neuron* b = &s->in[i];
            if (b->fired) {
                dvec_push(&s->sp_i[i], internal_time);
                int k = 0;
                for (int j = s->i2ra_ptr[i]; j < s->i2ra_ptr[i + 1]; ++j) {
                    s->scratch[k].t = internal_time + s->i2ra_d[j];
                    s->scratch[k].id = j - s->i2ra_ptr[i];
                    ++k;
                }
                queue_insert_batch(&s->q_i_ra[i], s->scratch, k);
            }
            queue* q = &s->q_i_ra[i];
            int k = 0;
            for (; k < q->n; ++k) {
                if (internal_time >= q->a[k].t) {
                    flag_ra_inh = 1;
                    int slot = s->i2ra_ptr[i] + q->a[k].id;
                    update_Gi_RA[s->i2ra_tgt[slot]] += s->i2ra_w[slot];
                } else
                    break;
            }
            queue_pop_front(q, k);
        }

        /* ---------------- synchronisation (NGS:9291-9564) ---------------- */
        if ((internal_time > network_time) || (t == num_steps - 1)) {
            if (flag_ra_inh) {
                for (int i = 0; i < nra; ++i) s->ra[i].Gi_d += update_Gi_RA[i];  /* raiseI */
                flag_ra_inh = 0;
                memset(update_Gi_RA, 0, sizeof(double) * (size_t)nra);
            }
            if (flag_i_exc) {
                for (int i = 0; i < ni; ++i) s->in[i].Ge += update_Ge_I[i];      /* HHI raiseE */
                flag_i_exc = 0;
                memset(update_Ge_I, 0, sizeof(double) * (size_t)ni);
            }
            if (flag_ra_exc) {
                for (int i = 0; i < nra; ++i) s->ra[i].Ge_d += update_Ge_RA[i] / 10000.0; /* raiseExcWeight */
                flag_ra_exc = 0;
                memset(update_Ge_RA, 0, sizeof(double) * (size_t)nra);
            }
            if (growth && flag_post) {
                flag_post = 0;
                for (int k = 0; k < new_post.n; ++k) dvec_push(&s->post_ev[new_post.a[k]], internal_time);
                for (int i = 0; i < nra; ++i) {
                    int base = s->rr_ptr[i];
                    int saturated = (s->sup_cnt[i] == Nss);
                    for (int k = 0; k < new_post.n; ++k) {
                        int post = new_post.a[k];
                        if (post == i) continue;
                        int e;
                        if (saturated) {
                            int found = 0;
                            for (int q = 0; q < s->sup_cnt[i]; ++q)
                                if (s->sup_list[base + q] == post) { found = 1; break; }
                            if (!found) continue;
                        }
                        e = edge_of(s, i, post);
                        if (e < 0) continue;
                        dvec* pv = &s->pre_ev[e];
                        if (pv->n == 0) continue;
                        int it = lower_bound_d(pv->a, pv->n, internal_time - STDP_WINDOW);
                        for (; it < pv->n; ++it) {
                            double dt = internal_time - pv->a[it];
                            if (dt <= d->syn.T_0)
                                do_ltd(s, e, dt);
                            else
                                do_ltp(s, e, dt);
                            update_synapse(s, i, e);
                        }
                    }
                }
                for (int i = 0; i < nra; ++i)
                    if ((s->sup_cnt[i] == Nss) && (s->remodeled[i] == 0)) axon_remodeling(s, i);
                new_post.n = 0;
            }
            network_time += d->sync_ms;
        }
        if (s->probe_out) s->probe_out[t] = probe_value(s);
    }

    if (growth) {
        potentiation_decay(s);
        for (int i = 0; i < nra; ++i) update_row(s, i);
    }
    free(update_Ge_RA); free(update_Gi_RA); free(update_Ge_I); free(injected); free(new_post.a);
    for (int i = 0; i < nra; ++i)
        if (ora_ra_check_bad_values(&s->ra[i]) == -1) return -1;
    return 0;
}

int ora_update_synapse(ora_sim* s, int i, int j)
{
    if (i < 0 || i >= s->d.n_ra) return -1;
    int e = edge_of(s, i, j);
    if (e < 0) return -1;
    update_synapse(s, i, e);
    return 0;
}

int ora_axon_remodeling(ora_sim* s, int i)
{
    if (i < 0 || i >= s->d.n_ra) return -1;
    axon_remodeling(s, i);
    return 0;
}

int ora_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ora_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int ora_spike_counts(const ora_sim* s, int kind, int* counts)
{
    int n = (kind == 2) ? s->d.n_i : s->d.n_ra;
    const dvec* v = (kind == 0) ? s->sp_soma : (kind == 1) ? s->sp_dend : s->sp_i;
    int tot = 0;
    for (int i = 0; i < n; ++i) { if (counts) counts[i] = v[i].n; tot += v[i].n; }
    return tot;
}

int ora_spike_times(const ora_sim* s, int kind, double* times)
{
    int n = (kind == 2) ? s->d.n_i : s->d.n_ra;
    const dvec* v = (kind == 0) ? s->sp_soma : (kind == 1) ? s->sp_dend : s->sp_i;
    int o = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < v[i].n; ++k) times[o++] = v[i].a[k];
    return o;
}

int ora_warnings(const ora_sim* s, int* out3)
{
    out3[0] = (int)s->warn_ltp; out3[1] = (int)s->warn_ltd; out3[2] = (int)s->warn_missing;
    return 0;
}

int ora_final_state(const ora_sim* s, double* ra, double* in)
{
    for (int i = 0; i < s->d.n_ra; ++i) ora_ra_get_state(&s->ra[i], ra + (size_t)i * ORA_RA_NSTATE);
    for (int i = 0; i < s->d.n_i; ++i) ora_i_get_state(&s->in[i], in + (size_t)i * ORA_I_NSTATE);
    return 0;
}
