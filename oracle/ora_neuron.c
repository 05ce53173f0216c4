/*
 * ora_neuron.c -- oracle restatement of the reference neuron classes
 * (test infrastructure, see hvc_oracle.h).
 *
 *   HVC-RA drop-in  : HH2_test        (c++/src/HH2_test.cpp, c++/include/HH2_test.h)
 *   HVC-RA RK4 form : HH2_buffer      (c++/src/HH2_buffer.cpp, c++/include/HH2_buffer.h)
 *   HVC-I           : HHI_buffer      (c++/src/HHI_buffer.cpp, c++/include/HHI_buffer.h)
 *   Noise draws     : Poisson_noise   (c++/src/poisson_noise.cpp:9-26)
 *
 * Each expression keeps the reference's operand order and association
 * (left-to-right C++ evaluation, integer literals converted to double).
 */
#include <math.h>
#include <stdint.h>

#include "hvc_oracle.h"
#include "ora_neuron.h"

/* ---------------- Poisson_noise draws (poisson_noise.cpp:9-21) ---------------- */
static double draw_u(ora_stream* s, uint32_t stream)
{
    uint32_t u = ora_poisson_u32(s->seed_lo, s->seed_hi, s->trial, stream, s->neuron, s->draw++);
    return (double)u / 4294967295.0;
}

/* get_spike_time(lambda) = -log(1 - random) / lambda */
static double get_spike_time(ora_stream* s, uint32_t stream, double lambda)
{
    double random = draw_u(s, stream);
    return -ora_log(1.0 - random) / lambda;
}

/* random(G) = (double) G * generator() / generator.max() */
static double random_G(ora_stream* s, uint32_t stream, double G)
{
    uint32_t u = ora_poisson_u32(s->seed_lo, s->seed_hi, s->trial, stream, s->neuron, s->draw++);
    return G * (double)u / 4294967295.0;
}

/* ======================= HVC-RA ======================= */
/* model parameters: HH2_test.cpp:262-289 (set_original_model_parameters) and the
 * constructor defaults 31-35; HH2_buffer.cpp has the same values without GsKLT. */
#define RA_CM 1.0
#define RA_AS 5000.0
#define RA_AD 10000.0
#define RA_RC 55.0
#define RA_GSKLT 3.5
#define RA_GSK 8.0
#define RA_GSNA 60.0
#define RA_GDCAK 150.0
#define RA_GSL 0.1
#define RA_GDL 0.1
#define RA_EDCA 120.0
#define RA_EDK (-90.0)
#define RA_ESNA 55.0
#define RA_ESK (-90.0)
#define RA_EGABA (-80.0)
#define RA_TEXC 5.0
#define RA_TINH 5.0
/* HH2_test.cpp:14-28 */
#define RA_GS_NOISE_INH 0.035
#define RA_GD_NOISE_INH 0.045
#define RA_GS_NOISE_EXC 0.035
#define RA_GD_NOISE_EXC 0.045
#define RA_LAMBDA_EXC 100.0
#define RA_LAMBDA_INH 100.0
#define RA_THRESHOLD_SPIKE (-30.0)
#define RA_THRESHOLD_BURST (-30.0)
#define RA_SPIKE_MARGIN 5.0
#define RA_BURST_MARGIN 5.0

/* HH2_test.cpp:33-35, point_distribution; sqrt(6) = 0x1.3988e1409212ep+1 */
static const double POINT_DISTRIBUTION[30] = {
    0x1.3988e1409212ep+1, -0x1.3988e1409212ep+1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

/* gating functions, HH2_test.h:265-276 (HH2_buffer.h:260-268) */
static double nInf(double v) { return 1 / (1 + ora_exp(-(v + 35) / 10)); }
static double tauN(double v) { return 0.1 + 0.5 / (1 + ora_exp((v + 27) / 15)); }
static double hInf(double v) { return 1 / (1 + ora_exp((v + 45) / 7)); }
static double tauH(double v) { return 0.1 + 0.75 / (1 + ora_exp((v + 40.5) / 6)); }
static double mInf(double v) { return 1 / (1 + ora_exp(-(v + 30) / 9.5)); }
static double lInf(double v) { return 1 / (1 + ora_exp(-(v + 40) / 5)); }
static double rInf(double v) { return 1 / (1 + ora_exp(-(v + 5) / 10)); }
static double cInf(double v) { return 1 / (1 + ora_exp(-(v - 10) / 7)); }
/* tauL = 10.0, tauR = 1, tauC = 15 (HH2_test) or 10 (HH2_buffer) */

/* conductances, HH2_test.cpp:861-864 */
static double Gi_s(const ora_ra_neuron* a, double t) { return a->Gi_s * ora_exp(-(t - a->time) / RA_TINH); }
static double Gi_d(const ora_ra_neuron* a, double t) { return a->Gi_d * ora_exp(-(t - a->time) / RA_TINH); }
static double Ge_d(const ora_ra_neuron* a, double t) { return a->Ge_d * ora_exp(-(t - a->time) / RA_TEXC); }
static double Ge_s(const ora_ra_neuron* a, double t) { return a->Ge_s * ora_exp(-(t - a->time) / RA_TEXC); }

/* kVs: HH2_test.cpp:866-874 (K-LT term present for model 0 only, HH2_buffer.cpp:851-859) */
static double kVs(const ora_ra_neuron* a, double vs, double vd, double n, double h, double l, double t)
{
    double m3, n4;
    m3 = mInf(vs) * mInf(vs) * mInf(vs);
    n4 = n * n * n * n;
    double acc = -RA_GSL * (vs - a->EL) - RA_GSNA * m3 * h * (vs - RA_ESNA) - RA_GSK * n4 * (vs - RA_ESK);
    if (a->model == 0) acc = acc - RA_GSKLT * l * (vs - RA_ESK);
    acc = acc - Gi_s(a, t) * (vs - RA_EGABA) - Ge_s(a, t) * vs
          + 100000 * (0.0 + a->mu_s) / RA_AS + 100000 * (vd - vs) / (RA_RC * RA_AS);
    return acc / RA_CM;
}

/* kVd: HH2_test.cpp:875-882 */
static double kVd(const ora_ra_neuron* a, double vs, double vd, double r, double c, double ca, double t)
{
    double r2;
    r2 = r * r;
    return (-RA_GDL * (vd - a->EL) - a->GCa * r2 * (vd - RA_EDCA) - RA_GDCAK * (vd - RA_EDK) * c / (1 + 6 / ca)
            - Ge_d(a, t) * vd - Gi_d(a, t) * (vd - RA_EGABA) + 100000 * (0.0 + a->mu_d) / RA_AD
            + 100000 * (vs - vd) / (RA_RC * RA_AD)) / RA_CM;
}

/* HH2_test.cpp:884-890 */
static double kn(double vs, double n) { return (nInf(vs) - n) / tauN(vs); }
static double kl(double vs, double l) { return (lInf(vs) - l) / 10.0; }
static double kh(double vs, double h) { return (hInf(vs) - h) / tauH(vs); }
static double kr(double vd, double r) { return (rInf(vd) - r) / 1; }
static double kc(const ora_ra_neuron* a, double vd, double c)
{
    return (cInf(vd) - c) / (a->model == 0 ? 15 : 10);
}
static double kCa(const ora_ra_neuron* a, double vd, double r, double ca)
{
    return (-0.1 * a->GCa * r * r * (vd - RA_EDCA) - 0.02 * ca);
}

void ora_ra_init(ora_ra_neuron* a, int model, double dt, double EL, double GCa,
                 double mu_s, double sigma_s, double mu_d, double sigma_d, int poisson)
{
    a->model = model;
    a->dt = dt;
    a->EL = EL;
    a->GCa = GCa;
    a->mu_s = mu_s; a->sigma_s = sigma_s; a->mu_d = mu_d; a->sigma_d = sigma_d;
    a->poisson = poisson;
    a->fired_soma = a->fired_dend = 0;
    a->noise_exc_soma = a->noise_inh_soma = a->noise_exc_dend = a->noise_inh_dend = 0.0;
    a->step = 1;
}

/* HH2_test.cpp:148-165 initialize_poisson_noise */
static void ra_init_noise(ora_ra_neuron* a, double* noise_time, double lambda)
{
    *noise_time = 1000.0 * get_spike_time(&a->st, ORA_STREAM_RA_POISSON, lambda);
}

/* HH2_test.cpp:191-242 */
void ora_ra_set_to_rest(ora_ra_neuron* a)
{
    a->time = 0.0;
    a->Vs = -79.97619025;
    a->Vd = -79.97268759;
    a->l = 0.0;
    a->n = 0.01101284;
    a->h = 0.9932845;
    a->r = 0.00055429;
    a->c = 0.00000261762353;
    a->Ca = 0.01689572;
    a->Gi_s = 0.0; a->Ge_s = 0.0; a->Gi_d = 0.0; a->Ge_d = 0.0;
    a->flag_soma = 0;
    a->flag_dend = 0;
    if (a->poisson) {
        ra_init_noise(a, &a->noise_exc_soma, RA_LAMBDA_EXC);
        ra_init_noise(a, &a->noise_inh_soma, RA_LAMBDA_INH);
        ra_init_noise(a, &a->noise_exc_dend, RA_LAMBDA_EXC);
        ra_init_noise(a, &a->noise_inh_dend, RA_LAMBDA_INH);
    }
}

int ora_ra_check_bad_values(const ora_ra_neuron* a)
{
    if ((a->Vs > 100) || (a->Vs < -200)) return -1;
    if (isnan(a->Vs)) return -1;
    if ((a->Vd > 100) || (a->Vd < -200)) return -1;
    if (isnan(a->Vd)) return -1;
    return 0;
}

/* HH2_test.cpp:561-594 */
static void ra_state_check(ora_ra_neuron* a)
{
    if ((a->flag_soma == 1) && (a->Vs < RA_THRESHOLD_SPIKE - RA_SPIKE_MARGIN)) {
        a->flag_soma = 0;
        a->fired_soma = 1;
    } else {
        a->fired_soma = 0;
        if ((a->flag_soma == 0) && (a->Vs > RA_THRESHOLD_SPIKE)) a->flag_soma = 1;
    }
    if ((a->flag_dend == 0) && (a->Vd >= RA_THRESHOLD_BURST)) {
        a->flag_dend = 1;
        a->fired_dend = 1;
    } else {
        a->fired_dend = 0;
        if ((a->flag_dend == 1) && (a->Vd < RA_THRESHOLD_BURST - RA_BURST_MARGIN)) a->flag_dend = 0;
    }
}

/* HH2_test.cpp:596-611 */
static void ra_noise_check(ora_ra_neuron* a, double* G, double G_noise, double lambda, double* noise_time)
{
    if (a->time > *noise_time) {
        *G += random_G(&a->st, ORA_STREAM_RA_POISSON, G_noise);
        double random = 1000.0 * get_spike_time(&a->st, ORA_STREAM_RA_POISSON, lambda);
        while (random < a->dt) {
            random = 1000.0 * get_spike_time(&a->st, ORA_STREAM_RA_POISSON, lambda);
            *G += random_G(&a->st, ORA_STREAM_RA_POISSON, G_noise);
        }
        *noise_time = *noise_time + random;
    }
}

/* HH2_test.cpp:613-754 */
static void ra_debraband(ora_ra_neuron* a)
{
    int ind[4];
    ora_white_indices(a->st.seed_lo, a->st.seed_hi, a->st.trial, a->st.neuron, a->step, ind);
    double J1s = POINT_DISTRIBUTION[ind[0]];
    double J2s = POINT_DISTRIBUTION[ind[1]];
    double J1d = POINT_DISTRIBUTION[ind[2]];
    double J2d = POINT_DISTRIBUTION[ind[3]];
    const double dt = a->dt;
    const double sq = sqrt(dt);

    double H1Vs, H1Vd, H1Ca, H1l, H1n, H1h, H1c, H1r;
    double a1Vs, a1Vd, a1n, a1l, a1h, a1r, a1c, a1Ca;
    H1Vs = a->Vs + sq * a->sigma_s * (-0.01844540496323970 * J1s - 0.1866426386543421 * J2s) / (RA_AS * RA_CM);
    H1Vd = a->Vd + sq * a->sigma_d * (-0.01844540496323970 * J1d - 0.1866426386543421 * J2d) / (RA_AD * RA_CM);
    H1Ca = a->Ca; H1l = a->l; H1n = a->n; H1h = a->h; H1c = a->c; H1r = a->r;
    a1Vs = kVs(a, H1Vs, H1Vd, H1n, H1h, H1l, a->time);
    a1Vd = kVd(a, H1Vs, H1Vd, H1r, H1c, H1Ca, a->time);
    a1n = kn(H1Vs, H1n); a1l = kl(H1Vs, H1l); a1h = kh(H1Vs, H1h);
    a1r = kr(H1Vd, H1r); a1c = kc(a, H1Vd, H1c); a1Ca = kCa(a, H1Vd, H1r, H1Ca);

    double H2Vs, H2Vd, H2Ca, H2n, H2l, H2h, H2c, H2r;
    double a2Vs, a2Vd, a2n, a2l, a2h, a2r, a2c, a2Ca;
    H2Vs = a->Vs + dt * a1Vs + sq * a->sigma_s * (0.8017012756521233 * J1s - 0.8575745885712401 * J2s) / (RA_AS * RA_CM);
    H2Vd = a->Vd + dt * a1Vd + sq * a->sigma_d * (0.8017012756521233 * J1d - 0.8575745885712401 * J2d) / (RA_AD * RA_CM);
    H2Ca = a->Ca + dt * a1Ca; H2n = a->n + dt * a1n; H2l = a->l + dt * a1l;
    H2h = a->h + dt * a1h; H2c = a->c + dt * a1c; H2r = a->r + dt * a1r;
    a2Vs = kVs(a, H2Vs, H2Vd, H2n, H2h, H2l, a->time + dt);
    a2Vd = kVd(a, H2Vs, H2Vd, H2r, H2c, H2Ca, a->time + dt);
    a2n = kn(H2Vs, H2n); a2l = kl(H2Vs, H2l); a2h = kh(H2Vs, H2h);
    a2r = kr(H2Vd, H2r); a2c = kc(a, H2Vd, H2c); a2Ca = kCa(a, H2Vd, H2r, H2Ca);

    double H3Vs, H3Vd, H3Ca, H3n, H3l, H3h, H3c, H3r;
    double a3Vs, a3Vd, a3n, a3l, a3h, a3r, a3c, a3Ca;
    H3Vs = a->Vs + dt * (3 * a1Vs + a2Vs) / 8 + sq * a->sigma_s * (0.5092227024816198 * J1s - 0.4723392695015512 * J2s) / (RA_AS * RA_CM);
    H3Vd = a->Vd + dt * (3 * a1Vd + a2Vd) / 8 + sq * a->sigma_d * (0.5092227024816198 * J1d - 0.4723392695015512 * J2d) / (RA_AD * RA_CM);
    H3Ca = a->Ca + dt * (3 * a1Ca + a2Ca) / 8;
    H3n = a->n + dt * (3 * a1n + a2n) / 8;
    H3l = a->l + dt * (3 * a1l + a2l) / 8;
    H3h = a->h + dt * (3 * a1h + a2h) / 8;
    H3c = a->c + dt * (3 * a1c + a2c) / 8;
    H3r = a->r + dt * (3 * a1r + a2r) / 8;
    a3Vs = kVs(a, H3Vs, H3Vd, H3n, H3h, H3l, a->time + 0.5 * dt);
    a3Vd = kVd(a, H3Vs, H3Vd, H3r, H3c, H3Ca, a->time + 0.5 * dt);
    a3n = kn(H3Vs, H3n); a3l = kl(H3Vs, H3l); a3h = kh(H3Vs, H3h);
    a3r = kr(H3Vd, H3r); a3c = kc(a, H3Vd, H3c); a3Ca = kCa(a, H3Vd, H3r, H3Ca);

    double H4Vs, H4Vd, H4Ca, H4n, H4l, H4h, H4c, H4r;
    double a4Vs, a4Vd, a4n, a4l, a4h, a4r, a4c, a4Ca;
    H4Vs = a->Vs + dt * (-0.4526683126055039 * a1Vs - 0.4842227708685013 * a2Vs + 1.9368910834740051 * a3Vs)
           + sq * a->sigma_s * (0.9758794209767762 * J1s + 0.3060354860326548 * J2s) / (RA_AS * RA_CM);
    H4Vd = a->Vd + dt * (-0.4526683126055039 * a1Vd - 0.4842227708685013 * a2Vd + 1.9368910834740051 * a3Vd)
           + sq * a->sigma_d * (0.9758794209767762 * J1d + 0.3060354860326548 * J2d) / (RA_AD * RA_CM);
    H4Ca = a->Ca + dt * (-0.4526683126055039 * a1Ca - 0.4842227708685013 * a2Ca + 1.9368910834740051 * a3Ca);
    H4n = a->n + dt * (-0.4526683126055039 * a1n - 0.4842227708685013 * a2n + 1.9368910834740051 * a3n);
    H4l = a->l + dt * (-0.4526683126055039 * a1l - 0.4842227708685013 * a2l + 1.9368910834740051 * a3l);
    H4h = a->h + dt * (-0.4526683126055039 * a1h - 0.4842227708685013 * a2h + 1.9368910834740051 * a3h);
    H4c = a->c + dt * (-0.4526683126055039 * a1c - 0.4842227708685013 * a2c + 1.9368910834740051 * a3c);
    H4r = a->r + dt * (-0.4526683126055039 * a1r - 0.4842227708685013 * a2r + 1.9368910834740051 * a3r);
    a4Vs = kVs(a, H4Vs, H4Vd, H4n, H4h, H4l, a->time + dt);
    a4Vd = kVd(a, H4Vs, H4Vd, H4r, H4c, H4Ca, a->time + dt);
    a4n = kn(H4Vs, H4n); a4l = kl(H4Vs, H4l); a4h = kh(H4Vs, H4h);
    a4r = kr(H4Vd, H4r); a4c = kc(a, H4Vd, H4c); a4Ca = kCa(a, H4Vd, H4r, H4Ca);

    a->Vs = a->Vs + dt * (a1Vs / 6 - 0.005430430675258792 * a2Vs + 2 * a3Vs / 3 + 0.1720970973419255 * a4Vs)
            + sq * a->sigma_s * J1s * 100000 / (RA_AS * RA_CM);
    a->Vd = a->Vd + dt * (a1Vd / 6 - 0.005430430675258792 * a2Vd + 2 * a3Vd / 3 + 0.1720970973419255 * a4Vd)
            + sq * a->sigma_d * J1d * 100000 / (RA_AD * RA_CM);
    a->Ca = a->Ca + dt * (a1Ca / 6 - 0.005430430675258792 * a2Ca + 2 * a3Ca / 3 + 0.1720970973419255 * a4Ca);
    a->n = a->n + dt * (a1n / 6 - 0.005430430675258792 * a2n + 2 * a3n / 3 + 0.1720970973419255 * a4n);
    a->l = a->l + dt * (a1l / 6 - 0.005430430675258792 * a2l + 2 * a3l / 3 + 0.1720970973419255 * a4l);
    a->h = a->h + dt * (a1h / 6 - 0.005430430675258792 * a2h + 2 * a3h / 3 + 0.1720970973419255 * a4h);
    a->r = a->r + dt * (a1r / 6 - 0.005430430675258792 * a2r + 2 * a3r / 3 + 0.1720970973419255 * a4r);
    a->c = a->c + dt * (a1c / 6 - 0.005430430675258792 * a2c + 2 * a3c / 3 + 0.1720970973419255 * a4c);

    double t1 = a->time + dt;
    a->Ge_d = Ge_d(a, t1);
    a->Ge_s = Ge_s(a, t1);
    a->Gi_d = Gi_d(a, t1);
    a->Gi_s = Gi_s(a, t1);
    a->time = a->time + dt;
}

/* HH2_buffer.cpp:741-843 (Kutta's 3/8 rule) */
static void ra_runge4(ora_ra_neuron* a)
{
    const double dt = a->dt;
    double n1, h1, c1, r1, vts, vtd, Cat, t;
    double k1Vs, k2Vs, k3Vs, k4Vs, k1Vd, k2Vd, k3Vd, k4Vd, k1Ca, k2Ca, k3Ca, k4Ca;
    double k1n, k2n, k3n, k4n, k1h, k2h, k3h, k4h, k1c, k2c, k3c, k4c, k1r, k2r, k3r, k4r;

    vts = a->Vs; n1 = a->n; h1 = a->h; vtd = a->Vd; r1 = a->r; c1 = a->c; Cat = a->Ca; t = a->time;
    k1Vs = kVs(a, vts, vtd, n1, h1, 0.0, t); k1n = kn(vts, n1); k1h = kh(vts, h1);
    k1Vd = kVd(a, vts, vtd, r1, c1, Cat, t); k1r = kr(vtd, r1); k1c = kc(a, vtd, c1); k1Ca = kCa(a, vtd, r1, Cat);

    vts = a->Vs + dt * k1Vs / 3; n1 = a->n + dt * k1n / 3; h1 = a->h + dt * k1h / 3;
    vtd = a->Vd + dt * k1Vd / 3; r1 = a->r + dt * k1r / 3; c1 = a->c + dt * k1c / 3; Cat = a->Ca + dt * k1Ca / 3;
    t = a->time + dt / 3;
    k2Vs = kVs(a, vts, vtd, n1, h1, 0.0, t); k2n = kn(vts, n1); k2h = kh(vts, h1);
    k2Vd = kVd(a, vts, vtd, r1, c1, Cat, t); k2r = kr(vtd, r1); k2c = kc(a, vtd, c1); k2Ca = kCa(a, vtd, r1, Cat);

    vts = a->Vs + dt * (-k1Vs / 3 + k2Vs); n1 = a->n + dt * (-k1n / 3 + k2n); h1 = a->h + dt * (-k1h / 3 + k2h);
    vtd = a->Vd + dt * (-k1Vd / 3 + k2Vd); r1 = a->r + dt * (-k1r / 3 + k2r); c1 = a->c + dt * (-k1c / 3 + k2c);
    Cat = a->Ca + dt * (-k1Ca / 3 + k2Ca);
    t = a->time + 2 * dt / 3;
    k3Vs = kVs(a, vts, vtd, n1, h1, 0.0, t); k3n = kn(vts, n1); k3h = kh(vts, h1);
    k3Vd = kVd(a, vts, vtd, r1, c1, Cat, t); k3r = kr(vtd, r1); k3c = kc(a, vtd, c1); k3Ca = kCa(a, vtd, r1, Cat);

    vts = a->Vs + dt * (k1Vs - k2Vs + k3Vs); n1 = a->n + dt * (k1n - k2n + k3n); h1 = a->h + dt * (k1h - k2h + k3h);
    vtd = a->Vd + dt * (k1Vd - k2Vd + k3Vd); r1 = a->r + dt * (k1r - k2r + k3r); c1 = a->c + dt * (k1c - k2c + k3c);
    Cat = a->Ca + dt * (k1Ca - k2Ca + k3Ca);
    t = a->time + dt;
    k4Vs = kVs(a, vts, vtd, n1, h1, 0.0, t); k4n = kn(vts, n1); k4h = kh(vts, h1);
    k4Vd = kVd(a, vts, vtd, r1, c1, Cat, t); k4r = kr(vtd, r1); k4c = kc(a, vtd, c1); k4Ca = kCa(a, vtd, r1, Cat);

    double t1 = a->time + dt;
    a->Ge_d = Ge_d(a, t1);
    a->Ge_s = Ge_s(a, t1);
    a->Gi_d = Gi_d(a, t1);
    a->Gi_s = Gi_s(a, t1);

    a->Vs = a->Vs + dt * (k1Vs + 3 * k2Vs + 3 * k3Vs + k4Vs) / 8;
    a->n = a->n + dt * (k1n + 3 * k2n + 3 * k3n + k4n) / 8;
    a->h = a->h + dt * (k1h + 3 * k2h + 3 * k3h + k4h) / 8;
    a->Vd = a->Vd + dt * (k1Vd + 3 * k2Vd + 3 * k3Vd + k4Vd) / 8;
    a->r = a->r + dt * (k1r + 3 * k2r + 3 * k3r + k4r) / 8;
    a->c = a->c + dt * (k1c + 3 * k2c + 3 * k3c + k4c) / 8;
    a->Ca = a->Ca + dt * (k1Ca + 3 * k2Ca + 3 * k3Ca + k4Ca) / 8;
    a->time = a->time + dt;
}

/* HH2_test.cpp:535-559 (model 0), HH2_buffer.cpp:521-528 (model 1) */
void ora_ra_step(ora_ra_neuron* a)
{
    ra_state_check(a);
    if (a->poisson) {
        ra_noise_check(a, &a->Gi_s, RA_GS_NOISE_INH, RA_LAMBDA_INH, &a->noise_inh_soma);
        ra_noise_check(a, &a->Gi_d, RA_GD_NOISE_INH, RA_LAMBDA_INH, &a->noise_inh_dend);
        ra_noise_check(a, &a->Ge_s, RA_GS_NOISE_EXC, RA_LAMBDA_EXC, &a->noise_exc_soma);
        ra_noise_check(a, &a->Ge_d, RA_GD_NOISE_EXC, RA_LAMBDA_EXC, &a->noise_exc_dend);
    }
    if (a->model == 0)
        ra_debraband(a);
    else
        ra_runge4(a);
    a->step += 1;
}

void ora_ra_get_state(const ora_ra_neuron* a, double* s)
{
    s[0] = a->Vs; s[1] = a->Vd; s[2] = a->n; s[3] = a->l; s[4] = a->h; s[5] = a->r;
    s[6] = a->c; s[7] = a->Ca; s[8] = a->Ge_s; s[9] = a->Gi_s; s[10] = a->Ge_d; s[11] = a->Gi_d;
}

/* ======================= HVC-I ======================= */
/* HHI_buffer.cpp:13-37 */
#define I_CM 1.0
#define I_A 6000.0
#define I_ENA 55.0
#define I_EK (-80.0)
#define I_EL (-65.0)
#define I_EI (-75.0)
#define I_GNA 100.0
#define I_GKDR 20.0
#define I_GKHT 500.0
#define I_GL 0.1
#define I_TEXC 2.0
#define I_TINH 5.0
#define I_THRESHOLD (-20.0)
#define I_SPIKE_MARGIN 5.0
#define I_G_NOISE 0.45
#define I_LAMBDA 250.0

/* HHI_buffer.h:173-180 */
static double an(double v) { return 0.15 * (v + 15) / (1 - ora_exp(-(v + 15) / 10)); }
static double bn(double v) { return 0.2 * ora_exp(-(v + 25) / 80); }
static double am(double v) { return (v + 22) / (1 - ora_exp(-(v + 22) / 10)); }
static double bm(double v) { return 40 * ora_exp(-(v + 47) / 18); }
static double ah(double v) { return 0.7 * ora_exp(-(v + 34) / 20); }
static double bh(double v) { return 10 / (1 + ora_exp(-(v + 4) / 10)); }
static double wInf(double v) { return 1 / (1 + ora_exp(-v / 5)); }

/* HHI_buffer.cpp:739-747 */
static double Ge(const ora_i_neuron* b, double t) { return b->Ge * ora_exp(-(t - b->time) / I_TEXC); }
static double Gi(const ora_i_neuron* b, double t) { return b->Gi * ora_exp(-(t - b->time) / I_TINH); }
static double kV(const ora_i_neuron* b, double v, double t, double h, double w, double m3, double n4)
{
    return (-I_GL * (v - I_EL) - I_GNA * h * m3 * (v - I_ENA) - I_GKDR * n4 * (v - I_EK)
            - I_GKHT * w * (v - I_EK) - Ge(b, t) * v - Gi(b, t) * (v - I_EI) + 100000 * 0.0 / I_A) / I_CM;
}
static double kni(double v, double n) { return an(v) * (1 - n) - bn(v) * n; }
static double kmi(double v, double m) { return am(v) * (1 - m) - bm(v) * m; }
static double khi(double v, double h) { return ah(v) * (1 - h) - bh(v) * h; }
static double kwi(double v, double w) { return (wInf(v) - w) / 1; }

void ora_i_init(ora_i_neuron* b, int model, double dt, int poisson)
{
    b->model = model;
    b->dt = dt;
    b->poisson = poisson;
    b->fired = 0;
    b->noise_exc = b->noise_inh = 0.0;
}

/* HHI_buffer.cpp:217-239; init_dt is the neuron's _timestep at the time of the call */
static void i_init_noise(ora_i_neuron* b, double* noise_time, double init_dt)
{
    *noise_time = 1000 * get_spike_time(&b->st, ORA_STREAM_I_POISSON, I_LAMBDA);
    while (*noise_time < init_dt)
        *noise_time = 1000 * get_spike_time(&b->st, ORA_STREAM_I_POISSON, I_LAMBDA);
}

/* HHI_buffer.cpp:178-200 */
void ora_i_set_to_rest(ora_i_neuron* b, double init_dt)
{
    b->time = 0.0;
    b->V = -66;
    b->n = 0.125;
    b->m = 0.0;
    b->h = 0.99;
    b->w = 0.0;
    b->Ge = 0.0;
    b->Gi = 0.0;
    b->flag = 0;
    if (b->poisson) {
        i_init_noise(b, &b->noise_exc, init_dt);
        i_init_noise(b, &b->noise_inh, init_dt);
    }
}

/* HHI_buffer.cpp:260-276 */
static void i_state_check(ora_i_neuron* b)
{
    if ((b->flag == 1) && (b->V < I_THRESHOLD - I_SPIKE_MARGIN)) {
        b->flag = 0;
        b->fired = 1;
    } else {
        b->fired = 0;
        if ((b->flag == 0) && (b->V > I_THRESHOLD)) b->flag = 1;
    }
}

/* HHI_buffer.cpp:278-294 */
static void i_noise_check(ora_i_neuron* b, double* G, double* noise_time)
{
    if (b->time > *noise_time) {
        *G += random_G(&b->st, ORA_STREAM_I_POISSON, I_G_NOISE);
        double random = 1000 * get_spike_time(&b->st, ORA_STREAM_I_POISSON, I_LAMBDA);
        while (random < b->dt) {
            random = 1000 * get_spike_time(&b->st, ORA_STREAM_I_POISSON, I_LAMBDA);
            *G += random_G(&b->st, ORA_STREAM_I_POISSON, I_G_NOISE);
        }
        *noise_time = *noise_time + random;
    }
}

#define I_STAGE(kV_, kn_, km_, kh_, kw_)            \
    n4 = n1 * n1 * n1 * n1;                          \
    m3 = m1 * m1 * m1;                               \
    kV_ = kV(b, v, t, h1, w1, m3, n4);               \
    kn_ = kni(v, n1);                                \
    km_ = kmi(v, m1);                                \
    kh_ = khi(v, h1);                                \
    kw_ = kwi(v, w1);

/* HHI_buffer.cpp:426-691 */
static void i_dp8(ora_i_neuron* b)
{
    const double dt = b->dt;
    double m1, n1, h1, w1, m3, n4, v, t;
    double k1V, k2V, k3V, k4V, k5V, k6V, k7V, k8V, k9V, k10V, k11V, k12V;
    double k1n, k2n, k3n, k4n, k5n, k6n, k7n, k8n, k9n, k10n, k11n, k12n;
    double k1m, k2m, k3m, k4m, k5m, k6m, k7m, k8m, k9m, k10m, k11m, k12m;
    double k1h, k2h, k3h, k4h, k5h, k6h, k7h, k8h, k9h, k10h, k11h, k12h;
    double k1w, k2w, k3w, k4w, k5w, k6w, k7w, k8w, k9w, k10w, k11w, k12w;

    m1 = b->m; n1 = b->n; h1 = b->h; w1 = b->w; v = b->V; t = b->time;
    I_STAGE(k1V, k1n, k1m, k1h, k1w)

#define UPD(expr_m, expr_n, expr_h, expr_w, expr_v) \
    m1 = b->m + dt * (expr_m); n1 = b->n + dt * (expr_n); h1 = b->h + dt * (expr_h); \
    w1 = b->w + dt * (expr_w); v = b->V + dt * (expr_v);

    m1 = b->m + dt * k1m / 18; n1 = b->n + dt * k1n / 18; h1 = b->h + dt * k1h / 18;
    w1 = b->w + dt * k1w / 18; v = b->V + dt * k1V / 18;
    t = b->time + dt / 18;
    I_STAGE(k2V, k2n, k2m, k2h, k2w)

    m1 = b->m + dt * (k1m + 3 * k2m) / 48; n1 = b->n + dt * (k1n + 3 * k2n) / 48;
    h1 = b->h + dt * (k1h + 3 * k2h) / 48; w1 = b->w + dt * (k1w + 3 * k2w) / 48;
    v = b->V + dt * (k1V + 3 * k2V) / 48;
    t = b->time + dt / 12;
    I_STAGE(k3V, k3n, k3m, k3h, k3w)

    m1 = b->m + dt * (k1m + 3 * k3m) / 32; n1 = b->n + dt * (k1n + 3 * k3n) / 32;
    h1 = b->h + dt * (k1h + 3 * k3h) / 32; w1 = b->w + dt * (k1w + 3 * k3w) / 32;
    v = b->V + dt * (k1V + 3 * k3V) / 32;
    t = b->time + dt / 8;
    I_STAGE(k4V, k4n, k4m, k4h, k4w)

    m1 = b->m + dt * (20 * k1m - 75 * k3m + 75 * k4m) / 64; n1 = b->n + dt * (20 * k1n - 75 * k3n + 75 * k4n) / 64;
    h1 = b->h + dt * (20 * k1h - 75 * k3h + 75 * k4h) / 64; w1 = b->w + dt * (20 * k1w - 75 * k3w + 75 * k4w) / 64;
    v = b->V + dt * (20 * k1V - 75 * k3V + 75 * k4V) / 64;
    t = b->time + 5 * dt / 16;
    I_STAGE(k5V, k5n, k5m, k5h, k5w)

    m1 = b->m + dt * (3 * k1m + 15 * k4m + 12 * k5m) / 80; n1 = b->n + dt * (3 * k1n + 15 * k4n + 12 * k5n) / 80;
    h1 = b->h + dt * (3 * k1h + 15 * k4h + 12 * k5h) / 80; w1 = b->w + dt * (3 * k1w + 15 * k4w + 12 * k5w) / 80;
    v = b->V + dt * (3 * k1V + 15 * k4V + 12 * k5V) / 80;
    t = b->time + 3 * dt / 8;
    I_STAGE(k6V, k6n, k6m, k6h, k6w)

#define S7(k1, k4, k5, k6) \
    (29443841 * k1 / 614563906 + 77736538 * k4 / 692538347 - 28693883 * k5 / 1125000000 + 23124283 * k6 / 1800000000)
    UPD(S7(k1m, k4m, k5m, k6m), S7(k1n, k4n, k5n, k6n), S7(k1h, k4h, k5h, k6h), S7(k1w, k4w, k5w, k6w),
        S7(k1V, k4V, k5V, k6V))
    t = b->time + 59 * dt / 400;
    I_STAGE(k7V, k7n, k7m, k7h, k7w)

#define S8(k1, k4, k5, k6, k7)                                                                   \
    (16016141 * k1 / 946692911 + 61564180 * k4 / 158732637 + 22789713 * k5 / 633445777 +      \
     545815736 * k6 / 2771057229 - 180193667 * k7 / 1043307555)
    UPD(S8(k1m, k4m, k5m, k6m, k7m), S8(k1n, k4n, k5n, k6n, k7n), S8(k1h, k4h, k5h, k6h, k7h),
        S8(k1w, k4w, k5w, k6w, k7w), S8(k1V, k4V, k5V, k6V, k7V))
    t = b->time + 93 * dt / 200;
    I_STAGE(k8V, k8n, k8m, k8h, k8w)

#define S9(k1, k4, k5, k6, k7, k8)                                                                \
    (39632708 * k1 / 573591083 - 433636366 * k4 / 683701615 - 421739975 * k5 / 2616292301 +   \
     100302831 * k6 / 723423059 + 790204164 * k7 / 839813087 + 800635310 * k8 / 3783071287)
    UPD(S9(k1m, k4m, k5m, k6m, k7m, k8m), S9(k1n, k4n, k5n, k6n, k7n, k8n), S9(k1h, k4h, k5h, k6h, k7h, k8h),
        S9(k1w, k4w, k5w, k6w, k7w, k8w), S9(k1V, k4V, k5V, k6V, k7V, k8V))
    t = b->time + 5490023248 * dt / 9719169821;
    I_STAGE(k9V, k9n, k9m, k9h, k9w)

#define S10(k1, k4, k5, k6, k7, k8, k9)                                                            \
    (246121993 * k1 / 1340847787 - 37695042795 * k4 / 15268766246 - 309121744 * k5 / 1061227803 - \
     12992083 * k6 / 490766935 + 6005943493 * k7 / 2108947869 + 393006217 * k8 / 1396673457 +    \
     123872331 * k9 / 1001029789)
    UPD(S10(k1m, k4m, k5m, k6m, k7m, k8m, k9m), S10(k1n, k4n, k5n, k6n, k7n, k8n, k9n),
        S10(k1h, k4h, k5h, k6h, k7h, k8h, k9h), S10(k1w, k4w, k5w, k6w, k7w, k8w, k9w),
        S10(k1V, k4V, k5V, k6V, k7V, k8V, k9V))
    t = b->time + 13 * dt / 20;
    I_STAGE(k10V, k10n, k10m, k10h, k10w)

#define S11(k1, k4, k5, k6, k7, k8, k9, k10)                                                       \
    (-1028468189 * k1 / 846180014 + 8478235783 * k4 / 508512852 + 1311729495 * k5 / 1432422823 -  \
     10304129995 * k6 / 1701304382 - 48777925059 * k7 / 3047939560 + 15336726248 * k8 / 1032824649 - \
     45442868181 * k9 / 3398467696 + 3065993473 * k10 / 597172653)
    UPD(S11(k1m, k4m, k5m, k6m, k7m, k8m, k9m, k10m), S11(k1n, k4n, k5n, k6n, k7n, k8n, k9n, k10n),
        S11(k1h, k4h, k5h, k6h, k7h, k8h, k9h, k10h), S11(k1w, k4w, k5w, k6w, k7w, k8w, k9w, k10w),
        S11(k1V, k4V, k5V, k6V, k7V, k8V, k9V, k10V))
    t = b->time + 1201146811 * dt / 12990119798;
    I_STAGE(k11V, k11n, k11m, k11h, k11w)

#define S12(k1, k4, k5, k6, k7, k8, k9, k10, k11)                                                   \
    (185892177 * k1 / 718116043 - 3185094517 * k4 / 667107341 - 477755414 * k5 / 1098053517 -      \
     703635378 * k6 / 230739211 + 5731566787 * k7 / 1027545527 + 5232866602 * k8 / 850066563 -     \
     4093664535 * k9 / 808688257 + 3962137247 * k10 / 1805957418 + 65686358 * k11 / 487910083)
    UPD(S12(k1m, k4m, k5m, k6m, k7m, k8m, k9m, k10m, k11m), S12(k1n, k4n, k5n, k6n, k7n, k8n, k9n, k10n, k11n),
        S12(k1h, k4h, k5h, k6h, k7h, k8h, k9h, k10h, k11h), S12(k1w, k4w, k5w, k6w, k7w, k8w, k9w, k10w, k11w),
        S12(k1V, k4V, k5V, k6V, k7V, k8V, k9V, k10V, k11V))
    t = b->time + dt;
    I_STAGE(k12V, k12n, k12m, k12h, k12w)

#define SF(k1, k6, k7, k8, k9, k10, k11, k12)                                                        \
    (13451932 * k1 / 455176623 - 808719846 * k6 / 976000145 + 1757004468 * k7 / 5645159321 +       \
     656045339 * k8 / 265891186 - 3867574721 * k9 / 1518517206 + 465885868 * k10 / 322736535 +     \
     53011238 * k11 / 667516719 + 2 * k12 / 45)
    b->V = b->V + dt * SF(k1V, k6V, k7V, k8V, k9V, k10V, k11V, k12V);
    b->n = b->n + dt * SF(k1n, k6n, k7n, k8n, k9n, k10n, k11n, k12n);
    b->m = b->m + dt * SF(k1m, k6m, k7m, k8m, k9m, k10m, k11m, k12m);
    b->h = b->h + dt * SF(k1h, k6h, k7h, k8h, k9h, k10h, k11h, k12h);
    b->w = b->w + dt * SF(k1w, k6w, k7w, k8w, k9w, k10w, k11w, k12w);
    b->Ge = Ge(b, b->time + dt);
    b->Gi = Gi(b, b->time + dt);
    b->time = b->time + dt;
#undef UPD
#undef S7
#undef S8
#undef S9
#undef S10
#undef S11
#undef S12
#undef SF
}

/* HHI_buffer.cpp:341-424 */
static void i_runge4(ora_i_neuron* b)
{
    const double dt = b->dt;
    double m1, n1, h1, w1, m3, n4, v, t;
    double k1V, k2V, k3V, k4V, k1n, k2n, k3n, k4n, k1m, k2m, k3m, k4m;
    double k1h, k2h, k3h, k4h, k1w, k2w, k3w, k4w;

    m1 = b->m; n1 = b->n; h1 = b->h; w1 = b->w; v = b->V; t = b->time;
    I_STAGE(k1V, k1n, k1m, k1h, k1w)
    m1 = b->m + dt * k1m / 3; n1 = b->n + dt * k1n / 3; h1 = b->h + dt * k1h / 3;
    w1 = b->w + dt * k1w / 3; v = b->V + dt * k1V / 3;
    t = b->time + dt / 3;
    I_STAGE(k2V, k2n, k2m, k2h, k2w)
    m1 = b->m + dt * (-k1m / 3 + k2m); n1 = b->n + dt * (-k1n / 3 + k2n); h1 = b->h + dt * (-k1h / 3 + k2h);
    w1 = b->w + dt * (-k1w / 3 + k2w); v = b->V + dt * (-k1V / 3 + k2V);
    t = b->time + 2 * dt / 3;
    I_STAGE(k3V, k3n, k3m, k3h, k3w)
    m1 = b->m + dt * (k1m - k2m + k3m); n1 = b->n + dt * (k1n - k2n + k3n); h1 = b->h + dt * (k1h - k2h + k3h);
    w1 = b->w + dt * (k1w - k2w + k3w); v = b->V + dt * (k1V - k2V + k3V);
    t = b->time + dt;
    I_STAGE(k4V, k4n, k4m, k4h, k4w)
    b->V = b->V + dt * (k1V + 3 * k2V + 3 * k3V + k4V) / 8;
    b->n = b->n + dt * (k1n + 3 * k2n + 3 * k3n + k4n) / 8;
    b->m = b->m + dt * (k1m + 3 * k2m + 3 * k3m + k4m) / 8;
    b->h = b->h + dt * (k1h + 3 * k2h + 3 * k3h + k4h) / 8;
    b->w = b->w + dt * (k1w + 3 * k2w + 3 * k3w + k4w) / 8;
    b->Ge = Ge(b, b->time + dt);
    b->Gi = Gi(b, b->time + dt);
    b->time = b->time + dt;
}

/* HHI_buffer.cpp:323-339 */
void ora_i_step(ora_i_neuron* b)
{
    i_state_check(b);
    if (b->poisson) {
        i_noise_check(b, &b->Ge, &b->noise_exc);
        i_noise_check(b, &b->Gi, &b->noise_inh);
    }
    if (b->model == 0)
        i_dp8(b);
    else
        i_runge4(b);
}

void ora_i_get_state(const ora_i_neuron* b, double* s)
{
    s[0] = b->V; s[1] = b->n; s[2] = b->m; s[3] = b->h; s[4] = b->w; s[5] = b->Ge; s[6] = b->Gi;
}
