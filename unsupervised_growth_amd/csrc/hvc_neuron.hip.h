// hvc_neuron.hip.h -- FP64 neuron right-hand sides and steppers for gfx950.
//
// HVC-RA, model 0: HH2_test::Debraband_step (c++/src/HH2_test.cpp:613-754) with
//                  the RHS of HH2_test.cpp:866-890 and gates of HH2_test.h:265-276.
// HVC-RA, model 1: HH2_buffer::Runge4_step (c++/src/HH2_buffer.cpp:741-843),
//                  RHS HH2_buffer.cpp:846-874 (no K-LT current, tauC = 10).
// HVC-I,  model 0: HHI_buffer::DP8_step (c++/src/HHI_buffer.cpp:426-691).
// HVC-I,  model 1: HHI_buffer::Runge4_step (c++/src/HHI_buffer.cpp:341-424).
//
// State lives in registers for a whole coupling window.  Each expression keeps
// the reference's operand order (the build disables FMA contraction), so the
// FP64 results equal the CPU evaluation of the same trees bit for bit.  The
// conductance decay factors exp(-((time + c dt) - time) / tau) depend only on
// the shared clock, so they come from a per-step table (DecayRow) computed once
// per trial length with the same exp.
#pragma once
#include "hvc_math.hip.h"

namespace hvc {

// constant-divisor quotient inside a stepper: IEEE division when EXACT, else the
// branch-free Markstein form, whose guard `bad` tells when the step must be repeated
#define CD(a, B) (EXACT ? (a) / (double)(B) : ::hvc::cdiv_fast((a), (double)(B), 1.0 / (double)(B), bad))

// offsets c*dt of the stage times used by the steppers, in this order
enum DecayOffset {
    kOffDt = 0,      // dt
    kOffHalf,        // 0.5 * dt
    kOffThird,       // dt / 3
    kOffTwoThird,    // 2 * dt / 3
    kOffDp2, kOffDp3, kOffDp4, kOffDp5, kOffDp6, kOffDp7, kOffDp8, kOffDp9, kOffDp10, kOffDp11,
    kNumOffsets
};

// one row per step: E5[o] = exp(-(fl(c+off_o)-c)/5), E2[o] = same with tau 2
struct DecayRow {
    double e5[kNumOffsets];
    double e2[kNumOffsets];
};

// ---------------- HVC-RA ----------------
struct RAState {
    double Vs, Vd, n, l, h, r, c, Ca;
    double Ge_s, Gi_s, Ge_d, Gi_d;
};

// per-network constants derived once on the host with the reference's own expressions
struct RAUniform {
    double sq_sig_s;   // sqrt(_timestep) * _sigma_soma
    double sq_sig_d;   // sqrt(_timestep) * _sigma_dend
    double mu_term_s;  // 100000 * (IsExt(t) + _mu_soma) / As
    double mu_term_d;  // 100000 * (IdExt(t) + _mu_dend) / Ad
    double dt;
};

constexpr double kRA_As = 5000.0, kRA_Ad = 10000.0, kRA_Rc = 55.0, kRA_cm = 1.0;
constexpr double kRA_GsKLT = 3.5, kRA_GsK = 8.0, kRA_GsNa = 60.0, kRA_GdCaK = 150.0;
constexpr double kRA_GsL = 0.1, kRA_GdL = 0.1, kRA_EdCa = 120.0, kRA_EdK = -90.0;
constexpr double kRA_EsNa = 55.0, kRA_EsK = -90.0, kRA_Egaba = -80.0;

template <bool EXACT>
__device__ __forceinline__ double nInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-(v + 35), 10))); }
template <bool EXACT>
__device__ __forceinline__ double tauN(double v, ::hvc::DivGuard& bad) { return 0.1 + 0.5 / (1 + dexp(CD(v + 27, 15))); }
template <bool EXACT>
__device__ __forceinline__ double hInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(v + 45, 7))); }
template <bool EXACT>
__device__ __forceinline__ double tauH(double v, ::hvc::DivGuard& bad) { return 0.1 + 0.75 / (1 + dexp(CD(v + 40.5, 6))); }
template <bool EXACT>
__device__ __forceinline__ double mInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-(v + 30), 9.5))); }
template <bool EXACT>
__device__ __forceinline__ double lInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-(v + 40), 5))); }
template <bool EXACT>
__device__ __forceinline__ double rInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-(v + 5), 10))); }
template <bool EXACT>
__device__ __forceinline__ double cInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-(v - 10), 7))); }

struct RAK {
    double Vs, Vd, n, l, h, r, c, Ca;
};

// all eight derivatives at one stage; E = decay factor of the stage time
template <int MODEL, bool EXACT>
__device__ __forceinline__ RAK ra_rhs(::hvc::DivGuard& bad, const RAState& s, double EL, double GCa, const RAUniform& u,
                                      double vs, double vd, double n, double h, double l, double r,
                                      double c, double ca, double E)
{
    RAK k;
    double m = mInf<EXACT>(vs, bad);
    double m3 = m * m * m;
    double n4 = n * n * n * n;
    double acc = -kRA_GsL * (vs - EL) - kRA_GsNa * m3 * h * (vs - kRA_EsNa) - kRA_GsK * n4 * (vs - kRA_EsK);
    if (MODEL == 0) acc = acc - kRA_GsKLT * l * (vs - kRA_EsK);
    acc = acc - (s.Gi_s * E) * (vs - kRA_Egaba) - (s.Ge_s * E) * vs + u.mu_term_s
          + CD(100000 * (vd - vs), 275000);
    k.Vs = acc;  // divided by cm = 1.0 (exact)
    double r2 = r * r;
    k.Vd = -kRA_GdL * (vd - EL) - GCa * r2 * (vd - kRA_EdCa) - kRA_GdCaK * (vd - kRA_EdK) * c / (1 + 6 / ca)
           - (s.Ge_d * E) * vd - (s.Gi_d * E) * (vd - kRA_Egaba) + u.mu_term_d
           + CD(100000 * (vs - vd), 550000);
    k.n = (nInf<EXACT>(vs, bad) - n) / tauN<EXACT>(vs, bad);
    k.l = (MODEL == 0) ? CD(lInf<EXACT>(vs, bad) - l, 10) : 0.0;
    k.h = (hInf<EXACT>(vs, bad) - h) / tauH<EXACT>(vs, bad);
    k.r = (rInf<EXACT>(vd, bad) - r);  // divided by tauR = 1 (exact)
    k.c = (MODEL == 0) ? CD(cInf<EXACT>(vd, bad) - c, 15) : CD(cInf<EXACT>(vd, bad) - c, 10);
    k.Ca = (-0.1 * GCa * r * r * (vd - kRA_EdCa) - 0.02 * ca);
    return k;
}

// HH2_test::Debraband_step; J* are the point-distribution values of this step
template <bool EXACT>
__device__ __forceinline__ void ra_debraband(::hvc::DivGuard& bad, RAState& s, double EL, double GCa, const RAUniform& u,
                                             double J1s, double J2s, double J1d, double J2d,
                                             double E_dt, double E_half)
{
    const double dt = u.dt;
    double H1Vs = s.Vs + CD(u.sq_sig_s * (-0.01844540496323970 * J1s - 0.1866426386543421 * J2s), 5000);
    double H1Vd = s.Vd + CD(u.sq_sig_d * (-0.01844540496323970 * J1d - 0.1866426386543421 * J2d), 10000);
    RAK a1 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H1Vs, H1Vd, s.n, s.h, s.l, s.r, s.c, s.Ca, 1.0);

    double H2Vs = s.Vs + dt * a1.Vs + CD(u.sq_sig_s * (0.8017012756521233 * J1s - 0.8575745885712401 * J2s), 5000);
    double H2Vd = s.Vd + dt * a1.Vd + CD(u.sq_sig_d * (0.8017012756521233 * J1d - 0.8575745885712401 * J2d), 10000);
    RAK a2 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H2Vs, H2Vd, s.n + dt * a1.n, s.h + dt * a1.h, s.l + dt * a1.l,
                       s.r + dt * a1.r, s.c + dt * a1.c, s.Ca + dt * a1.Ca, E_dt);

    double H3Vs = s.Vs + dt * (3 * a1.Vs + a2.Vs) * 0.125
                  + CD(u.sq_sig_s * (0.5092227024816198 * J1s - 0.4723392695015512 * J2s), 5000);
    double H3Vd = s.Vd + dt * (3 * a1.Vd + a2.Vd) * 0.125
                  + CD(u.sq_sig_d * (0.5092227024816198 * J1d - 0.4723392695015512 * J2d), 10000);
    RAK a3 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H3Vs, H3Vd, s.n + dt * (3 * a1.n + a2.n) * 0.125,
                       s.h + dt * (3 * a1.h + a2.h) * 0.125, s.l + dt * (3 * a1.l + a2.l) * 0.125,
                       s.r + dt * (3 * a1.r + a2.r) * 0.125, s.c + dt * (3 * a1.c + a2.c) * 0.125,
                       s.Ca + dt * (3 * a1.Ca + a2.Ca) * 0.125, E_half);

#define HVC_H4(x, k) (x + dt * (-0.4526683126055039 * a1.k - 0.4842227708685013 * a2.k + 1.9368910834740051 * a3.k))
    double H4Vs = HVC_H4(s.Vs, Vs) + CD(u.sq_sig_s * (0.9758794209767762 * J1s + 0.3060354860326548 * J2s), 5000);
    double H4Vd = HVC_H4(s.Vd, Vd) + CD(u.sq_sig_d * (0.9758794209767762 * J1d + 0.3060354860326548 * J2d), 10000);
    RAK a4 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H4Vs, H4Vd, HVC_H4(s.n, n), HVC_H4(s.h, h), HVC_H4(s.l, l),
                       HVC_H4(s.r, r), HVC_H4(s.c, c), HVC_H4(s.Ca, Ca), E_dt);
#undef HVC_H4

#define HVC_FIN(k) (dt * (CD(a1.k, 6) - 0.005430430675258792 * a2.k + CD(2 * a3.k, 3) + 0.1720970973419255 * a4.k))
    s.Vs = s.Vs + HVC_FIN(Vs) + CD(u.sq_sig_s * J1s * 100000, 5000);
    s.Vd = s.Vd + HVC_FIN(Vd) + CD(u.sq_sig_d * J1d * 100000, 10000);
    s.Ca = s.Ca + HVC_FIN(Ca);
    s.n = s.n + HVC_FIN(n);
    s.l = s.l + HVC_FIN(l);
    s.h = s.h + HVC_FIN(h);
    s.r = s.r + HVC_FIN(r);
    s.c = s.c + HVC_FIN(c);
#undef HVC_FIN
    s.Ge_d = s.Ge_d * E_dt;
    s.Ge_s = s.Ge_s * E_dt;
    s.Gi_d = s.Gi_d * E_dt;
    s.Gi_s = s.Gi_s * E_dt;
}

// HH2_buffer::Runge4_step (Kutta 3/8 rule)
template <bool EXACT>
__device__ __forceinline__ void ra_runge4(::hvc::DivGuard& bad, RAState& s, double EL, double GCa, const RAUniform& u,
                                          double E_third, double E_twothird, double E_dt)
{
    const double dt = u.dt;
    RAK k1 = ra_rhs<1, EXACT>(bad, s, EL, GCa, u, s.Vs, s.Vd, s.n, s.h, 0.0, s.r, s.c, s.Ca, 1.0);
    RAK k2 = ra_rhs<1, EXACT>(bad, s, EL, GCa, u, s.Vs + CD(dt * k1.Vs, 3), s.Vd + CD(dt * k1.Vd, 3), s.n + CD(dt * k1.n, 3),
                       s.h + CD(dt * k1.h, 3), 0.0, s.r + CD(dt * k1.r, 3), s.c + CD(dt * k1.c, 3),
                       s.Ca + CD(dt * k1.Ca, 3), E_third);
#define HVC_S3(x, k) (x + dt * (CD(-k1.k, 3) + k2.k))
    RAK k3 = ra_rhs<1, EXACT>(bad, s, EL, GCa, u, HVC_S3(s.Vs, Vs), HVC_S3(s.Vd, Vd), HVC_S3(s.n, n), HVC_S3(s.h, h), 0.0,
                       HVC_S3(s.r, r), HVC_S3(s.c, c), HVC_S3(s.Ca, Ca), E_twothird);
#undef HVC_S3
#define HVC_S4(x, k) (x + dt * (k1.k - k2.k + k3.k))
    RAK k4 = ra_rhs<1, EXACT>(bad, s, EL, GCa, u, HVC_S4(s.Vs, Vs), HVC_S4(s.Vd, Vd), HVC_S4(s.n, n), HVC_S4(s.h, h), 0.0,
                       HVC_S4(s.r, r), HVC_S4(s.c, c), HVC_S4(s.Ca, Ca), E_dt);
#undef HVC_S4
    s.Ge_d = s.Ge_d * E_dt;
    s.Ge_s = s.Ge_s * E_dt;
    s.Gi_d = s.Gi_d * E_dt;
    s.Gi_s = s.Gi_s * E_dt;
#define HVC_F(k) (dt * (k1.k + 3 * k2.k + 3 * k3.k + k4.k) * 0.125)
    s.Vs = s.Vs + HVC_F(Vs);
    s.n = s.n + HVC_F(n);
    s.h = s.h + HVC_F(h);
    s.Vd = s.Vd + HVC_F(Vd);
    s.r = s.r + HVC_F(r);
    s.c = s.c + HVC_F(c);
    s.Ca = s.Ca + HVC_F(Ca);
#undef HVC_F
}

// HH2_test::state_check (HH2_test.cpp:561-594); returns bit0 soma, bit1 dend
__device__ __forceinline__ int ra_state_check(double Vs, double Vd, int& flag_soma, int& flag_dend)
{
    int fired = 0;
    if ((flag_soma == 1) && (Vs < -30.0 - 5.0)) {
        flag_soma = 0;
        fired |= 1;
    } else if ((flag_soma == 0) && (Vs > -30.0)) {
        flag_soma = 1;
    }
    if ((flag_dend == 0) && (Vd >= -30.0)) {
        flag_dend = 1;
        fired |= 2;
    } else if ((flag_dend == 1) && (Vd < -30.0 - 5.0)) {
        flag_dend = 0;
    }
    return fired;
}

// ---------------- HVC-I ----------------
struct IState {
    double V, n, m, h, w, Ge, Gi;
};

constexpr double kI_gL = 0.1, kI_El = -65.0, kI_gNa = 100.0, kI_Ena = 55.0, kI_gKdr = 20.0;
constexpr double kI_Ek = -80.0, kI_gKHT = 500.0, kI_Ei = -75.0, kI_A = 6000.0;

template <bool EXACT>
__device__ __forceinline__ double i_an(double v, ::hvc::DivGuard& bad) { return 0.15 * (v + 15) / (1 - dexp(CD(-(v + 15), 10))); }
template <bool EXACT>
__device__ __forceinline__ double i_bn(double v, ::hvc::DivGuard& bad) { return 0.2 * dexp(CD(-(v + 25), 80)); }
template <bool EXACT>
__device__ __forceinline__ double i_am(double v, ::hvc::DivGuard& bad) { return (v + 22) / (1 - dexp(CD(-(v + 22), 10))); }
template <bool EXACT>
__device__ __forceinline__ double i_bm(double v, ::hvc::DivGuard& bad) { return 40 * dexp(CD(-(v + 47), 18)); }
template <bool EXACT>
__device__ __forceinline__ double i_ah(double v, ::hvc::DivGuard& bad) { return 0.7 * dexp(CD(-(v + 34), 20)); }
template <bool EXACT>
__device__ __forceinline__ double i_bh(double v, ::hvc::DivGuard& bad) { return 10 / (1 + dexp(CD(-(v + 4), 10))); }
template <bool EXACT>
__device__ __forceinline__ double i_wInf(double v, ::hvc::DivGuard& bad) { return 1 / (1 + dexp(CD(-v, 5))); }

struct IK {
    double V, n, m, h, w;
};

// HHI_buffer.cpp:741-747 at stage state (v, n1, m1, h1, w1); e2/e5 decay factors
template <bool EXACT>
__device__ __forceinline__ IK i_rhs(::hvc::DivGuard& bad, const IState& s, double v, double n1, double m1, double h1, double w1,
                                    double e2, double e5)
{
    IK k;
    double n4 = n1 * n1 * n1 * n1;
    double m3 = m1 * m1 * m1;
    k.V = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
          - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * 0.0 / kI_A;
    k.n = i_an<EXACT>(v, bad) * (1 - n1) - i_bn<EXACT>(v, bad) * n1;
    k.m = i_am<EXACT>(v, bad) * (1 - m1) - i_bm<EXACT>(v, bad) * m1;
    k.h = i_ah<EXACT>(v, bad) * (1 - h1) - i_bh<EXACT>(v, bad) * h1;
    k.w = (i_wInf<EXACT>(v, bad) - w1);  // divided by tauW = 1 (exact)
    return k;
}

// HHI_buffer::DP8_step; row = decay factors of this step
template <bool EXACT>
__device__ __forceinline__ void i_dp8(::hvc::DivGuard& bad, IState& s, double dt, const DecayRow& row)
{
    IK k1 = i_rhs<EXACT>(bad, s, s.V, s.n, s.m, s.h, s.w, 1.0, 1.0);
    IK k2 = i_rhs<EXACT>(bad, s, s.V + CD(dt * k1.V, 18), s.n + CD(dt * k1.n, 18), s.m + CD(dt * k1.m, 18), s.h + CD(dt * k1.h, 18),
                  s.w + CD(dt * k1.w, 18), row.e2[kOffDp2], row.e5[kOffDp2]);
    IK k3 = i_rhs<EXACT>(bad, s, s.V + CD(dt * (k1.V + 3 * k2.V), 48), s.n + CD(dt * (k1.n + 3 * k2.n), 48),
                  s.m + CD(dt * (k1.m + 3 * k2.m), 48), s.h + CD(dt * (k1.h + 3 * k2.h), 48),
                  s.w + CD(dt * (k1.w + 3 * k2.w), 48), row.e2[kOffDp3], row.e5[kOffDp3]);
    IK k4 = i_rhs<EXACT>(bad, s, s.V + dt * (k1.V + 3 * k3.V) * 0.03125, s.n + dt * (k1.n + 3 * k3.n) * 0.03125,
                  s.m + dt * (k1.m + 3 * k3.m) * 0.03125, s.h + dt * (k1.h + 3 * k3.h) * 0.03125,
                  s.w + dt * (k1.w + 3 * k3.w) * 0.03125, row.e2[kOffDp4], row.e5[kOffDp4]);
#define C5(x) (20 * k1.x - 75 * k3.x + 75 * k4.x)
    IK k5 = i_rhs<EXACT>(bad, s, s.V + dt * C5(V) * 0.015625, s.n + dt * C5(n) * 0.015625, s.m + dt * C5(m) * 0.015625, s.h + dt * C5(h) * 0.015625,
                  s.w + dt * C5(w) * 0.015625, row.e2[kOffDp5], row.e5[kOffDp5]);
#define C6(x) (3 * k1.x + 15 * k4.x + 12 * k5.x)
    IK k6 = i_rhs<EXACT>(bad, s, s.V + CD(dt * C6(V), 80), s.n + CD(dt * C6(n), 80), s.m + CD(dt * C6(m), 80), s.h + CD(dt * C6(h), 80),
                  s.w + CD(dt * C6(w), 80), row.e2[kOffDp6], row.e5[kOffDp6]);
#define C7(x) (CD(29443841 * k1.x, 614563906) + CD(77736538 * k4.x, 692538347) - CD(28693883 * k5.x, 1125000000) + \
               CD(23124283 * k6.x, 1800000000))
    IK k7 = i_rhs<EXACT>(bad, s, s.V + dt * C7(V), s.n + dt * C7(n), s.m + dt * C7(m), s.h + dt * C7(h), s.w + dt * C7(w),
                  row.e2[kOffDp7], row.e5[kOffDp7]);
#define C8(x) (CD(16016141 * k1.x, 946692911) + CD(61564180 * k4.x, 158732637) + CD(22789713 * k5.x, 633445777) + \
               CD(545815736 * k6.x, 2771057229) - CD(180193667 * k7.x, 1043307555))
    IK k8 = i_rhs<EXACT>(bad, s, s.V + dt * C8(V), s.n + dt * C8(n), s.m + dt * C8(m), s.h + dt * C8(h), s.w + dt * C8(w),
                  row.e2[kOffDp8], row.e5[kOffDp8]);
#define C9(x) (CD(39632708 * k1.x, 573591083) - CD(433636366 * k4.x, 683701615) - CD(421739975 * k5.x, 2616292301) + \
               CD(100302831 * k6.x, 723423059) + CD(790204164 * k7.x, 839813087) + CD(800635310 * k8.x, 3783071287))
    IK k9 = i_rhs<EXACT>(bad, s, s.V + dt * C9(V), s.n + dt * C9(n), s.m + dt * C9(m), s.h + dt * C9(h), s.w + dt * C9(w),
                  row.e2[kOffDp9], row.e5[kOffDp9]);
#define C10(x) (CD(246121993 * k1.x, 1340847787) - CD(37695042795 * k4.x, 15268766246) - CD(309121744 * k5.x, 1061227803) - \
                CD(12992083 * k6.x, 490766935) + CD(6005943493 * k7.x, 2108947869) + CD(393006217 * k8.x, 1396673457) + \
                CD(123872331 * k9.x, 1001029789))
    IK k10 = i_rhs<EXACT>(bad, s, s.V + dt * C10(V), s.n + dt * C10(n), s.m + dt * C10(m), s.h + dt * C10(h),
                   s.w + dt * C10(w), row.e2[kOffDp10], row.e5[kOffDp10]);
#define C11(x) (CD(-1028468189 * k1.x, 846180014) + CD(8478235783 * k4.x, 508512852) + CD(1311729495 * k5.x, 1432422823) - \
                CD(10304129995 * k6.x, 1701304382) - CD(48777925059 * k7.x, 3047939560) + \
                CD(15336726248 * k8.x, 1032824649) - CD(45442868181 * k9.x, 3398467696) + CD(3065993473 * k10.x, 597172653))
    IK k11 = i_rhs<EXACT>(bad, s, s.V + dt * C11(V), s.n + dt * C11(n), s.m + dt * C11(m), s.h + dt * C11(h),
                   s.w + dt * C11(w), row.e2[kOffDp11], row.e5[kOffDp11]);
#define C12(x) (CD(185892177 * k1.x, 718116043) - CD(3185094517 * k4.x, 667107341) - CD(477755414 * k5.x, 1098053517) - \
                CD(703635378 * k6.x, 230739211) + CD(5731566787 * k7.x, 1027545527) + CD(5232866602 * k8.x, 850066563) - \
                CD(4093664535 * k9.x, 808688257) + CD(3962137247 * k10.x, 1805957418) + CD(65686358 * k11.x, 487910083))
    IK k12 = i_rhs<EXACT>(bad, s, s.V + dt * C12(V), s.n + dt * C12(n), s.m + dt * C12(m), s.h + dt * C12(h),
                   s.w + dt * C12(w), row.e2[kOffDt], row.e5[kOffDt]);
#define CF(x) (CD(13451932 * k1.x, 455176623) - CD(808719846 * k6.x, 976000145) + CD(1757004468 * k7.x, 5645159321) + \
               CD(656045339 * k8.x, 265891186) - CD(3867574721 * k9.x, 1518517206) + CD(465885868 * k10.x, 322736535) + \
               CD(53011238 * k11.x, 667516719) + CD(2 * k12.x, 45))
    s.V = s.V + dt * CF(V);
    s.n = s.n + dt * CF(n);
    s.m = s.m + dt * CF(m);
    s.h = s.h + dt * CF(h);
    s.w = s.w + dt * CF(w);
    s.Ge = s.Ge * row.e2[kOffDt];
    s.Gi = s.Gi * row.e5[kOffDt];
#undef C5
#undef C6
#undef C7
#undef C8
#undef C9
#undef C10
#undef C11
#undef C12
#undef CF
}

// HHI_buffer::Runge4_step
template <bool EXACT>
__device__ __forceinline__ void i_runge4(::hvc::DivGuard& bad, IState& s, double dt, const DecayRow& row)
{
    IK k1 = i_rhs<EXACT>(bad, s, s.V, s.n, s.m, s.h, s.w, 1.0, 1.0);
    IK k2 = i_rhs<EXACT>(bad, s, s.V + CD(dt * k1.V, 3), s.n + CD(dt * k1.n, 3), s.m + CD(dt * k1.m, 3), s.h + CD(dt * k1.h, 3),
                  s.w + CD(dt * k1.w, 3), row.e2[kOffThird], row.e5[kOffThird]);
#define S3(x) (CD(-k1.x, 3) + k2.x)
    IK k3 = i_rhs<EXACT>(bad, s, s.V + dt * S3(V), s.n + dt * S3(n), s.m + dt * S3(m), s.h + dt * S3(h), s.w + dt * S3(w),
                  row.e2[kOffTwoThird], row.e5[kOffTwoThird]);
#undef S3
#define S4(x) (k1.x - k2.x + k3.x)
    IK k4 = i_rhs<EXACT>(bad, s, s.V + dt * S4(V), s.n + dt * S4(n), s.m + dt * S4(m), s.h + dt * S4(h), s.w + dt * S4(w),
                  row.e2[kOffDt], row.e5[kOffDt]);
#undef S4
#define F(x) (dt * (k1.x + 3 * k2.x + 3 * k3.x + k4.x) * 0.125)
    s.V = s.V + F(V);
    s.n = s.n + F(n);
    s.m = s.m + F(m);
    s.h = s.h + F(h);
    s.w = s.w + F(w);
#undef F
    s.Ge = s.Ge * row.e2[kOffDt];
    s.Gi = s.Gi * row.e5[kOffDt];
}

// HHI_buffer::state_check (HHI_buffer.cpp:260-276)
__device__ __forceinline__ int i_state_check(double V, int& flag)
{
    if ((flag == 1) && (V < -20.0 - 5.0)) {
        flag = 0;
        return 1;
    }
    if ((flag == 0) && (V > -20.0)) flag = 1;
    return 0;
}

// Poisson_noise::get_spike_time and random (poisson_noise.cpp:9-21) on one stream
struct PoissonStream {
    uint32_t seed_lo, seed_hi, trial, stream, neuron, draw;
    __device__ __forceinline__ double u()
    {
        return (double)poisson_u32(seed_lo, seed_hi, trial, stream, neuron, draw++) / 4294967295.0;
    }
    __device__ __forceinline__ double spike_time(double lambda) { return -dlog(1 - u()) / lambda; }
    __device__ __forceinline__ double random(double G)
    {
        return G * (double)poisson_u32(seed_lo, seed_hi, trial, stream, neuron, draw++) / 4294967295.0;
    }
};

// HHI_buffer::noise_check (HHI_buffer.cpp:278-294); time = neuron time before the step
__device__ __forceinline__ void i_noise_check(PoissonStream& ps, double time, double dt, double& G,
                                              double& noise_time)
{
    if (time > noise_time) {
        G += ps.random(0.45);
        double random = 1000 * ps.spike_time(250.0);
        while (random < dt) {
            random = 1000 * ps.spike_time(250.0);
            G += ps.random(0.45);
        }
        noise_time = noise_time + random;
    }
}

// HHI_buffer::initialize_noise (HHI_buffer.cpp:217-239)
__device__ __forceinline__ double i_init_noise(PoissonStream& ps, double init_dt)
{
    double t = 1000 * ps.spike_time(250.0);
    while (t < init_dt) t = 1000 * ps.spike_time(250.0);
    return t;
}

// HH2_buffer::noise_check for the RK4 form with Poisson noise (HH2_buffer.cpp:592-607)
__device__ __forceinline__ void ra_noise_check(PoissonStream& ps, double time, double dt, double& G,
                                               double G_noise, double& noise_time)
{
    if (time > noise_time) {
        G += ps.random(G_noise);
        double random = 1000.0 * ps.spike_time(100.0);
        while (random < dt) {
            random = 1000.0 * ps.spike_time(100.0);
            G += ps.random(G_noise);
        }
        noise_time = noise_time + random;
    }
}

__constant__ double kPointDistribution[30] = {
    0x1.3988e1409212ep+1, -0x1.3988e1409212ep+1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

}  // namespace hvc

namespace hvc {

// ---------------- HVC-I, five lanes per neuron ----------------
// DP8 spends its time in the gate rates (7 exp per stage) and in the stage
// combinations (one per state variable, up to 9 quotients each).  Lane role
// r = lane % 5 owns variable r of (V, n, m, h, w): it evaluates that variable's
// gate rates and its combination, and the five stage values are exchanged with
// cross-lane shuffles.  Every lane evaluates exactly the expression trees of
// HHI_buffer (HHI_buffer.cpp:426-691, 741-747), so results equal i_dp8 bit for
// bit.  Rate forms: A = c (v+d) / (1 - E), B = c E, C = c / (1 + E), with
// E = exp(-(v+d)/e):  an, am: A;  bn, bm, ah: B;  bh, wInf: C.
struct GateSlot {
    int type;        // 0: A, 1: B, 2: C
    double c, d, e, inv_e;
};

template <bool EXACT>
__device__ __forceinline__ double gate_rate(::hvc::DivGuard& bad, const GateSlot& g, double v)
{
    double vd = v + g.d;
    double E = dexp(EXACT ? -vd / g.e : cdiv_fast(-vd, g.e, g.inv_e, bad));
    double num = (g.type == 0) ? g.c * vd : g.c;
    double den = (g.type == 0) ? (1 - E) : (1 + E);
    double q = num / den;
    return (g.type == 1) ? g.c * E : q;
}

__constant__ GateSlot kIGates[5][2] = {
    {{1, 0.0, 0.0, 10.0, 0.1}, {1, 0.0, 0.0, 10.0, 0.1}},                    // V: no rates (unused)
    {{0, 0.15, 15.0, 10.0, 0.1}, {1, 0.2, 25.0, 80.0, 1.0 / 80.0}},          // n: an, bn
    {{0, 1.0, 22.0, 10.0, 0.1}, {1, 40.0, 47.0, 18.0, 1.0 / 18.0}},          // m: am, bm
    {{1, 0.7, 34.0, 20.0, 1.0 / 20.0}, {2, 10.0, 4.0, 10.0, 0.1}},           // h: ah, bh
    {{2, 1.0, 0.0, 5.0, 0.2}, {1, 0.0, 0.0, 10.0, 0.1}},                     // w: wInf (unused)
};

// derivative of the lane's variable at stage state (v, n1, m1, h1, w1)
template <bool EXACT>
__device__ __forceinline__ double i_lane_rhs(::hvc::DivGuard& bad, const IState& s, int role, double v, double n1, double m1, double h1,
                                             double w1, double e2, double e5)
{
    double n4 = n1 * n1 * n1 * n1;
    double m3 = m1 * m1 * m1;
    double kV = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
                - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * 0.0 / kI_A;
    double a = gate_rate<EXACT>(bad, kIGates[role][0], v);
    double b = gate_rate<EXACT>(bad, kIGates[role][1], v);
    double x = role == 1 ? n1 : (role == 2 ? m1 : (role == 3 ? h1 : w1));
    double kg = a * (1 - x) - b * x;
    double kw = (a - x);
    return role == 0 ? kV : (role == 4 ? kw : kg);
}

#define HVC_BCAST5(x)                       \
    v = __shfl(x, base + 0);                \
    n1 = __shfl(x, base + 1);               \
    m1 = __shfl(x, base + 2);               \
    h1 = __shfl(x, base + 3);               \
    w1 = __shfl(x, base + 4);

// HHI_buffer::DP8_step with five lanes per neuron; s is replicated in the group
template <bool EXACT>
__device__ __forceinline__ void i_dp8_lanes(::hvc::DivGuard& bad, IState& s, int role, int base, double dt, const DecayRow& row)
{
    const double x0 = role == 0 ? s.V : (role == 1 ? s.n : (role == 2 ? s.m : (role == 3 ? s.h : s.w)));
    double v = s.V, n1 = s.n, m1 = s.m, h1 = s.h, w1 = s.w;
    double x;
    double k1 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, 1.0, 1.0);
    x = x0 + CD(dt * k1, 18);
    HVC_BCAST5(x)
    double k2 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp2], row.e5[kOffDp2]);
    x = x0 + CD(dt * (k1 + 3 * k2), 48);
    HVC_BCAST5(x)
    double k3 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp3], row.e5[kOffDp3]);
    x = x0 + dt * (k1 + 3 * k3) * 0.03125;
    HVC_BCAST5(x)
    double k4 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp4], row.e5[kOffDp4]);
    x = x0 + dt * (20 * k1 - 75 * k3 + 75 * k4) * 0.015625;
    HVC_BCAST5(x)
    double k5 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp5], row.e5[kOffDp5]);
    x = x0 + CD(dt * (3 * k1 + 15 * k4 + 12 * k5), 80);
    HVC_BCAST5(x)
    double k6 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp6], row.e5[kOffDp6]);
    x = x0 + dt * (CD(29443841 * k1, 614563906) + CD(77736538 * k4, 692538347) - CD(28693883 * k5, 1125000000) +
                   CD(23124283 * k6, 1800000000));
    HVC_BCAST5(x)
    double k7 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp7], row.e5[kOffDp7]);
    x = x0 + dt * (CD(16016141 * k1, 946692911) + CD(61564180 * k4, 158732637) + CD(22789713 * k5, 633445777) +
                   CD(545815736 * k6, 2771057229) - CD(180193667 * k7, 1043307555));
    HVC_BCAST5(x)
    double k8 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp8], row.e5[kOffDp8]);
    x = x0 + dt * (CD(39632708 * k1, 573591083) - CD(433636366 * k4, 683701615) - CD(421739975 * k5, 2616292301) +
                   CD(100302831 * k6, 723423059) + CD(790204164 * k7, 839813087) + CD(800635310 * k8, 3783071287));
    HVC_BCAST5(x)
    double k9 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp9], row.e5[kOffDp9]);
    x = x0 + dt * (CD(246121993 * k1, 1340847787) - CD(37695042795 * k4, 15268766246) -
                   CD(309121744 * k5, 1061227803) - CD(12992083 * k6, 490766935) + CD(6005943493 * k7, 2108947869) +
                   CD(393006217 * k8, 1396673457) + CD(123872331 * k9, 1001029789));
    HVC_BCAST5(x)
    double k10 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp10], row.e5[kOffDp10]);
    x = x0 + dt * (CD(-1028468189 * k1, 846180014) + CD(8478235783 * k4, 508512852) + CD(1311729495 * k5, 1432422823) -
                   CD(10304129995 * k6, 1701304382) - CD(48777925059 * k7, 3047939560) +
                   CD(15336726248 * k8, 1032824649) - CD(45442868181 * k9, 3398467696) +
                   CD(3065993473 * k10, 597172653));
    HVC_BCAST5(x)
    double k11 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp11], row.e5[kOffDp11]);
    x = x0 + dt * (CD(185892177 * k1, 718116043) - CD(3185094517 * k4, 667107341) - CD(477755414 * k5, 1098053517) -
                   CD(703635378 * k6, 230739211) + CD(5731566787 * k7, 1027545527) + CD(5232866602 * k8, 850066563) -
                   CD(4093664535 * k9, 808688257) + CD(3962137247 * k10, 1805957418) + CD(65686358 * k11, 487910083));
    HVC_BCAST5(x)
    double k12 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDt], row.e5[kOffDt]);
    x = x0 + dt * (CD(13451932 * k1, 455176623) - CD(808719846 * k6, 976000145) + CD(1757004468 * k7, 5645159321) +
                   CD(656045339 * k8, 265891186) - CD(3867574721 * k9, 1518517206) + CD(465885868 * k10, 322736535) +
                   CD(53011238 * k11, 667516719) + CD(2 * k12, 45));
    HVC_BCAST5(x)
    s.V = v; s.n = n1; s.m = m1; s.h = h1; s.w = w1;
    s.Ge = s.Ge * row.e2[kOffDt];
    s.Gi = s.Gi * row.e5[kOffDt];
}
#undef HVC_BCAST5

}  // namespace hvc
