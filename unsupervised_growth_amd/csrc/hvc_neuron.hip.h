// hvc_neuron.hip.h -- neuron right-hand sides and steppers for gfx950, in the
// arithmetic of one precision form.
//
// HVC-RA, model 0: HH2_test::Debraband_step (c++/src/HH2_test.cpp:613-754) with
//                  the RHS of HH2_test.cpp:866-890 and gates of HH2_test.h:265-276.
// HVC-RA, model 1: HH2_buffer::Runge4_step (c++/src/HH2_buffer.cpp:741-843),
//                  RHS HH2_buffer.cpp:846-874 (no K-LT current, tauC = 10).
// HVC-I,  model 0: HHI_buffer::DP8_step (c++/src/HHI_buffer.cpp:426-691).
// HVC-I,  model 1: HHI_buffer::Runge4_step (c++/src/HHI_buffer.cpp:341-424).
//
// State lives in registers for a whole coupling window.  The conductance decay
// factors exp(-((time + c dt) - time) / tau) depend only on the shared clock, so
// they come from a per-step table (DecayRow) computed once per trial length.
//
// A translation unit includes this header once, in one precision form:
//   FP64 (default): real = double.  Each expression keeps the reference's
//     operand order (the build disables FMA contraction), literal divisions stay
//     divisions (correctly rounded, see cdiv_fast) and exp is the engine's own,
//     so results equal the CPU evaluation of the same trees bit for bit.
//     Names live in hvc::p64.
//   FP32 (HVC_FP32 defined): real = float for membrane, gating and conductance
//     state; exp is the hardware exp2 with the ln 2 scaling, divisions by a
//     constant multiply by the rounded reciprocal and other divisions use the
//     hardware reciprocal.  Noise event times and the clock stay FP64.
//     Names live in hvc::p32.
// HR(x) rounds a literal to the form's type, so an FP64 expression is the
// literal one and an FP32 expression never widens to FP64.
#pragma once
#include "hvc_math.hip.h"

namespace hvc {

#ifdef HVC_FP32
#define HVC_PREC_BEGIN inline namespace p32 {
HVC_PREC_BEGIN
typedef float real;
// no quotient needs a repeat in this form
struct Guard {
    static constexpr bool kActive = false;
    __device__ __forceinline__ bool failed() const { return false; }
    __device__ __forceinline__ void note_k(real) const {}
};
#define HR(x) ((real)(x))
#define CD(a, B) ((a) * (real)(1.0 / (double)(B)))
#define CDV(a, B, invB) ((a) * (invB))
#define CK(n, k, d) ((k) * (real)((double)(n) / (double)(d)))
#define CKN(n, k, d) CK(n, k, d)
#define DIVR(a, b) ((real)(a) * __builtin_amdgcn_rcpf((real)(b)))
// exp(a / B) as 2^(a * RN(log2(e) / B)): the argument a / B is already only
// known to |a / B| * 2^-24 in this form, and folding log2(e) into the constant
// keeps the error of that order with one multiplication and the hardware exp2
#define EXPCD(a, B) __builtin_amdgcn_exp2f((a) * (real)(1.4426950408889634 / (double)(B)))
#define EXPCDV(a, B, invB, inv2B) __builtin_amdgcn_exp2f((a) * (inv2B))
// the neuron arithmetic of this form may fuse a * b + c into one rounding
#define HVC_FP_FAST _Pragma("clang fp contract(fast)")
// keeps a value computed in every lane (the form's quotients are one
// reciprocal: cheaper than a divergent branch around them)
#define HVC_KEEP(x) asm volatile("" : "+v"(x))
// state update with compensated (Kahan) summation: st.f + st.f_lo carries the
// value, so increments far below half an ulp of st.f still accumulate
#define HVC_ACC(st, f, inc)                  \
    {                                        \
        real y_ = (inc) + st.f##_lo;         \
        real t_ = st.f + y_;                 \
        st.f##_lo = y_ - (t_ - st.f);        \
        st.f = t_;                           \
    }
#else
#define HVC_PREC_BEGIN inline namespace p64 {
HVC_PREC_BEGIN
typedef double real;
struct Guard : DivGuard {
    static constexpr bool kActive = true;
};
#define HR(x) ((real)(x))
// constant-divisor quotient inside a stepper: IEEE division when EXACT, else the
// branch-free Markstein form, whose guard `bad` tells when the step must be repeated
#define CD(a, B) (EXACT ? (a) / (double)(B) : ::hvc::cdiv_fast((a), (double)(B), 1.0 / (double)(B), bad))
#define CDV(a, B, invB) (EXACT ? (a) / (B) : ::hvc::cdiv_fast((a), (B), (invB), bad))
// tableau term (n k) / d, evaluated as the reference writes it
#define CK(n, k, d) CD((n) * (k), d)
// the same term for a k the step notes itself (DivGuard::note_k)
#define CKN(n, k, d) (EXACT ? ((n) * (k)) / (double)(d) : ::hvc::cdiv_plain((n) * (k), (double)(d), 1.0 / (double)(d)))
#define DIVR(a, b) ((a) / (b))
#define EXPR(x) ::hvc::dexp(x)
#define EXPCD(a, B) EXPR(CD(a, B))
#define EXPCDV(a, B, invB, inv2B) EXPR(CDV(a, B, invB))
#define HVC_FP_FAST
#define HVC_KEEP(x)
#define HVC_ACC(st, f, inc) st.f = st.f + inc;
#endif

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
// the row as the FP32 integration reads it from LDS: the stage factors
// rounded to float once, when the window is staged (they enter float
// expressions anyway), and the whole-step factors kept in FP64 for the
// conductance decays, which take one rounding to float per step
struct DecayRowF {
    float e5[kNumOffsets];
    float e2[kNumOffsets];
    double e5dt, e2dt;
};
// the whole-step decay factors of a row (FP64 in both row types)
__device__ __forceinline__ double row_e5dt(const DecayRow& r) { return r.e5[kOffDt]; }
__device__ __forceinline__ double row_e2dt(const DecayRow& r) { return r.e2[kOffDt]; }
__device__ __forceinline__ double row_e5dt(const DecayRowF& r) { return r.e5dt; }
__device__ __forceinline__ double row_e2dt(const DecayRowF& r) { return r.e2dt; }
#ifdef HVC_FP32
typedef DecayRowF WinRow;
#else
typedef DecayRow WinRow;
#endif

// ---------------- HVC-RA ----------------
struct RAState {
    real Vs, Vd, n, l, h, r, c, Ca;
    real Ge_s, Gi_s, Ge_d, Gi_d;
#ifdef HVC_FP32
    real Vs_lo, Vd_lo, n_lo, l_lo, h_lo, r_lo, c_lo, Ca_lo;
#endif
};

// per-network constants derived once on the host with the reference's own expressions
struct RAUniform {
    real sq_sig_s;   // sqrt(_timestep) * _sigma_soma
    real sq_sig_d;   // sqrt(_timestep) * _sigma_dend
    real mu_term_s;  // 100000 * (IsExt(t) + _mu_soma) / As
    real mu_term_d;  // 100000 * (IdExt(t) + _mu_dend) / Ad
    real dt;
};

constexpr real kRA_As = HR(5000.0), kRA_Ad = HR(10000.0), kRA_Rc = HR(55.0), kRA_cm = HR(1.0);
constexpr real kRA_GsKLT = HR(3.5), kRA_GsK = HR(8.0), kRA_GsNa = HR(60.0), kRA_GdCaK = HR(150.0);
constexpr real kRA_GsL = HR(0.1), kRA_GdL = HR(0.1), kRA_EdCa = HR(120.0), kRA_EdK = -HR(90.0);
constexpr real kRA_EsNa = HR(55.0), kRA_EsK = -HR(90.0), kRA_Egaba = -HR(80.0);

#ifdef HVC_FP32
// exp(sg (v + d) / B) as 2^(k1 v + k0): v + d folded into one multiply-add
#define HVC_EXPLIN(v, sg, d, B)                                                                          \
    __builtin_amdgcn_exp2f(__builtin_fmaf((v), (real)((sg) * 1.4426950408889634 / (double)(B)),         \
                                          (real)((sg) * (double)(d) * 1.4426950408889634 / (double)(B))))
#define HVC_SIGM(v, sg, d, B) __builtin_amdgcn_rcpf(1 + HVC_EXPLIN(v, sg, d, B))   // 1 / (1 + exp(sg (v + d) / B))
template <bool EXACT>
__device__ __forceinline__ real nInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, -1, 35, 10); }
template <bool EXACT>
__device__ __forceinline__ real tauN(real v, Guard& bad) { HVC_FP_FAST return __builtin_fmaf(HR(0.5), HVC_SIGM(v, 1, 27, 15), HR(0.1)); }
template <bool EXACT>
__device__ __forceinline__ real hInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, 1, 45, 7); }
template <bool EXACT>
__device__ __forceinline__ real tauH(real v, Guard& bad) { HVC_FP_FAST return __builtin_fmaf(HR(0.75), HVC_SIGM(v, 1, 40.5, 6), HR(0.1)); }
template <bool EXACT>
__device__ __forceinline__ real mInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, -1, 30, 9.5); }
template <bool EXACT>
__device__ __forceinline__ real lInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, -1, 40, 5); }
template <bool EXACT>
__device__ __forceinline__ real rInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, -1, 5, 10); }
template <bool EXACT>
__device__ __forceinline__ real cInf(real v, Guard& bad) { HVC_FP_FAST return HVC_SIGM(v, -1, -10, 7); }
// the denominators 1 + exp(...) of nInf, tauN, hInf, tauH, for the one-thread
// steppers: (x_inf - x) / tau then takes one reciprocal per gate instead of three
__device__ __forceinline__ real nInfDen(real v) { HVC_FP_FAST return 1 + HVC_EXPLIN(v, -1, 35, 10); }
__device__ __forceinline__ real tauNDen(real v) { HVC_FP_FAST return 1 + HVC_EXPLIN(v, 1, 27, 15); }
__device__ __forceinline__ real hInfDen(real v) { HVC_FP_FAST return 1 + HVC_EXPLIN(v, 1, 45, 7); }
__device__ __forceinline__ real tauHDen(real v) { HVC_FP_FAST return 1 + HVC_EXPLIN(v, 1, 40.5, 6); }
#undef HVC_SIGM
#undef HVC_EXPLIN
#else
template <bool EXACT>
__device__ __forceinline__ real nInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-(v + 35), 10)); }
template <bool EXACT>
__device__ __forceinline__ real tauN(real v, Guard& bad) { HVC_FP_FAST return HR(0.1) + DIVR(HR(0.5), 1 + EXPCD(v + 27, 15)); }
template <bool EXACT>
__device__ __forceinline__ real hInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(v + 45, 7)); }
template <bool EXACT>
__device__ __forceinline__ real tauH(real v, Guard& bad) { HVC_FP_FAST return HR(0.1) + DIVR(HR(0.75), 1 + EXPCD(v + HR(40.5), 6)); }
template <bool EXACT>
__device__ __forceinline__ real mInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-(v + 30), HR(9.5))); }
template <bool EXACT>
__device__ __forceinline__ real lInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-(v + 40), 5)); }
template <bool EXACT>
__device__ __forceinline__ real rInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-(v + 5), 10)); }
template <bool EXACT>
__device__ __forceinline__ real cInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-(v - 10), 7)); }
#endif

struct RAK {
    real Vs, Vd, n, l, h, r, c, Ca;
};

// gate values at one stage
struct RAGates {
    real m, nI, tN, hI, tH, lI, rI, cI;
};

// RAW (FP32 form, one thread per neuron): nI, tN, hI, tH hold the
// denominators of their sigmoids (nInfDen ...), which ra_rhs_g<..., true> expects
template <int MODEL, bool EXACT, bool RAW = false>
__device__ __forceinline__ RAGates ra_gates(Guard& bad, real vs, real vd)
{
    RAGates g;
    g.m = mInf<EXACT>(vs, bad);
#ifdef HVC_FP32
    if constexpr (RAW) {
        g.nI = nInfDen(vs);
        g.tN = tauNDen(vs);
        g.hI = hInfDen(vs);
        g.tH = tauHDen(vs);
    } else
#endif
    {
        g.nI = nInf<EXACT>(vs, bad);
        g.tN = tauN<EXACT>(vs, bad);
        g.hI = hInf<EXACT>(vs, bad);
        g.tH = tauH<EXACT>(vs, bad);
    }
    g.lI = (MODEL == 0) ? lInf<EXACT>(vs, bad) : HR(0.0);
    g.rI = rInf<EXACT>(vd, bad);
    g.cI = cInf<EXACT>(vd, bad);
    return g;
}

// all eight derivatives at one stage from its gate values; E = decay factor of the stage time
template <int MODEL, bool EXACT, bool RAW = false>
__device__ __forceinline__ RAK ra_rhs_g(Guard& bad, const RAState& s, real EL, real GCa, const RAUniform& u,
                                        const RAGates& g, real vs, real vd, real n, real h, real l, real r,
                                        real c, real ca, real E)
{
    HVC_FP_FAST
    RAK k;
    real m = g.m;
#ifdef HVC_FP32
    // one chain of multiply-adds per compartment (the soma's started from minus
    // its mean drive): the potassium conductances share (vs - E_K), the coupling terms
    // use RN(100000 / R) as one factor, c / (1 + 6 / [Ca]) is c [Ca] / ([Ca] + 6),
    // and the calcium current is shared by the dendrite and the calcium equation
    const real n2 = n * n;
    real gk = kRA_GsK * (n2 * n2);
    if (MODEL == 0) gk = __builtin_fmaf(kRA_GsKLT, l, gk);
    real t = __builtin_fmaf(kRA_GsL, vs - EL, -u.mu_term_s);
    t = __builtin_fmaf(kRA_GsNa * (m * m * m) * h, vs - kRA_EsNa, t);
    t = __builtin_fmaf(gk, vs - kRA_EsK, t);
    t = __builtin_fmaf(s.Gi_s * E, vs - kRA_Egaba, t);
    t = __builtin_fmaf(s.Ge_s * E, vs, t);
    k.Vs = __builtin_fmaf((real)(100000.0 / 275000.0), vd - vs, -t);   // divided by cm = 1.0
    const real ica = GCa * (r * r) * (vd - kRA_EdCa);
    real d = __builtin_fmaf(kRA_GdL, vd - EL, ica);
    d = __builtin_fmaf(kRA_GdCaK * (vd - kRA_EdK), c * ca * __builtin_amdgcn_rcpf(ca + 6), d);
    d = __builtin_fmaf(s.Ge_d * E, vd, d);
    d = __builtin_fmaf(s.Gi_d * E, vd - kRA_Egaba, d);
    k.Vd = __builtin_fmaf((real)(100000.0 / 550000.0), vs - vd, u.mu_term_d - d);
    if constexpr (RAW) {
        // (1/D - x) / (a + c/T) = (1 - x D) T / (D (a T + c)), D, T the sigmoids' denominators
        k.n = (__builtin_fmaf(-n, g.nI, HR(1.0)) * g.tN) *
              __builtin_amdgcn_rcpf(g.nI * __builtin_fmaf(HR(0.1), g.tN, HR(0.5)));
        k.h = (__builtin_fmaf(-h, g.hI, HR(1.0)) * g.tH) *
              __builtin_amdgcn_rcpf(g.hI * __builtin_fmaf(HR(0.1), g.tH, HR(0.75)));
    } else {
        k.n = DIVR(g.nI - n, g.tN);
        k.h = DIVR(g.hI - h, g.tH);
    }
    k.l = (MODEL == 0) ? CD(g.lI - l, 10) : HR(0.0);
    k.r = (g.rI - r);  // divided by tauR = 1 (exact)
    k.c = (MODEL == 0) ? CD(g.cI - c, 15) : CD(g.cI - c, 10);
    k.Ca = __builtin_fmaf(-HR(0.1), ica, -HR(0.02) * ca);
#else
    real m3 = m * m * m;
    real n4 = n * n * n * n;
    real acc = -kRA_GsL * (vs - EL) - kRA_GsNa * m3 * h * (vs - kRA_EsNa) - kRA_GsK * n4 * (vs - kRA_EsK);
    if (MODEL == 0) acc = acc - kRA_GsKLT * l * (vs - kRA_EsK);
    acc = acc - (s.Gi_s * E) * (vs - kRA_Egaba) - (s.Ge_s * E) * vs + u.mu_term_s
          + CD(100000 * (vd - vs), 275000);
    k.Vs = acc;  // divided by cm = 1.0 (exact)
    real r2 = r * r;
    k.Vd = -kRA_GdL * (vd - EL) - GCa * r2 * (vd - kRA_EdCa) - DIVR(kRA_GdCaK * (vd - kRA_EdK) * c, 1 + DIVR(6, ca))
           - (s.Ge_d * E) * vd - (s.Gi_d * E) * (vd - kRA_Egaba) + u.mu_term_d
           + CD(100000 * (vs - vd), 550000);
    k.n = DIVR(g.nI - n, g.tN);
    k.l = (MODEL == 0) ? CD(g.lI - l, 10) : HR(0.0);
    k.h = DIVR(g.hI - h, g.tH);
    k.r = (g.rI - r);  // divided by tauR = 1 (exact)
    k.c = (MODEL == 0) ? CD(g.cI - c, 15) : CD(g.cI - c, 10);
    k.Ca = (-HR(0.1) * GCa * r * r * (vd - kRA_EdCa) - HR(0.02) * ca);
#endif
    return k;
}

template <int MODEL, bool EXACT>
__device__ __forceinline__ RAK ra_rhs(Guard& bad, const RAState& s, real EL, real GCa, const RAUniform& u,
                                      real vs, real vd, real n, real h, real l, real r,
                                      real c, real ca, real E)
{
#ifdef HVC_FP32
    return ra_rhs_g<MODEL, EXACT, true>(bad, s, EL, GCa, u, ra_gates<MODEL, EXACT, true>(bad, vs, vd), vs, vd, n, h,
                                        l, r, c, ca, E);
#else
    return ra_rhs_g<MODEL, EXACT>(bad, s, EL, GCa, u, ra_gates<MODEL, EXACT>(bad, vs, vd), vs, vd, n, h, l, r, c,
                                  ca, E);
#endif
}

// ---- HVC-RA, RK4 form, four lanes per neuron ----
// The seven gates of HH2_buffer share one form, a + c / (1 + exp(+-(v + d) / e))
// on Vs or Vd.  Lane q of a neuron's quad evaluates gate slots 2q and 2q + 1,
// DPP quad broadcasts hand every lane all seven, and every lane then evaluates
// the derivatives and stage values itself: the state stays replicated across
// the quad and each value is the same expression tree as in ra_rhs.
struct RAGateSlot {
    real a, c, d, e, inv_e;   // inv_e = the form's rounding of 1 / e (as CD uses it)
    int neg, dend;            // argument -(v + d) / (v + d); v = Vd / Vs
    real inv2_e;              // RN(log2(e) / e) (FP32 form, EXPCDV)
};

__constant__ RAGateSlot kRAGates[8] = {
#define HVC_L2(e) (real)(1.4426950408889634 / (e))
    {HR(0.0), HR(1.0), HR(30.0), HR(9.5), (real)(1.0 / 9.5), 1, 0, HVC_L2(9.5)},     // mInf
    {HR(0.0), HR(1.0), HR(35.0), HR(10.0), (real)(1.0 / 10.0), 1, 0, HVC_L2(10.0)},  // nInf
    {HR(0.1), HR(0.5), HR(27.0), HR(15.0), (real)(1.0 / 15.0), 0, 0, HVC_L2(15.0)},  // tauN
    {HR(0.0), HR(1.0), HR(45.0), HR(7.0), (real)(1.0 / 7.0), 0, 0, HVC_L2(7.0)},     // hInf
    {HR(0.1), HR(0.75), HR(40.5), HR(6.0), (real)(1.0 / 6.0), 0, 0, HVC_L2(6.0)},    // tauH
    {HR(0.0), HR(1.0), HR(5.0), HR(10.0), (real)(1.0 / 10.0), 1, 1, HVC_L2(10.0)},   // rInf
    {HR(0.0), HR(1.0), -HR(10.0), HR(7.0), (real)(1.0 / 7.0), 1, 1, HVC_L2(7.0)},    // cInf
    {HR(0.0), HR(1.0), HR(0.0), HR(10.0), (real)(1.0 / 10.0), 1, 0},    // unused
};

template <bool EXACT>
__device__ __forceinline__ real ra_gate(Guard& bad, const RAGateSlot& g, real vs, real vd)
{
    HVC_FP_FAST
    real x = (g.dend ? vd : vs) + g.d;
    real q = DIVR(g.c, 1 + EXPCDV(g.neg ? -x : x, g.e, g.inv_e, g.inv2_e));
    return g.a + q;   // a = 0: 0 + q = q (q > 0)
}

// value of lane K of the quad, in every lane of the quad (DPP quad_perm)
template <int K>
__device__ __forceinline__ float quad_bcast(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), K * 0x55, 0xF, 0xF, true));
}
template <int K>
__device__ __forceinline__ double quad_bcast(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, K * 0x55, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), K * 0x55, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <bool EXACT, bool LANES>
__device__ __forceinline__ RAK ra_rhs4(Guard& bad, const RAState& s, real EL, real GCa, const RAUniform& u,
                                       const RAGateSlot* sl, real vs, real vd, real n, real h, real r, real c,
                                       real ca, real E)
{
    RAGates g;
    if constexpr (LANES) {
        const real g0 = ra_gate<EXACT>(bad, sl[0], vs, vd);
        const real g1 = ra_gate<EXACT>(bad, sl[1], vs, vd);
        g.m = quad_bcast<0>(g0);
        g.nI = quad_bcast<0>(g1);
        g.tN = quad_bcast<1>(g0);
        g.hI = quad_bcast<1>(g1);
        g.tH = quad_bcast<2>(g0);
        g.rI = quad_bcast<2>(g1);
        g.cI = quad_bcast<3>(g0);
        g.lI = HR(0.0);
    } else {
#ifdef HVC_FP32
        return ra_rhs_g<1, EXACT, true>(bad, s, EL, GCa, u, ra_gates<1, EXACT, true>(bad, vs, vd), vs, vd, n, h,
                                        HR(0.0), r, c, ca, E);
#else
        g = ra_gates<1, EXACT>(bad, vs, vd);
#endif
    }
    return ra_rhs_g<1, EXACT>(bad, s, EL, GCa, u, g, vs, vd, n, h, HR(0.0), r, c, ca, E);
}

// HH2_test::Debraband_step; J* are the point-distribution values of this step
template <bool EXACT>
__device__ __forceinline__ void ra_debraband(Guard& bad, RAState& s, real EL, real GCa, const RAUniform& u,
                                             real J1s, real J2s, real J1d, real J2d,
                                             real E_dt, real E_half)
{
    HVC_FP_FAST
    const real dt = u.dt;
    real H1Vs = s.Vs + CD(u.sq_sig_s * (-HR(0.01844540496323970) * J1s - HR(0.1866426386543421) * J2s), 5000);
    real H1Vd = s.Vd + CD(u.sq_sig_d * (-HR(0.01844540496323970) * J1d - HR(0.1866426386543421) * J2d), 10000);
    RAK a1 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H1Vs, H1Vd, s.n, s.h, s.l, s.r, s.c, s.Ca, HR(1.0));

    real H2Vs = s.Vs + dt * a1.Vs + CD(u.sq_sig_s * (HR(0.8017012756521233) * J1s - HR(0.8575745885712401) * J2s), 5000);
    real H2Vd = s.Vd + dt * a1.Vd + CD(u.sq_sig_d * (HR(0.8017012756521233) * J1d - HR(0.8575745885712401) * J2d), 10000);
    RAK a2 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H2Vs, H2Vd, s.n + dt * a1.n, s.h + dt * a1.h, s.l + dt * a1.l,
                       s.r + dt * a1.r, s.c + dt * a1.c, s.Ca + dt * a1.Ca, E_dt);

    real H3Vs = s.Vs + dt * (3 * a1.Vs + a2.Vs) * HR(0.125)
                  + CD(u.sq_sig_s * (HR(0.5092227024816198) * J1s - HR(0.4723392695015512) * J2s), 5000);
    real H3Vd = s.Vd + dt * (3 * a1.Vd + a2.Vd) * HR(0.125)
                  + CD(u.sq_sig_d * (HR(0.5092227024816198) * J1d - HR(0.4723392695015512) * J2d), 10000);
    RAK a3 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H3Vs, H3Vd, s.n + dt * (3 * a1.n + a2.n) * HR(0.125),
                       s.h + dt * (3 * a1.h + a2.h) * HR(0.125), s.l + dt * (3 * a1.l + a2.l) * HR(0.125),
                       s.r + dt * (3 * a1.r + a2.r) * HR(0.125), s.c + dt * (3 * a1.c + a2.c) * HR(0.125),
                       s.Ca + dt * (3 * a1.Ca + a2.Ca) * HR(0.125), E_half);

#define HVC_H4(x, k) (x + dt * (-HR(0.4526683126055039) * a1.k - HR(0.4842227708685013) * a2.k + HR(1.9368910834740051) * a3.k))
    real H4Vs = HVC_H4(s.Vs, Vs) + CD(u.sq_sig_s * (HR(0.9758794209767762) * J1s + HR(0.3060354860326548) * J2s), 5000);
    real H4Vd = HVC_H4(s.Vd, Vd) + CD(u.sq_sig_d * (HR(0.9758794209767762) * J1d + HR(0.3060354860326548) * J2d), 10000);
    RAK a4 = ra_rhs<0, EXACT>(bad, s, EL, GCa, u, H4Vs, H4Vd, HVC_H4(s.n, n), HVC_H4(s.h, h), HVC_H4(s.l, l),
                       HVC_H4(s.r, r), HVC_H4(s.c, c), HVC_H4(s.Ca, Ca), E_dt);
#undef HVC_H4

#define HVC_FIN(k) (dt * (CD(a1.k, 6) - HR(0.005430430675258792) * a2.k + CD(2 * a3.k, 3) + HR(0.1720970973419255) * a4.k))
    HVC_ACC(s, Vs, HVC_FIN(Vs) + CD(u.sq_sig_s * J1s * 100000, 5000))
    HVC_ACC(s, Vd, HVC_FIN(Vd) + CD(u.sq_sig_d * J1d * 100000, 10000))
    HVC_ACC(s, Ca, HVC_FIN(Ca))
    HVC_ACC(s, n, HVC_FIN(n))
    HVC_ACC(s, l, HVC_FIN(l))
    HVC_ACC(s, h, HVC_FIN(h))
    HVC_ACC(s, r, HVC_FIN(r))
    HVC_ACC(s, c, HVC_FIN(c))
#undef HVC_FIN
    s.Ge_d = s.Ge_d * E_dt;
    s.Ge_s = s.Ge_s * E_dt;
    s.Gi_d = s.Gi_d * E_dt;
    s.Gi_s = s.Gi_s * E_dt;
}

// HH2_buffer::Runge4_step (Kutta 3/8 rule); LANES: four lanes per neuron, sl = this lane's two gate slots
template <bool EXACT, bool LANES>
__device__ __forceinline__ void ra_runge4(Guard& bad, RAState& s, real EL, real GCa, const RAUniform& u,
                                          const RAGateSlot* sl, real E_third, real E_twothird, real E_dt)
{
    HVC_FP_FAST
    const real dt = u.dt;
#ifdef HVC_FP32
    // dt / 3 and dt / 8 once per step; k (dt / 8) is (dt k) / 8 exactly
    const real dt3 = dt * (real)(1.0 / 3.0), dt8 = dt * HR(0.125);
#define HVC_S2(x, k) __builtin_fmaf(k1.k, dt3, x)
#define HVC_F(k) ((k1.k + 3 * k2.k + 3 * k3.k + k4.k) * dt8)
#else
#define HVC_S2(x, k) (x + CD(dt * k1.k, 3))
#define HVC_F(k) (dt * (k1.k + 3 * k2.k + 3 * k3.k + k4.k) * HR(0.125))
#endif
    RAK k1 = ra_rhs4<EXACT, LANES>(bad, s, EL, GCa, u, sl, s.Vs, s.Vd, s.n, s.h, s.r, s.c, s.Ca, HR(1.0));
    RAK k2 = ra_rhs4<EXACT, LANES>(bad, s, EL, GCa, u, sl, HVC_S2(s.Vs, Vs), HVC_S2(s.Vd, Vd), HVC_S2(s.n, n),
                                   HVC_S2(s.h, h), HVC_S2(s.r, r), HVC_S2(s.c, c), HVC_S2(s.Ca, Ca), E_third);
#undef HVC_S2
#define HVC_S3(x, k) (x + dt * (CD(-k1.k, 3) + k2.k))
    RAK k3 = ra_rhs4<EXACT, LANES>(bad, s, EL, GCa, u, sl, HVC_S3(s.Vs, Vs), HVC_S3(s.Vd, Vd), HVC_S3(s.n, n),
                                   HVC_S3(s.h, h), HVC_S3(s.r, r), HVC_S3(s.c, c), HVC_S3(s.Ca, Ca), E_twothird);
#undef HVC_S3
#define HVC_S4(x, k) (x + dt * (k1.k - k2.k + k3.k))
    RAK k4 = ra_rhs4<EXACT, LANES>(bad, s, EL, GCa, u, sl, HVC_S4(s.Vs, Vs), HVC_S4(s.Vd, Vd), HVC_S4(s.n, n),
                                   HVC_S4(s.h, h), HVC_S4(s.r, r), HVC_S4(s.c, c), HVC_S4(s.Ca, Ca), E_dt);
#undef HVC_S4
    s.Ge_d = s.Ge_d * E_dt;
    s.Ge_s = s.Ge_s * E_dt;
    s.Gi_d = s.Gi_d * E_dt;
    s.Gi_s = s.Gi_s * E_dt;
    HVC_ACC(s, Vs, HVC_F(Vs))
    HVC_ACC(s, n, HVC_F(n))
    HVC_ACC(s, h, HVC_F(h))
    HVC_ACC(s, Vd, HVC_F(Vd))
    HVC_ACC(s, r, HVC_F(r))
    HVC_ACC(s, c, HVC_F(c))
    HVC_ACC(s, Ca, HVC_F(Ca))
#undef HVC_F
}

// HH2_test::state_check (HH2_test.cpp:561-594); returns bit0 soma, bit1 dend
__device__ __forceinline__ int ra_state_check(real Vs, real Vd, int& flag_soma, int& flag_dend)
{
    int fired = 0;
    if ((flag_soma == 1) && (Vs < -HR(30.0) - HR(5.0))) {
        flag_soma = 0;
        fired |= 1;
    } else if ((flag_soma == 0) && (Vs > -HR(30.0))) {
        flag_soma = 1;
    }
    if ((flag_dend == 0) && (Vd >= -HR(30.0))) {
        flag_dend = 1;
        fired |= 2;
    } else if ((flag_dend == 1) && (Vd < -HR(30.0) - HR(5.0))) {
        flag_dend = 0;
    }
    return fired;
}

// ---------------- HVC-I ----------------
struct IState {
    real V, n, m, h, w, Ge, Gi;
#ifdef HVC_FP32
    real V_lo, n_lo, m_lo, h_lo, w_lo;
#endif
};

constexpr real kI_gL = HR(0.1), kI_El = -HR(65.0), kI_gNa = HR(100.0), kI_Ena = HR(55.0), kI_gKdr = HR(20.0);
constexpr real kI_Ek = -HR(80.0), kI_gKHT = HR(500.0), kI_Ei = -HR(75.0), kI_A = HR(6000.0);

#ifdef HVC_FP32
// exp(-(v + d) / B) as 2^(k1 v + k0): v + d folded into one multiply-add
#define HVC_EXPLIN(v, d, B)                                                                          \
    __builtin_amdgcn_exp2f(__builtin_fmaf((v), (real)(-1.4426950408889634 / (double)(B)),           \
                                          (real)(-(double)(d) * 1.4426950408889634 / (double)(B))))
// x / (1 - exp(-x)) by its Taylor series 1 + x/2 + x^2/12 - x^4/720 +
// x^6/30240 - x^8/1209600 (for |x| <= 1/2 the remainder is below 2e-11)
__device__ __forceinline__ real xexpm1_series(real x)
{
    HVC_FP_FAST
    const real x2 = x * x;
    real p = __builtin_fmaf(x2, HR(-1.0 / 1209600.0), HR(1.0 / 30240.0));
    p = __builtin_fmaf(x2, p, HR(-1.0 / 720.0));
    p = __builtin_fmaf(x2, p, HR(1.0 / 12.0));
    return __builtin_fmaf(x2, p, __builtin_fmaf(x, HR(0.5), HR(1.0)));
}
// The rates an and am (HHI_buffer.h:173-176) have the form
// c (v + d) / (1 - exp(-(v + d) / e)).  Both factors vanish at v = -d, where
// the quotient cancels (and is 0/0 at v = -d exactly): within |v + d| <= e/2
// the rate is c e g(x), g = xexpm1_series, x = (v + d) / e (rates_vd_near).
// an (d 15, c 0.15) and am (d 22, c 1), both with e = 10, near their zeros:
// one test for the two (|v + 18.5| <= 8.5 covers both ranges), the series
// evaluated only when a lane of the wave is there
__device__ __forceinline__ bool rates_vd_near(real v, real& an, real& am)
{
    HVC_FP_FAST
    if (__any(__builtin_fabsf(v + HR(18.5)) <= HR(8.5))) {
        const real xn = (v + HR(15.0)) * HR(0.1), xm = (v + HR(22.0)) * HR(0.1);
        const real gn = xexpm1_series(xn), gm = xexpm1_series(xm);
        if (__builtin_fabsf(xn) <= HR(0.5)) an = HR(1.5) * gn;
        if (__builtin_fabsf(xm) <= HR(0.5)) am = HR(10.0) * gm;
        return true;
    }
    return false;
}
template <bool EXACT>
__device__ __forceinline__ real i_bn(real v, Guard& bad) { HVC_FP_FAST return HR(0.2) * HVC_EXPLIN(v, 25, 80); }
template <bool EXACT>
__device__ __forceinline__ real i_bm(real v, Guard& bad) { HVC_FP_FAST return 40 * HVC_EXPLIN(v, 47, 18); }
// an, am, bh, ah and wInf of one neuron (one thread per neuron) from two
// exponentials: with E = exp(-(v + 15) / 10) and F = exp(-(v + 34) / 20),
// exp(-(v + 22) / 10) = E e^-0.7, exp(-(v + 4) / 10) = E e^1.1 and
// exp(-v / 5) = F^4 e^6.8 (one exp2 each instead of five).  an and am as
// quotients hold away from their zeros only: within 5 mV of v = -15 and
// v = -22 rates_vd_near replaces them by their series.
struct IRates {
    real an, am, bh, ah, wInf;
};
__device__ __forceinline__ IRates i_rates_shared(real v)
{
    HVC_FP_FAST
    IRates q;
    const real vn = v + HR(15.0);
    const real e = __builtin_amdgcn_exp2f(vn * (real)(-1.4426950408889634 / 10.0));
    q.an = HR(0.15) * vn * __builtin_amdgcn_rcpf(1 - e);
    q.am = (v + HR(22.0)) * __builtin_amdgcn_rcpf(__builtin_fmaf(-e, HR(0.4965853037914095), HR(1.0)));  // e^-0.7
    q.bh = 10 * __builtin_amdgcn_rcpf(__builtin_fmaf(e, HR(3.0041660239464334), HR(1.0)));              // e^1.1
    const real f = HVC_EXPLIN(v, 34, 20);
    q.ah = HR(0.7) * f;
    const real f2 = f * f;
    q.wInf = __builtin_amdgcn_rcpf(__builtin_fmaf(f2 * f2, HR(897.8472916504176), HR(1.0)));          // e^6.8
    return q;
}
#undef HVC_EXPLIN
#else
template <bool EXACT>
__device__ __forceinline__ real i_an(real v, Guard& bad) { HVC_FP_FAST return DIVR(HR(0.15) * (v + 15), 1 - EXPCD(-(v + 15), 10)); }
template <bool EXACT>
__device__ __forceinline__ real i_bn(real v, Guard& bad) { HVC_FP_FAST return HR(0.2) * EXPCD(-(v + 25), 80); }
template <bool EXACT>
__device__ __forceinline__ real i_am(real v, Guard& bad) { HVC_FP_FAST return DIVR(v + 22, 1 - EXPCD(-(v + 22), 10)); }
template <bool EXACT>
__device__ __forceinline__ real i_bm(real v, Guard& bad) { HVC_FP_FAST return 40 * EXPCD(-(v + 47), 18); }
template <bool EXACT>
__device__ __forceinline__ real i_ah(real v, Guard& bad) { HVC_FP_FAST return HR(0.7) * EXPCD(-(v + 34), 20); }
template <bool EXACT>
__device__ __forceinline__ real i_bh(real v, Guard& bad) { HVC_FP_FAST return DIVR(10, 1 + EXPCD(-(v + 4), 10)); }
template <bool EXACT>
__device__ __forceinline__ real i_wInf(real v, Guard& bad) { HVC_FP_FAST return DIVR(1, 1 + EXPCD(-v, 5)); }
#endif

struct IK {
    real V, n, m, h, w;
};

// HHI_buffer.cpp:741-747 at stage state (v, n1, m1, h1, w1); e2/e5 decay factors
template <bool EXACT>
__device__ __forceinline__ IK i_rhs(Guard& bad, const IState& s, real v, real n1, real m1, real h1, real w1,
                                    real e2, real e5)
{
    HVC_FP_FAST
    IK k;
#ifdef HVC_FP32
    // one chain of multiply-adds, as in the quad form
    const real n2 = n1 * n1;
    const real gk = __builtin_fmaf(kI_gKdr, n2 * n2, kI_gKHT * w1);
    real t = __builtin_fmaf(kI_gL, v, -(kI_gL * kI_El));
    t = __builtin_fmaf(kI_gNa * h1 * (m1 * m1 * m1), v - kI_Ena, t);
    t = __builtin_fmaf(gk, v - kI_Ek, t);
    t = __builtin_fmaf(s.Ge * e2, v, t);
    t = __builtin_fmaf(s.Gi * e5, v - kI_Ei, t);
    k.V = -t;
    auto gate = [](real a, real b, real x) { return __builtin_fmaf(-(a + b), x, a); };   // a (1 - x) - b x
    IRates q = i_rates_shared(v);
    const real bn = i_bn<EXACT>(v, bad), bm = i_bm<EXACT>(v, bad);
    k.h = gate(q.ah, q.bh, h1);
    k.n = gate(q.an, bn, n1);
    k.m = gate(q.am, bm, m1);
    // an and am near their zeros: the series, after the rest of the stage
    if (rates_vd_near(v, q.an, q.am)) {
        k.n = gate(q.an, bn, n1);
        k.m = gate(q.am, bm, m1);
    }
    k.w = q.wInf - w1;   // divided by tauW = 1 (exact)
#else
    real n4 = n1 * n1 * n1 * n1;
    real m3 = m1 * m1 * m1;
    k.V = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
          - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * HR(0.0) / kI_A;
    k.n = i_an<EXACT>(v, bad) * (1 - n1) - i_bn<EXACT>(v, bad) * n1;
    k.m = i_am<EXACT>(v, bad) * (1 - m1) - i_bm<EXACT>(v, bad) * m1;
    k.h = i_ah<EXACT>(v, bad) * (1 - h1) - i_bh<EXACT>(v, bad) * h1;
    k.w = (i_wInf<EXACT>(v, bad) - w1);  // divided by tauW = 1 (exact)
#endif
    return k;
}

// dt times the common factor of DP8 stages 2-6 (HHI_buffer.cpp:866-890): the
// reference's order in the FP64 form; in the FP32 form dt and the factor are
// one loop-invariant product
#ifdef HVC_FP32
#define HVC_F2(x) ((x) * (dt * (real)(1.0 / 18)))
#define HVC_F3(x) ((x) * (dt * (real)(1.0 / 48)))
#define HVC_F4(x) ((x) * (dt * HR(0.03125)))
#define HVC_F5(x) ((x) * (dt * HR(0.015625)))
#define HVC_F6(x) ((x) * (dt * (real)(1.0 / 80)))
#else
#define HVC_F2(x) CD(dt * (x), 18)
#define HVC_F3(x) CD(dt * (x), 48)
#define HVC_F4(x) dt * (x) * HR(0.03125)
#define HVC_F5(x) dt * (x) * HR(0.015625)
#define HVC_F6(x) CD(dt * (x), 80)
#endif

// HHI_buffer::DP8_step; row = decay factors of this step
template <bool EXACT, class Row>
__device__ __forceinline__ void i_dp8(Guard& bad, IState& s, real dt, const Row& row)
{
    HVC_FP_FAST
    IK k1 = i_rhs<EXACT>(bad, s, s.V, s.n, s.m, s.h, s.w, HR(1.0), HR(1.0));
    IK k2 = i_rhs<EXACT>(bad, s, s.V + HVC_F2(k1.V), s.n + HVC_F2(k1.n), s.m + HVC_F2(k1.m), s.h + HVC_F2(k1.h),
                  s.w + HVC_F2(k1.w), row.e2[kOffDp2], row.e5[kOffDp2]);
    IK k3 = i_rhs<EXACT>(bad, s, s.V + HVC_F3(k1.V + 3 * k2.V), s.n + HVC_F3(k1.n + 3 * k2.n),
                  s.m + HVC_F3(k1.m + 3 * k2.m), s.h + HVC_F3(k1.h + 3 * k2.h),
                  s.w + HVC_F3(k1.w + 3 * k2.w), row.e2[kOffDp3], row.e5[kOffDp3]);
    IK k4 = i_rhs<EXACT>(bad, s, s.V + HVC_F4(k1.V + 3 * k3.V), s.n + HVC_F4(k1.n + 3 * k3.n),
                  s.m + HVC_F4(k1.m + 3 * k3.m), s.h + HVC_F4(k1.h + 3 * k3.h),
                  s.w + HVC_F4(k1.w + 3 * k3.w), row.e2[kOffDp4], row.e5[kOffDp4]);
#define C5(x) (20 * k1.x - 75 * k3.x + 75 * k4.x)
    IK k5 = i_rhs<EXACT>(bad, s, s.V + HVC_F5(C5(V)), s.n + HVC_F5(C5(n)), s.m + HVC_F5(C5(m)), s.h + HVC_F5(C5(h)),
                  s.w + HVC_F5(C5(w)), row.e2[kOffDp5], row.e5[kOffDp5]);
#define C6(x) (3 * k1.x + 15 * k4.x + 12 * k5.x)
    IK k6 = i_rhs<EXACT>(bad, s, s.V + HVC_F6(C6(V)), s.n + HVC_F6(C6(n)), s.m + HVC_F6(C6(m)), s.h + HVC_F6(C6(h)),
                  s.w + HVC_F6(C6(w)), row.e2[kOffDp6], row.e5[kOffDp6]);
#define C7(x) (CK(29443841, k1.x, 614563906) + CK(77736538, k4.x, 692538347) - CK(28693883, k5.x, 1125000000) + \
               CK(23124283, k6.x, 1800000000))
    IK k7 = i_rhs<EXACT>(bad, s, s.V + dt * C7(V), s.n + dt * C7(n), s.m + dt * C7(m), s.h + dt * C7(h), s.w + dt * C7(w),
                  row.e2[kOffDp7], row.e5[kOffDp7]);
#define C8(x) (CK(16016141, k1.x, 946692911) + CK(61564180, k4.x, 158732637) + CK(22789713, k5.x, 633445777) + \
               CK(545815736, k6.x, 2771057229) - CK(180193667, k7.x, 1043307555))
    IK k8 = i_rhs<EXACT>(bad, s, s.V + dt * C8(V), s.n + dt * C8(n), s.m + dt * C8(m), s.h + dt * C8(h), s.w + dt * C8(w),
                  row.e2[kOffDp8], row.e5[kOffDp8]);
#define C9(x) (CK(39632708, k1.x, 573591083) - CK(433636366, k4.x, 683701615) - CK(421739975, k5.x, 2616292301) + \
               CK(100302831, k6.x, 723423059) + CK(790204164, k7.x, 839813087) + CK(800635310, k8.x, 3783071287))
    IK k9 = i_rhs<EXACT>(bad, s, s.V + dt * C9(V), s.n + dt * C9(n), s.m + dt * C9(m), s.h + dt * C9(h), s.w + dt * C9(w),
                  row.e2[kOffDp9], row.e5[kOffDp9]);
#define C10(x) (CK(246121993, k1.x, 1340847787) - CK(37695042795, k4.x, 15268766246) - CK(309121744, k5.x, 1061227803) - \
                CK(12992083, k6.x, 490766935) + CK(6005943493, k7.x, 2108947869) + CK(393006217, k8.x, 1396673457) + \
                CK(123872331, k9.x, 1001029789))
    IK k10 = i_rhs<EXACT>(bad, s, s.V + dt * C10(V), s.n + dt * C10(n), s.m + dt * C10(m), s.h + dt * C10(h),
                   s.w + dt * C10(w), row.e2[kOffDp10], row.e5[kOffDp10]);
#define C11(x) (CK(-1028468189, k1.x, 846180014) + CK(8478235783, k4.x, 508512852) + CK(1311729495, k5.x, 1432422823) - \
                CK(10304129995, k6.x, 1701304382) - CK(48777925059, k7.x, 3047939560) + \
                CK(15336726248, k8.x, 1032824649) - CK(45442868181, k9.x, 3398467696) + CK(3065993473, k10.x, 597172653))
    IK k11 = i_rhs<EXACT>(bad, s, s.V + dt * C11(V), s.n + dt * C11(n), s.m + dt * C11(m), s.h + dt * C11(h),
                   s.w + dt * C11(w), row.e2[kOffDp11], row.e5[kOffDp11]);
#define C12(x) (CK(185892177, k1.x, 718116043) - CK(3185094517, k4.x, 667107341) - CK(477755414, k5.x, 1098053517) - \
                CK(703635378, k6.x, 230739211) + CK(5731566787, k7.x, 1027545527) + CK(5232866602, k8.x, 850066563) - \
                CK(4093664535, k9.x, 808688257) + CK(3962137247, k10.x, 1805957418) + CK(65686358, k11.x, 487910083))
    IK k12 = i_rhs<EXACT>(bad, s, s.V + dt * C12(V), s.n + dt * C12(n), s.m + dt * C12(m), s.h + dt * C12(h),
                   s.w + dt * C12(w), row.e2[kOffDt], row.e5[kOffDt]);
#define CF(x) (CK(13451932, k1.x, 455176623) - CK(808719846, k6.x, 976000145) + CK(1757004468, k7.x, 5645159321) + \
               CK(656045339, k8.x, 265891186) - CK(3867574721, k9.x, 1518517206) + CK(465885868, k10.x, 322736535) + \
               CK(53011238, k11.x, 667516719) + CK(2, k12.x, 45))
    HVC_ACC(s, V, dt * CF(V))
    HVC_ACC(s, n, dt * CF(n))
    HVC_ACC(s, m, dt * CF(m))
    HVC_ACC(s, h, dt * CF(h))
    HVC_ACC(s, w, dt * CF(w))
    s.Ge = s.Ge * row_e2dt(row);
    s.Gi = s.Gi * row_e5dt(row);
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
template <bool EXACT, class Row>
__device__ __forceinline__ void i_runge4(Guard& bad, IState& s, real dt, const Row& row)
{
    HVC_FP_FAST
    IK k1 = i_rhs<EXACT>(bad, s, s.V, s.n, s.m, s.h, s.w, HR(1.0), HR(1.0));
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
#define F(x) (dt * (k1.x + 3 * k2.x + 3 * k3.x + k4.x) * HR(0.125))
    HVC_ACC(s, V, F(V))
    HVC_ACC(s, n, F(n))
    HVC_ACC(s, m, F(m))
    HVC_ACC(s, h, F(h))
    HVC_ACC(s, w, F(w))
#undef F
    s.Ge = s.Ge * row_e2dt(row);
    s.Gi = s.Gi * row_e5dt(row);
}

// HHI_buffer::state_check (HHI_buffer.cpp:260-276)
__device__ __forceinline__ int i_state_check(real V, int& flag)
{
    if ((flag == 1) && (V < -HR(20.0) - HR(5.0))) {
        flag = 0;
        return 1;
    }
    if ((flag == 0) && (V > -HR(20.0))) flag = 1;
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
__device__ __forceinline__ void i_noise_check(PoissonStream& ps, double time, double dt, real& G,
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
__device__ __forceinline__ void ra_noise_check(PoissonStream& ps, double time, double dt, real& G,
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

__constant__ real kPointDistribution[30] = {
    HR(0x1.3988e1409212ep+1), -HR(0x1.3988e1409212ep+1), 1, 1, 1, 1, 1, 1, 1, 1, 1,
    -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};



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
    real c, d, e, inv_e;
    real inv2_e;     // RN(log2(e) / e) (FP32 form, EXPCDV)
};

template <bool EXACT, bool CAN_A = true>
__device__ __forceinline__ real gate_rate(Guard& bad, const GateSlot& g, real v)
{
    HVC_FP_FAST
    real vd = v + g.d;
    real E = EXPCDV(-vd, g.e, g.inv_e, g.inv2_e);
    real num = (g.type == 0) ? g.c * vd : g.c;
    real den = (g.type == 0) ? (1 - E) : (1 + E);
    real q = DIVR(num, den);
    HVC_KEEP(q);
#ifdef HVC_FP32
    // form A near v = -d: the series (as rates_vd_near)
    if constexpr (CAN_A) {
        const bool near = g.type == 0 && __builtin_fabsf(vd) <= HR(0.5) * g.e;
        if (__any(near)) {
            const real sr = (g.c * g.e) * xexpm1_series(vd * g.inv_e);
            if (near) q = sr;
        }
    }
#endif
    return (g.type == 1) ? g.c * E : q;
}

__constant__ GateSlot kIGates[5][2] = {
    {{1, HR(0.0), HR(0.0), HR(10.0), HR(0.1), HVC_L2(10.0)}, {1, HR(0.0), HR(0.0), HR(10.0), HR(0.1), HVC_L2(10.0)}},   // V: no rates (unused)
    {{0, HR(0.15), HR(15.0), HR(10.0), HR(0.1), HVC_L2(10.0)},
     {1, HR(0.2), HR(25.0), HR(80.0), HR(1.0) / HR(80.0), HVC_L2(80.0)}},                                                  // n: an, bn
    {{0, HR(1.0), HR(22.0), HR(10.0), HR(0.1), HVC_L2(10.0)},
     {1, HR(40.0), HR(47.0), HR(18.0), HR(1.0) / HR(18.0), HVC_L2(18.0)}},                                                 // m: am, bm
    {{1, HR(0.7), HR(34.0), HR(20.0), HR(1.0) / HR(20.0), HVC_L2(20.0)},
     {2, HR(10.0), HR(4.0), HR(10.0), HR(0.1), HVC_L2(10.0)}},                                                            // h: ah, bh
    {{2, HR(1.0), HR(0.0), HR(5.0), HR(0.2), HVC_L2(5.0)}, {1, HR(0.0), HR(0.0), HR(10.0), HR(0.1), HVC_L2(10.0)}},     // w: wInf (unused)
};

// derivative of the lane's variable at stage state (v, n1, m1, h1, w1)
template <bool EXACT>
__device__ __forceinline__ real i_lane_rhs(Guard& bad, const IState& s, int role, real v, real n1, real m1, real h1,
                                             real w1, real e2, real e5)
{
    HVC_FP_FAST
    real n4 = n1 * n1 * n1 * n1;
    real m3 = m1 * m1 * m1;
    real kV = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
                - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * HR(0.0) / kI_A;
    real a = gate_rate<EXACT>(bad, kIGates[role][0], v);
    real b = gate_rate<EXACT, false>(bad, kIGates[role][1], v);   // second slots are never form A
    real x = role == 1 ? n1 : (role == 2 ? m1 : (role == 3 ? h1 : w1));
    real kg = a * (1 - x) - b * x;
    real kw = (a - x);
    return role == 0 ? kV : (role == 4 ? kw : kg);
}

// DP8 stage values: the combination of stage derivatives P1.. of one
// variable (HHI_buffer.cpp:426-691), shared by the five-lane and quad forms;
// the tableau terms of stages 7-12 and the final combination are HVC_CKX
// (CK, or CKN where the step notes each stage derivative itself)
#define HVC_CKX CK
#define HVC_DP8_S2(P) HVC_F2(P##1)
#define HVC_DP8_S3(P) HVC_F3(P##1 + 3 * P##2)
#define HVC_DP8_S4(P) HVC_F4(P##1 + 3 * P##3)
#define HVC_DP8_S5(P) HVC_F5(20 * P##1 - 75 * P##3 + 75 * P##4)
#define HVC_DP8_S6(P) HVC_F6(3 * P##1 + 15 * P##4 + 12 * P##5)
#define HVC_DP8_S7(P) dt * (HVC_CKX(29443841, P##1, 614563906) + HVC_CKX(77736538, P##4, 692538347) - \
                            HVC_CKX(28693883, P##5, 1125000000) + HVC_CKX(23124283, P##6, 1800000000))
#define HVC_DP8_S8(P) dt * (HVC_CKX(16016141, P##1, 946692911) + HVC_CKX(61564180, P##4, 158732637) + \
                            HVC_CKX(22789713, P##5, 633445777) + HVC_CKX(545815736, P##6, 2771057229) - \
                            HVC_CKX(180193667, P##7, 1043307555))
#define HVC_DP8_S9(P) dt * (HVC_CKX(39632708, P##1, 573591083) - HVC_CKX(433636366, P##4, 683701615) - \
                            HVC_CKX(421739975, P##5, 2616292301) + HVC_CKX(100302831, P##6, 723423059) + \
                            HVC_CKX(790204164, P##7, 839813087) + HVC_CKX(800635310, P##8, 3783071287))
#define HVC_DP8_S10(P) dt * (HVC_CKX(246121993, P##1, 1340847787) - HVC_CKX(37695042795, P##4, 15268766246) - \
                             HVC_CKX(309121744, P##5, 1061227803) - HVC_CKX(12992083, P##6, 490766935) + \
                             HVC_CKX(6005943493, P##7, 2108947869) + HVC_CKX(393006217, P##8, 1396673457) + \
                             HVC_CKX(123872331, P##9, 1001029789))
#define HVC_DP8_S11(P) dt * (HVC_CKX(-1028468189, P##1, 846180014) + HVC_CKX(8478235783, P##4, 508512852) + \
                             HVC_CKX(1311729495, P##5, 1432422823) - HVC_CKX(10304129995, P##6, 1701304382) - \
                             HVC_CKX(48777925059, P##7, 3047939560) + HVC_CKX(15336726248, P##8, 1032824649) - \
                             HVC_CKX(45442868181, P##9, 3398467696) + HVC_CKX(3065993473, P##10, 597172653))
#define HVC_DP8_S12(P) dt * (HVC_CKX(185892177, P##1, 718116043) - HVC_CKX(3185094517, P##4, 667107341) - \
                             HVC_CKX(477755414, P##5, 1098053517) - HVC_CKX(703635378, P##6, 230739211) + \
                             HVC_CKX(5731566787, P##7, 1027545527) + HVC_CKX(5232866602, P##8, 850066563) - \
                             HVC_CKX(4093664535, P##9, 808688257) + HVC_CKX(3962137247, P##10, 1805957418) + \
                             HVC_CKX(65686358, P##11, 487910083))
#define HVC_DP8_FIN(P) dt * (HVC_CKX(13451932, P##1, 455176623) - HVC_CKX(808719846, P##6, 976000145) + \
                             HVC_CKX(1757004468, P##7, 5645159321) + HVC_CKX(656045339, P##8, 265891186) - \
                             HVC_CKX(3867574721, P##9, 1518517206) + HVC_CKX(465885868, P##10, 322736535) + \
                             HVC_CKX(53011238, P##11, 667516719) + HVC_CKX(2, P##12, 45))

#define HVC_BCAST5(x)                       \
    v = __shfl(x, base + 0);                \
    n1 = __shfl(x, base + 1);               \
    m1 = __shfl(x, base + 2);               \
    h1 = __shfl(x, base + 3);               \
    w1 = __shfl(x, base + 4);

// HHI_buffer::DP8_step with five lanes per neuron; s is replicated in the group
template <bool EXACT, class Row>
__device__ __forceinline__ void i_dp8_lanes(Guard& bad, IState& s, real& xlo, int role, int base, real dt,
                                            const Row& row)
{
    HVC_FP_FAST
    const real x0 = role == 0 ? s.V : (role == 1 ? s.n : (role == 2 ? s.m : (role == 3 ? s.h : s.w)));
    real v = s.V, n1 = s.n, m1 = s.m, h1 = s.h, w1 = s.w;
    real x;
    real k1 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, HR(1.0), HR(1.0));
    x = x0 + HVC_DP8_S2(k);
    HVC_BCAST5(x)
    real k2 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp2], row.e5[kOffDp2]);
    x = x0 + HVC_DP8_S3(k);
    HVC_BCAST5(x)
    real k3 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp3], row.e5[kOffDp3]);
    x = x0 + HVC_DP8_S4(k);
    HVC_BCAST5(x)
    real k4 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp4], row.e5[kOffDp4]);
    x = x0 + HVC_DP8_S5(k);
    HVC_BCAST5(x)
    real k5 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp5], row.e5[kOffDp5]);
    x = x0 + HVC_DP8_S6(k);
    HVC_BCAST5(x)
    real k6 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp6], row.e5[kOffDp6]);
    x = x0 + HVC_DP8_S7(k);
    HVC_BCAST5(x)
    real k7 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp7], row.e5[kOffDp7]);
    x = x0 + HVC_DP8_S8(k);
    HVC_BCAST5(x)
    real k8 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp8], row.e5[kOffDp8]);
    x = x0 + HVC_DP8_S9(k);
    HVC_BCAST5(x)
    real k9 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp9], row.e5[kOffDp9]);
    x = x0 + HVC_DP8_S10(k);
    HVC_BCAST5(x)
    real k10 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp10], row.e5[kOffDp10]);
    x = x0 + HVC_DP8_S11(k);
    HVC_BCAST5(x)
    real k11 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDp11], row.e5[kOffDp11]);
    x = x0 + HVC_DP8_S12(k);
    HVC_BCAST5(x)
    real k12 = i_lane_rhs<EXACT>(bad, s, role, v, n1, m1, h1, w1, row.e2[kOffDt], row.e5[kOffDt]);
    {
        struct { real x, x_lo; } acc{x0, xlo};
        HVC_ACC(acc, x, HVC_DP8_FIN(k))
        x = acc.x;
        xlo = acc.x_lo;
    }
    HVC_BCAST5(x)
    s.V = v; s.n = n1; s.m = m1; s.h = h1; s.w = w1;
    s.Ge = s.Ge * row_e2dt(row);
    s.Gi = s.Gi * row_e5dt(row);
}
#undef HVC_BCAST5

// ---- HVC-I, DP8, four lanes per neuron (quad) ----
// Lane q of a neuron's quad owns n, m or h (q = 1..3) and evaluates that
// variable's two gate rates; lane 0 owns V and w and evaluates wInf (V has no
// rates).  Only V and the three gates are needed by other lanes, so four
// stage values travel by DPP quad broadcasts and w stays in lane 0.  Every
// value is the expression tree of HHI_buffer, as in i_dp8 and the five-lane
// form.
__constant__ GateSlot kIQuadGates[4][2] = {
    {{2, HR(1.0), HR(0.0), HR(5.0), HR(0.2), HVC_L2(5.0)}, {1, HR(0.0), HR(0.0), HR(10.0), HR(0.1), HVC_L2(10.0)}},   // V, w: wInf (second slot unused)
    {{0, HR(0.15), HR(15.0), HR(10.0), HR(0.1), HVC_L2(10.0)},
     {1, HR(0.2), HR(25.0), HR(80.0), HR(1.0) / HR(80.0), HVC_L2(80.0)}},                                                // n: an, bn
    {{0, HR(1.0), HR(22.0), HR(10.0), HR(0.1), HVC_L2(10.0)},
     {1, HR(40.0), HR(47.0), HR(18.0), HR(1.0) / HR(18.0), HVC_L2(18.0)}},                                               // m: am, bm
    {{1, HR(0.7), HR(34.0), HR(20.0), HR(1.0) / HR(20.0), HVC_L2(20.0)},
     {2, HR(10.0), HR(4.0), HR(10.0), HR(0.1), HVC_L2(10.0)}},                                                          // h: ah, bh
};

// FP32 form of a gate slot: with per-lane constants the three rate forms are
// one expression, E = exp(-(v + d) / e), rate = cB E + (cA (v + d) + cC) / (1 + sg E)
// (A: cA = c, sg = -1; B: cB = c; C: cC = c), no lane-dependent selects.
// Form A cancels near v = -d (rates_vd_near): its lanes switch to c e g((v + d) / e)
// within |v + d| <= e/2 (lim; -1 in the other forms), when a lane of the wave is there.
struct GateK {
    real cA, cB, cC, sg, k1, d, ie, ce, lim;
};
__device__ __forceinline__ GateK gate_k(const GateSlot& g)
{
    const bool A = g.type == 0;
    return GateK{A ? g.c : HR(0.0), g.type == 1 ? g.c : HR(0.0), g.type == 2 ? g.c : HR(0.0),
                 A ? -HR(1.0) : HR(1.0), -g.inv2_e, g.d, g.inv_e, g.c * g.e, A ? HR(0.5) * g.e : -HR(1.0)};
}
#ifdef HVC_FP32
template <bool CAN_A>
__device__ __forceinline__ real gate_rate_k(const GateK& g, real v)
{
    HVC_FP_FAST
    const real vd = v + g.d;
    const real E = __builtin_amdgcn_exp2f(vd * g.k1);
    const real q = __builtin_fmaf(g.cA, vd, g.cC) * __builtin_amdgcn_rcpf(__builtin_fmaf(g.sg, E, HR(1.0)));
    real r = __builtin_fmaf(g.cB, E, q);
    if constexpr (CAN_A) {
        const bool near = __builtin_fabsf(vd) <= g.lim;
        if (__any(near)) {
            const real s = g.ce * xexpm1_series(vd * g.ie);
            if (near) r = s;
        }
    }
    return r;
}
#endif

// derivatives of the lane's variable (kx) and, in lane 0, of w (kw) at stage
// state (v, n1, m1, h1, w1); xo = the lane's own variable at the stage; w1 is
// valid in lane 0 only (kw of the other lanes is not used)
template <bool EXACT>
__device__ __forceinline__ void i_quad_rhs(Guard& bad, const IState& s, int role, const GateSlot& ga,
                                           const GateSlot& gb, real v, real n1, real m1, real h1, real w1, real xo,
                                           real e2, real e5, real& kx, real& kw)
{
    HVC_FP_FAST
#ifdef HVC_FP32
    // the same currents as one chain of multiply-adds: the two potassium
    // conductances share (v - E_K), the leak's reversal term is a constant
    // (the reference's closing "+ 100000 * 0 / A" only turns -0 into +0)
    const real n2 = n1 * n1;
    const real gk = __builtin_fmaf(kI_gKdr, n2 * n2, kI_gKHT * w1);
    real t = __builtin_fmaf(kI_gL, v, -(kI_gL * kI_El));
    t = __builtin_fmaf(kI_gNa * h1 * (m1 * m1 * m1), v - kI_Ena, t);
    t = __builtin_fmaf(gk, v - kI_Ek, t);
    t = __builtin_fmaf(s.Ge * e2, v, t);
    t = __builtin_fmaf(s.Gi * e5, v - kI_Ei, t);
    const real kV = -t;
    const GateK gka = gate_k(ga);
    real a = gate_rate_k<false>(gka, v);
    const real b = gate_rate_k<false>(gate_k(gb), v);   // second slots are never form A
    kw = a - w1;   // a = wInf in lane 0
    real kg = __builtin_fmaf(-(a + b), xo, a);   // a (1 - x) - b x
    // form A of the first slot near v = -d (rates_vd_near): the series, and the
    // derivatives again, when a lane of the wave is there (after the rest of
    // the stage, so that its work is not split by the branch)
    const real vd = v + gka.d;
    const bool near = __builtin_fabsf(vd) <= gka.lim;
    if (__any(near)) {
        const real sa = gka.ce * xexpm1_series(vd * gka.ie);
        if (near) {
            a = sa;
            kg = __builtin_fmaf(-(a + b), xo, a);
        }
    }
#else
    real n4 = n1 * n1 * n1 * n1;
    real m3 = m1 * m1 * m1;
    real kV = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
                - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * HR(0.0) / kI_A;
    real a = gate_rate<EXACT>(bad, ga, v);
    real b = gate_rate<EXACT>(bad, gb, v);
    kw = role == 0 ? (a - w1) : HR(0.0);   // a = wInf in lane 0
    real kg = a * (1 - xo) - b * xo;
#endif
    kx = role == 0 ? kV : kg;
}

#define HVC_BCAST4(x)              \
    v = quad_bcast<0>(x);          \
    n1 = quad_bcast<1>(x);         \
    m1 = quad_bcast<2>(x);         \
    h1 = quad_bcast<3>(x);

// HHI_buffer::DP8_step with four lanes per neuron; s is replicated in the quad
// except w (lane 0); xlo / wlo: low parts of the lane's variable and of w (FP32 form)
template <bool EXACT, class Row>
__device__ __forceinline__ void i_dp8_quad(Guard& bad, IState& s, real& xlo, real& wlo, int role, const GateSlot& ga,
                                           const GateSlot& gb, real dt, const Row& row)
{
    HVC_FP_FAST
    const real x0 = role == 0 ? s.V : (role == 1 ? s.n : (role == 2 ? s.m : s.h));
    const real y0 = s.w;
    real v = s.V, n1 = s.n, m1 = s.m, h1 = s.h, w1 = s.w;
    real x;
#define HVC_QSTAGE(N, e2, e5, SX)                                                          \
    real k_##N, q_##N;                                                                   \
    i_quad_rhs<EXACT>(bad, s, role, ga, gb, v, n1, m1, h1, w1, x, e2, e5, k_##N, q_##N); \
    x = x0 + SX(k_);                                                                     \
    w1 = y0 + SX(q_);                                                                    \
    HVC_BCAST4(x)
    real k_1, q_1;
    i_quad_rhs<EXACT>(bad, s, role, ga, gb, v, n1, m1, h1, w1, x0, HR(1.0), HR(1.0), k_1, q_1);
    x = x0 + HVC_DP8_S2(k_);
    w1 = y0 + HVC_DP8_S2(q_);
    HVC_BCAST4(x)
    HVC_QSTAGE(2, row.e2[kOffDp2], row.e5[kOffDp2], HVC_DP8_S3)
    HVC_QSTAGE(3, row.e2[kOffDp3], row.e5[kOffDp3], HVC_DP8_S4)
    HVC_QSTAGE(4, row.e2[kOffDp4], row.e5[kOffDp4], HVC_DP8_S5)
    HVC_QSTAGE(5, row.e2[kOffDp5], row.e5[kOffDp5], HVC_DP8_S6)
    HVC_QSTAGE(6, row.e2[kOffDp6], row.e5[kOffDp6], HVC_DP8_S7)
    HVC_QSTAGE(7, row.e2[kOffDp7], row.e5[kOffDp7], HVC_DP8_S8)
    HVC_QSTAGE(8, row.e2[kOffDp8], row.e5[kOffDp8], HVC_DP8_S9)
    HVC_QSTAGE(9, row.e2[kOffDp9], row.e5[kOffDp9], HVC_DP8_S10)
    HVC_QSTAGE(10, row.e2[kOffDp10], row.e5[kOffDp10], HVC_DP8_S11)
    HVC_QSTAGE(11, row.e2[kOffDp11], row.e5[kOffDp11], HVC_DP8_S12)
#undef HVC_QSTAGE
    real k_12, q_12;
    i_quad_rhs<EXACT>(bad, s, role, ga, gb, v, n1, m1, h1, w1, x, row.e2[kOffDt], row.e5[kOffDt], k_12, q_12);
    {
        struct { real x, x_lo; } acc{x0, xlo};
        HVC_ACC(acc, x, HVC_DP8_FIN(k_))
        x = acc.x;
        xlo = acc.x_lo;
        struct { real x, x_lo; } accw{y0, wlo};
        HVC_ACC(accw, x, HVC_DP8_FIN(q_))
        w1 = accw.x;
        wlo = accw.x_lo;
    }
    HVC_BCAST4(x)
    s.V = v; s.n = n1; s.m = m1; s.h = h1; s.w = w1;
    s.Ge = s.Ge * row_e2dt(row);
    s.Gi = s.Gi * row_e5dt(row);
}
#undef HVC_BCAST4

// ---- HVC-I, DP8, eight lanes per neuron (octet) ----
// Lanes q and 7 - q of a neuron's octet form a pair, exchanged by the DPP
// half-row mirror: (1, 6) carry n and evaluate an / bn, (2, 5) m with am / bm,
// (3, 4) h with ah / bh.  Each lane of a pair evaluates one of its variable's
// two rates, takes the other from its partner and advances the variable with
// the same operations on the same values, so both hold the same bits.  Lane 0
// owns V, lane 7 owns w and evaluates wInf.  A stage thus evaluates one rate
// per lane (the quad two) and one stage combination (the quad two); lane 0's
// membrane derivative gets n, m, h by quad broadcasts and w by the mirror.
// Every value is the expression tree of HHI_buffer, as in i_dp8.
__constant__ GateSlot kIOctGates[8] = {
    {1, HR(0.0), HR(0.0), HR(10.0), HR(0.1), HVC_L2(10.0)},                 // V: no rate (unused)
    {0, HR(0.15), HR(15.0), HR(10.0), HR(0.1), HVC_L2(10.0)},               // an
    {0, HR(1.0), HR(22.0), HR(10.0), HR(0.1), HVC_L2(10.0)},                // am
    {1, HR(0.7), HR(34.0), HR(20.0), HR(1.0) / HR(20.0), HVC_L2(20.0)},     // ah
    {2, HR(10.0), HR(4.0), HR(10.0), HR(0.1), HVC_L2(10.0)},                // bh
    {1, HR(40.0), HR(47.0), HR(18.0), HR(1.0) / HR(18.0), HVC_L2(18.0)},    // bm
    {1, HR(0.2), HR(25.0), HR(80.0), HR(1.0) / HR(80.0), HVC_L2(80.0)},     // bn
    {2, HR(1.0), HR(0.0), HR(5.0), HR(0.2), HVC_L2(5.0)},                   // wInf
};

// value of lane 7 - q of the octet (row_half_mirror), in lane q
__device__ __forceinline__ float oct_mirror(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x141, 0xF, 0xF, true));
}
__device__ __forceinline__ double oct_mirror(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x141, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x141, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// derivative of the lane's variable at stage state (v, n1, m1, h1, w1); xo =
// the lane's own variable at the stage; n1, m1, h1, w1 are needed in lane 0 only
template <bool EXACT>
__device__ __forceinline__ real i_oct_rhs(Guard& bad, const IState& s, int q, const GateSlot& g, real v, real n1,
                                          real m1, real h1, real w1, real xo, real e2, real e5)
{
    HVC_FP_FAST
#ifdef HVC_FP32
    const real n2 = n1 * n1;
    const real gk = __builtin_fmaf(kI_gKdr, n2 * n2, kI_gKHT * w1);
    real t = __builtin_fmaf(kI_gL, v, -(kI_gL * kI_El));
    t = __builtin_fmaf(kI_gNa * h1 * (m1 * m1 * m1), v - kI_Ena, t);
    t = __builtin_fmaf(gk, v - kI_Ek, t);
    t = __builtin_fmaf(s.Ge * e2, v, t);
    t = __builtin_fmaf(s.Gi * e5, v - kI_Ei, t);
    const real kV = -t;
    const GateK gk1 = gate_k(g);
    real r = gate_rate_k<false>(gk1, v);
    // form A (an, am) near v = -d: the series (rates_vd_near), when a lane of the wave is there
    const real vd = v + gk1.d;
    const bool near = __builtin_fabsf(vd) <= gk1.lim;
    if (__any(near)) {
        const real sr = gk1.ce * xexpm1_series(vd * gk1.ie);
        if (near) r = sr;
    }
    const real mr = oct_mirror(r);
    const real a = q < 4 ? r : mr, b = q < 4 ? mr : r;
    const real kg = __builtin_fmaf(-(a + b), xo, a);   // a (1 - x) - b x
#else
    real n4 = n1 * n1 * n1 * n1;
    real m3 = m1 * m1 * m1;
    real kV = -kI_gL * (v - kI_El) - kI_gNa * h1 * m3 * (v - kI_Ena) - kI_gKdr * n4 * (v - kI_Ek)
                - kI_gKHT * w1 * (v - kI_Ek) - (s.Ge * e2) * v - (s.Gi * e5) * (v - kI_Ei) + 100000 * HR(0.0) / kI_A;
    const real r = gate_rate<EXACT>(bad, g, v);
    const real mr = oct_mirror(r);
    const real a = q < 4 ? r : mr, b = q < 4 ? mr : r;
    const real kg = a * (1 - xo) - b * xo;
#endif
    const real kw = r - xo;   // lane 7: wInf - w
    return q == 0 ? kV : (q == 7 ? kw : kg);
}

// the stage values every lane needs: V in all lanes; n, m, h (quad
// broadcasts) and w (mirror of lane 7) in lane 0
// (both broadcasts of V run in every lane and a select picks one: a DPP move
// under a lane-dependent condition would split the wave's exec mask)
#define HVC_BCAST8(x)                                            \
    {                                                            \
        const real t_ = oct_mirror(x);                           \
        const real v_lo_ = quad_bcast<0>(x);                     \
        const real v_hi_ = quad_bcast<3>(t_);                    \
        v = q < 4 ? v_lo_ : v_hi_;                               \
        n1 = quad_bcast<1>(x);                                   \
        m1 = quad_bcast<2>(x);                                   \
        h1 = quad_bcast<3>(x);                                   \
        w1 = t_;                                                 \
    }

// HHI_buffer::DP8_step with eight lanes per neuron; xo / xlo: the lane's own
// variable and (FP32 form) its low part; s.V is current in every lane, s.n,
// s.m, s.h, s.w in lane 0 only
template <bool EXACT, class Row>
__device__ __forceinline__ void i_dp8_oct(Guard& bad, IState& s, real& xo, real& xlo, int q, const GateSlot& g,
                                          real dt, const Row& row)
{
    HVC_FP_FAST
    const real x0 = xo;
    real v = s.V, n1 = s.n, m1 = s.m, h1 = s.h, w1 = s.w;
    real x;
#ifdef HVC_FP32
    // the step's decay factors of stages 3-13, read from LDS together before
    // the stages (each stage's read would otherwise wait on its own round trip)
    real e2r[11], e5r[11];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        e2r[k] = row.e2[kOffDp2 + k];
        e5r[k] = row.e5[kOffDp2 + k];
    }
    e2r[10] = row.e2[kOffDt];
    e5r[10] = row.e5[kOffDt];
#define HVC_OE2(k, off) e2r[k]
#define HVC_OE5(k, off) e5r[k]
#else
#define HVC_OE2(k, off) row.e2[off]
#define HVC_OE5(k, off) row.e5[off]
#endif
// each stage derivative is noted once for all the tableau terms it enters
#undef HVC_CKX
#define HVC_CKX CKN
#define HVC_OSTAGE(N, e2, e5, SX)                                                   \
    const real k_##N = i_oct_rhs<EXACT>(bad, s, q, g, v, n1, m1, h1, w1, x, e2, e5); \
    bad.note_k(k_##N);                                                              \
    x = x0 + SX(k_);                                                                \
    HVC_BCAST8(x)
    const real k_1 = i_oct_rhs<EXACT>(bad, s, q, g, v, n1, m1, h1, w1, x0, HR(1.0), HR(1.0));
    bad.note_k(k_1);
    x = x0 + HVC_DP8_S2(k_);
    HVC_BCAST8(x)
    HVC_OSTAGE(2, HVC_OE2(0, kOffDp2), HVC_OE5(0, kOffDp2), HVC_DP8_S3)
    HVC_OSTAGE(3, HVC_OE2(1, kOffDp3), HVC_OE5(1, kOffDp3), HVC_DP8_S4)
    HVC_OSTAGE(4, HVC_OE2(2, kOffDp4), HVC_OE5(2, kOffDp4), HVC_DP8_S5)
    HVC_OSTAGE(5, HVC_OE2(3, kOffDp5), HVC_OE5(3, kOffDp5), HVC_DP8_S6)
    HVC_OSTAGE(6, HVC_OE2(4, kOffDp6), HVC_OE5(4, kOffDp6), HVC_DP8_S7)
    HVC_OSTAGE(7, HVC_OE2(5, kOffDp7), HVC_OE5(5, kOffDp7), HVC_DP8_S8)
    HVC_OSTAGE(8, HVC_OE2(6, kOffDp8), HVC_OE5(6, kOffDp8), HVC_DP8_S9)
    HVC_OSTAGE(9, HVC_OE2(7, kOffDp9), HVC_OE5(7, kOffDp9), HVC_DP8_S10)
    HVC_OSTAGE(10, HVC_OE2(8, kOffDp10), HVC_OE5(8, kOffDp10), HVC_DP8_S11)
    HVC_OSTAGE(11, HVC_OE2(9, kOffDp11), HVC_OE5(9, kOffDp11), HVC_DP8_S12)
#undef HVC_OSTAGE
    const real k_12 = i_oct_rhs<EXACT>(bad, s, q, g, v, n1, m1, h1, w1, x, HVC_OE2(10, kOffDt), HVC_OE5(10, kOffDt));
    bad.note_k(k_12);
#undef HVC_OE2
#undef HVC_OE5
    {
        struct { real x, x_lo; } acc{x0, xlo};
        HVC_ACC(acc, x, HVC_DP8_FIN(k_))
        x = acc.x;
        xlo = acc.x_lo;
    }
    xo = x;
    HVC_BCAST8(x)
    s.V = v; s.n = n1; s.m = m1; s.h = h1; s.w = w1;
    s.Ge = s.Ge * row_e2dt(row);
    s.Gi = s.Gi * row_e5dt(row);
}
#undef HVC_CKX
#define HVC_CKX CK
#undef HVC_BCAST8

}  // namespace p32 / p64
}  // namespace hvc
