// simuse_oracle.cpp — CPU oracle for the SiMuse per-timestep network step.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
// timed CPU baseline.  Nothing in the product library links or calls this.
//
// Every function cites the reference file:line it restates.  Path prefixes:
//   P/  = crates/SiMuse/crates/pillars/
//   PC/ = crates/SiMuse/crates/pillars-cuda/
//
// Build: see oracle/Makefile (g++ -O3 -ffp-contract=off, no fast-math), which
// keeps float arithmetic unfused exactly like the reference's default non-FMA
// build (P/src/math/fast_exp.rs:32-55, PC/src/engine.rs:762-770).

#include "simuse_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr uint32_t kSlots = 32;  // MAX_DELAY_SLOTS, P/src/api/net/graph.rs:33
constexpr float kClampLo = -100.0f, kClampHi = 100.0f;  // P/src/core/exec/population.rs:69

// ---------------------------------------------------------------------------
// Integer hashing
// ---------------------------------------------------------------------------

// poisson_spike — P/src/api/net/network/sources.rs:5-14 (SplitMix64 finaliser
// over ((seed<<32)|stream) ^ step*phi64, 24-bit uniform compared with p).
static inline bool poisson_spike(uint32_t seed, uint32_t stream, uint64_t step, float p) {
    uint64_t z = (static_cast<uint64_t>(seed) << 32) | static_cast<uint64_t>(stream);
    z ^= step * 0x9E3779B97F4A7C15ull;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    float u = static_cast<float>(z >> 40) / static_cast<float>(1ull << 24);
    return u < p;
}

// SplitMix64 step (Steele, Lea & Flood 2014) used for this build's own
// counter-based Gaussian noise; not a reference function.
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline float bits_to_f(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
static inline uint32_t f_to_bits(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }

// ---------------------------------------------------------------------------
// Deterministic float math built from IEEE +,-,*,/ only, so that the CPU and
// the GPU produce identical bits.  Accuracy is checked against libm in tests.
// ---------------------------------------------------------------------------

// ln(x) for finite normal x > 0: exponent split, m in (sqrt(1/2), sqrt(2)],
// ln(m) = 2*atanh(f/(2+f)) odd series to s^11.
static inline float ln_f32(float x) {
    uint32_t b = f_to_bits(x);
    int e = static_cast<int>((b >> 23) & 0xFFu) - 127;
    float m = bits_to_f((b & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = 0.181818187f;
    p = p * s2 + 0.222222224f;
    p = p * s2 + 0.285714298f;
    p = p * s2 + 0.400000006f;
    p = p * s2 + 0.666666687f;
    p = p * s2 + 2.0f;
    float lnm = s * p;
    float ef = static_cast<float>(e);
    return ef * 0.693145752f + (lnm + ef * 1.42860677e-06f);
}

// exp(x): Cody-Waite reduction x = k*ln2 + r, degree-7 Taylor on |r|<=0.35,
// scale by 2^k through the exponent field.  Results below 2^-126 flush to 0.
static inline float exp_f32(float x) {
    if (!(x == x)) return x;
    if (x > 88.72f) return std::numeric_limits<float>::infinity();
    if (x < -87.33f) return 0.0f;
    float kf = std::floor(x * 1.44269502f + 0.5f);
    float r = (x - kf * 0.693145752f) - kf * 1.42860677e-06f;
    float p = 1.98412701e-04f;
    p = p * r + 1.38888892e-03f;
    p = p * r + 8.33333377e-03f;
    p = p * r + 4.16666679e-02f;
    p = p * r + 0.166666672f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k = static_cast<int>(kf);
    if (k > 127) return p * bits_to_f(0x7F000000u) * 2.0f;
    if (k < -126) return 0.0f;
    return p * bits_to_f(static_cast<uint32_t>(k + 127) << 23);
}

// x^n for n >= 0 by right-to-left binary exponentiation, the expansion of
// f32::powi (compiler-rt __powisf2): r *= a on set bits, a *= a between.
template <class R>
static inline R powi_r(R a, int n) {
    R r = R(1);
    for (;;) {
        if (n & 1) r *= a;
        n /= 2;
        if (n == 0) break;
        a *= a;
    }
    return r;
}

struct TmProfile {
    float u0, a_rec, a_fac;
    bool is_fac;
};

// tm_spike, P/src/api/net/push.rs:60-86: lazy exponential recovery over the
// steps since the edge's last spike (capped at 1e6), then deliver
// weight * R*U with the pre-spike factor, then deplete R (and facilitate U).
template <class R>
static inline R tm_spike_r(const TmProfile& p, R& r, R& u, uint32_t& last, R weight, uint64_t step) {
    uint64_t gap = step > last ? step - last : 0;
    if (gap > 1000000u) gap = 1000000u;
    const int n = static_cast<int>(gap);
    if (n > 0) {
        const R ar = powi_r(R(p.a_rec), n);
        r = R(1) + (r - R(1)) * ar;
        if (p.is_fac) {
            const R af = powi_r(R(p.a_fac), n);
            u = R(p.u0) + (u - R(p.u0)) * af;
        }
    }
    last = static_cast<uint32_t>(step);
    const R scale = r * u;
    r *= R(1) - u;
    if (p.is_fac) u += R(p.u0) * (R(1) - u);
    return weight * scale;
}

// cos(2*pi*u) and sin(2*pi*u), u in [0,1) on a 2^-24 grid: exact quadrant
// split of 4u, then the Taylor series of sin/cos((pi/2) g), g in [0, 1/2],
// through x^9 / x^10, Horner in FP32 (same operation sequence as the device);
// the sine is the cosine of the quadrant before.
static inline void sincos2pi_f32(float u, float& cosv, float& sinv) {
    const float t = u * 4.0f;
    const float q = std::floor(t);
    const float f = t - q;
    const bool hi = f > 0.5f;
    const float g = hi ? 1.0f - f : f;
    const float x = g * 1.57079637f;
    const float x2 = x * x;
    const float sn = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f + x2 * (-0.000198412701f +
                                                                                     x2 * 2.75573188e-06f))));
    const float cs = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f +
                                                                     x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    const float c = hi ? sn : cs;
    const float s = hi ? cs : sn;
    const int qi = static_cast<int>(q);
    cosv = qi == 0 ? c : qi == 1 ? -s : qi == 2 ? -c : s;
    sinv = qi == 0 ? s : qi == 1 ? c : qi == 2 ? -s : -c;
}
static inline float cos2pi_f32(float u) {
    float c, s;
    sincos2pi_f32(u, c, s);
    return c;
}

// Thalamic Gaussian input sigma * N(0,1) per neuron and step: Box-Muller on
// one SplitMix64 hash of (seed, pair, step) with 24-bit u1 in (0,1], u2 in
// [0,1); neurons i and i ^ 1024 share the hash (key i & ~1024): bit 10 clear
// takes sqrt(-2 ln u1) cos(2 pi u2), set takes the sine.
static inline float gauss(uint64_t seed, uint32_t i, uint64_t step) {
    const uint32_t pair = i & ~1024u;
    const uint64_t base = mix64(seed ^ (static_cast<uint64_t>(pair) * 0xD1B54A32D192ED03ull));
    const uint64_t h = mix64(base ^ (step * 0xA24BAED4963EE407ull));
    const float u1 = static_cast<float>((h >> 40) + 1u) * 5.9604644775390625e-08f;
    const float u2 = static_cast<float>((h >> 16) & 0xFFFFFFu) * 5.9604644775390625e-08f;
    const float r = std::sqrt(-2.0f * ln_f32(u1));
    float c, s;
    sincos2pi_f32(u2, c, s);
    return (i & 1024u) ? r * s : r * c;
}

// vtrap(x, y) = x / (exp(x/y) - 1) — PC/kernels/hh.cu:28-34 (NEURON hh.mod).
// Near the removable 0/0 point (|x/y| < 0.25) the FP32 difference exp-1
// cancels; there the value is y * r/(e^r - 1) from its Bernoulli series,
// evaluated in FP64.
static inline float vtrap_f32(float x, float y) {
    float r = x / y;
    if (std::fabs(r) < 0.25f) {
        double rd = static_cast<double>(x) / static_cast<double>(y);
        double r2 = rd * rd;
        double ev = -691.0 / 1307674368000.0;
        ev = ev * r2 + 1.0 / 47900160.0;
        ev = ev * r2 - 1.0 / 1209600.0;
        ev = ev * r2 + 1.0 / 30240.0;
        ev = ev * r2 - 1.0 / 720.0;
        ev = ev * r2 + 1.0 / 12.0;
        double bern = (1.0 - 0.5 * rd) + r2 * ev;
        return static_cast<float>(static_cast<double>(y) * bern);
    }
    return x / (exp_f32(r) - 1.0f);
}
static inline double vtrap_f64(double x, double y) {
    double r = x / y;
    if (r == 0.0) return y;
    return x / std::expm1(r);
}

// ---------------------------------------------------------------------------
// Numeric twin policy
// ---------------------------------------------------------------------------
template <class R> struct Math;
// fast_exp, P/src/math/fast_exp.rs:21-63: exp(x) = 2^n * p(f), t = x*log2(e),
// n = floor(t), f = t - n, p a degree-6 polynomial; the Horner steps are
// rounded separately (the reference's build without the FMA target feature).
static inline float fast_exp_f32(float x) {
    const float t = x * 1.44269502f;
    // floor(t) as i32 (saturating); beyond +-1000 the result is 0 or inf anyway
    float tc = t;
    if (tc > 1000.0f) tc = 1000.0f;
    if (tc < -1000.0f) tc = -1000.0f;
    const int n = tc == tc ? static_cast<int>(std::floor(tc)) : 0;
    const float f = t - static_cast<float>(n);
    const float c1 = 0.693147182f, c2 = 0.2402265f, c3 = 0.05550411f, c4 = 0.009618129f, c5 = 0.001333356f,
                c6 = 0.0001540353f;
    float poly = (c6 * f + c5) * f + c4;
    poly = (poly * f + c3) * f + c2;
    poly = (poly * f + c1) * f + 1.0f;
    if (n < -126) return 0.0f;
    if (n > 127) return INFINITY;
    uint32_t bits = static_cast<uint32_t>(n + 127) << 23;
    float scale;
    std::memcpy(&scale, &bits, 4);
    return scale * poly;
}

template <> struct Math<float> {
    static float exp(float x) { return exp_f32(x); }
    static float vtrap(float x, float y) { return vtrap_f32(x, y); }
    static float fast_exp(float x) { return fast_exp_f32(x); }
    static float ln(float x) { return ln_f32(x); }
    static float mul_add(float a, float b, float c) { return std::fma(a, b, c); }
};
template <> struct Math<double> {
    static double exp(double x) { return std::exp(x); }
    static double vtrap(double x, double y) { return vtrap_f64(x, y); }
    static double fast_exp(double x) { return std::exp(x); }
    static double ln(double x) { return std::log(x); }
    static double mul_add(double a, double b, double c) { return std::fma(a, b, c); }
};

// f32::max / SimdLane::max for non-NaN operands.
template <class R> static inline R maxr(R a, R b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Conductance-based receptor synapses
// ---------------------------------------------------------------------------
// ReceptorType (P/src/kernels/synapse/receptor.rs:26-99): AMPA, NMDA, GABA_A,
// GABA_B, nAChR, GluCl, mAChR with decay time constants (ms) and reversal
// potentials (mV).
static const int kRec = 7;
static const float kRecTauF[kRec] = {2.0f, 50.0f, 10.0f, 200.0f, 3.0f, 5.0f, 150.0f};
static const float kRecErevF[kRec] = {0.0f, 0.0f, -65.0f, -90.0f, 0.0f, -70.0f, -90.0f};

// DecayFactors::new (kinetics.rs:24-38): fast_exp(-dt / tau) per receptor.
template <class R> static inline R rec_decay(R dt, int r) { return Math<R>::fast_exp(-dt / R(kRecTauF[r])); }

// nmda_mg_block (receptor.rs:126-150): 1 / (1 + 0.33 fast_exp(clamp(-0.06 v, -10, 10))).
template <class R> static inline R nmda_block(R v) {
    R x = R(-0.06f) * v;
    x = x < R(-10) ? R(-10) : x;
    x = x > R(10) ? R(10) : x;
    return R(1) / (R(1) + R(0.33f) * Math<R>::fast_exp(x));
}

// SynapticKinetics::compute_currents (kinetics.rs:157-205) for neuron i with
// conductances g[r] (already decayed), potential v, input `base`: neurons
// below n8 = n & ~7 run in 8-wide lanes, base + t_0 + ... + t_6 added in
// turn; the scalar tail adds base + (t_0 + ... + t_6).
template <class R> static inline R rec_current(const R* g, R v, R base, bool lane) {
    const R b = nmda_block<R>(v);
    R t[kRec];
    for (int r = 0; r < kRec; ++r) t[r] = (r == 1 ? g[r] * b : g[r]) * (R(kRecErevF[r]) - v);
    if (lane) {
        R cur = base;
        for (int r = 0; r < kRec; ++r) cur = cur + t[r];
        return cur;
    }
    R acc = R(0);
    for (int r = 0; r < kRec; ++r) acc = acc + t[r];
    return base + acc;
}

// ---------------------------------------------------------------------------
// Neuron models
// ---------------------------------------------------------------------------

// LIFModel::step — P/src/zoo/neurons/lif.rs:47-78.
// params: [0]dt [1]inv_tau_m [2]v_rest [3]v_thresh [4]v_reset [5]tau_ref
// state:  [0]v  [1]refr
template <class R>
static inline int lif_step(R& v, R& refr, R I, R dt, R inv_tau, R v_rest, R v_th,
                           R v_reset, R tau_ref) {
    const R zero = R(0);
    bool is_refr = refr > zero;
    R dv = ((v_rest - v) + I) * inv_tau * dt;
    R v_new = v + dv;
    bool spike_raw = v_new >= v_th;
    bool spike = is_refr ? false : spike_raw;
    R v_final = spike ? v_reset : v_new;
    v = is_refr ? v : v_final;
    R refr_dec = refr - dt;
    R refr_set = spike ? tau_ref : refr_dec;
    refr = maxr(refr_set, zero);
    return spike ? 1 : 0;
}

// IzhikevichModel::step — P/src/zoo/neurons/izhikevich.rs:71-125.
// params: [0]dt [1]a [2]b [3]c [4]d [5]tau_ref [6]v_thresh
// AlphaLifModel::step — P/src/zoo/neurons/alpha_lif.rs:99-136 (the scalar
// runner of pillars-connectome, runner.rs:1511-1600, in SIMD form):
// refractoriness masks only the delayed input, v integrates with the
// previous g through mul_add, strict threshold, a spike resets v and g.
// state: [0]v [1]g [2]refr_cnt; params: [1]time_factor [2]v_rest [3]v_thresh
// [4]v_reset [5]refr_steps [6]syn_decay.  #[no_clamp].
template <class R>
static inline int alpha_lif_step(R& v, R& g, R& refr_cnt, R I, R tf, R v_rest, R v_th, R v_reset, R refr_steps,
                                 R syn_decay) {
    const R g_old = g;
    const R delayed = refr_cnt >= refr_steps ? I : R(0);
    const R integrand = g_old - (v - v_rest);
    const R v_new = std::fma(tf, integrand, v);
    const bool spike = v_new > v_th;
    v = spike ? v_reset : v_new;
    const R g_new = std::fma(syn_decay, g_old, delayed);
    g = spike ? R(0) : g_new;
    refr_cnt = spike ? R(0) : refr_cnt + R(1);
    return spike ? 1 : 0;
}

// state:  [0]v [1]u [2]refr
template <class R>
static inline int izh_step(R& v, R& u, R& refr, R I, R dt, R a, R b, R c, R d,
                           R tau_ref, R v_th) {
    const R zero = R(0);
    bool is_refr = refr > zero;
    R vv = is_refr ? c : v;
    R half_dt = dt * R(0.5);
    const R c04 = R(0.04), c5 = R(5.0), c140 = R(140.0);
    R v_poly1 = vv * (c04 * vv + c5);
    R dv1 = (v_poly1 + c140 - u + I) * half_dt;
    R v_mid = vv + dv1;
    R v_poly2 = v_mid * (c04 * v_mid + c5);
    R dv2 = (v_poly2 + c140 - u + I) * half_dt;
    R dv_total = is_refr ? zero : (dv1 + dv2);
    R v_new = vv + dv_total;
    R du = a * (b * vv - u) * dt;
    R u_new = is_refr ? u : (u + du);
    bool spike = v_new >= v_th;
    v = spike ? c : v_new;
    u = spike ? (u_new + d) : u_new;
    R refr_dec = refr - dt;
    R refr_set = spike ? tau_ref : refr_dec;
    refr = maxr(refr_set, zero);
    return spike ? 1 : 0;
}

// Hodgkin–Huxley step — equations of PC/kernels/hh.cu:107-216 (the CPU HHModel
// at P/src/zoo/neurons/hh.rs:1-51 is a no-op stub, so hh.cu is the only live
// specification).  Column layout follows the stub's field order, with the
// refractory period appended as param 12.
// params: [0]dt [1]g_na [2]g_k [3]g_leak [4]e_na [5]e_k [6]e_leak [7]inv_cm
//         [8]q10 [9]v_thresh [10]g_ca_t [11]e_ca_t [12]tau_ref
// state:  [0]v [1]m [2]h [3]n [4]refr [5]c (T-type Ca gate)
template <class R>
static inline void hh_gate(R& g, R alpha, R beta, R dt, R q10) {
    R sum_q10 = (alpha + beta + R(1e-12)) * q10;       // hh.cu:174
    R m_inf = alpha / (alpha + beta + R(1e-12));        // hh.cu:175
    R decay = Math<R>::exp(-dt * sum_q10);              // hh.cu:176
    g = m_inf + (g - m_inf) * decay;                    // hh.cu:182
}
template <class R>
static inline int hh_step(R* st, const R* pr, R I) {
    using M = Math<R>;
    R dt = pr[0], g_na = pr[1], g_k = pr[2], g_l = pr[3], e_na = pr[4], e_k = pr[5];
    R e_l = pr[6], inv_cm = pr[7], q10 = pr[8], v_th = pr[9], g_cat = pr[10], e_cat = pr[11];
    R tau_ref = pr[12];
    R& v = st[0]; R& m = st[1]; R& h = st[2]; R& n = st[3]; R& refr = st[4]; R& cg = st[5];
    R ref = refr;                                        // hh.cu:142-146
    if (ref > R(0)) refr = maxr(R(0), ref - dt);
    R vi = v;
    // NaV m^3 h (hh.cu:40-53)
    R am = R(0.1) * M::vtrap(-(vi + R(40)), R(10));
    R bm = R(4) * M::exp(-(vi + R(65)) / R(18));
    hh_gate(m, am, bm, dt, q10);
    R ah = R(0.07) * M::exp(-(vi + R(65)) / R(20));
    R bh = R(1) / (R(1) + M::exp(-(vi + R(35)) / R(10)));
    hh_gate(h, ah, bh, dt, q10);
    R gna_ch = g_na;
    gna_ch *= m; gna_ch *= m; gna_ch *= m; gna_ch *= h;
    // KV n^4 (hh.cu:59-64)
    R an = R(0.01) * M::vtrap(-(vi + R(55)), R(10));
    R bn = R(0.125) * M::exp(-(vi + R(65)) / R(80));
    hh_gate(n, an, bn, dt, q10);
    R gk_ch = g_k;
    gk_ch *= n; gk_ch *= n; gk_ch *= n; gk_ch *= n;
    R g_ion = R(0) + gna_ch;
    R i_ion = R(0) + gna_ch * e_na;
    g_ion += gk_ch;
    i_ion += gk_ch * e_k;
    if (g_cat != R(0)) {  // CaV-T (hh.cu:73-80), present only when g_ca_t != 0
        R minf = R(1) / (R(1) + M::exp(-(vi + R(56)) / R(6.2)));
        R ac = minf / R(0.6);
        R bc = (R(1) - minf) / R(0.6);
        hh_gate(cg, ac, bc, dt, q10);
        R gc = g_cat * cg;
        g_ion += gc;
        i_ion += gc * e_cat;
    }
    // Exponential-Euler membrane (hh.cu:194-204), tau_v = C_m/g_tot.
    R g_tot = g_ion + g_l + R(1e-12);
    R v_inf = (I + i_ion + g_l * e_l) / g_tot;
    R decay_v = M::exp(-dt * (g_tot * inv_cm));
    R v_prev = vi;
    vi = v_inf + (v_prev - v_inf) * decay_v;
    int spiked = 0;                                      // hh.cu:206-213
    if (ref <= R(0) && v_prev < v_th && vi >= v_th) {
        spiked = 1;
        refr = tau_ref;
    }
    v = vi;
    return spiked;
}

// GpeaModel::step, P/src/zoo/neurons/gpea.rs:266-360 (port of pGPeA_k.mod):
// gates by exponential Euler at the old V (decay.mul_add(x - x_inf, x_inf),
// fused), r from the old cai, Ca by cnexp with I_Ca frozen from the old state,
// Nernst E_Na, E_K, E_Ca through ln, membrane by implicit Euler, rising-edge
// spike through v_thresh with the was_below flag.  #[no_clamp].
// params: [0]dt [1]celsius [2]inv_cm [3]v_thresh [4]area [5]gnabar [6]gkdrbar
//         [7]gl [8]el [9]gcatbar [10]gkcabar [11]kca [12]vol [13]caGain
// state:  [0]v [1]m [2]h [3]n [4]p [5]q [6]cai [7]r [8]was_below
template <class R>
struct Gpea {
    using M = Math<R>;
    // sigmoid(x, th, k) = 1/(1 + fast_exp((x - th) * (1/k)))   gpea.rs:130-133
    static R sigmoid(R x, float th, float k) {
        return R(1) / (R(1) + M::fast_exp((x - R(th)) * R(1.0f / k)));
    }
    // tau = t0 + t1/(fast_exp((v-tht1)*(-1/sig1)) + fast_exp(-(v-tht2)*(1/sig2)))   gpea.rs:148-180
    static R tau2(R v, float t0, float t1, float tht1, float sig1, float tht2, float sig2) {
        const R e1 = M::fast_exp((v - R(tht1)) * R(-1.0f / sig1));
        const R e2 = M::fast_exp(-(v - R(tht2)) * R(1.0f / sig2));
        return R(t0) + R(t1) / (e1 + e2);
    }
    static R tau_m(R v) {  // gpea.rs:141-145
        const R den = R(1) + M::fast_exp((v - R(-53.0f)) * R(-1.0f / -0.7f));
        return R(0.001f) + R(0.1f) / den;
    }
    static R tau_h(R v) { return tau2(v, 0.0f, 4.5f, -50.0f, -15.0f, -50.0f, 16.0f); }
    static R tau_n(R v) { return tau2(v, 0.0f, 2.4f, -40.0f, -40.0f, -40.0f, 50.0f); }
    static R tau_p(R v) { return tau2(v, 5.0f, 0.33f, -27.0f, -10.0f, -102.0f, 15.0f); }
    static R tau_q(R v) { return tau2(v, 0.0f, 400.0f, -50.0f, -15.0f, -50.0f, 16.0f); }
    static R m_inf(R v) { return sigmoid(v, -38.0f, -7.0f); }
    static R h_inf(R v) { return sigmoid(v, -45.5f, 6.4f); }
    static R n_inf(R v) { return sigmoid(v, -42.0f, -14.0f); }
    static R p_inf(R v) { return sigmoid(v, -56.0f, -6.7f); }
    static R q_inf(R v) { return sigmoid(v, -85.0f, 5.8f); }
    static R r_inf(R cai) { return sigmoid(cai, 0.17e-3f, -0.08e-3f); }
    static R exp_euler(R x, R x_inf, R tau, R dt) {  // gpea.rs:136-139
        const R decay = M::fast_exp(-dt / tau);
        return M::mul_add(decay, x - x_inf, x_inf);
    }
    static int step(R* st, const R* pr, R I) {
        const R dt = pr[0], celsius = pr[1], inv_cm = pr[2], v_th = pr[3], area = pr[4];
        const R gnabar = pr[5], gkdrbar = pr[6], gl = pr[7], el = pr[8], gcatbar = pr[9];
        const R gkcabar = pr[10], kca = pr[11], vol = pr[12], ca_gain = pr[13];
        const R F = R(96485.0f);
        const R rt = R(8.31441f) * ((R(273.0f) + celsius) - R(9.5f)) / F * R(1000.0f);
        const R ena = -rt * M::ln(R(10.0f) / R(140.0f));
        const R ek = rt * M::ln(R(5.0f) / R(140.0f));
        const R v = st[0];
        const R m_new = exp_euler(st[1], m_inf(v), tau_m(v), dt);
        const R h_new = exp_euler(st[2], h_inf(v), tau_h(v), dt);
        const R n_new = exp_euler(st[3], n_inf(v), tau_n(v), dt);
        const R p_new = exp_euler(st[4], p_inf(v), tau_p(v), dt);
        const R q_new = exp_euler(st[5], q_inf(v), tau_q(v), dt);
        const R cai = st[6];
        const R r_new = exp_euler(st[7], r_inf(cai), R(2.0f), dt);
        // Ca (cnexp): I_Ca from the old state
        const R eca_old = -rt * M::ln(cai / R(2.0f)) / R(2.0f);
        const R gcat_old = gcatbar * st[4] * st[4] * st[5];
        const R ica_old = gcat_old * (v - eca_old);
        const R j = -ica_old * area * R(1e-14f) / (R(2.0f) * F * vol);
        const R k_ca = ca_gain * kca;
        const R cai_decay = M::fast_exp(-dt * k_ca);
        const R cai_new = cai + (R(1) - cai_decay) * (ca_gain * j / k_ca - cai);
        const R eca = -rt * M::ln(cai_new / R(2.0f)) / R(2.0f);
        const R g_na = gnabar * (m_new * m_new * m_new) * h_new;
        const R n2 = n_new * n_new;
        const R g_k = gkdrbar * (n2 * n2);
        const R g_cat = gcatbar * p_new * p_new * q_new;
        const R g_ahp = gkcabar * r_new * r_new;
        const R g_sum = g_na + g_k + g_cat + g_ahp + gl;
        const R g_e_sum = g_na * ena + (g_k + g_ahp) * ek + g_cat * eca + gl * el;
        const R c_over_dt = R(1) / (inv_cm * dt);
        const R v_new = (c_over_dt * v + I + g_e_sum) / (c_over_dt + g_sum);
        // edge_spike, P/src/ops/spike.rs:70-74
        const bool above = v_new >= v_th;
        const int spike = (above && st[8] > R(0)) ? 1 : 0;
        st[0] = v_new; st[1] = m_new; st[2] = h_new; st[3] = n_new; st[4] = p_new; st[5] = q_new;
        st[6] = cai_new; st[7] = r_new; st[8] = above ? R(0) : R(1);
        return spike;
    }
};

static inline int n_state(int model) {
    return model == ORC_MODEL_LIF ? 2 : model == ORC_MODEL_IZHIKEVICH ? 3 : model == ORC_MODEL_HH ? 6
         : model == ORC_MODEL_ALPHA_LIF ? 3 : 9;
}
static inline int n_param(int model) {
    return model == ORC_MODEL_LIF ? 6 : model == ORC_MODEL_IZHIKEVICH ? 7 : model == ORC_MODEL_HH ? 13
         : model == ORC_MODEL_ALPHA_LIF ? 7 : 14;
}
static inline bool clamps(int model) {  // #[no_clamp] on HH (hh.rs:13) and GPEA (gpea.rs:213)
    return model == ORC_MODEL_LIF || model == ORC_MODEL_IZHIKEVICH;
}

// ---------------------------------------------------------------------------
// Network
// ---------------------------------------------------------------------------

struct NetBase {
    virtual ~NetBase() {}
    virtual int add_source(float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts,
                           const float* w, const uint8_t* d, uint32_t count) = 0;
    virtual int set_noise(int pop, const float* sigma, uint64_t seed) = 0;
    virtual int set_plasticity(const orc_stdp_params* p, const uint8_t* plastic) = 0;
    virtual int set_stp(const orc_stp_profile* profiles, int n_profiles, float dt, const uint8_t* syn_types,
                        const uint8_t* src_types, uint64_t n_src_edges) = 0;
    virtual int inject(int pop, const uint32_t* idx, const float* val, uint32_t count) = 0;
    virtual int inject_voltage(int pop, const uint32_t* idx, const float* val, uint32_t count) = 0;
    // electrical coupling, global neuron indices: kind 0 gap junction, 1 graded synapse
    virtual void add_coupling(int kind, uint32_t a, uint32_t b, float g) = 0;
    virtual int add_sigmoid_profile(float vth, float delta, float k, float erev, float cond, float dt) = 0;
    virtual int add_sigmoid_conn(int prof, uint32_t pre, uint32_t post, float weight) = 0;
    virtual int set_receptors(const uint8_t* types, float scale, const float* ext) = 0;
    virtual void conductances(double* out) const = 0;
    // Overwrite the dynamic state with a (device) snapshot: FP32 values are
    // exact in either twin.  Null pointers leave that part as it is.
    virtual int set_state(uint64_t step, const float* const* pop_states, const float* ring, const float* w,
                          const double* w64, const float* tr_pre, const float* tr_post, const float* g) = 0;
    virtual void step(uint32_t n) = 0;
    virtual void stdp_step(float decay, float lr_ltp, float lr_ltd, float w_min, float w_max) = 0;
    virtual uint64_t step_counter() const = 0;
    virtual void spikes(int32_t* out) const = 0;
    virtual uint32_t active(uint32_t* out) const = 0;
    virtual int state(int pop, int col, double* out) const = 0;
    virtual void ring(double* out) const = 0;
    virtual void weights(double* out) const = 0;
    virtual uint64_t total_spikes() const = 0;
    virtual uint64_t total_events() const = 0;
    virtual void record(bool on) = 0;
    virtual uint64_t raster_steps() const = 0;
    virtual void raster(int32_t* out) const = 0;
    int n_threads = 1;
};

template <class R>
struct Net final : NetBase {
    struct Pop {
        int model;
        uint32_t n, off, cap;
        std::vector<R> st, pr;  // [col*cap + i]
        bool noise = false;
        uint64_t seed = 0;
        std::vector<float> sigma;
    };
    struct Src {
        float p;
        uint32_t seed;
        std::vector<uint32_t> post;
        std::vector<float> w;
        std::vector<uint8_t> d;
    };
    std::vector<Pop> pops;
    uint32_t N = 0;
    std::vector<uint32_t> row_ptr, col;
    std::vector<R> w;            // per-edge weight in the twin's precision
    std::vector<uint8_t> d;
    std::vector<R> ringv;        // [32][N]
    std::vector<R> i_in;
    std::vector<int32_t> spk;
    std::vector<uint32_t> act;
    std::vector<Src> srcs;
    uint64_t stepc = 0, n_spikes = 0, n_events = 0;
    bool rec = false;
    std::vector<int32_t> rast;
    uint64_t rast_steps = 0;
    // pair-based STDP
    bool stdp = false;
    orc_stdp_params sp{};
    std::vector<uint8_t> plastic;
    std::vector<R> tr_pre, tr_post;
    std::vector<uint32_t> in_ptr, in_edge, edge_pre;
    std::vector<double> w64;  // FP64 master weights while plasticity is on
    // Tsodyks-Markram state per edge: synapses first, then source edges in
    // add order (Network::build, P/src/api/net/network/build.rs:70-84)
    std::vector<TmProfile> tm_prof;
    std::vector<uint8_t> tm_type;
    std::vector<R> tm_r, tm_u;
    std::vector<uint32_t> tm_last;
    // dense every-synapse STDP of CompiledNetwork::stdp_step (one shared trace)
    std::vector<R> dtrace;
    bool dense = false;
    // receptor synapses: per synapse 0 current-based or 1 + receptor type;
    // conductances [7][N]; constant external current
    bool rec_on = false;
    // Electrical coupling in the reference's list orders (step.rs:36-104,
    // core.rs:391-553): gap junctions, one-way graded synapses, and sigmoid
    // graded synapses per registered profile (sources.rs:91-130).
    struct Cpl {
        uint32_t a, b;
        R g;
    };
    std::vector<Cpl> gapj, graded;
    struct SigRt {
        R vth, inv_delta, k, erev, cond, dt;
        std::vector<uint32_t> pre, post;
        std::vector<R> w, s;
    };
    std::vector<SigRt> sig;
    std::vector<R> vscr;
    std::vector<uint8_t> rec_code;
    std::vector<R> gcond;
    std::vector<float> g_ext;
    float rec_scale = 1.0f;
    float d_lr_ltp = 0, d_lr_ltd = 0, d_wmin = 0, d_wmax = 0;

    int init(const orc_pop_desc* pd, int np_, const uint32_t* rp, const uint32_t* c,
             const float* wf, const uint8_t* dl, uint64_t n_syn) {
        uint32_t off = 0;
        for (int p = 0; p < np_; ++p) {
            Pop P;
            P.model = pd[p].model;
            P.n = pd[p].n;
            P.off = off;
            P.cap = (pd[p].n + 7u) / 8u * 8u;
            int ns = n_state(P.model), npar = n_param(P.model);
            if (pd[p].ns != static_cast<uint32_t>(ns) || pd[p].np != static_cast<uint32_t>(npar)) return -2;
            uint32_t hs = pd[p].stride;
            P.st.assign(static_cast<size_t>(ns) * P.cap, R(0));
            P.pr.assign(static_cast<size_t>(npar) * P.cap, R(0));
            for (int k = 0; k < ns; ++k)
                for (uint32_t i = 0; i < P.n; ++i)
                    P.st[static_cast<size_t>(k) * P.cap + i] = R(pd[p].states[static_cast<size_t>(k) * hs + i]);
            for (int k = 0; k < npar; ++k)
                for (uint32_t i = 0; i < P.n; ++i)
                    P.pr[static_cast<size_t>(k) * P.cap + i] = R(pd[p].params[static_cast<size_t>(k) * hs + i]);
            off += P.n;
            pops.push_back(std::move(P));
        }
        N = off;
        row_ptr.assign(rp, rp + N + 1);
        col.assign(c, c + n_syn);
        w.resize(n_syn);
        for (uint64_t e = 0; e < n_syn; ++e) w[e] = R(wf[e]);
        d.assign(dl, dl + n_syn);
        ringv.assign(static_cast<size_t>(kSlots) * N, R(0));
        i_in.assign(N, R(0));
        spk.assign(N, 0);
        return 0;
    }

    int add_source(float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts,
                   const float* wf, const uint8_t* dl, uint32_t count) override {
        // Network::add_poisson_source — P/src/api/net/network/core.rs:288-318
        Src s;
        s.p = std::min(std::max(rate_hz * dt_ms / 1000.0f, 0.0f), 1.0f);
        s.seed = seed;
        for (uint32_t i = 0; i < count; ++i) {
            if (posts[i] >= N || dl[i] >= kSlots) return -3;
            s.post.push_back(posts[i]);
            s.w.push_back(wf[i]);
            s.d.push_back(dl[i]);
        }
        srcs.push_back(std::move(s));
        return 0;
    }
    int set_noise(int p, const float* sigma, uint64_t seed) override {
        if (p < 0 || p >= static_cast<int>(pops.size())) return -3;
        pops[p].noise = true;
        pops[p].seed = seed;
        pops[p].sigma.assign(sigma, sigma + pops[p].n);
        return 0;
    }
    int set_plasticity(const orc_stdp_params* p, const uint8_t* pl) override {
        stdp = true;
        sp = *p;
        size_t ns = col.size();
        plastic.assign(pl, pl + ns);
        w64.resize(ns);
        for (size_t e = 0; e < ns; ++e) w64[e] = static_cast<double>(w[e]);
        tr_pre.assign(N, R(0));
        tr_post.assign(N, R(0));
        edge_pre.resize(ns);
        for (uint32_t i = 0; i < N; ++i)
            for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) edge_pre[e] = i;
        // in-edge lists: for each post, its in-edges in ascending edge order
        in_ptr.assign(N + 1, 0);
        for (size_t e = 0; e < ns; ++e) in_ptr[col[e] + 1]++;
        for (uint32_t i = 0; i < N; ++i) in_ptr[i + 1] += in_ptr[i];
        in_edge.assign(ns, 0);
        std::vector<uint32_t> cur(in_ptr.begin(), in_ptr.end() - 1);
        for (size_t e = 0; e < ns; ++e) in_edge[cur[col[e]]++] = static_cast<uint32_t>(e);
        return 0;
    }
    int set_stp(const orc_stp_profile* profiles, int n_profiles, float dt, const uint8_t* syn_types,
                const uint8_t* src_types, uint64_t n_src_edges) override {
        // Network::register_plasticity, P/src/api/net/network/core.rs:597-619
        tm_prof.clear();
        for (int k = 0; k < n_profiles; ++k) {
            TmProfile t;
            t.u0 = profiles[k].u0;
            t.a_rec = std::exp(-dt / profiles[k].tau_rec);
            t.is_fac = profiles[k].tau_fac > 0.0f;
            t.a_fac = t.is_fac ? std::exp(-dt / profiles[k].tau_fac) : 0.0f;
            tm_prof.push_back(t);
        }
        uint64_t n_src = 0;
        for (auto& s : srcs) n_src += s.post.size();
        if (src_types && n_src_edges != n_src) return -3;
        const size_t ne = col.size() + n_src;
        tm_type.assign(ne, 0);
        for (size_t e = 0; e < col.size(); ++e) tm_type[e] = syn_types ? syn_types[e] : 0;
        if (src_types)
            for (uint64_t e = 0; e < n_src; ++e) tm_type[col.size() + e] = src_types[e];
        for (uint8_t t : tm_type)
            if (t > tm_prof.size()) return -3;
        tm_r.assign(ne, R(1));
        tm_u.assign(ne, R(0));
        for (size_t e = 0; e < ne; ++e)
            if (tm_type[e]) tm_u[e] = R(tm_prof[tm_type[e] - 1].u0);
        tm_last.assign(ne, 0u);
        return 0;
    }
    R delivered(size_t e, R weight) {
        if (tm_type.empty() || tm_type[e] == 0) return weight;
        return tm_spike_r(tm_prof[tm_type[e] - 1], tm_r[e], tm_u[e], tm_last[e], weight, stepc);
    }
    int set_receptors(const uint8_t* types, float scale, const float* ext) override {
        // ReceptorType::from_u8 (receptor.rs:64-76): 0..6, anything else AMPA;
        // 255 keeps a synapse current-based
        rec_code.assign(col.size(), 0);
        for (size_t e = 0; e < col.size(); ++e) {
            const uint8_t r = types[e];
            rec_code[e] = r == 255 ? 0 : r < kRec ? static_cast<uint8_t>(r + 1) : 1;
        }
        rec_on = true;
        rec_scale = scale;
        gcond.assign(static_cast<size_t>(kRec) * N, R(0));
        g_ext.clear();
        if (ext) g_ext.assign(ext, ext + N);
        return 0;
    }
    void conductances(double* out) const override {
        for (size_t k = 0; k < gcond.size(); ++k) out[k] = static_cast<double>(gcond[k]);
    }
    int set_state(uint64_t step, const float* const* pop_states, const float* ring, const float* wf,
                  const double* w64v, const float* pre, const float* post, const float* g) override {
        stepc = step;
        if (pop_states)
            for (size_t p = 0; p < pops.size(); ++p) {
                Pop& P = pops[p];
                const int ns = n_state(P.model);
                for (int c = 0; c < ns; ++c)
                    for (uint32_t i = 0; i < P.n; ++i)
                        P.st[static_cast<size_t>(c) * P.cap + i] = R(pop_states[p][static_cast<size_t>(c) * P.n + i]);
            }
        if (ring)
            for (size_t k = 0; k < ringv.size(); ++k) ringv[k] = R(ring[k]);
        if (wf)
            for (size_t e = 0; e < w.size(); ++e) w[e] = R(wf[e]);
        if (w64v && stdp) {
            for (size_t e = 0; e < w64.size(); ++e) w64[e] = w64v[e];
            for (size_t e = 0; e < w.size(); ++e) w[e] = R(w64v[e]);
        }
        if (pre && stdp)
            for (uint32_t i = 0; i < N; ++i) tr_pre[i] = R(pre[i]);
        if (post && stdp)
            for (uint32_t i = 0; i < N; ++i) tr_post[i] = R(post[i]);
        if (g && rec_on)
            for (size_t k = 0; k < gcond.size(); ++k) gcond[k] = R(g[k]);
        return 0;
    }
    // Before the update (CN/src/sim/hh.rs:115-166): decay every conductance,
    // then the input becomes (input + external) plus the receptor currents
    // of the pre-update potential.
    void receptor_inputs() {
        const uint32_t n8 = N & ~7u;
        for (const Pop& P : pops) {
            R f[kRec];
            for (int r = 0; r < kRec; ++r) f[r] = rec_decay<R>(P.pr[0], r);
            for (uint32_t i = 0; i < P.n; ++i) {
                const uint32_t gi = P.off + i;
                R g[kRec];
                for (int r = 0; r < kRec; ++r) {
                    R& x = gcond[static_cast<size_t>(r) * N + gi];
                    x = x * f[r];
                    g[r] = x;
                }
                const R base = g_ext.empty() ? i_in[gi] : i_in[gi] + R(g_ext[gi]);
                i_in[gi] = rec_current<R>(g, P.st[i], base, gi < n8);  // state column 0 = v
            }
        }
    }
    int inject(int p, const uint32_t* idx, const float* val, uint32_t count) override {
        // Network::inject_current — P/src/api/net/network/build.rs:243-275
        if (p < 0 || p >= static_cast<int>(pops.size())) return -3;
        uint32_t slot = static_cast<uint32_t>(stepc % kSlots);
        for (uint32_t i = 0; i < count; ++i) {
            if (idx[i] >= pops[p].n) return -3;
            ringv[static_cast<size_t>(slot) * N + pops[p].off + idx[i]] += R(val[i]);
        }
        return 0;
    }

    void add_coupling(int kind, uint32_t a, uint32_t b, float g) override {
        (kind == 0 ? gapj : graded).push_back(Cpl{a, b, R(g)});
    }
    int add_sigmoid_profile(float vth, float delta, float k, float erev, float cond, float dt) override {
        SigRt r;
        r.vth = R(vth);
        r.inv_delta = R(1) / R(delta);  // SigmoidGradedRuntime::new, sources.rs:103-113
        r.k = R(k);
        r.erev = R(erev);
        r.cond = R(cond);
        r.dt = R(dt);
        sig.push_back(r);
        return static_cast<int>(sig.size() - 1);
    }
    int add_sigmoid_conn(int prof, uint32_t pre, uint32_t post, float weight) override {
        if (prof < 0 || prof >= static_cast<int>(sig.size())) return -3;
        SigRt& r = sig[prof];
        r.pre.push_back(pre);
        r.post.push_back(post);
        r.w.push_back(R(weight));
        r.s.push_back(R(0));
        return 0;
    }
    bool coupled() const { return !gapj.empty() || !graded.empty() || !sig.empty(); }
    // step.rs:36-104: V at the start of the step, then the three couplings
    // add into i_in in list order
    void coupling() {
        vscr.resize(N);
        for (auto& P : pops)
            for (uint32_t i = 0; i < P.n; ++i) vscr[P.off + i] = P.st[i];
        for (const Cpl& c : gapj) {
            const R dv = vscr[c.b] - vscr[c.a];
            i_in[c.a] = i_in[c.a] + c.g * dv;
            i_in[c.b] = i_in[c.b] - c.g * dv;
        }
        for (const Cpl& c : graded) {
            const R dv = vscr[c.a] - vscr[c.b];
            i_in[c.b] = i_in[c.b] + c.g * dv;
        }
        for (SigRt& r : sig)
            for (size_t q = 0; q < r.pre.size(); ++q) {  // SigmoidGradedRuntime::step, sources.rs:117-129
                const R vpre = vscr[r.pre[q]], vpost = vscr[r.post[q]];
                const R arg = (vpre - r.vth) * r.inv_delta;
                const R inf = R(1) / (R(1) + Math<R>::exp(-arg));
                const R tau = R(1) / (r.k * (R(1) + Math<R>::exp(arg)));
                const R decay = Math<R>::exp(-r.dt / tau);
                r.s[q] = r.s[q] + (inf - r.s[q]) * (R(1) - decay);
                const R current = r.w[q] * r.cond * r.s[q] * (r.erev - vpost);
                i_in[r.post[q]] = i_in[r.post[q]] + current;
            }
    }

    int inject_voltage(int p, const uint32_t* idx, const float* val, uint32_t count) override {
        // WgpuNetwork::inject_voltage + v_jump_kernel, PW/src/network.rs:1632-1659,
        // PW/shaders/csr_scatter.wgsl:117-124: v += value before the next step
        if (p < 0 || p >= static_cast<int>(pops.size())) return -3;
        Pop& P = pops[p];
        if (!idx && count != P.n) return -3;
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t i = idx ? idx[k] : k;
            if (i >= P.n) return -3;
            P.st[i] = P.st[i] + R(val[k]);
        }
        return 0;
    }

    void step_pop(Pop& P) {
        const uint32_t cap = P.cap;
        // dt broadcast from param column 0 of neuron 0 (population.rs:307,361)
        const R dt = P.pr[0];
        const bool cl = clamps(P.model);
        const int64_t n = P.n;
        const int nst = n_state(P.model), npar = n_param(P.model);
#pragma omp parallel for schedule(static) num_threads(n_threads) if (n_threads > 1 && n > 4096)
        for (int64_t ii = 0; ii < n; ++ii) {
            uint32_t i = static_cast<uint32_t>(ii);
            R I = i_in[P.off + i];
            if (cl) I = std::min(std::max(I, R(kClampLo)), R(kClampHi));  // population.rs:79-88
            R* s = P.st.data();
            const R* pr = P.pr.data();
            int spike;
            if (P.model == ORC_MODEL_ALPHA_LIF) {
                spike = alpha_lif_step<R>(s[i], s[cap + i], s[2 * cap + i], I, pr[cap + i], pr[2 * cap + i],
                                          pr[3 * cap + i], pr[4 * cap + i], pr[5 * cap + i], pr[6 * cap + i]);
            } else if (P.model == ORC_MODEL_LIF) {
                spike = lif_step<R>(s[i], s[cap + i], I, dt, pr[cap + i], pr[2 * cap + i],
                                    pr[3 * cap + i], pr[4 * cap + i], pr[5 * cap + i]);
            } else if (P.model == ORC_MODEL_IZHIKEVICH) {
                spike = izh_step<R>(s[i], s[cap + i], s[2 * cap + i], I, dt, pr[cap + i],
                                    pr[2 * cap + i], pr[3 * cap + i], pr[4 * cap + i],
                                    pr[5 * cap + i], pr[6 * cap + i]);
            } else {
                R stv[9], prv[14];
                for (int k = 0; k < nst; ++k) stv[k] = s[static_cast<size_t>(k) * cap + i];
                for (int k = 0; k < npar; ++k) prv[k] = pr[static_cast<size_t>(k) * cap + i];
                prv[0] = dt;
                spike = P.model == ORC_MODEL_HH ? hh_step<R>(stv, prv, I) : Gpea<R>::step(stv, prv, I);
                for (int k = 0; k < nst; ++k) s[static_cast<size_t>(k) * cap + i] = stv[k];
            }
            spk[P.off + i] = spike;
        }
    }

    // Pair-based STDP applied on the CSR, event-driven: rule of
    // P/src/ops/plasticity/stdp.rs:34-75 (traces decay by e^{-dt/tau} and add
    // the spike; LTP = A+ * pre-trace at a post spike, LTD = A- * post-trace at
    // a pre spike) with the spike-triggered in-edge/out-edge event scheme of
    // P/src/api/learn/brain.rs:204-286 (traces first, fused decay+add).
    // LTP pass (in-edges of spiking posts) then LTD pass (out-edges of
    // spiking pres), each clipped to [w_min, w_max].  The per-step decay
    // factors are e^{-dt/tau} in the twin's own precision.
    void stdp_update(R dt) {
        const R dpre = R(std::exp(-static_cast<double>(dt) / static_cast<double>(sp.tau_plus)));
        const R dpost = R(std::exp(-static_cast<double>(dt) / static_cast<double>(sp.tau_minus)));
        for (uint32_t i = 0; i < N; ++i) {
            R add = spk[i] ? R(1) : R(0);
            tr_pre[i] = std::fma(tr_pre[i], dpre, add);
            tr_post[i] = std::fma(tr_post[i], dpost, add);
        }
        // Weights accumulate in FP64 (w64) and are mirrored to the twin's
        // precision; the increment itself is formed in the twin's precision.
        const R ap = R(sp.a_plus), am = R(sp.a_minus);
        const double wlo = static_cast<double>(sp.w_min), whi = static_cast<double>(sp.w_max);
        for (uint32_t post : act)
            for (uint32_t k = in_ptr[post]; k < in_ptr[post + 1]; ++k) {
                uint32_t e = in_edge[k];
                if (!plastic[e]) continue;
                const R dw = ap * tr_pre[edge_pre[e]];
                const double nw = std::min(std::max(w64[e] + static_cast<double>(dw), wlo), whi);
                w64[e] = nw;
                w[e] = R(nw);
            }
        for (uint32_t pre : act)
            for (uint32_t e = row_ptr[pre]; e < row_ptr[pre + 1]; ++e) {
                if (!plastic[e]) continue;
                const R dw = am * tr_post[col[e]];
                const double nw = std::min(std::max(w64[e] - static_cast<double>(dw), wlo), whi);
                w64[e] = nw;
                w[e] = R(nw);
            }
    }

    void step_once() {
        // Network::step — P/src/api/net/network/step.rs:18-189
        if (N == 0) { ++stepc; return; }
        const uint32_t slot = static_cast<uint32_t>(stepc % kSlots);
        R* cur = ringv.data() + static_cast<size_t>(slot) * N;
        // Gaussian current noise enters the current slot like inject_current
        // (build.rs:243-275) issued just before the step.
        for (auto& P : pops)
            if (P.noise) {
                const int64_t pn = P.n;
                R* c = cur + P.off;
#pragma omp parallel for schedule(static) num_threads(n_threads) if (n_threads > 1 && pn > 4096)
                for (int64_t ii = 0; ii < pn; ++ii) {
                    const uint32_t i = static_cast<uint32_t>(ii);
                    c[i] += R(P.sigma[i] * gauss(P.seed, i, stepc));
                }
            }
        // 1-2: read slot -> i_in, clear slot (delay.rs:68-77)
        std::memcpy(i_in.data(), cur, sizeof(R) * N);
        std::fill(cur, cur + N, R(0));
        if (coupled()) coupling();
        if (rec_on) receptor_inputs();
        // 3: populations (population.rs:284-420)
        for (auto& P : pops) step_pop(P);
        // 4: compaction, ascending (compact.rs:30-61)
        act.clear();
        for (uint32_t i = 0; i < N; ++i)
            if (spk[i] != 0) act.push_back(i);
        n_spikes += act.size();
        if (rec) {
            rast.insert(rast.end(), spk.begin(), spk.end());
            ++rast_steps;
        }
        if (stdp) stdp_update(pops.empty() ? R(0) : pops[0].pr[0]);
        if (dense) dense_update();
        // 5: push, pre ascending then CSR row order (push.rs:126-173); the
        //    rayon variant (push.rs:184-237) applies triples in the same order.
        if (push_par()) {
            n_events += push_spikes_par();
        } else
        for (uint32_t pre : act) {
            for (uint32_t e = row_ptr[pre]; e < row_ptr[pre + 1]; ++e) {
                if (rec_on && rec_code[e]) {
                    // conductance jump, in push order (CN/src/propagate.rs:225-290)
                    R& x = gcond[static_cast<size_t>(rec_code[e] - 1) * N + col[e]];
                    x = x + delivered(e, w[e]) * R(rec_scale);
                    continue;
                }
                uint32_t dd = d[e] < 1 ? 1u : d[e];
                uint32_t ts = static_cast<uint32_t>((stepc + dd) % kSlots);
                ringv[static_cast<size_t>(ts) * N + col[e]] += delivered(e, w[e]);
            }
            n_events += row_ptr[pre + 1] - row_ptr[pre];
        }
        // 5b: Poisson sources (step.rs:166-185)
        size_t src_edge = col.size();
        for (auto& s : srcs) {
            const size_t first = src_edge;
            src_edge += s.post.size();
            if (s.p <= 0.0f || !poisson_spike(s.seed, 0, stepc, s.p)) continue;
            for (size_t i = 0; i < s.post.size(); ++i) {
                uint32_t dd = s.d[i] < 1 ? 1u : s.d[i];
                uint32_t ts = static_cast<uint32_t>((stepc + dd) % kSlots);
                ringv[static_cast<size_t>(ts) * N + s.post[i]] += delivered(first + i, R(s.w[i]));
            }
            n_events += s.post.size();
        }
        ++stepc;  // 6
    }
    void step(uint32_t n) override { for (uint32_t k = 0; k < n; ++k) step_once(); }

    // push_spikes_par, P/src/api/net/push.rs:184-237: without short-term
    // plasticity (or conductance jumps), with more than one thread and more
    // than 64 active neurons, the active list is cut into one chunk per thread
    // (chunk = max(ceil(active / threads), 8)); each thread collects
    // (post, slot, weight) triples of its chunk's rows in CSR order
    // (collect_push_triples, :246-269), then one thread adds them into the
    // ring in chunk order -- the sequential push's order, so the ring is
    // bit-identical to push_spikes.
    struct Triple {
        uint32_t post, slot;
        R w;
    };
    std::vector<std::vector<Triple>> trip;
    bool push_par() const {
        return n_threads > 1 && act.size() > 64 && tm_type.empty() && !rec_on;
    }
    uint64_t push_spikes_par() {
        const size_t na = act.size(), nt = static_cast<size_t>(n_threads);
        const size_t chunk = std::max<size_t>((na + nt - 1) / nt, 8);
        if (trip.size() < nt) trip.resize(nt);
        const uint64_t sc = stepc;
#pragma omp parallel for schedule(static, 1) num_threads(n_threads)
        for (int64_t tid = 0; tid < static_cast<int64_t>(nt); ++tid) {
            std::vector<Triple>& v = trip[tid];
            v.clear();
            const size_t a0 = static_cast<size_t>(tid) * chunk, a1 = std::min(a0 + chunk, na);
            for (size_t a = a0; a < a1; ++a) {
                const uint32_t pre = act[a];
                for (uint32_t e = row_ptr[pre]; e < row_ptr[pre + 1]; ++e) {
                    const uint32_t dd = d[e] < 1 ? 1u : d[e];
                    v.push_back(Triple{col[e], static_cast<uint32_t>((sc + dd) % kSlots), w[e]});
                }
            }
        }
        uint64_t events = 0;
        for (size_t tid = 0; tid < nt; ++tid) {
            for (const Triple& x : trip[tid]) ringv[static_cast<size_t>(x.slot) * N + x.post] += x.w;
            events += trip[tid].size();
        }
        return events;
    }

    // CompiledNetwork::stdp_step, PC/src/network.rs:495-718 with the kernels
    // of :53-83: trace = trace*decay + (spike of the previous step), then the
    // step itself with every synapse updated between the population update
    // and the push, w = min(w_max, max(w_min, (w + lr_ltp*tr[pre]) -
    // lr_ltd*tr[post])), each operation rounded separately.
    void stdp_step(float decay, float lr_ltp, float lr_ltd, float w_min, float w_max) override {
        if (dtrace.empty()) dtrace.assign(N, R(0));
        const R dc = R(decay);
        for (uint32_t i = 0; i < N; ++i) dtrace[i] = dtrace[i] * dc + (spk[i] != 0 ? R(1) : R(0));
        dense = true;
        d_lr_ltp = lr_ltp;
        d_lr_ltd = lr_ltd;
        d_wmin = w_min;
        d_wmax = w_max;
        step_once();
        dense = false;
    }
    void dense_update() {
        const R lp = R(d_lr_ltp), ld = R(d_lr_ltd), lo = R(d_wmin), hi = R(d_wmax);
        for (uint32_t pre = 0; pre < N; ++pre) {
            const R tp = dtrace[pre];
            for (uint32_t e = row_ptr[pre]; e < row_ptr[pre + 1]; ++e) {
                const R nw = (w[e] + lp * tp) - ld * dtrace[col[e]];
                w[e] = std::min(hi, std::max(lo, nw));
            }
        }
    }
    uint64_t step_counter() const override { return stepc; }
    void spikes(int32_t* out) const override { std::memcpy(out, spk.data(), 4 * static_cast<size_t>(N)); }
    uint32_t active(uint32_t* out) const override {
        if (out) std::memcpy(out, act.data(), 4 * act.size());
        return static_cast<uint32_t>(act.size());
    }
    int state(int p, int c, double* out) const override {
        if (p < 0 || p >= static_cast<int>(pops.size())) return -3;
        const Pop& P = pops[p];
        if (c < 0 || c >= n_state(P.model)) return -3;
        for (uint32_t i = 0; i < P.n; ++i) out[i] = static_cast<double>(P.st[static_cast<size_t>(c) * P.cap + i]);
        return 0;
    }
    void ring(double* out) const override {
        for (size_t k = 0; k < ringv.size(); ++k) out[k] = static_cast<double>(ringv[k]);
    }
    void weights(double* out) const override {
        for (size_t k = 0; k < w.size(); ++k) out[k] = static_cast<double>(w[k]);
    }
    uint64_t total_spikes() const override { return n_spikes; }
    uint64_t total_events() const override { return n_events; }
    // Turning recording on starts a fresh raster (the previous one is dropped).
    void record(bool on) override {
        rec = on;
        if (on) {
            rast.clear();
            rast_steps = 0;
        }
    }
    uint64_t raster_steps() const override { return rast_steps; }
    void raster(int32_t* out) const override { std::memcpy(out, rast.data(), 4 * rast.size()); }
};

// ---------------------------------------------------------------------------
// Kuramoto
// ---------------------------------------------------------------------------
// Coupling c_i = sum_j W_ij sin(theta_j - theta_i) over the CSR row of i
// (P/src/kernels/phase/sparse.rs:63-96, written there as
// sin_j cos_i - cos_j sin_i).  The reference integrates by forward Euler and
// wraps to [-pi, pi) (P/src/kernels/phase/kuramoto.rs:42-68,
// P/src/kernels/mod.rs:74-77); RK4 is this build's extension.  The oracle
// evaluates sin/cos with libm in its own precision.
struct KuraBase {
    virtual ~KuraBase() {}
    virtual void rk4(uint32_t n) = 0;
    virtual void euler(uint32_t n) = 0;
    virtual void phases(double* out) const = 0;
    virtual double order() const = 0;
};
template <class R>
struct Kura final : KuraBase {
    uint32_t n;
    std::vector<uint32_t> rp, col;
    std::vector<R> w, om, th, k1, k2, k3, k4, tmp, sn, cs;
    R dt;
    void deriv(const std::vector<R>& t, std::vector<R>& out) {
        for (uint32_t i = 0; i < n; ++i) { sn[i] = std::sin(t[i]); cs[i] = std::cos(t[i]); }
        for (uint32_t i = 0; i < n; ++i) {
            R acc = R(0);
            for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
                uint32_t j = col[e];
                acc += w[e] * (sn[j] * cs[i] - cs[j] * sn[i]);
            }
            out[i] = om[i] + acc;
        }
    }
    static R wrap(R p) {
        const R tau = R(6.283185307179586476925286766559);
        return p - tau * std::round(p / tau);
    }
    void rk4(uint32_t steps) override {
        for (uint32_t s = 0; s < steps; ++s) {
            deriv(th, k1);
            for (uint32_t i = 0; i < n; ++i) tmp[i] = th[i] + R(0.5) * dt * k1[i];
            deriv(tmp, k2);
            for (uint32_t i = 0; i < n; ++i) tmp[i] = th[i] + R(0.5) * dt * k2[i];
            deriv(tmp, k3);
            for (uint32_t i = 0; i < n; ++i) tmp[i] = th[i] + dt * k3[i];
            deriv(tmp, k4);
            for (uint32_t i = 0; i < n; ++i)
                th[i] = wrap(th[i] + dt / R(6) * (k1[i] + R(2) * k2[i] + R(2) * k3[i] + k4[i]));
        }
    }
    void euler(uint32_t steps) override {
        for (uint32_t s = 0; s < steps; ++s) {
            deriv(th, k1);
            for (uint32_t i = 0; i < n; ++i) th[i] = wrap(th[i] + dt * k1[i]);
        }
    }
    void phases(double* out) const override {
        for (uint32_t i = 0; i < n; ++i) out[i] = static_cast<double>(th[i]);
    }
    // phase_coherence with unit amplitudes (P/src/kernels/measure/mod.rs:76-106)
    double order() const override {
        double ss = 0, cc = 0;
        for (uint32_t i = 0; i < n; ++i) { ss += std::sin(static_cast<double>(th[i])); cc += std::cos(static_cast<double>(th[i])); }
        return std::sqrt(ss * ss + cc * cc) / (n ? n : 1);
    }
};

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// Creature pool: BrainPool (DE/src/brain/dynamics/gpu_pool.rs) over the
// WgpuNetwork step (PW/src/network.rs:802-885) and its shaders
// (PW/shaders/csr_scatter.wgsl), restated loop for loop in FP32.
// ---------------------------------------------------------------------------
struct OraclePool {
    orc_pool_params c;
    uint8_t delay;
    // BrainPool's concatenated CSR (gpu_pool.rs:101-128)
    std::vector<uint32_t> row_ptr{0}, col;
    std::vector<float> values, nt;
    std::vector<uint8_t> delays;
    std::vector<uint32_t> slot_off, slot_n, slot_syn;
    uint32_t n = 0;
    bool built = false;
    // WgpuNetwork state
    std::vector<uint32_t> rev_ptr, rev_src, rev_edge, rev_d;
    std::vector<float> ring, i_in, v, g, refr;
    std::vector<uint32_t> spikes, spikes_prev;
    uint32_t slot = 0;
    float refr_steps = 0, tf = 0, syn_decay = 0;

    void build() {  // ensure_built, gpu_pool.rs:205-260, and WgpuNetwork::new
        if (built) return;
        refr_steps = std::max(std::round(c.tau_ref / c.dt_ms), 1.0f);
        tf = c.dt_ms / c.tau_m;
        syn_decay = 1.0f - c.dt_ms / c.tau_syn;
        v.assign(n, c.v_rest);
        g.assign(n, 0.0f);
        refr.assign(n, refr_steps);
        const size_t ns = col.size();
        rev_ptr.assign(n + 1, 0);
        for (uint32_t x : col) rev_ptr[x + 1] += 1;
        for (uint32_t i = 0; i < n; ++i) rev_ptr[i + 1] += rev_ptr[i];
        std::vector<uint32_t> cursor(rev_ptr);
        rev_src.assign(ns, 0);
        rev_edge.assign(ns, 0);
        rev_d.assign(ns, 0);
        for (uint32_t s = 0; s < n; ++s)
            for (uint32_t e = row_ptr[s]; e < row_ptr[s + 1]; ++e) {
                const uint32_t k = cursor[col[e]]++;
                rev_src[k] = s;
                rev_edge[k] = e;
                rev_d[k] = delays[e];
            }
        ring.assign(static_cast<size_t>(kSlots) * n, 0.0f);
        i_in.assign(n, 0.0f);
        spikes.assign(n, 0);
        spikes_prev.assign(n, 0);
        slot = 0;
        built = true;
    }

    void substep(const std::vector<float>* jump) {
        if (jump)  // v_jump_kernel: v += value
            for (uint32_t i = 0; i < n; ++i) v[i] += (*jump)[i];
        for (uint32_t i = 0; i < n; ++i) i_in[i] = ring[static_cast<size_t>(slot) * n + i];  // copy_slot_kernel
        for (uint32_t i = 0; i < n; ++i) {
            float vv = v[i], gg = g[i], rr = refr[i];
            spikes[i] = static_cast<uint32_t>(
                alpha_lif_step<float>(vv, gg, rr, i_in[i], tf, c.v_rest, c.v_thresh, c.v_reset, refr_steps, syn_decay));
            v[i] = vv;
            g[i] = gg;
            refr[i] = rr;
        }
        if (!col.empty()) {  // scatter_kernel over the previous step's spikes
            for (uint32_t post = 0; post < n; ++post) {
                ring[static_cast<size_t>(slot) * n + post] = 0.0f;
                float acc[kSlots] = {};
                for (uint32_t k = rev_ptr[post]; k < rev_ptr[post + 1]; ++k) {
                    const uint32_t src = rev_src[k];
                    if (spikes_prev[src] == 0) continue;
                    const uint32_t d = std::max(rev_d[k], 1u);
                    acc[d % kSlots] = std::fma(values[rev_edge[k]], nt[src], acc[d % kSlots]);
                }
                for (uint32_t d = 0; d < kSlots; ++d)
                    if (acc[d] != 0.0f) ring[static_cast<size_t>((slot + d) % kSlots) * n + post] += acc[d];
            }
        }
        spikes_prev = spikes;
        slot = (slot + 1) % kSlots;
    }
};

extern "C" {

void* orc_net_create(int kind, const orc_pop_desc* pops, int32_t n_pops, const uint32_t* row_ptr,
                     const uint32_t* col, const float* w, const uint8_t* delay, uint64_t n_syn) {
    NetBase* nb = nullptr;
    int rc;
    if (kind == 64) {
        auto* n = new Net<double>();
        rc = n->init(pops, n_pops, row_ptr, col, w, delay, n_syn);
        nb = n;
    } else {
        auto* n = new Net<float>();
        rc = n->init(pops, n_pops, row_ptr, col, w, delay, n_syn);
        nb = n;
    }
    if (rc != 0) { delete nb; return nullptr; }
    return nb;
}
void orc_net_destroy(void* net) { delete static_cast<NetBase*>(net); }
int orc_net_add_poisson_source(void* net, float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts,
                               const float* w, const uint8_t* delay, uint32_t count) {
    return static_cast<NetBase*>(net)->add_source(rate_hz, dt_ms, seed, posts, w, delay, count);
}
int orc_net_set_noise(void* net, int32_t pop, const float* sigma, uint64_t seed) {
    return static_cast<NetBase*>(net)->set_noise(pop, sigma, seed);
}
int orc_net_set_plasticity(void* net, const orc_stdp_params* p, const uint8_t* plastic) {
    return static_cast<NetBase*>(net)->set_plasticity(p, plastic);
}
int orc_net_inject_current(void* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count) {
    return static_cast<NetBase*>(net)->inject(pop, idx, val, count);
}
int orc_net_inject_voltage(void* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count) {
    return static_cast<NetBase*>(net)->inject_voltage(pop, idx, val, count);
}
void orc_net_add_gap_junction(void* net, uint32_t a, uint32_t b, float g) {
    static_cast<NetBase*>(net)->add_coupling(0, a, b, g);
}
void orc_net_add_graded_synapse(void* net, uint32_t pre, uint32_t post, float g) {
    static_cast<NetBase*>(net)->add_coupling(1, pre, post, g);
}
int orc_net_add_sigmoid_profile(void* net, float vth, float delta, float k, float erev, float cond, float dt) {
    return static_cast<NetBase*>(net)->add_sigmoid_profile(vth, delta, k, erev, cond, dt);
}
int orc_net_add_sigmoid_conn(void* net, int32_t prof, uint32_t pre, uint32_t post, float weight) {
    return static_cast<NetBase*>(net)->add_sigmoid_conn(prof, pre, post, weight);
}
void orc_net_step(void* net, uint32_t n_steps) { static_cast<NetBase*>(net)->step(n_steps); }
int orc_net_set_receptors(void* net, const uint8_t* types, float scale, const float* ext) {
    return static_cast<NetBase*>(net)->set_receptors(types, scale, ext);
}
void orc_net_conductances(void* net, double* out) { static_cast<NetBase*>(net)->conductances(out); }
int orc_net_set_state(void* net, uint64_t step, const float* const* pop_states, const float* ring, const float* w,
                      const double* w64, const float* tr_pre, const float* tr_post, const float* g) {
    return static_cast<NetBase*>(net)->set_state(step, pop_states, ring, w, w64, tr_pre, tr_post, g);
}
void orc_rec_decay_factors(float dt, float* out) {
    for (int r = 0; r < kRec; ++r) out[r] = rec_decay<float>(dt, r);
}
float orc_nmda_block(float v) { return nmda_block<float>(v); }
void orc_rec_currents(const float* g, const float* v, float* i_out, uint32_t n) {
    const uint32_t n8 = n & ~7u;
    for (uint32_t i = 0; i < n; ++i) {
        float gi[kRec];
        for (int r = 0; r < kRec; ++r) gi[r] = g[static_cast<size_t>(r) * n + i];
        i_out[i] = rec_current<float>(gi, v[i], i_out[i], i < n8);
    }
}
int orc_net_set_stp(void* net, const orc_stp_profile* profiles, int32_t n_profiles, float dt,
                    const uint8_t* syn_types, const uint8_t* src_types, uint64_t n_src_edges) {
    return static_cast<NetBase*>(net)->set_stp(profiles, n_profiles, dt, syn_types, src_types, n_src_edges);
}
float orc_tm_spike(float u0, float a_rec, float a_fac, int32_t is_fac, float* r, float* u, uint32_t* last,
                   float weight, uint64_t step) {
    TmProfile p{u0, a_rec, a_fac, is_fac != 0};
    return tm_spike_r<float>(p, *r, *u, *last, weight, step);
}
float orc_powi_f32(float a, int32_t n) { return powi_r<float>(a, n); }
float orc_fast_exp(float x) { return fast_exp_f32(x); }
void orc_gpea_rates(float v, float cai, float* out) {
    using G = Gpea<float>;
    out[0] = G::m_inf(v); out[1] = G::h_inf(v); out[2] = G::n_inf(v); out[3] = G::p_inf(v);
    out[4] = G::q_inf(v); out[5] = G::r_inf(cai); out[6] = G::tau_m(v); out[7] = G::tau_h(v);
    out[8] = G::tau_n(v); out[9] = G::tau_p(v); out[10] = G::tau_q(v);
}
void orc_net_stdp_step(void* net, float decay, float lr_ltp, float lr_ltd, float w_min, float w_max) {
    static_cast<NetBase*>(net)->stdp_step(decay, lr_ltp, lr_ltd, w_min, w_max);
}
uint64_t orc_net_step_counter(void* net) { return static_cast<NetBase*>(net)->step_counter(); }
void orc_net_spikes(void* net, int32_t* out) { static_cast<NetBase*>(net)->spikes(out); }
uint32_t orc_net_active(void* net, uint32_t* out) { return static_cast<NetBase*>(net)->active(out); }
int orc_net_state(void* net, int32_t pop, int32_t col, double* out) {
    return static_cast<NetBase*>(net)->state(pop, col, out);
}
void orc_net_ring(void* net, double* out) { static_cast<NetBase*>(net)->ring(out); }
void orc_net_weights(void* net, double* out) { static_cast<NetBase*>(net)->weights(out); }
uint64_t orc_net_total_spikes(void* net) { return static_cast<NetBase*>(net)->total_spikes(); }
uint64_t orc_net_total_events(void* net) { return static_cast<NetBase*>(net)->total_events(); }
void orc_net_record(void* net, int32_t on) { static_cast<NetBase*>(net)->record(on != 0); }
uint64_t orc_net_raster_steps(void* net) { return static_cast<NetBase*>(net)->raster_steps(); }
void orc_net_raster(void* net, int32_t* out) { static_cast<NetBase*>(net)->raster(out); }
void orc_net_set_threads(void* net, int32_t n_threads) {
    static_cast<NetBase*>(net)->n_threads = n_threads < 1 ? 1 : n_threads;
}

int32_t orc_poisson_spike(uint32_t seed, uint32_t stream, uint64_t step, float p) {
    return poisson_spike(seed, stream, step, p) ? 1 : 0;
}
float orc_gauss(uint64_t seed, uint32_t i, uint64_t step) { return gauss(seed, i, step); }
float orc_cos2pi_f32(float u) { return cos2pi_f32(u); }
float orc_sin2pi_f32(float u) {
    float c, s;
    sincos2pi_f32(u, c, s);
    return s;
}
float orc_exp_f32(float x) { return exp_f32(x); }
float orc_ln_f32(float x) { return ln_f32(x); }
float orc_vtrap_f32(float x, float y) { return vtrap_f32(x, y); }
double orc_vtrap_f64(double x, double y) { return vtrap_f64(x, y); }

// compact_spikes_into — P/src/api/net/compact.rs:30-61 (ascending indices).
uint32_t orc_compact(const int32_t* spikes, uint32_t n, uint32_t* out) {
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (spikes[i] != 0) out[k++] = i;
    return k;
}

void* orc_kura_create(int kind, uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const float* w,
                      const float* omega, const float* theta0, float dt) {
    auto fill = [&](auto* k) {
        using R = typename std::remove_reference<decltype(k->w[0])>::type;
        k->n = n;
        k->rp.assign(row_ptr, row_ptr + n + 1);
        uint64_t ne = row_ptr[n];
        k->col.assign(col, col + ne);
        k->w.resize(ne);
        for (uint64_t e = 0; e < ne; ++e) k->w[e] = R(w[e]);
        k->om.resize(n); k->th.resize(n);
        for (uint32_t i = 0; i < n; ++i) { k->om[i] = R(omega[i]); k->th[i] = R(theta0[i]); }
        k->k1.assign(n, R(0)); k->k2 = k->k1; k->k3 = k->k1; k->k4 = k->k1;
        k->tmp = k->k1; k->sn = k->k1; k->cs = k->k1;
        k->dt = R(dt);
    };
    if (kind == 64) { auto* k = new Kura<double>(); fill(k); return static_cast<KuraBase*>(k); }
    auto* k = new Kura<float>(); fill(k); return static_cast<KuraBase*>(k);
}
void orc_kura_destroy(void* k) { delete static_cast<KuraBase*>(k); }
void orc_kura_step_rk4(void* k, uint32_t n) { static_cast<KuraBase*>(k)->rk4(n); }
void orc_kura_step_euler(void* k, uint32_t n) { static_cast<KuraBase*>(k)->euler(n); }
void orc_kura_phases(void* k, double* out) { static_cast<KuraBase*>(k)->phases(out); }
double orc_kura_order(void* k) { return static_cast<KuraBase*>(k)->order(); }

// firing_rates — P/src/kernels/stats/rates.rs:6-35 (dt in seconds).
void orc_firing_rates(const int32_t* raster, uint32_t n, uint32_t steps, float dt_s, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = 0.0f;
    for (uint32_t t = 0; t < steps; ++t)
        for (uint32_t i = 0; i < n; ++i)
            if (raster[static_cast<size_t>(t) * n + i] != 0) out[i] += 1.0f;
    float dur = static_cast<float>(steps) * dt_s;
    if (dur > 0.0f)
        for (uint32_t i = 0; i < n; ++i) out[i] /= dur;
}

// compute_cvs — P/src/kernels/stats/isi.rs:30-67 (n-1 variance, >= 2 ISIs).
void orc_isi_cv(const int32_t* raster, uint32_t n, uint32_t steps, float dt_s, float* out) {
    std::vector<float> isi;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = 0.0f;
        isi.clear();
        int64_t last = -1;
        for (uint32_t t = 0; t < steps; ++t) {
            if (raster[static_cast<size_t>(t) * n + i] != 0) {
                if (last >= 0) {
                    float iv = static_cast<float>(t - last) * dt_s * 1000.0f;
                    if (iv > 0.0f) isi.push_back(iv);
                }
                last = t;
            }
        }
        if (isi.size() >= 2) {
            float mean = 0.0f;
            for (float x : isi) mean += x;
            mean /= static_cast<float>(isi.size());
            if (mean > 0.0f) {
                float var = 0.0f;
                for (float x : isi) var += (x - mean) * (x - mean);
                var /= static_cast<float>(isi.size() - 1);
                out[i] = std::sqrt(var) / mean;
            }
        }
    }
}

void* orc_pool_create(const orc_pool_params* p, int32_t delay) {
    auto* P = new OraclePool();
    P->c = *p;
    // BrainPool::new, gpu_pool.rs:63-90: steps_delay + 1, capped at 31
    const float sd = std::max(std::round(p->t_delay / p->dt_ms), 0.0f);
    const uint32_t d = delay >= 0 ? static_cast<uint32_t>(delay) : static_cast<uint32_t>(sd) + 1u;
    P->delay = static_cast<uint8_t>(std::min(d, 31u));
    return P;
}
void orc_pool_destroy(void* pool) { delete static_cast<OraclePool*>(pool); }
int32_t orc_pool_register(void* pool, uint32_t n, const int32_t* row_ptr, const int32_t* col, const float* values,
                          const float* nt_types) {
    auto* P = static_cast<OraclePool*>(pool);
    const uint32_t off = P->n, syn_off = static_cast<uint32_t>(P->values.size());
    const uint32_t ns = static_cast<uint32_t>(row_ptr[n]);
    P->row_ptr.pop_back();
    for (uint32_t i = 0; i <= n; ++i) P->row_ptr.push_back(static_cast<uint32_t>(row_ptr[i]) + syn_off);
    for (uint32_t e = 0; e < ns; ++e) P->col.push_back(static_cast<uint32_t>(col[e]) + off);
    P->values.insert(P->values.end(), values, values + ns);
    P->delays.insert(P->delays.end(), ns, P->delay);
    for (uint32_t i = 0; i < n; ++i) P->nt.push_back(nt_types[i] < 0.0f ? -1.0f : 1.0f);
    P->slot_off.push_back(off);
    P->slot_n.push_back(n);
    P->slot_syn.push_back(syn_off);
    P->n += n;
    P->built = false;  // needs_rebuild: every creature starts over
    return static_cast<int32_t>(P->slot_off.size() - 1);
}
int orc_pool_set_weights(void* pool, int32_t k, const float* w, uint64_t n_syn) {
    auto* P = static_cast<OraclePool*>(pool);
    if (k < 0 || k >= static_cast<int32_t>(P->slot_off.size())) return -3;
    const uint32_t s0 = P->slot_syn[k];
    const uint32_t s1 = k + 1 < static_cast<int32_t>(P->slot_syn.size()) ? P->slot_syn[k + 1]
                                                                          : static_cast<uint32_t>(P->values.size());
    if (n_syn != s1 - s0) return -3;
    if (P->n) P->build();
    std::copy(w, w + n_syn, P->values.begin() + s0);
    return 0;
}
int orc_pool_set_delays(void* pool, int32_t k, const uint8_t* d, uint64_t n_syn) {
    auto* P = static_cast<OraclePool*>(pool);
    if (k < 0 || k >= static_cast<int32_t>(P->slot_off.size())) return -3;
    const uint32_t s0 = P->slot_syn[k];
    const uint32_t s1 = k + 1 < static_cast<int32_t>(P->slot_syn.size()) ? P->slot_syn[k + 1]
                                                                          : static_cast<uint32_t>(P->values.size());
    if (n_syn != s1 - s0) return -3;
    std::copy(d, d + n_syn, P->delays.begin() + s0);
    P->built = false;  // WgpuNetwork::new takes the delays: a new network
    return 0;
}
int orc_pool_step(void* pool, const float* ext, uint32_t spt, uint32_t* out) {
    auto* P = static_cast<OraclePool*>(pool);
    if (P->slot_off.empty()) return 0;
    P->build();
    std::vector<float> jump(P->n, 0.0f);
    if (ext) std::copy(ext, ext + P->n, jump.begin());
    for (uint32_t k = 0; k < spt; ++k) {
        P->substep(k % 64 == 0 ? &jump : nullptr);  // one v_jump per submit of <= 64 steps
        std::copy(P->spikes.begin(), P->spikes.end(), out + static_cast<size_t>(k) * P->n);
    }
    return 0;
}
int orc_pool_state(void* pool, int32_t col, float* out) {
    auto* P = static_cast<OraclePool*>(pool);
    if (P->n) P->build();
    const std::vector<float>* s = col == 0 ? &P->v : col == 1 ? &P->g : col == 2 ? &P->refr : nullptr;
    if (!s) return -3;
    std::copy(s->begin(), s->end(), out);
    return 0;
}
uint32_t orc_pool_n_total(void* pool) { return static_cast<OraclePool*>(pool)->n; }
int32_t orc_pool_delay(void* pool) { return static_cast<OraclePool*>(pool)->delay; }

}  // extern "C"
