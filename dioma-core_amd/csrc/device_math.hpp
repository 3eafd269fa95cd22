// device_math.hpp — integer hashing and deterministic float math for gfx950.
//
// Everything here is built from IEEE +, -, *, /, sqrt, floor, compares and
// integer ops only, and the library is compiled with -ffp-contract=off, so a
// CPU evaluating the same operation sequence gets identical bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace simuse {

// Reference Poisson roll: SplitMix64 finaliser over ((seed<<32)|stream) mixed
// with step*phi64, 24-bit uniform compared with p
// (P/src/api/net/network/sources.rs:5-14).
__device__ __forceinline__ bool poisson_roll(uint32_t seed, uint32_t stream, uint64_t step, float p) {
    uint64_t z = (static_cast<uint64_t>(seed) << 32) | static_cast<uint64_t>(stream);
    z ^= step * 0x9E3779B97F4A7C15ull;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    // 24-bit integers convert exactly from 32 bits (one instruction, not the 64-bit sequence)
    const float u = static_cast<float>(static_cast<uint32_t>(z >> 40)) / 16777216.0f;
    return u < p;
}

// SplitMix64 output function (Steele, Lea & Flood 2014).
__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Natural log of a normal positive float: x = 2^e * m with m folded into
// (1/sqrt2, sqrt2], ln m = 2 atanh(f/(2+f)) as an odd series in s = f/(2+f).
__device__ __forceinline__ float det_logf(float x) {
    const uint32_t b = __float_as_uint(x);
    int e = static_cast<int>((b >> 23) & 0xFFu) - 127;
    float m = __uint_as_float((b & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float s2 = s * s;
    float p = 0.181818187f;
    p = p * s2 + 0.222222224f;
    p = p * s2 + 0.285714298f;
    p = p * s2 + 0.400000006f;
    p = p * s2 + 0.666666687f;
    p = p * s2 + 2.0f;
    const float lnm = s * p;
    const float ef = static_cast<float>(e);
    return ef * 0.693145752f + (lnm + ef * 1.42860677e-06f);
}

// e^x: x = k ln2 + r with a two-part ln2 (the high part has 16 trailing zero
// bits, so k*ln2_hi is exact), degree-7 Taylor polynomial in r, 2^k through
// the exponent field; underflow below 2^-126 returns 0.
__device__ __forceinline__ float det_expf(float x) {
    if (!(x == x)) return x;
    if (x > 88.72f) return __builtin_inff();
    if (x < -87.33f) return 0.0f;
    const float kf = floorf(x * 1.44269502f + 0.5f);
    const float r = (x - kf * 0.693145752f) - kf * 1.42860677e-06f;
    float p = 1.98412701e-04f;
    p = p * r + 1.38888892e-03f;
    p = p * r + 8.33333377e-03f;
    p = p * r + 4.16666679e-02f;
    p = p * r + 0.166666672f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int k = static_cast<int>(kf);
    if (k > 127) return p * __uint_as_float(0x7F000000u) * 2.0f;
    if (k < -126) return 0.0f;
    return p * __uint_as_float(static_cast<uint32_t>(k + 127) << 23);
}

// fast_exp (P/src/math/fast_exp.rs:21-63): 2^n * p(f), t = x*log2(e),
// n = floor(t) (saturating), f = t - n, degree-6 Horner rounded step by step.
// Host and device evaluate the same operations (the host computes the
// per-population receptor decay factors with it).
__host__ __device__ __forceinline__ float fast_expf(float x) {
    const float t = x * 1.44269502f;
    float tc = t;
    if (tc > 1000.0f) tc = 1000.0f;
    if (tc < -1000.0f) tc = -1000.0f;
    const int n = tc == tc ? static_cast<int>(floorf(tc)) : 0;
    const float f = t - static_cast<float>(n);
    const float c1 = 0.693147182f, c2 = 0.2402265f, c3 = 0.05550411f, c4 = 0.009618129f, c5 = 0.001333356f,
                c6 = 0.0001540353f;
    float poly = (c6 * f + c5) * f + c4;
    poly = (poly * f + c3) * f + c2;
    poly = (poly * f + c1) * f + 1.0f;
    if (n < -126) return 0.0f;
    if (n > 127) return __builtin_inff();
    const uint32_t bits = static_cast<uint32_t>(n + 127) << 23;
    float scale;
    __builtin_memcpy(&scale, &bits, 4);
    return scale * poly;
}

// cos and sin of 2 pi u for u in [0, 1) on a 2^-24 grid: quadrant from 4u
// (exact), then sin and cos of (pi/2) g, g in [0, 1/2], by their Taylor
// series through x^9 / x^10 (truncation < 2e-9 on [0, pi/4]), Horner in FP32;
// sin(2 pi u) = cos(2 pi (u - 1/4)) is the quadrant before.
__device__ __forceinline__ void det_sincos2pi(float u, float& cosv, float& sinv) {
    const float t = u * 4.0f;
    const float q = floorf(t);
    const float f = t - q;
    const bool hi = f > 0.5f;
    const float g = hi ? 1.0f - f : f;
    const float x = g * 1.57079637f;
    const float x2 = x * x;
    const float sn = x * (1.0f + x2 * (-0.166666672f + x2 * (0.00833333377f + x2 * (-0.000198412701f +
                                                                                     x2 * 2.75573188e-06f))));
    const float cs = 1.0f + x2 * (-0.5f + x2 * (0.0416666679f + x2 * (-0.00138888892f +
                                                                     x2 * (2.48015876e-05f + x2 * -2.75573188e-07f))));
    const float c = hi ? sn : cs;  // cos((pi/2) f)
    const float s = hi ? cs : sn;  // sin((pi/2) f)
    const int qi = static_cast<int>(q);
    cosv = qi == 0 ? c : qi == 1 ? -s : qi == 2 ? -c : s;
    sinv = qi == 0 ? s : qi == 1 ? c : qi == 2 ? -s : -c;
}

// Standard normal draws keyed by (seed, neuron, step): Box & Muller (1958)
// on two 24-bit uniforms of one SplitMix64 hash of (seed, pair, step), u1 in
// (0, 1], u2 in [0, 1).  Neurons i and i ^ 1024 form a pair (key i & ~1024):
// the one with bit 10 clear gets sqrt(-2 ln u1) cos(2 pi u2), its partner
// the sine, two independent normals from one hash, one log and one sqrt.
// Counter-based and branch-free, so every lane of a wavefront draws in
// lockstep.
__device__ __forceinline__ void normal_pair(uint64_t seed, uint32_t pair, uint64_t step, float& zc, float& zs) {
    const uint64_t base = splitmix(seed ^ (static_cast<uint64_t>(pair) * 0xD1B54A32D192ED03ull));
    const uint64_t h = splitmix(base ^ (step * 0xA24BAED4963EE407ull));
    const float u1 = static_cast<float>(static_cast<uint32_t>(h >> 40) + 1u) * 5.9604644775390625e-08f;
    const float u2 = static_cast<float>(static_cast<uint32_t>(h >> 16) & 0xFFFFFFu) * 5.9604644775390625e-08f;
    const float r = __builtin_sqrtf(-2.0f * det_logf(u1));
    float c, s;
    det_sincos2pi(u2, c, s);
    zc = r * c;
    zs = r * s;
}

__device__ __forceinline__ float normal_draw(uint64_t seed, uint32_t i, uint64_t step) {
    float zc, zs;
    normal_pair(seed, i & ~1024u, step, zc, zs);
    return (i & 1024u) ? zs : zc;
}

// x / (e^{x/y} - 1) (NEURON hh.mod vtrap, PC/kernels/hh.cu:28-34).  For
// |x/y| < 0.25 the FP32 difference e^r - 1 loses up to ~20 bits, so the
// value there is y * B(r), B(r) = r/(e^r - 1) = sum B_n r^n / n!, in FP64.
__device__ __forceinline__ float vtrap(float x, float y) {
    const float r = x / y;
    if (fabsf(r) < 0.25f) {
        const double rd = static_cast<double>(x) / static_cast<double>(y);
        const double r2 = rd * rd;
        double ev = -691.0 / 1307674368000.0;
        ev = ev * r2 + 1.0 / 47900160.0;
        ev = ev * r2 - 1.0 / 1209600.0;
        ev = ev * r2 + 1.0 / 30240.0;
        ev = ev * r2 - 1.0 / 720.0;
        ev = ev * r2 + 1.0 / 12.0;
        const double bern = (1.0 - 0.5 * rd) + r2 * ev;
        return static_cast<float>(static_cast<double>(y) * bern);
    }
    return x / (det_expf(r) - 1.0f);
}

// AlphaLifModel::step, P/src/zoo/neurons/alpha_lif.rs:99-136: refractoriness
// masks only the delayed input; v integrates with the previous g through an
// explicit fused multiply-add, the threshold is strict (v > v_thresh), a
// spike resets v and g, and g decays by syn_decay with a fused multiply-add.
// (fmaf is the one contraction the library asks for: the reference's shader
// and CPU step both fuse exactly these two nodes.)
__device__ __forceinline__ int alpha_lif_update(float& v, float& g, float& refr_cnt, float I, float tf,
                                                float v_rest, float v_th, float v_reset, float refr_steps,
                                                float syn_decay) {
    const float g_old = g;
    const float delayed = refr_cnt >= refr_steps ? I : 0.0f;
    const float integrand = g_old - (v - v_rest);
    const float v_new = fmaf(tf, integrand, v);
    const bool spike = v_new > v_th;
    v = spike ? v_reset : v_new;
    const float g_new = fmaf(syn_decay, g_old, delayed);
    g = spike ? 0.0f : g_new;
    refr_cnt = spike ? 0.0f : refr_cnt + 1.0f;
    return spike ? 1 : 0;
}

}  // namespace simuse
