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
    const float u = static_cast<float>(z >> 40) / 16777216.0f;
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

// Standard normal draw keyed by (seed, neuron, step): Marsaglia's polar
// method on two 24-bit uniforms per attempt.  Counter-based, so any neuron
// of any step can be drawn independently and repeatably.
__device__ __forceinline__ float normal_draw(uint64_t seed, uint32_t i, uint64_t step) {
    const uint64_t base = splitmix(seed ^ (static_cast<uint64_t>(i) * 0xD1B54A32D192ED03ull));
    for (uint32_t k = 0; k < 64; ++k) {
        const uint64_t h = splitmix(base ^ (step * 0xA24BAED4963EE407ull + k));
        const float u1 = static_cast<float>(h >> 40) * 5.9604644775390625e-08f;
        const float u2 = static_cast<float>((h >> 16) & 0xFFFFFFu) * 5.9604644775390625e-08f;
        const float x = u1 * 2.0f - 1.0f;
        const float y = u2 * 2.0f - 1.0f;
        const float s = x * x + y * y;
        if (s >= 1.0f || s == 0.0f) continue;
        const float q = __builtin_sqrtf((-2.0f * det_logf(s)) / s);
        return x * q;
    }
    return 0.0f;
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

}  // namespace simuse
