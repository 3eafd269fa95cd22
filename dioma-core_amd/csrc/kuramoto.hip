// kuramoto.hip — Kuramoto phase coupling on a CSR graph, Euler or RK4.
//
// Coupling of node i: c_i = sum_j W_ij sin(theta_j - theta_i)
//                         = cos_i * sum_j W_ij sin_j - sin_i * sum_j W_ij cos_j
// (P/src/kernels/phase/sparse.rs:63-96 computes the same sum as
// w * (sin_j cos_i - cos_j sin_i); GPU twin PC/kernels/kuramoto.cu:32-75).
// The reference integrates with forward Euler and wraps to [-pi, pi)
// (P/src/kernels/phase/kuramoto.rs:42-68, P/src/kernels/mod.rs:74-77); the
// RK4 integrator is this engine's extension for long horizons.
//
// Precision: the phase is accumulated in FP64 (theta, 8 B/node); stage phases,
// their sin/cos and the derivatives are FP32.  One kernel per RK4 stage:
// it gathers the current stage's (sin, cos) over the row, forms the stage
// derivative, folds it into the RK4 accumulator, and writes (sin, cos) of the
// next stage's phase into the other buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/simuse.h"

namespace simuse {
int fail(int code, const char* fmt, ...);
}

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct KuraArgs {
    uint32_t n;
    uint32_t fixed_k;     // every row has this many edges (row i starts at i * fixed_k), or 0: read row_ptr
    const uint32_t* row_ptr;
    const uint32_t* col;
    const float* w;
    const float* w_row;   // every edge of row i has weight w_row[i] (bit-equal), or null: read w per edge
    const float* omega;
    double* theta;        // FP64 master phase, wrapped to [-pi, pi)
    double* acc;          // RK4 accumulator k1 + 2k2 + 2k3 (FP64)
    const float2* sc_in;  // (sin, cos) of the current stage phase
    float2* sc_out;       // (sin, cos) of the next stage phase
    double dt;
};

__device__ __forceinline__ double wrap_phase(double p) {
    // p - 2pi * round(p / 2pi), round half away from zero (Rust f32::round)
    return p - kTwoPi * round(p / kTwoPi);
}

// Gather sum_j w_ij (sin_j, cos_j) for row i with G lanes per row.  A graph
// whose rows each carry one weight (the modular small world's K / k) reads
// it once per row instead of once per edge, and one whose rows all have the
// same degree computes the row's extent: the same products and sums in the
// same order, 12 instead of 16 bytes per edge.
//
// A lane's edges are r0 + sub, r0 + sub + G, ...; they are taken kGatherBatch
// at a time: the column indices (and weights) first, then the neighbours'
// (sin, cos), then the products added in edge order, so a lane waits on two
// memory round trips per batch instead of two per edge (1M oscillators / 20M
// edges, RK4: 239 -> 154 us per step with 4 lanes per row, same bits).
constexpr uint32_t kGatherBatch = 8;

// Lanes per row: the fewest (a power of two, at most 64) whose batches cover
// a row of the mean degree in one pass.
int lanes_for_degree(double mean) {
    int g = 1;
    while (g < 64 && static_cast<double>(g) * kGatherBatch < mean) g <<= 1;
    return g;
}

template <int G>
__device__ __forceinline__ float2 row_gather(const KuraArgs& a, uint32_t i, uint32_t sub) {
    constexpr uint32_t kU = kGatherBatch;
    float s = 0.0f, c = 0.0f;
    const uint32_t r0 = a.fixed_k ? i * a.fixed_k : a.row_ptr[i];
    const uint32_t r1 = a.fixed_k ? r0 + a.fixed_k : a.row_ptr[i + 1];
    const bool row_w = a.w_row != nullptr;
    const float w_i = row_w ? a.w_row[i] : 0.0f;
    for (uint32_t e0 = r0 + sub; e0 < r1; e0 += kU * G) {
        uint32_t src[kU];
        float wv[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t e = e0 + u * G;
            src[u] = e < r1 ? a.col[e] : 0u;
            wv[u] = row_w ? w_i : (e < r1 ? a.w[e] : 0.0f);
        }
        float2 sc[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u)
            sc[u] = e0 + u * G < r1 ? a.sc_in[src[u]] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            if (e0 + u * G < r1) {
                s += wv[u] * sc[u].x;
                c += wv[u] * sc[u].y;
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, G);
        c += __shfl_xor(c, o, G);
    }
    return make_float2(s, c);
}

// STAGE: 0 = Euler step; 1..4 = RK4 stages.
template <int G, int STAGE>
__global__ __launch_bounds__(256) void kura_stage_kernel(KuraArgs a) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = gid / G;
    const uint32_t sub = gid % G;
    if (i >= a.n) return;  // whole groups exit together (blockDim % G == 0)
    // the row's own values load beside the gather, not after it
    float2 me = make_float2(0.0f, 0.0f);
    float om = 0.0f;
    double th = 0.0, acc = 0.0;
    if (sub == 0) {
        me = a.sc_in[i];
        om = a.omega[i];
        th = a.theta[i];
        if (STAGE >= 2) acc = a.acc[i];
    }
    const float2 sum = row_gather<G>(a, i, sub);
    if (sub != 0) return;
    const float k = om + (me.y * sum.x - me.x * sum.y);
    double next;
    if (STAGE == 0) {
        next = wrap_phase(th + a.dt * static_cast<double>(k));
        a.theta[i] = next;
    } else if (STAGE == 1) {
        a.acc[i] = static_cast<double>(k);
        next = th + 0.5 * a.dt * static_cast<double>(k);
    } else if (STAGE == 2) {
        a.acc[i] = acc + 2.0 * static_cast<double>(k);
        next = th + 0.5 * a.dt * static_cast<double>(k);
    } else if (STAGE == 3) {
        a.acc[i] = acc + 2.0 * static_cast<double>(k);
        next = th + a.dt * static_cast<double>(k);
    } else {
        const double inc = a.dt / 6.0 * (acc + static_cast<double>(k));
        next = wrap_phase(th + inc);
        a.theta[i] = next;
    }
    float sn, cs;
    sincosf(static_cast<float>(next), &sn, &cs);
    a.sc_out[i] = make_float2(sn, cs);
}

__global__ void kura_prepare_kernel(const double* theta, float2* sc, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float sn, cs;
    sincosf(static_cast<float>(theta[i]), &sn, &cs);
    sc[i] = make_float2(sn, cs);
}

// Order parameter R = |sum_i e^{i theta_i}| / n, deterministic two-pass sum
// in FP64 (P/src/kernels/measure/mod.rs:76-106 with unit amplitudes).
__global__ void kura_order_partial(const double* theta, const float* amp, uint32_t n, double2* part) {
    __shared__ double ss[256], cc[256];
    double s = 0, c = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double sn, cs;
        sincos(theta[i], &sn, &cs);
        const double a = amp ? static_cast<double>(amp[i]) : 1.0;
        s += a * sn;
        c += a * cs;
    }
    ss[threadIdx.x] = s;
    cc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < static_cast<uint32_t>(o)) {
            ss[threadIdx.x] += ss[threadIdx.x + o];
            cc[threadIdx.x] += cc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(ss[0], cc[0]);
}

// Directed modular small-world graph: node i of module m has k in-edges from
// its lattice neighbours i±1..±k/2 on the module ring; each is redrawn with
// probability `inter` from a uniform node of another module, otherwise with
// probability `beta` from a uniform node of its own module.  W = coupling/k.
__device__ __forceinline__ uint64_t kmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void kura_gen_kernel(uint32_t* row_ptr, uint32_t* col, float* w, float* omega, float* theta0, uint32_t n,
                                uint32_t modules, uint32_t k, float beta, float inter, float coupling, uint64_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { row_ptr[n] = n * k; return; }
    const uint32_t msize = n / modules;
    uint32_t m = i / msize;
    if (m >= modules) m = modules - 1;
    const uint32_t base = m * msize;
    const uint32_t size = (m == modules - 1) ? n - base : msize;
    const uint32_t l = i - base;
    row_ptr[i] = i * k;
    const float wv = coupling / static_cast<float>(k);
    for (uint32_t j = 0; j < k; ++j) {
        const int32_t o = (j < k / 2) ? -static_cast<int32_t>(j + 1) : static_cast<int32_t>(j - k / 2 + 1);
        uint32_t src = base + static_cast<uint32_t>((static_cast<int64_t>(l) + o + size) % size);
        const uint64_t h = kmix(seed ^ (static_cast<uint64_t>(i) * 0xD1B54A32D192ED03ull + j));
        const float u = static_cast<float>(h >> 40) * 5.9604644775390625e-08f;
        const uint64_t h2 = kmix(h);
        if (modules > 1 && u < inter) {
            const uint32_t om = (m + 1 + static_cast<uint32_t>(h2 % (modules - 1))) % modules;
            const uint32_t ob = om * msize;
            const uint32_t os = (om == modules - 1) ? n - ob : msize;
            src = ob + static_cast<uint32_t>((h2 >> 32) % os);
        } else if (u < inter + beta) {
            uint32_t r = static_cast<uint32_t>(h2 % (size - 1));
            r += r >= l ? 1u : 0u;
            src = base + r;
        }
        col[static_cast<size_t>(i) * k + j] = src;
        w[static_cast<size_t>(i) * k + j] = wv;
    }
    // omega ~ N(0,1) (Box-Muller on two 24-bit uniforms), theta0 ~ U(-pi, pi)
    const uint64_t a = kmix(seed ^ 0x5851F42D4C957F2Dull ^ (static_cast<uint64_t>(i) << 1));
    const uint64_t b = kmix(a);
    const float u1 = (static_cast<float>(a >> 40) + 1.0f) * 5.9604644775390625e-08f;
    const float u2 = static_cast<float>(b >> 40) * 5.9604644775390625e-08f;
    omega[i] = sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
    theta0[i] = -3.14159265f + 6.2831853f * static_cast<float>(kmix(b) >> 40) * 5.9604644775390625e-08f;
}

template <int G>
void launch_stage(int stage, const KuraArgs& a, hipStream_t s) {
    const uint64_t threads = static_cast<uint64_t>(a.n) * G;
    const uint32_t blocks = static_cast<uint32_t>((threads + 255) / 256);
    switch (stage) {
        case 0: kura_stage_kernel<G, 0><<<blocks, 256, 0, s>>>(a); break;
        case 1: kura_stage_kernel<G, 1><<<blocks, 256, 0, s>>>(a); break;
        case 2: kura_stage_kernel<G, 2><<<blocks, 256, 0, s>>>(a); break;
        case 3: kura_stage_kernel<G, 3><<<blocks, 256, 0, s>>>(a); break;
        default: kura_stage_kernel<G, 4><<<blocks, 256, 0, s>>>(a); break;
    }
}

void launch_stage_g(int g, int stage, const KuraArgs& a, hipStream_t s) {
    switch (g) {
        case 1: launch_stage<1>(stage, a, s); break;
        case 2: launch_stage<2>(stage, a, s); break;
        case 4: launch_stage<4>(stage, a, s); break;
        case 8: launch_stage<8>(stage, a, s); break;
        case 16: launch_stage<16>(stage, a, s); break;
        case 32: launch_stage<32>(stage, a, s); break;
        default: launch_stage<64>(stage, a, s); break;
    }
}

}  // namespace

using simuse::fail;

#define KHIP(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail(SIMUSE_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct simuse_kura {
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    uint64_t n_edges = 0;
    int method = SIMUSE_KURA_RK4;
    int group = 4;
    double dt = 0;
    std::vector<void*> allocs;
    uint32_t *row_ptr = nullptr, *col = nullptr;
    float *w = nullptr, *omega = nullptr;
    float* w_row = nullptr;  // per-row weight when every row's weights are bit-equal
    uint32_t fixed_k = 0;    // common out-degree, or 0
    double* acc = nullptr;
    double* theta = nullptr;
    float2* sc[2] = {nullptr, nullptr};
    double2* part = nullptr;
    float* amp = nullptr;  // amplitudes of phase_coherence
    int cur = 0;
    uint64_t steps = 0;
    uint32_t graph_steps = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool profiling = false;
    std::vector<hipEvent_t> ev;
    uint32_t prof_steps = 0;

    template <class T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
        }
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return SIMUSE_OK;
    }
    void drop_graph() {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
    ~simuse_kura() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    KuraArgs args(int in) const {
        KuraArgs a;
        a.n = n; a.fixed_k = fixed_k; a.row_ptr = row_ptr; a.col = col; a.w = w; a.w_row = w_row;
        a.omega = omega; a.theta = theta; a.acc = acc;
        a.sc_in = sc[in]; a.sc_out = sc[in ^ 1]; a.dt = dt;
        return a;
    }
    // One step = 1 (Euler) or 4 (RK4) stage kernels; the (sin, cos) buffers
    // alternate every stage, so after a step the parity of `cur` flips once
    // per stage.
    int stages() const { return method == SIMUSE_KURA_EULER ? 1 : 4; }
    void enqueue(uint32_t steps_n, int& c) {
        for (uint32_t k = 0; k < steps_n; ++k) {
            if (method == SIMUSE_KURA_EULER) {
                launch_stage_g(group, 0, args(c), stream);
                c ^= 1;
            } else {
                for (int s = 1; s <= 4; ++s) {
                    launch_stage_g(group, s, args(c), stream);
                    c ^= 1;
                }
            }
        }
    }
};

extern "C" {

int simuse_kura_create(const uint32_t* row_ptr, const uint32_t* col, const float* w, const float* omega,
                       const float* theta0, uint32_t n, float dt, int32_t method, simuse_kura** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "null output handle");
    *out = nullptr;
    if (n == 0) return fail(SIMUSE_E_INVALID, "kuramoto: no oscillators");
    if (!row_ptr || !omega || !theta0) return fail(SIMUSE_E_INVALID, "kuramoto: null arrays");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(SIMUSE_E_INVALID, "kuramoto: dt must be finite and > 0");
    if (method != SIMUSE_KURA_EULER && method != SIMUSE_KURA_RK4)
        return fail(SIMUSE_E_INVALID, "kuramoto: unknown method %d", method);
    if (row_ptr[0] != 0) return fail(SIMUSE_E_INVALID, "CSR row_ptr must start at 0");
    for (uint32_t i = 0; i < n; ++i)
        if (row_ptr[i] > row_ptr[i + 1]) return fail(SIMUSE_E_INVALID, "CSR row_ptr must be non-decreasing");
    const uint64_t ne = row_ptr[n];
    if (ne && (!col || !w)) return fail(SIMUSE_E_INVALID, "kuramoto: null CSR arrays");
    for (uint64_t e = 0; e < ne; ++e) {
        if (col[e] >= n) return fail(SIMUSE_E_INVALID, "col index %u >= %u", col[e], n);
        if (!std::isfinite(w[e])) return fail(SIMUSE_E_INVALID, "weight %llu must be finite",
                                              static_cast<unsigned long long>(e));
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(omega[i]) || !std::isfinite(theta0[i]))
            return fail(SIMUSE_E_INVALID, "kuramoto: non-finite omega/theta at %u", i);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "no HIP device available");
    }
    std::unique_ptr<simuse_kura> k(new (std::nothrow) simuse_kura());
    if (!k) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    KHIP(hipStreamCreate(&k->stream));
    k->n = n;
    k->n_edges = ne;
    k->method = method;
    k->dt = static_cast<double>(dt);
    k->group = lanes_for_degree(static_cast<double>(ne) / n);
    int rc;
    if ((rc = k->alloc(&k->row_ptr, n + 1))) return rc;
    if ((rc = k->alloc(&k->col, ne))) return rc;
    if ((rc = k->alloc(&k->w, ne))) return rc;
    if ((rc = k->alloc(&k->omega, n))) return rc;
    if ((rc = k->alloc(&k->acc, n))) return rc;
    if ((rc = k->alloc(&k->theta, n))) return rc;
    if ((rc = k->alloc(&k->sc[0], n))) return rc;
    if ((rc = k->alloc(&k->sc[1], n))) return rc;
    if ((rc = k->alloc(&k->part, 1024))) return rc;
    KHIP(hipMemcpy(k->row_ptr, row_ptr, 4ull * (n + 1), hipMemcpyHostToDevice));
    if (ne) {
        KHIP(hipMemcpy(k->col, col, 4 * ne, hipMemcpyHostToDevice));
        KHIP(hipMemcpy(k->w, w, 4 * ne, hipMemcpyHostToDevice));
    }
    // rows of one weight each (bit-equal) and one common degree
    bool uni_w = true, uni_k = true;
    std::vector<float> wr(n, 0.0f);
    const uint32_t k0 = row_ptr[1] - row_ptr[0];
    for (uint32_t i = 0; i < n && (uni_w || uni_k); ++i) {
        const uint32_t a = row_ptr[i], b = row_ptr[i + 1];
        uni_k = uni_k && b - a == k0;
        if (a < b) {
            wr[i] = w[a];
            for (uint32_t e = a + 1; e < b && uni_w; ++e) uni_w = std::memcmp(&w[e], &w[a], 4) == 0;
        }
    }
    if (uni_k) k->fixed_k = k0;
    if (uni_w && ne) {
        if ((rc = k->alloc(&k->w_row, n))) return rc;
        KHIP(hipMemcpy(k->w_row, wr.data(), 4ull * n, hipMemcpyHostToDevice));
    }
    KHIP(hipMemcpy(k->omega, omega, 4ull * n, hipMemcpyHostToDevice));
    std::vector<double> th(n);
    for (uint32_t i = 0; i < n; ++i) th[i] = static_cast<double>(theta0[i]);
    KHIP(hipMemcpy(k->theta, th.data(), 8ull * n, hipMemcpyHostToDevice));
    kura_prepare_kernel<<<(n + 255) / 256, 256, 0, k->stream>>>(k->theta, k->sc[0], n);
    KHIP(hipGetLastError());
    KHIP(hipStreamSynchronize(k->stream));
    *out = k.release();
    return SIMUSE_OK;
}

int simuse_kura_create_generated(const simuse_kura_gen_desc* g, int32_t method, simuse_kura** out) {
    if (!out || !g) return fail(SIMUSE_E_INVALID, "null argument");
    *out = nullptr;
    if (g->modules == 0 || g->k < 2 || g->k % 2 || g->n / g->modules < g->k + 2)
        return fail(SIMUSE_E_INVALID, "kuramoto generator: need even k >= 2 and modules of at least k+2 nodes");
    if (static_cast<uint64_t>(g->n) * g->k >= 0xFFFFFFFFull)
        return fail(SIMUSE_E_INVALID, "kuramoto generator: n*k exceeds the u32 CSR");
    if (!(g->dt > 0.0f) || !std::isfinite(g->coupling) || !(g->beta >= 0.0f) || !(g->inter >= 0.0f))
        return fail(SIMUSE_E_INVALID, "kuramoto generator: invalid parameters");
    if (method != SIMUSE_KURA_EULER && method != SIMUSE_KURA_RK4)
        return fail(SIMUSE_E_INVALID, "kuramoto: unknown method %d", method);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "no HIP device available");
    }
    std::unique_ptr<simuse_kura> k(new (std::nothrow) simuse_kura());
    if (!k) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    KHIP(hipStreamCreate(&k->stream));
    const uint32_t n = g->n;
    const uint64_t ne = static_cast<uint64_t>(n) * g->k;
    k->n = n;
    k->n_edges = ne;
    k->method = method;
    k->dt = static_cast<double>(g->dt);
    k->group = lanes_for_degree(static_cast<double>(g->k));
    int rc;
    float* th0 = nullptr;
    if ((rc = k->alloc(&k->row_ptr, n + 1))) return rc;
    if ((rc = k->alloc(&k->col, ne))) return rc;
    if ((rc = k->alloc(&k->w, ne))) return rc;
    if ((rc = k->alloc(&k->omega, n))) return rc;
    if ((rc = k->alloc(&k->acc, n))) return rc;
    if ((rc = k->alloc(&k->theta, n))) return rc;
    if ((rc = k->alloc(&k->sc[0], n))) return rc;
    if ((rc = k->alloc(&k->sc[1], n))) return rc;
    if ((rc = k->alloc(&k->part, 1024))) return rc;
    if ((rc = k->alloc(&th0, n))) return rc;
    kura_gen_kernel<<<(n + 256) / 256, 256, 0, k->stream>>>(k->row_ptr, k->col, k->w, k->omega, th0, n, g->modules,
                                                             g->k, g->beta, g->inter, g->coupling, g->seed);
    KHIP(hipGetLastError());
    // the generator gives every row k edges of weight coupling / k (the same
    // float division as the kernel's)
    k->fixed_k = g->k;
    if ((rc = k->alloc(&k->w_row, n))) return rc;
    {
        const std::vector<float> wr(n, g->coupling / static_cast<float>(g->k));
        KHIP(hipMemcpyAsync(k->w_row, wr.data(), 4ull * n, hipMemcpyHostToDevice, k->stream));
        KHIP(hipStreamSynchronize(k->stream));
    }
    std::vector<float> th(n);
    KHIP(hipMemcpyAsync(th.data(), th0, 4ull * n, hipMemcpyDeviceToHost, k->stream));
    KHIP(hipStreamSynchronize(k->stream));
    std::vector<double> thd(th.begin(), th.end());
    KHIP(hipMemcpy(k->theta, thd.data(), 8ull * n, hipMemcpyHostToDevice));
    kura_prepare_kernel<<<(n + 255) / 256, 256, 0, k->stream>>>(k->theta, k->sc[0], n);
    KHIP(hipGetLastError());
    KHIP(hipStreamSynchronize(k->stream));
    *out = k.release();
    return SIMUSE_OK;
}

int simuse_kura_download_graph(simuse_kura* k, uint32_t* row_ptr, uint32_t* col, float* w, float* omega) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    KHIP(hipStreamSynchronize(k->stream));
    if (row_ptr) KHIP(hipMemcpy(row_ptr, k->row_ptr, 4ull * (k->n + 1), hipMemcpyDeviceToHost));
    if (k->n_edges) {
        if (col) KHIP(hipMemcpy(col, k->col, 4 * k->n_edges, hipMemcpyDeviceToHost));
        if (w) KHIP(hipMemcpy(w, k->w, 4 * k->n_edges, hipMemcpyDeviceToHost));
    }
    if (omega) KHIP(hipMemcpy(omega, k->omega, 4ull * k->n, hipMemcpyDeviceToHost));
    return SIMUSE_OK;
}

int simuse_kura_info(simuse_kura* k, uint32_t* n, uint64_t* n_edges) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    if (n) *n = k->n;
    if (n_edges) *n_edges = k->n_edges;
    return SIMUSE_OK;
}

void simuse_kura_destroy(simuse_kura* k) { delete k; }

int simuse_kura_set_group(simuse_kura* k, int32_t lanes) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return fail(SIMUSE_E_INVALID, "lanes per row must be a power of two in 1..64");
    if (k->group != lanes) { k->group = lanes; k->drop_graph(); }
    return SIMUSE_OK;
}

int simuse_kura_step_async(simuse_kura* k, uint32_t n_steps) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    if (n_steps == 0) return SIMUSE_OK;
    if (k->profiling) {
        const int per = k->stages();
        const size_t need = static_cast<size_t>(k->prof_steps + n_steps) * 2;
        while (k->ev.size() < need) {
            hipEvent_t e;
            KHIP(hipEventCreate(&e));
            k->ev.push_back(e);
        }
        for (uint32_t s = 0; s < n_steps; ++s) {
            hipEvent_t* e = k->ev.data() + static_cast<size_t>(k->prof_steps + s) * 2;
            KHIP(hipEventRecord(e[0], k->stream));
            k->enqueue(1, k->cur);
            KHIP(hipEventRecord(e[1], k->stream));
        }
        (void)per;
        k->prof_steps += n_steps;
        k->steps += n_steps;
        KHIP(hipGetLastError());
        return SIMUSE_OK;
    }
    // The graph covers graph_steps steps; since the stage count per step is
    // even for RK4, the buffer parity at graph entry is the same every replay.
    uint32_t full = (k->stages() % 2 == 0) ? n_steps / k->graph_steps : 0;
    uint32_t rest = n_steps - full * k->graph_steps;
    if (full) {
        if (!k->gexec) {
            KHIP(hipStreamBeginCapture(k->stream, hipStreamCaptureModeThreadLocal));
            int c = k->cur;
            k->enqueue(k->graph_steps, c);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(k->stream, &g);
            if (e != hipSuccess) return fail(SIMUSE_E_HIP, "kuramoto capture: %s", hipGetErrorString(e));
            k->graph = g;
            e = hipGraphInstantiate(&k->gexec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) { k->drop_graph(); return fail(SIMUSE_E_HIP, "kuramoto instantiate: %s", hipGetErrorString(e)); }
        }
        for (uint32_t r = 0; r < full; ++r) KHIP(hipGraphLaunch(k->gexec, k->stream));
    }
    if (rest) k->enqueue(rest, k->cur);
    k->steps += n_steps;
    KHIP(hipGetLastError());
    return SIMUSE_OK;
}

int simuse_kura_sync(simuse_kura* k) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    KHIP(hipStreamSynchronize(k->stream));
    return SIMUSE_OK;
}

int simuse_kura_download_phases(simuse_kura* k, double* out, uint32_t n) {
    if (!k || !out || n != k->n) return fail(SIMUSE_E_INVALID, "phases buffer must hold n values");
    KHIP(hipMemcpyAsync(out, k->theta, 8ull * n, hipMemcpyDeviceToHost, k->stream));
    KHIP(hipStreamSynchronize(k->stream));
    return SIMUSE_OK;
}

static int kura_coherence(simuse_kura* k, const float* amp_dev, double* s_out, double* c_out) {
    const uint32_t blocks = std::min<uint32_t>(1024, (k->n + 255) / 256);
    kura_order_partial<<<blocks, 256, 0, k->stream>>>(k->theta, amp_dev, k->n, k->part);
    KHIP(hipGetLastError());
    std::vector<double2> p(blocks);
    KHIP(hipMemcpyAsync(p.data(), k->part, sizeof(double2) * blocks, hipMemcpyDeviceToHost, k->stream));
    KHIP(hipStreamSynchronize(k->stream));
    double s = 0, c = 0;
    for (const auto& x : p) { s += x.x; c += x.y; }
    *s_out = s;
    *c_out = c;
    return SIMUSE_OK;
}

int simuse_kura_order_parameter(simuse_kura* k, double* r) {
    if (!k || !r) return fail(SIMUSE_E_INVALID, "null argument");
    double s, c;
    int rc = kura_coherence(k, nullptr, &s, &c);
    if (rc) return rc;
    *r = std::sqrt(s * s + c * c) / k->n;
    return SIMUSE_OK;
}

int simuse_kura_phase_coherence(simuse_kura* k, const float* amplitudes, double* r) {
    if (!k || !r || !amplitudes) return fail(SIMUSE_E_INVALID, "null argument");
    if (!k->amp) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(float) * k->n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "phase_coherence: %s", hipGetErrorString(e));
        }
        k->allocs.push_back(q);
        k->amp = static_cast<float*>(q);
    }
    double sa = 0;
    for (uint32_t i = 0; i < k->n; ++i) sa += amplitudes[i];
    KHIP(hipMemcpyAsync(k->amp, amplitudes, sizeof(float) * k->n, hipMemcpyHostToDevice, k->stream));
    double s, c;
    int rc = kura_coherence(k, k->amp, &s, &c);
    if (rc) return rc;
    *r = std::sqrt(s * s + c * c) / std::max(sa, 1e-12);
    return SIMUSE_OK;
}

int simuse_kura_set_profiling(simuse_kura* k, int32_t on) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    KHIP(hipStreamSynchronize(k->stream));
    k->profiling = on != 0;
    k->prof_steps = 0;
    return SIMUSE_OK;
}

int simuse_kura_kernel_times(simuse_kura* k, float* ms, int32_t n) {
    if (!k || !ms || n < 1) return fail(SIMUSE_E_INVALID, "null argument");
    KHIP(hipStreamSynchronize(k->stream));
    double acc = 0;
    for (uint32_t s = 0; s < k->prof_steps; ++s) {
        float t = 0;
        KHIP(hipEventElapsedTime(&t, k->ev[2 * s], k->ev[2 * s + 1]));
        acc += t;
    }
    ms[0] = k->prof_steps ? static_cast<float>(acc / k->prof_steps) : 0.0f;
    k->prof_steps = 0;
    return SIMUSE_OK;
}

int simuse_kura_step_counter(simuse_kura* k, uint64_t* steps) {
    if (!k) return fail(SIMUSE_E_INVALID, "null handle");
    if (steps) *steps = k->steps;
    return SIMUSE_OK;
}

}  // extern "C"
