// pool.hip — the creature pool: many small alpha-LIF brains stepped as one
// block-diagonal network with a deterministic per-post pull scatter.
//
// Drop-in for the game engine's BrainPool (DE/src/brain/dynamics/
// gpu_pool.rs:1-260) over WgpuNetwork::step_many (PW/src/network.rs:802-885)
// and its shaders (PW/shaders/csr_scatter.wgsl).  DE/ = crates/dioma-engine/,
// PW/ = crates/SiMuse/crates/pillars-wgpu/.
//
// Every substep does, one thread per neuron (post):
//   1. at the first substep of a submit, v += ext (v_jump_kernel);
//   2. I = ring[cur][post] (copy_slot_kernel), alpha-LIF update, spike flag
//      written to the history row and, by wave ballot, to this step's bit
//      plane;
//   3. ring[cur][post] = 0, then a pull over the post's in-edges of the
//      PREVIOUS step's spike bits: per delay bucket b, acc = fma(w, sign, acc)
//      in ascending source order, ring[(cur + b) % 32][post] += acc when
//      acc != 0 (scatter_kernel).
// Step 3 reads only the previous substep's bit plane, and every ring cell is
// touched by its own post's thread, so the result is independent of the
// schedule and bit-identical to the reference's sequence of dispatches.
// When every creature fits one workgroup (at most 1024 neurons), a submit
// chunk of up to 64 substeps is ONE launch, one block per creature, with the
// substeps separated by block barriers (pool_tick_kernel); otherwise each
// substep is one launch over the whole pool (pool_step_kernel).  The
// signed weight w*sign is formed once per weight upload (exact: sign = +-1),
// so fma(w, sign, acc) = acc + w*sign is a single rounding either way.
//
// HBM layout: in-edges per post sorted by (delay bucket, source), packed as
// source | bucket << 27 (4 B), the signed weight in the same order (4 B);
// ring [32][n] fp32; state v, g, refr_cnt [n] each; spike bits [2][n/64]
// (by step parity; one-launch chunks keep each creature's last plane in its
// own bits instead, [creatures][16] u64); history [64][n] u32 (one row per
// substep of a submit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/simuse.h"
#include "device_math.hpp"

namespace simuse {
int fail(int code, const char* fmt, ...);
}
using simuse::fail;

namespace {

constexpr uint32_t kPoolTile = 256;
constexpr uint32_t kSlotsP = 32;                 // MAX_SLOTS, PW/src/network.rs:15
constexpr uint32_t kMaxSubmit = SIMUSE_POOL_MAX_STEPS_PER_SUBMIT;  // MAX_HIST / GPU_MAX_STEPS_PER_SUBMIT
constexpr uint32_t kSrcBits = 27;
constexpr uint32_t kSrcMask = (1u << kSrcBits) - 1u;
constexpr uint32_t kPullBatch = 8;  // in-edges whose loads a pull issues together

struct PoolArgs {
    uint32_t n, words;
    const uint32_t* rev_ptr;  // [n + 1]
    const uint32_t* rev;      // source | bucket << 27, by (post, bucket, source)
    const float* ws;          // signed weight of each in-edge
    float* ring;              // [32][n]
    float *v, *g, *refr;
    uint64_t* bits;           // [2][words]
    uint32_t* hist;           // [64][n]
    const float* ext;         // [n]
    const uint64_t* step;     // substeps done before this submit
    float tf, v_rest, v_th, v_reset, refr_steps, syn_decay;
};

__global__ __launch_bounds__(kPoolTile) void pool_step_kernel(PoolArgs a, uint32_t k) {
    const uint32_t i = blockIdx.x * kPoolTile + threadIdx.x;
    const uint64_t s = *a.step + k;
    const uint32_t cur = static_cast<uint32_t>(s & (kSlotsP - 1));
    const uint32_t par = static_cast<uint32_t>(s & 1);
    const bool live = i < a.n;
    const size_t ci = static_cast<size_t>(cur) * a.n + i;
    int spike = 0;
    if (live) {
        float v = a.v[i], g = a.g[i], r = a.refr[i];
        if (k == 0) v = v + a.ext[i];
        const float I = a.ring[ci];
        spike = simuse::alpha_lif_update(v, g, r, I, a.tf, a.v_rest, a.v_th, a.v_reset, a.refr_steps, a.syn_decay);
        a.v[i] = v;
        a.g[i] = g;
        a.refr[i] = r;
        a.hist[static_cast<size_t>(k) * a.n + i] = static_cast<uint32_t>(spike);
    }
    const uint64_t b = __ballot(spike);
    if ((threadIdx.x & 63) == 0 && (i >> 6) < a.words) a.bits[par * a.words + (i >> 6)] = b;
    if (!live) return;
    a.ring[ci] = 0.0f;
    const uint64_t* prev = a.bits + (par ^ 1u) * a.words;
    const uint32_t e1 = a.rev_ptr[i + 1];
    float acc = 0.0f;
    uint32_t cb = kSlotsP;  // no bucket yet
    // kPullBatch in-edges at a time: their packed sources and signed weights
    // load together, then the sources' spike words, then the adds in edge
    // order -- two dependent loads per batch instead of up to three per edge
    // (measured: 14.1 -> 11.0 us per substep on the 64 x 1000 pool; 16 per
    // batch costs occupancy, 13.3 us).
    for (uint32_t e0 = a.rev_ptr[i]; e0 < e1; e0 += kPullBatch) {
        uint32_t x[kPullBatch];
        float w[kPullBatch];
        uint64_t word[kPullBatch];
#pragma unroll
        for (uint32_t j = 0; j < kPullBatch; ++j) {
            x[j] = 0u;
            w[j] = 0.0f;
            if (e0 + j < e1) {
                x[j] = a.rev[e0 + j];
                w[j] = a.ws[e0 + j];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kPullBatch; ++j) word[j] = e0 + j < e1 ? prev[(x[j] & kSrcMask) >> 6] : 0ull;
#pragma unroll
        for (uint32_t j = 0; j < kPullBatch; ++j) {
            if (e0 + j >= e1) break;
            const uint32_t src = x[j] & kSrcMask, bk = x[j] >> kSrcBits;
            if (bk != cb) {
                if (acc != 0.0f) a.ring[static_cast<size_t>((cur + cb) & (kSlotsP - 1)) * a.n + i] += acc;
                acc = 0.0f;
                cb = bk;
            }
            if ((word[j] >> (src & 63)) & 1ull) acc = acc + w[j];
        }
    }
    if (acc != 0.0f) a.ring[static_cast<size_t>((cur + cb) & (kSlotsP - 1)) * a.n + i] += acc;
}

// One submit chunk of a pool whose creatures all fit one workgroup: block c
// runs creature c's `chunk` substeps in one launch.  The creatures are
// block-diagonal (every in-edge of a creature's neuron comes from the same
// creature), so a substep needs only the block's own previous spike plane,
// and substeps are separated by a block barrier instead of a kernel
// boundary.  Thread t is the creature's neuron t: v, g, refr_cnt in
// registers for the whole chunk, its column of the 32-slot ring in LDS
// (each ring cell is touched by its own post only), its first kEdgeRegs
// in-edges (local source | bucket, signed weight) in registers; the spike
// planes (two, by step parity) of the creature's local bits in LDS.  Every
// neuron does pool_step_kernel's arithmetic in pool_step_kernel's order, so
// the results are the same bits.  The ring, the state and the last plane go
// back to memory at the end; `lbits` carries the last plane to the next
// launch.
constexpr uint32_t kCreatureMax = 1024;           // neurons of a creature the one-launch chunk takes
constexpr uint32_t kCreatureWords = kCreatureMax / 64;
constexpr uint32_t kEdgeRegs = 32;                // in-edges per neuron held in registers
constexpr uint32_t kPlaneWords = 64;              // u32 per plane: 32 of spike bits, word 32 stays zero

// A register-held in-edge: bits 0-4 the source's bit in its 32-bit plane
// word, 5-9 the delay bucket, 16-31 the byte offset of that word in a plane.
// Unused slots point at the plane's zero word with the last edge's bucket.
__device__ __forceinline__ uint32_t tick_edge(uint32_t src, uint32_t bk) {
    return (src & 31u) | (bk << 5) | (((src >> 5) * 4u) << 16);
}

struct TickLds {
    uint64_t plane[2][kPlaneWords / 2];  // by step parity, the creature's local bits
    float ring[kSlotsP * kCreatureMax];  // [slot][neuron]
};

template <bool kOneBucket>
__global__ __launch_bounds__(kCreatureMax, 1) void pool_tick_kernel(PoolArgs a, const uint32_t* __restrict__ cre_off,
                                                                    uint64_t* __restrict__ lbits, uint32_t chunk,
                                                                    uint32_t bucket0) {
    __shared__ TickLds L;
    const uint32_t c = blockIdx.x, t = threadIdx.x, lane = t & 63u;
    const uint32_t off = cre_off[c], n = cre_off[c + 1] - off;
    const uint32_t nw = (n + 63) / 64;
    const bool live = t < n;
    const uint32_t i = off + t;
    const uint64_t s0 = *a.step;
#pragma unroll 4
    for (uint32_t q = 0; q < kSlotsP; ++q)
        L.ring[q * kCreatureMax + t] = live ? a.ring[static_cast<size_t>(q) * a.n + i] : 0.0f;
    if (t < kPlaneWords / 2) {
        // the plane of substep s0 - 1 (read by substep s0's pull) from the last launch; the rest zero
        const uint32_t pp = static_cast<uint32_t>((s0 - 1) & 1u);
        L.plane[pp][t] = t < nw ? lbits[static_cast<size_t>(c) * kCreatureWords + t] : 0ull;
        L.plane[pp ^ 1u][t] = 0ull;
    }
    float v = 0.0f, g = 0.0f, r = 0.0f;
    uint32_t e0 = 0u, e1 = 0u;
    if (live) {
        v = a.v[i];
        g = a.g[i];
        r = a.refr[i];
        e0 = a.rev_ptr[i];
        e1 = a.rev_ptr[i + 1];
    }
    const uint32_t nreg = min(e1 - e0, kEdgeRegs);
    uint32_t xs[kEdgeRegs];
    float wv[kEdgeRegs];
    uint32_t lastb = bucket0;
#pragma unroll
    for (uint32_t j = 0; j < kEdgeRegs; ++j) {
        wv[j] = 0.0f;
        if (j < nreg) {
            const uint32_t x = a.rev[e0 + j];
            lastb = x >> kSrcBits;
            xs[j] = tick_edge((x & kSrcMask) - off, lastb);
            wv[j] = a.ws[e0 + j];
        } else {
            xs[j] = tick_edge(32u * 32u, lastb);  // the zero word: never a spike
        }
    }
    __syncthreads();
    for (uint32_t k = 0; k < chunk; ++k) {
        const uint64_t s = s0 + k;
        const uint32_t cur = static_cast<uint32_t>(s & (kSlotsP - 1));
        const uint32_t par = static_cast<uint32_t>(s & 1);
        float* ring_t = L.ring + t;  // slot q of this neuron: ring_t[q * kCreatureMax]
        // the packed in-edges are opaque to the loop, so that the compiler
        // does not hoist their unpacked fields (several registers per edge)
        // out of it and spill them
#pragma unroll
        for (uint32_t j = 0; j < kEdgeRegs; ++j) asm volatile("" : "+v"(xs[j]));
        int spike = 0;
        if (live) {
            if (k == 0) v = v + a.ext[i];
            const float I = ring_t[cur * kCreatureMax];
            spike = simuse::alpha_lif_update(v, g, r, I, a.tf, a.v_rest, a.v_th, a.v_reset, a.refr_steps,
                                             a.syn_decay);
            a.hist[static_cast<size_t>(k) * a.n + i] = static_cast<uint32_t>(spike);
        }
        const uint64_t b = __ballot(spike);
        if (lane == 0) L.plane[par][t >> 6] = b;
        if (live) {
            ring_t[cur * kCreatureMax] = 0.0f;
            const char* prev = reinterpret_cast<const char*>(L.plane[par ^ 1u]);
            float acc = 0.0f;
            uint32_t cb = kOneBucket ? bucket0 : kSlotsP;  // kSlotsP: no bucket yet
#pragma unroll
            for (uint32_t j = 0; j < kEdgeRegs; ++j) {
                const uint32_t word = *reinterpret_cast<const uint32_t*>(prev + (xs[j] >> 16));
                if (!kOneBucket) {
                    const uint32_t bk = (xs[j] >> 5) & 31u;
                    if (bk != cb) {
                        if (acc != 0.0f) ring_t[((cur + cb) & (kSlotsP - 1)) * kCreatureMax] += acc;
                        acc = 0.0f;
                        cb = bk;
                    }
                }
                // bit field extract at offset xs[j] (the instruction reads the offset's low 5 bits)
                if (__builtin_amdgcn_ubfe(word, xs[j], 1u)) acc = acc + wv[j];
            }
            // in-edges beyond the registers' share, from memory in the same order
            const uint32_t* prev32 = reinterpret_cast<const uint32_t*>(prev);
            for (uint32_t e = e0 + kEdgeRegs; e < e1; ++e) {
                const uint32_t x = a.rev[e];
                const float w = a.ws[e];
                const uint32_t src = (x & kSrcMask) - off, bk = x >> kSrcBits;
                if (!kOneBucket && bk != cb) {
                    if (acc != 0.0f) ring_t[((cur + cb) & (kSlotsP - 1)) * kCreatureMax] += acc;
                    acc = 0.0f;
                    cb = bk;
                }
                if ((prev32[src >> 5] >> (src & 31u)) & 1u) acc = acc + w;
            }
            if (acc != 0.0f) ring_t[((cur + cb) & (kSlotsP - 1)) * kCreatureMax] += acc;
        }
        __syncthreads();  // the plane just read is the next substep's to write
    }
    if (live) {
        a.v[i] = v;
        a.g[i] = g;
        a.refr[i] = r;
    }
#pragma unroll 4
    for (uint32_t q = 0; q < kSlotsP; ++q)
        if (live) a.ring[static_cast<size_t>(q) * a.n + i] = L.ring[q * kCreatureMax + t];
    if (chunk && t < nw)
        lbits[static_cast<size_t>(c) * kCreatureWords + t] = L.plane[static_cast<uint32_t>((s0 + chunk - 1) & 1u)][t];
}

__global__ void pool_advance_kernel(uint64_t* step, uint32_t n) { *step += n; }

// ws[e] = w[rev_edge[e]] * sign[source]: exact, sign is +-1.
__global__ void pool_sign_weights_kernel(const float* w, const uint32_t* rev_edge, const uint32_t* rev,
                                         const float* sign, float* ws, uint32_t n_syn) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n_syn) ws[e] = w[rev_edge[e]] * sign[rev[e] & kSrcMask];
}

__global__ void pool_init_kernel(float* v, float* g, float* refr, uint32_t n, float v0, float r0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        v[i] = v0;
        g[i] = 0.0f;
        refr[i] = r0;
    }
}

#define PHIP(expr)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(SIMUSE_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct simuse_pool {
    simuse_pool_config c{};
    uint8_t delay = 0;
    // BrainPool's concatenated CSR and per-synapse delays (gpu_pool.rs:101-128)
    std::vector<uint32_t> row_ptr{0}, col;
    std::vector<float> values, nt;
    std::vector<uint8_t> delays;
    std::vector<simuse_pool_slot> slots;
    uint32_t n = 0;
    bool needs_rebuild = false;
    // device side
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    uint32_t *rev_ptr = nullptr, *rev = nullptr, *rev_edge = nullptr, *hist = nullptr;
    float *w = nullptr, *sign = nullptr, *ws = nullptr, *ring = nullptr, *ext = nullptr;
    float *v = nullptr, *g = nullptr, *refr = nullptr;
    uint64_t *bits = nullptr, *d_step = nullptr;
    uint32_t words = 0;
    // every creature within one workgroup: a submit chunk is one launch
    // (pool_tick_kernel) over the creatures' offsets, with each creature's
    // last spike plane in lbits; otherwise one launch per substep
    bool one_launch = false;
    bool one_bucket = false;  // every in-edge in delay bucket bucket0 (a pool-wide delay)
    uint32_t bucket0 = 0;
    uint32_t *cre_off = nullptr;
    uint64_t* lbits = nullptr;
    uint64_t steps = 0;
    float refr_steps = 0, tf = 0, syn_decay = 0;
    std::vector<float> ext_full;
    struct Graph {
        uint32_t chunk;
        hipGraph_t g;
        hipGraphExec_t x;
    };
    std::vector<Graph> graphs;

    template <class T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "pool: hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
        }
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return SIMUSE_OK;
    }
    void drop_graphs() {
        for (Graph& gr : graphs) {
            (void)hipGraphExecDestroy(gr.x);
            (void)hipGraphDestroy(gr.g);
        }
        graphs.clear();
    }
    void drop_device() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
    }
    ~simuse_pool() {
        drop_device();
        if (stream) (void)hipStreamDestroy(stream);
    }
    bool built() const { return d_step != nullptr && !needs_rebuild; }

    PoolArgs args() const {
        PoolArgs a;
        a.n = n; a.words = words;
        a.rev_ptr = rev_ptr; a.rev = rev; a.ws = ws; a.ring = ring;
        a.v = v; a.g = g; a.refr = refr; a.bits = bits; a.hist = hist; a.ext = ext; a.step = d_step;
        a.tf = tf; a.v_rest = c.v_rest; a.v_th = c.v_thresh; a.v_reset = c.v_reset;
        a.refr_steps = refr_steps; a.syn_decay = syn_decay;
        return a;
    }

    int upload_weights() {
        const uint32_t ns = static_cast<uint32_t>(values.size());
        if (ns == 0) return SIMUSE_OK;
        PHIP(hipMemcpyAsync(w, values.data(), sizeof(float) * ns, hipMemcpyHostToDevice, stream));
        pool_sign_weights_kernel<<<(ns + 255) / 256, 256, 0, stream>>>(w, rev_edge, rev, sign, ws, ns);
        PHIP(hipGetLastError());
        PHIP(hipStreamSynchronize(stream));
        return SIMUSE_OK;
    }

    // ensure_built (gpu_pool.rs:205-260) and WgpuNetwork::new's in-edge CSR
    // (PW/src/network.rs:231-257): a fresh network, every creature at the
    // initial state v = v_rest, g = 0, refr_cnt = refr_steps.
    int build() {
        if (built()) return SIMUSE_OK;
        if (n == 0) return fail(SIMUSE_E_STATE, "pool: no creature registered");
        drop_device();
        refr_steps = std::max(std::round(c.tau_ref / c.dt_ms), 1.0f);
        tf = c.dt_ms / c.tau_m;
        syn_decay = 1.0f - c.dt_ms / c.tau_syn;
        const uint32_t ns = static_cast<uint32_t>(values.size());
        // in-edges of each post: ascending source, then row order, stably
        // grouped by delay bucket max(d, 1) % 32
        auto bucket = [&](uint32_t e) { return std::max<uint32_t>(delays[e], 1u) % kSlotsP; };
        std::vector<uint32_t> h_rev(ns), h_edge(ns), h_ptr(static_cast<size_t>(n) + 1, 0);
        for (uint32_t e = 0; e < ns; ++e) h_ptr[col[e] + 1] += 1;
        for (uint32_t p = 0; p < n; ++p) h_ptr[p + 1] += h_ptr[p];
        {
            std::vector<uint32_t> cur(h_ptr.begin(), h_ptr.end() - 1);
            for (uint32_t s = 0; s < n; ++s)
                for (uint32_t e = row_ptr[s]; e < row_ptr[s + 1]; ++e) {
                    const uint32_t k = cur[col[e]]++;
                    h_rev[k] = s | (bucket(e) << kSrcBits);
                    h_edge[k] = e;
                }
        }
        // posts whose in-edges carry several buckets: stable sort by bucket
        std::vector<std::pair<uint32_t, uint32_t>> seg;
        for (uint32_t p = 0; p < n; ++p) {
            const uint32_t a = h_ptr[p], z = h_ptr[p + 1];
            bool mixed = false;
            for (uint32_t k = a + 1; k < z && !mixed; ++k) mixed = (h_rev[k] >> kSrcBits) != (h_rev[a] >> kSrcBits);
            if (!mixed) continue;
            seg.clear();
            for (uint32_t k = a; k < z; ++k) seg.emplace_back(h_rev[k], h_edge[k]);
            std::stable_sort(seg.begin(), seg.end(), [](const std::pair<uint32_t, uint32_t>& x,
                                                        const std::pair<uint32_t, uint32_t>& y) {
                return (x.first >> kSrcBits) < (y.first >> kSrcBits);
            });
            for (uint32_t k = a; k < z; ++k) {
                h_rev[k] = seg[k - a].first;
                h_edge[k] = seg[k - a].second;
            }
        }
        words = (n + 63) / 64;
        one_launch = true;
        for (const simuse_pool_slot& sl : slots) one_launch = one_launch && sl.n <= kCreatureMax;
        one_bucket = true;
        bucket0 = ns ? bucket(0) : 0u;
        for (uint32_t e = 1; e < ns && one_bucket; ++e) one_bucket = bucket(e) == bucket0;
        const uint32_t nc = static_cast<uint32_t>(slots.size());
        std::vector<uint32_t> h_off(static_cast<size_t>(nc) + 1, 0);
        for (uint32_t k = 0; k < nc; ++k) h_off[k + 1] = h_off[k] + slots[k].n;
        int rc;
        if ((rc = alloc(&rev_ptr, static_cast<size_t>(n) + 1)) || (rc = alloc(&rev, ns)) ||
            (rc = alloc(&rev_edge, ns)) || (rc = alloc(&w, ns)) || (rc = alloc(&ws, ns)) || (rc = alloc(&sign, n)) ||
            (rc = alloc(&ring, static_cast<size_t>(kSlotsP) * n)) || (rc = alloc(&ext, n)) || (rc = alloc(&v, n)) ||
            (rc = alloc(&g, n)) || (rc = alloc(&refr, n)) || (rc = alloc(&bits, 2 * static_cast<size_t>(words))) ||
            (rc = alloc(&hist, static_cast<size_t>(kMaxSubmit) * n)) || (rc = alloc(&d_step, 1)) ||
            (rc = alloc(&cre_off, static_cast<size_t>(nc) + 1)) ||
            (rc = alloc(&lbits, static_cast<size_t>(nc) * kCreatureWords)))
            return rc;
        PHIP(hipMemcpyAsync(cre_off, h_off.data(), 4 * h_off.size(), hipMemcpyHostToDevice, stream));
        PHIP(hipMemsetAsync(lbits, 0, 8 * static_cast<size_t>(nc) * kCreatureWords, stream));
        PHIP(hipMemcpyAsync(rev_ptr, h_ptr.data(), 4 * h_ptr.size(), hipMemcpyHostToDevice, stream));
        if (ns) {
            PHIP(hipMemcpyAsync(rev, h_rev.data(), 4 * static_cast<size_t>(ns), hipMemcpyHostToDevice, stream));
            PHIP(hipMemcpyAsync(rev_edge, h_edge.data(), 4 * static_cast<size_t>(ns), hipMemcpyHostToDevice, stream));
        }
        PHIP(hipMemcpyAsync(sign, nt.data(), 4 * static_cast<size_t>(n), hipMemcpyHostToDevice, stream));
        PHIP(hipMemsetAsync(ring, 0, sizeof(float) * kSlotsP * static_cast<size_t>(n), stream));
        PHIP(hipMemsetAsync(bits, 0, 16 * static_cast<size_t>(words), stream));
        PHIP(hipMemsetAsync(d_step, 0, 8, stream));
        pool_init_kernel<<<(n + 255) / 256, 256, 0, stream>>>(v, g, refr, n, c.v_rest, refr_steps);
        PHIP(hipGetLastError());
        steps = 0;
        ext_full.assign(n, 0.0f);
        needs_rebuild = false;
        return upload_weights();  // synchronises
    }

    void enqueue(uint32_t chunk) {
        const PoolArgs a = args();
        if (one_launch) {
            if (one_bucket)
                pool_tick_kernel<true><<<static_cast<uint32_t>(slots.size()), kCreatureMax, 0, stream>>>(
                    a, cre_off, lbits, chunk, bucket0);
            else
                pool_tick_kernel<false><<<static_cast<uint32_t>(slots.size()), kCreatureMax, 0, stream>>>(
                    a, cre_off, lbits, chunk, bucket0);
        } else {
            const uint32_t blocks = (n + kPoolTile - 1) / kPoolTile;
            for (uint32_t k = 0; k < chunk; ++k) pool_step_kernel<<<blocks, kPoolTile, 0, stream>>>(a, k);
        }
        pool_advance_kernel<<<1, 1, 0, stream>>>(d_step, chunk);
    }
    int run_chunk(uint32_t chunk) {
        for (Graph& gr : graphs)
            if (gr.chunk == chunk) {
                PHIP(hipGraphLaunch(gr.x, stream));
                return SIMUSE_OK;
            }
        PHIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        enqueue(chunk);
        hipGraph_t gg = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &gg);
        if (e != hipSuccess) return fail(SIMUSE_E_HIP, "pool capture: %s", hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, gg, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(gg);
            return fail(SIMUSE_E_HIP, "pool instantiate: %s", hipGetErrorString(e));
        }
        graphs.push_back({chunk, gg, x});
        PHIP(hipGraphLaunch(x, stream));
        return SIMUSE_OK;
    }
};

extern "C" {

int simuse_pool_create(const simuse_pool_config* cfg, int32_t delay_steps, simuse_pool** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "pool_create: null out");
    *out = nullptr;
    // pool_config, gpu_pool.rs:13-24
    simuse_pool_config c = {0.5f, 2.2f, 20.0f, -52.0f, -45.0f, -52.0f, 3.0f, 1.8f};
    if (cfg) c = *cfg;
    const float f[8] = {c.dt_ms, c.tau_ref, c.tau_m, c.v_rest, c.v_thresh, c.v_reset, c.tau_syn, c.t_delay};
    for (float x : f)
        if (!std::isfinite(x)) return fail(SIMUSE_E_INVALID, "pool_create: non-finite configuration value");
    if (!(c.dt_ms > 0.0f) || !(c.tau_m > 0.0f) || !(c.tau_syn > 0.0f))
        return fail(SIMUSE_E_INVALID, "pool_create: dt_ms, tau_m and tau_syn must be positive");
    // BrainPool::new, gpu_pool.rs:63-90: round(t_delay / dt) + 1 (or the
    // override), capped at 31
    const float sd = std::max(std::round(c.t_delay / c.dt_ms), 0.0f);
    if (delay_steps < 0 && !(sd < 1e6f)) return fail(SIMUSE_E_INVALID, "pool_create: t_delay / dt_ms too large");
    const uint32_t d = delay_steps >= 0 ? static_cast<uint32_t>(delay_steps) : static_cast<uint32_t>(sd) + 1u;
    simuse_pool* p = new (std::nothrow) simuse_pool();
    if (!p) return fail(SIMUSE_E_NOMEM, "pool_create: host allocation");
    p->c = c;
    p->delay = static_cast<uint8_t>(std::min(d, 31u));
    hipError_t e = hipStreamCreate(&p->stream);
    if (e != hipSuccess) {
        delete p;
        return fail(SIMUSE_E_HIP, "pool_create: hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = p;
    return SIMUSE_OK;
}

void simuse_pool_destroy(simuse_pool* p) { delete p; }

int simuse_pool_register(simuse_pool* p, uint32_t n, const int32_t* row_ptr, const int32_t* col_indices,
                         const float* values, const float* nt_types, simuse_pool_slot* slot) {
    if (!p) return fail(SIMUSE_E_INVALID, "pool_register: null pool");
    if (!row_ptr || (n && !nt_types)) return fail(SIMUSE_E_INVALID, "pool_register: null arrays");
    if (row_ptr[0] != 0) return fail(SIMUSE_E_INVALID, "pool_register: row_ptr[0] must be 0");
    for (uint32_t i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return fail(SIMUSE_E_INVALID, "pool_register: row_ptr decreases at %u", i);
    const uint64_t ns = static_cast<uint64_t>(row_ptr[n]);
    if (ns && (!col_indices || !values)) return fail(SIMUSE_E_INVALID, "pool_register: null arrays");
    for (uint64_t e = 0; e < ns; ++e)
        if (col_indices[e] < 0 || static_cast<uint64_t>(col_indices[e]) >= n)
            return fail(SIMUSE_E_INVALID, "pool_register: column %lld of edge %llu outside [0, %u)",
                        static_cast<long long>(col_indices[e]), static_cast<unsigned long long>(e), n);
    const uint64_t n_total = static_cast<uint64_t>(p->n) + n;
    const uint64_t s_total = p->values.size() + ns;
    if (n_total >= (1ull << kSrcBits))
        return fail(SIMUSE_E_UNSUPPORTED, "pool_register: %llu neurons exceed the pool's 2^27",
                    static_cast<unsigned long long>(n_total));
    if (s_total >= (1ull << 32))
        return fail(SIMUSE_E_UNSUPPORTED, "pool_register: %llu synapses exceed 2^32",
                    static_cast<unsigned long long>(s_total));
    const uint32_t off = p->n, syn_off = static_cast<uint32_t>(p->values.size());
    p->row_ptr.pop_back();
    for (uint32_t i = 0; i <= n; ++i) p->row_ptr.push_back(static_cast<uint32_t>(row_ptr[i]) + syn_off);
    for (uint64_t e = 0; e < ns; ++e) p->col.push_back(static_cast<uint32_t>(col_indices[e]) + off);
    p->values.insert(p->values.end(), values, values + ns);
    p->delays.insert(p->delays.end(), ns, p->delay);
    for (uint32_t i = 0; i < n; ++i) p->nt.push_back(nt_types[i] < 0.0f ? -1.0f : 1.0f);
    const simuse_pool_slot s = {off, n, syn_off, ns};
    p->slots.push_back(s);
    p->n = static_cast<uint32_t>(n_total);
    p->needs_rebuild = true;
    if (slot) *slot = s;
    return SIMUSE_OK;
}

int simuse_pool_set_delays(simuse_pool* p, uint32_t creature, const uint8_t* d, uint64_t n_syn) {
    if (!p) return fail(SIMUSE_E_INVALID, "pool_set_delays: null pool");
    if (creature >= p->slots.size()) return fail(SIMUSE_E_INVALID, "pool_set_delays: creature %u out of range", creature);
    const simuse_pool_slot& s = p->slots[creature];
    if (n_syn != s.n_syn) return fail(SIMUSE_E_INVALID, "pool_set_delays: %llu delays for %llu synapses",
                                      static_cast<unsigned long long>(n_syn), static_cast<unsigned long long>(s.n_syn));
    if (n_syn && !d) return fail(SIMUSE_E_INVALID, "pool_set_delays: null delays");
    std::copy(d, d + n_syn, p->delays.begin() + s.syn_off);
    p->needs_rebuild = true;
    return SIMUSE_OK;
}

int simuse_pool_n_creatures(simuse_pool* p, uint32_t* n) {
    if (!p || !n) return fail(SIMUSE_E_INVALID, "pool_n_creatures: null argument");
    *n = static_cast<uint32_t>(p->slots.size());
    return SIMUSE_OK;
}

int simuse_pool_slots(simuse_pool* p, simuse_pool_slot* out, uint32_t cap) {
    if (!p || (cap && !out)) return fail(SIMUSE_E_INVALID, "pool_slots: null argument");
    if (cap < p->slots.size()) return fail(SIMUSE_E_INVALID, "pool_slots: room for %u of %zu slots", cap, p->slots.size());
    std::copy(p->slots.begin(), p->slots.end(), out);
    return SIMUSE_OK;
}

int simuse_pool_sync_weights(simuse_pool* p, const float* const* mirrors, const uint8_t* dirty, uint32_t n_creatures) {
    if (!p) return fail(SIMUSE_E_INVALID, "pool_sync_weights: null pool");
    if (n_creatures != p->slots.size())
        return fail(SIMUSE_E_INVALID, "pool_sync_weights: %u mirrors for %zu creatures", n_creatures, p->slots.size());
    if (n_creatures && (!mirrors || !dirty)) return fail(SIMUSE_E_INVALID, "pool_sync_weights: null arrays");
    for (uint32_t k = 0; k < n_creatures; ++k)
        if (dirty[k] && p->slots[k].n_syn && !mirrors[k])
            return fail(SIMUSE_E_INVALID, "pool_sync_weights: creature %u dirty without weights", k);
    int rc = p->build();
    if (rc) return rc;
    bool any = false;
    for (uint32_t k = 0; k < n_creatures; ++k)
        if (dirty[k]) {
            const simuse_pool_slot& s = p->slots[k];
            std::copy(mirrors[k], mirrors[k] + s.n_syn, p->values.begin() + s.syn_off);
            any = true;
        }
    return any ? p->upload_weights() : SIMUSE_OK;
}

int simuse_pool_step(simuse_pool* p, const float* const* exts, uint32_t n_creatures, uint32_t spt, uint32_t* flat,
                     uint64_t flat_len) {
    if (!p) return fail(SIMUSE_E_INVALID, "pool_step: null pool");
    if (p->slots.empty()) return SIMUSE_OK;
    if (n_creatures != p->slots.size())
        return fail(SIMUSE_E_INVALID, "pool_step: %u inputs for %zu creatures", n_creatures, p->slots.size());
    if (flat && flat_len < static_cast<uint64_t>(spt) * p->n)
        return fail(SIMUSE_E_INVALID, "pool_step: flat holds %llu < spt * n_neurons = %llu",
                    static_cast<unsigned long long>(flat_len), static_cast<unsigned long long>(spt) * p->n);
    int rc = p->build();
    if (rc) return rc;
    std::fill(p->ext_full.begin(), p->ext_full.end(), 0.0f);
    for (uint32_t k = 0; k < n_creatures; ++k)
        if (exts && exts[k]) {
            const simuse_pool_slot& s = p->slots[k];
            std::copy(exts[k], exts[k] + s.n, p->ext_full.begin() + s.neuron_off);
        }
    PHIP(hipMemcpyAsync(p->ext, p->ext_full.data(), sizeof(float) * p->n, hipMemcpyHostToDevice, p->stream));
    for (uint32_t done = 0; done < spt;) {
        const uint32_t chunk = std::min(spt - done, kMaxSubmit);
        if ((rc = p->run_chunk(chunk))) return rc;
        if (flat)
            PHIP(hipMemcpyAsync(flat + static_cast<size_t>(done) * p->n, p->hist, sizeof(uint32_t) * chunk * p->n,
                                hipMemcpyDeviceToHost, p->stream));
        done += chunk;
    }
    p->steps += spt;
    PHIP(hipGetLastError());
    PHIP(hipStreamSynchronize(p->stream));
    return SIMUSE_OK;
}

int simuse_pool_download_state(simuse_pool* p, int32_t col, float* out, uint64_t n) {
    if (!p || !out) return fail(SIMUSE_E_INVALID, "pool_download_state: null argument");
    if (col < 0 || col > 2) return fail(SIMUSE_E_INVALID, "pool_download_state: column %d (0 v, 1 g, 2 refr_cnt)", col);
    int rc = p->build();
    if (rc) return rc;
    if (n != p->n) return fail(SIMUSE_E_INVALID, "pool_download_state: %llu values for %u neurons",
                               static_cast<unsigned long long>(n), p->n);
    const float* src = col == 0 ? p->v : col == 1 ? p->g : p->refr;
    PHIP(hipMemcpyAsync(out, src, sizeof(float) * n, hipMemcpyDeviceToHost, p->stream));
    PHIP(hipStreamSynchronize(p->stream));
    return SIMUSE_OK;
}

int simuse_pool_time_substeps(simuse_pool* p, uint32_t n, float* ms) {
    if (!p || !ms) return fail(SIMUSE_E_INVALID, "pool_time_substeps: null argument");
    if (p->slots.empty()) return fail(SIMUSE_E_STATE, "pool_time_substeps: no creature registered");
    int rc = p->build();
    if (rc) return rc;
    for (uint32_t done = 0; done < n;) {  // graphs of every chunk size first: no capture inside the region
        const uint32_t chunk = std::min(n - done, kMaxSubmit);
        bool have = false;
        for (const auto& gr : p->graphs) have = have || gr.chunk == chunk;
        if (!have) {
            PHIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
            p->enqueue(chunk);
            hipGraph_t gg = nullptr;
            hipError_t e = hipStreamEndCapture(p->stream, &gg);
            if (e != hipSuccess) return fail(SIMUSE_E_HIP, "pool capture: %s", hipGetErrorString(e));
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, gg, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(gg);
                return fail(SIMUSE_E_HIP, "pool instantiate: %s", hipGetErrorString(e));
            }
            p->graphs.push_back({chunk, gg, x});
        }
        done += chunk;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PHIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(SIMUSE_E_HIP, "pool_time_substeps: hipEventCreate");
    }
    hipError_t he = hipEventRecord(e0, p->stream);
    for (uint32_t done = 0; he == hipSuccess && done < n;) {
        const uint32_t chunk = std::min(n - done, kMaxSubmit);
        if ((rc = p->run_chunk(chunk))) break;
        done += chunk;
    }
    if (he == hipSuccess && !rc) he = hipEventRecord(e1, p->stream);
    if (he == hipSuccess && !rc) he = hipEventSynchronize(e1);
    if (he == hipSuccess && !rc) he = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (he != hipSuccess) return fail(SIMUSE_E_HIP, "pool_time_substeps: %s", hipGetErrorString(he));
    p->steps += n;
    return SIMUSE_OK;
}

int simuse_pool_info(simuse_pool* p, uint64_t* n_neurons, uint64_t* n_syn, uint64_t* steps, int32_t* delay) {
    if (!p) return fail(SIMUSE_E_INVALID, "pool_info: null pool");
    if (n_neurons) *n_neurons = p->n;
    if (n_syn) *n_syn = p->values.size();
    if (steps) *steps = p->built() ? p->steps : 0;
    if (delay) *delay = p->delay;
    return SIMUSE_OK;
}

}  // extern "C"
