// net_kernels.hip — per-timestep kernels of the SiMuse network step on gfx950.
//
// One step of one network is, in stream order:
//   1. neuron_kernel<MODEL>   (one launch per model present)
//        read ring slot s%32, zero it, add Gaussian noise, clamp (LIF/Izh),
//        advance the neuron, update STDP traces, wave-ballot the spike mask
//        and write the tile's spiking ids in ascending order.
//   2. ltp_kernel / ltd_kernel (only with pair-based STDP)
//   3. insert_kernel          every synaptic event (e = edge id) pushes
//        itself onto the chain of its (target slot, post) key with one
//        atomicExch; Poisson sources roll and push their edges the same way.
//   4. fold_kernel            the first event of each key walks the chain,
//        orders it by edge id — which is exactly the CPU push order: ascending
//        pre, CSR row order, then sources (P/src/api/net/push.rs:143-171,
//        P/src/api/net/network/step.rs:166-185) — and adds the weights into
//        the ring one by one.  No float atomics: the ring is bit-identical to
//        the sequential CPU push.  The last block advances the step counter.
//
// All arithmetic is compiled with -ffp-contract=off so every multiply and add
// rounds separately, as in the reference's non-FMA build.

#include "device_math.hpp"
#include "net_kernels.hpp"

namespace simuse {

__device__ __forceinline__ float par(const PopDev& P, int k, uint32_t li) {
    return ((P.uni_mask >> k) & 1u) ? P.puni[k] : P.pr[static_cast<size_t>(k) * P.dstride + li];
}

// Non-NaN max/min with the reference's argument order (f32::max / SIMD max).
__device__ __forceinline__ float fmax_ref(float a, float b) { return a > b ? a : b; }

// LIFModel::step, P/src/zoo/neurons/lif.rs:47-78.
__device__ __forceinline__ int lif_update(float& v, float& refr, float I, float dt, float inv_tau,
                                          float v_rest, float v_th, float v_reset, float tau_ref) {
    const bool is_refr = refr > 0.0f;
    const float dv = ((v_rest - v) + I) * inv_tau * dt;
    const float v_new = v + dv;
    const bool spike = !is_refr && (v_new >= v_th);
    const float v_final = spike ? v_reset : v_new;
    v = is_refr ? v : v_final;
    const float refr_set = spike ? tau_ref : (refr - dt);
    refr = fmax_ref(refr_set, 0.0f);
    return spike ? 1 : 0;
}

// IzhikevichModel::step, P/src/zoo/neurons/izhikevich.rs:71-125 (two 0.5*dt
// half-steps, Horner v*(0.04v+5)+140, du from the pre-integration v).
__device__ __forceinline__ int izh_update(float& v, float& u, float& refr, float I, float dt,
                                          float a, float b, float c, float d, float tau_ref,
                                          float v_th) {
    const bool is_refr = refr > 0.0f;
    const float vv = is_refr ? c : v;
    const float half_dt = dt * 0.5f;
    const float p1 = vv * (0.04f * vv + 5.0f);
    const float dv1 = (((p1 + 140.0f) - u) + I) * half_dt;
    const float vm = vv + dv1;
    const float p2 = vm * (0.04f * vm + 5.0f);
    const float dv2 = (((p2 + 140.0f) - u) + I) * half_dt;
    const float v_new = vv + (is_refr ? 0.0f : (dv1 + dv2));
    const float du = a * (b * vv - u) * dt;
    const float u_new = is_refr ? u : (u + du);
    const bool spike = v_new >= v_th;
    v = spike ? c : v_new;
    u = spike ? (u_new + d) : u_new;
    refr = fmax_ref(spike ? tau_ref : (refr - dt), 0.0f);
    return spike ? 1 : 0;
}

// Rush–Larsen gate update, PC/kernels/hh.cu:174-182.
__device__ __forceinline__ void hh_gate(float& g, float alpha, float beta, float dt, float q10) {
    const float sum_q10 = (alpha + beta + 1e-12f) * q10;
    const float g_inf = alpha / (alpha + beta + 1e-12f);
    const float decay = det_expf(-dt * sum_q10);
    g = g_inf + (g - g_inf) * decay;
}

// Hodgkin–Huxley with Na/K(/CaT) channels, equations of
// PC/kernels/hh.cu:107-216 (exponential-Euler membrane, rising edge through
// v_thresh gated by the refractory timer, the ODE keeps running while
// refractory).
__device__ __forceinline__ int hh_update(const PopDev& P, uint32_t li, float I, float dt) {
    float* s = P.st;
    const size_t ds = P.dstride;
    float v = s[li], m = s[ds + li], h = s[2 * ds + li], n = s[3 * ds + li], refr = s[4 * ds + li];
    const float g_na = par(P, 1, li), g_k = par(P, 2, li), g_l = par(P, 3, li);
    const float e_na = par(P, 4, li), e_k = par(P, 5, li), e_l = par(P, 6, li);
    const float inv_cm = par(P, 7, li), q10 = par(P, 8, li), v_th = par(P, 9, li);
    const float g_cat = par(P, 10, li), e_cat = par(P, 11, li), tau_ref = par(P, 12, li);
    const float ref = refr;
    if (ref > 0.0f) refr = fmax_ref(0.0f, ref - dt);
    const float vi = v;
    const float am = 0.1f * vtrap(-(vi + 40.0f), 10.0f);
    const float bm = 4.0f * det_expf(-(vi + 65.0f) / 18.0f);
    hh_gate(m, am, bm, dt, q10);
    const float ah = 0.07f * det_expf(-(vi + 65.0f) / 20.0f);
    const float bh = 1.0f / (1.0f + det_expf(-(vi + 35.0f) / 10.0f));
    hh_gate(h, ah, bh, dt, q10);
    float gna = g_na;
    gna *= m; gna *= m; gna *= m; gna *= h;
    const float an = 0.01f * vtrap(-(vi + 55.0f), 10.0f);
    const float bn = 0.125f * det_expf(-(vi + 65.0f) / 80.0f);
    hh_gate(n, an, bn, dt, q10);
    float gk = g_k;
    gk *= n; gk *= n; gk *= n; gk *= n;
    float g_ion = 0.0f + gna;
    float i_ion = 0.0f + gna * e_na;
    g_ion += gk;
    i_ion += gk * e_k;
    if (g_cat != 0.0f) {
        float cg = s[5 * ds + li];
        const float minf = 1.0f / (1.0f + det_expf(-(vi + 56.0f) / 6.2f));
        hh_gate(cg, minf / 0.6f, (1.0f - minf) / 0.6f, dt, q10);
        s[5 * ds + li] = cg;
        const float gc = g_cat * cg;
        g_ion += gc;
        i_ion += gc * e_cat;
    }
    const float g_tot = g_ion + g_l + 1e-12f;
    const float v_inf = (I + i_ion + g_l * e_l) / g_tot;
    const float decay_v = det_expf(-dt * (g_tot * inv_cm));
    const float vn = v_inf + (vi - v_inf) * decay_v;
    int spiked = 0;
    if (ref <= 0.0f && vi < v_th && vn >= v_th) {
        spiked = 1;
        refr = tau_ref;
    }
    s[li] = vn; s[ds + li] = m; s[2 * ds + li] = h; s[3 * ds + li] = n; s[4 * ds + li] = refr;
    return spiked;
}

// Exclusive prefix sum over one block of kTile threads; returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_wave, uint32_t& excl) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= static_cast<uint32_t>(o)) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWavesPerTile; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) base += t;
        total += t;
    }
    excl = base + inc - x;
    __syncthreads();
    return total;
}

template <int MODEL>
__global__ __launch_bounds__(kTile) void neuron_kernel(const NetArgs* __restrict__ nets,
                                                      const uint2* __restrict__ tiles) {
    const uint2 ref = tiles[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    const uint32_t tile = ref.y;
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;
    const PopDev& P = a.pops[a.tile_pop[tile]];
    const uint32_t slot = static_cast<uint32_t>(step & (kSlots - 1));
    const uint32_t li = (tile - P.tile0) * kTile + threadIdx.x;
    int spike = 0;
    int drive_now = 0;  // drive source of this neuron fires this step (one event)
    if (li < P.n) {
        const uint32_t gi = P.off + li;
        float* rs = a.ring + static_cast<size_t>(slot) * a.N + gi;
        float I = *rs;
        *rs = 0.0f;
        // A per-neuron Poisson source with one edge of delay <= 1 adds its
        // weight into the slot after every neuron push of the previous step
        // (P/src/api/net/network/step.rs:166-185), i.e. as the last add.
        if ((P.flags & kPopDrive) && P.drive_p > 0.0f) {
            const uint32_t sd = P.drive_seed0 + li * 2654435761u;
            if (step > 0 && poisson_roll(sd, 0u, step - 1, P.drive_p)) I = I + P.drive_w;
            drive_now = poisson_roll(sd, 0u, step, P.drive_p) ? 1 : 0;
        }
        if (P.flags & kPopNoise) {
            const float sg = (P.flags & kPopNoiseScalar) ? P.sigma_uni : P.sigma[li];
            I = I + sg * normal_draw(P.seed, li, step);
        }
        const float dt = P.puni[0];
        if (MODEL == 0) {
            I = fminf(fmaxf(I, -100.0f), 100.0f);  // population.rs:79-88
            float* s = P.st;
            float v = s[li];
            float refr = (P.flags & kPopRefrElided) ? 0.0f : s[P.dstride + li];
            spike = lif_update(v, refr, I, dt, par(P, 1, li), par(P, 2, li), par(P, 3, li),
                               par(P, 4, li), par(P, 5, li));
            s[li] = v;
            if (!(P.flags & kPopRefrElided)) s[P.dstride + li] = refr;
        } else if (MODEL == 1) {
            I = fminf(fmaxf(I, -100.0f), 100.0f);
            float* s = P.st;
            float v = s[li], u = s[P.dstride + li];
            float refr = (P.flags & kPopRefrElided) ? 0.0f : s[2 * P.dstride + li];
            spike = izh_update(v, u, refr, I, dt, par(P, 1, li), par(P, 2, li), par(P, 3, li),
                               par(P, 4, li), par(P, 5, li), par(P, 6, li));
            s[li] = v;
            s[P.dstride + li] = u;
            if (!(P.flags & kPopRefrElided)) s[2 * P.dstride + li] = refr;
        } else {
            spike = hh_update(P, li, I, dt);  // #[no_clamp] (hh.rs:13)
        }
        if (a.stdp) {
            const float add = spike ? 1.0f : 0.0f;
            a.tr_pre[gi] = fmaf(a.tr_pre[gi], a.stdp_dpre, add);
            a.tr_post[gi] = fmaf(a.tr_post[gi], a.stdp_dpost, add);
        }
    }
    // Ordered compaction: wave ballot + block prefix over the 4 waves.
    const uint64_t m = __ballot(spike);
    const uint64_t dm = __ballot(drive_now);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    __shared__ uint32_t s_cnt[kWavesPerTile], s_drv[kWavesPerTile];
    const uint32_t local_tile = tile - P.tile0;
    if (lane == 0) {
        s_cnt[wave] = static_cast<uint32_t>(__popcll(m));
        s_drv[wave] = static_cast<uint32_t>(__popcll(dm));
        const uint32_t word = P.word0 + local_tile * kWavesPerTile + wave;
        a.bits[word] = m;
        if (a.raster) {
            const uint64_t rs = step - *a.raster_base;
            if (rs < a.raster_cap) a.raster[rs * a.n_words + word] = m;
        }
    }
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWavesPerTile; ++w) {
        const uint32_t c = s_cnt[w];
        if (w < wave) base += c;
        total += c;
    }
    if (spike) {
        const uint32_t rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
        a.active[static_cast<size_t>(tile) * kTile + base + rank] = P.off + li;
    }
    if (threadIdx.x == 0) {
        a.tile_count[tile] = total;
        a.tile_spikes[tile] += total;
        // each firing drive source is one synaptic event, counted at its roll
        const uint32_t dtot = s_drv[0] + s_drv[1] + s_drv[2] + s_drv[3];
        if (dtot) a.tile_events[tile] += dtot;
    }
}

// Per-tile event enumeration shared by insert and fold.  A neuron tile holds
// up to kTile spiking neurons (ascending ids); a source tile holds up to kTile
// Poisson sources that fired this step.  `visit(eid, post, d)` is called for
// every outgoing edge; eid is the global event id (edge id, sources after
// n_syn).
struct TileRows {
    uint32_t cnt;
    uint32_t total;
    uint32_t base_eid;  // 0 for neuron edges, n_syn for source edges
};

template <int SCHED, class Visit>
__device__ __forceinline__ void for_tile_edges(const NetArgs& a, uint32_t b, uint64_t step, bool is_src,
                                               uint32_t* s_rs, uint32_t* s_pref, uint32_t* s_wave,
                                               uint32_t& out_total, Visit visit) {
    const uint32_t t = threadIdx.x;
    uint32_t cnt;
    uint32_t rs = 0, dg = 0;
    const uint32_t* rp = is_src ? a.src_row_ptr : a.row_ptr;
    if (!is_src) {
        cnt = a.tile_count[b];
        if (t < cnt) {
            const uint32_t id = a.active[static_cast<size_t>(b) * kTile + t];
            rs = rp[id];
            dg = rp[id + 1] - rs;
        }
    } else {
        const uint32_t s = (b - a.n_tiles) * kTile + t;
        int fired = 0;
        if (s < a.n_src) {
            const float p = a.src_p[s];
            fired = (p > 0.0f && poisson_roll(a.src_seed[s], 0u, step, p)) ? 1 : 0;
        }
        const uint64_t m = __ballot(fired);
        __shared__ uint32_t s_c[kWavesPerTile];
        const uint32_t lane = t & 63u, wave = t >> 6;
        if (lane == 0) s_c[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t base = 0;
        cnt = 0;
        for (uint32_t w = 0; w < kWavesPerTile; ++w) {
            if (w < wave) base += s_c[w];
            cnt += s_c[w];
        }
        if (fired) {
            const uint32_t rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
            s_rs[base + rank] = s;  // staged source ids
        }
        __syncthreads();
        if (t < cnt) {
            const uint32_t sid = s_rs[t];
            rs = rp[sid];
            dg = rp[sid + 1] - rs;
        }
        __syncthreads();
    }
    if (cnt == 0) { out_total = 0; return; }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(dg, s_wave, excl);
    if (t < cnt) { s_rs[t] = rs; s_pref[t] = excl; }
    __syncthreads();
    out_total = total;
    const uint32_t* col = is_src ? a.src_post : a.col;
    const uint8_t* dly = is_src ? a.src_delay : a.delay;
    const uint32_t base_eid = is_src ? static_cast<uint32_t>(a.n_syn) : 0u;
    if (SCHED == 2) {
        // one wavefront per spiking neuron, lanes stride over its row
        const uint32_t lane = t & 63u, wave = t >> 6;
        for (uint32_t k = wave; k < cnt; k += kWavesPerTile) {
            const uint32_t r0 = s_rs[k];
            const uint32_t len = (k + 1 < cnt ? s_pref[k + 1] : total) - s_pref[k];
            for (uint32_t j = lane; j < len; j += 64) {
                const uint32_t e = r0 + j;
                visit(base_eid + e, col[e], dly[e]);
            }
        }
    } else {
        // block-flattened: thread f takes edge f of the tile's concatenated rows
        for (uint32_t f = t; f < total; f += kTile) {
            uint32_t lo = 0, hi = cnt;  // largest k with s_pref[k] <= f
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pref[mid] <= f) lo = mid; else hi = mid;
            }
            const uint32_t e = s_rs[lo] + (f - s_pref[lo]);
            visit(base_eid + e, col[e], dly[e]);
        }
    }
}

template <int SCHED>
__global__ __launch_bounds__(kTile) void insert_kernel(const NetArgs* __restrict__ nets,
                                                      const uint2* __restrict__ blocks) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile];
    const uint2 ref = blocks[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    const uint32_t b = ref.y;
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;
    const bool is_src = b >= a.n_tiles;
    uint32_t total;
    uint32_t* head = a.head;
    uint32_t* next = a.next;
    const uint32_t N = a.N;
    for_tile_edges<SCHED>(a, b, step, is_src, s_rs, s_pref, s_wave, total,
                          [&](uint32_t eid, uint32_t post, uint8_t d) {
                              const uint32_t dd = d < 1 ? 1u : static_cast<uint32_t>(d);
                              const uint32_t ts = static_cast<uint32_t>((step + dd) & (kSlots - 1));
                              const size_t key = static_cast<size_t>(ts) * N + post;
                              next[eid] = atomicExch(&head[key], eid);
                          });
    if (threadIdx.x == 0 && total) a.tile_events[b] += total;
}

__device__ __forceinline__ float event_weight(const NetArgs& a, uint32_t eid) {
    return eid < a.n_syn ? a.w[eid] : a.src_w[eid - a.n_syn];
}

template <int SCHED>
__global__ __launch_bounds__(kTile) void fold_kernel(const NetArgs* __restrict__ nets,
                                                    const uint2* __restrict__ blocks) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile];
    constexpr int kChain = 8;
    __shared__ uint32_t s_chain[kTile * kChain];
    const uint2 ref = blocks[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    const uint32_t b = ref.y;
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;  // every block of the network sees the same value
    const bool is_src = b >= a.n_tiles;
    uint32_t total;
    uint32_t* head = a.head;
    const uint32_t* next = a.next;
    float* ring = a.ring;
    const uint32_t N = a.N;
    uint32_t* my = s_chain + threadIdx.x * kChain;
    for_tile_edges<SCHED>(a, b, step, is_src, s_rs, s_pref, s_wave, total,
                          [&](uint32_t eid, uint32_t post, uint8_t d) {
        if (next[eid] != kEmpty) return;  // not the first event of its key
        const uint32_t dd = d < 1 ? 1u : static_cast<uint32_t>(d);
        const uint32_t ts = static_cast<uint32_t>((step + dd) & (kSlots - 1));
        const size_t key = static_cast<size_t>(ts) * N + post;
        const uint32_t h = head[key];
        head[key] = kEmpty;
        float r = ring[key];
        if (h == eid) {  // single contribution
            ring[key] = r + event_weight(a, eid);
            return;
        }
        // Collect the chain (insertion order is arbitrary), order by edge id.
        int len = 0;
        bool overflow = false;
        for (uint32_t c = h; c != kEmpty; c = next[c]) {
            if (len == kChain) { overflow = true; break; }
            uint32_t x = c;
            int j = len++;
            while (j > 0 && my[j - 1] > x) { my[j] = my[j - 1]; --j; }
            my[j] = x;
        }
        if (!overflow) {
            for (int k = 0; k < len; ++k) r = r + event_weight(a, my[k]);
        } else {
            // Long chain: repeatedly take the smallest edge id above the last.
            uint32_t last = 0;
            bool first = true;
            for (;;) {
                uint32_t best = kEmpty;
                for (uint32_t c = h; c != kEmpty; c = next[c])
                    if ((first || c > last) && c < best) best = c;
                if (best == kEmpty) break;
                r = r + event_weight(a, best);
                last = best;
                first = false;
            }
        }
        ring[key] = r;
    });
    // The last of the network's blocks to finish advances its step counter.
    // Every block consumed its load of *a.step before reaching this point, so
    // a relaxed counter suffices; the ring and chain writes reach the next
    // kernel through the kernel boundary.  (No device-scope fence here: on
    // gfx950 a release fence writes back the XCD's L2, once per block.)
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = atomicAdd(a.done, 1u);
        if (prev == a.n_blocks - 1) {
            *a.done = 0u;
            *a.step = step + 1;
        }
    }
}

// Pair-based STDP, LTP pass: in-edges of every spiking post gain
// A+ * pre-trace.  Each edge has one post, so each edge is written by at most
// one thread.  FP64 master weight, FP32 mirror for the scatter.
__global__ __launch_bounds__(kTile) void ltp_kernel(const NetArgs* __restrict__ nets,
                                                  const uint2* __restrict__ tiles) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile];
    const uint2 ref = tiles[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    if (*a.step >= a.step_limit) return;
    const uint32_t b = ref.y, t = threadIdx.x;
    const uint32_t cnt = a.tile_count[b];
    if (cnt == 0) return;
    uint32_t rs = 0, dg = 0;
    if (t < cnt) {
        const uint32_t id = a.active[static_cast<size_t>(b) * kTile + t];
        rs = a.in_ptr[id];
        dg = a.in_ptr[id + 1] - rs;
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(dg, s_wave, excl);
    if (t < cnt) { s_rs[t] = rs; s_pref[t] = excl; }
    __syncthreads();
    const float ap = a.a_plus;
    for (uint32_t f = t; f < total; f += kTile) {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[mid] <= f) lo = mid; else hi = mid;
        }
        const uint32_t e = a.in_edge[s_rs[lo] + (f - s_pref[lo])];
        if (a.plastic && !a.plastic[e]) continue;
        const float dw = ap * a.tr_pre[a.edge_pre[e]];
        double nw = a.w64[e] + static_cast<double>(dw);
        nw = fmin(fmax(nw, static_cast<double>(a.w_min)), static_cast<double>(a.w_max));
        a.w64[e] = nw;
        const_cast<float*>(a.w)[e] = static_cast<float>(nw);
    }
}

// LTD pass: out-edges of every spiking pre lose A- * post-trace.
__global__ __launch_bounds__(kTile) void ltd_kernel(const NetArgs* __restrict__ nets,
                                                  const uint2* __restrict__ tiles) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile];
    const uint2 ref = tiles[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    if (*a.step >= a.step_limit) return;
    const uint32_t b = ref.y, t = threadIdx.x;
    const uint32_t cnt = a.tile_count[b];
    if (cnt == 0) return;
    uint32_t rs = 0, dg = 0;
    if (t < cnt) {
        const uint32_t id = a.active[static_cast<size_t>(b) * kTile + t];
        rs = a.row_ptr[id];
        dg = a.row_ptr[id + 1] - rs;
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(dg, s_wave, excl);
    if (t < cnt) { s_rs[t] = rs; s_pref[t] = excl; }
    __syncthreads();
    const float am = a.a_minus;
    for (uint32_t f = t; f < total; f += kTile) {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[mid] <= f) lo = mid; else hi = mid;
        }
        const uint32_t e = s_rs[lo] + (f - s_pref[lo]);
        if (a.plastic && !a.plastic[e]) continue;
        const float dw = am * a.tr_post[a.col[e]];
        double nw = a.w64[e] - static_cast<double>(dw);
        nw = fmin(fmax(nw, static_cast<double>(a.w_min)), static_cast<double>(a.w_max));
        a.w64[e] = nw;
        const_cast<float*>(a.w)[e] = static_cast<float>(nw);
    }
}

// inject_current into the slot the next step reads (build.rs:243-275).
__global__ void inject_kernel(float* ring, const uint64_t* step, uint32_t N, const uint32_t* idx,
                              const float* val, uint32_t count, int serial) {
    const size_t slot = static_cast<size_t>(*step & (kSlots - 1));
    if (serial) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (uint32_t i = 0; i < count; ++i) ring[slot * N + idx[i]] += val[i];
        return;
    }
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) ring[slot * N + idx[i]] += val[i];
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void widen_weights_kernel(const float* w, double* w64, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        w64[i] = static_cast<double>(w[i]);
}

// Firing rate and ISI CV per neuron from the bit raster
// (P/src/kernels/stats/rates.rs:6-35, P/src/kernels/stats/isi.rs:30-67):
// rate = count / (steps*dt); CV = sqrt(var_{n-1}) / mean over >= 2 ISIs.
// Two passes over the raster (mean, then centred variance), float sums in
// step order as in the reference.
__global__ void raster_stats_kernel(const uint64_t* raster, uint32_t n_words, uint32_t steps,
                                    const uint32_t* word_of, const uint32_t* bit_of, uint32_t count,
                                    float dt_s, float* rates, float* cvs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t wd = word_of[i];
    const uint64_t bm = 1ull << bit_of[i];
    float c = 0.0f, sum = 0.0f;
    int64_t last = -1;
    uint32_t nisi = 0;
    for (uint32_t t = 0; t < steps; ++t) {
        if (raster[static_cast<size_t>(t) * n_words + wd] & bm) {
            c += 1.0f;
            if (last >= 0) {
                const float iv = static_cast<float>(t - last) * dt_s * 1000.0f;
                if (iv > 0.0f) { sum += iv; ++nisi; }
            }
            last = t;
        }
    }
    const float dur = static_cast<float>(steps) * dt_s;
    rates[i] = dur > 0.0f ? c / dur : c;
    float cv = 0.0f;
    if (nisi >= 2) {
        const float mean = sum / static_cast<float>(nisi);
        if (mean > 0.0f) {
            float var = 0.0f;
            last = -1;
            for (uint32_t t = 0; t < steps; ++t) {
                if (raster[static_cast<size_t>(t) * n_words + wd] & bm) {
                    if (last >= 0) {
                        const float iv = static_cast<float>(t - last) * dt_s * 1000.0f;
                        if (iv > 0.0f) var += (iv - mean) * (iv - mean);
                    }
                    last = t;
                }
            }
            var /= static_cast<float>(nisi - 1);
            cv = __builtin_sqrtf(var) / mean;
        }
    }
    cvs[i] = cv;
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from net_api.hip)
// ---------------------------------------------------------------------------
void launch_neuron(int model, const NetArgs* nets, const uint2* tiles, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (model == 0) neuron_kernel<0><<<n, kTile, 0, s>>>(nets, tiles);
    else if (model == 1) neuron_kernel<1><<<n, kTile, 0, s>>>(nets, tiles);
    else neuron_kernel<2><<<n, kTile, 0, s>>>(nets, tiles);
}
void launch_stdp(const NetArgs* nets, const uint2* tiles, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    ltp_kernel<<<n, kTile, 0, s>>>(nets, tiles);
    ltd_kernel<<<n, kTile, 0, s>>>(nets, tiles);
}
void launch_insert(int sched, const NetArgs* nets, const uint2* blocks, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (sched == 2) insert_kernel<2><<<n, kTile, 0, s>>>(nets, blocks);
    else insert_kernel<1><<<n, kTile, 0, s>>>(nets, blocks);
}
void launch_fold(int sched, const NetArgs* nets, const uint2* blocks, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (sched == 2) fold_kernel<2><<<n, kTile, 0, s>>>(nets, blocks);
    else fold_kernel<1><<<n, kTile, 0, s>>>(nets, blocks);
}
void launch_inject(float* ring, const uint64_t* step, uint32_t N, const uint32_t* idx, const float* val,
                   uint32_t count, int serial, hipStream_t s) {
    if (count == 0) return;
    if (serial) inject_kernel<<<1, 1, 0, s>>>(ring, step, N, idx, val, count, 1);
    else inject_kernel<<<(count + 255) / 256, 256, 0, s>>>(ring, step, N, idx, val, count, 0);
}
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_u32_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(p, v, n);
}
void launch_widen(const float* w, double* w64, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    widen_weights_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(w, w64, n);
}
void launch_raster_stats(const uint64_t* raster, uint32_t n_words, uint32_t steps, const uint32_t* word_of,
                         const uint32_t* bit_of, uint32_t count, float dt_s, float* rates, float* cvs,
                         hipStream_t s) {
    if (count == 0) return;
    raster_stats_kernel<<<(count + 255) / 256, 256, 0, s>>>(raster, n_words, steps, word_of, bit_of, count,
                                                            dt_s, rates, cvs);
}

}  // namespace simuse
