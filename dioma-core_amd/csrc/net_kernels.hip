// net_kernels.hip — per-timestep kernels of the SiMuse network step on gfx950.
//
// One step of one network is, in stream order:
//   1. neuron_kernel<MODEL,SCHED>  (one launch per model present)
//        read ring slot s%32, zero it, add drive and Gaussian noise, clamp
//        (LIF/Izh), advance the neuron, update STDP traces, wave-ballot the
//        spike mask and write the tile's spiking ids in ascending order; for
//        networks without plasticity, scatter the tile's spikes right away.
//   2. stdp_kernel                 (only with pair-based STDP)
//   3. insert_kernel<SCHED>        scatter of plastic networks' spikes and of
//        the Poisson sources that fire this step
//   4. fixup_kernel                re-sums, in ascending event id, every ring
//        cell that received more than one event; the last block advances the
//        step counter.
// The ring ends each step bit-identical to the CPU's sequential push
// (P/src/api/net/push.rs:143-171): no float atomics.
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
// refractory).  State v, m, h, n, refr (and the CaT gate when g_cat != 0);
// p[k] = parameter column k.
__device__ __forceinline__ int hh_update(float& v, float& m, float& h, float& n, float& refr, float& cg,
                                         const float* p, float I, float dt) {
    const float g_na = p[1], g_k = p[2], g_l = p[3];
    const float e_na = p[4], e_k = p[5], e_l = p[6];
    const float inv_cm = p[7], q10 = p[8], v_th = p[9];
    const float g_cat = p[10], e_cat = p[11], tau_ref = p[12];
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
        const float minf = 1.0f / (1.0f + det_expf(-(vi + 56.0f) / 6.2f));
        hh_gate(cg, minf / 0.6f, (1.0f - minf) / 0.6f, dt, q10);
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
    v = vn;
    return spiked;
}

// Inclusive prefix sum over the 64 lanes of a wave, every lane active: four
// DPP row shifts scan each row of 16 lanes, then row_bcast:15 and
// row_bcast:31 carry the row totals forward (GFX9 data-parallel primitives,
// full-rate VALU; a __shfl_up ladder goes through the LDS crossbar instead).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    int v = static_cast<int>(x);
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return static_cast<uint32_t>(v);
}

// Exclusive prefix sum over one block of kTile threads; returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_wave, uint32_t& excl) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = x;
    inc = wave_incl_scan(inc);
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

// ---------------------------------------------------------------------------
// Scatter.  The CPU push adds each synaptic event's weight into ring cell
// key = ((step + max(d,1)) % 32) * N + post, pre ascending and in CSR row
// order, then the Poisson sources (P/src/api/net/push.rs:143-171,
// P/src/api/net/network/step.rs:166-185): per key the adds happen in
// ascending event id.  On the device every event exchanges itself into the
// key's tagged head.  The event that finds the head idle (another step's tag)
// is the key's first: it adds its weight into the ring at once and keeps the
// value it read.  An event that finds the head taken links to the event it
// displaced and enters its tile's collision list.  fixup_kernel re-sums each
// key that received more than one event, from the kept value in ascending
// event id, so the ring ends bit-identical to the sequential push.  An event
// alone on its key (the common case) costs one exchange and one ring
// read-modify-write; no float atomics anywhere.
// ---------------------------------------------------------------------------

// x^n, n >= 1, by right-to-left binary exponentiation (the expansion of
// f32::powi used by the reference's recovery step).
__device__ __forceinline__ float powi_f32(float a, int n) {
    float r = 1.0f;
    for (;;) {
        if (n & 1) r *= a;
        n /= 2;
        if (n == 0) break;
        a *= a;
    }
    return r;
}

// tm_spike (P/src/api/net/push.rs:60-86) for event e: lazy exponential
// recovery over the steps since the edge last fired (capped at 1e6), deliver
// weight * R*U with the pre-spike factor, then deplete R (facilitate U).  An
// edge fires at most once per step, so its state has a single writer.
__device__ __noinline__ float tm_deliver(const NetArgs& a, uint32_t e, float wt, uint64_t step) {
    const uint32_t t = a.tm_type[e];
    if (t == 0) return wt;
    const float4 p = a.tm_prof[t - 1];
    float r = a.tm_r[e], u = a.tm_u[e];
    const uint64_t last = a.tm_last[e];
    uint64_t gap = step > last ? step - last : 0u;
    if (gap > 1000000u) gap = 1000000u;
    const int n = static_cast<int>(gap);
    if (n > 0) {
        r = 1.0f + (r - 1.0f) * powi_f32(p.y, n);
        if (p.w != 0.0f) u = p.x + (u - p.x) * powi_f32(p.z, n);
    }
    const float scale = r * u;
    r *= 1.0f - u;
    if (p.w != 0.0f) u += p.x * (1.0f - u);
    a.tm_r[e] = r;
    a.tm_u[e] = u;
    a.tm_last[e] = static_cast<uint32_t>(step);
    return wt * scale;
}

struct ScatterCtx {
    uint64_t step;
    uint32_t tag;       // low 32 bits of step
    uint32_t ebase;     // collision list base of the tile
    uint32_t* s_ncoll;  // collisions of the block so far (LDS)
};

__device__ __forceinline__ void scatter_event(const NetArgs& a, const ScatterCtx& c, uint32_t eid, uint32_t post,
                                              uint32_t d, float wt) {
    const uint32_t dd = d < 1 ? 1u : d;
    const size_t key = static_cast<size_t>((c.step + dd) & (kSlots - 1)) * a.N + post;
    // The ring read does not depend on the exchange, so both round trips
    // overlap; only the key's first event uses the value.  When every delay
    // is <= 1 step, the target slot is untouched since it was consumed and
    // zeroed, so the read is skipped.
    const float r0 = a.ring_fresh ? 0.0f : a.ring[key];
    if (a.ev_w) a.ev_w[eid] = wt;
    const unsigned long long mine = (static_cast<unsigned long long>(c.tag) << 32) | eid;
    const unsigned long long old = atomicExch(a.head + key, mine);
    if (static_cast<uint32_t>(old >> 32) != c.tag) {
        a.ring[key] = r0 + wt;
        a.rec[eid] = make_uint2(kEmpty, __float_as_uint(r0));
    } else {
        a.rec[eid] = make_uint2(static_cast<uint32_t>(old), 0u);
        const uint32_t pos = atomicAdd(c.s_ncoll, 1u);
        a.coll[c.ebase + pos] = make_uint2(eid, static_cast<uint32_t>(key));
    }
}

// Visit every edge of `cnt` rows staged in LDS: row k starts at edge s_rs[k]
// and its edges are numbered from s_pref[k] in the concatenation (`total`
// edges).  SCHED 1: thread f takes edge f of the concatenation (binary search
// over the prefix); SCHED 2: one wavefront per row, lanes stride the row.
template <int SCHED, class Visit>
__device__ __forceinline__ void for_rows(uint32_t cnt, uint32_t total, const uint32_t* s_rs, const uint32_t* s_pref,
                                         Visit visit) {
    const uint32_t t = threadIdx.x;
    if (SCHED == 2) {
        const uint32_t lane = t & 63u, wave = t >> 6;
        for (uint32_t k = wave; k < cnt; k += kWavesPerTile) {
            const uint32_t r0 = s_rs[k];
            const uint32_t len = (k + 1 < cnt ? s_pref[k + 1] : total) - s_pref[k];
            for (uint32_t j = lane; j < len; j += 64) visit(r0 + j);
        }
    } else {
        for (uint32_t f = t; f < total; f += kTile) {
            uint32_t lo = 0, hi = cnt;  // largest k with s_pref[k] <= f
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pref[mid] <= f) lo = mid; else hi = mid;
            }
            visit(s_rs[lo] + (f - s_pref[lo]));
        }
    }
}

template <int SCHED>
__device__ __forceinline__ void scatter_rows(const NetArgs& a, const ScatterCtx& c, bool is_src, uint32_t cnt,
                                             uint32_t total, const uint32_t* s_rs, const uint32_t* s_pref) {
    if (!is_src) {
        const uint32_t* col = a.col;
        const uint8_t* dly = a.delay;
        const float* w = a.w;
        for_rows<SCHED>(cnt, total, s_rs, s_pref, [&](uint32_t e) {
            const float wt = a.tm_type ? tm_deliver(a, e, w[e], c.step) : w[e];
            scatter_event(a, c, e, col[e], dly[e], wt);
        });
    } else {
        const uint32_t base = static_cast<uint32_t>(a.n_syn);
        const uint32_t* col = a.src_post;
        const uint8_t* dly = a.src_delay;
        const float* w = a.src_w;
        for_rows<SCHED>(cnt, total, s_rs, s_pref, [&](uint32_t e) {
            const float wt = a.tm_type ? tm_deliver(a, base + e, w[e], c.step) : w[e];
            scatter_event(a, c, base + e, col[e], dly[e], wt);
        });
    }
}

// GpeaModel::step, P/src/zoo/neurons/gpea.rs:266-360 (pGPeA_k.mod): gates by
// exponential Euler at the old V (fused decay*(x - x_inf) + x_inf), r from the
// old cai, Ca by cnexp with I_Ca from the old state, Nernst potentials through
// ln, implicit-Euler membrane, rising-edge spike through v_thresh.  Same
// operation sequence as the oracle; #[no_clamp].
// params: [0]dt [1]celsius [2]inv_cm [3]v_thresh [4]area [5]gnabar [6]gkdrbar
//         [7]gl [8]el [9]gcatbar [10]gkcabar [11]kca [12]vol [13]caGain
// state:  [0]v [1]m [2]h [3]n [4]p [5]q [6]cai [7]r [8]was_below
__device__ __forceinline__ float gpea_sigmoid(float x, float th, float k) {
    return 1.0f / (1.0f + fast_expf((x - th) * (1.0f / k)));
}
__device__ __forceinline__ float gpea_tau2(float v, float t0, float t1, float tht1, float sig1, float tht2,
                                           float sig2) {
    const float e1 = fast_expf((v - tht1) * (-1.0f / sig1));
    const float e2 = fast_expf(-(v - tht2) * (1.0f / sig2));
    return t0 + t1 / (e1 + e2);
}
__device__ __forceinline__ float gpea_exp_euler(float x, float x_inf, float tau, float dt) {
    const float decay = fast_expf(-dt / tau);
    return fmaf(decay, x - x_inf, x_inf);
}
__device__ __forceinline__ int gpea_update(const PopDev& P, uint32_t li, float I, float dt) {
    float* s = P.st;
    const size_t ds = P.dstride;
    const float celsius = par(P, 1, li), inv_cm = par(P, 2, li), v_th = par(P, 3, li), area = par(P, 4, li);
    const float gnabar = par(P, 5, li), gkdrbar = par(P, 6, li), gl = par(P, 7, li), el = par(P, 8, li);
    const float gcatbar = par(P, 9, li), gkcabar = par(P, 10, li), kca = par(P, 11, li), vol = par(P, 12, li);
    const float ca_gain = par(P, 13, li);
    const float F = 96485.0f;
    const float rt = 8.31441f * ((273.0f + celsius) - 9.5f) / F * 1000.0f;
    const float ena = -rt * det_logf(10.0f / 140.0f);
    const float ek = rt * det_logf(5.0f / 140.0f);
    const float v = s[li];
    const float m_new = gpea_exp_euler(s[ds + li], gpea_sigmoid(v, -38.0f, -7.0f),
                                       0.001f + 0.1f / (1.0f + fast_expf((v - -53.0f) * (-1.0f / -0.7f))), dt);
    const float h_new = gpea_exp_euler(s[2 * ds + li], gpea_sigmoid(v, -45.5f, 6.4f),
                                       gpea_tau2(v, 0.0f, 4.5f, -50.0f, -15.0f, -50.0f, 16.0f), dt);
    const float n_new = gpea_exp_euler(s[3 * ds + li], gpea_sigmoid(v, -42.0f, -14.0f),
                                       gpea_tau2(v, 0.0f, 2.4f, -40.0f, -40.0f, -40.0f, 50.0f), dt);
    const float p_old = s[4 * ds + li], q_old = s[5 * ds + li];
    const float p_new = gpea_exp_euler(p_old, gpea_sigmoid(v, -56.0f, -6.7f),
                                       gpea_tau2(v, 5.0f, 0.33f, -27.0f, -10.0f, -102.0f, 15.0f), dt);
    const float q_new = gpea_exp_euler(q_old, gpea_sigmoid(v, -85.0f, 5.8f),
                                       gpea_tau2(v, 0.0f, 400.0f, -50.0f, -15.0f, -50.0f, 16.0f), dt);
    const float cai = s[6 * ds + li];
    const float r_new = gpea_exp_euler(s[7 * ds + li], gpea_sigmoid(cai, 0.17e-3f, -0.08e-3f), 2.0f, dt);
    const float eca_old = -rt * det_logf(cai / 2.0f) / 2.0f;
    const float gcat_old = gcatbar * p_old * p_old * q_old;
    const float ica_old = gcat_old * (v - eca_old);
    const float j = -ica_old * area * 1e-14f / (2.0f * F * vol);
    const float k_ca = ca_gain * kca;
    const float cai_decay = fast_expf(-dt * k_ca);
    const float cai_new = cai + (1.0f - cai_decay) * (ca_gain * j / k_ca - cai);
    const float eca = -rt * det_logf(cai_new / 2.0f) / 2.0f;
    const float g_na = gnabar * (m_new * m_new * m_new) * h_new;
    const float n2 = n_new * n_new;
    const float g_k = gkdrbar * (n2 * n2);
    const float g_cat = gcatbar * p_new * p_new * q_new;
    const float g_ahp = gkcabar * r_new * r_new;
    const float g_sum = g_na + g_k + g_cat + g_ahp + gl;
    const float g_e_sum = g_na * ena + (g_k + g_ahp) * ek + g_cat * eca + gl * el;
    const float c_over_dt = 1.0f / (inv_cm * dt);
    const float v_new = (c_over_dt * v + I + g_e_sum) / (c_over_dt + g_sum);
    const bool above = v_new >= v_th;  // edge_spike, P/src/ops/spike.rs:70-74
    const int spike = (above && s[8 * ds + li] > 0.0f) ? 1 : 0;
    s[li] = v_new;
    s[ds + li] = m_new;
    s[2 * ds + li] = h_new;
    s[3 * ds + li] = n_new;
    s[4 * ds + li] = p_new;
    s[5 * ds + li] = q_new;
    s[6 * ds + li] = cai_new;
    s[7 * ds + li] = r_new;
    s[8 * ds + li] = above ? 0.0f : 1.0f;
    return spike;
}

// Per-model register state: load everything a neuron-step reads, advance,
// store.  Splitting the three lets a block issue the loads of all its tiles
// before the first one is needed.
template <int MODEL>
struct Neuron;

template <>
struct Neuron<0> {  // LIF: states v, refr; params dt, inv_tau, v_rest, v_th, v_reset, tau_ref
    float v, refr, p1, p2, p3, p4, p5;
    __device__ __forceinline__ void load_state(const PopDev& P, uint32_t li) { v = P.st[li]; }
    __device__ __forceinline__ void load_params(const PopDev& P, uint32_t li) {
        refr = (P.flags & kPopRefrElided) ? 0.0f : P.st[P.dstride + li];
        p1 = par(P, 1, li); p2 = par(P, 2, li); p3 = par(P, 3, li); p4 = par(P, 4, li); p5 = par(P, 5, li);
    }
    __device__ __forceinline__ void load(const PopDev& P, uint32_t li) {
        load_state(P, li);
        load_params(P, li);
    }
    __device__ __forceinline__ float vm(const PopDev&, uint32_t) const { return v; }
    __device__ __forceinline__ int step(const PopDev&, uint32_t, float I, float dt) {
        I = fminf(fmaxf(I, -100.0f), 100.0f);  // population.rs:79-88
        return lif_update(v, refr, I, dt, p1, p2, p3, p4, p5);
    }
    __device__ __forceinline__ void store(const PopDev& P, uint32_t li) const {
        P.st[li] = v;
        if (!(P.flags & kPopRefrElided)) P.st[P.dstride + li] = refr;
    }
};

template <>
struct Neuron<1> {  // Izhikevich: states v, u, refr; params dt, a, b, c, d, tau_ref, v_th
    float v, u, refr, p1, p2, p3, p4, p5, p6;
    __device__ __forceinline__ void load_state(const PopDev& P, uint32_t li) {
        v = P.st[li];
        u = P.st[P.dstride + li];
    }
    __device__ __forceinline__ void load_params(const PopDev& P, uint32_t li) {
        refr = (P.flags & kPopRefrElided) ? 0.0f : P.st[2 * P.dstride + li];
        p1 = par(P, 1, li); p2 = par(P, 2, li); p3 = par(P, 3, li);
        p4 = par(P, 4, li); p5 = par(P, 5, li); p6 = par(P, 6, li);
    }
    __device__ __forceinline__ void load(const PopDev& P, uint32_t li) {
        load_state(P, li);
        load_params(P, li);
    }
    __device__ __forceinline__ float vm(const PopDev&, uint32_t) const { return v; }
    __device__ __forceinline__ int step(const PopDev&, uint32_t, float I, float dt) {
        I = fminf(fmaxf(I, -100.0f), 100.0f);
        return izh_update(v, u, refr, I, dt, p1, p2, p3, p4, p5, p6);
    }
    __device__ __forceinline__ void store(const PopDev& P, uint32_t li) const {
        P.st[li] = v;
        P.st[P.dstride + li] = u;
        if (!(P.flags & kPopRefrElided)) P.st[2 * P.dstride + li] = refr;
    }
};

template <>
struct Neuron<4> {  // alpha-LIF: states v, g, refr_cnt; params dt, tf, v_rest, v_th, v_reset, refr_steps, syn_decay
    float v, g, refr, p1, p2, p3, p4, p5, p6;
    __device__ __forceinline__ void load_state(const PopDev& P, uint32_t li) {
        v = P.st[li];
        g = P.st[P.dstride + li];
        refr = P.st[2 * P.dstride + li];
    }
    __device__ __forceinline__ void load_params(const PopDev& P, uint32_t li) {
        p1 = par(P, 1, li); p2 = par(P, 2, li); p3 = par(P, 3, li);
        p4 = par(P, 4, li); p5 = par(P, 5, li); p6 = par(P, 6, li);
    }
    __device__ __forceinline__ void load(const PopDev& P, uint32_t li) {
        load_state(P, li);
        load_params(P, li);
    }
    __device__ __forceinline__ float vm(const PopDev&, uint32_t) const { return v; }
    __device__ __forceinline__ int step(const PopDev&, uint32_t, float I, float) {
        return alpha_lif_update(v, g, refr, I, p1, p2, p3, p4, p5, p6);  // #[no_clamp]
    }
    __device__ __forceinline__ void store(const PopDev& P, uint32_t li) const {
        P.st[li] = v;
        P.st[P.dstride + li] = g;
        P.st[2 * P.dstride + li] = refr;
    }
};

template <>
struct Neuron<3> {  // GPEA: reads and writes its columns inside the update
    __device__ __forceinline__ void load_state(const PopDev&, uint32_t) {}
    __device__ __forceinline__ void load_params(const PopDev&, uint32_t) {}
    __device__ __forceinline__ void load(const PopDev&, uint32_t) {}
    __device__ __forceinline__ float vm(const PopDev& P, uint32_t li) const { return P.st[li]; }
    __device__ __forceinline__ int step(const PopDev& P, uint32_t li, float I, float dt) {
        return gpea_update(P, li, I, dt);
    }
    __device__ __forceinline__ void store(const PopDev&, uint32_t) const {}
};

template <>
struct Neuron<2> {  // Hodgkin-Huxley: states v, m, h, n, refr (, CaT gate); params 1..12
    float v, m, h, n, refr, cg;
    float p[13];
    __device__ __forceinline__ void load_state(const PopDev& P, uint32_t li) {
        const size_t ds = P.dstride;
        v = P.st[li];
        m = P.st[ds + li];
        h = P.st[2 * ds + li];
        n = P.st[3 * ds + li];
        refr = P.st[4 * ds + li];
    }
    __device__ __forceinline__ void load_params(const PopDev& P, uint32_t li) {
        p[0] = 0.0f;
#pragma unroll
        for (int k = 1; k < 13; ++k) p[k] = par(P, k, li);
        cg = p[10] != 0.0f ? P.st[5 * static_cast<size_t>(P.dstride) + li] : 0.0f;
    }
    __device__ __forceinline__ void load(const PopDev& P, uint32_t li) {
        load_state(P, li);
        load_params(P, li);
    }
    __device__ __forceinline__ float vm(const PopDev&, uint32_t) const { return v; }
    __device__ __forceinline__ int step(const PopDev&, uint32_t, float I, float dt) {
        return hh_update(v, m, h, n, refr, cg, p, I, dt);  // #[no_clamp] (hh.rs:13)
    }
    __device__ __forceinline__ void store(const PopDev& P, uint32_t li) const {
        const size_t ds = P.dstride;
        P.st[li] = v;
        P.st[ds + li] = m;
        P.st[2 * ds + li] = h;
        P.st[3 * ds + li] = n;
        P.st[4 * ds + li] = refr;
        if (p[10] != 0.0f) P.st[5 * ds + li] = cg;
    }
};

// ---------------------------------------------------------------------------
// Binned scatter for networks whose delays are all <= 1 step: every event of
// step s lands in slot (s+1) % 32.  A bin is kBinTiles consecutive tiles of
// the network (tiles never straddle populations); an event's cell is its
// post's tile within the bin * kTile + the post's index within the tile.  A
// network's scatter tiles are split among at most kMaxWriters writers, in
// ascending tile (so event id) order; writer j counts its events per bin in
// LDS, writes them grouped by bin from its first edge id (regions never
// overlap) as {event id, weight, cell}, and writes column j of bin_hdr:
// {first record, count} of each bin's group.  A consumer of bin k reads bin
// k's group of every writer in writer order (ascending event id), orders each
// cell's records by event id in LDS and adds them from zero: the sequential
// push.  Two layouts share this:
//   megastep   one kernel per step: block (network, bin k) gathers the
//              previous step's events of bin k's cells, advances those
//              neurons with them and writes its own spikes' events; records
//              and headers alternate between two buffers by step parity, and
//              the network's last block to finish advances the step counter;
//   two-kernel the writers are the neuron kernel's blocks (or bin_kernel after
//              the STDP kernels), then bin_apply writes each bin's slice of
//              ring slot (s+1) % 32 whole and advances the step.
// No global atomics beyond the megastep's block count, no random ring
// traffic.
// ---------------------------------------------------------------------------

// Exclusive prefix sum over a block of kThreads threads; returns the total.
// The wave totals (at most 16) are scanned once per wave in lanes 0..kW-1,
// one DPP row, and the wave's base and the block total come back through
// two readlanes: a handful of VALU instructions per wave instead of a
// compare, select and two adds per wave total in every lane.
template <uint32_t kThreads>
__device__ __forceinline__ uint32_t block_scan(uint32_t x, uint32_t* s_wave, uint32_t& excl) {
    constexpr uint32_t kW = kThreads / 64;
    static_assert(kW >= 1 && kW <= 16, "the wave totals fit one DPP row");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    const int v0 = lane < kW ? static_cast<int>(s_wave[lane]) : 0;
    int v = v0;
    if (kW > 1) v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    if (kW > 2) v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    if (kW > 4) v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    if (kW > 8) v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    const uint32_t base = static_cast<uint32_t>(__builtin_amdgcn_readlane(v - v0, static_cast<int>(wave)));
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(v, static_cast<int>(kW - 1)));
    excl = base + inc - x;
    __syncthreads();
    return total;
}

// Exclusive scan of n <= kThreads*kPer entries held in LDS (in place), thread
// t owning entries [t*kPer, (t+1)*kPer); returns the total.
template <uint32_t kThreads, uint32_t kPer>
__device__ __forceinline__ uint32_t lds_scan(uint32_t* v, uint32_t n, uint32_t* s_wave) {
    const uint32_t lo = threadIdx.x * kPer;
    uint32_t loc[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        loc[i] = lo + i < n ? v[lo + i] : 0u;
        sum += loc[i];
    }
    uint32_t excl;
    const uint32_t total = block_scan<kThreads>(sum, s_wave, excl);
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        if (lo + i < n) v[lo + i] = excl;
        excl += loc[i];
    }
    __syncthreads();
    return total;
}

template <uint32_t kPer>
__device__ __forceinline__ uint32_t lds_exclusive_scan(uint32_t* v, uint32_t n, uint32_t* s_wave) {
    return lds_scan<kTile, kPer>(v, n, s_wave);
}

// Largest index i in [lo0, hi) with v[i] <= x, given v[lo0] <= x and v
// nondecreasing; fixed trip count (steps of kSpan/2, ..., 1) so independent
// searches interleave.
template <uint32_t kSpan>
__device__ __forceinline__ uint32_t floor_index(const uint32_t* v, uint32_t lo0, uint32_t hi, uint32_t x) {
    uint32_t lo = lo0;
#pragma unroll
    for (uint32_t s = kSpan / 2; s > 0; s >>= 1)
        if (lo + s < hi && v[lo + s] <= x) lo += s;
    return lo;
}

// First step of a search over n >= 1 entries: the largest power of two
// below n (0 when n == 1), so that the steps top, top/2, ..., 1 reach every
// index of [0, n).
__device__ __forceinline__ uint32_t search_top(uint32_t n) {
    return n > 1u ? 1u << (31 - __builtin_clz(n - 1u)) : 0u;
}


// CSR extent of row i: first edge and out-degree.  A network whose rows all
// have the same out-degree d (uni_deg = d) computes it instead of reading
// row_ptr, which takes a dependent global load off the spike path.
__device__ __forceinline__ void row_extent(const NetArgs& a, uint32_t i, uint32_t& r0, uint32_t& d) {
    if (a.uni_deg) {
        r0 = i * a.uni_deg;
        d = a.uni_deg;
    } else {
        r0 = a.row_ptr[i];
        d = a.row_ptr[i + 1] - r0;
    }
}

// The network's population table in LDS: first neuron and first tile.
struct BinMap {
    uint32_t off[kMaxBinPops];
    uint32_t tile0[kMaxBinPops];
};

// Every thread of the block calls it; the table is valid after the caller's
// next barrier.
__device__ __forceinline__ void bin_map_load(const NetArgs& a, BinMap& M) {
    for (uint32_t p = threadIdx.x; p < a.n_pops; p += blockDim.x) {
        M.off[p] = a.pops[p].off;
        M.tile0[p] = a.pops[p].tile0;
    }
}

// (bin << 11) | cell of global post `post`.
__device__ __forceinline__ uint32_t post_bin_cell(const BinMap& M, uint32_t n_pops, uint32_t post) {
    uint32_t p = 0;  // last population starting at or before post
    for (uint32_t s = search_top(n_pops); s > 0; s >>= 1)
        if (p + s < n_pops && M.off[p + s] <= post) p += s;
    const uint32_t r = post - M.off[p];
    const uint32_t tile = M.tile0[p] + r / kTile;
    return ((tile / kBinTiles) << 11) | ((tile % kBinTiles) * kTile + r % kTile);
}

// Whether neuron gi spiked this step (its population's spike word).
__device__ __forceinline__ bool spiked_now(const NetArgs& a, uint32_t gi) {
    uint32_t p = 0;
    if (a.n_pops > 1) {  // last population starting at or before gi
        uint32_t lo = 0, hi = a.n_pops;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.pops[mid].off <= gi) lo = mid; else hi = mid;
        }
        p = lo;
    }
    const PopDev& P = a.pops[p];
    const uint32_t li = gi - P.off;
    return (a.bits[P.word0 + li / 64u] >> (li % 64u)) & 1ull;
}

// One STDP update of synapse e: FP64 master weight +- dw, clipped to
// [w_min, w_max], FP32 mirror for the scatter; returns the mirror.
__device__ __forceinline__ float stdp_apply(const NetArgs& a, uint32_t e, float dw, bool add) {
    double nw = add ? a.w64[e] + static_cast<double>(dw) : a.w64[e] - static_cast<double>(dw);
    nw = fmin(fmax(nw, static_cast<double>(a.w_min)), static_cast<double>(a.w_max));
    a.w64[e] = nw;
    const float f = static_cast<float>(nw);
    const_cast<dptr<float>>(a.w)[e] = f;
    return f;
}

// The STDP of K out-edges e[u] (distinct edges of rows whose pre spiked this
// step; u with on[u]), then their weights: LTP first when the post spiked
// too, then LTD (the order of the reference's two passes,
// P/src/api/learn/brain.rs:204-286), each an FP64 master update clipped to
// [w_min, w_max] with the FP32 mirror stored for the scatter; non-plastic
// edges keep their weight.  Every load of a level is issued before any is
// used -- the edge's flag and master, then the post's spike bit and trace --
// and the stores come last, so the K chains overlap instead of queueing
// behind each other's stores; the two updates of an edge are applied in
// registers (the arithmetic of two stdp_apply calls) and stored once.
template <uint32_t K>
__device__ __forceinline__ void stdp_out_edges(const NetArgs& a, const uint32_t (&e)[K], const uint32_t (&post)[K],
                                               const bool (&on)[K], float (&wout)[K]) {
    double m[K];
    bool pl[K], sp[K];
    float trq[K];
#pragma unroll
    for (uint32_t u = 0; u < K; ++u) {
        pl[u] = on[u] && (!a.plastic || a.plastic[e[u]]);
        m[u] = pl[u] ? a.w64[e[u]] : 0.0;
        if (on[u] && !pl[u]) wout[u] = a.w[e[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < K; ++u) {
        sp[u] = pl[u] && spiked_now(a, post[u]);
        trq[u] = pl[u] ? a.tr_post[post[u]] : 0.0f;
    }
    const double lo = static_cast<double>(a.w_min), hi = static_cast<double>(a.w_max);
#pragma unroll
    for (uint32_t u = 0; u < K; ++u) {
        if (!pl[u]) continue;
        double nw = m[u];
        if (sp[u]) nw = fmin(fmax(nw + static_cast<double>(a.a_plus * a.tr_pre[a.edge_pre[e[u]]]), lo), hi);
        nw = fmin(fmax(nw - static_cast<double>(a.a_minus * trq[u]), lo), hi);
        m[u] = nw;
        wout[u] = static_cast<float>(nw);
    }
#pragma unroll
    for (uint32_t u = 0; u < K; ++u) {
        if (!pl[u]) continue;
        a.w64[e[u]] = m[u];
        const_cast<dptr<float>>(a.w)[e[u]] = wout[u];
    }
}

// LTP on the in-edges of the spiking neurons of neuron tiles [t0, t1) whose
// pre did not spike this step (a spiking pre's out-edge walk takes the
// others).  Every thread of the block calls it; s_rs/s_pref hold kTile.
__device__ void stdp_in_edges(dptr<const NetArgs> nap, uint32_t t0, uint32_t t1, uint32_t* s_rs, uint32_t* s_pref,
                              uint32_t* s_wave, uint32_t* s_cnt) {
    const NetArgs& a = *nap;
    const uint32_t t = threadIdx.x;
    const float ap = a.a_plus;
    __syncthreads();  // the caller's use of s_cnt is over
    for (uint32_t b = t0 + t; b < t1; b += blockDim.x) s_cnt[b - t0] = a.tile_count[b];  // all tiles at once
    __syncthreads();
    for (uint32_t b = t0; b < t1; ++b) {
        const uint32_t cnt = s_cnt[b - t0];
        if (cnt == 0) continue;
        __syncthreads();  // the previous tile's rows are read
        uint32_t rs = 0, dg = 0;
        if (t < cnt) {
            const uint32_t id = a.active[static_cast<size_t>(b) * kTile + t];
            rs = a.in_ptr[id];
            dg = a.in_ptr[id + 1] - rs;
        }
        uint32_t excl;
        const uint32_t total = block_exclusive_scan(dg, s_wave, excl);
        if (t < cnt) {
            s_rs[t] = rs;
            s_pref[t] = excl;
        }
        __syncthreads();
        for (uint32_t f = t; f < total; f += blockDim.x) {
            uint32_t lo = 0, hi = cnt;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pref[mid] <= f) lo = mid; else hi = mid;
            }
            const uint32_t e = a.in_edge[s_rs[lo] + (f - s_pref[lo])];
            // the edge's flag, pre and master together, then the pre's spike
            // bit and trace together (stdp_apply's arithmetic, one store)
            const bool pl = !a.plastic || a.plastic[e];
            const uint32_t pre = a.edge_pre[e];
            const double m = a.w64[e];
            if (!pl) continue;
            const bool pre_spiked = spiked_now(a, pre);
            const float trp = a.tr_pre[pre];
            if (pre_spiked) continue;
            double nw = m + static_cast<double>(ap * trp);
            nw = fmin(fmax(nw, static_cast<double>(a.w_min)), static_cast<double>(a.w_max));
            a.w64[e] = nw;
            const_cast<dptr<float>>(a.w)[e] = static_cast<float>(nw);
        }
    }
}

// Records of buffer p (megastep networks keep two, by step parity).
__device__ __forceinline__ dptr<uint2> ev_buf(const NetArgs& a, uint32_t p) {
    return a.bin_ev + static_cast<size_t>(p) * a.n_ev * (a.ev_wide ? 2u : 1u);
}

// Record `pos` of a writer: event id eid (f-th of the writer's events this
// step), weight x, (bin << 11 | cell) bc, receptor code.  kNarrow: the
// network is known to use 8-byte records (megastep networks).
template <bool kRec, bool kNarrow = false>
__device__ __forceinline__ void put_record(const NetArgs& a, dptr<uint2> ev, uint32_t pos, uint32_t eid, uint32_t f,
                                           float x, uint32_t bc, uint32_t code) {
    if (kRec || (!kNarrow && a.ev_wide))
        reinterpret_cast<dptr<uint4>>(ev)[pos] = make_uint4(eid, __float_as_uint(x), bc & (kBinPosts - 1u), code);
    else
        ev[pos] = make_uint2(__float_as_uint(x), (f << kBinLog2) | (bc & (kBinPosts - 1u)));
}

// Count (kWrite false) or write out (kWrite true) the events of `cnt` staged
// rows (cnt <= 2048): row k starts at edge s_rs[k] and holds concatenation
// positions [s_pref[k], s_pref[k+1]).  Counting adds 1 to s_bin[bin]; writing
// puts the event at record s_bin[bin] of `ev` and advances it; the writer's
// events before these rows number fb.  Four events per thread are loaded
// before any is used.
template <bool kWrite, uint32_t kThreads, bool kRec = false, bool kStdp = false, bool kNarrow = false>
__device__ __forceinline__ void bin_events(const NetArgs& a, const BinMap& M, bool is_src, uint32_t cnt,
                                           uint32_t total, const uint32_t* s_rs, const uint32_t* s_pref,
                                           uint32_t* s_bin, dptr<uint2> ev, uint64_t step, uint32_t fb = 0u) {
    constexpr uint32_t kU = 4;
    const uint32_t* col = is_src ? a.src_post : a.col;
    const float* w = is_src ? a.src_w : a.w;
    const uint32_t eb = is_src ? static_cast<uint32_t>(a.n_syn) : 0u, np = a.n_pops;
    for (uint32_t f0 = threadIdx.x; f0 < total; f0 += kU * kThreads) {
        uint32_t e[kU], post[kU];
        float wt[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t f = f0 + u * kThreads;
            e[u] = 0u;
            if (f < total) {
                const uint32_t k = floor_index<2048>(s_pref, 0, cnt, f);
                e[u] = s_rs[k] + (f - s_pref[k]);
            }
        }
        if (kWrite && kStdp && !is_src) {
            bool on[kU];
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) {
                on[u] = f0 + u * kThreads < total;
                if (on[u]) post[u] = col[e[u]];
            }
            stdp_out_edges<kU>(a, e, post, on, wt);
        } else {
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) {
                if (f0 + u * kThreads < total) {
                    post[u] = col[e[u]];
                    if (kWrite) wt[u] = w[e[u]];
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            if (f0 + u * kThreads >= total) continue;
            const uint32_t bc = post_bin_cell(M, np, post[u]);
            if (!kWrite) {
                atomicAdd(&s_bin[bc >> 11], 1u);
            } else {
                const float x = a.tm_type ? tm_deliver(a, eb + e[u], wt[u], step) : wt[u];
                const uint32_t code = kRec && !is_src ? static_cast<uint32_t>(a.rec_code[e[u]]) : 0u;
                const uint32_t pos = atomicAdd(&s_bin[bc >> 11], 1u);
                put_record<kRec, kNarrow>(a, ev, pos, eb + e[u], fb + f0 + u * kThreads, x, bc, code);
            }
        }
    }
}

// The block's events when there are at most kEvCache per thread: loaded once
// (edge id, bin and cell, weight) and held in registers from the count to
// the write.
constexpr uint32_t kEvCache = 4;
struct EvCache {
    uint32_t e[kEvCache], bc[kEvCache];
    float w[kEvCache];
};

template <uint32_t kThreads, bool kStdp = false>
__device__ __forceinline__ void ev_cache_load(const NetArgs& a, const BinMap& M, bool is_src, uint32_t cnt,
                                              uint32_t total, const uint32_t* s_rs, const uint32_t* s_pref,
                                              EvCache& c) {
    const uint32_t* col = is_src ? a.src_post : a.col;
    const float* w = is_src ? a.src_w : a.w;
    uint32_t post[kEvCache];
    if (a.uni_deg && !is_src) {
        // every row holds d edges, so event f is edge f mod d of row f / d:
        // the quotient from one float multiply (error below one for f <
        // 2^22 d) corrected by one, in place of the search over s_pref
        const uint32_t d = a.uni_deg;
        const float inv = 1.0f / static_cast<float>(d);
#pragma unroll
        for (uint32_t u = 0; u < kEvCache; ++u) {
            const uint32_t f = threadIdx.x + u * kThreads;
            c.e[u] = 0u;
            if (f < total) {
                uint32_t k = static_cast<uint32_t>(static_cast<float>(f) * inv);
                if (k * d > f) --k;
                else if ((k + 1u) * d <= f) ++k;
                c.e[u] = s_rs[k] + (f - k * d);
            }
        }
    } else {
#pragma unroll
        for (uint32_t u = 0; u < kEvCache; ++u) {
            const uint32_t f = threadIdx.x + u * kThreads;
            c.e[u] = 0u;
            if (f < total) {
                const uint32_t k = floor_index<2048>(s_pref, 0, cnt, f);
                c.e[u] = s_rs[k] + (f - s_pref[k]);
            }
        }
    }
    if (kStdp && !is_src) {
        bool on[kEvCache];
#pragma unroll
        for (uint32_t u = 0; u < kEvCache; ++u) {
            on[u] = threadIdx.x + u * kThreads < total;
            if (on[u]) post[u] = col[c.e[u]];
        }
        stdp_out_edges<kEvCache>(a, c.e, post, on, c.w);
    } else {
#pragma unroll
        for (uint32_t u = 0; u < kEvCache; ++u) {
            if (threadIdx.x + u * kThreads < total) {
                post[u] = col[c.e[u]];
                c.w[u] = w[c.e[u]];
            }
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < kEvCache; ++u)
        if (threadIdx.x + u * kThreads < total) c.bc[u] = post_bin_cell(M, a.n_pops, post[u]);
}

template <uint32_t kThreads>
__device__ __forceinline__ void ev_cache_count(const EvCache& c, uint32_t total, uint32_t* s_bin) {
#pragma unroll
    for (uint32_t u = 0; u < kEvCache; ++u)
        if (threadIdx.x + u * kThreads < total) atomicAdd(&s_bin[c.bc[u] >> 11], 1u);
}

template <uint32_t kThreads, bool kRec = false, bool kNarrow = false>
__device__ __forceinline__ void ev_cache_write(const NetArgs& a, bool is_src, const EvCache& c, uint32_t total,
                                               uint32_t* s_bin, dptr<uint2> ev, uint64_t step) {
    const uint32_t eb = is_src ? static_cast<uint32_t>(a.n_syn) : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kEvCache; ++u) {
        const uint32_t f = threadIdx.x + u * kThreads;
        if (f >= total) continue;
        const uint32_t e = eb + c.e[u];
        const float x = a.tm_type ? tm_deliver(a, e, c.w[u], step) : c.w[u];
        const uint32_t code = kRec && !is_src ? static_cast<uint32_t>(a.rec_code[e]) : 0u;
        const uint32_t pos = atomicAdd(&s_bin[c.bc[u] >> 11], 1u);
        put_record<kRec, kNarrow>(a, ev, pos, e, f, x, c.bc[u], code);
    }
}

// Writer j's per-bin counts (s_bin) become the absolute first record of each
// bin's group, from `base`; column j of `hdr` gets {start, count} for each
// bin it sends events to (consumers zero the entries they read, so the
// others read as empty), or, in the one-kernel and split forms, for every
// bin (their consumers leave the row as read: one write per entry and step
// instead of a write and a clearing write for most).  Synchronises before
// and after.
template <uint32_t kThreads>
__device__ __forceinline__ void bin_offsets(const NetArgs& a, uint32_t* s_bin, uint32_t* s_wave, uint32_t j,
                                            uint32_t base, uint32_t events, uint2* hdr) {
    __syncthreads();
    const uint32_t nb = a.n_bins, t = threadIdx.x;
    if (nb <= kThreads) {  // one bin per thread: its count, offset and header at once
        const uint32_t c = t < nb ? s_bin[t] : 0u;
        uint32_t excl;
        (void)block_scan<kThreads>(c, s_wave, excl);
        if (t < nb) {
            s_bin[t] = base + excl;
            if (c || a.mega) hdr[static_cast<size_t>(t) * a.bin_rows + j] = make_uint2(base + excl, c);
        }
        __syncthreads();
        return;
    }
    if (nb <= 4 * kThreads) lds_scan<kThreads, 4>(s_bin, nb, s_wave);
    else lds_scan<kThreads, (kMaxBins + kThreads - 1) / kThreads>(s_bin, nb, s_wave);
    for (uint32_t k = t; k < nb; k += kThreads) {
        const uint32_t o = s_bin[k], e = k + 1 < nb ? s_bin[k + 1] : events;
        if (e != o || a.mega) hdr[static_cast<size_t>(k) * a.bin_rows + j] = make_uint2(base + o, e - o);
    }
    __syncthreads();
    for (uint32_t k = t; k < nb; k += kThreads) s_bin[k] += base;
    __syncthreads();
}

// Megastep and split form: the network's last block of the step (of n_done)
// moves the step counter on and clears the block count and the ring flag.
__device__ __forceinline__ void mega_done(const NetArgs& a, uint64_t step) {
    if (atomicAdd(a.bin_done, 1u) + 1u == a.n_done) {
        *a.bin_done = 0u;
        *a.bin_mat = 0u;
        *a.step = step + 1;
    }
}

// Conductance-based receptor synapses (SynapticKinetics, P/src/kernels/
// synapse/kinetics.rs:112-205, in the step order of CN/src/sim/hh.rs:115-166):
// every conductance decays by its population's factor, the constant external
// current joins the input, then each receptor adds g_r (E_r - v) with v the
// membrane potential before the update; NMDA's g is scaled by the Mg2+ block
// B(v) = 1 / (1 + 0.33 fast_exp(clamp(-0.06 v, -10, 10))) (receptor.rs:
// 126-150).  The reference adds neurons below n8 = N & ~7 in 8-wide lanes
// straight onto the input (input + t_0 + ... + t_6) and the rest in a scalar
// tail (input + (t_0 + ... + t_6)); both orders are kept.  The conductances
// are loaded (rec_load) with the neuron's state, before any store of the step.
struct RecIn {
    float g[kReceptors];
    float ext;
};
__device__ __forceinline__ void rec_load(const NetArgs& a, uint32_t gi, RecIn& r) {
#pragma unroll
    for (int k = 0; k < kReceptors; ++k)
        r.g[k] = ((a.rec_live >> k) & 1u) ? a.g[static_cast<size_t>(k) * a.N + gi] : 0.0f;
    r.ext = a.g_ext ? a.g_ext[gi] : 0.0f;
}
// v: the membrane potential before the update; r: the neuron's conductances
// and external current, loaded by rec_load
__device__ __forceinline__ float receptor_input(const NetArgs& a, const PopDev& P, uint32_t gi, float v,
                                                const RecIn& r, float I) {
    const float i0 = a.g_ext ? I + r.ext : I;
    float x = -0.06f * v;
    x = x < -10.0f ? -10.0f : x;
    x = x > 10.0f ? 10.0f : x;
    const float block = 1.0f / (1.0f + 0.33f * fast_expf(x));
    float t[kReceptors];
#pragma unroll
    for (int k = 0; k < kReceptors; ++k) {
        const float g = r.g[k] * P.gdec[k];
        if ((a.rec_live >> k) & 1u) a.g[static_cast<size_t>(k) * a.N + gi] = g;
        const float ge = k == kRecNmda ? g * block : g;
        t[k] = ge * (kRecErev[k] - v);
    }
    if (gi < a.n8) {
        float cur = i0;
#pragma unroll
        for (int k = 0; k < kReceptors; ++k) cur = cur + t[k];
        return cur;
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kReceptors; ++k) acc = acc + t[k];
    return i0 + acc;
}

// One step of up to kMaxGroup consecutive tiles of one population: read ring
// slot s%32 and zero it, add the Poisson drive and Gaussian noise, clamp
// (LIF/Izhikevich), advance the neuron, update STDP traces, then wave-ballot
// each tile's spike mask and write its spiking ids in ascending order.  For
// networks without plasticity the block's spikes are scattered here as well
// (the push only needs the spiking ids, which the block already holds); with
// plasticity the weights change between update and push, so insert_kernel
// scatters after the STDP kernels.  Work item: {network, first tile, tiles,
// population}.
template <int MODEL, int SCHED>
__global__ __launch_bounds__(kTile) void neuron_kernel(const NetArgs* __restrict__ nets,
                                                      const uint4* __restrict__ items, uint32_t defer_scatter) {
    __shared__ uint32_t s_rs[kMaxGroup * kTile], s_pref[kMaxGroup * kTile];
    __shared__ uint32_t s_cnt[kMaxGroup][kWavesPerTile], s_deg[kMaxGroup][kWavesPerTile];
    __shared__ uint32_t s_drv[kMaxGroup][kWavesPerTile];
    __shared__ uint32_t s_ncoll;
    __shared__ uint32_t s_bw[kWavesPerTile];
    __shared__ BinMap s_map;
    extern __shared__ uint32_t s_bin[];  // [n_bins] of binned writers (dynamic)
    const uint4 it = items[blockIdx.x];
    if (it.x == kEmpty) return;  // pad of the XCD-ordered list
    const NetArgs& a = nets[it.x];
    const PopDev& P = a.pops[it.w];
    const uint32_t tile0 = it.y, ntile = it.z & 15u, wj = it.z >> 4;
    const bool fused = a.stdp == 0 && !defer_scatter && a.n_bins == 0;  // chained scatter of the block's spikes
    const bool fused_bin = a.bin_fused && !defer_scatter;               // binned writer wj of its spikes
    const bool inline_push = fused || fused_bin;
    const uint32_t ebase = fused_bin ? a.tile_ebase[tile0] : 0u;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t li0 = (tile0 - P.tile0) * kTile + t;

    // Phase 1: every load of the tile(s).  The step counter comes first: the
    // ring slot's address depends on it, and the slot's read then overlaps
    // the state loads.
    const uint64_t step = *a.step;
    float* ring_slot = a.ring + static_cast<size_t>(step & (kSlots - 1)) * a.N + P.off;
    Neuron<MODEL> nrn[kMaxGroup];
    RecIn rin[kMaxGroup];
    float trp[kMaxGroup], trq[kMaxGroup];  // STDP traces
    uint32_t row0[kMaxGroup], deg[kMaxGroup];
    float cur[kMaxGroup];
#pragma unroll
    for (uint32_t g = 0; g < kMaxGroup; ++g) {
        row0[g] = deg[g] = 0u;
        cur[g] = 0.0f;
        const uint32_t li = li0 + g * kTile;
        if (g < ntile && li < P.n) {
            nrn[g].load(P, li);
            if (a.rec_code) rec_load(a, P.off + li, rin[g]);
            if (a.stdp) {
                trp[g] = a.tr_pre[P.off + li];
                trq[g] = a.tr_post[P.off + li];
            }
            if (inline_push) {
                row_extent(a, P.off + li, row0[g], deg[g]);
            }
            cur[g] = ring_slot[li];
        }
    }
    if (a.n_bins && tile0 == 0 && t == 0)  // the slot bin_apply fills
        *a.bin_slot = step < a.step_limit ? static_cast<uint32_t>((step + 1) & (kSlots - 1)) : kEmpty;
    if (step >= a.step_limit) return;
    if (fused_bin) {
        for (uint32_t k = t; k < a.n_bins; k += kTile) s_bin[k] = 0u;
        bin_map_load(a, s_map);
    }
#pragma unroll
    for (uint32_t g = 0; g < kMaxGroup; ++g) {
        const uint32_t li = li0 + g * kTile;
        if (g < ntile && li < P.n) ring_slot[li] = 0.0f;
    }

    // Phase 2: advance every neuron; per tile ballot the spikes and scan the
    // spiking rows' lengths within each wave.
    const float dt = P.puni[0];
    uint64_t m[kMaxGroup];
    uint32_t excl[kMaxGroup];  // edge prefix of the thread's row within its wave
#pragma unroll
    for (uint32_t g = 0; g < kMaxGroup; ++g) {
        const uint32_t li = li0 + g * kTile;
        int spike = 0, drive_now = 0;
        if (g < ntile && li < P.n) {
            float I = cur[g];
            // A per-neuron Poisson source with one edge of delay <= 1 adds
            // its weight into the slot after every neuron push of the previous
            // step (P/src/api/net/network/step.rs:166-185): the last add.
            if ((P.flags & kPopDrive) && P.drive_p > 0.0f) {
                const uint32_t sd = P.drive_seed0 + li * 2654435761u;
                if (step > 0 && poisson_roll(sd, 0u, step - 1, P.drive_p)) I = I + P.drive_w;
                drive_now = poisson_roll(sd, 0u, step, P.drive_p) ? 1 : 0;
            }
            // (coupled networks took their noise into the slot, before the coupling)
            if ((P.flags & kPopNoise) && !a.cp_ptr) {
                const float sg = (P.flags & kPopNoiseScalar) ? P.sigma_uni : P.sigma[li];
                I = I + sg * normal_draw(P.seed, li, step);
            }
            if (a.rec_code) I = receptor_input(a, P, P.off + li, nrn[g].vm(P, li), rin[g], I);
            spike = nrn[g].step(P, li, I, dt);
            nrn[g].store(P, li);
            if (a.stdp) {
                const uint32_t gi = P.off + li;
                const float add = spike ? 1.0f : 0.0f;
                a.tr_pre[gi] = fmaf(trp[g], a.stdp_dpre, add);
                a.tr_post[gi] = fmaf(trq[g], a.stdp_dpost, add);
            }
        }
        m[g] = __ballot(spike);
        const uint32_t drive_events = static_cast<uint32_t>(__popcll(__ballot(drive_now)));
        const uint32_t d = spike ? deg[g] : 0u;
        uint32_t inc = d;
        if (m[g]) {
            inc = wave_incl_scan(inc);
        }
        excl[g] = inc - d;
        if (g < ntile) {
            if (lane == 63) s_deg[g][wave] = inc;
            if (lane == 0) {
                s_cnt[g][wave] = static_cast<uint32_t>(__popcll(m[g]));
                s_drv[g][wave] = drive_events;
                const uint32_t word = P.word0 + (tile0 - P.tile0 + g) * kWavesPerTile + wave;
                a.bits[word] = m[g];
                if (a.raster) {
                    const uint64_t rs = step - *a.raster_base;
                    if (rs < a.raster_cap) a.raster[rs * a.n_words + word] = m[g];
                }
            }
        }
    }
    if (t == 0) s_ncoll = 0u;
    __syncthreads();

    // Phase 3: per-tile ordered compaction; the block's spiking rows are
    // staged in (tile, wave, lane) order with their edge prefix.
    uint32_t blk_sp = 0, blk_ed = 0, drv = 0;
#pragma unroll
    for (uint32_t g = 0; g < kMaxGroup; ++g) {
        if (g < ntile) {
            uint32_t base = 0, total = 0, dbase = 0, dtot = 0;
#pragma unroll
            for (uint32_t w = 0; w < kWavesPerTile; ++w) {
                const uint32_t c = s_cnt[g][w], d = s_deg[g][w];
                if (w < wave) { base += c; dbase += d; }
                total += c;
                dtot += d;
                drv += s_drv[g][w];
            }
            if ((m[g] >> lane) & 1ull) {
                const uint32_t rank = base + static_cast<uint32_t>(__popcll(m[g] & ((1ull << lane) - 1ull)));
                a.active[static_cast<size_t>(tile0 + g) * kTile + rank] = P.off + li0 + g * kTile;
                if (inline_push) {
                    s_rs[blk_sp + rank] = row0[g];
                    s_pref[blk_sp + rank] = blk_ed + dbase + excl[g];
                }
            }
            if (t == 0) a.tile_count[tile0 + g] = total;
            blk_sp += total;
            blk_ed += dtot;
        }
    }
    if (t == 0) {
        if (blk_sp) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_spikes + tile0), blk_sp);
        // each firing drive source is one synaptic event, counted at its roll;
        // the fused scatter's events are counted here too
        const uint32_t ev = drv + (inline_push ? blk_ed : 0u);
        if (ev) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_events + tile0), ev);
    }
    if (fused_bin) {
        __syncthreads();
        const bool cached = blk_ed <= kEvCache * kTile;
        EvCache ec;
        if (cached) {
            ev_cache_load<kTile>(a, s_map, false, blk_sp, blk_ed, s_rs, s_pref, ec);
            ev_cache_count<kTile>(ec, blk_ed, s_bin);
        } else {
            bin_events<false, kTile>(a, s_map, false, blk_sp, blk_ed, s_rs, s_pref, s_bin, a.bin_ev, step);
        }
        bin_offsets<kTile>(a, s_bin, s_bw, wj, ebase, blk_ed, a.bin_hdr);
        if (cached) ev_cache_write<kTile>(a, false, ec, blk_ed, s_bin, a.bin_ev, step);
        else bin_events<true, kTile>(a, s_map, false, blk_sp, blk_ed, s_rs, s_pref, s_bin, a.bin_ev, step);
        return;
    }
    if (!fused) return;
    // the block's collisions are listed from its first tile's base
    if (t < ntile && (t > 0 || blk_sp == 0)) a.tile_coll[tile0 + t] = 0u;
    if (blk_sp == 0) return;
    __syncthreads();
    const ScatterCtx c{step, static_cast<uint32_t>(step), a.tile_ebase[tile0], &s_ncoll};
    scatter_rows<SCHED>(a, c, false, blk_sp, blk_ed, s_rs, s_pref);
    __syncthreads();
    if (t == 0) a.tile_coll[tile0] = s_ncoll;
}

// Stage the spiking rows of scatter tile b in LDS: a neuron tile's ascending
// spiking ids (from the neuron kernel), or the sources of a source tile that
// fire this step (P/src/api/net/network/step.rs:166-185).  Row k starts at
// edge s_rs[k] (of the tile's own CSR), s_pref[k] is its first edge in the
// concatenation.  Returns the number of rows; *total gets the edge count.
// Every thread of the block must call it (it synchronises).
__device__ __forceinline__ uint32_t stage_tile_rows(const NetArgs& a, uint32_t b, uint64_t step, uint32_t* s_rs,
                                                    uint32_t* s_pref, uint32_t* s_wave, uint32_t* s_c,
                                                    uint32_t* total) {
    const uint32_t t = threadIdx.x;
    const bool is_src = b >= a.n_tiles;
    const uint32_t* rp = is_src ? a.src_row_ptr : a.row_ptr;
    uint32_t cnt, rs = 0, dg = 0;
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
        const uint32_t lane = t & 63u, wave = t >> 6;
        if (lane == 0) s_c[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t base = 0;
        cnt = 0;
        for (uint32_t w = 0; w < kWavesPerTile; ++w) {
            if (w < wave) base += s_c[w];
            cnt += s_c[w];
        }
        if (fired) s_rs[base + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)))] = s;
        __syncthreads();
        if (t < cnt) {
            const uint32_t sid = s_rs[t];
            rs = rp[sid];
            dg = rp[sid + 1] - rs;
        }
        __syncthreads();
    }
    uint32_t excl;
    *total = block_exclusive_scan(dg, s_wave, excl);  // uniform; also a barrier
    if (t < cnt) {
        s_rs[t] = rs;
        s_pref[t] = excl;
    }
    __syncthreads();
    return cnt;
}

// Scatter of the neuron tiles of networks with plasticity (after the STDP
// kernels) and of every Poisson-source tile, for networks on the chained
// scatter.
template <int SCHED>
__global__ __launch_bounds__(kTile) void insert_kernel(const NetArgs* __restrict__ nets,
                                                      const uint2* __restrict__ blocks) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile], s_c[kWavesPerTile];
    __shared__ uint32_t s_ncoll;
    const uint2 ref = blocks[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    const uint32_t b = ref.y, t = threadIdx.x;
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;
    if (t == 0) s_ncoll = 0u;
    uint32_t total;
    const uint32_t cnt = stage_tile_rows(a, b, step, s_rs, s_pref, s_wave, s_c, &total);
    if (cnt) {
        const ScatterCtx c{step, static_cast<uint32_t>(step), a.tile_ebase[b], &s_ncoll};
        scatter_rows<SCHED>(a, c, b >= a.n_tiles, cnt, total, s_rs, s_pref);
    }
    __syncthreads();
    if (t == 0) {
        a.tile_coll[b] = s_ncoll;
        if (total) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_events + b), total);
    }
}

// Stand-alone binned writer: work item {network, first scatter tile, tiles,
// writer index}; neuron tiles (of plastic networks, after the STDP kernels,
// or of every network in the dense STDP step) and/or Poisson-source tiles.
// The spiking rows of all of its neuron tiles are staged together.
__global__ __launch_bounds__(kTile) void bin_kernel(const NetArgs* __restrict__ nets,
                                                   const uint4* __restrict__ refs) {
    constexpr uint32_t kRows = 1024;  // spiking rows staged per chunk
    constexpr uint32_t kRowsPer = kRows / kTile;
    __shared__ uint32_t s_rs[kRows], s_pref[kRows];
    __shared__ uint32_t s_srs[kTile], s_spref[kTile];
    __shared__ uint32_t s_tp[kTile];  // first spiking row of each neuron tile of the block
    __shared__ uint32_t s_wave[kWavesPerTile], s_c[kWavesPerTile];
    __shared__ BinMap M;
    extern __shared__ uint32_t s_bin[];  // [n_bins] (dynamic)
    const uint4 ref = refs[blockIdx.x];
    if (ref.x == kEmpty) return;  // pad of the XCD-ordered list
    const NetArgs& a = nets[ref.x];
    const uint32_t b0 = ref.y, b1 = min(ref.y + ref.z, a.n_blocks), j = ref.w, t = threadIdx.x;
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;
    const uint32_t n_nt = b0 < a.n_tiles ? min(b1, a.n_tiles) - b0 : 0u;  // neuron tiles (<= kTile)
    if (j == kLtpItem) {  // plastic networks: the in-edge LTP of tiles [b0, b1), beside the writers
        stdp_in_edges((dptr<const NetArgs>)&a, b0, b0 + n_nt, s_rs, s_pref, s_wave, s_tp);
        if (a.mega) {  // split form: counted with the writers, so the step moves on after both
            __syncthreads();
            if (t == 0) mega_done(a, step);
        }
        return;
    }
    const uint32_t base = a.tile_ebase[b0];
    // megastep networks: this step's buffer of records and headers
    const uint32_t par = a.mega ? static_cast<uint32_t>(step & 1u) : 0u;
    const dptr<uint2> ev = ev_buf(a, par);
    uint2* hdr = a.bin_hdr + par * a.hdr_stride;
    for (uint32_t k = t; k < a.n_bins; k += kTile) s_bin[k] = 0u;
    bin_map_load(a, M);
    const uint32_t tc = t < n_nt ? a.tile_count[b0 + t] : 0u;
    uint32_t excl;
    const uint32_t n_rows = block_exclusive_scan(tc, s_wave, excl);
    if (t < n_nt) s_tp[t] = excl;
    __syncthreads();
    const bool one = n_rows <= kRows;  // rows staged once, for both passes
    uint32_t events = 0, kept = 0;
    bool cached = false;  // one chunk of at most kEvCache events per thread
    EvCache ec;
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t fb = 0;  // the writer's events before the current rows
        for (uint32_t r0 = 0; r0 < n_rows; r0 += kRows) {
            const uint32_t nr = min(kRows, n_rows - r0);
            uint32_t total = kept;
            if (!(one && pass == 1)) {
                uint32_t id[kRowsPer];
#pragma unroll
                for (uint32_t u = 0; u < kRowsPer; ++u) {
                    const uint32_t i = t + u * kTile;
                    id[u] = 0u;
                    if (i < nr) {
                        const uint32_t r = r0 + i;
                        const uint32_t lo = floor_index<kTile>(s_tp, 0, n_nt, r);  // tile holding row r
                        id[u] = a.active[static_cast<size_t>(b0 + lo) * kTile + (r - s_tp[lo])];
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < kRowsPer; ++u) {
                    const uint32_t i = t + u * kTile;
                    if (i < nr) {
                        uint32_t rs, dg;
                        row_extent(a, id[u], rs, dg);
                        s_rs[i] = rs;
                        s_pref[i] = dg;
                    }
                }
                __syncthreads();
                total = lds_exclusive_scan<kRowsPer>(s_pref, nr, s_wave);
                kept = total;
            }
            if (pass == 0) {
                events += total;
                cached = one && total <= kEvCache * kTile;
                if (cached) {
                    // plastic networks: each spiking row's STDP happens as its
                    // events are loaded, exactly once
                    if (a.stdp) ev_cache_load<kTile, true>(a, M, false, nr, total, s_rs, s_pref, ec);
                    else ev_cache_load<kTile>(a, M, false, nr, total, s_rs, s_pref, ec);
                    ev_cache_count<kTile>(ec, total, s_bin);
                } else {
                    bin_events<false, kTile>(a, M, false, nr, total, s_rs, s_pref, s_bin, ev, step);
                }
            } else if (cached) {
                // receptor networks tag each synapse event with its receptor
                if (a.rec_code) ev_cache_write<kTile, true>(a, false, ec, total, s_bin, ev, step);
                else ev_cache_write<kTile>(a, false, ec, total, s_bin, ev, step);
            } else if (a.stdp) {
                if (a.rec_code)
                    bin_events<true, kTile, true, true>(a, M, false, nr, total, s_rs, s_pref, s_bin, ev, step, fb);
                else bin_events<true, kTile, false, true>(a, M, false, nr, total, s_rs, s_pref, s_bin, ev, step, fb);
            } else if (a.rec_code) {
                bin_events<true, kTile, true>(a, M, false, nr, total, s_rs, s_pref, s_bin, ev, step, fb);
            } else {
                bin_events<true, kTile>(a, M, false, nr, total, s_rs, s_pref, s_bin, ev, step, fb);
            }
            fb += total;
            if (!one) __syncthreads();  // before the next chunk is staged
        }
        // source tiles (their rows are the sources that fire this step)
        for (uint32_t b = max(b0, a.n_tiles); b < b1; ++b) {
            uint32_t total;
            const uint32_t cnt = stage_tile_rows(a, b, step, s_srs, s_spref, s_wave, s_c, &total);
            if (pass == 0) {
                events += total;
                bin_events<false, kTile>(a, M, true, cnt, total, s_srs, s_spref, s_bin, ev, step);
            } else {
                bin_events<true, kTile>(a, M, true, cnt, total, s_srs, s_spref, s_bin, ev, step, fb);
            }
            fb += total;
            __syncthreads();
        }
        if (pass == 0) bin_offsets<kTile>(a, s_bin, s_wave, j, base, events, hdr);
    }
    if (t == 0 && events) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_events + b0), events);
    if (a.mega) {
        __syncthreads();
        if (t == 0) mega_done(a, step);
    }
}

// bin_apply and the megastep run kApply threads per block: two cells, two
// records and two writer headers per thread.
constexpr uint32_t kApply = 1024;
constexpr uint32_t kApplyWaves = kApply / 64;

// LDS of a consumer block.  A pass's records live in r_w and in `u`: while
// lists are built, u holds r_eid[kBinCap] (u32), nxt[kBinCap] (u16, the
// per-cell list links) and r_post[kBinCap] (u16, the cell); the sort of long
// lists reuses all of u as kBinCap 64-bit keys.
struct BinApplyLds {
    float slice[kBinPosts];
    uint32_t head[kBinPosts];
    float r_w[kBinCap];
    uint64_t u[kBinCap];
    uint32_t seg_base[kMaxWriters];
    uint32_t seg_pref[kMaxWriters + 1];
    uint32_t wave[kApplyWaves];
    uint16_t mq[kBinCap / 2];  // cells of the pass with two or more records (each holds two of <= kBinCap)
    uint32_t m, next, nm;
    __device__ uint32_t* r_eid() { return reinterpret_cast<uint32_t*>(u); }
    __device__ uint16_t* nxt() { return reinterpret_cast<uint16_t*>(u + kBinCap / 2); }
    __device__ uint16_t* r_post() { return reinterpret_cast<uint16_t*>(u + 3 * kBinCap / 4); }
};
constexpr uint32_t kListNil = 0xFFFFFFFFu, kListLong = 0xFFFFFFFEu;
constexpr uint16_t kLinkNil = 0xFFFFu;
constexpr int kShortList = 8;

// Record q of the pass: cell, event id, weight; receptor networks keep the
// record's receptor code (0 current, 1 + type) above the cell's 11 bits.
__device__ __forceinline__ void bin_link(BinApplyLds& L, uint32_t q, uint32_t cell, uint32_t eid, float w,
                                         uint32_t code = 0u) {
    L.r_eid()[q] = eid;
    L.r_w[q] = w;
    L.r_post()[q] = static_cast<uint16_t>(cell | (code << kBinLog2));
    L.nxt()[q] = static_cast<uint16_t>(atomicExch(&L.head[cell], q));  // kListNil -> kLinkNil
}

// Receptor networks: the bin's consumer block adds each conductance jump to
// g[r][post] (CN/src/propagate.rs:225-290: g += w * scale, in push order),
// starting from the stored conductance; `touched` marks the loaded ones.
struct RecCtx {
    dptr<const NetArgs> a;
    uint32_t k;                 // bin
    uint32_t c0 = 0u, c1 = kBinPosts;  // the cells this block applies (split form: part of the bin)
};

// Global neuron of cell c of bin k (a cell that received events).
__device__ __forceinline__ uint32_t cell_post(const NetArgs& a, uint32_t k, uint32_t c) {
    const uint32_t tile = k * kBinTiles + c / kTile;
    const PopDev& P = a.pops[a.tile_pop[tile]];
    return P.off + (tile - P.tile0) * kTile + c % kTile;
}

__device__ __forceinline__ void rec_jump(const NetArgs& a, uint32_t post, uint32_t code, float w,
                                         float (&g)[kReceptors], uint32_t& touched) {
    const float x = w * a.rec_scale;
#pragma unroll
    for (int r = 0; r < kReceptors; ++r)
        if (code == static_cast<uint32_t>(r + 1)) {
            if (!((touched >> r) & 1u)) {
                g[r] = a.g[static_cast<size_t>(r) * a.N + post];
                touched |= 1u << r;
            }
            g[r] = g[r] + x;
        }
}

__device__ __forceinline__ void rec_store(const NetArgs& a, uint32_t post, const float (&g)[kReceptors],
                                          uint32_t touched) {
#pragma unroll
    for (int r = 0; r < kReceptors; ++r)
        if ((touched >> r) & 1u) a.g[static_cast<size_t>(r) * a.N + post] = g[r];
}

// Insert (e, w) into the ascending ids[0..W) holding fewer than W entries
// (the free slots hold 0xFFFFFFFF): every slot holding a larger id takes its
// left neighbour's entry, or (e, w) if that one is smaller.
template <int W, int N>
__device__ __forceinline__ void sorted_insert(uint32_t (&ids)[N], float (&ws)[N], uint32_t e, float w) {
#pragma unroll
    for (int q = W - 1; q >= 0; --q) {
        const int qp = q > 0 ? q - 1 : 0;
        const bool prev_gt = q > 0 && ids[qp] > e;
        if (ids[q] > e) {
            ids[q] = prev_gt ? ids[qp] : e;
            ws[q] = prev_gt ? ws[qp] : w;
        }
    }
}

// The current-only form of the short lists: the first record needs no
// compare, a list of one (most cells that receive anything) is one add, and
// lists of two to four use a four-slot insertion, so that a wavefront runs
// the eight-slot network only when one of its cells has five or more.
// Returns whether the list was longer than kShortList (left for the sort).
__device__ __forceinline__ bool short_list_sum(BinApplyLds& L, uint32_t p, uint32_t h) {
    const uint32_t* r_eid = L.r_eid();
    const uint16_t* nxt = L.nxt();
    uint16_t nx = nxt[h];
    float w0 = L.r_w[h];
    if (nx == kLinkNil) {
        L.head[p] = kListNil;
        L.slice[p] = L.slice[p] + w0;
        return false;
    }
    uint32_t ids[kShortList];
    float ws[kShortList];
    ids[0] = r_eid[h];
    ws[0] = w0;
#pragma unroll
    for (int q = 1; q < kShortList; ++q) {
        ids[q] = 0xFFFFFFFFu;
        ws[q] = 0.0f;
    }
    int n = 1;
    while (nx != kLinkNil && n < 4) {
        const uint32_t i = nx;
        nx = nxt[i];
        sorted_insert<4>(ids, ws, r_eid[i], L.r_w[i]);
        ++n;
    }
    while (nx != kLinkNil) {
        if (n == kShortList) {
            L.head[p] = kListLong;
            return true;
        }
        const uint32_t i = nx;
        nx = nxt[i];
        sorted_insert<kShortList>(ids, ws, r_eid[i], L.r_w[i]);
        ++n;
    }
    L.head[p] = kListNil;
    float acc = L.slice[p];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < n) acc = acc + ws[q];
    if (n > 4) {
#pragma unroll
        for (int q = 4; q < kShortList; ++q)
            if (q < n) acc = acc + ws[q];
    }
    L.slice[p] = acc;
    return false;
}

// Add the pass's m records (linked per cell from head[]) into the slice,
// each cell in ascending event id, and leave every head empty.  Lists of up
// to kShortList records are insertion-sorted in registers by their owning
// thread; longer ones (hub posts) are sorted together by (cell, event id)
// with a block bitonic sort.
template <bool kRec>
__device__ void bin_accumulate(BinApplyLds& L, uint32_t m, RecCtx rc) {
    const uint32_t t = threadIdx.x;
    const uint32_t* r_eid = L.r_eid();
    const uint16_t* nxt = L.nxt();
    int longs = 0;
    if (!kRec) {
        // A list of one record (most cells that receive anything) is added
        // by its cell's thread; cells with more are queued (one LDS atomic
        // per wave) and summed by the first threads of the block, so that
        // the insertion runs in a few waves instead of in every one.
        const uint32_t lane = t & 63u;
#pragma unroll
        for (uint32_t pr = 0; pr < kBinPosts / kApply; ++pr) {
            const uint32_t p = t + pr * kApply;
            const uint32_t h = L.head[p];
            bool multi = false;
            if (h != kListNil) {
                if (nxt[h] == kLinkNil) {
                    L.head[p] = kListNil;
                    L.slice[p] = L.slice[p] + L.r_w[h];
                } else {
                    multi = true;
                }
            }
            const uint64_t mb = __ballot(multi);
            if (mb) {
                const uint32_t lead = static_cast<uint32_t>(__builtin_ctzll(mb));
                uint32_t base = 0u;
                if (lane == lead) base = atomicAdd(&L.nm, static_cast<uint32_t>(__popcll(mb)));
                base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(base), static_cast<int>(lead)));
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mb >> 32),
                                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mb), 0u));
                if (multi) L.mq[base + rank] = static_cast<uint16_t>(p);
            }
        }
        __syncthreads();
        const uint32_t nm = L.nm;
        for (uint32_t i = t; i < nm; i += kApply) {
            const uint32_t p = L.mq[i];
            if (short_list_sum(L, p, L.head[p])) longs = 1;
        }
    }
#pragma unroll
    for (uint32_t pr = 0; kRec && pr < kBinPosts / kApply; ++pr) {
        const uint32_t p = t + pr * kApply;
        const uint32_t h = L.head[p];
        if (h == kListNil) continue;
        if (p < rc.c0 || p >= rc.c1) {  // another block's cell: its conductances are not touched here
            L.head[p] = kListNil;
            continue;
        }
        uint32_t ids[kShortList], cs[kShortList];
        float ws[kShortList];
#pragma unroll
        for (int q = 0; q < kShortList; ++q) {
            ids[q] = 0xFFFFFFFFu;  // above every event id
            ws[q] = 0.0f;
            cs[q] = 0u;
        }
        int n = 0;
        bool fits = true;
        for (uint32_t i = h;;) {
            if (n == kShortList) {
                fits = false;
                break;
            }
            const uint32_t e = r_eid[i];
            const float w = L.r_w[i];
            const uint16_t nx = nxt[i];
            const uint32_t cd = kRec ? static_cast<uint32_t>(L.r_post()[i]) >> kBinLog2 : 0u;
            // insert (e, w) into the ascending ids[]: every slot holding a larger
            // id takes its left neighbour's entry, or (e, w) if that one is smaller
#pragma unroll
            for (int q = kShortList - 1; q >= 0; --q) {
                const int qp = q > 0 ? q - 1 : 0;
                const bool prev_gt = q > 0 && ids[qp] > e;
                if (ids[q] > e) {
                    ids[q] = prev_gt ? ids[qp] : e;
                    ws[q] = prev_gt ? ws[qp] : w;
                    if (kRec) cs[q] = prev_gt ? cs[qp] : cd;
                }
            }
            ++n;
            if (nx == kLinkNil) break;
            i = nx;
        }
        if (!fits) {
            L.head[p] = kListLong;
            longs = 1;
            continue;
        }
        L.head[p] = kListNil;
        float acc = L.slice[p];
        if (kRec) {
            float g[kReceptors];
            uint32_t touched = 0u;
            const uint32_t post = cell_post(*rc.a, rc.k, p);
#pragma unroll
            for (int q = 0; q < kShortList; ++q) {
                if (q >= n) continue;
                if (cs[q] == 0u) acc = acc + ws[q];
                else rec_jump(*rc.a, post, cs[q], ws[q], g, touched);
            }
            rec_store(*rc.a, post, g, touched);
        } else {
#pragma unroll
            for (int q = 0; q < kShortList; ++q)
                if (q < n) acc = acc + ws[q];
        }
        L.slice[p] = acc;
    }
    const int any_long = __syncthreads_or(longs);
    if (!kRec && t == 0) L.nm = 0u;  // every thread has read it
    if (!any_long) return;
    // the records of long lists, compacted as keys (cell << 43 | eid << 11 |
    // record; receptor networks: cell << 46 | eid << 14 | code << 11 |
    // record, the keys overwrite the records' codes), then sorted
    constexpr uint32_t kPer = kBinCap / kApply;
    constexpr uint32_t kCellSh = kRec ? 46u : 43u, kEidSh = kRec ? 14u : 11u;
    uint64_t kk[kPer];
    uint32_t nk = 0;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = t + r * kApply;
        kk[r] = ~0ull;
        if (i < m) {
            const uint32_t cell = static_cast<uint32_t>(L.r_post()[i]) & (kBinPosts - 1u);
            if (L.head[cell] == kListLong) {
                kk[r] = (static_cast<uint64_t>(cell) << kCellSh) | (static_cast<uint64_t>(r_eid[i]) << kEidSh) | i;
                if (kRec) kk[r] |= static_cast<uint64_t>(L.r_post()[i] >> kBinLog2) << 11;
                ++nk;
            }
        }
    }
    if (t == 0) L.m = 0u;
    __syncthreads();  // every key read; u is free
    uint32_t pos = nk ? atomicAdd(&L.m, nk) : 0u;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r)
        if (kk[r] != ~0ull) L.u[pos++] = kk[r];
    __syncthreads();
    const uint32_t n_long = L.m;
    uint32_t np2 = 1;
    while (np2 < n_long) np2 <<= 1;
    for (uint32_t i = n_long + t; i < np2; i += kApply) L.u[i] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= np2; k <<= 1) {
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = t; i < np2; i += kApply) {
                const uint32_t l = i ^ jj;
                if (l > i) {
                    const uint64_t x = L.u[i], y = L.u[l];
                    if ((x > y) == ((i & k) == 0)) {
                        L.u[i] = y;
                        L.u[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    // one thread per long list sums its sorted run
    for (uint32_t i = t; i < n_long; i += kApply) {
        const uint64_t k0 = L.u[i];
        const uint32_t cell = static_cast<uint32_t>(k0 >> kCellSh);
        if (i > 0 && static_cast<uint32_t>(L.u[i - 1] >> kCellSh) == cell) continue;
        float acc = L.slice[cell];
        if (kRec) {
            float g[kReceptors];
            uint32_t touched = 0u;
            const uint32_t post = cell_post(*rc.a, rc.k, cell);
            for (uint32_t q = i; q < n_long; ++q) {
                const uint64_t kq = L.u[q];
                if (static_cast<uint32_t>(kq >> kCellSh) != cell) break;
                const uint32_t r = static_cast<uint32_t>(kq & (kBinCap - 1));
                const uint32_t cd = static_cast<uint32_t>(kq >> 11) & 7u;
                if (cd == 0u) acc = acc + L.r_w[r];
                else rec_jump(*rc.a, post, cd, L.r_w[r], g, touched);
            }
            rec_store(*rc.a, post, g, touched);
        } else {
            for (uint32_t q = i; q < n_long; ++q) {
                const uint64_t kq = L.u[q];
                if (static_cast<uint32_t>(kq >> kCellSh) != cell) break;
                acc = acc + L.r_w[kq & (kBinCap - 1)];
            }
        }
        L.slice[cell] = acc;
        L.head[cell] = kListNil;
    }
    __syncthreads();
}

// Gather one bin's records of one step into L.slice (each cell from zero in
// ascending event id): hrow = the bin's row of headers (nw writers, two per
// thread; the entries read are zeroed), ev = the records (16-byte records
// when kRec, else 8-byte ones, whose per-cell order key is writer << 21 | f).
// Every thread of the block calls it after zeroing L.slice and emptying
// L.head; it ends with a barrier.
// Thread t's header entries of a row of nw writers: writer t's {first
// record, count} in x, y when nw <= kApply (z = w = 0), else writers 2t
// and 2t + 1.  One writer per thread spreads the listing of the records
// over all the block's waves.
__device__ __forceinline__ uint4 bin_header_load(const uint2* hrow2, uint32_t nw) {
    static_assert(kMaxWriters == 2 * kApply, "at most two writers per thread");
    const uint32_t t = threadIdx.x;
    uint4 hw = make_uint4(0u, 0u, 0u, 0u);
    if (nw <= kApply) {
        if (t < nw) {
            const uint2 h = hrow2[t];
            hw.x = h.x;
            hw.y = h.y;
        }
    } else if (2u * t < nw) {
        hw = reinterpret_cast<const uint4*>(hrow2)[t];  // rows hold an even count
    }
    return hw;
}

template <bool kRec>
__device__ __forceinline__ void bin_gather(BinApplyLds& L, uint2* hrow2, uint4 hw, dptr<const uint2> ev, uint32_t nw,
                                           RecCtx rc, bool zero_row = true) {
    const dptr<const uint4> ev4 = reinterpret_cast<dptr<const uint4>>(ev);
    const uint32_t t = threadIdx.x;
    const bool one_w = nw <= kApply;  // thread t holds writer t only
    const uint32_t w0 = one_w ? t : 2 * t;
    if (t == 0) L.nm = 0u;  // the accumulate's queue (read after the scan's barriers)
    if ((hw.y | hw.w) && zero_row) {  // consumed: empty for the next use
        if (one_w) hrow2[t] = make_uint2(0u, 0u);
        else reinterpret_cast<uint4*>(hrow2)[t] = make_uint4(0u, 0u, 0u, 0u);
    }
    uint32_t excl;
    const uint32_t total = block_scan<kApply>(hw.y + hw.w, L.wave, excl);
    // One pass takes every record when they fit (the common case): then
    // each thread lists its groups' records -- buffer index and writer -- at
    // their positions in the pass, so that a record's thread finds it with
    // one LDS read instead of a search over the groups.
    const bool one = total <= kBinCap;
    uint32_t* raddr = L.seg_base;
    uint32_t* rgrp = L.seg_pref;
    if (one) {
        for (uint32_t i = 0; i < hw.y; ++i) {
            raddr[excl + i] = hw.x + i;
            rgrp[excl + i] = w0;
        }
        for (uint32_t i = 0; i < hw.w; ++i) {  // hw.w == 0 with one writer per thread
            raddr[excl + hw.y + i] = hw.z + i;
            rgrp[excl + hw.y + i] = 2 * t + 1;
        }
    } else if (one_w) {  // groups [0, nw] of the searches below, nw <= kApply
        L.seg_base[t] = hw.x;
        L.seg_pref[t] = excl;
        if (t == kApply - 1) L.seg_pref[kApply] = total;
    } else {
        L.seg_base[2 * t] = hw.x;
        L.seg_base[2 * t + 1] = hw.z;
        L.seg_pref[2 * t] = excl;
        L.seg_pref[2 * t + 1] = excl + hw.y;
        if (t == 0) L.seg_pref[kMaxWriters] = total;
    }
    __syncthreads();
    uint32_t ja = 0;
    while (ja < nw) {
        const uint32_t f0 = one ? 0u : L.seg_pref[ja];
        if (f0 == total) break;  // nothing left
        // the longest run of whole groups [ja, jb) that fits one pass
        uint32_t lo = ja, hi = nw;  // largest jb with seg_pref[jb] <= f0 + kBinCap
        while (!one && lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (L.seg_pref[mid] - f0 <= kBinCap) lo = mid; else hi = mid - 1;
        }
        const uint32_t jb = one ? nw : lo;
        if (jb == ja) {
            // Group ja alone exceeds a pass; its records are unordered, so
            // take them in windows of kBinCap consecutive event ids (or
            // indices f) -- at most kBinCap records each -- lowest first.
            const uint32_t seg = L.seg_pref[ja + 1] - f0, rb = L.seg_base[ja];
            if (t == 0) L.next = kListNil;
            __syncthreads();
            for (uint32_t i = t; i < seg; i += kApply)
                atomicMin(&L.next, kRec ? ev4[rb + i].x : ev[rb + i].y >> kBinLog2);
            __syncthreads();
            uint32_t w0 = L.next;
            for (;;) {
                __syncthreads();
                if (t == 0) {
                    L.m = 0u;
                    L.next = kListNil;
                }
                __syncthreads();
                for (uint32_t i = t; i < seg; i += kApply) {
                    uint32_t key, cell, code = 0u;
                    float w;
                    if (kRec) {
                        const uint4 r = ev4[rb + i];
                        key = r.x;
                        cell = r.z;
                        code = r.w;
                        w = __uint_as_float(r.y);
                    } else {
                        const uint2 r = ev[rb + i];
                        key = r.y >> kBinLog2;
                        cell = r.y & (kBinPosts - 1u);
                        w = __uint_as_float(r.x);
                    }
                    if (key < w0) continue;
                    if (key - w0 < kBinCap) bin_link(L, atomicAdd(&L.m, 1u), cell, key, w, code);
                    else atomicMin(&L.next, key);
                }
                __syncthreads();
                bin_accumulate<kRec>(L, L.m, rc);
                if (L.next == kListNil) break;
                w0 = L.next;
            }
            __syncthreads();
            ++ja;
            continue;
        }
        const uint32_t m = one ? total : L.seg_pref[jb] - f0;
        constexpr uint32_t kRecs = kBinCap / kApply;
        if (kRec) {
            uint4 r[kRecs];
#pragma unroll
            for (uint32_t u = 0; u < kRecs; ++u) {
                const uint32_t f = t + u * kApply;
                if (f < m) {
                    if (one) {
                        r[u] = ev4[raddr[f]];
                    } else {
                        const uint32_t sg = floor_index<kMaxWriters>(L.seg_pref, ja, jb, f0 + f);  // its group
                        r[u] = ev4[L.seg_base[sg] + (f0 + f - L.seg_pref[sg])];
                    }
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kRecs; ++u) {
                const uint32_t f = t + u * kApply;
                if (f < m) bin_link(L, f, r[u].z, r[u].x, __uint_as_float(r[u].y), r[u].w);
            }
        } else {
            static_assert(kRecs == 2 && kMaxWriters <= (1u << 16), "two writer indices share a register");
            uint2 r[kRecs];
            uint32_t sgp = 0u;  // the records' groups (= writers), 16 bits each
#pragma unroll
            for (uint32_t u = 0; u < kRecs; ++u) {
                const uint32_t f = t + u * kApply;
                if (f < m) {
                    uint32_t sg, at;
                    if (one) {
                        sg = rgrp[f];
                        at = raddr[f];
                    } else {
                        sg = floor_index<kMaxWriters>(L.seg_pref, ja, jb, f0 + f);
                        at = L.seg_base[sg] + (f0 + f - L.seg_pref[sg]);
                    }
                    sgp |= sg << (16u * u);
                    r[u] = ev[at];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kRecs; ++u) {
                const uint32_t f = t + u * kApply;
                if (f < m)
                    bin_link(L, f, r[u].y & (kBinPosts - 1u),
                             (((sgp >> (16u * u)) & 0xFFFFu) << kRecFBits) | (r[u].y >> kBinLog2),
                             __uint_as_float(r[u].x));
            }
        }
        __syncthreads();
        bin_accumulate<kRec>(L, m, rc);
        ja = jb;
    }
    __syncthreads();
}

// Two-kernel consumer, work item {network, bin}; {kEmpty, 0} pads the
// XCD-ordered list.  mode 0: the step's apply (slot from bin_slot, then the
// step counter moves on); mode 1: a megastep network's pending events (the
// previous step's buffer) into ring slot step % 32, unless already there.
template <bool kRec>
__global__ __launch_bounds__(kApply) void bin_apply_kernel(const NetArgs* __restrict__ nets,
                                                          const uint2* __restrict__ refs, uint32_t mode) {
    __shared__ BinApplyLds L;
    const uint2 ref = refs[blockIdx.x];
    if (ref.x == kEmpty) return;
    const NetArgs& a = nets[ref.x];
    const uint32_t k = ref.y, t = threadIdx.x;
    uint32_t slot;
    uint2* hdr = a.bin_hdr;
    dptr<const uint2> ev = ev_buf(a, 0u);
    if (mode == 0) {
        slot = *a.bin_slot;
        if (slot == kEmpty) return;
    } else {
        if (*a.bin_mat) return;
        const uint64_t step = *a.step;
        slot = static_cast<uint32_t>(step & (kSlots - 1));
        const uint32_t par = static_cast<uint32_t>((step + 1) & 1u);
        hdr += par * a.hdr_stride;
        ev = ev_buf(a, par);
    }
#pragma unroll
    for (uint32_t pr = 0; pr < kBinPosts / kApply; ++pr) {
        L.slice[t + pr * kApply] = 0.0f;
        L.head[t + pr * kApply] = kListNil;
    }
    uint2* hrow = hdr + static_cast<size_t>(k) * a.bin_rows;
    bin_gather<kRec>(L, hrow, bin_header_load(hrow, a.n_bblk), ev, a.n_bblk, RecCtx{(dptr<const NetArgs>)&a, k});
    // cell c of the bin -> its post in the slot
    float* out = a.ring + static_cast<size_t>(slot) * a.N;
#pragma unroll
    for (uint32_t pr = 0; pr < kBinPosts / kApply; ++pr) {
        const uint32_t c = t + pr * kApply, tile = k * kBinTiles + c / kTile;
        if (tile < a.n_tiles) {
            const PopDev& P = a.pops[a.tile_pop[tile]];
            const uint32_t li = (tile - P.tile0) * kTile + c % kTile;
            if (li < P.n) out[P.off + li] = L.slice[c];
        }
    }
    if (mode == 0 && k == 0 && t == 0) *a.step = *a.step + 1;  // no other block of this kernel reads it
}

// Megastep: one kernel per step for a binned network whose populations share
// MODEL and have no pair-based plasticity.  Work item {network, bin k}: the
// block owns tiles [k*kBinTiles, (k+1)*kBinTiles), thread t neurons t and
// t + kApply of them (cells of bin k).  It gathers the previous step's events
// for its cells (or reads ring slot step % 32 when a host call has put them
// there), advances the neurons exactly as neuron_kernel does, and writes its
// spikes' events as writer k for the next step.
// LIF and Izhikevich fit two blocks per CU (eight waves per SIMD); HH and
// GPEA keep one, their updates need the registers.
template <int MODEL>
constexpr int kMegaWaves = MODEL <= 1 || MODEL == 4 ? 8 : 4;

template <int MODEL>
__global__ __launch_bounds__(kApply, kMegaWaves<MODEL>) void megastep_kernel(const NetArgs* __restrict__ nets,
                                                         const uint2* __restrict__ items) {
    __shared__ BinApplyLds L;
    __shared__ BinMap M;
    __shared__ uint32_t s_cnt[kBinTiles][kWavesPerTile], s_deg[kBinTiles][kWavesPerTile];
    __shared__ uint32_t s_drv[kBinTiles][kWavesPerTile];
    __shared__ uint32_t s_tp[kBinTiles + 1][3];  // per tile: first spiking row, first edge; then block totals
    const uint2 it = items[blockIdx.x];
    if (it.x == kEmpty) return;  // pad of the XCD-ordered list
    const NetArgs& a = nets[it.x];
    const uint32_t k = it.y, t = threadIdx.x, lane = t & 63u, wq = (t / 64u) % kWavesPerTile, tl = t % kTile;
    // the step counter and the materialised flag in one batch of scalar loads
    const uint64_t step = *a.step;
    const uint32_t mat = *a.bin_mat;
    if (step >= a.step_limit) return;
    const uint32_t wpar = static_cast<uint32_t>(step & 1u);

    // Thread t's neurons are cell t and t + 1024 of the bin: tiles w/4 and
    // w/4 + 4 of wave w.  Tiles and populations are wave-uniform, so the
    // population table and descriptors are read through the constant address
    // space (scalar loads into scalar registers, which neither wait behind
    // nor hold up the vector loads) and looked up again wherever needed.
    const uint32_t wv = __builtin_amdgcn_readfirstlane(t >> 6);
    auto tile_of = [&](uint32_t q) -> uint32_t { return k * kBinTiles + wv / kWavesPerTile + q * (kApply / kTile); };
    auto pop_of = [&](uint32_t q) -> uint32_t {
        const uint32_t tile = tile_of(q);
        return tile < a.n_tiles && a.n_pops > 1 ? ((cptr<uint32_t>)a.tile_pop)[tile] : 0u;
    };
    auto pop_dev = [&](uint32_t p) -> PopDev { return ((cptr<PopDev>)a.pops)[p]; };
    float cur[2] = {0.0f, 0.0f};
    PopDev PV[2];
    uint32_t li[2], tile[2], pidx[2];
    bool on[2];
#pragma unroll
    for (uint32_t q = 0; q < 2; ++q) {
        tile[q] = tile_of(q);
        pidx[q] = pop_of(q);
        PV[q] = pop_dev(pidx[q]);
        li[q] = (tile[q] - PV[q].tile0) * kTile + tl;
        on[q] = tile[q] < a.n_tiles && li[q] < PV[q].n;
    }
    // Gaussian input, drawn before the gather (it depends on (seed, neuron,
    // step) only): measured, the draw costs less here than after the gather,
    // where the state loads wait on it.  The thread's two neurons are li and li + 1024 of one
    // population whenever its bin lies inside that population, and then
    // share one Box-Muller draw (the pair of normal_pair).
    float nz[2] = {0.0f, 0.0f};
    if (on[0] && on[1] && pidx[0] == pidx[1] && (PV[0].flags & kPopNoise) && (li[0] & 1024u) == 0u &&
        li[1] == li[0] + 1024u) {
        const PopDev& P = PV[0];
        float zc, zs;
        normal_pair(P.seed, li[0], step, zc, zs);
        const bool uni = P.flags & kPopNoiseScalar;
        nz[0] = (uni ? P.sigma_uni : P.sigma[li[0]]) * zc;
        nz[1] = (uni ? P.sigma_uni : P.sigma[li[1]]) * zs;
    } else {
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            const PopDev& P = PV[q];
            if (on[q] && (P.flags & kPopNoise)) {
                const float sg = (P.flags & kPopNoiseScalar) ? P.sigma_uni : P.sigma[li[q]];
                nz[q] = sg * normal_draw(P.seed, li[q], step);
            }
        }
    }
    if (mat) {
        // a host call (current injection, ring download) has applied them
        float* ring_slot = a.ring + static_cast<size_t>(step & (kSlots - 1)) * a.N;
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            if (!on[q]) continue;
            const uint32_t i = PV[q].off + li[q];
            cur[q] = ring_slot[i];
            ring_slot[i] = 0.0f;
        }
        __syncthreads();
    } else {
#pragma unroll
        for (uint32_t pr = 0; pr < kBinPosts / kApply; ++pr) {
            L.slice[t + pr * kApply] = 0.0f;
            L.head[t + pr * kApply] = kListNil;
        }
        const uint32_t rpar = wpar ^ 1u;
        uint2* hrow = a.bin_hdr + rpar * a.hdr_stride + static_cast<size_t>(k) * a.bin_rows;
        // every writer fills its whole column each step: the row is left as read
        bin_gather<false>(L, hrow, bin_header_load(hrow, a.n_bblk), ev_buf(a, rpar), a.n_bblk, RecCtx{nullptr, 0u},
                          false);
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) cur[q] = L.slice[t + q * kApply];
    }
    uint32_t row0[2], deg[2];
    Neuron<MODEL> nrn[2];
#pragma unroll
    for (uint32_t q = 0; q < 2; ++q) {
        tile[q] = tile_of(q);
        pidx[q] = pop_of(q);
        PV[q] = pop_dev(pidx[q]);
        li[q] = (tile[q] - PV[q].tile0) * kTile + tl;
        on[q] = tile[q] < a.n_tiles && li[q] < PV[q].n;
        row0[q] = deg[q] = 0u;
        if (on[q]) nrn[q].load(PV[q], li[q]);
    }
    // the bin map (read by the writer phase) loads behind the state, not in
    // front of wave 0's header load; valid after the barrier below
    bin_map_load(a, M);

    // advance every neuron; per tile ballot the spikes and scan the spiking
    // rows' lengths within each wave
    uint64_t m[2];
    uint32_t excl[2];
#pragma unroll
    for (uint32_t q = 0; q < 2; ++q) {
        const PopDev& P = PV[q];
        const uint32_t ti = t / kTile + q * (kApply / kTile);
        int spike = 0, drive_now = 0;
        if (on[q]) {
            const uint32_t i = li[q];
            float I = cur[q];
            // A per-neuron Poisson source with one edge of delay <= 1 adds
            // its weight into the slot after every neuron push of the previous
            // step (P/src/api/net/network/step.rs:166-185): the last add.
            if ((P.flags & kPopDrive) && P.drive_p > 0.0f) {
                const uint32_t sd = P.drive_seed0 + i * 2654435761u;
                if (step > 0 && poisson_roll(sd, 0u, step - 1, P.drive_p)) I = I + P.drive_w;
                drive_now = poisson_roll(sd, 0u, step, P.drive_p) ? 1 : 0;
            }
            if (P.flags & kPopNoise) I = I + nz[q];
            spike = nrn[q].step(P, i, I, P.puni[0]);
            nrn[q].store(P, i);
            if (spike) {  // only spiking rows need their CSR extent
                row_extent(a, P.off + i, row0[q], deg[q]);
            }
        }
        m[q] = __ballot(spike);
        const uint32_t drive_events = static_cast<uint32_t>(__popcll(__ballot(drive_now)));
        const uint32_t d = spike ? deg[q] : 0u;
        uint32_t inc = d;
        if (m[q]) {
            inc = wave_incl_scan(inc);
        }
        excl[q] = inc - d;
        if (lane == 63) s_deg[ti][wq] = inc;
        if (lane == 0) {
            s_cnt[ti][wq] = static_cast<uint32_t>(__popcll(m[q]));
            s_drv[ti][wq] = drive_events;
            if (tile[q] < a.n_tiles) {
                const uint32_t word = P.word0 + (tile[q] - P.tile0) * kWavesPerTile + wq;
                a.bits[word] = m[q];
                if (a.raster) {
                    const uint64_t rs = step - *a.raster_base;
                    if (rs < a.raster_cap) a.raster[rs * a.n_words + word] = m[q];
                }
            }
        }
    }
    __syncthreads();

    // ordered compaction: the block's spiking rows in (tile, wave, lane)
    // order, i.e. ascending neuron, with their edge prefix
    if (t < kBinTiles) {  // lane t of wave 0: tile t's totals, then their prefix
        uint32_t sp = 0, ed = 0, dv = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWavesPerTile; ++w) {
            sp += s_cnt[t][w];
            ed += s_deg[t][w];
            dv += s_drv[t][w];
        }
        // inclusive scan over lanes 0..7 (one DPP row; a lane shifted out of
        // the row reads 0)
        static_assert(kBinTiles == 8, "one DPP row of eight lanes");
        int isp = static_cast<int>(sp), ied = static_cast<int>(ed), idv = static_cast<int>(dv);
        isp += __builtin_amdgcn_update_dpp(0, isp, 0x111, 0xf, 0xf, false);
        ied += __builtin_amdgcn_update_dpp(0, ied, 0x111, 0xf, 0xf, false);
        idv += __builtin_amdgcn_update_dpp(0, idv, 0x111, 0xf, 0xf, false);
        isp += __builtin_amdgcn_update_dpp(0, isp, 0x112, 0xf, 0xf, false);
        ied += __builtin_amdgcn_update_dpp(0, ied, 0x112, 0xf, 0xf, false);
        idv += __builtin_amdgcn_update_dpp(0, idv, 0x112, 0xf, 0xf, false);
        isp += __builtin_amdgcn_update_dpp(0, isp, 0x114, 0xf, 0xf, false);
        ied += __builtin_amdgcn_update_dpp(0, ied, 0x114, 0xf, 0xf, false);
        idv += __builtin_amdgcn_update_dpp(0, idv, 0x114, 0xf, 0xf, false);
        s_tp[t][0] = isp - sp;
        s_tp[t][1] = ied - ed;
        if (t == kBinTiles - 1) {
            s_tp[kBinTiles][0] = isp;
            s_tp[kBinTiles][1] = ied;
            s_tp[kBinTiles][2] = idv;
        }
    }
    __syncthreads();
    const uint32_t blk_sp = s_tp[kBinTiles][0], blk_ed = s_tp[kBinTiles][1], drv = s_tp[kBinTiles][2];
    uint32_t* s_rs = reinterpret_cast<uint32_t*>(L.u);  // the gather is done with u
    uint32_t* s_pref = s_rs + kBinCap;
#pragma unroll
    for (uint32_t q = 0; q < 2; ++q) {
        const uint32_t ti = t / kTile + q * (kApply / kTile);
        const uint32_t sp0 = s_tp[ti][0], ed0 = s_tp[ti][1];  // rows and edges of the tiles before ti
        if ((m[q] >> lane) & 1ull) {
            uint32_t base = 0, dbase = 0;
            for (uint32_t w = 0; w < wq; ++w) {
                base += s_cnt[ti][w];
                dbase += s_deg[ti][w];
            }
            const uint32_t rank = base + static_cast<uint32_t>(__popcll(m[q] & ((1ull << lane) - 1ull)));
            a.active[static_cast<size_t>(tile[q]) * kTile + rank] = PV[q].off + li[q];
            s_rs[sp0 + rank] = row0[q];
            s_pref[sp0 + rank] = ed0 + dbase + excl[q];
        }
        if (tl == 0 && tile[q] < a.n_tiles) a.tile_count[tile[q]] = s_tp[ti + 1][0] - sp0;
    }
    if (t == 0) {
        const uint32_t t0 = k * kBinTiles;
        if (blk_sp) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_spikes + t0), blk_sp);
        // each firing drive source is one synaptic event, counted at its roll
        if (drv + blk_ed) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_events + t0), drv + blk_ed);
    }
    uint32_t* s_bin = L.seg_base;  // n_bins <= kMaxWriters; the gather is done with it
    for (uint32_t b = t; b < a.n_bins; b += kApply) s_bin[b] = 0u;
    __syncthreads();

    // writer k of this step's events
    const dptr<uint2> ev = ev_buf(a, wpar);
    uint2* hdr = a.bin_hdr + wpar * a.hdr_stride;
    const bool cached = blk_ed <= kEvCache * kApply;
    EvCache ec;
    if (cached) {
        ev_cache_load<kApply>(a, M, false, blk_sp, blk_ed, s_rs, s_pref, ec);
        ev_cache_count<kApply>(ec, blk_ed, s_bin);
    } else {
        bin_events<false, kApply>(a, M, false, blk_sp, blk_ed, s_rs, s_pref, s_bin, ev, step);
    }
    bin_offsets<kApply>(a, s_bin, L.wave, k, a.tile_ebase[k * kBinTiles], blk_ed, hdr);
    if (cached) ev_cache_write<kApply, false, true>(a, false, ec, blk_ed, s_bin, ev, step);
    else bin_events<true, kApply, false, false, true>(a, M, false, blk_sp, blk_ed, s_rs, s_pref, s_bin, ev, step);
    __syncthreads();
    if (t == 0) mega_done(a, step);
}

// Split form, for binned networks the megastep cannot take (pair-based
// plasticity, receptor synapses): per step front_kernel, then bin_kernel's
// writers (with the out-edge STDP) and in-edge LTP blocks, the last of which
// moves the step counter on -- two launches instead of neuron_kernel,
// bin_kernel and bin_apply_kernel.  Work item {network, bin * kFrontParts +
// part}: the block gathers the previous step's records of the whole bin (from
// the parity buffers, leaving the header row as read, since the bin's other
// part reads it too) and applies those of its own cells, part * kApply + t for
// thread t: the current into the cell's input and, in receptor networks, each
// conductance jump into g in event order (bin_apply_kernel's arithmetic).
// Then it advances its neurons as neuron_kernel does (receptor currents, STDP
// traces) and leaves each tile's ascending spiking ids and count for the
// writers.  One neuron per thread keeps twice as many blocks at work as a
// block per bin would.
static_assert(kFrontParts * kApply == kBinPosts, "a front block's cells are one per thread");
constexpr uint32_t kFrontTiles = kApply / kTile;  // tiles of one part

template <int MODEL, bool kRec>
constexpr int kFrontWaves = !kRec && (MODEL <= 1 || MODEL == 4) ? 8 : 4;

template <int MODEL, bool kRec>
__global__ __launch_bounds__(kApply, (kFrontWaves<MODEL, kRec>)) void front_kernel(const NetArgs* __restrict__ nets,
                                                                          const uint2* __restrict__ items) {
    __shared__ BinApplyLds L;
    __shared__ uint32_t s_cnt[kFrontTiles][kWavesPerTile], s_drv[kFrontTiles][kWavesPerTile];
    const uint2 it = items[blockIdx.x];
    if (it.x == kEmpty) return;  // pad of the XCD-ordered list
    const NetArgs& a = nets[it.x];
    const uint32_t k = it.y / kFrontParts, part = it.y % kFrontParts;
    const uint32_t t = threadIdx.x, lane = t & 63u, wq = (t / 64u) % kWavesPerTile, tl = t % kTile;
    const uint64_t step = *a.step;
    const uint32_t mat = *a.bin_mat;
    if (step >= a.step_limit) return;

    // thread t: cell part * kApply + t of bin k, in the wave-uniform tile
    // tile_t; population descriptors through scalar loads, as in the megastep
    const uint32_t wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t ti = wv / kWavesPerTile;
    const uint32_t tile = k * kBinTiles + part * kFrontTiles + ti;
    const uint32_t pidx = tile < a.n_tiles && a.n_pops > 1 ? ((cptr<uint32_t>)a.tile_pop)[tile] : 0u;
    const PopDev P = ((cptr<PopDev>)a.pops)[pidx];
    const uint32_t li = (tile - P.tile0) * kTile + tl;
    const bool on = tile < a.n_tiles && li < P.n;
    const uint32_t gi = P.off + li;
    // The neuron's state and traces: with a block per CU (four waves per
    // SIMD) there are registers to load them before the gather, which
    // touches neither; with two, after it.
    constexpr bool kEarly = kFrontWaves<MODEL, kRec> == 4;
    Neuron<MODEL> nrn;
    float trp = 0.0f, trq = 0.0f;
    auto load_state = [&]() {
        if (on) {
            nrn.load(P, li);
            if (a.stdp) {
                trp = a.tr_pre[gi];
                trq = a.tr_post[gi];
            }
        }
    };
    if (kEarly) load_state();
    float nz = 0.0f;
    if (on && (P.flags & kPopNoise)) {
        const float sg = (P.flags & kPopNoiseScalar) ? P.sigma_uni : P.sigma[li];
        nz = sg * normal_draw(P.seed, li, step);
    }
    float cur = 0.0f;
    if (mat) {
        // a host call (current injection, ring download) has applied them
        float* ring_slot = a.ring + static_cast<size_t>(step & (kSlots - 1)) * a.N;
        if (on) {
            cur = ring_slot[gi];
            ring_slot[gi] = 0.0f;
        }
    } else {
#pragma unroll
        for (uint32_t pr = 0; pr < kBinPosts / kApply; ++pr) {
            L.slice[t + pr * kApply] = 0.0f;
            L.head[t + pr * kApply] = kListNil;
        }
        const uint32_t rpar = static_cast<uint32_t>(step & 1u) ^ 1u;
        uint2* hrow = a.bin_hdr + rpar * a.hdr_stride + static_cast<size_t>(k) * a.bin_rows;
        bin_gather<kRec>(L, hrow, bin_header_load(hrow, a.n_bblk), ev_buf(a, rpar), a.n_bblk,
                         RecCtx{(dptr<const NetArgs>)&a, k, part * kApply, (part + 1u) * kApply}, false);
        cur = L.slice[part * kApply + t];
    }

    // the update of neuron_kernel: conductances (after this step's jumps),
    // drive, noise, receptor currents, STDP traces
    if (!kEarly) load_state();
    RecIn rin;
    int spike = 0, drive_now = 0;
    if (on) {
        if (kRec) rec_load(a, gi, rin);
        float I = cur;
        if ((P.flags & kPopDrive) && P.drive_p > 0.0f) {
            const uint32_t sd = P.drive_seed0 + li * 2654435761u;
            if (step > 0 && poisson_roll(sd, 0u, step - 1, P.drive_p)) I = I + P.drive_w;
            drive_now = poisson_roll(sd, 0u, step, P.drive_p) ? 1 : 0;
        }
        if (P.flags & kPopNoise) I = I + nz;
        if (kRec) I = receptor_input(a, P, gi, nrn.vm(P, li), rin, I);
        spike = nrn.step(P, li, I, P.puni[0]);
        nrn.store(P, li);
        if (a.stdp) {
            const float add = spike ? 1.0f : 0.0f;
            a.tr_pre[gi] = fmaf(trp, a.stdp_dpre, add);
            a.tr_post[gi] = fmaf(trq, a.stdp_dpost, add);
        }
    }
    const uint64_t m = __ballot(spike);
    const uint32_t drive_events = static_cast<uint32_t>(__popcll(__ballot(drive_now)));
    if (lane == 0) {
        s_cnt[ti][wq] = static_cast<uint32_t>(__popcll(m));
        s_drv[ti][wq] = drive_events;
        if (tile < a.n_tiles) {
            const uint32_t word = P.word0 + (tile - P.tile0) * kWavesPerTile + wq;
            a.bits[word] = m;
            if (a.raster) {
                const uint64_t rs = step - *a.raster_base;
                if (rs < a.raster_cap) a.raster[rs * a.n_words + word] = m;
            }
        }
    }
    __syncthreads();

    // each tile's spiking ids in ascending order, and its count
    if ((m >> lane) & 1ull) {
        uint32_t base = 0;
        for (uint32_t w = 0; w < wq; ++w) base += s_cnt[ti][w];
        const uint32_t rank = base + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
        a.active[static_cast<size_t>(tile) * kTile + rank] = gi;
    }
    if (tl == 0 && tile < a.n_tiles) {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWavesPerTile; ++w) c += s_cnt[ti][w];
        a.tile_count[tile] = c;
    }
    if (t == 0) {
        uint32_t sp = 0, dv = 0;
#pragma unroll
        for (uint32_t x = 0; x < kFrontTiles; ++x)
#pragma unroll
            for (uint32_t w = 0; w < kWavesPerTile; ++w) {
                sp += s_cnt[x][w];
                dv += s_drv[x][w];
            }
        const uint32_t t0 = k * kBinTiles + part * kFrontTiles;
        if (sp) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_spikes + t0), sp);
        // each firing drive source is one synaptic event, counted at its roll
        if (dv) atomicAdd(reinterpret_cast<unsigned long long*>(a.tile_events + t0), dv);
    }
}

__device__ __forceinline__ float event_weight(const NetArgs& a, uint32_t eid) {
    const float* p = a.ev_w ? a.ev_w + eid : eid < a.n_syn ? a.w + eid : a.src_w + (eid - a.n_syn);
    return *p;
}

__device__ __forceinline__ void chain_insert(uint32_t* my, float* myw, int n, uint32_t e, float w) {
    int j = n;
    while (j > 0 && my[j - 1] > e) {
        my[j] = my[j - 1];
        myw[j] = myw[j - 1];
        --j;
    }
    my[j] = e;
    myw[j] = w;
}

// Re-sum one key that received more than one event this step.  `c` is one
// of the key's late events; only the key's last event (the current head)
// does the work.  The chain runs last -> ... -> first event; the first
// event's record holds the ring value before this step's adds.  Each hop
// loads an event's record and weight together.
__device__ __forceinline__ void fix_key(const NetArgs& a, uint32_t c, uint32_t key, uint32_t* my, float* myw) {
    const uint32_t last = reinterpret_cast<const uint32_t*>(a.head + key)[0];  // event id half
    uint2 r = a.rec[c];
    float we = event_weight(a, c);
    if (last != c) return;
    int len = 0;
    bool overflow = false;
    for (uint32_t e = c;;) {
        if (len < kChain) chain_insert(my, myw, len++, e, we);
        else overflow = true;
        if (r.x == kEmpty) break;
        e = r.x;
        r = a.rec[e];
        we = event_weight(a, e);
    }
    float acc = __uint_as_float(r.y);
    if (!overflow) {
        for (int k = 0; k < len; ++k) acc = acc + myw[k];
    } else {
        // Long chain: passes over the chain, each taking the kChain smallest
        // ids above the last one added.
        int64_t done_to = -1;
        for (;;) {
            int n = 0;
            for (uint32_t e = c;;) {
                if (static_cast<int64_t>(e) > done_to && (n < kChain || e < my[kChain - 1])) {
                    const float w = event_weight(a, e);
                    if (n < kChain) chain_insert(my, myw, n++, e, w);
                    else chain_insert(my, myw, kChain - 1, e, w);
                }
                const uint32_t nx = a.rec[e].x;
                if (nx == kEmpty) break;
                e = nx;
            }
            if (n == 0) break;
            for (int k = 0; k < n; ++k) acc = acc + myw[k];
            done_to = my[n - 1];
        }
    }
    a.ring[key] = acc;
}

// Fix-up of the step's collisions: block j of a network covers scatter tiles
// [j*kFixTiles, (j+1)*kFixTiles); block 0 advances the step counter.  `refs` null: single network, block j = blockIdx.x.
__global__ __launch_bounds__(kTile) void fixup_kernel(const NetArgs* __restrict__ nets,
                                                     const uint2* __restrict__ refs) {
    __shared__ uint32_t s_base[kFixTiles], s_pref[kFixTiles], s_wave[kWavesPerTile];
    __shared__ uint32_t s_chain[kTile * kChain];
    __shared__ float s_chain_w[kTile * kChain];
    const uint2 ref = refs ? refs[blockIdx.x] : make_uint2(0u, blockIdx.x);
    const NetArgs& a = nets[ref.x];
    const uint32_t t = threadIdx.x;
    const uint32_t t0 = ref.y * kFixTiles;
    const uint32_t nt = min(kFixTiles, a.n_blocks - t0);
    uint32_t cnt = 0, base = 0;
    if (t < nt) {
        cnt = a.tile_coll[t0 + t];
        base = a.tile_ebase[t0 + t];
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(cnt, s_wave, excl);
    if (t < nt) {
        s_base[t] = base;
        s_pref[t] = excl;
        if (cnt) a.tile_coll[t0 + t] = 0u;  // consumed: a network past its horizon re-reads zeros
    }
    __syncthreads();
    uint32_t* my = s_chain + t * kChain;
    float* myw = s_chain_w + t * kChain;
    for (uint32_t f = t; f < total; f += kTile) {
        uint32_t lo = 0, hi = nt;  // largest k with s_pref[k] <= f
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[mid] <= f) lo = mid; else hi = mid;
        }
        const uint2 ce = a.coll[s_base[lo] + (f - s_pref[lo])];
        fix_key(a, ce.x, ce.y, my, myw);
    }
    // No other block of this kernel reads the step counter, so block 0 of
    // the network advances it without waiting for the rest; the next step's
    // kernels see it, and the ring writes, through the kernel boundary.
    if (ref.y == 0 && t == 0) {
        const uint64_t step = *a.step;
        if (step < a.step_limit) *a.step = step + 1;
    }
}

// Heads tagged with a step before the current one are idle.  Rewriting them
// all with the tag of step-1 keeps that true for the next 2^32-1 steps; the
// runtime does it at reset, after restoring a snapshot and every 2^31 steps.
__global__ void head_clear_kernel(unsigned long long* head, const uint64_t* step, size_t n) {
    const unsigned long long v =
        (static_cast<unsigned long long>(static_cast<uint32_t>(*step - 1u)) << 32) | kEmpty;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        head[i] = v;
}

// Pair-based STDP of one step for the spiking neurons of tile b, in one pass
// that equals the reference's LTP pass followed by its LTD pass
// (P/src/api/learn/brain.rs:204-286): in-edges of every spiking post gain
// A+ * pre-trace, out-edges of every spiking pre lose A- * post-trace, each
// update clipped to [w_min, w_max] with the FP64 master weight.  A synapse
// whose pre and post both spiked takes both, LTP first; it is handled only
// by its pre's out-edge walk (the in-edge walk skips it), so no synapse is
// touched by two threads and the two walks need no ordering.
__global__ __launch_bounds__(kTile) void stdp_kernel(const NetArgs* __restrict__ nets,
                                                   const uint2* __restrict__ tiles) {
    __shared__ uint32_t s_rs[2][kTile], s_pref[2][kTile], s_id[kTile], s_wave[kWavesPerTile];
    const uint2 ref = tiles[blockIdx.x];
    const NetArgs& a = nets[ref.x];
    if (*a.step >= a.step_limit) return;
    const uint32_t b = ref.y, t = threadIdx.x;
    const uint32_t cnt = a.tile_count[b];
    if (cnt == 0) return;
    uint32_t id = 0, ri = 0, di = 0, ro = 0, dout = 0;
    if (t < cnt) {
        id = a.active[static_cast<size_t>(b) * kTile + t];
        ri = a.in_ptr[id];
        di = a.in_ptr[id + 1] - ri;
        ro = a.row_ptr[id];
        dout = a.row_ptr[id + 1] - ro;
    }
    uint32_t ei, eo;
    const uint32_t tin = block_exclusive_scan(di, s_wave, ei);
    __syncthreads();
    const uint32_t tout = block_exclusive_scan(dout, s_wave, eo);
    if (t < cnt) {
        s_rs[0][t] = ri;
        s_pref[0][t] = ei;
        s_rs[1][t] = ro;
        s_pref[1][t] = eo;
        s_id[t] = id;
    }
    __syncthreads();
    const float ap = a.a_plus, am = a.a_minus;
    for (uint32_t f = t; f < tin + tout; f += kTile) {
        const bool out = f >= tin;
        const uint32_t g = out ? f - tin : f, side = out ? 1u : 0u;
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[side][mid] <= g) lo = mid; else hi = mid;
        }
        if (!out) {  // LTP on an in-edge of spiking post s_id[lo]
            const uint32_t e = a.in_edge[s_rs[0][lo] + (g - s_pref[0][lo])];
            if (a.plastic && !a.plastic[e]) continue;
            const uint32_t pre = a.edge_pre[e];
            if (spiked_now(a, pre)) continue;  // its pre's out-edge walk takes it
            stdp_apply(a, e, ap * a.tr_pre[pre], true);
        } else {  // LTD (and LTP first if the post spiked too) on an out-edge of spiking pre s_id[lo]
            const uint32_t e = s_rs[1][lo] + (g - s_pref[1][lo]);
            if (a.plastic && !a.plastic[e]) continue;
            const uint32_t post = a.col[e];
            if (spiked_now(a, post)) stdp_apply(a, e, ap * a.tr_pre[s_id[lo]], true);
            stdp_apply(a, e, am * a.tr_post[post], false);
        }
    }
}

// Initial short-term plasticity state: R = 1, U = U0 of the edge's profile
// (0 for static edges), last = 0 (Network::build, build.rs:70-84).
__global__ void tm_init_kernel(const uint8_t* type, const float4* prof, float* r, float* u, uint32_t* last,
                               size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint32_t t = type[i];
        r[i] = 1.0f;
        u[i] = t ? prof[t - 1].x : 0.0f;
        last[i] = 0u;
    }
}

// Dense every-synapse STDP of CompiledNetwork::stdp_step
// (PC/src/network.rs:53-83, driven by :495-718), single network.  Trace
// update from the previous step's spike words: trace = trace*decay + spike.
__global__ __launch_bounds__(kTile) void dense_trace_kernel(const NetArgs* __restrict__ nets, float* trace,
                                                          float decay) {
    const NetArgs& a = nets[0];
    const uint32_t tile = blockIdx.x;
    const PopDev& P = a.pops[a.tile_pop[tile]];
    const uint32_t lt = tile - P.tile0, li = lt * kTile + threadIdx.x;
    if (li >= P.n) return;
    const uint64_t word = a.bits[P.word0 + lt * kWavesPerTile + (threadIdx.x >> 6)];
    const float s = ((word >> (threadIdx.x & 63u)) & 1ull) ? 1.0f : 0.0f;
    const uint32_t gi = P.off + li;
    trace[gi] = trace[gi] * decay + s;
}

// Every synapse of a tile of 256 pre neurons (rows flattened over the
// block): w = min(w_max, max(w_min, (w + lr_ltp*tr[pre]) - lr_ltd*tr[post])),
// between the population update and the push.
__global__ __launch_bounds__(kTile) void dense_stdp_kernel(const NetArgs* __restrict__ nets,
                                                         const float* __restrict__ trace, float lr_ltp,
                                                         float lr_ltd, float w_min, float w_max) {
    __shared__ uint32_t s_rs[kTile], s_pref[kTile], s_wave[kWavesPerTile];
    const NetArgs& a = nets[0];
    const uint32_t t = threadIdx.x, pre0 = blockIdx.x * kTile;
    const uint32_t cnt = min(kTile, a.N - pre0);
    uint32_t rs = 0, dg = 0;
    if (t < cnt) {
        rs = a.row_ptr[pre0 + t];
        dg = a.row_ptr[pre0 + t + 1] - rs;
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(dg, s_wave, excl);
    s_rs[t] = rs;
    s_pref[t] = excl;
    __syncthreads();
    const dptr<float> w = const_cast<dptr<float>>(a.w);
    for (uint32_t f = t; f < total; f += kTile) {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[mid] <= f) lo = mid; else hi = mid;
        }
        const uint32_t e = s_rs[lo] + (f - s_pref[lo]);
        const float nw = (w[e] + lr_ltp * trace[pre0 + lo]) - lr_ltd * trace[a.col[e]];
        const float cl = w_min < nw ? nw : w_min;  // max(w_min, nw), argument order kept
        w[e] = cl < w_max ? cl : w_max;            // min(w_max, cl)
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

// inject_voltage: v += value on state column 0 (the v_jump kernel of
// PW/shaders/csr_scatter.wgsl:117-124), between two steps.
__global__ void vjump_kernel(float* v, const uint32_t* idx, const float* val, uint32_t count, int serial) {
    if (serial) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (uint32_t i = 0; i < count; ++i) v[idx ? idx[i] : i] += val[i];
        return;
    }
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[idx ? idx[i] : i] += val[i];
}

// Electrical coupling, P/src/api/net/network/step.rs:36-104: with V at the
// start of the step (no neuron of this step is advanced before this kernel
// ends), every neuron adds its contributions to its input slot in the
// reference's order -- gap junctions g (V_other - V_self) at the first end
// and - g (V_self - V_other) at the second, graded synapses g (V_pre -
// V_post), then sigmoid graded synapses (SigmoidGradedRuntime::step,
// P/src/api/net/network/sources.rs:117-129) whose activation s is advanced
// here by its post neuron's thread, the only writer.  Gaussian noise (the
// reference's inject_current before the step) enters the slot first.  Work
// item {network, tile}.
__global__ __launch_bounds__(kTile) void coupling_kernel(const NetArgs* __restrict__ nets,
                                                         const uint2* __restrict__ items) {
    const uint2 it = items[blockIdx.x];
    if (it.x == kEmpty) return;
    const NetArgs& a = nets[it.x];
    const uint64_t step = *a.step;
    if (step >= a.step_limit) return;
    const uint32_t tile = it.y;
    const PopDev& P = a.pops[a.tile_pop[tile]];
    const uint32_t li = (tile - P.tile0) * kTile + threadIdx.x;
    if (li >= P.n) return;
    const uint32_t gi = P.off + li;
    const uint32_t e0 = a.cp_ptr[gi], e1 = a.cp_ptr[gi + 1];
    const bool noise = P.flags & kPopNoise;
    if (e0 == e1 && !noise) return;
    float* cell = a.ring + static_cast<size_t>(step & (kSlots - 1)) * a.N + gi;
    float I = *cell;
    if (noise) {
        const float sg = (P.flags & kPopNoiseScalar) ? P.sigma_uni : P.sigma[li];
        I = I + sg * normal_draw(P.seed, li, step);
    }
    const float vme = P.st[li];
    for (uint32_t e = e0; e < e1; ++e) {
        const uint4 c = a.cp[e];
        const float vo = a.pops[c.y & kCplPopMask].st[c.x];
        const uint32_t kind = (c.y >> 16) & 3u;
        const float g = __uint_as_float(c.z);
        if (kind == 0u) {
            const float dv = vo - vme;
            I = I + g * dv;
        } else if (kind == 1u) {
            const float dv = vme - vo;
            I = I - g * dv;
        } else if (kind == 2u) {
            const float dv = vo - vme;
            I = I + g * dv;
        } else {
            const dptr<const float> pf = a.sig_prof + static_cast<size_t>(kSigProf) * (c.y >> 18);
            const float arg = (vo - pf[0]) * pf[1];
            const float inf = 1.0f / (1.0f + det_expf(-arg));
            const float tau = 1.0f / (pf[2] * (1.0f + det_expf(arg)));
            const float decay = det_expf(-pf[5] / tau);
            float s = a.sig_s[c.w];
            s = s + (inf - s) * (1.0f - decay);
            a.sig_s[c.w] = s;
            const float current = g * pf[4] * s * (pf[3] - vme);
            I = I + current;
        }
    }
    *cell = I;
}

// Single population (the CompiledModel path, PC/src/engine.rs:98-423): one
// thread per neuron advances it with the device-held input current i_in
// (the same every step until the host uploads another) and the call's dt,
// with the network's update (Neuron<MODEL>, including the +-100 input clamp
// of LIF and Izhikevich, P/src/core/exec/population.rs:69-88), and writes
// the 0/1 spike flag and, by wave ballot, the step's spike bit word.
template <int MODEL>
__global__ __launch_bounds__(kTile) void pop_step_kernel(PopDev P, const float* __restrict__ i_in,
                                                         int32_t* __restrict__ spikes, uint64_t* __restrict__ bits,
                                                         float dt) {
    const uint32_t li = blockIdx.x * kTile + threadIdx.x;
    int spike = 0;
    if (li < P.n) {
        Neuron<MODEL> nrn;
        nrn.load(P, li);
        spike = nrn.step(P, li, i_in[li], dt);
        nrn.store(P, li);
        spikes[li] = spike;
    }
    const uint64_t m = __ballot(spike);
    if ((threadIdx.x & 63u) == 0u && li < P.n) bits[li >> 6] = m;
}

// The last step's spiking neurons in ascending order from its bit words
// (one block): thread t counts words [t * per, (t + 1) * per), a block scan
// gives each thread its first output slot.  The reference's compaction is an
// atomic counter (PC/src/engine.rs:40-52), so its order is unspecified; the
// set is the same.
__global__ __launch_bounds__(1024) void pop_compact_kernel(const uint64_t* __restrict__ bits, uint32_t n_words,
                                                           uint32_t* __restrict__ ids, uint32_t* __restrict__ count) {
    __shared__ uint32_t s_wave[1024 / 64];
    const uint32_t t = threadIdx.x, per = (n_words + 1023u) / 1024u;
    const uint32_t w0 = min(t * per, n_words), w1 = min(w0 + per, n_words);
    uint32_t c = 0;
    for (uint32_t w = w0; w < w1; ++w) c += static_cast<uint32_t>(__popcll(bits[w]));
    uint32_t excl;
    const uint32_t total = block_scan<1024>(c, s_wave, excl);
    for (uint32_t w = w0; w < w1; ++w) {
        uint64_t m = bits[w];
        while (m) {
            ids[excl++] = w * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1);
            m &= m - 1ull;
        }
    }
    if (t == 0) *count = total;
}

// Holds the stream for `ticks` of the constant-rate wall clock: work queued
// behind it starts without waiting for the host's submission.
__global__ void spacer_kernel(uint64_t ticks) {
    const uint64_t t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
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
void launch_neuron(int model, int sched, const NetArgs* nets, const uint4* items, uint32_t n, hipStream_t s,
                   uint32_t defer_scatter, uint32_t bins) {
    if (n == 0) return;
    const size_t lds = sizeof(uint32_t) * bins;
    if (sched == 2) {
        if (model == 0) neuron_kernel<0, 2><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 1) neuron_kernel<1, 2><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 2) neuron_kernel<2, 2><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 3) neuron_kernel<3, 2><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else neuron_kernel<4, 2><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
    } else {
        if (model == 0) neuron_kernel<0, 1><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 1) neuron_kernel<1, 1><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 2) neuron_kernel<2, 1><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else if (model == 3) neuron_kernel<3, 1><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
        else neuron_kernel<4, 1><<<n, kTile, lds, s>>>(nets, items, defer_scatter);
    }
}
void launch_dense_trace(const NetArgs* nets, uint32_t n_tiles, float* trace, float decay, hipStream_t s) {
    if (n_tiles == 0) return;
    dense_trace_kernel<<<n_tiles, kTile, 0, s>>>(nets, trace, decay);
}
void launch_dense_weights(const NetArgs* nets, uint32_t N, const float* trace, float lr_ltp, float lr_ltd,
                          float w_min, float w_max, hipStream_t s) {
    if (N == 0) return;
    dense_stdp_kernel<<<(N + kTile - 1) / kTile, kTile, 0, s>>>(nets, trace, lr_ltp, lr_ltd, w_min, w_max);
}
void launch_stdp(const NetArgs* nets, const uint2* tiles, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    stdp_kernel<<<n, kTile, 0, s>>>(nets, tiles);
}
void launch_insert(int sched, const NetArgs* nets, const uint2* blocks, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (sched == 2) insert_kernel<2><<<n, kTile, 0, s>>>(nets, blocks);
    else insert_kernel<1><<<n, kTile, 0, s>>>(nets, blocks);
}
void launch_bin(const NetArgs* nets, const uint4* refs, uint32_t n, uint32_t bins, hipStream_t s) {
    if (n == 0) return;
    bin_kernel<<<n, kTile, sizeof(uint32_t) * bins, s>>>(nets, refs);
}
void launch_bin_apply(const NetArgs* nets, const uint2* refs, uint32_t n, uint32_t mode, hipStream_t s) {
    if (n == 0) return;
    bin_apply_kernel<false><<<n, kApply, 0, s>>>(nets, refs, mode);
}
void launch_bin_apply_rec(const NetArgs* nets, const uint2* refs, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    bin_apply_kernel<true><<<n, kApply, 0, s>>>(nets, refs, 0u);
}
void launch_megastep(int model, const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (model == 0) megastep_kernel<0><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 1) megastep_kernel<1><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 2) megastep_kernel<2><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 3) megastep_kernel<3><<<n, kApply, 0, s>>>(nets, items);
    else megastep_kernel<4><<<n, kApply, 0, s>>>(nets, items);
}
template <bool kRec>
static void launch_front_r(int model, const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s) {
    if (model == 0) front_kernel<0, kRec><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 1) front_kernel<1, kRec><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 2) front_kernel<2, kRec><<<n, kApply, 0, s>>>(nets, items);
    else if (model == 3) front_kernel<3, kRec><<<n, kApply, 0, s>>>(nets, items);
    else front_kernel<4, kRec><<<n, kApply, 0, s>>>(nets, items);
}
void launch_front(int model, bool rec, const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    if (rec) launch_front_r<true>(model, nets, items, n, s);
    else launch_front_r<false>(model, nets, items, n, s);
}
void launch_bin_apply_rec_mode(const NetArgs* nets, const uint2* refs, uint32_t n, uint32_t mode, hipStream_t s) {
    if (n == 0) return;
    bin_apply_kernel<true><<<n, kApply, 0, s>>>(nets, refs, mode);
}
void launch_fixup(const NetArgs* nets, const uint2* refs, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    fixup_kernel<<<n, kTile, 0, s>>>(nets, refs);
}
void launch_head_clear(unsigned long long* head, const uint64_t* step, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    head_clear_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(head, step, n);
}
void launch_inject(float* ring, const uint64_t* step, uint32_t N, const uint32_t* idx, const float* val,
                   uint32_t count, int serial, hipStream_t s) {
    if (count == 0) return;
    if (serial) inject_kernel<<<1, 1, 0, s>>>(ring, step, N, idx, val, count, 1);
    else inject_kernel<<<(count + 255) / 256, 256, 0, s>>>(ring, step, N, idx, val, count, 0);
}
void launch_spacer(uint64_t ticks, hipStream_t s) { spacer_kernel<<<1, 64, 0, s>>>(ticks); }
void launch_pop_step(int model, const PopDev& P, const float* i_in, int32_t* spikes, uint64_t* bits, float dt,
                     hipStream_t s) {
    const uint32_t g = (P.n + kTile - 1) / kTile;
    if (g == 0) return;
    if (model == 0) pop_step_kernel<0><<<g, kTile, 0, s>>>(P, i_in, spikes, bits, dt);
    else if (model == 1) pop_step_kernel<1><<<g, kTile, 0, s>>>(P, i_in, spikes, bits, dt);
    else if (model == 2) pop_step_kernel<2><<<g, kTile, 0, s>>>(P, i_in, spikes, bits, dt);
    else if (model == 3) pop_step_kernel<3><<<g, kTile, 0, s>>>(P, i_in, spikes, bits, dt);
    else pop_step_kernel<4><<<g, kTile, 0, s>>>(P, i_in, spikes, bits, dt);
}
void launch_pop_compact(const uint64_t* bits, uint32_t n_words, uint32_t* ids, uint32_t* count, hipStream_t s) {
    pop_compact_kernel<<<1, 1024, 0, s>>>(bits, n_words, ids, count);
}
void launch_coupling(const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s) {
    if (n) coupling_kernel<<<n, kTile, 0, s>>>(nets, items);
}
void launch_vjump(float* v, const uint32_t* idx, const float* val, uint32_t count, int serial, hipStream_t s) {
    if (count == 0) return;
    if (serial) vjump_kernel<<<1, 1, 0, s>>>(v, idx, val, count, 1);
    else vjump_kernel<<<(count + 255) / 256, 256, 0, s>>>(v, idx, val, count, 0);
}
void launch_tm_init(const uint8_t* type, const float4* prof, float* r, float* u, uint32_t* last, size_t n,
                    hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    tm_init_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(type, prof, r, u, last, n);
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
