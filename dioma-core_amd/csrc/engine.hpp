// engine.hpp — host-side runtime shared by the single-network, batch and
// serving C APIs.
//
// NetState: every device buffer of one network plus its host bookkeeping.
// Engine:   one HIP stream stepping a set of networks together.  The step
//           kernels read a device array of NetArgs (one per network) and
//           per-launch work lists of (network, tile) pairs, so one launch
//           covers every network of the set.  Each network keeps its own
//           ring, step counter and RNG keys, so its results do not depend on
//           which other networks share the launch.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/simuse.h"
#include "net_kernels.hpp"

namespace simuse {

int fail(int code, const char* fmt, ...);

void launch_neuron(int model, int sched, const NetArgs* nets, const uint4* items, uint32_t n, hipStream_t s,
                   uint32_t defer_scatter, uint32_t bins);
void launch_dense_trace(const NetArgs* nets, uint32_t n_tiles, float* trace, float decay, hipStream_t s);
void launch_dense_weights(const NetArgs* nets, uint32_t N, const float* trace, float lr_ltp, float lr_ltd,
                          float w_min, float w_max, hipStream_t s);
void launch_stdp(const NetArgs* nets, const uint2* tiles, uint32_t n, hipStream_t s);
void launch_insert(int sched, const NetArgs* nets, const uint2* blocks, uint32_t n, hipStream_t s);
void launch_fixup(const NetArgs* nets, const uint2* refs, uint32_t n, hipStream_t s);
void launch_bin(const NetArgs* nets, const uint4* refs, uint32_t n, uint32_t bins, hipStream_t s);
void launch_bin_apply(const NetArgs* nets, const uint2* refs, uint32_t n, uint32_t mode, hipStream_t s);
void launch_bin_apply_rec(const NetArgs* nets, const uint2* refs, uint32_t n, hipStream_t s);
void launch_megastep(int model, const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s);
void launch_front(int model, bool rec, const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s);
void launch_bin_apply_rec_mode(const NetArgs* nets, const uint2* refs, uint32_t n, uint32_t mode, hipStream_t s);

void launch_head_clear(unsigned long long* head, const uint64_t* step, size_t n, hipStream_t s);
void launch_inject(float* ring, const uint64_t* step, uint32_t N, const uint32_t* idx, const float* val,
                   uint32_t count, int serial, hipStream_t s);
void launch_spacer(uint64_t ticks, hipStream_t s);
void launch_coupling(const NetArgs* nets, const uint2* items, uint32_t n, hipStream_t s);
void launch_vjump(float* v, const uint32_t* idx, const float* val, uint32_t count, int serial, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_tm_init(const uint8_t* type, const float4* prof, float* r, float* u, uint32_t* last, size_t n,
                    hipStream_t s);
void launch_widen(const float* w, double* w64, size_t n, hipStream_t s);
void launch_raster_stats(const uint64_t* raster, uint32_t n_words, uint32_t steps, const uint32_t* word_of,
                         const uint32_t* bit_of, uint32_t count, float dt_s, float* rates, float* cvs,
                         hipStream_t s);

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::simuse::fail(SIMUSE_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

inline int model_states(int m) {
    return m == SIMUSE_MODEL_LIF ? 2 : m == SIMUSE_MODEL_IZHIKEVICH ? 3 : m == SIMUSE_MODEL_HH ? 6
         : m == SIMUSE_MODEL_ALPHA_LIF ? 3 : 9;
}
inline int model_params(int m) {
    return m == SIMUSE_MODEL_LIF ? 6 : m == SIMUSE_MODEL_IZHIKEVICH ? 7 : m == SIMUSE_MODEL_HH ? 13
         : m == SIMUSE_MODEL_ALPHA_LIF ? 7 : 14;
}
// column of tau_ref and of the refractory state (LIF/Izhikevich only)
inline int model_tau_ref_col(int m) { return m == SIMUSE_MODEL_LIF || m == SIMUSE_MODEL_IZHIKEVICH ? 5 : -1; }
inline int model_refr_state(int m) { return m == SIMUSE_MODEL_LIF ? 1 : m == SIMUSE_MODEL_IZHIKEVICH ? 2 : -1; }

struct HostPop {
    int model = 0;
    uint32_t n = 0, ns = 0, np = 0, host_stride = 0;
    uint32_t off = 0, tile0 = 0, ntiles = 0, word0 = 0, dstride = 0;
    uint32_t uni_mask = 0, flags = 0;
    float puni[kMaxParams] = {};
    float* st = nullptr;
    float* pr = nullptr;
    float* sigma = nullptr;
    float sigma_uni = 0.0f;
    float drive_p = 0.0f, drive_w = 0.0f;
    uint32_t drive_seed0 = 0;
    uint64_t seed = 0;
};

struct SourceHost {
    float p;
    uint32_t seed;
    std::vector<uint32_t> post;
    std::vector<float> w;
    std::vector<uint8_t> d;
};

struct Engine;

// Validation shared by create and reset.
int validate_pops(const simuse_pop_desc* pops, int32_t n_pops, uint64_t* n_total);
int validate_csr(uint64_t n_total, const uint32_t* row_ptr, const uint32_t* col, const float* w,
                 const uint8_t* delay, uint64_t n_syn);

struct NetState {
    Engine* eng = nullptr;
    std::vector<HostPop> pops;
    uint32_t N = 0, n_tiles = 0, n_words = 0;
    uint64_t n_syn = 0;
    std::vector<void*> allocs;
    uint64_t bytes = 0;
    PopDev* d_pops = nullptr;
    uint32_t* d_tile_pop = nullptr;
    float* ring = nullptr;
    unsigned long long* head = nullptr;
    uint2* rec = nullptr;
    uint2* coll = nullptr;
    uint32_t* tile_coll = nullptr;
    uint32_t* tile_ebase = nullptr;
    uint32_t n_fix = 1;
    uint32_t max_delay = 0;    // largest synapse or source delay (steps)
    // binned scatter (every delay <= 1 step): bins of kBinPosts posts
    bool binned = false;
    uint32_t n_bins = 0, bin_rows = 0;
    uint2* bin_hdr = nullptr;
    // megastep (binned, one model, no pair-based plasticity): records and
    // headers double-buffered by step parity
    bool mega_ok = false;   // buffers for it exist
    bool no_mega = false;   // left for the two-kernel path (dense STDP step, plasticity)
    bool mega = false;      // running it (set with the work lists)
    int mega_model = 0;
    // split form (binned, one model, pair-based plasticity or receptor
    // synapses): front_kernel + bin_kernel, the same parity buffers
    bool split_ok = false;
    bool split = false;     // running it (set with the work lists)
    bool no_split = false;  // left for the three-kernel path (dense STDP step)
    bool hdr_dense = false;  // the header matrices hold whole columns (one-kernel or split form last)
    uint32_t nbuf() const { return mega_ok || split_ok ? 2u : 1u; }  // header and record buffers
    uint32_t* bin_done = nullptr;
    uint32_t* bin_mat = nullptr;
    std::vector<uint2> mat_items;      // (engine index, bin) of every bin, XCD-ordered
    uint2* d_mat_items = nullptr;      // [n_bins + kXcdPad]
    bool mat_dirty = true;
    std::vector<uint4> dense_writers;  // every writer as a bin_kernel item (dense STDP step)
    uint4* d_dense_writers = nullptr;  // [bin_rows + kXcdPad]
    bool dense_dirty = true;
    uint2* bin_ev = nullptr;   // [1 or 2][n_ev] records (two uint2 per record when ev_wide)
    bool ev_wide = false;      // 16-byte records: receptor synapses or a writer spanning >= 2^21 edges
    uint32_t* bin_slot = nullptr;
    uint64_t since_clear = 0;  // steps enqueued since the heads were last rewritten
    uint32_t* row_ptr = nullptr;
    uint32_t* col = nullptr;
    float* w = nullptr;
    uint8_t* delay = nullptr;
    uint32_t* tile_count = nullptr;
    uint32_t* active = nullptr;
    uint64_t* bits = nullptr;
    uint64_t* tile_spikes = nullptr;
    uint64_t* tile_events = nullptr;
    uint64_t* step = nullptr;  // [0] step, [1] raster base
    std::vector<SourceHost> srcs;
    uint32_t n_src = 0, n_src_tiles = 0;
    uint64_t n_src_edges = 0;
    uint32_t* src_row_ptr = nullptr;
    uint32_t* src_post = nullptr;
    float* src_w = nullptr;
    uint8_t* src_delay = nullptr;
    float* src_p = nullptr;
    uint32_t* src_seed = nullptr;
    uint64_t* raster = nullptr;
    uint32_t raster_cap = 0;
    uint64_t raster_start = 0;
    bool stdp = false;
    simuse_stdp_params sp{};
    double* w64 = nullptr;
    float* tr_pre = nullptr;
    float* tr_post = nullptr;
    uint32_t* in_ptr = nullptr;
    uint32_t* in_edge = nullptr;
    uint32_t* edge_pre = nullptr;
    uint8_t* plastic = nullptr;
    float stdp_dpre = 0, stdp_dpost = 0;
    float* dtrace = nullptr;      // shared trace of the dense stdp_step rule
    // Tsodyks-Markram short-term plasticity (host copies until finalize)
    std::vector<float4> h_tm_prof;
    std::vector<uint8_t> h_tm_type;  // synapses then source edges; empty = off
    uint8_t* tm_type = nullptr;
    float* tm_r = nullptr;
    float* tm_u = nullptr;
    uint32_t* tm_last = nullptr;
    float4* tm_prof = nullptr;
    float* ev_w = nullptr;
    // conductance-based receptor synapses (host copies until finalize)
    std::vector<uint8_t> h_rec_code;  // per synapse: 0 current-based, 1 + receptor type; empty = off
    std::vector<float> h_g_ext;       // constant external current per neuron (empty: none)
    float rec_scale = 1.0f;
    uint32_t rec_live = 0;  // bit r: receptor type r can hold a nonzero conductance
    uint8_t* rec_code = nullptr;
    float* g = nullptr;               // [kReceptors][N]
    float* g_ext = nullptr;
    // electrical coupling (host lists in call order until finalize)
    struct CplHost {
        uint32_t a, b;
        float g;
    };
    std::vector<CplHost> h_gap, h_graded;
    std::vector<std::array<float, kSigProf>> h_sig_prof;  // vth, 1/delta, k, erev, conductance, dt
    struct SigConnHost {
        uint32_t prof, pre, post;
        float w;
    };
    std::vector<SigConnHost> h_sig_conn;  // call order
    uint32_t* cp_ptr = nullptr;           // [N+1]
    uint4* cp = nullptr;
    float* sig_s = nullptr;               // [connections], sigmoid activation
    float* sig_prof = nullptr;
    uint32_t n_sig = 0;
    bool coupled() const { return !h_gap.empty() || !h_graded.empty() || !h_sig_conn.empty(); }
    uint2* d_all_blocks = nullptr;  // (0, b) for every scatter tile: dense-step insert list
    bool finalized = false;
    uint64_t host_step = 0;
    uint64_t step_limit = std::numeric_limits<uint64_t>::max();
    uint32_t* inj_idx = nullptr;
    float* inj_val = nullptr;
    uint32_t inj_cap = 0;
    std::vector<uint32_t> h_row_ptr;
    uint32_t uni_deg = 0;  // every row has this out-degree (0: rows differ)

    NetState() = default;
    NetState(const NetState&) = delete;
    NetState& operator=(const NetState&) = delete;
    ~NetState() {
        for (void* p : allocs) (void)hipFree(p);
    }

    template <class T>
    int alloc(T** p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                        hipGetErrorString(e));
        }
        allocs.push_back(q);
        bytes += count * sizeof(T);
        *p = static_cast<T*>(q);
        return SIMUSE_OK;
    }

    NetArgs args() const {
        NetArgs a{};
        a.pops = dev_ptr(d_pops);
        a.tile_pop = dev_ptr(d_tile_pop);
        a.N = N;
        a.n_tiles = n_tiles;
        a.n_src_tiles = n_src_tiles;
        a.n_src = n_src;
        a.n_blocks = n_tiles + n_src_tiles;
        a.n_fix = n_fix;
        a.n_syn = n_syn;
        a.step_limit = step_limit;
        a.ring = dev_ptr(ring);
        a.ring_fresh = max_delay <= 1 ? 1 : 0;
        a.n_bins = binned ? n_bins : 0u;
        a.bin_hdr = dev_ptr(bin_hdr);
        a.bin_rows = bin_rows;
        a.n_pops = static_cast<uint32_t>(pops.size());
        a.n_ev = n_syn + n_src_edges;
        a.hdr_stride = static_cast<uint64_t>(n_bins) * bin_rows;
        a.bin_done = dev_ptr(bin_done);
        a.bin_mat = dev_ptr(bin_mat);
        a.bin_ev = dev_ptr(bin_ev);
        a.ev_wide = ev_wide ? 1 : 0;
        a.bin_slot = dev_ptr(bin_slot);
        a.head = dev_ptr(head);
        a.rec = dev_ptr(rec);
        a.coll = dev_ptr(coll);
        a.tile_coll = dev_ptr(tile_coll);
        a.tile_ebase = dev_ptr(tile_ebase);
        a.row_ptr = dev_ptr(row_ptr);
        a.uni_deg = uni_deg;
        a.col = dev_ptr(col);
        a.w = dev_ptr(w);
        a.delay = dev_ptr(delay);
        a.src_row_ptr = dev_ptr(src_row_ptr);
        a.src_post = dev_ptr(src_post);
        a.src_w = dev_ptr(src_w);
        a.src_delay = dev_ptr(src_delay);
        a.src_p = dev_ptr(src_p);
        a.src_seed = dev_ptr(src_seed);
        a.tile_count = dev_ptr(tile_count);
        a.active = dev_ptr(active);
        a.bits = dev_ptr(bits);
        a.tile_spikes = dev_ptr(tile_spikes);
        a.tile_events = dev_ptr(tile_events);
        a.step = dev_ptr(step);
        a.raster = dev_ptr(raster);
        a.raster_cap = raster_cap;
        a.n_words = n_words;
        a.raster_base = dev_ptr(step + 1);
        a.stdp = stdp ? 1 : 0;
        a.w64 = dev_ptr(w64);
        a.tr_pre = dev_ptr(tr_pre);
        a.tr_post = dev_ptr(tr_post);
        a.in_ptr = dev_ptr(in_ptr);
        a.in_edge = dev_ptr(in_edge);
        a.edge_pre = dev_ptr(edge_pre);
        a.plastic = dev_ptr(plastic);
        a.stdp_dpre = stdp_dpre;
        a.stdp_dpost = stdp_dpost;
        a.a_plus = sp.a_plus;
        a.a_minus = sp.a_minus;
        a.w_min = sp.w_min;
        a.w_max = sp.w_max;
        a.tm_type = dev_ptr(tm_type);
        a.tm_r = dev_ptr(tm_r);
        a.tm_u = dev_ptr(tm_u);
        a.tm_last = dev_ptr(tm_last);
        a.tm_prof = dev_ptr(tm_prof);
        a.ev_w = dev_ptr(ev_w);
        a.rec_code = dev_ptr(rec_code);
        a.g = dev_ptr(g);
        a.g_ext = dev_ptr(g_ext);
        a.rec_scale = rec_scale;
        a.rec_live = rec_live;
        a.n8 = N & ~7u;
        a.cp_ptr = dev_ptr(static_cast<const uint32_t*>(cp_ptr));
        a.cp = dev_ptr(static_cast<const uint4*>(cp));
        a.sig_s = dev_ptr(sig_s);
        a.sig_prof = dev_ptr(static_cast<const float*>(sig_prof));
        return a;
    }

    // All creation work after validation.  `stream` orders the device fills.
    int init(const simuse_pop_desc* pd, int32_t n_pops, const uint32_t* rp, const uint32_t* c, const float* wf,
             const uint8_t* dl, uint64_t nsyn, hipStream_t stream);
    int finalize(hipStream_t stream);
    int finalize_coupling(hipStream_t stream);
    int upload_pops(const simuse_pop_desc* pd);
    int upload_pop_table();
    int reset_dynamic(hipStream_t stream);
    void clear_heads(hipStream_t stream) {
        if (head) launch_head_clear(head, step, static_cast<size_t>(kSlots) * N, stream);
        since_clear = 0;
    }
    int add_source(float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts, const float* wf,
                   const uint8_t* dl, uint32_t count);
    int set_noise(int32_t pop, const float* sigma, uint64_t seed);
    int set_drive(int32_t pop, float rate_hz, float dt_ms, float weight, uint32_t seed0);
    int set_plasticity(const simuse_stdp_params* p, const uint8_t* pl, hipStream_t stream);
    int set_stp(const simuse_stp_profile* profiles, int32_t n_profiles, float dt, const uint8_t* syn_types,
                const uint8_t* src_types, uint64_t n_src_edges_given);
    int set_receptors(const uint8_t* types, float scale, const float* ext);
    int add_coupling(int kind, const uint32_t* a, const uint32_t* b, const float* g, uint32_t count);
    int add_sigmoid_profile(float vth, float delta, float k, float erev, float cond, float dt_ms, int32_t* id);
    int add_sigmoid_conns(int32_t prof, const uint32_t* pre, const uint32_t* post, const float* w, uint32_t count);
    int download_conductances(float* out, uint64_t len, hipStream_t stream);
    int download_spikes(int32_t* out, uint64_t len, hipStream_t stream);
    int download_spike_ids(uint32_t* ids, uint32_t cap, uint32_t* count, hipStream_t stream);
    int download_state(int32_t pop, float* out, uint64_t len, hipStream_t stream);
    int download_ring(float* out, uint64_t len, hipStream_t stream);
    int upload_weights(const float* wf, uint64_t n, hipStream_t stream);
    int download_weights(float* wf, uint64_t n, hipStream_t stream);
    int inject(int32_t pop, const uint32_t* idx, const float* val, uint32_t count, hipStream_t stream);
    int inject_voltage(int32_t pop, const uint32_t* idx, const float* val, uint32_t count, hipStream_t stream);
    int reset(const simuse_pop_desc* pd, int32_t n_pops, hipStream_t stream);
    int counters(uint64_t* spikes, uint64_t* events, hipStream_t stream);
    int record(uint32_t max_steps, hipStream_t stream);
    int download_raster(uint64_t* out, uint64_t n_words_buf, uint32_t* steps, hipStream_t stream);
    int raster_stats(uint32_t first, uint32_t count, float dt_s, float* rates, float* cvs, hipStream_t stream);
    // checkpoint / resume as a .taas v2 file (snapshot.hip)
    int fingerprint(std::vector<uint64_t>& fp) const;
    int save_state(const char* path, hipStream_t stream);
    int load_state(const char* path, hipStream_t stream);
};

struct Engine {
    hipStream_t stream = nullptr;
    std::vector<NetState*> nets;
    bool dirty = true;
    // device tables
    NetArgs* d_nets = nullptr;
    uint32_t nets_cap = 0;
    static constexpr int kModels = 5;  // LIF, Izhikevich, HH, GPEA, alpha-LIF
    uint4* d_tiles[kModels] = {};      // neuron work items per model: {net, tile, tiles | writer << 4, pop}
    uint32_t n_tiles_m[kModels] = {}, cap_tiles_m[kModels] = {};
    uint2* d_mega[kModels] = {};       // megastep work items per model: {net, bin}
    uint32_t n_mega_m[kModels] = {}, cap_mega_m[kModels] = {};
    uint2* d_front[2 * kModels] = {};  // front_kernel items per (model, receptors): {net, bin * parts + part}
    uint32_t n_front_m[2 * kModels] = {}, cap_front_m[2 * kModels] = {};
    uint2* d_blocks = nullptr;  // insert_kernel work: plastic networks' tiles, source tiles
    uint32_t n_blocks = 0, cap_blocks = 0;
    uint2* d_stdp = nullptr;
    uint32_t n_stdp = 0, cap_stdp = 0;
    uint2* d_fix = nullptr;
    uint32_t n_fix = 0, cap_fix = 0;
    uint4* d_bin = nullptr;     // bin_kernel writers of binned networks
    uint32_t n_bin = 0, cap_bin = 0;
    uint32_t bins_fused = 0, bins_bin = 0;  // most bins of a network with neuron-kernel / bin_kernel writers
    uint2* d_apply = nullptr;   // bins of binned networks, XCD-ordered
    uint32_t n_apply = 0, cap_apply = 0;
    uint2* d_apply_rec = nullptr;  // the same for networks with receptor synapses
    uint32_t n_apply_rec = 0, cap_apply_rec = 0;
    uint2* d_cpl = nullptr;     // coupling_kernel work: {net, tile} of coupled networks
    uint32_t n_cpl = 0, cap_cpl = 0;
    bool direct_fix = false;
    // execution
    bool one_kernel = true;  // networks that can take the one-kernel step do (the autotuner may turn it off)
    int sched = SIMUSE_SCATTER_TILE;
    uint32_t group = 1;  // neuron tiles per block (1..kMaxGroup; the autotuner may change it)
    uint32_t graph_steps = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    static constexpr int kPhases = 4;  // neuron (+ fused scatter), stdp, insert + bin, apply + fix-up
    bool profiling = false;
    std::vector<hipEvent_t> ev;
    uint32_t prof_steps = 0;
    hipEvent_t tev[2] = {nullptr, nullptr};  // simuse_net_time_steps' pair, kept for the engine's life

    Engine() = default;
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    ~Engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : tev)
            if (e) (void)hipEventDestroy(e);
        for (uint4* t : d_tiles)
            if (t) (void)hipFree(t);
        for (uint2* t : d_mega)
            if (t) (void)hipFree(t);
        for (uint2* t : d_front)
            if (t) (void)hipFree(t);
        for (void* p : {static_cast<void*>(d_nets), static_cast<void*>(d_blocks), static_cast<void*>(d_stdp),
                        static_cast<void*>(d_fix), static_cast<void*>(d_bin), static_cast<void*>(d_apply),
                        static_cast<void*>(d_apply_rec), static_cast<void*>(d_cpl)})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int init() {
        // a blocking stream is ordered after the synchronous null-stream uploads
        HIP_TRY(hipStreamCreate(&stream));
        return SIMUSE_OK;
    }
    int timing_events() {
        for (hipEvent_t& e : tev)
            if (!e) HIP_TRY(hipEventCreate(&e));
        return SIMUSE_OK;
    }
    void drop_graph() {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
    void mark_dirty() {
        dirty = true;
        drop_graph();
    }

    template <class T>
    static int grow(T** p, uint32_t* cap, uint32_t need) {
        if (need <= *cap && *p) return SIMUSE_OK;
        uint32_t nc = std::max<uint32_t>(need, std::max<uint32_t>(64, *cap * 2));
        if (*p) { (void)hipFree(*p); *p = nullptr; *cap = 0; }
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * static_cast<size_t>(nc));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "engine table allocation failed: %s", hipGetErrorString(e));
        }
        *p = static_cast<T*>(q);
        *cap = nc;
        return SIMUSE_OK;
    }

    // Workgroup b of a launch runs on XCD b % kXcds: reorder a work list so
    // that runs of kXcdChunk consecutive items go to the XCDs in turn (run c
    // to XCD c % kXcds; padded with `pad`, which the kernels skip).  A run's
    // bins write their entries of every header row -- 16 adjacent 8-byte
    // entries, one 128-byte line -- from one L2, and populations whose bins
    // carry more work (more spikes to push) spread over the XCDs instead of
    // loading a few of them.
    template <class T>
    static void xcd_order(std::vector<T>& v, const T& pad) {
        const size_t n = v.size();
        std::vector<std::vector<T>> per_x(kXcds);
        for (size_t i = 0; i < n; ++i) per_x[(i / kXcdChunk) % kXcds].push_back(v[i]);
        size_t per = 0;
        for (const auto& q : per_x) per = std::max(per, q.size());
        std::vector<T> o(per * kXcds, pad);
        for (size_t b = 0; b < o.size(); ++b) {
            const auto& q = per_x[b % kXcds];
            if (b / kXcds < q.size()) o[b] = q[b / kXcds];
        }
        v.swap(o);
    }

    // Rebuild the NetArgs array and the (network, tile) work lists.
    int ensure_tables() {
        if (!dirty) return SIMUSE_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        std::vector<NetArgs> na;
        std::vector<uint4> tl[kModels], binl;
        std::vector<uint2> bl, sl, fl, apl, aplr, ml[kModels], fr[2 * kModels], cpl;
        bins_fused = bins_bin = 0;
        for (uint32_t k = 0; k < nets.size(); ++k) {
            NetState* s = nets[k];
            na.push_back(s->args());
            NetArgs& A = na.back();
            const uint32_t n_blk = s->n_tiles + s->n_src_tiles;
            if (s->cp_ptr)
                for (uint32_t t = 0; t < s->n_tiles; ++t) cpl.push_back(make_uint2(k, t));
            uint32_t ng = 0;  // neuron blocks
            for (const HostPop& P : s->pops) ng += (P.ntiles + group - 1) / group;
            const uint32_t ns = (s->n_src_tiles + kBinBlockTiles - 1) / kBinBlockTiles;
            // binned networks without pair-based plasticity: neuron block j is writer j
            // (receptor networks: bin_kernel writers, which tag each event with its receptor)
            const bool fz = s->binned && !s->stdp && !s->rec_code && ng + ns <= kMaxWriters;
            // megastep: bin b's block is writer b, the source blocks follow
            s->mega = s->binned && s->mega_ok && !s->no_mega && !s->stdp && one_kernel;
            s->split = s->binned && s->split_ok && !s->no_split && !s->mega;
            // The one-kernel and split forms fill whole header columns and
            // leave them as read; the binned form writes only the entries of
            // bins with events and reads every other entry as empty, so a
            // network that moves to it starts from cleared headers (its
            // pending events were applied to the ring before the move).
            const bool dense = s->mega || s->split;
            if (s->hdr_dense && !dense && s->bin_hdr)
                HIP_TRY(hipMemsetAsync(s->bin_hdr, 0, sizeof(uint2) * s->nbuf() * s->n_bins * s->bin_rows, stream));
            s->hdr_dense = dense;
            if (s->mega) {
                s->mat_items.clear();
                for (uint32_t b = 0; b < s->n_bins; ++b) {
                    ml[s->mega_model].push_back(make_uint2(k, b));
                    s->mat_items.push_back(make_uint2(k, b));
                }
                xcd_order(s->mat_items, make_uint2(kEmpty, 0u));
                s->mat_dirty = true;
                uint32_t jw = s->n_bins;
                for (uint32_t b = s->n_tiles; b < n_blk; b += kBinBlockTiles, ++jw)
                    binl.push_back(make_uint4(k, b, std::min(kBinBlockTiles, n_blk - b), jw));
                if (jw > s->bin_rows) return fail(SIMUSE_E_STATE, "megastep: %u writers, %u rows", jw, s->bin_rows);
                A.n_bblk = jw;
                A.n_done = jw;
                A.mega = 1;
                A.bin_fused = 0;
                if (s->n_src_tiles) bins_bin = std::max(bins_bin, s->n_bins);
                continue;
            }
            // split form: front blocks (kFrontParts per bin), then bin_kernel's
            // writers over every scatter tile and the in-edge LTP blocks, all
            // counted for the step's advance
            if (s->split) {
                s->mat_items.clear();
                const int fi = 2 * s->mega_model + (s->rec_code ? 1 : 0);
                for (uint32_t b = 0; b < s->n_bins; ++b) {
                    for (uint32_t h = 0; h < kFrontParts; ++h) fr[fi].push_back(make_uint2(k, b * kFrontParts + h));
                    s->mat_items.push_back(make_uint2(k, b));
                }
                xcd_order(s->mat_items, make_uint2(kEmpty, 0u));
                s->mat_dirty = true;
                const uint32_t T = std::max(kBinBlockTiles, (n_blk + kMaxWriters - 1) / kMaxWriters);
                uint32_t jw = 0, nd = 0;
                for (uint32_t b = 0; b < n_blk; b += T, ++jw, ++nd)
                    binl.push_back(make_uint4(k, b, std::min(T, n_blk - b), jw));
                if (jw > s->bin_rows) return fail(SIMUSE_E_STATE, "split form: %u writers, %u rows", jw, s->bin_rows);
                if (s->stdp)
                    for (uint32_t b = 0; b < s->n_tiles; b += kBinBlockTiles, ++nd)
                        binl.push_back(make_uint4(k, b, std::min(kBinBlockTiles, s->n_tiles - b), kLtpItem));
                A.n_bblk = jw;
                A.n_done = nd;
                A.mega = 1;
                A.split = 1;
                A.bin_fused = 0;
                bins_bin = std::max(bins_bin, s->n_bins);
                continue;
            }
            uint32_t j = 0;
            for (uint32_t p = 0; p < s->pops.size(); ++p) {
                const HostPop& P = s->pops[p];
                for (uint32_t t = 0; t < P.ntiles; t += group, ++j)
                    tl[P.model].push_back(make_uint4(k, P.tile0 + t, std::min(group, P.ntiles - t) | (j << 4), p));
            }
            // binned plastic networks run their STDP inside the writers (bin_kernel)
            if (s->stdp && !s->binned)
                for (uint32_t t = 0; t < s->n_tiles; ++t) sl.push_back(make_uint2(k, t));
            if (s->binned) {
                std::vector<uint4>& wl = s->dense_writers;
                wl.clear();
                j = 0;
                if (fz) {
                    for (const HostPop& P : s->pops)
                        for (uint32_t t = 0; t < P.ntiles; t += group, ++j)
                            wl.push_back(make_uint4(k, P.tile0 + t, std::min(group, P.ntiles - t), j));
                    for (uint32_t b = s->n_tiles; b < n_blk; b += kBinBlockTiles, ++j) {
                        wl.push_back(make_uint4(k, b, std::min(kBinBlockTiles, n_blk - b), j));
                        binl.push_back(wl.back());
                    }
                } else {
                    const uint32_t T = std::max(kBinBlockTiles, (n_blk + kMaxWriters - 1) / kMaxWriters);
                    for (uint32_t b = 0; b < n_blk; b += T, ++j) {
                        wl.push_back(make_uint4(k, b, std::min(T, n_blk - b), j));
                        binl.push_back(wl.back());
                    }
                }
                if (j > s->bin_rows) return fail(SIMUSE_E_STATE, "binned scatter: %u writers, %u rows", j, s->bin_rows);
                // plastic networks: the LTP of the spiking posts' in-edges runs in
                // blocks of its own beside the writers (which take the out-edges)
                if (s->stdp)
                    for (uint32_t b = 0; b < s->n_tiles; b += kBinBlockTiles)
                        binl.push_back(make_uint4(k, b, std::min(kBinBlockTiles, s->n_tiles - b), kLtpItem));
                xcd_order(wl, make_uint4(kEmpty, 0u, 0u, 0u));
                s->dense_dirty = true;
                A.n_bblk = j;
                A.bin_fused = fz ? 1 : 0;
                if (fz) bins_fused = std::max(bins_fused, s->n_bins);
                bins_bin = std::max(bins_bin, s->n_bins);
                for (uint32_t b = 0; b < s->n_bins; ++b) (s->ev_wide ? aplr : apl).push_back(make_uint2(k, b));
                continue;
            }
            // chained scatter: plastic networks scatter after the STDP kernels;
            // sources always
            for (uint32_t b = s->stdp ? 0 : s->n_tiles; b < s->n_tiles + s->n_src_tiles; ++b)
                bl.push_back(make_uint2(k, b));
            for (uint32_t j = 0; j < s->n_fix; ++j) fl.push_back(make_uint2(k, j));
        }
        direct_fix = nets.size() == 1;  // one network: fix-up block j is blockIdx.x
        int rc;
        if (!na.empty()) {
            if ((rc = grow(&d_nets, &nets_cap, static_cast<uint32_t>(na.size())))) return rc;
            HIP_TRY(hipMemcpy(d_nets, na.data(), sizeof(NetArgs) * na.size(), hipMemcpyHostToDevice));
        }
        for (int m = 0; m < kModels; ++m) {
            xcd_order(ml[m], make_uint2(kEmpty, 0u));
            n_mega_m[m] = static_cast<uint32_t>(ml[m].size());
            if (!ml[m].empty()) {
                if ((rc = grow(&d_mega[m], &cap_mega_m[m], n_mega_m[m]))) return rc;
                HIP_TRY(hipMemcpy(d_mega[m], ml[m].data(), sizeof(uint2) * ml[m].size(), hipMemcpyHostToDevice));
            }
            for (int r = 0; r < 2; ++r) {
                std::vector<uint2>& f = fr[2 * m + r];
                xcd_order(f, make_uint2(kEmpty, 0u));
                n_front_m[2 * m + r] = static_cast<uint32_t>(f.size());
                if (f.empty()) continue;
                if ((rc = grow(&d_front[2 * m + r], &cap_front_m[2 * m + r], n_front_m[2 * m + r]))) return rc;
                HIP_TRY(hipMemcpy(d_front[2 * m + r], f.data(), sizeof(uint2) * f.size(), hipMemcpyHostToDevice));
            }
            xcd_order(tl[m], make_uint4(kEmpty, 0u, 0u, 0u));
            n_tiles_m[m] = static_cast<uint32_t>(tl[m].size());
            if (tl[m].empty()) continue;
            if ((rc = grow(&d_tiles[m], &cap_tiles_m[m], n_tiles_m[m]))) return rc;
            HIP_TRY(hipMemcpy(d_tiles[m], tl[m].data(), sizeof(uint4) * tl[m].size(), hipMemcpyHostToDevice));
        }
        n_blocks = static_cast<uint32_t>(bl.size());
        if (n_blocks) {
            if ((rc = grow(&d_blocks, &cap_blocks, n_blocks))) return rc;
            HIP_TRY(hipMemcpy(d_blocks, bl.data(), sizeof(uint2) * bl.size(), hipMemcpyHostToDevice));
        }
        n_stdp = static_cast<uint32_t>(sl.size());
        if (n_stdp) {
            if ((rc = grow(&d_stdp, &cap_stdp, n_stdp))) return rc;
            HIP_TRY(hipMemcpy(d_stdp, sl.data(), sizeof(uint2) * sl.size(), hipMemcpyHostToDevice));
        }
        n_fix = static_cast<uint32_t>(fl.size());
        if (n_fix) {
            if ((rc = grow(&d_fix, &cap_fix, n_fix))) return rc;
            HIP_TRY(hipMemcpy(d_fix, fl.data(), sizeof(uint2) * fl.size(), hipMemcpyHostToDevice));
        }
        xcd_order(binl, make_uint4(kEmpty, 0u, 0u, 0u));
        n_bin = static_cast<uint32_t>(binl.size());
        if (n_bin) {
            if ((rc = grow(&d_bin, &cap_bin, n_bin))) return rc;
            HIP_TRY(hipMemcpy(d_bin, binl.data(), sizeof(uint4) * binl.size(), hipMemcpyHostToDevice));
        }
        n_cpl = static_cast<uint32_t>(cpl.size());
        if (n_cpl) {
            if ((rc = grow(&d_cpl, &cap_cpl, n_cpl))) return rc;
            HIP_TRY(hipMemcpy(d_cpl, cpl.data(), sizeof(uint2) * cpl.size(), hipMemcpyHostToDevice));
        }
        xcd_order(apl, make_uint2(kEmpty, 0u));
        n_apply = static_cast<uint32_t>(apl.size());
        if (n_apply) {
            if ((rc = grow(&d_apply, &cap_apply, n_apply))) return rc;
            HIP_TRY(hipMemcpy(d_apply, apl.data(), sizeof(uint2) * apl.size(), hipMemcpyHostToDevice));
        }
        xcd_order(aplr, make_uint2(kEmpty, 0u));
        n_apply_rec = static_cast<uint32_t>(aplr.size());
        if (n_apply_rec) {
            if ((rc = grow(&d_apply_rec, &cap_apply_rec, n_apply_rec))) return rc;
            HIP_TRY(hipMemcpy(d_apply_rec, aplr.data(), sizeof(uint2) * aplr.size(), hipMemcpyHostToDevice));
        }
        drop_graph();
        dirty = false;
        return SIMUSE_OK;
    }

    void launch_step(hipEvent_t* e) {
        if (e) (void)hipEventRecord(e[0], stream);
        launch_coupling(d_nets, d_cpl, n_cpl, stream);
        for (int m = 0; m < kModels; ++m) launch_megastep(m, d_nets, d_mega[m], n_mega_m[m], stream);
        for (int m = 0; m < kModels; ++m)
            for (int r = 0; r < 2; ++r) launch_front(m, r != 0, d_nets, d_front[2 * m + r], n_front_m[2 * m + r], stream);
        for (int m = 0; m < kModels; ++m)
            launch_neuron(m, sched, d_nets, d_tiles[m], n_tiles_m[m], stream, 0u, bins_fused);
        if (e) (void)hipEventRecord(e[1], stream);
        launch_stdp(d_nets, d_stdp, n_stdp, stream);
        if (e) (void)hipEventRecord(e[2], stream);
        launch_insert(sched, d_nets, d_blocks, n_blocks, stream);
        launch_bin(d_nets, d_bin, n_bin, bins_bin, stream);
        if (e) (void)hipEventRecord(e[3], stream);
        launch_bin_apply(d_nets, d_apply, n_apply, 0u, stream);
        launch_bin_apply_rec(d_nets, d_apply_rec, n_apply_rec, stream);
        launch_fixup(d_nets, direct_fix ? nullptr : d_fix, n_fix, stream);
        if (e) (void)hipEventRecord(e[4], stream);
    }

    int enqueue_steps(uint32_t n) {
        for (uint32_t k = 0; k < n; ++k) launch_step(nullptr);
        HIP_TRY(hipGetLastError());
        return SIMUSE_OK;
    }

    int build_graph() {
        drop_graph();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_steps(graph_steps);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc != SIMUSE_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(SIMUSE_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        graph = g;
        e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph();
            return fail(SIMUSE_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        return SIMUSE_OK;
    }

    // A megastep or split-form network keeps its pending events (the last
    // step's records) out of the ring; apply them into ring slot step % 32
    // (and receptor jumps into the conductances) before anything reads or
    // adds to that slot from the host side.  Idempotent.
    int materialize(NetState* s) {
        int rc;
        if ((rc = ensure_tables())) return rc;
        if (!s->finalized || !(s->mega || s->split)) return SIMUSE_OK;
        if (s->mat_dirty) {
            if (s->mat_items.size() > static_cast<size_t>(s->n_bins) + kXcdPad)
                return fail(SIMUSE_E_STATE, "materialize: %zu work items exceed the list's room", s->mat_items.size());
            HIP_TRY(hipMemcpy(s->d_mat_items, s->mat_items.data(), sizeof(uint2) * s->mat_items.size(),
                              hipMemcpyHostToDevice));
            s->mat_dirty = false;
        }
        if (s->ev_wide) launch_bin_apply_rec_mode(d_nets, s->d_mat_items, static_cast<uint32_t>(s->mat_items.size()), 1u, stream);
        else launch_bin_apply(d_nets, s->d_mat_items, static_cast<uint32_t>(s->mat_items.size()), 1u, stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(s->bin_mat), 1, 1, stream));
        return SIMUSE_OK;
    }
    // Move a megastep network to the two-kernel path for good (and, with
    // split, a split-form network to the three-kernel path).  The split
    // form's writers fill whole header columns, which that path would read
    // as events: both parities are cleared.
    int leave_megastep(NetState* s, bool split = false) {
        int rc;
        if ((rc = materialize(s))) return rc;
        if (s->mega_ok && !s->no_mega) {
            s->no_mega = true;
            mark_dirty();
        }
        if (split && s->split_ok && !s->no_split) {
            s->no_split = true;
            if (s->bin_hdr)
                HIP_TRY(hipMemsetAsync(s->bin_hdr, 0, sizeof(uint2) * s->nbuf() * s->n_bins * s->bin_rows, stream));
            mark_dirty();
        }
        return SIMUSE_OK;
    }

    int prepare() {
        int rc;
        for (NetState* s : nets)
            if (!s->finalized) {
                if ((rc = s->finalize(stream))) return rc;
                dirty = true;
            }
        return ensure_tables();
    }

    // Enqueue n steps of every network of the set.
    int step_async(uint32_t n) {
        int rc = prepare();
        if (rc) return rc;
        if (n == 0 || nets.empty()) return SIMUSE_OK;
        while (n > (1u << 30)) {
            if ((rc = step_async(1u << 30))) return rc;
            n -= 1u << 30;
        }
        // Scatter heads carry the low 32 bits of the step as a tag; rewrite
        // them before 2^31 steps pass so that no stale tag comes round again.
        for (NetState* s : nets) {
            if (s->since_clear + n >= (1ull << 31)) s->clear_heads(stream);
            s->since_clear += n;
        }
        if (profiling) {
            const size_t need = static_cast<size_t>(prof_steps + n) * (kPhases + 1);
            while (ev.size() < need) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                ev.push_back(e);
            }
            for (uint32_t k = 0; k < n; ++k)
                launch_step(ev.data() + static_cast<size_t>(prof_steps + k) * (kPhases + 1));
            prof_steps += n;
            HIP_TRY(hipGetLastError());
        } else {
            const uint32_t full = n / graph_steps, rest = n % graph_steps;
            if (full) {
                if (!gexec && (rc = build_graph())) return rc;
                for (uint32_t k = 0; k < full; ++k) HIP_TRY(hipGraphLaunch(gexec, stream));
            }
            if (rest && (rc = enqueue_steps(rest))) return rc;
        }
        for (NetState* s : nets) s->host_step = std::min(s->host_step + n, std::max(s->host_step, s->step_limit));
        return SIMUSE_OK;
    }

    int sync() {
        HIP_TRY(hipStreamSynchronize(stream));
        return SIMUSE_OK;
    }

    int set_profiling(bool on) {
        HIP_TRY(hipStreamSynchronize(stream));
        profiling = on;
        prof_steps = 0;
        return SIMUSE_OK;
    }

    int kernel_times(float* ms, int32_t n) {
        if (!ms || n < 1) return fail(SIMUSE_E_INVALID, "null output");
        HIP_TRY(hipStreamSynchronize(stream));
        double acc[kPhases] = {0, 0, 0, 0};
        for (uint32_t k = 0; k < prof_steps; ++k) {
            hipEvent_t* e = ev.data() + static_cast<size_t>(k) * (kPhases + 1);
            for (int p = 0; p < kPhases; ++p) {
                float t = 0;
                HIP_TRY(hipEventElapsedTime(&t, e[p], e[p + 1]));
                acc[p] += t;
            }
        }
        for (int p = 0; p < n && p < kPhases; ++p)
            ms[p] = prof_steps ? static_cast<float>(acc[p] / prof_steps) : 0.0f;
        prof_steps = 0;
        return SIMUSE_OK;
    }

    int set_sched(int mode) {
        if (mode != SIMUSE_SCATTER_TILE && mode != SIMUSE_SCATTER_WAVE)
            return fail(SIMUSE_E_INVALID, "unknown scatter mode %d", mode);
        if (sched != mode) {
            sched = mode;
            drop_graph();
        }
        return SIMUSE_OK;
    }

    int set_group(uint32_t g) {
        if (g < 1 || g > kMaxGroup) return fail(SIMUSE_E_INVALID, "tiles per block must be 1..%u", kMaxGroup);
        if (group != g) {
            group = g;
            mark_dirty();
        }
        return SIMUSE_OK;
    }

    // Networks that can take the one-kernel step take it (on) or the binned
    // form (off).  Either way a network's next input ends up in ring slot
    // step % 32 with the materialised flag set: a one-kernel network's
    // pending events are applied there, and a binned network's apply has
    // already written them there.
    int set_one_kernel(bool on) {
        int rc = prepare();
        if (rc) return rc;
        if (one_kernel == on) return SIMUSE_OK;
        for (NetState* s : nets) {
            if ((rc = materialize(s))) return rc;
            if (s->bin_mat && s->binned && s->mega_ok && !s->no_mega && !s->stdp && !s->mega)
                HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(s->bin_mat), 1, 1, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
        one_kernel = on;
        mark_dirty();
        return SIMUSE_OK;
    }

    int set_graph_steps(uint32_t steps) {
        if (steps < 1 || steps > 64) return fail(SIMUSE_E_INVALID, "graph steps must be 1..64");
        if (graph_steps != steps) {
            graph_steps = steps;
            drop_graph();
        }
        return SIMUSE_OK;
    }

    int autotune(uint32_t trial_steps);
};

}  // namespace simuse
