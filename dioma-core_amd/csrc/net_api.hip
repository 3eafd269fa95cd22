// net_api.hip — host runtime of the simuse_net_* C ABI (include/simuse.h).
//
// A handle owns one HIP stream and all device buffers of one network.  Steps
// are captured once into a hipGraph of `graph_steps` steps and replayed; the
// ring slot and step counter live on the device and advance inside the graph,
// so asynchronous steps need no host round trip (the reference patches a
// host-written slot scalar before every replay, PC/src/network.rs:431-460).

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
#include "net_kernels.hpp"

namespace simuse {
void launch_neuron(int model, const NetArgs& a, const uint32_t* tiles, uint32_t n, hipStream_t s);
void launch_stdp(const NetArgs& a, hipStream_t s);
void launch_insert(int sched, const NetArgs& a, hipStream_t s);
void launch_fold(int sched, const NetArgs& a, hipStream_t s);
void launch_inject(float* ring, const uint64_t* step, uint32_t N, const uint32_t* idx, const float* val,
                   uint32_t count, int serial, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_widen(const float* w, double* w64, size_t n, hipStream_t s);
void launch_raster_stats(const uint64_t* raster, uint32_t n_words, uint32_t steps, const uint32_t* word_of,
                         const uint32_t* bit_of, uint32_t count, float dt_s, float* rates, float* cvs,
                         hipStream_t s);

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::simuse::fail(SIMUSE_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int model_states(int m) { return m == SIMUSE_MODEL_LIF ? 2 : m == SIMUSE_MODEL_IZHIKEVICH ? 3 : 6; }
int model_params(int m) { return m == SIMUSE_MODEL_LIF ? 6 : m == SIMUSE_MODEL_IZHIKEVICH ? 7 : 13; }
// column index of tau_ref and of the refractory state per model
int model_tau_ref_col(int m) { return m == SIMUSE_MODEL_LIF ? 5 : m == SIMUSE_MODEL_IZHIKEVICH ? 5 : -1; }
int model_refr_state(int m) { return m == SIMUSE_MODEL_LIF ? 1 : m == SIMUSE_MODEL_IZHIKEVICH ? 2 : -1; }

}  // namespace simuse

using namespace simuse;

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
    uint64_t seed = 0;
};

struct SourceHost {
    float p;
    uint32_t seed;
    std::vector<uint32_t> post;
    std::vector<float> w;
    std::vector<uint8_t> d;
};

struct simuse_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostPop> pops;
    uint32_t N = 0, n_tiles = 0, n_words = 0;
    uint64_t n_syn = 0;
    std::vector<void*> allocs;
    uint64_t bytes = 0;
    // device buffers
    PopDev* d_pops = nullptr;
    uint32_t* d_tile_pop = nullptr;
    uint32_t* d_model_tiles[3] = {nullptr, nullptr, nullptr};
    uint32_t n_model_tiles[3] = {0, 0, 0};
    float* ring = nullptr;
    uint32_t* head = nullptr;
    uint32_t* next = nullptr;
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
    uint32_t* done = nullptr;
    // sources
    std::vector<SourceHost> srcs;
    uint32_t n_src = 0, n_src_tiles = 0;
    uint64_t n_src_edges = 0;
    uint32_t* src_row_ptr = nullptr;
    uint32_t* src_post = nullptr;
    float* src_w = nullptr;
    uint8_t* src_delay = nullptr;
    float* src_p = nullptr;
    uint32_t* src_seed = nullptr;
    // raster
    uint64_t* raster = nullptr;
    uint32_t raster_cap = 0;
    uint64_t raster_start = 0;
    // stdp
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
    // execution
    int sched = SIMUSE_SCATTER_TILE;
    bool finalized = false;
    uint64_t host_step = 0;
    uint32_t graph_steps = 8;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t* inj_idx = nullptr;
    float* inj_val = nullptr;
    uint32_t inj_cap = 0;
    // host copy of the CSR row_ptr (degree statistics for the autotuner)
    std::vector<uint32_t> h_row_ptr;

    template <class T>
    int alloc(T** p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        allocs.push_back(q);
        bytes += count * sizeof(T);
        *p = static_cast<T*>(q);
        return SIMUSE_OK;
    }

    void drop_graph() {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }

    ~simuse_net() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        drop_graph();
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    NetArgs args() const {
        NetArgs a{};
        a.pops = d_pops;
        a.tile_pop = d_tile_pop;
        a.N = N;
        a.n_tiles = n_tiles;
        a.n_src_tiles = n_src_tiles;
        a.n_src = n_src;
        a.n_syn = n_syn;
        a.ring = ring;
        a.head = head;
        a.next = next;
        a.row_ptr = row_ptr;
        a.col = col;
        a.w = w;
        a.delay = delay;
        a.src_row_ptr = src_row_ptr;
        a.src_post = src_post;
        a.src_w = src_w;
        a.src_delay = src_delay;
        a.src_p = src_p;
        a.src_seed = src_seed;
        a.tile_count = tile_count;
        a.active = active;
        a.bits = bits;
        a.tile_spikes = tile_spikes;
        a.tile_events = tile_events;
        a.step = step;
        a.done = done;
        a.raster = raster;
        a.raster_cap = raster_cap;
        a.n_words = n_words;
        a.raster_base = step + 1;
        a.stdp = stdp ? 1 : 0;
        a.w64 = w64;
        a.tr_pre = tr_pre;
        a.tr_post = tr_post;
        a.in_ptr = in_ptr;
        a.in_edge = in_edge;
        a.edge_pre = edge_pre;
        a.plastic = plastic;
        a.stdp_dpre = stdp_dpre;
        a.stdp_dpost = stdp_dpost;
        a.a_plus = sp.a_plus;
        a.a_minus = sp.a_minus;
        a.w_min = sp.w_min;
        a.w_max = sp.w_max;
        return a;
    }

    int enqueue_steps(uint32_t n) {
        const NetArgs a = args();
        for (uint32_t k = 0; k < n; ++k) {
            for (int m = 0; m < 3; ++m) launch_neuron(m, a, d_model_tiles[m], n_model_tiles[m], stream);
            if (stdp) launch_stdp(a, stream);
            launch_insert(sched, a, stream);
            launch_fold(sched, a, stream);
        }
        HIP_TRY(hipGetLastError());
        return SIMUSE_OK;
    }

    // Profiled steps: direct launches bracketed by HIP events on the
    // handle's stream; durations are read back by kernel_times().
    static constexpr int kPhases = 4;  // neuron, stdp, insert, fold
    bool profiling = false;
    std::vector<hipEvent_t> ev;       // (kPhases+1) events per profiled step
    uint32_t prof_steps = 0;
    int ensure_events(uint32_t steps) {
        const size_t need = static_cast<size_t>(steps) * (kPhases + 1);
        while (ev.size() < need) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        return SIMUSE_OK;
    }
    int enqueue_profiled(uint32_t n) {
        int rc = ensure_events(prof_steps + n);
        if (rc) return rc;
        const NetArgs a = args();
        for (uint32_t k = 0; k < n; ++k) {
            hipEvent_t* e = ev.data() + static_cast<size_t>(prof_steps + k) * (kPhases + 1);
            HIP_TRY(hipEventRecord(e[0], stream));
            for (int m = 0; m < 3; ++m) launch_neuron(m, a, d_model_tiles[m], n_model_tiles[m], stream);
            HIP_TRY(hipEventRecord(e[1], stream));
            if (stdp) launch_stdp(a, stream);
            HIP_TRY(hipEventRecord(e[2], stream));
            launch_insert(sched, a, stream);
            HIP_TRY(hipEventRecord(e[3], stream));
            launch_fold(sched, a, stream);
            HIP_TRY(hipEventRecord(e[4], stream));
        }
        prof_steps += n;
        HIP_TRY(hipGetLastError());
        return SIMUSE_OK;
    }

    int build_graph() {
        drop_graph();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_steps(graph_steps);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc != SIMUSE_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(SIMUSE_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        graph = g;
        e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) { drop_graph(); return fail(SIMUSE_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e)); }
        return SIMUSE_OK;
    }

    // Upload Poisson sources and size the chain links; runs before the first step.
    int finalize() {
        if (finalized) return SIMUSE_OK;
        n_src = static_cast<uint32_t>(srcs.size());
        n_src_tiles = (n_src + kTile - 1) / kTile;
        std::vector<uint32_t> rp(n_src + 1, 0), post;
        std::vector<float> sw, sprob;
        std::vector<uint8_t> sd;
        std::vector<uint32_t> sseed;
        for (uint32_t s = 0; s < n_src; ++s) {
            rp[s + 1] = rp[s] + static_cast<uint32_t>(srcs[s].post.size());
            post.insert(post.end(), srcs[s].post.begin(), srcs[s].post.end());
            sw.insert(sw.end(), srcs[s].w.begin(), srcs[s].w.end());
            sd.insert(sd.end(), srcs[s].d.begin(), srcs[s].d.end());
            sprob.push_back(srcs[s].p);
            sseed.push_back(srcs[s].seed);
        }
        n_src_edges = rp[n_src];
        if (n_syn + n_src_edges >= 0xFFFFFFFFull)
            return fail(SIMUSE_E_INVALID, "edges + source edges (%llu) exceed the u32 event id range",
                        static_cast<unsigned long long>(n_syn + n_src_edges));
        int rc;
        if ((rc = alloc(&next, n_syn + n_src_edges))) return rc;
        if ((rc = alloc(&tile_events, n_tiles + n_src_tiles))) return rc;
        HIP_TRY(hipMemsetAsync(tile_events, 0, sizeof(uint64_t) * (n_tiles + n_src_tiles), stream));
        if (n_src) {
            if ((rc = alloc(&src_row_ptr, n_src + 1))) return rc;
            if ((rc = alloc(&src_post, n_src_edges))) return rc;
            if ((rc = alloc(&src_w, n_src_edges))) return rc;
            if ((rc = alloc(&src_delay, n_src_edges))) return rc;
            if ((rc = alloc(&src_p, n_src))) return rc;
            if ((rc = alloc(&src_seed, n_src))) return rc;
            HIP_TRY(hipMemcpy(src_row_ptr, rp.data(), 4 * rp.size(), hipMemcpyHostToDevice));
            if (n_src_edges) {
                HIP_TRY(hipMemcpy(src_post, post.data(), 4 * post.size(), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(src_w, sw.data(), 4 * sw.size(), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(src_delay, sd.data(), sd.size(), hipMemcpyHostToDevice));
            }
            HIP_TRY(hipMemcpy(src_p, sprob.data(), 4 * sprob.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(src_seed, sseed.data(), 4 * sseed.size(), hipMemcpyHostToDevice));
        }
        finalized = true;
        return SIMUSE_OK;
    }

    int upload_pops(const simuse_pop_desc* pd) {
        for (size_t p = 0; p < pops.size(); ++p) {
            HostPop& P = pops[p];
            const uint32_t hs = pd[p].stride;
            if (P.n == 0) continue;
            HIP_TRY(hipMemcpy2D(P.st, sizeof(float) * P.dstride, pd[p].states, sizeof(float) * hs,
                                sizeof(float) * P.n, P.ns, hipMemcpyHostToDevice));
            // Uniform parameter columns become scalars; column 0 (dt) is
            // always taken from neuron 0 (population.rs:307,361).
            P.uni_mask = 0;
            for (uint32_t k = 0; k < P.np && k < static_cast<uint32_t>(kMaxParams); ++k) {
                const float* c = pd[p].params + static_cast<size_t>(k) * hs;
                bool uni = true;
                for (uint32_t i = 1; i < P.n && uni; ++i)
                    uni = std::memcmp(&c[i], &c[0], 4) == 0;
                P.puni[k] = c[0];
                if (uni || k == 0) P.uni_mask |= 1u << k;
            }
            HIP_TRY(hipMemcpy2D(P.pr, sizeof(float) * P.dstride, pd[p].params, sizeof(float) * hs,
                                sizeof(float) * P.n, P.np, hipMemcpyHostToDevice));
            // Refractory elision: with tau_ref uniformly <= 0, dt > 0 and every
            // timer exactly +0, refr = max(spike ? tau_ref : refr - dt, 0) stays 0.
            P.flags &= ~kPopRefrElided;
            const int tc = model_tau_ref_col(P.model), rs = model_refr_state(P.model);
            if (tc >= 0 && ((P.uni_mask >> tc) & 1u) && P.puni[tc] <= 0.0f && P.puni[0] > 0.0f) {
                const float* r = pd[p].states + static_cast<size_t>(rs) * hs;
                bool zero = true;
                for (uint32_t i = 0; i < P.n && zero; ++i) {
                    uint32_t b;
                    std::memcpy(&b, &r[i], 4);
                    zero = (b == 0u);
                }
                if (zero) P.flags |= kPopRefrElided;
            }
        }
        return upload_pop_table();
    }

    int upload_pop_table() {
        std::vector<PopDev> pd(pops.size());
        for (size_t p = 0; p < pops.size(); ++p) {
            const HostPop& P = pops[p];
            PopDev& D = pd[p];
            std::memset(&D, 0, sizeof(D));
            D.model = P.model; D.n = P.n; D.off = P.off; D.tile0 = P.tile0; D.word0 = P.word0;
            D.dstride = P.dstride; D.uni_mask = P.uni_mask; D.flags = P.flags;
            D.st = P.st; D.pr = P.pr; D.sigma = P.sigma; D.sigma_uni = P.sigma_uni; D.seed = P.seed;
            std::memcpy(D.puni, P.puni, sizeof(D.puni));
        }
        HIP_TRY(hipMemcpy(d_pops, pd.data(), sizeof(PopDev) * pd.size(), hipMemcpyHostToDevice));
        return SIMUSE_OK;
    }

    int reset_dynamic() {
        HIP_TRY(hipMemsetAsync(ring, 0, sizeof(float) * static_cast<size_t>(kSlots) * N, stream));
        launch_fill_u32(head, kEmpty, static_cast<size_t>(kSlots) * N, stream);
        HIP_TRY(hipMemsetAsync(step, 0, sizeof(uint64_t) * 2, stream));
        HIP_TRY(hipMemsetAsync(done, 0, sizeof(uint32_t), stream));
        HIP_TRY(hipMemsetAsync(tile_count, 0, sizeof(uint32_t) * n_tiles, stream));
        HIP_TRY(hipMemsetAsync(bits, 0, sizeof(uint64_t) * n_words, stream));
        HIP_TRY(hipMemsetAsync(tile_spikes, 0, sizeof(uint64_t) * n_tiles, stream));
        if (tile_events) HIP_TRY(hipMemsetAsync(tile_events, 0, sizeof(uint64_t) * (n_tiles + n_src_tiles), stream));
        if (stdp) {
            HIP_TRY(hipMemsetAsync(tr_pre, 0, sizeof(float) * N, stream));
            HIP_TRY(hipMemsetAsync(tr_post, 0, sizeof(float) * N, stream));
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream));
        host_step = 0;
        return SIMUSE_OK;
    }
};

static int check_net(simuse_net* net) {
    if (!net) return fail(SIMUSE_E_INVALID, "null network handle");
    return SIMUSE_OK;
}

static int validate_pops(const simuse_pop_desc* pops, int32_t n_pops, uint64_t* n_total) {
    if (n_pops <= 0 || !pops) return fail(SIMUSE_E_INVALID, "no populations");
    uint64_t tot = 0;
    for (int32_t p = 0; p < n_pops; ++p) {
        const simuse_pop_desc& d = pops[p];
        if (d.model < SIMUSE_MODEL_LIF || d.model > SIMUSE_MODEL_HH)
            return fail(SIMUSE_E_UNSUPPORTED, "population %d: unknown model %d", p, d.model);
        if (d.ns != static_cast<uint32_t>(model_states(d.model)) || d.np != static_cast<uint32_t>(model_params(d.model)))
            return fail(SIMUSE_E_INVALID, "population %d: model %d needs %d states and %d params (got %u, %u)", p,
                        d.model, model_states(d.model), model_params(d.model), d.ns, d.np);
        if (d.stride < d.n) return fail(SIMUSE_E_INVALID, "population %d: stride %u < n %u", p, d.stride, d.n);
        if (d.n > 0 && (!d.states || !d.params)) return fail(SIMUSE_E_INVALID, "population %d: null state/params", p);
        tot += d.n;
    }
    if (tot == 0) return fail(SIMUSE_E_INVALID, "CompiledNetwork: no neurons");
    if (tot >= 0xFFFFFFFFull) return fail(SIMUSE_E_INVALID, "total neuron count %llu exceeds u32",
                                          static_cast<unsigned long long>(tot));
    *n_total = tot;
    return SIMUSE_OK;
}

extern "C" {

const char* simuse_last_error(void) { return g_last_error.c_str(); }

int simuse_device_count(int32_t* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) { (void)hipGetLastError(); c = 0; }
    if (count) *count = c;
    return SIMUSE_OK;
}

int simuse_net_create(const simuse_pop_desc* pops, int32_t n_pops, const uint32_t* row_ptr, const uint32_t* col,
                      const float* w, const uint8_t* delay, uint64_t n_syn, float dt, simuse_net** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "null output handle");
    *out = nullptr;
    uint64_t n_total = 0;
    int rc = validate_pops(pops, n_pops, &n_total);
    if (rc) return rc;
    if (!row_ptr) return fail(SIMUSE_E_INVALID, "null row_ptr");
    if (n_syn && (!col || !w || !delay)) return fail(SIMUSE_E_INVALID, "null CSR arrays");
    if (n_syn >= 0xFFFFFFFFull) return fail(SIMUSE_E_INVALID, "too many synapses (%llu) for a u32 CSR",
                                           static_cast<unsigned long long>(n_syn));
    // SparseGraph::from_csr checks (P/src/api/net/graph.rs:217-280)
    if (row_ptr[0] != 0) return fail(SIMUSE_E_INVALID, "CSR row_ptr must start at 0");
    for (uint64_t i = 0; i < n_total; ++i)
        if (row_ptr[i] > row_ptr[i + 1])
            return fail(SIMUSE_E_INVALID, "CSR row_ptr must be monotonically non-decreasing: %u > %u at %llu",
                        row_ptr[i], row_ptr[i + 1], static_cast<unsigned long long>(i));
    if (row_ptr[n_total] != n_syn) return fail(SIMUSE_E_INVALID, "CSR row_ptr must end at col_idx.len()");
    for (uint64_t e = 0; e < n_syn; ++e) {
        if (col[e] >= n_total)
            return fail(SIMUSE_E_INVALID, "col index %u >= %llu", col[e], static_cast<unsigned long long>(n_total));
        if (!std::isfinite(w[e])) return fail(SIMUSE_E_INVALID, "weight %llu must be finite", static_cast<unsigned long long>(e));
        if (delay[e] >= kSlots) return fail(SIMUSE_E_INVALID, "delay %u >= MAX_DELAY_SLOTS (32)", delay[e]);
    }
    for (int32_t p = 0; p < n_pops; ++p) {
        if (pops[p].n == 0) continue;
        const float pdt = pops[p].params[0];
        if (dt > 0.0f && pdt != dt)
            return fail(SIMUSE_E_INVALID, "population %d: param dt %g differs from network dt %g", p, pdt, dt);
    }
    int dev = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "no HIP device available");
    }
    HIP_TRY(hipGetDevice(&dev));

    std::unique_ptr<simuse_net> net(new (std::nothrow) simuse_net());
    if (!net) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    net->device = dev;
    // A blocking stream is ordered after the synchronous (null-stream)
    // uploads the API performs.
    HIP_TRY(hipStreamCreate(&net->stream));
    net->N = static_cast<uint32_t>(n_total);
    net->n_syn = n_syn;
    uint32_t off = 0, tile = 0, word = 0;
    for (int32_t p = 0; p < n_pops; ++p) {
        HostPop P;
        P.model = pops[p].model;
        P.n = pops[p].n;
        P.ns = pops[p].ns;
        P.np = pops[p].np;
        P.host_stride = pops[p].stride;
        P.off = off;
        P.tile0 = tile;
        P.ntiles = (P.n + kTile - 1) / kTile;
        P.word0 = word;
        P.dstride = (P.n + 63u) / 64u * 64u;
        if (P.dstride == 0) P.dstride = 64;
        off += P.n;
        tile += P.ntiles;
        word += P.ntiles * kWavesPerTile;
        net->pops.push_back(P);
    }
    net->n_tiles = tile;
    net->n_words = word;
    simuse_net* n = net.get();
    for (auto& P : n->pops) {
        if ((rc = n->alloc(&P.st, static_cast<size_t>(P.ns) * P.dstride))) return rc;
        if ((rc = n->alloc(&P.pr, static_cast<size_t>(P.np) * P.dstride))) return rc;
        HIP_TRY(hipMemset(P.st, 0, sizeof(float) * P.ns * P.dstride));
        HIP_TRY(hipMemset(P.pr, 0, sizeof(float) * P.np * P.dstride));
    }
    if ((rc = n->alloc(&n->d_pops, n->pops.size()))) return rc;
    if ((rc = n->upload_pops(pops))) return rc;
    // tile tables
    std::vector<uint32_t> tile_pop(n->n_tiles);
    std::vector<uint32_t> mt[3];
    for (size_t p = 0; p < n->pops.size(); ++p)
        for (uint32_t t = 0; t < n->pops[p].ntiles; ++t) {
            tile_pop[n->pops[p].tile0 + t] = static_cast<uint32_t>(p);
            mt[n->pops[p].model].push_back(n->pops[p].tile0 + t);
        }
    if ((rc = n->alloc(&n->d_tile_pop, n->n_tiles))) return rc;
    HIP_TRY(hipMemcpy(n->d_tile_pop, tile_pop.data(), 4 * tile_pop.size(), hipMemcpyHostToDevice));
    for (int m = 0; m < 3; ++m) {
        n->n_model_tiles[m] = static_cast<uint32_t>(mt[m].size());
        if (mt[m].empty()) continue;
        if ((rc = n->alloc(&n->d_model_tiles[m], mt[m].size()))) return rc;
        HIP_TRY(hipMemcpy(n->d_model_tiles[m], mt[m].data(), 4 * mt[m].size(), hipMemcpyHostToDevice));
    }
    // CSR
    if ((rc = n->alloc(&n->row_ptr, n_total + 1))) return rc;
    if ((rc = n->alloc(&n->col, n_syn))) return rc;
    if ((rc = n->alloc(&n->w, n_syn))) return rc;
    if ((rc = n->alloc(&n->delay, n_syn))) return rc;
    HIP_TRY(hipMemcpy(n->row_ptr, row_ptr, 4 * (n_total + 1), hipMemcpyHostToDevice));
    if (n_syn) {
        HIP_TRY(hipMemcpy(n->col, col, 4 * n_syn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(n->w, w, 4 * n_syn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(n->delay, delay, n_syn, hipMemcpyHostToDevice));
    }
    n->h_row_ptr.assign(row_ptr, row_ptr + n_total + 1);
    // ring, chains, compaction
    if ((rc = n->alloc(&n->ring, static_cast<size_t>(kSlots) * n_total))) return rc;
    if ((rc = n->alloc(&n->head, static_cast<size_t>(kSlots) * n_total))) return rc;
    if ((rc = n->alloc(&n->tile_count, n->n_tiles))) return rc;
    if ((rc = n->alloc(&n->active, static_cast<size_t>(n->n_tiles) * kTile))) return rc;
    if ((rc = n->alloc(&n->bits, n->n_words))) return rc;
    if ((rc = n->alloc(&n->tile_spikes, n->n_tiles))) return rc;
    if ((rc = n->alloc(&n->step, 2))) return rc;
    if ((rc = n->alloc(&n->done, 1))) return rc;
    if ((rc = n->reset_dynamic())) return rc;
    HIP_TRY(hipDeviceSynchronize());
    *out = net.release();
    return SIMUSE_OK;
}

void simuse_net_destroy(simuse_net* net) { delete net; }

int simuse_net_add_poisson_source(simuse_net* net, float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts,
                                  const float* w, const uint8_t* delay, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    if (net->finalized) return fail(SIMUSE_E_STATE, "add_poisson_source must be called before the first step");
    if (!(std::isfinite(rate_hz) && rate_hz >= 0.0f))
        return fail(SIMUSE_E_INVALID, "add_poisson_source: rate_hz must be finite and >= 0, got %g", rate_hz);
    if (!(std::isfinite(dt_ms) && dt_ms > 0.0f))
        return fail(SIMUSE_E_INVALID, "add_poisson_source: dt_ms must be finite and > 0, got %g", dt_ms);
    if (count && (!posts || !w || !delay)) return fail(SIMUSE_E_INVALID, "null source arrays");
    SourceHost s;
    s.p = std::min(std::max(rate_hz * dt_ms / 1000.0f, 0.0f), 1.0f);  // core.rs:301
    s.seed = seed;
    for (uint32_t i = 0; i < count; ++i) {
        if (posts[i] >= net->N) return fail(SIMUSE_E_INVALID, "post index %u out of bounds", posts[i]);
        if (!std::isfinite(w[i])) return fail(SIMUSE_E_INVALID, "connect_poisson_source: weight %u must be finite", i);
        if (delay[i] >= kSlots) return fail(SIMUSE_E_INVALID, "connect_poisson_source: delay %u >= MAX_DELAY_SLOTS (32)", delay[i]);
    }
    s.post.assign(posts, posts + count);
    s.w.assign(w, w + count);
    s.d.assign(delay, delay + count);
    net->srcs.push_back(std::move(s));
    return SIMUSE_OK;
}

int simuse_net_set_noise(simuse_net* net, int32_t pop, const float* sigma, uint64_t seed) {
    int rc = check_net(net);
    if (rc) return rc;
    if (pop < 0 || pop >= static_cast<int32_t>(net->pops.size())) return fail(SIMUSE_E_INVALID, "pop %d out of range", pop);
    HostPop& P = net->pops[pop];
    if (!sigma) {
        P.flags &= ~kPopNoise;
    } else {
        for (uint32_t i = 0; i < P.n; ++i)
            if (!std::isfinite(sigma[i])) return fail(SIMUSE_E_INVALID, "noise sigma %u must be finite", i);
        bool uni = true;
        for (uint32_t i = 1; i < P.n && uni; ++i) uni = std::memcmp(&sigma[i], &sigma[0], 4) == 0;
        P.flags &= ~kPopNoiseScalar;
        if (uni) {
            P.sigma_uni = P.n ? sigma[0] : 0.0f;
            P.flags |= kPopNoiseScalar;
        } else {
            if (!P.sigma && (rc = net->alloc(&P.sigma, P.n))) return rc;
            HIP_TRY(hipMemcpy(P.sigma, sigma, 4 * static_cast<size_t>(P.n), hipMemcpyHostToDevice));
        }
        P.seed = seed;
        P.flags |= kPopNoise;
    }
    net->drop_graph();
    return net->upload_pop_table();
}

int simuse_net_set_plasticity(simuse_net* net, const simuse_stdp_params* p, const uint8_t* plastic) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!p) {
        net->stdp = false;
        net->drop_graph();
        return SIMUSE_OK;
    }
    if (!(p->tau_plus > 0 && p->tau_minus > 0 && std::isfinite(p->a_plus) && std::isfinite(p->a_minus) &&
          p->w_min <= p->w_max))
        return fail(SIMUSE_E_INVALID, "invalid STDP parameters");
    const uint32_t N = net->N;
    const uint64_t ns = net->n_syn;
    if (!net->w64) {
        if ((rc = net->alloc(&net->w64, ns))) return rc;
        if ((rc = net->alloc(&net->tr_pre, N))) return rc;
        if ((rc = net->alloc(&net->tr_post, N))) return rc;
        if ((rc = net->alloc(&net->in_ptr, N + 1))) return rc;
        if ((rc = net->alloc(&net->in_edge, ns))) return rc;
        if ((rc = net->alloc(&net->edge_pre, ns))) return rc;
        // in-edge lists (ascending edge id per post) and edge -> pre
        std::vector<uint32_t> colh(ns);
        if (ns) HIP_TRY(hipMemcpy(colh.data(), net->col, 4 * ns, hipMemcpyDeviceToHost));
        const auto& rp = net->h_row_ptr;
        std::vector<uint32_t> ip(N + 1, 0), ie(ns), ep(ns);
        for (uint64_t e = 0; e < ns; ++e) ip[colh[e] + 1]++;
        for (uint32_t i = 0; i < N; ++i) ip[i + 1] += ip[i];
        std::vector<uint32_t> cur(ip.begin(), ip.end() - 1);
        for (uint32_t i = 0; i < N; ++i)
            for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) ep[e] = i;
        for (uint64_t e = 0; e < ns; ++e) ie[cur[colh[e]]++] = static_cast<uint32_t>(e);
        HIP_TRY(hipMemcpy(net->in_ptr, ip.data(), 4 * ip.size(), hipMemcpyHostToDevice));
        if (ns) {
            HIP_TRY(hipMemcpy(net->in_edge, ie.data(), 4 * ns, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(net->edge_pre, ep.data(), 4 * ns, hipMemcpyHostToDevice));
        }
    }
    if (plastic) {
        if (!net->plastic && (rc = net->alloc(&net->plastic, ns))) return rc;
        if (ns) HIP_TRY(hipMemcpy(net->plastic, plastic, ns, hipMemcpyHostToDevice));
    } else if (net->plastic) {
        std::vector<uint8_t> ones(ns, 1);
        if (ns) HIP_TRY(hipMemcpy(net->plastic, ones.data(), ns, hipMemcpyHostToDevice));
    }
    launch_widen(net->w, net->w64, ns, net->stream);
    HIP_TRY(hipMemsetAsync(net->tr_pre, 0, 4 * static_cast<size_t>(N), net->stream));
    HIP_TRY(hipMemsetAsync(net->tr_post, 0, 4 * static_cast<size_t>(N), net->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(net->stream));
    net->sp = *p;
    // per-step trace decay e^{-dt/tau}, dt from population 0
    const double dt = static_cast<double>(net->pops[0].puni[0]);
    net->stdp_dpre = static_cast<float>(std::exp(-dt / static_cast<double>(p->tau_plus)));
    net->stdp_dpost = static_cast<float>(std::exp(-dt / static_cast<double>(p->tau_minus)));
    net->stdp = true;
    net->drop_graph();
    return SIMUSE_OK;
}

int simuse_net_set_scatter_mode(simuse_net* net, int32_t mode) {
    int rc = check_net(net);
    if (rc) return rc;
    if (mode != SIMUSE_SCATTER_TILE && mode != SIMUSE_SCATTER_WAVE && mode != SIMUSE_SCATTER_AUTO)
        return fail(SIMUSE_E_INVALID, "unknown scatter mode %d", mode);
    if (mode == SIMUSE_SCATTER_AUTO) return simuse_net_autotune(net, 16);
    if (net->sched != mode) { net->sched = mode; net->drop_graph(); }
    return SIMUSE_OK;
}

int simuse_net_get_scatter_mode(simuse_net* net, int32_t* mode) {
    int rc = check_net(net);
    if (rc) return rc;
    if (mode) *mode = net->sched;
    return SIMUSE_OK;
}

int simuse_net_set_graph_steps(simuse_net* net, uint32_t steps) {
    int rc = check_net(net);
    if (rc) return rc;
    if (steps < 1 || steps > 64) return fail(SIMUSE_E_INVALID, "graph steps must be 1..64");
    if (net->graph_steps != steps) { net->graph_steps = steps; net->drop_graph(); }
    return SIMUSE_OK;
}

int simuse_net_step_async(simuse_net* net, uint32_t n_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    if ((rc = net->finalize())) return rc;
    if (n_steps == 0) return SIMUSE_OK;
    if (net->profiling) {
        if ((rc = net->enqueue_profiled(n_steps))) return rc;
        net->host_step += n_steps;
        return SIMUSE_OK;
    }
    uint32_t full = n_steps / net->graph_steps, rest = n_steps % net->graph_steps;
    if (full) {
        if (!net->gexec && (rc = net->build_graph())) return rc;
        for (uint32_t k = 0; k < full; ++k) HIP_TRY(hipGraphLaunch(net->gexec, net->stream));
    }
    if (rest && (rc = net->enqueue_steps(rest))) return rc;
    net->host_step += n_steps;
    return SIMUSE_OK;
}

int simuse_net_sync(simuse_net* net) {
    int rc = check_net(net);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_step(simuse_net* net) {
    int rc = simuse_net_step_async(net, 1);
    if (rc) return rc;
    return simuse_net_sync(net);
}

int simuse_net_download_spikes(simuse_net* net, int32_t* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!out || len != net->N) return fail(SIMUSE_E_INVALID, "spikes len %llu != n_total %u",
                                           static_cast<unsigned long long>(len), net->N);
    std::vector<uint64_t> words(net->n_words);
    HIP_TRY(hipMemcpyAsync(words.data(), net->bits, 8 * words.size(), hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    for (const auto& P : net->pops)
        for (uint32_t i = 0; i < P.n; ++i)
            out[P.off + i] = static_cast<int32_t>((words[P.word0 + i / 64] >> (i % 64)) & 1ull);
    return SIMUSE_OK;
}

int simuse_net_download_spike_ids(simuse_net* net, uint32_t* ids, uint32_t cap, uint32_t* count) {
    int rc = check_net(net);
    if (rc) return rc;
    std::vector<uint32_t> cnt(net->n_tiles);
    HIP_TRY(hipMemcpyAsync(cnt.data(), net->tile_count, 4 * cnt.size(), hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    uint32_t total = 0;
    for (uint32_t t = 0; t < net->n_tiles; ++t) {
        const uint32_t c = cnt[t];
        if (c && ids && total < cap) {
            const uint32_t take = std::min(c, cap - total);
            HIP_TRY(hipMemcpyAsync(ids + total, net->active + static_cast<size_t>(t) * kTile, 4 * take,
                                   hipMemcpyDeviceToHost, net->stream));
        }
        total += c;
    }
    HIP_TRY(hipStreamSynchronize(net->stream));
    if (count) *count = total;
    return SIMUSE_OK;
}

int simuse_net_download_state(simuse_net* net, int32_t pop, float* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    if (pop < 0 || pop >= static_cast<int32_t>(net->pops.size()))
        return fail(SIMUSE_E_INVALID, "CompiledNetwork: pop %d out of range", pop);
    const HostPop& P = net->pops[pop];
    if (!out || len != static_cast<uint64_t>(P.ns) * P.host_stride)
        return fail(SIMUSE_E_INVALID, "states len %llu != ns*stride %llu", static_cast<unsigned long long>(len),
                    static_cast<unsigned long long>(P.ns) * P.host_stride);
    std::memset(out, 0, sizeof(float) * len);
    if (P.n == 0) return SIMUSE_OK;
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(float) * P.host_stride, P.st, sizeof(float) * P.dstride, sizeof(float) * P.n,
                             P.ns, hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_download_ring(simuse_net* net, float* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!out || len != static_cast<uint64_t>(kSlots) * net->N)
        return fail(SIMUSE_E_INVALID, "ring len must be 32*n_total");
    HIP_TRY(hipMemcpyAsync(out, net->ring, 4 * len, hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_upload_weights(simuse_net* net, const float* w, uint64_t n_syn) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n_syn != net->n_syn || (n_syn && !w)) return fail(SIMUSE_E_INVALID, "weights len mismatch");
    for (uint64_t e = 0; e < n_syn; ++e)
        if (!std::isfinite(w[e])) return fail(SIMUSE_E_INVALID, "weight %llu must be finite", static_cast<unsigned long long>(e));
    HIP_TRY(hipStreamSynchronize(net->stream));
    if (n_syn) HIP_TRY(hipMemcpy(net->w, w, 4 * n_syn, hipMemcpyHostToDevice));
    if (net->w64) {
        launch_widen(net->w, net->w64, n_syn, net->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(net->stream));
    }
    return SIMUSE_OK;
}

int simuse_net_download_weights(simuse_net* net, float* w, uint64_t n_syn) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n_syn != net->n_syn || (n_syn && !w)) return fail(SIMUSE_E_INVALID, "weights len mismatch");
    if (n_syn) HIP_TRY(hipMemcpyAsync(w, net->w, 4 * n_syn, hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_inject_current(simuse_net* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    if (pop < 0 || pop >= static_cast<int32_t>(net->pops.size()))
        return fail(SIMUSE_E_INVALID, "inject_current: population %d out of range (have %zu)", pop, net->pops.size());
    if (count == 0) return SIMUSE_OK;
    if (!idx || !val) return fail(SIMUSE_E_INVALID, "inject_current: null arrays");
    const HostPop& P = net->pops[pop];
    std::vector<uint32_t> g(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (idx[i] >= P.n)
            return fail(SIMUSE_E_INVALID, "inject_current: neuron index %u out of bounds (population has %u)", idx[i], P.n);
        if (!std::isfinite(val[i])) return fail(SIMUSE_E_INVALID, "inject_current: value %u must be finite", i);
        g[i] = P.off + idx[i];
    }
    std::vector<uint32_t> sorted(g);
    std::sort(sorted.begin(), sorted.end());
    const bool dup = std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end();
    if (count > net->inj_cap) {
        // grow the staging buffers (old ones stay owned by the handle)
        uint32_t cap = std::max(count, 2 * net->inj_cap);
        if ((rc = net->alloc(&net->inj_idx, cap))) return rc;
        if ((rc = net->alloc(&net->inj_val, cap))) return rc;
        net->inj_cap = cap;
    }
    HIP_TRY(hipStreamSynchronize(net->stream));
    HIP_TRY(hipMemcpy(net->inj_idx, g.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(net->inj_val, val, 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    launch_inject(net->ring, net->step, net->N, net->inj_idx, net->inj_val, count, dup ? 1 : 0, net->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_reset(simuse_net* net, const simuse_pop_desc* pops, int32_t n_pops) {
    int rc = check_net(net);
    if (rc) return rc;
    uint64_t nt = 0;
    if ((rc = validate_pops(pops, n_pops, &nt))) return rc;
    if (n_pops != static_cast<int32_t>(net->pops.size()) || nt != net->N)
        return fail(SIMUSE_E_INVALID, "reset: population layout differs from the network");
    for (int32_t p = 0; p < n_pops; ++p)
        if (pops[p].model != net->pops[p].model || pops[p].n != net->pops[p].n ||
            pops[p].stride != net->pops[p].host_stride)
            return fail(SIMUSE_E_INVALID, "reset: population %d differs from the network", p);
    HIP_TRY(hipStreamSynchronize(net->stream));
    if ((rc = net->upload_pops(pops))) return rc;
    net->drop_graph();
    return net->reset_dynamic();
}

int simuse_net_step_counter(simuse_net* net, uint64_t* step) {
    int rc = check_net(net);
    if (rc) return rc;
    if (step) *step = net->host_step;
    return SIMUSE_OK;
}
int simuse_net_n_total(simuse_net* net, uint64_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = net->N;
    return SIMUSE_OK;
}
int simuse_net_n_pops(simuse_net* net, int32_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = static_cast<int32_t>(net->pops.size());
    return SIMUSE_OK;
}
int simuse_net_pop_n(simuse_net* net, int32_t pop, uint64_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = (pop >= 0 && pop < static_cast<int32_t>(net->pops.size())) ? net->pops[pop].n : 0;
    return SIMUSE_OK;
}

int simuse_net_counters(simuse_net* net, uint64_t* spikes, uint64_t* events) {
    int rc = check_net(net);
    if (rc) return rc;
    std::vector<uint64_t> ts(net->n_tiles), te(net->n_tiles + net->n_src_tiles);
    HIP_TRY(hipMemcpyAsync(ts.data(), net->tile_spikes, 8 * ts.size(), hipMemcpyDeviceToHost, net->stream));
    if (net->tile_events)
        HIP_TRY(hipMemcpyAsync(te.data(), net->tile_events, 8 * te.size(), hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    uint64_t s = 0, e = 0;
    for (uint64_t x : ts) s += x;
    if (net->tile_events) for (uint64_t x : te) e += x;
    if (spikes) *spikes = s;
    if (events) *events = e;
    return SIMUSE_OK;
}

int simuse_net_record(simuse_net* net, uint32_t max_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    if (max_steps > net->raster_cap) {
        uint64_t* r = nullptr;
        if ((rc = net->alloc(&r, static_cast<size_t>(max_steps) * net->n_words))) return rc;
        net->raster = r;
        net->raster_cap = max_steps;
    }
    if (max_steps == 0) net->raster = nullptr;
    if (net->raster)
        HIP_TRY(hipMemset(net->raster, 0, 8 * static_cast<size_t>(net->raster_cap) * net->n_words));
    // recording starts at the current device step
    HIP_TRY(hipMemcpy(net->step + 1, net->step, 8, hipMemcpyDeviceToDevice));
    net->raster_start = net->host_step;
    net->drop_graph();
    return SIMUSE_OK;
}

int simuse_net_raster_words(simuse_net* net, uint64_t* words) {
    int rc = check_net(net);
    if (rc) return rc;
    if (words) *words = net->n_words;
    return SIMUSE_OK;
}

int simuse_net_download_raster(simuse_net* net, uint64_t* out, uint64_t n_words, uint32_t* steps) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!net->raster) return fail(SIMUSE_E_STATE, "recording is off");
    uint64_t s = std::min<uint64_t>(net->host_step - net->raster_start, net->raster_cap);
    if (steps) *steps = static_cast<uint32_t>(s);
    if (!out) return SIMUSE_OK;
    if (n_words < s * net->n_words) return fail(SIMUSE_E_INVALID, "raster buffer too small");
    HIP_TRY(hipMemcpyAsync(out, net->raster, 8 * s * net->n_words, hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_raster_stats(simuse_net* net, uint32_t first, uint32_t count, float dt_s, float* rates, float* cvs) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!net->raster) return fail(SIMUSE_E_STATE, "recording is off");
    if (static_cast<uint64_t>(first) + count > net->N) return fail(SIMUSE_E_INVALID, "neuron range out of bounds");
    if (count == 0) return SIMUSE_OK;
    uint32_t steps = static_cast<uint32_t>(std::min<uint64_t>(net->host_step - net->raster_start, net->raster_cap));
    std::vector<uint32_t> wo(count), bo(count);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t gi = first + k;
        const HostPop* P = nullptr;
        for (const auto& q : net->pops)
            if (gi >= q.off && gi < q.off + q.n) { P = &q; break; }
        const uint32_t li = gi - P->off;
        wo[k] = P->word0 + li / 64;
        bo[k] = li % 64;
    }
    uint32_t *dwo = nullptr, *dbo = nullptr;
    float *dr = nullptr, *dc = nullptr;
    HIP_TRY(hipMalloc(&dwo, 4 * static_cast<size_t>(count)));
    HIP_TRY(hipMalloc(&dbo, 4 * static_cast<size_t>(count)));
    HIP_TRY(hipMalloc(&dr, 4 * static_cast<size_t>(count)));
    HIP_TRY(hipMalloc(&dc, 4 * static_cast<size_t>(count)));
    HIP_TRY(hipMemcpy(dwo, wo.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dbo, bo.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    launch_raster_stats(net->raster, net->n_words, steps, dwo, dbo, count, dt_s, dr, dc, net->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
    if (e == hipSuccess && rates) e = hipMemcpy(rates, dr, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost);
    if (e == hipSuccess && cvs) e = hipMemcpy(cvs, dc, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost);
    (void)hipFree(dwo); (void)hipFree(dbo); (void)hipFree(dr); (void)hipFree(dc);
    if (e != hipSuccess) return fail(SIMUSE_E_HIP, "raster stats: %s", hipGetErrorString(e));
    return SIMUSE_OK;
}

int simuse_net_set_profiling(simuse_net* net, int32_t on) {
    int rc = check_net(net);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    net->profiling = on != 0;
    net->prof_steps = 0;
    return SIMUSE_OK;
}

int simuse_net_kernel_times(simuse_net* net, float* ms, int32_t n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!ms || n < 1) return fail(SIMUSE_E_INVALID, "null output");
    HIP_TRY(hipStreamSynchronize(net->stream));
    double acc[simuse_net::kPhases] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < net->prof_steps; ++k) {
        hipEvent_t* e = net->ev.data() + static_cast<size_t>(k) * (simuse_net::kPhases + 1);
        for (int p = 0; p < simuse_net::kPhases; ++p) {
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, e[p], e[p + 1]));
            acc[p] += t;
        }
    }
    for (int p = 0; p < n && p < simuse_net::kPhases; ++p)
        ms[p] = net->prof_steps ? static_cast<float>(acc[p] / net->prof_steps) : 0.0f;
    net->prof_steps = 0;
    return SIMUSE_OK;
}

// Snapshot of everything a step mutates, for autotuning trials.
struct Snapshot {
    std::vector<std::pair<void*, void*>> copies;  // (live, saved)
    std::vector<size_t> sizes;
    ~Snapshot() { for (auto& c : copies) (void)hipFree(c.second); }
};

static int snap_add(Snapshot& s, void* live, size_t bytes) {
    if (!live || bytes == 0) return SIMUSE_OK;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) { (void)hipGetLastError(); return fail(SIMUSE_E_NOMEM, "autotune snapshot: %s", hipGetErrorString(e)); }
    s.copies.push_back({live, q});
    s.sizes.push_back(bytes);
    return SIMUSE_OK;
}

static int snap_copy(simuse_net* net, Snapshot& s, bool save) {
    for (size_t k = 0; k < s.copies.size(); ++k) {
        void* dst = save ? s.copies[k].second : s.copies[k].first;
        void* src = save ? s.copies[k].first : s.copies[k].second;
        HIP_TRY(hipMemcpyAsync(dst, src, s.sizes[k], hipMemcpyDeviceToDevice, net->stream));
    }
    HIP_TRY(hipStreamSynchronize(net->stream));
    return SIMUSE_OK;
}

int simuse_net_autotune(simuse_net* net, uint32_t trial_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    if ((rc = net->finalize())) return rc;
    if (trial_steps == 0) trial_steps = 16;
    HIP_TRY(hipStreamSynchronize(net->stream));
    Snapshot snap;
    for (auto& P : net->pops) {
        if ((rc = snap_add(snap, P.st, sizeof(float) * P.ns * P.dstride))) return rc;
    }
    const size_t N = net->N;
    if ((rc = snap_add(snap, net->ring, sizeof(float) * kSlots * N))) return rc;
    if ((rc = snap_add(snap, net->step, sizeof(uint64_t) * 2))) return rc;
    if ((rc = snap_add(snap, net->tile_count, sizeof(uint32_t) * net->n_tiles))) return rc;
    if ((rc = snap_add(snap, net->active, sizeof(uint32_t) * kTile * net->n_tiles))) return rc;
    if ((rc = snap_add(snap, net->bits, sizeof(uint64_t) * net->n_words))) return rc;
    if ((rc = snap_add(snap, net->tile_spikes, sizeof(uint64_t) * net->n_tiles))) return rc;
    if ((rc = snap_add(snap, net->tile_events, sizeof(uint64_t) * (net->n_tiles + net->n_src_tiles)))) return rc;
    if (net->stdp) {
        if ((rc = snap_add(snap, net->w, sizeof(float) * net->n_syn))) return rc;
        if ((rc = snap_add(snap, net->w64, sizeof(double) * net->n_syn))) return rc;
        if ((rc = snap_add(snap, net->tr_pre, sizeof(float) * N))) return rc;
        if ((rc = snap_add(snap, net->tr_post, sizeof(float) * N))) return rc;
    }
    if ((rc = snap_copy(net, snap, true))) return rc;
    const int cand[2] = {SIMUSE_SCATTER_TILE, SIMUSE_SCATTER_WAVE};
    float best_ms = 0;
    int best = net->sched;
    const int saved_sched = net->sched;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int c = 0; c < 2; ++c) {
        net->sched = cand[c];
        if ((rc = snap_copy(net, snap, false))) break;
        if ((rc = net->enqueue_steps(2))) break;  // warm-up
        if ((rc = snap_copy(net, snap, false))) break;
        HIP_TRY(hipEventRecord(e0, net->stream));
        if ((rc = net->enqueue_steps(trial_steps))) break;
        HIP_TRY(hipEventRecord(e1, net->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (c == 0 || ms < best_ms) { best_ms = ms; best = cand[c]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int rc2 = snap_copy(net, snap, false);
    if (rc) { net->sched = saved_sched; return rc; }
    if (rc2) return rc2;
    if (best != saved_sched) net->drop_graph();
    net->sched = best;
    return SIMUSE_OK;
}

int simuse_net_device_bytes(simuse_net* net, uint64_t* bytes) {
    int rc = check_net(net);
    if (rc) return rc;
    if (bytes) *bytes = net->bytes;
    return SIMUSE_OK;
}

}  // extern "C"
