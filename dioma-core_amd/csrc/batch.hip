// batch.hip — many independent networks stepped by one engine, and
// on-device generation of seeded random networks.
//
// A batch is an Engine over many NetStates: every step is one launch per
// neuron model plus one insert and one fold launch covering all networks
// (captured in a hipGraph).  Networks keep their own rings, step counters and
// RNG keys, so a network's trajectory is bit-identical alone or in a batch.
// A network can carry a horizon (step limit); once reached, every kernel
// skips it, so networks of different lengths can share launches.
//
// The generator builds the CSR and the populations on the device from a
// simuse_gen_desc, so a request costs no host-side graph construction.

#include "device_math.hpp"
#include "engine.hpp"

namespace simuse {

namespace {

__device__ __forceinline__ float unit24(uint64_t h) {  // [0, 1) on 24 bits
    return static_cast<float>(h >> 40) * 5.9604644775390625e-08f;
}

// Edge e = i*k + j of neuron i: target uniform over the other n-1 neurons,
// weight U(0, w_exc) for excitatory rows and -U(0, w_inh) for inhibitory
// rows, delay uniform in [d_min, d_max].
__global__ void gen_csr_kernel(uint32_t* row_ptr, uint32_t* col, float* w, uint8_t* delay, uint32_t n, uint32_t k,
                               uint32_t n_exc, uint64_t seed, float w_exc, float w_inh, uint32_t d_min,
                               uint32_t d_span) {
    const uint64_t ne = static_cast<uint64_t>(n) * k;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t i = static_cast<uint32_t>(e / k);
        const uint64_t h = splitmix(seed ^ (e * 0x9E3779B97F4A7C15ull));
        const uint64_t h2 = splitmix(h ^ 0xD1B54A32D192ED03ull);
        uint32_t post = static_cast<uint32_t>(((h & 0xFFFFFFFFull) * static_cast<uint64_t>(n - 1)) >> 32);
        post += post >= i ? 1u : 0u;
        col[e] = post;
        const float u = unit24(h2);
        w[e] = i < n_exc ? u * w_exc : -(u * w_inh);
        delay[e] = static_cast<uint8_t>(d_min + static_cast<uint32_t>((h2 & 0xFFFFull) % d_span));
        if (e % k == 0) row_ptr[i] = static_cast<uint32_t>(e);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) row_ptr[n] = static_cast<uint32_t>(ne);
}

// LIF: v0 ~ U(-65, -50), refr 0 (params are uniform scalars).
__global__ void gen_lif_state_kernel(float* st, uint32_t n, uint32_t dstride, uint64_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = unit24(splitmix(seed ^ (static_cast<uint64_t>(i) * 0xA24BAED4963EE407ull)));
    st[i] = -65.0f + 15.0f * u;
    st[dstride + i] = 0.0f;
}

// Izhikevich (2003) RS (exc) and FS (inh) populations with r ~ U(0,1):
// RS c = -65 + 15 r^2, d = 8 - 6 r^2;  FS a = 0.02 + 0.08 r, b = 0.25 - 0.05 r.
// v0 = -65, u0 = b v0, refr 0.
__global__ void gen_izh_kernel(float* st, float* pr, uint32_t n, uint32_t dstride, uint64_t seed, int fs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = unit24(splitmix(seed ^ (static_cast<uint64_t>(i) * 0xA24BAED4963EE407ull)));
    float b;
    if (fs) {
        pr[1 * static_cast<size_t>(dstride) + i] = 0.02f + 0.08f * r;
        b = 0.25f - 0.05f * r;
        pr[2 * static_cast<size_t>(dstride) + i] = b;
    } else {
        b = 0.2f;
        pr[3 * static_cast<size_t>(dstride) + i] = -65.0f + 15.0f * (r * r);
        pr[4 * static_cast<size_t>(dstride) + i] = 8.0f - 6.0f * (r * r);
    }
    st[i] = -65.0f;
    st[dstride + i] = b * -65.0f;
    st[2 * static_cast<size_t>(dstride) + i] = 0.0f;
}

// Hodgkin-Huxley: v0 ~ U(-70, -60), gates at the standard resting values
// m 0.0529, h 0.596, n 0.318 (CN/src/sim/hh.rs:34-38), refr 0, CaT gate 0.
__global__ void gen_hh_state_kernel(float* st, uint32_t n, uint32_t dstride, uint64_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = unit24(splitmix(seed ^ (static_cast<uint64_t>(i) * 0xA24BAED4963EE407ull)));
    const size_t ds = dstride;
    st[i] = -70.0f + 10.0f * u;
    st[ds + i] = 0.0529f;
    st[2 * ds + i] = 0.596f;
    st[3 * ds + i] = 0.318f;
    st[4 * ds + i] = 0.0f;
    st[5 * ds + i] = 0.0f;
}

}  // namespace

// Build a generated network in place (replaces NetState::init for it).
int build_generated(NetState& s, const simuse_gen_desc& g, hipStream_t stream) {
    const uint32_t n = g.n;
    const uint32_t ne = static_cast<uint32_t>(std::floor(static_cast<double>(n) * g.exc_frac + 0.5));
    std::vector<simuse_pop_desc> descs;
    // populations: LIF one population; Izhikevich RS then FS
    struct Spec { uint32_t n; int fs; };
    std::vector<Spec> specs;
    if (g.model == SIMUSE_MODEL_LIF || g.model == SIMUSE_MODEL_HH) specs.push_back({n, 0});
    else {
        if (ne) specs.push_back({ne, 0});
        if (n - ne) specs.push_back({n - ne, 1});
    }
    uint32_t off = 0, tile = 0, word = 0;
    for (const Spec& sp : specs) {
        HostPop P;
        P.model = g.model;
        P.n = sp.n;
        P.ns = model_states(g.model);
        P.np = model_params(g.model);
        P.host_stride = (sp.n + 7u) / 8u * 8u;
        P.off = off;
        P.tile0 = tile;
        P.ntiles = (P.n + kTile - 1) / kTile;
        P.word0 = word;
        P.dstride = (P.n + 63u) / 64u * 64u;
        off += P.n;
        tile += P.ntiles;
        word += P.ntiles * kWavesPerTile;
        // parameters: uniform columns as scalars
        for (int k = 0; k < kMaxParams; ++k) P.puni[k] = 0.0f;
        P.puni[0] = g.dt;
        if (g.model == SIMUSE_MODEL_LIF) {
            const float vals[6] = {g.dt, 1.0f / 20.0f, -65.0f, -50.0f, -65.0f, 2.0f};
            for (int k = 0; k < 6; ++k) P.puni[k] = vals[k];
            P.uni_mask = 0x3Fu;
        } else if (g.model == SIMUSE_MODEL_HH) {
            // standard HH (CN/src/sim/hh.rs:26-31), threshold 0 mV, tau_ref 2 ms
            const float vals[13] = {g.dt, 120.0f, 36.0f, 0.3f, 50.0f, -77.0f, -65.0f, 1.0f, 1.0f, 0.0f,
                                    0.0f, 120.0f, 2.0f};
            for (int k = 0; k < 13; ++k) P.puni[k] = vals[k];
            P.uni_mask = 0x1FFFu;
        } else if (!sp.fs) {
            const float vals[7] = {g.dt, 0.02f, 0.2f, 0.0f, 0.0f, 0.0f, 30.0f};
            for (int k = 0; k < 7; ++k) P.puni[k] = vals[k];
            P.uni_mask = (1u << 0) | (1u << 1) | (1u << 2) | (1u << 5) | (1u << 6);
            P.flags |= kPopRefrElided;
        } else {
            const float vals[7] = {g.dt, 0.0f, 0.0f, -65.0f, 2.0f, 0.0f, 30.0f};
            for (int k = 0; k < 7; ++k) P.puni[k] = vals[k];
            P.uni_mask = (1u << 0) | (1u << 3) | (1u << 4) | (1u << 5) | (1u << 6);
            P.flags |= kPopRefrElided;
        }
        if (g.noise_sigma > 0.0f) {
            P.flags |= kPopNoise | kPopNoiseScalar;
            P.sigma_uni = g.noise_sigma;
            P.seed = g.seed * 0x9E3779B97F4A7C15ull + 17 + static_cast<uint64_t>(sp.fs);
        }
        if (g.drive_rate_hz > 0.0f) {
            P.flags |= kPopDrive;
            P.drive_p = std::min(std::max(g.drive_rate_hz * g.dt / 1000.0f, 0.0f), 1.0f);
            P.drive_w = g.drive_weight;
            P.drive_seed0 = static_cast<uint32_t>(g.seed * 2654435761ull + P.off);
        }
        s.pops.push_back(P);
    }
    s.N = off;
    s.n_tiles = tile;
    s.n_words = word;
    s.n_syn = static_cast<uint64_t>(n) * g.k;
    int rc;
    for (size_t p = 0; p < s.pops.size(); ++p) {
        HostPop& P = s.pops[p];
        if ((rc = s.alloc(&P.st, static_cast<size_t>(P.ns) * P.dstride))) return rc;
        if ((rc = s.alloc(&P.pr, static_cast<size_t>(P.np) * P.dstride))) return rc;
        HIP_TRY(hipMemsetAsync(P.pr, 0, sizeof(float) * P.np * P.dstride, stream));
        const uint64_t pseed = g.seed * 0xD1B54A32D192ED03ull + 1000003ull * (p + 1);
        if (g.model == SIMUSE_MODEL_LIF)
            gen_lif_state_kernel<<<(P.n + 255) / 256, 256, 0, stream>>>(P.st, P.n, P.dstride, pseed);
        else if (g.model == SIMUSE_MODEL_HH)
            gen_hh_state_kernel<<<(P.n + 255) / 256, 256, 0, stream>>>(P.st, P.n, P.dstride, pseed);
        else
            gen_izh_kernel<<<(P.n + 255) / 256, 256, 0, stream>>>(P.st, P.pr, P.n, P.dstride, pseed,
                                                                  specs[p].fs);
        HIP_TRY(hipGetLastError());
    }
    if ((rc = s.alloc(&s.d_pops, s.pops.size()))) return rc;
    if ((rc = s.upload_pop_table())) return rc;
    std::vector<uint32_t> tile_pop(s.n_tiles);
    for (size_t p = 0; p < s.pops.size(); ++p)
        for (uint32_t t = 0; t < s.pops[p].ntiles; ++t) tile_pop[s.pops[p].tile0 + t] = static_cast<uint32_t>(p);
    if ((rc = s.alloc(&s.d_tile_pop, s.n_tiles))) return rc;
    HIP_TRY(hipMemcpy(s.d_tile_pop, tile_pop.data(), 4 * tile_pop.size(), hipMemcpyHostToDevice));
    if ((rc = s.alloc(&s.row_ptr, static_cast<size_t>(n) + 1))) return rc;
    if ((rc = s.alloc(&s.col, s.n_syn))) return rc;
    if ((rc = s.alloc(&s.w, s.n_syn))) return rc;
    if ((rc = s.alloc(&s.delay, s.n_syn))) return rc;
    const uint32_t d_min = g.d_min, d_span = static_cast<uint32_t>(g.d_max) - g.d_min + 1u;
    s.max_delay = g.d_max;
    uint64_t blocks = (s.n_syn + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    gen_csr_kernel<<<static_cast<uint32_t>(blocks), 256, 0, stream>>>(s.row_ptr, s.col, s.w, s.delay, n, g.k, ne,
                                                                      g.seed, g.w_exc, g.w_inh, d_min, d_span);
    HIP_TRY(hipGetLastError());
    s.h_row_ptr.resize(static_cast<size_t>(n) + 1);
    for (uint32_t i = 0; i <= n; ++i) s.h_row_ptr[i] = i * g.k;
    s.uni_deg = g.k;  // row i holds edges [i k, (i + 1) k)
    if ((rc = s.alloc(&s.ring, static_cast<size_t>(kSlots) * s.N))) return rc;
    if ((rc = s.alloc(&s.tile_count, s.n_tiles))) return rc;
    if ((rc = s.alloc(&s.active, static_cast<size_t>(s.n_tiles) * kTile))) return rc;
    if ((rc = s.alloc(&s.bits, s.n_words))) return rc;
    if ((rc = s.alloc(&s.tile_spikes, s.n_tiles))) return rc;
    if ((rc = s.alloc(&s.step, 2))) return rc;
    return s.reset_dynamic(stream);
}

int validate_gen(const simuse_gen_desc* g) {
    if (!g) return fail(SIMUSE_E_INVALID, "null generator description");
    if (g->model != SIMUSE_MODEL_LIF && g->model != SIMUSE_MODEL_IZHIKEVICH && g->model != SIMUSE_MODEL_HH)
        return fail(SIMUSE_E_UNSUPPORTED, "generator supports LIF, Izhikevich and HH networks");
    if (g->n < 2 || g->n > kMaxNeurons) return fail(SIMUSE_E_INVALID, "generator: n must be in [2, 2^27]");
    if (static_cast<uint64_t>(g->n) * g->k >= 0xFFFFFFFFull)
        return fail(SIMUSE_E_INVALID, "generator: n*k exceeds the u32 CSR");
    if (!(g->dt > 0.0f) || !std::isfinite(g->dt)) return fail(SIMUSE_E_INVALID, "generator: dt must be > 0");
    if (!(g->exc_frac >= 0.0f && g->exc_frac <= 1.0f)) return fail(SIMUSE_E_INVALID, "generator: exc_frac in [0,1]");
    if (g->d_min > g->d_max || g->d_max >= kSlots) return fail(SIMUSE_E_INVALID, "generator: delays must be in 0..31");
    if (!std::isfinite(g->w_exc) || !std::isfinite(g->w_inh) || !std::isfinite(g->drive_weight) ||
        !std::isfinite(g->noise_sigma) || !(g->drive_rate_hz >= 0.0f))
        return fail(SIMUSE_E_INVALID, "generator: weights, noise and drive must be finite");
    return SIMUSE_OK;
}

// Estimated device bytes of a generated network (admission control).
uint64_t generated_bytes(const simuse_gen_desc& g) {
    const uint64_t n = g.n, ns = n * g.k;
    const uint64_t tiles = (n + kTile - 1) / kTile + 2;
    uint64_t b = ns * (4 + 4 + 1);                           // col, w, delay
    if (g.d_max <= 1)
        b += ns * 32 + n * 4 * 2 + n * 32 * 4;               // binned records (two buffers), row_ptr, ring
    else
        b += ns * (8 + 8) + n * 4 * 2 + n * 32 * (4 + 8);    // records, collision lists, row_ptr, ring, heads
    const uint64_t cols = static_cast<uint64_t>(model_states(g.model) + model_params(g.model));
    b += n * 4 * cols + tiles * (kTile * 4 + 4 + 8 + 8 + 8);  // state/params, compaction, counters
    return b + (1u << 20);
}

}  // namespace simuse

using namespace simuse;

struct simuse_batch {
    Engine eng;
    std::vector<std::unique_ptr<NetState>> nets;  // slot id -> network (null when removed)
    uint64_t bytes_live = 0;

    void rebuild_set() {
        eng.nets.clear();
        for (auto& p : nets)
            if (p) eng.nets.push_back(p.get());
        eng.mark_dirty();
    }
    NetState* get(int32_t id) {
        if (id < 0 || id >= static_cast<int32_t>(nets.size()) || !nets[id]) return nullptr;
        return nets[id].get();
    }
    int32_t place(std::unique_ptr<NetState> s) {
        for (size_t k = 0; k < nets.size(); ++k)
            if (!nets[k]) {
                nets[k] = std::move(s);
                return static_cast<int32_t>(k);
            }
        nets.push_back(std::move(s));
        return static_cast<int32_t>(nets.size() - 1);
    }
};

#define BATCH_NET(b, id, s)                                                          \
    if (!(b)) return fail(SIMUSE_E_INVALID, "null batch handle");                    \
    NetState* s = (b)->get(id);                                                      \
    if (!s) return fail(SIMUSE_E_INVALID, "network %d is not in the batch", (int)(id))

extern "C" {

int simuse_batch_create(simuse_batch** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "null output handle");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "no HIP device available");
    }
    std::unique_ptr<simuse_batch> b(new (std::nothrow) simuse_batch());
    if (!b) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    int rc = b->eng.init();
    if (rc) return rc;
    *out = b.release();
    return SIMUSE_OK;
}

void simuse_batch_destroy(simuse_batch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->eng.stream);
    b->eng.nets.clear();
    b->nets.clear();
    delete b;
}

int simuse_batch_add_network(simuse_batch* b, const simuse_pop_desc* pops, int32_t n_pops, const uint32_t* row_ptr,
                             const uint32_t* col, const float* w, const uint8_t* delay, uint64_t n_syn,
                             int32_t* net_id) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    uint64_t n_total = 0;
    int rc = validate_pops(pops, n_pops, &n_total);
    if (rc) return rc;
    if ((rc = validate_csr(n_total, row_ptr, col, w, delay, n_syn))) return rc;
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    std::unique_ptr<NetState> s(new (std::nothrow) NetState());
    if (!s) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    s->eng = &b->eng;
    if ((rc = s->init(pops, n_pops, row_ptr, col, w, delay, n_syn, b->eng.stream))) return rc;
    b->bytes_live += s->bytes;
    const int32_t id = b->place(std::move(s));
    if (net_id) *net_id = id;
    b->rebuild_set();
    return SIMUSE_OK;
}

int simuse_batch_add_generated(simuse_batch* b, const simuse_gen_desc* g, int32_t* net_id) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    int rc = validate_gen(g);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    std::unique_ptr<NetState> s(new (std::nothrow) NetState());
    if (!s) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    s->eng = &b->eng;
    if ((rc = build_generated(*s, *g, b->eng.stream))) return rc;
    b->bytes_live += s->bytes;
    const int32_t id = b->place(std::move(s));
    if (net_id) *net_id = id;
    b->rebuild_set();
    return SIMUSE_OK;
}

int simuse_batch_remove_network(simuse_batch* b, int32_t id) {
    BATCH_NET(b, id, s);
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    b->bytes_live -= s->bytes;
    b->nets[id].reset();
    b->rebuild_set();
    return SIMUSE_OK;
}

int simuse_batch_n_networks(simuse_batch* b, int32_t* n) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    if (n) *n = static_cast<int32_t>(b->eng.nets.size());
    return SIMUSE_OK;
}

int simuse_batch_add_poisson_source(simuse_batch* b, int32_t id, float rate_hz, float dt_ms, uint32_t seed,
                                    const uint32_t* posts, const float* w, const uint8_t* delay, uint32_t count) {
    BATCH_NET(b, id, s);
    return s->add_source(rate_hz, dt_ms, seed, posts, w, delay, count);
}

int simuse_batch_set_noise(simuse_batch* b, int32_t id, int32_t pop, const float* sigma, uint64_t seed) {
    BATCH_NET(b, id, s);
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    int rc = s->set_noise(pop, sigma, seed);
    b->eng.mark_dirty();
    return rc;
}

int simuse_batch_set_plasticity(simuse_batch* b, int32_t id, const simuse_stdp_params* p, const uint8_t* plastic) {
    BATCH_NET(b, id, s);
    if (p) {
        int rl = b->eng.leave_megastep(s);
        if (rl) return rl;
    }
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    int rc = s->set_plasticity(p, plastic, b->eng.stream);
    b->eng.mark_dirty();
    return rc;
}

int simuse_batch_set_short_term_plasticity(simuse_batch* b, int32_t id, const simuse_stp_profile* profiles,
                                           int32_t n_profiles, float dt, const uint8_t* syn_types,
                                           const uint8_t* src_types, uint64_t n_src_edges) {
    BATCH_NET(b, id, s);
    return s->set_stp(profiles, n_profiles, dt, syn_types, src_types, n_src_edges);
}

int simuse_batch_set_poisson_drive(simuse_batch* b, int32_t id, int32_t pop, float rate_hz, float dt_ms,
                                   float weight, uint32_t seed0) {
    BATCH_NET(b, id, s);
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    int rc = s->set_drive(pop, rate_hz, dt_ms, weight, seed0);
    b->eng.mark_dirty();
    return rc;
}

int simuse_batch_set_horizon(simuse_batch* b, int32_t id, uint64_t steps) {
    BATCH_NET(b, id, s);
    if (steps < s->host_step) return fail(SIMUSE_E_STATE, "horizon %llu is behind the network's step %llu",
                                          static_cast<unsigned long long>(steps),
                                          static_cast<unsigned long long>(s->host_step));
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    s->step_limit = steps;
    b->eng.mark_dirty();
    return SIMUSE_OK;
}

int simuse_batch_set_one_kernel(simuse_batch* b, int32_t on) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    if (on != 0 && on != 1) return fail(SIMUSE_E_INVALID, "one-kernel switch must be 0 or 1");
    return b->eng.set_one_kernel(on != 0);
}

int simuse_batch_get_one_kernel(simuse_batch* b, int32_t* on) {
    if (!b || !on) return fail(SIMUSE_E_INVALID, "null argument");
    *on = b->eng.one_kernel ? 1 : 0;
    return SIMUSE_OK;
}

int simuse_batch_set_graph_steps(simuse_batch* b, uint32_t steps) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.set_graph_steps(steps);
}

int simuse_batch_set_scatter_mode(simuse_batch* b, int32_t mode) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    if (mode == SIMUSE_SCATTER_AUTO) return b->eng.autotune(16);
    return b->eng.set_sched(mode);
}

int simuse_batch_set_block_tiles(simuse_batch* b, uint32_t tiles) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.set_group(tiles);
}

int simuse_batch_autotune(simuse_batch* b, uint32_t trial_steps, int32_t* mode) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    int rc = b->eng.autotune(trial_steps);
    if (mode) *mode = b->eng.sched;
    return rc;
}

int simuse_batch_step_async(simuse_batch* b, uint32_t n_steps) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.step_async(n_steps);
}

int simuse_batch_sync(simuse_batch* b) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.sync();
}

int simuse_batch_step_counter(simuse_batch* b, int32_t id, uint64_t* step) {
    BATCH_NET(b, id, s);
    if (step) *step = s->host_step;
    return SIMUSE_OK;
}

int simuse_batch_net_info(simuse_batch* b, int32_t id, uint64_t* n_total, uint64_t* n_syn, int32_t* n_pops) {
    BATCH_NET(b, id, s);
    if (n_total) *n_total = s->N;
    if (n_syn) *n_syn = s->n_syn;
    if (n_pops) *n_pops = static_cast<int32_t>(s->pops.size());
    return SIMUSE_OK;
}

int simuse_batch_pop_info(simuse_batch* b, int32_t id, int32_t pop, int32_t* model, uint64_t* n) {
    BATCH_NET(b, id, s);
    if (pop < 0 || pop >= static_cast<int32_t>(s->pops.size())) return fail(SIMUSE_E_INVALID, "pop out of range");
    if (model) *model = s->pops[pop].model;
    if (n) *n = s->pops[pop].n;
    return SIMUSE_OK;
}

int simuse_batch_download_state(simuse_batch* b, int32_t id, int32_t pop, float* out, uint64_t len) {
    BATCH_NET(b, id, s);
    return s->download_state(pop, out, len, b->eng.stream);
}

int simuse_batch_download_params(simuse_batch* b, int32_t id, int32_t pop, float* out, uint64_t len) {
    BATCH_NET(b, id, s);
    if (pop < 0 || pop >= static_cast<int32_t>(s->pops.size())) return fail(SIMUSE_E_INVALID, "pop out of range");
    const HostPop& P = s->pops[pop];
    if (!out || len != static_cast<uint64_t>(P.np) * P.host_stride)
        return fail(SIMUSE_E_INVALID, "params len must be np*stride");
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    std::vector<float> tmp(static_cast<size_t>(P.np) * P.dstride);
    HIP_TRY(hipMemcpy(tmp.data(), P.pr, 4 * tmp.size(), hipMemcpyDeviceToHost));
    std::memset(out, 0, 4 * len);
    for (uint32_t k = 0; k < P.np; ++k)
        for (uint32_t i = 0; i < P.n; ++i)
            out[static_cast<size_t>(k) * P.host_stride + i] =
                ((P.uni_mask >> k) & 1u) ? P.puni[k] : tmp[static_cast<size_t>(k) * P.dstride + i];
    return SIMUSE_OK;
}

int simuse_batch_download_csr(simuse_batch* b, int32_t id, uint32_t* row_ptr, uint32_t* col, float* w,
                              uint8_t* delay) {
    BATCH_NET(b, id, s);
    HIP_TRY(hipStreamSynchronize(b->eng.stream));
    if (row_ptr) HIP_TRY(hipMemcpy(row_ptr, s->row_ptr, 4 * (static_cast<size_t>(s->N) + 1), hipMemcpyDeviceToHost));
    if (s->n_syn) {
        if (col) HIP_TRY(hipMemcpy(col, s->col, 4 * s->n_syn, hipMemcpyDeviceToHost));
        if (w) HIP_TRY(hipMemcpy(w, s->w, 4 * s->n_syn, hipMemcpyDeviceToHost));
        if (delay) HIP_TRY(hipMemcpy(delay, s->delay, s->n_syn, hipMemcpyDeviceToHost));
    }
    return SIMUSE_OK;
}

int simuse_batch_download_spike_ids(simuse_batch* b, int32_t id, uint32_t* ids, uint32_t cap, uint32_t* count) {
    BATCH_NET(b, id, s);
    return s->download_spike_ids(ids, cap, count, b->eng.stream);
}

int simuse_batch_counters(simuse_batch* b, int32_t id, uint64_t* spikes, uint64_t* events) {
    BATCH_NET(b, id, s);
    return s->counters(spikes, events, b->eng.stream);
}

int simuse_batch_record(simuse_batch* b, int32_t id, uint32_t max_steps) {
    BATCH_NET(b, id, s);
    int rc = s->record(max_steps, b->eng.stream);
    b->eng.mark_dirty();
    return rc;
}

int simuse_batch_download_raster(simuse_batch* b, int32_t id, uint64_t* out, uint64_t n_words, uint32_t* steps,
                                 uint64_t* words_per_step) {
    BATCH_NET(b, id, s);
    if (words_per_step) *words_per_step = s->n_words;
    return s->download_raster(out, n_words, steps, b->eng.stream);
}

int simuse_batch_device_bytes(simuse_batch* b, uint64_t* bytes) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    if (bytes) *bytes = b->bytes_live;
    return SIMUSE_OK;
}

int simuse_batch_set_profiling(simuse_batch* b, int32_t on) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.set_profiling(on != 0);
}

int simuse_batch_kernel_times(simuse_batch* b, float* ms, int32_t n) {
    if (!b) return fail(SIMUSE_E_INVALID, "null batch handle");
    return b->eng.kernel_times(ms, n);
}

uint64_t simuse_gen_bytes(const simuse_gen_desc* g) { return g ? generated_bytes(*g) : 0; }

}  // extern "C"
