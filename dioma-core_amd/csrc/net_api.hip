// net_api.hip — per-network runtime (NetState), the autotuner, and the
// simuse_net_* C ABI of include/simuse.h.
//
// A simuse_net handle is an Engine with exactly one network: one HIP stream,
// all of the network's device buffers, and a hipGraph of `graph_steps` steps
// that is captured once and replayed.  The ring slot and the step counter
// live on the device and advance inside the graph, so asynchronous steps
// need no host round trip (the reference patches a host-written slot scalar
// before every replay, PC/src/network.rs:431-460).

#include <cmath>
#include <mutex>

#include "device_math.hpp"
#include "engine.hpp"

namespace simuse {

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

int validate_pops(const simuse_pop_desc* pops, int32_t n_pops, uint64_t* n_total) {
    if (n_pops <= 0 || !pops) return fail(SIMUSE_E_INVALID, "no populations");
    uint64_t tot = 0;
    for (int32_t p = 0; p < n_pops; ++p) {
        const simuse_pop_desc& d = pops[p];
        if (d.model < SIMUSE_MODEL_LIF || d.model > SIMUSE_MODEL_ALPHA_LIF)
            return fail(SIMUSE_E_UNSUPPORTED, "population %d: unknown model %d", p, d.model);
        if (d.ns != static_cast<uint32_t>(model_states(d.model)) ||
            d.np != static_cast<uint32_t>(model_params(d.model)))
            return fail(SIMUSE_E_INVALID, "population %d: model %d needs %d states and %d params (got %u, %u)", p,
                        d.model, model_states(d.model), model_params(d.model), d.ns, d.np);
        if (d.stride < d.n) return fail(SIMUSE_E_INVALID, "population %d: stride %u < n %u", p, d.stride, d.n);
        if (d.n > 0 && (!d.states || !d.params)) return fail(SIMUSE_E_INVALID, "population %d: null state/params", p);
        tot += d.n;
    }
    if (tot == 0) return fail(SIMUSE_E_INVALID, "CompiledNetwork: no neurons");
    // Ring cells are addressed as slot*N + post in 32 bits (32 slots).
    if (tot > kMaxNeurons)
        return fail(SIMUSE_E_INVALID, "total neuron count %llu exceeds %u per network",
                    static_cast<unsigned long long>(tot), kMaxNeurons);
    *n_total = tot;
    return SIMUSE_OK;
}

// SparseGraph::from_csr checks (P/src/api/net/graph.rs:217-280).
int validate_csr(uint64_t n_total, const uint32_t* row_ptr, const uint32_t* col, const float* w,
                 const uint8_t* delay, uint64_t n_syn) {
    if (!row_ptr) return fail(SIMUSE_E_INVALID, "null row_ptr");
    if (n_syn && (!col || !w || !delay)) return fail(SIMUSE_E_INVALID, "null CSR arrays");
    if (n_syn >= 0xFFFFFFFFull)
        return fail(SIMUSE_E_INVALID, "too many synapses (%llu) for a u32 CSR", static_cast<unsigned long long>(n_syn));
    if (row_ptr[0] != 0) return fail(SIMUSE_E_INVALID, "CSR row_ptr must start at 0");
    for (uint64_t i = 0; i < n_total; ++i)
        if (row_ptr[i] > row_ptr[i + 1])
            return fail(SIMUSE_E_INVALID, "CSR row_ptr must be monotonically non-decreasing: %u > %u at %llu",
                        row_ptr[i], row_ptr[i + 1], static_cast<unsigned long long>(i));
    if (row_ptr[n_total] != n_syn) return fail(SIMUSE_E_INVALID, "CSR row_ptr must end at col_idx.len()");
    for (uint64_t e = 0; e < n_syn; ++e) {
        if (col[e] >= n_total)
            return fail(SIMUSE_E_INVALID, "col index %u >= %llu", col[e], static_cast<unsigned long long>(n_total));
        if (!std::isfinite(w[e]))
            return fail(SIMUSE_E_INVALID, "weight %llu must be finite", static_cast<unsigned long long>(e));
        if (delay[e] >= kSlots) return fail(SIMUSE_E_INVALID, "delay %u >= MAX_DELAY_SLOTS (32)", delay[e]);
    }
    return SIMUSE_OK;
}

int NetState::init(const simuse_pop_desc* pd, int32_t n_pops, const uint32_t* rp, const uint32_t* c, const float* wf,
                   const uint8_t* dl, uint64_t nsyn, hipStream_t stream) {
    uint32_t off = 0, tile = 0, word = 0;
    for (int32_t p = 0; p < n_pops; ++p) {
        HostPop P;
        P.model = pd[p].model;
        P.n = pd[p].n;
        P.ns = pd[p].ns;
        P.np = pd[p].np;
        P.host_stride = pd[p].stride;
        P.off = off;
        P.tile0 = tile;
        P.ntiles = (P.n + kTile - 1) / kTile;
        P.word0 = word;
        P.dstride = (P.n + 63u) / 64u * 64u;
        if (P.dstride == 0) P.dstride = 64;
        off += P.n;
        tile += P.ntiles;
        word += P.ntiles * kWavesPerTile;
        pops.push_back(P);
    }
    N = off;
    n_tiles = tile;
    n_words = word;
    n_syn = nsyn;
    int rc;
    for (auto& P : pops) {
        if ((rc = alloc(&P.st, static_cast<size_t>(P.ns) * P.dstride))) return rc;
        if ((rc = alloc(&P.pr, static_cast<size_t>(P.np) * P.dstride))) return rc;
        HIP_TRY(hipMemset(P.st, 0, sizeof(float) * P.ns * P.dstride));
        HIP_TRY(hipMemset(P.pr, 0, sizeof(float) * P.np * P.dstride));
    }
    if ((rc = alloc(&d_pops, pops.size()))) return rc;
    if ((rc = upload_pops(pd))) return rc;
    std::vector<uint32_t> tile_pop(n_tiles);
    for (size_t p = 0; p < pops.size(); ++p)
        for (uint32_t t = 0; t < pops[p].ntiles; ++t) tile_pop[pops[p].tile0 + t] = static_cast<uint32_t>(p);
    if ((rc = alloc(&d_tile_pop, n_tiles))) return rc;
    if (n_tiles) HIP_TRY(hipMemcpy(d_tile_pop, tile_pop.data(), 4 * tile_pop.size(), hipMemcpyHostToDevice));
    if ((rc = alloc(&row_ptr, static_cast<size_t>(N) + 1))) return rc;
    if ((rc = alloc(&col, nsyn))) return rc;
    if ((rc = alloc(&w, nsyn))) return rc;
    if ((rc = alloc(&delay, nsyn))) return rc;
    HIP_TRY(hipMemcpy(row_ptr, rp, 4 * (static_cast<size_t>(N) + 1), hipMemcpyHostToDevice));
    if (nsyn) {
        HIP_TRY(hipMemcpy(col, c, 4 * nsyn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(w, wf, 4 * nsyn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(delay, dl, nsyn, hipMemcpyHostToDevice));
    }
    h_row_ptr.assign(rp, rp + N + 1);
    // every row with the same out-degree: the kernels compute row extents
    uni_deg = N && n_syn % N == 0 && n_syn / N > 0 && n_syn / N < (1ull << 20) ? static_cast<uint32_t>(n_syn / N) : 0u;
    for (uint32_t i = 0; i <= N && uni_deg; ++i)
        if (rp[i] != static_cast<uint64_t>(i) * uni_deg) uni_deg = 0u;
    max_delay = 0;
    for (uint64_t e = 0; e < nsyn; ++e) max_delay = std::max<uint32_t>(max_delay, dl[e]);
    if ((rc = alloc(&ring, static_cast<size_t>(kSlots) * N))) return rc;
    if ((rc = alloc(&tile_count, n_tiles))) return rc;
    if ((rc = alloc(&active, static_cast<size_t>(n_tiles) * kTile))) return rc;
    if ((rc = alloc(&bits, n_words))) return rc;
    if ((rc = alloc(&tile_spikes, n_tiles))) return rc;
    if ((rc = alloc(&step, 2))) return rc;
    if ((rc = reset_dynamic(stream))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return SIMUSE_OK;
}

int NetState::finalize(hipStream_t stream) {
    if (finalized) return SIMUSE_OK;
    n_src = static_cast<uint32_t>(srcs.size());
    n_src_tiles = (n_src + kTile - 1) / kTile;
    std::vector<uint32_t> rp(n_src + 1, 0), post, sseed;
    std::vector<float> sw, sprob;
    std::vector<uint8_t> sd;
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
    const uint64_t n_ev = n_syn + n_src_edges;
    const uint32_t n_blk = n_tiles + n_src_tiles;
    // Every delay <= 1 step (synapses and sources): the binned scatter;
    // otherwise the chained one with its tagged heads and event records.
    const uint32_t n_bins_t = (n_tiles + kBinTiles - 1) / kBinTiles;
    bool one_model = true;
    for (const HostPop& P : pops) one_model = one_model && P.model == pops[0].model;
    binned = max_delay <= 1 && n_tiles > 0 && n_bins_t <= kMaxBins && pops.size() <= kMaxBinPops;
    // First event id of every scatter tile's rows: a tile's events, and so its
    // collisions, fit in [tile_ebase, next tile's tile_ebase).
    std::vector<uint32_t> eb(n_blk + 1);
    for (const HostPop& P : pops)
        for (uint32_t t = 0; t < P.ntiles; ++t) eb[P.tile0 + t] = h_row_ptr[P.off + t * kTile];
    for (uint32_t j = 0; j < n_src_tiles; ++j) eb[n_tiles + j] = static_cast<uint32_t>(n_syn + rp[j * kTile]);
    eb[n_blk] = static_cast<uint32_t>(n_ev);
    if (binned) {
        // 8-byte records need every writer's events below 2^kRecFBits: every
        // writer the engine forms is a run of at most W consecutive scatter
        // tiles (a bin, a group of neuron tiles, a bin_kernel writer)
        const uint32_t W = std::max({kBinTiles, kBinBlockTiles, (n_blk + kMaxWriters - 1) / kMaxWriters});
        uint64_t span = 0;
        for (uint32_t b = 0; b < n_blk; ++b) span = std::max<uint64_t>(span, eb[std::min(b + W, n_blk)] - eb[b]);
        ev_wide = !h_rec_code.empty() || span > (1ull << kRecFBits);
        n_bins = n_bins_t;
        const uint32_t ns = (n_src_tiles + kBinBlockTiles - 1) / kBinBlockTiles;
        // writers: at most one per neuron tile plus one per kBinBlockTiles source tiles
        bin_rows = std::min(kMaxWriters, n_tiles + ns);
        bin_rows += bin_rows & 1u;  // even: consumers read two writers' headers at once
        // one kernel per step needs one model, no pair-based plasticity and a
        // writer per bin
        mega_ok = one_model && !stdp && !ev_wide && !coupled() && n_bins + ns <= kMaxWriters;
        mega_model = pops[0].model;
        // the split form takes the rest of the one-model networks (pair-based
        // plasticity, receptor synapses, wide records)
        split_ok = one_model && !mega_ok && !coupled();
        const size_t nb = nbuf();
        if ((rc = alloc(&bin_ev, nb * n_ev * (ev_wide ? 2 : 1)))) return rc;
        if ((rc = alloc(&bin_hdr, nb * n_bins * bin_rows))) return rc;
        HIP_TRY(hipMemsetAsync(bin_hdr, 0, sizeof(uint2) * nb * n_bins * bin_rows, stream));
        if ((rc = alloc(&d_dense_writers, bin_rows + kXcdPad))) return rc;  // XCD-ordered, padded
        if ((rc = alloc(&d_mat_items, n_bins + kXcdPad))) return rc;
        if ((rc = alloc(&bin_slot, 1))) return rc;
        if ((rc = alloc(&bin_done, 1))) return rc;
        if ((rc = alloc(&bin_mat, 1))) return rc;
        HIP_TRY(hipMemsetAsync(bin_done, 0, sizeof(uint32_t), stream));
        // before the first step the ring holds the step's input (zeros, or
        // currents injected so far): the megastep reads it from there
        HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(bin_mat), 1, 1, stream));
    } else {
        if ((rc = alloc(&head, static_cast<size_t>(kSlots) * N))) return rc;
        if ((rc = alloc(&rec, n_ev))) return rc;
        if ((rc = alloc(&coll, n_ev))) return rc;
        clear_heads(stream);
    }
    if ((rc = alloc(&tile_coll, n_blk))) return rc;
    if ((rc = alloc(&tile_ebase, n_blk))) return rc;
    HIP_TRY(hipMemsetAsync(tile_coll, 0, sizeof(uint32_t) * n_blk, stream));
    if (n_blk) HIP_TRY(hipMemcpy(tile_ebase, eb.data(), 4 * static_cast<size_t>(n_blk), hipMemcpyHostToDevice));
    n_fix = std::max<uint32_t>(1u, (n_blk + kFixTiles - 1) / kFixTiles);
    if (!h_tm_type.empty()) {
        h_tm_type.resize(n_ev, 0);  // source edges added after set_stp are static
        if ((rc = alloc(&tm_type, n_ev))) return rc;
        if ((rc = alloc(&tm_r, n_ev))) return rc;
        if ((rc = alloc(&tm_u, n_ev))) return rc;
        if ((rc = alloc(&tm_last, n_ev))) return rc;
        if (!binned && (rc = alloc(&ev_w, n_ev))) return rc;  // the fix-up re-reads delivered weights
        if ((rc = alloc(&tm_prof, h_tm_prof.size()))) return rc;
        HIP_TRY(hipMemcpy(tm_type, h_tm_type.data(), n_ev, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tm_prof, h_tm_prof.data(), sizeof(float4) * h_tm_prof.size(), hipMemcpyHostToDevice));
        launch_tm_init(tm_type, tm_prof, tm_r, tm_u, tm_last, n_ev, stream);
        HIP_TRY(hipGetLastError());
    }
    if (!h_rec_code.empty()) {
        // conductance jumps land in the step after the spike, like a delay of
        // one; the chained scatter (longer delays) has no receptor form
        if (!binned)
            return fail(SIMUSE_E_UNSUPPORTED,
                        "receptor synapses need every synapse and source delay <= 1 step and at most %u "
                        "populations (have max delay %u)",
                        kMaxBinPops, max_delay);
        if ((rc = alloc(&rec_code, n_syn))) return rc;
        if ((rc = alloc(&g, static_cast<size_t>(kReceptors) * N))) return rc;
        if (n_syn) HIP_TRY(hipMemcpy(rec_code, h_rec_code.data(), n_syn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(g, 0, sizeof(float) * kReceptors * N, stream));
        if (!h_g_ext.empty()) {
            if ((rc = alloc(&g_ext, N))) return rc;
            HIP_TRY(hipMemcpy(g_ext, h_g_ext.data(), 4 * static_cast<size_t>(N), hipMemcpyHostToDevice));
        }
    }
    if (coupled() && (rc = finalize_coupling(stream))) return rc;
    if ((rc = alloc(&tile_events, static_cast<size_t>(n_tiles) + n_src_tiles))) return rc;
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
    HIP_TRY(hipStreamSynchronize(stream));
    finalized = true;
    return SIMUSE_OK;
}

// Electrical coupling: per neuron, its contributions in the order the
// reference adds them (step.rs:55-104): every gap junction in call order
// (both ends), every graded synapse in call order, then the sigmoid graded
// synapses profile by profile, each profile's connections in call order.
int NetState::finalize_coupling(hipStream_t stream) {
    if (!h_rec_code.empty())
        return fail(SIMUSE_E_UNSUPPORTED, "electrical coupling and receptor synapses in one network");
    for (const HostPop& P : pops)
        if (P.flags & kPopDrive)
            return fail(SIMUSE_E_UNSUPPORTED,
                        "electrical coupling with the fused Poisson drive (use add_poisson_source instead)");
    if (pops.size() > kCplPopMask)
        return fail(SIMUSE_E_UNSUPPORTED, "electrical coupling: more than %u populations", kCplPopMask);
    auto pop_of = [&](uint32_t gi) -> uint32_t {
        uint32_t p = 0;
        while (p + 1 < pops.size() && gi >= pops[p + 1].off) ++p;
        return p;
    };
    auto entry = [&](uint32_t other, uint32_t kind, uint32_t prof, float gw, uint32_t conn) -> uint4 {
        const uint32_t p = pop_of(other);
        uint32_t gb;
        std::memcpy(&gb, &gw, 4);
        return make_uint4(other - pops[p].off, p | (kind << 16) | (prof << 18), gb, conn);
    };
    std::vector<std::vector<uint4>> per(N);
    for (const CplHost& c : h_gap) {
        per[c.a].push_back(entry(c.b, 0u, 0u, c.g, 0u));
        per[c.b].push_back(entry(c.a, 1u, 0u, c.g, 0u));
    }
    for (const CplHost& c : h_graded) per[c.b].push_back(entry(c.a, 2u, 0u, c.g, 0u));
    n_sig = 0;
    for (uint32_t q = 0; q < h_sig_prof.size(); ++q)
        for (const SigConnHost& c : h_sig_conn)
            if (c.prof == q) per[c.post].push_back(entry(c.pre, 3u, q, c.w, n_sig++));
    std::vector<uint32_t> ptr(static_cast<size_t>(N) + 1, 0);
    std::vector<uint4> flat;
    for (uint32_t i = 0; i < N; ++i) {
        ptr[i + 1] = ptr[i] + static_cast<uint32_t>(per[i].size());
        flat.insert(flat.end(), per[i].begin(), per[i].end());
    }
    int rc;
    if ((rc = alloc(&cp_ptr, static_cast<size_t>(N) + 1))) return rc;
    if ((rc = alloc(&cp, flat.size()))) return rc;
    if ((rc = alloc(&sig_s, n_sig))) return rc;
    if ((rc = alloc(&sig_prof, static_cast<size_t>(kSigProf) * h_sig_prof.size()))) return rc;
    HIP_TRY(hipMemcpy(cp_ptr, ptr.data(), 4 * ptr.size(), hipMemcpyHostToDevice));
    if (!flat.empty()) HIP_TRY(hipMemcpy(cp, flat.data(), sizeof(uint4) * flat.size(), hipMemcpyHostToDevice));
    if (!h_sig_prof.empty())
        HIP_TRY(hipMemcpy(sig_prof, h_sig_prof.data(), sizeof(float) * kSigProf * h_sig_prof.size(),
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(sig_s, 0, sizeof(float) * std::max<uint32_t>(n_sig, 1u), stream));
    return SIMUSE_OK;
}

int NetState::add_coupling(int kind, const uint32_t* a, const uint32_t* b, const float* gv, uint32_t count) {
    const char* what = kind == 0 ? "connect_gap_junction" : "connect_graded_synapse";
    if (finalized) return fail(SIMUSE_E_STATE, "%s must be called before the first step", what);
    if (count && (!a || !b || !gv)) return fail(SIMUSE_E_INVALID, "%s: null arrays", what);
    for (uint32_t i = 0; i < count; ++i) {
        if (a[i] >= N || b[i] >= N)
            return fail(SIMUSE_E_INVALID, "%s: neuron index out of bounds (%u, %u of %u)", what, a[i], b[i], N);
        if (!(std::isfinite(gv[i]) && gv[i] >= 0.0f))
            return fail(SIMUSE_E_INVALID, "%s: conductance must be finite and >= 0, got %g", what, gv[i]);
    }
    std::vector<CplHost>& L = kind == 0 ? h_gap : h_graded;
    for (uint32_t i = 0; i < count; ++i) L.push_back(CplHost{a[i], b[i], gv[i]});
    return SIMUSE_OK;
}

int NetState::add_sigmoid_profile(float vth, float delta, float k, float erev, float cond, float dt_ms, int32_t* id) {
    if (finalized) return fail(SIMUSE_E_STATE, "register_sigmoid_graded must be called before the first step");
    if (!(std::isfinite(dt_ms) && dt_ms > 0.0f))
        return fail(SIMUSE_E_INVALID, "register_sigmoid_graded: dt_ms must be finite and > 0, got %g", dt_ms);
    // SigmoidGradedSynapse::validate, P/src/api/net/network/sources.rs:71-86
    const float p[5] = {vth, delta, k, erev, cond};
    for (float x : p)
        if (!std::isfinite(x)) return fail(SIMUSE_E_INVALID, "SigmoidGradedSynapse: parameters must be finite");
    if (delta == 0.0f) return fail(SIMUSE_E_INVALID, "SigmoidGradedSynapse: delta must be non-zero");
    if (!(k > 0.0f)) return fail(SIMUSE_E_INVALID, "SigmoidGradedSynapse: k must be > 0");
    if (cond < 0.0f) return fail(SIMUSE_E_INVALID, "SigmoidGradedSynapse: conductance must be >= 0");
    if (h_sig_prof.size() >= (1u << 14)) return fail(SIMUSE_E_UNSUPPORTED, "register_sigmoid_graded: 16384 profiles");
    h_sig_prof.push_back({vth, 1.0f / delta, k, erev, cond, dt_ms});
    if (id) *id = static_cast<int32_t>(h_sig_prof.size() - 1);
    return SIMUSE_OK;
}

int NetState::add_sigmoid_conns(int32_t prof, const uint32_t* pre, const uint32_t* post, const float* wv,
                                uint32_t count) {
    if (finalized) return fail(SIMUSE_E_STATE, "connect_sigmoid_graded must be called before the first step");
    if (prof < 0 || static_cast<size_t>(prof) >= h_sig_prof.size())
        return fail(SIMUSE_E_INVALID, "sigmoid graded profile %d out of range (have %zu)", prof, h_sig_prof.size());
    if (count && (!pre || !post || !wv)) return fail(SIMUSE_E_INVALID, "connect_sigmoid_graded: null arrays");
    for (uint32_t i = 0; i < count; ++i) {
        if (pre[i] >= N || post[i] >= N)
            return fail(SIMUSE_E_INVALID, "connect_sigmoid_graded: neuron index out of bounds (%u, %u of %u)", pre[i],
                        post[i], N);
        if (!std::isfinite(wv[i]))
            return fail(SIMUSE_E_INVALID, "connect_sigmoid_graded: weight must be finite, got %g", wv[i]);
    }
    for (uint32_t i = 0; i < count; ++i)
        h_sig_conn.push_back(SigConnHost{static_cast<uint32_t>(prof), pre[i], post[i], wv[i]});
    return SIMUSE_OK;
}

int NetState::upload_pops(const simuse_pop_desc* pd) {
    for (size_t p = 0; p < pops.size(); ++p) {
        HostPop& P = pops[p];
        const uint32_t hs = pd[p].stride;
        if (P.n == 0) continue;
        HIP_TRY(hipMemcpy2D(P.st, sizeof(float) * P.dstride, pd[p].states, sizeof(float) * hs, sizeof(float) * P.n,
                            P.ns, hipMemcpyHostToDevice));
        // Uniform parameter columns become scalars; column 0 (dt) is always
        // taken from neuron 0 (population.rs:307,361).
        P.uni_mask = 0;
        for (uint32_t k = 0; k < P.np && k < static_cast<uint32_t>(kMaxParams); ++k) {
            const float* c = pd[p].params + static_cast<size_t>(k) * hs;
            bool uni = true;
            for (uint32_t i = 1; i < P.n && uni; ++i) uni = std::memcmp(&c[i], &c[0], 4) == 0;
            P.puni[k] = c[0];
            if (uni || k == 0) P.uni_mask |= 1u << k;
        }
        HIP_TRY(hipMemcpy2D(P.pr, sizeof(float) * P.dstride, pd[p].params, sizeof(float) * hs, sizeof(float) * P.n,
                            P.np, hipMemcpyHostToDevice));
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

int NetState::upload_pop_table() {
    std::vector<PopDev> pd(pops.size());
    for (size_t p = 0; p < pops.size(); ++p) {
        const HostPop& P = pops[p];
        PopDev& D = pd[p];
        std::memset(&D, 0, sizeof(D));
        D.model = P.model;
        D.n = P.n;
        D.off = P.off;
        D.tile0 = P.tile0;
        D.word0 = P.word0;
        D.dstride = P.dstride;
        D.uni_mask = P.uni_mask;
        D.flags = P.flags;
        D.st = dev_ptr(P.st);
        D.pr = dev_ptr(P.pr);
        D.sigma = dev_ptr(P.sigma);
        D.sigma_uni = P.sigma_uni;
        D.drive_p = P.drive_p;
        D.drive_w = P.drive_w;
        D.drive_seed0 = P.drive_seed0;
        D.seed = P.seed;
        std::memcpy(D.puni, P.puni, sizeof(D.puni));
        // DecayFactors::new(dt), P/src/kernels/synapse/kinetics.rs:24-38
        for (int r = 0; r < kReceptors; ++r) D.gdec[r] = fast_expf(-P.puni[0] / kRecTau[r]);
    }
    HIP_TRY(hipMemcpy(d_pops, pd.data(), sizeof(PopDev) * pd.size(), hipMemcpyHostToDevice));
    return SIMUSE_OK;
}

int NetState::reset_dynamic(hipStream_t stream) {
    HIP_TRY(hipMemsetAsync(ring, 0, sizeof(float) * static_cast<size_t>(kSlots) * N, stream));
    HIP_TRY(hipMemsetAsync(step, 0, sizeof(uint64_t) * 2, stream));
    clear_heads(stream);
    if (tile_coll) HIP_TRY(hipMemsetAsync(tile_coll, 0, sizeof(uint32_t) * (n_tiles + n_src_tiles), stream));
    if (bin_hdr)
        HIP_TRY(hipMemsetAsync(bin_hdr, 0, sizeof(uint2) * nbuf() * n_bins * bin_rows, stream));
    if (bin_done) HIP_TRY(hipMemsetAsync(bin_done, 0, sizeof(uint32_t), stream));
    if (bin_mat) HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(bin_mat), 1, 1, stream));  // zeroed ring
    HIP_TRY(hipMemsetAsync(tile_count, 0, sizeof(uint32_t) * n_tiles, stream));
    HIP_TRY(hipMemsetAsync(bits, 0, sizeof(uint64_t) * n_words, stream));
    HIP_TRY(hipMemsetAsync(tile_spikes, 0, sizeof(uint64_t) * n_tiles, stream));
    if (tile_events) HIP_TRY(hipMemsetAsync(tile_events, 0, sizeof(uint64_t) * (n_tiles + n_src_tiles), stream));
    if (g) HIP_TRY(hipMemsetAsync(g, 0, sizeof(float) * kReceptors * N, stream));
    if (n_sig) HIP_TRY(hipMemsetAsync(sig_s, 0, sizeof(float) * n_sig, stream));
    if (stdp) {
        HIP_TRY(hipMemsetAsync(tr_pre, 0, sizeof(float) * N, stream));
        HIP_TRY(hipMemsetAsync(tr_post, 0, sizeof(float) * N, stream));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    host_step = 0;
    return SIMUSE_OK;
}

int NetState::add_source(float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts, const float* wf,
                         const uint8_t* dl, uint32_t count) {
    if (finalized) return fail(SIMUSE_E_STATE, "add_poisson_source must be called before the first step");
    if (!(std::isfinite(rate_hz) && rate_hz >= 0.0f))
        return fail(SIMUSE_E_INVALID, "add_poisson_source: rate_hz must be finite and >= 0, got %g", rate_hz);
    if (!(std::isfinite(dt_ms) && dt_ms > 0.0f))
        return fail(SIMUSE_E_INVALID, "add_poisson_source: dt_ms must be finite and > 0, got %g", dt_ms);
    if (count && (!posts || !wf || !dl)) return fail(SIMUSE_E_INVALID, "null source arrays");
    SourceHost s;
    s.p = std::min(std::max(rate_hz * dt_ms / 1000.0f, 0.0f), 1.0f);  // core.rs:301
    s.seed = seed;
    for (uint32_t i = 0; i < count; ++i) {
        if (posts[i] >= N) return fail(SIMUSE_E_INVALID, "post index %u out of bounds", posts[i]);
        if (!std::isfinite(wf[i])) return fail(SIMUSE_E_INVALID, "connect_poisson_source: weight %u must be finite", i);
        if (dl[i] >= kSlots)
            return fail(SIMUSE_E_INVALID, "connect_poisson_source: delay %u >= MAX_DELAY_SLOTS (32)", dl[i]);
    }
    s.post.assign(posts, posts + count);
    s.w.assign(wf, wf + count);
    s.d.assign(dl, dl + count);
    for (uint32_t i = 0; i < count; ++i) max_delay = std::max<uint32_t>(max_delay, dl[i]);
    srcs.push_back(std::move(s));
    return SIMUSE_OK;
}

// Network::register_plasticity (P/src/api/net/network/core.rs:597-619) for
// every profile plus the per-edge profile ids (Edge::with_synapse_plasticity;
// source edges in add order).  a_rec = e^(-dt/tau_rec) in FP32, as the
// reference computes it.
int NetState::set_stp(const simuse_stp_profile* profiles, int32_t n_profiles, float dt, const uint8_t* syn_types,
                      const uint8_t* src_types, uint64_t n_src_edges_given) {
    if (finalized) return fail(SIMUSE_E_STATE, "short-term plasticity must be set before the first step");
    if (n_profiles < 1 || n_profiles > 255 || !profiles)
        return fail(SIMUSE_E_INVALID, "short-term plasticity: 1..255 profiles required");
    if (!(std::isfinite(dt) && dt > 0.0f)) return fail(SIMUSE_E_INVALID, "register_plasticity: dt must be > 0");
    if (!syn_types && n_syn) return fail(SIMUSE_E_INVALID, "short-term plasticity: null syn_types");
    uint64_t n_src = 0;
    for (const auto& s : srcs) n_src += s.post.size();
    if (src_types && n_src_edges_given != n_src)
        return fail(SIMUSE_E_INVALID, "short-term plasticity: %llu source edge types for %llu source edges",
                    static_cast<unsigned long long>(n_src_edges_given), static_cast<unsigned long long>(n_src));
    std::vector<float4> prof;
    for (int32_t k = 0; k < n_profiles; ++k) {
        const simuse_stp_profile& p = profiles[k];
        if (!std::isfinite(p.u0) || p.u0 <= 0.0f || p.u0 > 1.0f)
            return fail(SIMUSE_E_INVALID, "profile %d: init_release_prob must be in (0, 1], got %g", k, p.u0);
        if (!std::isfinite(p.tau_rec) || p.tau_rec <= 0.0f)
            return fail(SIMUSE_E_INVALID, "profile %d: tau_rec must be finite and > 0, got %g", k, p.tau_rec);
        if (!std::isfinite(p.tau_fac))
            return fail(SIMUSE_E_INVALID, "profile %d: tau_fac must be finite (<= 0: none)", k);
        const bool fac = p.tau_fac > 0.0f;
        prof.push_back(make_float4(p.u0, std::exp(-dt / p.tau_rec), fac ? std::exp(-dt / p.tau_fac) : 0.0f,
                                   fac ? 1.0f : 0.0f));
    }
    std::vector<uint8_t> types(n_syn + n_src, 0);
    for (uint64_t e = 0; e < n_syn; ++e) types[e] = syn_types[e];
    if (src_types)
        for (uint64_t e = 0; e < n_src; ++e) types[n_syn + e] = src_types[e];
    for (uint64_t e = 0; e < types.size(); ++e)
        if (types[e] > n_profiles)
            return fail(SIMUSE_E_INVALID, "edge %llu: plasticity profile %u not registered",
                        static_cast<unsigned long long>(e), types[e]);
    h_tm_prof = std::move(prof);
    h_tm_type = std::move(types);
    return SIMUSE_OK;
}

// Receptor types per synapse (ReceptorType::from_u8, P/src/kernels/synapse/
// receptor.rs:64-76: 0..6, anything else AMPA), stored as 1 + type; 255
// keeps a synapse current-based (0).  Jumps scale by `scale`
// (csr_propagate_conductance_from_queue, CN/src/propagate.rs:225-290).
int NetState::set_receptors(const uint8_t* types, float scale, const float* ext) {
    if (finalized) return fail(SIMUSE_E_STATE, "receptor synapses must be set before the first step");
    if (!types && n_syn) return fail(SIMUSE_E_INVALID, "receptor synapses: null receptor_types");
    if (!std::isfinite(scale)) return fail(SIMUSE_E_INVALID, "receptor synapses: scale must be finite");
    std::vector<uint8_t> code(n_syn);
    for (uint64_t e = 0; e < n_syn; ++e) {
        const uint8_t r = types[e];
        code[e] = r == 255 ? 0 : r < kReceptors ? static_cast<uint8_t>(r + 1) : 1;
    }
    std::vector<float> x;
    if (ext) {
        for (uint32_t i = 0; i < N; ++i)
            if (!std::isfinite(ext[i]))
                return fail(SIMUSE_E_INVALID, "receptor synapses: external current %u must be finite", i);
        x.assign(ext, ext + N);
    }
    rec_live = 0;
    for (const uint8_t c : code)
        if (c) rec_live |= 1u << (c - 1);
    h_rec_code = std::move(code);
    if (h_rec_code.empty()) h_rec_code.assign(1, 0);  // no synapses: still on (decay, external current)
    h_g_ext = std::move(x);
    rec_scale = scale;
    return SIMUSE_OK;
}

int NetState::download_conductances(float* out, uint64_t len, hipStream_t stream) {
    if (!out || len != static_cast<uint64_t>(kReceptors) * N)
        return fail(SIMUSE_E_INVALID, "conductances len must be 7*n_total");
    if (!g) {
        if (h_rec_code.empty()) return fail(SIMUSE_E_STATE, "receptor synapses are off");
        std::memset(out, 0, sizeof(float) * len);  // not stepped yet
        return SIMUSE_OK;
    }
    int rc;
    if (eng && (rc = eng->materialize(this))) return rc;  // the split form's pending jumps into g
    HIP_TRY(hipMemcpyAsync(out, g, sizeof(float) * len, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::set_noise(int32_t pop, const float* sigma, uint64_t seed) {
    if (pop < 0 || pop >= static_cast<int32_t>(pops.size())) return fail(SIMUSE_E_INVALID, "pop %d out of range", pop);
    HostPop& P = pops[pop];
    int rc;
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
            if (!P.sigma && (rc = alloc(&P.sigma, P.n))) return rc;
            HIP_TRY(hipMemcpy(P.sigma, sigma, 4 * static_cast<size_t>(P.n), hipMemcpyHostToDevice));
        }
        P.seed = seed;
        P.flags |= kPopNoise;
    }
    return upload_pop_table();
}

int NetState::set_drive(int32_t pop, float rate_hz, float dt_ms, float weight, uint32_t seed0) {
    if (pop < 0 || pop >= static_cast<int32_t>(pops.size())) return fail(SIMUSE_E_INVALID, "pop %d out of range", pop);
    if (!(std::isfinite(rate_hz) && rate_hz >= 0.0f) || !(std::isfinite(dt_ms) && dt_ms > 0.0f) ||
        !std::isfinite(weight))
        return fail(SIMUSE_E_INVALID, "poisson drive: rate >= 0, dt > 0 and a finite weight are required");
    HostPop& P = pops[pop];
    P.drive_p = std::min(std::max(rate_hz * dt_ms / 1000.0f, 0.0f), 1.0f);  // core.rs:301
    P.drive_w = weight;
    P.drive_seed0 = seed0;
    if (rate_hz > 0.0f) P.flags |= kPopDrive;
    else P.flags &= ~kPopDrive;
    return upload_pop_table();
}

int NetState::set_plasticity(const simuse_stdp_params* p, const uint8_t* pl, hipStream_t stream) {
    if (!p) {
        stdp = false;
        return SIMUSE_OK;
    }
    if (!(p->tau_plus > 0 && p->tau_minus > 0 && std::isfinite(p->a_plus) && std::isfinite(p->a_minus) &&
          p->w_min <= p->w_max))
        return fail(SIMUSE_E_INVALID, "invalid STDP parameters");
    int rc;
    const uint64_t ns = n_syn;
    HIP_TRY(hipStreamSynchronize(stream));
    if (!w64) {
        if ((rc = alloc(&w64, ns))) return rc;
        if ((rc = alloc(&tr_pre, N))) return rc;
        if ((rc = alloc(&tr_post, N))) return rc;
        if ((rc = alloc(&in_ptr, static_cast<size_t>(N) + 1))) return rc;
        if ((rc = alloc(&in_edge, ns))) return rc;
        if ((rc = alloc(&edge_pre, ns))) return rc;
        // in-edge lists (ascending edge id per post) and edge -> pre
        std::vector<uint32_t> colh(ns);
        if (ns) HIP_TRY(hipMemcpy(colh.data(), col, 4 * ns, hipMemcpyDeviceToHost));
        std::vector<uint32_t> ip(static_cast<size_t>(N) + 1, 0), ie(ns), ep(ns);
        for (uint64_t e = 0; e < ns; ++e) ip[colh[e] + 1]++;
        for (uint32_t i = 0; i < N; ++i) ip[i + 1] += ip[i];
        std::vector<uint32_t> cur(ip.begin(), ip.end() - 1);
        for (uint32_t i = 0; i < N; ++i)
            for (uint32_t e = h_row_ptr[i]; e < h_row_ptr[i + 1]; ++e) ep[e] = i;
        for (uint64_t e = 0; e < ns; ++e) ie[cur[colh[e]]++] = static_cast<uint32_t>(e);
        HIP_TRY(hipMemcpy(in_ptr, ip.data(), 4 * ip.size(), hipMemcpyHostToDevice));
        if (ns) {
            HIP_TRY(hipMemcpy(in_edge, ie.data(), 4 * ns, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(edge_pre, ep.data(), 4 * ns, hipMemcpyHostToDevice));
        }
    }
    if (pl) {
        if (!plastic && (rc = alloc(&plastic, ns))) return rc;
        if (ns) HIP_TRY(hipMemcpy(plastic, pl, ns, hipMemcpyHostToDevice));
    } else if (plastic) {
        std::vector<uint8_t> ones(ns, 1);
        if (ns) HIP_TRY(hipMemcpy(plastic, ones.data(), ns, hipMemcpyHostToDevice));
    }
    launch_widen(w, w64, ns, stream);
    HIP_TRY(hipMemsetAsync(tr_pre, 0, 4 * static_cast<size_t>(N), stream));
    HIP_TRY(hipMemsetAsync(tr_post, 0, 4 * static_cast<size_t>(N), stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    sp = *p;
    // per-step trace decay e^{-dt/tau}, dt of population 0
    const double dt = static_cast<double>(pops[0].puni[0]);
    stdp_dpre = static_cast<float>(std::exp(-dt / static_cast<double>(p->tau_plus)));
    stdp_dpost = static_cast<float>(std::exp(-dt / static_cast<double>(p->tau_minus)));
    stdp = true;
    return SIMUSE_OK;
}

int NetState::download_spikes(int32_t* out, uint64_t len, hipStream_t stream) {
    if (!out || len != N)
        return fail(SIMUSE_E_INVALID, "spikes len %llu != n_total %u", static_cast<unsigned long long>(len), N);
    std::vector<uint64_t> words(n_words);
    HIP_TRY(hipMemcpyAsync(words.data(), bits, 8 * words.size(), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    for (const auto& P : pops)
        for (uint32_t i = 0; i < P.n; ++i)
            out[P.off + i] = static_cast<int32_t>((words[P.word0 + i / 64] >> (i % 64)) & 1ull);
    return SIMUSE_OK;
}

int NetState::download_spike_ids(uint32_t* ids, uint32_t cap, uint32_t* count, hipStream_t stream) {
    std::vector<uint32_t> cnt(n_tiles);
    HIP_TRY(hipMemcpyAsync(cnt.data(), tile_count, 4 * cnt.size(), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    uint32_t total = 0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
        const uint32_t c = cnt[t];
        if (c && ids && total < cap) {
            const uint32_t take = std::min(c, cap - total);
            HIP_TRY(hipMemcpyAsync(ids + total, active + static_cast<size_t>(t) * kTile, 4 * take,
                                   hipMemcpyDeviceToHost, stream));
        }
        total += c;
    }
    HIP_TRY(hipStreamSynchronize(stream));
    if (count) *count = total;
    return SIMUSE_OK;
}

int NetState::download_state(int32_t pop, float* out, uint64_t len, hipStream_t stream) {
    if (pop < 0 || pop >= static_cast<int32_t>(pops.size()))
        return fail(SIMUSE_E_INVALID, "CompiledNetwork: pop %d out of range", pop);
    const HostPop& P = pops[pop];
    if (!out || len != static_cast<uint64_t>(P.ns) * P.host_stride)
        return fail(SIMUSE_E_INVALID, "states len %llu != ns*stride %llu", static_cast<unsigned long long>(len),
                    static_cast<unsigned long long>(P.ns) * P.host_stride);
    std::memset(out, 0, sizeof(float) * len);
    if (P.n == 0) return SIMUSE_OK;
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(float) * P.host_stride, P.st, sizeof(float) * P.dstride, sizeof(float) * P.n,
                             P.ns, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::download_ring(float* out, uint64_t len, hipStream_t stream) {
    if (!out || len != static_cast<uint64_t>(kSlots) * N) return fail(SIMUSE_E_INVALID, "ring len must be 32*n_total");
    int rc;
    if (eng && (rc = eng->materialize(this))) return rc;
    HIP_TRY(hipMemcpyAsync(out, ring, 4 * len, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::upload_weights(const float* wf, uint64_t n, hipStream_t stream) {
    if (n != n_syn || (n && !wf)) return fail(SIMUSE_E_INVALID, "weights len mismatch");
    for (uint64_t e = 0; e < n; ++e)
        if (!std::isfinite(wf[e]))
            return fail(SIMUSE_E_INVALID, "weight %llu must be finite", static_cast<unsigned long long>(e));
    HIP_TRY(hipStreamSynchronize(stream));
    if (n) HIP_TRY(hipMemcpy(w, wf, 4 * n, hipMemcpyHostToDevice));
    if (w64) {
        launch_widen(w, w64, n, stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream));
    }
    return SIMUSE_OK;
}

int NetState::download_weights(float* wf, uint64_t n, hipStream_t stream) {
    if (n != n_syn || (n && !wf)) return fail(SIMUSE_E_INVALID, "weights len mismatch");
    if (n) HIP_TRY(hipMemcpyAsync(wf, w, 4 * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::inject(int32_t pop, const uint32_t* idx, const float* val, uint32_t count, hipStream_t stream) {
    if (pop < 0 || pop >= static_cast<int32_t>(pops.size()))
        return fail(SIMUSE_E_INVALID, "inject_current: population %d out of range (have %zu)", pop, pops.size());
    if (count == 0) return SIMUSE_OK;
    if (!idx || !val) return fail(SIMUSE_E_INVALID, "inject_current: null arrays");
    const HostPop& P = pops[pop];
    std::vector<uint32_t> g(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (idx[i] >= P.n)
            return fail(SIMUSE_E_INVALID, "inject_current: neuron index %u out of bounds (population has %u)", idx[i],
                        P.n);
        if (!std::isfinite(val[i])) return fail(SIMUSE_E_INVALID, "inject_current: value %u must be finite", i);
        g[i] = P.off + idx[i];
    }
    std::vector<uint32_t> sorted(g);
    std::sort(sorted.begin(), sorted.end());
    const bool dup = std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end();
    int rc;
    if (eng && (rc = eng->materialize(this))) return rc;  // injection adds to the slot's event sums
    HIP_TRY(hipStreamSynchronize(stream));
    if (count > inj_cap) {
        // grow the staging buffers (old ones stay owned by the network)
        const uint32_t cap = std::max(count, 2 * inj_cap);
        if ((rc = alloc(&inj_idx, cap))) return rc;
        if ((rc = alloc(&inj_val, cap))) return rc;
        inj_cap = cap;
    }
    HIP_TRY(hipMemcpy(inj_idx, g.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(inj_val, val, 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    launch_inject(ring, step, N, inj_idx, inj_val, count, dup ? 1 : 0, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::inject_voltage(int32_t pop, const uint32_t* idx, const float* val, uint32_t count,
                             hipStream_t stream) {
    if (pop < 0 || pop >= static_cast<int32_t>(pops.size()))
        return fail(SIMUSE_E_INVALID, "inject_voltage: population %d out of range (have %zu)", pop, pops.size());
    if (count == 0) return SIMUSE_OK;
    const HostPop& P = pops[pop];
    if (!val) return fail(SIMUSE_E_INVALID, "inject_voltage: null values");
    if (!idx && count != P.n)
        return fail(SIMUSE_E_INVALID, "inject_voltage: %u values without indices for a population of %u", count, P.n);
    bool dup = false;
    if (idx) {
        std::vector<uint32_t> sorted(idx, idx + count);
        for (uint32_t i = 0; i < count; ++i)
            if (idx[i] >= P.n)
                return fail(SIMUSE_E_INVALID, "inject_voltage: neuron index %u out of bounds (population has %u)",
                            idx[i], P.n);
        std::sort(sorted.begin(), sorted.end());
        dup = std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end();
    }
    for (uint32_t i = 0; i < count; ++i)
        if (!std::isfinite(val[i])) return fail(SIMUSE_E_INVALID, "inject_voltage: value %u must be finite", i);
    HIP_TRY(hipStreamSynchronize(stream));
    int rc;
    if (count > inj_cap) {
        const uint32_t cap = std::max(count, 2 * inj_cap);
        if ((rc = alloc(&inj_idx, cap))) return rc;
        if ((rc = alloc(&inj_val, cap))) return rc;
        inj_cap = cap;
    }
    if (idx) HIP_TRY(hipMemcpy(inj_idx, idx, 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(inj_val, val, 4 * static_cast<size_t>(count), hipMemcpyHostToDevice));
    // state column 0 of the population is the membrane potential of every model
    launch_vjump(P.st, idx ? inj_idx : nullptr, inj_val, count, dup ? 1 : 0, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::reset(const simuse_pop_desc* pd, int32_t n_pops, hipStream_t stream) {
    uint64_t nt = 0;
    int rc;
    if ((rc = validate_pops(pd, n_pops, &nt))) return rc;
    if (n_pops != static_cast<int32_t>(pops.size()) || nt != N)
        return fail(SIMUSE_E_INVALID, "reset: population layout differs from the network");
    for (int32_t p = 0; p < n_pops; ++p)
        if (pd[p].model != pops[p].model || pd[p].n != pops[p].n || pd[p].stride != pops[p].host_stride)
            return fail(SIMUSE_E_INVALID, "reset: population %d differs from the network", p);
    HIP_TRY(hipStreamSynchronize(stream));
    if ((rc = upload_pops(pd))) return rc;
    if (dtrace) HIP_TRY(hipMemsetAsync(dtrace, 0, sizeof(float) * N, stream));
    if (tm_type) {
        launch_tm_init(tm_type, tm_prof, tm_r, tm_u, tm_last, n_syn + n_src_edges, stream);
        HIP_TRY(hipGetLastError());
    }
    return reset_dynamic(stream);
}

int NetState::counters(uint64_t* spikes, uint64_t* events, hipStream_t stream) {
    std::vector<uint64_t> ts(n_tiles), te(static_cast<size_t>(n_tiles) + n_src_tiles);
    HIP_TRY(hipMemcpyAsync(ts.data(), tile_spikes, 8 * ts.size(), hipMemcpyDeviceToHost, stream));
    if (tile_events) HIP_TRY(hipMemcpyAsync(te.data(), tile_events, 8 * te.size(), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    uint64_t s = 0, e = 0;
    for (uint64_t x : ts) s += x;
    if (tile_events)
        for (uint64_t x : te) e += x;
    if (spikes) *spikes = s;
    if (events) *events = e;
    return SIMUSE_OK;
}

int NetState::record(uint32_t max_steps, hipStream_t stream) {
    int rc;
    HIP_TRY(hipStreamSynchronize(stream));
    if (max_steps > raster_cap) {
        uint64_t* r = nullptr;
        if ((rc = alloc(&r, static_cast<size_t>(max_steps) * n_words))) return rc;
        raster = r;
        raster_cap = max_steps;
    }
    if (max_steps == 0) raster = nullptr;
    if (raster) HIP_TRY(hipMemset(raster, 0, 8 * static_cast<size_t>(raster_cap) * n_words));
    // recording starts at the current device step
    HIP_TRY(hipMemcpy(step + 1, step, 8, hipMemcpyDeviceToDevice));
    raster_start = host_step;
    return SIMUSE_OK;
}

int NetState::download_raster(uint64_t* out, uint64_t n_words_buf, uint32_t* steps, hipStream_t stream) {
    if (!raster) return fail(SIMUSE_E_STATE, "recording is off");
    const uint64_t s = std::min<uint64_t>(host_step - raster_start, raster_cap);
    if (steps) *steps = static_cast<uint32_t>(s);
    if (!out) return SIMUSE_OK;
    if (n_words_buf < s * n_words) return fail(SIMUSE_E_INVALID, "raster buffer too small");
    HIP_TRY(hipMemcpyAsync(out, raster, 8 * s * n_words, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

int NetState::raster_stats(uint32_t first, uint32_t count, float dt_s, float* rates, float* cvs, hipStream_t stream) {
    if (!raster) return fail(SIMUSE_E_STATE, "recording is off");
    if (static_cast<uint64_t>(first) + count > N) return fail(SIMUSE_E_INVALID, "neuron range out of bounds");
    if (count == 0) return SIMUSE_OK;
    const uint32_t steps = static_cast<uint32_t>(std::min<uint64_t>(host_step - raster_start, raster_cap));
    std::vector<uint32_t> wo(count), bo(count);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t gi = first + k;
        const HostPop* P = nullptr;
        for (const auto& q : pops)
            if (gi >= q.off && gi < q.off + q.n) {
                P = &q;
                break;
            }
        const uint32_t li = gi - P->off;
        wo[k] = P->word0 + li / 64;
        bo[k] = li % 64;
    }
    uint32_t *dwo = nullptr, *dbo = nullptr;
    float *dr = nullptr, *dc = nullptr;
    hipError_t e = hipMalloc(&dwo, 4 * static_cast<size_t>(count));
    if (e == hipSuccess) e = hipMalloc(&dbo, 4 * static_cast<size_t>(count));
    if (e == hipSuccess) e = hipMalloc(&dr, 4 * static_cast<size_t>(count));
    if (e == hipSuccess) e = hipMalloc(&dc, 4 * static_cast<size_t>(count));
    if (e == hipSuccess) e = hipMemcpy(dwo, wo.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dbo, bo.data(), 4 * static_cast<size_t>(count), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_raster_stats(raster, n_words, steps, dwo, dbo, count, dt_s, dr, dc, stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && rates) e = hipMemcpy(rates, dr, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost);
    if (e == hipSuccess && cvs) e = hipMemcpy(cvs, dc, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost);
    if (dwo) (void)hipFree(dwo);
    if (dbo) (void)hipFree(dbo);
    if (dr) (void)hipFree(dr);
    if (dc) (void)hipFree(dc);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "raster stats: %s", hipGetErrorString(e));
    }
    return SIMUSE_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: time each scatter schedule on a snapshot of every network of the
// engine, restore the snapshot, keep the faster schedule.
// ---------------------------------------------------------------------------
namespace {
struct Snapshot {
    std::vector<std::pair<void*, void*>> copies;  // (live, saved)
    std::vector<size_t> sizes;
    ~Snapshot() {
        for (auto& c : copies) (void)hipFree(c.second);
    }
    int add(void* live, size_t bytes) {
        if (!live || bytes == 0) return SIMUSE_OK;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_NOMEM, "autotune snapshot: %s", hipGetErrorString(e));
        }
        copies.push_back({live, q});
        sizes.push_back(bytes);
        return SIMUSE_OK;
    }
    int copy(hipStream_t s, bool save) {
        for (size_t k = 0; k < copies.size(); ++k) {
            void* dst = save ? copies[k].second : copies[k].first;
            void* src = save ? copies[k].first : copies[k].second;
            HIP_TRY(hipMemcpyAsync(dst, src, sizes[k], hipMemcpyDeviceToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        return SIMUSE_OK;
    }
};
}  // namespace

int Engine::autotune(uint32_t trial_steps) {
    int rc = prepare();
    if (rc) return rc;
    if (trial_steps == 0) trial_steps = 16;
    // megastep networks: the pending events into the ring, which the snapshot
    // keeps, with the flag that makes a one-kernel step read them there.  A
    // network that could take the one-kernel step but runs the binned form
    // has its next input in the ring already: the flag is set for it too, so
    // that a one-kernel trial (and the form chosen at the end) reads it.
    for (NetState* s : nets) {
        if ((rc = materialize(s))) return rc;
        if (s->bin_mat && s->binned && s->mega_ok && !s->no_mega && !s->stdp && !s->mega)
            HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(s->bin_mat), 1, 1, stream));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    Snapshot snap;
    for (NetState* s : nets) {
        if (s->bin_mat && (rc = snap.add(s->bin_mat, sizeof(uint32_t)))) return rc;
        for (auto& P : s->pops)
            if ((rc = snap.add(P.st, sizeof(float) * P.ns * P.dstride))) return rc;
        const size_t N = s->N;
        if ((rc = snap.add(s->ring, sizeof(float) * kSlots * N))) return rc;
        if ((rc = snap.add(s->step, sizeof(uint64_t) * 2))) return rc;
        if ((rc = snap.add(s->tile_count, sizeof(uint32_t) * s->n_tiles))) return rc;
        if ((rc = snap.add(s->active, sizeof(uint32_t) * kTile * s->n_tiles))) return rc;
        if ((rc = snap.add(s->bits, sizeof(uint64_t) * s->n_words))) return rc;
        if ((rc = snap.add(s->tile_spikes, sizeof(uint64_t) * s->n_tiles))) return rc;
        if ((rc = snap.add(s->tile_events, sizeof(uint64_t) * (s->n_tiles + s->n_src_tiles)))) return rc;
        if (s->raster) {
            if ((rc = snap.add(s->raster, sizeof(uint64_t) * s->raster_cap * s->n_words))) return rc;
        }
        if (s->tm_type) {
            const size_t n_ev = s->n_syn + s->n_src_edges;
            if ((rc = snap.add(s->tm_r, sizeof(float) * n_ev))) return rc;
            if ((rc = snap.add(s->tm_u, sizeof(float) * n_ev))) return rc;
            if ((rc = snap.add(s->tm_last, sizeof(uint32_t) * n_ev))) return rc;
        }
        if (s->g && (rc = snap.add(s->g, sizeof(float) * kReceptors * N))) return rc;
        if (s->n_sig && (rc = snap.add(s->sig_s, sizeof(float) * s->n_sig))) return rc;
        if (s->stdp) {
            if ((rc = snap.add(s->w, sizeof(float) * s->n_syn))) return rc;
            if ((rc = snap.add(s->w64, sizeof(double) * s->n_syn))) return rc;
            if ((rc = snap.add(s->tr_pre, sizeof(float) * N))) return rc;
            if ((rc = snap.add(s->tr_post, sizeof(float) * N))) return rc;
        }
    }
    if ((rc = snap.copy(stream, true))) return rc;
    // Restoring rewinds the step counters; the scatter heads written by the
    // trial steps carry later tags and are rewritten as idle.
    auto restore = [&]() {
        int r = snap.copy(stream, false);
        if (r) return r;
        for (NetState* s : nets) {
            s->clear_heads(stream);
            if (s->bin_hdr)  // the trial steps' pending events
                HIP_TRY(hipMemsetAsync(s->bin_hdr, 0, sizeof(uint2) * s->nbuf() * s->n_bins * s->bin_rows, stream));
            if (s->bin_done) HIP_TRY(hipMemsetAsync(s->bin_done, 0, sizeof(uint32_t), stream));
        }
        HIP_TRY(hipGetLastError());
        return SIMUSE_OK;
    };
    // Candidates: step form (networks that can take the one-kernel step may
    // instead run the binned form: neuron-kernel writers, then the per-bin
    // apply) x scatter schedule x neuron tiles per block.  The schedule and
    // the block size only shape the binned and chained forms, so with every
    // network in the one-kernel step a single trial stands for all six.
    bool can_mega = false, all_mega = true;
    for (NetState* s : nets) {
        const bool m = s->binned && s->mega_ok && !s->no_mega && !s->stdp;
        can_mega = can_mega || m;
        all_mega = all_mega && m;
    }
    const int cand_s[2] = {SIMUSE_SCATTER_TILE, SIMUSE_SCATTER_WAVE};
    const uint32_t cand_g[3] = {4, 2, 1};
    const int saved = sched;
    const uint32_t saved_g = group;
    const bool saved_one = one_kernel;
    float best_ms = 0;
    bool have = false;
    int best = sched;
    uint32_t best_g = group;
    bool best_one = one_kernel;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int c = 0; c < 12 && rc == SIMUSE_OK; ++c) {
        const bool one = c < 6;
        if (!one && !can_mega) continue;         // no network can tell the forms apart
        if (one && all_mega && c > 0) continue;  // the one-kernel step has neither knob
        if (cand_g[(c % 6) / 2] > kMaxGroup) continue;
        one_kernel = can_mega ? one : saved_one;
        mark_dirty();
        sched = cand_s[c % 2];
        if ((rc = set_group(cand_g[(c % 6) / 2]))) break;
        if ((rc = ensure_tables())) break;
        if ((rc = restore())) break;
        if ((rc = enqueue_steps(2))) break;  // warm-up
        if ((rc = restore())) break;
        hipError_t e = hipEventRecord(e0, stream);
        if (e == hipSuccess && (rc = enqueue_steps(trial_steps))) break;
        if (e == hipSuccess) e = hipEventRecord(e1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = fail(SIMUSE_E_HIP, "autotune timing: %s", hipGetErrorString(e));
            break;
        }
        if (!have || ms < best_ms) {
            have = true;
            best_ms = ms;
            best = sched;
            best_g = group;
            best_one = one_kernel;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int rc2 = restore();
    sched = saved;
    if (rc) {
        one_kernel = saved_one;
        mark_dirty();
        (void)set_group(saved_g);
        return rc;
    }
    if (rc2) return rc2;
    // the snapshot holds the pending events in the ring with the flag set,
    // which both forms read as the next step's input
    one_kernel = best_one;
    mark_dirty();
    if ((rc = set_group(best_g))) return rc;
    return set_sched(best);
}

}  // namespace simuse

using namespace simuse;

struct simuse_net {
    Engine eng;
    NetState st;
};

// Host work that builds vectors proportional to the caller's data runs
// through this, so that no C++ exception crosses the C ABI.
template <class F>
static int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(SIMUSE_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(SIMUSE_E_INVALID, "host error: %s", e.what());
    } catch (...) {
        return fail(SIMUSE_E_INVALID, "unexpected host error");
    }
}

static int check_net(simuse_net* net) {
    if (!net) return fail(SIMUSE_E_INVALID, "null network handle");
    return SIMUSE_OK;
}

extern "C" {

const char* simuse_last_error(void) { return g_last_error.c_str(); }

int simuse_device_count(int32_t* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
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
    if ((rc = validate_csr(n_total, row_ptr, col, w, delay, n_syn))) return rc;
    for (int32_t p = 0; p < n_pops; ++p) {
        if (pops[p].n == 0) continue;
        const float pdt = pops[p].params[0];
        if (dt > 0.0f && pdt != dt)
            return fail(SIMUSE_E_INVALID, "population %d: param dt %g differs from network dt %g", p, pdt, dt);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMUSE_E_HIP, "no HIP device available");
    }
    return guarded([&]() {
        std::unique_ptr<simuse_net> net(new (std::nothrow) simuse_net());
        if (!net) return fail(SIMUSE_E_NOMEM, "host allocation failed");
        int r;
        if ((r = net->eng.init())) return r;
        net->st.eng = &net->eng;
        if ((r = net->st.init(pops, n_pops, row_ptr, col, w, delay, n_syn, net->eng.stream))) return r;
        net->eng.nets.push_back(&net->st);
        *out = net.release();
        return SIMUSE_OK;
    });
}

void simuse_net_destroy(simuse_net* net) {
    if (!net) return;
    (void)hipStreamSynchronize(net->eng.stream);
    delete net;
}

int simuse_net_add_poisson_source(simuse_net* net, float rate_hz, float dt_ms, uint32_t seed, const uint32_t* posts,
                                  const float* w, const uint8_t* delay, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.add_source(rate_hz, dt_ms, seed, posts, w, delay, count); });
}

int simuse_net_set_noise(simuse_net* net, int32_t pop, const float* sigma, uint64_t seed) {
    int rc = check_net(net);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->eng.stream));
    rc = net->st.set_noise(pop, sigma, seed);
    net->eng.mark_dirty();
    return rc;
}

int simuse_net_set_poisson_drive(simuse_net* net, int32_t pop, float rate_hz, float dt_ms, float weight,
                                 uint32_t seed0) {
    int rc = check_net(net);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->eng.stream));
    rc = net->st.set_drive(pop, rate_hz, dt_ms, weight, seed0);
    net->eng.mark_dirty();
    return rc;
}

int simuse_net_set_plasticity(simuse_net* net, const simuse_stdp_params* p, const uint8_t* plastic) {
    int rc = check_net(net);
    if (rc) return rc;
    if (p && (rc = net->eng.leave_megastep(&net->st))) return rc;
    rc = net->st.set_plasticity(p, plastic, net->eng.stream);
    net->eng.mark_dirty();
    return rc;
}

int simuse_net_set_scatter_mode(simuse_net* net, int32_t mode) {
    int rc = check_net(net);
    if (rc) return rc;
    if (mode == SIMUSE_SCATTER_AUTO) return net->eng.autotune(16);
    return net->eng.set_sched(mode);
}

int simuse_net_get_scatter_mode(simuse_net* net, int32_t* mode) {
    int rc = check_net(net);
    if (rc) return rc;
    if (mode) *mode = net->eng.sched;
    return SIMUSE_OK;
}

int simuse_net_set_block_tiles(simuse_net* net, uint32_t tiles) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.set_group(tiles);
}

int simuse_net_get_block_tiles(simuse_net* net, uint32_t* tiles) {
    int rc = check_net(net);
    if (rc) return rc;
    if (tiles) *tiles = net->eng.group;
    return SIMUSE_OK;
}

int simuse_net_autotune(simuse_net* net, uint32_t trial_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.autotune(trial_steps);
}

int simuse_net_set_one_kernel(simuse_net* net, int32_t on) {
    int rc = check_net(net);
    if (rc) return rc;
    if (on != 0 && on != 1) return fail(SIMUSE_E_INVALID, "one-kernel switch must be 0 or 1");
    return guarded([&]() { return net->eng.set_one_kernel(on != 0); });
}

int simuse_net_set_graph_steps(simuse_net* net, uint32_t steps) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.set_graph_steps(steps);
}

int simuse_net_step_async(simuse_net* net, uint32_t n_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.step_async(n_steps);
}

int simuse_net_sync(simuse_net* net) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.sync();
}

int simuse_net_step(simuse_net* net) {
    int rc = simuse_net_step_async(net, 1);
    if (rc) return rc;
    return simuse_net_sync(net);
}

int simuse_net_set_short_term_plasticity(simuse_net* net, const simuse_stp_profile* profiles, int32_t n_profiles,
                                         float dt, const uint8_t* syn_types, const uint8_t* src_types,
                                         uint64_t n_src_edges) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.set_stp(profiles, n_profiles, dt, syn_types, src_types, n_src_edges); });
}

int simuse_net_stdp_step(simuse_net* net, float trace_decay, float lr_ltp, float lr_ltd, float w_min,
                         float w_max) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!std::isfinite(trace_decay) || !std::isfinite(lr_ltp) || !std::isfinite(lr_ltd) || !std::isfinite(w_min) ||
        !std::isfinite(w_max) || w_min > w_max)
        return fail(SIMUSE_E_INVALID, "stdp_step: parameters must be finite with w_min <= w_max");
    NetState& s = net->st;
    Engine& e = net->eng;
    if (s.stdp) return fail(SIMUSE_E_STATE, "stdp_step: pair-based plasticity is enabled on this network");
    if ((rc = e.prepare())) return rc;
    if ((rc = e.leave_megastep(&s, true))) return rc;  // this step's update, weights and push are separate kernels
    if ((rc = e.prepare())) return rc;
    const uint32_t n_blk = s.n_tiles + s.n_src_tiles;
    if (!s.dtrace) {
        if ((rc = s.alloc(&s.dtrace, s.N))) return rc;
        HIP_TRY(hipMemsetAsync(s.dtrace, 0, sizeof(float) * s.N, e.stream));
    }
    if (!s.d_all_blocks) {
        std::vector<uint2> bl(n_blk);
        for (uint32_t b = 0; b < n_blk; ++b) bl[b] = make_uint2(0u, b);
        if ((rc = s.alloc(&s.d_all_blocks, n_blk))) return rc;
        if (n_blk) HIP_TRY(hipMemcpy(s.d_all_blocks, bl.data(), sizeof(uint2) * n_blk, hipMemcpyHostToDevice));
    }
    if (s.step_limit != std::numeric_limits<uint64_t>::max() && s.host_step >= s.step_limit) return SIMUSE_OK;
    if (s.since_clear + 1 >= (1ull << 31)) s.clear_heads(e.stream);
    ++s.since_clear;
    // trace from the previous step's spikes; update without the inline
    // scatter; every synapse; then the push with the new weights.
    launch_dense_trace(e.d_nets, s.n_tiles, s.dtrace, trace_decay, e.stream);
    for (int m = 0; m < Engine::kModels; ++m)
        launch_neuron(m, e.sched, e.d_nets, e.d_tiles[m], e.n_tiles_m[m], e.stream, 1u, 0u);
    launch_dense_weights(e.d_nets, s.N, s.dtrace, lr_ltp, lr_ltd, w_min, w_max, e.stream);
    if (s.binned) {
        if (s.dense_dirty) {
            if (s.dense_writers.size() > static_cast<size_t>(s.bin_rows) + kXcdPad)
                return fail(SIMUSE_E_STATE, "dense step: %zu writers exceed the list's room", s.dense_writers.size());
            HIP_TRY(hipMemcpy(s.d_dense_writers, s.dense_writers.data(), sizeof(uint4) * s.dense_writers.size(),
                              hipMemcpyHostToDevice));
            s.dense_dirty = false;
        }
        launch_bin(e.d_nets, s.d_dense_writers, static_cast<uint32_t>(s.dense_writers.size()), s.n_bins, e.stream);
        launch_bin_apply(e.d_nets, e.d_apply, e.n_apply, 0u, e.stream);
        launch_bin_apply_rec(e.d_nets, e.d_apply_rec, e.n_apply_rec, e.stream);
    } else {
        launch_insert(e.sched, e.d_nets, s.d_all_blocks, n_blk, e.stream);
        launch_fixup(e.d_nets, nullptr, s.n_fix, e.stream);
    }
    HIP_TRY(hipGetLastError());
    ++s.host_step;
    return SIMUSE_OK;
}

int simuse_net_download_spikes(simuse_net* net, int32_t* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_spikes(out, len, net->eng.stream);
}

int simuse_net_download_spike_ids(simuse_net* net, uint32_t* ids, uint32_t cap, uint32_t* count) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_spike_ids(ids, cap, count, net->eng.stream);
}

int simuse_net_download_state(simuse_net* net, int32_t pop, float* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_state(pop, out, len, net->eng.stream);
}

int simuse_net_download_ring(simuse_net* net, float* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_ring(out, len, net->eng.stream);
}

int simuse_net_upload_weights(simuse_net* net, const float* w, uint64_t n_syn) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.upload_weights(w, n_syn, net->eng.stream);
}

int simuse_net_download_weights(simuse_net* net, float* w, uint64_t n_syn) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_weights(w, n_syn, net->eng.stream);
}

int simuse_net_inject_current(simuse_net* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.inject(pop, idx, val, count, net->eng.stream);
}

int simuse_net_inject_voltage(simuse_net* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.inject_voltage(pop, idx, val, count, net->eng.stream);
}

int simuse_net_reset(simuse_net* net, const simuse_pop_desc* pops, int32_t n_pops) {
    int rc = check_net(net);
    if (rc) return rc;
    rc = net->st.reset(pops, n_pops, net->eng.stream);
    net->eng.mark_dirty();
    return rc;
}

int simuse_net_set_receptors(simuse_net* net, const uint8_t* receptor_types, float scale, const float* ext) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.set_receptors(receptor_types, scale, ext); });
}

static int pop_index(const NetState& s, int32_t pop, uint32_t idx, const char* what, const char* end, uint32_t* gi) {
    if (pop < 0 || pop >= static_cast<int32_t>(s.pops.size()))
        return fail(SIMUSE_E_INVALID, "%s: %s population %d out of range (have %zu)", what, end, pop, s.pops.size());
    if (idx >= s.pops[pop].n) return fail(SIMUSE_E_INVALID, "%s: %s index %u out of bounds", what, end, idx);
    *gi = s.pops[pop].off + idx;
    return SIMUSE_OK;
}

int simuse_net_add_gap_junctions(simuse_net* net, const uint32_t* a, const uint32_t* b, const float* g,
                                 uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.add_coupling(0, a, b, g, count); });
}

int simuse_net_connect_gap_junction(simuse_net* net, int32_t from, uint32_t from_idx, int32_t to, uint32_t to_idx,
                                    float conductance) {
    int rc = check_net(net);
    if (rc) return rc;
    uint32_t a, b;
    if ((rc = pop_index(net->st, from, from_idx, "connect_gap_junction", "from", &a))) return rc;
    if ((rc = pop_index(net->st, to, to_idx, "connect_gap_junction", "to", &b))) return rc;
    return simuse_net_add_gap_junctions(net, &a, &b, &conductance, 1);
}

int simuse_net_add_graded_synapses(simuse_net* net, const uint32_t* pre, const uint32_t* post, const float* g,
                                   uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.add_coupling(1, pre, post, g, count); });
}

int simuse_net_connect_graded_synapse(simuse_net* net, int32_t from, uint32_t from_idx, int32_t to, uint32_t to_idx,
                                      float conductance) {
    int rc = check_net(net);
    if (rc) return rc;
    uint32_t a, b;
    if ((rc = pop_index(net->st, from, from_idx, "connect_graded_synapse", "from", &a))) return rc;
    if ((rc = pop_index(net->st, to, to_idx, "connect_graded_synapse", "to", &b))) return rc;
    return simuse_net_add_graded_synapses(net, &a, &b, &conductance, 1);
}

int simuse_net_register_sigmoid_graded(simuse_net* net, float vth, float delta, float k, float erev,
                                       float conductance, float dt_ms, int32_t* profile) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.add_sigmoid_profile(vth, delta, k, erev, conductance, dt_ms, profile); });
}

int simuse_net_add_sigmoid_graded(simuse_net* net, int32_t profile, const uint32_t* pre, const uint32_t* post,
                                  const float* weight, uint32_t count) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() { return net->st.add_sigmoid_conns(profile, pre, post, weight, count); });
}

int simuse_net_connect_sigmoid_graded(simuse_net* net, int32_t profile, int32_t from, uint32_t from_idx, int32_t to,
                                      uint32_t to_idx, float weight) {
    int rc = check_net(net);
    if (rc) return rc;
    uint32_t a, b;
    if ((rc = pop_index(net->st, from, from_idx, "connect_sigmoid_graded", "from", &a))) return rc;
    if ((rc = pop_index(net->st, to, to_idx, "connect_sigmoid_graded", "to", &b))) return rc;
    return simuse_net_add_sigmoid_graded(net, profile, &a, &b, &weight, 1);
}

int simuse_net_register_kinetics(simuse_net* net, int32_t kind, const float* params, int32_t n_params, float dt,
                                 int32_t* id) {
    int rc = check_net(net);
    if (rc) return rc;
    (void)kind;
    (void)params;
    (void)n_params;
    (void)dt;
    if (id) *id = 0;
    return fail(SIMUSE_E_UNSUPPORTED,
                "register_kinetics: kinetic synapses (expOne/expTwo/alpha/expThree waveforms with their own delay "
                "rings) are not implemented by this engine");
}

int simuse_net_download_conductances(simuse_net* net, float* out, uint64_t len) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_conductances(out, len, net->eng.stream);
}

int simuse_net_save_state(simuse_net* net, const char* path) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() {
        int r = net->eng.prepare();
        return r ? r : net->st.save_state(path, net->eng.stream);
    });
}

int simuse_net_load_state(simuse_net* net, const char* path) {
    int rc = check_net(net);
    if (rc) return rc;
    return guarded([&]() {
        int r = net->eng.prepare();
        if (r) return r;
        r = net->st.load_state(path, net->eng.stream);
        net->eng.mark_dirty();
        return r;
    });
}

int simuse_net_step_counter(simuse_net* net, uint64_t* step) {
    int rc = check_net(net);
    if (rc) return rc;
    if (step) *step = net->st.host_step;
    return SIMUSE_OK;
}

int simuse_net_n_total(simuse_net* net, uint64_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = net->st.N;
    return SIMUSE_OK;
}

int simuse_net_n_pops(simuse_net* net, int32_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = static_cast<int32_t>(net->st.pops.size());
    return SIMUSE_OK;
}

int simuse_net_pop_n(simuse_net* net, int32_t pop, uint64_t* n) {
    int rc = check_net(net);
    if (rc) return rc;
    if (n) *n = (pop >= 0 && pop < static_cast<int32_t>(net->st.pops.size())) ? net->st.pops[pop].n : 0;
    return SIMUSE_OK;
}

int simuse_net_counters(simuse_net* net, uint64_t* spikes, uint64_t* events) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.counters(spikes, events, net->eng.stream);
}

int simuse_net_record(simuse_net* net, uint32_t max_steps) {
    int rc = check_net(net);
    if (rc) return rc;
    const uint64_t* before = net->st.raster;
    const uint32_t cap = net->st.raster_cap;
    rc = net->st.record(max_steps, net->eng.stream);
    // the kernels see the raster through the network table: rebuild it (and
    // drop the captured graph) only when the buffer changed
    if (net->st.raster != before || net->st.raster_cap != cap) net->eng.mark_dirty();
    return rc;
}

int simuse_net_raster_words(simuse_net* net, uint64_t* words) {
    int rc = check_net(net);
    if (rc) return rc;
    if (words) *words = net->st.n_words;
    return SIMUSE_OK;
}

int simuse_net_download_raster(simuse_net* net, uint64_t* out, uint64_t n_words, uint32_t* steps) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.download_raster(out, n_words, steps, net->eng.stream);
}

int simuse_net_raster_stats(simuse_net* net, uint32_t first, uint32_t count, float dt_s, float* rates, float* cvs) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->st.raster_stats(first, count, dt_s, rates, cvs, net->eng.stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Raster statistics without allocation.  simuse_net_raster_stats stages the
// raster word and bit of each neuron and the two outputs in four buffers it
// allocates and frees per call; this variant keeps them in one file-scope
// device array, scratch = [word_of | bit_of | rates | cvs] of
// SIMUSE_RASTER_STATS_NOALLOC_MAX entries each (1 MiB, every GPU the library
// runs on holds its own copy until the process ends), and derives word and
// bit on the device from the population table instead of uploading them.
// ---------------------------------------------------------------------------
namespace {
constexpr uint32_t kStatsMax = SIMUSE_RASTER_STATS_NOALLOC_MAX;
static __device__ __align__(256) uint32_t scratch[4 * kStatsMax];
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;                       // serialises every use of scratch in the process
uint32_t* g_scratch[kScratchDevices] = {};     // device address of scratch, per device

// Device address of the current device's scratch (the name `scratch` is a
// host-side stand-in, not a device address).  Caller holds g_scratch_mu.
int scratch_of_current_device(uint32_t** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return fail(SIMUSE_E_UNSUPPORTED, "raster_stats_noalloc: device %d beyond the %d this library tracks", dev,
                    kScratchDevices);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(SIMUSE_E_HIP, "raster_stats_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        }
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return SIMUSE_OK;
}

// Raster word and bit of neurons first .. first+count-1: the population
// holding global neuron gi is the last one with off <= gi.
__global__ void raster_index_kernel(const PopDev* pops, int32_t n_pops, uint32_t first, uint32_t count,
                                    uint32_t* word_of, uint32_t* bit_of) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t gi = first + k;
    int32_t p = 0;
    while (p + 1 < n_pops && gi >= pops[p + 1].off) ++p;
    const uint32_t li = gi - pops[p].off;
    word_of[k] = pops[p].word0 + li / 64;
    bit_of[k] = li % 64;
}
}  // namespace

extern "C" {

int simuse_net_raster_stats_noalloc(simuse_net* net, uint32_t first, uint32_t count, float dt_s, float* rates,
                                    float* cvs) {
    int rc = check_net(net);
    if (rc) return rc;
    NetState& s = net->st;
    if (!s.raster) return fail(SIMUSE_E_STATE, "raster_stats_noalloc: recording is off");
    if (static_cast<uint64_t>(first) + count > s.N)
        return fail(SIMUSE_E_INVALID, "raster_stats_noalloc: neurons [%u, %llu) outside the network's %u", first,
                    static_cast<unsigned long long>(first) + count, s.N);
    if (count > kStatsMax)
        return fail(SIMUSE_E_INVALID, "raster_stats_noalloc: %u neurons in one call, at most %u", count, kStatsMax);
    if (count == 0) return SIMUSE_OK;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    uint32_t* base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return rc;
    uint32_t* word_of = base;
    uint32_t* bit_of = base + kStatsMax;
    float* r = reinterpret_cast<float*>(base + 2 * kStatsMax);
    float* c = reinterpret_cast<float*>(base + 3 * kStatsMax);
    const uint32_t steps = static_cast<uint32_t>(std::min<uint64_t>(s.host_step - s.raster_start, s.raster_cap));
    hipStream_t st = net->eng.stream;
    raster_index_kernel<<<(count + 255) / 256, 256, 0, st>>>(s.d_pops, static_cast<int32_t>(s.pops.size()), first,
                                                             count, word_of, bit_of);
    HIP_TRY(hipGetLastError());
    launch_raster_stats(s.raster, s.n_words, steps, word_of, bit_of, count, dt_s, r, c, st);
    HIP_TRY(hipGetLastError());
    if (rates) HIP_TRY(hipMemcpyAsync(rates, r, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost, st));
    if (cvs) HIP_TRY(hipMemcpyAsync(cvs, c, 4 * static_cast<size_t>(count), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SIMUSE_OK;
}

int simuse_net_set_profiling(simuse_net* net, int32_t on) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.set_profiling(on != 0);
}

int simuse_net_kernel_times(simuse_net* net, float* ms, int32_t n) {
    int rc = check_net(net);
    if (rc) return rc;
    return net->eng.kernel_times(ms, n);
}

static int time_steps_impl(simuse_net* net, uint32_t n, float* ms, bool queued) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!ms) return fail(SIMUSE_E_INVALID, "null output");
    return guarded([&]() {
        int r = net->eng.prepare();
        if (r) return r;
        if (net->eng.profiling) return fail(SIMUSE_E_STATE, "time_steps: turn profiling off first");
        // the graph's capture and instantiation are host work: done before the
        // first event, so the region holds only replays (and the rest steps)
        if (n >= net->eng.graph_steps && !net->eng.gexec && (r = net->eng.build_graph())) return r;
        // the pair of events is the engine's, created on the first call
        if ((r = net->eng.timing_events())) return r;
        hipEvent_t* e = net->eng.tev;
        hipError_t he = hipSuccess;
        if (queued) {
            // 1 ms of the constant-rate wall clock ahead of the first event
            int khz = 0, dev = 0;
            he = hipGetDevice(&dev);
            if (he == hipSuccess) he = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
            if (he == hipSuccess) {
                launch_spacer(static_cast<uint64_t>(khz > 0 ? khz : 100000), net->eng.stream);
                he = hipGetLastError();
            }
        }
        if (he == hipSuccess) he = hipEventRecord(e[0], net->eng.stream);
        if (he == hipSuccess) r = net->eng.step_async(n);
        if (he == hipSuccess && r == SIMUSE_OK) he = hipEventRecord(e[1], net->eng.stream);
        if (he == hipSuccess && r == SIMUSE_OK) he = hipEventSynchronize(e[1]);
        if (he == hipSuccess && r == SIMUSE_OK) he = hipEventElapsedTime(ms, e[0], e[1]);
        if (r) return r;
        if (he != hipSuccess) return fail(SIMUSE_E_HIP, "time_steps: %s", hipGetErrorString(he));
        return SIMUSE_OK;
    });
}

int simuse_net_time_steps(simuse_net* net, uint32_t n, float* ms) { return time_steps_impl(net, n, ms, false); }
int simuse_net_time_steps_queued(simuse_net* net, uint32_t n, float* ms) { return time_steps_impl(net, n, ms, true); }

int simuse_net_step_form(simuse_net* net, int32_t* form) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!form) return fail(SIMUSE_E_INVALID, "null output");
    return guarded([&]() {
        int r = net->eng.prepare();
        if (r) return r;
        const NetState& s = net->st;
        *form = s.mega    ? SIMUSE_FORM_ONE_KERNEL
                : s.split ? SIMUSE_FORM_SPLIT
                : s.binned ? SIMUSE_FORM_BINNED
                           : SIMUSE_FORM_CHAINED;
        return SIMUSE_OK;
    });
}

int simuse_net_device_bytes(simuse_net* net, uint64_t* bytes) {
    int rc = check_net(net);
    if (rc) return rc;
    if (bytes) *bytes = net->st.bytes;
    return SIMUSE_OK;
}

}  // extern "C"
