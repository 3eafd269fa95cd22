// snapshot.hip — checkpoint and resume of one network's device state as a
// `.taas` v2 file (crates/SiMuse/crates/taas-format/src/lib.rs:8-24).
//
// The reference resumes a network from explicit downloads and uploads:
// download_states / download_weights / upload_weights / reset(pops)
// (crates/SiMuse/crates/pillars-cuda/src/network.rs:481,721-748,827-842) and
// keeps them as named `.taas` sections (taas-format/src/lib.rs:238-263).
// Exact resume also needs what those calls leave out: the 32-slot input ring
// (events in flight for up to 31 steps), the step counter (the Poisson and
// noise streams are keyed by it), the STDP traces and FP64 weights, the
// Tsodyks-Markram state and the last step's spikes.  save_state writes all
// of it; load_state puts it into a network built the same way (same
// populations, CSR, sources, noise, drive, plasticity and short-term
// plasticity set-up), after which stepping continues bit for bit as the
// saved network would have.
//
// Sections (little-endian; a u32 array is stored as kind "i32"):
//   structure       u64  fingerprint of everything load requires to match
//   step            u64  [1] the device step counter
//   pop{p}.states   f32  [ns][n] state columns of population p
//   pop{p}.params   f32  [np][n] parameter columns of population p
//   ring            f32  [32][N] input ring, pending events applied
//   weights         f32  [n_syn]
//   spikes.count / spikes.active / spikes.bits / counters.spikes /
//   counters.events      the last step's spikes and the event counters
//   tm.r tm.u f32, tm.last i32        Tsodyks-Markram state (when enabled)
//   stdp.w64 u64 (FP64 bit patterns), stdp.pre stdp.post f32  (when enabled)
//   stdp.dense_trace f32              trace of the dense stdp_step rule
//   receptors.g f32   [7][N]          receptor conductances (when enabled)
//   coupling.s f32                    sigmoid graded activations (when any)
//
// The raster recorder is not part of the state: a network that records
// restarts its raster at the loaded step.
#include "engine.hpp"
#include "taas.hpp"

namespace simuse {
namespace {

constexpr uint64_t kStateFormat = 1;
constexpr size_t kFpHead = 19;  // fixed fingerprint entries before the per-population ones
constexpr size_t kFpPop = 8;

const char* fp_name(size_t i) {
    static const char* head[kFpHead] = {"format",           "n_total",          "n_synapses",
                                        "n_sources",        "n_source_edges",   "n_pops",
                                        "plasticity",       "short-term plasticity",
                                        "CSR row_ptr",      "CSR col",          "CSR delays",
                                        "Poisson sources",  "STDP parameters",  "plastic mask",
                                        "synapse types",    "synapse profiles", "receptor types",
                                        "receptor scale and external current", "electrical coupling"};
    static const char* pop[kFpPop] = {"model",       "size",        "state columns", "parameter columns",
                                      "host stride", "noise/drive flags", "noise/drive settings", "noise sigma"};
    return i < kFpHead ? head[i] : pop[(i - kFpHead) % kFpPop];
}

uint32_t f_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// CRC-32 of a device buffer, read back in pieces.
int dev_crc(const void* d, size_t bytes, uint32_t* out) {
    uint32_t c = 0;
    if (d && bytes) {
        const size_t chunk = size_t(64) << 20;
        std::vector<uint8_t> h(std::min(bytes, chunk));
        for (size_t o = 0; o < bytes; o += chunk) {
            const size_t n = std::min(chunk, bytes - o);
            HIP_TRY(hipMemcpy(h.data(), static_cast<const uint8_t*>(d) + o, n, hipMemcpyDeviceToHost));
            c = taas::crc32(h.data(), n, c);
        }
    }
    *out = c;
    return SIMUSE_OK;
}

template <class T>
int download(std::vector<T>& h, const void* d, size_t count) {
    h.resize(count);
    if (count) HIP_TRY(hipMemcpy(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost));
    return SIMUSE_OK;
}

template <class T>
void push(taas::Writer& w, const std::string& name, const char* kind, const std::vector<T>& v) {
    w.push(name, kind, v.size(), v.data(), sizeof(T) * v.size());
}

}  // namespace

int NetState::fingerprint(std::vector<uint64_t>& fp) const {
    fp.assign(kFpHead, 0);
    fp[0] = kStateFormat;
    fp[1] = N;
    fp[2] = n_syn;
    fp[3] = srcs.size();
    uint64_t n_src_e = 0;
    for (const SourceHost& s : srcs) n_src_e += s.post.size();
    fp[4] = n_src_e;
    fp[5] = pops.size();
    fp[6] = stdp ? 1 : 0;
    fp[7] = tm_type ? 1 : 0;
    fp[8] = taas::crc32(h_row_ptr.data(), 4 * h_row_ptr.size());
    uint32_t c;
    int rc;
    if ((rc = dev_crc(col, 4 * n_syn, &c))) return rc;
    fp[9] = c;
    if ((rc = dev_crc(delay, n_syn, &c))) return rc;
    fp[10] = c;
    c = 0;
    for (const SourceHost& s : srcs) {
        const uint32_t hdr[3] = {f_bits(s.p), s.seed, static_cast<uint32_t>(s.post.size())};
        c = taas::crc32(hdr, sizeof hdr, c);
        c = taas::crc32(s.post.data(), 4 * s.post.size(), c);
        c = taas::crc32(s.w.data(), 4 * s.w.size(), c);
        c = taas::crc32(s.d.data(), s.d.size(), c);
    }
    fp[11] = c;
    if (stdp) {
        const float sp_f[6] = {sp.a_plus, sp.a_minus, sp.tau_plus, sp.tau_minus, sp.w_min, sp.w_max};
        fp[12] = taas::crc32(sp_f, sizeof sp_f);
        if ((rc = dev_crc(plastic, plastic ? n_syn : 0, &c))) return rc;
        fp[13] = c;
    }
    if (tm_type) {
        fp[14] = taas::crc32(h_tm_type.data(), h_tm_type.size());
        fp[15] = taas::crc32(h_tm_prof.data(), sizeof(float4) * h_tm_prof.size());
    }
    if (g) {
        fp[16] = (uint64_t(1) << 32) | taas::crc32(h_rec_code.data(), h_rec_code.size());
        const uint32_t sc = f_bits(rec_scale);
        fp[17] = taas::crc32(h_g_ext.data(), 4 * h_g_ext.size(), taas::crc32(&sc, 4));
    }
    if (coupled()) {
        c = taas::crc32(h_gap.data(), sizeof(CplHost) * h_gap.size());
        c = taas::crc32(h_graded.data(), sizeof(CplHost) * h_graded.size(), c);
        c = taas::crc32(h_sig_prof.data(), sizeof(float) * kSigProf * h_sig_prof.size(), c);
        c = taas::crc32(h_sig_conn.data(), sizeof(SigConnHost) * h_sig_conn.size(), c);
        fp[18] = (uint64_t(1) << 32) | c;
    }
    for (const HostPop& P : pops) {
        fp.push_back(static_cast<uint64_t>(P.model));
        fp.push_back(P.n);
        fp.push_back(P.ns);
        fp.push_back(P.np);
        fp.push_back(P.host_stride);
        fp.push_back(P.flags & ~static_cast<uint32_t>(kPopRefrElided));  // the elision follows the loaded state
        const uint32_t cfg[6] = {f_bits(P.sigma_uni), f_bits(P.drive_p), f_bits(P.drive_w), P.drive_seed0,
                                 static_cast<uint32_t>(P.seed), static_cast<uint32_t>(P.seed >> 32)};
        fp.push_back(taas::crc32(cfg, sizeof cfg));
        const bool per_neuron = (P.flags & kPopNoise) && !(P.flags & kPopNoiseScalar) && P.sigma;
        if ((rc = dev_crc(P.sigma, per_neuron ? 4 * static_cast<size_t>(P.n) : 0, &c))) return rc;
        fp.push_back(c);
    }
    return SIMUSE_OK;
}

int NetState::save_state(const char* path, hipStream_t stream) {
    if (!path || !*path) return fail(SIMUSE_E_INVALID, "save_state: empty path");
    if (!finalized) return fail(SIMUSE_E_STATE, "save_state: the network is not finalized");
    int rc;
    if (eng && (rc = eng->materialize(this))) return rc;  // pending events into the ring
    HIP_TRY(hipStreamSynchronize(stream));
    std::vector<uint64_t> fp;
    if ((rc = fingerprint(fp))) return rc;

    taas::Writer out;
    out.game = "simuse";
    out.metadata["content"] = "network-state";
    out.metadata["n_total"] = std::to_string(N);
    out.metadata["state_format"] = std::to_string(kStateFormat);
    push(out, "structure", "u64", fp);
    std::vector<uint64_t> st;
    if ((rc = download(st, step, 1))) return rc;
    out.metadata["step"] = std::to_string(st[0]);
    push(out, "step", "u64", st);

    std::vector<float> f;
    std::vector<uint32_t> u;
    std::vector<uint64_t> q;
    for (size_t p = 0; p < pops.size(); ++p) {
        const HostPop& P = pops[p];
        const std::string base = "pop" + std::to_string(p);
        f.assign(static_cast<size_t>(P.ns) * P.n, 0.0f);
        if (P.n)
            HIP_TRY(hipMemcpy2D(f.data(), 4 * static_cast<size_t>(P.n), P.st, 4 * static_cast<size_t>(P.dstride),
                                4 * static_cast<size_t>(P.n), P.ns, hipMemcpyDeviceToHost));
        push(out, base + ".states", "f32", f);
        f.assign(static_cast<size_t>(P.np) * P.n, 0.0f);
        if (P.n)
            HIP_TRY(hipMemcpy2D(f.data(), 4 * static_cast<size_t>(P.n), P.pr, 4 * static_cast<size_t>(P.dstride),
                                4 * static_cast<size_t>(P.n), P.np, hipMemcpyDeviceToHost));
        push(out, base + ".params", "f32", f);
    }
    if ((rc = download(f, ring, static_cast<size_t>(kSlots) * N))) return rc;
    push(out, "ring", "f32", f);
    if ((rc = download(f, w, n_syn))) return rc;
    push(out, "weights", "f32", f);
    if ((rc = download(u, tile_count, n_tiles))) return rc;
    push(out, "spikes.count", "i32", u);
    if ((rc = download(u, active, static_cast<size_t>(kTile) * n_tiles))) return rc;
    push(out, "spikes.active", "i32", u);
    if ((rc = download(q, bits, n_words))) return rc;
    push(out, "spikes.bits", "u64", q);
    if ((rc = download(q, tile_spikes, n_tiles))) return rc;
    push(out, "counters.spikes", "u64", q);
    q.assign(static_cast<size_t>(n_tiles) + n_src_tiles, 0);
    if (tile_events && (rc = download(q, tile_events, q.size()))) return rc;
    push(out, "counters.events", "u64", q);
    if (tm_type) {
        const size_t n_ev = n_syn + n_src_edges;
        if ((rc = download(f, tm_r, n_ev))) return rc;
        push(out, "tm.r", "f32", f);
        if ((rc = download(f, tm_u, n_ev))) return rc;
        push(out, "tm.u", "f32", f);
        if ((rc = download(u, tm_last, n_ev))) return rc;
        push(out, "tm.last", "i32", u);
    }
    if (stdp) {
        if ((rc = download(q, w64, n_syn))) return rc;  // doubles as their bit patterns
        push(out, "stdp.w64", "u64", q);
        if ((rc = download(f, tr_pre, N))) return rc;
        push(out, "stdp.pre", "f32", f);
        if ((rc = download(f, tr_post, N))) return rc;
        push(out, "stdp.post", "f32", f);
    }
    if (dtrace) {
        if ((rc = download(f, dtrace, N))) return rc;
        push(out, "stdp.dense_trace", "f32", f);
    }
    if (g) {
        if ((rc = download(f, g, static_cast<size_t>(kReceptors) * N))) return rc;
        push(out, "receptors.g", "f32", f);
    }
    if (n_sig) {
        if ((rc = download(f, sig_s, n_sig))) return rc;
        push(out, "coupling.s", "f32", f);
    }
    if (out.to_file(path)) return fail(SIMUSE_E_INVALID, "save_state: cannot write %s", path);
    return SIMUSE_OK;
}

int NetState::load_state(const char* path, hipStream_t stream) {
    if (!path || !*path) return fail(SIMUSE_E_INVALID, "load_state: empty path");
    if (!finalized) return fail(SIMUSE_E_STATE, "load_state: the network is not finalized");
    taas::Reader r;
    if (!r.open(path)) return fail(SIMUSE_E_INVALID, "load_state %s: %s", path, r.error.c_str());
    auto it = r.metadata.find("content");
    if (it == r.metadata.end() || it->second != "network-state")
        return fail(SIMUSE_E_INVALID, "load_state %s: not a network state snapshot", path);

    // Everything is read and checked before the network is touched, so a
    // bad file leaves it as it was.
    int rc;
    std::vector<uint64_t> fp;
    if ((rc = fingerprint(fp))) return rc;
    const taas::Section* fs = r.find("structure");
    if (!fs || fs->kind != "u64") return fail(SIMUSE_E_INVALID, "load_state %s: missing structure section", path);
    const uint8_t* b = r.get("structure", "u64", fs->count);
    if (!b) return fail(SIMUSE_E_INVALID, "load_state %s: %s", path, r.error.c_str());
    std::vector<uint64_t> saved(fs->count);
    std::memcpy(saved.data(), b, 8 * saved.size());
    for (size_t i = 0; i < std::min(saved.size(), fp.size()); ++i)
        if (saved[i] != fp[i]) {
            if (i >= kFpHead)
                return fail(SIMUSE_E_INVALID, "load_state: population %zu differs from the snapshot (%s)",
                            (i - kFpHead) / kFpPop, fp_name(i));
            return fail(SIMUSE_E_INVALID, "load_state: the network differs from the snapshot (%s)", fp_name(i));
        }
    if (saved.size() != fp.size())
        return fail(SIMUSE_E_INVALID, "load_state: the network differs from the snapshot (structure length)");

    auto sec = [&](const std::string& name, const char* kind, uint64_t count) -> const uint8_t* {
        return r.get(name, kind, count);
    };
#define SNAP_GET(var, name, kind, count)                                                            \
    const uint8_t* var = sec(name, kind, count);                                                    \
    if (!var) return fail(SIMUSE_E_INVALID, "load_state %s: %s", path, r.error.c_str());
    SNAP_GET(s_step, "step", "u64", 1)
    std::vector<std::vector<float>> hs(pops.size()), hp(pops.size());
    std::vector<simuse_pop_desc> pd(pops.size(), simuse_pop_desc{});
    for (size_t p = 0; p < pops.size(); ++p) {
        const HostPop& P = pops[p];
        const std::string base = "pop" + std::to_string(p);
        const size_t n = P.n, stride = P.host_stride;
        SNAP_GET(s_st, base + ".states", "f32", static_cast<uint64_t>(P.ns) * n)
        SNAP_GET(s_pr, base + ".params", "f32", static_cast<uint64_t>(P.np) * n)
        hs[p].assign(static_cast<size_t>(P.ns) * stride, 0.0f);
        hp[p].assign(static_cast<size_t>(P.np) * stride, 0.0f);
        for (uint32_t k = 0; k < P.ns; ++k) std::memcpy(&hs[p][k * stride], s_st + 4 * k * n, 4 * n);
        for (uint32_t k = 0; k < P.np; ++k) std::memcpy(&hp[p][k * stride], s_pr + 4 * k * n, 4 * n);
        pd[p].model = P.model;
        pd[p].n = P.n;
        pd[p].ns = P.ns;
        pd[p].np = P.np;
        pd[p].stride = P.host_stride;
        pd[p].states = hs[p].data();
        pd[p].params = hp[p].data();
    }
    uint64_t nt = 0;
    if ((rc = validate_pops(pd.data(), static_cast<int32_t>(pd.size()), &nt))) return rc;
    SNAP_GET(s_ring, "ring", "f32", static_cast<uint64_t>(kSlots) * N)
    SNAP_GET(s_w, "weights", "f32", n_syn)
    for (uint64_t e = 0; e < n_syn; ++e) {
        float x;
        std::memcpy(&x, s_w + 4 * e, 4);
        if (!std::isfinite(x))
            return fail(SIMUSE_E_INVALID, "load_state: weight %llu is not finite", static_cast<unsigned long long>(e));
    }
    SNAP_GET(s_cnt, "spikes.count", "i32", n_tiles)
    SNAP_GET(s_act, "spikes.active", "i32", static_cast<uint64_t>(kTile) * n_tiles)
    SNAP_GET(s_bits, "spikes.bits", "u64", n_words)
    SNAP_GET(s_ts, "counters.spikes", "u64", n_tiles)
    SNAP_GET(s_te, "counters.events", "u64", static_cast<uint64_t>(n_tiles) + n_src_tiles)
    for (uint32_t t = 0; t < n_tiles; ++t) {
        uint32_t c;
        std::memcpy(&c, s_cnt + 4 * t, 4);
        if (c > kTile) return fail(SIMUSE_E_INVALID, "load_state: spike count of tile %u out of range", t);
    }
    const size_t n_ev = n_syn + n_src_edges;
    const uint8_t *s_tr = nullptr, *s_tu = nullptr, *s_tl = nullptr;
    if (tm_type) {
        SNAP_GET(a, "tm.r", "f32", n_ev)
        SNAP_GET(c, "tm.u", "f32", n_ev)
        SNAP_GET(d, "tm.last", "i32", n_ev)
        s_tr = a;
        s_tu = c;
        s_tl = d;
    }
    const uint8_t *s_w64 = nullptr, *s_pre = nullptr, *s_post = nullptr, *s_dt = nullptr;
    if (stdp) {
        SNAP_GET(a, "stdp.w64", "u64", n_syn)
        SNAP_GET(c, "stdp.pre", "f32", N)
        SNAP_GET(d, "stdp.post", "f32", N)
        s_w64 = a;
        s_pre = c;
        s_post = d;
    }
    if (r.find("stdp.dense_trace")) {
        SNAP_GET(a, "stdp.dense_trace", "f32", N)
        s_dt = a;
    }
    const uint8_t* s_g = nullptr;
    if (g) {
        SNAP_GET(a, "receptors.g", "f32", static_cast<uint64_t>(kReceptors) * N)
        s_g = a;
    }
    const uint8_t* s_sig = nullptr;
    if (n_sig) {
        SNAP_GET(a, "coupling.s", "f32", n_sig)
        s_sig = a;
    }
#undef SNAP_GET

    // Upload.  The work of earlier steps is finished first.
    HIP_TRY(hipStreamSynchronize(stream));
    if ((rc = upload_pops(pd.data()))) return rc;
    uint64_t st[2];
    std::memcpy(&st[0], s_step, 8);
    st[1] = st[0];
    HIP_TRY(hipMemcpy(step, st, sizeof st, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ring, s_ring, 4 * static_cast<size_t>(kSlots) * N, hipMemcpyHostToDevice));
    if (n_syn) HIP_TRY(hipMemcpy(w, s_w, 4 * n_syn, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tile_count, s_cnt, 4 * static_cast<size_t>(n_tiles), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(active, s_act, 4 * static_cast<size_t>(kTile) * n_tiles, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bits, s_bits, 8 * static_cast<size_t>(n_words), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tile_spikes, s_ts, 8 * static_cast<size_t>(n_tiles), hipMemcpyHostToDevice));
    if (tile_events)
        HIP_TRY(hipMemcpy(tile_events, s_te, 8 * (static_cast<size_t>(n_tiles) + n_src_tiles), hipMemcpyHostToDevice));
    if (tm_type && n_ev) {
        HIP_TRY(hipMemcpy(tm_r, s_tr, 4 * n_ev, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tm_u, s_tu, 4 * n_ev, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tm_last, s_tl, 4 * n_ev, hipMemcpyHostToDevice));
    }
    if (stdp) {
        if (n_syn) HIP_TRY(hipMemcpy(w64, s_w64, 8 * n_syn, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tr_pre, s_pre, 4 * static_cast<size_t>(N), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(tr_post, s_post, 4 * static_cast<size_t>(N), hipMemcpyHostToDevice));
    }
    if (s_g) {
        HIP_TRY(hipMemcpy(g, s_g, 4 * static_cast<size_t>(kReceptors) * N, hipMemcpyHostToDevice));
        // a receptor type without synapses keeps whatever conductance the file holds
        for (int r = 0; r < kReceptors; ++r) {
            const uint8_t* col = s_g + 4 * static_cast<size_t>(r) * N;
            for (uint32_t i = 0; i < N && !((rec_live >> r) & 1u); ++i) {
                uint32_t b;
                std::memcpy(&b, col + 4 * static_cast<size_t>(i), 4);
                if (b) rec_live |= 1u << r;  // the caller rebuilds the kernels' descriptors
            }
        }
    }
    if (s_sig) HIP_TRY(hipMemcpy(sig_s, s_sig, 4 * static_cast<size_t>(n_sig), hipMemcpyHostToDevice));
    if (s_dt) {
        if (!dtrace && (rc = alloc(&dtrace, N))) return rc;
        HIP_TRY(hipMemcpy(dtrace, s_dt, 4 * static_cast<size_t>(N), hipMemcpyHostToDevice));
    } else if (dtrace) {
        HIP_TRY(hipMemset(dtrace, 0, 4 * static_cast<size_t>(N)));
    }
    // Scatter bookkeeping as after a reset: idle heads tagged against the
    // loaded step, no pending binned records, the ring holds every event.
    clear_heads(stream);
    if (tile_coll) HIP_TRY(hipMemsetAsync(tile_coll, 0, sizeof(uint32_t) * (n_tiles + n_src_tiles), stream));
    if (bin_hdr)
        HIP_TRY(hipMemsetAsync(bin_hdr, 0, sizeof(uint2) * nbuf() * n_bins * bin_rows, stream));
    if (bin_done) HIP_TRY(hipMemsetAsync(bin_done, 0, sizeof(uint32_t), stream));
    if (bin_mat) HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(bin_mat), 1, 1, stream));
    host_step = st[0];
    if (raster) {
        HIP_TRY(hipMemsetAsync(raster, 0, 8 * static_cast<size_t>(raster_cap) * n_words, stream));
        raster_start = host_step;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return SIMUSE_OK;
}

}  // namespace simuse
