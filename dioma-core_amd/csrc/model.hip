// model.hip — one population without synapses, stepped with a device-held
// input current: the C ABI of the reference's single-population GPU path,
// pillars_cuda::CompiledModel (PC/src/engine.rs:98-423; PC/ =
// crates/SiMuse/crates/pillars-cuda/).
//
// CompiledModel keeps states, params and the input current i_in in device
// memory (upload once), advances every neuron per step with dt from the
// call, leaves the 0/1 spike flags of the last step on the device and
// compacts them on request.  Here the update is the network engine's own
// (Neuron<MODEL> in net_kernels.hip), one thread per neuron; the step also
// writes the spike bit words, from which the compaction lists the spiking
// neurons in ascending order (the reference's atomic compaction gives the
// same set in an unspecified order).
//
// Device layout: states [ns][dstride] and params [np][dstride] (SoA,
// dstride = n rounded up to 64; the host passes column stride `stride`),
// i_in [n], spikes [n] int32, bits [n / 64] u64, ids [n] u32, count u32.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "engine.hpp"

namespace simuse {
void launch_pop_step(int model, const PopDev& P, const float* i_in, int32_t* spikes, uint64_t* bits, float dt,
                     hipStream_t s);
void launch_pop_compact(const uint64_t* bits, uint32_t n_words, uint32_t* ids, uint32_t* count, hipStream_t s);
}  // namespace simuse

using namespace simuse;

struct simuse_pop {
    int32_t model = 0;
    uint32_t n = 0, ns = 0, np = 0, hstride = 0, dstride = 0, n_words = 0;
    hipStream_t stream = nullptr;
    float* st = nullptr;
    float* pr = nullptr;
    float* i_in = nullptr;
    int32_t* spikes = nullptr;
    uint64_t* bits = nullptr;
    uint32_t* ids = nullptr;
    uint32_t* count = nullptr;
    uint32_t compacted = 0;  // count of the last compact_spikes
    PopDev P{};
    hipEvent_t tev[2] = {nullptr, nullptr};  // simuse_pop_time_steps' pair

    ~simuse_pop() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : tev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {static_cast<void*>(st), static_cast<void*>(pr), static_cast<void*>(i_in),
                        static_cast<void*>(spikes), static_cast<void*>(bits), static_cast<void*>(ids),
                        static_cast<void*>(count)})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int alloc(void** p, size_t bytes) {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(SIMUSE_E_NOMEM, "population allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        }
        e = hipMemset(*p, 0, bytes);
        if (e != hipSuccess) return fail(SIMUSE_E_HIP, "hipMemset failed: %s", hipGetErrorString(e));
        return SIMUSE_OK;
    }

    int init(int32_t m, uint32_t ns_, uint32_t np_, uint32_t n_, uint32_t stride) {
        model = m;
        n = n_;
        ns = ns_;
        np = np_;
        hstride = stride;
        dstride = (n + 63u) / 64u * 64u;
        n_words = (n + 63u) / 64u;
        HIP_TRY(hipStreamCreate(&stream));
        int rc;
        if ((rc = alloc(reinterpret_cast<void**>(&st), sizeof(float) * ns * dstride))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&pr), sizeof(float) * np * dstride))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&i_in), sizeof(float) * n))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&spikes), sizeof(int32_t) * n))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&bits), sizeof(uint64_t) * n_words))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&ids), sizeof(uint32_t) * n))) return rc;
        if ((rc = alloc(reinterpret_cast<void**>(&count), sizeof(uint32_t)))) return rc;
        P.model = m;
        P.n = n;
        P.dstride = dstride;
        P.uni_mask = 0u;  // every parameter column is read per neuron
        P.flags = 0u;     // refractory timers loaded and stored
        P.st = dev_ptr(st);
        P.pr = dev_ptr(static_cast<const float*>(pr));
        return SIMUSE_OK;
    }

    int upload(const float* states, const float* params, const float* cur) {
        if (states)
            HIP_TRY(hipMemcpy2DAsync(st, sizeof(float) * dstride, states, sizeof(float) * hstride, sizeof(float) * n,
                                     ns, hipMemcpyHostToDevice, stream));
        if (params)
            HIP_TRY(hipMemcpy2DAsync(pr, sizeof(float) * dstride, params, sizeof(float) * hstride, sizeof(float) * n,
                                     np, hipMemcpyHostToDevice, stream));
        if (cur) HIP_TRY(hipMemcpyAsync(i_in, cur, sizeof(float) * n, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));  // the host buffers are borrowed for the call only
        return SIMUSE_OK;
    }

    int steps(float dt, uint32_t k) {
        for (uint32_t i = 0; i < k; ++i) launch_pop_step(model, P, i_in, spikes, bits, dt, stream);
        HIP_TRY(hipGetLastError());
        return SIMUSE_OK;
    }

    int download_states(float* out) {
        HIP_TRY(hipStreamSynchronize(stream));
        std::memset(out, 0, sizeof(float) * ns * hstride);
        HIP_TRY(hipMemcpy2D(out, sizeof(float) * hstride, st, sizeof(float) * dstride, sizeof(float) * n, ns,
                            hipMemcpyDeviceToHost));
        return SIMUSE_OK;
    }

    int download_spikes(int32_t* out) {
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(out, spikes, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        return SIMUSE_OK;
    }
};

#define POP_CHECK(p) \
    if (!(p)) return fail(SIMUSE_E_INVALID, "null population handle")

extern "C" {

int simuse_pop_create(int32_t model, uint32_t ns, uint32_t np, uint32_t n, uint32_t stride, simuse_pop** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "null output handle");
    *out = nullptr;
    if (model < SIMUSE_MODEL_LIF || model > SIMUSE_MODEL_ALPHA_LIF) return fail(SIMUSE_E_INVALID, "unknown model %d", model);
    if (ns != static_cast<uint32_t>(model_states(model)) || np != static_cast<uint32_t>(model_params(model)))
        return fail(SIMUSE_E_INVALID, "model %d has %d states and %d params, got ns=%u np=%u", model,
                    model_states(model), model_params(model), ns, np);
    if (n == 0 || n > kMaxNeurons) return fail(SIMUSE_E_INVALID, "population size %u outside 1..%u", n, kMaxNeurons);
    if (stride < n) return fail(SIMUSE_E_INVALID, "column stride %u below the population size %u", stride, n);
    simuse_pop* p = nullptr;
    try {
        p = new simuse_pop();
    } catch (...) {
        return fail(SIMUSE_E_NOMEM, "host allocation failed");
    }
    const int rc = p->init(model, ns, np, n, stride);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return SIMUSE_OK;
}

void simuse_pop_destroy(simuse_pop* p) { delete p; }

int simuse_pop_upload(simuse_pop* p, const float* states, const float* params, const float* i_in) {
    POP_CHECK(p);
    if (!states || !params || !i_in) return fail(SIMUSE_E_INVALID, "upload needs states, params and i_in");
    return p->upload(states, params, i_in);
}

int simuse_pop_upload_i_in(simuse_pop* p, const float* i_in) {
    POP_CHECK(p);
    if (!i_in) return fail(SIMUSE_E_INVALID, "null i_in");
    return p->upload(nullptr, nullptr, i_in);
}

int simuse_pop_step_async(simuse_pop* p, float dt, uint32_t n_steps) {
    POP_CHECK(p);
    return p->steps(dt, n_steps);
}

int simuse_pop_time_steps(simuse_pop* p, float dt, uint32_t n_steps, float* ms) {
    POP_CHECK(p);
    if (!ms) return fail(SIMUSE_E_INVALID, "null output");
    if (!p->tev[0]) {
        HIP_TRY(hipEventCreate(&p->tev[0]));
        HIP_TRY(hipEventCreate(&p->tev[1]));
    }
    // 1 ms of the constant-rate wall clock ahead of the first event: the host
    // enqueues the steps while it runs, so the events bracket the kernels
    int khz = 0, dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    launch_spacer(static_cast<uint64_t>(khz > 0 ? khz : 100000), p->stream);
    HIP_TRY(hipEventRecord(p->tev[0], p->stream));
    int rc = p->steps(dt, n_steps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->tev[1], p->stream));
    HIP_TRY(hipEventSynchronize(p->tev[1]));
    HIP_TRY(hipEventElapsedTime(ms, p->tev[0], p->tev[1]));
    return SIMUSE_OK;
}

int simuse_pop_sync(simuse_pop* p) {
    POP_CHECK(p);
    HIP_TRY(hipStreamSynchronize(p->stream));
    return SIMUSE_OK;
}

int simuse_pop_step_device(simuse_pop* p, float dt) {
    int rc = simuse_pop_step_async(p, dt, 1);
    if (rc) return rc;
    return simuse_pop_sync(p);
}

int simuse_pop_compact_spikes(simuse_pop* p, uint32_t* count) {
    POP_CHECK(p);
    if (!count) return fail(SIMUSE_E_INVALID, "null count");
    launch_pop_compact(p->bits, p->n_words, p->ids, p->count, p->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&p->compacted, p->count, sizeof(uint32_t), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    *count = p->compacted;
    return SIMUSE_OK;
}

int simuse_pop_download_compact(simuse_pop* p, int32_t* indices, uint32_t count) {
    POP_CHECK(p);
    if (count == 0) return SIMUSE_OK;
    if (!indices) return fail(SIMUSE_E_INVALID, "null indices");
    if (count > p->compacted)
        return fail(SIMUSE_E_INVALID, "%u indices asked, the last compaction found %u", count, p->compacted);
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(indices, p->ids, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));
    return SIMUSE_OK;
}

int simuse_pop_download_spikes(simuse_pop* p, int32_t* spikes, uint64_t len) {
    POP_CHECK(p);
    if (!spikes || len != p->n) return fail(SIMUSE_E_INVALID, "spikes len must be n = %u", p->n);
    return p->download_spikes(spikes);
}

int simuse_pop_download_states(simuse_pop* p, float* states, uint64_t len) {
    POP_CHECK(p);
    if (!states || len != static_cast<uint64_t>(p->ns) * p->hstride)
        return fail(SIMUSE_E_INVALID, "states len must be ns * stride = %u", p->ns * p->hstride);
    return p->download_states(states);
}

int simuse_pop_run(simuse_pop* p, float* states, const float* params, const float* i_in, int32_t* spikes, float dt,
                   uint32_t n_steps) {
    POP_CHECK(p);
    if (!states || !params || !i_in || !spikes) return fail(SIMUSE_E_INVALID, "run needs states, params, i_in, spikes");
    if (n_steps == 0) return SIMUSE_OK;  // the reference's loop runs no step and copies nothing
    int rc;
    if ((rc = p->upload(states, params, i_in))) return rc;
    if ((rc = p->steps(dt, n_steps))) return rc;
    if ((rc = p->download_states(states))) return rc;
    return p->download_spikes(spikes);
}

int simuse_pop_step(simuse_pop* p, float* states, const float* params, const float* i_in, int32_t* spikes, float dt) {
    return simuse_pop_run(p, states, params, i_in, spikes, dt, 1);
}

}  // extern "C"
