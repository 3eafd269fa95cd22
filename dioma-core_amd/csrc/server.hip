// server.hip — continuous-batching simulation server.
//
// Requests (a generated spiking network or a Kuramoto graph, each with a
// horizon in steps) are queued by simuse_server_submit from any thread.  A
// server thread loops:
//   1. admit queued requests in arrival order while the HBM budget allows,
//      building each on the device (no host-side graph construction);
//   2. step every live spiking network together through one simuse_batch
//      (one launch per kernel for all of them) and every live Kuramoto graph
//      on its own stream, for a chunk of steps;
//   3. retire requests whose horizon is reached, recording their results and
//      timestamps, and free their memory.
// Networks keep their own step counters and RNG keys, so a request's result is
// the same whether it ran alone or mixed with others.  No reference
// counterpart: the reference has no serving layer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/simuse.h"

namespace simuse {
int fail(int code, const char* fmt, ...);
}
using simuse::fail;

namespace {

uint64_t kura_bytes(const simuse_kura_gen_desc& g) {
    const uint64_t n = g.n, ne = n * g.k;
    return ne * 8 + n * (4 + 4 + 4 + 8 + 16) + (1u << 20);
}

struct Req {
    simuse_request r{};
    simuse_result res{};
    int32_t net = -1;
    simuse_kura* kura = nullptr;
    uint64_t bytes = 0;
    uint64_t kura_steps = 0;
};

}  // namespace

struct simuse_server {
    uint64_t budget = 0;
    uint32_t chunk = 16;
    simuse_batch* batch = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;  // owned; pointers stay valid
    std::vector<Req*> live;                                    // touched by the server thread only
    uint64_t next_id = 1;
    uint64_t live_bytes = 0;
    uint64_t done_count = 0;
    uint64_t live_count = 0;
    uint64_t peak_bytes = 0;
    uint32_t peak_live = 0;
    bool stop = false;
    bool kill = false;
    int error = SIMUSE_OK;
    std::thread th;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // Build one admitted request on the device.  Returns its status.
    int build(Req& q) {
        if (q.r.kind == SIMUSE_REQ_NETWORK) {
            int32_t id = -1;
            int rc = simuse_batch_add_generated(batch, &q.r.net, &id);
            if (rc) return rc;
            q.net = id;
            return simuse_batch_set_horizon(batch, id, q.r.horizon);
        }
        return simuse_kura_create_generated(&q.r.kura, SIMUSE_KURA_RK4, &q.kura);
    }

    void loop() {
        std::vector<Req*> admit;
        for (;;) {
            admit.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                if (kill) break;
                if (live.empty() && pending.empty()) {
                    if (stop) break;
                    cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
                while (!pending.empty()) {
                    Req* q = pending.front();
                    if (q->bytes > budget) {  // can never fit
                        q->res.status = SIMUSE_REQ_REJECTED;
                        q->res.t_done = now();
                        pending.pop_front();
                        ++done_count;
                        cv.notify_all();
                        continue;
                    }
                    if (live_bytes + q->bytes > budget) break;  // wait for memory to free up
                    live_bytes += q->bytes;
                    peak_bytes = std::max(peak_bytes, live_bytes);
                    admit.push_back(q);
                    pending.pop_front();
                }
            }
            for (Req* q : admit) {
                const int rc = build(*q);
                std::lock_guard<std::mutex> lk(mu);
                if (rc) {
                    q->res.status = SIMUSE_REQ_FAILED;
                    q->res.t_done = now();
                    live_bytes -= q->bytes;
                    ++done_count;
                    cv.notify_all();
                } else {
                    q->res.status = SIMUSE_REQ_RUNNING;
                    q->res.t_admit = now();
                    live.push_back(q);
                    ++live_count;
                }
            }
            if (live.empty()) continue;
            peak_live = std::max<uint32_t>(peak_live, static_cast<uint32_t>(live.size()));
            // step: the chunk is bounded by the nearest horizon so requests
            // retire promptly
            uint64_t nearest = ~0ull;
            for (Req* q : live) {
                uint64_t done_steps = 0;
                if (q->r.kind == SIMUSE_REQ_NETWORK) simuse_batch_step_counter(batch, q->net, &done_steps);
                else done_steps = q->kura_steps;
                nearest = std::min(nearest, q->r.horizon - std::min(q->r.horizon, done_steps));
            }
            const uint32_t n = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(chunk, nearest)));
            bool any_net = false;
            int rc = SIMUSE_OK;
            for (Req* q : live) {
                if (q->r.kind == SIMUSE_REQ_NETWORK) {
                    any_net = true;
                } else {
                    const uint32_t m = static_cast<uint32_t>(std::min<uint64_t>(n, q->r.horizon - q->kura_steps));
                    if (m && !rc) rc = simuse_kura_step_async(q->kura, m);
                    q->kura_steps += m;
                }
            }
            if (any_net && !rc) rc = simuse_batch_step_async(batch, n);
            if (any_net && !rc) rc = simuse_batch_sync(batch);
            for (Req* q : live)
                if (q->r.kind == SIMUSE_REQ_KURAMOTO && !rc) rc = simuse_kura_sync(q->kura);
            if (rc) {
                std::lock_guard<std::mutex> lk(mu);
                error = rc;
                stop = true;
                for (Req* q : live) {
                    q->res.status = SIMUSE_REQ_FAILED;
                    ++done_count;
                }
                for (Req* q : pending) {
                    q->res.status = SIMUSE_REQ_FAILED;
                    ++done_count;
                }
                live.clear();
                live_count = 0;
                pending.clear();
                cv.notify_all();
                break;
            }
            // retire
            std::vector<Req*> still;
            for (Req* q : live) {
                bool fin = false;
                uint64_t steps = 0, spikes = 0, events = 0;
                double order = 0.0;
                if (q->r.kind == SIMUSE_REQ_NETWORK) {
                    simuse_batch_step_counter(batch, q->net, &steps);
                    if (steps >= q->r.horizon) {
                        simuse_batch_counters(batch, q->net, &spikes, &events);
                        simuse_batch_remove_network(batch, q->net);
                        fin = true;
                    }
                } else if (q->kura_steps >= q->r.horizon) {
                    simuse_kura_order_parameter(q->kura, &order);
                    steps = q->kura_steps;
                    simuse_kura_destroy(q->kura);
                    q->kura = nullptr;
                    fin = true;
                }
                if (fin) {
                    std::lock_guard<std::mutex> lk(mu);
                    q->res.steps = steps;
                    q->res.spikes = spikes;
                    q->res.events = events;
                    q->res.order_parameter = order;
                    q->res.status = SIMUSE_REQ_DONE;
                    q->res.t_done = now();
                    live_bytes -= q->bytes;
                    ++done_count;
                    --live_count;
                    cv.notify_all();
                } else {
                    still.push_back(q);
                }
            }
            live.swap(still);
        }
    }
};

extern "C" {

int simuse_server_create(uint64_t hbm_budget_bytes, uint32_t chunk_steps, simuse_server** out) {
    if (!out) return fail(SIMUSE_E_INVALID, "null output handle");
    *out = nullptr;
    if (hbm_budget_bytes == 0) return fail(SIMUSE_E_INVALID, "server: HBM budget must be > 0");
    std::unique_ptr<simuse_server> s(new (std::nothrow) simuse_server());
    if (!s) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    s->budget = hbm_budget_bytes;
    s->chunk = chunk_steps ? chunk_steps : 16;
    int rc = simuse_batch_create(&s->batch);
    if (rc) return rc;
    simuse_batch_set_graph_steps(s->batch, std::min<uint32_t>(s->chunk, 64));
    simuse_server* raw = s.get();
    s->th = std::thread([raw] { raw->loop(); });
    *out = s.release();
    return SIMUSE_OK;
}

int simuse_server_submit(simuse_server* s, const simuse_request* r, uint64_t* id) {
    if (!s || !r) return fail(SIMUSE_E_INVALID, "null argument");
    Req q;
    q.r = *r;
    if (r->kind == SIMUSE_REQ_NETWORK) {
        if (r->net.model != SIMUSE_MODEL_LIF && r->net.model != SIMUSE_MODEL_IZHIKEVICH &&
            r->net.model != SIMUSE_MODEL_HH)
            return fail(SIMUSE_E_UNSUPPORTED, "server: network requests are LIF, Izhikevich or HH");
        q.bytes = simuse_gen_bytes(&r->net);
    } else if (r->kind == SIMUSE_REQ_KURAMOTO) {
        q.bytes = kura_bytes(r->kura);
    } else {
        return fail(SIMUSE_E_INVALID, "server: unknown request kind %d", r->kind);
    }
    if (r->horizon == 0) return fail(SIMUSE_E_INVALID, "server: horizon must be > 0");
    std::unique_ptr<Req> owned(new (std::nothrow) Req(q));
    if (!owned) return fail(SIMUSE_E_NOMEM, "host allocation failed");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(SIMUSE_E_STATE, "server is stopping (error %d)", s->error);
    const uint64_t rid = s->next_id++;
    owned->res.status = SIMUSE_REQ_QUEUED;
    owned->res.t_submit = s->now();
    owned->res.bytes = owned->bytes;
    Req* raw = owned.get();
    s->reqs.emplace(rid, std::move(owned));
    s->pending.push_back(raw);
    if (id) *id = rid;
    s->cv.notify_all();
    return SIMUSE_OK;
}

int simuse_server_result(simuse_server* s, uint64_t id, simuse_result* out) {
    if (!s || !out) return fail(SIMUSE_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(SIMUSE_E_INVALID, "unknown request %llu", static_cast<unsigned long long>(id));
    *out = it->second->res;
    return SIMUSE_OK;
}

int simuse_server_drain(simuse_server* s, double timeout_s) {
    if (!s) return fail(SIMUSE_E_INVALID, "null server handle");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (s->done_count < s->reqs.size()) {
        if (s->cv.wait_until(lk, deadline) == std::cv_status::timeout) break;
    }
    if (s->error) return fail(s->error, "server stopped on a device error");
    if (s->done_count < s->reqs.size()) return fail(SIMUSE_E_STATE, "drain timed out");
    return SIMUSE_OK;
}

int simuse_server_stats(simuse_server* s, simuse_server_stats_t* out) {
    if (!s || !out) return fail(SIMUSE_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->reqs.size();
    out->completed = s->done_count;
    out->live = s->live_count;
    out->queued = s->pending.size();
    out->live_bytes = s->live_bytes;
    out->peak_bytes = s->peak_bytes;
    out->peak_live = s->peak_live;
    out->now = s->now();
    return SIMUSE_OK;
}

void simuse_server_destroy(simuse_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->kill = true;
        s->pending.clear();
        s->cv.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    for (auto& kv : s->reqs)
        if (kv.second->kura) simuse_kura_destroy(kv.second->kura);
    simuse_batch_destroy(s->batch);
    delete s;
}

}  // extern "C"
