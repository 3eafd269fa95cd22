/*
 * simuse.h — C ABI of the MI355X (gfx950) SiMuse step engine.
 *
 * This is the drop-in boundary for the reference's GPU network API,
 * pillars_cuda::CompiledNetwork (crates/SiMuse/crates/pillars-cuda/src/
 * network.rs:182-882), with the step semantics of the CPU canon
 * pillars::api::Network::step (crates/SiMuse/crates/pillars/src/api/net/
 * network/step.rs:18-189).  Path prefixes in citations:
 *   P/  = crates/SiMuse/crates/pillars/
 *   PC/ = crates/SiMuse/crates/pillars-cuda/
 *
 * Conventions
 *   - Every call returns SIMUSE_OK (0) or a negative SIMUSE_E_* code; the
 *     message of the last failure on the calling thread is returned by
 *     simuse_last_error().  No exception or abort crosses the ABI.
 *   - Host pointers are borrowed for the duration of the call only.  A handle
 *     owns all of its device memory; destroy functions accept NULL.
 *   - Calls on different handles may run on different threads concurrently.
 *     Calls on one handle must be serialised by the caller.
 *   - Population state and parameters use the reference's column-major SoA:
 *     column k of population p lives at states[k*stride + i] with
 *     stride >= n (the reference pads to ceil8(n),
 *     P/src/core/exec/population.rs:106-131).
 */
#ifndef SIMUSE_H
#define SIMUSE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMUSE_OK 0
#define SIMUSE_E_INVALID (-1)     /* bad argument or malformed network */
#define SIMUSE_E_HIP (-2)         /* HIP runtime error */
#define SIMUSE_E_NOMEM (-3)       /* device allocation failed / budget exceeded */
#define SIMUSE_E_STATE (-4)       /* call not allowed in the handle's state */
#define SIMUSE_E_UNSUPPORTED (-5) /* feature outside this build's scope */
#define SIMUSE_E_DEVICE (-6)      /* a kernel raised its error flag */

#define SIMUSE_MAX_DELAY_SLOTS 32 /* P/src/api/net/graph.rs:33 */

/* Neuron models (replaces NetworkPop.source / kernel_name,
 * PC/src/network.rs:142-156, with an enum over hand-written kernels). */
#define SIMUSE_MODEL_LIF 0        /* P/src/zoo/neurons/lif.rs:14-78: 2 states, 6 params */
#define SIMUSE_MODEL_IZHIKEVICH 1 /* P/src/zoo/neurons/izhikevich.rs:30-125: 3 states, 7 params */
#define SIMUSE_MODEL_HH 2         /* PC/kernels/hh.cu:107-216: 6 states, 13 params */

/* Scatter schedules the autotuner chooses between. */
#define SIMUSE_SCATTER_AUTO 0
#define SIMUSE_SCATTER_TILE 1  /* all threads of a block share one tile's edges */
#define SIMUSE_SCATTER_WAVE 2  /* one wavefront per spiking neuron */

const char* simuse_last_error(void);
int simuse_device_count(int32_t* count);

/* Population descriptor — replaces NetworkPop (PC/src/network.rs:142-156). */
typedef struct simuse_pop_desc {
    int32_t model;         /* SIMUSE_MODEL_* */
    uint32_t n;            /* neurons */
    uint32_t ns;           /* state columns (must match the model) */
    uint32_t np;           /* parameter columns (must match the model) */
    uint32_t stride;       /* host column stride, >= n */
    const float* states;   /* [ns*stride] */
    const float* params;   /* [np*stride]; column 0 is dt (neuron 0 broadcast) */
} simuse_pop_desc;

/* Pair-based STDP (rule of P/src/ops/plasticity/stdp.rs:34-75, applied on the
 * CSR event-by-event as in P/src/api/learn/brain.rs:204-286). */
typedef struct simuse_stdp_params {
    float a_plus;    /* LTP amplitude, x pre-trace at a post spike */
    float a_minus;   /* LTD amplitude, x post-trace at a pre spike */
    float tau_plus;  /* ms */
    float tau_minus; /* ms */
    float w_min;
    float w_max;
} simuse_stdp_params;

typedef struct simuse_net simuse_net;

/* CompiledNetwork::new (PC/src/network.rs:221-381): populations concatenated
 * in order, global CSR (row_ptr[n_total+1], col/w/delay[n_syn]), delays in
 * steps 0..31 (0 is delivered next step, P/src/api/net/push.rs:161-165).
 * Validation follows SparseGraph::from_csr (P/src/api/net/graph.rs:217-280):
 * row_ptr starts at 0, is non-decreasing and ends at n_syn; col < n_total;
 * weights finite; delay < 32.  dt > 0 must equal every population's param
 * column 0 (dt <= 0 means "take dt from the params"). */
int simuse_net_create(const simuse_pop_desc* pops, int32_t n_pops,
                      const uint32_t* row_ptr, const uint32_t* col, const float* w,
                      const uint8_t* delay, uint64_t n_syn, float dt, simuse_net** out);
/* Drop (PC/src/network.rs:845-882). */
void simuse_net_destroy(simuse_net* net);

/* Network::add_poisson_source + connect_poisson_source
 * (P/src/api/net/network/core.rs:288-388), post indices global.  Allowed
 * until the first step. */
int simuse_net_add_poisson_source(simuse_net* net, float rate_hz, float dt_ms, uint32_t seed,
                                  const uint32_t* posts, const float* w, const uint8_t* delay,
                                  uint32_t count);
/* Gaussian current noise sigma[i]*N(0,1) added to the current slot of every
 * neuron of `pop` each step (thalamic input).  No reference counterpart: the
 * reference drives this through inject_current from the host every step. */
int simuse_net_set_noise(simuse_net* net, int32_t pop, const float* sigma, uint64_t seed);
/* Enables pair-based STDP; plastic[e] != 0 marks plastic edges (NULL = all).
 * Weights are accumulated in FP64 with an FP32 mirror used by the scatter. */
int simuse_net_set_plasticity(simuse_net* net, const simuse_stdp_params* p, const uint8_t* plastic);
int simuse_net_set_scatter_mode(simuse_net* net, int32_t mode);
int simuse_net_get_scatter_mode(simuse_net* net, int32_t* mode);
/* Times each scatter schedule on a snapshot of the live state for
 * `trial_steps` steps, restores the snapshot, keeps the faster schedule. */
int simuse_net_autotune(simuse_net* net, uint32_t trial_steps);
/* Steps captured per hipGraph replay (1..64). */
int simuse_net_set_graph_steps(simuse_net* net, uint32_t steps);

/* CompiledNetwork::step (PC/src/network.rs:411): one step, synchronous. */
int simuse_net_step(simuse_net* net);
/* step_graph (PC/src/network.rs:431) generalised to n steps: enqueue only.
 * The slot and step counter advance on the device, so back-to-back calls
 * need no host synchronisation. */
int simuse_net_step_async(simuse_net* net, uint32_t n_steps);
int simuse_net_sync(simuse_net* net);  /* PC/src/network.rs:463 */

/* download_spikes (PC/src/network.rs:471): 0/1 flags of the last step, n_total. */
int simuse_net_download_spikes(simuse_net* net, int32_t* out, uint64_t len);
/* Ascending spike indices of the last step (compact_spikes_into,
 * P/src/api/net/compact.rs:30-61).  *count receives the number of spikes;
 * at most `cap` are written. */
int simuse_net_download_spike_ids(simuse_net* net, uint32_t* ids, uint32_t cap, uint32_t* count);
/* download_states (PC/src/network.rs:481): column-major SoA with the
 * descriptor's host stride, len = ns*stride. */
int simuse_net_download_state(simuse_net* net, int32_t pop, float* out, uint64_t len);
/* Ring buffer [32][n_total] (P/src/api/net/delay.rs:30-101). */
int simuse_net_download_ring(simuse_net* net, float* out, uint64_t len);
/* upload_weights / download_weights (PC/src/network.rs:721-748). */
int simuse_net_upload_weights(simuse_net* net, const float* w, uint64_t n_syn);
int simuse_net_download_weights(simuse_net* net, float* w, uint64_t n_syn);
/* inject_current (PC/src/network.rs:755; CPU P/src/api/net/network/build.rs:243-275):
 * adds into the slot the next step reads. */
int simuse_net_inject_current(simuse_net* net, int32_t pop, const uint32_t* idx,
                              const float* val, uint32_t count);
/* reset (PC/src/network.rs:827-842): reload state and params, zero the ring,
 * step counter and traces. */
int simuse_net_reset(simuse_net* net, const simuse_pop_desc* pops, int32_t n_pops);

/* step_counter / n_total / n_pops / pop_n (PC/src/network.rs:389-401,820). */
int simuse_net_step_counter(simuse_net* net, uint64_t* step);
int simuse_net_n_total(simuse_net* net, uint64_t* n);
int simuse_net_n_pops(simuse_net* net, int32_t* n);
int simuse_net_pop_n(simuse_net* net, int32_t pop, uint64_t* n);
/* Device-side totals since creation/reset: spikes and synaptic events
 * (sum over spikes of out-degree, Poisson-source events included). */
int simuse_net_counters(simuse_net* net, uint64_t* spikes, uint64_t* events);

/* Spike raster recording on the device: one bit per neuron per step for at
 * most max_steps steps from now; words per step = simuse_net_raster_words. */
int simuse_net_record(simuse_net* net, uint32_t max_steps);
int simuse_net_raster_words(simuse_net* net, uint64_t* words_per_step);
int simuse_net_download_raster(simuse_net* net, uint64_t* out, uint64_t n_words, uint32_t* steps);
/* Firing statistics of the recorded raster for neurons [first, first+count):
 * per-neuron rate (Hz, P/src/kernels/stats/rates.rs:6-35) and ISI CV
 * (P/src/kernels/stats/isi.rs:30-67), computed on the device. */
int simuse_net_raster_stats(simuse_net* net, uint32_t first, uint32_t count, float dt_s,
                            float* rates, float* cvs);

/* Bytes of device memory the handle holds. */
int simuse_net_device_bytes(simuse_net* net, uint64_t* bytes);
/* Mean duration (ms) of each step kernel over the last simuse_net_step_async
 * window measured with HIP events on the handle's stream when profiling is
 * on: [neuron, insert, fold, stdp]. */
int simuse_net_set_profiling(simuse_net* net, int32_t on);
int simuse_net_kernel_times(simuse_net* net, float* ms, int32_t n);

/* ---------------------------------------------------------------------------
 * Kuramoto oscillators on a CSR graph.
 * Coupling c_i = sum_j W_ij sin(theta_j - theta_i) over row i; theta wraps to
 * [-pi, pi).  SIMUSE_KURA_EULER is the reference's integrator
 * (P/src/kernels/phase/sparse.rs:40-97, kuramoto.rs:42-68; GPU
 * PC/kernels/kuramoto.cu:9-75,163-195); SIMUSE_KURA_RK4 is classic RK4 with
 * the phase accumulated in FP64.
 * ------------------------------------------------------------------------- */
#define SIMUSE_KURA_EULER 0
#define SIMUSE_KURA_RK4 1

typedef struct simuse_kura simuse_kura;

int simuse_kura_create(const uint32_t* row_ptr, const uint32_t* col, const float* w, const float* omega,
                       const float* theta0, uint32_t n, float dt, int32_t method, simuse_kura** out);
void simuse_kura_destroy(simuse_kura* k);
/* lanes per CSR row in the gather (power of two 1..64; default ~ mean degree / 4) */
int simuse_kura_set_group(simuse_kura* k, int32_t lanes);
int simuse_kura_step_async(simuse_kura* k, uint32_t n_steps);
int simuse_kura_sync(simuse_kura* k);
int simuse_kura_download_phases(simuse_kura* k, double* out, uint32_t n);
/* R = |sum_i e^{i theta_i}| / n (phase_coherence with unit amplitudes,
 * P/src/kernels/measure/mod.rs:76-106), deterministic FP64 reduction. */
int simuse_kura_order_parameter(simuse_kura* k, double* r);
int simuse_kura_step_counter(simuse_kura* k, uint64_t* steps);
/* HIP-event timing of whole steps (all stage kernels of a step). */
int simuse_kura_set_profiling(simuse_kura* k, int32_t on);
int simuse_kura_kernel_times(simuse_kura* k, float* ms, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
