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
 *   S/  = crates/SiMuse/
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
#define SIMUSE_MODEL_GPEA 3       /* P/src/zoo/neurons/gpea.rs:266-360: 9 states, 14 params */
#define SIMUSE_MODEL_ALPHA_LIF 4  /* P/src/zoo/neurons/alpha_lif.rs:38-137: 3 states (v, g, refr_cnt),
                                     7 params (dt, time_factor, v_rest, v_thresh, v_reset, refr_steps,
                                     syn_decay); input not clamped */

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
/* Per-neuron Poisson drive: neuron i of `pop` gets its own Poisson source
 * (rate_hz, seed drive_seed0 + i*2654435761 mod 2^32, stream 0) with one edge
 * of weight `weight` and delay <= 1, exactly as add_poisson_source with one
 * post would give, but rolled inside the neuron kernel instead of scattered.
 * (Bit-identical to the explicit sources unless inject_current adds to the
 * same neuron in the same step.) */
int simuse_net_set_poisson_drive(simuse_net* net, int32_t pop, float rate_hz, float dt_ms, float weight,
                                 uint32_t drive_seed0);
/* Enables pair-based STDP; plastic[e] != 0 marks plastic edges (NULL = all).
 * Weights are accumulated in FP64 with an FP32 mirror used by the scatter. */
int simuse_net_set_plasticity(simuse_net* net, const simuse_stdp_params* p, const uint8_t* plastic);
/* Schedule of the chained scatter (networks with a delay > 1 step): one
 * spiking row per wave or a block-flattened row list.  No effect on results. */
int simuse_net_set_scatter_mode(simuse_net* net, int32_t mode);
int simuse_net_get_scatter_mode(simuse_net* net, int32_t* mode);
/* Neuron tiles (256 neurons each) one block of the neuron kernel advances,
 * 1..2 (the one-kernel step of networks with delays <= 1 always takes 8).
 * No effect on results. */
int simuse_net_set_block_tiles(simuse_net* net, uint32_t tiles);
int simuse_net_get_block_tiles(simuse_net* net, uint32_t* tiles);
/* Times each candidate on a snapshot of the live state for `trial_steps`
 * steps, restores the snapshot, keeps the fastest.  Candidates: the scatter
 * schedule x the block size, and, when the network can take the one-kernel
 * step, that step against the binned form (simuse_net_set_one_kernel).  No
 * effect on results. */
int simuse_net_autotune(simuse_net* net, uint32_t trial_steps);
/* on = 1 (the default): a network of one model without pair-based
 * plasticity, receptor synapses or delays above one step takes the
 * one-kernel step (megastep_kernel); on = 0: it takes the binned form
 * (neuron-kernel writers, then the per-bin apply) instead.  No effect on
 * results; the autotuner sets it.  No reference counterpart (the reference
 * has one GPU step form). */
int simuse_net_set_one_kernel(simuse_net* net, int32_t on);
/* Steps captured per hipGraph replay (1..64). */
int simuse_net_set_graph_steps(simuse_net* net, uint32_t steps);

/* Tsodyks-Markram short-term plasticity profile (PlasticityProfile,
 * P/src/api/net/synapse.rs:240-285).  tau_fac <= 0: depression only. */
typedef struct {
    float u0;       /* baseline release probability, (0, 1] */
    float tau_rec;  /* recovery time constant (ms), > 0 */
    float tau_fac;  /* facilitation time constant (ms), <= 0 for none */
} simuse_stp_profile;

/* Network::register_plasticity (P/src/api/net/network/core.rs:597-619) for
 * profiles 1..n_profiles, and per-edge profile ids (0 = static) for the
 * synapses (Edge::with_synapse_plasticity) and, optionally, for the Poisson
 * source edges in add order (n_src_edges must then equal their count).  The
 * push delivers weight*R*U and advances R, U (tm_spike, P/src/api/net/
 * push.rs:60-86); state starts at R = 1, U = U0 (build.rs:70-84).  Call after
 * add_poisson_source and before the first step. */
int simuse_net_set_short_term_plasticity(simuse_net* net, const simuse_stp_profile* profiles, int32_t n_profiles,
                                         float dt, const uint8_t* syn_types, const uint8_t* src_types,
                                         uint64_t n_src_edges);

/* CompiledNetwork::step (PC/src/network.rs:411): one step, synchronous. */
int simuse_net_step(simuse_net* net);
/* CompiledNetwork::stdp_step (PC/src/network.rs:495-718, kernels :53-83):
 * one step with the dense every-synapse rule.  trace = trace*decay + spike of
 * the previous step (one trace per neuron, zero at create/reset), then the
 * population update, then for every synapse
 * w = min(w_max, max(w_min, (w + lr_ltp*trace[pre]) - lr_ltd*trace[post])),
 * then the push with the updated weights.  Asynchronous like step_async.
 * Returns SIMUSE_E_STATE when pair-based plasticity is enabled. */
int simuse_net_stdp_step(simuse_net* net, float trace_decay, float lr_ltp, float lr_ltd, float w_min,
                         float w_max);
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
/* inject_voltage (WgpuNetwork::inject_voltage, PW/src/network.rs:1632-1659;
 * kernel v_jump_kernel, PW/shaders/csr_scatter.wgsl:117-124): adds val[k] to
 * the membrane potential (state column 0) of neuron idx[k] of population pop,
 * a local index; idx NULL means count == the population's size and neuron k
 * gets val[k].  The reference defers the jump to the start of the next step,
 * before that step reads its input; it is applied here at once, between two
 * steps, which is the same state change (two calls before one step both
 * apply here, where the reference keeps only the last).  Synchronous.
 * (PW/ = crates/SiMuse/crates/pillars-wgpu/.) */
int simuse_net_inject_voltage(simuse_net* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count);
/* reset (PC/src/network.rs:827-842): reload state and params, zero the ring,
 * step counter and traces. */
int simuse_net_reset(simuse_net* net, const simuse_pop_desc* pops, int32_t n_pops);

/* Conductance-based receptor synapses (SynapticKinetics,
 * P/src/kernels/synapse/kinetics.rs:67-233, receptor.rs:26-150; the step of
 * the reference's conductance simulations, CN/src/sim/hh.rs:115-166 with
 * CN = crates/SiMuse/crates/pillars-connectome/).  Synapse e gets receptor
 * type receptor_types[e]: 0 AMPA, 1 NMDA, 2 GABA_A, 3 GABA_B, 4 nAChR,
 * 5 GluCl, 6 mAChR (other values AMPA, as ReceptorType::from_u8), or 255 to
 * stay current-based.  Every step, before the update: each neuron's seven
 * conductances decay by fast_exp(-dt/tau_r) (dt of its population), the
 * constant external current ext[i] (NULL: none) is added to the input, then
 * sum_r g_r (E_r - v) with v the membrane potential before the update and
 * the NMDA Mg2+ block 1/(1 + 0.33 fast_exp(clamp(-0.06 v, -10, 10))).  A
 * spike jumps g[type][post] += w*scale in push order, seen from the next
 * step on (propagate.rs:225-290).  Neurons below (n_total & ~7) add the seven
 * currents onto the input one by one, the rest add their sum (the
 * reference's 8-lane / scalar-tail split).  Poisson sources stay
 * current-based.  Call before the first step; needs every delay <= 1
 * (SIMUSE_E_UNSUPPORTED at the first step otherwise). */
int simuse_net_set_receptors(simuse_net* net, const uint8_t* receptor_types, float scale, const float* ext);
/* Conductances [7][n_total] (receptor-major). */
int simuse_net_download_conductances(simuse_net* net, float* out, uint64_t len);

/* Electrical coupling inside the step (P/src/api/net/network/step.rs:36-104):
 * with every membrane potential V taken at the start of the step, after the
 * ring slot is read, each gap junction (a, b, g) adds g (V_b - V_a) to a's
 * input and subtracts it from b's, each graded synapse (pre, post, g) adds
 * g (V_pre - V_post) to post's, then each sigmoid graded synapse (profile
 * order, then connection order) advances its activation s (exponential
 * Euler) and adds weight * conductance * s * (erev - V_post) to post's; all
 * in call order.  Gaussian noise (set_noise) enters the input before them,
 * as the reference's inject_current before the step would.  Call before the
 * first step.  Networks with coupling run the two-kernel step; coupling is
 * refused (SIMUSE_E_UNSUPPORTED at the first step) with receptor synapses or
 * the fused Poisson drive.  The sigmoid's exponentials are this engine's
 * deterministic exp (within 2 ulp of the reference's f32::exp).
 * connect_gap_junction: Network::connect_gap_junction, core.rs:391-426.
 * connect_graded_synapse: Network::connect_graded_synapse, core.rs:438-473.
 * register_sigmoid_graded: Network::register_sigmoid_graded, core.rs:483-500,
 *   with SigmoidGradedSynapse::validate (sources.rs:71-86); returns the
 *   profile id in *profile.
 * connect_sigmoid_graded: Network::connect_sigmoid_graded, core.rs:510-553.
 * The add_* forms take global neuron indices and arrays (bulk wiring). */
int simuse_net_connect_gap_junction(simuse_net* net, int32_t from, uint32_t from_idx, int32_t to, uint32_t to_idx,
                                    float conductance);
int simuse_net_add_gap_junctions(simuse_net* net, const uint32_t* a, const uint32_t* b, const float* g,
                                 uint32_t count);
int simuse_net_connect_graded_synapse(simuse_net* net, int32_t from, uint32_t from_idx, int32_t to, uint32_t to_idx,
                                      float conductance);
int simuse_net_add_graded_synapses(simuse_net* net, const uint32_t* pre, const uint32_t* post, const float* g,
                                   uint32_t count);
int simuse_net_register_sigmoid_graded(simuse_net* net, float vth, float delta, float k, float erev,
                                       float conductance, float dt_ms, int32_t* profile);
int simuse_net_connect_sigmoid_graded(simuse_net* net, int32_t profile, int32_t from, uint32_t from_idx, int32_t to,
                                      uint32_t to_idx, float weight);
int simuse_net_add_sigmoid_graded(simuse_net* net, int32_t profile, const uint32_t* pre, const uint32_t* post,
                                  const float* weight, uint32_t count);
/* Network::register_kinetics (core.rs:564-586: SynapseKinetics waveforms
 * expOne/expTwo/alpha/expThree, P/src/api/net/synapse.rs:46-70, with their
 * own delay rings) is not implemented: always SIMUSE_E_UNSUPPORTED, so that a
 * network that needs it is refused rather than run without it. */
int simuse_net_register_kinetics(simuse_net* net, int32_t kind, const float* params, int32_t n_params, float dt,
                                 int32_t* id);

/* Checkpoint / resume: the reference saves download_states /
 * download_weights (PC/src/network.rs:481,736) as named sections of a
 * `.taas` v2 file (S/crates/taas-format/src/writer.rs:73-99, lib.rs:8-24)
 * and restores them with upload_weights and reset(pops)
 * (PC/src/network.rs:721,827-842).  save_state writes a `.taas` v2 file
 * (magic "TAAS2", u64 LE manifest length, JSON manifest, CRC-32 sections)
 * holding every piece of dynamic state: population states and parameters,
 * the 32-slot ring with the pending events, the step counter, weights,
 * last-step spikes, event counters, STDP traces and FP64 weights, synapse
 * resources.  load_state restores it into a network built the same way
 * (populations, CSR, sources, noise, drive and plasticity set-up; checked
 * against a fingerprint in the file, SIMUSE_E_INVALID naming the first
 * difference); stepping then continues bit for bit as the saved network
 * would have.  A bad file (magic, CRC, kinds, sizes) is rejected with
 * SIMUSE_E_INVALID before the network is changed.  Both finalize the
 * network if it has not stepped yet.  zstd-compressed sections are refused. */
int simuse_net_save_state(simuse_net* net, const char* path);
int simuse_net_load_state(simuse_net* net, const char* path);

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
/* No counterpart in the reference.  simuse_net_raster_stats without any
 * device allocation, for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc): the staging buffers that call allocates
 * and frees live instead in one file-scope device array holding
 * SIMUSE_RASTER_STATS_NOALLOC_MAX neurons (1 MiB; each GPU the library runs
 * on keeps its copy until the process ends).  count above that maximum
 * returns SIMUSE_E_INVALID.  Same results as simuse_net_raster_stats.  Every
 * call on a GPU shares that GPU's array and leaves its contents there; the
 * library serialises these calls within a process (calls that overlapped on
 * one GPU would race), and each call finishes its work before returning.
 * Call with the network's device current. */
#define SIMUSE_RASTER_STATS_NOALLOC_MAX 65536
int simuse_net_raster_stats_noalloc(simuse_net* net, uint32_t first, uint32_t count, float dt_s,
                                    float* rates, float* cvs);

/* Bytes of device memory the handle holds. */
int simuse_net_device_bytes(simuse_net* net, uint64_t* bytes);
/* Which kernel sequence steps the network (finalizes it if it has not stepped
 * yet).  No counterpart in the reference, whose network always takes the
 * same four launches (PC/src/network.rs:895-1035). */
#define SIMUSE_FORM_ONE_KERNEL 0 /* delays <= 1, one model, no pair STDP or receptors: megastep_kernel */
#define SIMUSE_FORM_BINNED 1     /* delays <= 1: neuron kernel / binned writers, then the per-bin apply */
#define SIMUSE_FORM_CHAINED 2    /* some delay > 1: chained scatter and its fix-up */
#define SIMUSE_FORM_SPLIT 3      /* delays <= 1, one model, pair STDP or receptors: front kernel, then writers */
int simuse_net_step_form(simuse_net* net, int32_t* form);
/* Mean duration (ms) of each phase of the step over the last
 * simuse_net_step_async window, measured with HIP events on the handle's
 * stream when profiling is on: [step kernels (the one-kernel step of networks
 * with delays <= 1, or the neuron update with its fused push), stdp, push
 * writers after STDP and of Poisson sources, binned apply / chained
 * fix-up]. */
int simuse_net_set_profiling(simuse_net* net, int32_t on);
int simuse_net_kernel_times(simuse_net* net, float* ms, int32_t n);
/* No counterpart in the reference (a measurement helper): run n steps as
 * simuse_net_step_async does (captured graphs) between two HIP events on the
 * handle's stream, wait for the second, and return the elapsed time in ms.
 * Profiling must be off. */
int simuse_net_time_steps(simuse_net* net, uint32_t n, float* ms);
/* time_steps with the steps queued behind a 1 ms device-side spacer, so
 * that the two events bracket the step kernels only (the host's submission of
 * the graph replays happens while the spacer runs): the steps' device time. */
int simuse_net_time_steps_queued(simuse_net* net, uint32_t n, float* ms);

/* ---------------------------------------------------------------------------
 * Single population without synapses: pillars_cuda::CompiledModel
 * (PC/src/engine.rs:98-423).  States, params and the input current i_in
 * stay in device memory; every step advances each neuron with the same
 * i_in (until the host uploads another) and the call's dt, with the network
 * engine's update for the model (LIF and Izhikevich clamp the input to
 * +-100 as the CPU population does, P/src/core/exec/population.rs:69-88).
 * The model is the enum of simuse_pop_desc instead of an NVRTC source
 * string.  Host arrays use the column-major SoA with column stride `stride`.
 * ------------------------------------------------------------------------- */
typedef struct simuse_pop simuse_pop;
/* CompiledModel::new (engine.rs:98-160): ns and np must be the model's state
 * and parameter counts, stride >= n.  Everything starts at zero. */
int simuse_pop_create(int32_t model, uint32_t ns, uint32_t np, uint32_t n, uint32_t stride, simuse_pop** out);
/* Drop (engine.rs:426-455): frees everything; NULL is accepted. */
void simuse_pop_destroy(simuse_pop* pop);
/* upload (engine.rs:191-220): states [ns*stride], params [np*stride], i_in [n]. */
int simuse_pop_upload(simuse_pop* pop, const float* states, const float* params, const float* i_in);
/* upload_i_in (engine.rs:222-236). */
int simuse_pop_upload_i_in(simuse_pop* pop, const float* i_in);
/* step_graph (engine.rs:253-286): n_steps steps queued, asynchronous. */
int simuse_pop_step_async(simuse_pop* pop, float dt, uint32_t n_steps);
/* sync (engine.rs:288-296). */
int simuse_pop_sync(simuse_pop* pop);
/* n_steps steps between two HIP events, queued behind a 1 ms device-side
 * spacer so that the events bracket the step kernels only; *ms = elapsed.
 * No reference counterpart (measurement helper). */
int simuse_pop_time_steps(simuse_pop* pop, float dt, uint32_t n_steps, float* ms);
/* step_device (engine.rs:238-251): one step, then sync. */
int simuse_pop_step_device(simuse_pop* pop, float dt);
/* compact_spikes (engine.rs:298-352): the last step's spiking neurons listed
 * on the device in ascending order (the reference's atomic list has the same
 * set in an unspecified order); *count receives how many. */
int simuse_pop_compact_spikes(simuse_pop* pop, uint32_t* count);
/* download_compact (engine.rs:354-373): the first `count` listed indices
 * (count at most that of the last compaction). */
int simuse_pop_download_compact(simuse_pop* pop, int32_t* indices, uint32_t count);
/* download_spikes (engine.rs:375-390): the last step's 0/1 flags, len = n. */
int simuse_pop_download_spikes(simuse_pop* pop, int32_t* spikes, uint64_t len);
/* download_states (engine.rs:392-407): len = ns*stride (columns padded with 0). */
int simuse_pop_download_states(simuse_pop* pop, float* states, uint64_t len);
/* step (engine.rs:163-189): upload, one step, states and spikes back. */
int simuse_pop_step(simuse_pop* pop, float* states, const float* params, const float* i_in, int32_t* spikes,
                    float dt);
/* run (engine.rs:409-423): n_steps steps with the same params and input;
 * states and the last step's spikes back (n_steps = 0 touches nothing). */
int simuse_pop_run(simuse_pop* pop, float* states, const float* params, const float* i_in, int32_t* spikes,
                   float dt, uint32_t n_steps);

/* ---------------------------------------------------------------------------
 * Batches: many independent networks stepped by one launch sequence.
 * No reference counterpart (the reference steps one CompiledNetwork per call);
 * each network keeps the simuse_net step semantics, its own ring, step
 * counter and RNG keys, so its results are identical alone or in a batch.
 * ------------------------------------------------------------------------- */
typedef struct simuse_batch simuse_batch;

/* Seeded random network built on the device.  Neurons [0, round(n*exc_frac))
 * are excitatory.  Each neuron has k outgoing synapses to uniform random
 * other neurons (multi-edges possible), weights U(0,w_exc) (exc rows) or
 * -U(0,w_inh) (inh rows), delays uniform in [d_min, d_max].
 * LIF: one population, tau_m 20 ms, v_rest = v_reset = -65, v_th -50,
 *      refractory 2 ms, v0 ~ U(-65,-50).
 * Izhikevich: RS population (exc) and FS population (inh) with the 2003
 *      randomisation, v_th 30, no refractory period, v0 -65, u0 = b v0.
 * HH: one population, standard Na/K/leak (CN/src/sim/hh.rs:26-38), v_th 0,
 *      refractory 2 ms, v0 ~ U(-70,-60), resting gates.
 * Optional per-neuron Poisson drive (simuse_net_set_poisson_drive semantics)
 * and Gaussian current noise. */
typedef struct simuse_gen_desc {
    int32_t model; /* SIMUSE_MODEL_LIF, SIMUSE_MODEL_IZHIKEVICH or SIMUSE_MODEL_HH */
    uint32_t n;
    uint32_t k;
    uint64_t seed;
    float dt;
    float exc_frac;
    float w_exc;
    float w_inh;
    uint8_t d_min;
    uint8_t d_max;
    float drive_rate_hz;
    float drive_weight;
    float noise_sigma;
} simuse_gen_desc;

int simuse_batch_create(simuse_batch** out);
void simuse_batch_destroy(simuse_batch* b);
int simuse_batch_add_network(simuse_batch* b, const simuse_pop_desc* pops, int32_t n_pops, const uint32_t* row_ptr,
                             const uint32_t* col, const float* w, const uint8_t* delay, uint64_t n_syn,
                             int32_t* net_id);
int simuse_batch_add_generated(simuse_batch* b, const simuse_gen_desc* g, int32_t* net_id);
/* Frees the network's device memory; its id may be reused. */
int simuse_batch_remove_network(simuse_batch* b, int32_t net_id);
int simuse_batch_n_networks(simuse_batch* b, int32_t* n);
int simuse_batch_add_poisson_source(simuse_batch* b, int32_t net_id, float rate_hz, float dt_ms, uint32_t seed,
                                    const uint32_t* posts, const float* w, const uint8_t* delay, uint32_t count);
int simuse_batch_set_noise(simuse_batch* b, int32_t net_id, int32_t pop, const float* sigma, uint64_t seed);
/* simuse_net_set_plasticity / simuse_net_set_short_term_plasticity for one
 * network of the batch (before its first step). */
int simuse_batch_set_plasticity(simuse_batch* b, int32_t net_id, const simuse_stdp_params* p, const uint8_t* plastic);
int simuse_batch_set_short_term_plasticity(simuse_batch* b, int32_t net_id, const simuse_stp_profile* profiles,
                                           int32_t n_profiles, float dt, const uint8_t* syn_types,
                                           const uint8_t* src_types, uint64_t n_src_edges);
int simuse_batch_set_poisson_drive(simuse_batch* b, int32_t net_id, int32_t pop, float rate_hz, float dt_ms,
                                   float weight, uint32_t drive_seed0);
/* The network stops advancing once its step counter reaches `steps`. */
int simuse_batch_set_horizon(simuse_batch* b, int32_t net_id, uint64_t steps);
int simuse_batch_set_graph_steps(simuse_batch* b, uint32_t steps);
/* simuse_net_set_one_kernel for every network of the batch, and the current
 * setting (the autotuner's choice after simuse_batch_autotune). */
int simuse_batch_set_one_kernel(simuse_batch* b, int32_t on);
int simuse_batch_get_one_kernel(simuse_batch* b, int32_t* on);
int simuse_batch_set_scatter_mode(simuse_batch* b, int32_t mode);
int simuse_batch_set_block_tiles(simuse_batch* b, uint32_t tiles);
int simuse_batch_autotune(simuse_batch* b, uint32_t trial_steps, int32_t* mode);
int simuse_batch_step_async(simuse_batch* b, uint32_t n_steps);
int simuse_batch_sync(simuse_batch* b);
int simuse_batch_step_counter(simuse_batch* b, int32_t net_id, uint64_t* step);
int simuse_batch_net_info(simuse_batch* b, int32_t net_id, uint64_t* n_total, uint64_t* n_syn, int32_t* n_pops);
int simuse_batch_pop_info(simuse_batch* b, int32_t net_id, int32_t pop, int32_t* model, uint64_t* n);
int simuse_batch_download_state(simuse_batch* b, int32_t net_id, int32_t pop, float* out, uint64_t len);
int simuse_batch_download_params(simuse_batch* b, int32_t net_id, int32_t pop, float* out, uint64_t len);
int simuse_batch_download_csr(simuse_batch* b, int32_t net_id, uint32_t* row_ptr, uint32_t* col, float* w,
                              uint8_t* delay);
int simuse_batch_download_spike_ids(simuse_batch* b, int32_t net_id, uint32_t* ids, uint32_t cap, uint32_t* count);
int simuse_batch_counters(simuse_batch* b, int32_t net_id, uint64_t* spikes, uint64_t* events);
int simuse_batch_record(simuse_batch* b, int32_t net_id, uint32_t max_steps);
int simuse_batch_download_raster(simuse_batch* b, int32_t net_id, uint64_t* out, uint64_t n_words, uint32_t* steps,
                                 uint64_t* words_per_step);
int simuse_batch_device_bytes(simuse_batch* b, uint64_t* bytes);
int simuse_batch_set_profiling(simuse_batch* b, int32_t on);
int simuse_batch_kernel_times(simuse_batch* b, float* ms, int32_t n);
/* Device bytes a generated network will take (admission control). */
uint64_t simuse_gen_bytes(const simuse_gen_desc* g);

/* ---------------------------------------------------------------------------
 * Serving: continuous batching of simulation requests.  No reference
 * counterpart.  A server thread admits queued requests in arrival order while
 * the HBM budget allows (estimated bytes of the request), builds them on the
 * device, steps every live network together (one simuse_batch) and every live
 * Kuramoto graph on its own stream, and retires each request at its horizon.
 * ------------------------------------------------------------------------- */
/* Directed modular small-world graph built on the device: `modules` rings of
 * n/modules nodes; node i takes k in-edges from its lattice neighbours
 * i±1..±k/2, each redrawn with probability `inter` from another module, else
 * with probability `beta` from its own module; W_ij = coupling/k;
 * omega ~ N(0,1), theta0 ~ U(-pi, pi). */
typedef struct simuse_kura_gen_desc {
    uint32_t n;
    uint32_t modules;
    uint32_t k;
    float beta;
    float inter;
    float coupling;
    float dt;
    uint64_t seed;
} simuse_kura_gen_desc;

#define SIMUSE_REQ_NETWORK 0
#define SIMUSE_REQ_KURAMOTO 1

#define SIMUSE_REQ_QUEUED 0
#define SIMUSE_REQ_RUNNING 1
#define SIMUSE_REQ_DONE 2
#define SIMUSE_REQ_REJECTED (-1) /* larger than the whole budget */
#define SIMUSE_REQ_FAILED (-2)

typedef struct simuse_server simuse_server;

typedef struct simuse_request {
    int32_t kind;                /* SIMUSE_REQ_NETWORK or SIMUSE_REQ_KURAMOTO */
    simuse_gen_desc net;         /* network requests */
    simuse_kura_gen_desc kura;   /* Kuramoto requests (RK4) */
    uint64_t horizon;            /* steps */
} simuse_request;

typedef struct simuse_result {
    int32_t status;              /* SIMUSE_REQ_* */
    uint64_t steps;
    uint64_t spikes;             /* network: spikes over the horizon */
    uint64_t events;             /* network: synaptic events over the horizon */
    double order_parameter;      /* Kuramoto: R at the horizon */
    double t_submit, t_admit, t_done; /* seconds since the server started */
    uint64_t bytes;              /* admission estimate */
} simuse_result;

typedef struct simuse_server_stats_t {
    uint64_t submitted, completed, live, queued;
    uint64_t live_bytes, peak_bytes;
    uint32_t peak_live;
    double now;
} simuse_server_stats_t;

int simuse_server_create(uint64_t hbm_budget_bytes, uint32_t chunk_steps, simuse_server** out);
int simuse_server_submit(simuse_server* s, const simuse_request* r, uint64_t* id);
int simuse_server_result(simuse_server* s, uint64_t id, simuse_result* out);
/* Waits until every submitted request has completed (or timeout_s). */
int simuse_server_drain(simuse_server* s, double timeout_s);
int simuse_server_stats(simuse_server* s, simuse_server_stats_t* out);
void simuse_server_destroy(simuse_server* s);

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

int simuse_kura_create_generated(const simuse_kura_gen_desc* g, int32_t method, simuse_kura** out);
int simuse_kura_download_graph(simuse_kura* k, uint32_t* row_ptr, uint32_t* col, float* w, float* omega);
int simuse_kura_info(simuse_kura* k, uint32_t* n, uint64_t* n_edges);
/* lanes per CSR row in the gather (power of two 1..64; default ~ mean degree / 4) */
int simuse_kura_set_group(simuse_kura* k, int32_t lanes);
int simuse_kura_step_async(simuse_kura* k, uint32_t n_steps);
int simuse_kura_sync(simuse_kura* k);
int simuse_kura_download_phases(simuse_kura* k, double* out, uint32_t n);
/* R = |sum_i e^{i theta_i}| / n (phase_coherence with unit amplitudes,
 * P/src/kernels/measure/mod.rs:76-106), deterministic FP64 reduction. */
int simuse_kura_order_parameter(simuse_kura* k, double* r);
/* phase_coherence (P/src/kernels/measure/mod.rs:76-106): |sum a_i e^{i theta_i}| /
 * max(sum a_i, 1e-12) with host amplitudes a (n values), summed in FP64. */
int simuse_kura_phase_coherence(simuse_kura* k, const float* amplitudes, double* r);
int simuse_kura_step_counter(simuse_kura* k, uint64_t* steps);
/* HIP-event timing of whole steps (all stage kernels of a step). */
int simuse_kura_set_profiling(simuse_kura* k, int32_t on);
int simuse_kura_kernel_times(simuse_kura* k, float* ms, int32_t n);

/* ---------------------------------------------------------------------------
 * Creature pool: many small alpha-LIF brains stepped together as one
 * block-diagonal network, with the reference's deterministic per-post pull
 * scatter and delay buckets.  Drop-in for the game engine's BrainPool
 * (DE/src/brain/dynamics/gpu_pool.rs:1-260) over WgpuNetwork::step_many
 * (PW/src/network.rs:802-885) and PW/shaders/csr_scatter.wgsl:26-59,117-124.
 * Per substep: v += ext at the first substep of every submit of up to
 * SIMUSE_POOL_MAX_STEPS_PER_SUBMIT substeps, input = ring slot, alpha-LIF
 * update (SIMUSE_MODEL_ALPHA_LIF), then the previous substep's spikes pulled
 * into ring slot (cur + max(d, 1)) % 32 per post, fma(w, sign(pre), acc) in
 * ascending source order.  Weights are read at the substep that pulls them,
 * so a weight sync between two steps applies to the spikes of the last
 * substep before it, as in the reference.
 * (DE/ = crates/dioma-engine/, PW/ = crates/SiMuse/crates/pillars-wgpu/.)
 * ------------------------------------------------------------------------- */
#define SIMUSE_POOL_MAX_STEPS_PER_SUBMIT 64 /* GPU_MAX_STEPS_PER_SUBMIT, gpu_pool.rs:26 */

typedef struct simuse_pool simuse_pool;

/* NeuronConfig::LifAlpha (crates/SiMuse/crates/pillars-connectome/src/runner.rs:114) */
typedef struct {
    float dt_ms, tau_ref, tau_m, v_rest, v_thresh, v_reset, tau_syn, t_delay;
} simuse_pool_config;

/* PoolSlot, gpu_pool.rs:5-11: a creature's first neuron and synapse */
typedef struct {
    uint64_t neuron_off, n, syn_off, n_syn;
} simuse_pool_slot;

/* BrainPool::new (gpu_pool.rs:57-90).  cfg NULL = pool_config()
 * (gpu_pool.rs:13-24: dt 0.5, tau_ref 2.2, tau_m 20, v_rest -52, v_thresh -45,
 * v_reset -52, tau_syn 3, t_delay 1.8).  delay_steps < 0: every synapse
 * delayed round(t_delay / dt) + 1 steps; >= 0: that value (the DIOMA_GPU_DELAY
 * override); capped at 31 either way. */
int simuse_pool_create(const simuse_pool_config* cfg, int32_t delay_steps, simuse_pool** out);
void simuse_pool_destroy(simuse_pool* p);
/* BrainPool::register (gpu_pool.rs:101-128): a creature's ConnectomeData CSR
 * with local indices (row_ptr: n + 1 entries, col_indices and values:
 * row_ptr[n]) and nt_types (n values; < 0 marks an inhibitory presynaptic
 * neuron, sign -1, else +1).  As in the reference, the pool is rebuilt at the
 * next step or weight sync, and every creature restarts from the initial
 * state (v = v_rest, g = 0, refr_cnt = refr_steps).  At most 2^27 neurons and
 * 2^32 synapses in all (SIMUSE_E_UNSUPPORTED beyond). */
int simuse_pool_register(simuse_pool* p, uint32_t n, const int32_t* row_ptr, const int32_t* col_indices,
                         const float* values, const float* nt_types, simuse_pool_slot* slot);
/* Per-synapse delays of a creature (n_syn values; the pool's delay until
 * set), as WgpuNetwork::new takes them (PW/src/network.rs:122-131).  BrainPool
 * itself always uses the one delay.  Rebuilds at the next step like register. */
int simuse_pool_set_delays(simuse_pool* p, uint32_t creature, const uint8_t* d, uint64_t n_syn);
int simuse_pool_n_creatures(simuse_pool* p, uint32_t* n);
int simuse_pool_slots(simuse_pool* p, simuse_pool_slot* out, uint32_t cap);
/* BrainPool::sync_weights (gpu_pool.rs:130-148): mirrors[k] (slot k's n_syn
 * weights) is taken for every creature with dirty[k] != 0. */
int simuse_pool_sync_weights(simuse_pool* p, const float* const* mirrors, const uint8_t* dirty, uint32_t n_creatures);
/* BrainPool::step (gpu_pool.rs:150-203): exts[k] is NULL or slot k's n voltage
 * inputs; spt substeps in submits of up to 64; flat receives the spikes,
 * spt rows of n_neurons u32 (1 spiked, 0 not), or is NULL to skip the
 * readback.  Synchronous.  A pool with no creature returns at once. */
int simuse_pool_step(simuse_pool* p, const float* const* exts, uint32_t n_creatures, uint32_t spt, uint32_t* flat,
                     uint64_t flat_len);
/* State column of every neuron: 0 v, 1 g, 2 refr_cnt
 * (WgpuNetwork::download_states, PW/src/network.rs:1140). */
int simuse_pool_download_state(simuse_pool* p, int32_t col, float* out, uint64_t n);
int simuse_pool_info(simuse_pool* p, uint64_t* n_neurons, uint64_t* n_syn, uint64_t* steps, int32_t* delay);
/* No counterpart in the reference (a measurement helper): n more substeps,
 * in submits of up to 64 with the last tick's voltage input applied at each
 * submit's first substep, between two HIP events on the pool's stream; no
 * readback.  The elapsed time in ms. */
int simuse_pool_time_substeps(simuse_pool* p, uint32_t n, float* ms);

#ifdef __cplusplus
}
#endif
#endif
