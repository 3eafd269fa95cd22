/*
 * simuse_oracle.h — C ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's CPU canon for the per-timestep spiking-network step
 * (pillars::api::Network::step, crates/SiMuse/crates/pillars/src/api/net/
 * network/step.rs:18-189).  It exists to check the HIP product path and to
 * time a CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Two numeric twins share one template:
 *   kind 32 — FP32 "reference-faithful": IEEE float, no FMA contraction
 *             (build with -ffp-contract=off), the reference's association
 *             order, ±100 input clamp, dt broadcast from param column 0.
 *   kind 64 — FP64 "truth": the same equations evaluated in double, libm
 *             exp/expm1, used to measure FP32 error.
 */
#ifndef SIMUSE_ORACLE_H
#define SIMUSE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model ids (same numbering as the product's simuse.h). */
#define ORC_MODEL_LIF 0
#define ORC_MODEL_IZHIKEVICH 1
#define ORC_MODEL_HH 2
#define ORC_MODEL_GPEA 3
#define ORC_MODEL_ALPHA_LIF 4

/* One population: column-major SoA with host stride ceil8(n)
 * (P/src/core/exec/population.rs:106-131). */
typedef struct {
    int32_t model;
    uint32_t n;
    uint32_t ns;
    uint32_t np;
    uint32_t stride;
    const float* states;
    const float* params;
} orc_pop_desc;

typedef struct {
    float a_plus;     /* LTP amplitude (per post spike, x pre-trace) */
    float a_minus;    /* LTD amplitude (per pre spike, x post-trace) */
    float tau_plus;   /* pre-trace time constant (ms) */
    float tau_minus;  /* post-trace time constant (ms) */
    float w_min;
    float w_max;
} orc_stdp_params;

/* Tsodyks-Markram short-term plasticity profile (PlasticityProfile,
 * P/src/api/net/synapse.rs:240-285); tau_fac <= 0 means depression only. */
typedef struct {
    float u0;
    float tau_rec;
    float tau_fac;
} orc_stp_profile;

void* orc_net_create(int kind, const orc_pop_desc* pops, int32_t n_pops,
                     const uint32_t* row_ptr, const uint32_t* col,
                     const float* w, const uint8_t* delay, uint64_t n_syn);
void orc_net_destroy(void* net);
int orc_net_add_poisson_source(void* net, float rate_hz, float dt_ms, uint32_t seed,
                               const uint32_t* posts, const float* w,
                               const uint8_t* delay, uint32_t count);
int orc_net_set_noise(void* net, int32_t pop, const float* sigma, uint64_t seed);
int orc_net_set_plasticity(void* net, const orc_stdp_params* p, const uint8_t* plastic);
/* Per-edge TM profile ids (0 = static) for the synapses and, in add order,
 * for the Poisson-source edges (src_types may be NULL); profiles are 1-based. */
int orc_net_set_stp(void* net, const orc_stp_profile* profiles, int32_t n_profiles, float dt,
                    const uint8_t* syn_types, const uint8_t* src_types, uint64_t n_src_edges);
/* Conductance-based receptor synapses (P/src/kernels/synapse/kinetics.rs,
 * receptor.rs; step order of CN/src/sim/hh.rs:115-166). */
int orc_net_set_receptors(void* net, const uint8_t* types, float scale, const float* ext);
void orc_net_conductances(void* net, double* out); /* [7][n_total] */
/* Overwrite the dynamic state (step counter, population states [ns][n] per
 * population, ring [32][N], weights, FP64 master weights, STDP traces,
 * conductances); NULL leaves a part unchanged. */
int orc_net_set_state(void* net, uint64_t step, const float* const* pop_states, const float* ring,
                      const float* w, const double* w64, const float* tr_pre, const float* tr_post,
                      const float* g);
void orc_rec_decay_factors(float dt, float* out);  /* 7 values */
float orc_nmda_block(float v);
/* compute_currents: i_out[i] += receptor currents of g[7][n] at v[i] */
void orc_rec_currents(const float* g, const float* v, float* i_out, uint32_t n);
/* v += val[k] on state column 0 of neuron idx[k] of pop (idx NULL: count ==
 * the population's size, neuron k gets val[k]), before the next step. */
int orc_net_inject_voltage(void* net, int32_t pop, const uint32_t* idx, const float* val, uint32_t count);
/* Electrical coupling, global neuron indices, in call order
 * (P/src/api/net/network/core.rs:391-553, step.rs:36-104). */
void orc_net_add_gap_junction(void* net, uint32_t a, uint32_t b, float g);
void orc_net_add_graded_synapse(void* net, uint32_t pre, uint32_t post, float g);
int orc_net_add_sigmoid_profile(void* net, float vth, float delta, float k, float erev, float cond, float dt);
int orc_net_add_sigmoid_conn(void* net, int32_t prof, uint32_t pre, uint32_t post, float weight);
int orc_net_inject_current(void* net, int32_t pop, const uint32_t* idx,
                           const float* val, uint32_t count);
void orc_net_step(void* net, uint32_t n_steps);
/* One step with the dense every-synapse STDP of CompiledNetwork::stdp_step. */
void orc_net_stdp_step(void* net, float decay, float lr_ltp, float lr_ltd, float w_min, float w_max);
uint64_t orc_net_step_counter(void* net);
void orc_net_spikes(void* net, int32_t* out);
uint32_t orc_net_active(void* net, uint32_t* out);
/* State column `col` of population `pop`, widened to double. */
int orc_net_state(void* net, int32_t pop, int32_t col, double* out);
/* Ring buffer [32][N] widened to double. */
void orc_net_ring(void* net, double* out);
void orc_net_weights(void* net, double* out);
uint64_t orc_net_total_spikes(void* net);
uint64_t orc_net_total_events(void* net);
/* Per-step spike recording (dense int32 raster [steps x N]). */
void orc_net_record(void* net, int32_t on);
uint64_t orc_net_raster_steps(void* net);
void orc_net_raster(void* net, int32_t* out);
void orc_net_set_threads(void* net, int32_t n_threads);

/* Stand-alone pieces for known-answer tests. */
int32_t orc_poisson_spike(uint32_t seed, uint32_t stream, uint64_t step, float p);
float orc_gauss(uint64_t seed, uint32_t i, uint64_t step);
/* tm_spike (P/src/api/net/push.rs:60-86) on one synapse: returns the delivered
 * weight and advances (r, u, last). */
float orc_tm_spike(float u0, float a_rec, float a_fac, int32_t is_fac, float* r, float* u, uint32_t* last,
                   float weight, uint64_t step);
float orc_powi_f32(float a, int32_t n);
/* fast_exp (P/src/math/fast_exp.rs:21-63, the separately rounded Horner). */
float orc_fast_exp(float x);
/* GPEA steady states and time constants at (v, cai): out = m_inf, h_inf, n_inf,
 * p_inf, q_inf, r_inf, tau_m, tau_h, tau_n, tau_p, tau_q. */
void orc_gpea_rates(float v, float cai, float* out);
float orc_cos2pi_f32(float u);
float orc_sin2pi_f32(float u);
float orc_exp_f32(float x);
float orc_ln_f32(float x);
float orc_vtrap_f32(float x, float y);
double orc_vtrap_f64(double x, double y);
uint32_t orc_compact(const int32_t* spikes, uint32_t n, uint32_t* out);

/* Kuramoto (P/src/kernels/phase/sparse.rs:40-97 in Euler form, plus RK4). */
void* orc_kura_create(int kind, uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                      const float* w, const float* omega, const float* theta0, float dt);
void orc_kura_destroy(void* k);
void orc_kura_step_rk4(void* k, uint32_t n_steps);
void orc_kura_step_euler(void* k, uint32_t n_steps);
void orc_kura_phases(void* k, double* out);
double orc_kura_order(void* k);

/* Creature pool of the game engine (crates/dioma-engine/src/brain/dynamics/
 * gpu_pool.rs) restated with the reference's own per-post pull scatter
 * (crates/SiMuse/crates/pillars-wgpu/shaders/csr_scatter.wgsl:26-59 over the
 * in-edge CSR of crates/SiMuse/crates/pillars-wgpu/src/network.rs:231-257):
 * creatures' CSR blocks concatenated block-diagonally, alpha-LIF neurons,
 * weights signed by the presynaptic neuron's transmitter, one axonal delay
 * (delay < 0: min(round(t_delay/dt) + 1, 31)). */
typedef struct {
    float dt_ms, tau_ref, tau_m, v_rest, v_thresh, v_reset, tau_syn, t_delay;
} orc_pool_params;
void* orc_pool_create(const orc_pool_params* p, int32_t delay);
void orc_pool_destroy(void* pool);
/* Appends a creature (its CSR with local indices); returns its index. */
int32_t orc_pool_register(void* pool, uint32_t n, const int32_t* row_ptr, const int32_t* col,
                          const float* values, const float* nt_types);
int orc_pool_set_weights(void* pool, int32_t creature, const float* w, uint64_t n_syn);
/* per-synapse delays of a creature (a new network, as WgpuNetwork::new takes them) */
int orc_pool_set_delays(void* pool, int32_t creature, const uint8_t* d, uint64_t n_syn);
/* spt substeps; ext (n_total values, NULL = zeros) is added to v once before
 * the first substep of every run of up to 64 substeps; spikes of each substep
 * into out[spt][n_total] (0/1). */
int orc_pool_step(void* pool, const float* ext, uint32_t spt, uint32_t* out);
/* state column (0 v, 1 g, 2 refr_cnt) of every neuron */
int orc_pool_state(void* pool, int32_t col, float* out);
uint32_t orc_pool_n_total(void* pool);
int32_t orc_pool_delay(void* pool);

/* Statistics (P/src/kernels/stats/rates.rs:6-35, isi.rs:30-67). */
void orc_firing_rates(const int32_t* raster, uint32_t n, uint32_t steps, float dt_s, float* out);
void orc_isi_cv(const int32_t* raster, uint32_t n, uint32_t steps, float dt_s, float* out);

#ifdef __cplusplus
}
#endif
#endif
