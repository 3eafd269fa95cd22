// net_kernels.hpp — device data layout shared by the network kernels and the
// host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace simuse {

// A pointer into device memory.  Device code sees a global (address space 1)
// pointer, so its loads, stores and atomics compile to global_* instructions
// (scalar base plus 32-bit offset; unlike flat_* they do not count against
// the LDS/scalar-load counter, so LDS work does not wait for them); host code
// sees a plain pointer.  Host code stores into these fields through dev_ptr.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using dptr = __attribute__((address_space(1))) T*;
#else
template <class T>
using dptr = T*;
#endif
template <class T>
__host__ __device__ inline dptr<T> dev_ptr(T* p) {
    return (dptr<T>)p;
}
// Read-only device data seen through the constant address space: uniform
// loads through it compile to scalar (s_load) loads.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using cptr = __attribute__((address_space(4))) const T*;
#else
template <class T>
using cptr = const T*;
#endif

constexpr uint32_t kTile = 256;          // neurons per tile = threads per block
constexpr uint32_t kWavesPerTile = kTile / 64;
constexpr uint32_t kSlots = 32;          // P/src/api/net/graph.rs:33
constexpr uint32_t kEmpty = 0xFFFFFFFFu; // end of a scatter chain
constexpr int kMaxParams = 16;
// Receptor types AMPA, NMDA, GABA_A, GABA_B, nAChR, GluCl, mAChR: decay time
// constant (ms) and reversal potential (mV), P/src/kernels/synapse/receptor.rs:26-99.
constexpr int kReceptors = 7;
constexpr float kRecTau[kReceptors] = {2.0f, 50.0f, 10.0f, 200.0f, 3.0f, 5.0f, 150.0f};
constexpr float kRecErev[kReceptors] = {0.0f, 0.0f, -65.0f, -90.0f, 0.0f, -70.0f, -90.0f};
constexpr int kRecNmda = 1;  // the receptor with the Mg2+ block
constexpr uint32_t kFixTiles = 8;        // scatter tiles per fix-up block
constexpr uint32_t kMaxGroup = 2;        // neuron tiles per block of the neuron kernel
constexpr uint32_t kMaxNeurons = 1u << 27;  // per network: slot*N + post fits 32 bits
constexpr int kChain = 8;                // chain entries sorted in LDS per thread
// Binned scatter (networks whose delays are all <= 1 step): posts per bin,
// most bins per network, events one apply block orders in LDS at a time, and
// most bin blocks (groups of consecutive scatter tiles) per network.
constexpr uint32_t kBinLog2 = 11;
constexpr uint32_t kBinPosts = 1u << kBinLog2;
constexpr uint32_t kMaxBins = 4096;
constexpr uint32_t kXcds = 8;          // MI355X: workgroups go round-robin over 8 XCDs
constexpr uint32_t kXcdChunk = 16;     // consecutive work items per XCD turn (see Engine::xcd_order)
constexpr uint32_t kXcdPad = kXcds * kXcdChunk;  // most pad items an XCD-ordered list gains
constexpr uint32_t kBinTiles = 8;       // tiles per bin
constexpr uint32_t kMaxBinPops = 64;    // populations of a binned network
constexpr uint32_t kBinCap = 2048;      // records bin_apply orders in one LDS pass
constexpr uint32_t kMaxWriters = 2048;  // binned writers per network
constexpr uint32_t kBinBlockTiles = 4;  // source tiles per bin_kernel writer
constexpr uint32_t kLtpItem = 0xFFFFFFFEu;  // bin_kernel item that runs in-edge LTP instead of writing
constexpr uint32_t kFrontParts = 2;  // split form: front_kernel blocks per bin, kBinPosts / kApply cells each
// 8-byte records: the event's index within its writer's events of the step
// takes the 21 bits above the cell; a writer spanning more edges than that
// makes its network use 16-byte records.  Writer index (< kMaxWriters = 2^11)
// and f together form a 32-bit key ascending with the event id.
constexpr uint32_t kRecFBits = 21;
constexpr uint32_t kCplPopMask = 0xFFFFu;  // coupling entry: population bits of .y
constexpr int kSigProf = 6;                // floats per sigmoid graded profile

enum : uint32_t {
    kPopRefrElided = 1u,  // refractory timer provably stays 0: not loaded/stored
    kPopNoise = 2u,       // Gaussian current noise enabled
    kPopNoiseScalar = 4u, // every neuron has the same sigma (sigma_uni)
    kPopDrive = 8u,       // per-neuron Poisson drive (one source per neuron)
};

// One population as the kernels see it.  Parameter column k is either a
// per-neuron array pr[k*dstride + i] or, when bit k of uni_mask is set, the
// single value puni[k] (uniform columns cost no HBM traffic).  puni[0]
// always holds dt of neuron 0, which the reference broadcasts to the whole
// population (P/src/core/exec/population.rs:307,361).
struct PopDev {
    int32_t model;
    uint32_t n;
    uint32_t off;       // global index of neuron 0
    uint32_t tile0;     // global index of the first tile
    uint32_t word0;     // first 64-bit spike word
    uint32_t dstride;   // device column stride (multiple of 64)
    uint32_t uni_mask;
    uint32_t flags;
    dptr<float> st;          // [ns][dstride]
    dptr<const float> pr;    // [np][dstride]
    dptr<const float> sigma; // [n] noise amplitude
    float sigma_uni;    // noise amplitude when kPopNoiseScalar
    float drive_p;      // Poisson drive: per-step probability
    float drive_w;      //                weight of one drive event
    uint32_t drive_seed0;  // seed of neuron i: drive_seed0 + i * 2654435761 (mod 2^32)
    uint64_t seed;
    float puni[kMaxParams];
    float gdec[kReceptors];  // receptor conductance decay per step, fast_exp(-dt / tau_r)
};

// Everything one step of one network touches.  An engine keeps an array of
// these in device memory; every kernel block is handed a (network, tile) pair.
struct NetArgs {
    dptr<const PopDev> pops;
    dptr<const uint32_t> tile_pop;   // tile -> population
    uint32_t N;                 // neurons
    uint32_t n_tiles;           // neuron tiles
    uint32_t n_src_tiles;       // Poisson-source tiles (after the neuron tiles)
    uint32_t n_src;
    uint32_t n_blocks;          // n_tiles + n_src_tiles (scatter tiles of this network)
    uint32_t n_fix;             // fix-up blocks: ceil(n_blocks / kFixTiles)
    uint64_t n_syn;
    uint64_t step_limit;        // kernels skip the network once *step reaches it
    dptr<float> ring;                // [32][N]
    int32_t ring_fresh;         // every delay <= 1: this step's target slot holds zeros
    // Binned scatter (n_bins > 0): bins of kBinTiles tiles; writer j (of
    // n_bblk, ascending event ids) puts its events grouped by bin from its
    // first edge id, and bin_hdr[k][j] = {first record, count} of its group
    // for bin k.  Megastep networks keep two buffers of each, by step parity:
    // records from ev_buf(a, p) and bin_hdr + p*hdr_stride.  A record is
    // 8 bytes, {weight bits, f << 11 | cell} with f the event's index in its
    // writer's concatenation of spiking rows (ascending event id within the
    // writer, below 2^kRecFBits), or, in networks with receptor synapses or a
    // writer spanning more edges (ev_wide), 16 bytes {event id, weight bits,
    // cell, receptor code}.
    uint32_t n_bins;
    uint32_t n_bblk;            // writers
    int32_t bin_fused;          // the neuron kernel's blocks are the neuron-tile writers
    int32_t mega;               // one-kernel or split form: parity buffers; writers write every
                                // header entry of their column, consumers leave the rows as read;
                                // the last of n_done blocks moves the step on
    int32_t split;              // split form (front_kernel + bin_kernel)
    uint32_t n_done;            // blocks of a step counted by mega_done
    uint32_t bin_rows;          // row length of bin_hdr (even, >= n_bblk)
    uint32_t n_pops;
    uint64_t n_ev;              // synapses + source edges
    uint64_t hdr_stride;        // n_bins * bin_rows
    dptr<uint2> bin_hdr;             // [1 or 2][n_bins][bin_rows]
    dptr<uint2> bin_ev;              // [1 or 2][n_ev] records, 8 B (or 16 B when ev_wide)
    int32_t ev_wide;            // 16-byte records
    dptr<uint32_t> bin_slot;         // ring slot bin_apply fills, kEmpty past the horizon
    dptr<uint32_t> bin_done;         // megastep: blocks of the step that finished
    dptr<uint32_t> bin_mat;          // megastep: the pending events are in ring slot step % 32
    // Scatter bookkeeping.  head[key] = (step tag << 32) | event id of the
    // last event that reached key = slot*N + post in the step with that tag;
    // a head with another tag is idle, so heads are never cleared per step.
    dptr<unsigned long long> head;   // [32][N]
    dptr<uint2> rec;                 // [events] per event: {earlier event of its key or kEmpty,
                                //   ring value before the first add (first event only)}
    dptr<uint2> coll;                // [events] {event, key} of events that found their key
                                //   taken, per block from tile_ebase[first tile]
    dptr<uint32_t> tile_coll;        // [n_blocks] how many, this step
    dptr<const uint32_t> tile_ebase; // [n_blocks] first event id of the tile's rows
    dptr<const uint32_t> row_ptr;    // [N+1]
    uint32_t uni_deg;                // every row's out-degree when they are all equal, else 0
    dptr<const uint32_t> col;        // [n_syn]
    dptr<const float> w;             // [n_syn]
    dptr<const uint8_t> delay;       // [n_syn]
    dptr<const uint32_t> src_row_ptr;
    dptr<const uint32_t> src_post;
    dptr<const float> src_w;
    dptr<const uint8_t> src_delay;
    dptr<const float> src_p;
    dptr<const uint32_t> src_seed;
    dptr<uint32_t> tile_count;       // [n_tiles] spikes of the last step per tile
    dptr<uint32_t> active;           // [n_tiles*kTile] ascending ids per tile
    dptr<uint64_t> bits;             // spike bit words of the last step
    dptr<uint64_t> tile_spikes;      // [n_tiles] running totals
    dptr<uint64_t> tile_events;      // [n_tiles + n_src_tiles] running totals
    dptr<uint64_t> step;             // device step counter
    dptr<uint64_t> raster;           // optional [raster_cap][n_words]
    uint32_t raster_cap;
    uint32_t n_words;
    dptr<uint64_t> raster_base;      // step at which recording started
    // pair-based STDP
    int32_t stdp;
    dptr<double> w64;                // FP64 master weights
    dptr<float> tr_pre;
    dptr<float> tr_post;
    dptr<const uint32_t> in_ptr;     // [N+1]
    dptr<const uint32_t> in_edge;    // [n_syn] in-edges per post, ascending edge id
    dptr<const uint32_t> edge_pre;   // [n_syn]
    dptr<const uint8_t> plastic;     // [n_syn] or null (all)
    float stdp_dpre, stdp_dpost, a_plus, a_minus, w_min, w_max;
    // Tsodyks-Markram short-term plasticity per event id (synapses, then
    // source edges); null when no edge carries a profile
    dptr<const uint8_t> tm_type;     // profile id, 0 = static
    dptr<float> tm_r;
    dptr<float> tm_u;
    dptr<uint32_t> tm_last;
    dptr<const float4> tm_prof;      // {u0, e^(-dt/tau_rec), e^(-dt/tau_fac), facilitating}
    dptr<float> ev_w;                // delivered weight of each event this step (for the fix-up)
    // conductance-based receptor synapses (rec_code == nullptr: off)
    dptr<const uint8_t> rec_code;    // per synapse: 0 current-based, 1 + receptor type
    dptr<float> g;                   // [kReceptors][N] conductances
    dptr<const float> g_ext;         // [N] constant external current, or nullptr
    float rec_scale;            // a jump adds weight * rec_scale
    uint32_t rec_live;          // bit r: some synapse has receptor type r (else g[r] stays 0)
    uint32_t n8;                // N & ~7: neurons below it sum their receptor currents onto the input
    // Electrical coupling (gap junctions, graded and sigmoid graded synapses;
    // cp_ptr == nullptr: none).  Neuron i's contributions cp[cp_ptr[i] ..
    // cp_ptr[i+1]) in the reference's order: {other's local index, other's
    // population | kind << 16 | sigmoid profile << 18, conductance or weight
    // bits, sigmoid connection}; kind 0 gap junction (i is its first end),
    // 1 gap junction (second end), 2 graded synapse, 3 sigmoid graded synapse.
    dptr<const uint32_t> cp_ptr;
    dptr<const uint4> cp;
    dptr<float> sig_s;               // sigmoid activation per connection
    dptr<const float> sig_prof;      // per profile: vth, 1/delta, k, erev, conductance, dt
};

}  // namespace simuse
