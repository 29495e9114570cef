// Device-side data layout and kernel entry points of the per-timestep path.
// Shared by the host engine (engine.cpp) and the HIP kernels (*.hip).
//
// Layout in HBM (one GPU, all resident):
//   node state     SoA float arrays per group: var[v][n], plus x (double),
//                  h (float), refractory (float) per RK5 neuron
//   accumulators   acc[port][n_total] float, cleared by the consumer kernel
//   spike ring     R slots x n_total entries: src (u32), height (f32),
//                  count (u32); per-slot counter ring_n[R]
//   synapses       CSR by source, delay-sorted within source:
//                  rec[s] = {u32 target | port << 26, f32 weight} (8 B)
//                  base[src] (u64) + seg[src][0..D] (u32, relative)
//   direct (Poisson) connections, CSR by target: dir_ptr[n_total + 1]
#pragma once
#include <stdint.h>

#include "conn_dist.h"

namespace ngpu {

constexpr int kMaxPort = 20;          // src/aeif_cond_beta.h:26
constexpr int kPortShift = 26;        // target in low 26 bits, port above
constexpr uint32_t kTargetMask = (1u << kPortShift) - 1u;
constexpr int kPoissonTableMax = 256;
constexpr int kMaxSynDt = 16384;      // src/syn_model.h:21
constexpr long long kLastSpikeTimeGuard = 0x70000000LL;
constexpr int kMaxAttemptsPerStep = 100000;
constexpr int kCounterSlots = 1024;   // spread statistic counters over L2
constexpr uint32_t kPoissonTag = 0x504F4953u;
constexpr uint32_t kLegacyPoissonTag = 0x4C504F49u;
constexpr int kStragRecBytes = 24;    // resume record of the two-pass update
// second RK5 passes overlap the next step for groups of up to this many
// receptor ports
constexpr int kOverlapMaxPort = 4;

// Connection rules generated on the device (connect_rules.hip).
enum ConnGenRule {
  kGenOneToOne = 0,
  kGenAllToAll = 1,
  kGenFixedTotalNumber = 2,
  kGenFixedIndegree = 3,
  kGenFixedOutdegree = 4,
};
constexpr size_t kConnGenMaxLds = 64 * 1024;  // sampler state per item
struct ConnGenArgs {
  int rule;
  const int *s_arr, *t_arr;     // node lists (nullptr: ranges s0.., t0..)
  int s0, ns, t0, nt;
  int K;                        // in/outdegree
  int cap;                      // rejection table size (set by the launcher)
  uint32_t port_word;           // port << kPortShift
  uint8_t sg_val;
  float w0, d0, dt;
  const float *w_arr, *d_arr;   // per connection, or nullptr
  DistSpec wd, dd;              // else drawn per connection (kind != 0)
  uint32_t k0, k1, call;        // stream key and call index
  unsigned long long m;         // connections
  uint32_t *src, *tp;           // outputs at the chunk's offset
  float *w;
  uint16_t *d;
  uint8_t *sg;                  // nullptr without plasticity
};
// LDS bytes of the sampler of an in/outdegree rule (0 for the others)
size_t ConnGenLdsBytes(int rule, int ns, int nt, int K, int *cap_out);
const char *LaunchConnGen(const ConnGenArgs &a, void *stream);

// Step clock in device memory.// Step clock in device memory.  Kernels launched with a non-null clock
// read the step-dependent values from it instead of from their arguments,
// so captured steps (hipGraph) replay unchanged.  Two clocks alternate
// (even and odd steps of a capture): the last kernel of a step writes the
// next step's values into the other one (clock_next, from block 0), which
// no kernel of its own step reads.  The arithmetic is the host's
// (SimulationStep):
//   T = T_base + it + 1,  t = t0 + dt * (it + 1)  (double, no fused ops)
struct StepClock {
  long long gs;       // global step index (spike ring slot gs % R)
  long long it;       // step index within the current Simulate call
  long long T;        // time index of the step
  double t;           // neural time at the end of the step (ms)
  long long pois_T;   // time index of the pending Poisson draw (<0 none)
  double pois_t;      // neural time of that step
  long long T_base;
  double t0, dt;
};

// A parameter that is either uniform over a group (val) or per neuron (ptr).
struct PField {
  const float *ptr;
  float val;
};

struct AeifParams {
  PField V_th, Delta_T, inv_Delta_T, g_L, E_L, inv_C_m, a, b, inv_tau_w, I_e,
      V_peak, V_reset, t_ref;
  PField E_rev[kMaxPort], inv_tau_rise[kMaxPort], inv_tau_decay[kMaxPort],
      g0[kMaxPort];
};

// Poisson direct connections feeding the targets of one group.
struct DirectIn {
  const uint32_t *ptr;        // [n_total+1] CSR by target (global node)
  const uint32_t *canon;      // canonical connection index (Philox counter)
  const uint32_t *gen;        // generator (global node) of the connection
  const float *weight;
  const float *delay;         // ms, dt*(d_int+1)
  const uint8_t *port;
  const float *gen_origin, *gen_start, *gen_stop;  // per global node
  const uint32_t *gen_table;  // per global node: table index
  const uint32_t *tables;     // [n_table][kPoissonTableMax]
  uint32_t key0, key1;
  int active;                 // any direct connections at all
  // Compact form (compact != 0), used when nodes lo .. hi-1 each have
  // exactly one direct connection, in canonical order (canonical index =
  // node - lo), all with one weight, delay and port: the per-connection
  // arrays above are not read.  The generator of node lo + j is gen0
  // (gen_mode 0), gen0 + j (1: one generator per node) or gen_of[j] (2).
  int compact;
  uint32_t lo, hi;
  int gen_mode;
  uint32_t gen0;
  const uint32_t *gen_of;
  float c_weight, c_delay;
  int c_port;
};

// Inputs to fold at the start of a node update (end-of-step port input of
// the previous step, src/get_spike.cu:116-140 and src/poiss_gen.cu:43-74).
struct FoldIn {
  float *acc;                 // [max_port][n_total]
  long long n_total;
  long long pois_T;           // time index of the pending Poisson draw, <0 none
  double pois_t;              // neural time (ms) of that step
  DirectIn dir;
};

struct SpikeOut {
  uint32_t *ring_src;         // this step's slot
  float *ring_h;
  uint32_t *ring_cnt;
  uint32_t *ring_n;           // counter of this step's slot
  // the whole ring, for a clocked launch (slot gs % R of [R][ring_cap])
  uint32_t *ring_src0;
  float *ring_h0;
  uint32_t *ring_cnt0;
  uint32_t *ring_n0;
  long long ring_cap;
  int R;
  long long *last_spike;      // [n_total]
  float *last_height;
  long long *last_rev;
  const int *den_idx;         // [n_total] dendritic delay in steps
  int *spike_count;           // per group, nullptr if off
  float *rec_times;           // per group, nullptr if off
  int *n_rec_times;
  int max_rec_times;
  long long T;
  double t;
};

struct AeifArgs {
  int n, i0;
  float *V, *w, *g[kMaxPort], *g1[kMaxPort];
  double *x;
  float *h;
  float *refr;
  AeifParams p;
  float h_min;
  double h_min_acc;           // h_min * (1 + 1e-5)
  float eps;
  float dt;
  double x1;                  // end of this step (neural time, ms)
  FoldIn fold;
  SpikeOut out;
  unsigned long long *cnt;    // [kCounterSlots][2] accepted, rejected
  unsigned long long *cnt1;   // the same for the first pass (nullptr: cnt)
  unsigned long long *forced; // [kCounterSlots] attempts accepted with a
                              // scaled error above 1 (at the minimum step)
  int do_step;                // 0: fold only (end-of-simulation flush)
  // two-pass update: first pass with at most `budget` RK5 attempts per
  // neuron, the rest finished by a second pass over 64 sub-lists of
  // resume records.  Four sets of lists serve the steps in turn by gs
  // mod 4 (strag: [4][2][64][strag_cap] records, strag_n2: [4][2][64]
  // lengths; per set, the deferred neurons' resume records, then the
  // catch-up list below); the first pass of a step clears the lengths of
  // the next step's set.  strag == nullptr selects the single-pass kernel
  void *strag;
  unsigned int *strag_n2;
  // Steps overlapped with the previous step's second pass.  The first
  // pass writes, per wave, the ballot of the lanes the second pass of its
  // step owns (mask: [2][n_waves] by gs parity).  With skip_prev set it
  // leaves the lanes the previous second pass owned alone (that pass may
  // still run): it only folds their input into carry ([2][P][n] by gs
  // parity) and lists them (catch-up); the second pass then runs their
  // whole step, and they stay with the second passes until a step runs
  // without skip_prev.
  unsigned long long *mask;
  unsigned long long n_waves;
  float *carry;
  int skip_prev;
  int mask_write;             // this step's second pass overlaps the next
  long long gs;
  const StepClock *clock;     // nullptr: step values from the arguments
  unsigned long long strag_cap;
  int budget;
  int pass2_grid;
  int uniform;                // every parameter PField of the group uniform
  // learned step-size proposer, second pass only: weights in the MFMA
  // lane layout (nullptr: plain controller) and the sample log of
  // (features, target) per attempt (nullptr: no logging)
  const unsigned char *prop_pk;
  float *prop_log;
  unsigned long long *prop_log_n;
  unsigned long long prop_log_max;
};

// Other RK5 models (rk5_models.hip): state and parameters in the group's
// SoA arrays [index][n], reference enum order.
enum Rk5Kind {
  kRk5PscAlpha = 0,   // aeif_psc_alpha
  kRk5PscExp,         // aeif_psc_exp
  kRk5PscDelta,       // aeif_psc_delta
  kRk5IzhCondBeta,    // izhikevich_cond_beta
};
constexpr int kRk5MaxPort = 8;  // receptor ports of these models

struct Rk5ModelArgs {
  int kind;                   // Rk5Kind
  int n, i0, n_port;
  float *var;                 // [n_var][n]
  float *param;               // [n_param][n], refractory_step written back
  double *x;
  float *h;
  float h_min;
  double h_min_acc;           // h_min * (1 + 1e-5)
  float eps;
  float dt;
  double x1;                  // end of this step (neural time, ms)
  FoldIn fold;
  SpikeOut out;
  unsigned long long *cnt;    // [kCounterSlots][2] accepted, rejected
  unsigned long long *forced; // [kCounterSlots] accepted above tolerance
  int do_step;                // 0: fold only (end-of-simulation flush)
  const StepClock *clock;
};

struct IafArgs {
  int n, i0;
  float *I_ex, *I_in, *V, *refr;
  PField I_e, Theta, V_reset, t_ref, P20, P11ex, P11in, P21ex, P21in, P22;
  float dt;
  FoldIn fold;
  SpikeOut out;
  int do_step;
  const StepClock *clock;
};

// Fixed-step point models with one input port folded into I_syn (weight 1)
// and group-wide parameters (point_models.hip).
enum PointKind {
  kPointIafG = 0,    // iaf_psc_exp_g: exact integration, propagators below
  kPointIafHc = 1,   // iaf_psc_exp_hc: the reference's fixed propagators
  kPointIzh = 2,     // izhikevich: forward Euler, input added to V_m
  kPointIzhPsc = 3   // izhikevich_psc_exp(_2s, _5s): Euler sub-steps, PSC
};
struct PointArgs {
  int kind;
  int n, i0;
  int substeps;                 // izhikevich_psc_exp: 1, 2 or 5
  float *I_syn, *V, *u, *refr;  // u: izhikevich models only
  PField I_e;
  float P11, P22, P21, P20;     // iaf_psc_exp_g
  float Theta, V_reset;         // iaf_psc_exp_g (relative to E_L)
  float V_th, a, b, c, d;       // izhikevich models
  float h, C_syn;               // Euler sub-step and its PSC decay
  int n_refr;                   // refractory steps set at a spike
  FoldIn fold;
  SpikeOut out;
  int do_step;
  const StepClock *clock;
};

// ext_neuron (src/ext_neuron.cu:23-39): per port, the input folded with
// port_weight is moved to port_value at each step, for the host to read.
struct ExtArgs {
  int n, i0, n_port;
  float *var;                   // [2 n_port][n]: port_input, port_value
  const float *port_weight;     // [n_port][n]
  FoldIn fold;
  int do_step;
  const StepClock *clock;
};

// spike_detector (src/spike_detector.cu:46-68) and parrot_neuron
// (src/parrot_neuron.cu:48-70): one input port folded into
// input_spike_height.
struct RelayArgs {
  int n, i0;
  int parrot;
  float *input, *out_var;     // detector: spike_height; parrot: V
  PField hold;
  FoldIn fold;
  SpikeOut out;
  int do_step;
  const StepClock *clock;
};

struct SpikeGenArgs {
  int n, i0;
  const int *n_spikes, *first;  // per node
  const int *time_idx;          // flattened spike time indexes
  const float *height;
  int *next;
  long long it;                 // step index within this Simulate call
  SpikeOut out;
  const StepClock *clock;
};

// nodes of CreatePoissonGenerator (src/poisson.cu:27-54): one Poisson
// count per node and step, dropped when above 5 lambda, sent as the height
// of a spike
struct LegacyPoisArgs {
  int n, i0;
  const uint32_t *tab;          // CDF thresholds of Poisson(lambda)
  float max_count;              // 5 lambda
  uint32_t key0, key1;          // Threefry key (the random seed)
  long long it;                 // step index within this Simulate call
  double t0, dt;                // neural time at the call's start, time step
  SpikeOut out;
  const StepClock *clock;
};

struct DeliveryArgs {
  const uint32_t *ring_src;     // [R][cap]
  const float *ring_h;
  const uint32_t *ring_cnt;
  uint32_t *ring_n;             // [R]
  long long cap;
  int R;
  long long gs;                 // global step index of this step
  int D;                        // number of delay slots (max d_int + 1)
  int tile_max;                 // largest (spike, delay) pairs per wave tile,
                                // a power of two in [1, 64]
  const unsigned long long *base;  // [n_total+1]
  const uint32_t *seg;          // [n_total][D+1]
  unsigned long long *rec;      // synapse records, weight in high word
  float *acc;
  long long n_total;
  // plasticity
  const uint8_t *syn_group;     // nullptr if no plastic synapse
  uint16_t *spike_time;
  const long long *last_rev;
  const float *stdp_param;      // [n_syn_group][8]
  const int *syn_type;          // [n_syn_group]
  long long T;
  float dt;
  unsigned long long *events;   // [kCounterSlots]
  const StepClock *clock;       // nullptr: gs and T from the arguments
  StepClock *clock_next;        // last launch of a captured step only
};

// Binned delivery (non-plastic networks): the step's synaptic events are
// counted per (block, accumulator range of kBinLen entries), placed
// range-major into an event buffer, summed per range in LDS and added to
// the accumulators once per entry instead of once per event.
constexpr int kBinShift = 13;
constexpr int kBinLen = 1 << kBinShift;   // accumulator entries per bin
constexpr int kMaxBins = 8192;            // sub-bins a block can count in LDS
constexpr int kSubBinsTarget = 64;        // preferred largest sub-bin count
constexpr int kReduceChunks = 512;        // target chunk count of a step
constexpr int kBinChunkMin = 4096;        // fewest events per chunk
constexpr int kSingleChunkBins = 128;     // from this many bins on, balanced
                                          // bins are reduced in one chunk each

struct BinArgs {
  // The count and scatter phases group events by sub-bin (2^sub_shift
  // accumulator entries, at most kBinLen), fine enough that the lanes of a
  // wavefront rarely hit the same LDS counter; a bin is a run of
  // kBinLen >> sub_shift consecutive sub-bins, contiguous in the buffer.
  uint32_t *hist;           // [nsub][grid] counts, then offsets within the
                            // sub-bin
  uint32_t *tot;            // [nsub] events per sub-bin
  uint32_t *bstart;         // [nb+1] first event of each bin
  uint32_t *cstart;         // [nb+1] first chunk of each bin
  unsigned long long *ev;   // [ev_cap] (index within bin, value bits)
  float *part;              // [chunk_cap][kBinLen] partial sums,
                            // chunk_cap = kReduceChunks + nb
  uint32_t *meta;           // [2] chunk size of the step, and whether
                            // some bin has more than one chunk
  long long ev_cap;         // above this many events the step falls back
                            // to atomic delivery
  int nb;                   // bins: ceil(acc_n / kBinLen)
  int nsub;                 // sub-bins: ceil(acc_n >> sub_shift)
  int sub_shift;
  int grid;                 // blocks of the count and scatter phases
  long long acc_n;          // accumulator entries (ports x nodes)
};

// Tiled delivery (non-plastic networks): a tile-major copy of the synapse
// store (records grouped by accumulator tile of 2^tile_shift entries, canonical
// order inside a tile) and, per (source, delay slot, tile), the first
// record of that group.  A block owns one tile in LDS: it walks every due
// (spike, delay) pair's records for its tile, sums them in LDS, and adds
// the tile to the accumulators (G == 1) or writes it as a partial that a
// second launch sums (G blocks per tile).  Each record is read once and
// the accumulators are written once per entry.  The tile holds
// 2^tile_shift entries (at most 2^15, 128 KB of LDS).
constexpr int kTileShiftMax = 15;
constexpr int kTiledBlock = 1024;

struct TileArgs {
  const unsigned long long *trec;  // [S] tile-major records
  const uint32_t *off;             // [(N*D + 1) * T]: off[(s*D + d)*T + t]
                                   // = first record of (s, d) in tile t;
                                   // the next row ends it
  // tile t is summed by G_t = tg[t+1] - tg[t] blocks; block b takes
  // share k = blk[b] >> 16 of the pairs for tile t = blk[b] & 0xFFFF (the
  // host orders the blocks so that the blocks of one share of every tile
  // run on one XCD and share its L2).  A tile of one block adds its sums
  // to the accumulators, else its block k writes partial tile tg[t] + k,
  // part[tg[t] + k][2^tile_shift]
  const uint32_t *tg;              // [T + 1]
  const uint32_t *blk;             // [n_blocks]
  float *part;
  long long *part_q;               // the same as fixed-point sums
  long long acc_n;                 // accumulator entries (ports x nodes)
  int T;                           // tiles
  int n_blocks;                    // tg[T]
  int split;                       // some tile has more than one block
  int tile_shift;
  // fixed != 0: the tile sums height*weight in 64-bit fixed point,
  // round(height*weight * q_scale) with q_scale = 2^F and q_inv = 2^-F
  // (exact sums, independent of the order of the events; F leaves the
  // largest possible sum of an entry below 2^62); an event with
  // |height| > h_max is added with a float atomic instead
  int fixed;
  float q_scale, q_inv, h_max;
};
constexpr float kTileHeightMax = 1024.0f;
constexpr int kTiledPairTile = 32;  // most pairs per wave tile, tiled mode

struct RevArgs {
  const uint32_t *ring_src;
  long long cap;
  int R;
  long long gs;
  const uint32_t *ring_n;
  const int *den_idx;
  int max_den;
  const unsigned long long *rev_ptr;  // [n_total+1]
  const uint32_t *rev_idx;            // synapse index
  unsigned long long *rec;
  const uint8_t *syn_group;
  const uint16_t *spike_time;
  const float *stdp_param;
  const int *syn_type;
  long long *last_rev;                // written by the dendritic-delay pass
  long long T;
  float dt;
  const StepClock *clock;
  StepClock *clock_next;              // last launch of a captured step only
};

struct RecordArgs {
  int n_col;
  float *const *ptr;            // per column; nullptr = spike activity
  const int *node;              // per column (global node) for spikes
  const long long *last_spike;
  const float *last_height;
  long long T;
  float *out;                   // row base
};

// launchers (kernels.hip); all asynchronous on `stream`
void LaunchAeif(const AeifArgs &a, int n_port, void *stream);
// learned step-size proposer on n feature vectors ([n][6]) -> out[n]
void LaunchPropEval(const unsigned char *pk, const float *feat, int n,
                    float *out, void *stream);
// the same for host arrays feat / out, staged through a file-scope device
// array of kPropEvalNoAllocMax vectors instead of allocated buffers (no
// hipMalloc / hipFree); synchronous.  Calls that overlap on one GPU share
// that array and race.  Returns an error message, or nullptr.
constexpr int kPropEvalNoAllocMax = 65536;
const char *PropEvalNoAlloc(const unsigned char *pk, const float *feat,
                            int n, float *out, void *stream);
// second pass of the two-pass update; with d != nullptr the delivery of
// the step (non-plastic) runs in the same launch on d_grid more blocks
void LaunchAeifPass2(const AeifArgs &a, int n_port, const DeliveryArgs *d,
                     int d_grid, void *stream);
void LaunchIaf(const IafArgs &a, void *stream);
// returns an error message, or nullptr
const char *LaunchRk5Model(const Rk5ModelArgs &a, void *stream);
void LaunchRelay(const RelayArgs &a, void *stream);
void LaunchSpikeGen(const SpikeGenArgs &a, void *stream);
void LaunchLegacyPoisson(const LegacyPoisArgs &a, void *stream);
const char *LaunchPoint(const PointArgs &a, void *stream);
void LaunchExt(const ExtArgs &a, void *stream);
void LaunchDelivery(const DeliveryArgs &a, int grid, void *stream);
void LaunchClockSet(StepClock *c, const StepClock &v, void *stream);
// Binned delivery: count, scan, scatter, reduce, sum (five launches).
// Requires a non-plastic network and b.nb <= kMaxBins.
void LaunchDeliveryTiled(const DeliveryArgs &a, const TileArgs &b,
                         void *stream);
void LaunchDeliveryBinned(const DeliveryArgs &a, const BinArgs &b,
                          int reduce_grid, void *stream);
void LaunchRevStdp(const RevArgs &a, int grid, void *stream);
void LaunchDenRev(const RevArgs &a, int grid, void *stream);
void LaunchRecord(const RecordArgs &a, void *stream);
void LaunchSpikeTimeWrap(uint16_t *spike_time, unsigned long long n, int up,
                         void *stream);
void LaunchFillU16(uint16_t *p, unsigned long long n, uint16_t v, void *stream);

}  // namespace ngpu
