// Internal state of the NeuronGPU engine (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "neurongpu.h"
#include "ngpu_kernels.h"

namespace ngpu {

#define NGPU_HIP(call)                                                      \
  do {                                                                      \
    hipError_t err_ = (call);                                               \
    if (err_ != hipSuccess)                                                 \
      throw ngpu_exception(std::string("HIP error: ") +                     \
                           hipGetErrorString(err_) + " in " #call);          \
  } while (0)

enum ModelId {
  kAeifCondBeta = 0,
  kAeifCondAlpha,
  kIafPscExp,
  kSpikeGenerator,
  kPoissonGenerator,
  kSpikeDetector,
  kParrotNeuron,
  kAeifPscAlpha,
  kAeifPscExp,
  kAeifPscDelta,
  kIzhikevichCondBeta,
  kUserM1,          // src/user_m1_kernel.h: aeif_cond_beta under another name
  kUserM2,          // src/user_m2_kernel.h: the same
  kIafPscExpG,      // fixed-step models of point_models.hip
  kIafPscExpHc,
  kIzhikevich,
  kIzhikevichPscExp,
  kIzhikevichPscExp2s,
  kIzhikevichPscExp5s,
  kExtNeuron,       // src/ext_neuron.cu
  kPoissonLegacy,   // CreatePoissonGenerator (src/poisson.cu); no model name
  kNumModels
};

// AdEx conductance models (two-pass RK5 kernels)
inline bool IsAeifCond(int m) {
  return m == kAeifCondBeta || m == kAeifCondAlpha || m == kUserM1 ||
         m == kUserM2;
}
inline bool IsCondBeta(int m) {
  return m == kAeifCondBeta || m == kUserM1 || m == kUserM2;
}
// the other RK5 models (rk5_models.hip): their Rk5Kind, or -1
inline int Rk5KindOf(int m) {
  switch (m) {
    case kAeifPscAlpha: return kRk5PscAlpha;
    case kAeifPscExp: return kRk5PscExp;
    case kAeifPscDelta: return kRk5PscDelta;
    case kIzhikevichCondBeta: return kRk5IzhCondBeta;
    default: return -1;
  }
}

// the fixed-step models of point_models.hip: their PointKind, or -1
inline int PointKindOf(int m) {
  switch (m) {
    case kIafPscExpG: return kPointIafG;
    case kIafPscExpHc: return kPointIafHc;
    case kIzhikevich: return kPointIzh;
    case kIzhikevichPscExp:
    case kIzhikevichPscExp2s:
    case kIzhikevichPscExp5s: return kPointIzhPsc;
    default: return -1;
  }
}

struct ModelDesc {
  const char *name;
  std::vector<std::string> scal_var, port_var, scal_param, port_param,
      group_param, array_param;
  std::vector<float> scal_param_default, port_param_default,
      group_param_default, scal_var_default, port_var_default;
  int fixed_ports;  // -1: chosen at Create
};

const ModelDesc &Model(ModelId id);
int FindModel(const std::string &name);
struct Engine;
// generates a rule-only chunk on the host (api.cpp)
void MaterializeRule(Engine &e, size_t chunk);
// iaf_psc_exp propagator P32 (engine.cpp)
double IafP32(double tau_syn, double tau, double C, double h);
// inversion thresholds of Poisson(lambda) (engine.cpp); returns the table
// length, or -1 when max_len entries do not reach the top of the CDF
int PoissonTable(double lambda, uint32_t *tab, int max_len);

// device buffer owned by the engine
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) NGPU_HIP(hipMalloc((void **)&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count, size_t off = 0) {
    if (count) NGPU_HIP(hipMemcpy(p + off, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(T *h, size_t count, size_t off = 0) const {
    if (count) NGPU_HIP(hipMemcpy(h, p + off, count * sizeof(T), hipMemcpyDeviceToHost));
  }
  ~DBuf() { release(); }
};

struct Group {
  ModelId model;
  int i0 = 0, n = 0, n_port = 0;
  int n_var = 0, n_param = 0;
  std::vector<float> var;          // host SoA [n_var][n]
  std::vector<float> param;        // host SoA [n_param][n]
  std::vector<float> group_param;
  // spike generator
  std::vector<std::vector<float> > sg_times, sg_heights;
  // recording of spikes
  bool spike_count_on = false;
  int max_rec_times = 0;
  std::vector<int> h_spike_count;
  std::vector<int> h_n_rec_times;
  std::vector<float> h_rec_times;
  // device
  DBuf<float> d_var, d_param;
  DBuf<double> d_x;
  DBuf<float> d_h;
  std::vector<DBuf<float> *> d_derived;  // per-neuron derived parameters
  DBuf<int> d_sg_n, d_sg_first, d_sg_tidx, d_sg_next;
  DBuf<float> d_sg_h;
  DBuf<int> d_spike_count, d_n_rec_times;
  DBuf<float> d_rec_times;
  DBuf<unsigned char> d_strag;      // resume records of the two-pass update
  DBuf<unsigned int> d_strag_n;     // [4][2][64] sub-list lengths
  DBuf<unsigned long long> d_mask;  // [2][n_waves] second-pass lanes
  DBuf<float> d_carry;              // [2][n_port][n] input of skipped neurons
  // synchronisation between host mirror and device
  bool host_newer = true;   // host arrays changed since last upload
  bool dev_newer = false;   // device advanced since last download
  bool kparam_dirty = true;
  // kernel parameters (rebuilt when dirty)
  AeifParams aeif;
  IafArgs iaf;
  PointArgs point;
  PField hold;
  float rk5_h_min = 0.f, rk5_h0 = 0.f, rk5_eps = 1e-6f;
  double rk5_h_min_acc = 0.0;

  int VarIdx(const std::string &name, int port) const;
  int ParamIdx(const std::string &name, int port) const;
  ~Group();
};

struct SynGroupDesc {
  int type;  // 1 test_syn_model, 2 stdp
  std::vector<std::string> names;
  std::vector<float> param;
};

struct Record {
  std::string file_name;
  std::vector<std::string> var_name;
  std::vector<int> node, port;
  std::vector<std::vector<float> > data;
  FILE *fp = nullptr;
};

class Engine {
 public:
  Engine();
  ~Engine();

  // kernel parameters (src/neurongpu.cu:59-85)
  float time_resolution = 0.1f;
  float sim_time = 1000.0f;
  int verbosity = 4;
  int max_spike_buffer_size = 20;
  float max_spike_num_fact = 1.0f;
  float max_spike_per_host_fact = 1.0f;
  unsigned long long seed = 54321ULL;
  int remote_spike_height_flag = 0;
  unsigned long long rng_calls = 0;  // host random stream counter

  // two-pass RK5 update (false: one neuron per lane until done)
  bool rk5_two_pass = true;
  int rk5_budget = 1;          // RK5 attempts per neuron in the first pass
  int rk5_pass2_grid = 256;   // blocks of the second pass
  bool calibrated = false;
  bool first_simulation = true;
  bool in_simulation = false;
  double t_min = 0.0, neural_time = 0.0, neur_t0 = 0.0;
  long long it = 0, Nt = 0, gs = 0;
  long long T_base = 0;

  std::vector<Group *> groups;
  std::vector<int> node_group;
  int n_node = 0;

  // A connection rule left for generation on the device at calibration
  // (connect_rules.hip): node lists and per-connection arrays are copies
  // of the caller's.  Queries made before calibration generate it on the
  // host instead (MaterializeRule), with the same result.
  struct ConnRule {
    int rule = 0;                 // ConnGenRule
    std::vector<int> s_arr, t_arr;
    int s0 = 0, ns = 0, t0 = 0, nt = 0, K = 0;
    int port = 0, sg = 0;
    float w0 = 0.0f, d0 = 0.0f;
    std::vector<float> w_arr, d_arr;
    DistSpec wd{}, dd{};          // distributions (kind kDistNone: none)
    unsigned long long seed = 0, call = 0;
  };
  // connections in coordinate form and insertion order, one chunk per
  // Connect call, kept until calibration
  struct CooChunk {
    size_t n = 0;
    std::unique_ptr<uint32_t[]> src, tp;  // tp = target | port << 26
    std::unique_ptr<float[]> w;
    std::unique_ptr<uint16_t[]> d;        // delay index d_int
    std::unique_ptr<uint8_t[]> sg;
    std::shared_ptr<ConnRule> rule;       // set: not generated yet
  };
  // 1: Connect leaves the rules above to the device when it can
  int device_connect = 1;
  std::vector<CooChunk> coo;
  size_t n_conn = 0;
  int coo_D = 0;                      // max delay index + 1
  bool any_plastic = false;

  // canonical CSR on the host (built at calibration, or on demand)
  bool csr_valid = false;
  int D = 0;                          // delay slots
  std::vector<unsigned long long> h_base;  // [n_node+1]
  std::vector<uint32_t> h_seg;        // [n_node][D+1]
  std::vector<unsigned long long> h_rec;
  std::vector<uint8_t> h_sg;
  bool h_rec_stale = false;           // device weights changed (plasticity)
  size_t n_syn = 0;                   // synapses in the device store

  std::vector<SynGroupDesc> syn_groups;

  // device state
  hipStream_t stream = nullptr;
  // second RK5 passes of this step, launched after all first passes
  struct DeferredPass2 {
    AeifArgs a;
    int n_port;
  };
  std::vector<DeferredPass2> deferred_pass2;
  bool fuse_delivery = true;  // delivery in the last second-pass launch
  // learned RK5 step-size proposer
  DBuf<float> d_prop_w, d_prop_m, d_prop_v, d_prop_grad, d_prop_loss;
  DBuf<uint16_t> d_prop_bf16;
  DBuf<unsigned char> d_prop_pk;   // MFMA-layout inference copy
  DBuf<float> d_prop_log;
  DBuf<unsigned long long> d_prop_log_n;
  unsigned long long prop_log_max = 0;
  bool prop_log_on = false, prop_use = false;
  uint32_t prop_iter = 0;
  void PropInit();
  int delivery_grid = 2048;   // blocks of the delivery kernel
  int delivery_tile_max = 64; // largest pairs per wave tile
  // 0: float atomics per event; 1: binned (events grouped by accumulator
  // range, summed in LDS); 3: tiled (PrepareTiled); non-plastic networks
  // only.  2: auto: once the recent steps average delivery_auto_events
  // synaptic events or more (read from the event counter with a lag),
  // tiled where TiledPays(), else binned
  int delivery_mode = 2;
  int delivery_bin_capacity = 1 << 27;  // events in the binned-mode buffer
  int delivery_auto_events = 500000;
  // 2: binned delivery on a second stream, beside the second RK5 pass
  // (they are independent without plastic synapses); 1: one stream
  int delivery_streams = 2;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void EnsureStream2();
  // k > 0: while delivery is binned (and nothing needs the end-of-step
  // state of every neuron), the second RK5 pass of a step runs on the
  // second stream beside the next step's first pass and this step's
  // delivery.  The next first pass leaves that pass's neurons to the next
  // second pass, for at most k steps in a row; then a first pass waits
  // for the second pass before it and takes every neuron again.
  // 0: the second pass ends each step.
  int step_overlap = 8;
  bool ov_async_prev = false;   // the last step's second pass may still run
  bool ov_skip = false;         // this step's first pass skips its neurons
  bool ov_mask = false;         // this step's second pass may overlap
  int ov_run = 0;               // steps in a row with ov_skip
  hipEvent_t ev_p1 = nullptr, ev_p2[2] = {nullptr, nullptr};
  int ev_p2_last = 0;
  unsigned long long *h_ev_snap = nullptr;  // pinned copy of the counters
  hipEvent_t ev_snap = nullptr;
  bool snap_pending = false;
  long long snap_gs = -1, prev_snap_gs = -1;
  unsigned long long prev_snap_total = 0;
  double recent_events = -1.0;  // events per step of the recent past (<0:
                                // not measured yet)
  void UpdateEventRate();
  // grid of the atomic delivery: a small one while recent steps carried
  // almost no events (results do not depend on the grid)
  int DeliveryGridNow() const {
    return (recent_events >= 0.0 && recent_events < 4096.0 &&
            delivery_grid > 256)
               ? 256
               : delivery_grid;
  }
  // hipGraph replay of whole steps: graph_steps steps per captured graph
  // (0: never, the default: on MI355X with ROCm 7.2 a replayed step of
  // the 1M-neuron network took longer than the same launches issued one by
  // one), one graph per delivery variant (atomic, binned); the captured
  // launches read their step values from d_clock
  int graph_steps = 0;
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    int steps = 0;
  };
  StepGraph graphs[4];  // atomic, binned, atomic with the quiet grid, tiled
  DBuf<StepClock> d_clock;
  bool clock_mode = false;    // launches being captured take d_clock
  int clock_slot = 0;         // clock of the step being captured
  bool clock_synced = false;  // d_clock[0] holds the host's next step
  const StepClock *StepClockArg() const {
    return clock_mode ? d_clock.p + clock_slot : nullptr;
  }
  StepClock *StepClockNext() const { return d_clock.p + (clock_slot ^ 1); }
  void DropGraphs();
  StepClock HostClock() const;  // values of the next step
  DBuf<uint32_t> d_bin_hist, d_bin_tot, d_bin_bstart, d_bin_cstart,
      d_bin_meta;
  DBuf<unsigned long long> d_bin_ev;
  DBuf<float> d_bin_part;
  // sets up the binned-delivery buffers; false if the network cannot use it
  bool PrepareBins(BinArgs &b);
  // tiled delivery: the tile-major store copy, built on first use (false
  // if the network cannot use it); the blocks per tile, 0: automatic
  DBuf<unsigned long long> d_trec;
  DBuf<uint32_t> d_toff;
  DBuf<float> d_tpart;
  DBuf<long long> d_tpart_q;
  DBuf<uint32_t> d_ttg, d_tblk;   // blocks of each tile, (tile, share) of each block
  std::vector<uint32_t> tiled_tg;        // the uploaded tg
  bool tiled_built = false, tiled_failed = false;
  int tiled_T = 0, tiled_shift = 0;
  bool tiled_bound_ok = false;      // the two below are for this copy
  unsigned long long tiled_max_deg = 0;   // >= most records onto one entry
  float tiled_max_w = 0.0f;         // largest |weight|
  int tiled_fixed = 0;              // fixed-point tile sums (where they fit)
  int delivery_tile_split = 0;
  int delivery_tile_shift = 15;   // log2 of the accumulator entries per tile
  bool PrepareTiled(TileArgs &t);
  // the tiled delivery visits every (spike, delay) pair once per tile: it
  // pays where the store's (source, delay, tile) segments average at least
  // kTiledMinSegment records (c2-shaped networks: ~15; the 1e6-neuron
  // README network: ~0.4)
  static constexpr double kTiledMinSegment = 4.0;
  bool TiledPays() const;
  int max_port = 1;
  int R = 0;
  long long ring_cap = 0;
  DBuf<float> d_acc;
  DBuf<uint32_t> d_ring_src, d_ring_cnt, d_ring_n;
  DBuf<float> d_ring_h;
  DBuf<long long> d_last_spike, d_last_rev;
  DBuf<float> d_last_height;
  DBuf<int> d_den_idx;
  int max_den = 0;
  DBuf<unsigned long long> d_base, d_rec;
  DBuf<uint32_t> d_seg;
  DBuf<uint8_t> d_sg;
  DBuf<uint16_t> d_spike_time;
  DBuf<unsigned long long> d_rev_ptr;
  DBuf<uint32_t> d_rev_idx;
  DBuf<float> d_stdp_param;
  DBuf<int> d_syn_type;
  // Poisson direct connections (CSR by target)
  long long n_dir = 0;
  DBuf<uint32_t> d_dir_ptr, d_dir_canon, d_dir_gen, d_gen_table, d_tables;
  DBuf<float> d_dir_w, d_dir_delay, d_gen_origin, d_gen_start, d_gen_stop;
  DBuf<uint8_t> d_dir_port;
  // compact form of the direct connections (DirectIn::compact)
  bool dir_compact = false;
  uint32_t dir_lo = 0, dir_hi = 0, dir_gen0 = 0;
  int dir_gen_mode = 0, dir_port = 0;
  float dir_w = 0.0f, dir_delay = 0.0f;
  DBuf<uint32_t> d_dir_gen_of;
  bool pois_dirty = true;
  // the group of CreatePoissonGenerator (src/poisson.cu:98-147): mean count
  // per step, its CDF table, and the count above which a draw is dropped
  int legacy_group = -1;
  float legacy_lambda = 0.0f;
  std::vector<uint32_t> legacy_tab;
  DBuf<uint32_t> d_legacy_tab;
  // the last GetExtNeuronInputSpikes result (library-owned arrays)
  std::vector<int> ext_node, ext_port;
  std::vector<float> ext_height;
  // statistics
  DBuf<unsigned long long> d_cnt_rk5, d_cnt_events;
  unsigned long long n_spikes_total = 0;
  double step_seconds = 0.0;
  // pending end-of-step input (Poisson draws of the last step)
  long long pending_T = -1;
  double pending_t = 0.0;
  // records
  std::vector<Record> records;
  int rec_cols = 0;
  DBuf<float *> d_rec_ptr;
  DBuf<int> d_rec_node;
  DBuf<float> d_rec_buf;
  long long rec_rows_cap = 0, rec_rows = 0;
  std::vector<double> rec_times;

  // optional per-kernel timing with HIP events on the engine stream
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_kind;
  size_t ev_used = 0;
  double kernel_ms[8] = {0};
  long long kernel_cnt[8] = {0};
  void TimeBegin(int kind, hipStream_t s = nullptr);
  void TimeEnd(hipStream_t s = nullptr);
  void SyncAll();
  void CollectTimes();

  // helpers
  int GroupOf(int i_node, int n_node_req) const;
  int GroupOfList(const int *nodes, int n) const;
  Group &G(int i) { return *groups[i]; }
  void CheckUncalibrated(const char *msg) const;
  void SyncHost(Group &g);
  void UploadGroup(Group &g);
  void BuildKernelParams(Group &g);
  void BuildCSR();
  void BuildPoisson();
  void AllocRecordBuffer(long long rows);
  void FlushRecords();
  void WriteHostRecordRow(double t);
  void LaunchGroup(Group &g, bool do_step, long long T, double t,
                   long long it_local);
  void EnsureHostCSR();
  void BuildDeviceStore();
  // the rows of sources s0 .. s0+n-1 in three copies: base[0..n] (offsets
  // into rec, relative to base[0]), seg[(D+1)*i ..] and rec
  void SourceRangeRows(int s0, int n, std::vector<unsigned long long> &base,
                       std::vector<uint32_t> &seg,
                       std::vector<unsigned long long> &rec);
  FoldIn MakeFold(long long pois_T, double pois_t);
  SpikeOut MakeSpikeOut(Group &g, long long T, double t);
};

}  // namespace ngpu
