// C++ API of the MI355X NeuronGPU engine.
//
// Source-compatible counterpart of the reference class API
// (src/neurongpu.h:94-679, src/connect_spec.h:46-92, src/connect.h:29-45,
// src/ngpu_exception.h): same class and method names, argument meaning,
// ownership of returned arrays and error behaviour.  The implementation
// behind it (neurongpu_amd/csrc/engine.cpp) is new: SoA device state, a
// spike ring and a delay-sorted CSR-by-source synapse store on one GPU.
// MPI members exist for source compatibility; with one process they report
// rank 0 of 1 and the remote-connection calls raise ngpu_exception.
#ifndef NGPU_AMD_NEURONGPU_H
#define NGPU_AMD_NEURONGPU_H

#include <exception>
#include <string>
#include <vector>

#define ON_EXCEPTION_EXIT 0
#define ON_EXCEPTION_HANDLE 1
#define STANDARD_SYNAPSE 0

class ngpu_exception : public std::exception {
  std::string message_;

 public:
  explicit ngpu_exception(const std::string &message) : message_(message) {}
  const char *what() const noexcept override { return message_.c_str(); }
};

// contiguous node range (reference src/neurongpu.h:43-80)
class NodeSeq {
 public:
  int i0;
  int n;
  NodeSeq(int i0 = 0, int n = 0) : i0(i0), n(n) {}
  int operator[](int i) const {
    if (i < 0) throw ngpu_exception("Sequence index cannot be negative");
    if (i >= n) throw ngpu_exception("Sequence index out of range");
    return i0 + i;
  }
  NodeSeq Subseq(int first, int last) const {
    if (first < 0 || first > last)
      throw ngpu_exception("Sequence subset range error");
    if (last >= n) throw ngpu_exception("Sequence subset out of range");
    return NodeSeq(i0 + first, last - first + 1);
  }
  std::vector<int> ToVector() const {
    std::vector<int> v(n);
    for (int i = 0; i < n; i++) v[i] = i0 + i;
    return v;
  }
};

class RemoteNodeSeq {
 public:
  int i_host;
  NodeSeq node_seq;
  RemoteNodeSeq(int i_host = 0, NodeSeq node_seq = NodeSeq())
      : i_host(i_host), node_seq(node_seq) {}
};

struct ConnectionId {
  int i_source_;
  int i_group_;
  int i_conn_;
};

struct ConnectionStatus {
  int i_source;
  int i_target;
  unsigned char port;
  unsigned char syn_group;
  float delay;
  float weight;
};

enum ConnectionRules {
  ONE_TO_ONE = 0,
  ALL_TO_ALL,
  FIXED_TOTAL_NUMBER,
  FIXED_INDEGREE,
  FIXED_OUTDEGREE,
  N_CONN_RULE
};

class NeuronGPU;

class ConnSpec {
  int rule_;
  int total_num_;
  int indegree_;
  int outdegree_;

 public:
  ConnSpec();
  ConnSpec(int rule, int degree = 0);
  int Init();
  int Init(int rule, int degree = 0);
  int SetParam(std::string param_name, int value);
  int GetParam(std::string param_name);
  static bool IsParam(std::string param_name);
  friend class NeuronGPU;
};

class SynSpec {
  unsigned char syn_group_;
  unsigned char port_;
  // distribution parameters of weight / delay (no reference counterpart):
  // mu, sigma, low, high, step; weight_distr_ / delay_distr_ select the
  // distribution (0 none, 1 uniform, 2 normal, 3 normal_clipped)
  float weight_dp_[5];
  float delay_dp_[5];

 public:
  int weight_distr_;
  float *weight_array_;
  float weight_;
  int delay_distr_;
  float *delay_array_;
  float delay_;

  SynSpec();
  SynSpec(float weight, float delay);
  SynSpec(int syn_group, float weight, float delay, int port = 0);
  int Init();
  int Init(float weight, float delay);
  int Init(int syn_group, float weight, float delay, int port = 0);
  int SetParam(std::string param_name, int value);
  int SetParam(std::string param_name, float value);
  int SetParam(std::string param_name, float *array_pt);
  float GetParam(std::string param_name);
  static bool IsIntParam(std::string param_name);
  static bool IsFloatParam(std::string param_name);
  static bool IsFloatPtParam(std::string param_name);
  friend class NeuronGPU;
};

namespace ngpu {
class Engine;
}

// Per-run statistics of the hot path (no reference counterpart).
struct NeuronGPUStats {
  unsigned long long rk5_accepted;
  unsigned long long rk5_rejected;
  unsigned long long syn_events;
  unsigned long long n_spikes;
  double step_seconds;   // wall time inside SimulationStep loops
  // the part of rk5_accepted / rk5_rejected made by the first pass of the
  // two-pass AdEx update (at most rk5_budget attempts per neuron-step)
  unsigned long long rk5_pass1_accepted;
  unsigned long long rk5_pass1_rejected;
  // accepted attempts whose scaled error exceeded 1 (the acceptance test
  // lets them pass only at the minimum step size)
  unsigned long long rk5_forced;
};

class NeuronGPU {
  ngpu::Engine *e_;
  bool error_flag_;
  std::string error_message_;
  unsigned char error_code_;
  int on_exception_;

 public:
  NeuronGPU();
  ~NeuronGPU();
  NeuronGPU(const NeuronGPU &) = delete;
  NeuronGPU &operator=(const NeuronGPU &) = delete;

  int SetRandomSeed(unsigned long long seed);
  int SetTimeResolution(float time_res);
  float GetTimeResolution();
  int SetSimTime(float sim_time);
  float GetSimTime();
  int SetVerbosityLevel(int verbosity_level);
  int SetMaxSpikeBufferSize(int max_size);
  int GetMaxSpikeBufferSize();

  int GetNFloatParam();
  std::vector<std::string> GetFloatParamNames();
  bool IsFloatParam(std::string param_name);
  int GetFloatParamIdx(std::string param_name);
  float GetFloatParam(std::string param_name);
  int SetFloatParam(std::string param_name, float val);
  int GetNIntParam();
  std::vector<std::string> GetIntParamNames();
  bool IsIntParam(std::string param_name);
  int GetIntParamIdx(std::string param_name);
  int GetIntParam(std::string param_name);
  int SetIntParam(std::string param_name, int val);

  NodeSeq Create(std::string model_name, int n_neuron = 1, int n_port = 1);
  NodeSeq CreatePoissonGenerator(int n_node, float rate);
  RemoteNodeSeq RemoteCreate(int i_host, std::string model_name,
                             int n_node = 1, int n_port = 1);

  int CreateRecord(std::string file_name, std::string *var_name_arr,
                   int *i_node_arr, int n_node);
  int CreateRecord(std::string file_name, std::string *var_name_arr,
                   int *i_node_arr, int *port_arr, int n_node);
  std::vector<std::vector<float> > *GetRecordData(int i_record);

  int SetNeuronParam(int i_node, int n_neuron, std::string param_name,
                     float val);
  int SetNeuronParam(int *i_node, int n_neuron, std::string param_name,
                     float val);
  int SetNeuronParam(int i_node, int n_neuron, std::string param_name,
                     float *param, int array_size);
  int SetNeuronParam(int *i_node, int n_neuron, std::string param_name,
                     float *param, int array_size);
  int SetNeuronParam(NodeSeq nodes, std::string param_name, float val) {
    return SetNeuronParam(nodes.i0, nodes.n, param_name, val);
  }
  int SetNeuronParam(NodeSeq nodes, std::string param_name, float *param,
                     int array_size) {
    return SetNeuronParam(nodes.i0, nodes.n, param_name, param, array_size);
  }
  int SetNeuronParam(std::vector<int> nodes, std::string param_name,
                     float val) {
    return SetNeuronParam(nodes.data(), (int)nodes.size(), param_name, val);
  }
  int SetNeuronParam(std::vector<int> nodes, std::string param_name,
                     float *param, int array_size) {
    return SetNeuronParam(nodes.data(), (int)nodes.size(), param_name, param,
                          array_size);
  }
  int SetNeuronIntVar(int i_node, int n_neuron, std::string var_name, int val);
  int SetNeuronIntVar(int *i_node, int n_neuron, std::string var_name,
                      int val);
  int SetNeuronIntVar(NodeSeq nodes, std::string var_name, int val) {
    return SetNeuronIntVar(nodes.i0, nodes.n, var_name, val);
  }
  int SetNeuronIntVar(std::vector<int> nodes, std::string var_name,
                      int val) {
    return SetNeuronIntVar(nodes.data(), (int)nodes.size(), var_name, val);
  }
  int SetNeuronVar(int i_node, int n_neuron, std::string var_name, float val);
  int SetNeuronVar(int *i_node, int n_neuron, std::string var_name, float val);
  int SetNeuronVar(int i_node, int n_neuron, std::string var_name, float *var,
                   int array_size);
  int SetNeuronVar(int *i_node, int n_neuron, std::string var_name,
                   float *var, int array_size);
  int SetNeuronVar(NodeSeq nodes, std::string var_name, float val) {
    return SetNeuronVar(nodes.i0, nodes.n, var_name, val);
  }
  int SetNeuronVar(NodeSeq nodes, std::string var_name, float *var,
                   int array_size) {
    return SetNeuronVar(nodes.i0, nodes.n, var_name, var, array_size);
  }
  int SetNeuronVar(std::vector<int> nodes, std::string var_name, float val) {
    return SetNeuronVar(nodes.data(), (int)nodes.size(), var_name, val);
  }
  int SetNeuronVar(std::vector<int> nodes, std::string var_name, float *var,
                   int array_size) {
    return SetNeuronVar(nodes.data(), (int)nodes.size(), var_name, var,
                        array_size);
  }
  int GetNeuronParamSize(int i_node, std::string param_name);
  int GetNeuronVarSize(int i_node, std::string var_name);
  float *GetNeuronParam(int i_node, int n_neuron, std::string param_name);
  float *GetNeuronParam(int *i_node, int n_neuron, std::string param_name);
  float *GetNeuronParam(NodeSeq nodes, std::string param_name) {
    return GetNeuronParam(nodes.i0, nodes.n, param_name);
  }
  float *GetNeuronParam(std::vector<int> nodes, std::string param_name) {
    return GetNeuronParam(nodes.data(), (int)nodes.size(), param_name);
  }
  float *GetArrayParam(int i_node, std::string param_name);
  int *GetNeuronIntVar(int i_node, int n_neuron, std::string var_name);
  int *GetNeuronIntVar(int *i_node, int n_neuron, std::string var_name);
  int *GetNeuronIntVar(NodeSeq nodes, std::string var_name) {
    return GetNeuronIntVar(nodes.i0, nodes.n, var_name);
  }
  int *GetNeuronIntVar(std::vector<int> nodes, std::string var_name) {
    return GetNeuronIntVar(nodes.data(), (int)nodes.size(), var_name);
  }
  float *GetNeuronVar(int i_node, int n_neuron, std::string var_name);
  float *GetNeuronVar(int *i_node, int n_neuron, std::string var_name);
  float *GetNeuronVar(NodeSeq nodes, std::string var_name) {
    return GetNeuronVar(nodes.i0, nodes.n, var_name);
  }
  float *GetNeuronVar(std::vector<int> nodes, std::string var_name) {
    return GetNeuronVar(nodes.data(), (int)nodes.size(), var_name);
  }
  float *GetArrayVar(int i_node, std::string var_name);
  int IsNeuronScalParam(int i_node, std::string param_name);
  int IsNeuronPortParam(int i_node, std::string param_name);
  int IsNeuronArrayParam(int i_node, std::string param_name);
  int IsNeuronIntVar(int i_node, std::string var_name);
  int IsNeuronScalVar(int i_node, std::string var_name);
  int IsNeuronPortVar(int i_node, std::string var_name);
  int IsNeuronArrayVar(int i_node, std::string var_name);
  int SetSpikeGenerator(int i_node, int n_spikes, float *spike_time,
                        float *spike_height);

  int Calibrate();
  int Simulate();
  int Simulate(float sim_time);
  int StartSimulation();
  int SimulationStep();
  int EndSimulation();

  int ConnectMpiInit(int argc, char *argv[]);
  int MpiId();
  int MpiNp();
  int ProcMaster();
  int MpiFinalize();
  std::string MpiRankStr();

  void SetErrorFlag(bool error_flag) { error_flag_ = error_flag; }
  void SetErrorMessage(std::string error_message) {
    error_message_ = error_message;
  }
  void SetErrorCode(unsigned char error_code) { error_code_ = error_code; }
  void SetOnException(int on_exception) { on_exception_ = on_exception; }
  bool GetErrorFlag() { return error_flag_; }
  char *GetErrorMessage() { return &error_message_[0]; }
  unsigned char GetErrorCode() { return error_code_; }
  int OnException() { return on_exception_; }

  unsigned int *RandomInt(size_t n);
  float *RandomUniform(size_t n);
  float *RandomNormal(size_t n, float mean, float stddev);
  // vstep defaults to 0 so that the reference's c++/examples/brunel_net.cpp,
  // which passes five arguments, compiles
  float *RandomNormalClipped(size_t n, float mean, float stddev, float vmin,
                             float vmax, float vstep = 0.0f);

  int Connect(int i_source_node, int i_target_node, unsigned char port,
              unsigned char syn_group, float weight, float delay);
  int Connect(int i_source, int n_source, int i_target, int n_target,
              ConnSpec &conn_spec, SynSpec &syn_spec);
  int Connect(int i_source, int n_source, int *target, int n_target,
              ConnSpec &conn_spec, SynSpec &syn_spec);
  int Connect(int *source, int n_source, int i_target, int n_target,
              ConnSpec &conn_spec, SynSpec &syn_spec);
  int Connect(int *source, int n_source, int *target, int n_target,
              ConnSpec &conn_spec, SynSpec &syn_spec);
  int Connect(NodeSeq source, NodeSeq target, ConnSpec &conn_spec,
              SynSpec &syn_spec) {
    return Connect(source.i0, source.n, target.i0, target.n, conn_spec,
                   syn_spec);
  }
  int Connect(NodeSeq source, std::vector<int> target, ConnSpec &conn_spec,
              SynSpec &syn_spec) {
    return Connect(source.i0, source.n, target.data(), (int)target.size(),
                   conn_spec, syn_spec);
  }
  int Connect(std::vector<int> source, NodeSeq target, ConnSpec &conn_spec,
              SynSpec &syn_spec) {
    return Connect(source.data(), (int)source.size(), target.i0, target.n,
                   conn_spec, syn_spec);
  }
  int Connect(std::vector<int> source, std::vector<int> target,
              ConnSpec &conn_spec, SynSpec &syn_spec) {
    return Connect(source.data(), (int)source.size(), target.data(),
                   (int)target.size(), conn_spec, syn_spec);
  }
  int RemoteConnect(int i_source_host, int i_source, int n_source,
                    int i_target_host, int i_target, int n_target,
                    ConnSpec &conn_spec, SynSpec &syn_spec);
  int RemoteConnect(int i_source_host, int i_source, int n_source,
                    int i_target_host, int *target, int n_target,
                    ConnSpec &conn_spec, SynSpec &syn_spec);
  int RemoteConnect(int i_source_host, int *source, int n_source,
                    int i_target_host, int i_target, int n_target,
                    ConnSpec &conn_spec, SynSpec &syn_spec);
  int RemoteConnect(int i_source_host, int *source, int n_source,
                    int i_target_host, int *target, int n_target,
                    ConnSpec &conn_spec, SynSpec &syn_spec);
  int RemoteConnect(int i_source_host, NodeSeq source, int i_target_host,
                    NodeSeq target, ConnSpec &conn_spec, SynSpec &syn_spec) {
    return RemoteConnect(i_source_host, source.i0, source.n, i_target_host,
                         target.i0, target.n, conn_spec, syn_spec);
  }
  int RemoteConnect(int i_source_host, NodeSeq source, int i_target_host,
                    std::vector<int> target, ConnSpec &conn_spec,
                    SynSpec &syn_spec) {
    return RemoteConnect(i_source_host, source.i0, source.n, i_target_host,
                         target.data(), (int)target.size(), conn_spec,
                         syn_spec);
  }
  int RemoteConnect(int i_source_host, std::vector<int> source,
                    int i_target_host, NodeSeq target, ConnSpec &conn_spec,
                    SynSpec &syn_spec) {
    return RemoteConnect(i_source_host, source.data(), (int)source.size(),
                         i_target_host, target.i0, target.n, conn_spec,
                         syn_spec);
  }
  int RemoteConnect(int i_source_host, std::vector<int> source,
                    int i_target_host, std::vector<int> target,
                    ConnSpec &conn_spec, SynSpec &syn_spec) {
    return RemoteConnect(i_source_host, source.data(), (int)source.size(),
                         i_target_host, target.data(), (int)target.size(),
                         conn_spec, syn_spec);
  }
  int BuildDirectConnections();

  std::vector<std::string> GetScalVarNames(int i_node);
  int GetNIntVar(int i_node);
  std::vector<std::string> GetIntVarNames(int i_node);
  int GetNScalVar(int i_node);
  std::vector<std::string> GetPortVarNames(int i_node);
  int GetNPortVar(int i_node);
  std::vector<std::string> GetScalParamNames(int i_node);
  int GetNScalParam(int i_node);
  std::vector<std::string> GetPortParamNames(int i_node);
  int GetNPortParam(int i_node);
  std::vector<std::string> GetArrayParamNames(int i_node);
  int GetNArrayParam(int i_node);
  std::vector<std::string> GetArrayVarNames(int i_node);
  std::vector<std::string> GetGroupParamNames(int i_node);
  int GetNGroupParam(int i_node);
  int GetNArrayVar(int i_node);

  ConnectionStatus GetConnectionStatus(ConnectionId conn_id);
  std::vector<ConnectionStatus> GetConnectionStatus(
      std::vector<ConnectionId> &conn_id_vect);
  std::vector<ConnectionId> GetConnections(int i_source, int n_source,
                                           int i_target, int n_target,
                                           int syn_group = -1);
  std::vector<ConnectionId> GetConnections(int *i_source, int n_source,
                                           int i_target, int n_target,
                                           int syn_group = -1);
  std::vector<ConnectionId> GetConnections(int i_source, int n_source,
                                           int *i_target, int n_target,
                                           int syn_group = -1);
  std::vector<ConnectionId> GetConnections(int *i_source, int n_source,
                                           int *i_target, int n_target,
                                           int syn_group = -1);
  std::vector<ConnectionId> GetConnections(NodeSeq source, NodeSeq target,
                                           int syn_group = -1) {
    return GetConnections(source.i0, source.n, target.i0, target.n,
                          syn_group);
  }
  std::vector<ConnectionId> GetConnections(std::vector<int> source,
                                           NodeSeq target,
                                           int syn_group = -1) {
    return GetConnections(source.data(), (int)source.size(), target.i0,
                          target.n, syn_group);
  }
  std::vector<ConnectionId> GetConnections(NodeSeq source,
                                           std::vector<int> target,
                                           int syn_group = -1) {
    return GetConnections(source.i0, source.n, target.data(),
                          (int)target.size(), syn_group);
  }
  std::vector<ConnectionId> GetConnections(std::vector<int> source,
                                           std::vector<int> target,
                                           int syn_group = -1) {
    return GetConnections(source.data(), (int)source.size(), target.data(),
                          (int)target.size(), syn_group);
  }

  int CreateSynGroup(std::string model_name);
  int GetSynGroupNParam(int syn_group);
  std::vector<std::string> GetSynGroupParamNames(int syn_group);
  bool IsSynGroupParam(int syn_group, std::string param_name);
  int GetSynGroupParamIdx(int syn_group, std::string param_name);
  float GetSynGroupParam(int syn_group, std::string param_name);
  int SetSynGroupParam(int syn_group, std::string param_name, float val);
  int SynGroupCalibrate();

  int ActivateSpikeCount(int i_node, int n_node);
  int ActivateSpikeCount(NodeSeq nodes) {
    return ActivateSpikeCount(nodes.i0, nodes.n);
  }
  int ActivateRecSpikeTimes(int i_node, int n_node, int max_n_rec_spike_times);
  int ActivateRecSpikeTimes(NodeSeq nodes, int max_n_rec_spike_times) {
    return ActivateRecSpikeTimes(nodes.i0, nodes.n, max_n_rec_spike_times);
  }
  int GetNRecSpikeTimes(int i_node);
  std::vector<float> GetRecSpikeTimes(int i_node);
  int PushSpikesToNodes(int n_spikes, int *node_id, float *spike_height);
  int PushSpikesToNodes(int n_spikes, int *node_id);
  int GetExtNeuronInputSpikes(int *n_spikes, int **node, int **port,
                              float **spike_height, bool include_zeros);
  int SetNeuronGroupParam(int i_node, int n_node, std::string param_name,
                          float val);
  int IsNeuronGroupParam(int i_node, std::string param_name);
  float GetNeuronGroupParam(int i_node, std::string param_name);

  // no reference counterpart: hot-path statistics of the runs so far
  NeuronGPUStats GetStats();
  // no reference counterpart: per-kernel device time with HIP events
  // (kinds 0 AdEx RK5, 1 iaf_psc_exp, 2 detector/parrot, 3 spike
  // generator, 4 delivery, 5 post-synaptic STDP, 6 record)
  int SetKernelTiming(int on);
  int GetKernelTimes(double *ms8, long long *cnt8);
  // no reference counterpart: one value per node for a scalar parameter or
  // variable of the nodes i_node .. i_node+n-1
  int SetNeuronScalArray(int i_node, int n, std::string name,
                         const float *vals);
  // no reference counterpart: bulk connection queries (canonical order of
  // GetConnections; n_source / n_target <= 0 mean every node) and synapse
  // counts per node and delay slot
  long long GetConnectionArrays(int i_source, int n_source, int i_target,
                                int n_target, long long n_max, int *source,
                                int *target, unsigned char *port,
                                unsigned char *syn_group, float *delay,
                                float *weight);
  int GetConnectionCounts(int i_node, int n_node, int port, int *in_degree,
                          int *out_degree, long long *delay_count,
                          int n_delay);
  // no reference counterpart: learned RK5 step-size proposer
  int ProposerLog(int on, long long max_samples);
  long long ProposerNSamples();
  int ProposerGetSamples(float *out, long long n);
  int ProposerSetSamples(const float *in, long long n);
  float ProposerTrain(int n_iter, int batch, float lr, int adam,
                      int full_batch);
  int ProposerGetWeights(float *w);
  int ProposerSetWeights(const float *w);
  int ProposerUse(int on);
  int ProposerEval(const float *feat, int n, float *out);
  int ProposerEvalNoAlloc(const float *feat, int n, float *out);
};

#endif
