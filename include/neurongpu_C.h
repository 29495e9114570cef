/*
 * C ABI of the MI355X NeuronGPU engine (libneurongpu.so).
 *
 * Drop-in for the reference C ABI src/neurongpu_C.h:22-324: the same 126
 * functions with the same signatures, argument meaning, ownership and
 * error convention.  Each declaration names the reference declaration it
 * replaces.
 *
 * Error convention (reference src/propagate_error.h:4-31): every call
 * clears the error state; on failure it returns 0/NULL, and
 * NeuronGPU_GetErrorCode() reports 2 (ngpu_exception), 1 (allocation
 * failure) or 255 (other), with NeuronGPU_GetErrorMessage() holding the
 * text.  With NeuronGPU_SetOnException(0) (the default, as in the
 * reference) a failing call prints the message and exits with the code;
 * the Python bindings select 1 (return to the caller) at import.
 *
 * Ownership: arrays returned by the library (parameter/variable values,
 * name lists, connection ids, recorded spike times, record rows) are
 * allocated by the library and are not freed by it.
 *
 * The single-process build reports MPI as unavailable, like a reference
 * build configured without MPI.
 */
#ifndef NGPU_AMD_NEURONGPU_C_H
#define NGPU_AMD_NEURONGPU_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

  // errors and kernel settings
  char *NeuronGPU_GetErrorMessage();  // src/neurongpu_C.h:22
  unsigned char NeuronGPU_GetErrorCode();  // src/neurongpu_C.h:24
  void NeuronGPU_SetOnException(int on_exception);  // src/neurongpu_C.h:26
  int NeuronGPU_SetRandomSeed(unsigned long long seed);  // src/neurongpu_C.h:28
  int NeuronGPU_SetTimeResolution(float time_res);  // src/neurongpu_C.h:30
  float NeuronGPU_GetTimeResolution();  // src/neurongpu_C.h:32
  int NeuronGPU_SetMaxSpikeBufferSize(int max_size);  // src/neurongpu_C.h:34
  int NeuronGPU_GetMaxSpikeBufferSize();  // src/neurongpu_C.h:36
  int NeuronGPU_SetSimTime(float sim_time);  // src/neurongpu_C.h:38
  int NeuronGPU_SetVerbosityLevel(int verbosity_level);  // src/neurongpu_C.h:40

  // node creation and records
  int NeuronGPU_Create(char *model_name, int n_neuron, int n_port);  // src/neurongpu_C.h:42
  int NeuronGPU_CreatePoissonGenerator(int n_node, float rate);  // src/neurongpu_C.h:44
  int NeuronGPU_CreateRecord(char *file_name, char *var_name_arr[], int *i_node_arr, int *port_arr, int n_node);  // src/neurongpu_C.h:46
  int NeuronGPU_GetRecordDataRows(int i_record);  // src/neurongpu_C.h:50
  int NeuronGPU_GetRecordDataColumns(int i_record);  // src/neurongpu_C.h:52
  float **NeuronGPU_GetRecordData(int i_record);  // src/neurongpu_C.h:54

  // node parameters and variables
  int NeuronGPU_SetNeuronScalParam(int i_node, int n_neuron, char *param_name, float val);  // src/neurongpu_C.h:56
  int NeuronGPU_SetNeuronArrayParam(int i_node, int n_neuron, char *param_name, float *param, int array_size);  // src/neurongpu_C.h:59
  int NeuronGPU_SetNeuronPtScalParam(int *i_node, int n_neuron, char *param_name, float val);  // src/neurongpu_C.h:63
  int NeuronGPU_SetNeuronPtArrayParam(int *i_node, int n_neuron, char *param_name, float *param, int array_size);  // src/neurongpu_C.h:66
  int NeuronGPU_IsNeuronScalParam(int i_node, char *param_name);  // src/neurongpu_C.h:70
  int NeuronGPU_IsNeuronPortParam(int i_node, char *param_name);  // src/neurongpu_C.h:72
  int NeuronGPU_IsNeuronArrayParam(int i_node, char *param_name);  // src/neurongpu_C.h:74
  int NeuronGPU_SetNeuronIntVar(int i_node, int n_neuron, char *var_name, int val);  // src/neurongpu_C.h:77
  int NeuronGPU_SetNeuronScalVar(int i_node, int n_neuron, char *var_name, float val);  // src/neurongpu_C.h:80
  int NeuronGPU_SetNeuronArrayVar(int i_node, int n_neuron, char *var_name, float *var, int array_size);  // src/neurongpu_C.h:83
  int NeuronGPU_SetNeuronPtIntVar(int *i_node, int n_neuron, char *var_name, int val);  // src/neurongpu_C.h:87
  int NeuronGPU_SetNeuronPtScalVar(int *i_node, int n_neuron, char *var_name, float val);  // src/neurongpu_C.h:90
  int NeuronGPU_SetNeuronPtArrayVar(int *i_node, int n_neuron, char *var_name, float *var, int array_size);  // src/neurongpu_C.h:93
  int NeuronGPU_IsNeuronIntVar(int i_node, char *var_name);  // src/neurongpu_C.h:97
  int NeuronGPU_IsNeuronScalVar(int i_node, char *var_name);  // src/neurongpu_C.h:99
  int NeuronGPU_IsNeuronPortVar(int i_node, char *var_name);  // src/neurongpu_C.h:101
  int NeuronGPU_IsNeuronArrayVar(int i_node, char *var_name);  // src/neurongpu_C.h:103
  int NeuronGPU_GetNeuronParamSize(int i_node, char *param_name);  // src/neurongpu_C.h:105
  int NeuronGPU_GetNeuronVarSize(int i_node, char *var_name);  // src/neurongpu_C.h:107
  float *NeuronGPU_GetNeuronParam(int i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:109
  float *NeuronGPU_GetNeuronPtParam(int *i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:112
  float *NeuronGPU_GetArrayParam(int i_node, char *param_name);  // src/neurongpu_C.h:115
  int *NeuronGPU_GetNeuronIntVar(int i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:117
  int *NeuronGPU_GetNeuronPtIntVar(int *i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:120
  float *NeuronGPU_GetNeuronVar(int i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:123
  float *NeuronGPU_GetNeuronPtVar(int *i_node, int n_neuron, char *param_name);  // src/neurongpu_C.h:126
  float *NeuronGPU_GetArrayVar(int i_node, char *var_name);  // src/neurongpu_C.h:129

  // the per-timestep path
  int NeuronGPU_Calibrate();  // src/neurongpu_C.h:131
  int NeuronGPU_Simulate();  // src/neurongpu_C.h:133
  int NeuronGPU_StartSimulation();  // src/neurongpu_C.h:135
  int NeuronGPU_SimulationStep();  // src/neurongpu_C.h:137
  int NeuronGPU_EndSimulation();  // src/neurongpu_C.h:139

  // MPI (single GPU: report that MPI is not available, as a non-MPI build does)
  int NeuronGPU_ConnectMpiInit(int argc, char *argv[]);  // src/neurongpu_C.h:141
  int NeuronGPU_MpiId();  // src/neurongpu_C.h:143
  int NeuronGPU_MpiNp();  // src/neurongpu_C.h:145
  int NeuronGPU_ProcMaster();  // src/neurongpu_C.h:147
  int NeuronGPU_MpiFinalize();  // src/neurongpu_C.h:149

  // random numbers
  unsigned int *NeuronGPU_RandomInt(size_t n);  // src/neurongpu_C.h:151
  float *NeuronGPU_RandomUniform(size_t n);  // src/neurongpu_C.h:153
  float *NeuronGPU_RandomNormal(size_t n, float mean, float stddev);  // src/neurongpu_C.h:155
  float *NeuronGPU_RandomNormalClipped(size_t n, float mean, float stddev, float vmin, float vmax, float vstep);  // src/neurongpu_C.h:157

  // connections
  int NeuronGPU_Connect(int i_source_node, int i_target_node, unsigned char port, unsigned char syn_group, float weight, float delay);  // src/neurongpu_C.h:160
  int NeuronGPU_ConnSpecInit();  // src/neurongpu_C.h:164
  int NeuronGPU_SetConnSpecParam(char *param_name, int value);  // src/neurongpu_C.h:166
  int NeuronGPU_ConnSpecIsParam(char *param_name);  // src/neurongpu_C.h:168
  int NeuronGPU_SynSpecInit();  // src/neurongpu_C.h:170
  int NeuronGPU_SetSynSpecIntParam(char *param_name, int value);  // src/neurongpu_C.h:172
  int NeuronGPU_SetSynSpecFloatParam(char *param_name, float value);  // src/neurongpu_C.h:174
  int NeuronGPU_SetSynSpecFloatPtParam(char *param_name, float *array_pt);  // src/neurongpu_C.h:176
  int NeuronGPU_SynSpecIsIntParam(char *param_name);  // src/neurongpu_C.h:178
  int NeuronGPU_SynSpecIsFloatParam(char *param_name);  // src/neurongpu_C.h:180
  int NeuronGPU_SynSpecIsFloatPtParam(char *param_name);  // src/neurongpu_C.h:182
  int NeuronGPU_ConnectSeqSeq(int i_source, int n_source, int i_target, int n_target);  // src/neurongpu_C.h:184
  int NeuronGPU_ConnectSeqGroup(int i_source, int n_source, int *i_target, int n_target);  // src/neurongpu_C.h:187
  int NeuronGPU_ConnectGroupSeq(int *i_source, int n_source, int i_target, int n_target);  // src/neurongpu_C.h:190
  int NeuronGPU_ConnectGroupGroup(int *i_source, int n_source, int *i_target, int n_target);  // src/neurongpu_C.h:193
  int NeuronGPU_RemoteConnectSeqSeq(int i_source_host, int i_source, int n_source, int i_target_host, int i_target, int n_target);  // src/neurongpu_C.h:196
  int NeuronGPU_RemoteConnectSeqGroup(int i_source_host, int i_source, int n_source, int i_target_host, int *i_target, int n_target);  // src/neurongpu_C.h:200
  int NeuronGPU_RemoteConnectGroupSeq(int i_source_host, int *i_source, int n_source, int i_target_host, int i_target, int n_target);  // src/neurongpu_C.h:204
  int NeuronGPU_RemoteConnectGroupGroup(int i_source_host, int *i_source, int n_source, int i_target_host, int *i_target, int n_target);  // src/neurongpu_C.h:208

  // model introspection
  char **NeuronGPU_GetIntVarNames(int i_node);  // src/neurongpu_C.h:212
  char **NeuronGPU_GetScalVarNames(int i_node);  // src/neurongpu_C.h:214
  int NeuronGPU_GetNIntVar(int i_node);  // src/neurongpu_C.h:216
  int NeuronGPU_GetNScalVar(int i_node);  // src/neurongpu_C.h:218
  char **NeuronGPU_GetPortVarNames(int i_node);  // src/neurongpu_C.h:220
  int NeuronGPU_GetNPortVar(int i_node);  // src/neurongpu_C.h:222
  char **NeuronGPU_GetScalParamNames(int i_node);  // src/neurongpu_C.h:224
  int NeuronGPU_GetNScalParam(int i_node);  // src/neurongpu_C.h:226
  char **NeuronGPU_GetPortParamNames(int i_node);  // src/neurongpu_C.h:228
  int NeuronGPU_GetNGroupParam(int i_node);  // src/neurongpu_C.h:230
  char **NeuronGPU_GetGroupParamNames(int i_node);  // src/neurongpu_C.h:232
  int NeuronGPU_GetNPortParam(int i_node);  // src/neurongpu_C.h:234
  char **NeuronGPU_GetArrayParamNames(int i_node);  // src/neurongpu_C.h:236
  int NeuronGPU_GetNArrayParam(int i_node);  // src/neurongpu_C.h:238
  char **NeuronGPU_GetArrayVarNames(int i_node);  // src/neurongpu_C.h:240
  int NeuronGPU_GetNArrayVar(int i_node);  // src/neurongpu_C.h:242

  // connection queries
  int *NeuronGPU_GetSeqSeqConnections(int i_source, int n_source, int i_target, int n_target, int syn_group, int *n_conn);  // src/neurongpu_C.h:244
  int *NeuronGPU_GetSeqGroupConnections(int i_source, int n_source, int *i_target, int n_target, int syn_group, int *n_conn);  // src/neurongpu_C.h:247
  int *NeuronGPU_GetGroupSeqConnections(int *i_source, int n_source, int i_target, int n_target, int syn_group, int *n_conn);  // src/neurongpu_C.h:251
  int *NeuronGPU_GetGroupGroupConnections(int *i_source, int n_source, int *i_target, int n_target, int syn_group, int *n_conn);  // src/neurongpu_C.h:255
  int NeuronGPU_GetConnectionStatus(int i_source, int i_group, int i_conn, int *i_target, unsigned char *port, unsigned char *syn_group, float *delay, float *weight);  // src/neurongpu_C.h:258

  // synapse groups
  int NeuronGPU_CreateSynGroup(char *model_name);  // src/neurongpu_C.h:263
  int NeuronGPU_GetSynGroupNParam(int i_syn_group);  // src/neurongpu_C.h:265
  char **NeuronGPU_GetSynGroupParamNames(int i_syn_group);  // src/neurongpu_C.h:267
  int NeuronGPU_IsSynGroupParam(int i_syn_group, char *param_name);  // src/neurongpu_C.h:269
  int NeuronGPU_GetSynGroupParamIdx(int i_syn_group, char *param_name);  // src/neurongpu_C.h:271
  float NeuronGPU_GetSynGroupParam(int i_syn_group, char *param_name);  // src/neurongpu_C.h:273
  int NeuronGPU_SetSynGroupParam(int i_syn_group, char *param_name, float val);  // src/neurongpu_C.h:275

  // spike recording and injection
  int NeuronGPU_ActivateSpikeCount(int i_node, int n_node);  // src/neurongpu_C.h:277
  int NeuronGPU_ActivateRecSpikeTimes(int i_node, int n_node, int max_n_rec_spike_times);  // src/neurongpu_C.h:279
  int NeuronGPU_GetNRecSpikeTimes(int i_node);  // src/neurongpu_C.h:282
  float* NeuronGPU_GetRecSpikeTimes(int i_node);  // src/neurongpu_C.h:284
  int NeuronGPU_PushSpikesToNodes(int n_spikes, int *node_id);  // src/neurongpu_C.h:286
  int NeuronGPU_GetExtNeuronInputSpikes(int *n_spikes, int **node, int **port, float **spike_height, int include_zeros);  // src/neurongpu_C.h:288

  // group and kernel parameters
  int NeuronGPU_SetNeuronGroupParam(int i_node, int n_node, char *param_name, float val);  // src/neurongpu_C.h:292
  int NeuronGPU_IsNeuronGroupParam(int i_node, char *param_name);  // src/neurongpu_C.h:295
  float NeuronGPU_GetNeuronGroupParam(int i_node, char *param_name);  // src/neurongpu_C.h:297
  int NeuronGPU_GetNFloatParam();  // src/neurongpu_C.h:299
  char **NeuronGPU_GetFloatParamNames();  // src/neurongpu_C.h:301
  int NeuronGPU_IsFloatParam(char *param_name);  // src/neurongpu_C.h:303
  int NeuronGPU_GetFloatParamIdx(char *param_name);  // src/neurongpu_C.h:305
  float NeuronGPU_GetFloatParam(char *param_name);  // src/neurongpu_C.h:307
  int NeuronGPU_SetFloatParam(char *param_name, float val);  // src/neurongpu_C.h:309
  int NeuronGPU_GetNIntParam();  // src/neurongpu_C.h:311
  char **NeuronGPU_GetIntParamNames();  // src/neurongpu_C.h:313
  int NeuronGPU_IsIntParam(char *param_name);  // src/neurongpu_C.h:315
  int NeuronGPU_GetIntParamIdx(char *param_name);  // src/neurongpu_C.h:317
  int NeuronGPU_GetIntParam(char *param_name);  // src/neurongpu_C.h:319
  int NeuronGPU_SetIntParam(char *param_name, int val);  // src/neurongpu_C.h:321
  int NeuronGPU_RemoteCreate(int i_host, char *model_name, int n_neuron, int n_port);  // src/neurongpu_C.h:323

  /* ---- no counterpart in the reference ---- */

  /* Discard the current network and all state; the next call starts from a
   * freshly constructed engine (tests and repeated runs in one process). */
  int NeuronGPU_ResetKernel();

  /* Hot-path statistics since calibration: out[0] RK5 accepted steps,
   * out[1] RK5 rejected attempts, out[2] synaptic events delivered. */
  int NeuronGPU_GetHotPathStats(unsigned long long *out3);

  /* The first pass's share of those counts (two-pass AdEx update, at most
   * rk5_budget attempts per neuron-step): out[0] accepted, out[1]
   * rejected. */
  int NeuronGPU_GetRk5Pass1Stats(unsigned long long *out2);

  /* RK5 attempts accepted with a scaled error above 1 since calibration:
   * the acceptance test lets them pass only at the minimum step size
   * (src/rk5.h:125), whatever step size the controller or the learned
   * proposer chose. */
  int NeuronGPU_GetRk5ForcedAccepts(unsigned long long *out1);

  /* Per-kernel device time measured with HIP events on the engine's stream
   * (on != 0 enables; enabling or disabling clears the totals).  Kinds:
   * 0 AdEx RK5 update (first pass), 1 iaf_psc_exp update, 2 spike
   * detector / parrot, 3 spike generator, 4 delivery, 5 post-synaptic STDP,
   * 6 record, 7 second RK5 pass (with the delivery when they share a
   * launch). */
  int NeuronGPU_SetKernelTiming(int on);
  int NeuronGPU_GetKernelTimes(double *ms8, long long *cnt8);

  /* One value per node (vals[0 .. n_node-1]) for the scalar parameter or
   * variable `name` of nodes i_node .. i_node + n_node - 1 (one node
   * group), where the reference's Python makes one call per node. */
  int NeuronGPU_SetNeuronScalArray(int i_node, int n_node, char *name,
                                   float *vals);

  /* Bulk form of GetConnections + GetConnectionStatus: the connections
   * with source in [i_source, i_source + n_source) and target in
   * [i_target, i_target + n_target) (n <= 0: every node), in the canonical
   * order of GetConnections.  Returns their number; the first n_max are
   * written into the caller's arrays (any may be NULL). */
  long long NeuronGPU_GetConnectionArrays(int i_source, int n_source,
                                          int i_target, int n_target,
                                          long long n_max, int *source,
                                          int *target, unsigned char *port,
                                          unsigned char *syn_group,
                                          float *delay, float *weight);

  /* Synapse counts for nodes i_node .. i_node + n_node - 1, counted on the
   * device after calibration: in_degree[i] / out_degree[i] = synapses with
   * node i_node + i as target / source, delay_count[d] = synapses with a
   * source in the range and a delay of d + 1 time steps (d < n_delay).
   * Only synapses on receptor `port` count (every port if port < 0).
   * Connections from poisson_generator nodes count like any other, as in
   * GetConnections.  Any output may be NULL. */
  int NeuronGPU_GetConnectionCounts(int i_node, int n_node, int port,
                                    int *in_degree, int *out_degree,
                                    long long *delay_count, int n_delay);

  /* Learned RK5 step-size proposer (6-32-32-1 MLP, BF16 inference, FP32
   * training), used for the RK5 steps of the second pass of the AdEx update
   * when enabled; the RK5 error test still accepts or rejects every
   * attempt.  Samples are 8 floats: 6 features, target, unused. */
  /* on != 0: log one sample per second-pass RK5 attempt, up to max_samples
   * (the buffer is kept until the next call with on != 0) */
  int NeuronGPU_ProposerLog(int on, long long max_samples);
  long long NeuronGPU_ProposerNSamples();
  int NeuronGPU_ProposerGetSamples(float *out, long long n);
  int NeuronGPU_ProposerSetSamples(float *in, long long n);
  /* n_iter minibatch steps of `batch` samples (full_batch != 0: samples
   * 0..batch-1 every step), SGD (adam == 0) or Adam with step size lr;
   * *loss receives the mean squared error of the last minibatch */
  int NeuronGPU_ProposerTrain(int n_iter, int batch, float lr, int adam,
                              int full_batch, float *loss);
  int NeuronGPU_ProposerGetWeights(float *w1313);
  int NeuronGPU_ProposerSetWeights(float *w1313);
  /* on != 0: the second RK5 pass starts its steps from the proposer's step
   * size instead of the plain controller's */
  int NeuronGPU_ProposerUse(int on);
  /* the proposer's output (log2 of the step in units of dt, before the
   * safety factor) for n feature vectors feat[n][6] */
  int NeuronGPU_ProposerEval(float *feat, int n, float *out);

  /* No counterpart in the reference.  NeuronGPU_ProposerEval for callers
   * that cannot allocate device memory (code under HIP graph capture, tight
   * serving loops, harnesses that forbid hipMalloc): the same kernel and
   * results, with the feature vectors and outputs staged through a
   * file-scope device array instead of buffers allocated per call.  At most
   * 65536 feature vectors per call (larger n is an error); the weights must
   * exist already (ProposerSetWeights, ProposerTrain or ProposerEval).
   * Calls that overlap on one GPU share the staging array and race. */
  int NeuronGPU_ProposerEvalNoAlloc(float *feat, int n, float *out);

#ifdef __cplusplus
}
#endif

#endif
