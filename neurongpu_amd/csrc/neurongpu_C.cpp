// C ABI wrappers (include/neurongpu_C.h) around the NeuronGPU class.
// Reference counterpart: src/neurongpu_C.cpp, error propagation
// src/propagate_error.h:4-31.  One library-owned engine instance, created
// on first use; ConnSpec/SynSpec instances hold the state of the
// Connect* calls between ConnSpecInit/SynSpecInit and the Connect call.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <string>
#include <vector>

#include "neurongpu.h"
#include "neurongpu_C.h"

namespace {

NeuronGPU *instance = nullptr;
ConnSpec conn_spec_instance;
SynSpec syn_spec_instance;

void CheckInstance() {
  if (!instance) instance = new NeuronGPU();
}

// Errors: code 2 for ngpu_exception, 1 for allocation failure, 255 other.
// With ON_EXCEPTION_EXIT a failing call prints the message and exits.
template <class F>
void Guard(F &&f) {
  CheckInstance();
  instance->SetErrorFlag(false);
  instance->SetErrorMessage("");
  instance->SetErrorCode(0);
  try {
    f();
  } catch (ngpu_exception &e) {
    instance->SetErrorFlag(true);
    instance->SetErrorMessage(e.what());
    instance->SetErrorCode(2);
  } catch (std::bad_alloc &) {
    instance->SetErrorFlag(true);
    instance->SetErrorMessage("Memory allocation error.");
    instance->SetErrorCode(1);
  } catch (...) {
    instance->SetErrorFlag(true);
    instance->SetErrorMessage("Error in NeuronGPU function.");
    instance->SetErrorCode(255);
  }
  if (instance->GetErrorFlag() && instance->OnException() == ON_EXCEPTION_EXIT) {
    std::cerr << instance->GetErrorMessage();
    exit(instance->GetErrorCode());
  }
}

char **NameArray(const std::vector<std::string> &v) {
  char **a = (char **)malloc((v.size() + 1) * sizeof(char *));
  for (size_t i = 0; i < v.size(); i++) {
    a[i] = (char *)malloc(v[i].size() + 1);
    std::strcpy(a[i], v[i].c_str());
  }
  a[v.size()] = nullptr;
  return a;
}

int *ConnIdArray(const std::vector<ConnectionId> &v, int *n_conn) {
  *n_conn = (int)v.size();
  int *a = (int *)malloc((v.size() * 3 + 1) * sizeof(int));
  for (size_t i = 0; i < v.size(); i++) {
    a[3 * i] = v[i].i_source_;
    a[3 * i + 1] = v[i].i_group_;
    a[3 * i + 2] = v[i].i_conn_;
  }
  return a;
}

}  // namespace

#define RET(type, init, body) \
  {                           \
    type ret = init;          \
    Guard([&] { body; });     \
    return ret;               \
  }

extern "C" {

char *NeuronGPU_GetErrorMessage() {
  CheckInstance();
  return instance->GetErrorMessage();
}
unsigned char NeuronGPU_GetErrorCode() {
  CheckInstance();
  return instance->GetErrorCode();
}
void NeuronGPU_SetOnException(int on_exception) {
  CheckInstance();
  instance->SetOnException(on_exception);
}
int NeuronGPU_SetRandomSeed(unsigned long long seed)
    RET(int, 0, ret = instance->SetRandomSeed(seed))
int NeuronGPU_SetTimeResolution(float time_res)
    RET(int, 0, ret = instance->SetTimeResolution(time_res))
float NeuronGPU_GetTimeResolution()
    RET(float, 0.0f, ret = instance->GetTimeResolution())
int NeuronGPU_SetMaxSpikeBufferSize(int max_size)
    RET(int, 0, ret = instance->SetMaxSpikeBufferSize(max_size))
int NeuronGPU_GetMaxSpikeBufferSize()
    RET(int, 0, ret = instance->GetMaxSpikeBufferSize())
int NeuronGPU_SetSimTime(float sim_time)
    RET(int, 0, ret = instance->SetSimTime(sim_time))
int NeuronGPU_SetVerbosityLevel(int v)
    RET(int, 0, ret = instance->SetVerbosityLevel(v))

int NeuronGPU_Create(char *model_name, int n_neuron, int n_port)
    RET(int, 0, ret = instance->Create(model_name, n_neuron, n_port)[0])
int NeuronGPU_CreatePoissonGenerator(int n_node, float rate)
    RET(int, 0, ret = instance->CreatePoissonGenerator(n_node, rate)[0])
int NeuronGPU_CreateRecord(char *file_name, char *var_name_arr[],
                           int *i_node_arr, int *port_arr, int n_node)
    RET(int, 0, {
      std::vector<std::string> names;
      for (int i = 0; i < n_node; i++) names.push_back(var_name_arr[i]);
      ret = instance->CreateRecord(file_name, names.data(), i_node_arr,
                                   port_arr, n_node);
    })
int NeuronGPU_GetRecordDataRows(int i_record)
    RET(int, 0, ret = (int)instance->GetRecordData(i_record)->size())
int NeuronGPU_GetRecordDataColumns(int i_record)
    RET(int, 0, {
      auto *d = instance->GetRecordData(i_record);
      if (d->empty()) throw ngpu_exception("Record has no data");
      ret = (int)d->at(0).size();
    })
float **NeuronGPU_GetRecordData(int i_record)
    RET(float **, nullptr, {
      auto *d = instance->GetRecordData(i_record);
      ret = new float *[d->size() + 1];
      for (size_t i = 0; i < d->size(); i++) ret[i] = (*d)[i].data();
    })

int NeuronGPU_SetNeuronScalParam(int i_node, int n, char *name, float val)
    RET(int, 0, ret = instance->SetNeuronParam(i_node, n, name, val))
int NeuronGPU_SetNeuronArrayParam(int i_node, int n, char *name, float *p,
                                  int size)
    RET(int, 0, ret = instance->SetNeuronParam(i_node, n, name, p, size))
int NeuronGPU_SetNeuronPtScalParam(int *i_node, int n, char *name, float val)
    RET(int, 0, ret = instance->SetNeuronParam(i_node, n, name, val))
int NeuronGPU_SetNeuronPtArrayParam(int *i_node, int n, char *name, float *p,
                                    int size)
    RET(int, 0, ret = instance->SetNeuronParam(i_node, n, name, p, size))
int NeuronGPU_IsNeuronScalParam(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronScalParam(i_node, name))
int NeuronGPU_IsNeuronPortParam(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronPortParam(i_node, name))
int NeuronGPU_IsNeuronArrayParam(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronArrayParam(i_node, name))
int NeuronGPU_SetNeuronIntVar(int i_node, int n, char *name, int val)
    RET(int, 0, ret = instance->SetNeuronIntVar(i_node, n, name, val))
int NeuronGPU_SetNeuronScalVar(int i_node, int n, char *name, float val)
    RET(int, 0, ret = instance->SetNeuronVar(i_node, n, name, val))
int NeuronGPU_SetNeuronArrayVar(int i_node, int n, char *name, float *v,
                                int size)
    RET(int, 0, ret = instance->SetNeuronVar(i_node, n, name, v, size))
int NeuronGPU_SetNeuronPtIntVar(int *i_node, int n, char *name, int val)
    RET(int, 0, ret = instance->SetNeuronIntVar(i_node, n, name, val))
int NeuronGPU_SetNeuronPtScalVar(int *i_node, int n, char *name, float val)
    RET(int, 0, ret = instance->SetNeuronVar(i_node, n, name, val))
int NeuronGPU_SetNeuronPtArrayVar(int *i_node, int n, char *name, float *v,
                                  int size)
    RET(int, 0, ret = instance->SetNeuronVar(i_node, n, name, v, size))
int NeuronGPU_IsNeuronIntVar(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronIntVar(i_node, name))
int NeuronGPU_IsNeuronScalVar(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronScalVar(i_node, name))
int NeuronGPU_IsNeuronPortVar(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronPortVar(i_node, name))
int NeuronGPU_IsNeuronArrayVar(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronArrayVar(i_node, name))
int NeuronGPU_GetNeuronParamSize(int i_node, char *name)
    RET(int, 0, ret = instance->GetNeuronParamSize(i_node, name))
int NeuronGPU_GetNeuronVarSize(int i_node, char *name)
    RET(int, 0, ret = instance->GetNeuronVarSize(i_node, name))
float *NeuronGPU_GetNeuronParam(int i_node, int n, char *name)
    RET(float *, nullptr, ret = instance->GetNeuronParam(i_node, n, name))
float *NeuronGPU_GetNeuronPtParam(int *i_node, int n, char *name)
    RET(float *, nullptr, ret = instance->GetNeuronParam(i_node, n, name))
float *NeuronGPU_GetArrayParam(int i_node, char *name)
    RET(float *, nullptr, ret = instance->GetArrayParam(i_node, name))
int *NeuronGPU_GetNeuronIntVar(int i_node, int n, char *name)
    RET(int *, nullptr, ret = instance->GetNeuronIntVar(i_node, n, name))
int *NeuronGPU_GetNeuronPtIntVar(int *i_node, int n, char *name)
    RET(int *, nullptr, ret = instance->GetNeuronIntVar(i_node, n, name))
float *NeuronGPU_GetNeuronVar(int i_node, int n, char *name)
    RET(float *, nullptr, ret = instance->GetNeuronVar(i_node, n, name))
float *NeuronGPU_GetNeuronPtVar(int *i_node, int n, char *name)
    RET(float *, nullptr, ret = instance->GetNeuronVar(i_node, n, name))
float *NeuronGPU_GetArrayVar(int i_node, char *name)
    RET(float *, nullptr, ret = instance->GetArrayVar(i_node, name))

int NeuronGPU_Calibrate() RET(int, 0, ret = instance->Calibrate())
int NeuronGPU_Simulate() RET(int, 0, ret = instance->Simulate())
int NeuronGPU_StartSimulation() RET(int, 0, ret = instance->StartSimulation())
int NeuronGPU_SimulationStep() RET(int, 0, ret = instance->SimulationStep())
int NeuronGPU_EndSimulation() RET(int, 0, ret = instance->EndSimulation())

int NeuronGPU_ConnectMpiInit(int argc, char *argv[])
    RET(int, 0, ret = instance->ConnectMpiInit(argc, argv))
int NeuronGPU_MpiId() RET(int, 0, ret = instance->MpiId())
int NeuronGPU_MpiNp() RET(int, 0, ret = instance->MpiNp())
int NeuronGPU_ProcMaster() RET(int, 0, ret = instance->ProcMaster())
int NeuronGPU_MpiFinalize() RET(int, 0, ret = instance->MpiFinalize())

unsigned int *NeuronGPU_RandomInt(size_t n)
    RET(unsigned int *, nullptr, ret = instance->RandomInt(n))
float *NeuronGPU_RandomUniform(size_t n)
    RET(float *, nullptr, ret = instance->RandomUniform(n))
float *NeuronGPU_RandomNormal(size_t n, float mean, float stddev)
    RET(float *, nullptr, ret = instance->RandomNormal(n, mean, stddev))
float *NeuronGPU_RandomNormalClipped(size_t n, float mean, float stddev,
                                     float vmin, float vmax, float vstep)
    RET(float *, nullptr,
        ret = instance->RandomNormalClipped(n, mean, stddev, vmin, vmax, vstep))

int NeuronGPU_Connect(int i_source, int i_target, unsigned char port,
                      unsigned char syn_group, float weight, float delay)
    RET(int, 0,
        ret = instance->Connect(i_source, i_target, port, syn_group, weight, delay))
int NeuronGPU_ConnSpecInit() RET(int, 0, ret = conn_spec_instance.Init())
int NeuronGPU_SetConnSpecParam(char *name, int value)
    RET(int, 0, ret = conn_spec_instance.SetParam(name, value))
int NeuronGPU_ConnSpecIsParam(char *name)
    RET(int, 0, ret = ConnSpec::IsParam(name))
int NeuronGPU_SynSpecInit() RET(int, 0, ret = syn_spec_instance.Init())
int NeuronGPU_SetSynSpecIntParam(char *name, int value)
    RET(int, 0, ret = syn_spec_instance.SetParam(name, value))
int NeuronGPU_SetSynSpecFloatParam(char *name, float value)
    RET(int, 0, ret = syn_spec_instance.SetParam(name, value))
int NeuronGPU_SetSynSpecFloatPtParam(char *name, float *array_pt)
    RET(int, 0, ret = syn_spec_instance.SetParam(name, array_pt))
int NeuronGPU_SynSpecIsIntParam(char *name)
    RET(int, 0, ret = SynSpec::IsIntParam(name))
int NeuronGPU_SynSpecIsFloatParam(char *name)
    RET(int, 0, ret = SynSpec::IsFloatParam(name))
int NeuronGPU_SynSpecIsFloatPtParam(char *name)
    RET(int, 0, ret = SynSpec::IsFloatPtParam(name))
int NeuronGPU_ConnectSeqSeq(int i_source, int n_source, int i_target,
                            int n_target)
    RET(int, 0,
        ret = instance->Connect(i_source, n_source, i_target, n_target,
                                conn_spec_instance, syn_spec_instance))
int NeuronGPU_ConnectSeqGroup(int i_source, int n_source, int *i_target,
                              int n_target)
    RET(int, 0,
        ret = instance->Connect(i_source, n_source, i_target, n_target,
                                conn_spec_instance, syn_spec_instance))
int NeuronGPU_ConnectGroupSeq(int *i_source, int n_source, int i_target,
                              int n_target)
    RET(int, 0,
        ret = instance->Connect(i_source, n_source, i_target, n_target,
                                conn_spec_instance, syn_spec_instance))
int NeuronGPU_ConnectGroupGroup(int *i_source, int n_source, int *i_target,
                                int n_target)
    RET(int, 0,
        ret = instance->Connect(i_source, n_source, i_target, n_target,
                                conn_spec_instance, syn_spec_instance))
int NeuronGPU_RemoteConnectSeqSeq(int hs, int i_source, int n_source, int ht,
                                  int i_target, int n_target)
    RET(int, 0,
        ret = instance->RemoteConnect(hs, i_source, n_source, ht, i_target,
                                      n_target, conn_spec_instance,
                                      syn_spec_instance))
int NeuronGPU_RemoteConnectSeqGroup(int hs, int i_source, int n_source, int ht,
                                    int *i_target, int n_target)
    RET(int, 0,
        ret = instance->RemoteConnect(hs, i_source, n_source, ht, i_target,
                                      n_target, conn_spec_instance,
                                      syn_spec_instance))
int NeuronGPU_RemoteConnectGroupSeq(int hs, int *i_source, int n_source, int ht,
                                    int i_target, int n_target)
    RET(int, 0,
        ret = instance->RemoteConnect(hs, i_source, n_source, ht, i_target,
                                      n_target, conn_spec_instance,
                                      syn_spec_instance))
int NeuronGPU_RemoteConnectGroupGroup(int hs, int *i_source, int n_source,
                                      int ht, int *i_target, int n_target)
    RET(int, 0,
        ret = instance->RemoteConnect(hs, i_source, n_source, ht, i_target,
                                      n_target, conn_spec_instance,
                                      syn_spec_instance))

char **NeuronGPU_GetIntVarNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetIntVarNames(i_node)))
char **NeuronGPU_GetScalVarNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetScalVarNames(i_node)))
int NeuronGPU_GetNIntVar(int i_node) RET(int, 0, ret = instance->GetNIntVar(i_node))
int NeuronGPU_GetNScalVar(int i_node) RET(int, 0, ret = instance->GetNScalVar(i_node))
char **NeuronGPU_GetPortVarNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetPortVarNames(i_node)))
int NeuronGPU_GetNPortVar(int i_node) RET(int, 0, ret = instance->GetNPortVar(i_node))
char **NeuronGPU_GetScalParamNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetScalParamNames(i_node)))
int NeuronGPU_GetNScalParam(int i_node)
    RET(int, 0, ret = instance->GetNScalParam(i_node))
char **NeuronGPU_GetPortParamNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetPortParamNames(i_node)))
int NeuronGPU_GetNGroupParam(int i_node)
    RET(int, 0, ret = instance->GetNGroupParam(i_node))
char **NeuronGPU_GetGroupParamNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetGroupParamNames(i_node)))
int NeuronGPU_GetNPortParam(int i_node)
    RET(int, 0, ret = instance->GetNPortParam(i_node))
char **NeuronGPU_GetArrayParamNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetArrayParamNames(i_node)))
int NeuronGPU_GetNArrayParam(int i_node)
    RET(int, 0, ret = instance->GetNArrayParam(i_node))
char **NeuronGPU_GetArrayVarNames(int i_node)
    RET(char **, nullptr, ret = NameArray(instance->GetArrayVarNames(i_node)))
int NeuronGPU_GetNArrayVar(int i_node)
    RET(int, 0, ret = instance->GetNArrayVar(i_node))

int *NeuronGPU_GetSeqSeqConnections(int i_source, int n_source, int i_target,
                                    int n_target, int syn_group, int *n_conn)
    RET(int *, nullptr,
        ret = ConnIdArray(instance->GetConnections(i_source, n_source, i_target,
                                                   n_target, syn_group),
                          n_conn))
int *NeuronGPU_GetSeqGroupConnections(int i_source, int n_source,
                                      int *i_target, int n_target,
                                      int syn_group, int *n_conn)
    RET(int *, nullptr,
        ret = ConnIdArray(instance->GetConnections(i_source, n_source, i_target,
                                                   n_target, syn_group),
                          n_conn))
int *NeuronGPU_GetGroupSeqConnections(int *i_source, int n_source,
                                      int i_target, int n_target,
                                      int syn_group, int *n_conn)
    RET(int *, nullptr,
        ret = ConnIdArray(instance->GetConnections(i_source, n_source, i_target,
                                                   n_target, syn_group),
                          n_conn))
int *NeuronGPU_GetGroupGroupConnections(int *i_source, int n_source,
                                        int *i_target, int n_target,
                                        int syn_group, int *n_conn)
    RET(int *, nullptr,
        ret = ConnIdArray(instance->GetConnections(i_source, n_source, i_target,
                                                   n_target, syn_group),
                          n_conn))
int NeuronGPU_GetConnectionStatus(int i_source, int i_group, int i_conn,
                                  int *i_target, unsigned char *port,
                                  unsigned char *syn_group, float *delay,
                                  float *weight)
    RET(int, 0, {
      ConnectionId cid;
      cid.i_source_ = i_source;
      cid.i_group_ = i_group;
      cid.i_conn_ = i_conn;
      ConnectionStatus st = instance->GetConnectionStatus(cid);
      *i_target = st.i_target;
      *port = st.port;
      *syn_group = st.syn_group;
      *delay = st.delay;
      *weight = st.weight;
    })

int NeuronGPU_CreateSynGroup(char *model_name)
    RET(int, 0, ret = instance->CreateSynGroup(model_name))
int NeuronGPU_GetSynGroupNParam(int sg)
    RET(int, 0, ret = instance->GetSynGroupNParam(sg))
char **NeuronGPU_GetSynGroupParamNames(int sg)
    RET(char **, nullptr, ret = NameArray(instance->GetSynGroupParamNames(sg)))
int NeuronGPU_IsSynGroupParam(int sg, char *name)
    RET(int, 0, ret = instance->IsSynGroupParam(sg, name))
int NeuronGPU_GetSynGroupParamIdx(int sg, char *name)
    RET(int, 0, ret = instance->GetSynGroupParamIdx(sg, name))
float NeuronGPU_GetSynGroupParam(int sg, char *name)
    RET(float, 0.0f, ret = instance->GetSynGroupParam(sg, name))
int NeuronGPU_SetSynGroupParam(int sg, char *name, float val)
    RET(int, 0, ret = instance->SetSynGroupParam(sg, name, val))

int NeuronGPU_ActivateSpikeCount(int i_node, int n_node)
    RET(int, 0, ret = instance->ActivateSpikeCount(i_node, n_node))
int NeuronGPU_ActivateRecSpikeTimes(int i_node, int n_node, int max_n)
    RET(int, 0, ret = instance->ActivateRecSpikeTimes(i_node, n_node, max_n))
int NeuronGPU_GetNRecSpikeTimes(int i_node)
    RET(int, 0, ret = instance->GetNRecSpikeTimes(i_node))
float *NeuronGPU_GetRecSpikeTimes(int i_node)
    RET(float *, nullptr, {
      std::vector<float> v = instance->GetRecSpikeTimes(i_node);
      ret = new float[v.size() + 1];
      for (size_t i = 0; i < v.size(); i++) ret[i] = v[i];
    })
int NeuronGPU_PushSpikesToNodes(int n_spikes, int *node_id)
    RET(int, 0, ret = instance->PushSpikesToNodes(n_spikes, node_id))
int NeuronGPU_GetExtNeuronInputSpikes(int *n_spikes, int **node, int **port,
                                      float **spike_height, int include_zeros)
    RET(int, 0,
        ret = instance->GetExtNeuronInputSpikes(n_spikes, node, port,
                                                spike_height, include_zeros != 0))

int NeuronGPU_SetNeuronGroupParam(int i_node, int n_node, char *name, float val)
    RET(int, 0, ret = instance->SetNeuronGroupParam(i_node, n_node, name, val))
int NeuronGPU_IsNeuronGroupParam(int i_node, char *name)
    RET(int, 0, ret = instance->IsNeuronGroupParam(i_node, name))
float NeuronGPU_GetNeuronGroupParam(int i_node, char *name)
    RET(float, 0.0f, ret = instance->GetNeuronGroupParam(i_node, name))
int NeuronGPU_GetNFloatParam() RET(int, 0, ret = instance->GetNFloatParam())
char **NeuronGPU_GetFloatParamNames()
    RET(char **, nullptr, ret = NameArray(instance->GetFloatParamNames()))
int NeuronGPU_IsFloatParam(char *name) RET(int, 0, ret = instance->IsFloatParam(name))
int NeuronGPU_GetFloatParamIdx(char *name)
    RET(int, 0, ret = instance->GetFloatParamIdx(name))
float NeuronGPU_GetFloatParam(char *name)
    RET(float, 0.0f, ret = instance->GetFloatParam(name))
int NeuronGPU_SetFloatParam(char *name, float val)
    RET(int, 0, ret = instance->SetFloatParam(name, val))
int NeuronGPU_GetNIntParam() RET(int, 0, ret = instance->GetNIntParam())
char **NeuronGPU_GetIntParamNames()
    RET(char **, nullptr, ret = NameArray(instance->GetIntParamNames()))
int NeuronGPU_IsIntParam(char *name) RET(int, 0, ret = instance->IsIntParam(name))
int NeuronGPU_GetIntParamIdx(char *name)
    RET(int, 0, ret = instance->GetIntParamIdx(name))
int NeuronGPU_GetIntParam(char *name) RET(int, 0, ret = instance->GetIntParam(name))
int NeuronGPU_SetIntParam(char *name, int val)
    RET(int, 0, ret = instance->SetIntParam(name, val))
int NeuronGPU_RemoteCreate(int i_host, char *model_name, int n_neuron,
                           int n_port)
    RET(int, 0,
        ret = instance->RemoteCreate(i_host, model_name, n_neuron, n_port)
                  .node_seq[0])

int NeuronGPU_ResetKernel() {
  int on_exc = instance ? instance->OnException() : ON_EXCEPTION_EXIT;
  delete instance;
  instance = new NeuronGPU();
  instance->SetOnException(on_exc);
  conn_spec_instance.Init();
  syn_spec_instance.Init();
  return 0;
}

int NeuronGPU_GetHotPathStats(unsigned long long *out3)
    RET(int, 0, {
      NeuronGPUStats s = instance->GetStats();
      out3[0] = s.rk5_accepted;
      out3[1] = s.rk5_rejected;
      out3[2] = s.syn_events;
    })

int NeuronGPU_GetRk5ForcedAccepts(unsigned long long *out1)
    RET(int, 0, out1[0] = instance->GetStats().rk5_forced;)

int NeuronGPU_GetRk5Pass1Stats(unsigned long long *out2)
    RET(int, 0, {
      NeuronGPUStats s = instance->GetStats();
      out2[0] = s.rk5_pass1_accepted;
      out2[1] = s.rk5_pass1_rejected;
    })

}  // extern "C"

extern "C" {
int NeuronGPU_SetKernelTiming(int on)
    RET(int, 0, ret = instance->SetKernelTiming(on))
int NeuronGPU_GetKernelTimes(double *ms8, long long *cnt8)
    RET(int, 0, ret = instance->GetKernelTimes(ms8, cnt8))
int NeuronGPU_SetNeuronScalArray(int i_node, int n_node, char *name,
                                 float *vals)
    RET(int, 0, ret = instance->SetNeuronScalArray(i_node, n_node, name, vals))
long long NeuronGPU_GetConnectionArrays(int i_source, int n_source,
                                        int i_target, int n_target,
                                        long long n_max, int *source,
                                        int *target, unsigned char *port,
                                        unsigned char *syn_group,
                                        float *delay, float *weight)
    RET(long long, 0,
        ret = instance->GetConnectionArrays(i_source, n_source, i_target,
                                            n_target, n_max, source, target,
                                            port, syn_group, delay, weight))
int NeuronGPU_GetConnectionCounts(int i_node, int n_node, int port,
                                  int *in_degree, int *out_degree,
                                  long long *delay_count, int n_delay)
    RET(int, 0,
        ret = instance->GetConnectionCounts(i_node, n_node, port, in_degree,
                                            out_degree, delay_count,
                                            n_delay))
int NeuronGPU_ProposerLog(int on, long long max_samples)
    RET(int, 0, ret = instance->ProposerLog(on, max_samples))
long long NeuronGPU_ProposerNSamples()
    RET(long long, 0, ret = instance->ProposerNSamples())
int NeuronGPU_ProposerGetSamples(float *out, long long n)
    RET(int, 0, ret = instance->ProposerGetSamples(out, n))
int NeuronGPU_ProposerSetSamples(float *in, long long n)
    RET(int, 0, ret = instance->ProposerSetSamples(in, n))
int NeuronGPU_ProposerTrain(int n_iter, int batch, float lr, int adam,
                            int full_batch, float *loss)
    RET(int, 0, {
      float l = instance->ProposerTrain(n_iter, batch, lr, adam, full_batch);
      if (loss) *loss = l;
    })
int NeuronGPU_ProposerGetWeights(float *w1313)
    RET(int, 0, ret = instance->ProposerGetWeights(w1313))
int NeuronGPU_ProposerSetWeights(float *w1313)
    RET(int, 0, ret = instance->ProposerSetWeights(w1313))
int NeuronGPU_ProposerUse(int on)
    RET(int, 0, ret = instance->ProposerUse(on))
int NeuronGPU_ProposerEval(float *feat, int n, float *out)
    RET(int, 0, ret = instance->ProposerEval(feat, n, out))
int NeuronGPU_ProposerEvalNoAlloc(float *feat, int n, float *out)
    RET(int, 0, ret = instance->ProposerEvalNoAlloc(feat, n, out))
}
