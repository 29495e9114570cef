// Builds (without simulating) the network of the reference's
// c++/examples/brunel_net.cpp through include/neurongpu.h, at a size given
// on the command line, and prints connection counts.  With "sim" as the
// second argument it also simulates 100 ms and prints the spike count.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "neurongpu.h"

int main(int argc, char *argv[]) {
  int n = argc > 1 ? std::atoi(argv[1]) : 1000;
  bool sim = argc > 2 && std::strcmp(argv[2], "sim") == 0;
  NeuronGPU ngpu;
  ngpu.SetOnException(1);
  ngpu.SetVerbosityLevel(0);
  ngpu.SetRandomSeed(1234ULL);
  int order = n / 5;
  int NE = 4 * order, NI = order, N = NE + NI;
  int CE = 800, CI = CE / 4;
  if (CE > NE) CE = NE;
  if (CI > NI) CI = NI;
  NodeSeq pg = ngpu.Create("poisson_generator");
  ngpu.SetNeuronParam(pg, "rate", 20000.0f);
  NodeSeq neuron = ngpu.Create("aeif_cond_beta", N, 2);
  NodeSeq exc = neuron.Subseq(0, NE - 1);
  NodeSeq inh = neuron.Subseq(NE, N - 1);
  float E_rev[] = {0.0, -85.0};
  float tau_decay[] = {1.0, 1.0};
  float tau_rise[] = {1.0, 1.0};
  ngpu.SetNeuronParam(neuron, "E_rev", E_rev, 2);
  ngpu.SetNeuronParam(neuron, "tau_decay", tau_decay, 2);
  ngpu.SetNeuronParam(neuron, "tau_rise", tau_rise, 2);
  float *exc_d = ngpu.RandomNormalClipped(CE * N, 0.5, 0.25, 0.1, 1.25);
  ConnSpec cs1(FIXED_INDEGREE, CE);
  SynSpec ss1;
  ss1.SetParam("receptor", 0);
  ss1.SetParam("weight", 0.05f);
  ss1.SetParam("delay_array", exc_d);
  ngpu.Connect(exc, neuron, cs1, ss1);
  delete[] exc_d;
  float *inh_d = ngpu.RandomNormalClipped(CI * N, 0.5, 0.25, 0.1, 1.25);
  ConnSpec cs2(FIXED_INDEGREE, CI);
  SynSpec ss2;
  ss2.SetParam("receptor", 1);
  ss2.SetParam("weight", 0.35f);
  ss2.SetParam("delay_array", inh_d);
  ngpu.Connect(inh, neuron, cs2, ss2);
  delete[] inh_d;
  ConnSpec cs3(ALL_TO_ALL);
  SynSpec ss3(STANDARD_SYNAPSE, 0.37, 0.2, 0);
  ngpu.Connect(pg, neuron, cs3, ss3);
  std::printf("exc %d\n", (int)ngpu.GetConnections(exc, neuron).size());
  std::printf("inh %d\n", (int)ngpu.GetConnections(inh, neuron).size());
  std::printf("pg %d\n", (int)ngpu.GetConnections(pg, neuron).size());
  if (sim) {
    ngpu.ActivateSpikeCount(neuron);
    ngpu.Simulate(100.0);
    int *cnt = ngpu.GetNeuronIntVar(neuron, "spike_count");
    long total = 0;
    for (int i = 0; i < N; i++) total += cnt[i];
    std::printf("spikes %ld\n", total);
  }
  return ngpu.GetErrorCode();
}
