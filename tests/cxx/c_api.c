/* Plain C client of include/neurongpu_C.h: host-side calls and the error
 * convention.  Prints one line per check. */
#include <stdio.h>

#include "neurongpu_C.h"

int main(void) {
  NeuronGPU_SetOnException(1);
  NeuronGPU_SetVerbosityLevel(0);
  int i0 = NeuronGPU_Create("aeif_cond_beta", 10, 2);
  printf("create %d err %d\n", i0, NeuronGPU_GetErrorCode());
  int bad = NeuronGPU_Create("no_such_model", 1, 1);
  printf("bad %d err %d msg %s\n", bad, NeuronGPU_GetErrorCode(),
         NeuronGPU_GetErrorMessage());
  NeuronGPU_ConnSpecInit();
  NeuronGPU_SetConnSpecParam("rule", 1);
  NeuronGPU_SynSpecInit();
  NeuronGPU_SetSynSpecFloatParam("weight", 0.5f);
  NeuronGPU_SetSynSpecFloatParam("delay", 1.0f);
  NeuronGPU_ConnectSeqSeq(i0, 10, i0, 10);
  printf("connect err %d\n", NeuronGPU_GetErrorCode());
  int n_conn = 0;
  int *conn = NeuronGPU_GetSeqSeqConnections(i0, 10, i0, 10, -1, &n_conn);
  printf("conns %d %s\n", n_conn, conn ? "ok" : "null");
  return 0;
}
