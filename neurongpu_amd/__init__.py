"""MI355X-native NeuronGPU hot path: HIP kernels for gfx950 behind the
reference's C ABI (include/neurongpu_C.h), C++ class (include/neurongpu.h)
and Python bindings (neurongpu_amd.neurongpu)."""
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libneurongpu.so")
