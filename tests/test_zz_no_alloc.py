"""The allocation-free proposer inference (NeuronGPU_ProposerEvalNoAlloc),
called through the C ABI, against NeuronGPU_ProposerEval on the same
feature vectors: identical outputs on two calls in a row, the second with
fewer vectors; sizes beyond its limit and a proposer without weights are
errors."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 6 * 32 + 32 + 32 * 32 + 32 + 32 + 1


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _eval(ngpu, name, x):
    """Calls NeuronGPU_<name> through the C ABI; returns the library's error
    code (0: success; the call itself returns 0 either way, as every entry
    point of the reference's C ABI does) and the outputs."""
    fn = getattr(ngpu._lib, "NeuronGPU_" + name)
    fn.restype = ctypes.c_int
    fn.argtypes = (ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)
    f = np.ascontiguousarray(x, dtype=np.float32)
    out = np.full(f.shape[0], np.nan, dtype=np.float32)
    fn(f.ctypes.data_as(ctypes.c_void_p), f.shape[0],
       out.ctypes.data_as(ctypes.c_void_p))
    return ngpu.GetErrorCode(), out


def test_no_alloc_inference_matches(ngpu):
    log("case no_alloc_inference")
    # no weights yet: the allocation-free call cannot create them
    rc, _ = _eval(ngpu, "ProposerEvalNoAlloc", np.zeros((4, 6)))
    assert rc != 0 and "no weights" in ngpu.GetErrorMessage()
    rng = np.random.default_rng(21)
    ngpu.ProposerSetWeights(rng.normal(0.0, 0.3, NP).astype(np.float32))
    for n in (65536, 1001):
        x = rng.normal(0.0, 1.0, (n, 6)).astype(np.float32)
        rc0, want = _eval(ngpu, "ProposerEval", x)
        rc1, got = _eval(ngpu, "ProposerEvalNoAlloc", x)
        log("n", n, "rc", rc0, rc1, "max |diff|",
            float(np.max(np.abs(got - want))))
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got, want)
    rc, _ = _eval(ngpu, "ProposerEvalNoAlloc", np.zeros((65537, 6)))
    assert rc != 0 and "allocation-free" in ngpu.GetErrorMessage()
