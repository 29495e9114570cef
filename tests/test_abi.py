"""The drop-in boundary: the shared library loads without a GPU, exports
every function include/neurongpu_C.h declares, and covers the reference's
C ABI (src/neurongpu_C.h) and Python API (pythonlib/neurongpu.py) by name.
No compute calls are made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    from neurongpu_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail("libneurongpu.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB_PATH)


def _declared():
    txt = open(os.path.join(ROOT, "include", "neurongpu_C.h")).read()
    return sorted(set(re.findall(r"\b(NeuronGPU_\w+)\s*\(", txt)))


def _golden(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return [l.strip() for l in f if l.strip()]


def test_library_exports_every_declared_symbol():
    lib = _lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_header_declares_reference_c_abi():
    ref = _golden("c_api_names.txt")
    assert len(ref) == 126
    missing = sorted(set(ref) - set(_declared()))
    assert not missing, missing


def test_header_cites_reference_for_each_reference_function():
    txt = open(os.path.join(ROOT, "include", "neurongpu_C.h")).read()
    ref = set(_golden("c_api_names.txt"))
    lines = txt.split("\n")
    for i, l in enumerate(lines):
        if l.lstrip().startswith(("*", "/")):
            continue
        m = re.search(r"\b(NeuronGPU_\w+)\s*\(", l)
        if m and m.group(1) in ref:
            ctx = "\n".join(lines[max(0, i - 3):i + 1])
            assert "src/neurongpu_C.h:" in ctx, m.group(1)


def test_python_bindings_cover_reference_api():
    from neurongpu_amd import neurongpu as ng
    ref = _golden("py_api_names.txt")
    missing = [n for n in ref if not callable(getattr(ng, n, None))]
    assert not missing, missing


def test_error_convention_without_gpu():
    """Errors come back as codes + message (src/propagate_error.h:4-31);
    no exit while the Python bindings have set the handle mode."""
    lib = _lib()
    from neurongpu_amd import neurongpu as ng  # sets on_exception = handle
    lib.NeuronGPU_Create.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int]
    lib.NeuronGPU_GetErrorMessage.restype = ctypes.c_char_p
    r = lib.NeuronGPU_Create(b"no_such_model", 1, 1)
    assert r == 0
    assert lib.NeuronGPU_GetErrorCode() == 2
    assert b"no_such_model" in lib.NeuronGPU_GetErrorMessage()
    lib.NeuronGPU_GetTimeResolution.restype = ctypes.c_float
    lib.NeuronGPU_GetTimeResolution()
    assert lib.NeuronGPU_GetErrorCode() == 0
    ng.ResetKernel()
