"""Python interface of the MI355X NeuronGPU engine.

Counterpart of the reference bindings pythonlib/neurongpu.py: the same
module-level functions (Create, Connect, SetStatus, GetStatus, Simulate,
CreateRecord, GetRecordData, CreateSynGroup, ...), the same argument
meaning, the same dictionaries (connection rules, synapse specs, value
distributions {"array"|"normal"|"normal_clipped"}) and the same error
behaviour (a ValueError carrying the library's message).

The shared library is loaded with ctypes from $NEURONGPU_LIB, or else from
the in-tree build neurongpu_amd/libneurongpu.so (the reference hard-codes
/usr/local/lib/libneurongpu.so, pythonlib/neurongpu.py:17-18).  There is no
fallback: a missing library raises ImportError.
"""
import ctypes
import gc
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.environ.get("NEURONGPU_LIB",
                          os.path.join(_HERE, "libneurongpu.so"))
if not os.path.exists(lib_path):
    raise ImportError("NeuronGPU library not found at %s; build it with "
                      "python -m neurongpu_amd.build" % lib_path)
_lib = ctypes.CDLL(lib_path)

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_char_p = ctypes.c_char_p
c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_float = ctypes.c_float
c_uint_p = ctypes.POINTER(ctypes.c_uint)
c_char_pp = ctypes.POINTER(ctypes.c_char_p)
c_float_pp = ctypes.POINTER(c_float_p)

conn_rule_name = ("one_to_one", "all_to_all", "fixed_total_number",
                  "fixed_indegree", "fixed_outdegree")

# name -> (restype, argtypes)
_SIG = {
    "GetErrorMessage": (c_char_p, ()),
    "GetErrorCode": (ctypes.c_ubyte, ()),
    "SetOnException": (None, (c_int,)),
    "SetRandomSeed": (c_int, (ctypes.c_ulonglong,)),
    "SetTimeResolution": (c_int, (c_float,)),
    "GetTimeResolution": (c_float, ()),
    "SetMaxSpikeBufferSize": (c_int, (c_int,)),
    "GetMaxSpikeBufferSize": (c_int, ()),
    "SetSimTime": (c_int, (c_float,)),
    "SetVerbosityLevel": (c_int, (c_int,)),
    "Create": (c_int, (c_char_p, c_int, c_int)),
    "CreatePoissonGenerator": (c_int, (c_int, c_float)),
    "CreateRecord": (c_int, (c_char_p, c_void_p, c_void_p, c_void_p, c_int)),
    "GetRecordDataRows": (c_int, (c_int,)),
    "GetRecordDataColumns": (c_int, (c_int,)),
    "GetRecordData": (c_float_pp, (c_int,)),
    "SetNeuronScalParam": (c_int, (c_int, c_int, c_char_p, c_float)),
    "SetNeuronArrayParam": (c_int, (c_int, c_int, c_char_p, c_void_p, c_int)),
    "SetNeuronPtScalParam": (c_int, (c_void_p, c_int, c_char_p, c_float)),
    "SetNeuronPtArrayParam": (c_int, (c_void_p, c_int, c_char_p, c_void_p,
                                      c_int)),
    "IsNeuronScalParam": (c_int, (c_int, c_char_p)),
    "IsNeuronPortParam": (c_int, (c_int, c_char_p)),
    "IsNeuronArrayParam": (c_int, (c_int, c_char_p)),
    "SetNeuronIntVar": (c_int, (c_int, c_int, c_char_p, c_int)),
    "SetNeuronScalVar": (c_int, (c_int, c_int, c_char_p, c_float)),
    "SetNeuronArrayVar": (c_int, (c_int, c_int, c_char_p, c_void_p, c_int)),
    "SetNeuronPtIntVar": (c_int, (c_void_p, c_int, c_char_p, c_int)),
    "SetNeuronPtScalVar": (c_int, (c_void_p, c_int, c_char_p, c_float)),
    "SetNeuronPtArrayVar": (c_int, (c_void_p, c_int, c_char_p, c_void_p,
                                    c_int)),
    "IsNeuronIntVar": (c_int, (c_int, c_char_p)),
    "IsNeuronScalVar": (c_int, (c_int, c_char_p)),
    "IsNeuronPortVar": (c_int, (c_int, c_char_p)),
    "IsNeuronArrayVar": (c_int, (c_int, c_char_p)),
    "GetNeuronParamSize": (c_int, (c_int, c_char_p)),
    "GetNeuronVarSize": (c_int, (c_int, c_char_p)),
    "GetNeuronParam": (c_float_p, (c_int, c_int, c_char_p)),
    "GetNeuronPtParam": (c_float_p, (c_void_p, c_int, c_char_p)),
    "GetArrayParam": (c_float_p, (c_int, c_char_p)),
    "GetNeuronIntVar": (c_int_p, (c_int, c_int, c_char_p)),
    "GetNeuronPtIntVar": (c_int_p, (c_void_p, c_int, c_char_p)),
    "GetNeuronVar": (c_float_p, (c_int, c_int, c_char_p)),
    "GetNeuronPtVar": (c_float_p, (c_void_p, c_int, c_char_p)),
    "GetArrayVar": (c_float_p, (c_int, c_char_p)),
    "Calibrate": (c_int, ()),
    "Simulate": (c_int, ()),
    "StartSimulation": (c_int, ()),
    "SimulationStep": (c_int, ()),
    "EndSimulation": (c_int, ()),
    "ConnectMpiInit": (c_int, (c_int, c_void_p)),
    "MpiId": (c_int, ()),
    "MpiNp": (c_int, ()),
    "ProcMaster": (c_int, ()),
    "MpiFinalize": (c_int, ()),
    "RandomInt": (c_uint_p, (ctypes.c_size_t,)),
    "RandomUniform": (c_float_p, (ctypes.c_size_t,)),
    "RandomNormal": (c_float_p, (ctypes.c_size_t, c_float, c_float)),
    "RandomNormalClipped": (c_float_p, (ctypes.c_size_t, c_float, c_float,
                                        c_float, c_float, c_float)),
    "Connect": (c_int, (c_int, c_int, ctypes.c_ubyte, ctypes.c_ubyte, c_float,
                        c_float)),
    "ConnSpecInit": (c_int, ()),
    "SetConnSpecParam": (c_int, (c_char_p, c_int)),
    "ConnSpecIsParam": (c_int, (c_char_p,)),
    "SynSpecInit": (c_int, ()),
    "SetSynSpecIntParam": (c_int, (c_char_p, c_int)),
    "SetSynSpecFloatParam": (c_int, (c_char_p, c_float)),
    "SetSynSpecFloatPtParam": (c_int, (c_char_p, c_void_p)),
    "SynSpecIsIntParam": (c_int, (c_char_p,)),
    "SynSpecIsFloatParam": (c_int, (c_char_p,)),
    "SynSpecIsFloatPtParam": (c_int, (c_char_p,)),
    "ConnectSeqSeq": (c_int, (c_int, c_int, c_int, c_int)),
    "ConnectSeqGroup": (c_int, (c_int, c_int, c_void_p, c_int)),
    "ConnectGroupSeq": (c_int, (c_void_p, c_int, c_int, c_int)),
    "ConnectGroupGroup": (c_int, (c_void_p, c_int, c_void_p, c_int)),
    "RemoteConnectSeqSeq": (c_int, (c_int, c_int, c_int, c_int, c_int, c_int)),
    "RemoteConnectSeqGroup": (c_int, (c_int, c_int, c_int, c_int, c_void_p,
                                      c_int)),
    "RemoteConnectGroupSeq": (c_int, (c_int, c_void_p, c_int, c_int, c_int,
                                      c_int)),
    "RemoteConnectGroupGroup": (c_int, (c_int, c_void_p, c_int, c_int,
                                        c_void_p, c_int)),
    "GetIntVarNames": (c_char_pp, (c_int,)),
    "GetScalVarNames": (c_char_pp, (c_int,)),
    "GetNIntVar": (c_int, (c_int,)),
    "GetNScalVar": (c_int, (c_int,)),
    "GetPortVarNames": (c_char_pp, (c_int,)),
    "GetNPortVar": (c_int, (c_int,)),
    "GetScalParamNames": (c_char_pp, (c_int,)),
    "GetNScalParam": (c_int, (c_int,)),
    "GetPortParamNames": (c_char_pp, (c_int,)),
    "GetNGroupParam": (c_int, (c_int,)),
    "GetGroupParamNames": (c_char_pp, (c_int,)),
    "GetNPortParam": (c_int, (c_int,)),
    "GetArrayParamNames": (c_char_pp, (c_int,)),
    "GetNArrayParam": (c_int, (c_int,)),
    "GetArrayVarNames": (c_char_pp, (c_int,)),
    "GetNArrayVar": (c_int, (c_int,)),
    "GetSeqSeqConnections": (c_int_p, (c_int, c_int, c_int, c_int, c_int,
                                       c_int_p)),
    "GetSeqGroupConnections": (c_int_p, (c_int, c_int, c_void_p, c_int, c_int,
                                         c_int_p)),
    "GetGroupSeqConnections": (c_int_p, (c_void_p, c_int, c_int, c_int, c_int,
                                         c_int_p)),
    "GetGroupGroupConnections": (c_int_p, (c_void_p, c_int, c_void_p, c_int,
                                           c_int, c_int_p)),
    "GetConnectionStatus": (c_int, (c_int, c_int, c_int, c_int_p,
                                    ctypes.POINTER(ctypes.c_ubyte),
                                    ctypes.POINTER(ctypes.c_ubyte), c_float_p,
                                    c_float_p)),
    "CreateSynGroup": (c_int, (c_char_p,)),
    "GetSynGroupNParam": (c_int, (c_int,)),
    "GetSynGroupParamNames": (c_char_pp, (c_int,)),
    "IsSynGroupParam": (c_int, (c_int, c_char_p)),
    "GetSynGroupParamIdx": (c_int, (c_int, c_char_p)),
    "GetSynGroupParam": (c_float, (c_int, c_char_p)),
    "SetSynGroupParam": (c_int, (c_int, c_char_p, c_float)),
    "ActivateSpikeCount": (c_int, (c_int, c_int)),
    "ActivateRecSpikeTimes": (c_int, (c_int, c_int, c_int)),
    "GetNRecSpikeTimes": (c_int, (c_int,)),
    "GetRecSpikeTimes": (c_float_p, (c_int,)),
    "PushSpikesToNodes": (c_int, (c_int, c_void_p)),
    "GetExtNeuronInputSpikes": (c_int, (c_int_p, c_void_p, c_void_p, c_void_p,
                                        c_int)),
    "SetNeuronGroupParam": (c_int, (c_int, c_int, c_char_p, c_float)),
    "IsNeuronGroupParam": (c_int, (c_int, c_char_p)),
    "GetNeuronGroupParam": (c_float, (c_int, c_char_p)),
    "GetNFloatParam": (c_int, ()),
    "GetFloatParamNames": (c_char_pp, ()),
    "IsFloatParam": (c_int, (c_char_p,)),
    "GetFloatParamIdx": (c_int, (c_char_p,)),
    "GetFloatParam": (c_float, (c_char_p,)),
    "SetFloatParam": (c_int, (c_char_p, c_float)),
    "GetNIntParam": (c_int, ()),
    "GetIntParamNames": (c_char_pp, ()),
    "IsIntParam": (c_int, (c_char_p,)),
    "GetIntParamIdx": (c_int, (c_char_p,)),
    "GetIntParam": (c_int, (c_char_p,)),
    "SetIntParam": (c_int, (c_char_p, c_int)),
    "RemoteCreate": (c_int, (c_int, c_char_p, c_int, c_int)),
    "ResetKernel": (c_int, ()),
    "GetHotPathStats": (c_int, (ctypes.POINTER(ctypes.c_ulonglong),)),
    "GetRk5Pass1Stats": (c_int, (ctypes.POINTER(ctypes.c_ulonglong),)),
    "GetRk5ForcedAccepts": (c_int, (ctypes.POINTER(ctypes.c_ulonglong),)),
    "SetKernelTiming": (c_int, (c_int,)),
    "GetKernelTimes": (c_int, (ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_longlong))),
    "ProposerLog": (c_int, (c_int, ctypes.c_longlong)),
    "ProposerNSamples": (ctypes.c_longlong, ()),
    "ProposerGetSamples": (c_int, (c_void_p, ctypes.c_longlong)),
    "ProposerSetSamples": (c_int, (c_void_p, ctypes.c_longlong)),
    "ProposerTrain": (c_int, (c_int, c_int, c_float, c_int, c_int,
                              ctypes.POINTER(c_float))),
    "ProposerGetWeights": (c_int, (c_void_p,)),
    "ProposerSetWeights": (c_int, (c_void_p,)),
    "ProposerUse": (c_int, (c_int,)),
    "ProposerEval": (c_int, (c_void_p, c_int, c_void_p)),
    "ProposerEvalNoAlloc": (c_int, (c_void_p, c_int, c_void_p)),
    "SetNeuronScalArray": (c_int, (c_int, c_int, c_char_p, c_void_p)),
    "GetConnectionArrays": (ctypes.c_longlong,
                            (c_int, c_int, c_int, c_int, ctypes.c_longlong,
                             c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p)),
    "GetConnectionCounts": (c_int, (c_int, c_int, c_int, c_void_p, c_void_p,
                                    c_void_p, c_int)),
}

_C = {}
for _name, (_res, _args) in _SIG.items():
    _f = getattr(_lib, "NeuronGPU_" + _name)
    _f.restype = _res
    _f.argtypes = _args
    _C[_name] = _f


def _b(s):
    if isinstance(s, bytes):
        return s
    if isinstance(s, str):
        return s.encode("ascii")
    raise ValueError("Variable cannot be converted to string")


to_byte_str = _b  # pythonlib/neurongpu.py:74-80


def to_def_str(s):
    "bytes -> str (pythonlib/neurongpu.py:82-86)"
    return s.decode("utf-8")


def waitenter(val):
    "Prompt and wait for Enter (pythonlib/neurongpu.py:88-92)"
    return input(val)


def GetErrorMessage():
    "Get error message from NeuronGPU exception"
    m = _C["GetErrorMessage"]()
    return m.decode() if m is not None else ""


def GetErrorCode():
    "Get error code from NeuronGPU exception"
    return _C["GetErrorCode"]()


def _check(ret=None):
    if GetErrorCode() != 0:
        raise ValueError(GetErrorMessage())
    return ret


def _call(name, *args):
    ret = _C[name](*args)
    return _check(ret)


def SetOnException(on_exception):
    "Define whether to handle exceptions (1) or exit (0) in case of errors"
    return _C["SetOnException"](on_exception)


SetOnException(1)


class NodeSeq(object):
    """Contiguous range of node indexes (pythonlib/neurongpu.py:25-57)."""

    def __init__(self, i0, n=1):
        if i0 is None:
            i0 = 0
            n = -1
        self.i0 = i0
        self.n = n

    def Subseq(self, first, last):
        if last < 0 and last >= -self.n:
            last = last % self.n
        if first < 0 or last < first:
            raise ValueError("Sequence subset range error")
        if last >= self.n:
            raise ValueError("Sequence subset out of range")
        return NodeSeq(self.i0 + first, last - first + 1)

    def __getitem__(self, i):
        if isinstance(i, slice):
            if i.step is not None:
                raise ValueError("Subsequence cannot have a step")
            return self.Subseq(i.start, i.stop - 1)
        if i < -self.n:
            raise ValueError("Sequence index error")
        if i >= self.n:
            raise ValueError("Sequence index out of range")
        if i < 0:
            i = i % self.n
        return self.i0 + i

    def ToList(self):
        return list(range(self.i0, self.i0 + self.n))

    def __len__(self):
        return self.n


class RemoteNodeSeq(object):
    def __init__(self, i_host=0, node_seq=NodeSeq(None)):
        self.i_host = i_host
        self.node_seq = node_seq


class ConnectionId(object):
    def __init__(self, i_source, i_group, i_conn):
        self.i_source = i_source
        self.i_group = i_group
        self.i_conn = i_conn


class SynGroup(object):
    def __init__(self, i_syn_group):
        self.i_syn_group = i_syn_group


def _int_array(nodes):
    a = np.ascontiguousarray(np.asarray(list(nodes) if not isinstance(
        nodes, np.ndarray) else nodes, dtype=np.int32))
    return a, a.ctypes.data_as(c_void_p)


def _float_array(vals):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.float32).ravel())
    return a, a.ctypes.data_as(c_void_p)


def _names(name, *args):
    pp = _C[name](*args)
    _check()
    out = []
    i = 0
    while pp[i] is not None:
        out.append(pp[i].decode())
        i += 1
    return out


# ---------------------------------------------------------------------
# kernel settings
def SetRandomSeed(seed):
    "Set seed for random number generation"
    return _call("SetRandomSeed", ctypes.c_ulonglong(seed))


def SetTimeResolution(time_res):
    "Set time resolution in ms"
    return _call("SetTimeResolution", time_res)


def GetTimeResolution():
    "Get time resolution in ms"
    return _call("GetTimeResolution")


def SetMaxSpikeBufferSize(max_size):
    "Set maximum size of spike buffer per node"
    return _call("SetMaxSpikeBufferSize", max_size)


def GetMaxSpikeBufferSize():
    "Get maximum size of spike buffer per node"
    return _call("GetMaxSpikeBufferSize")


def SetSimTime(sim_time):
    "Set neural activity simulated time in ms"
    return _call("SetSimTime", sim_time)


def SetVerbosityLevel(verbosity_level):
    "Set verbosity level"
    return _call("SetVerbosityLevel", verbosity_level)


# ---------------------------------------------------------------------
# creation and records
def Create(model_name, n_node=1, n_ports=1, status_dict=None):
    "Create a neuron group"
    if isinstance(status_dict, dict):
        node_group = Create(model_name, n_node, n_ports)
        SetStatus(node_group, status_dict)
        return node_group
    elif status_dict is not None:
        raise ValueError("Wrong argument in Create")
    i_node = _call("Create", _b(model_name), n_node, n_ports)
    return NodeSeq(i_node, n_node)


def CreatePoissonGenerator(n_node, rate):
    "Create a poisson-distributed spike generator"
    i_node = _call("CreatePoissonGenerator", n_node, rate)
    return NodeSeq(i_node, n_node)


def CreateRecord(file_name, var_name_list, i_node_list, i_port_list):
    "Create a record of neuron variables"
    n = len(i_node_list)
    names = (c_char_p * n)(*[_b(v) for v in var_name_list])
    nodes, nodes_p = _int_array(i_node_list)
    ports, ports_p = _int_array(i_port_list)
    return _call("CreateRecord", _b(file_name),
                 ctypes.cast(names, c_void_p), nodes_p, ports_p, n)


def GetRecordDataRows(i_record):
    "Get record n. of rows"
    return _call("GetRecordDataRows", i_record)


def GetRecordDataColumns(i_record):
    "Get record n. of columns"
    return _call("GetRecordDataColumns", i_record)


def GetRecordData(i_record):
    "Get record data as a list of rows [t, v1, ...]"
    pt = _call("GetRecordData", i_record)
    nr = GetRecordDataRows(i_record)
    nc = GetRecordDataColumns(i_record)
    return [[pt[ir][ic] for ic in range(nc)] for ir in range(nr)]


def GetRecordDataArray(i_record):
    "Record data as a numpy array (rows x columns); no reference counterpart"
    pt = _call("GetRecordData", i_record)
    nr = GetRecordDataRows(i_record)
    nc = GetRecordDataColumns(i_record) if nr else 0
    out = np.empty((nr, nc), dtype=np.float32)
    for ir in range(nr):
        out[ir] = np.ctypeslib.as_array(pt[ir], shape=(nc,))
    return out


# ---------------------------------------------------------------------
# neuron parameters and variables
def SetNeuronScalParam(i_node, n_node, param_name, val):
    "Set neuron scalar parameter value"
    return _call("SetNeuronScalParam", i_node, n_node, _b(param_name), val)


def SetNeuronArrayParam(i_node, n_node, param_name, param_list):
    "Set neuron array parameter value"
    a, p = _float_array(param_list)
    return _call("SetNeuronArrayParam", i_node, n_node, _b(param_name), p,
                 a.size)


def SetNeuronPtScalParam(nodes, param_name, val):
    "Set neuron list scalar parameter value"
    a, p = _int_array(nodes)
    return _call("SetNeuronPtScalParam", p, a.size, _b(param_name), val)


def SetNeuronPtArrayParam(nodes, param_name, param_list):
    "Set neuron list array parameter value"
    a, p = _int_array(nodes)
    v, vp = _float_array(param_list)
    return _call("SetNeuronPtArrayParam", p, a.size, _b(param_name), vp,
                 v.size)


def IsNeuronScalParam(i_node, param_name):
    "Check name of neuron scalar parameter"
    return _call("IsNeuronScalParam", i_node, _b(param_name)) != 0


def IsNeuronPortParam(i_node, param_name):
    "Check name of neuron port parameter"
    return _call("IsNeuronPortParam", i_node, _b(param_name)) != 0


def IsNeuronArrayParam(i_node, param_name):
    "Check name of neuron array parameter"
    return _call("IsNeuronArrayParam", i_node, _b(param_name)) != 0


def IsNeuronGroupParam(i_node, param_name):
    "Check name of neuron group parameter"
    return _call("IsNeuronGroupParam", i_node, _b(param_name)) != 0


def SetNeuronIntVar(i_node, n_node, var_name, val):
    "Set neuron integer variable value"
    return _call("SetNeuronIntVar", i_node, n_node, _b(var_name), int(val))


def SetNeuronScalVar(i_node, n_node, var_name, val):
    "Set neuron scalar variable value"
    return _call("SetNeuronScalVar", i_node, n_node, _b(var_name), val)


def SetNeuronArrayVar(i_node, n_node, var_name, var_list):
    "Set neuron array variable value"
    a, p = _float_array(var_list)
    return _call("SetNeuronArrayVar", i_node, n_node, _b(var_name), p, a.size)


def SetNeuronPtIntVar(nodes, var_name, val):
    "Set neuron list integer variable value"
    a, p = _int_array(nodes)
    return _call("SetNeuronPtIntVar", p, a.size, _b(var_name), int(val))


def SetNeuronPtScalVar(nodes, var_name, val):
    "Set neuron list scalar variable value"
    a, p = _int_array(nodes)
    return _call("SetNeuronPtScalVar", p, a.size, _b(var_name), val)


def SetNeuronPtArrayVar(nodes, var_name, var_list):
    "Set neuron list array variable value"
    a, p = _int_array(nodes)
    v, vp = _float_array(var_list)
    return _call("SetNeuronPtArrayVar", p, a.size, _b(var_name), vp, v.size)


def IsNeuronIntVar(i_node, var_name):
    return _call("IsNeuronIntVar", i_node, _b(var_name)) != 0


def IsNeuronScalVar(i_node, var_name):
    return _call("IsNeuronScalVar", i_node, _b(var_name)) != 0


def IsNeuronPortVar(i_node, var_name):
    return _call("IsNeuronPortVar", i_node, _b(var_name)) != 0


def IsNeuronArrayVar(i_node, var_name):
    return _call("IsNeuronArrayVar", i_node, _b(var_name)) != 0


def GetNeuronParamSize(i_node, param_name):
    return _call("GetNeuronParamSize", i_node, _b(param_name))


def GetNeuronVarSize(i_node, var_name):
    return _call("GetNeuronVarSize", i_node, _b(var_name))


def _rows(pt, n_node, size):
    if size > 1:
        return [[pt[i * size + k] for k in range(size)] for i in range(n_node)]
    return [pt[i] for i in range(n_node)]


def GetNeuronParam(i_node, n_node, param_name):
    "Get neuron parameter value"
    pt = _call("GetNeuronParam", i_node, n_node, _b(param_name))
    return _rows(pt, n_node, GetNeuronParamSize(i_node, param_name))


def GetNeuronPtParam(nodes, param_name):
    "Get neuron list parameter value"
    a, p = _int_array(nodes)
    pt = _call("GetNeuronPtParam", p, a.size, _b(param_name))
    return _rows(pt, a.size, GetNeuronParamSize(int(a[0]), param_name))


def GetArrayParam(i_node, n_node, param_name):
    "Get neuron array parameter"
    out = []
    for j in range(n_node):
        pt = _call("GetArrayParam", i_node + j, _b(param_name))
        size = GetNeuronParamSize(i_node + j, param_name)
        out.append([pt[k] for k in range(size)])
    return out


def GetNeuronListArrayParam(node_list, param_name):
    "Get neuron array parameter"
    out = []
    for i_node in node_list:
        pt = _call("GetArrayParam", i_node, _b(param_name))
        size = GetNeuronParamSize(i_node, param_name)
        out.append([pt[k] for k in range(size)])
    return out


def GetNeuronGroupParam(i_node, param_name):
    "Get neuron group parameter"
    return _call("GetNeuronGroupParam", i_node, _b(param_name))


def GetNeuronIntVar(i_node, n_node, var_name):
    "Get neuron integer variable value"
    pt = _call("GetNeuronIntVar", i_node, n_node, _b(var_name))
    return [pt[i] for i in range(n_node)]


def GetNeuronVar(i_node, n_node, var_name):
    "Get neuron variable value"
    pt = _call("GetNeuronVar", i_node, n_node, _b(var_name))
    return _rows(pt, n_node, GetNeuronVarSize(i_node, var_name))


def GetNeuronPtIntVar(nodes, var_name):
    a, p = _int_array(nodes)
    pt = _call("GetNeuronPtIntVar", p, a.size, _b(var_name))
    return [pt[i] for i in range(a.size)]


def GetNeuronPtVar(nodes, var_name):
    a, p = _int_array(nodes)
    pt = _call("GetNeuronPtVar", p, a.size, _b(var_name))
    return _rows(pt, a.size, GetNeuronVarSize(int(a[0]), var_name))


def GetArrayVar(i_node, n_node, var_name):
    out = []
    for j in range(n_node):
        pt = _call("GetArrayVar", i_node + j, _b(var_name))
        size = GetNeuronVarSize(i_node + j, var_name)
        out.append([pt[k] for k in range(size)])
    return out


def GetNeuronListArrayVar(node_list, var_name):
    out = []
    for i_node in node_list:
        pt = _call("GetArrayVar", i_node, _b(var_name))
        size = GetNeuronVarSize(i_node, var_name)
        out.append([pt[k] for k in range(size)])
    return out


def GetNeuronStatus(nodes, var_name):
    "Get neuron group scalar or array variable or parameter"
    if not isinstance(nodes, (list, tuple, NodeSeq)):
        raise ValueError("Unknown node type")
    if isinstance(nodes, NodeSeq):
        i0, n = nodes.i0, nodes.n
        if IsNeuronScalParam(i0, var_name) or IsNeuronPortParam(i0, var_name):
            return GetNeuronParam(i0, n, var_name)
        if IsNeuronArrayParam(i0, var_name):
            return GetArrayParam(i0, n, var_name)
        if IsNeuronIntVar(i0, var_name):
            return GetNeuronIntVar(i0, n, var_name)
        if IsNeuronScalVar(i0, var_name) or IsNeuronPortVar(i0, var_name):
            return GetNeuronVar(i0, n, var_name)
        if IsNeuronArrayVar(i0, var_name):
            return GetArrayVar(i0, n, var_name)
        if IsNeuronGroupParam(i0, var_name):
            return GetNeuronStatus(nodes.ToList(), var_name)
        raise ValueError("Unknown neuron variable or parameter")
    n0 = nodes[0]
    if IsNeuronScalParam(n0, var_name) or IsNeuronPortParam(n0, var_name):
        return GetNeuronPtParam(nodes, var_name)
    if IsNeuronArrayParam(n0, var_name):
        return GetNeuronListArrayParam(nodes, var_name)
    if IsNeuronIntVar(n0, var_name):
        return GetNeuronPtIntVar(nodes, var_name)
    if IsNeuronScalVar(n0, var_name) or IsNeuronPortVar(n0, var_name):
        return GetNeuronPtVar(nodes, var_name)
    if IsNeuronArrayVar(n0, var_name):
        return GetNeuronListArrayVar(nodes, var_name)
    if IsNeuronGroupParam(n0, var_name):
        return [GetNeuronGroupParam(i, var_name) for i in nodes]
    raise ValueError("Unknown neuron variable or parameter")


def GetNIntVar(i_node):
    return _call("GetNIntVar", i_node)


def GetNScalVar(i_node):
    return _call("GetNScalVar", i_node)


def GetIntVarNames(i_node):
    return _names("GetIntVarNames", i_node)


def GetScalVarNames(i_node):
    return _names("GetScalVarNames", i_node)


def GetNPortVar(i_node):
    return _call("GetNPortVar", i_node)


def GetPortVarNames(i_node):
    return _names("GetPortVarNames", i_node)


def GetNScalParam(i_node):
    return _call("GetNScalParam", i_node)


def GetScalParamNames(i_node):
    return _names("GetScalParamNames", i_node)


def GetNPortParam(i_node):
    return _call("GetNPortParam", i_node)


def GetPortParamNames(i_node):
    return _names("GetPortParamNames", i_node)


def GetNArrayParam(i_node):
    return _call("GetNArrayParam", i_node)


def GetArrayParamNames(i_node):
    return _names("GetArrayParamNames", i_node)


def GetNGroupParam(i_node):
    return _call("GetNGroupParam", i_node)


def GetGroupParamNames(i_node):
    return _names("GetGroupParamNames", i_node)


def GetNArrayVar(i_node):
    return _call("GetNArrayVar", i_node)


def GetArrayVarNames(i_node):
    return _names("GetArrayVarNames", i_node)


def SetNeuronGroupParam(nodes, param_name, val):
    "Set neuron group parameter value"
    if not isinstance(nodes, NodeSeq):
        raise ValueError("Wrong argument type in SetNeuronGroupParam")
    return _call("SetNeuronGroupParam", nodes.i0, nodes.n, _b(param_name),
                 val)


def DictToArray(param_dict, array_size):
    """Values from a distribution dictionary (pythonlib/neurongpu.py:1472-1514):
    {"array": values} | {"distribution": "normal", "mu", "sigma"} |
    {"distribution": "normal_clipped", "mu", "sigma", "low", "high",
    "step"}."""
    dist_name = None
    arr = None
    low, high, mu, sigma, vstep = -1.0e35, 1.0e35, None, None, 0
    for k, v in param_dict.items():
        if k == "array":
            dist_name, arr = "array", v
        elif k == "distribution":
            dist_name = v
        elif k == "low":
            low = v
        elif k == "high":
            high = v
        elif k == "mu":
            mu = v
        elif k == "sigma":
            sigma = v
        elif k == "step":
            vstep = v
        else:
            raise ValueError("Unknown parameter name in dictionary")
    if dist_name == "array":
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).ravel())
        if a.size != array_size:
            raise ValueError("Wrong array size.")
        return a
    if dist_name == "normal":
        return RandomNormal(array_size, mu, sigma)
    if dist_name == "normal_clipped":
        return RandomNormalClipped(array_size, mu, sigma, low, high, vstep)
    if dist_name == "uniform":
        u = RandomUniform(array_size)
        a = (np.float32(low) + np.float32(high - low) * u).astype(np.float32)
        if vstep > 0:
            n = int(np.floor((high - low) / vstep + 0.5)) + 1
            j = np.minimum(np.floor(u * n), n - 1)
            a = (np.float32(low) + np.float32(vstep) * j).astype(np.float32)
        return a
    raise ValueError("Unknown distribution")


def SetNeuronStatus(nodes, var_name, val):
    "Set neuron group scalar or array variable or parameter"
    if not isinstance(nodes, (list, tuple, NodeSeq)):
        raise ValueError("Unknown node type")
    if isinstance(val, dict):
        n = len(nodes)
        arr = DictToArray(val, n)
        if isinstance(nodes, NodeSeq):
            i0 = nodes.i0
            if IsNeuronScalParam(i0, var_name) or IsNeuronScalVar(i0, var_name):
                # one value per node in one call (same result as the
                # reference's per-node loop, pythonlib/neurongpu.py:1146-1151)
                a = np.ascontiguousarray(arr, dtype=np.float32)
                _call("SetNeuronScalArray", i0, nodes.n, _b(var_name),
                      a.ctypes.data_as(c_void_p))
                return
        for i in range(n):
            SetNeuronStatus([nodes[i]], var_name, float(arr[i]))
        return
    if isinstance(nodes, NodeSeq):
        i0, n = nodes.i0, nodes.n
        if IsNeuronGroupParam(i0, var_name):
            SetNeuronGroupParam(nodes, var_name, val)
        elif IsNeuronScalParam(i0, var_name):
            SetNeuronScalParam(i0, n, var_name, val)
        elif IsNeuronPortParam(i0, var_name) or IsNeuronArrayParam(i0,
                                                                    var_name):
            SetNeuronArrayParam(i0, n, var_name, val)
        elif IsNeuronIntVar(i0, var_name):
            SetNeuronIntVar(i0, n, var_name, val)
        elif IsNeuronScalVar(i0, var_name):
            SetNeuronScalVar(i0, n, var_name, val)
        elif IsNeuronPortVar(i0, var_name) or IsNeuronArrayVar(i0, var_name):
            SetNeuronArrayVar(i0, n, var_name, val)
        else:
            raise ValueError("Unknown neuron variable or parameter")
    else:
        n0 = nodes[0]
        if IsNeuronScalParam(n0, var_name):
            SetNeuronPtScalParam(nodes, var_name, val)
        elif IsNeuronPortParam(n0, var_name) or IsNeuronArrayParam(n0,
                                                                    var_name):
            SetNeuronPtArrayParam(nodes, var_name, val)
        elif IsNeuronIntVar(n0, var_name):
            SetNeuronPtIntVar(nodes, var_name, val)
        elif IsNeuronScalVar(n0, var_name):
            SetNeuronPtScalVar(nodes, var_name, val)
        elif IsNeuronPortVar(n0, var_name) or IsNeuronArrayVar(n0, var_name):
            SetNeuronPtArrayVar(nodes, var_name, val)
        else:
            raise ValueError("Unknown neuron variable or parameter")


def SetStatus(gen_object, params, val=None):
    "Set neuron or synapse group parameters or variables using dictionaries"
    if isinstance(gen_object, RemoteNodeSeq):
        if gen_object.i_host == 0:
            SetStatus(gen_object.node_seq, params, val)
        return
    gc.disable()
    try:
        if isinstance(gen_object, SynGroup):
            return SetSynGroupStatus(gen_object, params, val)
        nodes = gen_object
        if val is not None:
            SetNeuronStatus(nodes, params, val)
        elif isinstance(params, dict):
            for k in params:
                SetNeuronStatus(nodes, k, params[k])
        elif isinstance(params, (list, tuple)):
            if len(params) != len(nodes):
                raise ValueError("List should have the same size as nodes")
            for d in params:
                if not isinstance(d, dict):
                    raise ValueError("Type of list elements should be dict")
                for k in d:
                    SetNeuronStatus(nodes, k, d[k])
        else:
            raise ValueError("Wrong argument in SetStatus")
        _check()
    finally:
        gc.enable()


# ---------------------------------------------------------------------
# simulation
def Calibrate():
    "Calibrate simulation"
    return _call("Calibrate")


def Simulate(sim_time=1000.0):
    "Simulate neural activity"
    SetSimTime(sim_time)
    return _call("Simulate")


def StartSimulation():
    return _call("StartSimulation")


def SimulationStep():
    return _call("SimulationStep")


def EndSimulation():
    return _call("EndSimulation")


def ConnectMpiInit():
    "Initialize MPI connections (not available in the single-GPU build)"
    argv = (c_char_p * len(sys.argv))(*[_b(a) for a in sys.argv])
    return _call("ConnectMpiInit", len(sys.argv), ctypes.cast(argv, c_void_p))


def MpiId():
    return _call("MpiId")


def Rank():
    return MpiId()


def MpiNp():
    return _call("MpiNp")


def ProcMaster():
    return _call("ProcMaster")


def MpiFinalize():
    return _call("MpiFinalize")


# ---------------------------------------------------------------------
# random numbers
def RandomInt(n):
    pt = _call("RandomInt", n)
    return np.ctypeslib.as_array(pt, shape=(n,)).copy()


def RandomUniform(n):
    pt = _call("RandomUniform", n)
    return np.ctypeslib.as_array(pt, shape=(n,)).copy()


def RandomNormal(n, mean, stddev):
    pt = _call("RandomNormal", n, mean, stddev)
    return np.ctypeslib.as_array(pt, shape=(n,)).copy()


def RandomNormalClipped(n, mean, stddev, vmin, vmax, vstep=0):
    pt = _call("RandomNormalClipped", n, mean, stddev, vmin, vmax, vstep)
    return np.ctypeslib.as_array(pt, shape=(n,)).copy()


# ---------------------------------------------------------------------
# connections
def SingleConnect(i_source, i_target, i_port, weight, delay):
    "Connect two nodes"
    return _call("Connect", i_source, i_target, i_port, 0, weight, delay)


def ConnSpecInit():
    return _call("ConnSpecInit")


def SetConnSpecParam(param_name, val):
    return _call("SetConnSpecParam", _b(param_name), val)


def ConnSpecIsParam(param_name):
    return _call("ConnSpecIsParam", _b(param_name)) != 0


def SynSpecInit():
    return _call("SynSpecInit")


def SetSynSpecIntParam(param_name, val):
    return _call("SetSynSpecIntParam", _b(param_name), val)


def SetSynSpecFloatParam(param_name, val):
    return _call("SetSynSpecFloatParam", _b(param_name), val)


def SetSynSpecFloatPtParam(param_name, arr):
    """Pointer parameter (weight_array, delay_array); the caller keeps the
    array alive until Connect has returned (pythonlib/neurongpu.py:1424-1433).
    """
    if isinstance(arr, (list, tuple)):
        arr = (ctypes.c_float * len(arr))(*arr)
    if isinstance(arr, np.ndarray):
        arr = arr.ctypes.data_as(c_void_p)
    return _call("SetSynSpecFloatPtParam", _b(param_name), ctypes.cast(arr, c_void_p))


def SynSpecIsIntParam(param_name):
    return _call("SynSpecIsIntParam", _b(param_name)) != 0


def SynSpecIsFloatParam(param_name):
    return _call("SynSpecIsFloatParam", _b(param_name)) != 0


def SynSpecIsFloatPtParam(param_name):
    return _call("SynSpecIsFloatPtParam", _b(param_name)) != 0


def SetSynParamFromArray(param_name, par_dict, array_size):
    """Set <param_name>_array of the synapse spec from an array or
    distribution dictionary (pythonlib/neurongpu.py:1533-1541).  The array
    is returned so the caller can keep it alive until Connect."""
    arr_name = param_name + "_array"
    if not SynSpecIsFloatPtParam(arr_name):
        raise ValueError("Synapse parameter cannot be set by"
                         " arrays or distributions")
    arr = DictToArray(par_dict, array_size)
    SetSynSpecFloatPtParam(arr_name, arr)
    return arr


def RuleArraySize(conn_dict, source, target):
    rule = conn_dict["rule"]
    if rule == "one_to_one":
        return len(source)
    if rule == "all_to_all":
        return len(source) * len(target)
    if rule == "fixed_total_number":
        return conn_dict["total_num"]
    if rule == "fixed_indegree":
        return len(target) * conn_dict["indegree"]
    if rule == "fixed_outdegree":
        return len(source) * conn_dict["outdegree"]
    raise ValueError("Unknown number of connections for this rule")


_DIST_CODE = {"uniform": 1, "normal": 2, "normal_clipped": 3}


def _set_syn_distribution(name, d):
    """A weight or delay distribution of a synapse dictionary ({"distribution":
    "uniform" | "normal" | "normal_clipped", "mu", "sigma", "low", "high",
    "step"}) handed to the library, which draws one value per connection
    (on the device for rules generated there) instead of an array drawn
    here.  Returns False for an "array" dictionary."""
    dist = d.get("distribution")
    if "array" in d or dist is None:
        return False
    if dist not in _DIST_CODE:
        raise ValueError("Unknown distribution")
    SetSynSpecIntParam(name + "_distribution", _DIST_CODE[dist])
    for k, v in d.items():
        if k == "distribution":
            continue
        if k not in ("mu", "sigma", "low", "high", "step"):
            raise ValueError("Unknown parameter name in dictionary")
        SetSynSpecFloatParam(name + "_" + k, v)
    return True


def _connect_specs(source, target, conn_dict, syn_dict, keep):
    ConnSpecInit()
    SynSpecInit()
    for k, v in conn_dict.items():
        if k == "rule":
            if v not in conn_rule_name:
                raise ValueError("Unknown connection rule")
            SetConnSpecParam(k, conn_rule_name.index(v))
        elif ConnSpecIsParam(k):
            SetConnSpecParam(k, v)
        else:
            raise ValueError("Unknown connection parameter")
    array_size = RuleArraySize(conn_dict, source, target)
    for k, v in syn_dict.items():
        if SynSpecIsIntParam(k):
            if k == "synapse_group" and isinstance(v, SynGroup):
                v = v.i_syn_group
            SetSynSpecIntParam(k, v)
        elif SynSpecIsFloatParam(k):
            if isinstance(v, dict) and _set_syn_distribution(k, v):
                pass
            elif isinstance(v, dict):
                if not SynSpecIsFloatPtParam(k + "_array"):
                    raise ValueError("Synapse parameter cannot be set by"
                                     " arrays or distributions")
                arr = DictToArray(v, array_size)
                keep.append(arr)
                SetSynSpecFloatPtParam(k + "_array", arr)
            else:
                SetSynSpecFloatParam(k, v)
        elif SynSpecIsFloatPtParam(k):
            if isinstance(v, (list, tuple)):
                v = (ctypes.c_float * len(v))(*v)
            elif isinstance(v, np.ndarray):
                v = np.ascontiguousarray(v, dtype=np.float32)
            keep.append(v)
            SetSynSpecFloatPtParam(k, v)
        else:
            raise ValueError("Unknown synapse parameter")


def Connect(source, target, conn_dict, syn_dict):
    "Connect two node groups"
    if not isinstance(source, (list, tuple, NodeSeq, np.ndarray)):
        raise ValueError("Unknown source type")
    if not isinstance(target, (list, tuple, NodeSeq, np.ndarray)):
        raise ValueError("Unknown target type")
    gc.disable()
    keep = []
    try:
        _connect_specs(source, target, conn_dict, syn_dict, keep)
        if isinstance(source, NodeSeq) and isinstance(target, NodeSeq):
            ret = _C["ConnectSeqSeq"](source.i0, source.n, target.i0, target.n)
        else:
            if not isinstance(source, NodeSeq):
                sa, sp = _int_array(source)
            if not isinstance(target, NodeSeq):
                ta, tp = _int_array(target)
            if isinstance(source, NodeSeq):
                ret = _C["ConnectSeqGroup"](source.i0, source.n, tp, ta.size)
            elif isinstance(target, NodeSeq):
                ret = _C["ConnectGroupSeq"](sp, sa.size, target.i0, target.n)
            else:
                ret = _C["ConnectGroupGroup"](sp, sa.size, tp, ta.size)
        return _check(ret)
    finally:
        gc.enable()


def RemoteConnect(i_source_host, source, i_target_host, target, conn_dict,
                  syn_dict):
    "Connect node groups of different MPI hosts (host 0 only here)"
    keep = []
    _connect_specs(source, target, conn_dict, syn_dict, keep)
    if isinstance(source, NodeSeq) and isinstance(target, NodeSeq):
        ret = _C["RemoteConnectSeqSeq"](i_source_host, source.i0, source.n,
                                        i_target_host, target.i0, target.n)
    else:
        if not isinstance(source, NodeSeq):
            sa, sp = _int_array(source)
        if not isinstance(target, NodeSeq):
            ta, tp = _int_array(target)
        if isinstance(source, NodeSeq):
            ret = _C["RemoteConnectSeqGroup"](i_source_host, source.i0,
                                              source.n, i_target_host, tp,
                                              ta.size)
        elif isinstance(target, NodeSeq):
            ret = _C["RemoteConnectGroupSeq"](i_source_host, sp, sa.size,
                                              i_target_host, target.i0,
                                              target.n)
        else:
            ret = _C["RemoteConnectGroupGroup"](i_source_host, sp, sa.size,
                                                i_target_host, tp, ta.size)
    return _check(ret)


def GetConnections(source=None, target=None, syn_group=-1):
    "Get connections between two node groups"
    if source is None:
        source = NodeSeq(None)
    if target is None:
        target = NodeSeq(None)
    if isinstance(source, int):
        source = [source]
    if isinstance(target, int):
        target = [target]
    n = c_int(0)
    if isinstance(source, NodeSeq) and isinstance(target, NodeSeq):
        arr = _C["GetSeqSeqConnections"](source.i0, source.n, target.i0,
                                         target.n, syn_group, ctypes.byref(n))
    else:
        if not isinstance(source, NodeSeq):
            sa, sp = _int_array(source)
        if not isinstance(target, NodeSeq):
            ta, tp = _int_array(target)
        if isinstance(source, NodeSeq):
            arr = _C["GetSeqGroupConnections"](source.i0, source.n, tp,
                                               ta.size, syn_group,
                                               ctypes.byref(n))
        elif isinstance(target, NodeSeq):
            arr = _C["GetGroupSeqConnections"](sp, sa.size, target.i0,
                                               target.n, syn_group,
                                               ctypes.byref(n))
        else:
            arr = _C["GetGroupGroupConnections"](sp, sa.size, tp, ta.size,
                                                 syn_group, ctypes.byref(n))
    _check()
    return [ConnectionId(arr[3 * i], arr[3 * i + 1], arr[3 * i + 2])
            for i in range(n.value)]


def GetConnectionStatus(conn_id):
    i_target = c_int(0)
    port = ctypes.c_ubyte(0)
    syn = ctypes.c_ubyte(0)
    delay = c_float(0.0)
    weight = c_float(0.0)
    _call("GetConnectionStatus", conn_id.i_source, conn_id.i_group,
          conn_id.i_conn, ctypes.byref(i_target), ctypes.byref(port),
          ctypes.byref(syn), ctypes.byref(delay), ctypes.byref(weight))
    return {"source": conn_id.i_source, "target": i_target.value,
            "port": port.value, "syn": syn.value, "delay": delay.value,
            "weight": weight.value}


def GetStatus(gen_object, var_key=None):
    "Get neuron group, connection or synapse group status"
    if isinstance(gen_object, SynGroup):
        return GetSynGroupStatus(gen_object, var_key)
    if isinstance(gen_object, NodeSeq):
        gen_object = gen_object.ToList()
    if isinstance(gen_object, (list, tuple)):
        return [GetStatus(g, var_key) for g in gen_object]
    if isinstance(var_key, (list, tuple)):
        return [GetStatus(gen_object, k) for k in var_key]
    if var_key is None:
        if isinstance(gen_object, ConnectionId):
            return GetConnectionStatus(gen_object)
        if isinstance(gen_object, int):
            i = gen_object
            names = (GetIntVarNames(i) + GetScalVarNames(i) +
                     GetScalParamNames(i) + GetPortVarNames(i) +
                     GetPortParamNames(i) + GetArrayVarNames(i) +
                     GetArrayParamNames(i) + GetGroupParamNames(i))
            return {k: GetStatus(i, k) for k in names}
        raise ValueError("Unknown object type in GetStatus")
    if isinstance(var_key, (str, bytes)):
        if isinstance(gen_object, ConnectionId):
            return GetConnectionStatus(gen_object)[var_key]
        if isinstance(gen_object, int):
            return GetNeuronStatus([gen_object], var_key)[0]
        raise ValueError("Unknown object type in GetStatus")
    raise ValueError("Unknown key type in GetStatus", type(var_key))


# ---------------------------------------------------------------------
# synapse groups
def CreateSynGroup(model_name, status_dict=None):
    "Create a synapse group"
    if isinstance(status_dict, dict):
        sg = CreateSynGroup(model_name)
        SetStatus(sg, status_dict)
        return sg
    elif status_dict is not None:
        raise ValueError("Wrong argument in CreateSynGroup")
    return SynGroup(_call("CreateSynGroup", _b(model_name)))


def GetSynGroupNParam(syn_group):
    return _call("GetSynGroupNParam", syn_group.i_syn_group)


def GetSynGroupParamNames(syn_group):
    return _names("GetSynGroupParamNames", syn_group.i_syn_group)


def IsSynGroupParam(syn_group, param_name):
    return _call("IsSynGroupParam", syn_group.i_syn_group,
                 _b(param_name)) != 0


def GetSynGroupParam(syn_group, param_name):
    return _call("GetSynGroupParam", syn_group.i_syn_group, _b(param_name))


def SetSynGroupParam(syn_group, param_name, val):
    return _call("SetSynGroupParam", syn_group.i_syn_group, _b(param_name),
                 val)


def GetSynGroupStatus(syn_group, var_key=None):
    "Get synapse group status"
    if not isinstance(syn_group, SynGroup):
        raise ValueError("Wrong argument type in GetSynGroupStatus")
    if isinstance(var_key, (list, tuple)):
        return [GetSynGroupStatus(syn_group, k) for k in var_key]
    if var_key is None:
        return {k: GetSynGroupParam(syn_group, k)
                for k in GetSynGroupParamNames(syn_group)}
    if isinstance(var_key, (str, bytes)):
        return GetSynGroupParam(syn_group, var_key)
    raise ValueError("Unknown key type in GetSynGroupStatus", type(var_key))


def SetSynGroupStatus(syn_group, params, val=None):
    "Set synapse group parameters using dictionaries"
    if not isinstance(syn_group, SynGroup):
        raise ValueError("Wrong argument type in SetSynGroupStatus")
    if isinstance(params, dict) and val is None:
        for k in params:
            SetSynGroupStatus(syn_group, k, params[k])
    elif isinstance(params, str):
        return SetSynGroupParam(syn_group, params, val)
    else:
        raise ValueError("Wrong argument in SetSynGroupStatus")
    _check()


# ---------------------------------------------------------------------
# spikes
def ActivateSpikeCount(nodes):
    "Activate spike count for node group"
    if not isinstance(nodes, NodeSeq):
        raise ValueError("Argument type of ActivateSpikeCount must be NodeSeq")
    return _call("ActivateSpikeCount", nodes.i0, nodes.n)


def ActivateRecSpikeTimes(nodes, max_n_rec_spike_times):
    "Activate spike time recording for node group"
    if not isinstance(nodes, NodeSeq):
        raise ValueError("Argument type of ActivateRecSpikeTimes must be "
                         "NodeSeq")
    return _call("ActivateRecSpikeTimes", nodes.i0, nodes.n,
                 max_n_rec_spike_times)


def GetNRecSpikeTimes(i_node):
    return _call("GetNRecSpikeTimes", i_node)


def GetRecSpikeTimes(i_node):
    "Get recorded spike times for node"
    pt = _call("GetRecSpikeTimes", i_node)
    return [pt[i] for i in range(GetNRecSpikeTimes(i_node))]


def PushSpikesToNodes(nodes):
    a, p = _int_array(nodes)
    return _call("PushSpikesToNodes", a.size, p)


def GetExtNeuronInputSpikes(include_zeros=False):
    """(node, port, height) lists of the ext_neuron ports whose
    port_value is non-zero (all ports with include_zeros), through
    NeuronGPU_GetExtNeuronInputSpikes (src/neurongpu_C.h:288-289)."""
    n = ctypes.c_int(0)
    pn = ctypes.POINTER(ctypes.c_int)()
    pp = ctypes.POINTER(ctypes.c_int)()
    ph = ctypes.POINTER(ctypes.c_float)()
    _call("GetExtNeuronInputSpikes", ctypes.byref(n), ctypes.byref(pn),
          ctypes.byref(pp), ctypes.byref(ph), 1 if include_zeros else 0)
    k = n.value
    return ([pn[i] for i in range(k)], [pp[i] for i in range(k)],
            [ph[i] for i in range(k)])


# ---------------------------------------------------------------------
# kernel parameters
def GetNFloatParam():
    return _call("GetNFloatParam")


def GetFloatParamNames():
    return _names("GetFloatParamNames")


def IsFloatParam(param_name):
    return _call("IsFloatParam", _b(param_name)) != 0


def GetFloatParam(param_name):
    return _call("GetFloatParam", _b(param_name))


def SetFloatParam(param_name, val):
    return _call("SetFloatParam", _b(param_name), val)


def GetNIntParam():
    return _call("GetNIntParam")


def GetIntParamNames():
    return _names("GetIntParamNames")


def IsIntParam(param_name):
    return _call("IsIntParam", _b(param_name)) != 0


def GetIntParam(param_name):
    return _call("GetIntParam", _b(param_name))


def SetIntParam(param_name, val):
    return _call("SetIntParam", _b(param_name), val)


def GetKernelStatus(var_key=None):
    "Get kernel status"
    if isinstance(var_key, (list, tuple)):
        return [GetKernelStatus(k) for k in var_key]
    if var_key is None:
        return {k: GetKernelStatus(k)
                for k in GetFloatParamNames() + GetIntParamNames()}
    if isinstance(var_key, (str, bytes)):
        if IsFloatParam(var_key):
            return GetFloatParam(var_key)
        if IsIntParam(var_key):
            return GetIntParam(var_key)
        raise ValueError("Unknown parameter in GetKernelStatus", var_key)
    raise ValueError("Unknown key type in GetKernelStatus", type(var_key))


def SetKernelStatus(params, val=None):
    "Set kernel parameters using dictionaries"
    if isinstance(params, dict) and val is None:
        for k in params:
            SetKernelStatus(k, params[k])
    elif isinstance(params, str):
        if IsFloatParam(params):
            return SetFloatParam(params, val)
        if IsIntParam(params):
            return SetIntParam(params, val)
        raise ValueError("Unknown parameter in SetKernelStatus", params)
    else:
        raise ValueError("Wrong argument in SetKernelStatus")
    _check()


def RemoteCreate(i_host, model_name, n_node=1, n_ports=1, status_dict=None):
    "Create a remote neuron group (host 0 only here)"
    if isinstance(status_dict, dict):
        r = RemoteCreate(i_host, model_name, n_node, n_ports)
        SetStatus(r, status_dict)
        return r
    elif status_dict is not None:
        raise ValueError("Wrong argument in RemoteCreate")
    i_node = _call("RemoteCreate", i_host, _b(model_name), n_node, n_ports)
    return RemoteNodeSeq(i_host, NodeSeq(i_node, n_node))


# ---------------------------------------------------------------------
# no reference counterpart
def ResetKernel():
    "Discard the network and start from a fresh engine"
    return _C["ResetKernel"]()


def GetHotPathStats():
    "RK5 accepted/rejected step counts and delivered synaptic events"
    a = (ctypes.c_ulonglong * 3)()
    _call("GetHotPathStats", a)
    f = (ctypes.c_ulonglong * 1)()
    _call("GetRk5ForcedAccepts", f)
    return {"rk5_accepted": a[0], "rk5_rejected": a[1], "syn_events": a[2],
            "rk5_forced": f[0]}


def GetRk5Pass1Stats():
    "The first RK5 pass's share of the accepted/rejected counts"
    a = (ctypes.c_ulonglong * 2)()
    _call("GetRk5Pass1Stats", a)
    return {"rk5_pass1_accepted": a[0], "rk5_pass1_rejected": a[1]}


def _node_range(nodes):
    if nodes is None:
        return 0, 0
    if isinstance(nodes, NodeSeq):
        return nodes.i0, nodes.n
    if isinstance(nodes, int):
        return nodes, 1
    raise ValueError("A node range must be a NodeSeq or an int")


def GetConnectionArrays(source=None, target=None):
    """The connections from the node range `source` to the node range
    `target` (None: every node), in the canonical order of GetConnections,
    as numpy arrays: {"source", "target", "port", "syn", "delay",
    "weight"}.  One bulk call instead of one GetConnectionStatus per
    connection."""
    s0, ns = _node_range(source)
    t0, nt = _node_range(target)
    n = _call("GetConnectionArrays", s0, ns, t0, nt, 0, None, None, None,
              None, None, None)
    out = {"source": np.zeros(n, np.int32), "target": np.zeros(n, np.int32),
           "port": np.zeros(n, np.uint8), "syn": np.zeros(n, np.uint8),
           "delay": np.zeros(n, np.float32),
           "weight": np.zeros(n, np.float32)}
    m = _call("GetConnectionArrays", s0, ns, t0, nt, n,
              *[out[k].ctypes.data_as(c_void_p) for k in
                ("source", "target", "port", "syn", "delay", "weight")])
    if m != n:
        raise ValueError("The connections changed during the query")
    return out


def GetConnectionCounts(nodes, port=-1, n_delay=0):
    """Synapse counts of the node range `nodes` (a NodeSeq), counted on the
    device once the network is calibrated: {"in_degree", "out_degree"}
    per node and "delay_count"[d] = synapses from the range with a delay of
    d + 1 time steps (d < n_delay); only synapses on receptor `port`
    (every port if port < 0)."""
    i0, n = _node_range(nodes)
    ind = np.zeros(n, np.int32)
    outd = np.zeros(n, np.int32)
    dc = np.zeros(max(n_delay, 1), np.int64)
    _call("GetConnectionCounts", i0, n, port, ind.ctypes.data_as(c_void_p),
          outd.ctypes.data_as(c_void_p), dc.ctypes.data_as(c_void_p),
          n_delay)
    return {"in_degree": ind, "out_degree": outd,
            "delay_count": dc[:n_delay]}


KERNEL_KINDS = ("aeif_rk5", "iaf_psc_exp", "relay", "spike_generator",
                "delivery", "stdp_post", "record", "aeif_rk5_pass2")


def SetKernelTiming(on):
    "Enable per-kernel HIP-event timing (clears the totals)"
    return _call("SetKernelTiming", 1 if on else 0)


def GetKernelTimes():
    "{kind: (total_ms, launches)} of the timed kernels"
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_longlong * 8)()
    _call("GetKernelTimes", ms, cnt)
    return {KERNEL_KINDS[k]: (ms[k], cnt[k]) for k in range(8) if cnt[k]}


# ---------------------------------------------------------------------
# learned RK5 step-size proposer (no counterpart in the reference)
PROPOSER_N_PARAM = 1313
PROPOSER_SAMPLE = 8


def ProposerLog(on, max_samples=10000000):
    "Log (features, target) samples of second-pass RK5 attempts"
    return _call("ProposerLog", 1 if on else 0, int(max_samples))


def ProposerNSamples():
    return _call("ProposerNSamples")


def ProposerGetSamples(n=None):
    if n is None:
        n = ProposerNSamples()
    out = np.zeros((n, PROPOSER_SAMPLE), dtype=np.float32)
    if n:
        _call("ProposerGetSamples", out.ctypes.data_as(c_void_p), n)
    return out


def ProposerSetSamples(samples):
    a = np.ascontiguousarray(samples, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != PROPOSER_SAMPLE:
        raise ValueError("samples must have shape (n, %d)" % PROPOSER_SAMPLE)
    return _call("ProposerSetSamples", a.ctypes.data_as(c_void_p), a.shape[0])


def ProposerTrain(n_iter, batch=8192, lr=1e-3, adam=True, full_batch=False):
    "Train the proposer; returns the mean squared error of the last batch"
    loss = c_float(0.0)
    _call("ProposerTrain", int(n_iter), int(batch), float(lr),
          1 if adam else 0, 1 if full_batch else 0, ctypes.byref(loss))
    return loss.value


def ProposerGetWeights():
    w = np.zeros(PROPOSER_N_PARAM, dtype=np.float32)
    _call("ProposerGetWeights", w.ctypes.data_as(c_void_p))
    return w


def ProposerSetWeights(w):
    a = np.ascontiguousarray(w, dtype=np.float32)
    if a.size != PROPOSER_N_PARAM:
        raise ValueError("expected %d weights" % PROPOSER_N_PARAM)
    return _call("ProposerSetWeights", a.ctypes.data_as(c_void_p))


def ProposerEval(features):
    "Proposer output (log2 of the step / dt) for features of shape (n, 6)"
    f = np.ascontiguousarray(features, dtype=np.float32)
    out = np.zeros(f.shape[0], dtype=np.float32)
    _call("ProposerEval", f.ctypes.data_as(c_void_p), f.shape[0],
          out.ctypes.data_as(c_void_p))
    return out


def ProposerEvalNoAlloc(features):
    """ProposerEval without device allocation (at most 65536 vectors; the
    proposer's weights must exist already)"""
    f = np.ascontiguousarray(features, dtype=np.float32)
    out = np.zeros(f.shape[0], dtype=np.float32)
    _call("ProposerEvalNoAlloc", f.ctypes.data_as(c_void_p), f.shape[0],
          out.ctypes.data_as(c_void_p))
    return out


def ProposerUse(on):
    "Start second-pass RK5 steps from the proposer's step size"
    return _call("ProposerUse", 1 if on else 0)
