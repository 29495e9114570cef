"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/liboracle_f32.so or liboracle_f64.so (built by
oracle/Makefile) and exposes a small network object whose variable and
parameter layout follows the reference's enums:
  aeif_cond_beta   src/aeif_cond_beta_kernel.h:30-66
  aeif_cond_alpha  src/aeif_cond_alpha_kernel.h:30-66
  iaf_psc_exp      src/iaf_psc_exp.h:32-62
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

AEIF_COND_BETA = 1
AEIF_COND_ALPHA = 2
IAF_PSC_EXP = 3
SPIKE_GENERATOR = 4
POISSON_GENERATOR = 5
PARROT_NEURON = 6
AEIF_PSC_ALPHA = 7
AEIF_PSC_EXP = 8
AEIF_PSC_DELTA = 9
IZH_COND_BETA = 10
POISSON_LEGACY = 11   # nodes of CreatePoissonGenerator
IAF_PSC_EXP_G = 12
IAF_PSC_EXP_HC = 13
IZHIKEVICH = 14
IZH_PSC_EXP = 15
IZH_PSC_EXP_2S = 16
IZH_PSC_EXP_5S = 17
EXT_NEURON = 18

MODEL_IDS = {
    "aeif_cond_beta": AEIF_COND_BETA,
    "aeif_cond_alpha": AEIF_COND_ALPHA,
    "iaf_psc_exp": IAF_PSC_EXP,
    "spike_generator": SPIKE_GENERATOR,
    "poisson_generator": POISSON_GENERATOR,
    "parrot_neuron": PARROT_NEURON,
    "aeif_psc_alpha": AEIF_PSC_ALPHA,
    "aeif_psc_exp": AEIF_PSC_EXP,
    "aeif_psc_delta": AEIF_PSC_DELTA,
    "izhikevich_cond_beta": IZH_COND_BETA,
    # src/user_m1_kernel.h, src/user_m2_kernel.h: aeif_cond_beta renamed
    "user_m1": AEIF_COND_BETA,
    "user_m2": AEIF_COND_BETA,
    "legacy_poisson_generator": POISSON_LEGACY,
    "iaf_psc_exp_g": IAF_PSC_EXP_G,
    "iaf_psc_exp_hc": IAF_PSC_EXP_HC,
    "izhikevich": IZHIKEVICH,
    "izhikevich_psc_exp": IZH_PSC_EXP,
    "izhikevich_psc_exp_2s": IZH_PSC_EXP_2S,
    "izhikevich_psc_exp_5s": IZH_PSC_EXP_5S,
    "ext_neuron": EXT_NEURON,
}

# fixed-step point models: variables, then scalar + group parameters
POINT_VARS = {
    "iaf_psc_exp_g": ["I_syn", "V_m_rel", "refractory_step"],
    "iaf_psc_exp_hc": ["I_syn", "V_m_rel", "refractory_step"],
}
POINT_PARAMS = {
    "iaf_psc_exp_g": ["I_e", "tau_m", "C_m", "E_L", "Theta_rel",
                      "V_reset_rel", "tau_syn", "t_ref"],
    "iaf_psc_exp_hc": ["I_e"],
    "izhikevich": ["I_e", "den_delay", "V_th", "a", "b", "c", "d", "t_ref"],
}
for _m in ("izhikevich", "izhikevich_psc_exp", "izhikevich_psc_exp_2s",
           "izhikevich_psc_exp_5s"):
    POINT_VARS[_m] = ["I_syn", "V_m", "u", "refractory_step"]
    if _m != "izhikevich":
        POINT_PARAMS[_m] = ["I_e", "den_delay", "V_th", "a", "b", "c", "d",
                            "tau_syn", "t_ref"]

AEIF_SCAL_PARAM = ["V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w",
                   "I_e", "V_peak", "V_reset", "t_ref", "refractory_step",
                   "den_delay"]
BETA_PORT_PARAM = ["E_rev", "tau_rise", "tau_decay", "g0"]
ALPHA_PORT_PARAM = ["E_rev", "tau_syn", "g0"]
AEIF_SCAL_VAR = ["V_m", "w"]
AEIF_PORT_VAR = ["g", "g1"]
IZH_SCAL_PARAM = ["V_th", "a", "b", "c", "d", "I_e", "t_ref",
                  "refractory_step", "den_delay"]
IAF_VAR = ["I_syn_ex", "I_syn_in", "V_m_rel", "refractory_step"]
IAF_PARAM = ["tau_m", "C_m", "E_L", "I_e", "Theta_rel", "V_reset_rel",
             "tau_ex", "tau_in", "t_ref", "den_delay", "P20", "P11ex",
             "P11in", "P21ex", "P21in", "P22"]
STDP_PARAM = ["tau_plus", "tau_minus", "lambda", "alpha", "mu_plus",
              "mu_minus", "Wmax"]
STDP_DEFAULT = {"tau_plus": 20.0, "tau_minus": 20.0, "lambda": 1.0e-4,
                "alpha": 1.0, "mu_plus": 1.0, "mu_minus": 1.0,
                "Wmax": 100.0}  # src/stdp.cu:59-65


def ensure_built():
    for name in ("liboracle_f32.so", "liboracle_f64.so",
                 "liboracle_f32a64.so"):
        if not os.path.exists(os.path.join(HERE, name)):
            subprocess.check_call(["make", "-s", "-C", HERE])
            break


_LIBS = {}


def lib(prec="f32"):
    if prec not in _LIBS:
        ensure_built()
        L = ctypes.CDLL(os.path.join(HERE, "liboracle_%s.so" % prec))
        vp, i, d, f, l = (ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                          ctypes.c_float, ctypes.c_long)
        L.orc_net_create.restype = vp
        L.orc_net_create.argtypes = [d, d, ctypes.c_uint64]
        L.orc_net_free.argtypes = [vp]
        L.orc_net_add_group.argtypes = [vp, i, i, i]
        L.orc_net_n_var.argtypes = [vp, i]
        L.orc_net_n_param.argtypes = [vp, i]
        for fn in ("orc_net_set_var", "orc_net_get_var", "orc_net_set_param",
                   "orc_net_get_param"):
            getattr(L, fn).argtypes = [vp, i, i, vp]
        L.orc_net_set_rk5.argtypes = [vp, i, d, d, d]
        L.orc_net_spike_gen.argtypes = [vp, i, i, vp, vp]
        L.orc_net_poisson.argtypes = [vp, i, f, f, f, f]
        L.orc_net_legacy_rate.restype = i
        L.orc_net_legacy_rate.argtypes = [vp, i, f]
        L.orc_net_syn_group_stdp.argtypes = [vp, vp]
        L.orc_net_connect.argtypes = [vp, i, vp, vp, vp, vp, vp, vp]
        L.orc_net_record.argtypes = [vp, i, i, i]
        L.orc_net_calibrate.argtypes = [vp]
        L.orc_net_run.argtypes = [vp, l, i]
        L.orc_net_n_spikes.restype = l
        L.orc_net_n_spikes.argtypes = [vp]
        L.orc_net_get_spikes.argtypes = [vp, vp, vp, vp]
        L.orc_net_get_spike_heights.argtypes = [vp, vp]
        L.orc_net_n_records.argtypes = [vp]
        L.orc_net_n_record_rows.restype = l
        L.orc_net_n_record_rows.argtypes = [vp]
        L.orc_net_get_records.argtypes = [vp, vp]
        L.orc_net_n_conn.restype = l
        L.orc_net_n_conn.argtypes = [vp]
        L.orc_net_get_conn.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_net_counters.argtypes = [vp, vp]
        L.orc_expf.restype = f
        L.orc_expf.argtypes = [f]
        L.orc_logf.restype = f
        L.orc_logf.argtypes = [f]
        L.orc_pow_ctrl.restype = f
        L.orc_pow_ctrl.argtypes = [f, f]
        L.orc_stdp_update.restype = f
        L.orc_stdp_update.argtypes = [f, f, vp]
        L.orc_beta_g0.restype = d
        L.orc_beta_g0.argtypes = [d, d]
        L.orc_iaf_propagators.argtypes = [d, d, d, d, d, vp]
        L.orc_rk5_linear_fixed.restype = d
        L.orc_rk5_linear_fixed.argtypes = [d, d, d, i]
        L.orc_rk5_linear_adaptive.restype = d
        L.orc_rk5_linear_adaptive.argtypes = [d, d, d, d, d, d, vp, vp]
        L.orc_poisson_table.restype = i
        L.orc_poisson_table.argtypes = [d, vp, i]
        L.orc_poisson_sample.restype = i
        L.orc_poisson_sample.argtypes = [vp, ctypes.c_uint32]
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_threefry2x32_20.argtypes = [vp, vp, vp]
        _LIBS[prec] = L
    return _LIBS[prec]


def philox4x32_10(ctr, key):
    L = lib()
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    L.orc_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def threefry2x32_20(ctr, key):
    L = lib()
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(2, dtype=np.uint32)
    L.orc_threefry2x32_20(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def poisson_table(lam, max_len=256):
    L = lib()
    tab = np.zeros(max_len, dtype=np.uint32)
    n = L.orc_poisson_table(float(lam), tab.ctypes.data, max_len)
    return tab[:n] if n > 0 else None


def _ptr(a):
    return a.ctypes.data


class Group:
    def __init__(self, net, gid, model, i0, n, n_port):
        self.net, self.gid, self.model = net, gid, model
        self.i0, self.n, self.n_port = i0, n, n_port

    def var_names(self):
        if self.model in ("aeif_cond_beta", "aeif_cond_alpha", "user_m1",
                          "user_m2"):
            return AEIF_SCAL_VAR, AEIF_PORT_VAR
        if self.model == "aeif_psc_alpha":
            return AEIF_SCAL_VAR, ["I_syn", "I1_syn"]
        if self.model == "aeif_psc_exp":
            return AEIF_SCAL_VAR, ["I_syn"]
        if self.model == "aeif_psc_delta":
            return AEIF_SCAL_VAR, []
        if self.model == "izhikevich_cond_beta":
            return ["V_m", "u"], ["g", "g1"]
        if self.model == "iaf_psc_exp":
            return IAF_VAR, []
        if self.model == "parrot_neuron":
            return ["input_spike_height", "dummy_input", "V"], []
        if self.model in POINT_VARS:
            return POINT_VARS[self.model], []
        if self.model == "ext_neuron":
            return [], ["port_input", "port_value"]
        return [], []

    def param_names(self):
        if self.model in ("aeif_cond_beta", "user_m1", "user_m2"):
            return AEIF_SCAL_PARAM, BETA_PORT_PARAM
        if self.model == "aeif_psc_alpha":
            return AEIF_SCAL_PARAM, ["tau_syn", "I0"]
        if self.model == "aeif_psc_exp":
            return AEIF_SCAL_PARAM, ["tau_syn"]
        if self.model == "aeif_psc_delta":
            return AEIF_SCAL_PARAM, []
        if self.model == "izhikevich_cond_beta":
            return IZH_SCAL_PARAM, BETA_PORT_PARAM
        if self.model == "aeif_cond_alpha":
            return AEIF_SCAL_PARAM, ALPHA_PORT_PARAM
        if self.model == "iaf_psc_exp":
            return IAF_PARAM, []
        if self.model == "parrot_neuron":
            return ["hold_spike_height", "den_delay"], []
        if self.model in POINT_PARAMS:
            return POINT_PARAMS[self.model], []
        if self.model == "ext_neuron":
            return ["den_delay"], ["port_weight"]
        return [], []

    @staticmethod
    def _index(names, name, port):
        scal, portn = names
        if name in scal:
            return scal.index(name)
        if name in portn:
            return len(scal) + len(portn) * port + portn.index(name)
        raise KeyError(name)

    def var_index(self, name, port=0):
        return self._index(self.var_names(), name, port)

    def param_index(self, name, port=0):
        return self._index(self.param_names(), name, port)

    def _set(self, fn, idx, vals):
        a = np.empty(self.n, dtype=self.net.dtype)
        a[:] = vals
        getattr(self.net.L, fn)(self.net.h, self.gid, idx, _ptr(a))

    def _get(self, fn, idx):
        a = np.empty(self.n, dtype=self.net.dtype)
        getattr(self.net.L, fn)(self.net.h, self.gid, idx, _ptr(a))
        return a

    def set(self, name, vals, port=None):
        """Set a variable or parameter; port-array values set every port."""
        ports = [port] if port is not None else None
        scal_v, port_v = self.var_names()
        scal_p, port_p = self.param_names()
        if name in port_v or name in port_p:
            if ports is None:
                vals = np.asarray(vals, dtype=np.float64)
                if vals.ndim == 1 and vals.size == self.n_port:
                    for q in range(self.n_port):
                        self.set(name, vals[q], q)
                    return
                ports = range(self.n_port)
            for q in ports:
                if name in port_v:
                    self._set("orc_net_set_var", self.var_index(name, q), vals)
                else:
                    self._set("orc_net_set_param", self.param_index(name, q),
                              vals)
            return
        if name in scal_v:
            self._set("orc_net_set_var", self.var_index(name), vals)
        elif name in scal_p:
            self._set("orc_net_set_param", self.param_index(name), vals)
        else:
            raise KeyError(name)

    def get(self, name, port=0):
        scal_v, port_v = self.var_names()
        if name in scal_v or name in port_v:
            return self._get("orc_net_get_var", self.var_index(name, port))
        return self._get("orc_net_get_param", self.param_index(name, port))


class Net:
    """One simulated network in the oracle.

    dt and t0 in ms; seed keys the Poisson Threefry streams."""

    def __init__(self, dt=0.1, t0=0.0, seed=54321, prec="f32"):
        self.L = lib(prec)
        self.prec = prec
        # state precision (f32a64: float state, double input accumulators)
        self.dtype = np.float64 if prec == "f64" else np.float32
        self.h = self.L.orc_net_create(dt, t0, seed)
        self.groups = []
        self.dt = dt

    def __del__(self):
        try:
            self.L.orc_net_free(self.h)
        except Exception:
            pass

    def create(self, model, n=1, n_port=1):
        gid = len(self.groups)
        i0 = self.L.orc_net_add_group(self.h, MODEL_IDS[model], n, n_port)
        if model in ("iaf_psc_exp", "parrot_neuron"):
            n_port = 2
        if model in POINT_VARS:
            n_port = 1
        if model in ("spike_generator", "poisson_generator",
                     "legacy_poisson_generator"):
            n_port = 0
        g = Group(self, gid, model, i0, n, n_port)
        self.groups.append(g)
        return g

    def set_rk5(self, g, h_min_rel=1e-3, h0_rel=1e-2, eps=1e-6):
        self.L.orc_net_set_rk5(self.h, g.gid, h_min_rel, h0_rel, eps)

    def spike_generator_times(self, node, times, heights=None):
        t = np.ascontiguousarray(times, dtype=np.float32)
        hh = (np.ascontiguousarray(heights, dtype=np.float32)
              if heights is not None else np.ones(len(t), dtype=np.float32))
        self.L.orc_net_spike_gen(self.h, node, len(t), _ptr(t), _ptr(hh))

    def poisson(self, node, rate, origin=0.0, start=0.0, stop=1.0e30):
        self.L.orc_net_poisson(self.h, node, rate, origin, start, stop)

    def legacy_rate(self, g, rate):
        """Rate (Hz) of a legacy_poisson_generator group."""
        if self.L.orc_net_legacy_rate(self.h, g.gid, rate) != 0:
            raise ValueError("legacy Poisson rate out of range")

    def stdp_group(self, **params):
        p = dict(STDP_DEFAULT)
        p.update(params)
        a = np.array([p[k] for k in STDP_PARAM], dtype=np.float32)
        return self.L.orc_net_syn_group_stdp(self.h, _ptr(a))

    def connect(self, src, tgt, weight, delay, port=0, syn_group=0):
        src = np.ascontiguousarray(src, dtype=np.int32)
        n = src.size
        tgt = np.ascontiguousarray(np.broadcast_to(tgt, n), dtype=np.int32)
        w = np.ascontiguousarray(np.broadcast_to(weight, n), dtype=np.float32)
        d = np.ascontiguousarray(np.broadcast_to(delay, n), dtype=np.float32)
        p = np.ascontiguousarray(np.broadcast_to(port, n), dtype=np.uint8)
        s = np.ascontiguousarray(np.broadcast_to(syn_group, n), dtype=np.uint8)
        self.L.orc_net_connect(self.h, n, _ptr(src), _ptr(tgt), _ptr(p),
                               _ptr(s), _ptr(w), _ptr(d))

    def record(self, node, name, port=0):
        g = self.group_of(node)
        scal_v, port_v = g.var_names()
        if name in scal_v or name in port_v:
            return self.L.orc_net_record(self.h, node, g.var_index(name, port),
                                         0)
        return self.L.orc_net_record(self.h, node, g.param_index(name, port), 1)

    def group_of(self, node):
        for g in self.groups:
            if g.i0 <= node < g.i0 + g.n:
                return g
        raise KeyError(node)

    def calibrate(self):
        r = self.L.orc_net_calibrate(self.h)
        if r != 0:
            raise ValueError("oracle calibration failed (%d)" % r)

    def run(self, n_steps, n_threads=1):
        r = self.L.orc_net_run(self.h, int(n_steps), int(n_threads))
        if r != 0:
            raise ValueError("oracle run failed (%d)" % r)

    def spikes(self):
        n = self.L.orc_net_n_spikes(self.h)
        st = np.zeros(n, dtype=np.int64)
        nd = np.zeros(n, dtype=np.int32)
        ct = np.zeros(n, dtype=np.int32)
        if n:
            self.L.orc_net_get_spikes(self.h, _ptr(st), _ptr(nd), _ptr(ct))
        return st, nd, ct

    def spike_heights(self):
        """Summed height of each entry of spikes()."""
        n = self.L.orc_net_n_spikes(self.h)
        h = np.zeros(n, dtype=np.float64)
        if n:
            self.L.orc_net_get_spike_heights(self.h, _ptr(h))
        return h

    def records(self):
        nr = self.L.orc_net_n_records(self.h)
        rows = self.L.orc_net_n_record_rows(self.h)
        out = np.zeros((rows, 1 + nr), dtype=np.float64)
        if rows:
            self.L.orc_net_get_records(self.h, _ptr(out))
        return out

    def connections(self):
        n = self.L.orc_net_n_conn(self.h)
        src = np.zeros(n, np.int32)
        tgt = np.zeros(n, np.int32)
        port = np.zeros(n, np.int32)
        d = np.zeros(n, np.int32)
        w = np.zeros(n, np.float32)
        if n:
            self.L.orc_net_get_conn(self.h, _ptr(src), _ptr(tgt), _ptr(port),
                                    _ptr(d), _ptr(w))
        return src, tgt, port, d, w

    def counters(self):
        a = np.zeros(4, dtype=np.int64)
        self.L.orc_net_counters(self.h, _ptr(a))
        return dict(rk5_accepted=int(a[0]), rk5_rejected=int(a[1]),
                    events=int(a[2]), poisson_draws=int(a[3]))
