"""Builds one network description in both the engine (through the Python
bindings, i.e. the C ABI) and the CPU oracle, so GPU and oracle runs see the
same nodes, parameters, synapses and Poisson streams.

A description is a dict:
  dt, seed, groups: [ {model, n, n_port, params{name: scalar|list|array},
                       vars{...}, rk5: {h_min_rel, h0_rel, rk5_tol}} ],
          (model "legacy_poisson_generator" with "rate": the group of
          CreatePoissonGenerator)
  conns: [ {src, tgt, weight, delay, port, syn_group} ]   (arrays, one entry
          per synapse, connected one_to_one in the given order)
  poisson: [ {node, rate, origin, start, stop} ]
  spike_gen: [ {node, times, heights} ]
  stdp: [ {param dict} ]       (syn groups 1..)
  records: [ (node, var, port) ]
"""
import numpy as np

from neurongpu_amd.workloads import build_from_spec as build_engine  # noqa: F401
from oracle import pyoracle as po


def build_oracle(spec, prec="f32"):
    net = po.Net(dt=spec.get("dt", 0.1), seed=spec.get("seed", 54321),
                 prec=prec)
    groups = []
    for g in spec["groups"]:
        og = net.create(g["model"], g["n"], g.get("n_port", 1))
        groups.append(og)
        if g["model"] == "legacy_poisson_generator":
            net.legacy_rate(og, g["rate"])
        for k, v in list(g.get("params", {}).items()) + list(
                g.get("vars", {}).items()):
            og.set(k, v)
        if "rk5" in g:
            r = {"h_min_rel": 1e-3, "h0_rel": 1e-2, "rk5_tol": 1e-6}
            r.update(g["rk5"])
            net.set_rk5(og, np.float32(r["h_min_rel"]),
                        np.float32(r["h0_rel"]), np.float32(r["rk5_tol"]))
    for s in spec.get("stdp", []):
        net.stdp_group(**s)
    for p in spec.get("poisson", []):
        net.poisson(p["node"], p["rate"], p.get("origin", 0.0),
                    p.get("start", 0.0), p.get("stop", 1.0e30))
    for sg in spec.get("spike_gen", []):
        net.spike_generator_times(sg["node"], sg["times"], sg.get("heights"))
    for c in spec.get("conns", []):
        net.connect(c["src"], c["tgt"], c["weight"], c["delay"],
                    c.get("port", 0), c.get("syn_group", 0))
    for (node, var, port) in spec.get("records", []):
        net.record(node, var, port)
    return net, groups


def engine_spikes(ngpu, group):
    """(node, time) pairs from ActivateRecSpikeTimes of a group."""
    out = []
    for i in range(group.n):
        for t in ngpu.GetRecSpikeTimes(group[i]):
            out.append((group[i], t))
    return out


def oracle_spikes(net, dt, node_lo, node_hi):
    st, nd, ct = net.spikes()
    out = []
    for s, n, c in zip(st, nd, ct):
        if node_lo <= n < node_hi:
            # engine time stamp: t0 + (double)(float)dt * (step + 1)
            t = np.float32(np.float64(np.float32(dt)) * (s + 1))
            for _ in range(c):
                out.append((int(n), float(t)))
    return out


def c0_spec(n=1000, k_e=80, k_i=20, w_e=0.05, w_i=0.25, w_p=0.1,
            rate=8000.0, seed=1234, tol=1e-3, n_port=2):
    """Plain-method network of the small configuration: aeif_cond_beta, 2
    receptor ports (E_rev 0/-85 mV, tau_rise 0.2/0.5 ms, tau_decay 2/10 ms),
    fixed indegree k_e + k_i without replacement, integer delays 1-20
    steps, own Poisson generator per neuron."""
    rng = np.random.default_rng(seed)
    n_e = int(round(n * 0.8))
    v0 = rng.uniform(-70.0, -55.0, n).astype(np.float32)
    src, tgt, w, d, port = [], [], [], [], []
    for j in range(n):
        se = rng.choice(n_e, size=k_e, replace=False)
        si = n_e + rng.choice(n - n_e, size=k_i, replace=False)
        src += [se, si]
        tgt += [np.full(k_e, j), np.full(k_i, j)]
        w += [np.full(k_e, w_e), np.full(k_i, w_i)]
        d += [rng.integers(1, 21, k_e), rng.integers(1, 21, k_i)]
        port += [np.zeros(k_e, int), np.ones(k_i, int)]
    src = np.concatenate(src)
    tgt = np.concatenate(tgt)
    w = np.concatenate(w).astype(np.float32)
    d = (np.concatenate(d) * 0.1).astype(np.float32)
    port = np.concatenate(port)
    params = {"C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
              "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
              "V_reset": -60.0, "t_ref": 0.0,
              "E_rev": [0.0, -85.0], "tau_rise": [0.2, 0.5],
              "tau_decay": [2.0, 10.0]}
    spec = {"dt": 0.1, "seed": seed,
            "groups": [{"model": "aeif_cond_beta", "n": n, "n_port": n_port,
                        "params": params, "vars": {"V_m": v0},
                        "rk5": {"rk5_tol": tol}},
                       {"model": "poisson_generator", "n": n}],
            "conns": [], "poisson": []}
    for p in (0, 1):
        m = port == p
        spec["conns"].append({"src": src[m], "tgt": tgt[m], "weight": w[m],
                              "delay": d[m], "port": p})
    spec["conns"].append({"src": np.arange(n, 2 * n), "tgt": np.arange(n),
                          "weight": np.full(n, w_p, np.float32),
                          "delay": np.full(n, 0.1, np.float32), "port": 0})
    spec["poisson"] = [{"node": n + i, "rate": rate} for i in range(n)]
    return spec


def brunel_spec(order=200, seed=1234, w_ex=0.05, w_in=0.35, w_p=0.37,
                rate=20000.0, tol=1e-6):
    """The reference's network-statistics test (python/test/
    test_brunel_list.py:14-85): 4*order excitatory + order inhibitory
    aeif_cond_beta neurons, 2 ports (E_rev 0/-85 mV, tau_rise = tau_decay =
    1 ms), CE = 800 excitatory and CI = 200 inhibitory inputs per neuron
    (fixed indegree), delays normal N(0.5, 0.25) ms clipped to
    [0.1, 1.25], one 20 kHz Poisson generator connected to every neuron
    with weight 0.37 and delay 0.2 ms.  Its expected mean rate is 30.78
    spikes/s (test_brunel_list.py:11)."""
    rng = np.random.default_rng(seed)
    n_e, n_i = 4 * order, order
    n = n_e + n_i
    ce, ci = 800, 200

    def delays(m):
        d = rng.normal(0.5, 0.25, m)
        bad = (d < 0.1) | (d > 1.25)
        while bad.any():
            d[bad] = rng.normal(0.5, 0.25, bad.sum())
            bad = (d < 0.1) | (d > 1.25)
        return d.astype(np.float32)

    src_e = np.concatenate([rng.choice(n_e, ce, replace=False)
                            for _ in range(n)])
    tgt_e = np.repeat(np.arange(n), ce)
    src_i = np.concatenate([n_e + rng.choice(n_i, ci, replace=False)
                            for _ in range(n)])
    tgt_i = np.repeat(np.arange(n), ci)
    params = {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0],
              "tau_decay": [1.0, 1.0]}
    return {"dt": 0.1, "seed": seed,
            "groups": [{"model": "aeif_cond_beta", "n": n, "n_port": 2,
                        "params": params, "rk5": {"rk5_tol": tol}},
                       {"model": "poisson_generator", "n": 1}],
            "conns": [{"src": src_e, "tgt": tgt_e,
                       "weight": np.full(src_e.size, w_ex, np.float32),
                       "delay": delays(src_e.size), "port": 0},
                      {"src": src_i, "tgt": tgt_i,
                       "weight": np.full(src_i.size, w_in, np.float32),
                       "delay": delays(src_i.size), "port": 1},
                      {"src": np.full(n, n), "tgt": np.arange(n),
                       "weight": np.full(n, w_p, np.float32),
                       "delay": np.full(n, 0.2, np.float32), "port": 0}],
            "poisson": [{"node": n, "rate": rate}]}


def legacy_brunel_spec(order=200, seed=4321, w_ex=0.0625, w_in=0.375,
                       w_p=0.375, rate=20000.0, tol=1e-3):
    """The network of the reference's c++/examples/test_error.cpp:24-110
    (the Brunel network of brunel_spec, driven by CreatePoissonGenerator:
    one generator node per neuron, connected one to one with delay 0.2 ms),
    with dyadic weights so that every synaptic sum is exact."""
    spec = brunel_spec(order=order, seed=seed, w_ex=w_ex, w_in=w_in,
                       w_p=w_p, rate=rate, tol=tol)
    n = 5 * order
    spec["groups"][1] = {"model": "legacy_poisson_generator", "n": n,
                         "rate": rate}
    spec["conns"][2] = {"src": n + np.arange(n), "tgt": np.arange(n),
                        "weight": np.full(n, w_p, np.float32),
                        "delay": np.full(n, 0.2, np.float32), "port": 0}
    spec["poisson"] = []
    return spec


def ext_spec(port_weight=None):
    """python/test/test_ext_neuron.py:5-25: a spike generator (50, 100,
    400, 600 ms) into ports 0-2 of an ext_neuron with weights 0.1 / 0.2 /
    0.5 and delays 1 / 50 / 100 ms, plus one parrot neuron per delay as
    the timing reference."""
    ext = {"model": "ext_neuron", "n": 1, "n_port": 3}
    if port_weight is not None:
        ext["params"] = {"port_weight": port_weight}
    spec = {"groups": [ext,
                       {"model": "parrot_neuron", "n": 3,
                        "params": {"hold_spike_height": 1.0}},
                       {"model": "spike_generator", "n": 1}],
            "spike_gen": [{"node": 4, "times": [50.0, 100.0, 400.0, 600.0]}],
            "conns": [], "records": []}
    for p, (w, d) in enumerate(zip([0.1, 0.2, 0.5], [1.0, 50.0, 100.0])):
        spec["conns"].append({"src": [4], "tgt": [0], "weight": [w],
                              "delay": [d], "port": p})
        spec["conns"].append({"src": [4], "tgt": [1 + p], "weight": [w],
                              "delay": [d], "port": 0})
        spec["records"].append((0, "port_value", p))
    for p in range(3):
        spec["records"].append((1 + p, "V", 0))
    return spec


def load_stdp_cases():
    import json
    import os
    p = os.path.join(os.path.dirname(__file__), "golden", "stdp_cases.json")
    with open(p) as f:
        return json.load(f)


def stdp_case_spec(case, common):
    """Network of python/test/test_stdp/cases/case<k>.py: two parrot
    neurons driven by spike generators (weight 1, delay 1 ms), an STDP
    synapse pre -> post on receptor 1 (port weight 0, so the post neuron
    never relays the pre spikes), dendritic delay on the post neuron."""
    stdp = {k: common[k] for k in ("tau_plus", "tau_minus", "lambda", "alpha",
                                   "mu_plus", "mu_minus", "Wmax")}
    return {
        "dt": 0.1,
        "groups": [
            {"model": "parrot_neuron", "n": 1},                    # 0 pre
            {"model": "parrot_neuron", "n": 1,
             "params": {"den_delay": case["den_delay"]}},          # 1 post
            {"model": "spike_generator", "n": 1},                  # 2
            {"model": "spike_generator", "n": 1},                  # 3
        ],
        "spike_gen": [{"node": 2, "times": case["pre"]},
                      {"node": 3, "times": case["post"]}],
        "stdp": [stdp],
        "conns": [
            {"src": [2], "tgt": [0], "weight": 1.0,
             "delay": common["gen_delay"]},
            {"src": [3], "tgt": [1], "weight": 1.0,
             "delay": common["gen_delay"]},
            {"src": [0], "tgt": [1], "weight": common["weight"],
             "delay": case["delay"], "port": common["receptor"],
             "syn_group": 1},
        ],
    }


def stdp_closed_form(w, dts, common):
    """Weight after the sequence of updates of the reference's test
    (STDPUpdate of python/test/test_stdp/cases/case1.py:7-21)."""
    import math
    wplus = common["Wmax"] * common["lambda"]
    for dt in dts:
        if dt >= 0:
            fact = wplus * math.exp(-dt / common["tau_plus"])
            w = w + fact * math.pow(1.0 - w / common["Wmax"], common["mu_plus"])
            w = min(w, common["Wmax"])
        else:
            fact = -common["alpha"] * wplus * math.exp(dt / common["tau_minus"])
            w = w + fact * math.pow(w / common["Wmax"], common["mu_minus"])
            w = max(w, 0.0)
    return w


def nest_spec(model):
    """Single neuron with 3 ports driven by a spike generator at 10 and
    400 ms, delays 1/100/130 ms (python/test/test_aeif_cond_beta.py:4-30);
    the current-based AdEx models with 2 ports (1 for aeif_psc_delta),
    weights 1 / -2, delays 1 / 100 ms (python/test/test_aeif_psc_alpha.py:
    5-23, test_aeif_psc_exp.py, test_aeif_psc_delta.py)."""
    if model in ("aeif_psc_alpha", "aeif_psc_exp", "aeif_psc_delta"):
        params = {"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6,
                  "g_L": 300.0}
        n_port = 1
        if model == "aeif_psc_delta":
            params["C_m"] = 20000.0
        else:
            params["tau_syn"] = [40.0, 20.0]
            n_port = 2
        spec = {"groups": [{"model": model, "n": 1, "n_port": n_port,
                            "params": params},
                           {"model": "spike_generator", "n": 1}],
                "spike_gen": [{"node": 1, "times": [10.0, 400.0]}],
                "conns": [], "records": [(0, "V_m", 0)]}
        for syn, (w, d) in enumerate(zip([1.0, -2.0], [1.0, 100.0])):
            spec["conns"].append({"src": [1], "tgt": [0], "weight": [w],
                                  "delay": [d],
                                  "port": syn if n_port > 1 else 0})
        return spec
    if model in ("iaf_psc_exp_g", "iaf_psc_exp_hc", "izhikevich"):
        # python/test/test_iaf_psc_exp_g.py:8-27, test_iaf_psc_exp_hc.py,
        # test_izh.py:5-24: one port, weights 1 / -2
        izh = model == "izhikevich"
        times, delays = ([10.0, 40.0], [1.0, 10.0]) if izh else \
            ([10.0, 400.0], [1.0, 100.0])
        return {"groups": [{"model": model, "n": 1},
                           {"model": "spike_generator", "n": 1}],
                "spike_gen": [{"node": 1, "times": times}],
                "conns": [{"src": [1], "tgt": [0], "weight": [w],
                           "delay": [d], "port": 0}
                          for w, d in zip([1.0, -2.0], delays)],
                "records": [(0, "V_m" if izh else "V_m_rel", 0)]}
    if model == "iaf_psc_exp":
        return {"groups": [{"model": "iaf_psc_exp", "n": 1},
                           {"model": "spike_generator", "n": 1}],
                "spike_gen": [{"node": 1, "times": [10.0, 400.0]}],
                "conns": [{"src": [1], "tgt": [0], "weight": [1.0],
                           "delay": [1.0], "port": 0},
                          {"src": [1], "tgt": [0], "weight": [-2.0],
                           "delay": [100.0], "port": 1}],
                "records": [(0, "V_m_rel", 0)]}
    params = {"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0,
              "E_rev": [20.0, 0.0, -85.0]}
    if model == "aeif_cond_beta":
        params.update({"tau_decay": [40.0, 20.0, 30.0],
                       "tau_rise": [20.0, 10.0, 5.0]})
    else:
        params.update({"tau_syn": [40.0, 20.0, 30.0]})
    spec = {"groups": [{"model": model, "n": 1, "n_port": 3,
                        "params": params},
                       {"model": "spike_generator", "n": 1}],
            "spike_gen": [{"node": 1, "times": [10.0, 400.0]}],
            "conns": [], "records": [(0, "V_m", 0)]}
    for syn, (w, d) in enumerate(zip([0.1, 0.2, 0.5], [1.0, 100.0, 130.0])):
        spec["conns"].append({"src": [1], "tgt": [0], "weight": [w],
                              "delay": [d], "port": syn})
    return spec


def stdp_net_spec(n_src=200, n_post=100, k_plastic=20, seed=99,
                  den_delay=1.5, max_delay=5.0):
    """Plastic network whose plastic weights cannot feed back into the
    dynamics, so spike trains and final weights are comparable bit for bit
    with any accumulation order: aeif_cond_beta sources driven by constant
    current plus their own Poisson generator (weight 0.25 nS), parrot
    neurons relaying 3 sources each on port 0 (weight 1), and STDP synapses
    (fixed indegree k_plastic, random weights and delays) from the sources
    onto port 1 of the parrots, whose input weight is 0.  The parrots carry
    a dendritic delay."""
    rng = np.random.default_rng(seed)
    I_e = rng.uniform(450.0, 900.0, n_src).astype(np.float32)
    n_gen = n_src
    gen0 = n_src + n_post
    src_p = np.concatenate([rng.choice(n_src, k_plastic, replace=False)
                            for _ in range(n_post)])
    tgt_p = n_src + np.repeat(np.arange(n_post), k_plastic)
    w_p = rng.uniform(0.5, 5.0, src_p.size).astype(np.float32)
    d_p = (rng.integers(1, int(max_delay / 0.1) + 1, src_p.size)
           * 0.1).astype(np.float32)
    drv_src = np.concatenate([rng.choice(n_src, 3, replace=False)
                              for _ in range(n_post)])
    drv_tgt = n_src + np.repeat(np.arange(n_post), 3)
    return {
        "dt": 0.1, "seed": seed,
        "groups": [
            {"model": "aeif_cond_beta", "n": n_src, "n_port": 2,
             "params": {"I_e": I_e, "E_rev": [0.0, -85.0],
                        "tau_rise": [0.2, 0.5], "tau_decay": [2.0, 10.0]},
             "rk5": {"rk5_tol": 1e-3}},
            {"model": "parrot_neuron", "n": n_post,
             "params": {"den_delay": den_delay}},
            {"model": "poisson_generator", "n": n_gen},
        ],
        "poisson": [{"node": gen0 + i, "rate": 3000.0} for i in range(n_gen)],
        "stdp": [{"tau_plus": 20.0, "tau_minus": 20.0, "lambda": 0.01,
                  "alpha": 1.05, "mu_plus": 1.0, "mu_minus": 1.0,
                  "Wmax": 10.0}],
        "conns": [
            {"src": gen0 + np.arange(n_gen), "tgt": np.arange(n_src),
             "weight": 0.25, "delay": 0.1, "port": 0},
            {"src": drv_src, "tgt": drv_tgt, "weight": 1.0, "delay": 0.5,
             "port": 0},
            {"src": src_p, "tgt": tgt_p, "weight": w_p, "delay": d_p,
             "port": 1, "syn_group": 1},
        ],
    }


def spike_stats(times_by_neuron, n_neurons, t0, t1, n_pairs=500, bin_ms=5.0,
                seed=0):
    """Population statistics of a spike raster: mean rate (spikes/s per
    neuron), mean coefficient of variation of the inter-spike intervals
    (neurons with >= 3 spikes) and mean pairwise Pearson correlation of
    spike counts in bin_ms bins over n_pairs random neuron pairs."""
    dur = (t1 - t0) * 1e-3
    counts = np.array([len(ts) for ts in times_by_neuron], dtype=np.float64)
    rate = counts.sum() / n_neurons / dur
    cvs = []
    for ts in times_by_neuron:
        if len(ts) >= 3:
            isi = np.diff(np.sort(ts))
            if isi.mean() > 0:
                cvs.append(isi.std() / isi.mean())
    cv = float(np.mean(cvs)) if cvs else 0.0
    nb = int(round((t1 - t0) / bin_ms))
    rng = np.random.default_rng(seed)
    cc = []
    for _ in range(n_pairs):
        a, b = rng.choice(n_neurons, 2, replace=False)
        ha = np.histogram(times_by_neuron[a], bins=nb, range=(t0, t1))[0]
        hb = np.histogram(times_by_neuron[b], bins=nb, range=(t0, t1))[0]
        if ha.std() > 0 and hb.std() > 0:
            cc.append(np.corrcoef(ha, hb)[0, 1])
    corr = float(np.mean(cc)) if cc else 0.0
    return {"rate": float(rate), "cv_isi": cv, "corr": corr}


def oracle_times_by_neuron(net, dt, lo, hi):
    st, nd, ct = net.spikes()
    out = [[] for _ in range(hi - lo)]
    for s, n, c in zip(st, nd, ct):
        if lo <= n < hi:
            t = float(np.float64(np.float32(dt)) * (s + 1))
            out[n - lo] += [t] * int(c)
    return [np.array(x) for x in out]


def ensemble_instance_spec(n, seed, n_gen=40, k_rec=60, rate_hz=150.0,
                           t_max=300.0, w_gen=16.0):
    """One small AdEx network for the ensemble tests: n aeif_cond_beta
    neurons (2 ports), k_rec recurrent inputs each (80 % E with weight
    0.0625 nS on port 0, 20 % I with 0.375 nS on port 1, delays 1-20
    steps), driven by n_gen spike generators with seeded Poisson-like
    spike times (rate_hz) that connect to 30 random neurons each with
    weight w_gen nS.  All weights are dyadic, so every synaptic sum is exact
    in any order."""
    rng = np.random.default_rng(seed)
    n_e = int(0.8 * n)
    src, tgt, w, d, port = [], [], [], [], []
    for j in range(n):
        se = rng.choice(n_e, int(0.8 * k_rec), replace=False)
        si = n_e + rng.choice(n - n_e, k_rec - int(0.8 * k_rec),
                              replace=False)
        src += [se, si]
        tgt += [np.full(se.size, j), np.full(si.size, j)]
        w += [np.full(se.size, 0.0625), np.full(si.size, 0.375)]
        d += [rng.integers(1, 21, se.size), rng.integers(1, 21, si.size)]
        port += [np.zeros(se.size, int), np.ones(si.size, int)]
    src = np.concatenate(src).astype(np.int32)
    tgt = np.concatenate(tgt).astype(np.int32)
    w = np.concatenate(w).astype(np.float32)
    d = (np.concatenate(d) * 0.1).astype(np.float32)
    port = np.concatenate(port)
    gens = []
    g_src, g_tgt = [], []
    for g in range(n_gen):
        k = rng.poisson(rate_hz * t_max * 1e-3)
        times = np.unique(np.round(rng.uniform(1.0, t_max, k), 1))
        gens.append(times)
        tg = rng.choice(n, 30, replace=False)
        g_src.append(np.full(30, n + g))
        g_tgt.append(tg)
    return {"n": n, "n_gen": n_gen, "w_gen": w_gen,
            "rec": (src, tgt, w, d, port),
            "gen": (np.concatenate(g_src).astype(np.int32),
                    np.concatenate(g_tgt).astype(np.int32)),
            "gen_times": gens}


def ensemble_spec(instances):
    """Spec of several instances in one aeif group and one spike-generator
    group (node layout: all neurons, then all generators)."""
    n_tot = sum(x["n"] for x in instances)
    conns, spike_gen = [], []
    off, goff = 0, n_tot
    for x in instances:
        src, tgt, w, d, port = x["rec"]
        for p in (0, 1):
            m = port == p
            conns.append({"src": src[m] + off, "tgt": tgt[m] + off,
                          "weight": w[m], "delay": d[m], "port": p})
        gs, gt = x["gen"]
        conns.append({"src": gs - x["n"] + goff, "tgt": gt + off,
                      "weight": np.float32(x["w_gen"]),
                      "delay": np.float32(0.1),
                      "port": 0})
        for g, times in enumerate(x["gen_times"]):
            spike_gen.append({"node": goff + g, "times": times})
        off += x["n"]
        goff += x["n_gen"]
    params = {"E_rev": [0.0, -85.0], "tau_rise": [0.2, 0.5],
              "tau_decay": [2.0, 10.0]}
    return {"dt": 0.1, "seed": 99,
            "groups": [{"model": "aeif_cond_beta", "n": n_tot, "n_port": 2,
                        "params": params, "rk5": {"rk5_tol": 1e-3}},
                       {"model": "spike_generator",
                        "n": sum(x["n_gen"] for x in instances)}],
            "conns": conns, "spike_gen": spike_gen}
