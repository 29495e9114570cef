"""Seeded synthetic networks of the benchmark configurations, as explicit
specifications, and a builder that creates a specification through the
public Python API (i.e. through the C ABI).

A specification is a dict:
  dt, seed,
  groups:   [ {model, n, n_port, params{name: scalar | list | array},
               vars{...}, rk5: {h_min_rel, h0_rel, rk5_tol}} ]
  conns:    [ {src, tgt, weight, delay, port, syn_group} ]  (arrays, one
              entry per synapse, connected one_to_one in the given order)
  poisson:  [ {node, rate, origin, start, stop} ]
  spike_gen:[ {node, times, heights} ]
  stdp:     [ {param dict} ]        (synapse groups 1, 2, ...)
  records:  [ (node, var, port) ]

The same specification builds the CPU oracle's copy of a network in the
tests, so both see identical nodes, parameters, synapses and Poisson
streams.
"""
import numpy as np

PORT_PARAMS = {"E_rev", "tau_rise", "tau_decay", "tau_syn", "g0",
               "port_weight"}


def build_from_spec(ngpu, spec):
    """Reset the kernel and create the network of `spec`; returns the node
    sequences of the groups and the record ids."""
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetTimeResolution(spec.get("dt", 0.1))
    ngpu.SetRandomSeed(spec.get("seed", 54321))
    groups = []
    for g in spec["groups"]:
        if g["model"] == "legacy_poisson_generator":
            ns = ngpu.CreatePoissonGenerator(g["n"], g["rate"])
        else:
            ns = ngpu.Create(g["model"], g["n"], g.get("n_port", 1))
        groups.append(ns)
        for k, v in g.get("params", {}).items():
            _set(ngpu, ns, k, v)
        for k, v in g.get("vars", {}).items():
            _set(ngpu, ns, k, v)
        if "rk5" in g:
            for k, v in g["rk5"].items():
                ngpu.SetStatus(ns, {k: v})
    for s in spec.get("stdp", []):
        ngpu.CreateSynGroup("stdp", dict(s))
    for p in spec.get("poisson", []):
        ngpu.SetStatus([p["node"]], {"rate": p["rate"],
                                     "origin": p.get("origin", 0.0),
                                     "start": p.get("start", 0.0),
                                     "stop": p.get("stop", 1.0e30)})
    for sg in spec.get("spike_gen", []):
        d = {"spike_times": list(map(float, sg["times"]))}
        if sg.get("heights") is not None:
            d["spike_heights"] = list(map(float, sg["heights"]))
        ngpu.SetStatus([sg["node"]], d)
    for c in spec.get("conns", []):
        src = np.ascontiguousarray(c["src"], dtype=np.int32)
        n = src.size
        if n == 0:
            continue
        syn = {"weight": {"array": np.broadcast_to(
            np.asarray(c["weight"], np.float32), n)},
               "delay": {"array": np.broadcast_to(
                   np.asarray(c["delay"], np.float32), n)},
               "receptor": int(c.get("port", 0)),
               "synapse_group": int(c.get("syn_group", 0))}
        ngpu.Connect(src, np.ascontiguousarray(c["tgt"], dtype=np.int32),
                     {"rule": "one_to_one"}, syn)
    recs = []
    for (node, var, port) in spec.get("records", []):
        recs.append(ngpu.CreateRecord("", [var], [node], [port]))
    return groups, recs


def _set(ngpu, ns, k, v):
    if k in PORT_PARAMS and np.ndim(v) > 0:
        ngpu.SetStatus(ns, {k: list(map(float, v))})
    elif np.ndim(v) == 0:
        ngpu.SetStatus(ns, {k: float(v)})
    else:
        ngpu.SetStatus(ns, {k: {"array": np.asarray(v, np.float32)}})


def _normal_clipped(rng, mu, sigma, low, n, high=np.inf):
    x = rng.normal(mu, sigma, n)
    bad = (x < low) | (x > high)
    while bad.any():
        x[bad] = rng.normal(mu, sigma, int(bad.sum()))
        bad = (x < low) | (x > high)
    return x.astype(np.float32)


# ---------------------------------------------------------------------
# c1: balanced AdEx network at README scale (the headline workload)
# ---------------------------------------------------------------------
C1_NEURON = {"C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
             "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
             "V_reset": -60.0, "t_ref": 0.0, "E_rev": [0.0, -85.0],
             "tau_rise": [0.2, 0.5], "tau_decay": [2.0, 10.0]}


def build_c1(ngpu, n=1000000, seed=1234, tol=1e-3, k_e=800, k_i=200,
             w_e=0.05, w_i=0.25, w_p=0.1, rate=8000.0, log=None):
    """n aeif_cond_beta neurons (80 % E), 2 ports (E_rev 0 / -85 mV,
    tau_rise 0.2 / 0.5 ms, tau_decay 2 / 10 ms), V0 ~ U(-70, -55) mV;
    fixed in-degree k_e from the E neurons (w_e nS, port 0) plus k_i from
    the I neurons (w_i nS, port 1), drawn by the library's fixed_indegree
    rule (on the device at calibration); delays uniform over 1-20 time
    steps, drawn per synapse by the library (on the device with the rule);
    one Poisson generator at `rate` Hz connected to every neuron (w_p nS,
    port 0).  At n = 1e6 this is the 1e9-synapse configuration.  Returns
    the neurons' node sequence, the synapse count and V0."""
    import time
    rng = np.random.default_rng(seed)
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    n_e = int(round(0.8 * n))
    neuron = ngpu.Create("aeif_cond_beta", n, 2)
    ngpu.SetStatus(neuron, dict(C1_NEURON))
    ngpu.SetStatus(neuron, {"rk5_tol": tol})
    v0 = rng.uniform(-70.0, -55.0, n).astype(np.float32)
    ngpu.SetStatus(neuron, {"V_m": {"array": v0}})
    pg = ngpu.Create("poisson_generator", 1)
    ngpu.SetStatus(pg, {"rate": rate})
    t = time.time()
    delay = {"distribution": "uniform", "low": 0.1, "high": 2.0,
             "step": 0.1}
    ngpu.Connect(neuron[0:n_e], neuron,
                 {"rule": "fixed_indegree", "indegree": k_e},
                 {"weight": w_e, "delay": delay, "receptor": 0})
    ngpu.Connect(neuron[n_e:n], neuron,
                 {"rule": "fixed_indegree", "indegree": k_i},
                 {"weight": w_i, "delay": delay, "receptor": 1})
    ngpu.Connect(pg, neuron, {"rule": "all_to_all"},
                 {"weight": w_p, "delay": 0.1, "receptor": 0})
    if log:
        log("  connections in %.1f s" % (time.time() - t))
    return neuron, n * (k_e + k_i), v0


# ---------------------------------------------------------------------
# c2: layered cortical-column-shaped network of iaf_psc_exp neurons
# ---------------------------------------------------------------------
C2_SIZES = [20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948]
# external in-degrees of python/Potjans_2014_s/network_params.py:108,
# clipped to the configuration's range [1600, 2100]
C2_K_EXT = [1600, 1600, 2100, 1900, 2000, 1900, 2100, 2100]


def c2_spec(scale=1.0, seed=2014, integer_weights=False, dt=0.1,
            compensate=True):
    """8 populations (E, I alternating) of iaf_psc_exp neurons with the
    configuration's sizes times `scale`.  Connection probability p[j][i]
    ~ U(0.01, 0.15) per (target j, source i) population pair gives
    round(p N_i N_j) synapses with sources and targets drawn uniformly
    (fixed total number, multapses allowed).  Excitatory weights
    N(87.8, 8.8) pA on port 0, inhibitory -4 N(87.8, 8.8) pA on port 1;
    delays N(1.5, 0.75) ms (E) and N(0.8, 0.4) ms (I) clipped to >= 0.1 ms.
    One Poisson generator per population at 8 Hz x K_ext, weight 87.8 pA.
    With integer_weights the weights are rounded to whole pA, so every
    synaptic sum is exact in any order.  Below full size (scale < 1) the
    in-degrees shrink with the population sizes; with `compensate` the
    recurrent weights are divided by `scale` so that the mean recurrent
    input per neuron stays that of the full network."""
    rng = np.random.default_rng(seed)
    sizes = [max(1, int(round(n * scale))) for n in C2_SIZES]
    off = np.concatenate([[0], np.cumsum(sizes)])
    n_tot = int(off[-1])
    p = rng.uniform(0.01, 0.15, (8, 8))
    v0 = rng.uniform(0.0, 15.0, n_tot).astype(np.float32)
    conns = []
    for j in range(8):
        for i in range(8):
            k = int(round(p[j, i] * sizes[i] * sizes[j]))
            if k == 0:
                continue
            src = (off[i] + rng.integers(0, sizes[i], k)).astype(np.int32)
            tgt = (off[j] + rng.integers(0, sizes[j], k)).astype(np.int32)
            w = rng.normal(87.8, 8.8, k)
            if compensate and scale < 1.0:
                w = w / scale
            exc = i % 2 == 0
            if exc:
                d = _normal_clipped(rng, 1.5, 0.75, 0.1, k)
            else:
                w = -4.0 * w
                d = _normal_clipped(rng, 0.8, 0.4, 0.1, k)
            if integer_weights:
                w = np.round(w)
            conns.append({"src": src, "tgt": tgt,
                          "weight": w.astype(np.float32), "delay": d,
                          "port": 0 if exc else 1})
    w_ext = 88.0 if integer_weights else 87.8
    gen0 = n_tot
    for j in range(8):
        conns.append({"src": np.full(sizes[j], gen0 + j, np.int32),
                      "tgt": np.arange(off[j], off[j + 1], dtype=np.int32),
                      "weight": np.float32(w_ext), "delay": np.float32(0.1),
                      "port": 0})
    return {
        "dt": dt, "seed": seed,
        "groups": [{"model": "iaf_psc_exp", "n": n_tot,
                    "vars": {"V_m_rel": v0}},
                   {"model": "poisson_generator", "n": 8}],
        "poisson": [{"node": gen0 + j, "rate": 8.0 * C2_K_EXT[j]}
                    for j in range(8)],
        "conns": conns,
        "populations": [(int(off[j]), int(off[j + 1])) for j in range(8)],
    }


def n_synapses(spec, exclude_generators=True):
    gen_nodes = set(p["node"] for p in spec.get("poisson", []))
    tot = 0
    for c in spec["conns"]:
        src = np.asarray(c["src"])
        if exclude_generators and src.size and int(src[0]) in gen_nodes:
            continue
        tot += src.size
    return tot


# ---------------------------------------------------------------------
# c3: AdEx network with pair-based STDP on the E -> E synapses
# ---------------------------------------------------------------------
C3_STDP = {"tau_plus": 20.0, "tau_minus": 20.0, "lambda": 1.0e-3,
           "alpha": 1.05, "mu_plus": 1.0, "mu_minus": 1.0, "Wmax": 0.1}
C3_NEURON = {"C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
             "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
             "V_reset": -60.0, "t_ref": 0.0, "E_rev": [0.0, -85.0],
             "tau_syn": [2.0, 10.0]}


def c3_spec(scale=1.0, seed=4321, k=500, w_in=0.25, w_p=0.3, rate=8000.0,
            tol=1e-3, dt=0.1):
    """2e5 x scale aeif_cond_alpha neurons (80 % E), 2 ports (E_rev 0 /
    -85 mV, tau_syn 2 / 10 ms), fixed in-degree k (80 % from E, 20 % from
    I, without replacement).  E synapses start at w ~ U(0.02, 0.06) nS;
    those onto E neurons are plastic (pair-based STDP, tau+- 20 ms,
    lambda 1e-3, alpha 1.05, Wmax 0.1 nS, mu+- 1).  I synapses w_in nS on
    port 1.  Delays uniform 1-20 steps.  Each neuron has its own Poisson
    generator at `rate` Hz with weight w_p nS."""
    rng = np.random.default_rng(seed)
    n = max(10, int(round(200000 * scale)))
    n_e = int(round(0.8 * n))
    k_e, k_i = int(round(0.8 * k)), k - int(round(0.8 * k))
    src_e = np.concatenate([rng.choice(n_e, k_e, replace=False)
                            for _ in range(n)]).astype(np.int32)
    tgt_e = np.repeat(np.arange(n, dtype=np.int32), k_e)
    src_i = np.concatenate([n_e + rng.choice(n - n_e, k_i, replace=False)
                            for _ in range(n)]).astype(np.int32)
    tgt_i = np.repeat(np.arange(n, dtype=np.int32), k_i)
    w_e = rng.uniform(0.02, 0.06, src_e.size).astype(np.float32)
    d_e = (rng.integers(1, 21, src_e.size) * dt).astype(np.float32)
    d_i = (rng.integers(1, 21, src_i.size) * dt).astype(np.float32)
    ee = tgt_e < n_e
    v0 = rng.uniform(-70.0, -55.0, n).astype(np.float32)
    gen0 = n
    return {
        "dt": dt, "seed": seed,
        "groups": [{"model": "aeif_cond_alpha", "n": n, "n_port": 2,
                    "params": dict(C3_NEURON), "vars": {"V_m": v0},
                    "rk5": {"rk5_tol": tol}},
                   {"model": "poisson_generator", "n": n}],
        "stdp": [dict(C3_STDP)],
        "poisson": [{"node": gen0 + i, "rate": rate} for i in range(n)],
        "conns": [
            {"src": src_e[ee], "tgt": tgt_e[ee], "weight": w_e[ee],
             "delay": d_e[ee], "port": 0, "syn_group": 1},
            {"src": src_e[~ee], "tgt": tgt_e[~ee], "weight": w_e[~ee],
             "delay": d_e[~ee], "port": 0},
            {"src": src_i, "tgt": tgt_i,
             "weight": np.float32(w_in), "delay": d_i, "port": 1},
            {"src": np.arange(gen0, gen0 + n, dtype=np.int32),
             "tgt": np.arange(n, dtype=np.int32), "weight": np.float32(w_p),
             "delay": np.float32(dt), "port": 0},
        ],
        "populations": [(0, n_e), (n_e, n)],
    }


# ---------------------------------------------------------------------
# c4: batched ensemble of independent AdEx networks
# ---------------------------------------------------------------------
C4_NEURON = {"C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
             "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
             "V_reset": -60.0, "t_ref": 0.0, "E_rev": [0.0, -85.0],
             "tau_rise": [0.2, 0.5], "tau_decay": [2.0, 10.0]}


def c4_instances(n_inst=64, seed=6464, n_lo=5000, n_hi=50000):
    """Per-instance draws of the ensemble configuration: size log-uniform
    in [n_lo, n_hi], in-degree uniform in {100..1000}, E/I weight scale
    U(0.5, 2) times (0.05, 0.25) nS, Poisson rate U(4, 12) kHz."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        k = int(rng.integers(100, 1001))
        sc = float(rng.uniform(0.5, 2.0))
        rate = float(rng.uniform(4000.0, 12000.0))
        out.append({"n": n, "k": k, "w_e": 0.05 * sc, "w_i": 0.25 * sc,
                    "rate": rate})
    return out


def build_ensemble(ngpu, instances, seed=6464, tol=1e-3, w_p=0.1):
    """All instances in one aeif_cond_beta node group (one launch per
    kernel per step for the whole ensemble): instance k occupies a
    contiguous node range, draws 80 % of its in-degree from its own E
    neurons (port 0) and 20 % from its own I neurons (port 1), delays
    1-20 steps (drawn per synapse by the library), and has its own Poisson
    generator.  Returns the node sequence of the neurons, the instance
    ranges and the synapse count."""
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    rng = np.random.default_rng(seed)
    n_tot = sum(x["n"] for x in instances)
    neuron = ngpu.Create("aeif_cond_beta", n_tot, 2)
    ngpu.SetStatus(neuron, dict(C4_NEURON))
    ngpu.SetStatus(neuron, {"V_m": {"array": rng.uniform(
        -70.0, -55.0, n_tot).astype(np.float32)}, "rk5_tol": tol})
    pg = ngpu.Create("poisson_generator", len(instances))
    ranges, i0, n_syn = [], 0, 0
    for j, x in enumerate(instances):
        n, k = x["n"], x["k"]
        n_e = int(round(0.8 * n))
        k_e = min(int(round(0.8 * k)), n_e)
        k_i = min(k - int(round(0.8 * k)), n - n_e)
        net = neuron[i0:i0 + n]
        exc, inh = neuron[i0:i0 + n_e], neuron[i0 + n_e:i0 + n]
        delay = {"distribution": "uniform", "low": 0.1, "high": 2.0,
                 "step": 0.1}
        ngpu.Connect(exc, net, {"rule": "fixed_indegree", "indegree": k_e},
                     {"weight": x["w_e"], "delay": delay, "receptor": 0})
        ngpu.Connect(inh, net, {"rule": "fixed_indegree", "indegree": k_i},
                     {"weight": x["w_i"], "delay": delay, "receptor": 1})
        ngpu.SetStatus(pg[j:j + 1], {"rate": x["rate"]})
        ngpu.Connect(pg[j:j + 1], net, {"rule": "all_to_all"},
                     {"weight": w_p, "delay": 0.1, "receptor": 0})
        ranges.append((i0, i0 + n))
        n_syn += n * (k_e + k_i)
        i0 += n
    return neuron, ranges, n_syn
