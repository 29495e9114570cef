"""GPU parity of the benchmark configurations built by the benchmark's own
builders (neurongpu_amd.workloads), called through the C ABI:

  * the 1e6-neuron / 1e9-synapse AdEx network (in-degree 800 E + 200 I from
    the device fixed_indegree rule, delays 1-20 steps, 8 kHz Poisson drive,
    RK5 tolerance 1e-3) at 20 000 neurons with the full in-degree: synapse
    store properties, and every V_m, w, g, g1 bit-exact against the FP32
    oracle run on the synapses read back from the device after 200 ms;
  * the same network at full size: exact in-degree of every neuron per
    port, sources in range, no multapses, uniform sources (chi-square of
    the out-degrees) and delays (chi-square over the 20 slots), counted on
    the device;
  * instances of the batched-ensemble configuration built by
    build_ensemble, bit-exact against the FP32 oracle.
"""
import sys

import numpy as np
import pytest

import netbuild as nb
from neurongpu_amd import workloads as wl

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_from_engine(ngpu, seed, groups, rk5_tol, v0, poisson, params,
                        prec="f32"):
    """The network the engine holds, rebuilt in the oracle: node groups
    [(model, n, n_port)], the synapses read back from the device store in
    canonical order (generator connections included), V0 and the Poisson
    generators' rates."""
    c = ngpu.GetConnectionArrays()
    spec = {"dt": 0.1, "seed": seed,
            "groups": [], "poisson": poisson,
            "conns": [{"src": c["source"], "tgt": c["target"],
                       "weight": c["weight"], "delay": c["delay"],
                       "port": c["port"]}]}
    for model, n, n_port in groups:
        g = {"model": model, "n": n, "n_port": n_port}
        if model != "poisson_generator":
            g.update({"params": dict(params), "vars": {"V_m": v0},
                      "rk5": {"rk5_tol": rk5_tol}})
        spec["groups"].append(g)
    net, og = nb.build_oracle(spec, prec)
    return net, og, c


def _state(ngpu, i0, n):
    out = {"V_m": np.asarray(ngpu.GetNeuronVar(i0, n, "V_m"), np.float32),
           "w": np.asarray(ngpu.GetNeuronVar(i0, n, "w"), np.float32)}
    for name in ("g", "g1"):
        a = np.asarray(ngpu.GetNeuronVar(i0, n, name), np.float32)
        for q in range(2):
            out["%s%d" % (name, q)] = a[:, q]
    return out


def _oracle_state(og, i0, n):
    out = {"V_m": og.get("V_m")[i0:i0 + n], "w": og.get("w")[i0:i0 + n]}
    for name in ("g", "g1"):
        for q in range(2):
            out["%s%d" % (name, q)] = og.get(name, q)[i0:i0 + n]
    return out


def test_c1_reduced_n_bit_exact(ngpu):
    """The 1e9-synapse configuration's recipe at 20 000 neurons with the
    full in-degree (2e7 synapses): the device-built store has exactly 800 E
    sources on port 0 and 200 I sources on port 1 per neuron, no multapse,
    delays 1-20 steps, weights as specified; the device's connection counts
    equal those of the store read back; after 200 ms every neuron's V_m, w,
    g and g1 equal the FP32 oracle's bit for bit (as specified, the network
    is silent: the 8 kHz drive stays below the AdEx rheobase)."""
    n = 20000
    n_e = 16000
    log("case c1_reduced n", n)
    neuron, n_syn, v0 = wl.build_c1(ngpu, n=n)
    ngpu.ActivateSpikeCount(neuron)
    ngpu.Calibrate()
    # every connection onto the neurons: the recurrent ones and one from the
    # Poisson generator (node n) per neuron, as GetConnections lists them
    c_all = ngpu.GetConnectionArrays(target=neuron)
    c = ngpu.GetConnectionArrays(source=neuron, target=neuron)
    assert c["source"].size == n_syn == n * 1000
    assert c_all["source"].size == n_syn + n
    gen = c_all["source"] == n
    assert np.array_equal(np.sort(c_all["target"][gen]), np.arange(n))
    for port, lo, hi, k, w in ((0, 0, n_e, 800, 0.05), (1, n_e, n, 200,
                                                        0.25)):
        m = c["port"] == port
        src, tgt = c["source"][m], c["target"][m]
        assert np.all(np.bincount(tgt, minlength=n) == k)
        assert src.min() >= lo and src.max() < hi
        key = tgt.astype(np.int64) * n + src
        assert np.unique(key).size == key.size  # no multapse
        assert np.all(c["weight"][m] == np.float32(w))
        steps = np.rint(c["delay"][m] / np.float32(0.1)).astype(np.int64)
        assert steps.min() == 1 and steps.max() == 20
        # the device's counts agree with the store read back
        cnt = ngpu.GetConnectionCounts(neuron, port=port, n_delay=20)
        ma = c_all["port"] == port
        assert np.array_equal(cnt["in_degree"],
                              np.bincount(c_all["target"][ma], minlength=n))
        assert np.array_equal(cnt["out_degree"],
                              np.bincount(src, minlength=n))
        assert np.array_equal(cnt["delay_count"],
                              np.bincount(steps - 1, minlength=20))
    ngpu.Simulate(200.0)
    got = _state(ngpu, 0, n)
    spikes = int(np.sum(ngpu.GetNeuronIntVar(0, n, "spike_count")))
    net, og, _ = _oracle_from_engine(
        ngpu, 1234, [("aeif_cond_beta", n, 2), ("poisson_generator", 1, 1)],
        1e-3, v0, [{"node": n, "rate": 8000.0}], wl.C1_NEURON)
    net.run(2000, 16)
    exp = _oracle_state(og[0], 0, n)
    st, _, _ = net.spikes()
    log("spikes engine", spikes, "oracle", st.size, net.counters())
    assert spikes == 0 and st.size == 0
    for k in got:
        log(k, "max |diff|", float(np.max(np.abs(got[k] - exp[k]))))
        assert np.array_equal(got[k], exp[k]), k


def test_c1_full_size_store(ngpu):
    """The 1e6-neuron / 1e9-synapse store as the benchmark builds it (device
    fixed_indegree rule): every neuron has exactly 800 synapses on port 0,
    all from E neurons (plus the one from the Poisson generator), and 200
    on port 1, all from I neurons; the
    out-degrees of the sources are those of uniform sampling (chi-square
    and range); delays are uniform over the 20 slots (chi-square); the 200
    neurons of a sampled range have no multapse."""
    n, n_e = 1000000, 800000
    log("case c1_full_size_store")
    neuron, n_syn, _ = wl.build_c1(ngpu, n=n)
    ngpu.Calibrate()
    for port, lo, hi, k in ((0, 0, n_e, 800), (1, n_e, n, 200)):
        cnt = ngpu.GetConnectionCounts(neuron, port=port, n_delay=21)
        ind, outd, dc = cnt["in_degree"], cnt["out_degree"], cnt["delay_count"]
        n_gen = 1 if port == 0 else 0   # the Poisson generator's synapse
        assert np.all(ind == k + n_gen), (port, ind.min(), ind.max())
        # sources in range: no synapse on this port from the other class
        assert outd[:lo].sum() == 0 and outd[hi:].sum() == 0
        assert outd[lo:hi].sum() == n * k
        # source uniformity: out-degree ~ Binomial(n, k / n_src)
        m = n * k / (hi - lo)
        o = outd[lo:hi].astype(np.float64)
        chi2 = np.sum((o - m) ** 2) / m
        dof = hi - lo - 1
        var = 1.0 - k / (hi - lo)   # binomial variance factor
        log("port", port, "out-degree chi2/dof", chi2 / dof, "min",
            o.min(), "max", o.max())
        assert abs(chi2 - var * dof) < 6.0 * np.sqrt(2.0 * dof)
        assert np.max(np.abs(o - m)) < 7.0 * np.sqrt(m)
        # delays uniform over 1-20 steps (index 20 = 2.1 ms must be empty)
        assert dc[20] == 0 and dc.sum() == n * k
        e = n * k / 20.0
        chi2_d = np.sum((dc[:20] - e) ** 2) / e
        log("port", port, "delay chi2 (19 dof)", chi2_d)
        assert chi2_d < 60.0
    t0 = 123456
    c = ngpu.GetConnectionArrays(source=neuron, target=ngpu.NodeSeq(t0, 200))
    assert c["source"].size == 200 * 1000
    key = c["target"].astype(np.int64) * n + c["source"]
    assert np.unique(key).size == key.size
    for port, lo, hi, k in ((0, 0, n_e, 800), (1, n_e, n, 200)):
        m = c["port"] == port
        assert np.all(np.bincount(c["target"][m] - t0, minlength=200) == k)
        assert c["source"][m].min() >= lo and c["source"][m].max() < hi


def test_c4_ensemble_instances_bit_exact(ngpu):
    """Three instances of the batched-ensemble configuration (the smallest
    of its 64 seeded draws, each with its own in-degree, weight scale and
    Poisson rate) built into one node group by build_ensemble: every
    instance has its specified in-degree from its own E and I neurons, and
    after 200 ms every neuron's V_m, w, g and g1 equal the FP32 oracle's
    bit for bit (silent as specified)."""
    inst = sorted(wl.c4_instances(), key=lambda x: x["n"])[:3]
    log("case c4_instances", [(x["n"], x["k"]) for x in inst])
    neuron, ranges, n_syn = wl.build_ensemble(ngpu, inst)
    ngpu.ActivateSpikeCount(neuron)
    ngpu.Calibrate()
    n = neuron.n
    c = ngpu.GetConnectionArrays(source=neuron, target=neuron)
    assert c["source"].size == n_syn
    for x, (a, b) in zip(inst, ranges):
        n_i = b - a
        n_e = int(round(0.8 * n_i))
        k_e = int(round(0.8 * x["k"]))
        m = (c["target"] >= a) & (c["target"] < b)
        src, tgt, port = c["source"][m], c["target"][m], c["port"][m]
        e = port == 0
        assert np.all(np.bincount(tgt[e] - a, minlength=n_i) == k_e)
        assert np.all(np.bincount(tgt[~e] - a, minlength=n_i) ==
                      x["k"] - k_e)
        assert src[e].min() >= a and src[e].max() < a + n_e
        assert src[~e].min() >= a + n_e and src[~e].max() < b
        assert np.all(c["weight"][m][e] == np.float32(x["w_e"]))
        assert np.all(c["weight"][m][~e] == np.float32(x["w_i"]))
    v0 = np.asarray(ngpu.GetNeuronVar(0, n, "V_m"), np.float32)
    ngpu.Simulate(200.0)
    got = _state(ngpu, 0, n)
    spikes = int(np.sum(ngpu.GetNeuronIntVar(0, n, "spike_count")))
    net, og, _ = _oracle_from_engine(
        ngpu, 6464, [("aeif_cond_beta", n, 2),
                     ("poisson_generator", len(inst), 1)], 1e-3, v0,
        [{"node": n + j, "rate": x["rate"]} for j, x in enumerate(inst)],
        wl.C4_NEURON)
    net.run(2000, 16)
    exp = _oracle_state(og[0], 0, n)
    st, _, _ = net.spikes()
    log("spikes engine", spikes, "oracle", st.size, net.counters())
    assert spikes == st.size
    for k in got:
        log(k, "max |diff|", float(np.max(np.abs(got[k] - exp[k]))))
        assert np.array_equal(got[k], exp[k]), k
