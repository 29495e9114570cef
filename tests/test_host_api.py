"""Host-side API behaviour through the C ABI, no GPU needed: node creation,
parameter and variable access, connection rules, connection queries,
synapse groups and error handling happen on the host before calibration.

The scenario tests repeat the calls of the reference's own scripts
(python/test/test_connect.py, test_getarr.py, test_setvar2.py,
test_syn_group.py) and compare what they print with the reference's
golden logs (python/test/logp3_*.txt, banner removed), kept under
tests/golden/."""
import contextlib
import io
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_lines(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return f.read().rstrip("\n").split("\n")


def printed(fn, ng):
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        fn(ng)
    return buf.getvalue().rstrip("\n").split("\n")


def strip_trailing_blank(lines):
    while lines and lines[-1].strip() == "":
        lines = lines[:-1]
    return lines


# python/test/test_connect.py:1-94
def scenario_connect(ng):
    N = 5
    neuron = ng.Create("aeif_cond_beta", 2 * N)
    even = [neuron[2 * i] for i in range(N)]
    odd = [neuron[2 * i + 1] for i in range(N)]
    e2o_d, e2o_w, o2e_d, o2e_w = [], [], [], []
    for it in range(N):
        ite, ito = 2 * it, 2 * it + 1
        for isrc in range(N):
            ise, iso = 2 * isrc, 2 * isrc + 1
            e2o_d.append(2.0 * N * ise + ito)
            e2o_w.append(100.0 * (2.0 * N * ise + ito))
            o2e_d.append(2.0 * N * iso + ite)
            o2e_w.append(100.0 * (2.0 * N * iso + ite))
    conn = {"rule": "all_to_all"}
    ng.Connect(even, odd, conn, {"weight_array": e2o_w, "delay_array": e2o_d})
    ng.Connect(odd, even, conn, {"weight_array": o2e_w, "delay_array": o2e_d})

    def dump(title, src, tgt, keys=None):
        cid = ng.GetConnections(src, tgt)
        st = ng.GetStatus(cid) if keys is None else ng.GetStatus(cid, keys)
        print("########################################")
        print(title)
        for s in st:
            print(s)
        print()
        print()

    dump("Even to all", even, neuron)
    dump("Even to all weight, delat", even, neuron, ["weight", "delay"])
    dump("All to odd", neuron, odd)
    n36 = neuron[3:7]
    dump("Even to 3,4,5,6", even, n36)
    dump("3,4,5,6 to odd", n36, odd)


# python/test/test_getarr.py:1-50
def scenario_getarr(ng):
    spike = ng.Create("spike_generator", 4)
    s0, s1, s2 = spike[0:2], spike[2:3], spike[3:4]
    ng.SetStatus(s0, {"spike_times": [10.0, 400.0], "spike_heights": [1.0, 0.5]})
    ng.SetStatus(s1, {"spike_times": [4.0], "spike_heights": [2.0]})
    ng.SetStatus(s2, {"spike_times": [50.0, 20.0, 80.0],
                      "spike_heights": [0.1, 0.3, 0.2]})
    for s in (s0, s1, s2):
        print(ng.GetStatus(s, "spike_times"))
        print(ng.GetStatus(s, "spike_heights"))
        if s is not s2:
            print()
    print()
    print()
    nl = [spike[2], spike[3], spike[0], spike[1]]
    print(ng.GetStatus(nl, "spike_times"))
    print(ng.GetStatus(nl, "spike_heights"))
    print()
    print()
    print(ng.GetStatus(s1))
    print()
    print()
    print(ng.GetStatus(nl))


# python/test/test_setvar2.py:1-52
def scenario_setvar2(ng):
    neuron = ng.Create("aeif_cond_beta", 6, 2)
    even = [neuron[0], neuron[2], neuron[4]]
    odd = [neuron[3], neuron[5], neuron[1]]
    ng.SetStatus(even, {"tau_decay": [80.0, 40.0], "tau_rise": [60.0, 20.0]})
    ng.SetStatus(odd, {"tau_decay": [70.0, 30.0], "tau_rise": [50.0, 10.0]})
    ng.SetStatus(even, {"V_m": -80.0})
    ng.SetStatus(odd, {"V_m": -90.0})
    ng.SetStatus(even, {"g1": [0.4, 0.2]})
    ng.SetStatus(odd, {"g1": [0.3, 0.1]})
    ng.SetStatus(even, {"V_th": -40.0})
    ng.SetStatus(odd, {"V_th": -30.0})
    names = ["tau_decay", "tau_rise", "V_m", "V_th", "g1"]
    tags = ["td", "tr", "Vm", "Vth", "g1"]
    for tag, nm in zip(tags, names):
        print("read_" + tag, ng.GetNeuronStatus(neuron, nm))
    nl = [neuron[0], neuron[2], neuron[4], neuron[1], neuron[3], neuron[5]]
    for tag, nm in zip(tags, names):
        print("read1_" + tag, ng.GetNeuronStatus(nl, nm))


# python/test/test_syn_group.py:1-30
def scenario_syn_group(ng):
    sg = ng.CreateSynGroup("test_syn_model", {"fact": 1.0, "offset": 2.0})
    print(ng.GetStatus(sg))
    ng.SetStatus(sg, "fact", 3.0)
    ng.SetStatus(sg, "offset", 4.0)
    print(ng.GetStatus(sg))
    ng.SetStatus(sg, {"fact": 5.0, "offset": 6.0})
    print(ng.GetStatus(sg))
    print("fact: ", ng.GetSynGroupParam(sg, "fact"))
    print("offset: ", ng.GetSynGroupParam(sg, "offset"))
    print(ng.GetStatus(sg, "fact"))
    print(ng.GetStatus(sg, "offset"))
    print(ng.GetStatus(sg, ["fact", "offset"]))
    print(ng.GetStatus(sg, ["offset", "fact"]))


@pytest.mark.parametrize("name,fn", [
    ("connect", scenario_connect),
    ("getarr", scenario_getarr),
    ("setvar2", scenario_setvar2),
    ("syn_group", scenario_syn_group),
])
def test_reference_log(ngpu, name, fn):
    got = strip_trailing_blank(printed(fn, ngpu))
    want = strip_trailing_blank(golden_lines("logp3_%s.txt" % name))
    assert got == want


# ---------------------------------------------------------------------
# connection rules (src/connect_rules.h; counts follow the rule text)
# ---------------------------------------------------------------------
def _conns(ng, src, tgt):
    st = ng.GetStatus(ng.GetConnections(src, tgt))
    return [(d["source"], d["target"]) for d in st]


def test_rule_one_to_one(ngpu):
    a = ngpu.Create("aeif_cond_beta", 7)
    b = ngpu.Create("aeif_cond_beta", 7)
    ngpu.Connect(a, b, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 1.0})
    assert _conns(ngpu, a, b) == [(a[i], b[i]) for i in range(7)]


def test_rule_all_to_all(ngpu):
    a = ngpu.Create("aeif_cond_beta", 4)
    b = ngpu.Create("aeif_cond_beta", 3)
    ngpu.Connect(a, b, {"rule": "all_to_all"}, {"weight": 1.0, "delay": 1.0})
    got = sorted(_conns(ngpu, a, b))
    assert got == sorted((a[i], b[j]) for i in range(4) for j in range(3))


def test_rule_fixed_indegree(ngpu):
    a = ngpu.Create("aeif_cond_beta", 50)
    b = ngpu.Create("aeif_cond_beta", 20)
    ngpu.Connect(a, b, {"rule": "fixed_indegree", "indegree": 13},
                 {"weight": 1.0, "delay": 1.0})
    c = _conns(ngpu, a, b)
    assert len(c) == 20 * 13
    for j in range(20):
        srcs = [s for s, t in c if t == b[j]]
        assert len(srcs) == 13
        assert len(set(srcs)) == 13          # without replacement
        assert all(a[0] <= s < a[0] + 50 for s in srcs)


def test_rule_fixed_outdegree(ngpu):
    a = ngpu.Create("aeif_cond_beta", 20)
    b = ngpu.Create("aeif_cond_beta", 50)
    ngpu.Connect(a, b, {"rule": "fixed_outdegree", "outdegree": 11},
                 {"weight": 1.0, "delay": 1.0})
    c = _conns(ngpu, a, b)
    assert len(c) == 20 * 11
    for i in range(20):
        tg = [t for s, t in c if s == a[i]]
        assert len(tg) == 11 and len(set(tg)) == 11


def test_rule_fixed_total_number(ngpu):
    a = ngpu.Create("aeif_cond_beta", 30)
    b = ngpu.Create("aeif_cond_beta", 40)
    ngpu.Connect(a, b, {"rule": "fixed_total_number", "total_num": 321},
                 {"weight": 1.0, "delay": 1.0})
    c = _conns(ngpu, a, b)
    assert len(c) == 321
    assert all(a[0] <= s < a[0] + 30 and b[0] <= t < b[0] + 40 for s, t in c)


def test_weight_delay_distributions(ngpu):
    a = ngpu.Create("aeif_cond_beta", 100)
    ngpu.Connect(a, a, {"rule": "fixed_indegree", "indegree": 50},
                 {"weight": {"distribution": "normal_clipped", "mu": 1.0,
                             "sigma": 0.5, "low": 0.5, "high": 1.5},
                  "delay": {"distribution": "normal_clipped", "mu": 2.0,
                            "sigma": 1.0, "low": 0.5, "high": 3.0}})
    st = ngpu.GetStatus(ngpu.GetConnections(a, a))
    w = np.array([d["weight"] for d in st])
    d = np.array([d["delay"] for d in st])
    assert w.min() >= 0.5 and w.max() <= 1.5
    assert d.min() >= 0.5 - 1e-6 and d.max() <= 3.0 + 1e-6
    # delays are stored as whole time steps
    steps = d / 0.1
    assert np.allclose(steps, np.round(steps), atol=1e-3)
    assert abs(w.mean() - 1.0) < 0.05


def test_library_drawn_distributions(ngpu):
    """Weights and delays drawn by the library per connection (the synapse
    dictionary's distributions) follow the distribution: Kolmogorov-Smirnov
    against scipy's normal, truncated normal and uniform CDFs, and a
    chi-square over the 20 values of a stepped uniform delay."""
    from scipy import stats
    n1, n2 = 300000, 2000000

    def one(conn, syn, src, tgt):
        ngpu.ResetKernel()
        ngpu.SetVerbosityLevel(0)
        nodes = ngpu.Create("parrot_neuron", 3000)
        ngpu.Connect(nodes[src[0]:src[1]], nodes[tgt[0]:tgt[1]], conn, syn)
        return ngpu.GetConnectionArrays()

    c1 = one({"rule": "fixed_indegree", "indegree": 100},
             {"weight": {"distribution": "normal", "mu": 1.0, "sigma": 0.3},
              "delay": {"distribution": "uniform", "low": 0.1, "high": 2.0,
                        "step": 0.1}}, (0, 3000), (0, 3000))
    assert c1["weight"].size == n1
    p = stats.kstest(c1["weight"].astype(np.float64),
                     stats.norm(1.0, 0.3).cdf).pvalue
    steps = np.rint(c1["delay"] / np.float32(0.1)).astype(int)
    h = np.bincount(steps - 1, minlength=20)
    chi2 = np.sum((h - n1 / 20.0) ** 2) / (n1 / 20.0)
    assert p > 1e-3 and h.size == 20 and chi2 < 60.0
    c2 = one({"rule": "all_to_all"},
             {"weight": {"distribution": "normal_clipped", "mu": 0.5,
                         "sigma": 0.25, "low": 0.1, "high": 1.25},
              "delay": {"distribution": "normal_clipped", "mu": 1.5,
                        "sigma": 0.75, "low": 0.1, "high": 4.0}},
             (0, 1000), (1000, 3000))
    assert c2["weight"].size == n2
    tn = stats.truncnorm((0.1 - 0.5) / 0.25, (1.25 - 0.5) / 0.25, 0.5, 0.25)
    assert stats.kstest(c2["weight"].astype(np.float64), tn.cdf).pvalue > 1e-3
    # delays: round(d / dt) slots of a truncated normal
    td = stats.truncnorm((0.1 - 1.5) / 0.75, (4.0 - 1.5) / 0.75, 1.5, 0.75)
    k = np.rint(c2["delay"] / np.float32(0.1)).astype(int)
    edges = (np.arange(1, 42) - 0.5) * 0.1
    pk = np.diff(td.cdf(np.clip(edges, 0.1, 4.0)))
    hk = np.bincount(k - 1, minlength=40)[:40]
    ok = pk * n2 > 50
    chi2 = np.sum((hk[ok] - pk[ok] * n2) ** 2 / (pk[ok] * n2))
    assert chi2 < ok.sum() + 8.0 * np.sqrt(2.0 * ok.sum())
    c3 = one({"rule": "one_to_one"},
             {"weight": {"distribution": "uniform", "low": -0.5, "high": 0.7},
              "delay": 0.3}, (0, 500), (500, 1000))
    w3 = c3["weight"].astype(np.float64)
    assert w3.min() > -0.5 and w3.max() <= 0.7
    assert stats.kstest(w3, stats.uniform(-0.5, 1.2).cdf).pvalue > 1e-3


def test_distribution_errors(ngpu):
    a = ngpu.Create("parrot_neuron", 100)
    with pytest.raises(ValueError, match="time resolution"):
        # unbounded normal delays: some fall below the resolution
        ngpu.Connect(a, a, {"rule": "fixed_indegree", "indegree": 50},
                     {"weight": 1.0, "delay": {"distribution": "normal",
                                               "mu": 0.2, "sigma": 1.0}})
        ngpu.GetConnectionArrays()
    with pytest.raises(ValueError, match="probability"):
        ngpu.Connect(a, a, {"rule": "all_to_all"},
                     {"weight": {"distribution": "normal_clipped", "mu": 0.0,
                                 "sigma": 0.1, "low": 5.0, "high": 6.0},
                      "delay": 1.0})
    with pytest.raises(ValueError, match="Unknown distribution"):
        ngpu.Connect(a, a, {"rule": "all_to_all"},
                     {"weight": {"distribution": "lognormal"}, "delay": 1.0})


def test_connection_status_fields(ngpu):
    a = ngpu.Create("aeif_cond_beta", 2, 3)
    sg = ngpu.CreateSynGroup("stdp")
    ngpu.Connect([a[0]], [a[1]], {"rule": "one_to_one"},
                 {"weight": 2.5, "delay": 0.75, "receptor": 2,
                  "synapse_group": sg})
    (d,) = ngpu.GetStatus(ngpu.GetConnections(a[0], a[1]))
    assert d["port"] == 2 and d["syn"] == sg.i_syn_group
    assert d["weight"] == 2.5
    # delay rounded to whole steps: round(0.75 / 0.1) = 8 steps
    assert abs(d["delay"] - 0.8) < 1e-6


def test_bulk_connection_queries(ngpu):
    """GetConnectionArrays returns, in one call, what GetConnections and
    GetConnectionStatus give one connection at a time (same order); the
    host path of GetConnectionCounts equals counts taken from it."""
    rng = np.random.default_rng(3)
    a = ngpu.Create("aeif_cond_beta", 60, 2)
    b = ngpu.Create("iaf_psc_exp", 40)
    sg = ngpu.CreateSynGroup("stdp")
    ngpu.Connect(a, b, {"rule": "fixed_indegree", "indegree": 7},
                 {"weight": {"array": rng.normal(1.0, 0.2, 280)
                             .astype(np.float32)},
                  "delay": {"array": rng.integers(1, 9, 280)
                            .astype(np.float32) * 0.1}, "receptor": 1})
    ngpu.Connect(b[0:10], a, {"rule": "fixed_outdegree", "outdegree": 5},
                 {"weight": 0.5, "delay": 0.3, "synapse_group": sg})
    ngpu.Connect(a[0:20], a[20:40], {"rule": "one_to_one"},
                 {"weight": 0.25, "delay": 1.2, "receptor": 0})
    ids = ngpu.GetConnections()
    ref = [ngpu.GetConnectionStatus(c) for c in ids]
    got = ngpu.GetConnectionArrays()
    assert got["source"].size == len(ref) == 280 + 50 + 20
    for k, d in enumerate(ref):
        assert (got["source"][k], got["target"][k], got["port"][k],
                got["syn"][k]) == (d["source"], d["target"], d["port"],
                                   d["syn"])
        assert got["delay"][k] == np.float32(d["delay"])
        assert got["weight"][k] == np.float32(d["weight"])
    sub = ngpu.GetConnectionArrays(source=a, target=b)
    m = (got["source"] < 60) & (got["target"] >= 60)
    assert np.array_equal(sub["target"], got["target"][m])
    steps = np.rint(got["delay"] / np.float32(0.1)).astype(int)
    for port in (-1, 0, 1):
        m = got["port"] == port if port >= 0 else np.ones(got["port"].size,
                                                         bool)
        cnt = ngpu.GetConnectionCounts(a, port=port, n_delay=15)
        assert np.array_equal(cnt["in_degree"],
                              np.bincount(got["target"][m], minlength=100)
                              [:60])
        assert np.array_equal(cnt["out_degree"],
                              np.bincount(got["source"][m], minlength=100)
                              [:60])
        own = m & (got["source"] < 60)
        assert np.array_equal(cnt["delay_count"],
                              np.bincount(steps[own] - 1, minlength=15))


# ---------------------------------------------------------------------
# errors (src/connect.cu, src/neurongpu.cu messages)
# ---------------------------------------------------------------------
def test_error_delay_below_resolution(ngpu):
    a = ngpu.Create("aeif_cond_beta", 2)
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, {"rule": "all_to_all"},
                     {"weight": 1.0, "delay": 0.01})


def test_error_port_out_of_range(ngpu):
    a = ngpu.Create("aeif_cond_beta", 2, 2)
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, {"rule": "all_to_all"},
                     {"weight": 1.0, "delay": 1.0, "receptor": 2})


@pytest.mark.parametrize("conn", [
    {"rule": "fixed_indegree", "indegree": -1},
    {"rule": "fixed_outdegree", "outdegree": -1},
    {"rule": "fixed_total_number", "total_num": -1},
])
def test_error_negative_connection_count(ngpu, conn):
    a = ngpu.Create("aeif_cond_beta", 4)
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, conn, {"weight": 1.0, "delay": 1.0})


def test_error_unknown_model(ngpu):
    with pytest.raises(ValueError):
        ngpu.Create("no_such_model", 3)


@pytest.mark.parametrize("model,defaults,group", [
    # src/iaf_psc_exp_g.cu:96-118
    ("iaf_psc_exp_g", {"I_syn": 0.0, "V_m_rel": 0.0, "I_e": 0.0,
                       "tau_m": 10.0, "C_m": 250.0, "E_L": -65.0,
                       "Theta_rel": 15.0, "tau_syn": 0.5, "t_ref": 2.0},
     ["tau_m", "C_m", "E_L", "Theta_rel", "V_reset_rel", "tau_syn", "t_ref"]),
    # src/iaf_psc_exp_hc.cu:72-84
    ("iaf_psc_exp_hc", {"I_syn": 0.0, "V_m_rel": 0.0, "I_e": 0.0}, []),
    # src/izhikevich.cu:96-110
    ("izhikevich", {"V_m": -70.0, "u": -14.0, "V_th": 30.0, "c": -65.0,
                    "d": 8.0, "t_ref": 0.0, "den_delay": 0.0},
     ["V_th", "a", "b", "c", "d", "t_ref"]),
    # src/izhikevich_psc_exp_5s.cu:110-122
    ("izhikevich_psc_exp_5s", {"V_m": -70.0, "u": -14.0, "tau_syn": 2.0},
     ["V_th", "a", "b", "c", "d", "tau_syn", "t_ref"]),
])
def test_point_model_defaults(ngpu, model, defaults, group):
    """Fixed-step models: the reference's defaults and group parameters,
    one receptor port whatever n_port asks for."""
    ns = ngpu.Create(model, 3, 4)
    st = ngpu.GetStatus(ns[2])
    for k, v in defaults.items():
        assert st[k] == pytest.approx(v, rel=1e-7), k
    assert ngpu.GetGroupParamNames(ns[0]) == group
    with pytest.raises(ValueError):     # port 1 does not exist
        ngpu.Connect(ns, ns, {"rule": "all_to_all"},
                     {"weight": 1.0, "delay": 1.0, "receptor": 1})
    if group:
        ngpu.SetStatus(ns, {group[0]: 1.5})
        assert ngpu.GetStatus(ns[1], group[0]) == 1.5
        with pytest.raises(ValueError):  # a group parameter of part of it
            ngpu.SetNeuronGroupParam(ngpu.NodeSeq(ns.i0, 2), group[0], 1.0)


def test_ext_neuron_host(ngpu):
    """ext_neuron: two port variables and one port parameter per port
    (src/ext_neuron.h:46-60); GetExtNeuronInputSpikes lists every port
    with include_zeros."""
    ns = ngpu.Create("ext_neuron", 2, 3)
    assert ngpu.GetPortVarNames(ns[0]) == ["port_input", "port_value"]
    assert ngpu.GetStatus(ns[1], "port_weight") == [1.0, 1.0, 1.0]
    nodes, ports, hs = ngpu.GetExtNeuronInputSpikes(include_zeros=True)
    assert nodes == [0, 0, 0, 1, 1, 1] and ports == [0, 1, 2] * 2
    assert hs == [0.0] * 6
    assert ngpu.GetExtNeuronInputSpikes() == ([], [], [])


def test_create_poisson_generator(ngpu):
    """src/neurongpu.cu:245-265: one group, created before calibration,
    with no receptor ports; a second call or n_node <= 0 is an error."""
    a = ngpu.Create("aeif_cond_beta", 3, 2)
    with pytest.raises(ValueError):
        ngpu.CreatePoissonGenerator(0, 100.0)
    pg = ngpu.CreatePoissonGenerator(4, 20000.0)
    assert pg.i0 == 3 and pg.n == 4
    with pytest.raises(ValueError):
        ngpu.CreatePoissonGenerator(2, 100.0)
    with pytest.raises(ValueError):     # no model of that name
        ngpu.Create("base_neuron", 2)
    with pytest.raises(ValueError):     # no receptor port on the generator
        ngpu.Connect(a, pg, {"rule": "all_to_all"},
                     {"weight": 1.0, "delay": 1.0})
    ngpu.Connect(pg, a, {"rule": "all_to_all"}, {"weight": 0.5, "delay": 1.0})
    assert len(ngpu.GetConnections(pg, a)) == 12


def test_create_poisson_generator_rate_range(ngpu):
    with pytest.raises(ValueError):     # mean count per step above 150
        ngpu.CreatePoissonGenerator(4, 4.0e6)
    with pytest.raises(ValueError):
        ngpu.CreatePoissonGenerator(4, -5.0)


def test_error_unknown_parameter(ngpu):
    a = ngpu.Create("aeif_cond_beta", 2)
    with pytest.raises(ValueError):
        ngpu.SetStatus(a, {"no_such_param": 1.0})


def test_error_state_cleared_after_success(ngpu):
    a = ngpu.Create("aeif_cond_beta", 2)
    with pytest.raises(ValueError):
        ngpu.SetStatus(a, {"no_such_param": 1.0})
    ngpu.SetStatus(a, {"V_m": -60.0})
    assert ngpu.GetStatus(a, "V_m") == [-60.0, -60.0]


# ---------------------------------------------------------------------
# parameters, names, defaults
# ---------------------------------------------------------------------
def test_aeif_cond_beta_defaults(ngpu):
    # src/aeif_cond_beta.cu:25-56
    a = ngpu.Create("aeif_cond_beta", 1, 2)
    st = ngpu.GetStatus(a[0])
    want = {"V_th": -50.4, "Delta_T": 2.0, "g_L": 30.0, "E_L": -70.6,
            "C_m": 281.0, "a": 4.0, "b": 80.5, "tau_w": 144.0, "I_e": 0.0,
            "V_peak": 0.0, "V_reset": -60.0, "t_ref": 0.0}
    for k, v in want.items():
        assert st[k] == float(np.float32(v)), k
    assert st["E_rev"] == [0.0, 0.0]
    assert st["tau_decay"] == [20.0, 20.0]
    assert st["tau_rise"] == [2.0, 2.0]


def test_names(ngpu):
    a = ngpu.Create("aeif_cond_beta", 1, 2)
    assert ngpu.GetScalVarNames(a[0]) == ["V_m", "w"]
    assert ngpu.GetPortVarNames(a[0]) == ["g", "g1"]
    assert ngpu.GetPortParamNames(a[0]) == ["E_rev", "tau_rise", "tau_decay",
                                            "g0"]
    assert ngpu.GetScalParamNames(a[0])[:13] == [
        "V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w", "I_e",
        "V_peak", "V_reset", "t_ref", "refractory_step"]


def test_array_setters(ngpu):
    a = ngpu.Create("aeif_cond_beta", 5)
    vals = np.linspace(-70.0, -60.0, 5).astype(np.float32)
    ngpu.SetStatus(a, {"V_m": {"array": vals}})
    assert np.array_equal(np.float32(ngpu.GetStatus(a, "V_m")), vals)
    ngpu.SetStatus(a, {"V_m": {"distribution": "normal_clipped", "mu": 0.0,
                               "sigma": 3.0, "low": -1.0, "high": 1.0}})
    v = np.array(ngpu.GetStatus(a, "V_m"))
    assert v.min() >= -1.0 and v.max() <= 1.0


def test_random_helpers(ngpu):
    u = np.array(ngpu.RandomUniform(10000))
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.02
    n = np.array(ngpu.RandomNormal(10000, 1.0, 2.0))
    assert abs(n.mean() - 1.0) < 0.1 and abs(n.std() - 2.0) < 0.1
    c = np.array(ngpu.RandomNormalClipped(10000, 0.0, 1.0, -0.5, 0.5))
    assert c.min() >= -0.5 and c.max() <= 0.5


def test_kernel_status(ngpu):
    ngpu.SetTimeResolution(0.05)
    # stored as float (src/neurongpu.h: float time_resolution_)
    assert ngpu.GetTimeResolution() == float(np.float32(0.05))
    ngpu.SetMaxSpikeBufferSize(30)
    assert ngpu.GetMaxSpikeBufferSize() == 30


def test_launch_settings(ngpu):
    """Launch-shape settings of this implementation: defaults, round trip
    and range checks (reported like the reference's parameter errors)."""
    defaults = {"rk5_two_pass": 1, "rk5_budget": 1, "rk5_pass2_grid": 256,
                "fuse_delivery": 1, "delivery_grid": 2048,
                "delivery_tile_max": 64, "delivery_mode": 2,
                "delivery_bin_capacity": 1 << 27,
                "delivery_auto_events": 500000, "graph_steps": 0,
                "delivery_streams": 2, "step_overlap": 8,
                "device_connect": 1, "delivery_tile_split": 0,
                "delivery_tile_shift": 15, "delivery_tile_fixed": 0}
    for k, v in defaults.items():
        assert ngpu.GetKernelStatus(k) == v
    ngpu.SetKernelStatus("delivery_mode", 1)
    assert ngpu.GetKernelStatus("delivery_mode") == 1
    ngpu.SetKernelStatus("delivery_bin_capacity", 1000)
    assert ngpu.GetKernelStatus("delivery_bin_capacity") == 1000
    ngpu.SetKernelStatus("graph_steps", 16)
    assert ngpu.GetKernelStatus("graph_steps") == 16
    ngpu.SetKernelStatus("delivery_mode", 3)
    assert ngpu.GetKernelStatus("delivery_mode") == 3
    ngpu.SetKernelStatus("delivery_tile_split", 8)
    assert ngpu.GetKernelStatus("delivery_tile_split") == 8
    for k, bad in (("delivery_mode", 4), ("delivery_tile_max", 3),
                   ("delivery_tile_split", -1), ("delivery_tile_split", 257),
                   ("delivery_tile_shift", 9), ("delivery_tile_shift", 16),
                   ("delivery_grid", 0), ("rk5_budget", 0),
                   ("delivery_bin_capacity", 0), ("graph_steps", -1),
                   ("graph_steps", 7), ("delivery_streams", 3),
                   ("step_overlap", -1), ("step_overlap", 1025),
                   ("device_connect", 2), ("delivery_tile_fixed", 2)):
        with pytest.raises(ValueError):
            ngpu.SetKernelStatus(k, bad)


def test_autotuner_keeps_incumbent_within_noise():
    """The autotuner replaces a setting only when the candidate's median
    window is more than the margin faster; a fake engine whose windows
    scatter by up to +-1 % around the same mean keeps every default, and
    one where a value is 10 % faster takes it."""
    import random
    from neurongpu_amd import autotune as at

    class Fake:
        def __init__(self, fast=None):
            self.s = {k: v[0] for k, v in at.CANDIDATES.items()}
            self.fast = fast or {}
            self.rng = random.Random(7)

        def GetKernelStatus(self, k):
            return self.s[k]

        def SetKernelStatus(self, k, v):
            self.s[k] = v

        def GetTimeResolution(self):
            return 0.1

        def Simulate(self, t):
            f = 1.0 + self.rng.uniform(-0.01, 0.01)
            for k, v in self.fast.items():
                if self.s[k] == v:
                    f *= 0.9
            time_ = 1e-4 * f * t / 0.1
            end = at.time.perf_counter() + time_
            while at.time.perf_counter() < end:
                pass

    eng = Fake()
    out = at.autotune(eng, steps=4, reps=3)
    for k, v in at.CANDIDATES.items():
        assert out[k] == v[0], k
    eng = Fake(fast={"delivery_grid": 4096})
    out = at.autotune(eng, steps=4, reps=3)
    assert out["delivery_grid"] == 4096
    assert at.choose([1.0, 1.0, 1.0], [0.99, 0.99, 0.99], 0.02) is False
    assert at.choose([1.0, 1.0, 1.0], [0.9, 0.9, 2.0], 0.02) is True
