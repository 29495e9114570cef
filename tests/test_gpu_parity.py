"""GPU parity of the per-timestep path against the CPU oracle (and the
reference's own golden NEST traces), called through the C ABI.

Bars:
  * FP32 oracle (same operation sequence): traces, spike times and routing
    bit-exact where the accumulation order cannot matter (single neurons,
    dyadic synaptic weights); otherwise identical spike statistics.
  * FP64 oracle: single-neuron traces under fixed drive within relative
    error 1e-5 over 1 s.
  * reference golden NEST traces: the reference tests' own tolerance
    (std(dV)/|mean V| <= 5e-4 aeif, 5e-5 iaf_psc_exp).
"""
import os
import sys

import numpy as np
import pytest

import netbuild as nb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("model", ["aeif_cond_beta", "aeif_cond_alpha",
                                   "iaf_psc_exp", "aeif_psc_alpha",
                                   "aeif_psc_exp", "aeif_psc_delta"])
def test_nest_golden_trace(ngpu, model):
    log("case nest_golden", model)
    spec = nb.nest_spec(model)
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(800.0)
    data = ngpu.GetRecordDataArray(recs[0])
    V = data[:, 1].astype(np.float64)
    if model == "iaf_psc_exp":
        V = V - 70.0
    nest = np.loadtxt(os.path.join(GOLDEN, "%s_nest.txt" % model),
                      delimiter="\t")
    dV = np.array([V[i * 10 + 20] - nest[i, 1] for i in range(len(nest))])
    rmse = np.std(dV) / abs(np.mean(V))
    tol = 5e-5 if model == "iaf_psc_exp" else 5e-4
    log("rmse", rmse, "tol", tol)
    assert rmse <= tol
    # the same run in the FP32 oracle: identical trace
    net, _ = nb.build_oracle(spec, "f32")
    net.run(8000)
    orc = net.records()
    assert orc.shape == data.shape
    assert np.array_equal(orc[:, 1].astype(np.float32), data[:, 1])


@pytest.mark.parametrize("I_e,tol", [(500.0, 1e-6), (1000.0, 1e-6),
                                     (1000.0, 1e-3)])
def test_single_neuron_fixed_drive(ngpu, I_e, tol):
    """aeif_cond_beta under constant current for 1 s."""
    log("case fixed_drive I_e", I_e, "tol", tol)
    spec = {"groups": [{"model": "aeif_cond_beta", "n": 1, "n_port": 2,
                        "params": {"I_e": I_e}, "rk5": {"rk5_tol": tol}}],
            "records": [(0, "V_m", 0), (0, "w", 0)]}
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(1000.0)
    V = ngpu.GetRecordDataArray(recs[0])[:, 1]
    W = ngpu.GetRecordDataArray(recs[1])[:, 1]
    n32, _ = nb.build_oracle(spec, "f32")
    n32.run(10000)
    r32 = n32.records()
    assert np.array_equal(r32[:, 1].astype(np.float32), V)
    assert np.array_equal(r32[:, 2].astype(np.float32), W)
    if I_e < 546.0:   # below rheobase: smooth trace, compare with FP64
        n64, _ = nb.build_oracle(spec, "f64")
        n64.run(10000)
        r64 = n64.records()
        # relative error of each trace: max deviation over the trace's
        # magnitude (w starts at 0, so a pointwise ratio is ill-posed there)
        relV = np.max(np.abs(V - r64[:, 1])) / np.max(np.abs(r64[:, 1]))
        relW = np.max(np.abs(W - r64[:, 2])) / np.max(np.abs(r64[:, 2]))
        log("rel err vs FP64: V", relV, "w", relW)
        assert relV <= 1e-5
        assert relW <= 1e-5


@pytest.mark.parametrize("I_e", [700.0, 1000.0])
def test_spiking_neuron_fp64_spike_times(ngpu, I_e):
    """aeif_cond_beta above rheobase under constant current for 1 s at the
    configurations' RK5 tolerance 1e-3: the GPU's spike train (FP32 state)
    has the FP64 oracle's spike count, and every spike lies within one
    time step of the FP64 oracle's."""
    log("case fp64_spike_times I_e", I_e)
    spec = {"groups": [{"model": "aeif_cond_beta", "n": 1, "n_port": 2,
                        "params": {"I_e": I_e}, "rk5": {"rk5_tol": 1e-3}}]}
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.ActivateRecSpikeTimes(groups[0], 1000)
    ngpu.Simulate(1000.0)
    eng = np.array(ngpu.GetRecSpikeTimes(0), dtype=np.float64)
    n64, _ = nb.build_oracle(spec, "f64")
    n64.run(10000)
    orc = nb.oracle_times_by_neuron(n64, 0.1, 0, 1)[0]
    log("spikes gpu", eng.size, "fp64", orc.size)
    assert eng.size == orc.size and eng.size >= 5
    d = np.abs(eng - orc)
    isi_g, isi_o = np.diff(eng), np.diff(orc)
    log("max |dt| ms", d.max(), "mean ISI gpu", isi_g.mean(), "fp64",
        isi_o.mean())
    assert d.max() <= 0.1 + 1e-4
    assert abs(isi_g.mean() - isi_o.mean()) <= 1e-3 * isi_o.mean()


def test_spike_detector_delays(ngpu):
    """python/test/test_spike_detector.py: a spike recorded at row i with
    delay 10/20/30 ms appears at detector rows i+101/201/301 with heights
    1/2/3."""
    log("case spike_detector_delays")
    neuron = ngpu.Create("aeif_cond_beta", 3)
    ngpu.SetStatus(neuron, {"I_e": 1000.0})
    det = ngpu.Create("spike_detector")
    for k, (w, d) in enumerate([(1.0, 10.0), (2.0, 20.0), (3.0, 30.0)]):
        ngpu.Connect([neuron[k]], det, {"rule": "one_to_one"},
                     {"weight": w, "delay": d, "receptor": 0})
    rn = ngpu.CreateRecord("", ["spike"], [neuron[0]], [0])
    rd = ngpu.CreateRecord("", ["spike_height"], [det[0]], [0])
    ngpu.Simulate()
    sn = ngpu.GetRecordDataArray(rn)[:, 1]
    sd = ngpu.GetRecordDataArray(rd)[:, 1]
    n_checked = 0
    for i in range(len(sn) - 400):
        if sn[i] > 0.5:
            assert abs(sd[i + 101] - 1.0) < 1e-6
            assert abs(sd[i + 201] - 2.0) < 1e-6
            assert abs(sd[i + 301] - 3.0) < 1e-6
            n_checked += 1
    log("spikes checked", n_checked)
    assert n_checked > 10


def _run_both(ngpu, spec, steps, node_hi):
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    ngpu.Simulate(steps * spec.get("dt", 0.1))
    eng = sorted(nb.engine_spikes(ngpu, groups[0]))
    net, _ = nb.build_oracle(spec, "f32")
    net.run(steps)
    orc = sorted(nb.oracle_spikes(net, spec.get("dt", 0.1), 0, node_hi))
    return eng, orc, net


def test_network_dyadic_bit_exact(ngpu):
    """Reference Brunel test network with dyadic weights (all synaptic sums
    exact in any order): the GPU spike train equals the FP32 oracle's spike
    for spike over 200 ms."""
    log("case network_dyadic")
    spec = nb.brunel_spec(w_ex=0.0625, w_in=0.375, w_p=0.375, tol=1e-3)
    eng, orc, net = _run_both(ngpu, spec, 2000, 1000)
    log("spikes engine", len(eng), "oracle", len(orc), net.counters())
    assert len(orc) > 100
    assert eng == orc


def test_network_brunel_rate(ngpu):
    """Reference Brunel test network as specified: mean rate within 2 % of
    the FP32 oracle and within 10 % of the reference's expected 30.78
    spikes/s (different random graph, so only the statistic carries)."""
    log("case network_brunel_rate")
    spec = nb.brunel_spec()
    eng, orc, net = _run_both(ngpu, spec, 10000, 1000)
    r_eng = len(eng) / 1000.0
    r_orc = len(orc) / 1000.0
    log("rate engine", r_eng, "oracle", r_orc, net.counters())
    assert abs(r_eng - r_orc) <= 0.02 * r_orc
    assert abs(r_eng - 30.78) <= 0.1 * 30.78


def test_c0_subthreshold_bit_exact(ngpu):
    """Small configuration exactly as specified (silent: the Poisson drive
    stays below rheobase): every membrane potential after 200 ms equals the
    FP32 oracle's bit for bit (Threefry Poisson streams, fold, RK5)."""
    log("case c0_literal")
    spec = nb.c0_spec(n=1000, seed=1234)
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.Simulate(200.0)
    V = np.array(ngpu.GetNeuronVar(0, 1000, "V_m"), dtype=np.float32)
    g1 = np.array(ngpu.GetNeuronVar(0, 1000, "g1"), dtype=np.float32)
    net, og = nb.build_oracle(spec, "f32")
    net.run(2000)
    Vo = og[0].get("V_m")
    g1o = og[0].get("g1", 0)
    log("max |dV|", np.max(np.abs(V - Vo)), "V range", V.min(), V.max())
    assert np.array_equal(V, Vo)
    assert np.array_equal(g1[:, 0], g1o)


@pytest.mark.parametrize("case", [str(i) for i in range(1, 11)])
def test_stdp_reference_cases(ngpu, case):
    """python/test/test_stdp/cases/case<k>.py on the GPU: parrot neurons,
    dendritic delay, STDP synapse; the final weight equals the FP32
    oracle's bit for bit and the reference's closed form to 1e-6."""
    log("case stdp", case)
    cs = nb.load_stdp_cases()
    c = cs["cases"][case]
    spec = nb.stdp_case_spec(c, cs["common"])
    nb.build_engine(ngpu, spec)
    ngpu.Simulate(cs["common"]["sim_time"])
    (st,) = ngpu.GetStatus(ngpu.GetConnections(0, 1))
    w = np.float32(st["weight"])
    net, _ = nb.build_oracle(spec, "f32")
    net.run(int(round(cs["common"]["sim_time"] / 0.1)))
    src, tgt, port, d, wo = net.connections()
    (w_orc,) = wo[(src == 0) & (tgt == 1)]
    want = nb.stdp_closed_form(cs["common"]["weight"], c["Dt"], cs["common"])
    log("w gpu", float(w), "oracle", float(w_orc), "closed form", want)
    assert w == w_orc
    assert abs(float(w) - want) <= 1e-6 * want


@pytest.mark.parametrize("graph_steps", [0, 16])
def test_stdp_network_bit_exact(ngpu, graph_steps):
    """Plastic network (2000 STDP synapses, dendritic delay 1.5 ms) over
    4 s, past the 16-bit spike-time wrap at step 32768: parrot spike trains
    and every final plastic weight equal the FP32 oracle's, with steps
    launched from the host and replayed from captured hipGraphs (steps
    around the wrap run from the host)."""
    log("case stdp_network graph_steps", graph_steps)
    spec = nb.stdp_net_spec()
    nb.build_engine(ngpu, spec)
    ngpu.SetKernelStatus("graph_steps", graph_steps)
    ngpu.ActivateRecSpikeTimes(ngpu.NodeSeq(200, 100), 2000)
    ngpu.Simulate(4000.0)
    eng = []
    for i in range(200, 300):
        eng += [(i, t) for t in ngpu.GetRecSpikeTimes(i)]
    cid = ngpu.GetConnections(ngpu.NodeSeq(0, 200), ngpu.NodeSeq(200, 100))
    stl = ngpu.GetStatus(cid)
    wg = {(d["source"], d["target"], d["port"], round(d["delay"], 3)):
          np.float32(d["weight"]) for d in stl if d["port"] == 1}
    net, _ = nb.build_oracle(spec, "f32")
    net.run(40000, 8)
    orc = nb.oracle_spikes(net, 0.1, 200, 300)
    log("parrot spikes engine", len(eng), "oracle", len(orc))
    assert sorted(eng) == sorted(orc)
    src, tgt, port, d, wo = net.connections()
    n_eq = 0
    for s, t, p, dd, w in zip(src, tgt, port, d, wo):
        if p != 1:
            continue
        key = (int(s), int(t), 1, round(float(np.float32(0.1) * (dd + 1)), 3))
        assert wg[key] == w, key
        n_eq += 1
    log("plastic weights equal", n_eq)
    assert n_eq == 2000


def test_network_statistics_vs_fp64(ngpu):
    """Brunel network of the reference test at 10 000 neurons (order 2000),
    500 ms: per-population firing rate and mean CV of the inter-spike
    intervals within 2 % of the FP64 oracle, mean pairwise spike-count
    correlation (5 ms bins) within 0.02 absolute.  Spike trains diverge
    (chaotic dynamics, different rounding), so only the statistics are
    compared; the first 100 ms are discarded."""
    log("case network_statistics")
    order, steps = 2000, 5000
    spec = nb.brunel_spec(order=order)
    n, ne = 5 * order, 4 * order
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.ActivateRecSpikeTimes(groups[0], 1000)
    ngpu.Simulate(steps * 0.1)
    eng = [np.array(ngpu.GetRecSpikeTimes(i), dtype=np.float64)
           for i in range(n)]
    net, _ = nb.build_oracle(spec, "f64")
    net.run(steps, 16)
    orc = nb.oracle_times_by_neuron(net, 0.1, 0, n)
    t1 = steps * 0.1
    for name, sl, pairs in (("E", slice(0, ne), 500),
                            ("I", slice(ne, n), 200)):
        k = sl.stop - sl.start
        se = nb.spike_stats(eng[sl], k, 100.0, t1, n_pairs=pairs)
        so = nb.spike_stats(orc[sl], k, 100.0, t1, n_pairs=pairs)
        log(name, "gpu", se, "fp64 oracle", so)
        assert abs(se["rate"] - so["rate"]) <= 0.02 * so["rate"]
        assert abs(se["cv_isi"] - so["cv_isi"]) <= 0.02 * so["cv_isi"]
        assert abs(se["corr"] - so["corr"]) <= 0.02


@pytest.mark.parametrize("settings", [
    {"delivery_mode": 0},
    {"delivery_mode": 1},
    {"delivery_mode": 1, "delivery_bin_capacity": 50000},
    {"delivery_mode": 2, "delivery_auto_events": 100000},
    {"delivery_mode": 3},
    {"delivery_mode": 3, "delivery_tile_split": 3},
    {"delivery_mode": 3, "delivery_tile_fixed": 1,
     "delivery_tile_shift": 14},
])
def test_c2_iaf_network_bit_exact(ngpu, settings):
    """c2-shaped iaf_psc_exp network at 10 % size (7717 neurons, 5e6
    synapses, 8 Poisson populations), weights rounded to whole pA so that
    every synaptic sum is exact: spike trains equal the FP32 oracle's spike
    for spike over 100 ms.  Run with per-event atomics, with binned
    delivery (about 1e5 events per bin per step, so bins span several
    reduction chunks), with a binned-event buffer too small for the
    busy steps (those fall back to atomics) and in the automatic mode with
    a threshold the network crosses after its first few milliseconds."""
    from neurongpu_amd import workloads as wl
    log("case c2_scaled", settings)
    spec = wl.c2_spec(scale=0.1, integer_weights=True)
    n = spec["groups"][0]["n"]
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in settings.items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 200)
    ngpu.Simulate(100.0)
    eng = sorted(nb.engine_spikes(ngpu, groups[0]))
    net, _ = nb.build_oracle(spec, "f32")
    net.run(1000, 16)
    orc = sorted(nb.oracle_spikes(net, 0.1, 0, n))
    log("spikes engine", len(eng), "oracle", len(orc), net.counters())
    assert len(orc) > 10000
    assert eng == orc


def _wide_iaf_spec(n=540000, k_in=10, seed=11):
    """n iaf_psc_exp neurons (80 % excitatory), k_in uniformly drawn sources
    each, whole-pA weights (+88 pA on port 0 from E, -352 pA on port 1 from
    I), delays 1.0-2.0 ms in 0.1 ms steps, one 16 kHz Poisson generator at
    88 pA on every neuron.  2 ports x 540000 nodes = 132 accumulator bins of
    8192 entries."""
    rng = np.random.default_rng(seed)
    ne = (4 * n) // 5
    tgt = np.repeat(np.arange(n, dtype=np.int32), k_in)
    src = rng.integers(0, n, n * k_in).astype(np.int32)
    d = (10 + rng.integers(0, 11, n * k_in)) * 0.1
    exc = src < ne
    conns = [
        {"src": src[exc], "tgt": tgt[exc], "weight": np.float32(88.0),
         "delay": d[exc], "port": 0},
        {"src": src[~exc], "tgt": tgt[~exc], "weight": np.float32(-352.0),
         "delay": d[~exc], "port": 1},
        {"src": np.full(n, n, np.int32), "tgt": np.arange(n, dtype=np.int32),
         "weight": np.float32(88.0), "delay": np.float32(0.1), "port": 0},
    ]
    return {"dt": 0.1, "seed": seed,
            "groups": [{"model": "iaf_psc_exp", "n": n,
                        "vars": {"V_m_rel": rng.uniform(0.0, 15.0, n)
                                 .astype(np.float32)}},
                       {"model": "poisson_generator", "n": 1}],
            "poisson": [{"node": n, "rate": 16000.0}],
            "conns": conns}


def test_binned_delivery_one_chunk_per_bin(ngpu):
    """Binned delivery with more than 128 accumulator bins of about equal
    load, where each bin is reduced by one block and no partial sums are
    written: 540000 iaf_psc_exp neurons, 5.4e6 synapses with whole-pA
    weights, spike trains equal to the FP32 oracle's over 20 ms."""
    log("case binned_one_chunk_per_bin")
    spec = _wide_iaf_spec()
    n = spec["groups"][0]["n"]
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.SetKernelStatus("delivery_mode", 1)
    ngpu.ActivateRecSpikeTimes(groups[0], 50)
    ngpu.Simulate(20.0)
    eng = sorted(nb.engine_spikes(ngpu, groups[0]))
    net, _ = nb.build_oracle(spec, "f32")
    net.run(200, 16)
    orc = sorted(nb.oracle_spikes(net, 0.1, 0, n))
    log("spikes engine", len(eng), "oracle", len(orc))
    assert len(orc) > 10000
    assert eng == orc


def test_c3_stdp_network_statistics(ngpu):
    """c3-shaped plastic network at 2 % size (4000 aeif_cond_alpha neurons,
    2e6 synapses, STDP on E -> E), 500 ms: population rates within 2 % of
    the FP64 oracle; the final E -> E weights stay in [0, Wmax] and their
    mean and standard deviation agree with the oracle's to 0.5 %."""
    from neurongpu_amd import workloads as wl
    log("case c3_scaled")
    spec = wl.c3_spec(scale=0.02)
    n = spec["groups"][0]["n"]
    ne = spec["populations"][0][1]
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.ActivateSpikeCount(groups[0])
    ngpu.Simulate(500.0)
    cnt = np.array(ngpu.GetNeuronIntVar(0, n, "spike_count"), dtype=float)
    st = ngpu.GetStatus(ngpu.GetConnections(ngpu.NodeSeq(0, ne),
                                            ngpu.NodeSeq(0, ne)))
    wg = np.array([d["weight"] for d in st])
    net, _ = nb.build_oracle(spec, "f64")
    net.run(5000, 16)
    s_, nd, ct = net.spikes()
    src, tgt, port, d, wo = net.connections()
    wo = wo[(src < ne) & (tgt < ne)]
    for name, (a, b) in zip(("E", "I"), spec["populations"]):
        r_g = cnt[a:b].sum() / (b - a) / 0.5
        m = (nd >= a) & (nd < b)
        r_o = ct[m].sum() / (b - a) / 0.5
        log(name, "rate gpu", r_g, "fp64 oracle", r_o)
        assert abs(r_g - r_o) <= 0.02 * r_o
    log("EE weights gpu mean %.6f std %.6f, oracle mean %.6f std %.6f"
        % (wg.mean(), wg.std(), wo.mean(), wo.std()))
    assert wg.size == wo.size
    assert wg.min() >= 0.0 and wg.max() <= 0.1 + 1e-7
    assert abs(wg.mean() - wo.mean()) <= 0.005 * wo.mean()
    assert abs(wg.std() - wo.std()) <= 0.005 * wo.std()


def _ensemble_instances():
    return [nb.ensemble_instance_spec(n, seed)
            for n, seed in ((300, 11), (700, 12), (450, 13))]


def _run_spikes(ngpu, spec, t_ms, settings=None):
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in (settings or {}).items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    ngpu.Simulate(t_ms)
    return nb.engine_spikes(ngpu, groups[0])


def test_ensemble_routing_matches_instances_alone(ngpu):
    """Three AdEx networks of different sizes batched into one node group
    (one launch per kernel per step) spike exactly as each network run
    alone: same synapses, same delay slots, same inputs, node indices
    shifted by the instance offset (dyadic weights make every sum exact)."""
    log("case ensemble_routing")
    inst = _ensemble_instances()
    ens = sorted(_run_spikes(ngpu, nb.ensemble_spec(inst), 300.0))
    off = 0
    total = 0
    for x in inst:
        alone = sorted((i + off, t) for i, t in
                       _run_spikes(ngpu, nb.ensemble_spec([x]), 300.0))
        part = sorted(s for s in ens if off <= s[0] < off + x["n"])
        log("instance", x["n"], "spikes alone", len(alone), "batched",
            len(part))
        assert len(alone) > 50
        assert part == alone
        total += len(alone)
        off += x["n"]
    assert len(ens) == total


@pytest.mark.parametrize("settings", [
    {"rk5_two_pass": 0},
    {"rk5_budget": 3, "rk5_pass2_grid": 64},
    {"delivery_grid": 256, "delivery_tile_max": 8, "fuse_delivery": 0},
    {"delivery_grid": 4096, "delivery_tile_max": 1},
    {"delivery_mode": 1},
    {"delivery_mode": 1, "step_overlap": 1},
    {"delivery_mode": 1, "step_overlap": 1024},
    {"delivery_mode": 1, "rk5_budget": 1, "rk5_pass2_grid": 64},
    {"delivery_mode": 1, "step_overlap": 0},
    {"delivery_mode": 1, "step_overlap": 0, "delivery_streams": 1},
    {"delivery_mode": 1, "delivery_grid": 64, "delivery_tile_max": 4},
    {"delivery_mode": 3},
    {"delivery_mode": 3, "delivery_tile_split": 1, "step_overlap": 0},
    {"delivery_mode": 3, "delivery_tile_split": 7, "delivery_tile_max": 2},
    {"delivery_mode": 3, "delivery_tile_shift": 10},
    {"delivery_mode": 3, "delivery_tile_shift": 11, "delivery_tile_split": 1},
    {"delivery_mode": 3, "delivery_tile_fixed": 1,
     "delivery_tile_shift": 14},
    {"delivery_mode": 3, "delivery_tile_fixed": 1, "delivery_tile_split": 3,
     "delivery_tile_shift": 14},
])
def test_launch_shapes_do_not_change_results(ngpu, settings):
    """The autotuner's choices (RK5 two-pass budget and grid, delivery grid,
    tile size and mode, fusion) leave every spike unchanged."""
    log("case launch_shapes", settings)
    spec = nb.ensemble_spec(_ensemble_instances())
    ref = sorted(_run_spikes(ngpu, spec, 200.0))
    got = sorted(_run_spikes(ngpu, spec, 200.0, settings))
    assert len(ref) > 100
    assert got == ref


def _run_spikes_calls(ngpu, spec, calls, t_ms, settings):
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in settings.items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    for _ in range(calls):
        ngpu.Simulate(t_ms)
        ngpu.GetHotPathStats()   # host sync and counter read between calls
    return nb.engine_spikes(ngpu, groups[0])


@pytest.mark.parametrize("settings", [
    {"graph_steps": 16},
    {"graph_steps": 6},
    {"graph_steps": 16, "delivery_mode": 1},
    {"graph_steps": 4, "rk5_two_pass": 0},
    {"graph_steps": 16, "delivery_mode": 3},
])
def test_graph_steps_match_eager(ngpu, settings):
    """Steps replayed from captured hipGraphs (the step clock advanced on
    the device) spike exactly as steps launched one by one from the host,
    over four Simulate calls of 50 ms each (each call starts with an eager
    step; step counts that are not a multiple of the graph length end with
    eager steps)."""
    log("case graph_steps", settings)
    spec = nb.ensemble_spec(_ensemble_instances())
    ref = sorted(_run_spikes_calls(ngpu, spec, 4, 50.0, {"graph_steps": 0}))
    got = sorted(_run_spikes_calls(ngpu, spec, 4, 50.0, settings))
    assert len(ref) > 100
    assert got == ref


def test_step_overlap_over_calls(ngpu):
    """Binned delivery with each step's second RK5 pass overlapped with the
    next step (the default), over four Simulate calls with host reads in
    between: spikes identical to steps that end with their second pass."""
    log("case step_overlap_calls")
    spec = nb.ensemble_spec(_ensemble_instances())
    ref = sorted(_run_spikes_calls(ngpu, spec, 4, 50.0, {"step_overlap": 0}))
    got = sorted(_run_spikes_calls(ngpu, spec, 4, 50.0, {"delivery_mode": 1}))
    assert len(ref) > 100
    assert got == ref


def test_step_overlap_transitions(ngpu):
    """Overlapped steps (binned delivery) and steps that end with their
    second pass (atomic delivery) alternate inside one run of the stepping
    API, with a host read of the state in the middle: the first plain step
    after an overlapped one finishes the neurons the overlapped step left
    over, and the spikes equal those of plain steps throughout."""
    log("case step_overlap_transitions")
    spec = nb.ensemble_spec(_ensemble_instances())
    ref = sorted(_run_spikes(ngpu, spec, 200.0, {"step_overlap": 0}))
    groups, _ = nb.build_engine(ngpu, spec)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    ngpu.StartSimulation()
    for k in range(2000):
        if k % 400 == 0:
            ngpu.SetKernelStatus("delivery_mode", 1)
        elif k % 400 == 250:
            ngpu.SetKernelStatus("delivery_mode", 0)
        if k == 1111:
            ngpu.GetStatus(groups[0][0], "V_m")
        ngpu.SimulationStep()
    ngpu.EndSimulation()
    got = sorted(nb.engine_spikes(ngpu, groups[0]))
    assert len(ref) > 100
    assert got == ref


def test_overlapped_steps_then_graph_steps(ngpu):
    """Overlapped binned steps from the stepping API (the last one leaves
    its second pass running), then captured steps switched on with no host
    synchronisation in between, then Simulate: the spikes equal those of
    plain eager steps."""
    log("case overlap_then_graph_steps")
    spec = nb.ensemble_spec(_ensemble_instances())

    def run(first, then):
        groups, _ = nb.build_engine(ngpu, spec)
        ngpu.ActivateRecSpikeTimes(groups[0], 2000)
        for k, v in first.items():
            ngpu.SetKernelStatus(k, v)
        ngpu.StartSimulation()
        for _ in range(700):
            ngpu.SimulationStep()
        for k, v in then.items():
            ngpu.SetKernelStatus(k, v)
        ngpu.Simulate(130.0)
        return sorted(nb.engine_spikes(ngpu, groups[0]))

    ref = run({"step_overlap": 0}, {})
    got = run({"delivery_mode": 1}, {"graph_steps": 16})
    assert len(ref) > 100
    assert got == ref


def _build_rules(ngpu, device):
    """Every rule the device generates, with ranges and node lists, scalar
    and per-connection weights and delays, both sampler paths (n >= 5k:
    rejection; n < 5k: partial Fisher-Yates), a plastic synapse group, and
    an indegree whose sampler does not fit the LDS (left to the host)."""
    rng = np.random.default_rng(5)
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(777)
    ngpu.SetKernelStatus("device_connect", device)
    a = ngpu.Create("aeif_cond_beta", 1200, 2)
    b = ngpu.Create("iaf_psc_exp", 900)
    big = ngpu.Create("parrot_neuron", 20000)
    ngpu.CreateSynGroup("stdp", {})
    n_tot = 1200 + 900 + 20000
    m1 = 900 * 50
    ngpu.Connect(a, b, {"rule": "fixed_indegree", "indegree": 50},
                 {"weight": 0.5, "delay": {"array": rng.integers(1, 30, m1)
                                           .astype(np.float32) * 0.1}})
    ngpu.Connect(a[0:100], a, {"rule": "fixed_indegree", "indegree": 40},
                 {"weight": 0.25, "delay": 1.0, "receptor": 1})
    m3 = 500 * 30
    ngpu.Connect(a[0:500], b, {"rule": "fixed_outdegree", "outdegree": 30},
                 {"weight": {"array": rng.normal(1.0, 0.1, m3)
                             .astype(np.float32)}, "delay": 2.0})
    ngpu.Connect(b[0:60], a[0:60], {"rule": "fixed_outdegree", "outdegree": 20},
                 {"weight": 1.5, "delay": 0.5, "synapse_group": 1})
    src = rng.integers(0, 1200, 300).astype(np.int32)
    tgt = (1200 + rng.integers(0, 900, 300)).astype(np.int32)
    ngpu.Connect(list(map(int, src)), list(map(int, tgt)),
                 {"rule": "fixed_total_number", "total_num": 5000},
                 {"weight": 0.1, "delay": 0.3})
    ngpu.Connect(list(map(int, src[:30])), list(map(int, tgt[:40])),
                 {"rule": "all_to_all"}, {"weight": 0.2, "delay": 0.4})
    ngpu.Connect(list(map(int, src[:70])), list(map(int, tgt[:70])),
                 {"rule": "one_to_one"}, {"weight": 0.3, "delay": 0.6})
    ngpu.Connect(big[0:20000], b[0:2], {"rule": "fixed_indegree",
                                       "indegree": 3500},
                 {"weight": 0.05, "delay": 0.2})
    # weights and delays drawn by the library per connection
    ngpu.Connect(a, b[0:300], {"rule": "fixed_indegree", "indegree": 60},
                 {"weight": {"distribution": "normal", "mu": 1.0,
                             "sigma": 0.3},
                  "delay": {"distribution": "uniform", "low": 0.1,
                            "high": 2.0, "step": 0.1}})
    ngpu.Connect(a[0:200], a[200:400], {"rule": "all_to_all"},
                 {"weight": {"distribution": "uniform", "low": 0.2,
                             "high": 0.6},
                  "delay": {"distribution": "normal_clipped", "mu": 0.5,
                            "sigma": 0.25, "low": 0.1, "high": 1.25},
                  "receptor": 1})
    ngpu.Simulate(1.0)
    return n_tot


def test_device_connect_matches_host(ngpu):
    """Connection rules generated on the device at calibration give the
    synapse store the host generator gives: every connection's source,
    target, port, synapse group, delay and weight, in the same order."""
    log("case device_connect")
    n = _build_rules(ngpu, 0)
    host = ngpu.GetConnectionArrays()
    _build_rules(ngpu, 1)
    dev = ngpu.GetConnectionArrays()
    log("connections", host["source"].size)
    for k in host:
        assert np.array_equal(dev[k], host[k]), k
    host = list(zip(*[host[k] for k in ("source", "target", "port", "syn",
                                        "delay", "weight")]))
    assert len(host) == 900 * 50 + 1200 * 40 + 500 * 30 + 60 * 20 + 5000 + \
        30 * 40 + 70 + 2 * 3500 + 300 * 60 + 200 * 200


def test_graph_steps_stdp_network(ngpu):
    """Plastic network of the bit-exact STDP test (dyadic weights): parrot
    spikes and every plastic weight identical with and without captured
    steps, over two Simulate calls."""
    log("case graph_steps_stdp")
    spec = nb.stdp_net_spec()
    res = []
    for gsteps in (0, 16):
        nb.build_engine(ngpu, spec)
        ngpu.SetKernelStatus("graph_steps", gsteps)
        ngpu.ActivateRecSpikeTimes(ngpu.NodeSeq(200, 100), 2000)
        ngpu.Simulate(700.0)
        ngpu.Simulate(333.3)
        spikes = []
        for i in range(200, 300):
            spikes += [(i, t) for t in ngpu.GetRecSpikeTimes(i)]
        cid = ngpu.GetConnections(ngpu.NodeSeq(0, 200),
                                  ngpu.NodeSeq(200, 100))
        w = sorted((d["source"], d["target"], d["port"], d["weight"])
                   for d in ngpu.GetStatus(cid))
        res.append((sorted(spikes), w))
    assert len(res[0][0]) > 50
    assert res[1][0] == res[0][0]
    assert res[1][1] == res[0][1]


def _rk5_model_net_spec(model, n=300, k=20, seed=7):
    """Small recurrent network of one of the other RK5 models: constant
    drive above rheobase (spread per neuron), fixed in-degree k with
    dyadic weights (every synaptic sum exact in any order) on every port,
    delays 1-4 ms."""
    rng = np.random.default_rng(seed)
    izh = model == "izhikevich_cond_beta"
    n_port = 1 if model == "aeif_psc_delta" else 2
    I_e = rng.uniform(4.0, 12.0, n) if izh else rng.uniform(500.0, 900.0, n)
    params = {"I_e": I_e.astype(np.float32)}
    if model in ("aeif_psc_alpha", "aeif_psc_exp"):
        params["tau_syn"] = [2.0, 5.0]
    if izh:
        params["E_rev"] = [0.0, -85.0]
    src = rng.integers(0, n, n * k)
    tgt = np.repeat(np.arange(n), k)
    port = rng.integers(0, n_port, n * k)
    if izh:
        w = rng.choice([0.0625, 0.125, 0.25], n * k)
    elif model == "aeif_psc_delta":
        w = rng.choice([-0.5, 0.25, 0.5], n * k)
    else:
        w = rng.choice([-8.0, 4.0, 16.0], n * k)
    d = rng.integers(10, 41, n * k) * 0.1
    conns = [{"src": src[port == q].tolist(), "tgt": tgt[port == q].tolist(),
              "weight": w[port == q].tolist(), "delay": d[port == q].tolist(),
              "port": q} for q in range(n_port)]
    return {"groups": [{"model": model, "n": n, "n_port": n_port,
                        "params": params, "rk5": {"rk5_tol": 1e-4}}],
            "conns": conns, "records": []}


@pytest.mark.parametrize("model", ["aeif_psc_alpha", "aeif_psc_exp",
                                   "aeif_psc_delta", "izhikevich_cond_beta"])
def test_rk5_model_network_bit_exact(ngpu, model):
    """Recurrent 300-neuron network of each of the other RK5 models,
    200 ms: spike trains equal the FP32 oracle's spike for spike."""
    log("case rk5_model_network", model)
    spec = _rk5_model_net_spec(model)
    eng, orc, net = _run_both(ngpu, spec, 2000, 300)
    log("spikes engine", len(eng), "oracle", len(orc), net.counters())
    assert len(orc) > 100
    assert eng == orc


def test_izhikevich_trace_bit_exact(ngpu):
    """Single izhikevich_cond_beta neuron under constant current plus
    conductance input on two ports: V_m, u and g traces equal the FP32
    oracle's value for value over 300 ms."""
    log("case izhikevich_trace")
    spec = {"groups": [{"model": "izhikevich_cond_beta", "n": 1,
                        "n_port": 2,
                        "params": {"I_e": 5.0, "E_rev": [0.0, -85.0],
                                   "tau_rise": [1.0, 2.0],
                                   "tau_decay": [5.0, 20.0]}},
                       {"model": "spike_generator", "n": 1}],
            "spike_gen": [{"node": 1, "times": [20.0, 100.0, 180.0]}],
            "conns": [{"src": [1], "tgt": [0], "weight": [0.5],
                       "delay": [1.0], "port": 0},
                      {"src": [1], "tgt": [0], "weight": [0.3],
                       "delay": [30.0], "port": 1}],
            "records": [(0, "V_m", 0), (0, "u", 0), (0, "g", 0),
                        (0, "g", 1)]}
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(300.0)
    cols = [ngpu.GetRecordDataArray(r) for r in recs]
    data = np.column_stack([cols[0][:, 0]] + [c[:, 1] for c in cols])
    net, _ = nb.build_oracle(spec, "f32")
    net.run(3000)
    orc = net.records()
    assert orc.shape == data.shape
    assert np.sum(data[:, 1] >= 29.0) == 0     # reset below threshold
    assert np.array_equal(orc[:, 1:].astype(np.float32), data[:, 1:])


# python/test/test_iaf_psc_exp_g.py, test_iaf_psc_exp_hc.py, test_izh.py
POINT_NEST = {"iaf_psc_exp_g": ("fast_iaf_psc_exp_nest.txt", 800.0, -65.0, 5e-5),
              "iaf_psc_exp_hc": ("fast_iaf_psc_exp_nest.txt", 800.0, -65.0, 5e-5),
              "izhikevich": ("izhikevich_nest.txt", 80.0, 0.0, 5e-3)}


@pytest.mark.parametrize("model", list(POINT_NEST))
def test_point_model_nest_trace(ngpu, model):
    """The reference's golden NEST traces of iaf_psc_exp_g / _hc (5e-5)
    and izhikevich (its tolerance 5e-3); the trace equals the FP32
    oracle's value for value."""
    log("case point_nest", model)
    fname, t_sim, E_L, tol = POINT_NEST[model]
    spec = nb.nest_spec(model)
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(t_sim)
    data = ngpu.GetRecordDataArray(recs[0])
    V = data[:, 1].astype(np.float64) + E_L
    nest = np.loadtxt(os.path.join(GOLDEN, fname), delimiter="\t")
    dV = np.array([V[i * 10 + 20] - nest[i, 1] for i in range(len(nest))])
    rmse = np.std(dV) / abs(np.mean(V))
    log("rmse", rmse, "tol", tol)
    assert rmse <= tol
    net, _ = nb.build_oracle(spec, "f32")
    net.run(int(round(t_sim / 0.1)))
    orc = net.records()
    assert orc.shape == data.shape
    assert np.array_equal(orc[:, 1].astype(np.float32), data[:, 1])


def _point_net_spec(model, n=500, k=40, seed=11):
    """Recurrent network of one fixed-step model, 80 % excitatory, driven
    by per-neuron I_e; integer weights so that every input sum is exact."""
    rng = np.random.default_rng(seed)
    n_e = int(0.8 * n)
    params = {}
    if model.startswith("iaf"):
        params["I_e"] = rng.uniform(360.0, 420.0, n).astype(np.float32)
        w_e, w_i = 64.0, -256.0
    elif model == "izhikevich":
        params["I_e"] = rng.uniform(4.0, 12.0, n).astype(np.float32)
        w_e, w_i = 1.0, -4.0
    else:
        params["I_e"] = rng.uniform(4.0, 12.0, n).astype(np.float32)
        w_e, w_i = 2.0, -8.0
    src = rng.integers(0, n, n * k)
    tgt = np.repeat(np.arange(n), k)
    w = np.where(src < n_e, w_e, w_i).astype(np.float32)
    d = rng.integers(10, 41, n * k) * 0.1
    return {"groups": [{"model": model, "n": n, "params": params}],
            "conns": [{"src": src, "tgt": tgt, "weight": w, "delay": d,
                       "port": 0}],
            "records": []}


@pytest.mark.parametrize("model", ["iaf_psc_exp_g", "iaf_psc_exp_hc",
                                   "izhikevich", "izhikevich_psc_exp",
                                   "izhikevich_psc_exp_2s",
                                   "izhikevich_psc_exp_5s"])
def test_point_model_network_bit_exact(ngpu, model):
    """500-neuron recurrent network of each fixed-step model, 200 ms:
    spike trains equal the FP32 oracle's spike for spike."""
    log("case point_network", model)
    spec = _point_net_spec(model)
    eng, orc, net = _run_both(ngpu, spec, 2000, 500)
    log("spikes engine", len(eng), "oracle", len(orc))
    assert len(orc) > 200
    assert eng == orc


@pytest.mark.parametrize("settings", [
    {"delivery_mode": 3},
    {"delivery_mode": 3, "delivery_tile_fixed": 1,
     "delivery_tile_shift": 14},
    {"delivery_mode": 3, "delivery_tile_fixed": 1, "delivery_tile_split": 3,
     "delivery_tile_shift": 14},
    {"delivery_mode": 3, "delivery_tile_fixed": 1, "delivery_tile_shift": 10},
])
def test_tiled_delivery_spike_heights(ngpu, settings):
    """Tiled delivery with spike heights other than 1: the 500-neuron
    iaf_psc_exp_g network plus four spike generators onto every neuron,
    with heights 1, 3, 0.5 and 2000 (above the fixed-point tile's height
    bound, so added with a float atomic beside the tile's exact sums);
    integer and dyadic weights keep every sum exact, so the spike trains
    equal the FP32 oracle's spike for spike over 200 ms."""
    log("case tiled_spike_heights", settings)
    n = 500
    spec = _point_net_spec("iaf_psc_exp_g", n=n)
    spec["groups"].append({"model": "spike_generator", "n": 4})
    rng = np.random.default_rng(3)
    heights = [1.0, 3.0, 0.5, 2000.0]
    spec["spike_gen"] = []
    for g, h in enumerate(heights):
        times = np.sort(rng.choice(np.arange(5, 1900), 40, replace=False))
        times = [float(t) * 0.1 for t in times]
        spec["spike_gen"].append({"node": n + g, "times": times,
                                  "heights": [h] * len(times)})
        w = rng.integers(-8, 9, n).astype(np.float32) * np.float32(
            0.25 if h > 100 else 4.0)
        spec["conns"].append({"src": np.full(n, n + g), "tgt": np.arange(n),
                              "weight": w,
                              "delay": rng.integers(1, 31, n) * 0.1,
                              "port": 0})
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in settings.items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    ngpu.Simulate(200.0)
    eng = sorted(nb.engine_spikes(ngpu, groups[0]))
    net, _ = nb.build_oracle(spec, "f32")
    net.run(2000)
    orc = sorted(nb.oracle_spikes(net, 0.1, 0, n))
    log("spikes engine", len(eng), "oracle", len(orc))
    assert len(orc) > 200
    assert eng == orc


def test_tiled_fixed_point_sums_are_deterministic(ngpu):
    """Tiled delivery with weights that are not dyadic (float sums depend on
    the order of the adds): the fixed-point tile gives the same V_m, bit
    for bit, on two runs and with 1 or 5 blocks per tile, and stays within
    float rounding of the per-event atomic delivery."""
    log("case tiled_fixed_deterministic")
    spec = nb.ensemble_spec(_ensemble_instances())
    rng = np.random.default_rng(5)
    for c in spec["conns"]:
        w = np.asarray(c["weight"], np.float32)
        if w.ndim == 1:
            c["weight"] = w * rng.uniform(0.9, 1.1, w.size).astype(np.float32)

    def run(settings):
        groups, _ = nb.build_engine(ngpu, spec)
        for k, v in settings.items():
            ngpu.SetKernelStatus(k, v)
        ngpu.Simulate(10.0)
        n = groups[0].n
        return (np.asarray(ngpu.GetNeuronVar(0, n, "V_m"), np.float32),
                np.asarray(ngpu.GetNeuronVar(0, n, "g1"), np.float32))

    fx = {"delivery_mode": 3, "delivery_tile_fixed": 1,
          "delivery_tile_shift": 14}
    a = run(dict(fx, delivery_tile_split=1))
    b = run(dict(fx, delivery_tile_split=5))
    c = run(dict(fx, delivery_tile_split=5))
    z = run({"delivery_mode": 0})
    for x, y in ((a, b), (b, c)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    log("max |V_m - atomic|", float(np.max(np.abs(a[0] - z[0]))),
        "max |g1 - atomic|", float(np.max(np.abs(a[1] - z[1]))))
    np.testing.assert_allclose(a[0], z[0], rtol=0, atol=1e-3)
    np.testing.assert_allclose(a[1], z[1], rtol=1e-4, atol=1e-6)


def test_ext_neuron_bit_exact(ngpu):
    """ext_neuron (src/ext_neuron.cu:23-39) fed by a spike generator
    (python/test/test_ext_neuron.py's synapses) and a 5 kHz Poisson
    generator on ports 1 and 2: every port_value equals the FP32 oracle's,
    and GetExtNeuronInputSpikes (src/neurongpu.cu:1540-1573) returns the
    non-zero ports of the current step."""
    log("case ext_neuron")
    spec = nb.ext_spec([1.0, 2.0, 0.5])
    spec["groups"].append({"model": "poisson_generator", "n": 1})
    spec["poisson"] = [{"node": 5, "rate": 5000.0}]
    for p, w in ((1, 0.25), (2, 0.5)):
        spec["conns"].append({"src": [5], "tgt": [0], "weight": [w],
                              "delay": [0.3], "port": p})
    _, recs = nb.build_engine(ngpu, spec)
    n_ext, _, _ = ngpu.GetExtNeuronInputSpikes(include_zeros=True)
    assert n_ext == [0, 0, 0]
    ngpu.Simulate(51.2)
    nodes, ports, hs = ngpu.GetExtNeuronInputSpikes()
    ngpu.Simulate(748.8)
    data = np.column_stack([ngpu.GetRecordDataArray(r)[:, 1] for r in recs])
    # the oracle in the same two calls: spike_generator indexes spikes by
    # the step within the Simulate call (src/neurongpu.cu:482)
    net, _ = nb.build_oracle(spec, "f32")
    net.run(512)
    net.run(7488)
    orc = net.records()[:, 1:].astype(np.float32)
    assert orc.shape == data.shape
    assert np.array_equal(orc, data)
    assert np.count_nonzero(data[:, 1]) > 1000     # the Poisson input
    row = 512                                       # after Simulate(51.2)
    want = [(0, p, float(orc[row, p])) for p in range(3) if orc[row, p] != 0]
    assert list(zip(nodes, ports, hs)) == want
    assert (0, 0, float(np.float32(0.1))) in want


@pytest.mark.parametrize("launch", [{"delivery_mode": 0},
                                    {"delivery_mode": 1},
                                    {"delivery_mode": 3},
                                    {"delivery_mode": 3,
                                     "delivery_tile_fixed": 1,
                                     "delivery_tile_shift": 14},
                                    {"graph_steps": 16}])
def test_legacy_poisson_network_bit_exact(ngpu, launch):
    """Network of the reference's c++/examples/test_error.cpp (1000
    aeif_cond_beta neurons, CreatePoissonGenerator at 20 kHz, one node per
    neuron), dyadic weights, 100 ms: the neurons' and the generators' spike
    trains equal the FP32 oracle's spike for spike, with atomic, binned
    and tiled delivery (generator spikes have heights above 1).  Generator
    spikes carry the time of the step before (src/neurongpu.cu:461-470)."""
    log("case legacy_poisson_network", launch)
    spec = nb.legacy_brunel_spec()
    steps, n = 1000, 1000
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in launch.items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 2000)
    ngpu.ActivateRecSpikeTimes(groups[1], steps + 10)
    ngpu.Simulate(steps * 0.1)
    eng = sorted(nb.engine_spikes(ngpu, groups[0]))
    eng_gen = sorted(nb.engine_spikes(ngpu, groups[1]))
    net, _ = nb.build_oracle(spec, "f32")
    net.run(steps)
    orc = sorted(nb.oracle_spikes(net, 0.1, 0, n))
    st, nd, ct = net.spikes()
    dt64 = np.float64(np.float32(0.1))
    orc_gen = sorted((int(d), float(np.float32(dt64 * s)))
                     for s, d in zip(st, nd) if d >= n)
    log("spikes engine", len(eng), len(eng_gen), "oracle", len(orc),
        len(orc_gen))
    assert len(orc) > 100 and len(orc_gen) > 0.8 * n * steps
    assert eng_gen == orc_gen
    assert eng == orc


@pytest.mark.parametrize("model", ["user_m1", "user_m2"])
def test_user_models_match_aeif_cond_beta(ngpu, model):
    """user_m1 / user_m2 (aeif_cond_beta under other names in the
    reference) give the aeif_cond_beta trace value for value."""
    log("case user_model", model)
    spec = nb.nest_spec("aeif_cond_beta")
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(500.0)
    ref = ngpu.GetRecordDataArray(recs[0])
    spec["groups"][0]["model"] = model
    _, recs = nb.build_engine(ngpu, spec)
    ngpu.Simulate(500.0)
    got = ngpu.GetRecordDataArray(recs[0])
    assert np.array_equal(got, ref)
