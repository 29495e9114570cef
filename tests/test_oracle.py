"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Each oracle function is checked against something it does not compute
itself: the reference's golden vectors (NEST traces, STDP scenarios,
spike-detector timing), published known-answer tests (Philox, Threefry), library
routines (scipy's Poisson CDF, matrix exponential, root finder), closed
forms (beta-peak normalisation, STDP rule, AdEx fixed point) and
convergence order.  Runs on the CPU only."""
import math
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.optimize
import scipy.stats

import netbuild as nb
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------------
# Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32_10)
# ---------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2,
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
     [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, want):
    assert list(po.philox4x32_10(ctr, key)) == want


# ---------------------------------------------------------------------
# Threefry2x32-20 (the Poisson draws): Random123 known-answer vectors
# (kat_vectors, threefry2x32_20)
# ---------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0], [0, 0], [0x6b200159, 0x99ba4efe]),
    ([0xffffffff] * 2, [0xffffffff] * 2, [0x1cb996fc, 0xbb002be7]),
    ([0x243f6a88, 0x85a308d3], [0x13198a2e, 0x03707344],
     [0xc4923a9c, 0x483df7a0]),
])
def test_threefry_kat(ctr, key, want):
    assert list(po.threefry2x32_20(ctr, key)) == want


# ---------------------------------------------------------------------
# Poisson sampler: table of 32-bit CDF thresholds + inversion
# ---------------------------------------------------------------------
@pytest.mark.parametrize("lam", [1e-4, 0.01, 0.1, 0.8, 2.0, 8.0, 40.0])
def test_poisson_table_matches_cdf(lam):
    tab = po.poisson_table(lam)
    assert tab is not None
    k = np.arange(len(tab))
    cdf = scipy.stats.poisson.cdf(k, lam)
    # tab[k] = ceil(2^32 F(k)) - 1, clamped to the u32 range
    want = np.minimum(np.ceil(cdf * 2.0 ** 32) - 1, 2.0 ** 32 - 1)
    assert np.all(np.abs(tab.astype(np.float64) - want) <= 2.0)
    assert tab[-1] == 0xFFFFFFFF
    assert np.all(np.diff(tab.astype(np.int64)) >= 0)


def test_poisson_sample_boundaries():
    L = po.lib()
    tab = po.poisson_table(0.8)
    t = np.ascontiguousarray(tab, dtype=np.uint32)
    for k in range(len(tab) - 1):
        if tab[k] == tab[k + 1]:
            continue
        assert L.orc_poisson_sample(t.ctypes.data, int(tab[k])) == k
        assert L.orc_poisson_sample(t.ctypes.data, int(tab[k]) + 1) == k + 1
    assert L.orc_poisson_sample(t.ctypes.data, 0) == 0
    assert L.orc_poisson_sample(t.ctypes.data, 0xFFFFFFFF) <= len(tab) - 1


def test_poisson_zero_and_too_large_rate():
    L = po.lib()
    t0 = np.ascontiguousarray(po.poisson_table(0.0), dtype=np.uint32)
    for u in (0, 12345, 0xFFFFFFFF):
        assert L.orc_poisson_sample(t0.ctypes.data, u) == 0
    assert po.poisson_table(400.0) is None      # table longer than 256


def test_poisson_counts_distribution():
    """Counts drawn with the Threefry counter layout of the direct
    connections (counter (index, step), key (seed ^ tag, seed_hi)) follow
    Poisson(lambda) (chi-square, p > 1e-4)."""
    L = po.lib()
    lam = 0.8
    t = np.ascontiguousarray(po.poisson_table(lam), dtype=np.uint32)
    n = 200000
    counts = np.zeros(12, dtype=np.int64)
    for i in range(n):
        u = int(po.threefry2x32_20([i % 1000, i // 1000],
                                   [1234 ^ 0x504F4953, 0])[0])
        c = L.orc_poisson_sample(t.ctypes.data, u)
        counts[min(c, 11)] += 1
    pmf = scipy.stats.poisson.pmf(np.arange(11), lam)
    exp = np.append(pmf, 1.0 - pmf.sum()) * n
    mask = exp > 5
    chi2 = np.sum((counts[mask] - exp[mask]) ** 2 / exp[mask])
    p = scipy.stats.chi2.sf(chi2, mask.sum() - 1)
    assert p > 1e-4


# ---------------------------------------------------------------------
# CreatePoissonGenerator nodes (src/poisson.cu:27-54, src/neurongpu.cu:
# 245-265): every step each node sends a spike of height n ~ Poisson(lambda)
# unless n == 0 or n > 5 lambda, lambda = rate * dt / 1000.
# ---------------------------------------------------------------------
def _legacy_run(rate, n, steps, dt=0.1, seed=777):
    net = po.Net(dt=dt, seed=seed)
    g = net.create("legacy_poisson_generator", n)
    net.legacy_rate(g, rate)
    net.calibrate()
    net.run(steps)
    st, nd, ct = net.spikes()
    return st, nd, ct, net.spike_heights()


def test_legacy_poisson_heights_truncated_poisson():
    """lambda = 2 (20 kHz at 0.1 ms): heights 1..10 with the Poisson pmf,
    no spike with probability p(0) + p(k > 10) (chi-square, p > 1e-4)."""
    n, steps = 500, 400
    st, nd, ct, h = _legacy_run(20000.0, n, steps)
    assert np.all(ct == 1)
    assert np.all(h == np.round(h)) and h.min() >= 1 and h.max() <= 10
    counts = np.bincount(h.astype(np.int64), minlength=11)[:11]
    counts[0] = n * steps - len(h)
    lam = float(np.float32(np.float32(20000.0) * np.float32(0.1)) / 1000.0)
    pmf = scipy.stats.poisson.pmf(np.arange(11), lam)
    pmf[0] += scipy.stats.poisson.sf(10, lam)
    exp = pmf * n * steps
    mask = exp > 5
    chi2 = np.sum((counts[mask] - exp[mask]) ** 2 / exp[mask])
    assert scipy.stats.chi2.sf(chi2, mask.sum() - 1) > 1e-4
    # every (node, step) at most once, steps in range
    key = st * n + (nd - nd.min())
    assert len(np.unique(key)) == len(key)
    assert st.min() >= 0 and st.max() < steps


def test_legacy_poisson_cut_at_five_lambda():
    """lambda = 0.2 (2 kHz): 5 lambda == 1 in float, so only counts of one
    are kept (the cut is 'count > 5 lambda'); lambda = 0.1: 5 lambda < 1
    and the generator never fires."""
    n, steps = 1000, 300
    st, nd, ct, h = _legacy_run(2000.0, n, steps)
    assert np.all(h == 1.0)
    p1 = 0.2 * math.exp(-0.2)
    m = n * steps
    z = (len(h) - m * p1) / math.sqrt(m * p1 * (1 - p1))
    assert abs(z) < 5
    st, nd, ct, h = _legacy_run(1000.0, n, steps)
    assert len(h) == 0


def test_legacy_poisson_rate_range():
    net = po.Net(dt=0.1)
    g = net.create("legacy_poisson_generator", 4)
    with pytest.raises(ValueError):
        net.legacy_rate(g, 4.0e6)      # lambda 400: table longer than 256
    with pytest.raises(ValueError):
        net.legacy_rate(g, -1.0)
    net.legacy_rate(g, 0.0)


def test_legacy_poisson_delivery_through_parrots():
    """Spikes of height n reach a parrot neuron through a weight-1 synapse
    of delay 0.5 ms as an input of n, two steps after the delay (one step
    of the spike ring, one of the input fold), like any other spike."""
    n, steps = 64, 200
    net = po.Net(dt=0.1, seed=99)
    g = net.create("legacy_poisson_generator", n)
    net.legacy_rate(g, 20000.0)
    par = net.create("parrot_neuron", n)
    par.set("hold_spike_height", np.ones(n, np.float32))
    idx = np.arange(n, dtype=np.int32)
    net.connect(g.i0 + idx, par.i0 + idx, np.ones(n, np.float32),
                np.full(n, 0.5, np.float32))
    net.calibrate()
    net.run(steps)
    st, nd, ct = net.spikes()
    h = net.spike_heights()
    gen = {(int(s), int(d) - g.i0): v for s, d, v in zip(st, nd, h)
           if d < par.i0}
    out = {(int(s), int(d) - par.i0): v for s, d, v in zip(st, nd, h)
           if d >= par.i0}
    lag = 4 + 2
    want = {(s + lag, i): v for (s, i), v in gen.items() if s + lag < steps}
    assert len(want) > 1000
    assert out == want


# ---------------------------------------------------------------------
# own exp / log (FP32, shared with the kernels)
# ---------------------------------------------------------------------
def _ulp_err(got, ref):
    got = np.float32(got)
    spacing = np.spacing(np.abs(np.float32(ref))).astype(np.float64)
    return abs(np.float64(got) - ref) / spacing


def test_expf_accuracy():
    L = po.lib()
    xs = np.concatenate([np.linspace(-87.0, 88.0, 200001),
                         np.linspace(-1.0, 1.0, 20001),
                         [0.0, -0.0, 1e-8, -1e-8]]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = math.exp(float(x))
        worst = max(worst, _ulp_err(L.orc_expf(float(x)), ref))
    assert worst <= 2.0
    assert L.orc_expf(0.0) == 1.0
    assert L.orc_expf(-200.0) == 0.0
    assert math.isinf(L.orc_expf(200.0))


def test_logf_accuracy():
    L = po.lib()
    xs = np.concatenate([np.geomspace(1e-30, 1e30, 100001),
                         np.linspace(0.5, 2.0, 20001)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = math.log(float(x))
        if abs(ref) < 1e-6:
            continue
        worst = max(worst, _ulp_err(L.orc_logf(float(x)), ref))
    assert worst <= 2.0
    assert L.orc_logf(1.0) == 0.0


def test_pow_ctrl_accuracy():
    """The step-size controller's err^-0.2 / err^-0.25 (orc_pow_ctrl, the
    same sequence as dev_pow_ctrl) against math.pow, over the errors the
    controller sees (above err_min = 1.89e-4) and well beyond."""
    L = po.lib()
    xs = np.concatenate([np.geomspace(1e-4, 1e8, 200001),
                         np.linspace(0.5, 2.0, 20001)]).astype(np.float32)
    for c in (-0.2, -0.25):
        worst = 0.0
        for x in xs:
            ref = math.pow(float(x), c)
            got = L.orc_pow_ctrl(float(x), c)
            worst = max(worst, abs(got - ref) / ref)
        assert worst <= 4e-7, (c, worst)
        assert L.orc_pow_ctrl(float("inf"), c) == 0.0
        assert math.isnan(L.orc_pow_ctrl(float("nan"), c))


# ---------------------------------------------------------------------
# STDP rule (src/stdp.cu:24-51) against the reference's test formula
# ---------------------------------------------------------------------
def test_stdp_rule_closed_form():
    L = po.lib()
    rng = np.random.default_rng(7)
    for _ in range(2000):
        p = {"tau_plus": rng.uniform(5, 40), "tau_minus": rng.uniform(5, 40),
             "lambda": rng.uniform(1e-3, 0.1), "alpha": rng.uniform(0.5, 2),
             "mu_plus": rng.choice([0.0, 0.5, 1.0, 2.0]),
             "mu_minus": rng.choice([0.0, 0.5, 1.0, 2.0]),
             "Wmax": rng.uniform(1, 100)}
        w = float(np.float32(rng.uniform(0, p["Wmax"])))
        dt = float(np.float32(rng.uniform(-50, 50)))
        par = np.array([p[k] for k in po.STDP_PARAM], dtype=np.float32)
        got = L.orc_stdp_update(w, dt, par.ctypes.data)
        pf = {k: float(v) for k, v in zip(po.STDP_PARAM, par)}
        want = nb.stdp_closed_form(w, [dt], pf)
        assert abs(got - want) <= 2e-6 * max(1.0, abs(want)), (p, w, dt)


def test_stdp_clamps():
    L = po.lib()
    par = np.array([20, 20, 0.5, 1, 1, 1, 10], dtype=np.float32)
    assert L.orc_stdp_update(9.99, 0.1, par.ctypes.data) <= 10.0
    assert L.orc_stdp_update(0.01, -0.1, par.ctypes.data) >= 0.0


@pytest.mark.parametrize("case", [str(i) for i in range(1, 11)])
def test_stdp_reference_cases(case):
    """python/test/test_stdp/cases/case<k>.py: final weight equals the
    closed-form sequence of updates the reference test expects."""
    cs = nb.load_stdp_cases()
    c = cs["cases"][case]
    spec = nb.stdp_case_spec(c, cs["common"])
    want = nb.stdp_closed_form(cs["common"]["weight"], c["Dt"], cs["common"])
    for prec in ("f32", "f64"):
        net, _ = nb.build_oracle(spec, prec)
        net.run(int(round(cs["common"]["sim_time"] / 0.1)))
        src, tgt, port, d, w = net.connections()
        (got,) = w[(src == 0) & (tgt == 1)]
        assert abs(float(got) - want) <= 1e-6 * want, (prec, got, want)


# ---------------------------------------------------------------------
# calibration closed forms
# ---------------------------------------------------------------------
@pytest.mark.parametrize("tr,td", [(0.2, 2.0), (0.5, 10.0), (2.0, 20.0),
                                   (20.0, 40.0), (5.0, 5.0)])
def test_beta_g0_normalises_peak(tr, td):
    """g1(0) = g0 after a unit-weight spike; the conductance
    g(t) = g0 (e^-t/td - e^-t/tr)/(1/tr - 1/td) (t e^-t/td if tr = td)
    must peak at 1 (src/aeif_cond_beta.cu:66-85)."""
    L = po.lib("f64")
    g0 = L.orc_beta_g0(tr, td)
    t = np.linspace(1e-6, 10 * td, 2000001)
    if tr == td:
        g = g0 * t * np.exp(-t / td)
    else:
        g = g0 * (np.exp(-t / td) - np.exp(-t / tr)) / (1 / tr - 1 / td)
    assert abs(g.max() - 1.0) < 1e-9


@pytest.mark.parametrize("h,tm,cm,tex,tin", [(0.1, 10.0, 250.0, 2.0, 8.0),
                                             (0.1, 10.0, 250.0, 0.5, 0.5),
                                             (0.05, 20.0, 100.0, 3.0, 1.0)])
def test_iaf_propagators_match_expm(h, tm, cm, tex, tin):
    """Exact integration of (I_ex, I_in, V): propagators equal the matrix
    exponential of the linear system (src/iaf_psc_exp.cu:53-92)."""
    L = po.lib("f64")
    out = np.zeros(6, dtype=np.float32)
    L.orc_iaf_propagators(h, tm, cm, tex, tin, out.ctypes.data)
    P20, P11ex, P11in, P21ex, P21in, P22 = [float(v) for v in out]
    A = np.array([[-1 / tex, 0, 0], [0, -1 / tin, 0],
                  [1 / cm, 1 / cm, -1 / tm]])
    E = scipy.linalg.expm(A * h)
    want = [tm / cm * (1 - math.exp(-h / tm)), E[0, 0], E[1, 1], E[2, 0],
            E[2, 1], E[2, 2]]
    got = [P20, P11ex, P11in, P21ex, P21in, P22]
    for g, w in zip(got, want):
        assert abs(g - w) <= 2e-7 * max(abs(w), 1e-12) + 1e-12


# ---------------------------------------------------------------------
# RK5 (Cash-Karp) order and error control
# ---------------------------------------------------------------------
def test_rk5_fifth_order():
    L = po.lib("f64")
    errs = []
    for n in (32, 64, 128, 256):
        y = L.orc_rk5_linear_fixed(-2.0, 1.0, 2.0, n)
        errs.append(abs(y - math.exp(-4.0)))
    ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    # global error ~ h^5: halving h divides it by ~32 (asymptotic range;
    # a dropped or wrong tableau term shows up as order <= 4, ratio <= 16)
    for r in ratios:
        assert 27.0 < r < 34.0, ratios


@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-9])
def test_rk5_adaptive_tolerance(eps):
    L = po.lib("f64")
    na = np.zeros(1, np.int64)
    nr = np.zeros(1, np.int64)
    y = L.orc_rk5_linear_adaptive(-3.0, 1.0, 5.0, 0.01, 1e-8, eps,
                                  na.ctypes.data, nr.ctypes.data)
    want = math.exp(-15.0)
    # local error per step <= eps * (|y| + |h y'| + 0.1) scale; globally
    # bounded by steps * eps * (1 + 0.1)
    assert abs(y - want) <= na[0] * eps * 1.2
    assert na[0] > 0


# ---------------------------------------------------------------------
# neuron models against the reference's golden NEST traces
# ---------------------------------------------------------------------
@pytest.mark.parametrize("model", ["aeif_cond_beta", "aeif_cond_alpha",
                                   "iaf_psc_exp", "aeif_psc_alpha",
                                   "aeif_psc_exp", "aeif_psc_delta"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_nest_golden_trace_oracle(model, prec):
    """python/test/test_<model>.py: NEST row i <-> record row 10 i + 20,
    std(dV)/|mean V| <= 5e-4 (aeif) or 5e-5 (iaf_psc_exp)."""
    spec = nb.nest_spec(model)
    net, _ = nb.build_oracle(spec, prec)
    net.run(8000)
    V = net.records()[:, 1]
    if model == "iaf_psc_exp":
        V = V - 70.0
    nest = np.loadtxt(os.path.join(GOLDEN, "%s_nest.txt" % model),
                      delimiter="\t")
    dV = np.array([V[i * 10 + 20] - nest[i, 1] for i in range(len(nest))])
    rmse = np.std(dV) / abs(np.mean(V))
    tol = 5e-5 if model == "iaf_psc_exp" else 5e-4
    assert rmse <= tol


# python/test/test_iaf_psc_exp_g.py and test_iaf_psc_exp_hc.py (both against
# test_fast_iaf_psc_exp_nest.txt, 800 ms; tolerance 5e-5, tighter than the
# reference's 5e-4), and test_izh.py (test_izh_nest.txt, 80 ms; the
# reference's tolerance 5e-3)
POINT_NEST = {"iaf_psc_exp_g": ("fast_iaf_psc_exp_nest.txt", 8000, -65.0, 5e-5),
              "iaf_psc_exp_hc": ("fast_iaf_psc_exp_nest.txt", 8000, -65.0, 5e-5),
              "izhikevich": ("izhikevich_nest.txt", 800, 0.0, 5e-3)}


def point_nest_rmse(V, model):
    fname, _, E_L, _ = POINT_NEST[model]
    V = V + E_L
    nest = np.loadtxt(os.path.join(GOLDEN, fname), delimiter="\t")
    dV = np.array([V[i * 10 + 20] - nest[i, 1] for i in range(len(nest))])
    return np.std(dV) / abs(np.mean(V))


@pytest.mark.parametrize("model", list(POINT_NEST))
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_point_model_nest_trace_oracle(model, prec):
    spec = nb.nest_spec(model)
    net, _ = nb.build_oracle(spec, prec)
    net.run(POINT_NEST[model][1])
    V = net.records()[:, 1]
    assert point_nest_rmse(V, model) <= POINT_NEST[model][3]


def test_iaf_psc_exp_hc_constants():
    """iaf_psc_exp_hc's fixed propagators (src/iaf_psc_exp_hc_params.h) are
    those of iaf_psc_exp_g's defaults: tau_m 10, C_m 250, tau_syn 0.5,
    t_ref 2 at dt 0.1 (scipy expm of the linear system, 8 digits)."""
    h, tau_m, C_m, tau_s = 0.1, 10.0, 250.0, 0.5
    A = np.array([[-1.0 / tau_s, 0.0], [1.0 / C_m, -1.0 / tau_m]])
    E = scipy.linalg.expm(A * h)
    P11, P21, P22 = E[0, 0], E[1, 0], E[1, 1]
    P20 = tau_m / C_m * (1.0 - P22)
    for got, want in ((8.1873075E-01, P11), (3.6067175E-04, P21),
                      (9.9004983E-01, P22), (3.9800665E-04, P20)):
        assert abs(got - want) <= 5e-8 * abs(want) + 1e-15
    # and the g model with those parameters gives the hc trace to float
    # rounding
    sg = nb.nest_spec("iaf_psc_exp_g")
    sh = nb.nest_spec("iaf_psc_exp_hc")
    ng, _ = nb.build_oracle(sg, "f32")
    nh, _ = nb.build_oracle(sh, "f32")
    ng.run(8000)
    nh.run(8000)
    assert np.max(np.abs(ng.records()[:, 1] - nh.records()[:, 1])) < 1e-4


def _izh_psc_run(model, dt, t_end, I_e, w, tau_syn, prec="f64"):
    spec = {"dt": dt,
            "groups": [{"model": model, "n": 1,
                        "params": {"I_e": I_e, "tau_syn": tau_syn},
                        "vars": {"I_syn": w}}],
            "records": [(0, "V_m", 0), (0, "I_syn", 0)]}
    net, _ = nb.build_oracle(spec, prec)
    net.run(int(round(t_end / dt)))
    return net.records()


def test_izhikevich_psc_exp_euler_order():
    """izhikevich_psc_exp, _2s, _5s are forward Euler on 1, 2, 5 sub-steps
    of dt for dV = 0.04 V^2 + 5 V + 140 - u + I_syn + I_e, du = a (b V - u),
    dI_syn = -I_syn / tau_syn.  From V = -70, u = -14, I_syn = 2 (below
    threshold), the V_m error against scipy LSODA shrinks in proportion to
    the sub-step: first order."""
    import scipy.integrate
    I_e, w, tau_syn, t_end = 1.0, 2.0, 2.0, 40.0

    def f(t, y):
        V, u, I = y
        return [0.04 * V * V + 5 * V + 140 - u + I + I_e,
                0.02 * (0.2 * V - u), -I / tau_syn]
    sol = scipy.integrate.solve_ivp(f, (0, t_end), [-70.0, -14.0, w],
                                    method="LSODA", rtol=1e-12, atol=1e-12,
                                    dense_output=True).sol
    errs = {}
    for model, k in (("izhikevich_psc_exp", 1), ("izhikevich_psc_exp_2s", 2),
                     ("izhikevich_psc_exp_5s", 5)):
        r = _izh_psc_run(model, 0.1, t_end, I_e, w, tau_syn)
        assert r[:, 1].max() < -60.0
        errs[k] = np.max(np.abs(r[:, 1] - sol(r[:, 0])[0]))
    assert 1.8 < errs[1] / errs[2] < 2.2
    assert 4.5 < errs[1] / errs[5] < 5.5
    assert errs[5] < 0.02


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_izhikevich_psc_exp_substeps(prec):
    """_2s at dt takes the same sub-steps as the one-step model at dt/2
    (the current decays after every sub-step): V_m and I_syn agree bit for
    bit at the common times."""
    a = _izh_psc_run("izhikevich_psc_exp_2s", 0.1, 20.0, 1.0, 2.0, 2.0, prec)
    b = _izh_psc_run("izhikevich_psc_exp", 0.05, 20.0, 1.0, 2.0, 2.0, prec)
    assert np.array_equal(a[:, 1:], b[::2, 1:])


def test_izhikevich_spike_reset():
    """Suprathreshold current: each crossing of V_th = 30 resets V_m to c
    and adds d to u (src/izhikevich.cu:57-68); t_ref = 0 leaves no
    refractory steps."""
    spec = {"groups": [{"model": "izhikevich", "n": 1,
                        "params": {"I_e": 10.0}}],
            "records": [(0, "V_m", 0), (0, "u", 0)]}
    net, _ = nb.build_oracle(spec, "f32")
    net.run(3000)
    st, nd, ct = net.spikes()
    r = net.records()
    assert len(st) >= 5
    for s in st[:5]:
        # record row s + 1 is the end of the spiking step
        assert r[s + 1, 1] == np.float32(-65.0)
        assert r[s + 1, 2] > r[s, 2] + 7.0         # u jumped by d = 8


def test_aeif_fixed_point():
    """Constant I_e below rheobase: V, w converge to the root of
    dV = dw = 0 with w* = a (V* - E_L)."""
    I_e = 300.0
    spec = {"groups": [{"model": "aeif_cond_beta", "n": 1, "n_port": 1,
                        "params": {"I_e": I_e}}],
            "records": [(0, "V_m", 0), (0, "w", 0)]}
    net, _ = nb.build_oracle(spec, "f64")
    net.run(30000)                              # 3 s, ~20 tau_w
    r = net.records()
    V, W = r[-1, 1], r[-1, 2]
    gL, EL, VT, DT, a = 30.0, -70.6, -50.4, 2.0, 4.0

    def f(v):
        return (-gL * (v - EL) + gL * DT * math.exp((v - VT) / DT)
                - a * (v - EL) + I_e)
    vstar = scipy.optimize.brentq(f, -70.0, -55.0, xtol=1e-13)
    assert abs(V - vstar) < 1e-6 * abs(vstar)
    assert abs(W - a * (vstar - EL)) < 1e-5 * abs(a * (vstar - EL))


def test_rk5_tolerance_convergence():
    """The same neuron at tolerance 1e-6 and 1e-9: subthreshold traces agree
    to the looser tolerance's accuracy; a spiking neuron keeps its spike
    count and the tighter tolerance takes more RK5 attempts."""
    def run(I_e, tol, steps):
        spec = {"groups": [{"model": "aeif_cond_beta", "n": 1, "n_port": 1,
                            "params": {"I_e": I_e}, "rk5": {"rk5_tol": tol}}],
                "records": [(0, "V_m", 0)]}
        net, _ = nb.build_oracle(spec, "f64")
        net.run(steps)
        return net
    va = run(400.0, 1e-6, 5000).records()[:, 1]
    vb = run(400.0, 1e-9, 5000).records()[:, 1]
    assert np.max(np.abs(va - vb)) < 1e-4
    a, b = run(1000.0, 1e-6, 5000), run(1000.0, 1e-9, 5000)
    assert len(a.spikes()[0]) == len(b.spikes()[0]) > 5
    ca, cb = a.counters(), b.counters()
    assert (cb["rk5_accepted"] + cb["rk5_rejected"] >
            ca["rk5_accepted"] + ca["rk5_rejected"])


# ---------------------------------------------------------------------
# spike routing timing (python/test/test_spike_detector.py)
# ---------------------------------------------------------------------
def test_delivery_delay_rows():
    """A spike of the source at record row i reaches a relay with delay
    10/20/30 ms at row i + 101/201/301, with the synaptic weight as
    height (python/test/test_spike_detector.py:37-52)."""
    spec = {"groups": [
        {"model": "aeif_cond_beta", "n": 3, "n_port": 1,
         "params": {"I_e": 1000.0}},
        {"model": "parrot_neuron", "n": 3,
         "params": {"hold_spike_height": 1.0}}],
        "conns": [{"src": [0], "tgt": [3], "weight": 1.0, "delay": 10.0},
                  {"src": [1], "tgt": [4], "weight": 2.0, "delay": 20.0},
                  {"src": [2], "tgt": [5], "weight": 3.0, "delay": 30.0}],
        "records": [(3, "V", 0), (4, "V", 0), (5, "V", 0)]}
    net, _ = nb.build_oracle(spec, "f32")
    net.run(10000)
    rec = net.records()
    st, nd, ct = net.spikes()
    src_rows = sorted(set(int(s) + 1 for s, n in zip(st, nd) if n == 0))
    assert len(src_rows) > 5
    n_rows = rec.shape[0]
    for i in src_rows:
        for k, (off, h) in enumerate([(101, 1.0), (201, 2.0), (301, 3.0)]):
            if i + off < n_rows:
                assert rec[i + off, 1 + k] == h
    # and nothing arrives elsewhere
    for k, off in enumerate([101, 201, 301]):
        rows = set(np.nonzero(rec[:, 1 + k])[0])
        assert rows == set(i + off for i in src_rows if i + off < n_rows)


def test_ext_neuron_port_values():
    """ext_neuron (src/ext_neuron.cu:23-39) shows each input in port_value
    for exactly one step, at the step a parrot neuron with the same
    synapse shows it (the delivery timing pinned above), scaled by
    port_weight."""
    net, _ = nb.build_oracle(nb.ext_spec(), "f32")
    net.run(8000)
    r = net.records()
    for p in range(3):
        assert np.array_equal(r[:, 1 + p], r[:, 4 + p])
        assert np.count_nonzero(r[:, 1 + p]) == 4
    # the generator's first spike, at record row i, arrives through the
    # 1 / 50 / 100 ms synapses at rows i + 11 / 501 / 1001
    st, nd, _ = net.spikes()
    i = int(st[nd == 4][0]) + 1
    for p, (off, w) in enumerate([(11, 0.1), (501, 0.2), (1001, 0.5)]):
        assert r[i + off, 1 + p] == np.float32(w)
    net, _ = nb.build_oracle(nb.ext_spec([1.0, 2.0, 0.5]), "f32")
    net.run(8000)
    r2 = net.records()
    for p, pw in enumerate([1.0, 2.0, 0.5]):
        assert np.array_equal(r2[:, 1 + p],
                              (r[:, 1 + p] * np.float32(pw)).astype(np.float32))


# ---------------------------------------------------------------------
# network statistics (python/test/test_brunel_list.py:11,132-148)
# ---------------------------------------------------------------------
def test_brunel_rate():
    spec = nb.brunel_spec()
    net, _ = nb.build_oracle(spec, "f32")
    n_steps = 10000
    net.run(n_steps, n_threads=8)
    st, nd, ct = net.spikes()
    n_neur = spec["groups"][0]["n"]
    rate = np.sum(ct[nd < n_neur]) / n_neur / (n_steps * 0.1e-3)
    # reference expectation 30.78 Hz for its own random network
    assert abs(rate - 30.78) < 0.1 * 30.78, rate


def test_c2_spec_shape():
    """c2 recipe: population sizes, synapse count ~ sum p N_i N_j, weight
    signs by source population, delays >= 0.1 ms."""
    from neurongpu_amd import workloads as wl
    spec = wl.c2_spec(scale=0.05, integer_weights=True)
    pops = spec["populations"]
    assert [b - a for a, b in pops] == [
        int(round(n * 0.05)) for n in wl.C2_SIZES]
    for c in spec["conns"]:
        src = np.asarray(c["src"])
        if src[0] >= pops[-1][1]:
            continue                      # Poisson generator
        pop = [k for k, (a, b) in enumerate(pops) if a <= src[0] < b][0]
        w = np.asarray(c["weight"])
        assert (w > 0).all() if pop % 2 == 0 else (w < 0).all()
        assert np.all(np.asarray(c["delay"]) >= 0.1)
        assert np.all(w == np.round(w))


# ---------------------------------------------------------------------
# izhikevich_cond_beta and user_m1 / user_m2 (no golden vectors in the
# reference): closed forms, an independent integration, identities
# ---------------------------------------------------------------------
def test_izhikevich_default_state_is_fixed_point():
    """With I_e = 0 and no input, 0.04 V^2 + 5 V + 140 - u = 0 and
    b V - u = 0 hold at the defaults V = -70, u = b V = -14 (the stable
    root of 0.04 V^2 + 4.8 V + 140 = 0): the state stays there."""
    spec = {"groups": [{"model": "izhikevich_cond_beta", "n": 1,
                        "n_port": 1}],
            "records": [(0, "V_m", 0), (0, "u", 0)]}
    for prec in ("f64", "f32"):
        net, _ = nb.build_oracle(spec, prec)
        net.run(2000)
        r = net.records()
        assert np.allclose(r[:, 1], -70.0, atol=1e-5)
        assert np.allclose(r[:, 2], -14.0, atol=1e-5)


def _izh_scipy_spikes(I_e, t_end, a=0.02, b=0.2, c=-65.0, d=8.0,
                      V_th=30.0):
    """Spike times of the Izhikevich equations under constant current from
    scipy's LSODA with threshold events (independent of the oracle)."""
    from scipy.integrate import solve_ivp

    def f(t, y):
        V, u = y
        return [0.04 * V * V + 5.0 * V + 140.0 - u + I_e, a * (b * V - u)]

    def ev(t, y):
        return y[0] - V_th
    ev.terminal = True
    ev.direction = 1
    t, y = 0.0, [-70.0, b * -70.0]
    spikes = []
    while t < t_end:
        sol = solve_ivp(f, (t, t_end), y, method="LSODA", events=ev,
                        rtol=1e-10, atol=1e-10, max_step=0.05)
        if sol.status == 1:
            t = float(sol.t_events[0][0])
            spikes.append(t)
            y = [c, float(sol.y_events[0][0][1]) + d]
        else:
            break
    return np.array(spikes)


def test_izhikevich_spike_times_match_independent_integration():
    """Regular spiking Izhikevich neuron under I_e = 10 over 500 ms: the
    FP64 oracle (RK5, tolerance 1e-8) against an independent LSODA
    integration with exact threshold events.  The oracle resets at the end
    of the substep that crosses V_th (ExternalUpdate after each substep,
    src/rk5.h:149-162) and stamps spikes at the end of the time step, so
    the first spike agrees to within one step (0.1 ms) and every
    inter-spike interval to within one step plus that substep (0.12 ms);
    the spike counts are equal."""
    spec = {"groups": [{"model": "izhikevich_cond_beta", "n": 1,
                        "n_port": 1, "params": {"I_e": 10.0},
                        "rk5": {"rk5_tol": 1e-8}}],
            "records": []}
    net, _ = nb.build_oracle(spec, "f64")
    net.run(5000)
    st, nd, ct = net.spikes()
    t_orc = (np.asarray(st, dtype=np.float64) + 1) * 0.1
    t_ref = _izh_scipy_spikes(10.0, 500.0)
    assert len(t_ref) >= 5
    assert len(t_orc) == len(t_ref)
    assert 0.0 <= t_orc[0] - t_ref[0] <= 0.1 + 1e-9
    assert np.all(np.abs(np.diff(t_orc) - np.diff(t_ref)) <= 0.12)


def test_user_models_are_aeif_cond_beta():
    """user_m1 and user_m2 are aeif_cond_beta under other names in the
    reference (src/user_m1_kernel.h, src/user_m2_kernel.h): identical
    traces on the NEST test setup."""
    base = nb.nest_spec("aeif_cond_beta")
    net, _ = nb.build_oracle(base, "f32")
    net.run(3000)
    ref = net.records()
    for m in ("user_m1", "user_m2"):
        spec = nb.nest_spec("aeif_cond_beta")
        spec["groups"][0]["model"] = m
        net, _ = nb.build_oracle(spec, "f32")
        net.run(3000)
        assert np.array_equal(net.records(), ref)


def test_izhikevich_conductance_peak_is_weight():
    """izhikevich_cond_beta's g0 normalises the beta conductance to peak
    at the synaptic weight, t_p = td tr ln(td/tr) / (td - tr) after the
    spike arrives (src/izhikevich_cond_beta.cu:57-84): a single input of
    weight 0.5 nS on port 1 (tau_rise 2, tau_decay 20 ms) peaks at 0.5 nS,
    5.117 ms after its arrival."""
    spec = {"groups": [{"model": "izhikevich_cond_beta", "n": 1,
                        "n_port": 2,
                        "params": {"tau_rise": [1.0, 2.0],
                                   "tau_decay": [10.0, 20.0]}},
                       {"model": "spike_generator", "n": 1}],
            "spike_gen": [{"node": 1, "times": [10.0]}],
            "conns": [{"src": [1], "tgt": [0], "weight": [0.5],
                       "delay": [1.0], "port": 1}],
            "records": [(0, "g", 1)]}
    for prec in ("f64", "f32"):
        net, _ = nb.build_oracle(spec, prec)
        net.run(600)
        r = net.records()
        k = int(np.argmax(r[:, 1]))
        t_p = 20.0 * 2.0 * math.log(10.0) / 18.0
        # the input is folded at the end of step 110 (spike at 10 ms
        # delivered after 1 ms) and integrated from t = 11.1 ms
        assert abs(r[k, 1] - 0.5) <= 2e-4
        assert abs(r[k, 0] - (11.1 + t_p)) <= 0.1 + 1e-6


def test_float_vs_double_accumulators():
    """The reference accumulates each step's synaptic input in double
    (src/get_spike.cu:70-72); this build, like the configurations, in
    float.  Quantified on the oracle with float state and either
    accumulator:
      * c0 as specified is silent, so its accumulators receive no synaptic
        event and both runs are identical bit for bit;
      * on the reference's Brunel test network (1000 neurons, 1 s) the
        spike trains part once a sum rounds differently, and the mean rate
        and CV ISI of the two runs stay within 2 % of each other."""
    spec = nb.c0_spec(n=1000, seed=1234)
    runs = []
    for prec in ("f32", "f32a64"):
        net, og = nb.build_oracle(spec, prec)
        net.run(2000, n_threads=8)
        runs.append((og[0].get("V_m"), net.spikes()[0].size))
    assert runs[0][1] == 0 and runs[1][1] == 0
    assert np.array_equal(runs[0][0], runs[1][0])
    spec = nb.brunel_spec()
    n, steps = spec["groups"][0]["n"], 10000
    stats, trains = [], []
    for prec in ("f32", "f32a64"):
        net, _ = nb.build_oracle(spec, prec)
        net.run(steps, n_threads=8)
        tr = nb.oracle_times_by_neuron(net, 0.1, 0, n)
        trains.append(tr)
        stats.append(nb.spike_stats(tr, n, 100.0, steps * 0.1, n_pairs=0))
    first = min((float(a[0]) for a, b in zip(*trains)
                 if a.size and (a.size != b.size or not np.array_equal(a, b))),
                default=None)
    print("float vs double accumulators, Brunel 1000:", stats,
          "first differing spike train starts at", first, "ms")
    for k in ("rate", "cv_isi"):
        assert abs(stats[0][k] - stats[1][k]) <= 0.02 * stats[1][k]
