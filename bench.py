#!/usr/bin/env python3
"""Benchmark of the per-timestep spiking-network path on one MI355X.

A step is one simulation time step (dt = 0.1 ms) of the whole network:
the fused AdEx RK5 update (with the Poisson drive and the port-input fold
of the previous step), spike-ring append and the synapse delivery pass.
The timed region is `steps` steps run through NeuronGPU_Simulate, inputs
and synapses already resident in HBM; the reported value is wall seconds
per biological second.

Workloads:
  c1        1,000,000 aeif_cond_beta (800k E / 200k I), 2 ports
            (E_rev 0/-85 mV, tau_rise 0.2/0.5, tau_decay 2/10 ms), fixed
            indegree 1000 (800 E w=0.05 nS port 0, 200 I w=0.25 nS port 1),
            delays uniform 1-20 steps, 8 kHz Poisson drive per neuron
            (w=0.1 nS), V0 ~ U(-70,-55) mV, RK5 tolerance 1e-3, seed 1234.
            This is the headline workload (`value`).  With these
            parameters the Poisson drive stays below the AdEx rheobase and
            the network is silent.
  readme    the reference README's 1 M neuron / 1 B synapse benchmark network
            (c++/examples/brunel_net.cpp at n_neurons = 1e6): 20 kHz
            Poisson (w=0.37), tau_rise = tau_decay = 1 ms, Win = 0.35,
            delays N(0.5, 0.25) ms clipped [0.1, 1.25], RK5 tolerance 1e-6.
            An active network (~30 spikes/s); the reference quotes "little
            more than 70 s per second of neural activity" on an RTX 2080 Ti
            (README.md:11).  Reported in the same JSON line.
  active_1m the same 1 M / 1e9 network at RK5 tolerance 1e-3: an active
            1 M AdEx / 1 B synapse network at the tolerance the
            configurations use, reported beside the headline.

Active networks are run for --settle-ms of biological time (plain
stepping, not timed) before their warm-up and timed window, so that the
window sees the network's sustained activity rather than its silent
onset.  Each object reports the settle time it used.

CPU baseline: the FP64 CPU oracle (OpenMP over neurons) on bounded
samples of the c1 workload (reduced neuron count, same per-neuron
indegree, weights, delays and drive), on all the cores the process may
use and on one thread, scaled linearly to 1 M neurons and 1 s of
biological time; the host's lscpu topology is reported beside it.
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("wall s per bio s + synaptic events/s (1M AdEx, 1B syn); "
          "RK5 stages/step learned vs plain")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# build times of the last network, for the workload objects
BUILD = {}
# algorithmic work per unit (RK5 aeif P=2, delivery)
RK5_BYTES_PER_NEURON_STEP = 72.0   # 2*4*NVAR + 8 h + 8 refractory + 4*P
RK5_FLOP_ACCEPTED = 512.0
RK5_FLOP_REJECTED = 465.0
DELIVERY_BYTES_PER_EVENT = 9.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hip_sync():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize()


def build_c1(ngpu, n=1000000, seed=1234, tol=1e-3):
    from neurongpu_amd import workloads as wl
    t = time.time()
    neuron, n_syn, _ = wl.build_c1(ngpu, n=n, seed=seed, tol=tol)
    BUILD["connect_s"] = time.time() - t
    log("  network built in %.1f s" % BUILD["connect_s"])
    return neuron, n_syn


def build_readme(ngpu, n=1000000, seed=1234, tol=None):
    """c++/examples/brunel_net.cpp:31-99 at n_neurons = n (RK5 tolerance
    1e-6, the reference's fixed value, unless `tol` is given)."""
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    order = n // 5
    n_e, n_i = 4 * order, order
    ce, ci = 800, 200
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, {"rate": 20000.0})
    neuron = ngpu.Create("aeif_cond_beta", n_e + n_i, 2)
    exc = neuron[0:n_e]
    inh = neuron[n_e:n_e + n_i]
    ngpu.SetStatus(neuron, {"E_rev": [0.0, -85.0], "tau_decay": [1.0, 1.0],
                            "tau_rise": [1.0, 1.0]})
    if tol is not None:
        ngpu.SetStatus(neuron, {"rk5_tol": tol})
    t = time.time()
    # delays N(0.5, 0.25) ms clipped to [0.1, 0.5 + 3 * 0.25], drawn per
    # synapse by the library (brunel_net.cpp draws them with
    # RandomNormalClipped into host arrays)
    delay = {"distribution": "normal_clipped", "mu": 0.5, "sigma": 0.25,
             "low": 0.1, "high": 0.5 + 3 * 0.25}
    ngpu.Connect(exc, neuron, {"rule": "fixed_indegree", "indegree": ce},
                 {"weight": 0.05, "delay": delay, "receptor": 0})
    ngpu.Connect(inh, neuron, {"rule": "fixed_indegree", "indegree": ci},
                 {"weight": 0.35, "delay": delay, "receptor": 1})
    ngpu.Connect(pg, neuron, {"rule": "all_to_all"},
                 {"weight": 0.37, "delay": 0.2, "receptor": 0})
    BUILD["connect_s"] = time.time() - t
    log("  connections in %.1f s" % BUILD["connect_s"])
    return neuron, (n_e + n_i) * (ce + ci)


def build_c2(ngpu, scale=1.0, seed=2014):
    """c2: layered cortical-column-shaped iaf_psc_exp network (the recipe of
    neurongpu_amd.workloads.c2_spec, built with the library's own
    fixed_total_number rule and weight/delay distributions)."""
    from neurongpu_amd import workloads as wl
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    rng = np.random.default_rng(seed)
    sizes = [max(1, int(round(n * scale))) for n in wl.C2_SIZES]
    off = np.concatenate([[0], np.cumsum(sizes)])
    n_tot = int(off[-1])
    neuron = ngpu.Create("iaf_psc_exp", n_tot)
    ngpu.SetStatus(neuron, {"V_m_rel": {"array": rng.uniform(
        0.0, 15.0, n_tot).astype(np.float32)}})
    pg = ngpu.Create("poisson_generator", 8)
    p = rng.uniform(0.01, 0.15, (8, 8))
    pops = [neuron[int(off[j]):int(off[j + 1])] for j in range(8)]
    t = time.time()
    n_syn = 0
    for j in range(8):
        ngpu.SetStatus(pg[j:j + 1], {"rate": 8.0 * wl.C2_K_EXT[j]})
        for i in range(8):
            k = int(round(p[j, i] * sizes[i] * sizes[j]))
            if k == 0:
                continue
            exc = i % 2 == 0
            syn = {"weight": {"distribution": "normal",
                              "mu": 87.8 if exc else -351.2,
                              "sigma": 8.8 if exc else 35.2},
                   "delay": {"distribution": "normal_clipped",
                             "mu": 1.5 if exc else 0.8,
                             "sigma": 0.75 if exc else 0.4,
                             "low": 0.1, "high": 1.0e3},
                   "receptor": 0 if exc else 1}
            ngpu.Connect(pops[i], pops[j],
                         {"rule": "fixed_total_number", "total_num": k}, syn)
            n_syn += k
        ngpu.Connect(pg[j:j + 1], pops[j], {"rule": "all_to_all"},
                     {"weight": 87.8, "delay": 0.1, "receptor": 0})
    BUILD["connect_s"] = time.time() - t
    log("  connections in %.1f s (%d synapses)" % (BUILD["connect_s"], n_syn))
    return neuron, n_syn


def build_c3(ngpu, scale=1.0, seed=4321, k=500, w_in=0.25, w_p=0.3,
             rate=8000.0, tol=1e-3):
    """c3: aeif_cond_alpha network with STDP on E -> E (the recipe of
    neurongpu_amd.workloads.c3_spec, built with the library's
    fixed_indegree rule)."""
    from neurongpu_amd import workloads as wl
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    rng = np.random.default_rng(seed)
    n = max(10, int(round(200000 * scale)))
    n_e = int(round(0.8 * n))
    k_e, k_i = int(round(0.8 * k)), k - int(round(0.8 * k))
    neuron = ngpu.Create("aeif_cond_alpha", n, 2)
    exc, inh = neuron[0:n_e], neuron[n_e:n]
    ngpu.SetStatus(neuron, dict(wl.C3_NEURON))
    ngpu.SetStatus(neuron, {"V_m": {"array": rng.uniform(
        -70.0, -55.0, n).astype(np.float32)}, "rk5_tol": tol})
    pg = ngpu.Create("poisson_generator", n)
    ngpu.SetStatus(pg, {"rate": rate})
    sg = ngpu.CreateSynGroup("stdp", dict(wl.C3_STDP))
    t = time.time()

    def arrays(m):
        return ({"array": rng.uniform(0.02, 0.06, m).astype(np.float32)},
                {"array": (rng.integers(1, 21, m) * 0.1).astype(np.float32)})
    w, d = arrays(n_e * k_e)
    ngpu.Connect(exc, exc, {"rule": "fixed_indegree", "indegree": k_e},
                 {"weight": w, "delay": d, "receptor": 0,
                  "synapse_group": sg})
    w, d = arrays((n - n_e) * k_e)
    ngpu.Connect(exc, inh, {"rule": "fixed_indegree", "indegree": k_e},
                 {"weight": w, "delay": d, "receptor": 0})
    _, d = arrays(n * k_i)
    ngpu.Connect(inh, neuron, {"rule": "fixed_indegree", "indegree": k_i},
                 {"weight": w_in, "delay": d, "receptor": 1})
    ngpu.Connect(pg, neuron, {"rule": "one_to_one"},
                 {"weight": w_p, "delay": 0.1, "receptor": 0})
    BUILD["connect_s"] = time.time() - t
    log("  connections in %.1f s" % BUILD["connect_s"])
    return neuron, n * k


def window_stats(ngpu, neuron, steps):
    """Runs `steps` steps (host-synchronised on both sides) and returns
    their wall time, spikes, synaptic events and RK5 attempt counts."""
    dt = ngpu.GetTimeResolution()
    s0 = ngpu.GetHotPathStats()
    p0 = ngpu.GetRk5Pass1Stats()
    c0 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n,
                                         "spike_count")))
    hip_sync()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * dt)
    hip_sync()
    wall = time.perf_counter() - t0
    s1 = ngpu.GetHotPathStats()
    p1 = ngpu.GetRk5Pass1Stats()
    c1 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n,
                                         "spike_count")))
    bio_s = steps * dt * 1e-3
    acc = s1["rk5_accepted"] - s0["rk5_accepted"]
    rej = s1["rk5_rejected"] - s0["rk5_rejected"]
    return {
        "wall_s": wall, "bio_s": bio_s, "wall_s_per_bio_s": wall / bio_s,
        "ms_per_step": 1e3 * wall / steps,
        "syn_events": s1["syn_events"] - s0["syn_events"],
        "syn_events_per_s": (s1["syn_events"] - s0["syn_events"]) / wall,
        "rate_hz": (c1 - c0) / neuron.n / bio_s,
        "rk5_accepted": acc, "rk5_rejected": rej,
        "rk5_pass1_accepted": p1["rk5_pass1_accepted"] -
        p0["rk5_pass1_accepted"],
        "rk5_pass1_rejected": p1["rk5_pass1_rejected"] -
        p0["rk5_pass1_rejected"],
        "rk5_stages_per_neuron_step": (6.0 * acc + 5.0 * rej) /
        (neuron.n * steps)}


def measure(ngpu, neuron, n_syn, steps, warmup, timing_steps,
            after_warmup=None, settle_steps=0):
    """Calibrate, settle (untimed plain stepping), warm up, then time
    exactly `steps` steps; then a separate window of `timing_steps` steps
    with per-kernel HIP-event timing."""
    dt = ngpu.GetTimeResolution()
    ngpu.ActivateSpikeCount(neuron)
    t = time.time()
    ngpu.Calibrate()
    t_cal = time.time() - t
    log("  calibrate (CSR build + upload) in %.1f s" % t_cal)
    if settle_steps > 0:
        t = time.time()
        ngpu.Simulate(settle_steps * dt)
        hip_sync()
        log("  settled %.0f ms bio in %.1f s" % (settle_steps * dt,
                                                   time.time() - t))
    if warmup > 0:
        ngpu.Simulate(warmup * dt)
    extra = after_warmup() if after_warmup else None
    res = window_stats(ngpu, neuron, steps)
    res["settle_bio_ms"] = round(settle_steps * dt, 3)
    # per-kernel device time (HIP events on the engine stream); the second
    # RK5 pass and the delivery normally share one launch, so they are
    # timed here as separate launches
    fused = ngpu.GetKernelStatus("fuse_delivery")
    ngpu.SetKernelStatus("fuse_delivery", 0)
    ngpu.SetKernelTiming(1)
    s2 = ngpu.GetHotPathStats()
    p2 = ngpu.GetRk5Pass1Stats()
    ngpu.Simulate(timing_steps * dt)
    kt = ngpu.GetKernelTimes()
    s3 = ngpu.GetHotPathStats()
    p3 = ngpu.GetRk5Pass1Stats()
    ngpu.SetKernelTiming(0)
    ngpu.SetKernelStatus("fuse_delivery", fused)
    res["kernel_ms"] = {k: {"total_ms": v[0], "launches": v[1],
                            "avg_us": 1e3 * v[0] / max(v[1], 1)}
                        for k, v in kt.items()}
    acc1 = p3["rk5_pass1_accepted"] - p2["rk5_pass1_accepted"]
    rej1 = p3["rk5_pass1_rejected"] - p2["rk5_pass1_rejected"]
    ev_t = s3["syn_events"] - s2["syn_events"]
    rk = kt.get("aeif_rk5", (0.0, 0))
    dl = kt.get("delivery", (0.0, 0))
    n_launch = max(rk[1], 1)
    res["rk5_kernel"] = {
        "avg_us": 1e3 * rk[0] / n_launch,
        # first pass only: its own attempt counters
        "flop_per_launch": (RK5_FLOP_ACCEPTED * acc1 +
                            RK5_FLOP_REJECTED * rej1) / n_launch,
        "attempts_per_launch": (acc1 + rej1) / n_launch,
        "gbs": RK5_BYTES_PER_NEURON_STEP * neuron.n * rk[1] /
        max(rk[0] * 1e-3, 1e-12) / 1e9,
        "tflops": (RK5_FLOP_ACCEPTED * acc1 + RK5_FLOP_REJECTED * rej1) /
        max(rk[0] * 1e-3, 1e-12) / 1e12}
    res["delivery_kernel"] = {
        "avg_us": 1e3 * dl[0] / max(dl[1], 1),
        "events_per_launch": ev_t / max(dl[1], 1),
        "gbs": DELIVERY_BYTES_PER_EVENT * ev_t / max(dl[0] * 1e-3, 1e-12)
        / 1e9}
    res["n_neurons"] = neuron.n
    res["n_synapses"] = n_syn
    res["steps"] = steps
    # network construction: Connect calls (host checks; the rules are
    # generated on the device at calibration) and calibration
    c_s = BUILD.pop("connect_s", None)
    res["connect_s"] = None if c_s is None else round(c_s, 3)
    res["calibrate_s"] = round(t_cal, 3)
    if extra is not None:
        res["after_warmup"] = extra
    return res


def profiled_traffic(name="aeif_rk5_c1.md"):
    """HBM bytes per launch of the first RK5 pass on c1, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled against a copy kernel);
    None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"Scaled traffic: ([0-9.]+) MB per launch", txt)
    return float(m.group(1)) * 1e6 if m else None


def train_proposer(ngpu, neuron, log_steps, max_samples=10000000,
                   train_iter=5000, batch=8192, min_samples=100000):
    """Logs (state, accepted step) samples from the plain controller of the
    running network for up to `log_steps` steps (or `max_samples`), then
    trains the proposer; returns the weights and the training record, or
    None for the weights if the network gave fewer than `min_samples`."""
    dt = ngpu.GetTimeResolution()
    t = time.time()
    ngpu.ProposerLog(1, max_samples)
    done = 0
    while done < log_steps and ngpu.ProposerNSamples() < max_samples:
        k = min(200, log_steps - done)
        ngpu.Simulate(k * dt)
        done += k
    ngpu.ProposerLog(0, 0)
    n_smp = ngpu.ProposerNSamples()
    t_log = time.time() - t
    rec = {"samples": n_smp, "log_steps": done, "log_s": round(t_log, 2)}
    if n_smp < min_samples:
        log("  proposer: only %d samples in %d steps" % (n_smp, done))
        return None, rec
    t = time.time()
    loss = ngpu.ProposerTrain(train_iter, batch=batch, lr=1e-3)
    rec.update({"train_iter": train_iter, "batch": batch,
                "train_s": round(time.time() - t, 2),
                "train_loss": round(loss, 5)})
    log("  proposer: %d samples in %d steps (%.1f s), %d Adam steps of %d "
        "in %.1f s, loss %.4f" % (n_smp, done, t_log, train_iter, batch,
                                  rec["train_s"], loss))
    return ngpu.ProposerGetWeights(), rec


def learned_vs_plain(ngpu, build, settle, window, log_steps=8000,
                     weights=None, weights_from=None):
    """Stage evaluations of the plain and the learned RK5 controller over
    the same simulated window.  Run 1: settle, log samples from the plain
    controller and train (unless `weights` are given), then time `window`
    plain steps.  Run 2: the same network rebuilt and stepped with the
    plain controller to the same biological time, then `window` steps with
    the proposer."""
    dt = ngpu.GetTimeResolution()
    neuron, _ = build()
    ngpu.ActivateSpikeCount(neuron)
    ngpu.Calibrate()
    ngpu.Simulate(settle * dt)
    rec = {}
    if weights is None:
        weights, rec = train_proposer(ngpu, neuron, log_steps)
        pre = settle + rec["log_steps"]
    else:
        pre = settle
        rec = {"weights_from": weights_from}
    if weights is None:
        return None
    plain = window_stats(ngpu, neuron, window)
    neuron, _ = build()
    ngpu.ActivateSpikeCount(neuron)
    ngpu.Calibrate()
    ngpu.Simulate(pre * dt)
    ngpu.ProposerSetWeights(weights)
    ngpu.ProposerUse(1)
    learned = window_stats(ngpu, neuron, window)
    ngpu.ProposerUse(0)
    n_ns = neuron.n * window

    def per(x, k):
        return x[k] / n_ns

    p = plain["rk5_stages_per_neuron_step"]
    q = learned["rk5_stages_per_neuron_step"]
    out = {
        "window_steps": window, "window_starts_at_bio_ms": round(pre * dt, 1),
        # every neuron-step takes at least one RK5 attempt (6 stage
        # evaluations): input arrives every step, so no step crosses dt
        "stage_floor_per_neuron_step": 6.0,
        "plain": round(p, 5), "learned": round(q, 5),
        "reduction_pct": round(100.0 * (1.0 - q / p), 2) if p > 0 else None,
        "above_floor_reduction_pct": round(
            100.0 * (1.0 - (q - 6.0) / (p - 6.0)), 2) if p > 6.0 else None,
        "rejected_per_neuron_step": {
            "plain": round(per(plain, "rk5_rejected"), 6),
            "learned": round(per(learned, "rk5_rejected"), 6)},
        "rate_hz": {"plain": round(plain["rate_hz"], 3),
                    "learned": round(learned["rate_hz"], 3)},
        "wall_s_per_bio_s": {"plain": round(plain["wall_s_per_bio_s"], 4),
                             "learned": round(learned["wall_s_per_bio_s"],
                                              4)},
        "proposer": rec}
    return out, weights


def roofline(res, traffic=None):
    """Dominant kernel of the step against the roof that bounds it: the
    first RK5 pass against the FP32 vector peak (its HBM figure beside
    it), the delivery against HBM bandwidth."""
    km = res["kernel_ms"]
    cand = [k for k in ("aeif_rk5", "delivery") if k in km]
    dom = max(cand, key=lambda k: km[k]["total_ms"])
    if dom == "delivery":
        per_launch = DELIVERY_BYTES_PER_EVENT * res["delivery_kernel"][
            "events_per_launch"]
        ach = res["delivery_kernel"]["gbs"]
        return {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_launch": per_launch,
                "avg_launch_us": round(km[dom]["avg_us"], 2),
                "traffic": None}
    rk = res["rk5_kernel"]
    gbs = rk["gbs"]
    return {"kernel": dom, "bound": "valu",
            "achieved": round(rk["tflops"], 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(rk["tflops"] /
                                             FP32_PEAK_TFLOPS, 4),
            "algorithmic_flop_per_launch": rk["flop_per_launch"],
            "avg_launch_us": round(km[dom]["avg_us"], 2),
            "traffic": traffic,
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "algorithmic_bytes_per_launch":
                    RK5_BYTES_PER_NEURON_STEP * res["n_neurons"]}}


def lscpu():
    """Socket x core x SMT topology of the host, from lscpu."""
    import subprocess
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    keys = {"Model name": "model", "Socket(s)": "sockets",
            "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    out = {}
    for line in txt.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            out[keys[k.strip()]] = v.strip()
    return out


def cpu_sample_net(n, prec, seed=1234):
    """The c1 workload at n neurons in the CPU oracle: in-degree 800 + 200
    (sources drawn with replacement), same weights, delays and drive."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    n_e = int(round(0.8 * n))
    k_e, k_i = 800, 200
    net = po.Net(dt=0.1, seed=seed, prec=prec)
    g = net.create("aeif_cond_beta", n, 2)
    for k, v in {"C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
                 "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
                 "V_reset": -60.0, "t_ref": 0.0}.items():
        g.set(k, v)
    g.set("E_rev", [0.0, -85.0])
    g.set("tau_rise", [0.2, 0.5])
    g.set("tau_decay", [2.0, 10.0])
    g.set("V_m", rng.uniform(-70.0, -55.0, n).astype(np.float32))
    net.set_rk5(g, np.float32(1e-3), np.float32(1e-2), np.float32(1e-3))
    pg = net.create("poisson_generator", 1)
    net.poisson(pg.i0, 8000.0)
    tgt_e = np.repeat(np.arange(n, dtype=np.int32), k_e)
    src_e = rng.integers(0, n_e, n * k_e).astype(np.int32)
    net.connect(src_e, tgt_e, 0.05,
                (rng.integers(1, 21, n * k_e) * 0.1).astype(np.float32), 0)
    del src_e, tgt_e
    tgt_i = np.repeat(np.arange(n, dtype=np.int32), k_i)
    src_i = (n_e + rng.integers(0, n - n_e, n * k_i)).astype(np.int32)
    net.connect(src_i, tgt_i, 0.25,
                (rng.integers(1, 21, n * k_i) * 0.1).astype(np.float32), 1)
    del src_i, tgt_i
    net.connect(np.full(n, pg.i0, np.int32), np.arange(n, dtype=np.int32),
                0.1, 0.1, 0)
    net.calibrate()
    return net


def cpu_time(n, steps, warmup, threads, prec):
    net = cpu_sample_net(n, prec)
    net.run(warmup, threads)
    t0 = time.perf_counter()
    net.run(steps, threads)
    wall = time.perf_counter() - t0
    bio_s = steps * 0.1e-3
    return wall, wall / bio_s * (1.0e6 / n)


def cpu_baseline(n_all=100000, steps_all=1000, n_one=10000, steps_one=500,
                 warmup=100):
    """FP64 CPU oracle (OpenMP over neurons) on c1 samples: all the cores
    this process may use, and one thread."""
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    w_all, v_all = cpu_time(n_all, steps_all, warmup, cores, "f64")
    w_one, v_one = cpu_time(n_one, steps_one, warmup, 1, "f64")
    return {"value": round(v_all, 2), "unit": "wall-s/bio-s",
            "cores": cores, "kind": "port",
            "sample": ("FP64 CPU oracle (OpenMP over neurons), c1 workload "
                       "at %d neurons (indegree 800+200, same weights/"
                       "delays/drive; sources drawn with replacement), %d "
                       "steps (%.0f ms bio) after %d warm-up steps on %d "
                       "threads: %.2f s wall, scaled x%.0f in neurons to "
                       "1e6 and to 1 s bio"
                       % (n_all, steps_all, steps_all * 0.1, warmup, cores,
                          w_all, 1.0e6 / n_all)),
            "one_thread": {"value": round(v_one, 2),
                           "unit": "wall-s/bio-s",
                           "sample": "%d neurons, %d steps: %.2f s wall, "
                                     "scaled x%.0f" % (n_one, steps_one,
                                                       w_one, 1.0e6 / n_one)},
            "lscpu": lscpu()}


def active_object(r, name, workload, settle_ms):
    """Fields common to the active-network objects."""
    fired = r["rate_hz"] > 0.0 and r["syn_events"] > 0
    return {"workload": workload,
            "neurons": r["n_neurons"], "synapses": r["n_synapses"],
            "settle_bio_ms": settle_ms,
            "wall_s_per_bio_s": round(r["wall_s_per_bio_s"], 4),
            "ms_per_step": round(r["ms_per_step"], 4),
            "syn_events_per_s": round(r["syn_events_per_s"], 1),
            "rate_hz": round(r["rate_hz"], 3),
            "window_fired": fired,
            "rk5_stages_per_neuron_step": round(
                r["rk5_stages_per_neuron_step"], 4),
            "kernels": r["kernel_ms"],
            "roofline": roofline(r),
            "build_s": {"connect": r["connect_s"],
                        "calibrate": r["calibrate_s"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--neurons", type=int, default=1000000)
    ap.add_argument("--timing-steps", type=int, default=300)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="biological time the active networks run before "
                         "their warm-up and timed window")
    ap.add_argument("--stage-steps", type=int, default=1000,
                    help="window of the learned vs plain stage comparison")
    ap.add_argument("--no-readme", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned vs plain comparisons")
    ap.add_argument("--only", choices=["c1", "readme", "active_1m", "c2",
                                       "c3", "c4"],
                    help="measure one workload and print its object only "
                         "(profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workloads", default="active_1m,c2,c3,c4",
                    help="further configurations to measure after c1 and "
                         "the README network (comma list of active_1m, c2,"
                         " c3, c4; empty for none)")
    ap.add_argument("--set", action="append", default=[],
                    metavar="NAME=VALUE",
                    help="kernel int parameter set after each network is "
                         "built (SetKernelStatus), e.g. delivery_mode=3")
    ap.add_argument("--c4-instances", type=int, default=64)
    ap.add_argument("--c2-scale", type=float, default=1.0)
    ap.add_argument("--c3-scale", type=float, default=1.0)
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="also run a plain device copy of this size (MiB) "
                         "with torch, to scale rocprofv3 FETCH_SIZE/"
                         "WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU "
              "path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from neurongpu_amd import neurongpu as ngpu
    from neurongpu_amd import workloads as wl
    dt = 0.1
    settle = int(round(args.settle_ms / dt))
    wls = [w for w in args.workloads.split(",") if w]

    def c4_build():
        inst = wl.c4_instances()[:args.c4_instances]
        BUILD["c4_instances"] = len(inst)
        t = time.time()
        r = wl.build_ensemble(ngpu, inst)
        BUILD["connect_s"] = time.time() - t
        return r[0], r[2]

    builders = {
        "c1": lambda: build_c1(ngpu, args.neurons),
        "readme": lambda: build_readme(ngpu, args.neurons),
        "active_1m": lambda: build_readme(ngpu, args.neurons, tol=1e-3),
        "c2": lambda: build_c2(ngpu, args.c2_scale),
        "c3": lambda: build_c3(ngpu, args.c3_scale),
        "c4": c4_build,
    }

    settings = {}
    for kv in args.set:
        k, v = kv.split("=", 1)
        settings[k] = int(v)

    def run(name, after=None):
        log("building", name)
        neuron, n_syn = builders[name]()
        for k, v in settings.items():
            ngpu.SetKernelStatus(k, v)
        r = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                    args.timing_steps, after_warmup=after,
                    settle_steps=0 if name == "c1" else settle)
        log(name + ":", json.dumps(r))
        return r

    def c4_tune():
        from neurongpu_amd.autotune import autotune
        return autotune(ngpu, steps=40, log=log)

    if args.only:
        r = run(args.only, c4_tune if args.only == "c4" else None)
        r["roofline"] = roofline(r)
        out = {args.only: r}
        copy_calibration(args, out)
        print(json.dumps(out), flush=True)
        return
    c1 = run("c1")
    out = {
        "metric": METRIC,
        "value": round(c1["wall_s_per_bio_s"], 4),
        "unit": "wall-s/bio-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c1["ms_per_step"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded network, Threefry Poisson drive)",
        "config": {"workload": "c1: 1M aeif_cond_beta, 1e9 synapses "
                               "(800 E w=0.05, 200 I w=0.25), delays 1-20 "
                               "steps, 8 kHz Poisson w=0.1, RK5 tol 1e-3, "
                               "dt 0.1 ms (silent as specified: the drive "
                               "stays below rheobase)",
                   "neurons": c1["n_neurons"], "synapses": c1["n_synapses"]},
        "syn_events_per_s": round(c1["syn_events_per_s"], 1),
        "rate_hz": round(c1["rate_hz"], 3),
        "rk5_stages_per_neuron_step": round(
            c1["rk5_stages_per_neuron_step"], 4),
        "kernels": c1["kernel_ms"],
        "roofline": roofline(c1, profiled_traffic()),
        "build_s": {"connect": c1["connect_s"], "calibrate": c1["calibrate_s"]},
    }
    ngpu.ResetKernel()
    weights = {}
    if not args.no_readme:
        rd = run("readme")
        o = active_object(
            rd, "readme", "reference README benchmark network "
            "(c++/examples/brunel_net.cpp, 1M neurons, 1e9 synapses, 20 kHz "
            "Poisson, RK5 tol 1e-6)", args.settle_ms)
        o["reference_wall_s_per_bio_s"] = 70.0
        o["reference_hardware"] = "RTX 2080 Ti (README.md:11)"
        # the reference's figure is for an active network: compared only
        # when the timed window fired
        o["speedup_vs_reference"] = (round(70.0 / rd["wall_s_per_bio_s"], 2)
                                     if o["window_fired"] else None)
        if not args.no_learned:
            lr_ = learned_vs_plain(ngpu, builders["readme"], settle,
                                   args.stage_steps)
            if lr_:
                o["rk5_stages_learned_vs_plain"], weights["readme"] = lr_
                log("readme learned:",
                    json.dumps(o["rk5_stages_learned_vs_plain"]))
        out["readme_network"] = o
        ngpu.ResetKernel()
    if "active_1m" in wls:
        ra = run("active_1m")
        o = active_object(
            ra, "active_1m", "1M aeif_cond_beta / 1e9 synapses, active: the "
            "README network (20 kHz Poisson w 0.37, in-degree 800 E w 0.05 + "
            "200 I w 0.35) at RK5 tol 1e-3", args.settle_ms)
        o["target_wall_s_per_bio_s"] = 10.0
        out["active_1m"] = o
        ngpu.ResetKernel()
    if "c2" in wls:
        r2 = run("c2")
        out["c2"] = active_object(
            r2, "c2", "c2: 8 iaf_psc_exp populations (77169 neurons), "
            "p ~ U(0.01, 0.15), E w ~ N(87.8, 8.8) pA, I w = -4x, delays "
            "N(1.5, 0.75)/N(0.8, 0.4) ms, Poisson 8 Hz x K_ext, dt 0.1 ms",
            args.settle_ms)
        ngpu.ResetKernel()
    if "c3" in wls:
        r3 = run("c3")
        o = active_object(
            r3, "c3", "c3: 2e5 aeif_cond_alpha (80 % E), fixed in-degree "
            "500, STDP on E->E (tau 20 ms, lambda 1e-3, alpha 1.05, Wmax 0.1"
            " nS, w0 ~ U(0.02, 0.06)), own 8 kHz Poisson per neuron (w 0.3 "
            "nS), RK5 tol 1e-3", args.settle_ms)
        o["plastic_synapses"] = int(round(0.8 * r3["n_neurons"])) * 400
        if not args.no_learned:
            lr_ = learned_vs_plain(ngpu, builders["c3"], settle,
                                   args.stage_steps, log_steps=16000)
            if lr_:
                cmp3, weights["c3"] = lr_
                o["rk5_stages_learned_vs_plain"] = cmp3
                log("c3 learned:", json.dumps(cmp3))
                out["rk5_stages_learned_vs_plain"] = {
                    "workload": "c3", "plain": cmp3["plain"],
                    "learned": cmp3["learned"],
                    "stage_floor_per_neuron_step": 6.0,
                    "reduction_pct": cmp3["reduction_pct"],
                    "above_floor_reduction_pct":
                    cmp3["above_floor_reduction_pct"]}
        out["c3"] = o
        ngpu.ResetKernel()
    if "c4" in wls:
        r4 = run("c4", c4_tune)
        o = active_object(
            r4, "c4", "c4: %d independent aeif_cond_beta networks in one "
            "node group (sizes log-U[5e3, 5e4], in-degree U{100..1000}, "
            "weight scale U(0.5, 2) x (0.05, 0.25) nS, Poisson U(4, 12) kHz"
            " w 0.1), one launch per kernel per step, RK5 tol 1e-3"
            % BUILD.get("c4_instances", 0), args.settle_ms)
        o["instances"] = BUILD.get("c4_instances", 0)
        o["autotuned"] = r4.get("after_warmup")
        o["wall_s_per_bio_s_per_instance"] = round(
            r4["wall_s_per_bio_s"] / max(o["instances"], 1), 5)
        if not args.no_learned:
            # the ensemble is silent as specified (no second-pass steps to
            # log), so the proposer trained on the README network is used
            src = "readme" if "readme" in weights else (
                "c3" if "c3" in weights else None)
            lr_ = learned_vs_plain(ngpu, c4_build, settle, args.stage_steps,
                                   weights=weights.get(src),
                                   weights_from=src) if src else None
            if lr_:
                o["rk5_stages_learned_vs_plain"] = lr_[0]
        out["c4"] = o
        ngpu.ResetKernel()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    copy_calibration(args, out)
    print(json.dumps(out), flush=True)


def copy_calibration(args, out):
    """A plain device copy of --copy-calib-mib MiB (3 launches), whose PMC
    counters scale FETCH_SIZE / WRITE_SIZE to bytes."""
    if args.copy_calib_mib <= 0:
        return
    import torch
    n = args.copy_calib_mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    out["copy_calibration_bytes"] = 4 * n


if __name__ == "__main__":
    main()
