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

CPU baseline: the FP32 CPU oracle (OpenMP over neurons) on a bounded
sample of the c1 workload (reduced neuron count, same per-neuron
indegree, weights, delays and drive), scaled linearly to 1 M neurons and
1 s of biological time.
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
    rng = np.random.default_rng(seed)
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    n_e = int(round(0.8 * n))
    neuron = ngpu.Create("aeif_cond_beta", n, 2)
    ngpu.SetStatus(neuron, {
        "C_m": 281.0, "g_L": 30.0, "E_L": -70.6, "V_th": -50.4,
        "Delta_T": 2.0, "a": 4.0, "tau_w": 144.0, "b": 80.5,
        "V_reset": -60.0, "t_ref": 0.0, "E_rev": [0.0, -85.0],
        "tau_rise": [0.2, 0.5], "tau_decay": [2.0, 10.0]})
    ngpu.SetStatus(neuron, {"rk5_tol": tol})
    t = time.time()
    v0 = rng.uniform(-70.0, -55.0, n).astype(np.float32)
    ngpu.SetStatus(neuron, {"V_m": {"array": v0}})
    log("  V0 set in %.1f s" % (time.time() - t))
    pg = ngpu.Create("poisson_generator", 1)
    ngpu.SetStatus(pg, {"rate": 8000.0})
    exc = neuron[0:n_e]
    inh = neuron[n_e:n]
    k_e, k_i = 800, 200
    t = time.time()
    d_e = (rng.integers(1, 21, n * k_e, dtype=np.int8).astype(np.float32)
           * np.float32(0.1))
    ngpu.Connect(exc, neuron, {"rule": "fixed_indegree", "indegree": k_e},
                 {"weight": 0.05, "delay": {"array": d_e}, "receptor": 0})
    del d_e
    d_i = (rng.integers(1, 21, n * k_i, dtype=np.int8).astype(np.float32)
           * np.float32(0.1))
    ngpu.Connect(inh, neuron, {"rule": "fixed_indegree", "indegree": k_i},
                 {"weight": 0.25, "delay": {"array": d_i}, "receptor": 1})
    del d_i
    ngpu.Connect(pg, neuron, {"rule": "all_to_all"},
                 {"weight": 0.1, "delay": 0.1, "receptor": 0})
    BUILD["connect_s"] = time.time() - t
    log("  connections in %.1f s" % BUILD["connect_s"])
    return neuron, n * (k_e + k_i)


def build_readme(ngpu, n=1000000, seed=1234):
    """c++/examples/brunel_net.cpp:31-99 at n_neurons = n."""
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
    t = time.time()
    d_e = ngpu.RandomNormalClipped(ce * (n_e + n_i), 0.5, 0.25, 0.1,
                                   0.5 + 3 * 0.25)
    ngpu.Connect(exc, neuron, {"rule": "fixed_indegree", "indegree": ce},
                 {"weight": 0.05, "delay": {"array": d_e}, "receptor": 0})
    del d_e
    d_i = ngpu.RandomNormalClipped(ci * (n_e + n_i), 0.5, 0.25, 0.1,
                                   0.5 + 3 * 0.25)
    ngpu.Connect(inh, neuron, {"rule": "fixed_indegree", "indegree": ci},
                 {"weight": 0.35, "delay": {"array": d_i}, "receptor": 1})
    del d_i
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


def measure(ngpu, neuron, n_syn, steps, warmup, timing_steps,
            after_warmup=None):
    dt = ngpu.GetTimeResolution()
    ngpu.ActivateSpikeCount(neuron)
    t = time.time()
    ngpu.Calibrate()
    t_cal = time.time() - t
    log("  calibrate (CSR build + upload) in %.1f s" % t_cal)
    if warmup > 0:
        ngpu.Simulate(warmup * dt)
    extra = after_warmup() if after_warmup else None
    s0 = ngpu.GetHotPathStats()
    c0 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n,
                                         "spike_count")))
    hip_sync()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * dt)
    hip_sync()
    wall = time.perf_counter() - t0
    s1 = ngpu.GetHotPathStats()
    c1 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n,
                                         "spike_count")))
    bio_s = steps * dt * 1e-3
    events = s1["syn_events"] - s0["syn_events"]
    acc = s1["rk5_accepted"] - s0["rk5_accepted"]
    rej = s1["rk5_rejected"] - s0["rk5_rejected"]
    spikes = c1 - c0
    res = {
        "wall_s": wall, "bio_s": bio_s, "wall_s_per_bio_s": wall / bio_s,
        "ms_per_step": 1e3 * wall / steps, "syn_events": events,
        "syn_events_per_s": events / wall,
        "rate_hz": spikes / neuron.n / bio_s,
        "rk5_accepted": acc, "rk5_rejected": rej,
        "rk5_stages_per_neuron_step": (6.0 * acc + 5.0 * rej) /
        (neuron.n * steps),
    }
    # per-kernel device time (HIP events on the engine stream); the second
    # RK5 pass and the delivery normally share one launch, so they are
    # timed here as separate launches
    fused = ngpu.GetKernelStatus("fuse_delivery")
    ngpu.SetKernelStatus("fuse_delivery", 0)
    ngpu.SetKernelTiming(1)
    s2 = ngpu.GetHotPathStats()
    ngpu.Simulate(timing_steps * dt)
    kt = ngpu.GetKernelTimes()
    s3 = ngpu.GetHotPathStats()
    ngpu.SetKernelTiming(0)
    ngpu.SetKernelStatus("fuse_delivery", fused)
    res["kernel_ms"] = {k: {"total_ms": v[0], "launches": v[1],
                            "avg_us": 1e3 * v[0] / max(v[1], 1)}
                        for k, v in kt.items()}
    acc_t = s3["rk5_accepted"] - s2["rk5_accepted"]
    rej_t = s3["rk5_rejected"] - s2["rk5_rejected"]
    ev_t = s3["syn_events"] - s2["syn_events"]
    rk = kt.get("aeif_rk5", (0.0, 0))
    dl = kt.get("delivery", (0.0, 0))
    res["rk5_kernel"] = {
        "avg_us": 1e3 * rk[0] / max(rk[1], 1),
        "gbs": RK5_BYTES_PER_NEURON_STEP * neuron.n * rk[1] /
        max(rk[0] * 1e-3, 1e-12) / 1e9,
        "tflops": (RK5_FLOP_ACCEPTED * acc_t + RK5_FLOP_REJECTED * rej_t) /
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


def learned_vs_plain(ngpu, build, steps, warmup, max_samples=10000000,
                     max_log_steps=4000, train_iter=5000, batch=8192):
    """Train the learned RK5 step-size proposer on samples logged from the
    plain controller of the running network (after its plain measurement),
    then rebuild the same network, warm it up with the plain controller and
    time `steps` steps with the proposer on: the same simulated window as
    the plain measurement."""
    dt = ngpu.GetTimeResolution()
    t = time.time()
    ngpu.ProposerLog(1, max_samples)
    done = 0
    while done < max_log_steps and ngpu.ProposerNSamples() < max_samples:
        ngpu.Simulate(200 * dt)
        done += 200
    ngpu.ProposerLog(0, 0)
    n_smp = ngpu.ProposerNSamples()
    t_log = time.time() - t
    t = time.time()
    loss = ngpu.ProposerTrain(train_iter, batch=batch, lr=1e-3)
    t_train = time.time() - t
    w = ngpu.ProposerGetWeights()
    log("  proposer: %d samples in %d steps (%.1f s), %d Adam steps of %d "
        "in %.1f s, loss %.4f" % (n_smp, done, t_log, train_iter, batch,
                                  t_train, loss))
    neuron, _ = build()
    ngpu.ActivateSpikeCount(neuron)
    ngpu.Calibrate()
    if warmup > 0:
        ngpu.Simulate(warmup * dt)
    ngpu.ProposerSetWeights(w)
    ngpu.ProposerUse(1)
    s0 = ngpu.GetHotPathStats()
    c0 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n, "spike_count")))
    hip_sync()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * dt)
    hip_sync()
    wall = time.perf_counter() - t0
    s1 = ngpu.GetHotPathStats()
    c1 = int(np.sum(ngpu.GetNeuronIntVar(neuron.i0, neuron.n, "spike_count")))
    ngpu.ProposerUse(0)
    acc = s1["rk5_accepted"] - s0["rk5_accepted"]
    rej = s1["rk5_rejected"] - s0["rk5_rejected"]
    bio_s = steps * dt * 1e-3
    return {"samples": n_smp, "log_steps": done, "train_iter": train_iter,
            "batch": batch, "train_s": round(t_train, 2),
            "train_loss": round(loss, 5),
            "stages_per_neuron_step": (6.0 * acc + 5.0 * rej) /
            (neuron.n * steps),
            "rejected_per_neuron_step": rej / (neuron.n * steps),
            "wall_s_per_bio_s": wall / bio_s,
            "rate_hz": (c1 - c0) / neuron.n / bio_s}


def stage_comparison(plain, learned):
    p = plain["rk5_stages_per_neuron_step"]
    q = learned["stages_per_neuron_step"]
    return {"plain": round(p, 5), "learned": round(q, 5),
            "reduction_pct": round(100.0 * (1.0 - q / p), 2),
            "reduction_beyond_one_attempt_pct": round(
                100.0 * (1.0 - (q - 6.0) / (p - 6.0)), 2) if p > 6.0 else None,
            "rejected_per_neuron_step": {
                "plain": round(plain["rk5_rejected"] /
                               (plain["n_neurons"] * plain["steps"]), 6),
                "learned": round(learned["rejected_per_neuron_step"], 6)},
            "rate_hz": {"plain": round(plain["rate_hz"], 3),
                        "learned": round(learned["rate_hz"], 3)},
            "wall_s_per_bio_s": {
                "plain": round(plain["wall_s_per_bio_s"], 4),
                "learned": round(learned["wall_s_per_bio_s"], 4)},
            "proposer": {k: learned[k] for k in
                         ("samples", "log_steps", "train_iter", "batch",
                          "train_s", "train_loss")}}


def roofline(res, traffic=None):
    """Dominant kernel of the step and its algorithmic traffic."""
    km = res["kernel_ms"]
    # candidates with an algorithmic byte count per launch: the first RK5
    # pass (72 B per neuron-step) and the delivery (9 B per event); the
    # second RK5 pass is latency-bound on the slowest neurons
    cand = [k for k in ("aeif_rk5", "delivery") if k in km]
    dom = max(cand, key=lambda k: km[k]["total_ms"])
    if dom == "delivery":
        per_launch = DELIVERY_BYTES_PER_EVENT * res["delivery_kernel"][
            "events_per_launch"]
        ach = res["delivery_kernel"]["gbs"]
    else:
        per_launch = RK5_BYTES_PER_NEURON_STEP * res["n_neurons"]
        ach = res["rk5_kernel"]["gbs"]
    return {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_us": round(km[dom]["avg_us"], 2),
            "traffic": traffic if dom == "aeif_rk5" else None}


def cpu_baseline(n_sample=100000, steps=500, warmup=100, seed=1234):
    """FP32 oracle (OpenMP over neurons) on a c1 sample."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    n = n_sample
    n_e = int(round(0.8 * n))
    k_e, k_i = 800, 200
    net = po.Net(dt=0.1, seed=seed, prec="f32")
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
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    net.run(warmup, cores)
    t0 = time.perf_counter()
    net.run(steps, cores)
    wall = time.perf_counter() - t0
    bio_s = steps * 0.1e-3
    scaled = wall / bio_s * (1.0e6 / n)
    return {"value": round(scaled, 2), "unit": "wall-s/bio-s",
            "cores": cores, "kind": "port",
            "sample": ("FP32 CPU oracle, c1 workload at %d neurons "
                       "(indegree 800+200, same weights/delays/drive; sources"
                       " drawn with replacement), %d steps (%.0f ms bio) "
                       "after %d warm-up steps, %.2f s wall, scaled x%.0f "
                       "in neurons to 1e6 and to 1 s bio"
                       % (n, steps, bio_s * 1e3, warmup, wall, 1.0e6 / n))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--neurons", type=int, default=1000000)
    ap.add_argument("--timing-steps", type=int, default=300)
    ap.add_argument("--no-readme", action="store_true")
    ap.add_argument("--only-readme", action="store_true",
                    help="measure only the README network (profiling)")
    ap.add_argument("--only", choices=["c1", "readme", "c2", "c3", "c4"],
                    help="measure one workload and print its object only "
                         "(profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workloads", default="c2,c3,c4",
                    help="further configurations to measure after c1 "
                         "(comma list of c2, c3, c4; empty for none)")
    ap.add_argument("--c4-instances", type=int, default=64)
    ap.add_argument("--c2-scale", type=float, default=1.0)
    ap.add_argument("--c3-scale", type=float, default=1.0)
    ap.add_argument("--cpu-neurons", type=int, default=100000)
    ap.add_argument("--cpu-steps", type=int, default=2000)
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

    if args.only_readme:
        args.only = "readme"
    if args.only:
        from neurongpu_amd import workloads as wl
        after = None
        if args.only == "c1":
            neuron, n_syn = build_c1(ngpu, args.neurons)
        elif args.only == "readme":
            neuron, n_syn = build_readme(ngpu, args.neurons)
        elif args.only == "c2":
            neuron, n_syn = build_c2(ngpu, args.c2_scale)
        elif args.only == "c3":
            neuron, n_syn = build_c3(ngpu, args.c3_scale)
        else:
            from neurongpu_amd.autotune import autotune
            inst = wl.c4_instances()[:args.c4_instances]
            neuron, _, n_syn = wl.build_ensemble(ngpu, inst)
            after = (lambda: autotune(ngpu, steps=40, log=log))
        r = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                    args.timing_steps, after_warmup=after)
        r["roofline"] = roofline(r)
        out = {args.only: r}
        copy_calibration(args, out)
        print(json.dumps(out), flush=True)
        return
    log("building c1 (%d neurons)" % args.neurons)
    neuron, n_syn = build_c1(ngpu, args.neurons)
    c1 = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                 args.timing_steps)
    log("c1:", json.dumps(c1))
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
        "data": "synthetic (seeded network, Philox Poisson drive)",
        "config": {"workload": "c1: 1M aeif_cond_beta, 1e9 synapses "
                               "(800 E w=0.05, 200 I w=0.25), delays 1-20 "
                               "steps, 8 kHz Poisson w=0.1, RK5 tol 1e-3, "
                               "dt 0.1 ms",
                   "neurons": c1["n_neurons"], "synapses": c1["n_synapses"]},
        "syn_events_per_s": round(c1["syn_events_per_s"], 1),
        "rate_hz": round(c1["rate_hz"], 3),
        "rk5_stages_per_neuron_step": round(
            c1["rk5_stages_per_neuron_step"], 4),
        "kernels": c1["kernel_ms"],
        "rk5_valu": {"achieved": round(c1["rk5_kernel"]["tflops"], 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(c1["rk5_kernel"]["tflops"] /
                                   FP32_PEAK_TFLOPS, 4)},
        "roofline": roofline(c1, profiled_traffic()),
        "build_s": {"connect": c1["connect_s"], "calibrate": c1["calibrate_s"]},
    }
    if not args.no_readme:
        ngpu.ResetKernel()
        log("building readme network (%d neurons)" % args.neurons)
        neuron, n_syn = build_readme(ngpu, args.neurons)
        rd = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                     args.timing_steps)
        log("readme:", json.dumps(rd))
        lr_ = learned_vs_plain(ngpu, lambda: build_readme(ngpu, args.neurons),
                               args.steps, args.warmup)
        log("readme learned:", json.dumps(lr_))
        out["readme_network"] = {
            "rk5_stages_learned_vs_plain": stage_comparison(rd, lr_),
            "workload": "reference README benchmark network "
                        "(c++/examples/brunel_net.cpp, 1M neurons, 1e9 "
                        "synapses, 20 kHz Poisson, RK5 tol 1e-6)",
            "wall_s_per_bio_s": round(rd["wall_s_per_bio_s"], 4),
            "ms_per_step": round(rd["ms_per_step"], 4),
            "syn_events_per_s": round(rd["syn_events_per_s"], 1),
            "rate_hz": round(rd["rate_hz"], 3),
            "rk5_stages_per_neuron_step": round(
                rd["rk5_stages_per_neuron_step"], 4),
            "reference_wall_s_per_bio_s": 70.0,
            "reference_hardware": "RTX 2080 Ti (README.md:11)",
            "speedup_vs_reference": round(70.0 / rd["wall_s_per_bio_s"], 2),
            "kernels": rd["kernel_ms"],
            "roofline": roofline(rd),
            "build_s": {"connect": rd["connect_s"],
                        "calibrate": rd["calibrate_s"]}}
        ngpu.ResetKernel()
    if "c2" in args.workloads.split(","):
        log("building c2 (%.3g x 77169 iaf_psc_exp neurons)" % args.c2_scale)
        neuron, n_syn = build_c2(ngpu, args.c2_scale)
        r2 = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                     args.timing_steps)
        log("c2:", json.dumps(r2))
        out["c2"] = {
            "workload": "c2: 8 iaf_psc_exp populations (77169 neurons), "
                        "p ~ U(0.01, 0.15), E w ~ N(87.8, 8.8) pA, "
                        "I w = -4x, delays N(1.5, 0.75)/N(0.8, 0.4) ms, "
                        "Poisson 8 Hz x K_ext, dt 0.1 ms",
            "neurons": r2["n_neurons"], "synapses": r2["n_synapses"],
            "wall_s_per_bio_s": round(r2["wall_s_per_bio_s"], 4),
            "ms_per_step": round(r2["ms_per_step"], 4),
            "syn_events_per_s": round(r2["syn_events_per_s"], 1),
            "rate_hz": round(r2["rate_hz"], 3),
            "kernels": r2["kernel_ms"],
            "roofline": roofline(r2),
            "build_s": {"connect": r2["connect_s"],
                        "calibrate": r2["calibrate_s"]}}
        ngpu.ResetKernel()
    if "c3" in args.workloads.split(","):
        log("building c3 (%.3g x 2e5 aeif_cond_alpha neurons, STDP)"
            % args.c3_scale)
        neuron, n_syn = build_c3(ngpu, args.c3_scale)
        r3 = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                     args.timing_steps)
        log("c3:", json.dumps(r3))
        l3 = learned_vs_plain(ngpu, lambda: build_c3(ngpu, args.c3_scale),
                              args.steps, args.warmup)
        log("c3 learned:", json.dumps(l3))
        cmp3 = stage_comparison(r3, l3)
        out["rk5_stages_learned_vs_plain"] = dict(
            workload="c3", plain=cmp3["plain"], learned=cmp3["learned"],
            reduction_pct=cmp3["reduction_pct"])
        out["c3"] = {
            "rk5_stages_learned_vs_plain": cmp3,
            "workload": "c3: 2e5 aeif_cond_alpha (80 % E), fixed in-degree "
                        "500, STDP on E->E (tau 20 ms, lambda 1e-3, alpha "
                        "1.05, Wmax 0.1 nS, w0 ~ U(0.02, 0.06)), own 8 kHz "
                        "Poisson per neuron (w 0.3 nS), RK5 tol 1e-3",
            "neurons": r3["n_neurons"], "synapses": r3["n_synapses"],
            "plastic_synapses": int(round(0.8 * r3["n_neurons"])) * 400,
            "wall_s_per_bio_s": round(r3["wall_s_per_bio_s"], 4),
            "ms_per_step": round(r3["ms_per_step"], 4),
            "syn_events_per_s": round(r3["syn_events_per_s"], 1),
            "rate_hz": round(r3["rate_hz"], 3),
            "rk5_stages_per_neuron_step": round(
                r3["rk5_stages_per_neuron_step"], 4),
            "kernels": r3["kernel_ms"],
            "roofline": roofline(r3),
            "build_s": {"connect": r3["connect_s"],
                        "calibrate": r3["calibrate_s"]}}
        ngpu.ResetKernel()
    if "c4" in args.workloads.split(","):
        from neurongpu_amd import workloads as wl
        from neurongpu_amd.autotune import autotune
        inst = wl.c4_instances()
        if args.c4_instances < len(inst):
            inst = inst[:args.c4_instances]
        log("building c4 (%d AdEx networks, %d neurons)"
            % (len(inst), sum(x["n"] for x in inst)))
        t = time.time()
        neuron, ranges, n_syn = wl.build_ensemble(ngpu, inst)
        BUILD["connect_s"] = time.time() - t
        log("  connections in %.1f s" % BUILD["connect_s"])
        r4 = measure(ngpu, neuron, n_syn, args.steps, args.warmup,
                     args.timing_steps,
                     after_warmup=lambda: autotune(ngpu, steps=40, log=log))
        tuned = r4["after_warmup"]
        log("  autotuned:", tuned)
        log("c4:", json.dumps(r4))
        out["c4"] = {
            "workload": "c4: %d independent aeif_cond_beta networks in one "
                        "node group (sizes log-U[5e3, 5e4], in-degree "
                        "U{100..1000}, weight scale U(0.5, 2) x (0.05, "
                        "0.25) nS, Poisson U(4, 12) kHz w 0.1), one launch "
                        "per kernel per step, RK5 tol 1e-3" % len(inst),
            "instances": len(inst),
            "neurons": r4["n_neurons"], "synapses": r4["n_synapses"],
            "autotuned": tuned,
            "wall_s_per_bio_s": round(r4["wall_s_per_bio_s"], 4),
            "wall_s_per_bio_s_per_instance": round(
                r4["wall_s_per_bio_s"] / len(inst), 5),
            "ms_per_step": round(r4["ms_per_step"], 4),
            "syn_events_per_s": round(r4["syn_events_per_s"], 1),
            "rate_hz": round(r4["rate_hz"], 3),
            "rk5_stages_per_neuron_step": round(
                r4["rk5_stages_per_neuron_step"], 4),
            "kernels": r4["kernel_ms"],
            "roofline": roofline(r4),
            "build_s": {"connect": r4["connect_s"],
                        "calibrate": r4["calibrate_s"]}}
        ngpu.ResetKernel()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_neurons, args.cpu_steps)
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
