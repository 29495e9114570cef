"""Learned RK5 step-size proposer: FP32 training kernels against a numpy
FP64 restatement of the same MLP, convergence on a known function, and
the end-to-end invariant (the RK5 error test still judges every attempt,
so network statistics stay those of the plain controller)."""
import sys

import numpy as np
import pytest

import netbuild as nb

pytestmark = pytest.mark.gpu

NI, NH = 6, 32
W1, B1 = 0, NH * NI
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
NP = B3 + 1


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mlp_loss_grad(w, x, t):
    """Mean of 0.5 (y - t)^2 over the batch and its gradient (FP64)."""
    w = w.astype(np.float64)
    A1 = w[W1:B1].reshape(NH, NI)
    b1 = w[B1:W2]
    A2 = w[W2:B2].reshape(NH, NH)
    b2 = w[B2:W3]
    a3 = w[W3:B3]
    b3 = w[B3]
    z1 = x @ A1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ A2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ a3 + b3
    dy = (y - t) / len(t)
    g = np.zeros(NP)
    g[B3] = dy.sum()
    g[W3:B3] = h2.T @ dy
    d2 = np.outer(dy, a3) * (z2 > 0)
    g[B2:W3] = d2.sum(0)
    g[W2:B2] = (d2.T @ h1).ravel()
    d1 = (d2 @ A2) * (z1 > 0)
    g[B1:W2] = d1.sum(0)
    g[W1:B1] = (d1.T @ x).ravel()
    return 0.5 * np.mean((y - t) ** 2), g


def test_training_step_matches_numpy(ngpu):
    """One full-batch SGD step of the HIP training kernels equals
    w - lr * grad of the FP64 restatement."""
    log("case proposer_grad")
    rng = np.random.default_rng(5)
    w0 = rng.normal(0.0, 0.3, NP).astype(np.float32)
    x = rng.normal(0.0, 1.0, (512, NI)).astype(np.float32)
    t = rng.normal(0.0, 1.0, 512).astype(np.float32)
    smp = np.zeros((512, 8), np.float32)
    smp[:, :NI] = x
    smp[:, NI] = t
    ngpu.ProposerSetWeights(w0)
    ngpu.ProposerSetSamples(smp)
    lr = 0.05
    loss = ngpu.ProposerTrain(1, batch=512, lr=lr, adam=False,
                              full_batch=True)
    w1 = ngpu.ProposerGetWeights()
    l_ref, g = mlp_loss_grad(w0, x.astype(np.float64), t.astype(np.float64))
    want = w0.astype(np.float64) - lr * g
    log("loss", loss, "numpy", l_ref, "max |dw|",
        np.max(np.abs(w1 - want)))
    assert abs(loss - l_ref) <= 1e-4 * l_ref
    assert np.allclose(w1, want, rtol=1e-4, atol=2e-6)


def test_training_fits_known_function(ngpu):
    """Adam on samples of t = 0.5 x0 - 0.3 x4 + 0.2 |x1| reaches a mean
    squared error below 2e-3."""
    log("case proposer_fit")
    rng = np.random.default_rng(7)
    x = rng.normal(0.0, 1.0, (100000, NI)).astype(np.float32)
    t = 0.5 * x[:, 0] - 0.3 * x[:, 4] + 0.2 * np.abs(x[:, 1])
    smp = np.zeros((x.shape[0], 8), np.float32)
    smp[:, :NI] = x
    smp[:, NI] = t
    ngpu.ProposerSetSamples(smp)
    loss0 = ngpu.ProposerTrain(1, batch=4096, lr=1e-3)
    loss = ngpu.ProposerTrain(3000, batch=4096, lr=3e-3)
    log("loss", loss0, "->", loss)
    assert loss < 2e-3


def bf16(a):
    """Round FP32 values to BF16 (nearest even) and back."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_mfma_inference_matches_numpy(ngpu):
    """The matrix-core inference (BF16 weights and activations, FP32
    accumulation) matches a numpy evaluation with the same BF16 rounding:
    relative error below 5e-3 everywhere (an activation can round to the
    neighbouring BF16 value when FP32 and FP64 sums straddle a rounding
    boundary) and below 1e-4 on average, for a sample count that is not a
    multiple of the wave size."""
    log("case proposer_infer")
    rng = np.random.default_rng(9)
    w = rng.normal(0.0, 0.3, NP).astype(np.float32)
    ngpu.ProposerSetWeights(w)
    x = rng.normal(0.0, 1.0, (1000, NI)).astype(np.float32)
    got = ngpu.ProposerEval(x)
    A1 = bf16(w[W1:B1]).reshape(NH, NI).astype(np.float64)
    A2 = bf16(w[W2:B2]).reshape(NH, NH).astype(np.float64)
    h1 = np.maximum(bf16(x).astype(np.float64) @ A1.T + w[B1:W2], 0)
    h2 = np.maximum(bf16(h1).astype(np.float64) @ A2.T + w[B2:W3], 0)
    want = h2 @ w[W3:B3].astype(np.float64) + w[B3]
    rel = np.abs(got - want) / (1.0 + np.abs(want))
    log("max rel err", rel.max(), "mean", rel.mean())
    assert rel.max() < 5e-3
    assert rel.mean() < 1e-4


def _run_window(ngpu, spec, n, settle_ms, window_ms, weights=None,
                log_samples=0, settings=None):
    """Builds `spec`, runs `settle_ms` with the plain controller (logging
    up to `log_samples` proposer samples), then `window_ms` with the
    proposer when `weights` are given; returns the spike trains of the
    window, its RK5 counts and the logged samples' count."""
    groups, _ = nb.build_engine(ngpu, spec)
    for k, v in (settings or {}).items():
        ngpu.SetKernelStatus(k, v)
    ngpu.ActivateRecSpikeTimes(groups[0], 4000)
    if log_samples:
        ngpu.ProposerLog(1, log_samples)
    ngpu.Simulate(settle_ms)
    n_smp = ngpu.ProposerNSamples() if log_samples else 0
    if log_samples:
        ngpu.ProposerLog(0, 0)
    if weights is not None:
        ngpu.ProposerSetWeights(weights)
        ngpu.ProposerUse(1)
    s0 = ngpu.GetHotPathStats()
    ngpu.Simulate(window_ms)
    s1 = ngpu.GetHotPathStats()
    ngpu.ProposerUse(0)
    trains = [np.array([t for t in ngpu.GetRecSpikeTimes(i)
                        if t > settle_ms], dtype=np.float64)
              for i in range(n)]
    cnt = {k: s1[k] - s0[k] for k in ("rk5_accepted", "rk5_rejected",
                                      "rk5_forced")}
    return trains, cnt, n_smp


@pytest.mark.timeout(600)
@pytest.mark.parametrize("settings", [{}, {"delivery_mode": 1}])
def test_proposer_in_step_keeps_statistics(ngpu, settings):
    """The proposer inside the simulation step, on the reference's Brunel
    test network at 10 000 neurons (RK5 tolerance 1e-6, 20 kHz drive):
    samples logged from the plain controller, training, then the second RK5
    pass starting its steps from the proposer's step size.  The RK5 error
    estimate still judges every attempt: over the same 500 ms window the
    excitatory and inhibitory rates and mean CV ISI stay within 2 % of
    the plain controller's and of the FP64 oracle's, no attempt passes
    the error test above tolerance except at the minimum step size (a
    device counter; counted in both runs), and the proposer takes fewer
    stage evaluations than the plain controller.  With binned delivery
    (delivery_mode 1) the second pass, proposer included, overlaps the next
    step's first pass."""
    log("case proposer_in_step", settings)
    order, settle, window = 2000, 300.0, 500.0
    n, ne = 5 * order, 4 * order
    spec = nb.brunel_spec(order=order)
    plain, c_plain, n_smp = _run_window(ngpu, spec, n, settle, window,
                                        log_samples=2000000,
                                        settings=settings)
    log("samples", n_smp)
    assert n_smp >= 100000
    loss = ngpu.ProposerTrain(3000, batch=8192, lr=1e-3)
    w = ngpu.ProposerGetWeights()
    log("training loss", loss)
    learned, c_learned, _ = _run_window(ngpu, spec, n, settle, window,
                                        weights=w, settings=settings)
    net, _ = nb.build_oracle(spec, "f64")
    net.run(int(round((settle + window) / 0.1)), 16)
    orc = [ts[ts > settle] for ts in nb.oracle_times_by_neuron(net, 0.1, 0,
                                                                n)]
    steps = int(round(window / 0.1)) * n

    def stages(c):
        return (6.0 * c["rk5_accepted"] + 5.0 * c["rk5_rejected"]) / steps

    log("stages per neuron-step plain", stages(c_plain), "learned",
        stages(c_learned), "forced plain", c_plain["rk5_forced"],
        "learned", c_learned["rk5_forced"])
    t1 = settle + window
    for name, sl in (("E", slice(0, ne)), ("I", slice(ne, n))):
        k = sl.stop - sl.start
        sp = nb.spike_stats(plain[sl], k, settle, t1, n_pairs=0)
        sl_ = nb.spike_stats(learned[sl], k, settle, t1, n_pairs=0)
        so = nb.spike_stats(orc[sl], k, settle, t1, n_pairs=0)
        log(name, "plain", sp, "learned", sl_, "fp64 oracle", so)
        for ref in (sp, so):
            assert abs(sl_["rate"] - ref["rate"]) <= 0.02 * ref["rate"]
            assert abs(sl_["cv_isi"] - ref["cv_isi"]) <= 0.02 * ref["cv_isi"]
    # an attempt above tolerance passes only at the minimum step size (at
    # spike onsets, where the exponential term outruns any step at tolerance
    # 1e-6); the proposer does not make that more frequent
    f_p = c_plain["rk5_forced"] / steps
    f_l = c_learned["rk5_forced"] / steps
    assert f_p < 0.02
    assert f_l <= 1.1 * f_p
    assert stages(c_learned) < stages(c_plain)
