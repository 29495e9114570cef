"""Launch-shape autotuner.

Picks the kernel launch settings of the per-step path by timing short
stretches of the running simulation.  The settings are:
* the first-pass RK5 attempt budget and the second-pass grid;
* the delivery mode (one float atomic per synaptic event, or events
  binned by accumulator range and summed in LDS first), the delivery grid
  and the largest number of (spike, delay) pairs per wave tile;
* whether the delivery shares the second RK5 pass's launch, and how many
  steps in a row the second pass may overlap the next step;
* whether Simulate replays captured hipGraphs of 16 steps or launches
  every step from the host.

None of these changes results: each neuron's arithmetic and each synapse's
delivery are the same under every setting. The tests check this spike for
spike. The tuner advances the simulation while it measures, so call it
during warm-up.
"""
import time

CANDIDATES = {
    "rk5_budget": [1, 2],
    "rk5_pass2_grid": [256, 64, 1024],
    "delivery_mode": [2, 0, 1],
    "delivery_grid": [1024, 2048, 4096],
    "delivery_tile_max": [8, 16, 64],
    "fuse_delivery": [1, 0],
    "step_overlap": [2, 0, 1, 4],
    "graph_steps": [0, 16],
}


def _time_steps(ngpu, steps):
    dt = ngpu.GetTimeResolution()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * dt)   # returns after the stream has drained
    return (time.perf_counter() - t0) / steps


def autotune(ngpu, steps=40, rounds=1, candidates=None, log=None):
    """Coordinate descent over `candidates` (default CANDIDATES): for each
    setting in turn, time `steps` simulation steps per value and keep the
    fastest. Returns {setting: chosen value, "step_us": time per step}."""
    cand = dict(CANDIDATES if candidates is None else candidates)
    best = {k: ngpu.GetKernelStatus(k) for k in cand}
    _time_steps(ngpu, max(steps // 4, 1))       # settle
    t_best = _time_steps(ngpu, steps)
    for _ in range(rounds):
        for k, values in cand.items():
            for v in values:
                if v == best[k]:
                    continue
                ngpu.SetKernelStatus(k, v)
                t = _time_steps(ngpu, steps)
                if log:
                    log("  autotune %s=%s: %.1f us/step" % (k, v, t * 1e6))
                if t < t_best:
                    t_best, best[k] = t, v
                else:
                    ngpu.SetKernelStatus(k, best[k])
    for k, v in best.items():
        ngpu.SetKernelStatus(k, v)
    out = dict(best)
    out["step_us"] = round(t_best * 1e6, 2)
    return out
