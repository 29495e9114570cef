"""Launch-shape autotuner.

Picks the kernel launch settings of the per-step path by timing short
stretches of the running simulation.  The settings are:
* the first-pass RK5 attempt budget and the second-pass grid;
* the delivery mode (one float atomic per synaptic event, events binned
  by accumulator range and summed in LDS first, or tiled: one accumulator
  tile per block from a tile-major copy of the store), the delivery grid
  and the largest number of (spike, delay) pairs per wave tile;
* whether the delivery shares the second RK5 pass's launch, and how many
  steps in a row the second pass may overlap the next step;
* whether Simulate replays captured hipGraphs of 16 steps or launches
  every step from the host.

Each alternative is timed against the incumbent in interleaved windows and
kept only if its median is clearly faster, so that window-to-window noise
does not move the choice.

None of these changes results: each neuron's arithmetic and each synapse's
delivery are the same under every setting. The tests check this spike for
spike. The tuner advances the simulation while it measures, so call it
during warm-up.
"""
import statistics
import time

CANDIDATES = {
    "rk5_budget": [1, 2],
    "rk5_pass2_grid": [256, 64, 1024],
    "delivery_mode": [2, 0, 1, 3],
    "delivery_grid": [1024, 2048, 4096],
    "delivery_tile_max": [8, 16, 64],
    "fuse_delivery": [1, 0],
    "step_overlap": [8, 0, 2, 4, 16],
    "graph_steps": [0, 16],
}


def _time_steps(ngpu, steps):
    dt = ngpu.GetTimeResolution()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * dt)   # returns after the stream has drained
    return (time.perf_counter() - t0) / steps


def _window(ngpu, name, value, steps):
    """Time `steps` steps with setting `name` = value, after a short
    untimed stretch that absorbs one-time costs of the switch (a graph
    capture, the tile-major store copy)."""
    ngpu.SetKernelStatus(name, value)
    _time_steps(ngpu, max(steps // 4, 1))
    return _time_steps(ngpu, steps)


def choose(times_a, times_b, margin):
    """True when the candidate's median window (times_b) beats the
    incumbent's (times_a) by more than `margin` (relative)."""
    return statistics.median(times_b) < statistics.median(times_a) * (
        1.0 - margin)


def autotune(ngpu, steps=40, rounds=1, candidates=None, log=None, reps=3,
             margin=0.02):
    """Coordinate descent over `candidates` (default CANDIDATES): for each
    setting in turn, each other value is timed against the current one in
    `reps` interleaved pairs of `steps`-step windows, and replaces it only
    if its median is more than `margin` faster (single windows differ by
    a few per cent from one to the next).  Returns {setting: chosen value,
    "step_us": median time per step of the chosen settings}."""
    cand = dict(CANDIDATES if candidates is None else candidates)
    best = {k: ngpu.GetKernelStatus(k) for k in cand}
    _time_steps(ngpu, max(steps // 4, 1))       # settle
    for _ in range(rounds):
        for k, values in cand.items():
            for v in values:
                if v == best[k]:
                    continue
                ta, tb = [], []
                for _ in range(reps):
                    ta.append(_window(ngpu, k, best[k], steps))
                    tb.append(_window(ngpu, k, v, steps))
                if log:
                    log("  autotune %s: %s %.1f vs %s %.1f us/step (median "
                        "of %d)" % (k, best[k],
                                    statistics.median(ta) * 1e6, v,
                                    statistics.median(tb) * 1e6, reps))
                if choose(ta, tb, margin):
                    best[k] = v
                ngpu.SetKernelStatus(k, best[k])
    for k, v in best.items():
        ngpu.SetKernelStatus(k, v)
    t = [_time_steps(ngpu, steps) for _ in range(reps)]
    out = dict(best)
    out["step_us"] = round(statistics.median(t) * 1e6, 2)
    return out
