"""Autotuner of the launch-shape parameters (mr3_set_tuning).

The solver's results do not depend on these parameters, only its speed does: how many
streams carry the singleton phases, when the latency kernel takes over the last
singleton iterations, the lane budget of the multisection cluster refinements, the
level size from which dlarrf runs on one lane per cluster, the most waves of one
singleton launch, and the order from which the small route's matrices start with the
solve rather than after the large problems' root refinement.  `autotune` times a
prepared batch under candidate values, one parameter at a time (coordinate descent,
repeated while it improves), keeps the fastest, and leaves it applied.  The tuned values are plain JSON (`save` / `load` /
`apply`), so a search made once on a device and workload can be reused.

No counterpart in the reference, whose only knob is PMR_NUM_THREADS.
"""
import json
import time

import mr3smp_amd as M

# candidate values per parameter, searched in this order
SPACE = {
    "side_streams": [2, 3, 1],
    "tail_singletons": [8192, 4096, 16384, 32768],
    "group_lanes": [262144, 131072, 524288],
    "rrf_group_max": [65536, 16384, 262144],
    "singleton_waves": [8192, 4096, 16384],
    "small_first": [0, 350, 250],
}


def time_batch(arr, stream, reps=2):
    """Best wall time (s) of `reps` solves of a prepared batch (after one warm-up)."""
    import torch
    assert M.run_batch(arr, stream=stream) == 0
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = M.run_batch(arr, stream=stream)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
        assert rc == 0
    return best


def autotune(arr, stream=None, space=None, reps=2, rounds=2, log=None):
    """Coordinate descent over `space` (default SPACE) on the prepared batch `arr`.
    Returns (best tuning dict, its time in s, the list of (tuning, time) tried)."""
    space = space or SPACE
    best = M.get_tuning()
    M.set_tuning(**best)
    best_t = time_batch(arr, stream, reps)
    tried = [(dict(best), best_t)]
    for _ in range(rounds):
        improved = False
        for key, values in space.items():
            for v in values:
                if v == best[key]:
                    continue
                cand = dict(best, **{key: v})
                M.set_tuning(**cand)
                t = time_batch(arr, stream, reps)
                tried.append((cand, t))
                if log:
                    log(f"tune {key}={v}: {t * 1e3:.1f} ms (best {best_t * 1e3:.1f})")
                if t < best_t * 0.99:
                    best, best_t, improved = cand, t, True
            M.set_tuning(**best)
        if not improved:
            break
    M.set_tuning(**best)
    return best, best_t, tried


def save(tuning, path, note=""):
    with open(path, "w") as f:
        json.dump({"tuning": tuning, "note": note}, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)["tuning"]


def apply(path):
    """Applies a saved tuning; returns it."""
    return M.set_tuning(**load(path))
