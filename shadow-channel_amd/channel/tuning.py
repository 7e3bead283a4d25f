"""Launch-shape autotuner.

The library exposes a few launch shapes that do not change results (bit
for bit) but do change speed with the grid and the number of members
(sc_set_tuning): how many members' band solves share one block and its
staged factor image, and how many blocks the parity y-transforms spread
over.  autotune() times a short forward+adjoint window of the actual
configuration for each candidate, one shape at a time (the others held at
their best so far), keeps the fastest and leaves it set.  Results can be
kept in a JSON file keyed by grid, member count and precision, so that a
later process applies them without timing again.
"""
import json
import os
import time

import numpy as np

import c_channel

# the autotuner's choices for the configurations bench.py measures, timed on
# an MI355X (a cache in the format autotune() writes)
SHAPES_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shapes.json")

SHAPES = {
    "solve_members_per_block": (1, 2, 4),
    "rebuild_blocks": (32, 64, 128),
    "force_blocks": (64, 128, 256),
}


def _key(grid, members, mixed):
    return "%dx%dx%d/m%d/%s" % (grid[0], grid[1], grid[2], members, "mixed" if mixed else "fp64")


def _window(states, steps):
    """Seconds for one fwd+adj window of all members (states on the device)."""
    Cd = c_channel.DeviceArray(states)
    ACd = c_channel.DeviceArray(np.zeros_like(states))
    try:
        c_channel.sc_synchronize()
        t0 = time.perf_counter()
        c_channel.sc_primal_members(0, Cd)
        c_channel.sc_adjoint_members(steps, 0, ACd, 1.0, False)
        c_channel.sc_synchronize()
        return time.perf_counter() - t0
    finally:
        Cd.free()
        ACd.free()


def autotune(grid, box, Re, flux, dt, states, mixed=False, steps=3, reps=2, cache=None,
             shapes=None):
    """Choose the launch shapes for `states` (members x coefficient array) on
    the grid (Nx, Ny, Nz) and box (Lx, Lz); returns {name: value} and leaves
    it set.  cache: a JSON file read first and updated after timing."""
    states = np.ascontiguousarray(states, dtype=np.complex128)
    members = states.shape[0]
    key = _key(grid, members, mixed)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if key in known:
            for name, value in known[key].items():
                c_channel.sc_set_tuning(name, value)
            return dict(known[key])
    shapes = dict(SHAPES if shapes is None else shapes)
    c_channel.sc_set_members(members)
    c_channel.sc_set_precision(1 if mixed else 0)
    best = {name: 0 for name in shapes}
    timings = {}
    try:
        assert c_channel.c_init(*grid, *box, Re, flux, dt, steps) == 0
        for name in best:
            c_channel.sc_set_tuning(name, 0)
        _window(states, steps)                       # code and caches warm
        for name, values in shapes.items():
            choice, fastest = 0, None
            for value in values:
                if name == "solve_members_per_block" and value > members:
                    continue
                c_channel.sc_set_tuning(name, value)
                t = min(_window(states, steps) for _ in range(reps))
                timings["%s=%d" % (name, value)] = t
                if fastest is None or t < fastest:
                    choice, fastest = value, t
            c_channel.sc_set_tuning(name, choice)
            best[name] = choice
        c_channel.c_destroy()
    finally:
        c_channel.sc_set_members(1)
        c_channel.sc_set_precision(0)
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[key] = best
        with open(cache, "w") as f:
            json.dump(known, f, indent=1, sort_keys=True)
    autotune.last_timings = timings
    return best


autotune.last_timings = {}


def reset():
    """Back to the built-in launch shapes."""
    for name in SHAPES:
        c_channel.sc_set_tuning(name, 0)
