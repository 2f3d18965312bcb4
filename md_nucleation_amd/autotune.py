"""Launch-shape autotuner for the batched path.

Times short mdn_run_batched runs of each candidate shape on the engine's own
replicas and keeps the fastest (mdn_set_batch_shape).  The trial steps are
real steps: they advance every replica, so tuning doubles as warmup.

Candidates: split launches (the heaviest replicas in 512-thread workgroups
with a CU each, the rest in 128-thread workgroups that share CUs) or one
launch of 256- or 128-thread workgroups; LDS re-sizing every `chunk` steps.
An untimed run precedes the trials, so that none pays for the first
launches.
"""
from __future__ import annotations

import time

from . import _lib as M

DEFAULT_SHAPES = (
    (0, 0, 0, 1000),    # split, 1000-step launches
    (0, 0, 0, 250),     # split, 250-step launches
    (256, 0, 0, 1000),  # one launch, 256-thread workgroups
    (128, 0, 0, 1000),  # one launch, 128-thread workgroups
)


def _atom_steps(eng, n):
    return sum(eng.counters(r)[0] for r in range(n))


def tune_batched(eng, n_replicas, steps=100, precision=M.MIXED, shapes=DEFAULT_SHAPES, step0=0):
    """Returns (best_shape, {shape: atom-steps/s}); leaves the best applied.
    `step0` is the step index of the first step it runs (energy cadence); it
    runs (len(shapes) + 1) * steps steps."""
    results = {}
    s = step0
    # an untimed run first, so that no trial pays for the first launches
    eng.set_batch_shape(*shapes[0])
    eng.run_batched(steps, step0=s, precision=precision)
    eng.sync()
    s += steps
    for shape in shapes:
        eng.set_batch_shape(*shape)
        eng.sync()
        a0 = _atom_steps(eng, n_replicas)
        t0 = time.perf_counter()
        eng.run_batched(steps, step0=s, precision=precision)
        eng.sync()
        dt = time.perf_counter() - t0
        s += steps
        results[shape] = (_atom_steps(eng, n_replicas) - a0) / dt
    best = max(results, key=results.get)
    eng.set_batch_shape(*best)
    return best, results
