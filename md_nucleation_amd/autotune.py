"""Launch-shape autotuner for the batched path.

Times runs of `steps` real steps (they advance every replica, so tuning
doubles as warmup) of candidate launch shapes on the engine's own replicas,
rated by in-domain atom-steps per second of kernel time (the library's HIP
events around its launches), and keeps the fastest.

The knobs, searched by coordinate descent (each knob in turn over its
values, the others held at the best so far; two passes):
  * chunk: steps per launch (between launches the replicas are re-planned
    and the LDS capacity re-sized): 1000, 500, 250;
  * own_threads: workgroup size of the replicas with a CU of their own: 512,
    256;
  * shared_threads: workgroup size of the replicas that share CUs: 128, 64,
    256;
  * heavy_factor: the cost, as a multiple of the median, above which a
    replica may get a CU of its own: 2, 1.5, 3.
When there are no more replicas than CUs the split launch does not engage;
the workgroup size of the one launch (256, 128, 512, 64) then takes the
place of the three class knobs.  The dense pass's run length is not a knob:
it is chosen per replica from its block counts inside the kernel (a
run-time override cost the kernel 20-270 spilled registers; forced run
lengths of 8, 16 and 32 measured 3-16 % slower than the per-replica
choice).  A candidate replaces the incumbent only
when it is faster by more than `margin` (relative), so that run-to-run noise
does not steer the search.
"""
from __future__ import annotations

from . import _lib as M

SPLIT_KNOBS = (
    ("chunk", (1000, 500, 250)),
    ("own_threads", (512, 256)),
    ("shared_threads", (128, 64, 256)),
    ("heavy_factor", (2.0, 1.5, 3.0)),
)
UNIFORM_KNOBS = (
    ("chunk", (1000, 500, 250)),
    ("threads", (256, 128, 512, 64)),
)
DEFAULT = {"threads": 0, "chunk": 1000, "own_threads": 512, "shared_threads": 128, "heavy_factor": 2.0}


def apply(eng, cfg):
    """sets a configuration (a dict with DEFAULT's keys) on the engine"""
    eng.set_batch_shape(cfg["threads"], 0, 0, cfg["chunk"])
    eng.set_batch_tuning(cfg["own_threads"], cfg["shared_threads"], cfg["heavy_factor"])


def _atom_steps(eng, n):
    return sum(eng.counters(r)[0] for r in range(n))


def tune_batched(eng, n_replicas, steps=200, precision=M.MIXED, step0=0, passes=2, margin=0.01):
    """Returns (best configuration, trials, plan): trials lists (configuration,
    atom-steps per kernel-second) in the order run, plan the class counts
    of the best configuration's last launch.  Leaves the best applied."""
    s = step0
    trials = []

    def measure(cfg):
        nonlocal s
        apply(eng, cfg)
        eng.sync()
        a0 = _atom_steps(eng, n_replicas)
        eng.run_batched(steps, step0=s, precision=precision)
        eng.sync()
        s += steps
        rate = (_atom_steps(eng, n_replicas) - a0) / (eng.last_kernel_ms() * 1e-3)
        trials.append((dict(cfg), rate))
        return rate

    cur = dict(DEFAULT)
    # an untimed run first, so that no trial pays for the first launches
    apply(eng, cur)
    eng.run_batched(steps, step0=s, precision=precision)
    eng.sync()
    s += steps
    plan, _ = eng.batch_plan()
    knobs = UNIFORM_KNOBS if plan["uniform"] > 0 else SPLIT_KNOBS
    if plan["uniform"] > 0:
        cur["threads"] = 256
    best = measure(cur)
    for _ in range(passes):
        for knob, values in knobs:
            for v in values:
                if v == cur[knob]:
                    continue
                cand = dict(cur)
                cand[knob] = v
                r = measure(cand)
                if r > best * (1.0 + margin):
                    cur, best = cand, r
    apply(eng, cur)
    eng.run_batched(steps, step0=s, precision=precision)
    eng.sync()
    plan, _ = eng.batch_plan()
    return cur, trials, plan
