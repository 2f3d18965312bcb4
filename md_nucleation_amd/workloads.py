"""Synthetic workloads of the benchmark configurations, written as
reference-format input files (md_nucleation_amd.configs) and initialised by
libmdn's host front end (mdn_sysdef_init_state).

  batched_replicas  512 independent systems: ion drawn from {Na+, Cl-, the
                    150-atom AMBER chain}, 10-400 methanol or water vapor
                    molecules, 5,000 N2, 300 K, dt 1 fs
  single            one system from a condition file (e.g. Na+ + 20 MeOH +
                    10,000 N2)
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib as M
from . import configs


def batched_replicas(d, n_rep=512, n_gas=5000, seed=2024):
    """Returns [(SysDef, State, seed, kind)] for the batched-replica workload."""
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(seed)
    out = []
    cache = {}
    for r in range(n_rep):
        cond, kind = configs.config4_replica(d, r, rng, n_gas=n_gas)
        sd = M.SysDef(cond)
        st = sd.init_state(seed * 1000003 + r)
        out.append((sd, st, seed * 1000003 + r, kind))
        cache[kind] = cache.get(kind, 0) + 1
    return out


def cluster(d, n_vap=2000, n_gas=100000, seed=3):
    """Na+ with 2,000 methanol packed at 0.79 g/cm3 around it (12,001 atoms,
    all-pairs vv/vi) and 100,000 Ar: returns (SysDef, State, seed)."""
    os.makedirs(d, exist_ok=True)
    sd = M.SysDef(configs.config3(d, n_gas=n_gas, n_vap=n_vap))
    return sd, sd.init_state(seed), seed


def build_engine(reps, device=0):
    eng = M.Engine(device)
    for r, (sd, st, s, _) in enumerate(reps):
        eng.add(sd, st, seed=s, replica_id=r)
    eng.commit()
    # the constructor's first rebuild (initialization/md.cpp:34)
    for r in range(len(reps)):
        eng.make_pair(r)
    eng.sync()
    return eng
