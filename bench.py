"""Benchmark: in-domain atom-steps/s of the batched mixed-precision path.

Workload: 512 independent nucleation replicas in one persistent launch
(ion drawn from {Na+, Cl-, 150-atom AMBER chain}, 10-400 methanol or water
vapor molecules, 5,000 N2, 300 K, dt 1 fs), FP32 pair arithmetic with FP64
accumulation and integration.  Beside it, on the same workload:
  * the plain FP64 per-call GPU path (one replica after another, one launch
    per reference call, no graph), timed on a sample of replicas;
  * the CPU baseline: the FP64 oracle (a C restatement of the reference's
    loops, doing the same work as the GPU path, SHAKE included), one replica
    per OpenMP thread as the reference's main runs them, on the host cores
    given to this process, on a bounded sample; the reference's own stepping
    loop (built from its sources into oracle/_ref, where present) and the
    oracle without SHAKE on the same sample give the port/reference time
    ratio;
  * the force and energy error of the mixed path against the FP64 oracle on
    the timed launch (every replica) and on configs 1-3;
  * configs 1 and 2 (one large replica each): plain FP64 per-call GPU,
    batched and CPU-oracle throughput;
  * the 12,001-atom cluster and the differentiable vi term.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 200]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np
# torch (used by the differentiable-term section) brings its own HIP runtime;
# importing it before libmdn loads makes libmdn bind to that same runtime, so
# the process holds one HIP runtime rather than two
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s: batched mixed-prec vs plain FP64 GPU at ≤1e-5 rel force err"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per launch of the own-CU class (156 replicas with a CU each),
# from the FETCH_SIZE / WRITE_SIZE passes in
# profiles/k_batched_mixed_512replicas.md (copy-calibrated: FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0): 24.2 MB for a 200-step launch and the same for a 1-step
# one, i.e. the replicas' state read at the launch's start and written at its
# end; the steps themselves run out of LDS
HEAVY_HBM_BYTES_PER_LAUNCH = 24.2e6
# algorithmic FLOPs per unique pair / bonded term, as written in the reference
FLOP = {"ljcoul_w": 31, "ljcoul": 30, "lj": 24, "bond": 22, "angle": 70, "dihedral": 150}


def amber_long_count(sp):
    n = sp["mass"].size
    ex = set()
    for arr, w, a, b in ((sp["bond"], 3, 0, 1), (sp["rigid"], 3, 0, 1), (sp["angle"], 4, 0, 2), (sp["dih"], 5, 0, 3)):
        t = np.asarray(arr).reshape(-1, w)
        for row in t:
            i, j = int(row[a]), int(row[b])
            ex.add((min(i, j), max(i, j)))
    return n * (n - 1) // 2 - len(ex)


def step_flops(sd, st):
    """Algorithmic FLOPs of one step of one replica in state st."""
    from md_nucleation_amd import _lib as M
    v = sd.view
    fl = v["flags"]
    sp = v["species"]
    n_ion, na_g, na_v = sp[0]["mass"].size, sp[1]["mass"].size, sp[2]["mass"].size
    ng = int(st.region[1:1 + v["n_gas"]].sum())
    nv = int(st.region[1 + v["n_gas"]:].sum())
    f = 0
    if fl & M.MDN_VI:
        f += nv * na_v * n_ion * FLOP["ljcoul_w"]
    if fl & M.MDN_VV:
        f += nv * (nv - 1) // 2 * na_v * na_v * FLOP["ljcoul_w"]
    if fl & M.MDN_GI:
        f += st.c.n_gi * na_g * FLOP["lj"]
    if fl & M.MDN_GV:
        f += st.c.n_gv * na_g * min(na_g, na_v) * FLOP["lj"]
    if fl & M.MDN_ION_AMBER:
        f += amber_long_count(sp[0]) * FLOP["ljcoul"]
        f += (sp[0]["bond"].size // 3) * FLOP["bond"] + (sp[0]["angle"].size // 4) * FLOP["angle"]
        f += (sp[0]["dih"].size // 5) * FLOP["dihedral"]
    if fl & M.MDN_VAPOR_INTRA:
        f += nv * ((sp[2]["bond"].size // 3) * FLOP["bond"] + (sp[2]["angle"].size // 4) * FLOP["angle"]
                   + (sp[2]["dih"].size // 5) * FLOP["dihedral"])
    if fl & M.MDN_GAS_INTRA:
        f += ng * (sp[1]["bond"].size // 3) * FLOP["bond"]
    return f


def total_atom_steps(eng, n):
    return sum(eng.counters(r)[0] for r in range(n))


def _oracle_for(sd, seed, replica=0, shake=None):
    """The FP64 oracle on the same parsed system (the checker and the CPU
    baseline; never the measured path)."""
    from oracle import oracle as O
    v = sd.view
    sysd = {k: v[k] for k in ("pair_coeff", "btype", "ctype", "dmulti", "dcoeff", "species", "n_gas", "n_vap",
                              "ion_type_base", "dt", "T", "d_size", "margin", "ML2", "m_gas_kg", "m_vap_kg",
                              "gas_spin", "flags", "efield", "shake", "shake_max_iter")}
    sysd.update(seed=seed, replica=replica)
    if shake is not None:
        sysd["shake"] = shake
    return O.Oracle(sysd)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_threads():
    """The host threads this process may use: its CPU affinity, at most the
    16 cores a one-GPU share of the machine holds."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _oracle_sample(reps, k, shake=None):
    orcs, states = [], []
    for r in range(k):
        sd, st, seed, _ = reps[r]
        orc = _oracle_for(sd, seed, replica=r, shake=shake)
        S = orc.new_state(st.as_dict(), cap=20000)
        orc.make_pair(S)
        orcs.append(orc)
        states.append(S)
    return orcs, states


def _time_oracle(reps, k, seconds, shake=None, nsteps=None):
    from oracle import oracle as O
    orcs, states = _oracle_sample(reps, k, shake)
    if nsteps is None:
        n_cal = 50
        t0 = time.perf_counter()
        O.batch_verlet(orcs, states, n_cal, 0, k)
        dt_cal = time.perf_counter() - t0
        nsteps = int(max(50, min(20000, n_cal * seconds / max(dt_cal, 1e-6))))
    secs = []
    t0 = time.perf_counter()
    rc, acc = O.batch_verlet(orcs, states, nsteps, 0, k, seconds=secs)
    el = time.perf_counter() - t0
    # k cores kept busy (the 512-replica job over k cores): k x total
    # atom-steps / total replica-seconds, the measure of the reference loop
    return k * sum(acc) / sum(secs), nsteps, el


def cpu_baseline(reps, seconds=12.0):
    """The FP64 oracle on the host cores doing the GPU path's work (the same
    replicas, SHAKE as the GPU runs it, their own rebuilds): one replica per
    OpenMP thread, as the reference's main runs one replica per thread
    (main.cpp:21-29)."""
    k = cpu_threads()
    val, nsteps, el = _time_oracle(reps, k, seconds)
    info = cpu_info()
    return {"value": val, "unit": "atom-steps/s", "cores": k, "kind": "port",
            "cpu": info["model"], "nproc": info["nproc"], "affinity": info["affinity"],
            "sample": "the first %d replicas of the workload x %d steps (their rebuilds, SHAKE on as on the GPU), "
                      "FP64 oracle built with the reference's flags (-O3, OpenMP), one replica per thread, %.1f s "
                      "wall; value = %d x total atom-steps / total replica-seconds (cores kept busy, as for the "
                      "reference loop)" % (k, nsteps, el, k)}


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def cpu_reference_loop(d, n_rep, steps=20000):
    """The reference's own stepping loop (MD::verlet with its own rebuilds,
    built from the reference's sources by oracle/Makefile) on the first
    replicas of the workload: one single-threaded process per replica, as the
    reference's main runs one replica per OpenMP thread.  The reference runs
    no SHAKE (verlet.cpp:24).  Started before anything touches the GPU.
    None when the build is absent."""
    import subprocess
    if not os.path.exists(REF_HARNESS):
        return None
    k = min(cpu_threads(), n_rep)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    ps = [subprocess.Popen([REF_HARNESS, os.path.join(d, "c4_%d.cond" % r), "-", "time", str(steps)], cwd=d,
                           env=env, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for r in range(k)]
    acc, secs, wall = 0.0, 0.0, 0.0
    for p in ps:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        v = {ln.split()[0]: float(ln.split()[2]) for ln in out.strip().splitlines()}
        acc += v["atom_steps"]
        secs += v["seconds"]
        wall = max(wall, v["seconds"])
    return {"value": k * acc / secs, "unit": "atom-steps/s", "cores": k, "kind": "reference", "steps": steps,
            "sample": "the first %d replicas x %d steps each: the reference's own MD::verlet loop (its rebuilds, no "
                      "SHAKE), one single-threaded process per replica, %.1f s longest loop; value = %d x total "
                      "atom-steps / total process-seconds" % (k, steps, wall, k)}


def port_vs_reference(reps, ref):
    """The oracle with SHAKE off (the reference loop's work) on the sample the
    reference loop ran, for the port / reference time ratio on these cores."""
    k = ref["cores"]
    val, nsteps, el = _time_oracle(reps, k, 0.0, shake=0, nsteps=ref["steps"])
    return {"value": val, "unit": "atom-steps/s", "cores": k, "steps": nsteps,
            "port_over_reference_time": ref["value"] / val}


def _in_mask(sd, region):
    v = sd.view
    na = [sp["mass"].size for sp in v["species"]]
    return np.repeat(np.asarray(region, bool), [na[0]] + [na[1]] * v["n_gas"] + [na[2]] * v["n_vap"])


class ErrAcc:
    """Force and per-label energy error of the mixed path against the FP64
    oracle over one or more replicas: relative RMS over in-domain atoms, max
    abs, and per label the largest |dU| with the largest |U|."""

    def __init__(self):
        self.num = self.den = self.fmax = 0.0
        self.du = np.zeros(8)
        self.uref = np.zeros(8)
        self.n = 0

    def add(self, sd, st, S):
        m = _in_mask(sd, S.region)
        fa, fb = st.atom_f.reshape(-1, 3)[m], S.atom_f.reshape(-1, 3)[m]
        self.num += float(((fa - fb) ** 2).sum())
        self.den += float((fb ** 2).sum())
        self.fmax = max(self.fmax, float(np.abs(fa - fb).max())) if fa.size else self.fmax
        self.du = np.maximum(self.du, np.abs(st.U - S.U))
        self.uref = np.maximum(self.uref, np.abs(S.U))
        self.n += 1

    def result(self, what):
        from md_nucleation_amd import _lib as M
        return {"rel_rms": (self.num / max(self.den, 1e-300)) ** 0.5, "max_abs": self.fmax, "unit": "kcal/mol/A",
                "energy_max_abs_err": {lab: float(self.du[k]) for k, lab in enumerate(M.U_LABELS)},
                "energy_max_abs": {lab: float(self.uref[k]) for k, lab in enumerate(M.U_LABELS)},
                "replicas": self.n, "vs": what}


def _class_shapes(cfg, plan):
    """the chosen launch shape per class of the split launch"""
    if plan["uniform"]:
        return {"uniform": {"replicas": plan["uniform"], "threads": cfg["threads"]}}
    return {"own_cu": {"replicas": plan["own_cu"], "threads": cfg["own_threads"]},
            "three_per_cu": {"replicas": plan["three_per_cu"], "threads": cfg["shared_threads"]},
            "four_per_cu": {"replicas": plan["four_per_cu"], "threads": cfg["shared_threads"]},
            "dense_run": "per replica (in the kernel)", "heavy_factor": cfg["heavy_factor"],
            "chunk": cfg["chunk"]}


def batched_force_error(reps, cfg=None):
    """One mixed step of the timed launch shape (all replicas, split launch,
    the tuned configuration) against the FP64 oracle from the same states:
    both hold F(x1) for the same x1 (x1 depends only on the starting state),
    energies on that step."""
    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import autotune
    from md_nucleation_amd import workloads as W
    from oracle import oracle as O
    eng = W.build_engine(reps)
    if cfg:
        autotune.apply(eng, cfg)
    eng.run_batched(1, eflag_every=1, precision=M.MIXED)
    eng.sync()
    plan, _ = eng.batch_plan()
    orcs, states = _oracle_sample(reps, len(reps))
    O.batch_verlet(orcs, states, 1, 1, cpu_threads())
    acc = ErrAcc()
    for r, (sd, _, _, _) in enumerate(reps):
        acc.add(sd, eng.get_state(r), states[r])
    eng.close()
    out = acc.result("FP64 oracle, one step from identical states, every replica, the split launch")
    out["launch"] = plan
    return out


def single_configs(d, with_cpu=True):
    """Configs 1 and 2 (one replica each): the mixed force / energy error
    against the FP64 oracle, and throughput of the plain FP64 per-call GPU
    path, the batched path (mixed and FP64) and the FP64 oracle on one
    thread.  Config 3 (the cluster) adds its error here; its throughput is
    in cluster_bench."""
    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import configs
    cases = (("config1", "Na+ + 20 methanol + 10,000 N2 (SHAKE H)", lambda dd: configs.config1(dd), 2000, 2000, 300),
             ("config2", "150-atom AMBER chain + 200 water + 100,000 N2 (SHAKE H)", lambda dd: configs.config2(dd),
              1000, 1000, 100),
             ("config3", "Na+ + 2,000 methanol cluster (12,001 atoms, all-pairs) + 100,000 Ar",
              lambda dd: configs.config3(dd), 0, 0, 0))
    out = {}
    for name, what, fn, n_pc, n_b, n_cpu in cases:
        dd = os.path.join(d, name)
        os.makedirs(dd, exist_ok=True)
        sd = M.SysDef(fn(dd))
        seed = 17
        st0 = sd.init_state(seed)
        if name == "config3":
            st0.atom_w[:] = 1.0
        res = {"workload": what}
        # error: one mixed step against the oracle
        e = M.Engine(0)
        e.add(sd, st0, seed=seed)
        e.commit()
        e.make_pair(0)
        if name == "config3":
            e.verlet(0, 1, eflag_every=1, precision=M.MIXED)
        else:
            e.run_batched(1, eflag_every=1, precision=M.MIXED)
        e.sync()
        orc = _oracle_for(sd, seed)
        S = orc.new_state(st0.as_dict())
        orc.make_pair(S)
        orc.verlet(S, 1, eflag_every=1)
        acc = ErrAcc()
        acc.add(sd, e.get_state(0), S)
        res["force_err"] = acc.result("FP64 oracle, one step from an identical state")
        e.close()
        if n_pc:
            for label, path, prec, n in (("plain_fp64_gpu", "percall", M.FP64, n_pc),
                                         ("batched_fp64", "batched", M.FP64, n_b),
                                         ("batched_mixed", "batched", M.MIXED, n_b)):
                e = M.Engine(0)
                e.add(sd, st0, seed=seed)
                e.commit()
                e.make_pair(0)
                run = (lambda k, p=prec: e.verlet(0, k, precision=p)) if path == "percall" else \
                      (lambda k, p=prec: e.run_batched(k, precision=p))
                run(20)
                e.sync()
                a0 = e.counters(0)[0]
                t0 = time.perf_counter()
                run(n)
                e.sync()
                el = time.perf_counter() - t0
                res[label] = {"value": (e.counters(0)[0] - a0) / el, "unit": "atom-steps/s",
                              "ms_per_step": el * 1e3 / n, "steps": n}
                e.close()
            if with_cpu:
                S = orc.new_state(st0.as_dict())
                orc.make_pair(S)
                t0 = time.perf_counter()
                rc, acc_steps = orc.verlet(S, n_cpu)
                el = time.perf_counter() - t0
                res["cpu_oracle"] = {"value": acc_steps / el, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                                     "sample": "%d steps, FP64 oracle, one thread (one replica)" % n_cpu}
        out[name] = res
        print("%s done" % name, file=sys.stderr, flush=True)
    return out


def copy_check(nbytes):
    """A plain device-to-device copy of known size (the reference point for
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters), through the same HIP
    runtime as libmdn."""
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    print("copy_check: %d bytes device-to-device" % nbytes, file=sys.stderr, flush=True)


def cluster_bench(d, steps=40, cpu_steps=3, with_cpu=True):
    """The 12,001-atom cluster (Na+ + 2,000 methanol, 100,000 Ar) through the
    per-call path: mixed and FP64 throughput, and the all-pairs vv kernel
    alone against the FP32 VALU peak."""
    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import workloads as W
    sd, st0, seed = W.cluster(d)
    st0.atom_w[:] = 1.0  # all atoms within 100 A of the ion
    out = {"workload": "cluster: Na+ + 2000 methanol at 0.79 g/cm3 (12,001 atoms, all-pairs vv/vi) + 100,000 Ar"}
    for name, prec in (("mixed", M.MIXED), ("fp64", M.FP64)):
        eng = M.Engine(0)
        eng.add(sd, st0, seed=seed)
        eng.commit()
        eng.make_pair(0)
        eng.verlet(0, 2, precision=prec)
        eng.sync()
        a0 = eng.counters(0)[0]
        t0 = time.perf_counter()
        eng.verlet(0, steps, precision=prec)
        eng.sync()
        el = time.perf_counter() - t0
        out[name] = {"value": (eng.counters(0)[0] - a0) / el, "unit": "atom-steps/s", "ms_per_step": el * 1e3 / steps}
        # the vv kernel alone (one launch, 20 repeats)
        st = eng.get_state(0)
        nv = int(st.region[1 + sd.view["n_gas"]:].sum())
        na = sd.view["species"][2]["mass"].size
        pairs = nv * (nv - 1) // 2 * na * na
        ms = []
        for _ in range(20):
            eng.compute(0, M.MDN_VV, 0, prec)
            ms.append(eng.last_kernel_ms())
        t = float(np.median(ms)) * 1e-3
        peak = PEAK_FP32_TFLOPS if prec == M.MIXED else PEAK_FP64_TFLOPS
        ach = pairs * FLOP["ljcoul_w"] / t / 1e12
        out[name]["vv_kernel"] = {"kernel": "k_vv_n3<%s>" % ("float" if prec == M.MIXED else "double"),
                                  "unique_pairs": pairs, "ms": t * 1e3, "achieved_tflops": ach, "peak_tflops": peak,
                                  "frac": ach / peak}
        eng.close()
    if with_cpu:
        orc = _oracle_for(sd, seed)
        S = orc.new_state(st0.as_dict())
        orc.make_pair(S)
        t0 = time.perf_counter()
        rc, acc = orc.verlet(S, cpu_steps)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": acc / el, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                               "sample": "%d steps of the cluster, FP64 oracle, one thread (one replica)" % cpu_steps}
        out["mixed_vs_cpu"] = out["mixed"]["value"] / out["cpu_baseline"]["value"]
    out["mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def adjoint_bench(d, n_snap=256, repeats=5):
    """The differentiable vi term on the force-matching workload: 256 FP64
    frames of the 150-atom chain with 200 water packed around it; forward
    (E, F) and adjoint (gradients of a force-matching loss) wall time per
    call, host buffers in and out included."""
    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import adjoint as A
    from md_nucleation_amd import configs
    sd = M.SysDef(configs.config5(d, n_gas=5000))
    snaps = A.snapshots(sd, sd.init_state(11), 11, n=n_snap, every=10)
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    rng = np.random.default_rng(5)
    Ft = A.vi_forward(snaps, *[a * (1 + rng.uniform(-0.1, 0.1, a.size)) for a in (sig, eps, q)])[1]
    tf, ta = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, F = A.vi_forward(snaps, sig, eps, q)
        t1 = time.perf_counter()
        A.vi_adjoint(snaps, sig, eps, q, gF=2 * (F - Ft))
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        ta.append(t2 - t1)
    pairs = int(snaps.vap_in.sum()) * snaps.n_ion
    f, a = float(np.median(tf)), float(np.median(ta))
    return {"workload": "%d FP64 frames: 150-atom AMBER chain + 200 packed water (750 atoms)" % n_snap,
            "pairs_per_pass": pairs, "forward_ms": f * 1e3, "adjoint_ms": a * 1e3,
            "forward_pairs_per_s": pairs / f, "adjoint_pairs_per_s": pairs / a,
            "note": "wall time per C-ABI call including host-buffer copies",
            "segment": segment_bench(d)}


def segment_bench(d, nsteps=100, every=10):
    """Gradient of a 100-step trajectory loss (checkpoints every 10 steps)
    with respect to the ion's 450 vi parameters: wall time of one reverse
    pass (forward with checkpoints, recomputation, Hessian-vector products,
    adjoint-kernel calls)."""
    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import adjoint as A
    from md_nucleation_amd import configs
    from md_nucleation_amd import segment as SG
    sd = M.SysDef(configs.segment_condition(d, n_gas=200, seed=1))
    eng = M.Engine(0)
    eng.add(sd, sd.init_state(7), seed=7)
    eng.commit()
    eng.make_pair(0)
    eng.verlet(0, 100, 0, 0, M.FP64)
    st = eng.get_state(0)
    eng.close()
    seg = SG.Segment(sd, st)
    try:
        x0, v0 = seg.state(st)
        theta = tuple(t.numpy() for t in A.ion_params(sd))
        rng = np.random.default_rng(5)
        pert = tuple(a * (1 + rng.uniform(-0.1, 0.1, a.size)) for a in theta)
        xs, _, _ = seg.run(pert, x0, v0, nsteps)
        obs = set(range(every, nsteps + 1, every))
        tgt = {n: xs[n] for n in obs}
        # the device path (mdn_segment_grad): one untimed call, then the
        # median of three
        SG.loss_and_grad(seg, theta, x0, v0, nsteps, tgt, obs, every=every)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            SG.loss_and_grad(seg, theta, x0, v0, nsteps, tgt, obs, every=every)
            ts.append(time.perf_counter() - t0)
        dt = float(np.median(ts))
        # the same reverse pass over host arrays (a device call per force
        # evaluation and Hessian-vector product)
        seg.n_force_calls = 0
        seg.n_hvp_calls = 0
        t0 = time.perf_counter()
        SG.loss_and_grad_host(seg, theta, x0, v0, nsteps, tgt, obs, every=every)
        dt_host = time.perf_counter() - t0
        nf, nh = seg.n_force_calls, seg.n_hvp_calls
        # one exact Hessian-vector product alone (host copies included)
        u = rng.normal(0, 1, x0.shape)
        t1 = time.perf_counter()
        for _ in range(10):
            seg.hvp(theta, x0, u)
        hvp_ms = (time.perf_counter() - t1) * 1e2
        return {"workload": "%d-step segment of the chain + 200 water (750 atoms), checkpoints every %d" % (nsteps, every),
                "gradient_s": dt, "path": "mdn_segment_grad: state, checkpoints and reverse pass in device memory",
                "host_arrays_gradient_s": dt_host, "force_evaluations": nf, "hvp_calls": nh,
                "hvp": "exact: mdn_vi_jvp (closed form) + mdn_force_jvp (dual-number force kernels), FP64",
                "ms_per_hvp": hvp_ms}
    finally:
        seg.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-replicas", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--tune-steps", type=int, default=1000, help="steps per launch-shape trial (on a copy)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip configs 1-3 (errors, single-replica rates)")
    ap.add_argument("--no-cluster", action="store_true", help="skip the 12,001-atom cluster measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable vi term measurement")
    ap.add_argument("--copy-check", action="store_true",
                    help="also run a 256 MiB device-to-device copy (calibrates HBM byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only one GPU is measured; replicas would shard across GPUs as whole replicas "
              "with no exchange between them, but no multi-GPU run is implemented", file=sys.stderr)
        sys.exit(2)

    from md_nucleation_amd import _lib as M
    from md_nucleation_amd import workloads as W

    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        reps = W.batched_replicas(d, n_rep=args.replicas)
        # the reference's own loop runs in child processes started before
        # anything touches the GPU
        cpu_ref = None if args.no_cpu else cpu_reference_loop(d, len(reps))
        n = len(reps)
        # launch-shape autotuning on a copy of the workload: the timed engine
        # itself takes exactly the warmup steps before the timed region
        from md_nucleation_amd import autotune
        tuned = None
        cfg = dict(autotune.DEFAULT)
        if args.tune_steps > 0:
            et = W.build_engine(reps)
            tt = time.perf_counter()
            cfg, trials, tplan = autotune.tune_batched(et, n, steps=args.tune_steps, precision=M.MIXED)
            tuned = {"config": cfg, "classes": _class_shapes(cfg, tplan),
                     "trials": [[{k: v for k, v in c.items() if c[k] != autotune.DEFAULT[k]}, round(r / 1e9, 4)]
                                for c, r in trials],
                     "trial_unit": "1e9 atom-steps per kernel-second (changes from the default shown)",
                     "steps_per_trial": args.tune_steps, "seconds": time.perf_counter() - tt,
                     "on": "a copy of the workload"}
            et.close()
            print("autotune: %s" % cfg, file=sys.stderr, flush=True)
        eng = W.build_engine(reps)
        autotune.apply(eng, cfg)
        print("workload ready: %d replicas in %.1f s" % (n, time.perf_counter() - t0), file=sys.stderr, flush=True)
        if args.warmup:
            eng.run_batched(args.warmup, eflag_every=0, precision=M.MIXED)
        eng.sync()
        a0 = total_atom_steps(eng, n)
        t0 = time.perf_counter()
        eng.run_batched(args.steps, eflag_every=0, step0=args.warmup, precision=M.MIXED)
        eng.sync()
        el = time.perf_counter() - t0
        kern_ms = eng.last_kernel_ms()
        chunk = cfg["chunk"] if cfg["chunk"] > 0 else args.steps
        steps_per_launch = min(args.steps, chunk)
        lds_cap, lds_bytes, n_heavy = eng.batch_info()
        plan, _ = eng.batch_plan()
        a1 = total_atom_steps(eng, n)
        value = (a1 - a0) / el
        flops = 0
        for r in range(n):
            flops += step_flops(reps[r][0], eng.get_state(r))
        achieved = flops * args.steps / (kern_ms * 1e-3) / 1e12
        print("batched: %.3f s for %d steps, kernel %.3f ms" % (el, args.steps, kern_ms), file=sys.stderr, flush=True)
        eng.close()

        # mixed-precision force / energy error of the timed launch shape
        # against the FP64 oracle, every replica
        force_err = batched_force_error(reps, cfg)
        print("force error done", file=sys.stderr, flush=True)

        # plain FP64 per-call path on a sample of the same replicas
        pr = min(args.plain_replicas, n)
        ep = W.build_engine(reps[:pr])
        ep.verlet(0, 20, precision=M.FP64)
        ep.sync()
        p0 = total_atom_steps(ep, pr)
        t0 = time.perf_counter()
        for r in range(pr):
            ep.verlet(r, args.plain_steps, precision=M.FP64)
        ep.sync()
        pel = time.perf_counter() - t0
        plain = (total_atom_steps(ep, pr) - p0) / pel
        ep.close()

        cpu = None if args.no_cpu else cpu_baseline(reps, args.cpu_seconds)
        calib = port_vs_reference(reps, cpu_ref) if cpu_ref else None
        cfgs = None if args.no_configs else single_configs(os.path.join(d, "cfg"), with_cpu=not args.no_cpu)
        cl = None if args.no_cluster else cluster_bench(os.path.join(d, "cluster"), with_cpu=not args.no_cpu)
        adj = None if args.no_adjoint else adjoint_bench(os.path.join(d, "fm"))
        if args.copy_check:
            copy_check(256 << 20)

    n_in = (a1 - a0) / args.steps / n
    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 pair math, fp64 accumulation/integration", "data": "synthetic",
        "config": {"workload": "batched replicas: %d x (Na+ | Cl- | 150-atom AMBER chain) + 10-400 methanol|water "
                               "+ 5000 N2, 300 K, dt 1 fs, SHAKE H, gi/gv LJ, vi/vv LJcoul" % n,
                   "replicas": n, "mean_in_domain_atoms": n_in},
        "replica_steps_per_s": n * args.steps / el,
        "plain_fp64_gpu": {"value": plain, "unit": "atom-steps/s",
                           "sample": "%d replicas x %d steps, per-call kernels, one replica after another" %
                                     (pr, args.plain_steps)},
        "batched_vs_plain": value / plain,
        "force_err": force_err,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": HEAVY_HBM_BYTES_PER_LAUNCH,
                     "traffic_note": "HBM bytes per launch of the own-CU class (here %d steps; the same for 1 "
                                     "or 200 steps: the state in and out at the launch edges), counter passes "
                                     "in profiles/k_batched_mixed_512replicas.md" % steps_per_launch,
                     "kernel": "k_batched<float, false>", "kernel_ms": kern_ms, "lds_bytes_per_wg": lds_bytes,
                     "replicas_with_own_cu": n_heavy, "lds_cap_atoms": lds_cap, "launch_plan": plan,
                     "flops_per_step": flops},
        "cpu_baseline": cpu,
        "cpu_reference_loop": cpu_ref,
        "port_vs_reference": calib,
        "autotune": tuned,
        "configs": cfgs,
        "cluster": cl,
        "adjoint": adj,
    }
    if cpu:
        out["vs_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
