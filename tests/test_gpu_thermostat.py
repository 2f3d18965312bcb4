"""The relaxation-phase ion thermostat of MD::verlet (nvt.cpp, verlet.cpp:22,
27, 36) on the GPU.

  golden     the reference's own 30-step trajectories of the 40-atom AMBER
             chain + 20 water + 200 N2 under "NVTion scale" and "NVTion 250"
             (tests/golden/nvt.npz): per-call, graph and batched FP64 paths
             within 1e-12 relative, zeta_ion within 1e-12; batched mixed
             precision within 1e-6 relative of the same trajectory
  rebuilds   200 FP64 steps with pair-list rebuilds and SHAKE against the
             oracle under both modes
  setNVE     mdn_set_thermostat(MDN_NVE) mid-run (MD::setNVE, run.cpp:56)
             against the oracle switching modes at the same step, and
             mdn_set_state of one replica leaving the others' state alone
"""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from md_nucleation_amd import configs
from tests import engine_util as E
from tests import golden_util as G
from tests.golden.cases import write_case

pytestmark = pytest.mark.gpu

NVT = ["nvt_scale", "nvt_nh"]


def _engine(tmp_path, name, fx, n_rep=1):
    sd = M.SysDef(write_case(name, str(tmp_path)))
    sd.set_shake(0)  # MD::verlet as shipped: no SHAKE (verlet.cpp:24)
    arrays = G.state(fx, "init_")
    arrays["gi"] = fx["s_list_gi"].astype(np.int32)
    arrays["gv"] = fx["s_list_gv"].astype(np.int32)
    arrays.update(loopPair=0, loop_update=10 ** 8, zeta_ion=0.0)
    eng = M.Engine(0)
    for r in range(n_rep):
        eng.add(sd, E.state_from(sd, arrays), seed=7, replica_id=r)
    eng.commit()
    return eng, sd


@pytest.mark.parametrize("name", NVT)
@pytest.mark.parametrize("path,precision", [("percall", M.FP64), ("graph", M.FP64), ("batched", M.FP64),
                                            ("batched", M.MIXED)])
def test_thermostat_vs_reference(tmp_path, name, path, precision):
    fx = G.sub(G.load("nvt"), name + "_")
    eng, sd = _engine(tmp_path, name, fx, n_rep=2 if path == "batched" else 1)
    assert sd.view["nvt_ion"] == (1 if name == "nvt_scale" else 2)
    nsteps, every = (int(v) for v in fx["s_nsteps"])
    print("case", name, path, precision, flush=True)
    if path == "percall":
        eng.verlet(0, nsteps, eflag_every=every, precision=precision)
    elif path == "graph":
        eng.verlet_graph(0, nsteps, eflag_every=every, precision=precision)
    else:
        eng.run_batched(nsteps, eflag_every=every, precision=precision)
    eng.sync()
    tol = 1e-12 if precision == M.FP64 else 1e-6
    for r in range(2 if path == "batched" else 1):
        st = eng.get_state(r)
        errs = {k: E.rel(getattr(st, k), fx["s_" + k]) for k in ("atom_x", "atom_v", "atom_f", "mol_x", "mol_v")}
        z, zr = st.c.zeta_ion, fx["s_zeta_ion"][0]
        print(name, path, precision, r, errs, "zeta", z, zr)
        for k, e in errs.items():
            assert e < (1e-5 if precision == M.MIXED and k == "atom_f" else tol), k
        if name == "nvt_nh":
            assert abs(z - zr) <= tol * abs(zr) and zr != 0
        else:
            assert z == 0
        # mixed: the energy bar of the one-shot force test (rtol 1e-5, atol
        # 1e-4 kcal/mol: small buckets are sums of cancelling FP32 pair terms)
        if precision == M.FP64:
            np.testing.assert_allclose(st.U, fx["s_U"], rtol=1e-10, atol=1e-8)
        else:
            np.testing.assert_allclose(st.U, fx["s_U"], rtol=1e-5, atol=1e-4)


def _cond(d, nvt):
    m = configs.molecule_files(d)
    pep = configs.write_peptide(d + "/pep30.txt", seed=5, n=30)
    return configs.Condition(pep, m["methanol"], 12, "N2", 400, pressure=6e5, margin=1.0, shake_h=True,
                             ion_potential="AMBER", nvt_ion=nvt).write(d + "/t.cond")


@pytest.mark.parametrize("nvt", ["scale", "250"])
@pytest.mark.parametrize("path", ["percall", "batched"])
def test_thermostat_with_rebuilds_vs_oracle(tmp_path, nvt, path):
    sd = M.SysDef(_cond(str(tmp_path), nvt))
    eng = M.Engine(0)
    orcs, Ss = [], []
    for r, seed in enumerate((3, 4)):
        st0 = sd.init_state(seed)
        eng.add(sd, st0, seed=seed, replica_id=r)
        orc = E.oracle_for(sd, seed=seed, replica=r)
        orcs.append(orc)
        Ss.append(orc.new_state(st0.as_dict()))
    eng.commit()
    print("case", nvt, path, flush=True)
    for r in range(2):
        eng.make_pair(r)
        orcs[r].make_pair(Ss[r])
    if path == "percall":
        for r in range(2):
            eng.verlet(r, 200, eflag_every=10, precision=M.FP64)
    else:
        eng.run_batched(200, eflag_every=10, precision=M.FP64)
    eng.sync()
    for r in range(2):
        assert orcs[r].verlet(Ss[r], 200, eflag_every=10)[0] == 0
        st, S = eng.get_state(r), Ss[r]
        assert st.c.rebuild_count == S.c.rebuild_count > 1
        m = E.in_atom_mask(sd, st.region)
        x_err = np.abs(st.atom_x.reshape(-1, 3)[m] - S.atom_x.reshape(-1, 3)[m]).max()
        v_err = E.rel(st.atom_v.reshape(-1, 3)[m], S.atom_v.reshape(-1, 3)[m])
        print(nvt, path, r, "max |dx|", x_err, "rel v err", v_err, "zeta", st.c.zeta_ion, S.c.zeta_ion)
        assert x_err < 1e-8 and v_err < 1e-7
        assert abs(st.c.zeta_ion - S.c.zeta_ion) <= 1e-7 * max(abs(S.c.zeta_ion), 1e-12)
        if nvt == "250":
            assert S.c.zeta_ion != 0


def test_set_nve_mid_run_and_set_state_isolation(tmp_path):
    """Relaxation with velocity scaling, then MD::setNVE for the main run; and
    mdn_set_state of replica 1 leaves replica 0's state as it was."""
    sd = M.SysDef(_cond(str(tmp_path), "scale"))
    v = sd.view
    eng = M.Engine(0)
    orcs, Ss = [], []
    for r, seed in enumerate((5, 6)):
        st0 = sd.init_state(seed)
        eng.add(sd, st0, seed=seed, replica_id=r)
        orc = E.oracle_for(sd, seed=seed, replica=r)
        orcs.append(orc)
        Ss.append(orc.new_state(st0.as_dict()))
    eng.commit()
    for r in range(2):
        eng.make_pair(r)
        orcs[r].make_pair(Ss[r])
    eng.run_batched(40, eflag_every=0, precision=M.FP64)
    eng.set_thermostat(-1, M.NVE)
    eng.run_batched(40, eflag_every=20, step0=40, precision=M.FP64)
    eng.sync()
    from oracle import oracle as O
    for r in range(2):
        orcs[r].verlet(Ss[r], 40)
        # the same state under an NVE oracle (flags after setNVE)
        nve = O.Oracle(dict(E.oracle_sysd(v, seed=(5, 6)[r], replica=r), nvt_ion=0))
        S2 = nve.new_state(Ss[r].as_dict())
        assert nve.verlet(S2, 40, eflag_every=20, step0=40)[0] == 0
        st = eng.get_state(r)
        m = E.in_atom_mask(sd, st.region)
        x_err = np.abs(st.atom_x.reshape(-1, 3)[m] - S2.atom_x.reshape(-1, 3)[m]).max()
        print("setNVE replica", r, "max |dx|", x_err)
        assert x_err < 1e-8
        np.testing.assert_allclose(st.U, S2.U, rtol=1e-7, atol=1e-7)
    # set_state on replica 1 only
    before = eng.get_state(0)
    eng.set_state(1, sd.init_state(99))
    after = eng.get_state(0)
    for k in ("atom_x", "atom_v", "atom_f", "atom_w", "mol_x", "region"):
        np.testing.assert_array_equal(getattr(after, k), getattr(before, k), err_msg=k)
    assert after.c.rebuild_count == before.c.rebuild_count and after.c.loopPair == before.c.loopPair
    np.testing.assert_array_equal(after.list("gi"), before.list("gi"))
    a0, rb0, _, _ = eng.counters(0)
    assert a0 > 0 and rb0 == before.c.rebuild_count
    with pytest.raises(M.MdnError):
        eng.set_thermostat(0, 7)
