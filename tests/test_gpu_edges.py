"""Edge and failure paths of the step on the GPU, each against the oracle (or
the reference's own output) from the same state:

* gas-gas LJ (GasGas.cpp:4-63, the intended in-domain pair list): the
  reference's PotentialGasGas::compute on the golden state, per-call FP64
  and mixed; 200-step trajectories with rebuilds on the per-call and
  batched paths;
* SHAKE as shake.cpp writes it (mode 2: its own update, positions only,
  shake.cpp:44-92) on methanol, both paths;
* SHAKE that does not converge: isolated H-H constraints under mode 2
  oscillate with a growing amplitude (shake.cpp:67 multiplies by the inverse
  reduced mass), the iteration cap gives up, the molecule moves
  unconstrained for that step and the failure is counted -- the same steps
  and counts as the oracle's;
* a replica whose in-domain atoms outgrow the batched launch's LDS at a
  rebuild inside a step: the step is left unfinished, flagged, and finished
  by mdn_verlet (or mdn_verlet_graph), which then matches the oracle;
* a molecule leaving through r = 110 A and one entering through it, with a
  rebuild every step (Margin 1e-6 A: loop_update = 0, pairlist.cpp:40-46).
"""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from md_nucleation_amd import configs
from tests import engine_util as E
from tests import golden_util as G
from tests.golden.cases import write_case

pytestmark = pytest.mark.gpu


def _compare(eng, r, sd, S, tol_x=1e-8, tol_f=1e-7, what=""):
    st = eng.get_state(r)
    assert st.c.rebuild_count == S.c.rebuild_count, what
    np.testing.assert_array_equal(st.region, S.region, err_msg=what)
    m = E.in_atom_mask(sd, st.region)
    dx = float(np.abs(st.atom_x.reshape(-1, 3)[m] - S.atom_x.reshape(-1, 3)[m]).max())
    fe = E.rel(st.atom_f.reshape(-1, 3)[m], S.atom_f.reshape(-1, 3)[m])
    print(what, "max |dx| %.2e, rel F %.2e" % (dx, fe), flush=True)
    assert dx < tol_x, what
    assert fe < tol_f, what
    return st


@pytest.mark.parametrize("precision", [M.FP64, M.MIXED])
def test_gg_forces_vs_reference(tmp_path, precision):
    fx = G.load("gg")
    sd = M.SysDef(write_case("gg", str(tmp_path)))
    st = E.state_from(sd, G.state(fx), gg=fx["gg_list"].astype(np.int32), loopPair=0, loop_update=10 ** 8)
    eng = M.Engine(0)
    eng.add(sd, st)
    eng.commit()
    eng.zero_forces(0)
    eng.compute(0, M.MDN_GG, eflag=1, precision=precision)
    out = eng.get_state(0)
    assert out.list("gg").size == fx["gg_list"].size
    err = E.rel(out.atom_f, fx["gg_atom_f"])
    print("gg per-call precision", precision, "rel F vs reference", err)
    assert err < (1e-12 if precision == M.FP64 else 1e-5)
    # the reference adds gg energy to Uvg (GasGas.cpp:34); here it is Ugg
    np.testing.assert_allclose(out.U[4], fx["gg_U"][5], rtol=1e-12 if precision == M.FP64 else 1e-5)
    eng.close()


def _gg_cond(d, **kw):
    m = configs.molecule_files(d)
    return configs.Condition(m["na"], m["methanol"], 12, "N2", 400, pressure=6e5, margin=1.0,
                             interactions={"gg": "LJ", "gi": "LJ", "gv": "LJ", "vi": "LJcoul", "vv": "LJcoul"},
                             **kw).write(d + "/gg.cond")


def _case(tmp_path, cond_fn, seeds, shake=None, max_iter=1000):
    sds, orcs, Ss = [], [], []
    eng = M.Engine(0)
    for r, seed in enumerate(seeds):
        d = tmp_path / ("r%d" % r)
        d.mkdir()
        sd = M.SysDef(cond_fn(str(d)))
        if shake is not None:
            sd.set_shake(shake, max_iter)
        st0 = sd.init_state(seed)
        eng.add(sd, st0, seed=seed, replica_id=r)
        orc = E.oracle_for(sd, seed=seed, replica=r)
        S = orc.new_state(st0.as_dict())
        sds.append(sd)
        orcs.append(orc)
        Ss.append(S)
    eng.commit()
    for r in range(len(seeds)):
        eng.make_pair(r)
        assert orcs[r].make_pair(Ss[r]) == 0
    return eng, sds, orcs, Ss


def _run(eng, orcs, Ss, path, nsteps, precision=M.FP64):
    if path == "percall":
        for r in range(len(orcs)):
            eng.verlet(r, nsteps, eflag_every=10, precision=precision)
    else:
        eng.run_batched(nsteps, eflag_every=10, precision=precision)
    eng.sync()
    for r in range(len(orcs)):
        assert orcs[r].verlet(Ss[r], nsteps, eflag_every=10)[0] == 0


@pytest.mark.parametrize("path", ["percall", "batched"])
def test_gg_trajectory_vs_oracle(tmp_path, path):
    eng, sds, orcs, Ss = _case(tmp_path, _gg_cond, (31, 32))
    print("case gg", path, flush=True)
    _run(eng, orcs, Ss, path, 200)
    for r in range(2):
        st = _compare(eng, r, sds[r], Ss[r], what="gg %s replica %d" % (path, r))
        assert st.c.rebuild_count > 1
        gg = st.list("gg")
        assert gg.shape[0] > 100
        np.testing.assert_array_equal(gg, Ss[r].lists()["gg"])
        np.testing.assert_allclose(st.U, Ss[r].U, rtol=1e-6, atol=1e-6)
        assert abs(st.U[4]) > 0
    eng.close()


def test_gg_mixed_force_error(tmp_path):
    eng, sds, orcs, Ss = _case(tmp_path, _gg_cond, (33,))
    _run(eng, orcs, Ss, "batched", 1, precision=M.MIXED)
    st = eng.get_state(0)
    m = E.in_atom_mask(sds[0], st.region)
    err = E.rel(st.atom_f.reshape(-1, 3)[m], Ss[0].atom_f.reshape(-1, 3)[m])
    print("gg mixed rel RMS force error", err)
    assert err < 1e-5
    eng.close()


def _meoh_cond(d):
    m = configs.molecule_files(d)
    return configs.Condition(m["na"], m["methanol"], 12, "N2", 400, pressure=6e5, margin=1.0).write(d + "/m.cond")


@pytest.mark.parametrize("path", ["percall", "batched"])
def test_shake_as_written_vs_oracle(tmp_path, path):
    """Mode 2: shake.cpp's update (positions only) on methanol's X-H bonds."""
    eng, sds, orcs, Ss = _case(tmp_path, _meoh_cond, (41, 42), shake=2)
    assert sds[0].view["shake"] == 2
    print("case shake mode 2", path, flush=True)
    _run(eng, orcs, Ss, path, 200)
    for r in range(2):
        _compare(eng, r, sds[r], Ss[r], what="mode 2 %s replica %d" % (path, r))
        a, rb, it, nfail = eng.counters(r)
        assert nfail == Ss[r].c.shake_failures == 0 and it > 1
    eng.close()


def _h2_cond(d):
    m = configs.molecule_files(d)
    h2 = configs.write_molecule(d + "/h2.txt", [("#h1", 1.008, 0.0157, 2.47135)], [("h1-h1", 300.0, 0.74)], [], [],
                                [(1, 0.0, -0.37, 0.0, 0.0), (1, 0.0, 0.37, 0.0, 0.0)], [(1, 2, 1)], [], [])
    return configs.Condition(m["na"], h2, 20, "N2", 400, pressure=6e5, margin=1.0).write(d + "/h2.cond")


@pytest.mark.parametrize("path", ["percall", "batched"])
def test_shake_failure_is_counted_and_undone(tmp_path, path):
    """Isolated H-H constraints under mode 2 diverge (each sweep multiplies
    the constraint error by 1 - (1/m_H + 1/m_H)^2, about -2.9): after 20
    sweeps the cap gives up and the molecule moves unconstrained for the
    step (the oracle counts about 550 such failures in 30 steps)."""
    eng, sds, orcs, Ss = _case(tmp_path, _h2_cond, (51,), shake=2, max_iter=20)
    assert sds[0].view["species"][2]["rigid"].size == 3
    print("case shake failure", path, flush=True)
    _run(eng, orcs, Ss, path, 30)
    _compare(eng, 0, sds[0], Ss[0], what="shake failure %s" % path)
    a, rb, it, nfail = eng.counters(0)
    print("failures", nfail, "oracle", Ss[0].c.shake_failures, "max sweeps", it)
    assert nfail == Ss[0].c.shake_failures > 0
    assert it == 20
    eng.close()


def _overflow_cond(d):
    m = configs.molecule_files(d)
    pep = configs.write_peptide(d + "/pep30.txt", seed=5, n=30)
    return configs.Condition(pep, m["methanol"], 1, "N2", 100, pressure=1e6, ion_potential="AMBER").write(
        d + "/o.cond")


@pytest.mark.parametrize("path", ["percall", "graph"])
def test_lds_overflow_inside_a_step_is_finished_by_verlet(tmp_path, path):
    """The 30-atom chain, one methanol and 46 of its 100 N2 in the domain
    (128 atoms, the launch's LDS capacity) with 54 N2 inside the sphere but
    marked out: the rebuild of the first step lets them in, the replica
    outgrows the launch and is stopped inside the step.  mdn_verlet finishes
    that step and goes on; the trajectory is the oracle's."""
    d = str(tmp_path)
    sd = M.SysDef(_overflow_cond(d))
    v = sd.view
    assert v["d_size"] < 80  # every molecule within 110 A of the ion
    eng = M.Engine(0)
    eng.add(sd, sd.init_state(61), seed=61)
    eng.commit()
    eng.make_pair(0)
    st = eng.get_state(0, cap=20000)
    n_ion, na_v = v["species"][0]["mass"].size, v["species"][2]["mass"].size
    assert st.region.all()
    st.region[1 + 46:1 + v["n_gas"]] = 0
    n_in = n_ion + na_v + 2 * 46
    assert n_in == 128
    a = st.as_dict()
    a.update(loopPair=0, loop_update=0, gi=np.zeros(0, np.int32), gv=np.zeros(0, np.int32))
    st = E.state_from(sd, a)
    st.set_caps(20000)
    eng.set_state(0, st)
    orc = E.oracle_for(sd, seed=61)
    S = orc.new_state(st.as_dict())
    eng.set_batch_shape(256, 128, 0, 0)
    print("case LDS overflow inside a step", path, flush=True)
    eng.run_batched(5, precision=M.FP64)
    eng.sync()
    with pytest.raises(M.MdnError, match="status"):
        eng.counters(0)
    mid = eng.get_state(0, cap=20000)
    assert mid.region.all() and mid.c.rebuild_count == st.c.rebuild_count + 1
    assert np.all(mid.atom_f == 0)
    with pytest.raises(M.MdnError, match="unfinished"):
        eng.update_position(0)
    if path == "percall":
        eng.verlet(0, 5, eflag_every=5, precision=M.FP64)
    else:
        eng.verlet_graph(0, 5, eflag_every=5, precision=M.FP64)
    eng.sync()
    assert orc.verlet(S, 5, eflag_every=5)[0] == 0
    out = _compare(eng, 0, sd, S, what="LDS overflow %s" % path)
    np.testing.assert_allclose(out.U, S.U, rtol=1e-8, atol=1e-8)
    eng.close()


def _edge_cond(d):
    m = configs.molecule_files(d)
    return configs.Condition(m["na"], m["methanol"], 3, "N2", 2000, margin=1e-6).write(d + "/e.cond")


@pytest.mark.parametrize("path", ["percall", "batched"])
def test_molecules_cross_the_domain_boundary(tmp_path, path):
    """A methanol at 109.99 A moving out at 0.004 A/fs and one outside at
    110.002 A moving in (box 436 A, so no wrap), Na+ at rest at the origin;
    Margin 1e-6 A makes every step rebuild.  Step by step, classification,
    rebuild count and positions are the oracle's.  The outside molecule's
    first ballistic advance (pairlist.cpp:20-29) brings it in on step 1 and
    it stays in; the inside one leaves on step 3 (its centre at 110.002 A),
    after which its velocity is a fresh Maxwell draw at every rebuild
    (updateGasIn.cpp:43-45), so later steps are only compared, not
    predicted."""
    d = str(tmp_path)
    sd = M.SysDef(_edge_cond(d))
    v = sd.view
    assert v["d_size"] > 2 * 110.01
    eng = M.Engine(0)
    eng.add(sd, sd.init_state(71), seed=71)
    eng.commit()
    eng.make_pair(0)
    a = eng.get_state(0, cap=20000).as_dict()
    X, V = a["atom_x"].reshape(-1, 3), a["atom_v"].reshape(-1, 3)
    mx, mv = a["mol_x"].reshape(-1, 3), a["mol_v"].reshape(-1, 3)
    V[0] = 0.0
    n_ion, na_v = v["species"][0]["mass"].size, v["species"][2]["mass"].size
    va, vb = 1 + v["n_gas"], 2 + v["n_gas"]
    oa = n_ion + 2 * v["n_gas"]
    ua, ub = np.array([0.6, 0.8, 0.0]), np.array([0.0, -0.6, 0.8])
    tm = v["species"][2]["tmpl_x"].reshape(-1, 3)
    mx[va], mv[va] = 109.99 * ua, 0.004 * ua
    X[oa:oa + na_v] = tm + mx[va]
    V[oa:oa + na_v] = mv[va]
    a["region"][va] = 1
    mx[vb], mv[vb] = 110.002 * ub, -0.004 * ub
    a["region"][vb] = 0
    a["atom_f"][:] = 0
    # the first step rebuilds before any force: its lists are made then
    for k in ("gi", "gv", "gg"):
        a[k] = np.zeros(0, np.int32)
    st = E.state_from(sd, a)
    st.set_caps(20000)
    eng.set_state(0, st)
    orc = E.oracle_for(sd, seed=71)
    S = orc.new_state(st.as_dict())
    seen = []
    print("case boundary crossing", path, flush=True)
    for step in range(6):
        if path == "percall":
            eng.verlet(0, 1, precision=M.FP64)
        else:
            eng.run_batched(1, precision=M.FP64)
        assert orc.verlet(S, 1)[0] == 0
        out = _compare(eng, 0, sd, S, tol_x=1e-10, tol_f=1e-9, what="%s step %d" % (path, step + 1))
        assert out.c.rebuild_count == st.c.rebuild_count + step + 1
        seen.append((int(out.region[va]), int(out.region[vb])))
    print(path, "regions of the two molecules per step", seen)
    assert seen[:3] == [(1, 1), (1, 1), (0, 1)] and all(b == 1 for _, b in seen)
    eng.close()
