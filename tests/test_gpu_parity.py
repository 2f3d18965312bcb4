"""libmdn on the GPU against the reference's golden vectors and the oracle.

Bars: FP64 forces within 1e-12 relative of the reference's own forces on the
same state; pair lists and classification bit-identical; mixed precision
(FP32 pair arithmetic, FP64 accumulation) within 1e-5 relative RMS force
error of FP64; trajectories (with rebuilds) within the tolerances stated per
test against the oracle run from the same state with the same random stream.
"""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from oracle import oracle as O
from tests import engine_util as E
from tests import golden_util as G
from tests.golden.cases import write_case

pytestmark = pytest.mark.gpu

ALL_BITS = 0xFFF


def _engine_on_fixture(tmp_path, name, fx, prefix="init_", lists=True, shake=0, **extra):
    sd = M.SysDef(write_case(name, str(tmp_path)))
    sd.set_shake(shake)
    arrays = G.state(fx, prefix)
    if lists:
        arrays["gi"] = fx["f_list_gi"].astype(np.int32)
        arrays["gv"] = fx["f_list_gv"].astype(np.int32)
    arrays.update(extra)
    st = E.state_from(sd, arrays)
    eng = M.Engine(0)
    eng.add(sd, st, seed=7)
    eng.commit()
    return eng, sd


@pytest.mark.parametrize("name", ["c1", "pep"])
@pytest.mark.parametrize("precision", [M.FP64, M.MIXED])
def test_forces_vs_reference(tmp_path, name, precision):
    fx = G.load(name)
    eng, sd = _engine_on_fixture(tmp_path, name, fx)
    eng.zero_forces(0)
    eng.compute(0, sd.view["flags"], 1, precision)
    st = eng.get_state(0)
    ref = fx["f_atom_f"]
    err = E.rel(st.atom_f, ref)
    print(name, "precision", precision, "rel RMS force error", err)
    if precision == M.FP64:
        assert err < 1e-12
        np.testing.assert_allclose(st.U, fx["f_U"], rtol=1e-11, atol=1e-9)
    else:
        assert err < 1e-5
        np.testing.assert_allclose(st.U, fx["f_U"], rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("op,bit", [("tersoff", M.MDN_ION_TERSOFF), ("sw", M.MDN_ION_SW)])
def test_manybody_vs_reference(tmp_path, op, bit):
    fx = G.load("si")
    sd = M.SysDef(write_case(op, str(tmp_path)))
    n = fx[op + "_sys_ion_mass"].size
    st = sd.new_state()
    st.atom_x[: 3 * n] = fx[op + "_atom_x"][: 3 * n]
    st.region[:] = 1
    st.atom_w[:] = 1
    eng = M.Engine(0)
    eng.add(sd, st)
    eng.commit()
    eng.zero_forces(0)
    eng.compute(0, bit, 1, M.FP64)
    out = eng.get_state(0)
    assert E.rel(out.atom_f[: 3 * n], fx[op + "_ion_f"]) < 1e-12
    np.testing.assert_allclose(out.U[0], fx[op + "_U_raw"][0] * 23.06, rtol=1e-10)


def test_born_vs_reference(tmp_path):
    fx = G.load("bmh")
    sd = M.SysDef(write_case("bmh", str(tmp_path)))
    st = E.state_from(sd, G.state(fx))
    eng = M.Engine(0)
    eng.add(sd, st)
    eng.commit()
    eng.zero_forces(0)
    eng.compute(0, M.MDN_ION_BMH, 1, M.FP64)
    out = eng.get_state(0)
    n = G.sub(fx, "sys_")["ion_mass"].size
    assert E.rel(out.atom_f[: 3 * n], fx["born_ion_f"]) < 1e-12
    np.testing.assert_allclose(out.U[0], fx["born_U_ion"][0], rtol=1e-11)


@pytest.mark.parametrize("name", ["c1", "pep"])
def test_make_pair_bit_exact(tmp_path, name):
    """Classification, COMs, out-molecule advance and wrap, gi/gv lists of
    one MD::make_pair: bit-identical to the reference's on the same state."""
    fx = G.load(name)
    lp = int(fx["mp_loopPair"][0])
    eng, sd = _engine_on_fixture(tmp_path, name, fx, lists=False, loopPair=lp, loop_update=10 ** 8)
    eng.make_pair(0)
    st = eng.get_state(0)
    np.testing.assert_array_equal(st.region, fx["mp_region"].astype(np.uint8))
    np.testing.assert_array_equal(st.mol_x, fx["mp_mol_x"])
    np.testing.assert_array_equal(st.list("gi").reshape(-1), fx["mp_list_gi"].astype(np.int32))
    np.testing.assert_array_equal(st.list("gv").reshape(-1), fx["mp_list_gv"].astype(np.int32))
    # and the random part against the oracle (same Philox stream; device and
    # host libm may differ by an ulp in log/cos/sin)
    orc = E.oracle_for(sd, seed=7)
    S = orc.new_state(dict(G.state(fx), loopPair=lp, loop_update=10 ** 8))
    orc.make_pair(S)
    np.testing.assert_allclose(st.mol_v, S.mol_v, rtol=1e-13, atol=1e-17)
    np.testing.assert_allclose(st.atom_x, S.atom_x, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(st.atom_v, S.atom_v, rtol=1e-12, atol=1e-16)
    assert st.c.loop_update == S.c.loop_update


def _cond(tmp_path, margin=1.0, n_gas=400, n_vap=12, pressure=6e5, shake=True, ion="na", vap="methanol"):
    from md_nucleation_amd import configs
    d = str(tmp_path)
    m = configs.molecule_files(d)
    # ions: Na+ (one atom), a 30- and a 70-atom AMBER chain (the latter ends
    # in a partial 64-atom block), a 27-atom NaCl cluster (BMH, no AMBER pairs)
    pot = None
    if ion in m:
        ionf = m[ion]
    elif ion == "nacl":
        ionf, pot = configs.write_nacl_cluster(d + "/nacl27.txt", n_side=3), "Born-Mayer-Huggins-NaCl"
    else:
        n = 70 if ion == "pep70" else 30
        ionf, pot = configs.write_peptide(d + "/pep%d.txt" % n, seed=5, n=n), "AMBER"
    return configs.Condition(ionf, m[vap], n_vap, "N2", n_gas, pressure=pressure, margin=margin,
                             shake_h=shake, ion_potential=pot).write(d + "/t.cond")


def _trajectory_case(tmp_path, path, nsteps, precision, ion="na", shake=True, seeds=(3,), vap="methanol"):
    sds, sts, orcs, Ss = [], [], [], []
    eng = M.Engine(0)
    for r, seed in enumerate(seeds):
        sd = M.SysDef(_cond(tmp_path / ("r%d" % r), ion=ion, shake=shake, vap=vap))
        st0 = sd.init_state(seed)
        eng.add(sd, st0, seed=seed, replica_id=r)
        orc = E.oracle_for(sd, seed=seed, replica=r)
        S = orc.new_state(st0.as_dict())
        sds.append(sd); sts.append(st0); orcs.append(orc); Ss.append(S)
    eng.commit()
    # the constructor's rebuild (md.cpp:34) on both sides
    for r in range(len(seeds)):
        eng.make_pair(r)
        orcs[r].make_pair(Ss[r])
    if path == "percall":
        for r in range(len(seeds)):
            eng.verlet(r, nsteps, eflag_every=10, precision=precision)
    else:
        eng.run_batched(nsteps, eflag_every=10, precision=precision)
    eng.sync()
    for r in range(len(seeds)):
        rc, _ = orcs[r].verlet(Ss[r], nsteps, eflag_every=10)
        assert rc == 0
    return eng, sds, orcs, Ss


@pytest.mark.parametrize("path,ion", [("percall", "na"), ("batched", "na"), ("percall", "pep"), ("batched", "pep"),
                                      ("batched", "nacl"), ("batched", "pep70")])
def test_trajectory_fp64_vs_oracle(tmp_path, path, ion):
    """200 FP64 steps with SHAKE and several pair-list rebuilds (Margin 1 A)."""
    eng, sds, orcs, Ss = _trajectory_case(tmp_path, path, 200, M.FP64, ion=ion, seeds=(3, 4))
    for r in range(2):
        st = eng.get_state(r)
        S = Ss[r]
        assert st.c.rebuild_count == S.c.rebuild_count and st.c.rebuild_count > 1
        np.testing.assert_array_equal(st.region, S.region)
        m = E.in_atom_mask(sds[r], st.region)
        x_err = np.abs(st.atom_x.reshape(-1, 3)[m] - S.atom_x.reshape(-1, 3)[m]).max()
        f_err = E.rel(st.atom_f.reshape(-1, 3)[m], S.atom_f.reshape(-1, 3)[m])
        print(path, ion, "replica", r, "max |dx|", x_err, "rel force err", f_err)
        assert x_err < 1e-8
        assert f_err < 1e-7
        np.testing.assert_allclose(st.U, S.U, rtol=1e-6, atol=1e-6)
        a, rb, it, nfail = eng.counters(r)
        assert rb == S.c.rebuild_count


@pytest.mark.parametrize("ion", ["na", "pep", "nacl", "pep70"])
def test_batched_mixed_force_error_one_step(tmp_path, ion):
    """After one step both sides hold F(x1) for identical x1 (x1 depends only on
    the state's forces): the mixed-precision gate, relative RMS <= 1e-5."""
    eng, sds, orcs, Ss = _trajectory_case(tmp_path, "batched", 1, M.MIXED, ion=ion, seeds=(5, 6, 7))
    for r in range(3):
        st = eng.get_state(r)
        m = E.in_atom_mask(sds[r], st.region)
        err = E.rel(st.atom_f.reshape(-1, 3)[m], Ss[r].atom_f.reshape(-1, 3)[m])
        print(ion, r, "mixed rel RMS force error", err)
        assert err < 1e-5


def test_shake_groups_vs_oracle(tmp_path):
    """Ethylene glycol vapor: under SHAKE H its six X-H constraints form four
    groups that share no atom (O-H, C-H x2, C-H x2, O-H), which the batched
    kernel's overlapped SHAKE runs on four lanes per molecule.  200 FP64 steps
    against the oracle's single sequential loop in file order: the positions
    agree like those of methanol (two groups: C-H x3, O-H)."""
    eng, sds, orcs, Ss = _trajectory_case(tmp_path, "batched", 200, M.FP64, seeds=(3, 4), vap="glycol")
    rig = sds[0].view["species"][2]["rigid"].reshape(-1, 3)
    assert rig.shape[0] == 6
    for r in range(2):
        st = eng.get_state(r)
        S = Ss[r]
        assert st.c.rebuild_count == S.c.rebuild_count
        np.testing.assert_array_equal(st.region, S.region)
        m = E.in_atom_mask(sds[r], st.region)
        x_err = np.abs(st.atom_x.reshape(-1, 3)[m] - S.atom_x.reshape(-1, 3)[m]).max()
        f_err = E.rel(st.atom_f.reshape(-1, 3)[m], S.atom_f.reshape(-1, 3)[m])
        print("glycol replica", r, "max |dx|", x_err, "rel force err", f_err)
        assert x_err < 1e-8
        assert f_err < 1e-7
        a, rb, it, nfail = eng.counters(r)
        assert nfail == 0 and it > 1


@pytest.mark.parametrize("vap", ["methanol", "glycol"])
def test_shake_constraints_hold(tmp_path, vap):
    eng, sds, orcs, Ss = _trajectory_case(tmp_path, "batched", 100, M.FP64, seeds=(9,), vap=vap)
    st = eng.get_state(0)
    v = sds[0].view
    rig = v["species"][2]["rigid"].reshape(-1, 3)
    na_v = v["species"][2]["mass"].size
    off_v = v["species"][0]["mass"].size + v["n_gas"] * v["species"][1]["mass"].size
    X = st.atom_x.reshape(-1, 3)
    worst = 0.0
    for mol in np.flatnonzero(st.region[1 + v["n_gas"]:]):
        o = off_v + mol * na_v
        for i, j, t in rig:
            d = np.linalg.norm(X[o + i] - X[o + j])
            worst = max(worst, abs(d * d - v["btype"][2 * t + 1] ** 2))
    assert worst < 1e-9


def test_list_capacity_overflow_is_an_error(tmp_path):
    sd = M.SysDef(_cond(tmp_path))
    st = sd.init_state(3)
    st.set_caps(1)
    eng = M.Engine(0)
    eng.add(sd, st)
    eng.commit()
    eng.make_pair(0)
    eng.sync()
    with pytest.raises(M.MdnError, match="status"):
        eng.counters(0)


def test_lds_capacity_overflow_is_an_error(tmp_path):
    sd = M.SysDef(_cond(tmp_path))
    st = sd.init_state(3)
    eng = M.Engine(0)
    eng.add(sd, st)
    eng.commit()
    eng.make_pair(0)
    eng.set_batch_shape(256, 32, 0, 0)
    eng.run_batched(5)
    eng.sync()
    with pytest.raises(M.MdnError, match="status"):
        eng.counters(0)


def test_graph_replays_match_eager(tmp_path):
    """The captured per-call step, replayed many times with eager work and
    synchronisations in between, gives the eager path's trajectory."""
    sd = M.SysDef(_cond(tmp_path, margin=2.0))
    st0 = sd.init_state(21)
    engs = []
    for _ in range(2):
        e = M.Engine(0)
        e.add(sd, st0, seed=21)
        e.commit()
        e.make_pair(0)
        engs.append(e)
    eager, graphed = engs
    for chunk in range(4):
        eager.verlet(0, 25, eflag_every=10, step0=25 * chunk, precision=M.FP64)
        graphed.verlet_graph(0, 25, eflag_every=10, step0=25 * chunk, precision=M.FP64)
        graphed.sync()
        a, b = eager.get_state(0), graphed.get_state(0)
        np.testing.assert_array_equal(a.region, b.region)
        assert np.abs(a.atom_x - b.atom_x).max() < 1e-9, chunk
        np.testing.assert_allclose(a.U, b.U, rtol=1e-9, atol=1e-9)
        assert a.c.rebuild_count == b.c.rebuild_count
