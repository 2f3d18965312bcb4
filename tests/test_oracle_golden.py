"""The oracle against golden vectors written by the reference's own code.

tests/golden/make_golden.py ran the reference (compiled from its sources) on
the states stored in the fixtures; here the oracle runs on the same states.
A restatement that drops a term, flips a sign or swaps an index fails these.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_util as G

CASES = ["c1", "pep"]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_forces_and_energies(name):
    fx = G.load(name)
    orc, _ = G.oracle_for(fx)
    S = orc.new_state(G.state(fx))
    G.set_lists(S, fx)
    orc.forces(S, eflag=1)
    assert rel(S.atom_f, fx["f_atom_f"]) < 1e-13
    np.testing.assert_allclose(S.U, fx["f_U"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", CASES)
def test_trajectory_no_rebuild(name):
    fx = G.load(name)
    orc, _ = G.oracle_for(fx)
    S = orc.new_state(G.state(fx, loopPair=0, loop_update=10 ** 8))
    G.set_lists(S, fx)
    nsteps, every = (int(v) for v in fx["s_nsteps"])
    rc, _ = orc.verlet(S, nsteps, eflag_every=every)
    assert rc == 0
    for k in ("atom_x", "atom_v", "atom_f", "atom_w", "mol_x"):
        assert rel(getattr(S, k), fx["s_" + k]) < 1e-12, k
    np.testing.assert_allclose(S.U, fx["s_U"], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("name", CASES)
def test_make_pair_deterministic_part(name):
    """Classification, COMs, out-molecule advance + wrap, and the gi/gv lists
    of MD::make_pair (the random velocity/orientation draws are this build's
    own stream and are compared elsewhere)."""
    fx = G.load(name)
    orc, _ = G.oracle_for(fx)
    lp = int(fx["mp_loopPair"][0])
    S = orc.new_state(G.state(fx, loopPair=lp, loop_update=10 ** 8))
    before = S.region.copy()
    assert orc.make_pair(S) == 0
    np.testing.assert_array_equal(S.region, fx["mp_region"].astype(np.uint8))
    # molecule positions: COM for in-domain, advanced/wrapped for the others
    np.testing.assert_array_equal(S.mol_x, fx["mp_mol_x"])
    L = S.lists()
    np.testing.assert_array_equal(L["gi"].reshape(-1), fx["mp_list_gi"].astype(np.int32))
    np.testing.assert_array_equal(L["gv"].reshape(-1), fx["mp_list_gv"].astype(np.int32))
    # atoms of molecules that stayed in are untouched by the rebuild
    stayed = np.flatnonzero(before & S.region)
    assert stayed.size > 0
    n_in = int(S.region.sum())
    assert n_in == int(fx["mp_region"].sum())


def test_shake_fixed_point():
    for name in CASES:
        fx = G.load(name)
        # mode 2 is shake.cpp's own iteration
        orc, sysd = G.oracle_for(fx, shake=2)
        keys = [k for k in fx if k.startswith("shake") and k.endswith("_out")]
        assert keys
        for k in keys:
            mol = k[len("shake"):-len("_out")]
            old, new, out = fx["shake%s_old" % mol], fx["shake%s_new" % mol].copy(), fx[k]
            it = O.lib().orc_shake_molecule(orc.h, new.ctypes.data_as(O._dp), old.ctypes.data_as(O._dp), 2)
            assert it > 0
            # same iteration as the reference -> same fixed point to rounding
            np.testing.assert_allclose(new, out, rtol=0, atol=1e-12)


def test_shake_modes_share_the_fixed_point():
    """Mode 1 (textbook update) converges to the constraint solution of mode 2
    (shake.cpp's update) and needs fewer sweeps."""
    fx = G.load("c1")
    o1, _ = G.oracle_for(fx, shake=1)
    o2, _ = G.oracle_for(fx, shake=2)
    for k in [k for k in fx if k.startswith("shake") and k.endswith("_out")]:
        mol = k[len("shake"):-len("_out")]
        old = fx["shake%s_old" % mol]
        a, b = fx["shake%s_new" % mol].copy(), fx["shake%s_new" % mol].copy()
        i1 = O.lib().orc_shake_molecule(o1.h, a.ctypes.data_as(O._dp), old.ctypes.data_as(O._dp), 2)
        i2 = O.lib().orc_shake_molecule(o2.h, b.ctypes.data_as(O._dp), old.ctypes.data_as(O._dp), 2)
        assert 0 < i1 < i2
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9)


def test_born_mayer_huggins():
    fx = G.load("bmh")
    orc, _ = G.oracle_for(fx, flags=O.ORC_ION_BMH)
    S = orc.new_state(G.state(fx))
    S.atom_f[:] = 0
    orc.force_term(S, O.ORC_ION_BMH, 1)
    n = G.sub(fx, "sys_")["ion_mass"].size
    assert rel(S.atom_f[: 3 * n], fx["born_ion_f"]) < 1e-13
    np.testing.assert_allclose(S.U[0], fx["born_U_ion"][0], rtol=1e-12)


@pytest.mark.parametrize("op,bit", [("tersoff", O.ORC_ION_TERSOFF), ("sw", O.ORC_ION_SW)])
def test_manybody(op, bit):
    fx = G.load("si")
    orc, _ = G.oracle_for(fx, prefix=op + "_sys_", flags=bit)
    n = fx[op + "_sys_ion_mass"].size
    st = {k: np.zeros(3 * orc.n_atoms) for k in ("atom_x", "atom_v", "atom_f")}
    st["atom_x"][: 3 * n] = fx[op + "_atom_x"][: 3 * n]
    st.update(mol_x=np.zeros(3 * orc.n_mol), mol_v=np.zeros(3 * orc.n_mol), region=np.ones(orc.n_mol),
              atom_w=np.ones(orc.n_atoms))
    S = orc.new_state(st)
    orc.force_term(S, bit, 1)
    assert rel(S.atom_f[: 3 * n], fx[op + "_ion_f"]) < 1e-12
    # the reference scales forces by 23.06 (eV -> kcal/mol); the oracle puts
    # its energy in the same unit
    np.testing.assert_allclose(S.U[0], fx[op + "_U_raw"][0] * 23.06, rtol=1e-10)


def test_rotation():
    fx = G.load("rotation")
    rin = fx["rot_in"].reshape(-1, 6)
    out = []
    X, Y, Z = O.C.c_double(), O.C.c_double(), O.C.c_double()
    for r in rin:
        O.lib().orc_rotation(O.C.byref(X), O.C.byref(Y), O.C.byref(Z), *[float(v) for v in r])
        out += [X.value, Y.value, Z.value]
    np.testing.assert_array_equal(np.array(out), fx["rot_out"])
