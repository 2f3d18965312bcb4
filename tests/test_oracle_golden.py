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


# ---- ion thermostat of the relaxation phase (nvt.cpp) ----------------------
NVT = ["nvt_scale", "nvt_nh"]


@pytest.mark.parametrize("name", NVT)
def test_thermostat_trajectory(name):
    """MD::verlet with NVTion scale / NVTion 250: the reference's own 30-step
    trajectory (positions, velocities, forces, energies, zeta_ion)."""
    fx = G.sub(G.load("nvt"), name + "_")
    orc, sysd = G.oracle_for(fx)
    assert sysd["nvt_ion"] == (1 if name == "nvt_scale" else 2)
    S = orc.new_state(G.state(fx, loopPair=0, loop_update=10 ** 8))
    G.set_lists(S, fx, prefix="s_")
    nsteps, every = (int(v) for v in fx["s_nsteps"])
    rc, _ = orc.verlet(S, nsteps, eflag_every=every)
    assert rc == 0
    for k in ("atom_x", "atom_v", "atom_f", "atom_w", "mol_x", "mol_v"):
        assert rel(getattr(S, k), fx["s_" + k]) < 1e-12, k
    np.testing.assert_allclose(S.U, fx["s_U"], rtol=1e-10, atol=1e-10)
    assert abs(S.c.zeta_ion - fx["s_zeta_ion"][0]) <= 1e-12 * max(abs(fx["s_zeta_ion"][0]), 1e-6)
    # the thermostat changed the trajectory: NVE from the same state differs
    orc0, _ = G.oracle_for(fx, nvt_ion=0)
    S0 = orc0.new_state(G.state(fx, loopPair=0, loop_update=10 ** 8))
    G.set_lists(S0, fx, prefix="s_")
    orc0.verlet(S0, nsteps, eflag_every=every)
    assert rel(S0.atom_v, fx["s_atom_v"]) > 1e-6


def _kinetic_T(sysd, v):
    """Equipartition temperature of the ion from its atom velocities:
    K = 1/2 sum m v^2 (g/mol A^2/fs^2 -> kcal/mol x 1e4/4.184),
    T = 2 K / (3 N kB) with kB = 1.38e-23 * 6.02e23 / 4184 kcal/mol/K."""
    m = np.asarray(sysd["species"][0]["mass"], np.float64)
    n = m.size
    K = 0.5 * np.sum(m[:, None] * v[: 3 * n].reshape(-1, 3) ** 2) * 1e4 / 4.184
    return 2 * K / (3 * n * (1.38e-23 * 6.02e23 / 4184.0))


def test_thermostat_pieces_closed_forms():
    """Tion is the equipartition temperature; velocity scaling lands on
    exactly 300 K; the Nose-Hoover zeta step is (Tion - Tnh) * 3N * kB *
    1e-4 * dt and the velocity factor exp(-zeta dt / 2)."""
    fx = G.sub(G.load("nvt"), "nvt_nh_")
    orc, sysd = G.oracle_for(fx)
    S = orc.new_state(G.state(fx))
    v0 = S.atom_v.copy()
    T0 = orc.ion_temperature(S)
    assert abs(T0 - _kinetic_T(sysd, v0)) < 1e-10 * T0
    assert 50 < T0 < 2000
    n = sysd["species"][0]["mass"].size
    O.lib().orc_velocity_scaling(orc.h, O.C.byref(S.c))
    assert abs(orc.ion_temperature(S) - 300.0) < 1e-9
    ratio = S.atom_v[: 3 * n] / v0[: 3 * n]
    np.testing.assert_allclose(ratio, np.sqrt(300.0 / T0), rtol=1e-13)
    np.testing.assert_array_equal(S.atom_v[3 * n:], v0[3 * n:])  # only the ion
    S.atom_v[:] = v0
    S.c.zeta_ion = 0.01
    O.lib().orc_nosehoover_zeta(orc.h, O.C.byref(S.c))
    kB = 1.38e-23 * 6.02e23 / 4184.0
    want = 0.01 + (T0 - sysd["Tnh_ion"]) * 3 * n * kB * 1e-4 * sysd["dt"]
    assert abs(S.c.zeta_ion - want) < 1e-15
    O.lib().orc_nosehoover_ion(orc.h, O.C.byref(S.c))
    np.testing.assert_allclose(S.atom_v[: 3 * n], v0[: 3 * n] * np.exp(-want * 0.5 * sysd["dt"]), rtol=1e-14)
    np.testing.assert_array_equal(S.atom_v[3 * n:], v0[3 * n:])
