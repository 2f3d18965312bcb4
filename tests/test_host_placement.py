"""Host initialization (libmdn's placement, host C++, no GPU): the invariants
of the reference's initialization that the placement must keep, and its
failure mode.

* initialization_gas / initialization_vapor (initializationGas.cpp:24-90,
  initializationVapor.cpp:14-77): molecule centres uniform in the periodic
  box, every pair at least 10 A apart (periodic minimum image,
  initializationGas.cpp:25, 59-70), also from the ion's atoms and the
  origin; Maxwell centre velocities.  Where no place is left the reference
  retries without end; here that is an error (MDN_E_ARG).
* the pre-packed vapor cluster (condition keyword VaporCluster): molecules
  on the lattice sites nearest the ion at the given mass density, each a
  rigid copy of the template, whatever its atom count.
* the synthetic AMBER chain of the larger-ion workloads: the recipe's ranges.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from md_nucleation_amd import _lib as M
from md_nucleation_amd import configs

KB = 1.38e-23


def _mol_atoms(sd, st, mol):
    v = sd.view
    na = [s["mass"].size for s in v["species"]]
    if mol == 0:
        o, n = 0, na[0]
    elif mol <= v["n_gas"]:
        o, n = na[0] + (mol - 1) * na[1], na[1]
    else:
        o, n = na[0] + v["n_gas"] * na[1] + (mol - 1 - v["n_gas"]) * na[2], na[2]
    return st.atom_x.reshape(-1, 3)[o:o + n]


def test_gas_and_vapor_centres_keep_10_A(tmp_path):
    m = configs.molecule_files(str(tmp_path))
    sd = M.SysDef(configs.Condition(m["cl"], m["methanol"], 40, "N2", 3000, pressure=1e5).write(
        str(tmp_path / "g.cond")))
    v = sd.view
    st = sd.init_state(9)
    L = v["d_size"]
    assert L == pytest.approx((3000 * KB * 300 / 1e5) ** (1 / 3) * 1e10, rel=1e-12)
    X = st.mol_x.reshape(-1, 3)[1:]
    assert X.shape[0] == 3040
    assert np.all(np.abs(X) <= L / 2)
    pts = np.vstack([X, _mol_atoms(sd, st, 0), np.zeros((1, 3))])
    tree = cKDTree(np.mod(pts + L / 2, L), boxsize=L)
    # pairs among the ion's atoms and the origin (a one-atom ion sits on it)
    # are not placements
    pairs = [(i, j) for i, j in tree.query_pairs(10.0) if i < X.shape[0] or j < X.shape[0]]
    assert not pairs, sorted(pairs)[:5]
    # uniform in the box: each octant holds about an eighth
    octs = np.bincount(((X > 0) * [1, 2, 4]).sum(1), minlength=8)
    assert np.all(np.abs(octs - 380) < 5 * np.sqrt(380))
    # Maxwell centre velocities, sqrt(kB T / m) * 1e-5 A/fs per component
    Vg = st.mol_v.reshape(-1, 3)[1:1 + v["n_gas"]]
    sg = np.sqrt(KB * 300 / v["m_gas_kg"]) * 1e-5
    assert abs(Vg.std() / sg - 1) < 5 / np.sqrt(2 * Vg.size)


def test_crowded_box_is_an_error(tmp_path):
    """2,000 molecules in a 43.6 A box cannot keep 10 A apart: an error, not
    an endless retry and not silently overlapping molecules."""
    m = configs.molecule_files(str(tmp_path))
    sd = M.SysDef(configs.Condition(m["na"], m["methanol"], 1, "N2", 2000, pressure=1e8).write(
        str(tmp_path / "c.cond")))
    assert sd.view["d_size"] < 50
    with pytest.raises(M.MdnError, match="no place"):
        sd.init_state(1)


def test_packed_cluster_density_and_rigid_molecules(tmp_path):
    m = configs.molecule_files(str(tmp_path))
    n_vap, rho = 600, 0.79
    sd = M.SysDef(configs.Condition(m["na"], m["methanol"], n_vap, "Ar", 2000, cluster=(45.0, rho)).write(
        str(tmp_path / "k.cond")))
    v = sd.view
    st = sd.init_state(4)
    first = 1 + v["n_gas"]
    C = st.mol_x.reshape(-1, 3)[first:]
    assert np.all(st.region[first:] == 1)
    # lattice spacing of the density: one molecule per a^3
    vol = v["species"][2]["molmass"] / rho * 1.66054
    r = np.linalg.norm(C, axis=1)
    # the occupied sphere holds the molecules at the lattice's density
    r_occ = (3 * n_vap * vol / (4 * np.pi)) ** (1 / 3)
    assert r.max() < r_occ + 2 * vol ** (1 / 3)
    inner = (r < 0.8 * r_occ).sum() / (4 / 3 * np.pi * (0.8 * r_occ) ** 3)
    assert inner == pytest.approx(1 / vol, rel=0.1)
    # each molecule is the template rotated rigidly about its site
    tm = v["species"][2]["tmpl_x"].reshape(-1, 3)
    dref = np.linalg.norm(tm[:, None] - tm[None], axis=2)
    for k in range(0, n_vap, 37):
        A = _mol_atoms(sd, st, first + k)
        np.testing.assert_allclose(np.linalg.norm(A[:, None] - A[None], axis=2), dref, atol=1e-12)
        np.testing.assert_allclose((A * v["species"][2]["mass"][:, None]).sum(0) / v["species"][2]["mass"].sum(),
                                   C[k], atol=1e-9)
    # the gas keeps 10 A from the cluster's sites
    G = st.mol_x.reshape(-1, 3)[1:first]
    d = np.linalg.norm(G[:, None, :] - C[None, ::10, :], axis=2)
    assert d.min() > 10.0


def test_cluster_of_large_molecules_places_every_atom(tmp_path):
    """A 90-atom vapor molecule (a chain) in the packed cluster: every atom of
    every molecule placed, geometry kept."""
    d = str(tmp_path)
    m = configs.molecule_files(d)
    chain = configs.write_peptide(d + "/chain90.txt", seed=2, n=90)
    sd = M.SysDef(configs.Condition(m["na"], chain, 6, "Ar", 200, cluster=(60.0, 0.5), shake_h=False).write(
        d + "/l.cond"))
    v = sd.view
    assert v["species"][2]["mass"].size == 90
    st = sd.init_state(2)
    tm = v["species"][2]["tmpl_x"].reshape(-1, 3)
    dref = np.linalg.norm(tm[:, None] - tm[None], axis=2)
    for k in range(6):
        A = _mol_atoms(sd, st, 1 + v["n_gas"] + k)
        np.testing.assert_allclose(np.linalg.norm(A[:, None] - A[None], axis=2), dref, atol=1e-11)


def test_peptide_recipe_ranges(tmp_path):
    """The synthetic 150-atom chain: bonds between consecutive atoms at
    1.0-1.5 A with k_b 300-500, angles over consecutive triples with k_theta
    40-80, dihedrals over consecutive quads with V_n 0-2 and n in {1,2,3},
    charges summing to +1, sigma 2.5-3.9 A, eps 0.02-0.2 kcal/mol."""
    d = str(tmp_path)
    m = configs.molecule_files(d)
    pep = configs.write_peptide(d + "/p.txt", seed=1)
    sd = M.SysDef(configs.Condition(pep, m["water"], 10, "N2", 100, ion_potential="AMBER").write(d + "/p.cond"))
    v = sd.view
    I = v["species"][0]
    n = I["mass"].size
    assert n == 150
    b = I["bond"].reshape(-1, 3)
    np.testing.assert_array_equal(b[:, :2], np.stack([np.arange(n - 1), np.arange(1, n)], 1))
    bt = v["btype"].reshape(-1, 2)[b[:, 2]]
    assert np.all((bt[:, 0] >= 300) & (bt[:, 0] <= 500)) and np.all((bt[:, 1] >= 1.0) & (bt[:, 1] <= 1.5))
    # the built chain sits at its bond lengths
    X = I["tmpl_x"].reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(X[1:] - X[:-1], axis=1), bt[:, 1], rtol=1e-12)
    a = I["angle"].reshape(-1, 4)
    assert a.shape[0] == n - 2 and np.all(a[:, 1] == a[:, 0] + 1) and np.all(a[:, 2] == a[:, 0] + 2)
    ct = v["ctype"].reshape(-1, 2)[a[:, 3]]
    assert np.all((ct[:, 0] >= 40) & (ct[:, 0] <= 80))
    dh = I["dih"].reshape(-1, 5)
    assert dh.shape[0] == n - 3
    dc = v["dcoeff"].reshape(-1, 15)[dh[:, 4]]
    assert np.all(v["dmulti"][dh[:, 4]] == 1)
    assert np.all((dc[:, 0] >= 0) & (dc[:, 0] <= 2)) and set(np.unique(dc[:, 1])) <= {1.0, 2.0, 3.0}
    assert abs(I["charge"].sum() - 1.0) < 1e-9
    t = I["type"]
    sig, eps = v["type_sigma"][t], v["type_eps"][t]
    assert np.all((sig >= 2.5) & (sig <= 3.9)) and np.all((eps >= 0.02) & (eps <= 0.2))
