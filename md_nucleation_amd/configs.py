"""Input-file writers for the workloads of this package.

The files use the reference's own formats: the single-space condition file
read by ``MD::readCondFile`` (src/initialization/setCondition.cpp:5-254) and
the tab-separated molecule files read by ``Variables::readIonFile`` /
``readVaporFile`` (src/initialization/readIonFile.cpp:4-180,
readVaporFile.cpp:4-131).  Writing inputs, rather than arrays, means the same
files feed this package's parser and the reference's.

Molecules:
  * methanol: 6 sites (C .145, O -.683, HO .418, 3 H .04), GAFF-like bonded
    terms; X-H bonds become SHAKE constraints under ``SHAKE H``;
  * water: 3 sites (O -.8476, H .4238), O-H SHAKE constraints + HOH angle;
  * Na+ (sigma 2.35 A, eps 0.13 kcal/mol), Cl- (sigma 4.4, eps 0.1);
  * a synthetic 150-atom AMBER chain (bonds 1.0-1.5 A, k_b 300-500, k_theta
    40-80, V_n 0-2 with n in {1,2,3}, charges N(0,0.3) shifted to +1,
    sigma U(2.5,3.9), eps U(0.02,0.2)), built as a self-avoiding chain.
Hydrogen sites carry a small LJ core (sigma 1.0 A, eps 0.01 kcal/mol) so
random configurations stay bounded.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

_SECTIONS = ("atom type", "bond type", "angle type", "dihedral type",
             "atoms", "bonds", "angles", "dihedrals")


def _fmt(v):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_molecule(path, atom_types, bond_types, angle_types, dihedral_types,
                   atoms, bonds, angles, dihedrals):
    """Write a reference-format molecule file (all ids 1-based)."""
    rows = {
        "atom type": [(i + 1, n, m, e, s) for i, (n, m, e, s) in enumerate(atom_types)],
        "bond type": [(i + 1, n, k, r0) for i, (n, k, r0) in enumerate(bond_types)],
        "angle type": [(i + 1, n, k, th) for i, (n, k, th) in enumerate(angle_types)],
        "dihedral type": [(i + 1, n, m, *terms) for i, (n, m, terms) in enumerate(dihedral_types)],
        "atoms": [(i + 1, t, q, x, y, z) for i, (t, q, x, y, z) in enumerate(atoms)],
        "bonds": list(bonds),
        "angles": list(angles),
        "dihedrals": list(dihedrals),
    }
    with open(path, "w") as f:
        for sec in _SECTIONS:
            f.write(sec + "\n")
            for r in rows[sec]:
                f.write("\t".join(_fmt(v) if not isinstance(v, str) else v for v in r) + "\n")
            f.write("\n")
    return path


def _centre(atoms, masses):
    xyz = np.array([[a[2], a[3], a[4]] for a in atoms], float)
    m = np.asarray(masses, float)
    c = (xyz * m[:, None]).sum(0) / m.sum()
    return [(a[0], a[1], *(np.array(a[2:5]) - c)) for a in atoms]


def write_methanol(path):
    at = [("#c3", 12.01, 0.1094, 3.39967), ("#oh", 16.00, 0.2104, 3.06647),
          ("#ho", 1.008, 0.01, 1.0), ("#h1", 1.008, 0.0157, 2.47135)]
    bt = [("c3-oh", 314.1, 1.4233), ("oh-ho", 369.6, 0.9724), ("c3-h1", 335.9, 1.0969)]
    ct = [("h1-c3-h1", 39.18, 109.55), ("h1-c3-oh", 51.03, 109.88), ("c3-oh-ho", 47.09, 108.16)]
    dt = [("h1-c3-oh-ho", 1, (0.16667, 3, 0.0, 0.0, 0, 0.0, 0.0, 0, 0.0))]
    rco, roh, rch = 1.4233, 0.9724, 1.0969
    th_coh, th_hco = math.radians(108.16), math.radians(109.88)
    C = np.zeros(3)
    O = np.array([rco, 0, 0])
    HO = O + roh * np.array([math.cos(math.pi - th_coh), math.sin(math.pi - th_coh), 0.0])
    Hs = []
    for phi in (60.0, 180.0, 300.0):
        p = math.radians(phi)
        Hs.append(rch * np.array([math.cos(th_hco), math.sin(th_hco) * math.cos(p), math.sin(th_hco) * math.sin(p)]))
    pos = [C, O, HO] + Hs
    types = [1, 2, 3, 4, 4, 4]
    q = [0.145, -0.683, 0.418, 0.04, 0.04, 0.04]
    atoms = [(types[i], q[i], *pos[i]) for i in range(6)]
    atoms = _centre(atoms, [at[t - 1][1] for t in types])
    bonds = [(1, 2, 1), (2, 3, 2), (1, 4, 3), (1, 5, 3), (1, 6, 3)]
    angles = [(4, 1, 5, 1), (4, 1, 6, 1), (5, 1, 6, 1), (4, 1, 2, 2), (5, 1, 2, 2), (6, 1, 2, 2), (1, 2, 3, 3)]
    dih = [(4, 1, 2, 3, 1), (5, 1, 2, 3, 1), (6, 1, 2, 3, 1)]
    return write_molecule(path, at, bt, ct, dt, atoms, bonds, angles, dih)


def _bonded_site(a, b, r, theta, normal):
    """a point at distance r from atom a whose bond makes angle theta
    (radians) with the direction a -> b, in the plane of b - a and normal"""
    u = (b - a) / np.linalg.norm(b - a)
    w = np.cross(normal, u)
    w /= np.linalg.norm(w)
    return a + r * (math.cos(theta) * u + math.sin(theta) * w)


def write_glycol(path):
    """Ethylene glycol HO-CH2-CH2-OH with GAFF-like parameters (methanol's
    types plus c3-c3): under 'SHAKE H' its six X-H bonds fall into four
    groups that share no atom (O-H, the two C-H of each carbon, O-H)."""
    at = [("#c3", 12.01, 0.1094, 3.39967), ("#oh", 16.00, 0.2104, 3.06647),
          ("#ho", 1.008, 0.01, 1.0), ("#h1", 1.008, 0.0157, 2.47135)]
    bt = [("c3-oh", 314.1, 1.4233), ("oh-ho", 369.6, 0.9724), ("c3-h1", 335.9, 1.0969), ("c3-c3", 303.1, 1.5375)]
    ct = [("h1-c3-h1", 39.18, 109.55), ("h1-c3-oh", 51.03, 109.88), ("c3-oh-ho", 47.09, 108.16),
          ("c3-c3-oh", 67.7, 109.43), ("h1-c3-c3", 46.37, 110.07)]
    tet = math.radians(109.47)
    z, y = np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])
    C1, C2 = np.zeros(3), np.array([1.5375, 0.0, 0.0])
    O1 = _bonded_site(C1, C2, 1.4233, tet, z)
    O2 = _bonded_site(C2, C1, 1.4233, tet, z)
    HO1 = _bonded_site(O1, C1, 0.9724, math.radians(108.16), y)
    HO2 = _bonded_site(O2, C2, 0.9724, math.radians(108.16), y)
    def hydrogens(c, n1, n2):
        # the two tetrahedral sites of carbon c opposite its bonds to n1, n2
        u1, u2 = (n1 - c) / np.linalg.norm(n1 - c), (n2 - c) / np.linalg.norm(n2 - c)
        bis = -(u1 + u2) / np.linalg.norm(u1 + u2)
        out = np.cross(u1, u2) / np.linalg.norm(np.cross(u1, u2))
        h = math.radians(54.735)
        return [c + 1.0969 * (math.cos(h) * bis + sg * math.sin(h) * out) for sg in (1.0, -1.0)]
    H1a, H1b = hydrogens(C1, C2, O1)
    H2a, H2b = hydrogens(C2, C1, O2)
    pos = [C1, C2, O1, HO1, O2, HO2, H1a, H1b, H2a, H2b]
    types = [1, 1, 2, 3, 2, 3, 4, 4, 4, 4]
    q = [0.185, 0.185, -0.683, 0.418, -0.683, 0.418, 0.04, 0.04, 0.04, 0.04]
    atoms = _centre([(types[i], q[i], *pos[i]) for i in range(10)], [at[t - 1][1] for t in types])
    bonds = [(1, 2, 4), (1, 3, 1), (3, 4, 2), (2, 5, 1), (5, 6, 2), (1, 7, 3), (1, 8, 3), (2, 9, 3), (2, 10, 3)]
    angles = [(7, 1, 8, 1), (3, 1, 7, 2), (3, 1, 8, 2), (1, 3, 4, 3), (2, 5, 6, 3), (2, 1, 3, 4), (1, 2, 5, 4),
              (7, 1, 2, 5), (8, 1, 2, 5), (9, 2, 10, 1), (5, 2, 9, 2), (5, 2, 10, 2), (9, 2, 1, 5), (10, 2, 1, 5)]
    return write_molecule(path, at, bt, ct, [], atoms, bonds, angles, [])


def write_water(path):
    at = [("#ow", 15.999, 0.1553, 3.166), ("#hw", 1.008, 0.01, 1.0)]
    bt = [("ow-hw", 553.0, 1.0)]
    ct = [("hw-ow-hw", 100.0, 109.47)]
    th = math.radians(109.47) / 2
    pos = [np.zeros(3), np.array([math.sin(th), math.cos(th), 0.0]), np.array([-math.sin(th), math.cos(th), 0.0])]
    atoms = [(1, -0.8476, *pos[0]), (2, 0.4238, *pos[1]), (2, 0.4238, *pos[2])]
    atoms = _centre(atoms, [15.999, 1.008, 1.008])
    return write_molecule(path, at, bt, ct, [], atoms, [(1, 2, 1), (1, 3, 1)], [(2, 1, 3, 1)], [])


def write_single_ion(path, name, mass, eps, sigma, charge):
    return write_molecule(path, [(name, mass, eps, sigma)], [], [], [], [(1, charge, 0.0, 0.0, 0.0)], [], [], [])


def write_na(path):
    return write_single_ion(path, "Na", 22.99, 0.13, 2.35, 1.0)


def write_cl(path):
    return write_single_ion(path, "Cl", 35.45, 0.1, 4.4, -1.0)


def write_peptide(path, seed=1, n=150):
    """Synthetic AMBER chain of ``n`` atoms, one atom type per atom."""
    rng = np.random.default_rng(seed)
    sig = rng.uniform(2.5, 3.9, n)
    eps = rng.uniform(0.02, 0.2, n)
    mass = rng.choice([12.01, 14.01, 16.00], n)
    q = rng.normal(0.0, 0.3, n)
    q += (1.0 - q.sum()) / n
    r0 = rng.uniform(1.0, 1.5, n - 1)
    th0 = rng.uniform(100.0, 120.0, n - 2)
    pos = np.zeros((n, 3))
    pos[1] = [r0[0], 0, 0]
    a = math.radians(th0[0])
    pos[2] = pos[1] + r0[1] * np.array([-math.cos(a), math.sin(a), 0.0])
    for i in range(3, n):
        best, best_d = None, -1.0
        for _ in range(400):
            phi = rng.uniform(-math.pi, math.pi)
            # natural extension reference frame (NeRF) placement
            b, c, d = pos[i - 3], pos[i - 2], pos[i - 1]
            bc = c - b
            cd = d - c
            cd /= np.linalg.norm(cd)
            nrm = np.cross(bc, cd)
            nrm /= np.linalg.norm(nrm)
            m = np.stack([cd, np.cross(nrm, cd), nrm], 1)
            t = math.radians(th0[i - 2])
            loc = r0[i - 1] * np.array([-math.cos(t), math.sin(t) * math.cos(phi), math.sin(t) * math.sin(phi)])
            cand = d + m @ loc
            dmin = np.min(np.linalg.norm(pos[: i - 3] - cand, axis=1)) if i > 3 else 9.0
            if dmin > best_d:
                best, best_d = cand, dmin
            if dmin >= 3.2:
                break
        pos[i] = best
    pos -= (pos * mass[:, None]).sum(0) / mass.sum()
    at = [("#c3", mass[i], eps[i], sig[i]) for i in range(n)]
    kb = rng.uniform(300, 500, n - 1)
    bt = [("b%d" % i, kb[i], r0[i]) for i in range(n - 1)]
    kt = rng.uniform(40, 80, n - 2)
    # equilibrium angle = built angle, so the chain starts near its minimum
    ct = []
    for i in range(n - 2):
        v1, v2 = pos[i] - pos[i + 1], pos[i + 2] - pos[i + 1]
        ang = math.degrees(math.acos(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)))
        ct.append(("a%d" % i, kt[i], ang))
    vn = rng.uniform(0.0, 2.0, n - 3)
    nn = rng.integers(1, 4, n - 3)
    dd = rng.choice([0.0, 180.0], n - 3)
    dt = [("d%d" % i, 1, (vn[i], int(nn[i]), dd[i], 0.0, 0, 0.0, 0.0, 0, 0.0)) for i in range(n - 3)]
    atoms = [(i + 1, q[i], *pos[i]) for i in range(n)]
    bonds = [(i + 1, i + 2, i + 1) for i in range(n - 1)]
    angles = [(i + 1, i + 2, i + 3, i + 1) for i in range(n - 2)]
    dih = [(i + 1, i + 2, i + 3, i + 4, i + 1) for i in range(n - 3)]
    return write_molecule(path, at, bt, ct, dt, atoms, bonds, angles, dih)


def write_nacl_cluster(path, n_side=2, a=2.82):
    """Rock-salt NaCl cube (ion file: type 1 Na, type 2 Cl) for BMH."""
    at = [("Na", 22.99, 0.13, 2.35), ("Cl", 35.45, 0.1, 4.4)]
    atoms = []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                t = 1 if (i + j + k) % 2 == 0 else 2
                atoms.append((t, 1.0 if t == 1 else -1.0, i * a, j * a, k * a))
    atoms = _centre(atoms, [22.99 if t[0] == 1 else 35.45 for t in atoms])
    return write_molecule(path, at, [], [], [], atoms, [], [], [])


def write_si_cluster(path, n_side=2, a=5.431):
    """Diamond-lattice Si cluster (Tersoff / Stillinger-Weber ion)."""
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    pts = []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                for b in basis:
                    pts.append((np.array([i, j, k]) + b) * a)
    pts = np.array(pts)
    pts += np.random.default_rng(7).normal(0, 0.05, pts.shape)
    atoms = [(1, 0.0, *p) for p in pts]
    atoms = _centre(atoms, [28.0855] * len(atoms))
    return write_molecule(path, [("Si", 28.0855, 0.0, 0.0)], [], [], [], atoms, [], [], [])


@dataclass
class Condition:
    ion_file: str
    vapor_file: str | None
    n_vap: int
    gas: str | None = "N2"
    n_gas: int = 1000
    T: float = 300.0
    pressure: float = 1e5
    dt: float = 1.0
    total_steps: int = 1000
    relax_steps: int = 0
    cutoff: float | None = None
    margin: float | None = None
    shake_h: bool = True
    interactions: dict = field(default_factory=lambda: {"gg": "OFF", "gi": "LJ", "gv": "LJ", "vi": "LJcoul",
                                                        "vv": "LJcoul"})
    ion_potential: str | None = None
    efield: tuple | None = None
    cluster: tuple | None = None  # (radius A, density g/cm3)
    nvt_ion: str = "OFF"  # OFF, scale, or a Nose-Hoover temperature
    observe: int = 100000000  # Output ... OBSERVE cadence

    def write(self, path):
        lines = ["Input " + self.ion_file]
        if self.vapor_file:
            lines.append("Vapor %s %d" % (self.vapor_file, self.n_vap))
        if self.gas:
            lines.append("Gas %s %d" % (self.gas, self.n_gas))
        lines += ["Temperature %g" % self.T, "Pressure %g" % self.pressure, "dt %g" % self.dt,
                  "TotalSteps %d" % self.total_steps, "RelaxSteps %d" % self.relax_steps]
        if self.cutoff is not None:
            lines.append("CutOff %g" % self.cutoff)
        if self.margin is not None:
            lines.append("Margin %g" % self.margin)
        lines.append("Output out %d" % self.observe)
        lines.append("NVTion %s" % self.nvt_ion)
        if self.shake_h:
            lines.append("SHAKE H")
        if self.cluster:
            lines.append("VaporCluster %g %g" % self.cluster)
        lines.append("Interactions")
        for k, v in self.interactions.items():
            lines.append(" %s %s" % (k, v))
        if self.ion_potential:
            lines.append(" ion %s" % self.ion_potential)
        if self.efield:
            lines.append(" Efield %g %g %g" % tuple(self.efield))
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        return path


def molecule_files(d):
    """Write the standard molecule files into directory ``d``."""
    os.makedirs(d, exist_ok=True)
    return {
        "na": write_na(os.path.join(d, "na.txt")),
        "cl": write_cl(os.path.join(d, "cl.txt")),
        "methanol": write_methanol(os.path.join(d, "methanol.txt")),
        "water": write_water(os.path.join(d, "water.txt")),
        "glycol": write_glycol(os.path.join(d, "glycol.txt")),
    }


def config1(d, n_gas=10000, n_vap=20):
    """Na+ + 20 MeOH + 10,000 N2, dt 1 fs, gg OFF / gi LJ / gv LJ / vi, vv LJcoul."""
    m = molecule_files(d)
    return Condition(m["na"], m["methanol"], n_vap, "N2", n_gas, total_steps=1000).write(os.path.join(d, "c1.cond"))


def config2(d, n_gas=100000, n_vap=200, seed=1):
    """150-atom AMBER chain + 200 water + 100,000 N2, 10,000 steps."""
    m = molecule_files(d)
    pep = write_peptide(os.path.join(d, "peptide_%d.txt" % seed), seed)
    return Condition(pep, m["water"], n_vap, "N2", n_gas, total_steps=10000,
                     ion_potential="AMBER").write(os.path.join(d, "c2_%d.cond" % seed))


def config3(d, n_gas=100000, n_vap=2000, density=0.79, radius=45.0):
    """Na+ + 2,000 MeOH packed around it + 100,000 Ar, all-pairs vv/vi."""
    m = molecule_files(d)
    return Condition(m["na"], m["methanol"], n_vap, "Ar", n_gas, total_steps=1000,
                     cluster=(radius, density)).write(os.path.join(d, "c3.cond"))


def config5(d, n_gas=100000, n_vap=200, seed=1, radius=40.0, density=1.0):
    """Force-matching system: the 150-atom AMBER chain with 200 water packed
    around it (lattice sites nearest the chain at ``density`` g/cm3, all in
    the effective domain, 750 interacting ion + vapor atoms) + N2."""
    m = molecule_files(d)
    pep = write_peptide(os.path.join(d, "peptide_%d.txt" % seed), seed)
    return Condition(pep, m["water"], n_vap, "N2", n_gas, total_steps=10000, ion_potential="AMBER",
                     cluster=(radius, density)).write(os.path.join(d, "c5_%d.cond" % seed))


def segment_condition(d, n_gas=200, seed=1, radius=40.0, density=1.0):
    """The force-matching system for trajectory-segment gradients: the
    config-5 chain and 200 packed water, integrated as MD::verlet ships it
    (no SHAKE, so no 'SHAKE H' line: the water's O-H bonds keep their
    harmonic terms); the gas does not interact with ion or vapor."""
    m = molecule_files(d)
    pep = write_peptide(os.path.join(d, "peptide_%d.txt" % seed), seed)
    return Condition(pep, m["water"], 200, "N2", n_gas, total_steps=100, ion_potential="AMBER",
                     cluster=(radius, density), shake_h=False,
                     interactions={"gg": "OFF", "gi": "OFF", "gv": "OFF", "vi": "LJcoul", "vv": "LJcoul"}).write(
        os.path.join(d, "seg_%d.cond" % seed))


def config4_replica(d, r, rng, n_gas=5000):
    """One replica of the batched workload: ion in {Na+, Cl-, chain},
    vapor U[10,400] methanol or water, 5,000 N2."""
    m = molecule_files(d)
    kind = ("na", "cl", "pep")[int(rng.integers(0, 3))]
    if kind == "pep":
        ion = os.path.join(d, "peptide_1.txt")
        if not os.path.exists(ion):
            write_peptide(ion, 1)
        pot = "AMBER"
    else:
        ion, pot = m[kind], None
    vap = m["methanol"] if rng.integers(0, 2) == 0 else m["water"]
    n_vap = int(rng.integers(10, 401))
    return Condition(ion, vap, n_vap, "N2", n_gas, total_steps=10000, ion_potential=pot).write(
        os.path.join(d, "c4_%d.cond" % r)), kind
