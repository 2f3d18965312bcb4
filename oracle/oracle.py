"""ctypes front end of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / CPU baseline.  The product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

ORC_GI, ORC_VI, ORC_GG, ORC_GV, ORC_VV = 1, 2, 4, 8, 16
ORC_ION_AMBER, ORC_ION_BMH, ORC_ION_TERSOFF, ORC_ION_SW = 32, 64, 128, 256
ORC_EFIELD, ORC_VAPOR_INTRA, ORC_GAS_INTRA = 512, 1024, 2048
FLAG_ORDER = [ORC_GI, ORC_VI, ORC_GG, ORC_GV, ORC_VV, ORC_ION_AMBER, ORC_ION_BMH, ORC_ION_TERSOFF,
              ORC_ION_SW, ORC_EFIELD, ORC_VAPOR_INTRA, ORC_GAS_INTRA]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSystem(C.Structure):
    _fields_ = [
        ("n_types", C.c_int), ("pair_coeff", _dp),
        ("n_btypes", C.c_int), ("btype", _dp),
        ("n_ctypes", C.c_int), ("ctype", _dp),
        ("n_dtypes", C.c_int), ("dmulti", _ip), ("dcoeff", _dp),
        ("natoms", C.c_int * 3), ("mass", _dp * 3), ("charge", _dp * 3), ("type", _ip * 3),
        ("tmpl_x", _dp * 3),
        ("nbond", C.c_int * 3), ("bond", _ip * 3),
        ("nangle", C.c_int * 3), ("angle", _ip * 3),
        ("ndih", C.c_int * 3), ("dih", _ip * 3),
        ("nrigid", C.c_int * 3), ("rigid", _ip * 3),
        ("molmass", C.c_double * 3),
        ("n_gas", C.c_int), ("n_vap", C.c_int), ("ion_type_base", C.c_int),
        ("dt", C.c_double), ("T", C.c_double), ("d_size", C.c_double), ("margin", C.c_double),
        ("ML2", C.c_double), ("m_gas_kg", C.c_double), ("m_vap_kg", C.c_double),
        ("gas_spin", C.c_int), ("flags", C.c_uint32), ("efield", C.c_double * 3),
        ("shake", C.c_int), ("shake_max_iter", C.c_int),
        ("seed", C.c_uint64), ("replica", C.c_uint32), ("nvt_ion", C.c_int), ("Tnh_ion", C.c_double),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("mol_x", _dp), ("mol_v", _dp), ("region", C.POINTER(C.c_uint8)),
        ("atom_x", _dp), ("atom_v", _dp), ("atom_f", _dp), ("atom_w", _dp),
        ("loopPair", C.c_int), ("loop_update", C.c_int), ("rebuild_count", C.c_int64),
        ("cap_gi", C.c_int), ("cap_gv", C.c_int), ("cap_gg", C.c_int),
        ("n_gi", C.c_int), ("n_gv", C.c_int), ("n_gg", C.c_int),
        ("gi", _ip), ("gv", _ip), ("gg", _ip),
        ("U", C.c_double * 8), ("shake_iters_max", C.c_int), ("shake_failures", C.c_int64), ("status", C.c_int),
        ("zeta_ion", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", HERE])
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcSystem)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for fn in ("orc_make_pair", "orc_check_pairlist"):
            getattr(L, fn).argtypes = [C.c_void_p, C.POINTER(OrcState)]
            getattr(L, fn).restype = C.c_int
        for fn in ("orc_velocity_calculation", "orc_update_position", "orc_update_position_shake",
                   "orc_ion_com_weights", "orc_velocity_scaling", "orc_nosehoover_zeta", "orc_nosehoover_ion"):
            getattr(L, fn).argtypes = [C.c_void_p, C.POINTER(OrcState)]
            getattr(L, fn).restype = None
        L.orc_forces.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_int]
        L.orc_force_term.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_uint32, C.c_int]
        L.orc_verlet.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_int, C.c_int, C.c_int64,
                                 C.POINTER(C.c_int64)]
        L.orc_verlet_batch.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.POINTER(OrcState)), C.c_int, C.c_int,
                                       C.c_int, C.c_int64, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_double)]
        L.orc_shake_molecule.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_shake_molecule.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_cross_potentials.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_rotation.argtypes = [_dp, _dp, _dp] + [C.c_double] * 6
        L.orc_weight.argtypes = [C.c_double]
        L.orc_weight.restype = C.c_double
        L.orc_ion_temperature.argtypes = [C.c_void_p, C.POINTER(OrcState)]
        L.orc_ion_temperature.restype = C.c_double
        L.orc_n_long.argtypes = [C.c_void_p]
        L.orc_n_long.restype = C.c_int
        L.orc_rng_normal2.argtypes = [C.POINTER(OrcSystem), C.c_uint32, C.c_int64, C.c_uint32, _dp]
        L.orc_rng_uniform2.argtypes = [C.POINTER(OrcSystem), C.c_uint32, C.c_int64, C.c_uint32, _dp]
        L.orc_rng_many.argtypes = [C.POINTER(OrcSystem), C.c_uint32, C.c_int, C.c_int64, C.c_uint32, C.c_int, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(np.asarray(a).astype(np.int32))


class Oracle:
    """One replica's system for the oracle.

    ``sysd`` keys: pair_coeff, btype, ctype, dmulti, dcoeff, species list of
    3 dicts (mass, charge, type, tmpl_x, bond, angle, dih, rigid, molmass),
    n_gas, n_vap, ion_type_base, dt, T, d_size, margin, ML2, m_gas_kg,
    m_vap_kg, gas_spin, flags, efield, shake, shake_max_iter, seed, replica.
    """

    def __init__(self, sysd):
        self.sysd = sysd
        self._keep = []
        s = OrcSystem()
        pc = _d(sysd["pair_coeff"])
        s.n_types = int(round(np.sqrt(pc.size / 2)))
        s.pair_coeff = self._p(pc)
        bt, ct, dc = _d(sysd["btype"]), _d(sysd["ctype"]), _d(sysd["dcoeff"])
        s.n_btypes, s.btype = bt.size // 2, self._p(bt)
        s.n_ctypes, s.ctype = ct.size // 2, self._p(ct)
        dm = _i(sysd["dmulti"])
        s.n_dtypes, s.dmulti, s.dcoeff = dm.size, self._pi(dm), self._p(dc)
        self.natoms = []
        for k, sp in enumerate(sysd["species"]):
            m = _d(sp["mass"])
            s.natoms[k] = m.size
            self.natoms.append(m.size)
            s.mass[k] = self._p(m)
            s.charge[k] = self._p(_d(sp["charge"]))
            s.type[k] = self._pi(_i(sp["type"]))
            s.tmpl_x[k] = self._p(_d(sp["tmpl_x"]))
            for nm, w in (("bond", 3), ("angle", 4), ("dih", 5), ("rigid", 3)):
                arr = _i(np.asarray(sp[nm]).reshape(-1))
                getattr(s, "n" + nm)[k] = arr.size // w
                getattr(s, nm)[k] = self._pi(arr)
            s.molmass[k] = float(sp["molmass"])
        s.n_gas, s.n_vap = int(sysd["n_gas"]), int(sysd["n_vap"])
        s.ion_type_base = int(sysd.get("ion_type_base", 0))
        for k in ("dt", "T", "d_size", "margin", "ML2", "m_gas_kg", "m_vap_kg"):
            setattr(s, k, float(sysd[k]))
        s.gas_spin = int(sysd["gas_spin"])
        s.flags = int(sysd["flags"])
        for k in range(3):
            s.efield[k] = float(sysd.get("efield", (0, 0, 0))[k])
        s.shake = int(sysd.get("shake", 0))
        s.shake_max_iter = int(sysd.get("shake_max_iter", 1000))
        s.seed = int(sysd.get("seed", 1))
        s.replica = int(sysd.get("replica", 0))
        s.nvt_ion = int(sysd.get("nvt_ion", 0))
        s.Tnh_ion = float(sysd.get("Tnh_ion", 0.0))
        self.s = s
        self.h = lib().orc_create(C.byref(s))
        self.n_mol = 1 + s.n_gas + s.n_vap
        self.n_atoms = self.natoms[0] + s.n_gas * self.natoms[1] + s.n_vap * self.natoms[2]

    def _p(self, a):
        self._keep.append(a)
        return a.ctypes.data_as(_dp)

    def _pi(self, a):
        self._keep.append(a)
        return a.ctypes.data_as(_ip)

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def new_state(self, st=None, cap=200000):
        return OracleState(self, st, cap)

    def make_pair(self, S):
        rc = lib().orc_make_pair(self.h, C.byref(S.c))
        S.sync()
        return rc

    def forces(self, S, eflag=1, zero=True):
        if zero:
            S.atom_f[:] = 0
            S.c.U[:] = [0] * 8
        lib().orc_forces(self.h, C.byref(S.c), eflag)
        S.sync()

    def force_term(self, S, bit, eflag=1):
        lib().orc_force_term(self.h, C.byref(S.c), bit, eflag)
        S.sync()

    def verlet(self, S, nsteps, eflag_every=0, step0=0):
        a = C.c_int64(0)
        rc = lib().orc_verlet(self.h, C.byref(S.c), nsteps, eflag_every, step0, C.byref(a))
        S.sync()
        return rc, a.value

    def rng(self, n, rebuild=0, purpose=0, normal=True, mol0=0):
        """n draws (pairs) of the replica's stream for molecules mol0.. of one
        rebuild and purpose: array [n, 2]"""
        out = np.zeros(2 * n)
        lib().orc_rng_many(C.byref(self.s), mol0, n, rebuild, purpose, 1 if normal else 0, out.ctypes.data_as(_dp))
        return out.reshape(n, 2)

    def ion_temperature(self, S):
        return lib().orc_ion_temperature(self.h, C.byref(S.c))

    def velocity_calculation(self, S):
        lib().orc_velocity_calculation(self.h, C.byref(S.c))

    def update_position(self, S):
        if self.s.shake:
            lib().orc_update_position_shake(self.h, C.byref(S.c))
        else:
            lib().orc_update_position(self.h, C.byref(S.c))
        S.sync()


class OracleState:
    """numpy-backed orc_state."""

    def __init__(self, orc, st=None, cap=200000):
        nm, na = orc.n_mol, orc.n_atoms
        st = st or {}
        self.mol_x = _d(st.get("mol_x", np.zeros(3 * nm))).copy().reshape(-1)
        self.mol_v = _d(st.get("mol_v", np.zeros(3 * nm))).copy().reshape(-1)
        self.region = np.ascontiguousarray(np.asarray(st.get("region", np.zeros(nm))).astype(np.uint8)).copy()
        self.atom_x = _d(st.get("atom_x", np.zeros(3 * na))).copy().reshape(-1)
        self.atom_v = _d(st.get("atom_v", np.zeros(3 * na))).copy().reshape(-1)
        self.atom_f = _d(st.get("atom_f", np.zeros(3 * na))).copy().reshape(-1)
        self.atom_w = _d(st.get("atom_w", np.zeros(na))).copy().reshape(-1)
        assert self.mol_x.size == 3 * nm and self.atom_x.size == 3 * na, "state size mismatch"
        self.gi = np.zeros(2 * cap, np.int32)
        self.gv = np.zeros(2 * cap, np.int32)
        self.gg = np.zeros(2 * cap, np.int32)
        c = OrcState()
        c.mol_x, c.mol_v = self.mol_x.ctypes.data_as(_dp), self.mol_v.ctypes.data_as(_dp)
        c.region = self.region.ctypes.data_as(C.POINTER(C.c_uint8))
        c.atom_x, c.atom_v = self.atom_x.ctypes.data_as(_dp), self.atom_v.ctypes.data_as(_dp)
        c.atom_f, c.atom_w = self.atom_f.ctypes.data_as(_dp), self.atom_w.ctypes.data_as(_dp)
        c.loopPair = int(st.get("loopPair", 0))
        c.loop_update = int(st.get("loop_update", 0))
        c.rebuild_count = int(st.get("rebuild_count", 0))
        c.zeta_ion = float(st.get("zeta_ion", 0.0))
        c.cap_gi = c.cap_gv = c.cap_gg = cap
        c.gi, c.gv, c.gg = (a.ctypes.data_as(_ip) for a in (self.gi, self.gv, self.gg))
        for k, nmk in (("gi", "n_gi"), ("gv", "n_gv"), ("gg", "n_gg")):
            if k in st:
                arr = np.asarray(st[k], np.int32).reshape(-1)
                getattr(self, k)[: arr.size] = arr
                setattr(c, nmk, arr.size // 2)
        self.c = c

    def sync(self):
        pass

    @property
    def U(self):
        return np.array(self.c.U[:])

    def lists(self):
        c = self.c
        return {"gi": self.gi[: 2 * c.n_gi].reshape(-1, 2).copy(), "gv": self.gv[: 2 * c.n_gv].reshape(-1, 2).copy(),
                "gg": self.gg[: 2 * c.n_gg].reshape(-1, 2).copy()}

    def as_dict(self):
        d = {k: getattr(self, k).copy() for k in ("mol_x", "mol_v", "region", "atom_x", "atom_v", "atom_f", "atom_w")}
        d.update(loopPair=self.c.loopPair, loop_update=self.c.loop_update, rebuild_count=self.c.rebuild_count,
                 zeta_ion=self.c.zeta_ion)
        d.update({k: v.reshape(-1) for k, v in self.lists().items()})
        return d


def batch_verlet(oracles, states, nsteps, eflag_every=0, nthreads=0, seconds=None):
    """One replica per OpenMP thread (cpu baseline).  With a list for
    `seconds`, each replica's loop time is appended to it."""
    n = len(oracles)
    hs = (C.c_void_p * n)(*[o.h for o in oracles])
    ss = (C.POINTER(OrcState) * n)(*[C.pointer(s.c) for s in states])
    acc = (C.c_int64 * n)()
    sec = (C.c_double * n)()
    rc = lib().orc_verlet_batch(hs, ss, n, nsteps, eflag_every, 0, acc, nthreads, sec)
    if seconds is not None:
        seconds.extend(sec[:])
    return rc, list(acc)


# ---------------- reference harness helpers (only where it is built) -------------
def read_arrays(text):
    out = {}
    for line in text.splitlines():
        parts = line.split()
        if len(parts) < 2:
            continue
        n = int(parts[1])
        out[parts[0]] = np.array([float(v) for v in parts[2:2 + n]], np.float64)
    return out


def write_arrays(path, arrays):
    with open(path, "w") as f:
        for k, v in arrays.items():
            v = np.asarray(v, np.float64).reshape(-1)
            f.write("%s %d %s\n" % (k, v.size, " ".join(repr(float(x)) for x in v)))


def sysd_from_ref_dump(a, has_gas=True, seed=1, replica=0, shake=0):
    """Oracle system dict from the harness' ``sys`` dump (reference parse)."""
    sc = a["scalars"]
    fl = a["flags"]
    flags = 0
    for bit, v in zip(FLAG_ORDER, fl):
        if v:
            flags |= bit
    species = []
    for p, tm in (("ion", None), ("gas", "gas_tmpl"), ("vap", "vap_tmpl")):
        sp = {k: a[p + "_" + k] for k in ("mass", "charge", "type", "bond", "angle", "dih", "rigid")}
        sp["tmpl_x"] = a[tm] if tm else np.zeros(3 * a[p + "_mass"].size)
        sp["molmass"] = float(a[p + "_molmass"][0])
        species.append(sp)
    n_ion = species[0]["mass"].size
    return {
        "pair_coeff": a["pair_coeff"], "btype": a["btype"], "ctype": a["ctype"], "dmulti": a["dmulti"],
        "dcoeff": a["dcoeff"], "species": species, "n_gas": int(sc[10]), "n_vap": int(sc[11]),
        "ion_type_base": 1 if has_gas else 0, "dt": sc[0], "T": sc[1], "d_size": sc[3], "margin": sc[5],
        "ML2": sc[6], "m_gas_kg": sc[7], "m_vap_kg": sc[8], "gas_spin": 1 if int(sc[9]) == 2 else 0,
        "flags": flags, "efield": tuple(a.get("efield", np.zeros(3))), "shake": shake, "shake_max_iter": 1000, "seed": seed,
        "replica": replica, "n_ion": n_ion,
        # (mode, Tnh) of the NVTion line; absent from dumps that predate it
        "nvt_ion": int(a["nvt"][0]) if "nvt" in a else 0, "Tnh_ion": float(a["nvt"][1]) if "nvt" in a else 0.0,
    }


def run_harness(cond, state_path, op, arg=None, cwd=None):
    cmd = [REF_HARNESS, cond, state_path or "-", op] + ([str(arg)] if arg is not None else [])
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, check=True)
    return read_arrays(r.stdout)
