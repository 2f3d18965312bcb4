"""ctypes binding of libmdn (include/mdn.h).

The library is built in-tree (md_nucleation_amd/libmdn.so, see build.py).
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmdn.so")

MDN_GI, MDN_VI, MDN_GG, MDN_GV, MDN_VV = 1, 2, 4, 8, 16
MDN_ION_AMBER, MDN_ION_BMH, MDN_ION_TERSOFF, MDN_ION_SW = 32, 64, 128, 256
MDN_EFIELD, MDN_VAPOR_INTRA, MDN_GAS_INTRA = 512, 1024, 2048
FP64, MIXED = 0, 1
NVE, VSCALE, NOSE_HOOVER = 0, 1, 2  # ion thermostat modes (NVTion OFF / scale / <T>)
U_LABELS = ("Uion", "Ugas", "Uvap", "Ugi", "Ugg", "Uvg", "Uvi", "Uvv")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MdnSystem(C.Structure):
    _fields_ = [
        ("n_types", C.c_int32), ("pair_coeff", _dp), ("type_eps", _dp), ("type_sigma", _dp), ("type_mass", _dp),
        ("n_btypes", C.c_int32), ("btype", _dp), ("n_ctypes", C.c_int32), ("ctype", _dp),
        ("n_dtypes", C.c_int32), ("dmulti", _ip), ("dcoeff", _dp),
        ("natoms", C.c_int32 * 3), ("mass", _dp * 3), ("charge", _dp * 3), ("type", _ip * 3), ("tmpl_x", _dp * 3),
        ("nbond", C.c_int32 * 3), ("bond", _ip * 3), ("nangle", C.c_int32 * 3), ("angle", _ip * 3),
        ("ndih", C.c_int32 * 3), ("dih", _ip * 3), ("nrigid", C.c_int32 * 3), ("rigid", _ip * 3),
        ("molmass", C.c_double * 3), ("n_gas", C.c_int32), ("n_vap", C.c_int32), ("ion_type_base", C.c_int32),
        ("dt", C.c_double), ("T", C.c_double), ("pressure", C.c_double), ("d_size", C.c_double),
        ("cutoff", C.c_double), ("margin", C.c_double), ("ML2", C.c_double),
        ("m_gas_kg", C.c_double), ("m_vap_kg", C.c_double), ("gastype", C.c_int32), ("gas_spin", C.c_int32),
        ("flags", C.c_uint32), ("efield", C.c_double * 3), ("shake", C.c_int32), ("shake_max_iter", C.c_int32),
        ("total_steps", C.c_int64), ("relax_steps", C.c_int64), ("observe", C.c_int32),
        ("cluster_radius", C.c_double), ("nvt_ion", C.c_int32), ("Tnh_ion", C.c_double),
    ]


class MdnState(C.Structure):
    _fields_ = [
        ("mol_x", _dp), ("mol_v", _dp), ("region", C.POINTER(C.c_uint8)),
        ("atom_x", _dp), ("atom_v", _dp), ("atom_f", _dp), ("atom_w", _dp),
        ("loopPair", C.c_int32), ("loop_update", C.c_int32), ("rebuild_count", C.c_int64),
        ("cap_gi", C.c_int32), ("cap_gv", C.c_int32), ("cap_gg", C.c_int32),
        ("n_gi", C.c_int32), ("n_gv", C.c_int32), ("n_gg", C.c_int32),
        ("gi", _ip), ("gv", _ip), ("gg", _ip), ("U", C.c_double * 8), ("zeta_ion", C.c_double),
    ]


# every entry point of include/mdn.h with its argument types
class MdnVaporEvent(C.Structure):
    _fields_ = [("mol", C.c_int32), ("kind", C.c_int32), ("step", C.c_int64), ("dx", C.c_double * 3),
                ("dv", C.c_double * 3)]


_SIGS = {
    "mdn_last_error": (C.c_char_p, []),
    "mdn_sysdef_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mdn_sysdef_view": (C.c_int, [C.c_void_p, C.POINTER(MdnSystem)]),
    "mdn_sysdef_sizes": (C.c_int, [C.c_void_p, _ip, _ip]),
    "mdn_sysdef_set_shake": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "mdn_sysdef_set_flags": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mdn_sysdef_init_state": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(MdnState)]),
    "mdn_sysdef_free": (None, [C.c_void_p]),
    "mdn_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mdn_destroy": (None, [C.c_void_p]),
    "mdn_add_replica": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(MdnState), C.c_uint64, C.c_uint32, _ip]),
    "mdn_commit": (C.c_int, [C.c_void_p]),
    "mdn_set_state": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MdnState)]),
    "mdn_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MdnState)]),
    "mdn_set_thermostat": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "mdn_observe": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mdn_set_vapor_log": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_drain_vapor_log": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MdnVaporEvent), C.c_int32, _ip]),
    "mdn_set_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "mdn_recentre": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_run": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_uint32]),
    "mdn_set_positions": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mdn_get_forces": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mdn_force_jvp": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, _dp, _dp]),
    "mdn_segment_grad": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_int32), C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mdn_make_pair": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_zero_forces": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_compute": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_int32, C.c_int32]),
    "mdn_velocity_calculation": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_update_position": (C.c_int, [C.c_void_p, C.c_int32]),
    "mdn_verlet": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int32]),
    "mdn_verlet_graph": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int32]),
    "mdn_run_batched": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.c_int32]),
    "mdn_set_batch_shape": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mdn_set_batch_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "mdn_get_counters": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _ip,
                                   C.POINTER(C.c_int64)]),
    "mdn_get_batch_info": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "mdn_get_batch_plan": (C.c_int, [C.c_void_p, _ip, _ip, C.c_int32]),
    "mdn_vi_forward": (C.c_int, [_dp, _dp, C.c_int32, C.c_int32, C.c_int32] + [_dp] * 6 + [C.c_void_p, _dp, _dp]),
    "mdn_vi_adjoint": (C.c_int, [_dp, _dp, C.c_int32, C.c_int32, C.c_int32] + [_dp] * 6 + [C.c_void_p] + [_dp] * 5),
    "mdn_vi_jvp": (C.c_int, [_dp, _dp, C.c_int32, C.c_int32, C.c_int32] + [_dp] * 6 + [C.c_void_p] + [_dp] * 2),
    "mdn_vi_adjoint_noalloc": (C.c_int, [C.c_void_p] * 2 + [C.c_int32] * 3 + [C.c_void_p] * 12 + [C.c_void_p]),
    "mdn_qqrd2e": (C.c_double, []),
    "mdn_sync": (C.c_int, [C.c_void_p]),
    "mdn_last_kernel_ms": (C.c_int, [C.c_void_p, _dp]),
}

_lib = None


class MdnError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MdnError("libmdn.so is not built (run `python -m md_nucleation_amd.build`)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().mdn_last_error().decode(errors="replace")
        raise MdnError("%s failed (%d): %s" % (what, rc, msg))


def _arr(p, n, dt):
    if n <= 0 or not p:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()


class SysDef:
    """A parsed condition file (mdn_sysdef_load)."""

    def __init__(self, cond_path):
        h = C.c_void_p()
        check(lib().mdn_sysdef_load(os.fsencode(cond_path), C.byref(h)), "mdn_sysdef_load")
        self.h = h
        self.refresh()

    def refresh(self):
        v = MdnSystem()
        check(lib().mdn_sysdef_view(self.h, C.byref(v)), "mdn_sysdef_view")
        nm, na = C.c_int32(), C.c_int32()
        check(lib().mdn_sysdef_sizes(self.h, C.byref(nm), C.byref(na)), "mdn_sysdef_sizes")
        self.n_mol, self.n_atoms = nm.value, na.value
        self.view = self._to_dict(v)

    @staticmethod
    def _to_dict(v):
        nt = v.n_types
        d = {
            "n_types": nt, "pair_coeff": _arr(v.pair_coeff, 2 * nt * nt, np.float64),
            "type_eps": _arr(v.type_eps, nt, np.float64), "type_sigma": _arr(v.type_sigma, nt, np.float64),
            "type_mass": _arr(v.type_mass, nt, np.float64),
            "btype": _arr(v.btype, 2 * v.n_btypes, np.float64), "ctype": _arr(v.ctype, 2 * v.n_ctypes, np.float64),
            "dmulti": _arr(v.dmulti, v.n_dtypes, np.int32), "dcoeff": _arr(v.dcoeff, 15 * v.n_dtypes, np.float64),
            "species": [],
        }
        for k in range(3):
            n = v.natoms[k]
            d["species"].append({
                "mass": _arr(v.mass[k], n, np.float64), "charge": _arr(v.charge[k], n, np.float64),
                "type": _arr(v.type[k], n, np.int32), "tmpl_x": _arr(v.tmpl_x[k], 3 * n, np.float64),
                "bond": _arr(v.bond[k], 3 * v.nbond[k], np.int32), "angle": _arr(v.angle[k], 4 * v.nangle[k], np.int32),
                "dih": _arr(v.dih[k], 5 * v.ndih[k], np.int32), "rigid": _arr(v.rigid[k], 3 * v.nrigid[k], np.int32),
                "molmass": v.molmass[k],
            })
        for k in ("n_gas", "n_vap", "ion_type_base", "dt", "T", "pressure", "d_size", "cutoff", "margin", "ML2",
                  "m_gas_kg", "m_vap_kg", "gastype", "gas_spin", "flags", "shake", "shake_max_iter",
                  "total_steps", "relax_steps", "observe", "cluster_radius", "nvt_ion", "Tnh_ion"):
            d[k] = getattr(v, k)
        d["efield"] = tuple(v.efield)
        return d

    def set_shake(self, on, max_iter=1000):
        check(lib().mdn_sysdef_set_shake(self.h, int(on), int(max_iter)), "mdn_sysdef_set_shake")
        self.refresh()

    def set_flags(self, flags):
        check(lib().mdn_sysdef_set_flags(self.h, int(flags)), "mdn_sysdef_set_flags")
        self.refresh()

    def new_state(self):
        return State(self.n_mol, self.n_atoms)

    def init_state(self, seed):
        st = self.new_state()
        check(lib().mdn_sysdef_init_state(self.h, int(seed), C.byref(st.c)), "mdn_sysdef_init_state")
        return st

    def __del__(self):
        try:
            lib().mdn_sysdef_free(self.h)
        except Exception:
            pass


class State:
    """numpy-backed mdn_state (reference layout: global atom order, lists with
    global molecule ids)."""

    def __init__(self, n_mol, n_atoms, cap=0, arrays=None):
        a = arrays or {}
        self.mol_x = np.ascontiguousarray(a.get("mol_x", np.zeros(3 * n_mol)), np.float64).reshape(-1).copy()
        self.mol_v = np.ascontiguousarray(a.get("mol_v", np.zeros(3 * n_mol)), np.float64).reshape(-1).copy()
        self.region = np.ascontiguousarray(np.asarray(a.get("region", np.zeros(n_mol))).astype(np.uint8)).copy()
        self.atom_x = np.ascontiguousarray(a.get("atom_x", np.zeros(3 * n_atoms)), np.float64).reshape(-1).copy()
        self.atom_v = np.ascontiguousarray(a.get("atom_v", np.zeros(3 * n_atoms)), np.float64).reshape(-1).copy()
        self.atom_f = np.ascontiguousarray(a.get("atom_f", np.zeros(3 * n_atoms)), np.float64).reshape(-1).copy()
        self.atom_w = np.ascontiguousarray(a.get("atom_w", np.zeros(n_atoms)), np.float64).reshape(-1).copy()
        self.lists = {}
        c = MdnState()
        c.mol_x, c.mol_v = self.mol_x.ctypes.data_as(_dp), self.mol_v.ctypes.data_as(_dp)
        c.region = self.region.ctypes.data_as(C.POINTER(C.c_uint8))
        c.atom_x, c.atom_v = self.atom_x.ctypes.data_as(_dp), self.atom_v.ctypes.data_as(_dp)
        c.atom_f, c.atom_w = self.atom_f.ctypes.data_as(_dp), self.atom_w.ctypes.data_as(_dp)
        c.loopPair = int(a.get("loopPair", 0))
        c.loop_update = int(a.get("loop_update", 0))
        c.rebuild_count = int(a.get("rebuild_count", 0))
        c.zeta_ion = float(a.get("zeta_ion", 0.0))
        for k in ("gi", "gv", "gg"):
            arr = np.ascontiguousarray(np.asarray(a.get(k, np.zeros(0)), np.int32).reshape(-1))
            capk = max(cap, arr.size // 2)
            buf = np.zeros(2 * max(capk, 1), np.int32)
            buf[: arr.size] = arr
            self.lists[k] = buf
            setattr(c, k, buf.ctypes.data_as(_ip))
            setattr(c, "n_" + k, arr.size // 2)
            setattr(c, "cap_" + k, capk)
        self.c = c

    def set_caps(self, cap):
        for k in ("gi", "gv", "gg"):
            n = getattr(self.c, "n_" + k)
            buf = np.zeros(2 * max(cap, n, 1), np.int32)
            buf[: 2 * n] = self.lists[k][: 2 * n]
            self.lists[k] = buf
            setattr(self.c, k, buf.ctypes.data_as(_ip))
            setattr(self.c, "cap_" + k, max(cap, n))

    @property
    def U(self):
        return np.array(self.c.U[:])

    def list(self, k):
        return self.lists[k][: 2 * getattr(self.c, "n_" + k)].reshape(-1, 2).copy()

    def as_dict(self):
        d = {k: getattr(self, k).copy() for k in ("mol_x", "mol_v", "region", "atom_x", "atom_v", "atom_f", "atom_w")}
        d.update(loopPair=self.c.loopPair, loop_update=self.c.loop_update, rebuild_count=self.c.rebuild_count,
                 zeta_ion=self.c.zeta_ion)
        for k in ("gi", "gv", "gg"):
            d[k] = self.list(k).reshape(-1)
        return d


class Engine:
    """A libmdn device context holding one or more replicas."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().mdn_create(int(device), C.byref(h)), "mdn_create")
        self.h = h
        self.sys = []

    def add(self, sysdef: SysDef, state: State, seed=1, replica_id=0):
        r = C.c_int32()
        check(lib().mdn_add_replica(self.h, sysdef.h, C.byref(state.c), int(seed), int(replica_id), C.byref(r)),
              "mdn_add_replica")
        self.sys.append(sysdef)
        return r.value

    def commit(self):
        check(lib().mdn_commit(self.h), "mdn_commit")

    def set_state(self, r, st: State):
        check(lib().mdn_set_state(self.h, int(r), C.byref(st.c)), "mdn_set_state")

    def set_thermostat(self, r, mode, Tnh=0.0):
        """MD::setNVE (mode NVE) or an NVTion mode for replica r (-1: all)."""
        check(lib().mdn_set_thermostat(self.h, int(r), int(mode), float(Tnh)), "mdn_set_thermostat")

    def observe(self, r=-1):
        """Observer::computeProps: rows of (Kion, K_g, K_v, Tion, T_g, T_v)."""
        n = len(self.sys) if r == -1 else 1
        out = np.zeros(6 * n)
        check(lib().mdn_observe(self.h, int(r), out.ctypes.data_as(_dp)), "mdn_observe")
        return out.reshape(n, 6)

    def set_vapor_log(self, cap):
        check(lib().mdn_set_vapor_log(self.h, int(cap)), "mdn_set_vapor_log")

    def drain_vapor_log(self, r, cap=1 << 16):
        """The replica's Ovin/Ovout events since the last drain: a structured
        array (mol, kind, step, dx[3], dv[3])."""
        buf = (MdnVaporEvent * cap)()
        n = C.c_int32(0)
        check(lib().mdn_drain_vapor_log(self.h, int(r), buf, cap, C.byref(n)), "mdn_drain_vapor_log")
        dt = np.dtype([("mol", np.int32), ("kind", np.int32), ("step", np.int64), ("dx", np.float64, 3),
                       ("dv", np.float64, 3)])
        return np.frombuffer(bytes(buf)[: n.value * C.sizeof(MdnVaporEvent)], dtype=dt).copy()

    def set_step(self, r, step):
        check(lib().mdn_set_step(self.h, int(r), int(step)), "mdn_set_step")

    def recentre(self, r=-1):
        check(lib().mdn_recentre(self.h, int(r)), "mdn_recentre")

    def run(self, outdir, calc0=0, precision=FP64, verbose=False):
        """MD::run for every replica (relaxation, re-centre, make_pair, setNVE,
        main loop) writing the reference's output files under outdir."""
        check(lib().mdn_run(self.h, str(outdir).encode(), int(calc0), int(precision), 1 if verbose else 0),
              "mdn_run")

    def get_state(self, r, cap=None) -> State:
        sd = self.sys[r]
        st = State(sd.n_mol, sd.n_atoms, cap=cap if cap is not None else 4 * sd.n_mol + 1024)
        check(lib().mdn_get_state(self.h, int(r), C.byref(st.c)), "mdn_get_state")
        return st

    def set_positions(self, r, atom_x):
        x = np.ascontiguousarray(atom_x, np.float64).reshape(-1)
        if x.size != 3 * self.sys[r].n_atoms:
            raise ValueError("atom_x must hold 3 * n_atoms values")
        check(lib().mdn_set_positions(self.h, int(r), x.ctypes.data_as(_dp)), "mdn_set_positions")

    def get_forces(self, r):
        f = np.zeros(3 * self.sys[r].n_atoms)
        check(lib().mdn_get_forces(self.h, int(r), f.ctypes.data_as(_dp)), "mdn_get_forces")
        return f.reshape(-1, 3)

    def force_jvp(self, r, bits, u):
        """d/de F(x + e u) of the selected terms at the replica's positions
        (the Hessian of -U applied to u), FP64, [n_atoms, 3]"""
        uu = np.ascontiguousarray(np.asarray(u, np.float64).reshape(-1))
        assert uu.size == 3 * self.sys[r].n_atoms
        hu = np.zeros_like(uu)
        check(lib().mdn_force_jvp(self.h, int(r), int(bits), uu.ctypes.data_as(_dp), hu.ctypes.data_as(_dp)),
              "mdn_force_jvp")
        return hu.reshape(-1, 3)

    def segment_grad(self, r, nsteps, every, theta, obs, targets, x0=None, v0=None, grad=True, end=False):
        """mdn_segment_grad: (loss, (dL/dsigma, dL/deps, dL/dq) or None[, x_end, v_end])
        of the trajectory loss over ``nsteps`` MD::verlet steps of the
        replica's ion + vapor atoms, computed on the device; ``obs`` the
        observed steps (ascending), ``targets`` [len(obs), N, 3]"""
        sig, eps, q = (np.ascontiguousarray(np.asarray(a, np.float64)) for a in theta)
        n_ion = sig.size
        obs = np.ascontiguousarray(np.asarray(obs, np.int32))
        tg = np.ascontiguousarray(np.asarray(targets, np.float64))
        N = tg.shape[1] if tg.ndim == 3 and tg.shape[0] else None

        def opt(a):
            if a is None:
                return None, None
            a = np.ascontiguousarray(np.asarray(a, np.float64))
            return a, a.ctypes.data_as(_dp)
        x0a, px0 = opt(x0)
        v0a, pv0 = opt(v0)
        if N is None:
            N = (x0a.size // 3) if x0a is not None else None
        loss = C.c_double(0)
        gs = [np.zeros(n_ion) for _ in range(3)] if grad else None
        pg = [g.ctypes.data_as(_dp) for g in gs] if grad else [None] * 3
        xe = ve = None
        if end:
            if N is None:
                raise ValueError("segment_grad: give x0 or targets to size the end state")
            xe, ve = np.zeros((N, 3)), np.zeros((N, 3))
        check(lib().mdn_segment_grad(self.h, int(r), int(nsteps), int(every), px0, pv0,
                                     sig.ctypes.data_as(_dp), eps.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                                     obs.ctypes.data_as(C.POINTER(C.c_int32)), int(obs.size),
                                     tg.ctypes.data_as(_dp) if obs.size else None, C.byref(loss), *pg,
                                     xe.ctypes.data_as(_dp) if end else None,
                                     ve.ctypes.data_as(_dp) if end else None), "mdn_segment_grad")
        out = (loss.value, tuple(gs) if grad else None)
        return out + (xe, ve) if end else out

    def make_pair(self, r):
        check(lib().mdn_make_pair(self.h, int(r)), "mdn_make_pair")

    def zero_forces(self, r):
        check(lib().mdn_zero_forces(self.h, int(r)), "mdn_zero_forces")

    def compute(self, r, bits, eflag=1, precision=FP64):
        check(lib().mdn_compute(self.h, int(r), int(bits), int(eflag), int(precision)), "mdn_compute")

    def velocity_calculation(self, r):
        check(lib().mdn_velocity_calculation(self.h, int(r)), "mdn_velocity_calculation")

    def update_position(self, r):
        check(lib().mdn_update_position(self.h, int(r)), "mdn_update_position")

    def verlet(self, r, nsteps, eflag_every=0, step0=0, precision=FP64):
        check(lib().mdn_verlet(self.h, int(r), int(nsteps), int(eflag_every), int(step0), int(precision)),
              "mdn_verlet")

    def verlet_graph(self, r, nsteps, eflag_every=0, step0=0, precision=FP64):
        check(lib().mdn_verlet_graph(self.h, int(r), int(nsteps), int(eflag_every), int(step0), int(precision)),
              "mdn_verlet_graph")

    def run_batched(self, nsteps, eflag_every=0, step0=0, precision=MIXED):
        check(lib().mdn_run_batched(self.h, int(nsteps), int(eflag_every), int(step0), int(precision)),
              "mdn_run_batched")

    def set_batch_shape(self, threads=0, lds_cap_atoms=0, grid=0, chunk=1000):
        check(lib().mdn_set_batch_shape(self.h, int(threads), int(lds_cap_atoms), int(grid), int(chunk)),
              "mdn_set_batch_shape")

    def set_batch_tuning(self, own_threads=512, shared_threads=128, heavy_factor=2.0):
        check(lib().mdn_set_batch_tuning(self.h, int(own_threads), int(shared_threads), float(heavy_factor)),
              "mdn_set_batch_tuning")

    def counters(self, r):
        """(in-domain atom-steps, rebuilds, max SHAKE iterations, SHAKE failures)"""
        a, b, s, f = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        check(lib().mdn_get_counters(self.h, int(r), C.byref(a), C.byref(b), C.byref(s), C.byref(f)),
              "mdn_get_counters")
        return a.value, b.value, s.value, f.value

    def batch_info(self):
        """(LDS capacity in atoms, LDS bytes per workgroup, replicas with a CU of
        their own) of the last batched launch"""
        a, b, h = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().mdn_get_batch_info(self.h, C.byref(a), C.byref(b), C.byref(h)), "mdn_get_batch_info")
        return a.value, b.value, h.value

    def batch_plan(self):
        """Launch plan of the last batched chunk: ({"own_cu", "three_per_cu",
        "four_per_cu", "uniform"} -> replica count, replica ids in launch
        order)."""
        counts = np.zeros(4, np.int32)
        order = np.zeros(max(len(self.sys), 1), np.int32)
        check(lib().mdn_get_batch_plan(self.h, counts.ctypes.data_as(_ip), order.ctypes.data_as(_ip), order.size),
              "mdn_get_batch_plan")
        n = int(counts.sum())
        keys = ("own_cu", "three_per_cu", "four_per_cu", "uniform")
        return {k: int(v) for k, v in zip(keys, counts)}, order[:n].copy()

    def sync(self):
        check(lib().mdn_sync(self.h), "mdn_sync")

    def last_kernel_ms(self):
        ms = C.c_double()
        check(lib().mdn_last_kernel_ms(self.h, C.byref(ms)), "mdn_last_kernel_ms")
        return ms.value

    def close(self):
        if self.h:
            lib().mdn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
