"""Differentiable ion-vapor LJ + Coulomb term for fitting the ion's per-atom
sigma, eps and charge to target forces or energies.

The term is PotentialVaporIon::compute (src/potential/VaporIon.cpp:5-43)
with the Lorentz-Berthelot mixing of setCrossPotential.cpp:13-20 evaluated
per pair, so the ion's per-atom parameters are the inputs.  The forward pass
and the hand-written reverse pass run in FP64 on the GPU through libmdn's
C ABI (``mdn_vi_forward`` / ``mdn_vi_adjoint``); ``ViTerm`` puts them behind
``torch.autograd.Function`` so a fitting loop can use torch's optimisers.

Tensors cross the boundary as host float64 arrays (the C ABI copies them in
and out), so torch's own device runtime is never involved.

    snaps = snapshots(sd, st, seed, n=256, every=10)    # trajectory frames
    sig, eps, q = ion_params(sd, requires_grad=True)
    E, F = vi_energy_forces(sig, eps, q, snaps)
    loss = ((F - F_target) ** 2).sum()
    loss.backward()                                     # sig.grad, ...
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as M

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


@dataclass
class ViSnapshots:
    """S frames of ion + vapor atoms.

    x [S, n_ion + n_vap, 3] (ion atoms first, then vapor atoms in molecule
    order), w [S, n_ion + n_vap] blend weights, vap_in [S, n_vap] 1 where the
    vapor atom's molecule is inside the effective domain (only those interact,
    VaporIon.cpp:10), and the vapor's per-atom sigma, eps, q.
    """
    x: np.ndarray
    w: np.ndarray
    vap_in: np.ndarray
    n_ion: int
    vap_sigma: np.ndarray
    vap_eps: np.ndarray
    vap_q: np.ndarray

    @property
    def S(self):
        return self.x.shape[0]

    @property
    def n_vap(self):
        return self.x.shape[1] - self.n_ion

    def subset(self, idx):
        return ViSnapshots(self.x[idx].copy(), self.w[idx].copy(), self.vap_in[idx].copy(), self.n_ion,
                           self.vap_sigma, self.vap_eps, self.vap_q)


def _species_params(v, k):
    sp = v["species"][k]
    t = np.asarray(sp["type"], np.int64)
    return (np.ascontiguousarray(v["type_sigma"][t], np.float64), np.ascontiguousarray(v["type_eps"][t], np.float64),
            np.ascontiguousarray(sp["charge"], np.float64))


def ion_params(sd, requires_grad=False):
    """The ion's per-atom (sigma, eps, q) of a SysDef as float64 tensors."""
    return tuple(torch.tensor(a, dtype=torch.float64, requires_grad=requires_grad)
                 for a in _species_params(sd.view, 0))


def frame(sd, st):
    """(x, w, vap_in) of the ion and vapor atoms of one state."""
    v = sd.view
    n0, n1, n2 = (s["mass"].size for s in v["species"])
    g0 = n0 + v["n_gas"] * n1
    ax = st.atom_x.reshape(-1, 3)
    x = np.concatenate([ax[:n0], ax[g0:g0 + v["n_vap"] * n2]])
    w = np.concatenate([st.atom_w[:n0], st.atom_w[g0:g0 + v["n_vap"] * n2]])
    vin = np.repeat(np.asarray(st.region[1 + v["n_gas"]:], np.uint8) != 0, n2).astype(np.uint8)
    return x, w, vin


def snapshots(sd, st, seed, n=256, every=10, device=0, warmup=0):
    """``n`` frames ``every`` FP64 steps apart from one trajectory of the
    system (libmdn's per-call path, mdn_verlet)."""
    eng = M.Engine(device)
    try:
        eng.add(sd, st, seed=seed)
        eng.commit()
        eng.make_pair(0)
        if warmup:
            eng.verlet(0, warmup, 0, 0, M.FP64)
        xs, ws, vs = [], [], []
        for k in range(n):
            eng.verlet(0, every, 0, warmup + k * every, M.FP64)
            x, w, vin = frame(sd, eng.get_state(0))
            xs.append(x)
            ws.append(w)
            vs.append(vin)
    finally:
        eng.close()
    sv, ev, qv = _species_params(sd.view, 2)
    v = sd.view
    rep = v["n_vap"]
    return ViSnapshots(np.ascontiguousarray(np.stack(xs)), np.ascontiguousarray(np.stack(ws)),
                       np.ascontiguousarray(np.stack(vs)), v["species"][0]["mass"].size,
                       np.tile(sv, rep), np.tile(ev, rep), np.tile(qv, rep))


def _prep(snap, sig, eps, q):
    n_ion = snap.n_ion
    ps = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (sig, eps, q)]
    for a in ps:
        if a.size != n_ion:
            raise ValueError("ion parameter arrays must have n_ion = %d entries" % n_ion)
    x = np.ascontiguousarray(snap.x, np.float64)
    w = np.ascontiguousarray(snap.w, np.float64)
    vin = np.ascontiguousarray(snap.vap_in, np.uint8)
    vp = [np.ascontiguousarray(a, np.float64) for a in (snap.vap_sigma, snap.vap_eps, snap.vap_q)]
    if x.shape[1] != n_ion + vp[0].size or w.shape != x.shape[:2] or vin.shape != (x.shape[0], vp[0].size):
        raise ValueError("snapshot arrays have inconsistent shapes")
    return x, w, vin, ps, vp


def vi_forward(snap, sig, eps, q):
    """E [S], F [S, n_ion + n_vap, 3] (numpy, FP64) of the vi term."""
    x, w, vin, ps, vp = _prep(snap, sig, eps, q)
    S, n = x.shape[:2]
    E = np.zeros(S)
    F = np.zeros((S, n, 3))
    M.check(M.lib().mdn_vi_forward(_p(x), _p(w), S, snap.n_ion, n - snap.n_ion, *map(_p, ps), *map(_p, vp),
                                   vin.ctypes.data_as(C.c_void_p), _p(E), _p(F)), "mdn_vi_forward")
    return E, F


def vi_adjoint(snap, sig, eps, q, gE=None, gF=None):
    """dL/dsigma, dL/deps, dL/dq of the ion atoms given dL/dE [S] and
    dL/dF [S, n, 3] (either may be None)."""
    x, w, vin, ps, vp = _prep(snap, sig, eps, q)
    S, n = x.shape[:2]
    ge = None if gE is None else np.ascontiguousarray(np.asarray(gE, np.float64).reshape(S))
    gf = None if gF is None else np.ascontiguousarray(np.asarray(gF, np.float64).reshape(S, n, 3))
    out = [np.zeros(snap.n_ion) for _ in range(3)]
    M.check(M.lib().mdn_vi_adjoint(_p(x), _p(w), S, snap.n_ion, n - snap.n_ion, *map(_p, ps), *map(_p, vp),
                                   vin.ctypes.data_as(C.c_void_p), None if ge is None else _p(ge),
                                   None if gf is None else _p(gf), *map(_p, out)), "mdn_vi_adjoint")
    return tuple(out)


def vi_jvp(snap, sig, eps, q, U):
    """derivative of the vi forces along the displacement U [S, n, 3] of the
    snapshots' atoms (the Hessian of -E applied to U), [S, n, 3]"""
    x, w, vin, ps, vp = _prep(snap, sig, eps, q)
    S, n = x.shape[:2]
    u = np.ascontiguousarray(np.asarray(U, np.float64).reshape(S, n, 3))
    hu = np.zeros((S, n, 3))
    M.check(M.lib().mdn_vi_jvp(_p(x), _p(w), S, snap.n_ion, n - snap.n_ion, *map(_p, ps), *map(_p, vp),
                               vin.ctypes.data_as(C.c_void_p), _p(u), _p(hu)), "mdn_vi_jvp")
    return hu


def _host(t):
    return t.detach().to("cpu", torch.float64).numpy()


class ViTerm(torch.autograd.Function):
    """(sigma, eps, q) of the ion atoms -> (E [S], F [S, n, 3])."""

    @staticmethod
    def forward(ctx, sig, eps, q, snap):
        E, F = vi_forward(snap, _host(sig), _host(eps), _host(q))
        ctx.snap = snap
        ctx.save_for_backward(sig, eps, q)
        return torch.from_numpy(E).to(sig.device), torch.from_numpy(F).to(sig.device)

    @staticmethod
    def backward(ctx, gE, gF):
        sig, eps, q = ctx.saved_tensors
        g = vi_adjoint(ctx.snap, _host(sig), _host(eps), _host(q),
                       None if gE is None else _host(gE), None if gF is None else _host(gF))
        return tuple(torch.from_numpy(a).to(sig.device) for a in g) + (None,)


def vi_energy_forces(sig, eps, q, snap):
    return ViTerm.apply(sig, eps, q, snap)
