"""Gradients of a loss over a trajectory segment of MD::verlet with respect to
the ion's per-atom sigma, eps and q as the vapor sees them (the parameters of
the vi term, PotentialVaporIon::compute, VaporIon.cpp:5-43), by reverse-mode
differentiation through the integrator with checkpoints.

Dynamics: MD::verlet as the reference ships it (verlet.cpp:20-39): half
kick, drift, forces at the new positions, half kick; no SHAKE (the call is
commented out, verlet.cpp:24), no rebuild within a segment, blend weights 1
(every vapor molecule within 100 A of the ion).  The vi term takes the ion
parameters as inputs; every other selected term (AMBER, vv, vapor intra,
Efield) runs as ``mdn_compute`` runs it.

``loss_and_grad`` (and ``TrajectoryLoss``, its ``torch.autograd.Function``)
run the whole segment on the device through one C-ABI call,
``mdn_segment_grad``: state, checkpoints, recomputation, Hessian-vector
products and the parameter adjoint all in device memory.  ``Integrator`` /
``loss_and_grad_host`` are the same reverse pass written over host numpy
arrays, calling the device per force evaluation (``Segment``) or a numpy
pair model (the CPU tests); they check the device path.

Reverse pass.  One step n -> n + 1 is
    vh = v_n + c F(x_n) / m,  x_{n+1} = x_n + dt vh,  v_{n+1} = vh + c F(x_{n+1}) / m
with c = dt / 2 in the reference's kick units (0.5 dt 4.184e-4,
verlet.cpp:49).  With a_x the total adjoint of x_{n+1} and a_v that of
v_{n+1}, a step back is
    a_v <- a_v + dt a_x;            u = a_v / m
    a_x <- a_x + k c H(x_n) u + dL/dx_n
    g   <- g + k c (dF(x_n)/dtheta)^T u
where k = 2 for 0 < n (x_n enters the second half kick of step n - 1 and the
first of step n with the same u) and k = 1 for n = 0.  H(x) u, the Hessian
of -U applied to u, is the derivative of the forces along u, computed
exactly (to FP64 rounding): the vi term's closed form (``mdn_vi_jvp``) with
the parameters theta, and the force kernels instantiated with forward-mode
dual numbers (``mdn_force_jvp``) for AMBER, vv and vapor intra; the Efield
force does not depend on x.  (dF/dtheta)^T u is the hand-written vi adjoint
with gF = u, one call for all steps of a checkpoint interval.  The forward
pass keeps (x, v) every ``every`` steps; the reverse pass recomputes each
interval's states from its checkpoint.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as M
from . import adjoint as A

KICK_UNIT = 4.184e-4  # verlet.cpp:49
HVP_STEP = 1e-5       # largest displacement of a finite-difference Hessian-vector product, A


class Integrator:
    """MD::verlet over ``mass`` [n] with time step ``dt``; a subclass gives
    ``forces(theta, X [S, n, 3])`` and ``theta_vjp(theta, frames, us)``
    (the sum over frames of (dF(frame)/dtheta)^T u), and may give an exact
    ``hvp``."""

    def setup(self, mass, dt):
        self.mass = np.asarray(mass, np.float64)
        self.dt = float(dt)
        self.c = 0.5 * self.dt * KICK_UNIT
        self.n_force_calls = 0
        self.n_hvp_calls = 0

    def hvp(self, theta, x, u):
        """H(x) u; here by central differences of the forces"""
        return self.hvp_fd(theta, x, u)

    def hvp_fd(self, theta, x, u):
        """H(x) u by central differences of the forces (e |u|_max = 1e-5 A)"""
        um = np.abs(u).max()
        if um == 0:
            return np.zeros_like(u)
        e = HVP_STEP / um
        F = self.forces(theta, np.stack([x + e * u, x - e * u]))
        return (F[0] - F[1]) / (2 * e)

    def run(self, theta, x0, v0, nsteps, f0=None):
        """states x_0..x_nsteps, v_0..v_nsteps (and the last force)"""
        xs, vs = [np.array(x0, np.float64)], [np.array(v0, np.float64)]
        f = self.forces(theta, xs[0][None])[0] if f0 is None else f0
        m = self.mass[:, None]
        for _ in range(nsteps):
            vh = vs[-1] + self.c * f / m
            x = xs[-1] + self.dt * vh
            f = self.forces(theta, x[None])[0]
            xs.append(x)
            vs.append(vh + self.c * f / m)
        return xs, vs, f


class Segment(Integrator):
    """The ion and vapor atoms of one replica, integrated by MD::verlet with
    forces from libmdn.  ``x``/``v`` arrays are [n, 3] over the ion atoms then
    the vapor atoms (molecule order), like the snapshots of ``adjoint``."""

    def __init__(self, sd, st, device=0):
        v = sd.view
        if v["shake"]:
            raise ValueError("the segment follows MD::verlet without SHAKE: load a condition without 'SHAKE H'")
        n0, n1, n2 = (s["mass"].size for s in v["species"])
        self.n_ion, self.n_vap = n0, v["n_vap"] * n2
        g0 = n0 + v["n_gas"] * n1
        self.idx = np.concatenate([np.arange(n0), np.arange(g0, g0 + self.n_vap)])
        self.setup(np.concatenate([v["species"][0]["mass"], np.tile(v["species"][2]["mass"], v["n_vap"])]), v["dt"])
        if not np.all(np.asarray(st.region[1 + v["n_gas"]:]) != 0):
            raise ValueError("every vapor molecule must be in the effective domain")
        if not np.all(st.atom_w[self.idx] == 1.0):
            raise ValueError("every vapor atom must have blend weight 1 (within 100 A of the ion)")
        self.rest_bits = v["flags"] & (M.MDN_ION_AMBER | M.MDN_ION_BMH | M.MDN_VV | M.MDN_VAPOR_INTRA | M.MDN_EFIELD)
        self.vi = (v["flags"] & M.MDN_VI) != 0
        sv, ev, qv = A._species_params(v, 2)
        self.vap = tuple(np.tile(a, v["n_vap"]) for a in (sv, ev, qv))
        self.atom_x = st.atom_x.reshape(-1, 3).copy()  # gas atoms stay where they are
        self.eng = M.Engine(device)
        self.eng.add(sd, st, seed=1)
        self.eng.commit()

    def close(self):
        self.eng.close()

    def state(self, st):
        """(x, v) of the segment's atoms in a replica state"""
        return st.atom_x.reshape(-1, 3)[self.idx].copy(), st.atom_v.reshape(-1, 3)[self.idx].copy()

    def _snap(self, X):
        S, n = X.shape[:2]
        return A.ViSnapshots(np.ascontiguousarray(X), np.ones((S, n)), np.ones((S, self.n_vap), np.uint8),
                             self.n_ion, *self.vap)

    def forces(self, theta, X):
        """F [S, n, 3] at the frames X [S, n, 3] (FP64)"""
        X = np.asarray(X, np.float64).reshape(-1, self.idx.size, 3)
        F = A.vi_forward(self._snap(X), *theta)[1] if self.vi else np.zeros_like(X)
        for s in range(X.shape[0]):
            full = self.atom_x.copy()
            full[self.idx] = X[s]
            self.eng.set_positions(0, full)
            self.eng.zero_forces(0)
            if self.rest_bits:
                self.eng.compute(0, self.rest_bits, 0, M.FP64)
            F[s] += self.eng.get_forces(0)[self.idx]
        self.n_force_calls += X.shape[0]
        return F

    def hvp(self, theta, x, u):
        """H(x) u, exact to rounding: the derivative of the forces along u
        (mdn_vi_jvp for vi with the parameters theta, mdn_force_jvp for the
        other position-dependent terms)"""
        self.n_hvp_calls += 1
        x = np.asarray(x, np.float64)
        u = np.asarray(u, np.float64)
        out = A.vi_jvp(self._snap(x[None]), *theta, u[None])[0] if self.vi else np.zeros_like(x)
        bits = self.rest_bits & ~M.MDN_EFIELD
        if bits:
            full = self.atom_x.copy()
            full[self.idx] = x
            self.eng.set_positions(0, full)
            U = np.zeros_like(full)
            U[self.idx] = u
            out = out + self.eng.force_jvp(0, bits, U)[self.idx]
        return out

    def theta_vjp(self, theta, frames, us):
        """sum over frames of (dF/dtheta)^T u: the hand-written vi adjoint"""
        if not self.vi:
            return tuple(np.zeros(self.n_ion) for _ in range(3))
        return A.vi_adjoint(self._snap(np.stack(frames)), *theta, gF=np.stack(us))


def trajectory_loss(xs, targets, obs):
    """sum over observed steps n of mean |x_n - target_n|^2 (A^2), and its
    gradient with respect to each x_n"""
    L, dL = 0.0, {}
    N = xs[0].shape[0]
    for n in obs:
        d = xs[n] - targets[n]
        L += float((d * d).sum()) / N
        dL[n] = 2.0 * d / N
    return L, dL


def loss_and_grad(seg, theta, x0, v0, nsteps, targets, obs, every=10):
    """(L, (dL/dsigma, dL/deps, dL/dq)) of the trajectory loss over one
    segment of a ``Segment``, on the device (``mdn_segment_grad``): reverse
    mode with checkpoints every ``every`` steps"""
    obs, tg = _targets(seg, x0, v0, targets, obs)
    return seg.eng.segment_grad(0, nsteps, every, theta, obs, tg, x0=x0, v0=v0)


def _targets(seg, x0, v0, targets, obs):
    """observed steps in ascending order and their targets [n_obs, N, 3],
    shapes checked against the segment (the C ABI trusts them)"""
    N = seg.idx.size
    for a in (x0, v0):
        if np.shape(a) != (N, 3):
            raise ValueError("segment state must be [%d, 3]" % N)
    obs = sorted(obs)
    tg = np.stack([np.asarray(targets[n], np.float64) for n in obs]) if obs else np.zeros((0, N, 3))
    if tg.shape != (len(obs), N, 3):
        raise ValueError("targets must be [%d, 3] per observed step" % N)
    return obs, tg


def segment_loss(seg, theta, x0, v0, nsteps, targets, obs):
    """the trajectory loss alone, on the device"""
    obs, tg = _targets(seg, x0, v0, targets, obs)
    return seg.eng.segment_grad(0, nsteps, max(nsteps, 1), theta, obs, tg, x0=x0, v0=v0, grad=False)[0]


def loss_and_grad_host(seg, theta, x0, v0, nsteps, targets, obs, every=10):
    """``loss_and_grad`` over host arrays: the same reverse pass with an
    ``Integrator``'s forces, Hessian-vector products and parameter adjoint"""
    theta = tuple(np.asarray(a, np.float64) for a in theta)
    # forward: checkpoints only
    ck = {0: (np.array(x0, np.float64), np.array(v0, np.float64))}
    x, v = ck[0]
    f = None
    L = 0.0
    N = x.shape[0]
    for k0 in range(0, nsteps, every):
        n = min(every, nsteps - k0)
        xs, vs, f = seg.run(theta, x, v, n, f0=f)
        for j in range(1, n + 1):
            if k0 + j in obs:
                d = xs[j] - targets[k0 + j]
                L += float((d * d).sum()) / N
        x, v = xs[-1], vs[-1]
        ck[k0 + n] = (x, v)
    if 0 in obs:
        d = ck[0][0] - targets[0]
        L += float((d * d).sum()) / N
    # reverse, one checkpoint interval at a time
    m = seg.mass[:, None]
    a_x = np.zeros_like(x)
    a_v = np.zeros_like(x)
    if nsteps in obs:
        a_x += 2.0 * (ck[nsteps][0] - targets[nsteps]) / N
    g = [np.zeros_like(np.asarray(a, np.float64)) for a in theta]
    starts = list(range(0, nsteps, every))
    for k0 in reversed(starts):
        n = min(every, nsteps - k0)
        xs, _, _ = seg.run(theta, ck[k0][0], ck[k0][1], n - 1) if n > 1 else ([ck[k0][0]], None, None)
        us, coef, frames = [], [], []
        for j in range(n - 1, -1, -1):
            step = k0 + j
            a_v = a_v + seg.dt * a_x
            u = a_v / m
            kc = (2.0 if step > 0 else 1.0) * seg.c
            a_x = a_x + kc * seg.hvp(theta, xs[j], u)
            if step in obs:
                a_x = a_x + 2.0 * (xs[j] - targets[step]) / N
            us.append(kc * u)
            frames.append(xs[j])
        gi = seg.theta_vjp(theta, frames, us)
        for t in range(3):
            g[t] += gi[t]
    return L, tuple(g)


def segment_loss_host(seg, theta, x0, v0, nsteps, targets, obs):
    xs, _, _ = seg.run(theta, x0, v0, nsteps)
    return trajectory_loss(xs, targets, obs)[0]


class TrajectoryLoss(torch.autograd.Function):
    """(sigma, eps, q) of the ion atoms -> the segment's trajectory loss"""

    @staticmethod
    def forward(ctx, sig, eps, q, seg, x0, v0, nsteps, targets, obs, every):
        theta = tuple(A._host(t) for t in (sig, eps, q))
        L, g = loss_and_grad(seg, theta, x0, v0, nsteps, targets, obs, every)
        ctx.g = g
        ctx.dev = sig.device
        return torch.tensor(L, dtype=torch.float64, device=sig.device)

    @staticmethod
    def backward(ctx, gL):
        s = float(gL)
        return tuple(torch.from_numpy(s * a).to(ctx.dev) for a in ctx.g) + (None,) * 7


def trajectory_loss_fn(sig, eps, q, seg, x0, v0, nsteps, targets, obs, every=10):
    return TrajectoryLoss.apply(sig, eps, q, seg, x0, v0, nsteps, targets, obs, every)
