"""The differentiable vi term (force-matching workload): the 150-atom AMBER
chain with 200 water packed around it, FP64 snapshots of one trajectory of
that system (all 750 ion + water atoms interacting).

  forward      mdn_vi_forward E, F  ==  the oracle's vi term (VaporIon.cpp)
               on the same snapshots (FP64, 1e-12)
  adjoint      mdn_vi_adjoint  ==  central differences (h = 1e-6, FP64) of
               (i) the GPU forward and (ii) the oracle's energies / forces
               with the perturbed parameters mixed into the oracle's pair
               table by the Lorentz-Berthelot rule, for the energy and for a
               force-matching loss against targets from a +-10% perturbation
  autograd     ViTerm.backward  ==  mdn_vi_adjoint, and a fit lowers the loss
"""
import numpy as np
import pytest
import torch

from md_nucleation_amd import _lib as M
from md_nucleation_amd import adjoint as A
from md_nucleation_amd import configs
from oracle import oracle as O
from tests import engine_util as E

pytestmark = pytest.mark.gpu

H = 1e-6


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("fm"))
    sd = M.SysDef(configs.config5(d, n_gas=5000, n_vap=200, seed=1))
    seed = 11
    st = sd.init_state(seed)
    print("snapshots: chain + 200 water, 256 frames x 10 steps", flush=True)
    snaps = A.snapshots(sd, st, seed, n=256, every=10)
    return sd, snaps


def _oracle_vi(sd, snap, k, sig=None, eps=None, q=None):
    """The oracle's vi term on frame k with optional ion parameters: returns
    (E, F[n_ion + n_vap, 3]) in the snapshot's atom order."""
    v = dict(sd.view)
    species = [dict(s) for s in v["species"]]
    n0, n1, n2 = (s["mass"].size for s in species)
    pc = np.array(v["pair_coeff"], np.float64).reshape(-1, 2)
    if sig is not None:
        nt = int(round(np.sqrt(pc.shape[0])))
        ts, te = np.array(v["type_sigma"], np.float64), np.array(v["type_eps"], np.float64)
        it = np.asarray(species[0]["type"], np.int64)
        ts[it], te[it] = sig, eps
        vt = np.unique(np.asarray(species[2]["type"], np.int64))
        pc = pc.reshape(nt, nt, 2)
        for a in it:
            for b in vt:
                e, s = np.sqrt(te[a] * te[b]), (ts[a] + ts[b]) / 2
                pc[a, b] = pc[b, a] = (48 * e * s ** 12, 24 * e * s ** 6)
        pc = pc.reshape(-1)
    if q is not None:
        species[0]["charge"] = np.asarray(q, np.float64)
    v["species"] = species
    v["pair_coeff"] = pc.reshape(-1)
    orc = O.Oracle(E.oracle_sysd(v))
    ng, nv = v["n_gas"], v["n_vap"]
    na = n0 + ng * n1 + nv * n2
    ax = np.zeros((na, 3))
    aw = np.zeros(na)
    g0 = n0 + ng * n1
    ax[:n0], aw[:n0] = snap.x[k, :n0], snap.w[k, :n0]
    ax[g0:], aw[g0:] = snap.x[k, n0:], snap.w[k, n0:]
    # gas atoms far apart and away from everything (the vi term ignores them)
    ax[n0:g0] = 1e4 + np.arange(ng * n1)[:, None] * 10.0
    region = np.zeros(1 + ng + nv, np.uint8)
    region[0] = 1
    region[1 + ng:] = snap.vap_in[k, ::n2]
    S = orc.new_state({"atom_x": ax, "atom_w": aw, "region": region,
                       "mol_x": np.zeros(3 * (1 + ng + nv)), "mol_v": np.zeros(3 * (1 + ng + nv))}, cap=16)
    orc.force_term(S, O.ORC_VI, 1)
    f = S.atom_f.reshape(-1, 3)
    return S.U[6], np.concatenate([f[:n0], f[g0:]])


def test_snapshots_are_live(system):
    sd, snaps = system
    assert snaps.x.shape == (256, 150 + 600, 3)
    n_in = snaps.vap_in.sum(1) // 3
    print("in-domain water per frame: min %d max %d" % (n_in.min(), n_in.max()))
    assert n_in.min() == 200
    # frames differ (a trajectory, not copies)
    assert np.abs(snaps.x[-1] - snaps.x[0]).max() > 0.5


def test_forward_matches_oracle(system):
    sd, snaps = system
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    En, Fn = A.vi_forward(snaps, sig, eps, q)
    for k in (0, 1, 77, 255):
        Eo, Fo = _oracle_vi(sd, snaps, k)
        print("frame", k, "E", En[k], Eo, "rel F err", E.rel(Fn[k], Fo))
        assert abs(En[k] - Eo) <= 1e-12 * max(1.0, abs(Eo))
        assert E.rel(Fn[k], Fo) < 1e-12
        assert np.abs(Fo).max() > 1e-3


def test_pair_table_reconstruction(system):
    """The test's Lorentz-Berthelot re-mixing reproduces the parsed pair
    table at the unperturbed parameters (so the oracle FD below perturbs
    exactly the ion's parameters)."""
    sd, snaps = system
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    for k in (3, 200):
        e0, f0 = _oracle_vi(sd, snaps, k)
        e1, f1 = _oracle_vi(sd, snaps, k, sig, eps, q)
        assert abs(e0 - e1) <= 1e-12 * abs(e0)
        assert E.rel(f1, f0) < 1e-13


def _targets(sd, snaps, seed=5):
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    rng = np.random.default_rng(seed)
    pert = [a * (1 + rng.uniform(-0.1, 0.1, a.size)) for a in (sig, eps, q)]
    return (sig, eps, q), pert, A.vi_forward(snaps, *pert)[1]


def _loss_fm(F, Ft):
    return float(((F - Ft) ** 2).sum())


def test_adjoint_vs_gpu_finite_differences(system):
    sd, snaps = system
    theta, _, Ft = _targets(sd, snaps)
    En, Fn = A.vi_forward(snaps, *theta)
    # energy loss with per-frame weights and the force-matching loss
    c = np.random.default_rng(2).uniform(0.5, 1.5, snaps.S)
    gE_e = A.vi_adjoint(snaps, *theta, gE=c)
    gE_f = A.vi_adjoint(snaps, *theta, gF=2 * (Fn - Ft))
    rng = np.random.default_rng(3)
    atoms = rng.choice(snaps.n_ion, 12, replace=False)
    worst = 0.0
    for p in range(3):
        for i in atoms:
            tp = [a.copy() for a in theta]
            tm = [a.copy() for a in theta]
            tp[p][i] += H
            tm[p][i] -= H
            Ep, Fp = A.vi_forward(snaps, *tp)
            Em, Fm = A.vi_forward(snaps, *tm)
            fd_e = (c @ Ep - c @ Em) / (2 * H)
            fd_f = (_loss_fm(Fp, Ft) - _loss_fm(Fm, Ft)) / (2 * H)
            for g, fd, scale in ((gE_e[p][i], fd_e, np.abs(gE_e[p]).max()), (gE_f[p][i], fd_f, np.abs(gE_f[p]).max())):
                err = abs(g - fd) / scale
                worst = max(worst, err)
                assert err < 1e-6, (p, i, g, fd, scale)
    print("worst FD mismatch (relative to the largest gradient component):", worst)


def test_adjoint_vs_oracle_finite_differences(system):
    """The reference-derived gradient: central differences of the oracle's
    energy and forces (the reference's vi loop restated) under the same
    per-atom perturbations, on a subset of frames."""
    sd, snaps = system
    theta, _, Ft_all = _targets(sd, snaps)
    frames = np.array([0, 50, 128, 255])
    sub = snaps.subset(frames)
    Ft = Ft_all[frames]
    c = np.array([1.0, 0.7, 1.3, 0.9])
    F0 = np.stack([_oracle_vi(sd, sub, k, *theta)[1] for k in range(len(frames))])
    gE_e = A.vi_adjoint(sub, *theta, gE=c)
    gE_f = A.vi_adjoint(sub, *theta, gF=2 * (F0 - Ft))
    for p, i in ((0, 0), (0, 75), (1, 10), (1, 149), (2, 3), (2, 100)):
        vals = []
        for sgn in (1, -1):
            t = [a.copy() for a in theta]
            t[p][i] += sgn * H
            out = [_oracle_vi(sd, sub, k, *t) for k in range(len(frames))]
            vals.append((np.array([o[0] for o in out]), np.stack([o[1] for o in out])))
        fd_e = (c @ vals[0][0] - c @ vals[1][0]) / (2 * H)
        fd_f = (_loss_fm(vals[0][1], Ft) - _loss_fm(vals[1][1], Ft)) / (2 * H)
        print("param", p, "atom", i, "dE", gE_e[p][i], fd_e, "dL", gE_f[p][i], fd_f)
        assert abs(gE_e[p][i] - fd_e) <= 1e-6 * np.abs(gE_e[p]).max()
        assert abs(gE_f[p][i] - fd_f) <= 1e-6 * np.abs(gE_f[p]).max()


def test_autograd_function(system):
    sd, snaps = system
    sub = snaps.subset(np.arange(0, 256, 8))
    theta, _, Ft_all = _targets(sd, snaps)
    Ft = torch.from_numpy(Ft_all[::8].copy())
    sig, eps, q = A.ion_params(sd, requires_grad=True)
    En, Fn = A.vi_energy_forces(sig, eps, q, sub)
    loss = ((Fn - Ft) ** 2).sum() + 0.01 * En.sum()
    loss.backward()
    F0 = Fn.detach().numpy()
    want = A.vi_adjoint(sub, *theta, gE=np.full(sub.S, 0.01), gF=2 * (F0 - Ft.numpy()))
    for t, w in zip((sig, eps, q), want):
        np.testing.assert_allclose(t.grad.numpy(), w, rtol=1e-13, atol=1e-13 * np.abs(w).max())


def test_fit_lowers_force_matching_loss(system):
    """Starting from the true parameters, L-BFGS on the force-matching loss
    moves toward the perturbed parameters that made the targets (sigma and
    eps through log-multipliers, so they stay positive)."""
    sd, snaps = system
    sub = snaps.subset(np.arange(0, 256, 16))
    _, pert, Ft_all = _targets(sd, snaps)
    Ft = torch.from_numpy(Ft_all[::16].copy())
    sig0, eps0, q0 = A.ion_params(sd)
    a = torch.zeros_like(sig0, requires_grad=True)
    b = torch.zeros_like(eps0, requires_grad=True)
    q = q0.clone().requires_grad_(True)
    opt = torch.optim.LBFGS([a, b, q], lr=1.0, max_iter=60, history_size=20, line_search_fn="strong_wolfe")
    scale = float((Ft ** 2).sum())

    def closure():
        opt.zero_grad()
        _, F = A.vi_energy_forces(sig0 * torch.exp(a), eps0 * torch.exp(b), q, sub)
        loss = ((F - Ft) ** 2).sum() / scale
        loss.backward()
        return loss

    l0 = float(closure().detach())
    opt.step(closure)
    l1 = float(closure().detach())
    dq0 = np.abs(q0.numpy() - pert[2]).mean()
    dq1 = np.abs(q.detach().numpy() - pert[2]).mean()
    print("force-matching loss %.3e -> %.3e; mean |q - q_target| %.3e -> %.3e" % (l0, l1, dq0, dq1))
    assert np.isfinite(l1) and l1 < 0.1 * l0
    assert dq1 < dq0
