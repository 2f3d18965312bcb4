"""The reverse pass of the trajectory-segment gradient (segment.py) on the
CPU: a small numpy pair model integrated by the same MD::verlet stepper, its
loss gradient by the checkpointed reverse pass against central differences
of the loss, and checkpoint spacing not changing the result."""
import numpy as np

from md_nucleation_amd import segment as SG


class Toy(SG.Integrator):
    """6 atoms, LJ (eps_ij [(s/r)^12 - 2 (s/r)^6], LB mixing) + Coulomb;
    theta = per-atom (sigma, eps, q) of the first two atoms"""

    def __init__(self):
        self.setup([12.0, 16.0, 1.0, 14.0, 1.0, 12.0], 1.0)
        self.base = (np.array([3.0, 3.2, 2.5, 3.4, 2.4, 3.1]), np.array([0.2, 0.15, 0.05, 0.1, 0.03, 0.12]),
                     np.array([0.4, -0.6, 0.3, -0.2, 0.25, -0.15]))

    def params(self, theta):
        s, e, q = (b.copy() for b in self.base)
        s[:2], e[:2], q[:2] = theta
        return s, e, q

    def _force(self, theta, x):
        s, e, q = self.params(theta)
        F = np.zeros_like(x)
        for i in range(6):
            for j in range(i + 1, 6):
                d = x[i] - x[j]
                r2 = d @ d
                sg = 0.5 * (s[i] + s[j])
                ep = np.sqrt(e[i] * e[j])
                u = sg * sg / r2
                u3 = u ** 3
                fr = (12 * ep * (u3 * u3 - u3) + 332.0 * q[i] * q[j] / np.sqrt(r2)) / r2
                F[i] += fr * d
                F[j] -= fr * d
        return F

    def forces(self, theta, X):
        self.n_force_calls += len(X)
        return np.stack([self._force(theta, x) for x in X])

    def theta_vjp(self, theta, frames, us, h=1e-7):
        out = [np.zeros(2) for _ in range(3)]
        for p in range(3):
            for k in range(2):
                tp = [np.array(t, float) for t in theta]
                tm = [np.array(t, float) for t in theta]
                tp[p][k] += h
                tm[p][k] -= h
                out[p][k] = sum(((self._force(tp, x) - self._force(tm, x)) * u).sum() for x, u in zip(frames, us)) / (2 * h)
        return tuple(out)


def _setup():
    t = Toy()
    rng = np.random.default_rng(4)
    x0 = np.array([[0, 0, 0], [3.5, 0.3, 0], [0.2, 3.3, 0.4], [-3.4, 0.5, 0.2], [0.3, -3.2, -0.3], [0.4, 0.2, 3.6]])
    v0 = rng.normal(0, 0.002, (6, 3))
    theta = (np.array([3.0, 3.2]), np.array([0.2, 0.15]), np.array([0.4, -0.6]))
    pert = tuple(a * (1 + 0.08 * np.array([1, -1])) for a in theta)
    xs_t, _, _ = t.run(pert, x0, v0, 30)
    obs = {10, 20, 25, 30}
    return t, x0, v0, theta, {n: xs_t[n] for n in obs}, obs


def test_reverse_pass_matches_central_differences():
    t, x0, v0, theta, tgt, obs = _setup()
    L, g = SG.loss_and_grad_host(t, theta, x0, v0, 30, tgt, obs, every=7)
    assert L > 0
    h = 1e-6
    worst = 0.0
    for p in range(3):
        for k in range(2):
            tp = [a.copy() for a in theta]
            tm = [a.copy() for a in theta]
            tp[p][k] += h
            tm[p][k] -= h
            fd = (SG.segment_loss_host(t, tp, x0, v0, 30, tgt, obs) - SG.segment_loss_host(t, tm, x0, v0, 30, tgt, obs)) / (2 * h)
            scale = max(abs(fd), 1e-3 * max(abs(a).max() for a in g))
            worst = max(worst, abs(g[p][k] - fd) / scale)
    assert worst < 1e-5, worst


def test_checkpoint_spacing_does_not_change_the_gradient():
    t, x0, v0, theta, tgt, obs = _setup()
    L1, g1 = SG.loss_and_grad_host(t, theta, x0, v0, 30, tgt, obs, every=1)
    L2, g2 = SG.loss_and_grad_host(t, theta, x0, v0, 30, tgt, obs, every=10)
    L3, g3 = SG.loss_and_grad_host(t, theta, x0, v0, 30, tgt, obs, every=30)
    assert L1 == L2 == L3
    for a, b, c in zip(g1, g2, g3):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(a, c, rtol=1e-12, atol=1e-15)
