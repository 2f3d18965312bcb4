"""Host side of the differentiable vi term (no device calls): frame
extraction, parameter tables and argument checks."""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from md_nucleation_amd import adjoint as A
from md_nucleation_amd import configs


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("fm_host"))
    sd = M.SysDef(configs.config5(d, n_gas=200, n_vap=200, seed=1))
    return sd, sd.init_state(11)


def test_packed_water_all_in_domain(fm):
    sd, st = fm
    x, w, vin = A.frame(sd, st)
    assert x.shape == (750, 3) and w.shape == (750,) and vin.shape == (600,)
    assert vin.all()
    # ion atoms first, water after, in the state's atom order
    v = sd.view
    g0 = 150 + v["n_gas"] * 2
    np.testing.assert_array_equal(x[150:], st.atom_x.reshape(-1, 3)[g0:])
    # no water atom closer than 1.5 A to a chain atom
    d = np.linalg.norm(x[150:, None] - x[None, :150], axis=-1)
    assert d.min() > 1.5


def test_ion_params_follow_the_type_table(fm):
    sd, _ = fm
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    v = sd.view
    t = np.asarray(v["species"][0]["type"])
    assert len(set(t.tolist())) == 150  # one type per chain atom
    np.testing.assert_array_equal(sig, v["type_sigma"][t])
    np.testing.assert_array_equal(eps, v["type_eps"][t])
    assert abs(q.sum() - 1.0) < 1e-9
    assert ((sig > 2.5) & (sig < 3.9)).all() and ((eps > 0.02) & (eps < 0.2)).all()


def test_shape_errors_are_raised_before_any_device_call(fm):
    sd, st = fm
    x, w, vin = A.frame(sd, st)
    sv, ev, qv = A._species_params(sd.view, 2)
    snap = A.ViSnapshots(x[None], w[None], vin[None], 150, np.tile(sv, 200), np.tile(ev, 200), np.tile(qv, 200))
    sig, eps, q = (t.numpy() for t in A.ion_params(sd))
    with pytest.raises(ValueError):
        A.vi_forward(snap, sig[:-1], eps, q)
    bad = A.ViSnapshots(x[None], w[None, :-1], vin[None], 150, snap.vap_sigma, snap.vap_eps, snap.vap_q)
    with pytest.raises(ValueError):
        A.vi_adjoint(bad, sig, eps, q, gE=np.ones(1))
