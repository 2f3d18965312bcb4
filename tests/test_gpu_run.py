"""mdn_run (MD::run over the context) on the GPU against the reference's own
run (tests/golden/run.npz): every output file it writes, from the same
initial state.  FP64: all numbers within the printed precision (7
significant digits, dump columns to 2e-6 A), ids, counts and type names
identical.  Two replicas as calculation numbers 5 and 6 (main.cpp:25-29)
both reproduce it.  Mixed precision: the same files, numbers within 1e-3."""
import os

import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from tests import engine_util as E
from tests import golden_util as G
from tests import run_util as RU
from tests.golden.cases import write_case

pytestmark = pytest.mark.gpu


def _run(tmp_path, n_rep, calc0, precision):
    fx = G.load("run")
    sd = M.SysDef(write_case("run", str(tmp_path)))
    eng = M.Engine(0)
    for r in range(n_rep):
        eng.add(sd, E.state_from(sd, dict(G.state(fx), loopPair=0, loop_update=0)), seed=3, replica_id=r)
    eng.commit()
    eng.set_vapor_log(4096)
    eng.set_step(-1, 0)
    for r in range(n_rep):
        eng.make_pair(r)  # MD::MD's make_pair (md.cpp:34)
    out = tmp_path / "out"
    out.mkdir()
    print("mdn_run", n_rep, "replicas, precision", precision, flush=True)
    eng.run(str(out), calc0=calc0, precision=precision)
    want = {f: bytes(fx["file_" + f % 0]).decode() for f in RU.FILES}
    got = []
    for r in range(n_rep):
        g = {}
        for f in RU.FILES:
            p = out / (f % (calc0 + r))
            assert p.exists(), p
            g[f] = p.read_text()
        assert (out / ("vapor_collision_%d.dat" % (calc0 + r))).read_text() == ""
        got.append(g)
    return eng, sd, fx, got, want


def test_run_fp64_reproduces_reference_files(tmp_path):
    eng, sd, fx, got, want = _run(tmp_path, 1, 0, M.FP64)
    RU.compare_run_text(got[0], want)
    st = eng.get_state(0)
    assert E.rel(st.atom_x, fx["final_atom_x"]) < 1e-12
    np.testing.assert_array_equal(st.region, fx["final_region"].astype(np.uint8))


def test_run_two_replicas_as_calculation_numbers(tmp_path):
    eng, sd, fx, got, want = _run(tmp_path, 2, 5, M.FP64)
    for g in got:
        RU.compare_run_text(g, want)


def test_run_mixed_precision(tmp_path):
    eng, sd, fx, got, want = _run(tmp_path, 1, 0, M.MIXED)
    RU.compare_run_text(got[0], want, rel=1e-3, abs_e=1e-3, abs_f=1e-3)


def test_observe_matches_state(tmp_path):
    """mdn_observe against computeProps restated on the replica's state."""
    eng, sd, fx, got, want = _run(tmp_path, 2, 0, M.FP64)
    v = sd.view
    sp = v["species"]
    obs = eng.observe(-1)
    assert obs.shape == (2, 6)
    for r in range(2):
        st = eng.get_state(r)
        vi = st.atom_v[: 3 * sp[0]["mass"].size].reshape(-1, 3)
        r2k = 10000 / 4.184
        kion = np.sum(vi * vi * sp[0]["mass"][:, None]) * 0.5 * r2k
        mv = st.mol_v.reshape(-1, 3)
        kg = np.sum(mv[1:1 + v["n_gas"]] ** 2) * sp[1]["molmass"] * 0.5 * r2k
        kv = np.sum(mv[1 + v["n_gas"]:] ** 2) * sp[2]["molmass"] * 0.5 * r2k
        coeff = 1 / (1.38e-23 * 6.02e23 / 4184.0) / 1.5
        want_o = [kion, kg, kv, kion / sp[0]["mass"].size * coeff, kg / v["n_gas"] * coeff, kv / v["n_vap"] * coeff]
        np.testing.assert_allclose(obs[r], want_o, rtol=1e-12)
    np.testing.assert_allclose(eng.observe(1)[0], obs[1], rtol=0, atol=0)


def test_vapor_log_overflow_is_reported(tmp_path):
    fx = G.load("run")
    sd = M.SysDef(write_case("run", str(tmp_path)))
    eng = M.Engine(0)
    eng.add(sd, E.state_from(sd, dict(G.state(fx), loopPair=0, loop_update=0)), seed=3)
    eng.set_vapor_log(8)
    eng.commit()
    eng.make_pair(0)  # 20 water in the domain: 20 events, 8 kept
    with pytest.raises(M.MdnError, match="dropped"):
        eng.drain_vapor_log(0)
    assert eng.drain_vapor_log(0).size == 0
    eng.make_pair(0)
    eng.set_vapor_log(64)  # re-upload keeps the state; undrained events are gone
    eng.make_pair(0)
    ev = eng.drain_vapor_log(0)
    assert ev.size == 20 and (ev["kind"] == 0).all()
    assert sorted(ev["mol"].tolist()) == list(range(201, 221))
