"""Tersoff and Stillinger-Weber ions (Tersoff.cpp:13-297,
StillingerWeber.cpp:11-198) stepped by the batched kernel and by the per-call
path: a 64-atom diamond-lattice Si cluster, 300 FP64 steps against the oracle
(the single-evaluation forces of both terms are pinned against the
reference's own code by the golden 'si' case), and the mixed-precision step
against the oracle's forces (relative RMS <= 1e-5)."""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from tests import engine_util as E
from tests.golden.cases import write_case

pytestmark = pytest.mark.gpu

SEEDS = (3, 4)


def _case(tmp_path, op, path, nsteps, precision):
    sd = M.SysDef(write_case(op, str(tmp_path)))
    eng = M.Engine(0)
    orcs, Ss = [], []
    for r, seed in enumerate(SEEDS):
        st0 = sd.init_state(seed)
        eng.add(sd, st0, seed=seed, replica_id=r)
        orc = E.oracle_for(sd, seed=seed, replica=r)
        S = orc.new_state(st0.as_dict())
        orcs.append(orc)
        Ss.append(S)
    eng.commit()
    for r in range(len(SEEDS)):
        eng.make_pair(r)
        assert orcs[r].make_pair(Ss[r]) == 0
    print("case: %s %s %d steps precision %d" % (op, path, nsteps, precision), flush=True)
    if path == "percall":
        for r in range(len(SEEDS)):
            eng.verlet(r, nsteps, eflag_every=10, precision=precision)
    else:
        eng.run_batched(nsteps, eflag_every=10, precision=precision)
    eng.sync()
    for r in range(len(SEEDS)):
        assert orcs[r].verlet(Ss[r], nsteps, eflag_every=10)[0] == 0
    return sd, eng, Ss


@pytest.mark.parametrize("path", ["batched", "percall"])
@pytest.mark.parametrize("op", ["tersoff", "sw"])
def test_manybody_trajectory_vs_oracle(tmp_path, op, path):
    sd, eng, Ss = _case(tmp_path, op, path, 300, M.FP64)
    n = sd.view["species"][0]["mass"].size
    assert n == 64
    for r in range(len(SEEDS)):
        st = eng.get_state(r)
        S = Ss[r]
        dx = np.abs(st.atom_x[: 3 * n] - S.atom_x[: 3 * n]).max()
        dv = np.abs(st.atom_v[: 3 * n] - S.atom_v[: 3 * n]).max()
        fe = E.rel(st.atom_f[: 3 * n], S.atom_f[: 3 * n])
        print("%s %s replica %d: max |dx| %.2e A max |dv| %.2e rel F %.2e U_ion %.6f"
              % (op, path, r, dx, dv, fe, st.U[0]))
        assert dx < 1e-8 and dv < 1e-10 and fe < 1e-7
        np.testing.assert_allclose(st.U, S.U, rtol=1e-6, atol=1e-6)
        # the cluster moves: 300 steps are not a trivial check
        assert np.abs(S.atom_x[: 3 * n] - sd.init_state(SEEDS[r]).atom_x[: 3 * n]).max() > 1e-3


@pytest.mark.parametrize("op", ["tersoff", "sw"])
def test_manybody_batched_mixed_step(tmp_path, op):
    sd, eng, Ss = _case(tmp_path, op, "batched", 1, M.MIXED)
    n = sd.view["species"][0]["mass"].size
    for r in range(len(SEEDS)):
        st = eng.get_state(r)
        fe = E.rel(st.atom_f[: 3 * n], Ss[r].atom_f[: 3 * n])
        print("%s mixed replica %d: rel F %.2e" % (op, r, fe))
        assert fe < 1e-5
