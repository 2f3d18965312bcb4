"""The 12,001-atom cluster (Na+ + 2,000 methanol packed at liquid density,
100,000 Ar) through the per-call path: all-pairs vv (tiled kernel), vi, gi,
gv, intra terms, SHAKE; against the oracle on the same state."""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from md_nucleation_amd import workloads as W
from tests import engine_util as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cluster(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("cluster"))
    sd, st0, seed = W.cluster(d)
    # every atom lies within 100 A of the ion: weight 1 (what the first
    # update_position gives; molecules entering at a rebuild keep their stale
    # weight until then, which would switch vi/vv off in this snapshot)
    st0.atom_w[:] = 1.0
    orc = E.oracle_for(sd, seed=seed)
    S = orc.new_state(st0.as_dict())
    orc.make_pair(S)
    return sd, st0, seed, orc, S


def _engine(sd, st0, seed):
    eng = M.Engine(0)
    eng.add(sd, st0, seed=seed)
    eng.commit()
    eng.make_pair(0)
    return eng


@pytest.mark.parametrize("precision,tol", [(M.FP64, 1e-12), (M.MIXED, 1e-5)])
def test_cluster_forces(cluster, precision, tol):
    sd, st0, seed, orc, S = cluster
    eng = _engine(sd, st0, seed)
    st = eng.get_state(0)
    assert int(st.region.sum()) == int(S.region.sum()) > 2000
    np.testing.assert_array_equal(st.region, S.region)
    eng.zero_forces(0)
    eng.compute(0, sd.view["flags"], 1, precision)
    out = eng.get_state(0)
    orc.forces(S, eflag=1)
    m = E.in_atom_mask(sd, out.region)
    err = E.rel(out.atom_f.reshape(-1, 3)[m], S.atom_f.reshape(-1, 3)[m])
    print("cluster precision", precision, "rel RMS force error", err)
    assert err < tol
    # energies: FP64 sums of (FP32 in mixed mode) pair terms over ~7e7 pairs
    np.testing.assert_allclose(out.U, S.U, rtol=tol, atol=1e-6 if precision == M.FP64 else 2.0)
    # the all-pairs vv and vi terms are live in this snapshot
    assert abs(S.U[7]) > 100 and abs(S.U[6]) > 1


def test_cluster_steps_fp64(cluster):
    sd, st0, seed, orc, _ = cluster
    S = orc.new_state(st0.as_dict())
    orc.make_pair(S)
    eng = _engine(sd, st0, seed)
    eng.verlet(0, 5, eflag_every=5, precision=M.FP64)
    out = eng.get_state(0)
    orc.verlet(S, 5, eflag_every=5)
    m = E.in_atom_mask(sd, out.region)
    dx = np.abs(out.atom_x.reshape(-1, 3)[m] - S.atom_x.reshape(-1, 3)[m]).max()
    assert dx < 1e-10
    np.testing.assert_allclose(out.U, S.U, rtol=1e-9, atol=1e-6)
