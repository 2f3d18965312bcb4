"""NVE energy drift in a closed, fully weighted, rebuild-free subsystem (the
only setting where the reference's dynamics conserve energy: weights 1, no
resampling): the GPU paths' drift against the oracle's on the same start."""
import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from md_nucleation_amd import configs
from tests import engine_util as E

pytestmark = pytest.mark.gpu

KCAL = 1.0 / 4.184e-4  # g/mol A^2/fs^2 -> kcal/mol


def kinetic(sd, st):
    v = sd.view
    m = np.concatenate([v["species"][0]["mass"]] + [v["species"][1]["mass"]] * v["n_gas"]
                       + [v["species"][2]["mass"]] * v["n_vap"])
    vel = st.atom_v.reshape(-1, 3)
    mask = E.in_atom_mask(sd, st.region)
    return 0.5 * KCAL * float((m[mask, None] * vel[mask] ** 2).sum())


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("nve"))
    mf = configs.molecule_files(d)
    # 20 N2 in a 100 A box: every atom within 100 A of the ion (weight 1);
    # Margin 1000 A: no rebuild after the first; no SHAKE; dt 0.5 fs
    cond = configs.Condition(mf["na"], mf["methanol"], 10, "N2", 20, pressure=8.28e4, dt=0.5, margin=1000.0,
                             shake_h=False).write(d + "/nve.cond")
    sd = M.SysDef(cond)
    st0 = sd.init_state(5)
    st0.atom_w[:] = 1.0
    return sd, st0


def _energy_series(run, sd, eng, nchunks, every):
    E_ = []
    for k in range(nchunks):
        run(k)
        st = eng.get_state(0)
        E_.append(kinetic(sd, st) + float(st.U.sum()))
    return np.array(E_)


@pytest.mark.parametrize("path,precision", [("percall", M.FP64), ("batched", M.FP64), ("batched", M.MIXED)])
def test_nve_drift(system, path, precision):
    sd, st0 = system
    every, nchunks = 100, 20
    # oracle
    orc = E.oracle_for(sd, seed=5)
    S = orc.new_state(st0.as_dict())
    orc.make_pair(S)
    E_ref = []
    for k in range(nchunks):
        orc.verlet(S, every, eflag_every=every, step0=k * every)
        E_ref.append(kinetic(sd, S) + float(S.U.sum()))
    E_ref = np.array(E_ref)
    eng = M.Engine(0)
    eng.add(sd, st0, seed=5)
    eng.commit()
    eng.make_pair(0)

    def run(k):
        if path == "percall":
            eng.verlet(0, every, eflag_every=every, step0=k * every, precision=precision)
        else:
            eng.run_batched(every, eflag_every=every, step0=k * every, precision=precision)

    E_gpu = _energy_series(run, sd, eng, nchunks, every)
    assert eng.counters(0)[1] == 1  # no rebuild after the first
    drift_ref = E_ref[-1] - E_ref[0]
    drift_gpu = E_gpu[-1] - E_gpu[0]
    spread = E_ref.max() - E_ref.min()
    print(path, precision, "E0 %.6f drift oracle %.3e gpu %.3e spread %.3e" % (E_ref[0], drift_ref, drift_gpu,
                                                                                   spread))
    assert abs(E_gpu[0] - E_ref[0]) < 1e-6 * max(1.0, abs(E_ref[0]))
    if precision == M.FP64:
        np.testing.assert_allclose(E_gpu, E_ref, rtol=0, atol=1e-6 * max(1.0, abs(E_ref[0])))
    else:
        # mixed: drift within the FP64 run's own energy fluctuation (x3)
        assert abs(drift_gpu) <= 3 * max(abs(drift_ref), spread)
