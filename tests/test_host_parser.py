"""Host front end (libmdn's condition/molecule parser and parameter tables)
against the reference's own parse of the same files (golden sys dumps).
Host-only entry points: no GPU needed."""
import re

import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from tests import golden_util as G
from tests.golden.cases import write_case

ROOT_H = __import__("os").path.join(G.GOLDEN, "..", "..", "include", "mdn.h")


def test_library_exports_every_declared_symbol():
    text = open(ROOT_H).read()
    names = set(re.findall(r"\b(mdn_[a-z_0-9]+)\s*\(", text))
    L = M.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(M._SIGS) >= names


@pytest.mark.parametrize("name,fixture,prefix", [("c1", "c1", "sys_"), ("pep", "pep", "sys_"),
                                                 ("bmh", "bmh", "sys_"), ("tersoff", "si", "tersoff_sys_"),
                                                 ("sw", "si", "sw_sys_")])
def test_parse_matches_reference(tmp_path, name, fixture, prefix):
    fx = G.sub(G.load(fixture), prefix)
    sd = M.SysDef(write_case(name, str(tmp_path)))
    v = sd.view
    np.testing.assert_array_equal(v["pair_coeff"], fx["pair_coeff"])
    np.testing.assert_array_equal(v["type_eps"], fx["type_eps"])
    np.testing.assert_array_equal(v["type_sigma"], fx["type_sigma"])
    np.testing.assert_array_equal(v["btype"], fx["btype"])
    np.testing.assert_array_equal(v["ctype"], fx["ctype"])
    np.testing.assert_array_equal(v["dcoeff"], fx["dcoeff"])
    np.testing.assert_array_equal(v["dmulti"], fx["dmulti"])
    for k, p in enumerate(("ion", "gas", "vap")):
        sp = v["species"][k]
        if fx[p + "_mass"].size == 0 and sp["mass"].size == 0:
            continue
        for key in ("mass", "charge", "type", "bond", "angle", "dih", "rigid"):
            if p == "gas" and key == "mass" and v["gastype"] in (1, 3, 4):
                # the reference never records He / N2monoatomic / Ar atom masses
                # (setCondition.cpp:48-113); the parser gives them their mass
                assert sp["mass"].size == sp["type"].size and np.all(sp["mass"] > 0)
                continue
            np.testing.assert_array_equal(sp[key].astype(np.float64), fx[p + "_" + key], err_msg=p + key)
        if p != "ion":
            np.testing.assert_array_equal(sp["molmass"], fx[p + "_molmass"][0])
    np.testing.assert_array_equal(v["species"][1]["tmpl_x"], fx["gas_tmpl"])
    np.testing.assert_array_equal(v["species"][2]["tmpl_x"], fx["vap_tmpl"])
    sc = fx["scalars"]
    got = [v["dt"], v["T"], v["pressure"], v["d_size"], v["cutoff"], v["margin"], v["ML2"], v["m_gas_kg"],
           v["m_vap_kg"], v["gastype"], v["n_gas"], v["n_vap"]]
    np.testing.assert_array_equal(np.array(got, np.float64), sc)
    bits = [M.MDN_GI, M.MDN_VI, M.MDN_GG, M.MDN_GV, M.MDN_VV, M.MDN_ION_AMBER, M.MDN_ION_BMH, M.MDN_ION_TERSOFF,
            M.MDN_ION_SW, M.MDN_EFIELD, M.MDN_VAPOR_INTRA, M.MDN_GAS_INTRA]
    np.testing.assert_array_equal([1.0 if v["flags"] & b else 0.0 for b in bits], fx["flags"])


def test_parse_errors_are_reported(tmp_path):
    p = tmp_path / "bad.cond"
    p.write_text("Input missing_file.txt\n")
    with pytest.raises(M.MdnError, match="cannot open"):
        M.SysDef(str(p))
    p.write_text("Temperature 300\n")
    with pytest.raises(M.MdnError, match="no Input line"):
        M.SysDef(str(p))


@pytest.mark.parametrize("name", ["nvt_scale", "nvt_nh"])
def test_parse_nvtion_matches_reference(tmp_path, name):
    """NVTion scale / NVTion 250 as the reference parsed them
    (setCondition.cpp:158-172): (mode, Tnh_ion)."""
    fx = G.sub(G.load("nvt"), name + "_sys_")
    v = M.SysDef(write_case(name, str(tmp_path))).view
    np.testing.assert_array_equal([v["nvt_ion"], v["Tnh_ion"]], fx["nvt"])


def test_parse_nvtion_values(tmp_path):
    from md_nucleation_amd import configs
    m = configs.molecule_files(str(tmp_path))
    for text, mode, T in (("OFF", 0, 0.0), ("scale", 1, 0.0), ("300", 2, 300.0), ("412.5", 2, 412.5)):
        c = configs.Condition(m["na"], m["methanol"], 2, "N2", 10, nvt_ion=text).write(str(tmp_path / "n.cond"))
        v = M.SysDef(c).view
        assert (v["nvt_ion"], v["Tnh_ion"]) == (mode, T), text
    c = configs.Condition(m["na"], m["methanol"], 2, "N2", 10, nvt_ion="warm").write(str(tmp_path / "n.cond"))
    with pytest.raises(M.MdnError, match="NVTion"):
        M.SysDef(c)


def test_glycol_shake_constraints(tmp_path):
    """The ethylene glycol writer: under SHAKE H its six X-H bonds are the
    rigid pairs (four groups that share no atom), the template meets every
    constraint length, and the oracle's SHAKE keeps them through 20 steps."""
    from md_nucleation_amd import configs
    from tests import engine_util as E
    m = configs.molecule_files(str(tmp_path))
    cond = configs.Condition(m["na"], m["glycol"], 6, "N2", 100, shake_h=True).write(str(tmp_path / "g.cond"))
    sd = M.SysDef(cond)
    v = sd.view
    sp = v["species"][2]
    rig = sp["rigid"].reshape(-1, 3)
    assert sp["mass"].size == 10 and rig.shape[0] == 6
    # atom-disjoint groups of the constraint graph
    groups = []
    for i, j, _ in rig:
        hit = [g for g in groups if i in g or j in g]
        for g in hit:
            groups.remove(g)
        groups.append(set().union({int(i), int(j)}, *hit))
    assert len(groups) == 4
    tmpl = sp["tmpl_x"].reshape(-1, 3)
    for i, j, t in rig:
        assert abs(np.linalg.norm(tmpl[i] - tmpl[j]) - v["btype"][2 * t + 1]) < 1e-12
    st = sd.init_state(5)
    orc = E.oracle_for(sd, seed=5)
    S = orc.new_state(st.as_dict())
    orc.make_pair(S)
    rc, _ = orc.verlet(S, 20, eflag_every=10)
    assert rc == 0
    X = S.atom_x.reshape(-1, 3)
    off = v["species"][0]["mass"].size + v["n_gas"] * v["species"][1]["mass"].size
    for mol in np.flatnonzero(S.region[1 + v["n_gas"]:]):
        o = off + mol * 10
        for i, j, t in rig:
            d = np.linalg.norm(X[o + i] - X[o + j])
            assert abs(d * d - v["btype"][2 * t + 1] ** 2) < 1e-9
