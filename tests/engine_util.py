"""Shared set-up for the GPU tests: one SysDef feeds both libmdn and the
oracle, so both run on identical parameters and states."""
import numpy as np

from md_nucleation_amd import _lib as M
from oracle import oracle as O


def oracle_sysd(v, seed=1, replica=0):
    return {
        "pair_coeff": v["pair_coeff"], "btype": v["btype"], "ctype": v["ctype"], "dmulti": v["dmulti"],
        "dcoeff": v["dcoeff"], "species": v["species"], "n_gas": v["n_gas"], "n_vap": v["n_vap"],
        "ion_type_base": v["ion_type_base"], "dt": v["dt"], "T": v["T"], "d_size": v["d_size"],
        "margin": v["margin"], "ML2": v["ML2"], "m_gas_kg": v["m_gas_kg"], "m_vap_kg": v["m_vap_kg"],
        "gas_spin": v["gas_spin"], "flags": v["flags"], "efield": v["efield"], "shake": v["shake"],
        "shake_max_iter": v["shake_max_iter"], "seed": seed, "replica": replica,
        "nvt_ion": v.get("nvt_ion", 0), "Tnh_ion": v.get("Tnh_ion", 0.0),
    }


def oracle_for(sd, seed=1, replica=0):
    return O.Oracle(oracle_sysd(sd.view, seed, replica))


def state_from(sd, arrays, **extra):
    a = dict(arrays)
    a.update(extra)
    return M.State(sd.n_mol, sd.n_atoms, arrays=a)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def in_atom_mask(sd, region):
    """atoms of molecules with region set (the atoms the reference integrates)"""
    v = sd.view
    na = [s["mass"].size for s in v["species"]]
    counts = [na[0]] + [na[1]] * v["n_gas"] + [na[2]] * v["n_vap"]
    return np.repeat(np.asarray(region, bool), counts)
