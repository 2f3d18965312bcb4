"""Loading of the golden fixtures (tests/golden/*.npz) for the tests."""
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATE_KEYS = ("mol_x", "mol_v", "region", "atom_x", "atom_v", "atom_f", "atom_w")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def sub(fx, prefix):
    n = len(prefix)
    return {k[n:]: v for k, v in fx.items() if k.startswith(prefix)}


def oracle_for(fx, prefix="sys_", shake=0, seed=1, **over):
    sysd = O.sysd_from_ref_dump(sub(fx, prefix), shake=shake, seed=seed)
    sysd.update(over)
    return O.Oracle(sysd), sysd


def state(fx, prefix="init_", **extra):
    st = {k: fx[prefix + k] for k in STATE_KEYS}
    st.update(extra)
    return st


def set_lists(S, fx, prefix="f_"):
    for k in ("gi", "gv"):
        key = prefix + "list_" + k
        if key in fx:
            arr = fx[key].astype(np.int32)
            getattr(S, k)[: arr.size] = arr
            setattr(S.c, "n_" + k, arr.size // 2)
