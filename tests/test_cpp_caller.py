"""A C++ program (tests/cpp/mdn_caller.cpp) that binds include/mdn.h the way
the reference's MD class would: the end of MD::MD hands the state to the
device and builds the pair lists (md.cpp:34), MD::verlet (verlet.cpp:20-39)
becomes one mdn_verlet step, and main.cpp:21-30's replica fan-out becomes one
context stepped by mdn_run_batched.  Its trajectories are compared with the
oracle's from the same starting states; its error-contract mode needs no
device."""
import os
import subprocess

import numpy as np
import pytest

from md_nucleation_amd import _lib as M
from tests import engine_util as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "mdn_caller")


def build_caller():
    """tests/cpp/Makefile: g++ against the in-tree libmdn.so (also run by
    __graft_entry__.build())"""
    subprocess.check_call(["make", "-s", "-C", os.path.dirname(CALLER), "mdn_caller"])
    return CALLER


def _cond(d):
    from md_nucleation_amd import configs
    os.makedirs(d, exist_ok=True)
    m = configs.molecule_files(d)
    return configs.Condition(m["na"], m["methanol"], 12, "N2", 400, pressure=6e5, margin=1.0,
                             shake_h=True).write(os.path.join(d, "t.cond"))


def _run(args, timeout=120):
    assert os.path.exists(CALLER), "tests/cpp/mdn_caller is not built (__graft_entry__.build() builds it)"
    print("caller:", " ".join(args), flush=True)
    r = subprocess.run([CALLER] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


def _read(path):
    with open(path, "rb") as f:
        n_rep, n_mol, n_atoms = np.frombuffer(f.read(12), np.int32)
        out = []
        for _ in range(n_rep):
            rb = int(np.frombuffer(f.read(8), np.int64)[0])
            x, v, fo = (np.frombuffer(f.read(24 * n_atoms), np.float64).reshape(-1, 3) for _ in range(3))
            U = np.frombuffer(f.read(8 * len(M.U_LABELS)), np.float64)
            region = np.frombuffer(f.read(int(n_mol)), np.uint8)
            out.append(dict(rebuild_count=rb, atom_x=x, atom_v=v, atom_f=fo, U=U, region=region))
        assert f.read() == b""
    return out


def _oracle(sd, seed, replica, nsteps, every):
    orc = E.oracle_for(sd, seed=seed, replica=replica)
    S = orc.new_state(sd.init_state(seed).as_dict())
    assert orc.make_pair(S) == 0
    rc, _ = orc.verlet(S, nsteps, eflag_every=every)
    assert rc == 0
    return S


def _compare(sd, got, S, label):
    assert got["rebuild_count"] == S.c.rebuild_count and got["rebuild_count"] > 1, label
    np.testing.assert_array_equal(got["region"], S.region)
    m = E.in_atom_mask(sd, got["region"])
    dx = np.abs(got["atom_x"][m] - S.atom_x.reshape(-1, 3)[m]).max()
    dv = np.abs(got["atom_v"][m] - S.atom_v.reshape(-1, 3)[m]).max()
    fe = E.rel(got["atom_f"][m], S.atom_f.reshape(-1, 3)[m])
    print("%s: rebuilds %d max |dx| %.2e A max |dv| %.2e rel F %.2e" % (label, S.c.rebuild_count, dx, dv, fe))
    assert dx < 1e-8 and dv < 1e-10 and fe < 1e-7
    np.testing.assert_allclose(got["U"], S.U, rtol=1e-6, atol=1e-6)


def test_caller_error_contract(tmp_path):
    """host-side failures come back as a negative mdn_status with a message"""
    if not os.path.exists(CALLER):
        build_caller()
    r = _run(["errors", _cond(str(tmp_path))])
    assert r.returncode == 0 and "errors: ok" in r.stdout


@pytest.mark.gpu
def test_caller_percall_vs_oracle(tmp_path):
    """MD::verlet as mdn_verlet, one step per call, 200 FP64 steps with SHAKE
    and pair-list rebuilds"""
    cond = _cond(str(tmp_path))
    out = str(tmp_path / "percall.bin")
    r = _run(["percall", cond, "3", "200", "10", out])
    assert r.returncode == 0
    sd = M.SysDef(cond)
    _compare(sd, _read(out)[0], _oracle(sd, 3, 0, 200, 10), "percall")


@pytest.mark.gpu
def test_caller_batched_vs_oracle(tmp_path):
    """main's replica loop (calculation numbers 3, 4, 5) as one
    mdn_run_batched call of 200 FP64 steps"""
    cond = _cond(str(tmp_path))
    out = str(tmp_path / "batched.bin")
    r = _run(["batched", cond, "3", "3", "200", "10", out])
    assert r.returncode == 0
    sd = M.SysDef(cond)
    got = _read(out)
    assert len(got) == 3
    for i in range(3):
        _compare(sd, got[i], _oracle(sd, 3 + i, i, 200, 10), "batched replica %d" % i)
