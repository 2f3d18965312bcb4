"""Parsing and comparison of the output files MD::run writes (output.cpp):
K_<n>.dat, U_<n>.dat, ion_<T>_<n>.dat, <name>_<n>.dump, vapor_in/out_<n>.dat."""
import numpy as np

FILES = ("K_%d.dat", "U_%d.dat", "ion_300_%d.dat", "out_%d.dump", "vapor_in_%d.dat", "vapor_out_%d.dat")


def table(text):
    """whitespace table of numbers -> 2-D float array (0 rows for empty text)"""
    rows = [ln.split() for ln in text.splitlines() if ln.strip()]
    if not rows:
        return np.zeros((0, 0))
    return np.array([[float(t) for t in r] for r in rows], np.float64)


def dump_frames(text):
    """LAMMPS-style frames of export_dump_close: [(timestep, natoms, bounds[6],
    ids[n], types[n], values[n, 6])]"""
    lines = text.splitlines()
    out, k = [], 0
    while k < len(lines):
        assert lines[k] == "ITEM: TIMESTEP", lines[k]
        ts = int(lines[k + 1])
        assert lines[k + 2] == "ITEM: NUMBER OF ATOMS"
        n = int(lines[k + 3])
        assert lines[k + 4] == "ITEM: BOX BOUNDS pp pp pp"
        b = np.array([float(t) for ln in lines[k + 5:k + 8] for t in ln.split()])
        assert lines[k + 8] == "ITEM: ATOMS id type x y z vx vy vz"
        rows = [ln.split() for ln in lines[k + 9:k + 9 + n]]
        ids = np.array([int(r[0]) for r in rows])
        types = [r[1] for r in rows]
        vals = np.array([[float(t) for t in r[2:]] for r in rows]).reshape(n, 6)
        out.append((ts, n, b, ids, types, vals))
        k += 9 + n
    return out


def close(a, b, rel, abs_):
    """element-wise |a - b| <= abs_ + rel |b| (printed values: rounding of the
    last digit plus the arithmetic's own difference)"""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    bad = np.abs(a - b) > abs_ + rel * np.abs(b)
    return not bad.any(), (np.argwhere(bad)[:5].tolist(), a[bad][:5], b[bad][:5])


def compare_run_text(got, want, rel=2e-6, abs_e=1e-12, abs_f=2e-6, names=True):
    """got/want: dicts file name -> text.  %e columns within rel (7 printed
    digits), %f dump columns within abs_f; ids, counts, type names exact."""
    for f in want:
        if f.endswith(".dump"):
            A, B = dump_frames(got[f]), dump_frames(want[f])
            assert len(A) == len(B), (f, len(A), len(B))
            for fa, fb in zip(A, B):
                assert fa[0] == fb[0] and fa[1] == fb[1], (f, fa[:2], fb[:2])
                ok, why = close(fa[2], fb[2], rel, abs_e)
                assert ok, (f, "bounds", why)
                np.testing.assert_array_equal(fa[3], fb[3])
                if names:
                    assert fa[4] == fb[4], f
                ok, why = close(fa[5], fb[5], 0.0, abs_f)
                assert ok, (f, "atoms", why)
        else:
            A, B = table(got[f]), table(want[f])
            assert A.shape == B.shape, (f, A.shape, B.shape)
            if f.startswith("vapor") and A.size:
                np.testing.assert_array_equal(A[:, 0], B[:, 0])
            ok, why = close(A, B, rel, abs_e)
            assert ok, (f, why)
