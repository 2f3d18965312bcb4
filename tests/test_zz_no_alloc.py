"""mdn_vi_adjoint_noalloc (no reference counterpart: the allocation-free
variant of mdn_vi_adjoint for callers that cannot allocate) against
mdn_vi_adjoint, which stages host buffers and allocates its own partials.
Same kernels, same inputs: the gradients must agree to the last bit.  Two
calls in a row, the second with a smaller batch, through the C ABI with
device buffers the test allocates; and the size limit refused before any
launch."""
import ctypes as C

import numpy as np
import pytest

from md_nucleation_amd import _lib as M

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
_dp = C.POINTER(C.c_double)


def _hip():
    M.lib()  # the HIP runtime libmdn bound to (torch's, when torch loaded first)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype=np.float64):
        out = np.zeros(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _inputs(seed, S, n_ion, n_vap):
    rng = np.random.default_rng(seed)
    n = n_ion + n_vap
    x = np.zeros((S, n, 3))
    for s in range(S):
        pts = []
        while len(pts) < n:  # atoms at least 2 A apart in a 30 A box
            p = rng.uniform(-15, 15, 3)
            if all(np.sum((p - q) ** 2) > 4.0 for q in pts):
                pts.append(p)
        x[s] = pts
    return {
        "x": x.reshape(-1), "w": rng.uniform(0.5, 1.0, S * n),
        "si": rng.uniform(2.5, 3.9, n_ion), "ei": rng.uniform(0.02, 0.2, n_ion), "qi": rng.normal(0, 0.3, n_ion),
        "sv": rng.uniform(1.0, 3.4, n_vap), "ev": rng.uniform(0.01, 0.2, n_vap), "qv": rng.normal(0, 0.4, n_vap),
        "vin": (rng.uniform(size=S * n_vap) < 0.8).astype(np.uint8),
        "gE": rng.normal(size=S), "gF": rng.normal(size=S * n * 3), "S": S, "n_ion": n_ion, "n_vap": n_vap,
    }


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _host(inp, with_gE):
    gs, ge, gq = (np.zeros(inp["n_ion"]) for _ in range(3))
    M.check(M.lib().mdn_vi_adjoint(_ptr(inp["x"]), _ptr(inp["w"]), inp["S"], inp["n_ion"], inp["n_vap"],
                                   _ptr(inp["si"]), _ptr(inp["ei"]), _ptr(inp["qi"]), _ptr(inp["sv"]),
                                   _ptr(inp["ev"]), _ptr(inp["qv"]), inp["vin"].ctypes.data_as(C.c_void_p),
                                   _ptr(inp["gE"]) if with_gE else None, _ptr(inp["gF"]), _ptr(gs), _ptr(ge),
                                   _ptr(gq)), "mdn_vi_adjoint")
    return gs, ge, gq


KEYS = ("x", "w", "si", "ei", "qi", "sv", "ev", "qv", "vin", "gE", "gF")


def test_noalloc_matches_vi_adjoint_on_two_calls():
    hip = _hip()
    big = _inputs(1, 48, 40, 90)
    small = _inputs(2, 12, 15, 33)
    bufs = {k: DevBuf(hip, big[k].nbytes) for k in KEYS}
    outs = [DevBuf(hip, 8 * big["n_ion"]) for _ in range(3)]
    try:
        for inp, with_gE in ((big, True), (small, False)):
            for k in KEYS:
                bufs[k].put(inp[k])
            print("case: noalloc adjoint S=%d n_ion=%d n_vap=%d gE=%s" % (inp["S"], inp["n_ion"], inp["n_vap"],
                                                                          with_gE), flush=True)
            rc = M.lib().mdn_vi_adjoint_noalloc(
                bufs["x"].p, bufs["w"].p, inp["S"], inp["n_ion"], inp["n_vap"], bufs["si"].p, bufs["ei"].p,
                bufs["qi"].p, bufs["sv"].p, bufs["ev"].p, bufs["qv"].p, bufs["vin"].p,
                bufs["gE"].p if with_gE else None, bufs["gF"].p, outs[0].p, outs[1].p, outs[2].p, None)
            M.check(rc, "mdn_vi_adjoint_noalloc")
            assert hip.hipDeviceSynchronize() == 0
            got = [o.get(inp["n_ion"]) for o in outs]
            want = _host(inp, with_gE)
            for g, w, name in zip(got, want, ("sigma", "eps", "q")):
                print("  d/d%s max |diff| %.3e (max |g| %.3e)" % (name, np.abs(g - w).max(), np.abs(w).max()))
                assert np.abs(w).max() > 0
                np.testing.assert_array_equal(g, w)
    finally:
        for b in list(bufs.values()) + outs:
            b.free()


def test_noalloc_refuses_what_does_not_fit():
    hip = _hip()
    d = DevBuf(hip, 64)
    try:
        L = M.lib()
        # 1000 snapshots x 200 ion atoms x 3 partials exceed the 2^19 doubles
        rc = L.mdn_vi_adjoint_noalloc(*([d.p] * 2), 1000, 200, 10, *([d.p] * 7), None, None, d.p, d.p, d.p, None)
        assert rc == -1
        assert "exceeds the scratch capacity" in L.mdn_last_error().decode()
        rc = L.mdn_vi_adjoint_noalloc(*([d.p] * 2), 4, 2, 3, *([d.p] * 7), None, None, None, d.p, d.p, None)
        assert rc == -1
    finally:
        d.free()
