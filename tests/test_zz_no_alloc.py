"""edm_predict_noalloc (device-resident inputs and outputs, scratch in the library's static device array)
against edm_predict through the C ABI: two calls in a row on one context, the second with a smaller
task, must give edm_predict's outputs bit for bit (rc, k_used, predictions, S-map coefficients), and a
task whose scratch would exceed the array is refused with -12 before anything runs."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import load_product
from tests.tasks import random_task

pytestmark = pytest.mark.gpu

MISSING_I = -(2 ** 31)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself links (libamdhip64.so.7),
    so that the test's device buffers and the library share one runtime instance."""

    def __init__(self, fe):
        fe.load_library()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        self.bufs.append(p)
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _device_view(fe, hip, m):
    """An edm_manifold_view whose arrays are device copies of the host dict's."""
    v = fe.manifold_view(m, [])
    n = len(m["data"])
    for field in ("data", "targets", "point_times", "target_times"):
        a = np.ascontiguousarray(m.get(field, np.zeros(n)), np.float64)
        setattr(v, field, C.cast(hip.upload(a), fe._dp))
    v.panel = C.cast(hip.upload(np.ascontiguousarray(m["panel"], np.int32)), fe._ip) if m.get("panel") is not None \
        else None
    return v


def _noalloc(fe, M, Mp, opts, coeffs):
    hip = _Hip(fe)
    L = fe.load_library()
    hkeep = []
    try:
        mv = _device_view(fe, hip, M)
        mpv = _device_view(fe, hip, Mp)
        o = fe.options(opts, hkeep)
        NP, Ea = np.shape(Mp["data"])
        T = int(np.atleast_1d(opts["thetas"]).size)
        pred = hip.empty(8 * T * NP)
        rc = hip.empty(4 * T * NP)
        ku = hip.empty(4 * NP)
        co = hip.empty(8 * NP * (Ea + 1)) if coeffs else None
        status = L.edm_predict_noalloc(fe.context(), C.byref(mv), C.byref(mpv), C.byref(o), pred, rc, co, ku)
        out = (hip.download(pred, (T, NP), np.float64), hip.download(rc, (T, NP), np.int32),
               hip.download(co, (NP, Ea + 1), np.float64) if co is not None else None,
               hip.download(ku, (NP,), np.int32))
        return status, out
    finally:
        hip.free()


CASES = {
    # first call, second (smaller) call
    "simplex_list": (dict(NL=4000, NP=900, Ea=3, k=4, thetas=(1.0, 3.0)), dict(NL=700, NP=150, Ea=3, k=4)),
    "smap_list_coeffs": (dict(NL=2500, NP=600, Ea=6, k=9, algorithm=1, thetas=(0.0, 2.0), save_coeffs=True),
                         dict(NL=400, NP=80, Ea=6, k=9, algorithm=1, thetas=(1.0,), save_coeffs=True)),
    "options_list": (dict(NL=3000, NP=500, Ea=4, k=7, missing=0.05, missing_distance=0.3, panels=3, idw=0.5),
                     dict(NL=600, NP=90, Ea=4, k=5, panels=2, idw=-1.0)),
    "smap_all_dense": (dict(NL=1500, NP=300, Ea=4, k=-1, algorithm=1, thetas=(0, 1, 4), save_coeffs=True),
                       dict(NL=300, NP=40, Ea=4, k=-1, algorithm=1, thetas=(2.0,), save_coeffs=True)),
    "wide_select": (dict(NL=3000, NP=200, Ea=3, k=100, thetas=(0.5, 2.0)), dict(NL=900, NP=60, Ea=3, k=90)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_noalloc_matches_edm_predict(name):
    fe = load_product()
    for call, spec in enumerate(CASES[name]):
        M, Mp, opts = random_task(1000 + 17 * call + len(name), **spec)
        want_coeffs = bool(spec.get("save_coeffs"))
        print(f"case {name} call {call}: NL={spec['NL']} NP={spec['NP']} k={spec['k']}", flush=True)
        gp, gr, gc, gk, _, st = fe.predict(M, Mp, opts)
        status, (np_, nr, nc, nk) = _noalloc(fe, M, Mp, opts, want_coeffs)
        assert st == 0 and status == 0
        np.testing.assert_array_equal(nr, gr)
        np.testing.assert_array_equal(nk, gk)
        assert np.array_equal(np_.view(np.int64), gp.view(np.int64))
        if want_coeffs:
            assert np.array_equal(nc.view(np.int64), gc.view(np.int64))
        assert (gk != MISSING_I).any()


def test_noalloc_refuses_oversized_tasks():
    fe = load_product()
    M, Mp, opts = random_task(77, NL=400, NP=40000, Ea=2, k=64)
    status, _ = _noalloc(fe, M, Mp, opts, False)
    assert status == -12
    # the context still works afterwards
    M, Mp, opts = random_task(78, NL=500, NP=100, Ea=2, k=3)
    gp, gr, _, gk, _, st = fe.predict(M, Mp, opts)
    status, (np_, nr, _, nk) = _noalloc(fe, M, Mp, opts, False)
    assert st == 0 and status == 0
    np.testing.assert_array_equal(nk, gk)
    assert np.array_equal(np_.view(np.int64), gp.view(np.int64))
