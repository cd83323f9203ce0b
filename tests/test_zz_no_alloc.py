"""The allocation-free getKPE entry point (ltb_kpe_apply_noalloc) through the
C ABI, against ltb_kpe_apply on the same inputs.

Both run the same recursion kernels on the same start block (the FP32 form
rounds X to float in both), so the outputs agree bit for bit.  The inputs
and outputs are device buffers allocated here, as a caller would own them."""
import ctypes as C

import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


class Hip:
    """device buffers of the caller (the library allocates nothing here)"""

    H2D, D2H = 1, 2

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.live.append(p)
        return p

    def put(self, d, a):
        assert self.h.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0

    def get(self, a, d):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), d, a.nbytes, self.D2H) == 0

    def close(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


@pytest.fixture
def hip():
    h = Hip()
    yield h
    h.close()


def noalloc(ctx, hip, M, w, X, precision):
    dX, dY = hip.alloc(X.nbytes), hip.alloc(X.nbytes)
    hip.put(dX, np.ascontiguousarray(X))
    w = np.ascontiguousarray(w, np.float64)
    st = L.lib().ltb_kpe_apply_noalloc(ctx._h, M, w.ctypes.data_as(L.PD), X.shape[0], dX, dY, precision)
    msg = L.lib().ltb_last_error(ctx._h).decode()
    Y = np.empty_like(X)
    if st == L.LTB_OK:
        hip.get(Y, dY)
    return st, msg, Y


def prepare(ctx, s):
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    ctx.rescale(a, b)
    return a, b


@pytest.mark.parametrize("cells,M", [((3, 3, 2), 120), ((8, 8, 8), 64)])
def test_noalloc_apply_equals_kpe_apply(ctx, hip, cells, M):
    """Two calls in a row in each precision, the second with fewer columns
    (the scratch array still holds the first call's vectors), equal the
    allocating entry point bit for bit; FP64 also matches the oracle."""
    s = L.System.quartz(*cells, sigma_ang=0.08, seed=5)
    a, b = prepare(ctx, s)
    w = O.kpe_weights(O.fd_coeffs(M, -0.2, 1e-3, 8192), O.gmM(M))
    rng = np.random.default_rng(3)
    for prec in (L.FP64, L.FP32_ACC64):
        for ncols in (70, 5):
            X = rng.standard_normal((ncols, s.norb))
            ref = ctx.kpe_apply(M, w, X, prec)
            st, msg, Y = noalloc(ctx, hip, M, w, X, prec)
            assert st == L.LTB_OK, msg
            assert np.array_equal(Y, ref), (prec, ncols, float(np.abs(Y - ref).max()))
            if prec == L.FP64 and cells == (3, 3, 2):
                o = O.System.quartz(*cells, sigma_ang=0.08, seed=5)
                assert np.abs(Y[-1] - o.kpe_apply(a, b, w, X[-1])).max() < 1e-11


def test_noalloc_apply_rejects_calls_beyond_its_scratch(ctx, hip):
    """200 columns of 18,432 orbitals (ld 256) need 75.5 MB in FP32, which
    fits the 96 MiB array, and 113 MB in FP64, which does not: the FP64 call
    fails with LTB_ERR_ARG and names the largest column count that fits."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    prepare(ctx, s)
    M = 16
    w = O.kpe_weights(O.fd_coeffs(M, -0.2, 1e-3, 1024), O.gmM(M))
    X = np.random.default_rng(1).standard_normal((200, s.norb))
    st, msg, Y = noalloc(ctx, hip, M, w, X, L.FP64)
    assert st == L.LTB_ERR_ARG
    assert "maximum" in msg and "at most 192 columns" in msg, msg
    st, msg, Y = noalloc(ctx, hip, M, w, X, L.FP32_ACC64)
    assert st == L.LTB_OK, msg
    assert np.array_equal(Y, ctx.kpe_apply(M, w, X, L.FP32_ACC64))
