"""smalgo_apply_poly_block_noalloc, the filter entry point that allocates
nothing (its second ping-pong block is a fixed per-device scratch), against
smalgo_apply_poly_block on the same inputs.  Device buffers are allocated here,
by the test, through the HIP runtime; the entry point itself takes them."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DevArray:
    """n x ld FP64 row-major device block (test-side allocation)."""

    def __init__(self, n, ld):
        self.n, self.ld = n, ld
        self.nbytes = n * ld * 8
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), self.nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.n, self.ld)
        assert hip().hipMemcpy(self.p, a.ctypes.data, self.nbytes, 1) == 0   # host -> device

    def get(self):
        a = np.empty((self.n, self.ld))
        assert hip().hipMemcpy(a.ctypes.data, self.p, self.nbytes, 2) == 0   # device -> host
        return a

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(ctx, dm, b, dQ, co, smin, smax, dY):
    cp = np.ascontiguousarray(co, dtype=np.float64)
    return ctx.lib.smalgo_apply_poly_block_noalloc(
        ctx.h, dm.h, b, dQ.p, dQ.ld, cp.ctypes.data_as(C.POINTER(C.c_double)), len(cp) - 1,
        smin, smax, dY.p, dY.ld)


@pytest.fixture(scope="module")
def mat(ctx):
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd.lattices import anderson3d
    A = anderson3d(20, 16.5, 20)                      # n = 8000
    dm = sm.DeviceMatrix(A, ctx)
    return dm, dm.gershgorin_bounds()


def _check(ctx, dm, smin, smax, b, K, seed):
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd.lattices import gaussian_block
    n = dm.n
    print(f"noalloc case n={n} b={b} K={K} seed={seed}", flush=True)
    Q = gaussian_block(seed, n, b)
    co = sm.build_coeffs(K, smin, smax, 0.0, -0.3, 0.3, "step")
    Yref = sm.apply_poly_block(dm, Q, co, smin, smax, precision="fp64", ctx=ctx)
    ld = (b + 3) // 4 * 4
    Qrm = np.zeros((n, ld))
    Qrm[:, :b] = Q
    dQ, dY = DevArray(n, ld), DevArray(n, ld)
    try:
        dQ.put(Qrm)
        dY.put(np.full((n, ld), np.nan))             # the result must not depend on Y's contents
        rc = _noalloc(ctx, dm, b, dQ, co, smin, smax, dY)
        assert rc == 0, ctx.lib.smalgo_last_error(ctx.h).decode()
        Y = dY.get()[:, :b]
    finally:
        dQ.free()
        dY.free()
    scale = np.max(np.abs(Yref))
    err = np.max(np.abs(Y - Yref))
    assert np.all(np.isfinite(Y))
    assert err <= 1e-12 * scale, (b, K, err, scale)
    return err


def test_noalloc_filter_matches_apply_poly_block_two_calls(ctx, mat):
    dm, (smin, smax) = mat
    # two calls in a row, the second with a smaller input (fewer columns, lower degree):
    # the shared scratch keeps the first call's contents
    e1 = _check(ctx, dm, smin, smax, 48, 300, 3)
    e2 = _check(ctx, dm, smin, smax, 16, 150, 4)
    print(f"max |Y - Y_ref|: {e1:.1e}, {e2:.1e}")


@pytest.mark.parametrize("K", [0, 1, 2, 3, 4])
def test_noalloc_filter_low_degrees(ctx, mat, K):
    """Degrees 0-4 exercise every placement of Y in the ping-pong pair."""
    dm, (smin, smax) = mat
    _check(ctx, dm, smin, smax, 8, K, 10 + K)


def test_noalloc_filter_rejects_beyond_scratch(ctx, mat):
    """n * ldy * 8 above the 64 MiB scratch is an argument error, raised before
    the (tiny, never touched) device buffers are used."""
    from sparse_matrix_algorithms_amd import _lib as L
    dm, (smin, smax) = mat
    b = 1100                                          # 8000 * 1100 * 8 B = 70.4 MB > 64 MiB
    dQ, dY = DevArray(1, 4), DevArray(1, 4)
    dQ.ld = dY.ld = b
    try:
        rc = _noalloc(ctx, dm, b, dQ, np.ones(3), smin, smax, dY)
    finally:
        dQ.free()
        dY.free()
    assert rc == L.ERR_ARG
    assert "scratch" in ctx.lib.smalgo_last_error(ctx.h).decode()
