"""pd_compute_velocity_static, the allocation-free variant of
pd_compute_velocity: its intermediates live in a fixed per-GPU workspace of the
library, the mesh and state in the caller's device buffers.  Checked against
pd_compute_velocity on a context of the same mesh and state: two calls in a row
(the second on a smaller plate, so the workspace holds the first call's
leftovers), with bonds breaking in both; and the size limits of the workspace,
on the CPU."""
import ctypes as C

import numpy as np
import pytest

from conftest import say


def _static(pdh):
    fn = pdh.lib().pd_compute_velocity_static
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(pdh.Params), C.c_int, C.c_long, C.c_int] + [C.c_void_p] * 9 + [C.c_int, C.c_void_p]
    return fn


def test_static_rejects_meshes_beyond_the_workspace(pdpkg):
    pdh = pdpkg.pdhip
    fn = _static(pdh)
    p = pdh.Params.make()
    null = [None] * 9
    rc = fn(C.byref(p), 131072 + 1, 0, 28, *null, 0, None)
    assert rc == pdh.PD_ERR_INVALID
    assert b"131072" in pdh.lib().pd_last_error(None)
    rc = fn(C.byref(p), 1000, 524288 + 1, 28, *null, 0, None)
    assert rc == pdh.PD_ERR_INVALID
    assert b"524288" in pdh.lib().pd_last_error(None)
    rc = fn(C.byref(pdh.Params.make(precision=pdh.MIXED)), 1000, 1000, 28, *null, 0, None)
    assert rc == pdh.PD_ERR_UNSUPPORTED


def smooth_u(w, amp, seed):
    rng = np.random.default_rng(seed)
    P = w.X.reshape(-1, 2)
    L = np.ptp(P, axis=0).max()
    k = rng.normal(size=(3, 2)) * 2 * np.pi / L
    ph = rng.uniform(0, 2 * np.pi, size=3)
    c = rng.normal(size=(3, 2))
    u = sum(c[m] * np.sin(P @ k[m] + ph[m])[:, None] / np.linalg.norm(k[m]) for m in range(3))
    return (amp * u).reshape(-1)


class DeviceBuffers:
    """Device buffers from the HIP runtime lib2dpdgpu itself links
    (libamdhip64.so.7), freed on close."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 16)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.mark.gpu
def test_static_entry_point_matches_context_path(pdpkg):
    pdh = pdpkg.pdhip
    fn = _static(pdh)
    s0 = 2e-4
    for nx, ny, seed in ((48, 40, 3), (24, 20, 4)):
        say(f"static entry point: {nx} x {ny} plate, s0 = {s0}, against pd_compute_velocity")
        w = pdpkg.plates.single_plate(f"s{nx}", nx, ny, 20.0, notch_len_cells=nx // 3, seed=seed)
        rng = np.random.default_rng(seed)
        u = smooth_u(w, 4e-4, seed)
        v = rng.normal(size=2 * w.n)
        a = rng.normal(size=2 * w.n) * 1e6
        ctx = w.context(s0=s0)
        ctx.set_state(u=u, v=v, a=a)
        ctx.compute_velocity(write_acceleration=True)
        ref = ctx.get_state()
        ref_broken = ctx.get_broken()
        ctx.close()
        assert ref_broken.sum() > w.broken.sum()   # bonds broke in this call

        d = DeviceBuffers()
        try:
            X, V, row, col = d.put(w.X), d.put(w.V), d.put(w.row.astype(np.int32)), d.put(w.col.astype(np.int32))
            broken = d.put(w.broken.astype(np.uint8))
            ud, vd, ad = d.put(u), d.put(v), d.put(a)
            fd = d.put(np.full(2 * w.n, np.nan))
            p = w.params(precision=pdh.FP64, s0=s0)
            mf = int(np.diff(w.row).max())
            rc = fn(C.byref(p), w.n, w.nb, mf, X, V, row, col, broken, ud, vd, ad, fd, 1, None)
            assert rc == pdh.PD_OK, pdh.lib().pd_last_error(None)
            np.testing.assert_array_equal(d.get(broken, ref_broken), ref_broken)
            np.testing.assert_array_equal(d.get(fd, ref["f"]), ref["f"])
            np.testing.assert_array_equal(d.get(vd, ref["v"]), ref["v"])
            np.testing.assert_array_equal(d.get(ad, ref["a"]), ref["a"])
            np.testing.assert_array_equal(d.get(ud, u), u)
            # an asymmetric family is reported (the call waits for its mesh
            # check outside stream capture) and leaves the outputs untouched
            bad_col = w.col.astype(np.int32).copy()
            j0 = int(bad_col[0])
            bad_col[0] = j0 + 1 if j0 + 1 < w.n else j0 - 1
            bad = d.put(bad_col)
            f_before = d.get(fd, ref["f"])
            rc = fn(C.byref(p), w.n, w.nb, mf, X, V, row, bad, broken, ud, vd, ad, fd, 1, None)
            assert rc == pdh.PD_ERR_INVALID
            assert b"symmetric" in pdh.lib().pd_last_error(None)
            np.testing.assert_array_equal(d.get(fd, ref["f"]), f_before)
        finally:
            d.close()
