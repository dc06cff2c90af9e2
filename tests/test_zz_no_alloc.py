"""t3x_oc_noalloc (the allocation-free OC bisection on caller-owned device
arrays) against t3x_oc (the same bisection, TOP3D_XL.m:145-155, through a
handle that owns its device buffers), on two calls in a row, the second with
a smaller input and a passive set; then a refused oversize call.

The two entry points fold the volume sums over different index ranges (the
handle's padded cell grid against the compact array), so their sums may
differ in the last bits; the bisection decisions are far from such ties, so
the bar is: the same bisection count and x and change within 1e-12."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """Device buffers through the HIP runtime that libt3x itself is bound to
    (hipMalloc / hipMemcpy / hipFree looked up through the library's own
    handle), so the caller's memory and the library live in one runtime."""

    def __init__(self):
        import top3d_xl_amd as t3x
        self.hip = t3x.load_library()
        for f in (self.hip.hipMalloc, self.hip.hipMemcpy, self.hip.hipFree):
            f.restype = C.c_int
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 1))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p.value

    def get(self, ptr, n, dtype=np.float64):
        out = np.zeros(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    x = 0.05 + 0.9 * rng.random(n)
    xPhys = np.clip(x + 0.05 * rng.standard_normal(n), 0.0, 1.0)
    dc = -(0.01 + rng.random(n)) * 3 * x ** 2
    return x, xPhys, dc


def _case(shape, seed, V0, passive_ids=None):
    import top3d_xl_amd as t3x
    x, xPhys, dc = _inputs(shape, seed)
    P = t3x.Problem(np.ones(shape, bool))
    if passive_ids is not None:
        P.set_passive(passive_ids)
    x_ref, ch_ref, nb_ref = P.oc(x, xPhys, dc, V0)
    P.close()
    dev = DeviceArrays()
    try:
        xd, xp, dd = dev.put(x), dev.put(xPhys), dev.put(dc)
        pas = None
        if passive_ids is not None:
            m = np.zeros(x.size, np.uint8)
            m[np.asarray(passive_ids)] = 1  # compact cell index == voxel id on a full cuboid
            pas = dev.put(m)
        ch, nb = t3x.oc_noalloc(x.size, xd, xp, dd, pas, V0)
        got = dev.get(xd, x.size)
    finally:
        dev.free()
    print(f"  {shape} n={x.size}: nBisect {nb} (handle {nb_ref}), change {ch:.6e} (handle {ch_ref:.6e}), "
          f"max|dx| {np.abs(got - x_ref).max():.2e}", flush=True)
    assert nb == nb_ref
    assert abs(ch - ch_ref) <= 1e-12
    np.testing.assert_allclose(got, x_ref, rtol=0, atol=1e-12)
    if passive_ids is not None:
        assert np.all(got[np.asarray(passive_ids)] == 1.0)


def test_oc_noalloc_matches_handle_oc_twice():
    print("\n[gpu-noalloc] two calls, the second smaller", flush=True)
    _case((24, 48, 24), 3, 0.12)
    _case((10, 20, 10), 4, 0.3, passive_ids=list(range(0, 2000, 7)))


def test_oc_noalloc_refuses_oversize():
    import top3d_xl_amd as t3x
    with pytest.raises(t3x.T3XError, match="outside"):
        t3x.oc_noalloc(1 << 31, 1, 1, 1, None, 0.3)
    with pytest.raises(t3x.T3XError, match="outside"):
        t3x.oc_noalloc(0, 1, 1, 1, None, 0.3)
