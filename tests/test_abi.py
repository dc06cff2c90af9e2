"""CPU-side checks of the drop-in boundary: the C-ABI library loads, exports
every symbol include/*.h declares, and refuses to run without a GPU."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(ROOT, "include")
    src = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    return sorted(set(re.findall(r"\b(t3x_[A-Za-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    import top3d_xl_amd as t3x
    L = t3x.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(t3x.EXPORTED) == set(syms)


def test_default_params_are_initial_settings():
    import top3d_xl_amd as t3x
    p = t3x.default_params()
    assert (p.E0, p.nu, p.Emin, p.penal, p.tol, p.maxIT, p.omega, p.coarsestControl) == \
        (1.0, 0.3, 1e-6, 3.0, 1e-3, 800, 0.6, 50000)


def test_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import top3d_xl_amd as t3x
    with pytest.raises(t3x.T3XError, match="no GPU"):
        t3x.Problem(np.ones((4, 8, 4), bool))


def test_bad_inputs_raise_before_device():
    import top3d_xl_amd as t3x
    with pytest.raises(t3x.T3XError):
        t3x.Problem(np.ones((2, 8, 4), bool))   # cuboid thinner than 3 layers (532-534)
    with pytest.raises(t3x.T3XError, match="Wrong Input"):
        t3x.TOP3D_XL(np.ones((4, 8, 4), bool), "LOCAL", 0.5, 3, 1.7)   # LOCAL without rHat (16)


def test_server_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import top3d_xl_amd as t3x
    with pytest.raises(t3x.T3XError):
        t3x.Server(max_active=2)


def test_footprint_estimate_scales():
    import top3d_xl_amd as t3x
    a = t3x.footprint(32, 64, 32)
    b = t3x.footprint(64, 128, 64)
    assert 6 * a < b < 10 * a  # ~8x the nodes
    assert t3x.footprint(64, 128, 64, mode="LOCAL") > b  # MMA state
    assert t3x.footprint(64, 128, 64, precision=1) > b  # FP32 V-cycle copies
    std = dict(vcycle=1, omega=0.4)
    assert t3x.footprint(64, 128, 64, precision=2, **std) < t3x.footprint(64, 128, 64, precision=1, **std)  # FP16 Ks


def test_auto_depth_only_where_the_coarsest_band_is_costly():
    """coarsestControl <= 0 deepens the hierarchy while the coarsest band
    Cholesky would cost more than 1e11 flop: at 192x384x192 (coarsest 48x24x24
    cells, 91,875 DOF, bandwidth 1,953) one level more halves the footprint
    (the solve-block operators dominate it); at 256x512x256 (28,611 DOF,
    bandwidth 921) the hierarchy, and so the footprint, is unchanged."""
    import top3d_xl_amd as t3x
    form = dict(precision=1, vcycle=1, omega=0.4)
    ref192, auto192 = t3x.footprint(192, 384, 192, **form), t3x.footprint(192, 384, 192, coarsestControl=0, **form)
    assert auto192 < 0.6 * ref192
    assert t3x.footprint(256, 512, 256, coarsestControl=0, **form) == t3x.footprint(256, 512, 256, **form)


def test_create_multi_checks_arguments_before_device():
    """t3x_create_multi refuses null or negative boundary-condition arrays with
    T3X_ERR_INPUT before it looks for a device."""
    import ctypes as C
    import top3d_xl_amd as t3x
    L = t3x.load_library()
    vox = np.ones(4 * 8 * 4, dtype=np.uint8)
    P = t3x.default_params()
    h = C.c_void_p()
    fn = np.array([0, 1], dtype=np.int32)
    fs = np.ones(6, dtype=np.uint8)
    npc = np.array([1, -1], dtype=np.int32)
    ln = np.array([5], dtype=np.int32)
    lv = np.zeros(3)
    cases = [
        (2, None, fs.ctypes.data, 1, np.array([1], np.int32).ctypes.data, ln.ctypes.data, lv.ctypes.data),
        (2, fn.ctypes.data, fs.ctypes.data, 2, npc.ctypes.data, ln.ctypes.data, lv.ctypes.data),
        (2, fn.ctypes.data, fs.ctypes.data, 1, np.array([1], np.int32).ctypes.data, None, lv.ctypes.data),
        (2, fn.ctypes.data, fs.ctypes.data, 1, None, ln.ctypes.data, lv.ctypes.data),
    ]
    keep = [np.array([1], np.int32) for _ in cases]
    for (nf, f, s, nc, pc, l, v), k in zip(cases, keep):
        if pc is not None and nc == 1:
            pc = k.ctypes.data
        rc = L.t3x_create_multi(vox.ctypes.data, 4, 8, 4, nf, f, s, nc, pc, l, v, C.byref(P), C.byref(h))
        assert rc == 1, rc  # T3X_ERR_INPUT
        assert b"boundary-condition" in L.t3x_last_error(h)
        L.t3x_destroy(h)


def test_footprint_refuses_planes_beyond_launch_limits():
    """Kernels that write one reduction partial per work-group chunk z until
    the grid fits the partials buffer; a node plane with more tiles than that
    buffer even at one chunk (a 3840 x 3840 x 8 plate, which passes the 2^31
    DOF check) is refused with T3X_ERR_INPUT instead of launching a broken
    grid, while a 2048 x 2048 x 8 plate is accepted."""
    import top3d_xl_amd as t3x
    assert t3x.footprint(2048, 2048, 8) > 0
    with pytest.raises(t3x.T3XError):
        t3x.footprint(3840, 3840, 8)
    # the same voxel count as a tall column: the y-x plane is small, z is chunked
    assert t3x.footprint(8, 3840, 3840) > 0
