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
