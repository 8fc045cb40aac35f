"""The C-ABI library loads and exports every entry point its header declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dioptrin_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dio_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "dio_plan_create" in names and "dio_integrate_device" in names
    from dioptrin_amd import _native
    assert sorted(_native.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    from dioptrin_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dio_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_library_is_built_for_gfx950():
    from dioptrin_amd import _native
    data = open(_native.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_library_reports_version_and_no_device_cleanly():
    from dioptrin_amd import _native
    L = _native.lib()
    assert b"gfx950" in L.dio_version()
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = L.dio_open(0, ctypes.byref(h))
    assert rc == _native.DIO_ERR_UNAVAILABLE
    assert L.dio_last_error()


def test_invalid_configuration_fails_before_the_device():
    # Configuration errors come before any device work (cuda.rs:1001-1057).
    from dioptrin_amd import _native as N
    L = N.lib()
    p = N.Poni(1e-4, 1e-4, 0.1, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-10)
    d = N.PlanDesc()
    d.dim1, d.dim2, d.ndim, d.n_radial = 0, 10, 1, 10
    rc = L.dio_plan_create(None, ctypes.byref(p), ctypes.byref(d), ctypes.byref(ctypes.c_void_p()))
    assert rc == N.DIO_ERR_INVALID


def test_input_order_descriptor():
    # InputOrder (fixed.rs:20-24): the header's values, the Python names and
    # the descriptor field agree; an unknown order is rejected on the host.
    import re
    from dioptrin_amd import _native as N
    from dioptrin_amd.integrator import plan_args
    hdr = open(os.path.join(ROOT, "include", "dioptrin_hip.h")).read()
    m = re.search(r"DIO_ORDER_DETECTOR = (\d+), DIO_ORDER_FLIP_VERTICAL = (\d+)", hdr)
    assert m and N.ORDERS == {"detector": int(m.group(1)), "flip_vertical": int(m.group(2))}
    assert re.search(r"int32_t precision;[^\n]*\n\s*int32_t input_order;", hdr)
    assert [f for f, _ in N.PlanDesc._fields_][-2:] == ["precision", "input_order"]
    poni = dict(pixel1=1e-4, pixel2=1e-4, distance=0.1, wavelength=1e-10)
    _, d, _ = plan_args(poni, (8, 9), input_order="flip_vertical")
    assert d.input_order == N.ORDERS["flip_vertical"]
    _, d, _ = plan_args(poni, (8, 9))
    assert d.input_order == N.ORDERS["detector"]
    with pytest.raises(ValueError):
        plan_args(poni, (8, 9), input_order="sideways")


def test_image_of_2_29_pixels_is_rejected_before_the_device():
    # The tiled gather addresses frames with 32-bit byte offsets: images of
    # 2^29 pixels or more are refused with DIO_ERR_INVALID at plan creation.
    from dioptrin_amd import _native as N
    L = N.lib()
    p = N.Poni(1e-4, 1e-4, 0.1, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-10)
    d = N.PlanDesc()
    d.dim1, d.dim2, d.ndim, d.n_radial = 1 << 14, 1 << 15, 1, 10
    rc = L.dio_plan_create(None, ctypes.byref(p), ctypes.byref(d), ctypes.byref(ctypes.c_void_p()))
    assert rc == N.DIO_ERR_INVALID
    assert b"2^29" in L.dio_last_error()


class _Buf:
    def __init__(self, shape, dtype):
        self.shape, self.dtype, self.ptr = tuple(shape), dtype, 0


def test_device_buffers_are_checked_on_the_host():
    """GpuPlan validates device frames, outputs and masks before the C ABI,
    which trusts their sizes."""
    import numpy as np
    from dioptrin_amd.integrator import GpuPlan
    plan = GpuPlan.__new__(GpuPlan)
    plan.shape, plan.nbins = (8, 9), 5
    fr = _Buf((3, 8, 9), np.uint32)
    ok = _Buf((3, 5), np.float64)
    assert plan._check_device(fr, ok) == "uint32"
    assert plan._check_device(fr, ok, ok, errors=True) == "uint32"
    bad = [
        dict(frames=_Buf((3, 8, 10), np.uint32), out_i=ok),       # frame shape
        dict(frames=_Buf((3, 8, 9), np.int64), out_i=ok),         # dtype
        dict(frames=fr, out_i=_Buf((2, 5), np.float64)),          # too few rows
        dict(frames=fr, out_i=_Buf((3, 5), np.float32)),          # output dtype
        dict(frames=fr, out_i=ok, errors=True),                   # sigma missing
        dict(frames=fr, out_i=ok, out_s=_Buf((3, 4), np.float64), errors=True),
        dict(frames=fr, out_i=ok, azimuthal_std=True),            # std missing
        dict(frames=fr, out_i=ok, masks=_Buf((3, 8, 9), np.int8)),
        dict(frames=fr, out_i=None),                              # intensity missing
    ]
    for kw in bad:
        with pytest.raises((ValueError, TypeError)):
            plan._check_device(**kw)
