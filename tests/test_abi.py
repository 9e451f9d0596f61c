"""The C-ABI library loads and exports every function include/*.h declares
(no device calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ems_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared("ems_gpu.h") + declared("ems_host.h") + declared("ems_learn.h")
    assert len(names) > 20
    # the exact build and the one with FMA contraction export the same ABI
    for so in ("libems_gpu.so", "libems_gpu_fma.so"):
        lib = C.CDLL(os.path.join(ROOT, "ems_amd", so))
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (so, missing)


def test_invalid_arguments_are_rejected_without_a_device():
    import ems_amd._native as nat
    assert nat.ems_gpu_create(None, None) == nat.EMS_E_INVALID
    d = nat.GridDesc((C.c_int32 * 3)(1, 4, 4), 1e-12, 0)
    h = C.c_void_p()
    assert nat.ems_gpu_create(C.byref(d), C.byref(h)) == nat.EMS_E_INVALID
    d = nat.GridDesc((C.c_int32 * 3)(4, 4, 4), -1.0, 0)
    assert nat.ems_gpu_create(C.byref(d), C.byref(h)) == nat.EMS_E_INVALID
    assert nat.ems_gpu_run(None, 1) == nat.EMS_E_INVALID
    assert nat.ems_gpu_energy_noalloc(None, None) == nat.EMS_E_INVALID
    assert nat.ems_gpu_energy_in(None, 0, None) == nat.EMS_E_INVALID
    dil = (C.c_int32 * 2)(1, 2)
    assert nat.ems_cn_create(0, 2, 8, 2, dil, 0, C.byref(h)) == nat.EMS_E_INVALID
    assert nat.ems_cn_create(4, 2, 8, 1, dil, 0, C.byref(h)) == nat.EMS_E_INVALID
    assert nat.ems_cn_forward(None, None, 1, 1, None) == nat.EMS_E_INVALID
