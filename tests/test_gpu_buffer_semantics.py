"""The raw-buffer range check the STREAM kernels rely on, measured on the
device: (1) the SGPR offset is part of the checked offset, (2) a 16-byte load
that straddles the record count returns its in-range dwords and zeros beyond.
Every probed address lies inside a larger allocation, so the probes read
valid memory whatever the check does; only the returned values differ."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _launch(fn_name, records, vbase, soff, n_out):
    import torch

    from threadcoarsenedstencilgenerator_amd.build import PROBE_CO
    hip = C.CDLL("libamdhip64.so")
    mod, fn = C.c_void_p(), C.c_void_p()
    assert hip.hipModuleLoad(C.byref(mod), PROBE_CO.encode()) == 0
    assert hip.hipModuleGetFunction(C.byref(fn), mod, fn_name.encode()) == 0
    buf = torch.arange(1, (1 << 20) + 1, dtype=torch.int32, device="cuda")  # 4 MiB, value = index + 1
    out = torch.zeros(n_out, dtype=torch.int32, device="cuda")

    class A(C.Structure):
        _fields_ = [("p", C.c_void_p), ("records", C.c_uint32), ("vbase", C.c_uint32), ("soff", C.c_uint32),
                    ("pad", C.c_uint32), ("out", C.c_void_p)]

    a = A(buf.data_ptr(), records, vbase, soff, 0, out.data_ptr())
    sz = C.c_size_t(C.sizeof(a))
    extra = (C.c_void_p * 5)(C.c_void_p(1), C.cast(C.pointer(a), C.c_void_p), C.c_void_p(2),
                             C.cast(C.pointer(sz), C.c_void_p), C.c_void_p(3))
    assert hip.hipModuleLaunchKernel(fn, 1, 1, 1, 64, 1, 1, 0, None, None, extra) == 0
    torch.cuda.synchronize()
    hip.hipModuleUnload(mod)
    return out.cpu().numpy()


def test_sgpr_offset_is_range_checked():
    records = 1 << 20
    got = _launch("tcsg_probe_b32", records, records - 128, 64 * 1024, 64)
    # per-lane offsets below the record count, full offsets past it: all read 0
    assert np.all(got == 0)
    got = _launch("tcsg_probe_b32", records, 0, 4096, 64)
    assert got[:4].tolist() == [1025, 1026, 1027, 1028]


def test_straddling_b128_is_checked_per_dword():
    got = _launch("tcsg_probe_b128", 56, 0, 0, 256).reshape(64, 4)
    assert got[2].tolist() == [9, 10, 11, 12]
    assert got[3].tolist() == [13, 14, 0, 0]
    assert np.all(got[4:] == 0)
    got = _launch("tcsg_probe_b128", 56, 0, 16, 256).reshape(64, 4)
    assert got[2].tolist() == [13, 14, 0, 0]
