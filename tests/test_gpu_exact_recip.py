"""The exact kernels' reciprocal (recip_rn, tpr_amd/csrc/tpr_device.hpp)
against IEEE division, 1.0f / x, for every one of the 2^32 float inputs on
the GPU (tpr_amd/diag/exact_div.hip): the same bits for all of them, so
exact mode's answers stay the reference's bit for bit wherever the kernels
use it.  The unguarded four-instruction sequence is counted too, by range,
to show what the guard is for."""
import ctypes
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DIAG = os.path.join(os.path.dirname(HERE), "tpr_amd", "libtpr_diag.so")


@pytest.mark.gpu
def test_recip_rn_equals_ieee_division_on_every_float():
    import torch
    from tpr_amd import _lib
    _lib.lib()
    if not os.path.exists(DIAG):
        pytest.fail("tpr_amd/libtpr_diag.so is not built: run `python -m tpr_amd.build`")
    L = ctypes.CDLL(DIAG)
    L.tprh_diag_recip_check.restype = ctypes.c_int
    counts = torch.zeros(6, dtype=torch.int64, device="cuda")
    bad = torch.zeros(64, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("recip check over 2^32 inputs", flush=True)
    assert L.tprh_diag_recip_check(ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(bad.data_ptr()), stream) == 0
    torch.cuda.synchronize()
    c = counts.cpu().tolist()
    print("guarded %d, unguarded %d (subnormal %d, edge of range %d, other %d)" % tuple(c[:5]), flush=True)
    first = [hex(v & 0xFFFFFFFF) for v in bad.cpu().tolist()[:min(c[5], 8)]]
    assert c[0] == 0, first
    assert c[4] == 0
