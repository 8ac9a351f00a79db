"""The int8 MFMA calibration launch (measurement only): it runs, reports the
operations it launched, and all-zero operands leave its sink untouched."""
import ctypes

import pytest

import hdc_amd as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wps,qs,rs", [(1, 0, 0), (2, 15, 39)])
def test_calib_runs_and_counts_ops(wps, qs, rs):
    torch = H._torch()
    sink = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    ops = ctypes.c_double()
    iters = 16
    H.check(H.lib().hvk_calib_mfma_i8(wps, iters, qs, rs, H._ptr(sink), ctypes.byref(ops),
                                      H._stream()), "hvk_calib_mfma_i8")
    torch.cuda.synchronize()
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    assert ops.value == 2.0 * 32 ** 3 * 64 * iters * n_cu * wps * 4
    if qs == 0 and rs == 0:
        assert int(sink.item()) == 7  # zero products: the XOR is 0, nothing is stored
