"""The MFMA pipe calibration (ztm_calib_mfma): every shape returns a finite
rate below its published peak, low-entropy operands run at least as fast as
random ones on the FP16 shapes (the chip's power limit), and bad arguments
are refused."""
import math

import pytest

from ztm_amd import gfm

pytestmark = pytest.mark.gpu

PEAK = {"32x32x16_f16": 2500.0, "32x32x16_bf16": 2500.0, "16x16x32_f16": 2500.0}


@pytest.mark.parametrize("op", sorted(PEAK))
def test_calibration_rates(dev, op):
    rnd = gfm.calib_mfma(op, low_entropy=False, iters=200)
    low = gfm.calib_mfma(op, low_entropy=True, iters=200)
    for v in (rnd, low):
        assert math.isfinite(v) and 0.0 < v <= 1.05 * PEAK[op], (op, v)
    assert low >= 0.9 * rnd, (op, rnd, low)


def test_calibration_rejects(dev):
    import ctypes
    v = ctypes.c_double()
    assert gfm.lib().ztm_calib_mfma(3, 0, 10, ctypes.byref(v), None) == 1   # ZTM_EINVAL
    assert gfm.lib().ztm_calib_mfma(0, 0, 0, ctypes.byref(v), None) == 1   # ZTM_EINVAL
    assert gfm.lib().ztm_calib_mfma(0, 0, 10, None, None) == 1   # ZTM_EINVAL
