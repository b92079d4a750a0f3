"""Host NQT log10/pow10 (singularity-eos/base/fast-math/logs.hpp:188-214,
282-292; arXiv:2206.08957) for building table grids in NQT log space, exactly
as singularity-eos's Bounds::convertBoundsToLog_ does
(singularity-eos/base/spiner_table_utils.hpp:185-206). Integer arithmetic
on int64, identical to the device code's."""
import numpy as np

_ONE_AS_INT = (1 << 62) - (1 << 52)
_SCALE_DOWN = 2.22044604925031e-16
_MANT = (1 << 52) - 1
_LOW = (1 << 26) - 1


def _c_div3(v):
    # C integer division truncates toward zero
    q = abs(v) // 3
    return q if v >= 0 else -q


def lg(x):
    xi = int(np.float64(x).view(np.int64)) - _ONE_AS_INT
    fr = xi & _MANT
    hi = fr >> 26
    lo = fr & _LOW
    sq = hi * hi + ((hi * lo) >> 25)
    return float(np.float64(xi + _c_div3(fr - sq)) * np.float64(_SCALE_DOWN))


def log10(x):
    return float(np.float64(0.301029995663981195) * np.float64(lg(x)))
