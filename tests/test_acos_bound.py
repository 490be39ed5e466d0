"""Error bound of the force kernels' arccos (acos_as, oxdna_device.h), built for the host
from the same header (tests/cpp/libacos_check.so), against FP64 libm over [-1, 1].

The polynomial is Abramowitz & Stegun 4.4.46 (|error| <= 2e-8 rad in exact arithmetic);
evaluated in FP32 the bound is 5e-7 rad absolute (a few units in the last place of an
angle up to pi). The sine it returns, sqrt((1 - |x|)(1 + |x|)), is checked to 2e-7
relative where it exceeds 1e-3. A dense grid plus the points next to +-1 (where arccos
is steepest) and the reflection x < 0 -> pi - acos(-x) are all covered.
"""

import ctypes as C
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(__file__), "cpp", "libacos_check.so")


@pytest.fixture(scope="module")
def acos_as():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is missing: build it (make -C tests/cpp, part of __graft_entry__.build())")
    L = C.CDLL(LIB)
    fp = C.POINTER(C.c_float)
    L.acos_as_batch.argtypes = [fp, fp, fp, C.c_int]

    def call(x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        th, s = np.empty_like(x), np.empty_like(x)
        L.acos_as_batch(x.ctypes.data_as(fp), th.ctypes.data_as(fp), s.ctypes.data_as(fp), len(x))
        return th.astype(np.float64), s.astype(np.float64)
    return call


def _grid():
    one = np.float32(1.0)
    near = one - np.arange(4096, dtype=np.float32) * np.spacing(np.float32(0.5))
    x = np.concatenate([np.linspace(-1.0, 1.0, 1_000_001, dtype=np.float32), near, -near,
                        np.float32([0.0, -0.0, 0.5, -0.5])])
    return x[(x >= -1) & (x <= 1)]


def test_acos_error_bound(acos_as):
    x = _grid()
    th, s = acos_as(x)
    ref = np.arccos(x.astype(np.float64))
    err = np.abs(th - ref)
    assert err.max() <= 5e-7, (err.max(), x[err.argmax()])
    sref = np.sqrt(np.maximum(0.0, (1.0 - np.abs(x.astype(np.float64))) * (1.0 + np.abs(x.astype(np.float64)))))
    big = sref > 1e-3
    assert (np.abs(s - sref)[big] / sref[big]).max() <= 2e-7
    assert np.abs(s - sref).max() <= 1e-7


def test_acos_exact_points(acos_as):
    th, s = acos_as(np.float32([1.0, -1.0, 0.0]))
    assert th[0] == 0.0 and s[0] == 0.0
    assert abs(th[1] - np.pi) <= 5e-7 and s[1] == 0.0
    assert abs(th[2] - np.pi / 2) <= 5e-7 and s[2] == 1.0
