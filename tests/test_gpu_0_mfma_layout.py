"""In-bounds check of the lane layout the grouped GEMM assumes for the two
MFMA builtins it uses (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32):
A[i][k] from lane i+16k, B[k][j] from lane j+16k; D[i][j] in lane
j+16*(i%4), item i//4 for f64, and in lane j+16*(i//4), item i%4 for f32."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from nssdmrg_amd import _ffi as F

pytestmark = pytest.mark.gpu


def probe(prec, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    d = np.zeros(256)
    F.check(F.lib().nss_probe_mfma(prec, F.dptr(a), F.dptr(b), F.dptr(d)))
    return d


@pytest.mark.parametrize("prec", [0, 1])
def test_mfma_lane_layout(prec):
    rng = np.random.default_rng(prec)
    a = rng.integers(-4, 5, 64).astype(float)
    b = rng.integers(-4, 5, 64).astype(float)
    d = probe(prec, a, b)
    out = os.environ.get("NSS_PROBE_DUMP")
    if out:
        with open(out + f".{prec}.json", "w") as f:
            json.dump({"a": a.tolist(), "b": b.tolist(), "d": d.tolist()}, f)
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)])
    B = np.array([[b[j + 16 * k] for j in range(16)] for k in range(4)])
    D = A @ B
    if prec == 0:
        got = np.array([[d[4 * (j + 16 * (i % 4)) + i // 4] for j in range(16)] for i in range(16)])
    else:
        got = np.array([[d[4 * (j + 16 * (i // 4)) + i % 4] for j in range(16)] for i in range(16)])
    assert np.array_equal(got, D), (got, D)
