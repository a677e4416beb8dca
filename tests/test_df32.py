"""Double-float arithmetic of althea_amd/csrc/df32.hpp (the kernels' precise
addressing), built for the host (once as the host path, once as the device
path with a reciprocal one ulp off) and checked against FP64: atan2 to about
1e-13 rad over every octant, division and square root to ~2^-44 relative,
and the texel split (floor plus weights a, b = 1 - a, each one FP32 rounding
of the exact fraction) including coordinates just below and above integers."""
import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module", params=["host", "device_rcp"])
def lib(request, tmp_path_factory):
    """The header built for the host (df_rcp = 1 / x) and with the device code
    path (df_rcp = the hardware reciprocal, modelled as 1 / x one ulp off)."""
    so = tmp_path_factory.mktemp("df32") / f"libdf32_{request.param}.so"
    defs = ["-DDF32_DEVICE_RCP"] if request.param == "device_rcp" else []
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", *defs,
                    str(ROOT / "tests" / "cpp" / "df32_host.cpp"), "-o", str(so)], check=True, capture_output=True)
    L = C.CDLL(str(so))
    vp = C.c_void_p
    L.df32_atan2_batch.argtypes = [vp, vp, C.c_int, vp]
    L.df32_div_sqrt_batch.argtypes = [vp, vp, C.c_int, vp, vp]
    L.df32_texel_batch.argtypes = [vp, C.c_int, vp, vp, vp]
    return L


def _split(v):
    hi = np.asarray(v, np.float64).astype(np.float32)
    lo = (np.asarray(v, np.float64) - hi.astype(np.float64)).astype(np.float32)
    return np.ascontiguousarray(np.stack([hi, lo], -1))


def _val(pairs):
    return pairs[..., 0].astype(np.float64) + pairs[..., 1].astype(np.float64)


def test_atan2_every_octant(lib):
    rng = np.random.default_rng(0)
    ang = rng.uniform(-math.pi, math.pi, 200000)
    rad = np.exp(rng.uniform(-8, 8, ang.size))
    y, x = rad * np.sin(ang), rad * np.cos(ang)
    # exact axes, diagonals and tiny components
    extra = np.array([[0.0, 1.0], [1.0, 0.0], [0.0, -1.0], [-1.0, 0.0], [1.0, 1.0], [-1.0, 1.0], [1e-30, 1.0],
                      [1.0, -1e-30], [-3.0, -3.0], [2048.0, 1.0], [1.0, 2048.0]])
    y = np.concatenate([y, extra[:, 0]])
    x = np.concatenate([x, extra[:, 1]])
    ys, xs = _split(y), _split(x)
    out = np.empty_like(ys)
    lib.df32_atan2_batch(ys.ctypes.data, xs.ctypes.data, ys.shape[0], out.ctypes.data)
    ref = np.arctan2(_val(ys), _val(xs))
    err = np.abs(_val(out) - ref)
    err = np.minimum(err, np.abs(err - 2 * math.pi))  # +-pi on the negative x axis
    assert err.max() < 3e-13, err.max()
    # The FP32 pair itself: hi is the FP32 rounding of the result.
    assert np.all(np.abs(out[:, 1]) <= np.spacing(np.abs(out[:, 0])) + 1e-45)


def test_division_and_square_root(lib):
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, 100000) * np.exp(rng.uniform(-20, 20, 100000))
    b = rng.uniform(0.1, 1, 100000) * np.exp(rng.uniform(-20, 20, 100000)) * rng.choice([-1, 1], 100000)
    A, B = _split(a), _split(b)
    q, s = np.empty_like(A), np.empty_like(A)
    lib.df32_div_sqrt_batch(A.ctypes.data, B.ctypes.data, A.shape[0], q.ctypes.data, s.ctypes.data)
    rq = _val(A) / _val(B)
    assert np.max(np.abs(_val(q) - rq) / np.abs(rq)) < 2.0 ** -43
    rs = np.sqrt(np.abs(_val(A)))
    assert np.max(np.abs(_val(s) - rs) / rs) < 2.0 ** -43


def test_texel_split(lib):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-0.5, 4096.0, 100000),
                        np.arange(-1, 4097) + 1e-9, np.arange(-1, 4097) - 1e-9, np.arange(0, 4097, dtype=float)])
    X = _split(x)
    idx = np.empty(X.shape[0], np.int32)
    a = np.empty(X.shape[0], np.float32)
    b = np.empty(X.shape[0], np.float32)
    lib.df32_texel_batch(X.ctypes.data, X.shape[0], idx.ctypes.data, a.ctypes.data, b.ctypes.data)
    v = _val(X)
    f = np.floor(v)
    np.testing.assert_array_equal(idx, f.astype(np.int32))
    frac = v - f
    assert np.all((a >= 0) & (a <= 1) & (b > 0) & (b <= 1))
    # each weight is one rounding of the exact value: within half an FP32 ulp
    assert np.all(np.abs(a.astype(np.float64) - frac) <= np.spacing(np.maximum(a, 1e-30).astype(np.float32)) * 0.5001)
    assert np.all(np.abs(b.astype(np.float64) - (1 - frac)) <= np.spacing(b) * 0.5001)
