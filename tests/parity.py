"""Error metrics shared by the parity tests and the precision report.

Relative error is |x - ref| / max(|ref|, floor): the floor keeps near-zero
references (BRDF LUT B -> 0 at NdotV -> 1, dark texels) from dominating.
"""
from __future__ import annotations

import numpy as np

# Contract of the plain FP32 path against the FP64 oracle (radiance and LUT).
MAX_REL = 1e-4
# Floors: radiance outputs are O(1..1e3) with a log-normal background of
# median 1; the LUT lies in [0, 1].
RADIANCE_FLOOR = 1e-2
LUT_FLOOR = 1e-2


def rel_err(x, ref, floor: float) -> np.ndarray:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), floor)


def summary(err: np.ndarray) -> dict:
    e = np.asarray(err).ravel()
    return {"max": float(e.max()), "mean": float(e.mean()), "p99": float(np.percentile(e, 99)),
            "p999": float(np.percentile(e, 99.9)), "n": int(e.size)}


def cube_texel_ids(size: int, mips: int, per_face: int, seed: int = 0) -> np.ndarray:
    """(mip, face, x, y) ids: every face's four corners and edge midpoints plus
    `per_face` random texels per face and mip."""
    rng = np.random.default_rng(seed)
    ids = []
    for m in range(mips):
        s = max(1, size >> m)
        for f in range(6):
            fixed = {(0, 0), (s - 1, 0), (0, s - 1), (s - 1, s - 1), (s // 2, 0), (0, s // 2), (s - 1, s // 2),
                     (s // 2, s - 1), (s // 2, s // 2)}
            for x, y in fixed:
                ids.append((m, f, x, y))
            for _ in range(min(per_face, s * s)):
                ids.append((m, f, int(rng.integers(s)), int(rng.integers(s))))
    return np.array(ids, np.int32)


def cube_gather(cube: np.ndarray, size: int, ids: np.ndarray) -> np.ndarray:
    """RGB of cube texels (mip, face, x, y) from a flat (texels, 4) array."""
    out = np.empty((ids.shape[0], 3), np.float32)
    for k, (m, f, x, y) in enumerate(ids):
        s = max(1, size >> int(m))
        off = sum(6 * max(1, size >> j) ** 2 for j in range(int(m)))
        out[k] = cube[off + (int(f) * s + int(y)) * s + int(x), :3]
    return out
