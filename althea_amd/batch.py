"""Batched prefilter of many environment maps in one persistent,
graph-captured launch (althea_prefilter_batch_* of the C ABI)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

WIDTHS = (256, 512, 1024, 2048, 4096)


def _torch():
    import torch

    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


def batch_widths(count: int, seed_base: int = 0) -> list[int]:
    """Source widths of the batched configuration: drawn uniformly from
    {256, 512, 1024, 2048, 4096} by a SplitMix64 hash of the map index."""
    out = []
    for i in range(count):
        x = (seed_base + i + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 31
        out.append(WIDTHS[x % len(WIDTHS)])
    return out


class PrefilterBatch:
    """`envs`: list of (h, w, 4) float32 device tensors (kept alive here)."""

    def __init__(self, envs, face_sizes, mips: int = 5, samples: int = 1024, fp32_sums: bool = False):
        self.envs = list(envs)
        n = len(self.envs)
        views = (_lib.EnvViewC * n)()
        for i, e in enumerate(self.envs):
            assert e.is_cuda and e.is_contiguous()
            views[i] = _lib.EnvViewC(e.data_ptr(), int(e.shape[1]), int(e.shape[0]))
        sizes = (C.c_int * n)(*[int(s) for s in face_sizes])
        h = C.c_void_p()
        flags = _lib.PRECOMPUTE_FP32_SUMS if fp32_sums else 0
        check(lib().althea_prefilter_batch_create(views, n, sizes, mips, samples, flags, C.byref(h)),
              "althea_prefilter_batch_create")
        self._h = h
        self._destroy = lib().althea_prefilter_batch_destroy
        self.count = n
        self.mips = mips
        self.samples = samples

    def __del__(self):
        # The destroy function is bound at creation: module globals may be
        # gone when this runs at interpreter exit.
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    def run(self, stream=None) -> None:
        check(lib().althea_prefilter_batch_run(self._h, _stream(stream)), "althea_prefilter_batch_run")

    def autotune(self, stream=None) -> dict:
        t = _lib.TuningC()
        check(lib().althea_prefilter_batch_autotune(self._h, C.byref(t), _stream(stream)),
              "althea_prefilter_batch_autotune")
        return {"tile_w": t.tile_w, "tile_h": t.tile_h, "waves_per_cu": t.waves_per_cu}

    def set_tuning(self, tile_w: int, tile_h: int, waves_per_cu: int) -> None:
        t = _lib.TuningC(tile_w, tile_h, waves_per_cu, 0)
        check(lib().althea_prefilter_batch_set_tuning(self._h, C.byref(t)), "althea_prefilter_batch_set_tuning")

    def tuning(self) -> dict:
        t = _lib.TuningC()
        check(lib().althea_prefilter_batch_get_tuning(self._h, C.byref(t)), "althea_prefilter_batch_get_tuning")
        return {"tile_w": t.tile_w, "tile_h": t.tile_h, "waves_per_cu": t.waves_per_cu}

    def texels(self) -> int:
        return lib().althea_prefilter_batch_texels(self._h)

    def cube(self, index: int):
        """Copy of cube `index` as a (texels, 4) device tensor."""
        torch = _torch()
        from . import _hip_copy_dtod

        p, s, m = C.c_void_p(), C.c_int(), C.c_int()
        check(lib().althea_prefilter_batch_cube(self._h, index, C.byref(p), C.byref(s), C.byref(m)),
              "althea_prefilter_batch_cube")
        n = lib().althea_cube_texels(s.value, m.value)
        out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        _hip_copy_dtod(out.data_ptr(), p.value, n * 16)
        return out
