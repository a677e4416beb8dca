"""RGBE (.hdr) files through the C ABI: the reference's environment-map and
prefiltered-cache format (read and written with stb_image there)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib


def info(path: str | os.PathLike) -> tuple[int, int]:
    """(width, height) of an .hdr file."""
    w, h = C.c_int(0), C.c_int(0)
    check(lib().althea_hdr_info(os.fsencode(path), C.byref(w), C.byref(h)), "althea_hdr_info")
    return w.value, h.value


def read(path: str | os.PathLike) -> np.ndarray:
    """The image as float32 (height, width, 4), alpha 1."""
    w, h = info(path)
    out = np.empty((h, w, 4), np.float32)
    check(lib().althea_hdr_read(os.fsencode(path), out.ctypes.data, out.size), "althea_hdr_read")
    return out


def write(path: str | os.PathLike, rgba: np.ndarray) -> None:
    """Writes (height, width, 4) or (height, width, 3) float data as RGBE."""
    a = np.asarray(rgba, np.float32)
    if a.shape[-1] == 3:
        a = np.concatenate([a, np.ones(a.shape[:-1] + (1,), np.float32)], -1)
    a = np.ascontiguousarray(a)
    h, w = a.shape[:2]
    check(lib().althea_hdr_write(os.fsencode(path), a.ctypes.data, w, h), "althea_hdr_write")
