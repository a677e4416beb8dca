"""The learned prefiltered-radiance surrogate (BF16 MFMA MLP) through the C
ABI: create, train on (direction, roughness) -> radiance samples, infer."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch():
    import torch

    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


def _ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous()
    return t.data_ptr()


def param_count() -> int:
    return lib().althea_mlp_param_count()


class PrefilterMLP:
    """Owns an althea_mlp handle."""

    def __init__(self, seed: int = 0x4D4C50, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                 eps: float = 1e-8, init_output_bias: float = 0.0):
        d = _lib.MlpDescC()
        lib().althea_mlp_desc_default(C.byref(d))
        d.seed, d.learning_rate, d.beta1, d.beta2, d.epsilon = seed, lr, beta1, beta2, eps
        d.init_output_bias = init_output_bias
        h = C.c_void_p()
        check(lib().althea_mlp_create(C.byref(d), C.byref(h)), "althea_mlp_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().althea_mlp_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        out = np.empty(param_count(), np.float32)
        check(lib().althea_mlp_get_params(self._h, out.ctypes.data, _stream()), "althea_mlp_get_params")
        return out

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == param_count()
        check(lib().althea_mlp_set_params(self._h, p.ctypes.data, _stream()), "althea_mlp_set_params")

    def train_step(self, dir_rough, targets, loss_out=None, stream=None) -> None:
        check(lib().althea_mlp_train_step(self._h, _ptr(dir_rough), _ptr(targets), int(dir_rough.shape[0]),
                                          _ptr(loss_out), _stream(stream)), "althea_mlp_train_step")

    def gradients(self, dir_rough, targets, stream=None):
        torch = _torch()
        g = torch.empty(param_count(), dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        check(lib().althea_mlp_gradients(self._h, _ptr(dir_rough), _ptr(targets), int(dir_rough.shape[0]), _ptr(g),
                                         _ptr(loss), _stream(stream)), "althea_mlp_gradients")
        return g, loss

    def infer(self, dir_rough, out=None, stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty((dir_rough.shape[0], 4), dtype=torch.float32, device="cuda")
        check(lib().althea_mlp_infer(self._h, _ptr(dir_rough), int(dir_rough.shape[0]), _ptr(out), _stream(stream)),
              "althea_mlp_infer")
        return out

    def state_bytes(self) -> int:
        return lib().althea_mlp_state_bytes(self._h)


def mfma_probe(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """D = A . B through one BF16 MFMA with the MLP kernels' lane layout."""
    torch = _torch()
    ta = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    tb = torch.from_numpy(np.ascontiguousarray(b, np.float32)).cuda()
    td = torch.empty((16, 16), dtype=torch.float32, device="cuda")
    check(lib().althea_mfma_bf16_probe(_ptr(ta), _ptr(tb), _ptr(td), _stream()), "althea_mfma_bf16_probe")
    return td.cpu().numpy()
