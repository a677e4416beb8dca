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
        self._destroy = lib().althea_mlp_destroy

    def __del__(self):
        # The destroy function is bound at creation: module globals may be
        # gone when this runs at interpreter exit.
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
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

    def set_learning_rate(self, lr: float) -> None:
        check(lib().althea_mlp_set_learning_rate(self._h, float(lr)), "althea_mlp_set_learning_rate")

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

    def build_monitor(self, prefiltered, size: int, mips: int, threshold: float = 1e-3, stream=None):
        """Judge the MLP against every texel of the prefiltered cube (device
        tensor of cube texels x 4); returns (cell errors as a device tensor
        [mips, 6, 8, 8], number of accepted cells).  The handle keeps the
        acceptance for DeferredRendering.resolve_mlp."""
        torch = _torch()
        err = torch.empty((mips, 6, 8, 8), dtype=torch.float32, device="cuda")
        acc = C.c_int(0)
        check(lib().althea_mlp_build_monitor(self._h, _ptr(prefiltered), size, mips, float(threshold), _ptr(err),
                                             C.byref(acc), _stream(stream)), "althea_mlp_build_monitor")
        return err, acc.value


def mfma_probe(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """D = A . B through one BF16 MFMA with the MLP kernels' lane layout."""
    torch = _torch()
    ta = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    tb = torch.from_numpy(np.ascontiguousarray(b, np.float32)).cuda()
    td = torch.empty((16, 16), dtype=torch.float32, device="cuda")
    check(lib().althea_mfma_bf16_probe(_ptr(ta), _ptr(tb), _ptr(td), _stream()), "althea_mfma_bf16_probe")
    return td.cpu().numpy()


def training_set(pyramid, env_w: int, env_h: int, count: int, seed: int, samples: int = 1024, rmin: float = 0.0,
                 rmax: float = 1.0, stream=None):
    """(direction, roughness) queries, uniform on the sphere and in
    [rmin, rmax], labelled by the plain GPU prefilter with `samples`
    Hammersley samples (float4 radiance per query)."""
    torch = _torch()
    from . import prefilter_points, synth

    dr = torch.from_numpy(synth.dir_rough(seed=seed, n=count, rmin=rmin, rmax=rmax)).cuda()
    return dr, prefilter_points(pyramid, env_w, env_h, samples, dr, stream=stream)


def fit(net: PrefilterMLP, dir_rough, labels, batch: int = 65536, epochs: int = 8, lr: float | None = None,
        final_lr: float | None = None, stream=None) -> list:
    """Adam over the training set in contiguous batches (the queries are
    i.i.d., so a fixed order is a random order), `epochs` passes.  With `lr`
    and `final_lr` the step size decays geometrically from one to the other
    across the epochs.  Returns the loss of each epoch's first batch (mean
    squared log error before its step)."""
    torch = _torch()
    n = int(dir_rough.shape[0])
    steps = max(1, n // batch)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    history = []
    for ep in range(epochs):
        if lr is not None:
            f = ep / max(1, epochs - 1)
            net.set_learning_rate(lr * ((final_lr / lr) ** f if final_lr else 1.0))
        for k in range(steps):
            lo = k * batch
            hi = n if k == steps - 1 else lo + batch
            net.train_step(dir_rough[lo:hi], labels[lo:hi], loss_out=loss if k == 0 else None, stream=stream)
            if k == 0:
                history.append(loss)
                loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    return [float(h.item()) for h in history]
