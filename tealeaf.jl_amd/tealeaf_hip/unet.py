"""Learned warm start: host-side helpers over the U-Net entry points of
libtealeaf_hip (training data recording, training, weights I/O).

The network itself (forward, backward, Adam) runs in the library's HIP
kernels; nothing here computes on fields.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, load


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def param_count() -> int:
    return int(load().tl_unet_param_count())


def set_weights(chunk, w: Optional[np.ndarray]):
    """Load flat FP32 weights (None: the library's deterministic He init)."""
    L = load()
    if w is None:
        check(L.tl_unet_set_weights(chunk._ctx, None), chunk._ctx, "tl_unet_set_weights")
        return
    w = np.ascontiguousarray(w, dtype=np.float32)
    assert w.size == param_count()
    check(L.tl_unet_set_weights(chunk._ctx, _fp(w)), chunk._ctx, "tl_unet_set_weights")


def get_weights(chunk) -> np.ndarray:
    w = np.zeros(param_count(), dtype=np.float32)
    check(load().tl_unet_get_weights(chunk._ctx, _fp(w)), chunk._ctx, "tl_unet_get_weights")
    return w


def load_weights(chunk, path: str):
    set_weights(chunk, np.load(path, allow_pickle=False))


def save_weights(chunk, path: str):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    np.save(path, get_weights(chunk), allow_pickle=False)


def record(chunk, capacity: int, crops_per_step: int = 8):
    check(load().tl_unet_record(chunk._ctx, capacity, crops_per_step), chunk._ctx, "tl_unet_record")


def dataset_size(chunk) -> int:
    return int(load().tl_unet_dataset_size(chunk._ctx))


def train(chunk, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0, loss: str = "energy") -> np.ndarray:
    """Adam steps on the error e = pred - (u_new - u_prev)/s_u over the crops:
    loss "energy" minimises e^T A e (the norm CG minimises), "residual"
    ||A e||^2, "mse" ||e||^2."""
    tp = _lib.TrainParams()
    tp.steps, tp.batch, tp.crop, tp.lr, tp.seed = steps, batch, 128, lr, seed
    tp.reserved[0] = {"residual": 0, "mse": 1, "energy": 2}[loss]
    out = np.zeros(steps, dtype=np.float32)
    check(load().tl_unet_train(chunk._ctx, C.byref(tp), _fp(out)), chunk._ctx, "tl_unet_train")
    return out


def evaluate(chunk, x: np.ndarray, target: Optional[np.ndarray] = None):
    """Network output for x[B, H, W, 6]; with a target also the MSE loss and
    the flat gradient (diagnostics for the parity tests)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    B, H, W, _ = x.shape
    out = np.zeros((B, H, W), dtype=np.float32)
    L = load()
    if target is None:
        check(L.tl_unet_eval(chunk._ctx, _fp(x), None, B, H, W, _fp(out), None, None), chunk._ctx, "tl_unet_eval")
        return out
    t = np.ascontiguousarray(target, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    g = np.zeros(param_count(), dtype=np.float32)
    check(L.tl_unet_eval(chunk._ctx, _fp(x), _fp(t), B, H, W, _fp(out), _fp(loss), _fp(g)), chunk._ctx,
          "tl_unet_eval")
    return out, float(loss[0]), g


def predict(chunk, rx: float, ry: float, coefficient: int = 1) -> np.ndarray:
    """The warm-start prediction of u for the chunk's current state (host copy)."""
    check(load().tl_unet_predict(chunk._ctx, rx, ry, coefficient), chunk._ctx, "tl_unet_predict")
    return chunk.field("mi")
