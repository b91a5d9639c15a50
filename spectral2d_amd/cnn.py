"""The closure CNN (s2d_cnn_* in include/s2d.h): six periodic 3x3
convolutions (2 -> 64 x 5 -> 1 channels, GELU), BF16 weights and
activations with FP32 accumulation, trained with the library's own
backward kernels and Adam step. No reference counterpart.

Data arguments are device pointers (int addresses) to FP32 planes: sample
n's plane at ptr + n * stride floats, rows ld floats apart. torch tensors on
the GPU are convenient holders (``t.data_ptr()``); the arithmetic is the
library's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check_cnn


def param_count() -> int:
    return int(L.lib().s2d_cnn_param_count())


class ClosureNet:
    def __init__(self, H: int, W: int, batch_max: int):
        h = C.c_void_p()
        check_cnn(L.lib().s2d_cnn_create(int(H), int(W), int(batch_max), C.byref(h)))
        self._h = h
        self.H, self.W, self.batch_max = H, W, batch_max
        self.n = param_count()
        self.scales = (1.0, 1.0, 1.0)
        self.clip = False

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().s2d_cnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        check_cnn(rc, self._h)

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        self._c(L.lib().s2d_cnn_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n, np.float32)
        self._c(L.lib().s2d_cnn_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_scales(self, in0: float, in1: float, out: float):
        self.scales = (float(in0), float(in1), float(out))
        self._c(L.lib().s2d_cnn_set_scales(self._h, *self.scales))

    def set_clip(self, clip: bool):
        """Keep only the enstrophy-sink part of tau (omega * tau > 0)."""
        self._c(L.lib().s2d_cnn_set_clip(self._h, int(bool(clip))))
        self.clip = bool(clip)

    def set_launch_width(self, blocks: int):
        """Every kernel at the fewest blocks >= `blocks` (grid-stride ones at
        most `blocks`); 0 restores the defaults. Same results; for tests of
        the multi-row paths at small sizes."""
        self._c(L.lib().s2d_cnn_set_launch_width(self._h, int(blocks)))

    def forward(self, B, w_ptr, psi_ptr, stride, ld, tau_ptr, tau_stride, tau_ld):
        self._c(L.lib().s2d_cnn_forward(self._h, int(B), C.c_void_p(w_ptr), C.c_void_p(psi_ptr),
                                        int(stride), int(ld), C.c_void_p(tau_ptr),
                                        int(tau_stride), int(tau_ld)))

    def gradient(self, B, w_ptr, psi_ptr, stride, ld, tg_ptr, tg_stride, tg_ld):
        loss = C.c_double()
        g = np.zeros(self.n, np.float32)
        self._c(L.lib().s2d_cnn_gradient(self._h, int(B), C.c_void_p(w_ptr), C.c_void_p(psi_ptr),
                                         int(stride), int(ld), C.c_void_p(tg_ptr),
                                         int(tg_stride), int(tg_ld), C.byref(loss),
                                         g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train_step(self, B, w_ptr, psi_ptr, stride, ld, tg_ptr, tg_stride, tg_ld, lr):
        loss = C.c_double()
        self._c(L.lib().s2d_cnn_train_step(self._h, int(B), C.c_void_p(w_ptr),
                                           C.c_void_p(psi_ptr), int(stride), int(ld),
                                           C.c_void_p(tg_ptr), int(tg_stride), int(tg_ld),
                                           float(lr), C.byref(loss)))
        return loss.value


    def set_dataset(self, w: np.ndarray, psi: np.ndarray, tau: np.ndarray):
        """Host arrays (n, H, W) copied to the device for train_batch."""
        arrs = [np.ascontiguousarray(x, np.float32) for x in (w, psi, tau)]
        n = arrs[0].shape[0]
        fp = C.POINTER(C.c_float)
        self._c(L.lib().s2d_cnn_set_dataset(self._h, n, *[x.ctypes.data_as(fp) for x in arrs]))
        self.ds_n = n

    def train_batch(self, idx, lr: float) -> float:
        idx = np.ascontiguousarray(idx, np.int32)
        loss = C.c_double()
        self._c(L.lib().s2d_cnn_train_batch(self._h, int(idx.size),
                                            idx.ctypes.data_as(C.POINTER(C.c_int)), float(lr),
                                            C.byref(loss)))
        return loss.value

    def predict(self, i: int) -> np.ndarray:
        out = np.zeros((self.H, self.W), np.float32)
        self._c(L.lib().s2d_cnn_predict(self._h, int(i),
                                        out.ctypes.data_as(C.POINTER(C.c_float))))
        return out


__all__ = ["ClosureNet", "param_count"]
