"""Learned interface normal (include/simplic_learned.h): the network, its
GPU training set and GPU training.  All work runs in libsimplic_amd.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import SVP_HOST, _check, lib

N_PARAMS = 20483
SHAPE_DOUBLES = 19


class NormalModel:
    """27 -> 128 -> 128 -> 3 GELU network on one GPU."""

    def __init__(self, seed: int = 0, device: int = 0):
        h = C.c_void_p()
        _check(lib().svp_nm_create(int(device), C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.n_data = 0

    def close(self):
        if getattr(self, "h", None):
            lib().svp_nm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, np.float32)
        _check(lib().svp_nm_get_params(self.h, p.ctypes.data, N_PARAMS))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().svp_nm_set_params(self.h, p.ctypes.data, int(p.size)))

    def generate(self, n: int, seed: int = 1, depth: int = 5):
        _check(lib().svp_nm_generate(self.h, int(n), C.c_uint64(seed), int(depth)))
        self.n_data = int(n)

    def dataset(self, first: int, count: int):
        """dict of alpha27 (count,27), shape (count,19), x (count,32) uint16, y (count,4)."""
        a = np.zeros((count, 27))
        s = np.zeros((count, SHAPE_DOUBLES))
        x = np.zeros((count, 32), np.uint16)
        y = np.zeros((count, 4), np.float32)
        _check(lib().svp_nm_dataset_get(self.h, int(first), int(count), a.ctypes.data, s.ctypes.data,
                                        x.ctypes.data, y.ctypes.data))
        return {"alpha27": a, "shape": s, "x": x, "y": y}

    def train(self, epochs: int, batch: int = 4096, lr0: float = 3e-3, lr1: float = 3e-5, seed: int = 7):
        losses = np.zeros(max(epochs, 1), np.float32)
        sec = C.c_double()
        _check(lib().svp_nm_train(self.h, int(epochs), int(batch), C.c_float(lr0), C.c_float(lr1),
                                  C.c_uint64(seed), losses.ctypes.data, C.byref(sec)))
        return losses[:epochs], sec.value

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint16)
        out = np.zeros((x.shape[0], 3), np.float32)
        _check(lib().svp_nm_forward(self.h, x.ctypes.data, int(x.shape[0]), out.ctypes.data, SVP_HOST))
        return out

    def gradient(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.uint16)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().svp_nm_gradient(self.h, x.ctypes.data, y.ctypes.data, int(x.shape[0]), g.ctypes.data,
                                     C.byref(loss), SVP_HOST))
        return g, loss.value


def cube_fractions(planes: np.ndarray, device: int = 0) -> np.ndarray:
    """Liquid fraction of cubes [lo, lo+h]^3 on the side n.x + d < 0; rows nx ny nz d lox loy loz h."""
    planes = np.ascontiguousarray(planes, dtype=np.float64).reshape(-1, 8)
    out = np.zeros(planes.shape[0])
    _check(lib().svp_nm_cube_fractions(planes.ctypes.data, int(planes.shape[0]), out.ctypes.data, int(device)))
    return out
