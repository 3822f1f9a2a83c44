"""ctypes front-end of the Eigen CPU baseline (oracle/eigen_mlp.cpp -> oracle/_ref/libeigen_mlp.so).

TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg and by tests/.  The
library is built from the Eigen 3.4.0 headers the reference vendors, so it exists
only where it was built (oracle/Makefile target ``ref``) and travels with the tree.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libeigen_mlp.so")
REF = "/root/reference"


def build() -> bool:
    """Builds the library when the reference's Eigen is present; returns whether it exists."""
    if os.path.isdir(os.path.join(REF, "test", "eigen")):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)
    return os.path.exists(LIB)


def available() -> bool:
    return os.path.exists(LIB)


class EigenMLP:
    def __init__(self):
        if not available():
            raise RuntimeError(f"{LIB} is not built (oracle/Makefile target ref)")
        L = C.CDLL(LIB)
        P, I64 = C.c_void_p, C.c_int64
        L.eigen_set_threads.argtypes = [C.c_int]
        L.eigen_set_threads.restype = C.c_int
        L.eigen_forward_f16.argtypes = [P, C.c_int, P, I64, P]
        L.eigen_forward_f32.argtypes = [P, C.c_int, P, I64, P]
        L.eigen_gradient_l2_f32.argtypes = [P, C.c_int, P, P, I64, P]
        L.eigen_gradient_l2_f32.restype = C.c_double
        F = C.c_float
        L.eigen_train_step_l2_f32.argtypes = [P, P, P, C.c_int, P, P, I64, F, F, F, F, F, F]
        L.eigen_train_step_l2_f32.restype = C.c_double
        self.lib = L

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(C.c_void_p)

    def set_threads(self, n: int) -> int:
        return int(self.lib.eigen_set_threads(int(n)))

    def forward_f16(self, w, hidden: int, x) -> np.ndarray:
        w = np.ascontiguousarray(w, np.float16)
        x = np.ascontiguousarray(x, np.float16)
        out = np.zeros((x.shape[0], 3), np.float32)
        self.lib.eigen_forward_f16(self._p(w), hidden, self._p(x), x.shape[0], self._p(out))
        return out

    def forward_f32(self, w, hidden: int, x) -> np.ndarray:
        w = np.ascontiguousarray(w, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], 3), np.float32)
        self.lib.eigen_forward_f32(self._p(w), hidden, self._p(x), x.shape[0], self._p(out))
        return out

    def gradient_l2_f32(self, w, hidden: int, x, t):
        w = np.ascontiguousarray(w, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dw = np.zeros(w.size, np.float32)
        loss = self.lib.eigen_gradient_l2_f32(self._p(w), hidden, self._p(x), self._p(t), x.shape[0], self._p(dw))
        return dw, float(loss)

    def train_step_l2_f32(self, w, m, v, hidden: int, x, t, lr, b1, b2, eps, b1t, b2t) -> float:
        """One Adam step in place on float32 arrays w, m, v; b1t, b2t are beta^t of this step."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        return float(self.lib.eigen_train_step_l2_f32(self._p(w), self._p(m), self._p(v), hidden, self._p(x),
                                                      self._p(t), x.shape[0], lr, b1, b2, eps, b1t, b2t))
