"""ctypes front-end of the CPU oracle (nrc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, never by the product package.  ``Oracle('f64')``
is the FP64 checker; ``Oracle('f32')`` is the FP32 CPU baseline built from the
same source.

Record layouts (float32):
  query  [14]: position xyz, scattered_dir uv, normal uv, roughness, diffuse rgb, specular rgb
               (UnpackedNRCInput, shader/src/NRCRecord.glsl:40-45)
  train  [17]: query + target rgb
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WIDTH = 64
OUT = 3
REC = 14
TREC = 17
ENC_TRI12, ENC_SINCOS6, ENC_UV_SINCOS16 = 0, 1, 2
LOSS_REL_L2_LUM, LOSS_L2 = 0, 1


def weight_count(hidden: int) -> int:
    return hidden * WIDTH * WIDTH + OUT * WIDTH


def build() -> None:
    """Compile both oracle libraries (gcc, OpenMP), and the Eigen CPU baseline where the
    reference's vendored Eigen is present (oracle/eigen.py)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "all"], check=True)
    from oracle import eigen
    eigen.build()


class OptParams(C.Structure):
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double),
                ("ema_alpha", C.c_double), ("dynamic_scale", C.c_int), ("growth_interval", C.c_int)]


class OptState(C.Structure):
    _fields_ = [("t", C.c_double), ("beta1_t", C.c_double), ("beta2_t", C.c_double), ("alpha_t", C.c_double),
                ("alpha_t_1", C.c_double), ("loss_scale", C.c_double), ("clean_steps", C.c_double),
                ("skipped_steps", C.c_double)]


_P = C.c_void_p
_I64 = C.c_int64


class Oracle:
    """numpy wrapper; ``kind`` is 'f64' (oracle) or 'f32' (CPU baseline)."""

    def __init__(self, kind: str = "f64"):
        name = {"f64": "liboracle_f64.so", "f32": "libcpu_f32.so"}[kind]
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.dtype = np.float64 if kind == "f64" else np.float32
        self.kind = kind
        L = self.lib
        R = C.c_double if kind == "f64" else C.c_float
        self._R = R
        L.oracle_round_f16.restype = C.c_double
        L.oracle_round_f16.argtypes = [C.c_double]
        L.oracle_tri.restype = R
        L.oracle_tri.argtypes = [R]
        L.oracle_quartic_cdf.restype = R
        L.oracle_quartic_cdf.argtypes = [R]
        L.oracle_oneblob4.argtypes = [R, _P]
        L.oracle_sph_encode.argtypes = [_P, _P]
        L.oracle_encode_batch.argtypes = [_P, _I64, _I64, C.c_int, C.c_int, _P]
        L.oracle_infer.argtypes = [_P, C.c_int, _P, _I64, _I64, C.c_int, C.c_int, C.c_int, _P]
        L.oracle_infer_encoded.argtypes = [_P, C.c_int, _P, _I64, _P]
        L.oracle_loss_grad.argtypes = [C.c_int, _P, _P, R, _P, _P]
        L.oracle_gradient.argtypes = [_P, C.c_int, _P, _I64, _I64, C.c_int, C.c_int, C.c_int, R, _P, _P]
        L.oracle_gradient_encoded.argtypes = [_P, C.c_int, _P, _P, _I64, C.c_int, R, _P, _P]
        L.oracle_adam_step.restype = C.c_int
        L.oracle_adam_step.argtypes = [C.POINTER(OptParams), C.POINTER(OptState), _P, _I64, _I64, _P, _P, _P, _P]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int

    # -- helpers -------------------------------------------------------------
    def _arr(self, a, dtype=None):
        return np.ascontiguousarray(a, dtype=dtype or self.dtype)

    @staticmethod
    def _ptr(a):
        return a.ctypes.data_as(C.c_void_p)

    def set_threads(self, n: int) -> None:
        self.lib.oracle_set_threads(int(n))

    def threads(self) -> int:
        return int(self.lib.oracle_get_threads())

    # -- primitives ------------------------------------------------------------
    def round_f16(self, x: float) -> float:
        return float(self.lib.oracle_round_f16(float(x)))

    def tri(self, x: float) -> float:
        return float(self.lib.oracle_tri(x))

    def quartic_cdf(self, d: float) -> float:
        return float(self.lib.oracle_quartic_cdf(d))

    def oneblob4(self, x: float) -> np.ndarray:
        o = np.zeros(4, self.dtype)
        self.lib.oracle_oneblob4(self._R(x), self._ptr(o))
        return o

    def sph_encode(self, d) -> np.ndarray:
        d = self._arr(d)
        o = np.zeros(2, self.dtype)
        self.lib.oracle_sph_encode(self._ptr(d), self._ptr(o))
        return o

    def encode(self, recs: np.ndarray, enc: int, round_f16: bool = False) -> np.ndarray:
        recs = np.ascontiguousarray(recs, np.float32)
        n, stride = recs.shape
        out = np.zeros((n, WIDTH), self.dtype)
        self.lib.oracle_encode_batch(self._ptr(recs), n, stride, enc, int(round_f16), self._ptr(out))
        return out

    # -- MLP -------------------------------------------------------------------
    def infer(self, weights, hidden: int, recs, enc: int, round_f16: bool = False, clamp0: bool = True):
        w = self._arr(weights)
        assert w.size == weight_count(hidden)
        recs = np.ascontiguousarray(recs, np.float32)
        n, stride = recs.shape
        out = np.zeros((n, OUT), self.dtype)
        self.lib.oracle_infer(self._ptr(w), hidden, self._ptr(recs), stride, n, enc, int(round_f16), int(clamp0),
                              self._ptr(out))
        return out

    def infer_encoded(self, weights, hidden: int, feats):
        w = self._arr(weights)
        f = self._arr(feats)
        n = f.shape[0]
        out = np.zeros((n, OUT), self.dtype)
        self.lib.oracle_infer_encoded(self._ptr(w), hidden, self._ptr(f), n, self._ptr(out))
        return out

    def loss_grad(self, loss: int, o, t, scale: float = 1.0):
        o = self._arr(o)
        t = self._arr(t)
        g = np.zeros(3, self.dtype)
        v = np.zeros(1, self.dtype)
        self.lib.oracle_loss_grad(loss, self._ptr(o), self._ptr(t), self._R(scale), self._ptr(g), self._ptr(v))
        return g, float(v[0])

    def gradient(self, weights, hidden: int, recs, enc: int, loss: int, scale: float = 1.0, round_f16: bool = False):
        """Summed dW over the records and the summed loss value."""
        w = self._arr(weights)
        recs = np.ascontiguousarray(recs, np.float32)
        n, stride = recs.shape
        assert stride >= TREC
        dw = np.zeros(weight_count(hidden), self.dtype)
        ls = np.zeros(1, self.dtype)
        self.lib.oracle_gradient(self._ptr(w), hidden, self._ptr(recs), stride, n, enc, int(round_f16), loss,
                                 self._R(scale), self._ptr(dw), self._ptr(ls))
        return dw, float(ls[0])

    def gradient_encoded(self, weights, hidden: int, feats, targets, loss: int = LOSS_L2, scale: float = 1.0):
        w = self._arr(weights)
        f = self._arr(feats)
        t = self._arr(targets)
        dw = np.zeros(weight_count(hidden), self.dtype)
        ls = np.zeros(1, self.dtype)
        self.lib.oracle_gradient_encoded(self._ptr(w), hidden, self._ptr(f), self._ptr(t), f.shape[0], loss,
                                         self._R(scale), self._ptr(dw), self._ptr(ls))
        return dw, float(ls[0])


class OracleOptimizer:
    """Adam/EMA optimizer state of one cache, restated on the CPU
    (shader/src/nrc_optimize.comp:32-54, nrc_train_prepare.comp:16-29)."""

    def __init__(self, oracle: Oracle, weights, lr=0.002, beta1=0.9, beta2=0.999, eps=1e-8, ema_alpha=0.99,
                 loss_scale=1.0, dynamic_scale=False, growth_interval=2000):
        self.o = oracle
        dt = oracle.dtype
        self.w = np.array(weights, dtype=dt).copy()
        self.ema = self.w.copy()
        self.m = np.zeros_like(self.w)
        self.v = np.zeros_like(self.w)
        self.params = OptParams(lr, beta1, beta2, eps, ema_alpha, int(dynamic_scale), int(growth_interval))
        self.state = OptState(0.0, 1.0, 1.0, 1.0, 0.0, float(loss_scale), 0.0, 0.0)

    def step(self, G, count: int) -> bool:
        G = np.ascontiguousarray(G, self.o.dtype)
        p = Oracle._ptr
        return bool(self.o.lib.oracle_adam_step(C.byref(self.params), C.byref(self.state), p(G), G.size, int(count),
                                                p(self.m), p(self.v), p(self.w), p(self.ema)))
