"""FP8 (OCP e4m3) restatement of the FP8 inference forward, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The reference has no FP8 path (it runs FP16 throughout), so this restates the
arithmetic the FP8 path defines, on top of the reference's forward
(NN_nv.glsl:95-160: per layer y = W_l x, ReLU between layers, none after the
output layer; nrc_inference.comp:48 clamps the prediction at 0):

  * e4m3 (OCP 8-bit floating point, "E4M3FN"): 1 sign, 4 exponent (bias 7),
    3 mantissa bits; normals 2^-6 .. 448, subnormals k 2^-9; no infinities;
    round to nearest, ties to even.  Values beyond 448 saturate to +-448 (the
    kernel clamps before converting).
  * power-of-two scales C_0 = 1, C_1, .., C_H: the input features enter as
    e4m3(f); layer l's weights as e4m3(sat(c_l W_l)) with c_l = C_{l+1} / C_l;
    its output z (exact sum in the kernel's FP32 MFMA accumulator) becomes the
    next input e4m3(min(max(z, 0), 448)); the output layer's weights enter as
    e4m3(sat(c_H W_H)) and the prediction is (c_H W_H . a_H) * out_scale with
    out_scale = 1 / (c_H C_H).

Pinned by tests/test_oracle_fp8.py: the e4m3 rounding against the format's
defining values (largest normal, smallest subnormal, ties to even, saturation,
every representable value round-tripping), and the forward against the FP64
forward on inputs and weights that e4m3 holds exactly (where quantization is
the identity and only the scale bookkeeping remains).
"""
from __future__ import annotations

import numpy as np

E4M3_MAX = 448.0
E4M3_MIN_NORMAL_EXP = -6
E4M3_MANTISSA_BITS = 3


def e4m3_round(x) -> np.ndarray:
    """Nearest e4m3 value (ties to even), saturating at +-448; NaN stays NaN."""
    x = np.asarray(x, np.float64)
    a = np.minimum(np.abs(x), E4M3_MAX)
    # exponent of the binade, floored at the subnormal range's fixed exponent
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, E4M3_MIN_NORMAL_EXP)
    step = np.exp2(e - E4M3_MANTISSA_BITS)
    q = np.round(a / step) * step  # np.round: half to even
    return np.copysign(np.minimum(q, E4M3_MAX), x)


def e4m3_values() -> np.ndarray:
    """The 254 finite e4m3 values (every bit pattern but the two NaNs, +0 and -0 both), decoded from
    the bit layout."""
    vals = []
    for b in range(256):
        s, e, m = b >> 7, (b >> 3) & 15, b & 7
        if e == 15 and m == 7:
            continue  # NaN
        v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1.0 + m / 8.0) * 2.0 ** (e - 7)
        vals.append(-v if s else v)
    return np.array(vals, np.float64)


def infer_fp8(feats, weights, hidden: int, layer_scales, out_scale: float, clamp0: bool = True) -> np.ndarray:
    """FP8 forward of encoded features [n][64] with the FP32 use-weights (row-major, layer l at
    l*4096, W[row][col] multiplies input col into output row) and the scales of one inference."""
    f = np.asarray(feats, np.float64)
    w = np.asarray(weights, np.float64)
    C = np.asarray(layer_scales, np.float64)
    assert C.shape[0] == hidden + 1 and C[0] == 1.0
    x = e4m3_round(f)
    for l in range(hidden):
        W = w[l * 4096:(l + 1) * 4096].reshape(64, 64)
        c = C[l + 1] / C[l]
        Wq = e4m3_round(np.clip(W * c, -E4M3_MAX, E4M3_MAX))
        z = x @ Wq.T
        x = e4m3_round(np.clip(z, 0.0, E4M3_MAX))
    Wo = w[hidden * 4096:hidden * 4096 + 192].reshape(3, 64)
    cH = 1.0 / (out_scale * C[hidden])
    Wq = e4m3_round(np.clip(Wo * cH, -E4M3_MAX, E4M3_MAX))
    y = (x @ Wq.T) * out_scale
    return np.maximum(y, 0.0) if clamp0 else y
