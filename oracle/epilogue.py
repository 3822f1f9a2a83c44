"""CPU restatement of the inference epilogue (screen composite and training-target
back-fill), in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows nrc_inference.comp:50-72 and the destination codec of NRCRecord.glsl:12-33:
  * EncodeNRCEvalDstScreen(xy)   = (x | y << 15) << 1
  * EncodeNRCEvalDstTrain(b, l, r) = (b | l << 2 | r << 16) << 1 | 1
  * GetNRCEvalDstType(e)         = e & 1 (0 screen, 1 train); NRC_EVAL_INVALID_DST = ~0u
  * DecodeNRCEvalDstScreen(e)    : e >>= 1; (e & 0x7FFF, e >> 15)
  * DecodeNRCEvalDstTrain(e)     : e >>= 1; (e & 3, (e >> 2) & 0x3FFF, e >> 16)
  * screen: color = bias_factor_r.rgb + (bias_factor_r.a, factor_gb.rg) * predict,
            stored as (color, 0);
  * train:  for i in l..r: bias_i = bias_i + factor_i * predict.
Each product and sum is rounded to FP32 separately.  Pixels outside the image and
records past the batch are skipped (a robust store drops them).

Pinned by tests/test_oracle_epilogue.py: the codec against hand-evaluated codes
and round trips, the composite against closed forms on exactly representable
values, and the skip rules.
"""
from __future__ import annotations

import numpy as np

INVALID = 0xFFFFFFFF


def dst_screen(x, y):
    return ((np.asarray(x, np.uint64) | (np.asarray(y, np.uint64) << 15)) << 1).astype(np.uint32)


def dst_train(b, l, r):
    v = np.asarray(b, np.uint64) | (np.asarray(l, np.uint64) << 2) | (np.asarray(r, np.uint64) << 16)
    return ((v << 1) | 1).astype(np.uint32)


def decode(e):
    """(kind, a, b, c): kind 0 screen (x, y, -), 1 train (batch, l, r), -1 invalid."""
    e = int(e)
    if e == INVALID:
        return -1, 0, 0, 0
    v = e >> 1
    if e & 1 == 0:
        return 0, v & 0x7FFF, v >> 15, 0
    return 1, v & 3, (v >> 2) & 0x3FFF, v >> 16


def composite(rgb, dst, count, bias_factor_r, factor_gb, width, height, batches, factors, batch_stride):
    """Applies the epilogue in place to copies and returns (bias_factor_r, batches).  bias_factor_r is
    [height][pitch][4] float32, factor_gb [height][pitch][2]; batches [4 * batch_stride][17] float32
    (target in columns 14..16), factors [4 * batch_stride][3]."""
    f32 = np.float32
    bf = None if bias_factor_r is None else np.array(bias_factor_r, np.float32, copy=True)
    bt = None if batches is None else np.array(batches, np.float32, copy=True)
    for i in range(count):
        kind, a, b, c = decode(dst[i])
        p = np.asarray(rgb[i], np.float32)
        if kind == 0:
            if bf is None or a >= width or b >= height:
                continue
            px = bf[b, a]
            fgb = factor_gb[b, a]
            fac = np.array([px[3], fgb[0], fgb[1]], np.float32)
            col = (px[:3] + (fac * p).astype(f32)).astype(f32)
            bf[b, a, :3] = col
            bf[b, a, 3] = f32(0)
        elif kind == 1:
            if bt is None:
                continue
            for j in range(b, min(c, batch_stride - 1) + 1):
                row = a * batch_stride + j
                bt[row, 14:17] = (bt[row, 14:17] + (np.asarray(factors[row], np.float32) * p).astype(f32)).astype(f32)
    return bf, bt
