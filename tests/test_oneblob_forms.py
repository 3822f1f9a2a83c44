"""Pins the inference kernels' two-knot one-blob encoder (oneblob4_packed, nrc_mfma.hpp) against the
five-knot form of NRCOneBlob4Encode (shader/src/NRCRecord.glsl:51-63, the form of oneblob4_knots), both
restated here in FP32 with fused multiply-adds as the kernels compute them, and the byte-permute selector
table kOneBlobSel (read from the header) against the placement it must produce.

The claim under test: the two forms give the same FP16 features except where the five-knot form evaluates
a saturated knot (|u| just above 1) to one FP32 ulp short of 0 or 1; there its feature is the smallest
FP16 subnormal and the two-knot feature is the exact 0."""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "vknrc_amd", "csrc", "nrc_mfma.hpp")
F32, F64 = np.float32, np.float64


def fma32(a, b, c):
    """fmaf on float32 arrays, correctly rounded: the product is exact in float64; TwoSum gives the
    sum's float64 rounding error, which decides the float32 rounding where the float64 sum is a float32
    midpoint (the only case where rounding twice differs from rounding once)."""
    a, b, c = (np.asarray(x, F32) for x in (a, b, c))
    p = a.astype(F64) * b.astype(F64)
    cc = c.astype(F64)
    s = p + cc
    bb = s - p
    err = (p - (s - bb)) + (cc - bb)
    r = s.astype(F32)
    d = s - r.astype(F64)
    half = np.spacing(np.abs(r)).astype(F64) / 2
    away = (np.abs(d) == half) & (err != 0) & (np.sign(err) == np.sign(d))
    return np.where(away, np.nextafter(r, (s + err).astype(F32) * F32(2)), r)


def quartic(u):
    u = np.asarray(u, F32)
    u2 = (u * u).astype(F32)
    t = fma32(u2, F32(0.1875), F32(-0.625))
    t = fma32(u2, t, F32(0.9375))
    return np.clip(fma32(u, t, F32(0.5)), F32(0), F32(1))


def five_knot(x):
    """oneblob4_knots: the knots k - 4x, k = 0..4 (4x is exact, one rounding), differences in FP32."""
    x4 = (F32(4) * x).astype(F32)
    c = [quartic((F32(k) - x4).astype(F32)) for k in range(5)]
    return np.stack([(c[b + 1] - c[b]).astype(F32) for b in range(4)], -1)


def two_knot_halves(x):
    """oneblob4_packed without the placement: (j, [c0, c1 - c0, 1 - c1] as float16)."""
    x4 = (F32(4) * x).astype(F32)
    j = np.floor(x4)
    f = (x4 - j).astype(F32)
    c0, c1 = quartic(-f), quartic((F32(1) - f).astype(F32))
    tri = np.stack([c0, (c1 - c0).astype(F32), (F32(1) - c1).astype(F32)], -1).astype(np.float16)
    return j, tri


def selector_table():
    src = open(HEADER).read()
    body = re.search(r"kOneBlobSel\[8\]\[2\]\s*=\s*\{(.*?)\};", src, re.S).group(1)
    vals = [int(v, 16) for v in re.findall(r"0x([0-9A-Fa-f]+)u", body)]
    assert len(vals) == 16
    return np.array(vals, np.uint32).reshape(8, 2)


def perm(s0, s1, sel):
    """v_perm_b32: result byte i = byte sel_i of the 8 bytes (s1 bytes 0-3, s0 bytes 4-7); 12 -> 0x00."""
    src = np.concatenate([np.asarray(s1, np.uint32)[..., None].view(np.uint8),
                          np.asarray(s0, np.uint32)[..., None].view(np.uint8)], -1)
    out = np.zeros(np.shape(s0) + (4,), np.uint8)
    for i in range(4):
        b = (np.asarray(sel, np.uint32) >> (8 * i)) & 0xFF
        assert np.all((b <= 7) | (b == 12))
        out[..., i] = np.where(b == 12, 0, np.take_along_axis(src, np.minimum(b, 7)[..., None].astype(np.int64), -1)[..., 0])
    return out.view(np.uint32)[..., 0]


def two_knot(x):
    """oneblob4_packed: the three features placed by the two byte permutes of the selector table."""
    j, tri = two_knot_halves(x)
    bits = tri.view(np.uint16).astype(np.uint32)
    lo = bits[..., 0] | (bits[..., 1] << 16)
    hi = bits[..., 2]
    k = np.clip(np.where(np.isfinite(j), j, 0), -2, 5).astype(np.int64) + 2
    tab = selector_table()
    q = np.stack([perm(hi, lo, tab[k, 0]), perm(hi, lo, tab[k, 1])], -1)
    return q.view(np.uint16).reshape(len(x), 4)


def test_selector_table_places_the_three_features():
    tab = selector_table()
    lo = np.uint32(0x2222_1111)   # halves: A = 0x1111, B = 0x2222
    hi = np.uint32(0x0000_3333)   # D = 0x3333, then a zero half
    for j in range(-2, 6):
        q = np.array([perm(hi, lo, tab[j + 2, 0]), perm(hi, lo, tab[j + 2, 1])], np.uint32).view(np.uint16)
        want = np.zeros(4, np.uint16)
        for pos, v in zip((j - 1, j, j + 1), (0x1111, 0x2222, 0x3333)):
            if 0 <= pos < 4:
                want[pos] = v
        assert np.array_equal(q, want), (j, q, want)


def _compare(x):
    """Number of inputs whose quads differ; every difference must be the saturated-knot one."""
    a = five_knot(x).astype(np.float16).view(np.uint16)
    b = two_knot(x)
    diff = a != b
    # where they differ: the five-knot feature is the smallest subnormal (+-), the two-knot one exactly 0
    assert np.all((b[diff] == 0) & ((a[diff] & 0x7FFF) == 1)), x[diff.any(-1)][:5]
    return int(diff.any(-1).sum())


def test_two_knot_equals_five_knot_near_the_knots_and_on_a_sample():
    # every float within 2^17 ulps below and above each quarter point (where a knot saturates) ...
    pts = []
    for q in (0.25, 0.5, 0.75, 1.0):
        c = np.array(q, F32).view(np.uint32)
        pts.append((c + np.arange(-(1 << 17), (1 << 17) + 1, dtype=np.int64)).astype(np.uint32).view(F32))
    pts.append(np.arange(0, 1 << 17, dtype=np.uint32).view(F32))          # the smallest floats above 0
    near = np.concatenate(pts)
    near = near[(near >= 0) & (near <= 1)]
    n_near = _compare(near)
    # ... and a uniform sample of [0, 1]
    _compare(np.random.default_rng(11).uniform(0, 1, 2_000_000).astype(F32))
    # an exhaustive check over all 1.07e9 floats in [0, 1] (C, fmaf) finds 1,786 such inputs, all within
    # these neighbourhoods of the quarter points: the emulation here finds the same number
    assert n_near == 1786


def test_two_knot_outside_the_unit_interval_within_one_fp16_ulp():
    """Below 0 the two-knot form's f = 4x + 1 may round where the five-knot form's knot 0 - 4x is exact,
    so the features may differ there by one FP16 ulp (and by the saturated-knot subnormal); from 4x = 5
    on and below 4x = -1 the two-knot quad is the exact zero quad."""
    x = np.random.default_rng(12).uniform(-0.6, 1.6, 1_000_000).astype(F32)
    a = five_knot(x).astype(np.float16)
    b = two_knot(x).view(np.float16)
    ai, bi = a.view(np.uint16).astype(np.int64), b.view(np.uint16).astype(np.int64)
    same_sign = (ai >> 15) == (bi >> 15)
    assert np.all((np.abs(ai - bi) <= 1) & (same_sign | (a == 0) | (b == 0)))
    out = (x < -0.25) | (x >= 1.25)
    # there the two-knot quad is exactly zero; the five-knot one is zero up to the saturated-knot subnormal
    assert np.all(b[out].astype(np.float64) == 0)
    assert np.all((a[out].view(np.uint16) & 0x7FFF) <= 1)


def test_two_knot_exact_cases():
    x = np.array([0.0, 0.25, 0.5, 0.75, 1.0, 1.25, -0.25, 2.0, -1.0], F32)
    got = two_knot(x).view(np.float16).astype(np.float64)
    # at 4x = j the quad is (.., 0.5, 0.5, 0, ..) from position j-1 (C(0) = 1/2, C(1) - C(0) = 1/2, 1 - C(1) = 0)
    want = np.zeros((len(x), 4))
    for i, j in enumerate([0, 1, 2, 3, 4]):
        for pos, v in zip((j - 1, j), (0.5, 0.5)):
            if 0 <= pos < 4:
                want[i, pos] = v
    assert np.array_equal(got, want)
