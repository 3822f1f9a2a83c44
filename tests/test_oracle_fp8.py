"""Pins the FP8 (e4m3) oracle: the rounding against the format's definition and
a brute-force nearest-value search, the forward against the FP64 forward where
e4m3 is exact."""
import numpy as np

from oracle.fp8 import E4M3_MAX, e4m3_round, e4m3_values, infer_fp8
from oracle.oracle import Oracle, weight_count


def test_e4m3_value_table():
    v = e4m3_values()
    assert v.size == 254
    pos = np.unique(v[v > 0])
    assert pos.size == 126
    assert pos.max() == 448.0 and pos.min() == 2.0 ** -9      # largest normal, smallest subnormal
    assert 2.0 ** -6 in pos and 1.0 in pos and 1.125 in pos and 240.0 in pos


def test_e4m3_round_known_answers():
    cases = {
        0.0: 0.0, 1.0: 1.0, 448.0: 448.0, 449.0: 448.0, 464.0: 448.0, 1e30: 448.0, -1e30: -448.0,
        2.0 ** -9: 2.0 ** -9, 2.0 ** -10: 0.0,             # tie between 0 and 2^-9 -> even (0)
        1.5 * 2.0 ** -9: 2.0 ** -8,                        # tie between 1 and 2 steps -> even (2)
        1.0 + 1.0 / 16: 1.0, 1.1875: 1.25,                 # ties in [1, 2): step 1/8, to even mantissa
        1.0 + 1.0 / 16 + 1e-9: 1.125, 300.0: 288.0, 304.0: 320.0, 416.0: 416.0, 432.0: 448.0,
        -3.3: -3.25, 0.1: 0.1015625,
    }
    for x, want in cases.items():
        assert e4m3_round(x) == want, (x, e4m3_round(x), want)
    assert np.signbit(e4m3_round(-0.0))


def test_e4m3_round_is_the_nearest_representable_value():
    rng = np.random.default_rng(3)
    table = np.unique(e4m3_values())
    x = np.concatenate([rng.uniform(-500, 500, 4000), rng.normal(0, 1, 4000), rng.normal(0, 0.01, 2000)])
    x = np.clip(x, -E4M3_MAX, E4M3_MAX)
    q = e4m3_round(x)
    d = np.abs(table[None, :] - x[:, None])
    best = table[np.argmin(d, axis=1)]
    # where the nearest value is unique, it must be chosen (ties are covered by the known answers)
    dmin = d.min(axis=1)
    unique = (np.sort(d, axis=1)[:, 1] > dmin)
    assert np.array_equal(q[unique], best[unique])
    assert np.all(np.abs(q - x) <= dmin + 0.0)
    assert np.array_equal(e4m3_round(table), table)


def _signed_permutation_weights(rng, hidden):
    """Each layer maps input col p(row) into output row with weight +-2^k: every activation stays a
    single e4m3 product, so e4m3 rounding is the identity end to end."""
    w = np.zeros(weight_count(hidden))
    for l in range(hidden):
        W = np.zeros((64, 64))
        p = rng.permutation(64)
        W[np.arange(64), p] = rng.choice([0.5, 1.0, 2.0], 64) * rng.choice([-1.0, 1.0], 64)
        w[l * 4096:(l + 1) * 4096] = W.ravel()
    Wo = np.zeros((3, 64))
    Wo[np.arange(3), rng.choice(64, 3, replace=False)] = [1.0, 0.5, 2.0]
    w[hidden * 4096:] = Wo.ravel()
    return w


def test_infer_fp8_equals_fp64_where_e4m3_is_exact():
    rng = np.random.default_rng(5)
    o = Oracle("f64")
    table = np.unique(e4m3_values())
    small = table[(np.abs(table) >= 2.0 ** -3) & (np.abs(table) <= 1.0)]
    for hidden in (4, 5):
        w = _signed_permutation_weights(rng, hidden)
        f = rng.choice(small, (300, 64))
        ref = o.infer_encoded(w, hidden, f)
        # scales: C_{l+1}/C_l in {2, 4, 1/2}; out_scale undoes c_H C_H
        C = [1.0]
        for l in range(hidden):
            C.append(C[-1] * [2.0, 4.0, 0.5][l % 3])
        cH = 4.0
        y = infer_fp8(f, w, hidden, C, 1.0 / (cH * C[-1]), clamp0=True)
        assert np.array_equal(y, np.maximum(ref, 0.0))
        # a transposed weight layout breaks it
        wt = w.copy()
        for l in range(hidden):
            wt[l * 4096:(l + 1) * 4096] = w[l * 4096:(l + 1) * 4096].reshape(64, 64).T.ravel()
        assert not np.array_equal(infer_fp8(f, wt, hidden, C, 1.0 / (cH * C[-1])), y)
        # scales that push activations past 448 saturate them
        Cbig = [1.0] + [c * 2.0 ** 9 for c in C[1:]]
        ybig = infer_fp8(f, w, hidden, Cbig, 1.0 / (cH * Cbig[-1]))
        assert not np.allclose(ybig, y) and np.all(ybig <= 448.0 / Cbig[-1] * 2.0 + 1e-12)


def test_infer_fp8_error_on_a_random_network_is_e4m3_sized():
    rng = np.random.default_rng(6)
    o = Oracle("f64")
    w = rng.normal(0.0, np.sqrt(2.0 / 64.0), weight_count(4))
    f = rng.uniform(-1, 1, (2000, 64))
    ref = o.infer_encoded(w, 4, f)
    # scales from the exact activation maxima, 224 target as the kernel
    x, C = f, [1.0]
    for l in range(4):
        z = x @ w[l * 4096:(l + 1) * 4096].reshape(64, 64).T
        C.append(2.0 ** np.floor(np.log2(224.0 / z.max())))
        x = np.maximum(z, 0)
    cH = 2.0 ** np.floor(np.log2(448.0 / np.abs(w[4 * 4096:]).max()))
    y = infer_fp8(f, w, 4, C, 1.0 / (cH * C[-1]), clamp0=False)
    e = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert 0.005 < e < 0.3, e   # e4m3 keeps 3 mantissa bits: errors of a few percent per layer
