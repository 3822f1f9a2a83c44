"""Pins the CPU oracle (oracle/nrc_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip in the oracle fails it: closed-form
values of every encoder primitive, numpy's own float16 rounding, torch float64
autograd fixtures (tests/golden) for the forward and the weight gradient, central
finite differences of the loss, batch invariants, and closed forms of the first
Adam/EMA steps.  Reference semantics: shader/src/NRCRecord.glsl, NN_nv.glsl,
nrc_optimize.comp, nrc_train_prepare.comp.
"""
import math

import numpy as np
import pytest

from oracle.oracle import (ENC_SINCOS6, ENC_TRI12, ENC_UV_SINCOS16, LOSS_L2, LOSS_REL_L2_LUM, Oracle,
                           OracleOptimizer, weight_count)


# ---------------------------------------------------------------- encoders
@pytest.mark.parametrize("x,want", [(0.0, 0.0), (0.5, 1.0), (1.0, 0.0), (1.5, -1.0), (-0.5, -1.0), (2.0, 0.0),
                                    (0.25, 0.5), (-0.25, -0.5), (3.5, -1.0), (-3.0, 0.0), (1024.5, 1.0)])
def test_tri_closed_form(oracle64, x, want):
    # _nrc_tri, NRCRecord.glsl:65-68; GLSL mod is floor-based so negatives fold correctly
    assert oracle64.tri(x) == want


def test_tri_is_odd_and_2_periodic(oracle64):
    for x in np.linspace(-3, 3, 61):
        assert abs(oracle64.tri(x) + oracle64.tri(-x)) < 1e-12
        assert abs(oracle64.tri(x + 2.0) - oracle64.tri(x)) < 1e-12


def test_quartic_cdf_closed_form(oracle64):
    o = oracle64
    assert o.quartic_cdf(0.0) == 0.5
    assert o.quartic_cdf(-0.25) == 0.0 and o.quartic_cdf(-1.0) == 0.0
    assert o.quartic_cdf(0.25) == 1.0 and o.quartic_cdf(3.0) == 1.0
    # u = 1/2: 15/32 * (1 - 1/6 + 1/80) + 1/2 = 3045/7680 + 1/2
    assert abs(o.quartic_cdf(0.125) - (0.5 + 3045 / 7680)) < 1e-15
    for d in np.linspace(-0.3, 0.3, 31):
        assert abs(o.quartic_cdf(d) + o.quartic_cdf(-d) - 1.0) < 1e-15


def test_oneblob_bins(oracle64):
    o = oracle64
    # bins telescope: sum = C(1 - x) - C(-x) == 1 on [1/4, 3/4]
    for x in np.linspace(0.25, 0.75, 21):
        assert abs(o.oneblob4(x).sum() - 1.0) < 1e-14
    assert abs(o.oneblob4(0.0).sum() - 0.5) < 1e-15
    b = o.oneblob4(0.125)
    assert abs(b[0] - 2 * 3045 / 7680) < 1e-15
    assert b[2] == 0.0 and b[3] == 0.0
    assert np.allclose(o.oneblob4(0.5), [0, 0.5, 0.5, 0])


@pytest.mark.parametrize("d,want", [((1, 0, 0), (0.5, 0.5)), ((0, 1, 0), (0.75, 0.5)), ((-1, 0, 0), (1.0, 0.5)),
                                    ((0, -1, 0), (0.25, 0.5)), ((0, 0, 1), (0.5, 0.0)), ((0, 0, -1), (0.5, 1.0)),
                                    ((0, 0, 2), (0.5, 0.0))])
def test_sph_encode_axes(oracle64, d, want):
    # NRCSphEncode, NRCRecord.glsl:47-49 (xy == 0 -> 0.5; z clamped to [-1, 1])
    assert np.allclose(oracle64.sph_encode(np.array(d, float)), want, atol=1e-15)


def _rec(**kw):
    r = np.zeros((1, 14), np.float32)
    r[0, 0:3] = kw.get("pos", (0, 0, 0))
    r[0, 3:5] = kw.get("dir", (0.5, 0.5))
    r[0, 5:7] = kw.get("nrm", (0.5, 0.5))
    r[0, 7] = kw.get("rough", 0.0)
    r[0, 8:11] = kw.get("diff", (0, 0, 0))
    r[0, 11:14] = kw.get("spec", (0, 0, 0))
    return r


def test_tri12_feature_index_map(oracle64):
    # NRCInputEncode, NRCRecord.glsl:78-95: 3 x 12 tri waves, 5 one-blobs, albedos, (1, 1)
    r = _rec(pos=(0.5, 0.25, -0.5), dir=(0.125, 0.875), nrm=(0.5, 0.0), rough=math.log(2.0),
             diff=(0.1, 0.2, 0.3), spec=(0.4, 0.5, 0.6))
    f = oracle64.encode(r, ENC_TRI12)[0]
    want_x = [1.0] + [0.0] * 11             # tri(0.5)=1, tri(integer)=0
    want_y = [0.5, 1.0] + [0.0] * 10        # tri(0.25)=0.5, tri(0.5)=1
    want_z = [-1.0] + [0.0] * 11
    assert np.allclose(f[0:12], want_x, atol=1e-15)
    assert np.allclose(f[12:24], want_y, atol=1e-15)
    assert np.allclose(f[24:36], want_z, atol=1e-15)
    assert np.allclose(f[36:40], oracle64.oneblob4(0.125))
    assert np.allclose(f[40:44], oracle64.oneblob4(0.875))
    assert np.allclose(f[44:48], [0, 0.5, 0.5, 0])
    assert np.allclose(f[48:52], [0.5, 0, 0, 0])
    assert np.allclose(f[52:56], [0, 0.5, 0.5, 0])       # 1 - exp(-ln 2) = 0.5
    assert np.allclose(f[56:62], np.float32([0.1, 0.2, 0.3, 0.4, 0.5, 0.6]))
    assert f[62] == 1.0 and f[63] == 1.0


def test_sincos6_feature_index_map(oracle64):
    r = _rec(pos=(0.5, 1.0 / 6.0, 0.0))
    f = oracle64.encode(r, ENC_SINCOS6)[0]
    px = np.float32(0.5)
    py = float(np.float32(1.0 / 6.0))
    assert np.allclose(f[0:6], [1, 0, 0, 0, 0, 0], atol=1e-15)          # sin(2^k pi / 2)
    assert np.allclose(f[6:12], [0, -1, 1, 1, 1, 1], atol=1e-15)        # cos(2^k pi / 2)
    assert np.allclose(f[12:18], [math.sin(math.pi * py * 2 ** k) for k in range(6)], atol=1e-14)
    assert abs(f[12] - 0.5) < 1e-7
    assert np.allclose(f[24:30], 0, atol=1e-15) and np.allclose(f[30:36], 1, atol=1e-15)
    # sin^2 + cos^2 = 1 for every octave
    rng = np.random.default_rng(1)
    rr = np.zeros((50, 14), np.float32)
    rr[:, 0:3] = rng.uniform(-4, 4, (50, 3))
    ff = oracle64.encode(rr, ENC_SINCOS6)
    for a in range(3):
        s, c = ff[:, 12 * a:12 * a + 6], ff[:, 12 * a + 6:12 * a + 12]
        assert np.allclose(s * s + c * c, 1.0, atol=1e-13)
    assert px == 0.5


def test_uv_sincos16_closed_form(oracle64):
    r = _rec(pos=(0.25, 0.5, 0.0))
    f = oracle64.encode(r, ENC_UV_SINCOS16)[0]
    h = math.sqrt(0.5)
    assert np.allclose(f[0:4], [h, 1, 0, 0], atol=1e-15)
    assert np.allclose(f[16:20], [h, 0, -1, 1], atol=1e-15)
    assert np.allclose(f[32:35], [1, 0, 0], atol=1e-15)
    assert np.allclose(f[48:51], [0, -1, 1], atol=1e-15)


@pytest.mark.parametrize("enc", [ENC_TRI12, ENC_SINCOS6, ENC_UV_SINCOS16])
def test_encode_matches_torch_fixture(oracle64, golden, enc):
    f = oracle64.encode(golden["enc_records"][:, :14], enc)
    assert np.allclose(f, golden[f"enc{enc}_features"], rtol=0, atol=1e-12)


def test_round_f16_matches_numpy(oracle64):
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-5, 500), rng.normal(0, 3e4, 500),
                         [0.0, -0.0, 65504.0, 65519.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 1 + 2 ** -11]])
    for x in xs:
        with np.errstate(over="ignore"):
            want = float(np.float16(np.float64(x)))
        got = oracle64.round_f16(x)
        assert got == want or (math.isinf(got) and math.isinf(want)), (x, got, want)


# ---------------------------------------------------------------- forward
@pytest.mark.parametrize("hidden", [5, 4])
def test_forward_matches_torch_fixture(oracle64, golden, hidden):
    w = golden[f"h{hidden}_weights"]
    out = oracle64.infer(w, hidden, golden[f"h{hidden}_records"], ENC_TRI12, clamp0=False)
    want = golden[f"h{hidden}_out_raw"]
    assert np.allclose(out, want, rtol=1e-12, atol=1e-12)
    clamped = oracle64.infer(w, hidden, golden[f"h{hidden}_records"], ENC_TRI12, clamp0=True)
    assert np.array_equal(clamped, np.where(out > 0, out, 0.0))   # nrc_inference.comp:48


def test_forward_numpy_matmul_and_invariants(oracle64):
    rng = np.random.default_rng(3)
    w = rng.normal(0, 0.2, weight_count(5))
    x = rng.uniform(-1, 1, (77, 64))
    out = oracle64.infer_encoded(w, 5, x)
    a = x.T
    for l in range(5):
        a = np.maximum(w[l * 4096:(l + 1) * 4096].reshape(64, 64) @ a, 0)
    want = (w[5 * 4096:].reshape(3, 64) @ a).T
    assert np.allclose(out, want, rtol=1e-12, atol=1e-13)
    # no bias: positive homogeneity and zero input -> zero output
    assert np.allclose(oracle64.infer_encoded(w, 5, 3.5 * x), 3.5 * out, rtol=1e-12, atol=1e-13)
    assert np.all(oracle64.infer_encoded(w, 5, np.zeros((5, 64))) == 0)


def test_harness_fixture(oracle64, golden):
    # distribution of the reference harness, test/main.cpp:91-101,150-162
    out = oracle64.infer_encoded(golden["harness_weights"], 5, golden["harness_inputs"])
    assert np.allclose(out, golden["harness_out"], rtol=1e-12, atol=1e-14)
    dw, _ = oracle64.gradient_encoded(golden["harness_weights"], 5, golden["harness_inputs"],
                                      golden["harness_targets"], LOSS_L2)
    assert np.allclose(dw, golden["harness_dw"], rtol=1e-10, atol=1e-14)


# ---------------------------------------------------------------- loss + gradient
def test_loss_grad_closed_form(oracle64):
    o = np.array([1.0, 0.5, -2.0])
    t = np.array([0.5, 0.5, 1.0])
    lum = 0.299 * 1.0 + 0.587 * 0.5 + 0.0
    d = lum * lum + 0.01
    g, v = oracle64.loss_grad(LOSS_REL_L2_LUM, o, t, 4.0)
    assert np.allclose(g, 2 * 4.0 * (o - t) / d, rtol=1e-15)
    assert abs(v - ((o - t) ** 2).sum() / d) < 1e-12
    g, v = oracle64.loss_grad(LOSS_L2, o, t, 1.0)
    assert np.allclose(g, 2 * (o - t)) and abs(v - 9.25) < 1e-15
    # all-negative prediction: denominator is exactly 0.01
    g, _ = oracle64.loss_grad(LOSS_REL_L2_LUM, -o * (o > 0) - 1, t, 1.0)
    assert np.allclose(g, 2 * ((-o * (o > 0) - 1) - t) / 0.01)


@pytest.mark.parametrize("hidden", [5, 4])
@pytest.mark.parametrize("loss", [LOSS_REL_L2_LUM, LOSS_L2])
def test_gradient_matches_torch_autograd(oracle64, golden, hidden, loss):
    dw, ls = oracle64.gradient(golden[f"h{hidden}_weights"], hidden, golden[f"h{hidden}_records"], ENC_TRI12, loss)
    want = golden[f"h{hidden}_loss{loss}_dw"]
    assert np.allclose(dw, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())
    assert abs(ls - golden[f"h{hidden}_loss{loss}_sum"][0]) < 1e-9 * abs(ls)


def test_gradient_finite_differences(oracle64):
    rng = np.random.default_rng(11)
    hidden = 5
    w = rng.normal(0, 0.25, weight_count(hidden))
    x = rng.uniform(0, 1, (4, 64))
    t = rng.uniform(0, 1, (4, 3))
    o0 = oracle64.infer_encoded(w, hidden, x)
    lum = (np.maximum(o0, 0) * [0.299, 0.587, 0.114]).sum(1, keepdims=True)
    d0 = lum * lum + 0.01

    def L(wv):  # relative L2 with the denominator frozen at w (stop-gradient)
        o = oracle64.infer_encoded(wv, hidden, x)
        return (((o - t) ** 2) / d0).sum()

    dw, _ = oracle64.gradient_encoded(w, hidden, x, t, LOSS_REL_L2_LUM)
    idx = np.concatenate([rng.choice(weight_count(hidden) - 192, 40, replace=False), 20480 + np.arange(0, 192, 7)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (L(wp) - L(wm)) / (2 * h)
        assert abs(fd - dw[i]) < 1e-6 * max(1.0, abs(dw[i])), (i, fd, dw[i])


def test_gradient_batch_invariants(oracle64, golden):
    w = golden["h5_weights"]
    rec = golden["h5_records"]
    dw, ls = oracle64.gradient(w, 5, rec, ENC_TRI12, LOSS_REL_L2_LUM)
    dw2, ls2 = oracle64.gradient(w, 5, np.concatenate([rec, rec]), ENC_TRI12, LOSS_REL_L2_LUM)
    assert np.allclose(dw2, 2 * dw, rtol=1e-12, atol=1e-12)
    perm = np.random.default_rng(5).permutation(len(rec))
    dwp, _ = oracle64.gradient(w, 5, rec[perm], ENC_TRI12, LOSS_REL_L2_LUM)
    assert np.allclose(dwp, dw, rtol=1e-12, atol=1e-12)
    # zero-padded samples (nrc_gradient.comp:27-34): zero input and zero target contribute nothing
    feats = oracle64.encode(rec[:, :14], ENC_TRI12)
    a, _ = oracle64.gradient_encoded(w, 5, feats, rec[:, 14:17].astype(np.float64), LOSS_REL_L2_LUM)
    pad_f = np.concatenate([feats, np.zeros((9, 64))])
    pad_t = np.concatenate([rec[:, 14:17].astype(np.float64), np.zeros((9, 3))])
    b, _ = oracle64.gradient_encoded(w, 5, pad_f, pad_t, LOSS_REL_L2_LUM)
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    # loss scale is linear in the gradient
    c, _ = oracle64.gradient(w, 5, rec, ENC_TRI12, LOSS_REL_L2_LUM, scale=128.0)
    assert np.allclose(c, 128.0 * dw, rtol=1e-13)


# ---------------------------------------------------------------- optimizer
def test_adam_matches_torch_fixture(oracle64, golden):
    opt = OracleOptimizer(oracle64, golden["h5_weights"])
    G = golden["h5_loss0_dw"]
    for g, c in [(G, 40), (0.5 * G, 40), (-G, 20)]:
        assert opt.step(g, c)
    assert np.allclose(opt.w, golden["adam_w"], rtol=1e-13, atol=1e-15)
    assert np.allclose(opt.ema, golden["adam_ema"], rtol=1e-12, atol=1e-15)
    assert np.allclose(opt.m, golden["adam_m"], rtol=1e-13, atol=0)
    assert np.allclose(opt.v, golden["adam_v"], rtol=1e-13, atol=0)
    s = opt.state
    assert np.allclose([s.beta1_t, s.beta2_t, s.alpha_t, s.alpha_t_1], golden["adam_scalars"], rtol=1e-15)
    assert s.t == 3


def test_adam_first_step_closed_form(oracle64):
    rng = np.random.default_rng(2)
    w0 = rng.normal(0, 1, 500)
    G = rng.normal(0, 1, 500) * 10 ** rng.uniform(-6, 3, 500)
    opt = OracleOptimizer(oracle64, w0)
    opt.step(G, 7)
    g = G / 7
    assert np.allclose(opt.w, w0 - 0.002 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    # first step: eta = 1 - 0.99, eta_1 = 0, so the EMA equals the new weight exactly
    assert np.allclose(opt.ema, opt.w, rtol=1e-14, atol=0)


def test_adam_count_zero_and_nonfinite(oracle64):
    w0 = np.linspace(-1, 1, 64)
    opt = OracleOptimizer(oracle64, w0)
    assert not opt.step(np.ones(64), 0)          # nrc_optimize.comp:33-34 early-out
    assert opt.state.t == 0 and np.array_equal(opt.w, w0)
    opt.step(np.ones(64), 2)
    m1 = opt.m.copy()
    G = np.ones(64)
    G[3], G[9] = np.nan, np.inf                   # nrc_optimize.comp:37-38: NaN/Inf -> 0
    opt.step(G, 2)
    assert np.isclose(opt.m[3], 0.9 * m1[3]) and np.isclose(opt.m[9], 0.9 * m1[9])
    assert np.isclose(opt.m[0], 0.9 * m1[0] + 0.1 * 0.5)
    assert np.all(np.isfinite(opt.w))


def test_prepare_scalar_recurrences(oracle64):
    opt = OracleOptimizer(oracle64, np.zeros(8))
    for k in range(1, 60):
        opt.step(np.ones(8), 1)
        s = opt.state
        assert s.t == k
        assert abs(s.beta1_t - 0.9 ** k) < 1e-13 and abs(s.beta2_t - 0.999 ** k) < 1e-13
        assert abs(s.alpha_t - 0.99 ** k) < 1e-13 and abs(s.alpha_t_1 - 0.99 ** (k - 1)) < 1e-13


def test_dynamic_loss_scale_policy(oracle64):
    w0 = np.linspace(-1, 1, 16)
    opt = OracleOptimizer(oracle64, w0, loss_scale=128.0, dynamic_scale=True, growth_interval=3)
    G = np.ones(16) * 128.0
    G[5] = np.inf
    assert not opt.step(G, 4)
    assert opt.state.loss_scale == 64.0 and opt.state.t == 0 and np.array_equal(opt.w, w0)
    for _ in range(3):
        assert opt.step(np.ones(16) * 64.0, 4)
    assert opt.state.loss_scale == 128.0 and opt.state.t == 3
    # the scale divides out: g = G / count / s
    assert np.allclose(opt.m, 0.1 * 0.25 * (1 + 0.9 + 0.81))


def test_adam_constant_gradient_moves_by_lr_every_step(oracle64):
    """A gradient g held constant makes m_t = (1 - b1^t) g and v_t = (1 - b2^t) g^2, so the
    bias-corrected moments are exactly g and g^2 and every step t = 1..20 moves each weight by
    -lr g / (|g| + eps).  A missing or misplaced bias correction (m / (1 - b1) for instance, or a
    correction with b^(t-1)) changes the step size from t = 2 on."""
    rng = np.random.default_rng(41)
    w0 = rng.normal(0, 1, 400)
    g = rng.normal(0, 1, 400) * 10 ** rng.uniform(-7, 2, 400)
    count, scale = 9, 16.0
    opt = OracleOptimizer(oracle64, w0, loss_scale=scale)
    step = 0.002 * g / (np.abs(g) + 1e-8)
    for t in range(1, 21):
        assert opt.step(g * count * scale, count)     # G = count * s * g
        assert np.allclose(opt.m, (1 - 0.9 ** t) * g, rtol=1e-12, atol=0)
        assert np.allclose(opt.v, (1 - 0.999 ** t) * g * g, rtol=1e-11, atol=0)
        assert np.allclose(opt.w, w0 - t * step, rtol=1e-11, atol=1e-14), t


def test_adam_sign_flip_closed_form(oracle64):
    """g at step 1, -g at step 2: m_2 = 0.9 * 0.1 g - 0.1 g = -0.01 g, v_2 = 0.001 g^2 (0.999 + 1),
    so the second step is -lr (-0.01 g / 0.19) / (|g| sqrt(1.999e-3 / 1.999e-3) + eps)."""
    g = np.array([3.0, -0.5, 1e-3, 7e-9])
    opt = OracleOptimizer(oracle64, np.zeros(4))
    opt.step(g, 1)
    w1 = opt.w.copy()
    opt.step(-g, 1)
    m2 = -0.01 * g
    v2 = 0.001 * g * g * (0.999 + 1.0)
    assert np.allclose(opt.m, m2, rtol=1e-13) and np.allclose(opt.v, v2, rtol=1e-13)
    mh = m2 / (1 - 0.9 ** 2)
    vh = v2 / (1 - 0.999 ** 2)                         # = g^2 exactly
    assert np.allclose(vh, g * g, rtol=1e-12)
    assert np.allclose(opt.w, w1 - 0.002 * mh / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)


def test_ema_as_written_with_frozen_weights(oracle64):
    """lr = 0 keeps w fixed, so ema_t = c_t w with c_1 = 1 and, from nrc_optimize.comp:46-47 as
    written (the second term is not divided by 1 - alpha^t),
        c_2 = (1 - a) / (1 - a^2) + a (1 - a) = 1 / (1 + a) + a (1 - a),
    which for a = 0.99 is 0.512412...; a bias-corrected EMA would keep c_t = 1.  As alpha^t -> 0
    the recurrence becomes the plain EMA and c_t -> 1.  With a = 0 the EMA is the weight."""
    w = np.array([1.0, -2.5, 0.125, 3e-4])
    opt = OracleOptimizer(oracle64, w, lr=1e-300)       # a step of 1e-300 * O(1) leaves w as it is
    opt.step(np.ones(4), 1)
    assert np.array_equal(opt.w, w) and np.allclose(opt.ema, w, rtol=1e-15)       # c_1 = 1
    opt.step(np.ones(4), 1)
    a = 0.99
    c2 = 1.0 / (1.0 + a) + a * (1.0 - a)
    assert abs(c2 - 0.5124125628140703) < 1e-15
    assert np.allclose(opt.ema, c2 * w, rtol=1e-12)
    for _ in range(4000):                                # 0.99^4000 ~ 3.4e-18
        opt.step(np.ones(4), 1)
    assert np.allclose(opt.ema, w, rtol=1e-9)
    z = OracleOptimizer(oracle64, w, lr=1e-300, ema_alpha=0.0)
    z.step(np.ones(4), 1)
    z.step(np.ones(4), 1)
    assert np.array_equal(z.ema, w)
