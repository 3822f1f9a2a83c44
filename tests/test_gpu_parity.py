"""GPU parity of the HIP hot path against the CPU oracle, through the C ABI.

Bars: the plain FP32 path within FP32 rounding of the FP64 oracle; the fused
FP16-MFMA path within forward relative error 1e-2 and gradient cosine 0.999 of
the FP64 oracle run on the same FP16-rounded weights and features.  Reference
semantics: shader/src/NRCRecord.glsl, NN_nv.glsl, nrc_inference.comp,
nrc_gradient.comp, nrc_optimize.comp, nrc_train_prepare.comp.
"""
import numpy as np
import pytest

import vknrc_amd as v
from oracle.oracle import Oracle, OracleOptimizer
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"
FWD_REL_TOL = 1e-2       # fused forward relative error vs the FP64 oracle
GRAD_COSINE_MIN = 0.999  # fused weight-gradient cosine vs the FP64 oracle


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def f16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float64)


def rel_err(y, ref):
    return float(np.linalg.norm(np.asarray(y, np.float64) - ref) / max(np.linalg.norm(ref), 1e-300))


def cosine(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def make_cache(rng, **kw):
    cfg = v.NRCConfig(**kw)
    st = v.NRCState(cfg)
    w = rng.normal(0.0, np.sqrt(2.0 / 64.0), v.weight_count(cfg.hidden_layers)).astype(np.float32)
    st.set_state(weight=w)
    return st, w


def queries_for(enc, n, rng, spread=1.0):
    q = wl.make_queries(n, rng)
    q[:, 0:3] *= spread
    if enc == v.ENCODING_UV_SINCOS16:
        q[:, 0:2] = rng.uniform(0, 1, (n, 2))
    return q


# ------------------------------------------------------------------ encoders
@pytest.mark.parametrize("enc", [v.ENCODING_TRI12, v.ENCODING_SINCOS6, v.ENCODING_UV_SINCOS16])
def test_encode_matches_oracle(o64, enc):
    rng = np.random.default_rng(100 + enc)
    n = 5003
    q = queries_for(enc, n, rng, spread=1.7)
    out = torch.zeros((n, 64), dtype=torch.float32, device=DEV)
    v.encode(dev(q), n, out, enc)
    got = host(out)
    # the FP32 build of the oracle runs the reference's fp32 formula step for step
    err32 = np.abs(got - Oracle("f32").encode(q, enc)).max()
    assert err32 < 5e-6, err32
    # against FP64 the tri waves carry the fp32 rounding of x - 0.5 (x = 2^k p), at most
    # 2 ulp(|x| + 0.5) after the slope-2 fold; every other feature is within 5e-6
    err = np.abs(got - o64.encode(q, enc))
    bound = np.full(64, 5e-6)[None, :].repeat(n, 0)
    if enc == v.ENCODING_TRI12:
        x = np.abs(q[:, 0:3, None].astype(np.float64)) * (2.0 ** np.arange(12))[None, None, :]
        bound[:, :36] = 2 * np.spacing((x + 0.5).astype(np.float32)).reshape(n, 36) + 5e-6
    assert np.all(err <= bound), float((err - bound).max())


def test_sph_encode_matches_oracle(o64):
    rng = np.random.default_rng(9)
    d = np.concatenate([rng.normal(size=(997, 3)), np.eye(3), -np.eye(3)]).astype(np.float32)
    d[:997] /= np.linalg.norm(d[:997], axis=1, keepdims=True)
    uv = torch.zeros((d.shape[0], 2), dtype=torch.float32, device=DEV)
    v.sph_encode(dev(d), d.shape[0], uv)
    ref = np.array([o64.sph_encode(x.astype(np.float64)) for x in d])
    assert np.abs(host(uv) - ref).max() < 2e-6


# ------------------------------------------------------------------ inference
@pytest.mark.parametrize("hidden", [5, 4])
@pytest.mark.parametrize("enc", [v.ENCODING_TRI12, v.ENCODING_SINCOS6, v.ENCODING_UV_SINCOS16])
def test_plain_infer_matches_oracle(o64, hidden, enc):
    rng = np.random.default_rng(10 * hidden + enc)
    st, w = make_cache(rng, hidden_layers=hidden, encoding=enc)
    n = 3001
    q = queries_for(enc, n, rng)
    out = torch.full((n, 3), -7.0, dtype=torch.float32, device=DEV)
    st.infer(dev(q), out, path=v.PATH_PLAIN_F32)
    ref = o64.infer(w.astype(np.float64), hidden, q, enc, round_f16=False, clamp0=True)
    y = host(out)
    assert rel_err(y, ref) < 1e-5
    assert np.abs(y - ref).max() < 1e-4 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("hidden", [5, 4])
@pytest.mark.parametrize("enc", [v.ENCODING_TRI12, v.ENCODING_SINCOS6, v.ENCODING_UV_SINCOS16])
def test_fused_infer_matches_oracle(o64, hidden, enc):
    rng = np.random.default_rng(20 * hidden + enc)
    st, w = make_cache(rng, hidden_layers=hidden, encoding=enc)
    n = 70001  # several waves per block, ragged last tile
    q = queries_for(enc, n, rng)
    out = torch.full((n, 3), -7.0, dtype=torch.float32, device=DEV)
    st.infer(dev(q), out, path=v.PATH_FUSED_F16)
    y = host(out)
    ref16 = o64.infer(f16(w), hidden, q, enc, round_f16=True, clamp0=True)
    ref = o64.infer(w.astype(np.float64), hidden, q, enc, round_f16=False, clamp0=True)
    assert rel_err(y, ref16) < FWD_REL_TOL, rel_err(y, ref16)
    assert rel_err(y, ref) < FWD_REL_TOL, rel_err(y, ref)
    assert np.all(y >= 0)


def test_infer_respects_device_count():
    rng = np.random.default_rng(1)
    st, _ = make_cache(rng)
    q = dev(wl.make_queries(1000, rng))
    for path in (v.PATH_PLAIN_F32, v.PATH_FUSED_F16):
        full = torch.full((1000, 3), -7.0, dtype=torch.float32, device=DEV)
        st.infer(q, full, path=path)
        for cnt, want in ((100, 100), (0, 0), (5000, 1000), (1000, 1000), (65, 65)):
            out = torch.full((1000, 3), -7.0, dtype=torch.float32, device=DEV)
            c = torch.tensor([cnt], dtype=torch.int32, device=DEV)
            st.infer(q, out, count=c, path=path)
            y = host(out)
            assert np.all(y[want:] == -7.0)
            assert np.array_equal(y[:want], host(full)[:want])


def test_fused_infer_is_per_record(o64):
    """Each record's result is independent of its position in the batch (bit-exact)."""
    rng = np.random.default_rng(2)
    st, _ = make_cache(rng)
    n = 9000
    q = wl.make_queries(n, rng)
    perm = rng.permutation(n)
    a = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    b = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    st.infer(dev(q), a)
    st.infer(dev(q[perm]), b)
    assert np.array_equal(host(a)[perm], host(b))


def test_mlp_f16_entry_points_on_the_harness_distribution(o64, golden):
    """test/main.cpp:150-162 inputs through the test-kernel interface (pre-encoded fp16)."""
    for wdist in ("kaiming", "harness"):
        rng = np.random.default_rng(3)
        if wdist == "harness":
            w = golden["harness_weights"].astype(np.float32)
            x = golden["harness_inputs"].astype(np.float32)
            t = golden["harness_targets"].astype(np.float32)
        else:
            w = rng.normal(0, np.sqrt(2 / 64), v.weight_count(5)).astype(np.float32)
            x = rng.uniform(0, 1, (4000, 64)).astype(np.float32)
            t = rng.uniform(0, 1, (4000, 3)).astype(np.float32)
        n = x.shape[0]
        w16 = dev(w.astype(np.float16).view(np.uint16).view(np.int16))
        x16 = dev(x.astype(np.float16).view(np.uint16).view(np.int16))
        out = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
        v.mlp_forward_f16(w16, x16, n, out)
        ref = o64.infer_encoded(f16(w), 5, f16(x))
        tol = FWD_REL_TOL if wdist == "kaiming" else 5e-2  # harness outputs sit in the fp16 subnormal range
        assert rel_err(host(out), ref) < tol
        dw = torch.zeros(v.weight_count(5), dtype=torch.float32, device=DEV)
        v.mlp_gradient_f16(w16, x16, dev(t), n, dw)
        G, _ = o64.gradient_encoded(f16(w), 5, f16(x), t.astype(np.float64), loss=1)
        assert cosine(host(dw), G) > (GRAD_COSINE_MIN if wdist == "kaiming" else 0.99)


# ------------------------------------------------------------------ training
def train_records(n, seed, enc=v.ENCODING_TRI12):
    rng = np.random.default_rng(seed)
    lights = wl.GaussianLights(seed + 1)
    r = wl.make_train_records(n, rng, lights)
    if enc == v.ENCODING_UV_SINCOS16:
        r[:, 0:2] = rng.uniform(0, 1, (n, 2))
    return r


def one_step(o64, path, hidden, enc, loss, n, seed, chunks=0):
    rng = np.random.default_rng(seed)
    st, w = make_cache(rng, hidden_layers=hidden, encoding=enc, loss=loss)
    st.set_grad_chunks(chunks)
    r = train_records(n, seed, enc)
    st.train(dev(r), path=path)
    s = st.get_state()
    g_gpu = s["m"].astype(np.float64) / (1.0 - np.float32(0.9))   # m = (1 - beta1) g after one step
    is16 = path == v.PATH_FUSED_F16
    G, lsum = o64.gradient(f16(w) if is16 else w.astype(np.float64), hidden, r, enc, loss, round_f16=is16)
    return st, w, r, s, g_gpu, G / n, lsum / n


@pytest.mark.parametrize("hidden,enc,loss", [(5, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (5, v.ENCODING_SINCOS6, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (4, v.ENCODING_UV_SINCOS16, v.LOSS_L2),
                                             (5, v.ENCODING_TRI12, v.LOSS_L2)])
def test_plain_train_step_matches_oracle(o64, hidden, enc, loss):
    st, w, r, s, g, g_ref, l_ref = one_step(o64, v.PATH_PLAIN_F32, hidden, enc, loss, 3000, 31 + hidden + enc)
    assert cosine(g, g_ref) > 0.99999
    assert rel_err(g, g_ref) < 1e-4
    assert abs(s["state"]["last_loss"] - l_ref) < 1e-4 * abs(l_ref)
    assert s["state"]["t"] == 1 and s["state"]["last_count"] == 3000
    assert abs(s["state"]["beta1_t"] - 0.9) < 1e-7 and abs(s["state"]["alpha_t_1"] - 1.0) < 1e-7
    # weights: the oracle's first Adam step on the oracle gradient, where the sign is unambiguous
    opt = OracleOptimizer(o64, w.astype(np.float64))
    opt.step(g_ref * 3000, 3000)
    sure = np.abs(g_ref) > 1e-3 * np.abs(g_ref).max()
    assert np.allclose(s["weight"][sure], opt.w[sure], rtol=0, atol=2e-6)
    assert np.array_equal(s["weights_f16"], s["weight"].astype(np.float16))
    assert np.allclose(s["ema_weight"], s["weight"], rtol=0, atol=1e-6)   # first step: ema == w
    assert np.array_equal(s["use_weights_f16"], s["weights_f16"])


@pytest.mark.parametrize("hidden,enc,loss", [(5, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (5, v.ENCODING_SINCOS6, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (4, v.ENCODING_UV_SINCOS16, v.LOSS_L2),
                                             (4, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE)])
def test_fused_train_step_matches_oracle(o64, hidden, enc, loss):
    st, w, r, s, g, g_ref, l_ref = one_step(o64, v.PATH_FUSED_F16, hidden, enc, loss, 3000, 41 + hidden + enc)
    assert cosine(g, g_ref) > GRAD_COSINE_MIN, cosine(g, g_ref)
    assert rel_err(g, g_ref) < 5e-2
    assert abs(s["state"]["last_loss"] - l_ref) < 1e-2 * abs(l_ref)
    assert s["state"]["t"] == 1


@pytest.mark.parametrize("hidden,enc,loss", [(5, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (4, v.ENCODING_UV_SINCOS16, v.LOSS_L2),
                                             (4, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE),
                                             (5, v.ENCODING_SINCOS6, v.LOSS_L2)])
def test_fused_multi_chunk_step_matches_oracle_for_each_network_kind(o64, hidden, enc, loss):
    """4 chunks a workgroup over 3,000 records (12 workgroups, the last with 3 chunks, the
    last chunk partial): the multi-chunk kernel of each depth, encoding and loss."""
    st, w, r, s, g, g_ref, l_ref = one_step(o64, v.PATH_FUSED_F16, hidden, enc, loss, 3000, 51 + hidden + enc, chunks=4)
    assert cosine(g, g_ref) > GRAD_COSINE_MIN, cosine(g, g_ref)
    assert rel_err(g, g_ref) < 5e-2
    assert abs(s["state"]["last_loss"] - l_ref) < 1e-2 * abs(l_ref)
    assert s["state"]["t"] == 1 and s["state"]["last_count"] == 3000


@pytest.mark.parametrize("n", [40000, 65536])
def test_fused_multi_chunk_train_step_matches_oracle(o64, n):
    """Batches above 64 records per CU: workgroups walk several chunks and sum dW in registers."""
    rng = np.random.default_rng(90 + n % 7)
    st, w = make_cache(rng, train_batch_size=65536, encoding=v.ENCODING_SINCOS6)
    r = train_records(n, 91, v.ENCODING_SINCOS6)
    st.train(dev(r), path=v.PATH_FUSED_F16)
    s = st.get_state()
    g = s["m"].astype(np.float64) / (1.0 - np.float32(0.9))
    G, lsum = o64.gradient(f16(w), 5, r, v.ENCODING_SINCOS6, v.LOSS_RELATIVE_L2_LUMINANCE, round_f16=True)
    assert cosine(g, G / n) > GRAD_COSINE_MIN, cosine(g, G / n)
    assert rel_err(g, G / n) < 5e-2
    assert abs(s["state"]["last_loss"] - lsum / n) < 1e-2 * abs(lsum / n)
    assert s["state"]["last_count"] == n


@pytest.mark.parametrize("n,chunks", [(64, 1), (1000, 1), (40000, 1), (40000, 2), (40000, 16)])
def test_optimizer_reduction_over_any_partial_row_count(o64, n, chunks):
    """1, 16, 625, 313 and 40 partial rows: the optimizer's one-round reductions (1 to 16 loads a slice)
    and its multi-round loop (more than 256 rows) all give the oracle's gradient and loss."""
    rng = np.random.default_rng(70 + chunks)
    st, w = make_cache(rng, train_batch_size=65536, encoding=v.ENCODING_SINCOS6)
    st.set_grad_chunks(chunks)
    r = train_records(n, 71 + n % 5, v.ENCODING_SINCOS6)
    st.train(dev(r), path=v.PATH_FUSED_F16)
    s = st.get_state()
    g = s["m"].astype(np.float64) / (1.0 - np.float32(0.9))
    G, lsum = o64.gradient(f16(w), 5, r, v.ENCODING_SINCOS6, v.LOSS_RELATIVE_L2_LUMINANCE, round_f16=True)
    assert cosine(g, G / n) > GRAD_COSINE_MIN, cosine(g, G / n)
    assert rel_err(g, G / n) < 5e-2
    assert abs(s["state"]["last_loss"] - lsum / n) < 1e-2 * abs(lsum / n)
    assert s["state"]["last_count"] == n and s["state"]["t"] == 1


def test_train_is_deterministic():
    rng = np.random.default_rng(5)
    w = rng.normal(0, 0.17, v.weight_count(5)).astype(np.float32)
    r = dev(train_records(16384, 6))
    for path in (v.PATH_PLAIN_F32, v.PATH_FUSED_F16):
        res = []
        for _ in range(2):
            st = v.NRCState(v.NRCConfig())
            st.set_state(weight=w)
            for _ in range(3):
                st.train(r, path=path)
            res.append(st.get_state())
        assert np.array_equal(res[0]["weight"], res[1]["weight"])
        assert np.array_equal(res[0]["m"], res[1]["m"])


def test_empty_batch_changes_nothing():
    rng = np.random.default_rng(7)
    for path in (v.PATH_PLAIN_F32, v.PATH_FUSED_F16):
        st, w = make_cache(rng)
        before = st.get_state()
        c = torch.zeros(1, dtype=torch.int32, device=DEV)
        st.train(dev(train_records(512, 8)), count=c, path=path)
        after = st.get_state()
        for k in ("weight", "m", "v", "ema_weight", "weights_f16", "use_weights_f16"):
            assert np.array_equal(before[k], after[k]), k
        assert after["state"]["t"] == 0


def test_train_count_is_clamped_to_the_batch_size(o64):
    rng = np.random.default_rng(11)
    st, w = make_cache(rng, train_batch_size=1000)
    r = train_records(3000, 12)
    c = torch.tensor([2500], dtype=torch.int32, device=DEV)
    st.train(dev(r), count=c, max_count=3000, path=v.PATH_PLAIN_F32)
    s = st.get_state()
    assert s["state"]["last_count"] == 1000
    G, _ = o64.gradient(w.astype(np.float64), 5, r[:1000], v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE)
    assert cosine(s["m"] / 0.1, G) > 0.99999


@pytest.mark.parametrize("path", [v.PATH_PLAIN_F32, v.PATH_FUSED_F16])
def test_partial_batch_from_device_count(o64, path):
    """d_count well below max_count: workgroups past the count must contribute nothing,
    including on a cache whose partial buffers hold an earlier full step's values."""
    rng = np.random.default_rng(29)
    st, w = make_cache(rng)
    full = train_records(16384, 30)
    st.train(dev(full), path=path)            # fills every partial slot
    st.set_state(weight=w, state={"t": 0, "beta1_t": 1.0, "beta2_t": 1.0, "alpha_t": 1.0, "alpha_t_1": 0.0,
                                  "loss_scale": 1.0})   # back to w with fresh moments and scalars
    r = train_records(16384, 31)
    c = torch.tensor([1000], dtype=torch.int32, device=DEV)
    st.train(dev(r), count=c, path=path)
    s = st.get_state()
    assert s["state"]["last_count"] == 1000 and s["state"]["t"] == 1
    is16 = path == v.PATH_FUSED_F16
    G, _ = o64.gradient(f16(w) if is16 else w.astype(np.float64), 5, r[:1000], v.ENCODING_TRI12,
                        v.LOSS_RELATIVE_L2_LUMINANCE, round_f16=is16)
    g = s["m"].astype(np.float64) / (1.0 - np.float32(0.9))
    assert cosine(g, G / 1000) > (GRAD_COSINE_MIN if is16 else 0.99999)


def test_nonfinite_gradient_semantics():
    rng = np.random.default_rng(13)
    r = train_records(1024, 14)
    # static scale (reference): NaN/Inf gradient elements become 0, the step still counts
    st, w = make_cache(rng)
    bad = r.copy()
    bad[5, 14] = np.nan
    st.train(dev(bad), path=v.PATH_PLAIN_F32)
    s = st.get_state()
    assert s["state"]["t"] == 1 and np.isfinite(s["weight"]).all()
    # dynamic scale: a non-finite gradient skips the step and halves the scale
    st2, w2 = make_cache(rng, loss_scale=128.0, dynamic_loss_scale=True)
    huge = r.copy()
    huge[3, 14:17] = 1e38   # the FP32 output gradient itself overflows
    st2.train(dev(huge), path=v.PATH_FUSED_F16)
    s2 = st2.get_state()
    assert s2["state"]["t"] == 0 and s2["state"]["skipped_steps"] == 1 and s2["state"]["loss_scale"] == 64.0
    assert np.array_equal(s2["weight"], w2)
    mild = r.copy()
    mild[:, 14:17] = np.clip(mild[:, 14:17], 0.0, 1.0)
    st2.train(dev(mild), path=v.PATH_FUSED_F16)
    s3 = st2.get_state()
    assert s3["state"]["t"] == 1 and s3["state"]["loss_scale"] == 64.0 and s3["state"]["clean_steps"] == 1


@pytest.mark.parametrize("chunks", [0, 4])
def test_large_finite_gradients_are_scaled_not_skipped(o64, chunks):
    """An output gradient beyond FP16 range (2 * 128 * 3e5 / 0.01 here) puts the
    workgroup's deltas under a power-of-two scale instead of overflowing: the fused
    step is taken and matches the oracle like any other.  With 4 chunks a
    workgroup, record 3 scales the first chunk of workgroup 0 (three unscaled
    chunks follow) and record 2500 the last chunk of workgroup 9 (after three
    unscaled ones), so the dW sums carried across chunks are rescaled both ways."""
    rng = np.random.default_rng(15)
    r = train_records(3000, 16)
    r[3, 14:17] = 3e5
    r[2500, 14:17] = 1e6
    st, w = make_cache(rng, loss_scale=128.0, dynamic_loss_scale=True)
    st.set_grad_chunks(chunks)
    st.train(dev(r), path=v.PATH_FUSED_F16)
    s = st.get_state()
    assert s["state"]["t"] == 1 and s["state"]["skipped_steps"] == 0 and s["state"]["loss_scale"] == 128.0
    g = s["m"].astype(np.float64) / (1.0 - np.float32(0.9))
    G, _ = o64.gradient(f16(w), 5, r, v.ENCODING_TRI12, v.LOSS_RELATIVE_L2_LUMINANCE, round_f16=True)
    assert cosine(g, G / 3000) > GRAD_COSINE_MIN, cosine(g, G / 3000)


def test_frame_equals_infer_then_four_train_steps():
    rng = np.random.default_rng(17)
    w = rng.normal(0, 0.17, v.weight_count(5)).astype(np.float32)
    q = dev(wl.make_queries(20000, rng))
    b = dev(train_records(4 * 4096, 18))
    outs = []
    states = []
    for mode in ("frame", "manual"):
        st = v.NRCState(v.NRCConfig())
        st.set_state(weight=w)
        out = torch.zeros((20000, 3), dtype=torch.float32, device=DEV)
        if mode == "frame":
            st.frame(q, out, b, max_batch_records=4096)
        else:
            st.infer(q, out)
            for i in range(4):
                st.train(b[i * 4096:(i + 1) * 4096], write_use_weights=(i == 3))
        outs.append(host(out))
        states.append(st.get_state())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(states[0]["use_weights_f16"], states[1]["use_weights_f16"])
    assert states[0]["state"]["t"] == 4


def test_use_weights_follow_ema_when_enabled():
    rng = np.random.default_rng(19)
    st, w = make_cache(rng)
    st.set_use_ema_weights(True)
    r = dev(train_records(2048, 20))
    for i in range(3):
        st.train(r, write_use_weights=(i == 2), path=v.PATH_PLAIN_F32)
    s = st.get_state()
    assert np.array_equal(s["use_weights_f16"], s["ema_weight"].astype(np.float16))
    assert not np.array_equal(s["ema_weight"], s["weight"])


def test_full_size_fused_matches_plain_and_oracle(o64):
    """2,073,600 queries: fused vs plain at full size, and a sampled subset vs the oracle."""
    rng = np.random.default_rng(23)
    st, w = make_cache(rng, encoding=v.ENCODING_SINCOS6)
    n = 2073600
    q = wl.make_queries(n, rng)
    dq = dev(q)
    a = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    b = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    st.infer(dq, a, path=v.PATH_FUSED_F16)
    st.infer(dq, b, path=v.PATH_PLAIN_F32)
    ya, yb = host(a), host(b)
    assert rel_err(ya, yb) < FWD_REL_TOL
    idx = rng.choice(n, 8192, replace=False)
    ref = o64.infer(w.astype(np.float64), 5, q[idx], v.ENCODING_SINCOS6, clamp0=True)
    assert rel_err(yb[idx], ref) < 1e-5 and rel_err(ya[idx], ref) < FWD_REL_TOL
