"""The FP8 (OCP e4m3) inference forward, through the C ABI.

Per-layer power-of-two scales come from the largest activations of the cache's
last 16 FP8 inferences (delayed scaling), so the first FP8 inference after a
cache is created, or after its weights jump, runs on stale ranges and the next
ones on measured ranges.

Parity bar: the kernel against the FP8 oracle (oracle/fp8.py) run with the
scales the cache reports for that inference, relative L2 error within
FP8_ORACLE_TOL.  The two differ only where the MFMA's FP32 accumulation order
moves a sum across an e4m3 rounding boundary (one e4m3 step is 6-12 % of the
value, so such flips carry into later layers).  The error against the FP64
forward and against the FP16 path is e4m3's own (3 mantissa bits): it is
reported, and bounded loosely.
"""
import numpy as np
import pytest

import vknrc_amd as v
from oracle.fp8 import infer_fp8
from oracle.oracle import Oracle
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"
FP8_ORACLE_TOL = 0.02   # kernel vs FP8 oracle with the same scales, trained network
RANDOM_ORACLE_TOL = 0.05  # the same on an untrained Kaiming network, whose output cancels more
RANDOM_NET_TOL = 0.5    # FP8 vs FP64 on an untrained Kaiming network (e4m3 error compounds over 6 layers)
TRAINED_NET_TOL = 0.2   # FP8 vs FP64 / FP16 on the trained radiance cache


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


@pytest.fixture(scope="module")
def o32():
    return Oracle("f32")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def rel_err(y, ref):
    y = np.asarray(y, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(y - ref) / max(np.linalg.norm(ref), 1e-300))


def infer(st, q, path):
    out = torch.full((q.shape[0], 3), -7.0, dtype=torch.float32, device=DEV)
    st.infer(q, out, path=path)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def fp8_oracle(st, o32, q, enc, hidden, use_weights):
    sc = st.get_fp8_scales()
    feats = o32.encode(q, enc)
    return infer_fp8(feats, use_weights, hidden, sc["layer_scales"], sc["out_scale"], clamp0=True), sc


@pytest.mark.parametrize("hidden,enc", [(5, v.ENCODING_SINCOS6), (5, v.ENCODING_TRI12), (4, v.ENCODING_UV_SINCOS16)])
def test_fp8_infer_matches_the_fp8_oracle(o64, o32, hidden, enc):
    rng = np.random.default_rng(40 + hidden + enc)
    st = v.NRCState(v.NRCConfig(hidden_layers=hidden, encoding=enc))
    w = rng.normal(0.0, np.sqrt(2.0 / 64.0), v.weight_count(hidden)).astype(np.float32)
    st.set_state(weight=w)
    n = 70001  # a ragged last tile
    q = wl.make_queries(n, rng)
    if enc == v.ENCODING_UV_SINCOS16:
        q[:, 0:2] = rng.uniform(0, 1, (n, 2))
    dq = dev(q)
    infer(st, dq, v.PATH_FUSED_F8)          # first call: records the activation ranges
    y8 = infer(st, dq, v.PATH_FUSED_F8)
    y16 = infer(st, dq, v.PATH_FUSED_F16)
    yo, sc = fp8_oracle(st, o32, q, enc, hidden, w)
    ref = o64.infer(w.astype(np.float64), hidden, q, enc, round_f16=False, clamp0=True)
    e_or, e64, e16 = rel_err(y8, yo), rel_err(y8, ref), rel_err(y8, y16)
    same = float(np.mean(np.all(np.abs(y8 - yo) <= 1e-6 * np.abs(yo) + 1e-30, axis=1)))
    print(f"H={hidden} enc={enc} scales {sc['layer_scales']} out {sc['out_scale']}: fp8 vs fp8-oracle {e_or:.5f} "
          f"({same:.3f} of the records equal), vs fp64 {e64:.4f} (oracle {rel_err(yo, ref):.4f}), vs fp16 {e16:.4f}")
    assert np.isfinite(y8).all() and np.all(y8 >= 0)
    assert sc["frames"] == 1 and sc["layer_scales"][0] == 1.0
    assert e_or < RANDOM_ORACLE_TOL
    assert e64 < RANDOM_NET_TOL and e16 < RANDOM_NET_TOL
    # the same inputs, weights and history give the same bits
    y8b = infer(st, dq, v.PATH_FUSED_F8)
    assert st.get_fp8_scales()["frames"] == 2
    assert np.array_equal(y8b, y8)


def test_fp8_scales_follow_the_activation_ranges(o32):
    """Weights x2: the output of layer l grows by 2^(l+1).  The next inference runs on the old
    scales, so the activations past the first layer saturate at 448 and the deeper layers record
    too small maxima; each later inference corrects one more layer, and after H + 1 of them the
    scales are exactly the old ones divided by 2^l."""
    rng = np.random.default_rng(44)
    st = v.NRCState(v.NRCConfig(encoding=v.ENCODING_SINCOS6))
    w = rng.normal(0.0, np.sqrt(2.0 / 64.0), v.weight_count(5)).astype(np.float32)
    st.set_state(weight=w)
    q = wl.make_queries(20000, rng)
    dq = dev(q)
    for _ in range(3):
        infer(st, dq, v.PATH_FUSED_F8)
    old = st.get_fp8_scales()["layer_scales"].copy()
    w2 = (w * 2.0).astype(np.float32)
    st.set_state(weight=w2)                 # set_state keeps the FP8 history
    y_first = infer(st, dq, v.PATH_FUSED_F8)
    yo_first, sc_first = fp8_oracle(st, o32, q, v.ENCODING_SINCOS6, 5, w2)
    for _ in range(6):
        y_next = infer(st, dq, v.PATH_FUSED_F8)
    yo_next, sc_next = fp8_oracle(st, o32, q, v.ENCODING_SINCOS6, 5, w2)
    print("scales before", old, "first", sc_first["layer_scales"], "next", sc_next["layer_scales"])
    assert np.array_equal(sc_first["layer_scales"], old)
    assert np.array_equal(sc_next["layer_scales"], old / 2.0 ** np.arange(6))
    e_first, e_next = rel_err(y_first, yo_first), rel_err(y_next, yo_next)
    print(f"kernel vs fp8 oracle: first {e_first:.5f}, next {e_next:.5f}")
    assert e_first < RANDOM_ORACLE_TOL and e_next < RANDOM_ORACLE_TOL


def test_fp8_frames_of_training_and_inference(o64, o32):
    """500 frames of the radiance cache (FP8 inference of 8192 queries + 4 FP16 training steps of
    16384 records each): the per-frame error against FP16 inference with the same use weights, and
    every 50 frames against the FP8 oracle and the FP64 forward."""
    rng = np.random.default_rng(45)
    st = v.NRCState(v.NRCConfig(encoding=v.ENCODING_SINCOS6, loss_scale=128.0, dynamic_loss_scale=True,
                                seed=wl.workload_seed(wl.RADIANCE_FP8)))
    lights = wl.GaussianLights(wl.workload_seed(wl.RADIANCE_FP8, 2))
    pool = 32
    recs = dev(wl.make_train_records(pool * 4 * 16384, rng, lights))
    qe = wl.make_queries(8192, rng)
    dqe = dev(qe)
    out8 = torch.zeros((8192, 3), dtype=torch.float32, device=DEV)
    out16 = torch.zeros((8192, 3), dtype=torch.float32, device=DEV)
    e16, e_or, e64 = [], [], []
    for f in range(500):
        b = recs[(f % pool) * 4 * 16384:][:4 * 16384]
        st.infer(dqe, out16, path=v.PATH_FUSED_F16)
        st.frame(dqe, out8, b, max_batch_records=16384, path=v.PATH_FUSED_F8)
        torch.cuda.synchronize()
        e16.append(rel_err(out8.cpu().numpy(), out16.cpu().numpy()))
        if f % 50 == 49:
            y8 = infer(st, dqe, v.PATH_FUSED_F8)
            w = st.get_state()["weight"]
            yo, _ = fp8_oracle(st, o32, qe, v.ENCODING_SINCOS6, 5, w)
            ref = o64.infer(w.astype(np.float64), 5, qe, v.ENCODING_SINCOS6, round_f16=False, clamp0=True)
            e_or.append(rel_err(y8, yo))
            e64.append(rel_err(y8, ref))
    e16 = np.array(e16)
    print("fp8 vs fp16, max per 50 frames:", np.round(e16.reshape(-1, 50).max(1), 4))
    print("fp8 vs fp16, mean per 50 frames:", np.round(e16.reshape(-1, 50).mean(1), 4))
    print("fp8 vs fp8 oracle every 50 frames:", np.round(e_or, 5))
    print("fp8 vs fp64 every 50 frames:", np.round(e64, 4))
    print("loss", st.get_optimizer_state()["last_loss"])
    assert np.isfinite(e16).all()
    assert max(e_or) < FP8_ORACLE_TOL
    assert e16[50:].max() < TRAINED_NET_TOL
    assert max(e64) < TRAINED_NET_TOL
