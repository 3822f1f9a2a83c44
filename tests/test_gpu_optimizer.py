"""Multi-step parity of the GPU optimizer (Adam + EMA + the prepare scalars) with the oracle.

The reference's update is nrc_optimize.comp:32-54 with the running products of
nrc_train_prepare.comp:16-29.  Past the first step the bias corrections divide by
1 - beta^t with beta^t != beta, and the EMA (as written, nrc_optimize.comp:46-47)
takes its alpha^{t-1} term, so these tests run ten chained steps per cache and check
every one of them.

Each step's gradient is recovered from consecutive first moments,
g_t = (m_t - b1 m_{t-1}) / (1 - b1), with the cache's own FP32 constants.  The FP64
oracle (OracleOptimizer) then steps (i) from the GPU's previous state, which pins the
per-step update to FP32 rounding, and (ii) from its own chained state, which pins the
recurrence over all steps.  The scalars t, beta1^t, beta2^t, alpha^t and alpha^{t-1}
are compared bit for bit with FP32 running products.  With a dynamic loss scale, one
batch carries a target that overflows the FP32 output gradient: that step must change
no weight and no moment, halve the scale and advance neither t nor the products.
"""
import numpy as np
import pytest

import vknrc_amd as v
from oracle.oracle import Oracle, OptState, OracleOptimizer
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"
F32 = np.float32
B1, B2, AL, LR, EPS = F32(0.9), F32(0.999), F32(0.99), F32(0.002), F32(1e-8)
STEPS = 10
SKIP_AT = 4     # the forced overflow step of the dynamic-scale runs
N = 2048


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def oracle_from(o64, s, dynamic):
    """An FP64 oracle optimizer holding the GPU state s (entries and scalars)."""
    opt = OracleOptimizer(o64, s["weight"].astype(np.float64), lr=float(LR), beta1=float(B1), beta2=float(B2),
                          eps=float(EPS), ema_alpha=float(AL), loss_scale=float(s["state"]["loss_scale"]),
                          dynamic_scale=dynamic, growth_interval=2000)
    opt.m = s["m"].astype(np.float64)
    opt.v = s["v"].astype(np.float64)
    opt.ema = s["ema_weight"].astype(np.float64)
    st = s["state"]
    opt.state = OptState(float(st["t"]), float(st["beta1_t"]), float(st["beta2_t"]), float(st["alpha_t"]),
                         float(st["alpha_t_1"]), float(st["loss_scale"]), float(st["clean_steps"]),
                         float(st["skipped_steps"]))
    return opt


@pytest.mark.parametrize("dynamic", [False, True], ids=["static", "dynamic"])
@pytest.mark.parametrize("use_ema", [False, True], ids=["w", "ema"])
@pytest.mark.parametrize("path", [v.PATH_PLAIN_F32, v.PATH_FUSED_F16], ids=["plain", "fused"])
def test_ten_optimizer_steps_match_oracle(o64, path, use_ema, dynamic):
    rng = np.random.default_rng(300 + 4 * path + 2 * use_ema + dynamic)
    cfg = v.NRCConfig(encoding=v.ENCODING_SINCOS6, use_ema_weights=use_ema, dynamic_loss_scale=dynamic,
                      loss_scale=128.0 if dynamic else 1.0)
    st = v.NRCState(cfg)
    w0 = rng.normal(0.0, np.sqrt(2.0 / 64.0), v.weight_count(5)).astype(np.float32)
    st.set_state(weight=w0)
    lights = wl.GaussianLights(11)
    prev = st.get_state()
    chain = oracle_from(o64, prev, dynamic)
    b1t = b2t = at = F32(1.0)
    at1 = F32(0.0)
    t = 0
    for k in range(STEPS):
        r = wl.make_train_records(N, rng, lights)
        forced = dynamic and k == SKIP_AT
        if forced:
            r[17, 14:17] = 1e38          # the FP32 output gradient itself overflows
        write_use = (k % 3 == 2) or k == STEPS - 1
        st.train(dev(r), path=path, write_use_weights=write_use)
        cur = st.get_state()
        s0, s1 = prev["state"], cur["state"]
        if forced:
            for key in ("weight", "m", "v", "ema_weight", "weights_f16"):
                assert np.array_equal(cur[key], prev[key]), key
            assert s1["t"] == t and s1["skipped_steps"] == s0["skipped_steps"] + 1
            assert s1["loss_scale"] == s0["loss_scale"] / 2 and s1["clean_steps"] == 0
            assert (s1["beta1_t"], s1["beta2_t"], s1["alpha_t"], s1["alpha_t_1"]) == \
                   (s0["beta1_t"], s0["beta2_t"], s0["alpha_t"], s0["alpha_t_1"])
            G = np.ones(v.weight_count(5))
            G[3] = np.inf
            assert not chain.step(G, N)
            assert chain.state.loss_scale == s1["loss_scale"]
            prev = cur
            continue
        t += 1
        # scalars: FP32 running products, bit for bit (nrc_train_prepare.comp:23-27)
        b1t, b2t = F32(b1t * B1), F32(b2t * B2)
        at1, at = at, F32(at * AL)
        assert s1["t"] == t
        assert (F32(s1["beta1_t"]), F32(s1["beta2_t"]), F32(s1["alpha_t"]), F32(s1["alpha_t_1"])) == \
               (b1t, b2t, at, at1), (k, s1)
        assert s1["loss_scale"] == s0["loss_scale"] and s1["last_count"] == N
        # this step's gradient g = G / count / scale, recovered from the first moments
        m0, m1 = prev["m"].astype(np.float64), cur["m"].astype(np.float64)
        g = (m1 - float(B1) * m0) / float(F32(1.0) - B1)
        G = g * N * float(s0["loss_scale"])
        # (i) one oracle step from the GPU's previous state
        one = oracle_from(o64, prev, dynamic)
        assert one.step(G, N)
        assert np.allclose(cur["m"], one.m, rtol=1e-6, atol=1e-12)
        assert np.allclose(cur["v"], one.v, rtol=1e-5, atol=1e-14), k
        assert np.allclose(cur["weight"], one.w, rtol=2e-7, atol=2e-8), (k, np.abs(cur["weight"] - one.w).max())
        assert np.allclose(cur["ema_weight"], one.ema, rtol=3e-5, atol=1e-7), (k, np.abs(cur["ema_weight"] - one.ema).max())
        # (ii) the chained oracle from the initial state
        assert chain.step(G, N)
        assert np.allclose(cur["weight"], chain.w, rtol=1e-6, atol=2e-7), (k, np.abs(cur["weight"] - chain.w).max())
        assert np.allclose(cur["ema_weight"], chain.ema, rtol=1e-4, atol=1e-6), k
        assert np.allclose(cur["v"], chain.v, rtol=1e-4, atol=1e-13), k
        # the fp16 copies; the use-weights are written on the flagged steps only
        assert np.array_equal(cur["weights_f16"], cur["weight"].astype(np.float16))
        if write_use:
            want = (cur["ema_weight"] if use_ema else cur["weight"]).astype(np.float16)
            assert np.array_equal(cur["use_weights_f16"], want)
        else:
            assert np.array_equal(cur["use_weights_f16"], prev["use_weights_f16"])
        prev = cur
    assert t == (STEPS - 1 if dynamic else STEPS)
    # past t = 1 the EMA as written differs from the weight (it equals it only at t = 1)
    if t >= 2:
        assert not np.allclose(prev["ema_weight"], prev["weight"], rtol=1e-3)
    st.close()
