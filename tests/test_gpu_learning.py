"""Learning behaviour over 1000 optimizer steps, through the C ABI.

Two workloads, each trained step for step on identical batches by the plain
FP32 path and the fused FP16-MFMA path:

  * the image MLP (64 -> 4 x 64 -> 3, UV_SINCOS16 encoding, Xavier init, L2 loss,
    Adam lr 1e-2, beta2 0.99, batch 16384; the reference's
    test/mlp_learning_an_image/{gradient,optimize}.comp), also trained by the
    FP32 CPU port of the same algorithm (oracle/nrc_oracle.c built with
    REAL=float), then evaluated by inference on 65,536 fresh uv with the EMA
    weights (SetUseEMAWeights), which average out the last steps' noise;
  * the radiance cache (SINCOS6 encoding, relative-L2 loss, dynamic loss scale
    starting at 128, 5 hidden layers).

The bar for the fused path: its loss curve, as means over 100-step windows,
stays within 2 % of the plain path's.
"""
import os

import numpy as np
import pytest

import vknrc_amd as v
from oracle.oracle import Oracle, OracleOptimizer
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"
STEPS = 1000
BATCH = 16384
WINDOW = 100
CURVE_TOL = 0.02  # windowed mean loss, fused vs plain


def windowed(x):
    x = np.asarray(x, np.float64)
    return x[:len(x) // WINDOW * WINDOW].reshape(-1, WINDOW).mean(1)


def worst_ratio(a, b):
    return float(np.max(np.abs(windowed(a) / windowed(b) - 1.0)))


def image_mse(pred, target):
    return float(np.mean((np.asarray(pred, np.float64) - target) ** 2))


def test_image_mlp_learns_like_the_cpu_port_and_the_plain_path():
    img = wl.make_image(wl.workload_seed(wl.IMAGE_MLP, 1))
    kw = dict(hidden_layers=4, encoding=v.ENCODING_UV_SINCOS16, loss=v.LOSS_L2, init=v.INIT_XAVIER_UNIFORM,
              seed=wl.workload_seed(wl.IMAGE_MLP), learning_rate=1e-2, beta1=0.9, beta2=0.99, epsilon=1e-8,
              use_ema_weights=True)
    plain = v.NRCState(v.NRCConfig(**kw))
    fused = v.NRCState(v.NRCConfig(**kw))
    w0 = plain.get_state()["weight"]
    assert np.array_equal(w0, fused.get_state()["weight"])
    cpu = Oracle("f32")
    try:
        cpu.set_threads(min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        pass
    opt = OracleOptimizer(cpu, w0, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8)

    rng = np.random.default_rng(wl.workload_seed(wl.IMAGE_MLP, 2))
    uv_eval = np.zeros((65536, wl.QUERY_FLOATS), np.float32)
    uv_eval[:, 0:2] = rng.uniform(0, 1, (65536, 2))
    target_eval = wl.sample_bilinear(img, uv_eval[:, 0:2].astype(np.float64)).astype(np.float64)
    dq = torch.from_numpy(uv_eval).to(DEV)
    out = torch.zeros((65536, 3), dtype=torch.float32, device=DEV)

    def gpu_mse(st, path):
        st.infer(dq, out, path=path)
        torch.cuda.synchronize()
        return image_mse(out.cpu().numpy(), target_eval)

    mse0 = gpu_mse(plain, v.PATH_PLAIN_F32)
    lp, lf, lc = [], [], []
    for _ in range(STEPS):
        rec = wl.make_image_records(BATCH, rng, img)
        d = torch.from_numpy(rec).to(DEV)
        plain.train(d, path=v.PATH_PLAIN_F32)
        fused.train(d, path=v.PATH_FUSED_F16)
        G, L = cpu.gradient(opt.w, 4, rec, v.ENCODING_UV_SINCOS16, v.LOSS_L2)
        opt.step(G, BATCH)
        lc.append(L / BATCH)
        lp.append(plain.get_optimizer_state()["last_loss"])
        lf.append(fused.get_optimizer_state()["last_loss"])

    mse_plain = gpu_mse(plain, v.PATH_PLAIN_F32)
    mse_fused = gpu_mse(fused, v.PATH_FUSED_F16)
    mse_cpu = image_mse(cpu.infer(opt.ema, 4, uv_eval, v.ENCODING_UV_SINCOS16, clamp0=True), target_eval)
    print(f"image MSE: initial {mse0:.5f}  cpu {mse_cpu:.5f}  plain {mse_plain:.5f}  fused {mse_fused:.5f}")
    print("windowed loss cpu  ", np.round(windowed(lc), 5))
    print("windowed loss plain", np.round(windowed(lp), 5))
    print("windowed loss fused", np.round(windowed(lf), 5))
    assert mse_plain < 0.25 * mse0
    # the same FP32 algorithm on the GPU and the CPU: same learning, up to summation order
    assert worst_ratio(lp, lc) <= CURVE_TOL
    assert abs(mse_plain / mse_cpu - 1.0) <= 0.05
    # the fused FP16 path against the plain FP32 path
    assert worst_ratio(lf, lp) <= CURVE_TOL
    assert abs(mse_fused / mse_plain - 1.0) <= 0.05


def test_radiance_cache_loss_curve_fused_within_2pct_of_plain():
    kw = dict(encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE, loss_scale=128.0,
              dynamic_loss_scale=True, seed=wl.workload_seed(wl.RADIANCE))
    plain = v.NRCState(v.NRCConfig(**kw))
    fused = v.NRCState(v.NRCConfig(**kw))
    rng = np.random.default_rng(wl.workload_seed(wl.RADIANCE, 3))
    pool = 128
    recs = torch.from_numpy(wl.make_train_records(pool * BATCH, rng, wl.GaussianLights(wl.workload_seed(wl.RADIANCE, 2)))).to(DEV)
    lp, lf = [], []
    for s in range(STEPS):
        b = recs[(s % pool) * BATCH:][:BATCH]
        plain.train(b, path=v.PATH_PLAIN_F32, write_use_weights=False)
        fused.train(b, path=v.PATH_FUSED_F16, write_use_weights=False)
        lp.append(plain.get_optimizer_state()["last_loss"])
        lf.append(fused.get_optimizer_state()["last_loss"])
    sp, sf = plain.get_optimizer_state(), fused.get_optimizer_state()
    print("windowed loss plain", np.round(windowed(lp), 5))
    print("windowed loss fused", np.round(windowed(lf), 5))
    print(f"skipped steps plain {sp['skipped_steps']} fused {sf['skipped_steps']}; "
          f"loss scale plain {sp['loss_scale']} fused {sf['loss_scale']}")
    assert windowed(lp)[-1] < 0.5 * windowed(lp)[0]
    assert worst_ratio(lf, lp) <= CURVE_TOL
