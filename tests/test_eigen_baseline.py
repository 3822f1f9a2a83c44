"""The Eigen CPU baseline (oracle/eigen_mlp.cpp on the reference's vendored Eigen 3.4.0)
against the FP64 oracle: it is only a timing baseline, but what it times must be the
same MLP.  Skipped where the library was not built (no reference Eigen there)."""
import numpy as np
import pytest

from oracle import eigen
from oracle.oracle import LOSS_L2, Oracle, OracleOptimizer, weight_count

pytestmark = pytest.mark.skipif(not eigen.available(), reason="oracle/_ref/libeigen_mlp.so not built")


@pytest.fixture(scope="module")
def mlp():
    m = eigen.EigenMLP()
    m.set_threads(2)
    return m


def net(rng, hidden, n):
    w = rng.normal(0, np.sqrt(2 / 64), weight_count(hidden)).astype(np.float32)
    x = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    t = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    return w, x, t


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("hidden", [4, 5])
def test_eigen_forward_matches_oracle(mlp, oracle64, hidden):
    rng = np.random.default_rng(80 + hidden)
    w, x, _ = net(rng, hidden, 777)
    ref = oracle64.infer_encoded(w.astype(np.float64), hidden, x.astype(np.float64))
    assert rel(mlp.forward_f32(w, hidden, x), ref) < 1e-5
    # the reference test's _Float16 forward: every product and sum rounded to FP16
    w16, x16 = w.astype(np.float16), x.astype(np.float16)
    ref16 = oracle64.infer_encoded(w16.astype(np.float64), hidden, x16.astype(np.float64))
    e = rel(mlp.forward_f16(w16, hidden, x16), ref16)
    assert 1e-5 < e < 3e-2, e          # FP16 accumulation error, and not FP32 arithmetic


def test_eigen_gradient_and_adam_step_match_oracle(mlp, oracle64):
    rng = np.random.default_rng(88)
    w, x, t = net(rng, 4, 1500)
    dw, loss = mlp.gradient_l2_f32(w, 4, x, t)
    G, L = oracle64.gradient_encoded(w.astype(np.float64), 4, x.astype(np.float64), t.astype(np.float64), LOSS_L2)
    assert rel(dw, G) < 1e-4 and abs(loss - L) < 1e-4 * L
    w1, m, v = w.copy(), np.zeros_like(w), np.zeros_like(w)
    mlp.train_step_l2_f32(w1, m, v, 4, x, t, 1e-2, 0.9, 0.99, 1e-8, 0.9, 0.99)
    opt = OracleOptimizer(oracle64, w.astype(np.float64), lr=1e-2, beta2=0.99)
    opt.step(G, 1500)
    sure = np.abs(G) > 1e-3 * np.abs(G).max()
    assert np.allclose(w1[sure], opt.w[sure], rtol=0, atol=1e-6)
