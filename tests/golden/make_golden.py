"""Writes the golden fixtures in tests/golden/*.npz with torch float64 autograd.

The reference (AdamYuan/VkNRC) ships no golden vectors, no known-answer tests and
no Python, and its GPU path (GLSL + VK_NV_cooperative_matrix) cannot run here.
These fixtures are therefore minted from an independent float64 restatement in
torch: the forward is a chain of library matmuls and the weight gradient comes
from torch autograd, not from a hand-written backward, so they pin the oracle's
backward (dropped terms, transposed operands, wrong mask) as well as its forward.

Semantics restated (reference files):
  encoding        shader/src/NRCRecord.glsl:47-95 (tri waves use GLSL floor-mod)
  forward         shader/src/NN_nv.glsl:99-158, nrc_gradient.comp:38-45
  relative-L2     shader/src/NN_nv.glsl:178-196 (denominator held constant)
  L2              shader/src/NN_nv.glsl:162-177
  Adam + EMA      shader/src/nrc_optimize.comp:32-54, nrc_train_prepare.comp:16-29

Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os

import numpy as np
import torch

torch.set_default_dtype(torch.float64)
HERE = os.path.dirname(os.path.abspath(__file__))


def tri(x):
    a = x - 0.5
    return 2.0 * torch.abs(torch.remainder(a, 2.0) - 1.0) - 1.0


def quartic_cdf(d):
    u = 4.0 * d
    return torch.clamp(15.0 / 16.0 * u * (1 - 2.0 / 3.0 * u * u + 0.2 * u ** 4) + 0.5, 0.0, 1.0)


def oneblob4(x):
    ks = torch.arange(4, dtype=torch.float64)
    return quartic_cdf(0.25 * (ks + 1) - x[:, None]) - quartic_cdf(0.25 * ks - x[:, None])


def sinpi(x):
    return torch.sin(math.pi * torch.remainder(x, 2.0))


def cospi(x):
    return torch.cos(math.pi * torch.remainder(x, 2.0))


def encode(rec, enc):
    """rec: [n, >=14] float32 values held in float64."""
    n = rec.shape[0]
    if enc == 2:  # uv sin/cos, 16 octaves
        cols = []
        for a in range(2):
            x = rec[:, a:a + 1] * (2.0 ** torch.arange(16, dtype=torch.float64))
            cols += [sinpi(x), cospi(x)]
        return torch.cat(cols, 1)
    cols = []
    for a in range(3):
        if enc == 0:
            cols.append(tri(rec[:, a:a + 1] * (2.0 ** torch.arange(12, dtype=torch.float64))))
        else:
            x = rec[:, a:a + 1] * (2.0 ** torch.arange(6, dtype=torch.float64))
            cols += [sinpi(x), cospi(x)]
    cols += [oneblob4(rec[:, 3]), oneblob4(rec[:, 4]), oneblob4(rec[:, 5]), oneblob4(rec[:, 6]),
             oneblob4(1.0 - torch.exp(-rec[:, 7])), rec[:, 8:14], torch.ones(n, 2)]
    return torch.cat(cols, 1)


def split(w, hidden):
    mats = [w[l * 4096:(l + 1) * 4096].view(64, 64) for l in range(hidden)]
    mats.append(w[hidden * 4096:hidden * 4096 + 192].view(3, 64))
    return mats


def forward(w, hidden, x):
    a = x.T  # features x samples
    mats = split(w, hidden)
    for l in range(hidden):
        a = torch.relu(mats[l] @ a)
    return (mats[hidden] @ a).T


def loss_sum(o, t, kind, scale=1.0):
    if kind == 0:
        lum = (torch.clamp(o, min=0.0) * torch.tensor([0.299, 0.587, 0.114])).sum(1, keepdim=True)
        d = (lum * lum + 0.01).detach()
        return scale * (((o - t) ** 2) / d).sum(), (((o - t) ** 2) / d).sum()
    return scale * ((o - t) ** 2).sum(), ((o - t) ** 2).sum()


def grad(w, hidden, x, t, kind, scale=1.0):
    wv = w.clone().requires_grad_(True)
    o = forward(wv, hidden, x)
    scaled, raw = loss_sum(o, t, kind, scale)
    scaled.backward()
    return wv.grad.detach().clone(), float(raw.detach())


def adam_steps(w0, grads_and_counts, lr, b1, b2, eps, alpha):
    w = w0.clone()
    ema = w0.clone()
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    B1 = B2 = A = 1.0
    A1 = 0.0
    for G, c in grads_and_counts:
        B1 *= b1
        B2 *= b2
        A1 = A
        A *= alpha
        g = G / c
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        w = w - lr * (m / (1 - B1)) / (torch.sqrt(v / (1 - B2)) + eps)
        ema = (1 - alpha) / (1 - A) * w + alpha * (1 - A1) * ema
    return w, ema, m, v, np.array([B1, B2, A, A1])


def records(rng, n):
    r = np.zeros((n, 17), np.float32)
    r[:, 0:3] = rng.uniform(-1.5, 1.5, (n, 3))
    r[:, 3:7] = rng.uniform(0, 1, (n, 4))
    r[:, 7] = rng.uniform(0, 2, n)
    r[:, 8:14] = rng.uniform(0, 1, (n, 6))
    r[:, 14:17] = rng.lognormal(-1.0, 1.0, (n, 3))
    return r


def main():
    rng = np.random.default_rng(20240223)
    out = {}
    # encoder fixtures: random records plus hand-picked edge values
    rec = records(rng, 24)
    rec[0, 0:3] = [0.0, 0.5, -0.5]
    rec[1, 0:3] = [1.0, -1.0, 1.5]
    rec[2, 3:7] = [0.0, 1.0, 0.25, 0.75]
    rec[3, 7] = 0.0
    t = torch.from_numpy(rec.astype(np.float64))
    for enc in (0, 1, 2):
        out[f"enc{enc}_features"] = encode(t, enc).numpy()
    out["enc_records"] = rec

    # MLP fixtures for both hidden depths and both losses
    for hidden in (5, 4):
        wc = hidden * 4096 + 192
        w = torch.from_numpy(rng.normal(0.0, math.sqrt(2.0 / 64.0), wc))
        rec = records(rng, 40)
        t = torch.from_numpy(rec.astype(np.float64))
        x = encode(t, 0)
        tgt = t[:, 14:17]
        o = forward(w, hidden, x)
        out[f"h{hidden}_weights"] = w.numpy()
        out[f"h{hidden}_records"] = rec
        out[f"h{hidden}_out_raw"] = o.detach().numpy()
        for kind in (0, 1):
            G, L = grad(w, hidden, x, tgt, kind, scale=1.0)
            out[f"h{hidden}_loss{kind}_dw"] = G.numpy()
            out[f"h{hidden}_loss{kind}_sum"] = np.array([L])

    # the reference test harness's distribution (test/main.cpp:150-162): weights
    # U(-0.02, 0.02), pre-encoded inputs U(0, 1), targets U(0, 1), L2 loss
    w = torch.from_numpy(rng.uniform(-0.02, 0.02, 5 * 4096 + 192))
    x = torch.from_numpy(rng.uniform(0, 1, (128, 64)))
    tg = torch.from_numpy(rng.uniform(0, 1, (128, 3)))
    out["harness_weights"] = w.numpy()
    out["harness_inputs"] = x.numpy()
    out["harness_targets"] = tg.numpy()
    out["harness_out"] = forward(w, 5, x).detach().numpy()
    G, L = grad(w, 5, x, tg, 1)
    out["harness_dw"] = G.numpy()

    # three Adam steps of the reference optimizer on the h5 relative-L2 gradient
    w = torch.from_numpy(out["h5_weights"])
    G = torch.from_numpy(out["h5_loss0_dw"])
    seq = [(G, 40), (0.5 * G, 40), (-G, 20)]
    wf, ema, m, v, sc = adam_steps(w, seq, 0.002, 0.9, 0.999, 1e-8, 0.99)
    out["adam_w"], out["adam_ema"], out["adam_m"], out["adam_v"], out["adam_scalars"] = (
        wf.numpy(), ema.numpy(), m.numpy(), v.numpy(), sc)

    np.savez_compressed(os.path.join(HERE, "golden_v1.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_v1.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
