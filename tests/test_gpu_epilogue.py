"""The inference epilogue (nrc_composite) through the C ABI: screen composite and
training-target back-fill, bit-exact against the CPU restatement (oracle/epilogue.py)."""
import numpy as np
import pytest

import vknrc_amd as v
from oracle.epilogue import INVALID, composite
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def make_case(rng, n, W, H, pitch, S):
    """n queries with non-overlapping destinations: a quarter invalid, screen pixels (some off the
    image), and disjoint record ranges of the 4 batches (some running past the batch)."""
    kinds = rng.choice(3, n, p=[0.25, 0.45, 0.30])
    dst = np.full(n, INVALID, np.uint32)
    pix = rng.permutation(W * H + 64)            # the last 64 fall outside the image
    pi = 0
    nxt = np.zeros(4, np.int64)
    for i in range(n):
        if kinds[i] == 1 and pi < len(pix):
            p = int(pix[pi]); pi += 1
            x, y = (p % W, p // W) if p < W * H else (W + p % 5, rng.integers(0, H + 3))
            dst[i] = v.dst_screen(x, y)
        elif kinds[i] == 2:
            b = int(rng.integers(0, 4))
            l = int(nxt[b]); ln = int(rng.integers(1, 6))
            if l < S + 4:
                dst[i] = v.dst_train(b, l, l + ln - 1)
                nxt[b] = l + ln
    rgb = np.abs(rng.normal(0, 1, (n, 3))).astype(np.float32)
    bf = rng.normal(0, 1, (H, pitch, 4)).astype(np.float32)
    fgb = rng.normal(0, 1, (H, pitch, 2)).astype(np.float32)
    batches = rng.normal(0, 1, (4 * S, 17)).astype(np.float32)
    factors = rng.normal(0, 1, (4 * S, 3)).astype(np.float32)
    return dst, rgb, bf, fgb, batches, factors


@pytest.mark.parametrize("n,W,H,pitch,S", [(5000, 61, 47, 64, 2000), (70001, 320, 200, 320, 16384), (1, 1, 1, 1, 1)])
def test_composite_matches_the_oracle_bit_for_bit(n, W, H, pitch, S):
    rng = np.random.default_rng(n)
    dst, rgb, bf, fgb, batches, factors = make_case(rng, n, W, H, pitch, S)
    dbf, dbt = dev(bf), dev(batches)
    v.composite(dev(rgb), dev(dst), n, bias_factor_r=dbf, factor_gb=dev(fgb), width=W, height=H, pitch=pitch,
                batches=dbt, train_factors=dev(factors), batch_stride=S)
    torch.cuda.synchronize()
    obf, obt = composite(rgb, dst, n, bf, fgb, W, H, batches, factors, S)
    assert np.array_equal(dbf.cpu().numpy(), obf)
    assert np.array_equal(dbt.cpu().numpy(), obt)
    assert (not np.array_equal(obf, bf)) or n == 1


def test_composite_respects_the_device_count():
    rng = np.random.default_rng(9)
    n, W, H, S = 4000, 50, 40, 3000
    dst, rgb, bf, fgb, batches, factors = make_case(rng, n, W, H, W, S)
    for cnt in (0, 1, 1234, 4000, 9999):
        dbf, dbt = dev(bf), dev(batches)
        v.composite(dev(rgb), dev(dst), n, count=dev(np.array([cnt], np.uint32)), bias_factor_r=dbf,
                    factor_gb=dev(fgb), width=W, height=H, batches=dbt, train_factors=dev(factors), batch_stride=S)
        torch.cuda.synchronize()
        obf, obt = composite(rgb, dst, min(cnt, n), bf, fgb, W, H, batches, factors, S)
        assert np.array_equal(dbf.cpu().numpy(), obf) and np.array_equal(dbt.cpu().numpy(), obt)


def test_infer_then_composite_is_the_reference_epilogue():
    """nrc_infer followed by nrc_composite: the composite of the inferred radiance, bit-exact."""
    rng = np.random.default_rng(11)
    st = v.NRCState(v.NRCConfig(encoding=v.ENCODING_TRI12, seed=3))
    n, W, H, S = 30000, 160, 120, 16384
    q = wl.make_queries(n, rng)
    dst, _, bf, fgb, batches, factors = make_case(rng, n, W, H, W, S)
    rgb = torch.zeros((n, 3), dtype=torch.float32, device=DEV)
    st.infer(dev(q), rgb)
    dbf, dbt = dev(bf), dev(batches)
    v.composite(rgb, dev(dst), n, bias_factor_r=dbf, factor_gb=dev(fgb), width=W, height=H, batches=dbt,
                train_factors=dev(factors), batch_stride=S)
    torch.cuda.synchronize()
    pred = rgb.cpu().numpy()
    assert np.all(pred >= 0)
    obf, obt = composite(pred, dst, n, bf, fgb, W, H, batches, factors, S)
    assert np.array_equal(dbf.cpu().numpy(), obf) and np.array_equal(dbt.cpu().numpy(), obt)


def test_composite_argument_checks():
    rgb = torch.zeros((4, 3), dtype=torch.float32, device=DEV)
    dst = dev(np.full(4, INVALID, np.uint32))
    bf = torch.zeros((2, 2, 4), dtype=torch.float32, device=DEV)
    fgb = torch.zeros((2, 2, 2), dtype=torch.float32, device=DEV)
    with pytest.raises(v.NRCError):
        v.composite(rgb, dst, 4, bias_factor_r=bf, factor_gb=fgb, width=3, height=2, pitch=2)   # pitch < width
    with pytest.raises(v.NRCError):
        v.composite(rgb, dst, 4, batches=torch.zeros((4, 17), device=DEV), train_factors=None, batch_stride=1)
    v.composite(rgb, dst, 0)  # nothing to do
