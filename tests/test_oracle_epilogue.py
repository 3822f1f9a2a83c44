"""Pins the inference-epilogue oracle (oracle/epilogue.py): the destination codec and the
composite/back-fill arithmetic, on the CPU."""
import numpy as np

import vknrc_amd as v
from oracle.epilogue import INVALID, composite, decode, dst_screen, dst_train


def test_codec_known_answers():
    # EncodeNRCEvalDstScreen((3, 5)) = (3 | 5 << 15) << 1, evaluated by hand
    assert int(dst_screen(3, 5)) == 327686
    # EncodeNRCEvalDstTrain(2, 7, 9) = (2 | 7 << 2 | 9 << 16) << 1 | 1
    assert int(dst_train(2, 7, 9)) == 1179709
    assert decode(327686) == (0, 3, 5, 0)
    assert decode(1179709) == (1, 2, 7, 9)
    assert decode(INVALID)[0] == -1
    assert int(dst_screen(0x7FFF, 0x7FFF)) == 0x7FFFFFFE          # largest screen code
    assert int(dst_train(3, 0x3FFF, 0x3FFF)) == 0x7FFFFFFF        # largest train code


def test_codec_round_trips_and_matches_the_library_helpers():
    rng = np.random.default_rng(0)
    x, y = rng.integers(0, 1 << 15, 1000), rng.integers(0, 1 << 15, 1000)
    b, l, r = rng.integers(0, 4, 1000), rng.integers(0, 1 << 14, 1000), rng.integers(0, 1 << 14, 1000)
    es, et = dst_screen(x, y), dst_train(b, l, r)
    assert np.array_equal(es, v.dst_screen(x, y)) and np.array_equal(et, v.dst_train(b, l, r))
    for i in range(1000):
        assert decode(es[i]) == (0, x[i], y[i], 0)
        assert decode(et[i]) == (1, b[i], l[i], r[i])


def test_composite_closed_forms():
    W, H, S = 4, 3, 5
    bf = np.zeros((H, W, 4), np.float32)
    bf[1, 2] = [0.5, 1.0, 2.0, 4.0]              # bias rgb, factor r
    fgb = np.zeros((H, W, 2), np.float32)
    fgb[1, 2] = [0.25, 8.0]
    batches = np.zeros((4 * S, 17), np.float32)
    batches[2 * S + 1:2 * S + 4, 14:17] = 1.0
    factors = np.full((4 * S, 3), 2.0, np.float32)
    rgb = np.array([[1.0, 2.0, 0.5], [0.25, 0.5, 1.0], [9.0, 9.0, 9.0], [3.0, 3.0, 3.0]], np.float32)
    dst = np.array([dst_screen(2, 1), dst_train(2, 1, 3), INVALID, dst_screen(7, 0)], np.uint32)
    obf, obt = composite(rgb, dst, 4, bf, fgb, W, H, batches, factors, S)
    assert np.array_equal(obf[1, 2], [0.5 + 4.0 * 1.0, 1.0 + 0.25 * 2.0, 2.0 + 8.0 * 0.5, 0.0])
    assert np.array_equal(obf[0], bf[0]) and np.array_equal(obf[2], bf[2])     # x = 7 is off the image
    assert np.array_equal(obt[2 * S + 1:2 * S + 4, 14:17], np.tile([1.5, 2.0, 3.0], (3, 1)))
    changed = np.any(obt != batches, axis=1)
    assert changed.sum() == 3


def test_train_ranges_are_clipped_to_the_batch():
    S = 4
    batches = np.zeros((4 * S, 17), np.float32)
    factors = np.ones((4 * S, 3), np.float32)
    rgb = np.ones((2, 3), np.float32)
    dst = np.array([dst_train(1, 2, 100), dst_train(3, 5, 2)], np.uint32)   # r past the batch; l > r
    _, obt = composite(rgb, dst, 2, None, None, 0, 0, batches, factors, S)
    assert np.array_equal(np.nonzero(np.any(obt != 0, axis=1))[0], [1 * S + 2, 1 * S + 3])


def test_library_codec_matches_the_oracle():
    lib = v.load_library()
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y = (int(t) for t in rng.integers(0, 1 << 15, 2))
        b, l, r = int(rng.integers(0, 4)), int(rng.integers(0, 1 << 14)), int(rng.integers(0, 1 << 14))
        assert lib.nrc_dst_screen(x, y) == int(dst_screen(x, y))
        assert lib.nrc_dst_train(b, l, r) == int(dst_train(b, l, r))
