"""The 16-bit identities behind the fused gradient kernel's ReLU backward
(relu_bwd2 / act_masks in vknrc_amd/csrc/nrc_fused.hip), checked over every FP16
bit pattern against IEEE semantics. The reference semantics are NN_nv.glsl:198-220:
an inactive unit (activation not > 0) or a NaN delta gives 0.

- act_masks: sat16(0 - a) >> 15 (arithmetic) is all ones exactly when the FP16
  activation a is > 0, for every value a ReLU can output (no NaN; +-0 included).
- relu_bwd2's NaN test: (0x7C00 - (h & 0x7FFF)) >> 15 (int16, no overflow) is all
  ones exactly when the FP16 value h is NaN.
"""
import numpy as np

ALL = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)


def _as_f16(bits):
    return bits.view(np.float16)


def test_activation_mask_is_greater_than_zero():
    a = _as_f16(ALL)
    relu_outputs = ~np.isnan(a) & ((a >= 0) | (ALL == 0x8000))   # max(x, 0): +-0, positive, +Inf
    t = np.clip(0 - ALL.view(np.int16).astype(np.int32), -32768, 32767).astype(np.int16)
    mask = (t >> 15).view(np.uint16)
    want = np.where(a > 0, 0xFFFF, 0).astype(np.uint16)
    assert np.array_equal(mask[relu_outputs], want[relu_outputs])
    # the cases that matter: -0 and +0 are inactive, the smallest subnormal and +Inf active
    assert mask[0x0000] == 0 and mask[0x8000] == 0 and mask[0x0001] == 0xFFFF and mask[0x7C00] == 0xFFFF
    # a wrapping (non-saturating) subtraction would get -0 wrong: the clamp is needed
    wrapped = ((0 - ALL.view(np.int16).astype(np.int32)).astype(np.int16) >> 15).view(np.uint16)
    assert wrapped[0x8000] == 0xFFFF


def test_nan_mask_is_isnan():
    h = _as_f16(ALL)
    mag = (ALL & 0x7FFF).astype(np.int16)
    nan = ((np.int16(0x7C00) - mag) >> 15).view(np.uint16)
    assert np.array_equal(nan == 0xFFFF, np.isnan(h))
    assert np.all((nan == 0) | (nan == 0xFFFF))
    # +-Inf pass (an active unit's infinite delta is kept), NaN of either sign is masked
    assert nan[0x7C00] == 0 and nan[0xFC00] == 0 and nan[0x7E00] == 0xFFFF and nan[0xFE00] == 0xFFFF


def test_relu_bwd_pair_matches_the_reference_rule():
    """(fp16)v & act_mask & ~nan_mask equals (act > 0 && v == v) ? (fp16)v : +0 on a grid of deltas
    (finite, beyond the FP16 range, +-Inf, NaN, +-0) against every ReLU output."""
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(0, 1e3, 4000), [np.inf, -np.inf, np.nan, 0.0, -0.0, 7e4, -7e4, 65504.0,
                                                    6e-8, -6e-8]]).astype(np.float32)
    acts = ALL[~np.isnan(_as_f16(ALL)) & ((_as_f16(ALL) >= 0) | (ALL == 0x8000))]
    a = acts[rng.integers(0, len(acts), v.shape[0])]
    a[:16] = [0x0000, 0x8000, 0x0001, 0x7C00, 0x3C00] + [0x0000] * 11
    with np.errstate(over="ignore", invalid="ignore"):   # values beyond the FP16 range round to +-Inf
        h = v.astype(np.float16).view(np.uint16)
    t = np.clip(0 - a.view(np.int16).astype(np.int32), -32768, 32767).astype(np.int16)
    amask = (t >> 15).view(np.uint16)
    nan = ((np.int16(0x7C00) - (h & 0x7FFF).astype(np.int16)) >> 15).view(np.uint16)
    got = h & amask & ~nan
    act = _as_f16(a).astype(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        want = np.where((act > 0) & (v == v), v, np.float32(0.0)).astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
