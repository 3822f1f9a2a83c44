"""The allocation-free gradient entry point (nrc_mlp_gradient_f16_noalloc) against
nrc_mlp_gradient_f16, through the C ABI: the same kernels, so the same bits, on two
calls in a row (the second smaller, so the scratch holds the first call's rows and
fragments), on both network depths, and a clear capacity error past its limit."""
import ctypes as C

import numpy as np
import pytest

import vknrc_amd as v

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"


def case(rng, hidden, n):
    wc = v.weight_count(hidden)
    w = torch.from_numpy(rng.normal(0, 0.17, wc).astype(np.float16)).to(DEV)
    x = torch.from_numpy(rng.uniform(-1, 1, (n, 64)).astype(np.float16)).to(DEV)
    t = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)).to(DEV)
    return w, x, t


def both(hidden, w, x, t, n):
    wc = v.weight_count(hidden)
    a = torch.full((wc,), 7.0, dtype=torch.float32, device=DEV)
    b = torch.full((wc,), -7.0, dtype=torch.float32, device=DEV)
    v.mlp_gradient_f16_noalloc(w, x, t, n, a, hidden_layers=hidden)
    v.mlp_gradient_f16(w, x, t, n, b, hidden_layers=hidden)
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_gradient_equals_allocating_gradient_on_consecutive_calls():
    rng = np.random.default_rng(501)
    for hidden, n in ((5, 16384), (5, 3000), (4, 1000), (5, 64)):
        w, x, t = case(rng, hidden, n)
        a, b = both(hidden, w, x, t, n)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (hidden, n)


def test_noalloc_gradient_rejects_more_records_than_its_scratch_holds():
    rng = np.random.default_rng(502)
    w, x, t = case(rng, 5, 16385)
    out = torch.zeros(v.weight_count(5), dtype=torch.float32, device=DEV)
    with pytest.raises(v.NRCError) as e:
        v.mlp_gradient_f16_noalloc(w, x, t, 16385, out)
    assert e.value.status == 5  # NRC_ERROR_CAPACITY
    # the rejected call launched nothing; the next call still works
    a, b = both(5, w, x, t, 16384)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
