"""The serving loop on the GPU: a session's result is the same alone and in a mix of sessions
that arrive and leave around it, and admission keeps the HBM budget."""
import numpy as np
import pytest

import vknrc_amd as v
from vknrc_amd import workloads as wl
from vknrc_amd.serving import Server, SessionSpec, session_bytes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def pools():
    rng = np.random.default_rng(61)
    q = torch.from_numpy(wl.make_queries(80000, rng)).cuda()
    r = torch.from_numpy(wl.make_train_records(8 * 16384, rng, wl.GaussianLights(62))).cuda()
    return q, r


def make_config(spec):
    return v.NRCConfig(encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE, loss_scale=128.0,
                       dynamic_loss_scale=True, seed=1000 + spec.seed)


def serve(pools, specs, path, watch, budget=1 << 36, max_sessions=256):
    q, r = pools
    srv = Server(torch, q, r, make_config, budget_bytes=budget, max_sessions=max_sessions, path=path)
    got = {}

    def finish(spec, cache, rgb):
        if spec.sid in watch:
            torch.cuda.synchronize()
            got[spec.sid] = (cache.get_state()["weight"].copy(), rgb.cpu().numpy().copy())

    srv.on_finish = finish
    rep = srv.run_ticks(specs)
    return got, rep, srv


@pytest.mark.parametrize("path", [v.PATH_FUSED_F16, v.PATH_FUSED_F8])
def test_a_session_is_the_same_alone_and_in_the_mix(pools, path):
    x = SessionSpec(0, 0.0, 50000, 12, 5)
    alone, rep1, _ = serve(pools, [SessionSpec(0, 0.0, 50000, 12, 5)], path, {0})
    mix = [SessionSpec(1, 0.0, 30000, 5, 11), SessionSpec(0, 2.0, 50000, 12, 5),
           SessionSpec(2, 3.0, 70001, 20, 12), SessionSpec(3, 6.0, 1000, 4, 13),
           SessionSpec(4, 9.0, 65536, 3, 14)]
    mixed, rep2, _ = serve(pools, mix, path, {0})
    assert rep1["ticks"] == 12 and rep2["peak_sessions"] >= 3
    w1, y1 = alone[x.sid]
    w2, y2 = mixed[x.sid]
    assert np.array_equal(w1, w2)
    assert np.array_equal(y1, y2)
    assert np.isfinite(y1).all()


def test_admission_keeps_the_budget(pools):
    probe = make_config(SessionSpec(0, 0, 1, 1, 0))
    cache = v.cache_footprint(probe)
    one = session_bytes(SessionSpec(0, 0, 40000, 1, 0), cache)
    specs = [SessionSpec(i, float(i // 3), 40000, 4, 20 + i) for i in range(9)]
    specs.append(SessionSpec(9, 0.0, 10 ** 9, 2, 99))      # can never fit
    _, rep, srv = serve(pools, specs, v.PATH_FUSED_F16, set(), budget=2 * one + one // 2)
    assert rep["peak_sessions"] == 2
    assert rep["peak_bytes"] <= 2 * one + one // 2
    assert rep["rejected"] == 1 and rep["sessions"] == 9
    assert rep["frames"] == 9 * 4
    assert rep["frame_ms_p50"] > 0 and rep["queries"] == 9 * 4 * 40000
