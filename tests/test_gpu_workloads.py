"""The 16-cache and the served workloads at their own sizes, through the C ABI.

16 caches: the seeded workload of the benchmark (queries log-uniform in
[65,536, 2,073,600] a cache, batches in {4096, 8192, 16384}, relative-L2 with a
dynamic loss scale), autotuned gradient chunking and inference grid, run as
grouped frames, as single-cache frames and as replays of a captured graph: every
cache's radiance and optimizer state agree bit for bit, and a 4,096-query sample
of each cache's inference is within the forward bar of the FP64 oracle.

Served sessions: full-resolution sessions (720p to 2160p) with Poisson arrivals of
mean 20/s on a 60 frames/s clock, admitted by an HBM budget that makes some wait;
watched sessions give the same weights and radiance alone and in the mix, and
admission keeps the budget.
"""
import numpy as np
import pytest

import vknrc_amd as v
from oracle.oracle import Oracle
from vknrc_amd import autotune
from vknrc_amd import workloads as wl
from vknrc_amd.serving import Server, SessionSpec, poisson_sessions, session_bytes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"
FWD_REL_TOL = 1e-2


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def radiance_config(seed):
    return v.NRCConfig(seed=seed, encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE,
                       loss_scale=128.0, dynamic_loss_scale=True)


def caches16_set(spec, pool):
    out = []
    for c in spec:
        out.append(dict(cache=v.NRCState(radiance_config(c["seed"])), queries=pool[:c["queries"]],
                        max_queries=c["queries"],
                        rgb=torch.zeros((c["queries"], 3), dtype=torch.float32, device=DEV),
                        batches=dev(c["records"]), max_batch_records=c["batch"]))
    return out


def assert_same(xs, ys):
    torch.cuda.synchronize()
    for a, b in zip(xs, ys):
        assert torch.equal(a["rgb"], b["rgb"])
        sa, sb = a["cache"].get_state(), b["cache"].get_state()
        for k in ("weight", "m", "v", "ema_weight", "weights_f16", "use_weights_f16"):
            assert np.array_equal(sa[k], sb[k]), k
        assert sa["state"] == sb["state"]


def test_caches16_grouped_equals_alone_equals_graph_and_the_oracle():
    spec = wl.caches16()
    assert len(spec) == 16 and max(c["queries"] for c in spec) <= 2073600
    rng = np.random.default_rng(wl.workload_seed(wl.CACHES16, 1))
    qhost = wl.make_queries(2073600, rng)
    pool = dev(qhost)
    grouped, alone, graphed = caches16_set(spec, pool), caches16_set(spec, pool), caches16_set(spec, pool)
    tuned = autotune.tune_grad_chunks(torch, grouped)
    autotune.tune_infer_workgroups(torch, grouped)
    assert len(tuned["per_cache"]) == 16
    for i, e in enumerate(alone):
        e["cache"].set_grad_chunks(tuned["per_cache"][i])
        graphed[i]["cache"].set_grad_chunks(tuned["per_cache"][i])
    graph = v.GroupGraph(graphed)
    for _ in range(3):
        v.group_frame(grouped)
        for e in alone:
            e["cache"].frame(e["queries"], e["rgb"], e["batches"], max_queries=e["max_queries"],
                             max_batch_records=e["max_batch_records"])
        graph.launch()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device=DEV).sum().item()   # unrelated eager work between replays
        assert_same(grouped, alone)
        assert_same(grouped, graphed)
    graph.close()
    for e in grouped:
        assert e["cache"].get_state()["state"]["t"] + e["cache"].get_state()["state"]["skipped_steps"] == 12
    # inference with the current use-weights against the FP64 oracle on the same fp16 weights
    v.group_infer(grouped)
    torch.cuda.synchronize()
    o64 = Oracle("f64")
    for e in grouped:
        n = e["max_queries"]
        idx = rng.choice(n, 4096, replace=False)
        y = e["rgb"].cpu().numpy()[idx].astype(np.float64)
        w16 = e["cache"].get_state()["use_weights_f16"].astype(np.float64)
        ref = o64.infer(w16, 5, qhost[idx], v.ENCODING_SINCOS6, round_f16=True, clamp0=True)
        rel = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        assert rel <= FWD_REL_TOL, (n, rel)
    for e in grouped + alone + graphed:
        e["cache"].close()


FPS = 60


def full_resolution_sessions(n, frames, seed):
    """Poisson(20/s) sessions at 720p..2160p; arrivals converted to ticks of a 60 frames/s clock."""
    rng = np.random.default_rng(seed)
    specs = poisson_sessions(n, 20.0, rng, frames=frames, seed0=seed % 1000)
    for s in specs:
        s.arrival = float(np.floor(s.arrival * FPS))
    return specs


def test_served_full_resolution_sessions_alone_and_in_the_mix():
    rng = np.random.default_rng(wl.workload_seed(wl.SERVED, 5))
    qpool = dev(wl.make_queries(3840 * 2160, rng))
    rpool = dev(wl.make_train_records(6 * 16384, rng, wl.GaussianLights(wl.workload_seed(wl.SERVED, 6))))
    specs = full_resolution_sessions(10, 12, wl.workload_seed(wl.SERVED, 7) % 100000)
    sizes = {s.queries for s in specs}
    assert len(sizes) >= 2 and all(s.queries >= 1280 * 720 for s in specs)
    cache_b = v.cache_footprint(radiance_config(0))
    need = [session_bytes(s, cache_b) for s in specs]
    budget = max(need) + sorted(need)[len(need) // 2]     # room for about two sessions: later ones wait

    def make_config(spec):
        return radiance_config(7000 + spec.seed)

    watch = {specs[0].sid, specs[len(specs) // 2].sid, specs[-1].sid}

    def serve(run_specs):
        srv = Server(torch, qpool, rpool, make_config, budget_bytes=budget, max_sessions=256)
        got = {}

        def finish(spec, cache, rgb):
            if spec.sid in watch:
                torch.cuda.synchronize()
                got[spec.sid] = (cache.get_state()["weight"].copy(), rgb.cpu().numpy().copy())

        srv.on_finish = finish
        return got, srv.run_ticks(run_specs), srv

    mixed, rep, srv = serve([SessionSpec(**vars(s)) for s in specs])
    assert rep["rejected"] == 0 and rep["sessions"] == len(specs)
    assert rep["frames"] == sum(s.frames for s in specs)
    assert rep["peak_bytes"] <= budget and rep["peak_sessions"] >= 2
    assert max(srv.sched.waits) > 0                          # the budget made some session wait
    for sid in watch:
        s = next(x for x in specs if x.sid == sid)
        alone, _, _ = serve([SessionSpec(s.sid, 0.0, s.queries, s.frames, s.seed, s.width, s.height)])
        w1, y1 = alone[sid]
        w2, y2 = mixed[sid]
        assert np.array_equal(w1, w2), sid
        assert np.array_equal(y1, y2), sid
        assert np.isfinite(y1).all() and y1.shape[0] == s.queries
