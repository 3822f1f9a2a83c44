"""Grouped multi-cache launches and HIP-graph frames, through the C ABI.

A group runs many independent caches in shared launches (nrc_group_*); each
cache's outputs and optimizer state must be bit-identical to what the
single-cache calls give it alone, with uneven query and batch sizes, device
counts below the maximum, mixed network kinds, and more caches than one launch
takes.  A captured group frame (nrc_group_frame_graph) must equal the eager
frames on every replay.
"""
import numpy as np
import pytest

import vknrc_amd as v
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"

KINDS = [dict(encoding=v.ENCODING_SINCOS6, loss_scale=128.0, dynamic_loss_scale=True),
         dict(encoding=v.ENCODING_TRI12),
         dict(encoding=v.ENCODING_UV_SINCOS16, hidden_layers=4, loss=v.LOSS_L2, init=v.INIT_XAVIER_UNIFORM,
              learning_rate=1e-2, beta2=0.99),
         dict(encoding=v.ENCODING_SINCOS6, loss=v.LOSS_L2, use_ema_weights=True)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def build(n, queries, batches, seed0):
    """n caches (two identical sets come from the same seeds) with their inputs."""
    rng = np.random.default_rng(seed0)
    caches, entries = [], []
    for i in range(n):
        kind = KINDS[i % len(KINDS)]
        st = v.NRCState(v.NRCConfig(seed=seed0 * 100 + i, **kind))
        nq, nb = queries[i % len(queries)], batches[i % len(batches)]
        q = wl.make_queries(nq, rng)
        r = wl.make_train_records(4 * nb, rng, wl.GaussianLights(seed0 + i))
        if kind["encoding"] == v.ENCODING_UV_SINCOS16:
            q[:, 0:2] = rng.uniform(0, 1, (nq, 2))
            r[:, 0:2] = rng.uniform(0, 1, (4 * nb, 2))
        qc = np.array([max(nq - (i * 37) % 500, 0)], np.int32)      # device counts below the maximum
        bc = np.array([nb - (i * 13 + b) % 300 for b in range(4)], np.int32)
        caches.append(st)
        entries.append(dict(cache=st, queries=dev(q), query_count=dev(qc), max_queries=nq,
                            rgb=torch.zeros((nq, 3), dtype=torch.float32, device=DEV),
                            batches=dev(r), batch_counts=dev(bc), max_batch_records=nb))
    return caches, entries


def single_frame(e):
    e["cache"].frame(e["queries"], e["rgb"], e["batches"], query_count=e["query_count"],
                     batch_counts=e["batch_counts"], max_queries=e["max_queries"],
                     max_batch_records=e["max_batch_records"])


def assert_same(ea, eb):
    torch.cuda.synchronize()
    for a, b in zip(ea, eb):
        assert torch.equal(a["rgb"], b["rgb"])
        sa, sb = a["cache"].get_state(), b["cache"].get_state()
        for k in ("weight", "m", "v", "ema_weight", "weights_f16", "use_weights_f16"):
            assert np.array_equal(sa[k], sb[k]), k
        assert sa["state"] == sb["state"]


def test_group_frames_equal_single_cache_frames():
    queries, batches = [70001, 5000, 130000, 64, 20000], [16384, 4096, 8192, 1000, 16384, 300]
    n = 18  # more caches than one grouped launch takes
    _, grouped = build(n, queries, batches, 7)
    _, alone = build(n, queries, batches, 7)
    for _ in range(2):
        v.group_frame(grouped)
        for e in alone:
            single_frame(e)
        assert_same(grouped, alone)


def test_group_plain_path_equals_single_calls():
    _, grouped = build(3, [3000], [2048], 11)
    _, alone = build(3, [3000], [2048], 11)
    v.group_frame(grouped, path=v.PATH_PLAIN_F32)
    for e in alone:
        e["cache"].frame(e["queries"], e["rgb"], e["batches"], query_count=e["query_count"],
                         batch_counts=e["batch_counts"], max_queries=e["max_queries"],
                         max_batch_records=e["max_batch_records"], path=v.PATH_PLAIN_F32)
    assert_same(grouped, alone)


def test_graph_replays_equal_eager_frames():
    queries, batches = [40000, 9000, 100000], [16384, 4096, 8192]
    caches, graphed = build(5, queries, batches, 21)
    _, eager = build(5, queries, batches, 21)
    g = v.GroupGraph(graphed)
    rng = np.random.default_rng(3)
    for f in range(4):
        # new device counts each frame, written the same into both sets
        for a, b in zip(graphed, eager):
            qc = np.array([rng.integers(1, a["max_queries"] + 1)], np.int32)
            bc = rng.integers(0, a["max_batch_records"] + 1, 4).astype(np.int32)
            a["query_count"].copy_(dev(qc))
            b["query_count"].copy_(dev(qc))
            a["batch_counts"].copy_(dev(bc))
            b["batch_counts"].copy_(dev(bc))
        torch.cuda.synchronize()
        g.launch()
        v.group_frame(eager)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device=DEV).sum().item()  # unrelated eager work between replays
        assert_same(graphed, eager)
    # a step outside the graph changes the state buffer a replay would start from
    e0 = graphed[0]
    e0["cache"].train(e0["batches"][:1000], path=v.PATH_FUSED_F16)
    with pytest.raises(v.NRCError):
        g.launch()
    g.close()


def test_group_rejects_a_cache_twice():
    caches, entries = build(2, [1000], [1000], 31)
    with pytest.raises(v.NRCError):
        v.group_train([entries[0], entries[1], entries[0]], 0)
    # FP8 inference updates each cache's activation history and scales, so inference
    # takes distinct caches too; a rejected frame launches nothing (its inference included)
    for path in (v.PATH_FUSED_F16, v.PATH_FUSED_F8):
        with pytest.raises(v.NRCError):
            v.group_infer([entries[0], entries[0]], path=path)
    out0 = entries[0]["rgb"].clone()
    before = caches[0].get_state()
    with pytest.raises(v.NRCError):
        v.group_frame([entries[0], entries[1], entries[0]])
    torch.cuda.synchronize()
    assert torch.equal(entries[0]["rgb"], out0)
    after = caches[0].get_state()
    assert np.array_equal(before["weight"], after["weight"]) and after["state"]["t"] == before["state"]["t"]


@pytest.mark.parametrize("chunks", [2, 8, 16])
def test_grad_chunks_give_the_same_bits_alone_and_grouped(chunks):
    """A non-default gradient chunking is a property of the cache: grouped and single frames still
    agree bit for bit.  Against the default chunking only the FP32 summation order of dW differs;
    Adam's normalised step can turn that noise into up to about one learning-rate step on weights
    whose gradient is near zero, so the weights are compared as a whole (relative L2)."""
    queries, batches = [20000, 3000], [16384, 8192, 4096, 1000]
    n = 6
    _, grouped = build(n, queries, batches, 17)
    _, alone = build(n, queries, batches, 17)
    _, default = build(n, queries, batches, 17)
    for e in grouped + alone:
        e["cache"].set_grad_chunks(chunks)
    for _ in range(2):
        v.group_frame(grouped)
        v.group_frame(default)
        for e in alone:
            single_frame(e)
    assert_same(grouped, alone)
    for a, d in zip(grouped, default):
        wa, wd = a["cache"].get_state()["weight"], d["cache"].get_state()["weight"]
        assert np.linalg.norm(wa - wd) <= 2e-3 * np.linalg.norm(wd)


def test_grad_chunks_are_validated():
    st = v.NRCState(v.NRCConfig())
    with pytest.raises(v.NRCError):
        st.set_grad_chunks(17)
    st.set_grad_chunks(0)
    with pytest.raises(v.NRCError):
        v.set_infer_workgroups_per_cu(0)
    v.set_infer_workgroups_per_cu(2)
