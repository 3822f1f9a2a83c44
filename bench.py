"""NRC hot-path benchmark on one MI355X.

Workload (single cache, FP16 MFMA vs FP32 plain): per frame, inference over
2,073,600 query records (position U[-1,1]^3, direction and normal uniform on the
sphere, roughness, albedos; SINCOS6 encoding) followed by 4 Adam training steps
of 16,384 records each (relative-L2 loss, dynamic loss scale starting at 128),
the network 64 -> 5 x 64 -> 3.  A step is one frame.  Inputs are resident in HBM
before the timed region.

Prints one JSON line.  value = queries per second of whole frames (inference
plus the 4 training steps), measured between two device synchronisations.

The line also carries the 16-cache workload (`group16`): 16 independent caches
with distinct seeded weights and light sets, per-frame query counts
log-uniform in [65,536, 2,073,600] and training batches uniform in
{4096, 8192, 16384}, timed as sequential per-cache frames, as one grouped frame
and as a replayed HIP graph of the grouped frame; the FP8 inference (`fp8`); the
inference epilogue over one 1920x1080 frame (`epilogue`: screen composite and
training-target back-fill, with its HBM rate); and the served workload (`served`): render sessions arriving as a Poisson
process of 20/s at 720p-2160p, 300 frames of inference + one training step
each, admitted by an HBM budget and batched continuously into grouped launches
(p50/p99 frame latency, admission waits, aggregate queries/s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

QUERIES = 2073600
BATCH = 16384
BATCHES = 4
INFER_FLOP_PER_QUERY = 41344      # 2 * (5 * 64 * 64 + 3 * 64)
TRAIN_FLOP_PER_SAMPLE = 115840    # forward 41,344 + dW 41,344 + dA 33,152
QUERY_BYTES = 56 + 12             # 14-float record in, RGB out
PEAK_F16_TFLOPS = 2500.0
PEAK_F8_TFLOPS = 5000.0
PEAK_HBM_GBS = 8000.0
# HBM bytes per fused_infer_pipe_kernel launch at this workload: rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE
# (x1.0) KiB per dispatch, scaled by a 256 MiB copy kernel measured in the same runs (bench.py
# --copy-calibration); the commands and readings are in profiles/fused_infer_2073600q_hbm_traffic.md
INFER_TRAFFIC_BYTES = (57005.4 * 2.0 + 25325.5) * 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _rate(fn, items: int, budget_s: float) -> tuple[float, int]:
    """items/s of fn() repeated for about budget_s seconds (at least twice), and the repetitions."""
    fn()
    t0 = time.perf_counter()
    reps = 0
    while True:
        fn()
        reps += 1
        if time.perf_counter() - t0 > budget_s and reps >= 2:
            break
    return items * reps / (time.perf_counter() - t0), reps


def cpu_baseline(queries: np.ndarray, records: np.ndarray, budget_s: float) -> dict:
    """CPU baselines of the same MLP on the host cores, on bounded samples:
    the Eigen MLP (oracle/eigen_mlp.cpp on the Eigen 3.4.0 the reference vendors; the reference test's
    forward semantics, test/main.cpp:11-27) in FP32 on all cores and on one thread and in _Float16 on one
    thread, its config-1 training and inference rates, and the FP32 C port of the whole path
    (oracle/nrc_oracle.c REAL=float, encoding included)."""
    from oracle import eigen
    from oracle.oracle import ENC_SINCOS6, ENC_UV_SINCOS16, Oracle
    from vknrc_amd import nrc as vn
    from vknrc_amd import workloads as wl
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    port = Oracle("f32")
    w = vn.init_weights(vn.NRCConfig(seed=1, encoding=vn.ENCODING_SINCOS6)).astype(np.float32)
    n = 16384
    q = queries[:n]
    res = {"cores": cores, "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    share = budget_s / 7.0
    if eigen.available():
        mlp = eigen.EigenMLP()
        x = port.encode(q, ENC_SINCOS6).astype(np.float32)
        mlp.set_threads(cores)
        r_all, k_all = _rate(lambda: mlp.forward_f32(w, 5, x), n, share)
        mlp.set_threads(1)
        r_one, k_one = _rate(lambda: mlp.forward_f32(w, 5, x), n, share)
        n16 = 2048
        w16, x16 = w.astype(np.float16), x[:n16].astype(np.float16)
        r_16, k_16 = _rate(lambda: mlp.forward_f16(w16, 5, x16), n16, share)
        # config 1: the image MLP (UV sin/cos 16 octaves, 64 -> 4 x 64 -> 3, L2, Adam 1e-2 / 0.9 / 0.99)
        rng = np.random.default_rng(wl.workload_seed(wl.IMAGE_MLP, 9))
        img = wl.make_image(wl.workload_seed(wl.IMAGE_MLP))
        rec = wl.make_image_records(n, rng, img)
        xi = port.encode(rec, ENC_UV_SINCOS16).astype(np.float32)
        ti = np.ascontiguousarray(rec[:, 14:17])
        wi = vn.init_weights(vn.NRCConfig(hidden_layers=4, encoding=vn.ENCODING_UV_SINCOS16,
                                          init=vn.INIT_XAVIER_UNIFORM, seed=3)).astype(np.float32)
        mi, vi = np.zeros_like(wi), np.zeros_like(wi)
        mlp.set_threads(cores)
        step = [0]

        def train():
            step[0] += 1
            mlp.train_step_l2_f32(wi, mi, vi, 4, xi, ti, 1e-2, 0.9, 0.99, 1e-8, 0.9 ** step[0], 0.99 ** step[0])

        r_tr, k_tr = _rate(train, n, share)
        uv = wl.make_image_records(65536, rng, img)
        xu = port.encode(uv, ENC_UV_SINCOS16).astype(np.float32)
        r_inf1, k_inf1 = _rate(lambda: mlp.forward_f32(wi, 4, xu), 65536, share)
        res.update({
            "value": r_all, "unit": "queries/s", "kind": "port",
            "sample": f"Eigen 3.4.0 FP32 forward of the reference test MLP (test/main.cpp:11-27 semantics), "
                      f"{k_all} x {n} SINCOS6-encoded config-2 queries on {cores} threads (encoding excluded)",
            "eigen_f32_1thread_queries_per_s": r_one,
            "eigen_f16_1thread_queries_per_s": r_16,
            "eigen_samples": f"1 thread: {k_one} x {n} FP32, {k_16} x {n16} _Float16 (every product and sum "
                             f"rounded to FP16, as the reference test)",
            "config1_train_samples_per_s": r_tr,
            "config1_infer_queries_per_s": r_inf1,
            "config1_samples": f"{k_tr} Adam steps of {n} image samples, {k_inf1} x 65,536 uv inferences, "
                               f"{cores} threads, Eigen FP32 (encoding excluded)",
        })
    port.set_threads(cores)
    p_all, pk_all = _rate(lambda: port.infer(w, 5, q, ENC_SINCOS6, clamp0=True), n, share)
    port.set_threads(1)
    p_one, pk_one = _rate(lambda: port.infer(w, 5, q[:4096], ENC_SINCOS6, clamp0=True), 4096, share)
    port.set_threads(cores)
    p_tr, pk_tr = _rate(lambda: port.gradient(w, 5, records[:BATCH], ENC_SINCOS6, 0), BATCH, share)
    res.update({"port_queries_per_s": p_all, "port_1thread_queries_per_s": p_one,
                "port_train_samples_per_s": p_tr,
                "port_sample": f"FP32 C port (oracle/nrc_oracle.c REAL=float, OpenMP) with encoding: {pk_all} x {n} "
                               f"queries on {cores} threads, {pk_one} x 4096 on 1 thread, {pk_tr} x {BATCH}-record "
                               f"forward+backward on {cores} threads"})
    if "value" not in res:   # no Eigen library on this host: the C port is the baseline
        res.update({"value": p_all, "unit": "queries/s", "kind": "port", "sample": res["port_sample"]})
    return res


def epilogue_workload(torch, v, rgb, drec, stream, reps: int) -> dict:
    """nrc_composite over one 1920x1080 frame's inference output: the first 2,073,600 - 65,536 queries
    composite into their pixels (in pixel order), the last 65,536 back-fill one training record each
    (4 batches of 16,384)."""
    W, H = 1920, 1080
    n = rgb.shape[0]
    T = BATCHES * BATCH
    S = n - T
    p = np.arange(S, dtype=np.int64)
    j = np.arange(T, dtype=np.int64)
    dst = np.concatenate([v.dst_screen(p % W, p // W), v.dst_train(j // BATCH, j % BATCH, j % BATCH)])
    ddst = torch.from_numpy(dst).to("cuda")
    g = torch.Generator(device="cuda").manual_seed(1)
    bf = torch.rand((H, W, 4), device="cuda", generator=g)
    fgb = torch.rand((H, W, 2), device="cuda", generator=g)
    batches = drec[:T].clone()
    factors = torch.rand((T, 3), device="cuda", generator=g)

    def run():
        v.composite(rgb, ddst, n, bias_factor_r=bf, factor_gb=fgb, width=W, height=H, batches=batches,
                    train_factors=factors, batch_stride=BATCH, stream=stream)

    for _ in range(3):
        run()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        run()
        b.record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    # every query reads its code (4 B) and radiance (12 B); a pixel reads and writes bias_factor_r (16 + 16 B)
    # and reads factor_gb (8 B); a training record reads and writes its target (12 + 12 B) and reads its factor
    nbytes = n * 16 + S * 40 + T * 36
    return {"kernel": "composite_kernel", "ms_avg": ms, "queries": n, "screen_queries": S, "train_records": T,
            "bytes": nbytes, "GB_s": nbytes / (ms / 1e3) / 1e9, "frac_hbm": nbytes / (ms / 1e3) / 1e9 / PEAK_HBM_GBS}


def frontend_workload(torch, v, wl, reps: int) -> dict:
    """The record front-end over one 1920x1080 frame's records, on a synthetic scene of 65,536 triangles
    (8 instances, 32 materials, 8 RGBA8 sRGB textures of 256x256): 2,073,600 eval records unpacked
    into queries, 4 x 16,384 train records unpacked into training records, the epilogue over the wire
    records (2,008,064 screen codes in pixel order, 65,536 one-record back-fills) and the screen pass
    with accumulation.  Each kernel is timed alone with HIP events; GB/s count the record bytes each
    reads and writes (the scene gather, about 0.3 MB of triangles and 2 MB of textures, stays in L2)."""
    n = v.nrc
    W, H = 1920, 1080
    NQ, T = QUERIES, BATCHES * BATCH
    S = NQ - T
    s = wl.make_scene(wl.workload_seed(wl.SCENE, 1), primitives=65536, vertices=40000, instances=8, materials=32,
                      textures=8, texture_size=256)
    ds = n.DeviceScene(torch, s["vertices"], s["vertex_indices"], s["texcoords"], s["texcoord_indices"],
                       s["materials"], s["material_ids"], s["transforms"], s["textures"])
    rng = np.random.default_rng(wl.workload_seed(wl.SCENE, 2))
    rec = np.zeros(NQ, n.EVAL_RECORD)
    p = np.arange(S)
    rec["dst"][:S] = ((p % W) | ((p // W) << 15)) << 1
    j = np.arange(T)
    rec["dst"][S:] = ((((j // BATCH) | ((j % BATCH) << 2) | ((j % BATCH) << 16)) << 1) | 1).astype(np.uint32)
    rec["packed_input"] = wl.make_packed_inputs(s, NQ, rng)
    tr = np.zeros(T, n.WIRE_TRAIN_RECORD)
    tr["bias"] = rng.uniform(0, 1, (T, 3))
    tr["factor"] = rng.uniform(0, 1, (T, 3))
    tr["packed_input"] = wl.make_packed_inputs(s, T, rng)
    d_rec = torch.from_numpy(rec.view(np.uint8)).to("cuda")
    d_tr = torch.from_numpy(tr.view(np.uint8)).to("cuda")
    q = torch.zeros((NQ, 14), dtype=torch.float32, device="cuda")
    t17 = torch.zeros((T, 17), dtype=torch.float32, device="cuda")
    rgb = torch.rand((NQ, 3), device="cuda")
    bf = torch.rand((H, W, 4), device="cuda")
    fgb = torch.rand((H, W, 2), device="cuda")
    acc = torch.zeros((H, W, 4), device="cuda")
    out = torch.zeros((H, W, 4), device="cuda")
    stream = torch.cuda.current_stream()
    runs = {
        "unpack_eval": (lambda: v.unpack_eval_records(ds, d_rec, NQ, q, stream=stream), NQ * (20 + 56)),
        "unpack_train": (lambda: v.unpack_train_records(ds, d_tr, T, t17, stream=stream), T * (40 + 68)),
        "composite_records": (lambda: v.composite_records(rgb, d_rec, NQ, bias_factor_r=bf, factor_gb=fgb, width=W,
                                                          height=H, batches=d_tr, batch_stride=BATCH,
                                                          stream=stream), NQ * 16 + S * 40 + T * 36),
        "screen_accumulate_tonemap": (lambda: v.screen_tonemap(bf, W, H, out, accumulate_img=acc, accumulate=True,
                                                               accumulate_count=3, stream=stream), W * H * 64),
    }
    # the same eval records sorted by primitive: the most coherent order a renderer could hand over
    # (the order above, uniformly random primitives, is the least coherent)
    order = np.argsort(rec["packed_input"]["primitive_id"], kind="stable")
    d_rec_sorted = torch.from_numpy(rec[order].copy().view(np.uint8)).to("cuda")
    runs["unpack_eval_sorted"] = (lambda: v.unpack_eval_records(ds, d_rec_sorted, NQ, q, stream=stream), NQ * (20 + 56))
    # the same scene with its primitives packed one to a 64-byte record (nrc_scene_pack_primitives)
    dsp = n.DeviceScene(torch, s["vertices"], s["vertex_indices"], s["texcoords"], s["texcoord_indices"],
                        s["materials"], s["material_ids"], s["transforms"], s["textures"])
    dsp.pack_primitives(stream=stream)
    P = len(s["material_ids"])
    runs["pack_primitives"] = (lambda: dsp.pack_primitives(stream=stream), P * (24 + 4 + 64) + P * 3 * (12 + 8))
    runs["unpack_eval_packed"] = (lambda: v.unpack_eval_records(dsp, d_rec, NQ, q, stream=stream), NQ * (20 + 56))
    runs["unpack_eval_sorted_packed"] = (lambda: v.unpack_eval_records(dsp, d_rec_sorted, NQ, q, stream=stream),
                                         NQ * (20 + 56))
    runs["unpack_train_packed"] = (lambda: v.unpack_train_records(dsp, d_tr, T, t17, stream=stream), T * (40 + 68))
    res = {"scene": "65,536 triangles, 8 instances, 32 materials, 8 textures 256x256 RGBA8 sRGB; records on "
                    "uniformly random primitives (unpack_eval_sorted: the same records sorted by primitive; "
                    "*_packed: the scene's primitives packed one to a 64-byte record by nrc_scene_pack_primitives, "
                    "whose own time is pack_primitives)"}
    for name, (fn, nbytes) in runs.items():
        for _ in range(3):
            fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        res[name] = {"ms_avg": ms, "bytes": nbytes, "GB_s": nbytes / (ms / 1e3) / 1e9,
                     "frac_hbm": nbytes / (ms / 1e3) / 1e9 / PEAK_HBM_GBS}
    return res


def group_workload(torch, v, wl, q_host, frames: int) -> dict:
    """16 caches: sequential nrc_frame calls vs one nrc_group_frame vs a replayed graph."""
    spec = wl.caches16()
    n = len(spec)
    nq = np.array([c["queries"] for c in spec], np.int64)
    nb = np.array([c["batch"] for c in spec], np.int64)
    dq = torch.from_numpy(q_host).to("cuda")
    entries = []
    for c in spec:
        st = v.NRCState(v.NRCConfig(seed=c["seed"], encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE,
                                    loss_scale=128.0, dynamic_loss_scale=True))
        entries.append(dict(cache=st, queries=dq[:c["queries"]], max_queries=c["queries"],
                            rgb=torch.zeros((c["queries"], 3), dtype=torch.float32, device="cuda"),
                            batches=torch.from_numpy(c["records"]).to("cuda"), max_batch_records=c["batch"]))

    def sequential():
        for e in entries:
            e["cache"].frame(e["queries"], e["rgb"], e["batches"], max_queries=e["max_queries"],
                             max_batch_records=e["max_batch_records"])

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / frames

    out = {"sequential": timed(sequential)}     # default schedules: each cache's step fills the GPU alone
    from vknrc_amd.autotune import tune_grad_chunks, tune_infer_workgroups
    tuned_infer = tune_infer_workgroups(torch, entries)
    tuned_grad = tune_grad_chunks(torch, entries)
    out["grouped"] = timed(lambda: v.group_frame(entries))
    graph = v.GroupGraph(entries)
    out["graph"] = timed(graph.launch)
    graph.close()
    total_q = int(nq.sum())
    total_t = int(4 * nb.sum())
    return {"caches": n, "queries_per_frame": total_q, "train_samples_per_frame": total_t, "frames": frames,
            "sequential_ms_per_frame": out["sequential"], "grouped_ms_per_frame": out["grouped"],
            "graph_ms_per_frame": out["graph"], "graph_queries_per_s": total_q / (out["graph"] / 1e3),
            "grouped_vs_sequential": out["sequential"] / out["grouped"],
            "graph_vs_sequential": out["sequential"] / out["graph"],
            "autotuned": {"grad_chunks": tuned_grad["best"], "grad_chunks_per_cache": tuned_grad["per_cache"],
                          "grad_ms_per_step": {str(k): round(x, 4) for k, x in tuned_grad["ms_per_step"].items()},
                          "grad_ms_per_step_per_cache_tuned": round(tuned_grad["tuned_ms_per_step"], 4),
                          "infer_workgroups_per_cu": tuned_infer["best"],
                          "infer_ms": {str(k): round(x, 4) for k, x in tuned_infer["ms_per_launch"].items()}}}


def served_workload(torch, v, wl, dq, drec, sessions: int, budget_gib: float) -> dict:
    """Render sessions with Poisson arrivals (mean 20/s) at 720p/1080p/1440p/2160p (weights
    .3/.4/.2/.1), 300 frames each of full-resolution inference + one 16,384-record training step,
    admitted by an HBM budget and run by continuous batching into grouped launches."""
    from vknrc_amd.serving import Server, grouped_vs_sequential, poisson_sessions
    rng = np.random.default_rng(wl.workload_seed(wl.SERVED))
    specs = poisson_sessions(sessions, 20.0, rng, frames=300, seed0=wl.workload_seed(wl.SERVED, 1) % 100000)
    pool = torch.cat([dq] * 4)[:3840 * 2160]

    def make_config(spec):
        return v.NRCConfig(seed=spec.seed, encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE,
                           loss_scale=128.0, dynamic_loss_scale=True)

    srv = Server(torch, pool, drec, make_config, budget_bytes=int(budget_gib * (1 << 30)), max_sessions=256)
    # sessions are paced by a 60 frames/s clock: each lives 300 frames (5 s), so concurrency builds up
    # towards 20/s x 5 s until the HBM budget or the 256-session cap holds it
    rep = srv.run_realtime(specs, frame_interval=1.0 / 60.0)
    rep["arrival_rate_per_s"] = 20.0
    rep["hbm_budget_GiB"] = budget_gib
    # one frame of as many sessions as ran at once at the peak (same resolution mix), grouped launches
    # against per-session launches
    k = max(1, min(int(rep["peak_sessions"]), len(specs)))
    rep["vs_sequential"] = grouped_vs_sequential(torch, specs[:k], pool, drec, make_config)
    return rep


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-reps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=21.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--group-frames", type=int, default=30, help="timed frames of the 16-cache workload (0: skip)")
    ap.add_argument("--serve-sessions", type=int, default=120,
                    help="sessions of the served workload (0: skip)")
    ap.add_argument("--serve-budget-gib", type=float, default=24.0, help="HBM budget of the served workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after timing, run one elementwise copy (torch.neg) of 256 MiB fp32 -> 256 MiB, the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the NRC hot path is measured per GPU "
              "(caches shard across GPUs with no communication); use --gpus 1", file=sys.stderr)
        return 2

    import torch
    import vknrc_amd as v
    from vknrc_amd import workloads as wl

    dev = "cuda:0"
    torch.cuda.set_device(0)
    cfg = v.NRCConfig(seed=wl.workload_seed(wl.RADIANCE), encoding=v.ENCODING_SINCOS6, loss=v.LOSS_RELATIVE_L2_LUMINANCE,
                      loss_scale=128.0, dynamic_loss_scale=True)
    st = v.NRCState(cfg)
    rng = np.random.default_rng(wl.workload_seed(wl.RADIANCE, 1))
    log("generating synthetic data ...")
    q = wl.make_queries(QUERIES, rng)
    lights = wl.GaussianLights(wl.workload_seed(wl.RADIANCE, 2))
    pool_frames = 8
    recs = wl.make_train_records(pool_frames * BATCHES * BATCH, rng, lights)
    dq = torch.from_numpy(q).to(dev)
    drec = torch.from_numpy(recs).to(dev)
    out = torch.zeros((QUERIES, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    def frame(i):
        b = drec[(i % pool_frames) * BATCHES * BATCH:][:BATCHES * BATCH]
        st.frame(dq, out, b, max_batch_records=BATCH, path=v.PATH_FUSED_F16, stream=stream)

    log("warmup ...")
    for i in range(args.warmup):
        frame(i)
    torch.cuda.synchronize()

    K = args.steps
    log(f"timing {K} frames ...")
    fev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        fev[i][0].record(stream)
        frame(i)
        fev[i][1].record(stream)
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    frame_ms = np.array([a.elapsed_time(b) for a, b in fev])
    s = st.get_state()["state"]
    log(f"loss {s['last_loss']:.4f} scale {s['loss_scale']} t {s['t']} skipped {s['skipped_steps']}")

    # the inference kernel and the training steps alone (for the roofline and the split)
    iev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    for i in range(K):
        iev[i][0].record(stream)
        st.infer(dq, out, path=v.PATH_FUSED_F16, stream=stream)
        iev[i][1].record(stream)
    tev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    for i in range(K):
        b = drec[(i % pool_frames) * BATCHES * BATCH:][:BATCHES * BATCH]
        tev[i][0].record(stream)
        for j in range(BATCHES):
            st.train(b[j * BATCH:(j + 1) * BATCH], path=v.PATH_FUSED_F16, write_use_weights=(j == BATCHES - 1),
                     stream=stream)
        tev[i][1].record(stream)
    torch.cuda.synchronize()
    infer_ms = np.array([a.elapsed_time(b) for a, b in iev])
    train_ms = np.array([a.elapsed_time(b) for a, b in tev])

    # FP8 (e4m3) inference of the same queries with the same use weights: time and error against FP16
    st.infer(dq, out, path=v.PATH_FUSED_F16, stream=stream)
    out16 = out.clone()
    for _ in range(3):  # the activation-range history exists from the second call on
        st.infer(dq, out, path=v.PATH_FUSED_F8, stream=stream)
    f8ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    for i in range(K):
        f8ev[i][0].record(stream)
        st.infer(dq, out, path=v.PATH_FUSED_F8, stream=stream)
        f8ev[i][1].record(stream)
    torch.cuda.synchronize()
    f8_ms = np.array([a.elapsed_time(b) for a, b in f8ev])
    f8_err = float(torch.linalg.vector_norm((out - out16).double()) / torch.linalg.vector_norm(out16.double()))
    # both against an FP64 forward on the GPU (torch float64 matmuls on the library's FP32 features and the
    # FP32 use-weights, which equal the master weights after the last step wrote them) over a 65,536-query sample
    ns = 65536
    feats = torch.zeros((ns, 64), dtype=torch.float32, device=dev)
    v.encode(dq[:ns], ns, feats, encoding=v.ENCODING_SINCOS6, stream=stream)
    w64 = torch.from_numpy(st.get_state()["weight"].astype(np.float64)).to(dev)
    a = feats.double()
    for l in range(5):
        a = torch.relu(a @ w64[l * 4096:(l + 1) * 4096].reshape(64, 64).T)
    ref64 = torch.relu(a @ w64[5 * 4096:].reshape(3, 64).T)
    torch.cuda.synchronize()

    def err64(y):
        return float(torch.linalg.vector_norm(y[:ns].double() - ref64) / torch.linalg.vector_norm(ref64))
    f8_err64, f16_err64 = err64(out), err64(out16)

    # plain FP32 inference at the same size (the plain method)
    plain_ms = []
    for _ in range(args.plain_reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        st.infer(dq, out, path=v.PATH_PLAIN_F32, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms.append(e0.elapsed_time(e1))
    plain_ms = float(np.mean(plain_ms[1:]))

    epilogue = epilogue_workload(torch, v, out, drec, stream, K)

    infer_avg_s = float(infer_ms.mean()) / 1e3
    achieved = INFER_FLOP_PER_QUERY * QUERIES / infer_avg_s / 1e12
    res = {
        "metric": "fused-MLP infer queries/s + train samples/s (64-wide); p99 frame latency",
        "value": K * QUERIES / total_s,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32-acc",
        "data": "synthetic (seeded Gaussian-light radiance targets, random-init Kaiming weights)",
        "config": {"workload": "single cache, 2,073,600 queries + 4 x 16,384 training records per frame, "
                               "64->5x64->3, SINCOS6 encoding, relative-L2, dynamic loss scale 128",
                   "queries_per_frame": QUERIES, "train_batches": BATCHES, "train_batch": BATCH},
        "infer_queries_per_s": QUERIES / infer_avg_s,
        "train_samples_per_s": BATCHES * BATCH / (float(train_ms.mean()) / 1e3),
        "frame_ms_p50": float(np.percentile(frame_ms, 50)),
        "frame_ms_p99": float(np.percentile(frame_ms, 99)),
        "infer_ms_avg": float(infer_ms.mean()),
        "train_step_ms_avg": float(train_ms.mean()) / BATCHES,
        "plain_f32_infer_queries_per_s": QUERIES / (plain_ms / 1e3),
        "fused_vs_plain_infer": (QUERIES / infer_avg_s) / (QUERIES / (plain_ms / 1e3)),
        "roofline": {"bound": "mfma", "kernel": "fused_infer_pipe_kernel", "achieved": achieved,
                     "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_F16_TFLOPS,
                     "traffic": INFER_TRAFFIC_BYTES,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE then WRITE_SIZE of this kernel at this "
                                       "workload, scaled by a 256 MiB copy kernel measured in the same runs "
                                       "(profiles/fused_infer_2073600q_hbm_traffic.md); not read in this run",
                     "hbm_GBs": QUERY_BYTES * QUERIES / infer_avg_s / 1e9},
    }
    train_sps = BATCHES * BATCH / (float(train_ms.mean()) / 1e3)
    res["train_roofline"] = {"bound": "mfma", "kernels": "fused_grad_small_kernel + adam_kernel",
                             "achieved": train_sps * TRAIN_FLOP_PER_SAMPLE / 1e12, "peak": PEAK_F16_TFLOPS,
                             "unit": "TFLOP/s", "frac": train_sps * TRAIN_FLOP_PER_SAMPLE / 1e12 / PEAK_F16_TFLOPS,
                             "step_us": float(train_ms.mean()) * 1e3 / BATCHES}
    f8_s = float(f8_ms.mean()) / 1e3
    f8_tflops = INFER_FLOP_PER_QUERY * QUERIES / f8_s / 1e12
    res["fp8"] = {"infer_queries_per_s": QUERIES / f8_s, "infer_ms_avg": f8_s * 1e3,
                  "speedup_vs_f16": infer_avg_s / f8_s, "rel_err_vs_f16": f8_err,
                  "rel_err_vs_fp64": f8_err64, "f16_rel_err_vs_fp64": f16_err64,
                  "error_sample": "65,536 queries, FP64 forward with the FP32 use-weights on the GPU",
                  "achieved_TFLOPs": f8_tflops, "peak_TFLOPs": PEAK_F8_TFLOPS, "frac": f8_tflops / PEAK_F8_TFLOPS}
    res["epilogue"] = epilogue
    log("record front-end ...")
    res["frontend"] = frontend_workload(torch, v, wl, K)
    if args.group_frames > 0:
        log("16-cache workload ...")
        res["group16"] = group_workload(torch, v, wl, q, args.group_frames)
    if args.serve_sessions > 0:
        log("served workload ...")
        res["served"] = served_workload(torch, v, wl, dq, drec, args.serve_sessions, args.serve_budget_gib)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(q, recs, args.cpu_seconds)
    print(json.dumps(res), flush=True)
    st.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
