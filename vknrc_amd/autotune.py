"""Launch-schedule autotuner for grouped frames.

Two schedules are tuned by measurement on the workload itself (HIP events around
repeated launches, the fastest candidate wins):

  * the gradient chunking of each cache (nrc_cache_set_grad_chunks): how many
    64-record chunks one gradient workgroup accumulates before it writes its
    partial dW row.  It trades parallelism against partial-row traffic (one row
    is the whole 20,736-float gradient), and it fixes the cache's summation
    order, so it is measured on scratch caches built from the same descriptions
    and then set on the real caches, whose state the tuning never touches.  The
    search takes the best uniform chunking, then refines it cache by cache;
  * the persistent inference workgroups per CU (nrc_set_infer_workgroups_per_cu),
    measured on the FP16 inference, which changes no cache state and whose
    results do not depend on the grid.
"""
from __future__ import annotations

from . import nrc as _nrc

GRAD_CHUNK_CANDIDATES = (1, 2, 4, 8, 16)
INFER_WG_CANDIDATES = (1, 2, 3, 4)


def _time(torch, fn, reps: int, stream, trials: int = 3) -> float:
    """Median over `trials` of the mean time of `reps` back-to-back calls (ms), after one warm-up call."""
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(trials):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        out.append(e0.elapsed_time(e1) / reps)
    return sorted(out)[len(out) // 2]


def tune_grad_chunks(torch, entries, candidates=GRAD_CHUNK_CANDIDATES, reps: int = 8,
                     path: int = _nrc.PATH_FUSED_F16, apply: bool = True, per_cache: bool = True,
                     min_gain: float = 0.02) -> dict:
    """Times one grouped training step (nrc_group_train, batch 0) on scratch copies of the caches.

    First every cache takes the same chunking, each candidate in turn; then, from the best of those,
    one pass of coordinate descent over the caches (largest batch first) tries every other candidate
    for that cache alone and keeps it when the grouped step gets faster by more than `min_gain`.
    The chunkings found are set on the real caches when ``apply``."""
    stream = torch.cuda.current_stream()
    scratch = [dict(e, cache=_nrc.NRCState(e["cache"].config)) for e in entries]

    def step():
        _nrc.group_train(scratch, 0, path=path, stream=stream)

    ms = {}
    try:
        for c in candidates:
            for e in scratch:
                e["cache"].set_grad_chunks(c)
            ms[c] = _time(torch, step, reps, stream)
        best = min(ms, key=ms.get)
        chosen = [best] * len(scratch)
        for e in scratch:
            e["cache"].set_grad_chunks(best)
        cur = ms[best]
        if per_cache and len(scratch) > 1:
            order = sorted(range(len(scratch)), key=lambda i: -int(scratch[i].get("max_batch_records", 0)))
            for i in order:
                for c in candidates:
                    if c == chosen[i]:
                        continue
                    scratch[i]["cache"].set_grad_chunks(c)
                    t = _time(torch, step, reps, stream)
                    if t < cur * (1.0 - min_gain):
                        cur, chosen[i] = t, c
                    else:
                        scratch[i]["cache"].set_grad_chunks(chosen[i])
    finally:
        torch.cuda.synchronize()
        for e in scratch:
            e["cache"].close()
    if apply:
        for e, c in zip(entries, chosen):
            e["cache"].set_grad_chunks(c)
    return {"best": best, "ms_per_step": ms, "per_cache": chosen, "tuned_ms_per_step": cur}


def tune_infer_workgroups(torch, entries, candidates=INFER_WG_CANDIDATES, reps: int = 5,
                          apply: bool = True) -> dict:
    """Times one grouped FP16 inference (nrc_group_infer) per workgroups-per-CU candidate (median of
    three timings of `reps` launches each)."""
    stream = torch.cuda.current_stream()
    ms = {}
    for w in candidates:
        _nrc.set_infer_workgroups_per_cu(w)
        ms[w] = _time(torch, lambda: _nrc.group_infer(entries, path=_nrc.PATH_FUSED_F16, stream=stream), reps,
                      stream)
    best = min(ms, key=ms.get)
    _nrc.set_infer_workgroups_per_cu(best if apply else 2)
    return {"best": best, "ms_per_launch": ms}
