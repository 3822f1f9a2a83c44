"""Launch-schedule autotuner for grouped frames.

Two schedules are tuned by measurement on the workload itself (HIP events around
repeated launches, the fastest candidate wins):

  * the gradient chunking of each cache (nrc_cache_set_grad_chunks): how many
    64-record chunks one gradient workgroup accumulates before it writes its
    partial dW row.  It trades parallelism against partial-row traffic (one row
    is the whole 20,736-float gradient), and it fixes the cache's summation
    order, so it is measured on scratch caches built from the same descriptions
    and then set on the real caches, whose state the tuning never touches;
  * the persistent inference workgroups per CU (nrc_set_infer_workgroups_per_cu),
    measured on the FP16 inference, which changes no cache state and whose
    results do not depend on the grid.
"""
from __future__ import annotations

from . import nrc as _nrc

GRAD_CHUNK_CANDIDATES = (1, 2, 4, 8, 16)
INFER_WG_CANDIDATES = (1, 2, 3, 4)


def _time(torch, fn, reps: int, stream) -> float:
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_grad_chunks(torch, entries, candidates=GRAD_CHUNK_CANDIDATES, reps: int = 8,
                     path: int = _nrc.PATH_FUSED_F16, apply: bool = True) -> dict:
    """Times one grouped training step (nrc_group_train, batch 0) of scratch copies of the caches
    for each chunking; sets the fastest on the real caches when ``apply``."""
    stream = torch.cuda.current_stream()
    scratch = [dict(e, cache=_nrc.NRCState(e["cache"].config)) for e in entries]
    ms = {}
    try:
        for c in candidates:
            for e in scratch:
                e["cache"].set_grad_chunks(c)
            ms[c] = _time(torch, lambda: _nrc.group_train(scratch, 0, path=path, stream=stream), reps, stream)
    finally:
        torch.cuda.synchronize()
        for e in scratch:
            e["cache"].close()
    best = min(ms, key=ms.get)
    if apply:
        for e in entries:
            e["cache"].set_grad_chunks(best)
    return {"best": best, "ms_per_step": ms}


def tune_infer_workgroups(torch, entries, candidates=INFER_WG_CANDIDATES, reps: int = 5,
                          apply: bool = True) -> dict:
    """Times one grouped FP16 inference (nrc_group_infer) per workgroups-per-CU candidate."""
    stream = torch.cuda.current_stream()
    ms = {}
    for w in candidates:
        _nrc.set_infer_workgroups_per_cu(w)
        ms[w] = _time(torch, lambda: _nrc.group_infer(entries, path=_nrc.PATH_FUSED_F16, stream=stream), reps,
                      stream)
    best = min(ms, key=ms.get)
    _nrc.set_infer_workgroups_per_cu(best if apply else 2)
    return {"best": best, "ms_per_launch": ms}
