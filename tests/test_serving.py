"""The serving loop's arrival / admission / retirement logic (no GPU)."""
import numpy as np

from vknrc_amd.serving import (RESOLUTIONS, RESOLUTION_WEIGHTS, Scheduler, SessionSpec, poisson_sessions,
                               session_bytes)


def test_poisson_sessions_follow_the_rate_and_the_resolution_mix():
    rng = np.random.default_rng(1)
    s = poisson_sessions(20000, 20.0, rng, frames=300)
    t = np.array([x.arrival for x in s])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.02 / 20.0 * 10
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05          # exponential: cv = 1
    counts = np.array([sum(1 for x in s if (x.width, x.height) == r) for r in RESOLUTIONS]) / len(s)
    assert np.allclose(counts, RESOLUTION_WEIGHTS, atol=0.015)
    assert all(x.queries == x.width * x.height and x.frames == 300 for x in s)


def _run(sched, specs, ticks=10000):
    sched.submit(specs)
    trace = []
    for k in range(ticks):
        sched.arrive(float(k))
        sched.admit(float(k))
        assert sched.used_bytes <= sched.budget_bytes
        assert len(sched.active) <= sched.max_sessions
        trace.append([a.spec.sid for a in sched.active])
        if sched.active:
            sched.advance()
        if sched.idle():
            break
    return trace


def test_admission_keeps_the_budget_and_the_arrival_order():
    cache = 1 << 20
    one = session_bytes(SessionSpec(0, 0, 1920 * 1080, 5, 0), cache)
    sched = Scheduler(budget_bytes=3 * one, cache_bytes=cache, max_sessions=256)
    specs = [SessionSpec(i, float(i // 2), 1920 * 1080, 5, i) for i in range(12)]
    trace = _run(sched, specs)
    assert sched.idle() and not sched.rejected
    assert max(len(t) for t in trace) == 3 and sched.peak_sessions == 3
    # first come first served: admission order is the arrival order
    first_seen = {}
    for k, t in enumerate(trace):
        for sid in t:
            first_seen.setdefault(sid, k)
    order = sorted(first_seen, key=lambda s: (first_seen[s], s))
    assert order == list(range(12))
    # every session ran exactly its frame count
    for sid in range(12):
        assert sum(sid in t for t in trace) == 5
    assert len(sched.waits) == 12 and min(sched.waits) == 0.0


def test_session_cap_and_oversized_sessions():
    cache = 1 << 20
    sched = Scheduler(budget_bytes=1 << 40, cache_bytes=cache, max_sessions=4)
    specs = [SessionSpec(i, 0.0, 1000, 3, i) for i in range(10)]
    big = SessionSpec(99, 0.0, 10 ** 12, 3, 99)
    trace = _run(sched, specs + [big])
    assert max(len(t) for t in trace) == 4
    assert [s.sid for s in sched.rejected] == [99]
    assert sched.used_bytes == 0


def test_head_of_line_session_waits_for_room():
    """A large session at the head of the queue is not overtaken by later small ones."""
    cache = 1 << 20
    small = session_bytes(SessionSpec(0, 0, 1000, 1, 0), cache)
    large = session_bytes(SessionSpec(0, 0, 3840 * 2160, 1, 0), cache)
    sched = Scheduler(budget_bytes=large + small, cache_bytes=cache)
    specs = [SessionSpec(0, 0.0, 1000, 4, 0), SessionSpec(1, 0.0, 1000, 4, 1),
             SessionSpec(2, 1.0, 3840 * 2160, 2, 2), SessionSpec(3, 1.0, 1000, 2, 3)]
    trace = _run(sched, specs)
    start = {sid: next(k for k, t in enumerate(trace) if sid in t) for sid in range(4)}
    assert start[2] <= start[3]
