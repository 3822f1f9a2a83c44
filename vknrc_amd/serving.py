"""Served caches: render sessions admitted by an HBM budget and advanced together.

A session is one radiance cache serving one renderer: every frame it infers the
radiance of its full-resolution query buffer and takes one training step of
16,384 records (VkNRCState's per-frame use, NRCRenderGraph.cpp:57-80, with one
step instead of four).  The reference runs a single cache per process and has
no serving layer; this module is the request loop around the grouped launches:

  * arrivals: a Poisson process of sessions, each with a resolution and a frame
    count (``poisson_sessions``);
  * admission: first come first served, while the sessions' device bytes (the
    cache's own allocations from nrc_cache_footprint, plus 56 + 12 bytes a query
    for the query and radiance buffers) fit the HBM budget and fewer than
    ``max_sessions`` run; a session that could never fit is rejected;
  * continuous batching: every tick runs one frame of every admitted session as
    grouped launches (nrc_group_infer, then nrc_group_train with the use-weights
    written), so sessions that arrive while others are on the GPU join at the
    next tick and run in the same launches; a session leaves after its last
    frame and frees its budget;
  * measurement: HIP events around each tick give its device time, which is the
    frame latency of every session in it; admission waits, queries/s and the
    concurrency are reported beside the latency percentiles.

``Scheduler`` holds the arrival / admission / retirement logic and no device
state (it is tested on the CPU); ``Server`` runs it on the GPU.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import nrc as _nrc

RESOLUTIONS = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))
RESOLUTION_WEIGHTS = (0.3, 0.4, 0.2, 0.1)
QUERY_BYTES = 56  # nrc_query: 14 floats
RGB_BYTES = 12
TRAIN_STEP_RECORDS = 16384


@dataclass
class SessionSpec:
    """One render session: it arrives at ``arrival`` (seconds, or ticks for a tick clock) and
    runs ``frames`` frames over ``queries`` queries with a cache seeded by ``seed``."""
    sid: int
    arrival: float
    queries: int
    frames: int
    seed: int
    width: int = 0
    height: int = 0


def poisson_sessions(n: int, rate: float, rng: np.random.Generator, frames: int = 300,
                     resolutions=RESOLUTIONS, weights=RESOLUTION_WEIGHTS, seed0: int = 0) -> list[SessionSpec]:
    """n sessions with exponential inter-arrival times of mean 1/rate and resolutions drawn with
    the given weights."""
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    pick = rng.choice(len(resolutions), n, p=np.asarray(weights) / np.sum(weights))
    out = []
    for i in range(n):
        w, h = resolutions[pick[i]]
        out.append(SessionSpec(i, float(t[i]), w * h, frames, seed0 + i, w, h))
    return out


def session_bytes(spec: SessionSpec, cache_bytes: int) -> int:
    """Device bytes of a session: its cache plus its query and radiance buffers."""
    return int(cache_bytes) + spec.queries * (QUERY_BYTES + RGB_BYTES)


@dataclass
class _Active:
    spec: SessionSpec
    nbytes: int
    admitted_at: float
    done: int = 0


@dataclass
class Scheduler:
    """Arrival, admission and retirement decisions (no device work)."""
    budget_bytes: int
    cache_bytes: int
    max_sessions: int = 256
    pending: list = field(default_factory=list)    # not yet arrived, by arrival time
    waiting: list = field(default_factory=list)    # arrived, not admitted (FIFO)
    active: list = field(default_factory=list)
    rejected: list = field(default_factory=list)
    used_bytes: int = 0
    peak_bytes: int = 0
    peak_sessions: int = 0
    waits: list = field(default_factory=list)      # admission time - arrival time

    def submit(self, specs) -> None:
        self.pending.extend(specs)
        self.pending.sort(key=lambda s: s.arrival)

    def arrive(self, now: float) -> None:
        while self.pending and self.pending[0].arrival <= now:
            s = self.pending.pop(0)
            if session_bytes(s, self.cache_bytes) > self.budget_bytes:
                self.rejected.append(s)
            else:
                self.waiting.append(s)

    def admit(self, now: float) -> list:
        """Admits waiting sessions in arrival order while they fit; returns the new ones."""
        new = []
        while self.waiting and len(self.active) < self.max_sessions:
            s = self.waiting[0]
            nb = session_bytes(s, self.cache_bytes)
            if self.used_bytes + nb > self.budget_bytes:
                break
            self.waiting.pop(0)
            a = _Active(s, nb, now)
            self.active.append(a)
            self.used_bytes += nb
            self.waits.append(now - s.arrival)
            new.append(a)
        self.peak_bytes = max(self.peak_bytes, self.used_bytes)
        self.peak_sessions = max(self.peak_sessions, len(self.active))
        return new

    def advance(self) -> list:
        """One frame of every active session has run; returns (and retires) the finished ones."""
        done = []
        for a in self.active:
            a.done += 1
        for a in list(self.active):
            if a.done >= a.spec.frames:
                self.active.remove(a)
                self.used_bytes -= a.nbytes
                done.append(a)
        return done

    def idle(self) -> bool:
        return not (self.pending or self.waiting or self.active)

    def next_arrival(self):
        return self.pending[0].arrival if self.pending else None


class Server:
    """Runs a Scheduler's sessions on the GPU.

    ``query_pool`` is a device tensor [>= largest session's queries][14] that admission copies a
    session's query buffer from (at an offset given by its seed); ``record_pool`` is a device tensor
    [P * 16384][17] of training records, and frame f of a session trains on batch (seed + f) mod P.
    ``make_config(spec)`` gives the cache description of a session.
    """

    def __init__(self, torch, query_pool, record_pool, make_config, budget_bytes: int,
                 max_sessions: int = 256, path: int = _nrc.PATH_FUSED_F16, stream=None):
        self.torch = torch
        self.qpool = query_pool
        self.rpool = record_pool
        self.nbatches = record_pool.shape[0] // TRAIN_STEP_RECORDS
        self.make_config = make_config
        self.path = path
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        probe = make_config(SessionSpec(0, 0.0, 1, 1, 0))
        self.sched = Scheduler(budget_bytes, _nrc.cache_footprint(probe, path), max_sessions)
        self.res = {}           # sid -> dict(cache, queries, rgb)
        self.ticks = []         # (start event, end event, sessions, queries)
        self.finished = {}      # sid -> callback results
        self.on_finish = None   # optional callback(spec, cache, rgb) before a session's cache is freed

    # -- session resources
    def _open(self, a: _Active) -> None:
        s = a.spec
        cfg = self.make_config(s)
        cache = _nrc.NRCState(cfg)
        n = s.queries
        off = (s.seed * 7919) % max(1, self.qpool.shape[0] - n + 1)
        q = self.qpool[off:off + n].clone()
        rgb = self.torch.zeros((n, 3), dtype=self.torch.float32, device=self.qpool.device)
        self.res[s.sid] = dict(cache=cache, queries=q, rgb=rgb)

    def _close(self, a: _Active) -> None:
        r = self.res.pop(a.spec.sid)
        if self.on_finish is not None:
            self.on_finish(a.spec, r["cache"], r["rgb"])
        r["cache"].close()

    def _entries(self):
        ents = []
        for a in self.sched.active:
            r = self.res[a.spec.sid]
            b = (a.spec.seed + a.done) % self.nbatches
            ents.append(dict(cache=r["cache"], queries=r["queries"], rgb=r["rgb"], max_queries=a.spec.queries,
                             batches=self.rpool[b * TRAIN_STEP_RECORDS:(b + 1) * TRAIN_STEP_RECORDS],
                             max_batch_records=TRAIN_STEP_RECORDS))
        return ents

    def tick(self, now: float) -> int:
        """Admits what fits and runs one frame of every active session; returns the session count."""
        self.sched.arrive(now)
        for a in self.sched.admit(now):
            self._open(a)
        if not self.sched.active:
            return 0
        ents = self._entries()
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        _nrc.group_infer(ents, path=self.path, stream=self.stream)
        _nrc.group_train(ents, 0, path=self.path, write_use_weights=True, stream=self.stream)
        e1.record(self.stream)
        nq = sum(a.spec.queries for a in self.sched.active)
        self.ticks.append((e0, e1, len(ents), nq))
        for a in self.sched.advance():
            e1.synchronize()
            self._close(a)
        return len(ents)

    def run_realtime(self, specs, max_in_flight: int = 2, time_limit: float = 600.0,
                     frame_interval: float | None = None) -> dict:
        """Arrivals on the wall clock (SessionSpec.arrival in seconds from the start).  With a
        frame_interval (seconds, e.g. 1/60) the ticks follow a frame clock: a tick starts no earlier
        than one interval after the previous one, as a renderer presenting at that rate would ask for
        frames; a tick that starts later than its slot counts as late."""
        self.sched.submit(specs)
        t0 = time.perf_counter()
        next_slot = 0.0
        self.late_ticks = 0
        while not self.sched.idle():
            now = time.perf_counter() - t0
            if now > time_limit:
                raise RuntimeError("serving run exceeded its time limit")
            if frame_interval is not None and now < next_slot:
                time.sleep(next_slot - now)
                now = time.perf_counter() - t0
            n = self.tick(now)
            if n == 0:
                nxt = self.sched.next_arrival()
                if nxt is not None:
                    time.sleep(max(0.0, min(nxt - (time.perf_counter() - t0), 0.01)))
                next_slot = time.perf_counter() - t0
                continue
            if frame_interval is not None:
                if now > next_slot + frame_interval:
                    self.late_ticks += 1
                next_slot = max(next_slot + frame_interval, now)
            if len(self.ticks) > max_in_flight:  # keep the host at most a few ticks ahead of the device
                self.ticks[-1 - max_in_flight][1].synchronize()
        self.torch.cuda.synchronize()
        rep = self.report(time.perf_counter() - t0)
        if frame_interval is not None:
            rep["frame_interval_ms"] = frame_interval * 1e3
            rep["late_ticks"] = self.late_ticks
        return rep

    def run_ticks(self, specs) -> dict:
        """Arrivals on a tick clock (SessionSpec.arrival = the tick it arrives at)."""
        self.sched.submit(specs)
        k = 0
        t0 = time.perf_counter()
        while not self.sched.idle():
            self.tick(float(k))
            k += 1
        self.torch.cuda.synchronize()
        return self.report(time.perf_counter() - t0)

    def report(self, wall_s: float) -> dict:
        lat = []
        per_tick = []
        total_q = 0
        for e0, e1, n, nq in self.ticks:
            ms = e0.elapsed_time(e1)
            per_tick.append(ms)
            lat.extend([ms] * n)
            total_q += nq
        lat = np.asarray(lat) if lat else np.zeros(1)
        waits = np.asarray(self.sched.waits) if self.sched.waits else np.zeros(1)
        busy_s = float(np.sum(per_tick)) / 1e3
        return {"sessions": len(self.sched.waits), "rejected": len(self.sched.rejected), "ticks": len(self.ticks),
                "frames": int(len(lat)), "frame_ms_p50": float(np.percentile(lat, 50)),
                "frame_ms_p99": float(np.percentile(lat, 99)), "frame_ms_max": float(lat.max()),
                "admission_wait_s_p50": float(np.percentile(waits, 50)),
                "admission_wait_s_p99": float(np.percentile(waits, 99)),
                "peak_sessions": self.sched.peak_sessions, "peak_bytes": int(self.sched.peak_bytes),
                "mean_sessions_per_tick": float(len(lat) / max(1, len(self.ticks))),
                "queries": int(total_q), "wall_s": wall_s, "device_busy_s": busy_s,
                "queries_per_s": total_q / wall_s if wall_s > 0 else 0.0,
                "queries_per_device_s": total_q / busy_s if busy_s > 0 else 0.0}


def grouped_vs_sequential(torch, specs, query_pool, record_pool, make_config, ticks: int = 5,
                          path: int = _nrc.PATH_FUSED_F16) -> dict:
    """Device time of one frame of every session in `specs` as the serving loop runs it (grouped
    inference and training launches) against per-session launches (nrc_infer then nrc_train for
    each session), averaged over `ticks` frames after one warm-up frame, each on its own caches."""
    nb = record_pool.shape[0] // TRAIN_STEP_RECORDS

    def sessions():
        ents = []
        for s in specs:
            n = s.queries
            off = (s.seed * 7919) % max(1, query_pool.shape[0] - n + 1)
            b = s.seed % nb
            ents.append(dict(cache=_nrc.NRCState(make_config(s)), queries=query_pool[off:off + n],
                             rgb=torch.zeros((n, 3), dtype=torch.float32, device=query_pool.device),
                             max_queries=n, batches=record_pool[b * TRAIN_STEP_RECORDS:(b + 1) * TRAIN_STEP_RECORDS],
                             max_batch_records=TRAIN_STEP_RECORDS))
        return ents

    stream = torch.cuda.current_stream()

    def grouped(ents):
        _nrc.group_infer(ents, path=path, stream=stream)
        _nrc.group_train(ents, 0, path=path, write_use_weights=True, stream=stream)

    def sequential(ents):
        for e in ents:
            e["cache"].infer(e["queries"], e["rgb"], path=path, stream=stream)
            e["cache"].train(e["batches"], path=path, write_use_weights=True, stream=stream)

    out = {}
    for name, fn in (("grouped", grouped), ("sequential", sequential)):
        ents = sessions()
        fn(ents)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(ticks):
            fn(ents)
        e1.record(stream)
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) / ticks
        for e in ents:
            e["cache"].close()
    return {"sessions": len(specs), "queries_per_frame": int(sum(s.queries for s in specs)),
            "grouped_ms_per_frame": out["grouped"], "sequential_ms_per_frame": out["sequential"],
            "grouped_vs_sequential": out["sequential"] / out["grouped"]}

