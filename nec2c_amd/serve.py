"""Continuous-batching request loop for dielectric VIE solves on one GPU.

A request asks for the solution of one antenna model at one frequency for
one or more feed ports (one right-hand side each).  The server

* keeps prepared systems (kernel spectrum, permittivity, C/B/D blocks, all in
  HBM) in a cache keyed by (grid, frequency, permittivity key), so a request
  for a model already seen skips the spectrum build;
* admits requests in arrival order while the HBM footprint of the work in
  flight (cached systems + one pooled Krylov workspace per port) stays under a
  budget, evicting idle cached systems least-recently-used first;
* submits every port of an admitted request as a job of one
  ``vie.Batch`` (the ``vie_hip_batch_*`` C-ABI), so the Arnoldi steps of all
  requests in flight are interleaved on the device while new requests keep
  arriving;
* lets the ports of one model share a recycled Krylov subspace
  (``vie_hip_batch_share_subspaces``): the first port solved on a model
  publishes the subspace of its first cycle, later ports and later requests
  for the same cached model start from it; a request whose model has no
  subspace yet runs its first port alone and submits the others when it
  finishes (``seed_first``);
* runs new systems with the default launch shapes instead of tuning them
  (``tune``; a served model gets a few solves, fewer than tuning needs to
  pay for itself);
* records per-request latency (arrival -> last port finished) and service time
  (admission -> last port finished).

The reference has no server: it solves one frequency per forked process,
one right-hand side at a time (dielectric_wire.c:1296-1304, main.c:586-660).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Callable, Hashable, List, Optional, Sequence

import numpy as np

from . import vie

KRYLOV_SCRATCH = 1024 * 320 * 16  # reduction partials per job (krylov.h kReduceScratchBytes)


def system_device_bytes(shape, conductors, pec=False):
    """Device bytes of a prepared system: octant spectrum (FP64 + FP32), FFT
    workspace (FP64-sized), permittivity, cross blocks (FP64 + FP32)."""
    nx, ny, nz = shape
    kz = 2 * nz + 1 if pec else nz
    P = [1 << max(0, int(np.ceil(np.log2(max(2 * e - 1, 1))))) for e in (nx, ny, kz)]
    nbox = nx * ny * nz
    octant = 6 * (P[0] // 2 + 1) * (P[1] // 2 + 1) * (P[2] // 2 + 1)
    work = 3 * nx * ny * P[2] + 3 * nx * P[1] * P[2]
    c, nv = conductors, 3 * nbox
    blocks = c * c + 2 * c * nv
    return octant * 24 + work * 16 + nbox * 24 + blocks * 24 + (c + nv) * 16 * 4


def job_device_bytes(shape, conductors, restart, precision=vie.MIXED):
    """Device bytes one in-flight solve holds (Krylov basis, FP64 vectors,
    reduction scratch).  The pooled basis is FP64-sized, so it serves either
    precision (a mixed basis with its recycled subspace, restart + 1 + 2k FP32
    columns with k <= restart / 2, fits in it); a mixed solve that hands over
    to FP64 cycles keeps a recycled subspace of up to 32 vectors in FP64
    (krylov.h kMaxRecycle), so its basis may grow by 2 x 32 FP64 columns."""
    n = conductors + 3 * int(np.prod(shape))
    ldv = (n + 511) // 512 * 512   # krylov.h basis_ld
    extra = 2 * min(32, restart // 2) if precision == vie.MIXED else 0
    return (restart + 1 + extra) * ldv * 16 + 4 * n * 16 + KRYLOV_SCRATCH


def ready_system(system, tune=False):
    """What the server does to a freshly built system before its first job:
    without `tune` it fixes the default launch shape of every pass
    (vie_fft_set_variants with variant 0), so the system's first solve runs
    no tuning timings.  A served model gets a few solves, and tuning one costs
    more than it saves over them."""
    if not tune and hasattr(system, "set_variants"):
        system.set_variants([0] * 5, [0] * 5)
    return system


@dataclass
class Request:
    rid: int
    arrival: float                      # seconds after the server clock starts
    key: Hashable                       # (grid, frequency, permittivity key)
    shape: tuple
    conductors: int
    make_system: Callable[[], "vie.System"]
    rhs: Sequence[np.ndarray]           # one right-hand side per port
    # filled by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    cache_hit: bool = False
    results: list = field(default_factory=list)


class Server:
    def __init__(self, budget_bytes: int, precision=vie.MIXED, restart=120, maximum_iterations=2000,
                 relative_tolerance=1e-8, inner_reduction=0.0, recycle=0, share=True, seed_first=True,
                 tune=False, clock=time.perf_counter):
        self.budget = int(budget_bytes)
        self.precision = precision
        self.inner_reduction = inner_reduction
        self.recycle = recycle
        self.opts = dict(restart=restart, maximum_iterations=maximum_iterations,
                         relative_tolerance=relative_tolerance)
        self.restart = restart
        # recycled vectors of a solve (the library's default: 16 for MIXED, none for FP64)
        self.rk = recycle if recycle > 0 else (16 if precision == vie.MIXED and recycle == 0 else 0)
        self.share = share and self.rk > 0
        self.seed_first = seed_first and self.share
        self.batch = vie.Batch(share_subspaces=self.share)
        # key -> [system, bytes, users, seeded (a port has finished on it)]
        self.cache: "collections.OrderedDict[Hashable, list]" = collections.OrderedDict()
        self.deferred = {}       # rid -> ports waiting for the model's first port
        self.queue: "collections.deque[Request]" = collections.deque()
        self.inflight = {}       # job id -> (request, port)
        self.pending = {}        # rid -> ports still running
        self.used = 0            # bytes: cached systems + in-flight jobs
        self.done: List[Request] = []
        self.rejected: List[Request] = []
        self.clock = clock
        self.t0 = None
        self.builds = 0
        self.build_seconds = 0.0
        self.tune = tune         # tune each new system's passes (see ready_system)

    # ---------------------------------------------------------------- helpers
    def now(self):
        return self.clock() - self.t0

    def _job_bytes(self, req):
        return job_device_bytes(req.shape, req.conductors, self.restart, self.precision)

    def _system_bytes(self, req):
        """A cached system, with the shared subspace it holds once a port has
        published one (2 k FP32 basis vectors)."""
        extra = 0
        if self.share:
            n = req.conductors + 3 * int(np.prod(req.shape))
            extra = 2 * self.rk * ((n + 511) // 512 * 512) * 8
        return system_device_bytes(req.shape, req.conductors) + extra

    def _footprint(self, req):
        sys_bytes = 0 if req.key in self.cache else self._system_bytes(req)
        return sys_bytes + len(req.rhs) * self._job_bytes(req)

    def _evict_for(self, need, keep=None):
        """Closes idle cached systems, least recently used first, until `need`
        more bytes fit.  The system of key `keep` (the request being admitted,
        whose footprint counted it as cached) is never evicted."""
        for key in list(self.cache):
            if self.used + need <= self.budget:
                return True
            s, nbytes, users, _ = self.cache[key]
            if users == 0 and key != keep:
                s.close()
                del self.cache[key]
                self.used -= nbytes
        return self.used + need <= self.budget

    def _admit(self, req):
        now = self.now()
        entry = self.cache.get(req.key)
        if entry is None:
            t = self.clock()
            s = ready_system(req.make_system(), self.tune)
            self.build_seconds += self.clock() - t
            self.builds += 1
            entry = [s, self._system_bytes(req), 0, False]
            self.cache[req.key] = entry
            self.used += entry[1]
        else:
            req.cache_hit = True
            self.cache.move_to_end(req.key)
        entry[2] += 1
        req.admitted = now
        req.results = [None] * len(req.rhs)
        self.pending[req.rid] = len(req.rhs)
        ports = list(range(len(req.rhs)))
        if self.seed_first and len(ports) > 1 and not entry[3]:
            # the model's first port publishes its recycled subspace; the
            # others start from it (their memory is reserved now)
            self.deferred[req.rid] = ports[1:]
            self.used += len(ports[1:]) * self._job_bytes(req)
            ports = ports[:1]
        for port in ports:
            self._submit(req, entry, port)
            self.used += self._job_bytes(req)

    def _submit(self, req, entry, port):
        jid = self.batch.submit(entry[0], req.rhs[port], precision=self.precision,
                                inner_reduction=self.inner_reduction, recycle=self.recycle, **self.opts)
        self.inflight[jid] = (req, port)

    def _collect(self, jid):
        req, port = self.inflight.pop(jid)
        ok, x, rep, sec = self.batch.result(jid)
        req.results[port] = (ok, x, rep, sec)
        self.used -= self._job_bytes(req)
        entry = self.cache[req.key]
        entry[3] = True
        for q in self.deferred.pop(req.rid, ()):
            self._submit(req, entry, q)
        self.pending[req.rid] -= 1
        if self.pending[req.rid] == 0:
            del self.pending[req.rid]
            req.finished = self.now()
            self.cache[req.key][2] -= 1
            self.done.append(req)

    # ---------------------------------------------------------------- loop
    def step(self, arrived: "collections.deque[Request]"):
        """Admits what fits, advances every job once, collects finished jobs.
        Returns True when something happened."""
        busy = False
        now = self.now()
        while arrived and arrived[0].arrival <= now:
            self.queue.append(arrived.popleft())
        while self.queue:
            req = self.queue[0]
            need = self._footprint(req)
            if self.used + need > self.budget and not self._evict_for(need, keep=req.key):
                if not self.inflight and need > self.budget:
                    self.rejected.append(self.queue.popleft())  # can never fit
                    continue
                break  # wait for memory: arrival order is kept
            self._admit(self.queue.popleft())
            busy = True
        ready, _ = self.batch.progress()
        for jid in ready:
            self._collect(jid)
            busy = True
        return busy

    def run(self, requests: Sequence[Request]):
        """Serves `requests` (sorted by arrival) to completion in real time."""
        arrived = collections.deque(sorted(requests, key=lambda r: r.arrival))
        total = len(arrived) + len(self.done) + len(self.rejected)
        self.t0 = self.clock()
        while len(self.done) + len(self.rejected) < total:
            if not self.step(arrived) and not self.inflight and not self.queue and arrived:
                wait = arrived[0].arrival - self.now()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
        return self.report()

    def report(self):
        lat = np.array([r.finished - r.arrival for r in self.done])
        svc = np.array([r.finished - r.admitted for r in self.done])
        solves = sum(len(r.rhs) for r in self.done)
        span = max(r.finished for r in self.done) if self.done else 0.0
        reps = [res[2] for r in self.done for res in r.results]
        return {
            "requests": len(self.done),
            "rejected": len(self.rejected),
            "solves": solves,
            "seconds": span,
            "solves_per_s": solves / span if span else 0.0,
            "requests_per_s": len(self.done) / span if span else 0.0,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": 1e3 * float(np.percentile(lat, 99)) if len(lat) else None,
            "service_p50_ms": 1e3 * float(np.percentile(svc, 50)) if len(svc) else None,
            "service_p99_ms": 1e3 * float(np.percentile(svc, 99)) if len(svc) else None,
            "cache_hits": sum(r.cache_hit for r in self.done),
            "system_builds": self.builds,
            "system_build_seconds": self.build_seconds,
            "converged": sum(bool(res[0]) for r in self.done for res in r.results),
            "max_relative_residual": max((rep.relative_residual for rep in reps), default=None),
            "iterations_total": sum(rep.iterations for rep in reps),
            "failures": [{"request": r.rid, "key": list(r.key) if isinstance(r.key, tuple) else str(r.key),
                          "port": p, "status": int(res[2].status), "iterations": int(res[2].iterations),
                          "relative_residual": float(res[2].relative_residual)}
                         for r in self.done for p, res in enumerate(r.results) if not res[0]],
        }

    def timing_document(self, deck="served request stream"):
        """The run as a timing document (nec2c_amd/timing.py, schema version 3):
        the library's phases and this server's report."""
        from . import timing
        rep = self.report()
        return timing.document(deck, vie.phases(), rep["seconds"], serving=rep,
                               success=not self.rejected and rep["converged"] == rep["solves"])

    def close(self):
        self.batch.close()
        for s, *_ in self.cache.values():
            s.close()
        self.cache.clear()


def run_batch(jobs, precision=vie.MIXED, inflight=128, inner_reduction=0.0, batch=None, recycle=0, share=False,
              **opts):
    """Solves a closed batch: `jobs` is a list of (system, rhs).  At most
    `inflight` jobs are on the device at once; a finished job's slot is refilled
    at once (continuous batching).  `batch` may be a caller's ``vie.Batch``
    (e.g. with reserved workspaces).  With `share`, jobs of one system share
    the recycled subspace of the first of them to finish a cycle
    (vie_hip_batch_share_subspaces): put one right-hand side per system first.
    Returns (results in job order, seconds); each result is (converged, x,
    report, seconds in flight)."""
    own = batch is None
    if own:
        batch = vie.Batch()
    if share:
        batch.share_subspaces(True)
    try:
        results = [None] * len(jobs)
        ids = {}
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(jobs) or ids:
            while nxt < len(jobs) and len(ids) < inflight:
                s, b = jobs[nxt]
                ids[batch.submit(s, b, precision=precision, inner_reduction=inner_reduction, recycle=recycle,
                                 **opts)] = nxt
                nxt += 1
            ready, _ = batch.progress()
            for jid in ready:
                results[ids.pop(jid)] = batch.result(jid)
        return results, time.perf_counter() - t0
    finally:
        if own:
            batch.close()
