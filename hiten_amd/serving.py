"""Continuous-batching request loop for corrections and manifold fans.

A request is either
  * a single-shooting correction (corrector.CorrectionProblem), run by the
    corrector's resumable Newton state machine, or
  * a manifold fan: N event searches (DOP853 until y[index] = offset in a
    direction, dense-output bisection), as _integrate_dop853_until_event
    (integrators/rk.py:2678-2803) runs them per trajectory.

The loop is continuous batching, not a queue: requests are admitted as they
arrive while an HBM budget allows (each admitted job reserves the device
bytes of its buffers until it completes); every admitted fan is launched at
once on its own stream; all corrections that are waiting for the same kind of
work at a given moment (event searches -- residuals and every trial of their
line searches -- or STM propagations) go out together in one launch, on a
stream of their own, as soon as the previous batch that held them has
completed.  Kernels of different batches overlap on the device; the host only
polls completion events, moves results to the state machines and launches the
next batches.  Every result is a pure function of its own request: the
kernels integrate each trajectory independently of the batch it is in, so a
job gives the same bits alone and in the mix.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _lib, batch
from .corrector import CorrectionProblem, CorrectionResult, _Bank, halo_problem
from .singlehit import _default_windows

Y_INDEX = 1


@dataclass
class FanRequest:
    """N event searches from the rows of y0 (N, 6)."""
    mu: float
    y0: np.ndarray
    tmax: np.ndarray
    ev_index: int = 1
    ev_offset: float = 0.0
    direction: int = -1
    xtol: float = 1e-13
    gtol: float = 1e-13
    rtol: float = 1e-12
    atol: float = 1e-12

    @property
    def size(self) -> int:
        return int(self.y0.shape[0])


@dataclass
class FanResult:
    hit: np.ndarray
    t_hit: np.ndarray
    y_hit: np.ndarray
    n_accepted: np.ndarray


@dataclass
class Job:
    id: int
    kind: str                      # "correction" | "fan"
    request: object
    arrival: float                 # seconds since the loop started
    nbytes: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: object = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _pinned(t: torch.Tensor) -> torch.Tensor:
    """Asynchronous device -> pinned host copy on the current stream."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    return h


def _up(a) -> torch.Tensor:
    """Asynchronous host -> device copy through pinned memory on the current
    stream (a pageable copy would wait for the stream's earlier work)."""
    h = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).pin_memory()
    return h.to(batch._device(), non_blocking=True)


class _Packed:
    """Inputs of one launch in one pinned upload and its outputs in one device
    block per dtype: float64 rows ``fin`` are uploaded as a (rows, k) array;
    ``fout`` float64 and ``iout`` int32 rows are allocated on the device and
    copied back together once the launch is enqueued."""

    def __init__(self, fin: np.ndarray, n_fout: int, n_iout: int):
        self.k = fin.shape[1]
        self.din = _up(fin)
        dev = self.din.device
        self.fout = torch.empty((n_fout, self.k), dtype=torch.float64, device=dev)
        self.iout = torch.empty((n_iout, self.k), dtype=torch.int32, device=dev)

    @staticmethod
    def ptr(t: torch.Tensor) -> int:
        return t.data_ptr()

    def download(self, frows=None):
        f = self.fout if frows is None else self.fout[frows]
        self.hf = _pinned(f)
        self.hi = _pinned(self.iout)


def _launch_event(pk: _Packed, y_dev: torch.Tensor, stream, *, ev_index: int, ev_offset: float,
                  direction: int, xtol: float, gtol: float, rtol: float, atol: float) -> None:
    """hh_dop853_event on packed buffers: din rows [mu, t0, tmax, ...], y_dev
    (6, k) SoA start states; fout rows [t_hit, y_hit (6), y_last (6)], iout
    rows [hit, status, n_accepted]."""
    lib = _lib.load()
    P = _Packed.ptr
    rc = lib.hh_dop853_event(pk.k, P(pk.din[0]), P(pk.din[1]), P(pk.din[2]), float(rtol),
                             float(atol), math.inf, batch.MIN_STEP, batch.DEFAULT_MAX_STEPS,
                             int(ev_index), float(ev_offset), int(direction), float(xtol),
                             float(gtol), 1, 0, P(y_dev), P(pk.iout[0]), P(pk.fout[0]),
                             P(pk.fout[1]), P(pk.fout[7]), P(pk.iout[1]), P(pk.iout[2]), None, 0,
                             int(stream.cuda_stream))
    _lib.check(rc, "hh_dop853_event")


def _launch_final(pk: _Packed, dim: int, y_dev: torch.Tensor, y_out: torch.Tensor, stream, *,
                  forward: int, flip_indices, variant: int, max_step: float,
                  istatus: int, inacc: int) -> None:
    """hh_dop853_final on packed buffers: din rows [mu, t0, tf, ...]."""
    lib = _lib.load()
    P = _Packed.ptr
    rc = lib.hh_dop853_final(pk.k, dim, P(pk.din[0]), P(pk.din[1]), P(pk.din[2]), 1e-12, 1e-12,
                             float(max_step), batch.MIN_STEP, batch.DEFAULT_MAX_STEPS,
                             int(forward), batch.flip_mask(dim, flip_indices), P(y_dev), P(y_out),
                             P(pk.iout[istatus]), P(pk.iout[inacc]), None, int(variant), 0, 0,
                             None, None, int(stream.cuda_stream))
    _lib.check(rc, "hh_dop853_final")


class _Pending:
    """Device work on one stream whose host copies are complete once
    ``done`` has been recorded and has fired."""

    def __init__(self, stream: torch.cuda.Stream):
        self.stream = stream
        self.done = torch.cuda.Event()

    def ready(self) -> bool:
        return self.done.query()


class _FanLaunch(_Pending):
    def __init__(self, job: Job, stream):
        super().__init__(stream)
        self.job = job
        r: FanRequest = job.request
        k = r.size
        fin = np.empty((9, k))
        fin[0], fin[1], fin[2], fin[3:9] = r.mu, 0.0, r.tmax, r.y0.T
        with torch.cuda.stream(stream):
            self.pk = _Packed(fin, 13, 3)
            _launch_event(self.pk, self.pk.din[3], stream, ev_index=r.ev_index,
                          ev_offset=r.ev_offset, direction=r.direction, xtol=r.xtol, gtol=r.gtol,
                          rtol=r.rtol, atol=r.atol)
            self.pk.download(slice(0, 7))
            self.done.record(stream)

    def finish(self, server: "Server") -> None:
        f, i = self.pk.hf.numpy(), self.pk.hi.numpy()
        self.job.result = FanResult(i[0].astype(bool), f[0].copy(), f[1:7].T.copy(),
                                    i[2].copy())
        server.steps += int(i[2].sum())
        server._complete(self.job)


class _EventBatch(_Pending):
    """The correction residuals of section_crossing_batch (t_guess None) for
    bank rows waiting for event searches: the first window for every state,
    then alignment and the fallback window for the states without a hit
    (poincare/singlehit/backend.py:237-280)."""

    def __init__(self, bank: _Bank, rows: np.ndarray, forward: int, stream):
        super().__init__(stream)
        self.bank, self.rows, self.forward = bank, rows, forward
        self.states, counts = bank.event_inputs(rows)
        self.mu = np.repeat(bank.mu[rows], counts)
        (self.ts, self.span), (self.ts2, self.span2) = _default_windows()
        n = self.states.shape[0]
        self.ok = np.zeros(n, bool)
        self.t = np.full(n, np.nan)
        self.xe = np.full((n, 6), np.nan)
        self.stage = 1
        self._launch(np.arange(n), self.ts, self.span)

    def _launch(self, idx, ts, span):
        self.idx = idx
        self.cur_ts, self.cur_span = ts, span
        self.aligned = not np.isclose(0.0, ts)
        self.searched = not np.isclose(0.0, span)
        k = len(idx)
        fin = np.empty((10, k))
        fin[0], fin[1], fin[2], fin[3], fin[4:10] = self.mu[idx], 0.0, span, ts, self.states[idx].T
        with torch.cuda.stream(self.stream):
            # fout rows: t_hit, y_hit (6), y_last (6), aligned state (6);
            # iout rows: hit, status, n_accepted (event), status, n_accepted (alignment)
            pk = self.pk = _Packed(fin, 19, 5)
            y_dev = pk.din[4:10]
            if self.aligned:
                pk_al = _AlignView(pk)
                _launch_final(pk_al, 6, y_dev, pk.fout[13:19], self.stream,
                              forward=self.forward, flip_indices=None,
                              variant=_lib.HH_VARIANT_PLAIN, max_step=1e4, istatus=3, inacc=4)
                y_dev = pk.fout[13:19]
            if self.searched:
                _launch_event(pk, y_dev, self.stream, ev_index=Y_INDEX, ev_offset=0.0,
                              direction=0, xtol=1e-12, gtol=1e-12, rtol=1e-12, atol=1e-12)
            pk.download(slice(0, 13))
            self.done.record(self.stream)

    def finish(self, server: "Server") -> bool:
        """Consume the finished stage; returns True when the batch is complete."""
        f, i = self.pk.hf.numpy(), self.pk.hi.numpy()
        if self.aligned:
            server.steps += int(i[4].sum())
        if self.searched:
            server.steps += int(i[2].sum())
            hit = i[0].astype(bool)
            th, yh, yl = f[0], f[1:7].T, f[7:13].T
            t_rel = np.where(hit, th, self.cur_span)
            ok = hit & (t_rel < self.cur_span) & (t_rel >= 0.0)
            y_out = np.where(hit[:, None], yh, yl)
            sel = self.idx[ok]
            self.ok[sel] = True
            self.t[sel] = self.cur_ts + t_rel[ok]
            self.xe[sel] = y_out[ok]
        miss = np.nonzero(~self.ok)[0]
        if self.stage == 1 and miss.size:
            self.stage = 2
            self.done = torch.cuda.Event()
            self._launch(miss, self.ts2, self.span2)
            return False
        self.bank.feed_events(self.rows, self.t, self.xe, self.ok)
        server._advance(self.bank, self.rows)
        return True


class _AlignView:
    """The alignment launch reads mu from row 0, t0 (zeros) from row 1 and the
    alignment time from row 3 of an event batch's packed input."""

    def __init__(self, pk: _Packed):
        self.k = pk.k
        self.din = (pk.din[0], pk.din[1], pk.din[3])
        self.iout = pk.iout


class _StmBatch(_Pending):
    """Phi(t_event) of _compute_stm(var_dynsys, x, t_event, steps=500)
    (corrector/operators.py:298-307) for bank rows waiting for their
    Jacobian, one wave per trajectory."""

    def __init__(self, bank: _Bank, rows: np.ndarray, stream):
        super().__init__(stream)
        self.bank, self.rows = bank, rows
        n = len(rows)
        states, self.tev = bank.stm_inputs(rows)
        self.live = ~np.isclose(0.0, self.tev)
        self.h = None
        if self.live.any():
            PHI0 = np.zeros((n, 42))
            PHI0[:, :36] = np.eye(6).ravel()
            PHI0[:, 36:] = states
            mu = bank.mu[rows]
            k = int(self.live.sum())
            fin = np.empty((45, k))
            fin[0], fin[1], fin[2], fin[3:45] = mu[self.live], 0.0, self.tev[self.live], \
                PHI0[self.live].T
            with torch.cuda.stream(stream):
                pk = self.h = _Packed(fin, 42, 2)
                _launch_final(pk, 42, pk.din[3:45], pk.fout, stream, forward=1,
                              flip_indices=slice(36, 42), variant=_lib.HH_VARIANT_WIDE,
                              max_step=1e4, istatus=0, inacc=1)
                pk.download(slice(0, 36))
        self.done.record(stream)

    def finish(self, server: "Server") -> bool:
        phi = np.repeat(np.eye(6)[None], len(self.rows), axis=0)
        if self.h is not None:
            y, na = self.h.hf.numpy().T, self.h.hi.numpy()[1]
            phi[self.live] = y.reshape(-1, 6, 6)
            server.steps += int(na.sum())
        self.bank.feed_stm(self.rows, phi)
        server._advance(self.bank, self.rows)
        return True


class Server:
    """Continuous-batching loop; see the module docstring.

    ``hbm_budget`` bounds the device bytes reserved by admitted jobs; each
    launch goes to one of ``n_streams`` streams (default priority) that has
    no batch in flight when there is one.
    """

    CORRECTION_BYTES = 64 * 1024          # event batch rows (<= 15 per iteration) and STM

    def __init__(self, hbm_budget: int = 32 << 30, n_streams: int = 6):
        self.hbm_budget = int(hbm_budget)
        self.streams = [torch.cuda.Stream() for _ in range(max(2, n_streams))]
        self._rr = 0
        self.reset()

    def _stream(self) -> torch.cuda.Stream:
        """A stream with no batch in flight if there is one (so that the new
        batch does not queue behind an unrelated one), else the next in turn."""
        busy = {id(b.stream) for b in self.inflight}
        for k in range(len(self.streams)):
            s = self.streams[(self._rr + k) % len(self.streams)]
            if id(s) not in busy:
                self._rr = (self._rr + k + 1) % len(self.streams)
                return s
        s = self.streams[self._rr]
        self._rr = (self._rr + 1) % len(self.streams)
        return s

    def reset(self):
        self.used = 0
        self.peak_used = 0
        self.steps = 0
        self.jobs: List[Job] = []
        self.waiting = collections.deque()
        self.banks = {}                 # (n residuals, n controls) -> _Bank
        self.ready = {}                 # bank key -> list of rows ready to launch
        self.owner = {}                 # (bank key, row) -> Job
        self.inflight: List[_Pending] = []
        self.t0 = time.perf_counter()
        self.max_queue = 0

    def now(self) -> float:
        return time.perf_counter() - self.t0

    @staticmethod
    def fan_bytes(n: int) -> int:
        # y0, y_hit, y_last (6 doubles each, SoA and row copies), mu/t0/tmax/t_hit,
        # hit/status/counters
        return n * (6 * 8 * 4 + 4 * 8 + 4 * 4)

    def submit(self, kind: str, request, arrival: Optional[float] = None) -> Job:
        job = Job(len(self.jobs), kind, request, self.now() if arrival is None else arrival)
        job.nbytes = (self.fan_bytes(request.size) if kind == "fan" else self.CORRECTION_BYTES)
        if job.nbytes > self.hbm_budget:
            raise ValueError(f"job {job.id} needs {job.nbytes} B, more than the HBM budget")
        self.jobs.append(job)
        self.waiting.append(job)
        return job

    # -- internals ---------------------------------------------------------
    def _complete(self, job: Job) -> None:
        job.finished = self.now()
        self.used -= job.nbytes

    def _advance(self, bank: _Bank, rows: np.ndarray) -> None:
        """Rows whose batch completed: finished jobs out, the rest ready."""
        key = (bank.nr, bank.nc)
        done = bank.stage[rows] == bank.DONE
        for i in rows[done]:
            job = self.owner.pop((key, int(i)))
            job.result = bank.results[i]
            self._complete(job)
        if (~done).any():
            self.ready.setdefault(key, []).extend(rows[~done].tolist())

    def _admit(self) -> None:
        while self.waiting and self.used + self.waiting[0].nbytes <= self.hbm_budget:
            job = self.waiting.popleft()
            self.used += job.nbytes
            self.peak_used = max(self.peak_used, self.used)
            job.admitted = self.now()
            if job.kind == "fan":
                self.inflight.append(_FanLaunch(job, self._stream()))
            else:
                p = job.request
                key = (len(p.residual_indices), len(p.control_indices))
                bank = self.banks.get(key)
                if bank is None:
                    bank = self.banks[key] = _Bank(*key)
                i = bank.add(p)
                self.owner[(key, i)] = job
                self.ready.setdefault(key, []).append(i)
        self.max_queue = max(self.max_queue, len(self.waiting))

    def _launch_ready(self) -> None:
        for key, rows in self.ready.items():
            if not rows:
                continue
            bank = self.banks[key]
            rows = np.array(rows, dtype=int)
            self.ready[key] = []
            st = bank.stage[rows]
            ev = rows[st != bank.JAC]
            for f in np.unique(bank.fwd[ev]):
                self.inflight.append(_EventBatch(bank, ev[bank.fwd[ev] == f], int(f),
                                                 self._stream()))
            jac = rows[st == bank.JAC]
            if jac.size:
                self.inflight.append(_StmBatch(bank, jac, self._stream()))

    def _poll(self) -> bool:
        """Collect finished batches.  Work on one stream completes in launch
        order, so only the oldest batch of each stream is queried."""
        progressed = False
        seen = set()
        keep, relaunched = [], []
        for b in self.inflight:
            sid = id(b.stream)
            if sid in seen or not b.ready():
                seen.add(sid)
                keep.append(b)
                continue
            progressed = True
            if isinstance(b, _FanLaunch):
                b.finish(self)
            elif not b.finish(self):
                relaunched.append(b)       # its next stage is now last on its stream
                seen.add(sid)
        self.inflight = keep + relaunched
        return progressed

    def busy(self) -> bool:
        return bool(self.waiting or any(self.ready.values()) or self.inflight)

    def step(self) -> bool:
        """One pass of the loop: admit, launch what is ready, collect what has
        finished.  Returns True if anything moved."""
        n_before = len(self.inflight) + len(self.waiting)
        self._admit()
        self._launch_ready()
        moved = self._poll()
        return moved or n_before != len(self.inflight) + len(self.waiting)

    def run(self, stream_of_requests, time_limit: float = 120.0) -> List[Job]:
        """Serve requests (dicts with arrival time ``t``, as
        workloads.serving_stream draws them) in real time; returns the jobs."""
        reqs = sorted(stream_of_requests, key=lambda r: r["t"])
        self.reset()
        i = 0
        while i < len(reqs) or self.busy():
            now = self.now()
            if now > time_limit:
                raise TimeoutError(f"serving loop exceeded {time_limit} s")
            while i < len(reqs) and reqs[i]["t"] <= now:
                self.submit(*to_request(reqs[i]), arrival=reqs[i]["t"])
                i += 1
            if not self.step():
                time.sleep(20e-6)
        torch.cuda.synchronize()
        return self.jobs


def to_request(r: dict):
    """(kind, request) of a workloads.serving_stream entry."""
    if r["kind"] == "correction":
        return "correction", halo_problem(r["mu"], r["x0"], max_attempts=10, tol=1e-12)
    return "fan", FanRequest(r["mu"], np.asarray(r["y0"], float), np.asarray(r["tmax"], float))


def summarize(jobs: List[Job], wall: float, steps: int) -> dict:
    """Latency percentiles (ms) per kind and overall, and throughput."""
    out = {}
    for kind in ("correction", "fan", None):
        lat = np.array([j.latency for j in jobs if kind is None or j.kind == kind]) * 1e3
        if lat.size == 0:
            continue
        key = kind or "all"
        out[key] = dict(n=int(lat.size), p50_ms=float(np.percentile(lat, 50)),
                        p99_ms=float(np.percentile(lat, 99)), max_ms=float(lat.max()))
    out["trajectory_steps_per_s"] = steps / wall if wall > 0 else None
    out["wall_s"] = wall
    return out
