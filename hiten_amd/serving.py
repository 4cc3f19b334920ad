"""Continuous-batching request loop for corrections, continuations and
manifold fans.

A request is one of
  * a single-shooting correction (corrector.CorrectionProblem),
  * a continuation: one periodic-orbit family grown from a seed by the
    predictor-corrector loop of continuation/backends/pc.py:42-166 (as
    hiten_amd.continuation batches it: each round corrects a window of
    retry predictions); the loop runs each round's corrections as
    correction work of its own, beside every other request's, and starts
    the next round as soon as the results in hand fix the window's
    outcome, or
  * a manifold fan: N event searches (DOP853 until y[index] = offset in a
    direction, dense-output bisection), as _integrate_dop853_until_event
    (integrators/rk.py:2678-2803) runs them per trajectory.

The loop is continuous batching, not a queue: requests are admitted as they
arrive while an HBM budget allows (each admitted job reserves the device
bytes of its buffers until it completes), and every admitted fan is launched
at once on a stream of its own.

Corrections run device-resident by default (``corrector="device"``): all
corrections admitted since the last launch go out in one hh_correct launch,
one workgroup per correction running its whole Newton-Armijo loop, on a
dedicated stream -- a hardware queue of its own, so a launch that runs for
milliseconds holds up no fan and no other launch.  The kernel reads the
problems from and writes each result to mapped host memory, setting the
result's done flag last; the loop completes a job the moment its flag is
set.  With ``corrector="bank"`` the corrections advance instead through the
host-side state machine: all corrections waiting for the same kind of work
(event searches -- residuals and every trial of their line searches -- or
STM propagations) go out together in one launch as soon as the previous
batch that held them has completed.

Kernels of different launches overlap on the device; the host only polls
completion events and flags, moves results and launches.  Every result is
a pure function of its own request: the kernels integrate each trajectory
(and run each correction) independently of the launch it is in, so a job
gives the same bits alone and in the mix.
"""

from __future__ import annotations

import collections
import ctypes
import functools
import gc
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib, batch
from . import continuation as _cont
from . import corrector as _corr
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
class ContinuationRequest:
    """One family continued from ``seed`` (a corrected orbit):
    hiten_amd.continuation.continue_family's arguments; ``make_problem``
    turns a prediction into its correction problem."""
    seed: np.ndarray
    make_problem: Callable
    step: np.ndarray
    max_members: int = 8
    max_retries_per_step: int = 50
    stepper: str = "natural"
    state_indices: object = None
    shrink_policy: Optional[Callable] = None
    step_min: float = 1e-10
    step_max: float = 1.0
    window: int = 2

    def family_kwargs(self) -> dict:
        return dict(step=self.step, max_members=self.max_members,
                    max_retries_per_step=self.max_retries_per_step, stepper=self.stepper,
                    state_indices=self.state_indices, shrink_policy=self.shrink_policy,
                    step_min=self.step_min, step_max=self.step_max)


@dataclass
class FanResult:
    hit: np.ndarray
    t_hit: np.ndarray
    y_hit: np.ndarray
    n_accepted: np.ndarray


@dataclass
class Job:
    id: int
    kind: str                      # "correction" | "continuation" | "fan"
    request: object
    arrival: float                 # seconds since the loop started
    nbytes: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: object = None
    parent: object = None          # (run, round, index) of a continuation's correction

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Slot:
    """Reusable staging for one batch in flight: a stream, a completion event,
    pinned host and device buffers that grow to the largest batch seen.  A
    slot is handed to a new batch only after its previous batch completed, so
    its buffers are never overwritten while a copy or kernel still uses them.
    Copies and events go straight to the HIP runtime through the library."""

    def __init__(self, stream: torch.cuda.Stream):
        self.lib = _lib.load()
        self.stream = stream
        self.sh = int(stream.cuda_stream)
        self.ev = self.lib.hh_event_create()
        if not self.ev:
            raise _lib.HipLibraryError("hh_event_create failed")
        self.busy = False
        self.nf_in = self.nf_out = self.ni_out = 0

    def __del__(self):
        ev, self.ev = getattr(self, "ev", None), None
        if ev:
            self.lib.hh_event_destroy(ev)

    def record(self) -> None:
        _lib.check(self.lib.hh_event_record(self.ev, self.sh), "hh_event_record")

    def ready(self) -> bool:
        r = self.lib.hh_event_query(self.ev)
        if r < 0:
            _lib.check(-r, "hh_event_query")
        return r == 1

    def copy(self, dst: int, src: int, nbytes: int, to_device: bool) -> None:
        _lib.check(self.lib.hh_memcpy_async(dst, src, nbytes, int(to_device), self.sh),
                   "hh_memcpy_async")

    @staticmethod
    def _grow(n_have, n_need, dtype, pinned):
        n = max(n_need, 2 * n_have)
        if pinned:
            return n, torch.empty(n, dtype=dtype, pin_memory=True)
        return n, torch.empty(n, dtype=dtype, device=batch._device())

    def reserve(self, nf_in: int, nf_out: int, ni_out: int) -> None:
        if nf_in > self.nf_in:
            n, self.hin = self._grow(self.nf_in, nf_in, torch.float64, True)
            _, self.din = self._grow(self.nf_in, nf_in, torch.float64, False)
            self.hin_np = self.hin.numpy()
            self.nf_in = n
        if nf_out > self.nf_out:
            n, self.hf = self._grow(self.nf_out, nf_out, torch.float64, True)
            _, self.fout = self._grow(self.nf_out, nf_out, torch.float64, False)
            self.hf_np = self.hf.numpy()
            self.nf_out = n
        if ni_out > self.ni_out:
            n, self.hi = self._grow(self.ni_out, ni_out, torch.int32, True)
            _, self.iout = self._grow(self.ni_out, ni_out, torch.int32, False)
            self.hi_np = self.hi.numpy()
            self.ni_out = n


class _Packed:
    """Inputs of one launch in one pinned upload and its outputs in one device
    block per dtype, all in a slot's buffers: float64 rows ``fin`` (rows, k)
    go up in one copy; ``fout`` float64 and ``iout`` int32 rows come back in
    one copy each once the launch is enqueued (on the slot's stream)."""

    def __init__(self, fin: np.ndarray, n_fout: int, n_iout: int, slot: _Slot):
        r, k = fin.shape
        self.k, self.n_fout, self.n_iout = k, n_fout, n_iout
        slot.reserve(r * k, n_fout * k, n_iout * k)
        self.slot = slot
        slot.hin_np[:r * k] = fin.reshape(-1)
        slot.copy(slot.din.data_ptr(), slot.hin.data_ptr(), r * k * 8, True)
        self.din_ptr = slot.din.data_ptr()
        self.fout_ptr = slot.fout.data_ptr()
        self.iout_ptr = slot.iout.data_ptr()

    def din(self, row: int) -> int:
        """device address of input row ``row``"""
        return self.din_ptr + row * self.k * 8

    def fout(self, row: int) -> int:
        return self.fout_ptr + row * self.k * 8

    def iout(self, row: int) -> int:
        return self.iout_ptr + row * self.k * 4

    def download(self, frows=None):
        """Copy fout rows [frows] and all iout rows to the slot's pinned
        buffers; hf/hi are numpy views of them (valid once the event fired)."""
        a, b = (0, self.n_fout) if frows is None else (frows.start, frows.stop)
        k, sl = self.k, self.slot
        nf, ni = (b - a) * k, self.n_iout * k
        sl.copy(sl.hf.data_ptr(), self.fout(a), nf * 8, False)
        sl.copy(sl.hi.data_ptr(), self.iout_ptr, ni * 4, False)
        self.hf = sl.hf_np[:nf].reshape(b - a, k)
        self.hi = sl.hi_np[:ni].reshape(self.n_iout, k)


def _launch_event(pk: _Packed, y_row: int, stream_handle: int, *, ev_index: int,
                  ev_offset: float, direction: int, xtol: float, gtol: float, rtol: float,
                  atol: float, y_addr: Optional[int] = None, variant: int = 0) -> None:
    """hh_dop853_event on packed buffers: din rows [mu, t0, tmax, ...], start
    states at din row ``y_row`` (6 rows, SoA) or at ``y_addr``; fout rows
    [t_hit, y_hit (6), y_last (6)], iout rows [hit, status, n_accepted]."""
    lib = _lib.load()
    y = pk.din(y_row) if y_addr is None else y_addr
    rc = lib.hh_dop853_event(pk.k, pk.din(0), pk.din(1), pk.din(2), float(rtol),
                             float(atol), math.inf, batch.MIN_STEP, batch.DEFAULT_MAX_STEPS,
                             int(ev_index), float(ev_offset), int(direction), float(xtol),
                             float(gtol), 1, 0, y, pk.iout(0), pk.fout(0), pk.fout(1), pk.fout(7),
                             pk.iout(1), pk.iout(2), None, int(variant), 0, stream_handle)
    _lib.check(rc, "hh_dop853_event")


def _launch_final(pk: _Packed, dim: int, rows3, y_addr: int, out_addr: int, stream_handle: int, *,
                  forward: int, flip_mask: int, variant: int, max_step: float,
                  istatus: int, inacc: int) -> None:
    """hh_dop853_final on packed buffers: mu, t0, tf at din rows ``rows3``."""
    lib = _lib.load()
    r0, r1, r2 = rows3
    rc = lib.hh_dop853_final(pk.k, dim, pk.din(r0), pk.din(r1), pk.din(r2), 1e-12, 1e-12,
                             float(max_step), batch.MIN_STEP, batch.DEFAULT_MAX_STEPS,
                             int(forward), flip_mask, y_addr, out_addr, pk.iout(istatus),
                             pk.iout(inacc), None, int(variant), 0, 0, None, 0, None,
                             stream_handle)
    _lib.check(rc, "hh_dop853_final")


_FLIP_STM = batch.flip_mask(42, slice(36, 42))
_EYE36 = np.eye(6).ravel()


class _Pending:
    """Device work on a slot's stream whose host copies are complete once the
    slot's event, recorded after them, has fired."""

    def __init__(self, slot: _Slot):
        self.slot = slot
        self.stream = slot.stream
        slot.busy = True

    def ready(self) -> bool:
        return self.slot.ready()


class _FanLaunch(_Pending):
    def __init__(self, job: Job, slot: _Slot):
        super().__init__(slot)
        self.job = job
        r: FanRequest = job.request
        k = r.size
        fin = np.empty((9, k))
        fin[0], fin[1], fin[2], fin[3:9] = r.mu, 0.0, r.tmax, r.y0.T
        self.pk = _Packed(fin, 13, 3, slot)
        _launch_event(self.pk, 3, slot.sh, ev_index=r.ev_index, ev_offset=r.ev_offset,
                      direction=r.direction, xtol=r.xtol, gtol=r.gtol, rtol=r.rtol, atol=r.atol,
                      variant=_lib.HH_VARIANT_WIDE)
        self.pk.download(slice(0, 7))
        slot.record()

    def finish(self, server: "Server") -> None:
        f, i = self.pk.hf, self.pk.hi
        self.job.result = FanResult(i[0].astype(bool), f[0].copy(), f[1:7].T.copy(),
                                    i[2].copy())
        server.steps += int(i[2].sum())
        server._complete(self.job)


class _EventBatch(_Pending):
    """The correction residuals of section_crossing_batch (t_guess None) for
    bank rows waiting for event searches: both windows of
    poincare/singlehit/backend.py:237-280 in one hh_plane_crossing launch,
    one state per 8-lane group."""

    def __init__(self, bank: _Bank, rows: np.ndarray, forward: int, slot: _Slot,
                 ev_index: int = Y_INDEX):
        super().__init__(slot)
        self.bank, self.rows, self.forward = bank, rows, forward
        states, counts = bank.event_inputs(rows)
        mu = np.repeat(bank.mu[rows], counts)
        (ts, span), (ts2, span2) = _default_windows()
        k = states.shape[0]
        # din rows: mu, state (6); fout rows: t_event, x_event (6);
        # iout rows: ok, accepted steps
        fin = np.empty((7, k))
        fin[0], fin[1:7] = mu, states.T
        pk = self.pk = _Packed(fin, 7, 2, slot)
        lib = _lib.load()
        rc = lib.hh_plane_crossing(k, pk.din(0), pk.din(1), int(forward), int(ev_index), 0.0, 0,
                                   1e-12, 1e-12, ts, span, ts2, span2, pk.iout(0), pk.fout(0),
                                   pk.fout(1), pk.iout(1), slot.sh)
        _lib.check(rc, "hh_plane_crossing")
        pk.download(slice(0, 7))
        slot.record()

    def finish(self, server: "Server") -> bool:
        f, i = self.pk.hf, self.pk.hi
        server.steps += int(i[1].sum())
        ok = i[0].astype(bool)
        t = np.where(ok, f[0], np.nan)
        xe = np.where(ok[:, None], f[1:7].T, np.nan)
        self.bank.feed_events(self.rows, t, xe, ok)
        server._advance(self.bank, self.rows)
        return True


class _StmBatch(_Pending):
    """Phi(t_event) of _compute_stm(var_dynsys, x, t_event, steps=500)
    (corrector/operators.py:298-307) for bank rows waiting for their
    Jacobian, one wave per trajectory."""

    def __init__(self, bank: _Bank, rows: np.ndarray, slot: _Slot):
        super().__init__(slot)
        self.bank, self.rows = bank, rows
        states, self.tev = bank.stm_inputs(rows)
        tev = self.tev
        self.live = ~(np.abs(0.0 - tev) <= 1e-8 + 1e-5 * np.abs(tev))   # numpy.isclose(0, t)
        self.pk = None
        if self.live.any():
            k = int(self.live.sum())
            # din rows: mu, t0 (0), t_event, PHI0 (42: identity STM, state)
            fin = np.empty((45, k))
            fin[0], fin[1], fin[2] = bank.mu[rows][self.live], 0.0, tev[self.live]
            fin[3:39] = _EYE36[:, None]
            fin[39:45] = states[self.live].T
            pk = self.pk = _Packed(fin, 42, 2, slot)
            # two waves per trajectory (state and STM side by side) while the
            # batch leaves the GPU mostly idle, else one wave per trajectory
            var = _lib.HH_VARIANT_PAIR if k <= 1024 else _lib.HH_VARIANT_WIDE
            _launch_final(pk, 42, (0, 1, 2), pk.din(3), pk.fout(0), slot.sh, forward=1,
                          flip_mask=_FLIP_STM, variant=var, max_step=1e4,
                          istatus=0, inacc=1)
            pk.download(slice(0, 36))
        slot.record()

    def finish(self, server: "Server") -> bool:
        n = len(self.rows)
        phi = np.empty((n, 36))
        phi[:] = _EYE36
        if self.pk is not None:
            phi[self.live] = self.pk.hf.T
            server.steps += int(self.pk.hi[1].sum())
        self.bank.feed_stm(self.rows, phi.reshape(n, 6, 6))
        server._advance(self.bank, self.rows)
        return True


class _CorrSlot:
    """A dedicated stream (a hardware queue of its own, so a launch that runs
    for milliseconds holds up no other stream) with mapped, coherent host
    records for device-resident correction launches: the kernel reads the
    problems from and writes the results to host memory directly.  Buffers
    grow to the largest batch; a slot takes a new batch only after its
    previous launch completed."""

    def __init__(self):
        self.lib = _lib.load()
        self.sh = self.lib.hh_stream_create_dedicated()
        if not self.sh:
            raise _lib.HipLibraryError("hh_stream_create_dedicated failed")
        self.ev = self.lib.hh_event_create()
        if not self.ev:
            raise _lib.HipLibraryError("hh_event_create failed")
        self.busy = False
        self.cap = 0
        self._ptrs = []

    def _alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        _lib.check(self.lib.hh_host_alloc(nbytes, ctypes.byref(p)), "hh_host_alloc")
        self._ptrs.append(p.value)
        return p.value

    def reserve(self, n: int) -> None:
        if n <= self.cap:
            return
        self._free()
        cap = max(n, 2 * self.cap, 16)
        pb, rb = _corr.PROBLEM_DTYPE.itemsize * cap, _corr.RESULT_DTYPE.itemsize * cap
        self.pptr, self.rptr = self._alloc(pb), self._alloc(rb)
        self.prob = np.frombuffer((ctypes.c_char * pb).from_address(self.pptr),
                                  dtype=_corr.PROBLEM_DTYPE)
        self.res = np.frombuffer((ctypes.c_char * rb).from_address(self.rptr),
                                 dtype=_corr.RESULT_DTYPE)
        self.cap = cap

    def _free(self) -> None:
        self.prob = self.res = None
        for p in self._ptrs:
            self.lib.hh_host_free(p)
        self._ptrs = []

    def record(self) -> None:
        _lib.check(self.lib.hh_event_record(self.ev, self.sh), "hh_event_record")

    def ready(self) -> bool:
        r = self.lib.hh_event_query(self.ev)
        if r < 0:
            _lib.check(-r, "hh_event_query")
        return r == 1

    def close(self) -> None:
        """Release the stream, event and buffers (after the last launch)."""
        if getattr(self, "sh", None):
            self.lib.hh_stream_destroy(self.sh)
            self.sh = None
        if getattr(self, "ev", None):
            self.lib.hh_event_destroy(self.ev)
            self.ev = None
        self._free()


class _CorrLaunch:
    """Corrections run whole on the device (hh_correct: one workgroup per
    correction runs its Newton-Armijo loop to the end) on a dedicated
    stream; each job completes as soon as its own result's done flag is
    set, without waiting for the rest of its launch."""

    def __init__(self, jobs: list, slot: _CorrSlot):
        n = len(jobs)
        slot.reserve(n)
        slot.busy = True
        self.slot, self.jobs = slot, jobs
        slot.prob[:n] = _corr.pack_problems([j.request for j in jobs])
        slot.res["done"][:n] = 0
        _corr.launch_device(slot.pptr, slot.rptr, n, stream_handle=slot.sh)
        slot.record()
        self.pending = np.arange(n)

    def poll(self, server: "Server") -> bool:
        """Complete the jobs whose results are in; True once the launch is
        over (every result in and the kernel finished)."""
        if self.pending.size:
            res = self.slot.res
            flags = res["done"][self.pending]
            if flags.any():
                got = self.pending[flags != 0]
                recs = res[got].copy()
                for i, r in zip(got, _corr.results_from_records(
                        recs, [self.jobs[i].request for i in got])):
                    job = self.jobs[i]
                    job.result = r
                    server._complete(job)
                server.steps += int(recs["n_steps"].sum())
                self.pending = self.pending[flags == 0]
        if self.pending.size == 0 and self.slot.ready():
            self.slot.busy = False
            return True
        return False


class _ContRun:
    """A continuation job between rounds: the family's loop state, the
    window of predictions being corrected and the results that came back.
    Results of a superseded round (corrections the outcome no longer
    needed) are dropped."""

    def __init__(self, job: Job):
        r: ContinuationRequest = job.request
        if r.stepper not in ("natural", "secant"):
            raise ValueError("stepper must be 'natural' or 'secant'")
        if r.window < 1:
            raise ValueError("window must be >= 1")
        self.job = job
        self.fam = _cont._Family(r.seed, r.make_problem, **r.family_kwargs())
        self.round = 0

    def start_round(self, server: "Server") -> None:
        self.round += 1
        self.win = self.fam.window(int(self.job.request.window))
        self.res = [None] * len(self.win)
        self.fam.corrections += len(self.win)
        probs = _cont.make_problems(self.fam.make_problem, [pred for _, pred, _ in self.win])
        for k, p in enumerate(probs):
            if p is None:                     # its correction raised: not converged
                self.res[k] = _cont.FAILED
                continue
            sub = Job(-1, "correction", p, server.now())
            sub.admitted = sub.arrival
            sub.parent = (self, self.round, k)
            server._enqueue_correction(sub)
        if self.fam.decided(self.res) and self not in server.cont_dirty:
            server.cont_dirty.append(self)    # decided without a launch

    def on_result(self, rnd: int, k: int, result: CorrectionResult) -> bool:
        """Store a result; False if its round is over (or the family is)."""
        if rnd != self.round:
            return False
        self.res[k] = _cont.result_tuple(result, self.win[k][1])
        return True

    def advance(self, server: "Server") -> bool:
        """Consume the window once its outcome is fixed and start the next
        round; True once the family is complete."""
        n = self.fam.decided(self.res)
        if n == 0:
            return False
        self.fam.consume(self.win[:n], self.res[:n])
        if self.fam.growing:
            self.start_round(server)
            return False
        self.round = -1              # corrections still in flight are dropped
        self.job.result = self.fam.result()
        return True


class Server:
    """Continuous-batching loop; see the module docstring.

    ``hbm_budget`` bounds the device bytes reserved by admitted jobs.  Each
    batch in flight holds a staging slot (stream, event, reusable pinned and
    device buffers); slots are spread over ``n_streams`` default-priority
    streams (torch's pool, whose streams share the runtime's few hardware
    queues, or with ``dedicated_streams`` streams with queues of their own).
    Device-resident correction launches go to up to ``n_correction_streams``
    dedicated streams; at 2k requests/s four keep the correction latency low
    without delaying fans (more hardware queues than that make the
    scheduler time-slice them and fans wait).
    """

    CORRECTION_BYTES = 64 * 1024          # event batch rows (<= 15 per iteration) and STM

    def __init__(self, hbm_budget: int = 32 << 30, n_streams: int = 8,
                 corrector: str = "device", n_correction_streams: int = 4,
                 dedicated_streams: bool = False):
        if corrector not in ("device", "bank"):
            raise ValueError("corrector must be 'device' or 'bank'")
        self.hbm_budget = int(hbm_budget)
        self.corrector = corrector
        self._own_streams = []
        if dedicated_streams:
            lib = _lib.load()
            for _ in range(max(2, n_streams)):
                h = lib.hh_stream_create_dedicated()
                if not h:
                    raise _lib.HipLibraryError("hh_stream_create_dedicated failed")
                self._own_streams.append(h)
            self.streams = [torch.cuda.ExternalStream(h) for h in self._own_streams]
        else:
            self.streams = [torch.cuda.Stream() for _ in range(max(2, n_streams))]
        self.slots: List[_Slot] = []
        self.n_corr_slots = max(1, int(n_correction_streams))
        self.corr_slots: List[_CorrSlot] = []
        self.reset()

    def close(self) -> None:
        """Release the dedicated correction streams and their host records
        (waits for launches still running)."""
        torch.cuda.synchronize()
        for sl in self.corr_slots:
            sl.close()
        self.corr_slots = []
        if self._own_streams:
            # the staging slots run on these streams: the server is done
            lib = _lib.load()
            for h in self._own_streams:
                lib.hh_stream_destroy(h)
            self._own_streams = []
            self.slots, self.streams = [], []

    def __del__(self):
        try:
            if self.corr_slots or self._own_streams:
                self.close()
        except Exception:
            pass

    def _corr_slot(self) -> Optional[_CorrSlot]:
        for sl in self.corr_slots:
            if not sl.busy:
                return sl
        if len(self.corr_slots) < self.n_corr_slots:
            sl = _CorrSlot()
            self.corr_slots.append(sl)
            return sl
        return None

    def _slot(self) -> _Slot:
        """A slot whose previous batch has completed; a new one (on the next
        stream in turn) if all are in flight."""
        for sl in self.slots:
            if not sl.busy:
                return sl
        sl = _Slot(self.streams[len(self.slots) % len(self.streams)])
        self.slots.append(sl)
        return sl

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
        self.corr_waiting: list = []    # admitted corrections waiting for a device launch
        self.corr_inflight: List[_CorrLaunch] = []
        self.cont_runs: List[_ContRun] = []     # admitted continuations not yet complete
        self.cont_dirty: List[_ContRun] = []    # continuations with results to consume
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
        if kind not in ("correction", "continuation", "fan"):
            raise ValueError(f"unknown request kind {kind!r}")
        job = Job(len(self.jobs), kind, request, self.now() if arrival is None else arrival)
        if kind == "fan":
            job.nbytes = self.fan_bytes(request.size)
        elif kind == "continuation":
            job.nbytes = max(1, int(request.window)) * self.CORRECTION_BYTES
        else:
            job.nbytes = self.CORRECTION_BYTES
        if job.nbytes > self.hbm_budget:
            raise ValueError(f"job {job.id} needs {job.nbytes} B, more than the HBM budget")
        self.jobs.append(job)
        self.waiting.append(job)
        return job

    # -- internals ---------------------------------------------------------
    def _complete(self, job: Job) -> None:
        job.finished = self.now()
        self.used -= job.nbytes
        if job.parent is not None:
            run, rnd, k = job.parent
            if run.on_result(rnd, k, job.result) and run not in self.cont_dirty:
                self.cont_dirty.append(run)

    def _enqueue_correction(self, job: Job) -> None:
        """An admitted correction: to the next device-resident launch, or
        into the bank of its problem's shape."""
        if self.corrector == "device" and _corr.device_supported(job.request):
            self.corr_waiting.append(job)
            return
        p = job.request
        key = (len(p.residual_indices), len(p.control_indices))
        bank = self.banks.get(key)
        if bank is None:
            bank = self.banks[key] = _Bank(*key)
        i = bank.add(p)
        self.owner[(key, i)] = job
        self.ready.setdefault(key, []).append(i)

    def _advance_continuations(self) -> bool:
        """Continuations with new results: next round, or done."""
        runs, self.cont_dirty = self.cont_dirty, []
        for run in runs:
            if run.advance(self):
                self.cont_runs.remove(run)
                self._complete(run.job)
        return bool(runs)

    def _advance(self, bank: _Bank, rows: np.ndarray) -> None:
        """Rows whose batch completed: finished jobs out, the rest ready."""
        key = (bank.nr, bank.nc)
        done = bank.stages(rows) == bank.DONE
        for i in rows[done]:
            job = self.owner.pop((key, int(i)))
            job.result = bank.result(int(i))
            bank.release(int(i))
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
                self.inflight.append(_FanLaunch(job, self._slot()))
            elif job.kind == "continuation":
                run = _ContRun(job)
                if run.fam.growing:
                    self.cont_runs.append(run)
                    run.start_round(self)
                else:                       # max_members <= 1: the seed alone
                    job.result = run.fam.result()
                    self._complete(job)
            else:
                self._enqueue_correction(job)
        self.max_queue = max(self.max_queue, len(self.waiting))

    def _launch_ready(self) -> None:
        if self.corr_waiting:
            sl = self._corr_slot()
            if sl is not None:
                self.corr_inflight.append(_CorrLaunch(self.corr_waiting, sl))
                self.corr_waiting = []
        for key, rows in self.ready.items():
            if not rows:
                continue
            bank = self.banks[key]
            rows = np.array(rows, dtype=int)
            self.ready[key] = []
            st = bank.stages(rows)
            ev = rows[st != bank.JAC]
            fwd, evi = bank.fwd[ev], bank.evi[ev]
            for f, e in sorted(set(zip(fwd.tolist(), evi.tolist()))):
                self.inflight.append(_EventBatch(bank, ev[(fwd == f) & (evi == e)], int(f),
                                                 self._slot(), ev_index=int(e)))
            jac = rows[st == bank.JAC]
            if jac.size:
                self.inflight.append(_StmBatch(bank, jac, self._slot()))

    def _poll(self) -> bool:
        """Collect finished batches.  Work on one stream completes in launch
        order, so only the oldest batch of each stream is queried."""
        progressed = False
        n_jobs_open = sum(len(c.pending) for c in self.corr_inflight)
        self.corr_inflight = [c for c in self.corr_inflight if not c.poll(self)]
        if sum(len(c.pending) for c in self.corr_inflight) != n_jobs_open:
            progressed = True
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
                b.slot.busy = False
            elif not b.finish(self):
                relaunched.append(b)       # its next stage is now last on its stream
                seen.add(sid)
            else:
                b.slot.busy = False
        self.inflight = keep + relaunched
        return progressed

    def busy(self) -> bool:
        return bool(self.waiting or any(self.ready.values()) or self.inflight
                    or self.corr_waiting or self.corr_inflight or self.cont_runs)

    def step(self) -> bool:
        """One pass of the loop: admit, launch what is ready, collect what has
        finished.  Returns True if anything moved."""
        def count():
            return (len(self.inflight) + len(self.waiting) + len(self.corr_inflight)
                    + len(self.corr_waiting))
        n_before = count()
        self._admit()
        self._launch_ready()
        moved = self._poll()
        if self._advance_continuations():
            self._launch_ready()
            moved = True
        return moved or n_before != count()

    def run(self, stream_of_requests, time_limit: float = 120.0) -> List[Job]:
        """Serve requests (dicts with arrival time ``t``, as
        workloads.serving_stream draws them) in real time; returns the jobs."""
        reqs = sorted(stream_of_requests, key=lambda r: r["t"])
        # no cyclic garbage collection inside the loop: a full collection
        # over the loop's objects stalls every request in flight (one is run
        # before the clock starts)
        gc_was_on = gc.isenabled()
        gc.collect()
        gc.disable()
        self.reset()
        i = 0
        try:
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
        finally:
            if gc_was_on:
                gc.enable()
        return self.jobs


def serving_halo_problem(mu: float, x0) -> CorrectionProblem:
    """The stream's correction: Halo defaults, at most 10 iterations, tol 1e-12."""
    return halo_problem(mu, x0, max_attempts=10, tol=1e-12)


def to_request(r: dict):
    """(kind, request) of a workloads.serving_stream entry."""
    if r["kind"] == "correction":
        return "correction", serving_halo_problem(r["mu"], r["x0"])
    if r["kind"] == "continuation":
        return "continuation", ContinuationRequest(
            np.asarray(r["seed"], float), functools.partial(serving_halo_problem, r["mu"]),
            np.asarray(r["step"], float), max_members=int(r["members"]),
            max_retries_per_step=int(r["max_retries"]), state_indices=list(r["state_indices"]),
            window=int(r["window"]))
    return "fan", FanRequest(r["mu"], np.asarray(r["y0"], float), np.asarray(r["tmax"], float))


def correction_outcome(result) -> str:
    """converged / max_iterations (Newton stopped at max_attempts) /
    line_search (Armijo found no productive step) / no_crossing (no section
    crossing for an iterate) / other, from a CorrectionResult."""
    if result is None:
        return "other"
    if result.converged:
        return "converged"
    err = result.error or ""
    if err.startswith("Newton did not converge"):
        return "max_iterations"
    if "Armijo line search failed" in err:
        return "line_search"
    if err.startswith("no section crossing"):
        return "no_crossing"
    return "other"


def _percentiles(lat_ms: np.ndarray) -> dict:
    return dict(n=int(lat_ms.size), p50_ms=float(np.percentile(lat_ms, 50)),
                p99_ms=float(np.percentile(lat_ms, 99)), max_ms=float(lat_ms.max()))


def summarize(jobs: List[Job], wall: float, steps: int) -> dict:
    """Latency percentiles (ms) per kind and overall, the corrections' by
    outcome, and throughput."""
    out = {}
    for kind in ("correction", "continuation", "fan", None):
        lat = np.array([j.latency for j in jobs if kind is None or j.kind == kind]) * 1e3
        if lat.size == 0:
            continue
        key = kind or "all"
        out[key] = _percentiles(lat)
    corr = [j for j in jobs if j.kind == "correction"]
    if corr:
        by = {}
        for j in corr:
            by.setdefault(correction_outcome(j.result), []).append(j.latency * 1e3)
        out["correction"]["outcomes"] = {k: _percentiles(np.array(v)) for k, v in sorted(by.items())}
    out["trajectory_steps_per_s"] = steps / wall if wall > 0 else None
    out["wall_s"] = wall
    return out
