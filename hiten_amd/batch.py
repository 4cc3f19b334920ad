"""Batched GPU propagation: thin Python wrappers over the C-ABI.

Inputs may be NumPy arrays or torch tensors; trajectories are rows
(``y0.shape == (n, dim)``, the reference's per-trajectory vector layout).
Device buffers are structure-of-arrays (``[dim][n]``); results are returned as
torch tensors on the GPU (rows = trajectories) unless ``to_numpy=True``.

Every function here runs the HIP kernels of ``libhiten_hip.so``; there is no
CPU path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib

EPS = float(np.finfo(float).eps)
MIN_STEP = 10.0 * EPS                  # integrators/rk.py:833-834
DEFAULT_MAX_STEPS = 1_000_000
ALL_COMPONENTS = (1 << 64) - 1


def flip_mask(dim: int, flip_indices=None) -> int:
    """Bit mask of the components a backward-time _DirectedSystem negates.

    ``None`` flips every component (dynamics/base.py:298-300).
    """
    if flip_indices is None:
        return ALL_COMPONENTS
    if isinstance(flip_indices, slice):
        idx = range(dim)[flip_indices]
    else:
        idx = [int(i) for i in np.atleast_1d(flip_indices)]
    m = 0
    for i in idx:
        if not 0 <= i < dim:
            raise ValueError(f"flip index {i} out of range for dim {dim}")
        m |= 1 << i
    return m


_DEVICES: dict = {}


def _device() -> torch.device:
    lib = _lib.load()
    d = _DEVICES.get(id(lib))
    if d is None:
        d = _DEVICES[id(lib)] = torch.device("cuda", torch.cuda.current_device())
    return d


def _vec(x, n: int, name: str) -> torch.Tensor:
    """(n,) float64 device tensor from a scalar or array."""
    dev = _device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=torch.float64)
    else:
        t = torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)
    if t.ndim == 0:
        t = t.expand(n)
    if t.shape != (n,):
        raise ValueError(f"{name} must be a scalar or have shape ({n},), got {tuple(t.shape)}")
    return t.contiguous()


def _soa(y0, dim: int) -> torch.Tensor:
    """(n, dim) rows -> contiguous (dim, n) device tensor."""
    dev = _device()
    if isinstance(y0, torch.Tensor):
        y = y0.to(device=dev, dtype=torch.float64)
    else:
        y = torch.as_tensor(np.asarray(y0, dtype=np.float64), device=dev)
    if y.ndim == 1:
        y = y[None, :]
    if y.ndim != 2 or y.shape[1] != dim:
        raise ValueError(f"states must have shape (n, {dim}), got {tuple(y.shape)}")
    return y.t().contiguous()


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _out(t: torch.Tensor, to_numpy: bool):
    return t.cpu().numpy() if to_numpy else t


@dataclass
class FinalResult:
    """DOP853 final-value propagation of a batch."""
    y: torch.Tensor              # (n, dim) state at tf (dense-output value)
    status: torch.Tensor         # (n,) int32, 0 ok / 1 max_steps / 2 non-finite
    n_accepted: torch.Tensor     # (n,) int32
    n_rejected: torch.Tensor     # (n,) int32


@dataclass
class EventResult:
    hit: torch.Tensor            # (n,) bool
    t_hit: torch.Tensor          # (n,)
    y_hit: torch.Tensor          # (n, 6)
    y_last: torch.Tensor         # (n, 6)
    status: torch.Tensor
    n_accepted: torch.Tensor
    n_rejected: torch.Tensor


FIXED_ORDERS = (4, 6, 8)


def rk8_fixed(mu, y0, t0, tf, steps: int, **kw):
    """rk_fixed with the 13-stage RK8 tableau."""
    return rk_fixed(mu, y0, t0, tf, steps, order=8, **kw)


def rk_fixed(mu, y0, t0, tf, steps: int, *, order: int = 8, dim: int = 6, t_vals=None,
             forward: int = 1, flip_indices=None, return_trajectory: bool = False,
             return_derivatives: bool = False, variant: str = "auto", block: int = 0,
             stream=None):
    """Fixed-step RK of the given order (4, 6 or 8: the tableaux of
    rk.py:755-783) over numpy.linspace(t0, tf, steps), or over the explicit
    grid ``t_vals`` shared by all trajectories (rk.py:531-588).

    Returns the final states (n, dim); with ``return_trajectory`` also the
    grid states as an (n, steps, dim) view, and with ``return_derivatives``
    (which implies the grid) also the derivative at every grid state, as
    _integrate_fixed_rk returns them.  For dim 6, ``variant`` "wide"
    integrates one trajectory per 8-lane group (one component per lane, the
    latency form), "plain" one per lane; "auto" takes "wide" up to 65,536
    trajectories.  Both give the same bits.
    """
    if order not in FIXED_ORDERS:
        raise ValueError("RK order must be 4, 6, or 8")
    lib = _lib.load()
    y = _soa(y0, dim)
    n = y.shape[1]
    mu_t = _vec(mu, n, "mu")
    tv = None
    if t_vals is not None:
        tv = _vec(t_vals, len(t_vals), "t_vals")
        steps = tv.shape[0]
        t0_t = tf_t = None
    else:
        t0_t, tf_t = _vec(t0, n, "t0"), _vec(tf, n, "tf")
    out = torch.empty_like(y)
    grid = return_trajectory or return_derivatives
    traj = torch.empty((steps, dim, n), dtype=torch.float64, device=y.device) if grid else None
    der = torch.empty((steps, dim, n), dtype=torch.float64, device=y.device) \
        if return_derivatives else None
    if variant == "auto":
        variant = "wide" if n <= 65536 else "plain"
    rc = lib.hh_rk_fixed(int(order), n, dim, _ptr(mu_t), _ptr(t0_t), _ptr(tf_t), int(steps),
                         _ptr(tv), int(forward), flip_mask(dim, flip_indices), _ptr(y), _ptr(out),
                         _ptr(traj), _ptr(der), {"plain": _lib.HH_VARIANT_PLAIN,
                                                 "wide": _lib.HH_VARIANT_WIDE}[variant],
                         int(block), _lib.stream_handle(stream))
    _lib.check(rc, "hh_rk_fixed")
    if return_derivatives:
        return out.t(), traj.permute(2, 0, 1), der.permute(2, 0, 1)
    if return_trajectory:
        return out.t(), traj.permute(2, 0, 1)
    return out.t()


@dataclass
class FixedEventResult:
    hit: torch.Tensor            # (n,) bool
    t_hit: torch.Tensor          # (n,)
    y_hit: torch.Tensor          # (n, 6)
    y_last: torch.Tensor         # (n, 6)


def rk_fixed_event(mu, y0, t0, tf, steps: int, *, order: int = 8, t_vals=None,
                   ev_index: int = 1, ev_offset: float = 0.0, direction: int = 0,
                   xtol: float = 1e-12, gtol: float = 1e-12, forward: int = 1,
                   flip_indices=None, block: int = 0, stream=None) -> FixedEventResult:
    """Fixed-step RK (order 4, 6 or 8) until the first crossing of
    y[ev_index] = ev_offset, refined by bisection on the step's cubic Hermite
    interpolant (rk.py:655-718, 330-391); grid as rk_fixed.  Without a
    crossing ``t_hit`` is the last grid time and ``y_hit`` the last state."""
    if order not in FIXED_ORDERS:
        raise ValueError("RK order must be 4, 6, or 8")
    lib = _lib.load()
    y = _soa(y0, 6)
    n = y.shape[1]
    mu_t = _vec(mu, n, "mu")
    tv = None
    if t_vals is not None:
        tv = _vec(t_vals, len(t_vals), "t_vals")
        steps = tv.shape[0]
        t0_t = tf_t = None
    else:
        t0_t, tf_t = _vec(t0, n, "t0"), _vec(tf, n, "tf")
    hit = torch.empty(n, dtype=torch.int32, device=y.device)
    th = torch.empty(n, dtype=torch.float64, device=y.device)
    yh, yl = torch.empty_like(y), torch.empty_like(y)
    rc = lib.hh_rk_fixed_event(int(order), n, _ptr(mu_t), _ptr(t0_t), _ptr(tf_t), int(steps),
                               _ptr(tv), int(forward), flip_mask(6, flip_indices), int(ev_index),
                               float(ev_offset), int(direction), float(xtol), float(gtol), _ptr(y),
                               _ptr(hit), _ptr(th), _ptr(yh), _ptr(yl), int(block),
                               _lib.stream_handle(stream))
    _lib.check(rc, "hh_rk_fixed_event")
    return FixedEventResult(hit.bool(), th, yh.t(), yl.t())


@dataclass
class FinalBatch:
    """Device-resident inputs and outputs of one DOP853 final-value batch."""
    dim: int
    n: int
    mu: torch.Tensor
    t0: torch.Tensor
    tf: torch.Tensor
    y0: torch.Tensor             # (dim, n) structure-of-arrays
    y: torch.Tensor              # (dim, n)
    status: torch.Tensor
    n_accepted: torch.Tensor
    n_rejected: torch.Tensor
    order: Optional[torch.Tensor] = None
    workspace: Optional[torch.Tensor] = None   # grown on demand by launch_final

    def result(self) -> FinalResult:
        return FinalResult(self.y.t(), self.status, self.n_accepted, self.n_rejected)

    def workspace_for(self, variant_code: int):
        """(pointer, bytes) of a workspace large enough for ``variant_code``."""
        need = int(_lib.load().hh_dop853_final_workspace(self.n, self.dim, variant_code))
        if need == 0:
            return None, 0
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.y.device)
        return _ptr(self.workspace), int(self.workspace.numel())


def prepare_final(mu, y0, t0, tf, dim: int = 6, order: str = "span") -> FinalBatch:
    """Upload a batch once (rows -> SoA) and allocate its outputs.

    ``order`` sets the hand-out order of the persistent kernel: "span" gives
    the longest integration spans |tf - t0| first (longest-processing-time
    first, so the slowest trajectories do not start last); "interleave" also
    starts the longest eighth first but puts seven shorter spans beside each
    of them, so that a wave holding a long-running trajectory empties soon;
    "index" keeps the row order.
    """
    y = _soa(y0, dim)
    n = y.shape[1]
    i32 = dict(dtype=torch.int32, device=y.device)
    t0_t, tf_t = _vec(t0, n, "t0"), _vec(tf, n, "tf")
    perm = None
    if order in ("span", "interleave"):
        perm = torch.argsort(torch.abs(tf_t - t0_t), descending=True, stable=True)
        if order == "interleave" and n >= 8:
            # every aligned group of 8 hand-outs: one of the longest eighth of
            # the spans, then seven of the rest, both in decreasing span
            k = n // 8
            longest, rest = perm[:k], perm[k:]
            head = torch.cat([longest[:, None], rest[:7 * k].reshape(k, 7)], dim=1).reshape(-1)
            perm = torch.cat([head, rest[7 * k:]])
        perm = perm.to(torch.int32)
    elif order not in ("index", None):
        raise ValueError(f"unknown order {order!r}")
    return FinalBatch(dim, n, _vec(mu, n, "mu"), t0_t, tf_t, y, torch.empty_like(y),
                      torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(n, **i32), perm)


def launch_final(b: FinalBatch, *, rtol: float = 1e-12, atol: float = 1e-12,
                 max_step: float = 1e4, min_step: float = MIN_STEP,
                 max_steps: int = DEFAULT_MAX_STEPS, forward: int = 1, flip_indices=None,
                 variant: str = "plain", block: int = 0, grid: int = 0, stream=None) -> None:
    """Enqueue hh_dop853_final on the stream (asynchronous).

    ``variant``: "plain", "persistent" (dim 42), "persistent_pair" (dim 42),
    "wide", "pair" (dim 42), or "auto" ("wide" for dim 6 up to 65,536
    trajectories, else "plain").
    """
    lib = _lib.load()
    if variant == "auto":
        variant = "wide" if b.dim == 6 and b.n <= 65536 else "plain"
    var = {"plain": _lib.HH_VARIANT_PLAIN, "persistent": _lib.HH_VARIANT_PERSISTENT,
           "wide": _lib.HH_VARIANT_WIDE, "pair": _lib.HH_VARIANT_PAIR,
           "persistent_pair": _lib.HH_VARIANT_PERSISTENT_PAIR}[variant]
    ws, ws_bytes = b.workspace_for(var)
    rc = lib.hh_dop853_final(b.n, b.dim, _ptr(b.mu), _ptr(b.t0), _ptr(b.tf), float(rtol),
                             float(atol), float(max_step), float(min_step), int(max_steps),
                             int(forward), flip_mask(b.dim, flip_indices), _ptr(b.y0), _ptr(b.y),
                             _ptr(b.status), _ptr(b.n_accepted), _ptr(b.n_rejected), var,
                             int(block), int(grid), ws, ws_bytes, _ptr(b.order),
                             _lib.stream_handle(stream))
    _lib.check(rc, "hh_dop853_final")


class FinalGraph:
    """The launch sequence of one hh_dop853_final call on FinalBatch ``b``
    (the workspace reset kernels and the integration kernel) captured in a
    HIP graph: ``replay()`` re-runs it with one graph launch on the current
    stream, on the batch's current inputs.  The scheduling knobs in effect at
    capture time are the graph's.  The capture contains kernels only (the
    library resets its counters with a fill kernel, not a memset)."""

    def __init__(self, b: "FinalBatch", **launch_kw):
        import torch
        self.b = b
        self.launch_kw = dict(launch_kw)
        self.launch_kw.pop("stream", None)
        # one eager call first: allocates the workspace and caches the
        # library's device queries outside the capture
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            launch_final(b, **self.launch_kw)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            launch_final(b, **self.launch_kw)
        torch.cuda.synchronize()

    def replay(self) -> None:
        self.graph.replay()


def capture_final(b: "FinalBatch", **launch_kw) -> FinalGraph:
    """Capture ``launch_final(b, **launch_kw)`` in a HIP graph."""
    return FinalGraph(b, **launch_kw)


def dop853_final(mu, y0, t0, tf, *, dim: int = 6, rtol: float = 1e-12, atol: float = 1e-12,
                 max_step: float = 1e4, min_step: float = MIN_STEP,
                 max_steps: int = DEFAULT_MAX_STEPS, forward: int = 1, flip_indices=None,
                 variant: str = "auto", block: int = 0, grid: int = 0,
                 stream=None) -> FinalResult:
    """Adaptive DOP853 to tf; the state returned is the dense-output value at tf
    (the last row of _propagate_dynsys(..., steps=2)).
    """
    b = prepare_final(mu, y0, t0, tf, dim,
                      order="span" if variant.startswith("persistent") else None)
    launch_final(b, rtol=rtol, atol=atol, max_step=max_step, min_step=min_step,
                 max_steps=max_steps, forward=forward, flip_indices=flip_indices, variant=variant,
                 block=block, grid=grid, stream=stream)
    return b.result()


@dataclass
class GridResult:
    """DOP853 dense output on numpy.linspace(t0, tf, m) per trajectory."""
    states: torch.Tensor         # (n, m, dim)
    status: torch.Tensor
    n_accepted: torch.Tensor
    n_rejected: torch.Tensor


def dop853_grid(mu, y0, t0, tf, m: int, *, dim: int = 6, t_eval=None, rtol: float = 1e-12,
                atol: float = 1e-12, max_step: float = 1e4, min_step: float = MIN_STEP,
                max_steps: int = DEFAULT_MAX_STEPS, forward: int = 1, flip_indices=None,
                variant: str = "auto", block: int = 0, stream=None) -> GridResult:
    """States at the m points of numpy.linspace(t0, tf, m) -- or at the explicit
    times ``t_eval`` shared by all trajectories -- as the reference's
    dense-output phase produces them (rk.py:2497-2543).  ``variant`` "wide"
    (dim 6 only) integrates one trajectory per 8-lane group; "auto" takes it
    for dim 6 up to 65,536 trajectories."""
    lib = _lib.load()
    y = _soa(y0, dim)
    n = y.shape[1]
    te = None
    if t_eval is not None:
        te = _vec(t_eval, len(t_eval), "t_eval")
        m = te.shape[0]
        t0, tf = float(te[0]), float(te[-1])
    mu_t, t0_t, tf_t = _vec(mu, n, "mu"), _vec(t0, n, "t0"), _vec(tf, n, "tf")
    out = torch.empty((n, int(m), dim), dtype=torch.float64, device=y.device)
    i32 = dict(dtype=torch.int32, device=y.device)
    status, nacc, nrej = torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(n, **i32)
    if variant == "auto":
        variant = "wide" if dim == 6 and n <= 65536 else "plain"
    var = {"plain": _lib.HH_VARIANT_PLAIN, "wide": _lib.HH_VARIANT_WIDE}[variant]
    rc = lib.hh_dop853_grid(n, dim, _ptr(mu_t), _ptr(t0_t), _ptr(tf_t), int(m), _ptr(te), float(rtol),
                            float(atol), float(max_step), float(min_step), int(max_steps),
                            int(forward), flip_mask(dim, flip_indices), _ptr(y), _ptr(out),
                            _ptr(status), _ptr(nacc), _ptr(nrej), var, int(block),
                            _lib.stream_handle(stream))
    _lib.check(rc, "hh_dop853_grid")
    return GridResult(out, status, nacc, nrej)


def dop853_event(mu, y0, t0, tmax, *, ev_index: int = 1, ev_offset: float = 0.0,
                 direction: int = 0, rtol: float = 1e-12, atol: float = 1e-12,
                 max_step: float = math.inf, min_step: float = MIN_STEP,
                 max_steps: int = DEFAULT_MAX_STEPS, xtol: float = 1e-12, gtol: float = 1e-12,
                 forward: int = 1, flip_indices=None, variant: str = "auto", block: int = 0,
                 stream=None) -> EventResult:
    """DOP853 until the first crossing of y[ev_index] = ev_offset (rk.py:2678-2803).

    ``variant`` "plain" integrates one trajectory per lane (throughput);
    "wide" one per 8-lane group, one component per lane (about 2.5x shorter
    per trajectory; faster overall up to tens of thousands of trajectories);
    "auto" picks "wide" up to 65,536 trajectories.  All give the same bits."""
    lib = _lib.load()
    y = _soa(y0, 6)
    n = y.shape[1]
    mu_t, t0_t, tm_t = _vec(mu, n, "mu"), _vec(t0, n, "t0"), _vec(tmax, n, "tmax")
    i32 = dict(dtype=torch.int32, device=y.device)
    hit = torch.empty(n, **i32)
    th = torch.empty(n, dtype=torch.float64, device=y.device)
    yh = torch.empty_like(y)
    yl = torch.empty_like(y)
    status, nacc, nrej = torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(n, **i32)
    if variant == "auto":
        variant = "wide" if n <= 65536 else "plain"
    rc = lib.hh_dop853_event(n, _ptr(mu_t), _ptr(t0_t), _ptr(tm_t), float(rtol), float(atol),
                             float(max_step), float(min_step), int(max_steps), int(ev_index),
                             float(ev_offset), int(direction), float(xtol), float(gtol),
                             int(forward), flip_mask(6, flip_indices), _ptr(y), _ptr(hit),
                             _ptr(th), _ptr(yh), _ptr(yl), _ptr(status), _ptr(nacc), _ptr(nrej),
                             {"plain": _lib.HH_VARIANT_PLAIN, "wide": _lib.HH_VARIANT_WIDE}[variant],
                             int(block), _lib.stream_handle(stream))
    _lib.check(rc, "hh_dop853_event")
    return EventResult(hit.bool(), th, yh.t(), yl.t(), status, nacc, nrej)


def eval_rhs(mu, y, dim: int = 6, stream=None) -> torch.Tensor:
    """Right-hand side at each row of y (rtbp.py:30-74 for dim 6, 167-255 for 42)."""
    lib = _lib.load()
    ys = _soa(y, dim)
    n = ys.shape[1]
    mu_t = _vec(mu, n, "mu")
    out = torch.empty_like(ys)
    _lib.check(lib.hh_eval_rhs(n, dim, _ptr(mu_t), _ptr(ys), _ptr(out), _lib.stream_handle(stream)),
               "hh_eval_rhs")
    return out.t()


def eval_math(x, y, stream=None) -> torch.Tensor:
    """(6, n): sqrt(x), x/y, x**1.5, x**2.5, x**y as the kernels compute them
    and the kernels' own square root (hh_dev_sqrt) of x."""
    lib = _lib.load()
    xt = _vec(x, len(x), "x")
    yt = _vec(y, len(x), "y")
    n = xt.shape[0]
    out = torch.empty((6, n), dtype=torch.float64, device=xt.device)
    _lib.check(lib.hh_eval_math(n, _ptr(xt), _ptr(yt), _ptr(out), _lib.stream_handle(stream)),
               "hh_eval_math")
    return out
