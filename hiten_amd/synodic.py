"""Section crossings of stored trajectories (synodic Poincare maps) on the GPU.

Mirrors the reference's synodic detection backend
(src/hiten/algorithms/poincare/synodic/backend.py:624-887): given sampled
trajectories (times, states), find where they cross the affine section
n . x = offset, refine each crossing (linear, or cubic Hermite + Newton,
optionally on segment_refine sub-intervals), filter by direction,
deduplicate, and return the hits with their 2-d section points.  One
``hh_synodic_detect_ws`` call (two launches) handles a whole batch; trajectories stored by
``hh_dop853_grid`` ([n][m][6] on the device) are read where they lie.

Parity: hit counts and times equal the reference's; hit states agree to an
ulp (the reference's state interpolation squares through the C library's
pow, the kernel multiplies).  The plane value is summed left to right
(exact for coordinate-plane normals; the reference's BLAS dot product order
is implementation-defined for general normals).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib, batch

_IDX = {"x": 0, "y": 1, "z": 2, "vx": 3, "vy": 4, "vz": 5}


def axis_normal(coord) -> np.ndarray:
    """Normal of an axis-aligned plane (_AffinePlaneEvent.axis_plane,
    synodic/events.py:148-196): 1 at the coordinate, 0 elsewhere."""
    if isinstance(coord, str):
        try:
            i = _IDX[coord.lower()]
        except KeyError:
            raise ValueError(f"unknown coordinate {coord!r}") from None
    else:
        i = int(coord)
        if not 0 <= i < 6:
            raise ValueError("coordinate index must be 0..5")
    n = np.zeros(6)
    n[i] = 1.0
    return n


@dataclass
class SectionHit:
    """_SectionHit (poincare/core/types.py:45-86)."""
    time: float
    state: np.ndarray
    point2d: np.ndarray
    trajectory_index: int = 0


@dataclass
class SynodicBackendRequest:
    """The fields of the reference's SynodicBackendRequest that detection reads."""
    trajectories: Sequence[Tuple[np.ndarray, np.ndarray]]
    normal: Sequence[float]
    offset: float = 0.0
    plane_coords: Tuple[str, str] = ("y", "vy")
    interp_kind: str = "linear"
    segment_refine: int = 0
    tol_on_surface: float = 1e-12
    dedup_time_tol: float = 1e-9
    dedup_point_tol: float = 1e-12
    max_hits_per_traj: Optional[int] = None
    newton_max_iter: int = 4
    direction: Optional[int] = None
    trajectory_indices: Optional[Sequence[int]] = None


@dataclass
class SynodicBackendResponse:
    hits: List[List[SectionHit]]
    points: np.ndarray
    states: np.ndarray
    times: Optional[np.ndarray]
    trajectory_indices: np.ndarray
    metadata: dict = field(default_factory=dict)


@dataclass
class HitArrays:
    """Device-side result of one batched detection.  ``relaunch()`` repeats
    the detection into the same buffers (asynchronous; for timing)."""
    count: torch.Tensor      # (n,) int32, hits kept (may exceed cap)
    t: torch.Tensor          # (n, cap)
    x: torch.Tensor          # (n, cap, 6)
    p: torch.Tensor          # (n, cap, 2)
    relaunch: Optional[object] = None


def _plane_indices(plane_coords) -> Tuple[int, int]:
    if len(plane_coords) != 2:
        raise ValueError("plane_coords must name two coordinates")
    return tuple(_IDX[c.lower()] if isinstance(c, str) else int(c) for c in plane_coords)


def detect_arrays(times, states, *, normal, offset: float = 0.0, plane_coords=("y", "vy"),
                  interp_kind: str = "linear", segment_refine: int = 0,
                  tol_on_surface: float = 1e-12, dedup_time_tol: float = 1e-9,
                  dedup_point_tol: float = 1e-12, max_hits_per_traj: Optional[int] = None,
                  newton_max_iter: int = 4, direction: Optional[int] = None, lengths=None,
                  cap: int = 32, stream=None, two_pass: bool = True) -> HitArrays:
    """Detection for n trajectories of m samples: ``states`` (n, m, 6) and
    ``times`` (m,) shared or (n, m), as NumPy arrays or device tensors (read
    in place).  Every hit is kept: when a trajectory has more than ``cap``
    the call is repeated with room for all of them.  ``two_pass`` scans the
    parts of each trajectory with waves of their own and merges them
    (hh_synodic_detect_ws); otherwise one wave walks each trajectory
    (hh_synodic_detect).  Both give the same hits, bit for bit."""
    lib = _lib.load()
    dev = batch._device()
    st = torch.as_tensor(states, dtype=torch.float64, device=dev).contiguous()
    if st.ndim != 3 or st.shape[2] != 6:
        raise ValueError("states must be (n, m, 6)")
    n, m = int(st.shape[0]), int(st.shape[1])
    tm = torch.as_tensor(times, dtype=torch.float64, device=dev).contiguous()
    per = tm.ndim == 2
    if (per and tuple(tm.shape) != (n, m)) or (not per and tm.shape[0] != m):
        raise ValueError("times must be (m,) or (n, m)")
    ln = None
    if lengths is not None:
        ln = torch.as_tensor(lengths, dtype=torch.int32, device=dev).contiguous()
    nrm = np.ascontiguousarray(np.asarray(normal, dtype=np.float64))
    if nrm.shape != (6,) or not np.all(np.isfinite(nrm)):
        raise ValueError("normal must be 6 finite values")
    pi, pj = _plane_indices(plane_coords)
    dirn = 0 if direction is None else int(direction)
    if interp_kind not in ("linear", "cubic"):
        raise ValueError("interp_kind must be 'linear' or 'cubic'")

    if two_pass:
        # workspace of the two-pass form (torch's caching allocator: 256-byte aligned)
        wsb = int(lib.hh_synodic_workspace(n, m))
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)

    def launch(out: HitArrays) -> None:
        args = (n, m, batch._ptr(tm), int(per), batch._ptr(ln), batch._ptr(st),
                nrm.ctypes.data, float(offset), dirn, pi, pj, int(interp_kind == "cubic"),
                int(segment_refine), float(tol_on_surface), float(dedup_time_tol),
                float(dedup_point_tol), int(max_hits_per_traj or 0), int(newton_max_iter),
                int(out.t.shape[1]), batch._ptr(out.count), batch._ptr(out.t),
                batch._ptr(out.x), batch._ptr(out.p))
        if two_pass:
            rc = lib.hh_synodic_detect_ws(*args, ws.data_ptr(), wsb, _lib.stream_handle(stream))
            _lib.check(rc, "hh_synodic_detect_ws")
        else:
            rc = lib.hh_synodic_detect(*args, _lib.stream_handle(stream))
            _lib.check(rc, "hh_synodic_detect")

    while True:
        cap = max(int(cap), 1)
        out = HitArrays(torch.empty(n, dtype=torch.int32, device=dev),
                        torch.empty((n, cap), dtype=torch.float64, device=dev),
                        torch.empty((n, cap, 6), dtype=torch.float64, device=dev),
                        torch.empty((n, cap, 2), dtype=torch.float64, device=dev))
        launch(out)
        most = int(out.count.max().item()) if n else 0
        if most <= cap:
            out.relaunch = lambda: launch(out)
            return out
        cap = most


def hits_from_arrays(res: HitArrays, trajectory_indices=None) -> List[List[SectionHit]]:
    cnt = res.count.cpu().numpy()
    t, x, p = res.t.cpu().numpy(), res.x.cpu().numpy(), res.p.cpu().numpy()
    idx = np.arange(len(cnt)) if trajectory_indices is None else np.asarray(trajectory_indices)
    return [[SectionHit(float(t[i, h]), x[i, h].copy(), p[i, h].copy(), int(idx[i]))
             for h in range(int(cnt[i]))] for i in range(len(cnt))]


class _SynodicDetectionBackend:
    """The reference's detection backend on the GPU (backend.py:624-887)."""

    def detect_on_trajectory(self, times, states, *, normal, offset: float = 0.0,
                             plane_coords=("y", "vy"), interp_kind: str = "linear",
                             segment_refine: int = 0, tol_on_surface: float = 1e-12,
                             dedup_time_tol: float = 1e-9, dedup_point_tol: float = 1e-12,
                             max_hits_per_traj: Optional[int] = None,
                             newton_max_iter: int = 4, direction: Optional[int] = None,
                             trajectory_index: int = 0) -> List[SectionHit]:
        times = np.asarray(times, dtype=float)
        states = np.asarray(states, dtype=float)
        if times.size < 2:
            return []
        res = detect_arrays(times, states[None], normal=normal, offset=offset,
                            plane_coords=plane_coords, interp_kind=interp_kind,
                            segment_refine=segment_refine, tol_on_surface=tol_on_surface,
                            dedup_time_tol=dedup_time_tol, dedup_point_tol=dedup_point_tol,
                            max_hits_per_traj=max_hits_per_traj,
                            newton_max_iter=newton_max_iter, direction=direction)
        return hits_from_arrays(res, [trajectory_index])[0]

    def run(self, request: SynodicBackendRequest) -> SynodicBackendResponse:
        """All trajectories of the request in one launch (padded to the
        longest; each trajectory's own length bounds its segments)."""
        trajs = list(request.trajectories)
        idx = list(request.trajectory_indices) if request.trajectory_indices is not None \
            else list(range(len(trajs)))
        hits: List[List[SectionHit]] = []
        if trajs:
            lens = np.array([len(t) for t, _ in trajs], dtype=np.int32)
            m = max(int(lens.max()), 2)
            T = np.zeros((len(trajs), m))
            X = np.zeros((len(trajs), m, 6))
            for i, (t, s) in enumerate(trajs):
                T[i, :len(t)] = t
                X[i, :len(t)] = s
            res = detect_arrays(T, X, lengths=lens, normal=request.normal,
                                offset=request.offset, plane_coords=request.plane_coords,
                                interp_kind=request.interp_kind,
                                segment_refine=request.segment_refine,
                                tol_on_surface=request.tol_on_surface,
                                dedup_time_tol=request.dedup_time_tol,
                                dedup_point_tol=request.dedup_point_tol,
                                max_hits_per_traj=request.max_hits_per_traj,
                                newton_max_iter=request.newton_max_iter,
                                direction=request.direction)
            hits = hits_from_arrays(res, idx)
        flat = [h for hs in hits for h in hs]
        return SynodicBackendResponse(
            hits=hits,
            points=np.array([h.point2d for h in flat]) if flat else np.empty((0, 2)),
            states=np.array([h.state for h in flat]) if flat else np.empty((0, 6)),
            times=np.array([h.time for h in flat]) if flat else None,
            trajectory_indices=np.array([h.trajectory_index for h in flat], dtype=int)
            if flat else np.empty((0,), dtype=int))
