"""Trajectory container: a time grid and the matching states.

The result type of hiten's propagations and manifolds
(algorithms/types/states.py:613-830, ``Trajectory``): times (N,) strictly
monotonic, states (N, D), N >= 2, stored as float64; integer indexing gives
a (time, state) pair, slicing a Trajectory, iteration yields the pairs.
``from_solution`` wraps an integrator ``_Solution`` without copying
(states.py:806-814).  The state-vector class and frame are carried as given
(their domain classes are outside this package).
"""
from __future__ import annotations

from typing import Iterator, Optional, Tuple, Union

import numpy as np


class Trajectory:
    def __init__(self, times, states, state_vector_cls=None, frame=None):
        t_arr = np.asarray(times, dtype=np.float64)
        x_arr = np.asarray(states, dtype=np.float64)
        if t_arr.ndim != 1:
            raise ValueError("times must be a 1D sequence")
        if x_arr.ndim != 2:
            raise ValueError("states must be a 2D array-like of shape (N, D)")
        if len(t_arr) != len(x_arr):
            raise ValueError(f"times and states must have the same length: "
                             f"{len(t_arr)} != {len(x_arr)}")
        if len(t_arr) < 2:
            raise ValueError("trajectory must contain at least two samples")
        dt = np.diff(t_arr)
        if not (np.all(dt > 0.0) or np.all(dt < 0.0)):
            raise ValueError("times must be strictly monotonic (all increasing or all decreasing)")
        self._times = t_arr
        self._states = x_arr
        self._state_vector_cls = state_vector_cls
        self._frame = frame

    # -- data --------------------------------------------------------------
    @property
    def times(self) -> np.ndarray:
        return self._times

    @property
    def states(self) -> np.ndarray:
        return self._states

    @property
    def n_samples(self) -> int:
        return self._times.shape[0]

    @property
    def dim(self) -> int:
        return self._states.shape[1]

    @property
    def t0(self) -> float:
        return float(self._times[0])

    @property
    def tf(self) -> float:
        return float(self._times[-1])

    @property
    def duration(self) -> float:
        return float(self._times[-1] - self._times[0])

    @property
    def state_vector_cls(self):
        return self._state_vector_cls

    @property
    def frame(self):
        return self._frame

    def as_arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._times, self._states

    def to_arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._times, self._states

    # -- views -------------------------------------------------------------
    def with_state_vector(self, cls) -> "Trajectory":
        return Trajectory(self._times, self._states, state_vector_cls=cls, frame=self._frame)

    def with_frame(self, frame) -> "Trajectory":
        return Trajectory(self._times, self._states, state_vector_cls=self._state_vector_cls,
                          frame=frame)

    def vector_at(self, index: int, *, copy: bool = False):
        if self._state_vector_cls is None:
            raise ValueError("No state_vector_cls configured. Use with_state_vector(...) first.")
        return self._state_vector_cls(self._states[index], copy=copy, frame=self._frame)

    def iter_vectors(self, *, copy: bool = False):
        if self._state_vector_cls is None:
            raise ValueError("No state_vector_cls configured. Use with_state_vector(...) first.")
        for i in range(self.n_samples):
            yield self._state_vector_cls(self._states[i], copy=copy, frame=self._frame)

    def reversed(self) -> "Trajectory":
        return Trajectory(self._times[::-1].copy(), self._states[::-1].copy())

    # -- sequence protocol -------------------------------------------------
    def __len__(self) -> int:
        return self.n_samples

    def __iter__(self) -> Iterator[Tuple[float, np.ndarray]]:
        for i in range(self.n_samples):
            yield float(self._times[i]), self._states[i]

    def __getitem__(self, key: Union[int, slice]):
        if isinstance(key, (int, np.integer)):
            return float(self._times[key]), self._states[key]
        if isinstance(key, slice):
            return Trajectory(self._times[key], self._states[key],
                              state_vector_cls=self._state_vector_cls, frame=self._frame)
        raise TypeError(f"Trajectory indices must be integers or slices, not {type(key).__name__}")

    def __repr__(self) -> str:
        direction = "increasing" if self._times[1] > self._times[0] else "decreasing"
        return (f"Trajectory(N={self.n_samples}, D={self.dim}, t0={self.t0:.6g}, "
                f"tf={self.tf:.6g}, {direction})")

    # -- constructors -------------------------------------------------------
    @classmethod
    def from_arrays(cls, times, states, *, state_vector_cls=None,
                    frame=None) -> "Trajectory":
        return cls(times, states, state_vector_cls=state_vector_cls, frame=frame)

    @classmethod
    def from_solution(cls, solution, *, state_vector_cls=None,
                      frame=None) -> "Trajectory":
        return cls(solution.times, solution.states, state_vector_cls=state_vector_cls,
                   frame=frame)


__all__ = ["Trajectory"]
