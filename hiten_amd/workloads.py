"""Seeded synthetic inputs for the propagation workloads.

Each generator draws from ``numpy.random.default_rng(seed)`` in a fixed column
order (documented per function) so that the GPU runs, the parity tests and
the CPU baseline see identical inputs.

Workloads (trajectory counts are the full sizes; pass ``n`` to shrink):

* ``rk8_grid``       1024 planar-ish states near Earth-Moon L1, fixed-step
                     RK8 with 2000 grid points to t = 2*pi, no STM.
* ``stm_uneven``     4096 states as above with an identity-initialised STM
                     (42-d), DOP853 rtol = atol = 1e-12, t_f ~ U(2.5, 3.5).
* ``manifold_fan``   64 seeds x 512 unit directions on S^5 (+-1e-6 offsets)
                     = 32768 states, DOP853 until the first y = 0 crossing
                     with vy < 0 (direction -1), xtol = gtol = 1e-13,
                     t_max ~ U(2, 8).
* ``persistent_262k`` 262144 42-d states, mu drawn per trajectory from
                     {Earth-Moon, Sun-Earth, Saturn-Titan-like 2.528e-5},
                     states rescaled about each system's L1 by the Hill
                     ratio (mu / mu_EM)^(1/3), t_f log-uniform in [0.5, 20].
"""

from __future__ import annotations

import math

import numpy as np
from scipy.optimize import brentq

MU_EM = 0.012150585609624
MU_SET = (0.012150585609624, 3.0035e-6, 2.528e-5)
SEED = 20261202


def l1_position(mu: float) -> float:
    """x of L1: root of dOmega/dx on the x axis between the primaries.

    Same equation as the reference's collinear-point search
    (algorithms/types/services/libration.py:671-713, Brent on dOmega/dx).
    """
    def d_omega_dx(x):
        r1 = abs(x + mu)
        r2 = abs(x - 1.0 + mu)
        return x - (1.0 - mu) * (x + mu) / r1**3 - mu * (x - 1.0 + mu) / r2**3

    return brentq(d_omega_dx, -mu + 1e-3, 1.0 - mu - 1e-6, xtol=1e-15, rtol=4 * np.finfo(float).eps,
                  maxiter=500)


def planar_seeds(rng: np.random.Generator, n: int) -> np.ndarray:
    """x ~ U(0.80, 0.88), z ~ U(-0.2, 0.2), vy ~ U(-0.3, 0.3), y = vx = vz = 0.

    Draw order: x (n), then z (n), then vy (n).
    """
    y0 = np.zeros((n, 6))
    y0[:, 0] = rng.uniform(0.80, 0.88, n)
    y0[:, 2] = rng.uniform(-0.2, 0.2, n)
    y0[:, 4] = rng.uniform(-0.3, 0.3, n)
    return y0


def with_identity_stm(states: np.ndarray) -> np.ndarray:
    """(n, 6) -> (n, 42) rows [I6 row-major | x] (dynamics/rtbp.py:316-318)."""
    n = states.shape[0]
    out = np.zeros((n, 42))
    out[:, :36] = np.eye(6).ravel()
    out[:, 36:] = states
    return out


def rk8_grid(n: int = 1024, seed: int = SEED):
    rng = np.random.default_rng(seed)
    y0 = planar_seeds(rng, n)
    return dict(mu=MU_EM, y0=y0, t0=0.0, tf=2.0 * math.pi, steps=2000)


def stm_uneven(n: int = 4096, seed: int = SEED + 1):
    """Draw order: planar seeds, then t_f ~ U(2.5, 3.5)."""
    rng = np.random.default_rng(seed)
    x = planar_seeds(rng, n)
    tf = rng.uniform(2.5, 3.5, n)
    return dict(mu=MU_EM, y0=with_identity_stm(x), t0=np.zeros(n), tf=tf, rtol=1e-12, atol=1e-12)


def manifold_fan(n_seeds: int = 64, n_dirs: int = 512, seed: int = SEED + 2, eps: float = 1e-6):
    """Draw order: seeds (planar_seeds), then n_dirs/2 Gaussian 6-vectors
    (normalised; each used with + and - sign), then t_max ~ U(2, 8) per state.
    Row order: seed-major, then direction, then sign.
    """
    rng = np.random.default_rng(seed)
    seeds = planar_seeds(rng, n_seeds)
    g = rng.normal(size=(n_dirs // 2, 6))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    dirs = np.empty((n_dirs, 6))
    dirs[0::2] = u
    dirs[1::2] = -u
    y0 = (seeds[:, None, :] + eps * dirs[None, :, :]).reshape(-1, 6)
    tmax = rng.uniform(2.0, 8.0, y0.shape[0])
    return dict(mu=MU_EM, y0=y0, t0=np.zeros(y0.shape[0]), tmax=tmax, ev_index=1, ev_offset=0.0,
                direction=-1, rtol=1e-12, atol=1e-12, xtol=1e-13, gtol=1e-13)


def persistent_262k(n: int = 262144, seed: int = SEED + 3):
    """Draw order: mu index (n), planar seeds, t_f (log-uniform)."""
    rng = np.random.default_rng(seed)
    mu_idx = rng.integers(0, len(MU_SET), n)
    mu = np.asarray(MU_SET)[mu_idx]
    base = planar_seeds(rng, n)
    xl1_em = l1_position(MU_EM)
    x = np.zeros_like(base)
    for k, m in enumerate(MU_SET):
        sel = mu_idx == k
        s = (m / MU_EM) ** (1.0 / 3.0)
        x[sel, 0] = l1_position(m) + (base[sel, 0] - xl1_em) * s
        x[sel, 2] = base[sel, 2] * s
        x[sel, 4] = base[sel, 4] * s
    tf = np.exp(rng.uniform(math.log(0.5), math.log(20.0), n))
    return dict(mu=mu, y0=with_identity_stm(x), t0=np.zeros(n), tf=tf, rtol=1e-12, atol=1e-12)


def serving_stream(duration: float = 1.0, rate: float = 2000.0, seed: int = SEED + 4,
                   correction_fraction: float = 0.7, fan_min: int = 32, fan_max: int = 2048,
                   eps: float = 1e-6, continuation_fraction: float = 0.0,
                   continuation_seeds=None, members: int = 4, window: int = 2,
                   z_step: float = 1e-3, correction_seeds=None,
                   perturbation=(1e-4, 1e-3)):
    """Poisson request stream for the serving loop.  Draw order per request:
    inter-arrival time ~ Exp(rate), u ~ U(0, 1) for the kind (correction if
    u < ``correction_fraction``, continuation if below that plus
    ``continuation_fraction``, else fan), then
      correction:   x ~ U(0.80, 0.88), z ~ U(-0.2, 0.2), vy ~ U(-0.3, 0.3)
                    (a Halo-style single-shooting problem, <= 10 iterations);
                    or, with ``correction_seeds`` (corrected Halo orbits), a
                    seed drawn from them moved by a perturbation of size
                    log-uniform in ``perturbation`` along a Gaussian
                    direction in (x, z, vy): guesses that converge;
      continuation: a seed orbit drawn from ``continuation_seeds`` (corrected
                    Halo orbits), the sign of a natural-parameter step of
                    ``z_step`` in z; ``members`` members, retry window
                    ``window``, at most 4 retries per step;
      fan:          size ~ log-uniform [fan_min, fan_max], one planar seed,
                    then size/2 Gaussian 6-directions (+ and - each, eps
                    offset) and t_max ~ U(2, 8) per trajectory (stop at
                    y = 0 with vy < 0).
    With no continuations the draws are those of the corrections-and-fans
    stream.  Returns a list of dicts with an arrival time ``t`` in seconds."""
    rng = np.random.default_rng(seed)
    if continuation_fraction > 0 and not continuation_seeds:
        raise ValueError("continuations need seed orbits")
    out = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        u = rng.uniform()
        if u < correction_fraction:
            if correction_seeds is not None:
                x0 = np.asarray(correction_seeds[int(rng.integers(len(correction_seeds)))],
                                float).copy()
                size = math.exp(rng.uniform(math.log(perturbation[0]),
                                            math.log(perturbation[1])))
                g = rng.normal(size=3)
                x0[[0, 2, 4]] += size * g / np.linalg.norm(g)
                out.append(dict(t=t, kind="correction", mu=MU_EM, x0=x0))
                continue
            x = rng.uniform(0.80, 0.88)
            z = rng.uniform(-0.2, 0.2)
            vy = rng.uniform(-0.3, 0.3)
            out.append(dict(t=t, kind="correction", mu=MU_EM,
                            x0=np.array([x, 0.0, z, 0.0, vy, 0.0])))
        elif u < correction_fraction + continuation_fraction:
            s = np.asarray(continuation_seeds[int(rng.integers(len(continuation_seeds)))], float)
            sign = 1.0 if rng.uniform() < 0.5 else -1.0
            out.append(dict(t=t, kind="continuation", mu=MU_EM, seed=s.copy(),
                            step=np.array([sign * z_step]), state_indices=[2],
                            members=members, window=window, max_retries=4))
        else:
            size = int(round(math.exp(rng.uniform(math.log(fan_min), math.log(fan_max)))))
            size += size % 2
            s0 = planar_seeds(rng, 1)[0]
            g = rng.normal(size=(size // 2, 6))
            u = g / np.linalg.norm(g, axis=1, keepdims=True)
            dirs = np.empty((size, 6))
            dirs[0::2] = u
            dirs[1::2] = -u
            out.append(dict(t=t, kind="fan", mu=MU_EM, y0=s0[None, :] + eps * dirs,
                            tmax=rng.uniform(2.0, 8.0, size)))
    return out
