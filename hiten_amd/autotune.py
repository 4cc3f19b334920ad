"""Autotuner for the persistent 42-d DOP853 kernels.

Picks the kernel variant and its launch and scheduling knobs by timing passes
of the caller's own batch on the GPU:

- ``variant`` / ``block``: the work-stealing kernel with 4-wave workgroups
  (``persistent``) or the two-wave form with a resume queue
  (``persistent_pair``);
- ``waves_per_simd``: the resident grid and the register budget per lane
  (512 VGPRs at one wave per SIMD, 256 at two);
- ``refill_min``: how many idle groups a wave gathers before it refills;
- ``heavy_projected`` / ``min_age`` / ``pair_projected``: when a long-running
  trajectory leaves the 8-per-wave layout for one wave or a pair of waves;
- ``exclusive_projected``: when such a run takes its SIMD for itself.

Every knob changes scheduling only, so each trial's states, step counts and
statuses are compared bit for bit with the first trial's; a difference raises.
The search is coordinate descent from the defaults (each knob in turn, keeping
the best value), bounded by ``max_trials``.  Results are cached per device and
batch signature in a JSON file (``$HITEN_AMD_CACHE`` or
``~/.cache/hiten_amd/autotune.json``).

No counterpart in the reference (its integrators run on the CPU).
"""
from __future__ import annotations

import hashlib
import json
import os
from dataclasses import asdict, dataclass, field, replace
from typing import Callable, Dict, List, Optional, Sequence

# knobs in search order, with their candidate values (defaults first)
# (the pair-tail form's own knobs come before the variant: measured first,
# they are not skipped when the other form wins its first comparison by noise)
SPACE: Dict[str, Sequence] = {
    "heavy_projected": (2000, 1500, 2500, 3000),
    "min_age": (256, 128, 64),
    "pair_projected": (2000, 0, 6000),
    "refill_min": (4, 2, 1),
    "waves_per_simd": (2, 1),
    "variant": ("persistent_pair", "persistent"),
    "exclusive_projected": (0, 6000, 3000),
}

_TUNING_KEYS = ("heavy_projected", "min_age", "pair_projected", "refill_min", "waves_per_simd",
                "exclusive_projected")
_BLOCK = {"persistent": 256, "persistent_pair": 128}
# knobs a variant does not read (their trials are skipped)
_UNUSED = {"persistent": {"pair_projected", "refill_min", "waves_per_simd",
                          "exclusive_projected"}}


@dataclass(frozen=True)
class Config:
    """One point of SPACE.  The defaults are the library's, the best mean
    pass time measured on the 262,144-trajectory headline batch (20 passes
    per configuration: 94.9 ms against 100.3 ms without pairs and 100.9 ms
    for the 4-wave kernel)."""
    variant: str = "persistent_pair"
    heavy_projected: int = 2000
    min_age: int = 256
    pair_projected: int = 2000
    refill_min: int = 4
    waves_per_simd: int = 2
    exclusive_projected: int = 0

    @property
    def block(self) -> int:
        return _BLOCK[self.variant]

    def tuning(self) -> dict:
        return {k: getattr(self, k) for k in _TUNING_KEYS}


@dataclass
class Trial:
    config: Config
    seconds: float


@dataclass
class TuneResult:
    best: Config
    best_seconds: float
    trials: List[Trial] = field(default_factory=list)
    cached: bool = False

    def as_dict(self) -> dict:
        return {"best": asdict(self.best), "best_seconds": self.best_seconds,
                "cached": self.cached,
                "trials": [dict(asdict(t.config), seconds=t.seconds) for t in self.trials]}


def coordinate_descent(measure: Callable[[Config], float], start: Config = Config(),
                       space: Dict[str, Sequence] = SPACE,
                       max_trials: int = 20) -> TuneResult:
    """Minimise ``measure`` over ``space`` one knob at a time from ``start``.

    Knobs a variant does not read are skipped; a configuration is measured at
    most once.  Stops after ``max_trials`` measurements."""
    seen: Dict[Config, float] = {}
    trials: List[Trial] = []

    def m(c: Config) -> float:
        if c not in seen:
            seen[c] = float(measure(c))
            trials.append(Trial(c, seen[c]))
        return seen[c]

    best = start
    best_t = m(best)
    for knob, values in space.items():
        if knob in _UNUSED.get(best.variant, ()):
            continue
        for v in values:
            if len(trials) >= max_trials:
                return TuneResult(best, best_t, trials)
            c = replace(best, **{knob: v})
            if c in seen and c != best:
                continue
            t = m(c)
            if t < best_t:
                best, best_t = c, t
    return TuneResult(best, best_t, trials)


# --------------------------------------------------------------------------
# GPU measurement on a FinalBatch
# --------------------------------------------------------------------------

def _cache_path() -> str:
    base = os.environ.get("HITEN_AMD_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "hiten_amd")
    return os.path.join(base, "autotune.json")


def signature(b, step_kw: dict) -> str:
    """Cache key: device, batch size and a digest of the spans, mu and the
    step parameters (the schedule depends on all of them)."""
    import torch
    dev = torch.cuda.get_device_name(b.y.device) if torch.cuda.is_available() else "cpu"
    h = hashlib.sha1()
    for t in (b.t0, b.tf, b.mu):
        h.update(t.detach().cpu().numpy().tobytes())
    h.update(json.dumps(step_kw, sort_keys=True).encode())
    return f"{dev}|dim={b.dim}|n={b.n}|{h.hexdigest()[:16]}"


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, cfg: Config, seconds: float) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_cache()
        data[key] = {"config": asdict(cfg), "seconds": seconds}
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def apply(cfg: Config) -> dict:
    """Set the process-wide tuning of ``cfg``; returns the previous one."""
    from . import _lib
    return _lib.set_tuning(**cfg.tuning())


def launch(b, cfg: Config, **step_kw) -> None:
    """hh_dop853_final of ``b`` with ``cfg`` (tuning applied by the caller)."""
    from . import batch
    batch.launch_final(b, variant=cfg.variant, block=cfg.block, **step_kw)


def tune_final(b, *, reps: int = 2, max_trials: int = 16, use_cache: bool = True,
               start: Optional[Config] = None, **step_kw) -> TuneResult:
    """Autotune the persistent kernels on FinalBatch ``b`` (dim 42).

    Each trial: apply the configuration, one untimed pass, then ``reps``
    passes timed with HIP events on the current stream; the trial's time is
    their median.  Leaves the best configuration applied."""
    import torch

    from . import _lib

    if b.dim != 42:
        raise ValueError("the persistent kernels integrate the 42-d system")
    key = signature(b, step_kw)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            cfg = Config(**hit["config"])
            apply(cfg)
            return TuneResult(cfg, float(hit["seconds"]), [], cached=True)

    saved = _lib.get_tuning()
    ref: Dict[str, object] = {}

    def measure(cfg: Config) -> float:
        apply(cfg)
        launch(b, cfg, **step_kw)
        torch.cuda.synchronize()
        state = (b.y.clone(), b.n_accepted.clone(), b.n_rejected.clone(), b.status.clone())
        if not ref:
            ref["state"] = state
        elif not all(torch.equal(u, v) for u, v in zip(state, ref["state"])):
            raise RuntimeError(f"autotune: results of {cfg} differ from the first trial's")
        times = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            launch(b, cfg, **step_kw)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        times.sort()
        return times[len(times) // 2]

    try:
        res = coordinate_descent(measure, start or Config(), max_trials=max_trials)
    except BaseException:
        _lib.set_tuning(**saved)
        raise
    apply(res.best)
    if use_cache:
        _store_cache(key, res.best, res.best_seconds)
    return res
