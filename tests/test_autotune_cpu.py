"""CPU checks of the autotuner's search (no GPU: a synthetic cost model)."""
import pytest

from hiten_amd import autotune as A


def _cost(c: A.Config) -> float:
    # a separable bowl with its minimum away from the defaults
    t = 100.0
    t += {"persistent_pair": 0.0, "persistent": 7.0}[c.variant]
    t += abs(c.heavy_projected - 2500) / 100.0
    t += {64: 3.0, 128: 1.0, 256: 0.0}[c.min_age]
    t += {2: 2.0, 1: 0.0}[c.waves_per_simd]
    t += {1: 1.0, 2: 0.0, 4: 0.5}[c.refill_min]
    t += {0: 0.0, 2000: 2.0, 6000: 4.0}[c.pair_projected]
    return t


def test_coordinate_descent_finds_the_separable_minimum():
    res = A.coordinate_descent(_cost, max_trials=100)
    assert res.best == A.Config("persistent_pair", 2500, 256, 0, 2, 1)
    assert res.best_seconds == pytest.approx(100.0)
    # each configuration is measured once
    assert len({t.config for t in res.trials}) == len(res.trials)


def test_coordinate_descent_respects_the_trial_budget():
    res = A.coordinate_descent(_cost, max_trials=5)
    assert len(res.trials) == 5
    assert res.best_seconds == min(t.seconds for t in res.trials)


def test_knobs_the_variant_does_not_read_are_skipped():
    # the pair form's knobs are tuned before the variant; the 4-wave kernel,
    # which does not read them, is only tried with the values they ended at
    res = A.coordinate_descent(lambda c: 1.0 if c.variant == "persistent" else 2.0,
                               max_trials=100)
    assert res.best.variant == "persistent"
    d = A.Config()
    for t in res.trials:
        if t.config.variant == "persistent":
            assert (t.config.refill_min, t.config.waves_per_simd, t.config.pair_projected) == \
                (d.refill_min, d.waves_per_simd, d.pair_projected)


def test_config_maps_to_tuning_and_block():
    c = A.Config("persistent", 1500, 64, 0, 1, 2)
    assert c.block == 256 and A.Config().block == 128
    assert c.tuning() == {"heavy_projected": 1500, "min_age": 64, "pair_projected": 0,
                          "refill_min": 1, "waves_per_simd": 2, "exclusive_projected": 0}


def test_tuning_is_validated_by_the_library():
    from hiten_amd import _lib
    saved = _lib.get_tuning()
    try:
        _lib.set_tuning(refill_min=4, waves_per_simd=1)
        assert _lib.get_tuning()["refill_min"] == 4
        with pytest.raises(_lib.HipLibraryError):
            _lib.set_tuning(waves_per_simd=3)
        with pytest.raises(ValueError):
            _lib.set_tuning(nonsense=1)
    finally:
        _lib.set_tuning(**saved)
