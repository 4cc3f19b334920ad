"""GPU: the persistent pass replayed from a HIP graph, and the autotuner on a
real batch.  Scheduling never changes results, so both are checked bit for
bit against an eager launch."""
import sys

import numpy as np
import pytest
import torch

from hiten_amd import _lib, autotune, batch, workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    saved = _lib.get_tuning()
    yield
    _lib.set_tuning(**saved)


@pytest.mark.parametrize("variant", ["persistent", "persistent_pair", "plain"])
def test_graph_replay_matches_eager(gpu, variant):
    """Replays interleaved with eager work on other buffers; every replay is
    checked against the eager result."""
    say(f"FinalGraph {variant}: 1024 x 42-d, three replays")
    w = W.stm_uneven(n=1024)
    kw = dict(variant=variant, max_step=1e4)
    ref = batch.dop853_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42, **kw)
    want = (ref.y.clone(), ref.n_accepted.clone(), ref.status.clone())
    b = batch.prepare_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42,
                            order="span" if variant != "plain" else None)
    g = batch.capture_final(b, **kw)
    other = W.stm_uneven(n=256)
    for rep in range(3):
        b.y.zero_()
        b.n_accepted.zero_()
        g.replay()
        torch.cuda.synchronize()
        got = (b.y.t().contiguous(), b.n_accepted, b.status)
        assert torch.equal(got[0], want[0].contiguous()), f"replay {rep}: states differ"
        assert torch.equal(got[1], want[1]) and torch.equal(got[2], want[2]), f"replay {rep}"
        # eager work between replays (another batch, and torch ops)
        batch.dop853_final(other["mu"], other["y0"], other["t0"], other["tf"], dim=42,
                           variant="plain", max_step=1e4)
        torch.cuda.synchronize()


def test_graph_replays_new_inputs(gpu):
    """A replay reads the batch's current inputs: new initial states written
    into the captured buffers give the new trajectories' results."""
    say("FinalGraph persistent_pair: inputs replaced between replays")
    w = W.stm_uneven(n=512)
    w2 = W.stm_uneven(n=512, seed=7)
    b = batch.prepare_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42, order=None)
    g = batch.capture_final(b, variant="persistent_pair", max_step=1e4)
    b.y0.copy_(batch._soa(w2["y0"], 42))
    b.tf.copy_(torch.as_tensor(w2["tf"], device=b.tf.device))
    g.replay()
    torch.cuda.synchronize()
    ref = batch.dop853_final(w["mu"], w2["y0"], w["t0"], w2["tf"], dim=42, variant="plain",
                             max_step=1e4)
    assert torch.equal(b.y.t(), ref.y)


def test_autotune_on_a_batch(gpu, tmp_path, monkeypatch):
    say("autotune.tune_final: 2048 x 42-d, 6 trials, cache round trip")
    monkeypatch.setenv("HITEN_AMD_CACHE", str(tmp_path))
    w = W.stm_uneven(n=2048)
    b = batch.prepare_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42)
    res = autotune.tune_final(b, reps=1, max_trials=6, max_step=1e4)
    assert len(res.trials) == 6 and not res.cached
    assert res.best_seconds == min(t.seconds for t in res.trials)
    assert _lib.get_tuning() == res.best.tuning()
    again = autotune.tune_final(b, reps=1, max_trials=6, max_step=1e4)
    assert again.cached and again.best == res.best
    # the chosen configuration gives the plain kernel's bits
    autotune.launch(b, res.best, max_step=1e4)
    ref = batch.dop853_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42, variant="plain",
                             max_step=1e4)
    torch.cuda.synchronize()
    assert torch.equal(b.y.t(), ref.y)
