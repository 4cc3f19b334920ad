"""GPU: hh_dop853_final_static (no workspace argument; the persistent forms
take their scratch from the library's per-GPU static array) against
hh_dop853_final with a caller workspace, bit for bit, on two calls in a row
-- the second one smaller, on the array the first one left behind."""
import ctypes as C

import numpy as np
import pytest
import torch

from hiten_amd import _lib, batch, workloads as W

pytestmark = pytest.mark.gpu


def _static_call(w, variant_code):
    lib = _lib.load()
    y = batch._soa(w["y0"], 42)
    n = y.shape[1]
    mu = batch._vec(w["mu"], n, "mu")
    t0 = batch._vec(w["t0"], n, "t0")
    tf = batch._vec(w["tf"], n, "tf")
    out = torch.empty_like(y)
    i32 = dict(dtype=torch.int32, device=y.device)
    st, na, nr = torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(n, **i32)
    order = torch.argsort(torch.abs(tf - t0), descending=True, stable=True).to(torch.int32)
    rc = lib.hh_dop853_final_static(n, 42, mu.data_ptr(), t0.data_ptr(), tf.data_ptr(), 1e-12,
                                    1e-12, 1e4, batch.MIN_STEP, 200_000, 1, 0, y.data_ptr(),
                                    out.data_ptr(), st.data_ptr(), na.data_ptr(), nr.data_ptr(),
                                    variant_code, 0, 0, order.data_ptr(),
                                    _lib.stream_handle())
    _lib.check(rc, "hh_dop853_final_static")
    torch.cuda.synchronize()
    return (out.t().cpu().numpy(), st.cpu().numpy(), na.cpu().numpy(), nr.cpu().numpy())


@pytest.mark.parametrize("variant", ["persistent", "persistent_pair"])
def test_static_scratch_equals_workspace_form(gpu, variant):
    code = {"persistent": _lib.HH_VARIANT_PERSISTENT,
            "persistent_pair": _lib.HH_VARIANT_PERSISTENT_PAIR}[variant]
    old = _lib.set_tuning(heavy_projected=3, min_age=1, pair_projected=3 if variant ==
                          "persistent_pair" else 0)
    try:
        for n in (2048, 700):
            print(f"\n[case] hh_dop853_final_static {variant} n={n}", flush=True)
            w = W.stm_uneven(n=n)
            got = _static_call(w, code)
            r = batch.dop853_final(w["mu"], w["y0"], w["t0"], w["tf"], dim=42, max_step=1e4,
                                   max_steps=200_000, variant=variant)
            want = (r.y.cpu().numpy(), r.status.cpu().numpy(), r.n_accepted.cpu().numpy(),
                    r.n_rejected.cpu().numpy())
            for g, e, what in zip(got, want, ("y", "status", "n_accepted", "n_rejected")):
                assert np.array_equal(g, e), f"{variant} n={n}: {what} differs"
            assert (got[1] == 0).all()
    finally:
        _lib.set_tuning(**old)


def test_static_pair_rejects_batches_beyond_its_scratch(gpu):
    lib = _lib.load()
    n = 16385
    rc = lib.hh_dop853_final_static(n, 42, None, None, None, 1e-12, 1e-12, 1e4, batch.MIN_STEP,
                                    10, 1, 0, None, None, None, None, None,
                                    _lib.HH_VARIANT_PERSISTENT_PAIR, 0, 0, None, None)
    assert rc == _lib.HH_ERR_ARGS
    assert b"HH_STATIC_PAIR_MAX" in lib.hh_last_error()
