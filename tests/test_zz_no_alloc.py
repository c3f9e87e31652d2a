"""The no-allocation entry point bqp_propagate_batch_final_dev_noalloc
against bqp_propagate_batch_final_dev, through the C ABI.

Two calls in a row (4,096 rows, then 1,000), each bit-identical to the
allocating entry point on the same inputs (states, status, steps,
evaluations, totals); the device's free memory does not move across the
no-allocation calls; a batch above BQP_NOALLOC_MAX_ROWS is refused before
any device work."""
import ctypes as C

import numpy as np
import pytest

import bqp_amd
from bqp_amd import _native as nat
from bqp_amd import inputs
from bqp_amd.device import DeviceArray, hip, synchronize
from gpu_util import announce

pytestmark = pytest.mark.gpu


def _free_bytes():
    L = hip()
    L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert L.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _call(lib, fn, ctx, bufs, n):
    """One call of `fn`; returns (rc, stats, free device bytes just before and
    just after it, with no host copy in between), then the outputs."""
    y0, t0, tf, bc, yf, st, steps, evals = bufs
    s = nat.Stats()
    vp = lambda b: C.c_void_p(b.data_ptr())  # noqa: E731
    synchronize()
    before = _free_bytes()
    rc = getattr(lib, fn)(ctx, vp(y0), vp(t0), vp(tf), vp(bc), n, vp(yf), vp(st), vp(steps), vp(evals),
                          C.byref(s))
    synchronize()
    after = _free_bytes()
    return rc, s.as_dict(), (before, after), yf.numpy(), st.numpy(), steps.numpy(), evals.numpy()


def test_noalloc_entry_point_matches_allocating_one():
    lib = nat.load()
    w = inputs.workload("w1", n=4096, seed=21)
    field = bqp_amd.kaula(20)
    with bqp_amd.Propagator(field, method="rk87", rtol=1e-11, drag="vallado") as p:
        for k, n in enumerate((4096, 1000)):
            announce(f"no-alloc entry point, call {k + 1}: {n} rows x 1,800 s, degree 20, drag, vs the _dev call")
            tf_h = np.full(n, 1800.0)
            tf_h[::7] = 600.0  # ragged spans
            bufs = [DeviceArray.from_numpy(w["y0"][:n]), DeviceArray.from_numpy(np.zeros(n)),
                    DeviceArray.from_numpy(tf_h), DeviceArray.from_numpy(w["bc"][:n]), DeviceArray((n, 6)),
                    DeviceArray((n,), np.int32), DeviceArray((n,), np.int32), DeviceArray((n,), np.int32)]
            ref = _call(lib, "bqp_propagate_batch_final_dev", p._ctx, bufs, n)
            assert ref[0] == 0, p._lib.bqp_last_error(p._ctx)
            for b in bufs[4:]:  # clear the outputs between the two calls
                b.copy_from(np.zeros(b.shape, dtype=b.dtype))
            got = _call(lib, "bqp_propagate_batch_final_dev_noalloc", p._ctx, bufs, n)
            announce(f"free device bytes around the no-alloc call: {got[2][0]} -> {got[2][1]}")
            assert got[0] == 0, p._lib.bqp_last_error(p._ctx)
            assert got[2][1] == got[2][0]
            assert np.array_equal(got[3], ref[3])
            for a, b in zip(got[4:], ref[4:]):
                assert np.array_equal(a, b)
            assert np.all(got[4] == 0)
            for key in ("steps", "rejects", "evals", "rows_ok", "rows_failed", "first_failed_row"):
                assert got[1][key] == ref[1][key], key
            # the caller may leave the per-row counters to the library
            s = nat.Stats()
            vp = lambda b: C.c_void_p(b.data_ptr())  # noqa: E731
            bufs[4].copy_from(np.zeros((n, 6)))
            synchronize()
            before = _free_bytes()
            rc = lib.bqp_propagate_batch_final_dev_noalloc(p._ctx, vp(bufs[0]), vp(bufs[1]), vp(bufs[2]), vp(bufs[3]),
                                                           n, vp(bufs[4]), None, None, None, C.byref(s))
            synchronize()
            assert rc == 0 and _free_bytes() == before
            assert np.array_equal(bufs[4].numpy(), ref[3]) and s.as_dict()["evals"] == ref[1]["evals"]


def test_noalloc_refuses_oversized_batch():
    announce("no-alloc entry point refuses n > BQP_NOALLOC_MAX_ROWS before any device work")
    lib = nat.load()
    with bqp_amd.Propagator(bqp_amd.kaula(4)) as p:
        one = DeviceArray((1, 6))
        n = 262145
        rc = lib.bqp_propagate_batch_final_dev_noalloc(p._ctx, C.c_void_p(one.data_ptr()), C.c_void_p(one.data_ptr()),
                                                       C.c_void_p(one.data_ptr()), None, n,
                                                       C.c_void_p(one.data_ptr()), None, None, None, None)
        assert rc == nat.E_INVALID
        assert b"262144" in p._lib.bqp_last_error(p._ctx)
