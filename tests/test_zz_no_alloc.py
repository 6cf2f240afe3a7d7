"""pqs_run_batch_noalloc: the evolution entry point that allocates nothing
(instance table in the library's file-scope device array) gives the same
states as pqs_run_batch, bit for bit, on two calls in a row (the second with
a smaller batch), for a fixed-layout and a rolling-layout plan."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _plan(pyqc, n, flags, low):
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(4, n):
        qc.add_layer(l)
    return qc, E.Plan(n, qc.ops(), tile_bits=11, low_bits=low, flags=flags)


@pytest.mark.parametrize("rolling", [False, True])
def test_noalloc_matches_run_batch(pyqc, rolling):
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    n = 13
    qc, plan = _plan(pyqc, n, L.PQS_PLAN_ROLLING if rolling else 0, 3 if rolling else 4)
    assert plan.rolling == rolling and plan.n_sweeps >= 2
    ctx = E.context()
    dev = E.device()
    dim = 1 << n
    lib = L.lib()
    for B in (96, 37):  # the second call is smaller: stale table entries must not matter
        print(f"noalloc rolling={rolling} batch={B}", flush=True)
        th = torch.as_tensor(np.random.default_rng(B).random((B, qc.n_true_params)) * 2 * np.pi, device=dev)
        ref = E.run(plan, th, E.Init.zero())  # pqs_run_batch (also uploads the plan's device image)
        out = torch.full((B, dim), complex(np.nan, np.nan), dtype=torch.complex128, device=dev)
        alt = torch.empty_like(out) if rolling else None
        rc = lib.pqs_run_batch_noalloc(ctx.handle, plan.handle, E.ptr(out), dim, B, E.ptr(th), th.stride(0), 1,
                                       None, E.ptr(alt), None, E.stream_ptr())
        assert rc == 0, lib.pqs_last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        # and from a source state (init 0): the evolution of ref's rows
        src = ref[0].clone()
        out2 = torch.empty_like(out)
        rc = lib.pqs_run_batch_noalloc(ctx.handle, plan.handle, E.ptr(out2), dim, B, E.ptr(th), th.stride(0), 0,
                                       E.ptr(src), E.ptr(alt), None, E.stream_ptr())
        assert rc == 0, lib.pqs_last_error()
        ref2 = E.run(plan, th, E.Init.from_vector(src))
        torch.cuda.synchronize()
        assert torch.equal(out2, ref2)


def test_noalloc_rejects_what_it_cannot_do(pyqc):
    """Over the batch limit, a plan with tangents, a rolling plan without its
    second buffer: clear errors, nothing launched."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    n = 12
    qc, plan = _plan(pyqc, n, L.PQS_PLAN_ROLLING, 3)
    ctx = E.context()
    lib = L.lib()
    dev = E.device()
    th = torch.zeros((1, qc.n_true_params), dtype=torch.float64, device=dev)
    E.run(plan, th, E.Init.zero())
    out = torch.empty((1, 1 << n), dtype=torch.complex128, device=dev)
    rc = lib.pqs_run_batch_noalloc(ctx.handle, plan.handle, E.ptr(out), 1 << n, 16385, E.ptr(th), 0, 1, None,
                                   E.ptr(out), None, E.stream_ptr())
    assert rc == L.PQS_ERR_ARG and b"16384" in lib.pqs_last_error()
    rc = lib.pqs_run_batch_noalloc(ctx.handle, plan.handle, E.ptr(out), 1 << n, 1, E.ptr(th), 0, 1, None, None,
                                   None, E.stream_ptr())
    assert rc == L.PQS_ERR_ARG and b"alt" in lib.pqs_last_error()
    gplan = E.Plan(n, qc.ops(True), tile_bits=11, low_bits=3)
    rc = lib.pqs_run_batch_noalloc(ctx.handle, gplan.handle, E.ptr(out), 1 << n, 1, E.ptr(th), 0, 1, None, None,
                                   None, E.stream_ptr())
    assert rc == L.PQS_ERR_UNSUPPORTED
    fresh = E.Plan(n, qc.ops(), tile_bits=11, low_bits=4)  # never run: no device image yet
    rc = lib.pqs_run_batch_noalloc(ctx.handle, fresh.handle, E.ptr(out), 1 << n, 1, E.ptr(th), 0, 1, None, None,
                                   None, E.stream_ptr())
    assert rc == L.PQS_ERR_ARG and b"run it once" in lib.pqs_last_error()
    torch.cuda.synchronize()
