"""Host-to-device uploads (pqs_upload, engine.to_device): the bytes land
exactly for every dtype, size and alignment, in stream order behind a busy
stream, beyond the staging ring's buffer count and with the caller reusing
its array at once; a plan first launched on one stream runs correctly on
another stream while its arrays are still being copied on the first."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS
from tests.helpers import infidelity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from pyramaterised_amd import engine
    return engine


def _busy(x, k=12):
    """Queue about k passes over x (hundreds of MB) on the current stream."""
    for _ in range(k):
        x.mul_(1.0000001)


def test_to_device_dtypes_and_sizes(E):
    rng = np.random.default_rng(3)
    for dt in (np.float64, np.float32, np.complex128, np.complex64, np.int64, np.int32):
        for n in (0, 1, 3, 4, 1000, 1 << 20):
            if np.issubdtype(dt, np.complexfloating):
                a = (rng.random(n) + 1j * rng.random(n)).astype(dt)
            elif np.issubdtype(dt, np.integer):
                a = rng.integers(-2 ** 30, 2 ** 30, n).astype(dt)
            else:
                a = (rng.random(n) * 1e6).astype(dt)
            t = E.to_device(a)
            assert t.is_cuda and tuple(t.shape) == a.shape
            assert np.array_equal(t.cpu().numpy(), a), (dt, n)
    m = rng.random((7, 13))
    assert np.array_equal(E.to_device(m).cpu().numpy(), m)
    assert np.array_equal(E.to_device(m[:, ::2]).cpu().numpy(), m[:, ::2])  # strided input
    with pytest.raises(TypeError):
        E.to_device(np.zeros(3, dtype=np.uint16))


def test_upload_unaligned_destinations_and_errors(E):
    from pyramaterised_amd import _lib as L
    lib = L.lib()
    rng = np.random.default_rng(5)
    buf = torch.zeros(8192, dtype=torch.uint8, device="cuda")
    ctx = E.context().handle
    for off, nb in ((0, 16), (1, 1), (3, 17), (8, 40), (16, 4096), (5, 4000), (0, 4095)):
        src = rng.integers(1, 256, nb, dtype=np.uint8)
        buf.zero_()
        E.check(lib.pqs_upload(ctx, C.c_void_p(buf.data_ptr() + off), C.c_void_p(src.ctypes.data), nb,
                               E.stream_ptr()))
        got = buf.cpu().numpy()
        assert np.array_equal(got[off:off + nb], src), (off, nb)
        assert not got[:off].any() and not got[off + nb:].any(), (off, nb)
    src = np.ones(8, dtype=np.uint8)
    E.check(lib.pqs_upload(ctx, C.c_void_p(buf.data_ptr()), C.c_void_p(src.ctypes.data), 0, E.stream_ptr()))
    for args in ((None, C.c_void_p(buf.data_ptr()), C.c_void_p(src.ctypes.data), 8),
                 (ctx, C.c_void_p(0), C.c_void_p(src.ctypes.data), 8),
                 (ctx, C.c_void_p(buf.data_ptr()), C.c_void_p(0), 8),
                 (ctx, C.c_void_p(buf.data_ptr()), C.c_void_p(src.ctypes.data), -1)):
        with pytest.raises(L.EngineError):
            E.check(lib.pqs_upload(*args, E.stream_ptr()))


def test_uploads_behind_a_busy_stream_beyond_the_ring(E):
    s = torch.cuda.Stream()  # default priority
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    outs = []
    with torch.cuda.stream(s):
        _busy(x)
        for i in range(80):  # more uploads than staging buffers, all queued behind the busy work
            a = np.full(1000 + 7 * i, float(i))
            outs.append(E.to_device(a))
            a[:] = -1.0  # the caller reuses its array at once
        big = np.arange(1 << 22, dtype=np.float64)  # larger than any staging buffer so far
        t_big = E.to_device(big)
        big[:] = 0.0
        _busy(x, 2)
    torch.cuda.synchronize()
    for i, t in enumerate(outs):
        assert t.numel() == 1000 + 7 * i and bool((t == float(i)).all()), i
    assert np.array_equal(t_big.cpu().numpy(), np.arange(1 << 22, dtype=np.float64))


def test_plan_first_launched_on_another_stream(E):
    import pyramaterised_amd as pyqc
    from pyramaterised_amd import templates as T
    n, p = 10, 5
    rng = np.random.default_rng(9)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    ca, cb = E.Context(torch.cuda.current_device()), E.Context(torch.cuda.current_device())
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    og = OC.clifford_he(p, n)
    for jit in (False, True):
        qc = pyqc.PQC(n, jit=jit)
        for layer in T.clifford_HE_layers(p, n):
            qc.add_layer(layer)
        th = rng.random(qc.n_true_params) * 2 * np.pi
        plan, init = qc.plan(), qc.init()
        torch.cuda.synchronize()
        with E.using(ca, sa):
            _busy(x)  # the plan's arrays are copied behind this work on sa
            out_a = E.run(plan, th, init)
        with E.using(cb, sb):  # sb must wait for those copies
            out_b = E.run(plan, th, init)
            out_b2 = E.run(plan, th, init)
        torch.cuda.synchronize()
        want = OS.run(og, th, n)
        for out in (out_a, out_b, out_b2):
            assert infidelity(out.cpu().numpy().ravel(), want) < 1e-12, jit
