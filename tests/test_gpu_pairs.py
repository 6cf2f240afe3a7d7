"""Batched evolution with the overlaps of instance pairs (pqs_run_batch_pairs):
the states equal pqs_run_batch's, the overlaps equal pqs_inner's read-back of
those states and the oracle's, for specialised plans whose last sweep forms
the overlaps in registers (direct and LDS-staged final stores, rolling and
fixed layouts, both precisions) and for the plans that read the states back
(interpreted, persistent)."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _hea(pyqc, n, p, **kw):
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n, **kw)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    return qc


def _check(qc, params, tol_state, tol_ov, fused):
    """run_batch_pairs against run_batch + inner_pairs; returns overlaps."""
    from pyramaterised_amd import engine as E
    S = params.shape[0]
    plan = qc.plan(batch=S)
    srcs = [x.decode() for x in plan.jit_sweep_sources()] if plan.jitted else []
    assert any("pqs_jit_pair_" in x for x in srcs) == fused
    st, ov = qc.run_batch_pairs(params)
    ref_st = qc.run_batch(params)
    ref_ov = E.inner_pairs(ref_st, np.arange(0, S, 2), np.arange(1, S, 2))
    d_st = (st - ref_st).abs().max().item()
    d_ov = (ov - ref_ov).abs().max().item()
    assert d_st <= tol_state, d_st
    assert d_ov <= tol_ov, d_ov
    return st.cpu().numpy(), ov.cpu().numpy()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_pairs_fused_shapes_match_readback_and_oracle(pyqc, dtype):
    from pyramaterised_amd import _lib as L
    n, p, S = 16, 6, 8
    rng = np.random.default_rng(71)
    og = OC.clifford_he(p, n)
    c128 = dtype == "complex128"
    tol_ov = 1e-13 if c128 else 1e-6
    shapes = [(11, 3, L.PQS_PLAN_ROLLING), (12, 4, L.PQS_PLAN_ROLLING), (12, 4, 0),
              (12, 4, L.PQS_PLAN_ROLLING | L.PQS_PLAN_NO_DIRECT), (12, 4, L.PQS_PLAN_NO_DIRECT)]
    for k, low, fl in shapes:
        qc = _hea(pyqc, n, p, dtype=dtype, jit=True, tile_bits=k, low_bits=low, flags=fl)
        params = rng.random((S, qc.n_true_params)) * 2 * np.pi
        # the same sweep code apart from the overlap: equal states
        st, ov = _check(qc, params, 1e-14 if c128 else 1e-6, tol_ov, fused=True)
        for i in range(0, S, 2):
            a = OS.run(og, params[i], n)
            b = OS.run(og, params[i + 1], n)
            want = np.vdot(a, b)
            assert abs(ov[i // 2] - want) < (1e-12 if c128 else 2e-5), (k, low, fl, i)


def test_pairs_readback_paths_and_identical_rows(pyqc):
    """Interpreted and persistent plans read the states back; a pair of
    equal rows has overlap |psi|^2 = 1."""
    from pyramaterised_amd import _lib as L
    n, p, S = 14, 4, 6
    rng = np.random.default_rng(72)
    for kw, fused in ((dict(jit=False), False),
                      (dict(jit=True, tile_bits=12, low_bits=3, flags=L.PQS_PLAN_JIT_PERSISTENT), False),
                      (dict(jit=True, tile_bits=11, low_bits=3, flags=L.PQS_PLAN_ROLLING), True)):
        qc = _hea(pyqc, n, p, **kw)
        params = rng.random((S, qc.n_true_params)) * 2 * np.pi
        params[3] = params[2]
        _, ov = _check(qc, params, 1e-14, 1e-13, fused=fused)
        assert abs(ov[1] - 1.0) < 1e-12, ov[1]


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 2e-5)])
def test_pairs_single_sweep_plan(pyqc, dtype, tol):
    """A circuit that fits one tile: the pair kernel is also the first sweep
    (it generates |0...0> itself)."""
    n, p, S = 10, 5, 6
    qc = _hea(pyqc, n, p, dtype=dtype, jit=True, tile_bits=10, low_bits=3)
    plan = qc.plan(batch=S)
    assert plan.n_sweeps == 1
    og = OC.clifford_he(p, n)
    params = np.random.default_rng(74).random((S, qc.n_true_params)) * 2 * np.pi
    st, ov = _check(qc, params, 1e-14 if dtype == "complex128" else 1e-6, 1e-13 if dtype == "complex128" else 1e-6,
                    fused=True)
    for i in range(0, S, 2):
        want = np.vdot(OS.run(og, params[i], n), OS.run(og, params[i + 1], n))
        assert abs(ov[i // 2] - want) < tol


def test_pairs_initial_states_and_errors(pyqc):
    """Product (|+>) and vector initial states through the fused last sweep;
    an odd batch is refused."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import templates as T
    from pyramaterised_amd.qobj import Ket
    n, p, S = 14, 4, 4
    og = OC.clifford_he(p, n)
    rng = np.random.default_rng(73)
    plus = OS.product_state(np.array([1, 1]) / np.sqrt(2), n)
    qc = _hea(pyqc, n, p, jit=True, tile_bits=11, low_bits=3, flags=L.PQS_PLAN_ROLLING)
    params = rng.random((S, qc.n_true_params)) * 2 * np.pi
    qc.set_initial_state(T.plus_state())
    _, ov = _check(qc, params, 1e-14, 1e-13, fused=True)
    for i in range(0, S, 2):
        want = np.vdot(OS.run(og, params[i], n, plus), OS.run(og, params[i + 1], n, plus))
        assert abs(ov[i // 2] - want) < 1e-12
    v = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    v /= np.linalg.norm(v)
    qc.initial_state = Ket(torch.as_tensor(v, device="cuda"), n)
    _, ov = _check(qc, params, 1e-14, 1e-13, fused=True)
    for i in range(0, S, 2):
        want = np.vdot(OS.run(og, params[i], n, v), OS.run(og, params[i + 1], n, v))
        assert abs(ov[i // 2] - want) < 1e-12
    with pytest.raises(ValueError):
        qc.run_batch_pairs(params[:3])
