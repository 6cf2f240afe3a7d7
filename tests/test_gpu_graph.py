"""Batched runs captured in a HIP graph (pqs_batch_table + pqs_run_table):
every replay, with eager work and synchronisations in between, equals the
eager evolution of the same parameters."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


@pytest.mark.parametrize("jit,flags,init_kind", [(False, 0, "zero"), (True, 0, "product"), (True, 8, "zero"),
                                                 (True, 8, "vector")])
def test_graph_replays_match_eager(pyqc, jit, flags, init_kind):
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p, S = 14, 4, 8
    qc = pyqc.PQC(n, tile_bits=11, low_bits=3, flags=flags, jit=jit)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    plan = qc.plan(batch=S)
    assert plan.jitted == jit and plan.rolling == bool(flags & 8)
    rng = np.random.default_rng(50)
    if init_kind == "zero":
        init = E.Init.zero()
    elif init_kind == "product":
        init = E.Init.product([np.array([np.cos(0.3 * q), np.sin(0.3 * q)]) for q in range(n)])
    else:
        v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        init = E.Init.from_vector(torch.as_tensor(v / np.linalg.norm(v), device="cuda"))
    P = qc.n_true_params
    g = E.BatchGraph(plan, torch.as_tensor(rng.random((S, P)) * 2 * np.pi, device="cuda"), init)
    other = torch.as_tensor(rng.random((3, P)) * 2 * np.pi, device="cuda")
    for rep in range(3):
        prm = torch.as_tensor(rng.random((S, P)) * 2 * np.pi, device="cuda")
        g.params.copy_(prm)
        got = g.replay().clone()
        E.run(plan, other, init)  # eager work between replays
        torch.cuda.synchronize()
        ref = E.run(plan, prm, init)
        assert torch.max(torch.abs(got - ref)).item() < 1e-13, rep


def test_graph_single_sweep_plan_with_rolling_flag(pyqc):
    """A plan asked for rolling layouts whose ops fit one sweep runs in
    place like a fixed-layout plan: the graph form needs no second buffer,
    and replays equal the eager run."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    n, S = 13, 4
    layer = [pyqc.R_y(q, n) for q in range(2, n)] + [pyqc.CNOT([q, q + 1], n) for q in range(2, n - 1)]
    qc = pyqc.PQC(n, tile_bits=11, low_bits=3, flags=L.PQS_PLAN_ROLLING, jit=True)
    qc.add_layer(layer)
    plan = qc.plan(batch=S)
    assert plan.n_sweeps == 1 and not plan.rolling
    rng = np.random.default_rng(51)
    P = qc.n_true_params
    init = E.Init.zero()
    g = E.BatchGraph(plan, torch.as_tensor(rng.random((S, P)) * 2 * np.pi, device="cuda"), init)
    assert g.alt is None
    for rep in range(2):
        prm = torch.as_tensor(rng.random((S, P)) * 2 * np.pi, device="cuda")
        g.params.copy_(prm)
        got = g.replay().clone()
        torch.cuda.synchronize()
        ref = E.run(plan, prm, init)
        assert torch.max(torch.abs(got - ref)).item() < 1e-13, rep
