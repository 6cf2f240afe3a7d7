"""Product-start first sweep (pqs_jit_prod_0): runs whose instances start
from |0...0> or a product table apply every qubit's leading one-qubit gates
to that qubit's 2-vector and generate the tile as a product.  Checked against
the plain first sweep (the same start state passed as a vector, which takes
the plain kernel) and the oracle, for random mixes of R_x, R_y, R_z, fixed
gates, CNOT, CZ and R_zz (diagonal gates folded past CNOT controls, CZ and
R_zz), tile shapes, rolling and fixed layouts, both precisions."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS
from tests.helpers import infidelity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _mixed(pyqc, n, depth, rng):
    """A leading layer of one-qubit gates (some R_z after two-qubit gates on
    the same qubit), then random layers."""
    kinds1 = ["R_x", "R_y", "R_z", "H", "T", "S", "X"]
    kinds2 = ["CNOT", "CZ", "R_zz"]
    spec = []
    for q in range(n):
        spec.append((kinds1[rng.integers(len(kinds1))], q))
    for q in range(0, n - 1, 2):
        spec.append(("CNOT", (q, q + 1)))
        spec.append(("R_z", q))        # diagonal after a CNOT control: folds
        spec.append(("T", q))          # diagonal fixed gate: folds
        spec.append(("R_y", q + 1))    # after a CNOT target: stays in the sweep
    for _ in range(depth):
        for q in range(n):
            spec.append((kinds1[rng.integers(len(kinds1))], q))
        for _ in range(n // 2):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            spec.append((kinds2[rng.integers(len(kinds2))], (a, b)))
    pk, og = [], []
    for name, q in spec:
        args = list(q) if isinstance(q, tuple) else q
        pk.append(getattr(pyqc, name)(args, n))
        og.append(getattr(OC, name)(q, n))
    return pk, og


@pytest.mark.parametrize("dtype,tol,lim", [("complex128", 1e-12, 1e-13), ("complex64", 1e-5, 2e-6)])
def test_prod_start_matches_plain_first_sweep_and_oracle(pyqc, dtype, tol, lim):
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    rng = np.random.default_rng(81)
    n, S = 14, 3
    pk, og = _mixed(pyqc, n, 3, rng)
    zero = np.zeros(2 ** n, dtype=np.complex128)
    zero[0] = 1.0
    plus = OS.product_state(np.array([1, 1]) / np.sqrt(2), n)
    for k, low, fl in ((11, 3, L.PQS_PLAN_ROLLING), (12, 4, 0), (10, 3, L.PQS_PLAN_ROLLING | L.PQS_PLAN_NO_DIRECT)):
        qc = pyqc.PQC(n, jit=True, dtype=dtype, tile_bits=k, low_bits=low, flags=fl)
        qc.add_layer(pk)
        plan = qc.plan(batch=S)
        assert plan.jitted and plan.n_sweeps > 1
        assert "pqs_jit_prod_0" in plan.jit_sweep_sources()[0].decode()
        params = rng.random((S, qc.n_true_params)) * 2 * np.pi
        for start, ref0 in ((None, zero), (T.plus_state(), plus)):
            if start is not None:
                qc.set_initial_state(start)
            got = qc.run_batch(params)                          # product start (|0> or |+> table)
            vec = E.Init.from_vector(torch.as_tensor(ref0, device="cuda"))
            plain = E.run(plan, params, vec)                     # vector start: plain first sweep
            assert (got - plain).abs().max().item() < lim, (k, low, fl)
            got = got.cpu().numpy()
            for s in range(S):
                assert infidelity(got[s], OS.run(og, params[s], n, ref0)) < tol, (k, low, fl, s)


def test_prod_start_on_headline_circuit_single_and_pairs(pyqc):
    """The headline HEA: single runs, batches and pair runs (first sweep
    product-start, last sweep pair kernel) against the oracle."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import templates as T
    n, p, S = 16, 20, 4
    qc = pyqc.PQC(n, jit=True, tile_bits=11, low_bits=3, flags=L.PQS_PLAN_ROLLING)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    og = OC.clifford_he(p, n)
    rng = np.random.default_rng(82)
    params = rng.random((S, qc.n_true_params)) * 2 * np.pi
    st, ov = qc.run_batch_pairs(params)
    st = st.cpu().numpy()
    refs = [OS.run(og, params[s], n) for s in range(S)]
    for s in range(S):
        assert infidelity(st[s], refs[s]) < 1e-12
        assert np.max(np.abs(st[s] - refs[s])) < 1e-12
    for i in range(S // 2):
        assert abs(ov.cpu().numpy()[i] - np.vdot(refs[2 * i], refs[2 * i + 1])) < 1e-12
    one = qc.run(list(params[1])).full().ravel()
    assert np.max(np.abs(one - refs[1])) < 1e-12
