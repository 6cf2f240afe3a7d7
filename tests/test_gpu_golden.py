"""GPU results against the committed golden fixtures (tests/golden/*.npz,
written by the oracle) and size-independent checks at the full 28-qubit
size, where no CPU result exists."""
import os

import numpy as np
import pytest

from tests.helpers import infidelity, rel_fro

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def test_qg_golden(pyqc):
    g = load("qg")
    n = 4
    qc = pyqc.PQC(n)
    qc.add_layer([pyqc.fixed_R_y(i, n, np.pi / 4) for i in range(n)])
    qc.add_layer([pyqc.R_z(0, n), pyqc.R_x(1, n), pyqc.R_y(2, n), pyqc.R_z(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    qc.add_layer([pyqc.R_x(0, n), pyqc.R_x(1, n), pyqc.R_x(2, n), pyqc.R_y(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    qc.add_layer([pyqc.R_z(0, n), pyqc.R_x(1, n), pyqc.R_y(2, n), pyqc.R_y(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    th = list(g["params"])
    qc.state = qc.run(th)
    assert np.max(np.abs(qc.state.full().ravel() - g["state"])) < 1e-14
    assert abs(qc.cost(th) - float(g["energy"])) < 1e-14
    F = pyqc.measure.Measurements(qc).get_QFI()
    assert rel_fro(F, g["qfim"]) < 1e-9


@pytest.mark.parametrize("name", ["linear", "chain"])
def test_config1_golden(pyqc, name):
    from pyramaterised_amd import templates as T
    g = load("config1")
    n, p = 4, 3
    qc = pyqc.PQC(n)
    for l in (T.hea_linear_layers(p, n) if name == "linear" else T.clifford_HE_layers(p, n)):
        qc.add_layer(l)
    qc.state = qc.run(list(g[f"{name}_params"]))
    ref = g[f"{name}_state"]
    assert infidelity(qc.state.full(), ref) < 1e-12
    assert np.max(np.abs(qc.state.full().ravel() - ref)) < 1e-13
    m = pyqc.measure.Measurements(qc)
    assert abs(m.single_Q(qc.state, n) - float(g[f"{name}_Q"])) < 1e-13
    assert rel_fro(m.get_QFI(), g[f"{name}_qfim"]) < 1e-9


@pytest.mark.parametrize("n", [4, 6])
def test_npqc_golden(pyqc, n):
    from pyramaterised_amd import templates as T
    g = load("npqc")
    for P in range(1, 2 ** (n // 2) + 1):
        layers, th = T.NPQC_layers(P, n)
        assert np.allclose(th, g[f"n{n}_P{P}_angles"], atol=0)
        qc = pyqc.PQC(n)
        for l in layers:
            qc.add_layer(l)
        qc.state = qc.run(angles=th)
        F = pyqc.measure.Measurements(qc).get_QFI()
        assert np.max(np.abs(F - g[f"n{n}_P{P}_qfim"])) < 1e-12


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_config3_golden(pyqc, dtype, tol):
    from pyramaterised_amd import templates as T
    g = load("config3")
    n, p = 12, 10
    qc = pyqc.PQC(n, dtype=dtype)
    qc.set_initial_state(T.plus_state())
    for l in T.tfim_open_layers(p, n):
        qc.add_layer(l)
    qc.state = qc.run(list(g["params"]))
    assert infidelity(qc.state.full(), g["state"]) < tol
    m = pyqc.measure.Measurements(qc)
    F = m.get_QFI()
    assert rel_fro(F, g["qfim"]) < (1e-9 if dtype == "complex128" else 1e-4)
    if dtype == "complex128":
        assert m.get_effective_quantum_dimension(1e-12) == int(g["eqd"])


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 2e-6)])
def test_config4_22q_golden(pyqc, dtype, tol):
    from pyramaterised_amd import templates as T
    g = load("config4_22q")
    n, p = 22, 30
    qc = pyqc.PQC(n, dtype=dtype)
    for l in T.ry_cz_brickwork_layers(p, n):
        qc.add_layer(l)
    st = qc.run(list(g["params"])).t
    idx = g["sample_idx"]
    amps = st[idx].cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(amps - g["sample_amp"])) < tol
    prob = (st.abs() ** 2).double()
    assert abs(prob.sum().item() - float(g["norm"])) < (1e-12 if dtype == "complex128" else 1e-5)
    blocks = prob.reshape(64, -1).sum(dim=1).cpu().numpy()
    assert np.max(np.abs(blocks - g["block_prob"])) < (1e-12 if dtype == "complex128" else 1e-6)
    if dtype == "complex128":
        assert abs(complex(st.sum().item()) - complex(g["amp_sum"])) < 1e-9


def test_mw_golden(pyqc):
    g = load("mw")
    m = pyqc.measure.Measurements(pyqc.PQC(2))
    assert abs(m.single_Q(pyqc.bell_state("11"), 2) - float(g["bell_Q"])) < 1e-13
    s = g["product_singles"]
    prod = s[0]
    for v in s[1:]:
        prod = np.kron(prod, v)
    from pyramaterised_amd.qobj import to_ket
    assert abs(m.single_Q(to_ket(prod), 5) - float(g["product_Q"])) < 1e-13


def _inverse_ops(ops, L):
    """Ops of U^dagger: reversed order, rotation angles negated, fixed
    matrices conjugate-transposed (CNOT / CZ are self-inverse)."""
    out = []
    for o in reversed(ops):
        o2 = L.Op.from_buffer_copy(o)
        if o.kind in (L.OP_RX, L.OP_RY, L.OP_RZ, L.OP_RZZ, L.OP_RXX, L.OP_RYY):
            o2.scale, o2.offset = -o.scale, -o.offset
        elif o.kind == L.OP_MAT1:
            m = np.array(o.m[:8]).view(np.complex128).reshape(2, 2)
            md = np.conj(m.T).reshape(-1).view(np.float64)
            for i in range(8):
                o2.m[i] = float(md[i])
        elif o.kind not in (L.OP_CX, L.OP_CZ):
            raise NotImplementedError(o.kind)
        out.append(o2)
    return out


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_config4_28q_roundtrip_and_norm(pyqc, dtype):
    """Full 28-qubit size: the state has unit norm, and U^dagger U |0..0>
    returns |0..0> (amplitude 0 -> 1, the rest -> 0)."""
    import torch
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p = 28, 30
    qc = pyqc.PQC(n, dtype=dtype, tile_bits=13, low_bits=3)
    for l in T.ry_cz_brickwork_layers(p, n):
        qc.add_layer(l)
    th = np.random.default_rng(28).random(qc.n_true_params) * 2 * np.pi
    st = qc.run(list(th)).t
    tol = 1e-12 if dtype == "complex128" else 1e-5
    nrm = torch.linalg.vector_norm(st.reshape(-1)).item()
    assert abs(nrm - 1) < tol
    inv = E.Plan(n, _inverse_ops(qc.ops(), L), dtype=E.dtype_code(dtype), tile_bits=13, low_bits=3).jit()
    back = E.run(inv, torch.as_tensor(th, device="cuda"), E.Init.from_vector(st)).reshape(-1)
    a0 = complex(back[0].item())
    assert abs(a0 - 1) < (1e-11 if dtype == "complex128" else 2e-4)
    rest = torch.linalg.vector_norm(back[1:]).item()
    assert rest < (1e-7 if dtype == "complex128" else 1e-2)
    del st, back
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_28q_closed_form_product_state(pyqc, dtype):
    """Full 28-qubit size against a closed form: 30 layers of R_y on every
    qubit, each followed by the brickwork CZs applied twice (identity), give
    the product state (x)_q R_y(A_q)|0> with A_q the sum of qubit q's angles;
    sampled amplitudes and the 64 top-6-bit block probabilities are
    compared.  Pins qubit order and angle binding at full size."""
    import torch
    n, p = 28, 30
    qc = pyqc.PQC(n, dtype=dtype, tile_bits=13, low_bits=3)
    for layer in range(p):
        cz = [pyqc.CZ([i, i + 1], n) for i in range(layer % 2, n - 1, 2)]
        qc.add_layer([pyqc.R_y(q, n) for q in range(n)] + cz + [pyqc.CZ([i, i + 1], n) for i in range(layer % 2, n - 1, 2)])
    th = np.random.default_rng(29).random(qc.n_true_params) * 2 * np.pi
    st = qc.run(list(th)).t.reshape(-1)
    A = th.reshape(p, n).sum(axis=0)
    c, s = np.cos(A / 2), np.sin(A / 2)
    rng = np.random.default_rng(30)
    idx = np.sort(rng.choice(2 ** n, 4096, replace=False))
    bits = (idx[:, None] >> (n - 1 - np.arange(n))[None, :]) & 1
    ref = np.prod(np.where(bits == 1, s[None, :], c[None, :]), axis=1)
    got = st[torch.as_tensor(idx, device=st.device)].cpu().numpy()
    tol = 1e-13 if dtype == "complex128" else 1e-6
    assert np.max(np.abs(got - ref)) < tol
    # block probabilities over the top 6 qubits: prod of c^2 / s^2 of qubits 0..5
    bb = (np.arange(64)[:, None] >> (5 - np.arange(6))[None, :]) & 1
    pref = np.prod(np.where(bb == 1, s[None, :6] ** 2, c[None, :6] ** 2), axis=1)
    pgot = (st.abs() ** 2).double().reshape(64, -1).sum(dim=1).cpu().numpy()
    assert np.max(np.abs(pgot - pref)) < (1e-12 if dtype == "complex128" else 1e-5)
    del st
    torch.cuda.empty_cache()
