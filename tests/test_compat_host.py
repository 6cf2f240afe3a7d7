"""Host-side compatibility helpers that need no GPU: the reference's fSim
matrix helpers and its magic conversion tables."""
import numpy as np

from tests.helpers import sre_bruteforce


def test_fsim_helper_matrices():
    """gates.py:588-737.  The "derivative" matrices are the reference's as
    written: the differentiated block, with the other entries of fsim_gate
    kept.  Pinned by central differences of fsim_gate and by unitarity."""
    import pyramaterised_amd as pyqc
    h = 1e-6
    for th, ph in ((0.3, 1.1), (2.5, -0.7), (np.pi / 2, np.pi)):
        U = pyqc.fsim_gate(th, ph)
        assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-14)
        dt = pyqc.fsim_gate_d_theta(th, ph)
        fd = (pyqc.fsim_gate(th + h, ph) - pyqc.fsim_gate(th - h, ph)) / (2 * h)
        assert np.max(np.abs(dt[1:3, 1:3] - fd[1:3, 1:3])) < 1e-8
        keep = np.ones((4, 4), bool)
        keep[1:3, 1:3] = False
        assert np.array_equal(dt[keep], U[keep])
        dp = pyqc.fsim_gate_d_phi(th, ph)
        fd = (pyqc.fsim_gate(th, ph + h) - pyqc.fsim_gate(th, ph - h)) / (2 * h)
        assert abs(dp[3, 3] - fd[3, 3]) < 1e-8
        keep = np.ones((4, 4), bool)
        keep[3, 3] = False
        assert np.array_equal(dp[keep], U[keep])
        assert np.array_equal(pyqc.fixed_fsim_gate(th), pyqc.fsim_gate(th, 0.0))
        assert np.array_equal(pyqc.fixed_fsim_gate_d_theta(th), pyqc.fsim_gate_d_theta(th, 0.0))


def test_magic_conversion_tables_and_reference_formula():
    """measure.py:268-349: the index / sign tables, checked entry by entry
    against XOR and popcount parity, and the reference's O(4^n) Renyi
    formula over them against the sum over explicit Pauli strings."""
    import pyramaterised_amd as pyqc
    n = 3
    m = pyqc.measure.Measurements(pyqc.PQC(n))
    assert list(m.numberToBase(5, 2, n)) == [1, 0, 1]
    assert list(m.numberToBase(0, 2, n)) == [0, 0, 0]
    assert list(m.numberToBase(7, 3, 3)) == [0, 2, 1]
    try:
        m.numberToBase(8, 2, n)
        raise AssertionError("expected IndexError")
    except IndexError:
        pass
    add, sign = m.get_conversion_matrices()
    d = 2 ** n
    assert add.shape == sign.shape == (d, d)
    for i in range(d):
        for k in range(d):
            assert add[i, k] == i ^ k
            assert sign[i, k] == (-1) ** bin(i & k).count("1")
    rng = np.random.default_rng(3)
    psi = rng.normal(size=d) + 1j * rng.normal(size=d)
    psi /= np.linalg.norm(psi)
    for alpha in (2, 0.5):
        s = np.sum(np.abs(2 ** (-n / 2) * np.dot(np.conjugate(psi) * sign, psi[add])) ** (2 * alpha))
        val = 1 / (1 - alpha) * np.log(s) - np.log(d)
        assert abs(val - sre_bruteforce(psi, n, alpha)) < 1e-12


def test_qobj_data_view_matches_qutip_accessors():
    """``Qobj.data`` answers the CSR accessors the reference's code uses
    (``state.data.toarray()``, measure.py:335) with the same numbers."""
    import pyramaterised_amd as pyqc
    op = pyqc.Qobj(np.array([[0, 1j], [2.0, 0]]))
    d = op.data
    assert np.array_equal(d.toarray(), op.full())
    assert d.nnz == 2 and d.shape == (2, 2)
    csr = d.tocsr()
    assert csr.nnz == 2 and np.array_equal(csr.toarray(), op.full())


def test_clifford_circuit_layers_structure():
    """templates.py:17-48.  In 'random' mode each of the N draws from
    [H, S, CZ, CNOT] appends that gate: H and S (PRot subclasses,
    gates.py:211,286) on a random.randint qubit, CZ / CNOT (EntGate) on a
    random.sample pair; replaying Python's ``random`` with the same seed gives
    the same gates on the same qubits.  The other mode puts LEN_CLIFF_STRING
    H / S gates on each of N // 2 distinct qubits, then a CNOT CHAIN."""
    import random
    from pyramaterised_amd import templates as T
    n, p = 6, 4
    random.seed(17)
    layers = T.clifford_circuit_layers(p, n)
    random.seed(17)
    want = []
    for _ in range(p):
        layer = []
        for _g in range(n):
            name = random.choice(["H", "S", "CZ", "CNOT"])
            if name in ("H", "S"):
                layer.append((name, random.randint(0, n - 1)))
            else:
                layer.append((name, tuple(random.sample(range(n), k=2))))
        want.append(layer)
    got = [[(type(g).__name__, g.q_on if type(g).__name__ in ("H", "S") else (g.q1, g.q2)) for g in layer]
           for layer in layers]
    assert got == want
    random.seed(18)
    other = T.clifford_circuit_layers(2, n, method="strings")
    for layer in other:
        assert type(layer[-1]).__name__ == "CHAIN"
        singles = layer[:-1]
        assert len(singles) == (n // 2) * T.LEN_CLIFF_STRING
        assert {type(g).__name__ for g in singles} <= {"H", "S"}
        assert len({g.q_on for g in singles}) == n // 2


def test_to_device_rejects_unsupported_dtypes_before_any_device_call():
    """engine.to_device maps float32/64, complex64/128 and int32/64 arrays to
    torch dtypes; any other dtype raises TypeError on the host, before the
    engine asks for a device (so this runs without a GPU)."""
    import pytest
    from pyramaterised_amd import engine as E
    for dt in (np.uint8, np.uint16, np.int16, np.float16, np.bool_):
        with pytest.raises(TypeError):
            E.to_device(np.zeros(3, dtype=dt))
    assert set(E._NP_TO_TORCH) == {np.dtype(x) for x in (np.float64, np.float32, np.complex128, np.complex64,
                                                          np.int64, np.int32)}
