"""GPU parity of derivative states and QFIMs where the generator does not fit
one tile or is not a plain generator sum: the reference's TFIM templates at
13-16 qubits (shared R_x over all N qubits, templates.py:194-217), XXZ blocks
and genuinely non-commuting shared_parameter(commute=False) blocks
(gates.py:458-466), plus gates no random spec draws: ALLTOALL, I,
fixed_R_z, fixed_fSim.  complex128: derivative states to 1e-12 (max abs),
QFIM relative Frobenius < 1e-9; complex64: 1e-5."""
import numpy as np
import pytest

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS
from tests.helpers import infidelity, rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _layers(pyqc, n, layers, **kw):
    qc = pyqc.PQC(n, **kw)
    for l in layers:
        qc.add_layer(l)
    return qc


@pytest.mark.parametrize("p,n", [(3, 14), (2, 16)])
def test_tfim_templates_qfim_past_the_tile(pyqc, p, n):
    """TFIM_layers(p, n) with 12-bit derivative tiles: the shared R_x
    generator over n > 12 qubits is split into parts; the Jacobian (forward
    and plain), the QFIM and the EQD match the oracle."""
    from pyramaterised_amd import templates as T
    print(f"tfim p={p} n={n}", flush=True)
    qc = _layers(pyqc, n, T.TFIM_layers(p, n))
    og = OC.tfim(p, n)
    th = np.random.default_rng(n).random(qc.n_true_params) * 2 * np.pi
    qc.set_params(list(th))
    plan = qc.plan(with_generators=True, batch=qc.n_true_params)
    assert plan.tile_bits == 12
    assert max(plan.tangent_parts(t) for t in range(qc.n_true_params)) == 2
    psi, J = qc.jacobian()
    ref = OS.run(og, th, n)
    dref = OS.derivative_states(og, th, n)
    J = J.cpu().numpy()
    assert infidelity(psi.cpu().numpy(), ref) < 1e-12
    for t in range(qc.n_true_params):
        assert np.max(np.abs(J[t] - dref[t])) < 1e-12, t
    _, Jp = qc.jacobian(method="plain")
    assert np.max(np.abs(Jp.cpu().numpy() - J)) < 1e-12
    qc.state = qc.run(list(th))
    m = pyqc.measure.Measurements(qc)
    F = m.get_QFI()
    Fref = OM.qfim(dref, ref)
    assert rel_fro(F, Fref) < 1e-9
    assert m.get_effective_quantum_dimension(1e-12) == OM.effective_quantum_dimension(Fref, 1e-12)


def test_tfim_open_and_modified_past_the_tile(pyqc):
    """The config-3 ansatz (|+>^n start) at 14 qubits and modified_TFIM at 13
    qubits, complex128, plus the open chain in complex64 at 15 qubits
    (13-bit tiles: split there too)."""
    from pyramaterised_amd import templates as T
    n = 14
    qc = _layers(pyqc, n, T.tfim_open_layers(2, n))
    qc.set_initial_state(T.plus_state())
    og, init = OC.tfim_open(2, n)
    th = np.random.default_rng(5).random(qc.n_true_params) * 2 * np.pi
    qc.set_params(list(th))
    psi, J = qc.jacobian()
    dref = OS.derivative_states(og, th, n, init=init)
    assert np.max(np.abs(J.cpu().numpy() - np.array(dref))) < 1e-12

    n = 13
    qm = _layers(pyqc, n, T.modified_TFIM_layers(2, n))
    th = np.random.default_rng(6).random(qm.n_true_params) * 2 * np.pi
    qm.set_params(list(th))
    _, J = qm.jacobian()
    og = []
    for _ in range(2):
        og += [OC.RR_block("z", n), OC.shared_parameter([OC.R_x(i, n) for i in range(n)], n),
               OC.shared_parameter([OC.R_z(i, n) for i in range(n)], n)]
    dref = OS.derivative_states(og, th, n)
    assert np.max(np.abs(J.cpu().numpy() - np.array(dref))) < 1e-12

    n = 15
    q64 = _layers(pyqc, n, T.tfim_open_layers(2, n), dtype="complex64")
    q64.set_initial_state(T.plus_state())
    og, init = OC.tfim_open(2, n)
    th = np.random.default_rng(7).random(q64.n_true_params) * 2 * np.pi
    q64.set_params(list(th))
    plan = q64.plan(with_generators=True, batch=q64.n_true_params)
    assert plan.tile_bits == 13 and max(plan.tangent_parts(t) for t in range(4)) == 2
    psi, J = q64.jacobian()
    dref = np.array(OS.derivative_states(og, th, n, init=init))
    err = np.linalg.norm(J.cpu().numpy().astype(np.complex128) - dref) / np.linalg.norm(dref)
    assert err < 1e-5, err
    q64.state = q64.run(list(th))
    Fq = pyqc.measure.Measurements(q64).get_QFI()
    Fref = OM.qfim(list(dref), OS.run(og, th, n, init=init))
    assert rel_fro(Fq, Fref) < 1e-5


def test_xxz_gradients_and_qfim(pyqc):
    """XXZ_layers(2, 6) (commute=False by default, templates.py:230-257):
    every derivative state and the QFIM against the oracle's dense
    restatement of gates.py:458-466."""
    from pyramaterised_amd import templates as T
    n = 6
    # a fixed non-product start (XXZ blocks conserve the Z basis: from
    # |0..0> every derivative state is a phase of the state)
    pre = [pyqc.fixed_R_y(q, n, 0.4 + 0.3 * q) for q in range(n)] + [pyqc.CNOT([0, 3], n)]
    qc = _layers(pyqc, n, [pre] + T.XXZ_layers(2, n))
    assert qc.special_slots() == []  # symmetric, commuting Pauli strings: a generator sum
    og = [OC.fixed_R_y(q, n, 0.4 + 0.3 * q) for q in range(n)] + [OC.CNOT((0, 3), n)] + OC.xxz(2, n)
    th = np.random.default_rng(8).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    grads = qc.get_gradients()
    dref = OS.derivative_states(og, th, n)
    for t, g in enumerate(grads):
        assert np.max(np.abs(g.full().ravel() - dref[t])) < 1e-12, t
    F = pyqc.measure.Measurements(qc).get_QFI()
    Fref = OM.qfim(dref, OS.run(og, th, n))
    assert np.linalg.norm(Fref) > 0.1
    assert rel_fro(F, Fref) < 1e-9


def _noncommuting(pyqc, n):
    pk = [pyqc.R_y(q, n) for q in range(n)] + [pyqc.CNOT([q, q + 1], n) for q in range(n - 1)]
    pk += [pyqc.shared_parameter([pyqc.R_x(0, n), pyqc.R_y(0, n), pyqc.R_zz([0, 1], n)], n, commute=False)]
    pk += [pyqc.R_x(q, n) for q in range(n)] + [pyqc.CZ([0, n - 1], n)]
    og = [OC.R_y(q, n) for q in range(n)] + [OC.CNOT((q, q + 1), n) for q in range(n - 1)]
    og += [OC.shared_parameter([OC.R_x(0, n), OC.R_y(0, n), OC.R_zz((0, 1), n)], n, commute=False)]
    og += [OC.R_x(q, n) for q in range(n)] + [OC.CZ((0, n - 1), n)]
    return pk, og


def test_noncommuting_block_gradients(pyqc):
    """A commute=False block of non-commuting gates [R_x(0), R_y(0),
    R_zz(0, 1)] and one of R_y on two qubits (real, so conj(U) U = U^2):
    forward and plain Jacobians, take_derivative and the QFIM against the
    oracle; derivative() * gate applied to a ket gives the same state."""
    n = 5
    pk, og = _noncommuting(pyqc, n)
    pk += [pyqc.shared_parameter([pyqc.R_y(1, n), pyqc.R_y(3, n)], n, commute=False)]
    og += [OC.shared_parameter([OC.R_y(1, n), OC.R_y(3, n)], n, commute=False)]
    qc = pyqc.PQC(n)
    qc.add_layer(pk)
    assert len(qc.special_slots()) == 2
    th = np.random.default_rng(9).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    dref = OS.derivative_states(og, th, n)
    for method in ("forward", "plain"):
        _, J = qc.jacobian(method=method)
        assert np.max(np.abs(J.cpu().numpy() - np.array(dref))) < 1e-12, method
    for slot in qc.special_slots():
        g = next(x for x in qc.gates if qc._gate_slot.get(id(x)) == slot)
        d = qc.take_derivative(g)
        assert np.max(np.abs(d.full().ravel() - dref[slot])) < 1e-12
    F = pyqc.measure.Measurements(qc).get_QFI()
    assert rel_fro(F, OM.qfim(dref, OS.run(og, th, n))) < 1e-9
    # derivative() of the block, multiplied by the block, on the state before it
    blk = qc.gates[2 * n - 1]
    assert blk.special_derivative
    pre = pyqc.PQC(n)
    pre.add_layer(pk[:2 * n - 1])
    before = pre.run(list(th[:n]))
    blk.set_theta(th[n])
    out = blk.derivative() * (blk * before)
    og_pre = og[:2 * n - 1]
    psi_b = OS.run(og_pre, th[:n], n)
    want = OS.apply_deriv(OS.apply_prims(psi_b, og[2 * n - 1].prims(th[n]), n), og[2 * n - 1].deriv(th[n]), n)
    assert np.max(np.abs(out.full().ravel() - want)) < 1e-12


def test_alltoall_identity_fixed_gates(pyqc):
    """ALLTOALL (every ordered pair of permutations(range(N), 2),
    gates.py:382-401), I (consumes a parameter, D = -i/2 I, gates.py:150-156),
    fixed_R_z and fixed_fSim (gates.py:740-759): state and every derivative
    state against the oracle."""
    n = 4
    pk = [pyqc.R_y(q, n) for q in range(n)] + [pyqc.ALLTOALL(pyqc.CNOT, n)]
    pk += [pyqc.I(1, n), pyqc.fixed_R_z(2, n, 0.77), pyqc.R_x(0, n), pyqc.ALLTOALL(pyqc.CZ, n)]
    pk += [pyqc.fixed_fSim([0, 2], n), pyqc.fixed_fSim([3, 1], n), pyqc.R_z(3, n)]
    og = [OC.R_y(q, n) for q in range(n)] + [OC.ALLTOALL(OC.CNOT, n)]
    og += [OC.I(1, n), OC.fixed_R_z(2, n, 0.77), OC.R_x(0, n), OC.ALLTOALL(OC.CZ, n)]
    og += [OC.fixed_fSim((0, 2), n), OC.fixed_fSim((3, 1), n), OC.R_z(3, n)]
    qc = pyqc.PQC(n)
    qc.add_layer(pk)
    assert qc.n_true_params == OS.n_params(og) == 9
    th = np.random.default_rng(10).random(qc.n_true_params) * 2 * np.pi
    psi = qc.run(list(th)).full().ravel()
    ref = OS.run(og, th, n)
    assert infidelity(psi, ref) < 1e-12
    assert np.max(np.abs(psi - ref)) < 1e-13
    dref = OS.derivative_states(og, th, n)
    _, J = qc.jacobian()
    assert np.max(np.abs(J.cpu().numpy() - np.array(dref))) < 1e-12
    # the I gate's derivative state is -i/2 psi
    assert np.max(np.abs(J.cpu().numpy()[4] + 0.5j * ref)) < 1e-13
