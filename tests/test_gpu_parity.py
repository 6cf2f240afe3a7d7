"""GPU parity of the HIP engine against the CPU oracle (complex128: state
infidelity < 1e-12, QFIM relative Frobenius error < 1e-9; complex64:
infidelity < 1e-5)."""
import numpy as np
import pytest

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS
from tests.helpers import build_oracle, build_pkg, infidelity, random_spec, rel_fro, sre_bruteforce

pytestmark = pytest.mark.gpu

QG_ENERGY = 0.46135870050914374  # tests.py:207-209


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _pqc(pyqc, n, gates, **kw):
    qc = pyqc.PQC(n, **kw)
    qc.add_layer(gates)
    return qc


def test_config1_state_mw_qfim(pyqc):
    """4 qubits, 3 x (R_y, R_z all; CNOT 0->1->2->3): state, MW and 24x24 QFIM."""
    from pyramaterised_amd import templates as T
    n, p = 4, 3
    qc = pyqc.PQC(n)
    for l in T.hea_linear_layers(p, n):
        qc.add_layer(l)
    assert qc.n_true_params == 24
    th = np.random.default_rng(1).random(24) * 2 * np.pi
    qc.state = qc.run(list(th))
    og = OC.hea_linear(p, n)
    ref = OS.run(og, th, n)
    assert infidelity(qc.state.full(), ref) < 1e-12
    assert np.max(np.abs(qc.state.full().ravel() - ref)) < 1e-13
    m = pyqc.measure.Measurements(qc)
    assert abs(m.single_Q(qc.state, n) - OM.single_Q(ref, n)) < 1e-13
    F = m.get_QFI()
    Fref = OM.qfim(OS.derivative_states(og, th, n), ref)
    assert rel_fro(F, Fref) < 1e-9


def test_config1_chain_variant(pyqc):
    from pyramaterised_amd import templates as T
    n, p = 4, 3
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    th = np.random.default_rng(2).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    og = OC.clifford_he(p, n)
    ref = OS.run(og, th, n)
    assert infidelity(qc.state.full(), ref) < 1e-12
    F = pyqc.measure.Measurements(qc).get_QFI()
    assert rel_fro(F, OM.qfim(OS.derivative_states(og, th, n), ref)) < 1e-9


def test_qg_energy_and_eqd(pyqc):
    """tests.py:158-213 with the exact default_rng(1) draws."""
    n = 4
    qc = pyqc.PQC(n)
    qc.add_layer([pyqc.fixed_R_y(i, n, np.pi / 4) for i in range(n)])
    qc.add_layer([pyqc.R_z(0, n), pyqc.R_x(1, n), pyqc.R_y(2, n), pyqc.R_z(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    qc.add_layer([pyqc.R_x(0, n), pyqc.R_x(1, n), pyqc.R_x(2, n), pyqc.R_y(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    qc.add_layer([pyqc.R_z(0, n), pyqc.R_x(1, n), pyqc.R_y(2, n), pyqc.R_y(3, n), pyqc.CHAIN(pyqc.CNOT, n)])
    r = np.random.default_rng(1)
    th = [r.random(1)[0] * 2 * np.pi for _ in range(12)]
    energy = qc.cost(th)
    assert abs(energy - QG_ENERGY) < 1e-14
    eqd = pyqc.measure.Measurements(qc).get_effective_quantum_dimension(1e-12)
    assert eqd == 12


@pytest.mark.parametrize("n", [4, 6, 8])
def test_npqc_identity(pyqc, n):
    from pyramaterised_amd import templates as T
    for P in range(1, 2 ** (n // 2) + 1):
        layers, th = T.NPQC_layers(P, n)
        qc = pyqc.PQC(n)
        for l in layers:
            qc.add_layer(l)
        qc.state = pyqc.Qobj(qc.run(angles=th))
        F = pyqc.measure.Measurements(qc).get_QFI()
        assert np.max(np.abs(F - np.eye(len(th)))) < 1e-12


def test_basic_gates(pyqc):
    """tests.py:64-87."""
    t = pyqc.PQC(1)
    t.add_layer([pyqc.fixed_R_y(0, 1, np.pi / 2)])
    out = t.run("random")
    assert abs(np.real(out[1][0][0]) - 1 / np.sqrt(2)) < 1e-15
    t2 = pyqc.PQC(1)
    t2.add_layer([pyqc.H(0, 1)], n=2)
    assert t2.run("random") == pyqc.basis(2, 0)


def test_bell_mw_magic(pyqc):
    """tests.py:284-295: Q = 1, SRE = 0, GKP = 0 for bell_state('11')."""
    m = pyqc.measure.Measurements(pyqc.PQC(2))
    bell = pyqc.bell_state("11")
    assert abs(m.single_Q(bell, 2) - 1) < 1e-12
    assert abs(m.renyi_entropy_fast(bell)) < 1e-10
    assert abs(m.gkp_fast(bell)) < 1e-10


@pytest.mark.parametrize("n,depth,seed", [(3, 3, 0), (5, 4, 1), (9, 3, 2)])
def test_all_gate_kinds_state_and_derivatives(pyqc, n, depth, seed):
    rng = np.random.default_rng(seed)
    spec = random_spec(n, depth, rng)
    qc = _pqc(pyqc, n, build_pkg(spec, n, pyqc))
    og = build_oracle(spec, n)
    th = rng.random(qc.n_true_params) * 2 * np.pi
    assert OS.n_params(og) == len(th)
    qc.state = qc.run(list(th))
    ref = OS.run(og, th, n)
    assert infidelity(qc.state.full(), ref) < 1e-12
    grads = qc.get_gradients()
    dref = OS.derivative_states(og, th, n)
    assert len(grads) == len(dref)
    for g, d in zip(grads, dref):
        assert np.max(np.abs(g.full().ravel() - d)) < 1e-12
    plain = qc.get_gradients(method="plain")
    for g, d in zip(plain, dref):
        assert np.max(np.abs(g.full().ravel() - d)) < 1e-12


@pytest.mark.parametrize("n,tile", [(8, 5), (11, 6), (16, 13)])
def test_multi_sweep_matches_oracle(pyqc, n, tile):
    """Tiles smaller than the state: many sweeps, ops on qubits outside the
    tile; the Jacobian spawns tangents mid-plan."""
    rng = np.random.default_rng(10 + n)
    spec = random_spec(n, 3, rng, fsim=True)
    spec = [s for s in spec if s[0] != "shared_Rx"]
    qc = _pqc(pyqc, n, build_pkg(spec, n, pyqc), tile_bits=tile, low_bits=min(3, tile))
    assert qc.plan().n_sweeps > 1
    og = build_oracle(spec, n)
    th = rng.random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    ref = OS.run(og, th, n)
    assert infidelity(qc.state.full(), ref) < 1e-12
    if n <= 11:
        psi, J = qc.jacobian()
        dref = OS.derivative_states(og, th, n)
        Jh = J.cpu().numpy()
        for t in range(len(dref)):
            assert np.max(np.abs(Jh[t] - dref[t])) < 1e-12, t


def test_config3_tfim_qfim_adjoint_vs_reference_and_fd(pyqc):
    """12-qubit open-chain HVA, 10 layers, |+>^12: QFIM from the forward
    Jacobian against the oracle (reference algorithm) and central FD."""
    from pyramaterised_amd import templates as T
    n, p = 12, 10
    qc = pyqc.PQC(n)
    qc.set_initial_state(T.plus_state())
    for l in T.tfim_open_layers(p, n):
        qc.add_layer(l)
    assert qc.n_true_params == 20
    th = np.random.default_rng(3).random(20) * 2 * np.pi
    qc.state = qc.run(list(th))
    og, init = OC.tfim_open(p, n)
    ref = OS.run(og, th, n, init)
    assert infidelity(qc.state.full(), ref) < 1e-12
    F = pyqc.measure.Measurements(qc).get_QFI()
    Fref = OM.qfim(OS.derivative_states(og, th, n, init), ref)
    assert rel_fro(F, Fref) < 1e-9
    # central finite differences of the GPU state (h = 1e-6)
    h = 1e-6
    _, J = qc.jacobian()
    Jh = J.cpu().numpy()
    for t in (0, 7, 19):
        tp, tm = th.copy(), th.copy()
        tp[t] += h
        tm[t] -= h
        fd = (qc.run(list(tp)).full().ravel() - qc.run(list(tm)).full().ravel()) / (2 * h)
        assert np.max(np.abs(fd - Jh[t])) < 1e-8


def test_config3_complex64(pyqc):
    from pyramaterised_amd import templates as T
    n, p = 12, 10
    qc = pyqc.PQC(n, dtype="complex64")
    qc.set_initial_state(T.plus_state())
    for l in T.tfim_open_layers(p, n):
        qc.add_layer(l)
    th = np.random.default_rng(3).random(20) * 2 * np.pi
    st = qc.run(list(th))
    og, init = OC.tfim_open(p, n)
    ref = OS.run(og, th, n, init)
    assert infidelity(st.full(), ref) < 1e-5


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_config4_brickwork_22q(pyqc, dtype, tol):
    """R_y + brickwork CZ, 30 layers at 22 qubits against the einsum oracle."""
    from pyramaterised_amd import templates as T
    n, p = 22, 30
    qc = pyqc.PQC(n, dtype=dtype)
    for l in T.ry_cz_brickwork_layers(p, n):
        qc.add_layer(l)
    th = np.random.default_rng(4).random(qc.n_true_params) * 2 * np.pi
    st = qc.run(list(th)).full().ravel()
    ref = OS.run(OC.ry_cz_brickwork(p, n), th, n)
    assert infidelity(st, ref) < tol
    assert abs(np.vdot(st, st).real - 1) < (1e-12 if dtype == "complex128" else 1e-5)


def test_batched_expressibility_matches_sequential(pyqc):
    """Batch of random states == the same draws run one at a time; pair
    fidelities, histogram and KL equal the oracle's."""
    from pyramaterised_amd import gates as G
    from pyramaterised_amd import templates as T
    n, p, S = 6, 3, 40
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    G.rng = np.random.default_rng(5)
    params = qc.random_params(S)
    states = qc.run_batch(params).cpu().numpy()
    og = OC.clifford_he(p, n)
    refs = [OS.run(og, params[s], n) for s in range(S)]
    for s in range(S):
        assert np.max(np.abs(states[s] - refs[s])) < 1e-13
    m = pyqc.measure.Measurements(qc)
    import torch
    F = m.fidelities_all_pairs(torch.as_tensor(states, device="cuda")).cpu().numpy()
    Fref = np.array(OM.pair_fidelities(refs))
    assert np.max(np.abs(F - Fref)) < 1e-13
    e = m.expr(list(F), 2 ** n)
    assert abs(e - OM.expr(list(Fref), 2 ** n)) < 1e-12
    prob, mids = m._gen_histo(list(Fref))
    prob_ref, mids_ref = OM.gen_histo(list(Fref))
    assert np.array_equal(prob, prob_ref)
    # the random stream: S run("random") calls draw the same numbers
    G.rng = np.random.default_rng(5)
    seq = [qc.run("random").full().ravel() for _ in range(3)]
    for s in range(3):
        assert np.max(np.abs(seq[s] - states[s])) < 1e-13


def test_histogram_matches_numpy_edges(pyqc):
    import torch
    rng = np.random.default_rng(6)
    v = np.concatenate([rng.random(5000), np.linspace(0, 1, 76), [0.0, 1.0, -0.1, 1.1]])
    for bins in (1, 37, 75):
        edges = np.linspace(0, 1, bins + 1)
        from pyramaterised_amd import engine as E
        c = E.histogram(torch.as_tensor(v, device="cuda"), edges).cpu().numpy()
        ref, _ = np.histogram(v, bins=bins, range=(0, 1))
        assert np.array_equal(c, ref)


def test_expect_pauli_and_tfim_hamiltonian(pyqc):
    from pyramaterised_amd import templates as T
    n = 5
    rng = np.random.default_rng(7)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    H = T.TFIM_hamiltonian(n, g=0.7, h=0.3)
    ket = pyqc.qobj.to_ket(v)
    val = pyqc.expect(H, ket)
    ref = np.vdot(v, H.full() @ v).real
    assert abs(val - ref) < 1e-13
    Hy = pyqc.PauliSum(n, {(2, 1, 0, 3, 2): 0.5, (0, 2, 2, 0, 0): -1.25})
    assert abs(pyqc.expect(Hy, ket) - np.vdot(v, Hy.full() @ v).real) < 1e-13


def test_magic_matches_pauli_bruteforce(pyqc):
    rng = np.random.default_rng(8)
    n = 4
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    m = pyqc.measure.Measurements(pyqc.PQC(n))
    assert abs(m.renyi_entropy_fast(v) - sre_bruteforce(v, n, 2)) < 1e-12
    assert abs(m.gkp_fast(v) - sre_bruteforce(v, n, 0.5) / (2 * np.log(2))) < 1e-12
    # Haar average: log(3 + 2^n) - log 4 (tests.py:276-277), statistically
    states = rng.normal(size=(400, 2 ** n)) + 1j * rng.normal(size=(400, 2 ** n))
    states /= np.linalg.norm(states, axis=1, keepdims=True)
    from pyramaterised_amd.magic import renyi_entropy_batch
    import torch
    vals = renyi_entropy_batch(torch.as_tensor(states, device="cuda"), n, 2)
    assert abs(np.mean(vals) - (np.log(3 + 2 ** n) - np.log(4))) < 0.05


def test_take_derivative_and_qfim_from_grad_list(pyqc):
    from pyramaterised_amd import templates as T
    n = 5
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(2, n):
        qc.add_layer(l)
    th = np.random.default_rng(9).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    og = OC.clifford_he(2, n)
    dref = OS.derivative_states(og, th, n)
    pg = [g for g in qc.gates if g.param_count > 0]
    for k in (0, 3, len(pg) - 1):
        d = qc.take_derivative(pg[k]).full().ravel()
        assert np.max(np.abs(d - dref[k])) < 1e-13
    m = pyqc.measure.Measurements(qc)
    F1 = m.get_QFI()
    F2 = m.get_QFI(grad_list=qc.get_gradients(method="plain"))
    assert np.max(np.abs(F1 - F2)) < 1e-13


@pytest.mark.parametrize("n,p", [(8, 48), (8, 19), (10, 32)])
def test_qfim_real_gram_paths(pyqc, n, p):
    """pqs_qfim's real route: Re(J J^dag) over upper 128-row tiles on the
    matrix cores, split over K (P = 768: six full tiles; P = 304: a ragged
    last tile; P = 640); against the oracle QFIM and against the complex
    split-K Gram kernel (pqs_gram of J with itself, combined on the host)."""
    import torch
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    P = qc.n_true_params
    assert P == 2 * n * p
    th = np.random.default_rng(21).random(P) * 2 * np.pi
    qc.state = qc.run(list(th))
    psi, J = qc.jacobian()
    F = E.qfim(J, psi).cpu().numpy()
    og = OC.clifford_he(p, n)
    ref = OS.run(og, th, n)
    Fref = OM.qfim(OS.derivative_states(og, th, n), ref)
    assert rel_fro(F, Fref) < 1e-9
    G = E.gram(J, J).cpu().numpy()
    b = E.gram(psi.reshape(1, -1), J).cpu().numpy().ravel()
    F2 = 4 * np.real(G - np.outer(np.conj(b), b))
    assert np.max(np.abs(F - F2)) < 1e-10
    assert np.allclose(F, F.T, atol=0)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_qfim_few_amplitudes(pyqc, n):
    """QFIMs of 1-3 qubit circuits (2-8 amplitudes: the product kernel's k
    steps are 16 doubles, so 1-2 qubits take the zero-padded copy) against
    the oracle; one R_y on |0> has F = 1."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(3, n):
        qc.add_layer(l)
    th = np.random.default_rng(40 + n).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    psi, J = qc.jacobian()
    F = E.qfim(J, psi).cpu().numpy()
    og = OC.clifford_he(3, n)
    Fref = OM.qfim(OS.derivative_states(og, th, n), OS.run(og, th, n))
    assert np.max(np.abs(F - Fref)) < 1e-12
    q1 = pyqc.PQC(1)
    q1.add_layer([pyqc.gates.R_y(0, 1)])
    q1.state = q1.run([0.7])
    psi1, J1 = q1.jacobian()
    assert abs(E.qfim(J1, psi1).cpu().numpy()[0, 0] - 1.0) < 1e-14


@pytest.mark.parametrize("N,P,dtype", [(20, 160, "complex64"), (19, 600, "complex128")])
def test_qfim_column_blocks(pyqc, N, P, dtype):
    """Jacobians above 1 GiB go through the product kernel in column blocks
    copied to a compact buffer (three and ten blocks here, the last one
    narrower): F = 4 Re(G - conj(b) b^T) against a float64 torch matmul of
    the same random J and psi (rel-Frobenius 1e-11)."""
    import torch
    from pyramaterised_amd import engine as E
    dev = E.device()
    g = torch.Generator(device=dev).manual_seed(N)
    cd = getattr(torch, dtype)
    J = (torch.randn((P, 2 ** N), dtype=torch.complex128, device=dev, generator=g) / 2 ** (N / 2)).to(cd)
    psi = (torch.randn(2 ** N, dtype=torch.complex128, device=dev, generator=g) / 2 ** (N / 2)).to(cd)
    assert J.numel() * J.element_size() > 2 ** 30
    F = E.qfim(J, psi)
    Jd, pd = J.to(torch.complex128), psi.to(torch.complex128)
    G = Jd @ Jd.conj().T
    b = Jd.conj() @ pd
    Fr = 4 * (G - torch.outer(b.conj(), b)).real
    assert (torch.linalg.norm(F - Fr) / torch.linalg.norm(Fr)).item() < 1e-11
    assert torch.equal(F, F.T)


def test_config2_full_size_properties(pyqc):
    """The headline workload at its full size: 10,000 states of the 16-qubit,
    20-layer HEA (940 gates, P = 640) in one batched run.  Every state has
    norm 1; sampled states equal the oracle's; the 5000 independent pair
    fidelities average 1/2^16 (their Haar mean; the standard error of the
    mean is about 2e-7); the 75-bin histogram counts every pair once."""
    import torch
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p, S = 16, 20, 10000
    print(f"case n={n} layers={p} states={S}", flush=True)
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    params = np.random.default_rng(77).random((S, qc.n_true_params)) * 2 * np.pi
    st = qc.run_batch(params)
    assert qc.plan(batch=S).jitted
    norms = torch.linalg.vector_norm(st, dim=1).cpu().numpy()
    assert np.max(np.abs(norms - 1)) < 1e-12
    og = OC.clifford_he(p, n)
    for s in (0, 4321, S - 1):
        assert infidelity(st[s].cpu().numpy(), OS.run(og, params[s], n)) < 1e-12
    F = E.abs2(E.inner_pairs(st, np.arange(0, S, 2), np.arange(1, S, 2)))
    mean = float(F.mean().item())
    assert abs(mean - 2.0 ** -16) < 0.1 * 2.0 ** -16, mean
    counts = E.histogram(F, np.linspace(0, 1, 76)).cpu().numpy()
    assert counts.sum() == S // 2
    del st
    torch.cuda.empty_cache()
