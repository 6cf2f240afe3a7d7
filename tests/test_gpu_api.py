"""GPU tests of the drop-in API surface beyond the hot-path configs: every
generate_circuit type (templates.py:361-429), ``gate * ket`` and
``derivative()`` (gates.py:49-100), efficient_measurements in both angle
modes (measure.py:370-459), one gradient and one QNG step of ``train``
(measure.py:473-553), find_overparam_point and find_eff_H (measure.py:101-121,
199-224), and serving requests of 18-20 qubits and depth >= 30.
Oracle comparisons: complex128 states to 1e-12 (infidelity / max abs),
QFIMs to 1e-9 relative Frobenius."""
import random

import numpy as np
import pytest

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS
from tests.helpers import infidelity, rel_fro, sre_bruteforce
from tests.test_planner import GENERATE_CASES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


@pytest.mark.parametrize("case", GENERATE_CASES, ids=[c[0] + str(c[1].get("rotator", "")) for c in GENERATE_CASES])
def test_generate_circuit_states_and_gradients(pyqc, case):
    """The state and every derivative state of each generate_circuit type on
    the GPU against the oracle (the fSim family's "derivatives" as the
    reference writes them, gates.py:609-648, 719-737)."""
    from pyramaterised_amd import templates as T
    name, kw, n, p, build, start = case
    qc = T.generate_circuit(name, n, p, shuffle=False, **kw)
    og = build(p, n)
    init = OC.half_filled(n) if start == "half" else None
    th = np.random.default_rng(7 + len(name)).random(qc.n_true_params) * 2 * np.pi
    psi = qc.run(list(th)).full().ravel()
    ref = OS.run(og, th, n, init=init)
    assert infidelity(psi, ref) < 1e-12
    assert np.max(np.abs(psi - ref)) < 1e-13
    grads = qc.get_gradients()
    dref = OS.derivative_states(og, th, n, init=init)
    assert len(grads) == len(dref)
    for t, g in enumerate(grads):
        assert np.max(np.abs(g.full().ravel() - dref[t])) < 1e-12, (name, t)


def test_generate_circuit_shuffled_start_uses_random(pyqc):
    """shuffle=True permutes the half-filled start with random.shuffle, as
    the reference does (templates.py:377-380)."""
    from pyramaterised_amd import templates as T
    n = 6
    random.seed(12)
    qc = T.generate_circuit("fermionic", n, 1, shuffle=True)
    random.seed(12)
    bits = [1] * (n // 2) + [0] * (n - n // 2)
    random.shuffle(bits)
    want = np.array([1.0 + 0j])
    for b in bits:
        want = np.kron(want, np.array([0, 1]) if b else np.array([1, 0]))
    assert np.allclose(np.asarray(qc.initial_state.full()).ravel(), want)


def test_gate_times_ket_and_derivative(pyqc):
    """``gate * ket`` applies the gate's operation to a GPU ket and
    ``derivative() * (gate * ket)`` gives the reference's derivative state
    of that gate alone (gates.py:63-100, 133-138, 451-466, 519-522, 609-648),
    for rotations, RR gates, fixed and entangling gates, shared blocks (both
    commute modes), fSim and fixed_fSim."""
    n = 4
    rng = np.random.default_rng(3)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    ket = pyqc.qobj.to_ket(v, n)
    cases = [
        (pyqc.R_x(1, n), OC.R_x(1, n)),
        (pyqc.R_y(3, n), OC.R_y(3, n)),
        (pyqc.negative_R_z(0, n), OC.negative_R_z(0, n)),
        (pyqc.offset_R_z(2, n, 0.4), OC.offset_R_z(2, n, 0.4)),
        (pyqc.R_yy([0, 2], n), OC.R_yy((0, 2), n)),
        (pyqc.R_zz([3, 1], n), OC.R_zz((3, 1), n)),
        (pyqc.H(2, n), OC.H(2, n)),
        (pyqc.CNOT([2, 0], n), OC.CNOT((2, 0), n)),
        (pyqc.sqrtiSWAP([1, 3], n), OC.sqrtiSWAP((1, 3), n)),
        (pyqc.shared_parameter([pyqc.R_x(q, n) for q in range(n)], n),
         OC.shared_parameter([OC.R_x(q, n) for q in range(n)], n)),
        (pyqc.shared_parameter([pyqc.R_x(0, n), pyqc.R_y(0, n)], n, commute=False),
         OC.shared_parameter([OC.R_x(0, n), OC.R_y(0, n)], n, commute=False)),
        (pyqc.fSim([1, 2], n), OC.fSim((1, 2), n)),
        (pyqc.fixed_fSim([3, 0], n), OC.fixed_fSim((3, 0), n)),
    ]
    from pyramaterised_amd import gates as G
    for g, og in cases:
        for th, ph in ((1.234, 0.567), (-2.5, 3.0)):
            if g.is_param:
                g.set_theta(th)
                if g.param_count == 2:
                    g.set_phi(ph)
            n_plans = len(G._APPLY_PLANS)
            out = (g * ket).full().ravel()
            want = OS.apply_prims(v, og.prims(th, ph), n)
            assert np.max(np.abs(out - want)) < 1e-13, og.name
            if th < 0 and "fSim" not in og.name:
                # the same structure at other angles: the cached plan (an
                # fSim's angles live in its dense 4x4, so it is new)
                assert len(G._APPLY_PLANS) == n_plans, og.name
            if g.is_param:
                d = (g.derivative() * (g * ket)).full().ravel()
                dwant = OS.apply_deriv(want, og.deriv(th, ph, 1), n)
                assert np.max(np.abs(d - dwant)) < 1e-13, og.name
    # operation.full(): the dense operator, column by column through the
    # cached plan
    g = pyqc.R_y(1, 3)
    g.set_theta(0.7)
    U = np.asarray(g.operation.full())
    cols = [OS.apply_prims(np.eye(8, dtype=np.complex128)[:, j], OC.R_y(1, 3).prims(0.7, 0.0), 3) for j in range(8)]
    assert np.max(np.abs(U - np.stack(cols, axis=1))) < 1e-14
    # operator algebra: (g1 * g2) * ket applies g2 first, as QuTiP's product
    a, b = pyqc.R_x(0, n), pyqc.CNOT([0, 1], n)
    a.set_theta(0.3)
    out = ((a * b) * ket).full().ravel()
    want = OS.apply_prims(OS.apply_prims(v, OC.CNOT((0, 1), n).prims(), n), OC.R_x(0, n).prims(0.3), n)
    assert np.max(np.abs(out - want)) < 1e-13


@pytest.mark.parametrize("angles", ["random", "clifford"])
def test_efficient_measurements(pyqc, angles):
    """measure.py:370-459: one batch of states shared by the expressibility,
    entanglement, magic and GKP measures, both as summaries and with
    full_data; the oracle recomputes everything from the same parameters."""
    from pyramaterised_amd import templates as T
    n, p, S = 4, 2, 24
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    og = OC.clifford_he(p, n)
    P = qc.n_true_params
    m = pyqc.measure.Measurements(qc)
    results = {}
    for full in (False, True):
        if angles == "random":
            pyqc.gates.rng = np.random.default_rng(21)
            params = np.random.default_rng(21).random(S * P).reshape(S, P) * 2 * np.pi
        else:
            random.seed(21)
            cl = (0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi)
            init = [[random.choice(cl) for _ in range(qc.n_params)] for _ in range(S)]
            params = np.array([a[:P] for a in init])
            random.seed(21)
        results[full] = m.efficient_measurements(S, full_data=full, angles=angles)
    pyqc.gates.rng = np.random.default_rng(1)
    states = [OS.run(og, params[s], n) for s in range(S)]
    F = OM.pair_fidelities(states)
    qs = [OM.single_Q(s, n) for s in states]
    mag = [sre_bruteforce(s, n, 2) for s in states]
    gkp = [sre_bruteforce(s, n, 0.5) / (2 * np.log(2)) for s in states]
    summ, full = results[False], results[True]
    assert abs(summ["Expr"] - OM.expr(F, 2 ** n)) < 1e-9
    assert abs(summ["Ent"][0] - np.mean(qs)) < 1e-12 and abs(summ["Ent"][1] - np.std(qs)) < 1e-12
    assert abs(summ["Magic"][0] - np.mean(mag)) < 1e-10 and abs(summ["Magic"][1] - np.std(mag)) < 1e-10
    assert abs(summ["GKP"][0] - np.mean(gkp)) < 1e-10
    assert np.max(np.abs(np.array(full["Expr"]) - F)) < 1e-12
    assert np.max(np.abs(np.array(full["Ent"]) - qs)) < 1e-12
    assert np.max(np.abs(np.array(full["Magic"]) - mag)) < 1e-10
    assert np.max(np.abs(np.array(full["GKP"]) - gkp)) < 1e-10


def _zz_energy(psi, n):
    idx = np.arange(2 ** n)
    sign = np.where(((idx >> (n - 1)) ^ (idx >> (n - 2))) & 1, -1.0, 1.0)
    return float(np.sum(sign * np.abs(psi) ** 2))


def _oracle_grad(og, th, n):
    psi = OS.run(og, th, n)
    zz = np.where(((np.arange(2 ** n) >> (n - 1)) ^ (np.arange(2 ** n) >> (n - 2))) & 1, -1.0, 1.0)
    d = OS.derivative_states(og, th, n)
    return psi, np.array([2 * np.real(np.vdot(zz * psi, di)) for di in d]), d


@pytest.mark.parametrize("method", ["gradient", "QNG"])
def test_train_one_step(pyqc, method):
    """measure.py:473-553 with the default H = Z0 Z1: one gradient (or QNG)
    step, energies and the magic / entanglement / GKP trajectory against the
    oracle's step theta1 = theta0 - rate * g (or rate * pinv(F) g)."""
    from pyramaterised_amd import templates as T
    n, p = 4, 2
    qc = pyqc.PQC(n)
    for l in T.generic_HE_layers(p, n):
        qc.add_layer(l)
    og = OC.generic_he(p, n)
    th0 = np.random.default_rng(31).random(qc.n_true_params) * 2 * np.pi
    rate = 0.02
    m = pyqc.measure.Measurements(qc)
    energy, traj, magics, ents, gkps = m.train(epsilon=0.5, rate=rate, method=method, angles=list(th0))
    psi0, g, d = _oracle_grad(og, th0, n)
    if method == "gradient":
        th1 = th0 - rate * g
    else:
        F = OM.qfim(d, psi0)
        th1 = th0 - rate * np.linalg.pinv(F).dot(g)
    psi1 = OS.run(og, th1, n)
    assert len(traj) == 2
    assert abs(traj[0] - _zz_energy(psi0, n)) < 1e-12
    assert abs(traj[1] - _zz_energy(psi1, n)) < 1e-11
    assert abs(energy - traj[1]) < 1e-15
    assert np.max(np.abs(np.array(qc.get_params()) - th1)) < 1e-11
    assert abs(ents[1] - OM.single_Q(psi1, n)) < 1e-11
    assert abs(magics[1] - sre_bruteforce(psi1, n, 2)) < 1e-9
    assert abs(gkps[1] - sre_bruteforce(psi1, n, 0.5) / (2 * np.log(2))) < 1e-9


def test_serving_wide_and_deep_requests(pyqc):
    """Config-5 requests at the wide and deep end (18-20 qubits, 30-40
    layers, and a 30-layer QFIM) through the server: the oracle checks the
    18-qubit states and the QFIM; every result equals the same job run alone."""
    from pyramaterised_amd import serve as sv
    rng = np.random.default_rng(41)
    spec = [(18, 32, sv.STATE_MW), (20, 30, sv.STATE_MW), (19, 40, sv.STATE_MW), (18, 32, sv.STATE_MW),
            (10, 30, sv.QFIM), (12, 31, sv.QFIM)]
    reqs = [sv.Request(i, n, d, kind, rng.random(2 * n * d) * 2 * np.pi) for i, (n, d, kind) in enumerate(spec)]
    srv = sv.Server(n_streams=2, max_batch=8)
    mix = srv.run(reqs, realtime=False)
    srv.close()
    alone, _ = sv.run_one_at_a_time(reqs)
    for a, b in zip(mix, alone):
        if a.kind == sv.STATE_MW:
            assert np.max(np.abs(a.state.cpu().numpy() - b.state.cpu().numpy())) < 1e-12
            assert abs(a.Q - b.Q) < 1e-12
            assert abs(float(np.linalg.norm(a.state.cpu().numpy())) - 1) < 1e-12
        else:
            assert rel_fro(a.qfim, b.qfim) < 1e-12
    for r, res in zip(reqs, mix):
        if r.kind == sv.STATE_MW and r.n == 18:
            ref = OS.run(OC.clifford_he(r.depth, r.n), r.params, r.n)
            assert infidelity(res.state.cpu().numpy(), ref) < 1e-12
            assert abs(res.Q - OM.single_Q(ref, r.n)) < 1e-12
        if r.kind == sv.QFIM and r.n == 10:
            og = OC.clifford_he(r.depth, r.n)
            ref = OS.run(og, r.params, r.n)
            F = OM.qfim(OS.derivative_states(og, r.params, r.n), ref)
            assert rel_fro(res.qfim, F) < 1e-9


@pytest.mark.parametrize("n", [4, 7, 10])
def test_arbgate_dense_path(pyqc, n):
    """ARBGATE (gates.py:407-435): circuits with dense gates run in segments
    with U = expm(-i theta H) applied on the GPU; state, every derivative
    state (the reference's D = -i H / 2 after U) and the QFIM against the
    oracle's scipy-expm restatement; a Hermitian H given as a Qobj and a
    non-Hermitian one as an array (expm path)."""
    rng = np.random.default_rng(n)
    d = 2 ** n
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H1 = (A + A.conj().T) / (2 * np.sqrt(d))
    H2 = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / (4 * np.sqrt(d))
    pk = [pyqc.R_y(q, n) for q in range(n)] + [pyqc.CNOT([q, q + 1], n) for q in range(n - 1)]
    pk += [pyqc.ARBGATE(pyqc.Qobj(H1)), pyqc.R_x(0, n), pyqc.fSim([1, n - 1], n)]
    pk += [pyqc.ARBGATE(H2), pyqc.R_z(n - 1, n)]
    og = [OC.R_y(q, n) for q in range(n)] + [OC.CNOT((q, q + 1), n) for q in range(n - 1)]
    og += [OC.ARBGATE(H1, n), OC.R_x(0, n), OC.fSim((1, n - 1), n), OC.ARBGATE(H2, n), OC.R_z(n - 1, n)]
    qc = pyqc.PQC(n)
    qc.add_layer(pk)
    assert qc.has_dense() and qc.n_true_params == OS.n_params(og)
    th = rng.random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    ref = OS.run(og, th, n)
    psi = qc.state.full().ravel()
    assert np.max(np.abs(psi - ref)) < 1e-12
    dref = OS.derivative_states(og, th, n)
    _, J = qc.jacobian()
    assert np.max(np.abs(J.cpu().numpy() - np.array(dref))) < 1e-12
    slot = n  # the first ARBGATE's parameter
    assert np.max(np.abs(qc.take_derivative(qc.gates[2 * n - 1]).full().ravel() - dref[slot])) < 1e-12
    # not unitary (H2), so compare the QFIM formula on these states
    F = pyqc.measure.Measurements(qc).get_QFI()
    assert rel_fro(F, OM.qfim(dref, ref)) < 1e-9
    # batched evolution: each row its own angles
    prm = rng.random((3, qc.n_true_params)) * 2 * np.pi
    st = qc.run_batch(prm).cpu().numpy()
    for b in range(3):
        assert np.max(np.abs(st[b] - OS.run(og, prm[b], n))) < 1e-12
    # the gate on its own: operation * ket
    g = pyqc.ARBGATE(H1)
    g.set_theta(0.8)
    v = ref / np.linalg.norm(ref)
    out = (g * pyqc.qobj.to_ket(v, n)).full().ravel()
    assert np.max(np.abs(out - OS.apply_prims(v, OC.ARBGATE(H1, n).prims(0.8), n))) < 1e-12


def test_get_gradients_compat_index_lookup(pyqc):
    """get_gradients(compat=True) reproduces the reference's two-parameter
    branch (circuit.py:186-191: the second row of the fSim at parameterised
    index count is taken of self.gates[count]) against the oracle's
    restatement of that loop; compat=False gives the standard rows."""
    n = 3
    specs = [
        ([pyqc.H(0, n), pyqc.R_y(0, n), pyqc.fSim([0, 1], n), pyqc.CNOT([1, 2], n), pyqc.fSim([1, 2], n)],
         [OC.H(0, n), OC.R_y(0, n), OC.fSim((0, 1), n), OC.CNOT((1, 2), n), OC.fSim((1, 2), n)]),
        ([pyqc.CNOT([0, 1], n), pyqc.fSim([0, 1], n), pyqc.R_x(2, n)],
         [OC.CNOT((0, 1), n), OC.fSim((0, 1), n), OC.R_x(2, n)]),
        ([pyqc.R_y(0, n), pyqc.fSim([0, 1], n), pyqc.H(1, n), pyqc.R_z(1, n), pyqc.fSim([1, 2], n)],
         [OC.R_y(0, n), OC.fSim((0, 1), n), OC.H(1, n), OC.R_z(1, n), OC.fSim((1, 2), n)]),
    ]
    rng = np.random.default_rng(77)
    for gates, og in specs:
        qc = pyqc.PQC(n)
        qc.add_layer(gates)
        th = rng.random(qc.n_true_params) * 2 * np.pi
        qc.set_params(list(th))
        want = OS.derivative_states_compat(og, th, n)
        got = qc.get_gradients(compat=True)
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert np.max(np.abs(a.full().ravel() - b)) < 1e-12
        std = OS.derivative_states(og, th, n)
        for a, b in zip(qc.get_gradients(), std):
            assert np.max(np.abs(a.full().ravel() - b)) < 1e-12


def test_find_overparam_point_and_eff_H(pyqc):
    """find_overparam_point (measure.py:101-121) adds copies of a layer until
    the QFIM rank stops changing: replayed with the oracle on the same
    gates.rng stream, the count and every rank agree.  find_eff_H
    (measure.py:199-224) on samples of the Haar fidelity law for dimension 8
    lands where the same minimiser over the oracle's expr lands, near 8."""
    import scipy.optimize
    from pyramaterised_amd import gates as G
    from pyramaterised_amd import templates as T
    n = 3
    qc = pyqc.PQC(n)
    qc.add_layer(T.clifford_HE_layers(1, n)[0])
    m = pyqc.measure.Measurements(qc)
    saved = G.rng
    try:
        G.rng = np.random.default_rng(5)
        count = m.find_overparam_point([0])
    finally:
        G.rng = saved
    rng = np.random.default_rng(5)
    prev, diff, c, layers, ranks = 0, 1, 0, 1, []
    while diff > 1e-3 and c < 50:
        layers += 1
        og = OC.clifford_he(layers, n)
        th = np.array([rng.random(1)[0] * 2 * np.pi for _ in range(2 * n * layers)])
        F = OM.qfim(OS.derivative_states(og, th, n), OS.run(og, th, n))
        rank = np.linalg.matrix_rank(F)
        ranks.append(rank)
        diff, prev, c = abs(rank - prev), rank, c + 1
    assert count == c and ranks[-1] == 2 ** (n + 1) - 2, (count, c, ranks)
    assert qc.n_layers == layers
    # find_eff_H: F = 1 - U^(1/(N-1)) follows (N-1)(1-F)^(N-2)
    N = 8
    Fs = 1 - np.random.default_rng(6).random(10000) ** (1 / (N - 1))
    x = m.find_eff_H(Fs, n)
    ref = scipy.optimize.minimize(lambda d, F: OM.expr(F, d), [4], args=(Fs,), method="BFGS")
    assert ref.success and abs(x - ref.x[0]) < 1e-6, (x, ref.x[0])
    assert abs(x - N) < 1.0, x


def test_sampling_measures_replayed_with_the_oracle(pyqc):
    """expressibility, entanglement and entropy_of_magic (measure.py:182-197,
    239-249, 351-359) draw their states from gates.rng as successive
    run("random") calls would; replayed with the oracle on the same stream
    they agree.  theta_to_magic / theta_to_gkp (measure.py:251-266) against
    Pauli-string brute force; get_eigenvalues / new_measure (measure.py:73-99)
    on the GPU QFIM against the oracle's; the two setters."""
    from pyramaterised_amd import gates as G
    from pyramaterised_amd import templates as T
    n, p = 4, 2
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    og = OC.clifford_he(p, n)
    P = qc.n_true_params
    m = pyqc.measure.Measurements(qc)

    def replay(seed, S):
        rng = np.random.default_rng(seed)
        return [OS.run(og, np.array([rng.random(1)[0] * 2 * np.pi for _ in range(P)]), n) for _ in range(S)]

    saved = G.rng
    try:
        G.rng = np.random.default_rng(11)
        ex = m.expressibility(40)
        G.rng = np.random.default_rng(12)
        ent = m.entanglement(10)
        G.rng = np.random.default_rng(13)
        eom = m.entropy_of_magic(5)
    finally:
        G.rng = saved
    states = replay(11, 40)
    assert abs(ex - OM.expr(OM.pair_fidelities(states), 2 ** n)) < 1e-9
    assert np.max(np.abs(np.array(ent) - [OM.single_Q(s, n) for s in replay(12, 10)])) < 1e-12
    assert abs(eom - np.mean([sre_bruteforce(s, n, 2) for s in replay(13, 5)])) < 1e-10
    th = list(np.random.default_rng(14).random(P) * 2 * np.pi)
    ref = OS.run(og, np.array(th), n)
    assert abs(m.theta_to_magic(th) + sre_bruteforce(ref, n, 2)) < 1e-10
    assert abs(m.theta_to_gkp(th) + sre_bruteforce(ref, n, 0.5) / (2 * np.log(2))) < 1e-10
    qc.state = qc.run(th)
    F = m.get_QFI()
    Fref = OM.qfim(OS.derivative_states(og, np.array(th), n), ref)
    ev, _ = m.get_eigenvalues(F)
    assert np.max(np.abs(ev - np.linalg.eigvalsh(Fref))) < 1e-9
    assert abs(m.new_measure(F) - OM.new_measure(Fref)) < 1e-9
    f = lambda x: 0.0  # noqa: E731
    m.set_minimise_function(f)
    assert m.minimize_function is f
    mats = m.get_conversion_matrices()
    m.set_converstion_matrices(mats)
    assert m.conversion_matrices is mats


def test_pqc_layer_state_and_repr_methods(pyqc):
    """set_layer / get_layer (circuit.py:45-51) replace a layer, as the
    oracle's gate list with that layer replaced; update_state and fidelity
    (circuit.py:127-142); flip_deriv (circuit.py:144-147) leaves every
    derivative state as it was (the reference flips ``pauli`` after the
    operators are built); __repr__ (circuit.py:194-198) lists the layers."""
    from pyramaterised_amd import templates as T
    n = 4
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(2, n):
        qc.add_layer(l)
    ry_layer = [pyqc.R_y(q, n) for q in range(n)]
    qc.set_layer(ry_layer, 1)
    assert qc.get_layer(1) is not ry_layer and len(qc.get_layer(1)) == n  # a deep copy, as the reference
    og = OC.clifford_he(1, n) + [OC.R_y(q, n) for q in range(n)]
    th = np.random.default_rng(21).random(qc.n_true_params) * 2 * np.pi
    st = qc.update_state(list(th))
    assert qc.state is st
    ref = OS.run(og, th, n)
    assert infidelity(st.full().ravel(), ref) < 1e-12
    other = OS.run(og, th[::-1].copy(), n)
    assert abs(qc.fidelity(pyqc.qobj.to_ket(other, n)) - abs(np.vdot(ref, other)) ** 2) < 1e-12
    _, J0 = qc.jacobian()
    qc.flip_deriv()
    _, J1 = qc.jacobian()
    assert (J0 - J1).abs().max().item() == 0.0
    r = repr(qc)
    assert r.startswith(f"A {n} qubit, {qc.n_layers} layer deep PQC. \n")
    assert r.count("Layer ") == qc.n_layers and "R_y(" in r.split("Layer 1:")[1]
