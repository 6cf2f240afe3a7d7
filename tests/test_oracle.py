"""Pins of the CPU oracle against the reference's known answers, brute force
and closed forms.  CPU only."""
import numpy as np
import pytest

from oracle import circuits as C
from oracle import gates as OG
from oracle import measures as M
from oracle import sim as S

QG_ENERGY = 0.46135870050914374  # tests.py:207-209
QG_ROUNDED_ANGLES = [3.21587011, 5.97193953, 0.90578156, 5.96054027, 1.9592948,
                     2.65983852, 5.20060878, 2.571074, 3.45319898, 0.17315902,
                     4.73446249, 3.38125416]  # tests.py:192-205


def zz01(psi, n):
    """<Z0 Z1> (PQC.set_H('ZZ'), circuit.py:28-31; qt.expect at circuit.py:136)."""
    idx = np.arange(2 ** n)
    b0 = (idx >> (n - 1)) & 1
    b1 = (idx >> (n - 2)) & 1
    sign = np.where(b0 ^ b1, -1.0, 1.0)
    return float(np.sum(sign * np.abs(psi) ** 2))


def rng_draws(k, seed=1):
    """gates.rng = default_rng(1) (gates.py:10); one rng.random(1)[0]*2*pi per
    parameter (circuit.py:112)."""
    r = np.random.default_rng(seed)
    return [r.random(1)[0] * 2 * np.pi for _ in range(k)]


def test_rng_stream_matches_rounded_test_angles():
    exact = rng_draws(12)
    assert np.max(np.abs(np.array(exact) - QG_ROUNDED_ANGLES)) < 1e-8
    # one (k,) draw is bit-identical to k single draws
    batch = np.random.default_rng(1).random(12) * 2 * np.pi
    assert np.array_equal(batch, np.array(exact))


def test_qg_energy_golden():
    n, gates = C.qg_test_circuit()
    psi = S.run(gates, rng_draws(12), n)
    assert abs(zz01(psi, n) - QG_ENERGY) < 1e-14
    psi_r = S.run(gates, QG_ROUNDED_ANGLES, n)
    assert abs(zz01(psi_r, n) - QG_ENERGY) < 1e-5


def test_qg_energy_rejects_wrong_chain_order():
    """A sequential 0->1->2->3 chain gives a different energy: the golden value
    pins the CHAIN order of gates.py:368-375."""
    n, gates = C.qg_test_circuit()
    seq = S.OGate("seq", 0, lambda t, p: [((i, i + 1), OG.CNOT) for i in range(3)])
    gates2 = [seq if g.name == "CHAIN" else g for g in gates]
    psi = S.run(gates2, rng_draws(12), n)
    assert abs(zz01(psi, n) - QG_ENERGY) > 1e-2


def test_qg_eqd_golden():
    n, gates = C.qg_test_circuit()
    th = QG_ROUNDED_ANGLES
    psi = S.run(gates, th, n)
    F = M.qfim(S.derivative_states(gates, th, n), psi)
    assert M.effective_quantum_dimension(F, 1e-12) == 12


@pytest.mark.parametrize("n", [4, 6, 8])
def test_npqc_qfim_identity(n):
    """tests.py:103-128: QFIM == identity at the reference angles."""
    for p in range(1, 2 ** (n // 2) + 1):
        if n == 8 and p not in (1, 2, 5, 16):
            continue  # keep the CPU suite short; all P run at n = 4, 6
        gates, th = C.npqc(p, n)
        psi = S.run(gates, th, n)
        F = M.qfim(S.derivative_states(gates, th, n), psi)
        assert np.max(np.abs(F - np.eye(len(th)))) < 1e-12


def test_basic_gates():
    """tests.py:64-87."""
    psi = S.run([C.fixed_R_y(0, 1, np.pi / 2)], [], 1)
    assert abs(psi[1].real - 1 / np.sqrt(2)) < 1e-15
    psi = S.run([C.H(0, 1), C.H(0, 1)], [], 1)
    assert np.allclose(psi, [1, 0], atol=1e-15)


def test_textbook_identities():
    assert np.allclose(OG.hadamard(), np.array([[1, 1], [1, -1]]) / np.sqrt(2))
    assert np.allclose(OG.s_gate() @ OG.s_gate(), OG.SZ)
    assert np.allclose(OG.t_gate() @ OG.t_gate(), OG.s_gate())
    for f in (OG.rx, OG.ry, OG.rz):
        u = f(0.37)
        assert np.allclose(u.conj().T @ u, np.eye(2))
    for p in "xyz":
        u = OG.rpp(1.1, p)
        assert np.allclose(u.conj().T @ u, np.eye(4))
    # R_zz is diagonal with phases e^{-it/2} on even parity, e^{+it/2} on odd
    t = 0.8
    assert np.allclose(OG.rpp(t, "z"), np.diag(np.exp(-0.5j * t * np.array([1, -1, -1, 1]))))
    # rotations are exp(-i t P / 2)
    import scipy.linalg
    for a in "xyz":
        assert np.allclose(OG.ROT[a](0.9), scipy.linalg.expm(-0.45j * OG.PAULI[a]))


def test_fsim_matches_generator_form():
    """fSim(theta, phi) (gates.py:598-606) equals exp(-i theta (XX + YY) / 2)
    followed by a phase e^{-i phi} on |11>: the exchange block and the sign
    of both angles are pinned by this independent construction."""
    import scipy.linalg
    X, Y = OG.PAULI["x"], OG.PAULI["y"]
    gen = (np.kron(X, X) + np.kron(Y, Y)) / 2
    for t, f in ((0.0, 0.0), (0.37, 1.2), (2.9, -0.4), (np.pi / 2, np.pi / 6)):
        want = np.diag([1, 1, 1, np.exp(-1j * f)]) @ scipy.linalg.expm(-1j * t * gen)
        assert np.allclose(OG.fsim(t, f), want, atol=1e-14), (t, f)
        u = OG.fsim(t, f)
        assert np.allclose(u.conj().T @ u, np.eye(4))
    # iSWAP-like point: theta = pi/2 moves |01> to -i|10>
    assert np.allclose(OG.fsim(np.pi / 2, 0.0) @ np.array([0, 1, 0, 0]), [0, 0, -1j, 0])


def _random_circuit(n, rng):
    gates = []
    for _ in range(3):
        for q in range(n):
            gates.append([C.R_x, C.R_y, C.R_z][rng.integers(3)](q, n))
        a, b = rng.choice(n, 2, replace=False)
        gates.append([C.CNOT, C.CZ, C.sqrtiSWAP][rng.integers(3)]((int(a), int(b)), n))
        a, b = rng.choice(n, 2, replace=False)
        gates.append([C.R_zz, C.R_xx, C.R_yy][rng.integers(3)]((int(a), int(b)), n))
        gates.append(C.fSim((int(b), int(a)), n))
        gates += [C.H(rng.integers(n), n), C.T(rng.integers(n), n), C.S(rng.integers(n), n)]
        gates.append(C.negative_R_z(rng.integers(n), n))
        gates.append(C.offset_R_z(rng.integers(n), n, 0.3))
    gates.append(C.CHAIN(C.CNOT, n))
    gates.append(C.shared_parameter([C.R_x(q, n) for q in range(n)], n))
    return gates


@pytest.mark.parametrize("n", [2, 3, 5])
def test_einsum_sim_matches_bruteforce(n):
    rng = np.random.default_rng(7 + n)
    gates = _random_circuit(n, rng)
    th = rng.random(S.n_params(gates)) * 2 * np.pi
    psi = S.run(gates, th, n)
    u = S.dense_unitary(gates, th, n)
    assert np.allclose(u.conj().T @ u, np.eye(2 ** n), atol=1e-12)
    assert np.max(np.abs(psi - u[:, 0])) < 1e-13


def test_embed_matches_kron_adjacent():
    m = OG.SQRT_ISWAP @ OG.CNOT
    assert np.allclose(S.embed(m, (1, 2), 4), np.kron(np.kron(np.eye(2), m), np.eye(2)))
    # reversed order = swap-conjugated matrix
    sw = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(S.embed(m, (2, 1), 4), np.kron(np.kron(np.eye(2), sw @ m @ sw), np.eye(2)))


@pytest.mark.parametrize("maker", ["hea", "tfim", "npqc"])
def test_derivative_states_match_central_fd(maker):
    """For gates whose D satisfies D U = dU/dtheta the reference's derivative
    states are the true partial derivatives; central FD at h = 1e-6."""
    rng = np.random.default_rng(3)
    init = None
    if maker == "hea":
        n, gates = 4, C.hea_linear(2, 4)
    elif maker == "tfim":
        n = 5
        gates, init = C.tfim_open(2, n)
    else:
        n = 4
        gates, _ = C.npqc(3, n)
    th = rng.random(S.n_params(gates)) * 2 * np.pi
    ds = S.derivative_states(gates, th, n, init)
    h = 1e-6
    for p in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (S.run(gates, tp, n, init) - S.run(gates, tm, n, init)) / (2 * h)
        assert np.max(np.abs(fd - ds[p])) < 1e-8


def test_single_ry_qfim_is_one():
    gates = [C.R_y(0, 1)]
    for t in (0.1, 1.3, 2.9):
        psi = S.run(gates, [t], 1)
        F = M.qfim(S.derivative_states(gates, [t], 1), psi)
        assert abs(F[0, 0] - 1) < 1e-14


def test_mw_bell_and_product():
    bell11 = np.array([0, 1, -1, 0]) / np.sqrt(2)  # qutip bell_state('11')
    assert abs(M.single_Q(bell11, 2) - 1) < 1e-14  # tests.py:289-295
    prod = S.product_state(np.array([np.cos(0.3), np.exp(0.2j) * np.sin(0.3)]), 5)
    assert abs(M.single_Q(prod, 5)) < 1e-14
    ghz = np.zeros(8)
    ghz[0] = ghz[7] = 1 / np.sqrt(2)
    assert abs(M.single_Q(ghz, 3) - 1) < 1e-14


def test_reduced_matches_partial_trace_definition():
    rng = np.random.default_rng(5)
    n = 4
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    psi /= np.linalg.norm(psi)
    rho = np.outer(psi, psi.conj()).reshape([2] * 8)
    for k in range(n):
        # trace every axis but k by explicit einsum subscripts
        letters = "abcdefgh"
        row = list(letters[:n])
        col = list(letters[n:])
        for j in range(n):
            if j != k:
                col[j] = row[j]
        sub = "".join(row) + "".join(col) + "->" + row[k] + col[k]
        ref = np.einsum(sub, rho)
        assert np.allclose(M.reduced_1q(psi, k, n), ref, atol=1e-14)


def test_expr_idle_closed_form():
    """All F == 1 (idle circuit): all mass in the last bin, so
    KL = -log(P_haar[last])."""
    F = [1.0] * 300
    dim = 2
    bins = int(0.0075 * 300)
    _, mids = M.gen_histo(F)
    haar = (dim - 1) * (1 - mids) ** (dim - 2)
    ph = haar / haar.sum()
    assert len(mids) == bins
    assert abs(M.expr(F, dim) - (-np.log(ph[-1]))) < 1e-12


def test_expr_haar_samples_small():
    """Haar states in dimension 4: KL against the analytic density is small."""
    rng = np.random.default_rng(11)
    d = 4
    v = rng.normal(size=(2 * 4000, d)) + 1j * rng.normal(size=(2 * 4000, d))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    F = np.abs(np.sum(v[0::2].conj() * v[1::2], axis=1)) ** 2
    assert M.expr(list(F), d, bins=30) < 0.02
    assert M.expr(list(np.ones(4000)), d, bins=30) > 1.0


def test_noncommuting_shared_reduces_for_diagonal_block():
    """For a diagonal block conj(U) = U^dagger, so the commute=False derivative
    of gates.py:458-466 equals the commute=True one."""
    n = 3
    layer = [C.R_zz((0, 1), n), C.R_z(2, n), C.R_zz((1, 2), n)]
    a = C.shared_parameter(layer, n, commute=True)
    b = C.shared_parameter(layer, n, commute=False)
    psi = S.product_state(np.array([1, 1]) / np.sqrt(2), n)
    t = 0.77
    da = S.apply_deriv(S.apply_prims(psi, a.prims(t), n), a.deriv(t), n)
    db = S.apply_deriv(S.apply_prims(psi, b.prims(t), n), b.deriv(t), n)
    assert np.max(np.abs(da - db)) < 1e-14


def test_sqrt_iswap_pins():
    """sqrtiSWAP (gates.py:344-349, QuTiP sqrtiswap): its square is iSWAP,
    and it is exp(i pi/8 (XX + YY)), the half-way point of the XY exchange
    that reaches iSWAP at pi/4."""
    import scipy.linalg
    iswap = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
    assert np.allclose(OG.SQRT_ISWAP @ OG.SQRT_ISWAP, iswap, atol=1e-15)
    X, Y = OG.PAULI["x"], OG.PAULI["y"]
    xy = np.kron(X, X) + np.kron(Y, Y)
    assert np.allclose(OG.SQRT_ISWAP, scipy.linalg.expm(1j * np.pi / 8 * xy), atol=1e-15)
    assert np.allclose(iswap, scipy.linalg.expm(1j * np.pi / 4 * xy), atol=1e-15)


def test_rr_family_is_exp_of_pauli_pair():
    """R_xx / R_yy / R_zz (gates.py:509-517) = exp(-i t/2 P(x)P)."""
    import scipy.linalg
    for p in "xyz":
        PP = np.kron(OG.PAULI[p], OG.PAULI[p])
        for t in (0.0, 0.37, 2.1, -1.3, np.pi):
            assert np.allclose(OG.rpp(t, p), scipy.linalg.expm(-0.5j * t * PP), atol=1e-14), (p, t)


def test_fsim_derivative_matrices_against_central_differences():
    """The fSim-family "derivative" matrices as the reference writes them
    (gates.py:609-648, 719-737): in fsim_gate_d_theta the exchange block is
    d/dtheta of fSim's, and the corners are fSim's own (1 and e^{-i phi},
    kept, not differentiated); in fsim_gate_d_phi the |11> entry is
    d/dphi of fSim's and the rest is fSim's; fixed_fsim_gate_d_theta is the
    exchange-block derivative with both corners 1.  Central differences,
    h = 1e-6."""
    h = 1e-6
    for t, f in ((0.3, 0.9), (2.2, -1.1), (np.pi / 2, np.pi / 3), (5.0, 4.0)):
        fs = OG.fsim(t, f)
        dt = (OG.fsim(t + h, f) - OG.fsim(t - h, f)) / (2 * h)
        df = (OG.fsim(t, f + h) - OG.fsim(t, f - h)) / (2 * h)
        m = OG.fsim_d_theta(t, f)
        assert np.allclose(m[1:3, 1:3], dt[1:3, 1:3], atol=1e-9)
        assert m[0, 0] == 1 and m[3, 3] == fs[3, 3]
        off = np.ones((4, 4), bool)
        off[0, 0] = off[3, 3] = False
        off[1:3, 1:3] = False
        assert np.all(m[off] == 0)
        m2 = OG.fsim_d_phi(t, f)
        assert abs(m2[3, 3] - df[3, 3]) < 1e-9
        keep = np.ones((4, 4), bool)
        keep[3, 3] = False
        assert np.array_equal(m2[keep], fs[keep])
        m3 = OG.fixed_fsim_d_theta(t)
        dfix = (OG.fixed_fsim(t + h) - OG.fixed_fsim(t - h)) / (2 * h)
        assert np.allclose(m3[1:3, 1:3], dfix[1:3, 1:3], atol=1e-9)
        assert m3[0, 0] == 1 and m3[3, 3] == 1 and np.all(m3[off] == 0)


def test_alltoall_is_the_product_over_ordered_pairs():
    """ALLTOALL (gates.py:382-401) applies the entangler on every ordered pair
    of permutations(range(N), 2), the first listed first: against a product
    of CNOT permutation matrices built from bit flips (no embed)."""
    from itertools import permutations
    n = 4
    dim = 2 ** n
    want = np.eye(dim, dtype=complex)
    for a, b in permutations(range(n), 2):
        P = np.zeros((dim, dim))
        for i in range(dim):
            j = i ^ (1 << (n - 1 - b)) if (i >> (n - 1 - a)) & 1 else i
            P[j, i] = 1
        want = P @ want
    got = S.dense_unitary([C.ALLTOALL(C.CNOT, n)], [], n)
    assert np.array_equal(got, want)
    # the order matters: the reversed product differs
    rev = np.eye(dim, dtype=complex)
    for a, b in reversed(list(permutations(range(n), 2))):
        P = np.zeros((dim, dim))
        for i in range(dim):
            j = i ^ (1 << (n - 1 - b)) if (i >> (n - 1 - a)) & 1 else i
            P[j, i] = 1
        rev = P @ rev
    assert not np.array_equal(got, rev)


def test_identity_gate_and_new_measure_closed_forms():
    """I (gates.py:150-156): the state is unchanged and the derivative state
    is -i/2 psi (PRot.derivative with the identity, gates.py:133-138).
    new_measure (measure.py:89-99): the sum of eigenvalues capped at 1."""
    n = 3
    gates = [C.R_y(0, n), C.CNOT((0, 2), n), C.I(1, n), C.R_x(2, n)]
    th = np.array([0.7, 1.9, 2.4])
    psi = S.run(gates, th, n)
    assert np.allclose(psi, S.run([gates[0], gates[1], gates[3]], th[[0, 2]], n), atol=1e-15)
    d = S.derivative_states(gates, th, n)
    assert np.allclose(d[1], -0.5j * psi, atol=1e-15)
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)))
    ev = np.array([0.25, 3.0, 1.0, 0.0])
    F = Q @ np.diag(ev) @ Q.T
    assert abs(M.new_measure(F) - 2.25) < 1e-12
    assert abs(M.new_measure(np.eye(5)) - 5) < 1e-12


def test_arbgate_restatement_pins():
    """ARBGATE (gates.py:407-435) restated with scipy's expm: for a Pauli
    string P, expm(-i t P) = cos t I - i sin t P (note: no half angle, unlike
    the rotations), and the derivative state is -i/2 H U psi, half of the
    central difference of the state (the reference's factor)."""
    n = 3
    P = np.kron(np.kron(OG.SX, OG.SZ), OG.SY)
    g = C.ARBGATE(P, n)
    t = 0.73
    (qs, U), = g.prims(t)
    assert qs == (0, 1, 2)
    assert np.allclose(U, np.cos(t) * np.eye(8) - 1j * np.sin(t) * P, atol=1e-14)
    rng = np.random.default_rng(2)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    H = A + A.conj().T
    gates = [C.R_y(0, n), C.ARBGATE(H, n), C.R_x(2, n)]
    th = np.array([0.4, 0.9, 1.7])
    d = S.derivative_states(gates, th, n)
    h = 1e-6
    fd = (S.run(gates, th + [0, h, 0], n) - S.run(gates, th - [0, h, 0], n)) / (2 * h)
    assert np.max(np.abs(d[1] - 0.5 * fd)) < 1e-8


def _quirk_circuits():
    # (gates, n, expected map: row -> ("d", i) derivative_states row i, or "psi")
    n = 3
    a = [C.H(0, n), C.R_y(0, n), C.fSim((0, 1), n), C.CNOT((1, 2), n), C.fSim((1, 2), n)]
    # parameterised: R_y (rows 0), fSim01 (1, 2), fSim12 (3, 4); fSim01 is
    # parameterised gate 1 -> gates[1] = R_y; fSim12 is gate 2 -> gates[2] = fSim01
    b = [C.CNOT((0, 1), n), C.fSim((0, 1), n), C.R_x(2, n)]
    # fSim is parameterised gate 0 -> gates[0] = CNOT (fixed): the plain state
    c = [C.R_y(0, n), C.fSim((0, 1), n), C.H(1, n), C.R_z(1, n), C.fSim((1, 2), n)]
    # rows: R_y 0, fSim01 1-2, R_z 3, fSim12 4-5; fSim01 is parameterised
    # gate 1 -> gates[1] = itself; fSim12 is gate 3 -> gates[3] = R_z
    return [(a, n, {2: ("d", 0), 4: ("d", 2)}), (b, n, {1: "psi"}), (c, n, {5: ("d", 3)})]


def test_compat_gradients_restate_the_index_lookup():
    """derivative_states_compat against derivative_states: rows other than
    the looked-up ones are equal; the looked-up row of a two-parameter gate
    equals the derivative state of the gate at that circuit position (its
    phi state for a two-parameter gate) or the plain state for a fixed
    gate; with every gate before each fSim parameterised the two agree."""
    rng = np.random.default_rng(31)
    for gates, n, remap in _quirk_circuits():
        th = rng.random(S.n_params(gates)) * 2 * np.pi
        std = S.derivative_states(gates, th, n)
        cmp = S.derivative_states_compat(gates, th, n)
        psi = S.run(gates, th, n)
        assert len(cmp) == len(std)
        for r in range(len(std)):
            want = std[r] if r not in remap else (psi if remap[r] == "psi" else std[remap[r][1]])
            assert np.max(np.abs(cmp[r] - want)) < 1e-14, r
    gates = [C.fSim((0, 1), 2), C.R_y(0, 2), C.fSim((0, 1), 2)]  # positions and indices coincide
    th = rng.random(S.n_params(gates)) * 2 * np.pi
    for x, y in zip(S.derivative_states_compat(gates, th, 2), S.derivative_states(gates, th, 2)):
        assert np.max(np.abs(x - y)) < 1e-15
    # a looked-up R_y row: central differences of the state (the fSim rows
    # are the reference's derivative matrices, not derivatives; their
    # blocks are pinned by test_fsim_derivative_matrices_against_central_differences)
    gates, n, _ = _quirk_circuits()[0]
    th = rng.random(S.n_params(gates)) * 2 * np.pi
    h = 1e-6
    std = S.derivative_states_compat(gates, th, n)
    for row in (0, 2):  # R_y's own row, and fSim01's second row (looked up: R_y again)
        p = 0
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (S.run(gates, tp, n) - S.run(gates, tm, n)) / (2 * h)
        assert np.max(np.abs(fd - std[row])) < 1e-8
