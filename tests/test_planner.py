"""CPU tests of the planner (libpqs host code, no GPU) and of the compiled op
lists of the drop-in API against the oracle."""
import numpy as np
import pytest

from oracle import circuits as OC
from oracle import gates as OG
from oracle import sim as OS
from tests.helpers import build_oracle, build_pkg, random_spec

L = pytest.importorskip("pyramaterised_amd._lib")


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def op_matrix(o, params):
    """(qubits, local matrix) of one pqs_op, restated from include/pqs.h."""
    ang = o.offset + (o.scale * params[o.slot] if o.slot >= 0 else 0.0)
    k = o.kind
    m = np.array([complex(o.m[2 * i], o.m[2 * i + 1]) for i in range(16)])
    if k == L.OP_MAT1:
        return (o.q0,), m[:4].reshape(2, 2)
    if k in (L.OP_RX, L.OP_RY, L.OP_RZ):
        return (o.q0,), {L.OP_RX: OG.rx, L.OP_RY: OG.ry, L.OP_RZ: OG.rz}[k](ang)
    if k == L.OP_MAT2:
        return (o.q0, o.q1), m.reshape(4, 4)
    if k == L.OP_CX:
        return (o.q0, o.q1), OG.CNOT
    if k == L.OP_CZ:
        return (o.q0, o.q1), OG.CZ
    if k in (L.OP_RXX, L.OP_RYY, L.OP_RZZ):
        return (o.q0, o.q1), OG.rpp(ang, {L.OP_RXX: "x", L.OP_RYY: "y", L.OP_RZZ: "z"}[k])
    if k == L.OP_FSIM:
        ph = params[o.slot2] if o.slot2 >= 0 else 0.0
        return (o.q0, o.q1), OG.fsim(ang, ph)
    raise ValueError(k)


def simulate_ops(ops, order, n, params, init=None):
    psi = OS.zero_state(n) if init is None else init.copy()
    for i in order:
        qs, m = op_matrix(ops[i], params)
        psi = OS.apply(psi, qs, m, n)
    return psi


def active_qubits(o):
    k = o.kind
    if k in (L.OP_RZ, L.OP_CZ, L.OP_RZZ):
        return set()
    if k == L.OP_MAT1:
        return set() if (o.m[2] == 0 and o.m[3] == 0 and o.m[4] == 0 and o.m[5] == 0) else {o.q0}
    if k == L.OP_CX:
        return {o.q1}
    if k in (L.OP_RX, L.OP_RY):
        return {o.q0}
    if k == L.OP_GEN:  # generator term: Pauli on q0 (and q1)
        return set() if o.pauli in (0, 3) else {q for q in (o.q0, o.q1) if q >= 0}
    return {o.q0, o.q1}


def test_library_loads_and_exports_every_header_symbol():
    import os
    import re
    lib = L.lib()
    assert lib.pqs_abi_version() == 1
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "pqs.h")).read()
    names = set(re.findall(r"\b(pqs_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    for nm in names:
        assert hasattr(lib, nm), nm
    assert set(L.SIGNATURES) == names


def test_entry_points_reject_null_context_without_touching_the_gpu():
    """Every compute entry point checks its context / plan / buffer
    arguments first: a NULL context returns PQS_ERR_ARG (< 0) with a
    message, and nothing is launched (this runs without a GPU)."""
    import ctypes as C
    lib = L.lib()
    skip = {"pqs_abi_version", "pqs_last_error", "pqs_ctx_create", "pqs_ctx_destroy", "pqs_plan_destroy",
            "pqs_plan_create", "pqs_batch_table_bytes", "pqs_jit_prepare", "pqs_range_pop"}
    checked = 0
    for name, (res, args) in L.SIGNATURES.items():
        if name in skip:
            continue
        vals = []
        for t in args:
            if t in (C.c_int, C.c_int32, C.c_int64):
                vals.append(0)
            elif t is C.c_double:
                vals.append(0.0)
            else:
                vals.append(None)
        rc = getattr(lib, name)(*vals)
        assert rc < 0, name
        assert lib.pqs_last_error(), name
        checked += 1
    assert checked >= 20
    # destroying NULL is a no-op, as free(NULL) is
    assert lib.pqs_ctx_destroy(None) == 0
    assert lib.pqs_plan_destroy(None) == 0


def test_plan_rejects_bad_ops(pyqc):
    from pyramaterised_amd import engine as E
    from pyramaterised_amd.gates import _op
    with pytest.raises(L.EngineError):
        E.Plan(3, [_op(L.OP_RX, 3, slot=0)], tile_bits=3)
    with pytest.raises(L.EngineError):
        E.Plan(3, [_op(L.OP_CX, 1, 1)], tile_bits=3)
    with pytest.raises(L.EngineError):
        E.Plan(3, [_op(L.OP_RX, 0, slot=0)], tile_bits=14)
    # a single generator term wider than the tile cannot be scheduled
    from pyramaterised_amd.gates import _gen
    with pytest.raises(L.EngineError):
        E.Plan(3, [_gen(0, 1, [0, 1], -0.5j)], tile_bits=1, low_bits=1)


def _gen_parts(plan, ops):
    """Per tangent: the generator terms of each part, from the sweeps' op
    indices (a merged generator reports the index of its first term)."""
    reported = {i for _, idx in plan.sweeps() for i in idx}
    parts = {}
    i = 0
    while i < len(ops):
        if ops[i].kind != L.OP_GEN:
            i += 1
            continue
        j = i
        while j < len(ops) and ops[j].kind == L.OP_GEN and ops[j].tangent == ops[i].tangent:
            j += 1
        starts = [k for k in range(i, j) if k in reported]
        assert starts and starts[0] == i
        parts.setdefault(ops[i].tangent, []).extend(list(range(a, b)) for a, b in zip(starts, starts[1:] + [j]))
        i = j
    return parts


def _gen_term(o):
    code = o.pauli
    P = [OG.I2, OG.SX, OG.SY, OG.SZ][code]
    qs = (o.q0,) if o.q1 < 0 else (o.q0, o.q1)
    m = P if len(qs) == 1 else np.kron(P, P)
    return qs, complex(o.m[0], o.m[1]) * m


def _derivative_by_parts(plan, ops, n, params, t, parts, init=None):
    """d psi / d theta_t as the engine forms it: every part of tangent t's
    generator evolved through the sweep order on its own, then summed."""
    order = [i for _, idx in plan.sweeps() for i in idx]
    gen_first = {part[0]: part for ps in parts.values() for part in ps}
    total = np.zeros(2 ** n, dtype=complex)
    for part in parts[t]:
        psi = OS.zero_state(n) if init is None else init.copy()
        for i in order:
            o = ops[i]
            if o.kind == L.OP_GEN:
                if i != part[0]:
                    continue
                acc = np.zeros_like(psi)
                for k in gen_first[i]:
                    qs, m = _gen_term(ops[k])
                    acc += OS.apply(psi, qs, m, n)
                psi = acc
                continue
            qs, m = op_matrix(o, params)
            psi = OS.apply(psi, qs, m, n)
        total += psi
    return total


@pytest.mark.parametrize("rolling", [False, True])
def test_wide_generators_split_into_parts_match_oracle(pyqc, rolling):
    """A shared-parameter generator wider than the tile (TFIM_layers' shared
    R_x over all N qubits, templates.py:194-204; reference derivative
    gates.py:451-457) is split into parts that each fit the tile; summing the
    parts' derivative states, each evolved through the plan's sweep order,
    gives the oracle's derivative states."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p = 9, 2
    qc = pyqc.PQC(n)
    for l in T.TFIM_layers(p, n):
        qc.add_layer(l)
    ops = qc.ops(True)
    flags = L.PQS_PLAN_ROLLING if rolling else 0
    plan = E.Plan(n, ops, tile_bits=6, low_bits=2, flags=flags)
    parts = _gen_parts(plan, ops)
    for t in range(qc.n_true_params):
        want = 3 if t % 2 == 1 else 1  # R_x generators: 9 qubits in parts of <= k - 2 = 4; R_zz sums are diagonal
        assert plan.tangent_parts(t) == len(parts[t]) == want, t
    rng = np.random.default_rng(3)
    th = rng.random(qc.n_true_params) * 2 * np.pi
    og = OC.tfim(p, n)
    ref = OS.derivative_states(og, th, n)
    for t in range(qc.n_true_params):
        d = _derivative_by_parts(plan, ops, n, th, t, parts)
        assert np.max(np.abs(d - ref[t])) < 1e-12, t


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_tfim_derivative_plans_at_13_to_24_qubits(pyqc, dtype):
    """Derivative plans of the reference's TFIM templates plan at every width
    from 13 to 24 qubits with the derivative tile shapes (12-bit c128,
    13-bit c64): the shared R_x generators split into parts within the tile,
    every part's terms sit in the sweep that runs it, and the R_zz sums
    (diagonal) stay whole."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    for n in range(13, 25):
        for layers in (T.TFIM_layers(2, n), T.tfim_open_layers(2, n), T.modified_TFIM_layers(1, n)):
            qc = pyqc.PQC(n, dtype=dtype)
            for l in layers:
                qc.add_layer(l)
            ops = qc.ops(True)
            plan = E.Plan(n, ops, dtype=dtype, **qc.plan_options(True, dtype))
            k = plan.tile_bits
            parts = _gen_parts(plan, ops)
            sweeps = plan.sweeps()
            for t in range(qc.n_true_params):
                np_t = plan.tangent_parts(t)
                assert np_t == len(parts[t]) >= 1
                for part in parts[t]:
                    act = set()
                    for i in part:
                        act |= active_qubits(ops[i])
                    assert len(act) <= k
                    holders = [tq for tq, idx in sweeps if part[0] in idx]
                    assert len(holders) == 1 and act <= set(holders[0])
            wide = [t for t in range(qc.n_true_params) if plan.tangent_parts(t) > 1]
            assert bool(wide) == (n > k), (n, k)


@pytest.mark.parametrize("n,tile,low,seed", [(6, 3, 2, 0), (8, 5, 3, 1), (10, 4, 1, 2), (12, 6, 3, 3),
                                              (16, 13, 5, 4)])
def test_sweep_order_is_equivalent_to_program_order(pyqc, n, tile, low, seed):
    from pyramaterised_amd import engine as E
    rng = np.random.default_rng(seed)
    spec = [s for s in random_spec(n, 3, rng) if s[0] != "shared_Rx"]
    qc = pyqc.PQC(n, tile_bits=tile, low_bits=low)
    qc.add_layer(build_pkg(spec, n, pyqc))
    ops = qc.ops()
    plan = E.Plan(n, ops, tile_bits=tile, low_bits=low)
    sweeps = plan.sweeps()
    order = [i for _, idx in sweeps for i in idx]
    assert sorted(order) == list(range(len(ops)))
    for tq, idx in sweeps:
        assert len(tq) == min(tile, n)
        for i in idx:
            assert active_qubits(ops[i]) <= set(tq)
    params = rng.random(qc.n_true_params) * 2 * np.pi
    a = simulate_ops(ops, range(len(ops)), n, params)
    b = simulate_ops(ops, order, n, params)
    assert np.max(np.abs(a - b)) < 1e-12
    ref = OS.run(build_oracle(spec, n), params, n)
    assert np.max(np.abs(a - ref)) < 1e-12


def test_brickwork_28q_fuses_two_sweeps_per_layer(pyqc):
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(28, tile_bits=13, low_bits=5)
    for l in T.ry_cz_brickwork_layers(30, 28):
        qc.add_layer(l)
    plan = E.Plan(28, qc.ops(), tile_bits=13, low_bits=5)
    assert len(qc.ops()) == 1245
    assert plan.n_sweeps <= 62


def test_compiled_ops_match_oracle_templates(pyqc):
    """The package's templates compile to the oracle's gate sequence."""
    from pyramaterised_amd import templates as T
    cases = [
        (T.clifford_HE_layers(3, 5), OC.clifford_he(3, 5), 5, None),
        (T.hea_linear_layers(3, 4), OC.hea_linear(3, 4), 4, None),
        (T.generic_HE_layers(2, 6), OC.generic_he(2, 6), 6, None),
        (T.TFIM_layers(2, 4), OC.tfim(2, 4), 4, None),
        (T.ry_cz_brickwork_layers(4, 7), OC.ry_cz_brickwork(4, 7), 7, None),
        (T.NPQC_layers(4, 6)[0], OC.npqc(4, 6)[0], 6, None),
    ]
    qg_p = pyqc.PQC(4)
    for l in T.qg_circuit_layers(1, 4):
        qg_p.add_layer(l)
    rng = np.random.default_rng(0)
    for layers, og, n, init in cases:
        qc = pyqc.PQC(n)
        for l in layers:
            qc.add_layer(l)
        assert qc.n_true_params == OS.n_params(og)
        th = rng.random(qc.n_true_params) * 2 * np.pi
        a = simulate_ops(qc.ops(), range(len(qc.ops())), n, th)
        assert np.max(np.abs(a - OS.run(og, th, n))) < 1e-12


def test_pqc_bookkeeping_quirks(pyqc):
    """circuit.py:53-72: parameterised has a -1 after every gate, n_params is
    twice the parameter count; set_params binds in gate order."""
    n = 3
    qc = pyqc.PQC(n)
    qc.add_layer([pyqc.R_x(0, n), pyqc.H(1, n), pyqc.fSim([0, 2], n)])
    assert qc.parameterised == [2, -1, -1, 5, 6, -1]
    assert qc.n_params == 6
    assert qc.n_true_params == 3
    qc.set_params([0.1, 0.2, 0.3])
    assert qc.get_params() == [0.1, 0.2, 0.3]
    qc2 = pyqc.PQC(2)
    qc2.add_layer([pyqc.negative_R_z(0, 2), pyqc.offset_R_z(1, 2, 0.5)])
    qc2.set_params([0.4, 0.7])
    # get_params returns the stored (transformed) angles like the reference
    assert np.allclose(qc2.get_params(), [-0.4, 1.2])
    assert np.allclose(qc2._gate_angles(), [0.4, 0.7])


def test_random_params_equal_sequential_draws(pyqc):
    from pyramaterised_amd import gates as G
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(3)
    for l in T.clifford_HE_layers(2, 3):
        qc.add_layer(l)
    G.rng = np.random.default_rng(1)
    batch = qc.random_params(4)
    G.rng = np.random.default_rng(1)
    seq = []
    for _ in range(4):
        qc.set_params("random")
        seq.append(qc._angles.copy())
    assert np.array_equal(batch, np.array(seq))


def test_pauli_sum_algebra(pyqc):
    from pyramaterised_amd import templates as T
    H = T.TFIM_hamiltonian(4, g=0.6, h=0.2)
    Z = np.diag([1, -1])
    X = np.array([[0, 1], [1, 0]])

    def emb(m, q, n):
        return OS.kron_1q(m, q, n)
    ref = np.zeros((16, 16), dtype=complex)
    for i in range(4):
        ref += emb(Z, i, 4) @ emb(Z, (i + 1) % 4, 4) + 0.6 * emb(X, i, 4) + 0.2 * emb(Z, i, 4)
    assert np.allclose(H.full(), -ref)
    Y = pyqc.PauliSum.single(2, 0, 2)
    Xs = pyqc.PauliSum.single(2, 0, 1)
    Zs = pyqc.PauliSum.single(2, 0, 3)
    assert np.allclose((Xs * Y).full(), (1j * Zs).full())
    dense = pyqc.PauliSum.from_dense(pyqc.Qobj(H.full()))
    assert np.allclose(dense.full(), H.full())


def test_specialised_kernels_compile_without_gpu(pyqc, tmp_path, monkeypatch):
    """The run-time specialised sweep source compiles with hiprtc (no device
    needed) for a multi-sweep plan with every register op kind."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    monkeypatch.setenv("PQS_JIT_CACHE", str(tmp_path))
    n = 12
    qc = pyqc.PQC(n, jit=False)
    for l in T.clifford_HE_layers(2, n):
        qc.add_layer(l)
    qc.add_layer([pyqc.R_x(0, n), pyqc.H(3, n), pyqc.CZ([1, 9], n), pyqc.R_zz([2, 11], n), pyqc.T(5, n)])
    plan = E.Plan(n, qc.ops(), tile_bits=10, low_bits=3)
    srcs = plan.jit_sweep_sources()
    assert plan.n_sweeps > 1 and len(srcs) == plan.n_sweeps
    plan.jit(load=False, workers=1)


def _layout_checks(plan, n, k, low, ops):
    lay = plan.layouts()
    sweeps = plan.sweeps()
    std = [n - 1 - q for q in range(n)]
    assert lay[0][0] == std and lay[-1][1] == std
    for s, (r, w) in enumerate(lay):
        assert sorted(r) == list(range(n)) and sorted(w) == list(range(n))
        if s + 1 < len(lay):
            assert w == lay[s + 1][0]
        tq = sweeps[s][0]
        assert len(tq) == k
        if s > 0:
            assert sorted(r[q] for q in tq) == list(range(k))  # contiguous tile read
        # the lowest `low` write positions hold tile qubits (contiguous write runs)
        low_w = sorted(w[q] for q in tq)[:low]
        assert low_w == list(range(low)), (s, low_w)
        for i in sweeps[s][1]:
            assert active_qubits(ops[i]) <= set(tq)


@pytest.mark.parametrize("n,tile,low,seed", [(14, 10, 3, 5), (16, 12, 4, 6), (20, 12, 3, 7), (24, 13, 3, 8)])
def test_rolling_layouts_random_circuits(pyqc, n, tile, low, seed):
    """Rolling layouts: every layout a permutation, consecutive sweeps agree
    on the layout between them, standard at both ends, contiguous tile
    reads, at least `low` contiguous write bits; op order equivalent."""
    from pyramaterised_amd import engine as E
    rng = np.random.default_rng(seed)
    spec = [s for s in random_spec(n, 4, rng) if s[0] != "shared_Rx"]
    qc = pyqc.PQC(n)
    qc.add_layer(build_pkg(spec, n, pyqc))
    ops = qc.ops()
    plan = E.Plan(n, ops, tile_bits=tile, low_bits=low, flags=L.PQS_PLAN_ROLLING)
    assert plan.n_sweeps > 1 and plan.rolling
    _layout_checks(plan, n, tile, low, ops)
    order = [i for _, idx in plan.sweeps() for i in idx]
    assert sorted(order) == list(range(len(ops)))
    if n <= 16:
        params = rng.random(qc.n_true_params) * 2 * np.pi
        a = simulate_ops(ops, range(len(ops)), n, params)
        b = simulate_ops(ops, order, n, params)
        assert np.max(np.abs(a - b)) < 1e-12


def test_rolling_layouts_cut_sweeps_at_28q(pyqc):
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(28)
    for l in T.ry_cz_brickwork_layers(30, 28):
        qc.add_layer(l)
    ops = qc.ops()
    fixed = E.Plan(28, ops, tile_bits=12, low_bits=4)
    roll = E.Plan(28, ops, tile_bits=12, low_bits=4, flags=L.PQS_PLAN_ROLLING)
    assert not fixed.rolling and roll.rolling
    assert roll.n_sweeps < 0.6 * fixed.n_sweeps
    _layout_checks(roll, 28, 12, 4, ops)


def test_window_search_never_adds_sweeps(pyqc):
    """The window sweep search builds partitions with look-aheads of 1, 2
    and 3 sweeps and keeps one only when it has fewer sweeps than the plain
    greedy partition, so it never uses more; on nearest-neighbour circuits it
    uses far fewer.  Its plans keep the rolling-layout invariants and an op
    order equivalent to program order."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    R, G = L.PQS_PLAN_ROLLING, L.PQS_PLAN_GREEDY_SWEEPS

    def counts(qc, k, low):
        ops = qc.ops()
        a = E.Plan(qc.n_qubits, ops, tile_bits=k, low_bits=low, flags=R)
        b = E.Plan(qc.n_qubits, ops, tile_bits=k, low_bits=low, flags=R | G)
        _layout_checks(a, qc.n_qubits, k, low, ops)
        return a, b

    bw = pyqc.PQC(28)
    for l in T.ry_cz_brickwork_layers(30, 28):
        bw.add_layer(l)
    a, b = counts(bw, 12, 4)
    assert (a.n_sweeps, b.n_sweeps) == (11, 17)
    he = pyqc.PQC(20)
    for l in T.clifford_HE_layers(40, 20):
        he.add_layer(l)
    a, b = counts(he, 12, 4)
    assert a.n_sweeps <= 16 and b.n_sweeps >= 38
    he16 = pyqc.PQC(16)
    for l in T.clifford_HE_layers(20, 16):
        he16.add_layer(l)
    a, b = counts(he16, 11, 3)
    assert a.n_sweeps < b.n_sweeps
    ops = he16.ops()
    order = [i for _, idx in a.sweeps() for i in idx]
    assert sorted(order) == list(range(len(ops)))
    params = np.random.default_rng(9).random(he16.n_true_params) * 2 * np.pi
    x = simulate_ops(ops, range(len(ops)), 16, params)
    y = simulate_ops(ops, order, 16, params)
    assert np.max(np.abs(x - y)) < 1e-12
    for seed in range(4):
        rng = np.random.default_rng(50 + seed)
        spec = [s for s in random_spec(18, 5, rng) if s[0] != "shared_Rx"]
        qc = pyqc.PQC(18)
        qc.add_layer(build_pkg(spec, 18, pyqc))
        a, b = counts(qc, 12, 3)
        assert a.n_sweeps <= b.n_sweeps


def _phase_counts(plan):
    """Register phases per sweep, counted in the specialised sweep kernel's
    source (each register phase declares its pending-flip flags once; the
    last sweep's pair kernel, which repeats them, follows it)."""
    return [src.decode().split("void pqs_jit_pair_")[0].count("int xf0 = 0") for src in plan.jit_sweep_sources()]


def test_phase_search_never_adds_phases(pyqc):
    """The register-phase window search keeps a seeded register set only
    when the greedy pass then needs fewer phases for the rest of the run, so
    no sweep gets more phases than with greedy phases; on the HEA and
    brickwork circuits it removes 14% of them."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    R, GP = L.PQS_PLAN_ROLLING, L.PQS_PLAN_GREEDY_PHASES

    def both(qc, k, low):
        a = E.Plan(qc.n_qubits, qc.ops(), tile_bits=k, low_bits=low, flags=R)
        b = E.Plan(qc.n_qubits, qc.ops(), tile_bits=k, low_bits=low, flags=R | GP)
        assert a.n_sweeps == b.n_sweeps
        pa, pb = _phase_counts(a), _phase_counts(b)
        assert all(x <= y for x, y in zip(pa, pb)), (pa, pb)
        return sum(pa), sum(pb)

    he = pyqc.PQC(16)
    for l in T.clifford_HE_layers(20, 16):
        he.add_layer(l)
    assert both(he, 12, 4) == (60, 70)
    bw = pyqc.PQC(28)
    for l in T.ry_cz_brickwork_layers(30, 28):
        bw.add_layer(l)
    assert both(bw, 12, 4) == (65, 76)
    for seed in range(3):
        rng = np.random.default_rng(70 + seed)
        spec = [s for s in random_spec(14, 5, rng) if s[0] != "shared_Rx"]
        qc = pyqc.PQC(14)
        qc.add_layer(build_pkg(spec, 14, pyqc))
        a, b = both(qc, 11, 3)
        assert a <= b


def test_rolling_layouts_with_tangents(pyqc):
    """Derivative plans (generator terms for every parameter) under rolling
    layouts: the layout invariants hold, fewer sweeps than fixed layouts,
    and every tangent is spawned in a sweep that holds its generator."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(16)
    for l in T.clifford_HE_layers(10, 16):
        qc.add_layer(l)
    ops = qc.ops(True)
    fixed = E.Plan(16, ops, tile_bits=12, low_bits=3)
    plan = E.Plan(16, ops, tile_bits=12, low_bits=3, flags=L.PQS_PLAN_ROLLING)
    assert plan.rolling and not fixed.rolling
    assert plan.n_sweeps < fixed.n_sweeps
    _layout_checks(plan, 16, 12, 3, ops)
    sweeps = plan.sweeps()
    for t in range(qc.n_true_params):
        s0 = plan.tangent_sweep(t)
        gen = [i for i, o in enumerate(ops) if o.kind == L.OP_GEN and o.tangent == t]
        assert gen and all(i in sweeps[s0][1] for i in gen), t
    # the Python layer picks rolling derivative plans while the ping-pong
    # rows fit the cap
    assert qc.plan_options(True)["flags"] & L.PQS_PLAN_ROLLING
    big = pyqc.PQC(22)
    for l in T.clifford_HE_layers(40, 22):
        big.add_layer(l)
    assert not big.plan_options(True).get("flags", 0) & L.PQS_PLAN_ROLLING


def test_tuner_pick_prefers_earlier_candidates_within_tolerance():
    """The tuner takes the first candidate in list order whose time is within
    its tolerance of the fastest, so near-ties do not flip between boxes, and
    a clearly faster later candidate still wins."""
    from pyramaterised_amd import engine as E
    assert E.pick_candidate([10.0, 9.9, 12.0]) == 0          # 1% slower: a tie, the first wins
    assert E.pick_candidate([10.0, 9.5, 12.0]) == 1          # 5% faster later candidate wins
    assert E.pick_candidate([12.0, 10.1, 10.0]) == 1         # first within 1.5% of the fastest
    assert E.pick_candidate([10.0, 9.9], tolerance=0.0) == 1
    assert E.pick_candidate([7.0]) == 0
    # the c128 candidate list starts with the 11-bit rolling shape
    assert E.tile_candidates(16, 0)[0] == (11, 3, E.L.PQS_PLAN_ROLLING)


def test_arbgate_has_no_op_form(pyqc):
    """ARBGATE (gates.py:407-435) has no local-gate form: it yields no op
    records (PQC applies it as a dense operator between segment plans)."""
    qc = pyqc.PQC(3)
    H = np.diag(np.arange(8.0)).astype(complex)
    qc.add_layer([pyqc.R_y(0, 3), pyqc.ARBGATE(H)])
    with pytest.raises(NotImplementedError):
        qc.ops()


def test_planning_cost_stays_bounded(pyqc):
    """Plan creation stays fast where the searches have the most candidates:
    a 12-qubit, 100-layer circuit in one tile (a single long register-phase
    run) and a 24-qubit, 100-layer circuit under rolling layouts (many
    sweeps)."""
    import time
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    for n, d, k in ((12, 100, 12), (24, 100, 12)):
        qc = pyqc.PQC(n)
        for l in T.clifford_HE_layers(d, n):
            qc.add_layer(l)
        t = time.perf_counter()
        plan = E.Plan(n, qc.ops(), tile_bits=k, low_bits=4, flags=L.PQS_PLAN_ROLLING)
        assert time.perf_counter() - t < 10.0, (n, d)
        assert plan.n_sweeps >= 1


def _xxz_pair(pyqc, n):
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n)
    for l in T.XXZ_layers(2, n):
        qc.add_layer(l)
    return qc, OC.xxz(2, n)


def _noncommuting(pyqc, n):
    """A hardware-efficient layer, a genuinely non-commuting commute=False
    block [R_x(0), R_y(0), R_zz(0, 1)], then more gates."""
    pk = [pyqc.R_y(q, n) for q in range(n)] + [pyqc.CNOT([q, q + 1], n) for q in range(n - 1)]
    pk += [pyqc.shared_parameter([pyqc.R_x(0, n), pyqc.R_y(0, n), pyqc.R_zz([0, 1], n)], n, commute=False)]
    pk += [pyqc.R_x(q, n) for q in range(n)] + [pyqc.CZ([0, n - 1], n)]
    og = [OC.R_y(q, n) for q in range(n)] + [OC.CNOT((q, q + 1), n) for q in range(n - 1)]
    og += [OC.shared_parameter([OC.R_x(0, n), OC.R_y(0, n), OC.R_zz((0, 1), n)], n, commute=False)]
    og += [OC.R_x(q, n) for q in range(n)] + [OC.CZ((0, n - 1), n)]
    qc = pyqc.PQC(n)
    qc.add_layer(pk)
    return qc, og


@pytest.mark.parametrize("case", ["xxz", "noncommuting"])
def test_commute_false_derivatives_match_oracle(pyqc, case):
    """shared_parameter(commute=False) (gates.py:458-466): XXZ blocks (R_yy,
    R_xx on disjoint pairs: symmetric, commuting strings) compile to the
    generator sum; a genuinely non-commuting block gets a plan of its own
    (the circuit with conj(U) and the interleaved D_k inserted) whose
    tangent, evolved part by part through the sweep order, is the oracle's
    derivative state (dense restatement of the reference's formula)."""
    from pyramaterised_amd import engine as E
    n = 6 if case == "xxz" else 5
    qc, og = (_xxz_pair if case == "xxz" else _noncommuting)(pyqc, n)
    rng = np.random.default_rng(11)
    th = rng.random(qc.n_true_params) * 2 * np.pi
    ref = OS.derivative_states(og, th, n)
    specials = qc.special_slots()
    if case == "xxz":
        assert specials == []
    else:
        assert len(specials) == 1
    ops = qc.ops(True)
    plan = E.Plan(n, ops, tile_bits=4, low_bits=2)
    parts = _gen_parts(plan, ops)
    for t in range(qc.n_true_params):
        if t in specials:
            sops = qc.special_ops(t)
            sp = E.Plan(n, sops, tile_bits=4, low_bits=2)
            sparts = _gen_parts(sp, sops)
            assert sp.n_tangents == 1 and sp.tangent_parts(0) == len(sparts[0]) == 3
            d = _derivative_by_parts(sp, sops, n, th, 0, sparts)
        else:
            d = _derivative_by_parts(plan, ops, n, th, t, parts)
        assert np.max(np.abs(d - ref[t])) < 1e-12, t


GENERATE_CASES = [
    # (generate_circuit type, kwargs, n, p, oracle builder, start state)
    ("NPQC", {}, 4, 3, lambda p, n: OC.npqc(p, n)[0], None),
    ("TFIM", {}, 4, 2, OC.tfim, None),
    ("TFIM_modified", {}, 5, 2, OC.modified_tfim, None),
    ("XXZ", {}, 6, 2, OC.xxz, "half"),
    ("Circuit_1", {}, 4, 2, OC.circuit_1, None),
    ("Circuit_2", {}, 5, 2, OC.circuit_2, None),
    ("Circuit_9", {}, 4, 2, OC.circuit_9, None),
    ("qg_circuit", {}, 4, 2, OC.qg_circuit, None),
    ("generic_HE", {}, 5, 2, OC.generic_he, None),
    ("clifford", {}, 4, 2, OC.clifford_he, None),
    ("y_CPHASE", {}, 5, 3, OC.y_cphase, None),
    ("double_y_CPHASE", {}, 4, 2, OC.double_y_cphase, None),
    ("fermionic", {}, 6, 1, OC.fermionic, "half"),
    ("zfsim", {}, 4, 3, lambda p, n: OC.fsim_circuit(p, n, "z"), "half"),
    ("fsim", {}, 5, 4, lambda p, n: OC.fsim_circuit(p, n, "y"), "half"),
    ("fsim", {"rotator": "x"}, 4, 2, lambda p, n: OC.fsim_circuit(p, n, "x"), "half"),
    ("fixed_fsim", {}, 5, 3, lambda p, n: OC.fsim_circuit(p, n, "z", fixed_gate=True), "half"),
]


@pytest.mark.parametrize("case", GENERATE_CASES, ids=[c[0] + str(c[1].get("rotator", "")) for c in GENERATE_CASES])
def test_generate_circuit_types_match_oracle_templates(pyqc, case):
    """Every circuit type of generate_circuit (templates.py:361-429) compiles
    to the oracle's restated gate sequence and start state (shuffle=False)."""
    from pyramaterised_amd import templates as T
    name, kw, n, p, build, start = case
    qc = T.generate_circuit(name, n, p, shuffle=False, **kw)
    og = build(p, n)
    assert qc.n_true_params == OS.n_params(og)
    init = OC.half_filled(n) if start == "half" else None
    th = np.random.default_rng(len(name)).random(qc.n_true_params) * 2 * np.pi
    a = simulate_ops(qc.ops(), range(len(qc.ops())), n, th, init=init)
    assert np.max(np.abs(a - OS.run(og, th, n, init=init))) < 1e-12
    if start == "half":
        assert np.allclose(np.asarray(qc.initial_state.full()).ravel(), init)


def _adjoint_cpu(prog, psi, lam, params, n):
    """The adjoint backward pass with the program's op lists on the CPU."""
    grad = np.zeros(prog.P)
    for ops, group in prog.step_ops:
        lam = simulate_ops(ops, range(len(ops)), n, params, init=lam)
        psi = simulate_ops(ops, range(len(ops)), n, params, init=psi)
        for slot, terms in group:
            for c, x, z in terms:
                Pm = np.array([[1.0 + 0j]])
                for q in range(n):
                    bx, bz = (x >> q) & 1, (z >> q) & 1
                    Pm = np.kron(Pm, OG.I2 if not (bx or bz) else OG.SX if not bz else OG.SZ if not bx else OG.SY)
                grad[slot] += 2 * np.real(c * np.vdot(lam, Pm @ psi))
    return grad


@pytest.mark.parametrize("which", ["clifford", "tfim", "random"])
def test_adjoint_program_matches_oracle_gradient(pyqc, which):
    """The adjoint program's steps (inverse ops between measurement points,
    grouped generator terms), run on the CPU, give the oracle's energy
    gradient 2 Re <H psi|d_p psi> for H = Z0 Z1 (measure.py:461-471); an
    HEA layer needs two measurement points."""
    from pyramaterised_amd import adjoint
    from pyramaterised_amd import templates as T
    n = 5
    rng = np.random.default_rng(9)
    if which == "clifford":
        layers, og = T.clifford_HE_layers(3, n), OC.clifford_he(3, n)
    elif which == "tfim":
        layers, og = T.TFIM_layers(2, n), OC.tfim(2, n)
    else:
        spec = [s for s in random_spec(n, 3, rng, fsim=False) if s[0] != "sqrtH"]
        layers, og = [build_pkg(spec, n, pyqc)], build_oracle(spec, n)
    qc = pyqc.PQC(n)
    for l in layers:
        qc.add_layer(l)
    prog = adjoint.AdjointProgram(qc)
    if which == "clifford":
        assert len(prog.step_ops) == 2 * 3
    th = rng.random(qc.n_true_params) * 2 * np.pi
    psi = OS.run(og, th, n)
    zz = np.where(((np.arange(2 ** n) >> (n - 1)) ^ (np.arange(2 ** n) >> (n - 2))) & 1, -1.0, 1.0)
    g = _adjoint_cpu(prog, psi, zz * psi, th, n)
    d = OS.derivative_states(og, th, n)
    want = np.array([2 * np.real(np.vdot(zz * psi, di)) for di in d])
    assert np.max(np.abs(g - want)) < 1e-12


def test_adjoint_program_refuses_non_unitary_gates(pyqc):
    """sqrtH (element-wise square root of H, gates.py:235-242) is not
    unitary: the adjoint pass cannot uncompute it, so it declines (and the
    callers take the Jacobian route)."""
    from pyramaterised_amd import adjoint
    qc = pyqc.PQC(2)
    qc.add_layer([pyqc.R_x(0, 2), pyqc.sqrtH(0, 2), pyqc.R_y(1, 2)])
    with pytest.raises(NotImplementedError):
        adjoint.AdjointProgram(qc)


def test_exchange_transitions_compile_and_keep_layouts(pyqc, tmp_path, monkeypatch):
    """The headline circuit's 11-bit rolling plan trades register and lane
    bits between a dozen register phases (cross-lane swaps in the
    specialised kernels); its sources compile with hiprtc, and with
    PQS_PLAN_NO_EXCHANGE no phase exchanges and no swap is emitted."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    monkeypatch.setenv("PQS_JIT_CACHE", str(tmp_path))
    n = 16
    qc = pyqc.PQC(n, jit=False)
    for l in T.clifford_HE_layers(20, n):
        qc.add_layer(l)
    R = L.PQS_PLAN_ROLLING
    plan = E.Plan(n, qc.ops(), tile_bits=11, low_bits=3, flags=R)
    srcs = plan.jit_sweep_sources()
    n_swaps = sum(s.count(b"lswap<") for s in srcs)  # calls (the prelude defines lswap( )
    assert n_swaps >= 8 * 8  # at least 8 exchanged register bits (8 pairs each)
    plan2 = E.Plan(n, qc.ops(), tile_bits=11, low_bits=3, flags=R | L.PQS_PLAN_NO_EXCHANGE)
    assert sum(s.count(b"lswap<") for s in plan2.jit_sweep_sources()) == 0
    E._compile_in_children(srcs[:2], 2)
    plan.jit(load=False, workers=1)


def test_profiler_ranges_without_and_with_roctx():
    """pqs_range_push / pop: no-ops without PQS_ROCTX, a NULL name is
    rejected; with PQS_ROCTX=1 (a fresh process) the library loads
    the roctx library and nested ranges push and pop without a GPU."""
    import os
    import subprocess
    import sys
    lib = L.lib()
    assert lib.pqs_range_push(b"outer") == 0 and lib.pqs_range_pop() == 0
    assert lib.pqs_range_push(None) == L.PQS_ERR_ARG
    code = ("import ctypes, sys; lib = ctypes.CDLL(sys.argv[1]); "
            "rc = [lib.pqs_range_push(b'a'), lib.pqs_range_push(b'b'), lib.pqs_range_pop(), lib.pqs_range_pop()]; "
            "assert rc == [0, 0, 0, 0], rc; print('ok')")
    env = dict(os.environ, PQS_ROCTX="1")
    out = subprocess.run([sys.executable, "-c", code, L.LIB_PATH], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0 and "ok" in out.stdout, out.stderr
