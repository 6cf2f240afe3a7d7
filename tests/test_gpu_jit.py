"""Run-time specialised sweep kernels (hiprtc) against the CPU oracle and
against the interpreted kernels."""
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


def _build(pyqc, n, layers, jit, **kw):
    qc = pyqc.PQC(n, jit=jit, **kw)
    for l in layers:
        qc.add_layer(l)
    return qc


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_jit_brickwork_22q(pyqc, dtype, tol):
    from pyramaterised_amd import templates as T
    n, p = 22, 30
    qc = _build(pyqc, n, T.ry_cz_brickwork_layers(p, n), True, dtype=dtype)
    assert qc.plan().jitted
    th = np.random.default_rng(4).random(qc.n_true_params) * 2 * np.pi
    st = qc.run(list(th)).full().ravel()
    ref = OS.run(OC.ry_cz_brickwork(p, n), th, n)
    assert infidelity(st, ref) < tol


def test_jit_clifford_he_16q_batch_matches_interpreter_and_oracle(pyqc):
    from pyramaterised_amd import templates as T
    n, p, S = 16, 6, 6
    qj = _build(pyqc, n, T.clifford_HE_layers(p, n), True)
    qi = _build(pyqc, n, T.clifford_HE_layers(p, n), False)
    params = np.random.default_rng(5).random((S, qj.n_true_params)) * 2 * np.pi
    sj = qj.run_batch(params).cpu().numpy()
    si = qi.run_batch(params).cpu().numpy()
    og = OC.clifford_he(p, n)
    for s in range(S):
        ref = OS.run(og, params[s], n)
        assert infidelity(sj[s], ref) < 1e-12
        assert np.max(np.abs(sj[s] - si[s])) < 1e-13


def test_jit_all_reg_kinds_multi_sweep(pyqc):
    """R_x, R_y, R_z, fixed 1q, CNOT, CZ, R_zz with qubits inside and outside
    the tile and register bits, many sweeps (small tiles)."""
    rng = np.random.default_rng(6)
    n = 14
    kinds1 = ["R_x", "R_y", "R_z", "H", "T", "S", "X"]
    kinds2 = ["CNOT", "CZ", "R_zz"]
    spec = []
    for _ in range(6):
        for q in range(n):
            spec.append((kinds1[rng.integers(len(kinds1))], q))
        for _ in range(n):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            spec.append((kinds2[rng.integers(len(kinds2))], (a, b)))
    pk, og = [], []
    for name, q in spec:
        if isinstance(q, tuple):
            pk.append(getattr(pyqc, name)(list(q), n))
            og.append(getattr(OC, name)(q, n))
        else:
            pk.append(getattr(pyqc, name)(q, n))
            og.append(getattr(OC, name)(q, n))
    for tile in (10, 11, 13):
        qc = _build(pyqc, n, [pk], True, tile_bits=tile, low_bits=3)
        plan = qc.plan()
        assert plan.jitted and plan.n_sweeps > 1
        th = rng.random(qc.n_true_params) * 2 * np.pi
        st = qc.run(list(th)).full().ravel()
        ref = OS.run(og, th, n)
        assert infidelity(st, ref) < 1e-12
        assert np.max(np.abs(st - ref)) < 1e-12


def test_jit_jacobian_and_qfim(pyqc):
    """Forward Jacobian mixing specialised sweeps with interpreted generator
    sweeps; QFIM against the reference algorithm."""
    from pyramaterised_amd import templates as T
    n, p = 14, 2
    qc = _build(pyqc, n, T.clifford_HE_layers(p, n), True, tile_bits=11, low_bits=3)
    th = np.random.default_rng(7).random(qc.n_true_params) * 2 * np.pi
    qc.state = qc.run(list(th))
    psi, J = qc.jacobian()
    og = OC.clifford_he(p, n)
    ref = OS.run(og, th, n)
    dref = OS.derivative_states(og, th, n)
    Jh = J.cpu().numpy()
    for t in range(0, len(dref), 7):
        assert np.max(np.abs(Jh[t] - dref[t])) < 1e-12
    F = pyqc.measure.Measurements(qc).get_QFI()
    assert rel_fro(F, OM.qfim(dref, ref)) < 1e-9


def test_jit_product_init_and_edge_angles(pyqc):
    """|+>^n start (product init) and angles at multiples of pi, where
    cos(a/2) is ~0 and the scaled rotation form is clamped."""
    from pyramaterised_amd import templates as T
    n, p = 12, 3
    qc = _build(pyqc, n, T.clifford_HE_layers(p, n), True)
    qc.set_initial_state(T.plus_state())
    th = np.random.default_rng(8).choice([0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi, 3 * np.pi],
                                         size=qc.n_true_params)
    st = qc.run(list(th)).full().ravel()
    init = OS.product_state(np.array([1, 1]) / np.sqrt(2), n)
    ref = OS.run(OC.clifford_he(p, n), th, n, init)
    assert infidelity(st, ref) < 1e-12
    assert np.max(np.abs(st - ref)) < 1e-12


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_jit_many_rotations_at_pi_in_one_phase(pyqc, dtype, tol):
    """Deep stacks of rotations at and next to pi, where each unscaled
    rotation grows amplitudes by up to 1/cos(a/2): without intermediate
    scaling, c64 overflows after three of them in one register phase and
    c128 after about two dozen."""
    rng = np.random.default_rng(13)
    n = 12
    spec = []
    for _ in range(10):  # no entanglers: all layers fuse into the same phases
        spec += [("R_y", q) for q in range(n)]
    for layer in range(8):
        spec += [(("R_x", "R_y", "R_z")[layer % 3], q) for q in range(n)]
        spec += [("CZ", (q, q + 1)) for q in range(layer % 2, n - 1, 2)]
        spec += [("R_zz", (q, q + 1)) for q in range((layer + 1) % 2, n - 1, 2)]
    pk = [getattr(pyqc, nm)(list(q) if isinstance(q, tuple) else q, n) for nm, q in spec]
    og = [getattr(OC, nm)(q, n) for nm, q in spec]
    qc = _build(pyqc, n, [pk], True, dtype=dtype)
    assert qc.plan().jitted
    P = qc.n_true_params
    for th in (np.full(P, np.pi), np.pi + rng.choice([-1e-9, 0.0, 1e-9], size=P),
               rng.choice([np.pi, -np.pi, 3 * np.pi], size=P)):
        st = qc.run(list(th)).full().ravel()
        assert np.all(np.isfinite(st))
        ref = OS.run(og, th, n)
        assert infidelity(st, ref) < tol
        assert np.max(np.abs(st - ref)) < (1e-12 if dtype == "complex128" else 1e-5)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_jit_tile_shapes_persistent_and_tuner(pyqc, dtype, tol):
    """Every tile shape the tuner times, the persistent form, and the tuned
    default, against the oracle on a batch."""
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p, S = 20, 6, 6
    og = OC.ry_cz_brickwork(p, n)
    rng = np.random.default_rng(12)
    P = n * p
    params = rng.random((S, P)) * 2 * np.pi
    refs = [OS.run(og, params[s], n) for s in range(S)]
    cands = E.tile_candidates(n, E.dtype_code(dtype))
    # the tuner's shapes plus the persistent form (more tiles than resident
    # blocks: each block loops and prefetches)
    shapes = cands + [(13 if dtype == "complex128" else 14, 3, E.L.PQS_PLAN_JIT_PERSISTENT)]
    for k, low, fl in shapes:
        qc = _build(pyqc, n, T.ry_cz_brickwork_layers(p, n), True, dtype=dtype, tile_bits=k, low_bits=low, flags=fl)
        plan = qc.plan(batch=S)
        assert plan.jitted and plan.tile_bits == k
        st = qc.run_batch(params).cpu().numpy()
        for s in range(S):
            assert infidelity(st[s], refs[s]) < tol, (k, low, fl, s)
    qc = _build(pyqc, n, T.ry_cz_brickwork_layers(p, n), True, dtype=dtype)
    plan = qc.plan(batch=S)
    assert plan.jitted and plan.tile_bits in [c[0] for c in cands]
    st = qc.run_batch(params).cpu().numpy()
    for s in range(S):
        assert infidelity(st[s], refs[s]) < tol


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
@pytest.mark.parametrize("jit", [True, False])
def test_rolling_layouts_match_oracle(pyqc, dtype, tol, jit):
    """Rolling-layout plans (out of place, layout changes between sweeps)
    in batches and single runs, specialised and interpreted."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import templates as T
    n, p, S = 20, 8, 3
    qc = _build(pyqc, n, T.ry_cz_brickwork_layers(p, n), jit, dtype=dtype, tile_bits=12, low_bits=3,
                flags=L.PQS_PLAN_ROLLING)
    plan = qc.plan(batch=S)
    assert plan.rolling and plan.n_sweeps > 2 and plan.jitted == jit
    og = OC.ry_cz_brickwork(p, n)
    params = np.random.default_rng(40).random((S, n * p)) * 2 * np.pi
    st = qc.run_batch(params).cpu().numpy()
    for s in range(S):
        assert infidelity(st[s], OS.run(og, params[s], n)) < tol
    one = qc.run(list(params[1])).full().ravel()
    assert np.max(np.abs(one - st[1])) < (1e-13 if dtype == "complex128" else 1e-6)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_rolling_jacobian_matches_fixed_and_oracle(pyqc, dtype, tol):
    """Forward-mode Jacobian under rolling layouts (every tangent row
    ping-pongs with a workspace row) against the fixed-layout plan and the
    oracle's per-parameter re-simulation, from |0> and from |+>^n; and the
    QFIM from it."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p = 14, 8
    og = OC.clifford_he(p, n)
    th = np.random.default_rng(43).random(2 * n * p) * 2 * np.pi
    plus = OS.product_state(np.array([1, 1]) / np.sqrt(2), n)
    lim = 1e-12 if dtype == "complex128" else 2e-5
    for init in (None, plus):
        dref = OS.derivative_states(og, th, n, init)
        ref = OS.run(og, th, n, init)
        got = {}
        for fl in (L.PQS_PLAN_ROLLING, 0):
            qc = _build(pyqc, n, T.clifford_HE_layers(p, n), True, dtype=dtype, tile_bits=10, low_bits=3, flags=fl)
            if init is not None:
                qc.set_initial_state(T.plus_state())
            qc.state = qc.run(list(th))
            psi, J = qc.jacobian()
            plan = qc.plan(with_generators=True)
            assert plan.rolling == bool(fl) and plan.jitted
            got[fl] = (psi.cpu().numpy(), J.cpu().numpy())
            assert infidelity(got[fl][0], ref) < tol
            for t in range(0, len(dref), 5):
                assert np.max(np.abs(got[fl][1][t] - dref[t])) < lim, (fl, t)
            if dtype == "complex128":
                F = pyqc.measure.Measurements(qc).get_QFI()
                assert rel_fro(F, OM.qfim(dref, ref)) < 1e-9
        a, b = got.values()
        assert np.max(np.abs(a[1] - b[1])) < lim


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_direct_load_store_match_staged_tiles(pyqc, dtype, tol):
    """Sweeps whose first / last register phase loads / stores registers
    directly (no LDS staging) against the same plans staged through LDS
    (PQS_PLAN_NO_DIRECT) and the oracle: rolling and fixed layouts, |0>,
    product and vector inputs, a batch."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p, S = 18, 8, 3
    og = OC.ry_cz_brickwork(p, n)
    params = np.random.default_rng(42).random((S, n * p)) * 2 * np.pi
    plus = OS.product_state(np.array([1, 1]) / np.sqrt(2), n)
    for base in (L.PQS_PLAN_ROLLING, 0):
        outs = {}
        for fl in (base, base | L.PQS_PLAN_NO_DIRECT):
            qc = _build(pyqc, n, T.ry_cz_brickwork_layers(p, n), True, dtype=dtype, tile_bits=12, low_bits=4, flags=fl)
            plan = qc.plan(batch=S)
            assert plan.jitted
            srcs = [x.decode() for x in plan.jit_sweep_sources()]
            direct = sum("= pre[0];" in x for x in srcs), sum("dst[gt ^" in x for x in srcs)
            if fl & L.PQS_PLAN_NO_DIRECT:
                assert direct == (0, 0)
            else:  # (this fixed-layout plan has no sweep that can store directly)
                assert direct[0] > 0 and (direct[1] > 0 or not base), direct
            st = qc.run_batch(params).cpu().numpy()
            qc.set_initial_state(T.plus_state())
            sp = qc.run(list(params[0])).full().ravel()
            outs[fl] = (st, sp)
            for s in range(S):
                assert infidelity(st[s], OS.run(og, params[s], n)) < tol
            assert infidelity(sp, OS.run(og, params[0], n, plus)) < tol
        a, b = outs.values()
        lim = 1e-13 if dtype == "complex128" else 1e-6
        assert np.max(np.abs(a[0] - b[0])) < lim and np.max(np.abs(a[1] - b[1])) < lim


def test_rolling_all_reg_kinds_and_inplace_input(pyqc):
    """All register op kinds under rolling layouts; a vector input state;
    and the C ABI's in-place form (init 0, no source: the state buffer is
    the input) for plans with an odd and an even number of sweeps, from the
    window sweep search and from the plain greedy sweeps."""
    import ctypes as C
    import torch
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    rng = np.random.default_rng(41)
    n = 14
    kinds1 = ["R_x", "R_y", "R_z", "H", "T", "S", "X"]
    kinds2 = ["CNOT", "CZ", "R_zz"]
    parities = set()
    R, G = L.PQS_PLAN_ROLLING, L.PQS_PLAN_GREEDY_SWEEPS
    for depth, flags in ((2, R), (4, R), (7, R), (8, R | G)):
        srng = np.random.default_rng(100 + depth)
        spec = []
        for _ in range(depth):
            for q in range(n):
                spec.append((kinds1[srng.integers(len(kinds1))], q))
            for _ in range(n):
                a, b = (int(x) for x in srng.choice(n, 2, replace=False))
                spec.append((kinds2[srng.integers(len(kinds2))], (a, b)))
        pk = [getattr(pyqc, nm)(list(q) if isinstance(q, tuple) else q, n) for nm, q in spec]
        og = [getattr(OC, nm)(q, n) for nm, q in spec]
        qc = _build(pyqc, n, [pk], True, tile_bits=10, low_bits=3, flags=flags)
        plan = qc.plan()
        assert plan.rolling and plan.jitted
        parities.add(plan.n_sweeps % 2)
        th = rng.random(qc.n_true_params) * 2 * np.pi
        v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        v /= np.linalg.norm(v)
        st = E.run(plan, th, E.Init.from_vector(v))[0].cpu().numpy()
        ref = OS.run(og, th, n, v)
        assert infidelity(st, ref) < 1e-12
        # in place through the C ABI
        buf = torch.as_tensor(v, device="cuda").reshape(1, -1).contiguous()
        prm = torch.as_tensor(th, device="cuda")
        ctx = E.context()
        E.check(L.lib().pqs_run_batch(ctx.handle, plan.handle, C.c_void_p(buf.data_ptr()), 2 ** n, 1,
                                      C.c_void_p(prm.data_ptr()), len(th), None, 0, None, None, E.stream_ptr()))
        got = buf.cpu().numpy().ravel()
        assert np.max(np.abs(got - ref)) < 1e-12
    assert parities == {0, 1}
