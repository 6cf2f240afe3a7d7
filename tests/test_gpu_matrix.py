"""Random circuits with every gate family (dense two-qubit gates, fSim,
shared blocks, CHAIN, fixed and offset gates) across tile shapes, layouts
(fixed / rolling), barriers (wave-local or not) and executors (specialised
or interpreted), each against the oracle; plus edge cases of sizes."""
import numpy as np
import pytest

from oracle import sim as OS
from tests.helpers import build_oracle, build_pkg, infidelity, random_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


CASES = []
for n, tile in ((10, 10), (11, 10), (13, 11), (15, 11), (15, 12)):
    for flags in (0, 8, 24):
        for jit in (True, False):
            CASES.append((n, tile, flags, jit))


@pytest.mark.parametrize("n,tile,flags,jit", CASES)
def test_random_circuits_across_shapes(pyqc, n, tile, flags, jit):
    rng = np.random.default_rng(1000 + 17 * n + tile + flags + int(jit))
    spec = random_spec(n, 3, rng)
    qc = pyqc.PQC(n, tile_bits=tile, low_bits=3, flags=flags, jit=jit)
    qc.add_layer(build_pkg(spec, n, pyqc))
    S = 3
    params = rng.random((S, qc.n_true_params)) * 2 * np.pi
    st = qc.run_batch(params).cpu().numpy()
    og = build_oracle(spec, n)
    for s in range(S):
        ref = OS.run(og, params[s], n)
        assert infidelity(st[s], ref) < 1e-12
        assert np.max(np.abs(st[s] - ref)) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 9])
def test_small_states_and_empty_circuits(pyqc, n):
    """Tiles below the register executor (n < 10), a one-qubit circuit, and
    a circuit without gates (the initial state comes back)."""
    rng = np.random.default_rng(2000 + n)
    if n >= 2:
        spec = random_spec(n, 2, rng, fsim=n >= 2)
        qc = pyqc.PQC(n)
        qc.add_layer(build_pkg(spec, n, pyqc))
        th = rng.random(qc.n_true_params) * 2 * np.pi
        ref = OS.run(build_oracle(spec, n), th, n)
        assert infidelity(qc.run(list(th)).full().ravel(), ref) < 1e-12
    q1 = pyqc.PQC(1)
    q1.add_layer([pyqc.R_y(0, 1), pyqc.R_z(0, 1)])
    th = [0.7, 1.9]
    v = OS.run([__import__("oracle.circuits", fromlist=["R_y"]).R_y(0, 1),
                __import__("oracle.circuits", fromlist=["R_z"]).R_z(0, 1)], np.array(th), 1)
    assert np.max(np.abs(q1.run(th).full().ravel() - v)) < 1e-15
    empty = pyqc.PQC(n)
    out = empty.run([]).full().ravel()
    z = np.zeros(2 ** n)
    z[0] = 1
    assert np.max(np.abs(out - z)) == 0


@pytest.mark.parametrize("n,dtype,tol", [(30, "complex128", 1e-11), (31, "complex64", 2e-4)])
def test_largest_states_round_trip(pyqc, n, dtype, tol):
    """A 16 GiB state (30 qubits c128, 31 qubits c64): a brickwork circuit
    followed by its inverse (the same gates in reverse order, R_y angles
    negated; CZ is its own inverse) returns |0...0>.  Size-independent:
    amplitude 0 is 1 and every other amplitude 0, to rounding.  Covers
    64-bit addressing and the largest tiles' index maps."""
    import torch
    from pyramaterised_amd import engine as E
    print(f"case n={n} dtype={dtype}", flush=True)
    layers = 6
    gates, inv = [], []
    for lay in range(layers):
        for q in range(n):
            gates.append(("R_y", q))
        for q in range(lay % 2, n - 1, 2):
            gates.append(("CZ", q))
    qc = pyqc.PQC(n, dtype=dtype)
    fwd = [pyqc.R_y(q, n) if k == "R_y" else pyqc.CZ([q, q + 1], n) for k, q in gates]
    back = [pyqc.R_y(q, n) if k == "R_y" else pyqc.CZ([q, q + 1], n) for k, q in reversed(gates)]
    qc.add_layer(fwd + back)
    m = sum(1 for k, _ in gates if k == "R_y")
    th = np.random.default_rng(n).random(m) * 2 * np.pi
    params = np.concatenate([th, -th[::-1]])
    plan = qc.plan()
    out = torch.empty((1, 2 ** n), dtype=E.CDTYPE[plan.dtype], device="cuda")
    E.run(plan, torch.as_tensor(params, device="cuda"), qc.init(), out=out)
    a0 = complex(out[0, 0].item())
    assert abs(a0 - 1) < tol, a0
    out[0, 0] = 0
    assert float(out.abs().max().item()) < tol
    del out
    torch.cuda.empty_cache()
