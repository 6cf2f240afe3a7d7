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
