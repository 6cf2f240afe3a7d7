"""Random circuits over every gate family, through random tile shapes and
planner switches (rolling layouts, direct end loads / stores or LDS
staging, wave-local runs or barriers, lane order), in both precisions,
against the oracle.  Each case prints its parameters first."""
import numpy as np
import pytest

from oracle import sim as OS
from tests.helpers import build_oracle, build_pkg, infidelity, random_spec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", range(14))
def test_random_plan_matches_oracle(seed):
    import pyramaterised_amd as pyqc
    from pyramaterised_amd import _lib as L
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(10, 18))
    depth = int(rng.integers(2, 6))
    dtype = "complex128" if (seed // 2) % 2 == 0 else "complex64"
    kmax = 13 if dtype == "complex128" else 14
    tile = int(rng.integers(10, min(n, kmax) + 1))
    low = int(rng.integers(2, 5))
    flags = [0, L.PQS_PLAN_ROLLING, L.PQS_PLAN_ROLLING | L.PQS_PLAN_NO_DIRECT, L.PQS_PLAN_NO_WAVE_LOCAL,
             L.PQS_PLAN_ROLLING | L.PQS_PLAN_LANES_ASCENDING, L.PQS_PLAN_NO_DIRECT,
             L.PQS_PLAN_ROLLING | L.PQS_PLAN_GREEDY_SWEEPS | L.PQS_PLAN_GREEDY_PHASES][seed % 7]
    if seed % 2 == 0:  # only gates the specialised kernels take: every sweep is specialised
        depth = int(rng.integers(4, 9))
        spec = []
        for _ in range(depth):
            spec += [(("R_x", "R_y", "R_z", "H", "T", "S", "X")[rng.integers(7)], q, None) for q in range(n)]
            for _ in range(n):
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                spec.append((("CNOT", "CZ", "R_zz")[rng.integers(3)], (a, b), None))
    else:  # every gate family, dense two-qubit gates included
        spec = random_spec(n, depth, rng, fsim=(seed % 3 == 0))
    print(f"case seed={seed} n={n} depth={depth} dtype={dtype} tile={tile} low={low} flags={flags}", flush=True)
    qc = pyqc.PQC(n, jit=True, dtype=dtype, tile_bits=tile, low_bits=low, flags=flags)
    qc.add_layer(build_pkg(spec, n, pyqc))
    og = build_oracle(spec, n)
    S = 2
    assert qc.plan(batch=S).jitted or seed % 2 == 1  # (odd seeds may have no specialisable sweep)
    params = rng.random((S, qc.n_true_params)) * 2 * np.pi
    st = qc.run_batch(params).cpu().numpy()
    tol = 1e-12 if dtype == "complex128" else 1e-5
    for s in range(S):
        assert infidelity(st[s], OS.run(og, params[s], n)) < tol, (seed, s)
    one = qc.run(list(params[1])).full().ravel()
    assert np.max(np.abs(one - st[1])) < (1e-13 if dtype == "complex128" else 1e-6)
