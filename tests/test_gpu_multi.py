"""Multi-plan runs (pqs_run_multi): instances of circuits of uneven width
and depth in one launch per sweep round and tile size; every state equals
the same job run alone and the oracle (complex128, 1e-12)."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS
from tests.helpers import infidelity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _job(pyqc, n, depth, tile_bits=0, low_bits=0, flags=0, dtype="complex128"):
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n, dtype=dtype, tile_bits=tile_bits, low_bits=low_bits, flags=flags, jit=False)
    for l in T.clifford_HE_layers(depth, n):
        qc.add_layer(l)
    return qc


def test_mixed_widths_and_depths_match_alone_and_oracle(pyqc):
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    R = L.PQS_PLAN_ROLLING
    # (n, depth, batch, plan options): widths 6-9 share the per-op group;
    # 12, 14, 16 (12-bit tiles, one rolling) share a register-phase group
    spec = [(6, 3, 2, {}), (7, 9, 1, {}), (9, 2, 3, {}), (10, 4, 1, {}), (12, 5, 2, dict(tile_bits=12, low_bits=3)),
            (14, 3, 1, dict(tile_bits=12, low_bits=3, flags=R)), (16, 7, 2, dict(tile_bits=12, low_bits=4)),
            (16, 2, 1, dict(tile_bits=12, low_bits=3, flags=R))]
    rng = np.random.default_rng(4)
    jobs, circuits, params = [], [], []
    for n, d, B, opts in spec:
        qc = _job(pyqc, n, d, **opts)
        prm = rng.random((B, qc.n_true_params)) * 2 * np.pi
        init = qc.init()
        if n == 9:  # a product start state
            qc.set_initial_state((pyqc.basis(2, 0) + 1j * pyqc.basis(2, 1)) / np.sqrt(2))
            init = qc.init()
        jobs.append((qc.plan(), prm, init))
        circuits.append(qc)
        params.append(prm)
    groups = {p.tile_bits if p.tile_bits >= 10 else 0 for p, _, _ in jobs}
    assert len(groups) < len(jobs)  # several plans share launches
    outs = E.run_multi(jobs)
    for (n, d, B, _), qc, (plan, prm, init), out in zip(spec, circuits, jobs, outs):
        alone = E.run(plan, prm, init)
        assert torch.equal(out, alone), (n, d)
        og = OC.clifford_he(d, n)
        start = None
        if n == 9:
            start = OS.product_state(np.array([1, 1j]) / np.sqrt(2), n)
        for b in range(B):
            assert infidelity(out[b].cpu().numpy(), OS.run(og, prm[b], n, init=start)) < 1e-12, (n, d, b)


def test_multi_complex64_and_source_states(pyqc):
    """complex64 plans, and instances that start from a given state (init 0)."""
    from pyramaterised_amd import engine as E
    rng = np.random.default_rng(5)
    jobs, refs = [], []
    for n, d in ((8, 3), (11, 2), (11, 5)):
        qc = _job(pyqc, n, d, dtype="complex64")
        v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        v /= np.linalg.norm(v)
        prm = rng.random((2, qc.n_true_params)) * 2 * np.pi
        init = E.Init.from_vector(torch.as_tensor(v))
        jobs.append((qc.plan(), prm, init))
        refs.append([OS.run(OC.clifford_he(d, n), prm[b], n, init=v) for b in range(2)])
    outs = E.run_multi(jobs)
    for (plan, prm, init), out, ref in zip(jobs, outs, refs):
        assert torch.equal(out, E.run(plan, prm, init))
        for b in range(2):
            assert infidelity(out[b].cpu().numpy(), ref[b]) < 1e-5
