"""A block cap on a context (pqs_ctx_set_max_blocks) splits batched sweeps
over instance ranges: the states, Jacobians and QFIMs are bit-identical to
the uncapped launches, for interpreted and specialised plans, down to one
instance per launch, and they match the oracle."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS
from tests.helpers import infidelity

pytestmark = pytest.mark.gpu


def _hea(n, p, **kw):
    import pyramaterised_amd as pyqc
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n, **kw)
    for layer in T.clifford_HE_layers(p, n):
        qc.add_layer(layer)
    return qc


def test_capped_launches_equal_uncapped():
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import engine as E
    n, p, B = 12, 4, 6
    rng = np.random.default_rng(17)
    dev = torch.cuda.current_device()
    free, capped, single = E.Context(dev), E.Context(dev), E.Context(dev)
    capped.set_max_blocks(40)  # a few instances per launch
    single.set_max_blocks(1)   # one instance per launch
    s = torch.cuda.Stream()
    og = OC.clifford_he(p, n)
    for jit in (False, True):
        qc = _hea(n, p, jit=jit)
        th = rng.random(qc.n_true_params) * 2 * np.pi
        rows = rng.random((B, qc.n_true_params)) * 2 * np.pi
        gplan, plan, init = qc.plan(with_generators=True), qc.plan(), qc.init()
        got = []
        for ctx in (free, capped, single):
            with E.using(ctx, s):
                psi, J = E.jacobian(gplan, th, init)
                F = E.qfim(J, psi)
                st = E.run(plan, rows, init)
            s.synchronize()
            got.append((psi.cpu(), J.cpu(), F.cpu(), st.cpu()))
        for other in got[1:]:
            for a, b in zip(got[0], other):
                assert torch.equal(a, b), jit
        assert infidelity(got[0][0].numpy().ravel(), OS.run(og, th, n)) < 1e-12
        for i in range(B):
            assert infidelity(got[0][3][i].numpy(), OS.run(og, rows[i], n)) < 1e-12
    with pytest.raises(L.EngineError):
        free.set_max_blocks(-1)
