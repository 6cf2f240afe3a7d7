"""torch.autograd through the engine: gradients against central finite
differences of the engine's own states and against the oracle's derivative
states."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import sim as OS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pyqc():
    import pyramaterised_amd
    return pyramaterised_amd


def _qc(pyqc, n=6, p=3):
    from pyramaterised_amd import templates as T
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    return qc


def test_state_gradient_matches_oracle_vjp(pyqc):
    from pyramaterised_amd import autograd as AG
    n, p = 6, 3
    qc = _qc(pyqc, n, p)
    th0 = np.random.default_rng(31).random(qc.n_true_params) * 2 * np.pi
    th = torch.tensor(th0, requires_grad=True)
    psi = AG.state(qc, th)
    w = torch.as_tensor(np.random.default_rng(32).normal(size=2 ** n), device=psi.device)
    c = torch.as_tensor(np.random.default_rng(33).normal(size=2 ** n) + 0j, device=psi.device)
    L = (w * psi.abs() ** 2).sum() + (c * psi).real.sum()
    L.backward()
    # oracle: dL/dtheta_p = sum_i 2 w_i Re(conj(psi_i) dpsi_i) + Re(c_i dpsi_i)
    og = OC.clifford_he(p, n)
    ref = OS.run(og, th0, n)
    D = np.array(OS.derivative_states(og, th0, n))
    wn, cn = w.cpu().numpy(), c.cpu().numpy()
    g_ref = np.array([np.sum(2 * wn * np.real(np.conj(ref) * d)) + np.sum(np.real(cn * d)) for d in D])
    assert np.max(np.abs(th.grad.numpy() - g_ref)) < 1e-11


def test_expectation_gradient_and_value(pyqc):
    from pyramaterised_amd import autograd as AG
    qc = _qc(pyqc)
    th0 = np.random.default_rng(34).random(qc.n_true_params) * 2 * np.pi
    th = torch.tensor(th0, requires_grad=True)
    e = AG.expectation(qc, th)
    e.backward()
    assert abs(e.item() - qc.cost(list(th0))) < 1e-13
    gv = np.array(pyqc.measure.Measurements(qc).get_gradient_vector(list(th0)))
    assert np.max(np.abs(th.grad.numpy() - gv)) < 1e-12
    h = 1e-6
    for k in (0, 5, qc.n_true_params - 1):
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        fd = (qc.cost(list(tp)) - qc.cost(list(tm))) / (2 * h)
        assert abs(fd - th.grad[k].item()) < 1e-7
