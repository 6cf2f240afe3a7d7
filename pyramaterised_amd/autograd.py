"""torch.autograd entry points of the engine.

``state(qc, theta)`` returns the circuit's final state as a complex tensor
that is differentiable with respect to the float64 parameter tensor
``theta``; ``expectation(qc, theta, H)`` returns <psi|H|psi> the same way
(H: the circuit's Hamiltonian by default, a PauliSum or anything
``PQC.set_H`` accepts).  The reference differentiates by re-simulating the
circuit once per parameter (``PQC.take_derivative`` / ``get_gradients``,
circuit.py:149-192) and forms gradients on the host
(``Measurements.get_gradient_vector``, measure.py:461-471).

``state``: forward runs the engine's one-pass forward Jacobian (all
d psi / d theta_p in batched sweep launches, ``pqs_jacobian``); backward
contracts the incoming gradient with it on the GPU (``pqs_gram``):
dL/dtheta_p = Re <J_p | g> (g = dL/d conj(psi), torch's convention for
complex outputs).

``expectation``: forward evolves psi and lambda = H^dag psi; backward is the
adjoint pass of ``adjoint.py`` (dE/dtheta_p = 2 Re <lambda_g|D_p|psi_g> with
both states carried back through small plans of inverse ops), times dL/dE.
Circuits outside the adjoint pass (fSim gates, dense gates, non-generator
commute=False blocks) use the Jacobian: 2 Re <H psi | J_p> * dL/dE.
"""
import numpy as np
import torch

from . import engine as E


def _jacobian(qc, theta):
    th = theta.detach().to(device=E.device(), dtype=torch.float64).reshape(-1)
    if th.numel() != qc.n_true_params:
        raise ValueError(f"expected {qc.n_true_params} parameters, got {th.numel()}")
    plan = qc.plan(with_generators=True, batch=qc.n_true_params)
    return E.jacobian(plan, th, qc.init())


def vjp(J, g):
    """Re <J_p | g> for every row J_p of J (P, 2^N) and a cotangent g (2^N,)."""
    if J.shape[0] == 0:
        return torch.zeros(0, dtype=torch.float64, device=J.device)
    return E.gram(J, g.reshape(1, -1).to(J.dtype)).real.reshape(-1)


class _State(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, qc):
        psi, J = _jacobian(qc, theta)
        ctx.save_for_backward(J)
        ctx.theta_dtype = theta.dtype
        ctx.theta_device = theta.device
        return psi

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors
        grad = vjp(J, g.contiguous())
        return grad.to(dtype=ctx.theta_dtype, device=ctx.theta_device), None


class _Expectation(torch.autograd.Function):
    """<psi|H|psi>.  Where the circuit allows it (Pauli-sum generators, no
    dense gate), forward evolves psi only and backward is the adjoint pass
    (adjoint.py): lambda = H psi carried back with psi through small plans of
    inverse ops, 2 Re <lambda|D_p|psi> per parameter -- two states instead
    of P.  Otherwise forward keeps the Jacobian for backward."""

    @staticmethod
    def forward(ctx, theta, qc, H):
        from . import adjoint
        from .qobj import Ket, PauliSum
        th = theta.detach().to(device=E.device(), dtype=torch.float64).reshape(-1)
        if th.numel() != qc.n_true_params:
            raise ValueError(f"expected {qc.n_true_params} parameters, got {th.numel()}")
        ctx.theta_dtype = theta.dtype
        ctx.theta_device = theta.device
        ctx.qc = qc
        try:
            prog = adjoint.program(qc)
            Hps = H if isinstance(H, PauliSum) else PauliSum.from_dense(H)
            psi = E.run(qc.plan(), th, qc.init())[0]
            lam = adjoint.apply_hamiltonian(Hps, psi, qc.n_qubits, adjoint=True)
            ctx.adjoint = prog
            ctx.save_for_backward(psi, lam, th)
            e = E.inner(psi, lam)
        except NotImplementedError:
            psi, J = _jacobian(qc, theta)
            Hpsi = (H * Ket(psi, qc.n_qubits)).t.reshape(-1)
            ctx.adjoint = None
            ctx.save_for_backward(J, Hpsi)
            e = E.inner(psi, Hpsi)
        return torch.tensor(float(np.real(complex(e))), dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, gE):
        if ctx.adjoint is not None:
            psi, lam, th = ctx.saved_tensors
            g = torch.as_tensor(ctx.adjoint.gradient(psi, lam, th), dtype=torch.float64, device=psi.device)
            grad = g * gE.to(psi.device)
        else:
            J, Hpsi = ctx.saved_tensors
            grad = 2 * vjp(J, Hpsi) * gE.to(J.device)
        return grad.to(dtype=ctx.theta_dtype, device=ctx.theta_device), None, None


def state(qc, theta):
    """Final state of ``qc`` at parameters ``theta`` (1-D float tensor),
    differentiable with respect to ``theta``."""
    return _State.apply(theta, qc)


def expectation(qc, theta, H=None):
    """<psi(theta)|H|psi(theta)> as a differentiable float64 scalar."""
    if H is None:
        H = qc.hamiltonian()
    return _Expectation.apply(theta, qc, H)
