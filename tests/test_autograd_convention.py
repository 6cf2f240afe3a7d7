"""The backward formula of pyramaterised_amd.autograd (CPU): for a real loss
L of a complex state psi(theta), torch's gradient with respect to a real
theta_p equals Re <J_p | g>, where J_p = d psi / d theta_p and g is the
gradient torch hands to psi's backward.  Checked on a 2-qubit circuit written
in plain torch ops, with J from the closed-form derivative."""
import torch


def _ry(t):
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    return torch.stack([torch.stack([c, -s]), torch.stack([s, c])]).to(torch.complex128)


def _rz(t):
    e = torch.exp(-0.5j * t.to(torch.complex128))
    z = torch.zeros((), dtype=torch.complex128)
    return torch.stack([torch.stack([e, z]), torch.stack([z, torch.conj(e)])])


def _psi(th):
    a = torch.kron(_ry(th[0]), _rz(th[1]))
    cx = torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=torch.complex128)
    b = torch.kron(_rz(th[2]), _ry(th[3]))
    v = torch.zeros(4, dtype=torch.complex128)
    v[0] = 1
    return b @ (cx @ (a @ v))


def _loss(psi):
    w = torch.tensor([0.3, -1.2, 0.7, 2.0], dtype=torch.float64)
    return (w * psi.abs() ** 2).sum() + (torch.tensor(0.4 - 0.9j) * psi[1]).real + psi[2].imag ** 2


def test_vjp_convention_matches_torch():
    torch.manual_seed(0)
    th = (torch.rand(4, dtype=torch.float64) * 6.28).requires_grad_(True)
    L = _loss(_psi(th))
    (ref,) = torch.autograd.grad(L, th)
    # g: torch's incoming gradient at psi
    psi = _psi(th.detach()).requires_grad_(True)
    (g,) = torch.autograd.grad(_loss(psi), psi)
    # J rows: d psi / d theta_p by central differences of the closed form
    h = 1e-6
    J = []
    for p in range(4):
        tp, tm = th.detach().clone(), th.detach().clone()
        tp[p] += h
        tm[p] -= h
        J.append((_psi(tp) - _psi(tm)) / (2 * h))
    J = torch.stack(J)
    got = (J.conj() @ g).real  # Re <J_p | g>
    assert torch.max(torch.abs(got - ref)) < 1e-8
