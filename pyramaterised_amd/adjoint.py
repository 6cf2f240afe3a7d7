"""Reverse-mode (adjoint) energy gradients: dE/dtheta_p = 2 Re <psi|H|d_p psi>
for every parameter with two states and one backward pass.

Reference: ``Measurements.get_gradient_vector`` (measure.py:461-471) forms
H|d_p psi> for every derivative state, and the derivative states come from
one full re-simulation per parameter (``PQC.get_gradients``,
circuit.py:149-192).  Here, with psi = U_G..U_1 psi0 and lambda = H^dag psi,

    dE/dtheta_p = 2 Re <lambda_g | D_p | psi_g>,

where g is the gate of parameter p, D_p its generator (d U_g = D_p U_g) and
lambda_g, psi_g are lambda and psi carried back to the point right after
gate g (U_{g+1}^dag .. U_G^dag applied to both).  The backward pass walks
the gates in reverse: parameters whose generators commute with every gate
between them are measured at one point (a whole rotation layer of a
hardware-efficient ansatz at once), and between measurement points the pair
(lambda, psi) is evolved by a small plan of inverse ops (one batched run of
two instances).  Each measurement is one ``pqs_pauli_transition`` launch
over the group's Pauli terms.  Memory: the two states plus the workspace of
one small plan, against P states for the Jacobian route.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import engine as E


def _masks(o):
    """(active, diagonal) qubit masks of a primitive op (qubit q = bit q):
    qubits it acts on non-diagonally / only diagonally."""
    k, b0 = o.kind, 1 << o.q0
    b1 = (1 << o.q1) if o.q1 >= 0 else 0
    if k == L.OP_MAT1:
        diag = o.m[2] == 0 and o.m[3] == 0 and o.m[4] == 0 and o.m[5] == 0
        return (0, b0) if diag else (b0, 0)
    if k in (L.OP_RX, L.OP_RY):
        return b0, 0
    if k == L.OP_RZ:
        return 0, b0
    if k == L.OP_CX:
        return b1, b0
    if k in (L.OP_CZ, L.OP_RZZ):
        return 0, b0 | b1
    if k == L.OP_MAT2:
        m = np.array(o.m[:32]).view(np.complex128).reshape(4, 4)
        return (0, b0 | b1) if np.count_nonzero(m - np.diag(np.diag(m))) == 0 else (b0 | b1, 0)
    return b0 | b1, 0


def inverse_op(o):
    """The op of the primitive's inverse: rotation angles negated (R(a)^dag =
    R(-a)), fixed matrices conjugate-transposed; CNOT and CZ are their own
    inverses.  fSim with a phi parameter has no op form of its inverse, and a
    fixed matrix that is not unitary (sqrtH, the element-wise square root of
    H, gates.py:235-242) none the pass can share between lambda (carried
    back by U^dag) and psi (by U^-1)."""
    o2 = L.Op.from_buffer_copy(o)
    if o.kind in (L.OP_RX, L.OP_RY, L.OP_RZ, L.OP_RXX, L.OP_RYY, L.OP_RZZ):
        o2.scale, o2.offset = -o.scale, -o.offset
    elif o.kind == L.OP_FSIM and o.slot2 < 0:
        o2.scale, o2.offset = -o.scale, -o.offset  # fixed_fSim: fSim(a, 0)^dag = fSim(-a, 0)
    elif o.kind in (L.OP_MAT1, L.OP_MAT2):
        d = 2 if o.kind == L.OP_MAT1 else 4
        m = np.array(o.m[:2 * d * d]).view(np.complex128).reshape(d, d)
        if not np.allclose(m.conj().T @ m, np.eye(d), atol=1e-13, rtol=0):
            raise NotImplementedError("a non-unitary fixed gate takes the Jacobian route")
        md = np.ascontiguousarray(m.conj().T).reshape(-1).view(np.float64)
        for i in range(2 * d * d):
            o2.m[i] = float(md[i])
    elif o.kind not in (L.OP_CX, L.OP_CZ):
        raise NotImplementedError(f"no inverse op for kind {o.kind}")
    return o2


def generator_terms(gen_ops):
    """[(coef, xmask, zmask)] of a generator given as PQS_OP_GEN records
    (Pauli code on q0, and q1 for a two-qubit string)."""
    out = []
    for o in gen_ops:
        if o.kind != L.OP_GEN:
            raise NotImplementedError("the adjoint pass takes Pauli-sum generators only")
        x = z = 0
        for q in (o.q0, o.q1):
            if q < 0:
                continue
            if o.pauli in (1, 2):
                x |= 1 << q
            if o.pauli in (2, 3):
                z |= 1 << q
        out.append((complex(o.m[0], o.m[1]), x, z))
    return out


def _commutes(terms, ops):
    """Every term commutes with every op: its X/Y qubits avoid the op's
    qubits and its Z qubits avoid the op's non-diagonal ones."""
    for o in ops:
        act, dg = _masks(o)
        for _, x, z in terms:
            if x & (act | dg) or (z & ~x) & act:
                return False
    return True


class AdjointProgram:
    """The backward pass of a circuit: steps of (plan of inverse ops or None,
    [(slot, terms)] measured after it)."""

    def __init__(self, qc, dtype=None):
        dt = qc.dtype if dtype is None else E.dtype_code(dtype)
        if qc.has_dense() or qc.special_slots():
            raise NotImplementedError("dense gates and non-generator commute=False blocks take the Jacobian route")
        steps = []
        pending, since, group = [], [], []
        for g in reversed(qc.gates):
            slot = qc._gate_slot.get(id(g), -1)
            prims = g.prim_ops(max(slot, 0))
            if g.param_count > 0:
                if g.param_count != 1:
                    raise NotImplementedError("two-parameter gates (fSim) take the Jacobian route")
                terms = generator_terms(g.gen_ops(slot, slot))
                if group and _commutes(terms, since):
                    group.append((slot, terms))
                else:
                    if group:
                        steps.append((pending, group))
                        pending = []
                    pending = pending + since
                    since = []
                    group = [(slot, terms)]
            since += [inverse_op(o) for o in reversed(prims)]
        if group:
            steps.append((pending, group))
        self.n = qc.n_qubits
        self.dtype = dt
        self.P = qc.n_true_params
        self.step_ops = steps  # (inverse ops, group) per step, before planning
        self.steps = []
        for ops, grp in steps:
            plan = E.Plan(self.n, ops, dtype=dt, tile_bits=min(12, self.n), low_bits=3) if ops else None
            xs, zs, cs, owner = [], [], [], []
            for slot, terms in grp:
                for c, x, z in terms:
                    xs.append(x)
                    zs.append(z)
                    cs += [c.real, c.imag]
                    owner.append(slot)
            nt = len(xs)
            self.steps.append((plan, nt, (C.c_uint64 * max(1, nt))(*xs), (C.c_uint64 * max(1, nt))(*zs),
                               (C.c_double * max(2, 2 * nt))(*cs), np.array(owner, dtype=np.int64)))
        self.n_terms = sum(s[1] for s in self.steps)

    def gradient(self, psi, lam, params):
        """Energy gradient 2 Re <psi|H|d_p psi> from psi = U psi0 and lambda =
        H^dag psi (both (2^N,) tensors, left unchanged).  Returns a float64
        numpy array of P values."""
        ctx = E.context()
        pair = torch.stack([lam.reshape(-1), psi.reshape(-1)]).contiguous()
        out = torch.empty((max(1, self.n_terms), 2), dtype=torch.float64, device=pair.device)
        th = E._params_tensor(params, pair.device)
        off = 0
        for plan, nt, xs, zs, cs, _ in self.steps:
            if plan is not None:
                E.run_rows(plan, th, pair)
            E.check(L.lib().pqs_pauli_transition(ctx.handle, self.dtype, E.ptr(pair[0]), E.ptr(pair[1]), self.n,
                                                 xs, zs, cs, nt, E.ptr(out[off:off + nt]), E.stream_ptr()))
            off += nt
        vals = out[:off].cpu().numpy()
        grad = np.zeros(self.P)
        owner = np.concatenate([s[5] for s in self.steps]) if self.steps else np.zeros(0, dtype=np.int64)
        np.add.at(grad, owner, 2 * vals[:, 0])
        return grad


def apply_hamiltonian(H, psi, n, adjoint=False):
    """H psi (or H^dag psi) for a PauliSum H on one state tensor."""
    ctx = E.context()
    terms = H.terms()
    nt = len(terms)
    xm = (C.c_uint64 * max(1, nt))(*[int(t[1]) for t in terms])
    zm = (C.c_uint64 * max(1, nt))(*[int(t[2]) for t in terms])
    cs = [complex(t[0]).conjugate() if adjoint else complex(t[0]) for t in terms]
    cf = (C.c_double * max(2, 2 * nt))(*[v for c in cs for v in (c.real, c.imag)])
    x = psi.reshape(-1).contiguous()
    out = torch.empty_like(x)
    E.check(L.lib().pqs_apply_pauli(ctx.handle, E.dtype_code(x.dtype), E.ptr(x), n, xm, zm, cf, nt, E.ptr(out),
                                    E.stream_ptr()))
    return out


def program(qc):
    """The circuit's AdjointProgram, cached on the circuit (rebuilt when its
    gates change)."""
    key = ("adjoint", qc.dtype)
    p = qc._plans.get(key)
    if p is None:
        p = qc._plans[key] = AdjointProgram(qc)
    return p


def energy_and_gradient(qc, params, H=None):
    """(<psi|H|psi>, dE/dtheta) at ``params`` by the adjoint pass (H: the
    circuit's Hamiltonian as a PauliSum by default)."""
    from .qobj import PauliSum
    prog = program(qc)
    if H is None:
        H = qc.hamiltonian()
    elif not isinstance(H, PauliSum):
        H = PauliSum.from_dense(H)
    psi = E.run(qc.plan(), params, qc.init())[0]
    # lambda = H^dag psi, so that <lambda|d psi> = <psi|H|d psi> for any H
    lam = apply_hamiltonian(H, psi, qc.n_qubits, adjoint=True)
    energy = complex(E.inner(psi, lam).item()).real
    return energy, prog.gradient(psi, lam, params)
