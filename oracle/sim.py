"""FP64 numpy statevector simulator restating the reference's sequential evolution.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

A circuit is a list of ``OGate`` objects (one per reference ``Gate`` in the
flattened ``PQC.gates`` list, circuit.py:53-61).  Each gate yields, for its bound
angles, an ordered list of primitives ``(qubits, local_matrix)`` -- the
reference applies the gate's full tensored operator, and a product of local
operators on the listed qubits is the same linear map -- and a derivative
``D`` given as a sum of terms ``(coef, [(qubits, matrix), ...])`` such that the
reference's derivative state is ``U_G..U_{g+1} (D U_g) U_{g-1}..U_1 |psi0>``
(circuit.py:149-172).

State layout: amplitude index i = sum_k b_k 2^(N-1-k), i.e. qubit 0 is the most
significant bit, which is QuTiP's ``tensor`` order (gates.py:39-46,
circuit.py:22).  ``apply`` realises that by viewing the state as an N-axis
tensor whose axis k is qubit k.

Pins: ``dense_unitary`` builds full operators by explicit bit arithmetic (and
np.kron for one-qubit factors, the genFockOp construction of gates.py:39-42);
tests check ``run`` against it at N <= 5 and against the reference's qg energy.
``derivative_states_compat`` is checked against ``derivative_states`` (the same
states where every gate before each fSim is parameterised, the looked-up
gate's state or the plain state elsewhere) and against central differences.
"""
import numpy as np


class OGate:
    """One reference gate.

    prims(theta, phi) -> list[(qubits, matrix)]    applied in list order
    deriv(theta, phi, which) -> list[(coef, list[(qubits, matrix)])]
    n_params: 0, 1 or 2 (``param_count`` of gates.py).
    """

    def __init__(self, name, n_params, prims, deriv=None):
        self.name = name
        self.n_params = n_params
        self._prims = prims
        self._deriv = deriv

    def prims(self, theta=0.0, phi=0.0):
        return self._prims(theta, phi)

    def deriv(self, theta=0.0, phi=0.0, which=1):
        return self._deriv(theta, phi, which)

    def __repr__(self):
        return f"OGate({self.name})"


def zero_state(n):
    """|0...0> (circuit.py:22)."""
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def product_state(single, n):
    """tensor([s]*n) (circuit.py:35-36)."""
    out = np.array([1.0 + 0j])
    for _ in range(n):
        out = np.kron(out, np.asarray(single, dtype=np.complex128).ravel())
    return out


def apply(psi, qubits, m, n):
    """Apply a local 2^k x 2^k matrix on ``qubits`` (first listed = most significant
    factor of the local matrix) to an n-qubit state."""
    k = len(qubits)
    t = psi.reshape((2,) * n)
    mt = np.asarray(m, dtype=np.complex128).reshape((2,) * (2 * k))
    out = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), list(qubits)))
    # tensordot puts the k output axes first; move them back to ``qubits``.
    out = np.moveaxis(out, list(range(k)), list(qubits))
    return np.ascontiguousarray(out).reshape(-1)


def apply_prims(psi, prims, n):
    for qs, m in prims:
        psi = apply(psi, qs, m, n)
    return psi


def apply_deriv(psi, terms, n):
    """D psi for D = sum_t coef_t * (product of the term's local operators)."""
    out = np.zeros_like(psi)
    for coef, prims in terms:
        out += coef * apply_prims(psi, prims, n)
    return out


def bind(gates, params):
    """Per-gate (theta, phi) from a flat parameter vector, in the order of
    PQC.set_params (circuit.py:86-116): two values for a 2-parameter gate,
    one for a 1-parameter gate, none otherwise."""
    out, k = [], 0
    for g in gates:
        if g.n_params == 2:
            out.append((params[k], params[k + 1]))
            k += 2
        elif g.n_params == 1:
            out.append((params[k], 0.0))
            k += 1
        else:
            out.append(None)
    if k != len(params):
        raise ValueError(f"circuit has {k} parameters, got {len(params)}")
    return out


def n_params(gates):
    return sum(g.n_params for g in gates)


def run(gates, params, n, init=None):
    """PQC.run (circuit.py:118-125): |psi> = U_G ... U_1 |psi0>."""
    psi = zero_state(n) if init is None else np.asarray(init, dtype=np.complex128).copy()
    for g, ang in zip(gates, bind(gates, params)):
        th, ph = ang if ang is not None else (0.0, 0.0)
        psi = apply_prims(psi, g.prims(th, ph), n)
    return psi


def derivative_states(gates, params, n, init=None):
    """PQC.get_gradients (circuit.py:174-192) by the reference's algorithm: for
    every parameterised gate, re-simulate the whole circuit with that gate
    replaced by D*gate (circuit.py:149-172).  A 2-parameter gate yields two
    states, in (theta, phi) order, from parameterised_derivative(1) and (2).
    (The reference looks the second gate up by the wrong index,
    circuit.py:188; this restatement uses the same gate for both.)"""
    bound = bind(gates, params)
    psi0 = zero_state(n) if init is None else np.asarray(init, dtype=np.complex128)
    out = []
    for gi, g in enumerate(gates):
        if g.n_params == 0:
            continue
        for which in range(1, g.n_params + 1):
            psi = psi0.copy()
            for gj, (h, ang) in enumerate(zip(gates, bound)):
                th, ph = ang if ang is not None else (0.0, 0.0)
                psi = apply_prims(psi, h.prims(th, ph), n)
                if gj == gi:
                    psi = apply_deriv(psi, h.deriv(th, ph, which), n)
            out.append(psi)
    return out


def derivative_states_compat(gates, params, n, init=None):
    """PQC.get_gradients as written (circuit.py:174-192), its index lookup
    included: for the two-parameter gate at index ``count`` of the
    parameterised gates, the second state is take_derivative(self.gates[count],
    param=2) (circuit.py:188-189): the gate at position ``count`` of the
    circuit with parameterised_derivative(2) -- a two-parameter gate's phi
    derivative, a one-parameter gate's derivative (Gate.parameterised_derivative
    falls back to derivative(), gates.py:96-97), a fixed gate's identity
    (Gate.derivative, gates.py:93-94), which leaves the plain state.
    take_derivative puts a copy with the same angles back in the circuit
    (circuit.py:158,171), so the states do not depend on the copies."""
    bound = bind(gates, params)
    psi0 = zero_state(n) if init is None else np.asarray(init, dtype=np.complex128)

    def state_with(gi, which):
        psi = psi0.copy()
        for gj, (h, ang) in enumerate(zip(gates, bound)):
            th, ph = ang if ang is not None else (0.0, 0.0)
            psi = apply_prims(psi, h.prims(th, ph), n)
            if gj == gi and which > 0:
                psi = apply_deriv(psi, h.deriv(th, ph, which), n)
        return psi

    out = []
    for count, gi in enumerate([i for i, g in enumerate(gates) if g.n_params > 0]):
        out.append(state_with(gi, 1))
        if gates[gi].n_params == 2:
            x = gates[count]
            out.append(state_with(count, min(x.n_params, 2)))
    return out


# ---------------------------------------------------------------- brute force

def embed(m, qubits, n):
    """Full 2^n operator of a local matrix by explicit bit arithmetic:
    U[j, i] = m[r, c] where c = bits of i on ``qubits`` (first listed = most
    significant) and j = i with those bits replaced by r."""
    k = len(qubits)
    dim = 2 ** n
    full = np.zeros((dim, dim), dtype=np.complex128)
    pos = [n - 1 - q for q in qubits]
    for i in range(dim):
        c = 0
        for p in pos:
            c = (c << 1) | ((i >> p) & 1)
        base = i
        for p in pos:
            base &= ~(1 << p)
        for r in range(2 ** k):
            j = base
            for idx, p in enumerate(pos):
                if (r >> (k - 1 - idx)) & 1:
                    j |= 1 << p
            full[j, i] += m[r, c]
    return full


def kron_1q(m, q, n):
    """genFockOp (gates.py:39-42): tensor of identities with ``m`` at position q."""
    out = np.array([[1.0 + 0j]])
    for k in range(n):
        out = np.kron(out, m if k == q else np.eye(2))
    return out


def dense_unitary(gates, params, n):
    """Product of full operators, U = U_G ... U_1, built without ``apply``."""
    u = np.eye(2 ** n, dtype=np.complex128)
    for g, ang in zip(gates, bind(gates, params)):
        th, ph = ang if ang is not None else (0.0, 0.0)
        for qs, m in g.prims(th, ph):
            full = kron_1q(m, qs[0], n) if len(qs) == 1 else embed(m, qs, n)
            u = full @ u
    return u
