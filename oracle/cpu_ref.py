"""CPU baseline that restates the reference's *algorithm* (not its numbers).

TEST / BASELINE INFRASTRUCTURE ONLY (see ``oracle/__init__.py``); used by the
``cpu_baseline`` leg of bench.py.

The reference (pyramaterised 1.0.2 on QuTiP 4.7) keeps, for every gate, a full
2^N x 2^N sparse operator built by tensoring the local matrix with identities
(gates.py:39-46, 119-123), rebuilds it on every ``set_theta``
(gates.py:129-131), multiplies CHAIN entanglers into one operator once at
construction (gates.py:366-376), and runs a circuit as one sparse
matrix x ket product per gate (circuit.py:123-124).  QuTiP is not installed,
so this module does the same with scipy.sparse CSR matrices, single-threaded.
"""
import numpy as np
import scipy.sparse as sp

from . import gates as G


def full_1q(m, q, n):
    """I_{2^q} (x) m (x) I_{2^(n-q-1)} as CSR (genFockOp / qutip gate expansion)."""
    return sp.kron(sp.kron(sp.identity(2 ** q, format="csr"), sp.csr_matrix(m)),
                   sp.identity(2 ** (n - q - 1), format="csr"), format="csr")


def full_2q(m, a, b, n):
    """Full operator of a 4x4 local matrix on (a, b) by index arithmetic."""
    dim = 2 ** n
    idx = np.arange(dim)
    pa, pb = n - 1 - a, n - 1 - b
    ba, bb = (idx >> pa) & 1, (idx >> pb) & 1
    col_local = 2 * ba + bb
    rows, cols, vals = [], [], []
    base = idx & ~((1 << pa) | (1 << pb))
    for r in range(4):
        j = base | ((r >> 1) << pa) | ((r & 1) << pb)
        v = m[r, col_local]
        nz = v != 0
        rows.append(j[nz])
        cols.append(idx[nz])
        vals.append(v[nz])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(dim, dim))


class RefGate:
    """A gate holding its full operator, rebuilt on every angle change."""

    def __init__(self, n, kind, qubits, fixed=None, parts=None):
        self.n, self.kind, self.qubits = n, kind, qubits
        self.param = kind in ("rx", "ry", "rz")
        if fixed is not None:
            self.op = fixed
        elif parts is not None:  # CHAIN: product of entanglers, built once
            op = sp.identity(2 ** n, format="csr", dtype=np.complex128)
            for p in parts:
                op = p @ op
            self.op = op.tocsr()
        else:
            self.op = None

    def set_theta(self, t):
        self.op = full_1q(G.ROT[self.kind[1]](t), self.qubits[0], self.n)


def clifford_he_gates(p, n):
    """clifford_HE_layers(p, n) with CNOT CHAIN (templates.py:164-171)."""
    chain_pairs = [(2 * j, 2 * j + 1) for j in range(n // 2)] + [(2 * j + 1, 2 * j + 2) for j in range((n - 1) // 2)]
    gates = []
    for _ in range(p):
        gates += [RefGate(n, "ry", (q,)) for q in range(n)]
        gates += [RefGate(n, "rz", (q,)) for q in range(n)]
        gates.append(RefGate(n, "chain", None, parts=[full_2q(G.CNOT, a, b, n) for a, b in chain_pairs]))
    return gates


def run(gates, params, n):
    """set_params then the per-gate sparse products (circuit.py:118-125)."""
    k = 0
    for g in gates:
        if g.param:
            g.set_theta(params[k])
            k += 1
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    for g in gates:
        psi = g.op @ psi
    return psi
