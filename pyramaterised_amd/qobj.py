"""QuTiP-``Qobj``-shaped results and operators (the subset pyramaterised uses).

The reference returns QuTiP 4.7 ``Qobj`` kets and operators and its callers use
``overlap``, ``ptrace``, ``tr``, ``dims``, ``data.toarray()``, ``full()``,
``[i][0][0]`` indexing, ``==`` and ``qt.expect`` (circuit.py:136,141;
measure.py:52,58,135,233-235,335-336,418,469; tests.py:71-72,86).  QuTiP is not
a dependency here; these classes provide the same surface:

* ``Ket`` -- an N-qubit state held on the GPU (torch tensor); overlaps,
  partial traces and expectation values run as libpqs reductions.
* ``Qobj`` -- a small host-side object (numpy): one-qubit states, 2x2 reduced
  density matrices, user operators.  Kets built with ``tensor`` remember their
  one-qubit factors so that a circuit can start from them without a transfer.
* ``PauliSum`` -- Hamiltonians such as the default Z0 Z1 and the TFIM
  Hamiltonian (circuit.py:28-31, templates.py:220-227): sums of Pauli strings,
  evaluated on the GPU.
"""
import numbers

import numpy as np
import torch

from . import engine as E

ATOL = 1e-12  # qutip.settings.atol, used by Qobj.__eq__


class _Data:
    """The ``Qobj.data`` view: QuTiP 4 keeps a CSR matrix there; code that
    reads it (``state.data.toarray()``, measure.py:335) gets the same
    accessors over the dense array."""

    def __init__(self, arr):
        self._a = arr

    def toarray(self):
        return self._a.copy()

    def todense(self):
        return np.asmatrix(self._a)

    def tocsr(self):
        import scipy.sparse
        return scipy.sparse.csr_matrix(self._a)

    @property
    def nnz(self):
        return int(np.count_nonzero(self._a))

    @property
    def shape(self):
        return self._a.shape

    def __getitem__(self, i):
        return self._a[i]


def _qubits_of(dim):
    n = int(round(np.log2(dim)))
    if 2 ** n != dim:
        raise ValueError(f"dimension {dim} is not a power of two")
    return n


class Qobj:
    """Host-side QuTiP-shaped object (numpy storage)."""

    __array_priority__ = 100

    def __new__(cls, inpt=None, dims=None, type=None):
        if cls is Qobj and isinstance(inpt, Ket):
            return inpt.copy()
        return super().__new__(cls)

    def __init__(self, inpt=None, dims=None, type=None):
        if isinstance(inpt, Ket):
            return  # handled in __new__
        if isinstance(inpt, Qobj):
            self._a = inpt.full().copy()
            self.dims = [list(d) for d in inpt.dims]
            self.factors = getattr(inpt, "factors", None)
            return
        a = np.array(inpt if inpt is not None else [[0]], dtype=np.complex128)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        self._a = a
        self.factors = None
        if dims is None:
            r, c = a.shape
            if c == 1 and r > 1 and (r & (r - 1)) == 0:
                n = _qubits_of(r)
                dims = [[2] * n, [1] * n]
            elif r == c and r > 1 and (r & (r - 1)) == 0:
                n = _qubits_of(r)
                dims = [[2] * n, [2] * n]
            else:
                dims = [[r], [c]]
        self.dims = dims

    # -- shape / type
    @property
    def shape(self):
        return self._a.shape

    @property
    def type(self):
        r, c = self._a.shape
        if c == 1 and r >= 1 and self.dims[1] and all(d == 1 for d in self.dims[1]):
            return "ket"
        if r == 1 and c > 1:
            return "bra"
        return "oper"

    @property
    def isket(self):
        return self.type == "ket"

    @property
    def isoper(self):
        return self.type == "oper"

    @property
    def isherm(self):
        return self.isoper and np.allclose(self._a, self._a.conj().T, atol=ATOL)

    @property
    def data(self):
        return _Data(self._a)

    def full(self):
        return self._a.copy()

    def __array__(self, dtype=None, copy=None):
        return self._a if dtype is None else self._a.astype(dtype)

    def __getitem__(self, ind):
        out = self._a[ind]
        return np.atleast_2d(out) if np.ndim(out) == 1 else out

    # -- algebra
    def _wrap(self, a, dims=None):
        return Qobj(a, dims=dims or self.dims)

    def __mul__(self, other):
        if isinstance(other, Ket):
            return other.__rmul__(self)
        if isinstance(other, Qobj):
            a = self._a @ other._a
            dims = [self.dims[0], other.dims[1]]
            return Qobj(a, dims=dims)
        if isinstance(other, numbers.Number):
            return self._wrap(self._a * other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return self._wrap(self._a * other)
        return NotImplemented

    def __imul__(self, other):
        r = self.__mul__(other)
        if r is NotImplemented:
            return r
        self._a, self.dims = r._a, r.dims
        return self

    def __truediv__(self, other):
        return self._wrap(self._a / other)

    def __add__(self, other):
        if isinstance(other, Qobj) and not isinstance(other, Ket):
            return self._wrap(self._a + other._a)
        if isinstance(other, Ket):
            return other + self
        if isinstance(other, numbers.Number):
            if other == 0:
                return self._wrap(self._a.copy())
            return self._wrap(self._a + other * np.eye(self._a.shape[0]))
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-1) * other

    def __neg__(self):
        return self._wrap(-self._a)

    def __eq__(self, other):
        if isinstance(other, Ket):
            return other == self
        if not isinstance(other, Qobj):
            return False
        return self.dims == other.dims and self.shape == other.shape and \
            float(np.max(np.abs(self._a - other._a))) < ATOL

    def __hash__(self):
        return id(self)

    def dag(self):
        return Qobj(self._a.conj().T, dims=[self.dims[1], self.dims[0]])

    def conj(self):
        return self._wrap(self._a.conj())

    def tr(self):
        t = np.trace(self._a)
        return float(t.real) if self.isherm else complex(t)

    def norm(self):
        if self.isket or self.type == "bra":
            return float(np.linalg.norm(self._a))
        return float(np.sum(np.abs(np.linalg.eigvalsh(self._a)))) if self.isherm else \
            float(np.sum(np.linalg.svd(self._a, compute_uv=False)))

    def unit(self):
        return self / self.norm()

    def overlap(self, other):
        """<self|other> (QuTiP overlap of two kets)."""
        if isinstance(other, Ket):
            return np.conj(other.overlap(self))
        a = self._a.reshape(-1) if self.isket else self._a.conj().reshape(-1)
        return complex(np.vdot(a, other._a.reshape(-1)))

    def expm(self):
        import scipy.linalg
        return self._wrap(scipy.linalg.expm(self._a))

    def eigenenergies(self):
        return np.linalg.eigvalsh(self._a)

    def groundstate(self):
        w, v = np.linalg.eigh(self._a)
        return float(w[0]), Qobj(v[:, :1], dims=[self.dims[0], [1] * len(self.dims[0])])

    def ptrace(self, sel):
        if not self.isket:
            raise NotImplementedError("ptrace is provided for kets")
        n = len(self.dims[0])
        k = int(sel if np.isscalar(sel) else sel[0])
        t = self._a.reshape(2 ** k, 2, 2 ** (n - k - 1))
        return Qobj(np.einsum("iaj,ibj->ab", t, t.conj()), dims=[[2], [2]])

    def __repr__(self):
        return f"Qobj(dims={self.dims}, shape={self.shape}, type={self.type})\n{self._a}"


class Ket(Qobj):
    """An N-qubit ket resident on the GPU (complex128 or complex64 tensor)."""

    def __new__(cls, *a, **k):
        return object.__new__(cls)

    def __init__(self, tensor, n_qubits=None):
        if isinstance(tensor, Ket):
            return  # Qobj(ket) already returned a copy from __new__
        t = tensor.reshape(-1)
        self.t = t
        self.n = n_qubits if n_qubits is not None else _qubits_of(t.numel())
        self.dims = [[2] * self.n, [1] * self.n]
        self.factors = None

    @property
    def shape(self):
        return (self.t.numel(), 1)

    @property
    def type(self):
        return "ket"

    @property
    def isherm(self):
        return False

    @property
    def _a(self):
        return self.full()

    def full(self):
        return self.t.detach().to("cpu", torch.complex128).numpy().reshape(-1, 1).copy()

    def __getitem__(self, ind):
        v = self.t[ind].item() if isinstance(ind, (int, np.integer)) else self.full()[ind]
        return np.array([[v]]) if isinstance(ind, (int, np.integer)) else v

    def copy(self):
        return Ket(self.t.clone(), self.n)

    def overlap(self, other):
        if isinstance(other, Ket):
            return complex(E.inner(self.t, other.t).item())
        o = torch.as_tensor(np.asarray(other.full()).reshape(-1), device=self.t.device).to(self.t.dtype)
        return complex(E.inner(self.t, o).item())

    def ptrace(self, sel):
        k = int(sel if np.isscalar(sel) else sel[0])
        rho, _ = E.mw(self.t, self.n)
        r = rho[0, k].cpu().numpy()
        m = np.array([[r[0], r[2] + 1j * r[3]], [r[2] - 1j * r[3], r[1]]])
        return Qobj(m, dims=[[2], [2]])

    def norm(self):
        return float(np.sqrt(abs(self.overlap(self))))

    def unit(self):
        return Ket(self.t / self.norm(), self.n)

    def conj(self):
        return Ket(self.t.conj().resolve_conj(), self.n)

    def dag(self):
        return Qobj(self.full().conj().T, dims=[[1] * self.n, [2] * self.n])

    def _other_tensor(self, other):
        if isinstance(other, Ket):
            return other.t.to(self.t.dtype)
        return torch.as_tensor(np.asarray(other.full()).reshape(-1), device=self.t.device).to(self.t.dtype)

    def __add__(self, other):
        if isinstance(other, numbers.Number) and other == 0:
            return self.copy()
        return Ket(self.t + self._other_tensor(other), self.n)

    __radd__ = __add__

    def __sub__(self, other):
        return Ket(self.t - self._other_tensor(other), self.n)

    def __neg__(self):
        return Ket(-self.t, self.n)

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return Ket(self.t * other, self.n)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return Ket(self.t * other, self.n)
        apply = getattr(other, "apply_to", None)
        if apply is not None:
            return apply(self)
        if isinstance(other, Qobj):
            return PauliSum.from_dense(other).apply_to(self)
        return NotImplemented

    def __truediv__(self, other):
        return Ket(self.t / other, self.n)

    def __eq__(self, other):
        if not isinstance(other, Qobj):
            return False
        if self.dims != other.dims:
            return False
        d = self.t - self._other_tensor(other)
        return float(d.abs().max().item()) < ATOL

    def __hash__(self):
        return id(self)

    def __repr__(self):
        return f"Ket(n_qubits={self.n}, dtype={self.t.dtype}, device={self.t.device})"


# ------------------------------------------------------------------ constructors

def basis(n, k=0):
    v = np.zeros((n, 1), dtype=np.complex128)
    v[k, 0] = 1
    q = Qobj(v, dims=[[n], [1]])
    if n == 2:
        q.factors = [v.ravel().copy()]
    return q


class ProductKet(Qobj):
    """Tensor product of one-qubit kets kept in factored form (the full
    vector is built only when asked for)."""

    def __init__(self, factors):
        self.factors = [np.asarray(f, dtype=np.complex128).ravel() for f in factors]
        n = len(self.factors)
        self.dims = [[2] * n, [1] * n]

    @property
    def _a(self):
        a = np.array([1.0 + 0j])
        for f in self.factors:
            a = np.kron(a, f)
        return a.reshape(-1, 1)

    @_a.setter
    def _a(self, v):
        raise AttributeError("a ProductKet is immutable")

    @property
    def shape(self):
        return (2 ** len(self.factors), 1)

    def full(self):
        return self._a

    def _wrap(self, a, dims=None):
        return Qobj(a, dims=dims or self.dims)


def tensor(*args):
    items = args[0] if len(args) == 1 and isinstance(args[0], (list, tuple)) else args
    def _factors(it):
        f = getattr(it, "factors", None)
        if f is not None and it.dims[1] == [1] * len(it.dims[0]):
            return f
        if not isinstance(it, Ket) and tuple(it.shape) == (2, 1):
            return [np.asarray(it.full()).ravel()]
        return None

    fs = [_factors(it) for it in items]
    if items and all(f is not None for f in fs):
        return ProductKet([f for fl in fs for f in fl])
    a = np.array([[1.0 + 0j]])
    d0, d1, factors = [], [], []
    for it in items:
        a = np.kron(a, it.full())
        d0 += it.dims[0]
        d1 += it.dims[1]
        f = getattr(it, "factors", None)
        factors = factors + f if (factors is not None and f is not None) else None
    q = Qobj(a, dims=[d0, d1])
    q.factors = factors if factors and all(len(f) == 2 for f in factors) else None
    return q


def qeye(n):
    return Qobj(np.eye(n), dims=[[n], [n]])


def sigmax():
    return Qobj([[0, 1], [1, 0]])


def sigmay():
    return Qobj([[0, -1j], [1j, 0]])


def sigmaz():
    return Qobj([[1, 0], [0, -1]])


def bell_state(state="00"):
    """QuTiP bell_state: 00 -> (|00>+|11>), 01 -> (|00>-|11>),
    10 -> (|01>+|10>), 11 -> (|01>-|10>), each / sqrt(2)."""
    v = np.zeros(4, dtype=np.complex128)
    if state == "00":
        v[[0, 3]] = [1, 1]
    elif state == "01":
        v[[0, 3]] = [1, -1]
    elif state == "10":
        v[[1, 2]] = [1, 1]
    elif state == "11":
        v[[1, 2]] = [1, -1]
    return Qobj(v / np.sqrt(2), dims=[[2, 2], [1, 1]])


def to_ket(q, n_qubits=None, dtype=torch.complex128):
    """Upload a host ket to the GPU."""
    if isinstance(q, Ket):
        return q
    a = np.asarray(q.full() if isinstance(q, Qobj) else q, dtype=np.complex128).reshape(-1)
    return Ket(torch.as_tensor(a, device=E.device()).to(dtype), n_qubits)


def expect(op, state):
    """qt.expect(op, ket) for Pauli-sum (or small dense) operators."""
    ps = op if isinstance(op, PauliSum) else PauliSum.from_dense(op)
    ket = to_ket(state)
    val = complex(E.expect_pauli(ket.t, ket.n, ps.terms())[0].item())
    return val.real if ps.is_hermitian() else val


# ------------------------------------------------------------------ Pauli sums

_PMAT = {0: np.eye(2), 1: np.array([[0, 1], [1, 0]]), 2: np.array([[0, -1j], [1j, 0]]),
         3: np.array([[1, 0], [0, -1]])}
# product of single-qubit Paulis a*b = phase * c  (0=I 1=X 2=Y 3=Z)
_PMUL = {}
for _a in range(4):
    for _b in range(4):
        _m = _PMAT[_a] @ _PMAT[_b]
        for _c in range(4):
            _ph = np.trace(_PMAT[_c].conj().T @ _m) / 2
            if abs(_ph) > 0.5:
                _PMUL[(_a, _b)] = (_c, complex(_ph))


class PauliSum:
    """sum_t c_t P_t over n qubits; a string is a tuple of codes 0..3 per qubit."""

    def __init__(self, n, terms=None):
        self.n = n
        self.d = {}
        for s, c in (terms or {}).items() if isinstance(terms, dict) else (terms or []):
            self.d[tuple(s)] = self.d.get(tuple(s), 0) + c

    @staticmethod
    def single(n, q, code, coef=1.0):
        s = [0] * n
        s[q] = code
        return PauliSum(n, {tuple(s): coef})

    @property
    def dims(self):
        return [[2] * self.n, [2] * self.n]

    def copy(self):
        return PauliSum(self.n, dict(self.d))

    def __add__(self, other):
        if isinstance(other, numbers.Number):
            if other == 0:
                return self.copy()
            other = PauliSum(self.n, {(0,) * self.n: other})
        if not isinstance(other, PauliSum):
            return NotImplemented
        out = self.copy()
        for s, c in other.d.items():
            out.d[s] = out.d.get(s, 0) + c
        return out

    __radd__ = __add__

    def __neg__(self):
        return self * -1

    def __sub__(self, other):
        return self + (-1) * other

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return PauliSum(self.n, {s: c * other for s, c in self.d.items()})
        if isinstance(other, PauliSum):
            out = {}
            for s1, c1 in self.d.items():
                for s2, c2 in other.d.items():
                    ph, s = 1 + 0j, []
                    for a, b in zip(s1, s2):
                        c, p = _PMUL[(a, b)]
                        s.append(c)
                        ph *= p
                    out[tuple(s)] = out.get(tuple(s), 0) + ph * c1 * c2
            return PauliSum(self.n, out)
        if isinstance(other, Ket):
            return self.apply_to(other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return self * other
        return NotImplemented

    def is_hermitian(self):
        return all(abs(complex(c).imag) < 1e-14 for c in self.d.values())

    @property
    def isherm(self):
        return self.is_hermitian()

    def terms(self):
        """[(coef, xmask, zmask)] with qubit masks (bit q = qubit q)."""
        out = []
        for s, c in self.d.items():
            if c == 0:
                continue
            x = z = 0
            for q, code in enumerate(s):
                if code in (1, 2):
                    x |= 1 << q
                if code in (2, 3):
                    z |= 1 << q
            out.append((complex(c), x, z))
        return out

    def full(self):
        m = np.zeros((2 ** self.n, 2 ** self.n), dtype=np.complex128)
        for s, c in self.d.items():
            k = np.array([[1.0 + 0j]])
            for code in s:
                k = np.kron(k, _PMAT[code])
            m += c * k
        return m

    def expect(self, ket):
        return expect(self, ket)

    def apply_to(self, ket):
        """H|psi> on the GPU (one generator op of the sweep kernel)."""
        from .gates import _pauli_sum_plan
        return _pauli_sum_plan(self, ket)

    @staticmethod
    def from_dense(op):
        """Pauli decomposition of a dense operator (n <= 12), in O(n 4^n):
        with P = X^x Z^z, Tr(P^dag A) = sum_j (-1)^{z.j} A[j^x, j], so for
        each X mask x the coefficients over all Z masks are one
        Walsh-Hadamard transform of A's x-th off-diagonal; a Y on a qubit is
        i X Z there."""
        a = np.asarray(op.full() if hasattr(op, "full") else op, dtype=np.complex128)
        n = _qubits_of(a.shape[0])
        if n > 12:
            raise ValueError("dense operators above 12 qubits: pass a PauliSum")
        d = 2 ** n
        j = np.arange(d)
        m = a[j[None, :] ^ j[:, None], j[None, :]]  # m[x, j] = A[j ^ x, j]
        h = 1
        while h < d:  # Walsh-Hadamard transform along j
            m = m.reshape(d, -1, 2, h)
            m = np.stack([m[:, :, 0, :] + m[:, :, 1, :], m[:, :, 0, :] - m[:, :, 1, :]], axis=2)
            h *= 2
        m = m.reshape(d, d) / d  # m[x, z] = Tr((X^x Z^z)^dag A) / d
        xs, zs = np.nonzero(np.abs(m) > 1e-15)
        shifts = (n - 1 - np.arange(n))[None, :]
        bx, bz = (xs[:, None] >> shifts) & 1, (zs[:, None] >> shifts) & 1
        codes = np.where(bz == 0, bx, np.where(bx == 0, 3, 2))  # I=0 X=1 Y=2 Z=3
        ny = np.sum(bx & bz, axis=1)
        # a Y-string is i^ny X^x Z^z, so its coefficient is (-i)^ny m[x, z]
        coef = m[xs, zs] * (-1j) ** ny
        return PauliSum(n, dict(zip(map(tuple, codes.tolist()), coef.tolist())))

    def __repr__(self):
        return f"PauliSum(n={self.n}, terms={len(self.d)})"
