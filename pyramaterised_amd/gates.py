"""Gate library with the reference's names and protocol, as local operators.

Reference: pyramaterised/gates.py.  Every reference gate holds a full tensored
2^N x 2^N QuTiP operator rebuilt on each ``set_theta`` (gates.py:39-46,
119-131).  Here a gate is a short list of *primitive ops* -- a 2x2 / 4x4 local
matrix or a parameterised rotation on named qubits (``pqs_op`` of
include/pqs.h) -- plus the Pauli-sum generator D of its derivative, so that
``D U = dU/dtheta`` as in the reference's ``derivative()``
(gates.py:133-138, 185-187, 451-457, 519-522).  ``PQC`` compiles the flat
gate list into a plan that the HIP engine runs; the gate objects never build
a 2^N operator.

Compatibility surface kept from the reference (gates.py:49-100): attributes
``q_N, operation, param_count, is_param, theta, phi, q_on / q1, q2``; methods
``set_theta, set_phi, derivative, parameterised_derivative, flip_pauli`` and
``*`` / ``+`` through ``operation`` (here a lazily built ``LocalOperator``
that applies itself to a GPU ket).
"""
import ctypes as C
import operator
from functools import reduce
from itertools import permutations
from typing import Literal, Tuple, Type, TypeAlias, Union  # noqa: F401

import numpy as np

from . import _lib as L
from .qobj import Ket, PauliSum

rng = np.random.default_rng(1)  # gates.py:10

QubitIndex = int
QubitList = Union[list, tuple]
QubitNumber = int
Angle = Union[int, float]
Layer = list
RotationLayer = list
EntanglingLayer = list

_SQRT2 = np.sqrt(2.0)
_PAULI_CODE = {"i": 0, "x": 1, "y": 2, "z": 3}


def prod(factors):
    """gates.py:30-31."""
    return reduce(operator.mul, factors, 1)


def flatten(lst):
    """gates.py:34-35."""
    return [item for sub in lst for item in sub]


# ---------------------------------------------------------------- op records

def _op(kind, q0, q1=-1, slot=-1, slot2=-1, scale=1.0, offset=0.0, m=None, tangent=-1, pauli=0):
    o = L.Op()
    o.kind, o.q0, o.q1, o.slot, o.slot2 = kind, q0, q1, slot, slot2
    o.tangent, o.pauli, o.reserved = tangent, pauli, 0
    o.scale, o.offset = scale, offset
    if m is not None:
        flat = np.asarray(m, dtype=np.complex128).reshape(-1)
        for i, v in enumerate(flat):
            o.m[2 * i] = v.real
            o.m[2 * i + 1] = v.imag
    return o


def _gen(tangent, code, qs, coef):
    """One Pauli-string term coef * P(q0) [(x) P(q1)] of a generator."""
    o = _op(L.OP_GEN, qs[0], qs[1] if len(qs) > 1 else -1, tangent=tangent, pauli=code)
    o.m[0], o.m[1] = complex(coef).real, complex(coef).imag
    return o


# ------------------------------------------------------------ local operators

class LocalOperator:
    """Product (and, via ``+``, sum) of primitive ops on n qubits; the
    ``operation`` of a gate.  ``op * ket`` runs on the GPU."""

    def __init__(self, n, ops, terms=None):
        self.q_N = n
        self.ops = list(ops)  # applied in list order (first op first)
        self._terms = terms   # for sums: list of LocalOperator

    @property
    def dims(self):
        return [[2] * self.q_N, [2] * self.q_N]

    def __mul__(self, other):
        if isinstance(other, Ket):
            return self.apply_to(other)
        if isinstance(other, Gate):
            other = other.operation
        if isinstance(other, LocalOperator):
            if self._terms or other._terms:
                lt = self._terms or [self]
                rt = other._terms or [other]
                return LocalOperator(self.q_N, [], terms=[LocalOperator(self.q_N, r.ops + l.ops)
                                                          for l in lt for r in rt])
            return LocalOperator(self.q_N, other.ops + self.ops)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, Gate):
            return other.operation * self
        return NotImplemented

    def __add__(self, other):
        if isinstance(other, (int, float)) and other == 0:
            return self
        if isinstance(other, Gate):
            other = other.operation
        if isinstance(other, LocalOperator):
            return LocalOperator(self.q_N, [], terms=(self._terms or [self]) + (other._terms or [other]))
        return NotImplemented

    __radd__ = __add__

    def apply_to(self, ket):
        from . import engine as E
        if self._terms:
            out = None
            for t in self._terms:
                r = t.apply_to(ket)
                out = r if out is None else out + r
            return out
        if not self.ops:
            return ket.copy()
        n = ket.n
        ops, params = self.ops, None
        if _n_slots(ops) == 0:
            ops, params = _angles_as_params(ops)
        key = (n, str(ket.t.dtype), bytes((L.Op * len(ops))(*ops)))
        plan = _APPLY_PLANS.pop(key, None)
        if plan is None:
            plan = E.Plan(n, ops, dtype=ket.t.dtype)
            while len(_APPLY_PLANS) >= _APPLY_PLANS_MAX:
                _APPLY_PLANS.pop(next(iter(_APPLY_PLANS)))
        _APPLY_PLANS[key] = plan  # most recently used last
        has_gen = any(o.tangent >= 0 for o in ops)
        if params is None:
            params = np.zeros(max(1, _n_slots(ops)))
        out = E.run(plan, params, E.Init.from_vector(ket.t), tangents=[0] if has_gen else None)
        return Ket(out[0], n)

    def full(self):
        """Dense matrix, built by applying the operator to every basis ket on
        the GPU (compatibility only; small n)."""
        import torch
        from . import engine as E
        dim = 2 ** self.q_N
        cols = []
        for i in range(dim):
            v = torch.zeros(dim, dtype=torch.complex128, device=E.device())
            v[i] = 1
            cols.append(self.apply_to(Ket(v, self.q_N)).full().ravel())
        return np.array(cols).T


# Plans of ``op * ket`` products by op structure (n, dtype, ops with their
# constant rotation angles taken out), least recently used first: a gate
# applied again at other angles (``for g in gates: state = g * state``, the
# columns of ``full()``) reuses its plan.
_APPLY_PLANS = {}
_APPLY_PLANS_MAX = 256
_ROT_KINDS = (L.OP_RX, L.OP_RY, L.OP_RZ, L.OP_RZZ, L.OP_RXX, L.OP_RYY)


def _angles_as_params(ops):
    """Ops without parameter slots -> (the same ops with every rotation's
    constant angle as parameter k of a params vector, params)."""
    out, params = [], []
    for o in ops:
        c = L.Op.from_buffer_copy(o)
        if c.kind in _ROT_KINDS and c.slot < 0:
            c.slot, c.scale = len(params), 1.0
            params.append(c.offset)
            c.offset = 0.0
        out.append(c)
    return out, np.asarray(params if params else [0.0], dtype=np.float64)


def _n_slots(ops):
    m = -1
    for o in ops:
        m = max(m, o.slot, o.slot2)
    return m + 1


def _pauli_sum_plan(ps, ket):
    """Apply a PauliSum to a GPU ket with one generator op (tangent 0)."""
    ops = []
    for coef, x, z in ps.terms():
        qs = [q for q in range(ps.n) if (x >> q) & 1 or (z >> q) & 1]
        # split a multi-qubit string into generator terms of <= 2 qubits is not
        # possible (the sum needs each string as one term), so encode strings of
        # any length through the dense fallback below when they exceed 2 qubits.
        if len(qs) > 2:
            return _pauli_sum_dense(ps, ket)
        codes = [(1 if (x >> q) & 1 and not (z >> q) & 1 else 2 if (x >> q) & 1 else 3) for q in qs]
        if not qs:
            ops.append(_gen(0, 0, [0], coef))
        elif len(qs) == 1 or codes[0] == codes[1]:
            ops.append(_gen(0, codes[0], qs, coef))
        else:
            return _pauli_sum_dense(ps, ket)
    return LocalOperator(ps.n, ops).apply_to(ket)


def _pauli_sum_dense(ps, ket):
    terms = [LocalOperator(ps.n, [_op(L.OP_MAT1, q, m=_pmat(code)) for q, code in enumerate(s) if code]
                           + ([_op(L.OP_MAT1, 0, m=np.eye(2) * c)]))
             for s, c in ps.d.items()]
    return LocalOperator(ps.n, [], terms=terms).apply_to(ket)


def _pmat(code):
    return {1: [[0, 1], [1, 0]], 2: [[0, -1j], [1j, 0]], 3: [[1, 0], [0, -1]]}[code]


# ---------------------------------------------------------------------- gates

def _rot(axis, t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    if axis == "x":
        return np.array([[c, -1j * s], [-1j * s, c]])
    if axis == "y":
        return np.array([[c, -s], [s, c]])
    return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]])


class Gate:
    """gates.py:49-100."""

    def __init__(self, q_N):
        self.q_N = q_N
        self.param_count = 0
        self.is_param = False
        self.theta = 0
        self.phi = 0

    # -- engine interface
    def prim_ops(self, slot):
        """Primitive ops; parameterised ones read params[slot] (and slot+1)."""
        return []

    def gen_ops(self, tangent, slot, which=1):
        """Generator ops of derivative ``which`` as tangent ``tangent``."""
        return [_gen(tangent, 0, [0], 1.0)]

    def _angles(self):
        return float(self.theta), float(self.phi)

    # -- compat
    @property
    def operation(self):
        th, ph = self._angles()
        ops = self._const_ops(th, ph)
        return LocalOperator(self.q_N, ops)

    def get_op(self):
        """The gate's operator at its current angles (gates.py:122 and the
        other get_op methods); ``operation`` is the same, built on access."""
        return self.operation

    def _const_ops(self, th, ph):
        out = []
        for o in self.prim_ops(0):
            if o.slot >= 0:
                o.offset = o.scale * (th if o.slot == 0 else ph) + o.offset
                o.scale, o.slot = 0.0, -1
            if o.kind == L.OP_FSIM and o.slot2 >= 0:
                o.kind, o.slot2 = L.OP_MAT2, -1
                m = _fsim(o.offset, ph)
                for i, v in enumerate(m.reshape(-1)):
                    o.m[2 * i], o.m[2 * i + 1] = v.real, v.imag
            out.append(o)
        return out

    def __mul__(self, b):
        if isinstance(b, Gate):
            return self.operation * b.operation
        return self.operation * b

    def __rmul__(self, b):
        if isinstance(b, Gate):
            return b.operation * self.operation
        return b * self.operation

    def __add__(self, b):
        if isinstance(b, Gate):
            return self.operation + b.operation
        return self.operation + b

    def __radd__(self, b):
        if isinstance(b, Gate):
            return b.operation + self.operation
        return b + self.operation

    def set_theta(self, theta):
        return

    def set_phi(self, phi):
        return

    def derivative(self):
        return LocalOperator(self.q_N, self.gen_ops(0, 0))

    def parameterised_derivative(self, param):
        return self.derivative()

    def flip_pauli(self):
        pass


class PRot(Gate):
    """gates.py:106-147: rotation exp(-i theta/2 P) on q_on."""

    axis = "i"

    def __init__(self, q_on, q_N):
        self.q_on = q_on
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = True
        self.param_count = 1
        self.set_properties()

    def set_properties(self):
        self.pauli_sign = 1

    _scale = 1.0

    def prim_ops(self, slot):
        kind = {"x": L.OP_RX, "y": L.OP_RY, "z": L.OP_RZ}[self.axis]
        return [_op(kind, self.q_on, slot=slot, scale=self._scale, offset=self._offset())]

    def _offset(self):
        return 0.0

    def gen_ops(self, tangent, slot, which=1):
        # D = -i/2 * scale * sigma on q_on (gates.py:133-138; 185-187)
        return [_gen(tangent, _PAULI_CODE[self.axis], [self.q_on], -0.5j * self._scale)]

    def set_theta(self, theta):
        self.theta = theta

    def flip_pauli(self):
        # gates.py:140-141 negates ``pauli`` after ``fock`` was built from it,
        # so the derivative is unchanged; kept for API parity.
        self.pauli_sign = -getattr(self, "pauli_sign", 1)

    def __repr__(self):
        return f"{type(self).__name__}({self.theta:.2f})@q{self.q_on}"


class I(PRot):
    """gates.py:150-156: identity that consumes a parameter; D = -i/2 I."""

    def prim_ops(self, slot):
        return []

    def gen_ops(self, tangent, slot, which=1):
        return [_gen(tangent, 0, [self.q_on], -0.5j)]


class R_x(PRot):
    axis = "x"


class R_y(PRot):
    axis = "y"


class R_z(PRot):
    axis = "z"


class negative_R_z(R_z):
    """gates.py:180-187: theta -> -theta, D = +i/2 Z."""

    _scale = -1.0

    def set_theta(self, theta):
        self.theta = -1 * theta

    def _angles(self):
        # ``theta`` already holds the negated angle; the op negates again
        return -float(self.theta), float(self.phi)


class offset_R_z(R_z):
    """gates.py:190-205: theta -> theta + offset."""

    def __init__(self, q_on, q_N, offset):
        self.offset = offset
        super().__init__(q_on, q_N)

    def _offset(self):
        return float(self.offset)

    def set_theta(self, theta):
        self.theta = theta + self.offset

    def _angles(self):
        return float(self.theta) - float(self.offset), float(self.phi)


class _Fixed1(Gate):
    """Fixed one-qubit gate (is_param False)."""

    def __init__(self, q_on, q_N):
        self.q_on = q_on
        self.q_N = q_N
        self.theta = np.pi / 2
        self.phi = 0
        self.is_param = False
        self.param_count = 0

    def matrix(self):
        raise NotImplementedError

    def prim_ops(self, slot):
        return [_op(L.OP_MAT1, self.q_on, m=self.matrix())]

    def set_theta(self, angle):
        return None

    def __repr__(self):
        return f"{type(self).__name__}@q{self.q_on}"


class H(_Fixed1, PRot):
    """gates.py:211-232: H = X * R_y(pi/2)."""

    def matrix(self):
        return np.array([[0, 1], [1, 0]]) @ _rot("y", np.pi / 2)


class sqrtH(H):
    """gates.py:235-242: element-wise complex square root of H."""

    def matrix(self):
        return np.sqrt(H.matrix(self).astype(np.complex128))


class X(H):
    """gates.py:245-249."""

    def matrix(self):
        return np.array([[0, 1], [1, 0]])


class S(H):
    """gates.py:286-291: phasegate(pi/2)."""

    def matrix(self):
        return np.diag([1, 1j])


class T(H):
    """gates.py:294-300: t_gate = phasegate(pi/4)."""

    def matrix(self):
        return np.diag([1, np.exp(0.25j * np.pi)])


class fixed_R_y(_Fixed1, R_y):
    """gates.py:252-266."""

    def __init__(self, q_on, q_N, theta):
        super().__init__(q_on, q_N)
        self.theta = theta

    def matrix(self):
        return _rot("y", self.theta)


class fixed_R_z(_Fixed1, R_z):
    """gates.py:269-283."""

    def __init__(self, q_on, q_N, theta):
        super().__init__(q_on, q_N)
        self.theta = theta

    def matrix(self):
        return _rot("z", self.theta)


# ------------------------------------------------------------ entangling gates

class EntGate(Gate):
    """gates.py:306-324."""

    def __init__(self, qs_on, q_N):
        self.q1, self.q2 = qs_on[0], qs_on[1]
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = False
        self.param_count = 0

    def prim_ops(self, slot):
        return []

    def __repr__(self):
        return f"{type(self).__name__}@q{self.q1},q{self.q2}"


class CNOT(EntGate):
    def prim_ops(self, slot):
        return [_op(L.OP_CX, self.q1, self.q2)]


class CPHASE(EntGate):
    """CZ (gates.py:333-337)."""

    def prim_ops(self, slot):
        return [_op(L.OP_CZ, self.q1, self.q2)]


class CZ(CPHASE):
    pass


_SQRT_ISWAP = np.array([[1, 0, 0, 0], [0, 1 / _SQRT2, 1j / _SQRT2, 0],
                        [0, 1j / _SQRT2, 1 / _SQRT2, 0], [0, 0, 0, 1]])


class sqrtiSWAP(EntGate):
    def prim_ops(self, slot):
        return [_op(L.OP_MAT2, self.q1, self.q2, m=_SQRT_ISWAP)]


class CHAIN(EntGate):
    """gates.py:355-379: top pairs (2j, 2j+1) then bottom pairs (2j+1, 2j+2);
    the first listed pair is applied first (prod of the reversed list)."""

    def __init__(self, entangler, q_N):
        self.entangler = entangler
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = False
        self.param_count = 0
        N = q_N
        top = [[2 * j, 2 * j + 1] for j in range(N // 2)]
        bottom = [[2 * j + 1, 2 * j + 2] for j in range((N - 1) // 2)]
        self.pairs = top + bottom
        self.subgates = [entangler(p, N) for p in self.pairs]

    def prim_ops(self, slot):
        return [o for g in self.subgates for o in g.prim_ops(slot)]

    def __repr__(self):
        return f"CHAIN connected {self.entangler.__name__}s"


class ALLTOALL(CHAIN):
    """gates.py:382-401: every ordered pair of permutations(range(N), 2)."""

    def __init__(self, entangler, q_N):
        super().__init__(entangler, q_N)
        self.pairs = list(permutations(range(q_N), 2))
        self.subgates = [entangler(p, q_N) for p in self.pairs]

    def __repr__(self):
        return f"ALL connected {self.entangler.__name__}s"


class DenseOperator:
    """A dense 2^N x 2^N operator (host matrix) with the ``operation``
    protocol: ``op * ket`` applies it on the GPU (pqs_apply_dense),
    ``op * op`` multiplies, ``full()`` returns the matrix."""

    def __init__(self, n, matrix):
        self.q_N = n
        self.m = np.asarray(matrix, dtype=np.complex128)

    @property
    def dims(self):
        return [[2] * self.q_N, [2] * self.q_N]

    def full(self):
        return self.m.copy()

    def __mul__(self, other):
        from . import engine as E
        if isinstance(other, Ket):
            return Ket(E.apply_dense(self.m, other.t), other.n)
        if isinstance(other, Gate):
            other = other.operation
        if isinstance(other, DenseOperator):
            return DenseOperator(self.q_N, self.m @ other.m)
        if isinstance(other, LocalOperator):
            return DenseOperator(self.q_N, self.m @ other.full())
        if isinstance(other, (int, float, complex)):
            return DenseOperator(self.q_N, self.m * other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, (int, float, complex)):
            return DenseOperator(self.q_N, self.m * other)
        if isinstance(other, LocalOperator):
            return DenseOperator(self.q_N, other.full() @ self.m)
        return NotImplemented


class ARBGATE(Gate):
    """gates.py:407-435: U = expm(-i theta H) for a user operator H on all
    q_N qubits, and the reference's derivative D = -i H / 2 (half of
    dU/dtheta = -i H U; kept as written), so the derivative state applies
    D U.  A dense 2^N operator has no local-gate form: PQC runs a circuit
    that holds one in segments of local gates, with U (and D U) built on the
    host -- from an eigen-decomposition of H, computed once, when H is
    Hermitian, else scipy's expm -- and applied on the GPU in between
    (pqs_apply_dense, up to 14 qubits)."""

    dense = True

    def __init__(self, Ham):
        self._Ham = Ham
        self.theta = 0
        self.phi = 0
        self.is_param = True
        self.param_count = 1
        if hasattr(Ham, "full"):
            H = np.asarray(Ham.full(), dtype=np.complex128)
        else:
            H = np.asarray(Ham, dtype=np.complex128)
        if H.ndim != 2 or H.shape[0] != H.shape[1] or H.shape[0] & (H.shape[0] - 1):
            raise ValueError("ARBGATE needs a square 2^N x 2^N operator")
        self._H = H
        self.q_N = int(H.shape[0]).bit_length() - 1
        self._eig = None

    def prim_ops(self, slot):
        raise NotImplementedError("ARBGATE is a dense operator: PQC applies it between plans")

    def gen_ops(self, tangent, slot, which=1):
        raise NotImplementedError("ARBGATE is a dense operator: PQC applies its derivative between plans")

    def matrix(self, theta=None):
        """U = expm(-i theta H) (gates.py:416-419)."""
        import scipy.linalg
        th = float(self.theta if theta is None else theta)
        if self._eig is None:
            herm = np.allclose(self._H, self._H.conj().T, atol=1e-14, rtol=0)
            self._eig = np.linalg.eigh(self._H) if herm else False
        if self._eig is not False:
            w, V = self._eig
            return (V * np.exp(-1j * th * w)) @ V.conj().T
        return scipy.linalg.expm(-1j * th * self._H)

    def derivative_matrix(self, theta=None):
        """D U with the reference's D = -i H / 2 (gates.py:427-429)."""
        return (-0.5j * self._H) @ self.matrix(theta)

    @property
    def operation(self):
        return DenseOperator(self.q_N, self.matrix())

    def get_op(self):
        return self.operation

    def derivative(self):
        return DenseOperator(self.q_N, -0.5j * self._H)

    def set_theta(self, theta):
        self.theta = theta

    def flip_pauli(self):
        self.pauli = -1 * self.pauli  # AttributeError as in the reference (gates.py:420-421)

    def __repr__(self):
        return f"{type(self).__name__}({self.theta:.2f})"


# ------------------------------------------------------------ shared parameter

def _conj_op(o):
    """The element-wise complex conjugate of a primitive op, as an op of the
    same kind (rotations: angle -> -angle; R_y, CNOT and CZ are real)."""
    c = _op(o.kind, o.q0, o.q1, slot=o.slot, slot2=o.slot2, scale=o.scale, offset=o.offset,
            tangent=o.tangent, pauli=o.pauli)
    for i in range(32):
        c.m[i] = o.m[i]
    if o.kind in (L.OP_MAT1, L.OP_MAT2):
        for i in range(1, 32, 2):
            c.m[i] = -o.m[i]
    elif o.kind in (L.OP_RX, L.OP_RZ, L.OP_RXX, L.OP_RYY, L.OP_RZZ):
        c.scale, c.offset = -o.scale, -o.offset
    elif o.kind == L.OP_FSIM and o.slot2 < 0:
        c.scale, c.offset = -o.scale, -o.offset  # fixed_fSim: phi = 0
    elif o.kind not in (L.OP_RY, L.OP_CX, L.OP_CZ):
        raise NotImplementedError(f"no conjugate form for op kind {o.kind}")
    return c


def _pauli_string(g):
    """{qubit: 'x'|'y'|'z'} of a Pauli rotation gate, or None for any other
    gate (fixed gates, I, fSim, blocks)."""
    if isinstance(g, (_Fixed1, I, shared_parameter)) or not isinstance(g, PRot):
        return None
    if isinstance(g, RR):
        return {g.q1: g.axis, g.q2: g.axis}
    if g.axis in ("x", "y", "z") and type(g).prim_ops is PRot.prim_ops:
        return {g.q_on: g.axis}
    return None


class shared_parameter(PRot):
    """gates.py:441-484: one angle on every sub-gate, U = prod(layer[::-1])
    (first listed applied first).  With commute=True the derivative is the
    sum of the sub-gates' generators (gates.py:451-457).  With commute=False
    the reference's derivative is sum_k prod(layer with g_k -> D_k g_k)
    times conj(U) (gates.py:458-466), applied after U by take_derivative;
    where every sub-gate is a Pauli rotation whose string is symmetric (an
    even number of Y) and the strings commute pairwise, conj(U) = U^dag and
    each D_k commutes with the block, so that form is the generator sum too.
    Other commute=False blocks are differentiated through a plan of their own
    (``PQC`` builds it: the circuit with conj(U) and the interleaved
    D_k inserted)."""

    def __init__(self, layer, q_N, commute=True):
        self.layer = layer
        self.theta = 0
        self.phi = 0
        self.q_N = q_N
        self.is_param = True
        self.param_count = 1
        self.commute = commute

    def prim_ops(self, slot):
        return [o for g in self.layer for o in g.prim_ops(slot)]

    def generator_sum(self):
        """Whether the derivative is the sum of the sub-gates' generators
        inserted after the block (always for commute=True)."""
        if self.commute:
            return True
        strings = [_pauli_string(g) for g in self.layer]
        if any(s is None for s in strings):
            return False
        for s in strings:
            if sum(1 for a in s.values() if a == "y") % 2:
                return False  # P^T = -P: U is not symmetric
        for i in range(len(strings)):
            for j in range(i + 1, len(strings)):
                anti = sum(1 for q, a in strings[i].items() if q in strings[j] and strings[j][q] != a)
                if anti % 2:
                    return False
        return True

    @property
    def special_derivative(self):
        return not self.generator_sum()

    def gen_ops(self, tangent, slot, which=1):
        if not self.generator_sum():
            raise NotImplementedError(
                "the commute=False derivative of this block (gates.py:458-466) needs the circuit around it: "
                "use derivative_ops(), which PQC uses through a plan of its own")
        return [o for g in self.layer for o in g.gen_ops(tangent, slot)]

    def derivative_ops(self, tangent, slot):
        """Ops that map the state after the block, U psi, to the reference's
        derivative state of the block (gates.py:458-466 followed by the
        multiplication by U in take_derivative, circuit.py:165-169):
        conj(U) (the conjugated sub-gates in order), then the sub-gates again
        with D_k (tangent ``tangent``) right after g_k.  Each D_k is its own
        run of generator terms, so the engine sums the k terms (product
        rule over runs of one tangent)."""
        out = [_conj_op(o) for g in self.layer for o in g.prim_ops(slot)]
        for g in self.layer:
            out += g.prim_ops(slot)
            out += g.gen_ops(tangent, slot)
        return out

    def derivative(self):
        if self.generator_sum():
            return LocalOperator(self.q_N, self.gen_ops(0, 0))
        # D with D * U = the derivative: the terms act after U (the caller
        # multiplies D by the gate), with the angle bound to self.theta
        th = float(self.theta)
        terms = []
        conj = [_conj_op(o) for g in self.layer for o in g._const_ops(th, 0.0)]
        for k in range(len(self.layer)):
            ops = list(conj)
            for j, g in enumerate(self.layer):
                ops += g._const_ops(th, 0.0)
                if j == k:
                    for o in g.gen_ops(0, 0):
                        if o.slot >= 0:  # fSim-family derivative: bind its angle
                            o.offset = o.scale * th + o.offset
                            o.scale, o.slot = 0.0, -1
                        ops.append(o)
            terms.append(LocalOperator(self.q_N, ops))
        return LocalOperator(self.q_N, [], terms=terms)

    def set_theta(self, theta):
        self.theta = theta
        for g in self.layer:
            g.set_theta(theta)

    def flip_pauli(self):
        for g in self.layer:
            g.flip_pauli()

    def __repr__(self):
        return f"Block of {self.layer}"


class RR(PRot):
    """gates.py:492-527: cos(theta/2) I - i sin(theta/2) P(x)P; D = -i/2 P(x)P."""

    kind = None

    def __init__(self, qs_on, q_N):
        self.q1, self.q2 = qs_on[0], qs_on[1]
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = True
        self.param_count = 1
        self.set_properties()

    def prim_ops(self, slot):
        k = {"x": L.OP_RXX, "y": L.OP_RYY, "z": L.OP_RZZ}[self.axis]
        return [_op(k, self.q1, self.q2, slot=slot)]

    def gen_ops(self, tangent, slot, which=1):
        return [_gen(tangent, _PAULI_CODE[self.axis], [self.q1, self.q2], -0.5j)]

    def __repr__(self):
        return f"{type(self).__name__}({self.theta:.2f})@q{self.q1},q{self.q2}"


class R_zz(RR):
    axis = "z"


class R_xx(RR):
    axis = "x"


class R_yy(RR):
    axis = "y"


class RR_block(shared_parameter):
    """gates.py:554-585: rotator on ring pairs (i, (i+1) mod N), commuting."""

    def __init__(self, rotator, q_N):
        self.rotator = rotator
        self.theta = 0
        self.phi = 0
        self.q_N = q_N
        self.is_param = True
        self.param_count = 1
        self.layer = self.gen_layer()
        self.commute = True

    def gen_layer(self):
        N = self.q_N
        return [self.rotator([i, (i + 1) % N], N) for i in range(N)]

    def __repr__(self):
        return f"RR block of {self.layer}"


# ------------------------------------------------------------------------ fSim

def _fsim(t, f):
    return np.array([[1, 0, 0, 0], [0, np.cos(t), -1j * np.sin(t), 0],
                     [0, -1j * np.sin(t), np.cos(t), 0], [0, 0, 0, np.exp(-1j * f)]])


def fsim_gate(theta, phi, N=None, control=0, target=1):
    """gates.py:588-606 (local 4x4; the engine places it on its qubits)."""
    return _fsim(theta, phi)


def fsim_gate_d_theta(theta, phi, N=None, control=0, target=1):
    """gates.py:609-627, as written: the middle block differentiated in
    theta, the corners of fsim_gate kept (local 4x4)."""
    m = _fsim(theta, phi)
    m[1, 1] = m[2, 2] = -np.sin(theta)
    m[1, 2] = m[2, 1] = -1j * np.cos(theta)
    return m


def fsim_gate_d_phi(theta, phi, N=None, control=0, target=1):
    """gates.py:630-648, as written: fsim_gate with its last entry
    differentiated in phi (local 4x4)."""
    m = _fsim(theta, phi)
    m[3, 3] = -1j * np.exp(-1j * phi)
    return m


def fixed_fsim_gate(theta, N=None, control=0, target=1):
    """gates.py:700-716: fsim_gate at phi = 0 (local 4x4)."""
    return _fsim(theta, 0.0)


def fixed_fsim_gate_d_theta(theta, N=None, control=0, target=1):
    """gates.py:719-737, as written (local 4x4)."""
    return fsim_gate_d_theta(theta, 0.0)


class fSim(PRot):
    """gates.py:651-697: two parameters; the derivative states multiply the
    reference's fsim_gate_d_theta / d_phi matrices (gates.py:609-648) into the
    state after the gate."""

    def __init__(self, qs_on, q_N):
        self.q1, self.q2 = qs_on
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = True
        self.param_count = 2

    def prim_ops(self, slot):
        return [_op(L.OP_FSIM, self.q1, self.q2, slot=slot, slot2=slot + 1)]

    def gen_ops(self, tangent, slot, which=1):
        return [_op(L.OP_GEN_FSIM, self.q1, self.q2, slot=slot, slot2=slot + 1, tangent=tangent,
                    pauli=1 if which == 1 else 2)]

    def set_theta(self, theta):
        self.theta = theta

    def set_phi(self, phi):
        self.phi = phi

    def parameterised_derivative(self, param):
        return LocalOperator(self.q_N, self._const_gen(param))

    def _const_gen(self, which):
        o = self.gen_ops(0, 0, which)[0]
        o.offset, o.scale, o.slot = float(self.theta), 0.0, -1
        # phi from a constant: encode as a dense generator-free matrix
        th, ph = float(self.theta), float(self.phi)
        m = fsim_gate_d_theta(th, ph) if which == 1 else fsim_gate_d_phi(th, ph)
        return [_op(L.OP_MAT2, self.q1, self.q2, m=m)]

    def derivative(self):
        return self.parameterised_derivative(1)

    def flip_pauli(self):
        pass

    def __repr__(self):
        return f"{type(self).__name__}({self.theta:.2f},{self.phi:.2f})@q{self.q1, self.q2}"


class fixed_fSim(PRot):
    """gates.py:740-759: one parameter; derivative matrix as written
    (gates.py:719-737)."""

    def __init__(self, qs_on, q_N):
        self.q1, self.q2 = qs_on
        self.q_N = q_N
        self.theta = 0
        self.phi = 0
        self.is_param = True
        self.param_count = 1

    def prim_ops(self, slot):
        return [_op(L.OP_FSIM, self.q1, self.q2, slot=slot, slot2=-1)]

    def gen_ops(self, tangent, slot, which=1):
        return [_op(L.OP_GEN_FSIM, self.q1, self.q2, slot=slot, slot2=-1, tangent=tangent, pauli=3)]

    def derivative(self):
        m = fixed_fsim_gate_d_theta(float(self.theta))
        return LocalOperator(self.q_N, [_op(L.OP_MAT2, self.q1, self.q2, m=m)])

    def flip_pauli(self):
        pass


# ------------------------------------------------------------------ utilities

def genFockOp(op, position, size, levels=2, opdim=0):
    """gates.py:39-42 for Pauli operators: returns a one-term PauliSum."""
    a = np.asarray(op.full() if hasattr(op, "full") else op)
    for code, m in ((1, [[0, 1], [1, 0]]), (2, [[0, -1j], [1j, 0]]), (3, [[1, 0], [0, -1]]),
                    (0, [[1, 0], [0, 1]])):
        if np.allclose(a, m):
            return PauliSum.single(size, position, code)
    raise NotImplementedError("genFockOp supports Pauli operators and the identity")


def iden(N):
    """gates.py:45-46 as a PauliSum."""
    return PauliSum(N, {(0,) * N: 1.0})
