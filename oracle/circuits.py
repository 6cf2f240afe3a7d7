"""Restated reference gates and circuit templates as ``OGate`` lists.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each constructor follows the reference class it names (file:line); derivative
conventions follow the reference's ``derivative()`` methods:
  * PRot family: D = -i/2 * sigma on q_on (gates.py:133-138);
  * negative_R_z: theta -> -theta, D = +i/2 Z (gates.py:180-187);
  * offset_R_z: theta -> theta + offset, D as R_z (gates.py:190-205);
  * RR family: D = -i/2 P(x)P (gates.py:519-522);
  * shared_parameter(commute=True): D = sum of the sub-gates' D (gates.py:451-457);
  * shared_parameter(commute=False): D = sum_k prod(layer with k-th -> D_k g_k)
    times conj(U) (gates.py:458-466) -- restated with dense operators, tiny N only;
  * I: consumes a parameter, U = identity, D = -i/2 identity (gates.py:150-156
    with PRot.derivative; valid for N = 1, where the reference's fock is 2x2).
"""
from itertools import permutations

import numpy as np

from . import gates as G
from .sim import OGate, embed, product_state


def _rot_gate(axis, q, n, sign=1.0, offset=0.0, name=None):
    """R_x/R_y/R_z (gates.py:159-174) with theta -> sign*theta + offset."""
    f = G.ROT[axis]
    P = G.PAULI[axis]

    def prims(t, _p):
        return [((q,), f(sign * t + offset))]

    def deriv(_t, _p, _w):
        return [(-0.5j * sign, [((q,), P)])]

    return OGate(name or f"R_{axis}@{q}", 1, prims, deriv)


def R_x(q, n):
    return _rot_gate("x", q, n)


def R_y(q, n):
    return _rot_gate("y", q, n)


def R_z(q, n):
    return _rot_gate("z", q, n)


def negative_R_z(q, n):
    return _rot_gate("z", q, n, sign=-1.0, name=f"negative_R_z@{q}")


def offset_R_z(q, n, offset):
    return _rot_gate("z", q, n, offset=offset, name=f"offset_R_z@{q}")


def I(q, n):
    """I (gates.py:150-156)."""
    return OGate(f"I@{q}", 1, lambda t, p: [], lambda t, p, w: [(-0.5j, [])])


def fixed(m, q, name="fixed"):
    """Fixed one-qubit gate (is_param False, param_count 0)."""
    m = np.asarray(m, dtype=np.complex128)
    return OGate(f"{name}@{q}", 0, lambda t, p: [((q,), m)])


def H(q, n):
    return fixed(G.hadamard(), q, "H")


def sqrtH(q, n):
    return fixed(G.sqrt_h(), q, "sqrtH")


def X(q, n):
    return fixed(G.x_gate(), q, "X")


def S(q, n):
    return fixed(G.s_gate(), q, "S")


def T(q, n):
    return fixed(G.t_gate(), q, "T")


def fixed_R_y(q, n, theta):
    return fixed(G.ry(theta), q, "fixed_R_y")


def fixed_R_z(q, n, theta):
    return fixed(G.rz(theta), q, "fixed_R_z")


def ent(m, a, b, name):
    m = np.asarray(m, dtype=np.complex128)
    return OGate(f"{name}@{a},{b}", 0, lambda t, p: [((a, b), m)])


def CNOT(pair, n):
    return ent(G.CNOT, pair[0], pair[1], "CNOT")


def CZ(pair, n):
    return ent(G.CZ, pair[0], pair[1], "CZ")


CPHASE = CZ


def sqrtiSWAP(pair, n):
    return ent(G.SQRT_ISWAP, pair[0], pair[1], "sqrtiSWAP")


def chain_pairs(n):
    """CHAIN index order (gates.py:366-376): top pairs (2j, 2j+1), then bottom
    pairs (2j+1, 2j+2); prod(list[::-1]) applies the first listed pair first."""
    top = [(2 * j, 2 * j + 1) for j in range(n // 2)]
    bottom = [(2 * j + 1, 2 * j + 2) for j in range((n - 1) // 2)]
    return top + bottom


def CHAIN(entangler, n):
    """CHAIN (gates.py:355-379) as one gate whose primitives are its entanglers."""
    subs = [entangler(p, n) for p in chain_pairs(n)]
    return OGate("CHAIN", 0, lambda t, p: [pr for s in subs for pr in s.prims()])


def ALLTOALL(entangler, n):
    """ALLTOALL (gates.py:382-401): all ordered pairs of permutations(range(n), 2)."""
    subs = [entangler(p, n) for p in permutations(range(n), 2)]
    return OGate("ALLTOALL", 0, lambda t, p: [pr for s in subs for pr in s.prims()])


def _rr_gate(axis, pair, n):
    a, b = pair
    P = G.PAULI[axis]
    PP = np.kron(P, P)

    def prims(t, _p):
        return [((a, b), G.rpp(t, axis))]

    def deriv(_t, _p, _w):
        return [(-0.5j, [((a, b), PP)])]

    return OGate(f"R_{axis}{axis}@{a},{b}", 1, prims, deriv)


def R_zz(pair, n):
    return _rr_gate("z", pair, n)


def R_xx(pair, n):
    return _rr_gate("x", pair, n)


def R_yy(pair, n):
    return _rr_gate("y", pair, n)


def shared_parameter(layer, n, commute=True):
    """shared_parameter (gates.py:441-484): one angle on every sub-gate,
    U = product of sub-gates in list order."""

    def prims(t, _p):
        return [pr for g in layer for pr in g.prims(t, 0.0)]

    def deriv(t, _p, _w):
        if commute:
            return [term for g in layer for term in g.deriv(t, 0.0, 1)]
        # gates.py:458-466: sum_k prod(new_layer[::-1]) with new_layer[k] =
        # D_k * g_k, then multiplied by operation.conj() (element-wise conjugate
        # of the full block operator).  take_derivative then multiplies by U.
        u_full = np.eye(2 ** n, dtype=np.complex128)
        for qs, m in prims(t, 0.0):
            u_full = embed(m, qs, n) @ u_full
        conj_u = np.conj(u_full)
        terms = []
        all_q = tuple(range(n))
        for k, gk in enumerate(layer):
            for coef, dprims in gk.deriv(t, 0.0, 1):
                # operator applied to the state after the block:
                #   [g_n .. (D_k g_k) .. g_1] conj(U) U  then undo the trailing U
                # is expressed as: psi_after_block -> X psi_after_block with
                # X = [..] conj(U) (the caller applies U first, then X).
                op = np.eye(2 ** n, dtype=np.complex128)
                for j, gj in enumerate(layer):
                    for qs, m in gj.prims(t, 0.0):
                        op = embed(m, qs, n) @ op
                    if j == k:
                        for qs, m in dprims:
                            op = embed(m, qs, n) @ op
                terms.append((coef, [(all_q, op @ conj_u)]))
        return terms

    return OGate(f"shared[{len(layer)}]", 1, prims, deriv)


def RR_block(axis, n):
    """RR_block (gates.py:554-585): ring pairs (i, (i+1) mod n), commuting."""
    layer = [_rr_gate(axis, (i, (i + 1) % n), n) for i in range(n)]
    return shared_parameter(layer, n, commute=True)


def fSim(pair, n):
    """fSim (gates.py:651-697): theta, phi; parameterised_derivative returns the
    matrices of gates.py:609-648 (multiplied into the state after U)."""
    a, b = pair

    def prims(t, p):
        return [((a, b), G.fsim(t, p))]

    def deriv(t, p, which):
        m = G.fsim_d_theta(t, p) if which == 1 else G.fsim_d_phi(t, p)
        return [(1.0, [((a, b), m)])]

    return OGate(f"fSim@{a},{b}", 2, prims, deriv)


def fixed_fSim(pair, n):
    """fixed_fSim (gates.py:740-759): one parameter; derivative matrix as written."""
    a, b = pair

    def prims(t, _p):
        return [((a, b), G.fixed_fsim(t))]

    def deriv(t, _p, _w):
        return [(1.0, [((a, b), G.fixed_fsim_d_theta(t))])]

    return OGate(f"fixed_fSim@{a},{b}", 1, prims, deriv)


# ------------------------------------------------------------------ templates

def qg_test_circuit():
    """The 4-qubit circuit of tests.py:158-190 (flattened in add_layer order)."""
    n = 4
    gates = [fixed_R_y(i, n, np.pi / 4) for i in range(n)]
    gates += [R_z(0, n), R_x(1, n), R_y(2, n), R_z(3, n), CHAIN(CNOT, n)]
    gates += [R_x(0, n), R_x(1, n), R_x(2, n), R_y(3, n), CHAIN(CNOT, n)]
    gates += [R_z(0, n), R_x(1, n), R_y(2, n), R_y(3, n), CHAIN(CNOT, n)]
    return n, gates


def gen_shift_list(p, n):
    """templates.py:51-63."""
    A = list(range(n // 2))
    s = 1
    shift = np.zeros(2 ** (n // 2), dtype=np.int64)
    while A:
        r = A.pop(0)
        shift[s - 1] = r
        for q in range(1, s):
            shift[s + q - 1] = shift[q - 1]
        s *= 2
    return shift


def npqc(p, n):
    """NPQC_layers (templates.py:66-95): returns (gates, reference angles)."""
    gates = [R_y(i, n) for i in range(n)] + [R_z(i, n) for i in range(n)]
    angles = [np.pi / 2] * n + [0.0] * n
    shift = gen_shift_list(p, n)
    for i in range(p - 1):
        a_l = int(shift[i])
        rots, cph = [], []
        for k in range(1, 1 + n // 2):
            q = 2 * k - 2
            rots.append(fixed_R_y(q, n, np.pi / 2))
            cph.append(CZ((q, ((q + 1) + 2 * a_l) % n), n))
        layer = rots + cph
        for k in range(1, n // 2 + 1):
            q = 2 * k - 2
            layer += [R_y(q, n), R_z(q, n)]
            angles += [np.pi / 2, 0.0]
        gates += layer
    return gates, angles


def clifford_he(p, n):
    """clifford_HE_layers (templates.py:164-171): [Ry all, Rz all, CHAIN(CNOT)] x p."""
    gates = []
    for _ in range(p):
        gates += [R_y(i, n) for i in range(n)] + [R_z(i, n) for i in range(n)]
        gates += [CHAIN(CNOT, n)]
    return gates


def generic_he(p, n):
    """generic_HE_layers (templates.py:154-162)."""
    gates = [fixed_R_y(i, n, np.pi / 4) for i in range(n)]
    return gates + clifford_he(p, n)


def tfim(p, n):
    """TFIM_layers (templates.py:194-204): H layer, then p x [RR_block(R_zz),
    shared R_x]."""
    gates = [H(i, n) for i in range(n)]
    for _ in range(p):
        gates += [RR_block("z", n), shared_parameter([R_x(i, n) for i in range(n)], n)]
    return gates


def hea_linear(p, n):
    """Hardware-efficient ansatz with an explicit sequential CNOT chain
    0->1->...->n-1 after R_y, R_z on every qubit, p layers (the 4-qubit
    3-layer case has 24 parameters)."""
    gates = []
    for _ in range(p):
        gates += [R_y(i, n) for i in range(n)] + [R_z(i, n) for i in range(n)]
        gates += [CNOT((i, i + 1), n) for i in range(n - 1)]
    return gates


def tfim_open(p, n):
    """Hamiltonian-variational TFIM ansatz on an open chain: per layer one shared
    angle on R_zz over pairs (i, i+1), i < n-1, then one shared angle on R_x over
    all qubits; start state |+>^n."""
    gates = []
    for _ in range(p):
        gates += [shared_parameter([R_zz((i, i + 1), n) for i in range(n - 1)], n)]
        gates += [shared_parameter([R_x(i, n) for i in range(n)], n)]
    init = product_state(np.array([1, 1]) / np.sqrt(2), n)
    return gates, init


def ry_cz_brickwork(p, n):
    """p layers of (R_y on every qubit, then CZ on (i, i+1) for i even in even
    layers and i odd in odd layers)."""
    gates = []
    for layer in range(p):
        gates += [R_y(i, n) for i in range(n)]
        gates += [CZ((i, i + 1), n) for i in range(layer % 2, n - 1, 2)]
    return gates


def xxz(p, n, commute=False):
    """XXZ_layers (templates.py:230-257): per layer four shared blocks --
    R_zz on the odd links, [R_yy, R_xx] on the odd links (commute as given),
    R_zz on the even links, [R_yy, R_xx] on the even links.  Even links are
    (2i-2, 2i-1), odd links (2i-1, 2i mod n), i = 1..n//2."""
    even = [(2 * i - 2, 2 * i - 1) for i in range(1, n // 2 + 1)]
    odd = [(2 * i - 1, (2 * i) % n) for i in range(1, n // 2 + 1)]
    gates = []
    for _ in range(p):
        gates.append(shared_parameter([R_zz(pr, n) for pr in odd], n))
        gates.append(shared_parameter([R_yy(pr, n) for pr in odd] + [R_xx(pr, n) for pr in odd], n, commute))
        gates.append(shared_parameter([R_zz(pr, n) for pr in even], n))
        gates.append(shared_parameter([R_yy(pr, n) for pr in even] + [R_xx(pr, n) for pr in even], n, commute))
    return gates


# --------------------------------------------------- generate_circuit types
# Restated from templates.py:97-353 (gate order is the parameter-binding
# order, so it is the contract); tests build the package's circuits with
# generate_circuit (templates.py:361-429) and compare states.

def _ent(name):
    return {"cnot": CNOT, "cphase": CZ, "cz": CZ, "sqrtiswap": sqrtiSWAP}[name]


def circuit_1(p, n):
    """circuit_1_layers (templates.py:112-117): R_x all, R_y all, per layer."""
    return [g for _ in range(p) for g in [R_x(i, n) for i in range(n)] + [R_y(i, n) for i in range(n)]]


def circuit_2(p, n, ent="cnot"):
    """circuit_2_layers (templates.py:120-126): R_x all, R_z all, CHAIN."""
    return [g for _ in range(p)
            for g in [R_x(i, n) for i in range(n)] + [R_z(i, n) for i in range(n)] + [CHAIN(_ent(ent), n)]]


def circuit_9(p, n, ent="cphase"):
    """circuit_9_layers (templates.py:129-135): H all, CHAIN, R_x all."""
    return [g for _ in range(p)
            for g in [H(i, n) for i in range(n)] + [CHAIN(_ent(ent), n)] + [R_x(i, n) for i in range(n)]]


def qg_circuit(p, n, ent="cnot"):
    """qg_circuit_layers (templates.py:138-150): fixed R_y(pi/4) all, then p x
    [R_z all, CHAIN, R_x all, CHAIN, R_z all, CHAIN]."""
    gates = [fixed_R_y(i, n, np.pi / 4) for i in range(n)]
    for _ in range(p):
        for rot in (R_z, R_x, R_z):
            gates += [rot(i, n) for i in range(n)] + [CHAIN(_ent(ent), n)]
    return gates


def y_cphase(p, n):
    """y_CPHASE_layers (templates.py:173-180)."""
    return [g for _ in range(p) for g in [R_y(i, n) for i in range(n)] + [CHAIN(CZ, n)]]


def double_y_cphase(p, n):
    """double_y_CPHASE_layers (templates.py:182-189): R_y all twice, CHAIN(CPHASE)."""
    return [g for _ in range(p)
            for g in [R_y(i, n) for i in range(n)] + [R_y(i, n) for i in range(n)] + [CHAIN(CZ, n)]]


def modified_tfim(p, n):
    """modified_TFIM_layers (templates.py:207-217): RR_block(R_zz), shared R_x,
    shared R_z, per layer."""
    gates = []
    for _ in range(p):
        gates += [RR_block("z", n), shared_parameter([R_x(i, n) for i in range(n)], n),
                  shared_parameter([R_z(i, n) for i in range(n)], n)]
    return gates


def theta_block(a, b, n):
    """gen_theta_block (templates.py:260-274): sqrtiSWAP, shared
    [negative_R_z(a), offset_R_z(b, pi)], sqrtiSWAP, fixed_R_z(b, pi)."""
    return [sqrtiSWAP((a, b), n), shared_parameter([negative_R_z(a, n), offset_R_z(b, n, np.pi)], n),
            sqrtiSWAP((a, b), n), fixed_R_z(b, n, np.pi)]


def fermionic(p, n):
    """fermionic_circuit_layers (templates.py:282-304): per repetition a
    diamond of theta-block layers; layer m (1-based, m = 1..n//2, then back
    down to 1) pairs the qubits n//2-m .. n//2+m-1 (0-based) in neighbours."""
    gates = []
    widths = list(range(1, n // 2 + 1))
    widths = widths + widths[::-1][1:]
    for _ in range(p):
        for m in widths:
            qs = list(range(n // 2 - m, n // 2 + m))  # 0-based qubits of this layer
            for i in range(0, len(qs) - 1, 2):
                gates += theta_block(qs[i], qs[i + 1], n)
    return gates


def fsim_circuit(p, n, rotator="y", fixed_gate=False):
    """fSim_circuit_layers (templates.py:307-353): rotations on every qubit,
    then fSim on (i, i+1 mod n) for i of the layer's parity (even n), or, for
    odd n, one qubit left out per layer (index l mod n), pairs of the rest in
    order, the first and last paired when the left-out index is odd, and a
    closing rotation on the left-out qubit."""
    rot = {"x": R_x, "y": R_y, "z": R_z}[rotator]
    ent = fixed_fSim if fixed_gate else fSim
    gates = []
    for l in range(p):
        gates += [rot(i, n) for i in range(n)]
        if n % 2 == 0:
            gates += [ent((i, (i + 1) % n), n) for i in range(l % 2, n, 2)]
        else:
            skip = l % n
            rest = [i for i in range(n) if i != skip]
            pairs = [(rest[0], rest[-1])] if skip % 2 == 1 else []
            inner = rest[1:-1] if skip % 2 == 1 else rest
            pairs += [(inner[i], inner[i + 1]) for i in range(0, len(inner) - 1, 2)]
            gates += [ent(pr, n) for pr in pairs]
            gates.append(rot(skip, n))
    return gates


def half_filled(n):
    """The fermionic / XXZ / fSim start state of generate_circuit without
    shuffling (templates.py:377-380): |1> on the first n//2 qubits, |0> on
    the rest."""
    psi = np.array([1.0 + 0j])
    for q in range(n):
        psi = np.kron(psi, np.array([0, 1]) if q < n // 2 else np.array([1, 0]))
    return psi


def ARBGATE(H, n):
    """ARBGATE (gates.py:407-435): U = expm(-i theta H) over all n qubits
    (scipy.linalg.expm), and the reference's derivative D = -i H / 2 (half
    of dU/dtheta, as written), applied after U."""
    import scipy.linalg
    H = np.asarray(H, dtype=np.complex128)
    qs = tuple(range(n))

    def prims(t, _p):
        return [(qs, scipy.linalg.expm(-1j * t * H))]

    def deriv(_t, _p, _w):
        return [(-0.5j, [(qs, H)])]

    return OGate("ARBGATE", 1, prims, deriv)
