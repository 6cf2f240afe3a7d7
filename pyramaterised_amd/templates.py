"""Circuit templates with the reference's names and gate orderings.

Reference: pyramaterised/templates.py.  These builders only assemble gate
lists; each produces the same gates, in the same order and on the same qubits,
as the reference function it names, so that parameter binding and derivative
order (which follow gate order) match.  Builders whose reference draws from
Python's ``random`` module keep that (unseeded) behaviour.
"""
import random

import numpy as np

from .circuit import PQC
from .gates import (CHAIN, CNOT, CPHASE, CZ, H, R_x, R_xx, R_y, R_yy, R_z, R_zz, RR_block, S, X,
                    fixed_fSim, fixed_R_y, fixed_R_z, fSim, negative_R_z, offset_R_z, PRot, EntGate,
                    shared_parameter, sqrtiSWAP)
from .qobj import PauliSum, ProductKet, basis, tensor

LEN_CLIFF_STRING = 3


def clifford_circuit_layers(p, N, method="random"):
    """templates.py:17-48."""
    gates = [H, S, CZ, CNOT]
    qs = list(range(N))
    layers = []
    for _ in range(p):
        layer = []
        if method.lower() == "random":
            for _n in range(N):
                g = random.choice(gates)
                if issubclass(g, PRot):
                    layer.append(g(random.randint(0, N - 1), N))
                elif issubclass(g, EntGate):
                    layer.append(g(random.sample(qs, k=2), N))
        else:
            qs_on = random.sample(qs, k=N // 2)
            strings = [random.choices(gates[:2], k=LEN_CLIFF_STRING) for _q in qs_on]
            for q, string in zip(qs_on, strings):
                layer += [g(q, N) for g in string]
            layer.append(CHAIN(CNOT, N))
        layers.append(layer)
    return layers


def gen_shift_list(p, N):
    """templates.py:51-63: shift sequence a_s of the NPQC entangling layers."""
    todo = list(range(N // 2))
    shifts = np.zeros(2 ** (N // 2), dtype=np.int32)
    s = 1
    while todo:
        shifts[s - 1] = todo.pop(0)
        for q in range(1, s):
            shifts[s + q - 1] = shifts[q - 1]
        s *= 2
    return shifts


def NPQC_layers(p, N):
    """templates.py:66-95: returns (layers, reference angles)."""
    layers = [[R_y(i, N) for i in range(N)] + [R_z(i, N) for i in range(N)]]
    angles = [np.pi / 2] * N + [0] * N
    shifts = gen_shift_list(p, N)
    for i in range(p - 1):
        a = shifts[i]
        even = [2 * k - 2 for k in range(1, 1 + N // 2)]
        layer = [fixed_R_y(q, N, np.pi / 2) for q in even]
        layer += [CPHASE([q, ((q + 1) + 2 * a) % N], N) for q in even]
        for q in even:
            layer += [R_y(q, N), R_z(q, N)]
            angles += [np.pi / 2, 0]
        layers.append(layer)
    return layers, angles


def string_to_entangler(string):
    """templates.py:98-111."""
    table = {"cnot": CNOT, "cphase": CPHASE, "cz": CZ, "sqrtiswap": sqrtiSWAP}
    try:
        return table[string.lower()]
    except KeyError:
        raise Exception("Must supply a valid entangler!") from None


def _repeat(layer, p):
    return [layer for _ in range(p)]


def circuit_1_layers(p, N):
    """templates.py:114-119 (arXiv:1905.10876 circuit 1)."""
    return _repeat([R_x(i, N) for i in range(N)] + [R_y(i, N) for i in range(N)], p)


def circuit_2_layers(p, N, ent_str="cnot"):
    """templates.py:122-128."""
    ent = string_to_entangler(ent_str)
    return _repeat([R_x(i, N) for i in range(N)] + [R_z(i, N) for i in range(N)] + [CHAIN(ent, N)], p)


def circuit_9_layers(p, N, ent_str="cphase"):
    """templates.py:131-137."""
    ent = string_to_entangler(ent_str)
    return _repeat([H(i, N) for i in range(N)] + [CHAIN(ent, N)] + [R_x(i, N) for i in range(N)], p)


def qg_circuit_layers(p, N, ent_str="cnot"):
    """templates.py:140-151 (arXiv:2102.01659)."""
    ent = string_to_entangler(ent_str)
    body = []
    for rot in (R_z, R_x, R_z):
        body += [rot(i, N) for i in range(N)] + [CHAIN(ent, N)]
    return [[fixed_R_y(i, N, np.pi / 4) for i in range(N)]] + _repeat(body, p)


def clifford_HE_layers(p, N, ent_str="cnot"):
    """templates.py:164-171: [R_y all, R_z all, CHAIN] x p."""
    ent = string_to_entangler(ent_str)
    return _repeat([R_y(i, N) for i in range(N)] + [R_z(i, N) for i in range(N)] + [CHAIN(ent, N)], p)


def generic_HE_layers(p, N, ent_str="cnot"):
    """templates.py:154-162: fixed R_y(pi/4) layer, then clifford_HE layers."""
    return [[fixed_R_y(i, N, np.pi / 4) for i in range(N)]] + clifford_HE_layers(p, N, ent_str)


def y_CPHASE_layers(p, N):
    """templates.py:173-180."""
    return _repeat([R_y(i, N) for i in range(N)] + [CHAIN(CPHASE, N)], p)


def double_y_CPHASE_layers(p, N):
    """templates.py:182-189."""
    return _repeat([R_y(i, N) for i in range(N)] + [R_y(i, N) for i in range(N)] + [CHAIN(CPHASE, N)], p)


def TFIM_layers(p, N):
    """templates.py:194-204: H layer, then [RR_block(R_zz), shared R_x] x p."""
    layers = [[H(i, N) for i in range(N)]]
    for _ in range(p):
        layers.append([RR_block(R_zz, N), shared_parameter([R_x(i, N) for i in range(N)], N)])
    return layers


def modified_TFIM_layers(p, N):
    """templates.py:207-217."""
    return [[RR_block(R_zz, N), shared_parameter([R_x(i, N) for i in range(N)], N),
             shared_parameter([R_z(i, N) for i in range(N)], N)] for _ in range(p)]


def TFIM_hamiltonian(N, g, h=0):
    """templates.py:220-227: -sum_i (Z_i Z_{i+1 mod N} + g X_i + h Z_i)."""
    terms = {}

    def add(s, c):
        terms[s] = terms.get(s, 0) + c

    for i in range(N):
        zz = [0] * N
        zz[i] = 3
        zz[(i + 1) % N] = 3 if (i + 1) % N != i else 0
        if (i + 1) % N == i:
            zz[i] = 0  # Z_i Z_i = I for N = 1
        add(tuple(zz), -1.0)
        x = [0] * N
        x[i] = 1
        add(tuple(x), -g)
        z = [0] * N
        z[i] = 3
        add(tuple(z), -h)
    return PauliSum(N, terms)


def XXZ_layers(p, N, commute=False):
    """templates.py:230-257."""
    even = [((2 * i - 1) - 1, 2 * i - 1) for i in range(1, N // 2 + 1)]
    odd = [(2 * i - 1, ((2 * i + 1) - 1) % N) for i in range(1, N // 2 + 1)]
    layers = []
    for _ in range(p):
        layers.append([
            shared_parameter([R_zz(pr, N) for pr in odd], N),
            shared_parameter([R_yy(pr, N) for pr in odd] + [R_xx(pr, N) for pr in odd], N, commute=commute),
            shared_parameter([R_zz(pr, N) for pr in even], N),
            shared_parameter([R_yy(pr, N) for pr in even] + [R_xx(pr, N) for pr in even], N, commute=commute),
        ])
    return layers


def gen_theta_block(q1, q2, N):
    """templates.py:260-274."""
    return [sqrtiSWAP([q1, q2], N),
            shared_parameter([negative_R_z(q1, N), offset_R_z(q2, N, np.pi)], N),
            sqrtiSWAP([q1, q2], N),
            fixed_R_z(q2, N, np.pi)]


def list_to_pairs(x):
    """templates.py:277-279."""
    return [(x[i], x[i + 1]) for i in range(0, len(x) - 1, 2)]


def fermionic_circuit_layers(p, N):
    """templates.py:282-304: diamond of theta blocks."""
    layers = []
    for _ in range(p):
        blocks = []
        for n in range(1, 1 + N // 2):
            left = list(range(N // 2 - n + 1, N // 2 + 1))
            right = list(range(1 + N // 2, 1 + N // 2 + n))
            blocks.append(list_to_pairs(left + right))
        blocks = blocks + list(reversed(blocks))[1:]
        for b in blocks:
            layer = []
            for a, c in b:
                layer += gen_theta_block(a - 1, c - 1, N)
            layers.append(layer)
    return layers


def fSim_circuit_layers(p, N, rotator="y", fixed=False):
    """templates.py:307-353."""
    rot = {"y": R_y, "x": R_x, "z": R_z}.get(rotator.lower())
    if rot is None:
        raise Exception("Please supply a valid single qubit rotator")
    ent = fixed_fSim if fixed else fSim
    layers = []
    for l in range(p):
        layer = [rot(i, N) for i in range(N)]
        if N % 2 == 0:
            layer += [ent([i, (i + 1) % N], N) for i in range(l % 2, N, 2)]
        else:
            offset = l % N
            idx = [i for i in range(N) if i != offset]
            pairs = []
            if offset % 2 == 1:
                pairs.append((idx.pop(0), idx.pop(-1)))
            pairs += [(idx[i], idx[i + 1]) for i in range(0, len(idx), 2)]
            layer += [ent([a, b], N) for a, b in pairs]
            layer.append(rot(offset, N))
        layers.append(layer)
    return layers


def add_layers(circuit, layers):
    """templates.py:355-358."""
    for l in layers:
        circuit.add_layer(l)
    return circuit


def _half_filled(N, shuffle):
    st = [basis(2, 1) for _ in range(N // 2)] + [basis(2, 0) for _ in range(N // 2, N)]
    if shuffle:
        random.shuffle(st)
    return tensor(st)


def generate_circuit(circuit_type, N, p, hamiltonian="ZZ", rotator="", shuffle=True, **pqc_kwargs):
    """templates.py:361-429."""
    circuit = PQC(N, **pqc_kwargs)
    if circuit_type == "NPQC":
        layers, _ = NPQC_layers(p, N)
    elif circuit_type == "TFIM":
        layers = TFIM_layers(p, N)
    elif circuit_type == "TFIM_modified":
        layers = modified_TFIM_layers(p, N)
    elif circuit_type == "XXZ":
        layers = XXZ_layers(p, N)
        circuit.initial_state = _half_filled(N, shuffle)
    elif circuit_type == "Circuit_1":
        layers = circuit_1_layers(p, N)
    elif circuit_type == "Circuit_2":
        layers = circuit_2_layers(p, N)
    elif circuit_type == "Circuit_9":
        layers = circuit_9_layers(p, N)
    elif circuit_type == "qg_circuit":
        layers = qg_circuit_layers(p, N)
    elif circuit_type == "generic_HE":
        layers = generic_HE_layers(p, N)
    elif circuit_type == "clifford":
        layers = clifford_HE_layers(p, N)
    elif circuit_type == "y_CPHASE":
        layers = y_CPHASE_layers(p, N)
    elif circuit_type == "double_y_CPHASE":
        layers = double_y_CPHASE_layers(p, N)
    elif circuit_type == "fermionic":
        layers = fermionic_circuit_layers(p, N)
        circuit.initial_state = _half_filled(N, shuffle)
    elif circuit_type == "zfsim":
        layers = fSim_circuit_layers(p, N, rotator="z")
        circuit.initial_state = _half_filled(N, shuffle)
    elif circuit_type == "fsim":
        layers = fSim_circuit_layers(p, N, rotator=rotator) if rotator in ["x", "y", "z"] else \
            fSim_circuit_layers(p, N)
        circuit.initial_state = _half_filled(N, shuffle)
    elif circuit_type == "fixed_fsim":
        layers = fSim_circuit_layers(p, N, rotator="z", fixed=True)
        circuit.initial_state = _half_filled(N, shuffle)
    for l in layers:
        circuit.add_layer(l)
    return circuit


# ---------------------------------------------------------------- workloads
# Circuits of the benchmark workloads (not in the reference), built from the
# reference's gate classes.

def hea_linear_layers(p, N):
    """Hardware-efficient ansatz with an explicit sequential CNOT chain
    0->1->...->N-1 after R_y and R_z on every qubit."""
    return _repeat([R_y(i, N) for i in range(N)] + [R_z(i, N) for i in range(N)] +
                   [CNOT([i, i + 1], N) for i in range(N - 1)], p)


def tfim_open_layers(p, N):
    """HVA on an open chain: shared-angle R_zz on (i, i+1), i < N-1, then a
    shared-angle R_x on every qubit; use with a |+>^N initial state."""
    return [[shared_parameter([R_zz([i, i + 1], N) for i in range(N - 1)], N),
             shared_parameter([R_x(i, N) for i in range(N)], N)] for _ in range(p)]


def ry_cz_brickwork_layers(p, N):
    """R_y on every qubit, then CZ on (i, i+1) for even i in even layers and
    odd i in odd layers."""
    return [[R_y(i, N) for i in range(N)] + [CZ([i, i + 1], N) for i in range(l % 2, N - 1, 2)]
            for l in range(p)]


def plus_state():
    return (basis(2, 0) + basis(2, 1)) / np.sqrt(2)
