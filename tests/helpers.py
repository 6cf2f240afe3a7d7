"""Shared test helpers: matching package / oracle circuits from one spec."""
import numpy as np

from oracle import circuits as OC

# spec entries: (name, qubits-or-pair, extra)
ONE_Q_PARAM = ["R_x", "R_y", "R_z", "negative_R_z"]
TWO_Q_PARAM = ["R_xx", "R_yy", "R_zz"]
FIXED_1Q = ["H", "X", "S", "T", "sqrtH"]
FIXED_2Q = ["CNOT", "CZ", "sqrtiSWAP"]


def random_spec(n, depth, rng, fsim=True):
    spec = []
    for _ in range(depth):
        for q in range(n):
            spec.append((ONE_Q_PARAM[rng.integers(len(ONE_Q_PARAM))], q, None))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        spec.append((FIXED_2Q[rng.integers(len(FIXED_2Q))], (a, b), None))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        spec.append((TWO_Q_PARAM[rng.integers(len(TWO_Q_PARAM))], (a, b), None))
        spec.append((FIXED_1Q[rng.integers(len(FIXED_1Q))], int(rng.integers(n)), None))
        spec.append(("offset_R_z", int(rng.integers(n)), 0.37))
        if fsim:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            spec.append(("fSim", (a, b), None))
    spec.append(("CHAIN", None, "CNOT"))
    spec.append(("shared_Rx", None, None))
    spec.append(("fixed_R_y", int(rng.integers(n)), 0.9))
    return spec


def build_pkg(spec, n, pyqc):
    gates = []
    for name, q, extra in spec:
        if name in ONE_Q_PARAM or name in FIXED_1Q:
            gates.append(getattr(pyqc, name)(q, n))
        elif name == "offset_R_z":
            gates.append(pyqc.offset_R_z(q, n, extra))
        elif name == "fixed_R_y":
            gates.append(pyqc.fixed_R_y(q, n, extra))
        elif name in TWO_Q_PARAM or name in FIXED_2Q or name == "fSim":
            gates.append(getattr(pyqc, name)(list(q), n))
        elif name == "CHAIN":
            gates.append(pyqc.CHAIN(getattr(pyqc, extra), n))
        elif name == "shared_Rx":
            gates.append(pyqc.shared_parameter([pyqc.R_x(i, n) for i in range(n)], n))
        else:
            raise ValueError(name)
    return gates


def build_oracle(spec, n):
    gates = []
    for name, q, extra in spec:
        if name in ONE_Q_PARAM or name in FIXED_1Q:
            gates.append(getattr(OC, name)(q, n))
        elif name == "offset_R_z":
            gates.append(OC.offset_R_z(q, n, extra))
        elif name == "fixed_R_y":
            gates.append(OC.fixed_R_y(q, n, extra))
        elif name in TWO_Q_PARAM or name in FIXED_2Q or name == "fSim":
            gates.append(getattr(OC, name)(tuple(q), n))
        elif name == "CHAIN":
            gates.append(OC.CHAIN(getattr(OC, extra), n))
        elif name == "shared_Rx":
            gates.append(OC.shared_parameter([OC.R_x(i, n) for i in range(n)], n))
        else:
            raise ValueError(name)
    return gates


def infidelity(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return 1 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def rel_fro(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def sre_bruteforce(psi, n, alpha):
    """Stabiliser Renyi entropy of order alpha by explicit Pauli strings:
    1/(1-alpha) log sum_P (<P>^2 / d)^alpha - log d."""
    import itertools
    P = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.diag([1, -1])]
    d = 2 ** n
    s = 0.0
    for codes in itertools.product(range(4), repeat=n):
        m = np.array([[1.0 + 0j]])
        for c in codes:
            m = np.kron(m, P[c])
        e = np.vdot(psi, m @ psi).real
        s += (e * e / d) ** alpha
    return 1 / (1 - alpha) * np.log(s) - np.log(d)
