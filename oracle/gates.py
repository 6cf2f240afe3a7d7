"""Local gate matrices of the reference's gate library, restated from closed forms.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Conventions (QuTiP 4.7, which the reference calls):
  * a 2-qubit local matrix acts on (first listed qubit) (x) (second listed qubit),
    the first factor being the more significant bit -- ``gate_expand_2toN(U, N,
    control, target)`` at gates.py:595,705 and ``cnot(N, control, target)`` at
    gates.py:329-330;
  * rotations use half angles: R_P(t) = cos(t/2) I - i sin(t/2) P.

Pins: the qg-circuit energy and the NPQC identity in ``tests/test_oracle.py``
exercise rx/ry/rz/cnot/cz; unitarity and the textbook identities
(H = X.Ry(pi/2), H.H = I, S.S = Z, T^2 = S, R_zz diagonal, rotations and
R_PP as matrix exponentials, sqrt(iSWAP)^2 = iSWAP, fSim's generator form)
and central differences of the fSim-family derivative matrices are checked
there too.
"""
import numpy as np

SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)
PAULI = {"x": SX, "y": SY, "z": SZ, "i": I2}


def rx(t):
    """qutip rx (gates.py:161): [[c, -is], [-is, c]], c = cos t/2, s = sin t/2."""
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def ry(t):
    """qutip ry (gates.py:167): [[c, -s], [s, c]]."""
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz(t):
    """qutip rz (gates.py:173): diag(e^{-it/2}, e^{it/2})."""
    return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]], dtype=np.complex128)


ROT = {"x": rx, "y": ry, "z": rz}


def hadamard():
    """H = x_gate * ry(pi/2) (gates.py:226-232)."""
    return SX @ ry(np.pi / 2)


def sqrt_h():
    """sqrtH (gates.py:235-242): np.sqrt applied element-wise to the complex
    H operator.  Element-wise sqrt of I(x)..(x)H(x)..(x)I keeps the identity
    factors (entries 0 and 1), so the local matrix is the element-wise complex
    square root of H.  Parity unpinned."""
    return np.sqrt(hadamard().astype(np.complex128))


def phase(t):
    """qutip phasegate (gates.py:290-291): diag(1, e^{it})."""
    return np.array([[1, 0], [0, np.exp(1j * t)]], dtype=np.complex128)


def s_gate():
    """S = phasegate(pi/2) (gates.py:286-291)."""
    return phase(np.pi / 2)


def t_gate():
    """T = qutip t_gate = phasegate(pi/4) (gates.py:294-300)."""
    return phase(np.pi / 4)


def x_gate():
    """X (gates.py:245-249)."""
    return SX.copy()


CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
SQRT_ISWAP = np.array(
    [[1, 0, 0, 0],
     [0, 1 / np.sqrt(2), 1j / np.sqrt(2), 0],
     [0, 1j / np.sqrt(2), 1 / np.sqrt(2), 0],
     [0, 0, 0, 1]], dtype=np.complex128)


def rpp(t, p):
    """RR family (gates.py:509-517): cos(t/2) I - i sin(t/2) P(x)P, P in x/y/z."""
    P = PAULI[p]
    return np.cos(t / 2) * np.eye(4) - 1j * np.sin(t / 2) * np.kron(P, P)


def fsim(t, f):
    """fsim_gate (gates.py:598-606)."""
    return np.array([[1, 0, 0, 0],
                     [0, np.cos(t), -1j * np.sin(t), 0],
                     [0, -1j * np.sin(t), np.cos(t), 0],
                     [0, 0, 0, np.exp(-1j * f)]], dtype=np.complex128)


def fsim_d_theta(t, f):
    """fsim_gate_d_theta (gates.py:619-627), as written (not dU/dtheta): the
    exchange block differentiated, the corners fSim's.  Pinned by central
    differences of fsim."""
    return np.array([[1, 0, 0, 0],
                     [0, -np.sin(t), -1j * np.cos(t), 0],
                     [0, -1j * np.cos(t), -np.sin(t), 0],
                     [0, 0, 0, np.exp(-1j * f)]], dtype=np.complex128)


def fsim_d_phi(t, f):
    """fsim_gate_d_phi (gates.py:640-648), as written: the |11> entry
    differentiated, the rest fSim's.  Pinned by central differences."""
    return np.array([[1, 0, 0, 0],
                     [0, np.cos(t), -1j * np.sin(t), 0],
                     [0, -1j * np.sin(t), np.cos(t), 0],
                     [0, 0, 0, -1j * np.exp(-1j * f)]], dtype=np.complex128)


def fixed_fsim(t):
    """fixed_fsim_gate (gates.py:708-716)."""
    return fsim(t, 0.0)


def fixed_fsim_d_theta(t):
    """fixed_fsim_gate_d_theta (gates.py:729-737), as written: the exchange
    block differentiated, both corners 1.  Pinned by central differences."""
    m = fsim_d_theta(t, 0.0)
    m[3, 3] = 1
    return m
