"""Restated reference measures: QFIM, EQD, Meyer-Wallach Q, expressibility.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Pins: NPQC QFIM == identity and qg EQD == 12 (reference known answers), Bell
Q == 1 and product-state Q == 0 (closed forms), the QFIM of a single R_y on
|0> (== 1, textbook), and expressibility of Haar-distributed samples against
the analytic density (see ``tests/test_oracle.py``).
"""
from itertools import combinations

import numpy as np
import scipy.linalg
import scipy.special


def qfim(grads, psi):
    """Measurements.get_QFI (measure.py:33-71):
    F_pq = 4 Re(<d_p psi|d_q psi> - <d_p psi|psi><psi|d_q psi>)."""
    grads = [np.asarray(g) for g in grads]
    P = len(grads)
    single = np.array([np.vdot(psi, g) for g in grads])  # <psi|d_p psi>
    F = np.zeros((P, P))
    for p in range(P):
        for q in range(p, P):
            ov = np.vdot(grads[p], grads[q])
            F[p, q] = 4 * np.real(ov - np.conj(single[p]) * single[q])
    for p in range(P):
        for q in range(p + 1, P):
            F[q, p] = F[p, q]
    return F


def effective_quantum_dimension(F, cutoff):
    """measure.py:77-87: number of eigenvalues above ``cutoff``."""
    ev = scipy.linalg.eigh(F, eigvals_only=True)
    return int(np.sum(ev > cutoff))


def new_measure(F):
    """measure.py:89-99: sum of eigenvalues capped at 1."""
    ev = scipy.linalg.eigh(F, eigvals_only=True)
    return float(sum(1 if v > 1 else v for v in ev))


def reduced_1q(psi, k, n):
    """2x2 reduced density matrix of qubit k (QuTiP ptrace(k) of a ket):
    rho_ab = sum_rest psi[.., a, ..] conj(psi[.., b, ..])."""
    t = np.asarray(psi).reshape(2 ** k, 2, 2 ** (n - k - 1))
    return np.einsum("iaj,ibj->ab", t, np.conj(t))


def single_Q(psi, n):
    """Measurements.single_Q (measure.py:226-237): Q = 2 (1 - mean_k Tr rho_k^2)."""
    s = 0.0
    for k in range(n):
        r = reduced_1q(psi, k, n)
        s += np.real(np.trace(r @ r))
    return 2 * (1 - s / n)


def pair_fidelities(states):
    """_gen_f_samples (measure.py:131-137): |<psi|phi>|^2 over combinations."""
    return [abs(np.vdot(a, b)) ** 2 for a, b in combinations(states, 2)]


def gen_histo(F_samples, bins=None):
    """_gen_histo (measure.py:139-159): np.histogram on (0, 1) with
    int(0.0075 * len(F)) bins unless ``bins`` is given, normalised, midpoints."""
    if bins is None:
        bins = int((75 / 10000) * len(F_samples))
    prob, edges = np.histogram(F_samples, bins=bins, range=(0, 1))
    prob = prob / sum(prob)
    mids = np.array([(edges[i - 1] + edges[i]) / 2 for i in range(1, len(edges))])
    return prob, mids


def expr(F_samples, dim, bins=None):
    """Measurements.expr (measure.py:161-180): sum kl_div(P_pqc, P_haar) with the
    Haar density (d-1)(1-F)^(d-2) at the bin midpoints."""
    if len(F_samples) == 0:
        return 0
    P, F = gen_histo(F_samples, bins)
    haar = (dim - 1) * ((1 - F) ** (dim - 2))
    Ph = haar / sum(haar)
    return float(np.sum(scipy.special.kl_div(P, Ph)))
