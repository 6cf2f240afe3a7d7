"""Stabiliser Renyi entropy ("entropy of magic") and GKP magic on the GPU.

Reference: Measurements.renyi_entropy_fast / gkp_fast (measure.py:318-368)
build two 2^n x 2^n index and sign matrices and take an O(8^n) ``np.dot``.
Here ``pqs_magic`` forms, for each X-part j, conj(c_m) c_{m^j} in LDS and
Walsh-Hadamard transforms it, so every Pauli expectation value is produced in
O(4^n n) work with O(2^n) memory per workgroup.
"""
import numpy as np

from . import engine as E


def renyi_entropy_batch(states, n, alpha=2):
    """M_alpha = log(sum_P (<P>^2/2^n)^alpha) / (1 - alpha) - n log 2 per row."""
    s = E.magic_sum(states, n, alpha).cpu().numpy()
    return 1 / (1 - alpha) * np.log(s) - np.log(2 ** n)


def renyi_entropy(ket, alpha=2):
    return float(renyi_entropy_batch(ket.t, ket.n, alpha)[0])
