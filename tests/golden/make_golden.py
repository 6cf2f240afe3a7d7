"""Write the golden fixtures of the hot path from the FP64 oracle.

    python tests/golden/make_golden.py

Every fixture is data (inputs and expected outputs) produced by ``oracle/``,
which is itself pinned by the reference's known answers (qg energy and EQD,
tests.py:192-213; NPQC QFIM = identity, tests.py:103-128; Bell Q = 1,
tests.py:284-295) -- see tests/test_oracle.py.  The reference cannot run here
(QuTiP is absent), so no fixture comes from it directly.

Fixtures (numpy .npz, no pickles):
  qg.npz        4-qubit qg circuit (tests.py:158-190) at default_rng(1) draws:
                state, <Z0 Z1>, QFIM, EQD
  config1.npz   4 q x 3 layers of (R_y, R_z; CNOT 0->1->2->3), seed 1: state,
                Meyer-Wallach Q, 24x24 QFIM; and the CHAIN variant, seed 2
  npqc.npz      NPQC QFIMs at the reference angles for N = 4, 6 (all P)
  config3.npz   12 q TFIM HVA, |+>^12, seed 3: state, 20x20 QFIM, EQD
  config4_22q.npz  22 q R_y + brickwork CZ x 30, seed 4: norm, amplitude sum,
                index-weighted probability, probability of each of the 64
                top-6-bit blocks, and 2048 sampled amplitudes
  mw.npz        Bell(11) and a random product state: Q
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import circuits as OC  # noqa: E402
from oracle import measures as OM  # noqa: E402
from oracle import sim as OS  # noqa: E402


def zz_energy(psi, n):
    """<psi| Z_0 Z_1 |psi> (circuit.py:25-33 default H)."""
    idx = np.arange(2 ** n)
    z0 = 1 - 2 * ((idx >> (n - 1)) & 1)
    z1 = 1 - 2 * ((idx >> (n - 2)) & 1)
    return float(np.sum(z0 * z1 * np.abs(psi) ** 2))


def qg():
    n, gates = OC.qg_test_circuit()
    r = np.random.default_rng(1)
    th = np.array([r.random(1)[0] * 2 * np.pi for _ in range(12)])
    psi = OS.run(gates, th, n)
    F = OM.qfim(OS.derivative_states(gates, th, n), psi)
    return dict(params=th, state=psi, energy=zz_energy(psi, n), qfim=F,
                eqd=OM.effective_quantum_dimension(F, 1e-12))


def config1():
    n, p = 4, 3
    out = {}
    for name, og, seed in (("linear", OC.hea_linear(p, n), 1), ("chain", OC.clifford_he(p, n), 2)):
        th = np.random.default_rng(seed).random(OS.n_params(og)) * 2 * np.pi
        psi = OS.run(og, th, n)
        out[f"{name}_params"] = th
        out[f"{name}_state"] = psi
        out[f"{name}_Q"] = OM.single_Q(psi, n)
        out[f"{name}_qfim"] = OM.qfim(OS.derivative_states(og, th, n), psi)
    return out


def npqc():
    out = {}
    for n in (4, 6):
        for P in range(1, 2 ** (n // 2) + 1):
            og, ang = OC.npqc(P, n)
            ang = np.asarray(ang)
            psi = OS.run(og, ang, n)
            out[f"n{n}_P{P}_angles"] = ang
            out[f"n{n}_P{P}_qfim"] = OM.qfim(OS.derivative_states(og, ang, n), psi)
    return out


def config3():
    n, p = 12, 10
    og, init = OC.tfim_open(p, n)
    th = np.random.default_rng(3).random(20) * 2 * np.pi
    psi = OS.run(og, th, n, init)
    F = OM.qfim(OS.derivative_states(og, th, n, init), psi)
    return dict(params=th, state=psi, qfim=F, eqd=OM.effective_quantum_dimension(F, 1e-12))


def config4_summary(psi, n):
    """Size-independent summary of a large state (also used by the tests)."""
    prob = np.abs(psi) ** 2
    idx = np.arange(psi.size, dtype=np.float64)
    blocks = prob.reshape(64, -1).sum(axis=1)
    return dict(norm=float(prob.sum()), amp_sum=complex(psi.sum()),
                index_weighted=float((idx * prob).sum() / psi.size), block_prob=blocks)


def config4():
    n, p = 22, 30
    og = OC.ry_cz_brickwork(p, n)
    th = np.random.default_rng(4).random(OS.n_params(og)) * 2 * np.pi
    psi = OS.run(og, th, n)
    s = config4_summary(psi, n)
    sample = np.sort(np.random.default_rng(44).choice(psi.size, 2048, replace=False))
    return dict(params=th, sample_idx=sample, sample_amp=psi[sample], norm=s["norm"],
                amp_sum=s["amp_sum"], index_weighted=s["index_weighted"], block_prob=s["block_prob"])


def mw():
    bell = np.zeros(4, dtype=np.complex128)
    bell[1], bell[2] = 1 / np.sqrt(2), -1 / np.sqrt(2)  # QuTiP bell_state('11') = (|01> - |10>)/sqrt(2)
    rng = np.random.default_rng(9)
    singles = rng.normal(size=(5, 2)) + 1j * rng.normal(size=(5, 2))
    singles /= np.linalg.norm(singles, axis=1, keepdims=True)
    prod = singles[0]
    for s in singles[1:]:
        prod = np.kron(prod, s)
    return dict(bell_state=bell, bell_Q=OM.single_Q(bell, 2), product_singles=singles,
                product_Q=OM.single_Q(prod, 5))


def main(names=None):
    for name, fn in (("qg", qg), ("config1", config1), ("npqc", npqc), ("config3", config3),
                     ("config4_22q", config4), ("mw", mw)):
        if names and name not in names:
            continue
        d = fn()
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **{k: np.asarray(v) for k, v in d.items()})
        print("wrote", name, flush=True)


if __name__ == "__main__":
    import sys
    main(sys.argv[1:])
