"""CPU FP64 oracle for the statevector engine -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy, the algorithm of the reference library
pyramaterised 1.0.2 (``pyramaterised/{gates,circuit,measure,templates}.py``)
for the hot path: local gate matrices, sequential state evolution, per-parameter
derivative states, the QFIM, Meyer-Wallach entanglement and the expressibility
histogram / KL divergence.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or, for ``cpu_ref``, as the
timed CPU baseline).  The product package ``pyramaterised_amd`` never imports
it: its hot path runs on the HIP library and fails loudly without it.

Pinning (each function's docstring names its own pin):
  * reference known answers: qg-circuit energy 0.46135870050914374 and EQD 12
    (tests.py:192-213), NPQC QFIM == identity (tests.py:103-128), Bell Q == 1
    (tests.py:284-295), fixed_R_y(pi/2)|0> and H.H|0> == |0> (tests.py:64-87);
  * brute force: full 2^N operators built with np.kron at tiny N
    (``sim.dense_unitary``) against the einsum simulator; ALLTOALL against a
    product of CNOT permutation matrices built from bit flips;
  * closed forms and invariants: Bell/product-state purities, unitarity,
    central finite differences of the state for the derivative states, the
    fSim gate's generator form exp(-i theta (XX + YY) / 2) with the phase on
    |11>, sqrt(iSWAP)^2 = iSWAP = exp(i pi/4 (XX + YY)), R_PP(t) =
    exp(-i t/2 P(x)P), the fSim-family derivative matrices (central
    differences of their differentiated blocks; the kept entries are fSim's
    own), I's derivative -i/2 psi and new_measure's capped eigenvalue sum.

Parity unpinned (no reference output and no mathematical fixed point pins
it): ``gates.sqrt_h`` (element-wise square root applied to a QuTiP Qobj at
gates.py:239-242; only QuTiP's np.sqrt-on-Qobj semantics could pin it).
"""
