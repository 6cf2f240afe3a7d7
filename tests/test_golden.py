"""The committed golden fixtures (tests/golden/*.npz) against the reference's
known answers and against the oracle that wrote them (CPU only)."""
import os

import numpy as np
import pytest

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_qg_fixture_holds_reference_known_answers():
    g = load("qg")
    assert abs(float(g["energy"]) - 0.46135870050914374) < 1e-15  # tests.py:207-209
    assert int(g["eqd"]) == 12  # tests.py:210-213
    n, gates = OC.qg_test_circuit()
    assert np.max(np.abs(OS.run(gates, g["params"], n) - g["state"])) < 1e-15


def test_npqc_fixture_is_identity():
    g = load("npqc")
    keys = [k for k in g if k.endswith("_qfim")]
    assert len(keys) == 4 + 8
    for k in keys:
        F = g[k]
        assert np.max(np.abs(F - np.eye(len(F)))) < 1e-12  # tests.py:103-128


def test_mw_fixture_closed_forms():
    g = load("mw")
    assert abs(float(g["bell_Q"]) - 1) < 1e-12  # tests.py:284-295
    assert abs(float(g["product_Q"])) < 1e-12
    # QuTiP's bell_state('11') is the singlet (|01> - |10>)/sqrt(2): antisymmetric
    # under the swap of the two qubits, unit norm
    b = g["bell_state"]
    assert np.allclose(b, np.array([0, 1, -1, 0]) / np.sqrt(2), atol=1e-15)
    swap = b.reshape(2, 2).T.reshape(-1)
    assert np.allclose(swap, -b, atol=1e-15)


def test_package_bell_state_matches_fixture():
    from pyramaterised_amd import qobj
    g = load("mw")
    assert np.max(np.abs(qobj.bell_state("11").full().ravel() - g["bell_state"])) < 1e-15


def test_config1_and_config3_fixtures_reproduce():
    g = load("config1")
    n, p = 4, 3
    for name, og in (("linear", OC.hea_linear(p, n)), ("chain", OC.clifford_he(p, n))):
        psi = OS.run(og, g[f"{name}_params"], n)
        assert np.max(np.abs(psi - g[f"{name}_state"])) < 1e-14
        F = OM.qfim(OS.derivative_states(og, g[f"{name}_params"], n), psi)
        assert np.max(np.abs(F - g[f"{name}_qfim"])) < 1e-12
    g3 = load("config3")
    og, init = OC.tfim_open(10, 12)
    psi = OS.run(og, g3["params"], 12, init)
    assert np.max(np.abs(psi - g3["state"])) < 1e-13
    assert int(g3["eqd"]) == OM.effective_quantum_dimension(g3["qfim"], 1e-12)


def test_config4_fixture_consistent():
    g = load("config4_22q")
    assert abs(float(g["norm"]) - 1) < 1e-12
    assert abs(g["block_prob"].sum() - float(g["norm"])) < 1e-12
    assert len(g["params"]) == 22 * 30
    idx = g["sample_idx"]
    assert np.all(np.diff(idx) > 0) and idx[-1] < 2 ** 22
