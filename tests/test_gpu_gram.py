"""Complex Gram G = A^dag B (pqs_gram, reference measure.py:123-137's
pair overlaps) on the matrix cores: symmetric (A = B: upper tiles and a
conjugate mirror), general and ragged shapes, complex64 operands, operands
over 1 GiB in column blocks, and K below one load-ring step (the vector
kernel) against a float64 torch product (rel-Frobenius 1e-12); all-pairs
fidelities of circuit states against the oracle."""
import numpy as np
import pytest
import torch

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS

pytestmark = pytest.mark.gpu


def _rand(g, rows, dim, dt):
    x = torch.randn((rows, dim), dtype=torch.complex128, device="cuda", generator=g) / dim ** 0.5
    return x.to(dt)


def _ref(a, b):
    return a.to(torch.complex128).conj() @ b.to(torch.complex128).T


@pytest.mark.parametrize("m,nn,dim,same,dtype", [
    (200, 200, 4096, True, "complex128"),
    (70, 45, 2048, False, "complex128"),
    (333, 333, 1024, True, "complex64"),
    (64, 96, 8, False, "complex128"),
    (5, 5, 4, True, "complex128"),
    (2, 3, 65536, False, "complex64"),
])
def test_gram_shapes(m, nn, dim, same, dtype):
    from pyramaterised_amd import engine as E
    g = torch.Generator(device="cuda").manual_seed(m * 7 + nn)
    dt = getattr(torch, dtype)
    a = _rand(g, m, dim, dt)
    b = a if same else _rand(g, nn, dim, dt)
    G = E.gram(a, b)
    R = _ref(a, b)
    assert (torch.linalg.norm(G - R) / torch.linalg.norm(R)).item() < 1e-12
    if same:
        # exactly Hermitian on the matrix-core path (mirrored lower triangle,
        # real diagonal); the vector kernel (K < 8) computes both triangles
        if dim % 8 == 0:
            assert torch.equal(G, G.conj().T)
        else:
            assert torch.allclose(G, G.conj().T, rtol=0, atol=1e-15)


@pytest.mark.parametrize("same", [True, False])
def test_gram_column_blocks(same):
    """Operands above 1 GiB (complex64, 2^20 amplitudes): column blocks."""
    from pyramaterised_amd import engine as E
    g = torch.Generator(device="cuda").manual_seed(3)
    a = _rand(g, 160, 2 ** 20, torch.complex64)
    b = a if same else _rand(g, 40, 2 ** 20, torch.complex64)
    assert (a.numel() + (0 if same else b.numel())) * 8 > 2 ** 30
    G = E.gram(a, b)
    R = _ref(a, b)
    assert (torch.linalg.norm(G - R) / torch.linalg.norm(R)).item() < 1e-12


def test_all_pairs_fidelities_vs_oracle(pyqc_mod):
    """|<psi_a|psi_b>|^2 over combinations(range(S), 2) of S = 150 states of a
    10-qubit HEA (ragged 64 x 32 tiles) against the oracle's pair loop."""
    pyqc = pyqc_mod
    from pyramaterised_amd import templates as T
    n, p, S = 10, 2, 150
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    params = np.random.default_rng(8).random((S, qc.n_true_params)) * 2 * np.pi
    states = qc.run_batch(params)
    og = OC.clifford_he(p, n)
    refs = [OS.run(og, params[s], n) for s in range(S)]
    F = pyqc.measure.Measurements(qc).fidelities_all_pairs(states).cpu().numpy()
    Fref = np.array(OM.pair_fidelities(refs))
    assert F.shape == Fref.shape
    assert np.max(np.abs(F - Fref)) < 1e-13


@pytest.fixture(scope="module")
def pyqc_mod():
    import pyramaterised_amd
    return pyramaterised_amd
