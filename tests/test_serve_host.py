"""Host logic of the serving loop (no GPU): workload generation."""
import numpy as np
import pytest

from pyramaterised_amd import serve as SV


def test_workload_is_seeded_and_in_range():
    a = SV.make_workload(2000, rate=500.0, seed=3)
    b = SV.make_workload(2000, rate=500.0, seed=3)
    assert [r.n for r in a] == [r.n for r in b]
    assert all(np.array_equal(x.params, y.params) for x, y in zip(a, b))
    ns = np.array([r.n for r in a])
    ds = np.array([r.depth for r in a])
    assert ns.min() == 6 and ns.max() == 20
    assert ds.min() == 2 and ds.max() == 40
    frac = np.mean([r.kind == SV.QFIM for r in a])
    assert 0.26 < frac < 0.34
    for r in a[:50]:
        assert r.params.size == 2 * r.n * r.depth
        assert np.all((r.params >= 0) & (r.params < 2 * np.pi))
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) >= 0)
    # Poisson: mean gap 1/rate
    assert abs(np.mean(np.diff(t)) - 1 / 500.0) < 2e-4


def test_request_validation():
    with pytest.raises(ValueError):
        SV.Request(0, 4, 2, "other", np.zeros(16))
    with pytest.raises(ValueError):
        SV.Request(0, 4, 2, SV.STATE_MW, np.zeros(15))


def test_request_bytes_counts_rolling_jacobian_rows():
    """A QFIM request needs its Jacobian (P rows), psi and a spare state,
    plus P ping-pong rows when its derivative plan uses rolling layouts,
    which the circuit layer picks while those rows fit its cap."""
    import types
    from pyramaterised_amd import serve as SV
    from pyramaterised_amd import templates as T
    from pyramaterised_amd.circuit import PQC
    stub = types.SimpleNamespace(amp=16)
    for n, depth in ((10, 4), (20, 40)):
        P = 2 * n * depth
        req = SV.Request(0, n, depth, SV.QFIM, np.zeros(P), 0.0)
        qc = PQC(n)
        for l in T.clifford_HE_layers(depth, n):
            qc.add_layer(l)
        assert qc.n_true_params == P
        rows = P + 2 + (P if qc.rolling_jacobian() else 0)
        assert SV.Server.request_bytes(stub, req) == 16 * 2 ** n * rows + 48 * P * P
    assert PQC(10).rolling_jacobian() is True
