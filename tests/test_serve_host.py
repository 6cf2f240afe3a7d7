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


def test_stream_lanes_keep_kinds_apart():
    """QFIMs and state requests never share a stream once there are two;
    with three or more, large QFIMs run one at a time on the first stream,
    other QFIMs on the second and states round-robin over the rest."""
    import types
    from collections import deque

    def server(n_streams, max_inflight=2):
        return types.SimpleNamespace(slots=[types.SimpleNamespace(inflight=deque(), i=i) for i in range(n_streams)],
                                     max_inflight=max_inflight, _next_slot={})

    pick = SV.Server._free_slot
    s1 = server(1)
    assert all(pick(s1, lane) is s1.slots[0] for lane in ("large", "qfim", "state"))
    s1.slots[0].inflight.append(0)
    assert pick(s1, "large") is None and pick(s1, "state") is s1.slots[0]
    s2 = server(2)
    assert pick(s2, "large").i == 0 and pick(s2, "qfim").i == 0 and pick(s2, "state").i == 1
    s2.slots[0].inflight.append(0)
    assert pick(s2, "large") is None and pick(s2, "qfim").i == 0
    s2.slots[1].inflight.extend([0, 0])
    assert pick(s2, "state") is None  # full: a state request waits rather than join the QFIM stream
    s4 = server(4)
    assert pick(s4, "large").i == 0 and pick(s4, "qfim").i == 1
    got = []
    for _ in range(4):  # QFIM admissions in between do not reset the state lane's rotation
        got.append(pick(s4, "state").i)
        pick(s4, "qfim")
    assert got == [2, 3, 2, 3]
    s4.slots[1].inflight.extend([0, 0])
    assert pick(s4, "qfim") is None
    s4.slots[2].inflight.extend([0, 0])
    s4.slots[3].inflight.extend([0, 0])
    assert pick(s4, "state") is None and pick(s4, "large").i == 0
    stub = types.SimpleNamespace(large_bytes=1 << 30, amp=16)
    stub.is_large = lambda r: SV.Server.is_large(stub, r)
    stub.request_bytes = lambda r: SV.Server.request_bytes(stub, r)
    lane = SV.Server._lane
    assert lane(stub, SV.Request(0, 8, 2, SV.STATE_MW, np.zeros(32), 0.0)) == "state"
    assert lane(stub, SV.Request(0, 8, 2, SV.QFIM, np.zeros(32), 0.0)) == "qfim"
    assert lane(stub, SV.Request(0, 20, 40, SV.QFIM, np.zeros(1600), 0.0)) == "large"


def test_wire_format_round_trips_bit_exact():
    """Requests and results written as JSON lines read back bit-exact
    (parameters through float repr, arrays through base64 bytes), and a
    request read back is validated like a constructed one."""
    import json
    import torch
    reqs = SV.make_workload(60, rate=50.0, seed=9)
    for r in reqs:
        back = SV.request_from_json(SV.request_to_json(r))
        assert (back.rid, back.n, back.depth, back.kind, back.arrival) == (r.rid, r.n, r.depth, r.kind, r.arrival)
        assert back.params.dtype == np.float64 and np.array_equal(back.params.view(np.uint64), r.params.view(np.uint64))
    bad = json.loads(SV.request_to_json(reqs[0]))
    bad["params"] = bad["params"][:-1]
    with pytest.raises(ValueError):
        SV.request_from_json(json.dumps(bad))
    rng = np.random.default_rng(4)
    qreq = SV.Request(7, 3, 2, SV.QFIM, rng.random(12), 1.25)
    res = SV.Result(qreq)
    res.qfim = rng.normal(size=(12, 12))
    res.state = torch.as_tensor(rng.normal(size=8) + 1j * rng.normal(size=8))
    res.launched, res.done = 1.5, 2.0
    d = SV.result_from_json(SV.result_to_json(res, with_state=True))
    assert d["rid"] == 7 and d["kind"] == SV.QFIM and d["Q"] is None and d["done"] == 2.0
    assert d["qfim"].dtype == np.float64 and np.array_equal(d["qfim"], res.qfim)
    assert d["state"].dtype == np.complex128 and np.array_equal(d["state"], res.state.numpy())
    assert "state" not in SV.result_from_json(SV.result_to_json(res))
    sres = SV.Result(SV.Request(8, 2, 1, SV.STATE_MW, rng.random(4)))
    sres.Q = 0.123456789012345678
    assert SV.result_from_json(SV.result_to_json(sres))["Q"] == sres.Q


def test_server_circuits_equal_the_layer_by_layer_build():
    """The server builds each (n, depth) class with one template call per
    layer and no copies; its op list (with derivative ops) equals the
    reference-style build (add_layer per layer, deep copies)."""
    import pyramaterised_amd as pyqc
    from pyramaterised_amd import serve as SV
    from pyramaterised_amd import templates as T

    def key(ops):
        return [(o.kind, o.q0, o.q1, o.slot, o.slot2, o.tangent, o.pauli, o.scale, o.offset, tuple(o.m))
                for o in ops]

    for n, d in ((2, 1), (5, 3), (9, 7)):
        c = SV._Circuit(n, d, 0)
        ref = pyqc.PQC(n)
        for layer in T.clifford_HE_layers(d, n):
            ref.add_layer(layer)
        assert c.qc.n_true_params == ref.n_true_params == 2 * n * d
        assert key(c.qc.ops(True)) == key(ref.ops(True))
        assert len({id(g) for g in c.qc.gates}) == len(c.qc.gates)
