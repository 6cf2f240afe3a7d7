"""Serving loop on the GPU: each result is the same alone and in the mix,
and matches the oracle; admission respects the HBM budget."""
import numpy as np
import pytest

from oracle import circuits as OC
from oracle import measures as OM
from oracle import sim as OS
from tests.helpers import infidelity, rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    from pyramaterised_amd import serve
    return serve


def _check_same(a, b):
    assert a.kind == b.kind and a.rid == b.rid
    if a.kind == "state_mw":
        sa, sb = a.state.cpu().numpy().ravel(), b.state.cpu().numpy().ravel()
        assert np.max(np.abs(sa - sb)) < 1e-12
        assert abs(a.Q - b.Q) < 1e-12
    else:
        assert rel_fro(a.qfim, b.qfim) < 1e-12


def test_alone_vs_mix_and_oracle(sv):
    reqs = sv.make_workload(48, rate=0.0, seed=5, n_range=(6, 12), depth_range=(2, 6), qfim_frac=0.3)
    srv = sv.Server(n_streams=3, max_batch=8)
    mix = srv.run(reqs, realtime=False)
    assert all(r.done is not None for r in mix)
    assert max(srv.batch_sizes) > 1  # some classes were batched
    alone, _ = sv.run_one_at_a_time(reqs)
    for a, b in zip(mix, alone):
        _check_same(a, b)
    # oracle on a sample of each kind
    checked = {"state_mw": 0, "qfim": 0}
    for r, res in zip(reqs, mix):
        if checked[r.kind] >= 3:
            continue
        og = OC.clifford_he(r.depth, r.n)
        ref = OS.run(og, r.params, r.n)
        if r.kind == "state_mw":
            assert infidelity(res.state.cpu().numpy(), ref) < 1e-12
            assert abs(res.Q - OM.single_Q(ref, r.n)) < 1e-12
        else:
            F = OM.qfim(OS.derivative_states(og, r.params, r.n), ref)
            assert rel_fro(res.qfim, F) < 1e-9
        checked[r.kind] += 1
    assert checked == {"state_mw": 3, "qfim": 3}


def test_realtime_arrivals_and_tight_budget(sv):
    """Poisson arrivals; a budget that fits only about two of the largest
    requests at once, so admission has to wait for memory."""
    reqs = sv.make_workload(40, rate=400.0, seed=6, n_range=(8, 12), depth_range=(2, 4), qfim_frac=0.25)
    probe = sv.Server(n_streams=1)
    biggest = max(probe.request_bytes(r) for r in reqs)
    srv = sv.Server(hbm_budget=2 * biggest + 1, n_streams=4, max_batch=16, keep_states=False)
    res = srv.run(reqs, realtime=True)
    assert srv.peak_in_use <= srv.budget
    assert srv.in_use == 0
    alone, _ = sv.run_one_at_a_time(reqs)
    for a, b in zip(res, alone):
        assert a.done >= a.arrival
        if a.kind == "state_mw":
            assert abs(a.Q - b.Q) < 1e-12
        else:
            assert rel_fro(a.qfim, b.qfim) < 1e-12
    st = sv.stats(res, srv.wall)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_release_returns_budget(sv):
    reqs = sv.make_workload(6, rate=0.0, seed=7, n_range=(6, 8), depth_range=(2, 3), qfim_frac=0.0)
    srv = sv.Server(n_streams=2)
    res = srv.run(reqs, realtime=False)
    held = srv.in_use
    assert held == sum(16 * (1 << r.n) for r in reqs)
    for r in res:
        srv.release(r.rid)
    assert srv.in_use == 0


def test_batched_states_release_together(sv):
    """States of one batched launch are rows of one tensor: their bytes stay
    counted until the last row of the batch is released."""
    import numpy as np
    reqs = [sv.Request(i, 7, 3, sv.STATE_MW, np.random.default_rng(i).random(42) * 6.28) for i in range(3)]
    srv = sv.Server(n_streams=1)
    res = srv.run(reqs, realtime=False)
    assert srv.batch_sizes == [3]
    full = srv.in_use
    assert full == 3 * 16 * (1 << 7)
    srv.release(res[0].rid)
    srv.release(res[1].rid)
    assert srv.in_use == full
    srv.release(res[2].rid)
    assert srv.in_use == 0
    srv.release(res[2].rid)  # releasing twice is a no-op
    assert srv.in_use == 0


def test_specialised_class_matches_alone(sv):
    """Classes large enough for run-time specialised kernels: after the
    background compile is loaded, results equal the interpreted path."""
    import time
    from pyramaterised_amd import engine as E
    reqs = [r for r in sv.make_workload(30, rate=0.0, seed=8, n_range=(14, 14), depth_range=(2, 2),
                                        qfim_frac=0.5)][:8]
    srv = sv.Server(n_streams=2, specialise_qfim_from=14, specialise_state_from=14, keep_states=True)
    c = srv.circuit(14, 2)
    for which in ("plan", "gplan"):
        c.specialise(which, srv._pool, 4)
    t0 = time.time()
    while c._pending and time.time() - t0 < 120:
        c.poll()
        time.sleep(0.05)
    assert c._jit.get("plan") is not None and c._jit.get("gplan") is not None
    assert c.best("plan").jitted and c.best("gplan").jitted
    res = srv.run(reqs, realtime=False)
    srv.close()
    ref = {}
    for r in reqs:
        cc = sv._Circuit(r.n, r.depth, E.L.PQS_C128)  # interpreted only
        ref[r.rid] = sv.run_one_at_a_time([r], circuits={(r.n, r.depth): cc})[0][0]
    for a in res:
        b = ref[a.rid]
        if a.kind == "state_mw":
            assert np.max(np.abs(a.state.cpu().numpy() - b.state.cpu().numpy())) < 1e-12
            assert abs(a.Q - b.Q) < 1e-12
        else:
            assert rel_fro(a.qfim, b.qfim) < 1e-12


def test_classes_merge_into_multi_plan_launches(sv):
    """State requests of different widths and depths share launches
    (pqs_run_multi): far fewer launches than classes, every result equal to
    the same job run alone."""
    reqs = sv.make_workload(60, rate=0.0, seed=11, n_range=(6, 16), depth_range=(2, 12), qfim_frac=0.0)
    classes = {(r.n, r.depth) for r in reqs}
    srv = sv.Server(n_streams=2, max_batch=32, keep_states=True)
    mix = srv.run(reqs, realtime=False)
    srv.close()
    assert len(classes) > 20
    assert srv.n_batches <= 4
    assert max(srv.batch_sizes) >= 16
    alone, _ = sv.run_one_at_a_time(reqs)
    for a, b in zip(mix, alone):
        _check_same(a, b)


def test_cancellation_and_launch_failure_isolation(sv, monkeypatch):
    """Requests cancelled in the queue and while running end "cancelled";
    one QFIM request whose launch raises (the failure injected from the
    host into its Jacobian call) ends "error" without touching the other
    requests, which equal their alone results; the budget comes back to
    zero once every kept state is released."""
    import time
    reqs = sv.make_workload(40, rate=0.0, seed=9, n_range=(6, 10), depth_range=(2, 5), qfim_frac=0.3)
    alone, _ = sv.run_one_at_a_time(reqs)
    bad = next(r for r in reqs if r.kind == sv.QFIM)
    real = sv.E.jacobian

    def jacobian(plan, params, init, **kw):
        if np.array_equal(params.reshape(-1).cpu().numpy(), bad.params):
            raise sv.E.EngineError("injected launch failure")
        return real(plan, params, init, **kw)

    monkeypatch.setattr(sv.E, "jacobian", jacobian)
    srv = sv.Server(n_streams=2, max_batch=4)
    for r in reqs:
        srv.submit(r)
    queued = [reqs[3].rid, reqs[17].rid]
    assert bad.rid not in queued
    for rid in queued:
        assert srv.cancel(rid) == sv.CANCELLED
    t0 = time.perf_counter()
    finished, running = 0, None
    while finished < len(reqs):
        done, _ = srv.step(time.perf_counter() - t0)
        finished += done
        if running is None:
            cand = [r.rid for r in reqs if srv.results[r.rid].launched is not None
                    and srv.results[r.rid].status is None and r.rid != bad.rid]
            if cand:
                running = cand[0]
                assert srv.cancel(running) == sv.CANCELLED
    assert running is not None
    for r, a in zip(reqs, alone):
        res = srv.results[r.rid]
        if r.rid in queued:
            assert res.status == sv.CANCELLED and res.launched is None
        elif r.rid == running:
            assert res.status == sv.CANCELLED and res.state is None and res.Q is None and res.qfim is None
        elif r.rid == bad.rid:
            assert res.status == sv.ERROR and "injected" in res.error and res.qfim is None
        else:
            assert res.status == sv.OK
            _check_same(res, a)
    assert sv.result_from_json(sv.result_to_json(srv.results[bad.rid]))["status"] == sv.ERROR
    for r in reqs:
        srv.release(r.rid)
    assert srv.in_use == 0
    srv.close()


def test_wire_format_round_trip_of_served_results(sv):
    """Requests read from JSON lines are served on the GPU; every result
    written as a JSON line (state included) and read back equals the served
    one bit for bit, and matches the oracle."""
    reqs = sv.make_workload(12, rate=0.0, seed=9, n_range=(6, 10), depth_range=(2, 5), qfim_frac=0.4)
    lines = [sv.request_to_json(r) for r in reqs]
    back = [sv.request_from_json(x) for x in lines]
    srv = sv.Server(n_streams=2, max_batch=4)
    served = srv.run(back, realtime=False)
    kinds = set()
    for r, res in zip(reqs, served):
        d = sv.result_from_json(sv.result_to_json(res, with_state=True))
        assert d["rid"] == r.rid and d["status"] == sv.OK and d["kind"] == r.kind
        og = OC.clifford_he(r.depth, r.n)
        ref = OS.run(og, r.params, r.n)
        kinds.add(r.kind)
        if r.kind == "state_mw":
            st = res.state.detach().cpu().numpy().ravel()
            assert np.array_equal(d["state"].ravel(), st)          # bit-exact through the wire
            assert d["Q"] == res.Q
            assert infidelity(d["state"].ravel(), ref) < 1e-12
            assert abs(d["Q"] - OM.single_Q(ref, r.n)) < 1e-12
        else:
            assert np.array_equal(d["qfim"], np.asarray(res.qfim, dtype=np.float64))
            F = OM.qfim(OS.derivative_states(og, r.params, r.n), ref)
            assert rel_fro(d["qfim"], F) < 1e-9
    assert kinds == {"state_mw", "qfim"}
    for r in reqs:
        srv.release(r.rid)
    srv.close()
