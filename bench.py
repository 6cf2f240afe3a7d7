"""Benchmark of the MI355X statevector engine.

Headline workload (one step): batched expressibility of a 16-qubit
hardware-efficient ansatz, 20 layers of (R_y, R_z on every qubit, CHAIN of
CNOTs) = 940 gates and 640 parameters; 5000 pairs of uniform [0, 2pi)
parameter vectors (10,000 states, complex128) evolved in one batched call, the
5000 pair fidelities |<psi_a|psi_b>|^2 binned into 75 bins and compared by KL
divergence with the Haar density for 2^16 dimensions.  ``value`` is gates/s
over the whole job (10,000 x 940 gates per step), inputs resident in HBM.

Extra legs (reported in the same JSON line): single 28-qubit state (30 layers
of R_y + brickwork CZ, 1245 gates) in complex128 and complex64 with achieved
HBM GB/s; the adjoint/forward QFIM of the 12-qubit TFIM HVA; speedups of the
batched / forward paths over the plain one-circuit-per-launch GPU path at 16
qubits; and the CPU baseline (the reference's algorithm restated with scipy
CSR operators, single thread) on a bounded sample of the headline workload.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12   # MI355X HBM3E bytes/s
PEAK_FP64 = 78.6e12  # MI355X FP64 flop/s (vector and matrix rates are equal)
PEAK_FP32 = 157.3e12  # MI355X FP32 flop/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def build_c2(pyqc, n=16, p=20, shape=None):
    """The headline circuit; ``shape`` = (tile bits, kept bits) fixes a
    rolling-layout tile shape instead of the tuned one."""
    from pyramaterised_amd import _lib as L
    from pyramaterised_amd import templates as T
    kw = {} if shape is None else dict(tile_bits=shape[0], low_bits=shape[1], flags=L.PQS_PLAN_ROLLING)
    qc = pyqc.PQC(n, **kw)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    return qc


def algorithmic_flops(qc):
    """Flops of one state evolution by the per-gate counts of a dense
    statevector update: rotations (R_x, R_y, R_z, R_zz) 6 per amplitude,
    fixed dense 1q 14, dense 2q 30, CNOT / CZ 0 (permutations and signs)."""
    from pyramaterised_amd import _lib as L
    per = {L.OP_RX: 6, L.OP_RY: 6, L.OP_RZ: 6, L.OP_RZZ: 6, L.OP_RXX: 6, L.OP_RYY: 6,
           L.OP_MAT1: 14, L.OP_MAT2: 30, L.OP_FSIM: 30}
    return sum(per.get(o.kind, 0) for o in qc.ops()) * float(2 ** qc.n_qubits)


def sweep_bytes(plan, batch, amp_bytes, init_reads=False):
    """Algorithmic HBM bytes of one batched run: each sweep reads and writes
    every state once, except the first sweep, which generates |0..0> without
    reading it."""
    per_state = amp_bytes * (2 ** plan.n)
    sweeps = plan.n_sweeps
    return batch * per_state * (2 * sweeps - (0 if init_reads else 1))


class EventTimer:
    """HIP events on the current stream (the stream libpqs launches on)."""

    def __init__(self):
        import torch
        self.s = torch.cuda.Event(enable_timing=True)
        self.e = torch.cuda.Event(enable_timing=True)

    def start(self):
        self.s.record()

    def stop(self):
        self.e.record()

    def ms(self):
        self.e.synchronize()
        return self.s.elapsed_time(self.e)


def headline(pyqc, args):
    import torch
    from pyramaterised_amd import engine as E
    qc = build_c2(pyqc, shape=args.shape)
    P = qc.n_true_params
    n_pairs = args.pairs
    S = 2 * n_pairs
    rng = np.random.default_rng(2024)
    params = torch.as_tensor(rng.random((S, P)) * 2 * np.pi, device="cuda")
    t0 = time.perf_counter()
    plan = qc.plan(batch=S, pairs=True)
    jit_s = time.perf_counter() - t0
    init = qc.init()
    states = torch.empty((S, 2 ** qc.n_qubits), dtype=torch.complex128, device="cuda")
    m = pyqc.measure.Measurements(qc)
    a_idx, b_idx = np.arange(0, S, 2), np.arange(1, S, 2)
    edges = np.linspace(0, 1, 76)
    ov = torch.empty(n_pairs, dtype=torch.complex128, device="cuda")

    def step(timer=None):
        # evolution of all 10,000 states (written to `states`) with the pair
        # overlaps formed in the last sweep, then |.|^2 and the histogram
        if timer:
            timer.start()
        E.run_pairs(plan, params, init, out=states, overlaps=ov)
        if timer:
            timer.stop()
        F = E.abs2(ov)
        counts = E.histogram(F, edges).cpu().numpy()
        return counts, F

    for _ in range(args.warmup):
        step()
    sync()
    timers = [EventTimer() for _ in range(args.steps)]
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        counts, F = step(timers[k])
    sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    evolve_ms = float(np.mean([t.ms() for t in timers]))
    gates = S * 940
    expr = m.expr(F, 2 ** 16, bins=75)
    # the fused overlaps against a read-back of the written states
    readback = E.inner_pairs(states, a_idx, b_idx)
    pair_diff = float((readback - ov).abs().max().item())
    fused = bool(getattr(plan, "jitted", False)) and any(b"pqs_jit_pair_" in x for x in plan.jit_sweep_sources())
    bytes_run = sweep_bytes(plan, S, 16)
    flops_run = algorithmic_flops(qc) * S
    launches = plan.n_sweeps
    return {
        "ms_per_step": ms_step, "gates_per_s": gates / (ms_step * 1e-3), "evolve_ms": evolve_ms,
        "n_sweeps": plan.n_sweeps, "tile_bits": plan.tile_bits, "sweep_launch_ms": evolve_ms / launches,
        "bytes_per_launch": bytes_run / launches, "achieved_Bps": bytes_run / (evolve_ms * 1e-3),
        "flops_per_launch": flops_run / launches, "achieved_flops": flops_run / (evolve_ms * 1e-3),
        "expr_kl": float(expr), "mean_F": float(F.mean().item()), "counts_nonzero": int((counts > 0).sum()),
        "S": S, "P": P, "jit": bool(getattr(plan, "jitted", False)), "plan_and_tune_s": jit_s,
        "fused_pair_overlaps": fused, "pair_overlap_maxdiff_vs_readback": pair_diff,
    }


def leg_28q(pyqc, dtype, steps=3):
    import torch
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p = 28, 30
    qc = pyqc.PQC(n, dtype=dtype)
    for l in T.ry_cz_brickwork_layers(p, n):
        qc.add_layer(l)
    th = np.random.default_rng(7).random(qc.n_true_params) * 2 * np.pi
    plan = qc.plan()  # specialised kernels, tile shape tuned on first use
    amp = 16 if dtype == "complex128" else 8
    out = torch.empty((1, 2 ** n), dtype=E.CDTYPE[plan.dtype], device="cuda")
    params = torch.as_tensor(th, device="cuda")
    init = qc.init()
    E.run(plan, params, init, out=out)
    sync()
    tm = EventTimer()
    tm.start()
    for _ in range(steps):
        E.run(plan, params, init, out=out)
    tm.stop()
    ms = tm.ms() / steps
    b = sweep_bytes(plan, 1, amp)
    flops = algorithmic_flops(qc)
    norm = float(torch.linalg.vector_norm(out).item())
    del out
    torch.cuda.empty_cache()
    # both rooflines: the sweeps stream the state 11 times and carry 840
    # rotations (FP64 vector work for complex128, FP32 for complex64)
    peak_fl = PEAK_FP64 if dtype == "complex128" else PEAK_FP32
    return {"dtype": dtype, "ms": ms, "gates_per_s": 1245 / (ms * 1e-3), "n_sweeps": plan.n_sweeps,
            "tile_bits": plan.tile_bits, "sweep_launch_ms": ms / plan.n_sweeps,
            "jit": bool(getattr(plan, "jitted", False)),
            "GBps": b / (ms * 1e-3) / 1e9, "hbm_frac": b / (ms * 1e-3) / PEAK_HBM,
            "TFLOPs": flops / (ms * 1e-3) / 1e12, "flop_frac": flops / (ms * 1e-3) / peak_fl,
            "flop_peak": "fp64 78.6 TFLOP/s" if dtype == "complex128" else "fp32 157.3 TFLOP/s", "norm": norm}


def leg_qfim12(pyqc):
    import torch
    from pyramaterised_amd import templates as T
    n, p = 12, 10
    qc = pyqc.PQC(n)
    qc.set_initial_state(T.plus_state())
    for l in T.tfim_open_layers(p, n):
        qc.add_layer(l)
    th = list(np.random.default_rng(3).random(20) * 2 * np.pi)
    qc.state = qc.run(th)
    m = pyqc.measure.Measurements(qc)
    m.get_QFI()
    sync()
    t0 = time.perf_counter()
    for _ in range(5):
        F = m.get_QFI()
    sync()
    fwd = (time.perf_counter() - t0) / 5

    def plain_qfim():
        G = qc.get_gradients(method="plain")
        Fp = np.zeros((20, 20))
        single = [qc.state.overlap(g) for g in G]
        for a in range(20):
            for b in range(a, 20):
                Fp[a, b] = 4 * np.real(G[a].overlap(G[b]) - np.conj(single[a]) * single[b])
        return Fp

    # untimed: the plain path's plans, and their specialised kernels, which
    # plans get on their third use (timing earlier uses would include the
    # run-time compile)
    for _ in range(pyqc.PQC.JIT_AFTER_USES):
        plain_qfim()
    sync()
    t0 = time.perf_counter()
    for _ in range(3):
        Fp = plain_qfim()
    sync()
    plain = (time.perf_counter() - t0) / 3
    Fp = Fp + np.triu(Fp, 1).T
    import scipy.linalg
    ev = scipy.linalg.eigh(F, eigvals_only=True)
    return {"forward_ms": fwd * 1e3, "plain_ms": plain * 1e3, "speedup": plain / fwd,
            "eqd": int(np.sum(ev > 1e-12)), "max_diff_vs_plain": float(np.max(np.abs(F - Fp)))}


def speedups_16q(pyqc, args, head):
    """Batched expressibility and forward QFIM vs the plain path (one circuit
    per launch sequence, per-parameter re-simulation, one overlap per call) at
    16 qubits on the headline circuit; plain legs are timed on a sample and
    scaled linearly (their cost is a sum of identical independent calls)."""
    import torch
    from pyramaterised_amd import engine as E
    qc = build_c2(pyqc)
    qc.plan(batch=64)  # same specialised kernels for the plain path
    P = qc.n_true_params
    rng = np.random.default_rng(11)
    # plain expressibility: states one at a time, overlaps one at a time
    n_sample = args.plain_states
    prm = rng.random((n_sample, P)) * 2 * np.pi
    w = qc.run(list(prm[0]))  # untimed: the single-state plan
    abs(w.overlap(w))
    sync()
    t0 = time.perf_counter()
    sts = [qc.run(list(prm[i])) for i in range(n_sample)]
    fs = [abs(sts[i].overlap(sts[i + 1])) ** 2 for i in range(0, n_sample - 1, 2)]
    sync()
    t_plain_sample = time.perf_counter() - t0
    plain_expr_s = t_plain_sample * (head["S"] / n_sample)
    expr_speedup = plain_expr_s / (head["ms_per_step"] * 1e-3)
    # QFIM at 16 qubits, P = 640
    th = list(rng.random(P) * 2 * np.pi)
    qc.state = qc.run(th)
    m = pyqc.measure.Measurements(qc)
    F = m.get_QFI()
    sync()
    t0 = time.perf_counter()
    F = m.get_QFI()
    sync()
    fwd_s = time.perf_counter() - t0
    # plain: P re-simulations (all timed, after an untimed pass) + P(P+1)/2
    # overlaps (sampled)
    qc.jacobian(method="plain")
    sync()
    t0 = time.perf_counter()
    psi, J = qc.jacobian(method="plain")
    sync()
    t_resim = time.perf_counter() - t0
    G = [pyqc.Ket(J[i], qc.n_qubits) for i in range(P)]
    n_ov = args.plain_overlaps
    pairs = [(int(a), int(b)) for a, b in zip(rng.integers(0, P, n_ov), rng.integers(0, P, n_ov))]
    sync()
    t0 = time.perf_counter()
    for a, b in pairs:
        G[a].overlap(G[b])
    sync()
    t_ov = (time.perf_counter() - t0) / n_ov
    plain_qfim_s = t_resim + t_ov * (P * (P + 1) / 2 + P)
    # check forward J against plain J on this circuit
    _, Jf = qc.jacobian()
    jdiff = float((Jf - J).abs().max().item())
    del J, Jf, G
    torch.cuda.empty_cache()
    return {"expr_plain_s_est": plain_expr_s, "expr_speedup": expr_speedup, "plain_states_timed": n_sample,
            "qfim16_forward_s": fwd_s, "qfim16_plain_s_est": plain_qfim_s, "qfim16_speedup": plain_qfim_s / fwd_s,
            "qfim16_plain_resim_s": t_resim, "qfim16_overlap_call_us": t_ov * 1e6, "jac_forward_vs_plain_maxdiff": jdiff}


def events_ms(f, reps=5):
    """Best of ``reps`` HIP-event timings of f() on the current stream
    (after one untimed call)."""
    import torch
    f()
    sync()
    best = 1e30
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        f()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def leg_gram(pyqc):
    """The Gram steps alone on the headline circuit (16 qubits, P = 640): the
    QFIM from its Jacobian (real product on the matrix cores, b vector folded
    in) and the all-pairs fidelity Gram of 2000 sample states (complex
    product, upper triangle); HIP events, FP64 fractions of the upper
    triangle's flops."""
    import torch
    from pyramaterised_amd import engine as E
    qc = build_c2(pyqc)
    P, n = qc.n_true_params, qc.n_qubits
    rng = np.random.default_rng(16)
    qc.set_params(list(rng.random(P) * 2 * np.pi))
    psi, J = qc.jacobian()
    gram_ms = events_ms(lambda: E.qfim(J, psi))
    gram_flops = 2.0 * (P * (P + 1) / 2) * 2 * 2 ** n
    del J, psi
    torch.cuda.empty_cache()
    S2 = 2000
    sts = qc.run_batch(rng.random((S2, P)) * 2 * np.pi)
    ap_ms = events_ms(lambda: E.gram(sts, sts))
    ap_flops = 8.0 * (S2 * (S2 + 1) / 2) * 2 ** n
    del sts
    torch.cuda.empty_cache()
    return {"qfim16_gram_ms": gram_ms, "qfim16_gram_fp64_frac": gram_flops / (gram_ms * 1e-3) / 78.6e12,
            "allpairs2000_gram_ms": ap_ms, "allpairs2000_gram_fp64_frac": ap_flops / (ap_ms * 1e-3) / 78.6e12}


def leg_adjoint20(pyqc):
    """Energy gradient of a 20-qubit, 40-layer HEA (P = 1600, H = Z0 Z1):
    the adjoint pass (two states) against the forward-Jacobian route (P
    derivative states + one-row Gram); time, device memory high-water and
    the largest difference between the two gradients."""
    import torch
    from pyramaterised_amd import adjoint
    from pyramaterised_amd import engine as E
    from pyramaterised_amd import templates as T
    n, p = 20, 40
    qc = pyqc.PQC(n)
    for l in T.clifford_HE_layers(p, n):
        qc.add_layer(l)
    P = qc.n_true_params
    th = np.random.default_rng(20).random(P) * 2 * np.pi
    qc.set_params(list(th))
    H = qc.hamiltonian()

    def jac_route():
        psi, J = qc.jacobian()
        Hpsi = adjoint.apply_hamiltonian(H, psi, n, adjoint=True)
        g = E.gram(Hpsi.reshape(1, -1), J)[0]
        return (2 * g.real).cpu().numpy()

    out = {"n_qubits": n, "layers": p, "params": P}
    for name, f in (("adjoint", lambda: adjoint.energy_and_gradient(qc, th)[1]), ("jacobian", jac_route)):
        f()  # untimed: plans, adjoint program
        sync()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            g = f()
        sync()
        out[f"{name}_ms"] = (time.perf_counter() - t0) / reps * 1e3
        out[f"{name}_peak_GB"] = (torch.cuda.max_memory_allocated() - base) / 1e9
        out[f"{name}_grad"] = g
        torch.cuda.empty_cache()
    out["max_abs_diff"] = float(np.max(np.abs(out.pop("adjoint_grad") - out.pop("jacobian_grad"))))
    out["speedup"] = out["jacobian_ms"] / out["adjoint_ms"]
    out["measurement_points"] = len(adjoint.program(qc).steps)
    # the QFIM of the same circuit: the P = 1600 Gram of a 26.8 GB Jacobian
    # (column blocks through the matrix cores), HIP events around pqs_qfim
    psi, J = qc.jacobian()
    ms = events_ms(lambda: E.qfim(J, psi), reps=3)
    flops = 2.0 * (P * (P + 1) / 2) * 2 * 2 ** n  # upper triangle of Re(J J^dag)
    out["qfim_gram_ms"] = ms
    out["qfim_gram_fp64_frac"] = flops / (ms * 1e-3) / 78.6e12
    del J, psi
    torch.cuda.empty_cache()
    return out


def leg_serving(args):
    """Serving mix: Poisson requests (HEA, 6-20 qubits, depth 2-40, 70% state
    + Meyer-Wallach, 30% QFIM, complex128) through the continuous-batching
    server, then the same jobs one at a time; results must agree."""
    from pyramaterised_amd import serve as SV
    reqs = SV.make_workload(args.serve_requests, rate=args.serve_rate, seed=5)
    srv = SV.Server(keep_states=False)
    res = srv.run(reqs, realtime=True)
    st = SV.stats(res, srv.wall)
    srv.close()
    # saturation: every request queued at t = 0 (closed loop), same server
    # state (plans and specialised kernels already built)
    sat = SV.Server(keep_states=False)
    sat.circuits = srv.circuits
    at0 = [SV.Request(r.rid, r.n, r.depth, r.kind, r.params, 0.0) for r in reqs]
    res_sat = sat.run(at0, realtime=False)
    st_sat = SV.stats(res_sat, sat.wall)
    sat.close()
    alone, wall = SV.run_one_at_a_time(reqs, circuits=srv.circuits)
    dq, dF = 0.0, 0.0
    for a, b in zip(res, alone):
        if a.kind == SV.STATE_MW:
            dq = max(dq, abs(a.Q - b.Q))
        else:
            dF = max(dF, float(np.linalg.norm(a.qfim - b.qfim) / max(np.linalg.norm(b.qfim), 1e-300)))
    st.update({"rate_per_s": args.serve_rate, "batches": srv.n_batches,
               "mean_batch": float(np.mean(srv.batch_sizes)), "max_batch": int(np.max(srv.batch_sizes)),
               "specialised_classes": sum(1 for c in srv.circuits.values() for v in c._jit.values() if v is not None),
               "saturated_requests_per_s": st_sat["requests_per_s"], "saturated_p50_ms": st_sat["p50_ms"],
               "saturated_mean_batch": float(np.mean(sat.batch_sizes)),
               "one_at_a_time_s": wall, "one_at_a_time_requests_per_s": len(reqs) / wall,
               "max_abs_dQ_mix_vs_alone": dq, "max_rel_dQFIM_mix_vs_alone": dF,
               "qfim_requests": sum(r.kind == SV.QFIM for r in reqs)})
    for kind, name in ((SV.STATE_MW, "state"), (SV.QFIM, "qfim")):
        lat = np.array([r.latency for r in res if r.kind == kind])
        if lat.size:
            st[name + "_p50_ms"] = float(np.percentile(lat, 50) * 1e3)
            st[name + "_p99_ms"] = float(np.percentile(lat, 99) * 1e3)
    return st


def traffic_per_launch(head):
    """HBM bytes per sweep launch measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE on this workload with the same tile shape and sweep count as
    this run's plan, scaled by the same counters' reading for a plain copy
    of known size (profiles/traffic_headline.json, one entry per plan
    shape); None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "traffic_headline.json")
    try:
        with open(path) as f:
            t = json.load(f)["plans"][f"{head['tile_bits']}x{head['n_sweeps']}"]
        return float(t["bytes_per_launch"]) * head["S"] / float(t["states"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


def cpu_baseline(args):
    from oracle import cpu_ref
    n, p = 16, 20
    gates = cpu_ref.clifford_he_gates(p, n)
    rng = np.random.default_rng(2024)
    k = args.cpu_states
    prm = rng.random((k, 640)) * 2 * np.pi
    t0 = time.perf_counter()
    sts = [cpu_ref.run(gates, prm[i], n) for i in range(k)]
    for i in range(0, k - 1, 2):
        abs(np.vdot(sts[i], sts[i + 1])) ** 2
    dt = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": k * 940 / dt, "unit": "gates/s", "cores": 1, "kind": "port",
            "sample": f"{k} of the 10,000 states of the headline workload (16 qubits, 940 gates each) "
                      f"and their pair fidelity: per-gate CSR operators rebuilt per parameter + spmv, "
                      f"scipy single thread, {dt:.1f} s",
            "cpu_model": model, "host_cpus": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=5000)
    ap.add_argument("--plain-states", type=int, default=200)
    ap.add_argument("--plain-overlaps", type=int, default=2000)
    ap.add_argument("--cpu-states", type=int, default=12)
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-rate", type=float, default=25.0)
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--legs", default="28q,qfim12,adjoint20,gram,serving,speedups",
                    help="extra legs after the headline (comma list of 28q, qfim12, adjoint20, gram, serving, "
                         "speedups)")
    ap.add_argument("--only-legs", action="store_true",
                    help="skip the headline and the CPU baseline; print the legs' JSON only (for profiling a leg)")
    ap.add_argument("--shape", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    help="diagnostic: fix the headline's tile shape as 'TILE_BITS,KEPT_BITS' (rolling layouts) "
                         "instead of tuning it")
    ap.add_argument("--calib", action="store_true",
                    help="first copy 2 GiB (complex128, 2^27 elements) three times: a plain copy of known size "
                         "that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(the batched workloads shard as independent instances per GPU; one GPU is supported)")
        sys.exit(2)
    import torch
    import pyramaterised_amd as pyqc
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    if args.calib:
        a = torch.ones(2 ** 27, dtype=torch.complex128, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        sync()
        log("calibration copy: 3 x (read 2 GiB + write 2 GiB)")
        del a, b
        torch.cuda.empty_cache()
    legs = set(x for x in args.legs.split(",") if x)
    head = None
    if not args.only_legs:
        log("headline: batched expressibility, 16 qubits, 5000 pairs")
        head = headline(pyqc, args)
        log(json.dumps(head))
    extra = {}
    if not args.quick:
        if "28q" in legs:
            for dt in ("complex128", "complex64"):
                log(f"28-qubit single state, {dt}")
                extra[f"single_state_28q_{dt}"] = leg_28q(pyqc, dt)
                log(json.dumps(extra[f"single_state_28q_{dt}"]))
        if "qfim12" in legs:
            log("12-qubit TFIM QFIM")
            extra["qfim_12q_tfim"] = leg_qfim12(pyqc)
            log(json.dumps(extra["qfim_12q_tfim"]))
        if "adjoint20" in legs:
            log("20-qubit energy gradient: adjoint pass vs forward Jacobian")
            extra["gradient_20q_adjoint"] = leg_adjoint20(pyqc)
            log(json.dumps(extra["gradient_20q_adjoint"]))
        if "gram" in legs:
            log("16-qubit Gram steps: QFIM (P = 640) and all-pairs fidelities (2000 states)")
            extra["gram_16q"] = leg_gram(pyqc)
            log(json.dumps(extra["gram_16q"]))
        if "serving" in legs:
            log(f"serving mix: {args.serve_requests} requests at {args.serve_rate}/s")
            extra["serving"] = leg_serving(args)
            log(json.dumps(extra["serving"]))
        if "speedups" in legs and head is not None:
            log("16-qubit speedups vs plain path")
            extra["speedup_16q"] = speedups_16q(pyqc, args, head)
            log(json.dumps(extra["speedup_16q"]))
    if args.only_legs:
        print(json.dumps(extra))
        return
    log("cpu baseline")
    cpu = cpu_baseline(args)
    log(json.dumps(cpu))
    out = {
        "metric": "gates/s + HBM GB/s frac; batched+adjoint vs plain QFIM/expr speedup @1e-12",
        "value": head["gates_per_s"], "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched expressibility: 16-qubit HEA (R_y,R_z + CNOT CHAIN) x 20 layers, "
                               "940 gates, 640 params, 5000 pairs (10,000 states), 75-bin KL vs Haar(2^16)",
                   "n_qubits": 16, "layers": 20, "states": head["S"], "pairs": head["S"] // 2,
                   "amplitudes": "complex128"},
        # The headline's sweeps hold 10,000 resident 1 MiB states and run
        # ~10 register phases per LDS tile: bound by FP64 vector FMAs and LDS
        # exchange, so the FP64 compute roofline applies (MI355X's FP64
        # vector peak equals its FP64 matrix peak, 78.6 TFLOP/s).
        "roofline": {"bound": "fp64-valu", "achieved": head["achieved_flops"] / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": head["achieved_flops"] / PEAK_FP64,
                     "traffic": traffic_per_launch(head),
                     "kernel": "pqs_jit_sweep_* + pqs_jit_pair_* (run-time specialised sweeps, the last one forming the pair overlaps; all launches of one evolution)",
                     "launch_ms": head["sweep_launch_ms"], "flops_per_launch": head["flops_per_launch"],
                     "hbm_GBps": head["achieved_Bps"] / 1e9, "hbm_frac": head["achieved_Bps"] / PEAK_HBM},
        "cpu_baseline": cpu,
        "headline_detail": head,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
