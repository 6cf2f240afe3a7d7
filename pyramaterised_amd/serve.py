"""Request serving: continuous batching of mixed-width circuit jobs on one GPU.

A request is a hardware-efficient ansatz (``templates.clifford_HE_layers``:
R_y and R_z on every qubit, then a CNOT CHAIN, per layer; P = 2 n depth
parameters) with its own parameter vector.  It asks for either

  * the final state and its Meyer-Wallach entanglement
    (``PQC.run``, circuit.py:118-125; ``Measurements.single_Q``,
    measure.py:226-249), or
  * the full QFIM (``Measurements.get_QFI``, measure.py:33-71).

The reference evaluates one circuit at a time in one Python process and has
no serving layer.  Here a ``Server`` keeps a queue per width class (qubit
count, depth, request kind) and, at every turn of its loop:

  1. admits requests that have arrived, oldest class first, while their
     device memory fits the HBM budget (states stay resident until
     ``release``);
  2. launches one batched evolution per admitted class (all its requests'
     parameter vectors in one call per sweep), or one forward Jacobian + QFIM
     per QFIM request, on the next free HIP stream; each stream has its own
     engine context (a context's workspace serves one stream);
  3. polls the HIP events of launched batches and retires finished requests.

Requests that arrive while earlier batches are on the GPU are admitted at
the next turn as streams and memory free up, and batches of different
classes run concurrently on different streams.  A batch's results do not
depend on what else is in it: every instance runs the same kernels on its
own state.
"""
import base64
import json
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import engine as E
from . import templates as T
from .circuit import PQC

STATE_MW = "state_mw"
QFIM = "qfim"


class Request:
    """One job: n qubits, depth layers, kind, parameters, arrival time (s)."""

    __slots__ = ("rid", "n", "depth", "kind", "params", "arrival")

    def __init__(self, rid, n, depth, kind, params, arrival=0.0):
        if kind not in (STATE_MW, QFIM):
            raise ValueError(f"unknown request kind {kind!r}")
        params = np.asarray(params, dtype=np.float64).reshape(-1)
        if params.size != 2 * n * depth:
            raise ValueError(f"request {rid}: expected {2 * n * depth} parameters, got {params.size}")
        self.rid, self.n, self.depth, self.kind = int(rid), int(n), int(depth), kind
        self.params, self.arrival = params, float(arrival)


OK, CANCELLED, ERROR = "ok", "cancelled", "error"


class Result:
    """Outcome of a request.  ``state`` is a device tensor view (resident
    until ``Server.release``); ``Q`` the Meyer-Wallach value; ``qfim`` a
    (P, P) numpy array.  ``status``: None while queued or running, then
    "ok", "cancelled" (``Server.cancel``) or "error" (its launch raised;
    ``error`` holds the message).  Times are seconds on the server's
    clock."""

    __slots__ = ("rid", "kind", "n", "depth", "state", "Q", "qfim", "arrival", "launched", "done", "status",
                 "error")

    def __init__(self, req):
        self.rid, self.kind, self.n, self.depth = req.rid, req.kind, req.n, req.depth
        self.state = self.Q = self.qfim = None
        self.arrival = req.arrival
        self.launched = self.done = None
        self.status = self.error = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


# ------------------------------------------------------------ wire format
# One JSON object per line.  Floats of the parameter vector are written with
# Python's shortest round-trip repr, and arrays (QFIM, state) as base64 of
# their little-endian bytes, so a request or result read back is bit-exact.

def _pack_array(a):
    a = np.ascontiguousarray(a)
    le = a.astype(a.dtype.newbyteorder("<"), copy=False)
    return {"dtype": a.dtype.name, "shape": list(a.shape), "data": base64.b64encode(le.tobytes()).decode("ascii")}


def _unpack_array(d):
    dt = np.dtype(d["dtype"]).newbyteorder("<")
    return np.frombuffer(base64.b64decode(d["data"]), dtype=dt).reshape(d["shape"]).astype(dt.newbyteorder("="))


def request_to_json(req):
    """One request as a JSON line."""
    return json.dumps({"rid": req.rid, "n": req.n, "depth": req.depth, "kind": req.kind,
                       "params": [float(x) for x in req.params], "arrival": req.arrival})


def request_from_json(line):
    """A Request from ``request_to_json`` output; the same validation as the
    constructor (kind, parameter count)."""
    d = json.loads(line)
    return Request(d["rid"], d["n"], d["depth"], d["kind"], np.asarray(d["params"], dtype=np.float64),
                   d.get("arrival", 0.0))


def result_to_json(res, with_state=False):
    """One result as a JSON line: Q for state requests, the QFIM as a packed
    float64 array, the state (copied to the host) only with ``with_state``,
    and the server's times."""
    d = {"rid": res.rid, "kind": res.kind, "n": res.n, "depth": res.depth, "arrival": res.arrival,
         "launched": res.launched, "done": res.done, "Q": res.Q, "status": res.status}
    if res.error is not None:
        d["error"] = res.error
    if res.qfim is not None:
        d["qfim"] = _pack_array(np.asarray(res.qfim, dtype=np.float64))
    if with_state and res.state is not None:
        st = res.state.detach().cpu().numpy() if isinstance(res.state, torch.Tensor) else np.asarray(res.state)
        d["state"] = _pack_array(st)
    return json.dumps(d)


def result_from_json(line):
    """A dict with the fields of ``result_to_json``; arrays as numpy."""
    d = json.loads(line)
    for key in ("qfim", "state"):
        if key in d:
            d[key] = _unpack_array(d[key])
    return d


def make_workload(n_requests, rate, seed=0, n_range=(6, 20), depth_range=(2, 40), qfim_frac=0.3):
    """A Poisson stream of ``n_requests`` requests at ``rate`` per second:
    qubits uniform over n_range, depth uniform over depth_range (inclusive),
    a QFIM request with probability ``qfim_frac``, parameters uniform
    [0, 2 pi), all from one seeded generator."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests)) if rate > 0 else np.zeros(n_requests)
    out = []
    for i in range(n_requests):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        d = int(rng.integers(depth_range[0], depth_range[1] + 1))
        kind = QFIM if rng.random() < qfim_frac else STATE_MW
        out.append(Request(i, n, d, kind, rng.random(2 * n * d) * 2 * np.pi, float(t[i])))
    return out



class _Circuit:
    """Compiled plans of one (n, depth) class, shared by its requests.

    ``plan`` / ``gplan`` (with generators) run on the interpreted sweep
    kernels from the first request on.  ``specialise`` compiles run-time
    specialised kernels for them in the background (child processes, no GPU);
    ``poll`` loads finished ones on the calling (serving) thread, after which
    ``best`` returns them."""

    def __init__(self, n, depth, dtype):
        self.qc = PQC(n, dtype=dtype, jit=False)
        # one template call per layer: every layer has gate objects of its own
        # (the template repeats one layer object, which add_layer deep-copies)
        self.qc.add_layers([T.clifford_HE_layers(1, n)[0] for _ in range(depth)], copy=False)
        self.init = self.qc.init()
        self._plan = None
        self._gplan = None
        self._jit = {}
        self._pending = {}
        self._lock = threading.Lock()  # plans are built by the preparing thread or on demand

    def specialise(self, which, pool, workers):
        if which in self._jit or which in self._pending:
            return
        gen = which == "gplan"
        p = E.Plan(self.qc.n_qubits, self.qc.ops(gen), dtype=self.qc.dtype, **self.qc.plan_options(gen))
        srcs = p.jit_sweep_sources()
        if not srcs:
            self._jit[which] = None
            return
        self._pending[which] = (p, pool.submit(E._compile_in_children, srcs, workers))

    def poll(self):
        for which, (p, fut) in list(self._pending.items()):
            if not fut.done():
                continue
            del self._pending[which]
            try:
                fut.result()
                p.jit(workers=1)  # code objects come from the cache
                self._jit[which] = p
            except Exception:  # stay on the interpreted kernels
                self._jit[which] = None

    def best(self, which):
        j = self._jit.get(which)
        return j if j is not None else getattr(self, which)

    @property
    def plan(self):
        with self._lock:
            if self._plan is None:
                self._plan = self.qc.plan()
            return self._plan

    @property
    def gplan(self):
        with self._lock:
            if self._gplan is None:
                self._gplan = self.qc.plan(with_generators=True)
            return self._gplan


class _Slot:
    def __init__(self, dev):
        self.stream = torch.cuda.Stream(device=dev)  # default priority
        self.ctx = E.Context(dev.index)
        self.inflight = deque()


class _Batch:
    __slots__ = ("reqs", "results", "event", "bytes", "out", "q", "F", "slot")


class Server:
    """Continuous-batching request loop on the current GPU.

    hbm_budget: device bytes that admitted, unreleased requests may hold
    (default: half of the device memory; each stream also keeps one Jacobian
    buffer, sized to the largest QFIM it served, until ``close``).  n_streams: concurrent HIP streams;
    QFIMs and state requests do not share one (with three or more, the
    first carries large QFIMs one at a time, the second the other QFIMs and
    the rest the state requests; large_bytes is the device footprint from
    which a QFIM request counts as large; qfim_launch_blocks caps the blocks
    of one sweep launch on the QFIM streams, so state kernels do not wait
    behind whole Jacobian sweeps).
    max_batch: most requests of one class per launch.  keep_states: keep
    every state-request's state resident (counted in the budget) until
    ``release``; otherwise only Q is returned."""

    def __init__(self, hbm_budget=None, n_streams=4, max_batch=512, dtype="complex128", keep_states=True,
                 max_inflight_per_stream=2, specialise_qfim_from=14, specialise_state_from=18, compile_workers=8,
                 large_bytes=1 << 30, merge_classes=True, prepare_workers=4, qfim_launch_blocks=16384):
        self.dev = E.device()
        if hbm_budget is None:
            _, total = torch.cuda.mem_get_info(self.dev)
            hbm_budget = total // 2
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.dtype = E.dtype_code(dtype)
        self.amp = 16 if self.dtype == E.L.PQS_C128 else 8
        self.max_batch = int(max_batch)
        self.keep_states = keep_states
        self.max_inflight = int(max_inflight_per_stream)
        self.large_bytes = int(large_bytes)
        self.merge_classes = bool(merge_classes)
        self.slots = [_Slot(self.dev) for _ in range(max(1, n_streams))]
        if len(self.slots) > 1:
            # the streams that carry QFIMs split their Jacobian sweeps into
            # launches of at most this many blocks: a state batch's kernels
            # on another stream wait at most one such launch, not a whole
            # sweep of hundreds of rows (tens of milliseconds each)
            for s in self.slots[:1 if len(self.slots) == 2 else 2]:
                s.ctx.set_max_blocks(qfim_launch_blocks)
        self._next_slot = {}  # lane -> round-robin position among its streams
        self.circuits = {}
        self.queues = OrderedDict()  # class key -> deque of requests (FIFO)
        self.results = {}
        self._held = {}    # rid -> group of the tensor holding its resident state
        self._groups = {}  # group -> [rows not yet released, bytes held]
        self._next_group = 0
        self.n_batches = 0
        self.batch_sizes = []
        self.peak_in_use = 0
        # classes whose requests are large enough for specialised kernels to
        # pay back their compile: compiled in the background
        self.spec_qfim, self.spec_state = specialise_qfim_from, specialise_state_from
        self.compile_workers = compile_workers
        self._pool = ThreadPoolExecutor(max_workers=2)
        self._compiling = set()
        self._seen = {}
        self._n_pending = 0
        self._n_inflight = 0
        self._order = []          # class keys, oldest head-of-queue first
        self._order_dirty = False
        self._jscratch = {}       # slot -> reusable Jacobian buffer (stream-ordered reuse)
        self._cancel_running = set()  # rids cancelled while their batch runs
        # a class's circuit and plan are built in the background from its
        # first submitted request on (the planner's C++ part runs without
        # the GIL); admission takes only classes whose plan is ready
        self._prep_pool = ThreadPoolExecutor(max_workers=max(1, int(prepare_workers)))
        self._prep = {}           # (n, depth, plan name) -> future
        self._circuits_lock = threading.Lock()
        self._n_settled = 0       # requests cancelled in the queue or failed at launch, not yet counted

    # ---------------------------------------------------------------- sizing
    def request_bytes(self, req):
        """Device bytes a request needs while running (and, for a kept
        state, afterwards)."""
        dim = 1 << req.n
        if req.kind == STATE_MW:
            return self.amp * dim + 64 * req.n
        P = 2 * req.n * req.depth
        rows = P + 2
        if P * self.amp * dim <= PQC.ROLLING_JACOBIAN_BYTES:
            rows += P  # ping-pong rows of a rolling-layout Jacobian (workspace)
        return self.amp * dim * rows + 48 * P * P

    def circuit(self, n, depth):
        key = (n, depth)
        with self._circuits_lock:
            c = self.circuits.get(key)
            if c is None:
                c = self.circuits[key] = _Circuit(n, depth, self.dtype)
        return c

    @staticmethod
    def _plan_name(kind):
        return "gplan" if kind == QFIM else "plan"

    def _prepare(self, n, depth, which):
        getattr(self.circuit(n, depth), which)

    def _ready(self, key, now):
        """The class's plan is built.  A class whose preparation raised has
        its queued requests failed (status "error")."""
        n, depth, kind = key
        f = self._prep.get((n, depth, self._plan_name(kind)))
        if f is None or not f.done():
            return False
        err = f.exception()
        if err is None:
            return True
        q = self.queues[key]
        reqs = list(q)
        q.clear()
        self._n_pending -= len(reqs)
        self._order_dirty = True
        self._fail(reqs, 0, err, now)
        return False

    def _preparing(self):
        return any(not f.done() for f in self._prep.values())

    # ---------------------------------------------------------------- queue
    def submit(self, req):
        if self.request_bytes(req) > self.budget:
            raise ValueError(f"request {req.rid} needs {self.request_bytes(req)} bytes, over the HBM budget "
                             f"{self.budget}")
        self.results[req.rid] = Result(req)
        self.queues.setdefault((req.n, req.depth, req.kind), deque()).append(req)
        self._n_pending += 1
        self._order_dirty = True
        pk = (req.n, req.depth, self._plan_name(req.kind))
        if pk not in self._prep:
            self._prep[pk] = self._prep_pool.submit(self._prepare, *pk)

    def pending(self):
        return self._n_pending

    def inflight(self):
        return self._n_inflight

    def release(self, rid):
        """Drop a resident state.  States of one batched launch are rows of
        one device tensor, which is freed only when its last row is
        released, so their bytes return to the budget together then."""
        r = self.results.get(rid)
        if r is not None:
            r.state = None
        h = self._held.pop(rid, None)
        if h is None:
            return
        group = self._groups[h]
        group[0] -= 1
        if group[0] == 0:
            self.in_use -= group[1]
            del self._groups[h]

    def cancel(self, rid, now=None):
        """Cancel a request.  Queued: it leaves its queue at once.  Running:
        its batch completes (other requests of the batch are unaffected)
        and its result is dropped when the batch retires.  Finished: its
        resident state is released.  Returns the result's status after the
        call ("cancelled", or "ok" / "error" for a finished request).  Call
        it from the thread that drives the server (between ``step`` calls
        or from ``run``'s progress callback)."""
        res = self.results.get(rid)
        if res is None:
            raise KeyError(f"unknown request {rid}")
        if res.status is not None:
            if res.status == OK:
                self.release(rid)
            return res.status
        if res.launched is None:
            key = (res.n, res.depth, res.kind)
            q = self.queues.get(key)
            for r in q:
                if r.rid == rid:
                    q.remove(r)
                    break
            self._n_pending -= 1
            self._order_dirty = True
            res.status, res.done = CANCELLED, now
            self._n_settled += 1
            return CANCELLED
        self._cancel_running.add(rid)
        return CANCELLED

    def _settle(self):
        n, self._n_settled = self._n_settled, 0
        return n

    def _fail(self, reqs, need, err, now):
        """A launch raised before any work was queued: its requests end with
        status "error" and give their bytes back; the server goes on."""
        self.in_use -= need
        for r in reqs:
            res = self.results[r.rid]
            res.status, res.error, res.done = ERROR, f"{type(err).__name__}: {err}", now
            self._cancel_running.discard(r.rid)
        self._n_settled += len(reqs)

    # ---------------------------------------------------------------- loop
    def _free_slot(self, lane):
        """A stream with room for a launch of ``lane`` ("large": a QFIM that
        fills the GPU on its own; "qfim": any other QFIM; "state"), or None.
        Kinds do not share a stream, so a state batch never waits in stream
        order behind a Jacobian that takes hundreds of milliseconds:
        - one stream: everything;
        - two: QFIMs (large ones one at a time) on the first, states on the second;
        - three or more: large QFIMs one at a time on the first (they neither
          compete with each other nor scatter their multi-GiB buffers over
          several streams' allocator pools), other QFIMs on the second, and
          states round-robin over the rest."""
        S = len(self.slots)
        if S == 1:
            cands = self.slots
        elif lane == "state":
            cands = self.slots[1:] if S == 2 else self.slots[2:]
        elif lane == "large" or S == 2:
            cands = self.slots[:1]
        else:
            cands = self.slots[1:2]
        limit = 1 if lane == "large" else self.max_inflight
        if lane == "large":
            s = cands[0]
            return s if len(s.inflight) < limit else None
        for k in range(len(cands)):
            start = self._next_slot.get(lane, 0)
            s = cands[(start + k) % len(cands)]
            if len(s.inflight) < limit:
                self._next_slot[lane] = (start + k + 1) % len(cands)
                return s
        return None

    def _lane(self, req):
        if req.kind != QFIM:
            return "state"
        return "large" if self.is_large(req) else "qfim"

    def is_large(self, req):
        """A request whose launch alone fills the GPU (a QFIM whose Jacobian
        exceeds large_bytes)."""
        return req.kind == QFIM and self.request_bytes(req) >= self.large_bytes

    def _mergeable(self, key):
        """A state class that runs on the interpreted kernels (not yet or not
        to be specialised): it joins the multi-plan launch."""
        n, depth, kind = key
        if kind != STATE_MW or not self.merge_classes:
            return False
        c = self.circuits.get((n, depth))
        return c is None or c._jit.get("plan") is None

    def _admit_merged(self, now):
        """One multi-plan launch (pqs_run_multi) of the oldest admissible
        state requests of every mergeable class, widths and depths mixed."""
        slot = self._free_slot("state")
        if slot is None:
            return 0
        take, need = [], 0
        for key in self._order:
            if not self._mergeable(key) or not self.queues[key] or not self._ready(key, now):
                continue
            q = self.queues[key]
            while q and len(take) < self.max_batch:
                b = self.request_bytes(q[0])
                if self.in_use + need + b > self.budget:
                    break
                take.append(q.popleft())
                need += b
            if len(take) >= self.max_batch:
                break
        if not take:
            return 0
        self._n_pending -= len(take)
        self._order_dirty = True
        self.in_use += need
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        self._launch_merged(slot, take, need, now)
        return 1

    def _admit_and_launch(self, now):
        launched = 0
        if self._n_pending == 0:
            return 0
        # oldest head-of-queue first (re-sorted only when the queues change)
        if self._order_dirty:
            self._order = [k for _, k in sorted((q[0].arrival, k) for k, q in self.queues.items() if q)]
            self._order_dirty = False
        if self.merge_classes:
            launched += self._admit_merged(now)
        for key in self._order:
            if self._mergeable(key):
                continue
            q = self.queues[key]
            if not q or not self._ready(key, now):
                continue
            slot = self._free_slot(self._lane(q[0]))
            if slot is None:
                continue
            q = self.queues[key]
            take, need = [], 0
            while q and len(take) < self.max_batch:
                b = self.request_bytes(q[0])
                if self.in_use + need + b > self.budget:
                    break
                take.append(q.popleft())
                need += b
                if key[2] == QFIM:
                    break  # one Jacobian per launch
            if not take:
                continue
            self._n_pending -= len(take)
            self._order_dirty = True
            self.in_use += need
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self._launch(slot, key, take, need, now)
            launched += 1
        return launched

    def _launch(self, slot, key, reqs, need, now):
        n, depth, kind = key
        c = self.circuit(n, depth)
        which = "gplan" if kind == QFIM else "plan"
        seen = self._seen[key] = self._seen.get(key, 0) + len(reqs)
        if seen >= 2 and n >= (self.spec_qfim if kind == QFIM else self.spec_state):
            # a class that repeats: its compile is likely to pay back
            c.specialise(which, self._pool, self.compile_workers)
            self._compiling.add((n, depth))
        b = _Batch()
        b.reqs, b.bytes, b.slot = reqs, need, slot
        b.out = b.q = b.F = None
        try:
            with E.using(slot.ctx, slot.stream), E.marked(f"serve {kind} n={n} depth={depth} batch={len(reqs)}"):
                if kind == STATE_MW:
                    params = E.to_device(np.stack([r.params for r in reqs]), np.float64)
                    b.out = E.run(c.best("plan"), params, c.init)
                    _, b.q = E.mw(b.out, n)
                else:
                    r = reqs[0]
                    P, dim = 2 * r.n * r.depth, 1 << r.n
                    buf = self._jscratch.get(id(slot))
                    if buf is None or buf.numel() < P * dim:
                        # grown, never shrunk: launches on one stream reuse it in order
                        buf = self._jscratch[id(slot)] = torch.empty(P * dim, dtype=E.CDTYPE[self.dtype],
                                                                     device=self.dev)
                    J = buf[:P * dim].view(P, dim)
                    psi, J = E.jacobian(c.best("gplan"), E.to_device(r.params, np.float64), c.init, jac_out=J)
                    b.F = E.qfim(J, psi)
                    b.out = psi
                    del J
                b.event = torch.cuda.Event()
                b.event.record(slot.stream)
        except Exception as err:  # a launch-time failure ends this batch only
            self._fail(reqs, need, err, now)
            return
        for r in reqs:
            self.results[r.rid].launched = now
        slot.inflight.append(b)
        self._n_inflight += 1
        self.n_batches += 1
        self.batch_sizes.append(len(reqs))

    def _launch_merged(self, slot, reqs, need, now):
        """State requests of several classes: one pqs_run_multi over all of
        them (rows grouped per class), then Meyer-Wallach per class."""
        groups = OrderedDict()
        for r in reqs:
            groups.setdefault((r.n, r.depth), []).append(r)
        order = [r for g in groups.values() for r in g]
        for key, rs in groups.items():
            seen = self._seen[key + (STATE_MW,)] = self._seen.get(key + (STATE_MW,), 0) + len(rs)
            if seen >= 2 and key[0] >= self.spec_state:
                self.circuit(*key).specialise("plan", self._pool, self.compile_workers)
                self._compiling.add(key)
        b = _Batch()
        b.reqs, b.bytes, b.slot = order, need, slot
        b.F = None
        try:
            with E.using(slot.ctx, slot.stream), E.marked(f"serve merged classes={len(groups)} batch={len(order)}"):
                jobs = []
                for (n, depth), rs in groups.items():
                    c = self.circuit(n, depth)
                    jobs.append((c.plan, E.to_device(np.stack([r.params for r in rs]), np.float64), c.init))
                outs = E.run_multi(jobs)
                qs = []
                b.out = []
                for ((n, _), rs), out in zip(groups.items(), outs):
                    qs.append(E.mw(out, n)[1])
                    b.out += [out[i] for i in range(len(rs))]
                b.q = torch.cat(qs)
                b.event = torch.cuda.Event()
                b.event.record(slot.stream)
        except Exception as err:  # a launch-time failure ends this batch only
            self._fail(order, need, err, now)
            return
        for r in order:
            self.results[r.rid].launched = now
        slot.inflight.append(b)
        self._n_inflight += 1
        self.n_batches += 1
        self.batch_sizes.append(len(order))

    def _retire(self, now):
        done = 0
        for slot in self.slots:
            while slot.inflight and slot.inflight[0].event.query():
                b = slot.inflight.popleft()
                self._n_inflight -= 1
                kind = b.reqs[0].kind
                group = self._next_group
                self._next_group += 1
                held_total, n_held = 0, 0
                if kind == STATE_MW:
                    qs = b.q.cpu().numpy()
                    for i, r in enumerate(b.reqs):
                        res = self.results[r.rid]
                        res.done = now
                        cancelled = r.rid in self._cancel_running
                        if cancelled:
                            self._cancel_running.discard(r.rid)
                            res.status = CANCELLED
                        else:
                            res.Q = float(qs[i])
                            res.status = OK
                        if self.keep_states:
                            # a cancelled row stays part of the batch tensor:
                            # its bytes return with the group's last row
                            held = self.amp * (1 << r.n)
                            held_total += held
                            self.in_use -= self.request_bytes(r) - held
                            if not cancelled:
                                res.state = b.out[i]
                                self._held[r.rid] = group
                                n_held += 1
                        else:
                            self.in_use -= self.request_bytes(r)
                elif b.reqs[0].rid in self._cancel_running:
                    r = b.reqs[0]
                    self._cancel_running.discard(r.rid)
                    res = self.results[r.rid]
                    res.status, res.done = CANCELLED, now
                    self.in_use -= self.request_bytes(r)
                else:
                    r = b.reqs[0]
                    res = self.results[r.rid]
                    res.qfim = b.F.cpu().numpy()
                    res.done = now
                    res.status = OK
                    if self.keep_states:
                        res.state = b.out
                        held = self.amp * (1 << r.n)
                        self._held[r.rid] = group
                        held_total += held
                        n_held += 1
                        self.in_use -= self.request_bytes(r) - held
                    else:
                        self.in_use -= self.request_bytes(r)
                if n_held:
                    self._groups[group] = [n_held, held_total]
                else:
                    self.in_use -= held_total  # every row cancelled: the tensor goes now
                done += len(b.reqs)
        return done

    def _load_compiled(self):
        for key in list(self._compiling):
            c = self.circuits[key]
            c.poll()
            if not c._pending:
                self._compiling.discard(key)

    def close(self):
        self._prep_pool.shutdown(wait=True)
        self._pool.shutdown(wait=True)
        self._jscratch.clear()

    def step(self, now):
        """One turn of the loop at server time ``now``: retire finished
        batches, load kernels compiled in the background, admit and launch.
        Returns (requests that reached a final status, batches launched)."""
        done = self._retire(now) + self._settle()
        self._load_compiled()
        launched = self._admit_and_launch(now)
        return done + self._settle(), launched

    def run(self, requests, realtime=True, clock=time.perf_counter, progress=None, progress_every=1000):
        """Serve ``requests`` (sorted by arrival).  realtime: a request
        becomes visible at its arrival time on the server clock; otherwise
        all are visible at once.  progress(finished, total, seconds) is
        called every ``progress_every`` finished requests (it may call
        ``cancel``).  Returns the results in request order; each has a
        status: "ok", "cancelled" or "error" (a launch that raised fails its
        own batch only; a GPU fault is not recoverable in the process)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        i, total, finished = 0, len(reqs), 0
        while finished < total:
            now = clock() - t0
            while i < total and (not realtime or reqs[i].arrival <= now):
                self.submit(reqs[i])
                i += 1
            done, launched = self.step(now)
            if progress is not None and done and (finished + done) // progress_every > finished // progress_every:
                progress(finished + done, total, now)
            finished += done + self._settle()
            if self.pending() and not self.inflight() and not done and not launched and not self._preparing():
                # nothing running, nothing being prepared and nothing fits:
                # released states would have to make room
                raise RuntimeError("HBM budget exhausted by resident states; release results")
            if not done and not launched:
                time.sleep(50e-6)  # idle: yield the host thread briefly
        self.wall = clock() - t0
        return [self.results[r.rid] for r in requests]


def stats(results, wall):
    lat = np.array([r.latency for r in results])
    return {"requests": len(results), "wall_s": wall, "requests_per_s": len(results) / wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}


def run_one_at_a_time(requests, dtype="complex128", circuits=None):
    """The same jobs, one request per launch sequence and a device
    synchronisation after each (the unbatched comparator).  With a server's
    ``circuits`` it uses the same kernels the server had specialised.
    Returns (results, wall seconds)."""
    dev = E.device()
    dt = E.dtype_code(dtype)
    circuits = {} if circuits is None else circuits
    out = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in requests:
        c = circuits.get((r.n, r.depth))
        if c is None:
            c = circuits[(r.n, r.depth)] = _Circuit(r.n, r.depth, dt)
        res = Result(r)
        p = torch.as_tensor(r.params, device=dev)
        if r.kind == STATE_MW:
            st = E.run(c.best("plan"), p, c.init)
            _, q = E.mw(st, r.n)
            res.state, res.Q = st[0], float(q.item())
        else:
            psi, J = E.jacobian(c.best("gplan"), p, c.init)
            res.qfim = E.qfim(J, psi).cpu().numpy()
            res.state = psi
        torch.cuda.synchronize()
        res.done = time.perf_counter() - t0
        out.append(res)
    return out, time.perf_counter() - t0
