"""``PQC``: the reference's circuit container, evolved by the HIP engine.

Reference: pyramaterised/circuit.py.  The layer / gate bookkeeping, the
parameter binding order and the random-angle stream are the reference's
(circuit.py:16-116; gates.py:10).  What changes is state generation: instead
of ``for g in self.gates: circuit_state = g * circuit_state`` over full QuTiP
operators (circuit.py:118-125), the flat gate list is compiled once into a
plan (libpqs ``pqs_plan_create``) and every ``run`` is one batched GPU call;
``get_gradients`` yields all derivative states in one forward pass
(``pqs_jacobian``) instead of one full re-simulation per parameter
(circuit.py:149-192).
"""
from copy import deepcopy

import numpy as np
import torch

from . import engine as E
from .gates import *  # noqa: F401,F403  (circuit.py:9 re-exports the gate library)
from . import gates as _G
from .qobj import Ket, PauliSum, ProductKet, Qobj, basis, expect, tensor


class PQC:
    """An n-qubit, p-layer parameterised quantum circuit (circuit.py:12-198)."""

    def __init__(self, n_qubits, dtype="complex128", tile_bits=0, low_bits=0, block=0, jit="auto", flags=0):
        self.n_qubits = n_qubits
        self.n_layers = 0
        self.layers = []
        self.gates = []
        self.parameterised = []
        self.n_params = 0
        if n_qubits >= 2:
            self.set_H("ZZ")
        self.initial_state = tensor([basis(2, 0) for _ in range(self.n_qubits)])
        self.state = self.initial_state
        self.dtype = E.dtype_code(dtype)
        self.plan_opts = dict(tile_bits=tile_bits, low_bits=low_bits, block=block, flags=flags)
        # run-time specialised sweep kernels: "auto" uses them for batches and
        # large states, where their one-off compile pays back; PQS_JIT=0/1
        # overrides
        self.jit = jit
        self._angles = None
        self._plans = {}
        self._plan_uses = {}
        self.set_gates()

    # ------------------------------------------------------------ set-up
    def set_H(self, H):
        """circuit.py:25-33: 'ZZ' -> Z_0 Z_1, else a user operator.  A dense
        operator is decomposed into Pauli strings once, on first use
        (``hamiltonian``), not at every ``cost``."""
        if isinstance(H, str) and H == "ZZ":
            s = [0] * self.n_qubits
            s[0] = s[1] = 3
            self.H = PauliSum(self.n_qubits, {tuple(s): 1.0})
        else:
            self.H = H
        self._H_ps = (self.H, self.H) if isinstance(self.H, PauliSum) else None

    def hamiltonian(self):
        """self.H as a PauliSum (decomposed once per operator object)."""
        c = getattr(self, "_H_ps", None)
        if c is None or c[0] is not self.H:
            c = self._H_ps = (self.H, self.H if isinstance(self.H, PauliSum) else PauliSum.from_dense(self.H))
        return c[1]

    def set_initial_state(self, state):
        """circuit.py:35-36: tensor([state] * n)."""
        self.initial_state = tensor([state for _ in range(self.n_qubits)])

    def add_layer(self, layer, n=1):
        """circuit.py:38-43."""
        for _ in range(n):
            self.layers.append(deepcopy(layer))
        self.n_layers += n
        self.set_gates()

    def add_layers(self, layers, copy=True):
        """Append several layers with one gate-list rebuild (``add_layer``
        per layer rebuilds the list each time).  copy=False keeps the given
        gate objects instead of deep copies: only for layers whose gate
        objects are all distinct and held by nothing else (a template's
        list repeats one layer object, so it needs the copies)."""
        layers = list(layers)
        self.layers.extend(deepcopy(l) for l in layers) if copy else self.layers.extend(layers)
        self.n_layers += len(layers)
        self.set_gates()

    def set_layer(self, layer, pos):
        """circuit.py:45-48."""
        self.layers[pos] = deepcopy(layer)
        self.set_gates()

    def get_layer(self, pos):
        return self.layers[pos]

    def set_gates(self):
        """circuit.py:53-72, including its bookkeeping: ``parameterised`` gets a
        -1 after every gate (the for-else always runs) and ``n_params`` counts
        each parameter twice.  The true parameter count is ``len(self.slots)``."""
        layers = []
        for i in self.layers:
            layers = layers + i
        self.gates = layers
        self.parameterised = []
        total = 0
        for gate in self.gates:
            cur = gate.param_count
            total += cur
            for _ in range(cur):
                total += 1
                self.parameterised.append(total)
            else:
                self.parameterised.append(-1)
        self.n_params = total
        # slot of every parameterised gate, in binding order (circuit.py:89-116)
        self.slots = []
        self._gate_slot = {}
        k = 0
        for g in self.gates:
            if g.is_param:
                self._gate_slot[id(g)] = k
                self.slots.append(g)
                k += g.param_count
        self.n_true_params = k
        self._plans = {}
        self._plan_uses = {}

    # ------------------------------------------------------------ parameters
    def get_params(self):
        """circuit.py:74-84."""
        angles = []
        for g in self.gates:
            if g.param_count == 2:
                angles.append(g.theta)
                angles.append(g.phi)
            elif g.param_count == 1:
                angles.append(g.theta)
        return angles

    def set_params(self, angles):
        """circuit.py:86-116: explicit list, or "random" draws of
        gates.rng.random(1)[0] * 2 * pi in binding order."""
        parameterised = [g for g in self.gates if g.is_param]
        vec = np.zeros(self.n_true_params)
        k = 0
        for g in parameterised:
            if g.param_count == 2:
                if type(angles) != str:
                    a1, a2 = angles[k], angles[k + 1]
                elif angles == "random":
                    a1 = _G.rng.random(1)[0] * 2 * np.pi
                    a2 = _G.rng.random(1)[0] * 2 * np.pi
                else:
                    raise Exception("No parameters supplied!")
                g.set_theta(a1)
                g.set_phi(a2)
                vec[k], vec[k + 1] = a1, a2
                k += 2
            elif g.param_count == 1:
                if type(angles) != str:
                    a1 = angles[k]
                elif angles == "random":
                    a1 = _G.rng.random(1)[0] * 2 * np.pi
                else:
                    raise Exception("No parameters supplied!")
                g.set_theta(a1)
                vec[k] = a1
                k += 1
        self._angles = vec

    def random_params(self, n_samples):
        """n_samples parameter vectors from gates.rng: the same numbers that
        n_samples successive ``run("random")`` calls would draw."""
        P = self.n_true_params
        return _G.rng.random(n_samples * P).reshape(n_samples, P) * 2 * np.pi

    # ------------------------------------------------------------ compilation
    def ops(self, with_generators=False):
        """Flat primitive-op list in gate order; with generators, each
        parameterised gate's derivative op(s) follow its primitives.  A gate
        whose derivative is not a generator after it (a commute=False
        shared_parameter block, gates.py:458-466) gets none here: its
        tangent comes from ``special_ops``."""
        out = []
        for g in self.gates:
            slot = self._gate_slot.get(id(g), -1)
            out += g.prim_ops(max(slot, 0))
            if with_generators and g.param_count > 0 and not getattr(g, "special_derivative", False):
                for w in range(g.param_count):
                    out += g.gen_ops(slot + w, slot, which=w + 1)
        return out

    def special_slots(self):
        """Parameter slots whose derivative needs a plan of its own."""
        return [self._gate_slot[id(g)] for g in self.gates
                if g.param_count > 0 and getattr(g, "special_derivative", False)]

    def special_ops(self, slot):
        """The circuit with the derivative of the gate at ``slot`` inserted
        as the reference forms it (take_derivative, circuit.py:149-172, with
        a commute=False block's derivative, gates.py:458-466): the primitive
        ops of every gate, and after that gate its ``derivative_ops`` as
        tangent 0.  Its tangent-0 state is d psi / d theta_slot."""
        out = []
        for g in self.gates:
            s = self._gate_slot.get(id(g), -1)
            out += g.prim_ops(max(s, 0))
            if s == slot and g.param_count > 0:
                out += g.derivative_ops(0, s)
        return out

    def _special_plan(self, slot, dtype=None):
        dt = self.dtype if dtype is None else E.dtype_code(dtype)
        key = ("special", slot, dt)
        p = self._plans.get(key)
        if p is None:
            p = E.Plan(self.n_qubits, self.special_ops(slot), dtype=dt, **self.plan_options(True, dt))
            self._plans[key] = p
        return p

    # with jit="auto", a plan used this many times gets specialised kernels
    # even for small batches (repeated QFIMs, training loops)
    JIT_AFTER_USES = 3

    def plan(self, with_generators=False, dtype=None, batch=1, pairs=False):
        """The plan of this circuit (cached per generator mode and dtype);
        ``batch`` and ``pairs`` (the caller runs pair batches,
        ``run_batch_pairs``) describe the use the tile-shape tuner times."""
        dt = self.dtype if dtype is None else E.dtype_code(dtype)
        key = (with_generators, dt)
        p = self._plans.get(key)
        uses = self._plan_uses[key] = self._plan_uses.get(key, 0) + 1
        want = self._want_jit(batch)
        if (p is not None and not want and not getattr(p, "jitted", False) and self.jit == "auto"
                and uses >= self.JIT_AFTER_USES and self._jit_env() is None):
            p.jit()  # the same plan, specialised in place (no tile-shape tuning)
            return p
        if p is not None and (getattr(p, "jitted", False) or not want):
            return p
        ops = self.ops(with_generators)
        if want and self._want_tune(with_generators):
            # specialised kernels: time the candidate tile shapes once
            p = E.tuned_plan(self.n_qubits, ops, dt, batch, self.init(), pairs=pairs and not with_generators)
        else:
            if p is None:
                p = E.Plan(self.n_qubits, ops, dtype=dt, **self.plan_options(with_generators, dt))
            if want:
                p.jit()
        self._plans[key] = p
        return p

    # Jacobian plans use rolling layouts (fewer sweeps) when the ping-pong
    # rows they need beside the Jacobian stay within this many bytes
    ROLLING_JACOBIAN_BYTES = 4 << 30

    def rolling_jacobian(self, dtype=None):
        """Whether derivative plans use rolling layouts: each Jacobian row
        then needs a workspace row of the same size."""
        dt = self.dtype if dtype is None else E.dtype_code(dtype)
        amp = 16 if dt == E.L.PQS_C128 else 8
        return self.n_true_params * (amp << self.n_qubits) <= self.ROLLING_JACOBIAN_BYTES

    def plan_options(self, with_generators=False, dtype=None):
        """Plan options for an untuned plan: the constructor's, or for
        derivative plans 12-bit c128 / 13-bit c64 tiles with 128 B runs,
        which keep two workgroups per CU and leave the planner room
        (measured at 20 qubits, P = 1600: 332 ms against 427 ms for 13-bit
        tiles with 5 fixed low qubits), with rolling layouts when their
        ping-pong rows fit (rolling_jacobian)."""
        if any(self.plan_opts.values()) or not with_generators:
            return dict(self.plan_opts)
        dt = self.dtype if dtype is None else E.dtype_code(dtype)
        opts = dict(self.plan_opts)
        if dt == E.L.PQS_C128:
            opts.update(tile_bits=min(12, self.n_qubits), low_bits=3)
        else:
            opts.update(tile_bits=min(13, self.n_qubits), low_bits=4)
        if self.rolling_jacobian(dt):
            opts.update(flags=opts.get("flags", 0) | E.L.PQS_PLAN_ROLLING)
        return opts

    def _want_tune(self, with_generators):
        """Tile shapes are tuned unless the caller fixed them (constructor
        options) or PQS_TUNE=0; generator plans keep the default shape."""
        import os
        if os.environ.get("PQS_TUNE", "1") in ("0", "", "false", "False"):
            return False
        return not with_generators and not any(self.plan_opts.values())

    @staticmethod
    def _jit_env():
        import os
        return os.environ.get("PQS_JIT")

    def _want_jit(self, batch):
        env = self._jit_env()
        if env is not None:
            return env not in ("0", "", "false", "False")
        if self.jit == "auto":
            return batch >= 32 or self.n_qubits >= 20
        return bool(self.jit)

    def init(self):
        st = self.initial_state
        if isinstance(st, ProductKet) or getattr(st, "factors", None) is not None:
            return E.Init.product(st.factors)
        if isinstance(st, Ket):
            return E.Init.from_vector(st.t)
        return E.Init.from_vector(np.asarray(st.full()).reshape(-1))

    def _params_vec(self):
        P = self.n_true_params
        return self._angles if self._angles is not None and len(self._angles) == P else \
            np.array(self._gate_angles())

    def _gate_angles(self):
        """Raw bound angles from the gate objects (inverting the stored
        transforms of negative_R_z / offset_R_z)."""
        out = []
        for g in self.gates:
            if not g.is_param:
                continue
            th, ph = g._angles()
            out.append(th)
            if g.param_count == 2:
                out.append(ph)
        return out

    # ------------------------------------------------------------ dense gates
    # A gate with no local form (ARBGATE: a dense 2^N operator) splits the
    # circuit into segments of local gates, each run as its own plan, with
    # the dense operator applied on the GPU in between (pqs_apply_dense).
    def has_dense(self):
        return any(getattr(g, "dense", False) for g in self.gates)

    def _segments(self):
        """[(local gates, dense gate or None)] in gate order."""
        segs, cur = [], []
        for g in self.gates:
            if getattr(g, "dense", False):
                segs.append((cur, g))
                cur = []
            else:
                cur.append(g)
        segs.append((cur, None))
        return segs

    def _segment_plan(self, i, gates, with_generators, dt=None):
        dt = self.dtype if dt is None else dt
        key = ("segment", i, with_generators, dt)
        p = self._plans.get(key)
        if p is None:
            ops = []
            for g in gates:
                slot = self._gate_slot.get(id(g), -1)
                ops += g.prim_ops(max(slot, 0))
                if with_generators and g.param_count > 0:
                    if getattr(g, "special_derivative", False):
                        raise NotImplementedError("commute=False blocks that are not generator sums, in a circuit "
                                                  "with a dense gate")
                    for w in range(g.param_count):
                        ops += g.gen_ops(slot + w, slot, which=w + 1)
            p = E.Plan(self.n_qubits, ops, dtype=dt, **self.plan_options(with_generators, dt))
            self._plans[key] = p
        return p

    def _run_dense(self, params):
        """States (B, 2^N) of a circuit with dense gates, one row per
        parameter row."""
        params = np.atleast_2d(np.asarray(params, dtype=np.float64))
        B = params.shape[0]
        states = None
        for i, (gates, dense) in enumerate(self._segments()):
            if gates:
                plan = self._segment_plan(i, gates, False)
                if states is None:
                    states = E.run(plan, params, self.init())
                else:
                    E.run_rows(plan, params, states)
            if dense is not None:
                if states is None:
                    init = self.init()
                    states = E.run(E.Plan(self.n_qubits, [], dtype=self.dtype), np.zeros((B, 1)), init)
                t = self._gate_slot[id(dense)]
                for b in range(B):
                    states[b] = E.apply_dense(dense.matrix(params[b, t]), states[b])
        return states

    def _jacobian_dense(self):
        """(psi, J) of a circuit with dense gates, forward through the
        segments: rows spawned in a segment (pqs_jacobian from the state at
        its start) are carried through every later dense gate (U) and
        segment (plain evolution of the rows); a dense gate's own row is
        D U psi (gates.py:427-429)."""
        vec = self._params_vec()
        P = self.n_true_params
        dim = 2 ** self.n_qubits
        dev = E.device()
        cdt = E.CDTYPE[self.dtype]
        J = torch.zeros((P, dim), dtype=cdt, device=dev)
        live = []
        psi = None
        for i, (gates, dense) in enumerate(self._segments()):
            if gates:
                if live:
                    rows = J[live].contiguous()
                    E.run_rows(self._segment_plan(i, gates, False), vec, rows)
                    J[live] = rows
                init = self.init() if psi is None else E.Init.from_vector(psi)
                psi, Jseg = E.jacobian(self._segment_plan(i, gates, True), vec, init)
                for g in gates:
                    if g.param_count > 0:
                        t0 = self._gate_slot[id(g)]
                        for t in range(t0, t0 + g.param_count):
                            J[t] = Jseg[t]
                            live.append(t)
            if dense is not None:
                if psi is None:
                    psi = E.run(E.Plan(self.n_qubits, [], dtype=self.dtype), np.zeros(1), self.init())[0]
                t = self._gate_slot[id(dense)]
                th = vec[t]
                if live:
                    J[live] = E.apply_dense(dense.matrix(th), J[live].contiguous())
                J[t] = E.apply_dense(dense.derivative_matrix(th), psi)
                live.append(t)
                psi = E.apply_dense(dense.matrix(th), psi)
        return psi, J

    # ------------------------------------------------------------ evolution
    def run(self, angles):
        """circuit.py:118-125: |psi> = U_G ... U_1 |psi0> (one GPU call)."""
        self.set_params(angles=angles)
        if self.has_dense():
            return Ket(self._run_dense(self._params_vec())[0], self.n_qubits)
        states = E.run(self.plan(), self._params_vec(), self.init())
        return Ket(states[0], self.n_qubits)

    def run_batch(self, params):
        """Evolve many parameter vectors at once: (S, P) -> (S, 2^N) tensor.
        The gate objects keep the angles of the last row (as after S calls of
        ``run``)."""
        params = np.asarray(params, dtype=np.float64) if not torch.is_tensor(params) else params
        if len(params) == 0:
            return torch.empty((0, 2 ** self.n_qubits), dtype=E.CDTYPE[self.dtype], device=E.device())
        last = params[-1].tolist() if not torch.is_tensor(params) else params[-1].cpu().tolist()
        self.set_params(last)
        if self.has_dense():
            return self._run_dense(params.cpu().numpy() if torch.is_tensor(params) else params)
        return E.run(self.plan(batch=len(params)), params, self.init())

    def run_batch_pairs(self, params):
        """Evolve an even number of parameter vectors (S, P) and return
        (states (S, 2^N), overlaps (S/2,) complex128), overlaps[i] =
        <state_2i|state_2i+1>: the fidelity samples of independent pairs
        (measure.py:123-137) without reading the states back where the plan's
        sweeps are specialised.  The gate objects keep the last row's angles."""
        params = np.asarray(params, dtype=np.float64) if not torch.is_tensor(params) else params
        if len(params) % 2:
            raise ValueError("run_batch_pairs needs an even number of parameter rows")
        if len(params) == 0:
            return (torch.empty((0, 2 ** self.n_qubits), dtype=E.CDTYPE[self.dtype], device=E.device()),
                    torch.empty(0, dtype=torch.complex128, device=E.device()))
        last = params[-1].tolist() if not torch.is_tensor(params) else params[-1].cpu().tolist()
        self.set_params(last)
        if self.has_dense():
            st = self._run_dense(params.cpu().numpy() if torch.is_tensor(params) else params)
            S = st.shape[0]
            return st, E.inner_pairs(st, np.arange(0, S, 2), np.arange(1, S, 2))
        return E.run_pairs(self.plan(batch=len(params), pairs=True), params, self.init())

    def update_state(self, angles):
        """circuit.py:127-130."""
        self.state = self.run(angles)
        return self.state

    def cost(self, angles):
        """circuit.py:132-137: <psi|H|psi>."""
        self.state = self.run(angles=angles)
        return expect(self.hamiltonian(), self.state)

    def fidelity(self, target_state):
        """circuit.py:139-142: |<psi|phi>|^2."""
        return float(np.abs(self.state.overlap(target_state)) ** 2)

    def flip_deriv(self):
        """circuit.py:144-147."""
        for g in self.gates:
            if g.is_param:
                g.flip_pauli()

    # ------------------------------------------------------------ derivatives
    def _tangent_of(self, g_on, param):
        for i, g in enumerate(self.gates):
            if g is g_on:
                slot = self._gate_slot.get(id(g), -1)
                if slot < 0:
                    raise ValueError("gate has no parameter")
                return slot + (param - 1 if param >= 1 else 0)
        raise ValueError("gate is not in this circuit")  # list.index (circuit.py:156)

    def take_derivative(self, g_on, param=0):
        """circuit.py:149-172: the circuit with gate g_on replaced by D*gate,
        run on the initial state (one GPU run with the generator inserted)."""
        t = self._tangent_of(g_on, param)
        if self.has_dense():
            return Ket(self._jacobian_dense()[1][t].clone(), self.n_qubits)
        plan, tg = self._tangent_plan(t)
        out = E.run(plan, self._params_vec(), self.init(), tangents=[tg])
        return Ket(out[0], self.n_qubits)

    def _tangent_plan(self, t):
        """(plan, tangent index in it) that yields d psi / d theta_t."""
        if t in self.special_slots():
            return self._special_plan(t), 0
        return self.plan(with_generators=True), t

    def get_gradients(self, method="forward", compat=False):
        """circuit.py:174-192: every derivative state.  ``forward`` spawns all
        tangents from one forward pass (pqs_jacobian); ``plain`` re-simulates
        the circuit once per parameter, as the reference does.

        compat=True reproduces the reference's two-parameter branch
        (circuit.py:186-191): the second row of the two-parameter gate at
        index ``count`` of the parameterised gates is taken of
        ``self.gates[count]``, the gate at that position of the circuit,
        with its second-parameter derivative (``parameterised_derivative(2)``,
        which a one-parameter gate answers with its derivative and a fixed
        gate with the identity).  Any parameterised gate there other than
        the fSim itself was handled earlier in the loop, so the reference
        never fails on it; the row is then that gate's derivative state (the
        state itself for a fixed gate).  Without compat, the second row is
        the fSim's own phi derivative."""
        psi, J = self.jacobian(method=method)
        rows = [J[i] for i in range(J.shape[0])]
        if compat:
            params = [g for g in self.gates if g.param_count > 0]
            for count, g in enumerate(params):
                if g.param_count != 2:
                    continue
                x = self.gates[count]
                if x is g:
                    continue
                row = self._gate_slot[id(g)] + 1
                if x.param_count == 2:
                    rows[row] = J[self._gate_slot[id(x)] + 1]
                elif x.param_count == 1:
                    rows[row] = J[self._gate_slot[id(x)]]
                else:
                    rows[row] = psi.reshape(-1)
        return [Ket(r, self.n_qubits) for r in rows]

    def jacobian(self, method="forward"):
        """(psi, J) with J[t] = d psi / d theta_t as tensors.  Rows of
        commute=False blocks whose derivative is not a generator sum come
        from a plan of their own each (``special_ops``)."""
        if self.has_dense():
            if method not in ("forward", "plain"):
                raise ValueError(f"unknown method {method!r}")
            return self._jacobian_dense()
        plan = self.plan(with_generators=True, batch=self.n_true_params)
        specials = self.special_slots()
        if method == "forward":
            J = None
            if specials:  # the last rows may have no generator in the main plan
                J = torch.empty((self.n_true_params, 2 ** self.n_qubits), dtype=E.CDTYPE[plan.dtype],
                                device=E.device())
            psi, J = E.jacobian(plan, self._params_vec(), self.init(), jac_out=J)
            for t in specials:
                E.jacobian(self._special_plan(t), self._params_vec(), self.init(), jac_out=J[t:t + 1])
            return psi, J
        if method == "plain":
            P = self.n_true_params
            dev = E.device()
            J = torch.empty((P, 2 ** self.n_qubits), dtype=E.CDTYPE[plan.dtype], device=dev)
            for t in range(P):
                tp, tg = self._tangent_plan(t)
                E.run(tp, self._params_vec(), self.init(), tangents=[tg], out=J[t:t + 1])
            psi = E.run(self.plan(), self._params_vec(), self.init())[0]
            return psi, J
        raise ValueError(f"unknown method {method!r}")

    def __repr__(self):
        line = f"A {self.n_qubits} qubit, {self.n_layers} layer deep PQC. \n"
        for count, l in enumerate(self.layers):
            line += f"Layer {count}: {l} \n"
        return line


__all__ = [n for n in dir(_G) if not n.startswith("_")] + ["PQC", "Qobj", "Ket", "PauliSum", "tensor", "basis"]
