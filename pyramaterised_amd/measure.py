"""``Measurements``: capacity measures of a PQC on the GPU.

Reference: pyramaterised/measure.py.  The formulas, the histogram binning and
the Haar reference density are the reference's; the arithmetic runs in libpqs:

* QFIM (measure.py:33-71): all derivative states from one forward pass,
  then one Gram reduction F = 4 Re(<J_p|J_q> - <J_p|psi><psi|J_q>).
* expressibility (measure.py:123-197): sample states evolved as one batch,
  pair fidelities from a Gram (all pairs) or paired overlaps, histogram on the
  GPU; the 75-or-fewer-bin KL divergence is evaluated on the host.
* Meyer-Wallach Q (measure.py:226-249): one reduction over each state for
  every qubit's 2x2 reduced density matrix.
Eigen-decompositions of the P x P QFIM (measure.py:73-99) run on the host
(scipy), as in the reference.
"""
import random
from itertools import combinations  # noqa: F401
from typing import Callable, Dict, Literal, Tuple, Union  # noqa: F401

import numpy as np
import scipy
import scipy.linalg
import scipy.optimize
import scipy.special
import torch

from . import engine as E
from .circuit import PQC  # noqa: F401
from .qobj import Ket, Qobj, to_ket


def _as_ket(s, n=None):
    return s if isinstance(s, Ket) else to_ket(s, n)


def _stack(kets):
    return torch.stack([_as_ket(k).t.reshape(-1) for k in kets])


def pair_indices(S):
    """(a, b) index arrays of itertools.combinations(range(S), 2), in order."""
    a, b = np.triu_indices(S, k=1)
    return a, b


class Measurements:
    def __init__(self, QC):
        self.QC = QC
        try:
            self.minimize_function = QC.cost
        except AttributeError:
            self.minimize_function = lambda x: x

    def set_minimise_function(self, function):
        self.minimize_function = function

    # ------------------------------------------------------------------ QFIM
    def get_QFI(self, grad_list=[]):
        """measure.py:33-71.  Uses QC.state for <d_p psi|psi> as the reference."""
        n_params = len([i for i in self.QC.parameterised if i > -1])
        if len(grad_list) == 0:
            _, J = self.QC.jacobian()
        else:
            J = _stack(grad_list)
        psi = _as_ket(self.QC.state, self.QC.n_qubits).t
        F = E.qfim(J[:n_params], psi)
        return F.cpu().numpy()

    def get_eigenvalues(self, QFI):
        return scipy.linalg.eigh(QFI)

    def get_effective_quantum_dimension(self, cutoff_eigvals):
        """measure.py:77-87."""
        QFI = self.get_QFI()
        eigvals, _ = scipy.linalg.eigh(QFI)
        return len(eigvals[eigvals > cutoff_eigvals])

    def new_measure(self, QFI=None):
        """measure.py:89-99."""
        if QFI is None:
            QFI = self.get_QFI()
        eigvals, _ = self.get_eigenvalues(QFI)
        return sum(1 if v > 1 else v for v in eigvals)

    def find_overparam_point(self, layer_index_list, epsilon=1e-3):
        """measure.py:101-121."""
        layers_to_add = [self.QC.get_layer(i) for i in layer_index_list]
        prev_rank, rank_diff, count = 0, 1, 0
        while rank_diff > epsilon and count < 1e6:
            for l in layers_to_add:
                self.QC.add_layer(l)
            self.QC.update_state("random")
            QFI = self.get_QFI()
            rank = np.linalg.matrix_rank(QFI)
            rank_diff = np.abs(rank - prev_rank)
            print(f"Iteration {count}, r0={prev_rank}, r1={rank}, delta = {rank_diff}")
            prev_rank = rank
            count += 1
        return count

    # -------------------------------------------------------- expressibility
    def sample_states(self, sample_N, params=None):
        """sample_N states of ``run("random")`` evolved as one batch -> (S, 2^N)."""
        if params is None:
            params = self.QC.random_params(sample_N)
        if hasattr(self.QC, "run_batch"):
            return self.QC.run_batch(params)
        return _stack([self.QC.run("random") for _ in range(sample_N)])

    def fidelities_all_pairs(self, states):
        """|<psi_a|psi_b>|^2 over combinations(range(S), 2), on the GPU."""
        S = states.shape[0]
        if S < 2:
            return torch.empty(0, dtype=torch.float64, device=states.device)
        G = E.gram(states, states)
        a, b = pair_indices(S)
        idx = torch.as_tensor(a * S + b, device=states.device)
        return E.abs2(G.reshape(-1)[idx])

    def fidelities_pairs(self, states_a, states_b):
        """|<a_i|b_i>|^2 for paired rows."""
        S = states_a.shape[0]
        both = torch.cat([states_a, states_b])
        return E.abs2(E.inner_pairs(both, np.arange(S), np.arange(S) + S))

    def _gen_f_samples(self, sample_N):
        """measure.py:123-137 (list of floats, combinations order)."""
        states = self.sample_states(sample_N)
        return self.fidelities_all_pairs(states).cpu().tolist()

    def _gen_histo(self, F_samples, filt=0, bins=None):
        """measure.py:139-159: bins = int(75/10000 * len(F)) on (0, 1) unless
        given; ``filt`` is ignored by the reference's histogram call."""
        if bins is None:
            bins = int((75 / 10000) * len(F_samples))
        edges = np.linspace(0.0, 1.0, bins + 1)
        F = F_samples if torch.is_tensor(F_samples) else torch.as_tensor(
            np.asarray(F_samples, dtype=np.float64), device=E.device())
        counts = E.histogram(F, edges).cpu().numpy()
        prob = counts / sum(counts)
        mids = np.array([(edges[i - 1] + edges[i]) / 2 for i in range(1, len(edges))])
        return prob, mids

    def expr(self, F_samples, N, filt=0, bins=None):
        """measure.py:161-180."""
        if len(F_samples) == 0:
            return 0
        P_pqc, F = self._gen_histo(F_samples, filt, bins=bins)
        haar = (N - 1) * ((1 - F) ** (N - 2))
        P_haar = haar / sum(haar)
        return np.sum(scipy.special.kl_div(P_pqc, P_haar))

    def expressibility(self, sample_N):
        """measure.py:182-197."""
        N = 2 ** self.QC.n_qubits
        F_samples = self.fidelities_all_pairs(self.sample_states(sample_N))
        return self.expr(F_samples, N)

    def expressibility_pairs(self, n_pairs, bins=75, params=None):
        """Expressibility from n_pairs independent state pairs (2*n_pairs
        sample states; pair i = states 2i, 2i+1) with an explicit bin count.
        Returns (expr, F samples tensor)."""
        if params is None:
            params = self.QC.random_params(2 * n_pairs)
        if hasattr(self.QC, "run_batch_pairs"):
            _, ov = self.QC.run_batch_pairs(params)
            F = E.abs2(ov)
        else:
            states = self.sample_states(2 * n_pairs, params)
            F = E.abs2(E.inner_pairs(states, np.arange(0, 2 * n_pairs, 2), np.arange(1, 2 * n_pairs, 2)))
        return self.expr(F, 2 ** self.QC.n_qubits, bins=bins), F

    def find_eff_H(self, circuit_f_samples, n):
        """measure.py:199-224 (for n > 10 the reference calls a method it does
        not define, log_expr)."""
        def wrapper(x, F_samples=[]):
            return self.expr(F_samples, x, filt=0.2)

        if n > 10:
            raise NotImplementedError("find_eff_H for n > 10 (reference calls an undefined log_expr)")
        out = scipy.optimize.minimize(wrapper, [4], args=(circuit_f_samples), method="BFGS")
        if out.success is True:
            return out.x[0]
        print(out)
        return 0

    # ---------------------------------------------------------- entanglement
    def single_Q(self, system, n):
        """measure.py:226-237 on the GPU."""
        ket = _as_ket(system, n)
        _, q = E.mw(ket.t, n)
        return float(q[0].item())

    def entanglement_batch(self, states, n=None):
        """Q for every row of a (S, 2^N) state tensor -> float64 tensor."""
        n = n or self.QC.n_qubits
        return E.mw(states, n)[1]

    def entanglement(self, sample_N):
        """measure.py:239-249."""
        states = self.sample_states(sample_N)
        return self.entanglement_batch(states).cpu().tolist()

    # ----------------------------------------------------------------- magic
    def theta_to_magic(self, angles):
        state = self.QC.run(angles=angles)
        return -1 * self.renyi_entropy_fast(state)

    def theta_to_gkp(self, angles):
        state = self.QC.run(angles=angles)
        return -1 * self.gkp_fast(state)

    # The reference evaluates the stabiliser Renyi sum as an O(4^n) dot
    # product over two precomputed 2^n x 2^n tables (measure.py:268-313).  The
    # engine's kernel (magic.py) needs neither; the table builders are kept
    # for callers that build, pass or inspect them.
    def numberToBase(self, n, b, n_qubits):
        """measure.py:268-277: the n_qubits base-b digits of n, most
        significant first; IndexError when n needs more digits."""
        digits = np.zeros(n_qubits, dtype=int)
        k = n_qubits
        while n:
            k -= 1
            if k < 0:
                raise IndexError(f"{n} has more than {n_qubits} digits in base {b}")
            digits[k] = int(n % b)
            n //= b
        return digits

    def get_conversion_matrix_mod_add_index(self, base_states):
        """measure.py:279-291: entry [j, k] is the index (most significant
        digit first) of base state k plus base state j, digit-wise mod 2."""
        B = np.asarray(base_states, dtype=int)
        weights = 2 ** np.arange(B.shape[1])[::-1]
        out = np.zeros((len(B), len(B)), dtype=int)
        for j in range(len(B)):
            out[j] = ((B + B[j]) % 2) @ weights
        return out

    def get_conversion_matrix_binary_prod(self, base_states):
        """measure.py:293-302: entry [i, k] is (-1)^(b_i . b_k mod 2)."""
        B = np.asarray(base_states, dtype=int)
        return 1 - 2 * ((B @ B.T) % 2)

    def get_conversion_matrices(self):
        """measure.py:304-313: (mod-add index table, sign table) for this
        circuit's n, each 2^n x 2^n in host memory."""
        n = self.QC.n_qubits
        base = [self.numberToBase(i, 2, n) for i in range(2 ** n)]
        return (self.get_conversion_matrix_mod_add_index(base), self.get_conversion_matrix_binary_prod(base))

    def set_converstion_matrices(self, conv_mats):
        self.conversion_matrices = conv_mats

    def renyi_entropy_fast(self, state, conversion_matrices=None, alpha=2):
        """measure.py:318-349: stabiliser Renyi entropy of order alpha."""
        from .magic import renyi_entropy
        return renyi_entropy(_as_ket(state), alpha)

    def entropy_of_magic(self, sample_N):
        """measure.py:351-359."""
        from .magic import renyi_entropy_batch
        states = self.sample_states(sample_N)
        return float(np.mean(renyi_entropy_batch(states, self.QC.n_qubits, 2)))

    def gkp_fast(self, state, conversion_matrices=None):
        """measure.py:361-368."""
        return 1 / (2 * np.log(2)) * self.renyi_entropy_fast(state, conversion_matrices, alpha=1 / 2)

    # ----------------------------------------------------------- combined
    def efficient_measurements(self, sample_N, measure_expr=True, measure_ent=True, measure_eom=True,
                               measure_GKP=True, full_data=False, angles="random"):
        """measure.py:370-459: one batch of states reused for every measure."""
        n = self.QC.n_qubits
        if sample_N == 0:
            measure_expr = measure_ent = measure_eom = measure_GKP = False
        if angles == "clifford":
            cl = (0, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi)
            init_angles = [[random.choice(cl) for _ in range(self.QC.n_params)] for _ in range(sample_N)]
            P = self.QC.n_true_params
            params = np.array([a[:P] for a in init_angles]) if sample_N else np.zeros((0, P))
            states = self.sample_states(sample_N, params) if sample_N else None
        else:
            states = self.sample_states(sample_N) if sample_N else None
        overlaps, magics, gkps, q_vals = [], [], [], []
        if measure_expr and n < 12:
            F = self.fidelities_all_pairs(states)
            overlaps = F.cpu().tolist()
            expr = self.expr(F, 2 ** n) if n < 7 else -1
        else:
            expr = -1
        if measure_ent:
            q_vals = self.entanglement_batch(states).cpu().tolist()
            q, std = np.mean(q_vals), np.std(q_vals)
        else:
            q, std = -1, -1
        if measure_eom or measure_GKP:
            from .magic import renyi_entropy_batch
        if measure_eom:
            magics = list(renyi_entropy_batch(states, n, 2))
            magic_bar, magic_std = np.mean(magics), np.std(magics)
        else:
            magic_bar, magic_std = -1, -1
        if measure_GKP:
            gkps = list(renyi_entropy_batch(states, n, 0.5) / (2 * np.log(2)))
            gkp_bar, gkp_std = np.mean(gkps), np.std(gkps)
        else:
            gkp_bar, gkp_std = -1, -1
        if full_data is True:
            return {"Expr": overlaps, "Ent": q_vals, "Magic": magics, "GKP": gkps}
        return {"Expr": expr, "Ent": [q, std], "Magic": [magic_bar, magic_std], "GKP": [gkp_bar, gkp_std]}

    # ------------------------------------------------------------- training
    def get_gradient_vector(self, theta):
        """measure.py:461-471: dE/dtheta_i = 2 Re <psi|H|d_i psi>, by the
        adjoint pass (two states, one backward sweep of small plans) where
        the circuit allows it, else from the forward Jacobian.  The
        derivative states (``gradient_list``, used by QNG) are built on first
        access."""
        self.QC.state = self.QC.run(angles=theta)
        self._grad_states = None
        try:
            from . import adjoint
            _, g = adjoint.energy_and_gradient(self.QC, self.QC._params_vec())
            return list(g)
        except NotImplementedError:
            pass
        psi = self.QC.state
        J = self._jacobian_states()
        H = self.QC.hamiltonian() if hasattr(self.QC, "hamiltonian") else self.QC.H
        Hpsi = H * psi
        g = E.gram(Hpsi.t.reshape(1, -1), J)[0]
        return list((2 * g.real).cpu().numpy())

    def _jacobian_states(self):
        _, J = self.QC.jacobian()
        self._grad_states = [Ket(J[i], self.QC.n_qubits) for i in range(J.shape[0])]
        return J

    @property
    def gradient_list(self):
        """The derivative states at the circuit's current parameters
        (measure.py:468, where get_gradient_vector stores them)."""
        if getattr(self, "_grad_states", None) is None:
            self._jacobian_states()
        return self._grad_states

    @gradient_list.setter
    def gradient_list(self, states):
        self._grad_states = states

    def train(self, epsilon=1e-6, rate=0.001, method="gradient", angles="random", verbose=False):
        """measure.py:473-553."""
        quit_iterations = 100000
        count, diff = 0, 1
        traj, magics, gkps, ents = [], [], [], []

        def trajmaj(Xi):
            magics.append(self.renyi_entropy_fast(self.QC.state))
            traj.append(self.minimize_function(Xi))
            ents.append(self.single_Q(self.QC.state, self.QC.n_qubits))
            gkps.append(self.gkp_fast(self.QC.state))

        self.QC.state = self.QC.run(angles=angles)
        trajmaj(angles)
        if method.lower() in ["gradient", "qng"]:
            prev_energy = self.minimize_function(angles)
            while diff > epsilon and count < quit_iterations:
                theta = self.QC.get_params()
                gradients = self.get_gradient_vector(theta)
                if method == "gradient":
                    theta_update = list(np.array(theta) - rate * np.array(gradients))
                elif method == "QNG":
                    QFI = self.get_QFI(grad_list=self.gradient_list)
                    inverse = np.linalg.pinv(QFI)
                    theta_update = list(np.array(theta) - rate * inverse.dot(np.array(gradients)))
                if count % 100 == 0 and verbose is True:
                    print(f"On iteration {count}, energy = {prev_energy}, diff is {diff}")
                energy = self.minimize_function(theta_update)
                diff = np.abs(energy - prev_energy)
                trajmaj(theta_update)
                count += 1
                prev_energy = energy
        else:
            op_out = scipy.optimize.minimize(self.minimize_function, x0=angles, method=method, callback=trajmaj,
                                             tol=epsilon)
            energy = op_out.fun
        return (energy, traj, magics, ents, gkps)
