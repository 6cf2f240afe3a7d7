/*
 * pqs.h -- C ABI of the MI355X (gfx950) statevector engine behind the
 * pyramaterised-compatible Python API (package ``pyramaterised_amd``).
 *
 * The reference library (pyramaterised 1.0.2) has no native code and no FFI:
 * its hot path is a Python loop of QuTiP sparse operator x ket products.  Each
 * entry point below replaces one piece of that path and names the reference
 * code it stands in for.  A maintainer binds them with ctypes (the binding the
 * Python package itself uses is pyramaterised_amd/_lib.py).
 *
 * Conventions
 *   - State: 2^N complex amplitudes, interleaved (re, im); complex128 (16 B)
 *     or complex64 (8 B).  Amplitude index i = sum_k b_k 2^(N-1-k): qubit 0 is
 *     the most significant bit (QuTiP tensor order).
 *   - All pointers named *_dev are device pointers (e.g. torch tensors'
 *     data_ptr()); they are owned by the caller.  The library owns only its
 *     context workspace.
 *   - ``stream`` is a hipStream_t passed as void* (NULL = default stream).
 *     Every call is asynchronous on that stream unless stated otherwise.
 *   - Every function returns PQS_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is pqs_last_error().  No C++
 *     exception crosses the ABI.
 *   - One context per device per host thread; calls that share a context or
 *     a program must not run concurrently from several host threads.
 */
#ifndef PQS_H
#define PQS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQS_ABI_VERSION 1

enum pqs_status {
  PQS_OK = 0,
  PQS_ERR_ARG = -1,         /* invalid argument (bad qubit, size, pointer) */
  PQS_ERR_HIP = -2,         /* a HIP runtime call failed */
  PQS_ERR_UNSUPPORTED = -3, /* op / size the engine does not support */
  PQS_ERR_NOMEM = -4,       /* device allocation failed */
  PQS_ERR_NODEVICE = -5     /* no gfx950 device visible */
};

enum pqs_dtype { PQS_C128 = 0, PQS_C64 = 1 };

/* Primitive op kinds.  Angles are  scale * params[slot] + offset. */
enum pqs_op_kind {
  PQS_OP_MAT1 = 0,   /* fixed 2x2 on q0 (m: 4 complex, row-major)            */
  PQS_OP_RX = 1,     /* rx(angle) on q0                                      */
  PQS_OP_RY = 2,     /* ry(angle) on q0                                      */
  PQS_OP_RZ = 3,     /* rz(angle) on q0                                      */
  PQS_OP_MAT2 = 4,   /* fixed 4x4 on (q0, q1), q0 = more significant factor  */
  PQS_OP_CX = 5,     /* CNOT, control q0, target q1                          */
  PQS_OP_CZ = 6,     /* CZ on (q0, q1)                                       */
  PQS_OP_RXX = 7,    /* cos(a/2) I - i sin(a/2) X(x)X                        */
  PQS_OP_RYY = 8,    /* ... Y(x)Y                                            */
  PQS_OP_RZZ = 9,    /* ... Z(x)Z                                            */
  PQS_OP_FSIM = 10,  /* fsim(theta = angle(slot), phi = params[slot2]; 0 if
                        slot2 = -1, i.e. fixed_fSim)                          */
  /* Derivative ops: applied only by the instance whose tangent index equals
   * ``tangent`` (a no-op for every other instance).  Consecutive PQS_OP_GEN
   * terms with the same tangent form one generator D = sum_t coef_t P_t,
   * P_t = pauli(q0) (x) pauli(q1) (q1 = -1: one qubit; pauli 0=I 1=X 2=Y 3=Z),
   * coef_t = (m[0], m[1]).  A tangent whose terms come in several runs
   * (separated by other ops: one parameter on several gates) is
   * differentiated by the product rule: its state is the sum, over the runs,
   * of the evolution with that run's generator inserted. */
  PQS_OP_GEN = 11,
  /* fSim-family derivative matrices as the reference writes them
   * (gates.py:609-648, 719-737): pauli selects 1 = d_theta, 2 = d_phi,
   * 3 = fixed_fSim d_theta; angles from slot / slot2 as for PQS_OP_FSIM. */
  PQS_OP_GEN_FSIM = 12
};

typedef struct pqs_op {
  int32_t kind;
  int32_t q0, q1;       /* qubits (q1 = -1 for one-qubit ops)                 */
  int32_t slot, slot2;  /* parameter slots (-1: none)                         */
  int32_t tangent;      /* derivative ops: tangent index; else -1             */
  int32_t pauli;        /* PQS_OP_GEN / PQS_OP_GEN_FSIM selector              */
  int32_t reserved;
  double scale, offset; /* angle = scale * params[slot] + offset              */
  double m[32];         /* fixed matrix / coefficient, (re, im) interleaved  */
} pqs_op;

/* Planner options (zeros select defaults). */
typedef struct pqs_plan_opts {
  int32_t tile_bits;    /* qubits resident per tile (<= 13 for c128)          */
  int32_t low_bits;     /* lowest bit positions forced into every tile        */
  int32_t block;        /* threads per workgroup (256, 512 or 1024)           */
  int32_t dtype;        /* pqs_dtype                                          */
  int32_t flags;        /* PQS_PLAN_* tuning switches (0: defaults)           */
} pqs_plan_opts;

/* Planner tuning switches (pqs_plan_opts.flags). */
enum {
  PQS_PLAN_LANES_ASCENDING = 1, /* thread bits in ascending order, not bank-aware */
  PQS_PLAN_NO_SOFT_REGS = 2,    /* do not pull CNOT / CZ qubits into registers   */
  PQS_PLAN_JIT_PERSISTENT = 4,  /* specialised kernels: persistent blocks that
                                   prefetch their next tile into registers (pays
                                   off only at one resident block per CU)         */
  PQS_PLAN_NO_WAVE_LOCAL = 16,  /* specialised kernels: end every register phase
                                   with a workgroup barrier (default: phases that
                                   keep the same wave bits exchange through LDS
                                   within each wave, without one) */
  PQS_PLAN_NO_DIRECT = 32,      /* specialised kernels: stage every tile through
                                   LDS at both ends of a sweep (default: the first
                                   / last register phase loads / stores its
                                   registers directly when the tile bits at
                                   positions 0 and 1 are lane bits there) */
  PQS_PLAN_ROLLING = 8,         /* rolling layouts: each sweep writes the state
                                   with the qubit -> bit-position map under which
                                   the next sweep's tile is one contiguous block;
                                   low_bits is then the least overlap kept between
                                   consecutive tiles (contiguous write runs), not
                                   a fixed set of qubits.  Runs go out of place
                                   (ping-pong workspace of one state per
                                   instance).  Ignored for tiles below 10
                                   qubits. */
  PQS_PLAN_GREEDY_SWEEPS = 64,  /* rolling plans: take every sweep's tile from
                                   the greedy pass over the op order (default:
                                   also try tiles seeded with a window of
                                   consecutive qubits, scored by the greedy
                                   pass 1, 2 or 3 sweeps ahead, and keep the
                                   partition with the fewest sweeps) */
  PQS_PLAN_GREEDY_PHASES = 128, /* take every register phase's qubits from the
                                   greedy pass (default: also try register sets
                                   seeded with 4 consecutive tile qubits and
                                   keep the one after which the greedy pass
                                   needs the fewest phases) */
  PQS_PLAN_NO_EXCHANGE = 256    /* specialised kernels: every register-phase
                                   transition goes through LDS (default: a
                                   phase whose register bits differ from the
                                   previous phase's by one or two tile bits
                                   held on lane bits 5 / 4 there trades them
                                   with cross-lane swaps, no LDS round trip) */
};

/* One instance of a batched run. */
typedef struct pqs_instance {
  void *state_dev;         /* output state (2^N amplitudes)                   */
  const void *src_dev;     /* input state; NULL: start from the init table    */
  const double *params_dev;/* parameter vector of this instance               */
  int32_t tangent;         /* -1: plain evolution; t: insert generator t      */
  int32_t init;            /* 0: read src (or state if src NULL); 1: |0..0>;
                              2: product state from the init table            */
} pqs_instance;

typedef struct pqs_ctx pqs_ctx;
typedef struct pqs_plan pqs_plan;

int pqs_abi_version(void);
const char *pqs_last_error(void);

/* Named ranges for profilers (rocprofv3 --marker-trace).  With the
 * environment variable PQS_ROCTX=1 the library loads the rocprofiler-sdk
 * roctx library (or libroctx64) and marks
 * its compute entry points and the caller's ranges; otherwise these calls do
 * nothing.  Ranges nest per host thread.  No reference counterpart: the
 * serving loop marks its batches with them. */
int pqs_range_push(const char *name);
int pqs_range_pop(void);

/* Context on one device; owns a workspace that grows on demand. */
int pqs_ctx_create(int device, pqs_ctx **out);
int pqs_ctx_destroy(pqs_ctx *ctx);
int pqs_device_info(pqs_ctx *ctx, int32_t *n_cu, int32_t *max_lds_bytes, char *arch, int32_t arch_len);

/* Cap the blocks (workgroups) of one sweep launch on this context (0, the
 * default: no cap).  A batched sweep above the cap is launched over
 * instance ranges of at most `max_blocks` blocks each, with the same
 * results.  The blocks of a running launch take every slot that frees up
 * before a kernel of another stream gets one, so a cap keeps a long batched
 * run (a Jacobian of hundreds of rows) from holding back short kernels of
 * other streams for its whole length.  No reference counterpart: the
 * serving loop sets it on the contexts of its QFIM streams. */
int pqs_ctx_set_max_blocks(pqs_ctx *ctx, int64_t max_blocks);

/* Copy `bytes` host bytes at `src` to device memory at `dst` in the order
 * of `stream`.  The bytes are staged in the context's pinned buffers before
 * the call returns (the caller may reuse `src` at once), and a kernel on
 * `stream` copies them: the host does not wait for the stream's earlier
 * work, and the copy does not queue behind other streams' DMA copies.  No
 * reference counterpart: the bindings upload parameters with it. */
int pqs_upload(pqs_ctx *ctx, void *dst, const void *src, int64_t bytes, void *stream);

/* Compile a flat op list into sweeps (host only; no device needed).
 * Replaces the per-gate operator rebuild and the Python loop
 * ``for g in self.gates: circuit_state = g * circuit_state``
 * (reference circuit.py:118-125, gates.py:122-131). */
int pqs_plan_create(int32_t n_qubits, const pqs_op *ops, int32_t n_ops,
                    const pqs_plan_opts *opts, pqs_plan **out);
/* Frees the plan.  A plan that has run waits for the whole device to be idle
 * first (work queued on any stream may still read its arrays), so a loop
 * with other streams busy should keep its plans rather than drop them.
 * A plan's arrays are copied to the device at its first launch, in that
 * launch's stream order; launches on other streams wait for that copy. */
int pqs_plan_destroy(pqs_plan *plan);
/* Plan shape: number of sweeps, tile bits, ops after merging, tangents. */
int pqs_plan_info(const pqs_plan *plan, int32_t *n_sweeps, int32_t *tile_bits,
                  int32_t *n_exec_ops, int32_t *n_tangents);
/* Sweep s: the tile's qubits (ascending bit position; *n_tile entries) and the
 * indices, into the caller's op array, of the ops it executes in order.
 * A merged generator reports the index of its first term. */
int pqs_plan_sweep(const pqs_plan *plan, int32_t s, int32_t *tile_qubits,
                   int32_t *n_tile, int32_t *op_index, int32_t *n_op);
/* Bit positions of every qubit (n entries each) in the layout sweep s reads
 * and in the layout it writes.  Position n-1-q is the standard (QuTiP) order;
 * both are standard for every sweep of a plan without PQS_PLAN_ROLLING, and
 * the first read and the last write are standard in every plan. */
int pqs_plan_sweep_layout(const pqs_plan *plan, int32_t s, int32_t *read_pos, int32_t *write_pos);
/* First sweep that executes the generator of tangent t (-1 if none). */
int pqs_plan_tangent_sweep(const pqs_plan *plan, int32_t t, int32_t *sweep);
/* Number of parts the generator of tangent t was split into.  A merged
 * generator (a run of PQS_OP_GEN terms of one tangent) that acts
 * non-diagonally on more qubits than the tile is split into parts of
 * consecutive terms that each fit a tile, and every further run of a
 * tangent's terms (see PQS_OP_GEN) is a part too; pqs_jacobian and tangent
 * runs of pqs_run / pqs_run_batch evolve every part as its own instance
 * (extra rows from the context workspace) and sum them into the tangent's
 * state, which is exact by linearity (reference
 * shared_parameter.derivative, gates.py:451-466, for generators such as
 * TFIM_layers' shared R_x over all N qubits, templates.py:194-204).  1 for
 * an unsplit generator in one run. */
int pqs_plan_tangent_parts(const pqs_plan *plan, int32_t t, int32_t *n_parts);

/* Batched evolution: every instance runs the whole plan on its own state.
 * Replaces PQC.run (circuit.py:118-125), and with tangent >= 0 the
 * per-parameter re-simulation of PQC.take_derivative (circuit.py:149-172).
 * init_table_dev: 2N complex (per qubit the single-qubit state (s0, s1)) for
 * init == 2, in the plan's dtype. */
int pqs_run(pqs_ctx *ctx, pqs_plan *plan, const pqs_instance *insts,
            int64_t n_inst, const void *init_table_dev, void *stream);

/* pqs_run for a strided batch: instance i evolves states_dev + i*state_stride
 * (amplitudes) with params_dev + i*param_stride, tangent tangents_host[i]
 * (NULL: all plain), starting from src_dev (shared by all instances, init 0),
 * |0..0> (init 1) or the product table (init 2).  One launch per sweep for
 * the whole batch.  Replaces ``[self.QC.run("random") for i in
 * range(sample_N)]`` (measure.py:132, 247, 356, 407). */
int pqs_run_batch(pqs_ctx *ctx, pqs_plan *plan, void *states_dev, int64_t state_stride,
                  int64_t batch, const double *params_dev, int64_t param_stride,
                  const int32_t *tangents_host, int32_t init, const void *src_dev,
                  const void *init_table_dev, void *stream);

/* pqs_run_batch (plain evolution, no tangents) of an even batch whose
 * instances pair up as (2i, 2i+1), with overlaps_dev[i] = <psi_2i|psi_2i+1>
 * (complex128: 2 doubles per pair, batch / 2 pairs) besides the states.
 * Replaces the fidelity of a pair of random states in
 * Measurements._gen_f_samples (measure.py:123-137, psi.overlap(phi) at :135)
 * when states are sampled in independent pairs.  A plan with run-time
 * specialised sweeps forms the overlaps inside its last sweep (each block
 * evolves the same tile of both instances of a pair and pairs their final
 * amplitudes in registers), so the states are not read back; otherwise they
 * are read back after the sweeps.  PQS_ERR_ARG for an odd batch.  Arguments
 * otherwise as pqs_run_batch. */
int pqs_run_batch_pairs(pqs_ctx *ctx, pqs_plan *plan, void *states_dev, int64_t state_stride, int64_t batch,
                        const double *params_dev, int64_t param_stride, int32_t init, const void *src_dev,
                        const void *init_table_dev, double *overlaps_dev, void *stream);

/* Instances of several plans in one run: instance i evolves with plan
 * plans[inst_plan[i]] (fields as in pqs_run; init 2 takes the per-instance
 * product table init_tabs[i], init_tabs may be NULL otherwise).  The plans
 * may differ in width and depth but share one dtype.  Sweep round r is one
 * launch per group of plans with the same tile size (register-phase plans;
 * all plans with tiles below 10 qubits form one group) over sweep r of every
 * instance whose plan has more than r sweeps: its blocks are the
 * instances' tiles laid end to end, each block finding its instance by a
 * binary search.  Kernels are the interpreted sweep executors (run-time
 * specialised kernels are per plan).  Tangent instances are accepted for
 * plans without split generators.  No reference counterpart: it serves
 * callers that batch circuits of uneven width and depth (a serving loop's
 * width classes, the reference's sampling at measure.py:132 over many
 * circuits). */
int pqs_run_multi(pqs_ctx *ctx, pqs_plan *const *plans, int32_t n_plans, const pqs_instance *insts,
                  const int32_t *inst_plan, const void *const *init_tabs, int64_t n_inst, void *stream);

/* Graph-capturable batched runs.  pqs_batch_table writes the instance table
 * of a strided batch (arguments as pqs_run_batch; init 0 needs src_dev) into
 * the caller's device buffer table_dev (batch * pqs_batch_table_bytes(1)
 * bytes), with alt_dev as the second buffer of every instance when the plan
 * uses rolling layouts (batch * state_stride amplitudes, else NULL).
 * pqs_run_table then launches every sweep for that table: no allocation, no
 * host-to-device copy, nothing but kernel launches on `stream`, so it can
 * be captured in a HIP graph and replayed; the table records buffer
 * addresses, so replays see the current contents of states, params and src.
 * The plan must have been run or tabled once on ctx's device before a
 * capture (its device image and kernel attributes are set up then).  No
 * reference counterpart: it serves callers that replay one batch shape many
 * times (sampling loops, serving). */
int64_t pqs_batch_table_bytes(int64_t batch);
int pqs_batch_table(pqs_ctx *ctx, pqs_plan *plan, void *states_dev, int64_t state_stride, int64_t batch,
                    const double *params_dev, int64_t param_stride, int32_t init, const void *src_dev,
                    void *alt_dev, void *table_dev, void *stream);
int pqs_run_table(pqs_ctx *ctx, pqs_plan *plan, const void *table_dev, int64_t batch,
                  const void *init_table_dev, void *stream);

/* pqs_run_batch (plain evolution, no tangents) without any allocation.  No
 * counterpart in the reference: it serves callers that cannot allocate
 * device memory -- code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc.  The instance table is written by a
 * kernel into a file-scope device array of the library (one per GPU, held
 * until the process ends), so the call allocates nothing, copies nothing
 * from the host and only launches kernels on `stream`.  Limits: batch <=
 * 16384 instances (PQS_ERR_ARG beyond); the plan must have run once on
 * ctx's device (pqs_run_batch / pqs_run / pqs_batch_table upload its device
 * image; PQS_ERR_ARG otherwise) and must carry no tangents
 * (PQS_ERR_UNSUPPORTED); a rolling-layout plan needs alt_dev (batch *
 * state_stride amplitudes, caller-owned) and, with init 0, a src_dev that is
 * not the output.  Every call on one GPU shares the one array: calls that
 * overlap on one GPU (several streams or host threads) race and must be
 * serialised by the caller.  Arguments otherwise as pqs_run_batch. */
int pqs_run_batch_noalloc(pqs_ctx *ctx, pqs_plan *plan, void *states_dev, int64_t state_stride, int64_t batch,
                          const double *params_dev, int64_t param_stride, int32_t init, const void *src_dev,
                          void *alt_dev, const void *init_table_dev, void *stream);

/* All derivative states in one pass: psi_dev gets U|psi0>, jac_dev (row t =
 * tangent t, n_tangents rows of 2^N) gets every d psi / d theta_t.  Tangent t
 * is spawned from the plain state at the sweep that holds its generator and
 * carried to the end with the remaining sweeps (forward mode).
 * Replaces PQC.get_gradients (circuit.py:174-192). */
int pqs_jacobian(pqs_ctx *ctx, pqs_plan *plan, const double *params_dev,
                 int32_t init, const void *src_dev, const void *init_table_dev,
                 void *psi_dev, void *jac_dev, void *stream);

/* out[p] = <a_p|b_p> (complex, dtype of the states) for n_pairs pairs, each
 * pair given by element offsets (in amplitudes) from base_dev.
 * Replaces Qobj.overlap (circuit.py:141, measure.py:52,58,135,418). */
int pqs_inner(pqs_ctx *ctx, int32_t dtype, const void *base_dev, int64_t dim,
              const int64_t *a_off_dev, const int64_t *b_off_dev, int64_t n_pairs,
              double *out_dev, void *stream);

/* Per instance and qubit the 2x2 reduced density matrix (rho00, rho11,
 * Re rho01, Im rho01) into rho_dev[b][k][4] and the Meyer-Wallach
 * Q = 2(1 - mean_k Tr rho_k^2) into q_dev[b].  States are contiguous with
 * stride 2^N.  Replaces Measurements.single_Q / Qobj.ptrace
 * (measure.py:226-249). */
int pqs_mw(pqs_ctx *ctx, int32_t dtype, const void *states_dev, int32_t n_qubits,
           int64_t batch, double *rho_dev, double *q_dev, void *stream);

/* Complex Gram: out[i][j] = <a_i|b_j> for rows a_i (m rows) and b_j (n rows)
 * of length dim, row strides lda / ldb (amplitudes); out is m x n complex128
 * (re, im).  Used for the QFIM (measure.py:55-63) and all-pairs fidelities
 * (measure.py:133-136). */
int pqs_gram(pqs_ctx *ctx, int32_t dtype, const void *a_dev, int64_t lda, int64_t m,
             const void *b_dev, int64_t ldb, int64_t n, int64_t dim,
             double *out_dev, void *stream);

/* QFIM from a Jacobian and the state: F_pq = 4 Re(<J_p|J_q> - <J_p|psi><psi|J_q>)
 * into f_dev (P x P float64).  Replaces Measurements.get_QFI
 * (measure.py:33-71).  complex128 forms Re(J J^dag) = R R^T, R = J viewed as a
 * P x 2^(N+1) real matrix, one 64 x 64 upper tile per wave on the FP64 matrix
 * cores, split over K (each K chunk's tiles on one XCD, sharing its L2), the
 * diagonal tiles also forming <J_p|psi>; a Jacobian above 1 GiB goes through
 * in column blocks copied to a padded-row buffer; complex64 Jacobians take
 * the same path with their values widened to double in registers. */
int pqs_qfim(pqs_ctx *ctx, int32_t dtype, const void *jac_dev, int64_t P,
             const void *psi_dev, int64_t dim, double *f_dev, void *stream);

/* sum_t coef_t <psi|P_t|psi> for Pauli strings given as qubit masks
 * (bit q of xmask/zmask = X/Z component on qubit q; Y = both), coef complex
 * (re, im); one value per instance (contiguous states).  Replaces
 * qt.expect(H, psi) for Pauli-sum H (circuit.py:132-137). */
int pqs_expect_pauli(pqs_ctx *ctx, int32_t dtype, const void *states_dev, int32_t n_qubits,
                     int64_t batch, const uint64_t *xmask_host, const uint64_t *zmask_host,
                     const double *coef_host, int32_t n_terms, double *out_dev, void *stream);

/* out = H x for H = sum_t coef_t P_t (Pauli strings as in pqs_expect_pauli)
 * on one state (out must not alias x).  The lambda = H psi that starts the
 * adjoint gradient's backward pass (reference measure.py:461-471, which
 * forms H |d psi> per parameter). */
int pqs_apply_pauli(pqs_ctx *ctx, int32_t dtype, const void *x_dev, int32_t n_qubits, const uint64_t *xmask_host,
                    const uint64_t *zmask_host, const double *coef_host, int32_t n_terms, void *out_dev,
                    void *stream);

/* out[t] = <bra| coef_t P_t |ket> (complex128, (re, im)) for every term t
 * (Pauli strings as in pqs_expect_pauli), two-stage sums in a fixed order.
 * With bra = lambda and ket = psi at a point of the adjoint backward pass,
 * 2 Re of a parameter's terms is its energy gradient 2 Re <psi|H|d psi>
 * (reference Measurements.get_gradient_vector, measure.py:461-471). */
int pqs_pauli_transition(pqs_ctx *ctx, int32_t dtype, const void *bra_dev, const void *ket_dev, int32_t n_qubits,
                         const uint64_t *xmask_host, const uint64_t *zmask_host, const double *coef_host,
                         int32_t n_terms, double *out_dev, void *stream);

/* Histogram of values into n_bins int64 counts over the n_bins + 1 ascending
 * edges (device array, e.g. numpy.linspace(0, 1, n_bins + 1)) with
 * numpy.histogram's rules: values outside [edges[0], edges[n_bins]] are
 * dropped, the last bin is closed.  Replaces np.histogram in
 * Measurements._gen_histo (measure.py:153-155). */
int pqs_histogram(pqs_ctx *ctx, const double *values_dev, int64_t n, const double *edges_dev,
                  int32_t n_bins, int64_t *counts_dev, void *stream);

/* out_b = U x_b for a dense operator U (dim x dim complex128, row-major, device)
 * on a batch of states (row stride `stride` amplitudes, dtype of the
 * states; out must not alias them), dim <= 2^14.  Applies gates that have no
 * local form: ARBGATE's U = expm(-i theta H) and its derivative matrix
 * (reference gates.py:407-435), computed on the host. */
int pqs_apply_dense(pqs_ctx *ctx, int32_t dtype, const double *u_dev, int64_t dim, const void *states_dev,
                    int64_t stride, int64_t batch, void *out_dev, void *stream);

/* |z|^2 of complex128 values: out[i] = re^2 + im^2 (fidelity from overlap). */
int pqs_abs2(pqs_ctx *ctx, const double *z_dev, int64_t n, double *out_dev, void *stream);

/* Stabiliser Renyi sum per instance: out[b] = sum over Pauli strings P of
 * (<psi|P|psi>^2 / 2^n)^alpha, from which the entropy of magic is
 * log(out)/(1-alpha) - n log 2 (n <= 13).  Replaces the O(4^n)-memory index /
 * sign matrices and np.dot of Measurements.renyi_entropy_fast
 * (measure.py:268-349). */
int pqs_magic(pqs_ctx *ctx, int32_t dtype, const void *states_dev, int32_t n_qubits,
              int64_t batch, double alpha, double *out_dev, void *stream);

/* Run-time specialisation (hiprtc) of the plan's sweeps: every sweep made of
 * one-qubit rotations, fixed one-qubit matrices, CNOT, CZ and R_zz gets its
 * own straight-line kernel with register and bit indices as constants; other
 * sweeps stay interpreted.  load = 0 compiles only (no device needed, used
 * to check generated code); load = 1 also loads the module on ctx's device,
 * after which pqs_run / pqs_run_batch / pqs_jacobian launch the specialised
 * kernels.  No reference counterpart (the reference has no kernels). */
int pqs_plan_jit(pqs_ctx *ctx, pqs_plan *plan, int32_t load);
/* The generated source (for inspection); *len gets its length. */
int pqs_plan_jit_source(const pqs_plan *plan, char *buf, int64_t cap, int64_t *len,
                        int32_t *n_jit_sweeps);

/* Source of sweep s's specialised kernel (*len = 0 if it is interpreted). */
int pqs_plan_jit_sweep_source(const pqs_plan *plan, int32_t s, char *buf, int64_t cap, int64_t *len);
/* Compile one kernel source into the code-object cache (no device needed;
 * lets several processes share the compile work of pqs_plan_jit). */
int pqs_jit_prepare(const char *src, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* PQS_H */
