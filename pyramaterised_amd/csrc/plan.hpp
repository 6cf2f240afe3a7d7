// Internal structures shared by the planner (plan.cpp) and the HIP engine
// (engine.hip).  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pqs.h"

namespace pqs {

constexpr int kMaxQubits = 34;   // 2^34 complex128 = 256 GiB: one MI355X
constexpr int kMaxTileBits = 14; // 2^14 complex64 = 128 KiB of LDS

// A 64-bit qubit mask: bit q set <=> qubit q.
using qmask = uint64_t;

// One executable op after generator-term merging.
struct ExecOp {
  int kind = 0;          // pqs_op_kind, or kPauliSum for a merged PQS_OP_GEN
  int first = 0;         // index of the first source op
  int n_src = 1;         // number of source ops merged (generator terms)
  qmask active = 0;      // qubits the op acts on non-diagonally
  qmask diag = 0;        // qubits the op acts on diagonally (block-diagonal)
  int tangent = -1;      // derivative ops: tangent (or sub-tangent) index
};
constexpr int kPauliSum = 100;
// A one-term, one-qubit generator (coef * Pauli on q0, applied only by the
// instance whose tangent matches) placed in a register phase: device kind.
constexpr int kGen1 = 101;

struct Sweep {
  std::vector<int> tile_qubits; // ascending bit position in the read layout
  std::vector<int> ops;         // ExecOp indices in execution order
  std::vector<int> rpos, wpos;  // per qubit: bit position read / written
};

// ---- device image -------------------------------------------------------
// Op record as the kernels read it (uniform loads through the scalar cache).
struct alignas(16) DevOp {
  int32_t kind, tangent, slot, slot2;
  int32_t t0, t1;     // tile-local bit index of q0 / q1, -1 if outside the tile
  int32_t g0, g1;     // global bit position of q0 / q1, -1 if none
  int32_t term0, nterms;
  int32_t pauli, coef;  // coef: index of (cos, sin) in the sweep's table, -1 none
  int32_t r0, r1;       // register index of q0 / q1 in the op's phase, -1 if not
  double scale, offset;
  double m[32];
};

// One Pauli-string term of a merged generator, resolved against a tile:
// (P psi)[j] = coef * (-1)^{popc((j^xm) & zm_tile) + popc(base & zm_glob)} psi[j^xm]
struct alignas(16) DevTerm {
  uint32_t xm, zm_tile;
  uint64_t zm_glob;
  double cre, cim;  // coefficient including i^{#Y}
};

// Register-phase executor: each thread holds 2^kRegBits amplitudes of the
// tile, over kRegBits tile bits chosen per phase.
constexpr int kRegBits = 4;
constexpr int kMaxSweepOps = 896;  // op records staged in LDS per sweep (28 KiB)

// One phase of a sweep: either register resident (ops on rb[] bits run in
// VGPRs) or an LDS phase (one op that needs the whole tile).
struct DevPhase {
  int32_t op_begin, op_end;
  int32_t lds;               // 1: LDS phase
  int32_t rb[kRegBits];      // tile-local bit of register bit i
  int32_t ntb;               // number of thread bits
  int32_t tb[kMaxTileBits];  // tile-local bit of thread bit i
  // Exchange with the previous register phase: this phase's layout is the
  // previous one's with register bit xr[j] and thread (lane) bit xb[j]
  // (4 or 5) traded, j < xchg, and nothing else changed, so a specialised
  // kernel moves the amplitudes with cross-lane swaps instead of an LDS
  // round trip.  0: an ordinary transition.
  int32_t xchg;
  int32_t xr[2], xb[2];
};

struct DevSweep {
  int32_t op_begin, op_end;
  int32_t k;          // tile bits
  int32_t nruns;      // runs of tile positions
  int32_t run_pos[kMaxTileBits], run_len[kMaxTileBits], run_shift[kMaxTileBits];
  int32_t cnruns;     // runs of the complement (tile id) positions
  int32_t crun_pos[kMaxQubits], crun_len[kMaxQubits], crun_shift[kMaxQubits];
  int64_t n_tiles;
  int32_t phase_begin, phase_end;  // register phases (reg executor only)
  int32_t n_coef;                  // 1: the sweep has LDS phases (dense 2q ops, generators)
  // Write layout (rolling plans): the tile is stored through a permutation
  // of its bits, store index bit i <-> tile bit sperm[i], ordered by
  // ascending write position wpos[sperm[i]]; complement (tile id) bit j is
  // written at cwpos[j].  rolling = 0: write layout = read layout.
  int32_t rolling;
  int32_t n;  // qubits of the state
  int32_t sperm[kMaxTileBits], wpos[kMaxTileBits];
  int32_t cwpos[kMaxQubits];
  // The tile bits read (written) at positions 0 and 1 are lane bits of the
  // first (last) register phase, so the specialised kernel loads (stores)
  // that phase's registers directly, without staging the tile in LDS.
  int32_t direct_load, direct_store;
};

std::string plan_error(); // last planner error (thread local)

}  // namespace pqs

struct pqs_plan {
  int n = 0;
  int tile_bits = 0;
  int low_bits = 0;
  int block = 0;
  int dtype = 0;
  int flags = 0;
  bool rolling = false;  // rolling layouts in effect (PQS_PLAN_ROLLING honoured)
  int n_tangents = 0;     // tangents of the caller's op list
  // A generator wider than the tile is split into parts; every part after
  // the first runs as a sub-tangent n_tangents.. whose state is summed into
  // its parent's row at the end of a run.  n_sub counts all of them.
  int n_sub = 0;
  std::vector<int> tangent_parent;  // per tangent / sub-tangent: caller tangent
  std::vector<pqs_op> src_ops;
  std::vector<pqs::ExecOp> exec;
  std::vector<pqs::Sweep> sweeps;
  std::vector<int> tangent_sweep;  // first sweep holding (sub-)tangent t's generator

  // host copy of the device image
  std::vector<pqs::DevOp> dev_ops;
  std::vector<pqs::DevTerm> dev_terms;
  std::vector<pqs::DevSweep> dev_sweeps;
  std::vector<pqs::DevPhase> dev_phases;
  bool reg_exec = false;  // sweeps carry register phases

  // device copy (per plan, on the device of the first context that ran it)
  int dev_device = -1;
  void *d_ops = nullptr;
  void *d_terms = nullptr;
  void *d_phases = nullptr;
  void *d_sweeps = nullptr;  // DevSweep records (launches that mix plans)
  // the arrays are copied in stream order on the stream of the plan's first
  // launch; launches on other streams wait for this event until it has run
  void *up_event = nullptr;   // hipEvent_t
  void *up_stream = nullptr;  // hipStream_t
  bool up_pending = false;

  // run-time specialised sweep kernels (hiprtc), per plan on one device
  void *jit_module = nullptr;        // non-null once loaded
  std::vector<void *> jit_modules;   // one module per specialised sweep
  std::vector<void *> jit_fn;        // per sweep; null: interpreted
  void *jit_pair_fn = nullptr;       // pair kernel of the last sweep (pqs_run_batch_pairs)
};

namespace pqs {
// Build sweeps + the device image; returns PQS_OK or an error status.
int build_plan(pqs_plan &p, int n, const pqs_op *ops, int n_ops, const pqs_plan_opts *opts);
// Specialised kernel source for the sweeps that qualify (listed in sweeps).
std::string jit_source(const pqs_plan &p, std::vector<int> &sweeps);
std::string jit_source_sweep(const pqs_plan &p, int s);
bool jit_sweep_ok(const pqs_plan &p, int s);
int jit_compile(const std::string &src, std::vector<char> &code, std::string &log);
std::string jit_signature();  // compile options and hiprtc version
}
