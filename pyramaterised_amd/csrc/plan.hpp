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
};
constexpr int kPauliSum = 100;

struct Sweep {
  std::vector<int> tile_qubits; // ascending bit position (pos = N-1-q)
  std::vector<int> ops;         // ExecOp indices in execution order
};

// ---- device image -------------------------------------------------------
// Op record as the kernels read it (uniform loads through the scalar cache).
struct alignas(16) DevOp {
  int32_t kind, tangent, slot, slot2;
  int32_t t0, t1;     // tile-local bit index of q0 / q1, -1 if outside the tile
  int32_t g0, g1;     // global bit position of q0 / q1, -1 if none
  int32_t term0, nterms;
  int32_t pauli, pad;
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

struct DevSweep {
  int32_t op_begin, op_end;
  int32_t k;          // tile bits
  int32_t nruns;      // runs of tile positions
  int32_t run_pos[kMaxTileBits], run_len[kMaxTileBits], run_shift[kMaxTileBits];
  int32_t cnruns;     // runs of the complement (tile id) positions
  int32_t crun_pos[kMaxQubits], crun_len[kMaxQubits], crun_shift[kMaxQubits];
  int64_t n_tiles;
};

std::string plan_error(); // last planner error (thread local)

}  // namespace pqs

struct pqs_plan {
  int n = 0;
  int tile_bits = 0;
  int low_bits = 0;
  int block = 0;
  int dtype = 0;
  int n_tangents = 0;
  std::vector<pqs_op> src_ops;
  std::vector<pqs::ExecOp> exec;
  std::vector<pqs::Sweep> sweeps;
  std::vector<int> tangent_sweep;  // first sweep holding tangent t's generator

  // host copy of the device image
  std::vector<pqs::DevOp> dev_ops;
  std::vector<pqs::DevTerm> dev_terms;
  std::vector<pqs::DevSweep> dev_sweeps;

  // device copy (per plan, on the device of the first context that ran it)
  int dev_device = -1;
  void *d_ops = nullptr;
  void *d_terms = nullptr;
};

namespace pqs {
// Build sweeps + the device image; returns PQS_OK or an error status.
int build_plan(pqs_plan &p, int n, const pqs_op *ops, int n_ops, const pqs_plan_opts *opts);
}
