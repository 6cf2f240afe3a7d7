// Host-side planner and run-time kernel emitter under AddressSanitizer and
// UndefinedBehaviorSanitizer (no GPU): reads op lists from a file, plans each
// one, emits the specialised source of its sweeps and queries the plan.
//
// File: records of int32 n_qubits, int32 n_ops, pqs_plan_opts, n_ops pqs_op.
// Prints "cases C planned P refused R sources S" on success.
#include <cstdio>
#include <string>
#include <vector>

#include "pqs.h"
#include "plan.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int cases = 0, planned = 0, refused = 0;
  size_t source_bytes = 0;
  for (;;) {
    int32_t hdr[2];
    if (std::fread(hdr, sizeof(int32_t), 2, f) != 2) break;
    pqs_plan_opts o;
    if (std::fread(&o, sizeof o, 1, f) != 1) return 3;
    std::vector<pqs_op> ops(static_cast<size_t>(hdr[1]));
    if (hdr[1] > 0 && std::fread(ops.data(), sizeof(pqs_op), ops.size(), f) != ops.size()) return 3;
    ++cases;
    pqs_plan *p = nullptr;
    if (pqs_plan_create(hdr[0], ops.data(), hdr[1], &o, &p) != PQS_OK) {
      ++refused;
      continue;
    }
    ++planned;
    int32_t n_sweeps = 0, tile_bits = 0, n_exec = 0, n_tan = 0;
    if (pqs_plan_info(p, &n_sweeps, &tile_bits, &n_exec, &n_tan) != PQS_OK) return 4;
    for (int32_t s = 0; s < n_sweeps; ++s) {
      std::vector<int32_t> rd(static_cast<size_t>(hdr[0])), wr(static_cast<size_t>(hdr[0]));
      if (pqs_plan_sweep_layout(p, s, rd.data(), wr.data()) != PQS_OK) return 5;
    }
    for (int32_t t = 0; t < n_tan; ++t) {
      int32_t parts = 0;
      if (pqs_plan_tangent_parts(p, t, &parts) != PQS_OK || parts < 1) return 6;
    }
    std::vector<int> sw;
    source_bytes += pqs::jit_source(*p, sw).size();
    for (int s : sw) source_bytes += pqs::jit_source_sweep(*p, s).size();
    delete p;  // host-only: the plan never reached a device
  }
  std::fclose(f);
  std::printf("cases %d planned %d refused %d sources %zu\n", cases, planned, refused, source_bytes);
  return 0;
}
