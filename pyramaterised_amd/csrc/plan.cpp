// Planner: turns a flat op list (the reference's ``PQC.gates`` order, one
// primitive per local gate; reference circuit.py:53-61,118-125) into sweeps.
//
// A sweep keeps 2^k amplitudes of one tile in LDS: k "tile qubits" T, the
// rest fixed per tile.  An op can run in a sweep when every qubit it acts on
// non-diagonally is in T; qubits it touches only diagonally (a CNOT control,
// both CZ / R_zz qubits, R_z) may lie outside T because their bit is constant
// across the tile.  Ops keep the relative order of every pair that shares a
// qubit on which at least one of them is non-diagonal; all other pairs
// commute (both are block-diagonal in that qubit's basis), so reordering them
// does not change the linear map.  The greedy pass below grows T from the
// earliest dependency-free ops, so consecutive layers of 1- and 2-qubit gates
// fuse into as few full-state sweeps as the tile size allows.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "plan.hpp"

namespace pqs {

static thread_local std::string g_err;

void set_error(const std::string &s) { g_err = s; }
std::string plan_error() { return g_err; }

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

static inline int popc(qmask m) { return __builtin_popcountll(m); }
static inline qmask bitq(int q) { return qmask(1) << q; }

static bool is_two_qubit(int kind) {
  switch (kind) {
    case PQS_OP_MAT2: case PQS_OP_CX: case PQS_OP_CZ: case PQS_OP_RXX:
    case PQS_OP_RYY: case PQS_OP_RZZ: case PQS_OP_FSIM: case PQS_OP_GEN_FSIM:
      return true;
    default:
      return false;
  }
}

static bool has_angle(int kind) {
  switch (kind) {
    case PQS_OP_RX: case PQS_OP_RY: case PQS_OP_RZ: case PQS_OP_RXX:
    case PQS_OP_RYY: case PQS_OP_RZZ: case PQS_OP_FSIM: case PQS_OP_GEN_FSIM:
      return true;
    default:
      return false;
  }
}

static int validate(int n, const pqs_op &o, int idx) {
  char buf[256];
  auto bad = [&](const char *what) {
    std::snprintf(buf, sizeof buf, "op %d (kind %d): %s", idx, o.kind, what);
    return fail(PQS_ERR_ARG, buf);
  };
  if (o.kind < PQS_OP_MAT1 || o.kind > PQS_OP_GEN_FSIM) return bad("unknown kind");
  if (o.q0 < 0 || o.q0 >= n) return bad("q0 out of range");
  bool two = is_two_qubit(o.kind) || (o.kind == PQS_OP_GEN && o.q1 >= 0);
  if (two) {
    if (o.q1 < 0 || o.q1 >= n) return bad("q1 out of range");
    if (o.q1 == o.q0) return bad("q0 == q1");
  }
  if (has_angle(o.kind) && o.slot < 0 && o.scale != 0.0) return bad("angle op without a slot");
  if (o.kind == PQS_OP_GEN || o.kind == PQS_OP_GEN_FSIM) {
    if (o.tangent < 0) return bad("derivative op without a tangent index");
    if (o.kind == PQS_OP_GEN && (o.pauli < 0 || o.pauli > 3)) return bad("pauli must be 0..3");
    if (o.kind == PQS_OP_GEN_FSIM && (o.pauli < 1 || o.pauli > 3)) return bad("fSim derivative selector must be 1..3");
    if (o.kind == PQS_OP_GEN_FSIM && o.pauli != 3 && o.slot2 < 0) return bad("fSim derivative without a phi slot");
  } else if (o.tangent >= 0) {
    return bad("only derivative ops carry a tangent index");
  }
  return PQS_OK;
}

static void masks_of(int n, const pqs_op &o, qmask &active, qmask &diag) {
  (void)n;
  active = diag = 0;
  switch (o.kind) {
    case PQS_OP_MAT1: {
      bool d = o.m[2] == 0 && o.m[3] == 0 && o.m[4] == 0 && o.m[5] == 0;
      (d ? diag : active) |= bitq(o.q0);
      break;
    }
    case PQS_OP_RX: case PQS_OP_RY: active |= bitq(o.q0); break;
    case PQS_OP_RZ: diag |= bitq(o.q0); break;
    case PQS_OP_MAT2: {
      bool d = true;
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
          if (r != c && (o.m[2 * (4 * r + c)] != 0 || o.m[2 * (4 * r + c) + 1] != 0)) d = false;
      (d ? diag : active) |= bitq(o.q0) | bitq(o.q1);
      break;
    }
    case PQS_OP_CX: diag |= bitq(o.q0); active |= bitq(o.q1); break;
    case PQS_OP_CZ: case PQS_OP_RZZ: diag |= bitq(o.q0) | bitq(o.q1); break;
    case PQS_OP_RXX: case PQS_OP_RYY: case PQS_OP_FSIM: case PQS_OP_GEN_FSIM:
      active |= bitq(o.q0) | bitq(o.q1);
      break;
    case PQS_OP_GEN: {
      qmask qs = bitq(o.q0) | (o.q1 >= 0 ? bitq(o.q1) : 0);
      if (o.pauli == 1 || o.pauli == 2) active |= qs;
      else if (o.pauli == 3) diag |= qs;
      break;
    }
  }
}

int build_plan(pqs_plan &p, int n, const pqs_op *ops, int n_ops, const pqs_plan_opts *opts) {
  if (n < 1 || n > kMaxQubits) return fail(PQS_ERR_ARG, "n_qubits must be in 1..34");
  if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(PQS_ERR_ARG, "bad op array");
  pqs_plan_opts o{};
  if (opts) o = *opts;
  if (o.dtype != PQS_C128 && o.dtype != PQS_C64) return fail(PQS_ERR_ARG, "dtype must be PQS_C128 or PQS_C64");
  int kmax = o.dtype == PQS_C128 ? 13 : 14;
  int k = o.tile_bits > 0 ? o.tile_bits : kmax;
  if (k > kmax) return fail(PQS_ERR_ARG, "tile_bits exceeds the LDS tile limit (13 for c128, 14 for c64)");
  if (k < 1) return fail(PQS_ERR_ARG, "tile_bits must be >= 1");
  k = std::min(k, n);
  int low = o.low_bits > 0 ? o.low_bits : 5;
  low = std::min(low, k);
  int block = o.block > 0 ? o.block : 512;
  if (block != 256 && block != 512 && block != 1024) return fail(PQS_ERR_ARG, "block must be 256, 512 or 1024");

  p.n = n;
  p.tile_bits = k;
  p.low_bits = low;
  p.block = block;
  p.dtype = o.dtype;
  p.flags = o.flags;
  p.src_ops.assign(ops, ops + n_ops);
  p.exec.clear();
  p.sweeps.clear();
  p.n_tangents = 0;
  p.n_sub = 0;

  for (int i = 0; i < n_ops; ++i) {
    int rc = validate(n, ops[i], i);
    if (rc) return rc;
    if (ops[i].tangent >= 0) p.n_tangents = std::max(p.n_tangents, ops[i].tangent + 1);
  }

  // Merge consecutive generator terms of one tangent into a Pauli sum.  A sum
  // that acts non-diagonally on more qubits than a tile holds is split into
  // parts of consecutive terms, each within the tile; the first part keeps
  // the tangent, every further part becomes a sub-tangent whose state is
  // summed into the tangent's row after the run.  D psi = sum_parts D_part
  // psi, so the split is exact (reference shared_parameter.derivative with
  // commute=True sums the sub-gates' generators, gates.py:451-457).
  //
  // A tangent whose generator terms come in several runs (separated by other
  // ops: one parameter on several gates) is differentiated by the product
  // rule: every run after its first is a sub-tangent too, so the tangent's
  // state is the sum over the runs of the evolution with that run's
  // generator inserted.
  p.n_sub = p.n_tangents;
  p.tangent_parent.clear();
  for (int t = 0; t < p.n_tangents; ++t) p.tangent_parent.push_back(t);
  std::vector<char> gen_seen(size_t(p.n_tangents), 0);
  auto part_tangent = [&](int t, bool first_part) {
    if (first_part && !gen_seen[size_t(t)]) {
      gen_seen[size_t(t)] = 1;
      return t;
    }
    p.tangent_parent.push_back(t);
    return p.n_sub++;
  };
  for (int i = 0; i < n_ops;) {
    ExecOp e;
    e.first = i;
    e.tangent = ops[i].tangent;
    masks_of(n, ops[i], e.active, e.diag);
    if (ops[i].kind == PQS_OP_GEN) {
      int j = i + 1;
      while (j < n_ops && ops[j].kind == PQS_OP_GEN && ops[j].tangent == ops[i].tangent) {
        qmask a, d;
        masks_of(n, ops[j], a, d);
        e.active |= a;
        e.diag |= d;
        ++j;
      }
      e.diag &= ~e.active;
      e.kind = kPauliSum;
      e.n_src = j - i;
      if (popc(e.active) <= k) e.tangent = part_tangent(ops[i].tangent, true);
      if (popc(e.active) > k) {
        // parts of at most `cap` non-diagonal qubits, balanced over the
        // fewest parts that a cap of k - 2 allows (room left in the tile)
        const int wide = popc(e.active);
        const int cap0 = std::max(2, k - 2);
        const int n_parts = (wide + cap0 - 1) / cap0;
        const int cap = std::max(2, std::min(k, (wide + n_parts - 1) / n_parts));
        int first = i;
        bool first_part = true;
        while (first < j) {
          ExecOp part;
          part.kind = kPauliSum;
          part.first = first;
          int last = first;
          while (last < j) {
            qmask a, d;
            masks_of(n, ops[last], a, d);
            if (last > first && popc(part.active | a) > cap) break;
            part.active |= a;
            part.diag |= d;
            ++last;
          }
          part.diag &= ~part.active;
          part.n_src = last - first;
          if (popc(part.active) > k) {
            char buf[200];
            std::snprintf(buf, sizeof buf, "generator term %d acts non-diagonally on more than the %d-qubit tile",
                          first, k);
            return fail(PQS_ERR_UNSUPPORTED, buf);
          }
          part.tangent = part_tangent(ops[i].tangent, first_part);
          first_part = false;
          p.exec.push_back(part);
          first = last;
        }
        i = j;
        continue;
      }
      i = j;
    } else {
      e.kind = ops[i].kind;
      if (e.kind == PQS_OP_GEN_FSIM) e.tangent = part_tangent(ops[i].tangent, true);
      ++i;
    }
    if (popc(e.active) > k) {
      char buf[200];
      std::snprintf(buf, sizeof buf,
                    "op %d acts non-diagonally on %d qubits, more than the %d-qubit tile",
                    e.first, popc(e.active), k);
      return fail(PQS_ERR_UNSUPPORTED, buf);
    }
    p.exec.push_back(e);
  }

  // Positions: pos(q) = n-1-q.  Forced low positions 0..low-1 = qubits n-1..n-low.
  qmask forced = 0;
  for (int b = 0; b < low; ++b) forced |= bitq(n - 1 - b);
  qmask all = (n == 64) ? ~qmask(0) : ((qmask(1) << n) - 1);

  // Rolling layouts: the qubits kept from one tile to the next (at least
  // `low` of them) become the lowest bit positions of the layout in between,
  // so both the contiguous tile reads and the write runs are preserved while
  // the tile moves freely over the qubits.
  p.rolling = (o.flags & PQS_PLAN_ROLLING) && n > k && k >= kRegBits + 6;

  // One sweep: greedy over the ops `rem` with `want` in the tile first; if
  // `want` leaves no room for the next op, without it.
  auto plan_one = [&](const std::vector<int> &rem, qmask want, qmask &T, std::vector<int> &run,
                      std::vector<int> &rest) {
    run.clear();
    for (int attempt = 0; attempt < 2 && run.empty(); ++attempt) {
      T = (n <= k) ? all : (attempt == 0 ? want : 0);
      run.clear();
      rest.clear();
      qmask blocked_any = 0, blocked_nd = 0;
      int n_angle = 0;
      for (int ei : rem) {
        const ExecOp &e = p.exec[ei];
        qmask touched = e.active | e.diag;
        bool dep = (touched & blocked_nd) || (e.active & blocked_any);
        bool ok = false;
        if (n_angle >= kMaxSweepOps) dep = true;  // the sweep's LDS op table is full
        if (!dep) {
          qmask need = e.active & ~T;
          if (popc(T | need) <= k) {
            T |= need;
            ok = true;
            ++n_angle;
          }
        }
        if (ok) {
          run.push_back(ei);
        } else {
          rest.push_back(ei);
          blocked_any |= touched;
          blocked_nd |= e.active;
        }
      }
      if (rem.empty()) break;
    }
  };
  // qubits of `from` ordered by their first non-diagonal use in `ops_left`
  auto by_next_use = [&](qmask from, const std::vector<int> &ops_left) {
    std::vector<int> order;
    qmask seen = 0;
    for (int ei : ops_left) {
      qmask a = p.exec[ei].active & from & ~seen;
      for (int q = 0; q < n; ++q)
        if ((a >> q) & 1) order.push_back(q);
      seen |= a;
    }
    for (int q = 0; q < n; ++q)
      if (((from & ~seen) >> q) & 1) order.push_back(q);
    return order;
  };
  // the `low` previous-tile qubits needed soonest (kept in the next tile)
  auto carry_of = [&](qmask prevT, const std::vector<int> &rem) {
    qmask want = 0;
    for (int q : by_next_use(prevT, rem)) {
      if (popc(want) >= low) break;
      want |= bitq(q);
    }
    return want;
  };
  // One sweep over `rem` after tile prevT, seeded with `want`: the greedy
  // pass, the last sweep's layout rule and the fill of the tile to k qubits.
  auto sweep_from = [&](qmask want, qmask prevT, const std::vector<int> &rem, qmask &T,
                        std::vector<int> &run, std::vector<int> &rest) {
    plan_one(rem, want, T, run, rest);
    if (!run.empty() && p.rolling && rest.empty() && (T & forced) != forced &&
        popc(want | forced) <= k) {
      // the last sweep writes the standard layout: keep its lowest qubits in
      // the tile (a shorter last sweep and one more are cheaper than a
      // scattered final write)
      qmask T2;
      std::vector<int> run2, rest2;
      plan_one(rem, want | forced, T2, run2, rest2);
      if (!run2.empty() && (T2 & forced) == forced) {
        T = T2;
        run.swap(run2);
        rest.swap(rest2);
      }
    }
    if (p.rolling) {
      // fill: the standard lowest qubits if this is the last sweep, then the
      // previous tile's qubits (longer write runs), then the lowest positions
      if (rest.empty())
        for (int b = 0; b < n && popc(T) < k; ++b) T |= bitq(n - 1 - b) & forced;
      for (int q : by_next_use(prevT & ~T, rest)) {
        if (popc(T) >= k) break;
        T |= bitq(q);
      }
    }
    // Fill the tile to k qubits with the lowest free positions.
    for (int b = 0; b < n && popc(T) < k; ++b)
      if (!(T & bitq(n - 1 - b))) T |= bitq(n - 1 - b);
  };
  // The greedy pass over `rem` after tile prevT, at most `horizon` sweeps
  // ahead: the sweeps it needs (horizon + 1 if more) and the ops still left
  // then.  Fewer sweeps is better, then fewer ops left.
  struct Ahead {
    int sweeps;
    size_t left;
    bool operator<(const Ahead &b) const { return sweeps != b.sweeps ? sweeps < b.sweeps : left < b.left; }
  };
  auto greedy_ahead = [&](int horizon, qmask prevT, std::vector<int> rem) {
    int count = 0;
    std::vector<int> run, rest;
    while (!rem.empty()) {
      if (count == horizon) return Ahead{count + 1, rem.size()};
      qmask T;
      sweep_from(carry_of(prevT, rem), prevT, rem, T, run, rest);
      if (run.empty()) return Ahead{1 << 30, rem.size()};
      prevT = T;
      rem.swap(rest);
      ++count;
    }
    return Ahead{count, 0};
  };
  // Window search (rolling plans): besides the greedy tile, try tiles seeded
  // with the carried qubits plus a window of k - low consecutive qubits, and
  // keep the one after which the greedy pass, looking `horizon` sweeps
  // ahead, needs the fewest sweeps (then leaves the fewest ops); the greedy
  // tile wins ties.  On nearest-neighbour circuits the greedy tile follows
  // the op order and leaves ragged light-cone fronts; a window lets a sweep
  // run a deep light-cone block of its qubits.  Short horizons gave the
  // fewest sweeps (a full look-ahead was both slower and worse), so the
  // partition is built for horizons 1, 2 and 3 and the one with the fewest
  // sweeps is kept.
  const bool search = p.rolling && !(o.flags & PQS_PLAN_GREEDY_SWEEPS) && low < k;
  const int win = k - low;
  auto partition = [&](int horizon, std::vector<Sweep> &out) {
    out.clear();
    std::vector<int> rem(p.exec.size());
    for (size_t i = 0; i < rem.size(); ++i) rem[i] = int(i);
    qmask prevT = 0;
    while (!rem.empty() || out.empty()) {
      const bool first = out.empty();
      qmask T = 0;
      std::vector<int> run, rest;
      const qmask want = (p.rolling && !first) ? carry_of(prevT, rem) : forced;
      sweep_from(want, prevT, rem, T, run, rest);
      if (run.empty() && !rem.empty()) return false;
      if (horizon > 0 && !rest.empty()) {
        Ahead best = greedy_ahead(horizon, T, rest);
        qmask T2;
        std::vector<int> run2, rest2;
        for (int s0 = 0; s0 + win <= n; ++s0) {
          const qmask W = want | (((qmask(1) << win) - 1) << s0);
          if (popc(W) > k) continue;
          sweep_from(W, prevT, rem, T2, run2, rest2);
          if (run2.empty() || (!first && popc(T2 & prevT) < low)) continue;
          // a last sweep writes the standard layout: its tile must hold the
          // lowest qubits (contiguous write runs)
          if (rest2.empty() && (T2 & forced) != forced) continue;
          const Ahead c = greedy_ahead(horizon, T2, rest2);
          if (c < best) {
            best = c;
            T = T2;
            run.swap(run2);
            rest.swap(rest2);
          }
        }
      }
      Sweep sw;
      for (int b = 0; b < n; ++b)
        if (T & bitq(n - 1 - b)) sw.tile_qubits.push_back(n - 1 - b);
      sw.ops = run;
      out.push_back(sw);
      prevT = T;
      rem.swap(rest);
      if (rem.empty()) break;
    }
    return true;
  };
  if (!partition(0, p.sweeps)) return fail(PQS_ERR_UNSUPPORTED, "planner made no progress");
  if (search) {
    std::vector<Sweep> alt;
    for (int horizon = 1; horizon <= 3; ++horizon)
      if (partition(horizon, alt) && alt.size() < p.sweeps.size()) p.sweeps.swap(alt);
  }
  // A single sweep reads and writes the standard layout: it runs in place,
  // with no second buffer.
  if (p.sweeps.size() < 2) p.rolling = false;

  // Layouts: L_0 and L_S standard (pos = n-1-q); in between, L_s holds the
  // qubits kept from tile s-1 lowest, then the rest of tile s, then the
  // other qubits in qubit order.
  {
    const int S = int(p.sweeps.size());
    std::vector<std::vector<int>> L(S + 1, std::vector<int>(n));
    for (int q = 0; q < n; ++q) L[0][q] = L[S][q] = n - 1 - q;
    for (int s = 1; s < S; ++s) {
      if (!p.rolling) {
        L[s] = L[0];
        continue;
      }
      qmask Tp = 0, Tc = 0;
      for (int q : p.sweeps[s - 1].tile_qubits) Tp |= bitq(q);
      for (int q : p.sweeps[s].tile_qubits) Tc |= bitq(q);
      std::vector<int> keep;
      for (int q = 0; q < n; ++q)
        if ((Tp & Tc) & bitq(q)) keep.push_back(q);
      std::sort(keep.begin(), keep.end(), [&](int a, int b) { return L[s - 1][a] < L[s - 1][b]; });
      int pos = 0;
      for (int q : keep) L[s][q] = pos++;
      for (int q = 0; q < n; ++q)
        if ((Tc & ~Tp) & bitq(q)) L[s][q] = pos++;
      for (int q = 0; q < n; ++q)
        if (!(Tc & bitq(q))) L[s][q] = pos++;
    }
    for (int s = 0; s < S; ++s) {
      Sweep &sw = p.sweeps[s];
      sw.rpos = L[s];
      sw.wpos = L[s + 1];
      std::sort(sw.tile_qubits.begin(), sw.tile_qubits.end(),
                [&](int a, int b) { return sw.rpos[a] < sw.rpos[b]; });
    }
  }

  p.tangent_sweep.assign(p.n_sub, -1);
  for (size_t s = 0; s < p.sweeps.size(); ++s)
    for (int ei : p.sweeps[s].ops) {
      const int tg = p.exec[ei].tangent;
      if (tg >= 0 && p.tangent_sweep[tg] < 0) p.tangent_sweep[tg] = int(s);
    }

  // ---- register phases and device image
  // Tiles of >= 11 qubits use the register-phase executor: 2^kRegBits
  // amplitudes per thread, so at least one full wave per tile.
  p.reg_exec = p.tile_bits >= kRegBits + 6;  // at least one full wave per tile
  p.dev_ops.clear();
  p.dev_terms.clear();
  p.dev_sweeps.clear();
  p.dev_phases.clear();
  for (Sweep &sw : p.sweeps) {
    DevSweep ds{};
    int tile_index[kMaxQubits];
    for (int q = 0; q < n; ++q) tile_index[q] = -1;
    std::vector<int> tpos, cpos, cq;
    for (size_t i = 0; i < sw.tile_qubits.size(); ++i) {
      tile_index[sw.tile_qubits[i]] = int(i);
      tpos.push_back(sw.rpos[sw.tile_qubits[i]]);
    }
    {
      std::vector<int> qat(n);
      for (int q = 0; q < n; ++q) qat[sw.rpos[q]] = q;
      for (int b = 0; b < n; ++b)
        if (tile_index[qat[b]] < 0) {
          cpos.push_back(b);
          cq.push_back(qat[b]);
        }
    }
    auto runs = [](const std::vector<int> &pos, int32_t *rp, int32_t *rl, int32_t *rs) {
      int nr = 0;
      for (size_t i = 0; i < pos.size();) {
        size_t j = i + 1;
        while (j < pos.size() && pos[j] == pos[j - 1] + 1) ++j;
        rp[nr] = pos[i];
        rl[nr] = int(j - i);
        rs[nr] = int(i);
        ++nr;
        i = j;
      }
      return nr;
    };
    ds.k = int(tpos.size());
    ds.nruns = runs(tpos, ds.run_pos, ds.run_len, ds.run_shift);
    ds.cnruns = runs(cpos, ds.crun_pos, ds.crun_len, ds.crun_shift);
    ds.n_tiles = int64_t(1) << cpos.size();
    ds.rolling = p.rolling ? 1 : 0;
    ds.n = n;
    {
      for (int i = 0; i < ds.k; ++i) ds.wpos[i] = sw.wpos[sw.tile_qubits[i]];
      for (size_t j = 0; j < cq.size(); ++j) ds.cwpos[j] = sw.wpos[cq[j]];
      std::vector<int> perm(ds.k);
      for (int i = 0; i < ds.k; ++i) perm[i] = i;
      std::sort(perm.begin(), perm.end(), [&](int a, int b) { return ds.wpos[a] < ds.wpos[b]; });
      for (int i = 0; i < ds.k; ++i) ds.sperm[i] = perm[i];
    }

    // Phases: LDS ops (two-qubit dense, generators) run alone and order
    // everything around them; between them, greedy register phases of at
    // most kRegBits qubits with the same dependency rule as the sweeps.
    struct PhaseSpec {
      bool lds;
      std::vector<int> ops;
      qmask regs;
    };
    std::vector<PhaseSpec> phases;
    if (p.reg_exec) {
      auto reg_capable = [&](int ei) {
        const ExecOp &e = p.exec[ei];
        const int kind = e.kind;
        if (kind == kPauliSum) return e.n_src == 1 && ops[e.first].q1 < 0;  // one Pauli on one qubit
        return kind == PQS_OP_MAT1 || kind == PQS_OP_RX || kind == PQS_OP_RY || kind == PQS_OP_RZ ||
               kind == PQS_OP_CX || kind == PQS_OP_CZ || kind == PQS_OP_RZZ;
      };
      qmask tmask = 0;  // qubits of this sweep's tile
      for (int q : sw.tile_qubits) tmask |= bitq(q);
      size_t i = 0;
      while (i < sw.ops.size()) {
        if (!reg_capable(sw.ops[i])) {
          phases.push_back({true, {sw.ops[i]}, 0});
          ++i;
          continue;
        }
        size_t j = i;
        while (j < sw.ops.size() && reg_capable(sw.ops[j])) ++j;
        std::vector<int> rem(sw.ops.begin() + i, sw.ops.begin() + j);
        // One register phase over `rem`, its register set grown from `seed`
        // in op order with the sweep's dependency rule.
        auto phase_from = [&](const std::vector<int> &from, qmask seed, qmask &R, std::vector<int> &run,
                              std::vector<int> &rest) {
          R = seed;
          run.clear();
          rest.clear();
          qmask blocked_any = 0, blocked_nd = 0;
          for (int ei : from) {
            const ExecOp &e = p.exec[ei];
            qmask touched = e.active | e.diag;
            bool dep = (touched & blocked_nd) || (e.active & blocked_any);
            // A CNOT control or CZ qubit held in registers makes the op a
            // register rename; from thread bits it needs per-lane selects.
            const qmask soft = ((e.kind == PQS_OP_CX || e.kind == PQS_OP_CZ) && !(p.flags & PQS_PLAN_NO_SOFT_REGS))
                                   ? (e.diag & tmask)
                                   : 0;
            if (!dep && popc(R | e.active | soft) <= kRegBits) {
              R |= e.active | soft;
              run.push_back(ei);
            } else if (!dep && popc(R | e.active) <= kRegBits) {
              R |= e.active;
              run.push_back(ei);
            } else {
              rest.push_back(ei);
              blocked_any |= touched;
              blocked_nd |= e.active;
            }
          }
        };
        // The plain greedy pass over `from`, at most `lookahead` phases
        // ahead: the phases it needs (lookahead + 1 if more), then the ops
        // still left.  Runs of up to kLongRun ops are rolled out to the end;
        // longer ones (a whole circuit in one tile has thousands of ops) look
        // one phase ahead, which keeps planning linear in the run's length
        // and, measured on 10-16 qubit HEA circuits, also gave fewer phases
        // than the full roll-out there.
        struct PhAhead {
          int phases;
          size_t left;
          bool operator<(const PhAhead &b) const { return phases != b.phases ? phases < b.phases : left < b.left; }
        };
        constexpr size_t kLongRun = 512;
        const int lookahead = rem.size() > kLongRun ? 1 : (1 << 30);
        auto greedy_phases = [&](std::vector<int> from) {
          int count = 0;
          qmask R;
          std::vector<int> run, rest;
          while (!from.empty()) {
            if (count == lookahead) return PhAhead{count + 1, from.size()};
            phase_from(from, 0, R, run, rest);
            if (run.empty()) return PhAhead{1 << 30, from.size()};
            from.swap(rest);
            ++count;
          }
          return PhAhead{count, 0};
        };
        // Window phase search (as for sweeps): also try register sets seeded
        // with kRegBits consecutive tile qubits and keep the phase after
        // which the greedy pass needs the fewest phases; greedy wins ties.
        const bool phase_search = !(p.flags & PQS_PLAN_GREEDY_PHASES);
        while (!rem.empty()) {
          qmask R;
          std::vector<int> run, rest;
          phase_from(rem, 0, R, run, rest);
          if (run.empty()) return fail(PQS_ERR_UNSUPPORTED, "phase planner made no progress");
          if (phase_search && !rest.empty()) {
            PhAhead best = greedy_phases(rest);
            qmask R2;
            std::vector<int> run2, rest2;
            for (int q0 = 0; q0 + kRegBits <= n; ++q0) {
              const qmask W = ((qmask(1) << kRegBits) - 1) << q0;
              if ((W & tmask) != W) continue;
              phase_from(rem, W, R2, run2, rest2);
              if (run2.empty()) continue;
              const PhAhead c = greedy_phases(rest2);
              if (c < best) {
                best = c;
                R = R2;
                run.swap(run2);
                rest.swap(rest2);
              }
            }
          }
          phases.push_back({false, run, R});
          rem.swap(rest);
        }
        i = j;
      }
      if (int(p.dev_phases.size() + phases.size()) > (1 << 24)) return fail(PQS_ERR_UNSUPPORTED, "too many phases");
    } else {
      phases.push_back({false, sw.ops, 0});
    }

    ds.op_begin = int(p.dev_ops.size());
    ds.phase_begin = int(p.dev_phases.size());
    int n_coef = 0;
    std::vector<int> exec_order;
    // tile bit read / written at positions 0-3 (-1 if none); the first /
    // last register phase wants them as lane bits
    int rbit[4] = {-1, -1, -1, -1}, wbit[4] = {-1, -1, -1, -1};
    for (int t = 0; t < ds.k; ++t) {
      if (tpos[size_t(t)] < 4) rbit[tpos[size_t(t)]] = t;
      const int w = ds.rolling ? ds.wpos[t] : tpos[size_t(t)];
      if (w < 4) wbit[w] = t;
    }
    // direct loads / stores (specialised kernels; not with persistent blocks)
    const bool direct_ok = !(p.flags & (PQS_PLAN_JIT_PERSISTENT | PQS_PLAN_NO_DIRECT));
    // Wave bits: thread bits 6.. select the wave of a workgroup.  A run of
    // register phases that keeps the same tile bits there (none of them a
    // register bit) only permutes amplitudes within each wave, so the
    // specialised kernels exchange through LDS without a workgroup barrier
    // inside the run.  Per run, the wave bits are the tile bits needed
    // latest.
    const int nwb = std::max(0, ds.k - kRegBits - 6);
    std::vector<int> phase_w(phases.size(), 0);  // tile-bit mask of the wave bits
    if (p.reg_exec && nwb > 0 && !(p.flags & PQS_PLAN_NO_WAVE_LOCAL)) {
      std::vector<int> rt(phases.size(), 0);
      for (size_t i = 0; i < phases.size(); ++i)
        for (int q = 0; q < n; ++q)
          if (((phases[i].regs >> q) & 1) && tile_index[q] >= 0) rt[i] |= 1 << tile_index[q];
      size_t i = 0;
      while (i < phases.size()) {
        if (phases[i].lds) {
          ++i;
          continue;
        }
        std::vector<std::pair<size_t, int>> cand;  // (next use, tile bit)
        for (int t = 0; t < ds.k; ++t) {
          if ((rt[i] >> t) & 1) continue;
          size_t nu = phases.size();
          for (size_t j = i; j < phases.size() && !phases[j].lds; ++j)
            if ((rt[j] >> t) & 1) {
              nu = j;
              break;
            }
          cand.push_back({nu, t});
        }
        std::stable_sort(cand.begin(), cand.end(), [](const std::pair<size_t, int> &a, const std::pair<size_t, int> &b) {
          return a.first > b.first;
        });
        // the tile bits at positions 0 and 1 should stay lane bits in the
        // first phase (read layout) and the last (write layout), where the
        // specialised kernels load / store registers directly
        auto pick = [&](int avoid) {
          std::vector<std::pair<size_t, int>> c2 = cand;
          std::stable_partition(c2.begin(), c2.end(), [&](const std::pair<size_t, int> &x) { return !((avoid >> x.second) & 1); });
          int W = 0;
          for (int c = 0; c < nwb && c < int(c2.size()); ++c) W |= 1 << c2[size_t(c)].second;
          return W;
        };
        auto extent = [&](int W) {
          size_t j = i;
          while (j < phases.size() && !phases[j].lds && !(rt[j] & W)) ++j;
          return j;
        };
        auto bits01 = [](const int *b) { return (b[0] >= 0 ? 1 << b[0] : 0) | (b[1] >= 0 ? 1 << b[1] : 0); };
        const int avoid = (i == 0 && direct_ok) ? bits01(rbit) : 0;
        int W = pick(avoid);
        size_t j = extent(W);
        if (j == phases.size() && direct_ok) {
          const int W2 = pick(avoid | bits01(wbit));
          const size_t j2 = extent(W2);
          if (j2 > i) {
            W = W2;
            j = j2;
          }
        }
        for (size_t q = i; q < j; ++q) phase_w[q] = W;
        if (j == i) ++j;  // (W excludes phase i's register bits, so phase i always fits)
        i = j;
      }
    }
    // Register tile bits of phase j after a register phase with register
    // tile bits prev (0: none): the phase's qubits, then filler bits, taken
    // first from prev (amplitudes that stay in their registers across the
    // transition), then the lowest other tile bits; never wave bits.
    auto regset_of = [&](size_t j, int prev) {
      int m = 0, nr = 0;
      for (int q = 0; q < n; ++q)
        if (((phases[j].regs >> q) & 1) && tile_index[q] >= 0) {
          m |= 1 << tile_index[q];
          ++nr;
        }
      for (int pass = 0; pass < 2; ++pass)
        for (int t = 0; t < ds.k && nr < kRegBits; ++t)
          if (!((m >> t) & 1) && !((phase_w[j] >> t) & 1) && (pass == 1 || ((prev >> t) & 1))) {
            m |= 1 << t;
            ++nr;
          }
      return m;
    };
    // Exchange transitions (PQS_PLAN_NO_EXCHANGE off): a register phase whose
    // register set differs from the previous one's by at most two tile bits,
    // all thread bits of the previous phase below the wave bits, with the
    // same wave bits, takes the previous layout with those bits traded
    // against lane bits 5 and 4; the previous phase puts them there.
    const bool exch_ok = !(p.flags & PQS_PLAN_NO_EXCHANGE);
    std::vector<int> xnew(phases.size(), 0);  // per phase: planned exchange bits (tile-bit mask), 0 none
    int prev_regs = 0;                       // register tile bits of the previous register phase
    size_t phase_no = 0;
    for (const PhaseSpec &ph : phases) {
      const size_t idx = phase_no;
      const bool first = phase_no == 0, last = phase_no + 1 == phases.size();
      const int wmask = phase_w[phase_no++];
      DevPhase dp{};
      dp.op_begin = int(p.dev_ops.size());
      dp.lds = ph.lds ? 1 : 0;
      int reg_index[kMaxTileBits];
      for (int t = 0; t < kMaxTileBits; ++t) reg_index[t] = -1;
      if (p.reg_exec && !ph.lds && xnew[idx]) {
        // the previous layout with the planned bits traded (register bit of
        // a displaced tile bit <-> the lane bit holding a new one)
        const DevPhase &pv = p.dev_phases.back();
        const int rq = regset_of(idx, prev_regs);
        dp.ntb = pv.ntb;
        for (int i = 0; i < kRegBits; ++i) dp.rb[i] = pv.rb[i];
        for (int i = 0; i < pv.ntb; ++i) dp.tb[i] = pv.tb[i];
        // the lanes that hold the new register bits (5, 4, or both)
        int lanes[2], nl = 0;
        for (int l = 5; l >= 4; --l)
          if ((rq >> pv.tb[l]) & 1) lanes[nl++] = l;
        int nd = 0;
        for (int i = 0; i < kRegBits; ++i) {
          if ((rq >> pv.rb[i]) & 1) continue;  // stays a register bit
          if (nd >= nl) return fail(PQS_ERR_UNSUPPORTED, "internal planner error: exchange lanes");
          const int lane = lanes[nd];
          dp.rb[i] = pv.tb[lane];
          dp.tb[lane] = pv.rb[i];
          dp.xr[nd] = i;
          dp.xb[nd] = lane;
          ++nd;
        }
        dp.xchg = nd;
        int got = 0;
        for (int i = 0; i < kRegBits; ++i) got |= 1 << dp.rb[i];
        if (got != rq || nd < 1 || nd > 2)
          return fail(PQS_ERR_UNSUPPORTED, "internal planner error: exchange transition mismatch");
        for (int i = 0; i < kRegBits; ++i) reg_index[dp.rb[i]] = i;
        if (direct_ok && last && wbit[0] >= 0 && wbit[1] >= 0) {
          bool l0 = false, l1 = false;
          for (int i = 0; i < std::min(6, dp.ntb); ++i) {
            l0 = l0 || dp.tb[i] == wbit[0];
            l1 = l1 || dp.tb[i] == wbit[1];
          }
          if (l0 && l1) ds.direct_store = 1;
        }
      } else if (p.reg_exec && !ph.lds) {
        // register bits: the phase's qubits, then filler bits (regset_of)
        const int rmask = regset_of(idx, prev_regs);
        int nr = 0;
        for (int q = 0; q < n; ++q)
          if ((ph.regs >> q) & 1) {
            if (tile_index[q] < 0 || nr >= kRegBits)
              return fail(PQS_ERR_UNSUPPORTED, "internal planner error: register qubit outside the tile");
            reg_index[tile_index[q]] = nr;
            dp.rb[nr++] = tile_index[q];
          }
        for (int t = 0; t < ds.k && nr < kRegBits; ++t)
          if (reg_index[t] < 0 && ((rmask >> t) & 1)) {
            reg_index[t] = nr;
            dp.rb[nr++] = t;
          }
        if (nr != kRegBits) return fail(PQS_ERR_UNSUPPORTED, "internal planner error: register bits short");
        // Thread bits; the first ones become lane bits 0, 1, 2, ... and are
        // ordered so that their images under the LDS swizzle (tile bit t
        // lands on bank-group bit t mod 4) differ, spreading each group of
        // lanes over distinct banks; the wave bits (if any) come last.
        dp.ntb = 0;
        std::vector<int> rest;
        for (int t = 0; t < ds.k; ++t)
          if (reg_index[t] < 0 && !((wmask >> t) & 1)) rest.push_back(t);
        // A CNOT whose control is a lane bit and whose target is a register
        // bit of the same bank class becomes, in the specialised kernels, a
        // per-lane store-address flip that folds two lanes of a quarter-wave
        // onto one bank group.  Among the thread bits of a class, prefer one
        // that is no such control for lanes 0-3.
        {
          std::vector<int> bad(size_t(ds.k), 0);
          for (int ei : ph.ops) {
            const ExecOp &e = p.exec[ei];
            if (e.kind != PQS_OP_CX) continue;
            const pqs_op &so = ops[e.first];
            const int tc = tile_index[so.q0], tt = tile_index[so.q1];
            if (tc >= 0 && tt >= 0 && reg_index[tc] < 0 && reg_index[tt] >= 0) bad[size_t(tc)] |= 1 << (tt & 3);
          }
          std::stable_partition(rest.begin(), rest.end(), [&](int t) { return !((bad[size_t(t)] >> (t & 3)) & 1); });
        }
        // First / last register phase: the tile bits at read / write
        // positions 0-3 that are thread bits here become the first lane bits
        // (in bank-class order), provided positions 0 and 1 are among them:
        // then every 4 lanes of a direct load / store touch 64 contiguous
        // bytes.
        auto in_rest = [&](int t) { return t >= 0 && std::find(rest.begin(), rest.end(), t) != rest.end(); };
        const int *lowb = (first && !ph.lds) ? rbit : (last && !ph.lds) ? wbit : nullptr;
        if (lowb && in_rest(lowb[0]) && in_rest(lowb[1]) && direct_ok) {
          int want = 0;
          for (int i = 0; i < 4; ++i)
            if (in_rest(lowb[i])) want |= 1 << lowb[i];
          for (int cls = 0; cls < 4; ++cls)
            for (int t = 0; t < ds.k; ++t)
              if (((want >> t) & 1) && (t & 3) == cls) dp.tb[dp.ntb++] = t;
          rest.erase(std::remove_if(rest.begin(), rest.end(), [&](int t) { return (want >> t) & 1; }), rest.end());
        }
        // Lane bits 0-2 address one group of 8 lanes: a 16-byte LDS read is
        // conflict-free when the group's slots differ mod 16 (their bank
        // classes differ), a write when they differ mod 8 (class 3 is
        // bit 3 of the slot, which writes do not separate).  So the first
        // round takes classes 0, 1, 2 and then 3.
        if (!(p.flags & PQS_PLAN_LANES_ASCENDING) && dp.ntb < 4) {
          int used = 0;
          for (int i = 0; i < dp.ntb; ++i) used |= 1 << (dp.tb[i] & 3);
          for (int cls = 0; cls < 4; ++cls)
            for (size_t i = 0; i < rest.size() && !((used >> cls) & 1); ++i)
              if ((rest[i] & 3) == cls) {
                dp.tb[dp.ntb++] = rest[i];
                used |= 1 << cls;
                rest.erase(rest.begin() + long(i));
              }
        }
        for (int round = 1; round < 4 && !rest.empty() && !(p.flags & PQS_PLAN_LANES_ASCENDING); ++round) {
          int used = 0;
          for (size_t i = 0; i < rest.size();) {
            const int cls = rest[i] & 3;
            if (!(used & (1 << cls))) {
              used |= 1 << cls;
              dp.tb[dp.ntb++] = rest[i];
              rest.erase(rest.begin() + long(i));
            } else {
              ++i;
            }
          }
        }
        for (int t : rest) dp.tb[dp.ntb++] = t;
        for (int t = 0; t < ds.k; ++t)
          if ((wmask >> t) & 1) dp.tb[dp.ntb++] = t;
        // Plan an exchange into the next phase: its new register bits go to
        // lane bits 5 and 4 here (each may take either).  A bit already on
        // lane 3-5 moves freely; from lanes 0-2 a move is preferred against
        // a bit of its bank class (the lane order's conflict rule), else it
        // costs the phase's LDS accesses a possible conflict, which an
        // exchange more than repays (it saves a whole LDS round trip).  The
        // lanes a direct load of this phase needs stay put.
        if (exch_ok && idx + 1 < phases.size() && !phases[idx + 1].lds && phase_w[idx + 1] == wmask &&
            dp.ntb >= 6) {
          int rp = 0;
          for (int i = 0; i < kRegBits; ++i) rp |= 1 << dp.rb[i];
          const int rq = regset_of(idx + 1, rp);
          const int nb = rq & ~rp;
          int nbits[2], cnt = 0;
          for (int t = 0; t < ds.k; ++t)
            if ((nb >> t) & 1) {
              if (cnt < 2) nbits[cnt] = t;
              ++cnt;
            }
          int protect = 0;
          if (first && direct_ok)
            for (int i = 0; i < 4; ++i)
              if (rbit[i] >= 0) protect |= 1 << rbit[i];
          int best_cost = 1 << 30;
          DevPhase best = dp;
          const int lanes_of[2][2] = {{5, 4}, {4, 5}};
          for (int order = 0; cnt >= 1 && cnt <= 2 && order < 2; ++order) {
            DevPhase trial = dp;
            int cost = 0;
            bool ok = true;
            for (int j = 0; ok && j < cnt; ++j) {
              const int t = nbits[j], lane = lanes_of[order][j];
              int pos = -1;
              for (int i = 0; i < 6; ++i)
                if (trial.tb[i] == t) pos = i;
              if (pos < 0 || ((protect >> t) & 1)) {
                ok = false;
                break;
              }
              if (pos == lane) continue;
              const int other = trial.tb[lane];
              if ((protect >> other) & 1) {
                ok = false;
                break;
              }
              if (pos < 3 && (other & 3) != (t & 3)) ++cost;
              std::swap(trial.tb[pos], trial.tb[lane]);
            }
            if (ok && cost < best_cost) {
              best_cost = cost;
              best = trial;
            }
          }
          if (best_cost < (1 << 30)) {
            dp = best;
            xnew[idx + 1] = nb;
          }
        }
        if (direct_ok) {
          auto is_lane = [&](int t) {
            for (int i = 0; i < std::min(6, dp.ntb); ++i)
              if (dp.tb[i] == t) return true;
            return false;
          };
          if (first && rbit[0] >= 0 && rbit[1] >= 0 && is_lane(rbit[0]) && is_lane(rbit[1])) ds.direct_load = 1;
          if (last && wbit[0] >= 0 && wbit[1] >= 0 && is_lane(wbit[0]) && is_lane(wbit[1])) ds.direct_store = 1;
        }
        if (wmask && dp.ntb != ds.k - kRegBits)
          return fail(PQS_ERR_UNSUPPORTED, "internal planner error: wave bits overlap the register bits");
      }
      if (p.reg_exec && !ph.lds) {
        prev_regs = 0;
        for (int i2 = 0; i2 < kRegBits; ++i2) prev_regs |= 1 << dp.rb[i2];
      } else {
        prev_regs = 0;
      }
      if (p.reg_exec && !ph.lds) {
        // internal consistency: register + thread bits permute the tile bits
        int seen = 0;
        for (int i2 = 0; i2 < kRegBits; ++i2) seen |= 1 << dp.rb[i2];
        for (int i2 = 0; i2 < dp.ntb; ++i2) seen |= 1 << dp.tb[i2];
        if (seen != (1 << ds.k) - 1 || dp.ntb != ds.k - kRegBits)
          return fail(PQS_ERR_UNSUPPORTED, "internal planner error: phase bits do not cover the tile");
      }
      for (int ei : ph.ops) {
        exec_order.push_back(ei);
        const ExecOp &e = p.exec[ei];
        const pqs_op &so = ops[e.first];
        DevOp d{};
        d.kind = e.kind == kPauliSum ? ((p.reg_exec && !ph.lds) ? kGen1 : kPauliSum) : so.kind;
        d.tangent = e.tangent;
        d.slot = so.slot;
        d.slot2 = so.slot2;
        d.pauli = so.pauli;
        d.scale = so.scale;
        d.offset = so.offset;
        std::memcpy(d.m, so.m, sizeof d.m);
        d.g0 = sw.rpos[so.q0];
        d.t0 = tile_index[so.q0];
        d.g1 = so.q1 >= 0 ? sw.rpos[so.q1] : -1;
        d.t1 = so.q1 >= 0 ? tile_index[so.q1] : -1;
        d.r0 = d.t0 >= 0 ? reg_index[d.t0] : -1;
        d.r1 = d.t1 >= 0 ? reg_index[d.t1] : -1;
        d.coef = -1;
        if (p.reg_exec && !ph.lds) {
          const qmask act = e.active;
          if (((act >> so.q0) & 1) && d.r0 < 0) return fail(PQS_ERR_UNSUPPORTED, "internal planner error: active qubit not in registers");
          if (so.q1 >= 0 && ((act >> so.q1) & 1) && d.r1 < 0)
            return fail(PQS_ERR_UNSUPPORTED, "internal planner error: active qubit not in registers");
        }
        if (p.reg_exec && !ph.lds &&
            (d.kind == PQS_OP_RX || d.kind == PQS_OP_RY || d.kind == PQS_OP_RZ || d.kind == PQS_OP_RZZ))
          d.coef = n_coef++;
        if (d.kind == kPauliSum) {
          d.term0 = int(p.dev_terms.size());
          d.nterms = e.n_src;
          for (int t = e.first; t < e.first + e.n_src; ++t) {
            const pqs_op &g = ops[t];
            DevTerm dt{};
            int nq = g.q1 >= 0 ? 2 : 1;
            int qs[2] = {g.q0, g.q1};
            int ny = 0;
            for (int a = 0; a < nq; ++a) {
              int q = qs[a];
              bool x = g.pauli == 1 || g.pauli == 2;
              bool z = g.pauli == 2 || g.pauli == 3;
              if (g.pauli == 2) ++ny;
              if (x) dt.xm |= 1u << tile_index[q];  // active qubits are in the tile
              if (z) {
                if (tile_index[q] >= 0) dt.zm_tile |= 1u << tile_index[q];
                else dt.zm_glob |= uint64_t(1) << sw.rpos[q];
              }
            }
            double cr = g.m[0], ci = g.m[1];  // coef * i^ny
            for (int r = 0; r < ny; ++r) {
              double t2 = cr;
              cr = -ci;
              ci = t2;
            }
            dt.cre = cr;
            dt.cim = ci;
            p.dev_terms.push_back(dt);
          }
        }
        p.dev_ops.push_back(d);
      }
      dp.op_end = int(p.dev_ops.size());
      if (p.reg_exec) p.dev_phases.push_back(dp);
    }
    sw.ops = exec_order;  // report the execution order
    ds.op_end = int(p.dev_ops.size());
    ds.phase_end = int(p.dev_phases.size());
    if (!p.reg_exec) ds.phase_begin = ds.phase_end = 0;
    ds.n_coef = 0;
    for (const PhaseSpec &ph : phases)
      if (ph.lds) ds.n_coef = 1;
    (void)n_coef;
    p.dev_sweeps.push_back(ds);
  }
  return PQS_OK;
}

}  // namespace pqs

extern "C" {

const char *pqs_last_error(void) {
  static thread_local std::string copy;
  copy = pqs::g_err;
  return copy.c_str();
}

int pqs_abi_version(void) { return PQS_ABI_VERSION; }

int pqs_plan_create(int32_t n_qubits, const pqs_op *ops, int32_t n_ops,
                    const pqs_plan_opts *opts, pqs_plan **out) {
  if (!out) return pqs::fail(PQS_ERR_ARG, "out is NULL");
  *out = nullptr;
  pqs_plan *p = new (std::nothrow) pqs_plan;
  if (!p) return pqs::fail(PQS_ERR_NOMEM, "host allocation failed");
  int rc;
  try {
    rc = pqs::build_plan(*p, n_qubits, ops, n_ops, opts);
  } catch (...) {
    rc = pqs::fail(PQS_ERR_NOMEM, "host allocation failed while planning");
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return PQS_OK;
}

int pqs_plan_info(const pqs_plan *p, int32_t *n_sweeps, int32_t *tile_bits,
                  int32_t *n_exec_ops, int32_t *n_tangents) {
  if (!p) return pqs::fail(PQS_ERR_ARG, "plan is NULL");
  if (n_sweeps) *n_sweeps = int32_t(p->sweeps.size());
  if (tile_bits) *tile_bits = p->tile_bits;
  if (n_exec_ops) *n_exec_ops = int32_t(p->exec.size());
  if (n_tangents) *n_tangents = p->n_tangents;
  return PQS_OK;
}

int pqs_plan_sweep_layout(const pqs_plan *p, int32_t s, int32_t *read_pos, int32_t *write_pos) {
  if (!p) return pqs::fail(PQS_ERR_ARG, "plan is NULL");
  if (s < 0 || s >= int32_t(p->sweeps.size())) return pqs::fail(PQS_ERR_ARG, "sweep index out of range");
  const pqs::Sweep &sw = p->sweeps[s];
  for (int q = 0; q < p->n; ++q) {
    if (read_pos) read_pos[q] = sw.rpos[q];
    if (write_pos) write_pos[q] = sw.wpos[q];
  }
  return PQS_OK;
}

int pqs_plan_sweep(const pqs_plan *p, int32_t s, int32_t *tile_qubits, int32_t *n_tile,
                   int32_t *op_index, int32_t *n_op) {
  if (!p) return pqs::fail(PQS_ERR_ARG, "plan is NULL");
  if (s < 0 || s >= int32_t(p->sweeps.size())) return pqs::fail(PQS_ERR_ARG, "sweep index out of range");
  const pqs::Sweep &sw = p->sweeps[s];
  if (n_tile) *n_tile = int32_t(sw.tile_qubits.size());
  if (tile_qubits)
    for (size_t i = 0; i < sw.tile_qubits.size(); ++i) tile_qubits[i] = sw.tile_qubits[i];
  if (n_op) *n_op = int32_t(sw.ops.size());
  if (op_index)
    for (size_t i = 0; i < sw.ops.size(); ++i) op_index[i] = p->exec[sw.ops[i]].first;
  return PQS_OK;
}

int pqs_plan_tangent_parts(const pqs_plan *p, int32_t t, int32_t *n_parts) {
  if (!p || !n_parts) return pqs::fail(PQS_ERR_ARG, "NULL argument");
  if (t < 0 || t >= p->n_tangents) return pqs::fail(PQS_ERR_ARG, "tangent out of range");
  int c = 0;
  for (int u = 0; u < p->n_sub; ++u)
    if (p->tangent_parent[size_t(u)] == t && (u == t || u >= p->n_tangents)) ++c;
  *n_parts = c;
  return PQS_OK;
}

int pqs_plan_tangent_sweep(const pqs_plan *p, int32_t t, int32_t *sweep) {
  if (!p || !sweep) return pqs::fail(PQS_ERR_ARG, "NULL argument");
  if (t < 0 || t >= p->n_tangents) return pqs::fail(PQS_ERR_ARG, "tangent out of range");
  *sweep = p->tangent_sweep[t];
  return PQS_OK;
}

}  // extern "C"
