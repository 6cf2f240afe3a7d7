// Run-time specialisation of sweeps (hiprtc).
//
// The interpreter kernel (engine.hip, k_sweep_reg) dispatches every op at run
// time on its kind and register bits; with a register array of 16 complex
// amplitudes per thread, the compiler has to bring every specialised variant
// back to one register assignment at each dispatch join, which costs about
// as many register moves as the arithmetic itself.  For a plan that will run
// many times (batched sampling, Jacobians, benchmarks) the planner instead
// emits one straight-line kernel per sweep: register bits, tile / global bit
// positions, op kinds and fixed matrices are compile-time constants, so the
// whole op sequence is SSA code with no joins and no moves, and each
// rotation costs one FMA per output component (scaled form, below).
//
// Rotations are applied as f * M with f = cos(a/2) and
// M = [[1, -t], [t, 1]] (R_y), [[1, -i t], [-i t, 1]] (R_x) or
// diag(1 -/+ i t) (R_z, R_zz), t = tan(a/2).  The scalar f is the same for
// every amplitude; it is accumulated over the sweep's phases and applied
// once at the end (or at the end of a phase that flips its sign per lane).
// |cos| is clamped below at cmin (sign kept), which changes an entry of the
// matrix by at most cmin, and only when |cos(a/2)| < cmin.  The scaled
// matrix has norm 1/|cos|, so amplitudes grow by at most 1/|f| since f was
// last applied; near a = pi that is up to 1/cmin per rotation.  Every K
// rotations within a phase the kernel therefore applies f early when |f|
// has fallen below T.  Between two such checks |f| stays above T * cmin^K,
// which bounds the growth and keeps f a normal number:
//   double: cmin 1e-20, K 10, T 1e-100 -> |f| >= 1e-300;
//   float:  cmin 1e-7,  K 4,  T 1e-9   -> |f| >= 1e-37.
// With random angles |f| halves per rotation on average, so the early
// application is rare; the check itself is one compare per K rotations.
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>
#include <vector>

#include "plan.hpp"

namespace pqs {

void set_error(const std::string &s);

static const char *kPrelude = R"PQS(
typedef unsigned long long u64;
struct InstDev {
  void *state;
  const void *src;
  const double *params;
  void *alt;
  int tangent, init, start_sweep, pingpong;
};
#define NA 16
__device__ __forceinline__ V cmul(V a, V b) { return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ int swz(int j) { return j ^ (((j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15); }
// LDS tile element at a byte offset (phase addresses are XORs of byte offsets)
#define LDSB(o) (*reinterpret_cast<V *>(reinterpret_cast<char *>(tile) + (o)))
template <int B> __device__ __forceinline__ void ry(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    a[v] = V{fma(-t, y.x, x.x), fma(-t, y.y, x.y)};
    a[v | (1 << B)] = V{fma(t, x.x, y.x), fma(t, x.y, y.y)};
  }
}
template <int B> __device__ __forceinline__ void rx(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    a[v] = V{fma(t, y.y, x.x), fma(-t, y.x, x.y)};
    a[v | (1 << B)] = V{fma(t, x.y, y.x), fma(-t, x.x, y.y)};
  }
}
// x (1 - i t)
__device__ __forceinline__ V pm(V x, R t) { return V{fma(t, x.y, x.x), fma(-t, x.x, x.y)}; }
template <int B> __device__ __forceinline__ void rz(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = pm(a[v], (v & (1 << B)) ? -t : t);
}
template <int B0, int B1> __device__ __forceinline__ void rzz(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = pm(a[v], (((v >> B0) ^ (v >> B1)) & 1) ? -t : t);
}
__device__ __forceinline__ void pall(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = pm(a[v], t);
}
template <int B> __device__ __forceinline__ void mat1(V (&a)[NA], V m00, V m01, V m10, V m11) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    a[v] = cadd(cmul(m00, x), cmul(m01, y));
    a[v | (1 << B)] = cadd(cmul(m10, x), cmul(m11, y));
  }
}
__device__ __forceinline__ void scale(V (&a)[NA], V d) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = cmul(a[v], d);
}
__device__ __forceinline__ void rscale(V (&a)[NA], R f) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = V{a[v].x * f, a[v].y * f};
}
template <int BT, int BC> __device__ __forceinline__ void cxr(V (&a)[NA]) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if ((v & (1 << BT)) || !(v & (1 << BC))) continue;
    const V x = a[v];
    a[v] = a[v | (1 << BT)];
    a[v | (1 << BT)] = x;
  }
}
template <int BT> __device__ __forceinline__ void cxc(V (&a)[NA], int on) {
  // per-component selects (a select of whole structs can be lowered
  // through scratch memory)
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << BT)) continue;
    const V x = a[v], y = a[v | (1 << BT)];
    a[v].x = on ? y.x : x.x;
    a[v].y = on ? y.y : x.y;
    a[v | (1 << BT)].x = on ? x.x : y.x;
    a[v | (1 << BT)].y = on ? x.y : y.y;
  }
}
template <int B0, int B1> __device__ __forceinline__ void czr(V (&a)[NA]) {
#pragma unroll
  for (int v = 0; v < NA; ++v)
    if ((v & (1 << B0)) && (v & (1 << B1))) a[v] = V{-a[v].x, -a[v].y};
}
template <int B> __device__ __forceinline__ void czc(V (&a)[NA], int on) {
  const R s = on ? R(-1) : R(1);
#pragma unroll
  for (int v = 0; v < NA; ++v)
    if (v & (1 << B)) a[v] = V{a[v].x * s, a[v].y * s};
}
__device__ __forceinline__ void negall(V (&a)[NA], int on) { rscale(a, on ? R(-1) : R(1)); }
// Trade lane bit LB (5 or 4) with a register bit: x holds the register bit's
// 0 half, y its 1 half; y of a lane with the lane bit clear and x of its
// partner (the bit set) swap places (v_permlane32_swap / v_permlane16_swap
// on every 32-bit word).
template <int LB> __device__ __forceinline__ void lswap(V &x, V &y) {
  constexpr int nw = int(sizeof(V) / 4);
  unsigned xs[nw], ys[nw];
  __builtin_memcpy(xs, &x, sizeof(V));
  __builtin_memcpy(ys, &y, sizeof(V));
#pragma unroll
  for (int i = 0; i < nw; ++i) {
    if constexpr (LB == 5) {
      const auto r = __builtin_amdgcn_permlane32_swap(xs[i], ys[i], false, false);
      xs[i] = r[0];
      ys[i] = r[1];
    } else {
      const auto r = __builtin_amdgcn_permlane16_swap(xs[i], ys[i], false, false);
      xs[i] = r[0];
      ys[i] = r[1];
    }
  }
  __builtin_memcpy(&x, xs, sizeof(V));
  __builtin_memcpy(&y, ys, sizeof(V));
}
)PQS";

static constexpr int NA_() { return 1 << kRegBits; }
// cmin, K and T of the scaled rotation form (comment at the top of this file)
static double cos_min(const pqs_plan &p) { return p.dtype == PQS_C128 ? 1e-20 : 1e-7; }
static int rot_per_check(const pqs_plan &p) { return p.dtype == PQS_C128 ? 10 : 4; }
static double rescale_below(const pqs_plan &p) { return p.dtype == PQS_C128 ? 1e-100 : 1e-9; }
static int swz_host(int j) { return j ^ (((j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15); }

static std::string dbl(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", v);
  std::string s(b);
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

// bit expression of a qubit that is not a register bit in a phase
static std::string cbit_expr(const DevOp &d, int which) {
  const int t = which ? d.t1 : d.t0;
  const int g = which ? d.g1 : d.g0;
  std::ostringstream o;
  if (t >= 0) o << "((toff >> " << t << ") & 1)";
  else o << "int((base >> " << g << ") & 1ull)";
  return o.str();
}

bool jit_sweep_ok(const pqs_plan &p, int s) {
  if (!p.reg_exec) return false;
  const DevSweep &ds = p.dev_sweeps[s];
  if (ds.n_coef) return false;  // LDS phases (dense two-qubit ops, generators)
  for (int i = ds.op_begin; i < ds.op_end; ++i) {
    const int k = p.dev_ops[i].kind;
    if (k != PQS_OP_MAT1 && k != PQS_OP_RX && k != PQS_OP_RY && k != PQS_OP_RZ && k != PQS_OP_CX &&
        k != PQS_OP_CZ && k != PQS_OP_RZZ && k != kGen1)
      return false;
  }
  return true;
}

static void emit_sweep(const pqs_plan &p, int s, std::ostringstream &o);

static std::string prelude(const pqs_plan &p) {
  std::string t = p.dtype == PQS_C128 ? "typedef double R;\ntypedef double2 V;\n" : "typedef float R;\ntypedef float2 V;\n";
  return t + kPrelude;
}

std::string jit_source_sweep(const pqs_plan &p, int s) {
  std::ostringstream o;
  o << prelude(p);
  emit_sweep(p, s, o);
  return o.str();
}

std::string jit_source(const pqs_plan &p, std::vector<int> &sweeps_out) {
  std::ostringstream o;
  o << prelude(p);
  sweeps_out.clear();
  for (int s = 0; s < int(p.dev_sweeps.size()); ++s) {
    if (!jit_sweep_ok(p, s)) continue;
    sweeps_out.push_back(s);
    emit_sweep(p, s, o);
  }
  return o.str();
}

static void emit_sweep(const pqs_plan &p, int s, std::ostringstream &o) {
  const int S_total = int(p.dev_sweeps.size());
  {
    const DevSweep &ds = p.dev_sweeps[s];
    const int k = ds.k;
    const int S = 1 << k;
    const int nthr = S >> kRegBits;
    const int kt = k - kRegBits;
    // angle ops of this sweep, in op order
    std::vector<int> ang(ds.op_end - ds.op_begin, -1);
    std::vector<int> aops;
    for (int i = ds.op_begin; i < ds.op_end; ++i) {
      const int kd = p.dev_ops[i].kind;
      if (kd == PQS_OP_RX || kd == PQS_OP_RY || kd == PQS_OP_RZ || kd == PQS_OP_RZZ) {
        ang[i - ds.op_begin] = int(aops.size());
        aops.push_back(i);
      }
    }
    const int na = int(aops.size());
    if (na > 0) {
      o << "__device__ const int SL" << s << "[" << na << "] = {";
      for (int a = 0; a < na; ++a) o << (a ? "," : "") << p.dev_ops[aops[a]].slot;
      o << "};\n__device__ const double SC" << s << "[" << na << "] = {";
      for (int a = 0; a < na; ++a) o << (a ? "," : "") << dbl(p.dev_ops[aops[a]].scale);
      o << "};\n__device__ const double OF" << s << "[" << na << "] = {";
      for (int a = 0; a < na; ++a) o << (a ? "," : "") << dbl(p.dev_ops[aops[a]].offset);
      o << "};\n";
    }
    const bool persist = (p.flags & PQS_PLAN_JIT_PERSISTENT) != 0;
    // The last sweep of a plan also comes as a pair kernel (pqs_jit_pair_<s>)
    // when the plan has no tangent generators: block b evolves tile
    // (b mod n_tiles) of instances 2 (b / n_tiles) and 2 (b / n_tiles) + 1 one
    // after the other, keeps the first one's final amplitudes in registers
    // and writes the tile's share of <psi_2i|psi_2i+1> to pov[b].  Both
    // instances end in the same register / lane layout, so each thread pairs
    // its own amplitudes.
    const bool emit_pair = s == S_total - 1 && !persist && p.n_tangents == 0;
  for (int pm = 0; pm < (emit_pair ? 2 : 1); ++pm) {
    const bool pair = pm == 1;
    // Block b processes flat tile b (flat tile = instance * n_tiles + tile).
    // Persistent form (PQS_PLAN_JIT_PERSISTENT): block b processes tiles b,
    // b + G, b + 2G, ... and issues the loads of its next tile into
    // registers before the current tile's phases, so HBM traffic overlaps
    // compute when only one block fits per CU.
    o << "extern \"C\" __global__ __launch_bounds__(" << nthr << ") void " << (pair ? "pqs_jit_pair_" : "pqs_jit_sweep_")
      << s << "(const InstDev *__restrict__ insts, long long n_active, const V *__restrict__ init_tab"
      << (pair ? ", double2 *__restrict__ pov" : "") << ") {\n";
    o << "  __shared__ V tile[" << S << "];\n";
    o << "  __shared__ double2 tf[" << (na > 0 ? na : 1) << "];\n";
    o << "  const int tid = threadIdx.x;\n";
    if (pair) {
      o << "  __shared__ double2 pred[" << std::max(1, nthr / 64) << "];\n";
      o << "  const long long total = (n_active >> 1) << " << (p.n - k) << ";\n";
      o << "  const long long pb = blockIdx.x;\n";
      o << "  if (pb >= total) return;\n";
      // keep: the first instance's final amplitudes; (ore, oim): this
      // thread's share of the overlap.  Each pass takes its thread index
      // through an empty asm, so that nothing derived from it (lane
      // offsets, addresses) is hoisted out of the loop and held in
      // registers across both passes.
      o << "  V keep[NA];\n  double ore = 0.0, oim = 0.0;\n";
      o << "#pragma unroll 1\n  for (int half = 0; half < 2; ++half) {\n";
      o << "  int tidv = threadIdx.x;\n  asm volatile(\"\" : \"+v\"(tidv));\n  const int tid = tidv;\n";
      o << "  const long long t = ((((pb >> " << (p.n - k) << ") << 1) | half) << " << (p.n - k) << ") | (pb & "
        << (ds.n_tiles - 1) << "ll);\n";
    } else {
      o << "  const long long total = n_active << " << (p.n - k) << ";\n";
      o << "  long long t = blockIdx.x;\n";
      o << "  if (t >= total) return;\n";
    }
    // address / source helpers for a flat tile index
    o << "  auto tile_base = [&](long long ft) -> u64 {\n";
    o << "    const u64 tile_id = u64(ft) & " << (ds.n_tiles - 1) << "ull;\n";
    o << "    return 0ull";
    for (int r = 0; r < ds.cnruns; ++r)
      o << " | (((tile_id >> " << ds.crun_shift[r] << ") & " << ((1ull << ds.crun_len[r]) - 1) << "ull) << "
        << ds.crun_pos[r] << ")";
    o << ";\n  };\n";
    int tpos[kMaxTileBits];
    for (int r = 0; r < ds.nruns; ++r)
      for (int b = 0; b < ds.run_len[r]; ++b) tpos[ds.run_shift[r] + b] = ds.run_pos[r] + b;
    o << "  const u64 dtid = 0ull";
    for (int i = 0; i < kt; ++i) o << " | (u64((tid >> " << i << ") & 1) << " << tpos[i] << ")";
    o << ";\n";
    auto hp = [&](int m) {
      unsigned long long g = 0;
      for (int b = 0; b < kRegBits; ++b)
        if (m & (1 << b)) g |= 1ull << tpos[kt + b];
      return g;
    };
    o << "  auto io = [&](const InstDev &in, const V *&src, V *&dst, int &init) {\n";
    o << "    init = 0;\n";
    o << "    if (in.pingpong) {\n";
    o << "      dst = " << (((S_total - 1 - s) & 1) == 0 ? "(V *)in.state" : "(V *)in.alt") << ";\n";
    o << "      if (" << s << " == in.start_sweep) { src = (const V *)in.src; init = in.src ? 0 : in.init; }\n";
    o << "      else src = " << (((S_total - s) & 1) == 0 ? "(const V *)in.state" : "(const V *)in.alt") << ";\n";
    o << "    } else {\n";
    o << "      dst = (V *)in.state;\n";
    o << "      if (" << s << " == in.start_sweep) { src = in.src ? (const V *)in.src : dst; init = in.src ? 0 : in.init; }\n";
    o << "      else src = dst;\n";
    o << "    }\n  };\n";
    // Direct load / store: the first (last) phase's registers come from (go
    // to) HBM without staging the tile in LDS; the planner made the tile
    // bits at read (write) positions 0 and 1 lane bits of that phase, so
    // every 4 lanes touch 64 contiguous bytes.  Otherwise the tile is staged
    // with thread bit i = tile bit i and register bit b = tile bit kt + b.
    const bool dload = ds.direct_load && !persist && ds.phase_end > ds.phase_begin;
    const bool dstore = ds.direct_store && !persist && ds.phase_end > ds.phase_begin;
    int lthr[kMaxTileBits], lreg[kRegBits];  // global positions of the load's thread / register bits
    for (int i = 0; i < kt; ++i) lthr[i] = tpos[dload ? p.dev_phases[ds.phase_begin].tb[i] : i];
    for (int b = 0; b < kRegBits; ++b) lreg[b] = tpos[dload ? p.dev_phases[ds.phase_begin].rb[b] : kt + b];
    auto lhp = [&](int m) {
      unsigned long long g = 0;
      for (int b = 0; b < kRegBits; ++b)
        if (m & (1 << b)) g |= 1ull << lreg[b];
      return g;
    };
    o << "  const u64 ltid = 0ull";
    for (int i = 0; i < kt; ++i) o << " | (u64((tid >> " << i << ") & 1) << " << lthr[i] << ")";
    o << ";\n";
    o << "  auto fetch = [&](long long ft, V (&pre)[NA]) {\n";
    o << "    const InstDev in = insts[ft >> " << (p.n - k) << "];\n";
    o << "    const V *src; V *dst; int init;\n    io(in, src, dst, init);\n";
    o << "    const u64 dl = tile_base(ft) | ltid;\n";
    o << "    if (init == 0) {\n";
    for (int m = 0; m < NA_(); ++m) o << "      pre[" << m << "] = src[dl | " << lhp(m) << "ull];\n";
    o << "    } else if (init == 1) {\n";
    for (int m = 0; m < NA_(); ++m)
      o << "      pre[" << m << "] = V{(dl | " << lhp(m) << "ull) == 0ull ? R(1) : R(0), R(0)};\n";
    o << "    } else {\n";
    o << "#pragma unroll\n      for (int m = 0; m < NA; ++m) {\n";
    o << "        u64 gi = dl;\n";
    for (int b = 0; b < kRegBits; ++b) o << "        if (m & " << (1 << b) << ") gi |= " << (1ull << lreg[b]) << "ull;\n";
    o << "        V v{R(1), R(0)};\n";
    o << "        for (int q = 0; q < " << p.n << "; ++q) v = cmul(v, init_tab[2 * q + int((gi >> (" << (p.n - 1)
      << " - q)) & 1ull)]);\n";
    o << "        pre[m] = v;\n      }\n    }\n  };\n";
    o << "  V pre[NA];\n  fetch(t, pre);\n";
    // the final amplitude w of register / slot m: kept (first instance) or
    // paired with the kept one (second instance)
    auto capture = [&](const std::string &m, const std::string &w) {
      if (!pair) return;
      o << "    if (half == 0) keep[" << m << "] = " << w << ";\n"
        << "    else {\n"
        << "      ore += double(keep[" << m << "].x) * double(" << w << ".x) + double(keep[" << m << "].y) * double(" << w
        << ".y);\n"
        << "      oim += double(keep[" << m << "].x) * double(" << w << ".y) - double(keep[" << m << "].y) * double(" << w
        << ".x);\n"
        << "    }\n";
    };
    if (persist) o << "  for (; t < total; t += gridDim.x) {\n";
    else o << "  {\n";
    o << "  const InstDev in = insts[t >> " << (p.n - k) << "];\n";
    o << "  const V *src; V *dst; int init;\n  io(in, src, dst, init);\n";
    o << "  const u64 base = tile_base(t);\n";
    o << "  const u64 dl = base | dtid;\n";
    o << "  const double *params = in.params;\n";
    o << "  __syncthreads();\n";
    if (!dload)
      o << "#pragma unroll\n  for (int m = 0; m < NA; ++m) tile[swz(tid + m * " << nthr << ")] = pre[m];\n";
    if (persist) o << "  if (t + gridDim.x < total) fetch(t + gridDim.x, pre);\n";
    if (na > 0) {
      o << "  for (int i = tid; i < " << na << "; i += " << nthr << ") {\n";
      o << "    double sn, cs;\n";
      o << "    const int sl = SL" << s << "[i];\n";
      o << "    sincos(0.5 * (OF" << s << "[i] + (sl >= 0 ? SC" << s << "[i] * params[sl] : 0.0)), &sn, &cs);\n";
      o << "    if (fabs(cs) < " << dbl(cos_min(p)) << ") cs = cs < 0.0 ? -" << dbl(cos_min(p)) << " : "
        << dbl(cos_min(p)) << ";\n";
      o << "    tf[i] = double2{sn / cs, cs};\n  }\n";
    }
    o << "  __syncthreads();\n";
    // fs: the accumulated scale of the rotations (uniform over the block
    // unless a phase flips its sign per lane); it carries across phases and
    // is applied at the end of the sweep, at the end of a phase with
    // per-lane sign flips, or early by the growth check
    o << "  double fs = 1.0;\n";
    // registers carried into a phase that takes its layout by an exchange
    o << "  V areg[NA];\n";
    // (tan, cos) of the next phase's first rotations, read from the angle
    // table before this phase's LDS writes: LDS returns in issue order, so
    // a table read issued after the transition's 16 reads would hold the
    // next phase's first products until the whole transition has arrived
    constexpr int kPre = 4;
    o << "  double2 nxc[" << kPre << "];\n";
    std::vector<int> pre_ai;  // angle indices held in nxc for this phase
    auto aexpr = [&](int ai) {
      for (size_t j = 0; j < pre_ai.size(); ++j)
        if (pre_ai[j] == ai) return "nxc[" + std::to_string(j) + "]";
      return "tf[" + std::to_string(ai) + "]";
    };
    // emit the reads of the next register phase's first rotations
    auto preload = [&](int next) -> std::vector<int> {
      std::vector<int> v;
      if (next >= ds.phase_end) return v;
      const DevPhase &Q = p.dev_phases[next];
      for (int i = Q.op_begin; i < Q.op_end && int(v.size()) < kPre; ++i) {
        const int ai = ang[i - ds.op_begin];
        if (ai >= 0) v.push_back(ai);
      }
      for (size_t j = 0; j < v.size(); ++j) o << "    nxc[" << j << "] = tf[" << v[j] << "];\n";
      return v;
    };
    std::vector<int> next_pre;
    bool carry = false;
    int nrot = 0;  // rotations since the last check of the accumulated scale
    // phases
    for (int ph = ds.phase_begin; ph < ds.phase_end; ++ph) {
      const DevPhase &P = p.dev_phases[ph];
      pre_ai.swap(next_pre);
      next_pre.clear();
      bool lane_sign = false;  // this phase flips fs per lane
      o << "  {\n";
      // in a loop over tiles (persistent form, pair kernel) each phase takes
      // its lane index through an empty asm, so its lane offsets are not
      // hoisted out of the loop and held for the whole kernel; in one pass
      // the compiler may compute them early, before the previous phase's
      // LDS writes (12-bit headline plan -1.5%, others unchanged)
      if (persist || pair)
        o << "    int ptid = tid;\n    asm volatile(\"\" : \"+v\"(ptid));\n";
      else
        o << "    const int ptid = tid;\n";
      o << "    const int toff = 0";
      for (int i = 0; i < P.ntb; ++i) o << " | (((ptid >> " << i << ") & 1) << " << P.tb[i] << ")";
      o << ";\n    const int pt = swz(toff);\n";
      int kv[NA_()];
      for (int v = 0; v < NA_(); ++v) {
        int a = 0;
        for (int i = 0; i < kRegBits; ++i)
          if (v & (1 << i)) a ^= swz_host(1 << P.rb[i]);
        kv[v] = a;
      }
      o << "    const int ptb = pt * int(sizeof(V));\n    V a[NA];\n";
      if (carry) {
        for (int v = 0; v < NA_(); ++v) o << "    a[" << v << "] = areg[" << v << "];\n";
        carry = false;
      } else if (dload && ph == ds.phase_begin) {
        for (int v = 0; v < NA_(); ++v) o << "    a[" << v << "] = pre[" << v << "];\n";
      } else {
        for (int v = 0; v < NA_(); ++v)
          o << "    a[" << v << "] = LDSB(ptb ^ (" << kv[v] << " * int(sizeof(V))));\n";
      }
      // Pending X flips: a CNOT whose control is a thread or constant bit
      // does not move data; it XORs a per-lane flag xf<r> meaning "logical
      // register index bit r is stored flipped".  Later ops on r compensate
      // (X Ry(a) X = Ry(-a), X Rz(a) X = Rz(-a), X Rx X = Rx, CNOT with a
      // flipped control = CNOT then X on the target, X_a CZ X_a = CZ Z_b),
      // and the phase's LDS store addresses absorb the flips.
      bool fl[kRegBits] = {false, false, false, false};
      o << "    int xf0 = 0, xf1 = 0, xf2 = 0, xf3 = 0;\n";
      auto xf = [](int r) { return "xf" + std::to_string(r); };
      auto materialise = [&](int r) {  // apply a pending flip as data movement
        if (r >= 0 && fl[r]) {
          o << "    cxc<" << r << ">(a, " << xf(r) << ");\n    " << xf(r) << " = 0;\n";
          fl[r] = false;
        }
      };
      for (int i = P.op_begin; i < P.op_end; ++i) {
        const DevOp &d = p.dev_ops[i];
        const int ai = ang[i - ds.op_begin];
        const bool r0 = d.r0 >= 0, r1 = d.r1 >= 0;
        if (d.kind == PQS_OP_RX || d.kind == PQS_OP_RY || d.kind == PQS_OP_RZ || d.kind == PQS_OP_RZZ) {
          if (nrot == rot_per_check(p)) {
            o << "    if (fabs(fs) < " << dbl(rescale_below(p)) << ") { rscale(a, R(fs)); fs = 1.0; }\n";
            nrot = 0;
          }
          ++nrot;
        }
        // "(xf) ? -t : t" for a rotation whose sign flips with pending X flips
        auto tsign = [&](const std::string &flip, const std::string &t) {
          return flip.empty() ? t : "((" + flip + ") ? -" + t + " : " + t + ")";
        };
        switch (d.kind) {
          case PQS_OP_RY:
            o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; ry<" << d.r0 << ">(a, "
              << tsign(fl[d.r0] ? xf(d.r0) : "", "R(c.x)") << "); }\n";
            break;
          case PQS_OP_RX:
            o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; rx<" << d.r0 << ">(a, R(c.x)); }\n";
            break;
          case PQS_OP_RZ:
            if (r0)
              o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; rz<" << d.r0 << ">(a, "
                << tsign(fl[d.r0] ? xf(d.r0) : "", "R(c.x)") << "); }\n";
            else
              o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; pall(a, " << cbit_expr(d, 0)
                << " ? R(-c.x) : R(c.x)); }\n";
            break;
          case PQS_OP_RZZ: {
            std::string flip;
            if (r0 && fl[d.r0]) flip = xf(d.r0);
            if (r1 && fl[d.r1]) flip = flip.empty() ? xf(d.r1) : "(" + flip + " ^ " + xf(d.r1) + ")";
            if (r0 && r1)
              o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; rzz<" << d.r0 << "," << d.r1 << ">(a, "
                << tsign(flip, "R(c.x)") << "); }\n";
            else if (r0 || r1) {
              const std::string cb = cbit_expr(d, r0 ? 1 : 0);
              o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; rz<" << (r0 ? d.r0 : d.r1) << ">(a, ("
                << (flip.empty() ? cb : "(" + cb + " ^ " + flip + ")") << ") ? R(-c.x) : R(c.x)); }\n";
            } else
              o << "    { const double2 c = " << aexpr(ai) << "; fs *= c.y; pall(a, (" << cbit_expr(d, 0) << " ^ "
                << cbit_expr(d, 1) << ") ? R(-c.x) : R(c.x)); }\n";
            break;
          }
          case PQS_OP_MAT1: {
            auto cv = [&](int e) { return "V{R(" + dbl(d.m[2 * e]) + "), R(" + dbl(d.m[2 * e + 1]) + ")}"; };
            if (r0) {
              materialise(d.r0);
              o << "    mat1<" << d.r0 << ">(a, " << cv(0) << ", " << cv(1) << ", " << cv(2) << ", " << cv(3)
                << ");\n";
            } else
              o << "    scale(a, " << cbit_expr(d, 0) << " ? " << cv(3) << " : " << cv(0) << ");\n";
            break;
          }
          case kGen1: {
            // c * Pauli on q0, only in the instance carrying this tangent
            if (r0) materialise(d.r0);
            const double cr = d.m[0], ci = d.m[1];
            double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // m00, m01, m10, m11 (re, im)
            if (d.pauli == 0) { m[0] = cr; m[1] = ci; m[6] = cr; m[7] = ci; }
            else if (d.pauli == 1) { m[2] = cr; m[3] = ci; m[4] = cr; m[5] = ci; }
            else if (d.pauli == 2) { m[2] = ci; m[3] = -cr; m[4] = -ci; m[5] = cr; }
            else { m[0] = cr; m[1] = ci; m[6] = -cr; m[7] = -ci; }
            auto cv = [&](int e) { return "V{R(" + dbl(m[2 * e]) + "), R(" + dbl(m[2 * e + 1]) + ")}"; };
            o << "    if (in.tangent == " << d.tangent << ") ";
            if (r0)
              o << "mat1<" << d.r0 << ">(a, " << cv(0) << ", " << cv(1) << ", " << cv(2) << ", " << cv(3) << ");\n";
            else
              o << "scale(a, " << cbit_expr(d, 0) << " ? " << cv(3) << " : " << cv(0) << ");\n";
            break;
          }
          case PQS_OP_CX:
            // q0 = control, q1 = target (a register bit)
            if (r0) {
              o << "    cxr<" << d.r1 << "," << d.r0 << ">(a);\n";
              if (fl[d.r0]) {  // flipped control: CNOT then X on the target
                o << "    " << xf(d.r1) << " ^= " << xf(d.r0) << ";\n";
                fl[d.r1] = true;
              }
            } else {
              o << "    " << xf(d.r1) << " ^= " << cbit_expr(d, 0) << ";\n";
              fl[d.r1] = true;
            }
            break;
          case PQS_OP_CZ:
            if (r0 && r1) {
              o << "    czr<" << d.r0 << "," << d.r1 << ">(a);\n";
              // X_a CZ X_a = CZ Z_b; both flipped: -CZ Z_a Z_b
              if (fl[d.r0]) o << "    czc<" << d.r1 << ">(a, " << xf(d.r0) << ");\n";
              if (fl[d.r1]) o << "    czc<" << d.r0 << ">(a, " << xf(d.r1) << ");\n";
              if (fl[d.r0] && fl[d.r1]) {
                o << "    if (" << xf(d.r0) << " & " << xf(d.r1) << ") fs = -fs;\n";
                lane_sign = true;
              }
            } else if (r0 || r1) {
              const int rr = r0 ? d.r0 : d.r1;
              const std::string cb = cbit_expr(d, r0 ? 1 : 0);
              o << "    czc<" << rr << ">(a, " << cb << ");\n";
              // X_r CZ(r, c) X_r = CZ Z_c: a sign where the constant bit is set
              if (fl[rr]) {
                o << "    if (" << xf(rr) << " & " << cb << ") fs = -fs;\n";
                lane_sign = true;
              }
            } else {
              o << "    negall(a, " << cbit_expr(d, 0) << " & " << cbit_expr(d, 1) << ");\n";
            }
            break;
          default:
            break;
        }
      }
      // Exchange into the next phase: its layout is this one with register
      // bits traded against lane bits 5 / 4, so the amplitudes move with
      // cross-lane swaps and stay in registers (no LDS round trip, no
      // barrier).  Pending flips are applied as data first.
      const bool xnext = ph + 1 < ds.phase_end && p.dev_phases[ph + 1].xchg > 0 && !p.dev_phases[ph + 1].lds;
      if (xnext) {
        for (int r = 0; r < kRegBits; ++r) materialise(r);
        if (lane_sign) {
          o << "    rscale(a, R(fs));\n    fs = 1.0;\n";
          nrot = 0;
        }
        next_pre = preload(ph + 1);
        const DevPhase &Q = p.dev_phases[ph + 1];
        for (int j = 0; j < Q.xchg; ++j)
          for (int v = 0; v < NA_(); ++v)
            if (!(v & (1 << Q.xr[j])))
              o << "    lswap<" << Q.xb[j] << ">(a[" << v << "], a[" << (v | (1 << Q.xr[j])) << "]);\n";
        for (int v = 0; v < NA_(); ++v) o << "    areg[" << v << "] = a[" << v << "];\n";
        o << "  }\n";
        carry = true;
        continue;
      }
      // stores land where the flipped indices belong
      {
        bool any = false;
        for (int r = 0; r < kRegBits; ++r) any = any || fl[r];
        if (any) {
          o << "    const int pst = pt";
          for (int r = 0; r < kRegBits; ++r)
            if (fl[r]) o << " ^ (" << xf(r) << " ? " << swz_host(1 << P.rb[r]) << " : 0)";
          o << ";\n";
        } else {
          o << "    const int pst = pt;\n";
        }
      }
      if (lane_sign || ph + 1 == ds.phase_end) {
        o << "    rscale(a, R(fs));\n    fs = 1.0;\n";
        nrot = 0;
      }
      if (dstore && ph + 1 == ds.phase_end) {
        // direct store: tile bit j goes to write position wp[j]; flipped
        // register bits flip their position bit per lane
        int wp[kMaxTileBits];
        for (int j = 0; j < k; ++j) wp[j] = ds.rolling ? ds.wpos[j] : tpos[j];
        if (ds.rolling) {
          o << "    const u64 wtid = u64(t) & " << (ds.n_tiles - 1) << "ull;\n";
          o << "    const u64 gb = 0ull";
          for (int j = 0; j < p.n - k; ++j) o << " | (((wtid >> " << j << ") & 1ull) << " << ds.cwpos[j] << ")";
          o << ";\n";
        } else {
          o << "    const u64 gb = tile_base(t);\n";
        }
        o << "    const u64 gt = gb";
        for (int i = 0; i < P.ntb; ++i) o << " | (u64((ptid >> " << i << ") & 1) << " << wp[P.tb[i]] << ")";
        for (int r = 0; r < kRegBits; ++r)
          if (fl[r]) o << " ^ (" << xf(r) << " ? " << (1ull << wp[P.rb[r]]) << "ull : 0ull)";
        o << ";\n";
        for (int v = 0; v < NA_(); ++v) {
          unsigned long long cm = 0;
          for (int b = 0; b < kRegBits; ++b)
            if (v & (1 << b)) cm |= 1ull << wp[P.rb[b]];
          o << "    dst[gt ^ " << cm << "ull] = a[" << v << "];\n";
        }
        for (int v = 0; v < NA_(); ++v) capture(std::to_string(v), "a[" + std::to_string(v) + "]");
        o << "  }\n";
        continue;
      }
      next_pre = preload(ph + 1);
      o << "    const int pstb = pst * int(sizeof(V));\n";
      for (int v = 0; v < NA_(); ++v)
        o << "    LDSB(pstb ^ (" << kv[v] << " * int(sizeof(V)))) = a[" << v << "];\n";
      // the next phase keeps the wave bits (thread bits 6..): its LDS reads
      // touch only this wave's amplitudes, which this wave's own earlier
      // LDS writes (executed in order) already hold
      bool wave_local = false;
      if (ph + 1 < ds.phase_end && !(p.flags & PQS_PLAN_NO_WAVE_LOCAL)) {
        const DevPhase &Q = p.dev_phases[ph + 1];
        wave_local = !Q.lds && Q.ntb == P.ntb;
        for (int i = 6; wave_local && i < P.ntb; ++i) wave_local = Q.tb[i] == P.tb[i];
      }
      if (wave_local)
        o << "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");\n"
             "    __builtin_amdgcn_wave_barrier();\n"
             "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n  }\n";
      else
        o << "    __syncthreads();\n  }\n";
    }
    if (dstore) {
      // stored by the last phase
    } else if (!ds.rolling) {
      for (int m = 0; m < NA_(); ++m) {
        o << "  {\n    const V w = tile[swz(tid + " << m * nthr << ")];\n    dst[dl | " << hp(m) << "ull] = w;\n";
        capture(std::to_string(m), "w");
        o << "  }\n";
      }
    } else {
      // write layout: store index bit i is tile bit sperm[i], at write
      // position wpos[sperm[i]] (ascending), complement bit j at cwpos[j]
      o << "  {\n    const u64 wtid = u64(t) & " << (ds.n_tiles - 1) << "ull;\n";
      o << "    const u64 wdl = 0ull";
      for (int j = 0; j < p.n - k; ++j) o << " | (((wtid >> " << j << ") & 1ull) << " << ds.cwpos[j] << ")";
      for (int i = 0; i < kt; ++i) o << " | (u64((tid >> " << i << ") & 1) << " << ds.wpos[ds.sperm[i]] << ")";
      o << ";\n    const int jdl = 0";
      for (int i = 0; i < kt; ++i) o << " | (((tid >> " << i << ") & 1) << " << ds.sperm[i] << ")";
      o << ";\n";
      for (int m = 0; m < NA_(); ++m) {
        unsigned long long cm = 0;
        int jm = 0;
        for (int b = 0; b < kRegBits; ++b)
          if (m & (1 << b)) {
            cm |= 1ull << ds.wpos[ds.sperm[kt + b]];
            jm |= 1 << ds.sperm[kt + b];
          }
        o << "    {\n    const V w = tile[swz(jdl | " << jm << ")];\n    dst[wdl | " << cm << "ull] = w;\n";
        capture(std::to_string(m), "w");
        o << "    }\n";
      }
      o << "  }\n";
    }
    if (!pair) {
      o << "  }\n}\n";
      continue;
    }
    // close the tile scope and the loop over the pair; sum the threads'
    // shares in a fixed order (wave, then waves)
    o << "  }\n  }\n";
    o << "  for (int off = 32; off > 0; off >>= 1) {\n"
         "    ore += __shfl_xor(ore, off);\n    oim += __shfl_xor(oim, off);\n  }\n";
    o << "  if ((tid & 63) == 0) pred[tid >> 6] = double2{ore, oim};\n  __syncthreads();\n";
    o << "  if (tid == 0) {\n    double2 sum = pred[0];\n";
    if (nthr / 64 > 1)
      o << "    for (int w = 1; w < " << nthr / 64 << "; ++w) { sum.x += pred[w].x; sum.y += pred[w].y; }\n";
    o << "    pov[pb] = sum;\n  }\n}\n";
  }
  }
}

static const char *kJitOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                 "-Xclang", "-packed-fp32-ops"};

// What besides the source decides the code object: the compile options and
// the hiprtc version (part of the on-disk cache key).
std::string jit_signature() {
  std::string sig;
  for (const char *o : kJitOpts) {
    sig += o;
    sig += ' ';
  }
  int major = 0, minor = 0;
  hiprtcVersion(&major, &minor);
  sig += "hiprtc " + std::to_string(major) + "." + std::to_string(minor);
  return sig;
}

int jit_compile(const std::string &src, std::vector<char> &code, std::string &log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pqs_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return PQS_ERR_HIP;
  }
  hiprtcResult r = hiprtcCompileProgram(prog, int(sizeof kJitOpts / sizeof kJitOpts[0]), kJitOpts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return PQS_ERR_HIP;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return PQS_OK;
}

}  // namespace pqs
