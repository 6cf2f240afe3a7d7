// HIP kernels for gfx950 (MI355X) and the C ABI entry points of pqs.h.
//
// Hot path (reference: the Python loop ``for g in self.gates: circuit_state =
// g * circuit_state``, circuit.py:118-125, where every gate is a full 2^N x 2^N
// QuTiP CSR operator rebuilt on each parameter change, gates.py:122-131):
// a planned circuit runs as a sequence of *sweeps*.  One workgroup owns one
// tile of 2^k amplitudes of one instance: it gathers the tile from HBM into
// LDS (the low bit positions are always in the tile, so each wave reads
// contiguous runs), applies every op the planner put in that sweep, and
// writes the tile back.  Each sweep costs one read and one write of the state
// however many gates it fuses.  A batch of instances (random parameter
// vectors, derivative tangents) shares one launch per sweep: the grid is
// tiles x instances.
//
// Reductions (overlap, Meyer-Wallach, Gram / QFIM, Pauli expectation) are
// two-stage with a fixed order so that every result is bit-identical whatever
// else runs in the same batch.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plan.hpp"

namespace pqs {
void set_error(const std::string &s);
}

namespace {

using pqs::DevOp;
using pqs::DevSweep;
using pqs::DevTerm;
using pqs::DevPhase;

int fail(int code, const std::string &msg) {
  pqs::set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s failed: %s", what, hipGetErrorString(e));
  pqs::set_error(buf);
  return PQS_ERR_HIP;
}

#define HIPCHECK(x)                                   \
  do {                                                \
    hipError_t e_ = (x);                              \
    if (e_ != hipSuccess) return hip_fail(e_, #x);    \
  } while (0)

// roctx ranges around the entry points (seen by rocprofv3 --marker-trace):
// with PQS_ROCTX=1 the library loads the rocprofiler-sdk roctx library (the
// one rocprofv3 intercepts; the older libroctx64 otherwise) at first use;
// without it the ranges cost one branch.
struct Roctx {
  int (*push)(const char *) = nullptr;
  int (*pop)() = nullptr;
};
const Roctx &roctx() {
  static const Roctx r = [] {
    Roctx x;
    const char *e = std::getenv("PQS_ROCTX");
    if (e && e[0] == '1') {
      void *h = nullptr;
      for (const char *lib : {"librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1",
                              "libroctx64.so", "/opt/rocm/lib/libroctx64.so"})
        if (!h) h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
      if (h) {
        x.push = reinterpret_cast<int (*)(const char *)>(dlsym(h, "roctxRangePushA"));
        x.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
        if (!x.push || !x.pop) x.push = nullptr, x.pop = nullptr;
      }
    }
    return x;
  }();
  return r;
}
struct Range {
  const bool on;
  explicit Range(const char *name) : on(roctx().push != nullptr) {
    if (on) roctx().push(name);
  }
  ~Range() {
    if (on) roctx().pop();
  }
};

// Instance record as the sweep kernel reads it.
struct InstDev {
  void *state;          // destination state (and source once started)
  const void *src;      // source at the start sweep (NULL: state itself / init)
  const double *params; // parameter vector
  void *alt;            // second buffer of a ping-pong (Jacobian plain state)
  int32_t tangent;      // -1 or tangent index
  int32_t init;         // at the start sweep: 0 read, 1 |0..0>, 2 product table
  int32_t start_sweep;  // first sweep this instance takes part in
  int32_t pingpong;     // 1: out-of-place every sweep between state and alt
};

template <typename R> struct CT;
template <> struct CT<double> { using V = double2; };
template <> struct CT<float> { using V = float2; };

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
  return V{a.x + b.x, a.y + b.y};
}
// a * conj(b)
template <typename V>
__device__ __forceinline__ double2 cmulc(V a, V b) {
  return double2{double(a.x) * b.x + double(a.y) * b.y, double(a.y) * b.x - double(a.x) * b.y};
}

__device__ __forceinline__ int ins0(int u, int t) {
  return ((u >> t) << (t + 1)) | (u & ((1 << t) - 1));
}

__device__ __forceinline__ double op_angle(const DevOp *op, const double *params) {
  double a = op->offset;
  if (op->slot >= 0) a += op->scale * params[op->slot];
  return a;
}

// 2x2 complex matrix in double: m[0..3] = m00 m01 m10 m11
struct M2 {
  double2 m[4];
};
struct M4 {
  double2 m[16];
};

__device__ void mat1_of(const DevOp *op, const double *params, M2 &M) {
  const int kind = op->kind;
  if (kind == PQS_OP_MAT1) {
#pragma unroll
    for (int i = 0; i < 4; ++i) M.m[i] = double2{op->m[2 * i], op->m[2 * i + 1]};
    return;
  }
  double s, c;
  sincos(0.5 * op_angle(op, params), &s, &c);
  if (kind == PQS_OP_RX) {
    M.m[0] = double2{c, 0};
    M.m[1] = double2{0, -s};
    M.m[2] = double2{0, -s};
    M.m[3] = double2{c, 0};
  } else if (kind == PQS_OP_RY) {
    M.m[0] = double2{c, 0};
    M.m[1] = double2{-s, 0};
    M.m[2] = double2{s, 0};
    M.m[3] = double2{c, 0};
  } else {  // RZ
    M.m[0] = double2{c, -s};
    M.m[1] = double2{0, 0};
    M.m[2] = double2{0, 0};
    M.m[3] = double2{c, s};
  }
}

__device__ void mat2_of(const DevOp *op, const double *params, M4 &M) {
  const int kind = op->kind;
  if (kind == PQS_OP_MAT2) {
#pragma unroll
    for (int i = 0; i < 16; ++i) M.m[i] = double2{op->m[2 * i], op->m[2 * i + 1]};
    return;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) M.m[i] = double2{0, 0};
  if (kind == PQS_OP_RXX || kind == PQS_OP_RYY) {
    double s, c;
    sincos(0.5 * op_angle(op, params), &s, &c);
    for (int r = 0; r < 4; ++r) M.m[5 * r] = double2{c, 0};
    // X(x)X: 00<->11, 01<->10 with +1; Y(x)Y: 00<->11 with -1, 01<->10 with +1
    double sa = kind == PQS_OP_RXX ? 1.0 : -1.0;
    M.m[3] = double2{0, -s * sa};
    M.m[12] = double2{0, -s * sa};
    M.m[6] = double2{0, -s};
    M.m[9] = double2{0, -s};
    return;
  }
  // fSim family (gates.py:598-606, 619-627, 640-648, 708-716, 729-737)
  double th = op_angle(op, params);
  double ph = op->slot2 >= 0 ? params[op->slot2] : 0.0;
  double st, ct, sp, cp;
  sincos(th, &st, &ct);
  sincos(ph, &sp, &cp);
  M.m[0] = double2{1, 0};
  int sel = kind == PQS_OP_FSIM ? 0 : op->pauli;
  if (sel == 1 || sel == 3) {  // d_theta as written by the reference
    M.m[5] = double2{-st, 0};
    M.m[6] = double2{0, -ct};
    M.m[9] = double2{0, -ct};
    M.m[10] = double2{-st, 0};
  } else {
    M.m[5] = double2{ct, 0};
    M.m[6] = double2{0, -st};
    M.m[9] = double2{0, -st};
    M.m[10] = double2{ct, 0};
  }
  if (sel == 3) {
    M.m[15] = double2{1, 0};
  } else if (sel == 2) {  // -i e^{-i phi}
    M.m[15] = double2{-sp, -cp};
  } else {  // e^{-i phi}
    M.m[15] = double2{cp, -sp};
  }
}

template <typename R>
__device__ __forceinline__ typename CT<R>::V cvt(double2 a) {
  return typename CT<R>::V{R(a.x), R(a.y)};
}

__device__ __forceinline__ int swz(int j) { return j ^ (((j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15); }

template <typename R, int MAXPER, bool SWZ>
__device__ __forceinline__ void apply_lds_op(typename CT<R>::V *tile, const int S, const DevOp *op,
                                             const uint64_t base, const double *params,
                                             const DevTerm *terms, const int tid, const int nthr) {
  using V = typename CT<R>::V;
  const int kind = op->kind;
  const int t0 = op->t0, t1 = op->t1;
  auto IX = [](int j) { return SWZ ? swz(j) : j; };
  switch (kind) {
      case PQS_OP_MAT1:
      case PQS_OP_RX:
      case PQS_OP_RY:
      case PQS_OP_RZ: {
        M2 M;
        mat1_of(op, params, M);
        if (t0 < 0) {  // diagonal op on a bit constant over the tile
          const V d = cvt<R>(M.m[((base >> op->g0) & 1) ? 3 : 0]);
          for (int j = tid; j < S; j += nthr) tile[IX(j)] = cmul(tile[IX(j)], d);
        } else if (kind == PQS_OP_RZ) {
          const V d0 = cvt<R>(M.m[0]), d1 = cvt<R>(M.m[3]);
          for (int j = tid; j < S; j += nthr) tile[IX(j)] = cmul(tile[IX(j)], ((j >> t0) & 1) ? d1 : d0);
        } else {
          const V m00 = cvt<R>(M.m[0]), m01 = cvt<R>(M.m[1]), m10 = cvt<R>(M.m[2]), m11 = cvt<R>(M.m[3]);
          for (int u = tid; u < (S >> 1); u += nthr) {
            const int j0 = ins0(u, t0), j1 = j0 | (1 << t0);
            const V a = tile[IX(j0)], b = tile[IX(j1)];
            tile[IX(j0)] = cadd(cmul(m00, a), cmul(m01, b));
            tile[IX(j1)] = cadd(cmul(m10, a), cmul(m11, b));
          }
        }
        break;
      }
      case PQS_OP_CX: {
        const bool ctl_in = t0 >= 0;
        const bool ctl_on = !ctl_in && ((base >> op->g0) & 1);
        if (ctl_in || ctl_on) {
          for (int u = tid; u < (S >> 1); u += nthr) {
            const int j0 = ins0(u, t1);
            if (ctl_in && !((j0 >> t0) & 1)) continue;
            const int j1 = j0 | (1 << t1);
            const V a = tile[IX(j0)];
            tile[IX(j0)] = tile[IX(j1)];
            tile[IX(j1)] = a;
          }
        }
        break;
      }
      case PQS_OP_CZ:
      case PQS_OP_RZZ: {
        V d[4];
        if (kind == PQS_OP_CZ) {
          d[0] = d[1] = d[2] = V{R(1), R(0)};
          d[3] = V{R(-1), R(0)};
        } else {
          double s_, c_;
          sincos(0.5 * op_angle(op, params), &s_, &c_);
          d[0] = d[3] = V{R(c_), R(-s_)};
          d[1] = d[2] = V{R(c_), R(s_)};
        }
        const int c0 = t0 < 0 ? int((base >> op->g0) & 1) : -1;
        const int c1 = t1 < 0 ? int((base >> op->g1) & 1) : -1;
        for (int j = tid; j < S; j += nthr) {
          const int b0 = c0 >= 0 ? c0 : ((j >> t0) & 1);
          const int b1 = c1 >= 0 ? c1 : ((j >> t1) & 1);
          tile[IX(j)] = cmul(tile[IX(j)], d[2 * b0 + b1]);
        }
        break;
      }
      case PQS_OP_MAT2:
      case PQS_OP_RXX:
      case PQS_OP_RYY:
      case PQS_OP_FSIM:
      case PQS_OP_GEN_FSIM: {
        M4 M;
        mat2_of(op, params, M);
        V m[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = cvt<R>(M.m[i]);
        const int lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
        for (int u = tid; u < (S >> 2); u += nthr) {
          const int j = ins0(ins0(u, lo), hi);
          const int jj[4] = {j, j | (1 << t1), j | (1 << t0), j | (1 << t0) | (1 << t1)};
          V a[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) a[c] = tile[IX(jj[c])];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            V acc = cmul(m[4 * r], a[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(m[4 * r + c], a[c]));
            tile[IX(jj[r])] = acc;
          }
        }
        break;
      }
      case pqs::kPauliSum: {
                V acc[MAXPER];
        const int nt = op->nterms;
        const DevTerm *tm = terms + op->term0;
#pragma unroll
        for (int m = 0; m < MAXPER; ++m) {
          const int j = tid + m * nthr;
          V a{R(0), R(0)};
          if (j < S) {
            for (int t = 0; t < nt; ++t) {
              const int jj = j ^ int(tm[t].xm);
              const int par = __popc(uint32_t(jj) & tm[t].zm_tile) + __popcll(base & tm[t].zm_glob);
              V p = cmul(V{R(tm[t].cre), R(tm[t].cim)}, tile[IX(jj)]);
              if (par & 1) p = V{-p.x, -p.y};
              a = cadd(a, p);
            }
          }
          acc[m] = a;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < MAXPER; ++m) {
          const int j = tid + m * nthr;
          if (j < S) tile[IX(j)] = acc[m];
        }
        break;
      }
    default:
      break;
  }
}

// One tile of one instance through one sweep, op by op in LDS (tiles below
// 10 qubits).  `sw` is the sweep (a kernel argument, or a record in device
// memory for launches that mix plans).
template <typename R, int BLOCK>
__device__ __forceinline__ void sweep_lds_body(const DevSweep &sw, const DevOp *__restrict__ ops,
                                               const DevTerm *__restrict__ terms, const InstDev &in,
                                               const int64_t tile_id, const typename CT<R>::V *__restrict__ init_tab,
                                               int n, int s, int n_sweeps) {
  using V = typename CT<R>::V;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V *tile = reinterpret_cast<V *>(smem_raw);
  const int tid = threadIdx.x;
  const int S = 1 << sw.k;

  uint64_t base = 0;
  for (int r = 0; r < sw.cnruns; ++r)
    base |= (uint64_t(tile_id >> sw.crun_shift[r]) & ((uint64_t(1) << sw.crun_len[r]) - 1))
            << sw.crun_pos[r];

  const V *src;
  V *dst;
  int init = 0;
  if (in.pingpong) {
    dst = ((n_sweeps - 1 - s) & 1) == 0 ? (V *)in.state : (V *)in.alt;
    if (s == in.start_sweep) {
      src = (const V *)in.src;
      init = in.src ? 0 : in.init;
    } else {
      src = ((n_sweeps - s) & 1) == 0 ? (const V *)in.state : (const V *)in.alt;
    }
  } else {
    dst = (V *)in.state;
    if (s == in.start_sweep) {
      src = in.src ? (const V *)in.src : dst;
      init = in.src ? 0 : in.init;
    } else {
      src = dst;
    }
  }

  // ---- gather the tile
  for (int j = tid; j < S; j += BLOCK) {
    uint64_t gi = base;
    for (int r = 0; r < sw.nruns; ++r)
      gi |= uint64_t((j >> sw.run_shift[r]) & ((1 << sw.run_len[r]) - 1)) << sw.run_pos[r];
    V v;
    if (init == 0) {
      v = src[gi];
    } else if (init == 1) {
      v = V{gi == 0 ? R(1) : R(0), R(0)};
    } else {
      v = V{R(1), R(0)};
      for (int q = 0; q < n; ++q) v = cmul(v, init_tab[2 * q + int((gi >> (n - 1 - q)) & 1)]);
    }
    tile[j] = v;
  }

  const double *params = in.params;
  for (int o = sw.op_begin; o < sw.op_end; ++o) {
    const DevOp *op = ops + o;
    const int tg = op->tangent;
    if (tg >= 0 && tg != in.tangent) continue;
    __syncthreads();
    apply_lds_op<R, (1 << (sizeof(R) == 8 ? 13 : 14)) / BLOCK, false>(tile, S, op, base, params, terms, tid, BLOCK);
  }
  __syncthreads();

  // ---- scatter the tile
  for (int j = tid; j < S; j += BLOCK) {
    uint64_t gi = base;
    for (int r = 0; r < sw.nruns; ++r)
      gi |= uint64_t((j >> sw.run_shift[r]) & ((1 << sw.run_len[r]) - 1)) << sw.run_pos[r];
    dst[gi] = tile[j];
  }
}

template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sweep(const DevSweep sw, const DevOp *__restrict__ ops,
                                                 const DevTerm *__restrict__ terms,
                                                 const InstDev *__restrict__ insts, int64_t n_active,
                                                 const typename CT<R>::V *__restrict__ init_tab, int n,
                                                 int s, int n_sweeps) {
  const int64_t gid = blockIdx.x;
  const int64_t inst_idx = gid / sw.n_tiles;
  const int64_t tile_id = gid - inst_idx * sw.n_tiles;
  if (inst_idx >= n_active) return;
  const InstDev in = insts[inst_idx];
  sweep_lds_body<R, BLOCK>(sw, ops, terms, in, tile_id, init_tab, n, s, n_sweeps);
}

// ------------------------------------------------------- register executor
// Tiles of >= 10 qubits: each thread holds NA = 16 amplitudes (kRegBits = 4
// tile bits) in VGPRs; a 2^13 tile uses 512 threads (two waves per SIMD).
// A sweep is a list of phases; a register phase reads the thread's 16
// amplitudes from LDS for its register bits, applies all its ops in registers
// (the register bit of every op is dispatched to a template instance, so
// array indices stay compile-time constants), and writes them back.  LDS
// traffic is one read + one write of the tile per phase instead of per op.
// The sweep's ops are staged in LDS as compact records with their cos / sin
// precomputed, and the LDS tile is XOR-swizzled so that every phase's access
// pattern spreads over the banks.
constexpr int NA = 1 << pqs::kRegBits;

// Run a statement with a compile-time register bit B_ (or C_) equal to the
// runtime value r (0..3), so that register arrays keep constant indices.
#define SW4(r, ...)                                       \
  switch (r) {                                            \
    case 0: { constexpr int B_ = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int B_ = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int B_ = 2; __VA_ARGS__; } break; \
    default: { constexpr int B_ = 3; __VA_ARGS__; } break; \
  }
#define SW4C(r, ...)                                      \
  switch (r) {                                            \
    case 0: { constexpr int C_ = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int C_ = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int C_ = 2; __VA_ARGS__; } break; \
    default: { constexpr int C_ = 3; __VA_ARGS__; } break; \
  }

// Rotations in scaled form.  R_y(a) = c [[1, -t], [t, 1]] with t = s / c
// (|c| >= |s|, form A) or s [[u, -1], [1, u]] with u = c / s (form B); the
// scalar factor (c or s) is the same for every amplitude, so it is folded
// into a per-tile running factor and the pair update costs one FMA per
// output component.  R_x and the diagonal R_z / R_zz phases likewise.
template <int B, bool FB, typename V, typename R>
__device__ __forceinline__ void reg_ry(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    if constexpr (!FB) {  // [[1, -t], [t, 1]]
      a[v] = V{fma(-t, y.x, x.x), fma(-t, y.y, x.y)};
      a[v | (1 << B)] = V{fma(t, x.x, y.x), fma(t, x.y, y.y)};
    } else {  // [[u, -1], [1, u]]
      a[v] = V{fma(t, x.x, -y.x), fma(t, x.y, -y.y)};
      a[v | (1 << B)] = V{fma(t, y.x, x.x), fma(t, y.y, x.y)};
    }
  }
}

template <int B, bool FB, typename V, typename R>
__device__ __forceinline__ void reg_rx(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    if constexpr (!FB) {  // [[1, -i t], [-i t, 1]]
      a[v] = V{fma(t, y.y, x.x), fma(-t, y.x, x.y)};
      a[v | (1 << B)] = V{fma(t, x.y, y.x), fma(-t, x.x, y.y)};
    } else {  // [[u, -i], [-i, u]]
      a[v] = V{fma(t, x.x, y.y), fma(t, x.y, -y.x)};
      a[v | (1 << B)] = V{fma(t, y.x, x.y), fma(t, y.y, -x.x)};
    }
  }
}

// x * (1 - i t) (form A) or x * (u - i) (form B); conj for the other sign
template <bool FB, bool NEG, typename V, typename R>
__device__ __forceinline__ V phase_mul(V x, R t) {
  if constexpr (!FB) {
    return NEG ? V{fma(t, x.y, x.x), fma(-t, x.x, x.y)} : V{fma(-t, x.y, x.x), fma(t, x.x, x.y)};
  } else {
    return NEG ? V{fma(t, x.x, x.y), fma(t, x.y, -x.x)} : V{fma(t, x.x, -x.y), fma(t, x.y, x.x)};
  }
}

// diag(e^{-ia/2}, e^{+ia/2}) on register bit B (scaled)
template <int B, bool FB, typename V, typename R>
__device__ __forceinline__ void reg_rz(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = (v & (1 << B)) ? phase_mul<FB, false>(a[v], t) : phase_mul<FB, true>(a[v], t);
}

// e^{-ia/2} on even parity of register bits (B0, B1), e^{+ia/2} on odd (scaled)
template <int B0, int B1, bool FB, typename V, typename R>
__device__ __forceinline__ void reg_rzz(V (&a)[NA], R t) {
  if constexpr (B0 != B1) {
#pragma unroll
    for (int v = 0; v < NA; ++v)
      a[v] = (((v >> B0) ^ (v >> B1)) & 1) ? phase_mul<FB, false>(a[v], t) : phase_mul<FB, true>(a[v], t);
  }
}

// the same phase (sign NEG) on every amplitude (scaled)
template <bool FB, bool NEG, typename V, typename R>
__device__ __forceinline__ void reg_phase_all(V (&a)[NA], R t) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = phase_mul<FB, NEG>(a[v], t);
}

template <int B, typename V>
__device__ __forceinline__ void reg_mat1(V (&a)[NA], V m00, V m01, V m10, V m11) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << B)) continue;
    const V x = a[v], y = a[v | (1 << B)];
    a[v] = cadd(cmul(m00, x), cmul(m01, y));
    a[v | (1 << B)] = cadd(cmul(m10, x), cmul(m11, y));
  }
}

// multiply amplitudes with register bit B clear by d0, set by d1
template <int B, typename V>
__device__ __forceinline__ void reg_diag(V (&a)[NA], V d0, V d1) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = cmul(a[v], (v & (1 << B)) ? d1 : d0);
}

template <typename V>
__device__ __forceinline__ void reg_scale(V (&a)[NA], V d) {
#pragma unroll
  for (int v = 0; v < NA; ++v) a[v] = cmul(a[v], d);
}

template <typename V>
__device__ __forceinline__ V cneg(V x) {
  return V{-x.x, -x.y};
}

// CNOT with target register bit BT and control register bit BC
template <int BT, int BC, typename V>
__device__ __forceinline__ void reg_cx(V (&a)[NA]) {
  if constexpr (BT != BC) {
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      if ((v & (1 << BT)) || !(v & (1 << BC))) continue;
      const V x = a[v];
      a[v] = a[v | (1 << BT)];
      a[v | (1 << BT)] = x;
    }
  }
}

// X on register bit BT (a CNOT whose control is a set constant bit)
template <int BT, typename V>
__device__ __forceinline__ void reg_x(V (&a)[NA]) {
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    if (v & (1 << BT)) continue;
    const V x = a[v];
    a[v] = a[v | (1 << BT)];
    a[v | (1 << BT)] = x;
  }
}

// -1 where register bits B0 and B1 are both set
template <int B0, int B1, typename V>
__device__ __forceinline__ void reg_cz(V (&a)[NA]) {
  if constexpr (B0 != B1) {
#pragma unroll
    for (int v = 0; v < NA; ++v)
      if ((v & (1 << B0)) && (v & (1 << B1))) a[v] = cneg(a[v]);
  }
}

// -1 where register bit B is set
template <int B, typename V>
__device__ __forceinline__ void reg_z(V (&a)[NA]) {
#pragma unroll
  for (int v = 0; v < NA; ++v)
    if (v & (1 << B)) a[v] = cneg(a[v]);
}

// Compact op record staged in LDS (32 bytes).
struct alignas(16) ROp {
  int32_t kind;
  int32_t regs;  // r0 | r1 << 8 (0xff: not a register bit)
  int32_t t01;   // t0 | t1 << 8 (tile bit, 0xff: outside the tile)
  int32_t g01;   // g0 | g1 << 8 (global bit position)
  double t, f;   // rotations: ratio t (or u) and scalar factor; form B if regs bit 16
};

// One tile of one instance through one sweep of register phases (tiles of
// >= 10 qubits).  `sw` is the sweep (a kernel argument, or a record in
// device memory for launches that mix plans).
template <typename R, int MAXT, bool LDSOPS>
__device__ __forceinline__ void sweep_reg_body(const DevSweep &sw, const DevOp *__restrict__ ops,
                                               const DevTerm *__restrict__ terms,
                                               const DevPhase *__restrict__ phases, const InstDev &in,
                                               const int64_t tile_id,
                                               const typename CT<R>::V *__restrict__ init_tab, int n, int s,
                                               int n_sweeps) {
  using V = typename CT<R>::V;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V *tile = reinterpret_cast<V *>(smem_raw);
  const int S = 1 << sw.k;
  ROp *rops = reinterpret_cast<ROp *>(smem_raw + size_t(S) * sizeof(V));
  const int tid = threadIdx.x;
  const int nthr = S >> pqs::kRegBits;

  uint64_t base = 0;
  for (int r = 0; r < sw.cnruns; ++r)
    base |= (uint64_t(tile_id >> sw.crun_shift[r]) & ((uint64_t(1) << sw.crun_len[r]) - 1))
            << sw.crun_pos[r];

  const V *src;
  V *dst;
  int init = 0;
  if (in.pingpong) {
    dst = ((n_sweeps - 1 - s) & 1) == 0 ? (V *)in.state : (V *)in.alt;
    if (s == in.start_sweep) {
      src = (const V *)in.src;
      init = in.src ? 0 : in.init;
    } else {
      src = ((n_sweeps - s) & 1) == 0 ? (const V *)in.state : (const V *)in.alt;
    }
  } else {
    dst = (V *)in.state;
    if (s == in.start_sweep) {
      src = in.src ? (const V *)in.src : dst;
      init = in.src ? 0 : in.init;
    } else {
      src = dst;
    }
  }
  const double *params = in.params;

  // ---- gather the tile into LDS (coalesced; swizzled slots).  Tile index
  // j = tid + m * nthr: the low k-4 bits come from tid, the top 4 from m, so
  // the global index is base | dep(tid) | dep(m << (k-4)); all loads of a
  // thread are issued before any is consumed.
  const int kt = sw.k - pqs::kRegBits;
  uint64_t dl = base;
  uint64_t hp[pqs::kRegBits];
  for (int r = 0; r < sw.nruns; ++r) {
    for (int b = 0; b < sw.run_len[r]; ++b) {
      const int i = sw.run_shift[r] + b;  // tile bit
      const uint64_t gb = uint64_t(1) << (sw.run_pos[r] + b);
      if (i < kt) {
        if ((tid >> i) & 1) dl |= gb;
      } else {
        hp[i - kt] = gb;
      }
    }
  }
  auto gaddr = [&](int m) -> uint64_t {
    uint64_t g = dl;
#pragma unroll
    for (int b = 0; b < pqs::kRegBits; ++b)
      if (m & (1 << b)) g |= hp[b];
    return g;
  };
  {
    V tmp[NA];
    if (init == 0) {
#pragma unroll
      for (int m = 0; m < NA; ++m) tmp[m] = src[gaddr(m)];
    } else if (init == 1) {
#pragma unroll
      for (int m = 0; m < NA; ++m) tmp[m] = V{gaddr(m) == 0 ? R(1) : R(0), R(0)};
    } else {
#pragma unroll
      for (int m = 0; m < NA; ++m) {
        const uint64_t gi = gaddr(m);
        V v{R(1), R(0)};
        for (int q = 0; q < n; ++q) v = cmul(v, init_tab[2 * q + int((gi >> (n - 1 - q)) & 1)]);
        tmp[m] = v;
      }
    }
#pragma unroll
    for (int m = 0; m < NA; ++m) tile[swz(tid + m * nthr)] = tmp[m];
  }
  // ---- stage the sweep's op records (with cos / sin of half angles)
  for (int o = sw.op_begin + tid; o < sw.op_end; o += nthr) {
    const DevOp *d = ops + o;
    ROp r;
    r.kind = d->kind;
    r.regs = (d->r0 & 0xff) | ((d->r1 & 0xff) << 8);
    r.t01 = (d->t0 & 0xff) | ((d->t1 & 0xff) << 8);
    r.g01 = (d->g0 & 0xff) | ((d->g1 & 0xff) << 8);
    r.t = 0.0;
    r.f = 1.0;
    if (d->kind == pqs::kGen1) {
      // coefficient in (t, f); Pauli in regs bits 17-18; tangent split over
      // the high halves of t01 and g01
      r.t = d->m[0];
      r.f = d->m[1];
      r.regs |= (d->pauli & 3) << 17;
      const uint32_t tg = uint32_t(d->tangent);
      r.t01 = int32_t(uint32_t(r.t01) | ((tg & 0xffffu) << 16));
      r.g01 = int32_t(uint32_t(r.g01) | ((tg >> 16) << 16));
    }
    if (d->coef >= 0) {
      double sn, cs;
      sincos(0.5 * op_angle(d, params), &sn, &cs);
      if (fabs(cs) >= fabs(sn)) {
        r.t = sn / cs;
        r.f = cs;
      } else {
        r.t = cs / sn;
        r.f = sn;
        r.regs |= 1 << 16;
      }
    }
    rops[o - sw.op_begin] = r;
  }
  __syncthreads();

  for (int ph = sw.phase_begin; ph < sw.phase_end; ++ph) {
    const DevPhase *P = phases + ph;
    if (P->lds) {
      if constexpr (LDSOPS) {
        for (int o = P->op_begin; o < P->op_end; ++o) {
          const DevOp *op = ops + o;
          if (op->tangent >= 0 && op->tangent != in.tangent) continue;
          apply_lds_op<R, NA, true>(tile, S, op, base, params, terms, tid, nthr);
          __syncthreads();
        }
      }
      continue;
    }
    // thread part of the tile index and the swizzled register-bit offsets
    int toff = 0;
    for (int i = 0; i < P->ntb; ++i) toff |= ((tid >> i) & 1) << P->tb[i];
    const int pt = swz(toff);
    int sr[pqs::kRegBits];
#pragma unroll
    for (int i = 0; i < pqs::kRegBits; ++i) sr[i] = swz(1 << P->rb[i]);
    auto laddr = [&](int v) {
      int addr = pt;
#pragma unroll
      for (int i = 0; i < pqs::kRegBits; ++i)
        if (v & (1 << i)) addr ^= sr[i];
      return addr;
    };
    V a[NA];
#pragma unroll
    for (int v = 0; v < NA; ++v) a[v] = tile[laddr(v)];
    double fscale = 1.0;
    // constant bit of a qubit that is not a register bit
    auto cbit = [&](int t, int g) -> int {
      return t != 0xff ? ((toff >> t) & 1) : int((base >> g) & 1);
    };
    for (int o = P->op_begin; o < P->op_end; ++o) {
      const ROp op = rops[o - sw.op_begin];
      // the record is the same for every lane: keep its fields in SGPRs so
      // that dispatch is scalar branching, not exec-masked case walking
      const int kind = __builtin_amdgcn_readfirstlane(op.kind);
      const int regs = __builtin_amdgcn_readfirstlane(op.regs);
      const int t01 = __builtin_amdgcn_readfirstlane(op.t01);
      const int g01 = __builtin_amdgcn_readfirstlane(op.g01);
      const int r0 = regs & 0xff, r1 = (regs >> 8) & 0xff;
      const int t0 = t01 & 0xff, t1 = (t01 >> 8) & 0xff;
      const int g0 = g01 & 0xff, g1 = (g01 >> 8) & 0xff;
      const R t = R(op.t);
      const bool fb = (regs >> 16) & 1;
      switch (kind) {
        case PQS_OP_RY:
          fscale *= op.f;
          if (fb) {
            SW4(r0, reg_ry<B_, true>(a, t));
          } else {
            SW4(r0, reg_ry<B_, false>(a, t));
          }
          break;
        case PQS_OP_RX:
          fscale *= op.f;
          if (fb) {
            SW4(r0, reg_rx<B_, true>(a, t));
          } else {
            SW4(r0, reg_rx<B_, false>(a, t));
          }
          break;
        case PQS_OP_RZ:
        case PQS_OP_RZZ: {
          fscale *= op.f;
          // register bits involved and the constant bits' parity
          const bool q0r = r0 != 0xff;
          const bool q1r = kind == PQS_OP_RZZ && r1 != 0xff;
          int cpar = 0;
          if (!q0r) cpar ^= cbit(t0, g0);
          if (kind == PQS_OP_RZZ && !q1r) cpar ^= cbit(t1, g1);
          if (q0r && q1r) {
            if (fb) {
              SW4(r0, SW4C(r1, reg_rzz<B_, C_, true>(a, t)));
            } else {
              SW4(r0, SW4C(r1, reg_rzz<B_, C_, false>(a, t)));
            }
          } else if (q0r || q1r) {
            const int rr = q0r ? r0 : r1;
            // odd constant parity swaps the two phases: diag(e^{+}, e^{-})
            const R tt = cpar ? -t : t;
            if (!fb) {
              SW4(rr, reg_rz<B_, false>(a, tt));
            } else {
              // form B: x (u -/+ i); swapping the sign of i is conj of the phase
              if (cpar) {
                SW4(rr, reg_rz<B_, true>(a, -t));
                fscale = -fscale;
              } else {
                SW4(rr, reg_rz<B_, true>(a, t));
              }
            }
          } else {
            if (fb) {
              if (cpar) reg_phase_all<true, false>(a, t);
              else reg_phase_all<true, true>(a, t);
            } else {
              if (cpar) reg_phase_all<false, false>(a, t);
              else reg_phase_all<false, true>(a, t);
            }
          }
          break;
        }
        case PQS_OP_MAT1: {
          const DevOp *d = ops + o;
          const V m00{R(d->m[0]), R(d->m[1])}, m01{R(d->m[2]), R(d->m[3])};
          const V m10{R(d->m[4]), R(d->m[5])}, m11{R(d->m[6]), R(d->m[7])};
          if (r0 != 0xff) {
            SW4(r0, reg_mat1<B_>(a, m00, m01, m10, m11));
          } else {
            reg_scale(a, cbit(t0, g0) ? m11 : m00);  // diagonal by planning
          }
          break;
        }
        case pqs::kGen1: {
          const int tang = int((uint32_t(t01) >> 16) | ((uint32_t(g01) >> 16) << 16));
          if (tang != in.tangent) break;  // uniform per block: one instance per block
          const int pauli = (regs >> 17) & 3;
          const V c{R(op.t), R(op.f)};
          const V z{R(0), R(0)};
          V m00 = z, m01 = z, m10 = z, m11 = z;
          if (pauli == 0) {
            m00 = m11 = c;
          } else if (pauli == 1) {
            m01 = m10 = c;
          } else if (pauli == 2) {  // [[0, -i c], [i c, 0]]
            m01 = V{c.y, -c.x};
            m10 = V{-c.y, c.x};
          } else {
            m00 = c;
            m11 = cneg(c);
          }
          if (r0 != 0xff) {
            SW4(r0, reg_mat1<B_>(a, m00, m01, m10, m11));
          } else {
            reg_scale(a, cbit(t0 & 0xff, g0 & 0xff) ? m11 : m00);  // I or Z on a constant bit
          }
          break;
        }
        case PQS_OP_CX:
          if (r0 != 0xff) {
            SW4(r1, SW4C(r0, reg_cx<B_, C_>(a)));
          } else if (cbit(t0, g0)) {
            SW4(r1, reg_x<B_>(a));
          }
          break;
        case PQS_OP_CZ:
          if (r0 != 0xff && r1 != 0xff) {
            SW4(r0, SW4C(r1, reg_cz<B_, C_>(a)));
          } else if (r0 != 0xff) {
            if (cbit(t1, g1)) SW4(r0, reg_z<B_>(a));
          } else if (r1 != 0xff) {
            if (cbit(t0, g0)) SW4(r1, reg_z<B_>(a));
          } else if (cbit(t0, g0) && cbit(t1, g1)) {
            fscale = -fscale;
          }
          break;
        default:
          break;
      }
    }
    if (fscale != 1.0) {
      const R fs = R(fscale);
#pragma unroll
      for (int v = 0; v < NA; ++v) a[v] = V{a[v].x * fs, a[v].y * fs};
    }
#pragma unroll
    for (int v = 0; v < NA; ++v) tile[laddr(v)] = a[v];
    __syncthreads();
  }

  // ---- scatter the tile
  if (!sw.rolling) {
#pragma unroll
    for (int m = 0; m < NA; ++m) dst[gaddr(m)] = tile[swz(tid + m * nthr)];
  } else {
    // write layout: store index bits in ascending write position, so the
    // lanes of a wave cover the lowest (contiguous) write positions
    uint64_t wl = 0;
    for (int j = 0; j < sw.n - sw.k; ++j)
      if ((uint64_t(tile_id) >> j) & 1) wl |= uint64_t(1) << sw.cwpos[j];
    int jl = 0;
    for (int i = 0; i < kt; ++i)
      if ((tid >> i) & 1) {
        wl |= uint64_t(1) << sw.wpos[sw.sperm[i]];
        jl |= 1 << sw.sperm[i];
      }
#pragma unroll
    for (int m = 0; m < NA; ++m) {
      uint64_t wa = wl;
      int ja = jl;
#pragma unroll
      for (int b = 0; b < pqs::kRegBits; ++b)
        if (m & (1 << b)) {
          wa |= uint64_t(1) << sw.wpos[sw.sperm[kt + b]];
          ja |= 1 << sw.sperm[kt + b];
        }
      dst[wa] = tile[swz(ja)];
    }
  }
}

template <typename R, int MAXT, bool LDSOPS>
__global__ __launch_bounds__(MAXT) void k_sweep_reg(const DevSweep sw, const DevOp *__restrict__ ops,
                                                    const DevTerm *__restrict__ terms,
                                                    const DevPhase *__restrict__ phases,
                                                    const InstDev *__restrict__ insts, int64_t n_active,
                                                    const typename CT<R>::V *__restrict__ init_tab, int n, int s,
                                                    int n_sweeps) {
  const int64_t gid = blockIdx.x;
  const int64_t inst_idx = gid / sw.n_tiles;
  const int64_t tile_id = gid - inst_idx * sw.n_tiles;
  if (inst_idx >= n_active) return;
  const InstDev in = insts[inst_idx];
  sweep_reg_body<R, MAXT, LDSOPS>(sw, ops, terms, phases, in, tile_id, init_tab, n, s, n_sweeps);
}

// ------------------------------------------------- launches that mix plans
// A device record of a plan for launches that run instances of several
// plans (uneven widths, depths) at once, and one work item per instance of
// such a launch: its blocks are the instance's tiles, blocks tile_begin ..
// tile_begin + n_tiles - 1 of the grid (items sorted by tile_begin).
struct PlanDev {
  const DevSweep *sweeps;
  const DevOp *ops;
  const DevTerm *terms;
  const DevPhase *phases;
  int32_t n_sweeps, n;
};
struct MultiItem {
  int64_t tile_begin;
  int32_t inst, plan;
  const void *init_tab;
};

__device__ __forceinline__ int find_item(const MultiItem *__restrict__ items, int n_items, int64_t b) {
  int lo = 0, hi = n_items - 1;  // last item with tile_begin <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename R, int MAXT, bool LDSOPS>
__global__ __launch_bounds__(MAXT) void k_sweep_reg_multi(const PlanDev *__restrict__ plans,
                                                          const InstDev *__restrict__ insts,
                                                          const MultiItem *__restrict__ items, int n_items, int s) {
  const int64_t b = blockIdx.x;
  const MultiItem it = items[find_item(items, n_items, b)];
  const PlanDev pd = plans[it.plan];
  const InstDev in = insts[it.inst];
  sweep_reg_body<R, MAXT, LDSOPS>(pd.sweeps[s], pd.ops, pd.terms, pd.phases, in, b - it.tile_begin,
                                  static_cast<const typename CT<R>::V *>(it.init_tab), pd.n, s, pd.n_sweeps);
}

template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sweep_multi(const PlanDev *__restrict__ plans,
                                                      const InstDev *__restrict__ insts,
                                                      const MultiItem *__restrict__ items, int n_items, int s) {
  const int64_t b = blockIdx.x;
  const MultiItem it = items[find_item(items, n_items, b)];
  const PlanDev pd = plans[it.plan];
  const InstDev in = insts[it.inst];
  sweep_lds_body<R, BLOCK>(pd.sweeps[s], pd.ops, pd.terms, in, b - it.tile_begin,
                           static_cast<const typename CT<R>::V *>(it.init_tab), pd.n, s, pd.n_sweeps);
}

// --------------------------------------------------------------- reductions
constexpr int RB = 256;  // threads per reduction block

template <int N>
__device__ void block_sum(double (&v)[N], double *sh /* RB*N doubles */) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < N; ++i) sh[i * RB + tid] = v[i];
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int i = 0; i < N; ++i) sh[i * RB + tid] += sh[i * RB + tid + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = sh[i * RB];
  __syncthreads();
}

constexpr int kInnerChunk = 4096;

template <typename R>
__global__ __launch_bounds__(RB) void k_inner_partial(const typename CT<R>::V *__restrict__ base, int64_t dim,
                                                      const int64_t *__restrict__ a_off,
                                                      const int64_t *__restrict__ b_off, int64_t n_chunks,
                                                      double2 *__restrict__ part) {
  __shared__ double sh[2 * RB];
  const int64_t pair = blockIdx.x / n_chunks;
  const int64_t chunk = blockIdx.x - pair * n_chunks;
  const auto *a = base + a_off[pair];
  const auto *b = base + b_off[pair];
  const int64_t lo = chunk * kInnerChunk;
  const int64_t hi = lo + kInnerChunk < dim ? lo + kInnerChunk : dim;
  double v[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += RB) {
    // conj(a) * b
    const auto x = a[i], y = b[i];
    v[0] += double(x.x) * y.x + double(x.y) * y.y;
    v[1] += double(x.x) * y.y - double(x.y) * y.x;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = double2{v[0], v[1]};
}

// <a_i|b_i> partials of a strided batch: a_i = base + 2 i stride, b_i = a_i + stride.
template <typename R>
__global__ __launch_bounds__(RB) void k_pair_partial(const typename CT<R>::V *__restrict__ base, int64_t stride,
                                                     int64_t dim, int64_t n_chunks, double2 *__restrict__ part) {
  __shared__ double sh[2 * RB];
  const int64_t pair = blockIdx.x / n_chunks;
  const int64_t chunk = blockIdx.x - pair * n_chunks;
  const auto *a = base + 2 * pair * stride;
  const auto *b = a + stride;
  const int64_t lo = chunk * kInnerChunk;
  const int64_t hi = lo + kInnerChunk < dim ? lo + kInnerChunk : dim;
  double v[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += RB) {
    const auto x = a[i], y = b[i];
    v[0] += double(x.x) * y.x + double(x.y) * y.y;
    v[1] += double(x.x) * y.y - double(x.y) * y.x;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = double2{v[0], v[1]};
}

__global__ __launch_bounds__(RB) void k_sum_rows(const double2 *__restrict__ part, int64_t n_chunks,
                                                 double2 *__restrict__ out) {
  __shared__ double sh[2 * RB];
  const int64_t row = blockIdx.x;
  double v[2] = {0, 0};
  for (int64_t c = threadIdx.x; c < n_chunks; c += RB) {
    v[0] += part[row * n_chunks + c].x;
    v[1] += part[row * n_chunks + c].y;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) out[row] = double2{v[0], v[1]};
}

// Meyer-Wallach: per chunk of 2^c amplitudes, per qubit (rho00, rho11, rho01).
template <typename R>
__global__ __launch_bounds__(RB) void k_mw_partial(const typename CT<R>::V *__restrict__ states, int n, int c,
                                                   int64_t n_chunks, double *__restrict__ part) {
  using V = typename CT<R>::V;
  __shared__ double sh[4 * RB];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V *tile = reinterpret_cast<V *>(smem_raw);
  const int64_t b = blockIdx.x / n_chunks;
  const int64_t chunk = blockIdx.x - b * n_chunks;
  const int64_t dim = int64_t(1) << n;
  const V *st = states + b * dim;
  const int64_t cbase = chunk << c;
  const int CS = 1 << c;
  for (int j = threadIdx.x; j < CS; j += RB) tile[j] = st[cbase + j];
  __syncthreads();
  double *out = part + (b * n_chunks + chunk) * int64_t(n) * 4;
  // in-chunk qubits: bit position pos < c  <=>  qubit q = n-1-pos
  for (int pos = 0; pos < c; ++pos) {
    double v[4] = {0, 0, 0, 0};
    for (int u = threadIdx.x; u < (CS >> 1); u += RB) {
      const int j0 = ins0(u, pos), j1 = j0 | (1 << pos);
      const V a0 = tile[j0], a1 = tile[j1];
      v[0] += double(a0.x) * a0.x + double(a0.y) * a0.y;
      v[1] += double(a1.x) * a1.x + double(a1.y) * a1.y;
      const double2 r = cmulc(a0, a1);
      v[2] += r.x;
      v[3] += r.y;
    }
    block_sum<4>(v, sh);
    if (threadIdx.x == 0) {
      const int q = n - 1 - pos;
      out[4 * q + 0] = v[0];
      out[4 * q + 1] = v[1];
      out[4 * q + 2] = v[2];
      out[4 * q + 3] = v[3];
    }
  }
  if (c < n) {
    double tot[1] = {0};
    for (int j = threadIdx.x; j < CS; j += RB) tot[0] += double(tile[j].x) * tile[j].x + double(tile[j].y) * tile[j].y;
    block_sum<1>(tot, sh);
    for (int pos = c; pos < n; ++pos) {
      const int bit = int((cbase >> pos) & 1);
      double v[2] = {0, 0};
      if (!bit) {
        const V *pt = st + (cbase | (int64_t(1) << pos));
        for (int j = threadIdx.x; j < CS; j += RB) {
          const double2 r = cmulc(tile[j], pt[j]);
          v[0] += r.x;
          v[1] += r.y;
        }
        block_sum<2>(v, sh);
      }
      if (threadIdx.x == 0) {
        const int q = n - 1 - pos;
        out[4 * q + 0] = bit ? 0.0 : tot[0];
        out[4 * q + 1] = bit ? tot[0] : 0.0;
        out[4 * q + 2] = v[0];
        out[4 * q + 3] = v[1];
      }
    }
  }
}

__global__ __launch_bounds__(RB) void k_mw_final(const double *__restrict__ part, int n, int64_t n_chunks,
                                                 double *__restrict__ rho, double *__restrict__ qout) {
  __shared__ double sh[4 * RB];
  const int64_t b = blockIdx.x;
  double purity = 0;
  for (int q = 0; q < n; ++q) {
    double v[4] = {0, 0, 0, 0};
    for (int64_t ch = threadIdx.x; ch < n_chunks; ch += RB) {
      const double *p = part + ((b * n_chunks + ch) * n + q) * 4;
      v[0] += p[0];
      v[1] += p[1];
      v[2] += p[2];
      v[3] += p[3];
    }
    block_sum<4>(v, sh);
    if (threadIdx.x == 0) {
      if (rho) {
        rho[(b * n + q) * 4 + 0] = v[0];
        rho[(b * n + q) * 4 + 1] = v[1];
        rho[(b * n + q) * 4 + 2] = v[2];
        rho[(b * n + q) * 4 + 3] = v[3];
      }
      purity += v[0] * v[0] + v[1] * v[1] + 2 * (v[2] * v[2] + v[3] * v[3]);
    }
  }
  if (threadIdx.x == 0 && qout) qout[b] = 2 * (1 - purity / n);
}

// Complex Gram  out[i][j] = sum_d conj(a_i[d]) b_j[d]: 16x16 output tile per
// block, K split over blockIdx.z, partials reduced in order afterwards.
constexpr int GT = 16, GK = 16;

// One-row Gram (a GEMV): out[z][q] = sum over k in chunk z of conj(a_k) b_q,k
// for 8 rows q per block; a is read once per block and stays in L2 across
// blocks, each b row once.  Wave shuffles, then LDS across the 4 waves.
constexpr int BV_ROWS = 8;

template <typename R>
__global__ __launch_bounds__(256) void k_bvec_partial(const typename CT<R>::V *__restrict__ A,
                                                      const typename CT<R>::V *__restrict__ B, int64_t ldb,
                                                      int64_t nn, int64_t dim, int64_t kchunk,
                                                      double2 *__restrict__ part) {
  __shared__ double red[4][BV_ROWS][2];
  const int64_t q0 = int64_t(blockIdx.x) * BV_ROWS;
  const int64_t k0 = int64_t(blockIdx.y) * kchunk;
  const int64_t k1 = k0 + kchunk < dim ? k0 + kchunk : dim;
  const int nr = nn - q0 < BV_ROWS ? int(nn - q0) : BV_ROWS;
  double ar[BV_ROWS], ai[BV_ROWS];
#pragma unroll
  for (int r = 0; r < BV_ROWS; ++r) ar[r] = ai[r] = 0.0;
  for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const auto x = A[k];
#pragma unroll
    for (int r = 0; r < BV_ROWS; ++r)
      if (r < nr) {
        const auto y = B[(q0 + r) * ldb + k];
        ar[r] += double(x.x) * y.x + double(x.y) * y.y;
        ai[r] += double(x.x) * y.y - double(x.y) * y.x;
      }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < BV_ROWS; ++r) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ar[r] += __shfl_xor(ar[r], o);
      ai[r] += __shfl_xor(ai[r], o);
    }
    if (lane == 0) {
      red[w][r][0] = ar[r];
      red[w][r][1] = ai[r];
    }
  }
  __syncthreads();
  if (threadIdx.x < nr) {
    const int r = threadIdx.x;
    part[int64_t(blockIdx.y) * nn + q0 + r] =
        double2{red[0][r][0] + red[1][r][0] + red[2][r][0] + red[3][r][0],
                red[0][r][1] + red[1][r][1] + red[2][r][1] + red[3][r][1]};
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_gram_partial(const typename CT<R>::V *__restrict__ A, int64_t lda,
                                                      int64_t m, const typename CT<R>::V *__restrict__ B,
                                                      int64_t ldb, int64_t nn, int64_t dim, int64_t kchunk,
                                                      double2 *__restrict__ part) {
  using V = typename CT<R>::V;
  __shared__ V sa[GT][GK + 1];
  __shared__ V sb[GT][GK + 1];
  const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
  const int64_t i0 = int64_t(blockIdx.x) * GT, j0 = int64_t(blockIdx.y) * GT;
  const int64_t k0 = int64_t(blockIdx.z) * kchunk;
  const int64_t k1 = k0 + kchunk < dim ? k0 + kchunk : dim;
  double accr = 0, acci = 0;
  for (int64_t kk = k0; kk < k1; kk += GK) {
    // load A rows i0..i0+15, cols kk..kk+15 (thread (ty, tx) loads row ty, col tx)
    {
      const int64_t r = i0 + ty, k = kk + tx;
      sa[ty][tx] = (r < m && k < k1) ? A[r * lda + k] : V{0, 0};
      const int64_t r2 = j0 + ty;
      sb[ty][tx] = (r2 < nn && k < k1) ? B[r2 * ldb + k] : V{0, 0};
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      const V x = sa[ty][k], y = sb[tx][k];
      accr += double(x.x) * y.x + double(x.y) * y.y;
      acci += double(x.x) * y.y - double(x.y) * y.x;
    }
    __syncthreads();
  }
  const int64_t i = i0 + ty, j = j0 + tx;
  if (i < m && j < nn) part[(int64_t(blockIdx.z) * m + i) * nn + j] = double2{accr, acci};
}

// Real symmetric rank-K update C = R R^T (R: P x K doubles, row stride ld),
// upper 64 x 64 tiles only, split over K, on the FP64 matrix cores.
//
// v_mfma_f64_16x16x4f64 takes, in lane l, A[l % 16][l / 16] and
// B[l / 16][l % 16] and returns D[4 r + l / 16][l % 16] in element r.  The
// sum over k does not care which k sits in which of the four k slots, so
// lane l (row l % 16 of a 16-row block, k group g = l / 16) loads the pair
// R[row][kk + 2 g + {0, 1}] straight from memory into registers and feeds
// element s of the pair to the product of k slot g in sub-step s; B uses the
// same map, so the products pair up.  No LDS and no barriers: every wave is
// an independent unit of work (one 64 x 64 tile of C over one K chunk, 4 x 4
// blocks of 16 x 16, 128 accumulator registers), and each 8-wide k step's
// loads are issued a step ahead of its products.  Diagonal tiles load one
// operand, skip the 16 x 16 blocks below the diagonal and with the spare
// matrix-core time form the QFIM's b = <J_p|psi> for their 64 rows (psi as
// a second operand with two non-zero columns).
typedef double d4_t __attribute__((ext_vector_type(4)));
// Two waves per SIMD (249 registers each) and a two-step load ring: one
// wave per SIMD with a ring of 4 or 6 steps in the accumulator registers ran
// 5-10% slower, a ring of 3 at two waves spills.
constexpr int SYW = 64, SYK = 8, SYWAVES = 2, SYRING = 2;
// Jacobians above kQfimStageBytes go through the product in column blocks of
// kQfimBlockBytes copied to a compact buffer (below it the copy costs more
// than the shorter row stride gains).
constexpr size_t kQfimStageBytes = size_t(1) << 30, kQfimBlockBytes = size_t(512) << 20;

template <typename TIn>
__device__ __forceinline__ double2 syrk_ld2(const char *p) {
  if constexpr (sizeof(TIn) == 8) {
    return *reinterpret_cast<const double2 *>(p);
  } else {
    const float2 f = *reinterpret_cast<const float2 *>(p);
    return double2{double(f.x), double(f.y)};
  }
}

// Operand addresses are a uniform base (the k step) plus a 32-bit byte
// offset per lane (its row and k group; the host keeps R under 4 GiB per
// launch).  Every chunk is a whole number of ring turns (the host pads K),
// so no step needs masking; the last turn's prefetches (never used) re-read
// the chunk's first step instead of running past it.
template <typename TIn, bool DIAG>
__device__ __forceinline__ void syrk_wave_body(const TIn *R, const uint32_t (&oa)[4], const uint32_t (&ob)[4],
                                               const TIn *psi, int64_t k0, int64_t k1, int g, int col,
                                               d4_t (&acc)[4][4], d4_t (&accb)[4]) {
  double2 A[SYRING][4], B[SYRING][4];
  auto load = [&](double2(&X)[4], double2(&Y)[4], int64_t kk) {
    const int64_t kl = kk < k1 ? kk : k0;
    const char *base = reinterpret_cast<const char *>(R + kl);
#pragma unroll
    for (int a = 0; a < 4; ++a) X[a] = syrk_ld2<TIn>(base + oa[a]);
    if constexpr (!DIAG) {
#pragma unroll
      for (int b = 0; b < 4; ++b) Y[b] = syrk_ld2<TIn>(base + ob[b]);
    } else {
      // the b vector's operand: column 0 pairs (Re, Im) of a row with
      // (Re psi, Im psi), column 1 with (Im psi, -Re psi), the rest zero
      const double2 v = syrk_ld2<TIn>(reinterpret_cast<const char *>(psi + kl + 2 * g));
      Y[0] = col == 0 ? v : col == 1 ? double2{v.y, -v.x} : double2{0.0, 0.0};
    }
  };
  // the two sub-steps in turn over all blocks: products into one
  // accumulator are 16 (14) instructions apart
  auto mma = [&](const double2(&X)[4], const double2(&Y)[4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (DIAG && a > b) continue;
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[a].x, DIAG ? X[b].x : Y[b].x, acc[a][b], 0, 0, 0);
      }
    if constexpr (DIAG) {
#pragma unroll
      for (int a = 0; a < 4; ++a) accb[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[a].x, Y[0].x, accb[a], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (DIAG && a > b) continue;
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[a].y, DIAG ? X[b].y : Y[b].y, acc[a][b], 0, 0, 0);
      }
    if constexpr (DIAG) {
#pragma unroll
      for (int a = 0; a < 4; ++a) accb[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[a].y, Y[0].y, accb[a], 0, 0, 0);
    }
  };
#pragma unroll
  for (int r = 0; r < SYRING; ++r) load(A[r], B[r], k0 + r * SYK);
  for (int64_t kk = k0; kk < k1; kk += SYRING * SYK) {
#pragma unroll
    for (int r = 0; r < SYRING; ++r) {
      mma(A[r], B[r]);
      load(A[r], B[r], kk + (r + SYRING) * SYK);
    }
  }
}

// Work of one wave: the XCD-aware map from (block, wave).  The dispatcher
// deals consecutive workgroups round-robin to the 8 XCDs, each with its own
// L2.  Items (tile, K chunk) are numbered chunk-major and XCD x runs the
// contiguous range [x per_xcd, (x + 1) per_xcd): the tiles of a chunk run
// side by side on one XCD and walk its k range together, so a chunk's rows
// are fetched from HBM about once and re-read from that XCD's L2.
template <typename TIn>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SYWAVES))) void k_syrk_mfma(
    const TIn *__restrict__ R, int64_t ld, int64_t P, int64_t K, int64_t kchunk, int64_t n_items, int64_t per_xcd,
    const int2 *__restrict__ tiles, int64_t T, const TIn *__restrict__ psi, int accum, double *__restrict__ part,
    double2 *__restrict__ bpart) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r16 = l & 15;
  const int64_t slot = int64_t(blockIdx.x >> 3) * 4 + w;
  if (slot >= per_xcd) return;
  const int64_t m = int64_t(blockIdx.x & 7) * per_xcd + slot;
  if (m >= n_items) return;
  const int64_t chunk = m / T;
  const int2 tile = tiles[m - chunk * T];
  const int64_t i0 = int64_t(tile.x) * SYW, j0 = int64_t(tile.y) * SYW;
  const int64_t k0 = chunk * kchunk;
  if (k0 >= K) return;  // a narrower last column block has fewer chunks
  const int64_t k1 = k0 + kchunk < K ? k0 + kchunk : K;
  // rows past P read row P - 1: they only feed entries that are not stored
  uint32_t oa[4], ob[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t ra = i0 + 16 * a + r16, rb = j0 + 16 * a + r16;
    oa[a] = uint32_t(((ra < P ? ra : P - 1) * ld + 2 * g) * int64_t(sizeof(TIn)));
    ob[a] = uint32_t(((rb < P ? rb : P - 1) * ld + 2 * g) * int64_t(sizeof(TIn)));
  }
  d4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4_t{0.0, 0.0, 0.0, 0.0};
  d4_t accb[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) accb[a] = d4_t{0.0, 0.0, 0.0, 0.0};
  const bool diag = tile.x == tile.y;
  if (diag)
    syrk_wave_body<TIn, true>(R, oa, ob, psi, k0, k1, g, r16, acc, accb);
  else
    syrk_wave_body<TIn, false>(R, oa, ob, psi, k0, k1, g, r16, acc, accb);
  double *out = part + chunk * P * P;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (diag && a > b) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + 16 * a + 4 * r + g, j = j0 + 16 * b + r16;
        if (i < P && j < P) out[i * P + j] = accum ? out[i * P + j] + acc[a][b][r] : acc[a][b][r];
      }
    }
  if (diag && r16 < 2) {
    // <row|psi> over the chunk: column 0 (lanes 0, 16, 32, 48) the real
    // part, column 1 the imaginary part
    double *bo = reinterpret_cast<double *>(bpart + chunk * P);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + 16 * a + 4 * r + g;
        if (i < P) bo[2 * i + r16] = accum ? bo[2 * i + r16] + accb[a][r] : accb[a][r];
      }
  }
}

// Complex Gram G = A^dag B (A: m x K, B: nn x K complex, rows ld complex
// apart) on the FP64 matrix cores, split over K.  One wave = one 64 x 32
// tile of G (4 x 2 blocks of 16 x 16, real and imaginary accumulators: 128
// registers) over one K chunk.  Lane l (row l % 16, k slot g = l / 16) loads
// the complex amplitude kk + g of its row: Re G += Re a Re b + Im a Im b, Im G
// += Re a Im b - Im a Re b, four products per block and step of 4
// amplitudes.  Items, the XCD map and the load ring are those of the real
// product above.
constexpr int CGM = 64, CGN = 32, CGK = 4;

template <typename TIn>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SYWAVES))) void k_cgram_mfma(
    const TIn *__restrict__ A, int64_t lda, int64_t m, const TIn *__restrict__ B, int64_t ldb, int64_t nn, int64_t K,
    int64_t kchunk, int64_t n_items, int64_t per_xcd, const int2 *__restrict__ tiles, int64_t T, int accum,
    double2 *__restrict__ part) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r16 = l & 15;
  const int64_t slot = int64_t(blockIdx.x >> 3) * 4 + w;
  if (slot >= per_xcd) return;
  const int64_t mi = int64_t(blockIdx.x & 7) * per_xcd + slot;
  if (mi >= n_items) return;
  const int64_t chunk = mi / T;
  const int2 tile = tiles[mi - chunk * T];
  const int64_t i0 = int64_t(tile.x) * CGM, j0 = int64_t(tile.y) * CGN;
  const int64_t k0 = chunk * kchunk;
  if (k0 >= K) return;
  const int64_t k1 = k0 + kchunk < K ? k0 + kchunk : K;
  // byte offsets of the lane's rows (rows past the end re-read the last
  // row: they feed only entries that are not stored)
  constexpr int64_t cb = 2 * int64_t(sizeof(TIn));
  uint32_t oa[4], ob[2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t r = i0 + 16 * a + r16;
    oa[a] = uint32_t(((r < m ? r : m - 1) * lda + g) * cb);
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int64_t r = j0 + 16 * b + r16;
    ob[b] = uint32_t(((r < nn ? r : nn - 1) * ldb + g) * cb);
  }
  d4_t re[4][2], im[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) re[a][b] = im[a][b] = d4_t{0.0, 0.0, 0.0, 0.0};
  double2 X[SYRING][4], Y[SYRING][2];
  auto load = [&](double2(&x)[4], double2(&y)[2], int64_t kk) {
    const int64_t kl = kk < k1 ? kk : k0;
    const char *ba = reinterpret_cast<const char *>(A + 2 * kl), *bb = reinterpret_cast<const char *>(B + 2 * kl);
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = syrk_ld2<TIn>(ba + oa[a]);
#pragma unroll
    for (int b = 0; b < 2; ++b) y[b] = syrk_ld2<TIn>(bb + ob[b]);
  };
  auto mma = [&](const double2(&x)[4], const double2(&y)[2]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[a].x, y[b].x, re[a][b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[a].x, y[b].y, im[a][b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[a].y, y[b].y, re[a][b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-x[a].y, y[b].x, im[a][b], 0, 0, 0);
  };
#pragma unroll
  for (int r = 0; r < SYRING; ++r) load(X[r], Y[r], k0 + r * CGK);
  for (int64_t kk = k0; kk < k1; kk += SYRING * CGK) {
#pragma unroll
    for (int r = 0; r < SYRING; ++r) {
      mma(X[r], Y[r]);
      load(X[r], Y[r], kk + (r + SYRING) * CGK);
    }
  }
  double2 *out = part + chunk * m * nn;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + 16 * a + 4 * r + g, j = j0 + 16 * b + r16;
        if (i < m && j < nn) {
          double2 v{re[a][b][r], im[a][b][r]};
          if (accum) {
            const double2 o = out[i * nn + j];
            v.x += o.x;
            v.y += o.y;
          }
          out[i * nn + j] = v;
        }
      }
}

// out[i][j] = sum over the split of part[z][i][j]; for A = B only the tiles
// meeting the upper triangle were computed, out[i][j] = conj(out[j][i])
// below it and the diagonal is real.
__global__ void k_cgram_reduce(const double2 *__restrict__ part, int64_t m, int64_t nn, int64_t ks, int sym,
                               double2 *__restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * nn) return;
  const int64_t i = e / nn, j = e - i * nn;
  const bool lower = sym && i > j;
  const int64_t src = lower ? j * nn + i : e;
  double2 a{0, 0};
  for (int64_t z = 0; z < ks; ++z) {
    a.x += part[z * m * nn + src].x;
    a.y += part[z * m * nn + src].y;
  }
  // <a_i|a_i> is real: its products cancel term by term (as in vdot)
  out[e] = lower ? double2{a.x, -a.y} : sym && i == j ? double2{a.x, 0.0} : a;
}

// dst[r][0, w) = src[r][0, w) for P rows (16-byte units, row strides lds and
// ldd): a column block of the Jacobian into a compact buffer.
__global__ void k_copy_cols(const uint4 *__restrict__ src, int64_t lds, uint4 *__restrict__ dst, int64_t ldd,
                            int64_t w, int64_t units) {
  const int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const int64_t r = u / w, c = u - r * w;
  dst[r * ldd + c] = src[r * lds + c];
}

// dst[r][c] = src[r][c] for c < ld_src, zero up to ld_dst (units = P ld_dst
// elements): a short Jacobian padded to whole k steps.
template <typename T>
__global__ void k_pad_cols(const T *__restrict__ src, int64_t lds, T *__restrict__ dst, int64_t ldd, int64_t units) {
  const int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const int64_t r = u / ldd, c = u - r * ldd;
  dst[u] = c < lds ? src[r * lds + c] : T(0);
}

// Column-major C[i + j P] = sum over the split of part[z][i][j], for the
// upper triangle i <= j (every such entry lies in a computed block).
__global__ void k_syrk_reduce(const double *__restrict__ part, int64_t P, int64_t ks, double *__restrict__ C) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= P * P) return;
  const int64_t i = e / P, j = e - i * P;
  if (i > j) return;
  double a = 0;
  for (int64_t z = 0; z < ks; ++z) a += part[z * P * P + e];
  C[i + j * P] = a;
}

__global__ void k_gram_reduce(const double2 *__restrict__ part, int64_t mn, int64_t ks, double2 *__restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= mn) return;
  double2 a{0, 0};
  for (int64_t z = 0; z < ks; ++z) {
    a.x += part[z * mn + e].x;
    a.y += part[z * mn + e].y;
  }
  out[e] = a;
}

// F from the upper triangle of the real Gram C = Re(J J^dag) (column-major,
// C[p + q P] for p <= q) and b_p = <psi|J_p>.
__global__ void k_qfim_combine_sym(const double *__restrict__ C, const double2 *__restrict__ bvec, int64_t P,
                                   double *__restrict__ F) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= P * P) return;
  const int64_t p = e / P, q = e - p * P;
  const int64_t lo = p < q ? p : q, hi = p < q ? q : p;
  const double2 bp = bvec[p], bq = bvec[q];
  F[e] = 4 * (C[lo + hi * P] - (bp.x * bq.x + bp.y * bq.y));
}

struct PauliTerm {
  uint64_t x, z;
  double cre, cim;  // includes i^{#Y}
};

template <typename R>
__global__ __launch_bounds__(RB) void k_expect_partial(const typename CT<R>::V *__restrict__ states, int n,
                                                       int64_t n_chunks, const PauliTerm *__restrict__ terms,
                                                       int nt, double2 *__restrict__ part) {
  __shared__ double sh[2 * RB];
  const int64_t b = blockIdx.x / n_chunks;
  const int64_t chunk = blockIdx.x - b * n_chunks;
  const int64_t dim = int64_t(1) << n;
  const auto *st = states + b * dim;
  const int64_t lo = chunk * kInnerChunk;
  const int64_t hi = lo + kInnerChunk < dim ? lo + kInnerChunk : dim;
  double v[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += RB) {
    const auto a = st[i];
    for (int t = 0; t < nt; ++t) {
      const uint64_t ii = uint64_t(i) ^ terms[t].x;
      const auto y = st[ii];
      // conj(a) * y
      double re = double(a.x) * y.x + double(a.y) * y.y;
      double im = double(a.x) * y.y - double(a.y) * y.x;
      if (__popcll(ii & terms[t].z) & 1) {
        re = -re;
        im = -im;
      }
      v[0] += terms[t].cre * re - terms[t].cim * im;
      v[1] += terms[t].cre * im + terms[t].cim * re;
    }
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = double2{v[0], v[1]};
}

// out = sum_t c_t P_t x (Pauli strings as index masks; c_t includes i^{#Y}):
// (P x)[j] = (-1)^{popc((j ^ xm) & zm)} x[j ^ xm].  The Hamiltonian applied
// to the state at the start of the adjoint gradient's backward pass.
template <typename V>
__global__ __launch_bounds__(256) void k_apply_pauli(const V *__restrict__ x, int64_t dim,
                                                     const PauliTerm *__restrict__ terms, int nt,
                                                     V *__restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j >= dim) return;
  double re = 0.0, im = 0.0;
  for (int t = 0; t < nt; ++t) {
    const uint64_t jj = uint64_t(j) ^ terms[t].x;
    const V y = x[jj];
    double yr = double(y.x), yi = double(y.y);
    if (__popcll(jj & terms[t].z) & 1) {
      yr = -yr;
      yi = -yi;
    }
    re += terms[t].cre * yr - terms[t].cim * yi;
    im += terms[t].cre * yi + terms[t].cim * yr;
  }
  using R = decltype(V{}.x);
  out[j] = V{R(re), R(im)};
}

// <bra| c_t P_t |ket> for every term t: block (chunk, t) sums one chunk of
// one term; k_sum_rows reduces the chunks (fixed order).  The gradient
// terms of the adjoint backward pass, 2 Re <lambda| D_p |psi> per parameter.
template <typename V>
__global__ __launch_bounds__(RB) void k_transition_partial(const V *__restrict__ bra, const V *__restrict__ ket,
                                                           int64_t dim, int64_t n_chunks,
                                                           const PauliTerm *__restrict__ terms,
                                                           double2 *__restrict__ part) {
  __shared__ double sh[2 * RB];
  const int64_t t = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const PauliTerm pt = terms[t];
  const int64_t lo = chunk * kInnerChunk;
  const int64_t hi = lo + kInnerChunk < dim ? lo + kInnerChunk : dim;
  double v[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += RB) {
    const V a = bra[i];
    const uint64_t ii = uint64_t(i) ^ pt.x;
    const V y = ket[ii];
    double re = double(a.x) * y.x + double(a.y) * y.y;  // conj(a) * y
    double im = double(a.x) * y.y - double(a.y) * y.x;
    if (__popcll(ii & pt.z) & 1) {
      re = -re;
      im = -im;
    }
    v[0] += pt.cre * re - pt.cim * im;
    v[1] += pt.cre * im + pt.cim * re;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) part[t * n_chunks + chunk] = double2{v[0], v[1]};
}

__global__ void k_histogram(const double *__restrict__ v, int64_t n, const double *__restrict__ edges, int nb,
                            unsigned long long *__restrict__ counts) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = v[i];
  const double lo = edges[0], hi = edges[nb];
  if (!(x >= lo && x <= hi)) return;
  // numpy.histogram, uniform-bin path: index from the scaled offset, then
  // corrected against the explicit edges.
  const double f = ((x - lo) / (hi - lo)) * nb;
  long long idx = (long long)f;
  if (idx == nb) idx -= 1;
  if (x < edges[idx]) idx -= 1;
  if (x >= edges[idx + 1] && idx != nb - 1) idx += 1;
  atomicAdd(counts + idx, 1ull);
}

// Dense operator on a batch of states: out_b = U x_b (U: dim x dim complex128,
// row-major).  One wave per output row: lanes stride the row (coalesced) and
// the wave sums with cross-lane reductions; the row is re-read from L2 for
// every state of the batch.  Serves gates without a local form (ARBGATE).
template <typename V>
__global__ __launch_bounds__(256) void k_apply_dense(const double2 *__restrict__ U, int64_t dim,
                                                     const V *__restrict__ x, int64_t stride, int64_t batch,
                                                     V *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= dim) return;
  const double2 *u = U + row * dim;
  for (int64_t b = 0; b < batch; ++b) {
    const V *xb = x + b * stride;
    double re = 0.0, im = 0.0;
    for (int64_t k = lane; k < dim; k += 64) {
      const double2 a = u[k];
      const V v = xb[k];
      re += a.x * double(v.x) - a.y * double(v.y);
      im += a.x * double(v.y) + a.y * double(v.x);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_xor(re, off);
      im += __shfl_xor(im, off);
    }
    if (lane == 0) {
      using R = decltype(V{}.x);
      out[b * stride + row] = V{R(re), R(im)};
    }
  }
}

__global__ void k_abs2(const double2 *__restrict__ z, int64_t n, double *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = z[i].x * z[i].x + z[i].y * z[i].y;
}

template <typename V>
__global__ void k_fill_zero(V *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = V{0, 0};
}

// Sums of split-generator parts: row acc[y].dst += every source row
// srcs[first .. first + count) (in that order; one block row per target).
struct AccRow {
  void *dst;
  int32_t first, count;
};

template <typename V>
__global__ __launch_bounds__(256) void k_add_rows(const AccRow *__restrict__ acc, const void *const *__restrict__ srcs,
                                                  int64_t dim) {
  const AccRow a = acc[blockIdx.y];
  V *dst = static_cast<V *>(a.dst);
  for (int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x; j < dim; j += int64_t(gridDim.x) * 256) {
    V v = dst[j];
    for (int c = 0; c < a.count; ++c) {
      const V w = static_cast<const V *>(srcs[a.first + c])[j];
      v.x += w.x;
      v.y += w.y;
    }
    dst[j] = v;
  }
}

// Instance tables of pqs_run_batch_noalloc live in this file-scope device
// array: one per GPU, shared by every call on that GPU (calls that overlap
// on one GPU race for it).  Host code never passes its name to a kernel: it
// uses the per-device address from hipGetSymbolAddress (pqs_ctx_create).
constexpr int64_t kNoallocMaxBatch = 16384;
static __device__ __align__(256) unsigned char scratch[kNoallocMaxBatch * sizeof(InstDev)];

// Instance table of a strided batch, written on the device (no host copy).
__global__ __launch_bounds__(256) void k_fill_batch_table(InstDev *__restrict__ tab, char *states, int64_t state_bytes,
                                                          int64_t batch, const double *params, int64_t param_stride,
                                                          int32_t init, const void *src, char *alt,
                                                          int32_t pingpong) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= batch) return;
  InstDev d;
  d.state = states + i * state_bytes;
  d.src = src;
  d.params = params ? params + i * param_stride : nullptr;
  d.alt = pingpong ? alt + i * state_bytes : nullptr;
  d.tangent = -1;
  d.init = init;
  d.start_sweep = 0;
  d.pingpong = pingpong;
  tab[i] = d;
}

__global__ void k_zero_u64(unsigned long long *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Stabiliser Renyi entropy sum (reference measure.py:318-349): for every
// X-part j, f(m) = conj(c_m) c_{m^j}; a Walsh-Hadamard transform over m gives
// F(i) = sum_m (-1)^{i.m} f(m) = <psi|X^j Z^i|psi> up to a phase; the block
// returns sum_i |2^{-n/2} F(i)|^{2 alpha}.  One block per (instance, j).
template <typename R>
__global__ __launch_bounds__(RB) void k_magic_partial(const typename CT<R>::V *__restrict__ states, int n,
                                                      double alpha, double *__restrict__ part) {
  __shared__ double sh[RB];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2 *f = reinterpret_cast<double2 *>(smem_raw);
  const int64_t D = int64_t(1) << n;
  const int64_t b = blockIdx.x / D;
  const int64_t j = blockIdx.x - b * D;
  const auto *c = states + b * D;
  for (int64_t m = threadIdx.x; m < D; m += RB) {
    const auto x = c[m], y = c[m ^ j];
    f[m] = double2{double(x.x) * y.x + double(x.y) * y.y, double(x.x) * y.y - double(x.y) * y.x};
  }
  for (int bit = 0; bit < n; ++bit) {
    __syncthreads();
    for (int64_t u = threadIdx.x; u < D / 2; u += RB) {
      const int64_t i0 = ((u >> bit) << (bit + 1)) | (u & ((int64_t(1) << bit) - 1));
      const int64_t i1 = i0 | (int64_t(1) << bit);
      const double2 a = f[i0], d = f[i1];
      f[i0] = double2{a.x + d.x, a.y + d.y};
      f[i1] = double2{a.x - d.x, a.y - d.y};
    }
  }
  __syncthreads();
  const double scale = 1.0 / double(D);  // (2^{-n/2})^2 on |F|^2
  double v[1] = {0};
  for (int64_t i = threadIdx.x; i < D; i += RB) {
    const double a2 = (f[i].x * f[i].x + f[i].y * f[i].y) * scale;
    v[0] += alpha == 2.0 ? a2 * a2 : (alpha == 0.5 ? sqrt(a2) : pow(a2, alpha));
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(RB) void k_sum_rows_real(const double *__restrict__ part, int64_t n_chunks,
                                                      double *__restrict__ out) {
  __shared__ double sh[RB];
  const int64_t row = blockIdx.x;
  double v[1] = {0};
  for (int64_t c = threadIdx.x; c < n_chunks; c += RB) v[0] += part[row * n_chunks + c];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) out[row] = v[0];
}

inline unsigned grid_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

// ===================================================================== ctx
struct pqs_ctx {
  int device = 0;
  int n_cu = 0;
  int max_lds = 0;
  char arch[64] = {0};
  void *ws = nullptr;
  size_t ws_bytes = 0;
  hipStream_t ws_stream = nullptr;  // stream of the last call that used ws
  hipEvent_t ws_event = nullptr;
  void *scratch = nullptr;  // this device's address of the file-scope scratch array
  // Pinned staging for host-to-device uploads: a ring of buffers, each
  // reused only once its own last copy has run.  A busy buffer is never
  // waited for while the ring may still grow (up to kPinMax buffers), so an
  // upload does not hold the host thread behind the stream's earlier
  // kernels.  The copy itself is a kernel on the caller's stream that reads
  // the pinned buffer: a DMA-engine copy would queue behind other streams'
  // copies, which wait for those streams' kernels.
  struct PinSlot {
    void *host = nullptr;
    size_t bytes = 0;
    hipEvent_t event = nullptr;
    bool pending = false;
  };
  static constexpr int kPinMax = 32;
  std::vector<PinSlot> pins;
  int pin_next = 0;
  int64_t max_blocks = 0;  // cap on the blocks of one sweep launch (0: none)
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int d) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != d) (void)hipSetDevice(d);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// The workspace grows on demand and is allocated and released in stream
// order (hipMallocAsync / hipFreeAsync), so growing it never waits for the
// device.  Work on the stream that used it last may still read the old
// buffer: when that stream is not `st`, `st` first waits for it.
int ws_reserve(pqs_ctx *ctx, size_t bytes, hipStream_t st) {
  if (bytes <= ctx->ws_bytes) {
    ctx->ws_stream = st;
    return PQS_OK;
  }
  size_t nb = std::max<size_t>(bytes, size_t(64) << 20);
  nb = (nb + 4095) & ~size_t(4095);
  if (ctx->ws) {
    if (ctx->ws_stream != st) {
      HIPCHECK(hipEventRecord(ctx->ws_event, ctx->ws_stream));
      HIPCHECK(hipStreamWaitEvent(st, ctx->ws_event, 0));
    }
    HIPCHECK(hipFreeAsync(ctx->ws, st));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  hipError_t e = hipMallocAsync(&ctx->ws, nb, st);
  if (e != hipSuccess) {
    ctx->ws = nullptr;
    (void)hipGetLastError();
    return fail(PQS_ERR_NOMEM, "workspace allocation failed");
  }
  ctx->ws_bytes = nb;
  ctx->ws_stream = st;
  return PQS_OK;
}

__global__ void k_upload16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_upload1(const unsigned char *__restrict__ src, unsigned char *__restrict__ dst, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// true once the event's work has run; false while it is pending (the
// not-ready status is cleared so that a later launch check does not see it)
bool event_done(hipEvent_t ev) {
  const hipError_t e = hipEventQuery(ev);
  if (e == hipSuccess) return true;
  (void)hipGetLastError();
  return false;
}

// Copy host bytes to device memory in stream order: into a pinned staging
// buffer now, then by a copy kernel on `st`, which reads the pinned buffer.
int upload(pqs_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return PQS_OK;
  auto &pins = ctx->pins;
  const int n_slots = int(pins.size());
  int k = -1;
  for (int i = 0; i < n_slots && k < 0; ++i) {
    const int j = (ctx->pin_next + i) % n_slots;
    pqs_ctx::PinSlot &sl = pins[size_t(j)];
    if (sl.pending && event_done(sl.event)) sl.pending = false;
    if (!sl.pending && sl.bytes >= bytes) k = j;
  }
  if (k < 0 && n_slots < pqs_ctx::kPinMax) {
    pqs_ctx::PinSlot sl;
    HIPCHECK(hipEventCreateWithFlags(&sl.event, hipEventDisableTiming));
    const size_t nb = std::max<size_t>((bytes + 4095) & ~size_t(4095), size_t(1) << 20);
    const hipError_t e = hipHostMalloc(&sl.host, nb, hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipEventDestroy(sl.event);
      return hip_fail(e, "hipHostMalloc (upload staging)");
    }
    sl.bytes = nb;
    pins.push_back(sl);
    k = n_slots;
  }
  if (k < 0) {  // every buffer busy or too small and the ring is full: wait for the next one
    k = ctx->pin_next % n_slots;
    pqs_ctx::PinSlot &sl = pins[size_t(k)];
    if (sl.pending) HIPCHECK(hipEventSynchronize(sl.event));
    sl.pending = false;
    if (sl.bytes < bytes) {
      HIPCHECK(hipHostFree(sl.host));
      sl.host = nullptr;
      sl.bytes = 0;
      const size_t nb = (bytes + 4095) & ~size_t(4095);
      HIPCHECK(hipHostMalloc(&sl.host, nb, hipHostMallocDefault));
      sl.bytes = nb;
    }
  }
  ctx->pin_next = (k + 1) % int(pins.size());
  pqs_ctx::PinSlot &sl = pins[size_t(k)];
  std::memcpy(sl.host, src, bytes);
  const bool wide = (reinterpret_cast<uintptr_t>(dst) & 15) == 0 && (bytes & 15) == 0;
  const int64_t n = wide ? int64_t(bytes >> 4) : int64_t(bytes);
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
  if (wide)
    hipLaunchKernelGGL(k_upload16, dim3(unsigned(blocks)), dim3(256), 0, st, (const uint4 *)sl.host, (uint4 *)dst, n);
  else
    hipLaunchKernelGGL(k_upload1, dim3(unsigned(blocks)), dim3(256), 0, st, (const unsigned char *)sl.host,
                       (unsigned char *)dst, n);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(sl.event, st));
  sl.pending = true;
  return PQS_OK;
}

// A launch on `st` of a plan whose arrays may still be in flight on the
// stream of its first launch: `st` waits for that copy.
int plan_wait(pqs_plan *p, hipStream_t st) {
  if (p->up_pending) {
    if (event_done((hipEvent_t)p->up_event))
      p->up_pending = false;
    else if ((hipStream_t)p->up_stream != st)
      HIPCHECK(hipStreamWaitEvent(st, (hipEvent_t)p->up_event, 0));
  }
  return PQS_OK;
}

// The plan's arrays on the context's device, copied at its first launch in
// that launch's stream order (no host wait).
int plan_upload(pqs_ctx *ctx, pqs_plan *p, hipStream_t st) {
  if (p->dev_device >= 0 && p->dev_device != ctx->device)
    return fail(PQS_ERR_ARG, "plan already bound to another device");
  if (p->d_ops) return plan_wait(p, st);
  if (!p->up_event) {
    hipEvent_t ev;
    HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    p->up_event = (void *)ev;
  }
  const size_t ob = std::max<size_t>(1, p->dev_ops.size()) * sizeof(DevOp);
  const size_t tb = std::max<size_t>(1, p->dev_terms.size()) * sizeof(DevTerm);
  const size_t pb = std::max<size_t>(1, p->dev_phases.size()) * sizeof(pqs::DevPhase);
  const size_t sb = std::max<size_t>(1, p->dev_sweeps.size()) * sizeof(DevSweep);
  void *d[4] = {nullptr, nullptr, nullptr, nullptr};
  const size_t nb[4] = {ob, tb, pb, sb};
  for (int i = 0; i < 4; ++i) {
    const hipError_t e = hipMalloc(&d[i], nb[i]);
    if (e != hipSuccess) {
      for (int j = 0; j < i; ++j) (void)hipFree(d[j]);
      return hip_fail(e, "hipMalloc (plan arrays)");
    }
  }
  p->d_ops = d[0];
  p->d_terms = d[1];
  p->d_phases = d[2];
  p->d_sweeps = d[3];
  p->dev_device = ctx->device;
  int rc = upload(ctx, p->d_ops, p->dev_ops.data(), p->dev_ops.size() * sizeof(DevOp), st);
  if (!rc) rc = upload(ctx, p->d_terms, p->dev_terms.data(), p->dev_terms.size() * sizeof(DevTerm), st);
  if (!rc) rc = upload(ctx, p->d_phases, p->dev_phases.data(), p->dev_phases.size() * sizeof(pqs::DevPhase), st);
  if (!rc) rc = upload(ctx, p->d_sweeps, p->dev_sweeps.data(), p->dev_sweeps.size() * sizeof(DevSweep), st);
  if (rc) return rc;
  HIPCHECK(hipEventRecord((hipEvent_t)p->up_event, st));
  p->up_stream = (void *)st;
  p->up_pending = true;
  return PQS_OK;
}


// Compiled code objects, cached per source text in memory and, when
// PQS_JIT_CACHE (default $TMPDIR/pqs_jit_cache) is writable, on disk.
int jit_code_cached(const std::string &src, std::vector<char> &code, std::string &log) {
  static std::mutex mu;
  static std::unordered_map<std::string, std::vector<char>> mem;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = mem.find(src);
    if (it != mem.end()) {
      code = it->second;
      return PQS_OK;
    }
  }
  // FNV-1a of the source and of what else decides the code object (compile
  // options, hiprtc version)
  static const std::string sig = pqs::jit_signature();
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : src) h = (h ^ c) * 1099511628211ull;
  for (unsigned char c : sig) h = (h ^ c) * 1099511628211ull;
  const char *env = std::getenv("PQS_JIT_CACHE");
  const char *tmp = std::getenv("TMPDIR");
  std::string dir = env ? env : (std::string(tmp ? tmp : "/tmp") + "/pqs_jit_cache");
  char fname[64];
  std::snprintf(fname, sizeof fname, "/%016llx_%zu.co", (unsigned long long)h, src.size());
  const std::string path = dir + fname;
  if (FILE *f = std::fopen(path.c_str(), "rb")) {
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz > 0) {
      code.resize(size_t(sz));
      if (std::fread(code.data(), 1, code.size(), f) != code.size()) code.clear();
    }
    std::fclose(f);
    if (!code.empty()) {
      std::lock_guard<std::mutex> lk(mu);
      mem[src] = code;
      return PQS_OK;
    }
  }
  int rc = pqs::jit_compile(src, code, log);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(mu);
    mem[src] = code;
  }
  ::mkdir(dir.c_str(), 0755);
  const std::string tmpf = path + ".tmp" + std::to_string((unsigned long long)(uintptr_t)&code);
  if (FILE *f = std::fopen(tmpf.c_str(), "wb")) {
    const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
    std::fclose(f);
    if (ok) std::rename(tmpf.c_str(), path.c_str());
    else std::remove(tmpf.c_str());
  }
  return PQS_OK;
}

template <typename R, int BLOCK>
int launch_sweep_t(pqs_ctx *ctx, pqs_plan *p, int s, const InstDev *d_insts, int64_t n_active,
                   const void *init_tab, hipStream_t st) {
  static bool attr_set[64] = {false};
  const DevSweep &sw = p->dev_sweeps[s];
  size_t lds = (size_t(1) << sw.k) * sizeof(typename CT<R>::V);
  if (int(lds) > ctx->max_lds) return fail(PQS_ERR_UNSUPPORTED, "tile does not fit in LDS");
  auto kern = k_sweep<R, BLOCK>;
  if (lds > 65536 && ctx->device < 64 && !attr_set[ctx->device]) {
    HIPCHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    attr_set[ctx->device] = true;
  }
  int64_t grid = sw.n_tiles * n_active;
  if (grid <= 0) return PQS_OK;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "grid too large: split the batch");
  hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(BLOCK), lds, st, sw, (const DevOp *)p->d_ops,
                     (const DevTerm *)p->d_terms, d_insts, n_active,
                     (const typename CT<R>::V *)init_tab, p->n, s, int(p->dev_sweeps.size()));
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

template <typename R>
int launch_sweep_reg(pqs_ctx *ctx, pqs_plan *p, int s, const InstDev *d_insts, int64_t n_active,
                     const void *init_tab, hipStream_t st) {
  constexpr int MAXT = sizeof(R) == 8 ? 512 : 1024;
  static bool attr_set[64] = {false};
  const DevSweep &sw = p->dev_sweeps[s];
  const int S = 1 << sw.k;
  const int nthr = S >> pqs::kRegBits;
  if (nthr > MAXT || nthr < 64) return fail(PQS_ERR_UNSUPPORTED, "tile size outside the register executor's range");
  size_t lds = size_t(S) * sizeof(typename CT<R>::V) + size_t(sw.op_end - sw.op_begin) * sizeof(ROp);
  if (lds > size_t(ctx->max_lds)) return fail(PQS_ERR_UNSUPPORTED, "tile and coefficients do not fit in LDS");
  auto kern = sw.n_coef ? k_sweep_reg<R, MAXT, true> : k_sweep_reg<R, MAXT, false>;
  if (lds > 65536 && ctx->device < 64 && !attr_set[ctx->device]) {
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg<R, MAXT, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg<R, MAXT, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    attr_set[ctx->device] = true;
  }
  int64_t grid = sw.n_tiles * n_active;
  if (grid <= 0) return PQS_OK;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "grid too large: split the batch");
  hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(nthr), lds, st, sw, (const DevOp *)p->d_ops,
                     (const DevTerm *)p->d_terms, (const pqs::DevPhase *)p->d_phases, d_insts, n_active,
                     (const typename CT<R>::V *)init_tab, p->n, s, int(p->dev_sweeps.size()));
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int launch_sweep_one(pqs_ctx *ctx, pqs_plan *p, int s, const InstDev *d_insts, int64_t n_active,
                     const void *init_tab, hipStream_t st) {
  if (p->jit_module && p->dev_device == ctx->device && s < int(p->jit_fn.size()) && p->jit_fn[s]) {
    const DevSweep &sw = p->dev_sweeps[s];
    const int nthr = (1 << sw.k) >> pqs::kRegBits;
    int64_t grid = sw.n_tiles * n_active;
    if (grid <= 0) return PQS_OK;
    if (p->flags & PQS_PLAN_JIT_PERSISTENT) {
      // persistent kernels: as many blocks as can be resident at once
      int per_cu = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (hipFunction_t)p->jit_fn[s], nthr, 0) !=
              hipSuccess ||
          per_cu < 1)
        per_cu = 1;
      grid = std::min<int64_t>(grid, int64_t(ctx->n_cu) * per_cu);
    }
    if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "grid too large: split the batch");
    long long na = (long long)n_active;
    void *args[] = {(void *)&d_insts, (void *)&na, (void *)&init_tab};
    HIPCHECK(hipModuleLaunchKernel((hipFunction_t)p->jit_fn[s], unsigned(grid), 1, 1, unsigned(nthr), 1, 1, 0, st,
                                   args, nullptr));
    return PQS_OK;
  }
  if (p->reg_exec) {
    if (p->dtype == PQS_C128) return launch_sweep_reg<double>(ctx, p, s, d_insts, n_active, init_tab, st);
    return launch_sweep_reg<float>(ctx, p, s, d_insts, n_active, init_tab, st);
  }
  if (p->dtype == PQS_C128) {
    switch (p->block) {
      case 256: return launch_sweep_t<double, 256>(ctx, p, s, d_insts, n_active, init_tab, st);
      case 512: return launch_sweep_t<double, 512>(ctx, p, s, d_insts, n_active, init_tab, st);
      default: return launch_sweep_t<double, 1024>(ctx, p, s, d_insts, n_active, init_tab, st);
    }
  }
  switch (p->block) {
    case 256: return launch_sweep_t<float, 256>(ctx, p, s, d_insts, n_active, init_tab, st);
    case 512: return launch_sweep_t<float, 512>(ctx, p, s, d_insts, n_active, init_tab, st);
    default: return launch_sweep_t<float, 1024>(ctx, p, s, d_insts, n_active, init_tab, st);
  }
}

// One sweep over n_active instances.  With a block cap on the context the
// launch is split over instance ranges of at most max_blocks blocks each
// (an instance's tiles stay in one launch): the blocks of a launch that has
// started take every slot that frees up before a kernel of another stream
// gets one, so one long launch would hold back other streams' short kernels
// for its whole length.
int launch_sweep(pqs_ctx *ctx, pqs_plan *p, int s, const InstDev *d_insts, int64_t n_active,
                 const void *init_tab, hipStream_t st) {
  const int64_t tiles = p->dev_sweeps[size_t(s)].n_tiles;
  if (ctx->max_blocks <= 0 || tiles * n_active <= ctx->max_blocks)
    return launch_sweep_one(ctx, p, s, d_insts, n_active, init_tab, st);
  const int64_t per = std::max<int64_t>(1, ctx->max_blocks / std::max<int64_t>(1, tiles));
  for (int64_t i0 = 0; i0 < n_active; i0 += per) {
    const int rc = launch_sweep_one(ctx, p, s, d_insts + i0, std::min(per, n_active - i0), init_tab, st);
    if (rc) return rc;
  }
  return PQS_OK;
}

size_t amp_bytes(int dtype) { return dtype == PQS_C128 ? 16 : 8; }

}  // namespace

extern "C" {

int pqs_range_push(const char *name) {
  if (!name) return fail(PQS_ERR_ARG, "NULL range name");
  if (roctx().push) roctx().push(name);
  return PQS_OK;
}

int pqs_range_pop(void) {
  if (roctx().pop) roctx().pop();
  return PQS_OK;
}

int pqs_ctx_set_max_blocks(pqs_ctx *ctx, int64_t max_blocks) {
  if (!ctx) return fail(PQS_ERR_ARG, "ctx is NULL");
  if (max_blocks < 0) return fail(PQS_ERR_ARG, "negative block cap");
  ctx->max_blocks = max_blocks;
  return PQS_OK;
}

int pqs_upload(pqs_ctx *ctx, void *dst, const void *src, int64_t bytes, void *stream) {
  if (!ctx) return fail(PQS_ERR_ARG, "ctx is NULL");
  if (bytes < 0) return fail(PQS_ERR_ARG, "negative byte count");
  if (bytes == 0) return PQS_OK;
  if (!dst || !src) return fail(PQS_ERR_ARG, "NULL dst or src");
  DeviceGuard g(ctx->device);
  return upload(ctx, dst, src, size_t(bytes), (hipStream_t)stream);
}

int pqs_ctx_create(int device, pqs_ctx **out) {
  if (!out) return fail(PQS_ERR_ARG, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(PQS_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= count) return fail(PQS_ERR_ARG, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    std::string m = std::string("device is ") + prop.gcnArchName + ", the engine is built for gfx950";
    return fail(PQS_ERR_NODEVICE, m);
  }
  DeviceGuard g(device);
  pqs_ctx *c = new pqs_ctx;
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  c->max_lds = int(prop.sharedMemPerBlock);
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
    c->max_lds = v;
  std::snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
  {
    // the scratch array's address on this device, looked up once per device
    // with the device current (DeviceGuard above)
    static std::mutex mu;
    static void *addr[64] = {nullptr};
    std::lock_guard<std::mutex> lk(mu);
    if (device < 64 && !addr[device]) {
      void *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr) {
        delete c;
        return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
      }
      addr[device] = ptr;
    }
    c->scratch = device < 64 ? addr[device] : nullptr;
  }
  if (hipEventCreateWithFlags(&c->ws_event, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return fail(PQS_ERR_HIP, "hipEventCreate failed");
  }
  c->pins.reserve(pqs_ctx::kPinMax);
  *out = c;
  return PQS_OK;
}

int pqs_ctx_destroy(pqs_ctx *ctx) {
  if (!ctx) return PQS_OK;
  DeviceGuard g(ctx->device);
  (void)hipDeviceSynchronize();
  if (ctx->ws) {
    (void)hipFreeAsync(ctx->ws, ctx->ws_stream);
    (void)hipStreamSynchronize(ctx->ws_stream);
  }
  for (pqs_ctx::PinSlot &sl : ctx->pins) {
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.event) (void)hipEventDestroy(sl.event);
  }
  if (ctx->ws_event) (void)hipEventDestroy(ctx->ws_event);
  delete ctx;
  return PQS_OK;
}

int pqs_device_info(pqs_ctx *ctx, int32_t *n_cu, int32_t *max_lds, char *arch, int32_t arch_len) {
  if (!ctx) return fail(PQS_ERR_ARG, "ctx is NULL");
  if (n_cu) *n_cu = ctx->n_cu;
  if (max_lds) *max_lds = ctx->max_lds;
  if (arch && arch_len > 0) std::snprintf(arch, size_t(arch_len), "%s", ctx->arch);
  return PQS_OK;
}

int pqs_plan_destroy(pqs_plan *p) {
  if (!p) return PQS_OK;
  if (p->dev_device >= 0 && (p->d_ops || !p->jit_modules.empty())) {
    DeviceGuard g(p->dev_device);
    (void)hipDeviceSynchronize();
    if (p->d_ops) (void)hipFree(p->d_ops);
    if (p->d_terms) (void)hipFree(p->d_terms);
    if (p->d_phases) (void)hipFree(p->d_phases);
    if (p->d_sweeps) (void)hipFree(p->d_sweeps);
    for (void *m : p->jit_modules) (void)hipModuleUnload((hipModule_t)m);
  }
  if (p->up_event) (void)hipEventDestroy((hipEvent_t)p->up_event);
  delete p;
  return PQS_OK;
}

// Launch the accumulation of split-generator parts: for every AccRow of the
// table at d_acc (n_acc rows, sources at d_srcs), dst += sum of its sources.
static int launch_add_rows(pqs_plan *p, const AccRow *d_acc, const void *const *d_srcs, int64_t n_acc,
                           hipStream_t st) {
  if (n_acc <= 0) return PQS_OK;
  const int64_t dim = int64_t(1) << p->n;
  const unsigned gx = unsigned(std::min<int64_t>(grid_for(dim, 256), 2048));
  if (n_acc > 65535) return fail(PQS_ERR_UNSUPPORTED, "too many split-generator rows");
  if (p->dtype == PQS_C128)
    hipLaunchKernelGGL(k_add_rows<double2>, dim3(gx, unsigned(n_acc)), dim3(256), 0, st, d_acc, d_srcs, dim);
  else
    hipLaunchKernelGGL(k_add_rows<float2>, dim3(gx, unsigned(n_acc)), dim3(256), 0, st, d_acc, d_srcs, dim);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

// Run instances h of plan p (pqs_run, pqs_run_batch).
//  - An instance whose tangent's generator was split into parts gets one
//    extra instance per further part, evolving a workspace row from the same
//    input; after the sweeps the part rows are summed into the instance's
//    state.  A part that would read the instance's own buffer (in-place
//    input) starts from a copy of it.
//  - Rolling-layout plans change the bit layout from sweep to sweep, so every
//    instance runs out of place: its second buffer is a workspace row.  An
//    instance whose input is its own output buffer, while the first sweep
//    would write that buffer, first copies the input to its second buffer.
//  - pair_out (pqs_run_batch_pairs): also pair_out[i] = <psi_2i|psi_2i+1>
//    (complex128) for i < h.size() / 2.  The last sweep's pair kernel forms
//    the overlaps from its registers when the plan has one on this device;
//    otherwise they are read back from the states after the sweeps.
// Workspace: [alt rows | part rows | instance table | AccRow table | sources
// | overlap partials].
static int run_instances(pqs_ctx *ctx, pqs_plan *p, std::vector<InstDev> &h, const void *init_tab,
                         hipStream_t st, double *pair_out = nullptr) {
  const size_t sb = (size_t(1) << p->n) * amp_bytes(p->dtype);
  const size_t n_main = h.size();
  // extra instances for split generators: (owner, sub-tangent)
  std::vector<std::pair<size_t, int>> parts;
  if (p->n_sub > p->n_tangents)
    for (size_t i = 0; i < n_main; ++i) {
      const int t = h[i].tangent;
      if (t < 0) continue;
      for (int u = p->n_tangents; u < p->n_sub; ++u)
        if (p->tangent_parent[size_t(u)] == t) parts.push_back({i, u});
    }
  const size_t n_all = n_main + parts.size();
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t alt_off = 0;
  const size_t alt_b = p->rolling ? n_all * sb : 0;
  const size_t part_off = al(alt_off + alt_b);
  const size_t tab_off = al(part_off + parts.size() * sb);
  // one AccRow per owner with parts
  std::vector<AccRow> acc;
  std::vector<const void *> srcs;
  for (size_t k = 0; k < parts.size(); ++k) {
    void *row_dst = h[parts[k].first].state;
    if (k == 0 || parts[k - 1].first != parts[k].first)
      acc.push_back(AccRow{row_dst, int32_t(srcs.size()), 0});
    srcs.push_back(nullptr);  // filled below (workspace address)
    ++acc.back().count;
  }
  const size_t acc_off = al(tab_off + n_all * sizeof(InstDev));
  const size_t src_off = al(acc_off + acc.size() * sizeof(AccRow));
  const int S_sw = int(p->dev_sweeps.size());
  const int64_t n_pairs = int64_t(n_main / 2);
  const bool fused = pair_out && S_sw > 0 && parts.empty() && p->jit_pair_fn && p->jit_module &&
                     p->dev_device == ctx->device && S_sw - 1 < int(p->jit_fn.size()) && p->jit_fn[S_sw - 1];
  // partials: per pair and last-sweep tile (fused), else per pair and chunk
  const int64_t n_chunks = !pair_out ? 0
                           : fused   ? p->dev_sweeps[size_t(S_sw - 1)].n_tiles
                                     : ((int64_t(1) << p->n) + kInnerChunk - 1) / kInnerChunk;
  const size_t pov_off = al(src_off + srcs.size() * sizeof(void *));
  const size_t total = pov_off + size_t(n_pairs * n_chunks) * sizeof(double2);
  int rc = ws_reserve(ctx, total, st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  for (size_t k = 0; k < parts.size(); ++k) {
    const InstDev &own = h[parts[k].first];
    InstDev d = own;
    d.tangent = parts[k].second;
    d.state = ws + part_off + k * sb;
    srcs[k] = d.state;
    if (own.init == 0) {
      const void *in = own.src ? own.src : own.state;
      if (in == own.state) {
        HIPCHECK(hipMemcpyAsync(d.state, own.state, sb, hipMemcpyDeviceToDevice, st));
        d.src = nullptr;
      } else {
        d.src = in;
      }
    }
    h.push_back(d);
  }
  if (p->rolling) {
    const int S = int(p->dev_sweeps.size());
    const bool first_writes_state = ((S - 1) & 1) == 0;
    for (size_t i = 0; i < h.size(); ++i) {
      InstDev &d = h[i];
      d.alt = ws + alt_off + i * sb;
      d.pingpong = 1;
      if (d.init == 0 && (d.src == nullptr || d.src == d.state)) {
        if (first_writes_state) {
          HIPCHECK(hipMemcpyAsync(d.alt, d.state, sb, hipMemcpyDeviceToDevice, st));
          d.src = d.alt;
        } else {
          d.src = d.state;
        }
      }
    }
  }
  // one upload: instance table, AccRows, source pointers
  std::vector<char> blob(pov_off - tab_off, 0);
  std::memcpy(blob.data(), h.data(), h.size() * sizeof(InstDev));
  if (!acc.empty()) {
    std::memcpy(blob.data() + (acc_off - tab_off), acc.data(), acc.size() * sizeof(AccRow));
    std::memcpy(blob.data() + (src_off - tab_off), srcs.data(), srcs.size() * sizeof(void *));
  }
  rc = upload(ctx, ws + tab_off, blob.data(), blob.size(), st);
  if (rc) return rc;
  const InstDev *d_tab = (const InstDev *)(ws + tab_off);
  double2 *pov = (double2 *)(ws + pov_off);
  for (int s = 0; s < S_sw; ++s) {
    if (fused && s == S_sw - 1) {
      if (n_pairs == 0) break;
      const DevSweep &sw = p->dev_sweeps[size_t(s)];
      const int nthr = (1 << sw.k) >> pqs::kRegBits;
      const int64_t grid = n_pairs * sw.n_tiles;
      if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "grid too large: split the batch");
      long long na = (long long)h.size();
      void *args[] = {(void *)&d_tab, (void *)&na, (void *)&init_tab, (void *)&pov};
      HIPCHECK(hipModuleLaunchKernel((hipFunction_t)p->jit_pair_fn, unsigned(grid), 1, 1, unsigned(nthr), 1, 1, 0,
                                     st, args, nullptr));
      break;
    }
    rc = launch_sweep(ctx, p, s, d_tab, int64_t(h.size()), init_tab, st);
    if (rc) return rc;
  }
  rc = launch_add_rows(p, (const AccRow *)(ws + acc_off), (const void *const *)(ws + src_off),
                       int64_t(acc.size()), st);
  if (rc || !pair_out || n_pairs == 0) return rc;
  if (!fused) {
    // overlaps read back from the states: instance i's state is h[i].state
    // (a strided batch: pair i at h[2i].state, stride h[1].state - h[0].state)
    const int64_t dim = int64_t(1) << p->n;
    const size_t ab = amp_bytes(p->dtype);
    const int64_t stride = int64_t(((const char *)h[1].state - (const char *)h[0].state) / ptrdiff_t(ab));
    const int64_t grid = n_pairs * n_chunks;
    if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "too many pairs x chunks");
    if (p->dtype == PQS_C128)
      hipLaunchKernelGGL(k_pair_partial<double>, dim3(unsigned(grid)), dim3(RB), 0, st, (const double2 *)h[0].state,
                         stride, dim, n_chunks, pov);
    else
      hipLaunchKernelGGL(k_pair_partial<float>, dim3(unsigned(grid)), dim3(RB), 0, st, (const float2 *)h[0].state,
                         stride, dim, n_chunks, pov);
    HIPCHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sum_rows, dim3(unsigned(n_pairs)), dim3(RB), 0, st, pov, n_chunks, (double2 *)pair_out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_run(pqs_ctx *ctx, pqs_plan *p, const pqs_instance *insts, int64_t n_inst, const void *init_tab,
            void *stream) {
  Range rg_("pqs_run");
  if (!ctx || !p) return fail(PQS_ERR_ARG, "NULL ctx or plan");
  if (n_inst < 0 || (n_inst > 0 && !insts)) return fail(PQS_ERR_ARG, "bad instance array");
  if (n_inst == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  std::vector<InstDev> h((size_t)n_inst);
  for (int64_t i = 0; i < n_inst; ++i) {
    const pqs_instance &a = insts[i];
    if (!a.state_dev) return fail(PQS_ERR_ARG, "instance without a state buffer");
    if (a.init < 0 || a.init > 2) return fail(PQS_ERR_ARG, "instance init must be 0, 1 or 2");
    if (a.init == 2 && !init_tab) return fail(PQS_ERR_ARG, "product init needs an init table");
    if (a.tangent >= p->n_tangents) return fail(PQS_ERR_ARG, "tangent index beyond the plan's tangents");
    InstDev d{};
    d.state = a.state_dev;
    d.src = a.src_dev;
    d.params = a.params_dev;
    d.alt = nullptr;
    d.tangent = a.tangent < 0 ? -1 : a.tangent;
    d.init = a.init;
    d.start_sweep = 0;
    d.pingpong = 0;
    h[size_t(i)] = d;
  }
  int rc = plan_upload(ctx, p, st);
  if (rc) return rc;
  return run_instances(ctx, p, h, init_tab, st);
}

int pqs_run_batch(pqs_ctx *ctx, pqs_plan *p, void *states, int64_t state_stride, int64_t batch,
                  const double *params, int64_t param_stride, const int32_t *tangents, int32_t init,
                  const void *src, const void *init_tab, void *stream) {
  Range rg_("pqs_run_batch");
  if (!ctx || !p || !states) return fail(PQS_ERR_ARG, "NULL ctx, plan or states");
  if (batch < 0) return fail(PQS_ERR_ARG, "negative batch");
  if (batch == 0) return PQS_OK;
  if (state_stride < (int64_t(1) << p->n)) return fail(PQS_ERR_ARG, "state stride below 2^N");
  if (init < 0 || init > 2) return fail(PQS_ERR_ARG, "init must be 0, 1 or 2");
  if (init == 2 && !init_tab) return fail(PQS_ERR_ARG, "product init needs an init table");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t ab = amp_bytes(p->dtype);
  std::vector<InstDev> h((size_t)batch);
  for (int64_t i = 0; i < batch; ++i) {
    InstDev &d = h[(size_t)i];
    d.state = (char *)states + size_t(i) * size_t(state_stride) * ab;
    d.src = src;
    d.params = params ? params + i * param_stride : nullptr;
    d.alt = nullptr;
    d.tangent = tangents ? tangents[i] : -1;
    if (d.tangent >= p->n_tangents) return fail(PQS_ERR_ARG, "tangent index beyond the plan's tangents");
    if (d.tangent < 0) d.tangent = -1;
    d.init = init;
    d.start_sweep = 0;
    d.pingpong = 0;
  }
  int rc = plan_upload(ctx, p, st);
  if (rc) return rc;
  return run_instances(ctx, p, h, init_tab, st);
}

int pqs_run_batch_pairs(pqs_ctx *ctx, pqs_plan *p, void *states, int64_t state_stride, int64_t batch,
                        const double *params, int64_t param_stride, int32_t init, const void *src,
                        const void *init_tab, double *overlaps, void *stream) {
  Range rg_("pqs_run_batch_pairs");
  if (!ctx || !p || !states || !overlaps) return fail(PQS_ERR_ARG, "NULL ctx, plan, states or overlaps");
  if (batch < 0 || (batch & 1)) return fail(PQS_ERR_ARG, "batch must be even and not negative");
  if (batch == 0) return PQS_OK;
  if (state_stride < (int64_t(1) << p->n)) return fail(PQS_ERR_ARG, "state stride below 2^N");
  if (init < 0 || init > 2) return fail(PQS_ERR_ARG, "init must be 0, 1 or 2");
  if (init == 2 && !init_tab) return fail(PQS_ERR_ARG, "product init needs an init table");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t ab = amp_bytes(p->dtype);
  std::vector<InstDev> h((size_t)batch);
  for (int64_t i = 0; i < batch; ++i) {
    InstDev &d = h[(size_t)i];
    d.state = (char *)states + size_t(i) * size_t(state_stride) * ab;
    d.src = src;
    d.params = params ? params + i * param_stride : nullptr;
    d.alt = nullptr;
    d.tangent = -1;
    d.init = init;
    d.start_sweep = 0;
    d.pingpong = 0;
  }
  int rc = plan_upload(ctx, p, st);
  if (rc) return rc;
  return run_instances(ctx, p, h, init_tab, st, overlaps);
}

int pqs_run_multi(pqs_ctx *ctx, pqs_plan *const *plans, int32_t n_plans, const pqs_instance *insts,
                  const int32_t *inst_plan, const void *const *init_tabs, int64_t n_inst, void *stream) {
  Range rg_("pqs_run_multi");
  if (!ctx || !plans || n_plans <= 0) return fail(PQS_ERR_ARG, "NULL ctx or no plans");
  if (n_inst < 0 || (n_inst > 0 && (!insts || !inst_plan))) return fail(PQS_ERR_ARG, "bad instance arrays");
  if (n_inst == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int dtype = plans[0] ? plans[0]->dtype : 0;
  for (int32_t k = 0; k < n_plans; ++k) {
    pqs_plan *p = plans[k];
    if (!p) return fail(PQS_ERR_ARG, "NULL plan");
    if (p->dtype != dtype) return fail(PQS_ERR_ARG, "plans of one multi-plan run share one dtype");
    int rc = plan_upload(ctx, p, st);
    if (rc) return rc;
  }
  const size_t ab = amp_bytes(dtype);
  // launch groups: register-phase plans by tile size (one block size per
  // launch), the per-op LDS plans (tiles below 10 qubits) together
  auto group_of = [&](const pqs_plan *p) { return p->reg_exec ? p->tile_bits : 0; };
  std::vector<InstDev> h(static_cast<size_t>(n_inst));
  size_t alt_rows_b = 0;
  std::vector<size_t> alt_off(size_t(n_inst), 0);
  for (int64_t i = 0; i < n_inst; ++i) {
    const pqs_instance &a = insts[i];
    const int32_t pi = inst_plan[i];
    if (pi < 0 || pi >= n_plans) return fail(PQS_ERR_ARG, "instance plan index out of range");
    const pqs_plan *p = plans[pi];
    if (!a.state_dev) return fail(PQS_ERR_ARG, "instance without a state buffer");
    if (a.init < 0 || a.init > 2) return fail(PQS_ERR_ARG, "instance init must be 0, 1 or 2");
    if (a.init == 2 && (!init_tabs || !init_tabs[i])) return fail(PQS_ERR_ARG, "product init needs an init table");
    if (a.tangent >= p->n_tangents) return fail(PQS_ERR_ARG, "tangent index beyond the plan's tangents");
    if (a.tangent >= 0 && p->n_sub > p->n_tangents)
      return fail(PQS_ERR_UNSUPPORTED, "tangent runs of plans with split generators: use pqs_run");
    InstDev d{};
    d.state = a.state_dev;
    d.src = a.src_dev;
    d.params = a.params_dev;
    d.tangent = a.tangent < 0 ? -1 : a.tangent;
    d.init = a.init;
    if (p->rolling) {
      alt_off[size_t(i)] = alt_rows_b;
      alt_rows_b += (size_t(1) << p->n) * ab;
    }
    h[size_t(i)] = d;
  }
  // per (round, group): the active instances, their tiles laid end to end
  struct Launch {
    int round, group_k;
    size_t first, count;
    int64_t tiles;
    size_t lds;
    bool ldsops;
  };
  std::vector<MultiItem> items;
  std::vector<Launch> launches;
  std::vector<int> groups;
  int max_s = 0;
  for (int32_t k = 0; k < n_plans; ++k) {
    max_s = std::max(max_s, int(plans[k]->dev_sweeps.size()));
    const int gk = group_of(plans[k]);
    if (std::find(groups.begin(), groups.end(), gk) == groups.end()) groups.push_back(gk);
  }
  for (int r = 0; r < max_s; ++r)
    for (int gk : groups) {
      Launch L{r, gk, items.size(), 0, 0, 0, false};
      for (int64_t i = 0; i < n_inst; ++i) {
        const pqs_plan *p = plans[inst_plan[i]];
        if (group_of(p) != gk || r >= int(p->dev_sweeps.size())) continue;
        const DevSweep &sw = p->dev_sweeps[size_t(r)];
        MultiItem it{};
        it.tile_begin = L.tiles;
        it.inst = int32_t(i);
        it.plan = inst_plan[i];
        it.init_tab = init_tabs ? init_tabs[i] : nullptr;
        items.push_back(it);
        L.tiles += sw.n_tiles;
        size_t lds = (size_t(1) << sw.k) * ab;
        if (p->reg_exec) lds += size_t(sw.op_end - sw.op_begin) * sizeof(ROp);
        L.lds = std::max(L.lds, lds);
        L.ldsops = L.ldsops || sw.n_coef;
      }
      L.count = items.size() - L.first;
      if (L.count) launches.push_back(L);
    }
  if (items.size() > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "too many instances");
  std::vector<PlanDev> pdv(static_cast<size_t>(n_plans));
  for (int32_t k = 0; k < n_plans; ++k) {
    const pqs_plan *p = plans[k];
    pdv[size_t(k)] = PlanDev{(const DevSweep *)p->d_sweeps, (const DevOp *)p->d_ops, (const DevTerm *)p->d_terms,
                             (const DevPhase *)p->d_phases, int32_t(p->dev_sweeps.size()), p->n};
  }
  // workspace: [alt rows | instance table | plan records | items]
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t tab_off = al(alt_rows_b);
  const size_t pd_off = al(tab_off + h.size() * sizeof(InstDev));
  const size_t it_off = al(pd_off + pdv.size() * sizeof(PlanDev));
  const size_t total = it_off + items.size() * sizeof(MultiItem);
  int rc = ws_reserve(ctx, total, st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  for (int64_t i = 0; i < n_inst; ++i) {
    const pqs_plan *p = plans[inst_plan[i]];
    if (!p->rolling) continue;
    InstDev &d = h[size_t(i)];
    const size_t sb = (size_t(1) << p->n) * ab;
    d.alt = ws + alt_off[size_t(i)];
    d.pingpong = 1;
    const int S = int(p->dev_sweeps.size());
    if (d.init == 0 && (d.src == nullptr || d.src == d.state)) {
      if (((S - 1) & 1) == 0) {
        HIPCHECK(hipMemcpyAsync(d.alt, d.state, sb, hipMemcpyDeviceToDevice, st));
        d.src = d.alt;
      } else {
        d.src = d.state;
      }
    }
  }
  std::vector<char> blob(total - tab_off, 0);
  std::memcpy(blob.data(), h.data(), h.size() * sizeof(InstDev));
  std::memcpy(blob.data() + (pd_off - tab_off), pdv.data(), pdv.size() * sizeof(PlanDev));
  std::memcpy(blob.data() + (it_off - tab_off), items.data(), items.size() * sizeof(MultiItem));
  rc = upload(ctx, ws + tab_off, blob.data(), blob.size(), st);
  if (rc) return rc;
  const InstDev *d_tab = (const InstDev *)(ws + tab_off);
  const PlanDev *d_pd = (const PlanDev *)(ws + pd_off);
  const MultiItem *d_items = (const MultiItem *)(ws + it_off);
  static bool attr_set[64] = {false};
  if (ctx->device < 64 && !attr_set[ctx->device]) {
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg_multi<double, 512, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg_multi<double, 512, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg_multi<float, 1024, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    HIPCHECK(hipFuncSetAttribute((const void *)k_sweep_reg_multi<float, 1024, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
    attr_set[ctx->device] = true;
  }
  for (const Launch &L : launches) {
    if (L.tiles > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "grid too large: split the batch");
    if (L.lds > size_t(ctx->max_lds)) return fail(PQS_ERR_UNSUPPORTED, "tile and op records do not fit in LDS");
    const MultiItem *its = d_items + L.first;
    const int ni = int(L.count);
    if (L.group_k > 0) {
      const unsigned nthr = (1u << L.group_k) >> pqs::kRegBits;
      if (dtype == PQS_C128) {
        auto kern = L.ldsops ? k_sweep_reg_multi<double, 512, true> : k_sweep_reg_multi<double, 512, false>;
        hipLaunchKernelGGL(kern, dim3(unsigned(L.tiles)), dim3(nthr), L.lds, st, d_pd, d_tab, its, ni, L.round);
      } else {
        auto kern = L.ldsops ? k_sweep_reg_multi<float, 1024, true> : k_sweep_reg_multi<float, 1024, false>;
        hipLaunchKernelGGL(kern, dim3(unsigned(L.tiles)), dim3(nthr), L.lds, st, d_pd, d_tab, its, ni, L.round);
      }
    } else if (dtype == PQS_C128) {
      hipLaunchKernelGGL((k_sweep_multi<double, 256>), dim3(unsigned(L.tiles)), dim3(256), L.lds, st, d_pd, d_tab,
                         its, ni, L.round);
    } else {
      hipLaunchKernelGGL((k_sweep_multi<float, 256>), dim3(unsigned(L.tiles)), dim3(256), L.lds, st, d_pd, d_tab,
                         its, ni, L.round);
    }
    HIPCHECK(hipGetLastError());
  }
  return PQS_OK;
}

int64_t pqs_batch_table_bytes(int64_t batch) { return batch < 0 ? 0 : batch * int64_t(sizeof(InstDev)); }

int pqs_batch_table(pqs_ctx *ctx, pqs_plan *p, void *states, int64_t state_stride, int64_t batch,
                    const double *params, int64_t param_stride, int32_t init, const void *src, void *alt,
                    void *table, void *stream) {
  if (!ctx || !p || !states || !table) return fail(PQS_ERR_ARG, "NULL ctx, plan, states or table");
  if (batch <= 0) return fail(PQS_ERR_ARG, "batch must be positive");
  if (state_stride < (int64_t(1) << p->n)) return fail(PQS_ERR_ARG, "state stride below 2^N");
  if (init < 0 || init > 2) return fail(PQS_ERR_ARG, "init must be 0, 1 or 2");
  if (init == 0 && !src) return fail(PQS_ERR_ARG, "init 0 needs a source state for a tabled run");
  if (p->rolling && !alt) return fail(PQS_ERR_ARG, "rolling-layout plan needs the alt buffer");
  if (p->n_tangents > 0) return fail(PQS_ERR_UNSUPPORTED, "tabled runs take plans without tangents");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t ab = amp_bytes(p->dtype);
  std::vector<InstDev> h((size_t)batch);
  for (int64_t i = 0; i < batch; ++i) {
    InstDev &d = h[(size_t)i];
    d.state = (char *)states + size_t(i) * size_t(state_stride) * ab;
    d.src = src;
    d.params = params ? params + i * param_stride : nullptr;
    d.alt = p->rolling ? (char *)alt + size_t(i) * size_t(state_stride) * ab : nullptr;
    d.tangent = -1;
    d.init = init;
    d.start_sweep = 0;
    d.pingpong = p->rolling ? 1 : 0;
    if (p->rolling && src == d.state) return fail(PQS_ERR_ARG, "tabled rolling run: src must not be the output");
  }
  int rc = plan_upload(ctx, p, st);
  if (rc) return rc;
  return upload(ctx, table, h.data(), h.size() * sizeof(InstDev), st);
}

int pqs_run_table(pqs_ctx *ctx, pqs_plan *p, const void *table, int64_t batch, const void *init_tab,
                  void *stream) {
  Range rg_("pqs_run_table");
  if (!ctx || !p || !table) return fail(PQS_ERR_ARG, "NULL ctx, plan or table");
  if (batch <= 0) return fail(PQS_ERR_ARG, "batch must be positive");
  if (!p->d_ops || p->dev_device != ctx->device) return fail(PQS_ERR_ARG, "plan not on this device: table it first");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (int rc = plan_wait(p, st)) return rc;
  for (int s = 0; s < int(p->dev_sweeps.size()); ++s) {
    int rc = launch_sweep(ctx, p, s, (const InstDev *)table, batch, init_tab, st);
    if (rc) return rc;
  }
  return PQS_OK;
}

int pqs_run_batch_noalloc(pqs_ctx *ctx, pqs_plan *p, void *states, int64_t state_stride, int64_t batch,
                          const double *params, int64_t param_stride, int32_t init, const void *src, void *alt,
                          const void *init_tab, void *stream) {
  Range rg_("pqs_run_batch_noalloc");
  if (!ctx || !p || !states) return fail(PQS_ERR_ARG, "NULL ctx, plan or states");
  if (batch < 0) return fail(PQS_ERR_ARG, "negative batch");
  if (batch > kNoallocMaxBatch)
    return fail(PQS_ERR_ARG, "batch exceeds the no-allocation limit of 16384 instances: split it");
  if (batch == 0) return PQS_OK;
  if (state_stride < (int64_t(1) << p->n)) return fail(PQS_ERR_ARG, "state stride below 2^N");
  if (init < 0 || init > 2) return fail(PQS_ERR_ARG, "init must be 0, 1 or 2");
  if (init == 2 && !init_tab) return fail(PQS_ERR_ARG, "product init needs an init table");
  if (p->n_tangents > 0) return fail(PQS_ERR_UNSUPPORTED, "no-allocation runs take plans without tangents");
  if (!ctx->scratch) return fail(PQS_ERR_UNSUPPORTED, "no scratch array on this device");
  if (!p->d_ops || p->dev_device != ctx->device)
    return fail(PQS_ERR_ARG, "plan not on this device: run it once with pqs_run_batch first");
  if (p->rolling) {
    if (!alt) return fail(PQS_ERR_ARG, "rolling-layout plan needs the alt buffer (batch * state_stride amplitudes)");
    if (init == 0 && (!src || src == states)) return fail(PQS_ERR_ARG, "rolling-layout plan: src must not be the output");
  }
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (int rc = plan_wait(p, st)) return rc;
  InstDev *tab = static_cast<InstDev *>(ctx->scratch);
  const int64_t sbytes = state_stride * int64_t(amp_bytes(p->dtype));
  hipLaunchKernelGGL(k_fill_batch_table, dim3(grid_for(batch, 256)), dim3(256), 0, st, tab, (char *)states, sbytes,
                     batch, params, param_stride, init, src, (char *)alt, p->rolling ? 1 : 0);
  HIPCHECK(hipGetLastError());
  for (int s = 0; s < int(p->dev_sweeps.size()); ++s) {
    const int rc = launch_sweep(ctx, p, s, tab, batch, init_tab, st);
    if (rc) return rc;
  }
  return PQS_OK;
}

int pqs_jacobian(pqs_ctx *ctx, pqs_plan *p, const double *params, int32_t init, const void *src,
                 const void *init_tab, void *psi, void *jac, void *stream) {
  Range rg_("pqs_jacobian");
  if (!ctx || !p || !psi) return fail(PQS_ERR_ARG, "NULL ctx, plan or psi");
  if (p->n_tangents > 0 && !jac) return fail(PQS_ERR_ARG, "NULL jacobian buffer");
  if (init < 0 || init > 2) return fail(PQS_ERR_ARG, "init must be 0, 1 or 2");
  if (init == 0 && !src) return fail(PQS_ERR_ARG, "init 0 needs a source state");
  if (init == 2 && !init_tab) return fail(PQS_ERR_ARG, "product init needs an init table");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int S = int(p->dev_sweeps.size());
  const int64_t dim = int64_t(1) << p->n;
  const size_t ab = amp_bytes(p->dtype);
  int rc = plan_upload(ctx, p, st);
  if (rc) return rc;
  // Rolling layouts write every sweep out of place, so each tangent row
  // ping-pongs with a workspace row as the primal does; fixed layouts update
  // the rows in place.  The parts of split generators (sub-tangents
  // n_tangents..n_sub-1) evolve workspace rows that are summed into their
  // tangent's row at the end.  Workspace: [alt state | alt rows | part rows |
  // instance table | AccRow table | sources]
  const bool rows_pp = p->rolling && S > 1;
  if (rows_pp && init == 0 && src == psi)
    return fail(PQS_ERR_ARG, "with a rolling-layout plan the source state must not be the psi buffer");
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t sb = size_t(dim) * ab;
  const size_t alt_b = (S > 1) ? sb : 0;
  const size_t rows_off = al(alt_b);
  const size_t rows_b = rows_pp ? size_t(p->n_sub) * sb : 0;
  const int n_parts = p->n_sub - p->n_tangents;
  const size_t part_off = al(rows_off + rows_b);
  const size_t tab_off = al(part_off + size_t(n_parts) * sb);
  const size_t tb = size_t(1 + p->n_sub) * sizeof(InstDev);
  // AccRows: one per tangent with parts
  std::vector<AccRow> acc;
  std::vector<const void *> srcs;
  for (int t = 0; t < p->n_tangents && n_parts > 0; ++t) {
    AccRow a{(char *)jac + size_t(t) * sb, int32_t(srcs.size()), 0};
    for (int u = p->n_tangents; u < p->n_sub; ++u)
      if (p->tangent_parent[size_t(u)] == t) {
        srcs.push_back(nullptr);
        ++a.count;
      }
    if (a.count) acc.push_back(a);
  }
  const size_t acc_off = al(tab_off + tb);
  const size_t src_off = al(acc_off + acc.size() * sizeof(AccRow));
  const size_t total = src_off + srcs.size() * sizeof(void *);
  rc = ws_reserve(ctx, total, st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  void *alt = S > 1 ? (void *)ws : nullptr;
  auto W = [&](int s) -> void * { return ((S - 1 - s) & 1) == 0 ? psi : alt; };
  auto row_of = [&](int t) -> void * {
    return t < p->n_tangents ? (void *)((char *)jac + size_t(t) * sb)
                             : (void *)(ws + part_off + size_t(t - p->n_tangents) * sb);
  };
  {
    size_t k = 0;
    for (int t = 0; t < p->n_tangents && n_parts > 0; ++t)
      for (int u = p->n_tangents; u < p->n_sub; ++u)
        if (p->tangent_parent[size_t(u)] == t) srcs[k++] = row_of(u);
  }

  std::vector<int> order;
  for (int t = 0; t < p->n_sub; ++t)
    if (p->tangent_sweep[t] >= 0) order.push_back(t);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return p->tangent_sweep[a] < p->tangent_sweep[b]; });
  std::vector<InstDev> h;
  InstDev pr{};
  pr.state = psi;
  pr.alt = alt;
  pr.src = init == 0 ? src : nullptr;
  pr.params = params;
  pr.tangent = -1;
  pr.init = init;
  pr.start_sweep = 0;
  pr.pingpong = S > 1 ? 1 : 0;
  h.push_back(pr);
  for (int t : order) {
    InstDev d{};
    const int s0 = p->tangent_sweep[t];
    d.state = row_of(t);
    d.params = params;
    d.tangent = t;
    d.start_sweep = s0;
    d.pingpong = rows_pp ? 1 : 0;
    d.alt = rows_pp ? (void *)(ws + rows_off + size_t(t) * sb) : nullptr;
    if (s0 == 0) {
      d.src = init == 0 ? src : nullptr;
      d.init = init;
    } else {
      d.src = W(s0 - 1);
      d.init = 0;
    }
    h.push_back(d);
  }
  // tangents without a generator: zero rows
  for (int t = 0; t < p->n_tangents; ++t) {
    if (p->tangent_sweep[t] >= 0) continue;
    void *row = (char *)jac + size_t(t) * size_t(dim) * ab;
    if (p->dtype == PQS_C128)
      hipLaunchKernelGGL(k_fill_zero<double2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, (double2 *)row, dim);
    else
      hipLaunchKernelGGL(k_fill_zero<float2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, (float2 *)row, dim);
    HIPCHECK(hipGetLastError());
  }
  std::vector<char> blob(total - tab_off, 0);
  std::memcpy(blob.data(), h.data(), h.size() * sizeof(InstDev));
  if (!acc.empty()) {
    std::memcpy(blob.data() + (acc_off - tab_off), acc.data(), acc.size() * sizeof(AccRow));
    std::memcpy(blob.data() + (src_off - tab_off), srcs.data(), srcs.size() * sizeof(void *));
  }
  InstDev *d_tab = (InstDev *)(ws + tab_off);
  rc = upload(ctx, d_tab, blob.data(), blob.size(), st);
  if (rc) return rc;
  size_t active = 1;
  for (int s = 0; s < S; ++s) {
    while (active < h.size() && h[active].start_sweep <= s) ++active;
    rc = launch_sweep(ctx, p, s, d_tab, int64_t(active), init_tab, st);
    if (rc) return rc;
  }
  return launch_add_rows(p, (const AccRow *)(ws + acc_off), (const void *const *)(ws + src_off),
                         int64_t(acc.size()), st);
}

int pqs_inner(pqs_ctx *ctx, int32_t dtype, const void *base, int64_t dim, const int64_t *a_off,
              const int64_t *b_off, int64_t n_pairs, double *out, void *stream) {
  Range rg_("pqs_inner");
  if (!ctx || !base || !a_off || !b_off || !out) return fail(PQS_ERR_ARG, "NULL argument");
  if (dim <= 0 || n_pairs < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (n_pairs == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t nch = (dim + kInnerChunk - 1) / kInnerChunk;
  int rc = ws_reserve(ctx, size_t(n_pairs * nch) * sizeof(double2), st);
  if (rc) return rc;
  double2 *part = (double2 *)ctx->ws;
  const int64_t grid = n_pairs * nch;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "too many pairs x chunks");
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_inner_partial<double>, dim3(unsigned(grid)), dim3(RB), 0, st, (const double2 *)base, dim,
                       a_off, b_off, nch, part);
  else
    hipLaunchKernelGGL(k_inner_partial<float>, dim3(unsigned(grid)), dim3(RB), 0, st, (const float2 *)base, dim,
                       a_off, b_off, nch, part);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_rows, dim3(unsigned(n_pairs)), dim3(RB), 0, st, part, nch, (double2 *)out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_mw(pqs_ctx *ctx, int32_t dtype, const void *states, int32_t n, int64_t batch, double *rho, double *q,
           void *stream) {
  Range rg_("pqs_mw");
  if (!ctx || !states) return fail(PQS_ERR_ARG, "NULL argument");
  if (n < 1 || n > pqs::kMaxQubits || batch < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (batch == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int c = std::min(n, 11);
  const int64_t nch = int64_t(1) << (n - c);
  int rc = ws_reserve(ctx, size_t(batch * nch * n * 4) * sizeof(double), st);
  if (rc) return rc;
  double *part = (double *)ctx->ws;
  const int64_t grid = batch * nch;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "batch too large");
  const size_t lds = (size_t(1) << c) * amp_bytes(dtype);
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_mw_partial<double>, dim3(unsigned(grid)), dim3(RB), lds, st, (const double2 *)states, n, c,
                       nch, part);
  else
    hipLaunchKernelGGL(k_mw_partial<float>, dim3(unsigned(grid)), dim3(RB), lds, st, (const float2 *)states, n, c,
                       nch, part);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mw_final, dim3(unsigned(batch)), dim3(RB), 0, st, part, n, nch, rho, q);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

}  // extern "C"

namespace {

// K split of a Gram product: enough blocks to fill the device.
void gram_split(const pqs_ctx *ctx, int64_t m, int64_t nn, int64_t dim, int64_t &ks, int64_t &kchunk) {
  if (m == 1) {  // k_bvec_partial: BV_ROWS rows per block
    const int64_t tn = (nn + BV_ROWS - 1) / BV_ROWS;
    ks = std::max<int64_t>(1, (4 * int64_t(ctx->n_cu) + tn - 1) / tn);
    ks = std::min<int64_t>(ks, (dim + 1023) / 1024);
    ks = std::min<int64_t>(ks, 65535);
    kchunk = (dim + ks - 1) / ks;
    ks = (dim + kchunk - 1) / kchunk;
    return;
  }
  const int64_t tm = (m + GT - 1) / GT, tn = (nn + GT - 1) / GT;
  ks = std::max<int64_t>(1, (4 * int64_t(ctx->n_cu) + tm * tn - 1) / (tm * tn));
  ks = std::min<int64_t>(ks, (dim + 255) / 256);
  ks = std::min<int64_t>(ks, 65535);
  kchunk = (dim + ks - 1) / ks;
  kchunk = (kchunk + GK - 1) / GK * GK;
  ks = (dim + kchunk - 1) / kchunk;
}

size_t gram_part_bytes(const pqs_ctx *ctx, int64_t m, int64_t nn, int64_t dim) {
  int64_t ks, kc;
  gram_split(ctx, m, nn, dim, ks, kc);
  return size_t(ks * m * nn) * sizeof(double2);
}

// Gram into ``out`` with partials at ``part`` (>= gram_part_bytes).
int gram_impl(pqs_ctx *ctx, int32_t dtype, const void *a, int64_t lda, int64_t m, const void *b, int64_t ldb,
              int64_t nn, int64_t dim, double2 *out, double2 *part, hipStream_t st) {
  int64_t ks, kchunk;
  gram_split(ctx, m, nn, dim, ks, kchunk);
  if (m == 1) {
    const int64_t tb = (nn + BV_ROWS - 1) / BV_ROWS;
    if (tb > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "Gram too large");
    dim3 grid((unsigned)tb, (unsigned)ks);
    if (dtype == PQS_C128)
      hipLaunchKernelGGL(k_bvec_partial<double>, grid, dim3(256), 0, st, (const double2 *)a, (const double2 *)b, ldb,
                         nn, dim, kchunk, part);
    else
      hipLaunchKernelGGL(k_bvec_partial<float>, grid, dim3(256), 0, st, (const float2 *)a, (const float2 *)b, ldb,
                         nn, dim, kchunk, part);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(nn, 256)), dim3(256), 0, st, part, nn, ks, out);
    HIPCHECK(hipGetLastError());
    return PQS_OK;
  }
  const int64_t tm = (m + GT - 1) / GT, tn = (nn + GT - 1) / GT;
  if (tm > 0x7fffffff || tn > 65535) return fail(PQS_ERR_UNSUPPORTED, "Gram too large");
  dim3 grid((unsigned)tm, (unsigned)tn, (unsigned)ks);
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_gram_partial<double>, grid, dim3(256), 0, st, (const double2 *)a, lda, m,
                       (const double2 *)b, ldb, nn, dim, kchunk, part);
  else
    hipLaunchKernelGGL(k_gram_partial<float>, grid, dim3(256), 0, st, (const float2 *)a, lda, m, (const float2 *)b,
                       ldb, nn, dim, kchunk, part);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(m * nn, 256)), dim3(256), 0, st, part, m * nn, ks, out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

// A^dag B on the matrix cores (k_cgram_mfma).  Returns PQS_ERR_UNSUPPORTED,
// without touching the stream, for shapes it does not take: a K that is not
// a whole number of load-ring steps (fewer than 8 amplitudes) or operands
// whose rows cannot be addressed in 32 bits nor copied in 16-byte units.
int cgram_mfma(pqs_ctx *ctx, int32_t dtype, const void *a, int64_t lda, int64_t m, const void *b, int64_t ldb,
               int64_t nn, int64_t dim, double2 *out, hipStream_t st) {
  const int64_t cb = dtype == PQS_C128 ? 16 : 8;  // bytes per amplitude
  const int64_t ring = SYRING * CGK;
  if (dim % ring != 0) return PQS_ERR_UNSUPPORTED;
  const bool sym = a == b && lda == ldb && m == nn;
  // operands above 1 GiB in column blocks of at most 512 MiB copied to a
  // buffer with padded rows (as for the QFIM: TLB reach and L2 channels)
  const size_t rows = size_t(m) + (sym ? 0 : size_t(nn));
  const bool staged = rows * size_t(dim) * size_t(cb) > kQfimStageBytes;
  int64_t Kc = dim;
  if (staged) {
    if ((lda * cb) % 16 || (ldb * cb) % 16) return PQS_ERR_UNSUPPORTED;
    Kc = int64_t(kQfimBlockBytes / (rows * size_t(cb))) / ring * ring;
    Kc = std::max<int64_t>(ring, std::min<int64_t>(Kc, dim));
  }
  const int64_t Ks = Kc + 16;
  const int64_t la = staged ? Ks : lda, lb = staged ? Ks : ldb;
  if (size_t(m) * size_t(la) * size_t(cb) >= (size_t(1) << 32) || size_t(nn) * size_t(lb) * size_t(cb) >= (size_t(1) << 32))
    return PQS_ERR_UNSUPPORTED;
  const int64_t tm = (m + CGM - 1) / CGM, tn = (nn + CGN - 1) / CGN;
  std::vector<int2> tl;
  for (int64_t x = 0; x < tm; ++x)
    for (int64_t y = 0; y < tn; ++y)
      if (!sym || CGM * x <= CGN * y + CGN - 1) tl.push_back(int2{int(x), int(y)});
  const int64_t T = int64_t(tl.size());
  // the K split that leaves the fewest wave slots idle over whole rounds
  // (chunks of at least 256 amplitudes, partials within 2 GiB)
  const int64_t slots = 4 * SYWAVES * int64_t(ctx->n_cu);
  const int64_t ks_max = std::max<int64_t>(
      1, std::min<int64_t>({Kc / 256, int64_t(64), int64_t((size_t(2) << 30) / (size_t(m) * size_t(nn) * 16))}));
  int64_t ks = 1;
  double best = -1;
  for (int64_t k = 1; k <= ks_max; ++k) {
    const int64_t items = T * k, rounds = (items + slots - 1) / slots;
    const double fill = double(items) / double(rounds * slots);
    if (fill > best + 0.01) {
      best = fill;
      ks = k;
    }
  }
  int64_t kchunk = (Kc + ks - 1) / ks;
  kchunk = (kchunk + ring - 1) / ring * ring;
  ks = (Kc + kchunk - 1) / kchunk;
  const int64_t n_items = T * ks, per_xcd = (n_items + 7) / 8;
  const int64_t grid = 8 * ((per_xcd + 3) / 4);
  if (grid > 0x7fffffff) return PQS_ERR_UNSUPPORTED;
  const size_t al = 255;
  const size_t off_t = (size_t(ks) * size_t(m) * size_t(nn) * sizeof(double2) + al) & ~al;
  const size_t off_a = (off_t + tl.size() * sizeof(int2) + al) & ~al;
  const size_t off_b = (off_a + (staged ? size_t(m) * size_t(Ks) * size_t(cb) : 0) + al) & ~al;
  const size_t end = off_b + (staged && !sym ? size_t(nn) * size_t(Ks) * size_t(cb) : 0);
  int rc = ws_reserve(ctx, end, st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  double2 *part = (double2 *)ws;
  int2 *d_tiles = (int2 *)(ws + off_t);
  rc = upload(ctx, d_tiles, tl.data(), tl.size() * sizeof(int2), st);
  if (rc) return rc;
  for (int64_t c0 = 0; c0 < dim; c0 += Kc) {
    const int64_t w = std::min(Kc, dim - c0);
    const void *pa = (const char *)a + size_t(c0) * size_t(cb), *pb = (const char *)b + size_t(c0) * size_t(cb);
    if (staged) {
      const int64_t wu = w * cb / 16;
      hipLaunchKernelGGL(k_copy_cols, dim3(grid_for(m * wu, 256)), dim3(256), 0, st, (const uint4 *)pa,
                         lda * cb / 16, (uint4 *)(ws + off_a), Ks * cb / 16, wu, m * wu);
      HIPCHECK(hipGetLastError());
      pa = ws + off_a;
      if (sym) {
        pb = pa;
      } else {
        hipLaunchKernelGGL(k_copy_cols, dim3(grid_for(nn * wu, 256)), dim3(256), 0, st, (const uint4 *)pb,
                           ldb * cb / 16, (uint4 *)(ws + off_b), Ks * cb / 16, wu, nn * wu);
        HIPCHECK(hipGetLastError());
        pb = ws + off_b;
      }
    }
    const int accum = c0 > 0;
    if (dtype == PQS_C128)
      hipLaunchKernelGGL(k_cgram_mfma<double>, dim3(unsigned(grid)), dim3(256), 0, st, (const double *)pa, la, m,
                         (const double *)pb, lb, nn, w, kchunk, n_items, per_xcd, d_tiles, T, accum, part);
    else
      hipLaunchKernelGGL(k_cgram_mfma<float>, dim3(unsigned(grid)), dim3(256), 0, st, (const float *)pa, la, m,
                         (const float *)pb, lb, nn, w, kchunk, n_items, per_xcd, d_tiles, T, accum, part);
    HIPCHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_cgram_reduce, dim3(grid_for(m * nn, 256)), dim3(256), 0, st, (const double2 *)part, m, nn, ks,
                     int(sym), out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

}  // namespace

extern "C" {

int pqs_gram(pqs_ctx *ctx, int32_t dtype, const void *a, int64_t lda, int64_t m, const void *b, int64_t ldb,
             int64_t nn, int64_t dim, double *out, void *stream) {
  Range rg_("pqs_gram");
  if (!ctx || !a || !b || !out) return fail(PQS_ERR_ARG, "NULL argument");
  if (m <= 0 || nn <= 0 || dim <= 0 || lda < dim || ldb < dim) return fail(PQS_ERR_ARG, "bad sizes");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (m > 1) {
    const int rc = cgram_mfma(ctx, dtype, a, lda, m, b, ldb, nn, dim, (double2 *)out, st);
    if (rc != PQS_ERR_UNSUPPORTED) return rc;  // else: shapes the matrix-core path does not take
  }
  int rc = ws_reserve(ctx, gram_part_bytes(ctx, m, nn, dim), st);
  if (rc) return rc;
  return gram_impl(ctx, dtype, a, lda, m, b, ldb, nn, dim, (double2 *)out, (double2 *)ctx->ws, st);
}

int pqs_qfim(pqs_ctx *ctx, int32_t dtype, const void *jac, int64_t P, const void *psi, int64_t dim, double *F,
             void *stream) {
  Range rg_("pqs_qfim");
  if (!ctx || !jac || !psi || !F) return fail(PQS_ERR_ARG, "NULL argument");
  if (P <= 0 || dim <= 0) return fail(PQS_ERR_ARG, "bad sizes");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  {
    // Re(J J^dag) = R R^T, R = J as a real P x 2dim matrix: the upper
    // tiles of the symmetric product on the FP64 matrix cores, split over K
    // (rocBLAS dgemm on these tall-skinny shapes ran 2-10x slower, erratic
    // in P, and computes both triangles).  complex64 Jacobians are widened
    // to double in registers: the same products and sums.
    const int64_t K = 2 * dim;
    const size_t esz = dtype == PQS_C128 ? sizeof(double) : sizeof(float);
    const int64_t nt = (P + SYW - 1) / SYW;
    std::vector<int2> tl;
    for (int64_t a = 0; a < nt; ++a)
      for (int64_t b = a; b < nt; ++b) tl.push_back(int2{int(a), int(b)});
    const int64_t T = int64_t(tl.size());
    // A Jacobian above 1 GiB is taken in column blocks of at most 512 MiB,
    // each copied to a compact P x Kc buffer first: rows of 2^20 amplitudes
    // and more lie 16 MiB apart, and every load of the product kernel
    // touches 16 rows, so the TLB misses on nearly every access (measured:
    // 23% of translations missed, 12x the Jacobian fetched from HBM).  The
    // copy costs one extra read and write of J.
    const int64_t ring = SYRING * SYK;
    int64_t Kc = K;
    const bool pad = K % ring != 0;  // fewer than 8 amplitudes: zero-padded copy
    const bool staged = pad || size_t(P) * size_t(K) * esz > kQfimStageBytes;
    if (pad) {
      Kc = (K + ring - 1) / ring * ring;
    } else if (staged) {
      Kc = int64_t(kQfimBlockBytes / (size_t(P) * esz)) / ring * ring;
      Kc = std::max<int64_t>(ring, std::min<int64_t>(Kc, K));
    }
    // one item per wave slot (a single round), chunks of at least 1024 k, a
    // multiple of the load ring
    const int64_t slots = 4 * SYWAVES * int64_t(ctx->n_cu);
    int64_t ks = std::max<int64_t>(1, std::min<int64_t>(slots / T, Kc / 1024));
    int64_t kchunk = (Kc + ks - 1) / ks;
    kchunk = (kchunk + ring - 1) / ring * ring;
    ks = (Kc + kchunk - 1) / kchunk;
    const int64_t n_items = T * ks, per_xcd = (n_items + 7) / 8;
    const int64_t grid = 8 * ((per_xcd + 3) / 4);
    if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "QFIM too large");
    // workspace: [partials ks P^2 | C P^2 | b P | b partials ks P | tiles |
    // staged block (P Kc) | padded psi (Kc, short Jacobians only)]
    const size_t al = 255;
    const size_t off_c = (size_t(ks * P * P) * sizeof(double) + al) & ~al;
    const size_t off_b = (off_c + size_t(P * P) * sizeof(double) + al) & ~al;
    const size_t off_bp = (off_b + size_t(P) * sizeof(double2) + al) & ~al;
    const size_t off_t = (off_bp + size_t(ks * P) * sizeof(double2) + al) & ~al;
    const size_t off_s = (off_t + tl.size() * sizeof(int2) + al) & ~al;
    // staged rows are Kc + 32 elements apart: a power-of-two row stride
    // would put the 16 rows of every load on one L2 channel
    const int64_t Ks = Kc + 32;
    const size_t off_ps = (off_s + (staged ? size_t(P) * size_t(Ks) * esz : 0) + al) & ~al;
    if (size_t(P) * size_t(staged ? Ks : K) * esz >= (size_t(1) << 32))
      return fail(PQS_ERR_UNSUPPORTED, "QFIM too large");
    int rc = ws_reserve(ctx, off_ps + (pad ? size_t(Kc) * esz : 0), st);
    if (rc) return rc;
    char *ws = (char *)ctx->ws;
    double *C = (double *)(ws + off_c);
    double2 *bv = (double2 *)(ws + off_b);
    double2 *bpart = (double2 *)(ws + off_bp);
    int2 *d_tiles = (int2 *)(ws + off_t);
    void *stage = ws + off_s;
    rc = upload(ctx, d_tiles, tl.data(), tl.size() * sizeof(int2), st);
    if (rc) return rc;
    // column blocks after the first add to the partials (same stream, in
    // order); the diagonal tiles' waves also form b = <J_p|psi>
    for (int64_t c0 = 0; c0 < K; c0 += Kc) {
      const void *src = (const char *)jac + size_t(c0) * esz;
      const void *ps = (const char *)psi + size_t(c0) * esz;
      int64_t ld = K, w = std::min(Kc, K - c0);
      if (pad) {
        const int64_t units = P * Ks;
        void *pp = ws + off_ps;
        if (dtype == PQS_C128) {
          hipLaunchKernelGGL(k_pad_cols<double>, dim3(grid_for(units, 256)), dim3(256), 0, st, (const double *)src, K,
                             (double *)stage, Ks, units);
          hipLaunchKernelGGL(k_pad_cols<double>, dim3(1), dim3(256), 0, st, (const double *)ps, K, (double *)pp, Kc,
                             Kc);
        } else {
          hipLaunchKernelGGL(k_pad_cols<float>, dim3(grid_for(units, 256)), dim3(256), 0, st, (const float *)src, K,
                             (float *)stage, Ks, units);
          hipLaunchKernelGGL(k_pad_cols<float>, dim3(1), dim3(256), 0, st, (const float *)ps, K, (float *)pp, Kc, Kc);
        }
        HIPCHECK(hipGetLastError());
        src = stage;
        ps = pp;
        ld = Ks;
        w = Kc;
      } else if (staged) {
        const int64_t units = P * (w * int64_t(esz) / 16);
        hipLaunchKernelGGL(k_copy_cols, dim3(grid_for(units, 256)), dim3(256), 0, st, (const uint4 *)src,
                           K * int64_t(esz) / 16, (uint4 *)stage, Ks * int64_t(esz) / 16, w * int64_t(esz) / 16, units);
        HIPCHECK(hipGetLastError());
        src = stage;
        ld = Ks;
      }
      const int accum = c0 > 0;
      if (dtype == PQS_C128)
        hipLaunchKernelGGL(k_syrk_mfma<double>, dim3(unsigned(grid)), dim3(256), 0, st, (const double *)src, ld, P, w,
                           kchunk, n_items, per_xcd, d_tiles, T, (const double *)ps, accum, (double *)ws, bpart);
      else
        hipLaunchKernelGGL(k_syrk_mfma<float>, dim3(unsigned(grid)), dim3(256), 0, st, (const float *)src, ld, P, w,
                           kchunk, n_items, per_xcd, d_tiles, T, (const float *)ps, accum, (double *)ws, bpart);
      HIPCHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_syrk_reduce, dim3(grid_for(P * P, 256)), dim3(256), 0, st, (const double *)ws, P, ks, C);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(P, 256)), dim3(256), 0, st, bpart, P, ks, bv);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_qfim_combine_sym, dim3(grid_for(P * P, 256)), dim3(256), 0, st, C, bv, P, F);
    HIPCHECK(hipGetLastError());
    return PQS_OK;
  }
}

// Pauli terms from qubit masks (bit q = qubit q; Y = X and Z) to index masks
// (bit n-1-q), with i^{#Y} folded into the coefficient.
static std::vector<PauliTerm> pauli_terms(int n, const uint64_t *xm, const uint64_t *zm, const double *coef,
                                          int nt) {
  std::vector<PauliTerm> h(size_t(std::max(nt, 1)));
  for (int t = 0; t < nt; ++t) {
    uint64_t x = 0, z = 0;
    int ny = 0;
    for (int q = 0; q < n; ++q) {
      const bool bx = (xm[t] >> q) & 1, bz = (zm[t] >> q) & 1;
      if (bx) x |= uint64_t(1) << (n - 1 - q);
      if (bz) z |= uint64_t(1) << (n - 1 - q);
      if (bx && bz) ++ny;
    }
    double cr = coef[2 * t], ci = coef[2 * t + 1];
    for (int r = 0; r < ny; ++r) {
      const double t2 = cr;
      cr = -ci;
      ci = t2;
    }
    h[size_t(t)] = PauliTerm{x, z, cr, ci};
  }
  return h;
}

int pqs_apply_pauli(pqs_ctx *ctx, int32_t dtype, const void *x, int32_t n, const uint64_t *xm, const uint64_t *zm,
                    const double *coef, int32_t nt, void *out, void *stream) {
  Range rg_("pqs_apply_pauli");
  if (!ctx || !x || !out || (nt > 0 && (!xm || !zm || !coef))) return fail(PQS_ERR_ARG, "NULL argument");
  if (n < 1 || n > pqs::kMaxQubits || nt < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (x == out) return fail(PQS_ERR_ARG, "out must not be the input state");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t dim = int64_t(1) << n;
  std::vector<PauliTerm> h = pauli_terms(n, xm, zm, coef, nt);
  int rc = ws_reserve(ctx, h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  PauliTerm *d_terms = (PauliTerm *)ctx->ws;
  rc = upload(ctx, d_terms, h.data(), h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_apply_pauli<double2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, (const double2 *)x, dim,
                       d_terms, nt, (double2 *)out);
  else
    hipLaunchKernelGGL(k_apply_pauli<float2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, (const float2 *)x, dim,
                       d_terms, nt, (float2 *)out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_pauli_transition(pqs_ctx *ctx, int32_t dtype, const void *bra, const void *ket, int32_t n,
                         const uint64_t *xm, const uint64_t *zm, const double *coef, int32_t nt, double *out,
                         void *stream) {
  Range rg_("pqs_pauli_transition");
  if (!ctx || !bra || !ket || !out || (nt > 0 && (!xm || !zm || !coef))) return fail(PQS_ERR_ARG, "NULL argument");
  if (n < 1 || n > pqs::kMaxQubits || nt < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (nt == 0) return PQS_OK;
  if (nt > 65535) return fail(PQS_ERR_UNSUPPORTED, "too many terms");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t dim = int64_t(1) << n;
  const int64_t nch = (dim + kInnerChunk - 1) / kInnerChunk;
  std::vector<PauliTerm> h = pauli_terms(n, xm, zm, coef, nt);
  const size_t part_b = size_t(nt * nch) * sizeof(double2);
  const size_t off_t = (part_b + 255) & ~size_t(255);
  int rc = ws_reserve(ctx, off_t + h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  PauliTerm *d_terms = (PauliTerm *)(ws + off_t);
  rc = upload(ctx, d_terms, h.data(), h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  if (nch > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "state too large");
  const dim3 grid{unsigned(nch), unsigned(nt), 1};
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_transition_partial<double2>, grid, dim3(RB), 0, st, (const double2 *)bra,
                       (const double2 *)ket, dim, nch, d_terms, (double2 *)ws);
  else
    hipLaunchKernelGGL(k_transition_partial<float2>, grid, dim3(RB), 0, st, (const float2 *)bra,
                       (const float2 *)ket, dim, nch, d_terms, (double2 *)ws);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_rows, dim3(unsigned(nt)), dim3(RB), 0, st, (const double2 *)ws, nch, (double2 *)out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_expect_pauli(pqs_ctx *ctx, int32_t dtype, const void *states, int32_t n, int64_t batch, const uint64_t *xm,
                     const uint64_t *zm, const double *coef, int32_t nt, double *out, void *stream) {
  Range rg_("pqs_expect_pauli");
  if (!ctx || !states || !out || (nt > 0 && (!xm || !zm || !coef))) return fail(PQS_ERR_ARG, "NULL argument");
  if (n < 1 || n > pqs::kMaxQubits || batch < 0 || nt < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (batch == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t dim = int64_t(1) << n;
  const int64_t nch = (dim + kInnerChunk - 1) / kInnerChunk;
  std::vector<PauliTerm> h(size_t(std::max(nt, 1)));
  for (int t = 0; t < nt; ++t) {
    // convert qubit masks (bit q = qubit q) to index masks (bit n-1-q)
    uint64_t x = 0, z = 0;
    int ny = 0;
    for (int q = 0; q < n; ++q) {
      const bool bx = (xm[t] >> q) & 1, bz = (zm[t] >> q) & 1;
      if (bx) x |= uint64_t(1) << (n - 1 - q);
      if (bz) z |= uint64_t(1) << (n - 1 - q);
      if (bx && bz) ++ny;
    }
    double cr = coef[2 * t], ci = coef[2 * t + 1];
    for (int r = 0; r < ny; ++r) {
      const double t2 = cr;
      cr = -ci;
      ci = t2;
    }
    h[size_t(t)] = PauliTerm{x, z, cr, ci};
  }
  const size_t part_b = size_t(batch * nch) * sizeof(double2);
  const size_t off_t = (part_b + 255) & ~size_t(255);
  int rc = ws_reserve(ctx, off_t + h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  char *ws = (char *)ctx->ws;
  PauliTerm *d_terms = (PauliTerm *)(ws + off_t);
  rc = upload(ctx, d_terms, h.data(), h.size() * sizeof(PauliTerm), st);
  if (rc) return rc;
  const int64_t grid = batch * nch;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "batch too large");
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_expect_partial<double>, dim3(unsigned(grid)), dim3(RB), 0, st, (const double2 *)states, n,
                       nch, d_terms, nt, (double2 *)ws);
  else
    hipLaunchKernelGGL(k_expect_partial<float>, dim3(unsigned(grid)), dim3(RB), 0, st, (const float2 *)states, n,
                       nch, d_terms, nt, (double2 *)ws);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_rows, dim3(unsigned(batch)), dim3(RB), 0, st, (const double2 *)ws, nch, (double2 *)out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_histogram(pqs_ctx *ctx, const double *values, int64_t n, const double *edges, int32_t nb,
                  int64_t *counts, void *stream) {
  if (!ctx || !edges || !counts || (n > 0 && !values)) return fail(PQS_ERR_ARG, "NULL argument");
  if (nb < 1 || n < 0) return fail(PQS_ERR_ARG, "bad sizes");
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(nb, 256)), dim3(256), 0, st, (unsigned long long *)counts,
                     int64_t(nb));
  HIPCHECK(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_histogram, dim3(grid_for(n, 256)), dim3(256), 0, st, values, n, edges, nb,
                       (unsigned long long *)counts);
    HIPCHECK(hipGetLastError());
  }
  return PQS_OK;
}

int pqs_apply_dense(pqs_ctx *ctx, int32_t dtype, const double *u, int64_t dim, const void *states,
                    int64_t stride, int64_t batch, void *out, void *stream) {
  Range rg_("pqs_apply_dense");
  if (!ctx || !u || !states || !out) return fail(PQS_ERR_ARG, "NULL argument");
  if (dim <= 0 || batch < 0 || stride < dim) return fail(PQS_ERR_ARG, "bad sizes");
  if (dim > (int64_t(1) << 14)) return fail(PQS_ERR_UNSUPPORTED, "dense operators are limited to 14 qubits");
  if (states == out) return fail(PQS_ERR_ARG, "out must not be the input states");
  if (batch == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const unsigned grid = grid_for(dim, 4);
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_apply_dense<double2>, dim3(grid), dim3(256), 0, st, (const double2 *)u, dim,
                       (const double2 *)states, stride, batch, (double2 *)out);
  else
    hipLaunchKernelGGL(k_apply_dense<float2>, dim3(grid), dim3(256), 0, st, (const double2 *)u, dim,
                       (const float2 *)states, stride, batch, (float2 *)out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_abs2(pqs_ctx *ctx, const double *z, int64_t n, double *out, void *stream) {
  if (!ctx || (n > 0 && (!z || !out))) return fail(PQS_ERR_ARG, "NULL argument");
  if (n <= 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipLaunchKernelGGL(k_abs2, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, (const double2 *)z, n,
                     out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_magic(pqs_ctx *ctx, int32_t dtype, const void *states, int32_t n, int64_t batch, double alpha,
              double *out, void *stream) {
  Range rg_("pqs_magic");
  if (!ctx || !states || !out) return fail(PQS_ERR_ARG, "NULL argument");
  if (n < 1 || batch < 0) return fail(PQS_ERR_ARG, "bad sizes");
  if (n > 13) return fail(PQS_ERR_UNSUPPORTED, "stabiliser entropy is limited to 13 qubits (4^n work)");
  if (batch == 0) return PQS_OK;
  DeviceGuard g(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const int64_t D = int64_t(1) << n;
  int rc = ws_reserve(ctx, size_t(batch * D) * sizeof(double), st);
  if (rc) return rc;
  const size_t lds = size_t(D) * sizeof(double2);
  if (int(lds) > ctx->max_lds) return fail(PQS_ERR_UNSUPPORTED, "state does not fit in LDS");
  static bool attr[2] = {false, false};
  if (lds > 65536) {
    if (!attr[0]) {
      HIPCHECK(hipFuncSetAttribute((const void *)k_magic_partial<double>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
      HIPCHECK(hipFuncSetAttribute((const void *)k_magic_partial<float>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, ctx->max_lds));
      attr[0] = true;
    }
  }
  const int64_t grid = batch * D;
  if (grid > 0x7fffffff) return fail(PQS_ERR_UNSUPPORTED, "batch too large");
  double *part = (double *)ctx->ws;
  if (dtype == PQS_C128)
    hipLaunchKernelGGL(k_magic_partial<double>, dim3(unsigned(grid)), dim3(RB), lds, st, (const double2 *)states, n,
                       alpha, part);
  else
    hipLaunchKernelGGL(k_magic_partial<float>, dim3(unsigned(grid)), dim3(RB), lds, st, (const float2 *)states, n,
                       alpha, part);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_rows_real, dim3(unsigned(batch)), dim3(RB), 0, st, part, D, out);
  HIPCHECK(hipGetLastError());
  return PQS_OK;
}

int pqs_plan_jit_source(const pqs_plan *p, char *buf, int64_t cap, int64_t *len, int32_t *n_jit_sweeps) {
  if (!p) return fail(PQS_ERR_ARG, "plan is NULL");
  std::vector<int> sw;
  std::string src = pqs::jit_source(*p, sw);
  if (len) *len = int64_t(src.size());
  if (n_jit_sweeps) *n_jit_sweeps = int32_t(sw.size());
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(size_t(cap - 1), src.size());
    std::memcpy(buf, src.data(), n);
    buf[n] = 0;
  }
  return PQS_OK;
}

int pqs_plan_jit(pqs_ctx *ctx, pqs_plan *p, int32_t load) {
  if (!p) return fail(PQS_ERR_ARG, "plan is NULL");
  if (load && !ctx) return fail(PQS_ERR_ARG, "loading needs a context");
  if (p->jit_module) return PQS_OK;
  std::vector<int> sw;
  for (int s = 0; s < int(p->dev_sweeps.size()); ++s)
    if (pqs::jit_sweep_ok(*p, s)) sw.push_back(s);
  if (sw.empty()) return PQS_OK;
  std::vector<std::vector<char>> code(sw.size());
  std::vector<std::string> errs(sw.size());
  std::vector<int> rcs(sw.size(), PQS_OK);
  // compile the sweeps in parallel (independent hiprtc programs)
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= sw.size()) return;
      rcs[i] = jit_code_cached(pqs::jit_source_sweep(*p, sw[i]), code[i], errs[i]);
    }
  };
  unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
  nt = std::min<unsigned>(nt, unsigned(sw.size()));
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
  for (size_t i = 0; i < sw.size(); ++i)
    if (rcs[i]) return fail(PQS_ERR_HIP, "hiprtc compile failed: " + errs[i].substr(0, 2000));
  if (!load) return PQS_OK;
  DeviceGuard g(ctx->device);
  // bind the plan to this device; its arrays are copied at the first launch
  if (p->dev_device >= 0 && p->dev_device != ctx->device)
    return fail(PQS_ERR_ARG, "plan already bound to another device");
  p->dev_device = ctx->device;
  std::vector<void *> mods, fns(p->dev_sweeps.size(), nullptr);
  void *pair_fn = nullptr;
  const int last = int(p->dev_sweeps.size()) - 1;
  for (size_t i = 0; i < sw.size(); ++i) {
    hipModule_t mod;
    hipError_t e = hipModuleLoadData(&mod, code[i].data());
    if (e == hipSuccess) {
      hipFunction_t f;
      char name[64];
      std::snprintf(name, sizeof name, "pqs_jit_sweep_%d", sw[i]);
      e = hipModuleGetFunction(&f, mod, name);
      mods.push_back((void *)mod);
      if (e == hipSuccess) fns[sw[i]] = (void *)f;
      if (e == hipSuccess && sw[i] == last) {
        // the pair form exists only where the emitter wrote one
        std::snprintf(name, sizeof name, "pqs_jit_pair_%d", last);
        hipFunction_t fp;
        if (hipModuleGetFunction(&fp, mod, name) == hipSuccess) pair_fn = (void *)fp;
        (void)hipGetLastError();
      }
    }
    if (e != hipSuccess) {
      for (void *m : mods) (void)hipModuleUnload((hipModule_t)m);
      return hip_fail(e, "loading a specialised sweep");
    }
  }
  p->jit_modules = mods;
  p->jit_fn = fns;
  p->jit_pair_fn = pair_fn;
  p->jit_module = mods.empty() ? nullptr : mods[0];
  return PQS_OK;
}

int pqs_plan_jit_sweep_source(const pqs_plan *p, int32_t s, char *buf, int64_t cap, int64_t *len) {
  if (!p) return fail(PQS_ERR_ARG, "plan is NULL");
  if (s < 0 || s >= int32_t(p->dev_sweeps.size())) return fail(PQS_ERR_ARG, "sweep index out of range");
  if (!pqs::jit_sweep_ok(*p, s)) {
    if (len) *len = 0;
    return PQS_OK;
  }
  std::string src = pqs::jit_source_sweep(*p, s);
  if (len) *len = int64_t(src.size());
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(size_t(cap - 1), src.size());
    std::memcpy(buf, src.data(), n);
    buf[n] = 0;
  }
  return PQS_OK;
}

int pqs_jit_prepare(const char *src, int64_t len) {
  if (!src || len <= 0) return fail(PQS_ERR_ARG, "empty source");
  std::vector<char> code;
  std::string log;
  int rc = jit_code_cached(std::string(src, size_t(len)), code, log);
  if (rc) return fail(PQS_ERR_HIP, "hiprtc compile failed: " + log.substr(0, 2000));
  return PQS_OK;
}

}  // extern "C"


