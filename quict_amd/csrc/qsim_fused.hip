// Gate fusion: many gates per full-state pass.
//
// The reference applies one kernel per gate (QuICT/simulation/state_vector/
// statevector_simulator.py:155-164); at 30 qubits every gate streams the whole
// 16 GiB state through HBM.  Here a host planner groups gates into passes; a
// pass owns a set of K "tile" bits (the L lowest amplitude bits, always, plus
// up to K-L arbitrary higher bits).  The state splits into 2^(n-K) disjoint
// tiles of 2^K amplitudes; one workgroup loads one tile, applies every gate of
// the pass to it and writes it back: one HBM read+write per pass instead of
// per gate.
//
// Inside a tile the 2^K amplitudes live in registers: thread t of the 2^TB
// threads holds NR = 2^RB amplitudes.  A "layout" names which RB tile bits
// index the registers (the rest index the thread).  A gate whose non-diagonal
// targets are register bits is applied entirely in registers; controls and up
// to two diagonal target bits may sit on thread bits (lane predicates /
// selects).  When a gate needs a thread bit in registers the tile is
// re-laid-out through LDS (one ds write + one ds read per amplitude) — the
// only LDS traffic.  Loads/stores use a layout whose thread bits include the
// L low bits, so every wavefront touches contiguous 2^L-amplitude runs.
//
// The planner (plan_passes) is greedy with commutation by disjoint support:
// scanning the remaining gates in order, a gate joins the current pass if its
// bits fit in the tile and it touches no bit of a gate already deferred.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace qs {

enum FKind : int32_t {
  F_RELAYOUT = 0,
  F_DENSE1 = 1,
  F_PERM1 = 2,   // a0' = m0*a1, a1' = m1*a0 (pure exchange when pure != 0)
  F_DIAG = 3,
  F_DENSE2 = 4,
  F_SWAP2 = 5,   // a01' = m0*a10, a10' = m1*a01 (pure exchange when pure != 0)
  F_DENSE3 = 6,
  F_DENSE4 = 7,
};

constexpr int FMAX_TB = 10;
constexpr int FMAX_RB = 5;

struct FOp {
  int32_t kind;
  int32_t coef;          // complex offset of this op's payload in the pass coefficient array
  uint32_t creg;         // control mask in register-index space
  uint32_t cthr;         // control mask in thread-index space
  int32_t npos;          // register positions used (targets, MSB first)
  int32_t pos[4];
  int32_t ntt;           // DIAG: number of thread-bit targets (<= 2)
  int32_t tpos[2];       // DIAG: their thread-index bits, MSB first
  int32_t pure;          // PERM1/SWAP2 without phases
  int32_t rs[FMAX_RB];   // RELAYOUT: new register bits (tile-local), register bit i = rs[i]
  int32_t ts[FMAX_TB];   // RELAYOUT: new thread bits (tile-local), thread bit i = ts[i]
};

struct PassHdr {
  int32_t nops;
  int32_t op0;
  int32_t coef0;
  int32_t uses_lds;
  int32_t tbits[16];     // tile-local bit -> global bit
  int32_t sorted[16];    // tile bits ascending (global), for the tile base
  int32_t rs0[FMAX_RB], ts0[FMAX_TB];  // load layout (tile-local bits)
  int32_t rs1[FMAX_RB], ts1[FMAX_TB];  // store layout
};

// ---- device helpers ------------------------------------------------------------

// XOR fold of the 3-bit chunks above bit 3: spreads power-of-two strides over LDS banks.
__device__ __forceinline__ uint32_t lds_swz(uint32_t j) {
  uint32_t h = j >> 3, f = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    f ^= h & 7;
    h >>= 3;
  }
  return j ^ f;
}

template <int RB> __device__ __forceinline__ uint32_t reg_local(int r, const int (&rs)[RB]) {
  uint32_t j = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b) j |= (uint32_t)((r >> b) & 1) << rs[b];
  return j;
}

template <typename R> __device__ __forceinline__ cx<R> cload(const cx<R>* p) { return *p; }

template <typename R, int NR, int P>
__device__ __forceinline__ void do_dense1(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, bool tok) {
  const cx<R> m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & (1 << P)) continue;
    if (((uint32_t)r & creg) != creg) continue;
    if (tok) {
      const int r1 = r | (1 << P);
      const cx<R> x0 = a[r], x1 = a[r1];
      a[r] = cfma(mul(m00, x0), m01, x1);
      a[r1] = cfma(mul(m10, x0), m11, x1);
    }
  }
}

template <typename R, int NR, int P>
__device__ __forceinline__ void do_perm1(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, bool tok, int pure) {
  if (pure) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r & (1 << P)) continue;
      if (((uint32_t)r & creg) != creg) continue;
      if (tok) {
        const int r1 = r | (1 << P);
        const cx<R> t = a[r];
        a[r] = a[r1];
        a[r1] = t;
      }
    }
  } else {
    const cx<R> m0 = m[0], m1 = m[1];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r & (1 << P)) continue;
      if (((uint32_t)r & creg) != creg) continue;
      if (tok) {
        const int r1 = r | (1 << P);
        const cx<R> x0 = a[r], x1 = a[r1];
        a[r] = mul(m0, x1);
        a[r1] = mul(m1, x0);
      }
    }
  }
}

// positions P0 (MSB) > P1 (LSB)
template <typename R, int NR, int P0, int P1>
__device__ __forceinline__ void do_dense2(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, bool tok) {
  cx<R> mm[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) mm[i] = m[i];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & ((1 << P0) | (1 << P1))) continue;
    if (((uint32_t)r & creg) != creg) continue;
    if (tok) {
      const int idx[4] = {r, r | (1 << P1), r | (1 << P0), r | (1 << P0) | (1 << P1)};
      cx<R> x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = a[idx[j]];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cx<R> acc = mul(mm[4 * i], x[0]);
#pragma unroll
        for (int j = 1; j < 4; ++j) acc = cfma(acc, mm[4 * i + j], x[j]);
        a[idx[i]] = acc;
      }
    }
  }
}

template <typename R, int NR, int P0, int P1>
__device__ __forceinline__ void do_swap2(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, bool tok, int pure) {
  const cx<R> m0 = m[0], m1 = m[1];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & ((1 << P0) | (1 << P1))) continue;
    if (((uint32_t)r & creg) != creg) continue;
    if (tok) {
      const int i01 = r | (1 << P1), i10 = r | (1 << P0);
      const cx<R> x01 = a[i01], x10 = a[i10];
      if (pure) {
        a[i01] = x10;
        a[i10] = x01;
      } else {
        a[i01] = mul(m0, x10);
        a[i10] = mul(m1, x01);
      }
    }
  }
}

// Offset (register index bits) of matrix index q for targets given by mask PM,
// highest set bit = most significant matrix bit.
template <uint32_t PM, int KB> __host__ __device__ constexpr int densek_off(int q) {
  int o = 0, k = 0;
  for (int b = 31; b >= 0; --b)
    if ((PM >> b) & 1u) {
      if ((q >> (KB - 1 - k)) & 1) o |= 1 << b;
      ++k;
    }
  return o;
}

// dense on KB register bits given by mask PM
template <typename R, int NR, int KB, uint32_t PM>
__device__ __forceinline__ void do_densek(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, bool tok) {
  constexpr int D = 1 << KB;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if ((uint32_t)r & PM) continue;
    if (((uint32_t)r & creg) != creg) continue;
    if (tok) {
      cx<R> x[D];
#pragma unroll
      for (int q = 0; q < D; ++q) x[q] = a[r | densek_off<PM, KB>(q)];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        cx<R> acc = mul(m[i * D], x[0]);
#pragma unroll
        for (int j = 1; j < D; ++j) acc = cfma(acc, m[i * D + j], x[j]);
        a[r | densek_off<PM, KB>(i)] = acc;
      }
    }
  }
}

template <typename R, int NR>
__device__ __forceinline__ void do_diag(cx<R> (&a)[NR], const FOp& op, const cx<R>* d, uint32_t tid, bool tok) {
  const int nr = op.npos, nt = op.ntt;
  uint32_t tsel = 0;
  for (int k = 0; k < nt; ++k) tsel = (tsel << 1) | ((tid >> op.tpos[k]) & 1u);
  const uint32_t creg = op.creg;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (((uint32_t)r & creg) != creg) continue;
    uint32_t ridx = 0;
    for (int k = 0; k < nr; ++k) ridx = (ridx << 1) | (((uint32_t)r >> op.pos[k]) & 1u);
    const cx<R>* e = d + (ridx << nt);
    cx<R> val;
    if (nt == 0) {
      val = e[0];
    } else if (nt == 1) {
      const cx<R> e0 = e[0], e1 = e[1];
      val = tsel ? e1 : e0;
    } else {
      const cx<R> e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
      val = (tsel & 2) ? ((tsel & 1) ? e3 : e2) : ((tsel & 1) ? e1 : e0);
    }
    if (tok) a[r] = mul(val, a[r]);
  }
}

template <typename R, int RB, int TB>
__global__ __launch_bounds__(1 << TB) void k_fused(cx<R>* __restrict__ v, const PassHdr* __restrict__ H,
                                                   const FOp* __restrict__ ops, const cx<R>* __restrict__ coef) {
  constexpr int NR = 1 << RB;
  constexpr int K = RB + TB;
  extern __shared__ __align__(16) char lds_raw[];
  cx<R>* lds = reinterpret_cast<cx<R>*>(lds_raw);
  const uint32_t tid = threadIdx.x;

  const uint64_t tbase = insert_zeros((uint64_t)blockIdx.x, H->sorted, K);
  int rs[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) rs[b] = H->rs0[b];
  uint64_t gthr = 0;
  uint32_t thr_loc = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const uint32_t bit = (tid >> b) & 1u;
    gthr |= (uint64_t)bit << H->tbits[H->ts0[b]];
    thr_loc |= bit << H->ts0[b];
  }
  cx<R> a[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    uint64_t g = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b) g |= (uint64_t)((r >> b) & 1) << H->tbits[rs[b]];
    a[r] = ld(v + (tbase | gthr | g));
  }

  const int nops = H->nops;
  const FOp* P = ops + H->op0;
  const cx<R>* C = coef + H->coef0;
  for (int o = 0; o < nops; ++o) {
    const FOp& op = P[o];
    const int kind = op.kind;
    if (kind == F_RELAYOUT) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const uint32_t j = thr_loc | reg_local<RB>(r, rs);
        lds[lds_swz(j)] = a[r];
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < RB; ++b) rs[b] = op.rs[b];
      thr_loc = 0;
#pragma unroll
      for (int b = 0; b < TB; ++b) thr_loc |= ((tid >> b) & 1u) << op.ts[b];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const uint32_t j = thr_loc | reg_local<RB>(r, rs);
        a[r] = lds[lds_swz(j)];
      }
      continue;
    }
    const bool tok = (tid & op.cthr) == op.cthr;
    const cx<R>* m = C + op.coef;
    switch (kind) {
      case F_DENSE1:
        switch (op.pos[0]) {
#define QS_D1(P) \
  case P:        \
    if constexpr (P < RB) do_dense1<R, NR, P>(a, m, op.creg, tok); \
    break;
          QS_D1(0) QS_D1(1) QS_D1(2) QS_D1(3) QS_D1(4)
#undef QS_D1
        }
        break;
      case F_PERM1:
        switch (op.pos[0]) {
#define QS_P1(P) \
  case P:        \
    if constexpr (P < RB) do_perm1<R, NR, P>(a, m, op.creg, tok, op.pure); \
    break;
          QS_P1(0) QS_P1(1) QS_P1(2) QS_P1(3) QS_P1(4)
#undef QS_P1
        }
        break;
      case F_DIAG:
        do_diag<R, NR>(a, op, m, tid, tok);
        break;
      case F_DENSE2:
      case F_SWAP2: {
        const int code = op.pos[0] * 8 + op.pos[1];
        const bool sw = kind == F_SWAP2;
        switch (code) {
#define QS_D2(A, B)                                                         \
  case A * 8 + B:                                                           \
    if constexpr (A < RB) {                                                 \
      if (sw) do_swap2<R, NR, A, B>(a, m, op.creg, tok, op.pure);           \
      else do_dense2<R, NR, A, B>(a, m, op.creg, tok);                      \
    }                                                                       \
    break;
          QS_D2(1, 0) QS_D2(2, 0) QS_D2(2, 1) QS_D2(3, 0) QS_D2(3, 1) QS_D2(3, 2)
          QS_D2(4, 0) QS_D2(4, 1) QS_D2(4, 2) QS_D2(4, 3)
#undef QS_D2
        }
        break;
      }
      case F_DENSE3: {
        const uint32_t pm = (1u << op.pos[0]) | (1u << op.pos[1]) | (1u << op.pos[2]);
        switch (pm) {
#define QS_D3(M) \
  case M:        \
    if constexpr ((M >> RB) == 0) do_densek<R, NR, 3, M>(a, m, op.creg, tok); \
    break;
          QS_D3(7u) QS_D3(11u) QS_D3(13u) QS_D3(14u) QS_D3(19u) QS_D3(21u) QS_D3(22u) QS_D3(25u) QS_D3(26u)
          QS_D3(28u)
#undef QS_D3
        }
        break;
      }
      case F_DENSE4: {
        const uint32_t pm = (1u << op.pos[0]) | (1u << op.pos[1]) | (1u << op.pos[2]) | (1u << op.pos[3]);
        switch (pm) {
#define QS_D4(M) \
  case M:        \
    if constexpr ((M >> RB) == 0) do_densek<R, NR, 4, M>(a, m, op.creg, tok); \
    break;
          QS_D4(15u) QS_D4(23u) QS_D4(27u) QS_D4(29u) QS_D4(30u)
#undef QS_D4
        }
        break;
      }
      default:
        break;
    }
  }

  // store in the final layout (whose thread bits include the low contiguous bits)
  gthr = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) gthr |= (uint64_t)((tid >> b) & 1u) << H->tbits[H->ts1[b]];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    uint64_t g = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b) g |= (uint64_t)((r >> b) & 1) << H->tbits[H->rs1[b]];
    st(v + (tbase | gthr | g), a[r]);
  }
}

// ---- host planner --------------------------------------------------------------

struct Config {
  int RB, TB, L;
};

static bool gate_diag_like(const qs_gate& g) { return g.kind == QS_GATE_DIAG || g.kind == QS_GATE_PHASE; }

static uint64_t gate_mask(const qs_gate& g) {
  uint64_t m = 0;
  for (int i = 0; i < g.ncontrols + g.ntargets; ++i) m |= 1ull << g.bits[i];
  return m;
}

static bool fusable(const qs_gate& g, const Config& c) {
  const int K = c.RB + c.TB;
  if (g.ncontrols + g.ntargets > K) return false;
  if (g.kind == QS_GATE_DENSE && g.ntargets > std::min(4, c.RB)) return false;
  if (gate_diag_like(g) && g.ntargets > std::min(4, c.RB + 2)) return false;
  return true;
}

// Complex payload of a gate converted to the op form: returns entries in op order.
struct OpBuild {
  FOp op;
  std::vector<double> coef;  // (re, im) pairs
};

static void permute_matrix(const std::vector<double>& M, int k, const int* order, std::vector<double>& out) {
  // order[i] = index of the original target that becomes new target i (new target 0 = MSB)
  const int D = 1 << k;
  out.assign(2 * D * D, 0.0);
  auto remap = [&](int idx) {
    int o = 0;  // original index
    for (int i = 0; i < k; ++i) {
      const int bit = (idx >> (k - 1 - i)) & 1;
      o |= bit << (k - 1 - order[i]);
    }
    return o;
  };
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      const int oi = remap(i), oj = remap(j);
      out[2 * (i * D + j)] = M[2 * (oi * D + oj)];
      out[2 * (i * D + j) + 1] = M[2 * (oi * D + oj) + 1];
    }
}

struct Layout {
  std::vector<int> rs;  // register bit i -> tile-local bit
  std::vector<int> ts;  // thread bit i -> tile-local bit
};

static Layout make_layout(const std::vector<int>& regs, int K) {
  Layout L;
  L.rs = regs;
  std::vector<char> isreg(K, 0);
  for (int b : regs) isreg[b] = 1;
  for (int b = 0; b < K; ++b)
    if (!isreg[b]) L.ts.push_back(b);
  return L;
}

static int index_in(const std::vector<int>& v, int x) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == x) return (int)i;
  return -1;
}

// Tile-local bits that must be register bits for gate g under the op encoding.
static std::vector<int> needed_regs(const qs_gate& g, const std::vector<int>& loc, const Layout& cur) {
  std::vector<int> need;
  const int c = g.ncontrols, t = g.ntargets;
  if (gate_diag_like(g)) {
    int outside = 0;
    for (int i = 0; i < t; ++i)
      if (index_in(cur.rs, loc[c + i]) < 0) ++outside;
    if (outside <= 2) return need;
    for (int i = 0; i < t; ++i) need.push_back(loc[c + i]);
    return need;
  }
  for (int i = 0; i < t; ++i) need.push_back(loc[c + i]);
  return need;
}

template <typename R>
static int build_op(const qs_gate& g, const std::vector<int>& loc, const Layout& cur, OpBuild& ob) {
  FOp& op = ob.op;
  std::memset(&op, 0, sizeof(op));
  const int c = g.ncontrols, t = g.ntargets;
  // controls
  for (int i = 0; i < c; ++i) {
    const int b = loc[i];
    const int ri = index_in(cur.rs, b);
    if (ri >= 0) op.creg |= 1u << ri;
    else op.cthr |= 1u << index_in(cur.ts, b);
  }
  ob.coef.clear();
  if (g.kind == QS_GATE_PHASE || gate_diag_like(g)) {
    // Diagonal over the target bits.  PHASE = diag(1, .., 1, v): the last target
    // becomes a control-like selector: entries 1 except the all-ones one.
    const int D = 1 << t;
    std::vector<double> dg(2 * D);
    for (int m = 0; m < D; ++m) {
      if (g.kind == QS_GATE_PHASE) {
        dg[2 * m] = (m == D - 1) ? g.m[0] : 1.0;
        dg[2 * m + 1] = (m == D - 1) ? g.m[1] : 0.0;
      } else {
        dg[2 * m] = g.m[2 * m];
        dg[2 * m + 1] = g.m[2 * m + 1];
      }
    }
    // split targets into register and thread bits
    std::vector<int> rt, tt;  // original target indices
    for (int i = 0; i < t; ++i) {
      if (index_in(cur.rs, loc[c + i]) >= 0) rt.push_back(i);
      else tt.push_back(i);
    }
    if (tt.size() > 2) return QS_ERR_UNSUPPORTED;
    op.kind = F_DIAG;
    op.npos = (int)rt.size();
    for (size_t k = 0; k < rt.size(); ++k) op.pos[k] = index_in(cur.rs, loc[c + rt[k]]);
    op.ntt = (int)tt.size();
    for (size_t k = 0; k < tt.size(); ++k) op.tpos[k] = index_in(cur.ts, loc[c + tt[k]]);
    const int nr = (int)rt.size(), nt = (int)tt.size();
    ob.coef.assign(2 * D, 0.0);
    for (int ridx = 0; ridx < (1 << nr); ++ridx)
      for (int tsel = 0; tsel < (1 << nt); ++tsel) {
        int m = 0;
        for (int k = 0; k < nr; ++k)
          if ((ridx >> (nr - 1 - k)) & 1) m |= 1 << (t - 1 - rt[k]);
        for (int k = 0; k < nt; ++k)
          if ((tsel >> (nt - 1 - k)) & 1) m |= 1 << (t - 1 - tt[k]);
        const int e = (ridx << nt) | tsel;
        ob.coef[2 * e] = dg[2 * m];
        ob.coef[2 * e + 1] = dg[2 * m + 1];
      }
    return QS_OK;
  }
  // non-diagonal: all targets are register bits
  std::vector<int> rp(t);
  for (int i = 0; i < t; ++i) {
    rp[i] = index_in(cur.rs, loc[c + i]);
    if (rp[i] < 0) return QS_ERR_UNSUPPORTED;
  }
  if (g.kind == QS_GATE_XPERM) {
    op.kind = F_PERM1;
    op.npos = 1;
    op.pos[0] = rp[0];
    op.pure = (g.m[0] == 1.0 && g.m[1] == 0.0 && g.m[2] == 1.0 && g.m[3] == 0.0);
    ob.coef.assign(g.m, g.m + 4);
    return QS_OK;
  }
  if (g.kind == QS_GATE_SWAP) {
    op.kind = F_SWAP2;
    op.npos = 2;
    double m0r = g.m[0], m0i = g.m[1], m1r = g.m[2], m1i = g.m[3];
    int p0 = rp[0], p1 = rp[1];
    if (p0 < p1) {  // canonical: pos[0] > pos[1]; exchanging target roles swaps m0/m1
      std::swap(p0, p1);
      std::swap(m0r, m1r);
      std::swap(m0i, m1i);
    }
    op.pos[0] = p0;
    op.pos[1] = p1;
    op.pure = (m0r == 1.0 && m0i == 0.0 && m1r == 1.0 && m1i == 0.0);
    ob.coef = {m0r, m0i, m1r, m1i};
    return QS_OK;
  }
  // dense
  const int D = 1 << t;
  std::vector<double> M(2 * D * D);
  const double* src = (t <= 2) ? g.m : g.big;
  for (int i = 0; i < 2 * D * D; ++i) M[i] = src[i];
  // order targets by descending register position (MSB first)
  std::vector<int> order(t);
  for (int i = 0; i < t; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return rp[x] > rp[y]; });
  std::vector<double> Mp;
  permute_matrix(M, t, order.data(), Mp);
  op.npos = t;
  for (int i = 0; i < t; ++i) op.pos[i] = rp[order[i]];
  op.kind = (t == 1) ? F_DENSE1 : (t == 2) ? F_DENSE2 : (t == 3) ? F_DENSE3 : F_DENSE4;
  ob.coef = Mp;
  return QS_OK;
}

struct PassPlan {
  bool fused;
  int single;                      // gate index when !fused
  std::vector<int> gates;          // gate indices when fused
  std::vector<int> tbits;          // tile-local -> global
};

static void plan_groups(const qs_gate* gates, int ngates, int n, const Config& c, std::vector<PassPlan>& plans) {
  const int K = c.RB + c.TB;
  std::vector<int> remaining(ngates);
  for (int i = 0; i < ngates; ++i) remaining[i] = i;
  const uint64_t low = (1ull << c.L) - 1;
  while (!remaining.empty()) {
    PassPlan p;
    p.fused = true;
    p.single = -1;
    std::vector<int> rest;
    size_t i = 0;
    // A gate the tile kernel cannot take at the head of the list runs on its own.
    if (!fusable(gates[remaining[0]], c)) {
      p.fused = false;
      p.single = remaining[0];
      i = 1;
    } else {
      uint64_t S = low, blocked = 0;
      for (; i < remaining.size(); ++i) {
        const int gi = remaining[i];
        const qs_gate& g = gates[gi];
        const uint64_t gm = gate_mask(g);
        if ((gm & blocked) || !fusable(g, c) || __builtin_popcountll(S | gm) > K) {
          blocked |= gm;  // everything sharing a bit with a deferred gate must stay behind it
          rest.push_back(gi);
          continue;
        }
        S |= gm;
        p.gates.push_back(gi);
      }
      // fill the tile to K bits: prefer bits the deferred gates use first
      for (size_t j = 0; j < rest.size() && __builtin_popcountll(S) < K; ++j) {
        uint64_t gm = gate_mask(gates[rest[j]]);
        while (gm && __builtin_popcountll(S) < K) {
          S |= 1ull << __builtin_ctzll(gm);
          gm &= gm - 1;
        }
      }
      for (int b = 0; b < n && __builtin_popcountll(S) < K; ++b) S |= 1ull << b;
      for (int b = 0; b < n; ++b)
        if ((S >> b) & 1) p.tbits.push_back(b);  // ascending; the first L are the low bits
    }
    for (; i < remaining.size(); ++i) rest.push_back(remaining[i]);
    plans.push_back(std::move(p));
    remaining.swap(rest);
  }
}

// Choose the register set for a relayout: `need` first, then bits needed soonest.
static std::vector<int> choose_regs(const std::vector<int>& need, const std::vector<int>& lookahead,
                                    const std::vector<int>& cur, int RB, int K, int L) {
  std::vector<int> regs = need;
  auto has = [&](int b) { return std::find(regs.begin(), regs.end(), b) != regs.end(); };
  for (int b : lookahead) {
    if ((int)regs.size() >= RB) break;
    if (!has(b)) regs.push_back(b);
  }
  for (int b : cur) {
    if ((int)regs.size() >= RB) break;
    if (!has(b) && b >= L) regs.push_back(b);
  }
  for (int b = K - 1; b >= 0 && (int)regs.size() < RB; --b)
    if (!has(b)) regs.push_back(b);
  return regs;
}

template <typename R>
static int schedule_pass(const qs_gate* gates, const PassPlan& p, const Config& c, PassHdr& H,
                         std::vector<FOp>& ops, std::vector<double>& coef) {
  const int K = c.RB + c.TB;
  std::memset(&H, 0, sizeof(H));
  std::vector<int> g2l(64, -1);
  for (int j = 0; j < K; ++j) {
    H.tbits[j] = p.tbits[j];
    H.sorted[j] = p.tbits[j];
    g2l[p.tbits[j]] = j;
  }
  // per-gate local bits and lookahead list of needed (non-diagonal target) bits
  std::vector<std::vector<int>> loc(p.gates.size());
  for (size_t k = 0; k < p.gates.size(); ++k) {
    const qs_gate& g = gates[p.gates[k]];
    for (int i = 0; i < g.ncontrols + g.ntargets; ++i) loc[k].push_back(g2l[g.bits[i]]);
  }
  auto lookahead_from = [&](size_t k0) {
    std::vector<int> la;
    for (size_t k = k0; k < p.gates.size() && la.size() < 16; ++k) {
      const qs_gate& g = gates[p.gates[k]];
      if (gate_diag_like(g)) continue;
      for (int i = 0; i < g.ntargets; ++i) {
        const int b = loc[k][g.ncontrols + i];
        if (std::find(la.begin(), la.end(), b) == la.end()) la.push_back(b);
      }
    }
    return la;
  };
  // initial layout: registers on high tile bits needed first
  std::vector<int> init_need;
  for (int b : lookahead_from(0)) {
    if ((int)init_need.size() >= c.RB) break;
    if (b >= c.L) init_need.push_back(b);
  }
  std::vector<int> regs = choose_regs(init_need, {}, {}, c.RB, K, c.L);
  // ensure no low bit was picked by the fallback (K-1 downwards never reaches < L when K-L >= RB)
  Layout cur = make_layout(regs, K);
  for (int b = 0; b < c.RB; ++b) H.rs0[b] = cur.rs[b];
  for (int b = 0; b < c.TB; ++b) H.ts0[b] = cur.ts[b];
  const int op_start = (int)ops.size();
  const int coef_start = (int)(coef.size() / 2);
  for (size_t k = 0; k < p.gates.size(); ++k) {
    const qs_gate& g = gates[p.gates[k]];
    std::vector<int> need = needed_regs(g, loc[k], cur);
    bool ok = true;
    for (int b : need)
      if (index_in(cur.rs, b) < 0) ok = false;
    if (!ok) {
      std::vector<int> la = lookahead_from(k + 1);
      std::vector<int> nr = choose_regs(need, la, cur.rs, c.RB, K, c.L);
      cur = make_layout(nr, K);
      FOp rl;
      std::memset(&rl, 0, sizeof(rl));
      rl.kind = F_RELAYOUT;
      for (int b = 0; b < c.RB; ++b) rl.rs[b] = cur.rs[b];
      for (int b = 0; b < c.TB; ++b) rl.ts[b] = cur.ts[b];
      ops.push_back(rl);
      H.uses_lds = 1;
    }
    OpBuild ob;
    int rc = build_op<R>(g, loc[k], cur, ob);
    if (rc) {
      set_error("fusion planner produced an op the tile kernel cannot encode");
      return rc;
    }
    ob.op.coef = (int)(coef.size() / 2) - coef_start;
    coef.insert(coef.end(), ob.coef.begin(), ob.coef.end());
    ops.push_back(ob.op);
  }
  // final layout must keep the low L bits on threads
  bool low_in_regs = false;
  for (int b : cur.rs)
    if (b < c.L) low_in_regs = true;
  if (low_in_regs) {
    std::vector<int> keep;
    for (int b : cur.rs)
      if (b >= c.L) keep.push_back(b);
    std::vector<int> nr = choose_regs(keep, {}, {}, c.RB, K, c.L);
    cur = make_layout(nr, K);
    FOp rl;
    std::memset(&rl, 0, sizeof(rl));
    rl.kind = F_RELAYOUT;
    for (int b = 0; b < c.RB; ++b) rl.rs[b] = cur.rs[b];
    for (int b = 0; b < c.TB; ++b) rl.ts[b] = cur.ts[b];
    ops.push_back(rl);
    H.uses_lds = 1;
  }
  for (int b = 0; b < c.RB; ++b) H.rs1[b] = cur.rs[b];
  for (int b = 0; b < c.TB; ++b) H.ts1[b] = cur.ts[b];
  H.op0 = op_start;
  H.nops = (int)ops.size() - op_start;
  H.coef0 = coef_start;
  return QS_OK;
}

template <typename R, int RB, int TB>
static int run_fused_t(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* gates, int ngates, int L, int* passes,
                       hipStream_t s) {
  const Config c{RB, TB, L};
  const int K = RB + TB;
  std::vector<PassPlan> plans;
  plan_groups(gates, ngates, n, c, plans);
  std::vector<PassHdr> hdrs;
  std::vector<FOp> ops;
  std::vector<double> coef;
  std::vector<int> hdr_of(plans.size(), -1);
  for (size_t i = 0; i < plans.size(); ++i) {
    if (!plans[i].fused) continue;
    PassHdr H;
    int rc = schedule_pass<R>(gates, plans[i], c, H, ops, coef);
    if (rc) return rc;
    hdr_of[i] = (int)hdrs.size();
    hdrs.push_back(H);
  }
  if (const char* dbg = getenv("QS_FUSION_DEBUG")) {
    if (dbg[0] == '1') {
      int nrel = 0, kinds[8] = {0};
      for (const FOp& o : ops) {
        if (o.kind == F_RELAYOUT) ++nrel;
        if (o.kind >= 0 && o.kind < 8) ++kinds[o.kind];
      }
      int nf = 0, lds = 0;
      for (auto& h : hdrs) {
        ++nf;
        lds += h.uses_lds;
      }
      fprintf(stderr,
              "[qsim fusion] n=%d K=%d L=%d gates=%d passes=%zu fused=%d with_lds=%d ops=%zu relayouts=%d "
              "dense1=%d perm1=%d diag=%d dense2=%d swap2=%d dense3=%d dense4=%d\n",
              n, K, L, ngates, plans.size(), nf, lds, ops.size(), nrel, kinds[F_DENSE1], kinds[F_PERM1],
              kinds[F_DIAG], kinds[F_DENSE2], kinds[F_SWAP2], kinds[F_DENSE3], kinds[F_DENSE4]);
    }
  }
  // upload program: [hdrs][ops][coef in R]
  const uint64_t hb = ((hdrs.size() * sizeof(PassHdr)) + 255) & ~255ull;
  const uint64_t ob = ((ops.size() * sizeof(FOp)) + 255) & ~255ull;
  const uint64_t ncoef = coef.size() / 2;
  const uint64_t cb = ((ncoef * sizeof(cx<R>)) + 255) & ~255ull;
  int rc = ensure_prog(ctx, hb + ob + cb + 256);
  if (rc) return rc;
  std::vector<char> host(hb + ob + cb, 0);
  if (!hdrs.empty()) std::memcpy(host.data(), hdrs.data(), hdrs.size() * sizeof(PassHdr));
  if (!ops.empty()) std::memcpy(host.data() + hb, ops.data(), ops.size() * sizeof(FOp));
  cx<R>* hc = reinterpret_cast<cx<R>*>(host.data() + hb + ob);
  for (uint64_t i = 0; i < ncoef; ++i) hc[i] = mk<R>((R)coef[2 * i], (R)coef[2 * i + 1]);
  char* dprog = (char*)ctx->prog;
  if (!host.empty()) QS_CHECK_HIP(hipMemcpyAsync(dprog, host.data(), host.size(), hipMemcpyHostToDevice, s));
  const PassHdr* dH = (const PassHdr*)dprog;
  const FOp* dO = (const FOp*)(dprog + hb);
  const cx<R>* dC = (const cx<R>*)(dprog + hb + ob);
  const uint64_t tiles = 1ull << (n - K);
  const size_t lds_bytes = (size_t)(1u << K) * sizeof(cx<R>);
  int npass = 0;
  for (size_t i = 0; i < plans.size(); ++i) {
    if (!plans[i].fused) {
      rc = prof_begin(ctx, PROF_GATE, s);
      if (rc) return rc;
      rc = apply_gate(ctx, v, n, sizeof(R) == 8 ? QS_C128 : QS_C64, &gates[plans[i].single], s);
      if (rc) return rc;
      rc = prof_end(ctx, s);
      if (rc) return rc;
      ++npass;
      continue;
    }
    const int h = hdr_of[i];
    const size_t lds = hdrs[h].uses_lds ? lds_bytes : 0;
    rc = prof_begin(ctx, PROF_FUSED, s);
    if (rc) return rc;
    hipLaunchKernelGGL((k_fused<R, RB, TB>), dim3((unsigned)tiles), dim3(1 << TB), lds, s, v, dH + h, dO, dC);
    QS_CHECK_HIP(hipGetLastError());
    rc = prof_end(ctx, s);
    if (rc) return rc;
    ++npass;
  }
  if (passes) *passes = npass;
  // the host staging buffer must outlive the async upload
  QS_CHECK_HIP(hipStreamSynchronize(s));
  return QS_OK;
}

template <typename R, int RB, int TB> static void set_lds_attr() {
  static bool done = false;
  if (done) return;
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                      (int)((1u << (RB + TB)) * sizeof(cx<R>)));
  done = true;
}

int apply_gates_fused(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, int* passes,
                      hipStream_t s) {
  // Tile shapes: complex128 K = 12 (4 register bits x 256 threads, 64 KiB LDS);
  // complex64 K = 13 (4 x 512, 64 KiB).
  const int K = (dtype == QS_C128) ? 12 : 13;
  int L = ctx->low_bits ? ctx->low_bits : 4;
  if (n < K + 1 || ngates == 0) {
    // small registers: every pass is latency-bound anyway; one kernel per gate
    for (int i = 0; i < ngates; ++i) {
      int rc = apply_gate(ctx, vec, n, dtype, &gates[i], s);
      if (rc) return rc;
    }
    if (passes) *passes = ngates;
    return QS_OK;
  }
  if (dtype == QS_C128) {
    set_lds_attr<double, 4, 8>();
    return run_fused_t<double, 4, 8>(ctx, (cx<double>*)vec, n, gates, ngates, L, passes, s);
  }
  set_lds_attr<float, 4, 9>();
  return run_fused_t<float, 4, 9>(ctx, (cx<float>*)vec, n, gates, ngates, L, passes, s);
}

}  // namespace qs
