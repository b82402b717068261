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
// Complex coefficients one pass may carry (staged in LDS next to the relayout buffer).
constexpr int FMAX_PASS_COEF = 1024;

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
  int32_t has_dk;        // pass contains 3- or 4-target dense ops (kernel variant without prefetch)
  int32_t ncoef;         // complex coefficients of this pass (staged in LDS by the prefetching variant)
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

// Workgroup barrier for LDS traffic only: waits for this wave's LDS (and scalar)
// operations, not for outstanding global loads, so prefetched tile loads stay
// in flight across relayouts.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Non-temporal amplitude loads / stores for the streaming tile traffic.
template <typename R> __device__ __forceinline__ cx<R> ld_nt(const cx<R>* p) {
  return cx<R>{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}
template <typename R> __device__ __forceinline__ void st_nt(cx<R>* p, cx<R> a) {
  __builtin_nontemporal_store(a.re, &p->re);
  __builtin_nontemporal_store(a.im, &p->im);
}

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
        // one matrix row at a time keeps scalar-register pressure low
        cx<R> acc = mul(m[4 * i], x[0]);
#pragma unroll
        for (int j = 1; j < 4; ++j) acc = cfma(acc, m[4 * i + j], x[j]);
        a[idx[i]] = acc;
        __builtin_amdgcn_sched_barrier(0);
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
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// Diagonal op: the host resolved every register-bit dependence at plan time, so
// the payload is a table T[sel * NR + r] (sel = up to two thread-index bits, MSB
// first) and `creg` is the mask of registers whose entry is not 1 for some sel.
template <typename R, int NR, int NS>
__device__ __forceinline__ void do_diag_t(cx<R> (&a)[NR], const cx<R>* T, uint32_t rmask, uint32_t sel) {
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (!((rmask >> r) & 1u)) continue;
    cx<R> val;
    if constexpr (NS == 0) {
      val = T[r];
    } else if constexpr (NS == 1) {
      const cx<R> e0 = T[r], e1 = T[NR + r];
      val = sel ? e1 : e0;
    } else {
      const cx<R> e0 = T[r], e1 = T[NR + r], e2 = T[2 * NR + r], e3 = T[3 * NR + r];
      val = (sel & 2) ? ((sel & 1) ? e3 : e2) : ((sel & 1) ? e1 : e0);
    }
    a[r] = mul(val, a[r]);
  }
}

// Diagonal gate that depends on thread bits only: its value is a per-lane
// scalar, which commutes with every in-register op (those mix registers of one
// lane), so it is folded into the lane factor `f` and applied to the registers
// only before a relayout or the store.
template <typename R>
__device__ __forceinline__ void diag_lane(cx<R>& f, const FOp& op, const cx<R>* T, uint32_t tid) {
  const int ns = op.ntt;
  uint32_t sel = 0;
  for (int k = 0; k < ns; ++k) sel = (sel << 1) | ((tid >> op.tpos[k]) & 1u);
  // component-wise selects: selecting whole structs by a lane value makes the
  // compiler index a private (scratch) array
  R re = T[0].re, im = T[0].im;
  if (ns >= 1) {
    const bool b0 = sel & 1u;
    if (ns == 1) {
      re = b0 ? T[1].re : re;
      im = b0 ? T[1].im : im;
    } else {
      const bool b1 = sel & 2u;
      const R lo_re = b0 ? T[1].re : re, lo_im = b0 ? T[1].im : im;
      const R hi_re = b0 ? T[3].re : T[2].re, hi_im = b0 ? T[3].im : T[2].im;
      re = b1 ? hi_re : lo_re;
      im = b1 ? hi_im : lo_im;
    }
  }
  f = mul(mk<R>(re, im), f);
}

template <typename R, int NR>
__device__ __forceinline__ void flush_lane(cx<R> (&a)[NR], cx<R>& f) {
#pragma unroll
  for (int r = 0; r < NR; ++r) a[r] = mul(f, a[r]);
  f = mk<R>((R)1, (R)0);
}

template <typename R, int NR>
__device__ __forceinline__ void do_diag(cx<R> (&a)[NR], const FOp& op, const cx<R>* T, uint32_t tid) {
  const int ns = op.ntt;
  uint32_t sel = 0;
  for (int k = 0; k < ns; ++k) sel = (sel << 1) | ((tid >> op.tpos[k]) & 1u);
  // general table: per-lane loads from row `sel` (the address is lane-dependent,
  // so the 16..64 table entries never occupy scalar registers)
  const cx<R>* Tl = T + sel * NR;
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if ((op.creg >> r) & 1u) a[r] = mul(ld(Tl + r), a[r]);
}

// Apply a pass's op list to the register tile `a` (load layout on entry, store
// layout on exit).  DK: include the in-register 3- and 4-target dense ops.
template <typename R, int RB, int TB, bool DK>
__device__ __forceinline__ void run_ops(cx<R> (&a)[1 << RB], cx<R>* lds, const PassHdr* __restrict__ H,
                                        const FOp* __restrict__ P, const cx<R>* __restrict__ C, uint32_t tid) {
  constexpr int NR = 1 << RB;
  int rs[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) rs[b] = H->rs0[b];
  uint32_t thr_loc = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) thr_loc |= ((tid >> b) & 1u) << H->ts0[b];
  const int nops = H->nops;
  cx<R> f = mk<R>((R)1, (R)0);  // deferred per-lane diagonal factor
  bool fdirty = false;
  for (int o = 0; o < nops; ++o) {
    const FOp& op = P[o];
    const int kind = op.kind;
    if (kind == F_RELAYOUT) {
      if (fdirty) {
        flush_lane<R, NR>(a, f);
        fdirty = false;
      }
      lds_barrier();
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const uint32_t j = thr_loc | reg_local<RB>(r, rs);
        lds[lds_swz(j)] = a[r];
      }
      lds_barrier();
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
        if (op.pure) {
          diag_lane<R>(f, op, m, tid);
          fdirty = true;
        } else {
          do_diag<R, NR>(a, op, m, tid);
        }
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
        if constexpr (DK) {
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
        }
        break;
      }
      case F_DENSE4: {
        if constexpr (DK) {
          const uint32_t pm = (1u << op.pos[0]) | (1u << op.pos[1]) | (1u << op.pos[2]) | (1u << op.pos[3]);
          switch (pm) {
#define QS_D4(M) \
  case M:        \
    if constexpr ((M >> RB) == 0) do_densek<R, NR, 4, M>(a, m, op.creg, tok); \
    break;
            QS_D4(15u) QS_D4(23u) QS_D4(27u) QS_D4(29u) QS_D4(30u)
#undef QS_D4
          }
        }
        break;
      }
      default:
        break;
    }
  }
  if (fdirty) flush_lane<R, NR>(a, f);
}

// One pass over all 2^(n-K) tiles.  PF: persistent workgroups that prefetch
// the next tile into a second register set while the current one is computed
// (grid = resident workgroups); otherwise one tile per workgroup.
template <typename R, int RB, int TB, bool DK, bool PF>
__global__ __launch_bounds__(1 << TB) void k_fused(cx<R>* __restrict__ v, const PassHdr* __restrict__ H,
                                                   const FOp* __restrict__ ops, const cx<R>* __restrict__ coef,
                                                   uint64_t ntiles) {
  constexpr int NR = 1 << RB;
  constexpr int K = RB + TB;
  extern __shared__ __align__(16) char lds_raw[];
  cx<R>* lds = reinterpret_cast<cx<R>*>(lds_raw);
  const uint32_t tid = threadIdx.x;
  const FOp* P = ops + H->op0;
  const cx<R>* C;
  if constexpr (!DK) {
    // Stage the pass coefficients in LDS once per workgroup: op payload reads are
    // then LDS reads, which never wait on the prefetched global tile loads.
    cx<R>* lc = lds + (H->uses_lds ? (1 << K) : 0);
    const cx<R>* gc = coef + H->coef0;
    const int nc = H->ncoef;
    for (int i = (int)tid; i < nc; i += (int)blockDim.x) lc[i] = gc[i];
    __syncthreads();
    C = lc;
  } else {
    C = coef + H->coef0;
  }
  // per-lane parts of the load and store addresses, per-register parts as bit masks
  uint64_t gin = 0, gout = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) {
    const uint64_t bit = (tid >> b) & 1u;
    gin |= bit << H->tbits[H->ts0[b]];
    gout |= bit << H->tbits[H->ts1[b]];
  }
  uint64_t min_[RB], mout[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    min_[b] = 1ull << H->tbits[H->rs0[b]];
    mout[b] = 1ull << H->tbits[H->rs1[b]];
  }
  cx<R> a[NR];
  [[maybe_unused]] cx<R> nx[PF ? NR : 1];
  uint64_t t = blockIdx.x;
  if constexpr (PF) {
    if (t < ntiles) {
      const uint64_t base = insert_zeros(t, H->sorted, K) | gin;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= min_[b];
        a[r] = ld_nt(v + (base | g));
      }
    }
  }
  for (; t < ntiles; t += gridDim.x) {
    if constexpr (PF) {
      const uint64_t tn = t + gridDim.x;
      if (tn < ntiles) {
        const uint64_t base = insert_zeros(tn, H->sorted, K) | gin;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          uint64_t g = 0;
#pragma unroll
          for (int b = 0; b < RB; ++b)
            if ((r >> b) & 1) g |= min_[b];
          nx[r] = ld_nt(v + (base | g));
        }
      }
    } else {
      const uint64_t base = insert_zeros(t, H->sorted, K) | gin;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= min_[b];
        a[r] = ld_nt(v + (base | g));
      }
    }
    run_ops<R, RB, TB, DK>(a, lds, H, P, C, tid);
    const uint64_t base = insert_zeros(t, H->sorted, K) | gout;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      uint64_t g = 0;
#pragma unroll
      for (int b = 0; b < RB; ++b)
        if ((r >> b) & 1) g |= mout[b];
      st_nt(v + (base | g), a[r]);
    }
    if constexpr (PF) {
#pragma unroll
      for (int r = 0; r < NR; ++r) a[r] = nx[r];
    }
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

// Upper bound of the coefficient entries a gate's op can need.
static int coef_cost(const qs_gate& g, int NR) {
  if (gate_diag_like(g)) return 4 * NR;
  if (g.kind == QS_GATE_XPERM || g.kind == QS_GATE_SWAP) return 2;
  return 1 << (2 * g.ntargets);
}

static bool fusable(const qs_gate& g, const Config& c) {
  const int K = c.RB + c.TB;
  if (g.ncontrols + g.ntargets > K) return false;
  if (g.kind == QS_GATE_DENSE && g.ntargets > std::min(4, c.RB)) return false;
  if (gate_diag_like(g) && g.ncontrols + g.ntargets > c.RB + 2) return false;
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
    // at most two of the gate's bits (controls or targets) may stay thread bits
    int outside = 0;
    for (int i = 0; i < c + t; ++i)
      if (index_in(cur.rs, loc[i]) < 0) ++outside;
    if (outside <= 2) return need;
    for (int i = 0; i < t; ++i) need.push_back(loc[c + i]);
    for (int i = 0; i < c && (int)need.size() < c + t - 2; ++i) need.push_back(loc[i]);
    return need;
  }
  for (int i = 0; i < t; ++i) need.push_back(loc[c + i]);
  return need;
}

template <typename R>
static int build_op(const qs_gate& g, const std::vector<int>& loc, const Layout& cur, int NR, OpBuild& ob) {
  FOp& op = ob.op;
  std::memset(&op, 0, sizeof(op));
  const int c = g.ncontrols, t = g.ntargets;
  ob.coef.clear();
  if (gate_diag_like(g)) {
    // Diagonal gate (controls included): every gate bit that is a register bit is
    // resolved here per register; the (<= 2) thread bits select among 2^ns
    // values per register at run time.  Table T[sel * NR + r]; op.creg = mask
    // of registers that are not multiplied by 1.
    std::vector<int> selb;  // gate-bit indices living on thread bits
    for (int i = 0; i < c + t; ++i)
      if (index_in(cur.rs, loc[i]) < 0) selb.push_back(i);
    if (selb.size() > 2) return QS_ERR_UNSUPPORTED;
    const int ns = (int)selb.size();
    op.kind = F_DIAG;
    op.ntt = ns;
    for (int k = 0; k < ns; ++k) op.tpos[k] = index_in(cur.ts, loc[selb[k]]);
    const int D = 1 << t;
    ob.coef.assign(2 * (size_t)NR * (1 << ns), 0.0);
    uint32_t mask = 0;
    for (int sel = 0; sel < (1 << ns); ++sel)
      for (int r = 0; r < NR; ++r) {
        bool ctrl_ok = true;
        int m = 0;
        for (int i = 0; i < c + t; ++i) {
          int bit;
          const int si = index_in(selb, i);
          if (si >= 0) bit = (sel >> (ns - 1 - si)) & 1;
          else bit = (r >> index_in(cur.rs, loc[i])) & 1;
          if (i < c) {
            if (!bit) ctrl_ok = false;
          } else if (bit) {
            m |= 1 << (t - 1 - (i - c));
          }
        }
        double re = 1.0, im = 0.0;
        if (ctrl_ok) {
          if (g.kind == QS_GATE_PHASE) {
            if (m == D - 1) {
              re = g.m[0];
              im = g.m[1];
            }
          } else {
            re = g.m[2 * m];
            im = g.m[2 * m + 1];
          }
        }
        const size_t e = (size_t)sel * NR + r;
        ob.coef[2 * e] = re;
        ob.coef[2 * e + 1] = im;
        if (re != 1.0 || im != 0.0) mask |= 1u << r;
      }
    op.creg = mask;
    return QS_OK;
  }
  // controls of a non-diagonal gate: register-space mask + lane predicate
  for (int i = 0; i < c; ++i) {
    const int b = loc[i];
    const int ri = index_in(cur.rs, b);
    if (ri >= 0) op.creg |= 1u << ri;
    else op.cthr |= 1u << index_in(cur.ts, b);
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

// ---- peephole merging of built ops ---------------------------------------------
//
// Ops act on tile-local bit sets; two ops on disjoint sets commute.  A new op is
// merged into the most recent earlier op that shares a bit with it (everything
// in between commutes with it), never across a relayout:
//   diag  x diag           -> one diagonal table (thread selector union <= 2 bits)
//   dense1 x dense1        -> product matrix (same register position and controls)
//   dense1 x diag on that register bit (no controls) -> product matrix
// Random circuits put two or more gates on the same qubit inside most passes,
// so this removes a large share of the ops.

struct Built {
  FOp op;
  std::vector<double> coef;  // (re, im) pairs
  uint32_t bits = 0;         // tile-local bits the op depends on / changes
  std::vector<int> selb;     // DIAG: tile-local thread bits selecting the value, MSB first
  Layout lay;                // layout in effect when the op runs
};

using cplx = std::pair<double, double>;
static inline cplx cm_(cplx a, cplx b) { return {a.first * b.first - a.second * b.second, a.first * b.second + a.second * b.first}; }

static void mat2_mul(const double* A, const double* B, double* out) {  // out = A * B (2x2 complex)
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      double re = 0, im = 0;
      for (int k = 0; k < 2; ++k) {
        const double ar = A[2 * (i * 2 + k)], ai = A[2 * (i * 2 + k) + 1];
        const double br = B[2 * (k * 2 + j)], bi = B[2 * (k * 2 + j) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      out[2 * (i * 2 + j)] = re;
      out[2 * (i * 2 + j) + 1] = im;
    }
}

// DIAG op confined to one register bit: returns the register position and (d0, d1), else -1.
static int diag_single_reg(const Built& b, int NR, double* d2) {
  if (b.op.kind != F_DIAG || !b.selb.empty() || __builtin_popcount(b.bits) != 1) return -1;
  const int tb = __builtin_ctz(b.bits);
  const int p = index_in(b.lay.rs, tb);
  if (p < 0) return -1;
  for (int r = 0; r < NR; ++r) {
    const int r0 = r & (1 << p);
    if (b.coef[2 * r] != b.coef[2 * r0] || b.coef[2 * r + 1] != b.coef[2 * r0 + 1]) return -1;
  }
  d2[0] = b.coef[0];
  d2[1] = b.coef[1];
  d2[2] = b.coef[2 * (1 << p)];
  d2[3] = b.coef[2 * (1 << p) + 1];
  return p;
}

static void diag_to_dense1(const double* d2, double* M) {
  M[0] = d2[0]; M[1] = d2[1]; M[2] = 0; M[3] = 0;
  M[4] = 0; M[5] = 0; M[6] = d2[2]; M[7] = d2[3];
}

static bool merge_diag(Built& e, const Built& d, int NR) {
  std::vector<int> u = e.selb;
  for (int b : d.selb)
    if (index_in(u, b) < 0) u.push_back(b);
  if (u.size() > 2) return false;
  const int nu = (int)u.size();
  auto proj = [&](const std::vector<int>& sb, int usel) {
    int s = 0;
    for (size_t k = 0; k < sb.size(); ++k) {
      const int ui = index_in(u, sb[k]);
      s = (s << 1) | ((usel >> (nu - 1 - ui)) & 1);
    }
    return s;
  };
  std::vector<double> T(2 * (size_t)NR * (1 << nu));
  for (int us = 0; us < (1 << nu); ++us)
    for (int r = 0; r < NR; ++r) {
      const size_t ie = (size_t)proj(e.selb, us) * NR + r, id = (size_t)proj(d.selb, us) * NR + r;
      const cplx v = cm_({e.coef[2 * ie], e.coef[2 * ie + 1]}, {d.coef[2 * id], d.coef[2 * id + 1]});
      T[2 * ((size_t)us * NR + r)] = v.first;
      T[2 * ((size_t)us * NR + r) + 1] = v.second;
    }
  e.coef.swap(T);
  e.selb = u;
  e.bits |= d.bits;
  return true;
}

static bool try_merge(std::vector<Built>& out, Built& d, const Layout& cur, int NR) {
  (void)cur;
  if (d.op.kind != F_DIAG && d.op.kind != F_DENSE1) return false;
  int k = (int)out.size() - 1;
  for (; k >= 0; --k) {
    if (out[k].op.kind == F_RELAYOUT) return false;
    if (out[k].bits & d.bits) break;
  }
  if (k < 0) return false;
  Built& e = out[k];
  if (d.op.kind == F_DIAG && e.op.kind == F_DIAG) return merge_diag(e, d, NR);
  double d2[4], M[8];
  if (d.op.kind == F_DENSE1 && e.op.kind == F_DENSE1 && e.op.pos[0] == d.op.pos[0] && e.op.creg == d.op.creg &&
      e.op.cthr == d.op.cthr) {
    mat2_mul(d.coef.data(), e.coef.data(), M);
    e.coef.assign(M, M + 8);
    return true;
  }
  if (d.op.kind == F_DENSE1 && d.op.creg == 0 && d.op.cthr == 0) {
    const int p = diag_single_reg(e, NR, d2);
    if (p == d.op.pos[0]) {
      double D[8];
      diag_to_dense1(d2, D);
      mat2_mul(d.coef.data(), D, M);
      Built nb = d;
      nb.coef.assign(M, M + 8);
      out[k] = std::move(nb);
      return true;
    }
    return false;
  }
  if (d.op.kind == F_DIAG && e.op.kind == F_DENSE1 && e.op.creg == 0 && e.op.cthr == 0) {
    const int p = diag_single_reg(d, NR, d2);
    if (p == e.op.pos[0]) {
      double D[8];
      diag_to_dense1(d2, D);
      mat2_mul(D, e.coef.data(), M);
      e.coef.assign(M, M + 8);
      return true;
    }
  }
  return false;
}

// Final encoding of a DIAG op: tpos from the op's layout; register-independent
// tables collapse to one value per selector (op.pure = 1, lane mode).
static void finalize_diag(Built& b, int NR) {
  const int ns = (int)b.selb.size();
  b.op.ntt = ns;
  for (int k = 0; k < ns; ++k) b.op.tpos[k] = index_in(b.lay.ts, b.selb[k]);
  bool lane = true;
  uint32_t mask = 0;
  for (int s = 0; s < (1 << ns); ++s)
    for (int r = 0; r < NR; ++r) {
      const size_t i = (size_t)s * NR + r, i0 = (size_t)s * NR;
      if (b.coef[2 * i] != b.coef[2 * i0] || b.coef[2 * i + 1] != b.coef[2 * i0 + 1]) lane = false;
      if (b.coef[2 * i] != 1.0 || b.coef[2 * i + 1] != 0.0) mask |= 1u << r;
    }
  if (lane) {
    std::vector<double> T(2 * (1 << ns));
    for (int s = 0; s < (1 << ns); ++s) {
      T[2 * s] = b.coef[2 * (size_t)s * NR];
      T[2 * s + 1] = b.coef[2 * (size_t)s * NR + 1];
    }
    b.coef.swap(T);
    b.op.pure = 1;
    b.op.creg = mask ? (uint32_t)((1ull << NR) - 1) : 0u;
  } else {
    b.op.pure = 0;
    b.op.creg = mask;
  }
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
      int coef_budget = FMAX_PASS_COEF;  // LDS-staged payload per pass (complex entries)
      for (; i < remaining.size(); ++i) {
        const int gi = remaining[i];
        const qs_gate& g = gates[gi];
        const uint64_t gm = gate_mask(g);
        const int cost = coef_cost(g, 1 << c.RB);
        if ((gm & blocked) || !fusable(g, c) || __builtin_popcountll(S | gm) > K || cost > coef_budget) {
          blocked |= gm;  // everything sharing a bit with a deferred gate must stay behind it
          rest.push_back(gi);
          continue;
        }
        S |= gm;
        coef_budget -= cost;
        p.gates.push_back(gi);
      }
      // fill the tile to K bits with the lowest unused bits: that lengthens the
      // contiguous runs each workgroup reads and writes (DRAM page locality)
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
  const int NR = 1 << c.RB;
  std::vector<Built> built;
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
      Built rl;
      std::memset(&rl.op, 0, sizeof(rl.op));
      rl.op.kind = F_RELAYOUT;
      for (int b = 0; b < c.RB; ++b) rl.op.rs[b] = cur.rs[b];
      for (int b = 0; b < c.TB; ++b) rl.op.ts[b] = cur.ts[b];
      built.push_back(rl);
      H.uses_lds = 1;
    }
    OpBuild ob;
    int rc = build_op<R>(g, loc[k], cur, NR, ob);
    if (rc) {
      set_error("fusion planner produced an op the tile kernel cannot encode");
      return rc;
    }
    Built b;
    b.op = ob.op;
    b.coef = std::move(ob.coef);
    b.lay = cur;
    b.bits = 0;
    for (int i = 0; i < g.ncontrols + g.ntargets; ++i) b.bits |= 1u << loc[k][i];
    if (b.op.kind == F_DIAG)
      for (int i = 0; i < g.ncontrols + g.ntargets; ++i)
        if (index_in(cur.rs, loc[k][i]) < 0) b.selb.push_back(loc[k][i]);
    if (!try_merge(built, b, cur, NR)) built.push_back(std::move(b));
  }
  for (Built& b : built) {
    if (b.op.kind == F_DIAG) {
      finalize_diag(b, NR);
      if (b.op.creg == 0) continue;  // the merged product is the identity
    }
    if (b.op.kind == F_DENSE3 || b.op.kind == F_DENSE4) H.has_dk = 1;
    b.op.coef = (int)(coef.size() / 2) - coef_start;
    coef.insert(coef.end(), b.coef.begin(), b.coef.end());
    ops.push_back(b.op);
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
  H.ncoef = (int)(coef.size() / 2) - coef_start;
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
    const size_t lds_relayout = hdrs[h].uses_lds ? lds_bytes : 0;
    const size_t lds = hdrs[h].has_dk ? lds_relayout : lds_relayout + (size_t)hdrs[h].ncoef * sizeof(cx<R>);
    rc = prof_begin(ctx, PROF_FUSED, s);
    if (rc) return rc;
    if (hdrs[h].has_dk) {
      hipLaunchKernelGGL((k_fused<R, RB, TB, true, false>), dim3((unsigned)tiles), dim3(1 << TB), lds, s, v, dH + h,
                         dO, dC, tiles);
    } else if (ctx->fused_prefetch) {
      // persistent, prefetching: resident workgroups only
      const uint64_t grid = std::min<uint64_t>(tiles, (uint64_t)ctx->num_cus * ctx->fused_wg_per_cu);
      hipLaunchKernelGGL((k_fused<R, RB, TB, false, true>), dim3((unsigned)grid), dim3(1 << TB), lds, s, v, dH + h,
                         dO, dC, tiles);
    } else {
      hipLaunchKernelGGL((k_fused<R, RB, TB, false, false>), dim3((unsigned)tiles), dim3(1 << TB), lds, s, v, dH + h,
                         dO, dC, tiles);
    }
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
  const int bytes = (int)((1u << (RB + TB)) * sizeof(cx<R>) + FMAX_PASS_COEF * sizeof(cx<R>));
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, true, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, false, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, false, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
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
