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
// threads holds NR = 2^RB amplitudes.  A "layout" names which tile bit each
// register-index bit and each thread-index bit ("slot") carries.  A gate whose
// non-diagonal targets are register bits is applied entirely in registers;
// controls and up to two diagonal bits may sit on thread bits (lane predicates
// / per-lane tables).  When a gate needs a thread bit in registers the tile is
// re-laid-out through LDS (one ds write + one ds read per amplitude) — the
// only LDS traffic besides the pass coefficients.  Loads/stores use a layout
// whose thread bits include the L low bits, so every wavefront touches
// contiguous 2^L-amplitude runs.
//
// Gates that only relabel the amplitudes cost no instructions:
//  * an uncontrolled SWAP exchanges which tile bit two slots carry (the layout
//    changes, the registers do not);
//  * an uncontrolled X / Y flips the logical value of one slot: the planner
//    conjugates later ops by X on that slot, and the flip is folded into the
//    next LDS relayout address or the store address (one XOR per tile).
//
// The planner (plan_groups) is greedy with commutation by disjoint support:
// scanning the remaining gates in order, a gate joins the current pass if its
// bits fit in the tile and it touches no bit of a gate already deferred.
#include "qsim_common.hpp"
#include "qsim_fused_plan.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace qs {

// ---- device helpers ------------------------------------------------------------


// XOR of the register-slot offsets selected by the bits of r (compile-time r, uniform offsets).
template <int RB> __device__ __forceinline__ uint32_t reg_off(int r, const uint32_t (&lr)[RB]) {
  uint32_t o = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((r >> b) & 1) o ^= lr[b];
  return o;
}

template <int TB> __device__ __forceinline__ uint32_t thr_off(const uint32_t* lt, uint32_t tid) {
  uint32_t o = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) o ^= ((tid >> b) & 1u) ? lt[b] : 0u;
  return o;
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

template <typename R> __device__ __forceinline__ void st_lds(char* lds, uint32_t off, cx<R> a) {
  st(reinterpret_cast<cx<R>*>(lds + off), a);
}
template <typename R> __device__ __forceinline__ cx<R> ld_lds(const char* lds, uint32_t off) {
  return ld(reinterpret_cast<const cx<R>*>(lds + off));
}

// Every op below writes all the registers it may touch without branching on
// data-dependent conditions: controls select coefficients (the identity where a
// control fails) or select between old and new values.  Branches per register
// or lane-predicated regions make the compiler keep a second copy of the whole
// register tile and move it back after every op.
//
// CM (control mode): 0 no controls, 1 register-bit controls only (uniform per
// register pair), 2 thread-bit controls only (per lane), 3 both.

template <typename R> __device__ __forceinline__ cx<R> sel(bool c, cx<R> x, cx<R> y) {
  return cx<R>{c ? x.re : y.re, c ? x.im : y.im};
}
template <typename R> __device__ __forceinline__ cx<R> one() { return mk<R>((R)1, (R)0); }
template <typename R> __device__ __forceinline__ cx<R> zero() { return mk<R>((R)0, (R)0); }

// Uncontrolled dense1 with a real matrix (H, Ry, their products): half the products.
template <typename R, int NR, int P>
__device__ __forceinline__ void do_dense1_real(cx<R> (&a)[NR], const cx<R>* m) {
  const R m00 = m[0].re, m01 = m[1].re, m10 = m[2].re, m11 = m[3].re;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & (1 << P)) continue;
    const int r1 = r | (1 << P);
    const cx<R> x0 = a[r], x1 = a[r1];
    a[r] = mk<R>(rfma(m01, x1.re, m00 * x0.re), rfma(m01, x1.im, m00 * x0.im));
    a[r1] = mk<R>(rfma(m11, x1.re, m10 * x0.re), rfma(m11, x1.im, m10 * x0.im));
  }
}

template <typename R, int NR, int P, int CM>
__device__ __forceinline__ void do_dense1(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, uint32_t cregv, bool tok) {
  cx<R> m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
  if constexpr (CM >= 2) {  // lanes whose thread-bit controls fail apply the identity
    m00 = sel(tok, m00, one<R>());
    m01 = sel(tok, m01, zero<R>());
    m10 = sel(tok, m10, zero<R>());
    m11 = sel(tok, m11, one<R>());
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & (1 << P)) continue;
    cx<R> n00 = m00, n01 = m01, n10 = m10, n11 = m11;
    if constexpr (CM == 1 || CM == 3) {
      const bool on = ((uint32_t)r & creg) == cregv;
      n00 = sel(on, n00, one<R>());
      n01 = sel(on, n01, zero<R>());
      n10 = sel(on, n10, zero<R>());
      n11 = sel(on, n11, one<R>());
    }
    const int r1 = r | (1 << P);
    const cx<R> x0 = a[r], x1 = a[r1];
    a[r] = cfma(mul(n00, x0), n01, x1);
    a[r1] = cfma(mul(n10, x0), n11, x1);
  }
}

// a0' = m0 a1, a1' = m1 a0 (pure: plain exchange) where the controls hold.
// PURE: plain exchange (X, CX, ...): no phase products at all
template <typename R, int NR, int P, int CM, bool PURE>
__device__ __forceinline__ void do_perm1(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, uint32_t cregv, bool tok) {
  const cx<R> m0 = m[0], m1 = m[1];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & (1 << P)) continue;
    bool on = true;
    if constexpr (CM == 1 || CM == 3) on = ((uint32_t)r & creg) == cregv;
    if constexpr (CM >= 2) on = on && tok;
    const int r1 = r | (1 << P);
    const cx<R> x0 = a[r], x1 = a[r1];
    cx<R> y0 = x1, y1 = x0;
    if constexpr (!PURE) {
      y0 = mul(m0, x1);
      y1 = mul(m1, x0);
    }
    a[r] = sel(on, y0, x0);
    a[r1] = sel(on, y1, x1);
  }
}

// positions P0 (MSB) > P1 (LSB)
template <typename R, int NR, int P0, int P1, int CM>
__device__ __forceinline__ void do_dense2(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, uint32_t cregv, bool tok) {
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & ((1 << P0) | (1 << P1))) continue;
    bool on = true;
    if constexpr (CM == 1 || CM == 3) on = ((uint32_t)r & creg) == cregv;
    if constexpr (CM >= 2) on = on && tok;
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
      a[idx[i]] = (CM == 0) ? acc : sel(on, acc, x[i]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Diagonal + anti-diagonal two-target op (positions P0 > P1): per group
// y_s = d_s x_s + o_s x_(s^3), s = (bit P0, bit P1); REAL: real payload.
template <typename R, int NR, int P0, int P1, bool REAL>
__device__ __forceinline__ void do_xx2(cx<R> (&a)[NR], const cx<R>* m) {
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & ((1 << P0) | (1 << P1))) continue;
    const int idx[4] = {r, r | (1 << P1), r | (1 << P0), r | (1 << P0) | (1 << P1)};
    cx<R> x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = a[idx[j]];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if constexpr (REAL) {
        const R d = m[s].re, o = m[4 + s].re;
        a[idx[s]] = mk<R>(rfma(o, x[s ^ 3].re, d * x[s].re), rfma(o, x[s ^ 3].im, d * x[s].im));
      } else {
        a[idx[s]] = cfma(mul(m[s], x[s]), m[4 + s], x[s ^ 3]);
      }
    }
  }
}

template <typename R, int NR, int P0, int P1, int CM, bool PURE>
__device__ __forceinline__ void do_swap2(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, uint32_t cregv, bool tok) {
  const cx<R> m0 = m[0], m1 = m[1];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r & ((1 << P0) | (1 << P1))) continue;
    bool on = true;
    if constexpr (CM == 1 || CM == 3) on = ((uint32_t)r & creg) == cregv;
    if constexpr (CM >= 2) on = on && tok;
    const int i01 = r | (1 << P1), i10 = r | (1 << P0);
    const cx<R> x01 = a[i01], x10 = a[i10];
    cx<R> y01 = x10, y10 = x01;
    if constexpr (!PURE) {
      y01 = mul(m0, x10);
      y10 = mul(m1, x01);
    }
    a[i01] = sel(on, y01, x01);
    a[i10] = sel(on, y10, x10);
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
__device__ __forceinline__ void do_densek(cx<R> (&a)[NR], const cx<R>* m, uint32_t creg, uint32_t cregv, bool tok) {
  constexpr int D = 1 << KB;
  if (!tok) return;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if ((uint32_t)r & PM) continue;
    if (((uint32_t)r & creg) != cregv) continue;
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

// Diagonal gate that depends on thread bits only: one complex factor per lane
// (table entry `sel`, up to two thread-index bits t0 (MSB), t1), applied to every
// register at once.  One per-lane LDS read instead of selects between table
// entries: selects of loaded values become lane-divergent branches.
__device__ __forceinline__ uint32_t diag_sel(int ns, int t0, int t1, uint32_t tid) {
  const uint32_t b0 = (tid >> t0) & 1u, b1 = (tid >> t1) & 1u;
  return ns == 0 ? 0u : ns == 1 ? b0 : ((b0 << 1) | b1);
}

template <typename R, int NR>
__device__ __forceinline__ void do_diag_lane(cx<R> (&a)[NR], uint32_t sel, const cx<R>* T) {
  const cx<R> f = ld(T + sel);
#pragma unroll
  for (int r = 0; r < NR; ++r) a[r] = mul(f, a[r]);
}

// Diagonal op: the host resolved every register-bit dependence at plan time, so
// the payload is a table T[sel * NR + r] (sel = up to two thread-index bits, MSB
// first) and `creg` is the mask of registers whose entry is not 1 for some sel.
template <typename R, int NR>
__device__ __forceinline__ void do_diag(cx<R> (&a)[NR], uint32_t sel, const cx<R>* T) {
  // per-lane loads from row `sel` (the address is lane-dependent, so the 16..64
  // table entries never occupy scalar registers); every register is multiplied
  // (entries of untouched registers are 1): no per-register branches
  const cx<R>* Tl = T + sel * NR;
#pragma unroll
  for (int r = 0; r < NR; ++r) a[r] = mul(ld(Tl + r), a[r]);
}

// Relayout in two rounds through a half-tile LDS image (see the RELAYOUT
// payload): round k moves the registers whose bit P is k, writing them with the
// current layout (pending flips included) and reading the same registers back
// with the new layout.
template <typename R, int RB, int TB, int P>
__device__ __forceinline__ void relayout_half(cx<R> (&a)[1 << RB], char* lds, const uint32_t* __restrict__ x,
                                              uint32_t tid) {
  constexpr int NR = 1 << RB;
  uint32_t lw[RB], lq[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) lw[b] = x[1 + b];
  const uint32_t w = thr_off<TB>(x + 1 + RB, tid) ^ x[1 + RB + TB];
  const uint32_t* y = x + 2 + RB + TB;
#pragma unroll
  for (int b = 0; b < RB; ++b) lq[b] = y[b];
  const uint32_t rd = thr_off<TB>(y + RB, tid);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    lds_barrier();
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (((r >> P) & 1) == k) st_lds<R>(lds, w ^ reg_off<RB>(r, lw), a[r]);
    lds_barrier();
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (((r >> P) & 1) == k) a[r] = ld_lds<R>(lds, rd ^ reg_off<RB>(r, lq));
  }
}

// One op of a pass program applied to the register tile.
template <typename R, int RB, int TB, bool DK>
__device__ __forceinline__ void run_op(cx<R> (&a)[1 << RB], char* lds, const FOpW& e,
                                       const cx<R>* __restrict__ C, const cx<R>* __restrict__ G,
                                       const uint32_t* __restrict__ X, uint32_t tid) {
  constexpr int NR = 1 << RB;
  const int kind = (int)(e.w[0] & 0xff);
  const uint32_t opcoef = e.w[3];
  if (kind == F_RELAYOUT) {
    const uint32_t* x = X + opcoef;
    const uint32_t pk = x[0];
    if (pk == 0) relayout_half<R, RB, TB, 0>(a, lds, x, tid);
    else if (pk == 1) relayout_half<R, RB, TB, 1>(a, lds, x, tid);
    else if (pk == 2 || RB < 4) relayout_half<R, RB, TB, 2>(a, lds, x, tid);
    else relayout_half<R, RB, TB, (RB > 3 ? 3 : 0)>(a, lds, x, tid);
    return;
  }
  const int pure = (int)((e.w[0] >> 8) & 0xff);
  const int ntt = (int)(e.w[0] >> 24);
  const int pos0 = (int)(e.w[1] & 0xff), pos1 = (int)((e.w[1] >> 8) & 0xff);
  const int pos2 = (int)((e.w[1] >> 16) & 0xff), pos3 = (int)(e.w[1] >> 24);
  const uint32_t creg = e.w[4], cregv = e.w[5];
  const bool tok = (tid & e.w[6]) == e.w[7];
  // Payloads: diagonal tables are read per lane from the LDS copy (C); the
  // uniform payloads of dense / exchange ops come from global memory through
  // scalar loads (G), so control selects on them are scalar and they never
  // wait behind the tile's vector loads.
  const cx<R>* m = C + opcoef;
  const cx<R>* g = G + opcoef;
  // control mode: bit 0 register-bit controls, bit 1 thread-bit controls
  const int cm = (e.w[6] ? 2 : 0) | (creg ? 1 : 0);
  switch (kind) {
    case F_DENSE1:
      switch (pos0 * 4 + cm) {
#define QS_D1(P, M)                                                   \
  case P * 4 + M:                                                     \
    if constexpr (P < RB) {                                           \
      if (M == 0 && pure) do_dense1_real<R, NR, P>(a, g);              \
      else do_dense1<R, NR, P, M>(a, g, creg, cregv, tok);             \
    } else {                                                          \
      __builtin_unreachable();                                        \
    }                                                                 \
    break;
#define QS_D1P(P) QS_D1(P, 0) QS_D1(P, 1) QS_D1(P, 2) QS_D1(P, 3)
        QS_D1P(0) QS_D1P(1) QS_D1P(2) QS_D1P(3) QS_D1P(4)
#undef QS_D1P
#undef QS_D1
        default:
          __builtin_unreachable();
      }
      break;
    case F_PERM1:
      switch (pos0 * 4 + cm) {
#define QS_P1(P, M)                                                   \
  case P * 4 + M:                                                     \
    if constexpr (P < RB) {                                           \
      if (pure) do_perm1<R, NR, P, M, true>(a, g, creg, cregv, tok);   \
      else do_perm1<R, NR, P, M, false>(a, g, creg, cregv, tok);       \
    } else {                                                          \
      __builtin_unreachable();                                        \
    }                                                                 \
    break;
#define QS_P1P(P) QS_P1(P, 0) QS_P1(P, 1) QS_P1(P, 2) QS_P1(P, 3)
        QS_P1P(0) QS_P1P(1) QS_P1P(2) QS_P1P(3) QS_P1P(4)
#undef QS_P1P
#undef QS_P1
        default:
          __builtin_unreachable();
      }
      break;
    case F_DIAG: {
      const uint32_t sel = diag_sel(ntt, (int)(e.w[2] & 0xff), (int)((e.w[2] >> 8) & 0xff), tid);
      if (pure) do_diag_lane<R, NR>(a, sel, m);
      else do_diag<R, NR>(a, sel, m);
      break;
    }
    case F_DENSE2:
    case F_XX2:
    case F_SWAP2: {
      // controlled two-target ops are rare: one instance per pair, control mode 3
      // (selects) whenever any control is present; XX ops are never controlled
      const bool ctl = cm != 0;
      const int code = pos0 * 8 + pos1;
      const bool sw = kind == F_SWAP2, xx = kind == F_XX2;
      switch (code) {
#define QS_D2(A, B)                                                                  \
  case A * 8 + B:                                                                    \
    if constexpr (A < RB) {                                                          \
      if (sw) {                                                                      \
        if (ctl) {                                                                   \
          if (pure) do_swap2<R, NR, A, B, 3, true>(a, g, creg, cregv, tok);          \
          else do_swap2<R, NR, A, B, 3, false>(a, g, creg, cregv, tok);              \
        } else {                                                                     \
          if (pure) do_swap2<R, NR, A, B, 0, true>(a, g, creg, cregv, tok);          \
          else do_swap2<R, NR, A, B, 0, false>(a, g, creg, cregv, tok);              \
        }                                                                            \
      } else if (xx) {                                                               \
        if (pure) do_xx2<R, NR, A, B, true>(a, g);                                   \
        else do_xx2<R, NR, A, B, false>(a, g);                                       \
      } else {                                                                       \
        if (ctl) do_dense2<R, NR, A, B, 3>(a, g, creg, cregv, tok);                  \
        else do_dense2<R, NR, A, B, 0>(a, g, creg, cregv, tok);                      \
      }                                                                              \
    } else {                                                                         \
      __builtin_unreachable();                                                       \
    }                                                                                \
    break;
        QS_D2(1, 0) QS_D2(2, 0) QS_D2(2, 1) QS_D2(3, 0) QS_D2(3, 1) QS_D2(3, 2)
        QS_D2(4, 0) QS_D2(4, 1) QS_D2(4, 2) QS_D2(4, 3)
#undef QS_D2
        default:
          __builtin_unreachable();
      }
      break;
    }
    case F_DENSE3: {
      if constexpr (DK) {
        const uint32_t pm = (1u << pos0) | (1u << pos1) | (1u << pos2);
        switch (pm) {
#define QS_D3(M) \
case M:        \
  if constexpr ((M >> RB) == 0) do_densek<R, NR, 3, M>(a, m, creg, cregv, tok); \
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
        const uint32_t pm = (1u << pos0) | (1u << pos1) | (1u << pos2) | (1u << pos3);
        switch (pm) {
#define QS_D4(M) \
case M:        \
  if constexpr ((M >> RB) == 0) do_densek<R, NR, 4, M>(a, m, creg, cregv, tok); \
  break;
          QS_D4(15u) QS_D4(23u) QS_D4(27u) QS_D4(29u) QS_D4(30u)
#undef QS_D4
        }
      }
      break;
    }
    default:
      __builtin_unreachable();
  }
}

// Apply a pass's op list to the register tile `a` (load layout on entry, store
// layout on exit).  DK: include the in-register 3- and 4-target dense ops.
// U2: two ops per loop iteration.  With one op per iteration the tile's value at
// the loop head and after the op switch get different registers (an op writes
// its outputs while its inputs are still live), and the compiler copies the
// whole tile (64 registers at complex128) twice per op; with two ops per
// iteration the first op's outputs form their own register set and neither op
// needs a copy.  Every impossible switch arm is unreachable (the host checks
// each op before launch), so no arm passes the tile through unchanged.
template <typename R, int RB, int TB, bool DK, bool U2>
__device__ __forceinline__ void run_ops(cx<R> (&a)[1 << RB], char* lds, const PassHdr* __restrict__ H,
                                        const FOpW* __restrict__ P, const cx<R>* __restrict__ C,
                                        const cx<R>* __restrict__ G, const uint32_t* __restrict__ X,
                                        uint32_t tid) {
  const int nops = H->nops;
  if constexpr (U2) {
    int o = 0;
    for (; o + 1 < nops; o += 2) {
      run_op<R, RB, TB, DK>(a, lds, P[o], C, G, X, tid);
      run_op<R, RB, TB, DK>(a, lds, P[o + 1], C, G, X, tid);
    }
    if (o < nops) run_op<R, RB, TB, DK>(a, lds, P[o], C, G, X, tid);
  } else {
    for (int o = 0; o < nops; ++o) run_op<R, RB, TB, DK>(a, lds, P[o], C, G, X, tid);
  }
}

// One pass over all 2^(n-K) tiles.  PF: persistent workgroups that prefetch
// the next tile into a second register set while the current one is computed
// (grid = resident workgroups); otherwise one tile per workgroup.
template <typename R, int RB, int TB, bool DK, bool PF>
__global__ __launch_bounds__(1 << TB) void k_fused(cx<R>* __restrict__ v, const PassHdr* __restrict__ H,
                                                   const FOpW* __restrict__ ops, const cx<R>* __restrict__ coef,
                                                   const uint32_t* __restrict__ aux, uint64_t ntiles) {
  constexpr int NR = 1 << RB;
  constexpr int K = RB + TB;
  extern __shared__ __align__(16) char lds_raw[];
  const uint32_t tid = threadIdx.x;
  const FOpW* P = ops + H->op0;
  const cx<R>* C;
  if constexpr (!DK) {
    // Stage the pass coefficients in LDS once per workgroup: op payload reads are
    // then LDS reads, which never wait on the global tile loads.
    cx<R>* lc = reinterpret_cast<cx<R>*>(lds_raw) + (H->uses_lds ? (1 << (K - 1)) : 0);
    const cx<R>* gc = coef + H->coef0;
    const int nc = H->ncoef;
    for (int i = (int)tid; i < nc; i += (int)blockDim.x) lc[i] = gc[i];
    __syncthreads();
    C = lc;
  } else {
    C = coef + H->coef0;
  }
  cx<R> a[NR];
  if constexpr (!PF) {
    // One tile per workgroup, straight-line: the store address is derived from
    // the header after the ops, so only the tile base stays live across them
    // (the op interpreter needs every scalar register it can get).
    const uint64_t tb = insert_zeros((uint64_t)blockIdx.x, H->sorted, K);
    {
      uint64_t gin = 0;
#pragma unroll
      for (int b = 0; b < TB; ++b) gin |= (uint64_t)((tid >> b) & 1u) << H->tbits[H->ts0[b]];
      uint64_t min_[RB];
#pragma unroll
      for (int b = 0; b < RB; ++b) min_[b] = 1ull << H->tbits[H->rs0[b]];
      const uint64_t base = tb | gin;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= min_[b];
        a[r] = ld_nt(v + (base | g));
      }
    }
    run_ops<R, RB, TB, DK, !DK>(a, lds_raw, H, P, C, coef + H->coef0, aux, tid);
    if (H->sync_store) __syncthreads();
    // store: tile base (uniform) XOR per-lane part (with pending flips) XOR register part
    uint64_t gout = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b) gout |= (uint64_t)((tid >> b) & 1u) << H->tbo[H->ts1[b]];
    const uint64_t base = tb ^ gout ^ H->sflip;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      uint64_t g = 0;
#pragma unroll
      for (int b = 0; b < RB; ++b)
        if ((r >> b) & 1) g |= 1ull << H->tbo[H->rs1[b]];
      st_nt(v + (base ^ g), a[r]);
    }
    (void)ntiles;
  } else {
    // per-lane parts of the load and store addresses, per-register parts as bit masks
    uint64_t gin = 0, gout = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b) {
      const uint64_t bit = (tid >> b) & 1u;
      gin |= bit << H->tbits[H->ts0[b]];
      gout |= bit << H->tbo[H->ts1[b]];
    }
    gout ^= H->sflip;
    uint64_t min_[RB], mout[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      min_[b] = 1ull << H->tbits[H->rs0[b]];
      mout[b] = 1ull << H->tbo[H->rs1[b]];
    }
    cx<R> nx[NR];
    uint64_t t = blockIdx.x;
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
    for (; t < ntiles; t += gridDim.x) {
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
      run_ops<R, RB, TB, DK, false>(a, lds_raw, H, P, C, coef + H->coef0, aux, tid);
      if (H->sync_store) __syncthreads();
      const uint64_t base = insert_zeros(t, H->sorted, K) ^ gout;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        uint64_t g = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
          if ((r >> b) & 1) g |= mout[b];
        st_nt(v + (base ^ g), a[r]);
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) a[r] = nx[r];
    }
  }
}

// Device image of a pass program: [hdrs][ops as FOpW][aux][coef in R], each
// part 256-byte aligned.
struct ProgImage {
  std::vector<char> host;
  uint64_t hb = 0, ob = 0, xb = 0, cb = 0;
};

template <typename R> static void pack_program(const FusedProgram& prog, ProgImage& im) {
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  im.hb = al(prog.hdrs.size() * sizeof(PassHdr));
  im.ob = al(prog.ops.size() * sizeof(FOpW));
  im.xb = al(prog.aux.size() * sizeof(uint32_t));
  const uint64_t ncoef = prog.coef.size() / 2;
  im.cb = al(ncoef * sizeof(cx<R>));
  im.host.assign(im.hb + im.ob + im.xb + im.cb, 0);
  char* h = im.host.data();
  if (!prog.hdrs.empty()) std::memcpy(h, prog.hdrs.data(), prog.hdrs.size() * sizeof(PassHdr));
  for (size_t i = 0; i < prog.ops.size(); ++i) {
    const FOpW e = encode_op(prog.ops[i]);
    std::memcpy(h + im.hb + i * sizeof(FOpW), &e, sizeof(FOpW));
  }
  if (!prog.aux.empty()) std::memcpy(h + im.hb + im.ob, prog.aux.data(), prog.aux.size() * sizeof(uint32_t));
  cx<R>* hc = reinterpret_cast<cx<R>*>(h + im.hb + im.ob + im.xb);
  for (uint64_t i = 0; i < ncoef; ++i) hc[i] = mk<R>((R)prog.coef[2 * i], (R)prog.coef[2 * i + 1]);
}

// Device view of an uploaded image.
template <typename R> struct ProgDev {
  const PassHdr* H;
  const FOpW* O;
  const uint32_t* X;
  const cx<R>* C;
  ProgDev(const char* d, const ProgImage& im)
      : H((const PassHdr*)d), O((const FOpW*)(d + im.hb)), X((const uint32_t*)(d + im.hb + im.ob)),
        C((const cx<R>*)(d + im.hb + im.ob + im.xb)) {}
};

// One fused pass: host header hd, its device copy dh.  prof: bracket the
// launch with the context's timing events (never while capturing a graph).
template <typename R, int RB, int TB>
static int launch_fused(qs_ctx* ctx, cx<R>* v, int n, const PassHdr& hd, const PassHdr* dh, const ProgDev<R>& d,
                        hipStream_t s, bool prof) {
  constexpr int K = RB + TB;
  const uint64_t tiles = 1ull << (n - K);
  const size_t lds_bytes = (size_t)(1u << (K - 1)) * sizeof(cx<R>);  // half-tile relayout image
  const size_t lds_relayout = hd.uses_lds ? lds_bytes : 0;
  const size_t lds = hd.has_dk ? lds_relayout : lds_relayout + (size_t)hd.ncoef * sizeof(cx<R>);
  int rc = QS_OK;
  if (prof && (rc = prof_begin(ctx, PROF_FUSED, s))) return rc;
  if (hd.has_dk) {
    hipLaunchKernelGGL((k_fused<R, RB, TB, true, false>), dim3((unsigned)tiles), dim3(1 << TB), lds, s, v, dh, d.O,
                       d.C, d.X, tiles);
  } else if (ctx->fused_prefetch) {
    // persistent, prefetching: resident workgroups only
    const uint64_t grid = std::min<uint64_t>(tiles, (uint64_t)ctx->num_cus * ctx->fused_wg_per_cu);
    hipLaunchKernelGGL((k_fused<R, RB, TB, false, true>), dim3((unsigned)grid), dim3(1 << TB), lds, s, v, dh, d.O,
                       d.C, d.X, tiles);
  } else {
    hipLaunchKernelGGL((k_fused<R, RB, TB, false, false>), dim3((unsigned)tiles), dim3(1 << TB), lds, s, v, dh, d.O,
                       d.C, d.X, tiles);
  }
  QS_CHECK_HIP(hipGetLastError());
  if (prof && (rc = prof_end(ctx, s))) return rc;
  return QS_OK;
}

// A pass the tile kernel cannot take: the gate's own kernel.
template <typename R>
static int launch_single(qs_ctx* ctx, cx<R>* v, int n, const qs_gate& g, hipStream_t s, bool prof) {
  int rc = QS_OK;
  if (prof && (rc = prof_begin(ctx, PROF_GATE, s))) return rc;
  rc = apply_gate(ctx, v, n, sizeof(R) == 8 ? QS_C128 : QS_C64, &g, s);
  if (rc) return rc;
  if (prof && (rc = prof_end(ctx, s))) return rc;
  return QS_OK;
}

// Launch every pass of `prog` (image at dprog) on stream s.
template <typename R, int RB, int TB>
static int launch_passes(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* gates, const FusedProgram& prog,
                         const char* dprog, const ProgImage& im, hipStream_t s, bool prof, int* npass_out) {
  const ProgDev<R> d(dprog, im);
  int npass = 0, rc = QS_OK;
  for (size_t i = 0; i < prog.plans.size(); ++i) {
    if (!prog.plans[i].fused) {
      rc = launch_single<R>(ctx, v, n, prog.pgates[prog.plans[i].single], s, prof);
    } else {
      const int h = prog.hdr_of[i];
      rc = launch_fused<R, RB, TB>(ctx, v, n, prog.hdrs[h], d.H + h, d, s, prof);
    }
    if (rc) return rc;
    ++npass;
  }
  if (npass_out) *npass_out = npass;
  return QS_OK;
}

// Planner settings: LR = 0 (low bits stay on threads at the load and store);
// `moves` lets the planner move qubits between index bits when that saves at
// least 8% of the passes, for lists of more than 480 gates (shorter lists
// rarely make the 24 passes the moving plan is tried for; they plan without
// moves and so stream, see run_fused_t).  QS_FUSION_REMAP overrides it for
// every call: 0 never, 1 whenever it saves a pass.
static Config fused_config(int RB, int TB, int L, bool moves, int ngates) {
  Config c{RB, TB, L};
  c.remap = moves && ngates > 480 ? 1 : 0;
  if (const char* rm = getenv("QS_FUSION_REMAP")) c.remap = rm[0] == '1' ? 2 : 0;
  return c;
}

// Planning overlapped with the passes, for plans without qubit moves: each
// chunk of passes is uploaded and launched as soon as it is planned, so the
// device starts after the first pass is planned instead of after the whole
// program.  A planner error after the first chunk leaves the vector partly
// evolved (the planner's errors are internal: a gate the tile kernel takes
// always schedules).
template <typename R, int RB, int TB>
static int run_fused_streamed(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* gates, int ngates, const Config& c,
                              int* passes, hipStream_t s) {
  FusedProgram prog;
  int npass = 0;
  auto chunk = [&](const FusedProgram& sub, int) -> int {
    ProgImage im;
    pack_program<R>(sub, im);
    qs_ctx::ProgSlot* slot = nullptr;
    int rc = acquire_prog_slot(ctx, im.host.size() + 256, &slot);
    if (rc) return rc;
    std::memcpy(slot->host, im.host.data(), im.host.size());
    QS_CHECK_HIP(hipMemcpyAsync(slot->dev, slot->host, im.host.size(), hipMemcpyHostToDevice, s));
    const ProgDev<R> d((const char*)slot->dev, im);
    for (size_t h = 0; h < sub.hdrs.size() && !rc; ++h, ++npass)
      rc = launch_fused<R, RB, TB>(ctx, v, n, sub.hdrs[h], d.H + h, d, s, true);
    QS_CHECK_HIP(hipEventRecord(slot->done, s));
    slot->pending = true;
    return rc;
  };
  auto single = [&](const qs_gate& g) -> int {
    ++npass;
    return launch_single<R>(ctx, v, n, g, s, true);
  };
  int rc = build_streamed(gates, ngates, n, c, (uint32_t)sizeof(cx<R>), prog, chunk, single);
  if (!rc)
    if (const char* dbg = getenv("QS_FUSION_DEBUG"))
      if (dbg[0] == '1') print_program_stats(prog, n, RB + TB, c.L, ngates);
  if (passes) *passes = npass;
  return rc;
}

template <typename R, int RB, int TB>
static int run_fused_t(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* gates, int ngates, int L, bool moves,
                       int* passes, hipStream_t s) {
  const Config c = fused_config(RB, TB, L, moves, ngates);
  if (c.remap == 0) {  // plans without moves are final as they form: streamed
    const char* st = getenv("QS_FUSION_STREAM");
    if (!(st && st[0] == '0')) return run_fused_streamed<R, RB, TB>(ctx, v, n, gates, ngates, c, passes, s);
  }
  FusedProgram prog;
  int rc = build_program(gates, ngates, n, c, (uint32_t)sizeof(cx<R>), prog);
  if (rc) return rc;
  if (const char* dbg = getenv("QS_FUSION_DEBUG"))
    if (dbg[0] == '1') print_program_stats(prog, n, RB + TB, L, ngates);
  ProgImage im;
  pack_program<R>(prog, im);
  // upload from a pinned slot that stays untouched until this call's passes
  // have run (its event), so the call returns without waiting for the device
  qs_ctx::ProgSlot* slot = nullptr;
  rc = acquire_prog_slot(ctx, im.host.size() + 256, &slot);
  if (rc) return rc;
  char* dprog = (char*)slot->dev;
  if (!im.host.empty()) {
    std::memcpy(slot->host, im.host.data(), im.host.size());
    QS_CHECK_HIP(hipMemcpyAsync(dprog, slot->host, im.host.size(), hipMemcpyHostToDevice, s));
  }
  rc = launch_passes<R, RB, TB>(ctx, v, n, gates, prog, dprog, im, s, true, passes);
  QS_CHECK_HIP(hipEventRecord(slot->done, s));
  slot->pending = true;
  return rc;
}

}  // namespace qs

// A gate sequence planned and uploaded once, its passes captured in a HIP graph.
struct qs_graph {
  int device = 0;
  int passes = 0;
  void* dprog = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace qs {

template <typename R> __global__ __launch_bounds__(256) void k_basis0(cx<R>* __restrict__ v, uint64_t N) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride)
    st(v + i, mk<R>(i == 0 ? (R)1 : (R)0, (R)0));
}

template <typename R, int RB, int TB>
static int capture_t(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* gates, int ngates, int L, int init_zero,
                     qs_graph* G) {
  const Config c = fused_config(RB, TB, L, true, ngates);  // state vectors: plans as fuse=2
  FusedProgram prog;
  int rc = build_program(gates, ngates, n, c, (uint32_t)sizeof(cx<R>), prog);
  if (rc) return rc;
  for (const PassPlan& p : prog.plans)
    if (!p.fused) {
      set_error("qs_graph_capture: a gate the tile kernel cannot take would run as its own per-gate pass "
                "(a dense gate on more than 4 targets, a diagonal or phase gate on more than 6 bits, or a gate "
                "with more than tile_log2 bits, or more than tile_log2 - low_bits bits above the low bits); "
                "such sequences are not captured");
      return QS_ERR_UNSUPPORTED;
    }
  ProgImage im;
  pack_program<R>(prog, im);
  QS_CHECK_HIP(hipMalloc(&G->dprog, std::max<size_t>(im.host.size(), 256)));
  if (!im.host.empty()) QS_CHECK_HIP(hipMemcpy(G->dprog, im.host.data(), im.host.size(), hipMemcpyHostToDevice));
  hipStream_t s = nullptr;
  QS_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(s);
    set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    return QS_ERR_DEVICE;
  }
  if (init_zero) {
    const uint64_t N = 1ull << n;
    const uint64_t blocks = std::min<uint64_t>((N + 255) / 256, 1ull << 20);
    hipLaunchKernelGGL(k_basis0<R>, dim3((unsigned)blocks), dim3(256), 0, s, v, N);
  }
  int npass = 0;
  rc = launch_passes<R, RB, TB>(ctx, v, n, gates, prog, (const char*)G->dprog, im, s, false, &npass);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(s, &graph);
  (void)hipStreamDestroy(s);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return QS_ERR_DEVICE;
  }
  G->graph = graph;
  QS_CHECK_HIP(hipGraphInstantiate(&G->exec, graph, nullptr, nullptr, 0));
  G->passes = npass;
  return QS_OK;
}

template <typename R, int RB, int TB> static void set_lds_attr() {
  static bool done = false;
  if (done) return;
  const int bytes = (int)((1u << (RB + TB - 1)) * sizeof(cx<R>) + FMAX_PASS_COEF * sizeof(cx<R>));
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, true, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, false, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)k_fused<R, RB, TB, false, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  done = true;
}

static int apply_gates_fused_impl(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates,
                                  bool moves, int* passes, hipStream_t s);

// The planner is host C++: no exception may cross the C ABI.
int apply_gates_fused(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, bool moves,
                      int* passes, hipStream_t s) {
  try {
    return apply_gates_fused_impl(ctx, vec, n, dtype, gates, ngates, moves, passes, s);
  } catch (const std::exception& e) {
    set_error(std::string("fusion planner: ") + e.what());
    return QS_ERR_DEVICE;
  }
}

static int apply_gates_fused_impl(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates,
                                  bool moves, int* passes, hipStream_t s) {
  // Tile shapes (4 register bits per thread, 2^(K-4) threads, a half-tile LDS
  // image for relayouts): K = 12 by default for both precisions (256 threads;
  // 32 KiB of LDS for complex128, 16 KiB for complex64); complex128 also takes
  // 11 or 13, complex64 13.
  const int K = ctx->tile_log2 ? ctx->tile_log2 : 12;
  const int L = ctx->low_bits ? ctx->low_bits : 4;
  if (L > K - 4) {  // tile shape and low bits may be set in either order
    set_error("low_bits must not exceed tile_log2 - 4");
    return QS_ERR_ARG;
  }
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
    cx<double>* v = (cx<double>*)vec;
    switch (K) {
      case 11:
        set_lds_attr<double, 4, 7>();
        return run_fused_t<double, 4, 7>(ctx, v, n, gates, ngates, L, moves, passes, s);
      case 12:
        set_lds_attr<double, 4, 8>();
        return run_fused_t<double, 4, 8>(ctx, v, n, gates, ngates, L, moves, passes, s);
      case 13:
        set_lds_attr<double, 4, 9>();
        return run_fused_t<double, 4, 9>(ctx, v, n, gates, ngates, L, moves, passes, s);
      default:
        set_error("complex128 fusion tiles: tile_log2 must be 11, 12 or 13");
        return QS_ERR_UNSUPPORTED;
    }
  }
  cx<float>* v = (cx<float>*)vec;
  switch (K) {
    case 12:
      set_lds_attr<float, 4, 8>();
      return run_fused_t<float, 4, 8>(ctx, v, n, gates, ngates, L, moves, passes, s);
    case 13:
      set_lds_attr<float, 4, 9>();
      return run_fused_t<float, 4, 9>(ctx, v, n, gates, ngates, L, moves, passes, s);
    default:
      set_error("complex64 fusion tiles: tile_log2 must be 12 or 13");
      return QS_ERR_UNSUPPORTED;
  }
}

}  // namespace qs

namespace qs {

static int graph_capture_impl(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates,
                              int init_zero, qs_graph* G);

static int graph_capture(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, int init_zero,
                         qs_graph* G) {
  try {
    return graph_capture_impl(ctx, vec, n, dtype, gates, ngates, init_zero, G);
  } catch (const std::exception& e) {
    set_error(std::string("fusion planner: ") + e.what());
    return QS_ERR_DEVICE;
  }
}

static int graph_capture_impl(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates,
                              int init_zero, qs_graph* G) {
  const int K = ctx->tile_log2 ? ctx->tile_log2 : 12;
  const int L = ctx->low_bits ? ctx->low_bits : 4;
  if (L > K - 4) {
    set_error("low_bits must not exceed tile_log2 - 4");
    return QS_ERR_ARG;
  }
  if (n < K + 1) {
    set_error("qs_graph_capture: registers below 2^(tile+1) amplitudes run one kernel per gate and are not captured");
    return QS_ERR_UNSUPPORTED;
  }
  if (dtype == QS_C128) {
    cx<double>* v = (cx<double>*)vec;
    switch (K) {
      case 11: set_lds_attr<double, 4, 7>(); return capture_t<double, 4, 7>(ctx, v, n, gates, ngates, L, init_zero, G);
      case 12: set_lds_attr<double, 4, 8>(); return capture_t<double, 4, 8>(ctx, v, n, gates, ngates, L, init_zero, G);
      case 13: set_lds_attr<double, 4, 9>(); return capture_t<double, 4, 9>(ctx, v, n, gates, ngates, L, init_zero, G);
      default: set_error("complex128 fusion tiles: tile_log2 must be 11, 12 or 13"); return QS_ERR_UNSUPPORTED;
    }
  }
  cx<float>* v = (cx<float>*)vec;
  switch (K) {
    case 12: set_lds_attr<float, 4, 8>(); return capture_t<float, 4, 8>(ctx, v, n, gates, ngates, L, init_zero, G);
    case 13: set_lds_attr<float, 4, 9>(); return capture_t<float, 4, 9>(ctx, v, n, gates, ngates, L, init_zero, G);
    default: set_error("complex64 fusion tiles: tile_log2 must be 12 or 13"); return QS_ERR_UNSUPPORTED;
  }
}

}  // namespace qs

using namespace qs;

extern "C" {

int qs_graph_capture(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, int init_zero,
                     qs_graph** out) {
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(out != nullptr, "null output handle");
  *out = nullptr;
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ngates >= 0 && (ngates == 0 || gates != nullptr), "bad gate list");
  for (int i = 0; i < ngates; ++i) {
    rc = validate_gate(&gates[i], n);
    if (rc) return rc;
  }
  QS_CHECK_HIP(hipSetDevice(ctx->device));
  qs_graph* G = new qs_graph();
  G->device = ctx->device;
  rc = graph_capture(ctx, vec, n, dtype, gates, ngates, init_zero, G);
  if (rc) {
    qs_graph_destroy(G);
    return rc;
  }
  *out = G;
  return QS_OK;
}

int qs_graph_launch(qs_graph* g, void* stream) {
  QS_REQUIRE(g != nullptr && g->exec != nullptr, "null or empty graph");
  QS_CHECK_HIP(hipGraphLaunch(g->exec, as_stream(stream)));
  return QS_OK;
}

int qs_graph_passes(const qs_graph* g) { return g ? g->passes : -1; }

void qs_graph_destroy(qs_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->dprog) (void)hipFree(g->dprog);
  delete g;
}

}  // extern "C"
