// Host-side planner of the fused tile passes (shared structs, layout
// bookkeeping, op construction and merging, pass grouping).  Included by
// qsim_fused.hip, which holds the device kernels and the launch code.
#pragma once
#include "qsim_common.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

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
  F_XX2 = 8,     // two targets, entries only where row ^ col is 00 or 11 (noise superoperators,
                 // Rxx/Ryy/FSim): a_s' = d_s a_s + o_s a_(s^3), payload d_0..d_3, o_0..o_3;
                 // pure != 0: real payload.  Never controlled.
};

constexpr int FMAX_TB = 10;
constexpr int FMAX_RB = 5;
// Complex coefficients one pass may carry (staged in LDS next to the relayout
// buffer).  The planner charges every gate its worst case (a diagonal 4 x 2^RB
// entries), far above what the passes use after merging (at most 820 entries
// in a 26-qubit QFT, 288 in the 30-qubit random circuit), so the limit is set
// where it stops splitting passes; 32 KiB of complex128 in the worst case.
constexpr int FMAX_PASS_COEF = 2048;

// One op of a pass program, as the host planner builds it.
struct FOp {
  int kind;
  int pure;         // PERM1/SWAP2 without phases; DIAG: lane mode (one value per lane);
                    // DENSE1: uncontrolled real matrix; XX2: real payload
  int npos;         // register positions used (targets, MSB first)
  int ntt;          // DIAG: number of selecting thread bits (<= 2)
  int pos[4];
  int tpos[2];      // DIAG: selecting thread-index bits, MSB first
  int32_t coef;     // complex offset of the payload in the pass coefficients; RELAYOUT: index into aux
  uint32_t creg;    // control mask in register-index space ...
  uint32_t cregv;   // ... and the values those register bits must have
  uint32_t cthr;    // control mask in thread-index space ...
  uint32_t cthrv;   // ... and the values those thread bits must have
};

// Its device encoding: eight 32-bit words, read with one scalar load (every
// field a whole dword, so no sub-dword vector loads are needed):
//   w0 = kind | pure << 8 | npos << 16 | ntt << 24
//   w1 = pos[0] | pos[1] << 8 | pos[2] << 16 | pos[3] << 24
//   w2 = tpos[0] | tpos[1] << 8,  w3 = coef,  w4..w7 = creg, cregv, cthr, cthrv
struct FOpW {
  uint32_t w[8];
};
static_assert(sizeof(FOpW) == 32, "FOpW layout");

static FOpW encode_op(const FOp& o) {
  FOpW e;
  e.w[0] = (uint32_t)o.kind | (uint32_t)o.pure << 8 | (uint32_t)o.npos << 16 | (uint32_t)o.ntt << 24;
  e.w[1] = (uint32_t)o.pos[0] | (uint32_t)o.pos[1] << 8 | (uint32_t)o.pos[2] << 16 | (uint32_t)o.pos[3] << 24;
  e.w[2] = (uint32_t)o.tpos[0] | (uint32_t)o.tpos[1] << 8;
  e.w[3] = (uint32_t)o.coef;
  e.w[4] = o.creg;
  e.w[5] = o.cregv;
  e.w[6] = o.cthr;
  e.w[7] = o.cthrv;
  return e;
}

// RELAYOUT payload in the aux table (uint32): the register position p of the
// tile bit h that splits the exchange into two rounds; the RB register-slot and
// TB thread-slot LDS offsets of the current layout; the flip term of the
// current layout; then the RB + TB slot offsets of the new layout.  Offsets
// address a half-tile image indexed by the tile bits other than h: an offset is
// the byte offset the single tile bit its slot carries contributes under the
// relayout's XOR-linear placement (0 for the slot of h), so an amplitude's LDS
// address is the XOR of its slots' offsets.  The write side is carried per
// relayout because relabelings change the layout between relayouts.
struct PassHdr {
  int32_t nops;
  int32_t op0;
  int32_t coef0;
  int32_t uses_lds;
  int32_t has_dk;        // pass contains 3- or 4-target dense ops (kernel variant without LDS coefficients)
  int32_t ncoef;         // complex coefficients of this pass (staged in LDS)
  int32_t sync_store;    // store layout differs from the load layout and no relayout ordered the
                         // loads: workgroup barrier before the stores (another wave may still have
                         // to load an address this wave stores to)
  int32_t tbits[16];     // tile-local bit -> global bit
  int32_t tbo[16];       // tile-local bit -> global bit at the store: a permutation of tbits that moves
                         // qubits between the tile's index bits (the same set of addresses)
  int32_t sorted[16];    // tile bits ascending (global), for the tile base
  int32_t rs0[FMAX_RB], ts0[FMAX_TB];  // load layout (tile-local bits)
  int32_t rs1[FMAX_RB], ts1[FMAX_TB];  // store layout
  uint64_t sflip;        // global-index XOR applied to every store address (pending flips)
};

// ---- host planner --------------------------------------------------------------

struct Config {
  int RB, TB, L;
  // Low bits allowed in registers at the load and the store.  Measured on
  // MI355X (28 qubits, 2000-gate random circuit): 1 or 2 save about 40% of the
  // relayouts but slow every pass by ~20% (half-coalesced tile accesses).
  int LR = 0;
  // Qubits moved onto the low index bits between passes: 0 never; 1 when that
  // saves at least 8% of the passes; 2 whenever it saves a pass.  Measured on
  // MI355X, the 30-qubit 2000-gate random circuit needs 74 instead of 84
  // passes, each carrying more gates (8.19 against 7.65 ms): 606 against
  // 643 ms.  With 8% fewer passes the circuit was 2.3% faster; the noisy
  // density matrix (26 instead of 29 passes) is 5% slower, so its calls do
  // not ask for it.
  int remap = 0;
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
  if (g.kind == QS_GATE_XPERM && g.ncontrols == 0) {  // a flip, plus a diagonal phase table unless pure
    const bool pure = g.m[0] == 1.0 && g.m[1] == 0.0 && g.m[2] == 1.0 && g.m[3] == 0.0;
    return pure ? 0 : 2 * NR;
  }
  if (g.kind == QS_GATE_XPERM) return 2;
  if (g.kind == QS_GATE_SWAP) return 16;                         // may become a 4x4 under mixed flips
  return 1 << (2 * g.ntargets);
}

static bool fusable(const qs_gate& g, const Config& c) {
  const int K = c.RB + c.TB;
  if (g.ncontrols + g.ntargets > K) return false;
  if (g.kind == QS_GATE_DENSE && g.ntargets > std::min(4, c.RB)) return false;
  if (gate_diag_like(g) && g.ncontrols + g.ntargets > c.RB + 2) return false;
  return true;
}

// Complex payload of a gate converted to the op form.
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

// M'[i][j] = M[i ^ x][j ^ x]: conjugation by X on the flipped target bits.
static void xor_conjugate(std::vector<double>& M, int k, int x) {
  if (!x) return;
  const int D = 1 << k;
  std::vector<double> out(M.size());
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      out[2 * (i * D + j)] = M[2 * ((i ^ x) * D + (j ^ x))];
      out[2 * (i * D + j) + 1] = M[2 * ((i ^ x) * D + (j ^ x)) + 1];
    }
  M.swap(out);
}

static int index_in(const std::vector<int>& v, int x) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == x) return (int)i;
  return -1;
}

// Layout: which tile-local bit each register slot / thread slot carries, and
// which slots are flipped (the slot's physical index bit is the complement of
// its tile bit).  Slot numbering: register position p -> p, thread position q -> RB + q.
struct Layout {
  std::vector<int> rs;  // register bit i -> tile-local bit
  std::vector<int> ts;  // thread bit i -> tile-local bit
  uint32_t flip = 0;
};

// Threads take the non-register tile bits in the order `tord` (ascending when
// empty): the first thread bits are the ones the store writes to the low
// index bits.
static Layout make_layout(const std::vector<int>& regs, int K, const std::vector<int>& tord = {}) {
  Layout L;
  L.rs = regs;
  std::vector<char> isreg(K, 0);
  for (int b : regs) isreg[b] = 1;
  for (int i = 0; i < K; ++i) {
    const int b = tord.empty() ? i : tord[i];
    if (!isreg[b]) L.ts.push_back(b);
  }
  return L;
}

static int slot_of(const Layout& L, int b) {
  const int p = index_in(L.rs, b);
  if (p >= 0) return p;
  return (int)L.rs.size() + index_in(L.ts, b);
}
static int flipped(const Layout& L, int b) { return (L.flip >> slot_of(L, b)) & 1; }

// Tile-local index mask of the flipped slots.
static uint32_t flip_local(const Layout& L) {
  uint32_t m = 0;
  const int RB = (int)L.rs.size();
  for (int s = 0; s < RB + (int)L.ts.size(); ++s)
    if ((L.flip >> s) & 1) m |= 1u << (s < RB ? L.rs[s] : L.ts[s - RB]);
  return m;
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

// Diagonal table of a diagonal-like gate (or a plain per-bit diagonal given by
// `vals`, 2^t entries over the targets) under layout `cur`.
static int build_diag(const qs_gate& g, const std::vector<int>& loc, const Layout& cur, int NR, OpBuild& ob) {
  FOp& op = ob.op;
  const int c = g.ncontrols, t = g.ntargets;
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
        bit ^= flipped(cur, loc[i]);  // logical value of the gate bit
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

static int build_op(const qs_gate& g, const std::vector<int>& loc, const Layout& cur, int NR, OpBuild& ob) {
  FOp& op = ob.op;
  std::memset(&op, 0, sizeof(op));
  const int c = g.ncontrols, t = g.ntargets;
  ob.coef.clear();
  if (gate_diag_like(g)) return build_diag(g, loc, cur, NR, ob);
  // controls of a non-diagonal gate: register-space mask + lane predicate,
  // each with the physical value that means logical 1
  for (int i = 0; i < c; ++i) {
    const int b = loc[i];
    const int one = 1 ^ flipped(cur, b);
    const int ri = index_in(cur.rs, b);
    if (ri >= 0) {
      op.creg |= 1u << ri;
      op.cregv |= (uint32_t)one << ri;
    } else {
      const int ti = index_in(cur.ts, b);
      op.cthr |= 1u << ti;
      op.cthrv |= (uint32_t)one << ti;
    }
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
    double m0r = g.m[0], m0i = g.m[1], m1r = g.m[2], m1i = g.m[3];
    if (flipped(cur, loc[c])) {  // physical 0 is logical 1: the phases trade places
      std::swap(m0r, m1r);
      std::swap(m0i, m1i);
    }
    op.pure = (m0r == 1.0 && m0i == 0.0 && m1r == 1.0 && m1i == 0.0);
    ob.coef = {m0r, m0i, m1r, m1i};
    return QS_OK;
  }
  std::vector<double> M;
  if (g.kind == QS_GATE_SWAP) {
    const int f0 = flipped(cur, loc[c]), f1 = flipped(cur, loc[c + 1]);
    if (f0 == f1) {
      op.kind = F_SWAP2;
      op.npos = 2;
      double m0r = g.m[0], m0i = g.m[1], m1r = g.m[2], m1i = g.m[3];
      int p0 = rp[0], p1 = rp[1];
      if (p0 < p1) {  // canonical: pos[0] > pos[1]; exchanging target roles swaps m0/m1
        std::swap(p0, p1);
        std::swap(m0r, m1r);
        std::swap(m0i, m1i);
      }
      if (f0) {  // both flipped: physical 01 is logical 10
        std::swap(m0r, m1r);
        std::swap(m0i, m1i);
      }
      op.pos[0] = p0;
      op.pos[1] = p1;
      op.pure = (m0r == 1.0 && m0i == 0.0 && m1r == 1.0 && m1i == 0.0);
      ob.coef = {m0r, m0i, m1r, m1i};
      return QS_OK;
    }
    // one target flipped: the exchange becomes 00 <-> 11 physically; use the 4x4 form
    M.assign(32, 0.0);
    M[2 * 0] = 1.0;                                    // |00> -> |00>
    M[2 * (1 * 4 + 2)] = g.m[0], M[2 * (1 * 4 + 2) + 1] = g.m[1];  // a01' = m0 a10
    M[2 * (2 * 4 + 1)] = g.m[2], M[2 * (2 * 4 + 1) + 1] = g.m[3];  // a10' = m1 a01
    M[2 * 15] = 1.0;                                   // |11> -> |11>
  } else {
    const int D = 1 << t;
    M.resize(2 * D * D);
    const double* src = (t <= 2) ? g.m : g.big;
    for (int i = 0; i < 2 * D * D; ++i) M[i] = src[i];
  }
  // flips: conjugate by X on the flipped targets (gate target order, MSB first)
  int x = 0;
  for (int i = 0; i < t; ++i)
    if (flipped(cur, loc[c + i])) x |= 1 << (t - 1 - i);
  xor_conjugate(M, t, x);
  // order targets by descending register position (MSB first)
  std::vector<int> order(t);
  for (int i = 0; i < t; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return rp[a] > rp[b]; });
  std::vector<double> Mp;
  permute_matrix(M, t, order.data(), Mp);
  op.npos = t;
  for (int i = 0; i < t; ++i) op.pos[i] = rp[order[i]];
  op.kind = (t == 1) ? F_DENSE1 : (t == 2) ? F_DENSE2 : (t == 3) ? F_DENSE3 : F_DENSE4;
  ob.coef = Mp;
  if (t == 2 && c == 0) {
    // entries only on the diagonal and the anti-diagonal (row ^ col in {0, 3}):
    // half the coefficients, and real ones (noise superoperators) need no
    // imaginary products
    bool xx = true, real = true;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const double re = Mp[2 * (4 * i + j)], im = Mp[2 * (4 * i + j) + 1];
        if ((i ^ j) == 1 || (i ^ j) == 2) xx = xx && re == 0.0 && im == 0.0;
        else real = real && im == 0.0;
      }
    if (xx) {
      op.kind = F_XX2;
      op.pure = real;
      ob.coef.assign(16, 0.0);
      for (int s = 0; s < 4; ++s) {
        ob.coef[2 * s] = Mp[2 * (5 * s)];
        ob.coef[2 * s + 1] = Mp[2 * (5 * s) + 1];
        ob.coef[2 * (4 + s)] = Mp[2 * (4 * s + (s ^ 3))];
        ob.coef[2 * (4 + s) + 1] = Mp[2 * (4 * s + (s ^ 3)) + 1];
      }
    }
  }
  return QS_OK;
}

// ---- peephole merging of built ops ---------------------------------------------
//
// Ops act on physical slots (register positions and thread positions); two ops
// on disjoint slot sets commute.  A new op is merged into the most recent
// earlier op that shares a slot with it (everything in between commutes with
// it), never across a relayout:
//   diag  x diag           -> one diagonal table (thread selector union <= 2 bits)
//   dense1 x dense1        -> product matrix (same register position and controls)
//   dense1 x diag on that register bit (no controls) -> product matrix
// Relabelings (SWAP, X) move no data, so they never block a merge.

struct Built {
  FOp op;
  std::vector<double> coef;  // (re, im) pairs
  uint32_t slots = 0;        // physical footprint: bit p = register position p, bit RB+q = thread position q
  std::vector<int> selp;     // DIAG: thread positions selecting the value, MSB first
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

// DIAG op confined to one register position: returns the position and (d0, d1), else -1.
static int diag_single_reg(const Built& b, int NR, int RB, double* d2) {
  if (b.op.kind != F_DIAG || !b.selp.empty() || __builtin_popcount(b.slots) != 1) return -1;
  const int p = __builtin_ctz(b.slots);
  if (p >= RB) return -1;
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
  std::vector<int> u = e.selp;
  for (int b : d.selp)
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
      const size_t ie = (size_t)proj(e.selp, us) * NR + r, id = (size_t)proj(d.selp, us) * NR + r;
      const cplx v = cm_({e.coef[2 * ie], e.coef[2 * ie + 1]}, {d.coef[2 * id], d.coef[2 * id + 1]});
      T[2 * ((size_t)us * NR + r)] = v.first;
      T[2 * ((size_t)us * NR + r) + 1] = v.second;
    }
  e.coef.swap(T);
  e.selp = u;
  e.slots |= d.slots;
  return true;
}

static bool uncontrolled(const FOp& o) { return o.creg == 0 && o.cthr == 0; }

// Dense 4x4 form ((re, im) pairs) of an uncontrolled two-target op.
static void two_target_dense(const Built& b, double* M) {
  if (b.op.kind == F_DENSE2) {
    for (int i = 0; i < 32; ++i) M[i] = b.coef[i];
    return;
  }
  for (int i = 0; i < 32; ++i) M[i] = 0.0;
  for (int s = 0; s < 4; ++s) {  // F_XX2: diagonal d_s, anti-diagonal o_s
    M[2 * (5 * s)] = b.coef[2 * s];
    M[2 * (5 * s) + 1] = b.coef[2 * s + 1];
    M[2 * (4 * s + (s ^ 3))] = b.coef[2 * (4 + s)];
    M[2 * (4 * s + (s ^ 3)) + 1] = b.coef[2 * (4 + s) + 1];
  }
}

// Store the product of two same-position two-target ops in `e`, in the
// diagonal + anti-diagonal form when its entries allow.
static void set_two_target(Built& e, const double* M) {
  bool xx = true, real = true;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double re = M[2 * (4 * i + j)], im = M[2 * (4 * i + j) + 1];
      if ((i ^ j) == 1 || (i ^ j) == 2) xx = xx && re == 0.0 && im == 0.0;
      else real = real && im == 0.0;
    }
  if (xx) {
    e.op.kind = F_XX2;
    e.op.pure = real;
    e.coef.assign(16, 0.0);
    for (int s = 0; s < 4; ++s) {
      e.coef[2 * s] = M[2 * (5 * s)];
      e.coef[2 * s + 1] = M[2 * (5 * s) + 1];
      e.coef[2 * (4 + s)] = M[2 * (4 * s + (s ^ 3))];
      e.coef[2 * (4 + s) + 1] = M[2 * (4 * s + (s ^ 3)) + 1];
    }
  } else {
    e.op.kind = F_DENSE2;
    e.op.pure = 0;
    e.coef.assign(M, M + 32);
  }
}

static bool try_merge(std::vector<Built>& out, Built& d, int NR, int RB) {
  const bool two = (d.op.kind == F_DENSE2 || d.op.kind == F_XX2) && uncontrolled(d.op);
  if (d.op.kind != F_DIAG && d.op.kind != F_DENSE1 && !two) return false;
  int k = (int)out.size() - 1;
  for (; k >= 0; --k) {
    if (out[k].op.kind == F_RELAYOUT) return false;
    if (out[k].slots & d.slots) break;
  }
  if (k < 0) return false;
  Built& e = out[k];
  if (two) {
    // two-target x two-target on the same register positions: one product op
    if ((e.op.kind != F_DENSE2 && e.op.kind != F_XX2) || !uncontrolled(e.op) || e.op.pos[0] != d.op.pos[0] ||
        e.op.pos[1] != d.op.pos[1])
      return false;
    double A[32], B[32], P[32];
    two_target_dense(d, A);
    two_target_dense(e, B);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double re = 0.0, im = 0.0;
        for (int t = 0; t < 4; ++t) {
          const double ar = A[2 * (4 * i + t)], ai = A[2 * (4 * i + t) + 1];
          const double br = B[2 * (4 * t + j)], bi = B[2 * (4 * t + j) + 1];
          re += ar * br - ai * bi;
          im += ar * bi + ai * br;
        }
        P[2 * (4 * i + j)] = re;
        P[2 * (4 * i + j) + 1] = im;
      }
    set_two_target(e, P);
    return true;
  }
  if (d.op.kind == F_DIAG && e.op.kind == F_DIAG) return merge_diag(e, d, NR);
  double d2[4], M[8];
  if (d.op.kind == F_DENSE1 && e.op.kind == F_DENSE1 && e.op.pos[0] == d.op.pos[0] && e.op.creg == d.op.creg &&
      e.op.cregv == d.op.cregv && e.op.cthr == d.op.cthr && e.op.cthrv == d.op.cthrv) {
    mat2_mul(d.coef.data(), e.coef.data(), M);
    e.coef.assign(M, M + 8);
    return true;
  }
  if (d.op.kind == F_DENSE1 && uncontrolled(d.op)) {
    const int p = diag_single_reg(e, NR, RB, d2);
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
  if (d.op.kind == F_DIAG && e.op.kind == F_DENSE1 && uncontrolled(e.op)) {
    const int p = diag_single_reg(d, NR, RB, d2);
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

// Final encoding of a DIAG op: register-independent tables collapse to one
// value per selector (op.pure = 1, lane mode).
static void finalize_diag(Built& b, int NR) {
  const int ns = (int)b.selp.size();
  b.op.ntt = ns;
  for (int k = 0; k < ns; ++k) b.op.tpos[k] = b.selp[k];
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
  std::vector<int> tbo;            // tile-local -> global at the store (a permutation of tbits)
  int hdr = -1;                    // its header in the program (fused passes)
  bool move = false;               // the store moves the qubits on the first L lanes onto the low bits
  std::vector<int> pref;           // move: tile-local bits in the order relayouts give them lanes
};

// Pass grouping.  A gate may join the current pass when its bits fit in the K
// tile bits and no earlier gate still waiting shares a bit with it.  Among the
// gates that may join, the one adding the fewest new tile bits is taken first
// (gates inside the tile cost nothing), so the tile's bits go to the gates
// that share them most; this packs about a quarter more gates per pass than
// taking gates in circuit order.
//
// Every tile holds the L lowest index bits (coalesced accesses), so only K - L
// of its bits are chosen per pass.  With `remap`, a pass's store may permute
// the tile's own index bits (the same addresses): the pass's relayouts give
// the first lanes to the tile bits whose qubits the next gates use first, and
// the store moves whatever qubits sit on the first L lanes at the end onto the
// low bits (so the move itself costs no relayout), the qubits there taking the
// vacated bits; the next pass then has the moved qubits for free.  Gates are
// planned on physical bits (`pg`: the records with their
// bits translated through the map in force for their pass), and passes without
// gates at the end put every qubit back on its own bit.  Each fused pass goes
// to `on_pass` (the scheduler) as soon as it is formed: the move is known only
// after scheduling.
static void plan_groups(const qs_gate* gates, int ngates, int n, const Config& c, bool remap,
                        std::vector<PassPlan>& plans, std::vector<qs_gate>& pg,
                        const std::function<void(PassPlan&)>& on_pass) {
  const int K = c.RB + c.TB;
  const int L = c.L;
  pg.assign(gates, gates + ngates);
  std::vector<int> remaining(ngates);
  for (int i = 0; i < ngates; ++i) remaining[i] = i;
  const uint64_t low = (1ull << L) - 1;
  // phys[q]: index bit holding qubit bit q; at[b]: qubit bit held by index bit b
  std::vector<int> phys(n), at(n);
  for (int b = 0; b < n; ++b) phys[b] = at[b] = b;
  auto to_phys = [&](uint64_t m) {
    uint64_t r = 0;
    for (; m; m &= m - 1) r |= 1ull << phys[__builtin_ctzll(m)];
    return r;
  };
  // gm: every bit of the gate; gx: the bits it mixes (non-diagonal targets; all
  // bits of a gate the tile kernel cannot take).  Two gates commute when every
  // bit they share is used diagonally by both.  Logical masks (gmq, gxq) are
  // translated into gm, gx whenever the map changes.
  std::vector<uint64_t> gmq(ngates), gxq(ngates), gm(ngates), gx(ngates);
  std::vector<int> gcost(ngates);
  std::vector<char> ok(ngates);
  for (int i = 0; i < ngates; ++i) {
    const qs_gate& g = gates[i];
    gmq[i] = gate_mask(g);
    gcost[i] = coef_cost(g, 1 << c.RB);
    ok[i] = fusable(g, c);
    gxq[i] = 0;
    if (!ok[i]) gxq[i] = gmq[i];
    else if (!gate_diag_like(g))
      for (int k = 0; k < g.ntargets; ++k) gxq[i] |= 1ull << g.bits[g.ncontrols + k];
    gm[i] = gmq[i];
    gx[i] = gxq[i];
  }
  auto translate = [&](int gi) {
    qs_gate& g = pg[gi];
    for (int i = 0; i < g.ncontrols + g.ntargets; ++i) g.bits[i] = phys[gates[gi].bits[i]];
  };
  std::vector<char> taken(ngates, 0);
  const uint64_t all_bits = n >= 64 ? ~0ull : (1ull << n) - 1;
  while (!remaining.empty()) {
    PassPlan p;
    p.fused = true;
    p.single = -1;
    // A gate the tile kernel cannot take at the head of the list runs on its own.
    if (!ok[remaining[0]]) {
      p.fused = false;
      p.single = remaining[0];
      translate(p.single);
      remaining.erase(remaining.begin());
      plans.push_back(std::move(p));
      continue;
    }
    uint64_t S = low;
    int coef_budget = FMAX_PASS_COEF;  // LDS-staged payload per pass (complex entries)
    for (;;) {
      // a later gate may move ahead of the gates still waiting only if it
      // commutes with all of them
      uint64_t wait_all = 0, wait_mix = 0;
      int best = -1, best_new = 1 << 30;
      for (int gi : remaining) {
        if (taken[gi]) continue;
        const uint64_t m = gm[gi];
        const bool blocked = (m & wait_mix) || (gx[gi] & wait_all);
        if (!blocked && ok[gi] && gcost[gi] <= coef_budget && __builtin_popcountll(S | m) <= K) {
          const int nb = __builtin_popcountll(m & ~S);
          if (nb < best_new) {
            best_new = nb;
            best = gi;
            if (nb == 0) break;
          }
        }
        wait_all |= m;
        wait_mix |= gx[gi];
        if (wait_mix == all_bits) break;  // every later gate shares a bit some waiting gate mixes
      }
      if (best < 0) break;
      taken[best] = 1;
      S |= gm[best];
      coef_budget -= gcost[best];
      p.gates.push_back(best);
    }
    if (p.gates.empty()) {
      // No gate fits beside the L low bits (the head gate has more than K - L
      // bits above them, e.g. a 10-bit controlled gate on high qubits): the
      // head gate, which nothing precedes, runs as its own per-gate pass.
      p.fused = false;
      p.single = remaining[0];
      translate(p.single);
      remaining.erase(remaining.begin());
      plans.push_back(std::move(p));
      continue;
    }
    // circuit order within the pass (the scheduler reorders commuting gates itself)
    std::sort(p.gates.begin(), p.gates.end());
    for (int gi : p.gates) translate(gi);
    std::vector<int> rest;
    for (int gi : remaining)
      if (!taken[gi]) rest.push_back(gi);
    remaining.swap(rest);
    // fill the tile to K bits with the lowest unused bits: that lengthens the
    // contiguous runs each workgroup reads and writes (DRAM page locality)
    for (int b = 0; b < n && __builtin_popcountll(S) < K; ++b) S |= 1ull << b;
    for (int b = 0; b < n; ++b)
      if ((S >> b) & 1) p.tbits.push_back(b);  // ascending; the first L are the low bits
    p.tbo = p.tbits;
    if (remap && !remaining.empty()) {
      // lane order: tile bits by the first of the next 64 gates that uses them
      const int kt = (int)p.tbits.size();
      std::vector<int> first(kt, 1 << 30);
      int pos = 0;
      for (int gi : remaining) {
        if (pos == 64) break;
        for (int j = 0; j < kt; ++j)
          if (((gm[gi] >> p.tbits[j]) & 1) && first[j] > pos) first[j] = pos;
        ++pos;
      }
      p.pref.resize(kt);
      for (int j = 0; j < kt; ++j) p.pref[j] = j;
      std::stable_sort(p.pref.begin(), p.pref.end(), [&](int x, int y) { return first[x] < first[y]; });
      p.move = true;
    }
    on_pass(p);
    if (p.move) {
      std::vector<int> nat_ = at;
      bool changed = false;
      for (size_t j = 0; j < p.tbits.size(); ++j) {
        nat_[p.tbo[j]] = at[p.tbits[j]];
        changed = changed || p.tbo[j] != p.tbits[j];
      }
      if (changed) {
        at.swap(nat_);
        for (int b = 0; b < n; ++b) phys[at[b]] = b;
        for (int gi : remaining) {
          gm[gi] = to_phys(gmq[gi]);
          gx[gi] = to_phys(gxq[gi]);
        }
      }
    }
    plans.push_back(std::move(p));
  }
  if (!remap) return;
  // Passes without gates that return every qubit to its own bit.  Each moves
  // the qubits within one tile (the L low bits plus up to K - L others): whole
  // cycles of the permutation when they fit, else a run along one cycle (all
  // but its last qubit reach their own bits).
  for (;;) {
    std::vector<char> in_tile(n, 0), seen(n, 0);
    for (int b = 0; b < L; ++b) in_tile[b] = 1;
    int room = K - L;
    std::vector<std::vector<int>> cycles;
    for (int b = 0; b < n; ++b) {
      if (seen[b] || at[b] == b) continue;
      std::vector<int> cyc;
      for (int x = b; !seen[x]; x = at[x]) {  // index bit x's qubit belongs on bit at[x]
        seen[x] = 1;
        cyc.push_back(x);
      }
      cycles.push_back(cyc);
    }
    if (cycles.empty()) break;
    auto highs = [&](const std::vector<int>& cyc) {
      int h = 0;
      for (int x : cyc) h += (x >= L && !in_tile[x]);
      return h;
    };
    std::stable_sort(cycles.begin(), cycles.end(),
                     [&](const std::vector<int>& u, const std::vector<int>& v) { return highs(u) < highs(v); });
    bool any = false;
    for (const auto& cyc : cycles) {
      const int h = highs(cyc);
      if (h > room) continue;
      for (int x : cyc) in_tile[x] = 1;
      room -= h;
      any = true;
    }
    if (!any) {
      // a run along the first cycle, starting at one of its low bits if it has one
      std::vector<int> cyc = cycles[0];
      for (size_t i = 0; i < cyc.size(); ++i)
        if (cyc[i] < L) {
          std::rotate(cyc.begin(), cyc.begin() + i, cyc.end());
          break;
        }
      for (int x : cyc) {
        if (x >= L && !in_tile[x]) {
          if (room == 0) break;
          --room;
        }
        in_tile[x] = 1;
      }
    }
    // qubits whose own bit is in the tile go there; the others take the bits left
    PassPlan p;
    p.fused = true;
    p.single = -1;
    int cnt = 0;
    for (int b = 0; b < n; ++b) cnt += in_tile[b];
    for (int b = 0; b < n && cnt < K; ++b)
      if (!in_tile[b]) {
        in_tile[b] = 1;
        ++cnt;
      }
    for (int b = 0; b < n; ++b)
      if (in_tile[b]) p.tbits.push_back(b);
    const int kt = (int)p.tbits.size();
    std::vector<int> dest(kt, -1);
    std::vector<char> used(n, 0);
    for (int j = 0; j < kt; ++j) {
      const int q = at[p.tbits[j]];
      if (in_tile[q]) {
        dest[j] = q;
        used[q] = 1;
      }
    }
    int free_b = 0;
    for (int j = 0; j < kt; ++j) {
      if (dest[j] >= 0) continue;
      while (!(in_tile[p.tbits[free_b]] && !used[p.tbits[free_b]])) ++free_b;
      dest[j] = p.tbits[free_b];
      used[dest[j]] = 1;
    }
    p.tbo = dest;
    std::vector<int> nat_(n);
    for (int b = 0; b < n; ++b) nat_[b] = at[b];
    for (int j = 0; j < kt; ++j) nat_[dest[j]] = at[p.tbits[j]];
    at.swap(nat_);
    for (int b = 0; b < n; ++b) phys[at[b]] = b;
    on_pass(p);
    plans.push_back(std::move(p));
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

// A whole pass program: headers, ops, complex coefficients and the relayout
// constants, ready for upload.
struct FusedProgram {
  std::vector<PassPlan> plans;
  std::vector<qs_gate> pgates;  // the gate records on index bits (single-gate passes run these)
  std::vector<int> hdr_of;  // plan -> header (-1 for single-gate passes)
  std::vector<PassHdr> hdrs;
  std::vector<FOp> ops;
  std::vector<double> coef;  // (re, im)
  std::vector<uint32_t> aux;
  std::vector<int> aux0;     // header -> first aux word of its relayouts
};


// Tile-local index with bit h removed (higher bits shift down by one).
static uint32_t drop_bit(uint32_t m, int h) { return (m & ((1u << h) - 1)) | ((m >> (h + 1)) << h); }

// LDS offsets (bytes) of the slots of layout L in the half-tile image that
// leaves out tile bit h; the slot carrying h gets offset 0.
// XOR-linear placement of the half-tile image for one relayout from layout A
// to layout B (tile bit h left out): per tile bit, the amplitude index bits
// it contributes.  One LDS cycle serves 128 bytes, i.e. the lanes that differ
// in their first G = log2(128 / amp_bytes) thread bits; those lanes hit
// distinct banks when their bits' low G index parts are independent.  A's
// first G thread bits take index bits 0..G-1, every other bit its own index
// bit above, and B's first G thread bits that are not among A's get the free
// low bits XORed in.  The images of the K - 1 bits stay independent (each has
// its own index bit), so the placement is a bijection onto 2^(K-1) entries.
static std::vector<uint32_t> relayout_placement(const Layout& A, const Layout& B, int h, int K, uint32_t amp_bytes) {
  int G = 0;
  while ((amp_bytes << G) < 128u) ++G;
  std::vector<uint32_t> pl(K, 0);
  std::vector<char> placed(K, 0);
  placed[h] = 1;
  uint32_t low_used = 0;
  const int ga = std::min(G, (int)A.ts.size());
  for (int i = 0; i < ga; ++i) {
    pl[A.ts[i]] = 1u << i;
    placed[A.ts[i]] = 1;
  }
  int hi = ga;
  for (int b = 0; b < K; ++b)
    if (!placed[b]) pl[b] = 1u << hi++;
  const int gb = std::min(ga, (int)B.ts.size());
  for (int i = 0; i < gb; ++i)
    if (pl[B.ts[i]] < (1u << ga)) low_used |= pl[B.ts[i]];
  for (int i = 0; i < gb; ++i) {
    const int b = B.ts[i];
    if (pl[b] < (1u << ga)) continue;
    int k = 0;
    while ((low_used >> k) & 1) ++k;
    pl[b] |= 1u << k;
    low_used |= 1u << k;
  }
  return pl;
}

static void slot_offsets_placed(const Layout& L, const std::vector<uint32_t>& pl, uint32_t amp_bytes, uint32_t* lr,
                                uint32_t* lt) {
  for (size_t p = 0; p < L.rs.size(); ++p) lr[p] = pl[L.rs[p]] * amp_bytes;
  for (size_t q = 0; q < L.ts.size(); ++q) lt[q] = pl[L.ts[q]] * amp_bytes;
}

// With p.move, relayouts give lanes in the order p.pref and the store layout
// decides p.tbo: the qubits on the first L lanes go to the low bits.
static int schedule_pass(const qs_gate* gates, PassPlan& p, const Config& c, uint32_t amp_bytes, PassHdr& H,
                         FusedProgram& prog) {
  const int K = c.RB + c.TB;
  std::memset(&H, 0, sizeof(H));
  std::vector<int> g2l(64, -1);
  // tord: tile bits in the order of the index bits the store writes them to,
  // for the layouts relayouts build (the store's coalescing and contiguous
  // runs); the load layout stays ascending
  std::vector<int> tord;
  std::vector<char> out_low(K, 0);
  for (int j = 0; j < K; ++j) {
    H.tbits[j] = p.tbits[j];
    H.tbo[j] = p.tbo[j];
    H.sorted[j] = p.tbits[j];
    g2l[p.tbits[j]] = j;
    out_low[j] = p.tbo[j] < c.L;
  }
  if (p.move) {
    tord = p.pref;
  } else {
    for (int j = 0; j < K; ++j) tord.push_back(j);
    std::sort(tord.begin(), tord.end(), [&](int x, int y) { return p.tbo[x] < p.tbo[y]; });
  }
  // per-gate local bits and lookahead list of needed (non-diagonal target) bits
  std::vector<std::vector<int>> loc(p.gates.size());
  for (size_t k = 0; k < p.gates.size(); ++k) {
    const qs_gate& g = gates[p.gates[k]];
    for (int i = 0; i < g.ncontrols + g.ntargets; ++i) loc[k].push_back(g2l[g.bits[i]]);
  }
  auto relabel_swap = [&](const qs_gate& g, const std::vector<int>& lb) {
    // uncontrolled phase-free SWAP whose bits are both low or both high
    return g.kind == QS_GATE_SWAP && g.ncontrols == 0 && g.m[0] == 1.0 && g.m[1] == 0.0 && g.m[2] == 1.0 &&
           g.m[3] == 0.0 && ((lb[0] < c.L) == (lb[1] < c.L));
  };
  auto lookahead_from = [&](size_t k0) {
    std::vector<int> la;
    for (size_t k = k0; k < p.gates.size() && la.size() < 16; ++k) {
      const qs_gate& g = gates[p.gates[k]];
      if (gate_diag_like(g) || relabel_swap(g, loc[k])) continue;
      if (g.kind == QS_GATE_XPERM && g.ncontrols == 0) continue;
      for (int i = 0; i < g.ntargets; ++i) {
        const int b = loc[k][g.ncontrols + i];
        if (std::find(la.begin(), la.end(), b) == la.end()) la.push_back(b);
      }
    }
    return la;
  };
  // initial layout: registers on high tile bits needed first
  std::vector<int> init_need;
  int init_low = 0;
  for (int b : lookahead_from(0)) {
    if ((int)init_need.size() >= c.RB) break;
    if (b >= c.L) {
      init_need.push_back(b);
    } else if (init_low < c.LR) {
      init_need.push_back(b);
      ++init_low;
    }
  }
  std::vector<int> regs = choose_regs(init_need, {}, {}, c.RB, K, c.L);
  Layout cur = make_layout(regs, K);
  for (int b = 0; b < c.RB; ++b) H.rs0[b] = cur.rs[b];
  for (int b = 0; b < c.TB; ++b) H.ts0[b] = cur.ts[b];
  const int op_start = (int)prog.ops.size();
  const int coef_start = (int)(prog.coef.size() / 2);
  const int NR = 1 << c.RB;
  std::vector<Built> built;
  // Relayout through a half-tile LDS image: a tile bit h that is a register bit
  // of both layouts, at the same register position p, splits the exchange into
  // two rounds (registers with bit p = 0, then 1); each round writes 2^(K-1)
  // amplitudes and reads them back into the registers it just wrote, so the
  // LDS image needs only half the tile.  A flip of slot p carries over.
  auto relayout_one = [&](std::vector<int> nr) {
    int p = -1;
    for (int i = 0; i < c.RB && p < 0; ++i)
      if (index_in(nr, cur.rs[i]) >= 0) p = i;
    const int h = cur.rs[p];
    std::swap(nr[index_in(nr, h)], nr[p]);
    Built rl;
    std::memset(&rl.op, 0, sizeof(rl.op));
    rl.op.kind = F_RELAYOUT;
    rl.op.coef = (int)prog.aux.size();
    uint32_t lr[FMAX_RB], lt[FMAX_TB];
    prog.aux.push_back((uint32_t)p);
    Layout next = make_layout(nr, K, tord);
    const std::vector<uint32_t> pl = relayout_placement(cur, next, h, K, amp_bytes);
    slot_offsets_placed(cur, pl, amp_bytes, lr, lt);
    for (int b = 0; b < c.RB; ++b) prog.aux.push_back(lr[b]);
    for (int b = 0; b < c.TB; ++b) prog.aux.push_back(lt[b]);
    uint32_t fo = 0;
    const uint32_t fl = flip_local(cur);
    for (int b = 0; b < K; ++b)
      if (b != h && ((fl >> b) & 1)) fo ^= pl[b];
    prog.aux.push_back(fo * amp_bytes);
    const uint32_t fp = (cur.flip >> p) & 1u;
    cur = next;
    cur.flip = fp << p;
    slot_offsets_placed(cur, pl, amp_bytes, lr, lt);
    for (int b = 0; b < c.RB; ++b) prog.aux.push_back(lr[b]);
    for (int b = 0; b < c.TB; ++b) prog.aux.push_back(lt[b]);
    built.push_back(rl);
    H.uses_lds = 1;
  };
  auto relayout_to = [&](const std::vector<int>& nr) {
    bool shared = false;
    for (int b : cur.rs) shared = shared || index_in(nr, b) >= 0;
    if (!shared) {
      // no common register bit: go through a layout that keeps one old register bit
      std::vector<int> mid(nr.begin(), nr.end() - 1);
      mid.push_back(cur.rs[0]);
      relayout_one(mid);
    }
    relayout_one(nr);
  };
  auto add_built = [&](OpBuild& ob, const std::vector<int>& lb) {
    Built b;
    b.op = ob.op;
    b.coef = std::move(ob.coef);
    for (int bit : lb) b.slots |= 1u << slot_of(cur, bit);
    if (b.op.kind == F_DIAG)
      for (int k = 0; k < b.op.ntt; ++k) b.selp.push_back(b.op.tpos[k]);
    if (!try_merge(built, b, NR, c.RB)) built.push_back(std::move(b));
  };
  // ---- list scheduling of the pass's gates ----------------------------------
  // Gates j < k must stay in order when they share a bit that either of them
  // mixes (a non-diagonal target); bits both use diagonally (controls, bits of
  // diagonal gates) commute.  Among the ready gates, those whose targets are
  // already register bits run first (in circuit order); when none is, the tile
  // is re-laid-out for the ready gate whose targets are needed most by the
  // ready gates and the next ones.
  const size_t G = p.gates.size();
  std::vector<uint32_t> allm(G, 0), mixm(G, 0);
  for (size_t k = 0; k < G; ++k) {
    const qs_gate& g = gates[p.gates[k]];
    const int cc = g.ncontrols, tt = g.ntargets;
    for (int i = 0; i < cc + tt; ++i) allm[k] |= 1u << loc[k][i];
    if (!gate_diag_like(g))
      for (int i = 0; i < tt; ++i) mixm[k] |= 1u << loc[k][cc + i];
  }
  std::vector<std::vector<int>> succ(G);
  std::vector<int> indeg(G, 0);
  for (size_t k = 0; k < G; ++k)
    for (size_t j = 0; j < k; ++j)
      if ((allm[j] & allm[k]) & (mixm[j] | mixm[k])) {
        succ[j].push_back((int)k);
        ++indeg[k];
      }
  auto free_gate = [&](size_t k) {
    const qs_gate& g = gates[p.gates[k]];
    return relabel_swap(g, loc[k]) || (g.kind == QS_GATE_XPERM && g.ncontrols == 0);
  };
  auto fits = [&](size_t k) {
    if (free_gate(k)) return true;
    for (int b : needed_regs(gates[p.gates[k]], loc[k], cur))
      if (index_in(cur.rs, b) < 0) return false;
    return true;
  };
  auto emit = [&](size_t k) -> int {
    const qs_gate& g = gates[p.gates[k]];
    // a plain SWAP of two register slots also relabels: no bit moves between
    // registers and threads, so the load/store layouts are unaffected
    const bool reg_swap = g.kind == QS_GATE_SWAP && g.ncontrols == 0 && g.m[0] == 1.0 && g.m[1] == 0.0 &&
                          g.m[2] == 1.0 && g.m[3] == 0.0 && index_in(cur.rs, loc[k][0]) >= 0 &&
                          index_in(cur.rs, loc[k][1]) >= 0;
    if (relabel_swap(g, loc[k]) || reg_swap) {
      // the two slots exchange the tile bits they carry; flips stay with the slots
      const int a = loc[k][0], b = loc[k][1];
      auto& va = (index_in(cur.rs, a) >= 0) ? cur.rs : cur.ts;
      auto& vb = (index_in(cur.rs, b) >= 0) ? cur.rs : cur.ts;
      const int ia = index_in(va, a), ib = index_in(vb, b);
      va[ia] = b;
      vb[ib] = a;
      return QS_OK;
    }
    if (g.kind == QS_GATE_XPERM && g.ncontrols == 0) {
      // flip the slot; a0' = m0 a1, a1' = m1 a0 leaves diag(m0, m1) on the flipped bit
      const int b = loc[k][0];
      cur.flip ^= 1u << slot_of(cur, b);
      if (!(g.m[0] == 1.0 && g.m[1] == 0.0 && g.m[2] == 1.0 && g.m[3] == 0.0)) {
        qs_gate d;
        std::memset(&d, 0, sizeof(d));
        d.kind = QS_GATE_DIAG;
        d.ntargets = 1;
        d.bits[0] = g.bits[0];
        d.m[0] = g.m[0], d.m[1] = g.m[1], d.m[2] = g.m[2], d.m[3] = g.m[3];
        OpBuild ob;
        std::memset(&ob.op, 0, sizeof(ob.op));
        int rc = build_diag(d, loc[k], cur, NR, ob);
        if (rc) {
          set_error("fusion planner produced an op the tile kernel cannot encode");
          return rc;
        }
        add_built(ob, loc[k]);
      }
      return QS_OK;
    }
    if (!fits(k)) {
      set_error("fusion planner scheduled a gate whose targets are not register bits");
      return QS_ERR_UNSUPPORTED;
    }
    OpBuild ob;
    int rc = build_op(g, loc[k], cur, NR, ob);
    if (rc) {
      set_error("fusion planner produced an op the tile kernel cannot encode");
      return rc;
    }
    add_built(ob, loc[k]);
    return QS_OK;
  };
  std::vector<int> ready;
  for (size_t k = 0; k < G; ++k)
    if (indeg[k] == 0) ready.push_back((int)k);
  size_t done = 0;
  while (done < G) {
    std::sort(ready.begin(), ready.end());
    int pick = -1;
    for (int k : ready)
      if (fits((size_t)k)) {
        pick = k;
        break;
      }
    if (pick < 0) {
      // score tile bits: needed by ready gates (2), by gates in a lookahead window (1)
      std::vector<int> score(K, 0);
      std::vector<char> isready(G, 0);
      for (int k : ready) isready[k] = 1;
      for (int k : ready)
        for (int b : needed_regs(gates[p.gates[k]], loc[k], cur)) score[b] += 2;
      int window = 0;
      for (size_t k = 0; k < G && window < 24; ++k) {
        if (isready[k] || free_gate(k)) continue;
        bool pending = indeg[k] >= 0;  // not yet scheduled
        if (!pending) continue;
        for (int b : needed_regs(gates[p.gates[k]], loc[k], cur)) score[b] += 1;
        ++window;
      }
      int best = -1, best_score = -1;
      for (int k : ready) {
        const std::vector<int> need = needed_regs(gates[p.gates[k]], loc[k], cur);
        if ((int)need.size() > c.RB) continue;
        int sc = 0;
        for (int b : need) sc += score[b];
        if (sc > best_score) {
          best_score = sc;
          best = k;
        }
      }
      if (best < 0) {
        set_error("fusion planner: no schedulable gate");
        return QS_ERR_UNSUPPORTED;
      }
      std::vector<int> order(K);
      for (int b = 0; b < K; ++b) order[b] = b;
      std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return score[x] > score[y]; });
      std::vector<int> la;
      for (int b : order)
        if (score[b] > 0) la.push_back(b);
      relayout_to(choose_regs(needed_regs(gates[p.gates[best]], loc[best], cur), la, cur.rs, c.RB, K, c.L));
      pick = best;
    }
    int rc = emit((size_t)pick);
    if (rc) return rc;
    ready.erase(std::find(ready.begin(), ready.end(), pick));
    indeg[pick] = -1;  // scheduled
    ++done;
    for (int sidx : succ[pick])
      if (--indeg[sidx] == 0) ready.push_back(sidx);
  }
  // the store layout keeps at most LR of the tile bits stored to the low L
  // index bits in registers, and the others on the first thread bits
  int low_in_regs = 0;
  for (int b : cur.rs)
    if (out_low[b]) ++low_in_regs;
  bool low_lanes = true;
  for (int i = 0; i < c.L - low_in_regs && i < (int)cur.ts.size(); ++i) low_lanes = low_lanes && out_low[cur.ts[i]];
  if (p.move) {
    // the qubits on the first L lanes go to the low bits; the low bits' qubits
    // left elsewhere take the index bits those came from
    std::vector<char> on_low(K, 0);
    for (int i = 0; i < c.L; ++i) on_low[cur.ts[i]] = 1;
    std::vector<int> evicted;
    for (int j = 0; j < c.L; ++j)
      if (!on_low[j]) evicted.push_back(j);
    size_t e = 0;
    for (int i = 0; i < c.L; ++i) {
      const int m = cur.ts[i];
      p.tbo[m] = p.tbits[i];
      if (m >= c.L) p.tbo[evicted[e++]] = p.tbits[m];
    }
    for (int j = 0; j < K; ++j) H.tbo[j] = p.tbo[j];
  } else if (low_in_regs > c.LR || !low_lanes) {
    std::vector<int> keep;
    int kept_low = 0;
    for (int b : cur.rs)
      if (!out_low[b] || kept_low++ < c.LR) keep.push_back(b);
    for (int i = K - 1; i >= 0 && (int)keep.size() < c.RB; --i) {
      const int b = tord[i];
      if (!out_low[b] && index_in(keep, b) < 0) keep.push_back(b);
    }
    relayout_to(keep);
  }
  for (Built& b : built) {
    if (b.op.kind == F_DIAG) {
      finalize_diag(b, NR);
      if (b.op.creg == 0) continue;  // the merged product is the identity
    }
    if (b.op.kind == F_DENSE3 || b.op.kind == F_DENSE4) H.has_dk = 1;
    if (b.op.kind == F_DENSE1) {
      // uncontrolled with a real (merged) matrix: the kernel's half-cost form
      bool real = b.op.creg == 0 && b.op.cthr == 0;
      for (size_t i = 1; i < b.coef.size(); i += 2) real = real && b.coef[i] == 0.0;
      b.op.pure = real;
    }
    if (b.op.kind != F_RELAYOUT) {
      b.op.coef = (int)(prog.coef.size() / 2) - coef_start;
      prog.coef.insert(prog.coef.end(), b.coef.begin(), b.coef.end());
    }
    prog.ops.push_back(b.op);
  }
  for (int b = 0; b < c.RB; ++b) H.rs1[b] = cur.rs[b];
  for (int b = 0; b < c.TB; ++b) H.ts1[b] = cur.ts[b];
  // pending flips go into the store address
  const uint32_t fl = flip_local(cur);
  for (int j = 0; j < K; ++j)
    if ((fl >> j) & 1) H.sflip |= 1ull << p.tbo[j];
  bool same = H.sflip == 0;
  for (int j = 0; j < K; ++j) same = same && p.tbo[j] == p.tbits[j];
  for (int b = 0; b < c.RB; ++b) same = same && H.rs1[b] == H.rs0[b];
  for (int b = 0; b < c.TB; ++b) same = same && H.ts1[b] == H.ts0[b];
  H.sync_store = (!H.uses_lds && !same) ? 1 : 0;
  H.op0 = op_start;
  H.nops = (int)prog.ops.size() - op_start;
  H.coef0 = coef_start;
  H.ncoef = (int)(prog.coef.size() / 2) - coef_start;
  return QS_OK;
}

// Plan and schedule one program (with `remap`: moving qubits onto the low bits).
// `ready`, when given, sees each fused pass as soon as it is scheduled (its
// header is the last one; the plans before it are final, the pass itself is
// appended to prog.plans after the call); a nonzero return stops the build.
static int build_one(const qs_gate* gates, int ngates, int n, const Config& c, bool remap, uint32_t amp_bytes,
                     FusedProgram& prog, const std::function<int(PassPlan&)>& ready = {}) {
  int rc = QS_OK;
  auto on_pass = [&](PassPlan& p) {
    if (rc) return;
    PassHdr H;
    const int a0 = (int)prog.aux.size();
    rc = schedule_pass(prog.pgates.data(), p, c, amp_bytes, H, prog);
    if (rc) return;
    p.hdr = (int)prog.hdrs.size();
    prog.hdrs.push_back(H);
    prog.aux0.push_back(a0);
    if (ready) rc = ready(p);
  };
  plan_groups(gates, ngates, n, c, remap, prog.plans, prog.pgates, on_pass);
  if (rc) return rc;
  prog.hdr_of.assign(prog.plans.size(), -1);
  for (size_t i = 0; i < prog.plans.size(); ++i)
    if (prog.plans[i].fused) prog.hdr_of[i] = prog.plans[i].hdr;
  return QS_OK;
}

// The tile kernel's op switch treats every encoding outside these ranges as
// unreachable: a program is checked here before it can reach a launch.
static int check_passes(const FusedProgram& prog, const Config& c, int h0, int h1) {
  for (int hi = h0; hi < h1; ++hi) {
    const PassHdr& h = prog.hdrs[hi];
    for (int i = 0; i < h.nops; ++i) {
      const FOp& o = prog.ops[(size_t)h.op0 + i];
      bool ok = o.kind >= F_RELAYOUT && o.kind <= F_XX2;
      switch (o.kind) {
        case F_DENSE1:
        case F_PERM1:
          ok = o.pos[0] >= 0 && o.pos[0] < c.RB;
          break;
        case F_DIAG:
          ok = o.ntt >= 0 && o.ntt <= 2 && (o.ntt < 1 || (o.tpos[0] >= 0 && o.tpos[0] < c.TB)) &&
               (o.ntt < 2 || (o.tpos[1] >= 0 && o.tpos[1] < c.TB));
          break;
        case F_DENSE2:
        case F_SWAP2:
        case F_XX2:
          ok = o.pos[1] >= 0 && o.pos[0] > o.pos[1] && o.pos[0] < c.RB;
          break;
        case F_DENSE3:
        case F_DENSE4:
          ok = h.has_dk != 0;
          break;
        default:
          break;
      }
      if (!ok) {
        set_error("fusion planner: internal error: op outside the tile kernel's encodings");
        return QS_ERR_ARG;
      }
    }
  }
  return QS_OK;
}

static int check_program(const FusedProgram& prog, const Config& c) {
  return check_passes(prog, c, 0, (int)prog.hdrs.size());
}

static int check_config(const Config& c) {
  // every tile keeps its L low bits on thread slots at the load and the store
  if (c.L < 1 || c.L > c.TB) {
    set_error("fusion planner: low_bits must be in [1, tile_log2 - 4]");
    return QS_ERR_ARG;
  }
  return QS_OK;
}

static int build_program_unchecked(const qs_gate* gates, int ngates, int n, const Config& c, uint32_t amp_bytes,
                                   FusedProgram& prog) {
  int rc = check_config(c);
  if (rc) return rc;
  rc = build_one(gates, ngates, n, c, false, amp_bytes, prog);
  if (rc || c.remap == 0 || n < c.RB + c.TB + 2) return rc;
  // moving qubits onto the low bits pays when it saves enough passes (the
  // passes that put the qubits back included): the passes get heavier.  Short
  // programs (under 24 passes) rarely save 8%, and a second plan costs host time.
  if (c.remap == 1 && prog.plans.size() < 24) return QS_OK;
  FusedProgram moved;
  if (build_one(gates, ngates, n, c, true, amp_bytes, moved) != QS_OK) return QS_OK;
  const size_t a = moved.plans.size(), b = prog.plans.size();
  if (c.remap == 2 ? a < b : 100 * a <= 92 * b) prog = std::move(moved);
  return QS_OK;
}

static int build_program(const qs_gate* gates, int ngates, int n, const Config& c, uint32_t amp_bytes,
                         FusedProgram& prog) {
  const int rc = build_program_unchecked(gates, ngates, n, c, amp_bytes, prog);
  return rc ? rc : check_program(prog, c);
}

// The passes hdrs[h0, h1) of a program (possibly still being built) as a
// program of their own: headers rebased onto the chunk's ops, coefficients
// and relayout constants (a relayout op's coef is its first aux word).
static void chunk_program(const FusedProgram& prog, int h0, int h1, FusedProgram& sub) {
  const PassHdr& a = prog.hdrs[h0];
  const PassHdr& z = prog.hdrs[h1 - 1];
  const int o0 = a.op0, o1 = z.op0 + z.nops, c0 = a.coef0, c1 = z.coef0 + z.ncoef;
  const int x0 = prog.aux0[h0], x1 = h1 < (int)prog.aux0.size() ? prog.aux0[h1] : (int)prog.aux.size();
  sub.hdrs.assign(prog.hdrs.begin() + h0, prog.hdrs.begin() + h1);
  for (PassHdr& h : sub.hdrs) {
    h.op0 -= o0;
    h.coef0 -= c0;
  }
  sub.ops.assign(prog.ops.begin() + o0, prog.ops.begin() + o1);
  for (FOp& o : sub.ops)
    if (o.kind == F_RELAYOUT) o.coef -= x0;
  sub.aux.assign(prog.aux.begin() + x0, prog.aux.begin() + x1);
  sub.coef.assign(prog.coef.begin() + 2 * (size_t)c0, prog.coef.begin() + 2 * (size_t)c1);
}

// Plan without qubit moves (each pass is final once scheduled) and hand the
// passes over, in plan order, as they are planned: runs of fused passes as
// checked chunks of 1, 2, 4, ... 16 passes (`chunk(sub, first pass index)`, sub
// from chunk_program), passes the tile kernel cannot take one at a time
// (`single(gate record)`).  The device can start on the first chunk while the
// rest is planned.  A nonzero return of a callback stops the build.
static int build_streamed(const qs_gate* gates, int ngates, int n, const Config& c, uint32_t amp_bytes,
                          FusedProgram& prog, const std::function<int(const FusedProgram&, int)>& chunk,
                          const std::function<int(const qs_gate&)>& single) {
  int rc = check_config(c);
  if (rc) return rc;
  size_t next_plan = 0;  // plans before this one are handed over or in the open chunk
  int h0 = 0, h1 = 0;    // open chunk: headers [h0, h1)
  int want = 1;
  auto flush = [&]() -> int {
    if (h1 == h0) return QS_OK;
    int r = check_passes(prog, c, h0, h1);
    if (r) return r;
    FusedProgram sub;
    chunk_program(prog, h0, h1, sub);
    r = chunk(sub, h0);
    h0 = h1;
    want = std::min(2 * want, 16);
    return r;
  };
  auto singles = [&](size_t upto) -> int {  // single-gate passes planned since the last fused one
    for (; next_plan < upto; ++next_plan) {
      const PassPlan& q = prog.plans[next_plan];
      if (q.fused) continue;
      int r = flush();
      if (!r) r = single(prog.pgates[q.single]);
      if (r) return r;
    }
    return QS_OK;
  };
  auto ready = [&](PassPlan& p) -> int {
    const int r = singles(prog.plans.size());
    if (r) return r;
    next_plan = prog.plans.size() + 1;  // p itself is appended next
    h1 = p.hdr + 1;
    return h1 - h0 >= want ? flush() : QS_OK;
  };
  rc = build_one(gates, ngates, n, c, false, amp_bytes, prog, ready);
  if (!rc) rc = singles(prog.plans.size());
  if (!rc) rc = flush();
  return rc;
}

static void print_program_stats(const FusedProgram& prog, int n, int K, int L, int ngates) {
  // op counts by kind; dense1 / exchange by control mode (bit 0 register
  // controls, bit 1 thread controls); diagonals lane-mode vs table
  int nrel = 0, kinds[9] = {0}, d1cm[4] = {0}, p1cm[4] = {0}, dlane = 0, dtab = 0;
  for (const FOp& o : prog.ops) {
    if (o.kind == F_RELAYOUT) ++nrel;
    if (o.kind < 9) ++kinds[o.kind];
    const int cm = (o.cthr ? 2 : 0) | (o.creg ? 1 : 0);
    if (o.kind == F_DENSE1) ++d1cm[cm];
    if (o.kind == F_PERM1) ++p1cm[cm];
    if (o.kind == F_DIAG) (o.pure ? dlane : dtab)++;
  }
  int lds = 0;
  for (auto& h : prog.hdrs) lds += h.uses_lds;
  fprintf(stderr,
          "[qsim fusion] n=%d K=%d L=%d gates=%d passes=%zu fused=%zu with_lds=%d ops=%zu relayouts=%d "
          "dense1=%d perm1=%d diag=%d dense2=%d xx2=%d swap2=%d dense3=%d dense4=%d "
          "dense1_cm=%d,%d,%d,%d perm1_cm=%d,%d,%d,%d diag_lane=%d diag_table=%d\n",
          n, K, L, ngates, prog.plans.size(), prog.hdrs.size(), lds, prog.ops.size(), nrel, kinds[F_DENSE1],
          kinds[F_PERM1], kinds[F_DIAG], kinds[F_DENSE2], kinds[F_XX2], kinds[F_SWAP2], kinds[F_DENSE3],
          kinds[F_DENSE4], d1cm[0],
          d1cm[1], d1cm[2], d1cm[3], p1cm[0], p1cm[1], p1cm[2], p1cm[3], dlane, dtab);
}

}  // namespace qs
