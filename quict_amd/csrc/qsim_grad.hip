// Adjoint-method gradients (no reference counterpart: QuICT's QNN/VQE code
// lives in an external package).
//
// For E(theta) = <psi|H|psi>, psi = G_{L-1} ... G_0 |psi_0>, the caller hands in
// psi (forward state) and lam = H psi.  Walking j = L-1 .. 0 one kernel per gate
// reads the pair/quad of psi and lam amplitudes once and
//   psi <- G_j^dagger psi            (state before gate j, recomputed: nothing stored)
//   g_j += 2 Re <lam| dG_j |psi>     (block partials, reduced once at the end)
//   lam <- G_j^dagger lam
// so a whole gradient costs one read+write of both vectors per gate.  The
// per-gate records arrive already inverted (G_j^dagger) and, for parameterised
// gates, with the derivative dG_j/dtheta as a second record.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace qs {

constexpr int ADJ_THREADS = 256;
constexpr int ADJ_MAX_BLOCKS = 1024;

template <typename R> struct AdjArgs {
  uint64_t cmask;
  int nsorted;
  int sorted[QS_MAX_GATE_BITS];
  uint64_t off[4];
  cx<R> g[16];   // G^dagger, dense NA x NA
  cx<R> d[16];   // derivative, dense NA x NA
};

template <typename R, int NT, bool GRAD>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_step(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                          uint64_t ngroups, const AdjArgs<R> a,
                                                          double* __restrict__ partial) {
  constexpr int NA = 1 << NT;
  __shared__ double red[ADJ_THREADS / 64];
  double acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zeros(g, a.sorted, a.nsorted) | a.cmask;
    cx<R> p[NA], l[NA], pb[NA], lb[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      p[j] = ld(psi + (base | a.off[j]));
      l[j] = ld(lam + (base | a.off[j]));
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      cx<R> sp = mul(a.g[i * NA], p[0]), sl = mul(a.g[i * NA], l[0]);
#pragma unroll
      for (int j = 1; j < NA; ++j) {
        sp = cfma(sp, a.g[i * NA + j], p[j]);
        sl = cfma(sl, a.g[i * NA + j], l[j]);
      }
      pb[i] = sp;
      lb[i] = sl;
    }
    if (GRAD) {
      // Re <lam_after | D psi_before>
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        cx<R> mu = mul(a.d[i * NA], pb[0]);
#pragma unroll
        for (int j = 1; j < NA; ++j) mu = cfma(mu, a.d[i * NA + j], pb[j]);
        acc += (double)l[i].re * (double)mu.re + (double)l[i].im * (double)mu.im;
      }
    }
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      st(psi + (base | a.off[j]), pb[j]);
      st(lam + (base | a.off[j]), lb[j]);
    }
  }
  if (GRAD) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < ADJ_THREADS / 64; ++w) s += red[w];
      partial[blockIdx.x] = s;
    }
  }
}

// ---- several consecutive steps in one pass ------------------------------------
// A run of gates (in backward order) whose bits together span at most NB
// amplitude bits U: each thread loads groups of 2^NB amplitudes of psi and lam
// (the bits of U varying), applies every step of the run in registers and
// stores once.  Gate positions are local bit indices into U.
constexpr int MG_MAX_GATES = 8;
struct MGate {
  int32_t nt;          // 1 or 2 targets
  int32_t p0, p1;      // local positions of target 0 (matrix MSB) and target 1
  int32_t grad;        // parameterised: accumulate into partial row `row`
  int32_t row;
  uint32_t cmask, cval;  // local control bits and the values they need
  int32_t pad;
  double g[32];        // G^dagger, (re, im), 2^nt x 2^nt
  double d[32];        // dG/dtheta
};

template <typename R> __device__ __forceinline__ cx<R> cget(const double* m, int i) {
  return mk<R>((R)m[2 * i], (R)m[2 * i + 1]);
}

template <typename R, int NG, int P, bool GRAD>
__device__ __forceinline__ void mg_apply1(cx<R> (&p)[NG], cx<R> (&l)[NG], const MGate& G, double& acc) {
  const cx<R> g00 = cget<R>(G.g, 0), g01 = cget<R>(G.g, 1), g10 = cget<R>(G.g, 2), g11 = cget<R>(G.g, 3);
#pragma unroll
  for (int i = 0; i < NG; ++i) {
    if (i & (1 << P)) continue;
    const int i1 = i | (1 << P);
    if (((uint32_t)i & G.cmask) != G.cval) continue;
    const cx<R> a0 = cfma(mul(g00, p[i]), g01, p[i1]), a1 = cfma(mul(g10, p[i]), g11, p[i1]);
    const cx<R> b0 = cfma(mul(g00, l[i]), g01, l[i1]), b1 = cfma(mul(g10, l[i]), g11, l[i1]);
    if constexpr (GRAD) {
      const cx<R> m0 = cfma(mul(cget<R>(G.d, 0), a0), cget<R>(G.d, 1), a1);
      const cx<R> m1 = cfma(mul(cget<R>(G.d, 2), a0), cget<R>(G.d, 3), a1);
      acc += (double)l[i].re * (double)m0.re + (double)l[i].im * (double)m0.im +
             (double)l[i1].re * (double)m1.re + (double)l[i1].im * (double)m1.im;
    }
    p[i] = a0;
    p[i1] = a1;
    l[i] = b0;
    l[i1] = b1;
  }
}

template <typename R, int NG, int P0, int P1, bool GRAD>
__device__ __forceinline__ void mg_apply2(cx<R> (&p)[NG], cx<R> (&l)[NG], const MGate& G, double& acc) {
#pragma unroll
  for (int i = 0; i < NG; ++i) {
    if (i & ((1 << P0) | (1 << P1))) continue;
    if (((uint32_t)i & G.cmask) != G.cval) continue;
    const int idx[4] = {i, i | (1 << P1), i | (1 << P0), i | (1 << P0) | (1 << P1)};
    cx<R> a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      cx<R> sa = mul(cget<R>(G.g, 4 * r), p[idx[0]]), sb = mul(cget<R>(G.g, 4 * r), l[idx[0]]);
#pragma unroll
      for (int c = 1; c < 4; ++c) {
        sa = cfma(sa, cget<R>(G.g, 4 * r + c), p[idx[c]]);
        sb = cfma(sb, cget<R>(G.g, 4 * r + c), l[idx[c]]);
      }
      a[r] = sa;
      b[r] = sb;
    }
    if constexpr (GRAD) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cx<R> mu = mul(cget<R>(G.d, 4 * r), a[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) mu = cfma(mu, cget<R>(G.d, 4 * r + c), a[c]);
        acc += (double)l[idx[r]].re * (double)mu.re + (double)l[idx[r]].im * (double)mu.im;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      p[idx[r]] = a[r];
      l[idx[r]] = b[r];
    }
  }
}

template <typename R, int NB, bool GRAD>
__device__ __forceinline__ void mg_apply(cx<R> (&p)[1 << NB], cx<R> (&l)[1 << NB], const MGate& G, double& acc) {
  constexpr int NG = 1 << NB;
  if (G.nt == 1) {
    switch (G.p0) {
      case 0: mg_apply1<R, NG, 0, GRAD>(p, l, G, acc); break;
      case 1: if constexpr (NB > 1) mg_apply1<R, NG, 1, GRAD>(p, l, G, acc); break;
      case 2: if constexpr (NB > 2) mg_apply1<R, NG, 2, GRAD>(p, l, G, acc); break;
      default: break;
    }
  } else {
    switch (G.p0 * 4 + G.p1) {
      case 1: if constexpr (NB > 1) mg_apply2<R, NG, 0, 1, GRAD>(p, l, G, acc); break;
      case 4: if constexpr (NB > 1) mg_apply2<R, NG, 1, 0, GRAD>(p, l, G, acc); break;
      case 2: if constexpr (NB > 2) mg_apply2<R, NG, 0, 2, GRAD>(p, l, G, acc); break;
      case 8: if constexpr (NB > 2) mg_apply2<R, NG, 2, 0, GRAD>(p, l, G, acc); break;
      case 6: if constexpr (NB > 2) mg_apply2<R, NG, 1, 2, GRAD>(p, l, G, acc); break;
      case 9: if constexpr (NB > 2) mg_apply2<R, NG, 2, 1, GRAD>(p, l, G, acc); break;
      default: break;
    }
  }
}

struct MPass {
  int32_t ubits[3];   // amplitude bits of U, ascending (local bit b = ubits[b])
  int32_t nb;
  int32_t g0, ng;     // gates [g0, g0 + ng) of the descriptor table, in application order
};

template <typename R, int NB>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_multi(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                           uint64_t ngroups, const MPass P,
                                                           const MGate* __restrict__ gates,
                                                           double* __restrict__ partial) {
  constexpr int NG = 1 << NB;
  __shared__ double red[MG_MAX_GATES][ADJ_THREADS / 64];
  double acc[MG_MAX_GATES];
#pragma unroll
  for (int k = 0; k < MG_MAX_GATES; ++k) acc[k] = 0.0;
  uint64_t off[NG];
#pragma unroll
  for (int i = 0; i < NG; ++i) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if ((i >> b) & 1) o |= 1ull << P.ubits[b];
    off[i] = o;
  }
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t base = insert_zeros(g, P.ubits, NB);
    cx<R> p[NG], l[NG];
#pragma unroll
    for (int i = 0; i < NG; ++i) {
      p[i] = ld(psi + (base | off[i]));
      l[i] = ld(lam + (base | off[i]));
    }
#pragma unroll
    for (int k = 0; k < MG_MAX_GATES; ++k) {
      if (k >= P.ng) break;
      const MGate& G = gates[P.g0 + k];
      if (G.grad) mg_apply<R, NB, true>(p, l, G, acc[k]);
      else mg_apply<R, NB, false>(p, l, G, acc[k]);
    }
#pragma unroll
    for (int i = 0; i < NG; ++i) {
      st(psi + (base | off[i]), p[i]);
      st(lam + (base | off[i]), l[i]);
    }
  }
#pragma unroll
  for (int k = 0; k < MG_MAX_GATES; ++k) {
    if (k >= P.ng) break;
    double a = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  if (threadIdx.x < MG_MAX_GATES && (int)threadIdx.x < P.ng) {
    const MGate& G = gates[P.g0 + threadIdx.x];
    if (G.grad) {
      double s = 0;
      for (int w = 0; w < ADJ_THREADS / 64; ++w) s += red[threadIdx.x][w];
      partial[(uint64_t)G.row * ADJ_MAX_BLOCKS + blockIdx.x] = s;
    }
  }
}

// ---- many consecutive steps on a tile of both vectors in LDS -----------------
// A pass: up to AT_MAX_STEPS backward steps (AT_MAX_GRADS of them
// parameterised) whose bits, together with the AT_LOW lowest amplitude bits,
// fit in a tile of nb <= AT_MAX_BITS amplitude bits.  A workgroup loads the
// tile of psi and lam into LDS (runs of 2^AT_LOW contiguous amplitudes),
// applies every step there (one workgroup barrier per step), reduces each
// parameterised step's Re <lam|dG|psi> per wave into LDS and stores the tile
// once.  (Runs of steps on <= 3 bits applied in registers per LDS round trip
// were slower: half the threads idle on 1024-amplitude tiles.)
constexpr int AT_MAX_BITS = 11;
constexpr int AT_MAX_STEPS = 128;
constexpr int AT_MAX_GRADS = 64;
constexpr int AT_LOW = 3;

struct TStep {
  int32_t nt;            // 1 or 2 targets
  int32_t p0, p1;        // tile-local positions of target 0 (matrix MSB) and target 1
  int32_t gi;            // parameterised: accumulator slot in the pass, else -1
  uint32_t cmask, cval;  // tile-local control bits and the values they need
  int32_t twice;         // one target: the next step of the pass (the gate before, in circuit
                         // order), on the same bit and controls, shares the step (its G^dagger
                         // in g[8..15], dG in d[8..15]): one LDS round trip for both
  int32_t gi2;           // its accumulator slot, else -1
  double g[32];          // G^dagger, (re, im), 2^nt x 2^nt
  double d[32];          // dG/dtheta
};

struct TPass {
  int32_t ubits[AT_MAX_BITS];  // amplitude bits of the tile, ascending (tile-local bit b = ubits[b])
  int32_t nb;
  int32_t s0, ns;              // steps [s0, s0 + ns) of the step table, in sweep (backward) order
  int32_t ngrad;
  int32_t row[AT_MAX_GRADS];   // partial row of accumulator slot k
};

// LDS slot of tile amplitude i: the low 4 index bits XOR index bits 4-7 and
// 8-11 (a bijection), so that lanes whose pairs are spaced 2, 4 or 8
// amplitudes apart spread over the banks.
__device__ __forceinline__ uint32_t swz(uint32_t i) { return i ^ (((i >> 4) ^ (i >> 8)) & 15u); }

// Wave sum by DPP row shifts and row broadcasts (lane moves in the VALU, no
// LDS round trips): lanes whose source is outside the row add 0; lane 63
// ends with the sum of all 64 lanes.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xf, false),
                          __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ double wave_sum_to_lane63(double v) {
  v += dpp_move<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf>(v);  // row_shr:8: lane 15 of each row holds the row's sum
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Insert a zero bit at local position p.
__device__ __forceinline__ uint32_t ins0(uint32_t q, int p) {
  const uint32_t lo = q & ((1u << p) - 1u);
  return ((q - lo) << 1) | lo;
}

template <typename R>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_tile(cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                          uint64_t ntiles, const TPass P,
                                                          const TStep* __restrict__ steps,
                                                          double* __restrict__ partial) {
  extern __shared__ __align__(16) char at_lds[];
  const int nb = P.nb;
  const uint32_t NA = 1u << nb;
  cx<R>* sp = reinterpret_cast<cx<R>*>(at_lds);
  cx<R>* sl = sp + NA;
  double* red = reinterpret_cast<double*>(sl + NA);  // [AT_MAX_GRADS][ADJ_THREADS / 64]
  const uint32_t tid = threadIdx.x;
  const int wave = (int)(tid >> 6);
  constexpr int NW = ADJ_THREADS / 64;
  for (int i = (int)tid; i < AT_MAX_GRADS * NW; i += ADJ_THREADS) red[i] = 0.0;
  // address of local index i = tid + 256 k: deposit of the low 8 bits (per
  // thread, once) OR deposit of the bits above
  uint64_t dlo = 0;
  for (int b = 0; b < nb && b < 8; ++b)
    if ((tid >> b) & 1u) dlo |= 1ull << P.ubits[b];
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = insert_zeros(t, P.ubits, nb);
    __syncthreads();  // the previous tile's stores have read the LDS image
    for (uint32_t i = tid; i < NA; i += ADJ_THREADS) {
      uint64_t a = base | dlo;
      for (int b = 8; b < nb; ++b)
        if ((i >> b) & 1u) a |= 1ull << P.ubits[b];
      sp[swz(i)] = ld(psi + a);
      sl[swz(i)] = ld(lam + a);
    }
    __syncthreads();
    for (int k = 0; k < P.ns; ++k) {
      const TStep& S = steps[P.s0 + k];
      double acc = 0.0;
      const bool grad = S.gi >= 0;
      double acc2 = 0.0;
      const bool twice = S.twice != 0, grad2 = twice && S.gi2 >= 0;
      if (S.nt == 1) {
        const cx<R> g00 = cget<R>(S.g, 0), g01 = cget<R>(S.g, 1), g10 = cget<R>(S.g, 2), g11 = cget<R>(S.g, 3);
        for (uint32_t q = tid; q < NA / 2; q += ADJ_THREADS) {
          const uint32_t i0 = ins0(q, S.p0), i1 = i0 | (1u << S.p0);
          if ((i0 & S.cmask) != S.cval) continue;
          const uint32_t j0 = swz(i0), j1 = swz(i1);
          const cx<R> a0 = sp[j0], a1 = sp[j1], b0 = sl[j0], b1 = sl[j1];
          cx<R> pa = cfma(mul(g00, a0), g01, a1), pb = cfma(mul(g10, a0), g11, a1);
          cx<R> la = cfma(mul(g00, b0), g01, b1), lb = cfma(mul(g10, b0), g11, b1);
          if (grad) {
            const cx<R> m0 = cfma(mul(cget<R>(S.d, 0), pa), cget<R>(S.d, 1), pb);
            const cx<R> m1 = cfma(mul(cget<R>(S.d, 2), pa), cget<R>(S.d, 3), pb);
            acc += (double)b0.re * (double)m0.re + (double)b0.im * (double)m0.im + (double)b1.re * (double)m1.re +
                   (double)b1.im * (double)m1.im;
          }
          if (twice) {  // the following step, on the same pair
            const cx<R> h00 = cget<R>(S.g, 4), h01 = cget<R>(S.g, 5), h10 = cget<R>(S.g, 6), h11 = cget<R>(S.g, 7);
            const cx<R> qa = cfma(mul(h00, pa), h01, pb), qb = cfma(mul(h10, pa), h11, pb);
            if (grad2) {
              const cx<R> m0 = cfma(mul(cget<R>(S.d, 4), qa), cget<R>(S.d, 5), qb);
              const cx<R> m1 = cfma(mul(cget<R>(S.d, 6), qa), cget<R>(S.d, 7), qb);
              acc2 += (double)la.re * (double)m0.re + (double)la.im * (double)m0.im + (double)lb.re * (double)m1.re +
                      (double)lb.im * (double)m1.im;
            }
            const cx<R> ma = cfma(mul(h00, la), h01, lb), mb = cfma(mul(h10, la), h11, lb);
            pa = qa;
            pb = qb;
            la = ma;
            lb = mb;
          }
          sp[j0] = pa;
          sp[j1] = pb;
          sl[j0] = la;
          sl[j1] = lb;
        }
      } else {
        const int plo = S.p0 < S.p1 ? S.p0 : S.p1, phi = S.p0 < S.p1 ? S.p1 : S.p0;
        for (uint32_t q = tid; q < NA / 4; q += ADJ_THREADS) {
          const uint32_t i = ins0(ins0(q, plo), phi);
          if ((i & S.cmask) != S.cval) continue;
          const uint32_t idx[4] = {swz(i), swz(i | (1u << S.p1)), swz(i | (1u << S.p0)),
                                   swz(i | (1u << S.p0) | (1u << S.p1))};
          cx<R> a[4], b[4], pa[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            a[r] = sp[idx[r]];
            b[r] = sl[idx[r]];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            cx<R> x = mul(cget<R>(S.g, 4 * r), a[0]), y = mul(cget<R>(S.g, 4 * r), b[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) {
              x = cfma(x, cget<R>(S.g, 4 * r + c), a[c]);
              y = cfma(y, cget<R>(S.g, 4 * r + c), b[c]);
            }
            pa[r] = x;
            sp[idx[r]] = x;
            sl[idx[r]] = y;
          }
          if (grad) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              cx<R> mu = mul(cget<R>(S.d, 4 * r), pa[0]);
#pragma unroll
              for (int c = 1; c < 4; ++c) mu = cfma(mu, cget<R>(S.d, 4 * r + c), pa[c]);
              acc += (double)b[r].re * (double)mu.re + (double)b[r].im * (double)mu.im;
            }
          }
        }
      }
      if (grad) {
        acc = wave_sum_to_lane63(acc);
        if ((tid & 63u) == 63u) red[S.gi * NW + wave] += acc;  // one writer per (slot, wave)
      }
      if (grad2) {
        acc2 = wave_sum_to_lane63(acc2);
        if ((tid & 63u) == 63u) red[S.gi2 * NW + wave] += acc2;
      }
      __syncthreads();
    }
    for (uint32_t i = tid; i < NA; i += ADJ_THREADS) {
      uint64_t a = base | dlo;
      for (int b = 8; b < nb; ++b)
        if ((i >> b) & 1u) a |= 1ull << P.ubits[b];
      st(psi + a, sp[swz(i)]);
      st(lam + a, sl[swz(i)]);
    }
  }
  __syncthreads();
  if ((int)tid < P.ngrad) {
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[tid * NW + w];
    partial[(uint64_t)P.row[tid] * ADJ_MAX_BLOCKS + blockIdx.x] = s;
  }
}

// conj(m[i]): entries of a conjugate transpose.
template <typename R> __device__ __forceinline__ cx<R> cgetc(const double* m, int i) {
  return mk<R>((R)m[2 * i], -(R)m[2 * i + 1]);
}

// The forward state from the sweep's own plan: the tile passes (planned in
// backward order, as the sweep runs them) run last first, each applying its
// steps last first with G = (G^dagger)^dagger (the conjugate transpose of the
// step's matrix), on psi alone.  The grouping only exchanges commuting gates,
// so both directions compute the circuit's U.
template <typename R>
__global__ __launch_bounds__(ADJ_THREADS) void k_fwd_tile(cx<R>* __restrict__ psi, uint64_t ntiles, const TPass P,
                                                          const TStep* __restrict__ steps) {
  extern __shared__ __align__(16) char at_lds[];
  const int nb = P.nb;
  const uint32_t NA = 1u << nb;
  cx<R>* sp = reinterpret_cast<cx<R>*>(at_lds);
  const uint32_t tid = threadIdx.x;
  uint64_t dlo = 0;
  for (int b = 0; b < nb && b < 8; ++b)
    if ((tid >> b) & 1u) dlo |= 1ull << P.ubits[b];
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = insert_zeros(t, P.ubits, nb);
    __syncthreads();  // the previous tile's stores have read the LDS image
    for (uint32_t i = tid; i < NA; i += ADJ_THREADS) {
      uint64_t a = base | dlo;
      for (int b = 8; b < nb; ++b)
        if ((i >> b) & 1u) a |= 1ull << P.ubits[b];
      sp[swz(i)] = ld(psi + a);
    }
    __syncthreads();
    for (int k = P.ns - 1; k >= 0; --k) {  // the pass's steps are in sweep order
      const TStep& S = steps[P.s0 + k];
      if (S.nt == 1) {
        for (uint32_t q = tid; q < NA / 2; q += ADJ_THREADS) {
          const uint32_t i0 = ins0(q, S.p0), i1 = i0 | (1u << S.p0);
          if ((i0 & S.cmask) != S.cval) continue;
          const uint32_t j0 = swz(i0), j1 = swz(i1);
          cx<R> a0 = sp[j0], a1 = sp[j1];
          if (S.twice) {  // g[8..15] holds the earlier gate of the pair
            const cx<R> c0 = cfma(mul(cgetc<R>(S.g, 4), a0), cgetc<R>(S.g, 6), a1);
            const cx<R> c1 = cfma(mul(cgetc<R>(S.g, 5), a0), cgetc<R>(S.g, 7), a1);
            a0 = c0;
            a1 = c1;
          }
          sp[j0] = cfma(mul(cgetc<R>(S.g, 0), a0), cgetc<R>(S.g, 2), a1);
          sp[j1] = cfma(mul(cgetc<R>(S.g, 1), a0), cgetc<R>(S.g, 3), a1);
        }
      } else {
        const int plo = S.p0 < S.p1 ? S.p0 : S.p1, phi = S.p0 < S.p1 ? S.p1 : S.p0;
        for (uint32_t q = tid; q < NA / 4; q += ADJ_THREADS) {
          const uint32_t i = ins0(ins0(q, plo), phi);
          if ((i & S.cmask) != S.cval) continue;
          const uint32_t idx[4] = {swz(i), swz(i | (1u << S.p1)), swz(i | (1u << S.p0)),
                                   swz(i | (1u << S.p0) | (1u << S.p1))};
          cx<R> a[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) a[r] = sp[idx[r]];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            cx<R> x = mul(cgetc<R>(S.g, r), a[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) x = cfma(x, cgetc<R>(S.g, 4 * c + r), a[c]);
            sp[idx[r]] = x;
          }
        }
      }
      __syncthreads();
    }
    for (uint32_t i = tid; i < NA; i += ADJ_THREADS) {
      uint64_t a = base | dlo;
      for (int b = 8; b < nb; ++b)
        if ((i >> b) & 1u) a |= 1ull << P.ubits[b];
      st(psi + a, sp[swz(i)]);
    }
  }
}

// sums[k] = sum_b partial[k * stride + b], b < nb[k]
__global__ void k_adj_reduce(const double* __restrict__ partial, const int* __restrict__ nb, int stride, int nsteps,
                             double* __restrict__ sums) {
  const int k = blockIdx.x;
  if (k >= nsteps) return;
  double s = 0;
  for (int b = threadIdx.x; b < nb[k]; b += blockDim.x) s += partial[(uint64_t)k * stride + b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) sums[k] = red[0] + red[1] + red[2] + red[3];
}

// Dense 2^t x 2^t form of a record with t <= 2 (controls kept separately).
// Dense (re, im) matrix of a record with at most 2 targets into M[2 * 4^t].
static void record_dense(const qs_gate& g, double* M) {
  const int t = g.ntargets, D = 1 << t;
  for (int i = 0; i < 2 * D * D; ++i) M[i] = 0.0;
  auto set = [&](int i, int j, double re, double im) {
    M[2 * (i * D + j)] = re;
    M[2 * (i * D + j) + 1] = im;
  };
  switch (g.kind) {
    case QS_GATE_DENSE:
      for (int i = 0; i < 2 * D * D; ++i) M[i] = g.m[i];
      break;
    case QS_GATE_DIAG:
      for (int i = 0; i < D; ++i) set(i, i, g.m[2 * i], g.m[2 * i + 1]);
      break;
    case QS_GATE_PHASE:
      for (int i = 0; i < D - 1; ++i) set(i, i, 1, 0);
      set(D - 1, D - 1, g.m[0], g.m[1]);
      break;
    case QS_GATE_XPERM:
      set(0, 1, g.m[0], g.m[1]);
      set(1, 0, g.m[2], g.m[3]);
      break;
    case QS_GATE_SWAP:
      set(0, 0, 1, 0);
      set(3, 3, 1, 0);
      set(1, 2, g.m[0], g.m[1]);
      set(2, 1, g.m[2], g.m[3]);
      break;
  }
}

static void record_dense(const qs_gate& g, std::vector<double>& M) {
  M.assign(2 << (2 * g.ntargets), 0.0);
  record_dense(g, M.data());
}

// terms != nullptr: lam is computed here as H psi (and energy = <psi|H psi>)
// after the sweep is planned, so the planning overlaps the forward passes
// still running on the stream, and the energy comes back with the gradient
// (one host wait).
template <typename R>
static int adjoint_t(qs_ctx* ctx, cx<R>* psi, cx<R>* lam, int n, const qs_gate* gates, int ngates,
                     const qs_gate* dgates, const int32_t* slot, double* grads, int nslots, hipStream_t s,
                     const qs_pauli_term* terms = nullptr, int nterms = 0, double* energy = nullptr,
                     const qs_gate* fwd = nullptr) {
  int rc = QS_OK;
  std::vector<int> nb;
  std::vector<int> step_slot;
  // Plan (backward order): runs of consecutive steps become one pass: tile
  // passes in LDS (default: up to AT_MAX_STEPS steps whose bits fit in a tile
  // of `tile_bits` amplitude bits with the AT_LOW lowest ones), or register
  // passes spanning <= 3 amplitude bits (QS_ADJ_TILE=0); other gates run as
  // single steps.
  struct Item {
    int multi;   // 1: register pass index into passes, 2: tile pass index into tpasses, 0: single gate index
    int idx;
  };
  std::vector<Item> items;
  std::vector<MPass> passes;
  std::vector<MGate> mg;
  std::vector<TPass> tpasses;
  std::vector<TStep> ts;
  const char* tenv = getenv("QS_ADJ_TILE");
  const bool tile = !(tenv && tenv[0] == '0');
  int tile_bits = std::min(n, sizeof(R) == 8 ? 10 : 11);
  if (const char* tb = getenv("QS_ADJ_TILE_BITS")) tile_bits = std::min(n, std::max(AT_LOW + 2, std::min(AT_MAX_BITS, atoi(tb))));
  const uint64_t lowmask = (1ull << std::min(n, AT_LOW)) - 1;
  std::vector<int> cur;  // gate indices of the open pass, in application order
  uint64_t curbits = 0;
  std::vector<uint64_t> gmask(ngates);
  for (int j = 0; j < ngates; ++j) {
    uint64_t m = 0;
    for (int i = 0; i < gates[j].ncontrols + gates[j].ntargets; ++i) m |= 1ull << gates[j].bits[i];
    gmask[j] = m;
  }
  auto gbits = [&](int j) { return gmask[j]; };
  ts.reserve(ngates);
  const int dt = sizeof(R) == 8 ? QS_C128 : QS_C64;
  if (fwd && !tile) {
    // the forward state by the fused passes, before this call takes a program
    // slot (their uploads cycle through the slots)
    rc = qs_sv_init(ctx, psi, n, dt, 0, s);
    if (!rc) rc = apply_gates_fused(ctx, psi, n, dt, fwd, ngates, false, nullptr, s);
    if (rc) return rc;
  }
  // Every device buffer of the call lives in one pinned program slot, sized
  // before planning (at most one step or register-pass gate per gate, one
  // partial row per parameterised gate): the descriptors (pass steps, block
  // counts; uploaded from pinned memory, since a copy from pageable memory
  // would wait for forward passes still queued), the partial rows, sums and
  // the energy's partials (nothing else the call runs can move or overwrite
  // them: per-gate passes use the context's scratch); the slot's host side
  // takes the results back.
  int nparam = 0;
  for (int j = 0; j < ngates; ++j) nparam += slot[j] >= 0;
  const uint64_t rows = std::max(nparam, 1);
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t ts_off = al((tile ? 0 : (uint64_t)ngates) * sizeof(MGate));
  const uint64_t nb_off = al(ts_off + (tile ? (uint64_t)ngates : 0) * sizeof(TStep));
  const uint64_t up = nb_off + rows * sizeof(int), res_off = al(up);
  const uint64_t part_off = al(up), sums_off = al(part_off + rows * ADJ_MAX_BLOCKS * sizeof(double));
  const uint64_t ip_off = al(sums_off + rows * sizeof(double));
  const uint64_t dev_end = ip_off + (2 * (uint64_t)inner_partial_blocks() + 64) * sizeof(double);
  qs_ctx::ProgSlot* pslot = nullptr;
  rc = acquire_prog_slot(ctx, res_off + (rows + 2) * sizeof(double), dev_end, &pslot);
  if (rc) return rc;
  struct SlotGuard {  // the slot is busy until everything enqueued below has run
    qs_ctx::ProgSlot* sl;
    hipStream_t st;
    ~SlotGuard() {
      (void)hipEventRecord(sl->done, st);
      sl->pending = true;
    }
  } guard{pslot, s};
  char* hs = pslot->host;
  char* ds = (char*)pslot->dev;
  const MGate* d_mg = (const MGate*)ds;
  const TStep* d_ts = (const TStep*)(ds + ts_off);
  const int* d_nb = (const int*)(ds + nb_off);
  double* partial = (double*)(ds + part_off);
  double* sums = (double*)(ds + sums_off);
  double* ipart = (double*)(ds + ip_off);
  double* ires = ipart + 2 * inner_partial_blocks();
  double* h_res = (double*)(hs + res_off);  // grads rows, then the energy (re, im)
  auto eligible = [&](int j) {
    const qs_gate& g = gates[j];
    if (g.ntargets > 2) return false;
    if (tile ? __builtin_popcountll(gbits(j) | lowmask) > tile_bits : g.ncontrols + g.ntargets > 3) return false;
    if (slot[j] >= 0 && (dgates[j].ntargets != g.ntargets || dgates[j].ncontrols != g.ncontrols)) return false;
    return true;
  };
  auto flush_tile = [&]() {
    TPass P;
    std::memset(&P, 0, sizeof(P));
    uint64_t u = curbits | lowmask;
    for (int b = 0; b < n && __builtin_popcountll(u) < tile_bits; ++b) u |= 1ull << b;  // fill: longer runs
    int nbits = 0;
    for (int b = 0; b < n; ++b)
      if ((u >> b) & 1) P.ubits[nbits++] = b;
    P.nb = nbits;
    P.s0 = (int)ts.size();
    const uint64_t ntiles = 1ull << (n - nbits);
    const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, ntiles);
    auto local = [&](int bit) {
      for (int i = 0; i < nbits; ++i)
        if (P.ubits[i] == bit) return i;
      return -1;
    };
    for (int j : cur) {
      const qs_gate& g = gates[j];
      TStep S;
      std::memset(&S, 0, sizeof(S));
      S.nt = g.ntargets;
      S.p0 = local(g.bits[g.ncontrols]);
      S.p1 = g.ntargets > 1 ? local(g.bits[g.ncontrols + 1]) : 0;
      S.gi = -1;
      S.gi2 = -1;
      for (int i = 0; i < g.ncontrols; ++i) {
        S.cmask |= 1u << local(g.bits[i]);
        S.cval |= 1u << local(g.bits[i]);
      }
      double M[32], Dm[32];
      const int msz = 2 << (2 * g.ntargets);
      record_dense(g, M);
      if (slot[j] >= 0) record_dense(dgates[j], Dm);
      // a one-target step on the bit and controls of the previous one-target
      // step of the pass joins it (same pairs: one LDS round trip)
      TStep* prev = (int)ts.size() > P.s0 ? &ts.back() : nullptr;
      if (S.nt == 1 && prev && prev->nt == 1 && !prev->twice && prev->p0 == S.p0 && prev->cmask == S.cmask &&
          prev->cval == S.cval) {
        prev->twice = 1;
        for (int i = 0; i < msz; ++i) prev->g[8 + i] = M[i];
        if (slot[j] >= 0) {
          for (int i = 0; i < msz; ++i) prev->d[8 + i] = Dm[i];
          prev->gi2 = P.ngrad;
          P.row[P.ngrad++] = (int)nb.size();
          nb.push_back(blocks);
          step_slot.push_back(slot[j]);
        }
        continue;
      }
      for (int i = 0; i < msz; ++i) S.g[i] = M[i];
      if (slot[j] >= 0) {
        for (int i = 0; i < msz; ++i) S.d[i] = Dm[i];
        S.gi = P.ngrad;
        P.row[P.ngrad++] = (int)nb.size();
        nb.push_back(blocks);
        step_slot.push_back(slot[j]);
      }
      ts.push_back(S);
    }
    P.ns = (int)ts.size() - P.s0;
    items.push_back({2, (int)tpasses.size()});
    tpasses.push_back(P);
    cur.clear();
    curbits = 0;
  };
  auto flush = [&]() {
    if (cur.empty()) return;
    MPass P;
    std::memset(&P, 0, sizeof(P));
    int nbits = 0;
    for (int b = 0; b < 64; ++b)
      if ((curbits >> b) & 1) P.ubits[nbits++] = b;
    P.nb = nbits;
    P.g0 = (int)mg.size();
    P.ng = (int)cur.size();
    const uint64_t ngroups = 1ull << (n - nbits);
    const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, (ngroups + ADJ_THREADS - 1) / ADJ_THREADS);
    auto local = [&](int bit) {
      for (int i = 0; i < nbits; ++i)
        if (P.ubits[i] == bit) return i;
      return -1;
    };
    for (int j : cur) {
      const qs_gate& g = gates[j];
      MGate G;
      std::memset(&G, 0, sizeof(G));
      G.nt = g.ntargets;
      G.p0 = local(g.bits[g.ncontrols]);
      G.p1 = g.ntargets > 1 ? local(g.bits[g.ncontrols + 1]) : 0;
      for (int i = 0; i < g.ncontrols; ++i) {
        G.cmask |= 1u << local(g.bits[i]);
        G.cval |= 1u << local(g.bits[i]);
      }
      std::vector<double> M;
      record_dense(g, M);
      for (size_t i = 0; i < M.size(); ++i) G.g[i] = M[i];
      if (slot[j] >= 0) {
        record_dense(dgates[j], M);
        for (size_t i = 0; i < M.size(); ++i) G.d[i] = M[i];
        G.grad = 1;
        G.row = (int)nb.size();
        nb.push_back(blocks);
        step_slot.push_back(slot[j]);
      }
      mg.push_back(G);
    }
    items.push_back({1, (int)passes.size()});
    passes.push_back(P);
    cur.clear();
    curbits = 0;
  };
  std::vector<int> single_row(ngates, -1);
  // Tile passes group the steps like the forward fusion planner: within a run
  // of eligible steps (backward order), a step may join the current pass ahead
  // of steps still waiting if it shares no bit with them (disjoint gates
  // commute, so the same U and the gradient of the same function come out),
  // fewest new tile bits first; the pass applies its steps in backward order.
  // (Planned in circuit order instead, for a forward launched while the rest
  // is planned, the 20-qubit sweep needed 31 passes of 44.9 us against 32 of
  // 42.0 us.)
  std::vector<int> run;
  const uint64_t all_bits = n >= 64 ? ~0ull : (1ull << n) - 1;
  auto plan_run = [&]() -> int {
    std::vector<char> done(run.size(), 0);
    size_t left = run.size(), first = 0;  // steps before `first` are all placed
    while (left) {
      uint64_t S = lowmask;
      int ng = 0, np = 0;
      std::vector<char> pick(run.size(), 0);
      while (first < run.size() && done[first]) ++first;
      while (np < AT_MAX_STEPS) {
        uint64_t wait = 0;
        int best = -1, best_new = 1 << 30;
        for (size_t k = first; k < run.size() && wait != all_bits; ++k) {  // all bits waiting: nothing may join
          if (done[k] || pick[k]) continue;
          const int g = run[k];
          const uint64_t m = gbits(g);
          if (!(m & wait) && __builtin_popcountll(S | m) <= tile_bits && (slot[g] < 0 || ng < AT_MAX_GRADS)) {
            const int nn = __builtin_popcountll(m & ~S);
            if (nn < best_new) {
              best_new = nn;
              best = (int)k;
              if (nn == 0) break;
            }
          }
          wait |= m;
        }
        if (best < 0) break;
        pick[best] = 1;
        ++np;
        S |= gbits(run[best]);
        if (slot[run[best]] >= 0) ++ng;
      }
      cur.clear();
      curbits = 0;
      for (size_t k = 0; k < run.size(); ++k)
        if (pick[k]) {
          cur.push_back(run[k]);
          curbits |= gbits(run[k]);
          done[k] = 1;
          --left;
        }
      flush_tile();
    }
    run.clear();
    return QS_OK;
  };
  for (int j = ngates - 1; j >= 0; --j) {
    if (tile && eligible(j)) {
      run.push_back(j);
      continue;
    }
    if (tile && (rc = plan_run())) return rc;
    if (!eligible(j)) {
      flush();
      if (slot[j] >= 0) {
        const qs_gate& g = gates[j];
        if (g.ntargets > 2 || dgates[j].ntargets != g.ntargets || dgates[j].ncontrols != g.ncontrols) {
          set_error("adjoint gradient: parameterised gates must act on <= 2 targets");
          return QS_ERR_UNSUPPORTED;
        }
        const uint64_t ngroups = 1ull << (n - g.ncontrols - g.ntargets);
        single_row[j] = (int)nb.size();
        nb.push_back((int)std::min<uint64_t>(ADJ_MAX_BLOCKS, (ngroups + ADJ_THREADS - 1) / ADJ_THREADS));
        step_slot.push_back(slot[j]);
      }
      items.push_back({0, j});
      continue;
    }
    const uint64_t b = gbits(j);
    if (!cur.empty() && (__builtin_popcountll(curbits | b) > 3 || (int)cur.size() >= MG_MAX_GATES)) flush();
    cur.push_back(j);
    curbits |= b;
  }
  if (tile && (rc = plan_run())) return rc;
  flush();
  if ((int)nb.size() != nparam || mg.size() > (size_t)(tile ? 0 : ngates) || ts.size() > (size_t)ngates) {
    set_error("adjoint gradient: internal error: plan larger than its buffers");
    return QS_ERR_DEVICE;
  }
  // the descriptors: register-pass gates, steps, block counts
  if (!mg.empty()) std::memcpy(hs, mg.data(), mg.size() * sizeof(MGate));
  if (!ts.empty()) std::memcpy(hs + ts_off, ts.data(), ts.size() * sizeof(TStep));
  if (!nb.empty()) std::memcpy(hs + nb_off, nb.data(), nb.size() * sizeof(int));
  if (!mg.empty()) QS_CHECK_HIP(hipMemcpyAsync(ds, hs, mg.size() * sizeof(MGate), hipMemcpyHostToDevice, s));
  if (!ts.empty())
    QS_CHECK_HIP(hipMemcpyAsync(ds + ts_off, hs + ts_off, ts.size() * sizeof(TStep), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(ds + nb_off, hs + nb_off, rows * sizeof(int), hipMemcpyHostToDevice, s));
  if (fwd && tile) {
    // the forward state: |0...0>, then the plan's items last first
    rc = qs_sv_init(ctx, psi, n, dt, 0, s);
    for (auto it = items.rbegin(); !rc && it != items.rend(); ++it) {
      if (it->multi == 2) {
        const TPass& P = tpasses[it->idx];
        const uint64_t ntiles = 1ull << (n - P.nb);
        const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, ntiles);
        hipLaunchKernelGGL(k_fwd_tile<R>, dim3(blocks), dim3(ADJ_THREADS), ((size_t)1 << P.nb) * sizeof(cx<R>), s,
                           psi, ntiles, P, d_ts);
        QS_CHECK_HIP(hipGetLastError());
      } else {
        rc = apply_gate(ctx, psi, n, dt, &fwd[it->idx], s);  // register passes exist only without tiles
      }
    }
    if (rc) return rc;
  }
  if (energy) {
    rc = pauli_sum_async(ctx, psi, lam, n, dt, terms, nterms, s);
    if (rc) return rc;
    rc = inner_async(psi, lam, n, dt, ipart, ires, s);
    if (rc) return rc;
  }
  rc = prof_begin(ctx, PROF_GRAD, s);
  if (rc) return rc;
  for (const Item& it : items) {
    if (it.multi == 2) {
      const TPass& P = tpasses[it.idx];
      const uint64_t ntiles = 1ull << (n - P.nb);
      const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, ntiles);
      const size_t lds = 2 * ((size_t)1 << P.nb) * sizeof(cx<R>) + AT_MAX_GRADS * (ADJ_THREADS / 64) * sizeof(double);
      static bool lds_attr = false;  // tiles of 11 bits at complex128 need 65 KiB
      if (!lds_attr) {
        QS_CHECK_HIP(hipFuncSetAttribute((const void*)k_adj_tile<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(2 * ((size_t)1 << AT_MAX_BITS) * sizeof(cx<R>) +
                                               AT_MAX_GRADS * (ADJ_THREADS / 64) * sizeof(double))));
        lds_attr = true;
      }
      hipLaunchKernelGGL(k_adj_tile<R>, dim3(blocks), dim3(ADJ_THREADS), lds, s, psi, lam, ntiles, P, d_ts, partial);
      QS_CHECK_HIP(hipGetLastError());
      continue;
    }
    if (it.multi) {
      const MPass& P = passes[it.idx];
      const uint64_t ngroups = 1ull << (n - P.nb);
      const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, (ngroups + ADJ_THREADS - 1) / ADJ_THREADS);
      if (P.nb == 1)
        hipLaunchKernelGGL((k_adj_multi<R, 1>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, P, d_mg,
                           partial);
      else if (P.nb == 2)
        hipLaunchKernelGGL((k_adj_multi<R, 2>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, P, d_mg,
                           partial);
      else
        hipLaunchKernelGGL((k_adj_multi<R, 3>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, P, d_mg,
                           partial);
      QS_CHECK_HIP(hipGetLastError());
      continue;
    }
    const int j = it.idx;
    const qs_gate& g = gates[j];
    const bool grad = slot[j] >= 0;
    const int t = g.ntargets, c = g.ncontrols;
    if (t > 2) {
      rc = apply_gate(ctx, psi, n, sizeof(R) == 8 ? QS_C128 : QS_C64, &g, s);
      if (rc) return rc;
      rc = apply_gate(ctx, lam, n, sizeof(R) == 8 ? QS_C128 : QS_C64, &g, s);
      if (rc) return rc;
      continue;
    }
    AdjArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    int all[QS_MAX_GATE_BITS];
    for (int i = 0; i < c + t; ++i) all[i] = g.bits[i];
    std::sort(all, all + c + t);
    a.nsorted = c + t;
    for (int i = 0; i < c + t; ++i) a.sorted[i] = all[i];
    for (int i = 0; i < c; ++i) a.cmask |= 1ull << g.bits[i];
    for (int m = 0; m < (1 << t); ++m) {
      uint64_t o = 0;
      for (int k = 0; k < t; ++k)
        if ((m >> (t - 1 - k)) & 1) o |= 1ull << g.bits[c + k];
      a.off[m] = o;
    }
    std::vector<double> M;
    record_dense(g, M);
    for (size_t i = 0; i < M.size() / 2; ++i) a.g[i] = mk<R>((R)M[2 * i], (R)M[2 * i + 1]);
    if (grad) {
      record_dense(dgates[j], M);
      for (size_t i = 0; i < M.size() / 2; ++i) a.d[i] = mk<R>((R)M[2 * i], (R)M[2 * i + 1]);
    }
    const uint64_t ngroups = 1ull << (n - c - t);
    const int blocks = (int)std::min<uint64_t>(ADJ_MAX_BLOCKS, (ngroups + ADJ_THREADS - 1) / ADJ_THREADS);
    double* prow = partial + (uint64_t)(grad ? single_row[j] : 0) * ADJ_MAX_BLOCKS;
    if (t == 1) {
      if (grad) hipLaunchKernelGGL((k_adj_step<R, 1, true>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
      else hipLaunchKernelGGL((k_adj_step<R, 1, false>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
    } else {
      if (grad) hipLaunchKernelGGL((k_adj_step<R, 2, true>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
      else hipLaunchKernelGGL((k_adj_step<R, 2, false>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
    }
    QS_CHECK_HIP(hipGetLastError());
  }
  if (!nb.empty()) {
    hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)nb.size()), dim3(256), 0, s, partial, d_nb, ADJ_MAX_BLOCKS,
                       (int)nb.size(), sums);
    QS_CHECK_HIP(hipGetLastError());
  }
  rc = prof_end(ctx, s);
  if (rc) return rc;
  if (!nb.empty())
    QS_CHECK_HIP(hipMemcpyAsync(h_res, sums, nb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (energy) QS_CHECK_HIP(hipMemcpyAsync(h_res + rows, ires, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  if (energy) *energy = h_res[rows];
  for (size_t k = 0; k < nb.size(); ++k) {
    const int sl = step_slot[k];
    if (sl < nslots) grads[sl] += 2.0 * h_res[k];
  }
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

int qs_adjoint_gradient(qs_ctx* ctx, void* psi, void* lam, int n, int dtype, const qs_gate* gates, int ngates,
                        const qs_gate* dgates, const int32_t* param_slot, double* grads, int nslots, void* stream) {
  int rc = check_state_args(psi, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && lam && lam != psi, "bad state buffers");
  QS_REQUIRE(ngates >= 0 && (ngates == 0 || (gates && param_slot)), "bad gate list");
  QS_REQUIRE(nslots >= 0 && (nslots == 0 || grads), "bad gradient buffer");
  for (int j = 0; j < ngates; ++j) {
    rc = validate_gate(&gates[j], n);
    if (rc) return rc;
    if (param_slot[j] >= 0) {
      QS_REQUIRE(dgates != nullptr, "derivative records missing");
      QS_REQUIRE(param_slot[j] < nslots, "parameter slot out of range");
      rc = validate_gate(&dgates[j], n);
      if (rc) return rc;
    }
  }
  if (dtype == QS_C128)
    return adjoint_t<double>(ctx, (cx<double>*)psi, (cx<double>*)lam, n, gates, ngates, dgates, param_slot, grads,
                             nslots, as_stream(stream));
  return adjoint_t<float>(ctx, (cx<float>*)psi, (cx<float>*)lam, n, gates, ngates, dgates, param_slot, grads, nslots,
                          as_stream(stream));
}

int qs_adjoint_gradient_pauli(qs_ctx* ctx, void* psi, void* lam, int n, int dtype, const qs_pauli_term* terms,
                              int nterms, const qs_gate* fwd_gates, const qs_gate* gates, int ngates,
                              const qs_gate* dgates, const int32_t* param_slot, double* grads, int nslots,
                              double* energy, void* stream) {
  int rc = check_state_args(psi, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && lam && lam != psi && energy, "bad state buffers or energy pointer");
  QS_REQUIRE(nterms >= 1 && terms, "the Pauli sum needs at least one term");
  for (int k = 0; k < nterms; ++k) QS_REQUIRE(((terms[k].xmask | terms[k].zmask) >> n) == 0, "Pauli mask outside the register");
  QS_REQUIRE(ngates >= 0 && (ngates == 0 || (gates && param_slot)), "bad gate list");
  QS_REQUIRE(nslots >= 0 && (nslots == 0 || grads), "bad gradient buffer");
  for (int j = 0; j < ngates; ++j) {
    rc = validate_gate(&gates[j], n);
    if (rc) return rc;
    if (fwd_gates) {
      rc = validate_gate(&fwd_gates[j], n);
      if (rc) return rc;
    }
    if (param_slot[j] >= 0) {
      QS_REQUIRE(dgates != nullptr, "derivative records missing");
      QS_REQUIRE(param_slot[j] < nslots, "parameter slot out of range");
      rc = validate_gate(&dgates[j], n);
      if (rc) return rc;
    }
  }
  if (dtype == QS_C128)
    return adjoint_t<double>(ctx, (cx<double>*)psi, (cx<double>*)lam, n, gates, ngates, dgates, param_slot, grads,
                             nslots, as_stream(stream), terms, nterms, energy, fwd_gates);
  return adjoint_t<float>(ctx, (cx<float>*)psi, (cx<float>*)lam, n, gates, ngates, dgates, param_slot, grads, nslots,
                          as_stream(stream), terms, nterms, energy, fwd_gates);
}

}  // extern "C"
