// Dense gates on 4..6 targets as a complex matrix product on the matrix cores.
//
// Replaces QuICT/ops/linalg/gpu_calculator.py:343-478 (matrix_dot_vector: one
// thread per group of 2^k amplitudes, O(4^k) scalar FMAs, an out-of-place
// auxiliary vector) for the gates the register tile pass cannot hold
// (Unitary / QFT / IQFT / PermFx blocks on 5 or 6 targets, optionally
// controlled).
//
// With D = 2^k and the groups of the state (amplitudes that differ only in the
// target bits) as columns, the gate is Y = U X for X of shape D x #groups.  A
// workgroup owns a tile of G = 4096 / D groups (4096 amplitudes): its 12 tile
// bits are the k target bits and the log2(G) lowest other free bits, so the
// tile sweep reads and writes address-ordered runs.  Per tile:
//   1. gather X into LDS (separate re / im planes, row m = target combination,
//      column g = group), coalesced over the tile's address order;
//   2. each of the four waves multiplies its 16-row block of U (held in
//      registers for the whole launch, loaded once per workgroup) by four
//      16-column blocks of X with v_mfma_{f64,f32}_16x16x4: four real products
//      per complex step (re*re, -im*im, re*im, im*re);
//   3. the 16x16 results go back to LDS and are stored to the addresses they
//      were loaded from (in place: a tile's groups belong to this workgroup
//      only).
// Flops per amplitude: 8 * D; HBM bytes per amplitude touched: 2 * S.  At 6
// targets (512 flops per 32 bytes for complex128) the product is above the
// machine balance, so the bound is the FP64 matrix rate, not HBM.
// Controls fix their bits to 1; only the controlled subspace is touched.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstring>

namespace qs {

constexpr int DM_TILE_LOG2 = 12;  // amplitudes per workgroup tile
constexpr int DM_THREADS = 256;
constexpr int DM_PAD = 2;          // LDS row padding (elements) against bank conflicts

struct DenseTile {
  int32_t tbits[DM_TILE_LOG2];      // tile bits ascending (address bits of the local index)
  int32_t skip[QS_MAX_GATE_BITS + DM_TILE_LOG2];  // tile bits and control bits ascending (tile base)
  int32_t nskip;
  uint32_t mg[DM_TILE_LOG2];        // local bit j -> (1 << matrix bit) | (1 << (8 + group bit))
  uint64_t cmask;                   // control bits
};

// One 16x16x4 product-accumulate C += A B per wave.  Operands: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16].  Result: lane l, element r holds
// C[row(l / 16, r)][l % 16], where the row differs between the two
// instructions (measured on gfx950: f64 interleaves, f32 blocks).
template <typename R> struct mfma_acc;
template <> struct mfma_acc<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct mfma_acc<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// Local-index part of thread t's elements: e = t + 256 * j, j = 0..15.
__device__ __forceinline__ void dm_lane_parts(const DenseTile& P, uint32_t t, uint64_t& addr_t, uint32_t& mg_t) {
  addr_t = 0;
  mg_t = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b)
    if ((t >> b) & 1u) {
      addr_t |= 1ull << P.tbits[b];
      mg_t |= P.mg[b];
    }
}

template <typename R, int KB>
__global__ __launch_bounds__(DM_THREADS) void k_dense_mfma(cx<R>* __restrict__ v, const DenseTile P,
                                                          const R* __restrict__ Ure, const R* __restrict__ Uim,
                                                          uint64_t ntiles) {
  using A4 = mfma_acc<R>;
  using v4 = typename A4::v4;
  constexpr int D = 1 << KB;
  constexpr int G = (1 << DM_TILE_LOG2) / D;
  constexpr int GP = G + DM_PAD;
  constexpr int NRB = D / 16;       // 16-row blocks of U
  constexpr int KS = D / 4;         // k steps of 4
  static_assert(NRB >= 1 && NRB <= 4, "4..6 targets");
  extern __shared__ __align__(16) char dm_lds[];
  R* Xre = reinterpret_cast<R*>(dm_lds);
  R* Xim = Xre + D * GP;

  const uint32_t tid = threadIdx.x;
  const int lane = (int)(tid & 63), wave = (int)(tid >> 6);
  const int rb = wave % NRB;                 // this wave's row block
  const int ct0 = (wave / NRB) * 4;          // its first of four 16-column blocks
  const int lr = lane & 15, lk = lane >> 4;  // MFMA operand coordinates

  // U fragments of rows 16 rb + lr, columns 4 s + lk
  R are[KS], aim[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int i = 16 * rb + lr, k = 4 * s + lk;
    are[s] = Ure[i * D + k];
    aim[s] = Uim[i * D + k];
  }

  uint64_t addr_t;
  uint32_t mg_t;
  dm_lane_parts(P, tid, addr_t, mg_t);

  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = insert_zeros(tile, P.skip, P.nskip) | P.cmask | addr_t;
    // 1. gather
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      uint64_t aj = 0;
      uint32_t mgj = mg_t;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((j >> b) & 1) {
          aj |= 1ull << P.tbits[8 + b];
          mgj |= P.mg[8 + b];
        }
      const cx<R> x = ld(v + (base | aj));
      const int m = (int)(mgj & 0xffu), g = (int)(mgj >> 8);
      Xre[m * GP + g] = x.re;
      Xim[m * GP + g] = x.im;
    }
    __syncthreads();
    // 2. Y = U X on the matrix cores
    v4 cre[4], cim[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      cre[c] = v4{0, 0, 0, 0};
      cim[c] = v4{0, 0, 0, 0};
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int off = (4 * s + lk) * GP + 16 * (ct0 + c) + lr;
        const R bre = Xre[off], bim = Xim[off];
        cre[c] = A4::step(are[s], bre, cre[c]);
        cre[c] = A4::step(aim[s], -bim, cre[c]);
        cim[c] = A4::step(are[s], bim, cim[c]);
        cim[c] = A4::step(aim[s], bre, cim[c]);
      }
    }
    __syncthreads();  // every wave has read X before Y overwrites it
    // 3. results to LDS (row 16 rb + row(lk, r), column 16 (ct0 + c) + lr), then stored in address order
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int off = (16 * rb + A4::row(lk, r)) * GP + 16 * (ct0 + c) + lr;
        Xre[off] = cre[c][r];
        Xim[off] = cim[c][r];
      }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      uint64_t aj = 0;
      uint32_t mgj = mg_t;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((j >> b) & 1) {
          aj |= 1ull << P.tbits[8 + b];
          mgj |= P.mg[8 + b];
        }
      const int m = (int)(mgj & 0xffu), g = (int)(mgj >> 8);
      st(v + (base | aj), mk<R>(Xre[m * GP + g], Xim[m * GP + g]));
    }
    __syncthreads();  // the next tile's gather reuses the planes
  }
}

template <typename R, int KB> static size_t dense_mfma_lds() {
  constexpr int D = 1 << KB;
  constexpr int G = (1 << DM_TILE_LOG2) / D;
  return (size_t)2 * D * (G + DM_PAD) * sizeof(R);
}

// Whether the matrix-core path takes this gate: 4..6 targets and at least one
// full tile of groups outside the target and control bits.
bool dense_mfma_eligible(const qs_gate* g, int n) {
  const int t = g->ntargets, c = g->ncontrols;
  if (g->kind != QS_GATE_DENSE || t < 4 || t > 6) return false;
  const int lg = DM_TILE_LOG2 - t;
  if (n - t - c < lg) return false;
  if (const char* e = getenv("QS_DENSE_MFMA"))
    if (e[0] == '0') return false;
  return true;
}

template <typename R, int KB>
static int launch_dense_mfma_t(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* g, hipStream_t s) {
  constexpr int D = 1 << KB;
  constexpr int LG = DM_TILE_LOG2 - KB;
  const int c = g->ncontrols;
  DenseTile P;
  std::memset(&P, 0, sizeof(P));
  uint64_t used = 0;
  for (int i = 0; i < c + KB; ++i) used |= 1ull << g->bits[i];
  // tile bits: the targets and the LG lowest free bits
  int lbits[DM_TILE_LOG2];
  int nl = 0;
  for (int i = 0; i < KB; ++i) lbits[nl++] = g->bits[c + i];
  for (int b = 0; b < n && nl < DM_TILE_LOG2; ++b)
    if (!((used >> b) & 1)) lbits[nl++] = b;
  std::sort(lbits, lbits + DM_TILE_LOG2);
  int gcount = 0;
  for (int j = 0; j < DM_TILE_LOG2; ++j) {
    P.tbits[j] = lbits[j];
    int tpos = -1;
    for (int i = 0; i < KB; ++i)
      if (g->bits[c + i] == lbits[j]) tpos = i;
    if (tpos >= 0) P.mg[j] = 1u << (KB - 1 - tpos);  // target 0 = matrix MSB
    else P.mg[j] = 1u << (8 + gcount++);
  }
  if (gcount != LG) {
    set_error("dense matrix-core pass: tile construction failed");
    return QS_ERR_DEVICE;
  }
  int sk[QS_MAX_GATE_BITS + DM_TILE_LOG2];
  int ns = 0;
  for (int j = 0; j < DM_TILE_LOG2; ++j) sk[ns++] = lbits[j];
  for (int i = 0; i < c; ++i) {
    sk[ns++] = g->bits[i];
    P.cmask |= 1ull << g->bits[i];
  }
  std::sort(sk, sk + ns);
  for (int i = 0; i < ns; ++i) P.skip[i] = sk[i];
  P.nskip = ns;
  const uint64_t ntiles = 1ull << (n - ns);
  // the matrix as re / im planes (row-major), in the context's scratch
  const uint64_t plane = (uint64_t)D * D * sizeof(R);
  int rc = ensure_scratch(ctx, 2 * plane);
  if (rc) return rc;
  R* dre = (R*)ctx->scratch;
  R* dim = (R*)((char*)ctx->scratch + plane);
  std::vector<R> h(2 * (size_t)D * D);
  for (int i = 0; i < D * D; ++i) {
    h[i] = (R)g->big[2 * i];
    h[(size_t)D * D + i] = (R)g->big[2 * i + 1];
  }
  QS_CHECK_HIP(hipMemcpyAsync(dre, h.data(), 2 * plane, hipMemcpyHostToDevice, s));
  const size_t lds = dense_mfma_lds<R, KB>();
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_dense_mfma<R, KB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    attr_set = true;
  }
  const uint64_t grid = std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 4);
  hipLaunchKernelGGL((k_dense_mfma<R, KB>), dim3((unsigned)grid), dim3(DM_THREADS), lds, s, v, P, dre, dim, ntiles);
  QS_CHECK_HIP(hipGetLastError());
  // the host copy of the matrix must outlive the async upload
  QS_CHECK_HIP(hipStreamSynchronize(s));
  return QS_OK;
}

int launch_dense_mfma(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* g, hipStream_t s) {
  if (dtype == QS_C128) {
    cx<double>* v = (cx<double>*)vec;
    switch (g->ntargets) {
      case 4: return launch_dense_mfma_t<double, 4>(ctx, v, n, g, s);
      case 5: return launch_dense_mfma_t<double, 5>(ctx, v, n, g, s);
      case 6: return launch_dense_mfma_t<double, 6>(ctx, v, n, g, s);
    }
  } else {
    cx<float>* v = (cx<float>*)vec;
    switch (g->ntargets) {
      case 4: return launch_dense_mfma_t<float, 4>(ctx, v, n, g, s);
      case 5: return launch_dense_mfma_t<float, 5>(ctx, v, n, g, s);
      case 6: return launch_dense_mfma_t<float, 6>(ctx, v, n, g, s);
    }
  }
  set_error("dense matrix-core pass takes 4 to 6 targets");
  return QS_ERR_UNSUPPORTED;
}

}  // namespace qs
