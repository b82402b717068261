// Per-gate state-vector kernels: one full read+write pass per gate.
//
// Stand-in for the per-gate CUDA kernels of QuICT/ops/gate_kernel/gate_function.py
// (rows 1-18 of its kernel table) and the k-target mat-vec of
// QuICT/ops/linalg/gpu_calculator.py:343-478.  Instead of one kernel per
// (matrix type x arity) we have one template per structural kind; controls are
// folded in by inserting zero bits at every gate bit and OR-ing the control
// mask, so each thread touches only the amplitudes the gate changes.
//
// Layout: amplitudes are contiguous interleaved complex; consecutive threads
// own consecutive group indices, so for every target bit the loads of a
// wavefront fall on contiguous 2^b-amplitude runs (full 128-byte lines once
// b >= 3 for complex128).  Each thread handles G groups spaced one grid apart
// to keep several loads in flight per lane.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstring>

namespace qs {

template <typename R, int KIND, int NT, int G>
__global__ __launch_bounds__(256) void k_gate(cx<R>* __restrict__ v, uint64_t ngroups, const GateK<R> a) {
  constexpr int NA = (KIND == QS_GATE_PHASE) ? 1 : (KIND == QS_GATE_SWAP ? 2 : (1 << NT));
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t base[G];
  cx<R> x[G][NA];
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const uint64_t g = tid + r * stride;
    base[r] = (g < ngroups) ? (insert_zeros(g, a.sorted, a.nsorted) | a.cmask) : ~0ull;
  }
#pragma unroll
  for (int r = 0; r < G; ++r) {
    if (base[r] == ~0ull) continue;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      uint64_t off;
      if (KIND == QS_GATE_PHASE) off = a.off[(1 << NT) - 1];
      else if (KIND == QS_GATE_SWAP) off = a.off[1 + j];
      else off = a.off[j];
      x[r][j] = ld(v + (base[r] | off));
    }
  }
#pragma unroll
  for (int r = 0; r < G; ++r) {
    if (base[r] == ~0ull) continue;
    cx<R> y[NA];
    if constexpr (KIND == QS_GATE_DENSE) {
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        cx<R> acc = mk<R>(0, 0);
#pragma unroll
        for (int j = 0; j < NA; ++j) acc = cfma(acc, a.m[i * NA + j], x[r][j]);
        y[i] = acc;
      }
    } else if constexpr (KIND == QS_GATE_DIAG) {
#pragma unroll
      for (int i = 0; i < NA; ++i) y[i] = mul(a.m[i], x[r][i]);
    } else if constexpr (KIND == QS_GATE_PHASE) {
      y[0] = mul(a.m[0], x[r][0]);
    } else if constexpr (KIND == QS_GATE_XPERM) {
      y[0] = mul(a.m[0], x[r][1]);
      y[1] = mul(a.m[1], x[r][0]);
    } else {  // SWAP: x[0] = |01>, x[1] = |10>
      y[0] = mul(a.m[0], x[r][1]);
      y[1] = mul(a.m[1], x[r][0]);
    }
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      uint64_t off;
      if (KIND == QS_GATE_PHASE) off = a.off[(1 << NT) - 1];
      else if (KIND == QS_GATE_SWAP) off = a.off[1 + j];
      else off = a.off[j];
      st(v + (base[r] | off), y[j]);
    }
  }
}

// Three-target dense gate in place: one thread per group of 8 amplitudes,
// the 8 x 8 matrix and the target offsets in the kernel arguments (no
// scratch, nothing to keep alive on the host).
template <typename R> struct Dense3Args {
  cx<R> m[64];  // row-major; target 0 is the matrix index MSB
  uint64_t off[8];
  SortedBits sb;
  uint64_t cmask;
};

template <typename R>
__global__ __launch_bounds__(256) void k_dense3(cx<R>* __restrict__ v, uint64_t groups, const Dense3Args<R> a) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= groups) return;
  const uint64_t base = insert_zeros(g, a.sb.b, a.sb.n) | a.cmask;
  cx<R> x[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = ld(v + (base | a.off[c]));
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    cx<R> y = mul(a.m[8 * r], x[0]);
#pragma unroll
    for (int c = 1; c < 8; ++c) y = cfma(y, a.m[8 * r + c], x[c]);
    st(v + (base | a.off[r]), y);
  }
}

// k-target dense mat-vec (k >= 4 where the matrix-core kernel does not
// apply), out of place: one thread per output amplitude.
// out[base(g) | off(r)] = sum_c A[r][c] * in[base(g) | off(c)]
template <typename R>
__global__ __launch_bounds__(256) void k_dense_k(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                 uint64_t total, int k, const cx<R>* __restrict__ A,
                                                 const uint64_t* __restrict__ offs, SortedBits sb,
                                                 uint64_t cmask) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint64_t dim = 1ull << k;
  const uint64_t r = t & (dim - 1);
  const uint64_t g = t >> k;
  const uint64_t base = insert_zeros(g, sb.b, sb.n) | cmask;
  const cx<R>* row = A + r * dim;
  cx<R> acc = mk<R>(0, 0);
  for (uint64_t c = 0; c < dim; ++c) acc = cfma(acc, row[c], ld(in + (base | offs[c])));
  st(out + (base | offs[r]), acc);
}

template <typename R>
__global__ void k_copy_groups(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, uint64_t total, int k,
                              const uint64_t* __restrict__ offs, SortedBits sb, uint64_t cmask) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint64_t r = t & ((1ull << k) - 1);
  const uint64_t base = insert_zeros(t >> k, sb.b, sb.n) | cmask;
  const uint64_t i = base | offs[r];
  st(out + i, ld(in + i));
}

// ---- host side -----------------------------------------------------------------

int validate_gate(const qs_gate* g, int n) {
  QS_REQUIRE(g != nullptr, "null gate");
  const int c = g->ncontrols, t = g->ntargets;
  QS_REQUIRE(c >= 0 && t >= 1 && c + t <= QS_MAX_GATE_BITS, "gate arity out of range");
  QS_REQUIRE(c + t <= n, "gate wider than the register");
  uint64_t seen = 0;
  for (int i = 0; i < c + t; ++i) {
    QS_REQUIRE(g->bits[i] >= 0 && g->bits[i] < n, "gate bit outside the register");
    QS_REQUIRE(!(seen & (1ull << g->bits[i])), "gate bits must be distinct");
    seen |= 1ull << g->bits[i];
  }
  switch (g->kind) {
    case QS_GATE_DENSE:
      QS_REQUIRE(t <= QS_MAX_DENSE_TARGETS, "dense gate has too many targets");
      QS_REQUIRE(t <= 2 || g->big != nullptr, "dense gate with >= 3 targets needs `big`");
      break;
    case QS_GATE_DIAG:
    case QS_GATE_PHASE:
      QS_REQUIRE(t <= 4, "diagonal gate supports at most 4 targets");
      break;
    case QS_GATE_XPERM:
      QS_REQUIRE(t == 1, "xperm gate has exactly 1 target");
      break;
    case QS_GATE_SWAP:
      QS_REQUIRE(t == 2, "swap gate has exactly 2 targets");
      break;
    default:
      QS_REQUIRE(false, "unknown gate kind");
  }
  return QS_OK;
}

template <typename R> void build_gatek(const qs_gate* g, GateK<R>* k) {
  std::memset(k, 0, sizeof(*k));
  const int c = g->ncontrols, t = g->ntargets;
  int all[QS_MAX_GATE_BITS];
  for (int i = 0; i < c + t; ++i) all[i] = g->bits[i];
  std::sort(all, all + c + t);
  k->nsorted = c + t;
  for (int i = 0; i < c + t; ++i) k->sorted[i] = all[i];
  for (int i = 0; i < c; ++i) k->cmask |= 1ull << g->bits[i];
  const int tt = std::min(t, 4);
  for (int m = 0; m < (1 << tt); ++m) {
    uint64_t o = 0;
    for (int j = 0; j < tt; ++j)
      if ((m >> (tt - 1 - j)) & 1) o |= 1ull << g->bits[c + j];
    k->off[m] = o;
  }
  for (int i = 0; i < 16; ++i) k->m[i] = mk<R>((R)g->m[2 * i], (R)g->m[2 * i + 1]);
}

template <typename R, int KIND, int NT>
static int launch_kind(cx<R>* v, int n, const qs_gate* g, hipStream_t s) {
  constexpr int G = 4;
  GateK<R> k;
  build_gatek<R>(g, &k);
  const uint64_t ngroups = 1ull << (n - g->ncontrols - g->ntargets);
  const uint64_t threads = (ngroups + G - 1) / G;
  const uint64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL((k_gate<R, KIND, NT, G>), dim3((unsigned)blocks), dim3(256), 0, s, v, ngroups, k);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

template <typename R>
static int launch_dense_k(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* g, hipStream_t s) {
  const int c = g->ncontrols, t = g->ntargets;
  const uint64_t dim = 1ull << t;
  const uint64_t total = 1ull << (n - c);  // groups * dim
  const uint64_t state_bytes = (1ull << n) * sizeof(cx<R>);
  const uint64_t mat_bytes = dim * dim * sizeof(cx<R>);
  const uint64_t off_bytes = dim * sizeof(uint64_t);
  const uint64_t need = state_bytes + ((mat_bytes + 255) & ~255ull) + off_bytes;
  int rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  char* base = (char*)ctx->scratch;
  cx<R>* out = (cx<R>*)base;
  cx<R>* A = (cx<R>*)(base + state_bytes);
  uint64_t* offs = (uint64_t*)(base + state_bytes + ((mat_bytes + 255) & ~255ull));
  std::vector<cx<R>> hA(dim * dim);
  for (uint64_t i = 0; i < dim * dim; ++i) hA[i] = mk<R>((R)g->big[2 * i], (R)g->big[2 * i + 1]);
  std::vector<uint64_t> hoff(dim);
  for (uint64_t m = 0; m < dim; ++m) {
    uint64_t o = 0;
    for (int j = 0; j < t; ++j)
      if ((m >> (t - 1 - j)) & 1) o |= 1ull << g->bits[c + j];
    hoff[m] = o;
  }
  SortedBits sb;
  sb.n = c + t;
  for (int i = 0; i < c + t; ++i) sb.b[i] = g->bits[i];
  std::sort(sb.b, sb.b + sb.n);
  uint64_t cmask = 0;
  for (int i = 0; i < c; ++i) cmask |= 1ull << g->bits[i];
  QS_CHECK_HIP(hipMemcpyAsync(A, hA.data(), mat_bytes, hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(offs, hoff.data(), off_bytes, hipMemcpyHostToDevice, s));
  const uint64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_dense_k<R>, dim3((unsigned)blocks), dim3(256), 0, s, v, out, total, t, A, offs, sb, cmask);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_groups<R>, dim3((unsigned)blocks), dim3(256), 0, s, out, v, total, t, offs, sb, cmask);
  QS_CHECK_HIP(hipGetLastError());
  // host vectors must outlive the async copies
  QS_CHECK_HIP(hipStreamSynchronize(s));
  return QS_OK;
}

template <typename R> static int launch_dense3(cx<R>* v, int n, const qs_gate* g, hipStream_t s) {
  const int c = g->ncontrols;
  Dense3Args<R> a;
  std::memset(&a, 0, sizeof(a));
  for (int i = 0; i < 64; ++i) a.m[i] = mk<R>((R)g->big[2 * i], (R)g->big[2 * i + 1]);
  for (int m = 0; m < 8; ++m) {
    uint64_t o = 0;
    for (int j = 0; j < 3; ++j)
      if ((m >> (2 - j)) & 1) o |= 1ull << g->bits[c + j];
    a.off[m] = o;
  }
  a.sb.n = c + 3;
  for (int i = 0; i < c + 3; ++i) a.sb.b[i] = g->bits[i];
  std::sort(a.sb.b, a.sb.b + a.sb.n);
  for (int i = 0; i < c; ++i) a.cmask |= 1ull << g->bits[i];
  const uint64_t groups = 1ull << (n - c - 3);
  hipLaunchKernelGGL(k_dense3<R>, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, v, groups, a);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

template <typename R> int apply_gate_t(qs_ctx* ctx, cx<R>* v, int n, const qs_gate* g, hipStream_t s) {
  const int t = g->ntargets;
  switch (g->kind) {
    case QS_GATE_DENSE:
      if (t == 1) return launch_kind<R, QS_GATE_DENSE, 1>(v, n, g, s);
      if (t == 2) return launch_kind<R, QS_GATE_DENSE, 2>(v, n, g, s);
      if (t == 3) return launch_dense3<R>(v, n, g, s);
      if (dense_mfma_eligible(g, n))
        return launch_dense_mfma(ctx, v, n, sizeof(R) == 8 ? QS_C128 : QS_C64, g, s);
      return launch_dense_k<R>(ctx, v, n, g, s);
    case QS_GATE_DIAG:
      if (t == 1) return launch_kind<R, QS_GATE_DIAG, 1>(v, n, g, s);
      if (t == 2) return launch_kind<R, QS_GATE_DIAG, 2>(v, n, g, s);
      if (t == 3) return launch_kind<R, QS_GATE_DIAG, 3>(v, n, g, s);
      return launch_kind<R, QS_GATE_DIAG, 4>(v, n, g, s);
    case QS_GATE_PHASE:
      if (t == 1) return launch_kind<R, QS_GATE_PHASE, 1>(v, n, g, s);
      if (t == 2) return launch_kind<R, QS_GATE_PHASE, 2>(v, n, g, s);
      if (t == 3) return launch_kind<R, QS_GATE_PHASE, 3>(v, n, g, s);
      return launch_kind<R, QS_GATE_PHASE, 4>(v, n, g, s);
    case QS_GATE_XPERM:
      return launch_kind<R, QS_GATE_XPERM, 1>(v, n, g, s);
    case QS_GATE_SWAP:
      return launch_kind<R, QS_GATE_SWAP, 2>(v, n, g, s);
  }
  set_error("unknown gate kind");
  return QS_ERR_ARG;
}

int apply_gate(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* g, hipStream_t s) {
  int rc = validate_gate(g, n);
  if (rc) return rc;
  if (dtype == QS_C128) return apply_gate_t<double>(ctx, (cx<double>*)vec, n, g, s);
  return apply_gate_t<float>(ctx, (cx<float>*)vec, n, g, s);
}

}  // namespace qs
