// Measurement, reset, sampling and reduction kernels.
//
// Replaces QuICT/ops/gate_kernel/gate_function.py:1596-1847 (ProbAdd* kernels +
// ReductionKernel for P0, MeasureGate0/1, ResetGate0/1) and the per-shot
// collapse loop of StateVectorSimulator.sample (statevector_simulator.py:595-626).
//
// * P0 is reduced directly from the state in double (no 2^(n-1)-element
//   complex scratch as in the reference): wave-level shuffles, then one
//   partial per block, then a single-block final sum.  The grid is fixed, so
//   results are bit-reproducible run to run.
// * sample(): a pair-sum pyramid of |a|^2 (level l holds 2^(n-l) doubles) is
//   built once; every shot then descends from the root, one level per qubit,
//   with outcome 1 iff u > left/(left+right).  That is the reference's
//   sequential measure-and-renormalise rule evaluated without touching the
//   state, so a shot costs n loads instead of 3n+1 full-state passes.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

namespace qs {

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[RED_THREADS / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += part[i];
  }
  __syncthreads();
  return s;
}

// sum |a_i|^2 over i with (i & mask) == val ... expressed over half-index h:
// i = insert bit `bit` with value `bitval` into h.  bit < 0: every amplitude.
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_norm_partial(const cx<R>* __restrict__ v, uint64_t count, int bit,
                                                              int bitval, double* __restrict__ partial) {
  double s = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < count; h += stride) {
    uint64_t i = h;
    if (bit >= 0) i = ((h >> bit) << (bit + 1)) | (h & ((1ull << bit) - 1)) | ((uint64_t)bitval << bit);
    const cx<R> a = ld(v + i);
    s += (double)a.re * (double)a.re + (double)a.im * (double)a.im;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[0] = s;
}

template <typename R>
__global__ void k_collapse(cx<R>* __restrict__ v, uint64_t half, int bit, int outcome, R scale) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= half) return;
  const uint64_t i0 = ((h >> bit) << (bit + 1)) | (h & ((1ull << bit) - 1));
  const uint64_t i1 = i0 | (1ull << bit);
  const uint64_t keep = outcome ? i1 : i0, drop = outcome ? i0 : i1;
  st(v + keep, scale_(ld(v + keep), scale));
  st(v + drop, mk<R>(0, 0));
}

template <typename R>
__global__ void k_reset(cx<R>* __restrict__ v, uint64_t half, int bit, int move, R inv_alpha) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= half) return;
  const uint64_t i0 = ((h >> bit) << (bit + 1)) | (h & ((1ull << bit) - 1));
  const uint64_t i1 = i0 | (1ull << bit);
  if (move) {
    st(v + i0, ld(v + i1));
  } else {
    st(v + i0, scale_(ld(v + i0), inv_alpha));
  }
  st(v + i1, mk<R>(0, 0));
}

template <typename R>
__global__ void k_probs(const cx<R>* __restrict__ v, uint64_t N, double* __restrict__ p) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const cx<R> a = ld(v + i);
  p[i] = (double)a.re * (double)a.re + (double)a.im * (double)a.im;
}

// Leaf probabilities: |a_i|^2 of a state, or a stored probability vector.
template <typename R> struct LeafState {
  const cx<R>* v;
  __device__ __forceinline__ double operator()(uint64_t i) const {
    const cx<R> a = ld(v + i);
    return (double)a.re * (double)a.re + (double)a.im * (double)a.im;
  }
};
struct LeafProbs {
  const double* p;
  __device__ __forceinline__ double operator()(uint64_t i) const { return p[i]; }
};

// Pyramid level 1 from the leaves: L1[j] = p_2j + p_2j+1.
template <typename Leaf>
__global__ void k_pyr_first(Leaf leaf, uint64_t M, double* __restrict__ L1) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  L1[j] = leaf(2 * j) + leaf(2 * j + 1);
}

__global__ void k_pyr_next(const double* __restrict__ lo, uint64_t M, double* __restrict__ hi) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  hi[j] = lo[2 * j] + lo[2 * j + 1];
}

// One thread per shot.  level_ptr[l] = start of level l (l = 1..n) inside pyr;
// level 0 is read from the state.
template <typename Leaf>
__global__ void k_descend(Leaf leaf, const double* __restrict__ pyr,
                          const uint64_t* __restrict__ level_off, int n, const double* __restrict__ u, int shots,
                          uint64_t* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= shots) return;
  uint64_t node = 0;
  for (int q = 0; q < n; ++q) {
    const int l = n - 1 - q;  // level holding the children
    double left, right;
    if (l == 0) {
      left = leaf(2 * node);
      right = leaf(2 * node + 1);
    } else {
      const double* L = pyr + level_off[l];
      left = L[2 * node];
      right = L[2 * node + 1];
    }
    const double tot = left + right;
    const double p0 = tot > 0 ? left / tot : 1.0;
    const uint64_t b = u[(uint64_t)s * n + q] > p0 ? 1 : 0;
    node = 2 * node + b;
  }
  out[s] = node;
}

// ---- host ------------------------------------------------------------------------

template <typename R>
static int reduce_t(qs_ctx* ctx, const cx<R>* v, int n, int bit, int bitval, double* host_out, hipStream_t s) {
  int rc = ensure_scratch(ctx, (RED_BLOCKS + 64) * sizeof(double));
  if (rc) return rc;
  double* partial = (double*)ctx->scratch;
  double* res = partial + RED_BLOCKS;
  const uint64_t count = bit >= 0 ? (1ull << (n - 1)) : (1ull << n);
  uint64_t blocks = (count + RED_THREADS - 1) / RED_THREADS;
  if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
  hipLaunchKernelGGL(k_norm_partial<R>, dim3((unsigned)blocks), dim3(RED_THREADS), 0, s, v, count, bit, bitval,
                     partial);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_THREADS), 0, s, partial, (int)blocks, res);
  QS_CHECK_HIP(hipGetLastError());
  QS_CHECK_HIP(hipMemcpyAsync(ctx->host_pinned, res, sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  *host_out = ctx->host_pinned[0];
  return QS_OK;
}

int reduce_prob_zero(qs_ctx* ctx, const void* vec, int n, int dtype, int bit, double* out, hipStream_t s) {
  if (dtype == QS_C128) return reduce_t<double>(ctx, (const cx<double>*)vec, n, bit, 0, out, s);
  return reduce_t<float>(ctx, (const cx<float>*)vec, n, bit, 0, out, s);
}

template <typename R>
static int collapse_t(cx<R>* v, int n, int bit, int outcome, double scale, hipStream_t s) {
  const uint64_t half = 1ull << (n - 1);
  hipLaunchKernelGGL(k_collapse<R>, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, v, half, bit, outcome,
                     (R)scale);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

template <typename R> static int reset_t(cx<R>* v, int n, int bit, double p0, hipStream_t s) {
  const uint64_t half = 1ull << (n - 1);
  const double alpha = sqrt(p0);
  const int move = alpha < 1e-6;
  hipLaunchKernelGGL(k_reset<R>, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, v, half, bit, move,
                     (R)(move ? 0.0 : 1.0 / alpha));
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

// Scratch layout of one sampling call: pyramid levels 1..n (sizes 2^(n-1) .. 1,
// back to back), the level offsets, the uniforms and the outcomes.
struct SampleLayout {
  uint64_t off[64];
  uint64_t pyr_bytes, off_bytes, u_bytes, o_bytes;
  uint64_t total() const { return pyr_bytes + off_bytes + u_bytes + o_bytes; }
};

static SampleLayout sample_layout(int n, int shots) {
  SampleLayout L{};
  uint64_t total = 0;
  for (int l = 1; l <= n; ++l) {
    L.off[l] = total;
    total += 1ull << (n - l);
  }
  L.pyr_bytes = ((total * sizeof(double)) + 255) & ~255ull;
  L.off_bytes = 256 * 8;
  L.u_bytes = (((uint64_t)shots * n * sizeof(double)) + 255) & ~255ull;
  L.o_bytes = (uint64_t)shots * sizeof(uint64_t);
  return L;
}

// Build the pyramid and descend it for every shot, in `base` (at least
// L.total() bytes of device memory).  Synchronous: returns with `out` filled.
template <typename Leaf>
static int sample_in(Leaf leaf, int n, const double* uniforms, int shots, uint64_t* out, const SampleLayout& L,
                     char* base, hipStream_t s) {
  const uint64_t* off = L.off;
  const uint64_t pyr_bytes = L.pyr_bytes, off_bytes = L.off_bytes, u_bytes = L.u_bytes, o_bytes = L.o_bytes;
  double* pyr = (double*)base;
  uint64_t* d_off = (uint64_t*)(base + pyr_bytes);
  double* d_u = (double*)(base + pyr_bytes + off_bytes);
  uint64_t* d_out = (uint64_t*)(base + pyr_bytes + off_bytes + u_bytes);
  if (n >= 1) {
    const uint64_t M = 1ull << (n - 1);
    hipLaunchKernelGGL(k_pyr_first<Leaf>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, leaf, M, pyr + off[1]);
    QS_CHECK_HIP(hipGetLastError());
    for (int l = 2; l <= n; ++l) {
      const uint64_t Ml = 1ull << (n - l);
      hipLaunchKernelGGL(k_pyr_next, dim3((unsigned)((Ml + 255) / 256)), dim3(256), 0, s, pyr + off[l - 1], Ml,
                         pyr + off[l]);
      QS_CHECK_HIP(hipGetLastError());
    }
  }
  QS_CHECK_HIP(hipMemcpyAsync(d_off, off, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_u, uniforms, (uint64_t)shots * n * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_descend<Leaf>, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, s, leaf, pyr, d_off, n,
                     d_u, shots, d_out);
  QS_CHECK_HIP(hipGetLastError());
  QS_CHECK_HIP(hipMemcpyAsync(out, d_out, o_bytes, hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  return QS_OK;
}

template <typename Leaf>
static int sample_t(qs_ctx* ctx, Leaf leaf, int n, const double* uniforms, int shots, uint64_t* out,
                    hipStream_t s) {
  const SampleLayout L = sample_layout(n, shots);
  int rc = ensure_scratch(ctx, L.total());
  if (rc) return rc;
  return sample_in(leaf, n, uniforms, shots, out, L, (char*)ctx->scratch, s);
}

// Fixed device scratch of qs_sv_sample_noalloc: a file-scope array, one copy
// per device, held until the process ends.  Host code never uses the name
// itself as a device address: noalloc_scratch_init resolves each device's
// copy with hipGetSymbolAddress (the context keeps it) and only that address
// is passed to kernels and copies.
constexpr uint64_t NOALLOC_SAMPLE_BYTES = 40ull << 20;
static __device__ __align__(256) char scratch[NOALLOC_SAMPLE_BYTES];

int noalloc_scratch_init(qs_ctx* ctx) {
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return QS_ERR_DEVICE;
  }
  ctx->noalloc_scratch = ptr;
  ctx->noalloc_scratch_bytes = NOALLOC_SAMPLE_BYTES;
  return QS_OK;
}

template <typename R> static int probs_t(const cx<R>* v, int n, double* p, hipStream_t s) {
  const uint64_t N = 1ull << n;
  hipLaunchKernelGGL(k_probs<R>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, v, N, p);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

int qs_sv_prob_zero(qs_ctx* ctx, const void* vec, int n, int dtype, int bit, double* p0, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && p0, "null ctx or output");
  QS_REQUIRE(bit >= 0 && bit < n, "bit outside the register");
  return reduce_prob_zero(ctx, vec, n, dtype, bit, p0, as_stream(stream));
}

int qs_sv_norm2(qs_ctx* ctx, const void* vec, int n, int dtype, double* out, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && out, "null ctx or output");
  if (dtype == QS_C128) return reduce_t<double>(ctx, (const cx<double>*)vec, n, -1, 0, out, as_stream(stream));
  return reduce_t<float>(ctx, (const cx<float>*)vec, n, -1, 0, out, as_stream(stream));
}

int qs_sv_collapse(qs_ctx* ctx, void* vec, int n, int dtype, int bit, int outcome, double scale, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(bit >= 0 && bit < n, "bit outside the register");
  QS_REQUIRE(outcome == 0 || outcome == 1, "outcome must be 0 or 1");
  if (dtype == QS_C128) return collapse_t<double>((cx<double>*)vec, n, bit, outcome, scale, as_stream(stream));
  return collapse_t<float>((cx<float>*)vec, n, bit, outcome, scale, as_stream(stream));
}

int qs_sv_reset(qs_ctx* ctx, void* vec, int n, int dtype, int bit, double p0, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(bit >= 0 && bit < n, "bit outside the register");
  if (dtype == QS_C128) return reset_t<double>((cx<double>*)vec, n, bit, p0, as_stream(stream));
  return reset_t<float>((cx<float>*)vec, n, bit, p0, as_stream(stream));
}

int qs_sv_sample(qs_ctx* ctx, const void* vec, int n, int dtype, const double* uniforms, int shots, uint64_t* out,
                 void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && uniforms && out, "null argument");
  QS_REQUIRE(shots >= 0, "negative shot count");
  if (shots == 0) return QS_OK;
  if (dtype == QS_C128)
    return sample_t(ctx, LeafState<double>{(const cx<double>*)vec}, n, uniforms, shots, out, as_stream(stream));
  return sample_t(ctx, LeafState<float>{(const cx<float>*)vec}, n, uniforms, shots, out, as_stream(stream));
}

int qs_sv_sample_noalloc(qs_ctx* ctx, const void* vec, int n, int dtype, const double* uniforms, int shots,
                         uint64_t* out, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && uniforms && out, "null argument");
  QS_REQUIRE(shots >= 0, "negative shot count");
  if (shots == 0) return QS_OK;
  QS_REQUIRE(ctx->noalloc_scratch != nullptr, "qs_sv_sample_noalloc: the context has no fixed scratch");
  const SampleLayout L = sample_layout(n, shots);
  if (L.total() > ctx->noalloc_scratch_bytes) {
    set_error("qs_sv_sample_noalloc: the pyramid of 2^n - 1 doubles plus shots * (n + 1) * 8 bytes exceeds the "
              "fixed 40 MiB scratch (at most 22 qubits; fewer shots at 22)");
    return QS_ERR_ARG;
  }
  char* base = (char*)ctx->noalloc_scratch;
  if (dtype == QS_C128)
    return sample_in(LeafState<double>{(const cx<double>*)vec}, n, uniforms, shots, out, L, base,
                     as_stream(stream));
  return sample_in(LeafState<float>{(const cx<float>*)vec}, n, uniforms, shots, out, L, base, as_stream(stream));
}

int qs_sample_probs(qs_ctx* ctx, const double* probs_dev, int n, const double* uniforms, int shots, uint64_t* out,
                    void* stream) {
  QS_REQUIRE(ctx && probs_dev && uniforms && out, "null argument");
  QS_REQUIRE(n >= 1 && n <= 40, "qubit count must be in [1, 40]");
  QS_REQUIRE(shots >= 0, "negative shot count");
  if (shots == 0) return QS_OK;
  return sample_t(ctx, LeafProbs{probs_dev}, n, uniforms, shots, out, as_stream(stream));
}

int qs_sv_probabilities(qs_ctx* ctx, const void* vec, int n, int dtype, double* probs_dev, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && probs_dev, "null argument");
  if (dtype == QS_C128) return probs_t<double>((const cx<double>*)vec, n, probs_dev, as_stream(stream));
  return probs_t<float>((const cx<float>*)vec, n, probs_dev, as_stream(stream));
}

}  // extern "C"
