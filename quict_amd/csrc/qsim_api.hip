// C-ABI entry points: context, state initialisation, gate application,
// permutation, density-matrix readouts, Pauli sums and inner products.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace qs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int check_state_args(const void* vec, int n, int dtype) {
  QS_REQUIRE(vec != nullptr, "null state pointer");
  QS_REQUIRE(n >= 1 && n <= 40, "qubit count must be in [1, 40]");
  QS_REQUIRE(dtype == QS_C64 || dtype == QS_C128, "dtype must be QS_C64 or QS_C128");
  return QS_OK;
}

int ensure_scratch(qs_ctx* ctx, uint64_t bytes) {
  if (bytes <= ctx->scratch_bytes) return QS_OK;
  if (ctx->scratch) QS_CHECK_HIP(hipFree(ctx->scratch));
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  uint64_t want = std::max<uint64_t>(bytes, 1ull << 20);
  QS_CHECK_HIP(hipMalloc(&ctx->scratch, want));
  ctx->scratch_bytes = want;
  return QS_OK;
}

int acquire_prog_slot(qs_ctx* ctx, uint64_t host_bytes, uint64_t dev_bytes, qs_ctx::ProgSlot** out) {
  qs_ctx::ProgSlot& S = ctx->prog_slot[ctx->prog_next];
  ctx->prog_next ^= 1;
  if (!S.done) QS_CHECK_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
  if (S.pending) {  // the call before last: its upload and passes must be finished
    QS_CHECK_HIP(hipEventSynchronize(S.done));
    S.pending = false;
  }
  if (S.host_bytes < host_bytes) {
    if (S.host) QS_CHECK_HIP(hipHostFree(S.host));
    S.host = nullptr;
    S.host_bytes = 0;
    const uint64_t want = std::max<uint64_t>(host_bytes, 1ull << 16);
    QS_CHECK_HIP(hipHostMalloc((void**)&S.host, want, hipHostMallocDefault));
    S.host_bytes = want;
  }
  if (S.dev_bytes < dev_bytes) {
    if (S.dev) QS_CHECK_HIP(hipFree(S.dev));
    S.dev = nullptr;
    S.dev_bytes = 0;
    const uint64_t want = std::max<uint64_t>(dev_bytes, 1ull << 16);
    QS_CHECK_HIP(hipMalloc(&S.dev, want));
    S.dev_bytes = want;
  }
  *out = &S;
  return QS_OK;
}

int acquire_prog_slot(qs_ctx* ctx, uint64_t bytes, qs_ctx::ProgSlot** out) {
  return acquire_prog_slot(ctx, bytes, bytes, out);
}

int prof_begin(qs_ctx* ctx, int cls, hipStream_t s) {
  if (!ctx->profile) return QS_OK;
  const size_t i = ctx->ev_pending.size();
  if (i >= ctx->ev_pool.size()) {
    hipEvent_t a, b;
    QS_CHECK_HIP(hipEventCreate(&a));
    QS_CHECK_HIP(hipEventCreate(&b));
    ctx->ev_pool.push_back({a, b});
  }
  ctx->ev_pending.push_back({(int)i, cls});
  QS_CHECK_HIP(hipEventRecord(ctx->ev_pool[i].first, s));
  return QS_OK;
}

int prof_end(qs_ctx* ctx, hipStream_t s) {
  if (!ctx->profile || ctx->ev_pending.empty()) return QS_OK;
  QS_CHECK_HIP(hipEventRecord(ctx->ev_pool[ctx->ev_pending.back().first].second, s));
  return QS_OK;
}

int prof_collect(qs_ctx* ctx) {
  for (auto& pe : ctx->ev_pending) {
    float ms = 0;
    QS_CHECK_HIP(hipEventSynchronize(ctx->ev_pool[pe.first].second));
    QS_CHECK_HIP(hipEventElapsedTime(&ms, ctx->ev_pool[pe.first].first, ctx->ev_pool[pe.first].second));
    ctx->prof_ms[pe.second] += ms;
    ctx->prof_count[pe.second] += 1;
  }
  ctx->ev_pending.clear();
  return QS_OK;
}

// Streaming copy, 16 B per lane: every thread first issues its CP_U loads
// (strided by the block size, so each load instruction of a wave covers 1 KiB
// contiguously), then its CP_U stores.
constexpr int CP_U = 4;
template <bool NT>
__global__ __launch_bounds__(256) void k_copy16(const double2* __restrict__ src, double2* __restrict__ dst,
                                                uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * (256 * CP_U) + threadIdx.x;
  double2 v[CP_U];
#pragma unroll
  for (int k = 0; k < CP_U; ++k) {
    const uint64_t i = base + (uint64_t)k * 256;
    if (i < n) {
      if constexpr (NT) {
        v[k].x = __builtin_nontemporal_load(&src[i].x);
        v[k].y = __builtin_nontemporal_load(&src[i].y);
      } else {
        v[k] = src[i];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < CP_U; ++k) {
    const uint64_t i = base + (uint64_t)k * 256;
    if (i < n) {
      if constexpr (NT) {
        __builtin_nontemporal_store(v[k].x, &dst[i].x);
        __builtin_nontemporal_store(v[k].y, &dst[i].y);
      } else {
        dst[i] = v[k];
      }
    }
  }
}

template <typename R> __global__ void k_init(cx<R>* __restrict__ v, uint64_t N, uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride)
    st(v + i, mk<R>(i == index ? (R)1 : (R)0, (R)0));
}

// out[j] = in[src(j)], bit perm[b] of src = bit b of j.
struct PermBits {
  int b[64];
};

template <typename R>
__global__ void k_permute(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, uint64_t N, int n, PermBits pm) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  uint64_t src = 0;
  for (int b = 0; b < n; ++b) src |= ((j >> b) & 1ull) << pm.b[b];
  st(out + j, ld(in + src));
}

struct PauliTerms {
  int n;
  double coef[64];
  uint64_t x[64];
  uint64_t z[64];
};

// out[j] = sum_k c_k * phase_k(j ^ x_k) * v[j ^ x_k],   phase(i) = i^{|x&z|} (-1)^{|i&z|}
template <typename R>
__global__ void k_pauli_sum(const cx<R>* __restrict__ v, cx<R>* __restrict__ out, uint64_t N, PauliTerms pt,
                            int first, int accumulate) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double re = 0, im = 0;
  if (accumulate) {
    const cx<R> o = ld(out + j);
    re = o.re;
    im = o.im;
  }
  for (int k = 0; k < pt.n; ++k) {
    const uint64_t i = j ^ pt.x[k];
    const cx<R> a = ld(v + i);
    const int ny = __popcll(pt.x[k] & pt.z[k]) & 3;
    const int sign = __popcll(i & pt.z[k]) & 1;
    double ar = a.re, ai = a.im;
    // multiply by i^ny
    double pr, pi;
    switch (ny) {
      case 0: pr = ar; pi = ai; break;
      case 1: pr = -ai; pi = ar; break;
      case 2: pr = -ar; pi = -ai; break;
      default: pr = ai; pi = -ar; break;
    }
    const double c = sign ? -pt.coef[k] : pt.coef[k];
    re += c * pr;
    im += c * pi;
  }
  st(out + j, mk<R>((R)re, (R)im));
}

constexpr int IP_BLOCKS = 1024;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_inner_partial(const cx<R>* __restrict__ a, const cx<R>* __restrict__ b,
                                                       uint64_t N, double* __restrict__ part) {
  __shared__ double sre[4], sim[4];
  double re = 0, im = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const cx<R> x = ld(a + i), y = ld(b + i);
    re += (double)x.re * y.re + (double)x.im * y.im;
    im += (double)x.re * y.im - (double)x.im * y.re;
  }
  re = wsum(re);
  im = wsum(im);
  if ((threadIdx.x & 63) == 0) {
    sre[threadIdx.x >> 6] = re;
    sim[threadIdx.x >> 6] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0, m = 0;
    for (int w = 0; w < 4; ++w) {
      r += sre[w];
      m += sim[w];
    }
    part[2 * blockIdx.x] = r;
    part[2 * blockIdx.x + 1] = m;
  }
}

__global__ void k_inner_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sre[4], sim[4];
  double re = 0, im = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    re += part[2 * i];
    im += part[2 * i + 1];
  }
  re = wsum(re);
  im = wsum(im);
  if ((threadIdx.x & 63) == 0) {
    sre[threadIdx.x >> 6] = re;
    sim[threadIdx.x >> 6] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0, m = 0;
    for (int w = 0; w < 4; ++w) {
      r += sre[w];
      m += sim[w];
    }
    out[0] = r;
    out[1] = m;
  }
}

// rho as vec: index r*2^n + c.  diag[r] = Re rho[r, r].
template <typename R>
__global__ void k_dm_diag(const cx<R>* __restrict__ rho, int n, double* __restrict__ d, double* __restrict__ dim_) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t D = 1ull << n;
  if (r >= D) return;
  const cx<R> a = ld(rho + r * D + r);
  d[r] = a.re;
  if (dim_) dim_[r] = a.im;
}

template <typename R>
__global__ void k_dm_collapse(cx<R>* __restrict__ rho, int n, int bit, int outcome, R scale) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t N = 1ull << (2 * n);
  if (i >= N) return;
  const int rb = (i >> (n + bit)) & 1, cb = (i >> bit) & 1;
  if (rb != outcome || cb != outcome) {
    st(rho + i, mk<R>(0, 0));
  } else {
    st(rho + i, scale_(ld(rho + i), scale));
  }
}

template <typename R>
static int inner_t(qs_ctx* ctx, const cx<R>* a, const cx<R>* b, uint64_t N, double* re, double* im, hipStream_t s) {
  int rc = ensure_scratch(ctx, (2 * IP_BLOCKS + 64) * sizeof(double));
  if (rc) return rc;
  double* part = (double*)ctx->scratch;
  double* res = part + 2 * IP_BLOCKS;
  uint64_t blocks = std::min<uint64_t>(IP_BLOCKS, (N + 255) / 256);
  hipLaunchKernelGGL(k_inner_partial<R>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, N, part);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_inner_final, dim3(1), dim3(256), 0, s, part, (int)blocks, res);
  QS_CHECK_HIP(hipGetLastError());
  QS_CHECK_HIP(hipMemcpyAsync(ctx->host_pinned, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  *re = ctx->host_pinned[0];
  *im = ctx->host_pinned[1];
  return QS_OK;
}

int inner_partial_blocks() { return IP_BLOCKS; }

int inner_async(const void* a, const void* b, int n, int dtype, double* part, double* res, hipStream_t s) {
  const uint64_t N = 1ull << n;
  const uint64_t blocks = std::min<uint64_t>(IP_BLOCKS, (N + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_inner_partial<double>, dim3((unsigned)blocks), dim3(256), 0, s, (const cx<double>*)a,
                       (const cx<double>*)b, N, part);
  else
    hipLaunchKernelGGL(k_inner_partial<float>, dim3((unsigned)blocks), dim3(256), 0, s, (const cx<float>*)a,
                       (const cx<float>*)b, N, part);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_inner_final, dim3(1), dim3(256), 0, s, part, (int)blocks, res);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

int inner_product(qs_ctx* ctx, const void* a, const void* b, int n, int dtype, double* re, double* im,
                  hipStream_t s) {
  if (dtype == QS_C128) return inner_t<double>(ctx, (const cx<double>*)a, (const cx<double>*)b, 1ull << n, re, im, s);
  return inner_t<float>(ctx, (const cx<float>*)a, (const cx<float>*)b, 1ull << n, re, im, s);
}

}  // namespace qs

using namespace qs;

namespace qs {
int pauli_sum_async(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const qs_pauli_term* terms,
                    int nterms, hipStream_t s) {
  (void)ctx;
  const uint64_t N = 1ull << n;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  if (nterms == 0) {
    QS_CHECK_HIP(hipMemsetAsync(out, 0, N * amp_bytes(dtype), s));
    return QS_OK;
  }
  for (int first = 0; first < nterms; first += 64) {
    PauliTerms pt;
    pt.n = std::min(64, nterms - first);
    for (int k = 0; k < pt.n; ++k) {
      pt.coef[k] = terms[first + k].coef;
      pt.x[k] = terms[first + k].xmask;
      pt.z[k] = terms[first + k].zmask;
      QS_REQUIRE(((pt.x[k] | pt.z[k]) >> n) == 0, "Pauli mask outside the register");
    }
    if (dtype == QS_C128)
      hipLaunchKernelGGL(k_pauli_sum<double>, dim3(blocks), dim3(256), 0, s, (const cx<double>*)vec,
                         (cx<double>*)out, N, pt, first, first > 0);
    else
      hipLaunchKernelGGL(k_pauli_sum<float>, dim3(blocks), dim3(256), 0, s, (const cx<float>*)vec,
                         (cx<float>*)out, N, pt, first, first > 0);
    QS_CHECK_HIP(hipGetLastError());
  }
  return QS_OK;
}
}  // namespace qs

extern "C" {

const char* qs_last_error(void) { return g_last_error.c_str(); }

qs_ctx* qs_ctx_create(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    set_error("no such HIP device");
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return nullptr;
  }
  qs_ctx* ctx = new qs_ctx();
  ctx->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->num_cus = cus;
  if (hipHostMalloc((void**)&ctx->host_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    delete ctx;
    return nullptr;
  }
  if (noalloc_scratch_init(ctx) != QS_OK) {  // keeps its own error message
    (void)hipHostFree(ctx->host_pinned);
    delete ctx;
    return nullptr;
  }
  return ctx;
}

void qs_ctx_destroy(qs_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  for (auto& S : ctx->prog_slot) {
    if (S.dev) (void)hipFree(S.dev);
    if (S.host) (void)hipHostFree(S.host);
    if (S.done) (void)hipEventDestroy(S.done);
  }
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  for (auto& e : ctx->ev_pool) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  delete ctx;
}

int qs_profile_enable(qs_ctx* ctx, int on) {
  QS_REQUIRE(ctx, "null ctx");
  ctx->profile = on ? 1 : 0;
  return QS_OK;
}

int qs_profile_read(qs_ctx* ctx, double* ms, int64_t* launches, int reset) {
  QS_REQUIRE(ctx && ms && launches, "null argument");
  int rc = prof_collect(ctx);
  if (rc) return rc;
  for (int i = 0; i < 4; ++i) {
    ms[i] = ctx->prof_ms[i];
    launches[i] = ctx->prof_count[i];
    if (reset) {
      ctx->prof_ms[i] = 0;
      ctx->prof_count[i] = 0;
    }
  }
  return QS_OK;
}

uint64_t qs_ctx_scratch_bytes(const qs_ctx* ctx) {
  return ctx ? ctx->scratch_bytes + ctx->prog_slot[0].dev_bytes + ctx->prog_slot[1].dev_bytes : 0;
}

int qs_set_fusion(qs_ctx* ctx, int tile_log2, int low_bits) {
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(tile_log2 == 0 || (tile_log2 >= 11 && tile_log2 <= 13), "tile_log2 must be 0 or in [11, 13]");
  QS_REQUIRE(low_bits == 0 || (low_bits >= 2 && low_bits <= 8), "low_bits must be 0 or in [2, 8]");
  // a tile of 2^K amplitudes keeps its low bits on its K - 4 thread slots
  QS_REQUIRE(low_bits <= (tile_log2 ? tile_log2 : 12) - 4, "low_bits must not exceed tile_log2 - 4");
  ctx->tile_log2 = tile_log2;
  ctx->low_bits = low_bits;
  return QS_OK;
}

int qs_stream_copy(qs_ctx* ctx, void* dst, const void* src, uint64_t bytes, int nontemporal, void* stream) {
  QS_REQUIRE(ctx && dst && src && bytes % 16 == 0, "bad copy arguments (bytes must be a multiple of 16)");
  const uint64_t n16 = bytes / 16;
  const uint64_t blocks = (n16 + 256 * CP_U - 1) / (256 * CP_U);
  QS_REQUIRE(blocks <= 0x7fffffffull, "copy too large for one launch");
  int rc = prof_begin(ctx, PROF_GATE, as_stream(stream));
  if (rc) return rc;
  if (nontemporal)
    hipLaunchKernelGGL(k_copy16<true>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), (const double2*)src,
                       (double2*)dst, n16);
  else
    hipLaunchKernelGGL(k_copy16<false>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), (const double2*)src,
                       (double2*)dst, n16);
  QS_CHECK_HIP(hipGetLastError());
  return prof_end(ctx, as_stream(stream));
}

int qs_set_fusion_launch(qs_ctx* ctx, int prefetch, int wg_per_cu) {
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(wg_per_cu >= 1 && wg_per_cu <= 8, "wg_per_cu must be in [1, 8]");
  ctx->fused_prefetch = prefetch ? 1 : 0;
  ctx->fused_wg_per_cu = wg_per_cu;
  return QS_OK;
}

int qs_sv_init(qs_ctx* ctx, void* vec, int n, int dtype, uint64_t index, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  const uint64_t N = 1ull << n;
  QS_REQUIRE(index < N, "basis index outside the register");
  uint64_t blocks = std::min<uint64_t>((N + 255) / 256, 1ull << 20);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_init<double>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), (cx<double>*)vec, N,
                       index);
  else
    hipLaunchKernelGGL(k_init<float>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), (cx<float>*)vec, N,
                       index);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

int qs_sv_apply_gate(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gate, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  return apply_gate(ctx, vec, n, dtype, gate, as_stream(stream));
}

int qs_sv_apply_gates(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, int fuse,
                      int* passes, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(ngates >= 0 && (ngates == 0 || gates), "bad gate list");
  for (int i = 0; i < ngates; ++i) {
    rc = validate_gate(&gates[i], n);
    if (rc) return rc;
  }
  QS_REQUIRE(fuse >= 0 && fuse <= 2, "fuse must be 0, 1 or 2");
  if (fuse) return apply_gates_fused(ctx, vec, n, dtype, gates, ngates, fuse == 2, passes, as_stream(stream));
  for (int i = 0; i < ngates; ++i) {
    rc = prof_begin(ctx, PROF_GATE, as_stream(stream));
    if (rc) return rc;
    rc = apply_gate(ctx, vec, n, dtype, &gates[i], as_stream(stream));
    if (rc) return rc;
    rc = prof_end(ctx, as_stream(stream));
    if (rc) return rc;
  }
  if (passes) *passes = ngates;
  return QS_OK;
}

int qs_sv_permute(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const int32_t* perm, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && out && perm && out != vec, "bad arguments (out must differ from vec)");
  PermBits pm;
  uint64_t seen = 0;
  for (int b = 0; b < n; ++b) {
    QS_REQUIRE(perm[b] >= 0 && perm[b] < n && !(seen & (1ull << perm[b])), "perm must be a permutation");
    seen |= 1ull << perm[b];
    pm.b[b] = perm[b];
  }
  const uint64_t N = 1ull << n;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_permute<double>, dim3(blocks), dim3(256), 0, as_stream(stream), (const cx<double>*)vec,
                       (cx<double>*)out, N, n, pm);
  else
    hipLaunchKernelGGL(k_permute<float>, dim3(blocks), dim3(256), 0, as_stream(stream), (const cx<float>*)vec,
                       (cx<float>*)out, N, n, pm);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

int qs_sv_apply_pauli_sum(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const qs_pauli_term* terms,
                          int nterms, void* stream) {
  int rc = check_state_args(vec, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && out && out != vec && nterms >= 0 && (nterms == 0 || terms), "bad arguments");
  return pauli_sum_async(ctx, vec, out, n, dtype, terms, nterms, as_stream(stream));
}


int qs_sv_inner(qs_ctx* ctx, const void* a, const void* b, int n, int dtype, double* re, double* im, void* stream) {
  int rc = check_state_args(a, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && b && re && im, "null argument");
  return inner_product(ctx, a, b, n, dtype, re, im, as_stream(stream));
}

int qs_dm_diagonal(qs_ctx* ctx, const void* rho, int n, int dtype, double* diag_dev, void* stream) {
  int rc = check_state_args(rho, 2 * n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && diag_dev, "null argument");
  const uint64_t D = 1ull << n;
  const unsigned blocks = (unsigned)((D + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_dm_diag<double>, dim3(blocks), dim3(256), 0, as_stream(stream), (const cx<double>*)rho, n,
                       diag_dev, (double*)nullptr);
  else
    hipLaunchKernelGGL(k_dm_diag<float>, dim3(blocks), dim3(256), 0, as_stream(stream), (const cx<float>*)rho, n,
                       diag_dev, (double*)nullptr);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

int qs_dm_trace(qs_ctx* ctx, const void* rho, int n, int dtype, double* re, double* im, void* stream) {
  int rc = check_state_args(rho, 2 * n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && re && im, "null argument");
  const uint64_t D = 1ull << n;
  hipStream_t s = as_stream(stream);
  const uint64_t need = 2 * D * sizeof(double) + (2 * IP_BLOCKS + 64) * sizeof(double);
  rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  double* dre = (double*)ctx->scratch + (2 * IP_BLOCKS + 64);
  double* dim_ = dre + D;
  const unsigned blocks = (unsigned)((D + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_dm_diag<double>, dim3(blocks), dim3(256), 0, s, (const cx<double>*)rho, n, dre, dim_);
  else
    hipLaunchKernelGGL(k_dm_diag<float>, dim3(blocks), dim3(256), 0, s, (const cx<float>*)rho, n, dre, dim_);
  QS_CHECK_HIP(hipGetLastError());
  // sum of (re, im) pairs laid out as two arrays: reuse the inner-product reducer with
  // a = (1, 0) constant is awkward, so reduce each array separately via a tiny kernel.
  std::vector<double> h(2 * D);
  QS_CHECK_HIP(hipMemcpyAsync(h.data(), dre, 2 * D * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  double sr = 0, si = 0;
  for (uint64_t i = 0; i < D; ++i) {
    sr += h[i];
    si += h[D + i];
  }
  *re = sr;
  *im = si;
  return QS_OK;
}

int qs_dm_prob_zero(qs_ctx* ctx, const void* rho, int n, int dtype, int bit, double* p0, void* stream) {
  int rc = check_state_args(rho, 2 * n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx && p0, "null argument");
  QS_REQUIRE(bit >= 0 && bit < n, "bit outside the register");
  const uint64_t D = 1ull << n;
  hipStream_t s = as_stream(stream);
  rc = ensure_scratch(ctx, D * sizeof(double) + (2 * IP_BLOCKS + 64) * sizeof(double));
  if (rc) return rc;
  double* d = (double*)ctx->scratch + (2 * IP_BLOCKS + 64);
  const unsigned blocks = (unsigned)((D + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_dm_diag<double>, dim3(blocks), dim3(256), 0, s, (const cx<double>*)rho, n, d,
                       (double*)nullptr);
  else
    hipLaunchKernelGGL(k_dm_diag<float>, dim3(blocks), dim3(256), 0, s, (const cx<float>*)rho, n, d,
                       (double*)nullptr);
  QS_CHECK_HIP(hipGetLastError());
  std::vector<double> h(D);
  QS_CHECK_HIP(hipMemcpyAsync(h.data(), d, D * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  double acc = 0;
  for (uint64_t r = 0; r < D; ++r)
    if (!((r >> bit) & 1)) acc += h[r];
  *p0 = acc;
  return QS_OK;
}

int qs_dm_collapse(qs_ctx* ctx, void* rho, int n, int dtype, int bit, int outcome, double scale, void* stream) {
  int rc = check_state_args(rho, 2 * n, dtype);
  if (rc) return rc;
  QS_REQUIRE(ctx, "null ctx");
  QS_REQUIRE(bit >= 0 && bit < n && (outcome == 0 || outcome == 1), "bad bit/outcome");
  const uint64_t N = 1ull << (2 * n);
  const unsigned blocks = (unsigned)((N + 255) / 256);
  if (dtype == QS_C128)
    hipLaunchKernelGGL(k_dm_collapse<double>, dim3(blocks), dim3(256), 0, as_stream(stream), (cx<double>*)rho, n, bit,
                       outcome, (double)scale);
  else
    hipLaunchKernelGGL(k_dm_collapse<float>, dim3(blocks), dim3(256), 0, as_stream(stream), (cx<float>*)rho, n, bit,
                       outcome, (float)scale);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

}  // extern "C"
