// Shared device/host helpers for libqsim_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/qsim.h"

namespace qs {

// ---- complex arithmetic on interleaved (re, im) -----------------------------
template <typename R> struct cx {
  R re, im;
};

template <typename R> __host__ __device__ __forceinline__ cx<R> mk(R re, R im) { return cx<R>{re, im}; }
template <typename R> __host__ __device__ __forceinline__ cx<R> add(cx<R> a, cx<R> b) {
  return cx<R>{a.re + b.re, a.im + b.im};
}
template <typename R> __host__ __device__ __forceinline__ cx<R> mul(cx<R> a, cx<R> b) {
  return cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ double rfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__host__ __device__ __forceinline__ float rfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// acc + a*b
template <typename R> __host__ __device__ __forceinline__ cx<R> cfma(cx<R> acc, cx<R> a, cx<R> b) {
  acc.re = rfma(a.re, b.re, acc.re);
  acc.re = rfma(-a.im, b.im, acc.re);
  acc.im = rfma(a.re, b.im, acc.im);
  acc.im = rfma(a.im, b.re, acc.im);
  return acc;
}
template <typename R> __host__ __device__ __forceinline__ cx<R> scale_(cx<R> a, R s) {
  return cx<R>{a.re * s, a.im * s};
}
template <typename R> __host__ __device__ __forceinline__ R norm2(cx<R> a) { return a.re * a.re + a.im * a.im; }

// 16-byte / 8-byte vector loads and stores of one amplitude.
template <typename R> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

template <typename R> __device__ __forceinline__ cx<R> ld(const cx<R>* p) {
  using V = typename vec2<R>::type;
  V v = *reinterpret_cast<const V*>(p);
  return cx<R>{v.x, v.y};
}
template <typename R> __device__ __forceinline__ void st(cx<R>* p, cx<R> a) {
  using V = typename vec2<R>::type;
  V v;
  v.x = a.re;
  v.y = a.im;
  *reinterpret_cast<V*>(p) = v;
}

// Insert a zero bit at each position of `sorted` (ascending) into g.
__host__ __device__ __forceinline__ uint64_t insert_zeros(uint64_t g, const int* sorted, int k) {
  for (int j = 0; j < k; ++j) {
    const int b = sorted[j];
    g = ((g >> b) << (b + 1)) | (g & ((1ull << b) - 1));
  }
  return g;
}

// Parallel bit deposit of the low bits of v into the positions of `bits` (ascending order of v's bits).
__host__ __device__ __forceinline__ uint64_t deposit(uint64_t v, const int* bits, int k) {
  uint64_t r = 0;
  for (int j = 0; j < k; ++j) r |= ((v >> j) & 1ull) << bits[j];
  return r;
}

// ---- error handling ----------------------------------------------------------
void set_error(const std::string& msg);

#define QS_CHECK_HIP(expr)                                                                   \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      ::qs::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
      return QS_ERR_DEVICE;                                                                  \
    }                                                                                        \
  } while (0)

#define QS_REQUIRE(cond, msg)           \
  do {                                  \
    if (!(cond)) {                      \
      ::qs::set_error(msg);             \
      return QS_ERR_ARG;                \
    }                                   \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace qs

// ---- context -------------------------------------------------------------------
struct qs_ctx {
  int device = 0;
  int tile_log2 = 0;   // 0 = default per dtype
  int low_bits = 0;    // 0 = default
  int num_cus = 256;   // compute units of the device
  int fused_wg_per_cu = 2;  // resident workgroups per CU for the persistent tile kernel
  int fused_prefetch = 0;   // 1: persistent prefetching tile kernel, 0: one tile per workgroup
  // device scratch, grown on demand
  void* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  // pinned host staging for small results
  double* host_pinned = nullptr;
  // this device's copy of the fixed scratch of qs_sv_sample_noalloc (never freed)
  void* noalloc_scratch = nullptr;
  uint64_t noalloc_scratch_bytes = 0;
  // fused-pass programs: two slots used in turn, each a pinned host staging
  // buffer, a device buffer and an event recorded after the call's last pass,
  // so a call returns without waiting and the next-but-one call reuses a slot
  // only after its passes have finished
  struct ProgSlot {
    char* host = nullptr;
    uint64_t host_bytes = 0;
    void* dev = nullptr;
    uint64_t dev_bytes = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
  };
  ProgSlot prog_slot[2];
  int prog_next = 0;
  // kernel timing (HIP events recorded on the launch stream around each launch)
  int profile = 0;
  double prof_ms[4] = {0, 0, 0, 0};       // per kernel class, see qs_profile_read
  int64_t prof_count[4] = {0, 0, 0, 0};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<std::pair<int, int>> ev_pending;  // (pool index, class)
};

namespace qs {
// Grow-only scratch buffers owned by the context.
int ensure_scratch(qs_ctx* ctx, uint64_t bytes);
// A free program slot of at least `bytes` (host and device), waiting for its
// previous use to finish.
int acquire_prog_slot(qs_ctx* ctx, uint64_t bytes, qs_ctx::ProgSlot** out);
// The same with separate pinned-host and device sizes.
int acquire_prog_slot(qs_ctx* ctx, uint64_t host_bytes, uint64_t dev_bytes, qs_ctx::ProgSlot** out);
int check_state_args(const void* vec, int n, int dtype);
// Profiling: bracket a launch of kernel class `cls` with events (no-op unless enabled).
int prof_begin(qs_ctx* ctx, int cls, hipStream_t s);
int prof_end(qs_ctx* ctx, hipStream_t s);
// Fold completed event pairs into the totals (call after the stream is synchronised).
int prof_collect(qs_ctx* ctx);
enum ProfClass { PROF_FUSED = 0, PROF_GATE = 1, PROF_GRAD = 2, PROF_QNN = 3 };
}  // namespace qs
