// Dense 2^n x 2^n circuit matrices: identity, in-place transpose and
// matrix-vector products.
//
// The circuit unitary U is built column-major: the buffer holds W = U^T, whose
// row c is the state U|c>.  Every gate then acts on the low n bits of a 2n-bit
// vector, so the fused tile passes (qs_sv_apply_gates on 2n bits) apply the
// circuit to all 2^n columns at once with tiles made of gate bits only.
// qs_mat_transpose turns W into the row-major U when the caller wants the
// matrix itself; qs_mat_vec(trans=1) multiplies with W directly.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>

namespace qs {

// ---- identity ----------------------------------------------------------------
// One 16-byte chunk per thread and iteration (1 complex128 or 2 complex64).
template <typename R>
__global__ __launch_bounds__(256) void k_identity(cx<R>* __restrict__ m, int n) {
  constexpr int E = 16 / (2 * sizeof(R));
  const uint64_t D = 1ull << n;
  const uint64_t chunks = (D * D) / E;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < chunks; j += stride) {
    const uint64_t i0 = j * E;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint64_t i = i0 + e;
      st(m + i, mk<R>((i >> n) == (i & (D - 1)) ? (R)1 : (R)0, (R)0));
    }
  }
}

// ---- in-place square transpose -----------------------------------------------
// Workgroup (bi, bj), bi <= bj, swaps tile (bi, bj) with tile (bj, bi) through
// LDS; diagonal tiles transpose in place.  Tiles are T x T (t x t when the
// matrix is smaller); a tile row is T*S contiguous bytes.
template <typename R, int T>
__global__ __launch_bounds__(256) void k_transpose(cx<R>* __restrict__ m, int n, int t) {
  __shared__ cx<R> ta[T][T + 1];
  __shared__ cx<R> tb[T][T + 1];
  const uint64_t D = 1ull << n;
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  const uint64_t r0 = (uint64_t)bi * t, c0 = (uint64_t)bj * t;
  const int tt = t * t;
  // load A = m[r0.., c0..] and B = m[c0.., r0..]
  for (int k = threadIdx.x; k < tt; k += blockDim.x) {
    const int i = k / t, j = k % t;
    ta[i][j] = ld(m + (r0 + i) * D + c0 + j);
    if (bi != bj) tb[i][j] = ld(m + (c0 + i) * D + r0 + j);
  }
  __syncthreads();
  // m[r0+i][c0+j] <- B[j][i] (= old m[c0+j][r0+i]);  m[c0+i][r0+j] <- A[j][i]
  for (int k = threadIdx.x; k < tt; k += blockDim.x) {
    const int i = k / t, j = k % t;
    if (bi != bj) {
      st(m + (r0 + i) * D + c0 + j, tb[j][i]);
      st(m + (c0 + i) * D + r0 + j, ta[j][i]);
    } else {
      st(m + (r0 + i) * D + c0 + j, ta[j][i]);
    }
  }
}

// ---- matrix-vector products ---------------------------------------------------
__device__ __forceinline__ double wsum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// y = M x, M row-major.  One wave per MV_RW consecutive rows; lane l sums the
// columns l, l+64, ...; x is loaded once per column for all the wave's rows.
constexpr int MV_RW = 4;
template <typename R>
__global__ __launch_bounds__(256) void k_matvec_rows(const cx<R>* __restrict__ m, const cx<R>* __restrict__ x,
                                                     cx<R>* __restrict__ y, int n) {
  const uint64_t D = 1ull << n;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const uint64_t r0 = wave * MV_RW;
  if (r0 >= D) return;
  const int nr = (D - r0) < (uint64_t)MV_RW ? (int)(D - r0) : MV_RW;
  double ar[MV_RW], ai[MV_RW];
#pragma unroll
  for (int k = 0; k < MV_RW; ++k) ar[k] = ai[k] = 0.0;
#pragma unroll 4
  for (uint64_t c = lane; c < D; c += 64) {
    const cx<R> xv = ld(x + c);
#pragma unroll
    for (int k = 0; k < MV_RW; ++k) {
      if (k < nr) {
        const cx<R> a = ld(m + (r0 + k) * D + c);
        ar[k] = rfma((double)a.re, (double)xv.re, ar[k]);
        ar[k] = rfma(-(double)a.im, (double)xv.im, ar[k]);
        ai[k] = rfma((double)a.re, (double)xv.im, ai[k]);
        ai[k] = rfma((double)a.im, (double)xv.re, ai[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < MV_RW; ++k) {
    const double sr = wsum64(ar[k]), si = wsum64(ai[k]);
    if (lane == 0 && k < nr) st(y + r0 + k, mk<R>((R)sr, (R)si));
  }
}

// y = M^T x: y[r] = sum_c M[c][r] x[c].  Workgroup (strip, part): its 256
// threads own the 256 consecutive entries r of one strip and sum the rows c of
// column-part `part` (x[c] is uniform across the workgroup); partial sums go
// to part_out[part][r] in double and k_matvec_reduce adds them in order.
constexpr int MT_STRIP = 256;
template <typename R>
__global__ __launch_bounds__(256) void k_matvec_cols(const cx<R>* __restrict__ m, const cx<R>* __restrict__ x,
                                                     double* __restrict__ part_out, int n, uint64_t rows_per_part) {
  const uint64_t D = 1ull << n;
  const uint64_t r = (uint64_t)blockIdx.x * MT_STRIP + threadIdx.x;
  const uint64_t c_begin = (uint64_t)blockIdx.y * rows_per_part;
  const uint64_t c_end = (c_begin + rows_per_part) < D ? c_begin + rows_per_part : D;
  double sr = 0.0, si = 0.0;
  if (r < D) {
#pragma unroll 4
    for (uint64_t c = c_begin; c < c_end; ++c) {
      const cx<R> xv = ld(x + c);
      const cx<R> a = ld(m + c * D + r);
      sr = rfma((double)a.re, (double)xv.re, sr);
      sr = rfma(-(double)a.im, (double)xv.im, sr);
      si = rfma((double)a.re, (double)xv.im, si);
      si = rfma((double)a.im, (double)xv.re, si);
    }
    double* o = part_out + 2 * ((uint64_t)blockIdx.y * D + r);
    o[0] = sr;
    o[1] = si;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_matvec_reduce(const double* __restrict__ part, int nparts,
                                                       cx<R>* __restrict__ y, uint64_t D) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D) return;
  double sr = 0.0, si = 0.0;
  for (int p = 0; p < nparts; ++p) {
    sr += part[2 * ((uint64_t)p * D + r)];
    si += part[2 * ((uint64_t)p * D + r) + 1];
  }
  st(y + r, mk<R>((R)sr, (R)si));
}

static int check_mat_args(const void* m, int n, int dtype) {
  QS_REQUIRE(m != nullptr, "null matrix pointer");
  QS_REQUIRE(n >= 1 && n <= 20, "matrix qubit count must be in [1, 20]");
  QS_REQUIRE(dtype == QS_C64 || dtype == QS_C128, "dtype must be QS_C64 or QS_C128");
  return QS_OK;
}

template <typename R> static int mat_identity_t(void* m, int n, hipStream_t s) {
  constexpr int E = 16 / (2 * sizeof(R));
  const uint64_t D = 1ull << n;
  const uint64_t chunks = std::max<uint64_t>((D * D) / E, 1);
  const uint64_t blocks = std::min<uint64_t>((chunks + 255) / 256, 1ull << 16);
  hipLaunchKernelGGL(k_identity<R>, dim3((unsigned)blocks), dim3(256), 0, s, (cx<R>*)m, n);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

template <typename R, int T> static int mat_transpose_t(void* m, int n, hipStream_t s) {
  const uint64_t D = 1ull << n;
  const int t = (int)std::min<uint64_t>(T, D);
  const unsigned nt = (unsigned)(D / t);
  hipLaunchKernelGGL((k_transpose<R, T>), dim3(nt, nt), dim3(256), 0, s, (cx<R>*)m, n, t);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

template <typename R>
static int mat_vec_t(qs_ctx* ctx, const void* m, int n, int trans, const void* x, void* y, hipStream_t s) {
  const uint64_t D = 1ull << n;
  if (!trans) {
    const uint64_t waves = (D + MV_RW - 1) / MV_RW;
    const uint64_t blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(k_matvec_rows<R>, dim3((unsigned)blocks), dim3(256), 0, s, (const cx<R>*)m, (const cx<R>*)x,
                       (cx<R>*)y, n);
    QS_CHECK_HIP(hipGetLastError());
    return QS_OK;
  }
  // Enough workgroups to fill the device: strips x parts >= ~8 per CU.
  const uint64_t strips = (D + MT_STRIP - 1) / MT_STRIP;
  const uint64_t want = (uint64_t)ctx->num_cus * 8;
  uint64_t parts = std::max<uint64_t>(1, std::min<uint64_t>(D, want / strips));
  const uint64_t rows_per_part = (D + parts - 1) / parts;
  parts = (D + rows_per_part - 1) / rows_per_part;
  int rc = ensure_scratch(ctx, parts * D * 2 * sizeof(double));
  if (rc) return rc;
  double* part = (double*)ctx->scratch;
  hipLaunchKernelGGL(k_matvec_cols<R>, dim3((unsigned)strips, (unsigned)parts), dim3(MT_STRIP), 0, s,
                     (const cx<R>*)m, (const cx<R>*)x, part, n, rows_per_part);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_matvec_reduce<R>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, s, (const double*)part,
                     (int)parts, (cx<R>*)y, D);
  QS_CHECK_HIP(hipGetLastError());
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

int qs_mat_identity(qs_ctx* ctx, void* m, int n, int dtype, void* stream) {
  QS_REQUIRE(ctx, "null ctx");
  int rc = check_mat_args(m, n, dtype);
  if (rc) return rc;
  return dtype == QS_C128 ? mat_identity_t<double>(m, n, as_stream(stream))
                          : mat_identity_t<float>(m, n, as_stream(stream));
}

int qs_mat_transpose(qs_ctx* ctx, void* m, int n, int dtype, void* stream) {
  QS_REQUIRE(ctx, "null ctx");
  int rc = check_mat_args(m, n, dtype);
  if (rc) return rc;
  return dtype == QS_C128 ? mat_transpose_t<double, 32>(m, n, as_stream(stream))
                          : mat_transpose_t<float, 32>(m, n, as_stream(stream));
}

int qs_mat_vec(qs_ctx* ctx, const void* m, int n, int dtype, int trans, const void* x, void* y, void* stream) {
  QS_REQUIRE(ctx, "null ctx");
  int rc = check_mat_args(m, n, dtype);
  if (rc) return rc;
  QS_REQUIRE(x != nullptr && y != nullptr, "null vector pointer");
  QS_REQUIRE(x != y, "qs_mat_vec is out of place: x and y must differ");
  QS_REQUIRE(trans == 0 || trans == 1, "trans must be 0 or 1");
  return dtype == QS_C128 ? mat_vec_t<double>(ctx, m, n, trans, x, y, as_stream(stream))
                          : mat_vec_t<float>(ctx, m, n, trans, x, y, as_stream(stream));
}

}  // extern "C"
