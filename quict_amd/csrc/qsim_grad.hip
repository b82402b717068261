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
static void record_dense(const qs_gate& g, std::vector<double>& M) {
  const int t = g.ntargets, D = 1 << t;
  M.assign(2 * D * D, 0.0);
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

template <typename R>
static int adjoint_t(qs_ctx* ctx, cx<R>* psi, cx<R>* lam, int n, const qs_gate* gates, int ngates,
                     const qs_gate* dgates, const int32_t* slot, double* grads, int nslots, hipStream_t s) {
  // one partial row per parameterised step
  int nsteps = 0;
  for (int j = 0; j < ngates; ++j)
    if (slot[j] >= 0) ++nsteps;
  const uint64_t need = (uint64_t)std::max(nsteps, 1) * ADJ_MAX_BLOCKS * sizeof(double) +
                        (uint64_t)std::max(nsteps, 1) * (sizeof(double) + sizeof(int)) + 1024;
  int rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  double* partial = (double*)ctx->scratch;
  double* sums = partial + (uint64_t)std::max(nsteps, 1) * ADJ_MAX_BLOCKS;
  int* d_nb = (int*)(sums + std::max(nsteps, 1));
  std::vector<int> nb;
  std::vector<int> step_slot;
  rc = prof_begin(ctx, PROF_GRAD, s);
  if (rc) return rc;
  for (int j = ngates - 1; j >= 0; --j) {
    const qs_gate& g = gates[j];
    const bool grad = slot[j] >= 0;
    const int t = g.ntargets, c = g.ncontrols;
    if (t > 2 || (grad && (dgates[j].ntargets != t || dgates[j].ncontrols != c))) {
      if (grad) {
        set_error("adjoint gradient: parameterised gates must act on <= 2 targets");
        return QS_ERR_UNSUPPORTED;
      }
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
    double* prow = partial + (uint64_t)nb.size() * ADJ_MAX_BLOCKS;
    if (t == 1) {
      if (grad) hipLaunchKernelGGL((k_adj_step<R, 1, true>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
      else hipLaunchKernelGGL((k_adj_step<R, 1, false>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
    } else {
      if (grad) hipLaunchKernelGGL((k_adj_step<R, 2, true>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
      else hipLaunchKernelGGL((k_adj_step<R, 2, false>), dim3(blocks), dim3(ADJ_THREADS), 0, s, psi, lam, ngroups, a, prow);
    }
    QS_CHECK_HIP(hipGetLastError());
    if (grad) {
      nb.push_back(blocks);
      step_slot.push_back(slot[j]);
    }
  }
  if (!nb.empty()) {
    QS_CHECK_HIP(hipMemcpyAsync(d_nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)nb.size()), dim3(256), 0, s, partial, d_nb, ADJ_MAX_BLOCKS,
                       (int)nb.size(), sums);
    QS_CHECK_HIP(hipGetLastError());
  }
  rc = prof_end(ctx, s);
  if (rc) return rc;
  std::vector<double> h(nb.size());
  if (!nb.empty()) QS_CHECK_HIP(hipMemcpyAsync(h.data(), sums, nb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  for (size_t k = 0; k < nb.size(); ++k) {
    const int sl = step_slot[k];
    if (sl < nslots) grads[sl] += 2.0 * h[k];
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

}  // extern "C"
