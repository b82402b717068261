// Batched variational circuits: one persistent launch runs forward + adjoint
// backward for every circuit of a batch (no reference counterpart).
//
// Circuit c (width w = widths[c], qubit 0 = most significant amplitude bit):
//   angle encoding   Ry(x[c, j]) on qubit j mod w, j < nfeat
//   `layers` times   Ry(theta[l,0,q]) Rz(theta[l,1,q]) on every qubit q, then CX(q, q+1), q = 0..w-2
//   readout          E_c = <Z_0>,  loss_c = (E_c - target_c)^2,  loss = mean_c loss_c
//
// Mechanism.  Ry rotations about one axis add, so the encoding is the product
// state prod_q (cos(phi_q/2)|0> + sin(phi_q/2)|1>) with phi_q = sum_{j = q mod w} x_j:
// one pass.  A CX ladder is the GF(2)-linear permutation "prefix xor" of the
// qubit bits, so no amplitude is moved: the kernel keeps, per logical qubit q,
// the physical index mask c_q = P(e_q) and the parity mask r_q (logical bit q of
// a physical index p is parity(p & r_q)), and updates both after every ladder.
// A rotation on logical qubit q then pairs physical p with p ^ c_q.  Every
// circuit therefore costs w rotation passes per layer forward and w fused
// (gradient + inverse) passes per layer backward, on a state that lives in LDS
// (psi and lam, 2 x 2^w amplitudes) when w <= lds_max_width, otherwise in a
// per-workgroup global slot.  Workgroups pull circuits from an atomic queue
// ordered by decreasing width, so the widest circuits start first.
//
// Fusion depth D (1..3, a launch parameter the autotuner picks): the rotations
// of D consecutive qubits of a layer run in one pass over groups of 2^D
// amplitudes (a coset of the span of their D column masks, enumerated by
// inserting zeros at the masks' elimination pivots), so a layer costs
// ceil(w / D) passes forward and backward instead of w: D times fewer sweeps
// of a wide circuit's state through the memory of its slot.
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace qs {

constexpr int QNN_MAX_W = 20;
constexpr int QNN_MAX_LAYERS = 16;

struct QnnArgs {
  int fuse;               // rotations of `fuse` qubits per pass (1..3)
  int ncirc;
  int nfeat;
  int layers;
  int max_width;
  int lds_max_width;
  const int32_t* order;    // circuits, widest first
  const int32_t* widths;
  const double* x;         // ncirc x nfeat
  const double* target;    // ncirc
  const double* theta;     // layers x 2 x max_width
  double* expval;          // ncirc
  double* loss_terms;      // ncirc
  double* grad_circ;       // ncirc x (layers x 2 x max_width): each circuit's term of d(loss)/d(theta)
  double* grad_x;          // ncirc x nfeat
  void* slots;             // global per-workgroup state slots (2 x 2^max_width amplitudes each)
  int* counter;
  double inv_ncirc;
};

__device__ __forceinline__ int parity64(uint64_t v) { return __popcll(v) & 1; }

template <typename R>
__device__ __forceinline__ double block_reduce(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

template <typename R>
__device__ __forceinline__ void block_reduce2(double& a, double& b, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = a;
    red[2 * (threadIdx.x >> 6) + 1] = b;
  }
  __syncthreads();
  double sa = 0, sb = 0;
  for (int i = 0; i < nw; ++i) {
    sa += red[2 * i];
    sb += red[2 * i + 1];
  }
  a = sa;
  b = sb;
}

// Pair (p0, p1) of physical indices for logical qubit with column mask m and
// row mask r; h enumerates 2^(w-1) pairs.  Returns p0 with logical bit 0.
__device__ __forceinline__ void pair_of(uint64_t h, uint64_t m, uint64_t r, uint64_t& p0, uint64_t& p1) {
  const int pivot = 63 - __clzll(m);
  uint64_t a = ((h >> pivot) << (pivot + 1)) | (h & ((1ull << pivot) - 1));
  uint64_t b = a ^ m;
  if (parity64(a & r)) {
    p0 = b;
    p1 = a;
  } else {
    p0 = a;
    p1 = b;
  }
}

// Coset representatives of span{m_0..m_{D-1}}: indices with zeros at the
// masks' elimination pivots (ascending in piv); returns D.
template <int D>
__device__ __forceinline__ void group_pivots(const uint64_t* m, int* piv) {
  uint64_t red[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    uint64_t v = m[i];
#pragma unroll
    for (int j = 0; j < i; ++j)
      if ((v >> piv[j]) & 1) v ^= red[j];
    red[i] = v;
    piv[i] = 63 - __clzll(v);
  }
  // ascending order for zero insertion
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i + 1; j < D; ++j)
      if (piv[j] < piv[i]) {
        const int t = piv[i];
        piv[i] = piv[j];
        piv[j] = t;
      }
}

__device__ __forceinline__ uint64_t insert0(uint64_t h, int p) {
  return ((h >> p) << (p + 1)) | (h & ((1ull << p) - 1));
}

// Forward: apply U_i (2x2, row-major u[4 i .. 4 i + 3]) on logical qubit i of
// the group, i < D, to every group of psi.
template <typename R, int D>
__device__ __forceinline__ void qnn_fwd_group(cx<R>* psi, uint64_t ngroups, const uint64_t* m, const uint64_t* r,
                                              const cx<R>* u, int tid, int nt) {
  constexpr int G = 1 << D;
  int piv[D];
  group_pivots<D>(m, piv);
  uint64_t off[G];
#pragma unroll
  for (int sI = 0; sI < G; ++sI) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
      if ((sI >> i) & 1) o ^= m[i];
    off[sI] = o;
  }
  for (uint64_t h = tid; h < ngroups; h += nt) {
    uint64_t base = h;
#pragma unroll
    for (int i = 0; i < D; ++i) base = insert0(base, piv[i]);
    cx<R> a[G];
#pragma unroll
    for (int sI = 0; sI < G; ++sI) a[sI] = ld(psi + (base ^ off[sI]));
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const bool flip = parity64(base & r[i]);  // member with s_i = 1 holds logical 0
#pragma unroll
      for (int sI = 0; sI < G; ++sI) {
        if ((sI >> i) & 1) continue;
        const int s1 = sI | (1 << i);
        const cx<R> x0 = flip ? a[s1] : a[sI], x1 = flip ? a[sI] : a[s1];
        const cx<R> y0 = cfma(mul(u[4 * i], x0), u[4 * i + 1], x1), y1 = cfma(mul(u[4 * i + 2], x0), u[4 * i + 3], x1);
        a[sI] = flip ? y1 : y0;
        a[s1] = flip ? y0 : y1;
      }
    }
#pragma unroll
    for (int sI = 0; sI < G; ++sI) st(psi + (base ^ off[sI]), a[sI]);
  }
}

// Backward over one group of D qubits: for each qubit i the Rz step (gradient
// gz_i = Im <lam|Z|psi>, then Rz^dagger on both) and the Ry step (gy_i =
// Im <lam|Y|psi>, then Ry^dagger); different qubits commute.
template <typename R, int D>
__device__ __forceinline__ void qnn_bwd_group(cx<R>* psi, cx<R>* lam, uint64_t ngroups, const uint64_t* m,
                                              const uint64_t* r, const double* cy, const double* sy,
                                              const double* cz, const double* sz, double* gy, double* gz, int tid,
                                              int nt) {
  constexpr int G = 1 << D;
  int piv[D];
  group_pivots<D>(m, piv);
  uint64_t off[G];
#pragma unroll
  for (int sI = 0; sI < G; ++sI) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
      if ((sI >> i) & 1) o ^= m[i];
    off[sI] = o;
  }
  for (uint64_t h = tid; h < ngroups; h += nt) {
    uint64_t base = h;
#pragma unroll
    for (int i = 0; i < D; ++i) base = insert0(base, piv[i]);
    cx<R> a[G], b[G];
#pragma unroll
    for (int sI = 0; sI < G; ++sI) {
      a[sI] = ld(psi + (base ^ off[sI]));
      b[sI] = ld(lam + (base ^ off[sI]));
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const bool flip = parity64(base & r[i]);
      const cx<R> ep = mk<R>((R)cz[i], (R)sz[i]), em = mk<R>((R)cz[i], (R)-sz[i]);
#pragma unroll
      for (int sI = 0; sI < G; ++sI) {
        if ((sI >> i) & 1) continue;
        const int s1 = sI | (1 << i);
        const int j0 = flip ? s1 : sI, j1 = flip ? sI : s1;  // logical 0 / 1 members
        cx<R> a0 = a[j0], a1 = a[j1], b0 = b[j0], b1 = b[j1];
        gz[i] += ((double)b0.re * a0.im - (double)b0.im * a0.re) - ((double)b1.re * a1.im - (double)b1.im * a1.re);
        a0 = mul(ep, a0);
        b0 = mul(ep, b0);
        a1 = mul(em, a1);
        b1 = mul(em, b1);
        gy[i] += -((double)b0.re * a1.re + (double)b0.im * a1.im) + ((double)b1.re * a0.re + (double)b1.im * a0.im);
        const R c = (R)cy[i], sn = (R)sy[i];
        a[j0] = mk<R>(c * a0.re + sn * a1.re, c * a0.im + sn * a1.im);
        a[j1] = mk<R>(-sn * a0.re + c * a1.re, -sn * a0.im + c * a1.im);
        b[j0] = mk<R>(c * b0.re + sn * b1.re, c * b0.im + sn * b1.im);
        b[j1] = mk<R>(-sn * b0.re + c * b1.re, -sn * b0.im + c * b1.im);
      }
    }
#pragma unroll
    for (int sI = 0; sI < G; ++sI) {
      st(psi + (base ^ off[sI]), a[sI]);
      st(lam + (base ^ off[sI]), b[sI]);
    }
  }
}

template <typename R>
__global__ __launch_bounds__(1024) void k_qnn(QnnArgs A) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ double red[64];
  __shared__ uint64_t cm[QNN_MAX_W], rm[QNN_MAX_W];
  __shared__ double gacc[QNN_MAX_LAYERS * 2 * QNN_MAX_W];
  __shared__ double enc_c[QNN_MAX_W], enc_s[QNN_MAX_W];
  __shared__ int cur;
  const int tid = threadIdx.x, nt = blockDim.x;
  cx<R>* slot_psi = reinterpret_cast<cx<R>*>(A.slots) + (uint64_t)blockIdx.x * 2 * (1ull << A.max_width);
  cx<R>* slot_lam = slot_psi + (1ull << A.max_width);

  for (;;) {
    if (tid == 0) cur = atomicAdd(A.counter, 1);
    __syncthreads();
    const int k = cur;
    __syncthreads();
    if (k >= A.ncirc) break;
    const int c = A.order[k];
    const int w = A.widths[c];
    const uint64_t N = 1ull << w, half = N >> 1;
    cx<R>* psi;
    cx<R>* lam;
    if (w <= A.lds_max_width) {
      psi = reinterpret_cast<cx<R>*>(lds_raw);
      lam = psi + N;
    } else {
      psi = slot_psi;
      lam = slot_lam;
    }
    const double* th = A.theta;
    const int MW = A.max_width;
    // encoding angles and product state
    for (int i = tid; i < 2 * A.layers * MW; i += nt) gacc[i] = 0;
    if (tid < w) {
      cm[tid] = 1ull << (w - 1 - tid);
      rm[tid] = 1ull << (w - 1 - tid);
      double phi = 0;
      for (int j = tid; j < A.nfeat; j += w) phi += A.x[(uint64_t)c * A.nfeat + j];
      enc_c[tid] = cos(0.5 * phi);
      enc_s[tid] = sin(0.5 * phi);
    }
    __syncthreads();
    for (uint64_t i = tid; i < N; i += nt) {
      double re = 1;
      for (int q = 0; q < w; ++q) re *= ((i >> (w - 1 - q)) & 1) ? enc_s[q] : enc_c[q];
      st(psi + i, mk<R>((R)re, (R)0));
    }
    __syncthreads();
    // forward
    for (int l = 0; l < A.layers; ++l) {
      for (int q0 = 0; q0 < w; q0 += A.fuse) {
        const int d = min(A.fuse, w - q0);
        uint64_t m[3], r[3];
        cx<R> u[12];
        for (int i = 0; i < d; ++i) {
          const int q = q0 + i;
          const double ty = th[(l * 2 + 0) * MW + q], tz = th[(l * 2 + 1) * MW + q];
          const double cy = cos(0.5 * ty), sy = sin(0.5 * ty), cz = cos(0.5 * tz), sz = sin(0.5 * tz);
          // U = Rz(tz) Ry(ty) = [[e^{-i tz/2} cy, -e^{-i tz/2} sy], [e^{i tz/2} sy, e^{i tz/2} cy]]
          u[4 * i] = mk<R>((R)(cz * cy), (R)(-sz * cy));
          u[4 * i + 1] = mk<R>((R)(-cz * sy), (R)(sz * sy));
          u[4 * i + 2] = mk<R>((R)(cz * sy), (R)(sz * sy));
          u[4 * i + 3] = mk<R>((R)(cz * cy), (R)(sz * cy));
          m[i] = cm[q];
          r[i] = rm[q];
        }
        if (d == 1) qnn_fwd_group<R, 1>(psi, N >> 1, m, r, u, tid, nt);
        else if (d == 2) qnn_fwd_group<R, 2>(psi, N >> 2, m, r, u, tid, nt);
        else qnn_fwd_group<R, 3>(psi, N >> 3, m, r, u, tid, nt);
        __syncthreads();
      }
      // CX ladder: r'_k = r_0 ^ .. ^ r_k ; c'_q = c_q ^ c_{q+1}
      if (tid == 0) {
        for (int q = 1; q < w; ++q) rm[q] ^= rm[q - 1];
        for (int q = 0; q + 1 < w; ++q) cm[q] ^= cm[q + 1];
      }
      __syncthreads();
    }
    // E = <Z_0>, lam = Z_0 psi
    double e = 0;
    {
      const uint64_t r0 = rm[0];
      for (uint64_t i = tid; i < N; i += nt) {
        const cx<R> a = ld(psi + i);
        const double p = (double)a.re * a.re + (double)a.im * a.im;
        const bool one = parity64(i & r0);
        e += one ? -p : p;
        st(lam + i, one ? mk<R>(-a.re, -a.im) : a);
      }
    }
    e = block_reduce<R>(e, red);
    const double tgt = A.target[c];
    const double dl = 2.0 * (e - tgt) * A.inv_ncirc;  // d(mean loss)/dE_c
    // backward
    for (int l = A.layers - 1; l >= 0; --l) {
      if (tid == 0) {
        for (int q = w - 1; q >= 1; --q) rm[q] ^= rm[q - 1];
        for (int q = w - 2; q >= 0; --q) cm[q] ^= cm[q + 1];
      }
      __syncthreads();
      for (int q0 = 0; q0 < w; q0 += A.fuse) {
        const int d = min(A.fuse, w - q0);
        uint64_t m[3], r[3];
        double cy[3], sy[3], cz[3], sz[3], gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
        for (int i = 0; i < d; ++i) {
          const int q = q0 + i;
          const double ty = th[(l * 2 + 0) * MW + q], tz = th[(l * 2 + 1) * MW + q];
          cy[i] = cos(0.5 * ty);
          sy[i] = sin(0.5 * ty);
          cz[i] = cos(0.5 * tz);
          sz[i] = sin(0.5 * tz);
          m[i] = cm[q];
          r[i] = rm[q];
        }
        // per qubit: gz = Im <lam|Z|psi>, Rz^dagger = diag(e^{i tz/2}, e^{-i tz/2}) on both,
        // gy = Im <lam|Y|psi> (Y psi = (-i a1, i a0)), Ry^dagger = [[cy, sy], [-sy, cy]] on both
        if (d == 1) qnn_bwd_group<R, 1>(psi, lam, N >> 1, m, r, cy, sy, cz, sz, gy, gz, tid, nt);
        else if (d == 2) qnn_bwd_group<R, 2>(psi, lam, N >> 2, m, r, cy, sy, cz, sz, gy, gz, tid, nt);
        else qnn_bwd_group<R, 3>(psi, lam, N >> 3, m, r, cy, sy, cz, sz, gy, gz, tid, nt);
        for (int i = 0; i < d; ++i) {
          block_reduce2<R>(gz[i], gy[i], red);
          if (tid == 0) {
            gacc[(l * 2 + 0) * MW + q0 + i] += gy[i];
            gacc[(l * 2 + 1) * MW + q0 + i] += gz[i];
          }
        }
        __syncthreads();
      }
    }
    // encoding gradient: dE/dphi_q = Im <lam| Y_q |psi> (identity map again)
    for (int q = 0; q < w; ++q) {
      const uint64_t m = 1ull << (w - 1 - q);
      double gy = 0, dummy = 0;
      for (uint64_t h = tid; h < half; h += nt) {
        uint64_t p0, p1;
        pair_of(h, m, m, p0, p1);
        const cx<R> a0 = ld(psi + p0), a1 = ld(psi + p1), b0 = ld(lam + p0), b1 = ld(lam + p1);
        gy += -((double)b0.re * a1.re + (double)b0.im * a1.im) + ((double)b1.re * a0.re + (double)b1.im * a0.im);
      }
      block_reduce2<R>(gy, dummy, red);
      if (tid == 0)
        for (int j = q; j < A.nfeat; j += w) A.grad_x[(uint64_t)c * A.nfeat + j] = dl * gy;
      __syncthreads();
    }
    if (tid == 0) {
      A.expval[c] = e;
      A.loss_terms[c] = (e - tgt) * (e - tgt);
    }
    // this circuit's row of the per-circuit gradient terms (summed in circuit
    // order by k_qnn_reduce: the result does not depend on which workgroup ran
    // which circuit)
    for (int i = tid; i < A.layers * 2 * MW; i += nt) {
      const int q = i % MW;
      A.grad_circ[(uint64_t)c * A.layers * 2 * MW + i] = q < w ? dl * gacc[i] : 0.0;
    }
    __syncthreads();
  }
}

__global__ void k_qnn_clear(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// grad_theta[i] = sum over circuits c = 0, 1, ... (in that order) of grad_circ[c][i]
__global__ void k_qnn_reduce(const double* __restrict__ grad_circ, int ncirc, int nth, double* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nth) return;
  double s = 0;
  for (int c = 0; c < ncirc; ++c) s += grad_circ[(uint64_t)c * nth + i];
  g[i] = s;
}

struct QnnPlan {
  int dtype = -1, ncirc = 0, nfeat = 0, layers = 0, max_width = 0, lds_max_width = 0, threads = 0, blocks = 0,
      fuse = 0;
  bool slots = false;  // the buffer holds the global state slots (some circuit wider than the LDS limit)
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t stream = nullptr;
  void* buf = nullptr;
  uint64_t buf_bytes = 0;
};

static QnnPlan g_plan;  // one cached plan (graph + buffers) per process

template <typename R>
static int qnn_t(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths, int nfeat, const double* x,
                 const double* target, int layers, int max_width, const double* theta, double* expval, double* loss,
                 double* grad_theta, double* grad_x, const qs_qnn_config* cfg, hipStream_t user_stream) {
  int lds_w = cfg ? cfg->lds_max_width : 11;
  int threads = cfg ? cfg->block_threads : 256;
  int per_cu = cfg ? cfg->blocks_per_cu : 2;
  const int use_graph = cfg ? cfg->use_graph : 1;
  const int fuse = (cfg && cfg->fuse_qubits) ? cfg->fuse_qubits : 1;
  QS_REQUIRE(fuse >= 1 && fuse <= 3, "fuse_qubits must be in [1, 3] (0: 1)");
  QS_REQUIRE(threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024,
             "block_threads must be a power of two in [64, 1024]");
  QS_REQUIRE(lds_w >= 0 && lds_w <= 13, "lds_max_width must be in [0, 13]");
  QS_REQUIRE(per_cu >= 1 && per_cu <= 8, "blocks_per_cu must be in [1, 8]");
  const size_t lds_bytes = lds_w > 0 ? (size_t)2 * (1ull << lds_w) * sizeof(cx<R>) : 0;
  QS_REQUIRE(lds_bytes <= 150 * 1024, "LDS state too large for one workgroup");
  int dev = ctx->device, cus = 0;
  QS_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = std::max(1, cus * per_cu);
  // device buffers: order, widths, x, target, theta, expval, loss, grad_theta, grad_x, counter,
  // per-circuit gradient terms, slots
  // global state slots only when some circuit is wider than the LDS limit
  int wmax = 0;
  for (int i = 0; i < ncirc; ++i) wmax = std::max(wmax, (int)widths[i]);
  const uint64_t slot_bytes = wmax > lds_w ? (uint64_t)blocks * 2 * (1ull << max_width) * sizeof(cx<R>) : 0;
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t nth = (uint64_t)layers * 2 * max_width;
  const uint64_t sz_i = al(ncirc * sizeof(int32_t)), sz_x = al((uint64_t)ncirc * nfeat * sizeof(double)),
                 sz_c = al(ncirc * sizeof(double)), sz_t = al(nth * sizeof(double));
  const uint64_t sz_gc = al((uint64_t)ncirc * nth * sizeof(double));
  const uint64_t total = 2 * sz_i + 2 * sz_x + 3 * sz_c + 2 * sz_t + 256 + sz_gc + slot_bytes;
  QnnPlan& P = g_plan;
  const bool same = P.exec && P.dtype == dtype && P.ncirc == ncirc && P.nfeat == nfeat && P.layers == layers &&
                    P.max_width == max_width && P.lds_max_width == lds_w && P.threads == threads &&
                    P.blocks == blocks && P.fuse == fuse && P.slots == (slot_bytes != 0);
  if (!same) {
    if (P.exec) (void)hipGraphExecDestroy(P.exec);
    if (P.graph) (void)hipGraphDestroy(P.graph);
    P.exec = nullptr;
    P.graph = nullptr;
    if (!P.stream) QS_CHECK_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    if (P.buf_bytes < total) {
      if (P.buf) QS_CHECK_HIP(hipFree(P.buf));
      P.buf = nullptr;
      QS_CHECK_HIP(hipMalloc(&P.buf, total));
      P.buf_bytes = total;
    }
  }
  char* b = (char*)P.buf;
  int32_t* d_order = (int32_t*)b;
  int32_t* d_w = (int32_t*)(b + sz_i);
  double* d_x = (double*)(b + 2 * sz_i);
  double* d_gx = (double*)(b + 2 * sz_i + sz_x);
  double* d_t = (double*)(b + 2 * sz_i + 2 * sz_x);
  double* d_e = (double*)(b + 2 * sz_i + 2 * sz_x + sz_c);
  double* d_l = (double*)(b + 2 * sz_i + 2 * sz_x + 2 * sz_c);
  double* d_th = (double*)(b + 2 * sz_i + 2 * sz_x + 3 * sz_c);
  double* d_gth = (double*)(b + 2 * sz_i + 2 * sz_x + 3 * sz_c + sz_t);
  int* d_cnt = (int*)(b + 2 * sz_i + 2 * sz_x + 3 * sz_c + 2 * sz_t);
  double* d_gc = (double*)(b + 2 * sz_i + 2 * sz_x + 3 * sz_c + 2 * sz_t + 256);
  void* d_slots = b + 2 * sz_i + 2 * sz_x + 3 * sz_c + 2 * sz_t + 256 + sz_gc;

  // widest first
  std::vector<int32_t> order(ncirc);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c2) { return widths[a] > widths[c2]; });
  hipStream_t s = P.stream;
  QS_CHECK_HIP(hipMemcpyAsync(d_order, order.data(), ncirc * sizeof(int32_t), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_w, widths, ncirc * sizeof(int32_t), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_x, x, (uint64_t)ncirc * nfeat * sizeof(double), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_t, target, ncirc * sizeof(double), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_th, theta, nth * sizeof(double), hipMemcpyHostToDevice, s));

  QnnArgs A;
  A.fuse = fuse;
  A.ncirc = ncirc;
  A.nfeat = nfeat;
  A.layers = layers;
  A.max_width = max_width;
  A.lds_max_width = lds_w;
  A.order = d_order;
  A.widths = d_w;
  A.x = d_x;
  A.target = d_t;
  A.theta = d_th;
  A.expval = d_e;
  A.loss_terms = d_l;
  A.grad_circ = d_gc;
  A.grad_x = d_gx;
  A.slots = d_slots;
  A.counter = d_cnt;
  A.inv_ncirc = 1.0 / ncirc;
  if (lds_bytes > 64 * 1024)
    QS_CHECK_HIP(hipFuncSetAttribute((const void*)k_qnn<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_bytes));
  int rc = prof_begin(ctx, PROF_QNN, s);
  if (rc) return rc;
  if (use_graph) {
    if (!same) {
      // queue reset + persistent kernel + ordered reduction, captured once and
      // replayed for every batch of this shape
      QS_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(k_qnn_clear, dim3(1), dim3(64), 0, s, d_cnt);
      hipLaunchKernelGGL(k_qnn<R>, dim3(blocks), dim3(threads), lds_bytes, s, A);
      hipLaunchKernelGGL(k_qnn_reduce, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, s, d_gc, ncirc, (int)nth,
                         d_gth);
      QS_CHECK_HIP(hipStreamEndCapture(s, &P.graph));
      QS_CHECK_HIP(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
    }
    QS_CHECK_HIP(hipGraphLaunch(P.exec, s));
  } else {
    hipLaunchKernelGGL(k_qnn_clear, dim3(1), dim3(64), 0, s, d_cnt);
    QS_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_qnn<R>, dim3(blocks), dim3(threads), lds_bytes, s, A);
    QS_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_qnn_reduce, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, s, d_gc, ncirc, (int)nth,
                       d_gth);
    QS_CHECK_HIP(hipGetLastError());
  }
  rc = prof_end(ctx, s);
  if (rc) return rc;
  P.dtype = dtype;
  P.ncirc = ncirc;
  P.nfeat = nfeat;
  P.layers = layers;
  P.max_width = max_width;
  P.lds_max_width = lds_w;
  P.threads = threads;
  P.blocks = blocks;
  P.fuse = fuse;
  P.slots = slot_bytes != 0;
  std::vector<double> lt(ncirc);
  QS_CHECK_HIP(hipMemcpyAsync(expval, d_e, ncirc * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipMemcpyAsync(lt.data(), d_l, ncirc * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipMemcpyAsync(grad_theta, d_gth, nth * sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad_x)
    QS_CHECK_HIP(hipMemcpyAsync(grad_x, d_gx, (uint64_t)ncirc * nfeat * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  (void)user_stream;
  double L = 0;
  for (int i = 0; i < ncirc; ++i) L += lt[i];
  *loss = L / ncirc;
  return QS_OK;
}

// ---- per-circuit gate programs --------------------------------------------------
// Each circuit runs its own op list (qs_qnn_op): Rx / Ry / Rz on a logical qubit
// with a shared trainable angle theta[param] or a fixed value, Ry(x[c, j])
// encodings, CX (a GF(2) update of the qubit masks, no amplitude moves) and CZ
// (one sign pass).  Forward, E_c = <Z_0>, lam = Z_0 psi, then the ops in reverse:
// a trainable rotation exp(-i t P / 2) contributes Im <lam|P|psi> (both at the
// point after the gate) to d E / d t before its inverse is applied to both;
// an encoding contributes the same to d E / d x[c, j].  One rotation per pass.
struct QnnProgArgs {
  int ncirc, nfeat, nparams, lds_max_width, max_width;
  const int32_t* order;   // circuits, widest first
  const int32_t* widths;
  const int32_t* off;     // ncirc + 1 op offsets
  const qs_qnn_op* ops;
  const double* x;        // ncirc x nfeat
  const double* target;
  const double* theta;    // nparams
  double* expval;
  double* loss_terms;
  double* grad_circ;      // ncirc x nparams
  double* grad_x;         // ncirc x nfeat
  void* slots;
  int* counter;
  double inv_ncirc;
};

// 2x2 rotation exp(-i t P / 2), P = X, Y, Z (kinds QS_QNN_RX, RY / ENC_RY, RZ)
template <typename R> __device__ __forceinline__ void rot_matrix(int axis, double t, cx<R>* u) {
  const double c = cos(0.5 * t), s = sin(0.5 * t);
  if (axis == 0) {  // X: [[c, -i s], [-i s, c]]
    u[0] = mk<R>((R)c, (R)0);
    u[1] = mk<R>((R)0, (R)-s);
    u[2] = mk<R>((R)0, (R)-s);
    u[3] = mk<R>((R)c, (R)0);
  } else if (axis == 1) {  // Y: [[c, -s], [s, c]]
    u[0] = mk<R>((R)c, (R)0);
    u[1] = mk<R>((R)-s, (R)0);
    u[2] = mk<R>((R)s, (R)0);
    u[3] = mk<R>((R)c, (R)0);
  } else {  // Z: diag(e^{-i t/2}, e^{i t/2})
    u[0] = mk<R>((R)c, (R)-s);
    u[1] = mk<R>((R)0, (R)0);
    u[2] = mk<R>((R)0, (R)0);
    u[3] = mk<R>((R)c, (R)s);
  }
}

__device__ __forceinline__ int op_axis(int kind) { return kind == QS_QNN_RX ? 0 : (kind == QS_QNN_RZ ? 2 : 1); }

template <typename R>
__global__ __launch_bounds__(1024) void k_qnn_prog(QnnProgArgs A) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ double red[64];
  __shared__ uint64_t cm[QNN_MAX_W], rm[QNN_MAX_W];
  __shared__ int cur;
  const int tid = threadIdx.x, nt = blockDim.x;
  cx<R>* slot_psi = reinterpret_cast<cx<R>*>(A.slots) + (uint64_t)blockIdx.x * 2 * (1ull << A.max_width);
  cx<R>* slot_lam = slot_psi + (1ull << A.max_width);
  for (;;) {
    if (tid == 0) cur = atomicAdd(A.counter, 1);
    __syncthreads();
    const int k = cur;
    __syncthreads();
    if (k >= A.ncirc) break;
    const int c = A.order[k];
    const int w = A.widths[c];
    const uint64_t N = 1ull << w, half = N >> 1;
    cx<R>* psi;
    cx<R>* lam;
    if (w <= A.lds_max_width) {
      psi = reinterpret_cast<cx<R>*>(lds_raw);
      lam = psi + N;
    } else {
      psi = slot_psi;
      lam = slot_lam;
    }
    const qs_qnn_op* ops = A.ops + A.off[c];
    const int nops = A.off[c + 1] - A.off[c];
    if (tid < w) cm[tid] = rm[tid] = 1ull << (w - 1 - tid);
    for (uint64_t i = tid; i < N; i += nt) st(psi + i, mk<R>((R)(i == 0 ? 1 : 0), (R)0));
    __syncthreads();
    // forward
    for (int o = 0; o < nops; ++o) {
      const qs_qnn_op op = ops[o];
      if (op.kind == QS_QNN_CX) {
        if (tid == 0) {
          rm[op.q1] ^= rm[op.q0];
          cm[op.q0] ^= cm[op.q1];
        }
      } else if (op.kind == QS_QNN_CZ) {
        const uint64_t ra = rm[op.q0], rb = rm[op.q1];
        for (uint64_t i = tid; i < N; i += nt)
          if (parity64(i & ra) & parity64(i & rb)) {
            const cx<R> a = ld(psi + i);
            st(psi + i, mk<R>(-a.re, -a.im));
          }
      } else {
        const double t = op.kind == QS_QNN_ENC_RY ? A.x[(uint64_t)c * A.nfeat + op.param]
                                                  : (op.param >= 0 ? A.theta[op.param] : op.value);
        cx<R> u[4];
        rot_matrix<R>(op_axis(op.kind), t, u);
        const uint64_t m = cm[op.q0], r = rm[op.q0];
        for (uint64_t h = tid; h < half; h += nt) {
          uint64_t p0, p1;
          pair_of(h, m, r, p0, p1);
          const cx<R> a0 = ld(psi + p0), a1 = ld(psi + p1);
          st(psi + p0, cfma(mul(u[0], a0), u[1], a1));
          st(psi + p1, cfma(mul(u[2], a0), u[3], a1));
        }
      }
      __syncthreads();
    }
    double e = 0;
    {
      const uint64_t r0 = rm[0];
      for (uint64_t i = tid; i < N; i += nt) {
        const cx<R> a = ld(psi + i);
        const double pr = (double)a.re * a.re + (double)a.im * a.im;
        const bool one = parity64(i & r0);
        e += one ? -pr : pr;
        st(lam + i, one ? mk<R>(-a.re, -a.im) : a);
      }
    }
    e = block_reduce<R>(e, red);
    const double tgt = A.target[c];
    const double dl = 2.0 * (e - tgt) * A.inv_ncirc;
    double* grow = A.grad_circ + (uint64_t)c * A.nparams;
    for (int i = tid; i < A.nparams; i += nt) grow[i] = 0.0;
    for (int j = tid; j < A.nfeat; j += nt) A.grad_x[(uint64_t)c * A.nfeat + j] = 0.0;
    __syncthreads();
    // backward
    for (int o = nops - 1; o >= 0; --o) {
      const qs_qnn_op op = ops[o];
      if (op.kind == QS_QNN_CX) {
        if (tid == 0) {
          rm[op.q1] ^= rm[op.q0];
          cm[op.q0] ^= cm[op.q1];
        }
      } else if (op.kind == QS_QNN_CZ) {
        const uint64_t ra = rm[op.q0], rb = rm[op.q1];
        for (uint64_t i = tid; i < N; i += nt)
          if (parity64(i & ra) & parity64(i & rb)) {
            const cx<R> a = ld(psi + i), b = ld(lam + i);
            st(psi + i, mk<R>(-a.re, -a.im));
            st(lam + i, mk<R>(-b.re, -b.im));
          }
      } else {
        const bool enc = op.kind == QS_QNN_ENC_RY;
        const bool train = enc || op.param >= 0;
        const int axis = op_axis(op.kind);
        const double t = enc ? A.x[(uint64_t)c * A.nfeat + op.param] : (op.param >= 0 ? A.theta[op.param] : op.value);
        cx<R> u[4];
        rot_matrix<R>(axis, -t, u);  // the inverse
        const uint64_t m = cm[op.q0], r = rm[op.q0];
        double g = 0;
        for (uint64_t h = tid; h < half; h += nt) {
          uint64_t p0, p1;
          pair_of(h, m, r, p0, p1);
          const cx<R> a0 = ld(psi + p0), a1 = ld(psi + p1), b0 = ld(lam + p0), b1 = ld(lam + p1);
          if (train) {
            // Im <lam|P|psi>: X psi = (a1, a0); Y psi = (-i a1, i a0); Z psi = (a0, -a1)
            if (axis == 0)
              g += ((double)b0.re * a1.im - (double)b0.im * a1.re) + ((double)b1.re * a0.im - (double)b1.im * a0.re);
            else if (axis == 1)
              g += -((double)b0.re * a1.re + (double)b0.im * a1.im) + ((double)b1.re * a0.re + (double)b1.im * a0.im);
            else
              g += ((double)b0.re * a0.im - (double)b0.im * a0.re) - ((double)b1.re * a1.im - (double)b1.im * a1.re);
          }
          st(psi + p0, cfma(mul(u[0], a0), u[1], a1));
          st(psi + p1, cfma(mul(u[2], a0), u[3], a1));
          st(lam + p0, cfma(mul(u[0], b0), u[1], b1));
          st(lam + p1, cfma(mul(u[2], b0), u[3], b1));
        }
        if (train) {
          double dummy = 0;
          block_reduce2<R>(g, dummy, red);
          if (tid == 0) {
            if (enc) A.grad_x[(uint64_t)c * A.nfeat + op.param] += dl * g;
            else grow[op.param] += dl * g;
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      A.expval[c] = e;
      A.loss_terms[c] = (e - tgt) * (e - tgt);
    }
    __syncthreads();
  }
}

template <typename R>
static int qnn_prog_t(qs_ctx* ctx, int ncirc, const int32_t* widths, const int32_t* off, const qs_qnn_op* ops,
                      int nparams, int nfeat, const double* x, const double* target, const double* theta,
                      double* expval, double* loss, double* grad_theta, double* grad_x, const qs_qnn_config* cfg,
                      hipStream_t s) {
  const int lds_w = cfg ? cfg->lds_max_width : 11;
  const int threads = cfg ? cfg->block_threads : 256;
  const int per_cu = cfg ? cfg->blocks_per_cu : 2;
  QS_REQUIRE(threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024,
             "block_threads must be a power of two in [64, 1024]");
  QS_REQUIRE(lds_w >= 0 && lds_w <= 13, "lds_max_width must be in [0, 13]");
  QS_REQUIRE(per_cu >= 1 && per_cu <= 8, "blocks_per_cu must be in [1, 8]");
  const size_t lds_bytes = lds_w > 0 ? (size_t)2 * (1ull << lds_w) * sizeof(cx<R>) : 0;
  QS_REQUIRE(lds_bytes <= 150 * 1024, "LDS state too large for one workgroup");
  int max_width = 1;
  for (int i = 0; i < ncirc; ++i) max_width = std::max(max_width, (int)widths[i]);
  int cus = 0;
  QS_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int blocks = std::max(1, cus * per_cu);
  const int nops = off[ncirc];
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t np = std::max(nparams, 1), nf = std::max(nfeat, 1);
  const uint64_t sz_i = al((uint64_t)ncirc * sizeof(int32_t)), sz_o = al((uint64_t)(ncirc + 1) * sizeof(int32_t)),
                 sz_ops = al((uint64_t)std::max(nops, 1) * sizeof(qs_qnn_op)),
                 sz_x = al((uint64_t)ncirc * nf * sizeof(double)), sz_c = al((uint64_t)ncirc * sizeof(double)),
                 sz_t = al(np * sizeof(double)), sz_gc = al((uint64_t)ncirc * np * sizeof(double));
  const uint64_t slot_bytes = max_width > lds_w ? (uint64_t)blocks * 2 * (1ull << max_width) * sizeof(cx<R>) : 0;
  const uint64_t total = 2 * sz_i + sz_o + sz_ops + 2 * sz_x + 3 * sz_c + 2 * sz_t + 256 + sz_gc + slot_bytes;
  int rc = ensure_scratch(ctx, total);
  if (rc) return rc;
  char* b = (char*)ctx->scratch;
  int32_t* d_order = (int32_t*)b;
  b += sz_i;
  int32_t* d_w = (int32_t*)b;
  b += sz_i;
  int32_t* d_off = (int32_t*)b;
  b += sz_o;
  qs_qnn_op* d_ops = (qs_qnn_op*)b;
  b += sz_ops;
  double* d_x = (double*)b;
  b += sz_x;
  double* d_gx = (double*)b;
  b += sz_x;
  double* d_t = (double*)b;
  b += sz_c;
  double* d_e = (double*)b;
  b += sz_c;
  double* d_l = (double*)b;
  b += sz_c;
  double* d_th = (double*)b;
  b += sz_t;
  double* d_gth = (double*)b;
  b += sz_t;
  int* d_cnt = (int*)b;
  b += 256;
  double* d_gc = (double*)b;
  b += sz_gc;
  void* d_slots = b;
  std::vector<int32_t> order(ncirc);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c2) { return widths[a] > widths[c2]; });
  QS_CHECK_HIP(hipMemcpyAsync(d_order, order.data(), ncirc * sizeof(int32_t), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_w, widths, ncirc * sizeof(int32_t), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_off, off, (ncirc + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (nops) QS_CHECK_HIP(hipMemcpyAsync(d_ops, ops, nops * sizeof(qs_qnn_op), hipMemcpyHostToDevice, s));
  if (nfeat) QS_CHECK_HIP(hipMemcpyAsync(d_x, x, (uint64_t)ncirc * nfeat * sizeof(double), hipMemcpyHostToDevice, s));
  QS_CHECK_HIP(hipMemcpyAsync(d_t, target, ncirc * sizeof(double), hipMemcpyHostToDevice, s));
  if (nparams) QS_CHECK_HIP(hipMemcpyAsync(d_th, theta, nparams * sizeof(double), hipMemcpyHostToDevice, s));
  QnnProgArgs A;
  A.ncirc = ncirc;
  A.nfeat = nfeat;
  A.nparams = nparams;
  A.lds_max_width = lds_w;
  A.max_width = max_width;
  A.order = d_order;
  A.widths = d_w;
  A.off = d_off;
  A.ops = d_ops;
  A.x = d_x;
  A.target = d_t;
  A.theta = d_th;
  A.expval = d_e;
  A.loss_terms = d_l;
  A.grad_circ = d_gc;
  A.grad_x = d_gx;
  A.slots = d_slots;
  A.counter = d_cnt;
  A.inv_ncirc = 1.0 / ncirc;
  if (lds_bytes > 64 * 1024)
    QS_CHECK_HIP(hipFuncSetAttribute((const void*)k_qnn_prog<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_bytes));
  rc = prof_begin(ctx, PROF_QNN, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_qnn_clear, dim3(1), dim3(64), 0, s, d_cnt);
  QS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_qnn_prog<R>, dim3(blocks), dim3(threads), lds_bytes, s, A);
  QS_CHECK_HIP(hipGetLastError());
  if (nparams) {
    hipLaunchKernelGGL(k_qnn_reduce, dim3((unsigned)((nparams + 255) / 256)), dim3(256), 0, s, d_gc, ncirc, nparams,
                       d_gth);
    QS_CHECK_HIP(hipGetLastError());
  }
  rc = prof_end(ctx, s);
  if (rc) return rc;
  std::vector<double> lt(ncirc);
  QS_CHECK_HIP(hipMemcpyAsync(expval, d_e, ncirc * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipMemcpyAsync(lt.data(), d_l, ncirc * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nparams) QS_CHECK_HIP(hipMemcpyAsync(grad_theta, d_gth, nparams * sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad_x && nfeat)
    QS_CHECK_HIP(hipMemcpyAsync(grad_x, d_gx, (uint64_t)ncirc * nfeat * sizeof(double), hipMemcpyDeviceToHost, s));
  QS_CHECK_HIP(hipStreamSynchronize(s));
  double L = 0;
  for (int i = 0; i < ncirc; ++i) L += lt[i];
  *loss = L / ncirc;
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" int qs_qnn_batch(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths, int nfeat, const double* x,
                            const double* target, int layers, int max_width, const double* theta, double* expval,
                            double* loss, double* grad_theta, double* grad_x, const qs_qnn_config* cfg,
                            void* stream) {
  QS_REQUIRE(ctx && widths && x && target && theta && expval && loss && grad_theta, "null argument");
  QS_REQUIRE(dtype == QS_C64 || dtype == QS_C128, "dtype must be QS_C64 or QS_C128");
  QS_REQUIRE(ncirc >= 1 && nfeat >= 0 && layers >= 1 && layers <= QNN_MAX_LAYERS, "bad batch shape");
  QS_REQUIRE(max_width >= 1 && max_width <= QNN_MAX_W, "max_width must be in [1, 20]");
  for (int i = 0; i < ncirc; ++i) QS_REQUIRE(widths[i] >= 2 && widths[i] <= max_width, "width out of range");
  if (dtype == QS_C128)
    return qnn_t<double>(ctx, dtype, ncirc, widths, nfeat, x, target, layers, max_width, theta, expval, loss,
                         grad_theta, grad_x, cfg, as_stream(stream));
  return qnn_t<float>(ctx, dtype, ncirc, widths, nfeat, x, target, layers, max_width, theta, expval, loss, grad_theta,
                      grad_x, cfg, as_stream(stream));
}

extern "C" int qs_qnn_batch_programs(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths,
                                     const int32_t* prog_offsets, const qs_qnn_op* ops, int nparams, int nfeat,
                                     const double* x, const double* target, const double* theta, double* expval,
                                     double* loss, double* grad_theta, double* grad_x, const qs_qnn_config* cfg,
                                     void* stream) {
  QS_REQUIRE(ctx && widths && prog_offsets && target && expval && loss, "null argument");
  QS_REQUIRE(dtype == QS_C64 || dtype == QS_C128, "dtype must be QS_C64 or QS_C128");
  QS_REQUIRE(ncirc >= 1 && nparams >= 0 && nfeat >= 0, "bad batch shape");
  QS_REQUIRE(nparams == 0 || (theta && grad_theta), "theta / grad_theta missing");
  QS_REQUIRE(nfeat == 0 || x, "x missing");
  QS_REQUIRE(prog_offsets[0] == 0, "prog_offsets[0] must be 0");
  for (int c = 0; c < ncirc; ++c) {
    const int w = widths[c];
    QS_REQUIRE(w >= 1 && w <= QNN_MAX_W, "width out of range [1, 20]");
    QS_REQUIRE(prog_offsets[c + 1] >= prog_offsets[c], "prog_offsets must not decrease");
    QS_REQUIRE(prog_offsets[c + 1] == prog_offsets[c] || ops, "ops missing");
    for (int o = prog_offsets[c]; o < prog_offsets[c + 1]; ++o) {
      const qs_qnn_op& op = ops[o];
      QS_REQUIRE(op.kind >= QS_QNN_RX && op.kind <= QS_QNN_ENC_RY, "unknown op kind");
      QS_REQUIRE(op.q0 >= 0 && op.q0 < w, "op qubit out of range");
      if (op.kind == QS_QNN_CX || op.kind == QS_QNN_CZ)
        QS_REQUIRE(op.q1 >= 0 && op.q1 < w && op.q1 != op.q0, "two-qubit op needs two distinct qubits");
      if (op.kind == QS_QNN_ENC_RY) QS_REQUIRE(op.param >= 0 && op.param < nfeat, "encoding feature out of range");
      else if (op.kind <= QS_QNN_RZ) QS_REQUIRE(op.param < nparams, "parameter index out of range");
    }
  }
  if (dtype == QS_C128)
    return qnn_prog_t<double>(ctx, ncirc, widths, prog_offsets, ops, nparams, nfeat, x, target, theta, expval, loss,
                              grad_theta, grad_x, cfg, as_stream(stream));
  return qnn_prog_t<float>(ctx, ncirc, widths, prog_offsets, ops, nparams, nfeat, x, target, theta, expval, loss,
                           grad_theta, grad_x, cfg, as_stream(stream));
}
