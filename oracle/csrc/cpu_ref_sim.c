/*
 * C restatement of the reference's CPU state-vector kernels — the timed CPU
 * baseline.  TEST / MEASUREMENT INFRASTRUCTURE ONLY (never linked into the
 * product library).
 *
 * Follows QuICT/ops/linalg/cpu_calculator.py with the same per-gate structure:
 *   matrix_dot_vector (:139-186): precompute the 2^k offsets of a group, walk
 *     the 2^(n-k) groups (opening a zero bit at each sorted argument), gather
 *     the group, multiply by the dense matrix, scatter;
 *   diagonal_matrix (:189-235): multiply the touched amplitudes (the last 2^t
 *     diagonal entries, or only the all-ones entry for control-type gates);
 *   swap_matrix (:238-280) and reverse_matrix (:283-326): exchange a pair with
 *     the anti-diagonal values;
 *   measure_gate_apply (:329-363): P0 over the bit-0 half, collapse.
 * One full-state pass per gate, as in the reference.  The gate list uses the
 * qs_gate record of include/qsim.h (kind, controls, targets, bit positions,
 * payload), so it runs the exact workload the GPU path runs.
 *
 * The reference's loops are numba @njit without parallel=True (serial);
 * `threads` > 1 splits the group loop with OpenMP for an all-core figure.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/qsim.h"

typedef double _Complex zd;

static uint64_t open_zeros(uint64_t i, const int* sorted, int k) {
  for (int j = 0; j < k; ++j) {
    const int b = sorted[j];
    i = ((i >> b) << (b + 1)) | (i & ((1ull << b) - 1));
  }
  return i;
}

static void sort_bits(const int* in, int k, int* out) {
  memcpy(out, in, sizeof(int) * k);
  for (int a = 1; a < k; ++a)
    for (int b = a; b > 0 && out[b - 1] > out[b]; --b) {
      int t = out[b];
      out[b] = out[b - 1];
      out[b - 1] = t;
    }
}

static zd cz(const double* m, int i) { return m[2 * i] + I * m[2 * i + 1]; }

/* gate-matrix index bit (target j, MSB first) -> amplitude offset */
static uint64_t target_offset(const qs_gate* g, int mi) {
  const int t = g->ntargets, c = g->ncontrols;
  uint64_t o = 0;
  for (int j = 0; j < t; ++j)
    if ((mi >> (t - 1 - j)) & 1) o |= 1ull << g->bits[c + j];
  return o;
}

static void apply_one(zd* vec, int n, const qs_gate* g, int threads) {
  const int c = g->ncontrols, t = g->ntargets, k = c + t;
  int sorted[QS_MAX_GATE_BITS];
  sort_bits(g->bits, k, sorted);
  uint64_t cmask = 0;
  for (int i = 0; i < c; ++i) cmask |= 1ull << g->bits[i];
  const uint64_t repeat = 1ull << (n - k);
  const int D = 1 << t;
  uint64_t offs[1 << QS_MAX_DENSE_TARGETS];
  for (int m = 0; m < D; ++m) offs[m] = target_offset(g, m);
  (void)threads;
  switch (g->kind) {
    case QS_GATE_DENSE: {
      const double* src = (t <= 2) ? g->m : g->big;
      zd* mat = (zd*)malloc(sizeof(zd) * D * D);
      for (int i = 0; i < D * D; ++i) mat[i] = cz(src, i);
#pragma omp parallel for num_threads(threads) schedule(static)
      for (uint64_t r = 0; r < repeat; ++r) {
        zd buf[1 << QS_MAX_DENSE_TARGETS], out[1 << QS_MAX_DENSE_TARGETS];
        const uint64_t base = open_zeros(r, sorted, k) | cmask;
        for (int m = 0; m < D; ++m) buf[m] = vec[base | offs[m]];
        for (int i = 0; i < D; ++i) {
          zd acc = 0;
          for (int j = 0; j < D; ++j) acc += mat[i * D + j] * buf[j];
          out[i] = acc;
        }
        for (int m = 0; m < D; ++m) vec[base | offs[m]] = out[m];
      }
      free(mat);
      break;
    }
    case QS_GATE_DIAG: {
      zd d[16];
      for (int m = 0; m < D; ++m) d[m] = cz(g->m, m);
#pragma omp parallel for num_threads(threads) schedule(static)
      for (uint64_t r = 0; r < repeat; ++r) {
        const uint64_t base = open_zeros(r, sorted, k) | cmask;
        for (int m = 0; m < D; ++m) vec[base | offs[m]] *= d[m];
      }
      break;
    }
    case QS_GATE_PHASE: {
      const zd v = cz(g->m, 0);
      const uint64_t o = offs[D - 1];
#pragma omp parallel for num_threads(threads) schedule(static)
      for (uint64_t r = 0; r < repeat; ++r) vec[(open_zeros(r, sorted, k) | cmask) | o] *= v;
      break;
    }
    case QS_GATE_XPERM: {
      const zd a = cz(g->m, 0), b = cz(g->m, 1);
#pragma omp parallel for num_threads(threads) schedule(static)
      for (uint64_t r = 0; r < repeat; ++r) {
        const uint64_t i0 = open_zeros(r, sorted, k) | cmask, i1 = i0 | offs[1];
        const zd x0 = vec[i0];
        vec[i0] = vec[i1] * a;
        vec[i1] = x0 * b;
      }
      break;
    }
    case QS_GATE_SWAP: {
      const zd a = cz(g->m, 0), b = cz(g->m, 1);
#pragma omp parallel for num_threads(threads) schedule(static)
      for (uint64_t r = 0; r < repeat; ++r) {
        const uint64_t base = open_zeros(r, sorted, k) | cmask;
        const uint64_t i01 = base | offs[1], i10 = base | offs[2];
        const zd x = vec[i01];
        vec[i01] = vec[i10] * a;
        vec[i10] = x * b;
      }
      break;
    }
  }
}

/* Apply `ngates` records in order; returns 0. */
int cpu_apply_gates(double* vec, int n, const qs_gate* gates, int ngates, int threads) {
  for (int i = 0; i < ngates; ++i) apply_one((zd*)vec, n, &gates[i], threads);
  return 0;
}

/* P(bit == 0) */
double cpu_prob_zero(const double* vec, int n, int bit) {
  const zd* v = (const zd*)vec;
  double s = 0;
  const uint64_t half = 1ull << (n - 1);
  for (uint64_t h = 0; h < half; ++h) {
    const uint64_t i = ((h >> bit) << (bit + 1)) | (h & ((1ull << bit) - 1));
    s += creal(v[i]) * creal(v[i]) + cimag(v[i]) * cimag(v[i]);
  }
  return s;
}

/* Sequential measure-and-collapse of every qubit per shot, state restored per
   shot (statevector_simulator.py:595-626); uniforms[s*n + q]. */
int cpu_sample(const double* vec, int n, const double* uniforms, int shots, uint64_t* out) {
  const uint64_t N = 1ull << n;
  zd* w = (zd*)malloc(sizeof(zd) * N);
  for (int s = 0; s < shots; ++s) {
    memcpy(w, vec, sizeof(zd) * N);
    uint64_t value = 0;
    for (int q = 0; q < n; ++q) {
      const int bit = n - 1 - q;
      const double p0 = cpu_prob_zero((const double*)w, n, bit);
      const int one = uniforms[(uint64_t)s * n + q] > p0;
      const double alpha = one ? 1.0 / sqrt(1 - p0) : 1.0 / sqrt(p0);
      for (uint64_t i = 0; i < N; ++i) {
        const int b = (i >> bit) & 1;
        w[i] = (b == one) ? w[i] * alpha : 0;
      }
      value = (value << 1) | (uint64_t)one;
    }
    out[s] = value;
  }
  free(w);
  return 0;
}
