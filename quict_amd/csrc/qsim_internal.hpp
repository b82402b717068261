// Internal (non-ABI) declarations shared by the libqsim_hip translation units.
#pragma once

#include "qsim_common.hpp"

namespace qs {

struct SortedBits {
  int n;
  int b[QS_MAX_GATE_BITS];
};

// Kernel-argument form of one gate (built on the host from qs_gate).
template <typename R> struct GateK {
  uint64_t cmask;                 // OR of control bits
  int nsorted;                    // controls + targets
  int sorted[QS_MAX_GATE_BITS];   // all gate bits, ascending
  uint64_t off[16];               // offset of target combination m (target 0 = MSB of m)
  cx<R> m[16];                    // payload converted to R
};

int validate_gate(const qs_gate* g, int n);
int apply_gate(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* g, hipStream_t s);

// Fused multi-gate passes (qsim_fused.hip).
int apply_gates_fused(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, bool moves,
                      int* passes, hipStream_t s);

// Double-precision reductions (qsim_measure.hip).
int reduce_prob_zero(qs_ctx* ctx, const void* vec, int n, int dtype, int bit, double* out, hipStream_t s);

// Dense gates on 4..6 targets on the matrix cores (qsim_dense.hip).
bool dense_mfma_eligible(const qs_gate* g, int n);
int launch_dense_mfma(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* g, hipStream_t s);

// Resolve this device's copy of the fixed sampling scratch (qsim_measure.hip);
// called once per context, with the context's device current.
int noalloc_scratch_init(qs_ctx* ctx);

// out = H vec for a Pauli sum (qsim_api.hip); enqueued on s, no host wait.
int pauli_sum_async(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const qs_pauli_term* terms,
                    int nterms, hipStream_t s);
// <a|b> into res[0..1] (device), using part (device, 2 * inner_partial_blocks()
// doubles); enqueued on s, no host wait (qsim_api.hip).
int inner_async(const void* a, const void* b, int n, int dtype, double* part, double* res, hipStream_t s);
int inner_partial_blocks();

inline size_t amp_bytes(int dtype) { return dtype == QS_C128 ? 16 : 8; }

}  // namespace qs
