/*
 * qsim.h — C ABI of libqsim_hip.so, the MI355X (gfx950) circuit-simulation
 * library behind quict_amd.
 *
 * The reference's drop-in boundary for this path is the function table that
 * QuICT binds through LinAlgLoader (QuICT/ops/utils/algorithm_loader.py:27-60):
 * the per-gate kernels exported by QuICT/ops/gate_kernel/gate_function.py:6-28
 * plus QuICT/ops/linalg/gpu_calculator.py (dot, matrix_dot_vector, ...), each
 * called in place on a device vector with reversed qubit indices (n-1-q).  The
 * entry points below replace that table; each one names the reference
 * function(s) it stands in for.  A maintainer binds them from Python with
 * ctypes (see quict_amd/_native.py and INTEGRATION.md).
 *
 * Conventions
 *  - Every function returns an int status: QS_OK (0) or a negative QS_ERR_*.
 *    qs_last_error() returns a message for the most recent failure on the
 *    calling thread.
 *  - Amplitude/density buffers are caller-owned device memory (for example a
 *    torch tensor's data_ptr()).  The library never frees caller memory.
 *  - "bits" are amplitude-index bit positions (0 = least significant).  QuICT
 *    qubit q of an n-qubit register is bit n-1-q.
 *  - stream is a hipStream_t (NULL = the null stream).  Calls that return host
 *    values synchronise that stream.  Gate calls (qs_sv_apply_gate(s),
 *    qs_graph_launch) return once their kernels are queued, except a dense
 *    gate on 4 or more targets run as its own pass (its matrix upload waits
 *    for the stream; 3-target gates carry theirs in the kernel arguments);
 *    the fused program of a qs_sv_apply_gates call stays in
 *    one of two per-context slots until its passes finish, so the call after
 *    next on the same context may wait for them.  The adjoint, QNN and
 *    circuit-matrix calls return host values and synchronise.
 *  - A context is bound to one device and is not thread-safe.
 */
#ifndef QSIM_H
#define QSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QS_OK 0
#define QS_ERR_ARG (-1)         /* invalid argument: bad bit index, kind, size, null pointer */
#define QS_ERR_DEVICE (-2)      /* HIP runtime failure: launch error, out of memory */
#define QS_ERR_UNSUPPORTED (-3) /* valid request this build does not implement */

#define QS_C64 0  /* complex<float>  amplitudes, interleaved (re, im) */
#define QS_C128 1 /* complex<double> amplitudes, interleaved (re, im) */

#define QS_MAX_GATE_BITS 16
#define QS_MAX_DENSE_TARGETS 10

/* Gate kinds.  Matrix payload m[] holds complex numbers as (re, im) doubles. */
enum qs_gate_kind {
  QS_GATE_DENSE = 0, /* 2^t x 2^t dense on the targets (row-major), applied where all
                        controls are 1.  t <= 2: m[]; t >= 3: big (host pointer). */
  QS_GATE_DIAG = 1,  /* diagonal, 2^t entries in m[] (t <= 4), under controls */
  QS_GATE_PHASE = 2, /* multiply the amplitude whose controls and targets are all 1 by m[0] */
  QS_GATE_XPERM = 3, /* 1 target: a0' = m[0]*a1, a1' = m[1]*a0, under controls (X, Y, CX, CY, CCX) */
  QS_GATE_SWAP = 4   /* 2 targets: a01' = m[0]*a10, a10' = m[1]*a01, under controls (Swap, iSwap, CSwap) */
};

typedef struct qs_gate {
  int32_t kind;
  int32_t ncontrols;
  int32_t ntargets;
  int32_t bits[QS_MAX_GATE_BITS]; /* controls first, then targets; target 0 = MSB of the matrix index */
  double m[32];                   /* kind-specific payload (up to 16 complex) */
  const double* big;              /* DENSE, ntargets >= 3: host 2^t x 2^t complex, row-major */
} qs_gate;

typedef struct qs_ctx qs_ctx;

/* ---- context ------------------------------------------------------------ */
/* Create a context on HIP device `device`.  Returns NULL on failure. */
qs_ctx* qs_ctx_create(int device);
void qs_ctx_destroy(qs_ctx* ctx);
const char* qs_last_error(void);
/* Bytes of device scratch the context currently holds (measurement pyramids,
   fusion programs, reduction partials). */
uint64_t qs_ctx_scratch_bytes(const qs_ctx* ctx);

/* Kernel timing for measurement: when enabled, every tile-pass / per-gate /
   gradient / batch launch is bracketed by HIP events on its own stream.
   qs_profile_read returns, per class [fused tile pass, per-gate pass,
   adjoint, qnn batch], the summed kernel milliseconds and launch counts. */
int qs_profile_enable(qs_ctx* ctx, int on);
int qs_profile_read(qs_ctx* ctx, double* ms /*[4]*/, int64_t* launches /*[4]*/, int reset);

/* Streaming copy of `bytes` (a multiple of 16) from src to dst, 16 B per lane,
   optionally with non-temporal loads/stores: the achievable-bandwidth
   reference that roofline figures are compared with (no reference
   counterpart). */
int qs_stream_copy(qs_ctx* ctx, void* dst, const void* src, uint64_t bytes, int nontemporal, void* stream);

/* ---- state vector ------------------------------------------------------- */
/* |index> : replaces StateVectorSimulator.initial_state_vector
   (QuICT/simulation/state_vector/statevector_simulator.py:122-129). */
int qs_sv_init(qs_ctx* ctx, void* vec, int n, int dtype, uint64_t index, void* stream);

/* One gate, one full-state pass (no fusion).  Replaces the per-gate table of
   QuICT/ops/gate_kernel/gate_function.py:1034-1590 (diagonal_targ, normal_targ,
   control_targ, swap_targ, reverse_targ, *_targs, *_ctargs, reverse_more,
   diagonal_more, swap_tmore) and, for >= 3 targets,
   QuICT/ops/linalg/gpu_calculator.py:430-478 (matrix_dot_vector). */
int qs_sv_apply_gate(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gate, void* stream);

/* A gate sequence.  fuse=0: one pass per gate (same kernels as qs_sv_apply_gate).
   fuse=1: the fusion planner groups gates into LDS/register-resident tiles and
   runs each group in one read+write pass.  fuse=2: as 1, and the planner may
   also move qubits onto the low index bits between passes (every qubit is back
   on its own bit when the call returns) when that saves at least 8% of the
   passes, for lists of more than 480 gates; faster on long state-vector
   circuits, not on density matrices.  Plans without moves (fuse=1, and fuse=2
   lists of up to 480 gates) are launched pass by pass while they are planned.
   Replaces the per-gate loop of StateVectorSimulator.run
   (statevector_simulator.py:155-164, 202-315).
   *passes (may be NULL) receives the number of full-state passes executed. */
int qs_sv_apply_gates(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates,
                      int fuse, int* passes, void* stream);

/* Fusion tile shape: log2 amplitudes per tile (0 = built-in default 12; also
   11 or 13 for complex128, 13 for complex64) and the number of low bits always
   kept contiguous (0 = 4, or 2..8, at most tile_log2 - 4; checked again when
   a fused call runs, since the two may be set in either order). */
int qs_set_fusion(qs_ctx* ctx, int tile_log2, int low_bits);

/* Launch shape of the fused tile pass: prefetch != 0 selects persistent
   workgroups (wg_per_cu per CU) that load the next tile while computing the
   current one; prefetch == 0 launches one workgroup per tile. */
int qs_set_fusion_launch(qs_ctx* ctx, int prefetch, int wg_per_cu);

/* Captured gate sequences (HIP graphs).  qs_graph_capture plans the sequence
   once (the same fused passes as qs_sv_apply_gates with fuse=2), uploads the
   program into memory the graph owns and captures the passes on `vec` in a HIP
   graph; init_zero != 0 puts a |0...0> initialisation in front.  Every
   qs_graph_launch then replays them on `stream` with no host-side planning:
   repeated runs of one circuit (sampling loops, serving) cost only the GPU
   passes.  The graph is bound to `vec` (same buffer, size and dtype) until
   qs_graph_destroy.  Sequences with a gate the tile kernel cannot take (dense
   on more than 4 targets, diagonal or phase on more than 6 bits, more than
   tile_log2 bits, or more than tile_log2 - low_bits bits above the low bits)
   and registers smaller than 2^(tile+1) amplitudes are not captured
   (QS_ERR_UNSUPPORTED).  Replaces the per-run gate loop of
   StateVectorSimulator.run (statevector_simulator.py:131-166) for a fixed
   circuit; the reference has no captured form. */
typedef struct qs_graph qs_graph;
int qs_graph_capture(qs_ctx* ctx, void* vec, int n, int dtype, const qs_gate* gates, int ngates, int init_zero,
                     qs_graph** out);
int qs_graph_launch(qs_graph* g, void* stream);
/* Full-state passes one launch executes (-1 for NULL). */
int qs_graph_passes(const qs_graph* g);
void qs_graph_destroy(qs_graph* g);

/* P(bit == 0) = sum |a|^2 over the half with that bit clear, reduced in double.
   Replaces measured_prob_calculate (gate_function.py:1751-1776). */
int qs_sv_prob_zero(qs_ctx* ctx, const void* vec, int n, int dtype, int bit, double* p0, void* stream);

/* Keep the `outcome` half scaled by `scale`, zero the other half.  Replaces the
   MeasureGate0/1 kernels of apply_measuregate (gate_function.py:1779-1815). */
int qs_sv_collapse(qs_ctx* ctx, void* vec, int n, int dtype, int bit, int outcome, double scale, void* stream);

/* Reset `bit` to 0 given p0 = P(bit == 0) (reset_gate_apply semantics,
   QuICT/ops/linalg/cpu_calculator.py:366-384). */
int qs_sv_reset(qs_ctx* ctx, void* vec, int n, int dtype, int bit, double p0, void* stream);

/* Sample `shots` full-register outcomes.  uniforms: host (shots x n) doubles,
   row s holds the draws for qubits 0..n-1 (bit n-1 first).  Outcome rule per
   qubit: 1 iff u > P(0 | outcomes so far), the rule of
   StateVectorSimulator.sample (statevector_simulator.py:595-626).  out: host
   uint64 amplitude indices.  Builds a pair-sum pyramid of |a|^2 in double. */
int qs_sv_sample(qs_ctx* ctx, const void* vec, int n, int dtype, const double* uniforms, int shots,
                 uint64_t* out, void* stream);

/* No counterpart in the reference: qs_sv_sample for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same arguments, rule and outputs as
   qs_sv_sample; its pyramid, uniforms and outcomes live in a fixed 40 MiB
   device array per GPU (resolved once per device when the context is
   created, held until the process ends) instead of the context's grow-on-
   demand scratch, and it calls no hipMalloc / hipFree.  Limit: (2^n - 1) * 8
   + 2048 + shots * (n + 1) * 8 bytes (rounded up to 256) must fit in 40 MiB,
   i.e. at most 22 qubits; larger requests return QS_ERR_ARG.  Every call on
   one GPU shares that GPU's array: calls that overlap on one GPU (several host
   threads, or streams of other contexts of the same device) race. */
int qs_sv_sample_noalloc(qs_ctx* ctx, const void* vec, int n, int dtype, const double* uniforms, int shots,
                         uint64_t* out, void* stream);

/* Same sampling rule over a caller device buffer of 2^n probabilities (used
   for the diagonal of a density matrix, DensityMatrixSimulator.sample,
   QuICT/simulation/density_matrix/density_matrix_simulator.py:210-240). */
int qs_sample_probs(qs_ctx* ctx, const double* probs_dev, int n, const double* uniforms, int shots, uint64_t* out,
                    void* stream);

/* |a_i|^2 into a caller device buffer of 2^n doubles. */
int qs_sv_probabilities(qs_ctx* ctx, const void* vec, int n, int dtype, double* probs_dev, void* stream);

/* sum |a|^2 (host result). */
int qs_sv_norm2(qs_ctx* ctx, const void* vec, int n, int dtype, double* out, void* stream);

/* Qubit permutation, out of place: out[j] = vec[src(j)] with
   bit perm[b] of src taken from bit b of j.  Replaces VectorPermutation
   (QuICT/ops/linalg/gpu_calculator.py:170-250). */
int qs_sv_permute(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const int32_t* perm, void* stream);

/* ---- density matrix (vec(rho) as a 2n-bit state: row bits n..2n-1) ------- */
/* diag(rho) into a caller device buffer of 2^n doubles (real parts). */
int qs_dm_diagonal(qs_ctx* ctx, const void* rho, int n, int dtype, double* diag_dev, void* stream);
/* trace (host). */
int qs_dm_trace(qs_ctx* ctx, const void* rho, int n, int dtype, double* re, double* im, void* stream);
/* P(qubit bit == 0) from the diagonal (host).  Replaces the projector trace of
   DensityMatrixSimulator.apply_measure (density_matrix_simulator.py:182-208). */
int qs_dm_prob_zero(qs_ctx* ctx, const void* rho, int n, int dtype, int bit, double* p0, void* stream);
/* rho <- P rho P / p: zero rows/cols whose `bit` != outcome, scale the rest. */
int qs_dm_collapse(qs_ctx* ctx, void* rho, int n, int dtype, int bit, int outcome, double scale, void* stream);

/* ---- circuit matrices (2^n x 2^n complex, n <= 20) ------------------------ */
/* The circuit unitary is built column-major: start from the identity and apply
   the circuit's gates with qs_sv_apply_gates on the 2n-bit vector, bits 0..n-1
   (qubit q = bit n-1-q, as for a state vector).  The buffer then holds
   W = U^T (row c = U|c>).  Together these replace
   CircuitMatrix.get_unitary_matrix (QuICT/core/utils/circuit_matrix.py:69-125,
   reached from Circuit.matrix, QuICT/core/utils/circuit_info.py:177-200) and
   its MatrixTensorI / MatrixPermutation / tensor / dot expansions
   (QuICT/ops/linalg/gpu_calculator.py:13-330). */
/* m <- identity (the empty circuit's matrix, circuit_info.py:187-195). */
int qs_mat_identity(qs_ctx* ctx, void* m, int n, int dtype, void* stream);
/* In-place transpose of a square 2^n x 2^n matrix (turns W into U). */
int qs_mat_transpose(qs_ctx* ctx, void* m, int n, int dtype, void* stream);
/* y <- M x (trans = 0, M row-major) or y <- M^T x (trans = 1, i.e. U x when M
   holds W).  x and y are distinct 2^n-amplitude device vectors.  Replaces
   dot(U, vec) of UnitarySimulator.run (QuICT/simulation/unitary/
   unitary_simulator.py:100-104, gpu_calculator.py:13-38).  Accumulates in
   double for both dtypes; deterministic (fixed summation order). */
int qs_mat_vec(qs_ctx* ctx, const void* m, int n, int dtype, int trans, const void* x, void* y, void* stream);

/* ---- expectation values and gradients (no reference counterpart) --------- */
/* Observable sum_k coef_k * P_k with P_k a Pauli string given as two bit masks
   (x-mask: bits carrying X or Y, z-mask: bits carrying Z or Y). */
typedef struct qs_pauli_term {
  double coef;
  uint64_t xmask;
  uint64_t zmask;
} qs_pauli_term;

/* out <- H vec (caller device buffer, same size/dtype as vec). */
int qs_sv_apply_pauli_sum(qs_ctx* ctx, const void* vec, void* out, int n, int dtype, const qs_pauli_term* terms,
                          int nterms, void* stream);
/* <a|b> (host, complex). */
int qs_sv_inner(qs_ctx* ctx, const void* a, const void* b, int n, int dtype, double* re, double* im, void* stream);

/* Adjoint-method gradient of E = <psi|H|psi> for psi = G_{L-1}...G_0 |psi_in>.
   On entry psi holds the forward state, lam holds H psi.  For gate j with
   param_slot[j] >= 0, gen[j] is its generator as a (re, im) 2^t x 2^t matrix D_j
   with dG_j/dtheta = D_j (stored in gen_mats at offset gen_off[j]).  The
   routine walks j = L-1..0 applying G_j^dagger to psi and lam and
   accumulating grads[param_slot[j]] += 2 Re <lam| D_j |psi_before_j>.
   No per-layer states are stored: psi is recomputed by inverse gates.
   psi and lam are overwritten. grads: host, nslots doubles (accumulated). */
int qs_adjoint_gradient(qs_ctx* ctx, void* psi, void* lam, int n, int dtype, const qs_gate* gates, int ngates,
                        const qs_gate* dgates, const int32_t* param_slot, double* grads, int nslots, void* stream);

/* The same gradient for H given as a Pauli sum, with the energy: lam is
   computed here as H psi (lam's prior contents are ignored) and *energy =
   <psi|H|psi> (host) comes back with the gradient; one host wait at the end.
   fwd_gates NULL: psi holds the forward state on entry (the sweep is planned
   before any work is enqueued, so the planning overlaps forward passes still
   running on the stream).  fwd_gates given (ngates records, fwd_gates[j] the
   gate whose inverse is gates[j]): psi is set to |0...0> and the forward
   state is computed here from the sweep's own plan: its LDS tile passes run
   last first with the conjugate transposes of the gates they hold (gates
   outside the passes run from fwd_gates): one plan serves both directions.
   No counterpart in the reference, whose gradient code lives in the
   external QuICT_ml. */
int qs_adjoint_gradient_pauli(qs_ctx* ctx, void* psi, void* lam, int n, int dtype, const qs_pauli_term* terms,
                              int nterms, const qs_gate* fwd_gates, const qs_gate* gates, int ngates,
                              const qs_gate* dgates, const int32_t* param_slot, double* grads, int nslots,
                              double* energy, void* stream);

/* ---- batched variational circuits (no reference counterpart) ------------- */
/* Launch shape of the batched QNN kernel (picked by the autotuner):
   widths <= lds_max_width keep psi and lam in LDS (2 x 2^w amplitudes),
   wider circuits use a global slot per workgroup; block_threads threads per
   workgroup; blocks_per_cu persistent workgroups per CU; use_graph != 0
   captures the launch in a hipGraph on first use and replays it for every
   later batch of the same shape. */
typedef struct qs_qnn_config {
  int32_t lds_max_width;
  int32_t block_threads;
  int32_t blocks_per_cu;
  int32_t use_graph;
  int32_t fuse_qubits; /* rotations of this many qubits (1..3) per pass over a circuit's state; 0 means 1 */
} qs_qnn_config;

/* Hardware-efficient-ansatz QNN batch, one persistent launch for forward and
   adjoint backward of every circuit.  Circuit c has width widths[c] (<= 20),
   angle-encodes x[c*nfeat + j] with Ry on qubit j mod width, then `layers`
   layers of (Ry(theta), Rz(theta) on every qubit; CX(i, i+1) ladder).  theta is
   shared: theta[(l*2 + g)*max_width + q].  Loss: mean over circuits of
   (<Z_0> - target[c])^2.  Outputs (host): expval[c], loss, grad_theta
   (d loss / d theta), grad_x (per circuit, nfeat; may be NULL).  dtype selects
   complex64 / complex128 amplitudes (reductions in double).  cfg may be NULL.
   grad_theta is the sum of per-circuit terms taken in circuit order, so it is
   bit-identical across runs and launch shapes of the same block_threads. */
int qs_qnn_batch(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths, int nfeat, const double* x,
                 const double* target, int layers, int max_width, const double* theta, double* expval,
                 double* loss, double* grad_theta, double* grad_x, const qs_qnn_config* cfg, void* stream);

/* One op of a per-circuit QNN program (no reference counterpart: QuICT's QNN
   code lives in the external QuICT_ml package).  kind: QS_QNN_RX / _RY / _RZ,
   exp(-i t P / 2) on qubit q0 with t = theta[param] (trainable) or, when
   param < 0, t = value; QS_QNN_ENC_RY, Ry(x[c*nfeat + param]) on q0 (an input
   encoding: its gradient goes to grad_x); QS_QNN_CX, control q0, target q1;
   QS_QNN_CZ on q0, q1.  Qubit 0 is the most significant amplitude bit. */
#define QS_QNN_RX 0
#define QS_QNN_RY 1
#define QS_QNN_RZ 2
#define QS_QNN_CX 3
#define QS_QNN_CZ 4
#define QS_QNN_ENC_RY 5
typedef struct qs_qnn_op {
  int32_t kind;
  int32_t q0;
  int32_t q1;
  int32_t param;
  double value;
} qs_qnn_op;

/* Batch of circuits with their own gate programs, forward and adjoint backward
   in one persistent launch (no reference counterpart).  Circuit c has width
   widths[c] (1..20) and runs ops[prog_offsets[c] .. prog_offsets[c+1]) from
   |0..0>; E_c = <Z_0>, loss = mean over circuits of (E_c - target[c])^2.
   theta (nparams values) is shared by every circuit.  Outputs (host):
   expval[c], loss, grad_theta (d loss / d theta, summed over circuits in
   circuit order: bit-identical across runs), grad_x (ncirc x nfeat, d loss /
   d x for the encoding ops; may be NULL).  cfg: lds_max_width, block_threads,
   blocks_per_cu (use_graph and fuse_qubits are ignored: one rotation per
   pass); NULL for defaults.  Replaces no reference function. */
int qs_qnn_batch_programs(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths, const int32_t* prog_offsets,
                          const qs_qnn_op* ops, int nparams, int nfeat, const double* x, const double* target,
                          const double* theta, double* expval, double* loss, double* grad_theta, double* grad_x,
                          const qs_qnn_config* cfg, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QSIM_H */
