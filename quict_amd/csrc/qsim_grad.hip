// Adjoint-method gradients and the batched QNN launch (no reference counterpart).
#include "qsim_common.hpp"
#include "qsim_internal.hpp"

using namespace qs;

extern "C" {

int qs_adjoint_gradient(qs_ctx* ctx, void* psi, void* lam, int n, int dtype, const qs_gate* gates, int ngates,
                        const qs_gate* dgates, const int32_t* param_slot, double* grads, int nslots, void* stream) {
  (void)ctx; (void)psi; (void)lam; (void)n; (void)dtype; (void)gates; (void)ngates; (void)dgates;
  (void)param_slot; (void)grads; (void)nslots; (void)stream;
  set_error("qs_adjoint_gradient: not built yet");
  return QS_ERR_UNSUPPORTED;
}

int qs_qnn_batch(qs_ctx* ctx, int dtype, int ncirc, const int32_t* widths, int nfeat, const double* x,
                 const double* target, int layers, int max_width, const double* theta, double* expval, double* loss,
                 double* grad_theta, double* grad_x, int tile_log2, void* stream) {
  (void)ctx; (void)dtype; (void)ncirc; (void)widths; (void)nfeat; (void)x; (void)target; (void)layers;
  (void)max_width; (void)theta; (void)expval; (void)loss; (void)grad_theta; (void)grad_x; (void)tile_log2;
  (void)stream;
  set_error("qs_qnn_batch: not built yet");
  return QS_ERR_UNSUPPORTED;
}

}  // extern "C"
