// Learned-surrogate form of the persistent propagation kernel with the BF16
// network: exact degree <= nlow part in FP64 plus the MFMA network residual,
// with the exact FP64 field outside the radius band, on sampled checks and
// after a failed check.  The FP8 network's kernels are in prop_sur_fp8.hip.
#include "prop_forms.hpp"

namespace bqp {
BQP_FORM_UNIT(pick_sur_bf16, upload_sur, 2)

prop_fn pick_sur_bf16(int method, int G) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return pick_w<BQP_METHOD_DOPRI5, 2>(G);
    case BQP_METHOD_DOP853: return pick_w<BQP_METHOD_DOP853, 2>(G);
    case BQP_METHOD_RK87: return pick_w<BQP_METHOD_RK87, 2>(G);
    default: return nullptr;
  }
}

prop_fn pick_sur(int method, int prec, int G) {
  return prec == 3 ? pick_sur_fp8(method, G) : pick_sur_bf16(method, G);
}
}  // namespace bqp
