// Split-precision form of the persistent propagation kernel: degree <= 4 in
// FP64, the higher-degree terms in FP32 with compensated column sums.
#include "prop_forms.hpp"

namespace bqp {
BQP_FORM_UNIT(pick_split, upload_split, 1)

prop_fn pick_split(int method, int G) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return pick_w<BQP_METHOD_DOPRI5, 1>(G);
    case BQP_METHOD_DOP853: return pick_w<BQP_METHOD_DOP853, 1>(G);
    case BQP_METHOD_RK87: return pick_w<BQP_METHOD_RK87, 1>(G);
    default: return nullptr;
  }
}

prop_fn pick_pair_split(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return propagate_pair_kernel<BQP_METHOD_DOPRI5, 1>;
    case BQP_METHOD_DOP853: return propagate_pair_kernel<BQP_METHOD_DOP853, 1>;
    case BQP_METHOD_RK87: return propagate_pair_kernel<BQP_METHOD_RK87, 1>;
    default: return nullptr;
  }
}
}  // namespace bqp
