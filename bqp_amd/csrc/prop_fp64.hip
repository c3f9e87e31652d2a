// Plain FP64 form of the persistent propagation kernel (the correctness
// reference of the other forms).
#include "prop_forms.hpp"

namespace bqp {
BQP_FORM_UNIT(pick_fp64, upload_fp64, 0)

prop_fn pick_fp64(int method, int G) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return pick_w<BQP_METHOD_DOPRI5, 0>(G);
    case BQP_METHOD_DOP853: return pick_w<BQP_METHOD_DOP853, 0>(G);
    case BQP_METHOD_RK87: return pick_w<BQP_METHOD_RK87, 0>(G);
    default: return nullptr;
  }
}

prop_fn pick_team_fp64(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return propagate_team_kernel<BQP_METHOD_DOPRI5, 0>;
    case BQP_METHOD_DOP853: return propagate_team_kernel<BQP_METHOD_DOP853, 0>;
    case BQP_METHOD_RK87: return propagate_team_kernel<BQP_METHOD_RK87, 0>;
    default: return nullptr;
  }
}

prop_fn pick_pair_fp64(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return propagate_pair_kernel<BQP_METHOD_DOPRI5, 0>;
    case BQP_METHOD_DOP853: return propagate_pair_kernel<BQP_METHOD_DOP853, 0>;
    case BQP_METHOD_RK87: return propagate_pair_kernel<BQP_METHOD_RK87, 0>;
    default: return nullptr;
  }
}
}  // namespace bqp
