// Encke form of the persistent propagation kernel (FP64): the deviation from
// a two-body baseline is integrated with Battin's correction and rebased every
// 5,400 s, the reference's native formulation.
#include "prop_forms.hpp"

namespace bqp {
BQP_FORM_UNIT(pick_encke, upload_encke, 4)

prop_fn pick_encke(int method, int G) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return pick_w<BQP_METHOD_DOPRI5, 4>(G);
    case BQP_METHOD_DOP853: return pick_w<BQP_METHOD_DOP853, 4>(G);
    case BQP_METHOD_RK87: return pick_w<BQP_METHOD_RK87, 4>(G);
    default: return nullptr;
  }
}

prop_fn pick_team_encke(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return propagate_team_kernel<BQP_METHOD_DOPRI5, 4>;
    case BQP_METHOD_DOP853: return propagate_team_kernel<BQP_METHOD_DOP853, 4>;
    case BQP_METHOD_RK87: return propagate_team_kernel<BQP_METHOD_RK87, 4>;
    default: return nullptr;
  }
}

prop_fn pick_pair_encke(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return propagate_pair_kernel<BQP_METHOD_DOPRI5, 4>;
    case BQP_METHOD_DOP853: return propagate_pair_kernel<BQP_METHOD_DOP853, 4>;
    case BQP_METHOD_RK87: return propagate_pair_kernel<BQP_METHOD_RK87, 4>;
    default: return nullptr;
  }
}
}  // namespace bqp
