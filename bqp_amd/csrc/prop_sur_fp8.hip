// Learned-surrogate form of the persistent propagation kernel with the FP8
// E4M3 network (hidden layers FP8 with per-layer scales; see prop_sur.hip).
#include "prop_forms.hpp"

namespace bqp {
BQP_FORM_UNIT(pick_sur_fp8, upload_sur_fp8, 3)

prop_fn pick_sur_fp8(int method, int G) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return pick_w<BQP_METHOD_DOPRI5, 3>(G);
    case BQP_METHOD_DOP853: return pick_w<BQP_METHOD_DOP853, 3>(G);
    case BQP_METHOD_RK87: return pick_w<BQP_METHOD_RK87, 3>(G);
    default: return nullptr;
  }
}
}  // namespace bqp
