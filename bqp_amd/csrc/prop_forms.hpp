// Per-form translation units of the persistent propagation kernel.
//
// Each precision form (FP64, split, surrogate) instantiates its kernels in a
// translation unit of its own (prop_fp64.hip, prop_split.hip, prop_sur.hip, prop_sur_fp8.hip, prop_encke.hip)
// so the forms compile in parallel.  A unit has its own copy of the
// __constant__ tables of sectoral factors, which bqp_create fills through its
// upload function (the tableaus are compiled-in constants).
#pragma once

#include <hip/hip_runtime.h>

#include "bqp_kernels.hpp"
#include "bqp_team.hpp"

namespace bqp {

typedef void (*prop_fn)(const PropParams);

// kernel for (method, precision, waves per group), or nullptr
prop_fn pick_fp64(int method, int G);
prop_fn pick_split(int method, int G);
prop_fn pick_sur(int method, int prec, int G);  // prec 2 (BF16 network) or 3 (FP8 network)
prop_fn pick_sur_bf16(int method, int G);
prop_fn pick_sur_fp8(int method, int G);
prop_fn pick_encke(int method, int G);
// team kernels (bqp_team.hpp) for a method, FP64 Cowell or Encke; pair
// kernels also for the split form
prop_fn pick_team_fp64(int method);
prop_fn pick_team_encke(int method);
prop_fn pick_pair_fp64(int method);
prop_fn pick_pair_split(int method);
prop_fn pick_pair_encke(int method);

hipError_t upload_fp64(const double* sect, const float* sectf);
hipError_t upload_split(const double* sect, const float* sectf);
hipError_t upload_sur(const double* sect, const float* sectf);
hipError_t upload_sur_fp8(const double* sect, const float* sectf);
hipError_t upload_encke(const double* sect, const float* sectf);

// Body of a form's unit: kernel selection over the four group widths and
// the constant upload.
#define BQP_FORM_UNIT(PICK, UPLOAD, PREC_EXPR)                                                   \
  namespace {                                                                                  \
  template <int M, int P>                                                                      \
  prop_fn pick_w(int G) {                                                                      \
    switch (G) {                                                                               \
      case 1: return propagate_kernel<M, P, 1>;                                                \
      case 2: return propagate_kernel<M, P, 2>;                                                \
      case 4: return propagate_kernel<M, P, 4>;                                                \
      case 8: return propagate_kernel<M, P, 8>;                                                \
      case 16: return propagate_kernel<M, P, 16>;                                              \
      default: return nullptr;                                                                 \
    }                                                                                          \
  }                                                                                            \
  }                                                                                            \
  hipError_t UPLOAD(const double* sect, const float* sectf) {                                  \
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_sect), sect, SECT_MAX * sizeof(double));     \
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_sectf), sectf, SECT_MAX * sizeof(float)); \
    return e;                                                                                  \
  }

}  // namespace bqp
