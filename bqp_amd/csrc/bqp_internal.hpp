// Internal host-side state shared by the library's translation units
// (bqp.hip: propagation runtime; mlp.hip: learned gravity surrogate).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include <vector>

#include "../../include/bqp.h"
#include "bqp_device.hpp"
#include "mlp_device.hpp"

// Learned-surrogate state of a context (FP32 master parameters, Adam
// moments, packed inference weights, training scratch, the exact low-degree
// table, and the in-integrator settings).
struct bqp_mlp_state {
  int nlow = 4;
  double r_lo = 0.0, r_hi = 0.0;
  double out_scale[3] = {1.0, 1.0, 1.0};
  float* d_params = nullptr;  // MLP_NPARAMS
  float* d_grad = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  int64_t adam_t = 0;
  bqp::GravEntryL* d_tab_low = nullptr;  // degree <= nlow, rescaled form like the full table
  int low_jmax = 0;
  int packed_fmt = -1;  // -1: nothing packed
  void* d_packed = nullptr;
  float* d_bias = nullptr;
  bqp::MlpPacked packed{};
  void* d_work = nullptr;  // training / loss scratch
  size_t work_bytes = 0;
  int check_every = 0;        // 0: never compare against the exact field
  double threshold = 1e-9;    // m/s^2
  unsigned long long* d_sur_totals = nullptr;  // net, exact, checks, failures, band
  bqp_surrogate_stats last{};
};

struct GraphKey;  // identity of a captured launch sequence (bqp.hip)

struct bqp_ctx {
  bqp_config cfg{};
  bqp_launch launch{};
  hipStream_t stream = nullptr;
  int device = 0;
  int cus = 0;
  // field
  int nmax = -1, mmax = -1, jmax = -1;
  double mu = 0.0, radius = 0.0;
  bqp::GravEntry* d_tab = nullptr;
  bqp::GravEntryL* d_tabl = nullptr;
  bqp::GravEntryLF* d_tablf = nullptr;  // FP32 rescaled table (split precision, degree 30..90)
  bqp::GravEntryF* d_tabf = nullptr;
  size_t n_entries = 0;
  std::vector<double> h_cbar, h_sbar;  // host copy of the loaded coefficients
  // scratch (grown on demand, freed by bqp_destroy)
  unsigned int* d_queue = nullptr;
  unsigned long long* d_totals = nullptr;
  int* d_first = nullptr;
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  void* d_sort = nullptr;
  size_t sort_bytes = 0;
  void* d_part = nullptr;  // surrogate up-front check: flags and the two row lists
  void* d_cost = nullptr;  // learned cost model (bqp::CostTable), allocated with the context
  size_t part_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  GraphKey* graph_key = nullptr;
  int64_t graph_builds = 0;
  bqp_mlp_state* mlp = nullptr;
  bqp_launch_info last_launch{};  // shape of the last propagation call
  std::string err;
};


// helpers shared by the translation units
int bqp_fail(bqp_ctx* c, int code, const std::string& msg);
int bqp_ensure(bqp_ctx* c, void** ptr, size_t* have, size_t need);
void bqp_drop_graph(bqp_ctx* c);
void bqp_mlp_release(bqp_ctx* c);
int bqp_mlp_alloc_init(bqp_ctx* c, uint64_t seed);

#define BQP_HIP_TRY(ctx, expr)                                                          \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bqp_fail(ctx, BQP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
