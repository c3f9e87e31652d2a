// bqp host runtime and C ABI (include/bqp.h) over the gfx950 kernels.
//
// Boundary mirrored: LightyearSession::integrate_variable_final_into
// (dust_transfer/crates/lightyear_odeint_rs/src/session.rs:2325-2436) and
// integrate_batch_native_into (.../src/batch.rs:247-390): caller-owned flat
// buffers, rows independent, failed rows NaN-filled with a status code, the
// first failed row reported.  The field store mirrors the process-global
// coefficient pack of .../src/config.rs:24-223, but is per context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bqp_internal.hpp"
#include "bqp_aux.hpp"
#include "prop_forms.hpp"

using namespace bqp;

// ---------------------------------------------------------------- static scratch
// Device memory of the no-allocation entry point
// (bqp_propagate_batch_final_dev_noalloc): one array per GPU, held until the
// process ends.  Layout for a call of n <= BQP_NOALLOC_MAX_ROWS rows:
//   [0, 256)        work-queue head (u32)
//   [256, 512)      row totals (5 x u64)
//   [512, 1024)     first failed row (i32)
//   [1024, ...)     hand-out tables: histogram (SORT_BUCKETS x u32) | order (n x i32) | bucket (n x u16)
//   NOALLOC_ROWS_OFF  per-row status | steps | evals | rejects (4 x n x i32)
// Host code never names `scratch` as an address: hipGetSymbolAddress gives
// each device's copy (noalloc_base), and every launch takes that pointer.
constexpr size_t NOALLOC_SORT_OFF = 1024;
constexpr size_t NOALLOC_ROWS_OFF =
    (NOALLOC_SORT_OFF + (size_t)SORT_BUCKETS * 4 + (size_t)BQP_NOALLOC_MAX_ROWS * 6 + 255) / 256 * 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_ROWS_OFF + (size_t)BQP_NOALLOC_MAX_ROWS * 16;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static_assert(NOALLOC_BYTES <= BQP_NOALLOC_SCRATCH_BYTES, "include/bqp.h states the scratch size per GPU");
static_assert(SORT_BUCKETS == 4096, "BQP_NOALLOC_SCRATCH_BYTES counts 4096 histogram buckets");

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_base[NOALLOC_MAX_DEVICES] = {};

// Device address of this device's scratch array, looked up once per device
// with that device current (called from bqp_create).  nullptr on failure.
void* noalloc_base(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_base[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_noalloc_base[dev] = ptr;
  }
  return g_noalloc_base[dev];
}

// The address looked up at set-up, without a new lookup.
void* noalloc_base_known(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return g_noalloc_base[dev];
}
}  // namespace

// Identity of a captured launch sequence: a replay is valid only for the same
// buffers, sizes, settings and field.
struct GraphKey {
  PropParams p;
  int64_t n;
  bqp_config cfg;
  int G, teams, blocks, sort;
  const void* tab;
  const void* sort_buf;  // the hand-out tables' buffer (hist, order, bucket), not in p
  bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

int bqp_fail(bqp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

namespace {

int fail(bqp_ctx* c, int code, const std::string& msg) { return bqp_fail(c, code, msg); }

#define HIP_TRY(ctx, expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(ctx, BQP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ensure(bqp_ctx* c, void** ptr, size_t* have, size_t need) { return bqp_ensure(c, ptr, have, need); }

}  // namespace

// Drops the captured launch sequence: it holds the addresses of the buffers
// it was captured with, so it must not be replayed once one of them is freed.
void bqp_drop_graph(bqp_ctx* c) {
  if (c->graph_exec) {
    (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
}

int bqp_ensure(bqp_ctx* c, void** ptr, size_t* have, size_t need) {
  if (*have >= need && *ptr) return BQP_OK;
  if (*ptr) {
    (void)hipStreamSynchronize(c->stream);
    bqp_drop_graph(c);
    (void)hipFree(*ptr);
    *ptr = nullptr;
    *have = 0;
  }
  size_t cap = std::max(need, (size_t)4096);
  HIP_TRY(c, hipMalloc(ptr, cap));
  *have = cap;
  return BQP_OK;
}

namespace {

bool valid_config(const bqp_config& c, std::string& why) {
  if (c.method < 0 || c.method > 2) { why = "method must be 0 (dopri5), 1 (dop853) or 2 (rk87)"; return false; }
  if (c.control < 0 || c.control > 2) { why = "control must be 0, 1 or 2"; return false; }
  if (c.precision < 0 || c.precision > 2) { why = "precision must be 0 (fp64), 1 (split) or 2 (surrogate)"; return false; }
  if (c.drag < 0 || c.drag > 2) { why = "drag must be 0, 1 or 2"; return false; }
  if (c.formulation < 0 || c.formulation > 1) { why = "formulation must be 0 (Cowell) or 1 (Encke)"; return false; }
  if (c.formulation == BQP_FORMULATION_ENCKE && c.precision != BQP_PRECISION_FP64) {
    why = "the Encke formulation runs in FP64 (precision 0)";
    return false;
  }
  if (c.control == BQP_CONTROL_SCALED &&
      !(c.rtol >= 0.0 && c.atol_pos >= 0.0 && c.atol_vel >= 0.0 && (c.rtol > 0 || (c.atol_pos > 0 && c.atol_vel > 0)))) {
    why = "scaled control needs rtol >= 0, atol >= 0, not all zero";
    return false;
  }
  if (c.control != BQP_CONTROL_SCALED && !(c.eps > 0.0 && std::isfinite(c.eps))) {
    why = "absolute/compat control needs eps > 0";
    return false;
  }
  if (!std::isfinite(c.omega) || !std::isfinite(c.theta0) || !std::isfinite(c.r_atm) || !(c.r_min >= 0.0)) {
    why = "omega, theta0, r_atm must be finite and r_min >= 0";
    return false;
  }
  if (!(c.ground_flattening >= 0.0 && c.ground_flattening < 1.0)) {
    why = "ground_flattening must be in [0, 1)";
    return false;
  }
  if (c.ground_flattening > 0.0 && c.r_min > 0.0 && !(c.r_atm > 0.0)) {
    why = "an ellipsoidal ground needs r_atm > 0 (its equatorial radius)";
    return false;
  }
  return true;
}

// ---------------------------------------------------------------- packing

// Builds the regrouped entry table (see bqp_device.hpp) from normalised
// coefficients; GM/R^2 is folded into the contribution coefficients.
void pack_field(int N, int M, const double* C, const double* S, double mu, double R,
                std::vector<GravEntry>& out, int& jmax) {
  jmax = std::min(N + 1, M + 1);
  const double gm = mu / (R * R);
  auto tri = [](int n, int m) { return n * (n + 1) / 2 + m; };
  auto cc = [&](int n, int m) -> double {
    if (n < 0 || n > N || m < 0 || m > n || m > M) return 0.0;
    return C[tri(n, m)];
  };
  auto ss = [&](int n, int m) -> double {
    if (n < 0 || n > N || m < 1 || m > n || m > M) return 0.0;
    return S[tri(n, m)];
  };
  auto kk = [](int n) { return (2.0 * n + 1.0) / (2.0 * n + 3.0); };
  // f1 for m = 0 is the zonal factor sqrt(k (n+1)(n+2) / 2)
  auto f1 = [&](int n, int m) {
    return m == 0 ? std::sqrt(kk(n) * (n + 1.0) * (n + 2.0) / 2.0)
                  : 0.5 * std::sqrt(kk(n) * (n + m + 1.0) * (n + m + 2.0));
  };
  auto f2 = [&](int n, int m) {
    return 0.5 * std::sqrt(kk(n) * (n - m + 1.0) * (n - m + 2.0)) * (m == 1 ? std::sqrt(2.0) : 1.0);
  };
  auto f3 = [&](int n, int m) { return std::sqrt(kk(n) * (n + m + 1.0) * (n - m + 1.0)); };
  out.clear();
  out.reserve((size_t)col_offset(N, jmax + 1));
  for (int j = 0; j <= jmax; ++j) {
    const int nlo = j == 0 ? 1 : j;
    for (int n = nlo; n <= N + 1; ++n) {
      GravEntry e{};
      if (n == j) {
        e.a = (j == 1) ? std::sqrt(3.0) : std::sqrt((2.0 * j + 1.0) / (2.0 * j));
        e.b = 0.0;
      } else {
        const double dn = n, dm = j;
        e.a = std::sqrt((2 * dn + 1) * (2 * dn - 1) / ((dn - dm) * (dn + dm)));
        e.b = (n >= j + 2) ? std::sqrt((2 * dn + 1) * (dn + dm - 1) * (dn - dm - 1) /
                                       ((dn - dm) * (dn + dm) * (2 * dn - 3)))
                           : 0.0;
      }
      const int np = n - 1;  // field degree whose terms read V/W(n, .)
      double al = 0, ga = 0, be = 0, de = 0, zc = 0, zs = 0;
      if (j - 1 >= 0 && j - 1 <= np) {
        al = f1(np, j - 1) * cc(np, j - 1);
        ga = (j - 1 >= 1) ? f1(np, j - 1) * ss(np, j - 1) : 0.0;
      }
      if (j + 1 <= np) {
        be = f2(np, j + 1) * cc(np, j + 1);
        de = f2(np, j + 1) * ss(np, j + 1);
      }
      if (j <= np) {
        zc = f3(np, j) * cc(np, j);
        zs = f3(np, j) * ss(np, j);
      }
      e.xv = gm * (be - al);
      e.xw = gm * (de - ga);
      e.yv = gm * (ga + de);
      e.yw = gm * (-al - be);
      e.zv = gm * (-zc);
      e.zw = gm * (-zs);
      out.push_back(e);
    }
  }
}

// Rescaled copy of an (a, b) table (the GravEntryL form): in each column s_n
// is the running product of the recursion coefficients a_n (exact formulas in
// long double), started at the column's sectoral factor product
// S_j = s_1 ... s_j (S_0 = 1) on its first entry; g_n = b_n / (a_n a_{n-1})
// and the summation coefficients are multiplied by s_n.  The kernel then
// starts column j from (x~ + i y~)^j R/r without the sectoral factors.
void pack_lambda(int N, int jmax, const std::vector<GravEntry>& ab, std::vector<GravEntryL>& out) {
  out.clear();
  out.reserve(ab.size());
  auto a_nm = [](int n, int m) -> long double {
    const long double dn = n, dm = m;
    return sqrtl((2 * dn + 1) * (2 * dn - 1) / ((dn - dm) * (dn + dm)));
  };
  auto b_nm = [](int n, int m) -> long double {
    if (n < m + 2) return 0.0L;
    const long double dn = n, dm = m;
    return sqrtl((2 * dn + 1) * (dn + dm - 1) * (dn - dm - 1) / ((dn - dm) * (dn + dm) * (2 * dn - 3)));
  };
  size_t idx = 0;
  long double sect = 1.0L;  // S_j = s_1 ... s_j: the column's sectoral factor product
  for (int j = 0; j <= jmax; ++j) {
    const int nlo = j == 0 ? 1 : j;
    if (j >= 1) sect *= (j == 1) ? sqrtl(3.0L) : sqrtl((2.0L * j + 1.0L) / (2.0L * j));
    long double s = sect, aprev = 0.0L;
    for (int n = nlo; n <= N + 1; ++n, ++idx) {
      const GravEntry& e = ab[idx];
      GravEntryL L{};
      long double scale = sect;
      if (!(j >= 1 && n == j)) {  // recursion entry (the sectoral entry keeps S_j)
        const long double a = a_nm(n, j), b = b_nm(n, j);
        s *= a;
        scale = s;
        L.g = aprev > 0.0L ? (double)(b / (a * aprev)) : 0.0;
        aprev = a;
      }
      L.xv = (double)((long double)e.xv * scale);
      L.xw = (double)((long double)e.xw * scale);
      L.yv = (double)((long double)e.yv * scale);
      L.yw = (double)((long double)e.yw * scale);
      L.zv = (double)((long double)e.zv * scale);
      L.zw = (double)((long double)e.zw * scale);
      out.push_back(L);
    }
  }
}

FieldView field_view(const bqp_ctx* c) {
  FieldView f;
  f.tab = c->d_tab;
  f.tabf = c->d_tabf;
  f.tabl = c->d_tabl;
  f.tablf = c->d_tablf;
  f.nmax = c->nmax;
  f.jmax = c->jmax;
  f.radius = c->radius;
  return f;
}

// ---------------------------------------------------------------- launch

prop_fn pick_kernel(int method, int prec, int G) {
  if (prec == 0) return pick_fp64(method, G);
  if (prec == 1) return pick_split(method, G);
  if (prec == 4) return pick_encke(method, G);
  return pick_sur(method, prec, G);
}

int default_g(const bqp_ctx* c, int64_t n) {
  // Split each 64-trajectory group over enough waves that about four waves
  // are resident per SIMD (the occupancy the kernel is compiled for); measured
  // at 65,536 rows x degree 70: G=1 8.8, G=2 19.9, G=4 29.3, G=8 28.6 TFLOP/s.
  // Larger batches keep G = 4 (the persistent grid then refills lanes): at
  // 262,144 rows x 6 h x degree 70, G = 1 / 2 / 4 / 8 measured 47.7k / 49.6k /
  // 55.3k / 51.6k orbit-days/s (the G = 1 kernel holds only 2 waves per SIMD:
  // it needs 232 VGPRs).
  const int64_t groups = (n + 63) / 64;
  const int64_t simds = (int64_t)c->cus * 4;
  if (groups >= simds / 2) return 4;
  // Fewer groups than CUs: each group's workgroup has a CU to itself, and 16
  // waves give it four per SIMD instead of two.
  if (groups > (int64_t)c->cus) return 8;
  return 16;
}

// First queue position: the first 64 * blocks positions are dealt to the
// blocks by index (one group each).
unsigned queue_start(int64_t n, int blocks) { return (unsigned)std::min<int64_t>(n, 64LL * blocks); }

// The learned cost model when the hand-out order uses it, else nullptr.
CostTable* cost_table(bqp_ctx* c) {
  return c->launch.sort_by_cost == 1 ? (CostTable*)c->d_cost : nullptr;
}

// Adds the call's finished rows to the learned cost model (final-state calls
// that use it).
void enqueue_cost_learn(bqp_ctx* c, const PropParams& p, int64_t n) {
  CostTable* tab = cost_table(c);
  if (!tab || p.t_eval != nullptr || !p.tf || !p.t0) return;
  hipLaunchKernelGGL(cost_learn_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p.y0, p.t0,
                     p.tf, p.status, p.evals, c->mu, c->radius, (int)n, tab);
}

// Forgets what the cost model learned (new field or configuration).
int reset_cost_model(bqp_ctx* c) {
  hipLaunchKernelGGL(cost_clear_kernel, dim3(1), dim3(256), 0, c->stream, (CostTable*)c->d_cost);
  return hipGetLastError() == hipSuccess ? BQP_OK : BQP_E_DEVICE;
}

// Kernel and launch shape of one propagation over n rows.
struct LaunchShape {
  prop_fn kern = nullptr;
  int threads = 0;       // workgroup size
  int blocks = 0;        // persistent grid
  int G = 0;             // fixed-G kernel: waves per group (0 for the team kernel)
  int teams = 0;         // team kernel: groups per workgroup (0 for the fixed-G kernel)
  unsigned queue0 = 0;   // first queue position: the groups of the dealt first round come before it
};

// Device memory a propagation call uses beyond its inputs and outputs: the
// work-queue head, the row totals, the first failed row and the hand-out
// tables (histogram | order | bucket).  The context's own buffers, or the
// library's static scratch array (the no-allocation entry point).
struct CallScratch {
  unsigned int* queue;
  unsigned long long* totals;
  int* first;
  void* sort;
};

CallScratch context_scratch(bqp_ctx* c) { return CallScratch{c->d_queue, c->d_totals, c->d_first, c->d_sort}; }

// Enqueues the whole call (order, queue reset, persistent kernel, row
// statistics) on the context stream.  Pure kernel launches: no allocation,
// no memset, no synchronisation, so the sequence can be captured in a graph.
void enqueue_propagation(bqp_ctx* c, const PropParams& p, int64_t n, const LaunchShape& ls, bool sort,
                         bool events, const CallScratch& sc) {
  PropParams q = p;
  if (sort) {
    char* base = (char*)sc.sort;
    unsigned int* hist = (unsigned int*)base;
    int32_t* order = (int32_t*)(base + SORT_BUCKETS * sizeof(unsigned int));
    unsigned short* bucket = (unsigned short*)(base + SORT_BUCKETS * sizeof(unsigned int) + n * sizeof(int32_t));
    const unsigned nb = (unsigned)((n + 255) / 256);
    CostTable* tab = cost_table(c);
    if (tab) hipLaunchKernelGGL(cost_fill_kernel, dim3(1), dim3(1024), 0, c->stream, tab);
    hipLaunchKernelGGL(sort_clear_kernel, dim3(1), dim3(1024), 0, c->stream, hist);
    hipLaunchKernelGGL(sort_hist_kernel, dim3(nb), dim3(256), 0, c->stream, p.y0, p.t0, p.tf, c->mu, c->radius,
                       tab, (int)n, hist, bucket);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, c->stream, hist);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(nb), dim3(256), 0, c->stream, bucket, (int)n, hist, order);
    q.order = order;
  }
  q.deal_cus = c->cus;
  q.teams = ls.teams;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, c->stream, sc.queue, ls.queue0);
  if (p.sur_totals) hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, p.sur_totals, 5);
  if (events) (void)hipEventRecord(c->ev0, c->stream);
  hipLaunchKernelGGL(ls.kern, dim3(ls.blocks), dim3(ls.threads), 0, c->stream, q);
  if (events) (void)hipEventRecord(c->ev1, c->stream);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, sc.totals, 5);
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, c->stream, sc.first, 0x7fffffff);
  hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0,
                     c->stream, q.status, q.steps, q.evals, q.rejects, (int)n, sc.totals);
  hipLaunchKernelGGL(first_failed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, q.status,
                     (int)n, sc.first);
  enqueue_cost_learn(c, p, n);
}

// Algorithmic FLOPs of one gravity evaluation in the rescaled recursion
// (FMA = 2, other operations 1): set-up 16, column 0 9 per entry (g rho, U,
// three sums), column j >= 1 18 for its sectoral value (one complex
// multiply) and either the fold of its six sums times the sectoral value
// (factored sweep) or the sectoral entry's six terms, plus per recursion
// entry 16 (factored sweep: g rho, Q, six sums) or 19 (the split form's
// (a, b) table: g rho, V and W, six sums).  The per-wave powers z^gap (at
// most 2G complex multiplies per wave) are not counted.
double grav_flops(int N, int jmax, bool factored = true) {
  const double per_entry = factored ? 16.0 : 19.0;
  double f = 16.0 + 9.0 * (N + 1.0);
  for (int j = 1; j <= jmax; ++j) f += 18.0 + per_entry * (N + 1.0 - j);
  return f;
}

int grid_for(bqp_ctx* c, prop_fn kern, int G, int64_t n, int* blocks) {
  const int64_t groups = (n + 63) / 64;
  int b = c->launch.blocks;
  if (b <= 0) {
    int per_cu = 0;
    HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, 64 * G, 0));
    if (per_cu < 1) per_cu = 1;
    b = (int)std::min<int64_t>(groups, (int64_t)per_cu * c->cus);
  }
  *blocks = (int)std::max<int64_t>(1, std::min<int64_t>(b, groups));
  return BQP_OK;
}

// Records a launch shape as the context's last one (bqp_last_launch).
void note_launch(bqp_ctx* c, const LaunchShape& ls) {
  bqp_launch_info& li = c->last_launch;
  li.kernel = ls.teams == 0 ? BQP_KERNEL_FIXED : (ls.G == PAIR_WAVES ? BQP_KERNEL_PAIR : BQP_KERNEL_TEAM);
  li.waves_per_group = ls.G;
  li.teams = ls.teams;
  li.blocks = ls.blocks;
  li.threads = ls.threads;
}

// Whether an automatic launch (waves_per_group = teams_per_block = 0) takes
// the pair kernel: fields of degree >= PAIR_AUTO_MIN_N whose batch is one
// round of G = 4 groups (each lane holds one row, so the launch ends with
// groups running alone on their CUs, which the pair kernel runs on eight
// waves instead of four).  Measured on 65,536 LEO rows x 1 d with drag, pair
// against G = 4: degree 40 -18%, 50 -3%, 60 +5%, 70 +12% (7 d: +12.4%); at
// degree 70 with 32,768 rows +14%, with 98,304 rows (lanes refill) -2%;
// Encke at degree 70 +7%; split form at degree 70 +10% (1 d), +11% (7 d),
// degree 50 +4%.  The pair kernel's results equal G = 8 bit for bit.
bool auto_pair(const bqp_ctx* c, const bqp_launch& L, int64_t n) {
  if (L.waves_per_group != 0 || L.teams_per_block != 0 || c->nmax < PAIR_AUTO_MIN_N) return false;
  const int64_t groups = (n + 63) / 64;
  return default_g(c, n) == 4 && groups <= (int64_t)c->cus * 4;
}

// The team kernel (bqp_team.hpp) takes FP64 Cowell and Encke calls that ask
// for it (teams_per_block 1, 2 or 4 with an automatic group width) on fields
// of degree >= PIPE_MIN_N: one 16-wave workgroup per CU carrying that many
// groups.  The pair kernel takes FP64 and split Cowell and Encke calls with
// waves_per_group 8 and teams_per_block 2: two groups per 8-wave workgroup.
// Other calls take the fixed-G kernel.
int shape_for(bqp_ctx* c, int kprec, int64_t n, LaunchShape* out) {
  const bqp_launch& L = c->launch;
  const int64_t groups = (n + 63) / 64;
  *out = LaunchShape{};
  const bool pair_ok = (kprec == 0 || kprec == 1 || kprec == 4);
  const bool pair_asked = L.waves_per_group == PAIR_WAVES && L.teams_per_block == PAIR_GROUPS;
  if ((pair_asked && pair_ok) || (pair_ok && auto_pair(c, L, n))) {
    out->kern = (kprec == 4)   ? pick_pair_encke(c->cfg.method)
                : (kprec == 1) ? pick_pair_split(c->cfg.method)
                               : pick_pair_fp64(c->cfg.method);
    if (!out->kern) return fail(c, BQP_E_INVALID, "unknown method");
    const int64_t need = (groups + PAIR_GROUPS - 1) / PAIR_GROUPS;
    int64_t b = L.blocks;
    if (b <= 0) {
      int per_cu = 0;
      HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)out->kern, 64 * PAIR_WAVES, 0));
      b = (int64_t)std::max(per_cu, 1) * c->cus;
    }
    b = std::max<int64_t>(1, std::min<int64_t>(b, need));
    out->threads = 64 * PAIR_WAVES;
    out->blocks = (int)b;
    out->G = PAIR_WAVES;
    out->teams = PAIR_GROUPS;
    out->queue0 = (unsigned)std::min<int64_t>(n, 64LL * PAIR_GROUPS * b);
    return BQP_OK;
  }
  const bool team = L.waves_per_group == 0 && L.teams_per_block > 0 && (kprec == 0 || kprec == 4) &&
                    c->nmax >= PIPE_MIN_N;
  if (team) {
    const int nt = L.teams_per_block;
    out->kern = (kprec == 4) ? pick_team_encke(c->cfg.method) : pick_team_fp64(c->cfg.method);
    if (!out->kern) return fail(c, BQP_E_INVALID, "unknown method");
    const int64_t need = (groups + nt - 1) / nt;
    int64_t b = L.blocks > 0 ? L.blocks : std::min<int64_t>(need, c->cus);
    b = std::max<int64_t>(1, std::min<int64_t>(b, need));
    out->threads = 64 * TEAM_WAVES;
    out->blocks = (int)b;
    out->teams = nt;
    out->queue0 = (unsigned)std::min<int64_t>(n, 64LL * nt * b);
    return BQP_OK;
  }
  out->G = L.waves_per_group > 0 ? L.waves_per_group : default_g(c, n);
  out->kern = pick_kernel(c->cfg.method, kprec, out->G);
  if (!out->kern) return fail(c, BQP_E_INVALID, "waves_per_group must be 1, 2, 4, 8 or 16");
  out->threads = 64 * out->G;
  const int rc = grid_for(c, out->kern, out->G, n, &out->blocks);
  if (rc) return rc;
  out->queue0 = queue_start(n, out->blocks);
  return BQP_OK;
}

// Checked surrogate propagation with the first check made up front: the
// rows whose initial state fails it run the plain FP64 kernel (the same
// arithmetic as the surrogate kernel's exact path, so the same results), the
// others the surrogate kernel.  Both groups keep the longest-first order.
int run_surrogate_partitioned(bqp_ctx* c, PropParams p, int64_t n, int kprec, bqp_stats* stats) {
  int rc = ensure(c, &c->d_part, &c->part_bytes, (size_t)(3 * n + 64) * sizeof(int));
  if (rc) return rc;
  int* flags = (int*)c->d_part;
  int* exact_rows = flags + n;
  int* sur_rows = exact_rows + n;
  int* counts = sur_rows + n;
  // longest-first hand-out order over all rows, then a stable split of it
  const bool sort = c->launch.sort_by_cost != 0 && p.t_eval == nullptr && n > 64;
  int32_t* order = nullptr;
  if (sort) {
    const size_t need = (size_t)SORT_BUCKETS * sizeof(unsigned int) + (size_t)n * sizeof(int32_t) +
                        (size_t)n * sizeof(unsigned short) + 512;
    rc = ensure(c, &c->d_sort, &c->sort_bytes, need);
    if (rc) return rc;
    char* base = (char*)c->d_sort;
    unsigned int* hist = (unsigned int*)base;
    order = (int32_t*)(base + SORT_BUCKETS * sizeof(unsigned int));
    unsigned short* bucket = (unsigned short*)(base + SORT_BUCKETS * sizeof(unsigned int) + n * sizeof(int32_t));
    const unsigned nbs = (unsigned)((n + 255) / 256);
    CostTable* tab = cost_table(c);
    if (tab) hipLaunchKernelGGL(cost_fill_kernel, dim3(1), dim3(1024), 0, c->stream, tab);
    hipLaunchKernelGGL(sort_clear_kernel, dim3(1), dim3(1024), 0, c->stream, hist);
    hipLaunchKernelGGL(sort_hist_kernel, dim3(nbs), dim3(256), 0, c->stream, p.y0, p.t0, p.tf, c->mu, c->radius,
                       tab, (int)n, hist, bucket);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, c->stream, hist);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(nbs), dim3(256), 0, c->stream, bucket, (int)n, hist, order);
  }
  const unsigned nb = (unsigned)((n + 63) / 64);
  if (kprec == 3)
    hipLaunchKernelGGL(sur_precheck_kernel<3>, dim3(nb), dim3(64), 0, c->stream, p, (int)n, flags);
  else
    hipLaunchKernelGGL(sur_precheck_kernel<2>, dim3(nb), dim3(64), 0, c->stream, p, (int)n, flags);
  hipLaunchKernelGGL(partition_kernel, dim3(1), dim3(1024), 0, c->stream, flags, order, (int)n, exact_rows,
                     sur_rows, counts);
  HIP_TRY(c, hipGetLastError());
  int cnt[2] = {0, 0};
  HIP_TRY(c, hipMemcpyAsync(cnt, counts, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // the failing rows take the kernel the plain FP64 form would take for them
  LaunchShape se, ss;
  if (cnt[0] > 0 && (rc = shape_for(c, 0, cnt[0], &se))) return rc;
  if (cnt[1] > 0 && (rc = shape_for(c, kprec, cnt[1], &ss))) return rc;
  note_launch(c, cnt[0] > 0 ? se : ss);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, p.sur_totals, 5);
  HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
  if (cnt[0] > 0) {
    PropParams q = p;
    q.order = exact_rows;
    q.n = cnt[0];
    q.sur_totals = nullptr;
    q.deal_cus = c->cus;
    q.teams = se.teams;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, c->stream, c->d_queue, se.queue0);
    hipLaunchKernelGGL(se.kern, dim3(se.blocks), dim3(se.threads), 0, c->stream, q);
  }
  if (cnt[1] > 0) {
    PropParams q = p;
    q.order = sur_rows;
    q.n = cnt[1];
    q.deal_cus = c->cus;
    q.teams = ss.teams;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, c->stream, c->d_queue, ss.queue0);
    hipLaunchKernelGGL(ss.kern, dim3(ss.blocks), dim3(ss.threads), 0, c->stream, q);
  }
  HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, c->d_totals, 5);
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, c->stream, c->d_first, 0x7fffffff);
  hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0,
                     c->stream, p.status, p.steps, p.evals, p.rejects, (int)n, c->d_totals);
  hipLaunchKernelGGL(first_failed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p.status,
                     (int)n, c->d_first);
  enqueue_cost_learn(c, p, n);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  unsigned long long tot[5] = {0, 0, 0, 0, 0}, st[5] = {0, 0, 0, 0, 0};
  int first = 0x7fffffff;
  HIP_TRY(c, hipMemcpy(tot, c->d_totals, sizeof(tot), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(&first, c->d_first, sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(st, c->mlp->d_sur_totals, sizeof(st), hipMemcpyDeviceToHost));
  // the rows that failed up front: one check each, every evaluation exact
  bqp_surrogate_stats& L = c->mlp->last;
  L.net_evals = (int64_t)st[0];
  L.exact_evals = (int64_t)tot[1] - (int64_t)st[0];
  L.checks = (int64_t)st[2] + cnt[0];
  L.check_failures = (int64_t)st[3] + cnt[0];
  L.band_fallbacks = (int64_t)st[4];
  if (stats) {
    float ms = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const double full = bqp_flops_per_rhs(c);
    const double gfull = grav_flops(c->nmax, c->jmax);
    const double low = grav_flops(std::min(c->mlp->nlow, c->nmax), c->mlp->low_jmax);
    stats->steps = (int64_t)tot[0];
    stats->evals = (int64_t)tot[1];
    stats->rejects = (int64_t)tot[2];
    stats->rows_ok = (int64_t)tot[3];
    stats->rows_failed = (int64_t)tot[4];
    stats->first_failed_row = (tot[4] > 0) ? first : -1;
    stats->kernel_ms = ms;
    stats->flops = (double)st[0] * (MLP_FLOPS + low + (full - gfull)) + (double)st[1] * (full + low) +
                   ((double)tot[1] - (double)st[0] - (double)st[1]) * full;
  }
  if (tot[4] > 0) {
    char buf[160];
    snprintf(buf, sizeof(buf), "%llu row(s) failed; first failed row %d", tot[4], first);
    c->err = buf;
    return BQP_E_ROWS_FAILED;
  }
  return BQP_OK;
}

// Runs one propagation over device buffers; fills per-row outputs.  With
// `fixed` the call uses that scratch and allocates nothing: direct launches
// only (no graph), and a checked surrogate keeps its checks in the kernel
// instead of the up-front partition.
int run_propagation(bqp_ctx* c, PropParams p, int64_t n, bqp_stats* stats, const CallScratch* fixed) {
  if (c->nmax < 0 || !c->d_tab) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  if (n <= 0) return BQP_OK;
  if (n > 0x7fffffffLL) return fail(c, BQP_E_INVALID, "n exceeds 2^31-1 rows");
  const bqp_config& cfg = c->cfg;
  const bool sur = cfg.precision == BQP_PRECISION_SURROGATE;
  int kprec = cfg.precision;
  if (sur) {
    const bqp_mlp_state* s = c->mlp;
    if (!s || !s->d_tab_low) return fail(c, BQP_E_INVALID, "surrogate precision needs bqp_surrogate_init");
    if (s->packed_fmt != BQP_MLP_BF16 && s->packed_fmt != BQP_MLP_FP8)
      return fail(c, BQP_E_INVALID, "surrogate precision needs packed weights (bqp_surrogate_pack)");
    kprec = s->packed_fmt == BQP_MLP_FP8 ? 3 : 2;
  }
  if (cfg.formulation == BQP_FORMULATION_ENCKE) kprec = 4;
  if (c->launch.waves_per_group != 0 && !pick_kernel(cfg.method, kprec, c->launch.waves_per_group))
    return fail(c, BQP_E_INVALID, "waves_per_group must be 1, 2, 4, 8 or 16");
  if (cfg.precision == BQP_PRECISION_SPLIT && !c->d_tabf)
    return fail(c, BQP_E_NO_FIELD, "split precision needs the FP32 table (reload the field)");

  // all allocation happens here, before any launch is enqueued or captured
  const bool sort = c->launch.sort_by_cost != 0 && p.t_eval == nullptr && n > 64;
  if (sort && !fixed) {
    const size_t need = (size_t)SORT_BUCKETS * sizeof(unsigned int) + (size_t)n * sizeof(int32_t) +
                        (size_t)n * sizeof(unsigned short) + 512;
    int rc = ensure(c, &c->d_sort, &c->sort_bytes, need);
    if (rc) return rc;
  }
  const CallScratch sc = fixed ? *fixed : context_scratch(c);
  p.order = nullptr;
  p.queue = sc.queue;
  p.n = (int)n;
  p.field = field_view(c);
  p.mu = c->mu;
  p.model.omega = cfg.omega;
  p.model.theta0 = cfg.theta0;
  p.model.r_atm = cfg.r_atm;
  p.model.r_min = cfg.r_min;
  p.model.flat = cfg.ground_flattening;
  p.model.drag = cfg.drag;
  p.control = cfg.control;
  p.max_steps = (int)std::min<int64_t>(cfg.max_steps > 0 ? cfg.max_steps : 500000, 1 << 29);
  p.max_rejects = cfg.max_rejects > 0 ? cfg.max_rejects : 50;
  p.rtol = cfg.rtol;
  p.atol_pos = cfg.atol_pos;
  p.atol_vel = cfg.atol_vel;
  p.eps = cfg.eps;
  p.h0 = cfg.h0;
  p.h_min = cfg.h_min;
  p.h_max = cfg.h_max;
  if (sur) {
    const bqp_mlp_state* s = c->mlp;
    p.low = field_view(c);
    p.low.tab = nullptr;
    p.low.tabl = s->d_tab_low;
    p.low.tabf = nullptr;
    p.low.tablf = nullptr;
    p.low.nmax = std::min(s->nlow, c->nmax);
    p.low.jmax = s->low_jmax;
    p.mlp = s->packed;
    p.sur_r_lo = s->r_lo;
    p.sur_r_hi = s->r_hi;
    p.sur_thresh2 = s->threshold * s->threshold;
    p.sur_check_every = s->check_every;
    p.sur_totals = s->d_sur_totals;
  } else {
    p.low = FieldView{};
    p.mlp = MlpPacked{};
    p.sur_r_lo = p.sur_r_hi = p.sur_thresh2 = 0.0;
    p.sur_check_every = 0;
    p.sur_totals = nullptr;
  }

  if (sur && c->mlp->check_every > 0 && !c->launch.use_graph && !fixed)
    return run_surrogate_partitioned(c, p, n, kprec, stats);

  LaunchShape ls;
  {
    const int rc = shape_for(c, kprec, n, &ls);
    if (rc) return rc;
  }
  note_launch(c, ls);

  if (c->launch.use_graph && !fixed) {
    GraphKey key;
    std::memset(&key, 0, sizeof(key));
    key.p = p;
    key.n = n;
    key.cfg = cfg;
    key.G = ls.G;
    key.teams = ls.teams;
    key.blocks = ls.blocks;
    key.sort = (sort ? 8 : 0) | c->launch.sort_by_cost;  // the mode decides the cost-model kernels
    key.tab = c->d_tab;
    key.sort_buf = sort ? c->d_sort : nullptr;
    if (!c->graph_key) {
      c->graph_key = new (std::nothrow) GraphKey();
      if (!c->graph_key) return fail(c, BQP_E_DEVICE, "out of host memory");
    }
    if (!c->graph_exec || !(key == *c->graph_key)) {
      bqp_drop_graph(c);
      HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_propagation(c, p, n, ls, sort, false, sc);
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
      if (ce != hipSuccess || !graph) return fail(c, BQP_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        c->graph_exec = nullptr;
        return fail(c, BQP_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
      }
      std::memcpy((void*)c->graph_key, &key, sizeof(key));  // padding included (the key compares bytes)
      ++c->graph_builds;
    }
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
  } else {
    enqueue_propagation(c, p, n, ls, sort, true, sc);
  }
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  unsigned long long tot[5] = {0, 0, 0, 0, 0};
  int first = 0x7fffffff;
  HIP_TRY(c, hipMemcpy(tot, sc.totals, sizeof(tot), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(&first, sc.first, sizeof(int), hipMemcpyDeviceToHost));
  if (stats) {
    float ms = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    stats->steps = (int64_t)tot[0];
    stats->evals = (int64_t)tot[1];
    stats->rejects = (int64_t)tot[2];
    stats->rows_ok = (int64_t)tot[3];
    stats->rows_failed = (int64_t)tot[4];
    stats->first_failed_row = (tot[4] > 0) ? first : -1;
    stats->kernel_ms = ms;
    stats->flops = bqp_flops_per_rhs(c) * (double)tot[1];
  }
  if (sur) {
    unsigned long long st[5] = {0, 0, 0, 0, 0};
    HIP_TRY(c, hipMemcpy(st, c->mlp->d_sur_totals, sizeof(st), hipMemcpyDeviceToHost));
    bqp_surrogate_stats& L = c->mlp->last;
    L.net_evals = (int64_t)st[0];
    L.exact_evals = (int64_t)st[1];
    L.checks = (int64_t)st[2];
    L.check_failures = (int64_t)st[3];
    L.band_fallbacks = (int64_t)st[4];
    if (stats) {
      // network evaluation: MLP (MFMA) + exact low-degree part + frame/drag;
      // exact evaluation: low-degree part (always computed) + the full RHS
      const double full = bqp_flops_per_rhs(c);
      const double gfull = grav_flops(c->nmax, c->jmax);
      const double low = grav_flops(std::min(c->mlp->nlow, c->nmax), c->mlp->low_jmax);
      stats->flops = (double)st[0] * (MLP_FLOPS + low + (full - gfull)) + (double)st[1] * (full + low);
    }
  }
  if (tot[4] > 0) {
    char buf[160];
    snprintf(buf, sizeof(buf), "%llu row(s) failed; first failed row %d", tot[4], first);
    c->err = buf;
    return BQP_E_ROWS_FAILED;
  }
  return BQP_OK;
}

}  // namespace

// ================================================================ C ABI

extern "C" {

int32_t bqp_default_config(bqp_config* out) {
  if (!out) return BQP_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->method = BQP_METHOD_RK87;
  out->control = BQP_CONTROL_SCALED;
  out->precision = BQP_PRECISION_FP64;
  out->drag = BQP_DRAG_NONE;
  out->rtol = 1e-12;
  out->atol_pos = 1e-12 * 6378137.0;
  out->atol_vel = 1e-12 * std::sqrt(3.986004418e14 / 6378137.0);
  out->eps = 1e-8;
  out->max_steps = 500000;
  out->max_rejects = 50;
  out->omega = 7.292115e-5;
  out->theta0 = 0.0;
  out->r_atm = 6378137.0;
  out->r_min = 0.0;
  out->ground_flattening = 0.0;
  return BQP_OK;
}

int32_t bqp_create(const bqp_config* cfg, bqp_ctx** out) {
  if (!out) return BQP_E_INVALID;
  *out = nullptr;
  bqp_config def;
  bqp_default_config(&def);
  const bqp_config& use = cfg ? *cfg : def;
  std::string why;
  if (!valid_config(use, why)) return BQP_E_INVALID;
  bqp_ctx* c = new (std::nothrow) bqp_ctx();
  if (!c) return BQP_E_DEVICE;
  c->cfg = use;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    delete c;
    return BQP_E_DEVICE;
  }
  c->device = dev;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipMalloc(&c->d_queue, 256) != hipSuccess || hipMalloc(&c->d_totals, 256) != hipSuccess ||
      hipMalloc(&c->d_first, 256) != hipSuccess || hipMalloc(&c->d_cost, sizeof(CostTable)) != hipSuccess ||
      reset_cost_model(c) != BQP_OK) {
    bqp_destroy(c);
    return BQP_E_DEVICE;
  }
  c->cus = prop.multiProcessorCount;
  c->launch.sort_by_cost = 1;
  // the no-allocation entry point's scratch address on this device (the
  // lookup allocates nothing; a failure only disables that entry point)
  (void)noalloc_base(dev);
  {
    double sect[SECT_MAX];
    float sectf[SECT_MAX];
    sect[0] = 0.0;
    for (int j = 1; j < SECT_MAX; ++j) sect[j] = (j == 1) ? std::sqrt(3.0) : std::sqrt((2.0 * j + 1.0) / (2.0 * j));
    for (int j = 0; j < SECT_MAX; ++j) sectf[j] = (float)sect[j];
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_sect), sect, sizeof(sect)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_sectf), sectf, sizeof(sectf)) != hipSuccess ||
        upload_fp64(sect, sectf) != hipSuccess || upload_split(sect, sectf) != hipSuccess ||
        upload_sur(sect, sectf) != hipSuccess || upload_sur_fp8(sect, sectf) != hipSuccess ||
        upload_encke(sect, sectf) != hipSuccess) {
      bqp_destroy(c);
      return BQP_E_DEVICE;
    }
  }
  *out = c;
  return BQP_OK;
}

int32_t bqp_destroy(bqp_ctx* c) {
  if (!c) return BQP_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->d_tab);
  (void)hipFree(c->d_tabf);
  (void)hipFree(c->d_tabl);
  (void)hipFree(c->d_tablf);
  (void)hipFree(c->d_queue);
  (void)hipFree(c->d_totals);
  (void)hipFree(c->d_first);
  (void)hipFree(c->d_stage);
  (void)hipFree(c->d_sort);
  (void)hipFree(c->d_part);
  (void)hipFree(c->d_cost);
  bqp_mlp_release(c);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c->graph_key;
  delete c;
  return BQP_OK;
}

int32_t bqp_set_config(bqp_ctx* c, const bqp_config* cfg) {
  if (!c || !cfg) return BQP_E_INVALID;
  std::string why;
  if (!valid_config(*cfg, why)) return fail(c, BQP_E_INVALID, why);
  const bool changed = std::memcmp(&c->cfg, cfg, sizeof(bqp_config)) != 0;
  c->cfg = *cfg;
  if (changed && reset_cost_model(c) != BQP_OK) return fail(c, BQP_E_DEVICE, "cost model reset failed");
  return BQP_OK;
}

int32_t bqp_set_launch(bqp_ctx* c, const bqp_launch* l) {
  if (!c || !l) return BQP_E_INVALID;
  if (l->waves_per_group != 0 && l->waves_per_group != 1 && l->waves_per_group != 2 && l->waves_per_group != 4 &&
      l->waves_per_group != 8 && l->waves_per_group != 16)
    return fail(c, BQP_E_INVALID, "waves_per_group must be 0 (auto), 1, 2, 4, 8 or 16");
  if (l->blocks < 0) return fail(c, BQP_E_INVALID, "blocks must be >= 0");
  if (l->use_graph != 0 && l->use_graph != 1) return fail(c, BQP_E_INVALID, "use_graph must be 0 or 1");
  if (l->sort_by_cost < 0 || l->sort_by_cost > 2) return fail(c, BQP_E_INVALID, "sort_by_cost must be 0, 1 or 2");
  if (l->teams_per_block != 0 && l->teams_per_block != 1 && l->teams_per_block != 2 && l->teams_per_block != 4)
    return fail(c, BQP_E_INVALID, "teams_per_block must be 0 (the fixed-G kernel), 1, 2 or 4");
  if (l->teams_per_block != 0 && l->waves_per_group != 0 &&
      !(l->waves_per_group == PAIR_WAVES && l->teams_per_block == PAIR_GROUPS))
    return fail(c, BQP_E_INVALID,
                "with teams_per_block > 0, waves_per_group must be 0 (the team kernel) or 8 with teams_per_block 2 "
                "(the pair kernel)");
  (void)hipStreamSynchronize(c->stream);
  bqp_drop_graph(c);
  c->launch = *l;
  return BQP_OK;
}

const char* bqp_last_error(const bqp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int32_t bqp_last_launch(const bqp_ctx* c, bqp_launch_info* out) {
  if (!c || !out) return BQP_E_INVALID;
  *out = c->last_launch;
  return BQP_OK;
}

int32_t bqp_get_cost_model(bqp_ctx* c, int64_t* evals, double* revolutions, int64_t* rows) {
  if (!c || !c->d_cost) return BQP_E_INVALID;
  std::vector<unsigned long long> e(COST_CELLS), r(COST_CELLS);
  std::vector<unsigned long long> k(COST_CELLS);
  const CostTable* t = (const CostTable*)c->d_cost;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(e.data(), t->evals, COST_CELLS * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(r.data(), t->revs, COST_CELLS * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(k.data(), t->rows, COST_CELLS * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int i = 0; i < COST_CELLS; ++i) {
    if (evals) evals[i] = (int64_t)e[i];
    if (revolutions) revolutions[i] = (double)r[i] / COST_REV_UNIT;
    if (rows) rows[i] = (int64_t)k[i];
  }
  return BQP_OK;
}

int32_t bqp_load_field(bqp_ctx* c, int32_t nmax, int32_t mmax, const double* cbar, const double* sbar,
                       double mu, double radius) {
  if (!c) return BQP_E_INVALID;
  if (!cbar || !sbar) return fail(c, BQP_E_INVALID, "cbar/sbar must not be NULL");
  if (nmax < 0 || nmax > 200 || mmax < 0 || mmax > nmax)
    return fail(c, BQP_E_INVALID, "need 0 <= mmax <= nmax <= 200");
  if (!(mu > 0.0) || !(radius > 0.0) || !std::isfinite(mu) || !std::isfinite(radius))
    return fail(c, BQP_E_INVALID, "mu and radius must be finite and > 0");
  const int nt = (nmax + 1) * (nmax + 2) / 2;
  for (int i = 0; i < nt; ++i)
    if (!std::isfinite(cbar[i]) || !std::isfinite(sbar[i]))
      return fail(c, BQP_E_INVALID, "non-finite coefficient");
  std::vector<GravEntry> tab;
  int jmax = 0;
  pack_field(nmax, mmax, cbar, sbar, mu, radius, tab, jmax);
  std::vector<GravEntryL> tabl;
  pack_lambda(nmax, jmax, tab, tabl);
  std::vector<GravEntryF> tabf(tab.size());
  for (size_t i = 0; i < tab.size(); ++i) {
    const GravEntry& e = tab[i];
    tabf[i] = GravEntryF{(float)e.a, (float)e.b, (float)e.xv, (float)e.xw,
                         (float)e.yv, (float)e.yw, (float)e.zv, (float)e.zw};
  }
  (void)hipStreamSynchronize(c->stream);
  bqp_drop_graph(c);
  bqp_mlp_release(c);  // a surrogate belongs to the field it was set up for
  (void)hipFree(c->d_tab);
  (void)hipFree(c->d_tabf);
  (void)hipFree(c->d_tabl);
  (void)hipFree(c->d_tablf);
  c->d_tab = nullptr;
  c->d_tabf = nullptr;
  c->d_tabl = nullptr;
  c->d_tablf = nullptr;
  c->nmax = -1;
  c->h_cbar.assign(cbar, cbar + nt);
  c->h_sbar.assign(sbar, sbar + nt);
  // zero entries of padding: the unrolled column loops load entries in
  // pairs, and the pipelined FP64 loops read up to two entries past the
  // last column
  tab.push_back(GravEntry{});
  tab.push_back(GravEntry{});
  tabf.push_back(GravEntryF{});
  tabf.push_back(GravEntryF{});
  for (int k = 0; k < 4; ++k) tabl.push_back(GravEntryL{});
  HIP_TRY(c, hipMalloc(&c->d_tab, tab.size() * sizeof(GravEntry)));
  HIP_TRY(c, hipMalloc(&c->d_tabf, tabf.size() * sizeof(GravEntryF)));
  HIP_TRY(c, hipMemcpy(c->d_tab, tab.data(), tab.size() * sizeof(GravEntry), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_tabf, tabf.data(), tabf.size() * sizeof(GravEntryF), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMalloc(&c->d_tabl, tabl.size() * sizeof(GravEntryL)));
  HIP_TRY(c, hipMemcpy(c->d_tabl, tabl.data(), tabl.size() * sizeof(GravEntryL), hipMemcpyHostToDevice));
  if (nmax >= PIPE_MIN_N && nmax <= SPLIT_RESCALED_MAX_N) {
    // FP32 copy of the rescaled table for the split-precision form; its
    // pipelined loops read up to four entries past the last column
    std::vector<GravEntryLF> tablf(tabl.size() + 4, GravEntryLF{});
    for (size_t i = 0; i < tabl.size(); ++i) {
      const GravEntryL& e = tabl[i];
      tablf[i] = GravEntryLF{(float)e.g, (float)e.xv, (float)e.xw, (float)e.yv, (float)e.yw, (float)e.zv, (float)e.zw};
    }
    HIP_TRY(c, hipMalloc(&c->d_tablf, tablf.size() * sizeof(GravEntryLF)));
    HIP_TRY(c, hipMemcpy(c->d_tablf, tablf.data(), tablf.size() * sizeof(GravEntryLF), hipMemcpyHostToDevice));
  }
  c->n_entries = tab.size();
  c->nmax = nmax;
  c->mmax = mmax;
  c->jmax = jmax;
  c->mu = mu;
  c->radius = radius;
  if (reset_cost_model(c) != BQP_OK) return fail(c, BQP_E_DEVICE, "cost model reset failed");
  return BQP_OK;
}

int32_t bqp_gravity_accel_batch_dev(bqp_ctx* c, const double* d_pos, int64_t n, double* d_acc) {
  if (!c) return BQP_E_INVALID;
  if (c->nmax < 0) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  if (n < 0 || (n > 0 && (!d_pos || !d_acc))) return fail(c, BQP_E_INVALID, "bad buffers");
  if (n == 0) return BQP_OK;
  const FieldView f = field_view(c);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (c->cfg.precision == BQP_PRECISION_SPLIT)
    hipLaunchKernelGGL(gravity_kernel<1>, dim3(blocks), dim3(256), 0, c->stream, f, d_pos, d_acc, (long long)n);
  else
    hipLaunchKernelGGL(gravity_kernel<0>, dim3(blocks), dim3(256), 0, c->stream, f, d_pos, d_acc, (long long)n);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BQP_OK;
}

int32_t bqp_gravity_accel_batch(bqp_ctx* c, const double* pos, int64_t n, double* acc) {
  if (!c) return BQP_E_INVALID;
  if (c->nmax < 0) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  if (n < 0 || (n > 0 && (!pos || !acc))) return fail(c, BQP_E_INVALID, "bad buffers");
  if (n == 0) return BQP_OK;
  const size_t bytes = (size_t)n * 3 * sizeof(double);
  int rc = ensure(c, &c->d_stage, &c->stage_bytes, 2 * bytes);
  if (rc) return rc;
  double* dp = (double*)c->d_stage;
  double* da = dp + 3 * n;
  HIP_TRY(c, hipMemcpyAsync(dp, pos, bytes, hipMemcpyHostToDevice, c->stream));
  rc = bqp_gravity_accel_batch_dev(c, dp, n, da);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpy(acc, da, bytes, hipMemcpyDeviceToHost));
  return BQP_OK;
}

int32_t bqp_propagate_batch_final_dev(bqp_ctx* c, const double* d_y0, const double* d_t0, const double* d_tf,
                                      const double* d_bc, int64_t n, double* d_yf, int32_t* d_status,
                                      int32_t* d_steps, int32_t* d_evals, bqp_stats* stats) {
  if (!c) return BQP_E_INVALID;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n < 0 || (n > 0 && (!d_y0 || !d_t0 || !d_tf || !d_yf)))
    return fail(c, BQP_E_INVALID, "y0, t0, tf and yf must not be NULL");
  if (n == 0) return BQP_OK;
  // per-row counters live in the staging area when the caller passes none
  const size_t cnt_bytes = (size_t)n * sizeof(int32_t);
  int rc = ensure(c, &c->d_stage, &c->stage_bytes, 4 * cnt_bytes);
  if (rc) return rc;
  int32_t* s0 = (int32_t*)c->d_stage;
  PropParams p{};
  p.y0 = d_y0;
  p.t0 = d_t0;
  p.tf = d_tf;
  p.bc = d_bc;
  p.yf = d_yf;
  p.status = d_status ? d_status : s0;
  p.steps = d_steps ? d_steps : s0 + n;
  p.evals = d_evals ? d_evals : s0 + 2 * n;
  p.rejects = s0 + 3 * n;
  return run_propagation(c, p, n, stats, nullptr);
}

int32_t bqp_propagate_batch_final_dev_noalloc(bqp_ctx* c, const double* d_y0, const double* d_t0,
                                              const double* d_tf, const double* d_bc, int64_t n, double* d_yf,
                                              int32_t* d_status, int32_t* d_steps, int32_t* d_evals,
                                              bqp_stats* stats) {
  if (!c) return BQP_E_INVALID;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n < 0 || (n > 0 && (!d_y0 || !d_t0 || !d_tf || !d_yf)))
    return fail(c, BQP_E_INVALID, "y0, t0, tf and yf must not be NULL");
  if (n > BQP_NOALLOC_MAX_ROWS)
    return fail(c, BQP_E_INVALID, "n = " + std::to_string(n) + " exceeds the no-allocation entry point's maximum of " +
                                      std::to_string(BQP_NOALLOC_MAX_ROWS) + " rows (use bqp_propagate_batch_final_dev)");
  if (n == 0) return BQP_OK;
  char* base = (char*)noalloc_base_known(c->device);
  if (!base) return fail(c, BQP_E_DEVICE, "the static scratch array's device address is unavailable");
  CallScratch sc;
  sc.queue = (unsigned int*)base;
  sc.totals = (unsigned long long*)(base + 256);
  sc.first = (int*)(base + 512);
  sc.sort = base + NOALLOC_SORT_OFF;
  int32_t* rows = (int32_t*)(base + NOALLOC_ROWS_OFF);
  PropParams p{};
  p.y0 = d_y0;
  p.t0 = d_t0;
  p.tf = d_tf;
  p.bc = d_bc;
  p.yf = d_yf;
  p.status = d_status ? d_status : rows;
  p.steps = d_steps ? d_steps : rows + n;
  p.evals = d_evals ? d_evals : rows + 2 * n;
  p.rejects = rows + 3 * n;
  return run_propagation(c, p, n, stats, &sc);
}

int32_t bqp_propagate_batch_final(bqp_ctx* c, const double* y0, const double* t0, const double* tf,
                                  const double* bc, int64_t n, double* yf, int32_t* row_status,
                                  int32_t* row_steps, int32_t* row_evals, bqp_stats* stats) {
  if (!c) return BQP_E_INVALID;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n < 0 || (n > 0 && (!y0 || !t0 || !tf || !yf)))
    return fail(c, BQP_E_INVALID, "y0, t0, tf and yf must not be NULL");
  if (n == 0) return BQP_OK;
  if (c->nmax < 0) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  // staging: y0 | t0 | tf | bc | yf | status | steps | evals | rejects
  const size_t d6 = (size_t)n * 6 * sizeof(double), d1 = (size_t)n * sizeof(double),
               i1 = (size_t)n * sizeof(int32_t);
  const size_t need = 2 * d6 + 3 * d1 + 4 * i1 + 256;
  int rc = ensure(c, &c->d_stage, &c->stage_bytes, need);
  if (rc) return rc;
  char* b = (char*)c->d_stage;
  double* dy0 = (double*)b;
  double* dt0 = (double*)(b + d6);
  double* dtf = (double*)(b + d6 + d1);
  double* dbc = (double*)(b + d6 + 2 * d1);
  double* dyf = (double*)(b + d6 + 3 * d1);
  int32_t* dst = (int32_t*)(b + 2 * d6 + 3 * d1);
  int32_t* dsp = dst + n;
  int32_t* dev = dsp + n;
  int32_t* drj = dev + n;
  HIP_TRY(c, hipMemcpyAsync(dy0, y0, d6, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dt0, t0, d1, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dtf, tf, d1, hipMemcpyHostToDevice, c->stream));
  if (bc) HIP_TRY(c, hipMemcpyAsync(dbc, bc, d1, hipMemcpyHostToDevice, c->stream));
  PropParams p{};
  p.y0 = dy0;
  p.t0 = dt0;
  p.tf = dtf;
  p.bc = bc ? dbc : nullptr;
  p.yf = dyf;
  p.status = dst;
  p.steps = dsp;
  p.evals = dev;
  p.rejects = drj;
  const int rc2 = run_propagation(c, p, n, stats, nullptr);
  if (rc2 != BQP_OK && rc2 != BQP_E_ROWS_FAILED) return rc2;
  HIP_TRY(c, hipMemcpy(yf, dyf, d6, hipMemcpyDeviceToHost));
  if (row_status) HIP_TRY(c, hipMemcpy(row_status, dst, i1, hipMemcpyDeviceToHost));
  if (row_steps) HIP_TRY(c, hipMemcpy(row_steps, dsp, i1, hipMemcpyDeviceToHost));
  if (row_evals) HIP_TRY(c, hipMemcpy(row_evals, dev, i1, hipMemcpyDeviceToHost));
  return rc2;
}

int32_t bqp_propagate_batch_sampled(bqp_ctx* c, const double* y0, double t0, const double* t_eval, int64_t m,
                                    const double* bc, int64_t n, double* out, int32_t* row_status,
                                    bqp_stats* stats) {
  if (!c) return BQP_E_INVALID;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n < 0 || m <= 0 || (n > 0 && (!y0 || !t_eval || !out)))
    return fail(c, BQP_E_INVALID, "need m >= 1 and non-NULL y0, t_eval, out");
  if (!std::isfinite(t0)) return fail(c, BQP_E_INVALID, "t0 must be finite");
  for (int64_t i = 0; i < m; ++i) {
    if (!std::isfinite(t_eval[i]) || t_eval[i] < t0 || (i > 0 && !(t_eval[i] > t_eval[i - 1])))
      return fail(c, BQP_E_INVALID, "t_eval must be finite, >= t0 and strictly increasing");
  }
  if (n == 0) return BQP_OK;
  if (c->nmax < 0) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  const size_t d6 = (size_t)n * 6 * sizeof(double), d1 = (size_t)n * sizeof(double),
               dm = (size_t)m * sizeof(double), ds = (size_t)n * m * 6 * sizeof(double),
               i1 = (size_t)n * sizeof(int32_t);
  const size_t need = d6 + d1 + dm + ds + 4 * i1 + 512;
  int rc = ensure(c, &c->d_stage, &c->stage_bytes, need);
  if (rc) return rc;
  char* b = (char*)c->d_stage;
  double* dy0 = (double*)b;
  double* dbc = (double*)(b + d6);
  double* dte = (double*)(b + d6 + d1);
  double* dout = (double*)(b + d6 + d1 + dm);
  int32_t* dst = (int32_t*)(b + d6 + d1 + dm + ds);
  int32_t* dsp = dst + n;
  int32_t* dev = dsp + n;
  int32_t* drj = dev + n;
  HIP_TRY(c, hipMemcpyAsync(dy0, y0, d6, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dte, t_eval, dm, hipMemcpyHostToDevice, c->stream));
  if (bc) HIP_TRY(c, hipMemcpyAsync(dbc, bc, d1, hipMemcpyHostToDevice, c->stream));
  PropParams p{};
  p.y0 = dy0;
  p.bc = bc ? dbc : nullptr;
  p.t_eval = dte;
  p.m_eval = (int)m;
  p.t0c = t0;
  p.ysamp = dout;
  p.status = dst;
  p.steps = dsp;
  p.evals = dev;
  p.rejects = drj;
  const int rc2 = run_propagation(c, p, n, stats, nullptr);
  if (rc2 != BQP_OK && rc2 != BQP_E_ROWS_FAILED) return rc2;
  HIP_TRY(c, hipMemcpy(out, dout, ds, hipMemcpyDeviceToHost));
  if (row_status) HIP_TRY(c, hipMemcpy(row_status, dst, i1, hipMemcpyDeviceToHost));
  return rc2;
}

int32_t bqp_get_tableau(int32_t method, bqp_tableau* out) {
  static const bqp_tableau tabs[3] = {BQP_DOPRI5_TABLEAU, BQP_DOP853_TABLEAU, BQP_RK87_TABLEAU};
  if (!out || method < 0 || method > 2) return BQP_E_INVALID;
  *out = tabs[method];
  return BQP_OK;
}

int32_t bqp_surrogate_dataset_dev(bqp_ctx* c, int64_t n, uint64_t seed, double* d_pos, float* d_feat,
                                  float* d_targ) {
  if (!c) return BQP_E_INVALID;
  bqp_mlp_state* s = c->mlp;
  if (!s || !s->d_tab_low) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (n < 0 || (n > 0 && (!d_feat || !d_targ))) return fail(c, BQP_E_INVALID, "bad buffers");
  if (n == 0) return BQP_OK;
  FieldView full = field_view(c);
  FieldView low = full;
  low.tab = nullptr;
  low.tabl = s->d_tab_low;
  low.tabf = nullptr;
  low.tablf = nullptr;
  low.nmax = std::min(s->nlow, c->nmax);
  low.jmax = s->low_jmax;
  const double rmid = 0.5 * (s->r_lo + s->r_hi), rhalf = 0.5 * (s->r_hi - s->r_lo);
  hipLaunchKernelGGL(surrogate_data_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, full, low,
                     (long long)n, (unsigned long long)seed, s->r_lo, s->r_hi, (float)rmid, (float)(1.0 / rhalf),
                     1.0 / s->out_scale[0], 1.0 / s->out_scale[1], 1.0 / s->out_scale[2], d_pos, d_feat, d_targ);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BQP_OK;
}

int32_t bqp_surrogate_init(bqp_ctx* c, int32_t nlow, double r_lo, double r_hi, uint64_t seed) {
  if (!c) return BQP_E_INVALID;
  if (c->nmax < 0) return fail(c, BQP_E_NO_FIELD, "no gravity field loaded (call bqp_load_field)");
  if (nlow < 1 || nlow > 8 || nlow >= c->nmax)
    return fail(c, BQP_E_INVALID, "need 1 <= nlow <= 8 and nlow < field degree");
  if (!(r_lo > 0.0) || !(r_hi > r_lo) || !std::isfinite(r_hi))
    return fail(c, BQP_E_INVALID, "need 0 < r_lo < r_hi (finite)");
  if (!c->mlp) {
    c->mlp = new (std::nothrow) bqp_mlp_state();
    if (!c->mlp) return fail(c, BQP_E_DEVICE, "out of host memory");
  }
  bqp_mlp_state* s = c->mlp;
  s->nlow = nlow;
  s->r_lo = r_lo;
  s->r_hi = r_hi;
  s->check_every = 100;
  s->threshold = 1e-9;
  std::vector<GravEntry> tab;
  std::vector<GravEntryL> tabl;
  int jmax = 0;
  pack_field(nlow, std::min(c->mmax, nlow), c->h_cbar.data(), c->h_sbar.data(), c->mu, c->radius, tab, jmax);
  pack_lambda(nlow, jmax, tab, tabl);
  for (int k = 0; k < 4; ++k) tabl.push_back(GravEntryL{});  // padding, as for the field table
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(s->d_tab_low);
  s->d_tab_low = nullptr;
  HIP_TRY(c, hipMalloc(&s->d_tab_low, tabl.size() * sizeof(GravEntryL)));
  HIP_TRY(c, hipMemcpy(s->d_tab_low, tabl.data(), tabl.size() * sizeof(GravEntryL), hipMemcpyHostToDevice));
  s->low_jmax = jmax;
  int rc = bqp_mlp_alloc_init(c, seed);
  if (rc) return rc;
  // output scales: per-component RMS of the residual over a seeded pilot sample
  for (int i = 0; i < 3; ++i) s->out_scale[i] = 1.0;
  const int64_t npilot = 65536;
  float* d_f = nullptr;
  float* d_t = nullptr;
  HIP_TRY(c, hipMalloc(&d_f, (size_t)npilot * MLP_IN * sizeof(float)));
  if (hipMalloc(&d_t, (size_t)npilot * 3 * sizeof(float)) != hipSuccess) {
    (void)hipFree(d_f);
    return fail(c, BQP_E_DEVICE, "hipMalloc failed");
  }
  rc = bqp_surrogate_dataset_dev(c, npilot, mix64(seed ^ 0x9170ull), nullptr, d_f, d_t);
  std::vector<float> t((size_t)npilot * 3);
  if (rc == BQP_OK && hipMemcpy(t.data(), d_t, t.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(c, BQP_E_DEVICE, "copy of the pilot sample failed");
  (void)hipFree(d_f);
  (void)hipFree(d_t);
  if (rc) return rc;
  for (int k = 0; k < 3; ++k) {
    double ss = 0.0;
    for (int64_t i = 0; i < npilot; ++i) ss += (double)t[3 * i + k] * (double)t[3 * i + k];
    const double rms = std::sqrt(ss / (double)npilot);
    s->out_scale[k] = (rms > 0.0 && std::isfinite(rms)) ? rms : 1.0;
  }
  return BQP_OK;
}

// Algorithmic FLOPs of one RHS evaluation: the gravity count (grav_flops)
// plus rotation 14 and, when enabled, drag 30.
double bqp_flops_per_rhs(const bqp_ctx* c) {
  if (!c || c->nmax < 0) return 0.0;
  // every form runs the factored sweep except the split form's (a, b) table
  // (fields outside the degree range of its rescaled FP32 copy)
  const bool factored = c->cfg.precision != BQP_PRECISION_SPLIT || c->d_tablf != nullptr;
  double f = grav_flops(c->nmax, c->jmax, factored);
  f += 14.0;
  if (c->cfg.drag != BQP_DRAG_NONE) f += 30.0;
  return f;
}

}  // extern "C"
