/*
 * libhexdyn.so C ABI (declared in include/hexdyn.h): handle lifetime, entity
 * renumbering and row layout, state transfer, phase dispatch and the
 * publish/validation semantics of one large step.
 *
 * Reference boundary: tree/src/hexcore/cuda_driver.py CudaDryDycoreDriver
 * (from_host :2826-3025, _step_device_v841 :4854-5537, step_device :5751,
 * step :6112-6149).  A step never mutates the published state: it works on
 * working copies and publishes by swapping buffers only when the sticky
 * validation flag stayed clear (cuda_driver.py:5355-5403, :5407-5427).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hexdyn.h"
#include "hexdyn_internal.h"
#include "hexdyn_powf.h"

using namespace hexdyn;

namespace {

#define HX_CHECK(expr)                                                              \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      if (h) h->error = std::string(#expr) + ": " + hipGetErrorString(_e);          \
      return HEXDYN_E_HIP;                                                          \
    }                                                                               \
  } while (0)

struct Schedule {
  double large[3], acoustic[3];
  int steps[3];
};

// integration.py:110-160 (RKSchedule.from_mpas)
Schedule rk_schedule(double dt, int order, int substeps, int splits) {
  Schedule s;
  const double dt_dyn = dt / splits;
  if (order == 3) {
    s.large[0] = dt_dyn / 3.0; s.large[1] = dt_dyn / 2.0; s.large[2] = dt_dyn;
    s.acoustic[0] = dt_dyn / 3.0; s.acoustic[1] = dt_dyn / substeps; s.acoustic[2] = dt_dyn / substeps;
    s.steps[0] = 1; s.steps[1] = std::max(1, substeps / 2); s.steps[2] = substeps;
  } else {
    s.large[0] = dt_dyn / 2.0; s.large[1] = dt_dyn / 2.0; s.large[2] = dt_dyn;
    for (int i = 0; i < 3; ++i) s.acoustic[i] = dt_dyn / substeps;
    s.steps[0] = std::max(1, substeps / 2); s.steps[1] = std::max(1, substeps / 2); s.steps[2] = substeps;
  }
  return s;
}

// 3-D Hilbert key (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004).
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

uint64_t point_key(double x, double y, double z) {
  const double r = std::sqrt(x * x + y * y + z * z);
  const int bits = 16;
  const double s = (double)((1u << bits) - 1);
  auto q = [&](double v) { return (uint32_t)std::lround((v / r * 0.5 + 0.5) * s); };
  return hilbert3(q(x), q(y), q(z), bits);
}

std::vector<int> order_by_keys(const std::vector<uint64_t> &keys) {
  std::vector<int> idx(keys.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
  return idx;  // idx[new] = old
}

// logical [level][entity] <-> rows of rs floats in curve order (levels
// nlev .. rs-1 of a row are zero)
__global__ void k_import(float *__restrict__ dst, const float *__restrict__ src, const int *__restrict__ old_of_new,
                         int n, int nlev, int rs) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = threadIdx.x & 63;
  if (e >= n || k >= rs) return;
  dst[(size_t)e * rs + k] = k < nlev ? src[(size_t)k * n + old_of_new[e]] : 0.0f;
}

__global__ void k_export(float *__restrict__ dst, const float *__restrict__ src, const int *__restrict__ old_of_new,
                         int n, int nlev, int rs) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = threadIdx.x & 63;
  if (e >= n || k >= nlev) return;
  dst[(size_t)k * n + old_of_new[e]] = src[(size_t)e * rs + k];
}

struct ProfEntry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

}  // namespace

namespace hexdyn {
int g_prof_on = 0;
static std::map<std::string, ProfEntry> g_prof;
static std::map<std::string, hipEvent_t> g_prof_open;

void prof_mark(const char *name, hipStream_t s, int start) {
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  (void)hipEventRecord(ev, s);
  if (start) {
    g_prof_open[name] = ev;
  } else {
    auto it = g_prof_open.find(name);
    if (it == g_prof_open.end()) {
      (void)hipEventDestroy(ev);
      return;
    }
    g_prof[name].events.push_back({it->second, ev});
    g_prof_open.erase(it);
  }
}
}  // namespace hexdyn

struct hexdyn_handle {
  Instance I;
  hexdyn_config cfg;
  Schedule dyn, sca;
  int fast = 0;
  int in_step = 0;
  int cur_sub = 1;
  // Buffer roles within a step (no copies): the first subcycle starts from
  // the published buffers, later subcycles from the previous subcycle's
  // current buffers (s_* and c_* swap roles); until a subcycle's first
  // recovery the current fields are read where the start fields live, and
  // until the step's first recovery u and w are read from the published
  // buffers.  exner and p' change only in a stage-3 recovery: until then
  // exner reads as the subcycle-start exner and p' as the published p'.
  // The published state is read-only during a step.
  int start_is_pub = 0;
  int fresh_sub = 0;
  int fresh_step = 0;
  int exner_start = 0;
  int pp_pub = 0;
  double time_s = 0.0;
  std::string error;
  std::vector<void *> allocs;
  int *old_c = nullptr, *old_e = nullptr, *old_v = nullptr;  // device: old index of each new entity
  double *int_partial = nullptr, *int_out = nullptr;          // integral partials and results
  int device = 0;                                             // the HIP device current at create
  LaunchTune tune;                                            // gather kernels' wavefronts per block
  // hipGraph-captured large steps, one executable per buffer-role layout at step start
  int use_graphs = 1;
  int post_clamp = 1;  // end-of-step q < 0 -> 0 (0: scalars handed over unclamped)
  hipStream_t capture_stream = nullptr;
  Side side{};  // the whole step's second stream and its fork/join events (created with the first step)
  struct Roles {  // host-side buffer roles and step flags
    Instance I;
    int in_step, cur_sub, start_is_pub, fresh_sub, fresh_step, exner_start, pp_pub;
  };
  std::map<std::string, std::pair<hipGraphExec_t, Roles>> graphs;
  Roles roles() const { return {I, in_step, cur_sub, start_is_pub, fresh_sub, fresh_step, exner_start, pp_pub}; }
  void set_roles(const Roles &r) {
    I = r.I; in_step = r.in_step; cur_sub = r.cur_sub; start_is_pub = r.start_is_pub; fresh_sub = r.fresh_sub;
    fresh_step = r.fresh_step; exner_start = r.exner_start; pp_pub = r.pp_pub;
  }
  std::vector<int> old_c_h, old_e_h, old_v_h;

  template <class T> int alloc(T **p, size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      error = "hipMalloc failed";
      return HEXDYN_E_HIP;
    }
    if (hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return HEXDYN_E_HIP;
    allocs.push_back(q);
    *p = (T *)q;
    return 0;
  }
  template <class T> int upload(const T **p, const std::vector<T> &h) {
    T *q = nullptr;
    int rc = alloc(&q, h.size());
    if (rc) return rc;
    if (hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return HEXDYN_E_HIP;
    *p = q;
    return 0;
  }
};

namespace {

// Bitmap over the binary32 arguments [HEXDYN_POW_LO, HEXDYN_POW_HI) of the
// ones where the host C library's powf(x, y) (the GNU powf the reference's
// v8.4.1 exner selects, cuda_driver.py:1466-1472) is not the correctly rounded
// power, i.e. differs from hexdyn_powf_eval.  Built once per process, in
// word-aligned slices over the host's cores; every marked argument is checked
// to be the neighbour hexdyn_gnu_powf_fix will choose.
const std::vector<uint32_t> &gnu_powf_ties(float y, std::string *failure, double *near_tie = nullptr) {
  static std::once_flag once;
  static std::vector<uint32_t> bits;
  static std::string fail;
  static float built_for = 0.0f;
  static double nearest = 0.5;  // smallest |d - r| in ulps over the marked arguments
  std::call_once(once, [y] {
    built_for = y;
    const uint32_t n = HEXDYN_POW_HI - HEXDYN_POW_LO;
    const uint32_t words = (n + 31u) / 32u;
    bits.assign(words, 0u);
    const unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::string> errs(nt);
    std::vector<double> near(nt, 0.5);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        const uint32_t w0 = (uint32_t)((uint64_t)words * t / nt), w1 = (uint32_t)((uint64_t)words * (t + 1) / nt);
        for (uint32_t w = w0; w < w1; ++w) {
          uint32_t acc = 0;
          for (uint32_t b = 0; b < 32u; ++b) {
            const uint32_t i = w * 32u + b;
            if (i >= n) break;
            const float x = hexdyn_f32_from_bits(HEXDYN_POW_LO + i);
            const double d = hexdyn_pow_d(x, y);
            const float r = (float)d;
            const float g = ::powf(x, y);
            if (hexdyn_f32_bits(g) == hexdyn_f32_bits(r)) continue;
            acc |= 1u << b;
            const uint32_t ir = hexdyn_f32_bits(r);
            const uint32_t fixed = d > (double)r ? ir + 1u : ir - 1u;  // hexdyn_gnu_powf_fix's choice
            const double ulp = (double)hexdyn_f32_from_bits(ir + 1u) - (double)r;
            near[t] = std::min(near[t], std::fabs(d - (double)r) / ulp);
            if (fixed != hexdyn_f32_bits(g) && errs[t].empty()) {
              char msg[160];
              std::snprintf(msg, sizeof msg, "host powf(%a, %a) = %a is not the neighbour of %a the table encodes",
                            (double)x, (double)y, (double)g, (double)r);
              errs[t] = msg;
            }
          }
          bits[w] = acc;
        }
      });
    for (auto &th : pool) th.join();
    for (double v : near) nearest = std::min(nearest, v);
    nearest = nearest * (1.0 - 1e-6);  // margin for the ulp and difference roundings
    for (const auto &e : errs)
      if (!e.empty()) {
        fail = e;
        break;
      }
  });
  if (built_for != y && fail.empty()) fail = "the GNU powf table serves one exponent per process";
  if (failure) *failure = fail;
  if (near_tie) *near_tie = nearest;
  return bits;
}

int run_phase_begin_step(hexdyn_handle *h, hipStream_t s, bool fused);

// The device pointers the kernels of the current phase see (buffer roles above).
Instance view(const hexdyn_handle *h, bool recovery) {
  Instance J = h->I;
  // the phase entry points keep every exportable working field current;
  // the whole-step sequence (step_view) stores only what a later kernel reads
  J.k.store_flux = 1;
  Work &B = J.b;
  if (h->start_is_pub) {
    B.s_rho = B.rho; B.s_rt = B.rt; B.s_ru = B.ru; B.s_rw = B.rw;
    B.s_theta = B.p_theta; B.s_exner = B.p_exner; B.s_rho_p = B.p_rho_p; B.s_rtheta_p = B.p_rtheta_p;
  }
  if (!recovery && h->fresh_sub) {
    B.c_rho = B.s_rho; B.c_rt = B.s_rt; B.c_ru = B.s_ru; B.c_rw = B.s_rw;
    B.theta = B.s_theta; B.rho_p = B.s_rho_p; B.rtheta_p = B.s_rtheta_p;
  }
  if (!recovery && h->exner_start) B.exner = B.s_exner;
  if (!recovery && h->pp_pub) B.pp = B.p_pp;
  if (!recovery && h->fresh_step) {
    B.u = B.p_u;
    B.w = B.p_w;
  }
  return J;
}

// Subcycle start: a later subcycle starts from the current buffers, which
// become the start buffers; the spare start buffers take the next currents.
void rotate_subcycle(hexdyn_handle *h, int sub) {
  Work &B = h->I.b;
  if (sub > 1) {
    std::swap(B.s_rho, B.c_rho); std::swap(B.s_rt, B.c_rt); std::swap(B.s_ru, B.c_ru); std::swap(B.s_rw, B.c_rw);
    std::swap(B.s_theta, B.theta); std::swap(B.s_exner, B.exner); std::swap(B.s_rho_p, B.rho_p);
    std::swap(B.s_rtheta_p, B.rtheta_p);
    h->start_is_pub = 0;
  } else {
    h->start_is_pub = 1;
  }
  h->fresh_sub = 1;
  h->exner_start = 1;
}

// after a recovery of `stage` the fields it wrote are the current ones
void recovered(hexdyn_handle *h, int stage) {
  h->fresh_sub = 0;
  h->fresh_step = 0;
  if (stage == 3) h->exner_start = h->pp_pub = 0;
}

#define NS_CALL(h, fn, ...) ((h)->fast ? hexdyn::fast::fn(__VA_ARGS__) : hexdyn::strict::fn(__VA_ARGS__))

int set_error(hexdyn_handle *h, int rc, const std::string &msg) {
  if (h) h->error = msg;
  return rc;
}

}  // namespace

extern "C" {

int hexdyn_abi_version(void) { return HEXDYN_ABI_VERSION; }

float hexdyn_powf(float x, float y) {
  const float rcv = 287.0f / (1004.5f - 287.0f);
  const double d = hexdyn_pow_d(x, y);
  const float r = (float)d;
  if (y != rcv || !(x > 0.0f)) return r;
  std::string why;
  double near_tie = 0.0;
  const std::vector<uint32_t> &ties = gnu_powf_ties(rcv, &why, &near_tie);
  return why.empty() ? hexdyn_gnu_powf_fix(x, r, d, ties.data(), near_tie) : r;
}

namespace {
thread_local std::string t_create_error;  // why this thread's last hexdyn_create failed

// Scratch of hexdyn_integrals_noalloc: the per-block partial sums, held in a
// file-scope device array so that entry point allocates nothing.  Each GPU
// the library runs on holds its own copy (8 MiB) until the process ends.
// Host code never passes the array's name to a kernel (on the host it is
// only a stand-in for the symbol): hexdyn_create resolves the array's device
// address once per device with hipGetSymbolAddress and every later use goes
// through that address.
static __device__ __align__(256) double scratch[HEXDYN_NOALLOC_MAX_PARTIALS];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// Resolve this device's scratch address (device `dev` must be current).
static int resolve_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch_ptr[dev] = (double *)ptr;
  return 0;
}

static double *scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}
}

int hexdyn_last_create_error(char *buf, size_t len) {
  if (!buf || len == 0) return HEXDYN_E_INVALID;
  std::snprintf(buf, len, "%s", t_create_error.c_str());
  return 0;
}

int hexdyn_last_error(const hexdyn_handle *h, char *buf, size_t len) {
  if (!h || !buf || len == 0) return HEXDYN_E_INVALID;
  std::snprintf(buf, len, "%s", h->error.c_str());
  return 0;
}

double hexdyn_model_time(const hexdyn_handle *h) { return h ? h->time_s : 0.0; }

void hexdyn_destroy(hexdyn_handle *h) {
  if (!h) return;
  for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.second.first);
  if (h->capture_stream) (void)hipStreamDestroy(h->capture_stream);
  if (h->side.s) (void)hipStreamDestroy(h->side.s);
  if (h->side.fork) (void)hipEventDestroy(h->side.fork);
  if (h->side.join) (void)hipEventDestroy(h->side.join);
  for (void *p : h->allocs) (void)hipFree(p);
  delete h;
}

static int create_impl(const hexdyn_mesh *m, const hexdyn_vertical *vt, const hexdyn_config *c, hexdyn_handle **out);

int hexdyn_create(const hexdyn_mesh *m, const hexdyn_vertical *vt, const hexdyn_config *c, hexdyn_handle **out) {
  t_create_error.clear();
  if (!m || !vt || !c || !out) {
    t_create_error = "hexdyn_create: null argument";
    return HEXDYN_E_INVALID;
  }
  *out = nullptr;
  return create_impl(m, vt, c, out);
}

static int create_impl(const hexdyn_mesh *m, const hexdyn_vertical *vt, const hexdyn_config *c, hexdyn_handle **out) {
  hexdyn_handle *h = new hexdyn_handle();
  h->I.tune = &h->tune;
  h->cfg = *c;
  Sizes &n = h->I.n;
  n.C = m->n_cells; n.E = m->n_edges; n.V = m->n_vertices; n.L = vt->n_levels;
  n.ME = m->max_edges; n.ME2 = m->max_edges2; n.MA = m->max_adv; n.NT = c->n_scalars; n.NM = c->n_moist;
  auto refuse = [&](const std::string &msg) {
    t_create_error = msg;
    int rc = HEXDYN_E_REFUSED;
    hexdyn_destroy(h);
    return rc;
  };
  // a HIP failure during set-up: keep the reason readable after the handle is gone
  auto hip_fail = [&](const char *what) {
    t_create_error = std::string(what) + (h->error.empty() ? "" : ": " + h->error);
    hexdyn_destroy(h);
    return HEXDYN_E_HIP;
  };
  if (n.C <= 0 || n.E <= 0 || n.V <= 0) return refuse("empty mesh");
  if (n.L < 4 || n.L + 1 > HX_ROW) return refuse("n_levels must satisfy 4 <= L and L + 1 <= 64");
  n.RS = std::min(HX_ROW, (n.L + 1 + 7) / 8 * 8);
  if (n.ME < 6 || n.ME > HEXDYN_MAX_EDGES) return refuse("max_edges must be 6 or 7");
  // edge stencils are stored at the fixed width 2*maxEdges-2 (kernels unroll over it)
  if (m->max_edges2 > 2 * n.ME - 2 || m->max_adv > 2 * n.ME - 2)
    return refuse("edgesOnEdge / advCellsForEdge wider than 2*maxEdges-2");
  if (m->vertex_degree != 3) return refuse("vertex_degree must be 3");
  if (n.NT < 0 || n.NM < 0 || n.NM > n.NT) return refuse("n_moist must lie in [0, n_scalars]");
  {
    // kernels address every row as a 32-bit byte offset from its array's
    // base (row * 256 + level * 4): the largest row index of any buffer must
    // stay below 2^32 / 256 rows
    const int64_t me_max = std::max(m->max_edges, 6);
    const int64_t rows[] = {(int64_t)n.C, (int64_t)n.E, (int64_t)n.V, (int64_t)n.C * me_max,
                            (int64_t)std::max(n.NT, 1) * n.C, (int64_t)std::max(n.NT, 1) * n.E};
    for (int64_t r : rows)
      if (r * HX_ROW * (int64_t)sizeof(float) >= ((int64_t)1 << 32))
        return refuse("mesh too large for 32-bit row offsets: an array would reach 4 GiB "
                      "(rows x 256 B with rows = cells, edges, vertices, cells x maxEdges, "
                      "scalars x cells or scalars x edges)");
  }
  if (c->config_time_integration_order != 2 && c->config_time_integration_order != 3)
    return refuse("config_time_integration_order must be 2 or 3");
  if (c->config_dynamics_split_steps != 1 && c->config_dynamics_split_steps != 3)
    return refuse("config_dynamics_split_steps must be 1 or 3");
  if (c->config_dynamics_split_steps == 3 && !c->config_split_dynamics_transport)
    return refuse("config_dynamics_split_steps=3 requires config_split_dynamics_transport");
  if (c->config_number_of_sub_steps < 1) return refuse("config_number_of_sub_steps must be positive");
  if (!(c->config_dt > 0.0)) return refuse("config_dt must be positive");
  if (c->config_apvm_upwinding < 0.0f) return refuse("config_apvm_upwinding must be non-negative");
  if (c->config_divergence_damping && !(c->config_len_disp > 0.0f)) return refuse("config_len_disp must be positive");
  if (c->config_horiz_mixing != HEXDYN_MIXING_OFF && c->config_horiz_mixing != HEXDYN_MIXING_2D_SMAGORINSKY)
    return refuse("config_horiz_mixing must be off or 2d_smagorinsky");
  const bool mixing = c->config_horiz_mixing == HEXDYN_MIXING_2D_SMAGORINSKY;
  if (mixing) {
    if (!(c->config_len_disp > 0.0f)) return refuse("config_len_disp must be positive");
    if (!(c->config_visc4_2dsmag >= 0.0f) || !(c->config_smagorinsky_coef >= 0.0f))
      return refuse("config_visc4_2dsmag and config_smagorinsky_coef must be non-negative");
    if (!(c->config_del4u_div_factor > 0.0f)) return refuse("config_del4u_div_factor must be positive");
    if (!(c->config_mpas_cam_coef >= 0.0f)) return refuse("config_mpas_cam_coef must be non-negative");
    if (c->config_mpas_cam_coef > 0.0f &&
        (c->config_number_cam_damping_levels < 1 || c->config_number_cam_damping_levels > n.L))
      return refuse("config_number_cam_damping_levels must lie in 1..nVertLevels");
    if (!m->deformation_coef_c2 || !m->deformation_coef_s2 || !m->deformation_coef_cs)
      return refuse("2d_smagorinsky mixing needs the deformation weights");
    if (!m->area_cell || !m->area_triangle) return refuse("mixing needs areaCell and areaTriangle");
  }
  if (c->config_rayleigh_damp_u) {
    if (c->config_number_rayleigh_damp_u_levels < 1 || c->config_number_rayleigh_damp_u_levels > n.L)
      return refuse("config_number_rayleigh_damp_u_levels must lie in 1..nVertLevels");
    if (!(c->config_rayleigh_damp_u_timescale_days > 0.0f) || !std::isfinite(c->config_rayleigh_damp_u_timescale_days))
      return refuse("config_rayleigh_damp_u_timescale_days must be finite and positive");
  }
  h->fast = c->numerics == HEXDYN_NUMERICS_FAST;
  const int in_me2 = m->max_edges2, in_ma = m->max_adv;
  n.ME2 = n.MA = 2 * n.ME - 2;
  const int C = n.C, E = n.E, V = n.V, L = n.L, ME = n.ME, ME2 = n.ME2, MA = n.MA;

  // ---- renumbering along a 3-D Hilbert curve
  std::vector<uint64_t> kc(C), ke(E), kv(V);
  const bool have_xyz = m->x_cell && m->y_cell && m->z_cell;
  for (int i = 0; i < C; ++i) kc[i] = have_xyz ? point_key(m->x_cell[i], m->y_cell[i], m->z_cell[i]) : (uint64_t)i;
  for (int e = 0; e < E; ++e) {
    const int a = m->cells_on_edge[2 * e], b = m->cells_on_edge[2 * e + 1];
    if (a < 0 || b < 0 || a >= C || b >= C) return refuse("cellsOnEdge out of range");
    ke[e] = have_xyz ? point_key(m->x_cell[a] + m->x_cell[b], m->y_cell[a] + m->y_cell[b], m->z_cell[a] + m->z_cell[b])
                     : (uint64_t)e;
  }
  for (int v = 0; v < V; ++v) {
    double x = 0, y = 0, z = 0;
    for (int j = 0; j < 3; ++j) {
      const int cc = m->cells_on_vertex[3 * v + j];
      if (cc < 0 || cc >= C) return refuse("cellsOnVertex out of range");
      if (have_xyz) { x += m->x_cell[cc]; y += m->y_cell[cc]; z += m->z_cell[cc]; }
    }
    kv[v] = have_xyz ? point_key(x, y, z) : (uint64_t)v;
  }
  h->old_c_h = order_by_keys(kc);
  h->old_e_h = order_by_keys(ke);
  h->old_v_h = order_by_keys(kv);
  std::vector<int> new_c(C), new_e(E), new_v(V);
  for (int i = 0; i < C; ++i) new_c[h->old_c_h[i]] = i;
  for (int i = 0; i < E; ++i) new_e[h->old_e_h[i]] = i;
  for (int i = 0; i < V; ++i) new_v[h->old_v_h[i]] = i;
  auto chk = [](int x, int lim) { return x >= 0 && x < lim; };

  // ---- connectivity and geometry in new numbering
  std::vector<int> coe(2 * E), voe(2 * E), eoc((size_t)C * ME, 0), coc((size_t)C * ME, 0), voc((size_t)C * ME, 0),
      neoc(C), kfc((size_t)C * ME, 0), eov(3 * V), eoe((size_t)E * ME2, 0), neoe(E), nadv(E), advc((size_t)E * MA, 0);
  std::vector<float> csign((size_t)C * ME, 0.0f), vsign(3 * V), woe((size_t)E * ME2, 0.0f), dc(E), dv(E), dcdv(E),
      invdc(E), invdv(E), invarea(C), invtri(V), kite(3 * V), fedge(E), fvert(V), cosang(E), sinang(E),
      adv((size_t)E * MA, 0.0f), adv3((size_t)E * MA, 0.0f);
  for (int ne = 0; ne < E; ++ne) {
    const int e = h->old_e_h[ne];
    for (int j = 0; j < 2; ++j) {
      const int cc = m->cells_on_edge[2 * e + j], vv = m->vertices_on_edge[2 * e + j];
      if (!chk(vv, V)) return refuse("verticesOnEdge out of range");
      coe[2 * ne + j] = new_c[cc];
      voe[2 * ne + j] = new_v[vv];
    }
    dc[ne] = m->dc_edge[e]; dv[ne] = m->dv_edge[e]; dcdv[ne] = dc[ne] * dv[ne];
    invdc[ne] = m->inv_dc_edge[e]; invdv[ne] = m->inv_dv_edge[e];
    fedge[ne] = m->f_edge[e]; cosang[ne] = m->cos_angle_edge[e]; sinang[ne] = m->sin_angle_edge[e];
    neoe[ne] = m->n_edges_on_edge[e];
    if (neoe[ne] < 0 || neoe[ne] > in_me2) return refuse("nEdgesOnEdge out of range");
    for (int s = 0; s < neoe[ne]; ++s) {
      const int o = m->edges_on_edge[(size_t)e * in_me2 + s];
      if (!chk(o, E)) return refuse("edgesOnEdge out of range");
      eoe[(size_t)ne * ME2 + s] = new_e[o];
      woe[(size_t)ne * ME2 + s] = m->weights_on_edge[(size_t)e * in_me2 + s];
    }
    nadv[ne] = m->n_adv_cells_for_edge[e];
    if (nadv[ne] < 0 || nadv[ne] > in_ma) return refuse("nAdvCellsForEdge out of range");
    for (int s = 0; s < nadv[ne]; ++s) {
      const int o = m->adv_cells_for_edge[(size_t)e * in_ma + s];
      if (!chk(o, C)) return refuse("advCellsForEdge out of range");
      advc[(size_t)ne * MA + s] = new_c[o];
      adv[(size_t)ne * MA + s] = m->adv_coefs[(size_t)e * in_ma + s];
      adv3[(size_t)ne * MA + s] = m->adv_coefs_3rd[(size_t)e * in_ma + s];
    }
  }
  for (int nc = 0; nc < C; ++nc) {
    const int cc = h->old_c_h[nc];
    neoc[nc] = m->n_edges_on_cell[cc];
    if (neoc[nc] < 3 || neoc[nc] > ME) return refuse("nEdgesOnCell out of range");
    invarea[nc] = m->inv_area_cell[cc];
    for (int s = 0; s < neoc[nc]; ++s) {
      const size_t i = (size_t)cc * ME + s, o = (size_t)nc * ME + s;
      const int e = m->edges_on_cell[i], nb = m->cells_on_cell[i], vv = m->vertices_on_cell[i];
      if (!chk(e, E) || !chk(nb, C) || !chk(vv, V)) return refuse("cell connectivity out of range");
      eoc[o] = new_e[e]; coc[o] = new_c[nb]; voc[o] = new_v[vv];
      kfc[o] = m->kite_for_cell[i];
      if (kfc[o] < 0 || kfc[o] > 2) return refuse("kiteForCell out of range");
      csign[o] = m->cells_on_edge[2 * e] == cc ? 1.0f : -1.0f;
    }
  }
  for (int nv = 0; nv < V; ++nv) {
    const int vv = h->old_v_h[nv];
    invtri[nv] = m->inv_area_triangle[vv];
    fvert[nv] = m->f_vertex[vv];
    for (int j = 0; j < 3; ++j) {
      const int e = m->edges_on_vertex[3 * vv + j];
      if (!chk(e, E)) return refuse("edgesOnVertex out of range");
      eov[3 * nv + j] = new_e[e];
      vsign[3 * nv + j] = m->vertices_on_edge[2 * e + 1] == vv ? 1.0f : -1.0f;
      kite[3 * nv + j] = m->kite_areas_on_vertex[3 * vv + j];
    }
  }
  // per-slot copies of edge/vertex geometry
  std::vector<float> csdv((size_t)C * ME, 0.0f), cdcdv((size_t)C * ME, 0.0f), ckite((size_t)C * ME, 0.0f),
      vdc(3 * V), vdcdv(3 * V), ecos((size_t)E * ME2, 0.0f), esin((size_t)E * ME2, 0.0f);
  std::vector<int> ocell((size_t)C * ME, 0);
  int ocp = 1;
  for (int nc = 0; nc < C; ++nc)
    for (int s = 0; s < neoc[nc]; ++s) {
      const size_t o = (size_t)nc * ME + s;
      const int a0 = coe[2 * eoc[o]], a1 = coe[2 * eoc[o] + 1];
      ocell[o] = a0 == nc ? a1 : a0;
      if (ocell[o] != coc[o]) ocp = 0;
      csdv[o] = csign[o] * dv[eoc[o]];
      cdcdv[o] = dcdv[eoc[o]];
      ckite[o] = kite[3 * voc[o] + kfc[o]];
    }
  for (int nv = 0; nv < V; ++nv)
    for (int j = 0; j < 3; ++j) {
      vdc[3 * nv + j] = vsign[3 * nv + j] * dc[eov[3 * nv + j]];
      vdcdv[3 * nv + j] = dcdv[eov[3 * nv + j]];
    }
  for (int ne = 0; ne < E; ++ne)
    for (int s = 0; s < neoe[ne]; ++s) {
      const size_t o = (size_t)ne * ME2 + s;
      ecos[o] = cosang[eoe[o]];
      esin[o] = sinang[eoe[o]];
    }
  Mesh &M = h->I.m;
  int rc = 0;
  rc |= h->upload(&M.csdv, csdv); rc |= h->upload(&M.cdcdv, cdcdv); rc |= h->upload(&M.ckite, ckite);
  rc |= h->upload(&M.vdc, vdc); rc |= h->upload(&M.vdcdv, vdcdv);
  rc |= h->upload(&M.ecos, ecos); rc |= h->upload(&M.esin, esin); rc |= h->upload(&M.ocell, ocell);
  n.OCP = ocp;
  if (mixing) {
    // binary32 scalars staged as the reference host does (cuda_horizontal_v841.py:1336-1350)
    const float len = c->config_len_disp, csm = c->config_smagorinsky_coef;
    const float h4 = c->config_visc4_2dsmag * ((len * len) * len);
    std::vector<float> del2(E, 1.0f), del4(E, 1.0f);
    if (c->config_h_ScaleWithMesh) {
      for (int ne = 0; ne < E; ++ne) {
        const int e = h->old_e_h[ne];
        const int a0 = m->cells_on_edge[2 * e], a1 = m->cells_on_edge[2 * e + 1];
        const float d0 = m->mesh_density ? m->mesh_density[a0] : 1.0f;
        const float d1 = m->mesh_density ? m->mesh_density[a1] : 1.0f;
        const float mean = 0.5f * (d0 + d1);
        if (!(mean > 0.0f)) return refuse("meshDensity must be positive");
        // the reference's mesh-scaling power: binary64 pow of the binary32
        // mean, rounded once (mixing.py:313-318 through rkind_libm.powf_rkind)
        del2[ne] = 1.0f / (float)std::pow((double)mean, (double)0.25f);
        del4[ne] = 1.0f / (float)std::pow((double)mean, (double)0.75f);
      }
    }
    std::vector<float> c2s((size_t)C * ME, 0.0f), s2s((size_t)C * ME, 0.0f), css((size_t)C * ME, 0.0f),
        fdiv((size_t)C * ME, 0.0f), fth2((size_t)C * ME, 0.0f), fw2((size_t)C * ME, 0.0f), f4((size_t)C * ME, 0.0f),
        fvort(3 * (size_t)V, 0.0f), idc(E), idv(E), rdc(E), rdv(E);
    for (int nc = 0; nc < C; ++nc) {
      const int cc = h->old_c_h[nc];
      const float ia = 1.0f / m->area_cell[cc];
      for (int s = 0; s < neoc[nc]; ++s) {
        const size_t i = (size_t)cc * ME + s, o = (size_t)nc * ME + s;
        const int e = m->edges_on_cell[i];
        const float sg = m->cells_on_edge[2 * e] == cc ? 1.0f : -1.0f;
        const float dve = m->dv_edge[e], idce = 1.0f / m->dc_edge[e];
        c2s[o] = m->deformation_coef_c2[i];
        s2s[o] = m->deformation_coef_s2[i];
        css[o] = m->deformation_coef_cs[i];
        fdiv[o] = (ia * dve) * sg;
        fth2[o] = ((ia * sg) * dve) * idce;
        fw2[o] = (((0.5f * ia) * sg) * dve) * idce;
        f4[o] = (((del4[new_e[e]] * (h4 * ia)) * dve) * sg) * idce;
      }
    }
    for (int nv = 0; nv < V; ++nv) {
      const int vv = h->old_v_h[nv];
      const float it = 1.0f / m->area_triangle[vv];
      for (int j = 0; j < 3; ++j) {
        const int e = m->edges_on_vertex[3 * vv + j];
        const float vs = m->vertices_on_edge[2 * e] == vv ? -1.0f : 1.0f;
        fvort[3 * nv + j] = (it * m->dc_edge[e]) * vs;
      }
    }
    for (int ne = 0; ne < E; ++ne) {
      const int e = h->old_e_h[ne];
      idc[ne] = 1.0f / m->dc_edge[e];
      idv[ne] = std::fmin(1.0f / m->dv_edge[e], 4.0f * idc[ne]);
      const float scale = del4[ne] * h4;
      rdc[ne] = (scale * c->config_del4u_div_factor) * idc[ne];
      rdv[ne] = scale * idv[ne];
    }
    rc |= h->upload(&M.defc2, c2s); rc |= h->upload(&M.defs2, s2s); rc |= h->upload(&M.defcs, css);
    rc |= h->upload(&M.mx_fdiv, fdiv); rc |= h->upload(&M.mx_fth2, fth2); rc |= h->upload(&M.mx_fw2, fw2);
    rc |= h->upload(&M.mx_f4, f4); rc |= h->upload(&M.mx_fvort, fvort);
    rc |= h->upload(&M.mx_idc, idc); rc |= h->upload(&M.mx_idv, idv); rc |= h->upload(&M.mx_rdc, rdc);
    rc |= h->upload(&M.mx_rdv, rdv); rc |= h->upload(&M.mx_del2, del2);
    Consts &K0 = h->I.k;
    K0.mixing = 1;
    K0.mix4 = h4 > 0.0f;
    K0.strain_scale = (csm * len) * (csm * len);
    const float inv_dt = 1.0f / (float)c->config_dt;
    K0.ceiling = (0.01f * (len * len)) * inv_dt;
    if (c->config_mpas_cam_coef > 0.0f) {
      K0.cam_levels = c->config_number_cam_damping_levels;
      K0.cam_base = ((4.0f * 2.0833f) * len) * c->config_mpas_cam_coef;
    }
  }
  rc |= h->upload(&M.coe, coe); rc |= h->upload(&M.voe, voe); rc |= h->upload(&M.eoc, eoc);
  rc |= h->upload(&M.coc, coc); rc |= h->upload(&M.voc, voc); rc |= h->upload(&M.neoc, neoc);
  rc |= h->upload(&M.kfc, kfc); rc |= h->upload(&M.eov, eov); rc |= h->upload(&M.eoe, eoe);
  rc |= h->upload(&M.neoe, neoe); rc |= h->upload(&M.nadv, nadv); rc |= h->upload(&M.advc, advc);
  rc |= h->upload(&M.csign, csign); rc |= h->upload(&M.vsign, vsign); rc |= h->upload(&M.woe, woe);
  rc |= h->upload(&M.dc, dc); rc |= h->upload(&M.dv, dv); rc |= h->upload(&M.dcdv, dcdv);
  rc |= h->upload(&M.invdc, invdc); rc |= h->upload(&M.invdv, invdv); rc |= h->upload(&M.invarea, invarea);
  rc |= h->upload(&M.invtri, invtri); rc |= h->upload(&M.kite, kite); rc |= h->upload(&M.fedge, fedge);
  rc |= h->upload(&M.fvert, fvert); rc |= h->upload(&M.cosang, cosang); rc |= h->upload(&M.sinang, sinang);
  rc |= h->upload(&M.adv, adv); rc |= h->upload(&M.adv3, adv3);
  {
    // edge tiles of the staged flux gathers: distinct advection cells in
    // order of first use, and each slot's index among them
    const int ntiles = (E + HX_TILE_E - 1) / HX_TILE_E;
    std::vector<int> tcell((size_t)ntiles * HX_TILE_U, 0), ntcell(ntiles, 0), advl((size_t)E * MA, 0);
    int tu = 0;
    bool fits = true;
    for (int t = 0; t < ntiles && fits; ++t) {
      int cnt = 0;
      int *row = &tcell[(size_t)t * HX_TILE_U];
      for (int ne = t * HX_TILE_E; ne < std::min(E, (t + 1) * HX_TILE_E) && fits; ++ne)
        for (int sl = 0; sl < MA; ++sl) {
          const int cell = advc[(size_t)ne * MA + sl];  // padded slots hold cell 0 and are staged too
          int j = 0;
          while (j < cnt && row[j] != cell) ++j;
          if (j == cnt) {
            if (cnt == HX_TILE_U) {
              fits = false;
              break;
            }
            row[cnt++] = cell;
          }
          advl[(size_t)ne * MA + sl] = j;
        }
      for (int j = cnt; j < HX_TILE_U; ++j) row[j] = row[0];
      ntcell[t] = cnt;
      tu = std::max(tu, cnt);
    }
    n.TU = fits ? tu : 0;
    rc |= h->upload(&M.tcell, tcell);
    rc |= h->upload(&M.ntcell, ntcell);
    rc |= h->upload(&M.advl, advl);
  }
  {
    std::vector<int> a(h->old_c_h), b(h->old_e_h), cvec(h->old_v_h);
    const int *pa, *pb, *pc;
    rc |= h->upload(&pa, a); rc |= h->upload(&pb, b); rc |= h->upload(&pc, cvec);
    h->old_c = (int *)pa; h->old_e = (int *)pb; h->old_v = (int *)pc;
  }
  if (rc) return hip_fail("device allocation or upload failed");

  // ---- vertical metrics, padded per-level arrays and row-layout static fields
  Vert &Z = h->I.z;
  auto pad = [&](const float *src, int count) {
    std::vector<float> v(HX_ROW, 0.0f);
    for (int i = 0; i < count; ++i) v[i] = src[i];
    return v;
  };
  const int RS = n.RS;
  auto rows_cells = [&](const float *src, int nlev) {  // [nlev][C] old -> [C][RS] new
    std::vector<float> v((size_t)C * HX_ROW, 0.0f);
    for (int nc = 0; nc < C; ++nc)
      for (int k = 0; k < nlev; ++k) v[(size_t)nc * RS + k] = src[(size_t)k * C + h->old_c_h[nc]];
    return v;
  };
  auto rows_edges = [&](const float *src, int nlev) {
    std::vector<float> v((size_t)E * HX_ROW, 0.0f);
    for (int ne = 0; ne < E; ++ne)
      for (int k = 0; k < nlev; ++k) v[(size_t)ne * RS + k] = src[(size_t)k * E + h->old_e_h[ne]];
    return v;
  };
  rc |= h->upload(&Z.rdzw, pad(vt->rdzw, L)); rc |= h->upload(&Z.rdzu, pad(vt->rdzu, L));
  rc |= h->upload(&Z.fzm, pad(vt->fzm, L)); rc |= h->upload(&Z.fzp, pad(vt->fzp, L));
  rc |= h->upload(&Z.etp, pad(vt->etp, L)); rc |= h->upload(&Z.etm, pad(vt->etm, L));
  rc |= h->upload(&Z.ewp, pad(vt->ewp, L + 1)); rc |= h->upload(&Z.ewm, pad(vt->ewm, L + 1));
  rc |= h->upload(&Z.uinit, pad(vt->u_init, L)); rc |= h->upload(&Z.vinit, pad(vt->v_init, L));
  Z.cf1 = vt->cf1; Z.cf2 = vt->cf2; Z.cf3 = vt->cf3;
  rc |= h->upload(&Z.zz, rows_cells(vt->zz, L)); rc |= h->upload(&Z.dss, rows_cells(vt->dss, L));
  rc |= h->upload(&Z.rb, rows_cells(vt->rho_base, L)); rc |= h->upload(&Z.rtb, rows_cells(vt->rtheta_base, L));
  rc |= h->upload(&Z.exb, rows_cells(vt->exner_base, L));
  rc |= h->upload(&Z.zxu, rows_edges(vt->zxu, L));
  if (vt->zgrid) {
    std::vector<float> zm((size_t)C * HX_ROW, 0.0f);
    for (int nc = 0; nc < C; ++nc)
      for (int k = 0; k < L; ++k) {
        const int cc = h->old_c_h[nc];
        zm[(size_t)nc * RS + k] =
            (float)(0.5 * ((double)vt->zgrid[(size_t)k * C + cc] + (double)vt->zgrid[(size_t)(k + 1) * C + cc]));
      }
    rc |= h->upload(&Z.zmid, zm);
  }
  {
    std::vector<float> zb((size_t)C * ME * HX_ROW, 0.0f), zb3((size_t)C * ME * HX_ROW, 0.0f);
    for (int nc = 0; nc < C; ++nc) {
      const int cc = h->old_c_h[nc];
      for (int s = 0; s < ME; ++s)
        for (int k = 0; k <= L; ++k) {
          const size_t src = ((size_t)k * C + cc) * ME + s;
          const size_t dst = ((size_t)nc * ME + s) * RS + k;
          zb[dst] = vt->zb_cell[src];
          zb3[dst] = vt->zb3_cell[src];
        }
    }
    rc |= h->upload(&Z.zb, zb); rc |= h->upload(&Z.zb3, zb3);
  }
  if (rc) return hip_fail("device allocation or upload failed");

  // ---- working set
  Work &B = h->I.b;
  const size_t cr = (size_t)C * HX_ROW, er = (size_t)E * HX_ROW, vr = (size_t)V * HX_ROW;
  const size_t sr = (size_t)std::max(n.NT, 1) * cr, ser = (size_t)std::max(n.NT, 1) * er;
  float **cell_bufs[] = {&B.rho, &B.rt, &B.rw, &B.theta, &B.exner, &B.rho_p, &B.rtheta_p, &B.pp, &B.w,
                         &B.p_theta, &B.p_exner, &B.p_rho_p, &B.p_rtheta_p, &B.p_pp, &B.p_w,
                         &B.c_rho, &B.c_rt, &B.c_rw, &B.s_rho, &B.s_rt, &B.s_rw, &B.s_theta, &B.s_rho_p,
                         &B.s_rtheta_p, &B.s_exner, &B.qtot, &B.cqw, &B.cofwz, &B.cofwt, &B.coftz,
                         &B.atri, &B.alpha, &B.gamma, &B.dpdz, &B.euler_rw, &B.tend_rho, &B.tend_rt, &B.tend_rw,
                         &B.hdiv, &B.rw_p, &B.ww_avg, &B.rho_pp, &B.rtheta_pp, &B.rtheta_pp_old, &B.ke,
                         &B.divergence, &B.pv_cell, &B.flux_w, &B.split_w, &B.target};
  for (float **p : cell_bufs) rc |= h->alloc(p, cr);
  float **edge_bufs[] = {&B.ru, &B.u, &B.p_u, &B.c_ru, &B.s_ru, &B.cqu, &B.euler_ru, &B.tend_ru, &B.fth,
                         &B.fcorr, &B.fw, &B.ru_p, &B.ru_avg, &B.v, &B.pv_edge, &B.flux_u, &B.split_u};
  for (float **p : edge_bufs) rc |= h->alloc(p, er);
  float **vert_bufs[] = {&B.vorticity, &B.ke_vertex, &B.pv_vertex};
  for (float **p : vert_bufs) rc |= h->alloc(p, vr);
  float **scal_bufs[] = {&B.q, &B.q_stage, &B.q_next, &B.fct_mn, &B.fct_mx, &B.fct_mass, &B.fct_sin,
                         &B.fct_sout, &B.fct_vres};
  for (float **p : scal_bufs) rc |= h->alloc(p, sr);
  float **sedge_bufs[] = {&B.ev, &B.fct_hres};
  for (float **p : sedge_bufs) rc |= h->alloc(p, ser);
  if (mixing) {
    float **mix_cell[] = {&B.kdiff, &B.delsq_div, &B.delsq_theta, &B.delsq_w, &B.mix_tw, &B.euler_theta};
    for (float **p : mix_cell) rc |= h->alloc(p, cr);
    rc |= h->alloc(&B.delsq_u, er);
    rc |= h->alloc(&B.delsq_vort, vr);
  }
  rc |= h->alloc(&B.flag, 16);
  rc |= h->alloc(&h->int_partial, (size_t)((C + 3) / 4) * (3 + n.NT));
  rc |= h->alloc(&h->int_out, (size_t)(3 + n.NT));
  if (rc) return hip_fail("device allocation or upload failed");
  if (hipGetDevice(&h->device) != hipSuccess || resolve_scratch(h->device) != 0)
    return hip_fail("resolving the device scratch of hexdyn_integrals_noalloc failed");
  if (n.NM == 0) {
    std::vector<float> ones(std::max(cr, er), 1.0f);
    if (hipMemcpy(B.cqw, ones.data(), cr * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B.cqu, ones.data(), er * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return hip_fail("dry moist-coefficient upload failed");
  }

  // ---- constants in the reference's binary32 forms
  Consts &K = h->I.k;
  K.gravity = (float)9.80616;
  K.rcv = 287.0f / (1004.5f - 287.0f);
  K.c2 = 1004.5f * K.rcv;
  K.cfac = (float)(287.0 / 100000.0);
  K.rgas = 287.0f;
  if (K.rcv != (float)(287.0 / (1004.5 - 287.0))) return refuse("inconsistent binary32 R/cv");
  {
    std::string why;
    const std::vector<uint32_t> &ties = gnu_powf_ties(K.rcv, &why, &K.powf_near_tie);
    if (!why.empty()) return refuse(why.c_str());
    if (h->upload(&h->I.m.powf_ties, ties)) return hip_fail("powf correction table upload failed");
  }
  if (c->config_rayleigh_damp_u) {
    const int nr = c->config_number_rayleigh_damp_u_levels;
    K.ray_levels = nr;
    K.ray_inverse = 1.0f / ((float)nr * (c->config_rayleigh_damp_u_timescale_days * 86400.0f));
  }
  K.coef3 = c->config_coef_3rd_order;
  K.apvm_on = c->config_apvm_upwinding > 0.0f;
  K.apvm_scale = c->config_apvm_upwinding * (float)c->config_dt;
  K.divdamp_base = (2.0f * c->config_smdiv) * c->config_len_disp;
  h->dyn = rk_schedule(c->config_dt, c->config_time_integration_order, c->config_number_of_sub_steps,
                       c->config_dynamics_split_steps);
  h->sca = rk_schedule(c->config_dt, c->config_time_integration_order, c->config_number_of_sub_steps, 1);
  for (int i = 0; i < 3; ++i) {
    K.dts[i] = (float)h->dyn.acoustic[i];
    K.dts_inv_coef[i] = K.divdamp_base * (1.0f / K.dts[i]);
    K.steps[i] = h->dyn.steps[i];
    K.inv_steps[i] = 1.0f / (float)h->dyn.steps[i];
    K.sca_dt[i] = (float)h->sca.large[i];
    K.wn[i] = 1.0f;
  }
  if (c->config_split_dynamics_transport) {
    K.wn[0] = (float)(1.0 / (c->config_time_integration_order == 3 ? 3.0 : 2.0));
    K.wn[1] = 0.5f;
  }
  K.split_inv = 1.0f / (float)c->config_dynamics_split_steps;
  K.order = c->config_time_integration_order;
  K.splits = c->config_dynamics_split_steps;
  K.monotonic = c->config_monotonic;
  K.divdamp = c->config_divergence_damping;
  K.split_transport = c->config_split_dynamics_transport;
  K.scalar_adv = c->config_scalar_advection;
  *out = h;
  return 0;
}

}  // extern "C"

namespace {

struct FieldRef {
  float *ptr;
  int n;       // entity count
  int nlev;    // levels
  const int *old;
  int planes;  // scalar planes
};

FieldRef field_ref(hexdyn_handle *h, int id) {
  Instance J = view(h, false);  // mid-step fields as the next phase would read them
  Work &B = J.b;
  const Sizes &n = h->I.n;
  const int C = n.C, E = n.E, V = n.V, L = n.L;
  const bool in = h->in_step;
  auto cell = [&](float *p, int lv) { return FieldRef{p, C, lv, h->old_c, 1}; };
  auto edge = [&](float *p) { return FieldRef{p, E, L, h->old_e, 1}; };
  switch (id) {
    case HEXDYN_F_RHO_ZZ: return cell(in ? B.c_rho : B.rho, L);
    case HEXDYN_F_RTHETA_M: return cell(in ? B.c_rt : B.rt, L);
    case HEXDYN_F_RU: return edge(in ? B.c_ru : B.ru);
    case HEXDYN_F_RW: return cell(in ? B.c_rw : B.rw, L + 1);
    case HEXDYN_F_SCALARS: return FieldRef{B.q, C, L, h->old_c, n.NT};
    case HEXDYN_F_SCALARS_STAGE: return FieldRef{B.q_stage, C, L, h->old_c, n.NT};
    case HEXDYN_F_THETA_M: return cell(in ? B.theta : B.p_theta, L);
    case HEXDYN_F_EXNER: return cell(in ? B.exner : B.p_exner, L);
    case HEXDYN_F_RHO_P: return cell(in ? B.rho_p : B.p_rho_p, L);
    case HEXDYN_F_RTHETA_P: return cell(in ? B.rtheta_p : B.p_rtheta_p, L);
    case HEXDYN_F_PRESSURE_P: return cell(in ? B.pp : B.p_pp, L);
    case HEXDYN_F_U: return edge(in ? B.u : B.p_u);
    case HEXDYN_F_W: return cell(in ? B.w : B.p_w, L + 1);
    case HEXDYN_F_V: return edge(B.v);
    case HEXDYN_F_PV_EDGE: return edge(B.pv_edge);
    case HEXDYN_F_KE: return cell(B.ke, L);
    case HEXDYN_F_DIVERGENCE: return cell(B.divergence, L);
    case HEXDYN_F_VORTICITY: return FieldRef{B.vorticity, V, L, h->old_v, 1};
    case HEXDYN_F_TEND_RU: return edge(B.tend_ru);
    case HEXDYN_F_TEND_RT: return cell(B.tend_rt, L);
    case HEXDYN_F_TEND_RW: return cell(B.tend_rw, L + 1);
    case HEXDYN_F_TEND_RHO: return cell(B.tend_rho, L);
    case HEXDYN_F_RU_P: return edge(B.ru_p);
    case HEXDYN_F_RW_P: return cell(B.rw_p, L + 1);
    case HEXDYN_F_RHO_PP: return cell(B.rho_pp, L);
    case HEXDYN_F_RTHETA_PP: return cell(B.rtheta_pp, L);
    case HEXDYN_F_RU_AVG: return edge(B.ru_avg);
    case HEXDYN_F_WW_AVG: return cell(B.ww_avg, L + 1);
    case HEXDYN_F_A_TRI: return cell(B.atri, L);
    case HEXDYN_F_ALPHA_TRI: return cell(B.alpha, L);
    case HEXDYN_F_GAMMA_TRI: return cell(B.gamma, L);
    case HEXDYN_F_COFWZ: return cell(B.cofwz, L);
    case HEXDYN_F_COFWT: return cell(B.cofwt, L);
    case HEXDYN_F_COFTZ: return cell(B.coftz, L + 1);
    case HEXDYN_F_EULER_RU: return edge(B.euler_ru);
    case HEXDYN_F_EULER_RW: return cell(B.euler_rw, L + 1);
    case HEXDYN_F_KDIFF: return cell(B.kdiff, L);
    case HEXDYN_F_EULER_THETA: return cell(B.euler_theta, L);
    case HEXDYN_F_FLUX_U: return edge(B.flux_u);
    case HEXDYN_F_FLUX_W: return cell(B.flux_w, L + 1);
    case HEXDYN_F_QTOT: return cell(B.qtot, L);
    case HEXDYN_F_CQW: return cell(B.cqw, L);
    case HEXDYN_F_CQU: return edge(B.cqu);
    default: return FieldRef{nullptr, 0, 0, nullptr, 0};
  }
}

int transfer(hexdyn_handle *h, const FieldRef &f, float *dev, bool to_internal, hipStream_t s) {
  for (int p = 0; p < f.planes; ++p) {
    float *in_rows = f.ptr + (size_t)p * f.n * h->I.n.RS;
    float *logical = dev + (size_t)p * f.nlev * f.n;
    const int nb = (f.n + 3) / 4;
    if (to_internal)
      hipLaunchKernelGGL(k_import, dim3(nb), dim3(256), 0, s, in_rows, (const float *)logical, f.old, f.n, f.nlev,
                         h->I.n.RS);
    else
      hipLaunchKernelGGL(k_export, dim3(nb), dim3(256), 0, s, logical, (const float *)in_rows, f.old, f.n, f.nlev,
                         h->I.n.RS);
  }
  return hipGetLastError() == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int copy_rows(float *dst, const float *src, size_t rows, hipStream_t s) {
  return hipMemcpyAsync(dst, src, rows * HX_ROW * sizeof(float), hipMemcpyDeviceToDevice, s) == hipSuccess
             ? 0
             : HEXDYN_E_HIP;
}

__global__ void k_clear_flag(int *flag) {
  if (threadIdx.x == 0) flag[0] = 0;
}

int run_phase_begin_step(hexdyn_handle *h, hipStream_t s, bool fused) {
  Work &B = h->I.b;
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, B.flag);
  if (hipGetLastError() != hipSuccess) return HEXDYN_E_HIP;
  h->in_step = 1;
  h->start_is_pub = 1;
  h->fresh_sub = 1;
  h->fresh_step = 1;
  h->exner_start = 1;
  h->pp_pub = 1;
  Instance J = view(h, false);
  J.k.fused = fused ? 1 : 0;
  J.k.store_flux = fused ? 0 : 1;
  return NS_CALL(h, begin_step, J, s);
}

bool transported(const hexdyn_handle *h) {
  return h->cfg.config_scalar_advection && h->cfg.config_split_dynamics_transport && h->I.n.NT > 0;
}

// device work that closes a step (before the validation flag is read): the
// q < 0 -> 0 clamp of every species (cuda_driver.py:5306-5311, atm_srk3's
// unconditional clamp), a where-assignment that keeps -0.0 and NaN bits.
// Monotonic transport already ends every value in max(0, .) >= +0, so the
// clamp would change nothing there.  Without transport the published scalars
// are clamped into the candidate buffer, which the publish then takes.
int end_step_device(hexdyn_handle *h, hipStream_t s) {
  if (!h->post_clamp || h->I.n.NT == 0) return 0;
  if (transported(h)) {
    if (!h->cfg.config_monotonic && NS_CALL(h, clamp_scalars, h->I, h->I.b.q_stage, h->I.b.q_stage, s))
      return HEXDYN_E_HIP;
  } else if (NS_CALL(h, clamp_scalars, h->I, h->I.b.q, h->I.b.q_stage, s)) {
    return HEXDYN_E_HIP;
  }
  return 0;
}

// the candidate scalars (q_stage) replace the published ones at publish
bool scalars_in_stage(const hexdyn_handle *h) { return transported(h) || (h->post_clamp && h->I.n.NT > 0); }

// read the flag; publish by swapping buffers when it is clear
int end_step_publish(hexdyn_handle *h, hipStream_t s) {
  Work &B = h->I.b;
  int flag = 0;
  if (hipMemcpyAsync(&flag, B.flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return HEXDYN_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return HEXDYN_E_HIP;
  h->in_step = 0;
  h->start_is_pub = h->fresh_sub = h->fresh_step = h->exner_start = h->pp_pub = 0;
  if (flag) {
    h->error = "validation flag set: non-finite or non-positive state; step not published";
    return HEXDYN_E_NONFINITE;
  }
  std::swap(B.rho, B.c_rho); std::swap(B.rt, B.c_rt); std::swap(B.ru, B.c_ru); std::swap(B.rw, B.c_rw);
  std::swap(B.p_theta, B.theta); std::swap(B.p_exner, B.exner); std::swap(B.p_rho_p, B.rho_p);
  std::swap(B.p_rtheta_p, B.rtheta_p); std::swap(B.p_pp, B.pp); std::swap(B.p_u, B.u); std::swap(B.p_w, B.w);
  if (scalars_in_stage(h)) std::swap(B.q, B.q_stage);
  h->time_s += h->cfg.config_dt;
  return 0;
}

int run_end_step(hexdyn_handle *h, hipStream_t s) {
  if (end_step_device(h, s)) return HEXDYN_E_HIP;
  return end_step_publish(h, s);
}

}  // namespace

extern "C" {

int hexdyn_set_state(hexdyn_handle *h, const hexdyn_state_view *st, int rebuild, void *stream) {
  if (!h || !st) return HEXDYN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  h->in_step = 0;
  h->start_is_pub = h->fresh_sub = h->fresh_step = h->exner_start = h->pp_pub = 0;
  const struct { int id; float *p; } prog[] = {
      {HEXDYN_F_RHO_ZZ, st->rho_zz}, {HEXDYN_F_RTHETA_M, st->rtheta_m}, {HEXDYN_F_RU, st->ru},
      {HEXDYN_F_RW, st->rw}, {HEXDYN_F_SCALARS, st->scalars}};
  for (auto &f : prog) {
    if (!f.p) {
      if (f.id == HEXDYN_F_SCALARS && h->I.n.NT == 0) continue;
      return set_error(h, HEXDYN_E_INVALID, "missing prognostic field");
    }
    if (transfer(h, field_ref(h, f.id), f.p, true, s)) return HEXDYN_E_HIP;
  }
  Work &B = h->I.b;
  if (rebuild) {
    if (NS_CALL(h, rebuild_saved, h->I, s)) return HEXDYN_E_HIP;
    const Sizes &n = h->I.n;
    int rc = 0;
    rc |= copy_rows(B.p_theta, B.theta, n.C, s); rc |= copy_rows(B.p_exner, B.exner, n.C, s);
    rc |= copy_rows(B.p_rho_p, B.rho_p, n.C, s); rc |= copy_rows(B.p_rtheta_p, B.rtheta_p, n.C, s);
    rc |= copy_rows(B.p_pp, B.pp, n.C, s); rc |= copy_rows(B.p_u, B.u, n.E, s);
    rc |= copy_rows(B.p_w, B.w, n.C, s);
    if (rc) return HEXDYN_E_HIP;
  } else {
    const struct { int id; float *p; } diag[] = {
        {HEXDYN_F_THETA_M, st->theta_m}, {HEXDYN_F_EXNER, st->exner}, {HEXDYN_F_RHO_P, st->rho_p},
        {HEXDYN_F_RTHETA_P, st->rtheta_p}, {HEXDYN_F_PRESSURE_P, st->pressure_p}, {HEXDYN_F_U, st->u},
        {HEXDYN_F_W, st->w}};
    for (auto &f : diag) {
      if (!f.p) return set_error(h, HEXDYN_E_INVALID, "missing saved diagnostic field");
      if (transfer(h, field_ref(h, f.id), f.p, true, s)) return HEXDYN_E_HIP;
    }
  }
  return hipStreamSynchronize(s) == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int hexdyn_get_state(hexdyn_handle *h, const hexdyn_state_view *st, void *stream) {
  if (!h || !st) return HEXDYN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int saved = h->in_step;
  h->in_step = 0;
  const struct { int id; float *p; } all[] = {
      {HEXDYN_F_RHO_ZZ, st->rho_zz}, {HEXDYN_F_RTHETA_M, st->rtheta_m}, {HEXDYN_F_RU, st->ru},
      {HEXDYN_F_RW, st->rw}, {HEXDYN_F_SCALARS, st->scalars}, {HEXDYN_F_THETA_M, st->theta_m},
      {HEXDYN_F_EXNER, st->exner}, {HEXDYN_F_RHO_P, st->rho_p}, {HEXDYN_F_RTHETA_P, st->rtheta_p},
      {HEXDYN_F_PRESSURE_P, st->pressure_p}, {HEXDYN_F_U, st->u}, {HEXDYN_F_W, st->w}};
  int rc = 0;
  for (auto &f : all)
    if (f.p && transfer(h, field_ref(h, f.id), f.p, false, s)) rc = HEXDYN_E_HIP;
  h->in_step = saved;
  if (rc) return rc;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int hexdyn_field_shape(const hexdyn_handle *h, int field, int64_t *rows, int64_t *cols) {
  if (!h || !rows || !cols) return HEXDYN_E_INVALID;
  FieldRef f = field_ref(const_cast<hexdyn_handle *>(h), field);
  if (!f.ptr) return HEXDYN_E_INVALID;
  *rows = (int64_t)f.planes * f.nlev;
  *cols = f.n;
  return 0;
}

int hexdyn_export_field(hexdyn_handle *h, int field, float *dst, void *stream) {
  if (!h || !dst) return HEXDYN_E_INVALID;
  FieldRef f = field_ref(h, field);
  if (!f.ptr) return HEXDYN_E_INVALID;
  if (transfer(h, f, dst, false, (hipStream_t)stream)) return HEXDYN_E_HIP;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int hexdyn_import_field(hexdyn_handle *h, int field, const float *src, void *stream) {
  if (!h || !src) return HEXDYN_E_INVALID;
  FieldRef f = field_ref(h, field);
  if (!f.ptr) return HEXDYN_E_INVALID;
  if (transfer(h, f, const_cast<float *>(src), true, (hipStream_t)stream)) return HEXDYN_E_HIP;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int hexdyn_begin_step(hexdyn_handle *h, void *stream) {
  if (!h) return HEXDYN_E_INVALID;
  return run_phase_begin_step(h, (hipStream_t)stream, false);
}

int hexdyn_begin_subcycle(hexdyn_handle *h, int subcycle, void *stream) {
  if (!h || subcycle < 1 || subcycle > h->cfg.config_dynamics_split_steps) return HEXDYN_E_INVALID;
  if (!h->in_step) return set_error(h, HEXDYN_E_INVALID, "hexdyn_begin_subcycle outside a step");
  h->cur_sub = subcycle;
  rotate_subcycle(h, subcycle);
  return NS_CALL(h, begin_subcycle, view(h, false), subcycle, (hipStream_t)stream) ? HEXDYN_E_HIP : 0;
}

int hexdyn_dyn_tend(hexdyn_handle *h, int stage, void *stream) {
  if (!h || stage < 1 || stage > 3) return HEXDYN_E_INVALID;
  return NS_CALL(h, dyn_tend, view(h, false), stage, (hipStream_t)stream) ? HEXDYN_E_HIP : 0;
}

int hexdyn_acoustic_substep(hexdyn_handle *h, int stage, int small_step, void *stream) {
  if (!h || stage < 1 || stage > 3 || small_step < 1 || small_step > h->I.k.steps[stage - 1]) return HEXDYN_E_INVALID;
  return NS_CALL(h, acoustic_substep, view(h, false), stage, small_step, 0, (hipStream_t)stream) ? HEXDYN_E_HIP
                                                                                                   : 0;
}

int hexdyn_divergence_damping(hexdyn_handle *h, int stage, void *stream) {
  if (!h || stage < 1 || stage > 3) return HEXDYN_E_INVALID;
  return NS_CALL(h, divergence_damping, view(h, false), stage, (hipStream_t)stream) ? HEXDYN_E_HIP : 0;
}

int hexdyn_recover(hexdyn_handle *h, int stage, void *stream) {
  if (!h || stage < 1 || stage > 3) return HEXDYN_E_INVALID;
  // At stage 3 the recovered fluxes are also accumulated into the split-transport
  // sums (copy on the first subcycle, add afterwards: integration.py:27-56).
  Instance J = view(h, true);
  const int rc = NS_CALL(h, recover, J, stage, h->cur_sub, 0, (hipStream_t)stream);
  recovered(h, stage);
  return rc ? HEXDYN_E_HIP : 0;
}

int hexdyn_solve_diagnostics(hexdyn_handle *h, int rk_step, void *stream) {
  if (!h || rk_step < 1 || rk_step > 3) return HEXDYN_E_INVALID;
  Instance J = view(h, false);
  return NS_CALL(h, solve_diagnostics, J, rk_step, (hipStream_t)stream) ? HEXDYN_E_HIP : 0;
}

int hexdyn_end_subcycle(hexdyn_handle *h, int subcycle, void *stream) {
  (void)stream;
  if (!h || subcycle != h->cur_sub) return HEXDYN_E_INVALID;
  return 0;  // carried diagnostics are updated in place; split sums were accumulated by hexdyn_recover
}

int hexdyn_scalars(hexdyn_handle *h, int stage, void *stream) {
  if (!h || stage < 1 || stage > 3) return HEXDYN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  Work &B = h->I.b;
  // stage 1 reads the published scalars in place; stages 2-3 read the previous stage's result
  if (NS_CALL(h, scalars_stage, view(h, false), stage, s)) return HEXDYN_E_HIP;
  std::swap(B.q_stage, B.q_next);
  return 0;
}

int hexdyn_end_step(hexdyn_handle *h, void *stream) {
  if (!h) return HEXDYN_E_INVALID;
  return run_end_step(h, (hipStream_t)stream);
}

namespace {

// The instance as the whole-step sequence sees it: fused launches on.
Instance step_view(const hexdyn_handle *h, bool recovery) {
  Instance J = view(h, recovery);
  J.k.store_flux = 0;
  J.k.fused = 1;
  return J;
}

// Every device operation of one large step, in order, on stream s.  Each
// stage's theta/w edge fluxes are formed by the diagnostics pass before it
// (begin_step's for the first stage of the first subcycle).
int enqueue_step(hexdyn_handle *h, hipStream_t s) {
  const hexdyn_config &c = h->cfg;
  if (!h->side.s &&
      (hipStreamCreateWithFlags(&h->side.s, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&h->side.fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&h->side.join, hipEventDisableTiming) != hipSuccess))
    return HEXDYN_E_HIP;
  int rc = run_phase_begin_step(h, s, true);
  // profiled steps stay on one stream, so each kernel's HIP-event time is its own
  const Side *side = g_prof_on ? nullptr : &h->side;
  for (int sub = 1; rc == 0 && sub <= c.config_dynamics_split_steps; ++sub) {
    h->cur_sub = sub;
    rotate_subcycle(h, sub);
    rc |= NS_CALL(h, begin_subcycle, step_view(h, false), sub, s, side);
    for (int st = 1; rc == 0 && st <= 3; ++st) {
      rc |= NS_CALL(h, dyn_tend, step_view(h, false), st, s, side);
      // divergence damping of substep ss is fused into the next edge pass:
      // the acoustic edge update of ss+1, or the edge recovery after the last substep
      const int dd = c.config_divergence_damping ? 1 : 0;
      for (int ss = 1; rc == 0 && ss <= h->I.k.steps[st - 1]; ++ss)
        rc |= NS_CALL(h, acoustic_substep, step_view(h, false), st, ss, (dd && ss > 1) ? 1 : 0, s);
      rc |= NS_CALL(h, recover, step_view(h, true), st, sub, dd, s);
      recovered(h, st);
      // the last stage's diagnostics would only feed a subcycle that does not
      // follow: the next step's begin_step recomputes them from the published state
      if (!(sub == c.config_dynamics_split_steps && st == 3))
        rc |= NS_CALL(h, solve_diagnostics, step_view(h, false), st, s);
    }
  }
  if (rc == 0 && c.config_split_dynamics_transport && transported(h))
    for (int st = 1; rc == 0 && st <= 3; ++st) rc |= hexdyn_scalars(h, st, s);
  if (rc == 0) rc = end_step_device(h, s);
  return rc;
}

// Key of what a step's launches depend on besides the static mesh: the
// buffer-role layout at step start and the run-time switches.  Work holds
// only pointers and Consts one double and 4-byte scalars (asserted at its
// definition), so neither has padding bytes.
std::string layout_key(const hexdyn_handle *h) {
  return std::string(reinterpret_cast<const char *>(&h->I.b), sizeof(Work)) +
         std::string(reinterpret_cast<const char *>(&h->I.k), sizeof(Consts)) +
         std::string(1, h->post_clamp ? '1' : '0');
}

// Enqueue one step as a graph launch on s.  The first step from a given
// buffer-role layout is captured (on a private stream) and stored with the
// roles it leaves behind; later steps from that layout replay the graph and
// take the stored roles, which the launches determine completely.
int enqueue_step_graph(hexdyn_handle *h, hipStream_t s) {
  const std::string key = layout_key(h);
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    if (!h->capture_stream && hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking) != hipSuccess)
      return HEXDYN_E_HIP;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return HEXDYN_E_HIP;
    const int rc = enqueue_step(h, h->capture_stream);
    const hipError_t e = hipStreamEndCapture(h->capture_stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return HEXDYN_E_HIP;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return HEXDYN_E_HIP;
    it = h->graphs.emplace(key, std::make_pair(x, h->roles())).first;
  }
  if (hipGraphLaunch(it->second.first, s) != hipSuccess) return HEXDYN_E_HIP;
  h->set_roles(it->second.second);
  return 0;
}

}  // namespace

int hexdyn_step(hexdyn_handle *h, int n_steps, void *stream) {
  if (!h || n_steps < 0) return HEXDYN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  for (int it = 0; it < n_steps; ++it) {
    const bool graph = h->use_graphs && !g_prof_on;
    int rc = graph ? enqueue_step_graph(h, s) : enqueue_step(h, s);
    if (rc) {
      h->in_step = 0;
      h->start_is_pub = h->fresh_sub = h->fresh_step = h->exner_start = h->pp_pub = 0;
      if (h->error.empty()) h->error = "kernel launch or graph capture failed";
      return HEXDYN_E_HIP;
    }
    rc = end_step_publish(h, s);
    if (rc) return rc;
  }
  return 0;
}

int hexdyn_set_physics_tendencies(hexdyn_handle *h, const hexdyn_physics_view *p, void *stream) {
  if (!h) return HEXDYN_E_INVALID;
  if (h->in_step) return set_error(h, HEXDYN_E_INVALID, "physics tendencies change between steps only");
  Work &B = h->I.b;
  if (!p) {
    h->I.k.physics = 0;
    h->I.k.physics_q = 0;
    return 0;
  }
  const Sizes &n = h->I.n;
  if (!B.phys_rho) {
    int rc = h->alloc(&B.phys_rho, (size_t)n.C * HX_ROW);
    rc |= h->alloc(&B.phys_ru, (size_t)n.E * HX_ROW);
    rc |= h->alloc(&B.phys_rt, (size_t)n.C * HX_ROW);
    if (rc) return set_error(h, HEXDYN_E_HIP, "physics buffer allocation failed");
  }
  hipStream_t s = (hipStream_t)stream;
  const struct { const float *src; float *dst; int n; const int *old; } f[] = {
      {p->rho, B.phys_rho, n.C, h->old_c}, {p->rho_u, B.phys_ru, n.E, h->old_e}, {p->rho_theta, B.phys_rt, n.C, h->old_c}};
  for (const auto &x : f) {
    if (x.src) {
      if (transfer(h, FieldRef{x.dst, x.n, n.L, x.old, 1}, const_cast<float *>(x.src), true, s)) return HEXDYN_E_HIP;
    } else if (hipMemsetAsync(x.dst, 0, (size_t)x.n * HX_ROW * sizeof(float), s) != hipSuccess) {
      return HEXDYN_E_HIP;
    }
  }
  h->I.k.physics = 1;
  h->I.k.physics_q = 0;
  if (p->scalars && n.NT > 0) {
    if (!B.phys_q) {
      int rc = h->alloc(&B.phys_q, (size_t)n.NT * n.C * HX_ROW);
      rc |= h->alloc(&B.q_src, (size_t)n.NT * n.C * HX_ROW);
      if (rc) return set_error(h, HEXDYN_E_HIP, "physics buffer allocation failed");
    }
    if (transfer(h, FieldRef{B.phys_q, n.C, n.L, h->old_c, n.NT}, const_cast<float *>(p->scalars), true, s))
      return HEXDYN_E_HIP;
    h->I.k.physics_q = 1;
  }
  return hipStreamSynchronize(s) == hipSuccess ? 0 : HEXDYN_E_HIP;
}

int hexdyn_graph_count(const hexdyn_handle *h) { return h ? (int)h->graphs.size() : 0; }

int hexdyn_set_graphs(hexdyn_handle *h, int enable) {
  if (!h) return HEXDYN_E_INVALID;
  h->use_graphs = enable ? 1 : 0;
  return 0;
}

namespace {
void drop_graphs(hexdyn_handle *h) {
  (void)hipDeviceSynchronize();
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second.first);
  h->graphs.clear();
}
// kernels with a fixed block shape: the column kernels (16 or 8 columns per
// block through LDS) and the edge-tiled flux pass (HX_TILE_E edges per block)
bool column_kernel(const char *name) {
  return std::strcmp(name, "k_vcoef") == 0 || std::strcmp(name, "k_ac_cell") == 0 ||
         std::strcmp(name, "k_diag_edge_flux_t") == 0;
}
}  // namespace

int hexdyn_set_launch(hexdyn_handle *h, const char *kernel, int waves_per_block) {
  if (!h || !kernel) return HEXDYN_E_INVALID;
  if (h->in_step) return set_error(h, HEXDYN_E_INVALID, "hexdyn_set_launch inside a step");
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4 && waves_per_block != 8 &&
      waves_per_block != 16)
    return set_error(h, HEXDYN_E_INVALID, "waves_per_block must be 1, 2, 4, 8 or 16");
  LaunchTune &t = h->tune;
  if (std::strcmp(kernel, "*") == 0) {
    t.n = 0;
    t.def = waves_per_block;
  } else {
    if (column_kernel(kernel) || std::strlen(kernel) >= 32 || std::strncmp(kernel, "k_", 2) != 0)
      return set_error(h, HEXDYN_E_INVALID, std::string("not a tunable gather kernel: ") + kernel);
    int i = 0;
    while (i < t.n && std::strncmp(t.names[i], kernel, 31) != 0) ++i;
    if (i == t.n) {
      if (t.n == LaunchTune::kMax) return set_error(h, HEXDYN_E_INVALID, "launch table full");
      std::snprintf(t.names[i], sizeof t.names[i], "%s", kernel);
      ++t.n;
    }
    t.wpb[i] = waves_per_block;
  }
  drop_graphs(h);
  return 0;
}

int hexdyn_get_launch(const hexdyn_handle *h, const char *kernel) {
  if (!h || !kernel) return -1;
  return h->tune.waves_per_block(kernel);
}

int hexdyn_set_post_transport_clamp(hexdyn_handle *h, int enable) {
  if (!h) return HEXDYN_E_INVALID;
  if (h->in_step) return set_error(h, HEXDYN_E_INVALID, "hexdyn_set_post_transport_clamp inside a step");
  h->post_clamp = enable ? 1 : 0;
  return 0;
}

int hexdyn_integrals(hexdyn_handle *h, double *out, int n_out, void *stream) {
  if (!h || !out) return HEXDYN_E_INVALID;
  const int NQ = 3 + h->I.n.NT;
  if (n_out < NQ) return set_error(h, HEXDYN_E_INVALID, "hexdyn_integrals: n_out < 3 + n_scalars");
  if (h->in_step) return set_error(h, HEXDYN_E_INVALID, "hexdyn_integrals: a step is in progress");
  hipStream_t s = (hipStream_t)stream;
  const Work &B = h->I.b;
  if (hexdyn::integrals(h->I, B.rho, B.rt, B.q, B.p_theta, B.p_exner, B.p_u, h->I.z.zmid, h->int_partial,
                        h->int_out, s))
    return set_error(h, HEXDYN_E_HIP, "integral kernels failed to launch");
  if (hipMemcpyAsync(out, h->int_out, NQ * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return set_error(h, HEXDYN_E_HIP, "integral copy failed");
  return 0;
}

int hexdyn_integrals_noalloc(hexdyn_handle *h, double *out_device, int n_out, void *stream) {
  if (!h || !out_device) return HEXDYN_E_INVALID;
  const int NQ = 3 + h->I.n.NT;
  if (n_out < NQ) return set_error(h, HEXDYN_E_INVALID, "hexdyn_integrals_noalloc: n_out < 3 + n_scalars");
  if (h->in_step) return set_error(h, HEXDYN_E_INVALID, "hexdyn_integrals_noalloc: a step is in progress");
  const size_t need = (size_t)((h->I.n.C + 3) / 4) * (size_t)NQ;
  if (need > (size_t)HEXDYN_NOALLOC_MAX_PARTIALS)
    return set_error(h, HEXDYN_E_INVALID,
                     "hexdyn_integrals_noalloc: ceil(n_cells/4)*(3+n_scalars) = " + std::to_string(need) +
                         " partial sums exceed the scratch of " + std::to_string(HEXDYN_NOALLOC_MAX_PARTIALS));
  double *part = scratch_ptr(h->device);
  if (!part) return set_error(h, HEXDYN_E_HIP, "hexdyn_integrals_noalloc: device scratch not resolved");
  const Work &B = h->I.b;
  if (hexdyn::integrals(h->I, B.rho, B.rt, B.q, B.p_theta, B.p_exner, B.p_u, h->I.z.zmid, part, out_device,
                        (hipStream_t)stream))
    return set_error(h, HEXDYN_E_HIP, "integral kernels failed to launch");
  return 0;
}

// ---------------------------------------------------------------- state copy
namespace {
// Buffers of one state copy (every row is 256-byte aligned: float4 moves).
struct CopyList {
  int n;
  float *dst[16];
  const float *src[16];
  size_t count[16];  // float4 elements
};

// all buffers of the list in one launch: each buffer is split over the grid
__global__ void __launch_bounds__(256) k_copy_many(const CopyList cl) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int b = 0; b < cl.n; ++b) {
    const float4 *src = (const float4 *)cl.src[b];
    float4 *dst = (float4 *)cl.dst[b];
    for (size_t i = tid; i < cl.count[b]; i += stride) dst[i] = src[i];
  }
}

bool same_mesh(const hexdyn_handle *a, const hexdyn_handle *b) {
  const Sizes &x = a->I.n, &y = b->I.n;
  return x.C == y.C && x.E == y.E && x.V == y.V && x.L == y.L && x.ME == y.ME && x.NT == y.NT &&
         a->old_c_h == b->old_c_h && a->old_e_h == b->old_e_h && a->old_v_h == b->old_v_h;
}
}  // namespace

int hexdyn_copy_state(hexdyn_handle *dst, hexdyn_handle *src, void *stream) {
  if (!dst || !src) return HEXDYN_E_INVALID;
  if (dst == src) return 0;
  if (dst->in_step || src->in_step) return set_error(dst, HEXDYN_E_INVALID, "hexdyn_copy_state inside a step");
  if (!same_mesh(dst, src)) return set_error(dst, HEXDYN_E_INVALID, "hexdyn_copy_state: handles of different meshes");
  hipStream_t s = (hipStream_t)stream;
  const Sizes &n = src->I.n;
  const Work &S = src->I.b;
  Work &D = dst->I.b;
  CopyList cl{};
  auto add = [&](float *d, const float *p, size_t rows) {
    cl.dst[cl.n] = d;
    cl.src[cl.n] = p;
    cl.count[cl.n] = rows * HX_ROW / 4;  // float4 per row quarter
    ++cl.n;
  };
  add(D.rho, S.rho, n.C); add(D.rt, S.rt, n.C); add(D.ru, S.ru, n.E); add(D.rw, S.rw, n.C);
  if (n.NT > 0) add(D.q, S.q, (size_t)n.NT * n.C);
  add(D.p_theta, S.p_theta, n.C); add(D.p_exner, S.p_exner, n.C); add(D.p_rho_p, S.p_rho_p, n.C);
  add(D.p_rtheta_p, S.p_rtheta_p, n.C); add(D.p_pp, S.p_pp, n.C); add(D.p_u, S.p_u, n.E);
  add(D.p_w, S.p_w, n.C);
  hipLaunchKernelGGL(k_copy_many, dim3(2048), dim3(256), 0, s, cl);
  if (hipGetLastError() != hipSuccess) return set_error(dst, HEXDYN_E_HIP, "hexdyn_copy_state: copy failed");
  dst->time_s = src->time_s;
  return 0;
}

// ---------------------------------------------------------------- corrector
struct hexdyn_net {
  hexdyn::net::Net *impl = nullptr;
  hexdyn_handle *h = nullptr;  // the mesh the network was created on
  int layers = 0;
  std::string error;
};

namespace {
int net_error(hexdyn_net *n, int rc, const std::string &msg) {
  n->error = msg;
  return rc;
}
int net_mesh_ok(hexdyn_net *n, hexdyn_handle *h) {
  if (!h) return net_error(n, HEXDYN_E_INVALID, "null handle");
  if (h->in_step) return net_error(n, HEXDYN_E_INVALID, "a step is in progress on the handle");
  if (h != n->h && !same_mesh(h, n->h)) return net_error(n, HEXDYN_E_INVALID, "handle of another mesh");
  return 0;
}
}  // namespace

int hexdyn_net_create(hexdyn_handle *h, int layers, hexdyn_net **out) {
  if (!h || !out || layers < 1 || layers > 8) return HEXDYN_E_INVALID;
  *out = nullptr;
  std::string err;
  hexdyn::net::Net *impl = hexdyn::net::create(h->I, layers, &err);
  if (!impl) return set_error(h, HEXDYN_E_HIP, err);
  hexdyn_net *n = new hexdyn_net();
  n->impl = impl;
  n->h = h;
  n->layers = layers;
  *out = n;
  return 0;
}

void hexdyn_net_destroy(hexdyn_net *net) {
  if (!net) return;
  (void)hipDeviceSynchronize();
  hexdyn::net::destroy(net->impl);
  delete net;
}

int64_t hexdyn_net_param_count(const hexdyn_net *net) {
  return net ? (int64_t)hexdyn::net::param_count(net->impl) : 0;
}

int hexdyn_net_set_params(hexdyn_net *net, const float *host, int64_t n, void *stream) {
  if (!net || !host) return HEXDYN_E_INVALID;
  if (n != (int64_t)hexdyn::net::param_count(net->impl)) return net_error(net, HEXDYN_E_INVALID, "parameter count");
  return hexdyn::net::set_params(*net->impl, host, (hipStream_t)stream) ? net_error(net, HEXDYN_E_HIP, "upload") : 0;
}

int hexdyn_net_get_params(hexdyn_net *net, float *host, int64_t n, void *stream) {
  if (!net || !host) return HEXDYN_E_INVALID;
  if (n != (int64_t)hexdyn::net::param_count(net->impl)) return net_error(net, HEXDYN_E_INVALID, "parameter count");
  return hexdyn::net::get_vec(*net->impl, 0, host, (hipStream_t)stream) ? net_error(net, HEXDYN_E_HIP, "download") : 0;
}

int hexdyn_net_get_grads(hexdyn_net *net, float *host, int64_t n, void *stream) {
  if (!net || !host) return HEXDYN_E_INVALID;
  if (n != (int64_t)hexdyn::net::param_count(net->impl)) return net_error(net, HEXDYN_E_INVALID, "parameter count");
  return hexdyn::net::get_vec(*net->impl, 1, host, (hipStream_t)stream) ? net_error(net, HEXDYN_E_HIP, "download") : 0;
}

int hexdyn_net_set_scales(hexdyn_net *net, const float *in_mean, const float *in_scale, const float *out_scale,
                          void *stream) {
  if (!net || !in_mean || !in_scale || !out_scale) return HEXDYN_E_INVALID;
  return hexdyn::net::set_scales(*net->impl, in_mean, in_scale, out_scale, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "upload")
             : 0;
}

int hexdyn_net_forward(hexdyn_net *net, hexdyn_handle *h, int keep, void *stream) {
  if (!net) return HEXDYN_E_INVALID;
  if (int rc = net_mesh_ok(net, h)) return rc;
  return hexdyn::net::forward(*net->impl, h->I, keep != 0, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "forward kernels failed to launch")
             : 0;
}

int hexdyn_net_set_target(hexdyn_net *net, hexdyn_handle *coarse, hexdyn_handle *fine, void *stream) {
  if (!net) return HEXDYN_E_INVALID;
  if (int rc = net_mesh_ok(net, coarse)) return rc;
  if (int rc = net_mesh_ok(net, fine)) return rc;
  return hexdyn::net::target(*net->impl, coarse->I, fine->I, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "target kernels failed to launch")
             : 0;
}

int hexdyn_net_loss(hexdyn_net *net, double *out5, void *stream) {
  if (!net || !out5) return HEXDYN_E_INVALID;
  return hexdyn::net::loss(*net->impl, net->h->I, out5, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "loss kernels failed")
             : 0;
}

int hexdyn_net_backward(hexdyn_net *net, void *stream) {
  if (!net) return HEXDYN_E_INVALID;
  return hexdyn::net::backward(*net->impl, net->h->I, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "backward kernels failed to launch")
             : 0;
}

int hexdyn_net_adam(hexdyn_net *net, float lr, float beta1, float beta2, float eps, int step, void *stream) {
  if (!net || step < 1) return HEXDYN_E_INVALID;
  return hexdyn::net::adam(*net->impl, lr, beta1, beta2, eps, step, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "adam kernels failed to launch")
             : 0;
}

int hexdyn_net_apply(hexdyn_net *net, hexdyn_handle *h, float gain, void *stream) {
  if (!net) return HEXDYN_E_INVALID;
  if (int rc = net_mesh_ok(net, h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (hexdyn::net::apply(*net->impl, h->I, gain, s)) return net_error(net, HEXDYN_E_HIP, "apply kernels failed");
  // the carried cell diagnostics recomputed from the corrected prognostics,
  // written straight into the published buffers (the rebuild reads only
  // prognostics); the edge pass of the apply already formed u from ru
  Instance J = h->I;
  Work &B = J.b;
  B.theta = B.p_theta; B.exner = B.p_exner; B.rho_p = B.p_rho_p; B.rtheta_p = B.p_rtheta_p;
  B.pp = B.p_pp; B.u = B.p_u; B.w = B.p_w;
  if (NS_CALL(h, rebuild_cells, J, s)) return net_error(net, HEXDYN_E_HIP, "diagnostic rebuild failed");
  return 0;
}

int hexdyn_net_checks(hexdyn_net *net, hexdyn_handle *h, float *out3, double *dry_mass, void *stream) {
  if (!net || !out3 || !dry_mass) return HEXDYN_E_INVALID;
  if (int rc = net_mesh_ok(net, h)) return rc;
  return hexdyn::net::checks(*net->impl, h->I, (float)h->cfg.config_dt, out3, dry_mass, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "check kernels failed")
             : 0;
}

int hexdyn_net_features(hexdyn_net *net, hexdyn_handle *h, float *dst, void *stream) {
  if (!net || !dst) return HEXDYN_E_INVALID;
  if (int rc = net_mesh_ok(net, h)) return rc;
  return hexdyn::net::raw_features(*net->impl, h->I, h->old_c, dst, (hipStream_t)stream)
             ? net_error(net, HEXDYN_E_HIP, "feature kernels failed")
             : 0;
}

int hexdyn_net_export(hexdyn_net *net, int what, float *dst, void *stream) {
  if (!net || !dst) return HEXDYN_E_INVALID;
  const int rc = hexdyn::net::export_buf(*net->impl, what, net->h->old_c, net->h->old_e, dst, (hipStream_t)stream);
  if (rc > 0) return net_error(net, HEXDYN_E_INVALID, "unknown export");
  return rc ? net_error(net, HEXDYN_E_HIP, "export failed") : 0;
}

int hexdyn_net_last_error(const hexdyn_net *net, char *buf, size_t len) {
  if (!net || !buf || len == 0) return HEXDYN_E_INVALID;
  std::snprintf(buf, len, "%s", net->error.c_str());
  return 0;
}

int hexdyn_profile_enable(hexdyn_handle *h, int enable) {
  if (!h) return HEXDYN_E_INVALID;
  (void)hipDeviceSynchronize();
  for (auto &kv : g_prof)
    for (auto &p : kv.second.events) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  g_prof.clear();
  g_prof_open.clear();
  g_prof_on = enable ? 1 : 0;
  return 0;
}

int hexdyn_profile_read(hexdyn_handle *h, const char *kernel, double *avg_ms, int64_t *launches) {
  if (!h || !kernel || !avg_ms || !launches) return HEXDYN_E_INVALID;
  (void)hipDeviceSynchronize();
  auto it = g_prof.find(kernel);
  if (it == g_prof.end()) {
    *avg_ms = 0.0;
    *launches = 0;
    return HEXDYN_E_INVALID;
  }
  double total = 0.0;
  for (auto &p : it->second.events) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, p.first, p.second);
    total += ms;
  }
  *launches = (int64_t)it->second.events.size();
  *avg_ms = *launches ? total / (double)*launches : 0.0;
  return 0;
}

int hexdyn_profile_kernels(hexdyn_handle *h, char *buf, size_t len) {
  if (!h || !buf || len == 0) return HEXDYN_E_INVALID;
  std::string names;
  for (auto &kv : g_prof) {
    if (!names.empty()) names += ",";
    names += kv.first;
  }
  std::snprintf(buf, len, "%s", names.c_str());
  return 0;
}

}  // extern "C"
