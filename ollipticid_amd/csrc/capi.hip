// C ABI of include/olliptic_amd.h over the engine (engine.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/olliptic_amd.h"
#include "engine.hpp"

using namespace oll;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define CAPI_CHECK(expr)                                                             \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(OLL_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

#define CAPI_RC(expr)                   \
  do {                                  \
    std::string e_;                     \
    const int rc_ = (expr);             \
    if (rc_ != OLL_OK) return fail(rc_, e_); \
  } while (0)

}  // namespace

// Device scratch of oll_adm_mass_noalloc: the sphere samples of up to
// kScratchBoxes boxes (kAdmNt^2 doubles each), the box centres and the
// puncture points and values, 5.2 MB in all.  Each GPU the library runs on
// holds its own copy until the process ends; every call on a GPU reuses it,
// so calls that overlap on one GPU race (the entry point's contract).  In
// host code `scratch` only names the symbol: its device address on the
// current GPU comes from hipGetSymbolAddress (scratch_address), looked up
// once per device.
constexpr int kScratchBoxes = 64;
constexpr size_t kScratchSph = size_t(kScratchBoxes) * kAdmNt * kAdmNt;
constexpr size_t kScratchDoubles = kScratchSph + 3 * kScratchBoxes + 4 * OLL_MAX_BH + OLL_MAX_BH;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch;  // device id -> address of `scratch` there

int scratch_address(double **out) {
  int dev = 0;
  CAPI_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it == g_scratch.end()) {
    double *p = nullptr;
    CAPI_CHECK(hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch)));
    it = g_scratch.emplace(dev, p).first;
  }
  *out = it->second;
  return OLL_OK;
}
}  // namespace

struct oll_hier {
  Hier H;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Schedule vcycle;   // top-level V-cycle (graph-replayed)
  Schedule scratch;  // single stages
  Timing timing;
  double last_ms = 0;
  bool graphs = true;
};

namespace {

int hier_check(const oll_hier *h, int level) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  if (level < 0 || level >= h->H.geo.levels) return fail(OLL_EINVAL, "level out of range");
  return OLL_OK;
}

int run_stage(oll_hier *h, int level, int stage, int nu, double *linf, int box = -1) {
  std::string err;
  int rc = build_stage(h->H, level, stage, nu, h->scratch, err, box);
  if (!rc) rc = run_schedule(h->scratch, h->stream, nullptr, false, err);
  if (!rc && linf) {
    rc = read_norms(h->H, h->stream, err);
    if (!rc) {
      if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(OLL_EDEVICE, "sync");
      decode_norms(h->H);
      *linf = h->H.last_slots[size_t(h->H.slot_scratch())];
      if (std::isinf(*linf)) return fail(OLL_ENAN, "NAN check error in the smoother");
    }
  } else if (!rc) {
    if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(OLL_EDEVICE, "sync");
  }
  return rc ? fail(rc, err) : OLL_OK;
}

}  // namespace

extern "C" {

const char *oll_last_error(void) { return g_err.c_str(); }

void oll_problem_default(oll_problem *p) {
  std::memset(p, 0, sizeof *p);
  p->eq = OLL_EQ_PUNCTURES;
  p->num_bh = 1;
  p->L = 2.0;
  p->N = 30;
  p->levels = 1;
  p->order = 4;
  p->nu1 = 3;
  p->nu2 = 3;
  p->tol = 1e-8;
  p->robin_a = 0.0;
  p->robin_n = 1;
  p->ps_case = 1;
  p->ps_r0 = 120;
  p->ps_sigma = 8;
  p->ps_phi0 = 0.025;
  p->ps_a2 = 1;
  p->precision = OLL_PREC_FP64;
}

int oll_hier_create(const oll_problem *p, oll_hier **out) {
  g_err.clear();
  if (!p || !out) return fail(OLL_EINVAL, "null argument");
  *out = nullptr;
  kernels_init();
  oll_hier *h = new (std::nothrow) oll_hier();
  if (!h) return fail(OLL_ENOMEM, "host allocation failed");
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(OLL_EDEVICE, "hipStreamCreate failed");
  }
  (void)hipEventCreate(&h->ev0);
  (void)hipEventCreate(&h->ev1);
  std::string err;
  const int rc = hier_build(h->H, *p, h->stream, err);
  if (rc) {
    oll_hier_destroy(h);
    return fail(rc, err);
  }
  *out = h;
  return OLL_OK;
}

void oll_hier_destroy(oll_hier *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  release_schedule(h->vcycle);
  release_schedule(h->scratch);
  hier_release(h->H);
  for (hipEvent_t e : h->timing.ev) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int oll_num_levels(const oll_hier *h) { return h ? h->H.geo.levels : 0; }
int oll_mgrid_levels(const oll_hier *h) { return h ? h->H.geo.mgrid : 0; }
int oll_num_boxes(const oll_hier *h) { return h ? int(h->H.geo.boxes.size()) : 0; }

int oll_level_boxes(const oll_hier *h, int level, int *first, int *count) {
  int rc = hier_check(h, level);
  if (rc) return rc;
  if (first) *first = h->H.geo.level_first[size_t(level)];
  if (count) *count = h->H.geo.level_count[size_t(level)];
  return OLL_OK;
}

static void fill_desc(const Box &b, int colour_order, oll_box_desc *out) {
  out->colour_order = colour_order;
  out->level = b.level;
  out->order = b.order;
  out->ghost = b.R;
  for (int a = 0; a < 3; ++a) {
    out->n[a] = int(b.n[a]);
    out->origin[a] = b.lo[a];
    out->d[a] = b.d[a];
    out->centre[a] = b.c[a];
    out->end[a] = b.hi[a];
  }
}

int oll_box_info(const oll_hier *h, int ibox, oll_box_desc *out) {
  if (!h || !out) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->H.geo.boxes.size())) return fail(OLL_EINVAL, "box out of range");
  fill_desc(h->H.geo.boxes[size_t(ibox)], h->H.hbox[size_t(ibox)].flip, out);
  return OLL_OK;
}

size_t oll_device_bytes(const oll_hier *h) { return h ? h->H.bytes : 0; }

int oll_problem_boxes(const oll_problem *p, oll_box_desc *out, int cap, int *nbox, int *mgrid) {
  if (!p || !nbox) return fail(OLL_EINVAL, "null argument");
  PunctureSet ps;
  for (int i = 0; i < p->num_bh && i < OLL_MAX_BH; ++i) {
    ps.x.push_back(p->bh[i].x);
    ps.y.push_back(p->bh[i].y);
    ps.z.push_back(p->bh[i].z);
  }
  Hierarchy g;
  const std::string err =
      build_hierarchy(size_t(p->N), p->L, p->levels, p->order, ps, g, p->method != OLL_METHOD_APP);
  if (!err.empty()) return fail(err.find("enough") != std::string::npos ? OLL_EGRID : OLL_EINVAL, err);
  *nbox = int(g.boxes.size());
  if (mgrid) *mgrid = g.mgrid;
  for (int i = 0; i < *nbox && i < cap && out; ++i) {
    const Box &b = g.boxes[size_t(i)];
    fill_desc(b, colour_flip(g, b, colours_of(b.order, p->eq)), &out[i]);
  }
  return OLL_OK;
}

int oll_problem_footprint(const oll_problem *p, size_t *bytes) {
  if (!p || !bytes) return fail(OLL_EINVAL, "null argument");
  std::string err;
  const int rc = problem_footprint(*p, *bytes, err);
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_get_field(oll_hier *h, int ibox, int field, double *host) {
  if (!h || !host) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->H.hbox.size()) || field < 0 || field >= fields_of(h->H.prob.eq))
    return fail(OLL_EINVAL, "box or field out of range");
  const DevBox &b = h->H.hbox[size_t(ibox)];
  std::vector<double> st(size_t(b.size));
  CAPI_CHECK(hipMemcpyAsync(st.data(), b.f[field], st.size() * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  // colour blocks -> plain lexicographic order (i fastest)
  size_t q = 0;
  for (int k = 0; k < b.n[2]; ++k)
    for (int j = 0; j < b.n[1]; ++j)
      for (int i = 0; i < b.n[0]; ++i) host[q++] = st[size_t(pidx(b, i, j, k))];
  return OLL_OK;
}

int oll_set_field(oll_hier *h, int ibox, int field, const double *host) {
  if (!h || !host) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->H.hbox.size()) || field < 0 || field >= fields_of(h->H.prob.eq))
    return fail(OLL_EINVAL, "box or field out of range");
  const DevBox &b = h->H.hbox[size_t(ibox)];
  std::vector<double> st(size_t(b.size), 0.0);
  size_t q = 0;
  for (int k = 0; k < b.n[2]; ++k)
    for (int j = 0; j < b.n[1]; ++j)
      for (int i = 0; i < b.n[0]; ++i) st[size_t(pidx(b, i, j, k))] = host[q++];
  CAPI_CHECK(hipMemcpyAsync(b.f[field], st.data(), st.size() * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  if (field == F_RHO) {  // the caller's rho may be nonzero anywhere in the box
    std::string err;
    const int rc = hier_rho_whole(h->H, ibox, 1, h->stream, err);
    if (rc) return fail(rc, err);
  }
  return OLL_OK;
}

int oll_transfer(oll_hier *h, int from_level, int from_field, int to_level, int to_field,
                 int region, int method, int order, int all, int to_zero) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  const int nf = fields_of(h->H.prob.eq);
  if (from_field < 0 || from_field >= nf || to_field < 0 || to_field >= nf)
    return fail(OLL_EINVAL, "field out of range");
  if (from_level < 0 || to_level < 0) return fail(OLL_EINVAL, "level out of range");
  if (region < OLL_REGION_ALL || region > OLL_REGION_BOUNDARY)
    return fail(OLL_EINVAL, "region must be OLL_REGION_ALL, _INSIDE or _BOUNDARY");
  if (method != OLL_INTERP_LINEAR && method != OLL_INTERP_LAGRANGE)
    return fail(OLL_EINVAL, "method must be OLL_INTERP_LINEAR or OLL_INTERP_LAGRANGE");
  std::string err;
  const int rc = transfer_fields(h->H, h->stream, from_level, from_field, to_level, to_field, region,
                                 method, order, all, to_zero, err);
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_smooth(oll_hier *h, int level, int nu, double *linf) {
  int rc = hier_check(h, level);
  if (rc) return rc;
  double v = 0;
  rc = run_stage(h, level, STAGE_SMOOTH, nu, &v);
  if (linf) *linf = v;
  return rc;
}

int oll_residual(oll_hier *h, int level, double *linf) {
  int rc = hier_check(h, level);
  if (rc) return rc;
  double v = 0;
  rc = run_stage(h, level, STAGE_RESIDUAL, 0, &v);
  if (linf) *linf = v;
  return rc;
}

int oll_robin(oll_hier *h, int level) {
  int rc = hier_check(h, level);
  return rc ? rc : run_stage(h, level, STAGE_ROBIN, 0, nullptr);
}

int oll_robin_box(oll_hier *h, int ibox) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  if (ibox < 0 || ibox >= int(h->H.hbox.size())) return fail(OLL_EINVAL, "box out of range");
  return run_stage(h, h->H.hbox[size_t(ibox)].level, STAGE_ROBIN, 0, nullptr, ibox);
}

int oll_operator(oll_hier *h, int level, int derivative, double *linf) {
  int rc = hier_check(h, level);
  if (rc) return rc;
  if (derivative != 0 && derivative != 1) return fail(OLL_EINVAL, "derivative must be 0 or 1");
  double v = 0;
  rc = run_stage(h, level, derivative ? STAGE_DERIVATIVE : STAGE_RESIDUAL, 0, &v);
  if (linf) *linf = v;
  return rc;
}

int oll_boundary_operator(oll_hier *h, int level) {
  int rc = hier_check(h, level);
  return rc ? rc : run_stage(h, level, STAGE_BOUNDARY_OP, 0, nullptr);
}

int oll_fas(oll_hier *h, int level, double *linf) {
  int rc = hier_check(h, level);
  if (rc) return rc;
  if (level < 1) return fail(OLL_EINVAL, "FAS step needs level >= 1");
  std::string err;
  std::vector<Hier *> M{&h->H};
  if ((rc = build_vcycle_from(M, {level}, h->scratch, err))) return fail(rc, err);
  if ((rc = run_schedule(h->scratch, h->stream, nullptr, false, err))) return fail(rc, err);
  if ((rc = read_norms(h->H, h->stream, err))) return fail(rc, err);
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  decode_norms(h->H);
  const double v = h->H.last_slots[size_t(Hier::slot_post(level))];
  if (std::isinf(v)) return fail(OLL_ENAN, "NAN check error in the smoother");
  if (linf) *linf = v;
  return OLL_OK;
}

int oll_solve(oll_hier *h, oll_result *res, double *hist, int hist_cap) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  std::string err;
  int rc;
  if (h->H.prob.method == OLL_METHOD_APP) {
    // ApproximateID: the approximation is the answer (hier_reset writes it)
    CAPI_CHECK(hipEventRecord(h->ev0, h->stream));
    if ((rc = hier_reset(h->H, h->stream, err))) return fail(rc, err);
    CAPI_CHECK(hipEventRecord(h->ev1, h->stream));
    CAPI_CHECK(hipEventSynchronize(h->ev1));
    float ms = 0;
    CAPI_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    if (res) *res = oll_result{0, 0.0, 0, 1};
    return OLL_OK;
  }
  if (h->H.top() < 1) return fail(OLL_EINVAL, "hierarchy has a single level");
  if (h->vcycle.launches.empty()) {
    std::vector<Hier *> M{&h->H};
    if ((rc = build_vcycle(M, h->vcycle, err))) return fail(rc, err);
  }
  Hier &H = h->H;
  H.cycle = 0;
  H.res = H.res_prev = 1e12;
  H.best = 1e300;
  H.best_cycle = 0;
  H.done = false;
  H.status = 0;
  H.hist.clear();
  H.coarse_sweeps = 0;
  CAPI_CHECK(hipEventRecord(h->ev0, h->stream));
  while (!H.done) {
    if ((rc = run_schedule(h->vcycle, h->stream, &h->timing, h->graphs, err))) return fail(rc, err);
    if ((rc = finish_cycle(H, h->stream, err))) return fail(rc, err);
  }
  CAPI_CHECK(hipEventRecord(h->ev1, h->stream));
  CAPI_CHECK(hipEventSynchronize(h->ev1));
  float ms = 0;
  CAPI_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  const int nh = int(H.hist.size() / 2);
  for (int i = 0; hist && i < nh && i < hist_cap; ++i) {
    hist[2 * i] = H.hist[size_t(2 * i)];
    hist[2 * i + 1] = H.hist[size_t(2 * i + 1)];
  }
  if (res) {
    res->cycles = H.cycle;
    res->residual = H.res;
    res->hist_len = nh;
    res->converged = H.res <= H.eps_solv ? 1 : 0;
  }
  return OLL_OK;
}

int oll_last_norms(const oll_hier *h, double *pre, double *post, int nlevels) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  const std::vector<double> &s = h->H.last_slots;
  for (int k = 0; k < nlevels && k < h->H.geo.levels; ++k) {
    if (pre) pre[k] = s.empty() ? 0.0 : s[size_t(Hier::slot_pre(k))];
    if (post) post[k] = s.empty() ? 0.0 : s[size_t(Hier::slot_post(k))];
  }
  return OLL_OK;
}

int oll_last_solve_ms(const oll_hier *h, double *ms) {
  if (!h || !ms) return fail(OLL_EINVAL, "null argument");
  *ms = h->last_ms;
  return OLL_OK;
}

int oll_last_solve_coarse_sweeps(const oll_hier *h, long long *sweeps) {
  if (!h || !sweeps) return fail(OLL_EINVAL, "null argument");
  *sweeps = h->H.coarse_sweeps;
  return OLL_OK;
}

int oll_reset(oll_hier *h) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  std::string err;
  const int rc = hier_reset(h->H, h->stream, err);
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_set_kernel_timing(oll_hier *h, int on) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  h->timing.on = on != 0;
  h->timing.used = 0;
  h->timing.bytes = 0;
  h->timing.launch_bytes.clear();
  h->timing.launch_kind.clear();
  return OLL_OK;
}

int oll_kernel_timing(oll_hier *h, int *launches, double *total_ms, double *bytes) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  double tot = 0;
  for (size_t i = 0; i + 1 < h->timing.used; i += 2) {
    float ms = 0;
    CAPI_CHECK(hipEventElapsedTime(&ms, h->timing.ev[i], h->timing.ev[i + 1]));
    tot += ms;
  }
  if (launches) *launches = int(h->timing.used / 2);
  if (total_ms) *total_ms = tot;
  if (bytes) *bytes = h->timing.bytes;
  return OLL_OK;
}

int oll_set_graphs(oll_hier *h, int on) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  h->graphs = on != 0;
  return OLL_OK;
}

int oll_adm_mass(oll_hier *h, double *puncture, double *surface) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  std::string err;
  const int rc = adm_mass(h->H, h->stream, puncture, surface, err);
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_adm_mass_noalloc(oll_hier *h, double *puncture, double *surface) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  const int nb = int(h->H.hbox.size());
  if (surface && nb > kScratchBoxes)
    return fail(OLL_EINVAL, "oll_adm_mass_noalloc: " + std::to_string(nb) +
                                " boxes; the device scratch holds the surface samples of at most " +
                                std::to_string(kScratchBoxes));
  if (h->H.prob.num_bh > OLL_MAX_BH) return fail(OLL_EINVAL, "too many punctures");
  double *base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  AdmBuffers d;
  d.sph = base;
  d.ctr = base + kScratchSph;
  d.pxyz = d.ctr + 3 * kScratchBoxes;
  d.up = d.pxyz + 3 * OLL_MAX_BH;
  d.pbox = reinterpret_cast<int *>(d.up + OLL_MAX_BH);
  std::string err;
  rc = adm_mass_with(h->H, h->stream, d, puncture, surface, err);
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_puncture_approx(const oll_hier *h, double *u_at_p) {
  if (!h || !u_at_p) return fail(OLL_EINVAL, "null argument");
  puncture_approx(h->H, u_at_p);
  return OLL_OK;
}

int oll_interp(oll_hier *h, int field, int npts, const int *box, const double *xyz, int ord,
               double *out) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  if (npts < 0 || (npts > 0 && (!box || !xyz || !out))) return fail(OLL_EINVAL, "null argument");
  if (field < 0 || field >= fields_of(h->H.prob.eq)) return fail(OLL_EINVAL, "field out of range");
  if (ord < 1 || ord > 9) return fail(OLL_EINVAL, "interpolation order must be 1..9");
  const int nb = int(h->H.hbox.size());
  for (int q = 0; q < npts; ++q)
    if (box[q] < 0 || box[q] >= nb) return fail(OLL_EINVAL, "box index out of range");
  std::string err;
  const int rc = interp_points(h->H, h->stream, field, npts, box, xyz, ord, out, err);
  return rc ? fail(rc, err) : OLL_OK;
}

// ---------------------------------------------------------------------
// engine: many resident requests, V-cycles in lockstep
// ---------------------------------------------------------------------
}  // extern "C"

struct oll_engine {
  hipStream_t stream = nullptr;
  size_t budget = 0, in_use = 0;
  bool graphs = true;
  struct Member {
    std::unique_ptr<Hier> H;
    long long tag;
  };
  std::vector<Member> active;
  std::deque<Member> done;
  // one schedule per (nu1, nu2) class of the active members
  std::map<std::pair<int, int>, Schedule> sched;
  bool dirty = true;
  long long cycles_run = 0;
  Timing timing;
  // host wall time (ms) in schedule builds, cycle runs (launch to norms on
  // the host), admissions; rebuild count
  double build_ms = 0, run_ms = 0, admit_ms = 0;
  long long rebuilds = 0;
};

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
}  // namespace

extern "C" {

int oll_engine_create(size_t hbm_budget, int use_graphs, oll_engine **out) {
  g_err.clear();
  if (!out) return fail(OLL_EINVAL, "null argument");
  kernels_init();
  oll_engine *e = new (std::nothrow) oll_engine();
  if (!e) return fail(OLL_ENOMEM, "host allocation failed");
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return fail(OLL_EDEVICE, "hipStreamCreate failed");
  }
  e->budget = hbm_budget;
  e->graphs = use_graphs != 0;
  *out = e;
  return OLL_OK;
}

void oll_engine_destroy(oll_engine *e) {
  if (!e) return;
  (void)hipStreamSynchronize(e->stream);
  for (auto &kv : e->sched) release_schedule(kv.second);
  for (auto &m : e->active) hier_release(*m.H);
  for (auto &m : e->done) hier_release(*m.H);
  for (hipEvent_t ev : e->timing.ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(e->stream);
  delete e;
}

int oll_engine_admit(oll_engine *e, const oll_problem *p, long long tag) {
  if (!e || !p) return fail(OLL_EINVAL, "null argument");
  if (p->method != OLL_METHOD_VCYCLE) return fail(OLL_EINVAL, "the engine runs V-cycle requests");
  const double t0 = now_ms();
  struct AdmitClock {
    oll_engine *e;
    double t0;
    ~AdmitClock() { e->admit_ms += now_ms() - t0; }
  } clock{e, t0};
  std::string err;
  size_t need = 0;
  int rc = problem_footprint(*p, need, err);
  if (rc) return fail(rc, err);
  if (e->budget && e->in_use + need > e->budget) return fail(OLL_EBUSY, "HBM budget exhausted");
  oll_engine::Member m;
  m.H.reset(new (std::nothrow) Hier());
  m.tag = tag;
  if (!m.H) return fail(OLL_ENOMEM, "host allocation failed");
  if ((rc = hier_build(*m.H, *p, e->stream, err))) {
    hier_release(*m.H);
    return fail(rc, err);
  }
  e->in_use += m.H->bytes;
  e->active.push_back(std::move(m));
  e->dirty = true;
  return OLL_OK;
}

int oll_engine_admit_hier(oll_engine *e, oll_hier *h, long long tag) {
  if (!e || !h) return fail(OLL_EINVAL, "null argument");
  if (!h->H.slab) return fail(OLL_EINVAL, "empty hierarchy handle");
  if (h->H.prob.method != OLL_METHOD_VCYCLE) return fail(OLL_EINVAL, "the engine runs V-cycle requests");
  if (e->budget && e->in_use + h->H.bytes > e->budget) return fail(OLL_EBUSY, "HBM budget exhausted");
  CAPI_CHECK(hipStreamSynchronize(h->stream));
  release_schedule(h->vcycle);
  release_schedule(h->scratch);
  oll_engine::Member m;
  m.H.reset(new (std::nothrow) Hier());
  if (!m.H) return fail(OLL_ENOMEM, "host allocation failed");
  m.tag = tag;
  *m.H = std::move(h->H);
  h->H = Hier();  // the handle keeps no device memory
  std::string err;
  const int rc = hier_reset(*m.H, e->stream, err);
  e->in_use += m.H->bytes;
  e->active.push_back(std::move(m));
  e->dirty = true;
  return rc ? fail(rc, err) : OLL_OK;
}

int oll_engine_step(oll_engine *e, int *active_after) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  std::string err;
  int rc;
  if (e->dirty) {
    const double t0 = now_ms();
    std::map<std::pair<int, int>, std::vector<Hier *>> cls;
    for (auto &m : e->active) cls[{m.H->prob.nu1, m.H->prob.nu2}].push_back(m.H.get());
    // schedules keep their device buffers across rebuilds
    for (auto it = e->sched.begin(); it != e->sched.end();) {
      if (cls.count(it->first)) {
        it->second.clear();
        ++it;
      } else {
        release_schedule(it->second);
        it = e->sched.erase(it);
      }
    }
    for (auto &kv : cls)
      if ((rc = build_vcycle(kv.second, e->sched[kv.first], err))) return fail(rc, err);
    e->dirty = false;
    ++e->rebuilds;
    e->build_ms += now_ms() - t0;
  }
  const double t1 = now_ms();
  for (auto &kv : e->sched)
    if ((rc = run_schedule(kv.second, e->stream, &e->timing, e->graphs, err))) return fail(rc, err);
  for (auto &m : e->active)
    if ((rc = read_norms(*m.H, e->stream, err))) return fail(rc, err);
  CAPI_CHECK(hipStreamSynchronize(e->stream));
  e->run_ms += now_ms() - t1;
  ++e->cycles_run;
  std::vector<oll_engine::Member> still;
  for (auto &m : e->active) {
    std::string merr;
    apply_cycle(*m.H, merr);
    if (m.H->done) {
      e->done.push_back(std::move(m));
      e->dirty = true;
    } else {
      still.push_back(std::move(m));
    }
  }
  e->active = std::move(still);
  if (active_after) *active_after = int(e->active.size());
  return OLL_OK;
}

int oll_engine_pop_done(oll_engine *e, long long *tag, oll_result *res, double *adm) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  if (e->done.empty()) return 0;
  oll_engine::Member m = std::move(e->done.front());
  e->done.pop_front();
  Hier &H = *m.H;
  if (tag) *tag = m.tag;
  if (res) {
    res->cycles = H.cycle;
    res->residual = H.res;
    res->hist_len = int(H.hist.size() / 2);
    res->converged = H.status == 0 && H.res <= H.eps_solv ? 1 : 0;
  }
  int rc = H.status;
  if (!rc && adm) {
    std::string err;
    rc = adm_mass(H, e->stream, adm, nullptr, err);
    if (rc) g_err = err;
  }
  e->in_use -= std::min(e->in_use, H.bytes);
  hier_release(H);
  if (H.status) return fail(H.status, "NAN check error in the smoother");
  return rc ? rc : 1;
}

int oll_engine_pop_hier(oll_engine *e, long long *tag, oll_result *res, oll_hier **out) {
  if (!e || !out) return fail(OLL_EINVAL, "null argument");
  *out = nullptr;
  if (e->done.empty()) return 0;
  oll_hier *h = new (std::nothrow) oll_hier();
  if (!h) return fail(OLL_ENOMEM, "host allocation failed");
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(OLL_EDEVICE, "hipStreamCreate failed");
  }
  (void)hipEventCreate(&h->ev0);
  (void)hipEventCreate(&h->ev1);
  oll_engine::Member m = std::move(e->done.front());
  e->done.pop_front();
  if (tag) *tag = m.tag;
  if (res) {
    res->cycles = m.H->cycle;
    res->residual = m.H->res;
    res->hist_len = int(m.H->hist.size() / 2);
    res->converged = m.H->status == 0 && m.H->res <= m.H->eps_solv ? 1 : 0;
  }
  e->in_use -= std::min(e->in_use, m.H->bytes);
  const int status = m.H->status;
  h->H = std::move(*m.H);  // device memory changes owner; nothing is freed
  *m.H = Hier();
  *out = h;
  if (status) return fail(status, "NAN check error in the smoother");
  return 1;
}

int oll_engine_set_kernel_timing(oll_engine *e, int on) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  CAPI_CHECK(hipStreamSynchronize(e->stream));
  e->timing.on = on != 0;
  e->timing.used = 0;
  e->timing.bytes = 0;
  e->timing.launch_bytes.clear();
  e->timing.launch_kind.clear();
  return OLL_OK;
}

int oll_engine_kernel_launches(oll_engine *e, int cap, double *bytes, double *ms, int *kind,
                               int *n) {
  if (!e || !n) return fail(OLL_EINVAL, "null argument");
  CAPI_CHECK(hipStreamSynchronize(e->stream));
  const Timing &t = e->timing;
  const int total = int(std::min(t.launch_bytes.size(), t.used / 2));
  *n = total;
  for (int i = 0; i < total && i < cap; ++i) {
    float m = 0;
    CAPI_CHECK(hipEventElapsedTime(&m, t.ev[size_t(2 * i)], t.ev[size_t(2 * i + 1)]));
    if (bytes) bytes[i] = t.launch_bytes[size_t(i)];
    if (ms) ms[i] = m;
    if (kind) kind[i] = t.launch_kind[size_t(i)];
  }
  return OLL_OK;
}

int oll_engine_kernel_timing(oll_engine *e, int *launches, double *total_ms, double *bytes) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  CAPI_CHECK(hipStreamSynchronize(e->stream));
  double tot = 0;
  for (size_t i = 0; i + 1 < e->timing.used; i += 2) {
    float ms = 0;
    CAPI_CHECK(hipEventElapsedTime(&ms, e->timing.ev[i], e->timing.ev[i + 1]));
    tot += ms;
  }
  if (launches) *launches = int(e->timing.used / 2);
  if (total_ms) *total_ms = tot;
  if (bytes) *bytes = e->timing.bytes;
  return OLL_OK;
}

int oll_set_sweep_threads(int precision, int order, int threads) {
  if (precision != OLL_PREC_FP64 && precision != OLL_PREC_MIXED)
    return fail(OLL_EINVAL, "precision must be OLL_PREC_FP64 or OLL_PREC_MIXED");
  if (!set_sweep_threads(precision == OLL_PREC_MIXED, order, threads))
    return fail(OLL_EINVAL, "order must be 2, 4, 6 or 8 and threads a multiple of 64 in 64..512");
  return OLL_OK;
}

int oll_get_sweep_threads(int precision, int order) {
  return sweep_threads(precision == OLL_PREC_MIXED, order);
}

int oll_set_fused_sweeps(int on) {
  set_fused_sweeps(on != 0);
  return OLL_OK;
}

int oll_get_fused_sweeps(void) { return fused_sweeps() ? 1 : 0; }

int oll_set_shared_loads(int on) {
  set_shared_loads(on != 0);
  return OLL_OK;
}

int oll_get_shared_loads(void) { return shared_loads() ? 1 : 0; }

int oll_set_fused_shape(int order, int ty, int tz, int threads) {
  if (!set_fused_shape(order, ty, tz, threads))
    return fail(OLL_EINVAL,
                "order must be 2, 4, 6 or 8, ty and tz 1..64 and threads a multiple of 64 in 64..512");
  return OLL_OK;
}

int oll_get_fused_shape(int order, int *ty, int *tz, int *threads) {
  if (!(order == 2 || order == 4 || order == 6 || order == 8))
    return fail(OLL_EINVAL, "order must be 2, 4, 6 or 8");
  const FusedShape fs = fused_shape(order);
  if (ty) *ty = fs.ty;
  if (tz) *tz = fs.tz;
  if (threads) *threads = fs.threads;
  return OLL_OK;
}

int oll_engine_profile(const oll_engine *e, double *build_ms, double *run_ms, double *admit_ms,
                       long long *rebuilds, long long *cycles) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  if (build_ms) *build_ms = e->build_ms;
  if (run_ms) *run_ms = e->run_ms;
  if (admit_ms) *admit_ms = e->admit_ms;
  if (rebuilds) *rebuilds = e->rebuilds;
  if (cycles) *cycles = e->cycles_run;
  return OLL_OK;
}

int oll_engine_stats(const oll_engine *e, size_t *bytes_in_use, int *active, int *done_pending) {
  if (!e) return fail(OLL_EINVAL, "null handle");
  if (bytes_in_use) *bytes_in_use = e->in_use;
  if (active) *active = int(e->active.size());
  if (done_pending) *done_pending = int(e->done.size());
  return OLL_OK;
}

}  // extern "C"
