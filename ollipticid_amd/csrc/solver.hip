// Host driver: hierarchy object resident in HBM, the FAS V-cycle, the
// C ABI of include/olliptic_amd.h.
//
// The control flow of one FAS step follows elliptic.cc:1469-1968 (FASN);
// the outer loop follows elliptic.cc:1251-1342 (Vcycle).  Every grid
// operation is a kernel launch on the handle's stream; the host reads back
// one residual norm per V-cycle (the stopping test) and nothing else.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/olliptic_amd.h"
#include "device.hpp"
#include "geometry.hpp"
#include "kernels.hpp"

using namespace oll;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(OLL_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

double dulap_of(int order, const double idx2[3]) {
  const double s = idx2[0] + idx2[1] + idx2[2];
  switch (order) {
    case 2: return -2.0 * s;
    case 4: return -2.5 * s;
    case 6: return -490.0 * s / 180.0;
    default: return -205.0 * s / 72.0;
  }
}
}  // namespace

struct oll_hier {
  oll_problem prob{};
  Hierarchy geo;
  std::vector<DevBox> hbox;
  DevBox *dbox = nullptr;
  double *slab = nullptr;
  Fields F{};
  long long total = 0;
  int nfields = 0;
  // transfer pools
  int *ipool = nullptr;
  unsigned char *bpool = nullptr;
  double *dpool = nullptr;
  DevXfer *dxfer = nullptr;
  std::vector<LevelView> lv;
  std::vector<XferView> ru, rw, pv, gf;  // indexed by the fine level k
  unsigned long long *norms = nullptr;   // 2 per level + scratch
  int *iters = nullptr;
  int nslots = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  PunctureArgs pa{};
  double ipn = 1.0;
  double eps_solv = 0, eps_coarse = 0;
  double last_ms = 0;
  size_t device_bytes = 0;
  int coarse_max_iter = 100000;
  double coarse_omega = 1.0;
  int coarse_too_big = 0;
  long long coarse_sweeps = 0;  // coarsest-level sweeps of the last oll_solve
  std::vector<double> last_slots;  // norms of the last FAS step / V-cycle
  // optional per-launch timing of the smoother (events around each launch)
  int timing = 0;
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  double t_bytes = 0;
  std::vector<std::vector<double>> colour_pts;  // [level][colour] interior points
};

namespace {

int slot_pre(int k) { return 2 * k; }
int slot_post(int k) { return 2 * k + 1; }
unsigned long long *slot_ptr(oll_hier &h, int slot) { return h.norms + size_t(slot) * kNormSubSlots; }

LevelView make_level_view(const oll_hier &h, int level) {
  LevelView v{};
  v.nbox = h.geo.level_count[size_t(level)];
  int rows = 0;
  long long pts = 0, shell = 0;
  int maxrow = 0;
  for (int b = 0; b < v.nbox; ++b) {
    const int ib = h.geo.level_first[size_t(level)] + b;
    const DevBox &B = h.hbox[size_t(ib)];
    v.box[b] = ib;
    v.row0[b] = rows;
    v.pt0[b] = pts;
    v.shell0[b] = shell;
    rows += (B.n[1] - 2 * B.R) * (B.n[2] - 2 * B.R);
    pts += (long long)B.n[0] * B.n[1] * B.n[2];
    shell += (long long)B.n[0] * B.n[1] * B.n[2] -
             (long long)(B.n[0] - 2 * B.R) * (B.n[1] - 2 * B.R) * (B.n[2] - 2 * B.R);
    maxrow = std::max(maxrow, B.n[0] - 2 * B.R);
  }
  v.row0[v.nbox] = rows;
  v.pt0[v.nbox] = pts;
  v.shell0[v.nbox] = shell;
  v.max_row_len = maxrow;
  return v;
}

struct PoolBuilder {
  std::vector<int> ipool;
  std::vector<unsigned char> bpool;
  std::vector<double> dpool;
  std::vector<DevXfer> xfers;

  DevAxis add_axis(const AxisTable &t) {
    DevAxis a{};
    a.n = int(t.start.size());
    a.k = t.k;
    a.start_off = int(ipool.size());
    ipool.insert(ipool.end(), t.start.begin(), t.start.end());
    a.coinc_off = int(ipool.size());
    ipool.insert(ipool.end(), t.coinc.begin(), t.coinc.end());
    a.flags_off = int(bpool.size());
    bpool.insert(bpool.end(), t.flags.begin(), t.flags.end());
    a.w_off = int(dpool.size());
    dpool.insert(dpool.end(), t.w.begin(), t.w.end());
    return a;
  }

  // transfers into every box of `tl` from every box of `sl`; the XferView's
  // xfer pointer is fixed up after upload (stores the index meanwhile)
  XferView add(const oll_hier &h, int tl, int sl, int method, int ord) {
    XferView xv{};
    xv.tgt = make_level_view(h, tl);
    const int ft = h.geo.level_first[size_t(tl)], nt = h.geo.level_count[size_t(tl)];
    const int fs = h.geo.level_first[size_t(sl)], ns = h.geo.level_count[size_t(sl)];
    const long long base = (long long)xfers.size();
    for (int t = 0; t < nt; ++t) {
      xv.x0[t] = int(xfers.size() - base);
      for (int s = 0; s < ns; ++s) {
        DevXfer x{};
        x.target = ft + t;
        x.source = fs + s;
        const Box &T = h.geo.boxes[size_t(ft + t)], &S = h.geo.boxes[size_t(fs + s)];
        for (int a = 0; a < 3; ++a) x.ax[a] = add_axis(build_axis_table(T, S, a, method, ord));
        xfers.push_back(x);
      }
    }
    xv.x0[nt] = int(xfers.size() - base);
    xv.xfer = reinterpret_cast<const DevXfer *>(base);  // fixed up later
    return xv;
  }
};

int level_order(const oll_hier &h, int level) {
  return h.hbox[size_t(h.geo.level_first[size_t(level)])].order;
}

// ---------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------
void robin_level(oll_hier &h, int level, int mode) {
  const int f = h.geo.level_first[size_t(level)], n = h.geo.level_count[size_t(level)];
  for (int b = f; b < f + n; ++b) launch_robin(h.stream, h.dbox, b, h.hbox[size_t(b)], h.F, mode, h.ipn);
}

hipEvent_t timing_event(oll_hier &h) {
  if (h.tev_used == h.tev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h.tev.push_back(e);
  }
  return h.tev[h.tev_used++];
}

// Algorithmic bytes of one colour pass: read u, b, c, rho (+ d, e) and write
// u for every point of the colour, 40 B (56 B with the scalar field).
double smooth_bytes(const oll_hier &h, int level, int colour) {
  const double per = h.prob.eq == OLL_EQ_PUNCTURES_SCALAR ? 56.0 : 40.0;
  return per * h.colour_pts[size_t(level)][size_t(colour)];
}

void smooth_level(oll_hier &h, int level, int nu, unsigned long long *slot) {
  const int order = level_order(h, level);
  const int nc = order / 2 + 1;
  for (int s = 0; s < nu; ++s)
    for (int c = 0; c < nc; ++c) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (h.timing) {
        e0 = timing_event(h);
        e1 = timing_event(h);
        if (e0) (void)hipEventRecord(e0, h.stream);
      }
      launch_smooth(h.stream, h.lv[size_t(level)], h.dbox, h.F, order, h.prob.eq, c, slot);
      if (h.timing && e1) {
        (void)hipEventRecord(e1, h.stream);
        h.t_bytes += smooth_bytes(h, level, c);
      }
    }
}

void fas(oll_hier &h, int k) {
  const int M = h.geo.mgrid;
  const bool mg = k <= M;
  // pre-smoothing
  if (mg) robin_level(h, k, ROBIN_UPDATE);
  smooth_level(h, k, h.prob.nu1, slot_ptr(h, slot_pre(k)));
  // fine defect W = S - R (interior; Robin shell on multigrid levels)
  launch_residual(h.stream, h.lv[size_t(k)], h.dbox, h.F, level_order(h, k), h.prob.eq, 1, nullptr);
  if (mg) robin_level(h, k, ROBIN_DEFECT_W);
  // I u, then the coarse FAS source and ww = u
  launch_restrict_u(h.stream, h.ru[size_t(k)], h.dbox, h.F);
  launch_restrict_defect(h.stream, h.rw[size_t(k)], h.dbox, h.F, level_order(h, k - 1), h.prob.eq,
                         mg ? 1 : 0, h.ipn);
  if (k == 1) {
    const int b0 = h.geo.level_first[0];
    if (!launch_coarse_solve(h.stream, h.dbox, b0, h.hbox[size_t(b0)], h.F, h.prob.eq, h.ipn,
                             h.eps_coarse, h.coarse_omega, h.coarse_max_iter,
                             slot_ptr(h, slot_pre(0)), h.iters))
      h.coarse_too_big = 1;
  } else {
    fas(h, k - 1);
  }
  // coarse correction, ghost shell, post-smoothing
  launch_prolong(h.stream, h.pv[size_t(k)], h.dbox, h.F, mg ? REG_ALL : REG_INSIDE);
  if (!mg) launch_ghost_fill(h.stream, h.gf[size_t(k)], h.dbox, h.F);
  if (mg) robin_level(h, k, ROBIN_UPDATE);
  smooth_level(h, k, h.prob.nu2, slot_ptr(h, slot_post(k)));
}

void zero_norms(oll_hier &h) { launch_zero_u64(h.stream, h.norms, h.nslots * kNormSubSlots); }

double slot_value(unsigned long long v) {
  double d;
  std::memcpy(&d, &v, sizeof d);
  return d;
}

int read_slots(oll_hier &h, std::vector<double> &out) {
  std::vector<unsigned long long> raw(size_t(h.nslots) * kNormSubSlots);
  HIP_CHECK(hipMemcpyAsync(raw.data(), h.norms, raw.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, h.stream));
  HIP_CHECK(hipStreamSynchronize(h.stream));
  if (h.coarse_too_big) return fail(OLL_EINVAL, "coarsest box exceeds the LDS of one workgroup");
  out.assign(size_t(h.nslots), 0.0);
  for (int sl = 0; sl < h.nslots; ++sl) {
    unsigned long long m = 0;
    for (int q = 0; q < kNormSubSlots; ++q) m = std::max(m, raw[size_t(sl) * kNormSubSlots + q]);
    out[size_t(sl)] = slot_value(m);
  }
  return OLL_OK;
}

int build(oll_hier &h) {
  const oll_problem &p = h.prob;
  PunctureSet ps;
  for (int i = 0; i < p.num_bh; ++i) {
    ps.x.push_back(p.bh[i].x);
    ps.y.push_back(p.bh[i].y);
    ps.z.push_back(p.bh[i].z);
  }
  const std::string err = build_hierarchy(size_t(p.N), p.L, p.levels, p.order, ps, h.geo);
  if (!err.empty()) return fail(err.find("enough") != std::string::npos ? OLL_EGRID : OLL_EINVAL, err);
  for (int l = 0; l < h.geo.levels; ++l)
    if (h.geo.level_count[size_t(l)] > kMaxLevelBoxes)
      return fail(OLL_EINVAL, "too many boxes on one level");
  // device boxes
  long long off = 0;
  for (const Box &b : h.geo.boxes) {
    DevBox d{};
    for (int a = 0; a < 3; ++a) {
      d.n[a] = int(b.n[a]);
      d.lo[a] = b.lo[a];
      d.d[a] = b.d[a];
      d.idx[a] = 1.0 / b.d[a];
      d.idx2[a] = d.idx[a] * d.idx[a];
    }
    d.R = b.R;
    d.order = b.order;
    d.level = b.level;
    d.off = off;
    d.dulap = dulap_of(b.order, d.idx2);
    off += (long long)b.npts();
    h.hbox.push_back(d);
  }
  h.total = off;
  h.nfields = p.eq == OLL_EQ_PUNCTURES_SCALAR ? NFIELDS : NFIELDS;
  const size_t slab_bytes = size_t(h.total) * sizeof(double) * size_t(NFIELDS);
  hipError_t e = hipMalloc(&h.slab, slab_bytes);
  if (e != hipSuccess) return fail(OLL_ENOMEM, std::string("hipMalloc slab: ") + hipGetErrorString(e));
  HIP_CHECK(hipMemsetAsync(h.slab, 0, slab_bytes, h.stream));
  for (int f = 0; f < NFIELDS; ++f) h.F.f[f] = h.slab + (size_t)f * size_t(h.total);
  HIP_CHECK(hipMalloc(&h.dbox, h.hbox.size() * sizeof(DevBox)));
  HIP_CHECK(hipMemcpyAsync(h.dbox, h.hbox.data(), h.hbox.size() * sizeof(DevBox),
                           hipMemcpyHostToDevice, h.stream));
  h.device_bytes = slab_bytes + h.hbox.size() * sizeof(DevBox);

  // level views and transfer tables
  for (int l = 0; l < h.geo.levels; ++l) {
    h.lv.push_back(make_level_view(h, l));
    // the smoother maps threads to points with 24-bit exact float division
    const LevelView &v = h.lv.back();
    const int nc = level_order(h, l) / 2 + 1;
    if ((long long)v.row0[v.nbox] * ((v.max_row_len + nc - 1) / nc) >= (1LL << 24))
      return fail(OLL_EINVAL, "a level exceeds 2^24 points per colour");
  }
  for (int l = 0; l < h.geo.levels; ++l) {
    const int nc = level_order(h, l) / 2 + 1;
    std::vector<double> cnt(size_t(nc), 0.0);
    for (int b = 0; b < h.geo.level_count[size_t(l)]; ++b) {
      const DevBox &B = h.hbox[size_t(h.geo.level_first[size_t(l)] + b)];
      for (int k = B.R; k < B.n[2] - B.R; ++k)
        for (int j = B.R; j < B.n[1] - B.R; ++j)
          for (int i = B.R; i < B.n[0] - B.R; ++i) cnt[size_t((i + j + k) % nc)] += 1.0;
    }
    h.colour_pts.push_back(cnt);
  }
  PoolBuilder pb;
  h.ru.resize(size_t(h.geo.levels));
  h.rw.resize(size_t(h.geo.levels));
  h.pv.resize(size_t(h.geo.levels));
  h.gf.resize(size_t(h.geo.levels));
  for (int k = 1; k < h.geo.levels; ++k) {
    h.ru[size_t(k)] = pb.add(h, k - 1, k, 1, 3);
    h.rw[size_t(k)] = pb.add(h, k - 1, k, 1, 2);
    h.pv[size_t(k)] = pb.add(h, k, k - 1, 0, 1);
    h.gf[size_t(k)] = pb.add(h, k, k - 1, 1, p.order + 1);
  }
  auto upload = [&](auto **dst, const auto &vec) -> int {
    using T = typename std::decay<decltype(vec)>::type::value_type;
    const size_t bytes = std::max<size_t>(1, vec.size()) * sizeof(T);
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(dst), bytes));
    if (!vec.empty())
      HIP_CHECK(hipMemcpyAsync(*dst, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, h.stream));
    h.device_bytes += bytes;
    return OLL_OK;
  };
  int rc;
  if ((rc = upload(&h.ipool, pb.ipool)) != OLL_OK) return rc;
  if ((rc = upload(&h.bpool, pb.bpool)) != OLL_OK) return rc;
  if ((rc = upload(&h.dpool, pb.dpool)) != OLL_OK) return rc;
  if ((rc = upload(&h.dxfer, pb.xfers)) != OLL_OK) return rc;
  for (auto *vec : {&h.ru, &h.rw, &h.pv, &h.gf})
    for (XferView &x : *vec) {
      x.xfer = h.dxfer + reinterpret_cast<long long>(x.xfer);
      x.ipool = h.ipool;
      x.bpool = h.bpool;
      x.dpool = h.dpool;
    }
  h.nslots = 2 * h.geo.levels + 4;
  HIP_CHECK(hipMalloc(&h.norms, size_t(h.nslots) * kNormSubSlots * sizeof(unsigned long long)));
  HIP_CHECK(hipMalloc(&h.iters, 4 * sizeof(int)));
  zero_norms(h);
  {
    // over-relaxation of the coarsest-level loop, tuned for the 13^3-21^3
    // coarsest boxes the hierarchy produces
    const int n = h.hbox[size_t(h.geo.level_first[0])].n[0];
    h.coarse_omega = 2.0 - 1.6 / double(n - 1);
  }

  // coefficients and initial state (operator.cc:752-868, boundary.cc:8-84)
  h.pa.n = p.num_bh;
  h.pa.eq = p.eq;
  for (int i = 0; i < p.num_bh; ++i) {
    const oll_puncture &q = p.bh[i];
    h.pa.p[i] = Puncture{q.m, q.x, q.y, q.z, q.px, q.py, q.pz, q.sx, q.sy, q.sz};
  }
  h.pa.ps_case = p.ps_case;
  h.pa.ps_r0 = p.ps_r0;
  h.pa.ps_isigma = 1 / p.ps_sigma;
  h.pa.ps_phi0 = (p.ps_a2 * p.ps_a2 / (1 + p.ps_a2 * p.ps_a2)) * p.ps_phi0;
  h.pa.ps_d1 = M_PI * h.pa.ps_phi0 * h.pa.ps_phi0 * h.pa.ps_isigma * h.pa.ps_isigma;
  h.pa.ps_e1 = 1 / (16 * p.ps_a2);
  h.pa.ps_e2 = 4 * std::sqrt(M_PI / 3);
  h.pa.ps_de1 = 4 * M_PI * h.pa.ps_phi0 * h.pa.ps_phi0 * h.pa.ps_isigma * h.pa.ps_isigma;
  for (const DevBox &b : h.hbox) launch_coefficients_box(h.stream, b, h.F, h.pa, p.robin_a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h.stream));
  return OLL_OK;
}

void release(oll_hier *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  (void)hipFree(h->slab);
  (void)hipFree(h->dbox);
  (void)hipFree(h->ipool);
  (void)hipFree(h->bpool);
  (void)hipFree(h->dpool);
  (void)hipFree(h->dxfer);
  (void)hipFree(h->norms);
  (void)hipFree(h->iters);
  for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int check_level(const oll_hier *h, int level) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  if (level < 0 || level >= h->geo.levels) return fail(OLL_EINVAL, "level out of range");
  return OLL_OK;
}

}  // namespace

// =====================================================================
// C ABI
// =====================================================================
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
  if (p->num_bh < 1 || p->num_bh > OLL_MAX_BH) return fail(OLL_EINVAL, "NumBH out of range");
  if (!(p->order == 2 || p->order == 4 || p->order == 6 || p->order == 8))
    return fail(OLL_EINVAL, "order must be [2,4,6,8].");
  if (p->levels < 1) return fail(OLL_EINVAL, "levels must be >= 1");
  if (p->N < 3) return fail(OLL_EINVAL, "N too small");
  if (p->nu1 < 0 || p->nu2 < 0) return fail(OLL_EINVAL, "negative sweep count");
  if (p->robin_n < 1) return fail(OLL_EINVAL, "Robin fall-off power must be >= 1");
  if (p->eq == OLL_EQ_PUNCTURES)
    for (int i = 0; i < p->num_bh; ++i)
      if (p->bh[i].m <= 0) return fail(OLL_EINVAL, "Wrong punture mass parameters");
  if (p->precision != OLL_PREC_FP64) return fail(OLL_EINVAL, "only OLL_PREC_FP64 is available");
  oll_hier *h = new (std::nothrow) oll_hier();
  if (!h) return fail(OLL_ENOMEM, "host allocation failed");
  h->prob = *p;
  h->ipn = 1.0 / double(p->robin_n);
  h->eps_coarse = 2 * DBL_EPSILON;
  h->eps_solv = p->tol < h->eps_coarse ? h->eps_coarse : p->tol;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(OLL_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  (void)hipEventCreate(&h->ev0);
  (void)hipEventCreate(&h->ev1);
  const int rc = build(*h);
  if (rc != OLL_OK) {
    const std::string keep = g_err;
    release(h);
    g_err = keep;
    return rc;
  }
  *out = h;
  return OLL_OK;
}

void oll_hier_destroy(oll_hier *h) { release(h); }

int oll_num_levels(const oll_hier *h) { return h ? h->geo.levels : 0; }
int oll_mgrid_levels(const oll_hier *h) { return h ? h->geo.mgrid : 0; }
int oll_num_boxes(const oll_hier *h) { return h ? int(h->geo.boxes.size()) : 0; }

int oll_level_boxes(const oll_hier *h, int level, int *first, int *count) {
  int rc = check_level(h, level);
  if (rc) return rc;
  if (first) *first = h->geo.level_first[size_t(level)];
  if (count) *count = h->geo.level_count[size_t(level)];
  return OLL_OK;
}

int oll_box_info(const oll_hier *h, int ibox, oll_box_desc *out) {
  if (!h || !out) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->geo.boxes.size())) return fail(OLL_EINVAL, "box out of range");
  const Box &b = h->geo.boxes[size_t(ibox)];
  out->level = b.level;
  out->order = b.order;
  out->ghost = b.R;
  for (int a = 0; a < 3; ++a) {
    out->n[a] = int(b.n[a]);
    out->origin[a] = b.lo[a];
    out->d[a] = b.d[a];
    out->centre[a] = b.c[a];
  }
  return OLL_OK;
}

size_t oll_device_bytes(const oll_hier *h) { return h ? h->device_bytes : 0; }

int oll_problem_boxes(const oll_problem *p, oll_box_desc *out, int cap, int *nbox, int *mgrid) {
  if (!p || !nbox) return fail(OLL_EINVAL, "null argument");
  PunctureSet ps;
  for (int i = 0; i < p->num_bh && i < OLL_MAX_BH; ++i) {
    ps.x.push_back(p->bh[i].x);
    ps.y.push_back(p->bh[i].y);
    ps.z.push_back(p->bh[i].z);
  }
  Hierarchy g;
  const std::string err = build_hierarchy(size_t(p->N), p->L, p->levels, p->order, ps, g);
  if (!err.empty()) return fail(err.find("enough") != std::string::npos ? OLL_EGRID : OLL_EINVAL, err);
  *nbox = int(g.boxes.size());
  if (mgrid) *mgrid = g.mgrid;
  for (int i = 0; i < *nbox && i < cap && out; ++i) {
    const Box &b = g.boxes[size_t(i)];
    out[i].level = b.level;
    out[i].order = b.order;
    out[i].ghost = b.R;
    for (int a = 0; a < 3; ++a) {
      out[i].n[a] = int(b.n[a]);
      out[i].origin[a] = b.lo[a];
      out[i].d[a] = b.d[a];
      out[i].centre[a] = b.c[a];
    }
  }
  return OLL_OK;
}

int oll_problem_footprint(const oll_problem *p, size_t *bytes) {
  if (!p || !bytes) return fail(OLL_EINVAL, "null argument");
  PunctureSet ps;
  for (int i = 0; i < p->num_bh && i < OLL_MAX_BH; ++i) {
    ps.x.push_back(p->bh[i].x);
    ps.y.push_back(p->bh[i].y);
    ps.z.push_back(p->bh[i].z);
  }
  Hierarchy g;
  const std::string err = build_hierarchy(size_t(p->N), p->L, p->levels, p->order, ps, g);
  if (!err.empty()) return fail(OLL_EGRID, err);
  size_t pts = 0, tables = 0;
  for (const Box &b : g.boxes) pts += b.npts();
  for (size_t i = 0; i + 1 < g.boxes.size(); ++i)
    tables += 4 * 3 * (g.boxes[i + 1].n[0] * (size_t(p->order) + 4) * 10);
  *bytes = pts * sizeof(double) * size_t(NFIELDS) + tables;
  return OLL_OK;
}

int oll_get_field(oll_hier *h, int ibox, int field, double *host) {
  if (!h || !host) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->hbox.size()) || field < 0 || field >= NFIELDS)
    return fail(OLL_EINVAL, "box or field out of range");
  const DevBox &b = h->hbox[size_t(ibox)];
  const size_t n = size_t(b.n[0]) * size_t(b.n[1]) * size_t(b.n[2]);
  HIP_CHECK(hipMemcpyAsync(host, h->F.f[field] + b.off, n * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return OLL_OK;
}

int oll_set_field(oll_hier *h, int ibox, int field, const double *host) {
  if (!h || !host) return fail(OLL_EINVAL, "null argument");
  if (ibox < 0 || ibox >= int(h->hbox.size()) || field < 0 || field >= NFIELDS)
    return fail(OLL_EINVAL, "box or field out of range");
  const DevBox &b = h->hbox[size_t(ibox)];
  const size_t n = size_t(b.n[0]) * size_t(b.n[1]) * size_t(b.n[2]);
  HIP_CHECK(hipMemcpyAsync(h->F.f[field] + b.off, host, n * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return OLL_OK;
}

int oll_smooth(oll_hier *h, int level, int nu, double *linf) {
  int rc = check_level(h, level);
  if (rc) return rc;
  const int scratch = h->nslots - 1;
  zero_norms(*h);
  smooth_level(*h, level, nu, slot_ptr(*h, scratch));
  HIP_CHECK(hipGetLastError());
  std::vector<double> s;
  if ((rc = read_slots(*h, s))) return rc;
  if (std::isinf(s[size_t(scratch)])) return fail(OLL_ENAN, "NAN check error in the smoother");
  if (linf) *linf = s[size_t(scratch)];
  return OLL_OK;
}

int oll_residual(oll_hier *h, int level, double *linf) {
  int rc = check_level(h, level);
  if (rc) return rc;
  const int scratch = h->nslots - 1;
  zero_norms(*h);
  launch_residual(h->stream, h->lv[size_t(level)], h->dbox, h->F, level_order(*h, level), h->prob.eq,
                  0, slot_ptr(*h, scratch));
  HIP_CHECK(hipGetLastError());
  std::vector<double> s;
  if ((rc = read_slots(*h, s))) return rc;
  if (linf) *linf = s[size_t(scratch)];
  return OLL_OK;
}

int oll_robin(oll_hier *h, int level) {
  int rc = check_level(h, level);
  if (rc) return rc;
  robin_level(*h, level, ROBIN_UPDATE);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return OLL_OK;
}

int oll_boundary_operator(oll_hier *h, int level) {
  int rc = check_level(h, level);
  if (rc) return rc;
  robin_level(*h, level, ROBIN_OPERATOR_R);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return OLL_OK;
}

int oll_fas(oll_hier *h, int level, double *linf) {
  int rc = check_level(h, level);
  if (rc) return rc;
  if (level < 1) return fail(OLL_EINVAL, "FAS step needs level >= 1");
  zero_norms(*h);
  fas(*h, level);
  HIP_CHECK(hipGetLastError());
  std::vector<double> s;
  if ((rc = read_slots(*h, s))) return rc;
  h->last_slots = s;
  if (std::isinf(s[size_t(slot_post(level))])) return fail(OLL_ENAN, "NAN check error in the smoother");
  if (linf) *linf = s[size_t(slot_post(level))];
  return OLL_OK;
}

int oll_solve(oll_hier *h, oll_result *res, double *hist, int hist_cap) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  const int top = h->geo.levels - 1;
  if (top < 1) return fail(OLL_EINVAL, "hierarchy has a single level");
  double r = 1e12, r_prev;
  int cycle = 0, nh = 0;
  std::vector<double> s;
  h->coarse_sweeps = 0;
  HIP_CHECK(hipEventRecord(h->ev0, h->stream));
  do {
    r_prev = r;
    zero_norms(*h);
    fas(*h, top);
    HIP_CHECK(hipGetLastError());
    int rc = read_slots(*h, s);
    if (rc) return rc;
    {
      int it = 0;
      HIP_CHECK(hipMemcpy(&it, h->iters, sizeof(int), hipMemcpyDeviceToHost));
      h->coarse_sweeps += it;
    }
    for (int k = top; k >= 0; --k) {
      if (hist && nh < hist_cap) {
        hist[2 * nh] = cycle;
        hist[2 * nh + 1] = s[size_t(slot_pre(k))];
      }
      ++nh;
    }
    r = s[size_t(slot_post(top))];
    h->last_slots = s;
    ++cycle;
    for (int k = 0; k <= top; ++k)
      if (std::isinf(s[size_t(slot_pre(k))]) || std::isinf(s[size_t(slot_post(k))]))
        return fail(OLL_ENAN, "NAN check error in the smoother");
    if (h->prob.max_cycles > 0 && cycle >= h->prob.max_cycles) break;
  } while (r > h->eps_solv && r < r_prev);
  HIP_CHECK(hipEventRecord(h->ev1, h->stream));
  HIP_CHECK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  if (res) {
    res->cycles = cycle;
    res->residual = r;
    res->hist_len = nh;
  }
  return OLL_OK;
}

int oll_reset(oll_hier *h) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  for (const DevBox &b : h->hbox) launch_coefficients_box(h->stream, b, h->F, h->pa, h->prob.robin_a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return OLL_OK;
}

int oll_set_kernel_timing(oll_hier *h, int on) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  HIP_CHECK(hipStreamSynchronize(h->stream));
  h->timing = on ? 1 : 0;
  h->tev_used = 0;
  h->t_bytes = 0;
  return OLL_OK;
}

int oll_kernel_timing(oll_hier *h, int *launches, double *total_ms, double *bytes) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  HIP_CHECK(hipStreamSynchronize(h->stream));
  double tot = 0;
  for (size_t i = 0; i + 1 < h->tev_used; i += 2) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, h->tev[i], h->tev[i + 1]));
    tot += ms;
  }
  if (launches) *launches = int(h->tev_used / 2);
  if (total_ms) *total_ms = tot;
  if (bytes) *bytes = h->t_bytes;
  return OLL_OK;
}

int oll_last_solve_ms(const oll_hier *h, double *ms) {
  if (!h || !ms) return fail(OLL_EINVAL, "null argument");
  *ms = h->last_ms;
  return OLL_OK;
}

int oll_last_norms(const oll_hier *h, double *pre, double *post, int nlevels) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  for (int k = 0; k < nlevels && k < h->geo.levels; ++k) {
    const bool have = !h->last_slots.empty();
    if (pre) pre[k] = have ? h->last_slots[size_t(slot_pre(k))] : 0.0;
    if (post) post[k] = have ? h->last_slots[size_t(slot_post(k))] : 0.0;
  }
  return OLL_OK;
}

int oll_last_solve_coarse_sweeps(const oll_hier *h, long long *sweeps) {
  if (!h || !sweeps) return fail(OLL_EINVAL, "null argument");
  *sweeps = h->coarse_sweeps;
  return OLL_OK;
}

int oll_adm_mass(oll_hier *h, double *puncture, double *surface) {
  if (!h) return fail(OLL_EINVAL, "null handle");
  const oll_problem &p = h->prob;
  const int top = h->geo.levels - 1;
  const int nb = int(h->hbox.size());
  const int nt = 4 * 25 + 1;
  // u at each puncture from its finest box (Lagrange order 4);
  // multigrid.cc:97-116 falls back to the level's first box
  std::vector<int> pbox(size_t(p.num_bh));
  std::vector<double> pxyz(size_t(3 * p.num_bh));
  for (int i = 0; i < p.num_bh; ++i) {
    pbox[size_t(i)] = h->geo.level_first[size_t(top)] + (i < h->geo.level_count[size_t(top)] ? i : 0);
    pxyz[size_t(3 * i)] = p.bh[i].x;
    pxyz[size_t(3 * i + 1)] = p.bh[i].y;
    pxyz[size_t(3 * i + 2)] = p.bh[i].z;
  }
  std::vector<double> ctr(size_t(3 * nb));
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < 3; ++a) ctr[size_t(3 * b + a)] = h->geo.boxes[size_t(b)].c[a];
  int *d_pbox = nullptr;
  double *d_pxyz = nullptr, *d_up = nullptr, *d_ctr = nullptr, *d_sph = nullptr;
  HIP_CHECK(hipMalloc(&d_pbox, pbox.size() * sizeof(int)));
  HIP_CHECK(hipMalloc(&d_pxyz, pxyz.size() * sizeof(double)));
  HIP_CHECK(hipMalloc(&d_up, pbox.size() * sizeof(double)));
  HIP_CHECK(hipMalloc(&d_ctr, ctr.size() * sizeof(double)));
  HIP_CHECK(hipMalloc(&d_sph, size_t(nb) * nt * nt * sizeof(double)));
  HIP_CHECK(hipMemcpyAsync(d_pbox, pbox.data(), pbox.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(d_pxyz, pxyz.data(), pxyz.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipMemcpyAsync(d_ctr, ctr.data(), ctr.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  launch_interp_points(h->stream, h->dbox, h->F, F_U, d_pbox, d_pxyz, p.num_bh, 4, d_up);
  launch_adm_sphere(h->stream, h->dbox, nb, h->F, d_ctr, d_sph, nt);
  std::vector<double> up(pbox.size()), sph(size_t(nb) * nt * nt);
  HIP_CHECK(hipMemcpyAsync(up.data(), d_up, up.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipMemcpyAsync(sph.data(), d_sph, sph.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  (void)hipFree(d_pbox);
  (void)hipFree(d_pxyz);
  (void)hipFree(d_up);
  (void)hipFree(d_ctr);
  (void)hipFree(d_sph);
  // elliptic.cc:2448-2469
  for (int i = 0; i < p.num_bh && puncture; ++i) {
    double bhmj = 0;
    for (int j = 0; j < p.num_bh; ++j)
      if (j != i) {
        const double dx = p.bh[i].x - p.bh[j].x, dy = p.bh[i].y - p.bh[j].y, dz = p.bh[i].z - p.bh[j].z;
        bhmj += 0.5 * p.bh[j].m / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    puncture[i] = p.bh[i].m * (1 + bhmj + up[size_t(i)]);
  }
  // Boole's rule over theta and phi, elliptic.cc:2560-2604
  if (surface) {
    const double dtheta = M_PI / (nt - 1), dphi = 2 * M_PI / (nt - 1);
    for (int b = 0; b < nb; ++b) {
      const double *S = &sph[size_t(b) * nt * nt];
      const Box &B = h->geo.boxes[size_t(b)];
      const double r0 = 0.375 * std::min(B.len[0], std::min(B.len[1], B.len[2]));
      double adm = 0;
      for (int i = 0; i < nt - 3; i += 4) {
        double I[5] = {0, 0, 0, 0, 0};
        for (int j = 0; j < nt - 3; j += 4)
          for (int q = 0; q < 5; ++q) {
            const double *row = S + size_t(i + q) * nt;
            I[q] += 7 * row[j] + 32 * row[j + 1] + 12 * row[j + 2] + 32 * row[j + 3] + 7 * row[j + 4];
          }
        for (int q = 0; q < 5; ++q) I[q] *= 2 * dphi / 45;
        adm += 7 * I[0] + 32 * I[1] + 12 * I[2] + 32 * I[3] + 7 * I[4];
      }
      adm *= -r0 * r0 * dtheta / (45 * M_PI);
      surface[b] = adm;
    }
  }
  return OLL_OK;
}

}  // extern "C"
