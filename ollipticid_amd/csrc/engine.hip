// Host engine; see engine.hpp.
#include "engine.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

// host arithmetic that must equal the reference's (coordinates, the
// approximation's puncture values) is rounded operation by operation
#pragma clang fp contract(off)

namespace oll {

namespace {

#define ENG_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                             \
      return OLL_EDEVICE;                                                                  \
    }                                                                                      \
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

long long npts(const DevBox &b) { return (long long)b.n[0] * b.n[1] * b.n[2]; }
long long shell_pts(const DevBox &b) {
  return npts(b) - (long long)(b.n[0] - 2 * b.R) * (b.n[1] - 2 * b.R) * (b.n[2] - 2 * b.R);
}
int interior_rows(const DevBox &b) { return (b.n[1] - 2 * b.R) * (b.n[2] - 2 * b.R); }
int face_columns(const DevBox &b) {
  const int ni = b.n[0] - 2 * b.R, nj = b.n[1] - 2 * b.R, nk = b.n[2] - 2 * b.R;
  return 2 * (nj * nk + ni * nk + ni * nj);
}

int validate(const oll_problem &p, std::string &err) {
  if (p.num_bh < 1 || p.num_bh > OLL_MAX_BH) return err = "NumBH out of range", OLL_EINVAL;
  if (!(p.order == 2 || p.order == 4 || p.order == 6 || p.order == 8))
    return err = "order must be [2,4,6,8].", OLL_EINVAL;
  if (p.levels < 1) return err = "levels must be >= 1", OLL_EINVAL;
  if (p.N < 3) return err = "N too small", OLL_EINVAL;
  if (p.nu1 < 0 || p.nu2 < 0) return err = "negative sweep count", OLL_EINVAL;
  if (p.robin_n < 1) return err = "Robin fall-off power must be >= 1", OLL_EINVAL;
  if (p.eq < OLL_EQ_PUNCTURES || p.eq > OLL_EQ_PUNCTURES_EM) return err = "unknown equation", OLL_EINVAL;
  if (p.eq == OLL_EQ_PUNCTURES_EM && p.precision != OLL_PREC_FP64)
    return err = "Punctures_EM runs in FP64 only", OLL_EINVAL;
  if (p.eq == OLL_EQ_PUNCTURES)
    for (int i = 0; i < p.num_bh; ++i)
      if (p.bh[i].m <= 0) return err = "Wrong punture mass parameters", OLL_EINVAL;
  if (p.stop_rule != OLL_STOP_REFERENCE && p.stop_rule != OLL_STOP_TOLERANCE)
    return err = "unknown stopping rule", OLL_EINVAL;
  if (p.stop_rule == OLL_STOP_TOLERANCE && p.max_cycles <= 0 && p.patience <= 0)
    return err = "OLL_STOP_TOLERANCE needs max_cycles or patience", OLL_EINVAL;
  if (p.precision != OLL_PREC_FP64 && p.precision != OLL_PREC_MIXED)
    return err = "precision must be OLL_PREC_FP64 or OLL_PREC_MIXED", OLL_EINVAL;
  if (p.method != OLL_METHOD_VCYCLE && p.method != OLL_METHOD_APP)
    return err = "unknown method", OLL_EINVAL;
  if (p.method == OLL_METHOD_APP && p.eq != OLL_EQ_PUNCTURES)
    return err = "App puncture only works with puncture equation", OLL_EINVAL;
  if (p.initial_guess != OLL_GUESS_ZERO && p.initial_guess != OLL_GUESS_APPROX)
    return err = "unknown initial guess", OLL_EINVAL;
  if (p.initial_guess == OLL_GUESS_APPROX && p.eq != OLL_EQ_PUNCTURES && p.eq != OLL_EQ_PUNCTURES_SCALAR)
    return err = "the approximate initial guess is for the puncture equations", OLL_EINVAL;
  return OLL_OK;
}

int geometry(const oll_problem &p, Hierarchy &g, std::string &err) {
  PunctureSet ps;
  for (int i = 0; i < p.num_bh; ++i) {
    ps.x.push_back(p.bh[i].x);
    ps.y.push_back(p.bh[i].y);
    ps.z.push_back(p.bh[i].z);
  }
  err = build_hierarchy(size_t(p.N), p.L, p.levels, p.order, ps, g, p.method != OLL_METHOD_APP);
  if (!err.empty()) return err.find("enough") != std::string::npos ? OLL_EGRID : OLL_EINVAL;
  return OLL_OK;
}

// transfer tables of one hierarchy, built on the host
struct Pools {
  std::vector<int> ip;
  std::vector<unsigned char> bp;
  std::vector<double> dp;
  struct X {
    int src;
    int start[3], coinc[3], flags[3], w[3], k[3], n[3];
    int aff[3], affine;
    int filt;  // a Lagrange defect restriction reaching non-coincident nodes
  };

  std::vector<X> xs;
  // per source box: bounding box [lo, hi) of the nodes linear transfers
  // (prolongations) read from it, with one node of margin
  std::map<int, std::array<int, 6>> src_nodes;

  Hier::XRef add(const Hierarchy &g, int tbox, int slevel, int method, int ord) {
    Hier::XRef r;
    r.off = int(xs.size());
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = 1 << 30;
      r.hi[a] = 0;
    }
    const int fs = g.level_first[size_t(slevel)], ns = g.level_count[size_t(slevel)];
    for (int s = 0; s < ns; ++s) {
      X x{};
      x.src = fs + s;
      int slo[3], shi[3];
      for (int a = 0; a < 3; ++a) {
        const AxisTable t = build_axis_table(g.boxes[size_t(tbox)], g.boxes[size_t(fs + s)], a, method, ord);
        slo[a] = int(t.flags.size());
        shi[a] = 0;
        for (int q = 0; q < int(t.flags.size()); ++q)
          if (t.flags[size_t(q)] & AX_NB) {
            slo[a] = std::min(slo[a], q);
            shi[a] = q + 1;
          }
        if (method == 0) {
          const Box &S = g.boxes[size_t(fs + s)];
          int lo = 1 << 30, hi = -1;
          for (int q = 0; q < int(t.flags.size()); ++q) {
            const uint8_t fq = t.flags[size_t(q)];
            if (!(fq & (AX_NB | AX_IN))) continue;
            const int n0 = (fq & AX_COINC) ? t.coinc[size_t(q)] : t.start[size_t(q)];
            lo = std::min(lo, n0 - 1);
            hi = std::max(hi, n0 + 3);
          }
          if (hi > lo) {
            auto it = src_nodes.find(fs + s);
            if (it == src_nodes.end())
              it = src_nodes.emplace(fs + s, std::array<int, 6>{1 << 30, 1 << 30, 1 << 30, 0, 0, 0}).first;
            it->second[size_t(a)] = std::max(0, std::min(it->second[size_t(a)], lo));
            it->second[size_t(3 + a)] = std::min(int(S.n[a]), std::max(it->second[size_t(3 + a)], hi));
          }
        }
        x.start[a] = int(ip.size());
        ip.insert(ip.end(), t.start.begin(), t.start.end());
        x.coinc[a] = int(ip.size());
        ip.insert(ip.end(), t.coinc.begin(), t.coinc.end());
        x.flags[a] = int(bp.size());
        bp.insert(bp.end(), t.flags.begin(), t.flags.end());
        x.w[a] = int(dp.size());
        dp.insert(dp.end(), t.w.begin(), t.w.end());
        x.k[a] = t.k;
        x.n[a] = int(t.start.size());
        if (method == 0 && !nested_axis(t, x.aff[a])) x.affine = -1;
        // a target coordinate inside the source that is no source node:
        // the defect restriction interpolates there (k_filter_defect)
        if (method == 1 && ord == 2)
          for (uint8_t f : t.flags)
            if ((f & (AX_NB | AX_IN)) && !(f & AX_COINC)) x.filt = 1;
      }
      x.affine = (method == 0 && x.affine == 0) ? 1 : 0;
      xs.push_back(x);
      if (slo[0] < shi[0] && slo[1] < shi[1] && slo[2] < shi[2])
        for (int a = 0; a < 3; ++a) {
          r.lo[a] = std::min(r.lo[a], slo[a]);
          r.hi[a] = std::max(r.hi[a], shi[a]);
        }
    }
    r.cnt = int(xs.size()) - r.off;
    return r;
  }
};

template <typename T>
int upload(T **dst, const std::vector<T> &v, size_t &bytes, std::string &err) {
  const size_t n = std::max<size_t>(1, v.size()) * sizeof(T);
  ENG_CHECK(hipMalloc(reinterpret_cast<void **>(dst), n));
  if (!v.empty()) ENG_CHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  bytes += n;
  return OLL_OK;
}

// adds to cnt[q] the interior points of X inside [lo, hi) that colour
// pass q updates (flipped boxes update colour nc-1-q): indices counted by
// residue mod nc per axis, the colour (i + j + k) mod nc from the three
// residues
void colour_counts(const DevBox &X, const int lo[3], const int hi[3], std::vector<double> &cnt) {
  const int nc = X.nc;
  std::vector<long long> hist[3];
  for (int a = 0; a < 3; ++a) {
    hist[a].assign(size_t(nc), 0);
    for (int x = std::max(X.R, lo[a]); x < std::min(X.n[a] - X.R, hi[a]); ++x)
      ++hist[a][size_t(x % nc)];
  }
  for (int ri = 0; ri < nc; ++ri)
    for (int rj = 0; rj < nc; ++rj)
      for (int rk = 0; rk < nc; ++rk) {
        const int c = (ri + rj + rk) % nc;
        cnt[size_t(X.flip ? nc - 1 - c : c)] +=
            double(hist[0][size_t(ri)] * hist[1][size_t(rj)] * hist[2][size_t(rk)]);
      }
}

}  // namespace

// per colour pass, the points of level l's boxes inside their rho boxes
std::vector<double> level_rho_points(const Hier &H, int l) {
  const int f = H.geo.level_first[size_t(l)];
  std::vector<double> cnt(size_t(H.hbox[size_t(f)].nc), 0.0);
  for (int b = 0; b < H.geo.level_count[size_t(l)]; ++b) {
    const DevBox &X = H.hbox[size_t(f + b)];
    colour_counts(X, X.rlo, X.rhi, cnt);
  }
  return cnt;
}

int hier_rho_whole(Hier &H, int first, int count, hipStream_t s, std::string &err) {
  bool changed = false;
  for (int b = first; b < first + count; ++b) {
    DevBox &d = H.hbox[size_t(b)];
    for (int a = 0; a < 3; ++a) {
      changed |= d.rlo[a] != 0 || d.rhi[a] != d.n[a];
      d.rlo[a] = 0;
      d.rhi[a] = d.n[a];
    }
  }
  if (!changed) return OLL_OK;
  ENG_CHECK(hipMemcpyAsync(H.dbox + first, H.hbox.data() + first, size_t(count) * sizeof(DevBox),
                           hipMemcpyHostToDevice, s));
  ENG_CHECK(hipStreamSynchronize(s));
  for (int l = 0; l < H.geo.levels; ++l) H.colour_rho[size_t(l)] = level_rho_points(H, l);
  return OLL_OK;
}

int problem_footprint(const oll_problem &p, size_t &bytes, std::string &err) {
  int rc = validate(p, err);
  if (rc) return rc;
  Hierarchy g;
  if ((rc = geometry(p, g, err))) return rc;
  size_t pts = 0, tables = 0;
  for (const Box &b : g.boxes) {
    const int n[3] = {int(b.n[0]), int(b.n[1]), int(b.n[2])};
    pts += size_t((layout_size(n, colours_of(int(b.order), p.eq)) + 31) & ~31LL);
  }
  for (int l = 1; l < g.levels; ++l)
    for (int b = 0; b < g.level_count[size_t(l)]; ++b) {
      const Box &B = g.boxes[size_t(g.level_first[size_t(l)] + b)];
      const size_t nsrc = size_t(g.level_count[size_t(l - 1)]);
      tables += nsrc * (B.n[0] + B.n[1] + B.n[2]) * (9 + 8 * size_t(p.order + 2)) * 2;
    }
  bytes = pts * (sizeof(double) * size_t(fields_of(p.eq)) + (p.precision == OLL_PREC_MIXED ? sizeof(float) * NFIELDS32 : 0)) +
          tables + (64 << 10);
  return OLL_OK;
}

int hier_build(Hier &H, const oll_problem &p, hipStream_t s, std::string &err) {
  int rc = validate(p, err);
  if (rc) return rc;
  H.prob = p;
  if ((rc = geometry(p, H.geo, err))) return rc;
  const Hierarchy &g = H.geo;
  // slab: field-major, boxes back to back (colour blocks, kernels.hpp)
  long long total = 0;
  std::vector<long long> off;
  for (const Box &b : g.boxes) {
    if (b.n[0] >= size_t(kMaxAxisPoints) || b.n[1] >= size_t(kMaxAxisPoints) ||
        b.n[2] >= size_t(kMaxAxisPoints))
      return err = "a box exceeds 4095 points per axis", OLL_EINVAL;
    const int n[3] = {int(b.n[0]), int(b.n[1]), int(b.n[2])};
    const int nc = colours_of(int(b.order), p.eq);
    if (layout_size(n, nc) * (long long)sizeof(double) >= kMaxFieldBytes)
      return err = "a box field exceeds 4 GiB", OLL_EINVAL;
    // the ghost fill splits shell point indices with 24-bit float division
    const long long inner = (long long)(n[0] - 2 * b.R) * (n[1] - 2 * b.R) * (n[2] - 2 * b.R);
    if ((long long)n[0] * n[1] * n[2] - inner >= (1LL << 24))
      return err = "a box exceeds 2^24 ghost-shell points", OLL_EINVAL;
    off.push_back(total);
    total += (layout_size(n, nc) + 31) & ~31LL;  // 256-byte aligned boxes
  }
  const int nf = fields_of(p.eq);
  const size_t slab_bytes = size_t(total) * sizeof(double) * size_t(nf);
  hipError_t e = hipMalloc(&H.slab, slab_bytes);
  if (e != hipSuccess) {
    err = std::string("hipMalloc slab: ") + hipGetErrorString(e);
    return OLL_ENOMEM;
  }
  H.bytes += slab_bytes;
  ENG_CHECK(hipMemsetAsync(H.slab, 0, slab_bytes, s));
  if (p.precision == OLL_PREC_MIXED) {
    // e must read zero on the ghost shell: the whole slab starts zeroed and
    // the kernels write interior points only
    const size_t b32 = size_t(total) * sizeof(float) * NFIELDS32;
    e = hipMalloc(&H.slab32, b32);
    if (e != hipSuccess) {
      err = std::string("hipMalloc slab32: ") + hipGetErrorString(e);
      return OLL_ENOMEM;
    }
    H.bytes += b32;
    ENG_CHECK(hipMemsetAsync(H.slab32, 0, b32, s));
  }
  const double ipn = 1.0 / double(p.robin_n);
  for (size_t ib = 0; ib < g.boxes.size(); ++ib) {
    const Box &b = g.boxes[ib];
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
    d.dulap = dulap_of(b.order, d.idx2);
    d.ipn = ipn;
    set_layout(d, colours_of(b.order, p.eq));
    d.flip = colour_flip(g, b, d.nc);
    for (int a = 0; a < 3; ++a) {  // the whole box until the transfer tables narrow it
      d.rlo[a] = 0;
      d.rhi[a] = d.n[a];
    }
    for (int f = 0; f < NFIELDS; ++f)
      d.f[f] = f < nf ? H.slab + size_t(f) * size_t(total) + size_t(off[ib]) : nullptr;
    for (int f = 0; f < NFIELDS32; ++f)
      d.g[f] = H.slab32 ? H.slab32 + size_t(f) * size_t(total) + size_t(off[ib]) : nullptr;
    H.hbox.push_back(d);
  }
  for (int l = 0; l < g.levels; ++l) {
    const DevBox &B = H.hbox[size_t(g.level_first[size_t(l)])];
    const int nc = B.nc;
    std::vector<double> cnt(size_t(nc), 0.0);
    int rows = 0;
    for (int b = 0; b < g.level_count[size_t(l)]; ++b) {
      const DevBox &X = H.hbox[size_t(g.level_first[size_t(l)] + b)];
      // rows longer than a launch's threads are split into segments (64
      // slots at the smallest launch shape)
      rows += interior_rows(X) * ((X.rl + 63) / 64);
      const int lo[3] = {0, 0, 0};
      colour_counts(X, lo, X.n, cnt);
    }
    // row kernels divide row indices with 24-bit exact float division
    if ((long long)rows * nc >= (1LL << 24))
      return err = "a level exceeds 2^24 interior rows", OLL_EINVAL;
    H.colour_pts.push_back(cnt);
  }
  if ((rc = upload(&H.dbox, H.hbox, H.bytes, err))) return rc;

  // transfer tables
  Pools P;
  const int L = g.levels;
  H.ru.assign(size_t(L), {});
  H.rw.assign(size_t(L), {});
  H.pv.assign(size_t(L), {});
  H.gf.assign(size_t(L), {});
  for (int k = 1; k < L; ++k) {
    for (int t = 0; t < g.level_count[size_t(k - 1)]; ++t) {
      const int tb = g.level_first[size_t(k - 1)] + t;
      H.ru[size_t(k)].push_back(P.add(g, tb, k, 1, 3));
      H.rw[size_t(k)].push_back(P.add(g, tb, k, 1, 2));
    }
    for (int t = 0; t < g.level_count[size_t(k)]; ++t) {
      const int tb = g.level_first[size_t(k)] + t;
      H.pv[size_t(k)].push_back(P.add(g, tb, k - 1, 0, 1));
      H.gf[size_t(k)].push_back(P.add(g, tb, k - 1, 1, p.order + 1));
    }
  }
  // rho boxes: rho stays +0 at the interior points no restriction writes.
  // A box of the top level is no restriction's target; a multigrid level's
  // box is covered whole by the finer one; otherwise the target points
  // inside some finer box's non-ghost part (XRef lo, hi) bound the FAS
  // source.  Poisson's rho is its source term, nonzero everywhere
  if (p.eq != 2)
    for (int l = 0; l < L; ++l)
      for (int t = 0; t < g.level_count[size_t(l)]; ++t) {
        DevBox &d = H.hbox[size_t(g.level_first[size_t(l)] + t)];
        if (l + 1 <= g.mgrid) continue;
        for (int a = 0; a < 3; ++a) {
          d.rlo[a] = 0;
          d.rhi[a] = 0;
        }
        if (l + 1 >= L) continue;
        const Hier::XRef &r = H.rw[size_t(l + 1)][size_t(t)];
        if (r.lo[0] < r.hi[0] && r.lo[1] < r.hi[1] && r.lo[2] < r.hi[2])
          for (int a = 0; a < 3; ++a) {
            d.rlo[a] = r.lo[a];
            d.rhi[a] = r.hi[a];
          }
      }
  ENG_CHECK(hipMemcpy(H.dbox, H.hbox.data(), H.hbox.size() * sizeof(DevBox), hipMemcpyHostToDevice));
  H.colour_rho.clear();
  for (int l = 0; l < L; ++l) H.colour_rho.push_back(level_rho_points(H, l));
  H.prolong_nodes.assign(g.boxes.size(), std::array<int, 6>{1 << 30, 1 << 30, 1 << 30, 0, 0, 0});
  for (const auto &kv : P.src_nodes) H.prolong_nodes[size_t(kv.first)] = kv.second;
  if ((rc = upload(&H.ipool, P.ip, H.bytes, err))) return rc;
  if ((rc = upload(&H.bpool, P.bp, H.bytes, err))) return rc;
  if ((rc = upload(&H.dpool, P.dp, H.bytes, err))) return rc;
  std::vector<DevXfer> xf;
  H.filt_src.assign(g.boxes.size(), 0);
  for (const Pools::X &x : P.xs) {
    DevXfer d{};
    d.src = H.dbox + x.src;
    for (int a = 0; a < 3; ++a) {
      d.ax[a].start = H.ipool + x.start[a];
      d.ax[a].coinc = H.ipool + x.coinc[a];
      d.ax[a].flags = H.bpool + x.flags[a];
      d.ax[a].w = H.dpool + x.w[a];
      d.ax[a].k = x.k[a];
      d.ax[a].n = x.n[a];
      d.aff[a] = x.aff[a];
    }
    d.affine = x.affine;
    d.filt = x.filt;
    if (x.filt) H.filt_src[size_t(x.src)] = 1;
    xf.push_back(d);
  }
  if ((rc = upload(&H.dxfer, xf, H.bytes, err))) return rc;

  H.nslots = 2 * L + 2;
  ENG_CHECK(hipMalloc(&H.norms, size_t(H.nslots) * kNormWords * sizeof(unsigned long long)));
  ENG_CHECK(hipMalloc(&H.iters, sizeof(int)));
  ENG_CHECK(hipMemsetAsync(H.iters, 0, sizeof(int), s));
  H.bytes += size_t(H.nslots) * kNormWords * sizeof(unsigned long long) + sizeof(int);
  launch_zero_u64(s, H.norms, (long long)H.nslots * kNormWords);
  H.eps_coarse = 2 * DBL_EPSILON;
  H.eps_solv = p.tol < H.eps_coarse ? H.eps_coarse : p.tol;
  // over-relaxation of the coarsest-level loop, tuned on the 13^3-21^3
  // coarsest boxes the hierarchy produces (17^3: 1.9)
  H.coarse_omega = 2.0 - 1.6 / double(H.hbox[size_t(g.level_first[0])].n[0] - 1);

  PunctureArgs &pa = H.pa;
  pa.n = p.num_bh;
  pa.eq = p.eq;
  for (int i = 0; i < p.num_bh; ++i) {
    const oll_puncture &q = p.bh[i];
    pa.p[i] = Puncture{q.m, q.x, q.y, q.z, q.px, q.py, q.pz, q.sx, q.sy, q.sz, q.q};
  }
  pa.ps_case = p.ps_case;
  pa.ps_r0 = p.ps_r0;
  pa.ps_isigma = 1 / p.ps_sigma;
  pa.ps_phi0 = (p.ps_a2 * p.ps_a2 / (1 + p.ps_a2 * p.ps_a2)) * p.ps_phi0;
  pa.ps_d1 = M_PI * pa.ps_phi0 * pa.ps_phi0 * pa.ps_isigma * pa.ps_isigma;
  pa.ps_e1 = 1 / (16 * p.ps_a2);
  pa.ps_e2 = 4 * std::sqrt(M_PI / 3);
  pa.ps_de1 = 4 * M_PI * pa.ps_phi0 * pa.ps_phi0 * pa.ps_isigma * pa.ps_isigma;
  return hier_reset(H, s, err);
}

int hier_reset(Hier &H, hipStream_t s, std::string &err) {
  for (const DevBox &b : H.hbox) launch_coefficients_box(s, b, H.pa, H.prob.robin_a);
  if (H.prob.method == OLL_METHOD_APP || H.prob.initial_guess == OLL_GUESS_APPROX)
    for (const DevBox &b : H.hbox) launch_approx_box(s, b, H.pa);
  ENG_CHECK(hipGetLastError());
  ENG_CHECK(hipStreamSynchronize(s));
  H.cycle = 0;
  H.res = H.res_prev = 1e12;
  H.best = 1e300;
  H.best_cycle = 0;
  H.done = false;
  H.status = 0;
  H.hist.clear();
  H.last_slots.clear();
  H.coarse_sweeps = 0;
  return OLL_OK;
}

void hier_release(Hier &H) {
  (void)hipFree(H.slab);
  (void)hipFree(H.slab32);
  (void)hipFree(H.dbox);
  (void)hipFree(H.ipool);
  (void)hipFree(H.bpool);
  (void)hipFree(H.dpool);
  (void)hipFree(H.dxfer);
  (void)hipFree(H.norms);
  (void)hipFree(H.iters);
  H.slab = nullptr;
  H.slab32 = nullptr;
  H.dbox = nullptr;
  H.ipool = nullptr;
  H.bpool = nullptr;
  H.dpool = nullptr;
  H.dxfer = nullptr;
  H.norms = nullptr;
  H.iters = nullptr;
}

// ---------------------------------------------------------------------
// schedules
// ---------------------------------------------------------------------

int transfer_fields(Hier &H, hipStream_t s, int from, int ffield, int to, int tfield, int region,
                    int method, int ord, int all, int to_zero, std::string &err) {
  const Hierarchy &g = H.geo;
  const int L = g.levels;
  from = std::max(0, std::min(from, L - 1));
  to = std::max(0, std::min(to, L - 1));
  // the kernel reads the source while it writes the target: a field onto
  // itself on one level would depend on the order of the points
  if (from == to && ffield == tfield)
    return err = "a field cannot be transferred onto itself on one level", OLL_EINVAL;
  if (method == 1 && (ord < 2 || ord > 9))
    return err = "Lagrange transfers take order 2..9", OLL_EINVAL;
  const int nodes = method == 0 ? 2 : ord + 1;
  Pools P;
  std::vector<Hier::XRef> refs;
  const int ft = g.level_first[size_t(to)], nt = g.level_count[size_t(to)];
  if (tfield == F_RHO) {  // rho may now be nonzero anywhere on the target level
    const int rc = hier_rho_whole(H, ft, nt, s, err);
    if (rc) return rc;
  }
  for (int t = 0; t < nt; ++t) refs.push_back(P.add(g, ft + t, from, method, method == 0 ? 1 : ord));
  int *ip = nullptr;
  unsigned char *bp = nullptr;
  double *dp = nullptr;
  DevXfer *dx = nullptr;
  PtBlock *db = nullptr;
  size_t bytes = 0;
  int rc = upload(&ip, P.ip, bytes, err);
  if (!rc) rc = upload(&bp, P.bp, bytes, err);
  if (!rc) rc = upload(&dp, P.dp, bytes, err);
  std::vector<DevXfer> xf;
  for (const Pools::X &x : P.xs) {
    DevXfer d{};
    d.src = H.dbox + x.src;
    for (int a = 0; a < 3; ++a) {
      d.ax[a].start = ip + x.start[a];
      d.ax[a].coinc = ip + x.coinc[a];
      d.ax[a].flags = bp + x.flags[a];
      d.ax[a].w = dp + x.w[a];
      d.ax[a].k = x.k[a];
      d.ax[a].n = x.n[a];
      d.aff[a] = x.aff[a];
    }
    d.affine = x.affine;
    xf.push_back(d);
  }
  if (!rc) rc = upload(&dx, xf, bytes, err);
  std::vector<PtBlock> blocks;
  const int mode = (region & 3) | (all ? 4 : 0) | (to_zero ? 8 : 0);
  for (int t = 0; t < nt; ++t) {
    const DevBox &B = H.hbox[size_t(ft + t)];
    const long long np = npts(B), plane = (long long)B.n[0] * B.n[1];
    for (long long q0 = 0; q0 < np; q0 += kPtThreads)
      blocks.push_back(PtBlock{H.dbox + ft + t, dx + refs[size_t(t)].off, refs[size_t(t)].cnt, mode,
                               unsigned(q0), unsigned(std::min<long long>(kPtThreads, np - q0)),
                               int(q0 % B.n[0]), int((q0 % plane) / B.n[0]), int(q0 / plane)});
  }
  if (!rc) rc = upload(&db, blocks, bytes, err);
  if (!rc && !launch_transfer_fields(s, db, int(blocks.size()), ffield, tfield, nodes))
    rc = (err = "unsupported interpolation", OLL_EINVAL);
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = (err = "transfer failed", OLL_EDEVICE);
  (void)hipFree(ip);
  (void)hipFree(bp);
  (void)hipFree(dp);
  (void)hipFree(dx);
  (void)hipFree(db);
  return rc;
}

// fused double-buffered sweeps (k_smooth_fused): off by default.  On
// MI355X they run slower than the colour-by-colour passes at every stencil
// order and tile shape measured (4th order, N=65: 1.1-3x; 6th order,
// N=161: 1.5x): the first pass's recomputed halo rows (1 + R per tile side)
// add a third to a half of the work, and the tiles' working sets exceed
// the L2, so the second pass gains no reuse.
static bool g_fused_sweeps = false;

// Two-colour boxes sweep first the colour of their points that coincide
// with nodes of the parent box (the box of the coarser level that holds
// the box's centre): flip = 1 when those points have odd i + j + k.  A box
// whose origin lies off the parent's lattice (an offset that is not a
// whole number of fine spacings along some axis) has no coincident points
// and sweeps in ascending order.
int colour_flip(const Hierarchy &g, const Box &b, int nc) {
  if (nc != 2 || b.level == 0) return 0;
  const int f = g.level_first[size_t(b.level - 1)], n = g.level_count[size_t(b.level - 1)];
  const Box *P = &g.boxes[size_t(f)];
  for (int q = f; q < f + n; ++q) {
    const Box &C = g.boxes[size_t(q)];
    bool in = true;
    for (int a = 0; a < 3; ++a) in = in && b.c[a] >= C.lo[a] && b.c[a] <= C.hi[a];
    if (in) {
      P = &C;
      break;
    }
  }
  int sum = 0;
  for (int a = 0; a < 3; ++a) {
    const double t = (b.lo[a] - P->lo[a]) / b.d[a];
    const double m = std::nearbyint(t);
    if (std::fabs(t - m) > 1e-6) return 0;
    sum += int(std::fmod(std::fabs(m), 2.0));
  }
  return sum & 1;
}
void set_fused_sweeps(bool on) { g_fused_sweeps = on; }
static bool g_shared_loads = true;
void set_shared_loads(bool on) { g_shared_loads = on; }
bool shared_loads() { return g_shared_loads; }
// columns pay at orders 6 and 8 (3 and 4 points per column share 8^3 and
// 10^3 node stencils: ghost fill of the N=161 6th-order workload 766 ->
// 502 us per launch); at orders 2 and 4 the halved thread count and lower
// occupancy cost more than the saved loads (4th-order binary: 10% slower
// alone), so those fill point by point
bool ghost_columns_for(int order) { return g_shared_loads && order >= 6; }
bool fused_sweeps() { return g_fused_sweeps; }

void Schedule::clear() {
  launches.clear();
  blob.clear();
  uploaded = false;
  runs = 0;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

void release_schedule(Schedule &S) {
  S.clear();
  (void)hipFree(S.dblob);
  S.dblob = nullptr;
  S.dcap = 0;
}

namespace {

struct Builder {
  Schedule &S;

  // threads per row of a level: colour points per interior row (colour
  // passes) or slots per colour row (whole-row kernels)
  static int per_row(const Hier &H, int level, bool subrows) {
    const int f = H.geo.level_first[size_t(level)], n = H.geo.level_count[size_t(level)];
    int m = 0;
    for (int b = f; b < f + n; ++b) {
      const DevBox &B = H.hbox[size_t(b)];
      m = std::max(m, subrows ? B.rl : (B.n[0] - 2 * B.R + B.nc - 1) / B.nc);
    }
    return m;
  }

  // row blocks of one level of one member appended to `out`: the interior
  // rows (or colour rows) of each box in runs of kRowThreads / pr
  // (pr: the level's longest row in threads; a launch runs min(pr,
  // threads) threads per row, splitting longer rows into segments)
  static void row_blocks(const Hier &H, int level, unsigned long long *norm, bool subrows, int pr,
                         std::vector<RowBlock> &out, int threads = kRowThreads, int reps = 1) {
    const int seg = std::min(pr, threads);
    const int rpb = std::max(1, threads / seg) * reps;
    const int f = H.geo.level_first[size_t(level)], n = H.geo.level_count[size_t(level)];
    for (int b = f; b < f + n; ++b) {
      const DevBox &B = H.hbox[size_t(b)];
      const int len = subrows ? B.rl : (B.n[0] - 2 * B.R + B.nc - 1) / B.nc;
      const int nseg = (len + seg - 1) / seg;
      const int rows = interior_rows(B) * (subrows ? B.nc : 1) * nseg;
      for (int r0 = 0; r0 < rows; r0 += rpb)
        out.push_back(RowBlock{H.dbox + b, norm, r0, std::min(rpb, rows - r0), nseg});
    }
  }

  // blocks row_blocks() gives at reps = 1
  static long long row_block_count(const Hier &H, int level, bool subrows, int pr,
                                   int threads = kRowThreads) {
    const int seg = std::min(pr, threads);
    const int rpb = std::max(1, threads / seg);
    const int f = H.geo.level_first[size_t(level)], n = H.geo.level_count[size_t(level)];
    long long nb = 0;
    for (int b = f; b < f + n; ++b) {
      const DevBox &B = H.hbox[size_t(b)];
      const int len = subrows ? B.rl : (B.n[0] - 2 * B.R + B.nc - 1) / B.nc;
      const long long rows = (long long)interior_rows(B) * (subrows ? B.nc : 1) * ((len + seg - 1) / seg);
      nb += (rows + rpb - 1) / rpb;
    }
    return nb;
  }

  // tiles of the fused passes over the interior rows of every box of a
  // level of one member
  static void tile_blocks(const Hier &H, int level, unsigned long long *norm, int ty, int tz,
                          std::vector<TileBlock> &out) {
    const int f = H.geo.level_first[size_t(level)], n = H.geo.level_count[size_t(level)];
    for (int b = f; b < f + n; ++b) {
      const DevBox &B = H.hbox[size_t(b)];
      for (int k0 = B.R; k0 < B.n[2] - B.R; k0 += tz)
        for (int j0 = B.R; j0 < B.n[1] - B.R; j0 += ty)
          out.push_back(TileBlock{H.dbox + b, norm, j0, k0, std::min(ty, B.n[1] - B.R - j0),
                                  std::min(tz, B.n[2] - B.R - k0)});
    }
  }

  // point blocks over the ghost shells of every box of a level of one member
  static void shell_blocks(const Hier &H, int level, std::vector<PtBlock> &out) {
    const int f = H.geo.level_first[size_t(level)], n = H.geo.level_count[size_t(level)];
    for (int b = f; b < f + n; ++b) {
      const long long np = shell_pts(H.hbox[size_t(b)]);
      for (long long q0 = 0; q0 < np; q0 += kPtThreads)
        out.push_back(PtBlock{H.dbox + b, nullptr, 0, 0, unsigned(q0),
                              unsigned(std::min<long long>(kPtThreads, np - q0)), 0, 0, 0});
    }
  }

  // clears every member's norm slots, each with its own slot count
  void zero(const std::vector<Hier *> &M) {
    std::vector<ZeroItem> items;
    for (Hier *H : M) items.push_back(ZeroItem{H->norms, (long long)H->nslots * kNormWords});
    Launch L{Launch::ZERO};
    L.off = S.add(items);
    L.n = int(items.size());
    S.launches.push_back(L);
  }

  // one smoothing phase (nu sweeps) of level lv[h] of each member; FP64
  // members get nu * (R+1) colour passes, mixed-precision members a defect
  // launch, nu * (R+1) FP32 colour passes and an update launch
  void smooth(const std::vector<Hier *> &M, const std::vector<int> &lv, int nu, bool post,
              int norm_slot = -1) {
    if (nu <= 0) return;
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;  // (order, eq, mixed) -> members
    for (size_t h = 0; h < M.size(); ++h)
      if (lv[h] >= 0)
        groups[{M[h]->level_order(lv[h]), M[h]->prob.eq, M[h]->mixed() ? 1 : 0}].push_back(int(h));
    for (auto &g : groups) {
      const int order = std::get<0>(g.first), eq = std::get<1>(g.first);
      const bool mixed = std::get<2>(g.first) != 0;
      const int nc = colours_of(order, eq);
      auto slot = [&](const Hier &H, int l) {
        return H.slot(norm_slot >= 0 ? norm_slot : post ? Hier::slot_post(l) : Hier::slot_pre(l));
      };
      int pc = 0, ps = 0;  // threads per row: colour passes, whole-row kernels
      for (int h : g.second) {
        pc = std::max(pc, per_row(*M[size_t(h)], lv[size_t(h)], false));
        ps = std::max(ps, per_row(*M[size_t(h)], lv[size_t(h)], true));
      }
      std::vector<RowBlock> cb, sb;
      double bytes[5] = {0, 0, 0, 0, 0};
      std::set<const DevBox *> boxes;
      // colour-pass workgroup size from the launch-shape table, at least
      // one interior row
      int threads = sweep_threads(mixed ? 1 : 0, order);
      while (threads < pc && threads < kMaxRowThreads) threads *= 2;
      int sreps = 1;
      if (mixed) {
        long long nsb = 0;
        for (int h : g.second) nsb += row_block_count(*M[size_t(h)], lv[size_t(h)], true, ps);
        sreps = block_reps(g.second.size(), nsb);
      }
      for (int h : g.second) {
        const Hier &H = *M[size_t(h)];
        const int l = lv[size_t(h)];
        row_blocks(H, l, mixed ? nullptr : slot(H, l), false, pc, cb, threads);
        if (mixed) row_blocks(H, l, slot(H, l), true, ps, sb, kRowThreads, sreps);
        for (int b = 0; b < H.geo.level_count[size_t(l)]; ++b)
          boxes.insert(H.dbox + H.geo.level_first[size_t(l)] + b);
        // algorithmic bytes per updated point: FP64 reads u, b, c, rho
        // (+ d, e) and writes u, rho only inside the box's rho box; FP32
        // reads e, q, r and writes e
        const double per = mixed ? 16.0
                           : (eq == OLL_EQ_PUNCTURES_SCALAR ? 56.0 : (eq == OLL_EQ_PUNCTURES_EM ? 88.0 : 40.0));
        for (int q = 0; q < nc; ++q) {
          const double pts = H.colour_pts[size_t(l)][size_t(q)];
          bytes[q] += per * pts - (mixed ? 0.0 : 8.0 * (pts - H.colour_rho[size_t(l)][size_t(q)]));
        }
      }
      // FP64 sweeps in pairs as fused double-buffered launches (an even
      // count, so u ends in its own field); the rest one colour at a time
      const bool fuse = fused_sweeps() && !mixed && eq != OLL_EQ_PUNCTURES_EM;
      const int nfused = fuse ? nu - nu % 2 : 0;
      if (nfused > 0) {
        std::vector<TileBlock> tb;
        std::vector<PtBlock> sh;
        const FusedShape fs = fused_shape(order);
        for (int h : g.second) {
          const Hier &H = *M[size_t(h)];
          const int l = lv[size_t(h)];
          tile_blocks(H, l, slot(H, l), fs.ty, fs.tz, tb);
          shell_blocks(H, l, sh);
        }
        Launch C{Launch::COPY_SHELL};
        C.off = S.add(sh);
        C.n = int(sh.size());
        S.launches.push_back(C);
        const size_t toff = S.add(tb);
        for (int sw = 0; sw < nfused; ++sw)
          for (int q = 0; q < nc; q += 2) {
            Launch L{Launch::SMOOTH_FUSED};
            L.order = order;
            L.eq = eq;
            L.colour = q;
            L.pair = q + 1 < nc ? 1 : 0;
            L.parity = sw & 1;
            L.off = toff;
            L.n = int(tb.size());
            L.threads = fs.threads;
            L.bytes = bytes[q] + (L.pair ? bytes[q + 1] : 0.0);
            S.launches.push_back(L);
          }
      }
      const int nu_single = nu - nfused;
      const size_t coff = S.add(cb);
      const size_t soff = mixed ? S.add(sb) : 0;
      auto whole = [&](Launch::Kind kind) {
        Launch L{kind};
        L.order = order;
        L.eq = eq;
        L.off = soff;
        L.n = int(sb.size());
        L.max_row = std::min(ps, kRowThreads);
        S.launches.push_back(L);
      };
      if (mixed) whole(Launch::DEFECT32);
      for (int s = 0; s < nu_single; ++s)
        for (int q = 0; q < nc; ++q) {
          Launch L{mixed ? Launch::SMOOTH32 : Launch::SMOOTH};
          L.order = order;
          L.eq = eq;
          L.colour = q;
          L.off = coff;
          L.n = int(cb.size());
          L.max_row = std::min(pc, threads);
          L.threads = threads;
          L.bytes = bytes[q];
          if (boxes.size() == 1) L.box1 = *boxes.begin();
          S.launches.push_back(L);
        }
      if (mixed) whole(Launch::UPDATE32);
    }
  }

  void residual(const std::vector<Hier *> &M, const std::vector<int> &lv, int write_defect,
                bool with_norm) {
    std::map<std::pair<int, int>, std::vector<int>> groups;
    for (size_t h = 0; h < M.size(); ++h)
      if (lv[h] >= 0) groups[{M[h]->level_order(lv[h]), M[h]->prob.eq}].push_back(int(h));
    for (auto &g : groups) {
      int ps = 0;
      for (int h : g.second) ps = std::max(ps, per_row(*M[size_t(h)], lv[size_t(h)], true));
      std::vector<RowBlock> blocks;
      long long nb = 0;
      for (int h : g.second) nb += row_block_count(*M[size_t(h)], lv[size_t(h)], true, ps);
      const int reps = block_reps(g.second.size(), nb);
      for (int h : g.second) {
        const Hier &H = *M[size_t(h)];
        row_blocks(H, lv[size_t(h)], with_norm ? H.slot(H.slot_scratch()) : nullptr, true, ps,
                   blocks, kRowThreads, reps);
      }
      Launch L{Launch::RESIDUAL};
      L.order = g.first.first;
      L.eq = g.first.second;
      L.mode = write_defect;
      L.off = S.add(blocks);
      L.n = int(blocks.size());
      L.max_row = std::min(ps, kRowThreads);
      S.launches.push_back(L);
    }
  }

  // only_box >= 0: that box of the level alone (one member)
  void robin(const std::vector<Hier *> &M, const std::vector<int> &lv, int mode, int only_box = -1) {
    std::vector<RobinItem> items;
    int total = 0, edges = 0;
    for (size_t h = 0; h < M.size(); ++h) {
      if (lv[h] < 0) continue;
      const Hier &H = *M[h];
      const int f = H.geo.level_first[size_t(lv[h])], n = H.geo.level_count[size_t(lv[h])];
      for (int b = f; b < f + n; ++b) {
        if (only_box >= 0 && b != only_box) continue;
        const int nc = face_columns(H.hbox[size_t(b)]);
        items.push_back(RobinItem{H.dbox + b, total, nc, edges});
        total += nc;
        edges += robin_edge_sections(H.hbox[size_t(b)]);
      }
    }
    if (items.empty()) return;
    Launch L{Launch::ROBIN};
    L.mode = mode;
    L.off = S.add(items);
    L.n = int(items.size());
    L.total = total;
    L.max_int = edges;
    S.launches.push_back(L);
  }

  // point blocks of the target boxes of level tl with transfer refs
  // X[k][t] (k = the member's fine level of the pair)
  // (shell: the ghost shell's points, or with columns its columns of R
  // points across the face slabs)
  static long long point_count(const DevBox &B, bool shell, bool columns) {
    // shell points by column slot (k_ghost_fill: R points a slot)
    return shell ? (columns ? ghost_column_slots(B) : ghost_column_slots(B) * B.R) : npts(B);
  }
  // blocks points() gives level tl of H at reps = 1
  static long long point_blocks(const Hier &H, int tl, bool shell, bool columns = false) {
    const int f = H.geo.level_first[size_t(tl)], n = H.geo.level_count[size_t(tl)];
    long long nb = 0;
    for (int t = 0; t < n; ++t)
      nb += (point_count(H.hbox[size_t(f + t)], shell, columns) + kPtThreads - 1) / kPtThreads;
    return nb;
  }
  static void points(const Hier &H, int tl, int k, const std::vector<std::vector<Hier::XRef>> &X,
                     bool shell, int mg, std::vector<PtBlock> &out, bool columns = false,
                     int reps = 1) {
    const int f = H.geo.level_first[size_t(tl)], n = H.geo.level_count[size_t(tl)];
    const long long per = (long long)kPtThreads * reps;
    for (int t = 0; t < n; ++t) {
      const DevBox &B = H.hbox[size_t(f + t)];
      const Hier::XRef r = X[size_t(k)][size_t(t)];
      const long long np = point_count(B, shell, columns);
      const long long plane = (long long)B.n[0] * B.n[1];
      for (long long q0 = 0; q0 < np; q0 += per)
        out.push_back(PtBlock{H.dbox + f + t, H.dxfer + r.off, r.cnt, mg, unsigned(q0),
                              unsigned(std::min<long long>(per, np - q0)),
                              int(q0 % B.n[0]), int((q0 % plane) / B.n[0]), int(q0 / plane)});
    }
  }

  // point blocks, patches of whole rows, over the region of a refinement
  // pair's coarse box where a restriction acts: the interior points inside
  // some fine box's non-ghost part (XRef lo, hi), and with `ww` also the
  // nodes the pair's prolongation reads back (u - ww: k_restrict_defect
  // saves ww there; elsewhere no kernel reads ww)
  static void segments(const Hier &H, int tl, int k, const std::vector<std::vector<Hier::XRef>> &X,
                       std::vector<PtBlock> &out, bool ww) {
    const int f = H.geo.level_first[size_t(tl)], n = H.geo.level_count[size_t(tl)];
    for (int t = 0; t < n; ++t) {
      const DevBox &B = H.hbox[size_t(f + t)];
      const Hier::XRef r = X[size_t(k)][size_t(t)];
      int lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::max(r.lo[a], B.R);
        hi[a] = std::min(r.hi[a], B.n[a] - B.R);
      }
      if (ww) {
        const std::array<int, 6> &pn = H.prolong_nodes[size_t(f + t)];
        if (pn[0] < pn[3] && pn[1] < pn[4] && pn[2] < pn[5])
          for (int a = 0; a < 3; ++a) {
            if (lo[a] >= hi[a]) {
              lo[a] = pn[size_t(a)];
              hi[a] = pn[size_t(3 + a)];
            } else {
              lo[a] = std::min(lo[a], pn[size_t(a)]);
              hi[a] = std::max(hi[a], pn[size_t(3 + a)]);
            }
          }
      }
      row_patches(H.dbox + f + t, H.dxfer + r.off, r.cnt, B, lo, hi, out);
    }
  }

  // blocks over the sub-box [lo, hi) of box B: whole rows, as many of one
  // plane as fill kPtThreads threads (one row a block when a row is longer)
  static void row_patches(const DevBox *dbox, const DevXfer *src, int nsrc, const DevBox &B,
                          const int lo[3], const int hi[3], std::vector<PtBlock> &out) {
    if (lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]) return;
    const int w = hi[0] - lo[0], m = std::max(1, kPtThreads / w);
    const long long plane = (long long)B.n[0] * B.n[1];
    for (int kk = lo[2]; kk < hi[2]; ++kk)
      for (int j = lo[1]; j < hi[1]; j += m) {
        const int rows = std::min(m, hi[1] - j);
        const long long q0 = kk * plane + (long long)j * B.n[0] + lo[0];
        out.push_back(PtBlock{dbox, src, nsrc, 0, unsigned(q0), unsigned(rows * w), lo[0], j, kk, w,
                              1.0f / float(w)});
      }
  }

  // Blocks per kPtThreads points (or per workgroup of rows) for a launch
  // over one request; a launch over several requests gives each block reps
  // times that many, which the kernels' threads step through, so that its
  // table (built and uploaded again whenever the set of requests changes)
  // is a fraction of the size while the launch keeps >= 4096 blocks.
  static int block_reps(size_t members, long long blocks) {
    if (members < 2) return 1;
    return int(std::max<long long>(1, std::min<long long>(16, blocks / 4096)));
  }

  // v = u - ww on the nodes of the coarse boxes of level lv[h] - 1 that the
  // prolongation into level lv[h] reads (Hier::prolong_nodes), in patches
  // of whole rows
  void corr_v(const std::vector<Hier *> &M, const std::vector<int> &lv) {
    std::vector<PtBlock> items;
    for (size_t h = 0; h < M.size(); ++h) {
      const int k = lv[h];
      if (k < 1) continue;
      const Hier &H = *M[h];
      const int f = H.geo.level_first[size_t(k - 1)], n = H.geo.level_count[size_t(k - 1)];
      for (int b = f; b < f + n; ++b) {
        const std::array<int, 6> &pn = H.prolong_nodes[size_t(b)];
        if (!(pn[0] < pn[3] && pn[1] < pn[4] && pn[2] < pn[5])) continue;
        const int lo[3] = {pn[0], pn[1], pn[2]}, hi[3] = {pn[3], pn[4], pn[5]};
        row_patches(H.dbox + b, nullptr, 0, H.hbox[size_t(b)], lo, hi, items);
      }
    }
    if (items.empty()) return;
    Launch L{Launch::CORR_V};
    L.off = S.add(items);
    L.n = int(items.size());
    S.launches.push_back(L);
  }

  void transfer(Launch::Kind kind, const std::vector<Hier *> &M, const std::vector<int> &lv) {
    // lv[h] = the member's fine level k of the pair (k-1, k); restrictions
    // are grouped by the coarse level's (order, eq), ghost fills by the
    // fine level's order (their Lagrange stencil)
    std::map<std::pair<int, int>, std::vector<PtBlock>> groups;
    // point blocks of this stage over all members at one block per
    // kPtThreads points (row-patch blocks are never joined)
    long long nb = 0;
    size_t members = 0;
    for (size_t h = 0; h < M.size(); ++h) {
      const int k = lv[h];
      if (k < 1) continue;
      const Hier &H = *M[h];
      const int mg = k <= H.geo.mgrid;
      ++members;
      if ((kind == Launch::RESTRICT_U || kind == Launch::RESTRICT_DEFECT) && mg)
        nb += point_blocks(H, k - 1, false);
      if (kind == Launch::PROLONG) nb += point_blocks(H, k, false);
      if (kind == Launch::GHOST && !mg)
        nb += point_blocks(H, k, true, ghost_columns_for(H.level_order(k)));
    }
    const int reps = block_reps(members, nb);
    for (size_t h = 0; h < M.size(); ++h) {
      const int k = lv[h];
      if (k < 1) continue;
      const Hier &H = *M[h];
      const int mg = k <= H.geo.mgrid;
      if (kind == Launch::RESTRICT_U) {
        if (mg)
          points(H, k - 1, k, H.ru, false, mg, groups[{0, 0}], false, reps);
        else
          segments(H, k - 1, k, H.ru, groups[{0, 0}], false);
      }
      if (kind == Launch::RESTRICT_DEFECT) {
        if (mg)
          points(H, k - 1, k, H.rw, false, mg, groups[{H.level_order(k - 1), H.prob.eq}], false, reps);
        else
          segments(H, k - 1, k, H.rw, groups[{H.level_order(k - 1), H.prob.eq}], true);
      }
      if (kind == Launch::PROLONG) points(H, k, k, H.pv, false, mg, groups[{0, 0}], false, reps);
      if (kind == Launch::GHOST && !mg)
        points(H, k, k, H.gf, true, 0, groups[{H.level_order(k), 0}],
               ghost_columns_for(H.level_order(k)), reps);
    }
    for (auto &g : groups) {
      if (g.second.empty()) continue;
      Launch L{kind};
      L.order = g.first.first;
      L.eq = g.first.second;
      L.mode = kind == Launch::GHOST && ghost_columns_for(L.order) ? 1 : 0;
      L.off = S.add(g.second);
      L.n = int(g.second.size());
      S.launches.push_back(L);
    }
  }

  // the defect filter of every box of level lv[h] that is the source of a
  // defect restriction off the coarser lattice
  void filter(const std::vector<Hier *> &M, const std::vector<int> &lv) {
    std::vector<PtBlock> items;
    for (size_t h = 0; h < M.size(); ++h) {
      const int k = lv[h];
      if (k < 1) continue;
      const Hier &H = *M[h];
      const int f = H.geo.level_first[size_t(k)], n = H.geo.level_count[size_t(k)];
      for (int b = f; b < f + n; ++b) {
        if (!H.filt_src[size_t(b)]) continue;
        const DevBox &B = H.hbox[size_t(b)];
        const int run = filter_run_length();
        const long long np = shared_loads() ? (long long)((B.n[0] + run - 1) / run) * B.n[1] * B.n[2]
                                            : npts(B);
        const long long plane = (long long)B.n[0] * B.n[1];
        for (long long q0 = 0; q0 < np; q0 += kPtThreads)
          items.push_back(PtBlock{H.dbox + b, nullptr, 0, 0, unsigned(q0),
                                  unsigned(std::min<long long>(kPtThreads, np - q0)),
                                  int(q0 % B.n[0]), int((q0 % plane) / B.n[0]), int(q0 / plane)});
      }
    }
    if (items.empty()) return;
    Launch L{Launch::FILTER};
    L.mode = shared_loads() ? 1 : 0;
    L.off = S.add(items);
    L.n = int(items.size());
    S.launches.push_back(L);
  }

  void coarse(const std::vector<Hier *> &M) {
    std::map<int, std::vector<CoarseItem>> groups;  // eq -> items
    std::map<int, std::pair<int, int>> sizes;
    for (Hier *H : M) {
      const int b0 = H->geo.level_first[0];
      const DevBox &B = H->hbox[size_t(b0)];
      groups[H->prob.eq].push_back(CoarseItem{H->dbox + b0, H->slot(Hier::slot_pre(0)), H->iters,
                                              H->coarse_omega});
      auto &sz = sizes[H->prob.eq];
      sz.first = std::max<int>(sz.first, int(B.size));
      sz.second = std::max<int>(sz.second, int(coarse_colour_slots(B.n)));
    }
    for (auto &g : groups) {
      Launch L{Launch::COARSE};
      L.eq = g.first;
      L.off = S.add(g.second);
      L.n = int(g.second.size());
      L.max_pts = sizes[g.first].first;
      L.max_int = sizes[g.first].second;
      S.launches.push_back(L);
    }
  }
};

}  // namespace

int build_vcycle_from(const std::vector<Hier *> &M, const std::vector<int> &start, Schedule &S,
                      std::string &err) {
  S.clear();
  if (M.empty()) return OLL_OK;
  for (Hier *H : M)
    if (H->prob.nu1 != M[0]->prob.nu1 || H->prob.nu2 != M[0]->prob.nu2)
      return err = "members of one schedule must share nu1 and nu2", OLL_EINVAL;
  for (Hier *H : M) {
    const DevBox &B = H->hbox[size_t(H->geo.level_first[0])];
    if (!coarse_box_fits(B.n, B.size))
      return err = "coarsest box exceeds the LDS of one workgroup", OLL_EINVAL;
    // k_coarse_solve's Robin shell and sweeps are those of a 2nd-order box
    if (B.nc != 2 || B.R != 1) return err = "coarsest box must be 2nd order", OLL_EINVAL;
  }
  Builder b{S};
  b.zero(M);
  int dmax = 0;
  for (size_t h = 0; h < M.size(); ++h) dmax = std::max(dmax, start[h]);
  const int nu1 = M[0]->prob.nu1, nu2 = M[0]->prob.nu2;
  std::vector<int> lv(M.size()), mglv(M.size());
  for (int d = 0; d < dmax; ++d) {
    // level of each member at this depth (-1: idle)
    for (size_t h = 0; h < M.size(); ++h) {
      lv[h] = (start[h] - d >= 1) ? start[h] - d : -1;
      mglv[h] = (lv[h] >= 1 && lv[h] <= M[h]->geo.mgrid) ? lv[h] : -1;
    }
    b.robin(M, mglv, ROBIN_UPDATE);
    b.smooth(M, lv, nu1, false);
    b.residual(M, lv, RES_DEFECT, false);
    b.robin(M, mglv, ROBIN_DEFECT_W);
    b.transfer(Launch::RESTRICT_U, M, lv);
    b.filter(M, lv);
    b.transfer(Launch::RESTRICT_DEFECT, M, lv);
  }
  b.coarse(M);
  for (int d = dmax - 1; d >= 0; --d) {
    for (size_t h = 0; h < M.size(); ++h) {
      lv[h] = (start[h] - d >= 1) ? start[h] - d : -1;
      mglv[h] = (lv[h] >= 1 && lv[h] <= M[h]->geo.mgrid) ? lv[h] : -1;
    }
    b.corr_v(M, lv);
    b.transfer(Launch::PROLONG, M, lv);
    b.transfer(Launch::GHOST, M, lv);
    b.robin(M, mglv, ROBIN_UPDATE);
    b.smooth(M, lv, nu2, true);
  }
  return OLL_OK;
}

int build_vcycle(const std::vector<Hier *> &M, Schedule &S, std::string &err) {
  std::vector<int> start;
  for (Hier *H : M) start.push_back(H->top());
  return build_vcycle_from(M, start, S, err);
}

int build_stage(Hier &H, int level, int stage, int nu, Schedule &S, std::string &err, int box) {
  (void)err;
  S.clear();
  std::vector<Hier *> M{&H};
  std::vector<int> lv{level};
  Builder b{S};
  if (stage == STAGE_SMOOTH || stage == STAGE_RESIDUAL || stage == STAGE_DERIVATIVE) b.zero(M);
  if (stage == STAGE_SMOOTH) {
    // the scratch slot collects the norm of these sweeps
    b.smooth(M, lv, nu, false, H.slot_scratch());
  } else if (stage == STAGE_RESIDUAL) {
    b.residual(M, lv, RES_OPERATOR, true);
  } else if (stage == STAGE_DERIVATIVE) {
    b.residual(M, lv, RES_DERIVATIVE, true);
  } else if (stage == STAGE_ROBIN) {
    b.robin(M, lv, ROBIN_UPDATE, box);
  } else {
    b.robin(M, lv, ROBIN_OPERATOR_R, box);
  }
  return OLL_OK;
}

int upload_schedule(Schedule &S, hipStream_t s, std::string &err) {
  if (S.blob.size() > S.dcap) {
    (void)hipFree(S.dblob);
    S.dblob = nullptr;
    const size_t cap = std::max<size_t>(S.blob.size() * 2, 4096);
    ENG_CHECK(hipMalloc(&S.dblob, cap));
    S.dcap = cap;
  }
  if (!S.blob.empty())
    ENG_CHECK(hipMemcpyAsync(S.dblob, S.blob.data(), S.blob.size(), hipMemcpyHostToDevice, s));
  S.uploaded = true;
  return OLL_OK;
}

namespace {

hipEvent_t timing_event(Timing &t) {
  if (t.used == t.ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    t.ev.push_back(e);
  }
  return t.ev[t.used++];
}

int issue(const Schedule &S, hipStream_t s, Timing *t, std::string &err) {
  for (const Launch &L : S.launches) {
    const void *items = S.dblob + L.off;
    // timed smoother launches carry their events into the dispatch: the
    // runtime stamps them at the kernel's start and end
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (t && t->on &&
        (L.kind == Launch::SMOOTH || L.kind == Launch::SMOOTH32 || L.kind == Launch::SMOOTH_FUSED)) {
      e0 = timing_event(*t);
      e1 = e0 ? timing_event(*t) : nullptr;
      if (e0 && !e1) --t->used;  // keep the events paired
      if (!e1) e0 = nullptr;
    }
    switch (L.kind) {
      case Launch::ZERO:
        launch_zero_lists(s, reinterpret_cast<const ZeroItem *>(items), L.n);
        break;
      case Launch::ROBIN:
        launch_robin(s, reinterpret_cast<const RobinItem *>(items), L.n, int(L.total), L.max_int,
                     L.mode);
        break;
      case Launch::SMOOTH:
        launch_smooth(s, reinterpret_cast<const RowBlock *>(items), L.n, L.max_row, L.order, L.eq,
                      L.colour, L.threads, e0, e1, L.box1);
        break;
      case Launch::DEFECT32:
        launch_defect32(s, reinterpret_cast<const RowBlock *>(items), L.n, L.max_row, L.order, L.eq);
        break;
      case Launch::SMOOTH32:
        launch_smooth32(s, reinterpret_cast<const RowBlock *>(items), L.n, L.max_row, L.order,
                        L.colour, L.threads, e0, e1);
        break;
      case Launch::UPDATE32:
        launch_update32(s, reinterpret_cast<const RowBlock *>(items), L.n, L.max_row, L.order);
        break;
      case Launch::RESIDUAL:
        launch_residual(s, reinterpret_cast<const RowBlock *>(items), L.n, L.max_row, L.order, L.eq,
                        L.mode);
        break;
      case Launch::RESTRICT_U:
        launch_restrict_u(s, reinterpret_cast<const PtBlock *>(items), L.n);
        break;
      case Launch::RESTRICT_DEFECT:
        launch_restrict_defect(s, reinterpret_cast<const PtBlock *>(items), L.n, L.order, L.eq);
        break;
      case Launch::COARSE:
        if (!launch_coarse_solve(s, reinterpret_cast<const CoarseItem *>(items), L.n, L.max_pts,
                                 L.max_int, L.eq, 2 * DBL_EPSILON, 100000))
          return err = "coarsest box exceeds the LDS of one workgroup", OLL_EINVAL;
        break;
      case Launch::CORR_V:
        launch_corr_v(s, reinterpret_cast<const PtBlock *>(items), L.n);
        break;
      case Launch::PROLONG:
        launch_prolong(s, reinterpret_cast<const PtBlock *>(items), L.n);
        break;
      case Launch::FILTER:
        launch_filter_defect(s, reinterpret_cast<const PtBlock *>(items), L.n, L.mode);
        break;
      case Launch::SMOOTH_FUSED:
        launch_smooth_fused(s, reinterpret_cast<const TileBlock *>(items), L.n, L.order, L.eq,
                            L.colour, L.pair, L.parity, L.threads, e0, e1);
        break;
      case Launch::COPY_SHELL:
        launch_copy_shell(s, reinterpret_cast<const PtBlock *>(items), L.n);
        break;
      case Launch::GHOST:
        launch_ghost_fill(s, reinterpret_cast<const PtBlock *>(items), L.n, L.order, L.mode);
        break;
    }
    if (e1) {
      t->bytes += L.bytes;
      t->launch_bytes.push_back(L.bytes);
      t->launch_kind.push_back(2 * L.order + (L.kind == Launch::SMOOTH32 ? 1 : 0));
    }
  }
  ENG_CHECK(hipGetLastError());
  return OLL_OK;
}

}  // namespace

int run_schedule(Schedule &S, hipStream_t s, Timing *t, bool graphs, std::string &err) {
  int rc;
  if (!S.uploaded && (rc = upload_schedule(S, s, err))) return rc;
  const bool timed = t && t->on;
  if (!graphs || timed || S.runs++ == 0) return issue(S, s, t, err);
  if (!S.exec) {
    ENG_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = issue(S, s, nullptr, err);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (e != hipSuccess) {
      err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
      return OLL_EDEVICE;
    }
    S.graph = g;
    ENG_CHECK(hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
  }
  ENG_CHECK(hipGraphLaunch(S.exec, s));
  return OLL_OK;
}

int read_norms(Hier &H, hipStream_t s, std::string &err) {
  H.raw.resize(size_t(H.nslots) * kNormWords);
  ENG_CHECK(hipMemcpyAsync(H.raw.data(), H.norms, H.raw.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
  ENG_CHECK(hipMemcpyAsync(&H.raw_iters, H.iters, sizeof(int), hipMemcpyDeviceToHost, s));
  return OLL_OK;
}

void decode_norms(Hier &H) {
  H.last_slots.assign(size_t(H.nslots), 0.0);
  for (int sl = 0; sl < H.nslots; ++sl) {
    unsigned long long m = 0;
    for (int q = 0; q < kNormSubSlots; ++q)
      m = std::max(m, H.raw[size_t(sl) * kNormWords + size_t(q) * kNormStride]);
    double d;
    std::memcpy(&d, &m, sizeof d);
    H.last_slots[size_t(sl)] = d;
  }
}

int finish_cycle(Hier &H, hipStream_t s, std::string &err) {
  int rc = read_norms(H, s, err);
  if (rc) return rc;
  ENG_CHECK(hipStreamSynchronize(s));
  return apply_cycle(H, err);
}

int apply_cycle(Hier &H, std::string &err) {
  decode_norms(H);
  H.coarse_sweeps += H.raw_iters;
  const int top = H.top();
  for (int k = top; k >= 0; --k) {
    H.hist.push_back(H.cycle);
    H.hist.push_back(H.last_slots[size_t(Hier::slot_pre(k))]);
  }
  H.res_prev = H.res;
  H.res = H.last_slots[size_t(Hier::slot_post(top))];
  ++H.cycle;
  for (int k = 0; k <= top; ++k)
    if (std::isinf(H.last_slots[size_t(Hier::slot_pre(k))]) ||
        std::isinf(H.last_slots[size_t(Hier::slot_post(k))])) {
      H.status = OLL_ENAN;
      H.done = true;
      err = "NAN check error in the smoother";
      return OLL_ENAN;
    }
  if (H.res < H.best) {
    H.best = H.res;
    H.best_cycle = H.cycle;
  }
  if (H.prob.stop_rule == OLL_STOP_TOLERANCE) {
    if (!(H.res > H.eps_solv)) H.done = true;
    if (H.prob.patience > 0 && H.cycle - H.best_cycle >= H.prob.patience) H.done = true;
  } else if (!(H.res > H.eps_solv && H.res < H.res_prev)) {
    H.done = true;
  }
  if (H.prob.max_cycles > 0 && H.cycle >= H.prob.max_cycles) H.done = true;
  return OLL_OK;
}

// device buffers of one host call, freed on every return path
struct TempBuffers {
  std::vector<void *> p;
  ~TempBuffers() {
    for (void *q : p) (void)hipFree(q);
  }
  template <typename T>
  hipError_t alloc(T **out, size_t count) {
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T *>(q);
    return e;
  }
};

int adm_mass(Hier &H, hipStream_t s, double *puncture, double *surface, std::string &err) {
  const int nb = int(H.hbox.size());
  const size_t nsph = surface ? size_t(nb) * kAdmNt * kAdmNt : 0;
  AdmBuffers d;
  TempBuffers tmp;
  ENG_CHECK(tmp.alloc(&d.pbox, size_t(H.prob.num_bh)));
  ENG_CHECK(tmp.alloc(&d.pxyz, size_t(3 * H.prob.num_bh)));
  ENG_CHECK(tmp.alloc(&d.up, size_t(H.prob.num_bh)));
  if (surface) {
    ENG_CHECK(tmp.alloc(&d.ctr, size_t(3 * nb)));
    ENG_CHECK(tmp.alloc(&d.sph, nsph));
  }
  return adm_mass_with(H, s, d, puncture, surface, err);
}

int adm_mass_with(Hier &H, hipStream_t s, const AdmBuffers &d, double *puncture, double *surface,
                  std::string &err) {
  const oll_problem &p = H.prob;
  const int top = H.top();
  const int nb = int(H.hbox.size());
  const int nt = kAdmNt;
  // u at each puncture from its finest box (Lagrange order 4);
  // multigrid.cc:97-116 falls back to the level's first box
  std::vector<int> pbox(size_t(p.num_bh));
  std::vector<double> pxyz(size_t(3 * p.num_bh));
  for (int i = 0; i < p.num_bh; ++i) {
    pbox[size_t(i)] = H.geo.level_first[size_t(top)] + (i < H.geo.level_count[size_t(top)] ? i : 0);
    pxyz[size_t(3 * i)] = p.bh[i].x;
    pxyz[size_t(3 * i + 1)] = p.bh[i].y;
    pxyz[size_t(3 * i + 2)] = p.bh[i].z;
  }
  std::vector<double> ctr(size_t(3 * nb));
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < 3; ++a) ctr[size_t(3 * b + a)] = H.geo.boxes[size_t(b)].c[a];
  const size_t nsph = surface ? size_t(nb) * nt * nt : 0;
  ENG_CHECK(hipMemcpyAsync(d.pbox, pbox.data(), pbox.size() * sizeof(int), hipMemcpyHostToDevice, s));
  ENG_CHECK(hipMemcpyAsync(d.pxyz, pxyz.data(), pxyz.size() * sizeof(double), hipMemcpyHostToDevice, s));
  launch_interp_points(s, H.dbox, F_U, d.pbox, d.pxyz, p.num_bh, 4, d.up);
  std::vector<double> up(pbox.size()), sph(nsph);
  if (surface) {
    ENG_CHECK(hipMemcpyAsync(d.ctr, ctr.data(), ctr.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_adm_sphere(s, H.dbox, nb, d.ctr, d.sph, nt);
    ENG_CHECK(hipMemcpyAsync(sph.data(), d.sph, nsph * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  ENG_CHECK(hipMemcpyAsync(up.data(), d.up, up.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  ENG_CHECK(hipStreamSynchronize(s));
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
      const double *Sd = &sph[size_t(b) * nt * nt];
      const Box &B = H.geo.boxes[size_t(b)];
      const double r0 = 0.375 * std::min(B.len[0], std::min(B.len[1], B.len[2]));
      double adm = 0;
      for (int i = 0; i < nt - 3; i += 4) {
        double I[5] = {0, 0, 0, 0, 0};
        for (int j = 0; j < nt - 3; j += 4)
          for (int q = 0; q < 5; ++q) {
            const double *row = Sd + size_t(i + q) * nt;
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

void puncture_approx(const Hier &H, double *u_at_p) {
  const oll_problem &p = H.prob;
  int n = 0;
  while (n < p.num_bh && p.bh[n].m != 0.0) ++n;
  for (int j = 0; j < p.num_bh; ++j) u_at_p[j] = 0.0;
  const Hierarchy &g = H.geo;
  const int top = g.levels - 1;
  for (int b = g.level_first[size_t(top)]; b < g.level_first[size_t(top)] + g.level_count[size_t(top)]; ++b) {
    const Box &B = g.boxes[size_t(b)];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double pj[3] = {p.bh[j].x, p.bh[j].y, p.bh[j].z};
        if (!(B.in_box(0, pj[0]) && B.in_box(1, pj[1]) && B.in_box(2, pj[2]))) continue;
        double X[3];
        for (int a = 0; a < 3; ++a) X[a] = B.x(a, long(B.cell(a, pj[a])));
        u_at_p[j] += approx_puncture_u(H.pa.p[i], X[0], X[1], X[2]);
      }
  }
}

int interp_points(Hier &H, hipStream_t s, int field, int npts, const int *box, const double *xyz,
                  int ord, double *out, std::string &err) {
  if (npts <= 0) return OLL_OK;
  int *d_box = nullptr;
  double *d_xyz = nullptr, *d_out = nullptr;
  TempBuffers tmp;
  ENG_CHECK(tmp.alloc(&d_box, size_t(npts)));
  ENG_CHECK(tmp.alloc(&d_xyz, size_t(3 * npts)));
  ENG_CHECK(tmp.alloc(&d_out, size_t(npts)));
  ENG_CHECK(hipMemcpyAsync(d_box, box, size_t(npts) * sizeof(int), hipMemcpyHostToDevice, s));
  ENG_CHECK(hipMemcpyAsync(d_xyz, xyz, size_t(3 * npts) * sizeof(double), hipMemcpyHostToDevice, s));
  launch_interp_points(s, H.dbox, field, d_box, d_xyz, npts, ord, d_out);
  ENG_CHECK(hipGetLastError());
  ENG_CHECK(hipMemcpyAsync(out, d_out, size_t(npts) * sizeof(double), hipMemcpyDeviceToHost, s));
  ENG_CHECK(hipStreamSynchronize(s));
  return OLL_OK;
}

}  // namespace oll
