/*
 * Oracle: plain-C restatement of the reference FAS multigrid solve of the
 * puncture Hamiltonian constraint.  TEST INFRASTRUCTURE ONLY: only the
 * test-suite, the smoke check and the benchmark's CPU-baseline leg use it,
 * and only as the checker.  See olliptic_oracle.h.
 *
 * Citations are to files of the reference tree (WuShichao/OllipticID).
 * Floating-point expressions keep the reference's evaluation order so that,
 * in lexicographic mode, iterates agree with the reference binary to
 * rounding (build with -ffp-contract=off).
 */
#include "olliptic_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EPS DBL_EPSILON
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static char g_err[512];

static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char *oracle_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
/* Geometry: domain.cc / domain.h / interface.cc                       */
/* ------------------------------------------------------------------ */

typedef struct {
  size_t n[3];      /* points per axis incl. ghost layers */
  size_t b[3];      /* buffer (ghost) width per axis, >= 1 */
  size_t order;
  double d[3];      /* spacing */
  double c[3];      /* centre */
  double g0[3], g1[3]; /* "global" extents gxi/gxf */
  double lo[3], hi[3]; /* extents xi/xf after Set_Boundary */
  double len[3];    /* Lx = xf - xi */
  int level;
} odom;

/* interface.cc:883-949 */
size_t oracle_valid_n(size_t N) {
  static const size_t n_V[63] = {
      3,   4,   5,   6,   7,   8,   9,   10,  11,  12,  13,  14,  15,
      17,  19,  21,  23,  25,  27,  29,  33,  37,  41,  45,  49,  53,
      57,  65,  73,  81,  89,  97,  105, 113, 129, 145, 161, 177, 193,
      209, 225, 257, 289, 321, 353, 385, 417, 449, 513, 577, 641, 705,
      769, 833, 897, 1153, 1281, 1409, 1665, 1793, 2305, 2817, 3329};
  int valid = 0;
  size_t sug_i = n_V[61], sug_f = n_V[62], i;
  for (int j = 0; j < 63; j++)
    if (n_V[j] == N) valid = 1;
  if (!valid) {
    i = 0;
    while (i < 62 && n_V[i] < N) {
      sug_i = n_V[i];
      sug_f = n_V[i + 1];
      i++;
    }
    if (N - sug_i < sug_f - N)
      N = sug_i;
    else
      N = sug_f;
  }
  return N;
}

/* domain.cc:451-525 (one rank: k = 0, c = 1, no symmetry) */
static void set_boundary_axis(double *ti, double *tf, size_t *Nt, double gti,
                              double gtf, double Lt, double dt) {
  const size_t k = 0, c = 1;
  double lti = gti + (double)k * Lt / (double)c;
  double ltf = gtf - (double)(c - 1 - k) * Lt / (double)c;
  while (*ti + (double)(*Nt - 1) * dt > ltf) (*Nt)--;
  if (*ti + (double)(*Nt - 1) * dt < ltf - 0.5 * dt) (*Nt)++;
  *tf = *ti + (double)(*Nt - 1) * dt;
  while (*tf - (double)(*Nt - 1) * dt < lti) (*Nt)--;
  if (*tf - (double)(*Nt - 1) * dt > lti + 0.5 * dt) (*Nt)++;
  *ti = *tf - (double)(*Nt - 1) * dt;
}

/* domain.cc:27-99 (Make), :111-286 (Set), :293-445 (Set_N), :536-539 */
static void dom_make(odom *D, const size_t Nin[3], const double d[3],
                     const double c[3], size_t ord) {
  memset(D, 0, sizeof *D);
  D->order = ord;
  for (int a = 0; a < 3; a++) {
    size_t N = Nin[a];
    D->d[a] = d[a];
    D->c[a] = c[a];
    D->b[a] = ord / 2 >= 1 ? ord / 2 : 1;
    N = oracle_valid_n(N);
    double gi = c[a] - 0.5 * (double)(N - 1) * d[a] - (double)(D->b[a] - 1) * d[a];
    N += D->b[a] - 1;
    double gf = gi + (double)(N - 1) * d[a] + (double)(D->b[a] - 1) * d[a];
    N += D->b[a] - 1;
    D->g0[a] = gi;
    D->g1[a] = gf;
    D->n[a] = N;
    D->lo[a] = gi;
    D->hi[a] = gf;
    set_boundary_axis(&D->lo[a], &D->hi[a], &D->n[a], gi, gf, gf - gi, d[a]);
    D->len[a] = D->hi[a] - D->lo[a];
  }
}

/* domain.h:871-875 */
static int dcomp(double x, double y) { return fabs(x - y) <= 3 * EPS; }

/* domain.h:701-717 Get_x */
static double dom_x(const odom *D, int a, long i) {
  return D->lo[a] + (double)i * D->d[a];
}

/* domain.cc:1493-1675 Get_i/Get_j/Get_k (binary search) */
static size_t dom_get_i(const odom *D, int a, double x) {
  size_t i = 0;
  int sup = (int)D->n[a];
  int inf = 0;
  int mid = D->n[a] % 2 == 0 ? (int)(D->n[a] / 2) : (int)((D->n[a] - 1) / 2);
  if (D->lo[a] < x && x < D->hi[a]) {
    i = (size_t)mid;
    do {
      if (D->lo[a] + (double)i * D->d[a] > x || dcomp(D->lo[a] + (double)i * D->d[a], x)) {
        sup = (int)i;
        i = (sup + inf) % 2 == 0 ? (size_t)((sup + inf) / 2) : (size_t)((sup + inf - 1) / 2);
      } else {
        inf = (int)i;
        i = (sup + inf) % 2 == 0 ? (size_t)((sup + inf) / 2) : (size_t)((sup + inf - 1) / 2);
      }
    } while (inf + 1 < sup);
  } else
    i = (D->lo[a] >= x) ? 0 : D->n[a] - 2;
  i++;
  return i;
}

/* domain.cc:1682-1706 */
static void dom_find_cell(const odom *D, const double p[3], double lo[3],
                          double hi[3]) {
  for (int a = 0; a < 3; a++) {
    size_t i = dom_get_i(D, a, p[a]);
    lo[a] = dom_x(D, a, (long)i - 1);
    hi[a] = dom_x(D, a, (long)i);
  }
}

/* domain.cc:1811-1871 operator+ */
static void dom_add(odom *out, const odom *A, const odom *B) {
  size_t n[3];
  double c[3], d[3];
  for (int a = 0; a < 3; a++) {
    double xi = A->g0[a] < B->g0[a] ? A->g0[a] : B->g0[a];
    double xf = A->g1[a] > B->g1[a] ? A->g1[a] : B->g1[a];
    c[a] = 0.5 * (xi + xf);
    n[a] = (size_t)((xf - xi) / A->d[a] + 1);
    d[a] = A->d[a];
  }
  dom_make(out, n, d, c, A->order);
}

/* domain.cc:1878-1922 operator-, :1972-1980 operator>=, :1998-2005 operator&&.
 * A && B evaluates B >= (B - A): true when the boxes overlap or touch. */
static int dom_touch(const odom *A, const odom *B) {
  for (int a = 0; a < 3; a++) {
    double ci = B->lo[a] > A->lo[a] ? B->lo[a] : A->lo[a];
    double cf = B->hi[a] < A->hi[a] ? B->hi[a] : A->hi[a];
    double xi = ci < cf ? ci : cf;
    double xf = ci > cf ? ci : cf;
    if (!(B->lo[a] <= xi && xf <= B->hi[a])) return 0;
  }
  return 1;
}

/* domain.h:767-775 Is_In */
static int dom_is_in(const odom *D, const double p[3]) {
  for (int a = 0; a < 3; a++)
    if (!(D->lo[a] - D->d[a] * 0.1 <= p[a] && p[a] <= D->hi[a] + D->d[a] * 0.1))
      return 0;
  return 1;
}

/* domain.h:839-850 is_in_nb */
static int dom_is_in_nb(const odom *D, const double p[3]) {
  for (int a = 0; a < 3; a++)
    if (!(D->lo[a] + D->d[a] * ((double)D->b[a] - 0.1) < p[a] &&
          p[a] < D->hi[a] - D->d[a] * ((double)D->b[a] - 0.1)))
      return 0;
  return 1;
}

/* ------------------------------------------------------------------ */
/* Solver state                                                        */
/* ------------------------------------------------------------------ */

typedef struct {
  odom D;
  size_t npts;
  double *f[OF_NFIELDS];
  double idx[3], idx2[3];
} obox;

struct oracle_solver {
  oracle_params p;
  int levels;        /* total levels incl. multigrid coarsenings */
  int mgrid;         /* number of coarsenings (base box = level mgrid) */
  int nbox;
  obox *box;
  int *lvl_first, *lvl_count;
  /* scalar-field constants, elliptic.cc:71-81 */
  double ps_r0, ps_isigma, ps_phi0, ps_d1, ps_e1, ps_e2, ps_de1;
  double eps_solv, eps_coarse;
  int failed;
  long coarse_iters;  /* sweeps spent in the coarsest-level loop */
};

#define IDX(B, i, j, k) ((size_t)(i) + (B)->D.n[0] * ((size_t)(j) + (B)->D.n[1] * (size_t)(k)))

void oracle_default_params(oracle_params *p) {
  memset(p, 0, sizeof *p);
  p->eq = ORACLE_PUNCTURES;
  p->nbh = 1;
  p->mp[0] = 1.0;
  p->L = 2.0;
  p->N = 30;
  p->levels = 1;
  p->order = 4;
  p->nu1 = 3;
  p->nu2 = 3;
  p->tol = 1e-8;
  p->ps_case = 1;
  p->ps_r0 = 120;
  p->ps_sigma = 8;
  p->ps_phi0 = 0.025;
  p->ps_a2 = 1;
  p->sweep_order = ORACLE_LEX;
}

/* elliptic.cc:300-921 Set_Grids (full domain, V-cycle method) */
static int build_grids(oracle_solver *s) {
  const oracle_params *p = &s->p;
  size_t N = oracle_valid_n((size_t)p->N);
  double Dx = fabs(p->L) / (double)(N - 1.0); /* interface.cc:1061-1065 */
  size_t cap = 64;
  size_t *nx = malloc(cap * sizeof *nx);
  double *dx = malloc(cap * sizeof *dx);
  size_t nl = 1;
  nx[0] = N;
  dx[0] = Dx;
  const size_t base = 12;
  int limit = 0;
  for (; !p->method_app;) {
    size_t ntmp = (nx[nl - 1] + 1) / 2;
    if (ntmp < base) {
      limit = 1;
      ntmp = nx[nl - 1];
    }
    if (nl == cap) break;
    nx[nl] = ntmp;
    dx[nl] = (nx[nl] == nx[nl - 1]) ? dx[nl - 1] : dx[nl - 1] * 2.0;
    nl++;
    if (limit) { /* the three axes are identical here */
      nl--;
      break;
    }
  }
  if (nl <= 1 && !p->method_app) {
    set_err("There are not enough grid-points for multigrid");
    free(nx);
    free(dx);
    return -1;
  }
  s->mgrid = (int)(nl - 1);
  s->levels = p->levels + s->mgrid;
  int L = s->levels;
  int nb = p->nbh;
  /* worst case: one box per puncture per refinement level */
  size_t maxbox = (size_t)(s->mgrid + 1) + (size_t)(L) * (size_t)nb + 4;
  odom *D = calloc(maxbox, sizeof *D);
  int *ndl = calloc((size_t)L + 1, sizeof *ndl);
  size_t nd = 0;
  for (int l = s->mgrid; l >= 0; l--) {
    size_t gz = l == 0 ? (size_t)p->order : 2;
    size_t n3[3] = {nx[l], nx[l], nx[l]};
    double d3[3] = {dx[l], dx[l], dx[l]};
    double c3[3] = {0, 0, 0};
    dom_make(&D[nd++], n3, d3, c3, gz);
    ndl[s->mgrid - l] = 1;
  }
  for (int l = s->mgrid + 1; l < L; l++) {
    for (int i = 0; i < nb; i++) {
      double pot = l - s->mgrid;
      double _d = Dx / pow(2.0, pot);
      double pos[3] = {p->x[i], p->y[i], p->z[i]};
      double lo[3], hi[3], xc[3];
      dom_find_cell(&D[s->mgrid], pos, lo, hi);
      for (int a = 0; a < 3; a++) {
        if (dcomp(lo[a], pos[a]) || dcomp(hi[a], pos[a]))
          xc[a] = pos[a];
        else {
          while (lo[a] + _d < pos[a]) lo[a] += _d;
          while (hi[a] - _d > pos[a]) hi[a] -= _d;
          xc[a] = fabs(lo[a] - pos[a]) < fabs(hi[a] - pos[a]) ? lo[a] : hi[a];
        }
      }
      size_t n3[3] = {N, N, N};
      double d3[3] = {_d, _d, _d};
      dom_make(&D[nd++], n3, d3, xc, (size_t)p->order);
    }
    ndl[l] = nb;
  }
  /* merge overlapping boxes, elliptic.cc:839-896 */
  for (int l = s->mgrid + 1; l < L; l++) {
    int restore = 0;
    for (int i = 0; i < ndl[l]; i++)
      for (int j = 0; j < ndl[l]; j++)
        if (i != j) {
          size_t ib = 0, jb = 0;
          for (int ll = 0; ll < l; ll++) ib += (size_t)ndl[ll];
          jb = ib + (size_t)j;
          ib += (size_t)i;
          /* box i of level l is addressed in the list of every level's
           * boxes: after a merge erased an earlier box, D[ib] may be a box
           * of level l + 1, as in the reference; past the end of the list
           * (where the reference reads past its vector) the set-up fails */
          if (ib >= nd || jb >= nd) {
            set_err("box merging at level %d reaches past the box list", l);
            free(D);
            free(ndl);
            free(nx);
            free(dx);
            return -1;
          }
          if (dom_touch(&D[ib], &D[jb])) {
            odom m;
            dom_add(&m, &D[ib], &D[jb]);
            D[ib] = m;
            memmove(&D[jb], &D[jb + 1], (nd - jb - 1) * sizeof *D);
            nd--;
            ndl[l]--;
            restore = 1;
          }
        }
    if (restore) l--;
  }
  s->nbox = (int)nd;
  s->box = calloc(nd, sizeof *s->box);
  s->lvl_first = calloc((size_t)L, sizeof(int));
  s->lvl_count = calloc((size_t)L, sizeof(int));
  size_t ib = 0;
  for (int l = 0; l < L; l++) {
    s->lvl_first[l] = (int)ib;
    s->lvl_count[l] = ndl[l];
    for (int i = 0; i < ndl[l]; i++) {
      D[ib].level = l;
      s->box[ib].D = D[ib];
      ib++;
    }
  }
  free(D);
  free(ndl);
  free(nx);
  free(dx);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Coefficients: operator.cc:752-868, 2193-2380; elliptic.h:1099-1147  */
/* ------------------------------------------------------------------ */

/* operator.cc:2193-2223 */
static double punctures_c(const oracle_solver *s, double X, double Y, double Z) {
  double cval = 1;
  for (int i = 0; i < s->p.nbh; i++) {
    double x = X - s->p.x[i], y = Y - s->p.y[i], z = Z - s->p.z[i];
    double r = sqrt(x * x + y * y + z * z);
    if (r < 2 * EPS) r = 2 * EPS;
    cval += 0.5 * s->p.mp[i] / r;
  }
  return cval;
}

/* operator.cc:3043-3310 EM_BH_phi / EM_BH_eta and their gradients: the
 * charge and mass potentials phi = sum q / r, eta = sum m / (2 r); g[0] is
 * the potential, g[1..3] its x, y, z derivatives (r and r^3 clamped at
 * 2 eps) */
static void em_potential(const oracle_solver *s, const double *w, double scale, double X,
                         double Y, double Z, double g[4]) {
  g[0] = g[1] = g[2] = g[3] = 0;
  for (int i = 0; i < s->p.nbh; i++) {
    const double x = X - s->p.x[i], y = Y - s->p.y[i], z = Z - s->p.z[i];
    double r = sqrt(x * x + y * y + z * z);
    if (r < 2 * EPS) r = 2 * EPS;
    g[0] += w[i] / r;
    double r3 = pow(x * x + y * y + z * z, 1.5);
    if (r3 < 2 * EPS) r3 = 2 * EPS;
    g[1] -= w[i] * x / r3;
    g[2] -= w[i] * y / r3;
    g[3] -= w[i] * z / r3;
  }
  for (int a = 0; a < 4; a++) g[a] *= scale;
}

/* elliptic.h:1458-1530 EM_BH_A1 .. EM_BH_E: the thirteen coefficients of
 * Punctures_EM at one point, in oracle_field order from OF_A1 */
static void em_coefficients(const oracle_solver *s, double x, double y, double z, double out[13]) {
  double e[4], f[4];
  em_potential(s, s->p.mp, 0.5, x, y, z, e);
  em_potential(s, s->p.qp, 1.0, x, y, z, f);
  const double eta = e[0], phi = f[0];
  const double f2 = pow(f[1], 2) + pow(f[2], 2) + pow(f[3], 2);
  const double fe = f[1] * e[1] + f[2] * e[2] + f[3] * e[3];
  out[0] = 3 * eta;
  out[1] = 3 * pow(eta, 2) - 0.25 * pow(phi, 2);
  out[2] = pow(eta, 3) - 0.25 * eta * phi * phi;
  out[3] = -0.25 * pow(phi, 2);
  for (int a = 0; a < 3; a++) {
    out[4 + a] = 0.5 * phi * f[1 + a];
    out[7 + a] = 0.5 * phi * (eta * f[1 + a] - phi * e[1 + a]);
  }
  out[10] = -0.25 * f2;
  out[11] = 0.5 * phi * fe - 0.5 * eta * f2;
  out[12] = -0.25 * phi * phi * (e[1] * e[1] + e[2] * e[2] + e[3] * e[3]) + 0.5 * phi * eta * fe +
            0.25 * (1 - eta * eta) * (f[1] * f[1] + f[2] * f[2] + f[3] * f[3]);
}

/* operator.cc:2228-2333 */
static double punctures_b(const oracle_solver *s, double X, double Y, double Z) {
  double A[3][3] = {{0}};
  const oracle_params *p = &s->p;
  for (int i = 0; i < p->nbh; i++) {
    double x = X - p->x[i], y = Y - p->y[i], z = Z - p->z[i];
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double r2 = x2 + y2 + z2;
    if (r2 < 2 * EPS) r2 = 2 * EPS;
    double r = sqrt(r2);
    double ir5 = 1.0 / (r2 * r2 * r);
    A[0][0] += 1.5 * ir5 * (p->px[i] * x * (x2 + r2) + p->py[i] * y * (x2 - r2) +
                            p->pz[i] * z * (x2 - r2) +
                            4.0 * x * (p->sy[i] * z - p->sz[i] * y));
    A[1][1] += 1.5 * ir5 * (p->px[i] * x * (y2 - r2) + p->py[i] * y * (y2 + r2) +
                            p->pz[i] * z * (y2 - r2) +
                            4.0 * y * (p->sz[i] * x - p->sx[i] * z));
    A[2][2] += 1.5 * ir5 * (p->px[i] * x * (z2 - r2) + p->py[i] * y * (z2 - r2) +
                            p->pz[i] * z * (z2 + r2) +
                            4.0 * z * (p->sx[i] * y - p->sy[i] * x));
    A[0][1] += 1.5 * ir5 * (p->px[i] * y * (r2 + x2) + p->py[i] * x * (r2 + y2) +
                            p->pz[i] * x * y * z - 2.0 * p->sx[i] * x * z +
                            2.0 * p->sy[i] * y * z + 2.0 * p->sz[i] * (x2 - y2));
    A[0][2] += 1.5 * ir5 * (p->px[i] * z * (r2 + x2) + p->py[i] * x * y * z +
                            p->pz[i] * x * (r2 + z2) + 2.0 * p->sx[i] * x * y +
                            2.0 * p->sy[i] * (z2 - x2) - 2.0 * p->sz[i] * y * z);
    A[1][2] += 1.5 * ir5 * (p->px[i] * x * y * z + p->py[i] * z * (r2 + y2) +
                            p->pz[i] * y * (r2 + z2) + 2.0 * p->sx[i] * (y2 - z2) -
                            2.0 * p->sy[i] * x * y + 2.0 * p->sz[i] * x * z);
  }
  A[1][0] = A[0][1];
  A[2][0] = A[0][2];
  A[2][1] = A[1][2];
  double bval = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) bval += A[i][j] * A[i][j];
  return 0.125 * bval;
}

/* elliptic.h:1099-1147 */
static double sf_phi(const oracle_solver *s, double r) {
  switch (s->p.ps_case) {
  case 1: return s->ps_phi0 * tanh((r - s->ps_r0) * s->ps_isigma);
  case 2: return s->ps_phi0 * exp(-(r - s->ps_r0) * (r - s->ps_r0) * s->ps_isigma);
  case 3: {
    double x = r, l2 = 0.6;
    return (0.0481646 * exp(-0.0581545 * (x - 1.8039e-08) * (x - 1.8039e-08) / (l2)) +
            0.298408 * exp(-0.111412 * (x + 9.6741e-09) * (x + 9.6741e-09) / l2) +
            0.42755 * exp(-0.207156 * (x - 1.09822e-08) * (x - 1.09822e-08) / l2) +
            0.204229 * exp(-0.37742 * (x + 2.13778e-08) * (x + 2.13778e-08) / l2) +
            0.021649 * exp(-0.68406 * (x - 8.78608e-08) * (x - 8.78608e-08) / l2)) * l2;
  }
  default:
    return s->ps_phi0 * 0.5 * (tanh((r + s->ps_r0) * s->ps_isigma) - tanh((r - s->ps_r0) * s->ps_isigma));
  }
}

static double sf_dphi(const oracle_solver *s, double r) {
  switch (s->p.ps_case) {
  case 1: return s->ps_d1 / pow(cosh((r - s->ps_r0) * s->ps_isigma), 4);
  case 2:
    return s->ps_de1 * (r - s->ps_r0) * (r - s->ps_r0) *
           exp(-2 * s->ps_isigma * (r - s->ps_r0) * (r - s->ps_r0));
  case 3: {
    double x = r, l2 = 10000;
    return M_PI * pow(-2 * (0.0581545 * (x - 1.8039e-08) * 0.0481646 * exp(-0.0581545 * (x - 1.8039e-08) * (x - 1.8039e-08) / (l2)) +
                            0.111412 * (x + 9.6741e-09) * 0.298408 * exp(-0.111412 * (x + 9.6741e-09) * (x + 9.6741e-09) / l2) +
                            0.207156 * (x - 1.09822e-08) * 0.42755 * exp(-0.207156 * (x - 1.09822e-08) * (x - 1.09822e-08) / l2) +
                            0.37742 * (x + 2.13778e-08) * 0.204229 * exp(-0.37742 * (x + 2.13778e-08) * (x + 2.13778e-08) / l2) +
                            0.68406 * (x - 8.78608e-08) * 0.021649 * exp(-0.68406 * (x - 8.78608e-08) * (x - 8.78608e-08) / l2)) /
                          (l2 * l2), 2);
  }
  default:
    return s->ps_de1 * 0.25 *
           pow(pow(1. / cosh((r + s->ps_r0) * s->ps_isigma), 2) - pow(1. / cosh((r - s->ps_r0) * s->ps_isigma), 2), 2);
  }
}

/* operator.cc:2339-2380 */
static double punctures_scalar_d(const oracle_solver *s, double X, double Y, double Z) {
  double r = sqrt(X * X + Y * Y + Z * Z);
  return sf_dphi(s, r);
}

static double punctures_scalar_e(const oracle_solver *s, double X, double Y, double Z) {
  if (s->p.ps_case == 2 || s->p.ps_case == 3 || s->p.ps_case == 4) {
    double r = sqrt(X * X + Y * Y + Z * Z);
    double ff = s->ps_e2 * sf_phi(s, r);
    return s->ps_e1 * (1 - exp(ff)) * (1 - exp(ff)) * exp(-2 * ff);
  }
  return 0.0;
}

/* ------------------------------------------------------------------ */
/* Stencils: ffunction.h:1290-1524, diff_4th.cc:392-412,               */
/* diff_8th.cc:596-676, diff_2nd.cc:237-335, diff_4th.cc:6-89          */
/* ------------------------------------------------------------------ */

static double lap_at(const obox *B, const double *u, size_t i, size_t j, size_t k) {
  const size_t sx = 1, sy = B->D.n[0], sz = B->D.n[0] * B->D.n[1];
  const size_t p = IDX(B, i, j, k);
  const double u0 = u[p];
  const double ix2 = B->idx2[0], iy2 = B->idx2[1], iz2 = B->idx2[2];
#define U(o) u[p + (o)]
#define UM(o) u[p - (o)]
  switch (B->D.order) {
  case 2:
    return (U(sx) - 2.0 * u0 + UM(sx)) * ix2 + (U(sy) - 2.0 * u0 + UM(sy)) * iy2 +
           (U(sz) - 2.0 * u0 + UM(sz)) * iz2;
  case 4:
    return ((-(U(2 * sx) + UM(2 * sx)) + 16.0 * (U(sx) + UM(sx)) - 30.0 * u0) * ix2 +
            (-(U(2 * sy) + UM(2 * sy)) + 16.0 * (U(sy) + UM(sy)) - 30.0 * u0) * iy2 +
            (-(U(2 * sz) + UM(2 * sz)) + 16.0 * (U(sz) + UM(sz)) - 30.0 * u0) * iz2) /
           12.0;
  case 6: {
    const double i180 = 1.0 / 180;
    return (2.0 * (U(3 * sx) + UM(3 * sx)) -
            27.0 * (U(2 * sx) + UM(2 * sx) - 10.0 * (U(sx) + UM(sx))) - 490.0 * u0) * ix2 * i180 +
           (2.0 * (U(3 * sy) + UM(3 * sy)) -
            27.0 * (U(2 * sy) + UM(2 * sy) - 10.0 * (U(sy) + UM(sy))) - 490.0 * u0) * iy2 * i180 +
           (2.0 * (U(3 * sz) + UM(3 * sz)) -
            27.0 * (U(2 * sz) + UM(2 * sz) - 10.0 * (U(sz) + UM(sz))) - 490.0 * u0) * iz2 * i180;
  }
  default:
    return (-9.0 * (U(4 * sx) + UM(4 * sx)) + 128.0 * (U(3 * sx) + UM(3 * sx)) -
            1008.0 * (U(2 * sx) + UM(2 * sx)) + 8064.0 * (U(sx) + UM(sx)) - 14350.0 * u0) * ix2 / 5040.0 +
           (-9.0 * (U(4 * sy) + UM(4 * sy)) + 128.0 * (U(3 * sy) + UM(3 * sy)) -
            1008.0 * (U(2 * sy) + UM(2 * sy)) + 8064.0 * (U(sy) + UM(sy)) - 14350.0 * u0) * iy2 / 5040.0 +
           (-9.0 * (U(4 * sz) + UM(4 * sz)) + 128.0 * (U(3 * sz) + UM(3 * sz)) -
            1008.0 * (U(2 * sz) + UM(2 * sz)) + 8064.0 * (U(sz) + UM(sz)) - 14350.0 * u0) * iz2 / 5040.0;
  }
#undef U
#undef UM
}

static double dulap(const obox *B) {
  const double s = B->idx2[0] + B->idx2[1] + B->idx2[2];
  switch (B->D.order) {
  case 2: return -2.0 * s;
  case 4: return -2.5 * s;
  case 6: return -490 * s * (1.0 / 180);
  default: return -205 * s / 72.0;
  }
}

/* First derivative along axis a at (i,j,k): dx2 (order 2) or dx4 (orders
 * 4,6,8; ffunction.cc:55-314 selects dx4 for all three).  *coef receives
 * the coefficient of u at the point itself (dudx2/dudx4). */
static double deriv(const obox *B, const double *u, int a, const size_t ijk[3],
                    double *coef) {
  const size_t st[3] = {1, B->D.n[0], B->D.n[0] * B->D.n[1]};
  const size_t p = IDX(B, ijk[0], ijk[1], ijk[2]);
  const size_t s = st[a];
  const long n = (long)B->D.n[a];
  const long i = (long)ijk[a];
  const double id = B->idx[a];
  if (B->D.order == 2) {
    if (i == 0) {
      if (coef) *coef = -3.0 * 0.5 * id;
      return (-3.0 * u[p] + 4.0 * u[p + s] - u[p + 2 * s]) * 0.5 * id;
    } else if (i == n - 1) {
      if (coef) *coef = 3.0 * 0.5 * id;
      return (3.0 * u[p] - 4.0 * u[p - s] + u[p - 2 * s]) * 0.5 * id;
    } else {
      if (coef) *coef = 0.0;
      return (u[p + s] - u[p - s]) * 0.5 * id;
    }
  } else {
    const long nm = n % 2 == 0 ? n / 2 : (n - 1) / 2;
    if (i <= nm) {
      if (coef) *coef = -25 * id / 12.0;
      return -(25 * u[p] - 48 * u[p + s] + 36 * u[p + 2 * s] - 16 * u[p + 3 * s] +
               3 * u[p + 4 * s]) * id / 12.0;
    } else {
      if (coef) *coef = 25 * id / 12.0;
      return (25 * u[p] - 48 * u[p - s] + 36 * u[p - 2 * s] - 16 * u[p - 3 * s] +
              3 * u[p - 4 * s]) * id / 12.0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Pointwise operators: elliptic.h:405-445                             */
/* ------------------------------------------------------------------ */

static void op_at(const oracle_solver *s, const obox *B, size_t i, size_t j,
                  size_t k, double *N, double *dN) {
  const double *u = B->f[OF_U];
  const size_t p = IDX(B, i, j, k);
  const double bb = B->f[OF_B][p], cu = B->f[OF_C][p] + u[p];
  double lap = lap_at(B, u, i, j, k);
  if (s->p.eq == ORACLE_PUNCTURES_EM) {
    /* elliptic.h:510-626 puncture_EM_OP / puncture_EM_duOP:
     *   N = Lap U + [B |grad U|^2 + (B1 U + B2) . grad U + C U^2 + D U + E]
     *               / (U^3 + A1 U^2 + A2 U + A3)
     * and its derivative in U at the point.  Dx, Dy, Dz and their
     * coefficients at the point are deriv's (centred at order 2, one-sided
     * five-point at orders 4-8: the reference's Dx as its solver runs it) */
    const double uu = u[p], uu2 = uu * uu;
    const size_t ijk[3] = {i, j, k};
    double du[3];
    const double udx = deriv(B, u, 0, ijk, &du[0]), udy = deriv(B, u, 1, ijk, &du[1]),
                 udz = deriv(B, u, 2, ijk, &du[2]);
    double c[13];
    for (int q = 0; q < 13; q++) c[q] = B->f[OF_A1 + q][p];
    const double num = c[3] * (udx * udx + udy * udy + udz * udz) + (c[4] * uu + c[7]) * udx +
                       (c[5] * uu + c[8]) * udy + (c[6] * uu + c[9]) * udz + c[10] * uu2 +
                       c[11] * uu + c[12];
    const double den = uu2 * uu + c[0] * uu2 + c[1] * uu + c[2];
    *N = lap + num / den;
    if (dN) {
      const double dnum = 2 * c[3] * (udx * du[0] + udy * du[1] + udz * du[2]) +
                          (c[4] * uu + c[7]) * du[0] + c[4] * udx + (c[5] * uu + c[8]) * du[1] +
                          c[5] * udy + (c[6] * uu + c[9]) * du[2] + c[6] * udz + 2 * c[10] * uu +
                          c[11];
      *dN = dulap(B) + (dnum * den - num * (3 * uu2 + 2 * c[0] * uu + c[1])) / pow(den, 2);
    }
    return;
  }
  if (s->p.eq == ORACLE_PUNCTURES) {
    *N = lap + bb / pow(cu, 7);
    if (dN) *dN = dulap(B) - 7.0 * bb / pow(cu, 8);
  } else if (s->p.eq == ORACLE_POISSON) {
    /* elliptic.h:357-370 poisson_OP / poisson_duOP (c = 0: cu = u) */
    *N = lap + cu * cu;
    if (dN) *dN = dulap(B) + 2.0 * cu;
  } else if (s->p.eq == ORACLE_BRILL) {
    /* elliptic.h:381-398 brill_OP / brill_duOP, b = Lq (c = 0) */
    *N = lap + bb * cu;
    if (dN) *dN = dulap(B) + bb;
  } else {
    const double dd = B->f[OF_D][p], ee = B->f[OF_E][p];
    *N = lap + bb / pow(cu, 7) + dd * cu + ee * pow(cu, 5);
    if (dN) *dN = dulap(B) - 7.0 * bb / pow(cu, 8) + dd + 5 * ee * pow(cu, 4);
  }
}

static double smooth_omega(oracle_solver *s, int ibox, int nu, double omega);

/* elliptic.cc:1126-1242 Smooth; lexicographic or multi-colour order */
double oracle_smooth(oracle_solver *s, int ibox, int nu) { return smooth_omega(s, ibox, nu, 1.0); }

/* colour order of a two-colour box (the GPU's choice): when the box's
 * points that coincide with nodes of its parent (the coarser box holding
 * its centre) are of colour 1, its sweeps take colour 1 first; a box off
 * its parent's lattice sweeps in ascending order */
static int colour_flip(const oracle_solver *s, const obox *B) {
  const int lev = B->D.level;
  if (lev == 0) return 0;
  /* the parent: the coarser box holding the box's centre */
  const obox *C = &s->box[s->lvl_first[lev - 1]];
  for (int q = 0; q < s->lvl_count[lev - 1]; q++) {
    const obox *P = &s->box[s->lvl_first[lev - 1] + q];
    int in = 1;
    for (int a = 0; a < 3; a++) in = in && B->D.c[a] >= P->D.lo[a] && B->D.c[a] <= P->D.hi[a];
    if (in) {
      C = P;
      break;
    }
  }
  int sum = 0;
  for (int a = 0; a < 3; a++) {
    const double t = (dom_x(&B->D, a, 0) - dom_x(&C->D, a, 0)) / B->D.d[a];
    const double m = nearbyint(t);
    if (fabs(t - m) > 1e-6) return 0; /* off the parent's lattice: no coincident points */
    sum += (int)fmod(fabs(m), 2.0);
  }
  return sum & 1;
}

int oracle_colour_flip(const oracle_solver *s, int ibox) {
  if (ibox < 0 || ibox >= s->nbox) return -1;
  const obox *B = &s->box[ibox];
  const int ncol = (s->p.eq == ORACLE_PUNCTURES_EM && B->D.order >= 4) ? 5 : (int)B->D.b[0] + 1;
  return ncol == 2 ? colour_flip(s, B) : 0;
}

static double smooth_omega(oracle_solver *s, int ibox, int nu, double omega) {
  obox *B = &s->box[ibox];
  double *u = B->f[OF_U], *r = B->f[OF_R], *rho = B->f[OF_RHO];
  const size_t R = B->D.b[0];
  const size_t n0 = B->D.n[0], n1 = B->D.n[1], n2 = B->D.n[2];
  /* colours: R + 1; 5 for Punctures_EM at orders 4-8, whose one-sided
   * first derivatives reach 4 points (the GPU's choice) */
  const int ncol = s->p.sweep_order != ORACLE_COLOUR ? 1
                   : (s->p.eq == ORACLE_PUNCTURES_EM && B->D.order >= 4) ? 5 : (int)R + 1;
  double gLinf = 0;
  for (int kk = 0; kk < nu; kk++) {
    double Linf = 0;
    for (int q = 0; q < ncol; q++)
      for (int col = (ncol == 2 && colour_flip(s, B)) ? ncol - 1 - q : q; col >= 0; col = -1)
      for (size_t k = R; k < n2 - R; k++)
        for (size_t j = R; j < n1 - R; j++)
          for (size_t i = R; i < n0 - R; i++) {
            if (ncol > 1 && (int)((i + j + k) % (size_t)ncol) != col) continue;
            const size_t p = IDX(B, i, j, k);
            double N, dN;
            op_at(s, B, i, j, k, &N, &dN);
            r[p] = N - rho[p];
            if (omega == 1.0)
              u[p] -= r[p] / dN;
            else
              u[p] -= omega * (r[p] / dN);
            double f = fabs(r[p]);
            if (Linf < f) Linf = f;
            if (!(u[p] == u[p])) {
              set_err("NAN check error: box %d point (%zu,%zu,%zu)", ibox, i, j, k);
              s->failed = 1;
              return NAN;
            }
          }
    gLinf = Linf > gLinf ? Linf : gLinf;
  }
  return gLinf;
}

/* elliptic.cc:944-991 App_Ell_Operator */
double oracle_residual(oracle_solver *s, int ibox) {
  obox *B = &s->box[ibox];
  double *r = B->f[OF_R];
  const size_t R = B->D.b[0];
  double Linf = 0;
  for (size_t k = R; k < B->D.n[2] - R; k++)
    for (size_t j = R; j < B->D.n[1] - R; j++)
      for (size_t i = R; i < B->D.n[0] - R; i++) {
        double N;
        op_at(s, B, i, j, k, &N, NULL);
        r[IDX(B, i, j, k)] = N;
        double f = fabs(N);
        if (Linf < f) Linf = f;
      }
  return Linf;
}

/* ------------------------------------------------------------------ */
/* Robin boundary: boundary.cc:447-800, 915-1101, 1118-1710            */
/* Region iteration order: domain.cc:575-1388 (Fix_Points)             */
/* ------------------------------------------------------------------ */

enum { S_LO = 0, S_HI = 1, S_MID = 2 };

typedef struct {
  int side[3];   /* per axis */
  int loop[3];   /* axes, outermost first */
} oregion;

/* faces Xi,Xf,Yi,Yf,Zi,Zf; edges XY(4),XZ(4),YZ(4); corners (8),
 * in the call order of App_Bound / App_Boundary_Operator */
static const oregion REGIONS[26] = {
    {{S_LO, S_MID, S_MID}, {0, 2, 1}}, {{S_HI, S_MID, S_MID}, {0, 2, 1}},
    {{S_MID, S_LO, S_MID}, {1, 2, 0}}, {{S_MID, S_HI, S_MID}, {1, 2, 0}},
    {{S_MID, S_MID, S_LO}, {2, 1, 0}}, {{S_MID, S_MID, S_HI}, {2, 1, 0}},
    {{S_LO, S_LO, S_MID}, {0, 1, 2}},  {{S_LO, S_HI, S_MID}, {0, 1, 2}},
    {{S_HI, S_LO, S_MID}, {0, 1, 2}},  {{S_HI, S_HI, S_MID}, {0, 1, 2}},
    {{S_LO, S_MID, S_LO}, {0, 2, 1}},  {{S_LO, S_MID, S_HI}, {0, 2, 1}},
    {{S_HI, S_MID, S_LO}, {0, 2, 1}},  {{S_HI, S_MID, S_HI}, {0, 2, 1}},
    {{S_MID, S_LO, S_LO}, {2, 1, 0}},  {{S_MID, S_LO, S_HI}, {2, 1, 0}},
    {{S_MID, S_HI, S_LO}, {2, 1, 0}},  {{S_MID, S_HI, S_HI}, {2, 1, 0}},
    {{S_LO, S_LO, S_LO}, {2, 1, 0}},   {{S_LO, S_LO, S_HI}, {2, 1, 0}},
    {{S_HI, S_LO, S_LO}, {2, 1, 0}},   {{S_HI, S_LO, S_HI}, {2, 1, 0}},
    {{S_LO, S_HI, S_LO}, {2, 1, 0}},   {{S_LO, S_HI, S_HI}, {2, 1, 0}},
    {{S_HI, S_HI, S_LO}, {2, 1, 0}},   {{S_HI, S_HI, S_HI}, {2, 1, 0}},
};

/* position t (0-based) in the iteration range of one axis side */
static size_t side_len(const obox *B, int a, int side) {
  return side == S_MID ? B->D.n[a] - 2 * B->D.b[a] : B->D.b[a];
}
static size_t side_idx(const obox *B, int a, int side, size_t t) {
  if (side == S_LO) return B->D.b[a] - 1 - t;            /* b-1 ... 0 */
  if (side == S_HI) return B->D.n[a] - B->D.b[a] + t;    /* n-b ... n-1 */
  return B->D.b[a] + t;                                  /* b ... n-b-1 */
}

static void robin_region(oracle_solver *s, obox *B, const oregion *rg, int apply) {
  double *u = B->f[OF_U];
  double *S = B->f[OF_RHO];
  double *R = B->f[OF_R];
  const double ipn = 1.0 / (double)1; /* falloff_n = 1, operator.cc:785-791 */
  (void)s;
  size_t len[3], t[3];
  for (int a = 0; a < 3; a++) len[a] = side_len(B, a, rg->side[a]);
  const int A0 = rg->loop[0], A1 = rg->loop[1], A2 = rg->loop[2];
  for (t[A0] = 0; t[A0] < len[A0]; t[A0]++)
    for (t[A1] = 0; t[A1] < len[A1]; t[A1]++)
      for (t[A2] = 0; t[A2] < len[A2]; t[A2]++) {
        size_t ijk[3];
        for (int a = 0; a < 3; a++) ijk[a] = side_idx(B, a, rg->side[a], t[a]);
        const size_t p = IDX(B, ijk[0], ijk[1], ijk[2]);
        double X[3];
        for (int a = 0; a < 3; a++) X[a] = dom_x(&B->D, a, (long)ijk[a]);
        double x2 = X[0] * X[0], y2 = X[1] * X[1], z2 = X[2] * X[2];
        double r2 = x2 + y2 + z2;
        double sF = 0, sdF = 0, dv1 = 0, cf1 = 0;
        int nd = 0, a1 = 0;
        for (int a = 0; a < 3; a++) {
          if (rg->side[a] == S_MID) continue;
          double cf;
          double dv = deriv(B, u, a, ijk, &cf);
          if (nd == 0) {
            dv1 = dv;
            cf1 = cf;
            a1 = a;
            sF = dv / X[a];
            sdF = cf / X[a];
          } else {
            sF += dv / X[a];
            sdF += cf / X[a];
          }
          nd++;
        }
        /* faces evaluate r2*ipn*du/x (boundary.cc:926-930); edges and
         * corners r2*ipn*(du1/x1 + du2/x2 [+ du3/x3]) (:961-968, :1003-1012) */
        double tF = nd == 1 ? r2 * ipn * dv1 / X[a1] : r2 * ipn * sF;
        double tdF = nd == 1 ? r2 * ipn * cf1 / X[a1] : r2 * ipn * sdF;
        if (apply) {
          double F = tF + u[p] - S[p];
          double dF = tdF + 1.0;
          u[p] = u[p] - F / dF;
        } else {
          R[p] = tF + u[p];
        }
      }
}

/* boundary.cc:447-800 App_Bound (Robin branch) */
void oracle_robin(oracle_solver *s, int ibox) {
  for (int r = 0; r < 26; r++) robin_region(s, &s->box[ibox], &REGIONS[r], 1);
}

/* boundary.cc:1118-1535 App_Boundary_Operator (Robin branch) */
void oracle_boundary_operator(oracle_solver *s, int ibox) {
  for (int r = 0; r < 26; r++) robin_region(s, &s->box[ibox], &REGIONS[r], 0);
}

/* ------------------------------------------------------------------ */
/* Interpolation: interpolation.cc:125-375                             */
/* ------------------------------------------------------------------ */

static double get_val(const obox *B, const double *f, long i, long j, long k) {
  return f[IDX(B, i, j, k)];
}

/* interpolation.cc:125-182 */
static double linear_interp(const obox *B, const double *f, double x, double y, double z) {
  size_t i = dom_get_i(&B->D, 0, x), j = dom_get_i(&B->D, 1, y), k = dom_get_i(&B->D, 2, z);
  if (dcomp(x, dom_x(&B->D, 0, (long)i)) && dcomp(y, dom_x(&B->D, 1, (long)j)) &&
      dcomp(z, dom_x(&B->D, 2, (long)k)))
    return get_val(B, f, (long)i, (long)j, (long)k);
  i = i == 0 ? 1 : i;
  j = j == 0 ? 1 : j;
  k = k == 0 ? 1 : k;
  double dxf = (x - dom_x(&B->D, 0, (long)i - 1)) * B->idx[0];
  double dyf = (y - dom_x(&B->D, 1, (long)j - 1)) * B->idx[1];
  double dzf = (z - dom_x(&B->D, 2, (long)k - 1)) * B->idx[2];
  double dxi = (1 - dxf), dyi = (1 - dyf), dzi = (1 - dzf);
  long I = (long)i, J = (long)j, K = (long)k;
  return ((get_val(B, f, I - 1, J - 1, K - 1) * dxi + get_val(B, f, I, J - 1, K - 1) * dxf) * dyi +
          (get_val(B, f, I - 1, J, K - 1) * dxi + get_val(B, f, I, J, K - 1) * dxf) * dyf) * dzi +
         ((get_val(B, f, I - 1, J - 1, K) * dxi + get_val(B, f, I, J - 1, K) * dxf) * dyi +
          (get_val(B, f, I - 1, J, K) * dxi + get_val(B, f, I, J, K) * dxf) * dyf) * dzf;
}

/* interpolation.cc:309-375 */
static double lagrange_poly(const obox *B, int a, double t, long i, long ni, long nf) {
  double lp = 1.0;
  const double ti = dom_x(&B->D, a, i);
  for (long j = ni; j < i; j++) lp *= (t - dom_x(&B->D, a, j)) / (ti - dom_x(&B->D, a, j));
  for (long j = i + 1; j <= nf; j++) lp *= (t - dom_x(&B->D, a, j)) / (ti - dom_x(&B->D, a, j));
  return lp;
}

/* interpolation.cc:189-303 */
static double lagrange_interp(const obox *B, const double *f, int ord, double x,
                              double y, double z) {
  size_t i = dom_get_i(&B->D, 0, x), j = dom_get_i(&B->D, 1, y), k = dom_get_i(&B->D, 2, z);
  if (dcomp(x, dom_x(&B->D, 0, (long)i)) && dcomp(y, dom_x(&B->D, 1, (long)j)) &&
      dcomp(z, dom_x(&B->D, 2, (long)k)))
    return get_val(B, f, (long)i, (long)j, (long)k);
  long st[3];
  const size_t ijk[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    long s0 = ord % 2 == 0 ? (long)ijk[a] - ord / 2 : (long)ijk[a] - (ord + 1) / 2;
    if (s0 < 0)
      s0 = 0;
    else if (s0 + ord >= (long)B->D.n[a])
      s0 = (long)B->D.n[a] - ord - 1;
    st[a] = s0;
  }
  double uk[10][10];
  for (int I = 0; I <= ord; I++)
    for (int J = 0; J <= ord; J++) {
      uk[I][J] = 0;
      for (int K = 0; K <= ord; K++)
        uk[I][J] += get_val(B, f, st[0] + I, st[1] + J, st[2] + K) *
                    lagrange_poly(B, 2, z, st[2] + K, st[2], st[2] + ord);
    }
  double ujk[10];
  for (int I = 0; I <= ord; I++) {
    ujk[I] = 0;
    for (int J = 0; J <= ord; J++)
      ujk[I] += uk[I][J] * lagrange_poly(B, 1, y, st[1] + J, st[1], st[1] + ord);
  }
  double fxyz = 0;
  for (int I = 0; I <= ord; I++)
    fxyz += ujk[I] * lagrange_poly(B, 0, x, st[0] + I, st[0], st[0] + ord);
  return fxyz;
}

static double interp(const obox *B, const double *f, int method, int ord, double x,
                     double y, double z) {
  return method == OI_LINEAR ? linear_interp(B, f, x, y, z) : lagrange_interp(B, f, ord, x, y, z);
}

/* elliptic.cc:2037-2213 (and 2286-2419): the term of puncture i of the
 * approximate solution at (x, y, z) */
static double approx_term(const oracle_params *p, int i, double x, double y, double z) {
  const double sbl = sqrt(p->sx[i] * p->sx[i] + p->sy[i] * p->sy[i] + p->sz[i] * p->sz[i]);
  const double imass = 1.0 / p->mp[i];
  const double imass2 = imass * imass;
  const double Jx = 4.0 * p->sx[i] * imass2, Jy = 4.0 * p->sy[i] * imass2,
               Jz = 4.0 * p->sz[i] * imass2;
  const double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
  const double pbl = sqrt(p->px[i] * p->px[i] + p->py[i] * p->py[i] + p->pz[i] * p->pz[i]);
  const double Px = 2.0 * p->px[i] * imass, Py = 2.0 * p->py[i] * imass,
               Pz = 2.0 * p->pz[i] * imass;
  const double P2 = Px * Px + Py * Py + Pz * Pz;
  const double xi = x - p->x[i], yi = y - p->y[i], zi = z - p->z[i];
  double r = sqrt(xi * xi + yi * yi + zi * zi);
  if (fabs(r) <= 0.00000000001) r = 0.0000000001;
  const double Rx = 2.0 * xi * imass, Ry = 2.0 * yi * imass, Rz = 2.0 * zi * imass;
  const double R = 2.0 * r * imass;
  const double l1 = 1.0 / (1.0 + R);
  const double l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1;
  const double u0j = 0.025 * (l1 + l2 + l3 - 4.0 * l4 + 2.0 * l5);
  const double u2j = -0.05 * l5;
  const double muj = (sbl != 0.0 && r != 0.0)
                         ? (p->sx[i] * xi + p->sy[i] * yi + p->sz[i] * zi) / (r * sbl) : 0.0;
  const double PL2j = 0.5 * (3.0 * muj * muj - 1.0);
  const double uj = J2 * (u0j + u2j * R * R * PL2j);
  const double u0p = 0.15625 * (l1 - 2.0 * l2 + 2.0 * l3 - l4 + 0.2 * l5);
  const double u2p = 0.0125 * (15.0 * l1 + 132.0 * l2 + 53.0 * l3 + 96.0 * l4 + 82 * l5 +
                               84.0 * (l5 + log(l1) / R) / R) / R;
  const double mup = (pbl != 0.0 && r != 0.0)
                         ? (p->px[i] * xi + p->py[i] * yi + p->pz[i] * zi) / (r * pbl) : 0.0;
  const double PL2p = 0.5 * (3.0 * mup * mup - 1.0);
  const double up = P2 * (u0p + u2p * PL2p);
  const double c0 = Py * Jz - Pz * Jy, c1 = Pz * Jx - Px * Jz, c2 = Px * Jy - Py * Jx;
  const double uc = 0.0125 * (c0 * Rx + c1 * Ry + c2 * Rz) * (1.0 + 5.0 * R + 10.0 * R * R) * l5;
  return uj + up + uc;
}

void oracle_approx_id(oracle_solver *s) {
  const oracle_params *p = &s->p;
  for (int b = 0; b < s->nbox; b++) {
    obox *B = &s->box[b];
    double *u = B->f[OF_U];
    for (size_t k = 0; k < B->D.n[2]; k++)
      for (size_t j = 0; j < B->D.n[1]; j++)
        for (size_t i = 0; i < B->D.n[0]; i++) {
          const double x = dom_x(&B->D, 0, (long)i), y = dom_x(&B->D, 1, (long)j),
                       z = dom_x(&B->D, 2, (long)k);
          double v = 0.0;
          for (int q = 0; q < p->nbh && p->mp[q] != 0.0; q++) v += approx_term(p, q, x, y, z);
          u[IDX(B, i, j, k)] = v;
        }
  }
}

void oracle_puncture_approx(const oracle_solver *s, double *u_at_p) {
  const oracle_params *p = &s->p;
  int n = 0;
  while (n < p->nbh && p->mp[n] != 0.0) n++;
  for (int j = 0; j < n; j++) u_at_p[j] = 0.0;
  const int top = s->levels - 1;
  for (int b = s->lvl_first[top]; b < s->lvl_first[top] + s->lvl_count[top]; b++) {
    const odom *D = &s->box[b].D;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double pj[3] = {p->x[j], p->y[j], p->z[j]};
        if (!dom_is_in(D, pj)) continue;
        const double x = dom_x(D, 0, (long)dom_get_i(D, 0, pj[0]));
        const double y = dom_x(D, 1, (long)dom_get_i(D, 1, pj[1]));
        const double z = dom_x(D, 2, (long)dom_get_i(D, 2, pj[2]));
        u_at_p[j] += approx_term(p, i, x, y, z);
      }
  }
}

double oracle_interp_point(oracle_solver *s, int ibox, int field, int method, int ord,
                           double x, double y, double z) {
  const obox *B = &s->box[ibox];
  return interp(B, B->f[field], method, ord, x, y, z);
}

static int in_region(const obox *B, size_t i, size_t j, size_t k, int region) {
  if (region == OR_ALL) return 1;
  const size_t ijk[3] = {i, j, k};
  int inside = 1;
  for (int a = 0; a < 3; a++)
    if (ijk[a] < B->D.b[a] || ijk[a] >= B->D.n[a] - B->D.b[a]) inside = 0;
  return region == OR_INSIDE ? inside : !inside;
}

/* ffunction.cc:2436-2475 */
static void transfer_box(obox *T, double *tf, const obox *S, const double *sf, int region,
                         int method, int ord, int all, int to_zero) {
  for (size_t k = 0; k < T->D.n[2]; k++)
    for (size_t j = 0; j < T->D.n[1]; j++)
      for (size_t i = 0; i < T->D.n[0]; i++) {
        if (!in_region(T, i, j, k, region)) continue;
        double X[3] = {dom_x(&T->D, 0, (long)i), dom_x(&T->D, 1, (long)j), dom_x(&T->D, 2, (long)k)};
        double *dst = &tf[IDX(T, i, j, k)];
        if (dom_is_in_nb(&S->D, X))
          *dst = interp(S, sf, method, ord, X[0], X[1], X[2]);
        else if (all && dom_is_in(&S->D, X))
          *dst = interp(S, sf, method, ord, X[0], X[1], X[2]);
        else if (to_zero)
          *dst = 0.0;
      }
}

/* multigrid.cc:122-141 */
void oracle_transfer(oracle_solver *s, int from, int ff, int to, int tfld, int region,
                     int method, int ord, int all, int to_zero) {
  from = from < s->levels ? from : s->levels - 1;
  to = to < s->levels ? to : s->levels - 1;
  for (int ib = 0; ib < s->lvl_count[from]; ib++)
    for (int jb = 0; jb < s->lvl_count[to]; jb++) {
      obox *T = &s->box[s->lvl_first[to] + jb];
      const obox *S = &s->box[s->lvl_first[from] + ib];
      transfer_box(T, T->f[tfld], S, S->f[ff], region, method, ord, all, to_zero);
    }
}

/* ffunction.cc:2808-2849 Set_In(W, Dom, all=false), called as S->Set_In */
static void set_in(obox *B, const odom *fine) {
  double *S = B->f[OF_RHO], *W = B->f[OF_W];
  for (size_t k = 0; k < B->D.n[2]; k++)
    for (size_t j = 0; j < B->D.n[1]; j++)
      for (size_t i = 0; i < B->D.n[0]; i++) {
        double X[3] = {dom_x(&B->D, 0, (long)i), dom_x(&B->D, 1, (long)j), dom_x(&B->D, 2, (long)k)};
        if (dom_is_in_nb(&B->D, X) && dom_is_in_nb(fine, X))
          S[IDX(B, i, j, k)] = linear_interp(B, W, X[0], X[1], X[2]);
      }
}

/* Defect restriction used with multi-colour sweeps: the same target points,
 * precedence and Lagrange-2 nodes as the reference's Transferir(W, ALL,
 * LAGRANGE 2, all) (elliptic.cc:1660-1685), applied to the filtered defect:
 * at every fine node whose 3x3x3 neighbourhood lies in the fine interior the
 * defect is replaced by its full-weighting average (1/4,1/2,1/4)^3, elsewhere
 * kept.  Where the coarse point coincides with a fine node this is full
 * weighting instead of injection; where a merged fine box lies off the
 * coarse lattice, the filter removes the (i+j+k)-parity pattern that the
 * multi-colour sweeps leave in the defect and that the interpolation would
 * alias into the coarse correction (without it the V-cycle of such
 * hierarchies diverges, even for the Laplacian).  A converged FAS solution
 * has zero interior defect, so the fixed point is the reference's. */
static int coincident_index(const odom *D, int a, double x, long *idx) {
  size_t i = dom_get_i(D, a, x);
  for (long c = (long)i - 1; c <= (long)i; c++) {
    if (c < 0 || c >= (long)D->n[a]) continue;
    if (fabs(dom_x(D, a, c) - x) <= 1e-6 * D->d[a]) {
      *idx = c;
      return 1;
    }
  }
  return 0;
}

static void restrict_defect_fw(oracle_solver *s, int k, int all) {
  static const double w[3] = {0.25, 0.5, 0.25};
  for (int ib = 0; ib < s->lvl_count[k]; ib++) {
    const obox *S = &s->box[s->lvl_first[k] + ib];
    const double *sf = S->f[OF_W];
    /* the filtered defect of the fine box */
    double *flt = malloc(S->npts * sizeof(double));
    const long n0 = (long)S->D.n[0], n1 = (long)S->D.n[1], n2 = (long)S->D.n[2];
    const long R0 = (long)S->D.b[0], R1 = (long)S->D.b[1], R2 = (long)S->D.b[2];
    for (long kk = 0; kk < n2; kk++)
      for (long j = 0; j < n1; j++)
        for (long i = 0; i < n0; i++) {
          double val = sf[IDX(S, i, j, kk)];
          if (i - 1 >= R0 && i + 1 < n0 - R0 && j - 1 >= R1 && j + 1 < n1 - R1 && kk - 1 >= R2 &&
              kk + 1 < n2 - R2) {
            val = 0;
            for (int c = -1; c <= 1; c++)
              for (int b = -1; b <= 1; b++)
                for (int a = -1; a <= 1; a++)
                  val += w[a + 1] * w[b + 1] * w[c + 1] * sf[IDX(S, i + a, j + b, kk + c)];
          }
          flt[IDX(S, i, j, kk)] = val;
        }
    for (int jb = 0; jb < s->lvl_count[k - 1]; jb++) {
      obox *T = &s->box[s->lvl_first[k - 1] + jb];
      for (size_t kk = 0; kk < T->D.n[2]; kk++)
        for (size_t j = 0; j < T->D.n[1]; j++)
          for (size_t i = 0; i < T->D.n[0]; i++) {
            double X[3] = {dom_x(&T->D, 0, (long)i), dom_x(&T->D, 1, (long)j), dom_x(&T->D, 2, (long)kk)};
            int take = dom_is_in_nb(&S->D, X) || (all && dom_is_in(&S->D, X));
            if (!take) continue;
            /* within 1e-6 dx of a node whose neighbourhood is interior: the
             * full-weighting value of that node */
            long fi[3];
            int fw = 1;
            for (int a = 0; a < 3 && fw; a++) {
              if (!coincident_index(&S->D, a, X[a], &fi[a])) fw = 0;
              else if (fi[a] - 1 < (long)S->D.b[a] || fi[a] + 1 >= (long)(S->D.n[a] - S->D.b[a])) fw = 0;
            }
            T->f[OF_W][IDX(T, i, j, kk)] = fw ? flt[IDX(S, fi[0], fi[1], fi[2])]
                                              : lagrange_interp(S, flt, 2, X[0], X[1], X[2]);
          }
    }
    free(flt);
  }
}

/* ------------------------------------------------------------------ */
/* FAS cycle: elliptic.cc:1251-1342 (Vcycle), 1469-1968 (FASN)         */
/* ------------------------------------------------------------------ */

typedef struct {
  double *h;
  int cap, n;
  int cycle;
} ohist;

static void hist_add(ohist *H, double v) {
  if (H && H->h && H->n < H->cap) {
    H->h[2 * H->n] = H->cycle;
    H->h[2 * H->n + 1] = v;
  }
  if (H) H->n++;
}

/* Relaxation of the coarsest-level loop: 1 (the reference), a given value,
 * or (negative setting) 2 - 1.6/(n-1) for an n^3 coarsest box, the GPU's
 * choice. */
static double coarse_omega(const oracle_solver *s, int ibox) {
  if (s->p.coarse_omega > 0) return s->p.coarse_omega;
  if (s->p.coarse_omega < 0) return 2.0 - 1.6 / (double)(s->box[ibox].D.n[0] - 1);
  return 1.0;
}

static double fasn(oracle_solver *s, int k, ohist *H) {
  const int M = s->mgrid;
  double norm_k = 0, norm_km1 = 0;
  const int fk = s->lvl_first[k], nk = s->lvl_count[k];
  const int fc = s->lvl_first[k - 1], nc = s->lvl_count[k - 1];
  for (int i = 0; i < nk; i++) {
    if (k <= M) oracle_robin(s, fk + i);
    double nrm = oracle_smooth(s, fk + i, s->p.nu1);
    norm_k = nrm > norm_k ? nrm : norm_k;
  }
  hist_add(H, norm_k);
  /* restrict u (injection at coincident points) */
  oracle_transfer(s, k, OF_U, k - 1, OF_U, OR_INSIDE, OI_LAGRANGE, 3, 0, 0);
  for (int j = 0; j < nc; j++) {
    oracle_residual(s, fc + j);
    if (k <= M) oracle_boundary_operator(s, fc + j);
  }
  for (int i = 0; i < nk; i++) {
    oracle_residual(s, fk + i);
    if (k <= M) oracle_boundary_operator(s, fk + i);
  }
  for (int i = 0; i < nk; i++) {
    obox *B = &s->box[fk + i];
    for (size_t q = 0; q < B->npts; q++) B->f[OF_W][q] = B->f[OF_RHO][q] - B->f[OF_R][q];
  }
  if (s->p.restrict_fw)
    restrict_defect_fw(s, k, k - 1 <= M);
  else if (k - 1 <= M)
    oracle_transfer(s, k, OF_W, k - 1, OF_W, OR_ALL, OI_LAGRANGE, 2, 1, 0);
  else
    oracle_transfer(s, k, OF_W, k - 1, OF_W, OR_ALL, OI_LAGRANGE, 2, 0, 0);
  for (int j = 0; j < nc; j++) {
    obox *B = &s->box[fc + j];
    if (k <= M) {
      for (size_t q = 0; q < B->npts; q++) B->f[OF_RHO][q] = B->f[OF_W][q] + B->f[OF_R][q];
    } else {
      for (size_t q = 0; q < B->npts; q++) B->f[OF_W][q] += B->f[OF_R][q];
      for (int i = 0; i < nk; i++) set_in(B, &s->box[fk + i].D);
    }
  }
  for (int j = 0; j < nc; j++) {
    obox *B = &s->box[fc + j];
    memcpy(B->f[OF_WW], B->f[OF_U], B->npts * sizeof(double));
  }
  if (k == 1) {
    for (int j = 0; j < nc; j++) {
      double norm_prev;
      norm_km1 = 1;
      do {
        norm_prev = norm_km1;
        oracle_robin(s, fc + j);
        norm_km1 = smooth_omega(s, fc + j, 1, coarse_omega(s, fc + j));
        s->coarse_iters++;
        if (s->failed) return NAN;
      } while (norm_km1 > s->eps_coarse && norm_prev != norm_km1);
    }
    hist_add(H, norm_km1);
  } else {
    norm_km1 = fasn(s, k - 1, H);
  }
  if (s->failed) return NAN;
  for (int j = 0; j < nc; j++) {
    obox *B = &s->box[fc + j];
    for (size_t q = 0; q < B->npts; q++) B->f[OF_V][q] = B->f[OF_U][q] - B->f[OF_WW][q];
  }
  if (k <= M)
    oracle_transfer(s, k - 1, OF_V, k, OF_V, OR_ALL, OI_LINEAR, 3, 1, 1);
  else
    oracle_transfer(s, k - 1, OF_V, k, OF_V, OR_INSIDE, OI_LINEAR, s->p.order + 1, 1, 0);
  for (int i = 0; i < nk; i++) {
    obox *B = &s->box[fk + i];
    for (size_t q = 0; q < B->npts; q++) B->f[OF_U][q] += B->f[OF_V][q];
  }
  if (k > M)
    oracle_transfer(s, k - 1, OF_U, k, OF_U, OR_BOUNDARY, OI_LAGRANGE, s->p.order + 1, 1, 0);
  norm_k = 0;
  for (int i = 0; i < nk; i++) {
    if (k <= M) oracle_robin(s, fk + i);
    double nrm = oracle_smooth(s, fk + i, s->p.nu2);
    norm_k = nrm > norm_k ? nrm : norm_k;
  }
  (void)norm_km1;
  return norm_k;
}

double oracle_fasn(oracle_solver *s, int k) {
  if (k < 1 || k >= s->levels) return NAN;
  return fasn(s, k, NULL);
}

int oracle_solve(oracle_solver *s, double *hist, int hist_cap, int *n_hist, double *final_res) {
  ohist H = {hist, hist_cap, 0, 0};
  double res = 1e12, res_prev;
  int cycle = 0;
  do {
    res_prev = res;
    H.cycle = cycle;
    res = fasn(s, s->levels - 1, &H);
    cycle++;
    if (s->failed) break;
    if (s->p.max_cycles > 0 && cycle >= s->p.max_cycles) break;
  } while (res > s->eps_solv && res < res_prev);
  if (n_hist) *n_hist = H.n;
  if (final_res) *final_res = res;
  return s->failed ? -1 : cycle;
}

/* ------------------------------------------------------------------ */
/* ADM mass: elliptic.cc:2447-2611                                     */
/* ------------------------------------------------------------------ */

static double dU_at(const obox *B, long i, long j, long k) {
  const size_t ijk[3] = {(size_t)i, (size_t)j, (size_t)k};
  double X[3];
  for (int a = 0; a < 3; a++) X[a] = dom_x(&B->D, a, (long)ijk[a]) - B->D.c[a];
  double ir = 1 / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  double g[3];
  for (int a = 0; a < 3; a++)
    g[a] = deriv(B, B->f[OF_C], a, ijk, NULL) + deriv(B, B->f[OF_U], a, ijk, NULL);
  return ir * (X[0] * g[0] + X[1] * g[1] + X[2] * g[2]);
}

/* interpolation.cc:125-182 applied to the pointwise dU field */
static double dU_interp(const obox *B, double x, double y, double z) {
  size_t i = dom_get_i(&B->D, 0, x), j = dom_get_i(&B->D, 1, y), k = dom_get_i(&B->D, 2, z);
  if (dcomp(x, dom_x(&B->D, 0, (long)i)) && dcomp(y, dom_x(&B->D, 1, (long)j)) &&
      dcomp(z, dom_x(&B->D, 2, (long)k)))
    return dU_at(B, (long)i, (long)j, (long)k);
  i = i == 0 ? 1 : i;
  j = j == 0 ? 1 : j;
  k = k == 0 ? 1 : k;
  double dxf = (x - dom_x(&B->D, 0, (long)i - 1)) * B->idx[0];
  double dyf = (y - dom_x(&B->D, 1, (long)j - 1)) * B->idx[1];
  double dzf = (z - dom_x(&B->D, 2, (long)k - 1)) * B->idx[2];
  double dxi = (1 - dxf), dyi = (1 - dyf), dzi = (1 - dzf);
  long I = (long)i, J = (long)j, K = (long)k;
  return ((dU_at(B, I - 1, J - 1, K - 1) * dxi + dU_at(B, I, J - 1, K - 1) * dxf) * dyi +
          (dU_at(B, I - 1, J, K - 1) * dxi + dU_at(B, I, J, K - 1) * dxf) * dyf) * dzi +
         ((dU_at(B, I - 1, J - 1, K) * dxi + dU_at(B, I, J - 1, K) * dxf) * dyi +
          (dU_at(B, I - 1, J, K) * dxi + dU_at(B, I, J, K) * dxf) * dyf) * dzf;
}

void oracle_adm_mass(oracle_solver *s, double *puncture, double *surface) {
  const oracle_params *p = &s->p;
  const int top = s->levels - 1;
  for (int i = 0; i < p->nbh; i++) {
    /* multigrid.cc:97-116 Get_dom_index: a box index beyond the level's
     * count falls back to the level's first box */
    int ib = s->lvl_first[top] + (i < s->lvl_count[top] ? i : 0);
    const obox *B = &s->box[ib];
    double bhmj = 0;
    for (int j = 0; j < p->nbh; j++)
      if (j != i)
        bhmj += 0.5 * p->mp[j] /
                sqrt((p->x[i] - p->x[j]) * (p->x[i] - p->x[j]) + (p->y[i] - p->y[j]) * (p->y[i] - p->y[j]) +
                     (p->z[i] - p->z[j]) * (p->z[i] - p->z[j]));
    double up = lagrange_interp(B, B->f[OF_U], 4, p->x[i], p->y[i], p->z[i]);
    puncture[i] = p->mp[i] * (1 + bhmj + up);
  }
  if (!surface) return;
  for (int ib = 0; ib < s->nbox; ib++) {
    const obox *B = &s->box[ib];
    const double cx = B->D.c[0], cy = B->D.c[1], cz = B->D.c[2];
    double lmin = B->D.len[1] < B->D.len[2] ? B->D.len[1] : B->D.len[2];
    lmin = B->D.len[0] < lmin ? B->D.len[0] : lmin;
    const double r0 = 0.375 * lmin;
    enum { NT = 4 * 25 + 1 };
    const int N_theta = NT, N_phi = NT;
    const double dtheta = M_PI / (N_theta - 1);
    const double dphi = 2 * M_PI / (N_phi - 1);
    static double Sph[NT][NT];
    for (int it = 0; it < N_theta; it++) {
      double theta = it * dtheta;
      for (int ip = 0; ip < N_phi; ip++) {
        double phi = ip * dphi;
        double sxp = cx + r0 * cos(phi) * sin(theta);
        double syp = cy + r0 * sin(phi) * sin(theta);
        double szp = cz + r0 * cos(theta);
        Sph[it][ip] = dU_interp(B, sxp, syp, szp) * sin(theta);
      }
    }
    double ADM = 0;
    for (int i = 0; i < N_theta - 3; i += 4) {
      double I0 = 0, I1 = 0, I2 = 0, I3 = 0, I4 = 0;
      for (int j = 0; j < N_phi - 3; j += 4) {
        I0 += 7 * Sph[i][j] + 32 * Sph[i][j + 1] + 12 * Sph[i][j + 2] + 32 * Sph[i][j + 3] + 7 * Sph[i][j + 4];
        I1 += 7 * Sph[i + 1][j] + 32 * Sph[i + 1][j + 1] + 12 * Sph[i + 1][j + 2] + 32 * Sph[i + 1][j + 3] + 7 * Sph[i + 1][j + 4];
        I2 += 7 * Sph[i + 2][j] + 32 * Sph[i + 2][j + 1] + 12 * Sph[i + 2][j + 2] + 32 * Sph[i + 2][j + 3] + 7 * Sph[i + 2][j + 4];
        I3 += 7 * Sph[i + 3][j] + 32 * Sph[i + 3][j + 1] + 12 * Sph[i + 3][j + 2] + 32 * Sph[i + 3][j + 3] + 7 * Sph[i + 3][j + 4];
        I4 += 7 * Sph[i + 4][j] + 32 * Sph[i + 4][j + 1] + 12 * Sph[i + 4][j + 2] + 32 * Sph[i + 4][j + 3] + 7 * Sph[i + 4][j + 4];
      }
      I0 *= 2 * dphi / 45;
      I1 *= 2 * dphi / 45;
      I2 *= 2 * dphi / 45;
      I3 *= 2 * dphi / 45;
      I4 *= 2 * dphi / 45;
      ADM += 7 * I0 + 32 * I1 + 12 * I2 + 32 * I3 + 7 * I4;
    }
    ADM *= -r0 * r0 * dtheta / (45 * M_PI);
    surface[ib] = ADM;
  }
}

/* ------------------------------------------------------------------ */
/* Construction: elliptic.cc:31-221 (Make), operator.cc:629-868        */
/* (Set_Coefficients), boundary.cc:8-84 (Set_Boundary)                 */
/* ------------------------------------------------------------------ */

oracle_solver *oracle_create(const oracle_params *p) {
  g_err[0] = 0;
  if (!(p->order == 2 || p->order == 4 || p->order == 6 || p->order == 8)) {
    set_err("order must be [2,4,6,8].");
    return NULL;
  }
  if (p->levels < 1) {
    set_err("levels must be >= 1");
    return NULL;
  }
  if (p->nbh < 1 || p->nbh > ORACLE_MAX_BH) {
    set_err("NumBH out of range");
    return NULL;
  }
  oracle_solver *s = calloc(1, sizeof *s);
  s->p = *p;
  s->eps_coarse = 2 * EPS;
  s->eps_solv = p->tol < s->eps_coarse ? s->eps_coarse : p->tol;
  s->ps_r0 = p->ps_r0;
  s->ps_isigma = 1 / p->ps_sigma;
  s->ps_phi0 = (p->ps_a2 * p->ps_a2 / (1 + p->ps_a2 * p->ps_a2)) * p->ps_phi0;
  s->ps_d1 = M_PI * s->ps_phi0 * s->ps_phi0 * s->ps_isigma * s->ps_isigma;
  s->ps_e1 = 1 / (16 * p->ps_a2);
  s->ps_e2 = 4 * sqrt(M_PI / 3);
  s->ps_de1 = 4 * M_PI * s->ps_phi0 * s->ps_phi0 * s->ps_isigma * s->ps_isigma;
  if (build_grids(s) != 0) {
    free(s);
    return NULL;
  }
  for (int ib = 0; ib < s->nbox; ib++) {
    obox *B = &s->box[ib];
    B->npts = B->D.n[0] * B->D.n[1] * B->D.n[2];
    for (int f = 0; f < OF_NFIELDS; f++)
      B->f[f] = (f < OF_NBASE || p->eq == ORACLE_PUNCTURES_EM) ? calloc(B->npts, sizeof(double)) : NULL;
    for (int a = 0; a < 3; a++) {
      B->idx[a] = 1.0 / B->D.d[a];
      B->idx2[a] = B->idx[a] * B->idx[a];
    }
    for (size_t k = 0; k < B->D.n[2]; k++)
      for (size_t j = 0; j < B->D.n[1]; j++)
        for (size_t i = 0; i < B->D.n[0]; i++) {
          double x = dom_x(&B->D, 0, (long)i), y = dom_x(&B->D, 1, (long)j), z = dom_x(&B->D, 2, (long)k);
          size_t q = IDX(B, i, j, k);
          const int shell = i < B->D.b[0] || i >= B->D.n[0] - B->D.b[0] || j < B->D.b[1] ||
                            j >= B->D.n[1] - B->D.b[1] || k < B->D.b[2] || k >= B->D.n[2] - B->D.b[2];
          if (p->eq == ORACLE_POISSON) {
            /* operator.cc:640-662: rho0 = Gaussian_source, u = Gaussian,
             * Robin value 0 on the shell (boundary.cc:8-84) */
            const double r2 = x * x + y * y + z * z, a = 0.004;
            B->f[OF_RHO][q] = shell ? 0.0 : a * exp(-r2) * (a + exp(0.5 * r2) * (r2 - 3.0));
            B->f[OF_U][q] = a * exp(-0.5 * r2);
            continue;
          }
          if (p->eq == ORACLE_PUNCTURES_EM) {
            /* operator.cc:1010-1100: the thirteen coefficients, rho0 = 0,
             * u = 1 (elliptic::one), Robin value 1 on the shell */
            double cf[13];
            em_coefficients(s, x, y, z, cf);
            for (int c = 0; c < 13; c++) B->f[OF_A1 + c][q] = cf[c];
            B->f[OF_RHO][q] = shell ? 1.0 : 0.0;
            B->f[OF_U][q] = 1.0;
            continue;
          }
          if (p->eq == ORACLE_BRILL) {
            /* operator.cc:707-725 and brill (:1688-1698): Lq, rho0 = 0,
             * u = 1, Robin value 1 on the shell */
            const double Amp = 1.9, z2 = z * z, rho2 = x * x + y * y;
            const double qrr = 2.0 * Amp * exp(-rho2 - z2) * (1.0 - 5.0 * rho2 + 2.0 * rho2 * rho2);
            const double qzz = 2.0 * Amp * exp(-rho2 - z2) * rho2 * (2.0 * z2 - 1.0);
            B->f[OF_B][q] = 0.25 * (qrr + qzz);
            B->f[OF_RHO][q] = shell ? 1.0 : 0.0;
            B->f[OF_U][q] = 1.0;
            continue;
          }
          B->f[OF_B][q] = punctures_b(s, x, y, z);
          B->f[OF_C][q] = punctures_c(s, x, y, z);
          if (p->eq == ORACLE_PUNCTURES_SCALAR) {
            B->f[OF_D][q] = punctures_scalar_d(s, x, y, z);
            B->f[OF_E][q] = punctures_scalar_e(s, x, y, z);
          }
          /* rho0 = 0 (operator.cc:766-771 and the Robin shell value
           * boundary_coeff = 0, boundary.cc:8-40); u = 0 (boundary.cc:58-80) */
        }
  }
  return s;
}

void oracle_destroy(oracle_solver *s) {
  if (!s) return;
  for (int ib = 0; ib < s->nbox; ib++)
    for (int f = 0; f < OF_NFIELDS; f++) free(s->box[ib].f[f]);
  free(s->box);
  free(s->lvl_first);
  free(s->lvl_count);
  free(s);
}

int oracle_num_levels(const oracle_solver *s) { return s->levels; }
long oracle_coarse_iterations(const oracle_solver *s) { return s->coarse_iters; }
int oracle_mgrid_levels(const oracle_solver *s) { return s->mgrid; }
int oracle_num_boxes(const oracle_solver *s) { return s->nbox; }
int oracle_level_first_box(const oracle_solver *s, int level) { return s->lvl_first[level]; }
int oracle_level_num_boxes(const oracle_solver *s, int level) { return s->lvl_count[level]; }

void oracle_box_info(const oracle_solver *s, int ibox, int *level, int *order, int *R,
                     int n[3], double origin[3], double d[3], double ext[6]) {
  const odom *D = &s->box[ibox].D;
  if (level) *level = D->level;
  if (order) *order = (int)D->order;
  if (R) *R = (int)D->b[0];
  for (int a = 0; a < 3; a++) {
    if (n) n[a] = (int)D->n[a];
    if (origin) origin[a] = D->lo[a];
    if (d) d[a] = D->d[a];
    if (ext) {
      ext[2 * a] = D->lo[a];
      ext[2 * a + 1] = D->hi[a];
    }
  }
}

double *oracle_field(oracle_solver *s, int ibox, int field) {
  if (ibox < 0 || ibox >= s->nbox || field < 0 || field >= OF_NFIELDS) return NULL;
  return s->box[ibox].f[field];
}
