/*
 * olliptic_amd: MI355X-native FAS multigrid solve of the puncture
 * Hamiltonian constraint  Lap u + b (c + u)^-7 [+ d (c+u) + e (c+u)^5] = 0
 * on nested fixed refinement boxes with 2nd-8th order finite differences and
 * Robin outer boundaries.
 *
 * C ABI: plain pointers and sizes only.  One handle (oll_hier) owns one box
 * hierarchy resident in HBM and one HIP stream; calls on one handle are not
 * reentrant, distinct handles may be used concurrently from different host
 * threads.  Every function returns 0 on success and a negative code on
 * failure; oll_last_error() returns the calling thread's message.
 *
 * Each entry point names the reference interface it replaces
 * (files of WuShichao/OllipticID).
 */
#ifndef OLLIPTIC_AMD_H
#define OLLIPTIC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OLL_MAX_BH 16

/* error codes */
#define OLL_OK 0
#define OLL_EINVAL (-1)     /* bad argument / configuration */
#define OLL_EGRID (-2)      /* not enough grid points for multigrid */
#define OLL_EDEVICE (-3)    /* HIP runtime / kernel failure */
#define OLL_ENAN (-4)       /* NaN produced by the smoother */
#define OLL_ENOMEM (-5)     /* device allocation failed */
#define OLL_EBUSY (-6)      /* engine: admitting would exceed the HBM budget */

/* equations (interface.cc:1184-1323 names "Punctures", "Punctures_scalar_field",
 * "Poisson", "Brill-waves", "Punctures_EM"), all N(u) = Lap u + terms of u
 * (and, for PUNCTURES_EM, of grad u) on the same kernels:
 *   PUNCTURES         b (c+u)^-7                     (elliptic.h:405-421)
 *   PUNCTURES_SCALAR  + d (c+u) + e (c+u)^5          (elliptic.h:429-445)
 *   POISSON           u^2, rho = Gaussian source, u0 = Gaussian, Robin value 0
 *                     (elliptic.h:357-370, operator.cc:640-662, 1865-1892)
 *   BRILL             Lq u with Lq the Brill-wave seed, rho = 0, u0 = 1, Robin
 *                     value 1 (elliptic.h:381-398, operator.cc:707-725, 1688-1698)
 *   PUNCTURES_EM      [B |grad u|^2 + (B1 u + B2).grad u + C u^2 + D u + E]
 *                     / (u^3 + A1 u^2 + A2 u + A3), coefficients from the mass
 *                     and charge potentials of the punctures (oll_puncture.q),
 *                     rho = 0, u0 = 1, Robin value 1 (elliptic.h:510-626,
 *                     1458-1530, operator.cc:1010-1100, 3043-3310); FP64
 *                     only.  Its first derivatives are those the reference
 *                     evaluates: centred at order 2, one-sided five-point
 *                     (four points to one side) at orders 4-8, where its
 *                     boxes are swept in 5 colours instead of order/2 + 1
 * The Robin value of the shell is oll_problem.robin_a (set it to 1 for BRILL
 * and PUNCTURES_EM, as the reference's boundary_coeff is). */
#define OLL_EQ_PUNCTURES 0
#define OLL_EQ_PUNCTURES_SCALAR 1
#define OLL_EQ_POISSON 2
#define OLL_EQ_BRILL 3
#define OLL_EQ_PUNCTURES_EM 4

/* arithmetic of the smoothing blocks.  FP64: the reference's Newton-Gauss-
 * Seidel sweeps in float64.  MIXED: each block of nu sweeps computes the
 * float64 defect rho - N(u) and dN/du once, relaxes the linearised
 * correction equation nu times in float32, and adds the correction to u in
 * float64; restriction, prolongation, Robin, the coarsest-level solve and
 * the coefficients stay float64.  Both converge to the same discrete
 * solution (the float64 defect vanishes there). */
#define OLL_PREC_FP64 0
#define OLL_PREC_MIXED 1

/* stopping rule of the V-cycle loop.  REFERENCE (elliptic.cc:1330): go on
 * while res > tol and res < the previous cycle's res, i.e. stop at the first
 * cycle that does not lower the residual.  TOLERANCE: go on while res > tol,
 * stopping only at max_cycles or after `patience` cycles without a new
 * lowest residual (for streams whose solves must reach tol). */
#define OLL_STOP_REFERENCE 0
#define OLL_STOP_TOLERANCE 1

/* solve method (-m).  VCYCLE: FAS V-cycles (elliptic.cc:1251-1342).  APP
 * ("AppPuncture", elliptic.cc:1977-2222): u is the analytic approximation of
 * each puncture's spin, momentum and cross terms, with no iteration; its
 * grids have no multigrid coarsening (elliptic.cc:510, 678); Punctures only
 * (interface.cc:1154). */
#define OLL_METHOD_VCYCLE 0
#define OLL_METHOD_APP 1

/* initial u of the V-cycles.  ZERO: the reference's u = 0 (boundary.cc:8-84).
 * APPROX: the analytic approximation of OLL_METHOD_APP on every point of
 * every box (no reference counterpart; the iterates start closer, the
 * converged solution is the same). */
#define OLL_GUESS_ZERO 0
#define OLL_GUESS_APPROX 1

/* field ids (reference field names, operator.cc:103-124,509-567) */
#define OLL_F_U 0
#define OLL_F_B 1
#define OLL_F_C 2
#define OLL_F_D 3
#define OLL_F_E 4
#define OLL_F_RHO 5
#define OLL_F_R 6
#define OLL_F_V 7
#define OLL_F_W 8
#define OLL_F_WW 9
/* PUNCTURES_EM only: with eta = sum m / (2r) in OLL_F_B and phi = sum q / r in
 * OLL_F_D, their gradients (operator.cc:3043-3310) */
#define OLL_F_ETA_X 10
#define OLL_F_ETA_Y 11
#define OLL_F_ETA_Z 12
#define OLL_F_PHI_X 13
#define OLL_F_PHI_Y 14
#define OLL_F_PHI_Z 15
#define OLL_NFIELDS 10

/* One puncture: bare mass, position, linear momentum, spin, charge
 * (reference -bhmp<i> -bhx<i> ... -bhsz<i> -bhqp<i>, interface.cc:748-864;
 * the charge enters PUNCTURES_EM only). */
typedef struct oll_puncture {
  double m, x, y, z, px, py, pz, sx, sy, sz;
  double q;
} oll_puncture;

/* Problem description: the reference's hot-path CLI parameters
 * (interface.cc:35-355 defaults, elliptic.cc:31-221 Make). */
typedef struct oll_problem {
  int eq;               /* OLL_EQ_* (-eq) */
  int num_bh;           /* -NumBH */
  oll_puncture bh[OLL_MAX_BH];
  double L;             /* -L */
  int N;                /* -N, snapped to the reference's valid sizes */
  int levels;           /* -l refinement levels */
  int order;            /* -O 2,4,6,8 */
  int nu1, nu2;         /* -nu1 -nu2 */
  double tol;           /* -s */
  double robin_a;       /* -BA Robin value at infinity (reference: 0) */
  int robin_n;          /* -Bn fall-off power (reference: 1) */
  int ps_case;          /* -psc */
  double ps_r0, ps_sigma, ps_phi0, ps_a2;  /* -r0 -sigma -phi0 -a2 */
  int max_cycles;       /* > 0: also stop after this many V-cycles */
  int precision;        /* OLL_PREC_* */
  int stop_rule;        /* OLL_STOP_* */
  int patience;         /* OLL_STOP_TOLERANCE: stop after this many cycles without a new best */
  int method;           /* OLL_METHOD_* (-m) */
  int initial_guess;    /* OLL_GUESS_* */
} oll_problem;

typedef struct oll_hier oll_hier;

typedef struct oll_box_desc {
  int level;            /* 0 = coarsest multigrid level */
  int order;            /* stencil order of this box (2 on coarsened levels) */
  int ghost;            /* ghost width R = max(1, order/2) */
  int n[3];             /* points per axis incl. ghosts */
  double origin[3];     /* coordinate of index 0 per axis */
  double d[3];          /* spacing */
  double centre[3];
  double end[3];        /* coordinate of the last point as the reference's domain keeps
                           it (domain.cc:326-351); may differ from origin + (n-1) d in
                           the last bit */
  int colour_order;     /* two-colour boxes: 1 when the multi-colour sweeps take
                           colour 1 first (the box's points that coincide with nodes of
                           its parent, the coarser box holding its centre, have odd
                           i + j + k), else 0 (also off the parent's lattice) */
} oll_box_desc;

typedef struct oll_result {
  int cycles;           /* V-cycles run */
  double residual;      /* last top-level post-smoothing norm */
  int hist_len;         /* entries written to history (cycle, norm) pairs */
  int converged;        /* 1 when the solve ended with residual <= tol (the
                           reference's stopping rule may end it above tol, at
                           the first cycle that does not lower the residual) */
} oll_result;

/* Defaults of the reference CLI (interface.cc:35-355). */
void oll_problem_default(oll_problem *p);

/* Builds the box hierarchy (elliptic.cc:300-921 Set_Grids), allocates it in
 * HBM and evaluates the coefficients b, c [,d, e] on the device
 * (operator.cc:629-868 Set_Coefficients, boundary.cc:8-84 Set_Boundary). */
int oll_hier_create(const oll_problem *p, oll_hier **out);
void oll_hier_destroy(oll_hier *h);
const char *oll_last_error(void);

/* Hierarchy shape (multigrid.h:49-551 Get_num_dom / Get_dom_index). */
int oll_num_levels(const oll_hier *h);
int oll_mgrid_levels(const oll_hier *h);
int oll_num_boxes(const oll_hier *h);
int oll_level_boxes(const oll_hier *h, int level, int *first, int *count);
int oll_box_info(const oll_hier *h, int ibox, oll_box_desc *out);
/* Bytes of HBM the hierarchy holds. */
size_t oll_device_bytes(const oll_hier *h);
/* Box table of the hierarchy a problem would build, host only (no device
 * needed): elliptic.cc:300-921 Set_Grids.  Writes min(cap, *nbox) boxes. */
int oll_problem_boxes(const oll_problem *p, oll_box_desc *out, int cap, int *nbox,
                      int *mgrid);
/* Footprint oll_hier_create would allocate, computed on the host only. */
int oll_problem_footprint(const oll_problem *p, size_t *bytes);

/* Field access, lexicographic (i fastest) float64 (ffunction.h Get_val/Set). */
int oll_get_field(oll_hier *h, int ibox, int field, double *host);
int oll_set_field(oll_hier *h, int ibox, int field, const double *host);

/* multigrid::Transferir (multigrid.cc:122-141, ffunction.cc:2436-2475):
 * field from_field of the boxes of level from_level interpolated into
 * field to_field of the boxes of level to_level, at the points of `region`
 * (OLL_REGION_*: all, interior, ghost shell); OLL_INTERP_LINEAR (trilinear)
 * or OLL_INTERP_LAGRANGE of `order` (2..9, order + 1 nodes per axis; a point
 * on a source node takes its value).  A point takes the last source box
 * whose non-ghost part holds it (with `all`, whose extent holds it); with
 * `to_zero` the last source box alone decides and points it does not hold
 * are set to 0, as the reference's loop over source boxes leaves them.
 * Levels beyond the last are clamped to it; a field onto itself on one
 * level is refused (OLL_EINVAL).  Transfer tables are built for
 * the call (device memory is allocated and freed). */
enum { OLL_REGION_ALL = 0, OLL_REGION_INSIDE = 1, OLL_REGION_BOUNDARY = 2 };
enum { OLL_INTERP_LINEAR = 0, OLL_INTERP_LAGRANGE = 1 };
int oll_transfer(oll_hier *h, int from_level, int from_field, int to_level, int to_field,
                 int region, int method, int order, int all, int to_zero);

/* Single stages, all boxes of one level (for tests and tools). */
/* elliptic.cc:1126-1242 Smooth (multi-colour order); linf = max |N(u)-rho|. */
int oll_smooth(oll_hier *h, int level, int nu, double *linf);
/* elliptic.cc:944-991 App_Ell_Operator: r = N(u) on the interior. */
int oll_residual(oll_hier *h, int level, double *linf);
/* boundary.cc:447-800 App_Bound (Robin update of the ghost shell). */
int oll_robin(oll_hier *h, int level);
/* boundary.cc:447-800 App_Bound(l, ibox): the Robin update of one box's
 * ghost shell (elliptic.h:301). */
int oll_robin_box(oll_hier *h, int ibox);
/* elliptic.h:337-345 OPERATOR / duOPERATOR over the interior of every box of
 * one level: derivative 0 writes N(u) into OLL_F_R (as oll_residual),
 * derivative 1 writes dN/du at the point itself (the Newton diagonal of
 * elliptic.cc:1126-1242).  linf = max |N(u)| either way. */
int oll_operator(oll_hier *h, int level, int derivative, double *linf);
/* boundary.cc:1118-1535 App_Boundary_Operator: r = Robin operator on shell. */
int oll_boundary_operator(oll_hier *h, int level);
/* elliptic.cc:1469-1968 FASN(k): one FAS step of level k >= 1 (recursive). */
int oll_fas(oll_hier *h, int level, double *linf);
/* Per-level norms of the last oll_fas / last V-cycle of oll_solve: pre[k]
 * the pre-smoothing norm of level k (the coarsest-loop norm for k = 0),
 * post[k] the post-smoothing norm (output.cc:1355-1594 Print_Pre/Post_Cycle). */
int oll_last_norms(const oll_hier *h, double *pre, double *post, int nlevels);
/* elliptic.cc:1251-1342 Vcycle loop; hist (2*cap doubles) may be NULL.
 * On an OLL_METHOD_APP hierarchy: elliptic.cc:1977-2222 ApproximateID (u set
 * to the analytic approximation; 0 cycles, residual 0). */
int oll_solve(oll_hier *h, oll_result *res, double *hist, int hist_cap);
/* elliptic.cc:2232-2444 Get_Puncture_App: per puncture j (up to the first
 * with zero mass), the approximation summed over the punctures at the grid
 * node Get_i/j/k of j's position, over every finest-level box holding it
 * (0.1 dx tolerance); u_at_p has num_bh entries, the rest left 0.  Host
 * arithmetic; the printed values of -m AppPuncture and -fapp. */
int oll_puncture_approx(const oll_hier *h, double *u_at_p);
/* elliptic.cc:2447-2611 Compute_ADM_Mass: per puncture and per box. */
int oll_adm_mass(oll_hier *h, double *puncture, double *surface);
/* The ADM masses of oll_adm_mass without allocating device memory (no
 * reference counterpart: for callers that cannot allocate, such as code
 * under stream capture or serving loops that forbid hipMalloc).  The
 * device staging lives in a fixed 5.2 MB per-GPU scratch array that holds
 * the surface samples of at most 64 boxes: a hierarchy with more boxes
 * (and surface != NULL) is refused with OLL_EINVAL.  Every call on one GPU
 * reuses that array, so calls that overlap on one GPU race; results equal
 * oll_adm_mass's bit for bit. */
int oll_adm_mass_noalloc(oll_hier *h, double *puncture, double *surface);
/* ffunction::Interpol_In with LAGRANGE of order ord (interpolation.cc:189-303,
 * set by ffunction::Set_Interpol, interpolation.cc:32-60): field at npts
 * points, point q = (xyz[3q], xyz[3q+1], xyz[3q+2]) taken from box box[q]
 * (ord + 1 stencil points per axis, shifted to stay inside the box; a point
 * within 3 DBL_EPSILON of a grid point on every axis returns that grid
 * value).  Host arrays; ord 1..9.
 * The multigrid writers (multigrid.cc:151-958) choose the box per point. */
int oll_interp(oll_hier *h, int field, int npts, const int *box, const double *xyz, int ord,
               double *out);

/* Device time of the last oll_solve (HIP events on the handle's stream). */
int oll_last_solve_ms(const oll_hier *h, double *ms);
/* Sweeps spent in the coarsest-level loop during the last oll_solve. */
int oll_last_solve_coarse_sweeps(const oll_hier *h, long long *sweeps);
/* Restores the initial state u = 0, rho = Robin value on the shell
 * (boundary.cc:8-84 Set_Boundary) so the same hierarchy can be solved again. */
int oll_reset(oll_hier *h);
/* Smoother launch timing: when on, every smoother launch is bracketed by HIP
 * events on the handle's stream; oll_kernel_timing returns the number of
 * launches, their summed device time and their algorithmic bytes (40 B per
 * updated point: u, b, c, rho read and u written; 56 B with d, e; 16 B for
 * the FP32 correction sweeps of OLL_PREC_MIXED: e, q, r read, e written). */
int oll_set_kernel_timing(oll_hier *h, int on);
int oll_kernel_timing(oll_hier *h, int *launches, double *total_ms, double *bytes);
/* V-cycles of oll_solve are captured once as a HIP graph and replayed
 * (default on); off issues every launch from the host. */
int oll_set_graphs(oll_hier *h, int on);

/* ------------------------------------------------------------------
 * Engine: many resident requests solved together (no reference
 * counterpart: the reference solves one problem per process).  Requests of
 * any N, level count, order and equation are admitted while their HBM
 * footprint fits the budget; oll_engine_step runs one V-cycle of every
 * active request, one launch per stage covering all of them, and moves
 * requests that met the stopping rule to the done queue.
 * ------------------------------------------------------------------ */
typedef struct oll_engine oll_engine;
/* hbm_budget 0: no limit. */
int oll_engine_create(size_t hbm_budget, int use_graphs, oll_engine **out);
void oll_engine_destroy(oll_engine *e);
/* OLL_EBUSY when the request does not fit the remaining budget. */
int oll_engine_admit(oll_engine *e, const oll_problem *p, long long tag);
/* Moves a built hierarchy (oll_hier_create, or one handed out by
 * oll_engine_pop_hier) into the engine as a new request, reset to the
 * initial guess u = 0: its HBM stays resident, nothing is rebuilt.  h keeps
 * no device memory afterwards but must still be destroyed by the caller. */
int oll_engine_admit_hier(oll_engine *e, oll_hier *h, long long tag);
int oll_engine_step(oll_engine *e, int *active_after);
/* Pops one finished request: returns 1 and fills tag, result and the
 * puncture ADM masses (num_bh values), 0 when none is finished, <0 on error
 * (the request's own failure, e.g. OLL_ENAN).  Frees its HBM. */
int oll_engine_pop_done(oll_engine *e, long long *tag, oll_result *res, double *adm);
/* Like oll_engine_pop_done, but hands the solved hierarchy to the caller
 * (oll_get_field, oll_adm_mass, ...) instead of freeing it; *out is then
 * destroyed with oll_hier_destroy.  Its HBM leaves the engine's budget. */
int oll_engine_pop_hier(oll_engine *e, long long *tag, oll_result *res, oll_hier **out);
int oll_engine_stats(const oll_engine *e, size_t *bytes_in_use, int *active, int *done_pending);
/* host wall time since creation (ms) spent building V-cycle schedules,
 * running cycles (launches to norms back on the host) and admitting
 * requests; schedule rebuilds; engine steps */
int oll_engine_profile(const oll_engine *e, double *build_ms, double *run_ms, double *admit_ms,
                       long long *rebuilds, long long *cycles);
/* smoother launch timing over the engine's V-cycles, as oll_set_kernel_timing
 * (while on, V-cycles are issued launch by launch instead of graph replays);
 * bytes are algorithmic: 40 B (FP64) or 16 B (FP32: e, q, r read, e written)
 * per updated point */
int oll_engine_set_kernel_timing(oll_engine *e, int on);
int oll_engine_kernel_timing(oll_engine *e, int *launches, double *total_ms, double *bytes);
/* The timed smoother launches one by one, in issue order: algorithmic bytes,
 * device ms and kind (2 * stencil order + 0 for an FP64 sweep colour, + 1
 * for an FP32 correction colour) of the first cap of them; *n is how many
 * there are. */
int oll_engine_kernel_launches(oll_engine *e, int cap, double *bytes, double *ms, int *kind,
                               int *n);

/* Launch-shape table (no reference counterpart; filled by the autotuner,
 * ollipticid_amd/autotune.py): threads per workgroup of the smoothers'
 * colour passes for one precision (OLL_PREC_*) and order, a multiple of 64
 * from 64 to 512 (default 256; raised to hold one interior row).
 * Process-wide; schedules built afterwards use it.  Results do not depend
 * on it.  The setter returns OLL_EINVAL for other values. */
int oll_set_sweep_threads(int precision, int order, int threads);
int oll_get_sweep_threads(int precision, int order);
/* Fused FP64 sweeps (no reference counterpart): on, FP64 smoothing phases
 * run their sweeps in pairs as double-buffered launches that each do two
 * colour passes on a tile of ty x tz interior rows (u and the r field swap
 * roles from sweep to sweep), an odd last sweep colour by colour; off (the
 * default), every sweep colour by colour, which measures faster on MI355X.
 * The results are the same bits either way.  Process-wide; schedules
 * built afterwards use it. */
int oll_set_fused_sweeps(int on);
int oll_get_fused_sweeps(void);
/* Shared-load transfer kernels (no reference counterpart): on (the
 * default), the ghost-shell fill (elliptic.cc:1905-1908) of boxes of order
 * 6 and 8 runs one thread per column of R shell points across a face slab,
 * loading the interpolation nodes once for the whole column (orders 2 and
 * 4 stay point by point, which measures faster there), and the defect
 * filter ahead of an off-lattice restriction runs one thread per run of 4
 * points along x, loading the neighbour rows once for the run; off, both
 * go point by point.  The same bits either way.  Process-wide; schedules
 * built afterwards use it. */
int oll_set_shared_loads(int on);
int oll_get_shared_loads(void);
/* tile of the fused sweeps for one order: ty x tz interior rows (1..64
 * each) and threads per workgroup (a multiple of 64, 64..512; default
 * 8 x 8 rows, 256 threads); OLL_EINVAL for other values */
int oll_set_fused_shape(int order, int ty, int tz, int threads);
int oll_get_fused_shape(int order, int *ty, int *tz, int *threads);

#ifdef __cplusplus
}
#endif
#endif
