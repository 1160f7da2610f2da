/*
 * Oracle: plain-C restatement of the reference FAS multigrid solver for the
 * puncture Hamiltonian constraint (Punctures, Punctures_scalar_field,
 * Punctures_EM, and the Poisson and Brill-waves test equations; Robin outer
 * boundary, V-cycles).
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, and only as the
 * checker.  The product never links it.
 *
 * Every function in olliptic_oracle.c cites the reference file:line it
 * restates.  Two sweep orders are provided:
 *   ORACLE_LEX    - lexicographic point Gauss-Seidel-Newton, the reference's
 *                   own order (elliptic.cc:1126-1242);
 *   ORACLE_COLOUR - (i+j+k) mod (R+1) multi-colour order, R = stencil radius
 *                   (mod 5 for Punctures_EM at orders 4-8, whose first
 *                   derivatives reach 4 points to one side), which is the
 *                   order the GPU kernels use.  Within one colour
 *                   no point reads another point of the same colour, so the
 *                   result does not depend on the order inside a colour.
 * Pinning: LEX reproduces the reference's residual histories bit for bit on
 * every fixture of tests/golden (every equation); COLOUR with full-weighting
 * restriction and the over-relaxed coarsest loop converges to the
 * reference's converged fields.  Parity unpinned: the COLOUR iteration of
 * Punctures_EM at orders 4-8, whose reference runs stop before converging.
 */
#ifndef OLLIPTIC_ORACLE_H
#define OLLIPTIC_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_BH 16

enum oracle_eq {
  ORACLE_PUNCTURES = 0, ORACLE_PUNCTURES_SCALAR = 1, ORACLE_POISSON = 2, ORACLE_BRILL = 3,
  ORACLE_PUNCTURES_EM = 4
};
enum oracle_order { ORACLE_LEX = 0, ORACLE_COLOUR = 1 };
/* OF_A1 .. OF_EME: the thirteen Punctures_EM coefficients A1, A2, A3, B,
 * B1x, B1y, B1z, B2x, B2y, B2z, C, D, E (operator.cc:191-221), allocated
 * only for that equation (oracle_field returns NULL otherwise) */
enum oracle_field {
  OF_U = 0, OF_B, OF_C, OF_D, OF_E, OF_RHO, OF_R, OF_V, OF_W, OF_WW,
  OF_A1, OF_A2, OF_A3, OF_EMB, OF_B1X, OF_B1Y, OF_B1Z, OF_B2X, OF_B2Y, OF_B2Z,
  OF_EMC, OF_EMD, OF_EME, OF_NFIELDS
};
#define OF_NBASE OF_A1
enum oracle_region { OR_ALL = 0, OR_INSIDE = 1, OR_BOUNDARY = 2 };
enum oracle_interp { OI_LINEAR = 0, OI_LAGRANGE = 1 };

typedef struct oracle_params {
  int eq;               /* oracle_eq */
  int nbh;
  double mp[ORACLE_MAX_BH];
  double x[ORACLE_MAX_BH], y[ORACLE_MAX_BH], z[ORACLE_MAX_BH];
  double px[ORACLE_MAX_BH], py[ORACLE_MAX_BH], pz[ORACLE_MAX_BH];
  double sx[ORACLE_MAX_BH], sy[ORACLE_MAX_BH], sz[ORACLE_MAX_BH];
  double L;             /* domain length (-L) */
  int N;                /* points per side (-N), snapped by Valid_N */
  int levels;           /* refinement levels (-l) */
  int order;            /* 2,4,6,8 (-O) */
  int nu1, nu2;         /* pre/post sweeps */
  double tol;           /* -s */
  int ps_case;          /* scalar field: -psc */
  double ps_r0, ps_sigma, ps_phi0, ps_a2;
  int sweep_order;      /* oracle_order */
  int max_cycles;       /* 0 = no limit */
  int restrict_fw;      /* 1: full-weighting restriction of the FAS defect
                           where the 27-point stencil lies in the fine
                           interior (the GPU's choice); 0: the reference's
                           injection */
  double coarse_omega;  /* over-relaxation of the coarsest-level loop:
                           0 plain Gauss-Seidel (the reference), >0 that
                           factor, <0 2 - 1.6/(n-1) (the GPU's choice) */
  int method_app;       /* 1: -m AppPuncture, whose grids have no multigrid
                           coarsening (elliptic.cc:510, 678) */
  double qp[ORACLE_MAX_BH]; /* puncture charges -bhqp (Punctures_EM) */
} oracle_params;

typedef struct oracle_solver oracle_solver;

void oracle_default_params(oracle_params *p);
/* Returns NULL on error (message via oracle_last_error). */
oracle_solver *oracle_create(const oracle_params *p);
void oracle_destroy(oracle_solver *s);
const char *oracle_last_error(void);

int oracle_num_levels(const oracle_solver *s);   /* incl. multigrid levels */
/* Sweeps spent in the coarsest-level loop since creation. */
long oracle_coarse_iterations(const oracle_solver *s);
int oracle_mgrid_levels(const oracle_solver *s);
int oracle_num_boxes(const oracle_solver *s);
int oracle_level_first_box(const oracle_solver *s, int level);
int oracle_level_num_boxes(const oracle_solver *s, int level);
/* n[3], origin[3] (x_i of index 0 per axis), d[3], ext[6] = xi,xf,yi,yf,zi,zf */
void oracle_box_info(const oracle_solver *s, int ibox, int *level, int *order,
                     int *R, int n[3], double origin[3], double d[3],
                     double ext[6]);
double *oracle_field(oracle_solver *s, int ibox, int field);

/* Single operations on one box (ibox global index). */
double oracle_smooth(oracle_solver *s, int ibox, int nu);
/* 1 when a two-colour box's multi-colour sweeps take colour 1 first (the
 * GPU's order), 0 otherwise; -1 for a box index out of range */
int oracle_colour_flip(const oracle_solver *s, int ibox);
double oracle_residual(oracle_solver *s, int ibox);
void oracle_robin(oracle_solver *s, int ibox);
void oracle_boundary_operator(oracle_solver *s, int ibox);
/* multigrid::Transferir between two levels for one field. */
void oracle_transfer(oracle_solver *s, int from_level, int from_field,
                     int to_level, int to_field, int region, int method,
                     int ord, int all, int to_zero);
/* The FAS step of level k (k >= 1); returns the post-smoothing norm. */
double oracle_fasn(oracle_solver *s, int k);
/* Full solve. hist receives (cycle, norm) pairs as written to Loo.gp;
 * returns number of V-cycles, <0 on error. */
int oracle_solve(oracle_solver *s, double *hist, int hist_cap, int *n_hist,
                 double *final_res);
/* ADM masses: per puncture (nbh values) and per box surface integral. */
void oracle_adm_mass(oracle_solver *s, double *puncture, double *surface);
/* elliptic.cc:1977-2222 ApproximateID: u on every point of every box set
 * to the analytic approximation (spin, momentum and cross terms of each
 * puncture up to the first with zero mass). */
void oracle_approx_id(oracle_solver *s);
/* elliptic.cc:2232-2444 Get_Puncture_App: per puncture, the approximation
 * summed over the punctures at the grid node Get_i/j/k of its position in
 * every finest-level box that holds it (0.1 dx tolerance). */
void oracle_puncture_approx(const oracle_solver *s, double *u_at_p);

/* Lagrange/linear point interpolation of a field of one box (testing). */
double oracle_interp_point(oracle_solver *s, int ibox, int field, int method,
                           int ord, double x, double y, double z);
/* Valid_N snap (interface.cc:883-949). */
size_t oracle_valid_n(size_t n);

#ifdef __cplusplus
}
#endif
#endif
