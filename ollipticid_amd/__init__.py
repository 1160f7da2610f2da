"""MI355X-native FAS multigrid solver of the puncture Hamiltonian constraint.

The solver is the HIP library libolliptic_amd.so (C ABI in
include/olliptic_amd.h); this package is its Python binding and serving
front end.  Build with `python -m ollipticid_amd.build`.
"""
from .native import (  # noqa: F401
    EQ_BRILL, EQ_POISSON, EQ_PUNCTURES, EQ_PUNCTURES_EM, EQ_PUNCTURES_SCALAR, F_B,
    F_ETA_X, F_ETA_Y, F_ETA_Z, F_PHI_X, F_PHI_Y, F_PHI_Z, F_C, F_D, F_E, F_R, F_RHO, F_U, F_V, F_W,
    F_WW, GUESS_APPROX, GUESS_ZERO, METHOD_APP, METHOD_VCYCLE, PREC_FP64, PREC_MIXED,
    STOP_REFERENCE, STOP_TOLERANCE, Engine, Hierarchy, OllError,
    Problem, apply_launch_shapes, default_problem, load, set_sweep_threads, sweep_threads,
    problem_boxes, problem_footprint, problem_from_argv,
)
