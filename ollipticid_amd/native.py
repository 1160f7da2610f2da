"""ctypes binding of libolliptic_amd.so (the C ABI of include/olliptic_amd.h).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback in the product path.
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libolliptic_amd.so")

MAX_BH = 16
EQ_PUNCTURES, EQ_PUNCTURES_SCALAR, EQ_POISSON, EQ_BRILL, EQ_PUNCTURES_EM = 0, 1, 2, 3, 4
EQ_NAMES = {"Punctures": EQ_PUNCTURES, "Punctures_scalar_field": EQ_PUNCTURES_SCALAR,
            "Poisson": EQ_POISSON, "Brill-waves": EQ_BRILL, "Punctures_EM": EQ_PUNCTURES_EM}
PREC_FP64, PREC_MIXED = 0, 1
STOP_REFERENCE, STOP_TOLERANCE = 0, 1
METHOD_VCYCLE, METHOD_APP = 0, 1
GUESS_ZERO, GUESS_APPROX = 0, 1
F_U, F_B, F_C, F_D, F_E, F_RHO, F_R, F_V, F_W, F_WW = range(10)
# Punctures_EM only (eta in F_B, phi in F_D): their gradients
F_ETA_X, F_ETA_Y, F_ETA_Z, F_PHI_X, F_PHI_Y, F_PHI_Z = range(10, 16)
NFIELDS = 10
ERRORS = {-1: "EINVAL", -2: "EGRID", -3: "EDEVICE", -4: "ENAN", -5: "ENOMEM", -6: "EBUSY"}
EBUSY = -6


class Puncture(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("m", "x", "y", "z", "px", "py", "pz", "sx", "sy", "sz", "q")]


class Problem(C.Structure):
    _fields_ = [
        ("eq", C.c_int), ("num_bh", C.c_int), ("bh", Puncture * MAX_BH),
        ("L", C.c_double), ("N", C.c_int), ("levels", C.c_int), ("order", C.c_int),
        ("nu1", C.c_int), ("nu2", C.c_int), ("tol", C.c_double),
        ("robin_a", C.c_double), ("robin_n", C.c_int),
        ("ps_case", C.c_int), ("ps_r0", C.c_double), ("ps_sigma", C.c_double),
        ("ps_phi0", C.c_double), ("ps_a2", C.c_double),
        ("max_cycles", C.c_int), ("precision", C.c_int),
        ("stop_rule", C.c_int), ("patience", C.c_int),
        ("method", C.c_int), ("initial_guess", C.c_int),
    ]


class BoxDesc(C.Structure):
    _fields_ = [("level", C.c_int), ("order", C.c_int), ("ghost", C.c_int),
                ("n", C.c_int * 3), ("origin", C.c_double * 3), ("d", C.c_double * 3),
                ("centre", C.c_double * 3), ("end", C.c_double * 3), ("colour_order", C.c_int)]

    def as_dict(self):
        return {"level": self.level, "order": self.order, "R": self.ghost,
                "n": list(self.n), "origin": list(self.origin), "d": list(self.d),
                "centre": list(self.centre), "end": list(self.end),
                "colour_order": self.colour_order}


class Result(C.Structure):
    _fields_ = [("cycles", C.c_int), ("residual", C.c_double), ("hist_len", C.c_int),
                ("converged", C.c_int)]


# the exported C ABI: name -> (restype, argtypes)
_vp = C.c_void_p
_P = C.POINTER(Problem)
SIGNATURES = {
    "oll_problem_default": (None, [_P]),
    "oll_hier_create": (C.c_int, [_P, C.POINTER(_vp)]),
    "oll_hier_destroy": (None, [_vp]),
    "oll_last_error": (C.c_char_p, []),
    "oll_num_levels": (C.c_int, [_vp]),
    "oll_mgrid_levels": (C.c_int, [_vp]),
    "oll_num_boxes": (C.c_int, [_vp]),
    "oll_level_boxes": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oll_box_info": (C.c_int, [_vp, C.c_int, C.POINTER(BoxDesc)]),
    "oll_device_bytes": (C.c_size_t, [_vp]),
    "oll_problem_boxes": (C.c_int, [_P, C.POINTER(BoxDesc), C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    "oll_problem_footprint": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "oll_get_field": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "oll_set_field": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "oll_smooth": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "oll_residual": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    "oll_robin": (C.c_int, [_vp, C.c_int]),
    "oll_boundary_operator": (C.c_int, [_vp, C.c_int]),
    "oll_robin_box": (C.c_int, [_vp, C.c_int]),
    "oll_operator": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "oll_fas": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    "oll_solve": (C.c_int, [_vp, C.POINTER(Result), C.POINTER(C.c_double), C.c_int]),
    "oll_adm_mass": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "oll_adm_mass_noalloc": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "oll_puncture_approx": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "oll_interp": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                             C.c_int, C.POINTER(C.c_double)]),
    "oll_last_solve_ms": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "oll_reset": (C.c_int, [_vp]),
    "oll_set_graphs": (C.c_int, [_vp, C.c_int]),
    "oll_engine_create": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(_vp)]),
    "oll_engine_destroy": (None, [_vp]),
    "oll_engine_admit": (C.c_int, [_vp, _P, C.c_longlong]),
    "oll_engine_admit_hier": (C.c_int, [_vp, _vp, C.c_longlong]),
    "oll_engine_step": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "oll_engine_pop_done": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(Result),
                                      C.POINTER(C.c_double)]),
    "oll_engine_pop_hier": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(Result),
                                      C.POINTER(_vp)]),
    "oll_engine_set_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "oll_engine_kernel_timing": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]),
    "oll_engine_kernel_launches": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]),
    "oll_set_sweep_threads": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "oll_get_sweep_threads": (C.c_int, [C.c_int, C.c_int]),
    "oll_set_fused_sweeps": (C.c_int, [C.c_int]),
    "oll_get_fused_sweeps": (C.c_int, []),
    "oll_transfer": (C.c_int, [_vp] + [C.c_int] * 9),
    "oll_set_shared_loads": (C.c_int, [C.c_int]),
    "oll_get_shared_loads": (C.c_int, []),
    "oll_set_fused_shape": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "oll_get_fused_shape": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "oll_engine_profile": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_longlong)]),
    "oll_engine_stats": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]),
    "oll_last_norms": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    "oll_last_solve_coarse_sweeps": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "oll_set_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "oll_kernel_timing": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
}

_lib = None


def load():
    """Loads the HIP library; raises when it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libolliptic_amd.so is not built: run `python -m ollipticid_amd.build` "
                "(the solver has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        apply_launch_shapes()
    return _lib


SHAPES_PATH = os.path.join(HERE, "launch_shapes.json")
PRECISION_NAMES = {"fp64": PREC_FP64, "mixed": PREC_MIXED}


def set_sweep_threads(precision, order, threads):
    """Threads per workgroup of the colour passes for one precision and
    order (a multiple of 64, 64..512); schedules built afterwards use it."""
    _check(load().oll_set_sweep_threads(precision, order, threads))


def sweep_threads(precision, order):
    return load().oll_get_sweep_threads(precision, order)


def set_fused_sweeps(on):
    """FP64 sweeps in pairs as fused double-buffered launches, or colour by
    colour (the default, faster on MI355X); the same bits either way."""
    _check(load().oll_set_fused_sweeps(1 if on else 0))


def fused_sweeps():
    return bool(load().oll_get_fused_sweeps())


def set_shared_loads(on):
    """Ghost fills of 6th- and 8th-order boxes one thread per shell column
    and defect filters one thread per run of points along x (the default),
    or both point by point; the same bits either way."""
    _check(load().oll_set_shared_loads(1 if on else 0))


def shared_loads():
    return bool(load().oll_get_shared_loads())


def set_fused_shape(order, ty, tz, threads):
    """Tile (ty x tz interior rows) and workgroup size of the fused sweeps
    for one order; schedules built afterwards use it."""
    _check(load().oll_set_fused_shape(order, ty, tz, threads))


def fused_shape(order):
    ty, tz, th = C.c_int(), C.c_int(), C.c_int()
    _check(load().oll_get_fused_shape(order, C.byref(ty), C.byref(tz), C.byref(th)))
    return ty.value, tz.value, th.value


def apply_launch_shapes(path=SHAPES_PATH):
    """Applies the autotuner's table (ollipticid_amd/autotune.py writes it)
    when the file exists: colour-pass workgroup sizes, fused-sweep tile
    shapes and the fused-sweep switch; returns the applied
    {precision: {order: threads}}."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        full = json.load(f)
    table = full.get("sweep_threads", {})
    for pname, per_order in table.items():
        for order, threads in per_order.items():
            set_sweep_threads(PRECISION_NAMES[pname], int(order), int(threads))
    for order, (ty, tz, threads) in full.get("fused_shape", {}).items():
        set_fused_shape(int(order), int(ty), int(tz), int(threads))
    set_fused_sweeps(bool(full.get("fused_sweeps", False)))
    return table


class OllError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


def _check(rc):
    if rc != 0:
        raise OllError(rc, load().oll_last_error().decode())
    return rc


def default_problem():
    p = Problem()
    load().oll_problem_default(C.byref(p))
    return p


def problem_from_argv(argv):
    """Maps a reference command line onto Problem with the reference's rules
    (interface.cc:366-874): flags match exactly, a repeated flag keeps its
    last value, unknown flags are ignored, nothing is parsed when the line has
    fewer than two arguments.  -BA/-Bn (Robin value at infinity and fall-off
    power) are flags of this solver; without -BA the Robin value is the one
    the reference hard-codes per equation (boundary_coeff, operator.cc:665,
    744, 1093): 1 for Brill-waves and Punctures_EM, 0 otherwise."""
    p = default_problem()
    a = list(argv)
    if len(a) < 2:
        return p
    eq = None
    robin_a = None
    nbh = p.num_bh
    bh = {}
    for i, tok in enumerate(a):
        if i + 1 >= len(a):
            break
        v = a[i + 1]
        if tok in ("-eq", "--equation"):
            eq = v
        elif tok in ("-NumBH", "--num-bh"):
            nbh = int(float(v))
        elif tok in ("-L", "--length", "-Lx", "--length-x"):
            p.L = float(v)
        elif tok in ("-N", "-n"):
            p.N = int(float(v))
        elif tok in ("-l", "--levels"):
            p.levels = int(float(v))
        elif tok in ("-O", "--order"):
            p.order = int(float(v))
        elif tok in ("-nu1", "--pre-smooth"):
            p.nu1 = int(float(v))
        elif tok in ("-nu2", "--post-smooth"):
            p.nu2 = int(float(v))
        elif tok in ("-s", "--tolerance"):
            p.tol = float(v)
        elif tok in ("-BA", "--boundary-value"):
            robin_a = float(v)
        elif tok == "-Bn":
            p.robin_n = int(float(v))
        elif tok in ("-m", "--method"):
            if v not in ("Vcycle", "AppPuncture"):
                raise ValueError("method %r is outside this solver (Vcycle, AppPuncture)" % v)
            p.method = METHOD_APP if v == "AppPuncture" else METHOD_VCYCLE
        elif tok in ("-ig", "--initial-guess"):
            if v not in ("zero", "approx"):
                raise ValueError("initial guess must be zero or approx")
            p.initial_guess = GUESS_APPROX if v == "approx" else GUESS_ZERO
        elif tok in ("-prec", "--precision"):
            if v not in ("fp64", "mixed"):
                raise ValueError("precision must be fp64 or mixed")
            p.precision = PREC_MIXED if v == "mixed" else PREC_FP64
        elif tok in ("-psc", "-ps_case"):
            p.ps_case = int(float(v))
        elif tok in ("-r0", "--radius_sf_r0"):
            p.ps_r0 = float(v)
        elif tok in ("-sigma", "--sigma_sf"):
            p.ps_sigma = float(v)
        elif tok in ("-phi0", "--phi0_sf"):
            p.ps_phi0 = float(v)
        elif tok in ("-a2", "--a2_sf"):
            p.ps_a2 = float(v)
        elif tok.startswith("-bh"):
            for key in ("mp", "qp", "px", "py", "pz", "sx", "sy", "sz", "x", "y", "z"):
                rest = tok[3 + len(key):]
                if tok.startswith("-bh" + key) and rest.isdigit():
                    bh[(key, int(rest) - 1)] = float(v)
                    break
    if eq is not None:
        if eq not in EQ_NAMES:
            raise ValueError("equation %r is outside this solver (%s)" % (eq, ", ".join(EQ_NAMES)))
        p.eq = EQ_NAMES[eq]
    if robin_a is not None:
        p.robin_a = robin_a
    elif p.eq in (EQ_BRILL, EQ_PUNCTURES_EM):
        p.robin_a = 1.0
    if not 1 <= nbh <= MAX_BH:
        raise ValueError("NumBH must be in 1..%d" % MAX_BH)
    p.num_bh = nbh
    names = {"mp": "m", "x": "x", "y": "y", "z": "z", "px": "px", "py": "py", "pz": "pz",
             "sx": "sx", "sy": "sy", "sz": "sz", "qp": "q"}
    for j in range(MAX_BH):
        for key, field in names.items():
            setattr(p.bh[j], field, bh.get((key, j), 0.0))
    return p


def problem_boxes(p):
    """Box table of the hierarchy (host only, no device needed)."""
    L = load()
    n, mg = C.c_int(), C.c_int()
    _check(L.oll_problem_boxes(C.byref(p), None, 0, C.byref(n), C.byref(mg)))
    arr = (BoxDesc * n.value)()
    _check(L.oll_problem_boxes(C.byref(p), arr, n.value, C.byref(n), C.byref(mg)))
    return [b.as_dict() for b in arr], mg.value


def problem_footprint(p):
    out = C.c_size_t()
    _check(load().oll_problem_footprint(C.byref(p), C.byref(out)))
    return out.value


class Hierarchy:
    """One box hierarchy resident in HBM (oll_hier)."""

    def __init__(self, problem, _handle=None):
        self.L = load()
        self.problem = problem
        if _handle is None:
            h = _vp()
            _check(self.L.oll_hier_create(C.byref(problem), C.byref(h)))
        else:
            h = _handle
        self.h = h
        self.nbox = self.L.oll_num_boxes(self.h)
        self.levels = self.L.oll_num_levels(self.h)
        self.mgrid = self.L.oll_mgrid_levels(self.h)
        self._info = [self.box_info(i) for i in range(self.nbox)]

    def close(self):
        if getattr(self, "h", None):
            self.L.oll_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def box_info(self, ib):
        d = BoxDesc()
        _check(self.L.oll_box_info(self.h, ib, C.byref(d)))
        return d.as_dict()

    def boxes(self):
        return list(self._info)

    def level_boxes(self, level):
        f, n = C.c_int(), C.c_int()
        _check(self.L.oll_level_boxes(self.h, level, C.byref(f), C.byref(n)))
        return list(range(f.value, f.value + n.value))

    def get_field(self, ib, fid):
        n = self._info[ib]["n"]
        out = np.empty((n[2], n[1], n[0]), dtype=np.float64)
        _check(self.L.oll_get_field(self.h, ib, fid, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_field(self, ib, fid, arr):
        n = self._info[ib]["n"]
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(n[2], n[1], n[0])
        _check(self.L.oll_set_field(self.h, ib, fid, a.ctypes.data_as(C.POINTER(C.c_double))))

    def smooth(self, level, nu):
        out = C.c_double()
        _check(self.L.oll_smooth(self.h, level, nu, C.byref(out)))
        return out.value

    def residual(self, level):
        out = C.c_double()
        _check(self.L.oll_residual(self.h, level, C.byref(out)))
        return out.value

    def robin(self, level):
        _check(self.L.oll_robin(self.h, level))

    def boundary_operator(self, level):
        _check(self.L.oll_boundary_operator(self.h, level))

    def robin_box(self, ibox):
        _check(self.L.oll_robin_box(self.h, ibox))

    def operator(self, level, derivative=False):
        """N(u) (or dN/du with derivative=True) into F_R on the level's
        interior points; returns max |N(u)|."""
        out = C.c_double()
        _check(self.L.oll_operator(self.h, level, int(bool(derivative)), C.byref(out)))
        return out.value

    def transfer(self, from_level, from_field, to_level, to_field, region=0, method=0, order=1,
                 all_=True, to_zero=False):
        """multigrid::Transferir: field from_field of level from_level into
        to_field of level to_level (region 0 all, 1 interior, 2 ghost shell;
        method 0 linear, 1 Lagrange of `order`)."""
        _check(self.L.oll_transfer(self.h, from_level, from_field, to_level, to_field, region,
                                   method, order, 1 if all_ else 0, 1 if to_zero else 0))

    def fas(self, level):
        out = C.c_double()
        _check(self.L.oll_fas(self.h, level, C.byref(out)))
        return out.value

    def solve(self, hist_cap=100000):
        res = Result()
        hist = (C.c_double * (2 * hist_cap))()
        _check(self.L.oll_solve(self.h, C.byref(res), hist, hist_cap))
        h = np.array(hist[: 2 * min(res.hist_len, hist_cap)]).reshape(-1, 2)
        return res.cycles, h, res.residual

    def last_solve_ms(self):
        out = C.c_double()
        _check(self.L.oll_last_solve_ms(self.h, C.byref(out)))
        return out.value

    def adm_mass(self):
        nbh = self.problem.num_bh
        p = (C.c_double * nbh)()
        s = (C.c_double * self.nbox)()
        _check(self.L.oll_adm_mass(self.h, p, s))
        return list(p), list(s)

    def adm_mass_noalloc(self):
        """adm_mass through oll_adm_mass_noalloc: no device allocation (the
        per-GPU scratch array of the library)."""
        nbh = self.problem.num_bh
        p = (C.c_double * nbh)()
        s = (C.c_double * self.nbox)()
        _check(self.L.oll_adm_mass_noalloc(self.h, p, s))
        return list(p), list(s)

    def puncture_approx(self):
        """Get_Puncture_App values (elliptic.cc:2232-2444), one per puncture."""
        out = (C.c_double * self.problem.num_bh)()
        _check(self.L.oll_puncture_approx(self.h, out))
        return list(out)

    def interp(self, field, box, xyz, order):
        """Lagrange interpolation (order + 1 points per axis) of `field` at
        points xyz[q] of boxes box[q] (ffunction::Interpol_In,
        interpolation.cc:189-303)."""
        box = np.ascontiguousarray(box, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        assert len(box) == len(xyz)
        out = np.empty(len(box))
        _check(self.L.oll_interp(self.h, field, len(box),
                                 box.ctypes.data_as(C.POINTER(C.c_int)),
                                 xyz.ctypes.data_as(C.POINTER(C.c_double)), order,
                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def last_norms(self):
        pre = (C.c_double * self.levels)()
        post = (C.c_double * self.levels)()
        _check(self.L.oll_last_norms(self.h, pre, post, self.levels))
        return list(pre), list(post)

    def last_solve_coarse_sweeps(self):
        out = C.c_longlong()
        _check(self.L.oll_last_solve_coarse_sweeps(self.h, C.byref(out)))
        return out.value

    def set_graphs(self, on):
        _check(self.L.oll_set_graphs(self.h, 1 if on else 0))

    def reset(self):
        _check(self.L.oll_reset(self.h))

    def set_kernel_timing(self, on):
        _check(self.L.oll_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        n, ms, by = C.c_int(), C.c_double(), C.c_double()
        _check(self.L.oll_kernel_timing(self.h, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value

    def device_bytes(self):
        return self.L.oll_device_bytes(self.h)


class Engine:
    """Many resident requests solved together (oll_engine)."""

    def __init__(self, hbm_budget=0, graphs=True):
        self.L = load()
        h = _vp()
        _check(self.L.oll_engine_create(hbm_budget, 1 if graphs else 0, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.oll_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, problem, tag):
        """True when admitted, False when the HBM budget is exhausted."""
        rc = self.L.oll_engine_admit(self.h, C.byref(problem), tag)
        if rc == EBUSY:
            return False
        _check(rc)
        return True

    def admit_hier(self, hier, tag):
        """Moves a resident Hierarchy into the engine as request `tag` (reset
        to u = 0); False when the HBM budget is exhausted.  `hier` is left
        empty and closed."""
        rc = self.L.oll_engine_admit_hier(self.h, hier.h, tag)
        if rc == EBUSY:
            return False
        _check(rc)
        hier.close()
        return True

    def step(self):
        n = C.c_int()
        _check(self.L.oll_engine_step(self.h, C.byref(n)))
        return n.value

    def pop_done(self):
        """List of (tag, cycles, residual, adm list or error code)."""
        out = []
        while True:
            tag = C.c_longlong()
            res = Result()
            adm = (C.c_double * MAX_BH)()
            rc = self.L.oll_engine_pop_done(self.h, C.byref(tag), C.byref(res), adm)
            if rc == 0:
                return out
            if rc < 0:
                out.append((tag.value, res.cycles, res.residual, rc))
            else:
                out.append((tag.value, res.cycles, res.residual, list(adm)))

    def pop_hier(self, problems=None):
        """List of (tag, cycles, residual, Hierarchy) of the finished requests,
        each hierarchy handed over with its fields (problems: optional
        tag -> Problem map kept on the returned objects)."""
        out = []
        while True:
            tag = C.c_longlong()
            res = Result()
            h = _vp()
            rc = self.L.oll_engine_pop_hier(self.h, C.byref(tag), C.byref(res), C.byref(h))
            if rc == 0:
                return out
            hier = Hierarchy((problems or {}).get(tag.value), _handle=h) if h.value else None
            out.append((tag.value, res.cycles, res.residual, hier if rc > 0 else rc))

    def set_kernel_timing(self, on):
        _check(self.L.oll_engine_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        """(smoother launches, their summed device ms, algorithmic bytes)"""
        n, ms, b = C.c_int(), C.c_double(), C.c_double()
        _check(self.L.oll_engine_kernel_timing(self.h, C.byref(n), C.byref(ms), C.byref(b)))
        return n.value, ms.value, b.value

    def kernel_launches(self):
        """per timed smoother launch: (algorithmic bytes, device ms, kind) arrays,
        kind = 2 * stencil order + (0 for FP64 sweep colours, 1 for FP32
        correction colours)"""
        n = C.c_int()
        _check(self.L.oll_engine_kernel_launches(self.h, 0, None, None, None, C.byref(n)))
        b, ms = np.zeros(n.value), np.zeros(n.value)
        k = np.zeros(n.value, dtype=np.int32)
        _check(self.L.oll_engine_kernel_launches(
            self.h, n.value, b.ctypes.data_as(C.POINTER(C.c_double)),
            ms.ctypes.data_as(C.POINTER(C.c_double)), k.ctypes.data_as(C.POINTER(C.c_int)),
            C.byref(n)))
        return b, ms, k

    def profile(self):
        """host ms in schedule builds, cycle runs, admissions; rebuilds; steps"""
        b, r, a = C.c_double(), C.c_double(), C.c_double()
        nb, nc = C.c_longlong(), C.c_longlong()
        _check(self.L.oll_engine_profile(self.h, C.byref(b), C.byref(r), C.byref(a), C.byref(nb),
                                         C.byref(nc)))
        return {"build_ms": b.value, "run_ms": r.value, "admit_ms": a.value,
                "rebuilds": nb.value, "steps": nc.value}

    def stats(self):
        b, a, d = C.c_size_t(), C.c_int(), C.c_int()
        _check(self.L.oll_engine_stats(self.h, C.byref(b), C.byref(a), C.byref(d)))
        return {"bytes_in_use": b.value, "active": a.value, "done_pending": d.value}
