"""ctypes binding of the oracle (liboracle.so, the plain-C restatement of the
reference FAS solver).  Test infrastructure only: used by tests/, the smoke
check and the benchmark's CPU-baseline leg as the checker, never by the
product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "Olliptic.x")

MAX_BH = 16
PUNCTURES, PUNCTURES_SCALAR, POISSON, BRILL, PUNCTURES_EM = 0, 1, 2, 3, 4
EQ_NAMES = {"Punctures": PUNCTURES, "Punctures_scalar_field": PUNCTURES_SCALAR,
            "Poisson": POISSON, "Brill-waves": BRILL, "Punctures_EM": PUNCTURES_EM}
LEX, COLOUR = 0, 1
F_U, F_B, F_C, F_D, F_E, F_RHO, F_R, F_V, F_W, F_WW = range(10)
# Punctures_EM coefficients A1, A2, A3, B, B1x, B1y, B1z, B2x, B2y, B2z, C, D, E
F_EM = list(range(10, 23))
R_ALL, R_INSIDE, R_BOUNDARY = 0, 1, 2
I_LINEAR, I_LAGRANGE = 0, 1

_D16 = C.c_double * MAX_BH


class Params(C.Structure):
    _fields_ = [
        ("eq", C.c_int), ("nbh", C.c_int),
        ("mp", _D16), ("x", _D16), ("y", _D16), ("z", _D16),
        ("px", _D16), ("py", _D16), ("pz", _D16),
        ("sx", _D16), ("sy", _D16), ("sz", _D16),
        ("L", C.c_double), ("N", C.c_int), ("levels", C.c_int),
        ("order", C.c_int), ("nu1", C.c_int), ("nu2", C.c_int),
        ("tol", C.c_double), ("ps_case", C.c_int),
        ("ps_r0", C.c_double), ("ps_sigma", C.c_double),
        ("ps_phi0", C.c_double), ("ps_a2", C.c_double),
        ("sweep_order", C.c_int), ("max_cycles", C.c_int),
        ("restrict_fw", C.c_int), ("coarse_omega", C.c_double),
        ("method_app", C.c_int), ("qp", _D16),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Params)
        vp = C.c_void_p
        L.oracle_default_params.argtypes = [P]
        L.oracle_create.argtypes = [P]
        L.oracle_create.restype = vp
        L.oracle_destroy.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        for fn in ("oracle_num_levels", "oracle_mgrid_levels", "oracle_num_boxes"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = C.c_int
        for fn in ("oracle_level_first_box", "oracle_level_num_boxes"):
            getattr(L, fn).argtypes = [vp, C.c_int]
            getattr(L, fn).restype = C.c_int
        L.oracle_box_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.oracle_field.argtypes = [vp, C.c_int, C.c_int]
        L.oracle_field.restype = C.POINTER(C.c_double)
        L.oracle_smooth.argtypes = [vp, C.c_int, C.c_int]
        L.oracle_smooth.restype = C.c_double
        L.oracle_colour_flip.argtypes = [vp, C.c_int]
        L.oracle_colour_flip.restype = C.c_int
        L.oracle_residual.argtypes = [vp, C.c_int]
        L.oracle_residual.restype = C.c_double
        L.oracle_robin.argtypes = [vp, C.c_int]
        L.oracle_boundary_operator.argtypes = [vp, C.c_int]
        L.oracle_transfer.argtypes = [vp] + [C.c_int] * 9
        L.oracle_fasn.argtypes = [vp, C.c_int]
        L.oracle_fasn.restype = C.c_double
        L.oracle_solve.argtypes = [vp, C.POINTER(C.c_double), C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.oracle_solve.restype = C.c_int
        L.oracle_adm_mass.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_interp_point.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double]
        L.oracle_interp_point.restype = C.c_double
        L.oracle_approx_id.argtypes = [vp]
        L.oracle_puncture_approx.argtypes = [vp, C.POINTER(C.c_double)]
        L.oracle_coarse_iterations.argtypes = [vp]
        L.oracle_coarse_iterations.restype = C.c_long
        L.oracle_valid_n.argtypes = [C.c_size_t]
        L.oracle_valid_n.restype = C.c_size_t
        _lib = L
    return _lib


def params_from_argv(argv, sweep_order=LEX, max_cycles=0, restrict_fw=None):
    """Maps a reference-style command line (the hot-path flags) to Params,
    following the reference's defaults (interface.cc:35-355) and its parsing
    rules (a repeated flag takes its last value)."""
    p = Params()
    lib().oracle_default_params(C.byref(p))
    eq = "none"
    method = "Vcycle"
    bh = {}
    nbh = 1
    a = list(argv)
    for i, tok in enumerate(a[:-1]):
        v = a[i + 1]
        if tok == "-eq":
            eq = v
        elif tok == "-m":
            method = v
        elif tok == "-NumBH":
            nbh = int(v)
        elif tok == "-L":
            p.L = float(v)
        elif tok in ("-N", "-n"):
            p.N = int(v)
        elif tok == "-l":
            p.levels = int(v)
        elif tok == "-O":
            p.order = int(v)
        elif tok == "-nu1":
            p.nu1 = int(v)
        elif tok == "-nu2":
            p.nu2 = int(v)
        elif tok == "-s":
            p.tol = float(v)
        elif tok == "-psc":
            p.ps_case = int(v)
        elif tok == "-r0":
            p.ps_r0 = float(v)
        elif tok == "-sigma":
            p.ps_sigma = float(v)
        elif tok == "-phi0":
            p.ps_phi0 = float(v)
        elif tok == "-a2":
            p.ps_a2 = float(v)
        elif tok.startswith("-bh"):
            for key in ("mp", "qp", "px", "py", "pz", "sx", "sy", "sz", "x", "y", "z"):
                if tok.startswith("-bh" + key) and tok[3 + len(key):].isdigit():
                    bh[(key, int(tok[3 + len(key):]) - 1)] = float(v)
                    break
    p.eq = EQ_NAMES.get(eq, PUNCTURES)
    p.nbh = nbh
    for key in ("mp", "qp", "x", "y", "z", "px", "py", "pz", "sx", "sy", "sz"):
        arr = getattr(p, key)
        for j in range(MAX_BH):
            arr[j] = bh.get((key, j), 0.0)
    p.method_app = int(method == "AppPuncture")
    p.sweep_order = sweep_order
    p.max_cycles = max_cycles
    p.restrict_fw = (sweep_order == COLOUR) if restrict_fw is None else int(restrict_fw)
    # the GPU's coarsest-level over-relaxation goes with its colour order
    p.coarse_omega = -1.0 if sweep_order == COLOUR else 0.0
    return p


class Solver:
    def __init__(self, params):
        self.L = lib()
        self.h = self.L.oracle_create(C.byref(params))
        if not self.h:
            raise RuntimeError(self.L.oracle_last_error().decode())
        self.params = params
        self.nbox = self.L.oracle_num_boxes(self.h)
        self.levels = self.L.oracle_num_levels(self.h)
        self.mgrid = self.L.oracle_mgrid_levels(self.h)

    @classmethod
    def from_argv(cls, argv, **kw):
        return cls(params_from_argv(argv, **kw))

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def box_info(self, ib):
        lv, od, R = C.c_int(), C.c_int(), C.c_int()
        n = (C.c_int * 3)()
        org = (C.c_double * 3)()
        d = (C.c_double * 3)()
        ext = (C.c_double * 6)()
        self.L.oracle_box_info(self.h, ib, C.byref(lv), C.byref(od), C.byref(R), n, org, d, ext)
        return {"level": lv.value, "order": od.value, "R": R.value, "n": list(n),
                "origin": list(org), "d": list(d), "ext": list(ext)}

    def boxes(self):
        return [self.box_info(i) for i in range(self.nbox)]

    def level_boxes(self, level):
        f = self.L.oracle_level_first_box(self.h, level)
        return list(range(f, f + self.L.oracle_level_num_boxes(self.h, level)))

    def field(self, ib, fid):
        """numpy view (k, j, i) of one field of one box (shares memory)."""
        n = self.box_info(ib)["n"]
        ptr = self.L.oracle_field(self.h, ib, fid)
        if not ptr:
            raise ValueError("field %d is not held for this equation" % fid)
        return np.ctypeslib.as_array(ptr, shape=(n[2], n[1], n[0]))

    def colour_flip(self, ib):
        """1 when box ib's two-colour sweeps take colour 1 first, else 0."""
        return self.L.oracle_colour_flip(self.h, ib)

    def smooth(self, ib, nu):
        return self.L.oracle_smooth(self.h, ib, nu)

    def residual(self, ib):
        return self.L.oracle_residual(self.h, ib)

    def robin(self, ib):
        self.L.oracle_robin(self.h, ib)

    def boundary_operator(self, ib):
        self.L.oracle_boundary_operator(self.h, ib)

    def transfer(self, frm, ffld, to, tfld, region, method, ord_, all_=1, to_zero=0):
        self.L.oracle_transfer(self.h, frm, ffld, to, tfld, region, method, ord_, all_, to_zero)

    def fasn(self, k):
        return self.L.oracle_fasn(self.h, k)

    def solve(self, hist_cap=100000):
        hist = (C.c_double * (2 * hist_cap))()
        nh = C.c_int()
        res = C.c_double()
        cyc = self.L.oracle_solve(self.h, hist, hist_cap, C.byref(nh), C.byref(res))
        if cyc < 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        h = np.array(hist[: 2 * min(nh.value, hist_cap)]).reshape(-1, 2)
        return cyc, h, res.value

    def adm_mass(self):
        nbh = self.params.nbh
        p = (C.c_double * nbh)()
        s = (C.c_double * self.nbox)()
        self.L.oracle_adm_mass(self.h, p, s)
        return list(p), list(s)

    def coarse_iterations(self):
        return self.L.oracle_coarse_iterations(self.h)

    def approx_id(self):
        self.L.oracle_approx_id(self.h)

    def puncture_approx(self):
        out = (C.c_double * self.params.nbh)()
        self.L.oracle_puncture_approx(self.h, out)
        return list(out)

    def interp(self, ib, fid, method, ord_, x, y, z):
        return self.L.oracle_interp_point(self.h, ib, fid, method, ord_, x, y, z)


def valid_n(n):
    return lib().oracle_valid_n(n)
