"""The oracle's Lagrange point interpolation (interpolation.cc:189-303), which
the output writers and the ADM mass use, pinned by what the mathematics
fixes: with ord + 1 points per axis it reproduces every polynomial of degree
<= ord in each variable (inside the box and, with the stencil shifted to the
edge, outside it), returns the grid value at a grid point, and a polynomial
of degree ord + 1 is not reproduced (so a dropped or repeated stencil point
fails)."""
import numpy as np
import pytest

from oracle import oracle as O

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
ARGV = PUNC + ["-NumBH", "1", "-bhmp1", "0.8", "-bhx1", "0.5", "-bhpy1", "0.15",
               "-L", "50", "-N", "25", "-l", "2", "-O", "8", "-nu2", "4", "-s", "1e-12"]


def grid(b):
    n, o, d = b["n"], b["origin"], b["d"]
    ax = [o[a] + np.arange(n[a]) * d[a] for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def poly(x, y, z, deg, c):
    """sum of c-weighted monomials x^i y^j z^k with i, j, k <= deg."""
    out = np.zeros_like(np.asarray(x, dtype=float))
    for i in range(deg + 1):
        for j in range(deg + 1):
            for k in range(deg + 1):
                out = out + c[i, j, k] * x ** i * y ** j * z ** k
    return out


@pytest.fixture(scope="module")
def solver():
    s = O.Solver.from_argv(ARGV)
    yield s
    s.close()


@pytest.mark.parametrize("ord_", [1, 2, 3, 5, 7, 9])
def test_reproduces_polynomials(solver, ord_):
    ib = solver.nbox - 1
    b = solver.box_info(ib)
    assert min(b["n"]) > ord_
    rng = np.random.default_rng(ord_)
    # scaled coordinates keep the monomials of one size
    ctr = [b["origin"][a] + 0.5 * (b["n"][a] - 1) * b["d"][a] for a in range(3)]
    half = [0.5 * (b["n"][a] - 1) * b["d"][a] for a in range(3)]
    c = rng.standard_normal((ord_ + 1,) * 3) / 3.0 ** np.add.outer(
        np.add.outer(np.arange(ord_ + 1), np.arange(ord_ + 1)), np.arange(ord_ + 1))
    x, y, z = grid(b)
    s = lambda x, y, z: poly((x - ctr[0]) / half[0], (y - ctr[1]) / half[1],  # noqa: E731
                             (z - ctr[2]) / half[2], ord_, c)
    solver.field(ib, O.F_U)[:] = s(x, y, z)
    pts = np.stack([ctr[a] + half[a] * rng.uniform(-1.05, 1.05, 64) for a in range(3)], 1)
    scale = np.abs(solver.field(ib, O.F_U)).max()
    for p in pts:
        got = solver.interp(ib, O.F_U, O.I_LAGRANGE, ord_, *p)
        assert abs(got - s(*p)) <= 1e-11 * scale, (ord_, p)
    # one degree more is not reproduced
    c2 = np.zeros((ord_ + 2,) * 3)
    c2[ord_ + 1, 0, 0] = 1.0
    solver.field(ib, O.F_U)[:] = poly((x - ctr[0]) / half[0], 0 * y, 0 * z, ord_ + 1, c2)
    p = pts[0]
    exact = ((p[0] - ctr[0]) / half[0]) ** (ord_ + 1)
    err = abs(solver.interp(ib, O.F_U, O.I_LAGRANGE, ord_, *p) - exact)
    assert err > 1e-9


def test_grid_points_return_grid_values(solver):
    ib = solver.nbox - 1
    b = solver.box_info(ib)
    u = np.random.default_rng(1).standard_normal(solver.field(ib, O.F_U).shape)
    solver.field(ib, O.F_U)[:] = u
    x, y, z = grid(b)
    for (k, j, i) in [(0, 0, 0), (3, 7, 11), (b["n"][2] - 1, b["n"][1] - 1, b["n"][0] - 1)]:
        got = solver.interp(ib, O.F_U, O.I_LAGRANGE, 9, x[k, j, i], y[k, j, i], z[k, j, i])
        assert got == u[k, j, i]
