"""Output writers on the GPU: the point interpolation behind them (oll_interp)
against the oracle, and the files of the drop-in command line against the
reference's own files (tests/golden/out_*, written by the reference binary).

oll_interp must equal the oracle's Lagrange interpolation bit for bit on the
same field values (both evaluate interpolation.cc:189-303 in the same order
without contraction) and reproduce polynomials of degree <= ord.

The files must have the reference's names, blocks and coordinates (the
coordinate columns exactly, as printed); values within 1e-9 of the
reference's, the bar of the converged field itself.  Both runs are ones the
reference iterates to convergence on every level (its stop rule watches the
finest level only; where it stops with coarser levels still at 1e-6 its
field depends on the sweep order, which the GPU's multi-colour sweeps do not
share).  Values the writers take
from a coarsened multigrid level (the first and last R points of its 1-D
cuts, which lie in its Robin shell, and composite points that no refinement
box holds) carry FAS states, not the solution: they are compared to 1e-6."""
import os
import subprocess

import numpy as np
import pytest

import ollipticid_amd as oa
from oracle import oracle as O
from tests.util import load_golden

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "ollipticid_amd", "Olliptic.x")

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
O8 = PUNC + ["-NumBH", "1", "-bhmp1", "0.8", "-bhx1", "0.5", "-bhpy1", "0.15",
             "-L", "50", "-N", "25", "-l", "2", "-O", "8", "-nu2", "4", "-s", "1e-12"]


def grid_axes(b):
    return [b["origin"][a] + np.arange(b["n"][a]) * b["d"][a] for a in range(3)]


def sample_points(b, rng, m):
    """Points inside the box, a little outside it, and exactly on grid points."""
    ax = grid_axes(b)
    lo = np.array([ax[a][0] for a in range(3)])
    hi = np.array([ax[a][-1] for a in range(3)])
    span = hi - lo
    inside = lo + span * rng.uniform(0, 1, (m, 3))
    outside = lo + span * rng.uniform(-0.05, 1.05, (m // 4, 3))
    idx = rng.integers(0, np.array(b["n"]), (m // 4, 3))
    nodes = np.stack([ax[a][idx[:, a]] for a in range(3)], 1)
    # grid lines: two coordinates on grid points, one between
    lines = nodes.copy()
    lines[:, 1] = lo[1] + span[1] * rng.uniform(0, 1, m // 4)
    return np.concatenate([inside, outside, nodes, lines])


@pytest.fixture(scope="module")
def pair():
    print("interp pair argv:", " ".join(O8), flush=True)
    h = oa.Hierarchy(oa.problem_from_argv(O8))
    s = O.Solver.from_argv(O8)
    assert [b["n"] for b in h.boxes()] == [b["n"] for b in s.boxes()]
    for bh, bs in zip(h.boxes(), s.boxes()):
        assert bh["origin"] == bs["origin"] and bh["d"] == bs["d"]
    yield h, s
    h.close()
    s.close()


@pytest.mark.parametrize("ord_", [3, 5, 7, 9])
def test_interp_matches_oracle_bitwise(pair, ord_):
    h, s = pair
    rng = np.random.default_rng(ord_)
    boxes, pts = [], []
    for ib, b in enumerate(h.boxes()):
        if min(b["n"]) <= ord_:
            continue
        x, y, z = np.meshgrid(*grid_axes(b)[::-1], indexing="ij")[::-1]
        u = np.sin(0.3 * x + 0.1) * np.cos(0.2 * y) * (1 + 0.05 * z) + 1e-3 * rng.standard_normal(x.shape)
        h.set_field(ib, oa.F_U, u)
        s.field(ib, O.F_U)[:] = u
        p = sample_points(b, rng, 64)
        boxes += [ib] * len(p)
        pts.append(p)
    assert boxes
    pts = np.concatenate(pts)
    got = h.interp(oa.F_U, boxes, pts, ord_)
    want = np.array([s.interp(ib, O.F_U, O.I_LAGRANGE, ord_, *p) for ib, p in zip(boxes, pts)])
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (ord_, bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("ord_", [3, 9])
def test_interp_reproduces_polynomials(pair, ord_):
    h, _ = pair
    ib = h.nbox - 1
    b = h.boxes()[ib]
    ax = grid_axes(b)
    ctr = np.array([0.5 * (a[0] + a[-1]) for a in ax])
    half = np.array([0.5 * (a[-1] - a[0]) for a in ax])
    rng = np.random.default_rng(11)
    c = rng.standard_normal(3 * (ord_ + 1)).reshape(3, ord_ + 1)

    def f(x, y, z):
        # degree ord in each variable on its own
        t = [(x - ctr[0]) / half[0], (y - ctr[1]) / half[1], (z - ctr[2]) / half[2]]
        return np.polyval(c[0], t[0]) * np.polyval(c[1], t[1]) + np.polyval(c[2], t[2])

    x, y, z = np.meshgrid(*ax[::-1], indexing="ij")[::-1]
    h.set_field(ib, oa.F_U, f(x, y, z))
    p = sample_points(b, rng, 128)
    got = h.interp(oa.F_U, [ib] * len(p), p, ord_)
    want = f(p[:, 0], p[:, 1], p[:, 2])
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_interp_rejects_bad_arguments(pair):
    h, _ = pair
    with pytest.raises(oa.OllError):
        h.interp(oa.F_U, [h.nbox], np.zeros((1, 3)), 3)
    with pytest.raises(oa.OllError):
        h.interp(oa.F_U, [0], np.zeros((1, 3)), 10)


# ----------------------------------------------------------------- files
def parse_blocks(path):
    rows, blocks, cur = [], [], 0
    with open(path) as f:
        for ln in f:
            if ln.strip():
                rows.append([float(v) for v in ln.split()])
                cur += 1
            else:
                blocks.append(cur)
                cur = 0
    if cur:
        blocks.append(cur)
    return np.array(rows), blocks


def coarse_points(boxes, mgrid, pts):
    """Composite points read from a coarsened multigrid level: those no
    refinement box holds in its interior less its buffer (the writers'
    box rule, multigrid.cc:223-241)."""
    out = np.zeros(len(pts), dtype=bool)
    for q, p in enumerate(pts):
        lev = 0
        for b in boxes:
            ax = grid_axes(b)
            w = b["R"] - 0.1
            if all(ax[a][0] + b["d"][a] * w < p[a] < b["end"][a] - b["d"][a] * w for a in range(3)):
                lev = b["level"]
        out[q] = lev < mgrid
    return out


@pytest.mark.parametrize("case", ["out_cfg1", "out_o2_4bh"])
def test_cli_output_files_match_reference(tmp_path, case):
    meta, ref = load_golden(case)
    print("cli outputs", case, flush=True)
    out = subprocess.run([EXE] + meta["argv"] + ["-o", "run"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    for ln in meta["stdout"]:
        assert ln in out.stdout
    run = tmp_path / "run"
    files = sorted(f for f in os.listdir(run) if f.startswith("u-l") and f[-2:] in (".x", ".y", ".z")
                   or f in ("u.X", "u.Y", "u.Z", "u.XY"))
    assert files == meta["files"]
    with oa.Hierarchy(oa.problem_from_argv(meta["argv"])) as h:
        boxes, mgrid = h.boxes(), h.mgrid
    worst, worst_coarse = 0.0, 0.0
    for f in files:
        got, blocks = parse_blocks(run / f)
        assert blocks == meta["blocks"][f], f
        want = ref[f]
        assert got.shape == want.shape, f
        ncoord = want.shape[1] - 1
        np.testing.assert_array_equal(got[:, :ncoord], want[:, :ncoord], err_msg=f)
        coarse = np.zeros(len(want), dtype=bool)
        if f.startswith("u-l"):
            lv = int(f[3:f.index("b")])
            if lv < mgrid:
                R = boxes[[b["level"] for b in boxes].index(lv)]["R"]
                coarse[:R] = coarse[-R:] = True
        else:
            axes = {"u.X": [0], "u.Y": [1], "u.Z": [2], "u.XY": [0, 1]}[f]
            pts = np.zeros((len(want), 3))
            for c, a in enumerate(axes):
                pts[:, a] = want[:, c]
            coarse = coarse_points(boxes, mgrid, pts)
        d = np.abs(got[:, -1] - want[:, -1])
        worst = max(worst, d[~coarse].max(initial=0))
        worst_coarse = max(worst_coarse, d[coarse].max(initial=0))
        assert d[~coarse].max(initial=0) <= 1e-9, (f, d[~coarse].max())
        assert d[coarse].max(initial=0) <= 1e-6, (f, d[coarse].max())
    print("%s: max |du| %.3g (solution points), %.3g (coarsened-level points)"
          % (case, worst, worst_coarse), flush=True)
