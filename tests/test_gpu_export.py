"""The BAM (-iob) and Zcode (-ioZ) exports of the drop-in command line against
the reference's own runs on the same point files (tests/golden/io_cfg1.json,
written by the reference binary after its cfg1 solve iterated to
convergence).

Per run: the same exit status, the reference's "File found" / "reading from" /
error lines in order, and the same output files with the same lines: every
coordinate field exactly as the reference printed it, the layout (field
widths of the in-place BAM rewrite, the repeated last record of a Zcode file
that ends in a newline, the unrewritten last BAM record when the file does not
end in one), and each interpolated value within 1e-9 of the reference's (1e-6
where the point is read from a coarsened multigrid level, whose values carry
FAS states rather than the solution).

The reference's BAM loop never ends after its first file (the test's
fixtures stop it once its records are written); this build returns."""
import json
import os
import subprocess

import numpy as np
import pytest

import ollipticid_amd as oa
from tests.test_gpu_output import coarse_points

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "ollipticid_amd", "Olliptic.x")


def load_meta():
    with open(os.path.join(REPO, "tests", "golden", "io_cfg1.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def grid():
    meta = load_meta()
    with oa.Hierarchy(oa.problem_from_argv(meta["argv"])) as h:
        return h.boxes(), h.mgrid


def split_line(ln, bam):
    """(coordinate text, value) of one record line."""
    if bam:
        return ln[:50], float(ln[50:])
    t = ln.split()
    return " ".join(t[:3]), float(t[3])


@pytest.mark.parametrize("run", range(5))
def test_export_matches_reference(tmp_path, grid, run):
    meta = load_meta()
    r = meta["runs"][run]
    print("export run", r["prefix"], r["flag"], flush=True)
    for f, text in r["inputs"].items():
        (tmp_path / f).write_text(text)
    cmd = [EXE] + meta["argv"] + ["-o", "run", "-p1d", "no", "-pint", "no", r["flag"], "yes",
                                  "-iof", r["prefix"]]
    out = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    want_rc = 0 if r["rc"] is None else r["rc"]
    assert out.returncode == want_rc, out.stdout[-2000:] + out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines()
             if ln.startswith(("File found", "reading from", "Error", "X: "))]
    assert lines == r["stdout"]
    boxes, mgrid = grid
    bam = r["flag"] == "-iob"
    worst = 0.0
    for f, want in r["outputs"].items():
        got = (tmp_path / f).read_text()
        if f in r["inputs"] and not bam:
            assert got == want, f  # Zcode inputs are left alone
            continue
        assert got.endswith("\n") == want.endswith("\n"), f
        gl, wl = got.splitlines(), want.splitlines()
        assert len(gl) == len(wl), f
        rec_in = r["inputs"].get(f, "").splitlines()
        for q, (g, w) in enumerate(zip(gl, wl)):
            if bam and q < len(rec_in) and w == rec_in[q]:
                assert g == w, (f, q)  # a record the reference leaves as it was
                continue
            assert len(g) == len(w) or not bam, (f, q)
            gc, gv = split_line(g, bam)
            wc, wv = split_line(w, bam)
            assert gc == wc, (f, q)
            p = np.array([[float(v) for v in wc.split()]])
            tol = 1e-6 if coarse_points(boxes, mgrid, p)[0] else 1e-9
            assert abs(gv - wv) <= tol, (f, q, gv, wv)
            worst = max(worst, abs(gv - wv))
    print("%s: max |du| %.3g" % (r["prefix"], worst), flush=True)
