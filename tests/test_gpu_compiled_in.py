"""The reference's build layout on the GPU: the generated kernels.cu compiled
into the launch unit that generateLaunchUpdateGrid.py prints (as
hip/generateLaunchUpdateGrid.py:55-58 compiles its kernels.cu in), driven

- through that unit's launchUpdateGrid.h entry points, set up as
  hip/benchmark.cpp sets up its grids (the reference's initial values, weights
  1/n), by a harness of our own (harness/compiled_in_check.cpp, mode "unit"):
  bit for bit against the oracle after three steps;
- through the kernel file's own symbols (updateGrid, init, d_mask), launched
  the way the reference's generated unit launches them (mode "plugin"): bit for
  bit as well, also with a block shape the fast body cannot run, which takes
  the reference-shaped body; and
- by the reference's own harness (hip/main.cpp, benchmark.cpp, primes.cpp,
  compiled from its sources into oracle/_ref by oracle/Makefile), whose
  block-shape sweep runs the generated kernel and prints its TSV.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import build, spec

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# grid size incl. halo and a block shape for each compiled-in kernel
CASES = {
    "star3d_r4_2x4x128_dppw": ((72, 36, 264), (128, 2, 1)),
    "box2d_r1_1x1_plain": ((130, 66), (64, 4, 1)),
    "box3d_r2_2x3x2_dpp_read": ((23, 19, 17), (16, 4, 2)),
}
NAMES = {spec.BOX: "Box", spec.STAR: "Star"}
COEFFS = {spec.JACOBI: "Jacobi", spec.VARIABLE: "Variable"}
ALIGNS = ["None", "Read", "Write"]


def _check_args(name, size, block, steps, out, mode="unit"):
    d = build.compiled_in_desc(name)
    g = spec.geometry(d, size)
    return d, g, [os.path.join(build.COMPILED_IN_DIR, name, "compiled_in_check"), mode, NAMES[d.neighbourhood],
                  COEFFS[d.coeff], str(d.range), str(d.dimension), ALIGNS[d.alignment], str(g.xdim), str(g.ydim),
                  str(g.zdim), str(g.ro), *map(str, block), *map(str, g.threads), str(steps), out]


RUNS = [(name, "unit", CASES[name][1]) for name in CASES] + [(name, "plugin", CASES[name][1]) for name in CASES]
# the wave-aligned DPP body needs block rows of whole waves: updateGrid takes the reference-shaped body
RUNS.append(("star3d_r4_2x4x128_dppw", "plugin", (96, 2, 1)))


@pytest.mark.parametrize("name,mode,block", RUNS)
def test_compiled_in_kernel_matches_oracle(name, mode, block, tmp_path):
    size, _ = CASES[name]
    out = str(tmp_path / "grid.bin")
    d, g, argv = _check_args(name, size, block, 3, out, mode)
    print(f"compiled-in {name} {mode} size {size} block {block}", flush=True)
    r = subprocess.run(argv, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if mode == "unit":
        assert int(r.stdout.split()[1]) <= d.max_threads   # maxPotentialBlockSize within the launch bound
    got = np.fromfile(out, dtype=np.float64)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size, d.coarsening)
    assert got.size == og.n_points
    a = np.zeros(og.n_points)
    oracle.init_reference(og, a)
    ref = oracle.run_f64(d, og, np.full(d.n, 1.0 / d.n), a.copy(), a.copy(), 3)
    bad = np.flatnonzero(got.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, f"{bad.size} points differ from the oracle, first {bad[:5]}"


def test_compiled_in_unit_refuses_block_shapes_the_kernel_cannot_run(tmp_path):
    """A wave-aligned DPP kernel with block rows that are not whole waves:
    benchmarkUpdateGrid throws hipErrorInvalidValue without launching, which
    the reference's sweep reports as a failed block shape."""
    size, _ = CASES["star3d_r4_2x4x128_dppw"]
    _, _, argv = _check_args("star3d_r4_2x4x128_dppw", size, (96, 2, 1), 1, str(tmp_path / "x.bin"))
    r = subprocess.run(argv, capture_output=True, text=True, timeout=120)
    assert r.returncode == 3 and "threw" in r.stdout, r.stdout + r.stderr


def _sweep(name, grid, max_time="0.002"):
    exe = os.path.join(ROOT, "oracle", "_ref", f"refbench_{name}")
    if not os.path.exists(exe):
        pytest.skip("the reference's harness was not built (its source tree was absent at build time)")
    d = build.compiled_in_desc(name)
    argv = [exe, NAMES[d.neighbourhood], f"{d.dimension}D", str(d.range), COEFFS[d.coeff],
            "Row" if d.orientation == spec.ROW else "Column", ALIGNS[d.alignment], str(d.cl_size or 1), max_time,
            str(d.xc), str(d.yc)] + ([str(d.zc)] if d.dimension == 3 else []) + [str(v) for v in grid]
    print("reference harness:", " ".join(argv[1:]), flush=True)
    r = subprocess.run(argv, capture_output=True, text=True, timeout=240)
    rows = [ln.split("\t") for ln in r.stdout.splitlines() if ln and not ln.startswith("#")]
    return d, r, rows


def test_reference_harness_sweeps_the_headline_kernel():
    """The reference's bench, unchanged, over the 512^3 3D Star r=4 Variable
    production kernel: every block shape it reports is one the kernel runs
    (x extent a multiple of 64, at most its 256-thread launch bound), the
    others are reported as failed, and its Lup/s column is positive."""
    d, r, rows = _sweep("star3d_r4_2x4x128_dppw", (520, 520, 520))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "#grid size excluding border: x 512 y 512 z 512" in r.stdout
    assert rows, r.stdout
    for row in rows:
        x, y, z = (int(v) for v in row[:3])
        assert x % 64 == 0 and x * y * z <= 256, row
        assert float(row[5]) > 0
    assert "kernel failed for blockSize" in r.stderr
    best = max(rows, key=lambda row: float(row[5]))
    print(f"best block {best[:3]}: {float(best[5]) / 1e9:.1f} GLup/s (reference protocol, one sync per step)")


@pytest.mark.parametrize("name,grid", [("box2d_r1_1x1_plain", (1026, 1026)),
                                       ("box3d_r2_2x3x2_dpp_read", (100, 90, 80))])
def test_reference_harness_sweeps_other_kernels(name, grid):
    d, r, rows = _sweep(name, grid, "0.0005")
    assert r.returncode == 0, r.stderr[-2000:]
    assert rows and all(float(row[5]) > 0 for row in rows)
