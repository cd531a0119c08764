"""The C++ `bench` harness (reference argv and TSV output) and a harness
written against the reference's launchUpdateGrid.h, on the GPU."""
import os
import subprocess

import pytest

from threadcoarsenedstencilgenerator_amd import build, manifest, tuned

pytestmark = pytest.mark.gpu


def _bench(kernel_desc, *args):
    env = dict(os.environ, TCSG_KERNEL=build.kernel_path(kernel_desc))
    return subprocess.run([build.BENCH_BIN, *args], capture_output=True, text=True, env=env, timeout=120)


def test_bench_sweep_output_format():
    d = manifest.cfg1_desc("plain")  # 2D Box Jacobi r=1, 1x1, Row, None
    r = _bench(d, "Box", "2D", "1", "Jacobi", "Row", "None", "0", "0.0", "1", "1", "130", "70")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[0].startswith("#GPU: ")
    assert "#grid size excluding border: x 128 y 68" in r.stdout
    hdr = [l for l in lines if l.startswith("#xthreads")]
    assert hdr and hdr[0].split("\t")[:3] == ["#xthreads", "ythreads", "zthreads"]
    rows = [l.split("\t") for l in lines if not l.startswith("#")]
    # every swept shape of 64..1024 threads (hip/benchmark.cpp:221-249), one step each
    assert len(rows) == 230
    assert all(row[2] == "0" and int(row[7]) == 1 and float(row[5]) > 0 for row in rows)


def test_bench_single_block_and_mismatch():
    d, (bx, by, bz) = tuned.CFG2_F64["desc"], tuned.CFG2_F64["block"]
    r = _bench(d, "Star", "3D", "4", "Variable", "Row", "Write", "16", "0.0", "2", "4", "128", "264", "72", "264",
               str(bx), str(by), str(bz))
    assert r.returncode == 0, r.stderr
    rows = [l for l in r.stdout.split("\n") if l and not l.startswith("#")]
    assert len(rows) == 1 and rows[0].startswith(f"{bx}\t{by}\t{bz}\t")
    # parameters that differ from the kernel's are refused (the reference's sanity check)
    r = _bench(d, "Star", "3D", "3", "Variable", "Row", "Write", "16", "0.0", "2", "4", "128", "136", "72", "264")
    assert r.returncode != 0 and "different stencil parameters" in r.stderr


def test_launchupdategrid_harness():
    d = manifest.cfg1_desc("plain")
    env = dict(os.environ, TCSG_KERNEL=build.kernel_path(d))
    r = subprocess.run([build.SHIM_DEMO, "Box", "Jacobi", "1", "2", "None", "258", "130"], capture_output=True,
                       text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "steps 1" in r.stdout
    r = subprocess.run([build.SHIM_DEMO, "Star", "Jacobi", "1", "2", "None", "258", "130"], capture_output=True,
                       text=True, env=env, timeout=60)
    assert r.returncode == 3  # sanity check throws hipErrorInvalidValue


def test_run_benchmark_script_small_grid(tmp_path):
    """RunBenchmark.sh end to end on a small 2D grid: it generates and compiles the
    kernel with the reference's argv, then `bench` sweeps every block shape."""
    script = os.path.join(os.path.dirname(build.BENCH_BIN), "..", "RunBenchmark.sh")
    env = dict(os.environ, X2D="130", Y2D="70", TIME="0.0")
    r = subprocess.run(["bash", script, "Box", "2D", "1", "Jacobi", "Row", "None", "0", "2", "2"], cwd=tmp_path,
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (tmp_path / "kernels.co").exists()
    assert "#grid size excluding border: x 128 y 68" in r.stdout and "#coarsening factors: x 2 y 2" in r.stdout
    rows = [l.split("\t") for l in r.stdout.strip().split("\n") if l and not l.startswith("#")]
    assert len(rows) == 230 and all(float(row[5]) > 0 for row in rows)


def _init_cases():
    """One built PLAIN kernel per (dimension, alignment) of the parity matrix, with its grid."""
    from threadcoarsenedstencilgenerator_amd import spec
    seen, out = set(), []
    for d, size in manifest.parity_cases():
        key = (d.dimension, d.alignment)
        if d.variant == spec.PLAIN and d.order == "point" and key not in seen:
            seen.add(key)
            out.append((d, size))
    return out


@pytest.mark.parametrize("case", _init_cases(), ids=lambda c: f"{c[0].dimension}d_align{c[0].alignment}_{c[0].key()}")
def test_launch_init_reference_matches_oracle(case):
    """tcsg_launch_init(TCSG_INIT_REFERENCE), the replacement of launchInit
    (hip/generateLaunchUpdateGrid.py:150-158), writes the reference's initial values
    (hip/generate2DStencil_hip.py:322-351, hip/generate3DStencil_hip.py:387-431) over the
    whole padded buffer, bit for bit as the oracle's init_reference on the columns the
    reference writes, and 0 in the Read/Write padding it leaves unwritten."""
    import numpy as np
    import torch

    import oracle
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d, size = case
    k = rt.Kernel(d)
    g = k.geometry(size)
    grids = rt.Grids(g, k.dtype)
    grids.upload(0, np.full(g.n_points, np.nan))
    k.init(g, grids.ptr(0), rt.INIT_REFERENCE)
    torch.cuda.synchronize()
    got = grids.host(0)
    og = oracle.Geometry(d.dimension, d.range, d.alignment, d.cl_size, size, d.coarsening)
    ref = np.full(og.n_points, np.nan)
    oracle.init_reference(og, ref)
    print(f"init reference {d.key()} size {size}: pitch {og.xdim}, lo {og.lo}, ro {og.ro}", flush=True)
    assert got.shape == ref.shape
    # the reference writes the columns [lo, xdim - ro) of every row; the padding columns,
    # which it leaves unwritten, are 0 here
    x = np.arange(got.size) % og.xdim
    written = (x >= og.lo) & (x < og.xdim - og.ro)
    assert np.array_equal(got[written].view(np.uint64), ref[written].view(np.uint64))
    assert not np.isnan(ref[written]).any() and np.isnan(ref[~written]).all()
    assert (got[~written] == 0.0).all()
    k.close()
