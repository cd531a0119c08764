"""CPU checks of the drop-in surfaces: the generator command lines (the
reference's argv and error behaviour), the C ABI exports of libtcsg.so, the
launchUpdateGrid.h compatibility unit, and geometry validation through the C
ABI (no GPU call)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "threadcoarsenedstencilgenerator_amd")


def _gen(script, *args):
    return subprocess.run([sys.executable, os.path.join(PKG, script), *args], capture_output=True, text=True)


def test_generator_argv_like_reference():
    r = _gen("generate2DStencil.py", "Box", "Row", "Jacobi", "1", "1", "1")
    assert r.returncode == 0 and "tcsg_update" in r.stdout
    # the reference kernel file's symbols (hip/generate2DStencil_hip.py:70-76,221-226,322-330)
    assert ("__global__ void __launch_bounds__(1024) updateGrid(const unsigned long xdim, const unsigned long ydim, "
            "const double * const __restrict__ array0, double * const __restrict__ array1)") in r.stdout
    assert "#define range 1" in r.stdout and "__global__ void init(double * array, const unsigned long xdim, " \
        "const unsigned long ydim)" in r.stdout and "d_mask" not in r.stdout
    r = _gen("generate3DStencil.py", "Star", "Row", "Variable", "4", "1", "2", "8", "Write", "16")
    assert ("updateGrid(const unsigned long xdim, const unsigned long ydim, const unsigned long zdim, "
            "double * const __restrict__ array0, double * const __restrict__ array1, "
            "const unsigned int rightOffset)") in r.stdout
    assert "__constant__ double d_mask[25];" in r.stdout and "+ 12;" in r.stdout
    r = _gen("generate3DStencil.py", "Star", "Row", "Variable", "4", "1", "2", "8", "--dtype=f32c")
    assert r.returncode == 0 and "updateGrid" not in r.stdout   # FP32 storage: no FP64 interface
    r = _gen("generate3DStencil.py", "Star", "Column", "Variable", "4", "1", "2", "8", "Write", "16")
    assert r.returncode == 0 and "tcsg_params[24] = {1413698375, 1, 3, 1, 1, 1, 4, 1, 2, 8, 2, 16," in r.stdout
    # wrong argument counts / values: usage on stderr, status 1 (hip/generate2DStencil_hip.py:28-60)
    for args in (("Box", "Row", "Jacobi", "1", "1"), ("Boxx", "Row", "Jacobi", "1", "1", "1"),
                 ("Box", "Row", "Jacobi", "0", "1", "1"), ("Box", "Row", "Jacobi", "1", "0", "1"),
                 ("Box", "Row", "Jacobi", "1", "1", "1", "Write", "0"),
                 ("Box", "Row", "Jacobi", "1", "1", "1", "Left", "16")):
        r = _gen("generate2DStencil.py", *args)
        assert r.returncode == 1 and "usage" in r.stderr
    r = _gen("generate3DStencil.py", "Box", "Row", "Jacobi", "1", "1", "1")
    assert r.returncode == 1
    # MI355X options: remainder handling (params slot 18 bit 2) and the PLAIN statement order
    r = _gen("generate3DStencil.py", "Star", "Row", "Variable", "4", "2", "4", "8", "Write", "16", "--xext=dppw",
             "--remainder")
    assert r.returncode == 0 and "remainder=1" in r.stdout and "tcsg_st2(rout0, sfull" in r.stdout
    assert re.search(r"tcsg_params\[24\] = \{(\d+, ){18}5,", r.stdout)   # nt_store | remainder << 2
    r = _gen("generate2DStencil.py", "Box", "Row", "Variable", "8", "4", "4", "--variant=plain", "--order=offset")
    assert r.returncode == 0 and "order=offset" in r.stdout
    # paired edge-column loads: the DPP row shifts that move the second row's columns into place
    r = _gen("generate3DStencil.py", "Star", "Row", "Variable", "4", "2", "4", "8", "Write", "16", "--xext=dppw",
             "--edge-pairs")
    assert r.returncode == 0 and "eoffP" in r.stdout and "tcsg_rowshl2(eP_" in r.stdout   # H = ceil(4 / 2)
    # the offset order in a 2D STREAM kernel (register prefetch off by default there)
    r = _gen("generate2DStencil.py", "Box", "Row", "Variable", "8", "4", "4", "--xext=dppw", "--order=offset")
    assert r.returncode == 0 and "order=offset" in r.stdout and "tcsg_shr1_or" in r.stdout
    for script, extra in (("generate3DStencil.py", ("1", "--order=offset")), ("generate2DStencil.py", ("--edge-pairs",)),
                          ("generate2DStencil.py", ("--order=offset", "--prefetch=1"))):
        r = _gen(script, "Box", "Row", "Jacobi", "1", "2", "1", *extra)
        assert r.returncode == 1 and "usage" in r.stderr


def test_launch_interface_generator():
    # default: the reference's launchUpdateGrid.cu layout, kernels.cu compiled in, parameters fixed
    r = _gen("generateLaunchUpdateGrid.py", "Star", "Variable", "4", "3D", "Write")
    assert r.returncode == 0 and '#include "kernels.cu"' in r.stdout and '#include "launchUpdateGrid.h"' in r.stdout
    assert 'alignment == "Write"' in r.stdout and "stencilRange == 4u" in r.stdout and "@" not in r.stdout
    for fn in ("benchmarkUpdateGrid(", "launchInit(", "maxPotentialBlockSize()", "copyMask("):
        assert fn in r.stdout
    # --shim: the same entry points over libtcsg for kernels loaded as code objects
    r = _gen("generateLaunchUpdateGrid.py", "Star", "Variable", "4", "3D", "Write", "--shim")
    assert r.returncode == 0 and '#define TCSG_EXPECT_ALIGNMENT "Write"' in r.stdout
    assert _gen("generateLaunchUpdateGrid.py", "Star", "Variable", "4", "4D").returncode == 1


def test_c_abi_exports_every_declared_symbol():
    from threadcoarsenedstencilgenerator_amd.build import LIB
    header = open(os.path.join(ROOT, "include", "tcsg.h")).read()
    declared = set(re.findall(r"\b(tcsg_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 18
    lib = C.CDLL(LIB)
    for name in declared:
        assert hasattr(lib, name), name


def test_launchupdategrid_header_compiles_and_links():
    """A harness written against the reference's four launchUpdateGrid.h
    signatures builds against the compatibility library."""
    from threadcoarsenedstencilgenerator_amd.build import PKG as P, SHIM_LIB
    src = r'''
#include "launchUpdateGrid.h"
int main() {
  std::vector<double> (*a)(std::string, std::string, unsigned, unsigned, std::string, double, dim3, dim3, dim3,
                           double*, double*, unsigned) = benchmarkUpdateGrid;
  void (*b)(dim3, dim3, dim3, double*, unsigned) = launchInit;
  int (*c)() = maxPotentialBlockSize;
  void (*d)(double*, unsigned) = copyMask;
  return (a && b && c && d) ? 0 : 1;
}'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "h.cpp")
        open(f, "w").write(src)
        r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", f"-I{ROOT}/include", f,
                            f"-L{os.path.dirname(SHIM_LIB)}", "-llaunchUpdateGrid", f"-L{P}", "-ltcsg",
                            "-o", os.path.join(td, "h")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[:3000]


def test_geometry_check_via_c_abi():
    from threadcoarsenedstencilgenerator_amd import manifest
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    L = rt.library()
    d = manifest.cfg2_desc(2, 4, 128, shape="dppw")
    g = rt.geometry(d, (520, 520, 520))
    assert L.tcsg_geometry_check(C.byref(rt.to_cdesc(d)), C.byref(g)) == 0
    g.right_offset += 1
    assert L.tcsg_geometry_check(C.byref(rt.to_cdesc(d)), C.byref(g)) == 1
    # a STREAM kernel addresses one block's window of planes with 32-bit offsets: a 4.5 GB
    # grid of 8.7 MB planes fits, planes of 17 MB (two 136-plane windows past 4 GiB) do not
    big = rt.geometry(d, (1032, 1032, 520))
    assert big.n_points * 8 > 4 << 30
    assert L.tcsg_geometry_check(C.byref(rt.to_cdesc(d)), C.byref(big)) == 0
    wide = rt.geometry(d, (2056, 1032, 136))
    assert L.tcsg_geometry_check(C.byref(rt.to_cdesc(d)), C.byref(wide)) == 5
    # grids smaller than the halo are rejected
    with pytest.raises(rt.TcsgError):
        rt.geometry(d, (8, 520, 520))


def test_compiled_in_launch_unit_compiles(tmp_path):
    """The generated kernels.cu (with the reference kernel file's symbols) and
    the launch unit generateLaunchUpdateGrid.py prints compile together, with
    both launch paths of the check harness, for a 2D Read-aligned Variable
    kernel (host and gfx950 syntax)."""
    from threadcoarsenedstencilgenerator_amd.build import HIP_FLAGS, HIPCC, INCLUDE
    k = _gen("generate2DStencil.py", "Box", "Column", "Variable", "2", "2", "2", "Read", "8")
    u = _gen("generateLaunchUpdateGrid.py", "Box", "Variable", "2", "2D", "Read")
    assert k.returncode == 0 and u.returncode == 0
    (tmp_path / "kernels.cu").write_text(k.stdout)
    (tmp_path / "launchUpdateGrid.cu").write_text(u.stdout)
    r = subprocess.run([HIPCC, *HIP_FLAGS, "-Wall", "-fsyntax-only", f"-I{INCLUDE}", f"-I{tmp_path}",
                        os.path.join(PKG, "harness", "compiled_in_check.cpp")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
