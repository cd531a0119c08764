"""Builds the native parts in-tree: the runtime library (libtcsg.so), the
`bench` harness binary and generated kernel code objects (kernels/*.co).

Every HIP compile targets gfx950 only, with -ffp-contract=off (the stencil
arithmetic is written with explicit FMAs) and packed FP32 disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

from . import isa_check, spec
from .generator.emit import emit_source

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
KDIR = os.path.join(PKG, "kernels")
LIB = os.path.join(PKG, "libtcsg.so")
BENCH_BIN = os.path.join(PKG, "harness", "bench")
SHIM_LIB = os.path.join(PKG, "harness", "liblaunchUpdateGrid.so")
SHIM_DEMO = os.path.join(PKG, "harness", "shim_demo")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_runtime() -> str:
    srcs = [os.path.join(PKG, "csrc", f) for f in ("tcsg_runtime.hip", "tcsg_mlp.hip", "tcsg_hbm_probe.hip")]
    _run([HIPCC, *HIP_FLAGS, "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", *srcs, "-o", LIB])
    return LIB


PROBE_CO = os.path.join(KDIR, "buffer_probe.co")


def build_probe() -> str:
    """Code object of the buffer range-check probe the GPU tests load (a test fixture:
    tests/fixtures/buffer_probe.hip)."""
    os.makedirs(KDIR, exist_ok=True)
    src = os.path.join(os.path.dirname(PKG), "tests", "fixtures", "buffer_probe.hip")
    _run([HIPCC, *HIP_FLAGS, "--genco", "--no-gpu-bundle-output", src, "-o", PROBE_CO])
    return PROBE_CO


def kernel_path(desc: spec.StencilDesc) -> str:
    return os.path.join(KDIR, desc.key() + ".co")


def build_kernel(desc: spec.StencilDesc, force: bool = True) -> str:
    """Emit and compile one kernel; returns the code-object path.  The
    generated source is compiled from a temporary file; a .sha1 sidecar
    records which source the code object was built from."""
    import hashlib
    import tempfile
    os.makedirs(KDIR, exist_ok=True)
    co = kernel_path(desc)
    text = emit_source(desc)
    digest = hashlib.sha1(text.encode()).hexdigest()
    side = co[:-3] + ".sha1"
    if not force and os.path.exists(co) and os.path.exists(side) and open(side).read().strip() == digest:
        return co
    with tempfile.TemporaryDirectory(prefix="tcsg_") as td:
        src = os.path.join(td, desc.key() + ".hip")
        with open(src, "w") as f:
            f.write(text)
        tmp = co + f".{os.getpid()}.tmp"
        _run([HIPCC, *HIP_FLAGS, "--genco", "--no-gpu-bundle-output", "-x", "hip", src, "-o", tmp])
    try:
        isa_check.check_code_object(tmp)
    except RuntimeError:
        os.remove(tmp)
        raise
    os.replace(tmp, co)
    with open(side, "w") as f:
        f.write(digest + "\n")
    return co


def build_kernels(descs, force: bool = True, jobs: int | None = None):
    """Compile many kernels in parallel; returns {desc: path}."""
    descs = list(dict.fromkeys(descs))
    jobs = jobs or min(8, os.cpu_count() or 1)
    out = {}
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(build_kernel, d, force): d for d in descs}
        for f in cf.as_completed(futs):
            out[futs[f]] = f.result()
    return out


def prune_kernels(descs) -> int:
    """Delete code objects (and their .sha1 sidecars) in the kernel directory
    that belong to none of `descs`, the range-check probe kept; returns how
    many code objects went."""
    keep = {d.key() for d in descs} | {os.path.splitext(os.path.basename(PROBE_CO))[0]}
    gone = 0
    for f in os.listdir(KDIR):
        stem, ext = os.path.splitext(f)
        if ext in (".co", ".sha1") and stem not in keep:
            os.remove(os.path.join(KDIR, f))
            gone += ext == ".co"
    return gone


# Kernels built the reference's way: kernels.cu (the kernel generator's output, with the
# reference kernel file's symbols) compiled into the launch unit generateLaunchUpdateGrid.py
# prints, under harness/compiled_in/<name>/, with a check harness of our own that drives both
# the unit's launchUpdateGrid.h entry points and the kernel file's symbols (compiled_in_check).
COMPILED_IN_DIR = os.path.join(PKG, "harness", "compiled_in")
COMPILED_IN = {
    "star3d_r4_2x4x128_dppw": dict(dim=3, kernel=["Star", "Row", "Variable", "4", "2", "4", "128", "Write", "16",
                                                  "--xext=dppw", "--max-threads=256", "--prefetch=0"],
                                   launch=["Star", "Variable", "4", "3D", "Write"]),
    "box2d_r1_1x1_plain": dict(dim=2, kernel=["Box", "Row", "Jacobi", "1", "1", "1", "--variant=plain"],
                               launch=["Box", "Jacobi", "1", "2D"]),
    "box3d_r2_2x3x2_dpp_read": dict(dim=3, kernel=["Box", "Column", "Variable", "2", "2", "3", "2", "Read", "8",
                                                   "--xext=dpp"],
                                    launch=["Box", "Variable", "2", "3D", "Read"]),
}


def compiled_in_desc(name: str) -> spec.StencilDesc:
    from .generator import cli
    c = COMPILED_IN[name]
    return cli.desc_from_argv(c["dim"], c["kernel"])[0]


def build_compiled_in(jobs: int | None = None) -> dict:
    """Writes kernels.cu and launchUpdateGrid.cu for each COMPILED_IN entry
    and compiles compiled_in_check with them; returns {name: directory}."""
    from .generator import cli
    hd = os.path.join(PKG, "harness")

    def one(name):
        c = COMPILED_IN[name]
        d = os.path.join(COMPILED_IN_DIR, name)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "kernels.cu"), "w") as f:
            f.write(emit_source(compiled_in_desc(name), plugin=True))
        with open(os.path.join(d, "launchUpdateGrid.cu"), "w") as f:
            f.write(cli.launch_unit(c["launch"]))
        _run([HIPCC, *HIP_FLAGS, "-Wall", f"-I{INCLUDE}", f"-I{d}", os.path.join(hd, "compiled_in_check.cpp"),
              "-o", os.path.join(d, "compiled_in_check")])
        return d

    with cf.ThreadPoolExecutor(max_workers=jobs or min(4, os.cpu_count() or 1)) as ex:
        return dict(zip(COMPILED_IN, ex.map(one, COMPILED_IN)))


def build_harness() -> str:
    """The `bench` CLI (same argv as the reference's bench) and the
    launchUpdateGrid.h compatibility library."""
    hd = os.path.join(PKG, "harness")
    srcs = [os.path.join(hd, n) for n in ("main.cpp", "benchmark.cpp", "primes.cpp")]
    _run([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{hd}", *srcs,
          "-o", BENCH_BIN, f"-L{PKG}", "-ltcsg", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."])
    _run([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared", f"-I{INCLUDE}",
          os.path.join(hd, "launchUpdateGrid.cpp"), "-o", SHIM_LIB, f"-L{PKG}", "-ltcsg",
          "-Wl,-rpath,$ORIGIN/.."])
    _run([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", f"-I{INCLUDE}", os.path.join(hd, "shim_demo.cpp"),
          "-o", SHIM_DEMO, f"-L{hd}", "-llaunchUpdateGrid", f"-L{PKG}", "-ltcsg", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath,$ORIGIN/.."])
    return BENCH_BIN
