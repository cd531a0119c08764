"""Command lines of the generator, with the reference's positional argv.

    generate2DStencil.py  Box|Star Row|Column Jacobi|Variable range xC yC [None|Read|Write CLsize]
    generate3DStencil.py  Box|Star Row|Column Jacobi|Variable range xC yC zC [None|Read|Write CLsize]
    generateLaunchUpdateGrid.py  Box|Star Jacobi|Variable range 2D|3D [None|Read|Write] [--shim]

(hip/generate2DStencil_hip.py:28-60, hip/generate3DStencil_hip.py:29-63,
hip/generateLaunchUpdateGrid.py:24-45.)  The kernel generators print gfx950
HIP source on stdout, as the reference prints kernels.cu; `-o FILE.co`
compiles it to a code object instead.  Options that select the MI355X code
shape (no reference counterpart) are given as --name=value:
--variant=plain|stream|temporal, --xext=load|dpp|dppw, --dtype=f64|f32c,
--tsteps=N (temporal variant, 2D: N time steps per launch through LDS),
--max-threads=N, --prefetch=N, --waves-per-eu=N, --comp-group=N,
--order=point|offset (plain variant; 2D stream kernels without --prefetch), --no-xcd-remap,
--no-nt-store, --zfirst,
--remainder (grids whose interior is not a multiple of the coarsening are
updated whole instead of truncated), --edge-pairs (--xext=dppw: the wave-edge
lanes fetch two adjacent rows' extra columns per load), --stage=BXxBYxBZ (3D,
--xext=dppw: LDS halo staging for blocks of BX x BY x BZ threads, the only
block shape such a kernel is then launched with; with --remainder the grid's
interior must hold one block's tile in y and z; its source has no reference
updateGrid, which takes any block shape).
Invalid arguments print the usage on stderr and exit with status 1, as the
reference does.
"""
from __future__ import annotations

import sys

from .. import spec

USAGE2 = ("generate2DStencil.py usage:\n\tstencil neighbourhood\t\t'Box' or 'Star'\n\taccess orientation\t'Row' or "
          "'Column'\n\tcoefficients\t'Jacobi' or 'Variable'\n\trange\n\tcoarsening-factor in x\n\tcoarsening-factor in "
          "y\nx-/y coarsening-factors[>0] are elements processed by a single thread\n\n\toptional:\n\talignment\t\t"
          "'Write' or 'Read', default is 'None'\n\tsize of alignment (in units of grid-points)\n")
USAGE3 = USAGE2.replace("generate2DStencil.py", "generate3DStencil.py").replace(
    "\tcoarsening-factor in y\n", "\tcoarsening-factor in y\n\tcoarsening-factor in z\n")
USAGE_L = ("generateLaunchUpdateGrid.py usage:\n\tstencil neighbourhood\t\teither 'Box' or 'Star'\n\tcoefficients\t\t"
           "'Jacobi' or 'Variable'(coefficients have to be set by host)\n\trange\n\tdimension\n\n\toptional:\n"
           "\talignment\t\t'Write' or 'Read'\n")


def _err(msg, usage):
    print(msg + "\n" + usage, file=sys.stderr)
    sys.exit(1)


def _split(argv):
    """Positional arguments, --name=value options and the -o output path."""
    opts = {}
    out = None
    i = 0
    pos = []
    while i < len(argv):
        a = argv[i]
        if a == "-o":
            out = argv[i + 1]
            i += 2
            continue
        if a.startswith("--"):
            k, _, v = a[2:].partition("=")
            opts[k] = v
        else:
            pos.append(a)
        i += 1
    return pos, opts, out


def _int(s, what, usage):
    try:
        return int(s)
    except ValueError:
        _err(f"{what} must be an integer", usage)


def desc_from_argv(dim: int, argv):
    usage = USAGE2 if dim == 2 else USAGE3
    pos, opts, out = _split(argv)
    nfix = 6 if dim == 2 else 7
    if len(pos) not in (nfix, nfix + 2):
        _err("Unknown arguments!", usage)
    nb, orient, coeff = pos[0], pos[1], pos[2]
    r = _int(pos[3], "range", usage)
    c = [_int(v, "coarsening factor", usage) for v in pos[4:4 + dim - 1 + 1]]
    align, cl = "None", 0
    if len(pos) == nfix + 2:
        align, cl = pos[nfix], _int(pos[nfix + 1], "size of alignment", usage)
        if cl <= 0 and align != "None":
            _err("size of alignment has to be > 0", usage)
    if nb not in spec.NB_NAMES:
        _err("Unknown stencil neighbourhood!", usage)
    if orient not in spec.ORIENT_NAMES:
        _err("Unknown access orientation!", usage)
    if coeff not in spec.COEFF_NAMES:
        _err("Unknown coefficient type!", usage)
    if r < 1:
        _err("Range has to be at least 1!", usage)
    if min(c) < 1:
        _err("x- or y-coarsening factors have to be at least 1!", usage)
    if align not in spec.ALIGN_NAMES:
        _err("Unkown alignment! Default is 'None'.", usage)
    kw = dict(dimension=dim, neighbourhood=spec.NB_NAMES[nb], orientation=spec.ORIENT_NAMES[orient],
              coeff=spec.COEFF_NAMES[coeff], range=r, xc=c[0], yc=c[1], zc=c[2] if dim == 3 else 1,
              alignment=spec.ALIGN_NAMES[align], cl_size=cl if align != "None" else 0,
              variant=spec.VARIANT_NAMES[opts.get("variant", "stream")],
              dtype=spec.DTYPE_NAMES[opts.get("dtype", "f64")],
              xext=opts.get("xext", "dpp"), xcd_remap="no-xcd-remap" not in opts, nt_store="no-nt-store" not in opts,
              zfirst="zfirst" in opts, remainder="remainder" in opts, edge_pairs="edge-pairs" in opts)
    for k, f in (("max-threads", "max_threads"), ("prefetch", "prefetch"), ("waves-per-eu", "waves_per_eu"),
                 ("comp-group", "comp_group"), ("tsteps", "tsteps")):
        if k in opts:
            kw[f] = _int(opts[k], k, usage)
    if "stage" in opts:
        parts = opts["stage"].split("x")
        if len(parts) != 3:
            _err("--stage takes a block shape BXxBYxBZ", usage)
        kw["stage"] = tuple(_int(v, "stage", usage) for v in parts)
    if kw["variant"] == spec.PLAIN:
        kw.update(xext="load", xcd_remap=False, nt_store=False, zfirst=False, order=opts.get("order", "point"))
    elif kw["variant"] == spec.TEMPORAL:
        kw.update(xext="load", xcd_remap=False, nt_store=False)
        if "order" in opts:
            _err("--order applies to the plain variant", usage)
    elif "order" in opts:
        if dim != 2:
            _err("--order applies to the plain variant and to 2D stream kernels", usage)
        kw["order"] = opts["order"]
        if opts["order"] == "offset" and "prefetch" not in opts:
            kw["prefetch"] = 0
    try:
        return spec.StencilDesc(**kw), out
    except ValueError as e:
        _err(str(e), usage)


def generate(dim: int, argv=None) -> int:
    from .emit import emit_source
    d, out = desc_from_argv(dim, sys.argv[1:] if argv is None else argv)
    if out:
        from ..build import HIP_FLAGS, HIPCC, _run
        import os
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            src = os.path.join(td, "kernels.hip")
            open(src, "w").write(emit_source(d))
            _run([HIPCC, *HIP_FLAGS, "--genco", "--no-gpu-bundle-output", src, "-o", out])
        from .. import isa_check
        isa_check.check_code_object(out)
    else:
        # the printed kernels.cu also carries the reference kernel file's symbols (FP64)
        sys.stdout.write(emit_source(d, plugin=d.dtype == spec.F64 and d.variant != spec.TEMPORAL and not d.staged))
    return 0


def generate_launch(argv=None) -> int:
    """Prints launchUpdateGrid.cu as the reference does
    (hip/generateLaunchUpdateGrid.py): the four launchUpdateGrid.h entry
    points over the generated kernels.cu compiled into the same unit, with
    the expected parameters fixed, so the reference's Makefile builds its
    harness unchanged.  With --shim it prints instead a C++ unit that builds
    those entry points over libtcsg, for kernels loaded as code objects at
    run time."""
    sys.stdout.write(launch_unit(sys.argv[1:] if argv is None else argv))
    return 0


def launch_unit(argv) -> str:
    """Text of the launch unit for the reference's generateLaunchUpdateGrid.py
    argv (plus --shim); invalid arguments exit with status 1 and the usage."""
    shim = "--shim" in argv
    argv = [a for a in argv if a != "--shim"]
    if len(argv) not in (4, 5):
        _err("Unknown arguments!", USAGE_L)
    nb, coeff, r, dim = argv[0], argv[1], _int(argv[2], "range", USAGE_L), argv[3]
    align = argv[4] if len(argv) == 5 else "None"
    if nb not in spec.NB_NAMES:
        _err("Unknown stencil neighbourhood!", USAGE_L)
    if coeff not in spec.COEFF_NAMES:
        _err("Unknown coefficient-type!", USAGE_L)
    if r < 1:
        _err("Range has to be at least 1!", USAGE_L)
    if dim not in ("2D", "3D"):
        _err("Unknown dimension!", USAGE_L)
    if align not in spec.ALIGN_NAMES:
        _err("Unknown alignment!", USAGE_L)
    import os
    hdir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "harness")
    if not shim:
        with open(os.path.join(hdir, "compiled_in_launch.cpp.in")) as f:
            text = f.read()
        subst = {"@NB@": nb, "@COEFF@": coeff, "@RANGE@": str(r), "@DIM@": dim[0], "@ALIGN@": align,
                 "@NB_CODE@": str(spec.NB_NAMES[nb]), "@COEFF_CODE@": str(spec.COEFF_NAMES[coeff]),
                 "@ALIGN_CODE@": str(spec.ALIGN_NAMES[align])}
        for k, v in subst.items():
            text = text.replace(k, v)
        return text
    shim = os.path.join(hdir, "launchUpdateGrid.cpp")
    return "\n".join(["// Generated by threadcoarsenedstencilgenerator_amd.generator (launch interface).",
                      f'#define TCSG_EXPECT_NEIGHBOURHOOD "{nb}"',
                      f'#define TCSG_EXPECT_COEFF "{coeff}"',
                      f"#define TCSG_EXPECT_RANGE {r}",
                      f"#define TCSG_EXPECT_DIMENSION {dim[0]}",
                      f'#define TCSG_EXPECT_ALIGNMENT "{align}"',
                      f'#include "{shim}"', ""])
