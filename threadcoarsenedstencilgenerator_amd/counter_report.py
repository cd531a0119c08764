"""Counter report of the measured workloads from rocprofv3 counter passes.

Input: one directory per `rocprofv3 --pmc` pass over
``python -m threadcoarsenedstencilgenerator_amd.profile_run --workload all``
(each holding rocprofv3's ``*counter_collection.csv``), the dispatch table that
run writes (``--meta``), and a kernel trace of the same run
(``*kernel_trace.csv``) for the dispatch durations.  Dispatches are matched to
workloads by kernel symbol, grid and workgroup size; each counter is averaged
over a workload's dispatches.  FETCH_SIZE counts what L2 fetches from the data fabric,
which the 256 MB Infinity Cache in front of HBM may serve: for the workloads here (1-9 GB)
it is taken as HBM traffic.

Output (markdown table and JSON), per workload and per time step:
FP64 flops (measured, ``SQ_INSTS_VALU_FLOPS_FP64``) against the reference's
count, FP64 TFLOP/s against the 78.6 TFLOP/s vector peak, L1 (TCP) bytes per
FLOP from the tag accesses (128-byte lines), LDS bytes per FLOP
(``SQ_INSTS_LDS_*_BANDWIDTH``, 64-byte units), HBM bytes from ``FETCH_SIZE`` and
``WRITE_SIZE`` (each scaled by what it reads for the 2^30-byte calibration copy)
and the achieved HBM GB/s against 8 TB/s, plus the issue and memory-pipeline
figures that tell latency-, VALU- and TA-bound kernels apart.

    python -m threadcoarsenedstencilgenerator_amd.counter_report --meta M.json --trace TRACE_DIR \
        PASS_DIR [PASS_DIR ...] [--md OUT.md] [--json OUT.json]

With --tsv it instead appends counter columns to the `bench` harness's TSV (the reference's
printResults rows): the passes are `rocprofv3 --pmc ... --kernel-trace` runs of the same
harness command, whose kernel traces give each dispatch's block shape.

    python -m threadcoarsenedstencilgenerator_amd.counter_report --tsv BENCH.tsv PASS_DIR [PASS_DIR ...]
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
from collections import defaultdict

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
L1_LINE = 128       # bytes per TCP tag access (gfx950 vector L1 line)
LDS_UNIT = 64       # SQ_INSTS_LDS_*_BANDWIDTH count 64-byte units
COPY_BYTES = 1 << 30
WAVE = 64


def _kernel_map(rows, meta):
    """{Kernel_Id: workload}: the generated kernels all carry the symbol tcsg_update, so
    the k-th code object of a symbol to be dispatched (the run loads them in the order
    of the dispatch table) is the k-th workload of that symbol; the copy is matched by
    its symbol."""
    first = {}
    for row in rows:
        kid, did = row["Kernel_Id"], int(row["Dispatch_Id"])
        if kid not in first or did < first[kid][0]:
            first[kid] = (did, row["Kernel_Name"])
    by_sym = defaultdict(list)
    for kid, (did, name) in sorted(first.items(), key=lambda t: t[1][0]):
        by_sym[name].append(kid)
    out = {}
    for sym in {m["symbol"] for m in meta}:
        wls = [m["name"] for m in meta if m["symbol"] == sym]
        kids = by_sym.get(sym, [])
        if sym.startswith("__amd_rocclr_copyBuffer"):
            for kid in kids:
                out[kid] = wls[0]
            continue
        if len(kids) != len(wls):
            raise ValueError(f"{len(kids)} code objects dispatch {sym}, the dispatch table lists {len(wls)}")
        out.update(zip(kids, wls))
    return out


def _rows(d, pattern):
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not files:
        raise FileNotFoundError(f"no {pattern} under {d}")
    rows = []
    for f in files:
        rows += list(csv.DictReader(open(f)))
    return rows


def read_pass(d, meta):
    """{workload: {counter: mean value per dispatch}} of one pass directory."""
    rows = _rows(d, "*counter_collection.csv")
    kmap = _kernel_map(rows, meta)
    per = defaultdict(lambda: defaultdict(float))   # (workload, dispatch) -> counter -> value
    for row in rows:
        wl = kmap.get(row["Kernel_Id"])
        if wl is not None:
            per[(wl, row["Dispatch_Id"])][row["Counter_Name"]] += float(row["Counter_Value"])
    out = defaultdict(lambda: defaultdict(list))
    for (wl, _), cs in per.items():
        for c, v in cs.items():
            out[wl][c].append(v)
    # copyBuffer also serves the runtime's small copies: the calibration copy is the largest
    return {wl: {c: (max(v) if wl == "copy_2^30B" else sum(v) / len(v)) for c, v in cs.items()}
            for wl, cs in out.items()}


def read_trace(d, meta):
    """{workload: mean dispatch duration in ns} from a kernel trace of the same run."""
    rows = _rows(d, "*kernel_trace.csv")
    kmap = _kernel_map(rows, meta)
    dur = defaultdict(list)
    for row in rows:
        wl = kmap.get(row["Kernel_Id"])
        if wl is not None and wl != "copy_2^30B":
            dur[wl].append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
    return {wl: sum(v) / len(v) for wl, v in dur.items()}


def report(meta, passes, trace):
    c = defaultdict(dict)
    for p in passes:
        for wl, cs in p.items():
            c[wl].update(cs)
    copy = c.get("copy_2^30B", {})
    fscale = COPY_BYTES / (copy["FETCH_SIZE"] * 1024) if copy.get("FETCH_SIZE") else None
    wscale = COPY_BYTES / (copy["WRITE_SIZE"] * 1024) if copy.get("WRITE_SIZE") else None
    rows = []
    for m in meta:
        if m["name"] == "copy_2^30B" or m["name"] not in c:
            continue
        k = c[m["name"]]
        ns = trace.get(m["name"])
        upd = m["interior_points"] * m["steps_per_launch"]   # lattice updates per dispatch
        g = lambda n: k.get(n)
        r = dict(workload=m["name"], kernel=m["kernel"], dispatch_ns=ns, updates_per_dispatch=upd,
                 flops_per_update_reference=m["flops_per_update"])
        # the FLOPS counters count per wave instruction (FMA 2, add/mul 1): times 64 lanes
        fl = g("SQ_INSTS_VALU_FLOPS_FP64") and g("SQ_INSTS_VALU_FLOPS_FP64") * WAVE
        if fl:
            r["fp64_flops_per_update"] = fl / upd
            if ns:
                r["fp64_tflops"] = fl / ns / 1e3
                r["fp64_frac_of_peak"] = r["fp64_tflops"] / FP64_PEAK_TFLOPS
        fl32 = g("SQ_INSTS_VALU_FLOPS_FP32") and g("SQ_INSTS_VALU_FLOPS_FP32") * WAVE
        if fl32:
            r["fp32_flops_per_update"] = fl32 / upd
        flops = fl if fl else fl32
        if g("TCP_TOTAL_CACHE_ACCESSES_sum"):
            l1 = g("TCP_TOTAL_CACHE_ACCESSES_sum") * L1_LINE
            r["l1_bytes_per_update"] = l1 / upd
            if flops:
                r["l1_bytes_per_flop"] = l1 / flops
        lds = (g("SQ_INSTS_LDS_LOAD_BANDWIDTH") or 0) + (g("SQ_INSTS_LDS_STORE_BANDWIDTH") or 0)
        if g("SQ_INSTS_LDS_LOAD_BANDWIDTH") is not None:
            r["lds_bytes_per_update"] = lds * LDS_UNIT / upd
            if flops:
                r["lds_bytes_per_flop"] = lds * LDS_UNIT / flops
        if fscale and wscale and g("FETCH_SIZE") is not None and g("WRITE_SIZE") is not None:
            hbm = g("FETCH_SIZE") * 1024 * fscale + g("WRITE_SIZE") * 1024 * wscale
            r["hbm_bytes_per_dispatch"] = hbm
            r["hbm_bytes_per_update"] = hbm / upd
            r["hbm_vs_algorithmic"] = hbm / (m["bytes_per_update"] * m["interior_points"])
            if ns:
                r["hbm_gbs"] = hbm / ns
                r["hbm_frac_of_peak"] = r["hbm_gbs"] / HBM_PEAK_GBS
        if g("SQ_WAVE_CYCLES"):
            r["wait_inst_frac"] = (g("SQ_WAIT_INST_ANY") or 0) / g("SQ_WAVE_CYCLES")
            r["valu_active_frac_of_wave_time"] = (g("SQ_ACTIVE_INST_VALU") or 0) / g("SQ_WAVE_CYCLES")
            r["vmem_active_frac_of_wave_time"] = (g("SQ_ACTIVE_INST_VMEM") or 0) / g("SQ_WAVE_CYCLES")
        for n in ("SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_INSTS_LDS", "SQ_INSTS_SALU"):
            if g(n) is not None:
                r[n.lower() + "_per_update"] = g(n) / upd
        if g("TCC_HIT_sum") is not None and g("TCC_MISS_sum"):
            r["l2_hit_rate"] = g("TCC_HIT_sum") / (g("TCC_HIT_sum") + g("TCC_MISS_sum"))
        if g("GRBM_GUI_ACTIVE") and g("TA_TA_BUSY_sum") is not None:
            # TA_TA_BUSY summed over the 256 CUs' TA instances; GRBM_GUI_ACTIVE over the 8 XCDs
            r["ta_busy_frac"] = g("TA_TA_BUSY_sum") / 256 / (g("GRBM_GUI_ACTIVE") / 8)
        if g("GRBM_GUI_ACTIVE") and g("TD_TD_BUSY_sum") is not None:
            r["td_busy_frac"] = g("TD_TD_BUSY_sum") / 256 / (g("GRBM_GUI_ACTIVE") / 8)
        cyc = g("GRBM_GUI_ACTIVE") and g("GRBM_GUI_ACTIVE") / 8   # cycles of one XCD's clock domain
        if cyc and g("TCP_PENDING_STALL_CYCLES_sum") is not None:
            # per-CU fractions of the dispatch's cycles (256 CUs)
            r["l1_waiting_on_l2_frac"] = g("TCP_PENDING_STALL_CYCLES_sum") / 256 / cyc
        if cyc and g("TA_DATA_STALLED_BY_TC_CYCLES_sum") is not None:
            r["ta_data_stalled_frac"] = g("TA_DATA_STALLED_BY_TC_CYCLES_sum") / 256 / cyc
        if g("TCP_TCC_READ_REQ_sum"):
            r["l2_read_requests_per_update"] = g("TCP_TCC_READ_REQ_sum") / upd
            if g("TCP_TCC_READ_REQ_LATENCY_sum"):
                lat = g("TCP_TCC_READ_REQ_LATENCY_sum") / g("TCP_TCC_READ_REQ_sum")
                r["l2_read_latency_cycles"] = lat
                if cyc:
                    # Little's law: L2 reads in flight per CU
                    r["l2_reads_in_flight_per_cu"] = g("TCP_TCC_READ_REQ_sum") / 256 / cyc * lat
        rows.append(r)
    return dict(fetch_scale=fscale, write_scale=wscale, rows=rows)


def _f(v, spec):
    return "" if v is None else format(v, spec)


def markdown(rep) -> str:
    cols = [("workload", "workload", "s"), ("dispatch_ns", "dispatch ns", ",.0f"),
            ("fp64_flops_per_update", "FP64 flop/update (ref.)", ".1f"),
            ("fp64_tflops", "FP64 TFLOP/s", ".1f"), ("fp64_frac_of_peak", "of 78.6", ".3f"),
            ("l1_bytes_per_flop", "L1 B/flop", ".2f"), ("lds_bytes_per_flop", "LDS B/flop", ".2f"),
            ("hbm_bytes_per_update", "HBM B/update", ".2f"), ("hbm_vs_algorithmic", "HBM / alg.", ".3f"),
            ("hbm_gbs", "HBM GB/s", ",.0f"), ("hbm_frac_of_peak", "of 8 TB/s", ".3f")]
    out = ["| " + " | ".join(h for _, h, _ in cols) + " |", "|" + "---|" * len(cols)]
    for r in rep["rows"]:
        cells = []
        for key, _, sp in cols:
            v = r.get(key)
            if key == "fp64_flops_per_update" and v is not None:
                cells.append(f"{v:.1f} ({r['flops_per_update_reference']})")
            elif key == "workload":
                cells.append(r["workload"])
            else:
                cells.append(_f(v, sp))
        out.append("| " + " | ".join(cells) + " |")
    cols2 = [("workload", "workload", "s"), ("sq_insts_valu_per_update", "VALU inst/update", ".3f"),
             ("sq_insts_vmem_rd_per_update", "VMEM rd/update", ".4f"), ("l1_bytes_per_update", "L1 B/update", ".1f"),
             ("lds_bytes_per_update", "LDS B/update", ".1f"), ("l2_hit_rate", "L2 hit", ".3f"),
             ("wait_inst_frac", "wave time waiting", ".3f"), ("valu_active_frac_of_wave_time", "VALU active", ".3f"),
             ("vmem_active_frac_of_wave_time", "VMEM active", ".3f"), ("ta_busy_frac", "TA busy", ".3f"),
             ("td_busy_frac", "TD busy", ".3f"), ("ta_data_stalled_frac", "TA stalled on L1", ".3f"),
             ("l1_waiting_on_l2_frac", "L1 waiting on L2", ".3f"),
             ("l2_read_requests_per_update", "L2 reads/update", ".3f"),
             ("l2_read_latency_cycles", "L2 read latency [cyc]", ".0f"),
             ("l2_reads_in_flight_per_cu", "L2 reads in flight/CU", ".0f")]
    out += ["", "| " + " | ".join(h for _, h, _ in cols2) + " |", "|" + "---|" * len(cols2)]
    for r in rep["rows"]:
        out.append("| " + " | ".join(r["workload"] if k == "workload" else _f(r.get(k), sp)
                                     for k, _, sp in cols2) + " |")
    return "\n".join(out) + "\n"


TSV_COLUMNS = ("HBM [B/Lup]", "HBM [GB/s]", "FP64 [Flop/Lup] counted", "FP64 counted [Flop/s]", "L1 [B/Flop]",
               "LDS [B/Flop]")


def shape_counters(pass_dirs, symbol: str = "tcsg_update"):
    """{(bx, by, bz): {counter: mean per dispatch}} over the dispatches of `symbol` in counter
    passes each recorded with a kernel trace of the same run (`rocprofv3 --pmc ... --kernel-trace`):
    the trace gives every dispatch's block shape, the counter rows only its thread count."""
    out = defaultdict(lambda: defaultdict(list))
    for d in pass_dirs:
        shape = {}
        for row in _rows(d, "*kernel_trace.csv"):
            if row["Kernel_Name"] == symbol:
                shape[row["Dispatch_Id"]] = (int(row["Workgroup_Size_X"]), int(row["Workgroup_Size_Y"]),
                                             int(row["Workgroup_Size_Z"]))
        per = defaultdict(lambda: defaultdict(float))
        for row in _rows(d, "*counter_collection.csv"):
            if row["Kernel_Name"] == symbol and row["Dispatch_Id"] in shape:
                per[row["Dispatch_Id"]][row["Counter_Name"]] += float(row["Counter_Value"])
        for did, cs in per.items():
            for c, v in cs.items():
                out[shape[did]][c].append(v)
    return {b: {c: sum(v) / len(v) for c, v in cs.items()} for b, cs in out.items()}


def tsv_with_counters(tsv_text: str, counters, fetch_scale: float, write_scale: float,
                      steps_per_launch: int = 1) -> str:
    """The harness's TSV (the reference's printResults layout, hip/benchmark.cpp:352-426) with
    counter columns appended to every row whose block shape the passes measured: HBM bytes per
    lattice update (FETCH_SIZE and WRITE_SIZE in KiB, scaled by what they read for a copy of known
    size) and the GB/s they make at the row's Lup/s (the HIP-event column where the TSV has one,
    else the mean of the reference's protocol), FP64 flops per update as counted
    (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes) and the Flop/s they make, and L1 (TCP tag accesses x
    128 B) and LDS (SQ_INSTS_LDS_*_BANDWIDTH x 64 B) bytes per counted flop; '-' where a counter
    is missing.  Rows of shapes the passes did not see are left as they are."""
    lines = tsv_text.rstrip("\n").split("\n")
    pts = None
    for ln in lines:
        if ln.startswith("#grid size excluding border:"):
            v = ln.split(":", 1)[1].split()
            dims = [int(v[i + 1]) for i in range(0, len(v), 2)]
            pts = 1
            for x in dims:
                pts *= x
    if pts is None:
        raise ValueError("no '#grid size excluding border' line in the TSV")
    upd = pts * steps_per_launch
    out = []
    rate_col = 5   # Perf [Lup/s]; the HIP-event rate instead where the TSV has it (TCSG_EXTRA_COLUMNS)
    for ln in lines:
        if ln.startswith("#xthreads"):
            cols = ln.lstrip("#").split("\t")
            if "event Perf [Lup/s]" in cols:
                rate_col = cols.index("event Perf [Lup/s]")
            out.append(ln + "".join("\t" + c for c in TSV_COLUMNS))
            continue
        if ln.startswith("#") or not ln.strip():
            out.append(ln)
            continue
        f = ln.split("\t")
        b = (int(f[0]), int(f[1]), max(1, int(f[2])))
        k = counters.get(b)
        if not k:
            out.append(ln)
            continue
        lups = float(f[rate_col])
        g = k.get
        fmt = lambda v: "-" if v is None else f"{v:.6e}"
        hbm = (g("FETCH_SIZE") * 1024 * fetch_scale + g("WRITE_SIZE") * 1024 * write_scale) / upd \
            if g("FETCH_SIZE") is not None and g("WRITE_SIZE") is not None else None
        fl = g("SQ_INSTS_VALU_FLOPS_FP64") * WAVE / upd if g("SQ_INSTS_VALU_FLOPS_FP64") else None
        l1 = g("TCP_TOTAL_CACHE_ACCESSES_sum") * L1_LINE / upd if g("TCP_TOTAL_CACHE_ACCESSES_sum") is not None else None
        lds = ((g("SQ_INSTS_LDS_LOAD_BANDWIDTH") or 0) + (g("SQ_INSTS_LDS_STORE_BANDWIDTH") or 0)) * LDS_UNIT / upd \
            if g("SQ_INSTS_LDS_LOAD_BANDWIDTH") is not None else None
        vals = (hbm, hbm * lups / 1e9 if hbm is not None else None, fl, fl * lups if fl else None,
                l1 / fl if l1 is not None and fl else None, lds / fl if lds is not None and fl else None)
        out.append(ln + "".join("\t" + fmt(v) for v in vals))
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("passes", nargs="+")
    ap.add_argument("--tsv", help="instead: append counter columns to this harness TSV; the passes are "
                                  "`rocprofv3 --pmc ... --kernel-trace` runs of the same harness command")
    ap.add_argument("--fetch-scale", type=float, default=2.0, help="--tsv: bytes per FETCH_SIZE KiB unit, "
                    "as the 2^30-byte calibration copy reads it")
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--steps-per-launch", type=int, default=1)
    ap.add_argument("--meta")
    ap.add_argument("--trace")
    ap.add_argument("--md")
    ap.add_argument("--json")
    ap.add_argument("--traffic", help="update this JSON (kernel key -> HBM bytes per launch, as bench.py "
                                      "reads it) with the workloads' FETCH_SIZE/WRITE_SIZE traffic")
    a = ap.parse_args()
    if a.tsv:
        text = tsv_with_counters(open(a.tsv).read(), shape_counters(a.passes), a.fetch_scale, a.write_scale,
                                 a.steps_per_launch)
        print(text, end="")
        if a.md:
            open(a.md, "w").write(text)
        return
    if not a.meta or not a.trace:
        ap.error("--meta and --trace are required without --tsv")
    meta = json.load(open(a.meta))
    rep = report(meta, [read_pass(p, meta) for p in a.passes], read_trace(a.trace, meta))
    md = markdown(rep)
    print(md)
    if a.md:
        open(a.md, "w").write(md)
    if a.json:
        json.dump(rep, open(a.json, "w"), indent=1)
    if a.traffic:
        update_traffic(a.traffic, meta, rep)


def update_traffic(path, meta, rep):
    t = json.load(open(path)) if os.path.exists(path) else {}
    by_name = {m["name"]: m for m in meta}
    for r in rep["rows"]:
        if "hbm_bytes_per_dispatch" not in r:
            continue
        m = by_name[r["workload"]]
        t[m["kernel"]] = {"block": m["workgroup"], "hbm_bytes_per_launch": int(r["hbm_bytes_per_dispatch"]),
                          "fetch_scale": rep["fetch_scale"], "write_scale": rep["write_scale"],
                          "algorithmic_bytes_per_launch": m["bytes_per_update"] * m["interior_points"],
                          "source": "profiles/counters_measured_workloads.md", "workload": m["name"]}
    with open(path, "w") as f:
        json.dump(t, f, indent=1)


if __name__ == "__main__":
    main()
