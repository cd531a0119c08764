"""The counter report's arithmetic on small rocprofv3-shaped fixtures (CPU): the
TSV join that appends counter columns to the harness's printResults rows
(hip/benchmark.cpp:352-426), matched to block shapes through the kernel trace of
the same counter run."""
import csv
import os

import pytest

from threadcoarsenedstencilgenerator_amd import counter_report as cr

TSV = """#GPU: test
#grid size excluding border: x 64 y 32 z 8
#coarsening factors: x 2 y 4 z 8
#alignment: None
#xthreads\tythreads\tzthreads\tPerf [Flop/s]\tPerf dev [Flop/s]\tPerf [Lup/s]\tPerf dev [Lup/s]\titerations\ttotal time[s]
64\t2\t1\t4.9e+12\t1.0e+10\t1.0e+11\t1.0e+08\t3\t1.0e-3
128\t1\t1\t4.9e+12\t1.0e+10\t2.0e+11\t1.0e+08\t3\t1.0e-3
32\t4\t1\t4.9e+12\t1.0e+10\t1.0e+11\t1.0e+08\t3\t1.0e-3
"""


def _write(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def _pass(d, counters):
    """A pass directory: dispatches 1-2 of block 64x2x1, 3 of 128x1x1, 4 another kernel."""
    os.makedirs(d)
    shapes = {"1": (64, 2, 1), "2": (64, 2, 1), "3": (128, 1, 1), "4": (64, 2, 1)}
    names = {"1": "tcsg_update", "2": "tcsg_update", "3": "tcsg_update", "4": "init_kernel"}
    _write(os.path.join(d, "x_kernel_trace.csv"),
           ["Dispatch_Id", "Kernel_Name", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z"],
           [{"Dispatch_Id": k, "Kernel_Name": names[k], "Workgroup_Size_X": s[0], "Workgroup_Size_Y": s[1],
             "Workgroup_Size_Z": s[2]} for k, s in shapes.items()])
    rows = []
    for k in shapes:
        for c, per in counters.items():
            # two counter instances per dispatch (the report sums them)
            rows += [{"Dispatch_Id": k, "Kernel_Name": names[k], "Counter_Name": c, "Counter_Value": per[k] / 2}] * 2
    _write(os.path.join(d, "x_counter_collection.csv"), ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           rows)


def test_tsv_counter_columns(tmp_path):
    upd = 64 * 32 * 8
    # per dispatch: 64x2x1 dispatches read 10 and 30 KiB (mean 20), 128x1x1 reads 40; init reads 999
    _pass(str(tmp_path / "p1"), {"FETCH_SIZE": {"1": 10.0, "2": 30.0, "3": 40.0, "4": 999.0},
                                 "WRITE_SIZE": {"1": 16.0, "2": 16.0, "3": 16.0, "4": 999.0}})
    _pass(str(tmp_path / "p2"), {"SQ_INSTS_VALU_FLOPS_FP64": {"1": upd / 64 * 49, "2": upd / 64 * 49,
                                                              "3": upd / 64 * 98, "4": 5.0},
                                 "TCP_TOTAL_CACHE_ACCESSES_sum": {"1": upd / 128 * 49, "2": upd / 128 * 49,
                                                                  "3": 0.0, "4": 5.0},
                                 "SQ_INSTS_LDS_LOAD_BANDWIDTH": {"1": upd / 64 * 49 * 2, "2": upd / 64 * 49 * 2,
                                                                 "3": 0.0, "4": 5.0},
                                 "SQ_INSTS_LDS_STORE_BANDWIDTH": {"1": 0.0, "2": 0.0, "3": 0.0, "4": 5.0}})
    k = cr.shape_counters([str(tmp_path / "p1"), str(tmp_path / "p2")])
    assert set(k) == {(64, 2, 1), (128, 1, 1)}
    out = cr.tsv_with_counters(TSV, k, fetch_scale=2.0, write_scale=1.0).splitlines()
    assert out[4].endswith("\t".join(("", ) + cr.TSV_COLUMNS))
    r64 = out[5].split("\t")
    assert r64[:9] == TSV.splitlines()[5].split("\t")
    hbm = (20 * 1024 * 2 + 16 * 1024) / upd
    assert float(r64[9]) == pytest.approx(hbm)
    assert float(r64[10]) == pytest.approx(hbm * 1e11 / 1e9)
    assert float(r64[11]) == pytest.approx(49.0)
    assert float(r64[12]) == pytest.approx(49.0 * 1e11)
    assert float(r64[13]) == pytest.approx(1.0)   # 49 L1 bytes per update against 49 flops
    assert float(r64[14]) == pytest.approx(2.0)   # 98 LDS bytes per update against 49 flops
    r128 = out[6].split("\t")
    assert float(r128[9]) == pytest.approx((40 * 1024 * 2 + 16 * 1024) / upd)
    assert float(r128[11]) == pytest.approx(98.0)
    assert float(r128[13]) == 0.0
    # a shape the passes did not see keeps its row as it was
    assert out[7] == TSV.splitlines()[7]


def test_tsv_prefers_the_event_rate(tmp_path):
    """With TCSG_EXTRA_COLUMNS the rates come from the HIP-event column."""
    _pass(str(tmp_path / "p1"), {"FETCH_SIZE": {"1": 8.0, "2": 8.0, "3": 8.0, "4": 0.0},
                                 "WRITE_SIZE": {"1": 8.0, "2": 8.0, "3": 8.0, "4": 0.0}})
    lines = TSV.splitlines()
    lines[4] += "\tevent Perf [Lup/s]\tevent [GB/s]"
    lines[5] += "\t3.0e+11\t4.8e+03"
    out = cr.tsv_with_counters("\n".join(lines[:6]), cr.shape_counters([str(tmp_path / "p1")]), 2.0, 1.0)
    r = out.splitlines()[5].split("\t")
    hbm = (8 * 1024 * 2 + 8 * 1024) / (64 * 32 * 8)
    assert float(r[11]) == pytest.approx(hbm) and float(r[12]) == pytest.approx(hbm * 3.0e11 / 1e9)


def test_tsv_needs_the_grid_line():
    with pytest.raises(ValueError):
        cr.tsv_with_counters("#xthreads\ty\n64\t1\t1\t0\t0\t1\t0\t1\t1\n", {}, 2.0, 1.0)
