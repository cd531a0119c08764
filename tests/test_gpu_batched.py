"""Batched uneven grids on the GPU: the persistent kernel (with and without
a captured hipGraph of the step loop) and the separate-launch baseline give
every grid exactly the oracle's result, also with remainder handling on grid
sizes that are not rounded to the coarsening."""
import numpy as np
import pytest

import oracle
from threadcoarsenedstencilgenerator_amd import manifest, spec
from threadcoarsenedstencilgenerator_amd.spec import F32C, F64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [F64, F32C])
@pytest.mark.parametrize("mode", ["graph", "launches", "separate"])
@pytest.mark.parametrize("shape,block,workgroups", [("dppw", (64, 2, 1), 7), ("plain", (32, 4, 2), 7),
                                                    ("plain", (32, 4, 2), 0), ("plain", (32, 4, 2), -1),
                                                    ("plain+remainder", (32, 4, 2), 0),
                                                    ("dppw+remainder", (64, 2, 1), 7),
                                                    ("stream+remainder+pack", (32, 4, 1), 7),
                                                    ("stream+remainder+pack", (32, 4, 1), 0),
                                                    ("stream+remainder+pack", (16, 8, 2), -1),
                                                    ("plain+remainder+pack", (32, 4, 2), 0)])
def test_batched_matches_oracle(dtype, mode, shape, block, workgroups):
    """workgroups: 7 (fewer than 8: every workgroup strides the whole table),
    0 (as many per CU as fit: the per-XCD ranges), -1 (one per tile).  pack:
    tiles are runs of block x * block y consecutive threads of a grid's xy
    thread space (x fastest, continuing on the next thread row)."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    shape, *flags = shape.split("+")
    rem = "remainder" in flags
    d = manifest.cfg4_desc(shape=shape, batched=True, dtype=dtype, max_threads=1024 if shape == "plain" else 256)
    d = d.replace(remainder=rem, pack="pack" in flags)
    k = rt.Kernel(d)
    # remainder kernels take any size: interiors not rounded to the coarsening
    sizes = rt.batch_sizes(6, 3, 8, 40, (1, 1, 1) if rem else d.coarsening, d.range)
    print(f"batched {mode} {d.key()} sizes {sizes}", flush=True)
    geoms, grids, init = [], [], []
    for i, size in enumerate(sizes):
        g = k.geometry(size)
        gr = rt.Grids(g, k.dtype)
        for j in range(2):
            k.init(g, gr.ptr(j), rt.INIT_U01, 20 + i)
        geoms.append(g)
        grids.append(gr)
    torch.cuda.synchronize()
    init = [gr.host(0) for gr in grids]
    if workgroups < 0 and d.pack:
        per = block[0] * block[1]
        workgroups = sum(-(-g.threads.x * g.threads.y // per) * -(-g.threads.z // block[2]) for g in geoms)
    elif workgroups < 0:
        workgroups = sum(int(np.prod(rt.grid_blocks(g, block))) for g in geoms)
    b = rt.Batch(k, geoms, [gr.ptr(0) for gr in grids], [gr.ptr(1) for gr in grids], block, workgroups=workgroups)
    steps = 5
    if mode == "separate":
        fin = b.run_separate(steps)
    else:
        fin = b.run(steps, graph=(mode == "graph"))
        if mode == "graph":  # replay the cached graph from the restored state
            for gr, a in zip(grids, init):
                gr.upload(0, a)
                gr.upload(1, a)
            torch.cuda.synchronize()
            fin = b.run(steps, graph=True)
    torch.cuda.synchronize()
    run = oracle.run_f32c if dtype == F32C else oracle.run_f64
    for size, gr, a in zip(sizes, grids, init):
        og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, (1, 1, 1) if rem else d.coarsening)
        ref = run(d, og, None, a.copy(), a.copy(), steps)
        got = gr.host(fin)
        ui = np.uint32 if dtype == F32C else np.uint64
        assert np.array_equal(got.view(ui), ref.view(ui)), f"grid {size} differs"
    b.close()


@pytest.mark.parametrize("dtype", [F64, F32C])
@pytest.mark.parametrize("block,workgroups", [((32, 8, 1), 0), ((64, 2, 2), 5)])
def test_packed_star_variable_batch_matches_oracle(dtype, block, workgroups):
    """Packed tiles with a Variable mask (set once for the kernel, shared by every grid), range
    3, Write alignment and coarsening 2x3x5 with remainder handling, on uneven grids not rounded
    to the coarsening, 4 steps through the graph: every grid equals the oracle."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d = manifest.packed_star_desc(dtype)
    k = rt.Kernel(d)
    mask = rt.mask_seeded(8, d.n)
    k.set_mask(mask)
    sizes = rt.batch_sizes(7, 5, 12, 60, (1, 1, 1), d.range)
    print(f"packed star batch {d.key()} block {block} sizes {sizes}", flush=True)
    geoms, grids = [], []
    for i, size in enumerate(sizes):
        g = k.geometry(size)
        gr = rt.Grids(g, k.dtype)
        for j in range(2):
            k.init(g, gr.ptr(j), rt.INIT_NORMAL, 50 + i)
        geoms.append(g)
        grids.append(gr)
    torch.cuda.synchronize()
    init = [gr.host(0) for gr in grids]
    b = rt.Batch(k, geoms, [gr.ptr(0) for gr in grids], [gr.ptr(1) for gr in grids], block, workgroups=workgroups)
    steps = 4
    fin = b.run(steps, graph=True)
    torch.cuda.synchronize()
    run = oracle.run_f32c if dtype == F32C else oracle.run_f64
    ui = np.uint32 if dtype == F32C else np.uint64
    for size, gr, a in zip(sizes, grids, init):
        og = oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, (1, 1, 1))
        ref = run(d, og, mask, a.copy(), a.copy(), steps)
        assert np.array_equal(gr.host(fin).view(ui), ref.view(ui)), f"grid {size} differs"
    b.close()


@pytest.mark.parametrize("dtype", [F64, F32C])
def test_graph_replays_with_work_between(dtype):
    """The captured step loop replayed four times from the same initial state,
    with other GPU work (the plain kernel on another grid, a torch reduction)
    and host synchronisation between replays: every replay equals the oracle,
    so each launch found the work queue cleared by the previous one."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    d = manifest.cfg4_desc(shape="plain", batched=True, dtype=dtype, max_threads=1024)
    k = rt.Kernel(d)
    sizes = rt.batch_sizes(5, 7, 8, 40, d.coarsening, d.range)
    print(f"graph replays {d.key()} sizes {sizes}", flush=True)
    geoms, grids = [], []
    for i, size in enumerate(sizes):
        g = k.geometry(size)
        gr = rt.Grids(g, k.dtype)
        for j in range(2):
            k.init(g, gr.ptr(j), rt.INIT_U01, 40 + i)
        geoms.append(g)
        grids.append(gr)
    torch.cuda.synchronize()
    init = [gr.host(0) for gr in grids]
    steps = 3
    run = oracle.run_f32c if dtype == F32C else oracle.run_f64
    refs = [run(d, oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, d.coarsening), None, a.copy(),
                a.copy(), steps) for size, a in zip(sizes, init)]
    other = manifest.cfg4_desc(shape="plain", dtype=dtype)
    ko = rt.Kernel(other)
    go = ko.geometry((28, 20, 24))
    gro = rt.Grids(go, ko.dtype)
    b = rt.Batch(k, geoms, [gr.ptr(0) for gr in grids], [gr.ptr(1) for gr in grids], (32, 8, 2), workgroups=0)
    ui = np.uint32 if dtype == F32C else np.uint64
    for replay in range(4):
        for gr, a in zip(grids, init):
            gr.upload(0, a)
            gr.upload(1, a)
        torch.cuda.synchronize()
        fin = b.run(steps, graph=True)
        torch.cuda.synchronize()
        for gr, ref, size in zip(grids, refs, sizes):
            assert np.array_equal(gr.host(fin).view(ui), ref.view(ui)), f"replay {replay}: grid {size} differs"
        # unrelated work between replays
        ko.run(go, (32, 4, 2), rt.grid_blocks(go, (32, 4, 2)), gro.ptr(0), gro.ptr(1), 2)
        float(gro.bufs[0].sum())
        torch.cuda.synchronize()
    b.close()


@pytest.mark.parametrize("dtype", [F64, F32C])
def test_benchmark_batch_matches_oracle(dtype):
    """The batch bench.py times: 64 grids with interior axes log-uniform in [32, 256]
    (rounded to the 2x2x4 coarsening), the benchmarked kernel (z march over packed tiles with
    remainder handling) and block, the occupancy-sized persistent grid (one range of the
    tile table per XCD, workgroups draining other XCDs' ranges, the last workgroup clearing
    the queue), 3 steps through the captured graph, replayed twice from the restored state:
    every grid equals the oracle."""
    import torch

    from threadcoarsenedstencilgenerator_amd import runtime as rt
    from threadcoarsenedstencilgenerator_amd import tuned
    cfg = tuned.best("batched", dtype)
    d, block = cfg["desc"], tuple(cfg["block"])
    k = rt.Kernel(d)
    sizes = rt.batch_sizes(64, 1, 32, 256, tuned.BATCH_ROUNDING, d.range)
    geoms, grids = [], []
    for i, size in enumerate(sizes):
        g = k.geometry(size)
        gr = rt.Grids(g, k.dtype)
        for j in range(2):
            k.init(g, gr.ptr(j), rt.INIT_U01, 100 + i)
        geoms.append(g)
        grids.append(gr)
    torch.cuda.synchronize()
    init = [gr.host(0) for gr in grids]
    b = rt.Batch(k, geoms, [gr.ptr(0) for gr in grids], [gr.ptr(1) for gr in grids], block, workgroups=0)
    pts, tiles, nwg = b.info()
    print(f"benchmark batch {d.key()}: {len(sizes)} grids, {pts} points, {tiles} tiles, {nwg} workgroups",
          flush=True)
    assert len(sizes) == 64 and nwg >= 256 and tiles > nwg
    steps = 3
    run = oracle.run_f32c if dtype == F32C else oracle.run_f64
    refs = [run(d, oracle.Geometry(3, d.range, d.alignment, d.cl_size, size, tuned.BATCH_ROUNDING), None, a.copy(),
                a.copy(), steps) for size, a in zip(sizes, init)]
    ui = np.uint32 if dtype == F32C else np.uint64
    for replay in range(2):
        for gr, a in zip(grids, init):
            gr.upload(0, a)
            gr.upload(1, a)
        torch.cuda.synchronize()
        fin = b.run(steps, graph=True)
        torch.cuda.synchronize()
        for gr, ref, size in zip(grids, refs, sizes):
            assert np.array_equal(gr.host(fin).view(ui), ref.view(ui)), f"replay {replay}: grid {size} differs"
    b.close()
