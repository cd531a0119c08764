"""Learned block-shape tuner.

The reference finds a kernel's block shape by measuring every legal one: the
harness sweeps total thread counts 64, 128, ... up to the launch bound and
every factorisation of each into (x, y[, z]) (hip/benchmark.cpp:221-263 with
hip/primes.cpp).  This module predicts each candidate's throughput with a
small cost model, measures only the top k predictions and keeps the best;
without a trained model it falls back to the exhaustive sweep.

The cost model is a 3-layer MLP (128 hidden units, ReLU, FP32) trained with
Adam by the HIP kernels of csrc/tcsg_mlp.hip.  Its inputs are the stencil's
parameters, the grid shape, the block shape, the coarsening and the loaded
kernel's resources (``features``); its target is log GLup/s.

Data (tuner_data/, measured on MI355X by ``collect``): configurations are
drawn by ``sample_config`` from the tuning workload (2D/3D, Box/Star,
Jacobi/Variable, range 1-12, coarsening 1-16 per axis, 2^26-2^27 interior
points, launch bounds up to 1024 threads).  The training set times 16
random launchable block shapes of each of 250 configurations (about 3800
samples; its first 125 configurations, about 1900 samples, form the smaller
training set the evaluation also reports); the held-out set times every
launchable shape of 200 further configurations whose kernels the training
set does not contain.  That pair is the evaluation protocol (train on the
first set, score on the second); the committed model that ``tune`` uses is
trained on both sets (about 50,800 samples) and carries the model-free
ranking of their block shapes, which gets one of the k measurements.  The
live check (``live``) sweeps 40 further configurations on the GPU it runs on
and scores that model's picks against those sweeps.

    python -m threadcoarsenedstencilgenerator_amd.tuner build               # compile both sets' kernels
    python -m threadcoarsenedstencilgenerator_amd.tuner collect train       # measure (GPU), resumable
    python -m threadcoarsenedstencilgenerator_amd.tuner collect heldout
    python -m threadcoarsenedstencilgenerator_amd.tuner evaluate --json R   # top-k regret on held-out (GPU)
    python -m threadcoarsenedstencilgenerator_amd.tuner fit --train T H     # train the committed model (GPU)
    python -m threadcoarsenedstencilgenerator_amd.tuner live --json R       # sweep 40 fresh configurations now (GPU)
"""
from __future__ import annotations

import csv
import ctypes as C
import gzip
import json
import math
import os
import sys
import time

import numpy as np

from .spec import (ALIGN_NONE, BOX, COLUMN, F32C, F64, PLAIN, ROW, STAR,
                   STREAM, VARIABLE, StencilDesc)

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(HERE, "tuner_data")
TRAIN_DATA = os.path.join(DATA_DIR, "train_mi355x.csv.gz")
HELDOUT_DATA = os.path.join(DATA_DIR, "heldout_mi355x.csv.gz")
DATA = {"train": TRAIN_DATA, "heldout": HELDOUT_DATA}
DEFAULT_MODEL = os.path.join(DATA_DIR, "cost_model_mi355x.npz")
WARP = 64


# ---------------------------------------------------------------- candidates
def divisors_in_sweep_order(n: int):
    """Divisors of n in the order the reference's sweep visits them
    (hip/primes.cpp): per prime factor p^m, the list so far times p^0..p^m."""
    out = [1]
    rest = n
    p = 2
    while rest > 1:
        if p * p > rest:
            p = rest
        m = 0
        while rest % p == 0:
            rest //= p
            m += 1
        if m:
            out = [d * p ** e for e in range(m + 1) for d in out]
        p += 1
    return out


def sweep_shapes(dim: int, max_threads: int):
    """Block shapes of the exhaustive sweep: every factorisation of each
    multiple of 64 up to max_threads (hip/benchmark.cpp:221-263)."""
    shapes = []
    for n in range(WARP, max_threads + 1, WARP):
        for x in divisors_in_sweep_order(n):
            if dim == 2:
                shapes.append((x, n // x, 1))
            else:
                for y in divisors_in_sweep_order(n // x):
                    shapes.append((x, y, n // x // y))
    return shapes


def candidates(kernel, dim: int):
    """The sweep's shapes that this kernel can launch."""
    _, largest = kernel.max_block_size()
    return [s for s in sweep_shapes(dim, largest) if kernel.block_ok(s)]


# ------------------------------------------------------------------ features
FEATURES = ("is3d", "star", "variable", "column", "range", "log2_points", "log2_xc", "log2_yc", "log2_zc",
            "log2_nx", "log2_ny", "log2_nz", "log2_bx", "log2_by", "log2_bz", "log2_threads", "log2_tile_x",
            "log2_tile_y", "log2_tile_z", "log2_workgroups", "log2_cover", "f32c", "stream", "xext_load",
            "xext_dpp", "xext_dppw", "log2_max_threads", "zfirst", "prefetch", "waves_per_eu", "bytes_per_point",
            "log2_tiles_x", "log2_tiles_yz", "log2_vgprs", "log2_scratch", "wg_per_cu", "waves_per_simd",
            "wave_rounds", "last_round_fill", "log2_terms", "row_bytes_per_wave")
CUS, SIMDS_PER_CU = 256, 4


def features(desc: StencilDesc, size, block, res=(0, 0, 0)) -> np.ndarray:
    """Input vector of the cost model for one (kernel, grid, block) triple.
    size is the grid extent including the halo, as the harness takes it;
    res = (VGPRs per lane, scratch bytes per lane, workgroups per CU at this
    block size) of the loaded kernel (runtime Kernel.resources)."""
    r = desc.range
    dim = desc.dimension
    size = tuple(size) + (1,) * (3 - len(tuple(size)))
    interior = [max(1, size[0] - 2 * r), max(1, size[1] - 2 * r), max(1, size[2] - 2 * r) if dim == 3 else 1]
    co = (desc.xc, desc.yc, desc.zc if dim == 3 else 1)
    bl = tuple(block) + (1,) * (3 - len(tuple(block)))
    tile = [bl[i] * co[i] for i in range(3)]
    tiles = [math.ceil(interior[i] / tile[i]) for i in range(3)]
    cover = 1.0
    for i in range(3):
        cover *= tiles[i] * tile[i] / interior[i]
    l2 = math.log2
    stream = desc.variant == STREAM
    threads = bl[0] * bl[1] * bl[2]
    vgprs, scratch, wgpc = (int(v) for v in res)
    wg = tiles[0] * tiles[1] * tiles[2]
    waves_wg = (threads + 63) // 64
    per_round = CUS * max(1, wgpc)
    rounds = math.ceil(wg / per_round)
    es = 4 if desc.dtype == F32C else 8
    f = [
        float(dim == 3), float(desc.neighbourhood == STAR), float(desc.coeff == VARIABLE),
        float(desc.orientation == COLUMN), float(r), l2(desc.n), l2(co[0]), l2(co[1]), l2(co[2]),
        l2(interior[0]), l2(interior[1]), l2(interior[2]), l2(bl[0]), l2(bl[1]), l2(bl[2]),
        l2(threads), l2(tile[0]), l2(tile[1]), l2(tile[2]), l2(wg),
        l2(cover), float(desc.dtype == F32C), float(stream), float(stream and desc.xext == "load"),
        float(stream and desc.xext == "dpp"), float(stream and desc.xext == "dppw"), l2(desc.max_threads),
        float(desc.zfirst), float(desc.prefetch if stream else 0), float(desc.waves_per_eu),
        float(es), l2(tiles[0]), l2(tiles[1] * tiles[2]),
        l2(max(1, vgprs)), l2(1 + scratch), float(wgpc), wgpc * waves_wg / SIMDS_PER_CU,
        l2(rounds), wg / (rounds * per_round), l2(co[0] * co[1] * co[2] * desc.n),
        # bytes one wave's row segment spans in x (coalescing)
        l2(min(bl[0], 64) * co[0] * es),
    ]
    return np.asarray(f, dtype=np.float64)


# -------------------------------------------------------------------- corpus
TUNE_SHAPES = ("plain", "stream", "dpp", "dppw")
MAX_TERMS = 4096          # outputs per thread x stencil size of a sampled kernel
TRAIN_CONFIGS, HELDOUT_CONFIGS = 250, 200
TRAIN_SHAPES_PER_CONFIG = 16
TRAIN_SEED, HELDOUT_SEED = 11, 12
# live check: further configurations, none of whose kernels either corpus contains, swept and
# ranked on the GPU the check runs on (nothing about them was measured beforehand)
LIVE_CONFIGS, LIVE_SEED = 40, 20261205


def _desc(dim, nb, coeff, r, co, shape, mt, align=ALIGN_NONE, cl=0, dtype=F64, orient=ROW):
    kw = dict(variant=PLAIN)
    if shape != "plain":
        kw = dict(variant=STREAM, xext={"stream": "load"}.get(shape, shape), xcd_remap=True, nt_store=True,
                  zfirst=shape == "dppw")
    return StencilDesc(dimension=dim, neighbourhood=nb, orientation=orient, coeff=coeff, range=r, xc=co[0],
                       yc=co[1], zc=co[2] if dim == 3 else 1, alignment=align, cl_size=cl, dtype=dtype,
                       max_threads=mt, **kw)


def sample_config(rng):
    """One stencil configuration of the tuning workload: 2D/3D, Box/Star,
    Jacobi/Variable, Row/Column, range 1-12 (3D Box up to 4, so one step
    stays within a few milliseconds), coarsening 1-16 per axis (powers of
    two, at most MAX_TERMS accumulate statements per thread), code shape,
    launch bound 256-1024, alignment, FP64 or FP32-compensated storage, and
    a grid of 2^26-2^27 interior points with a random aspect ratio.
    Returns (desc, grid size incl. halo)."""
    while True:
        dim = int(rng.integers(2, 4))
        nb = int(rng.integers(0, 2))
        coeff = int(rng.integers(0, 2))
        orient = int(rng.integers(0, 2))
        r = int(rng.integers(1, (4 if dim == 3 and nb == BOX else 12) + 1))
        shape = TUNE_SHAPES[int(rng.choice(4, p=[0.4, 0.2, 0.2, 0.2]))]
        dtype = F32C if rng.random() < 0.2 else F64
        co = [2 ** int(rng.integers(0, 5)) for _ in range(3)]
        if dim == 2:
            co[2] = 1
        if dtype == F32C:
            co[0] = min(co[0], 2)
        n = (2 * r + 1) ** dim if nb == BOX else 2 * dim * r + 1
        if co[0] * co[1] * co[2] * n > MAX_TERMS:
            continue
        mt = int(rng.choice([256, 512, 1024]))
        align = int(rng.integers(0, 3))
        cl = (16 if dtype == F64 else 32) if align != ALIGN_NONE else 0
        e = rng.uniform(26.0, 27.0)
        lo = 6.0 if dim == 3 else 9.0
        frac = rng.dirichlet([3.0] * dim)
        interior = []
        for a in range(dim):
            length = 2.0 ** (lo + frac[a] * (e - dim * lo))
            interior.append(max(co[a], int(round(length / co[a])) * co[a]))
        if dim == 3 and interior[2] < 2 * r:
            continue
        d = _desc(dim, nb, coeff, r, co, shape, mt, align, cl, dtype, orient)
        return d, tuple(v + 2 * r for v in interior)


def sample_configs(n: int, seed: int, exclude=()):
    """n configurations with distinct kernels, none of them in `exclude`
    (kernel keys)."""
    rng = np.random.default_rng(seed)
    seen = set(exclude)
    out = []
    while len(out) < n:
        d, size = sample_config(rng)
        if d.key() in seen:
            continue
        seen.add(d.key())
        out.append((d, size))
    return out


def train_configs():
    return sample_configs(TRAIN_CONFIGS, TRAIN_SEED)


def heldout_configs():
    """Held-out configurations: kernels that no training configuration uses."""
    return sample_configs(HELDOUT_CONFIGS, HELDOUT_SEED, exclude={d.key() for d, _ in train_configs()})


def live_configs():
    """Configurations of the live check: kernels in neither the training nor the held-out set."""
    return sample_configs(LIVE_CONFIGS, LIVE_SEED,
                          exclude={d.key() for d, _ in train_configs() + heldout_configs()})


def config_set(name: str):
    return {"train": train_configs, "heldout": heldout_configs, "live": live_configs}[name]()


# -------------------------------------------------------------- measurements
COLUMNS = ("key", "dimension", "neighbourhood", "orientation", "coeff", "range", "xc", "yc", "zc", "alignment",
           "cl_size", "dtype", "variant", "max_threads", "xext", "xcd_remap", "nt_store", "comp_group",
           "prefetch", "waves_per_eu", "zfirst", "sx", "sy", "sz", "bx", "by", "bz", "vgprs", "scratch",
           "wg_per_cu", "steps", "ms_per_step", "glups")
_DESC_INT = ("dimension", "neighbourhood", "orientation", "coeff", "range", "xc", "yc", "zc", "alignment",
             "cl_size", "dtype", "variant", "max_threads", "comp_group", "prefetch", "waves_per_eu")
_DESC_BOOL = ("xcd_remap", "nt_store", "zfirst")


def row_desc(row) -> StencilDesc:
    kw = {k: int(row[k]) for k in _DESC_INT}
    kw.update({k: bool(int(row[k])) for k in _DESC_BOOL})
    return StencilDesc(xext=row["xext"], **kw)


SLOW_MS = 5.0   # a shape whose first step takes longer is timed by that step alone


def measure_sweep(desc: StencilDesc, size, shapes=None, mask_seed: int = 2, sample=None):
    """Time candidate block shapes of one kernel on one grid (HIP events on
    back-to-back steps): per shape one step, which also warms the shape up,
    then two timed steps unless the first took longer than SLOW_MS.  shapes
    defaults to every launchable shape of the sweep; sample=(n, seed) times
    n of them drawn at random instead.  Returns corpus rows."""
    import torch

    from . import runtime as rt
    k = rt.Kernel(desc)
    g = k.geometry(size)
    if desc.coeff == VARIABLE:
        k.set_mask(rt.mask_seeded(mask_seed, desc.n))
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_U01, 1 + i)
    torch.cuda.synchronize()
    pts = g.interior_points
    rows = []
    shapes = shapes if shapes is not None else candidates(k, desc.dimension)
    if sample is not None and len(shapes) > sample[0]:
        pick = np.random.default_rng(sample[1]).choice(len(shapes), size=sample[0], replace=False)
        shapes = [shapes[i] for i in sorted(pick)]
    base = {f: getattr(desc, f) for f in _DESC_INT + _DESC_BOOL}
    base.update(key=desc.key(), xext=desc.xext)
    sz = tuple(size) + (1,) * (3 - len(size))
    occ = {}
    for s in shapes:
        blocks = rt.grid_blocks(g, s)
        t = s[0] * s[1] * s[2]
        if t not in occ:
            occ[t] = k.resources(t)
        vg, sc, wgpc = occ[t]
        try:
            one = k.time_steps(g, s, blocks, grids.ptr(0), grids.ptr(1), 0, 1)
        except rt.TcsgError as e:   # a shape the launch rejects is left out, as the harness prints and moves on
            print(f"    block {s}: {e}", flush=True)
            continue
        n = 1 if one > SLOW_MS else 2
        ms = one if n == 1 else k.time_steps(g, s, blocks, grids.ptr(0), grids.ptr(1), 0, n) / n
        rows.append(dict(base, sx=sz[0], sy=sz[1], sz=sz[2], bx=s[0], by=s[1], bz=s[2], vgprs=vg, scratch=sc,
                         wg_per_cu=wgpc, steps=n, ms_per_step=round(ms, 6),
                         glups=round(pts / (ms * 1e-3) / 1e9, 4)))
    del grids
    k.close()
    torch.cuda.empty_cache()
    return rows


def write_rows(path: str, rows, append=False):
    opener = gzip.open if path.endswith(".gz") else open
    new = not (append and os.path.exists(path))
    with opener(path, "at" if append else "wt", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COLUMNS)
        if new:
            w.writeheader()
        for r in rows:
            w.writerow({c: (int(v) if isinstance(v, bool) else v) for c, v in ((c, r[c]) for c in COLUMNS)})


def read_rows(path: str):
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt", newline="") as f:
        return list(csv.DictReader(f))


class Corpus:
    """Rows grouped by (kernel, grid): one group is one exhaustive sweep."""

    def __init__(self, rows):
        self.rows = rows
        descs = {}
        self.groups = {}
        for i, r in enumerate(rows):
            gk = (r["key"], int(r["sx"]), int(r["sy"]), int(r["sz"]))
            self.groups.setdefault(gk, []).append(i)
            if r["key"] not in descs:
                descs[r["key"]] = row_desc(r)
        self.descs = descs
        self.X = np.stack([features(descs[r["key"]], self._size(r), (int(r["bx"]), int(r["by"]), int(r["bz"])),
                                    (int(r["vgprs"]), int(r["scratch"]), int(r["wg_per_cu"])))
                           for r in rows]) if rows else np.zeros((0, len(FEATURES)))
        self.glups = np.array([float(r["glups"]) for r in rows])
        self.ms = np.array([float(r["ms_per_step"]) for r in rows])
        self.steps = np.array([int(r["steps"]) for r in rows], dtype=np.int64)
        self.shape = [(int(r["bx"]), int(r["by"]), int(r["bz"])) for r in rows]
        self.y = np.log(self.glups)

    @staticmethod
    def _size(r):
        s = (int(r["sx"]), int(r["sy"]), int(r["sz"]))
        return s if int(r["dimension"]) == 3 else s[:2]

    def kernels(self):
        return sorted(self.descs)


# ----------------------------------------------------------------- the model
def _lib():
    from . import runtime as rt
    return rt.library()


class CostModel:
    """The HIP-trained MLP (csrc/tcsg_mlp.hip) with input standardisation."""

    def __init__(self, in_dim: int = len(FEATURES), hidden: int = 128, max_batch: int = 4096, seed: int = 1):
        from . import runtime as rt
        self.in_dim, self.hidden, self.max_batch = in_dim, hidden, max_batch
        h = C.c_void_p()
        rt._check(_lib().tcsg_mlp_create(C.byref(h), in_dim, hidden, max_batch, seed), "mlp_create")
        self._h = h
        self.n_params = int(_lib().tcsg_mlp_num_params(in_dim, hidden))
        self.mean = np.zeros(in_dim)
        self.std = np.ones(in_dim)
        self.y_mean = 0.0
        # the model-free ranking of the corpora the model was trained on (shape_scores) and how
        # many of tune()'s k measurements go to it (mixed_order)
        self.fixed = {}
        self.mix = 0

    def close(self):
        if getattr(self, "_h", None):
            _lib().tcsg_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # raw parameter access (float32, the layout of tcsg_mlp.hip)
    def params(self) -> np.ndarray:
        from . import runtime as rt
        p = np.empty(self.n_params, dtype=np.float32)
        rt._check(_lib().tcsg_mlp_get_params(self._h, p.ctypes.data, p.size), "mlp_get_params")
        return p

    def set_params(self, p: np.ndarray):
        from . import runtime as rt
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter count mismatch")
        rt._check(_lib().tcsg_mlp_set_params(self._h, p.ctypes.data, p.size), "mlp_set_params")

    def _norm(self, X):
        return ((np.asarray(X, dtype=np.float64) - self.mean) / self.std).astype(np.float32)

    def forward_raw(self, X32: np.ndarray) -> np.ndarray:
        """Network output for already standardised float32 inputs."""
        import torch

        from . import runtime as rt
        X32 = np.ascontiguousarray(X32, dtype=np.float32)
        out = torch.empty(max(1, X32.shape[0]), dtype=torch.float32, device="cuda")
        xd = torch.from_numpy(X32).to("cuda")
        for s in range(0, X32.shape[0], self.max_batch):
            b = min(self.max_batch, X32.shape[0] - s)
            rt._check(_lib().tcsg_mlp_forward(self._h, xd.data_ptr() + 4 * s * self.in_dim, b,
                                              out.data_ptr() + 4 * s, None), "mlp_forward")
        torch.cuda.synchronize()
        return out.cpu().numpy()[:X32.shape[0]]

    def gradients_raw(self, X32, y32):
        """(gradient, loss) of the MSE at the current parameters."""
        import torch

        from . import runtime as rt
        xd = torch.from_numpy(np.ascontiguousarray(X32, dtype=np.float32)).to("cuda")
        yd = torch.from_numpy(np.ascontiguousarray(y32, dtype=np.float32)).to("cuda")
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_float()
        rt._check(_lib().tcsg_mlp_gradients(self._h, xd.data_ptr(), yd.data_ptr(), len(y32), g.ctypes.data,
                                            C.byref(loss)), "mlp_gradients")
        return g, loss.value

    def train_raw(self, X32, y32, lr: float):
        """One Adam step on one batch; returns the loss before the step."""
        import torch

        from . import runtime as rt
        xd = torch.from_numpy(np.ascontiguousarray(X32, dtype=np.float32)).to("cuda")
        yd = torch.from_numpy(np.ascontiguousarray(y32, dtype=np.float32)).to("cuda")
        ld = torch.zeros(1, dtype=torch.float32, device="cuda")
        rt._check(_lib().tcsg_mlp_train_step(self._h, xd.data_ptr(), yd.data_ptr(), len(y32), lr, ld.data_ptr(),
                                             None), "mlp_train_step")
        torch.cuda.synchronize()
        return float(ld.item())

    def fit(self, X, y, epochs: int = 60, batch: int = 64, lr: float = 2e-3, seed: int = 0):
        """Standardise the inputs on X, then Adam over shuffled minibatches
        with a cosine learning-rate decay.  Returns the per-epoch mean loss."""
        import torch

        from . import runtime as rt
        X = np.asarray(X, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        self.mean = X.mean(axis=0)
        self.std = X.std(axis=0)
        self.std[self.std < 1e-9] = 1.0
        self.y_mean = float(y.mean())
        Xn = self._norm(X)
        yn = (y - self.y_mean).astype(np.float32)
        n = len(y)
        batch = min(batch, self.max_batch, n)
        per_epoch = (n + batch - 1) // batch
        total = epochs * per_epoch
        rng = np.random.default_rng(seed)
        xd = torch.empty((n, self.in_dim), dtype=torch.float32, device="cuda")
        yd = torch.empty(n, dtype=torch.float32, device="cuda")
        losses = torch.zeros(total, dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        step = 0
        for _ in range(epochs):
            perm = rng.permutation(n)
            xd.copy_(torch.from_numpy(np.ascontiguousarray(Xn[perm])))
            yd.copy_(torch.from_numpy(np.ascontiguousarray(yn[perm])))
            for s in range(0, n, batch):
                b = min(batch, n - s)
                cur = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * step / total)))
                rt._check(_lib().tcsg_mlp_train_step(self._h, xd.data_ptr() + 4 * s * self.in_dim,
                                                     yd.data_ptr() + 4 * s, b, cur,
                                                     losses.data_ptr() + 4 * step, stream), "mlp_train_step")
                step += 1
        torch.cuda.synchronize()
        return losses.cpu().numpy().reshape(epochs, per_epoch).mean(axis=1)

    def predict(self, X) -> np.ndarray:
        """Predicted log GLup/s."""
        X = np.asarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[None, :]
        return self.forward_raw(self._norm(X)).astype(np.float64) + self.y_mean

    def save(self, path: str):
        shapes = sorted(self.fixed)
        np.savez(path, params=self.params(), mean=self.mean, std=self.std, y_mean=np.array(self.y_mean),
                 hidden=np.array(self.hidden), features=np.array(FEATURES), mix=np.array(self.mix),
                 fixed_shapes=np.array(shapes, dtype=np.int64).reshape(-1, 3),
                 fixed_scores=np.array([self.fixed[t] for t in shapes], dtype=np.float64))

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL):
        z = np.load(path, allow_pickle=False)
        if tuple(str(f) for f in z["features"]) != FEATURES:
            raise ValueError(f"{path}: feature set differs from this version")
        m = cls(hidden=int(z["hidden"]))
        m.set_params(z["params"])
        m.mean, m.std, m.y_mean = z["mean"], z["std"], float(z["y_mean"])
        if "mix" in z.files:
            m.mix = int(z["mix"])
            m.fixed = {tuple(int(v) for v in t): float(sc) for t, sc in zip(z["fixed_shapes"], z["fixed_scores"])}
        return m


# ---------------------------------------------------------------- evaluation
def random_k_regret(glups: np.ndarray, k: int) -> float:
    """Expected regret best/max(sample) - 1 of measuring k shapes drawn
    uniformly without replacement: the i-th largest value is the sample's
    maximum with probability C(n-i, k-1) / C(n, k)."""
    v = np.sort(np.asarray(glups, dtype=np.float64))[::-1]
    n = len(v)
    k = min(k, n)
    tot = math.comb(n, k)
    e = 0.0
    for i in range(1, n - k + 2):
        e += math.comb(n - i, k - 1) / tot * (v[0] / v[i - 1])
    return e - 1.0


def shape_cost_ms(ms: np.ndarray, steps: np.ndarray) -> np.ndarray:
    """GPU time measure_sweep spends on each shape: the warm-up step plus
    the timed steps (one step alone for a slow shape)."""
    ms = np.asarray(ms, dtype=np.float64)
    steps = np.asarray(steps)
    return ms * np.where(steps > 1, 1 + steps, 1)


def top_k_outcome(glups: np.ndarray, cost: np.ndarray, order: np.ndarray, k: int):
    """(regret, tuning-time fraction) of measuring the first k of `order`:
    regret = sweep best / best of the k - 1, time = their measuring cost
    over the whole sweep's."""
    top = order[:k]
    regret = glups.max() / glups[top].max() - 1.0
    frac = cost[top].sum() / cost.sum()
    return regret, frac


def shape_scores(corpus: Corpus):
    """Mean throughput of each block shape relative to the best measured
    shape of its (kernel, grid) group: the model-free ranking."""
    score, cnt = {}, {}
    for idx in corpus.groups.values():
        best = corpus.glups[idx].max()
        for i in idx:
            s = corpus.shape[i]
            score[s] = score.get(s, 0.0) + corpus.glups[i] / best
            cnt[s] = cnt.get(s, 0) + 1
    return {s: score[s] / cnt[s] for s in score}


def mixed_order(pred_order, fixed_order, k: int, mix: int):
    """The shapes to measure: the model's first k - mix, then the fixed ranking's first `mix`
    not already chosen, then the rest in the model's order."""
    head = list(pred_order[:k - mix])
    seen = set(head)
    for i in fixed_order:
        if len(head) >= k:
            break
        if i not in seen:
            head.append(i)
            seen.add(i)
    head += [i for i in pred_order if i not in seen]
    return np.array(head)


def evaluate(train: Corpus, heldout: Corpus, k: int = 8, seed: int = 0, epochs: int = 60, hidden: int = 128,
             model: "CostModel | None" = None, log=None, mix: int = 0, scores=None):
    """Train the cost model on `train` (or use `model`), then for every
    held-out (kernel, grid) group rank its candidates, 'measure' the top k
    (look them up in the group's exhaustive sweep) and record the regret
    against the sweep's best and the fraction of the sweep's GPU time the
    k measurements plus the model's inference cost.  Next to it: two
    model-free strategies with the same k, k shapes at random (expected
    regret) and the k shapes with the best mean relative throughput over
    the training groups.  mix > 0 gives that many of the k measurements to
    the second ranking (mixed_order); scores replaces the model-free ranking
    computed from `train`."""
    fit_s = 0.0
    own = model is None
    if own:
        model = CostModel(hidden=hidden, seed=seed + 1)
        t0 = time.perf_counter()
        model.fit(train.X, train.y, epochs=epochs, seed=seed)
        fit_s = time.perf_counter() - t0
        if log:
            log(f"trained on {len(train.rows)} rows from {len(train.groups)} groups in {fit_s:.1f}s")
    scores = shape_scores(train) if scores is None else scores
    per_group = []
    spent = {"model": 0.0, "fixed": 0.0, "sweep": 0.0}
    for gk, idx in heldout.groups.items():
        idx = np.array(idx)
        g = heldout.glups[idx]
        cost = shape_cost_ms(heldout.ms[idx], heldout.steps[idx])
        t0 = time.perf_counter()
        pred = model.predict(heldout.X[idx])
        order = np.argsort(-pred, kind="stable")
        infer_ms = (time.perf_counter() - t0) * 1e3
        fixed = np.argsort(-np.array([scores.get(heldout.shape[i], 0.0) for i in idx]), kind="stable")
        if mix:
            order = mixed_order(order, fixed, k, mix)
        reg, frac = top_k_outcome(g, cost, order, k)
        frac += infer_ms / cost.sum()
        freg, ffrac = top_k_outcome(g, cost, fixed, k)
        spent["model"] += cost[order[:k]].sum() + infer_ms
        spent["fixed"] += cost[fixed[:k]].sum()
        spent["sweep"] += cost.sum()
        per_group.append(dict(kernel=gk[0], size=list(gk[1:]), candidates=len(idx), best_glups=float(g.max()),
                              best_shape=list(heldout.shape[idx[int(np.argmax(g))]]),
                              model_shape=list(heldout.shape[idx[order[:k][int(np.argmax(g[order[:k]]))]]]),
                              model_regret=reg, model_time_frac=frac, random_regret=random_k_regret(g, k),
                              fixed_regret=freg, fixed_time_frac=ffrac, inference_ms=infer_ms,
                              sweep_ms=float(cost.sum()), rank_corr=float(np.corrcoef(pred, np.log(g))[0, 1])
                              if len(idx) > 2 else 1.0))
    if own:
        model.close()

    def agg(key, sel=lambda d: True):
        v = np.array([d[key] for d in per_group if sel(d)])
        return dict(mean=float(v.mean()), median=float(np.median(v)), p90=float(np.quantile(v, 0.9)),
                    max=float(v.max())) if v.size else None

    return dict(k=k, train_rows=len(train.rows), train_groups=len(train.groups), heldout_groups=len(per_group),
                heldout_rows=len(heldout.rows), fit_seconds=fit_s,
                model_regret=agg("model_regret"), random_regret=agg("random_regret"),
                fixed_regret=agg("fixed_regret"), model_time_frac=agg("model_time_frac"),
                fixed_time_frac=agg("fixed_time_frac"), rank_corr=agg("rank_corr"),
                # GPU time of tuning all held-out configurations, over the exhaustive sweeps' time
                model_time_frac_total=spent["model"] / spent["sweep"] if spent["sweep"] else None,
                fixed_time_frac_total=spent["fixed"] / spent["sweep"] if spent["sweep"] else None,
                sweep_seconds_total=spent["sweep"] * 1e-3,
                within_5pct=float(np.mean([d["model_regret"] <= 0.05 for d in per_group])) if per_group else None,
                per_group=per_group)


def live_evaluate(configs=None, k: int = 8, model: "CostModel | None" = None, log=None, rows_out=None):
    """The held-out evaluation re-measured where it runs: sweep every launchable block shape
    of each configuration now (measure_sweep), rank the same candidates with the committed
    model, and score the k shapes tune() would measure (the model's ranking mixed with the
    model-free one it carries) against that sweep (regret, share of the sweep's GPU time),
    next to k random shapes and the model-free ranking's k."""
    configs = live_configs() if configs is None else configs
    rows = []
    t0 = time.perf_counter()
    for j, (d, s) in enumerate(configs):
        got = measure_sweep(d, s)
        rows += got
        if log:
            best = max(got, key=lambda r: r["glups"])
            log(f"[{j}] {d.key()} {s}: {len(got)} shapes, best {best['glups']:.1f} GLup/s "
                f"at {best['bx']}x{best['by']}x{best['bz']}")
    sweep_s = time.perf_counter() - t0
    if rows_out:
        write_rows(rows_out, rows)
    own = model is None
    model = CostModel.load() if own else model
    rep = evaluate(Corpus([]), Corpus(rows), k=k, model=model, mix=model.mix, scores=model.fixed)
    for key in ("train_rows", "train_groups", "fit_seconds"):   # the committed model is not trained here
        rep.pop(key)
    rep["mix"] = model.mix
    if own:
        model.close()
    rep["wall_seconds_sweeps"] = sweep_s
    return rep


# ------------------------------------------------------------------- tuning
def rank(kernel, size, model: "CostModel", k: int = 8):
    """The kernel's launchable sweep shapes in the order tune() measures them: best predicted
    first, with `model.mix` of the first k taken from the model-free ranking the model
    carries (mixed_order)."""
    desc = kernel.desc
    shapes = candidates(kernel, desc.dimension)
    occ = {}
    X = []
    for s in shapes:
        t = s[0] * s[1] * s[2]
        if t not in occ:
            occ[t] = kernel.resources(t)
        X.append(features(desc, size, s, occ[t]))
    pred = model.predict(np.stack(X))
    order = np.argsort(-pred, kind="stable")
    if model.mix and model.fixed:
        fixed = np.argsort(-np.array([model.fixed.get(tuple(s), 0.0) for s in shapes]), kind="stable")
        order = mixed_order(order, fixed, k, model.mix)
    return [shapes[i] for i in order]


def tune(kernel, size, k: int = 8, model: "CostModel | None" = None, d0=None, d1=None, steps: int = 2):
    """Block shape for `kernel` on a grid of `size` (incl. halo): rank the
    sweep's shapes with the cost model and time the top k, or, with no
    trained model in the tree, time every shape (the reference's exhaustive
    sweep).  Returns (best block, GLup/s, [(block, GLup/s) measured])."""
    import torch

    from . import runtime as rt
    if model is None and os.path.exists(DEFAULT_MODEL):
        model = CostModel.load()
    shapes = rank(kernel, size, model, k)[:k] if model is not None else candidates(kernel, kernel.desc.dimension)
    g = kernel.geometry(size)
    grids = None
    if d0 is None:
        grids = rt.Grids(g, kernel.dtype)
        for i in range(2):
            kernel.init(g, grids.ptr(i), rt.INIT_U01, 1 + i)
        d0, d1 = grids.ptr(0), grids.ptr(1)
    res = []
    for s in shapes:
        ms = kernel.time_steps(g, s, rt.grid_blocks(g, s), d0, d1, 1, steps) / steps
        res.append((s, g.interior_points / (ms * 1e-3) / 1e9))
    del grids
    torch.cuda.synchronize()
    best = max(res, key=lambda t: t[1])
    return best[0], best[1], res


# ---------------------------------------------------------------------- CLI
def _done_groups(path):
    if not os.path.exists(path):
        return set()
    try:
        return {(r["key"], int(r["sx"]), int(r["sy"]), int(r["sz"])) for r in read_rows(path)}
    except (EOFError, OSError):   # a gzip member cut off by a killed run
        raise SystemExit(f"{path} is truncated; remove its last member or the file")


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("build")
    c = sub.add_parser("collect")
    c.add_argument("set", choices=("train", "heldout"))
    c.add_argument("--out", default=None)
    c.add_argument("--budget", type=float, default=600.0, help="stop starting new groups after this many seconds")
    e = sub.add_parser("evaluate")
    e.add_argument("--train", default=TRAIN_DATA)
    e.add_argument("--heldout", default=HELDOUT_DATA)
    e.add_argument("--json", default=None)
    e.add_argument("--k", type=int, default=8)
    e.add_argument("--epochs", type=int, default=60)
    e.add_argument("--train-configs", type=int, default=0,
                   help="train on the first N training configurations only (0: all)")
    e.add_argument("--seed", type=int, default=0)
    t = sub.add_parser("tune", help="block shape for one kernel: time the model's top k shapes")
    t.add_argument("dim", choices=("2D", "3D"))
    t.add_argument("kernel", help="the generator's argv in one string, e.g. "
                                  "'Star Row Variable 4 2 4 128 Write 16 --xext=dppw'")
    t.add_argument("--grid", type=int, nargs="+", required=True, help="grid size incl. halo, x y [z]")
    t.add_argument("--k", type=int, default=8)
    lv = sub.add_parser("live", help="sweep fresh configurations now and score the committed model's top k")
    lv.add_argument("--json", default=None)
    lv.add_argument("--k", type=int, default=8)
    lv.add_argument("--rows", default=None, help="also write the sweeps' rows (csv.gz) here")
    f = sub.add_parser("fit")
    f.add_argument("--train", nargs="+", default=[TRAIN_DATA])
    f.add_argument("--model", default=DEFAULT_MODEL)
    f.add_argument("--epochs", type=int, default=60)
    f.add_argument("--mix", type=int, default=1, help="of tune()'s k measurements, how many go to the "
                                                      "model-free ranking of the training corpora")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "build":
        from . import build as B
        B.build_kernels([d for d, _ in train_configs() + heldout_configs() + live_configs()], force=False)
        return 0
    if a.cmd == "live":
        import torch
        torch.cuda.set_device(0)
        rep = live_evaluate(k=a.k, log=log, rows_out=a.rows)
        log(json.dumps({kk: v for kk, v in rep.items() if kk != "per_group"}, indent=1))
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(rep, fh, indent=1)
        return 0
    if a.cmd == "collect":
        import torch
        torch.cuda.set_device(0)
        out = a.out or DATA[a.set]
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        done = _done_groups(out)
        todo = [(j, d, s) for j, (d, s) in enumerate(config_set(a.set))
                if (d.key(), *(tuple(s) + (1,) * (3 - len(s)))) not in done]
        log(f"{a.set}: {len(done)} groups measured, {len(todo)} to go")
        t_start = time.perf_counter()
        for j, d, s in todo:
            if time.perf_counter() - t_start > a.budget:
                log("budget reached")
                break
            t0 = time.perf_counter()
            log(f"[{j}] {d.key()} size {s} ...")
            rows = measure_sweep(d, s, sample=(TRAIN_SHAPES_PER_CONFIG, 1000 + j) if a.set == "train" else None)
            write_rows(out, rows, append=True)
            best = max(rows, key=lambda r: r["glups"])
            log(f"    {len(rows)} shapes in {time.perf_counter() - t0:.1f}s, best {best['glups']:.1f} GLup/s "
                f"at {best['bx']}x{best['by']}x{best['bz']}")
        return 0
    if a.cmd == "evaluate":
        import torch
        torch.cuda.set_device(0)
        rows = read_rows(a.train)
        if a.train_configs:
            keep = {d.key() for d, _ in train_configs()[:a.train_configs]}
            rows = [r for r in rows if r["key"] in keep]
        rep = evaluate(Corpus(rows), Corpus(read_rows(a.heldout)), k=a.k, epochs=a.epochs, seed=a.seed, log=log)
        short = {kk: v for kk, v in rep.items() if kk != "per_group"}
        log(json.dumps(short, indent=1))
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(rep, fh, indent=1)
        return 0
    if a.cmd == "tune":
        import torch

        from . import build as B
        from . import runtime as rt
        from .generator.cli import desc_from_argv
        torch.cuda.set_device(0)
        dim = int(a.dim[0])
        if len(a.grid) != dim:
            raise SystemExit(f"--grid takes {dim} sizes for a {a.dim} kernel")
        d, _ = desc_from_argv(dim, a.kernel.split())
        path = B.build_kernel(d, force=False)   # compiled here unless the tree already holds it
        k = rt.Kernel(d, path)
        if d.coeff == VARIABLE:
            k.set_mask([1.0 / d.n] * d.n)
        block, glups, tried = tune(k, tuple(a.grid), k=a.k)
        log(json.dumps({"kernel": d.key(), "grid": a.grid, "block": list(block), "glups": round(glups, 2),
                        "model": os.path.exists(DEFAULT_MODEL),
                        "tried": [[list(b), round(v, 2)] for b, v in tried]}))
        return 0
    if a.cmd == "fit":
        import torch
        torch.cuda.set_device(0)
        cp = Corpus([r for p in a.train for r in read_rows(p)])
        m = CostModel()
        losses = m.fit(cp.X, cp.y, epochs=a.epochs)
        m.fixed, m.mix = shape_scores(cp), a.mix
        m.save(a.model)
        log(f"trained on {len(cp.rows)} rows: loss {losses[0]:.4f} -> {losses[-1]:.4f}; saved {a.model}")
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
