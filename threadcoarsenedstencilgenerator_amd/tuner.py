"""Learned block-shape tuner.

The reference finds a kernel's block shape by measuring every legal one: the
harness sweeps total thread counts 64, 128, ... up to the launch bound and
every factorisation of each into (x, y[, z]) (hip/benchmark.cpp:221-263 with
hip/primes.cpp).  This module predicts each candidate's throughput with a
small cost model, measures only the top k predictions and keeps the best.

The cost model is a 3-layer MLP (128 hidden units, ReLU, FP32) trained with
Adam by the HIP kernels of csrc/tcsg_mlp.hip.  Its inputs are the stencil's
parameters, the grid shape, the block shape and the coarsening
(``features``); its target is log GLup/s.  The training corpus is a set of
full sweeps measured on MI355X (``tuner_data/``), gathered by
``collect`` over the kernels of ``corpus()``.

    python -m threadcoarsenedstencilgenerator_amd.tuner build            # compile the corpus kernels
    python -m threadcoarsenedstencilgenerator_amd.tuner collect OUT.csv  # measure the sweeps (GPU)
    python -m threadcoarsenedstencilgenerator_amd.tuner evaluate DATA    # k-fold regret report (GPU)
    python -m threadcoarsenedstencilgenerator_amd.tuner fit DATA MODEL   # train on all of DATA (GPU)
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

from .spec import (ALIGN_NONE, ALIGN_READ, ALIGN_WRITE, BOX, COLUMN, F32C, F64, JACOBI, PLAIN, ROW, STAR,
                   STREAM, VARIABLE, StencilDesc)

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(HERE, "tuner_data")
DEFAULT_DATA = os.path.join(DATA_DIR, "sweeps_mi355x.csv.gz")
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
            "log2_tiles_x", "log2_tiles_yz")


def features(desc: StencilDesc, size, block) -> np.ndarray:
    """Input vector of the cost model for one (kernel, grid, block) triple.
    size is the grid extent including the halo, as the harness takes it."""
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
    f = [
        float(dim == 3), float(desc.neighbourhood == STAR), float(desc.coeff == VARIABLE),
        float(desc.orientation == COLUMN), float(r), l2(desc.n), l2(co[0]), l2(co[1]), l2(co[2]),
        l2(interior[0]), l2(interior[1]), l2(interior[2]), l2(bl[0]), l2(bl[1]), l2(bl[2]),
        l2(bl[0] * bl[1] * bl[2]), l2(tile[0]), l2(tile[1]), l2(tile[2]), l2(tiles[0] * tiles[1] * tiles[2]),
        l2(cover), float(desc.dtype == F32C), float(stream), float(stream and desc.xext == "load"),
        float(stream and desc.xext == "dpp"), float(stream and desc.xext == "dppw"), l2(desc.max_threads),
        float(desc.zfirst), float(desc.prefetch if stream else 0), float(desc.waves_per_eu),
        4.0 if desc.dtype == F32C else 8.0, l2(tiles[0]), l2(tiles[1] * tiles[2]),
    ]
    return np.asarray(f, dtype=np.float64)


# -------------------------------------------------------------------- corpus
def _desc(dim, nb, coeff, r, co, shape, mt, align=ALIGN_NONE, cl=0, dtype=F64, orient=ROW):
    kw = dict(variant=PLAIN)
    if shape != "plain":
        kw = dict(variant=STREAM, xext={"stream": "load"}.get(shape, shape), xcd_remap=True, nt_store=True,
                  zfirst=shape == "dppw")
    return StencilDesc(dimension=dim, neighbourhood=nb, orientation=orient, coeff=coeff, range=r, xc=co[0],
                       yc=co[1], zc=co[2] if dim == 3 else 1, alignment=align, cl_size=cl, dtype=dtype,
                       max_threads=mt, **kw)


def corpus():
    """(kernel, grid sizes incl. halo) pairs whose full sweeps form the
    training corpus: 3D grids of 128^3, 256^3 and 512^3 interior points, 2D
    grids of 1024^2, 4096^2 and 16384^2, over stencil shapes, coarsenings
    and code shapes."""
    W, R = ALIGN_WRITE, ALIGN_READ
    k3 = [
        _desc(3, STAR, VARIABLE, 4, (1, 1, 1), "plain", 512, W, 16),
        _desc(3, STAR, VARIABLE, 4, (2, 4, 8), "plain", 512, W, 16),
        _desc(3, STAR, VARIABLE, 4, (1, 2, 4), "stream", 256, W, 16),
        _desc(3, STAR, VARIABLE, 4, (1, 4, 16), "dpp", 256, W, 16),
        _desc(3, STAR, VARIABLE, 4, (2, 4, 32), "dppw", 512, W, 16),
        _desc(3, STAR, JACOBI, 1, (1, 1, 1), "plain", 512),
        _desc(3, STAR, JACOBI, 1, (1, 2, 8), "dpp", 256),
        _desc(3, STAR, JACOBI, 1, (2, 2, 16), "dppw", 512),
        _desc(3, STAR, VARIABLE, 2, (1, 2, 4), "plain", 512, W, 16),
        _desc(3, STAR, VARIABLE, 2, (1, 4, 16), "dppw", 512, W, 16),
        _desc(3, STAR, VARIABLE, 3, (1, 2, 8), "dpp", 256, R, 8, orient=COLUMN),
        _desc(3, BOX, JACOBI, 2, (2, 2, 4), "plain", 512),
        _desc(3, BOX, JACOBI, 2, (2, 2, 4), "dpp", 256),
        _desc(3, BOX, JACOBI, 2, (2, 2, 4), "dppw", 512),
        _desc(3, BOX, VARIABLE, 1, (1, 1, 1), "plain", 512, W, 16),
        _desc(3, BOX, VARIABLE, 1, (1, 2, 8), "dpp", 256, W, 16),
        _desc(3, BOX, VARIABLE, 1, (2, 2, 8), "dppw", 512, W, 16),
        _desc(3, STAR, VARIABLE, 4, (2, 4, 32), "dppw", 512, W, 16, dtype=F32C),
        _desc(3, BOX, JACOBI, 2, (2, 2, 4), "plain", 512, dtype=F32C),
    ]
    k2 = [
        _desc(2, BOX, JACOBI, 1, (1, 1, 1), "plain", 512),
        _desc(2, BOX, JACOBI, 1, (1, 1, 1), "stream", 256),
        _desc(2, BOX, JACOBI, 1, (1, 1, 1), "dppw", 512),
        _desc(2, BOX, JACOBI, 1, (2, 4, 1), "dppw", 512),
        _desc(2, BOX, VARIABLE, 8, (1, 1, 1), "plain", 512, R, 16),
        _desc(2, BOX, VARIABLE, 8, (2, 2, 1), "dpp", 256, R, 16),
        _desc(2, BOX, VARIABLE, 8, (1, 4, 1), "dppw", 512, R, 16),
        _desc(2, STAR, VARIABLE, 4, (1, 4, 1), "dpp", 256, W, 16),
        _desc(2, STAR, VARIABLE, 4, (2, 8, 1), "dppw", 512, W, 16),
        _desc(2, STAR, JACOBI, 2, (1, 8, 1), "stream", 256),
        _desc(2, BOX, VARIABLE, 8, (2, 2, 1), "dppw", 512, R, 16, dtype=F32C),
    ]
    out = []
    for d in k3:
        out.append((d, [(n + 2 * d.range,) * 3 for n in (128, 256, 512)]))
    for d in k2:
        out.append((d, [(n + 2 * d.range,) * 2 for n in (1024, 4096, 16384)]))
    return out


# -------------------------------------------------------------- measurements
COLUMNS = ("key", "dimension", "neighbourhood", "orientation", "coeff", "range", "xc", "yc", "zc", "alignment",
           "cl_size", "dtype", "variant", "max_threads", "xext", "xcd_remap", "nt_store", "comp_group",
           "prefetch", "waves_per_eu", "zfirst", "sx", "sy", "sz", "bx", "by", "bz", "steps", "ms_per_step",
           "glups")
_DESC_INT = ("dimension", "neighbourhood", "orientation", "coeff", "range", "xc", "yc", "zc", "alignment",
             "cl_size", "dtype", "variant", "max_threads", "comp_group", "prefetch", "waves_per_eu")
_DESC_BOOL = ("xcd_remap", "nt_store", "zfirst")


def row_desc(row) -> StencilDesc:
    kw = {k: int(row[k]) for k in _DESC_INT}
    kw.update({k: bool(int(row[k])) for k in _DESC_BOOL})
    return StencilDesc(xext=row["xext"], **kw)


def _steps_for(points: int) -> int:
    return int(min(40, max(2, round(2e8 / points))))


def measure_sweep(desc: StencilDesc, size, shapes=None):
    """Time every candidate block shape of one kernel on one grid (HIP
    events, back-to-back steps).  Returns rows for the corpus."""
    import torch

    from . import runtime as rt
    k = rt.Kernel(desc)
    g = k.geometry(size)
    if desc.coeff == VARIABLE:
        k.set_mask(rt.mask_seeded(2, desc.n))
    grids = rt.Grids(g, k.dtype)
    for i in range(2):
        k.init(g, grids.ptr(i), rt.INIT_U01, 1 + i)
    torch.cuda.synchronize()
    pts = g.interior_points
    steps = _steps_for(pts)
    rows = []
    shapes = shapes if shapes is not None else candidates(k, desc.dimension)
    base = {f: getattr(desc, f) for f in _DESC_INT + _DESC_BOOL}
    base.update(key=desc.key(), xext=desc.xext)
    sz = tuple(size) + (1,) * (3 - len(size))
    for s in shapes:
        blocks = rt.grid_blocks(g, s)
        try:
            one = k.time_steps(g, s, blocks, grids.ptr(0), grids.ptr(1), 0, 1)
        except rt.TcsgError as e:   # a shape the launch rejects is left out, as the harness prints and moves on
            print(f"    block {s}: {e}", flush=True)
            continue
        n = steps if one < 20.0 else 1
        ms = one if n == 1 else k.time_steps(g, s, blocks, grids.ptr(0), grids.ptr(1), 1, n) / n
        rows.append(dict(base, sx=sz[0], sy=sz[1], sz=sz[2], bx=s[0], by=s[1], bz=s[2], steps=n,
                         ms_per_step=round(ms, 6), glups=round(pts / (ms * 1e-3) / 1e9, 4)))
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


def read_rows(path: str = DEFAULT_DATA):
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
        self.X = np.stack([features(descs[r["key"]], self._size(r), (int(r["bx"]), int(r["by"]), int(r["bz"])))
                           for r in rows]) if rows else np.zeros((0, len(FEATURES)))
        self.glups = np.array([float(r["glups"]) for r in rows])
        self.ms = np.array([float(r["ms_per_step"]) for r in rows])
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

    def fit(self, X, y, epochs: int = 60, batch: int = 256, lr: float = 2e-3, seed: int = 0):
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
        np.savez(path, params=self.params(), mean=self.mean, std=self.std, y_mean=np.array(self.y_mean),
                 hidden=np.array(self.hidden), features=np.array(FEATURES))

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL):
        z = np.load(path, allow_pickle=False)
        if tuple(str(f) for f in z["features"]) != FEATURES:
            raise ValueError(f"{path}: feature set differs from this version")
        m = cls(hidden=int(z["hidden"]))
        m.set_params(z["params"])
        m.mean, m.std, m.y_mean = z["mean"], z["std"], float(z["y_mean"])
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


def top_k_outcome(glups: np.ndarray, ms: np.ndarray, order: np.ndarray, k: int):
    """(regret, tuning-time fraction) of measuring the first k of `order`."""
    top = order[:k]
    regret = glups.max() / glups[top].max() - 1.0
    frac = ms[top].sum() / ms.sum()
    return regret, frac


def _folds(keys, folds: int, seed: int):
    rng = np.random.default_rng(seed)
    keys = list(keys)
    rng.shuffle(keys)
    return [keys[i::folds] for i in range(folds)]


def evaluate(corpus: Corpus, folds: int = 5, k: int = 8, seed: int = 0, epochs: int = 60, hidden: int = 128,
             log=None):
    """Leave-kernels-out cross-validation: the kernels are split into
    `folds` groups; for each, a model trained on the other kernels' sweeps
    ranks the held-out kernels' candidates (every grid size) and the top k
    are 'measured' (looked up in the sweep).  Reports regret against the
    exhaustive sweep's best and the fraction of the sweep's timed work the
    top k cost, next to two model-free strategies: k shapes at random
    (expected value) and the k shapes with the best mean relative throughput
    over the training kernels."""
    per_group = []
    fit_s = 0.0
    for fi, held in enumerate(_folds(corpus.kernels(), folds, seed)):
        held = set(held)
        tr = np.array([i for i, r in enumerate(corpus.rows) if r["key"] not in held])
        m = CostModel(hidden=hidden, seed=seed + 1 + fi)
        t0 = time.perf_counter()
        m.fit(corpus.X[tr], corpus.y[tr], epochs=epochs, seed=seed + fi)
        fit_s += time.perf_counter() - t0
        # model-free baseline: mean throughput relative to each training sweep's best
        score, cnt = {}, {}
        for gk, idx in corpus.groups.items():
            if gk[0] in held:
                continue
            best = corpus.glups[idx].max()
            for i in idx:
                s = corpus.shape[i]
                score[s] = score.get(s, 0.0) + corpus.glups[i] / best
                cnt[s] = cnt.get(s, 0) + 1
        for gk, idx in corpus.groups.items():
            if gk[0] not in held:
                continue
            idx = np.array(idx)
            g = corpus.glups[idx]
            ms = corpus.ms[idx]
            t0 = time.perf_counter()
            pred = m.predict(corpus.X[idx])
            infer_s = time.perf_counter() - t0
            reg, frac = top_k_outcome(g, ms, np.argsort(-pred, kind="stable"), k)
            fixed = np.argsort(-np.array([score.get(corpus.shape[i], 0.0) / max(1, cnt.get(corpus.shape[i], 0))
                                          for i in idx]), kind="stable")
            freg, ffrac = top_k_outcome(g, ms, fixed, k)
            per_group.append(dict(kernel=gk[0], size=list(gk[1:]), candidates=len(idx), best_glups=float(g.max()),
                                  model_regret=reg, model_time_frac=frac, random_regret=random_k_regret(g, k),
                                  fixed_regret=freg, fixed_time_frac=ffrac, inference_s=infer_s))
        m.close()
        if log:
            log(f"fold {fi + 1}/{folds}: {len(held)} kernels held out")

    def agg(key, sel=lambda d: True):
        v = np.array([d[key] for d in per_group if sel(d)])
        return dict(mean=float(v.mean()), median=float(np.median(v)), max=float(v.max())) if v.size else None

    big = lambda d: d["candidates"] > 4 * k   # noqa: E731
    return dict(k=k, folds=folds, groups=len(per_group), kernels=len(corpus.kernels()), rows=len(corpus.rows),
                fit_seconds_total=fit_s,
                model_regret=agg("model_regret"), random_regret=agg("random_regret"),
                fixed_regret=agg("fixed_regret"), model_time_frac=agg("model_time_frac"),
                model_regret_large=agg("model_regret", big), random_regret_large=agg("random_regret", big),
                fixed_regret_large=agg("fixed_regret", big), large_groups=sum(1 for d in per_group if big(d)),
                per_group=per_group)


# ------------------------------------------------------------------- tuning
def tune(kernel, size, k: int = 8, model: CostModel | None = None, d0=None, d1=None, steps: int = 5):
    """Rank the kernel's candidate block shapes with the cost model, time
    the top k on the grid and return (best block, GLup/s, timings)."""
    import torch

    from . import runtime as rt
    model = model or CostModel.load()
    desc = kernel.desc
    shapes = candidates(kernel, desc.dimension)
    pred = model.predict(np.stack([features(desc, size, s) for s in shapes]))
    top = [shapes[i] for i in np.argsort(-pred, kind="stable")[:k]]
    g = kernel.geometry(size)
    grids = None
    if d0 is None:
        grids = rt.Grids(g, kernel.dtype)
        for i in range(2):
            kernel.init(g, grids.ptr(i), rt.INIT_U01, 1 + i)
        d0, d1 = grids.ptr(0), grids.ptr(1)
    res = []
    for s in top:
        ms = kernel.time_steps(g, s, rt.grid_blocks(g, s), d0, d1, 1, steps) / steps
        res.append((s, g.interior_points / (ms * 1e-3) / 1e9))
    del grids
    torch.cuda.synchronize()
    best = max(res, key=lambda t: t[1])
    return best[0], best[1], res


# ---------------------------------------------------------------------- CLI
def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("build")
    c = sub.add_parser("collect")
    c.add_argument("out")
    c.add_argument("--part", default="0/1", help="i/n: measure every n-th corpus entry starting at i")
    e = sub.add_parser("evaluate")
    e.add_argument("data", nargs="?", default=DEFAULT_DATA)
    e.add_argument("--json", default=None)
    e.add_argument("--k", type=int, default=8)
    e.add_argument("--epochs", type=int, default=60)
    f = sub.add_parser("fit")
    f.add_argument("data", nargs="?", default=DEFAULT_DATA)
    f.add_argument("model", nargs="?", default=DEFAULT_MODEL)
    f.add_argument("--epochs", type=int, default=60)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "build":
        from . import build as B
        B.build_kernels([d for d, _ in corpus()], force=False)
        return 0
    if a.cmd == "collect":
        import torch
        torch.cuda.set_device(0)
        i, n = (int(v) for v in a.part.split("/"))
        entries = [(d, s) for d, sizes in corpus() for s in sizes][i::n]
        for j, (d, s) in enumerate(entries):
            t0 = time.perf_counter()
            log(f"[{j + 1}/{len(entries)}] {d.key()} size {s} ...")
            rows = measure_sweep(d, s)
            write_rows(a.out, rows, append=True)
            best = max(rows, key=lambda r: r["glups"])
            log(f"    {len(rows)} shapes in {time.perf_counter() - t0:.1f}s, best {best['glups']:.1f} GLup/s "
                f"at {best['bx']}x{best['by']}x{best['bz']}")
        return 0
    if a.cmd == "evaluate":
        import torch
        torch.cuda.set_device(0)
        cp = Corpus(read_rows(a.data))
        rep = evaluate(cp, k=a.k, epochs=a.epochs, log=log)
        short = {kk: v for kk, v in rep.items() if kk != "per_group"}
        log(json.dumps(short, indent=1))
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(rep, fh, indent=1)
        return 0
    if a.cmd == "fit":
        import torch
        torch.cuda.set_device(0)
        cp = Corpus(read_rows(a.data))
        m = CostModel()
        losses = m.fit(cp.X, cp.y, epochs=a.epochs)
        m.save(a.model)
        log(f"trained on {len(cp.rows)} rows: loss {losses[0]:.4f} -> {losses[-1]:.4f}; saved {a.model}")
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
