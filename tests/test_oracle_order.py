"""Pins the oracle's offset set, mask layout and per-output summation order.

The order is derived here by replaying the loop nests of the reference's
generators -- which grid point each thread loads in which order
(hip/generate2DStencil_hip.py:245-298, hip/generate3DStencil_hip.py:266-361)
and which result registers each load feeds, with the reference's
initialise-or-add rule (hip/generate2DStencil_hip.py:168-198,
hip/generate3DStencil_hip.py:134-159) -- for several coarsening factors, and
compared with the oracle's canonical order.  The 3D Column Star nest is
replayed with the slice loop enclosing the cross, the evident intent of the
reference lines that do not parse (hip/generate3DStencil_hip.py:295-296).
"""
import itertools

import pytest

import oracle
from oracle import BOX, COLUMN, ROW, STAR


def _traversal(dim, nb, orient, r, C):
    """Grid points (thread-relative) in the order the reference kernel loads them."""
    xc, yc = C[0], C[1]
    zc = C[2] if dim == 3 else 1
    pts = []
    if dim == 2:
        if nb == BOX:
            if orient == COLUMN:
                pts = [(x, y, 0) for x in range(-r, r + xc) for y in range(-r, r + yc)]
            else:
                pts = [(x, y, 0) for y in range(-r, r + yc) for x in range(-r, r + xc)]
        else:
            if orient == COLUMN:
                pts += [(x, y, 0) for x in range(-r, 0) for y in range(0, yc)]
                pts += [(x, y, 0) for x in range(0, xc) for y in range(-r, r + yc)]
                pts += [(x, y, 0) for x in range(xc, r + xc) for y in range(0, yc)]
            else:
                pts += [(x, y, 0) for y in range(-r, 0) for x in range(0, xc)]
                pts += [(x, y, 0) for y in range(0, yc) for x in range(-r, r + xc)]
                pts += [(x, y, 0) for y in range(yc, r + yc) for x in range(0, xc)]
        return pts
    if nb == BOX:
        for z in range(-r, r + zc):
            if orient == COLUMN:
                pts += [(x, y, z) for x in range(-r, r + xc) for y in range(-r, r + yc)]
            else:
                pts += [(x, y, z) for y in range(-r, r + yc) for x in range(-r, r + xc)]
        return pts
    # 3D Star
    if orient == COLUMN:
        pts += [(x, y, z) for z in range(-r, 0) for x in range(0, xc) for y in range(0, yc)]
        for z in range(0, zc):
            pts += [(x, y, z) for x in range(-r, 0) for y in range(0, yc)]
            pts += [(x, y, z) for x in range(0, xc) for y in range(-r, r + yc)]
            pts += [(x, y, z) for x in range(xc, r + xc) for y in range(0, yc)]
        pts += [(x, y, z) for z in range(zc, r + zc) for x in range(0, xc) for y in range(0, yc)]
    else:
        pts += [(x, y, z) for z in range(-r, 0) for y in range(0, yc) for x in range(0, xc)]
        for z in range(0, zc):
            pts += [(x, y, z) for y in range(-r, 0) for x in range(0, xc)]
            pts += [(x, y, z) for y in range(0, yc) for x in range(-r, r + xc)]
            pts += [(x, y, z) for y in range(yc, r + yc) for x in range(0, xc)]
        pts += [(x, y, z) for z in range(zc, r + zc) for y in range(0, yc) for x in range(0, xc)]
    return pts


def _contributes(dim, nb, r, p, e):
    """printAdditions' membership test for grid point p and result e."""
    d = [p[i] - e[i] for i in range(3)]
    if nb == BOX:
        return all(abs(v) <= r for v in d[:dim])
    if dim == 2:
        return (abs(d[0]) <= r and d[1] == 0) or (abs(d[1]) <= r and d[0] == 0)
    return ((abs(d[0]) <= r and d[1] == 0 and d[2] == 0) or (abs(d[1]) <= r and d[0] == 0 and d[2] == 0)
            or (abs(d[2]) <= r and d[0] == 0 and d[1] == 0))


def _is_init(dim, nb, orient, r, p, e):
    """printAdditions' initialise condition (first contributor)."""
    x, y, z = p
    if nb == BOX:
        return x == e[0] - r and y == e[1] - r and (dim == 2 or z == e[2] - r)
    if dim == 2:
        if orient == COLUMN:
            return y == e[1] and x == e[0] - r
        return x == e[0] and y == e[1] - r
    return x == e[0] and y == e[1] and z == e[2] - r


def _replay(dim, nb, orient, r, C):
    """Per result register: list of (offset, is_init) in emission order."""
    zc = C[2] if dim == 3 else 1
    results = [(i, j, k) for k in range(zc) for j in range(C[1]) for i in range(C[0])]
    seq = {e: [] for e in results}
    for p in _traversal(dim, nb, orient, r, C):
        for e in results:
            if _contributes(dim, nb, r, p, e):
                off = (p[0] - e[0], p[1] - e[1], p[2] - e[2] if dim == 3 else 0)
                seq[e].append((off, _is_init(dim, nb, orient, r, p, e)))
    return seq


CASES = [(dim, nb, orient, r, C)
         for dim, nb, orient in itertools.product((2, 3), (BOX, STAR), (ROW, COLUMN))
         for r, C in ((1, (1, 1, 1)), (2, (3, 2, 2)), (3, (1, 4, 3)), (1, (4, 1, 5)))]


@pytest.mark.parametrize("dim,nb,orient,r,C", CASES)
def test_order_matches_reference_loop_nests(dim, nb, orient, r, C):
    off, midx = oracle.offsets(dim, nb, orient, r)
    canon = [tuple(int(v) for v in o) for o in off]
    for e, seq in _replay(dim, nb, orient, r, C).items():
        offs = [s[0] for s in seq]
        assert offs == canon, f"result {e}: reference order differs from oracle order"
        inits = [s[1] for s in seq]
        assert inits[0] and not any(inits[1:]), f"result {e}: first contributor is not the initialiser"


@pytest.mark.parametrize("dim,nb", list(itertools.product((2, 3), (BOX, STAR))))
def test_mask_layout_is_dzdydx_rank(dim, nb):
    """Mask index = rank of the offset in (dz,dy,dx) lexicographic order over
    the neighbourhood, for both orientations."""
    for r in (1, 2, 3):
        for orient in (ROW, COLUMN):
            off, midx = oracle.offsets(dim, nb, orient, r)
            pts = sorted((tuple(int(v) for v in o) for o in off), key=lambda o: (o[2], o[1], o[0]))
            rank = {p: i for i, p in enumerate(pts)}
            for o, m in zip(off, midx):
                assert rank[tuple(int(v) for v in o)] == m
            assert sorted(midx.tolist()) == list(range(len(pts)))


def test_stencil_sizes():
    assert oracle.stencil_size(2, BOX, 1) == 9
    assert oracle.stencil_size(3, STAR, 4) == 25
    assert oracle.stencil_size(2, BOX, 8) == 289
    assert oracle.stencil_size(3, BOX, 2) == 125
    assert oracle.stencil_size(2, STAR, 3) == 13


def test_product_spec_matches_oracle():
    """The generator's own offset/mask tables agree with the oracle's."""
    from threadcoarsenedstencilgenerator_amd import spec
    for dim, nb, orient in itertools.product((2, 3), (BOX, STAR), (ROW, COLUMN)):
        for r in (1, 2, 4):
            d = spec.StencilDesc(dimension=dim, neighbourhood=nb, orientation=orient, range=r,
                                 zc=1)
            off, midx = oracle.offsets(dim, nb, orient, r)
            mine = spec.offsets(d)
            assert [tuple(int(v) for v in o) for o in off] == mine
            assert [spec.mask_index(d, *o) for o in mine] == midx.tolist()
