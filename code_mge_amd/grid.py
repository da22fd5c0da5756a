"""Cartesian level hierarchy and the cut-cell geometry of each level.

Each multigrid level is one block covering the whole domain (the
configuration every benchmark case uses).  Vertices follow
init_metric.f90:59-72: x(i) = h*(i-2) + origin for i = 1..ib, with
x(0) left at 0.  Coarsening halves a direction while it keeps >= 2 cells
(psmoo_old needs two cells per line, psmoo.f90:132,166), otherwise the
direction is kept (semi-coarsening).

A stretched tensor level (the reference block is uniform, with one
cellvolume) gives its interior vertex coordinates per direction instead;
its ghost vertices repeat the nearest interior spacing, each cell's volume
is the product of its vertex differences (cellvolume() returns 0 to say
so), and a coarser level takes every other vertex.
"""
import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _ffi


@dataclass
class LevelSpec:
    n: tuple               # cells (ni, nj, nk)
    lo: tuple              # domain lower corner
    hi: tuple              # domain upper corner
    ratio: tuple           # coarsening ratio to the next coarser level
    bctype: tuple          # faces 1..6: i-lo, j-lo, k-lo, i-hi, j-hi, k-hi
    xs: Optional[tuple] = None   # stretched level: interior vertices x(2..n+2) per direction

    @property
    def stretched(self):
        return self.xs is not None

    def vertices(self):
        out = []
        for a in range(3):
            n = self.n[a]
            v = np.zeros(n + 4, dtype=np.float64)
            if self.xs is not None:
                xa = np.asarray(self.xs[a], dtype=np.float64)
                assert xa.shape == (n + 1,)
                v[2:n + 3] = xa
                v[1] = xa[0] - (xa[1] - xa[0])
                v[n + 3] = xa[n] + (xa[n] - xa[n - 1])
            else:
                h = (self.hi[a] - self.lo[a]) / n
                idx = np.arange(1, n + 4, dtype=np.float64)
                v[1:] = h * (idx - 2.0) + self.lo[a]
            out.append(v)
        return out

    def spacing(self):
        return tuple((self.hi[a] - self.lo[a]) / self.n[a] for a in range(3))

    def cellvolume(self):
        """the block's one cell volume; 0 for a stretched level, whose cells
        each take the product of their vertex differences"""
        if self.xs is not None:
            return 0.0
        hx, hy, hz = self.spacing()
        return hx * hy * hz

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]


def hierarchy(n_fine, lo, hi, nlevel, bctype, min_cells=2) -> List[LevelSpec]:
    """Levels coarsest first (index 0 = reference level 1)."""
    specs = []
    n = list(n_fine)
    for _ in range(nlevel):
        specs.append([tuple(n)])
        nxt = []
        for a in range(3):
            nxt.append(n[a] // 2 if (n[a] % 2 == 0 and n[a] // 2 >= min_cells) else n[a])
        n = nxt
    levels = []
    for li in range(nlevel):
        nn = specs[li][0]
        if li + 1 < nlevel:
            nc = specs[li + 1][0]
            ratio = tuple(nn[a] // nc[a] for a in range(3))
        else:
            ratio = (2, 2, 2)
        levels.append(LevelSpec(n=nn, lo=tuple(lo), hi=tuple(hi), ratio=ratio, bctype=tuple(bctype)))
    levels.reverse()
    return levels


def stretched_hierarchy(xs_fine, nlevel, bctype, min_cells=2) -> List[LevelSpec]:
    """Levels coarsest first for a stretched tensor grid given by its finest
    interior vertices per direction; a direction is halved (every other
    vertex) while it keeps >= min_cells cells, like hierarchy()."""
    xs = [np.asarray(x, dtype=np.float64) for x in xs_fine]
    specs = []
    for _ in range(nlevel):
        specs.append(tuple(xs))
        nxt = []
        for a in range(3):
            n = len(xs[a]) - 1
            nxt.append(xs[a][::2] if (n % 2 == 0 and n // 2 >= min_cells) else xs[a])
        xs = nxt
    levels = []
    for li in range(nlevel):
        cur = specs[li]
        n = tuple(len(x) - 1 for x in cur)
        if li + 1 < nlevel:
            ratio = tuple(n[a] // (len(specs[li + 1][a]) - 1) for a in range(3))
        else:
            ratio = (2, 2, 2)
        levels.append(LevelSpec(n=n, lo=tuple(float(x[0]) for x in cur), hi=tuple(float(x[-1]) for x in cur),
                                ratio=ratio, bctype=tuple(bctype), xs=cur))
    levels.reverse()
    return levels


def geometric_vertices(n, half, ratio):
    """2*(n//2) ... n cells on [-half, half] whose widths grow by `ratio`
    from the centre outward (symmetric about 0); n even"""
    m = n // 2
    w = ratio ** np.arange(m, dtype=np.float64)
    w *= half / w.sum()
    right = np.concatenate([[0.0], np.cumsum(w)])
    right[-1] = half
    return np.concatenate([-right[:0:-1], right])


def body(kind, center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5), axis=2, rot=None):
    b = _ffi.MgeBody()
    b.kind = {"none": 0, "cylinder": 1, "sphere": 2, "ellipsoid": 3}[kind] if isinstance(kind, str) else kind
    b.axis = axis
    for a in range(3):
        b.center[a] = center[a]
        b.semi[a] = semi[a]
    R = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64)
    for r in range(3):
        for s in range(3):
            b.rot[r * 3 + s] = R[r, s]
    return b


class TriBody:
    """a closed triangulated surface: tri is (ntri, 9) float64, the three
    vertices of every triangle (e.g. the quads of PLOT3D surface patches,
    each split in two)"""

    def __init__(self, tri):
        self.tri = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(-1, 9))

    @property
    def ntri(self):
        return self.tri.shape[0]


@dataclass
class LevelGeom:
    spec: LevelSpec
    xv: np.ndarray
    yv: np.ndarray
    zv: np.ndarray
    blank: np.ndarray       # (nk, nj, ni) uint8, 1 fluid
    recs: object            # ctypes array of MgeCutRec
    nmir: int


def level_geometry(spec: LevelSpec, bdy) -> LevelGeom:
    """Run the geometry preprocessor (host code of libmge) for one level;
    bdy is an analytic body (MgeBody), a TriBody or None."""
    L = _ffi.lib()
    xv, yv, zv = spec.vertices()
    ni, nj, nk = spec.n
    blank = np.ones((nk, nj, ni), dtype=np.uint8)
    nm = C.c_int(0)
    dp = C.POINTER(C.c_double)
    args = (ni, nj, nk, xv.ctypes.data_as(dp), yv.ctypes.data_as(dp), zv.ctypes.data_as(dp),
            blank.ctypes.data_as(C.c_void_p))
    if isinstance(bdy, TriBody):
        def call(recs, cap):
            return L.mge_geom_level_tri(bdy.tri.ctypes.data_as(dp), bdy.ntri, *args, recs, cap, C.byref(nm))
        what = "mge_geom_level_tri"
    else:
        bptr = C.byref(bdy) if bdy is not None else None

        def call(recs, cap):
            return L.mge_geom_level(bptr, *args, recs, cap, C.byref(nm))
        what = "mge_geom_level"
    _ffi.check(call(None, 0), what)
    recs = (_ffi.MgeCutRec * max(nm.value, 1))()
    _ffi.check(call(recs, nm.value), what)
    return LevelGeom(spec, xv, yv, zv, blank, recs, nm.value)
