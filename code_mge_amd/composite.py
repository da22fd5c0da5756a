"""Embedded multi-block hierarchies (composite grids) on the HIP path
(mge_comp_*): several blocks per multigrid level, each finer level covering
part of the coarser one, as the reference's embedded grids
(bc_embed.f90:50-143, bc_overlap.f90:32-90, collc.f90:58-146,
addw.f90:84-188, the block loops of solver.f90:119-261).

`embedded_hierarchy` turns the blocks of a .pin file (GRID.NAME = LNN: level
L, serial NN) into block specs with their composite level, parent and
lattice position; `Composite` runs the solve over one-level `Case`s built
from them."""
import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import _ffi
from .grid import LevelSpec

_dp = C.POINTER(C.c_double)


@dataclass
class BlockSpec:
    name: int
    level: int            # 1 = coarsest
    parent: int           # index of the block of level - 1 holding it (-1 on level 1)
    g0: tuple             # index of the block's cell 2 on its level's lattice
    spec: LevelSpec       # the block as a one-level case


def embedded_hierarchy(pin_blocks, dom_lo, dom_hi, dom_bctype, tol=1e-9) -> List[BlockSpec]:
    """Block specs of a .pin's grids (dicts of inputs.read_pin).  Level-1
    grids fix the lattice origin and spacing; every level halves the
    spacing; each block's parent is the block of the next coarser level
    containing it.  A face on the domain boundary takes the domain's
    condition (dom_bctype, faces i-lo, j-lo, k-lo, i-hi, j-hi, k-hi), any
    other face is embedded (bctype 0)."""
    blocks = sorted(pin_blocks, key=lambda g: g["name"])
    lev = [g["name"] // 100 for g in blocks]
    if min(lev) != 1:
        raise ValueError("a hierarchy needs a level-1 grid")
    base = [g for g, l in zip(blocks, lev) if l == 1][0]
    h1 = np.array([base["length"][a] / (base["points"][a] - 1) for a in range(3)])
    org = np.array(base["origin"], dtype=float)
    out = []
    for g, L in zip(blocks, lev):
        n = tuple(int(p) - 1 for p in g["points"])
        h = h1 / 2.0 ** (L - 1)
        hb = np.array([g["length"][a] / n[a] for a in range(3)])
        if np.max(np.abs(hb - h) / h) > tol:
            raise ValueError(f"grid {g['name']}: spacing {hb} is not the level-{L} lattice's {h}")
        x0 = (np.array(g["origin"]) - org) / h
        g0 = tuple(int(round(v)) for v in x0)
        if np.max(np.abs(x0 - np.array(g0))) > 1e-6:
            raise ValueError(f"grid {g['name']}: origin off the level-{L} lattice")
        lo = tuple(float(v) for v in g["origin"])
        hi = tuple(float(g["origin"][a] + g["length"][a]) for a in range(3))
        bct = []
        for f in range(6):
            a = f % 3
            on = abs((lo[a] if f < 3 else hi[a]) - (dom_lo[a] if f < 3 else dom_hi[a])) <= 1e-9 * max(1.0, abs(hi[a]))
            bct.append(int(dom_bctype[f]) if on else 0)
        out.append(BlockSpec(g["name"], L, -1, g0, LevelSpec(n=n, lo=lo, hi=hi, ratio=(2, 2, 2), bctype=tuple(bct))))
    for i, b in enumerate(out):
        if b.level == 1:
            continue
        for j, p in enumerate(out):
            if p.level != b.level - 1:
                continue
            inside = all(2 * p.g0[a] <= b.g0[a] and b.g0[a] + b.spec.n[a] <= 2 * (p.g0[a] + p.spec.n[a])
                         for a in range(3))
            if inside:
                b.parent = j
                break
        if b.parent < 0:
            raise ValueError(f"grid {b.name}: no grid of level {b.level - 1} holds it")
    return out


class Composite:
    """mge_comp over one-level cases (the blocks, in the order of `blocks`);
    while it lives the cases run on its stream and are not used alone."""

    def __init__(self, cases: Sequence, blocks: Sequence[BlockSpec]):
        self.lib = _ffi.lib()
        self.cases = list(cases)
        self.blocks = list(blocks)
        nb = len(cases)
        hs = (C.c_void_p * nb)(*[c.h for c in cases])
        lev = (C.c_int * nb)(*[b.level for b in blocks])
        par = (C.c_int * nb)(*[b.parent for b in blocks])
        g0 = (C.c_int * (3 * nb))(*[v for b in blocks for v in b.g0])
        h = C.c_void_p()
        _ffi.check(self.lib.mge_comp_create(nb, hs, lev, par, g0, C.byref(h)), "mge_comp_create")
        self.h = h
        self.nlevel = max(b.level for b in blocks)

    def solve(self, mcyc: Sequence[int], cap: int = 100000):
        """solver.f90's FMG ramp over the composite: (level, cycle, residual)
        triples and the cell-stage count"""
        mc = (C.c_int * self.nlevel)(*mcyc)
        log = np.zeros(3 * cap, dtype=np.float64)
        nlog = C.c_int(0)
        cs = C.c_longlong(0)
        _ffi.check(self.lib.mge_comp_solve(self.h, mc, log.ctypes.data_as(_dp), cap, C.byref(nlog), C.byref(cs)),
                   "mge_comp_solve")
        n = min(nlog.value, cap)
        return log[: 3 * n].reshape(n, 3), cs.value

    def close(self):
        if self.h:
            self.lib.mge_comp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_embedded_job(jobsheet, pin):
    """The jobsheet's parameters, surface, FMG ramp and domain conditions
    (inputs.load_job) with the grids of a multi-block .pin: (params, block
    specs, surface body, mcyc, reference area)"""
    from .inputs import load_job, read_pin
    job = load_job(jobsheet)
    raw = job.raw
    blocks = embedded_hierarchy(read_pin(pin), tuple(raw.dom_lo), tuple(raw.dom_hi), tuple(raw.bctype))
    nlevel = max(b.level for b in blocks)
    mcyc = list(job.mcyc[:nlevel]) + [job.mcyc[-1]] * max(0, nlevel - len(job.mcyc))
    return job.params, blocks, job.body, mcyc, job.ref_area
