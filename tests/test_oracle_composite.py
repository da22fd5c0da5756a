"""The oracle's embedded multi-block restatement (orc_comp_*), pinned by
what the mathematics fixes and by the single-case solve:

- bc_embed.f90:50-143, bc_overlap.f90:32-90, collc.f90:58-146 and
  addw.f90:84-188 with block offsets: a field linear in the cell centres
  moves between blocks of a three-block hierarchy exactly (the 1/4-3/4
  prolongation and the 8-cell mean reproduce linear fields), cells that no
  block covers keep their values, and each block receives data exactly where
  the lattice positions say it lies;
- the solver.f90:119-261 block loops: a two-block composite whose child
  covers its whole parent is the single-case two-level solve, bit for bit;
- the reference's seven-block cylinder hierarchy (cylinder05_11122.pin):
  block positions and parents worked out by hand from the .pin text.
"""
import os

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.composite import BlockSpec, embedded_hierarchy, load_embedded_job
from code_mge_amd.grid import LevelSpec, level_geometry
from code_mge_amd.inputs import read_pin
from code_mge_amd.params import Params
from oracle.oracle import OracleCase, comp_op, comp_solve

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cylinder")
FAR = configs.FAR
PRM = Params(mach=0.3, alpha_deg=0.0)

# level 1: 12 x 10 x 8 unit cells on [0,12]x[0,10]x[0,8]; level 2 (h = 1/2):
# A at lattice (4,2,2), 8x8x6 cells, and B at (10,4,2), 8x6x6 cells, both
# inside the level-1 block and overlapping in lattice i = 10..11
H1 = 1.0
P_N, A_N, B_N = (12, 10, 8), (8, 8, 6), (8, 6, 6)
A_G0, B_G0 = (4, 2, 2), (10, 4, 2)


def _spec(g0, n, h, bct):
    lo = tuple(g0[a] * h for a in range(3))
    hi = tuple((g0[a] + n[a]) * h for a in range(3))
    return LevelSpec(n=n, lo=lo, hi=hi, ratio=(2, 2, 2), bctype=bct)


def _hierarchy():
    blocks = [BlockSpec(1, 1, -1, (0, 0, 0), _spec((0, 0, 0), P_N, H1, (FAR,) * 6)),
              BlockSpec(201, 2, 0, A_G0, _spec(A_G0, A_N, H1 / 2, (0,) * 6)),
              BlockSpec(202, 2, 0, B_G0, _spec(B_G0, B_N, H1 / 2, (0,) * 6))]
    cases = [OracleCase(PRM, [level_geometry(b.spec, None)], mxcyc=4) for b in blocks]
    for c in cases:
        c.init_flow()
    return blocks, cases


COEF = np.array([[1.0, 0.010, -0.020, 0.005],     # rho
                 [0.3, 0.004, 0.002, -0.001],     # rho u
                 [0.05, -0.003, 0.001, 0.002],    # rho v
                 [0.02, 0.001, -0.002, 0.003],    # rho w
                 [2.5, 0.020, 0.010, -0.015]])    # rho E


def _lattice(b, h):
    """global lattice index and cell-centre coordinates (i fastest) of every
    cell of a block, ghosts included; cell 2 is lattice position g0"""
    n = b.spec.n
    idx = [b.g0[a] + np.arange(n[a] + 4) - 2 for a in range(3)]
    X = [(idx[a] + 0.5) * h for a in range(3)]
    G = np.meshgrid(*idx[::-1], indexing="ij")[::-1]
    Z, Y, Xc = np.meshgrid(X[2], X[1], X[0], indexing="ij")
    return G, (Xc, Y, Z)


def _linear(xyz, coef=COEF):
    x, y, z = xyz
    return np.stack([c[0] + c[1] * x + c[2] * y + c[3] * z for c in coef])


def _eos(w, gamma):
    qq = 0.5 * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2) / w[0]
    return (gamma - 1.0) * (w[4] - qq)


def _interior(b):
    n = b.spec.n
    m = np.zeros((n[2] + 4, n[1] + 4, n[0] + 4), dtype=bool)
    m[2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 2] = True
    return m


def _inside(G, g0, n):
    return np.logical_and.reduce([(G[a] >= g0[a]) & (G[a] < g0[a] + n[a]) for a in range(3)])


def test_embed_linear_field_with_offsets():
    blocks, cases = _hierarchy()
    _, xp = _lattice(blocks[0], H1)
    cases[0].upload(1, "w", _linear(xp))
    sentinel = [c.download(1, "w") for c in cases]
    comp_op(cases, blocks, "embed", 2)
    for b, c, s0 in zip(blocks[1:], cases[1:], sentinel[1:]):
        G, xf = _lattice(b, H1 / 2)
        w = c.download(1, "w")
        ghost = ~_interior(b)
        # a fine ghost cell takes the parent's data when its centre lies in
        # a physical cell of the parent (level-1 cells 0..11 in lattice units)
        inpar = _inside([g // 2 for g in G], (0, 0, 0), P_N) & ghost
        assert inpar.sum() == ghost.sum()            # both children lie well inside the parent
        np.testing.assert_allclose(w[:, inpar], _linear(xf)[:, inpar], rtol=1e-14, atol=1e-15)
        assert np.array_equal(w[:, ~ghost], s0[:, ~ghost])
        p = c.download(1, "p")
        np.testing.assert_allclose(p[inpar], _eos(w, PRM.gamma)[inpar], rtol=1e-13)
    for c in cases:
        c.close()


def test_embed_leaves_cells_outside_the_parent():
    """a child on the parent's lower corner: its ghost cells beyond the
    parent's cells are not embedded"""
    blocks = [BlockSpec(1, 1, -1, (0, 0, 0), _spec((0, 0, 0), P_N, H1, (FAR,) * 6)),
              BlockSpec(201, 2, 0, (0, 0, 0), _spec((0, 0, 0), (6, 4, 4), H1 / 2, (FAR, FAR, FAR, 0, 0, 0)))]
    cases = [OracleCase(PRM, [level_geometry(b.spec, None)], mxcyc=4) for b in blocks]
    for c in cases:
        c.init_flow()
    _, xp = _lattice(blocks[0], H1)
    cases[0].upload(1, "w", _linear(xp))
    s0 = cases[1].download(1, "w")
    comp_op(cases, blocks, "embed", 2)
    G, xf = _lattice(blocks[1], H1 / 2)
    w = cases[1].download(1, "w")
    ghost = ~_interior(blocks[1])
    inpar = np.logical_and.reduce([G[a] >= 0 for a in range(3)]) & ghost
    assert 0 < inpar.sum() < ghost.sum()
    np.testing.assert_allclose(w[:, inpar], _linear(xf)[:, inpar], rtol=1e-14, atol=1e-15)
    assert np.array_equal(w[:, ~inpar], s0[:, ~inpar])
    for c in cases:
        c.close()


def test_overlap_copies_from_the_neighbour():
    blocks, cases = _hierarchy()
    before = []
    for b, c in zip(blocks[1:], cases[1:]):
        G, xf = _lattice(b, H1 / 2)
        w = _linear(xf)
        w[:, ~_interior(b)] = _linear(xf, COEF * 1.5)[:, ~_interior(b)]   # ghosts: another field
        c.upload(1, "w", w)
        before.append(w)
    comp_op(cases, blocks, "overlap", 2)
    for k, (b, c) in enumerate(zip(blocks[1:], cases[1:])):
        other = blocks[2 - k]
        G, xf = _lattice(b, H1 / 2)
        w = c.download(1, "w")
        cov = _inside(G, other.g0, other.spec.n) & ~_interior(b)
        assert cov.sum() > 0
        np.testing.assert_array_equal(w[:, cov], _linear(xf)[:, cov])
        assert np.array_equal(w[:, ~cov], before[k][:, ~cov])
        p = c.download(1, "p")
        np.testing.assert_allclose(p[cov], _eos(w, PRM.gamma)[cov], rtol=1e-13)
    for c in cases:
        c.close()


def test_restriction_with_offsets():
    """movco: the parent's cells under a child take the 8-cell mean of the
    child's state (exact for a linear field) and its residual forcing from
    the mean of dw/vol times the coarse volume; other cells keep theirs"""
    blocks, cases = _hierarchy()
    for b, c in zip(blocks[1:], cases[1:]):
        _, xf = _lattice(b, H1 / 2)
        c.upload(1, "w", _linear(xf))
        c.upload(1, "dw", _linear(xf, COEF[::-1]) * (H1 / 2) ** 3)
    w0 = cases[0].download(1, "w")
    wr0 = cases[0].download(1, "wr")
    comp_op(cases, blocks, "movco", 1)
    G, xp = _lattice(blocks[0], H1)
    cov = np.zeros_like(G[0], dtype=bool)
    for b in blocks[1:]:
        cov |= _inside(G, tuple(g // 2 for g in b.g0), tuple(n // 2 for n in b.spec.n))
    w = cases[0].download(1, "w")
    wr = cases[0].download(1, "wr")
    np.testing.assert_allclose(w[:, cov], _linear(xp)[:, cov], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(wr[:, cov], (_linear(xp, COEF[::-1]) * H1 ** 3)[:, cov], rtol=1e-13, atol=1e-15)
    rest = ~cov & _interior(blocks[0])
    assert rest.sum() > 0
    assert np.array_equal(w[:, rest], w0[:, rest]) and np.array_equal(wr[:, rest], wr0[:, rest])
    np.testing.assert_array_equal(cases[0].download(1, "wn")[:, _interior(blocks[0])], w[:, _interior(blocks[0])])
    for c in cases:
        c.close()


@pytest.mark.parametrize("fmg", [True, False])
def test_prolongation_with_offsets(fmg):
    """movfin: after the FMG step a child holds the parent's linear field at
    its own centres everywhere (interior by prolongation, ghosts by embedding
    and overlap); inside a cycle it adds the prolonged correction w - ws"""
    blocks, cases = _hierarchy()
    _, xp = _lattice(blocks[0], H1)
    lin = _linear(xp)
    corr = COEF.copy()
    corr[:, 0] = 0.0
    corr *= 0.1
    cases[0].upload(1, "w", lin)
    cases[0].upload(1, "ws", lin - _linear(xp, corr))
    old = [c.download(1, "w") for c in cases[1:]]
    comp_op(cases, blocks, "movfin", 2, mlvl=1 if fmg else 2)
    for b, c, w0 in zip(blocks[1:], cases[1:], old):
        _, xf = _lattice(b, H1 / 2)
        w = c.download(1, "w")
        if fmg:
            # bcond's bc_corner (bcond.f90:42) then rewrites the eight corner
            # cells of the first ghost layer by (a+b+c)*2/3 - d of their
            # neighbours, which no linear field satisfies
            m = np.ones(w.shape[1:], dtype=bool)
            ie, je, ke = b.spec.n[0] + 2, b.spec.n[1] + 2, b.spec.n[2] + 2
            for k in (1, ke):
                for j in (1, je):
                    for i in (1, ie):
                        m[k, j, i] = False
            np.testing.assert_allclose(w[:, m], _linear(xf)[:, m], rtol=1e-14, atol=1e-15)
        else:
            m = _interior(b)
            np.testing.assert_allclose(w[:, m], (w0 + _linear(xf, corr))[:, m], rtol=1e-14, atol=1e-15)
    for c in cases:
        c.close()


def test_full_coverage_composite_is_the_single_case():
    prm, levels, bdy, _ = configs.cylinder2d(n=32, nz=4, nlevel=2)
    geoms = [level_geometry(s, bdy) for s in levels]
    one = OracleCase(prm, geoms, mxcyc=20)
    one.init_flow()
    ls, css = one.solve([2, 6])
    blocks = [BlockSpec(1, 1, -1, (0, 0, 0), levels[0]), BlockSpec(2, 2, 0, (0, 0, 0), levels[1])]
    cases = [OracleCase(prm, [g], mxcyc=20) for g in geoms]
    for c in cases:
        c.init_flow()
    lc, csc = comp_solve(cases, blocks, [2, 6])
    assert css == csc and np.array_equal(ls, lc)
    assert np.array_equal(cases[0].download(1, "w"), one.download(1, "w"))
    assert np.array_equal(cases[1].download(1, "w"), one.download(2, "w"))
    for c in cases + [one]:
        c.close()


def test_cylinder_11122_hierarchy():
    """lattice positions from the .pin origins: level-1 spacing 28/112 =
    0.25, origin (-10, 0, -12); grid 401 (level 4, h = 0.03125) at
    x (-1.625 + 10) / 0.03125 = 268, z (-1.75 + 12) / 0.03125 = 328; grid
    502 (level 5) at x 10.75 / 0.015625 = 688 = 2 x 344, the i-lo corner of
    grid 402, so 502 lies in 402 and 501 (576 + 116 <= 2 x (268 + 80)) in 401"""
    prm, blocks, body, mcyc, aref = load_embedded_job(os.path.join(GOLDEN, "jobsheet"),
                                                      os.path.join(GOLDEN, "cylinder05_11122.pin"))
    assert [b.name for b in blocks] == [101, 201, 301, 401, 402, 501, 502]
    assert [b.level for b in blocks] == [1, 2, 3, 4, 4, 5, 5]
    assert [b.parent for b in blocks] == [-1, 0, 1, 2, 2, 3, 4]
    assert [b.g0 for b in blocks] == [(0, 0, 0), (48, 0, 48), (128, 0, 144), (268, 0, 328), (344, 0, 328),
                                      (576, 0, 704), (688, 0, 704)]
    assert [b.spec.n for b in blocks] == [(112, 2, 96), (136, 4, 96), (128, 8, 96), (80, 16, 112), (80, 16, 112),
                                          (116, 32, 128), (128, 32, 128)]
    # level 1 keeps the domain's conditions; finer blocks only on the span faces
    assert blocks[0].spec.bctype == tuple(blocks[0].spec.bctype)
    for b in blocks[1:]:
        assert b.spec.bctype[0] == b.spec.bctype[2] == b.spec.bctype[3] == b.spec.bctype[5] == 0
        assert b.spec.bctype[1] == blocks[0].spec.bctype[1] and b.spec.bctype[4] == blocks[0].spec.bctype[4]
    assert len(mcyc) == 5


def test_hierarchy_rejects_off_lattice_grids():
    g = read_pin(os.path.join(GOLDEN, "cylinder05_11122.pin"))
    bad = [dict(x) for x in g]
    bad[3]["origin"] = (bad[3]["origin"][0] + 0.01, bad[3]["origin"][1], bad[3]["origin"][2])
    with pytest.raises(ValueError):
        embedded_hierarchy(bad, (-10.0, 0.0, -12.0), (18.0, 0.5, 12.0), (FAR,) * 6)
