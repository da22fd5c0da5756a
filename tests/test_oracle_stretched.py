"""Pins of the oracle on stretched tensor levels (cellvolume 0: each cell's
volume is the product of its vertex differences; no GPU).  The reference
block is uniform with one cellvolume, so this extension is defined by this
build: its volumes, transfers and fluxes are pinned here by what the
mathematics fixes -- volumes that tile the box, a preserved freestream,
telescoping on a periodic box, linear fields restricted and prolonged
exactly, the restricted residual summing to the fine one."""
import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import LevelSpec, geometric_vertices, level_geometry, stretched_hierarchy
from code_mge_amd.params import Params
from oracle.oracle import OracleCase


def _levels(n=(16, 12, 8), ratio=1.1, nlevel=3, bctype=(-20,) * 6):
    xs = (geometric_vertices(n[0], 1.0, ratio), 0.5 * geometric_vertices(n[1], 1.0, ratio ** 0.5),
          0.1 * np.arange(n[2] + 1, dtype=np.float64) ** 1.2)
    return stretched_hierarchy(xs, nlevel, bctype)


def _centres(spec):
    """centres of cells 0..ib (cells 0 and ib, which have no second vertex,
    continue the spacing of their neighbours)"""
    out = []
    for v in spec.vertices():
        c = np.empty(len(v))
        c[1:-1] = 0.5 * (v[1:-1] + v[2:])
        c[0] = 2.0 * c[1] - c[2]
        c[-1] = 2.0 * c[-2] - c[-3]
        out.append(c)
    return out


def test_volumes_tile_the_box():
    levels = _levels()
    prm = Params(mach=0.3)
    oc = OracleCase(prm, [level_geometry(s, None) for s in levels])
    for li, s in enumerate(levels):
        vol = oc.download(li + 1, "vol")[2:-2, 2:-2, 2:-2]
        dx, dy, dz = (np.diff(v)[2:-1] for v in s.vertices())
        assert np.array_equal(vol, (dz[:, None, None] * 0 + (dx[None, None, :] * dy[None, :, None]) * dz[:, None, None]))
        box = np.prod([s.xs[a][-1] - s.xs[a][0] for a in range(3)])
        assert abs(vol.sum() - box) < 1e-13 * box
    oc.close()


def test_stretched_freestream_preserved():
    """no body, alpha and beta non-zero, JST with smoothing, W-cycles on
    three stretched levels: the freestream stays (every face pair of a
    tensor cell cancels, every difference vanishes, transfer weights sum
    to one)"""
    levels = _levels()
    prm = Params(mach=0.4, alpha_deg=10.0, beta_deg=5.0)
    oc = OracleCase(prm, [level_geometry(s, None) for s in levels], mxcyc=10)
    oc.init_flow()
    w0 = oc.download(3, "w")
    log, _ = oc.solve([0, 1, 3])
    assert np.max(np.abs(oc.download(3, "w") - w0)) < 1e-13 * np.abs(w0).max()
    assert np.all(log[:, 2] < 1e-12)
    oc.close()


def test_stretched_periodic_box_telescopes():
    """a random state on a periodic stretched box: the convective residual
    of all cells sums to zero (every face vector is shared by its two cells;
    the JST dissipation's end conventions -- dflux.f90:78-86 and the clamped
    ghost radii -- do not telescope across a periodic wrap on any grid)"""
    levels = _levels(nlevel=1, bctype=(-60,) * 6)
    prm = Params(mach=0.3, alpha_deg=20.0, beta_deg=10.0, smoo=0.0, vis2=0.0, vis4=0.0)
    d = prm.derived()
    oc = OracleCase(prm, [level_geometry(levels[0], None)], mxcyc=4)
    n = levels[0].n
    shape = (n[2] + 4, n[1] + 4, n[0] + 4)
    rng = np.random.default_rng(4)
    rho = 1.0 + 0.1 * rng.uniform(-1, 1, shape)
    u, v, w = (d[k] + 0.1 * rng.uniform(-1, 1, shape) for k in ("u0", "v0", "w0"))
    p = d["p0"] * (1.0 + 0.1 * rng.uniform(-1, 1, shape))
    st = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
    oc.upload(1, "w", st)
    oc.upload(1, "p", p)
    oc.bcond(1)
    oc.timestep(1)
    oc.euler(1, 1, 1, 1, lres=True, lfir=False)
    dw = oc.download(1, "dw")[:, 2:-2, 2:-2, 2:-2]
    for c in range(5):
        assert abs(dw[c].sum()) < 1e-13 * np.abs(dw[c]).sum(), c
    oc.close()


def test_stretched_transfers_exact_on_linear_fields():
    """restriction (movco) of a linear field gives its values at the coarse
    centres; the restricted residual sums to the fine residual; the
    full-multigrid prolongation (movfin) of a linear field gives its values
    at the fine centres"""
    levels = _levels(nlevel=2)
    prm = Params(mach=0.3)
    oc = OracleCase(prm, [level_geometry(s, None) for s in levels], mxcyc=4)
    oc.init_flow()
    a = np.array([[1.0, 0.1, -0.2, 0.05], [0.2, 0.3, 0.1, -0.1], [-0.1, 0.05, 0.2, 0.1],
                  [0.05, -0.1, 0.1, 0.2], [40.0, 1.0, 2.0, -1.5]])

    def lin(spec):
        xc, yc, zc = _centres(spec)
        Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
        return np.stack([a[c, 0] + a[c, 1] * X + a[c, 2] * Y + a[c, 3] * Z for c in range(5)])
    fine, coarse = levels[1], levels[0]
    oc.upload(2, "w", lin(fine))
    rng = np.random.default_rng(8)
    dwf = rng.uniform(-1, 1, oc.download(2, "dw").shape)
    oc.upload(2, "dw", dwf)
    oc.movco(1)
    s = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    wc = oc.download(1, "w")
    assert np.allclose(wc[s], lin(coarse)[s], rtol=1e-13, atol=1e-13)
    wr = oc.download(1, "wr")
    for c in range(5):
        assert abs(wr[s][c].sum() - dwf[s][c].sum()) < 1e-12 * np.abs(dwf[s][c]).sum()
    # full-multigrid step: the fine level takes the interpolated coarse solution
    oc.upload(1, "w", lin(coarse))
    oc.movfin(2, 1)
    assert np.allclose(oc.download(2, "w")[s], lin(fine)[s], rtol=1e-13, atol=1e-13)
    oc.close()


def test_stretched_laminar_cylinder_converges_symmetric():
    """the laminar cylinder on a geometrically clustered grid (reduced
    size): W-cycles reduce the residual and keep the lift of the symmetric
    flow at rounding level (smoothing off: psmoo_old reuses full-line pivots
    after a blank cell, which breaks the mirror symmetry on any grid)"""
    prm, levels, bdy, aref = configs.cylinder2d_stretched(n=64, nz=4, nlevel=3, ratio=1.05, half=10.0,
                                                          mach=0.2, smoo=0.0)
    oc = OracleCase(prm, [level_geometry(s, bdy) for s in levels], mxcyc=80)
    oc.init_flow()
    log, _ = oc.solve([0, 0, 60])
    r = log[log[:, 0] == 3][:, 2]
    assert r[-1] < 0.05 * r[0], (r[0], r[-1])
    F = oc.wall_force(3)
    assert abs(F[1]) < 1e-8 * abs(F[0]), F
    oc.close()
