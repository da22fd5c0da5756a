"""Benchmark and test cases (synthetic inputs; no reference data needed).

cylinder2d: inviscid flow past a circular cylinder (radius 0.5, axis z) in
  [-10,10]^2 x [0, 4h], h = 20/n, periodic in z, far field on x/y faces;
  M = 0.2, alpha = 0; JST k2 = 0.5, k4 = 1/32; 5-stage RK, CFL 2.5 (local
  time step); smoothing eps = 0.6; W-cycle.  n = 128 is the 128x128x4
  three-level case of the measurement plan.
sphere3d: sphere radius 0.5 in [-8,8]^3 with far field on all faces.
ellipsoid_sweep: the batched-throughput sweep of seeded random ellipsoids.
"""
import math

import numpy as np

from .grid import body, geometric_vertices, hierarchy, stretched_hierarchy
from .params import Params

FAR, SYM, PER, WALL = -20, -50, -60, -80


def cylinder2d(n=128, nz=4, nlevel=3, mach=0.2, alpha=0.0, half=10.0, **kw):
    h = 2.0 * half / n
    lo = (-half, -half, 0.0)
    hi = (half, half, nz * h)
    levels = hierarchy((n, n, nz), lo, hi, nlevel, bctype=(FAR, FAR, PER, FAR, FAR, PER))
    prm = Params(mach=mach, alpha_deg=alpha, **kw)
    bdy = body("cylinder", center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5), axis=2)
    ref_area = 1.0 * (nz * h)  # diameter x span
    return prm, levels, bdy, ref_area


def cylinder2d_stretched(n=512, nz=4, nlevel=4, mach=0.1, alpha=0.0, half=20.0, ratio=1.02, **kw):
    """The laminar-cylinder grid: n x n x nz cells on [-half, half]^2 whose
    widths grow geometrically (`ratio`) away from the body in x and y, z
    uniform with the mean x spacing and periodic; far field on x/y faces.
    Laminar Navier-Stokes at Re = 40, Pr = 0.72 unless overridden."""
    xv = geometric_vertices(n, half, ratio)
    hz = 2.0 * half / n
    zv = hz * np.arange(nz + 1, dtype=np.float64)
    levels = stretched_hierarchy((xv, xv.copy(), zv), nlevel, bctype=(FAR, FAR, PER, FAR, FAR, PER))
    kw.setdefault("eqntype", "n")
    kw.setdefault("reynolds", 40.0)
    prm = Params(mach=mach, alpha_deg=alpha, **kw)
    bdy = body("cylinder", center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5), axis=2)
    ref_area = 1.0 * (nz * hz)  # diameter x span
    return prm, levels, bdy, ref_area


def sphere3d(n=64, nlevel=3, mach=0.5, alpha=2.0, box=8.0, **kw):
    lo = (-box, -box, -box)
    hi = (box, box, box)
    levels = hierarchy((n, n, n), lo, hi, nlevel, bctype=(FAR,) * 6)
    prm = Params(mach=mach, alpha_deg=alpha, **kw)
    bdy = body("sphere", center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5))
    ref_area = math.pi * 0.25
    return prm, levels, bdy, ref_area


def ellipsoid3d(n=48, nlevel=3, semi=(0.6, 0.4, 0.3), rot=None, mach=0.5, alpha=2.0, box=8.0, **kw):
    lo = (-box, -box, -box)
    hi = (box, box, box)
    levels = hierarchy((n, n, n), lo, hi, nlevel, bctype=(FAR,) * 6)
    prm = Params(mach=mach, alpha_deg=alpha, **kw)
    bdy = body("ellipsoid", semi=semi, rot=rot)
    ref_area = math.pi * semi[0] * semi[1]
    return prm, levels, bdy, ref_area


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def quaternion_rotation(q):
    """rotation matrix of the unit quaternion q = (w, x, y, z)"""
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def ellipsoid_sweep(count=64, seed=20261202, nlevel=3, box=8.0, **kw):
    """`count` Euler cases from numpy PCG64(seed), drawn per case in the
    order: semi-axes a, b, c ~ U[0.2, 1.0]; a uniformly random rotation
    (unit quaternion of 4 normals); N ~ exp U[ln 48, ln 160] rounded to a
    multiple of 8 cells per edge; M ~ U[0.1, 0.8]; alpha ~ U[-5, 15] deg.
    Box [-box, box]^3, far field on every face, `nlevel` levels.  Returns a
    list of (params, levels, body, ref_area, description)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        semi = tuple(float(v) for v in rng.uniform(0.2, 1.0, size=3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        n = int(8 * round(math.exp(rng.uniform(math.log(48.0), math.log(160.0))) / 8.0))
        mach = float(rng.uniform(0.1, 0.8))
        alpha = float(rng.uniform(-5.0, 15.0))
        prm, levels, bdy, aref = ellipsoid3d(n=n, nlevel=nlevel, semi=semi, rot=quaternion_rotation(q), mach=mach,
                                             alpha=alpha, box=box, **kw)
        out.append((prm, levels, bdy, aref, dict(n=n, semi=semi, mach=mach, alpha=alpha)))
    return out
