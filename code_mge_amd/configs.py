"""Benchmark and test cases (synthetic inputs; no reference data needed).

cylinder2d: inviscid flow past a circular cylinder (radius 0.5, axis z) in
  [-10,10]^2 x [0, 4h], h = 20/n, periodic in z, far field on x/y faces;
  M = 0.2, alpha = 0; JST k2 = 0.5, k4 = 1/32; 5-stage RK, CFL 2.5 (local
  time step); smoothing eps = 0.6; W-cycle.  n = 128 is the 128x128x4
  three-level case of the measurement plan.
sphere3d: sphere radius 0.5 in [-8,8]^3 with far field on all faces.
"""
import math

import numpy as np

from .grid import body, hierarchy
from .params import Params

FAR, SYM, PER, WALL = -20, -50, -60, -80


def cylinder2d(n=128, nz=4, nlevel=3, mach=0.2, alpha=0.0, **kw):
    h = 20.0 / n
    lo = (-10.0, -10.0, 0.0)
    hi = (10.0, 10.0, nz * h)
    levels = hierarchy((n, n, nz), lo, hi, nlevel, bctype=(FAR, FAR, PER, FAR, FAR, PER))
    prm = Params(mach=mach, alpha_deg=alpha, **kw)
    bdy = body("cylinder", center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5), axis=2)
    ref_area = 1.0 * (nz * h)  # diameter x span
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
