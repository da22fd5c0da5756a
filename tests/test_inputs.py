"""The input files of the reference's cylinder case (jobsheet, the .pin
grid and the PLOT3D surface, kept as fixtures in tests/golden/cylinder):
the readers return what the files say, the library's parameter derivation
equals the Python host's bit for bit, and the triangulated surface gives
the same cut-cell geometry as the analytic r = 0.5 cylinder up to the
polygon's deviation from the circle."""
import ctypes as C
import math
import os

import numpy as np

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.inputs import job_params_c, load_job, read_p3d, read_pin

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cylinder")


def test_jobsheet_values():
    job = load_job(os.path.join(HERE, "jobsheet"))
    p = job.params
    assert job.name == "cylinder_ibm_re20_test"
    assert job.mcyc == [100, 100, 120, 150, 180]
    assert (p.mach, p.reynolds, p.gamma, p.prlmean) == (0.5, 200.0, 1.4, 0.712)
    assert (p.alpha_deg, p.beta_deg, p.cfl, p.vis2, p.vis4, p.smoo, p.mgtype) == (0.0, 0.0, -2.0, 1.0, 2.0, 0.0, 1)
    assert tuple(p.cstp) == (0.25, 0.16667, 0.375, 0.5, 1.0)
    assert tuple(p.cdis) == (1.0, 0.0, 0.56, 0.0, 0.44)
    assert (p.eqntype, p.distype, p.dissip, p.epsentr) == ("n", "c", "scalar", 0.05)
    # the single-block pin: 449 x 9 x 257 points from (-10, 0, -8), extent (28, 0.5, 16)
    fine = job.levels[-1]
    assert fine.n == (448, 8, 256)
    assert fine.lo == (-10.0, 0.0, -8.0) and fine.hi == (18.0, 0.5, 8.0)
    # coarsening halves while >= 2 cells: y stops at 2
    assert [lv.n for lv in job.levels] == [(28, 2, 16), (56, 2, 32), (112, 2, 64), (224, 4, 128), (448, 8, 256)]
    assert [lv.ratio for lv in job.levels][1:] == [(2, 1, 2), (2, 1, 2), (2, 2, 2), (2, 2, 2)]
    # DOMAIN: x bounds -10/18 far field, y 0/0.5 periodic; the block's z faces
    # (+-8) are inside the z bounds (+-12): far field
    assert fine.bctype == (-20, -60, -20, -20, -60, -20)


def test_parameter_derivation_matches_host():
    job = load_job(os.path.join(HERE, "jobsheet"))
    a = bytes(job_params_c(job, 64))
    b = bytes(job.params.to_c())
    assert a == b


def test_pin_files():
    g = read_pin(os.path.join(HERE, "cylinder05_1.pin"))
    assert g == [dict(name=101, points=(449, 9, 257), origin=(-10.0, 0.0, -8.0), length=(28.0, 0.5, 16.0))]
    g = read_pin(os.path.join(HERE, "cylinder05_11122.pin"))
    assert [x["name"] for x in g] == [101, 201, 301, 401, 402, 501, 502]
    assert g[5]["points"] == (117, 33, 129) and g[5]["origin"] == (-1.0, 0.0, -1.0)


def test_p3d_surface_is_the_closed_cylinder():
    tri = read_p3d(os.path.join(HERE, "cylinder.p3d"))
    # 97x11 lateral patch and two 47x5 caps, two triangles per quad
    assert len(tri) == 2 * (96 * 10 + 2 * 46 * 4)
    v = tri.reshape(-1, 3)
    r = np.hypot(v[:, 0], v[:, 2])
    assert r.max() < 0.5 + 1e-9
    assert v[:, 1].min() == 0.0 and v[:, 1].max() == 0.5
    # the lateral patch (the first 2*96*10 triangles) goes round the full circle
    side = tri[:2 * 96 * 10].reshape(-1, 3)
    assert np.all(np.abs(np.hypot(side[:, 0], side[:, 2]) - 0.5) < 1e-9)
    ang = np.sort(np.unique(np.round(np.arctan2(side[:, 2], side[:, 0]), 9)))
    assert ang[0] < -math.pi + 1e-8 and ang[-1] > math.pi - 1e-8   # from -pi to pi (seam at pi)
    assert np.max(np.diff(ang)) < 2 * math.pi / 96 + 1e-6


def test_surface_geometry_matches_analytic_cylinder():
    """cut cells of the triangulated surface against the analytic cylinder
    (axis along y, r = 0.5) on the same level: the masks agree except where
    a cell centre lies between the polygon and the circle, and the surface
    points and normals of the common records agree to the polygon's
    deviation (r (1 - cos(pi/96)) = 2.7e-4)"""
    job = load_job(os.path.join(HERE, "jobsheet"))
    spec = job.levels[-2]   # 224 x 4 x 128
    gt = level_geometry(spec, job.body)
    from code_mge_amd.grid import body
    ga = level_geometry(spec, body("cylinder", center=(0.0, 0.0, 0.0), semi=(0.5, 0.5, 0.5), axis=1))
    diff = np.argwhere(gt.blank != ga.blank)
    xv, yv, zv = spec.vertices()
    cx = 0.5 * (xv[1:-1] + xv[2:])
    cz = 0.5 * (zv[1:-1] + zv[2:])
    for k, j, i in diff:
        rr = math.hypot(cx[i + 1], cz[k + 1])
        assert abs(rr - 0.5) < 3e-4, (i, j, k, rr)
    assert gt.nmir > 0 and abs(gt.nmir - ga.nmir) <= 4 * len(diff) + 4
    ka = {(tuple(r.ccid), r.faceid): r for r in ga.recs[:ga.nmir]}
    n = 0
    for r in gt.recs[:gt.nmir]:
        a = ka.get((tuple(r.ccid), r.faceid))
        if a is None:
            continue
        n += 1
        # on the polygon: radius between its inscribed and circumscribed circles
        assert 0.5 * math.cos(math.pi / 96) - 1e-12 <= math.hypot(r.psurf[0], r.psurf[2]) <= 0.5 + 1e-12
        assert max(abs(r.nvec[d] - a.nvec[d]) for d in range(3)) < 0.04   # facet vs exact normal
    assert n > 0.9 * gt.nmir
    assert abs(job.ref_area - 0.5) < 1e-12   # diameter x span
