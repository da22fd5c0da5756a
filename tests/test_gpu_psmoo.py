"""Parity of the variable-coefficient smoother (smoother = 1: psmoo.f90:1-95
with progon3.f90:1-62, coefficients timestep.f90:286-309) on the HIP path
against the oracle.  The GPU sweep is the Thomas form of progon3's
recurrences (x_r = v_{r+1} x_{r+1} + u_{r+1} instead of s_r x_n + t_r), equal
in exact arithmetic, so single calls agree to rounding (1e-13) and solves to
the 1e-10 per-cell bar after 50 cycles."""
import dataclasses

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Batch, Case, force_coefficients
from tests.gpu_common import cell_rel_err, make_pair, max_rel_to_field, perturbed_state, rel_err

pytestmark = pytest.mark.gpu

SINGLE = 1e-13
CYCLES = 1e-10


def test_psmoo_coefficients_and_stage_parity(require_gpu):
    """timestep's coefficients and two smoothed relax steps on a perturbed
    cylinder state (blank cells inside the body)"""
    gpu, orc, prm, _ = make_pair(n=32, nz=4, nlevel=1, smoother=1)
    gpu.init_flow(); orc.init_flow()
    w, p = perturbed_state(gpu, 1, prm)
    for c in (gpu, orc):
        c.upload(1, "w", w)
        c.upload(1, "p", p)
        c.bc_cutcell(1)
        c.bcond(1)
    for cyc in (1, 2):
        for c in (gpu, orc):
            c.timestep(1)
        rg = gpu.euler(1, cyc, 1, 1, 0, 0)
        ro = orc.euler(1, cyc, 1, 1, 0, 0)
        print(f"step {cyc}: rtrms gpu {rg:.15e} oracle {ro:.15e}", flush=True)
        assert rel_err(rg, ro) < 1e-12
        for f in ("w", "dw"):
            assert max_rel_to_field(gpu.download(1, f), orc.download(1, f)) < SINGLE, f


def test_psmoo_cylinder_50_cycles_parity(require_gpu):
    """the 128x128x4 cylinder, 3-level W-cycle, 50 cycles with psmoo:
    per-cell state within 1e-10, same residual history, Cd/Cl within 1e-9"""
    gpu, orc, prm, aref = make_pair(n=128, nz=4, nlevel=3, smoother=1)
    for c in (gpu, orc):
        c.init_flow()
    lg, csg = gpu.solve([0, 0, 50])
    lo, cso = orc.solve([0, 0, 50])
    assert csg == cso and lg.shape == lo.shape
    assert rel_err(lg[:, 2], lo[:, 2]) < 1e-9
    r = lg[lg[:, 0] == 3][:, 2]
    print(f"psmoo cylinder: residual {r[0]:.3e} -> {r[-1]:.3e}", flush=True)
    assert r[-1] < 0.1 * r[0]
    for lvl in (1, 2, 3):
        e = cell_rel_err(gpu.download(lvl, "w"), orc.download(lvl, "w"), prm)
        print(f"level {lvl}: per-cell relative state error {e:.3e}", flush=True)
        assert e < CYCLES
    q = prm.dynamic_pressure()
    assert np.max(np.abs(gpu.wall_force(3) - orc.wall_force(3))) / (q * aref) < 1e-9


def test_psmoo_sphere_roe_parity(require_gpu):
    """3-D body, Roe/MUSCL scheme, psmoo, ten cycles of an FMG ramp"""
    gpu, orc, prm, _ = make_pair(cfg="sphere3d", n=32, nlevel=3, box=4.0, distype="r2", smoother=1)
    for c in (gpu, orc):
        c.init_flow()
    lg, _ = gpu.solve([0, 2, 10])
    lo, _ = orc.solve([0, 2, 10])
    assert rel_err(lg[:, 2], lo[:, 2]) < 1e-9
    assert cell_rel_err(gpu.download(3, "w"), orc.download(3, "w"), prm) < CYCLES


def test_psmoo_fp32_forces_and_batch(require_gpu):
    """FP32 bulk with psmoo tracks FP64 in Cd/Cl (1e-5 after 100 cycles of
    the 64x64x4 cylinder), and two psmoo cases batched end bit for bit where
    their own runs end"""
    prm, levels, bdy, aref = configs.cylinder2d(n=64, nz=4, nlevel=3, smoother=1)
    geoms = [level_geometry(s, bdy) for s in levels]
    cdcl = {}
    for prec in (64, 32):
        with Case(dataclasses.replace(prm, precision=prec), levels, geoms=geoms, mxcyc=8) as c:
            c.init_flow()
            c.cycles(3, 100)
            cdcl[prec] = force_coefficients(c, aref)
    print(f"Cd/Cl fp64 {cdcl[64]} fp32 {cdcl[32]}", flush=True)
    assert abs(cdcl[32][0] - cdcl[64][0]) < 1e-5 and abs(cdcl[32][1] - cdcl[64][1]) < 1e-5
    specs = [(prm, levels, geoms)]
    p2, l2, b2, _ = configs.ellipsoid3d(n=32, nlevel=3, semi=(0.9, 0.6, 0.5), box=4.0, smoother=1)
    specs.append((p2, l2, [level_geometry(s, b2) for s in l2]))
    ref = []
    for sp, lv, ge in specs:
        with Case(sp, lv, geoms=ge, mxcyc=8) as c:
            c.init_flow()
            res = [c.cycles(3, 1, ncyc0=k)[0] for k in range(1, 13)]
            ref.append((np.array(res), c.download(3, "w")))
    cases = []
    with Batch(use_graph=True, log_cap=16) as b:
        slots = []
        for sp, lv, ge in specs:
            c = Case(sp, lv, geoms=ge, mxcyc=8)
            c.init_flow()
            cases.append(c)
            slots.append(b.add(c, 1e30, 12))
        while b.run(2) > 0:
            pass
        for i, c in enumerate(cases):
            st = b.status(slots[i])
            assert st["ncyc"] == 12
            assert np.array_equal(b.log(slots[i]), ref[i][0])
            assert np.array_equal(c.download(3, "w"), ref[i][1])
            b.remove(slots[i])
    for c in cases:
        c.close()
