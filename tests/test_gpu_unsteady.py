"""Parity of the HIP unsteady dual time stepping (execmode 1: the BDF2
physical time derivative in euler, the point-implicit factor on coarse
levels and in the update, the physical cap on the pseudo step, the
time-level shift; euler.f90:249-275, 320-331, 383-393, timestep.f90:86-99,
solver.f90:263-276) against the oracle, over several physical steps so
that every time-level form (itime 0, 1 and >= 2) runs.  FP64 to the 1e-10
per-cell bar; the FP32 bulk path within 1e-5 in Cd/Cl.  The Fortran driver
runs the reference's cylinder job with runtype = 1 and must equal the
library's step loop bit for bit."""
import dataclasses
import re
import subprocess

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.inputs import load_job
from code_mge_amd.solver import Case
from oracle.oracle import OracleCase
from tests.gpu_common import cell_rel_err, make_pair, rel_err
from tests.test_gpu_fortran import EXE, _job_dir

pytestmark = pytest.mark.gpu

MCYC = [1, 2, 4]


def _steps(gpu, orc, prm, nsteps=4, check_levels=(1, 2, 3)):
    for c in (gpu, orc):
        c.init_flow()
    for it in range(nsteps):
        lg, csg = gpu.solve_step(MCYC, it)
        lo, cso = orc.solve_step(MCYC, it)
        assert csg == cso and lg.shape == lo.shape
        assert np.array_equal(lg[:, :2], lo[:, :2])
        # relative to the step's largest residual: on the first physical step
        # the coarse levels start from the freestream and log rounding noise
        # (1e-16), whose own relative error means nothing
        e_log = float(np.abs(lg[:, 2] - lo[:, 2]).max() / np.abs(lo[:, 2]).max())
        errs = []
        for lvl in check_levels:
            for f in ("w", "w1", "w2"):
                errs.append(cell_rel_err(gpu.download(lvl, f), orc.download(lvl, f), prm))
        print(f"step {it}: residual log {e_log:.3e}, per-cell state {max(errs):.3e}, "
              f"last residual {lg[-1, 2]:.6e}", flush=True)
        assert e_log < 1e-9
        assert max(errs) < 1e-10


@pytest.mark.parametrize("distype", ["c", "r2"])
def test_unsteady_euler_parity(require_gpu, distype):
    gpu, orc, prm, aref = make_pair(n=48, nz=2, nlevel=3, mach=0.3, alpha=3.0, distype=distype,
                                    execmode=1, dtstep=0.4)
    _steps(gpu, orc, prm)
    fg = gpu.wall_force()
    fo = orc.wall_force()
    print("wall force gpu", fg, "oracle", fo)
    assert rel_err(fg[:2], fo[:2]) < 1e-9


def test_unsteady_ns_parity(require_gpu):
    """laminar flow: dtstep = u0/L/5 from input_param.f90:93-96, the
    viscous pseudo-step term and the cap together"""
    gpu, orc, prm, aref = make_pair(n=48, nz=2, nlevel=3, mach=0.2, eqntype="n", reynolds=40.0,
                                    execmode=1, dtstep=123.0)
    assert prm.derived()["dtstep"] == prm.derived()["u0"] / 5.0
    _steps(gpu, orc, prm, nsteps=3)


def test_unsteady_global_time_step_parity(require_gpu):
    """vt = 1 (jobsheet cfl > 0): lam from the global time step"""
    gpu, orc, prm, aref = make_pair(n=32, nz=2, nlevel=3, mach=0.3, cfl=1.5, execmode=1, dtstep=0.3)
    _steps(gpu, orc, prm, nsteps=3)


def test_unsteady_fp32_forces(require_gpu):
    prm, levels, bdy, aref = configs.cylinder2d(n=48, nz=2, nlevel=3, mach=0.3, alpha=3.0, execmode=1,
                                                dtstep=0.4)
    geoms = [level_geometry(s, bdy) for s in levels]
    orc = OracleCase(prm, geoms, mxcyc=200)
    g32 = Case(dataclasses.replace(prm, precision=32), levels, geoms=geoms, mxcyc=200)
    for c in (g32, orc):
        c.init_flow()
    for it in range(4):
        g32.solve_step(MCYC, it)
        orc.solve_step(MCYC, it)
    q = prm.dynamic_pressure()
    f32 = g32.wall_force()
    f64 = orc.wall_force()
    cd32, cl32 = f32[0] / (q * aref), f32[1] / (q * aref)
    cd64, cl64 = f64[0] / (q * aref), f64[1] / (q * aref)
    print(f"FP32 unsteady: Cd err {abs(cd32 - cd64):.3e}, Cl err {abs(cl32 - cl64):.3e}")
    assert abs(cd32 - cd64) < 1e-5 and abs(cl32 - cl64) < 1e-5
    g32.close()


def test_unsteady_batch_is_refused(require_gpu):
    from code_mge_amd import _ffi
    from code_mge_amd.solver import Batch
    gpu, orc, prm, aref = make_pair(n=16, nz=2, nlevel=2, execmode=1, dtstep=0.4)
    gpu.init_flow()
    with Batch(use_graph=False) as b:
        with pytest.raises(_ffi.MgeError):
            b.add(gpu, 1e5, 10)
    gpu.close()


def _parse_log(path):
    rows, steps = [], 0
    for line in open(path):
        if line.startswith(" force"):
            continue
        f = line.split()
        if len(f) == 1:
            steps += 1
            continue
        rows.append([float(x) for x in f])
    return np.array(rows), steps


def test_fortran_driver_unsteady(require_gpu, tmp_path):
    """runtype = 1, period = 1, step = 3: three physical steps of the FMG
    ramp, each ended by mge_shift_time_levels and the "(I6)" line"""
    d = _job_dir(tmp_path)
    txt = (d / "jobsheet").read_text()
    txt = re.sub(r"(?m)^(\s*runtype\s*=\s*)0", r"\g<1>1", txt)
    txt = re.sub(r"(?m)^(\s*step\s*=\s*)2000", r"\g<1>3", txt)
    txt = re.sub(r"(?m)^(\s*runtype\s*=\s*1)", r"\g<1>\n  dt = 0.002", txt)
    assert re.search(r"(?m)^\s*runtype\s*=\s*1", txt) and re.search(r"(?m)^\s*step\s*=\s*3", txt)
    (d / "jobsheet").write_text(txt)
    p = subprocess.run([EXE, "jobsheet", "64"], cwd=d, capture_output=True, text=True, timeout=600)
    print(p.stdout[-2500:], p.stderr[-1500:])
    assert p.returncode == 0
    flog, nsep = _parse_log(d / "cylinder_ibm_re20_test_res_full.dat")
    job = load_job(str(d / "jobsheet"))
    assert job.params.execmode == 1 and job.nsteps == 3 == nsep
    geoms = [level_geometry(s, job.body) for s in job.levels]
    with Case(job.params, job.levels, geoms=geoms, mxcyc=max(job.mcyc) + 2) as c:
        c.init_flow()
        logs, _ = c.solve_unsteady(job.mcyc, job.nsteps)
        fp = c.wall_force()
    log = np.concatenate(logs)
    force_line = [l for l in open(d / "cylinder_ibm_re20_test_res_full.dat") if l.startswith(" force")][0]
    ff = np.array([float(x) for x in force_line.split()[1:]])
    print("Fortran log", flog[-2:], "library log", log[-2:])
    print("Fortran force", ff, "library force", fp)
    assert flog.shape == log.shape
    assert np.array_equal(flog, log)
    assert np.array_equal(ff, fp)
