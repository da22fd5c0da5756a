"""The reference's input files for the solve step -- jobsheet, .pin grid and
PLOT3D .p3d surface -- read by the library's readers (mge_read_job,
mge_read_p3d: the role of make_fsijob, srcf90/main.f90:44), turned into the
host objects of this package.  The same readers serve the Fortran driver
(code_mge_amd/fortran/mgsol.f90), so both hosts build identical cases."""
import ctypes as C
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _ffi
from .grid import LevelSpec, TriBody
from .params import Params


@dataclass
class Job:
    name: str
    params: Params
    levels: List[LevelSpec]     # coarsest first
    body: object                # TriBody or None
    mcyc: List[int]             # cycles per level of the FMG ramp
    nsteps: int                 # physical time steps (period x step; 1 for steady jobs)
    ref_area: float             # projected bounding-box area of the surface normal to x
    raw: object                 # the mge_job structure


def _err(rc, what):
    if rc != 0:
        msg = _ffi.lib().mge_input_error().decode()
        raise _ffi.MgeError(rc, what, msg)


def read_p3d(path):
    """triangles (ntri, 9) of the surface patches of a PLOT3D file"""
    L = _ffi.lib()
    n = C.c_int(0)
    _err(L.mge_read_p3d(path.encode(), 0, None, C.byref(n)), "mge_read_p3d")
    tri = np.zeros((max(n.value, 1), 9))
    _err(L.mge_read_p3d(path.encode(), n.value, tri.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)),
         "mge_read_p3d")
    return tri[:n.value]


def read_pin(path):
    L = _ffi.lib()
    n = C.c_int(0)
    _err(L.mge_read_pin(path.encode(), 0, None, C.byref(n)), "mge_read_pin")
    out = (_ffi.MgePinGrid * max(n.value, 1))()
    _err(L.mge_read_pin(path.encode(), n.value, out, C.byref(n)), "mge_read_pin")
    return [dict(name=g.name, points=tuple(g.points), origin=tuple(g.origin), length=tuple(g.length))
            for g in out[:n.value]]


def frontal_area(tri):
    """projected bounding-box area of the surface normal to x (the
    reference area of the force coefficients)"""
    if tri is None or len(tri) == 0:
        return 0.0
    y = tri[:, 1::3]
    z = tri[:, 2::3]
    return float((y.max() - y.min()) * (z.max() - z.min()))


def load_job(path, precision=64) -> Job:
    L = _ffi.lib()
    j = _ffi.MgeJob()
    _err(L.mge_read_job(path.encode(), C.byref(j)), "mge_read_job")
    ns = j.mstage
    Pc = _ffi.MgeParams()
    _err(L.mge_job_params(C.byref(j), precision, C.byref(Pc)), "mge_job_params")
    prm = Params(mach=j.mach, alpha_deg=j.alphad, beta_deg=j.betad, gamma=j.gamma, cfl=j.cflf, vis2=j.vis2,
                 vis4=j.vis4, smoo=j.smoo, mgtype=j.mgtype, cstp=tuple(j.cstp[:ns]), cdis=tuple(j.cdis[:ns]),
                 eqntype="n" if j.eqntype == 1 else "e", dissip="matrix" if j.dissip == 1 else "scalar",
                 distype={0: "c", 1: "r1", 2: "r2"}[j.distype], reynolds=j.reynolds, prlmean=j.prlmean,
                 epsentr=j.epsentr, precision=precision, execmode=j.execmode, dtstep=Pc.dtstep)
    levels = []
    for lvl in range(1, j.nlevel + 1):
        n = (C.c_int * 3)()
        r = (C.c_int * 3)()
        _err(L.mge_job_level(C.byref(j), lvl, n, r), "mge_job_level")
        levels.append(LevelSpec(n=tuple(n), lo=tuple(j.lo), hi=tuple(j.hi), ratio=tuple(r),
                                bctype=tuple(j.bctype)))
    surface = j.surface.decode()
    tri = read_p3d(surface) if surface else None
    return Job(j.job.decode(), prm, levels, TriBody(tri) if tri is not None else None,
               list(j.mcyc[:j.nlevel]), max(j.nperiod, 1) * max(j.nstep, 1) if j.execmode == 1 else 1,
               frontal_area(tri), j)


def job_params_c(job: Job, precision=64):
    """the library's own derivation of the job's parameters (mge_job_params)"""
    P = _ffi.MgeParams()
    _err(_ffi.lib().mge_job_params(C.byref(job.raw), precision, C.byref(P)), "mge_job_params")
    return P
