"""Host-side mirror of the reference solve-step interface over libmge.

`Case` keeps the reference's call surface (srcf90/solver.f90 calls
timestep, euler, bc_overlap, movco, movfin per level) so a driver can run
the FMG/FAS cycle call by call, and `Case.solve` runs the same cycle inside
the library (mge_solve).  All state lives on the GPU; arrays cross the ABI
only in download/upload.
"""
import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _ffi
from .grid import LevelGeom, LevelSpec, level_geometry
from .params import Params

_dp = C.POINTER(C.c_double)

# upload_solut's variable names (upload_solut.f90:206-234)
SOLUT_NAMES = ("rho", "u", "v", "w", "cp", "mach", "pt-loss", "energy", "gradn1", "gradn2", "gradn3")


class Case:
    def __init__(self, params: Params, levels: Sequence[LevelSpec], body=None,
                 mxcyc: int = 1000, geoms: Optional[List[LevelGeom]] = None):
        self.lib = _ffi.lib()
        self.params = params
        self.levels = list(levels)
        self.nlevel = len(self.levels)
        self.geoms = geoms if geoms is not None else [level_geometry(s, body) for s in self.levels]
        self.mxcyc = mxcyc
        self._p = params.to_c()
        h = C.c_void_p()
        _ffi.check(self.lib.mge_case_create(C.byref(self._p), self.nlevel, C.byref(h)), "mge_case_create")
        self.h = h
        for li, g in enumerate(self.geoms):
            s = g.spec
            bct = (C.c_int * 6)(*s.bctype)
            rat = (C.c_int * 3)(*s.ratio)
            _ffi.check(self.lib.mge_level_define(
                self.h, li + 1, s.n[0], s.n[1], s.n[2],
                g.xv.ctypes.data_as(_dp), g.yv.ctypes.data_as(_dp), g.zv.ctypes.data_as(_dp),
                s.cellvolume(), bct, rat, g.blank.ctypes.data_as(C.c_void_p), g.nmir,
                g.recs, mxcyc), "mge_level_define")

    # ---- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.mge_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- reference call surface
    def init_flow(self):
        return _ffi.check(self.lib.mge_init_flow(self.h), "mge_init_flow")

    def timestep(self, lvl):
        return _ffi.check(self.lib.mge_timestep(self.h, lvl, 1), "mge_timestep")

    def euler(self, lvl, ncyc, mgl, mlvl, lres, lfir, itime=0):
        r = C.c_double(0.0)
        _ffi.check(self.lib.mge_euler(self.h, lvl, 1, itime, ncyc, mgl, mlvl, int(lres), int(lfir),
                                      C.byref(r)), "mge_euler")
        return r.value

    def bcond(self, lvl):
        return _ffi.check(self.lib.mge_bcond(self.h, lvl, 1), "mge_bcond")

    def bc_cutcell(self, lvl):
        return _ffi.check(self.lib.mge_bc_cutcell(self.h, lvl, 1), "mge_bc_cutcell")

    def bc_overlap(self, lvl):
        return _ffi.check(self.lib.mge_bc_overlap(self.h, lvl), "mge_bc_overlap")

    def movco(self, mgl):
        return _ffi.check(self.lib.mge_movco(self.h, mgl), "mge_movco")

    def movfin(self, mgl, mlvl):
        return _ffi.check(self.lib.mge_movfin(self.h, mgl, mlvl), "mge_movfin")

    def solve(self, mcyc: Sequence[int], cap: int = 100000):
        mc = (C.c_int * self.nlevel)(*mcyc)
        log = np.zeros(3 * cap, dtype=np.float64)
        nlog = C.c_int(0)
        cs = C.c_longlong(0)
        _ffi.check(self.lib.mge_solve(self.h, mc, log.ctypes.data_as(_dp), cap, C.byref(nlog), C.byref(cs)),
                   "mge_solve")
        n = min(nlog.value, cap)
        return log[: 3 * n].reshape(n, 3), cs.value

    def solve_step(self, mcyc: Sequence[int], itime: int, cap: int = 100000):
        """one physical time step (solver.f90's alop/rstep body): the FMG ramp
        at time level itime, then (unsteady cases) w2 = w1, w1 = w"""
        mc = (C.c_int * self.nlevel)(*mcyc)
        log = np.zeros(3 * cap, dtype=np.float64)
        nlog = C.c_int(0)
        cs = C.c_longlong(0)
        _ffi.check(self.lib.mge_solve_step(self.h, mc, itime, log.ctypes.data_as(_dp), cap, C.byref(nlog),
                                           C.byref(cs)), "mge_solve_step")
        n = min(nlog.value, cap)
        return log[: 3 * n].reshape(n, 3), cs.value

    def solve_unsteady(self, mcyc: Sequence[int], nsteps: int, itime0: int = 0):
        """nsteps physical time steps from time level itime0 (solver.f90's
        period x step loop); the residual log of each step"""
        logs, cs = [], 0
        for it in range(itime0, itime0 + nsteps):
            log, c = self.solve_step(mcyc, it)
            logs.append(log)
            cs += c
        return logs, cs

    def shift_time_levels(self):
        return _ffi.check(self.lib.mge_shift_time_levels(self.h), "mge_shift_time_levels")

    def cycles(self, mlvl, count, ncyc0=1):
        """`count` cycles at finest active level mlvl without the drop test"""
        r = C.c_double(0.0)
        cs = C.c_longlong(0)
        _ffi.check(self.lib.mge_cycles(self.h, mlvl, ncyc0, count, C.byref(r), C.byref(cs)), "mge_cycles")
        return r.value, cs.value

    # ---- data
    def dims(self, lvl):
        ni, nj, nk = self.levels[lvl - 1].n
        return ni + 4, nj + 4, nk + 4

    def download(self, lvl, field):
        ib1, jb1, kb1 = self.dims(lvl)
        nc = 5 if field in ("w", "dw", "fw", "ws", "wr", "qf", "w1", "w2") else (15 if field == "lim" else 1)
        out = np.empty((nc, kb1, jb1, ib1), dtype=np.float64)
        _ffi.check(self.lib.mge_download(self.h, lvl, 1, field.encode(), out.ctypes.data_as(_dp)), "mge_download")
        return out if nc > 1 else out[0]

    def upload(self, lvl, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _ffi.check(self.lib.mge_upload(self.h, lvl, 1, field.encode(), a.ctypes.data_as(_dp)), "mge_upload")

    def upload_solut(self, lvl, p_tot=0.0, noalloc=False):
        """upload_solut's fields of level lvl (upload_solut.f90:89-241): a dict
        name -> (nk, nj, ni) array over the interior cells; p_tot <= 0 takes
        the freestream total pressure.  noalloc stages them through the
        library's fixed device scratch (mge_upload_solut_noalloc)."""
        fn = self.lib.mge_upload_solut_noalloc if noalloc else self.lib.mge_upload_solut
        what = "mge_upload_solut_noalloc" if noalloc else "mge_upload_solut"
        nv = C.c_int(0)
        _ffi.check(fn(self.h, lvl, 1, p_tot, None, C.byref(nv)), what)
        ni, nj, nk = self.levels[lvl - 1].n
        out = np.empty((nv.value, nk, nj, ni), dtype=np.float64)
        _ffi.check(fn(self.h, lvl, 1, p_tot, out.ctypes.data_as(_dp), C.byref(nv)), what)
        return dict(zip(SOLUT_NAMES[:nv.value], out))

    def cuts(self, lvl):
        g = self.geoms[lvl - 1]
        out = (_ffi.MgeCutRec * max(g.nmir, 1))()
        n = self.lib.mge_get_cuts(self.h, lvl, 1, out)
        _ffi.check(n, "mge_get_cuts")
        return out, n

    def wall_force(self, lvl=None):
        lvl = self.nlevel if lvl is None else lvl
        out = (C.c_double * 3)()
        _ffi.check(self.lib.mge_wall_force(self.h, lvl, out), "mge_wall_force")
        return np.array(out[:])

    def dtim(self, lvl):
        return self.lib.mge_get_dtim(self.h, lvl, 1)

    def device_bytes(self):
        return int(self.lib.mge_case_bytes(self.h))

    def write_solut(self, path, binary=False):
        """restart file of every level (write_solut: the reference's text
        format, or raw doubles with binary)"""
        _ffi.check(self.lib.mge_write_solut(self.h, str(path).encode(), 1 if binary else 0), "mge_write_solut")

    def read_solut(self, path, binary=False):
        """state of every level from a restart file (read_solut); p is
        recomputed from w"""
        _ffi.check(self.lib.mge_read_solut(self.h, str(path).encode(), 1 if binary else 0), "mge_read_solut")

    def autotune(self, levels=None):
        """pick the stage kernels' tile rows and k-chunk length per level by
        timing (mge_autotune); returns {level: (tile rows, kc, best launch us)}"""
        out = {}
        for lvl in (levels or range(1, len(self.levels) + 1)):
            kc = C.c_int(0)
            us = C.c_double(0.0)
            _ffi.check(self.lib.mge_autotune(self.h, lvl, C.byref(kc), C.byref(us)), "mge_autotune")
            out[lvl] = (self.stage_tile(lvl)[0], kc.value, us.value)
        return out

    def set_kc(self, lvl, kc):
        _ffi.check(self.lib.mge_set_kc(self.h, lvl, kc), "mge_set_kc")

    def set_freeze(self, far=0, lim=0):
        """set the far-field and limiter freeze flags directly (mge_set_freeze)"""
        _ffi.check(self.lib.mge_set_freeze(self.h, int(far), int(lim)), "mge_set_freeze")

    def stage_tile(self, lvl):
        """(tile rows, k-chunk length) the stage kernel of level lvl launches with"""
        tj, kc = C.c_int(0), C.c_int(0)
        _ffi.check(self.lib.mge_stage_tile(self.h, lvl, C.byref(tj), C.byref(kc)), "mge_stage_tile")
        return tj.value, kc.value

    def set_tile(self, lvl, tj, kc=0):
        """force the stage kernel's tile rows (4 or 8; 0: the precision's
        default) and k-chunk length (0: the fill rule) of level lvl"""
        _ffi.check(self.lib.mge_set_tile(self.h, lvl, tj, kc), "mge_set_tile")

    def sync(self):
        return _ffi.check(self.lib.mge_synchronize(self.h), "mge_synchronize")

    def timing(self, enable=True, name=None):
        """(mean ms, launches, mean algorithmic bytes) of kernel `name`;
        name None resets the statistics."""
        ms = C.c_double(0.0)
        cnt = C.c_longlong(0)
        by = C.c_double(0.0)
        _ffi.check(self.lib.mge_kernel_timing(self.h, int(enable), name.encode() if name else None,
                                              C.byref(ms), C.byref(cnt), C.byref(by)), "mge_kernel_timing")
        return ms.value, cnt.value, by.value


def force_coefficients(case: Case, ref_area: float):
    """Cd, Cl from the FP64 wall force of the finest level: drag along
    (cos a cos b, sin a cos b, sin b), lift along (-sin a, cos a, 0),
    normalised by q_inf * ref_area."""
    d = case.params.derived()
    a, b = d["alpha"], d["beta"]
    F = case.wall_force()
    q = case.params.dynamic_pressure()
    dd = np.array([np.cos(a) * np.cos(b), np.sin(a) * np.cos(b), np.sin(b)])
    ll = np.array([-np.sin(a), np.cos(a), 0.0])
    return float(F @ dd / (q * ref_area)), float(F @ ll / (q * ref_area))


class Batch:
    """Many cases advanced together (mge_batch_*): one launch per operation
    for all cases, replayed as a hipGraph.  Each added case (initialised,
    same levels/scheme/stages/cycle/precision as the others) runs multigrid
    cycles on its finest level until its residual has dropped by `drop`,
    or maxcyc cycles; it then holds the state its own run would hold."""

    STATUS = {0: "running", 1: "converged", 2: "cycle limit", 3: "removed"}

    def __init__(self, use_graph=True, log_cap=2000):
        self.lib = _ffi.lib()
        h = C.c_void_p()
        _ffi.check(self.lib.mge_batch_create(int(use_graph), log_cap, C.byref(h)), "mge_batch_create")
        self.h = h
        self.cases = {}

    def close(self):
        if self.h:
            self.lib.mge_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, case: Case, drop: float, maxcyc: int) -> int:
        slot = _ffi.check(self.lib.mge_batch_add(self.h, case.h, float(drop), int(maxcyc)), "mge_batch_add")
        self.cases[slot] = case
        return slot

    def remove(self, slot: int):
        _ffi.check(self.lib.mge_batch_remove(self.h, slot), "mge_batch_remove")
        self.cases.pop(slot, None)

    def run(self, periods: int = 1) -> int:
        n = C.c_int(0)
        _ffi.check(self.lib.mge_batch_run(self.h, int(periods), C.byref(n)), "mge_batch_run")
        return n.value

    def status(self, slot: int):
        st, nc = C.c_int(0), C.c_int(0)
        ini, last = C.c_double(0.0), C.c_double(0.0)
        _ffi.check(self.lib.mge_batch_status(self.h, slot, C.byref(st), C.byref(nc), C.byref(ini), C.byref(last)),
                   "mge_batch_status")
        return dict(status=st.value, ncyc=nc.value, ini=ini.value, last=last.value)

    def log(self, slot: int, cap: int = 100000):
        out = np.zeros(cap)
        n = _ffi.check(self.lib.mge_batch_log(self.h, slot, out.ctypes.data_as(_dp), cap), "mge_batch_log")
        return out[:n]
