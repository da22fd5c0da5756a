"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
cpu_baseline leg of bench.py, always as the checker.  It mirrors
code_mge_amd.solver.Case so parity tests read call for call.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("mstage", C.c_int), ("cstp", C.c_double * 8), ("cdis", C.c_double * 8),
        ("cfl", C.c_double), ("vt", C.c_int), ("vis2", C.c_double), ("vis4", C.c_double),
        ("smoo", C.c_double), ("mgtype", C.c_int), ("gamma", C.c_double),
        ("rho0", C.c_double), ("p0", C.c_double), ("c0", C.c_double), ("ei0", C.c_double),
        ("u0", C.c_double), ("v0", C.c_double), ("w0", C.c_double), ("h0", C.c_double),
        ("eqntype", C.c_int), ("dissip", C.c_int),
        ("viscl", C.c_double), ("prlmean", C.c_double), ("gas_const", C.c_double),
        ("drop", C.c_double),
        ("distype", C.c_int), ("epsentr", C.c_double),
        ("execmode", C.c_int), ("dtstep", C.c_double),
        ("smoother", C.c_int),
        ("ffreeze", C.c_double),
        ("lfreeze", C.c_double),
    ]


class OrcCut(C.Structure):
    _fields_ = [
        ("faceid", C.c_int), ("zz", C.c_int),
        ("ccid", C.c_int * 3), ("scid", C.c_int * 3),
        ("psurf", C.c_double * 3), ("nvec", C.c_double * 3),
        ("radcur", C.c_double * 2), ("rr", C.c_double),
        ("npoint", C.c_int), ("ijk", (C.c_int * 3) * 8), ("coef", C.c_double * 8),
        ("psol", C.c_double), ("wsol", C.c_double * 5),
    ]


_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "liboracle.so"], cwd=HERE, check=True)
    L = C.CDLL(LIB_PATH)
    P, I = C.c_void_p, C.c_int
    L.orc_case_new.restype = P
    L.orc_case_new.argtypes = [C.POINTER(OrcParams), I]
    L.orc_case_free.argtypes = [P]
    L.orc_set_level.argtypes = [P, I, I, I, I, _dp, _dp, _dp, C.c_double, C.POINTER(I), C.c_void_p, I,
                                C.c_void_p, C.POINTER(I), I]
    for n in ("orc_init_flow",):
        getattr(L, n).argtypes = [P]
    L.orc_timestep.argtypes = [P, I]
    L.orc_euler.argtypes = [P, I, I, I, I, I, I, I]
    L.orc_bcond.argtypes = [P, I]
    L.orc_bc_cutcell.argtypes = [P, I]
    L.orc_movco.argtypes = [P, I]
    L.orc_movfin.argtypes = [P, I, I]
    L.orc_solve.argtypes = [P, C.POINTER(I), _dp, I, C.POINTER(I), C.POINTER(C.c_longlong)]
    L.orc_solve_step.argtypes = [P, C.POINTER(I), I, _dp, I, C.POINTER(I), C.POINTER(C.c_longlong)]
    L.orc_shift_time_levels.argtypes = [P]
    L.orc_upload_solut.argtypes = [P, I, C.c_double, _dp]
    L.orc_get.argtypes = [P, I, C.c_char_p, _dp]
    L.orc_set_frozen.argtypes = [P, I, I]
    L.orc_cycles.argtypes = [P, I, I, I, _dp, C.POINTER(C.c_longlong)]
    L.orc_comp_solve.argtypes = [C.POINTER(P), I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I), _dp, I,
                                 C.POINTER(I), C.POINTER(C.c_longlong)]
    L.orc_comp_op.argtypes = [C.POINTER(P), I, C.POINTER(I), C.POINTER(I), C.POINTER(I), I, I, I]
    L.orc_set_frozen.restype = None
    L.orc_set.argtypes = [P, I, C.c_char_p, _dp]
    L.orc_get_cuts.argtypes = [P, I, C.c_void_p]
    L.orc_get_rtrms.argtypes = [P, I, I]
    L.orc_get_rtrms.restype = C.c_double
    L.orc_get_dtim.argtypes = [P, I]
    L.orc_get_dtim.restype = C.c_double
    L.orc_wall_force.argtypes = [P, I, _dp]
    L.orc_psmoo.argtypes = [P, I]
    L.orc_psmoo_var.argtypes = [P, I]
    L.orc_progon3.argtypes = [_dp, _dp, _dp, _dp, _dp, I]
    L.orc_dflux_matrix.argtypes = [C.c_double, _dp, _dp, _dp, _dp, _dp]
    L.orc_visc_face.argtypes = [I, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                C.c_double, _dp]
    L.orc_visc.argtypes = [P, I, C.c_double]
    _lib = L
    return L


def params_from(prm):
    """OrcParams from a code_mge_amd.params.Params (same derivations)."""
    d = prm.derived()
    P = OrcParams()
    P.mstage = len(prm.cstp)
    for i, v in enumerate(prm.cstp):
        P.cstp[i] = v
    for i, v in enumerate(prm.cdis):
        P.cdis[i] = v
    P.cfl = d["cfl"]
    P.vt = d["vt"]
    P.vis2 = prm.vis2
    P.vis4 = d["vis4"]
    P.smoo = prm.smoo
    P.mgtype = prm.mgtype
    P.gamma = prm.gamma
    for k in ("rho0", "p0", "c0", "ei0", "u0", "v0", "w0", "h0"):
        setattr(P, k, d[k])
    P.eqntype = 1 if prm.eqntype.lower() == "n" else 0
    P.dissip = 1 if prm.dissip.lower() == "matrix" else 0
    P.viscl = d["viscl"]
    P.prlmean = prm.prlmean
    P.gas_const = 1.0 / prm.gamma
    P.drop = prm.drop
    P.distype = {"c": 0, "r1": 1, "r2": 2}[prm.distype.lower()]
    P.epsentr = prm.epsentr
    P.execmode = d["execmode"]
    P.dtstep = d["dtstep"]
    P.smoother = prm.smoother
    P.ffreeze = prm.ffreeze
    P.lfreeze = prm.lfreeze
    return P


def roe_face(gamma, eps, cdis, qL, qR, S, signed=False):
    """one Roe face of the oracle: (central part e, upwind part d)"""
    L = lib()
    e = (C.c_double * 5)()
    d = (C.c_double * 5)()
    L.orc_roe_face(C.c_double(gamma), C.c_double(eps), C.c_double(cdis), (C.c_double * 5)(*qL),
                   (C.c_double * 5)(*qR), (C.c_double * 3)(*S), int(signed), e, d)
    return np.array(e[:]), np.array(d[:])


def muscl(order, qm2, qm1, qp0, qp1, fl=(1, 1, 1, 1)):
    """MUSCL-minmod face states (qL, qR) of the oracle"""
    L = lib()
    qL = (C.c_double * 5)()
    qR = (C.c_double * 5)()
    L.orc_muscl(int(order), (C.c_double * 5)(*qm2), (C.c_double * 5)(*qm1), (C.c_double * 5)(*qp0),
                (C.c_double * 5)(*qp1), (C.c_int * 4)(*fl), qL, qR)
    return np.array(qL[:]), np.array(qR[:])


class OracleCase:
    """Same calls as code_mge_amd.solver.Case, on the CPU restatement."""

    def __init__(self, prm, geoms, mxcyc=1000):
        self.L = lib()
        self.prm = prm
        self.geoms = geoms
        self.nlevel = len(geoms)
        self._p = params_from(prm)
        self.h = self.L.orc_case_new(C.byref(self._p), self.nlevel)
        for li, g in enumerate(geoms):
            s = g.spec
            cuts = (OrcCut * max(g.nmir, 1))()
            C.memmove(cuts, g.recs, C.sizeof(OrcCut) * g.nmir)
            bct = (C.c_int * 6)(*s.bctype)
            rat = (C.c_int * 3)(*s.ratio)
            rc = self.L.orc_set_level(self.h, li + 1, s.n[0], s.n[1], s.n[2],
                                      g.xv.ctypes.data_as(_dp), g.yv.ctypes.data_as(_dp),
                                      g.zv.ctypes.data_as(_dp), s.cellvolume(), bct,
                                      g.blank.ctypes.data_as(C.c_void_p), g.nmir, cuts, rat, mxcyc)
            assert rc == 0, rc

    def close(self):
        if self.h:
            self.L.orc_case_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"oracle {what} failed: {rc}")
        return rc

    def init_flow(self):
        return self._chk(self.L.orc_init_flow(self.h), "init_flow")

    def timestep(self, lvl):
        return self._chk(self.L.orc_timestep(self.h, lvl), "timestep")

    def euler(self, lvl, ncyc, mgl, mlvl, lres, lfir, itime=0):
        self._chk(self.L.orc_euler(self.h, lvl, itime, ncyc, mgl, mlvl, int(lres), int(lfir)), "euler")
        return self.L.orc_get_rtrms(self.h, lvl, ncyc)

    def bcond(self, lvl):
        return self._chk(self.L.orc_bcond(self.h, lvl), "bcond")

    def bc_cutcell(self, lvl):
        return self._chk(self.L.orc_bc_cutcell(self.h, lvl), "bc_cutcell")

    def movco(self, mgl):
        return self._chk(self.L.orc_movco(self.h, mgl), "movco")

    def movfin(self, mgl, mlvl):
        return self._chk(self.L.orc_movfin(self.h, mgl, mlvl), "movfin")

    def solve(self, mcyc, cap=100000):
        mc = (C.c_int * self.nlevel)(*mcyc)
        log = np.zeros(3 * cap)
        nlog = C.c_int(0)
        cs = C.c_longlong(0)
        self._chk(self.L.orc_solve(self.h, mc, log.ctypes.data_as(_dp), cap, C.byref(nlog), C.byref(cs)),
                  "solve")
        n = min(nlog.value, cap)
        return log[: 3 * n].reshape(n, 3), cs.value

    def solve_step(self, mcyc, itime, cap=100000):
        mc = (C.c_int * self.nlevel)(*mcyc)
        log = np.zeros(3 * cap)
        nlog = C.c_int(0)
        cs = C.c_longlong(0)
        self._chk(self.L.orc_solve_step(self.h, mc, itime, log.ctypes.data_as(_dp), cap, C.byref(nlog),
                                        C.byref(cs)), "solve_step")
        n = min(nlog.value, cap)
        return log[: 3 * n].reshape(n, 3), cs.value

    def upload_solut(self, lvl, p_tot=0.0):
        ni, nj, nk = self.geoms[lvl - 1].spec.n
        nv = 11 if self.prm.eqntype.lower() == "n" else 7
        out = np.empty((nv, nk, nj, ni))
        self._chk(self.L.orc_upload_solut(self.h, lvl, p_tot, out.ctypes.data_as(_dp)), "upload_solut")
        names = ("rho", "u", "v", "w", "cp", "mach", "pt-loss", "energy", "gradn1", "gradn2", "gradn3")
        return dict(zip(names[:nv], out))

    def shift_time_levels(self):
        return self._chk(self.L.orc_shift_time_levels(self.h), "shift_time_levels")

    def dims(self, lvl):
        ni, nj, nk = self.geoms[lvl - 1].spec.n
        return ni + 4, nj + 4, nk + 4

    def download(self, lvl, field):
        ib1, jb1, kb1 = self.dims(lvl)
        nc = {"w": 5, "dw": 5, "fw": 5, "ws": 5, "wn": 5, "wr": 5, "qf": 5, "w1": 5, "w2": 5,
              "si": 3, "sj": 3, "sk": 3, "lim": 15}.get(field, 1)
        out = np.empty((nc, kb1, jb1, ib1))
        self._chk(self.L.orc_get(self.h, lvl, field.encode(), out.ctypes.data_as(_dp)), "get")
        return out if nc > 1 else out[0]

    def upload(self, lvl, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.orc_set(self.h, lvl, field.encode(), a.ctypes.data_as(_dp)), "set")

    def cycles(self, mlvl, count, ncyc0=1):
        """count W-cycles at finest active level mlvl without the FMG
        transfer (numbered from ncyc0, as mge_cycles)"""
        r = C.c_double(0.0)
        cs = C.c_longlong(0)
        self._chk(self.L.orc_cycles(self.h, mlvl, ncyc0, count, C.byref(r), C.byref(cs)), "cycles")
        return r.value, cs.value

    def set_frozen(self, far=0, lim=0):
        """set the far-field and limiter freeze flags directly"""
        self.L.orc_set_frozen(self.h, int(far), int(lim))

    def cuts(self, lvl):
        g = self.geoms[lvl - 1]
        out = (OrcCut * max(g.nmir, 1))()
        n = self.L.orc_get_cuts(self.h, lvl, out)
        return out, n

    def wall_force(self, lvl=None):
        lvl = self.nlevel if lvl is None else lvl
        out = (C.c_double * 3)()
        self.L.orc_wall_force(self.h, lvl, out)
        return np.array(out[:])

    def dtim(self, lvl):
        return self.L.orc_get_dtim(self.h, lvl)

    def psmoo(self, lvl):
        return self._chk(self.L.orc_psmoo(self.h, lvl), "psmoo")

    def psmoo_var(self, lvl):
        return self._chk(self.L.orc_psmoo_var(self.h, lvl), "psmoo_var")

    def visc(self, lvl, cdis=1.0):
        """one viscous pass into fw (zeroed first)"""
        return self._chk(self.L.orc_visc(self.h, lvl, cdis), "visc")


def progon3(a, b, c, d):
    """the oracle's progon3 (0-based arrays of length n)"""
    L = lib()
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d)]
    x = np.zeros(len(arr[0]))
    L.orc_progon3(*[v.ctypes.data_as(_dp) for v in arr], x.ctypes.data_as(_dp), len(x))
    return x


def dflux_matrix(gamma, u1, u2, dw, ss):
    L = lib()
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (u1, u2, dw, ss)]
    ff = np.zeros(5)
    L.orc_dflux_matrix(gamma, *[x.ctypes.data_as(_dp) for x in a], ff.ctypes.data_as(_dp))
    return ff


def visc_face(direction, pL, pR, t1L, t1R, t2L, t2R, hn, S, mu, kap):
    """one viscous face flux of the oracle (5 components)"""
    L = lib()
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (pL, pR, t1L, t1R, t2L, t2R)]
    g = np.zeros(5)
    L.orc_visc_face(direction, *[x.ctypes.data_as(_dp) for x in a], hn, S, mu, kap, g.ctypes.data_as(_dp))
    return g


def comp_solve(cases, blocks, mcyc, cap=100000):
    """the embedded multi-block solve (orc_comp_solve) over one-level
    OracleCases; blocks carry level, parent and g0 as
    code_mge_amd.composite.BlockSpec"""
    L = lib()
    nb = len(cases)
    hs = (C.c_void_p * nb)(*[c.h for c in cases])
    lev = (C.c_int * nb)(*[b.level for b in blocks])
    par = (C.c_int * nb)(*[b.parent for b in blocks])
    g0 = (C.c_int * (3 * nb))(*[v for b in blocks for v in b.g0])
    nlevel = max(b.level for b in blocks)
    mc = (C.c_int * nlevel)(*mcyc)
    log = np.zeros(3 * cap)
    nlog = C.c_int(0)
    cs = C.c_longlong(0)
    rc = L.orc_comp_solve(hs, nb, lev, par, g0, mc, log.ctypes.data_as(_dp), cap, C.byref(nlog), C.byref(cs))
    if rc < 0:
        raise RuntimeError(f"oracle comp_solve failed: {rc}")
    n = min(nlog.value, cap)
    return log[: 3 * n].reshape(n, 3), cs.value


def comp_op(cases, blocks, op, L, mlvl=0):
    """one composite transfer alone (orc_comp_op): "overlap" or "embed" on
    level L, "movco" to coarse level L, "movfin" to fine level L"""
    L_ = lib()
    nb = len(cases)
    hs = (C.c_void_p * nb)(*[c.h for c in cases])
    lev = (C.c_int * nb)(*[b.level for b in blocks])
    par = (C.c_int * nb)(*[b.parent for b in blocks])
    g0 = (C.c_int * (3 * nb))(*[v for b in blocks for v in b.g0])
    code = {"overlap": 0, "embed": 1, "movco": 2, "movfin": 3}[op]
    rc = L_.orc_comp_op(hs, nb, lev, par, g0, code, L, mlvl)
    if rc < 0:
        raise RuntimeError(f"oracle comp_op {op} failed: {rc}")
