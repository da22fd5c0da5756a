"""Scheme and freestream parameters: the derivations of input_param.

Restates srcf90/input_param.f90:13-115 (non-dimensionalisation with
rho0 = 1, |U_inf| = 1, p0 = 1/(gamma M^2)) with two recorded deviations:
 - the laminar Reynolds number is taken from the input instead of the
   hard-coded 500 of input_param.f90:111-115;
 - angles are given in degrees (the jobsheet's alphad/betad) and converted.
"""
import math
from dataclasses import dataclass, field
from typing import Sequence

from . import _ffi


@dataclass
class Params:
    mach: float = 0.2
    alpha_deg: float = 0.0
    beta_deg: float = 0.0
    gamma: float = 1.4
    cfl: float = -2.5                 # jobsheet sign: negative = local time step
    vis2: float = 0.5
    vis4: float = 1.0                 # divided by 32 below (input_param.f90:40)
    smoo: float = 0.6
    mgtype: int = 2
    cstp: Sequence[float] = (0.25, 0.16667, 0.375, 0.5, 1.0)
    cdis: Sequence[float] = (1.0, 0.0, 0.56, 0.0, 0.44)
    eqntype: str = "e"
    dissip: str = "scalar"
    distype: str = "c"
    reynolds: float = 500.0
    prlmean: float = 0.72
    ref_leng: float = 1.0
    drop: float = 1.0e6
    precision: int = 64
    epsentr: float = 0.05             # Roe entropy fix (jobsheet r2_epsentr)
    execmode: int = 0                 # 0 steady, 1 unsteady dual time stepping
    dtstep: float = 0.0               # physical time step in units of ref_leng/|U_inf|
    smoother: int = 0                 # 0 psmoo_old (constant coefficient), 1 psmoo + progon3
    ffreeze: float = 0.0              # > 0: freeze the far-field inflow/outflow choice at this drop
    lfreeze: float = 0.0              # > 0: freeze the MUSCL-minmod slope choices at this drop (Roe r2)
    extra: dict = field(default_factory=dict)

    def derived(self):
        """Return the derived typepara values as a dict."""
        g = self.gamma
        rm = self.mach
        alpha = math.radians(self.alpha_deg)
        beta = math.radians(self.beta_deg)
        vt = 0 if self.cfl < 0.0 else 1
        cfl = abs(self.cfl)
        vis4 = self.vis4 / 32.0
        rho0 = 1.0
        p0 = 1.0 / (g * (rm * rm))
        c0 = math.sqrt(g * p0 / rho0)
        ei0 = p0 / ((g - 1.0) * rho0)
        u0 = rm * c0 * math.cos(beta) * math.cos(alpha)
        v0 = rm * c0 * math.cos(beta) * math.sin(alpha)
        w0 = rm * c0 * math.sin(beta)
        h0 = g * ei0 + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0)
        viscl = 0.0
        if self.eqntype.lower() == "n":
            viscl = rho0 * u0 * self.ref_leng / self.reynolds
        # laminar unsteady flow takes a fixed physical step (input_param.f90:93-96)
        dtstep = self.dtstep
        if self.eqntype.lower() == "n" and self.execmode == 1:
            dtstep = u0 / self.ref_leng / 5.0
        return dict(vt=vt, cfl=cfl, vis4=vis4, rho0=rho0, p0=p0, c0=c0, ei0=ei0,
                    u0=u0, v0=v0, w0=w0, h0=h0, viscl=viscl, alpha=alpha, beta=beta,
                    execmode=self.execmode, dtstep=dtstep)

    def to_c(self):
        d = self.derived()
        P = _ffi.MgeParams()
        P.mstage = len(self.cstp)
        for i, v in enumerate(self.cstp):
            P.cstp[i] = v
        for i, v in enumerate(self.cdis):
            P.cdis[i] = v
        P.cfl = d["cfl"]
        P.vt = d["vt"]
        P.vis2 = self.vis2
        P.vis4 = d["vis4"]
        P.smoo = self.smoo
        P.mgtype = self.mgtype
        P.gamma = self.gamma
        for k in ("rho0", "p0", "c0", "ei0", "u0", "v0", "w0", "h0"):
            setattr(P, k, d[k])
        P.eqntype = 1 if self.eqntype.lower() == "n" else 0
        P.dissip = 1 if self.dissip.lower() == "matrix" else 0
        P.viscl = d["viscl"]
        P.prlmean = self.prlmean
        P.gas_const = 1.0 / self.gamma
        P.drop = self.drop
        P.distype = {"c": 0, "r1": 1, "r2": 2}[self.distype.lower()]
        P.precision = self.precision
        P.epsentr = self.epsentr
        P.execmode = d["execmode"]
        P.dtstep = d["dtstep"]
        P.smoother = self.smoother
        P.ffreeze = self.ffreeze
        P.lfreeze = self.lfreeze
        return P

    def dynamic_pressure(self):
        d = self.derived()
        return 0.5 * d["rho0"] * (d["u0"] ** 2 + d["v0"] ** 2 + d["w0"] ** 2)
