"""Pins of the oracle's laminar viscous flux.

euler.f90:137-181 calls fgradi/fgradj/fgradk and vflux of the viscpak
module, which has no source in the reference tree; the flux is this
build's own (see oracle.c, "viscous fluxes").  These closed forms pin it:

- Couette flow (u = U y/H, p constant, T(y) = T0 - mu U^2/(2 kappa H^2) y^2
  + c1 y) is an exact steady solution: every component of the viscous
  balance vanishes.  A wrong heat-flux sign, conductivity, work term or
  face velocity leaves an energy residual (shown by planting a wrong T
  curvature);
- Poiseuille profile u = U (1 - (y/H)^2): the x-momentum balance is
  -mu u'' V = 2 mu U V / H^2 in every cell, the other momentum components
  vanish;
- rigid rotation and uniform dilatation have zero stress (tau = mu (G +
  G^T) - 2/3 mu div(u) I), so their balances vanish; using G alone or a
  wrong 2/3 fails these.
"""
import numpy as np
import pytest

from code_mge_amd.grid import LevelSpec, level_geometry
from code_mge_amd.params import Params
from oracle.oracle import OracleCase

N = (8, 10, 4)
HI = (1.0, 1.25, 0.5)


def _case():
    spec = LevelSpec(n=N, lo=(0.0, 0.0, 0.0), hi=HI, ratio=(2, 2, 2), bctype=(-20,) * 6)
    prm = Params(mach=0.3, eqntype="n", reynolds=50.0)
    oc = OracleCase(prm, [level_geometry(spec, None)], mxcyc=4)
    return oc, prm, spec


def _centres(spec):
    """cell-centre coordinates of every array index (ghosts included)"""
    out = []
    for a in range(3):
        h = (spec.hi[a] - spec.lo[a]) / spec.n[a]
        out.append(spec.lo[a] + h * (np.arange(spec.n[a] + 4) - 1.5))
    z, y, x = np.meshgrid(out[2], out[1], out[0], indexing="ij")
    return x, y, z


def _run(oc, prm, rho, u, v, w, p):
    g = prm.gamma
    W = np.stack([rho, rho * u, rho * v, rho * w, p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])
    oc.upload(1, "w", W)
    oc.upload(1, "p", p)
    oc.visc(1, 1.0)
    fw = oc.download(1, "fw")
    return fw[:, 2:N[2] + 2, 2:N[1] + 2, 2:N[0] + 2]


def _consts(prm):
    d = prm.derived()
    R = 1.0 / prm.gamma
    mu = d["viscl"]
    kap = mu * prm.gamma * R / ((prm.gamma - 1.0) * prm.prlmean)
    return d, R, mu, kap


def test_couette_exact_solution():
    oc, prm, spec = _case()
    d, R, mu, kap = _consts(prm)
    x, y, z = _centres(spec)
    U, H = 0.8, HI[1]
    p0 = d["p0"]
    T0 = p0 / (d["rho0"] * R)
    c1 = 0.3 * T0
    a = mu * U * U / (2.0 * kap * H * H)
    area = (HI[0] / N[0]) * (HI[2] / N[2])
    scale = (mu * U / H + kap * abs(c1)) * area
    zero = np.zeros_like(x)
    for curv, expect_zero in ((a, True), (2.0 * a, False)):
        T = T0 - curv * y * y + c1 * y
        p = np.full_like(x, p0)
        fw = _run(oc, prm, p / (R * T), U * y / H, zero, zero, p)
        err = np.abs(fw).max() / scale
        print(f"Couette, T curvature {curv / a:.0f}x exact: max |balance| / flux scale = {err:.3e}")
        if expect_zero:
            assert err < 1e-12
        else:
            assert np.abs(fw[4]).max() / scale > 1e-3   # the energy balance sees the wrong profile
    oc.close()


def test_poiseuille_momentum_balance():
    oc, prm, spec = _case()
    d, R, mu, kap = _consts(prm)
    x, y, z = _centres(spec)
    U, H = 0.6, 0.9
    zero = np.zeros_like(x)
    fw = _run(oc, prm, np.full_like(x, d["rho0"]), U * (1.0 - (y / H) ** 2), zero, zero, np.full_like(x, d["p0"]))
    V = spec.cellvolume()
    expect = 2.0 * mu * U * V / (H * H)
    scale = expect
    print("Poiseuille x-momentum balance / closed form:", fw[1].min() / expect, fw[1].max() / expect)
    assert np.abs(fw[1] - expect).max() / scale < 1e-9
    for n in (0, 2, 3):
        assert np.abs(fw[n]).max() / scale < 1e-9


@pytest.mark.parametrize("field", ["rotation", "dilatation"])
def test_stress_free_fields(field):
    oc, prm, spec = _case()
    d, R, mu, kap = _consts(prm)
    x, y, z = _centres(spec)
    om = 0.7
    if field == "rotation":
        u, v, w = -om * (y - 0.6), om * (x - 0.5), np.zeros_like(x)
    else:
        u, v, w = om * (x - 0.5), om * (y - 0.6), om * (z - 0.25)
    fw = _run(oc, prm, np.full_like(x, d["rho0"]), u, v, w, np.full_like(x, d["p0"]))
    area = (HI[0] / N[0]) * (HI[2] / N[2])
    scale = mu * om * area
    err = np.abs(fw).max() / scale
    print(f"{field}: max |balance| / (mu omega A) = {err:.3e}")
    assert err < 1e-11
