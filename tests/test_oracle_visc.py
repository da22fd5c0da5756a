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


def test_wall_face_no_slip_adiabatic():
    """The wall-record face between a cut cell and its no-slip mirror
    (velocity reversed, same temperature): the face velocity is zero, so the
    energy flux (work + heat) vanishes and only the stress remains."""
    from oracle.oracle import visc_face
    cut = np.array([0.3, -0.2, 0.1, 7.0])
    ghost = np.array([-0.3, 0.2, -0.1, 7.0])
    z = np.zeros(3)
    for d in range(3):
        g = visc_face(d, ghost, cut, z, z, z, z, 0.1, -0.01, 0.02, 0.05)
        assert g[0] == 0.0 and g[4] == 0.0
        # tangential shear mu * du_t/dn * S, normal stress 4/3 mu du_n/dn * S
        dn = (cut[:3] - ghost[:3]) / 0.1
        for a in range(3):
            expect = (4.0 / 3.0 if a == d else 1.0) * 0.02 * dn[a] * -0.01
            assert abs(g[1 + a] - expect) < 1e-15


def _ns_cylinder(**kw):
    from code_mge_amd import configs
    prm, levels, bdy, aref = configs.cylinder2d(n=32, nz=4, nlevel=1, mach=0.2, eqntype="n", reynolds=40.0,
                                                smoo=0.0, **kw)
    geoms = [level_geometry(s, bdy) for s in levels]
    return prm, geoms, aref


def test_ns_cylinder_symmetry_and_friction_drag():
    """Navier-Stokes cylinder at alpha = 0 without smoothing: the state stays
    mirror symmetric in y and the lift zero (a lo/hi side error in
    the wall records breaks this); the viscous part of the wall force (total
    minus the same state's force at vanishing viscosity) points downstream."""
    prm, geoms, aref = _ns_cylinder()
    oc = OracleCase(prm, geoms, mxcyc=50)
    oc.init_flow()
    for c in range(1, 6):
        oc.timestep(1)
        oc.euler(1, c, 1, 1, 0, 0)
    w = oc.download(1, "w")
    n = geoms[0].spec.n
    wm = w[:, :, ::-1, :].copy()
    wm[2] *= -1
    inner = w[:, 2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 2] - wm[:, 2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 2]
    assert np.max(np.abs(inner)) < 1e-13
    F = oc.wall_force()
    assert abs(F[1]) < 1e-12 * abs(F[0])   # mirror pairs cancel up to the summation order
    # same state and mirror at vanishing viscosity: the convective part only
    import dataclasses
    p0 = dataclasses.replace(prm, reynolds=1e300)
    oc0 = OracleCase(p0, geoms, mxcyc=4)
    oc0.upload(1, "w", w)
    oc0.upload(1, "p", oc.download(1, "p"))
    oc0.bc_cutcell(1)
    F0 = oc0.wall_force()
    fv = F - F0
    print("wall force total", F, "viscous part", fv)
    assert fv[0] > 0.0 and abs(fv[0]) > 1e-3 * abs(F[0])
    oc.close()
    oc0.close()
