"""Pins of the oracle's unsteady dual time stepping (euler.f90:249-275,
320-331, 383-393; timestep.f90:86-99; solver.f90:263-276; no GPU).

The reference has no test or fixture for this path, so the oracle is pinned
by closed forms: in a periodic box holding a uniform state every spatial
residual is exactly zero, so a stage reduces to the scalar recurrence
w(m) = wn - tau (0.5/dt) X / (1 + 1.5 tau/dt), tau = fn*vdtim, with
X = 3 w(m-1) - 4 w1 + w2 (X = w(m-1) - w1 on the first physical step), and
the dual-time iteration converges to the BDF2 fixed point 3w - 4w1 + w2 = 0.
A sign, factor or time-level slip in any of the three terms fails one of
them.  A physical step so large that every transient term vanishes must
reproduce the steady solve bit for bit."""
import dataclasses

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import LevelSpec, hierarchy, level_geometry
from code_mge_amd.params import Params
from oracle.oracle import OracleCase

CSTP = (0.25, 0.16667, 0.375, 0.5, 1.0)


def _box(nlevel=1, dt=0.01, cfl=-2.5, n=(8, 8, 8), drop=1e6):
    levels = hierarchy(n, (0.0, 0.0, 0.0), (0.8, 0.8, 0.8), nlevel, bctype=(-60,) * 6)
    prm = Params(mach=0.5, alpha_deg=30.0, beta_deg=10.0, smoo=0.0, cfl=cfl, cstp=CSTP,
                 execmode=1, dtstep=dt, drop=drop)
    geoms = [level_geometry(s, None) for s in levels]
    return OracleCase(prm, geoms, mxcyc=400), prm, geoms


def _uniform(oc, lvl, v):
    ib1, jb1, kb1 = oc.dims(lvl)
    return np.broadcast_to(np.asarray(v, float)[:, None, None, None], (5, kb1, jb1, ib1)).copy()


def _set_levels(oc, nlevel, a, a1, a2, gamma):
    for lvl in range(1, nlevel + 1):
        w = _uniform(oc, lvl, a)
        oc.upload(lvl, "w", w)
        s = w[1] * w[1] + w[2] * w[2] + w[3] * w[3]
        oc.upload(lvl, "p", (gamma - 1.0) * (w[4] - 0.5 * s / w[0]))
        oc.upload(lvl, "w1", _uniform(oc, lvl, a1))
        oc.upload(lvl, "w2", _uniform(oc, lvl, a2))
        oc.bcond(lvl)


def _states(prm):
    d = prm.derived()
    base = np.array([d["rho0"], d["rho0"] * d["u0"], d["rho0"] * d["v0"], d["rho0"] * d["w0"],
                     d["rho0"] * d["h0"] - d["p0"]])
    a = base * np.array([1.01, 0.97, 1.03, 0.95, 1.02])
    a1 = base * np.array([1.00, 1.00, 1.00, 1.00, 1.00])
    a2 = base * np.array([0.98, 1.04, 0.96, 1.05, 0.99])
    return a, a1, a2


@pytest.mark.parametrize("itime", [0, 1, 2, 5])
def test_stage_recurrence_closed_form(itime):
    """one relax of a uniform state: the five stages follow the scalar
    recurrence exactly (to rounding), for each time level's form."""
    oc, prm, geoms = _box()
    a, a1, a2 = _states(prm)
    _set_levels(oc, 1, a, a1, a2, prm.gamma)
    oc.timestep(1)
    vd = oc.download(1, "vdtim")[2:-2, 2:-2, 2:-2]
    d = prm.derived()
    dt = d["dtstep"]
    # the physical step caps the pseudo step: vdtim = 2 dt / (3 cfl) exactly
    # (pressure switch r = 1 on a uniform state)
    assert np.all(vd == 2.0 * dt / (3.0 * d["cfl"]))
    oc.euler(1, 1, 1, 1, 0, 0, itime=itime)
    got = oc.download(1, "w")[:, 2:-2, 2:-2, 2:-2]
    w = a.copy()
    tau_v = vd[0, 0, 0]
    for m in range(5):
        fn = CSTP[m] * d["cfl"]
        tau = fn * tau_v
        if itime > 1:
            x = 3.0 * w - 4.0 * a1 + a2
        elif itime == 1:
            x = w - a1
        else:
            x = 0.0 * w
        w = a - tau * (0.5 / dt) * x / (1.0 + 1.5 * tau / dt)
    for n in range(5):
        err = np.abs(got[n] - w[n]).max() / abs(w[n])
        assert err < 1e-14, (n, err)
    if itime == 0:
        assert np.all(got == a[:, None, None, None])   # no transient term, R = 0
    else:
        assert np.abs(got[0] - a[0]).max() > 1e-6       # the term moved the state


def test_bdf2_fixed_point_through_multigrid():
    """W-cycles on three levels converge the uniform state to the BDF2
    solution w = (4 w1 - w2)/3 (coarse levels carry the forcing and their
    own lam division), here past the reference's 1e6 drop to 1e11 (the rounding floor is
    near 1e-13)."""
    oc, prm, geoms = _box(nlevel=3, dt=0.05, n=(16, 16, 16), drop=1e11)
    a, a1, a2 = _states(prm)
    _set_levels(oc, 3, a, a1, a2, prm.gamma)
    log, _ = oc.solve_step([0, 0, 60], itime=3)
    target = (4.0 * a1 - a2) / 3.0
    # the step ends with the time-level shift: w1 holds the converged state
    w1 = oc.download(3, "w1")[:, 2:-2, 2:-2, 2:-2]
    w2 = oc.download(3, "w2")[:, 2:-2, 2:-2, 2:-2]
    for n in range(5):
        assert np.abs(w1[n] - target[n]).max() < 1e-13 * abs(target[n]), n
        assert np.all(w2[n] == a1[n])
    assert log[-1, 2] < 1e-11 * log[0, 2]
    assert len(log) < 40                      # multigrid: a few cycles per decade


def test_huge_physical_step_is_the_steady_solve():
    """dtstep = 1e300: every transient term underflows against the spatial
    residual, lam = 1 and the cap never binds, so three unsteady steps equal
    three consecutive steady solves bit for bit."""
    prm, levels, bdy, aref = configs.cylinder2d(n=32, nz=2, nlevel=2, mach=0.3)
    geoms = [level_geometry(s, bdy) for s in levels]
    ocs = OracleCase(prm, geoms, mxcyc=20)
    ocu = OracleCase(dataclasses.replace(prm, execmode=1, dtstep=1e300), geoms, mxcyc=20)
    ocs.init_flow()
    ocu.init_flow()
    for it in range(3):
        ls, _ = ocs.solve([2, 6])
        lu, _ = ocu.solve_step([2, 6], itime=it)
        assert np.array_equal(ls, lu), it
        for lvl in (1, 2):
            assert np.array_equal(ocs.download(lvl, "w"), ocu.download(lvl, "w")), (it, lvl)


def test_time_levels_start_from_initial_state_and_shift():
    """init_flow copies the initial state into both earlier time levels
    (init_flow.f90:79-86); a step ends with w2 = w1, w1 = w."""
    prm, levels, bdy, aref = configs.cylinder2d(n=16, nz=2, nlevel=2, mach=0.3)
    geoms = [level_geometry(s, bdy) for s in levels]
    oc = OracleCase(dataclasses.replace(prm, execmode=1, dtstep=0.5), geoms, mxcyc=20)
    oc.init_flow()
    w0 = oc.download(2, "w")
    assert np.array_equal(oc.download(2, "w1")[:, 2:-2, 2:-2, 2:-2], w0[:, 2:-2, 2:-2, 2:-2])
    assert np.array_equal(oc.download(2, "w2")[:, 2:-2, 2:-2, 2:-2], w0[:, 2:-2, 2:-2, 2:-2])
    oc.solve_step([1, 3], itime=0)
    w_a = oc.download(2, "w")
    assert np.array_equal(oc.download(2, "w1"), w_a)
    oc.solve_step([1, 3], itime=1)
    assert np.array_equal(oc.download(2, "w2"), w_a)
    assert not np.array_equal(oc.download(2, "w1"), w_a)


def test_laminar_unsteady_step_is_fixed():
    """input_param.f90:93-96: laminar unsteady flow takes dtstep = u0/L/5
    whatever the input says."""
    d = Params(mach=0.2, alpha_deg=5.0, eqntype="n", execmode=1, dtstep=3.0).derived()
    assert d["dtstep"] == d["u0"] / 1.0 / 5.0
    d = Params(mach=0.2, alpha_deg=5.0, eqntype="e", execmode=1, dtstep=3.0).derived()
    assert d["dtstep"] == 3.0
