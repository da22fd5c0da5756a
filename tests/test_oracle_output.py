"""Pins of the oracle's upload_solut fields (upload_solut.f90:89-241; no
GPU) by gas-dynamics closed forms rather than by its own formulas: every
state on the freestream's isentrope with the freestream total pressure has
zero total-pressure loss, the Mach number it was built with, and the
isentropic pressure coefficient ((1 + 0.2 M^2)/(1 + 0.2 M_l^2))^3.5 - 1 scaled
by 2/(gamma M^2) (gamma = 1.4, whose (gamma-1)/2 and gamma/(gamma-1) the
reference writes as 0.2 and 3.5); a stagnation state has Mach 0."""
import numpy as np
import pytest

from code_mge_amd.grid import LevelSpec, level_geometry
from code_mge_amd.params import Params
from oracle.oracle import OracleCase


def _case(eqntype="e", mach=0.6):
    spec = LevelSpec(n=(6, 5, 4), lo=(0, 0, 0), hi=(0.6, 0.5, 0.4), ratio=(2, 2, 2), bctype=(-20,) * 6)
    prm = Params(mach=mach, alpha_deg=7.0, beta_deg=-4.0, eqntype=eqntype)
    oc = OracleCase(prm, [level_geometry(spec, None)])
    oc.init_flow()
    return oc, prm, spec


def _isentropic_states(prm, spec, seed=3):
    """states with the freestream's total pressure and entropy, local Mach
    numbers 0..1.6 and random directions"""
    d = prm.derived()
    g, M = prm.gamma, prm.mach
    pt = d["p0"] * (1.0 + 0.2 * M * M) ** 3.5
    rt = d["rho0"] * (1.0 + 0.2 * M * M) ** 2.5
    ni, nj, nk = spec.n
    rng = np.random.default_rng(seed)
    ml = rng.uniform(0.0, 1.6, size=(nk, nj, ni))
    ml[0, 0, 0] = 0.0
    ml[0, 0, 1] = M
    p = pt / (1.0 + 0.2 * ml * ml) ** 3.5
    rho = rt * (p / pt) ** (1.0 / g)
    speed = ml * np.sqrt(g * p / rho)
    v = rng.normal(size=(3, nk, nj, ni))
    v /= np.linalg.norm(v, axis=0)
    u = speed * v
    w = np.empty((5, nk, nj, ni))
    w[0] = rho
    w[1:4] = rho * u
    w[4] = p / (g - 1.0) + 0.5 * rho * speed * speed
    return w, ml, p, u


@pytest.mark.parametrize("eqntype", ["e", "n"])
def test_isentropic_states(eqntype):
    oc, prm, spec = _case(eqntype)
    w, ml, p, u = _isentropic_states(prm, spec)
    full = oc.download(1, "w")
    full[:, 2:-2, 2:-2, 2:-2] = w
    oc.upload(1, "w", full)
    out = oc.upload_solut(1)
    assert len(out) == (11 if eqntype == "n" else 7)
    M = prm.mach
    assert np.allclose(out["rho"], w[0], rtol=0, atol=0)
    for n, key in enumerate(("u", "v", "w")):
        assert np.abs(out[key] - u[n]).max() < 1e-13
    assert np.abs(out["mach"] - ml).max() < 1e-13
    assert np.abs(out["pt-loss"]).max() < 1e-13
    cp = (((1.0 + 0.2 * M * M) / (1.0 + 0.2 * ml * ml)) ** 3.5 - 1.0) * 2.0 / (prm.gamma * M * M)
    assert np.abs(out["cp"] - cp).max() < 1e-12
    # the stagnation cell: Mach 0, cp of the stagnation pressure; the
    # freestream-Mach cell: cp 0
    assert out["mach"][0, 0, 0] == 0.0
    assert abs(out["cp"][0, 0, 1]) < 1e-12
    if eqntype == "n":
        e = w[4] / w[0]
        assert np.abs(out["energy"] - e).max() < 1e-14 * np.abs(e).max()
        for key in ("gradn1", "gradn2", "gradn3"):
            assert np.all(out[key] == 0.0)


def test_total_pressure_argument():
    """a given p_tot shifts only the loss: loss = p (1+0.2M^2)^3.5 / p_tot - 1"""
    oc, prm, spec = _case()
    d = prm.derived()
    M = prm.mach
    pt = d["p0"] * (1.0 + 0.2 * M * M) ** 3.5
    a = oc.upload_solut(1)
    b = oc.upload_solut(1, p_tot=2.0 * pt)
    assert np.abs(b["pt-loss"] - (0.5 * (a["pt-loss"] + 1.0) - 1.0)).max() < 1e-14
    for key in ("rho", "u", "cp", "mach"):
        assert np.array_equal(a[key], b[key])
    assert np.abs(a["mach"] - M).max() < 1e-14 and np.abs(a["cp"]).max() < 1e-12
