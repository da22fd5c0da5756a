"""The far-field freeze (ffreeze > 0; no counterpart in the reference,
whose bc_far.f90:58-60 chooses inflow or outflow by the sign of the normal
velocity at every call) in the oracle (no GPU): until the freeze cycle the
solve is bit for bit the reference's; after it the ghost points keep their
last choice; on a case that converges either way the forces agree."""
import dataclasses

import numpy as np

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from oracle.oracle import OracleCase


def _solve(ffreeze, ncyc, **kw):
    prm, levels, bdy, aref = configs.cylinder2d(n=32, nz=4, nlevel=3, alpha=3.0, ffreeze=ffreeze, **kw)
    oc = OracleCase(prm, [level_geometry(s, bdy) for s in levels], mxcyc=ncyc + 2)
    oc.init_flow()
    log, _ = oc.solve([0, 0, ncyc])
    return oc, log, prm


def test_freeze_is_the_reference_until_it_triggers():
    """the freeze test runs after each finest-level cycle: with a drop
    factor first exceeded at cycle k, cycles 1..k are identical to the
    unfrozen solve"""
    ref, lr, prm = _solve(0.0, 40)
    r = lr[lr[:, 0] == 3][:, 2]
    k = int(np.argmax(r[0] / r > 30.0))          # first cycle past a drop of 30
    assert 3 < k < 39
    fz, lf, _ = _solve(30.0, 40)
    rf = lf[lf[:, 0] == 3][:, 2]
    assert np.array_equal(rf[:k + 2], r[:k + 2])   # cycle k+1 still used the choice of cycle k+1's own calls
    ref.close()
    fz.close()


def test_frozen_choices_stay_and_forces_agree():
    """on the converging cylinder the frozen and the unfrozen solve reach the
    same forces to 1e-6 of q A, both with the residual down by 1e3"""
    ref, lr, prm = _solve(0.0, 200)
    fz, lf, _ = _solve(100.0, 200)
    r0 = lr[lr[:, 0] == 3][:, 2]
    r1 = lf[lf[:, 0] == 3][:, 2]
    assert r0[-1] < 1e-3 * r0[0] and r1[-1] < 1e-3 * r1[0]
    q = prm.dynamic_pressure()
    aref = 1.0 * 4 * 20.0 / 32
    dF = np.abs(ref.wall_force(3) - fz.wall_force(3)).max() / (q * aref)
    print(f"force difference frozen vs reference: {dF:.2e} of q A")
    assert dF < 1e-6
    ref.close()
    fz.close()
