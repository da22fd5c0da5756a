"""Limiter freezing (lfreeze > 0, Roe + MUSCL-minmod) on the HIP path:
recorded limiter values and the solve through the freeze against the oracle,
and batched cases (freeze test in the device-side convergence kernel) bit
for bit equal to their own runs (freeze test after each cycle on the case
stream)."""
import dataclasses

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Batch, Case
from oracle.oracle import OracleCase
from tests.gpu_common import cell_rel_err, rel_err

pytestmark = pytest.mark.gpu


def _prims(w, p):
    rr = 1.0 / w[0]
    return np.stack([w[0], w[1] * rr, w[2] * rr, w[3] * rr, p])


def test_lfreeze_parity_sphere(require_gpu):
    """24^3 sphere, 3 levels.  Recording: after an FMG ramp and 20 W-cycles
    the GPU's codes equal the oracle's wherever the oracle's two slope
    candidates are not tied to rounding (equal magnitudes or a zero); where
    the flow is still uniform they are rounding noise on both sides, so the
    frozen phase starts from the GPU's words copied into the oracle: 20
    frozen W-cycles (no FMG transfer) then give the same residual history and
    state"""
    prm, levels, bdy, _ = configs.sphere3d(n=24, nlevel=3, box=4.0, mach=0.5, alpha=2.0, distype="r2",
                                           lfreeze=1e30)
    geoms = [level_geometry(s, bdy) for s in levels]
    gpu = Case(prm, levels, geoms=geoms, mxcyc=45)
    orc = OracleCase(prm, geoms, mxcyc=45)
    for c in (gpu, orc):
        c.init_flow()
        c.solve([0, 2, 20])
        c.timestep(3)
        c.euler(3, 21, 3, 3, 1, 0)      # one residual evaluation: codes of the current state
    # recorded limiter bytes of the finest level, against the oracle's state
    cg, co = gpu.download(3, "lim"), orc.download(3, "lim")
    q = _prims(orc.download(3, "w"), orc.download(3, "p"))
    scale = np.abs(q).reshape(5, -1).max(axis=1)[:, None, None, None]
    noise = np.zeros(cg.shape, bool)      # a difference at the rounding level of the state
    for d in range(3):
        ax = 3 - d
        back = q - np.roll(q, 1, axis=ax)
        fwd = np.roll(q, -1, axis=ax) - q
        noise[5 * d:5 * d + 5] = (np.abs(back) < 1e-11 * scale) | (np.abs(fwd) < 1e-11 * scale)
    diff = np.abs(cg - co)
    print(f"limiter bytes differing {int((diff > 0).sum())} of {cg.size}: by one {int((diff == 1).sum())}, "
          f"by more at rounding-level differences {int(((diff > 1) & noise).sum())}, "
          f"by more elsewhere {int(((diff > 1) & ~noise).sum())}", flush=True)
    for idx in np.argwhere((diff > 1) & ~noise)[:8]:
        b, k, j, i = idx
        print(f"  cell ({i},{j},{k}) dir {b // 5} comp {b % 5}: gpu {cg[tuple(idx)]} oracle {co[tuple(idx)]}",
              flush=True)
    assert not ((diff > 1) & ~noise).any()
    assert (diff == 1).mean() < 1e-3
    # frozen phase from the same words
    for lvl in (1, 2, 3):
        orc.upload(lvl, "lim", gpu.download(lvl, "lim"))
    gpu.set_freeze(0, 1)
    orc.set_frozen(0, 1)
    lg = np.array([gpu.cycles(3, 1, ncyc0=k)[0] for k in range(22, 42)])
    lo = np.array([orc.cycles(3, 1, ncyc0=k)[0] for k in range(22, 42)])
    e = rel_err(lg, lo)
    es = cell_rel_err(gpu.download(3, "w"), orc.download(3, "w"), prm)
    print(f"frozen phase: residual history rel err {e:.3e}, state {es:.3e}", flush=True)
    assert e < 1e-9 and es < 1e-10
    assert np.array_equal(gpu.download(3, "lim"), orc.download(3, "lim"))   # frozen words are not rewritten
    gpu.close()
    orc.close()


def test_lfreeze_batch_equals_own_runs(require_gpu):
    specs = []
    for n, a in ((24, 2.0), (32, 6.0)):
        prm, levels, bdy, _ = configs.sphere3d(n=n, nlevel=3, box=4.0, mach=0.5, alpha=a, distype="r2", lfreeze=5.0)
        specs.append((prm, levels, [level_geometry(s, bdy) for s in levels]))
    ncyc = 24
    ref = []
    for prm, levels, geoms in specs:
        with Case(prm, levels, geoms=geoms, mxcyc=4) as c:
            c.init_flow()
            res = [c.cycles(3, 1, ncyc0=k)[0] for k in range(1, ncyc + 1)]
            assert res[0] / min(res) > 5.0       # the freeze triggers within the run
            ref.append((np.array(res), c.download(3, "w"), c.download(3, "lim")))
    cases, slots = [], []
    with Batch(use_graph=True, log_cap=ncyc + 2) as b:
        for prm, levels, geoms in specs:
            c = Case(prm, levels, geoms=geoms, mxcyc=4)
            c.init_flow()
            cases.append(c)
            slots.append(b.add(c, 1e30, ncyc))
        while b.run(4) > 0:
            pass
        for i, c in enumerate(cases):
            assert b.status(slots[i])["ncyc"] == ncyc
            assert np.array_equal(b.log(slots[i]), ref[i][0])
            assert np.array_equal(c.download(3, "w"), ref[i][1])
            assert np.array_equal(c.download(3, "lim"), ref[i][2])
            b.remove(slots[i])
    for c in cases:
        c.close()


def test_lfreeze_fp32_values_and_ghosts(require_gpu):
    """FP32 records limiter values too; ghost cells stay unrecorded (255)"""
    prm, levels, bdy, _ = configs.sphere3d(n=24, nlevel=2, box=4.0, mach=0.5, alpha=2.0, distype="r2",
                                           lfreeze=1e30)
    geoms = [level_geometry(s, bdy) for s in levels]
    with Case(dataclasses.replace(prm, precision=32), levels, geoms=geoms, mxcyc=8) as c:
        c.init_flow()
        c.cycles(2, 3)
        w = c.download(2, "lim")
    assert (w[:, :2] == 255).all() and (w[:, -2:] == 255).all()
    assert (w[:, :, :2] == 255).all() and (w[..., -2:] == 255).all()
    inner = w[:, 2:-2, 2:-2, 2:-2]
    assert (inner < 255).mean() > 0.9   # recorded on the fluid cells
