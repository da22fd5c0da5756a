"""Edge cases of the HIP path against the oracle: blocks whose sizes fit
neither the 64x4 workgroup tile nor the k chunks (odd sizes, so the
hierarchy semi-coarsens), an off-centre rotated ellipsoid (wall records in
every orientation), symmetry and wall faces besides the far field, and the
smoothing dimension error of a one-cell direction."""
import numpy as np
import pytest

from code_mge_amd import _ffi, configs
from code_mge_amd.grid import body, hierarchy, level_geometry
from code_mge_amd.params import Params
from code_mge_amd.solver import Case
from oracle.oracle import OracleCase
from tests.gpu_common import cell_rel_err

pytestmark = pytest.mark.gpu

FAR, SYM, PER, WALL = -20, -50, -60, -80


def _pair(n, lo, hi, nlevel, bctype, bdy, **kw):
    levels = hierarchy(n, lo, hi, nlevel, bctype=bctype)
    prm = Params(**kw)
    geoms = [level_geometry(s, bdy) for s in levels]
    return Case(prm, levels, geoms=geoms, mxcyc=60), OracleCase(prm, geoms, mxcyc=60), prm, levels


def _solve_and_compare(gpu, orc, prm, levels, mcyc, tol=1e-10):
    for c in (gpu, orc):
        c.init_flow()
    lg, csg = gpu.solve(mcyc)
    lo, cso = orc.solve(mcyc)
    assert csg == cso and lg.shape == lo.shape
    assert np.array_equal(lg[:, :2], lo[:, :2])
    e_log = float(np.abs(lg[:, 2] - lo[:, 2]).max() / np.abs(lo[:, 2]).max())
    print(f"  residual log {e_log:.3e}; levels {[s.n for s in levels]}; last residual {lg[-1, 2]:.4e}")
    assert e_log < 1e-9
    for lvl in range(1, len(levels) + 1):
        e = cell_rel_err(gpu.download(lvl, "w"), orc.download(lvl, "w"), prm)
        print(f"  level {lvl}: per-cell relative state error {e:.3e}")
        assert e < tol
    fg, fo = gpu.wall_force(), orc.wall_force()
    print(f"  wall force gpu {fg} oracle {fo}")
    scale = max(np.abs(fo).max(), 1e-12)
    assert np.abs(np.asarray(fg) - np.asarray(fo)).max() < 1e-9 * scale


@pytest.mark.parametrize("distype,eqntype", [("c", "e"), ("r2", "e"), ("c", "n")])
def test_ragged_block_rotated_ellipsoid(require_gpu, distype, eqntype):
    """70 x 37 x 19 cells (not multiples of the 64 x 4 tile; 37 and 19 odd,
    so coarser levels keep them: ratio 1 in y and z), an ellipsoid rotated
    off the axes and off the box centre, 3 levels, 12 W-cycles"""
    rot = configs.quaternion_rotation(np.array([0.8, 0.3, -0.4, 0.33]) / np.linalg.norm([0.8, 0.3, -0.4, 0.33]))
    bdy = body("ellipsoid", center=(0.13, -0.07, 0.05), semi=(0.9, 0.45, 0.35), rot=rot)
    gpu, orc, prm, levels = _pair((70, 37, 19), (-3.5, -1.85, -1.25), (3.5, 1.85, 1.25), 3, (FAR,) * 6, bdy,
                                  mach=0.4, alpha_deg=6.0, beta_deg=-3.0, distype=distype, eqntype=eqntype,
                                  reynolds=60.0)
    assert [s.n for s in levels] == [(35, 37, 19), (35, 37, 19), (70, 37, 19)]   # ratio 1 between levels 1 and 2
    print(f"ragged {distype}/{eqntype}:")
    _solve_and_compare(gpu, orc, prm, levels, [0, 0, 12])
    gpu.close()


def test_symmetry_wall_and_periodic_faces(require_gpu):
    """bcond with every face kind (bc_symm, bc_wall, bc_periodic and the far
    field) around a sphere, 3 levels, 12 W-cycles"""
    bdy = body("sphere", center=(0.0, 0.3, 0.0), semi=(0.5, 0.5, 0.5))
    bct = (FAR, SYM, PER, FAR, WALL, PER)
    gpu, orc, prm, levels = _pair((48, 40, 24), (-3.0, -2.0, -1.5), (3.0, 3.0, 1.5), 3, bct, bdy,
                                  mach=0.3, alpha_deg=4.0)
    print("face kinds:")
    _solve_and_compare(gpu, orc, prm, levels, [0, 0, 12])
    gpu.close()


@pytest.mark.parametrize("n", [(8, 257, 256), (256, 3, 3), (300, 5, 3), (60, 3, 3), (129, 3, 3)])
def test_smoothing_line_shapes(require_gpu, n):
    """One smoothed stage against the oracle at 1e-13 relative to each
    field's largest value, for the line shapes the sweeps treat apart:
    y lines of 257 cells (beyond the 256 the checkpointed three-pass sweep
    takes: the four-pass sweep runs) and z lines of exactly 256 (16 full
    checkpoint chunks, no tail); x lines of 256 cells (4 per lane, four
    lines per wave, 45 lines: the last wave has one), of 300 (8 per lane,
    one line per wave), of 60 (1 per lane, four lines per wave) and of 129
    (rows of odd length: every other line starts off a pair boundary and
    takes single loads, and the last lane holds one cell)"""
    from tests.gpu_common import max_rel_to_field, perturbed_state
    levels = hierarchy(n, (0.0, 0.0, 0.0), tuple(0.01 * m for m in n), 1, bctype=(FAR,) * 6)
    prm = Params(mach=0.4, alpha_deg=3.0, beta_deg=1.0, smoo=0.6)
    geoms = [level_geometry(s, None) for s in levels]
    orc = OracleCase(prm, geoms, mxcyc=4)
    with Case(prm, levels, geoms=geoms, mxcyc=4) as gpu:
        gpu.init_flow()
        orc.init_flow()
        w, p = perturbed_state(gpu, 1, prm, seed=11)
        for c in (gpu, orc):
            c.upload(1, "w", w)
            c.upload(1, "p", p)
            c.bcond(1)
            c.timestep(1)
            c.euler(1, 1, 1, 1, 0, 0)
        for f in ("w", "dw", "p"):
            e = max_rel_to_field(gpu.download(1, f), orc.download(1, f))
            print(f"  {f}: {e:.3e}")
            assert e < 1e-13, f


def test_single_cell_direction_with_smoothing_is_an_error(require_gpu):
    """psmoo_old needs two cells per line (psmoo.f90:132,166): the GPU
    reports MGE_ERR_SMOO_DIM, as the restatement does"""
    levels = hierarchy((16, 12, 1), (0, 0, 0), (1.6, 1.2, 0.1), 1, bctype=(FAR, FAR, PER, FAR, FAR, PER))
    prm = Params(mach=0.3, smoo=0.6)
    geoms = [level_geometry(s, None) for s in levels]
    with Case(prm, levels, geoms=geoms, mxcyc=4) as c:
        c.init_flow()
        c.timestep(1)
        with pytest.raises(_ffi.MgeError) as ei:
            c.euler(1, 1, 1, 1, 0, 0)
        assert ei.value.code == -16
    orc = OracleCase(prm, geoms, mxcyc=4)
    orc.init_flow()
    orc.timestep(1)
    with pytest.raises(RuntimeError):
        orc.euler(1, 1, 1, 1, 0, 0)
