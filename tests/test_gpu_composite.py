"""Embedded multi-block hierarchies (mge_comp_*) on the HIP path: the
reference's own seven-block cylinder hierarchy (cylinder05_11122.pin: five
levels, two overlapping blocks on each of the two finest) against the
oracle's restatement, a composite whose child covers its whole parent
against the single-case two-level solve, and a preserved freestream."""
import dataclasses
import os

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.composite import BlockSpec, Composite, load_embedded_job
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Case
from oracle.oracle import OracleCase, comp_solve
from tests.gpu_common import cell_rel_err, rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cylinder")


def _cylinder_11122():
    return load_embedded_job(os.path.join(GOLDEN, "jobsheet"), os.path.join(GOLDEN, "cylinder05_11122.pin"))


def test_full_coverage_composite_is_the_single_case(require_gpu):
    """two blocks, the child covering its whole parent: the composite solve
    equals the single-case two-level solve bit for bit (FP64 and FP32)"""
    for prec in (64, 32):
        prm, levels, bdy, _ = configs.cylinder2d(n=64, nz=4, nlevel=2)
        prm = dataclasses.replace(prm, precision=prec)
        geoms = [level_geometry(s, bdy) for s in levels]
        with Case(prm, levels, geoms=geoms, mxcyc=20) as c:
            c.init_flow()
            ls, css = c.solve([2, 8])
            w1, w2 = c.download(1, "w"), c.download(2, "w")
        blocks = [BlockSpec(1, 1, -1, (0, 0, 0), levels[0]), BlockSpec(2, 2, 0, (0, 0, 0), levels[1])]
        cases = [Case(prm, [s], geoms=[g], mxcyc=20) for s, g in zip(levels, geoms)]
        for c in cases:
            c.init_flow()
        with Composite(cases, blocks) as comp:
            lc, csc = comp.solve([2, 8])
        print(f"FP{prec}: single {ls[-1]} composite {lc[-1]}", flush=True)
        assert csg_equal(ls, lc) and css == csc
        assert np.array_equal(cases[0].download(1, "w"), w1)
        assert np.array_equal(cases[1].download(1, "w"), w2)
        for c in cases:
            c.close()


def csg_equal(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("distype", ["c", "r2"])
def test_cylinder_11122_parity(require_gpu, distype):
    """the reference's seven-block hierarchy of the cylinder job (the
    jobsheet's parameters: JST; and the same with Roe/MUSCL), FMG ramp
    1:1:1:1:2: the residual history and the state of every block against the
    oracle"""
    prm, blocks, body, mcyc, aref = _cylinder_11122()
    prm = dataclasses.replace(prm, distype=distype)
    geoms = [level_geometry(b.spec, body) for b in blocks]
    print("blocks", [(b.name, b.level, b.parent, b.g0, b.spec.n) for b in blocks], flush=True)
    gcases = [Case(prm, [b.spec], geoms=[g], mxcyc=8) for b, g in zip(blocks, geoms)]
    ocases = [OracleCase(prm, [g], mxcyc=8) for g in geoms]
    for c in gcases + ocases:
        c.init_flow()
    ramp = [1, 1, 1, 1, 2]
    with Composite(gcases, blocks) as comp:
        lg, csg = comp.solve(ramp)
    print("gpu done; oracle (about 20 s)", flush=True)
    lo, cso = comp_solve(ocases, blocks, ramp)
    print("residuals gpu", lg[:, 2], "oracle", lo[:, 2], flush=True)
    assert csg == cso and lg.shape == lo.shape
    assert rel_err(lg[:, 2], lo[:, 2]) < 1e-9
    for b, g, o in zip(blocks, gcases, ocases):
        e = cell_rel_err(g.download(1, "w"), o.download(1, "w"), prm)
        print(f"block {b.name}: per-cell relative state error {e:.3e}", flush=True)
        assert e < 1e-10
    for c in gcases + ocases:
        c.close()


def test_cylinder_11122_freestream(require_gpu):
    """without the body the composite keeps the freestream in every block
    (embedding, overlap, restriction and prolongation of a constant state)"""
    prm, blocks, body, mcyc, aref = _cylinder_11122()
    geoms = [level_geometry(b.spec, None) for b in blocks]
    cases = [Case(prm, [b.spec], geoms=[g], mxcyc=8) for b, g in zip(blocks, geoms)]
    for c in cases:
        c.init_flow()
    w0 = [c.download(1, "w") for c in cases]
    with Composite(cases, blocks) as comp:
        log, _ = comp.solve([1, 1, 1, 1, 2])
    for c, w in zip(cases, w0):
        assert np.max(np.abs(c.download(1, "w") - w)) < 1e-12 * np.abs(w).max()
    assert np.max(log[:, 2]) < 1e-12
    for c in cases:
        c.close()
