"""Batched execution (mge_batch_*): every case of a batch must end with
exactly the state, residual history and cycle count of its own
single-case run with the same stopping rule (stop at the first cycle whose
residual has dropped by more than `drop` from the first cycle, or at
maxcyc).  The per-case arithmetic is the same kernels on the same blocks,
so the comparison is bit for bit."""
import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import LevelSpec, level_geometry
from code_mge_amd.params import Params
from code_mge_amd.solver import Batch, Case

pytestmark = pytest.mark.gpu

DROP = 300.0
MAXCYC = 60


def _cases(precision=64, distype="c", **kw):
    """mixed shapes: ellipsoids of three sizes, a cylinder (semi-coarsened
    z) and a box without a body (no wall records)"""
    out = []
    rng = np.random.default_rng(4)
    for n, semi, mach, alpha in ((32, (0.9, 0.6, 0.5), 0.5, 2.0), (40, (0.7, 0.7, 0.4), 0.3, 8.0),
                                 (48, (1.0, 0.5, 0.6), 0.6, -3.0)):
        prm, levels, bdy, aref = configs.ellipsoid3d(n=n, nlevel=3, semi=semi, box=4.0, mach=mach, alpha=alpha,
                                                     rot=configs.random_rotation(rng), precision=precision,
                                                     distype=distype, **kw)
        out.append((prm, levels, [level_geometry(s, bdy) for s in levels]))
    prm, levels, bdy, aref = configs.cylinder2d(n=64, nz=4, nlevel=3, precision=precision, distype=distype, **kw)
    out.append((prm, levels, [level_geometry(s, bdy) for s in levels]))
    spec = [LevelSpec(n=(8 * m, 6 * m, 4 * m), lo=(0, 0, 0), hi=(1.6, 1.2, 0.8), ratio=(2, 2, 2), bctype=(-20,) * 6)
            for m in (1, 2, 4)]
    prm = Params(mach=0.4, alpha_deg=5.0, precision=precision, distype=distype, **kw)
    out.append((prm, spec, [level_geometry(s, None) for s in spec]))
    return out


def _alone(prm, levels, geoms):
    with Case(prm, levels, geoms=geoms, mxcyc=4) as c:
        c.init_flow()
        res = []
        for k in range(1, MAXCYC + 1):
            r, _ = c.cycles(len(levels), 1, ncyc0=k)
            res.append(r)
            if res[0] / r > DROP:
                break
        print(f"  alone: {levels[-1].n} {len(res)} cycles, residual {res[0]:.3e} -> {res[-1]:.3e}", flush=True)
        return np.array(res), c.download(len(levels), "w"), c.download(len(levels), "p")


def _batch(specs, use_graph, late=()):
    cases, slots = [], []
    with Batch(use_graph=use_graph, log_cap=MAXCYC + 2) as b:
        for i, (prm, levels, geoms) in enumerate(specs):
            c = Case(prm, levels, geoms=geoms, mxcyc=4)
            c.init_flow()
            cases.append(c)
            if i not in late:
                slots.append(b.add(c, DROP, MAXCYC))
            else:
                slots.append(None)
        print(f"  batch: {len(slots) - len(late)} cases, graph={use_graph}", flush=True)
        n = b.run(3)   # the late cases join while the others are running
        print(f"  batch: after 3 periods {n} active", flush=True)
        for i in late:
            slots[i] = b.add(cases[i], DROP, MAXCYC)
        for _ in range(MAXCYC):   # each period runs 2 cycles of every active case
            n = b.run(4)
            print(f"  batch: {n} active, cycles {[b.status(s)['ncyc'] for s in slots]}", flush=True)
            if n == 0:
                break
        assert n == 0
        out = []
        for i, c in enumerate(cases):
            st = b.status(slots[i])
            out.append((st, b.log(slots[i]), c.download(len(specs[i][1]), "w"), c.download(len(specs[i][1]), "p")))
            b.remove(slots[i])
    for c in cases:
        c.close()
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_alone_fp64(require_gpu, use_graph):
    specs = _cases()
    ref = [_alone(*s) for s in specs]
    got = _batch(specs, use_graph, late=(2,) if use_graph else ())
    for i, ((res, w, p), (st, log, wb, pb)) in enumerate(zip(ref, got)):
        print(f"case {i}: alone {len(res)} cycles, batch {st}")
        assert st["status"] in (1, 2)
        assert st["ncyc"] == len(res)
        assert np.array_equal(log, res)
        assert np.array_equal(wb, w) and np.array_equal(pb, p)


def test_batch_equals_alone_fp32_roe(require_gpu):
    specs = _cases(precision=32, distype="r2")
    ref = [_alone(*s) for s in specs]
    got = _batch(specs, True)
    for i, ((res, w, p), (st, log, wb, pb)) in enumerate(zip(ref, got)):
        print(f"case {i}: alone {len(res)} cycles, batch {st}")
        assert st["ncyc"] == len(res)
        assert np.array_equal(log, res)
        assert np.array_equal(wb, w) and np.array_equal(pb, p)


def test_batch_equals_alone_ns_fp32(require_gpu):
    """laminar Navier-Stokes cases (viscous kernels, no-slip mirror) batched"""
    specs = _cases(precision=32, eqntype="n", reynolds=50.0)
    ref = [_alone(*s) for s in specs]
    got = _batch(specs, True)
    for i, ((res, w, p), (st, log, wb, pb)) in enumerate(zip(ref, got)):
        print(f"case {i}: alone {len(res)} cycles, batch {st}")
        assert st["ncyc"] == len(res)
        assert np.array_equal(log, res)
        assert np.array_equal(wb, w) and np.array_equal(pb, p)


def test_batch_equals_alone_four_stages(require_gpu):
    """an even stage count: the stage rotation must return each level's
    state to the buffer the recorded program starts from after two cycles,
    or the replays from the second period on would read a stage
    intermediate (CFL 1.5: this four-stage set diverges at 2.5 on these
    bodies, in the oracle too)"""
    kw = dict(cstp=(0.25, 0.3333, 0.5, 1.0), cdis=(1.0, 0.0, 0.56, 0.44), cfl=-1.5)
    specs = _cases(**kw)[:3]
    ref = [_alone(*s) for s in specs]
    got = _batch(specs, True)
    for i, ((res, w, p), (st, log, wb, pb)) in enumerate(zip(ref, got)):
        print(f"case {i}: alone {len(res)} cycles, batch {st}")
        assert st["ncyc"] == len(res) and st["ncyc"] > 4
        assert np.array_equal(log, res)
        assert np.array_equal(wb, w) and np.array_equal(pb, p)
