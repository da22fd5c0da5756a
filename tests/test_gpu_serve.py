"""The serving loop: requests arriving while others run are admitted by
HBM budget and batched with them; every result (cycle count, residual
ratio, Cd, Cl) is the same as the case's own run with the same stopping
rule, bit for bit."""
import dataclasses
import time

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.serve import Request, Server, level_bytes, percentiles
from code_mge_amd.solver import Case, force_coefficients

pytestmark = pytest.mark.gpu

DROP = 1e3
MAXCYC = 400


def _requests():
    out = []
    arrivals = (0.0, 0.0, 0.2, 0.5, 0.9, 1.2)
    for rid, (t, n, mach) in enumerate(zip(arrivals, (32, 40, 32, 48, 40, 32), (0.5, 0.3, 0.7, 0.4, 0.6, 0.2))):
        prm, levels, bdy, aref = configs.sphere3d(n=n, nlevel=3, mach=mach, alpha=3.0, box=4.0)
        out.append(Request(rid, t, prm, levels, bdy, aref, dict(n=n)))
    return out


def _alone(req, precision):
    prm = dataclasses.replace(req.params, precision=precision)
    geoms = [level_geometry(s, req.body) for s in req.levels]
    with Case(prm, req.levels, geoms=geoms, mxcyc=4) as c:
        c.init_flow()
        res = []
        for k in range(1, MAXCYC + 1):
            r, _ = c.cycles(len(req.levels), 1, ncyc0=k)
            res.append(r)
            if res[0] / r > DROP:
                break
        cd, cl = force_coefficients(c, req.ref_area)
        return len(res), res[-1] / res[0], cd, cl, c.device_bytes(), geoms


@pytest.mark.parametrize("budget_cases", [6, 2])
def test_serve_results_equal_alone(require_gpu, budget_cases):
    reqs = _requests()
    ref = [_alone(r, 32) for r in reqs]
    per_case = max(x[4] for x in ref)
    budget = per_case * budget_cases + 1
    peak = [0]
    srv = Server(hbm_budget=budget, drop=DROP, maxcyc=MAXCYC, precision=32, periods=2,
                 log=lambda m: print(m, flush=True))
    orig_run = srv.run

    res = orig_run(reqs)
    assert len(res) == len(reqs)
    by = {r.rid: r for r in res}
    for rid, (ncyc, ratio, cd, cl, nbytes, geoms) in enumerate(ref):
        r = by[rid]
        print(f"request {rid}: alone {ncyc} cycles Cd {cd:.8f}; served {r.ncyc} cycles Cd {r.cd:.8f} "
              f"ttc {r.time_to_convergence:.2f}s")
        assert r.status == 1 and r.ncyc == ncyc
        assert r.residual_ratio == ratio and r.cd == cd and r.cl == cl
        assert r.device_bytes == nbytes
        # the admission estimate bounds the allocation from above, within 5%
        est = srv.estimate(reqs[rid], geoms)
        assert nbytes <= est <= 1.05 * nbytes
    p50, p99 = percentiles(res)
    print(f"p50 {p50:.2f}s p99 {p99:.2f}s, peak reserved {srv.peak_used / 2**20:.0f} MiB of "
          f"{budget / 2**20:.0f} MiB")
    assert srv.peak_used <= budget
    if budget_cases == 2:
        # with room for two cases at a time, requests wait for admission
        assert max(r.admitted - r.arrival for r in res) > 0.0
