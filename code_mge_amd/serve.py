"""Serving loop for steady solves: continuous batching on one GPU.

Requests (a case description plus an arrival time) queue in arrival order.
A request is admitted as soon as its device memory fits the HBM budget: its
geometry is built, its multigrid state is allocated and initialised on the
GPU and it joins the batch of its class (cases of one batch share levels,
scheme, stages, cycle type, smoothing and precision).  Every tick advances
each batch by a few periods (two multigrid cycles each, all cases in one
launch per operation); a case leaves the batch at the cycle its residual
has dropped by `drop` (or at maxcyc), its Cd/Cl are read, its memory is
freed and waiting requests are admitted in its place.  Admission (geometry,
device state, init_flow, autotune) runs on its own thread while the GPU
works, and the batches of different classes (Euler and Navier-Stokes, say)
advance concurrently on their own streams.  A case computes exactly what
it computes alone.

No counterpart in the reference (it runs one case per process).
"""
import dataclasses
import math
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np

from .grid import level_geometry
from .params import Params
from .solver import Batch, Case, force_coefficients

DEVREC_BYTES = 192


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    params: Params
    levels: list              # LevelSpec, coarsest first
    body: Any                 # analytic body or None
    ref_area: float
    info: Dict = field(default_factory=dict)


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    completed: float
    status: int               # 1 converged, 2 cycle limit, < 0 error
    ncyc: int
    residual_ratio: float
    cd: float
    cl: float
    device_bytes: int

    @property
    def time_to_convergence(self):
        return self.completed - self.arrival


def level_bytes(n, es, nmir=0, mxcyc=4, viscous=False, coarse=False, psmoo=False, lfreeze=False):
    """device bytes of one level as the library allocates them (every array
    padded by 64 bytes); wall-record arrays at their upper bound; a level
    below the finest also holds the prolongation's correction; the
    variable-coefficient smoother holds its coefficients and sweep scratch;
    limiter freezing three 8-byte words per cell; the wall-record workgroup
    ranges at one per touched cell at most"""
    ni, nj, nk = n
    E = (ni + 4) * (nj + 4) * (nk + 4)
    nmax = max(ni, nj, nk) + 5
    ngrid = -(-ni // 64) * -(-nj // 4) * -(-nk // 8)   # residual-norm partials (8 planes per block)
    sizes = [es * 6 * E] * 3 + [es * 5 * E] * 5 + [es * E] * 2 + [E] * 2
    sizes += [es * nmax] * 3 + [es * 9 * nmax, es * nmax, 8 * ngrid, 8, 8, 8 * (mxcyc + 1), 24]
    if viscous:
        sizes += [es * 4 * E]           # velocity and temperature
    if coarse:
        sizes += [es * 5 * E]           # correction w - ws
    if psmoo:
        sizes += [es * 4 * E]           # psmoo coefficients (3) and sweep scratch
    if lfreeze:
        sizes += [24 * E]               # limiter words (three per cell)
    if nmir:
        ntouch, nops = 2 * nmir, 6 * nmir
        sizes += [DEVREC_BYTES * nmir, 40 * nmir, 8 * nmir, 8 * ntouch, 4 * (ntouch + 1), 8 * (nops + 1),
                  8 * (ntouch + 1), 80 * (nops + 1)]
        sizes += [8 * nmir, 4 * (nmir + 1), 4 * nmir]   # cut cells with their record lists
    return sum(s + 64 for s in sizes)


class Server:
    def __init__(self, hbm_budget: float, drop: float = 1e5, maxcyc: int = 3000, precision: int = 32,
                 use_graph: bool = True, periods: int = 4, log=None, autotune: bool = True,
                 ffreeze: Optional[float] = None):
        self.budget = float(hbm_budget)
        self.drop = drop
        self.maxcyc = maxcyc
        self.precision = precision
        self.use_graph = use_graph
        self.periods = periods
        self.log = log
        self.autotune = autotune
        self.ffreeze = ffreeze   # None: each request's own far-field freeze setting
        self.batches: Dict[tuple, Batch] = {}
        self.used = 0
        self.peak_used = 0     # most device bytes reserved at once

    def _class(self, prm: Params, levels) -> tuple:
        return (prm.eqntype.lower(), prm.distype.lower(), prm.dissip.lower(), len(levels), len(prm.cstp),
                prm.mgtype, prm.smoo > 0.0, prm.smoother, prm.lfreeze > 0.0, self.precision)

    def estimate(self, req: Request, geoms=None) -> int:
        es = 4 if self.precision == 32 else 8
        total = 0
        for li, spec in enumerate(req.levels):
            nmir = geoms[li].nmir if geoms is not None else 0
            total += level_bytes(spec.n, es, nmir, viscous=req.params.eqntype.lower() == "n",
                                 coarse=li < len(req.levels) - 1,
                                 psmoo=req.params.smoother == 1 and req.params.smoo > 0.0,
                                 lfreeze=req.params.lfreeze > 0.0 and req.params.distype.lower() == "r2")
        return total

    def _prepare(self, req: Request):
        """host geometry, device state, init_flow and autotune of one
        request: the admission work, done off the tick thread"""
        geoms = [level_geometry(s, req.body) for s in req.levels]
        need = self.estimate(req, geoms)
        with self._cv:
            # wait for room in the HBM budget (an empty server always admits)
            while self.used + need > self.budget and self._resident > 0 and not self._stop:
                self._cv.wait(0.05)
            if self._stop:
                return None
            self.used += need
            self._resident += 1
            self.peak_used = max(self.peak_used, self.used)
        prm = dataclasses.replace(req.params, precision=self.precision)
        if self.ffreeze is not None:
            prm = dataclasses.replace(prm, ffreeze=self.ffreeze)
        case = Case(prm, req.levels, geoms=geoms, mxcyc=4)
        case.init_flow()
        if self.autotune:
            case.autotune()   # cached per level shape after the first request
        got = case.device_bytes()
        with self._cv:
            self.used += got - need
        return req, case, got

    def _preparer(self, pending, t0, clock):
        """admits requests in arrival order: sleeps until each arrives, then
        prepares it and hands it to the tick thread"""
        for req in pending:
            while not self._stop:
                wait = req.arrival - (clock() - t0)
                if wait <= 0:
                    break
                time.sleep(min(wait, 0.05))
            if self._stop:
                return
            try:
                item = self._prepare(req)
            except Exception as e:   # reported by the tick thread
                item = (req, e, 0)
            if item is None:
                return
            with self._cv:
                self._ready.append(item)
                self._cv.notify_all()

    def run(self, requests: List[Request], clock=time.monotonic) -> List[Result]:
        """Serve `requests` (arrival times relative to the start).  A preparer
        thread admits them in arrival order by the HBM budget and builds
        their device state while the GPU runs; the tick thread adds ready
        cases to the batch of their class, advances every class batch by
        `periods` periods -- the batches of different classes on their own
        streams from their own threads, so they overlap on the device --
        and retires finished cases, which frees their budget."""
        import threading
        from concurrent.futures import ThreadPoolExecutor
        pending = sorted(requests, key=lambda r: r.arrival)
        self._cv = threading.Condition()
        self._ready = []
        self._stop = False
        self._resident = 0
        active = {}            # (class, slot) -> (req, case, admitted, bytes)
        results: List[Result] = []
        t0 = clock()
        prep = threading.Thread(target=self._preparer, args=(pending, t0, clock), daemon=True)
        prep.start()
        pool = ThreadPoolExecutor(max_workers=4)
        try:
            while len(results) < len(pending):
                with self._cv:
                    ready, self._ready = self._ready, []
                    if not ready and not active:
                        self._cv.wait(0.05)
                        ready, self._ready = self._ready, []
                for req, case, got in ready:
                    if isinstance(case, Exception):
                        raise case
                    key = self._class(case.params, req.levels)
                    b = self.batches.get(key)
                    if b is None:
                        b = Batch(use_graph=self.use_graph, log_cap=8)
                        self.batches[key] = b
                    slot = b.add(case, self.drop, self.maxcyc)
                    active[(key, slot)] = (req, case, clock() - t0, got)
                    if self.log:
                        self.log(f"[serve] t={clock() - t0:7.2f}s admit {req.rid} ({got / 2**20:.0f} MiB, "
                                 f"{len(active)} active)")
                if not active:
                    continue
                keys = {k for (k, _) in active}
                runs = [pool.submit(self.batches[k].run, self.periods) for k in keys]
                for f in runs:
                    f.result()
                for (key, slot) in list(active):
                    b = self.batches[key]
                    st = b.status(slot)
                    if st["status"] == 0:
                        continue
                    req, case, adm, got = active.pop((key, slot))
                    cd, cl = force_coefficients(case, req.ref_area) if st["status"] > 0 else (math.nan, math.nan)
                    b.remove(slot)
                    case.close()
                    with self._cv:
                        self.used -= got
                        self._resident -= 1
                        self._cv.notify_all()
                    done = clock() - t0
                    ratio = st["last"] / st["ini"] if st["ini"] else math.nan
                    results.append(Result(req.rid, req.arrival, adm, done, st["status"], st["ncyc"], ratio, cd, cl,
                                          got))
                    if self.log:
                        self.log(f"[serve] t={done:7.2f}s done {req.rid} after {st['ncyc']} cycles "
                                 f"(status {st['status']}, {done - req.arrival:.2f}s)")
        finally:
            with self._cv:
                self._stop = True
                self._cv.notify_all()
            prep.join()
            pool.shutdown()
            for b in self.batches.values():
                b.close()
            self.batches.clear()
        return results


def percentiles(results: List[Result]):
    """p50 and p99 time-to-convergence over the requests that converged
    (status 1); NaN when none did.  Requests stopped at the cycle limit or
    by an error never converged and are counted by outcome_counts."""
    t = np.array([r.time_to_convergence for r in results if r.status == 1])
    if t.size == 0:
        return math.nan, math.nan
    return float(np.percentile(t, 50)), float(np.percentile(t, 99))


def outcome_counts(results: List[Result]):
    """requests per outcome: converged, stopped at the cycle limit, failed"""
    return {"converged": sum(1 for r in results if r.status == 1),
            "cycle_limit": sum(1 for r in results if r.status == 2),
            "errors": sum(1 for r in results if r.status < 0)}


def poisson_requests(count=256, rate=4.0, seed=20261203, nlevel=3, box=8.0, euler_only=False):
    """Requests with exponential inter-arrival gaps (rate per second) from
    numpy PCG64(seed); per request in order: type ~ Bernoulli(0.5)
    Euler/Navier-Stokes, grid N ~ exp U[ln 32, ln 192] rounded to a multiple
    of 8 cells per edge, M ~ U[0.1, 0.8], Re ~ exp U[ln 20, ln 200] (drawn
    for the Navier-Stokes type only), then the body as in the ellipsoid
    sweep: semi-axes ~ U[0.2, 1.0], a uniformly random rotation (unit
    quaternion of 4 normals), alpha ~ U[-5, 15] deg; box [-box, box]^3.
    With euler_only the Navier-Stokes draws are kept (same stream) but
    served as Euler requests."""
    from . import configs
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        ns = bool(rng.uniform() < 0.5)
        n = int(8 * round(math.exp(rng.uniform(math.log(32.0), math.log(192.0))) / 8.0))
        mach = float(rng.uniform(0.1, 0.8))
        re = float(math.exp(rng.uniform(math.log(20.0), math.log(200.0)))) if ns else None
        semi = tuple(float(v) for v in rng.uniform(0.2, 1.0, size=3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        alpha = float(rng.uniform(-5.0, 15.0))
        kw = {}
        if ns and not euler_only:
            kw = dict(eqntype="n", reynolds=re)
        prm, levels, bdy, aref = configs.ellipsoid3d(n=n, nlevel=nlevel, semi=semi,
                                                     rot=configs.quaternion_rotation(q), mach=mach, alpha=alpha,
                                                     box=box, **kw)
        out.append(Request(rid, t, prm, levels, bdy, aref,
                           dict(n=n, mach=mach, ns=ns and not euler_only, re=re, semi=semi, alpha=alpha)))
    return out
