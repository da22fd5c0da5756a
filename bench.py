#!/usr/bin/env python
"""Benchmark of the multigrid Jameson-RK solve step on one MI355X.

Workload: steady inviscid flow past a sphere (radius 0.5) in a 256^3
Cartesian box [-8,8]^3, M = 0.5, alpha = 2 deg, cut-cell wall, Roe flux
with MUSCL-minmod reconstruction, 5-stage RK with residual smoothing,
4-level W-cycle.  A step is one multigrid cycle on the finest level; the
unit of work is a cell-stage (one RK-stage residual + update of one cell,
a residual-only pass counting 1).

The headline value is the plain FP64 path; the same workload is then run
on the FP32 bulk path and reported beside it.  Prints one JSON line with
the throughput (timed without instrumentation), the roofline of the
dominant kernel (the finest-level stage kernel; HIP events recorded on the
library's stream around each launch in a separate pass of a few cycles
after the timed ones) and the CPU baseline (the FP64 CPU restatement, one
core, bounded sample).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(n, prm_kw, nz_slab=48):
    """FP64 CPU restatement on one core: timestep + one 5-stage relax on a
    n x n x nz_slab slab (z in [-nz_slab*h/2, nz_slab*h/2]) of the same grid
    and body; returns Mcell-stage/s."""
    from code_mge_amd import configs
    from code_mge_amd.grid import LevelSpec, level_geometry
    from oracle.oracle import OracleCase
    prm, levels, bdy, _ = configs.sphere3d(n=n, nlevel=1, **prm_kw)
    h = 16.0 / n
    spec = LevelSpec(n=(n, n, nz_slab), lo=(-8.0, -8.0, -nz_slab * h / 2), hi=(8.0, 8.0, nz_slab * h / 2),
                     ratio=(2, 2, 2), bctype=(-20,) * 6)
    geoms = [level_geometry(spec, bdy)]
    oc = OracleCase(prm, geoms, mxcyc=4)
    oc.init_flow()
    t0 = time.perf_counter()
    oc.timestep(1)
    oc.euler(1, 1, 1, 1, 0, 0)
    dt = time.perf_counter() - t0
    cs = prm_stages(prm) * n * n * nz_slab
    oc.close()
    return cs / dt / 1e6, dt, f"timestep + one {prm_stages(prm)}-stage relax on a {n}x{n}x{nz_slab} slab " \
                             f"through the sphere of the {n}^3 grid ({cs/1e6:.1f} M cell-stages, {dt:.1f} s)"


def cfg1_leg(cycles=500):
    """The reference's CPU-runnable case in full (128x128x4 cylinder, M=0.2,
    alpha=0, JST, CFL 2.5, smoothing 0.6, 3-level W-cycle, `cycles` cycles
    after a 2-cycle FMG ramp per coarse level): the CPU restatement on one
    core and the HIP path (FP64) on the same case; wall time, throughput,
    Cd and the residual drop of each."""
    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry
    from code_mge_amd.solver import Case
    from oracle.oracle import OracleCase
    prm, levels, bdy, aref = configs.cylinder2d(n=128, nz=4, nlevel=3)
    assert prm.alpha_deg == 0.0 and prm.beta_deg == 0.0
    geoms = [level_geometry(s, bdy) for s in levels]
    mcyc = [2, 2, cycles]
    out = {"workload": f"cylinder 128x128x4, M=0.2, JST, 3-level W-cycle, FMG 2:2:{cycles}", "cycles": cycles}
    for kind in ("gpu", "cpu"):
        if kind == "gpu":
            c = Case(prm, levels, geoms=geoms, mxcyc=cycles + 4)
        else:
            c = OracleCase(prm, geoms, mxcyc=cycles + 4)
        c.init_flow()
        if kind == "gpu":
            c.sync()
        t0 = time.perf_counter()
        log, cs = c.solve(mcyc)
        dt = time.perf_counter() - t0
        F = c.wall_force(3)          # FP64 wall force; drag along the stream (alpha = beta = 0 here)
        cd = float(F[0] / (prm.dynamic_pressure() * aref))
        c.close()
        fine = log[log[:, 0] == 3]
        out[kind] = {"s": round(dt, 3), "mcell_stage_per_s": round(cs / dt / 1e6, 2), "cell_stages": int(cs),
                     "cd": cd, "residual_drop": float(fine[0, 2] / fine[-1, 2])}
        print(f"[bench] cfg1 {kind}: {dt:.2f} s, {cs / dt / 1e6:.1f} Mcs/s, Cd {cd:.8f}", file=sys.stderr,
              flush=True)
    out["cpu"]["cores"] = 1
    out["speedup_gpu_vs_cpu"] = round(out["cpu"]["s"] / out["gpu"]["s"], 1)
    return out


def _cpu_sweep_case(item):
    """one sweep case on the CPU restatement: geometry and setup, then two
    W-cycles on the finest level (timed)"""
    import dataclasses
    from code_mge_amd.grid import level_geometry
    from oracle.oracle import OracleCase
    prm, levels, bdy = item
    t0 = time.perf_counter()
    geoms = [level_geometry(s, bdy) for s in levels]
    oc = OracleCase(dataclasses.replace(prm, precision=64, drop=1e30), geoms, mxcyc=4)
    oc.init_flow()
    t1 = time.perf_counter()
    oc.solve([0] * (len(levels) - 1) + [2])
    t2 = time.perf_counter()
    oc.close()
    return t1 - t0, (t2 - t1) / 2.0


def cpu_sweep(sweep, cycles, threads):
    """all-core CPU estimate of the sweep's solves per hour: every case's
    setup and two W-cycles measured on the CPU restatement (FP64, one case
    per thread, `threads` at once; the restatement keeps no global state and
    its ctypes calls release the interpreter lock), each case's time to its
    drop extrapolated as setup + cycles_needed x (cycle time), with
    cycles_needed the FP64 count of the same case on the GPU, and the cases
    packed longest first onto the threads"""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        meas = list(ex.map(_cpu_sweep_case, [(prm, levels, bdy) for prm, levels, bdy, aref, info in sweep]))
    wall = time.perf_counter() - t0
    est = sorted((s + c * n for (s, c), n in zip(meas, cycles)), reverse=True)
    load = [0.0] * threads
    for e in est:
        load[load.index(min(load))] += e
    makespan = max(load)
    return {"threads": threads, "kind": "port", "cpu": _cpu_model(),
            "sample": f"setup + 2 W-cycles of each of the {len(sweep)} cases measured ({wall:.1f} s wall), "
                      "time to the drop extrapolated with each case's FP64 cycle count",
            "cycle_s": [round(c, 3) for s, c in meas], "setup_s": [round(s, 3) for s, c in meas],
            "est_makespan_s": round(makespan, 1), "est_solves_per_h": round(len(sweep) / makespan * 3600.0, 1)}


def prm_stages(prm):
    return len(prm.cstp)


def load_traffic(tag):
    """HBM bytes per k_stage launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", "k_stage_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(tag, {}).get("bytes_per_launch")
    except (OSError, ValueError, AttributeError):
        return None


# MI355X: 256 CUs x 4 SIMDs; a 64-lane wave instruction issues in 4 clocks
# on a 16-lane SIMD (FP64 vector at full rate); peak engine clock 2.4 GHz
SIMDS = 1024
CLOCK_GHZ = 2.4


def limiter(kname, precision, avg_launch_ms, hbm_frac):
    """what bounds the kernel, from the committed counter summary
    (rocprofv3 --pmc SQ_INSTS_VALU SQ_WAVES on the same launch): the time
    its VALU instructions alone need at full issue rate, as a fraction of
    the measured launch time, beside the HBM fraction"""
    p = os.path.join(ROOT, "profiles", "k_stage_counters.json")
    try:
        with open(p) as f:
            d = json.load(f)[f"{kname}_fp{precision}"]
    except (OSError, ValueError, KeyError):
        return None
    valu_ms = d["valu_instr_per_wave"] * d["waves_per_launch"] * 4 / (SIMDS * CLOCK_GHZ * 1e9) * 1e3
    valu_frac = valu_ms / avg_launch_ms
    kind = "valu" if valu_frac > max(hbm_frac, 0.5) else ("hbm" if hbm_frac > 0.5 else "latency")
    return {"measured": kind, "valu_floor_ms": round(valu_ms, 4), "valu_frac": round(valu_frac, 4),
            "hbm_frac": round(hbm_frac, 4), "waves_per_simd": d.get("waves_per_simd"),
            "source": "profiles/k_stage_counters.json"}


def run_leg(prm, levels, geoms, aref, args, precision):
    """warmup, the timed cycles (no instrumentation), then a separate pass of
    a few cycles with HIP events around every launch for the roofline of
    the finest-level stage kernel; returns the measurements"""
    import dataclasses
    from code_mge_amd.solver import Case, force_coefficients
    pp = dataclasses.replace(prm, precision=precision)
    nroof = 0 if args.no_timing else args.roofline_steps
    case = Case(pp, levels, geoms=geoms, mxcyc=args.warmup + args.steps + nroof + 2)
    case.init_flow()
    tuned = case.autotune()   # tile rows and k-chunk length of the stage kernel per level
    print(f"[bench] FP{precision}: autotuned (tile rows, k-chunk, us) {tuned}", file=sys.stderr, flush=True)
    nl = args.levels
    print(f"[bench] FP{precision}: warmup {args.warmup} cycles", file=sys.stderr, flush=True)
    r0, _ = case.cycles(nl, args.warmup, ncyc0=1)
    case.sync()
    t0 = time.perf_counter()
    r1, cs = case.cycles(nl, args.steps, ncyc0=args.warmup + 1)   # ends with a stream sync
    t1 = time.perf_counter()
    dt = t1 - t0
    cd, cl = force_coefficients(case, aref)
    print(f"[bench] FP{precision}: {args.steps} cycles in {dt:.3f} s, {cs/1e6:.1f} M cell-stages, "
          f"res {r0:.3e} -> {r1:.3e}", file=sys.stderr, flush=True)
    roofline = None
    if nroof > 0:
        case.timing(enable=True, name=None)
        case.cycles(nl, nroof, ncyc0=args.warmup + args.steps + 1)
        ms, cnt, by = case.timing(enable=False, name=f"k_stage@{nl}")
        ms_all, cnt_all, by_all = case.timing(enable=False, name="k_stage")
        if cnt > 0 and ms > 0:
            achieved = by / (ms * 1e-3) / 1e9
            kname = {"r1": "k_stage_roe", "r2": "k_stage_roe"}.get(prm.distype.lower(), "k_stage2")
            # the roofline of this path is HBM bandwidth (no contraction: no MFMA);
            # `limiter` names what the counters say holds the kernel below it
            roofline = {"kernel": f"{kname}<{'double' if precision == 64 else 'float'}> on the finest level "
                                  f"({levels[-1].n[0]}x{levels[-1].n[1]}x{levels[-1].n[2]})",
                        "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "limiter": limiter(kname, precision, ms, achieved / PEAK_HBM_GBS),
                        "frac": round(achieved / PEAK_HBM_GBS, 4),
                        "traffic": load_traffic(("jst_" if kname == "k_stage2" else "") + f"fp{precision}"),
                        "avg_launch_ms": round(ms, 5), "launches": cnt, "alg_bytes_per_launch": round(by),
                        "timed_cycles": nroof,
                        "all_levels": {"avg_launch_ms": round(ms_all, 5), "launches": cnt_all,
                                       "frac": round(by_all / (ms_all * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
                                       if ms_all > 0 else None}}
    kernels = finest_kernels(case, levels[-1].n, nl, precision) if nroof > 0 else None
    case.close()
    return dict(value=cs / dt / 1e6, dt=dt, cs=cs, cd=cd, cl=cl, roofline=roofline, r0=r0, r1=r1,
                tile={lvl: [tj, kc] for lvl, (tj, kc, us) in tuned.items()}, kernels=kernels)


def finest_kernels(case, n, lvl, precision):
    """HIP-event mean time and HBM fraction of the finest level's kernels in
    the roofline pass, from their algorithmic bytes per launch (s bytes per
    value, N interior cells, E cells with the two ghost layers):
    stage kernel as counted by the library; x/y/z smoothing read and write
    dw (10 s N) and read the flags (N); the post-smoothing update reads W0
    and dw and writes W and p on interior cells (16 s N); on this workload
    (six far-field faces, interior clouds) it carries no ghost shell"""
    s = 8 if precision == 64 else 4
    N = n[0] * n[1] * n[2]
    out = {}
    for name, by in (("k_stage", None), ("k_smooth_x", (10 * s + 1) * N), ("k_smooth_y", (10 * s + 1) * N),
                     ("k_smooth_z", (10 * s + 1) * N), ("k_update_ghost", 16 * s * N)):
        ms, cnt, b = case.timing(enable=False, name=f"{name}@{lvl}")
        if cnt == 0 or ms <= 0:
            continue
        b = b if by is None else by
        gbs = b / (ms * 1e-3) / 1e9
        out[name] = {"avg_us": round(ms * 1e3, 1), "launches": cnt, "alg_bytes": int(b),
                     "achieved_gbs": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


def jst_leg(args):
    """The same sphere with the reference's own JST scheme (the 256^3 grid,
    levels and flow of the headline legs; distype c): Mcell-stage/s in FP64
    and FP32 and the finest level's per-kernel HBM fractions"""
    import copy
    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry
    a = copy.copy(args)
    a.steps, a.warmup = min(args.steps, 5), min(args.warmup, 2)
    prm, levels, bdy, aref = configs.sphere3d(n=args.n, nlevel=args.levels, mach=0.5, alpha=2.0, distype="c")
    geoms = [level_geometry(s, bdy) for s in levels]
    out = {"workload": f"sphere {args.n}^3, M=0.5, alpha=2, {args.levels}-level W-cycle, JST (k2=0.5, k4=1/32)"}
    for precision in (64, 32):
        print(f"[bench] JST leg FP{precision}", file=sys.stderr, flush=True)
        leg = run_leg(prm, levels, geoms, aref, a, precision)
        out[f"fp{precision}"] = {"value": round(leg["value"], 2), "unit": "Mcell-stage/s",
                                 "ms_per_step": round(leg["dt"] / a.steps * 1e3, 3),
                                 "kernels": leg["kernels"], "cd": round(leg["cd"], 6)}
    return out


def sweep_leg(k, drop=1e5, maxcyc=3000, ffreeze=1e3, cpu_threads=0):
    """Converged solves per hour on the first k cases of the seeded ellipsoid
    sweep: FP32 served (continuous batching, all requests at t = 0) against
    the plain FP64 path one case at a time (solver.f90's loop with the drop
    test), both to a residual drop of `drop`; per-case setup (geometry,
    device state) is inside both timings."""
    import dataclasses
    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry
    from code_mge_amd.serve import Request, Server, outcome_counts, percentiles
    from code_mge_amd.solver import Case, force_coefficients
    sweep = configs.ellipsoid_sweep()[:k]
    reqs = [Request(i, 0.0, prm, levels, bdy, aref, info) for i, (prm, levels, bdy, aref, info) in enumerate(sweep)]
    print(f"[bench] sweep: {k} cases, FP32 served", file=sys.stderr, flush=True)
    srv = Server(hbm_budget=200e9, drop=drop, maxcyc=maxcyc, precision=32, periods=4, ffreeze=ffreeze,
                 log=lambda m: print(m, file=sys.stderr, flush=True) if " done " in m else None)
    t0 = time.perf_counter()
    served = srv.run(reqs)
    t32 = time.perf_counter() - t0
    served = {r.rid: r for r in served}
    print(f"[bench] sweep: FP32 served in {t32:.1f} s; FP64 one at a time", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    alone = {}
    for i, (prm, levels, bdy, aref, info) in enumerate(sweep):
        p64 = dataclasses.replace(prm, precision=64, drop=drop, ffreeze=ffreeze)
        geoms = [level_geometry(s, bdy) for s in levels]
        with Case(p64, levels, geoms=geoms, mxcyc=maxcyc + 2) as c:
            c.init_flow()
            log, _ = c.solve([0] * (len(levels) - 1) + [maxcyc])
            alone[i] = (len(log), force_coefficients(c, aref), log[-1, 2] / log[0, 2])
        print(f"[bench] sweep: FP64 case {i} ({levels[-1].n[0]}^3) {len(log)} cycles, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    t64 = time.perf_counter() - t0
    print(f"[bench] sweep: FP64 one at a time in {t64:.1f} s", file=sys.stderr, flush=True)
    cpu = None
    if cpu_threads > 0:
        print(f"[bench] sweep: CPU restatement, {cpu_threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_sweep(sweep, [alone[i][0] for i in range(len(sweep))], cpu_threads)
    conv32 = sum(1 for r in served.values() if r.status == 1)
    conv64 = sum(1 for v in alone.values() if v[2] < 1.0 / drop)
    dcd = max(abs(served[i].cd - alone[i][1][0]) for i in alone)
    dcl = max(abs(served[i].cl - alone[i][1][1]) for i in alone)
    p50, p99 = percentiles(list(served.values()))
    s32 = conv32 / t32 * 3600.0
    s64 = conv64 / t64 * 3600.0
    return {"sample": f"first {k} cases of the seeded ellipsoid sweep (PCG64(20261202), N 48..160, 3 levels), "
                      f"JST, residual drop {drop:g}, far-field choice frozen at a drop of {ffreeze:g} (both sides)",
            "fp32_served_solves_per_h": round(s32, 1), "fp64_one_at_a_time_solves_per_h": round(s64, 1),
            "ratio": round(s32 / s64, 2) if s64 > 0 else None,
            "fp32_converged": conv32, "fp64_converged": conv64, "cases": k,
            "fp32_wall_s": round(t32, 2), "fp64_wall_s": round(t64, 2),
            "max_cd_err_fp32_vs_fp64": dcd, "max_cl_err_fp32_vs_fp64": dcl,
            "ttc_p50_s": round(p50, 3), "ttc_p99_s": round(p99, 3),
            "ttc_over": "converged requests only", "fp32_outcomes": outcome_counts(list(served.values())),
            "cycles_fp32": [served[i].ncyc for i in sorted(served)], "cycles_fp64": [alone[i][0] for i in sorted(alone)],
            "cpu_all_cores": cpu}


def traffic_probe(args, precision):
    """A short program for rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE):
    a 1 GiB device-to-device copy (the calibration dispatch: 1 GiB read,
    1 GiB written) and then, after the autotuner's launches and a small fill
    dispatch that marks the start, one cycle of the bench workload: its
    first six stage-kernel dispatches (the finest level's five stages and
    its residual evaluation) give the HBM bytes and counters per launch."""
    import dataclasses
    import torch
    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry
    from code_mge_amd.solver import Case
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    prm, levels, bdy, aref = configs.sphere3d(n=args.n, nlevel=args.levels, mach=0.5, alpha=2.0,
                                              distype=args.probe_scheme)
    geoms = [level_geometry(s, bdy) for s in levels]
    with Case(dataclasses.replace(prm, precision=precision), levels, geoms=geoms, mxcyc=4) as c:
        c.init_flow()
        c.autotune()       # the launch shapes of the timed bench
        c.sync()
        mark = torch.zeros(64, device="cuda")
        mark.fill_(1.0)    # a fill dispatch marks where the cycle's dispatches begin
        torch.cuda.synchronize()
        c.cycles(args.levels, 1, ncyc0=1)
        c.sync()
    print(f"[bench] traffic probe FP{precision} done", file=sys.stderr, flush=True)


def converge_leg(prm, levels, geoms, aref, maxcyc, lfreeze, ffreeze=1e3, drop64=1e8, drop32=1e5):
    """The bench workload run from freestream to its stopping residual drop
    (1e-8 FP64, 1e-5 FP32; the W-cycle on the finest level after the FMG
    ramp, 2 cycles per coarser level) or maxcyc cycles, with the MUSCL-minmod
    limiter values frozen once the residual has dropped by lfreeze (0: never;
    without the freeze the limiter's switching stalls the residual near a
    drop of 3e2) and the far-field branches at ffreeze; time, cycles and
    Cd/Cl of each, and the FP32 error against FP64."""
    import dataclasses
    from code_mge_amd.solver import Case, force_coefficients
    out = {"workload": "the bench workload to convergence", "maxcyc": maxcyc, "lfreeze": lfreeze, "ffreeze": ffreeze}
    nl = len(levels)
    for prec, drop in ((64, drop64), (32, drop32)):
        pp = dataclasses.replace(prm, precision=prec, drop=drop, lfreeze=lfreeze, ffreeze=ffreeze)
        with Case(pp, levels, geoms=geoms, mxcyc=maxcyc + 8) as c:
            c.init_flow()
            c.autotune()
            c.sync()
            t0 = time.perf_counter()
            log, cs = c.solve([2] * (nl - 1) + [maxcyc])
            dt = time.perf_counter() - t0
            cd, cl = force_coefficients(c, aref)
        fine = log[log[:, 0] == nl]
        out[f"fp{prec}"] = {"s": round(dt, 2), "cycles": int(len(fine)), "residual_drop": float(fine[0, 2] / fine[-1, 2]),
                            "mcell_stage_per_s": round(cs / dt / 1e6, 1), "cd": cd, "cl": cl}
        print(f"[bench] converge FP{prec}: {len(fine)} cycles, {dt:.1f} s, drop {fine[0, 2] / fine[-1, 2]:.2e}, "
              f"Cd {cd:.8f}", file=sys.stderr, flush=True)
    out["cd_err_fp32"] = abs(out["fp32"]["cd"] - out["fp64"]["cd"])
    out["cl_err_fp32"] = abs(out["fp32"]["cl"] - out["fp64"]["cl"])
    return out


def ns_leg(cycles):
    """Laminar Navier-Stokes cylinder (Re = 40 on the diameter, M = 0.1,
    Pr = 0.72, constant viscosity) on 512 x 512 x 4 cells over [-20, 20]^2
    clustered geometrically toward the body (ratio 1.02; z uniform,
    periodic), 4-level W-cycle, JST + viscous fluxes: `cycles` W-cycles of
    the plain FP64 path and of the FP32 bulk path from freestream;
    throughput, residual drop and Cd/Cl of each, and the FP32 error against
    FP64 at the same iterate.  With the default 4000 cycles both precisions
    pass a residual drop of 1e5; the wake then still carries a slowly growing
    asymmetric mode (Cl about -1e-3 in FP64: psmoo_old's full-line pivots
    reused past the body's blank cells are not mirror-symmetric), which sets
    the size of the FP32-FP64 difference at a fixed iterate."""
    import dataclasses
    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry
    from code_mge_amd.solver import Case, force_coefficients
    prm, levels, bdy, aref = configs.cylinder2d_stretched(n=512, nz=4, nlevel=4, mach=0.1, half=20.0, ratio=1.02,
                                                          eqntype="n", reynolds=40.0, prlmean=0.72)
    geoms = [level_geometry(s, bdy) for s in levels]
    out = {"workload": "NS cylinder Re=40 M=0.1 Pr=0.72, 512x512x4 on [-20,20]^2 clustered toward the body at "
                       "ratio 1.02, 4-level W-cycle, JST + laminar viscous fluxes, smoothing 0.6", "cycles": cycles}
    for prec in (64, 32):
        with Case(dataclasses.replace(prm, precision=prec), levels, geoms=geoms, mxcyc=cycles + 2) as c:
            c.init_flow()
            c.autotune()
            r0, _ = c.cycles(len(levels), 1, ncyc0=1)
            c.sync()
            t0 = time.perf_counter()
            r1, cs = c.cycles(len(levels), cycles - 1, ncyc0=2)
            dt = time.perf_counter() - t0
            cd, cl = force_coefficients(c, aref)
        out[f"fp{prec}"] = {"mcell_stage_per_s": round(cs / dt / 1e6, 1), "s": round(dt, 2),
                            "residual_drop": r0 / r1, "cd": cd, "cl": cl}
        print(f"[bench] NS cylinder FP{prec}: {cs / dt / 1e6:.0f} Mcs/s, drop {r0 / r1:.3e}, Cd {cd:.6f}",
              file=sys.stderr, flush=True)
    out["cd_err_fp32"] = abs(out["fp32"]["cd"] - out["fp64"]["cd"])
    out["cl_err_fp32"] = abs(out["fp32"]["cl"] - out["fp64"]["cl"])
    return out


def serve_leg(count, rate=4.0, budget_gb=200.0, drop=1e5, maxcyc=3000, ffreeze=1e3):
    """Poisson serving stream (PCG64(20261203), `rate` requests/s, half of
    them laminar Navier-Stokes): ellipsoids, grids 32^3..192^3, 3 levels,
    FP32 served under an HBM budget to a residual drop of `drop`;
    p50/p99 time-to-convergence (completion - arrival) and solves/h."""
    from code_mge_amd.serve import Server, outcome_counts, percentiles, poisson_requests
    reqs = poisson_requests(count=count, rate=rate)
    srv = Server(hbm_budget=budget_gb * 1e9, drop=drop, maxcyc=maxcyc, precision=32, periods=4,
                 ffreeze=ffreeze, log=lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    p50, p99 = percentiles(res)
    conv = sum(1 for r in res if r.status == 1)
    ns = [r for r in res if reqs[r.rid].info["ns"]]
    return {"workload": f"{count} Poisson requests at {rate}/s (PCG64(20261203)), Euler/NS ellipsoids "
                        f"32^3..192^3, 3 levels, FP32, drop {drop:g}, HBM budget {budget_gb:g} GB",
            "wall_s": round(wall, 2), "converged": conv, "requests": count, "ns_requests": len(ns),
            "solves_per_h": round(conv / wall * 3600.0, 1), "ttc_p50_s": round(p50, 3), "ttc_p99_s": round(p99, 3),
            "ttc_over": "converged requests only", "outcomes": outcome_counts(res),
            "last_arrival_s": round(reqs[-1].arrival, 2),
            "cycles_max": max(r.ncyc for r in res)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip the HIP-event roofline pass")
    ap.add_argument("--roofline-steps", type=int, default=2, help="cycles of the HIP-event roofline pass")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 leg")
    ap.add_argument("--sweep", type=int, default=16, help="cases of the solves/h leg (0: skip)")
    ap.add_argument("--no-jst", action="store_true", help="skip the JST sphere leg")
    ap.add_argument("--ns-leg", type=int, default=4000,
                    help="W-cycles of the NS cylinder leg (0: skip); 4000 take both precisions past a 1e5 drop")
    ap.add_argument("--converge", type=int, default=0,
                    help="run the workload to its residual drop, at most this many cycles (0: skip)")
    ap.add_argument("--serve-leg", type=int, default=32, help="requests of the Poisson serving leg (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="threads of the all-core CPU estimate of the sweep (0: skip)")
    ap.add_argument("--cfg1-leg", type=int, default=0,
                    help="cycles of the full cfg1 cylinder on the CPU restatement and the HIP path (0: skip)")
    ap.add_argument("--lfreeze", type=float, default=100.0,
                    help="residual drop at which --converge freezes the limiter values (0: never)")
    ap.add_argument("--traffic-probe", type=int, default=0, choices=(0, 32, 64),
                    help="run only the short program profiled with rocprofv3 --pmc (precision 32 or 64)")
    ap.add_argument("--probe-scheme", default="r2", choices=("r2", "c"),
                    help="scheme of the --traffic-probe cycle: r2 (Roe/MUSCL, the headline) or c (JST)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args, args.traffic_probe)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the solve step runs on one MI355X; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)

    from code_mge_amd import configs
    from code_mge_amd.grid import level_geometry

    prm_kw = dict(mach=0.5, alpha=2.0, distype="r2")
    prm, levels, bdy, aref = configs.sphere3d(n=args.n, nlevel=args.levels, **prm_kw)
    print(f"[bench] building {args.n}^3 sphere case, {args.levels} levels", file=sys.stderr, flush=True)
    geoms = [level_geometry(s, bdy) for s in levels]
    main_leg = run_leg(prm, levels, geoms, aref, args, 64)
    fp32 = None
    if not args.no_fp32:
        leg = run_leg(prm, levels, geoms, aref, args, 32)
        fp32 = {"value": round(leg["value"], 2), "unit": "Mcell-stage/s", "dtype": "f32",
                "ms_per_step": round(leg["dt"] / args.steps * 1e3, 3), "roofline": leg["roofline"],
                "cd": round(leg["cd"], 6), "cl": round(leg["cl"], 6),
                "speedup_vs_fp64": round(leg["value"] / main_leg["value"], 3), "kernels": leg["kernels"]}
    value, dt, cs = main_leg["value"], main_leg["dt"], main_leg["cs"]
    cd, cl, roofline = main_leg["cd"], main_leg["cl"], main_leg["roofline"]

    jst = jst_leg(args) if not args.no_jst else None
    solves = sweep_leg(args.sweep, cpu_threads=args.cpu_threads if not args.no_cpu_baseline else 0) \
        if args.sweep > 0 else None
    ns = ns_leg(args.ns_leg) if args.ns_leg > 0 else None
    conv = converge_leg(prm, levels, geoms, aref, args.converge, args.lfreeze) if args.converge > 0 else None
    served = serve_leg(args.serve_leg) if args.serve_leg > 0 else None
    cfg1 = cfg1_leg(args.cfg1_leg) if args.cfg1_leg > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        v, t, sample = cpu_baseline(args.n, prm_kw)
        cpu = {"value": round(v, 3), "unit": "Mcell-stage/s", "cores": 1, "kind": "port",
               "sample": sample, "cpu": _cpu_model()}

    out = {
        "metric": "Mcell-stage updates/s",
        "value": round(value, 2),
        "unit": "Mcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (freestream start, analytic sphere geometry)",
        "config": {"workload": f"sphere {args.n}^3 in [-8,8]^3, M=0.5, alpha=2, {args.levels}-level W-cycle, "
                               "5-stage RK with residual smoothing, Roe flux + MUSCL-minmod, FP64",
                   "n": args.n, "levels": args.levels, "cell_stages_per_step": cs // args.steps,
                   "cd": round(cd, 6), "cl": round(cl, 6), "stage_tile_per_level": main_leg["tile"]},
        "roofline": roofline,
        "kernels": main_leg["kernels"],
        "cpu_baseline": cpu,
        "fp32": fp32,
        "solves": solves,
        "jst": jst,
    }
    if ns is not None:
        out["ns_cylinder"] = ns
    if cfg1 is not None:
        out["cfg1"] = cfg1
    if conv is not None:
        out["converged"] = conv
    if served is not None:
        out["serving"] = served
    print(json.dumps(out))


if __name__ == "__main__":
    main()
