"""Helpers shared by the GPU parity tests: build one case on the HIP path
and the same case on the oracle, from the same geometry."""
import numpy as np

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Case
from oracle.oracle import OracleCase


def make_pair(cfg="cylinder2d", mxcyc=200, **kw):
    prm, levels, bdy, aref = getattr(configs, cfg)(**kw)
    geoms = [level_geometry(s, bdy) for s in levels]
    gpu = Case(prm, levels, geoms=geoms, mxcyc=mxcyc)
    orc = OracleCase(prm, geoms, mxcyc=mxcyc)
    return gpu, orc, prm, aref


def perturbed_state(case_like, lvl, prm, seed=7, amp=0.01):
    """freestream x (1 + amp*xi), xi ~ U[-1,1] from PCG64(seed), with the
    pressure from the reference EOS (depvars.f90:37-52), same bits on both
    sides."""
    ib1, jb1, kb1 = case_like.dims(lvl)
    d = prm.derived()
    rng = np.random.Generator(np.random.PCG64(seed))
    w = np.empty((5, kb1, jb1, ib1))
    base = [d["rho0"], d["rho0"] * d["u0"], d["rho0"] * d["v0"], d["rho0"] * d["w0"],
            d["rho0"] * d["h0"] - d["p0"]]
    for n in range(5):
        w[n] = base[n] * (1.0 + amp * rng.uniform(-1.0, 1.0, size=(kb1, jb1, ib1)))
    s = 0.0 + w[1] * w[1]
    s = s + w[2] * w[2]
    s = s + w[3] * w[3]
    qq = 0.5 * s / w[0]
    p = (prm.gamma - 1.0) * np.where(w[4] > qq, w[4] - qq, 0.0)
    return w, p


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def max_rel_to_field(a, b, comp_axis=0):
    """max |a-b| / max|b| per component (robust where fields cross zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 4:
        out = 0.0
        for n in range(a.shape[0]):
            s = max(np.abs(b[n]).max(), 1e-300)
            out = max(out, float(np.abs(a[n] - b[n]).max() / s))
        return out
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s)


def cell_rel_err(a, b, prm):
    """per-cell relative error of a 5-component state: |a-b| / (|b| + s_n)
    with s_n the freestream magnitude of component n (rho, rho|U| for the
    three momenta, rho*E), so components that cross zero are measured
    against their physical scale."""
    d = prm.derived()
    U = (d["u0"] ** 2 + d["v0"] ** 2 + d["w0"] ** 2) ** 0.5
    scale = [d["rho0"], d["rho0"] * U, d["rho0"] * U, d["rho0"] * U, d["rho0"] * d["h0"] - d["p0"]]
    out = 0.0
    for n in range(5):
        e = np.abs(a[n] - b[n]) / (np.abs(b[n]) + scale[n])
        out = max(out, float(e.max()))
    return out
