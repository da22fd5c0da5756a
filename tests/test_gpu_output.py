"""Parity of the HIP upload_solut kernel (upload_solut.f90:89-241: rho, u,
v, w, cp, Mach, total-pressure loss, and for laminar flow the energy and
gradient slots) against the oracle on a flowed state: FP64 to rounding, the
FP32 bulk path within 1e-5 in cp and Mach."""
import dataclasses

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Case
from oracle.oracle import OracleCase
from tests.gpu_common import make_pair

pytestmark = pytest.mark.gpu


def _cmp(fg, fo, tol):
    assert fg.keys() == fo.keys()
    worst = 0.0
    for k in fo:
        scale = max(np.abs(fo[k]).max(), 1.0)
        e = float(np.abs(fg[k] - fo[k]).max() / scale)
        print(f"  {k:8s} {e:.3e}")
        worst = max(worst, e)
    assert worst < tol


@pytest.mark.parametrize("eqntype", ["e", "n"])
def test_upload_solut_parity(require_gpu, eqntype):
    gpu, orc, prm, aref = make_pair(n=64, nz=2, nlevel=3, mach=0.4, alpha=2.0, eqntype=eqntype,
                                    reynolds=40.0)
    for c in (gpu, orc):
        c.init_flow()
        c.solve([0, 0, 20])
    for lvl in (2, 3):
        fg = gpu.upload_solut(lvl)
        fo = orc.upload_solut(lvl)
        assert len(fg) == (11 if eqntype == "n" else 7)
        print(f"level {lvl}:")
        _cmp(fg, fo, 1e-12)
    # a given total pressure
    _cmp(gpu.upload_solut(3, p_tot=9.0), orc.upload_solut(3, p_tot=9.0), 1e-12)
    fg = gpu.upload_solut(3)
    assert fg["cp"].min() < -0.2 and fg["cp"].max() > 0.5     # suction and stagnation on the cylinder


def test_upload_solut_fp32(require_gpu):
    prm, levels, bdy, aref = configs.cylinder2d(n=64, nz=2, nlevel=3, mach=0.4, alpha=2.0)
    geoms = [level_geometry(s, bdy) for s in levels]
    orc = OracleCase(prm, geoms, mxcyc=50)
    with Case(dataclasses.replace(prm, precision=32), levels, geoms=geoms, mxcyc=50) as g32:
        for c in (g32, orc):
            c.init_flow()
            c.solve([0, 0, 20])
        fg = g32.upload_solut(3)
    fo = orc.upload_solut(3)
    for k in ("cp", "mach", "pt-loss"):
        e = float(np.abs(fg[k] - fo[k]).max())
        print(f"FP32 {k}: {e:.3e}")
        assert e < 1e-5
