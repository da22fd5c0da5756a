"""mge_upload_solut_noalloc against mge_upload_solut through the C ABI: the
allocation-free entry point stages the output fields through the library's
fixed device scratch array in chunks of whole k-planes and must return the
same bits.  Two calls in a row on one GPU share that array: the first
(a 96^3 level, several chunks) and the second (a smaller laminar FP32
level, one chunk) are both checked, then the first case's coarse level."""
import numpy as np
import pytest

from code_mge_amd import _ffi, configs
from code_mge_amd.grid import LevelSpec, level_geometry
from code_mge_amd.params import Params
from code_mge_amd.solver import Case

pytestmark = pytest.mark.gpu

SCRATCH = 32 << 20


def _same(case, lvl):
    a = case.upload_solut(lvl)
    b = case.upload_solut(lvl, noalloc=True)
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), (lvl, k)
    return a


def test_noalloc_upload_solut_equals_allocating_entry_point(require_gpu):
    prm, levels, bdy, _ = configs.sphere3d(n=96, nlevel=3)
    geoms = [level_geometry(s, bdy) for s in levels]
    n = levels[-1].n
    assert n[0] * n[1] * n[2] * 7 * 8 > SCRATCH      # the first call takes several chunks
    print("case 1: 96^3 sphere, FP64 Euler, 3 W-cycles", flush=True)
    with Case(prm, levels, geoms=geoms, mxcyc=8) as big:
        big.init_flow()
        big.cycles(3, 3)
        f1 = _same(big, 3)
        assert np.isfinite(f1["mach"]).all() and f1["mach"].max() > 0.3
        print("case 2: 64x64x4 cylinder, FP32 laminar", flush=True)
        prm2, lv2, b2, _ = configs.cylinder2d(n=64, nz=4, nlevel=2, precision=32, eqntype="n", reynolds=40.0)
        with Case(prm2, lv2, geoms=[level_geometry(s, b2) for s in lv2], mxcyc=8) as small:
            small.init_flow()
            small.cycles(2, 2)
            f2 = _same(small, 2)
            assert len(f2) == 11
        print("case 1 again: its 48^3 level", flush=True)
        _same(big, 2)


def test_noalloc_rejects_a_plane_beyond_the_scratch(require_gpu):
    """a 640x640 laminar level: one k-plane of its 11 fields is 36 MB"""
    spec = [LevelSpec(n=(640, 640, 2), lo=(0, 0, 0), hi=(1, 1, 0.01), ratio=(2, 2, 1), bctype=(-20,) * 6)]
    prm = Params(mach=0.3, eqntype="n", reynolds=100.0)
    print("case 3: 640x640x2 laminar level, expect MGE_ERR_ARG", flush=True)
    with Case(prm, spec, geoms=[level_geometry(spec[0], None)], mxcyc=4) as c:
        c.init_flow()
        with pytest.raises(_ffi.MgeError) as ei:
            c.upload_solut(1, noalloc=True)
        assert ei.value.code == -20 and "scratch" in str(ei.value)
        f = c.upload_solut(1)        # the allocating entry point still serves it
        assert f["rho"].shape == (2, 640, 640)
