"""The stage-kernel autotuner (mge_autotune): it picks tile rows (4 or 8)
and a k-chunk length in [1, nk] for every level, caches them per shape, and
the launch shape changes no arithmetic -- the same cycles with any forced
tile rows and chunk length give bit-identical states (JST and Roe, FP64 and
FP32)."""
import dataclasses

import numpy as np
import pytest

from code_mge_amd import configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("distype,precision", [("c", 64), ("c", 32), ("r2", 64), ("r2", 32)])
def test_autotune_picks_valid_chunks_and_changes_no_bits(require_gpu, distype, precision):
    prm, levels, bdy, _ = configs.sphere3d(n=48, nlevel=3, box=4.0, distype=distype)
    prm = dataclasses.replace(prm, precision=precision)
    geoms = [level_geometry(s, bdy) for s in levels]
    states = []
    for mode in ("tuned", (4, 1), (8, 3), (4, "nk"), (8, "nk")):
        with Case(prm, levels, geoms=geoms, mxcyc=8) as c:
            c.init_flow()
            if mode == "tuned":
                t = c.autotune()
                print(f"{distype} FP{precision} tuned:", t)
                for lvl, (tj, kc, us) in t.items():
                    assert tj in (4, 8) and 1 <= kc <= levels[lvl - 1].n[2] and us > 0.0
                    assert c.stage_tile(lvl) == (tj, kc)
                assert c.autotune() == t   # cached per shape
            else:
                tj, kc = mode
                for lvl in range(1, 4):
                    c.set_tile(lvl, tj, levels[lvl - 1].n[2] if kc == "nk" else kc)
                    assert c.stage_tile(lvl)[0] == tj
            r, _ = c.cycles(3, 3, ncyc0=1)
            states.append((r, c.download(3, "w"), c.download(3, "p")))
    for r, w, p in states[1:]:
        assert r == states[0][0]
        assert np.array_equal(w, states[0][1]) and np.array_equal(p, states[0][2])
