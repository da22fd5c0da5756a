"""Restart files (mge_write_solut / mge_read_solut; write_solut and
read_solut of srcf90/wsolut.f90:47-93, srcf90/rsolut.f90:1-41):

- the binary variant restores every level exactly: a run restarted from it
  continues bit for bit like the run that wrote it;
- the text variant has the reference's layout -- a "(5I5)" header (block
  code, low(4) = 0 0 0 1, upp(4) = ib jb kb 5) and w in "(5f16.7)", i
  fastest -- restores w to its 7 decimals, p by the EOS, and a run
  restarted from it stays close to the continued one;
- a file of another grid is refused with the reference's "INCONSISTENT
  FILE" condition."""
import numpy as np
import pytest

from code_mge_amd import _ffi, configs
from code_mge_amd.grid import level_geometry
from code_mge_amd.solver import Case

pytestmark = pytest.mark.gpu


def _case(n=64, **kw):
    prm, levels, bdy, aref = configs.cylinder2d(n=n, nz=4, nlevel=3, **kw)
    geoms = [level_geometry(s, bdy) for s in levels]
    return Case(prm, levels, geoms=geoms, mxcyc=40), levels


def test_binary_restart_continues_bit_for_bit(require_gpu, tmp_path):
    a, levels = _case()
    a.init_flow()
    a.cycles(3, 10, ncyc0=1)
    a.write_solut(tmp_path / "cyl_restart.bin", binary=True)
    ra, _ = a.cycles(3, 5, ncyc0=11)
    b, _ = _case()
    b.read_solut(tmp_path / "cyl_restart.bin", binary=True)
    rb, _ = b.cycles(3, 5, ncyc0=11)
    print("continued", ra, "restarted", rb)
    assert ra == rb
    for lvl in (1, 2, 3):
        assert np.array_equal(a.download(lvl, "w"), b.download(lvl, "w"))
        assert np.array_equal(a.download(lvl, "p"), b.download(lvl, "p"))
    a.close()
    b.close()


def test_text_restart_layout_and_values(require_gpu, tmp_path):
    a, levels = _case()
    a.init_flow()
    a.cycles(3, 6, ncyc0=1)
    path = tmp_path / "cyl_restart.dat"
    a.write_solut(path)
    lines = open(path).read().splitlines()
    ni, nj, nk = levels[0].n
    assert lines[0] == "%5d%5d%5d%5d%5d" % (101, 0, 0, 0, 1)
    assert lines[1] == "%5d%5d%5d%5d" % (ni + 3, nj + 3, nk + 3, 5)
    assert all(len(l) == 80 for l in lines[2:6])      # five f16.7 fields per line
    b, _ = _case()
    b.read_solut(path)
    for lvl in (1, 2, 3):
        wa, wb = a.download(lvl, "w"), b.download(lvl, "w")
        assert np.max(np.abs(wa - wb)) <= 5.1e-8       # rounded to 7 decimals
        g = 1.4
        pb = b.download(lvl, "p")
        q = 0.5 * (wb[1] ** 2 + wb[2] ** 2 + wb[3] ** 2) / wb[0]
        assert np.allclose(pb, (g - 1.0) * np.maximum(wb[4] - q, 0.0), rtol=1e-14, atol=1e-14)
    ra, _ = a.cycles(3, 5, ncyc0=7)
    rb, _ = b.cycles(3, 5, ncyc0=7)
    print("continued", ra, "restarted from text", rb)
    assert abs(ra - rb) < 1e-3 * ra
    c, _ = _case(n=32)
    with pytest.raises(_ffi.MgeError) as e:
        c.read_solut(path)
    assert "INCONSISTENT" in str(e.value)
    for x in (a, b, c):
        x.close()
