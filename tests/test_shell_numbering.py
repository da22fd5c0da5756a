"""The ghost-shell numbering of k_bc_faces (shell_cell_by_depth in
code_mge_amd/csrc/mge_kernels.h) checked on the host: for block shapes with
one-cell, odd, even and large directions, thread t < shell_cells(d) maps to
a ghost cell inside the array, no two threads share a cell, together they
cover the whole shell (every cell with a coordinate in {0, 1, e, b}), and
t = shell_cells(d) maps to nothing.  A wrong mapping would make the faces'
launch miss ghost cells or write outside the shell."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "mge_kernels.h"
using namespace mge;
static Dims mk(int ni, int nj, int nk) {
  Dims d{};
  d.ni = ni; d.nj = nj; d.nk = nk;
  d.il = ni + 1; d.jl = nj + 1; d.kl = nk + 1;
  d.ie = ni + 2; d.je = nj + 2; d.ke = nk + 2;
  d.ib = ni + 3; d.jb = nj + 3; d.kb = nk + 3;
  d.sy = d.ib + 1; d.sz = (long)(d.ib + 1) * (d.jb + 1); d.E = d.sz * (d.kb + 1);
  return d;
}
int main() {
  const int dims[][3] = {{32, 32, 32}, {64, 48, 16}, {5, 7, 3}, {2, 2, 2}, {1, 9, 4},
                         {12, 1, 6}, {8, 6, 1}, {1, 1, 1}, {130, 66, 34}};
  for (const auto &n : dims) {
    const Dims d = mk(n[0], n[1], n[2]);
    const long N = shell_cells(d);
    std::vector<char> seen(d.E, 0);
    long ghosts = 0;
    for (int z = 0; z <= d.kb; z++)
      for (int y = 0; y <= d.jb; y++)
        for (int x = 0; x <= d.ib; x++)
          if (x <= 1 || x >= d.ie || y <= 1 || y >= d.je || z <= 1 || z >= d.ke) ghosts++;
    if (N != ghosts) { printf("FAIL %dx%dx%d: shell_cells %ld, ghost cells %ld\n", n[0], n[1], n[2], N, ghosts); return 1; }
    for (long t = 0; t < N; t++) {
      int x = -1, y = -1, z = -1;
      if (!shell_cell_by_depth(d, t, x, y, z)) { printf("FAIL %dx%dx%d: t=%ld unmapped\n", n[0], n[1], n[2], t); return 1; }
      const bool in = x >= 0 && x <= d.ib && y >= 0 && y <= d.jb && z >= 0 && z <= d.kb;
      const bool ghost = x <= 1 || x >= d.ie || y <= 1 || y >= d.je || z <= 1 || z >= d.ke;
      if (!in || !ghost) { printf("FAIL %dx%dx%d: t=%ld -> (%d,%d,%d)\n", n[0], n[1], n[2], t, x, y, z); return 1; }
      const long q = ix(d, x, y, z);
      if (seen[q]++) { printf("FAIL %dx%dx%d: (%d,%d,%d) twice\n", n[0], n[1], n[2], x, y, z); return 1; }
    }
    int x, y, z;
    if (shell_cell_by_depth(d, N, x, y, z)) { printf("FAIL %dx%dx%d: t=N mapped\n", n[0], n[1], n[2]); return 1; }
    printf("%dx%dx%d: %ld ghost cells, each once\n", n[0], n[1], n[2], N);
  }
  printf("OK\n");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc is not installed")
def test_shell_cell_by_depth_is_a_bijection_onto_the_shell(tmp_path):
    src = tmp_path / "shell_numbering.hip"
    src.write_text(PROGRAM)
    exe = tmp_path / "shell_numbering"
    inc = os.path.join(ROOT, "code_mge_amd", "csrc")
    subprocess.run(["hipcc", "-O1", "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
