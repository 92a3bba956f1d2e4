"""The backprojection's tile launch order (wrt_adj_tile_order in csrc/wrt_fast.h, CPU only):
a permutation of the tiles of an n^3 volume, the tiles with a voxel at radius 1 - 1e-4 or
beyond first and the others after them, each part in index order.  The launches take every
tile from the list, so a tile missing or listed twice would leave voxels unwritten or write
them twice; the check runs the header's C under the address and undefined-behaviour
sanitizers."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wrt-project_amd", "csrc")

PROGRAM = r"""
#include <stdio.h>
#include <stdlib.h>
#include "wrt_fast.h"
int main(int argc, char** argv) {
  const int n = atoi(argv[1]), vpt = atoi(argv[2]);
  const long long nt = wrt_adj_tiles(n, vpt);
  int* order = (int*)malloc(sizeof(int) * (size_t)nt);
  if (!order) return 1;
  wrt_adj_tile_order(n, vpt, order);
  printf("%lld\n", nt);
  for (long long i = 0; i < nt; ++i) printf("%d\n", order[i]);
  free(order);
  return 0;
}
"""


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    d = tmp_path_factory.mktemp("order")
    src, exe = d / "order.c", d / "order"
    src.write_text(PROGRAM)
    cmd = ["gcc", "-std=gnu11", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", CSRC,
           "-I", "/opt/rocm/include", str(src), "-o", str(exe)]
    if subprocess.run(cmd + ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
                      capture_output=True).returncode != 0:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    return str(exe)


def _inside(n, vpt, t):
    # the tile's voxels (indices clamped into the volume, as the kernel clamps them) all
    # within radius 1 - 1e-4, from every voxel centre: brute force
    ntx = (n + 3) // 4
    tz, rem = divmod(t, ntx * ntx)
    tx, ty = divmod(rem, ntx)
    c = -1.0 + 1.0 / n + np.arange(n) * (2.0 / n)
    xs = c[np.minimum(np.arange(4 * tx, 4 * tx + 4), n - 1)]
    ys = c[np.minimum(np.arange(4 * ty, 4 * ty + 4), n - 1)]
    zs = c[np.minimum(np.arange(4 * vpt * tz, 4 * vpt * tz + 4 * vpt), n - 1)]
    r2 = xs[:, None, None] ** 2 + ys[None, :, None] ** 2 + zs[None, None, :] ** 2
    return bool(np.all(r2 <= (1.0 - 1e-4) ** 2))


@pytest.mark.parametrize("n,vpt", [(8, 2), (20, 2), (37, 2), (64, 2), (45, 3)])
def test_tile_order_is_a_boundary_first_permutation(program, n, vpt):
    r = subprocess.run([program, str(n), str(vpt)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    vals = [int(v) for v in r.stdout.split()]
    nt, order = vals[0], np.array(vals[1:])
    ntx, ntz = (n + 3) // 4, (n + 4 * vpt - 1) // (4 * vpt)
    assert nt == ntx * ntx * ntz
    assert np.array_equal(np.sort(order), np.arange(nt))
    inside = np.array([_inside(n, vpt, int(t)) for t in order])
    k = int(np.argmax(inside)) if inside.any() else nt
    assert not inside[:k].any() and inside[k:].all()  # boundary tiles, then inside ones
    assert np.all(np.diff(order[:k]) > 0) and np.all(np.diff(order[k:]) > 0)
    if n >= 20:
        assert 0 < k < nt  # both kinds occur
