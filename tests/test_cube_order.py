"""The engine's tiled cube order (gc_kernels.h: cube_tile_shift, cube_row,
cube_coords), compiled for the host and checked without a GPU:

* key = cube_off + x + nx * cube_row(y, z) is a bijection of the grid onto
  [cube_off, cube_off + nx ny nz), and cube_coords inverts it;
* x is fastest: the three cubes of a row, x - 1 .. x + 1, are consecutive
  keys (the neighbour scans read them as one slot range);
* in a tiled grid the rows at z - 1 and z + 1 are at most CUBE_TILE_CUBES
  cubes away, and a grid whose y range fits one tile keeps the plain
  x + nx (y + ny z) order (the reference's);
over the benchmark grids, grids with a short last tile, and random grids."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CHECKER = r"""
#include "gc_kernels.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv)
{
    int bad = 0, tiled = 0;
    for (int a = 1; a + 2 < argc; a += 3) {
        const int nx = std::atoi(argv[a]), ny = std::atoi(argv[a + 1]), nz = std::atoi(argv[a + 2]);
        gc::GcRep R{};
        R.nc[0] = nx; R.nc[1] = ny; R.nc[2] = nz; R.cube_off = 11;
        R.tys = gc::cube_tile_shift(nx, ny);
        const bool tile = (1 << R.tys) < ny;
        tiled += tile;
        const long K = (long)nx * ny * nz;
        std::vector<char> seen(K, 0);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                const int row = gc::cube_row(R, y, z);
                if (!tile && row != y + ny * z) ++bad;
                if (tile && z + 1 < nz && (gc::cube_row(R, y, z + 1) - row) * (long)nx > gc::CUBE_TILE_CUBES) ++bad;
                for (int x = 0; x < nx; ++x) {
                    const long k = R.cube_off + x + (long)nx * row;
                    if (k < R.cube_off || k >= R.cube_off + K || seen[k - R.cube_off]) { ++bad; continue; }
                    seen[k - R.cube_off] = 1;
                    int cx, cy, cz;
                    gc::cube_coords(R, (int)k, cx, cy, cz);
                    if (cx != x || cy != y || cz != z) ++bad;
                }
            }
    }
    std::printf("%d %d\n", bad, tiled);
    return 0;
}
"""


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("cube_order")
    src = d / "check.cpp"
    src.write_text(CHECKER)
    exe = d / "check"
    r = subprocess.run([HIPCC, "-std=c++17", "--offload-arch=gfx950", "-O1",
                        "-I", os.path.join(ROOT, "gpucell_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run(exe, grids):
    args = [str(v) for g in grids for v in g]
    out = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), int(out[1])


def test_benchmark_grids(checker):
    # 10M duct (skin and per-step cubes), 1M duct, stenosis, 32^3 box, the
    # reference's default domain, a tile's worth of rows exactly
    grids = [(242, 92, 92), (274, 104, 104), (136, 36, 36), (153, 42, 42), (176, 47, 47),
             (21, 21, 21), (20, 22, 22), (256, 16, 5), (256, 17, 5)]
    bad, tiled = run(checker, grids)
    assert bad == 0
    assert tiled >= 5


def test_random_grids(checker):
    rng = np.random.default_rng(7)
    grids = [(int(rng.integers(1, 700)), int(rng.integers(1, 120)), int(rng.integers(1, 40)))
             for _ in range(60)]
    bad, tiled = run(checker, grids)
    assert bad == 0
    assert 0 < tiled < len(grids)
