"""CPU: the order in which the lane-row pass 4 kernels' blocks take the grid rows
(plan_block_order / emit_block_order, miniisosurface_amd/csrc/fe_rows.h) is a permutation of the
blocks for every row width and plane size, so that every row is written exactly once; and the
row kernels' (j, k) of grid row r0 + q from one division per wavefront (split_row) equals
r mod ny, r / ny. The header is compiled for the host with hipcc; the kernels' own parity tests
cover both on the GPU."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r"""
#include "fe_rows.h"
#include <cstdio>
#include <vector>
int main()
{
    long tiled = 0;
    for (int64_t W = 1; W <= 64; ++W)
        for (int64_t ny : {1, 7, 64, 96, 128, 256, 384, 512, 640, 1000, 1024, 2048, 3072, 4096, 24576})
            for (int64_t nz : {1, 2, 3, 5, 7, 9, 64, 127, 1024}) {
                int64_t wp = 1;
                while (wp < W) wp <<= 1;
                const int64_t per_block = 4 * (64 / wp);
                const int64_t n = (ny * nz + per_block - 1) / per_block;
                miso::Geom g{};
                g.ny = ny;
                g.nz = nz;
                g.W = W;
                g.bo = miso::plan_block_order(ny, nz, W);
                if (g.bo.n0 > n) { printf("W=%ld ny=%ld nz=%ld: n0 %u > %ld blocks\n", W, ny, nz, g.bo.n0, n); return 1; }
                std::vector<char> seen(n, 0);
                for (int64_t b = 0; b < n; ++b) {
                    const int64_t a = b < (int64_t)g.bo.n0 ? miso::apply_block_order(g.bo, (uint32_t)b) : b;
                    if (a < 0 || a >= n || seen[a]) {
                        printf("W=%ld ny=%ld nz=%ld: block %ld -> %ld\n", W, ny, nz, b, a);
                        return 1;
                    }
                    seen[a] = 1;
                    tiled += a != b;
                }
            }
    // split_row against division, across 32-bit and 64-bit row indices and small ny (carries)
    for (int64_t ny : {1ll, 2ll, 3ll, 7ll, 63ll, 64ll, 65ll, 1000ll, 4096ll, 70000ll, 5000000000ll})
        for (int64_t r0 : {0ll, 1ll, 63ll, 64ll, 4095ll, 123456ll, 4294967295ll, 4294967296ll, 9000000000123ll})
            for (int q = 0; q < 64; ++q) {
                const miso::RowJK p = miso::split_row(r0, q, ny);
                const int64_t r = r0 + q;
                if (p.j != r % ny || p.k != r / ny) {
                    printf("split_row(%ld, %d, %ld) = (%ld, %ld)\n", r0, q, ny, p.j, p.k);
                    return 1;
                }
            }
    printf("tiled %ld\n", tiled);
    return 0;
}
"""


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_block_order_is_a_permutation():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "order.hip")
        exe = os.path.join(tmp, "order")
        with open(src, "w") as f:
            f.write(PROGRAM)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
                        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "miniisosurface_amd", "csrc"),
                        "-o", exe, src], check=True, capture_output=True, timeout=300)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stdout
        # the tile order is in use for most of these geometries, not only the identity
        assert int(out.stdout.split()[-1]) > 10 ** 8
