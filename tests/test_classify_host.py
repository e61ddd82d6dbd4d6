"""The reduction's exit classification equals the path kernels' (host build, no GPU).

The reduction finds an FP64 exit's side again from its exit point with classify_coord_c, which
takes the global rectangle as centre and half-sides (computed once on the host, as the engine
does) and tests max(|x - cx| - hx, |y - cy| - hy) >= -1e-8 in place of rect_dist's signed
distance.  A small program compiled from the engine's own header compares it with classify_coord
(the rule the oracle restates, Stencil::G_update Meshes/stencil.cpp:246-301 after the global test
of Solvers/EMFKACSolver.hpp:559) on points placed where a difference would show: on the edges
and corners of both rectangles, a few ulps and 1e-8 either side of them, inside and outside.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include "engine_kernels.h"
using namespace pdds;

static uint64_t st = 0x9E3779B97F4A7C15ULL;
static uint64_t next_u64() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static double uni(double a, double b) { return a + (b - a) * ((next_u64() >> 11) * (1.0 / 9007199254740992.0)); }
static double ulps(double v, int k) {
    for (int i = 0; i < (k < 0 ? -k : k); ++i) v = std::nextafter(v, k < 0 ? -INFINITY : INFINITY);
    return v;
}
// a coordinate near one of the edges e[0..n), or anywhere in [lo, hi]
static double coord(const double *e, int n, double lo, double hi) {
    const int c = (int)(next_u64() % 6);
    const double b = e[next_u64() % n];
    switch (c) {
        case 0: return b;
        case 1: return ulps(b, (int)(next_u64() % 9) - 4);
        case 2: return ulps(b + 1e-8, (int)(next_u64() % 9) - 4);
        case 3: return ulps(b - 1e-8, (int)(next_u64() % 9) - 4);
        case 4: return b + uni(-3e-8, 3e-8);
        default: return uni(lo, hi);
    }
}

int main() {
    const double globs[][4] = {{-1, -1, 1, 1}, {0, 0, 2, 1}, {-0.3, 0.1, 0.7, 0.45}, {1e-3, -5, 3.0, 7.25}};
    long n = 0, diff = 0, global = 0, sides = 0;
    for (const auto &g : globs) {
        const double gc[4] = {0.5 * (g[0] + g[2]), 0.5 * (g[1] + g[3]), g[2] - 0.5 * (g[0] + g[2]),
                              g[3] - 0.5 * (g[1] + g[3])};
        for (int r = 0; r < 400; ++r) {
            // a stencil rectangle inside glob, sharing some of its edges
            double s[4];
            s[0] = (next_u64() & 1) ? g[0] : uni(g[0], 0.5 * (g[0] + g[2]));
            s[1] = (next_u64() & 1) ? g[1] : uni(g[1], 0.5 * (g[1] + g[3]));
            s[2] = (next_u64() & 1) ? g[2] : uni(0.5 * (g[0] + g[2]), g[2]);
            s[3] = (next_u64() & 1) ? g[3] : uni(0.5 * (g[1] + g[3]), g[3]);
            const double ex[5] = {g[0], g[2], s[0], s[2], gc[0]};
            const double ey[5] = {g[1], g[3], s[1], s[3], gc[1]};
            for (int i = 0; i < 2500; ++i) {
                const double x = coord(ex, 5, g[0] - 0.1, g[2] + 0.1), y = coord(ey, 5, g[1] - 0.1, g[3] + 0.1);
                const int a = classify_coord(s, g, x, y), b = classify_coord_c(s, gc, x, y);
                ++n;
                diff += a != b;
                global += a == 4;
                sides += a >= 0 && a < 4;
            }
        }
    }
    std::printf("%ld %ld %ld %ld\n", n, diff, global, sides);
    return 0;
}
"""


def test_classify_coord_c_equals_classify_coord(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    src = tmp_path / "classify.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "classify"
    # host code only, with the engine's host flags (no contraction: the same roundings)
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "pdds_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=120).stdout
    n, diff, glob_hits, side_hits = (int(v) for v in out.split())
    assert n == 4 * 400 * 2500
    assert diff == 0
    # the samples reach both outcomes the rule separates
    assert glob_hits > n // 20 and side_hits > n // 20
