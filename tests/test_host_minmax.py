"""The host builders' inline min / max (min_num, max_num in
rendiation_amd/csrc/bvh_build.h) against the C library's fminf / fmaxf,
whose results they must reproduce bit for bit (the box bounds of every
tree, and so the node-identical parity with the oracle, depend on them):
every pair of special values -- signed zeros, infinities, NaNs of either
sign, subnormals, ordinary numbers -- compiled with the library's flags."""

import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

SRC = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include "bvh_build.h"
int main() {
  const float inf = std::numeric_limits<float>::infinity(), nan = std::numeric_limits<float>::quiet_NaN();
  volatile float v[] = {0.0f, -0.0f, 1.0f, -1.0f, 2.5f, inf, -inf, nan, -nan, 1e-40f, -1e-40f,
                        std::numeric_limits<float>::max(), std::numeric_limits<float>::lowest()};
  const int n = sizeof(v) / sizeof(v[0]);
  int bad = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const float a = v[i], b = v[j];
      const float m1 = std::fmin(a, b), m2 = rdrt::min_num(a, b), x1 = std::fmax(a, b), x2 = rdrt::max_num(a, b);
      bad += std::memcmp(&m1, &m2, 4) != 0;
      bad += std::memcmp(&x1, &x2, 4) != 0;
    }
  std::printf("%d %d\n", n * n, bad);
  return 0;
}
"""


@pytest.mark.parametrize("opt", ["-O0", "-O3"])
def test_min_num_max_num_equal_libm(opt):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "mm.cpp"
        exe = Path(d) / "mm"
        src.write_text(SRC)
        subprocess.run([gxx, opt, "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                        "-I", str(ROOT / "rendiation_amd" / "csrc"), str(src), "-o", str(exe)], check=True)
        pairs, bad = (int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                                    check=True).stdout.split())
    assert pairs == 169
    assert bad == 0
