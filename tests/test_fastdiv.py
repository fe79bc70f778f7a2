"""The kernels' division by a launch-time divisor (natrium_amd/csrc/nat_fastdiv.hpp) against
the C++ `/` operator: every numerator below 2^22 for divisors 1..257, the cell counts of the
configured meshes, powers of two and their neighbours; random 32-bit numerators (including
2^32 - 1) for divisors up to 2^31 - 1. Compiled with g++ from the same header the kernels use
(host branch of the multiply-high)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>
#include "nat_fastdiv.hpp"
int main() {
  long bad = 0, checked = 0;
  auto check = [&](uint32_t d, uint32_t n) {
    static nat::FastDivU f;
    if (f.d != d) f.set(d);
    if (f.div(n) != n / d) {
      if (bad < 5) std::printf("d=%u n=%u got %u want %u\n", d, n, f.div(n), n / d);
      bad++;
    }
    checked++;
  };
  std::vector<uint32_t> ds;
  for (uint32_t d = 1; d <= 257; d++) ds.push_back(d);
  for (uint32_t d : {24u, 48u, 128u, 255u, 256u, 1000u, 4096u, 65535u, 65536u, 65537u}) ds.push_back(d);
  for (int k = 2; k < 31; k++) { ds.push_back((1u << k) - 1); ds.push_back(1u << k); ds.push_back((1u << k) + 1); }
  ds.push_back(2147483647u);
  for (uint32_t d : ds) {
    const uint32_t lim = d <= 257 ? (1u << 22) : (1u << 16);
    for (uint32_t n = 0; n < lim; n++) check(d, n);
    check(d, 0xffffffffu);
    check(d, 0x80000000u);
    check(d, d * (0xffffffffu / d));
    check(d, d * (0xffffffffu / d) - 1);
  }
  std::mt19937_64 rng(1234);
  for (int i = 0; i < 4000000; i++) {
    uint32_t d = (uint32_t)(rng() % 2147483647u) + 1u;
    if (i % 3 == 0) d = (uint32_t)(rng() % 5000u) + 1u;
    check(d, (uint32_t)rng());
  }
  std::printf("%ld %ld\n", checked, bad);
  return bad != 0;
}
"""


def test_fastdiv_matches_integer_division(tmp_path):
    src = tmp_path / "fd.cpp"
    src.write_text(SRC)
    exe = tmp_path / "fd"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "natrium_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    checked, bad = map(int, r.stdout.split()[-2:])
    assert bad == 0 and checked > 10 ** 9 // 1000
