"""Host-side shape limits of the trainer, checked without a GPU: the helper
header is compiled with g++ into a tiny program that prints its decisions."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include "limits.h"
int main() {
  const long long cases[][3] = {
    {10, 4000000, 1},           // C3 shape: 4e7 positions, batched
    {16, 268435455, 1},         // 16 * (2^28 - 1) < 2^32: batched
    {16, 268435456, 1},         // 16 * 2^28 = 2^32: one tree at a time
    {16, 270000000, 1},         // the wrap case: 4.32e9 positions
    {2, 2147483647, 1},         // 2 * (2^31 - 1) < 2^32
    {3, 2147483647, 1},         // 3 * (2^31 - 1) >= 2^32
    {10, 1000, 0},              // batching not wanted
    {1, 3000000000LL, 1},       // one output never batches
    {40, 1000000, 1},           // more classes than a batch: batches of 16
    {17, 268435455, 1},         // batches of 16 that fit
    {40, 268435456, 1},         // 16 * 2^28 = 2^32: one tree at a time
  };
  for (auto& c : cases)
    std::printf("%d\n", bst::class_batch_trees((int)c[0], c[1], c[2] != 0));
  return 0;
}
"""


def test_class_batching_keeps_positions_32_bit(tmp_path):
    src = tmp_path / "limits_check.cpp"
    src.write_text(PROG)
    exe = tmp_path / "limits_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "boosters_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [10, 16, 1, 1, 2, 1, 1, 1, 16, 16, 1]
