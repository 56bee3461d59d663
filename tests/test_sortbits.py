"""Key-width helpers of the device sorts (csrc/sortbits.hpp), on the CPU.

The one-pass rank sort compares keys under a mask of whole 4-bit digits;
for 29..32 key bits that mask is all 32 bits.  A 32-bit shift by 32 is
undefined (on x86 it gives a zero mask, which compares no bits and leaves the
keys unsorted), so the harness is built with UBSan and must run clean.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "inmyroom_vulkan_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_rank_sort_mask_and_key_bits(tmp_path):
    src = tmp_path / "bits.cpp"
    src.write_text(r'''
#include "sortbits.hpp"
#include <cstdio>
int main() {
    for (uint32_t b = 1; b <= 32; ++b) std::printf("m %u %u\n", b, cdk::rank_sort_mask(b));
    const unsigned long long bounds[] = {0, 1, 2, 3, 4, 5, 255, 256, 257, 4096, 4097, 65535ull << 16,
                                        4097ull << 16, 32768ull << 16, 1ull << 31, (1ull << 31) + 1,
                                        1ull << 32, 1ull << 40};
    for (unsigned long long x : bounds) std::printf("k %llu %u\n", x, cdk::key_bits_below(x));
    return 0;
}
''')
    exe = tmp_path / "bits"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=undefined", "-fno-sanitize-recover=all", "-I", CSRC,
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    assert "runtime error" not in out.stderr
    masks, kb = {}, {}
    for line in out.stdout.splitlines():
        tag, a, b = line.split()
        (masks if tag == "m" else kb)[int(a)] = int(b)
    for b in range(1, 33):
        width = min(32, (b + 3) // 4 * 4)
        assert masks[b] == (1 << width) - 1, b
        assert masks[b] >> (b - 1) & 1, "the mask keeps the key's top bit"
    assert masks[29] == masks[31] == 0xFFFFFFFF
    for x, bits in kb.items():
        want = 1
        while want < 32 and (1 << want) < x:
            want += 1
        assert bits == want, x
    # the frame's pair key (first << 16 | second) for 4,097..32,768 entities
    # needs 29..31 bits: the case the full mask must cover
    assert kb[4097 << 16] == 29 and kb[32768 << 16] == 31
