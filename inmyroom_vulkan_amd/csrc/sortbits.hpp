// Key-width arithmetic of the device sorts (host code, no HIP dependency).
#pragma once

#include <stdint.h>

namespace cdk {

// Bits that hold every value below `bound` (at least 1, at most 32).
inline uint32_t key_bits_below(uint64_t bound) {
    uint32_t b = 1;
    while (b < 32 && (uint64_t(1) << b) < bound) ++b;
    return b;
}

// Mask of the key bits the one-pass rank sort compares: key_bits rounded up
// to whole 4-bit digits.  Formed in 64 bits, since for key_bits 29..32 the
// rounded width is 32 and a 32-bit shift by 32 is undefined.
inline uint32_t rank_sort_mask(uint32_t key_bits) {
    const uint32_t width = key_bits >= 32 ? 32u : (key_bits + 3) / 4 * 4;
    return uint32_t((uint64_t(1) << width) - 1u);
}

}  // namespace cdk
