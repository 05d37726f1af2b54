// XCD-aware block numbering for the fused sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace armon {

// MI355X dispatches the blocks of a grid round-robin over its 8 XCDs (block b runs on XCD
// b % 8), each XCD with its own L2. xcd_block renumbers blocks so that every XCD works through
// one contiguous eighth of the logical blocks: neighbouring blocks (consecutive tiles of a row,
// bands of neighbouring rows) then share an L2, which serves their overlapping halo cells,
// and each XCD streams its own contiguous part of the arrays. A trailing partial group of 8
// keeps its numbers.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int q = nb / kXcds;
    return b < q * kXcds ? (b % kXcds) * q + b / kXcds : b;
}

}  // namespace armon
