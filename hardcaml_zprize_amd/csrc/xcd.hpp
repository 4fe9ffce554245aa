// Workgroup id -> tile, so that the workgroups the dispatcher sends to one
// XCD (round robin over MI355X's 8) take adjacent tiles: lines that adjacent
// tiles share (narrow column reads, runs written next to each other) meet in
// one L2 instead of being fetched or written back partly by two.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zp {

__device__ __forceinline__ uint32_t xcd_tile(uint32_t g, uint32_t tiles) {
  const uint32_t per = tiles >> 3;
  if ((tiles & 7) != 0 || per == 0) return g;
  return (g & 7) * per + (g >> 3);
}

}  // namespace zp
