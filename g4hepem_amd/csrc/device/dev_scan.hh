// dev_scan.hh -- block-wide exclusive scans for the stack compaction kernels
// (the slab shower's and the general track stack's).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hepem {

// Exclusive scans of a and b over the block (blockDim.x == NT) into xa, xb and
// the block sums of a, b and w. Every thread of the block must call it; a
// second call in the same block needs a barrier between the two (both use the
// same LDS words).
template <int NT>
__device__ inline void BlockScan2(uint32_t a, uint32_t b, uint32_t w, uint32_t& xa, uint32_t& xb, uint32_t& sa, uint32_t& sb,
                                  uint32_t& sw) {
  constexpr int kW = NT / 64;
  __shared__ uint32_t sh[3][kW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = a, y = b, z = w;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t ux = __shfl_up(x, o, 64), uy = __shfl_up(y, o, 64);
    if (lane >= o) {
      x += ux;
      y += uy;
    }
  }
  for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
  if (lane == 63) {
    sh[0][wv] = x;
    sh[1][wv] = y;
  }
  if (lane == 0) sh[2][wv] = z;
  __syncthreads();
  uint32_t ba = 0, bb = 0, ta = 0, tb = 0, tw = 0;
  for (int k = 0; k < kW; ++k) {
    const uint32_t u = sh[0][k], v = sh[1][k];
    if (k < wv) {
      ba += u;
      bb += v;
    }
    ta += u;
    tb += v;
    tw += sh[2][k];
  }
  xa = ba + x - a;
  xb = bb + y - b;
  sa = ta;
  sb = tb;
  sw = tw;
}

}  // namespace hepem
