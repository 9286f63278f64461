// Device helpers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "bh_common.hpp"

namespace bh {

// Block-wide exclusive scan of one u32 per thread (blockDim.x <= 1024, a multiple of
// 64).  Returns the exclusive prefix; *total receives the block sum.  Wave scans by
// shuffles, one wave scans the wave totals (scratch[32 .. 48)); three barriers.
__device__ inline u32 block_exscan(u32 v, u32 *scratch, u32 *total)
{
    u32 const lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    u32 inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        u32 const up = __shfl_up(inc, o);
        if (static_cast<int>(lane) >= o) inc += up;
    }
    if (lane == 63u) scratch[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        u32 x = lane < nw ? scratch[lane] : 0u;
        for (int o = 1; o < 16; o <<= 1) {
            u32 const up = __shfl_up(x, o);
            if (static_cast<int>(lane) >= o) x += up;
        }
        if (lane < nw) scratch[32 + lane] = x; // inclusive over waves
    }
    __syncthreads();
    u32 const base = wave ? scratch[32 + wave - 1] : 0u;
    *total = scratch[32 + nw - 1];
    __syncthreads(); // scratch free for the caller's next scan
    return base + inc - v;
}

} // namespace bh
