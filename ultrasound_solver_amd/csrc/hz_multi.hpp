// Fused multi-scene launches.
//
// A kernel that takes part in lockstep batches is written as a descriptor K:
//   K::Args     its launch arguments as one struct;
//   K::THREADS  its workgroup size;
//   K::run(a, b) the body, with the workgroup's coordinates b in its own grid
//                (blockIdx / gridDim of a single launch);
//   K::single(grid, stream, a) launches the named single-scene kernel.
// k_multi<K> runs several scenes' launches of K as one grid: workgroup
// blockIdx.x belongs to scene s when first[s] <= blockIdx.x < first[s + 1],
// and runs K::run on the scene's own arguments (read from a table in device
// memory) with its coordinates in the scene's own grid.  Every scene's range
// starts at a multiple of 8, so a workgroup's XCD (its index modulo 8) is
// the one it has in the scene's single launch, and XCD-aware tile orders
// keep their meaning.
#pragma once

#include <hip/hip_runtime.h>

namespace hz {

// Unsigned, as blockIdx and gridDim are, so that a body's index arithmetic
// is the same as in a kernel that reads them directly.
struct Blk {
    unsigned x, y, z;     // blockIdx in the scene's grid
    unsigned gx, gy, gz;  // the scene's gridDim
};

__device__ __forceinline__ Blk blk_here() {
    return {blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z};
}

constexpr int MULTI_MAX = 32;  // scenes per fused launch

// Passed by value (kernel arguments live in SGPRs), so the scene lookup is
// scalar work.
struct MultiGrid {
    int count;
    int first[MULTI_MAX + 1];  // first workgroup of each scene; first[count] = total
    int gx[MULTI_MAX], gy[MULTI_MAX], gz[MULTI_MAX];
};

template <class K>
__global__ __launch_bounds__(K::THREADS) void k_multi(const typename K::Args *__restrict__ table, MultiGrid mg) {
    const int bid = (int)blockIdx.x;
    int s = 0;
#pragma unroll
    for (int i = 1; i < MULTI_MAX; ++i) s += (i < mg.count && bid >= mg.first[i]) ? 1 : 0;
    const unsigned local = (unsigned)(bid - mg.first[s]);
    const unsigned gx = (unsigned)mg.gx[s], gy = (unsigned)mg.gy[s], gz = (unsigned)mg.gz[s];
    if (local >= gx * gy * gz) return;  // padding to the next multiple of 8
    const Blk b{local % gx, (local / gx) % gy, local / (gx * gy), gx, gy, gz};
    K::run(table[s], b);
}

}  // namespace hz
