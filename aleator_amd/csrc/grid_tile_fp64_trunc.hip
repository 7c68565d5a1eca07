// grid_tile_fp64_trunc.hip -- grid tile kernels for FP64 precision with truncated LJ (grid_tile.hpp).
#include "grid_tile.hpp"

namespace aleator_gpu {

void launch_tiles_fp64_trunc(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                             hipStream_t s) {
    dispatch_deg<true, 0>(shape, ng, deg, d, prefix, n_desc, blocks, s);
}

}  // namespace aleator_gpu
