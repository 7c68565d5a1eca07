// grid_tile_fp64_lfs.hip -- grid tile kernels for FP64 precision with linear-force-shifted LJ (grid_tile.hpp).
#include "grid_tile.hpp"

namespace aleator_gpu {

void launch_tiles_fp64_lfs(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                           hipStream_t s) {
    dispatch_deg<true, 2>(shape, ng, deg, d, prefix, n_desc, blocks, s);
}

}  // namespace aleator_gpu
