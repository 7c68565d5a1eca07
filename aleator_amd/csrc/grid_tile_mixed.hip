// grid_tile_mixed.hip -- grid tile kernels for MIXED precision with truncated or shifted LJ (grid_tile.hpp).
#include "grid_tile.hpp"

namespace aleator_gpu {

void launch_tiles_mixed(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                        hipStream_t s) {
    dispatch_deg<false, false>(shape, ng, deg, d, prefix, n_desc, blocks, s);
}

}  // namespace aleator_gpu
