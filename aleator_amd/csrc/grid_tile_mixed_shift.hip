// grid_tile_mixed_shift.hip -- grid tile kernels for MIXED precision with shifted LJ (grid_tile.hpp).
#include "grid_tile.hpp"

namespace aleator_gpu {

void launch_tiles_mixed_shift(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                              hipStream_t s) {
    dispatch_deg<false, 1>(shape, ng, deg, d, prefix, n_desc, blocks, s);
}

}  // namespace aleator_gpu
