// widom_mixed_lj2.hip -- Widom direct-sum kernels, MIXED precision, LJ truncation kind 2 (widom.hpp).
#include <type_traits>

#include "widom.hpp"

namespace aleator_gpu {

void widom_direct_mixed_lj2(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                          bool wide, unsigned blocks, int deg, hipStream_t s) {
    widom_detail::launch_direct<false, 2>(a, c, n_cells, wide, blocks, deg, s);
}

}  // namespace aleator_gpu
