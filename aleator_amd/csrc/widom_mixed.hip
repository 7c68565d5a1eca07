// widom_mixed.hip -- Widom direct-sum kernels for MIXED precision (widom.hpp).
#include <type_traits>

#include "widom.hpp"

namespace aleator_gpu {

void widom_direct_mixed(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                        bool wide, unsigned blocks, int deg, bool lfs, hipStream_t s) {
    if (lfs) widom_detail::launch_direct<false, true>(a, c, n_cells, wide, blocks, deg, s);
    else widom_detail::launch_direct<false, false>(a, c, n_cells, wide, blocks, deg, s);
}

}  // namespace aleator_gpu
