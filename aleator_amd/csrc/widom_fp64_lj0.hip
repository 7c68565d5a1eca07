// widom_fp64_lj0.hip -- Widom direct-sum kernels, FP64 precision, LJ truncation kind 0 (widom.hpp).
#include <type_traits>

#include "widom.hpp"

namespace aleator_gpu {

void widom_direct_fp64_lj0(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                          bool wide, unsigned blocks, int deg, hipStream_t s) {
    widom_detail::launch_direct<true, 0>(a, c, n_cells, wide, blocks, deg, s);
}

}  // namespace aleator_gpu
