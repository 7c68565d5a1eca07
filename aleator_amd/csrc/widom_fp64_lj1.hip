// widom_fp64_lj1.hip -- Widom direct-sum kernels, FP64 precision, LJ truncation kind 1 (widom.hpp).
#include <type_traits>

#include "widom.hpp"

namespace aleator_gpu {

void widom_direct_fp64_lj1(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                          bool wide, unsigned blocks, int deg, hipStream_t s) {
    widom_detail::launch_direct<true, 1>(a, c, n_cells, wide, blocks, deg, s);
}

}  // namespace aleator_gpu
