// interp.hip -- FrameworkEnergyGrid::interpolate on the device
// (engines/monte_carlo/framework_energy_grid.cc:251-284; the arithmetic is grid_interpolate in
// interp.hpp, bitwise the reference's double).  One thread per position; the table stays in
// HBM/L2 (8 gathers per position).
#include <hip/hip_runtime.h>

#include <cmath>

#include "aleator_gpu.h"
#include "internal.hpp"
#include "interp.hpp"

namespace aleator_gpu {
namespace {

__global__ void interpolate_kernel(const GridGeometry a, const double* __restrict__ table,
                                   const double* __restrict__ pos, uint64_t count, double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    out[i] = grid_interpolate(a, table, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

}  // namespace
}  // namespace aleator_gpu

using namespace aleator_gpu;

extern "C" aleator_status aleator_gpu_grid_interpolate(aleator_ctx* ctx_in, const aleator_lattice* lattice,
                                                       const int32_t n_points[3], const double* table_device,
                                                       const double* positions_device, uint64_t n,
                                                       double* out_device) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Context*>(ctx_in);
        require(ctx && lattice && n_points, ALEATOR_INVALID_ARGUMENT, "null argument");
        if (n == 0) return;
        require(table_device && positions_device && out_device, ALEATOR_INVALID_ARGUMENT, "null argument");
        GridGeometry a{};
        for (int i = 0; i < 3; ++i) {
            require(n_points[i] >= 1, ALEATOR_INVALID_ARGUMENT, "interpolate: point counts must be >= 1");
            a.n[i] = n_points[i];
        }
        const HostLattice h = make_lattice(lattice->m);
        require(std::abs(h.volume) >= 1e-10, ALEATOR_INVALID_ARGUMENT,
                "Lattice::cartesianToFractional: degenerate lattice");
        grid_cofactors(lattice->m, a);
        a.volume = h.volume;
        const uint64_t blocks = (n + 255) / 256;
        require(blocks <= 0x7fffffffULL, ALEATOR_INVALID_ARGUMENT, "interpolate: too many positions");
        hipLaunchKernelGGL(interpolate_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a, table_device,
                           positions_device, n, out_device);
        ALEATOR_HIP_CHECK(hipGetLastError());
    });
}

extern "C" aleator_status aleator_gpu_grid_interpolate_host(aleator_ctx* ctx_in, const aleator_lattice* lattice,
                                                            const int32_t n_points[3], const double* table,
                                                            const double* positions, uint64_t n, double* out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Context*>(ctx_in);
        require(ctx && lattice && n_points && table && (n == 0 || (positions && out)), ALEATOR_INVALID_ARGUMENT,
                "null argument");
        for (int i = 0; i < 3; ++i)
            require(n_points[i] >= 1, ALEATOR_INVALID_ARGUMENT, "interpolate: point counts must be >= 1");
        if (n == 0) return;
        const size_t nodes = (size_t)n_points[0] * n_points[1] * n_points[2];
        double *dt = nullptr, *dp = nullptr, *dout = nullptr;
        auto release = [&] {
            for (double* p : {dt, dp, dout})
                if (p) (void)hipFree(p);
        };
        try {
            ALEATOR_HIP_CHECK(hipMalloc(&dt, nodes * sizeof(double)));
            ALEATOR_HIP_CHECK(hipMalloc(&dp, 3 * n * sizeof(double)));
            ALEATOR_HIP_CHECK(hipMalloc(&dout, n * sizeof(double)));
            ALEATOR_HIP_CHECK(hipMemcpyAsync(dt, table, nodes * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            ALEATOR_HIP_CHECK(hipMemcpyAsync(dp, positions, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            const aleator_status st = aleator_gpu_grid_interpolate(ctx_in, lattice, n_points, dt, dp, n, dout);
            if (st != ALEATOR_OK) fail(st, last_error_slot());
            ALEATOR_HIP_CHECK(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            release();
            throw;
        }
        release();
    });
}
