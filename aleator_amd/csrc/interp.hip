// interp.hip -- FrameworkEnergyGrid::interpolate on the device
// (engines/monte_carlo/framework_energy_grid.cc:251-284): fractional coordinates by
// Lattice::cartesianToFractional (lattice.cc:167-175: dot with the cofactor rows, divided by
// the volume), wrapped into [0, 1), then trilinear interpolation between the 8 surrounding
// nodes (i/n) with periodic wrap, in the reference's operation order and without contraction,
// so every output is the reference's double.  One thread per position; the table stays in
// HBM/L2 (8 gathers per position).
#include <hip/hip_runtime.h>

#include <cmath>

#include "aleator_gpu.h"
#include "internal.hpp"

namespace aleator_gpu {
namespace {

struct InterpArgs {
    double c12[3], c20[3], c01[3];  // a1 x a2, a2 x a0, a0 x a1
    double volume;
    int n[3];
};

__global__ void interpolate_kernel(const InterpArgs a, const double* __restrict__ table,
                                   const double* __restrict__ pos, uint64_t count, double* __restrict__ out) {
#pragma clang fp contract(off)
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    double f[3] = {(x * a.c12[0] + y * a.c12[1] + z * a.c12[2]) / a.volume,
                   (x * a.c20[0] + y * a.c20[1] + z * a.c20[2]) / a.volume,
                   (x * a.c01[0] + y * a.c01[1] + z * a.c01[2]) / a.volume};
    int i0[3], i1[3];
    double t[3];
    for (int ax = 0; ax < 3; ++ax) {
        f[ax] = f[ax] - floor(f[ax]);
        const int n = a.n[ax];
        const double g = f[ax] * (double)n;
        const int base = (int)floor(g);
        t[ax] = g - (double)base;
        const int wrapped = ((base % n) + n) % n;
        i0[ax] = wrapped;
        i1[ax] = (wrapped + 1) % n;
    }
    auto v = [&](int ix, int iy, int iz) {
        return table[((size_t)iz * (size_t)a.n[1] + (size_t)iy) * (size_t)a.n[0] + (size_t)ix];
    };
    const double c00 = v(i0[0], i0[1], i0[2]) * (1.0 - t[0]) + v(i1[0], i0[1], i0[2]) * t[0];
    const double c10 = v(i0[0], i1[1], i0[2]) * (1.0 - t[0]) + v(i1[0], i1[1], i0[2]) * t[0];
    const double c01 = v(i0[0], i0[1], i1[2]) * (1.0 - t[0]) + v(i1[0], i0[1], i1[2]) * t[0];
    const double c11 = v(i0[0], i1[1], i1[2]) * (1.0 - t[0]) + v(i1[0], i1[1], i1[2]) * t[0];
    const double c0 = c00 * (1.0 - t[1]) + c10 * t[1];
    const double c1 = c01 * (1.0 - t[1]) + c11 * t[1];
    out[i] = c0 * (1.0 - t[2]) + c1 * t[2];
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
        InterpArgs a{};
        for (int i = 0; i < 3; ++i) {
            require(n_points[i] >= 1, ALEATOR_INVALID_ARGUMENT, "interpolate: point counts must be >= 1");
            a.n[i] = n_points[i];
        }
        const HostLattice h = make_lattice(lattice->m);
        require(std::abs(h.volume) >= 1e-10, ALEATOR_INVALID_ARGUMENT,
                "Lattice::cartesianToFractional: degenerate lattice");
        // cofactor rows: inv[r][j] * volume is column j of the inverse, i.e. row r of (a_{j+1} x a_{j+2})
        const double(*m)[3] = lattice->m;
        auto cross = [](const double* p, const double* q, double* o) {
#pragma clang fp contract(off)
            o[0] = p[1] * q[2] - p[2] * q[1];
            o[1] = p[2] * q[0] - p[0] * q[2];
            o[2] = p[0] * q[1] - p[1] * q[0];
        };
        cross(m[1], m[2], a.c12);
        cross(m[2], m[0], a.c20);
        cross(m[0], m[1], a.c01);
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
