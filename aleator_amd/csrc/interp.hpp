// interp.hpp -- FrameworkEnergyGrid::interpolate (engines/monte_carlo/framework_energy_grid.cc:251-284)
// as a device function, shared by the interpolation entry point and the GCMC chains in grid mode.
//
// Fractional coordinates by Lattice::cartesianToFractional (lattice.cc:167-175: dot with the
// cofactor rows, divided by the signed volume), wrapped into [0, 1), then trilinear interpolation
// between the 8 surrounding nodes (i/n) with periodic wrap, in the reference's operation order and
// without contraction, so every output is the reference's double.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace aleator_gpu {

struct GridGeometry {
    double c12[3], c20[3], c01[3];  // a1 x a2, a2 x a0, a0 x a1
    double volume;                  // signed, a0 . (a1 x a2)
    int n[3];                       // nodes per axis, table index (iz * n1 + iy) * n0 + ix
};

// Cofactor rows of the lattice rows m (the same products Lattice::cartesianToFractional uses).
inline void grid_cofactors(const double m[3][3], GridGeometry& g) {
#pragma clang fp contract(off)
    auto cross = [](const double* p, const double* q, double* o) {
        o[0] = p[1] * q[2] - p[2] * q[1];
        o[1] = p[2] * q[0] - p[0] * q[2];
        o[2] = p[0] * q[1] - p[1] * q[0];
    };
    cross(m[1], m[2], g.c12);
    cross(m[2], m[0], g.c20);
    cross(m[0], m[1], g.c01);
}

__device__ inline double grid_interpolate(const GridGeometry& a, const double* __restrict__ table, double x, double y,
                                          double z) {
#pragma clang fp contract(off)
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
    return c0 * (1.0 - t[2]) + c1 * t[2];
}

}  // namespace aleator_gpu
