// synth.cpp -- synthetic frameworks for the benchmark workloads (no reference
// counterpart).  Placement draws three uniforms per candidate from
// Philox4x32Rng(seed, stream) (the reference generator, counter_based_rng.cc:56-79)
// as fractional coordinates and rejects a candidate closer than `min_dist` to an
// earlier atom; the separation test folds the fractional difference to the nearest
// image, which is exact below half the smallest perpendicular width.
#include <cmath>
#include <cstdint>

#include "aleator_gpu.h"
#include "device.hpp"

#pragma clang fp contract(off)

namespace {
struct HostRng {
    uint64_t seed, stream, call = 0;
    double next() { return aleator_gpu::philox_uniform(seed, stream, call++); }
};
}  // namespace

extern "C" aleator_status aleator_synth_framework(const aleator_lattice* lattice, uint64_t n_atoms,
                                                  uint64_t seed, uint64_t stream, double min_dist,
                                                  const double* proportions, uint32_t n_types,
                                                  double charge_sigma, uint64_t charge_stream, double* x,
                                                  double* y, double* z, double* charge, uint32_t* species) {
    if (!lattice || (n_atoms && (!x || !y || !z || !charge || !species)) || n_types == 0 || !proportions)
        return ALEATOR_INVALID_ARGUMENT;
    const double(*m)[3] = lattice->m;
    double* f = new double[3 * (n_atoms ? n_atoms : 1)];
    HostRng rng{seed, stream};
    const double md2 = min_dist * min_dist;
    uint64_t placed = 0, attempts = 0;
    while (placed < n_atoms) {
        if (++attempts > 1000ull * n_atoms + 1000ull) {
            delete[] f;
            return ALEATOR_RUNTIME_ERROR;
        }
        const double f0 = rng.next();
        const double f1 = rng.next();
        const double f2 = rng.next();
        bool ok = true;
        for (uint64_t j = 0; j < placed && ok; ++j) {
            double d0 = f[3 * j] - f0, d1 = f[3 * j + 1] - f1, d2 = f[3 * j + 2] - f2;
            d0 -= std::rint(d0);
            d1 -= std::rint(d1);
            d2 -= std::rint(d2);
            const double cx = d0 * m[0][0] + d1 * m[1][0] + d2 * m[2][0];
            const double cy = d0 * m[0][1] + d1 * m[1][1] + d2 * m[2][1];
            const double cz = d0 * m[0][2] + d1 * m[1][2] + d2 * m[2][2];
            if (cx * cx + cy * cy + cz * cz < md2) ok = false;
        }
        if (!ok) continue;
        f[3 * placed] = f0;
        f[3 * placed + 1] = f1;
        f[3 * placed + 2] = f2;
        x[placed] = f0 * m[0][0] + f1 * m[1][0] + f2 * m[2][0];
        y[placed] = f0 * m[0][1] + f1 * m[1][1] + f2 * m[2][1];
        z[placed] = f0 * m[0][2] + f1 * m[1][2] + f2 * m[2][2];
        ++placed;
    }
    delete[] f;
    uint64_t w = 0;
    for (uint32_t t = 0; t < n_types; ++t) {
        uint64_t cnt;
        if (t + 1 < n_types) {
            cnt = (uint64_t)std::llround(proportions[t] * (double)n_atoms);
            if (w + cnt > n_atoms) cnt = n_atoms - w;
        } else {
            cnt = n_atoms - w;
        }
        for (uint64_t i = 0; i < cnt; ++i) species[w++] = t;
    }
    if (charge_sigma > 0.0) {
        HostRng crng{seed, charge_stream};
        double mean = 0.0;
        for (uint64_t i = 0; i < n_atoms; ++i) {
            const double u1 = crng.next();
            const double u2 = crng.next();
            const double g = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
            charge[i] = charge_sigma * g;
            mean += charge[i];
        }
        mean /= (double)n_atoms;
        for (uint64_t i = 0; i < n_atoms; ++i) charge[i] -= mean;
    } else {
        for (uint64_t i = 0; i < n_atoms; ++i) charge[i] = 0.0;
    }
    return ALEATOR_OK;
}
