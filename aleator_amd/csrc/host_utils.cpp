// host_utils.cpp -- host-side pieces of the path that carry no device work: supercell
// replication, the Peng-Robinson fugacity of a chain's reservoir and the analytic LJ tail
// correction of one insertion.  Restated in the reference's arithmetic order.
#include <cmath>
#include <cstdint>
#include <vector>

#include "aleator_gpu.h"
#include "device.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace {

double perpendicular_width(const double m[3][3], int axis) {
    const double* o0 = m[(axis + 1) % 3];
    const double* o1 = m[(axis + 2) % 3];
    const double c[3] = {o0[1] * o1[2] - o0[2] * o1[1], o0[2] * o1[0] - o0[0] * o1[2],
                         o0[0] * o1[1] - o0[1] * o1[0]};
    const double face = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    // Lattice::volume (lattice.hpp:35-41)
    const double cx = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double cy = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double cz = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double vol = m[0][0] * cx + m[0][1] * cy + m[0][2] * cz;
    return face > 0.0 ? std::abs(vol) / face : 0.0;
}

double volume(const double m[3][3]) {
    const double cx = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double cy = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double cz = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    return m[0][0] * cx + m[0][1] * cy + m[0][2] * cz;
}

// tailEnergyCorrection (lennard_jones.cc:218-257) from per-species counts
double tail_energy(const aleator_lj_params* lj, const std::vector<double>& counts, double vol) {
    const double rc = lj->cutoff;
    double total = 0.0;
    for (uint32_t a = 0; a < lj->n_species; ++a) {
        if (counts[a] == 0.0) continue;
        for (uint32_t b = 0; b < lj->n_species; ++b) {
            if (counts[b] == 0.0) continue;
            double eps, sig;
            if (a == b) {
                eps = lj->epsilon[a];
                sig = lj->sigma[a];
            } else {
                eps = std::sqrt(lj->epsilon[a] * lj->epsilon[b]);
                sig = lj->mixing == ALEATOR_MIX_LORENTZ_BERTHELOT ? 0.5 * (lj->sigma[a] + lj->sigma[b])
                                                                 : std::sqrt(lj->sigma[a] * lj->sigma[b]);
            }
            const double sr3 = sig * sig * sig;
            const double sr9 = sr3 * sr3 * sr3 / (rc * rc * rc * rc * rc * rc * rc * rc * rc);
            const double sr3_over_rc3 = sr3 / (rc * rc * rc);
            total += counts[a] * counts[b] * (eps * sr3 * ((4.0 / 9.0) * sr9 - (4.0 / 3.0) * sr3_over_rc3));
        }
    }
    return (2.0 * 3.14159265358979323846 / std::abs(vol)) * total;
}

}  // namespace

namespace aleator_gpu {

double lj_tail_delta(const Framework& fw, const uint32_t* site_species, int n_sites) {
    aleator_lj_params lj{};
    lj.epsilon = fw.eps.data();
    lj.sigma = fw.sigma.data();
    lj.n_species = (uint32_t)fw.eps.size();
    lj.cutoff = fw.rc;
    lj.truncation = fw.trunc;
    lj.mixing = fw.mixing;
    std::vector<double> before(lj.n_species, 0.0);
    for (uint32_t sp : fw.species) before[sp] += 1.0;
    std::vector<double> after = before;
    for (int s = 0; s < n_sites; ++s) {
        if (site_species[s] >= lj.n_species) fail(ALEATOR_OUT_OF_RANGE, "tail correction: species out of range");
        after[site_species[s]] += 1.0;
    }
    const double vol = fw.lat.volume;
    return tail_energy(&lj, after, vol) - tail_energy(&lj, before, vol);
}

}  // namespace aleator_gpu

extern "C" {

aleator_status aleator_philox_uniforms(uint64_t seed, uint64_t stream, uint64_t first_call_index, uint64_t n,
                                      double* out) {
    if (n && !out) return ALEATOR_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < n; ++i) out[i] = aleator_gpu::philox_uniform(seed, stream, first_call_index + i);
    return ALEATOR_OK;
}

aleator_status aleator_supercell_replication(const aleator_lattice* lattice, double cutoff, int32_t n[3]) {
    if (!lattice || !n || !(cutoff > 0.0)) return ALEATOR_INVALID_ARGUMENT;
    for (int axis = 0; axis < 3; ++axis) {
        const double width = perpendicular_width(lattice->m, axis);
        if (!(width > 0.0)) return ALEATOR_INVALID_ARGUMENT;
        n[axis] = 1;
        while ((double)n[axis] * width <= 2.0 * cutoff) ++n[axis];
    }
    return ALEATOR_OK;
}

aleator_status aleator_replicate_supercell(const aleator_particles* in, const aleator_lattice* lattice, int32_t nx,
                                           int32_t ny, int32_t nz, double* x, double* y, double* z, double* charge,
                                           uint32_t* species, aleator_lattice* out_lattice) {
    if (!in || !lattice || nx < 1 || ny < 1 || nz < 1 || !x || !y || !z || !charge || !species || !out_lattice)
        return ALEATOR_INVALID_ARGUMENT;
    const double(*m)[3] = lattice->m;
    uint64_t w = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const double ox = i * m[0][0] + j * m[1][0] + k * m[2][0];
                const double oy = i * m[0][1] + j * m[1][1] + k * m[2][1];
                const double oz = i * m[0][2] + j * m[1][2] + k * m[2][2];
                for (uint64_t p = 0; p < in->n; ++p, ++w) {
                    x[w] = in->x[p] + ox;
                    y[w] = in->y[p] + oy;
                    z[w] = in->z[p] + oz;
                    charge[w] = in->charge ? in->charge[p] : 0.0;
                    species[w] = in->species[p];
                }
            }
    for (int d = 0; d < 3; ++d) {
        out_lattice->m[0][d] = nx * m[0][d];
        out_lattice->m[1][d] = ny * m[1][d];
        out_lattice->m[2][d] = nz * m[2][d];
    }
    return ALEATOR_OK;
}

aleator_status aleator_peng_robinson_fugacity(double tc, double pc, double omega, double t, double p,
                                              double* fugacity_pa) {
    if (!fugacity_pa || !(tc > 0.0) || !(pc > 0.0) || !(t > 0.0) || !(p > 0.0)) return ALEATOR_INVALID_ARGUMENT;
    const double r = 8.31446261815324;  // PengRobinson::kGasConstant
    const double a = 0.45724 * r * r * tc * tc / pc;
    const double b = 0.07780 * r * tc / pc;
    const double kappa = 0.37464 + 1.54226 * omega - 0.26992 * omega * omega;
    const double term = 1.0 + kappa * (1.0 - std::sqrt(t / tc));
    const double a_alpha = a * (term * term);
    const double A = a_alpha * p / (r * r * t * t);
    const double B = b * p / (r * t);
    const double c2 = -(1.0 - B);
    const double c1 = A - 3.0 * B * B - 2.0 * B;
    const double c0 = -(A * B - B * B - B * B * B);
    double z = 1.0;
    bool converged = false;
    for (int iter = 0; iter < 100; ++iter) {
        const double fz = z * z * z + c2 * z * z + c1 * z + c0;
        if (std::abs(fz) < 1e-14) {
            converged = true;
            break;
        }
        const double fpz = 3.0 * z * z + 2.0 * c2 * z + c1;
        if (fpz == 0.0) break;
        z -= fz / fpz;
        if (!(z > B)) z = B + 0.5 * (1.0 - B > 0 ? 1.0 - B : 1.0);
    }
    if (!converged && std::abs(z * z * z + c2 * z * z + c1 * z + c0) > 1e-8) return ALEATOR_RUNTIME_ERROR;
    const double sqrt2 = 1.4142135623730951;
    const double ln_phi = (z - 1.0) - std::log(z - B) -
                          (A / (2.0 * sqrt2 * B)) * std::log((z + (1.0 + sqrt2) * B) / (z + (1.0 - sqrt2) * B));
    *fugacity_pa = std::exp(ln_phi) * p;
    return ALEATOR_OK;
}

aleator_status aleator_lj_tail_insertion_delta(const aleator_lj_params* lj, const aleator_lattice* lattice,
                                               const uint32_t* fw_species, uint64_t n_fw,
                                               const aleator_molecule_site* sites, int32_t n_sites, double* delta) {
    if (!lj || !lattice || !delta || (n_fw && !fw_species) || n_sites < 0 || (n_sites && !sites) ||
        !(lj->cutoff > 0.0))
        return ALEATOR_INVALID_ARGUMENT;
    std::vector<double> before(lj->n_species, 0.0);
    for (uint64_t i = 0; i < n_fw; ++i) {
        if (fw_species[i] >= lj->n_species) return ALEATOR_OUT_OF_RANGE;
        before[fw_species[i]] += 1.0;
    }
    std::vector<double> after = before;
    for (int s = 0; s < n_sites; ++s) {
        if (sites[s].lj_species >= lj->n_species) return ALEATOR_OUT_OF_RANGE;
        after[sites[s].lj_species] += 1.0;
    }
    const double vol = volume(lattice->m);
    *delta = tail_energy(lj, after, vol) - tail_energy(lj, before, vol);
    return ALEATOR_OK;
}

aleator_status aleator_erf_kernel_poly(double alpha, double rmax, double tol, int32_t min_degree, double* coef,
                                       double* xs, int32_t* degree) {
    if (!coef || !xs || !degree || !(alpha > 0.0) || !(rmax > 0.0) || !(tol > 0.0))
        return ALEATOR_INVALID_ARGUMENT;
    for (int k = 0; k <= aleator_gpu::kErfPolyMax; ++k) coef[k] = 0.0;
    *degree = aleator_gpu::erf_kernel_poly(alpha, rmax, tol, min_degree, coef, xs);
    return ALEATOR_OK;
}

}  // extern "C"
