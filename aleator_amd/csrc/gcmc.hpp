// gcmc.hpp -- device layout of the batched GCMC chains (see gcmc.hip).
#pragma once

#include <cstdint>

#include "internal.hpp"
#include "interp.hpp"

namespace aleator_gpu {

constexpr int kMaxGuestCand = 4096;  // guest sites (max molecules x sites) the guest cell filter holds
constexpr int kGuestCellMinSites = 1024;  // default: guest sites from which a step uses the cells

struct ChainParams {
    double fugacity_internal;  // K/A^3 (fugacity in Pa times 1/(k_B 1e30), monte_carlo_engine.hpp:184)
    double temperature;
    uint64_t seed, stream;
};

struct ChainState {
    uint64_t call;             // next Philox call index
    int32_t n_mol;
    int32_t moves_since_resync;
    double self_energy, exclusion_energy;
    double sum_n;              // sum over steps of the molecule count after the step
    uint64_t steps;
    uint64_t attempted[4], accepted[4];  // translate, rotate, insert, delete
    int32_t error;             // 1: molecule capacity exceeded
    int32_t pad;
    // consistency check at every resync: largest |fresh - committed| / max(1, |fresh|) of the
    // Ewald energy (reciprocal + self + exclusion), and the number of resyncs it covers
    double max_drift;
    uint64_t resyncs;
};

struct GcmcArgs {
    // framework
    const double* fx;
    const double* fy;
    const double* fz;
    const double* q;
    const int32_t* type;
    int natoms;
    double m[9], inv[9], volume;
    // Lennard-Jones
    double rc_lj, rc2_lj;
    int trunc;
    const double* ljtab;  // [4][kMaxSites][kMaxTypes]: site vs framework type
    const double* ljgg;   // [4][kMaxSites][kMaxSites]: site vs guest site
    // Ewald
    int charged;
    double alpha, rc2_real, kc, recip_scale, self_pref, fw_self;
    // real-space pair kernel erfc(alpha r)/r = 1/r - erf(alpha r)/r, erf(alpha r)/r as a polynomial
    // of degree erf_deg (18, 24 or 28) in x = r^2 erf_xs - 1 over r in [0, rc_real]
    // (erf_kernel_poly, cut at 1e-17)
    double erf_poly[kErfPolyMax + 1], erf_xs;
    int erf_deg;
    const int4* kn;
    const double* kw;
    const double* fw_sre;
    const double* fw_sim;
    int nk_half, kmax;
    // species
    int ns;
    double local[kMaxSites][3];
    double charge[kMaxSites];
    double max_translation, max_rotation;
    // chains
    int max_mol;
    const ChainParams* params;
    ChainState* state;
    double* mol_pos;  // [chain][max_mol][ns][3]
    double* sre;      // [chain][nk_half]
    double* sim;
    uint64_t n_steps;
    int32_t* n_trace;  // optional [chain][n_steps]
    // insertion density (null: uniform insertion); bins over the cell divided by prop_reps
    const double* prop_cdf;  // [prop_nbins] cumulative bin probabilities, last = 1
    const double* prop_fac;  // [prop_nbins] n_bins * (cdf[b] - cdf[b-1]) = V p(r) in bin b
    int prop_nb[3], prop_reps[3], prop_nbins;
    // framework atoms within rc + half the longest bin diagonal of each bin centre, for bins
    // fb[0] x fb[1] x fb[2] of the cell (CSR; null: every site scans every atom)
    const int32_t* bin_start;  // [n_bins + 1]
    const double4* bin_pos;    // list entries: fractional x, y, z and charge of the atom
    const int32_t* bin_type;   // list entries: compact LJ type of the atom
    int fb[3];
    // guest cells per lattice axis for the guest-guest candidate filter (gc[0] == 0: every
    // guest site is a candidate); each cell is at least the guest cutoff wide
    int gc[3];
    int gc_min_sites;  // guest sites from which a step uses the cells
    // energy grid (null: framework LJ summed over the atoms).  With a grid, a site's framework
    // LJ energy is table grid_table[s] interpolated at its Cartesian position, the framework scan
    // keeps only real-space Coulomb, and guest-guest LJ is unchanged (moleculeDispersionEnergy
    // with a FrameworkEnergyGrid, monte_carlo_engine.cc:341-374)
    const double* grid;  // [tables][n2][n1][n0]
    size_t grid_cells;
    GridGeometry grid_geo;
    int grid_table[kMaxSites];
};

// Builds the framework bin lists of `a` (a.fb set) into device buffers owned by the caller:
// start [n_bins + 1], then per entry the atom's (fx, fy, fz, q) and type, in bin order.
void build_framework_bins(const GcmcArgs& a, double reach, int32_t** start, double4** pos, int32_t** type,
                          hipStream_t s);

// Launches the chains with `a`, which must already be in the device copy d_args (the kernel reads
// its arguments from there).
void launch_gcmc(const GcmcArgs& a, const GcmcArgs* d_args, int n_chains, hipStream_t s);

}  // namespace aleator_gpu
