// gcmc.hpp -- device layout of the batched GCMC chains (see gcmc.hip).
#pragma once

#include <cstdint>

#include "internal.hpp"

namespace aleator_gpu {

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
};

void launch_gcmc(const GcmcArgs& a, int n_chains, hipStream_t s);

}  // namespace aleator_gpu
