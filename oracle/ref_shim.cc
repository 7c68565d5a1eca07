// ref_shim.cc -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" wrapper around the reference's own hot-path classes, compiled
// together with the reference's translation units (oracle/Makefile, target `ref`)
// into oracle/_ref/libaleator_ref.so.  It lets the tests compare the oracle's FP64
// restatement bit-for-bit with the reference, and lets bench.py time the
// reference's CPU path as the baseline.  Nothing in the product links it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <vector>

#include "core/geometry/lattice.hpp"
#include "core/geometry/supercell.hpp"
#include "core/math/counter_based_rng.hpp"
#include "core/math/particle_data.hpp"
#include "engines/monte_carlo/framework_energy_grid.hpp"
#include "engines/monte_carlo/monte_carlo_engine.hpp"
#include "engines/monte_carlo/peng_robinson.hpp"
#include "forcefield/electrostatics/ewald.hpp"
#include "forcefield/electrostatics/ewald_incremental_state.hpp"
#include "forcefield/pairwise/lennard_jones.hpp"
#include "io/cif_document.hpp"
#include "io/cif_error.hpp"
#include "io/element_table.hpp"
#include "io/structure_io.hpp"
#include "io/symmetry_operation.hpp"

using namespace aleator;

namespace {

core::Lattice makeLattice(const double* m) {
    std::array<std::array<double, 3>, 3> a{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = m[3 * i + j];
    return core::Lattice(a);
}

core::ParticleData makeParticles(uint64_t n, const double* x, const double* y, const double* z,
                                 const double* q, const uint32_t* species) {
    core::ParticleData p;
    for (uint64_t i = 0; i < n; ++i)
        p.push_back(x[i], y[i], z[i], 0.0, 0.0, 0.0, 1.0, q ? q[i] : 0.0, species ? species[i] : 0u);
    return p;
}

std::shared_ptr<forcefield::LennardJones> makeLj(uint32_t ns, const double* eps, const double* sig,
                                                 double rc, int trunc, int mixing) {
    std::vector<forcefield::LennardJonesParameters> per(ns);
    for (uint32_t i = 0; i < ns; ++i) per[i] = {eps[i], sig[i]};
    return std::make_shared<forcefield::LennardJones>(
        per, rc, static_cast<forcefield::LennardJonesTruncation>(trunc),
        static_cast<forcefield::MixingRule>(mixing));
}

std::vector<forcefield::ChargedSite> sites(const double* s, int n) {
    std::vector<forcefield::ChargedSite> v(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) v[i] = {s[4 * i], s[4 * i + 1], s[4 * i + 2], s[4 * i + 3]};
    return v;
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_shim: %s\n", e.what());
        return -1;
    }
}

} // namespace

extern "C" {

void ref_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    const auto r = core::philox4x32_10({ctr[0], ctr[1], ctr[2], ctr[3]}, {key[0], key[1]});
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void ref_uniforms(uint64_t seed, uint64_t stream, uint64_t n, double* out) {
    core::Philox4x32Rng rng;
    rng.seed(seed, stream);
    for (uint64_t i = 0; i < n; ++i) out[i] = rng.nextUniform();
}

int ref_minimum_image(const double* m, uint64_t n, const double* a, const double* b, double* out) {
    return guarded([&] {
        const auto lat = makeLattice(m);
        for (uint64_t i = 0; i < n; ++i) {
            const auto d = lat.minimumImageDisplacement({a[3 * i], a[3 * i + 1], a[3 * i + 2]},
                                                        {b[3 * i], b[3 * i + 1], b[3 * i + 2]});
            out[3 * i] = d[0];
            out[3 * i + 1] = d[1];
            out[3 * i + 2] = d[2];
        }
    });
}

int ref_perpendicular_widths(const double* m, double* out) {
    return guarded([&] {
        const auto lat = makeLattice(m);
        for (int a = 0; a < 3; ++a) out[a] = lat.perpendicularWidth(a);
    });
}

int ref_min_supercell(const double* m, double rc, int* n) {
    return guarded([&] {
        const auto r = core::minimumSupercellReplication(makeLattice(m), rc);
        for (int a = 0; a < 3; ++a) n[a] = r[a];
    });
}

int ref_lj_particle_energy(uint32_t ns, const double* eps, const double* sig, double rc, int trunc,
                           int mixing, const double* m, uint64_t n, const double* x,
                           const double* y, const double* z, const uint32_t* species,
                           uint64_t index, double* out) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        const auto p = makeParticles(n, x, y, z, nullptr, species);
        *out = lj->computeParticleEnergy(index, p, makeLattice(m));
    });
}

int ref_lj_tail(uint32_t ns, const double* eps, const double* sig, double rc, int trunc,
                int mixing, const double* m, uint64_t n, const uint32_t* species, double* out) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        std::vector<double> zeros(n, 0.0);
        const auto p = makeParticles(n, zeros.data(), zeros.data(), zeros.data(), nullptr, species);
        *out = lj->tailEnergyCorrection(p, makeLattice(m));
    });
}

int ref_reciprocal_terms(const double* m, double alpha, int kmax, uint64_t cap, double* k,
                         double* w, uint64_t* n_out) {
    return guarded([&] {
        forcefield::Ewald ew(alpha, 1.0, kmax);
        const auto terms = ew.reciprocalTerms(makeLattice(m));
        *n_out = terms.size();
        for (std::size_t t = 0; t < terms.size() && t < cap; ++t) {
            k[3 * t] = terms[t].k[0];
            k[3 * t + 1] = terms[t].k[1];
            k[3 * t + 2] = terms[t].k[2];
            w[t] = terms[t].weight;
        }
    });
}

int ref_ewald_real_particle(double alpha, double rc, const double* m, uint64_t n, const double* x,
                            const double* y, const double* z, const double* q, uint64_t index,
                            double* out) {
    return guarded([&] {
        forcefield::Ewald ew(alpha, rc, 1);
        const auto p = makeParticles(n, x, y, z, q, nullptr);
        *out = ew.realSpaceParticleEnergy(index, p, makeLattice(m));
    });
}

// ---- EwaldIncrementalState (opaque handle keeps Ewald + Lattice alive) ----
struct RefEwaldState {
    forcefield::Ewald ewald;
    core::Lattice lattice;
    std::unique_ptr<forcefield::EwaldIncrementalState> state;
    RefEwaldState(double a, double rc, int k, const core::Lattice& l)
        : ewald(a, rc, k), lattice(l) {}
};

void* ref_ewald_state_create(double alpha, int kmax, const double* m, uint64_t n, const double* x,
                             const double* y, const double* z, const double* q) {
    try {
        auto* s = new RefEwaldState(alpha, 1.0, kmax, makeLattice(m));
        const auto p = makeParticles(n, x, y, z, q, nullptr);
        s->state = std::make_unique<forcefield::EwaldIncrementalState>(s->ewald, p, s->lattice);
        return s;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_shim: %s\n", e.what());
        return nullptr;
    }
}

void ref_ewald_state_destroy(void* h) { delete static_cast<RefEwaldState*>(h); }

double ref_ewald_state_propose(void* h, const double* removed, int nr, const double* added, int na) {
    auto* s = static_cast<RefEwaldState*>(h);
    if (nr == 0) return s->state->proposeInsertion(sites(added, na));
    if (na == 0) return s->state->proposeDeletion(sites(removed, nr));
    return s->state->proposeMove(sites(removed, nr), sites(added, na));
}

void ref_ewald_state_commit(void* h, const double* removed, int nr, const double* added, int na) {
    auto* s = static_cast<RefEwaldState*>(h);
    if (nr == 0)
        s->state->commitInsertion(sites(added, na));
    else if (na == 0)
        s->state->commitDeletion(sites(removed, nr));
    else
        s->state->commitMove(sites(removed, nr), sites(added, na));
}

double ref_ewald_state_committed(void* h) {
    return static_cast<RefEwaldState*>(h)->state->committedEnergy();
}

uint64_t ref_ewald_state_raw(void* h, double* out, uint64_t cap) {
    const auto raw = static_cast<RefEwaldState*>(h)->state->rawStateForTesting();
    for (std::size_t i = 0; i < raw.size() && i < cap; ++i) out[i] = raw[i];
    return raw.size();
}

static std::vector<std::vector<std::size_t>> molecules(const uint64_t* off, const uint64_t* s,
                                                       uint64_t nm) {
    std::vector<std::vector<std::size_t>> mol(nm);
    for (uint64_t m = 0; m < nm; ++m)
        for (uint64_t i = off[m]; i < off[m + 1]; ++i) mol[m].push_back(s[i]);
    return mol;
}

void ref_ewald_state_resync(void* h, uint64_t n, const double* x, const double* y, const double* z,
                            const double* q, const uint64_t* off, const uint64_t* s, uint64_t nm) {
    auto* st = static_cast<RefEwaldState*>(h);
    st->state->resync(makeParticles(n, x, y, z, q, nullptr), molecules(off, s, nm));
}

double ref_ewald_state_recompute(void* h, uint64_t n, const double* x, const double* y,
                                 const double* z, const double* q, const uint64_t* off,
                                 const uint64_t* s, uint64_t nm) {
    auto* st = static_cast<RefEwaldState*>(h);
    return st->state->recomputeFromScratch(makeParticles(n, x, y, z, q, nullptr),
                                           molecules(off, s, nm));
}

// ---- FrameworkEnergyGrid: build with the reference constructor, then read the
// node values back from the reference's own cache file (format v1). ----
int ref_grid_build(uint32_t ns, const double* eps, const double* sig, double rc, int trunc,
                   int mixing, const double* m, uint64_t n, const double* x, const double* y,
                   const double* z, const uint32_t* species, const uint32_t* guests, int n_guest,
                   double spacing, const char* cache_dir, int32_t* counts, double* out,
                   uint64_t cap) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        const auto p = makeParticles(n, x, y, z, nullptr, species);
        const auto lat = makeLattice(m);
        std::vector<uint32_t> g(guests, guests + n_guest);
        engines::FrameworkEnergyGridCacheRequest req{cache_dir, g, spacing};
        bool hit = false;
        const auto grid = engines::loadOrBuildFrameworkEnergyGrid(*lj, p, lat, req, &hit);
        const auto pc = grid.pointCounts();
        for (int a = 0; a < 3; ++a) counts[a] = pc[a];
        const auto key = engines::makeFrameworkEnergyGridCacheKey(*lj, p, lat, g, spacing);
        const auto path = engines::FrameworkEnergyGrid::cacheFilePath(cache_dir, key);
        // Parse: version, key, counts, nTables, (species, len, values)*
        std::ifstream in(path, std::ios::binary);
        uint32_t version = 0;
        in.read(reinterpret_cast<char*>(&version), 4);
        auto skipVec = [&](std::size_t elem) {
            uint64_t k = 0;
            in.read(reinterpret_cast<char*>(&k), 8);
            in.seekg(static_cast<std::streamoff>(k * elem), std::ios::cur);
        };
        skipVec(8);
        skipVec(8);
        skipVec(8);
        skipVec(4);
        in.seekg(9 * 8, std::ios::cur);
        uint64_t nsp = 0;
        in.read(reinterpret_cast<char*>(&nsp), 8);
        in.seekg(static_cast<std::streamoff>(nsp * 16), std::ios::cur);
        in.seekg(8 + 1, std::ios::cur);
        skipVec(4);
        in.seekg(16, std::ios::cur);
        in.seekg(12, std::ios::cur);
        uint64_t ntab = 0;
        in.read(reinterpret_cast<char*>(&ntab), 8);
        const std::size_t total = static_cast<std::size_t>(pc[0]) * pc[1] * pc[2];
        for (uint64_t t = 0; t < ntab; ++t) {
            uint32_t sp = 0;
            uint64_t len = 0;
            in.read(reinterpret_cast<char*>(&sp), 4);
            in.read(reinterpret_cast<char*>(&len), 8);
            std::vector<double> vals(len);
            in.read(reinterpret_cast<char*>(vals.data()), static_cast<std::streamsize>(len * 8));
            int slot = -1;
            for (int gi = 0; gi < n_guest; ++gi)
                if (guests[gi] == sp) slot = gi;
            if (slot < 0) throw std::runtime_error("unexpected table species");
            for (std::size_t i = 0; i < len && (slot * total + i) < cap; ++i)
                out[slot * total + i] = vals[i];
        }
        if (!in) throw std::runtime_error("cache parse failed");
        (void)version;
    });
}

// ---- FrameworkEnergyGrid cache: the reference's key hash, and its loader followed by its
// interpolate at caller positions (a file that loads and interpolates bit-identically holds
// the same key and tables). ----
int ref_grid_cache_hash(uint32_t ns, const double* eps, const double* sig, double rc, int trunc, int mixing,
                        const double* m, uint64_t n, const double* x, const double* y, const double* z,
                        const uint32_t* species, const uint32_t* guests, int n_guest, double spacing,
                        uint64_t* hash) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        const auto p = makeParticles(n, x, y, z, nullptr, species);
        const auto lat = makeLattice(m);
        std::vector<uint32_t> g(guests, guests + n_guest);
        *hash = engines::makeFrameworkEnergyGridCacheKey(*lj, p, lat, g, spacing).hash();
    });
}

int ref_grid_cache_load_interpolate(uint32_t ns, const double* eps, const double* sig, double rc, int trunc,
                                    int mixing, const double* m, uint64_t n, const double* x, const double* y,
                                    const double* z, const uint32_t* species, const uint32_t* guests, int n_guest,
                                    double spacing, const char* cache_dir, uint32_t guest, const double* pos,
                                    uint64_t n_pos, double* out, int32_t* counts, int* hit) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        const auto p = makeParticles(n, x, y, z, nullptr, species);
        const auto lat = makeLattice(m);
        std::vector<uint32_t> g(guests, guests + n_guest);
        const auto key = engines::makeFrameworkEnergyGridCacheKey(*lj, p, lat, g, spacing);
        const auto grid = engines::FrameworkEnergyGrid::loadFromCache(cache_dir, key);
        *hit = grid.has_value() ? 1 : 0;
        if (!grid) return;
        const auto pc = grid->pointCounts();
        for (int a = 0; a < 3; ++a) counts[a] = pc[a];
        for (uint64_t i = 0; i < n_pos; ++i) out[i] = grid->interpolate(guest, {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
    });
}

// Widom through the reference engine: returns K_H.
int ref_widom(uint32_t ns, const double* eps, const double* sig, double rc, int trunc, int mixing,
              const double* m, uint64_t n, const double* x, const double* y, const double* z,
              const double* q, const uint32_t* species, int n_sites, const double* site_xyz,
              const uint32_t* site_species, const double* site_q, int charged, double alpha,
              double rc_real, int kmax, uint64_t seed, uint64_t stream, uint64_t n_trials,
              double temperature, double* kh) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        auto rng = std::make_unique<core::Philox4x32Rng>();
        rng->seed(seed, stream);
        engines::MoleculeSpecies sp;
        for (int s = 0; s < n_sites; ++s)
            sp.sites.push_back({site_xyz[3 * s], site_xyz[3 * s + 1], site_xyz[3 * s + 2],
                                site_species[s], site_q[s], 1.0});
        std::shared_ptr<const forcefield::Ewald> ew;
        if (charged) ew = std::make_shared<forcefield::Ewald>(alpha, rc_real, kmax);
        engines::MonteCarloEngine eng(makeParticles(n, x, y, z, q, species), makeLattice(m), lj,
                                      std::move(rng), temperature, sp, 1.0, 1.5, 0.5, ew);
        *kh = eng.widomInsertionHenryCoefficient(n_trials);
    });
}

// The FrameworkEnergyGrid constructor's inner loop (framework_energy_grid.cc:227-243) with
// the reference's own Lattice/LennardJones calls, at explicit nodes of a (possibly primitive)
// grid lattice `gm` while atoms are periodic in `m`: the CPU baseline of one LJ table.
int ref_grid_nodes_lj(uint32_t ns, const double* eps, const double* sig, double rc, int trunc,
                      int mixing, const double* m, uint64_t n, const double* x, const double* y,
                      const double* z, const uint32_t* species, const double* gm,
                      const int32_t* counts, uint32_t guest, const int64_t* nodes,
                      uint64_t n_nodes, double cap, double* out) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        const auto lat = makeLattice(m);
        const auto glat = makeLattice(gm);
        const auto p = makeParticles(n, x, y, z, nullptr, species);
        for (uint64_t i = 0; i < n_nodes; ++i) {
            const int64_t ix = nodes[i] % counts[0];
            const int64_t iy = (nodes[i] / counts[0]) % counts[1];
            const int64_t iz = nodes[i] / (static_cast<int64_t>(counts[0]) * counts[1]);
            const auto cart = glat.fractionalToCartesian(
                {static_cast<double>(ix) / counts[0], static_cast<double>(iy) / counts[1],
                 static_cast<double>(iz) / counts[2]});
            double energy = 0.0;
            for (std::size_t f = 0; f < p.size(); ++f) {
                const auto params = lj->mixedParameters(guest, p.species[f]);
                const auto d = lat.minimumImageDisplacement(cart, {p.x[f], p.y[f], p.z[f]});
                energy += lj->pairEnergy(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), params);
            }
            out[i] = std::min(energy, cap);
        }
    });
}

// Ewald potential per unit charge at explicit nodes through the reference's own calls:
// real space = Ewald::realSpaceParticleEnergy of a unit charge appended to the framework
// (ewald.cc:223-249); reciprocal = the part of EwaldIncrementalState::proposeInsertion linear
// in the inserted charge, (propose(+1) - propose(-1)) / 2 (ewald_incremental_state.cc:95-136).
int ref_coulomb_nodes(double alpha, double rc_real, int kmax, const double* m, uint64_t n,
                      const double* x, const double* y, const double* z, const double* q,
                      const double* gm, const int32_t* counts, const int64_t* nodes,
                      uint64_t n_nodes, double* out_real, double* out_recip) {
    return guarded([&] {
        const auto lat = makeLattice(m);
        const auto glat = makeLattice(gm);
        forcefield::Ewald ew(alpha, rc_real, kmax);
        auto p = makeParticles(n, x, y, z, q, nullptr);
        forcefield::EwaldIncrementalState state(ew, p, lat);
        for (uint64_t i = 0; i < n_nodes; ++i) {
            const int64_t ix = nodes[i] % counts[0];
            const int64_t iy = (nodes[i] / counts[0]) % counts[1];
            const int64_t iz = nodes[i] / (static_cast<int64_t>(counts[0]) * counts[1]);
            const auto c = glat.fractionalToCartesian(
                {static_cast<double>(ix) / counts[0], static_cast<double>(iy) / counts[1],
                 static_cast<double>(iz) / counts[2]});
            const std::size_t idx = p.push_back(c[0], c[1], c[2], 0, 0, 0, 1.0, 1.0, 0u);
            out_real[i] = ew.realSpaceParticleEnergy(idx, p, lat);
            p.popBack();
            const double plus = state.proposeInsertion({{c[0], c[1], c[2], 1.0}});
            const double minus = state.proposeInsertion({{c[0], c[1], c[2], -1.0}});
            out_recip[i] = 0.5 * (plus - minus);
        }
    });
}

// Runs the reference engine's single-species GCMC for `steps` steps and copies out the
// final particle arrays (framework + guests), molecule count, and committed Ewald energy.
int ref_gcmc_run(uint32_t ns, const double* eps, const double* sig, double rc, int trunc, int mixing,
                 const double* m, uint64_t n, const double* x, const double* y, const double* z,
                 const double* q, const uint32_t* species, int n_sites, const double* site_xyz,
                 const uint32_t* site_species, const double* site_q, double fugacity_pa,
                 double temperature, double max_translation, double max_rotation, int charged,
                 double alpha, double rc_real, int kmax, uint64_t seed, uint64_t stream,
                 uint64_t steps, uint64_t cap, double* ox, double* oy, double* oz, double* oq,
                 uint64_t* n_particles, uint64_t* n_molecules, double* elec_energy) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        auto rng = std::make_unique<core::Philox4x32Rng>();
        rng->seed(seed, stream);
        engines::MoleculeSpecies sp;
        for (int s = 0; s < n_sites; ++s)
            sp.sites.push_back({site_xyz[3 * s], site_xyz[3 * s + 1], site_xyz[3 * s + 2],
                                site_species[s], site_q[s], 1.0});
        std::shared_ptr<const forcefield::Ewald> ew;
        if (charged) ew = std::make_shared<forcefield::Ewald>(alpha, rc_real, kmax);
        engines::MonteCarloEngine eng(makeParticles(n, x, y, z, q, species), makeLattice(m), lj,
                                      std::move(rng), temperature, sp, fugacity_pa,
                                      max_translation, max_rotation, ew);
        eng.run(steps);
        const auto& p = eng.particles();
        *n_particles = p.size();
        *n_molecules = eng.moleculeCount();
        *elec_energy = eng.electrostaticEnergyForTesting();
        for (std::size_t i = 0; i < p.size() && i < cap; ++i) {
            ox[i] = p.x[i];
            oy[i] = p.y[i];
            oz[i] = p.z[i];
            oq[i] = p.charge[i];
        }
    });
}

// ref_gcmc_run with the reference's FrameworkEnergyGrid (loadOrBuildFrameworkEnergyGrid in
// cache_dir, so ref_grid_build on the same directory returns the very tables the engine used).
int ref_gcmc_run_grid(uint32_t ns, const double* eps, const double* sig, double rc, int trunc, int mixing,
                 const double* m, uint64_t n, const double* x, const double* y, const double* z,
                 const double* q, const uint32_t* species, int n_sites, const double* site_xyz,
                 const uint32_t* site_species, const double* site_q, double fugacity_pa,
                 double temperature, double max_translation, double max_rotation, int charged,
                 double alpha, double rc_real, int kmax, uint64_t seed, uint64_t stream,
                 uint64_t steps, uint64_t cap, double* ox, double* oy, double* oz, double* oq,
                 uint64_t* n_particles, uint64_t* n_molecules, double* elec_energy, double spacing,
                      const char* cache_dir, const uint32_t* guests, int n_guest) {
    return guarded([&] {
        const auto lj = makeLj(ns, eps, sig, rc, trunc, mixing);
        auto rng = std::make_unique<core::Philox4x32Rng>();
        rng->seed(seed, stream);
        engines::MoleculeSpecies sp;
        for (int s = 0; s < n_sites; ++s)
            sp.sites.push_back({site_xyz[3 * s], site_xyz[3 * s + 1], site_xyz[3 * s + 2],
                                site_species[s], site_q[s], 1.0});
        std::shared_ptr<const forcefield::Ewald> ew;
        if (charged) ew = std::make_shared<forcefield::Ewald>(alpha, rc_real, kmax);
        std::vector<uint32_t> gids(guests, guests + n_guest);
        engines::FrameworkEnergyGridCacheRequest req{cache_dir, gids, spacing};
        const auto grid = std::make_shared<const engines::FrameworkEnergyGrid>(engines::loadOrBuildFrameworkEnergyGrid(
            *lj, makeParticles(n, x, y, z, q, species), makeLattice(m), req, nullptr));
        engines::MonteCarloEngine eng(makeParticles(n, x, y, z, q, species), makeLattice(m), lj,
                                      std::move(rng), temperature, sp, fugacity_pa,
                                      max_translation, max_rotation, ew, grid);
        eng.run(steps);
        const auto& p = eng.particles();
        *n_particles = p.size();
        *n_molecules = eng.moleculeCount();
        *elec_energy = eng.electrostaticEnergyForTesting();
        for (std::size_t i = 0; i < p.size() && i < cap; ++i) {
            ox[i] = p.x[i];
            oy[i] = p.y[i];
            oz[i] = p.z[i];
            oq[i] = p.charge[i];
        }
    });
}

int ref_pr_fugacity(double tc, double pc, double omega, double t, double p, double* out) {
    return guarded([&] {
        engines::PengRobinson pr({tc, pc, omega});
        *out = pr.fugacityPascal(t, p);
    });
}

// ---- CIF reader (io/) ---------------------------------------------------
// Status convention of the CIF wrappers: 0 ok, 1 io::CifParseError, 2 any other exception.
} // extern "C"

namespace {
template <typename F>
int cif_guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const io::CifParseError&) {
        return 1;
    } catch (const std::exception&) {
        return 2;
    }
}
void put_text(const std::string& s, char* buf, uint64_t cap, uint64_t* len) {
    *len = s.size();
    if (buf && cap > s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
}
} // namespace

extern "C" {

// io::readCif: lattice (9), per-atom arrays (when n <= cap), species symbols and source
// labels each joined with '\n'.
int ref_read_cif(const char* path, double* lattice9, uint64_t* n_atoms, uint64_t cap, double* x, double* y,
                 double* z, double* mass, double* charge, double* occupancy, uint32_t* species, char* symbols,
                 uint64_t symbols_cap, uint64_t* symbols_len, char* labels, uint64_t labels_cap,
                 uint64_t* labels_len) {
    return cif_guarded([&] {
        const io::StructureData s = io::readCif(path);
        const auto& m = s.lattice.matrix();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) lattice9[3 * i + j] = m[i][j];
        const std::size_t n = s.particles.size();
        *n_atoms = n;
        if (n <= cap) {
            for (std::size_t i = 0; i < n; ++i) {
                x[i] = s.particles.x[i];
                y[i] = s.particles.y[i];
                z[i] = s.particles.z[i];
                mass[i] = s.particles.mass[i];
                charge[i] = s.particles.charge[i];
                species[i] = s.particles.species[i];
                occupancy[i] = s.occupancy[i];
            }
        }
        std::string sym, lab;
        for (std::size_t i = 0; i < s.speciesSymbols.size(); ++i) sym += (i ? "\n" : "") + s.speciesSymbols[i];
        for (std::size_t i = 0; i < s.sourceLabel.size(); ++i) lab += (i ? "\n" : "") + s.sourceLabel[i];
        put_text(sym, symbols, symbols_cap, symbols_len);
        put_text(lab, labels, labels_cap, labels_len);
    });
}

// io::parseCifDocument, flattened to text: "block\x1f<name>", then one "tag\x1f<t>\x1f<v>" per
// single tag in sorted order, then per loop "loop\x1f<ncols>\x1f<nrows>", its columns and its
// values, every record ended by '\x1e'.
int ref_cif_document(const char* text, uint64_t length, char* out, uint64_t cap, uint64_t* len) {
    return cif_guarded([&] {
        const io::CifDocument d = io::parseCifDocument(std::string(text, length));
        std::string r = "block\x1f" + d.blockName + "\x1e";
        std::vector<std::pair<std::string, std::string>> tags(d.tags.begin(), d.tags.end());
        std::sort(tags.begin(), tags.end());
        for (const auto& [t, v] : tags) r += "tag\x1f" + t + "\x1f" + v + "\x1e";
        for (const auto& l : d.loops) {
            r += "loop\x1f" + std::to_string(l.columns.size()) + "\x1f" + std::to_string(l.rows.size()) + "\x1e";
            for (const auto& c : l.columns) r += c + "\x1e";
            for (const auto& row : l.rows)
                for (const auto& v : row) r += v + "\x1e";
        }
        put_text(r, out, cap, len);
    });
}

int ref_cif_try_number(const char* token, double* value, int* has_value) {
    return cif_guarded([&] {
        const auto v = io::tryParseCifNumber(token);
        *has_value = v.has_value();
        *value = v.value_or(0.0);
    });
}

int ref_symmetry_operation(const char* text, double* matrix9, double* translation) {
    return cif_guarded([&] {
        const io::SymmetryOperation op = io::parseSymmetryOperation(text);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) matrix9[3 * i + j] = op.matrix[i][j];
            translation[i] = op.translation[i];
        }
    });
}

int ref_element_from_label(const char* label, char* symbol4, int* found) {
    return cif_guarded([&] {
        const auto e = io::elementFromLabel(label);
        *found = e.has_value();
        std::memset(symbol4, 0, 4);
        if (e) std::memcpy(symbol4, e->c_str(), std::min<std::size_t>(e->size(), 3));
    });
}

} // extern "C"
