// aleator_gpu.hpp -- C++23 host layer over the C ABI of include/aleator_gpu.h.
//
// Mirrors the reference's interfaces on the adsorption hot path so that a caller written
// against them reads the same on the device:
//   engines::FrameworkEnergyGrid and its cache (engines/monte_carlo/framework_energy_grid.hpp:23-220),
//   forcefield::EwaldIncrementalState (forcefield/electrostatics/ewald_incremental_state.hpp:60-140),
//   MonteCarloEngine::widomInsertionHenryCoefficient (engines/monte_carlo/monte_carlo_engine.hpp:198),
//   MonteCarloEngine's GCMC run, batched over chains, with or without a FrameworkEnergyGrid
//   (GcmcChains; engines/monte_carlo/monte_carlo_engine.hpp:129-151),
// with the reference's exception types for its error cases (std::invalid_argument,
// std::out_of_range, std::runtime_error; aleator::gpu::NotImplemented stands for
// aleator::NotImplemented, core/exceptions.hpp:13-17).
//
// The templates take the reference's own value types, or anything with the same members:
//   lattice:    .matrix() -> std::array<std::array<double, 3>, 3> (core::Lattice)
//   particles:  .x .y .z .charge (double ranges, charge may be empty) .species (u32 range)
//               (core::ParticleData)
//   LJ:         .perSpecies() -> range of {.epsilon, .sigma}; .cutoff(); .truncation();
//               .mixingRule() (forcefield::LennardJones; enums in the reference's order)
//   Ewald:      .alpha() .realSpaceCutoff() .kMax() (forcefield::Ewald)
//   sites:      {.x .y .z .charge} (ChargedSite); {.x .y .z .ljSpecies .charge} (MoleculeSite)
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <filesystem>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "aleator_gpu.h"

namespace aleator::gpu {

/// Stand-in for aleator::NotImplemented when built outside the reference.
struct NotImplemented : std::logic_error {
    using std::logic_error::logic_error;
};
/// A HIP runtime failure (no reference counterpart: the reference has no device).
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
/// Stand-in for aleator::io::CifParseError (io/cif_error.hpp:8-16).
struct CifParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

/// Maps a status of the C ABI onto the reference's exception types.
inline void check(aleator_status s) {
    if (s == ALEATOR_OK) return;
    const char* m = aleator_gpu_last_error();
    const std::string msg = m ? m : "aleator_gpu error";
    switch (s) {
        case ALEATOR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case ALEATOR_OUT_OF_RANGE: throw std::out_of_range(msg);
        case ALEATOR_NOT_IMPLEMENTED: throw NotImplemented(msg);
        case ALEATOR_DEVICE_ERROR: throw DeviceError(msg);
        case ALEATOR_PARSE_ERROR: throw CifParseError(msg);
        default: throw std::runtime_error(msg);
    }
}

/// One GPU and its stream (no reference counterpart).
class Device {
public:
    explicit Device(int device = 0) { check(aleator_gpu_init(device, &ctx_)); }
    ~Device() {
        if (ctx_) aleator_gpu_destroy(ctx_);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    [[nodiscard]] aleator_ctx* handle() const noexcept { return ctx_; }

private:
    aleator_ctx* ctx_ = nullptr;
};

namespace detail {

template <class Lat>
aleator_lattice lattice(const Lat& l) {
    aleator_lattice out{};
    const auto& m = l.matrix();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out.m[i][j] = m[i][j];
    return out;
}

// owning copies of particle arrays behind an aleator_particles view
struct Particles {
    std::vector<double> x, y, z, q;
    std::vector<std::uint32_t> species;
    aleator_particles view{};
    template <class P>
    explicit Particles(const P& p)
        : x(p.x.begin(), p.x.end()), y(p.y.begin(), p.y.end()), z(p.z.begin(), p.z.end()),
          q(p.charge.begin(), p.charge.end()), species(p.species.begin(), p.species.end()) {
        view = {x.data(), y.data(), z.data(), q.empty() ? nullptr : q.data(), species.data(), x.size()};
    }
    Particles(const Particles&) = delete;
};

struct LjParams {
    std::vector<double> eps, sig;
    aleator_lj_params view{};
    template <class LJ>
    explicit LjParams(const LJ& lj) {
        for (const auto& s : lj.perSpecies()) {
            eps.push_back(s.epsilon);
            sig.push_back(s.sigma);
        }
        view = {eps.data(), sig.data(), (uint32_t)eps.size(), lj.cutoff(), (int32_t)lj.truncation(),
                (int32_t)lj.mixingRule()};
    }
    LjParams(const LjParams&) = delete;
};

template <class Site>
std::vector<aleator_charged_site> charged(const std::vector<Site>& s) {
    std::vector<aleator_charged_site> out;
    out.reserve(s.size());
    for (const auto& v : s) out.push_back({v.x, v.y, v.z, v.charge});
    return out;
}

}  // namespace detail

/// A rigid framework resident on the device: atoms, mixed LJ constants and, with Ewald
/// parameters, the reference k-set and its structure factor.
class Framework {
public:
    template <class LJ, class P, class Lat>
    Framework(Device& dev, const LJ& forceField, const P& particles, const Lat& lattice) : dev_(&dev) {
        create(forceField, particles, lattice, nullptr);
    }
    template <class LJ, class E, class P, class Lat>
    Framework(Device& dev, const LJ& forceField, const E& ewald, const P& particles, const Lat& lattice) : dev_(&dev) {
        const aleator_ewald_params e{ewald.alpha(), ewald.realSpaceCutoff(), ewald.kMax()};
        create(forceField, particles, lattice, &e);
    }
    ~Framework() {
        if (fw_) aleator_gpu_framework_destroy(fw_);
    }
    Framework(const Framework&) = delete;
    Framework& operator=(const Framework&) = delete;
    [[nodiscard]] aleator_framework* handle() const noexcept { return fw_; }
    [[nodiscard]] Device& device() const noexcept { return *dev_; }
    /// NaN / Inf sentinel: non-finite table values written by this framework's grid builds.
    [[nodiscard]] std::uint64_t nonfiniteCount() const {
        std::uint64_t n = 0;
        check(aleator_gpu_framework_nonfinite_count(fw_, &n));
        return n;
    }

private:
    template <class LJ, class P, class Lat>
    void create(const LJ& forceField, const P& particles, const Lat& lattice, const aleator_ewald_params* e) {
        const detail::Particles p(particles);
        const detail::LjParams lj(forceField);
        const aleator_lattice lat = detail::lattice(lattice);
        check(aleator_gpu_framework_create(dev_->handle(), &lat, &p.view, &lj.view, e, &fw_));
    }
    Device* dev_;
    aleator_framework* fw_ = nullptr;
};

/// engines::FrameworkEnergyGridCacheKey (framework_energy_grid.hpp:23-60), as
/// makeFrameworkEnergyGridCacheKey builds it.
class FrameworkEnergyGridCacheKey {
public:
    template <class LJ, class P, class Lat>
    FrameworkEnergyGridCacheKey(const LJ& forceField, const P& frameworkParticles, const Lat& lattice,
                                std::vector<std::uint32_t> guestSpeciesIds, double spacingAngstrom)
        : p_(frameworkParticles), lj_(forceField), lat_(detail::lattice(lattice)), guests_(std::move(guestSpeciesIds)) {
        key_ = {&lat_, &p_.view, &lj_.view, guests_.empty() ? nullptr : guests_.data(), (int32_t)guests_.size(),
                spacingAngstrom};
    }
    FrameworkEnergyGridCacheKey(const FrameworkEnergyGridCacheKey&) = delete;
    [[nodiscard]] std::uint64_t hash() const {
        std::uint64_t h = 0;
        check(aleator_grid_cache_key_hash(&key_, &h));
        return h;
    }
    [[nodiscard]] const aleator_grid_cache_key* c_key() const noexcept { return &key_; }
    [[nodiscard]] const aleator_lattice& lattice() const noexcept { return lat_; }
    [[nodiscard]] double spacingAngstrom() const noexcept { return key_.spacing; }

private:
    detail::Particles p_;
    detail::LjParams lj_;
    aleator_lattice lat_;
    std::vector<std::uint32_t> guests_;
    aleator_grid_cache_key key_{};
};

/// engines::FrameworkEnergyGrid (framework_energy_grid.hpp:94-220): tables built on the
/// device, kept on the host (and uploaded once for interpolation).
class FrameworkEnergyGrid {
public:
    static constexpr double kEnergyCapKelvin = 1.0e5;

    /// Builds one table per distinct guest species at nodes (i/n) of `lattice`, n by the
    /// reference's sizing rule; precision FP64 is the reference's arithmetic.
    template <class LJ, class P, class Lat>
    FrameworkEnergyGrid(Device& dev, const LJ& forceField, const P& frameworkParticles, const Lat& lattice,
                        const std::vector<std::uint32_t>& guestSpeciesIds, double spacingAngstrom,
                        aleator_precision precision = ALEATOR_PRECISION_FP64)
        : dev_(&dev), lattice_(detail::lattice(lattice)), spacing_(spacingAngstrom) {
        if (!(spacingAngstrom > 0.0))
            throw std::invalid_argument("FrameworkEnergyGrid: spacingAngstrom must be positive");
        check(aleator_gpu_grid_point_counts(&lattice_, spacing_, counts_.data()));
        std::vector<std::uint32_t> g = guestSpeciesIds;
        std::sort(g.begin(), g.end());
        g.erase(std::unique(g.begin(), g.end()), g.end());
        Framework fw(dev, forceField, frameworkParticles, lattice);
        const size_t total = nodes();
        std::vector<double> out(total * g.size());
        if (!g.empty())
            check(aleator_gpu_grid_build(fw.handle(), nullptr, g.data(), (int32_t)g.size(), 0, counts_.data(), precision,
                                         kEnergyCapKelvin, out.data()));
        for (size_t t = 0; t < g.size(); ++t)
            tables_.emplace(g[t], std::vector<double>(out.begin() + t * total, out.begin() + (t + 1) * total));
    }

    /// Trilinearly interpolated energy (K) of `guestSpeciesId` at a Cartesian position, on the
    /// device; std::out_of_range for a species without a table.
    [[nodiscard]] double interpolate(std::uint32_t guestSpeciesId, const std::array<double, 3>& position) const {
        return interpolate(guestSpeciesId, std::vector<std::array<double, 3>>{position}).front();
    }
    /// Many positions in one launch.
    [[nodiscard]] std::vector<double> interpolate(std::uint32_t guestSpeciesId,
                                                  const std::vector<std::array<double, 3>>& positions) const;

    [[nodiscard]] double spacingAngstrom() const noexcept { return spacing_; }
    [[nodiscard]] std::array<int, 3> pointCounts() const noexcept { return {counts_[0], counts_[1], counts_[2]}; }
    /// Species with a table, ascending (no reference counterpart).
    [[nodiscard]] std::vector<std::uint32_t> speciesIds() const {
        std::vector<std::uint32_t> ids;
        for (const auto& [s, t] : tables_) ids.push_back(s);
        return ids;
    }
    /// Node values [nz][ny][nx] of one table (no reference counterpart; the reference keeps them private).
    [[nodiscard]] const std::vector<double>& table(std::uint32_t guestSpeciesId) const {
        const auto it = tables_.find(guestSpeciesId);
        if (it == tables_.end()) throw std::out_of_range("FrameworkEnergyGrid: no table for this guest species");
        return it->second;
    }

    void serialize(const std::filesystem::path& path, const FrameworkEnergyGridCacheKey& key) const {
        std::vector<std::uint32_t> sp;
        std::vector<double> all;
        for (const auto& [s, t] : tables_) {
            sp.push_back(s);
            all.insert(all.end(), t.begin(), t.end());
        }
        check(aleator_grid_cache_serialize(path.string().c_str(), key.c_key(), counts_.data(), sp.data(),
                                           (int32_t)sp.size(), all.data()));
    }
    [[nodiscard]] static std::filesystem::path cacheFilePath(const std::filesystem::path& cacheDirectory,
                                                              const FrameworkEnergyGridCacheKey& key) {
        std::string buf(4096, '\0');
        check(aleator_grid_cache_file_path(cacheDirectory.string().c_str(), key.c_key(), buf.data(), buf.size()));
        return std::filesystem::path(buf.c_str());
    }
    /// std::nullopt on a miss; std::runtime_error for a version or key mismatch or a corrupt file.
    [[nodiscard]] static std::optional<FrameworkEnergyGrid> loadFromCache(Device& dev,
                                                                         const std::filesystem::path& cacheDirectory,
                                                                         const FrameworkEnergyGridCacheKey& key) {
        int32_t n[3], nt = 0, hit = 0;
        const std::string dir = cacheDirectory.string();
        check(aleator_grid_cache_load(dir.c_str(), key.c_key(), n, nullptr, &nt, nullptr, 0, &hit));
        if (!hit) return std::nullopt;
        FrameworkEnergyGrid g(dev, key.lattice(), key.spacingAngstrom());
        const size_t total = (size_t)n[0] * n[1] * n[2];
        std::vector<std::uint32_t> sp(std::max(nt, 1));
        std::vector<double> all(total * (size_t)nt);
        check(aleator_grid_cache_load(dir.c_str(), key.c_key(), n, sp.data(), &nt, all.data(), all.size(), &hit));
        for (int i = 0; i < 3; ++i) g.counts_[i] = n[i];
        for (int t = 0; t < nt; ++t)
            g.tables_.emplace(sp[t], std::vector<double>(all.begin() + t * total, all.begin() + (t + 1) * total));
        return g;
    }

private:
    FrameworkEnergyGrid(Device& dev, const aleator_lattice& lat, double spacing) : dev_(&dev), lattice_(lat), spacing_(spacing) {}
    [[nodiscard]] size_t nodes() const noexcept { return (size_t)counts_[0] * counts_[1] * counts_[2]; }
    Device* dev_;
    aleator_lattice lattice_;
    double spacing_;
    std::array<int32_t, 3> counts_{1, 1, 1};
    std::map<std::uint32_t, std::vector<double>> tables_;
};

inline std::vector<double> FrameworkEnergyGrid::interpolate(std::uint32_t guestSpeciesId,
                                                            const std::vector<std::array<double, 3>>& positions) const {
    const auto& t = table(guestSpeciesId);
    std::vector<double> pos;
    pos.reserve(3 * positions.size());
    for (const auto& p : positions) pos.insert(pos.end(), p.begin(), p.end());
    std::vector<double> out(positions.size());
    check(aleator_gpu_grid_interpolate_host(dev_->handle(), &lattice_, counts_.data(), t.data(), pos.data(),
                                            positions.size(), out.data()));
    return out;
}

/// engines::FrameworkEnergyGridCacheRequest (framework_energy_grid.hpp:208-212).
struct FrameworkEnergyGridCacheRequest {
    std::filesystem::path cacheDirectory;
    std::vector<std::uint32_t> guestSpeciesIds;
    double spacingAngstrom = 0.0;
};

/// engines::loadOrBuildFrameworkEnergyGrid (framework_energy_grid.cc:410-435): load the cache
/// entry of these inputs, or build on the device and write it.
template <class LJ, class P, class Lat>
FrameworkEnergyGrid loadOrBuildFrameworkEnergyGrid(Device& dev, const LJ& forceField, const P& frameworkParticles,
                                                   const Lat& lattice, const FrameworkEnergyGridCacheRequest& request,
                                                   bool* wasHit = nullptr) {
    const FrameworkEnergyGridCacheKey key(forceField, frameworkParticles, lattice, request.guestSpeciesIds,
                                          request.spacingAngstrom);
    if (auto cached = FrameworkEnergyGrid::loadFromCache(dev, request.cacheDirectory, key)) {
        if (wasHit) *wasHit = true;
        return std::move(*cached);
    }
    if (wasHit) *wasHit = false;
    FrameworkEnergyGrid grid(dev, forceField, frameworkParticles, lattice, request.guestSpeciesIds,
                             request.spacingAngstrom);
    grid.serialize(FrameworkEnergyGrid::cacheFilePath(request.cacheDirectory, key), key);
    return grid;
}

/// One chain of MonteCarloEngine's GCMC ensemble (monte_carlo_engine.hpp:143-151): fugacity in
/// Pa, temperature, and the Philox4x32Rng (seed, stream) it draws from.
struct ChainConfig {
    double fugacityPascal = 0.0;
    double temperatureKelvin = 0.0;
    std::uint64_t seed = 0, stream = 0;
};

/// Many single-species GCMC chains of MonteCarloEngine::run (monte_carlo_engine.cc:405-671) on
/// one device framework, advanced together; with a FrameworkEnergyGrid the framework LJ is the
/// interpolated grid, as the engine's frameworkEnergyGrid argument makes it
/// (monte_carlo_engine.hpp:129-151, monte_carlo_engine.cc:341-374).  Each chain follows the
/// reference engine's trajectory for its (seed, stream).
class GcmcChains {
public:
    template <class Site>
    GcmcChains(const Framework& fw, const std::vector<Site>& species, const std::vector<ChainConfig>& chains,
               double maxTranslationAngstrom = 1.5, double maxRotationRadians = 0.5, int maxMolecules = 1024,
               const FrameworkEnergyGrid* frameworkEnergyGrid = nullptr)
        : n_chains_((int)chains.size()), n_sites_((int)species.size()) {
        std::vector<aleator_molecule_site> s;
        for (const auto& v : species) s.push_back({v.x, v.y, v.z, v.ljSpecies, v.charge});
        std::vector<aleator_gcmc_chain> c;
        for (const auto& v : chains) c.push_back({v.fugacityPascal, v.temperatureKelvin, v.seed, v.stream});
        check(aleator_gpu_gcmc_create(fw.handle(), s.data(), (int32_t)s.size(), c.data(), (int32_t)c.size(),
                                      maxTranslationAngstrom, maxRotationRadians, maxMolecules, &g_));
        try {
            setEnergyGrid(frameworkEnergyGrid);
        } catch (...) {
            aleator_gpu_gcmc_destroy(g_);
            throw;
        }
    }
    ~GcmcChains() {
        if (g_) aleator_gpu_gcmc_destroy(g_);
    }
    GcmcChains(const GcmcChains&) = delete;
    GcmcChains& operator=(const GcmcChains&) = delete;

    /// Framework LJ from `grid` from the next run on (std::out_of_range if a site's species has
    /// no table); nullptr restores the direct sum.
    void setEnergyGrid(const FrameworkEnergyGrid* grid) {
        if (!grid) {
            check(aleator_gpu_gcmc_set_energy_grid(g_, nullptr, nullptr, 0, nullptr));
            return;
        }
        const auto ids = grid->speciesIds();
        if (ids.empty()) throw std::out_of_range("FrameworkEnergyGrid: no tables");
        std::vector<double> all;
        for (auto id : ids) all.insert(all.end(), grid->table(id).begin(), grid->table(id).end());
        const auto n = grid->pointCounts();
        const int32_t counts[3] = {n[0], n[1], n[2]};
        check(aleator_gpu_gcmc_set_energy_grid_host(g_, counts, ids.data(), (int32_t)ids.size(), all.data()));
    }
    // Guest-guest candidates through guest cells (the engine's guestCellList_): 0 full scan
    // (the default), 1 cells when every axis has at least 4 and a step starts with 1024+ guest
    // sites, 2 cells in every step whenever the sites fit.
    void setGuestCells(int mode) { check(aleator_gpu_gcmc_set_guest_cells(g_, mode)); }
    std::array<int, 3> guestCells() const {
        int32_t n[3];
        check(aleator_gpu_gcmc_guest_cells(g_, n));
        return {n[0], n[1], n[2]};
    }
    /// Advances every chain by numSteps (MonteCarloEngine::run).
    void run(std::uint64_t numSteps) { check(aleator_gpu_gcmc_run(g_, numSteps, nullptr)); }
    /// Per chain: molecules now, steps and summed molecule count since the last reset, and the
    /// attempted / accepted counts per move kind.
    [[nodiscard]] std::vector<aleator_gcmc_stats> statistics() const {
        std::vector<aleator_gcmc_stats> st(n_chains_);
        check(aleator_gpu_gcmc_stats(g_, st.data()));
        for (const auto& c : st)
            if (c.error) throw std::runtime_error("GcmcChains: a chain exceeded its molecule capacity");
        return st;
    }
    void resetStatistics() { check(aleator_gpu_gcmc_reset_stats(g_)); }
    /// Site positions of one chain, molecule by molecule in the engine's order.
    [[nodiscard]] std::vector<std::array<double, 3>> sitePositions(int chain) const {
        int32_t n = 0;
        check(aleator_gpu_gcmc_positions(g_, chain, nullptr, 0, &n));
        std::vector<double> xyz((size_t)n * n_sites_ * 3);
        check(aleator_gpu_gcmc_positions(g_, chain, xyz.data(), xyz.size(), &n));
        std::vector<std::array<double, 3>> out(xyz.size() / 3);
        for (size_t i = 0; i < out.size(); ++i) out[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        return out;
    }
    /// EwaldIncrementalState::committedEnergy of the chain (0 for an uncharged framework).
    [[nodiscard]] double electrostaticEnergy(int chain) const {
        double e = 0.0;
        check(aleator_gpu_gcmc_electrostatic_energy(g_, chain, &e));
        return e;
    }
    /// Largest relative Ewald-energy drift seen at a resync, per chain (the reference's 1e-10
    /// drift gate, test_ewald_incremental_state.cc:137-222, checked inside the chains).
    [[nodiscard]] std::vector<double> maxResyncDrift() const {
        std::vector<double> drift(n_chains_);
        std::vector<std::uint64_t> count(n_chains_);
        check(aleator_gpu_gcmc_resync_check(g_, drift.data(), count.data()));
        return drift;
    }
    /// One chain's state as a checkpoint record; loadChain restores it.
    [[nodiscard]] std::vector<unsigned char> saveChain(int chain) const {
        std::uint64_t n = 0, written = 0;
        check(aleator_gpu_gcmc_checkpoint_size(g_, chain, &n));
        std::vector<unsigned char> rec(n);
        check(aleator_gpu_gcmc_save_chain(g_, chain, rec.data(), rec.size(), &written));
        rec.resize(written);
        return rec;
    }
    void loadChain(int chain, const std::vector<unsigned char>& rec) {
        check(aleator_gpu_gcmc_load_chain(g_, chain, rec.data(), rec.size()));
    }
    [[nodiscard]] aleator_gcmc* handle() const noexcept { return g_; }

private:
    aleator_gcmc* g_ = nullptr;
    int n_chains_, n_sites_;
};

/// forcefield::EwaldIncrementalState (ewald_incremental_state.hpp:60-140): S(k) of the
/// particles resident on the device; proposals never mutate, commits add the same delta.
class EwaldIncrementalState {
public:
    template <class E, class P, class Lat>
    EwaldIncrementalState(Device& dev, const E& ewald, const P& particles, const Lat& lattice) {
        // the framework object needs LJ constants; a zero-depth table stands in (unused here)
        struct Zero {
            struct S {
                double epsilon, sigma;
            };
            std::vector<S> s;
            double rc;
            const std::vector<S>& perSpecies() const { return s; }
            double cutoff() const { return rc; }
            int truncation() const { return ALEATOR_LJ_TRUNCATED; }
            int mixingRule() const { return ALEATOR_MIX_LORENTZ_BERTHELOT; }
        } zero;
        std::uint32_t ns = 1;
        for (auto s : particles.species) ns = std::max<std::uint32_t>(ns, s + 1);
        zero.s.assign(ns, {0.0, 1.0});
        zero.rc = ewald.realSpaceCutoff();
        fw_.emplace(dev, zero, ewald, particles, lattice);
        check(aleator_gpu_ewald_state_create(fw_->handle(), &st_));
    }
    ~EwaldIncrementalState() {
        if (st_) aleator_gpu_ewald_state_destroy(st_);
    }
    EwaldIncrementalState(const EwaldIncrementalState&) = delete;
    EwaldIncrementalState& operator=(const EwaldIncrementalState&) = delete;

    [[nodiscard]] double committedEnergy() const {
        double e = 0.0;
        check(aleator_gpu_ewald_state_committed_energy(st_, &e));
        return e;
    }
    template <class Site>
    [[nodiscard]] double proposeMove(const std::vector<Site>& oldSites, const std::vector<Site>& newSites) const {
        return propose(detail::charged(oldSites), detail::charged(newSites));
    }
    template <class Site>
    [[nodiscard]] double proposeInsertion(const std::vector<Site>& newSites) const {
        return propose({}, detail::charged(newSites));
    }
    template <class Site>
    [[nodiscard]] double proposeDeletion(const std::vector<Site>& existingSites) const {
        return propose(detail::charged(existingSites), {});
    }
    template <class Site>
    void commitMove(const std::vector<Site>& oldSites, const std::vector<Site>& newSites) {
        commit(detail::charged(oldSites), detail::charged(newSites));
    }
    template <class Site>
    void commitInsertion(const std::vector<Site>& newSites) {
        commit({}, detail::charged(newSites));
    }
    template <class Site>
    void commitDeletion(const std::vector<Site>& existingSites) {
        commit(detail::charged(existingSites), {});
    }
    template <class P>
    [[nodiscard]] double recomputeFromScratch(const P& particles,
                                              const std::vector<std::vector<std::size_t>>& molecules) const {
        const detail::Particles p(particles);
        const auto [off, sites] = csr(molecules);
        double e = 0.0;
        check(aleator_gpu_ewald_state_recompute(st_, &p.view, off.data(), sites.data(), molecules.size(), &e));
        return e;
    }
    template <class P>
    void resync(const P& particles, const std::vector<std::vector<std::size_t>>& molecules) {
        const detail::Particles p(particles);
        const auto [off, sites] = csr(molecules);
        check(aleator_gpu_ewald_state_resync(st_, &p.view, off.data(), sites.data(), molecules.size()));
    }
    [[nodiscard]] std::vector<double> rawStateForTesting() const {
        uint64_t nk = 0, written = 0;
        check(aleator_gpu_framework_num_k(fw_->handle(), &nk));
        std::vector<double> out(2 * nk + 2);
        check(aleator_gpu_ewald_state_raw(st_, out.data(), out.size(), &written));
        out.resize(written);
        return out;
    }

private:
    double propose(const std::vector<aleator_charged_site>& rem, const std::vector<aleator_charged_site>& add) const {
        double d = 0.0;
        check(aleator_gpu_ewald_state_propose(st_, rem.data(), (int32_t)rem.size(), add.data(), (int32_t)add.size(), &d));
        return d;
    }
    void commit(const std::vector<aleator_charged_site>& rem, const std::vector<aleator_charged_site>& add) {
        check(aleator_gpu_ewald_state_commit(st_, rem.data(), (int32_t)rem.size(), add.data(), (int32_t)add.size()));
    }
    static std::pair<std::vector<uint64_t>, std::vector<uint64_t>> csr(const std::vector<std::vector<std::size_t>>& m) {
        std::vector<uint64_t> off{0}, sites;
        for (const auto& mol : m) {
            sites.insert(sites.end(), mol.begin(), mol.end());
            off.push_back(sites.size());
        }
        return {off, sites};
    }
    std::optional<Framework> fw_;
    aleator_ewald_state* st_ = nullptr;
};

/// MonteCarloEngine::widomInsertionHenryCoefficient (monte_carlo_engine.cc:850-917) of one
/// rigid species against a device framework: K_H = <exp(-dU/T)>/T over numTrials insertions
/// drawn from Philox4x32Rng(seed, stream) from call index firstCallIndex; dU per trial into
/// `du` when not null.
template <class Site>
double widomInsertionHenryCoefficient(const Framework& fw, const std::vector<Site>& sites, std::size_t numTrials,
                                      double temperature, std::uint64_t seed, std::uint64_t stream,
                                      std::uint64_t firstCallIndex = 0,
                                      aleator_precision precision = ALEATOR_PRECISION_FP64,
                                      std::vector<double>* du = nullptr) {
    std::vector<aleator_molecule_site> s;
    for (const auto& v : sites) s.push_back({v.x, v.y, v.z, v.ljSpecies, v.charge});
    if (du) du->assign(numTrials, 0.0);
    double sum = 0.0;
    check(aleator_gpu_widom(fw.handle(), s.data(), (int32_t)s.size(), seed, stream, firstCallIndex, numTrials,
                            temperature, precision, du ? du->data() : nullptr, &sum));
    return sum / ((double)numTrials * temperature);
}

/// core::Lattice as returned by the reader: rows are the lattice vectors (lattice.hpp:21-30).
struct Lattice {
    std::array<std::array<double, 3>, 3> m{};
    [[nodiscard]] const std::array<std::array<double, 3>, 3>& matrix() const noexcept { return m; }
};

/// The core::ParticleData members a structure file fills (particle_data.hpp:20-99).
struct ParticleData {
    std::vector<double> x, y, z, mass, charge;
    std::vector<std::uint32_t> species;
    [[nodiscard]] std::size_t size() const noexcept { return x.size(); }
};

/// io::StructureData (io/structure_io.hpp:18-37).
struct StructureData {
    ParticleData particles;
    Lattice lattice;
    std::vector<std::string> speciesSymbols;
    std::vector<double> occupancy;
    std::vector<std::string> sourceLabel;
};

namespace detail {
inline StructureData take_structure(aleator_structure* h) {
    struct Guard {
        aleator_structure* h;
        ~Guard() { aleator_structure_destroy(h); }
    } guard{h};
    StructureData s;
    std::uint64_t n = 0;
    std::uint32_t ns = 0;
    check(aleator_structure_size(h, &n, &ns));
    aleator_lattice lat{};
    check(aleator_structure_lattice(h, &lat));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s.lattice.m[i][j] = lat.m[i][j];
    auto& p = s.particles;
    for (auto* v : {&p.x, &p.y, &p.z, &p.mass, &p.charge, &s.occupancy}) v->resize(n);
    p.species.resize(n);
    check(aleator_structure_atoms(h, p.x.data(), p.y.data(), p.z.data(), p.mass.data(), p.charge.data(),
                                  p.species.data(), s.occupancy.data()));
    const char* str = nullptr;
    for (std::uint32_t i = 0; i < ns; ++i) {
        check(aleator_structure_species_symbol(h, i, &str));
        s.speciesSymbols.emplace_back(str);
    }
    for (std::uint64_t i = 0; i < n; ++i) {
        check(aleator_structure_source_label(h, i, &str));
        s.sourceLabel.emplace_back(str);
    }
    return s;
}
}  // namespace detail

/// io::readCif (io/structure_io.hpp:39-50): throws CifParseError on malformed input or a file
/// that cannot be opened.
inline StructureData readCif(const std::filesystem::path& file) {
    aleator_structure* h = nullptr;
    check(aleator_read_cif(file.string().c_str(), &h));
    return detail::take_structure(h);
}

/// readCif from CIF text in memory.
inline StructureData parseCif(const std::string& text) {
    aleator_structure* h = nullptr;
    check(aleator_parse_cif(text.data(), text.size(), &h));
    return detail::take_structure(h);
}

}  // namespace aleator::gpu
