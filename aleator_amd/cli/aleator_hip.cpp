// aleator_hip.cpp -- command-line caller of the device path: `aleator-hip gcmc run <config.toml>`
// runs the reference CLI's GCMC protocol (cli/main.cc:173-588) on one MI355X, and
// `aleator-hip widom run <config.toml>` computes the same configuration's Widom Henry coefficient
// (MonteCarloEngine::widomInsertionHenryCoefficient, monte_carlo_engine.cc:850-917) there.
//
// The configuration schema, its validation and error messages, the resolved-configuration
// artifact and the printed results follow the reference (io/config.cc:116-350, cli/main.cc), so
// a config written for the reference CLI runs here unchanged and, with the same seed, prints the
// same <N>: the device chain takes the reference engine's draws in the reference's order.
// With gcmc.energy_grid_spacing_angstrom the framework LJ energy comes from the adsorbate's grid
// table, loaded from gcmc.energy_grid_cache_directory (the reference's cache file format, so a
// grid the reference cached loads here and the other way round) or built on the device and
// cached.  Not part of this build, and reported as such: [pore] configs and `bench`.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "aleator_gpu.h"
#include "aleator_gpu.hpp"
#include "toml_lite.hpp"

namespace fs = std::filesystem;
namespace ag = aleator::gpu;

namespace {

constexpr int kExitSuccess = 0;
constexpr int kExitConfigError = 1;
constexpr int kExitNotInThisBuild = 2;
constexpr const char* kVersion = "0.1.0-hip";
constexpr double kAmuToKg = 1.66053906660e-27;  // CODATA 2018
constexpr double kAvogadro = 6.02214076e23;     // SI 2019
constexpr double kPascalToInternal = 1.0 / (1.380649e-23 * 1.0e30);  // Pa -> K / A^3

// ---------------------------------------------------------------- configuration
// io::ConfigError (io/config_error.hpp:11-30): "<file>:<line>:<column>: [<key>] <problem>".
struct ConfigError : std::runtime_error {
    ConfigError(const std::string& file, const std::string& key, const std::string& problem, int line, int column)
        : std::runtime_error(file + ":" + std::to_string(line) + ":" + std::to_string(column) + ": [" + key + "] " +
                             problem) {}
};

struct RunConfig {
    std::string name;
    fs::path output_directory = ".";
    std::uint64_t rng_seed = 0;
    unsigned thread_count = 1;
};
struct FrameworkSpecies {
    std::string element;
    double epsilon, sigma;
};
struct Critical {
    double tc, pc, omega;
};
struct Adsorbate {
    std::string name;
    double epsilon = 0, sigma = 0, mass = 0, charge = 0;
    std::optional<Critical> critical;
};
struct GcmcConfig {
    fs::path framework_cif;
    double temperature = 0, pressure_bar = 0, cutoff = 12.0;
    std::size_t equilibration_steps = 10000, production_steps = 20000;
    std::vector<FrameworkSpecies> framework_lj;
    Adsorbate adsorbate;
    std::optional<double> grid_spacing;
    fs::path grid_cache_directory = "energy_grid_cache";
    std::optional<std::array<int, 3>> supercell;
};
struct Config {
    RunConfig run;
    GcmcConfig gcmc;
};

using toml_lite::Node;

// Typed access with the reference loader's messages (io/config.cc:30-95).
class Keys {
public:
    explicit Keys(std::string file) : file_(std::move(file)) {}

    const Node& table(const Node& parent, const std::string& key, const std::string& path) const {
        const Node* n = parent.find(key);
        if (!n || n->kind != Node::Table)
            throw ConfigError(file_, path, n ? std::string("expected a table, found ") + n->kind_name()
                                             : "missing required table",
                              n ? n->line : parent.line, n ? n->column : parent.column);
        return *n;
    }
    template <class T>
    T need(const Node& parent, const std::string& key, const std::string& path) const {
        const Node* n = parent.find(key);
        if (!n) throw ConfigError(file_, path, "missing required key (expected " + what<T>() + ")", parent.line, parent.column);
        return read<T>(*n, path);
    }
    template <class T>
    T opt(const Node& parent, const std::string& key, T fallback, const std::string& path) const {
        const Node* n = parent.find(key);
        return n ? read<T>(*n, path) : fallback;
    }
    void positive(double v, const Node& parent, const std::string& key, const std::string& path) const {
        if (!(v > 0.0)) {
            const Node* n = parent.find(key);
            throw ConfigError(file_, path, "must be positive (got " + std::to_string(v) + ")", n ? n->line : parent.line,
                              n ? n->column : parent.column);
        }
    }
    [[noreturn]] void fail(const Node& at, const std::string& path, const std::string& problem) const {
        throw ConfigError(file_, path, problem, at.line, at.column);
    }

private:
    template <class T>
    static std::string what() {
        if constexpr (std::is_same_v<T, std::string>) return "a string";
        else if constexpr (std::is_same_v<T, double>) return "a number";
        else return "an integer";
    }
    template <class T>
    T read(const Node& n, const std::string& path) const {
        std::optional<T> v;
        if constexpr (std::is_same_v<T, std::string>) v = n.as_string();
        else if constexpr (std::is_same_v<T, double>) v = n.as_double();
        else v = n.as_int();
        if (!v) throw ConfigError(file_, path, "expected " + what<T>() + ", found " + n.kind_name(), n.line, n.column);
        return *v;
    }
    std::string file_;
};

Node parse_config(const fs::path& file) {
    try {
        return toml_lite::parse_file(file.string());
    } catch (const toml_lite::ParseError& e) {
        throw ConfigError(file.string(), "<parse>", e.what(), e.line, e.column);
    }
}

fs::path relative_to(const fs::path& config, const std::string& p) {
    const fs::path c(p);
    return c.is_absolute() ? c.lexically_normal() : (config.parent_path() / c).lexically_normal();
}

// parseRunTable (io/config.cc:116-150)
RunConfig run_table(const Node& root, const fs::path& file) {
    const Keys k(file.string());
    const Node& run = k.table(root, "run", "run");
    RunConfig r;
    r.name = k.need<std::string>(run, "name", "run.name");
    r.output_directory = relative_to(file, k.opt<std::string>(run, "output_directory", ".", "run.output_directory"));
    r.rng_seed = (std::uint64_t)k.opt<std::int64_t>(run, "rng_seed", 0, "run.rng_seed");
    r.thread_count = (unsigned)k.opt<std::int64_t>(run, "thread_count", 1, "run.thread_count");
    if (r.thread_count != 1)
        k.fail(*run.find("thread_count"), "run.thread_count",
               "not supported in this build: only thread_count = 1 is accepted (one device chain per run)");
    return r;
}

// loadGcmcConfig (io/config.cc:158-316)
Config load_gcmc_config(const fs::path& file) {
    const Node root = parse_config(file);
    const Keys k(file.string());
    Config c;
    c.run = run_table(root, file);
    const Node& g = k.table(root, "gcmc", "gcmc");
    GcmcConfig& cfg = c.gcmc;
    cfg.framework_cif = relative_to(file, k.need<std::string>(g, "framework_cif", "gcmc.framework_cif"));
    cfg.temperature = k.need<double>(g, "temperature_kelvin", "gcmc.temperature_kelvin");
    k.positive(cfg.temperature, g, "temperature_kelvin", "gcmc.temperature_kelvin");
    cfg.pressure_bar = k.need<double>(g, "pressure_bar", "gcmc.pressure_bar");
    k.positive(cfg.pressure_bar, g, "pressure_bar", "gcmc.pressure_bar");
    cfg.cutoff = k.opt<double>(g, "cutoff_angstrom", cfg.cutoff, "gcmc.cutoff_angstrom");
    k.positive(cfg.cutoff, g, "cutoff_angstrom", "gcmc.cutoff_angstrom");
    const auto eq = k.opt<std::int64_t>(g, "equilibration_steps", (std::int64_t)cfg.equilibration_steps,
                                        "gcmc.equilibration_steps");
    const auto pr = k.opt<std::int64_t>(g, "production_steps", (std::int64_t)cfg.production_steps,
                                        "gcmc.production_steps");
    if (eq < 0) k.fail(*g.find("equilibration_steps"), "gcmc.equilibration_steps", "must not be negative");
    if (pr <= 0) k.fail(g.find("production_steps") ? *g.find("production_steps") : g, "gcmc.production_steps", "must be positive");
    cfg.equilibration_steps = (std::size_t)eq;
    cfg.production_steps = (std::size_t)pr;
    if (g.find("energy_grid_spacing_angstrom")) {
        cfg.grid_spacing = k.need<double>(g, "energy_grid_spacing_angstrom", "gcmc.energy_grid_spacing_angstrom");
        k.positive(*cfg.grid_spacing, g, "energy_grid_spacing_angstrom", "gcmc.energy_grid_spacing_angstrom");
    }
    cfg.grid_cache_directory = relative_to(file, k.opt<std::string>(g, "energy_grid_cache_directory", "energy_grid_cache",
                                                                     "gcmc.energy_grid_cache_directory"));
    if (const Node* sc = g.find("supercell")) {
        if (sc->kind != Node::Array || sc->items.size() != 3)
            k.fail(*sc, "gcmc.supercell", "expected an array of exactly 3 positive integers [nx, ny, nz]");
        std::array<int, 3> v{};
        for (int i = 0; i < 3; ++i) {
            const auto x = sc->items[i].as_int();
            if (!x || *x < 1) k.fail(sc->items[i], "gcmc.supercell", "each entry must be a positive integer (>= 1)");
            v[i] = (int)*x;
        }
        cfg.supercell = v;
    }
    const Node* lj = g.find("framework_lj");
    if (!lj || lj->kind != Node::Array || lj->items.empty())
        k.fail(lj ? *lj : g, "gcmc.framework_lj", "missing or empty; at least one [[gcmc.framework_lj]] entry is required");
    for (size_t i = 0; i < lj->items.size(); ++i) {
        const Node& e = lj->items[i];
        const std::string p = "gcmc.framework_lj[" + std::to_string(i) + "]";
        if (e.kind != Node::Table) k.fail(e, p, "expected a table");
        FrameworkSpecies s;
        s.element = k.need<std::string>(e, "element", p + ".element");
        s.epsilon = k.need<double>(e, "epsilon_kelvin", p + ".epsilon_kelvin");
        k.positive(s.epsilon, e, "epsilon_kelvin", p + ".epsilon_kelvin");
        s.sigma = k.need<double>(e, "sigma_angstrom", p + ".sigma_angstrom");
        k.positive(s.sigma, e, "sigma_angstrom", p + ".sigma_angstrom");
        cfg.framework_lj.push_back(s);
    }
    const Node& a = k.table(g, "adsorbate", "gcmc.adsorbate");
    Adsorbate& ad = cfg.adsorbate;
    ad.name = k.need<std::string>(a, "name", "gcmc.adsorbate.name");
    ad.epsilon = k.need<double>(a, "epsilon_kelvin", "gcmc.adsorbate.epsilon_kelvin");
    k.positive(ad.epsilon, a, "epsilon_kelvin", "gcmc.adsorbate.epsilon_kelvin");
    ad.sigma = k.need<double>(a, "sigma_angstrom", "gcmc.adsorbate.sigma_angstrom");
    k.positive(ad.sigma, a, "sigma_angstrom", "gcmc.adsorbate.sigma_angstrom");
    ad.mass = k.need<double>(a, "mass_amu", "gcmc.adsorbate.mass_amu");
    k.positive(ad.mass, a, "mass_amu", "gcmc.adsorbate.mass_amu");
    ad.charge = k.opt<double>(a, "charge_e", 0.0, "gcmc.adsorbate.charge_e");
    const bool has_tc = a.find("critical_temperature_kelvin"), has_pc = a.find("critical_pressure_pascal"),
               has_w = a.find("acentric_factor");
    if (has_tc || has_pc || has_w) {
        if (!(has_tc && has_pc && has_w))
            k.fail(a, "gcmc.adsorbate",
                   "critical_temperature_kelvin, critical_pressure_pascal, and acentric_factor must all be given "
                   "together, or all omitted (omitted -> ideal-gas fugacity approximation)");
        Critical cp;
        cp.tc = k.need<double>(a, "critical_temperature_kelvin", "gcmc.adsorbate.critical_temperature_kelvin");
        k.positive(cp.tc, a, "critical_temperature_kelvin", "gcmc.adsorbate.critical_temperature_kelvin");
        cp.pc = k.need<double>(a, "critical_pressure_pascal", "gcmc.adsorbate.critical_pressure_pascal");
        k.positive(cp.pc, a, "critical_pressure_pascal", "gcmc.adsorbate.critical_pressure_pascal");
        cp.omega = k.need<double>(a, "acentric_factor", "gcmc.adsorbate.acentric_factor");
        ad.critical = cp;
    }
    return c;
}

enum class Kind { Gcmc, Pore, Unknown };
Kind detect_kind(const fs::path& file) {
    const Node root = parse_config(file);
    const Node* g = root.find("gcmc");
    const Node* p = root.find("pore");
    const bool hg = g && g->kind == Node::Table, hp = p && p->kind == Node::Table;
    return hg && !hp ? Kind::Gcmc : hp && !hg ? Kind::Pore : Kind::Unknown;
}

// ---------------------------------------------------------------- prepared system
// prepareGcmc (cli/main.cc:173-280): framework from the CIF, replicated to the minimum image
// supercell (or the config's larger override), LJ parameters by element, adsorbate last.
struct Prepared {
    ag::StructureData structure;  // after replication
    std::array<int, 3> supercell{1, 1, 1};
    std::vector<double> eps, sig;
    std::uint32_t adsorbate_species = 0;
    double fugacity = 0;
    bool ideal_gas = true;
    double mass_amu = 0;
};

Prepared prepare(const Config& cfg) {
    Prepared p;
    p.structure = ag::readCif(cfg.gcmc.framework_cif);
    aleator_lattice lat{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) lat.m[i][j] = p.structure.lattice.m[i][j];
    std::int32_t minimum[3];
    ag::check(aleator_supercell_replication(&lat, cfg.gcmc.cutoff, minimum));
    std::array<int, 3> reps = {minimum[0], minimum[1], minimum[2]};
    if (cfg.gcmc.supercell) {
        const auto& r = *cfg.gcmc.supercell;
        static const char* axis[3] = {"x", "y", "z"};
        for (int i = 0; i < 3; ++i)
            if (r[i] < minimum[i])
                throw std::runtime_error(
                    "gcmc.supercell requests " + std::to_string(r[0]) + "x" + std::to_string(r[1]) + "x" +
                    std::to_string(r[2]) + ", but the minimum replication required to satisfy minimum image at "
                    "cutoff_angstrom=" + std::to_string(cfg.gcmc.cutoff) + " for this framework's lattice is " +
                    std::to_string(minimum[0]) + "x" + std::to_string(minimum[1]) + "x" + std::to_string(minimum[2]) +
                    " (axis " + axis[i] + ": requested " + std::to_string(r[i]) + " < required " +
                    std::to_string(minimum[i]) + ") -- gcmc.supercell may only override the computed minimum upward, "
                    "never downward");
        reps = r;
    }
    p.supercell = reps;
    if (reps != std::array<int, 3>{1, 1, 1}) {
        auto& s = p.structure;
        const size_t n = s.particles.size(), cells = (size_t)reps[0] * reps[1] * reps[2];
        std::vector<double> x(n * cells), y(n * cells), z(n * cells), q(n * cells);
        std::vector<std::uint32_t> sp(n * cells);
        const aleator_particles in{s.particles.x.data(), s.particles.y.data(), s.particles.z.data(),
                                   s.particles.charge.data(), s.particles.species.data(), n};
        aleator_lattice out{};
        ag::check(aleator_replicate_supercell(&in, &lat, reps[0], reps[1], reps[2], x.data(), y.data(), z.data(),
                                              q.data(), sp.data(), &out));
        // per-atom metadata re-tiled in the same cell-then-atom order
        std::vector<double> mass, occ;
        std::vector<std::string> label;
        for (size_t c = 0; c < cells; ++c)
            for (size_t i = 0; i < n; ++i) {
                mass.push_back(s.particles.mass[i]);
                occ.push_back(s.occupancy[i]);
                label.push_back(s.sourceLabel[i]);
            }
        s.particles.x = std::move(x);
        s.particles.y = std::move(y);
        s.particles.z = std::move(z);
        s.particles.charge = std::move(q);
        s.particles.species = std::move(sp);
        s.particles.mass = std::move(mass);
        s.occupancy = std::move(occ);
        s.sourceLabel = std::move(label);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) s.lattice.m[i][j] = out.m[i][j];
    }
    for (const std::string& sym : p.structure.speciesSymbols) {
        const auto it = std::find_if(cfg.gcmc.framework_lj.begin(), cfg.gcmc.framework_lj.end(),
                                     [&](const FrameworkSpecies& f) { return f.element == sym; });
        if (it == cfg.gcmc.framework_lj.end())
            throw std::runtime_error("gcmc.framework_lj has no entry for element \"" + sym + "\", which is present in " +
                                     cfg.gcmc.framework_cif.string() + " -- add a [[gcmc.framework_lj]] entry for it");
        p.eps.push_back(it->epsilon);
        p.sig.push_back(it->sigma);
    }
    p.adsorbate_species = (std::uint32_t)p.eps.size();
    p.eps.push_back(cfg.gcmc.adsorbate.epsilon);
    p.sig.push_back(cfg.gcmc.adsorbate.sigma);
    const double pressure = cfg.gcmc.pressure_bar * 1.0e5;
    p.fugacity = pressure;
    if (cfg.gcmc.adsorbate.critical) {
        const Critical& c = *cfg.gcmc.adsorbate.critical;
        ag::check(aleator_peng_robinson_fugacity(c.tc, c.pc, c.omega, cfg.gcmc.temperature, pressure, &p.fugacity));
        p.ideal_gas = false;
    }
    for (double m : p.structure.particles.mass) p.mass_amu += m;
    return p;
}

std::string json_escape(std::string_view s) {
    std::string o;
    for (char c : s) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c == '\n') o += "\\n";
        else o += c;
    }
    return o;
}

// resolvedConfigJson (cli/main.cc:344-395), plus "backend":"hip"
std::string resolved_json(const Config& cfg, const std::array<int, 3>& sc, const std::optional<bool>& grid_hit) {
    const auto& r = cfg.run;
    const auto& g = cfg.gcmc;
    std::ostringstream o;
    o << std::setprecision(17);
    o << "{\"run\":{\"name\":\"" << json_escape(r.name) << "\",\"output_directory\":\""
      << json_escape(r.output_directory.string()) << "\",\"rng_seed\":" << r.rng_seed
      << ",\"thread_count\":" << r.thread_count << "},\"gcmc\":{\"framework_cif\":\""
      << json_escape(g.framework_cif.string()) << "\",\"temperature_kelvin\":" << g.temperature
      << ",\"pressure_bar\":" << g.pressure_bar << ",\"cutoff_angstrom\":" << g.cutoff
      << ",\"equilibration_steps\":" << g.equilibration_steps << ",\"production_steps\":" << g.production_steps
      << ",\"supercell\":[" << sc[0] << "," << sc[1] << "," << sc[2] << "],\"energy_grid_spacing_angstrom\":"
      << (g.grid_spacing ? std::to_string(*g.grid_spacing) : "null") << ",\"energy_grid_cache_directory\":\""
      << json_escape(g.grid_cache_directory.string()) << "\",\"energy_grid_cache_status\":"
      << (grid_hit ? (*grid_hit ? "\"hit\"" : "\"will_build\"") : "null")
      << ",\"adsorbate\":{\"name\":\"" << json_escape(g.adsorbate.name) << "\",\"epsilon_kelvin\":"
      << g.adsorbate.epsilon << ",\"sigma_angstrom\":" << g.adsorbate.sigma << ",\"mass_amu\":" << g.adsorbate.mass
      << ",\"charge_e\":" << g.adsorbate.charge << ",\"uses_peng_robinson\":"
      << (g.adsorbate.critical ? "true" : "false") << "},\"framework_lj\":[";
    for (size_t i = 0; i < g.framework_lj.size(); ++i)
        o << (i ? "," : "") << "{\"element\":\"" << json_escape(g.framework_lj[i].element)
          << "\",\"epsilon_kelvin\":" << g.framework_lj[i].epsilon << ",\"sigma_angstrom\":" << g.framework_lj[i].sigma
          << "}";
    o << "]},\"backend\":\"hip\"}";
    return o.str();
}

void print_config(const Config& cfg, const Prepared& p, const std::optional<bool>& grid_hit, bool json) {
    if (json) {
        std::cout << resolved_json(cfg, p.supercell, grid_hit) << "\n";
        return;
    }
    const auto& r = cfg.run;
    const auto& g = cfg.gcmc;
    std::cout << std::setprecision(17);
    std::cout << "run.name                       = " << r.name << "\n"
              << "run.output_directory           = " << r.output_directory.string() << "\n"
              << "run.rng_seed                   = " << r.rng_seed << "\n"
              << "run.thread_count               = " << r.thread_count << "\n"
              << "gcmc.framework_cif             = " << g.framework_cif.string() << "\n"
              << "gcmc.temperature_kelvin        = " << g.temperature << "\n"
              << "gcmc.pressure_bar              = " << g.pressure_bar << "\n"
              << "gcmc.cutoff_angstrom           = " << g.cutoff << "\n"
              << "gcmc.supercell                 = " << p.supercell[0] << "x" << p.supercell[1] << "x"
              << p.supercell[2] << (g.supercell ? " (from gcmc.supercell)" : " (auto-computed minimum)") << "\n"
              << "gcmc.equilibration_steps       = " << g.equilibration_steps << "\n"
              << "gcmc.production_steps          = " << g.production_steps << "\n"
              << "gcmc.energy_grid_spacing_angstrom = "
              << (g.grid_spacing ? std::to_string(*g.grid_spacing) : "not set (direct guest-host sums on the device)")
              << "\n";
    if (g.grid_spacing)
        std::cout << "gcmc.energy_grid_cache_directory = " << g.grid_cache_directory.string() << "\n"
                  << "gcmc.energy_grid_cache_status  = "
                  << (grid_hit ? (*grid_hit ? "hit (will load, not rebuild)" : "miss (will build and cache)") : "unknown")
                  << "\n";
    std::cout << "gcmc.adsorbate.name            = " << g.adsorbate.name << "\n"
              << "gcmc.adsorbate.epsilon_kelvin  = " << g.adsorbate.epsilon << "\n"
              << "gcmc.adsorbate.sigma_angstrom  = " << g.adsorbate.sigma << "\n"
              << "gcmc.adsorbate.mass_amu        = " << g.adsorbate.mass << "\n"
              << "gcmc.adsorbate.charge_e        = " << g.adsorbate.charge << "\n"
              << "gcmc.adsorbate.fugacity_model  = "
              << (g.adsorbate.critical ? "Peng-Robinson" : "ideal gas (no critical properties given -- fugacity = pressure)")
              << "\n"
              << "gcmc.framework_lj (" << g.framework_lj.size() << " species):\n";
    for (const auto& s : g.framework_lj)
        std::cout << "  " << s.element << ": epsilon=" << s.epsilon << " K, sigma=" << s.sigma << " Ang\n";
    std::cout << "backend                        = hip (one MI355X chain)\n";
}

// The energy-grid cache request of a prepared run (frameworkEnergyGridCacheRequest,
// cli/main.cc:278-286): the adsorbate's table at the configured spacing, keyed on the replicated
// framework, its lattice and the force field (makeFrameworkEnergyGridCacheKey).
struct GridKey {
    aleator_lattice lat{};
    aleator_particles atoms{};
    aleator_lj_params lj{};
    std::uint32_t guest = 0;
    aleator_grid_cache_key key{};
    std::string directory;
    GridKey(const Config& cfg, const Prepared& p) : guest(p.adsorbate_species) {
        const auto& s = p.structure;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) lat.m[i][j] = s.lattice.m[i][j];
        atoms = {s.particles.x.data(), s.particles.y.data(), s.particles.z.data(), nullptr, s.particles.species.data(),
                 s.particles.size()};
        lj = {p.eps.data(), p.sig.data(), (std::uint32_t)p.eps.size(), cfg.gcmc.cutoff, ALEATOR_LJ_SHIFTED,
              ALEATOR_MIX_LORENTZ_BERTHELOT};
        key = {&lat, &atoms, &lj, &guest, 1, *cfg.gcmc.grid_spacing};
        directory = cfg.gcmc.grid_cache_directory.string();
    }
    GridKey(const GridKey&) = delete;
    GridKey& operator=(const GridKey&) = delete;
};

// frameworkEnergyGridCacheHitForDryRun (cli/main.cc:331-341): whether the cache holds this
// run's grid, from the file header only; a mismatching or corrupt file is an error, as there.
std::optional<bool> grid_cache_status(const Config& cfg, const Prepared& p) {
    if (!cfg.gcmc.grid_spacing) return std::nullopt;
    const GridKey k(cfg, p);
    std::int32_t n[3], nt = 0, hit = 0;
    ag::check(aleator_grid_cache_load(k.directory.c_str(), &k.key, n, nullptr, &nt, nullptr, 0, &hit));
    return hit != 0;
}
// A device framework of the prepared system: LJ only, shifted at the cutoff, Lorentz-Berthelot
// (the reference CLI runs without electrostatics, cli/main.cc:524-528).
struct DeviceRun {
    ag::Device dev;
    aleator_framework* fw = nullptr;
    DeviceRun(const Config& cfg, const Prepared& p) : dev(0) {
        const auto& s = p.structure;
        aleator_lattice lat{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) lat.m[i][j] = s.lattice.m[i][j];
        const aleator_particles atoms{s.particles.x.data(), s.particles.y.data(), s.particles.z.data(), nullptr,
                                      s.particles.species.data(), s.particles.size()};
        const aleator_lj_params lj{p.eps.data(), p.sig.data(), (std::uint32_t)p.eps.size(), cfg.gcmc.cutoff,
                                   ALEATOR_LJ_SHIFTED, ALEATOR_MIX_LORENTZ_BERTHELOT};
        ag::check(aleator_gpu_framework_create(dev.handle(), &lat, &atoms, &lj, nullptr, &fw));
    }
    ~DeviceRun() {
        if (fw) aleator_gpu_framework_destroy(fw);
    }
};

double cell_volume(const ag::Lattice& l) {
    const auto& m = l.m;
    return std::abs(m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                    m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]));
}

// ---------------------------------------------------------------- commands
int gcmc_run(const fs::path& path, bool dry_run, bool json, int max_molecules) {
    const Config cfg = load_gcmc_config(path);
    const Prepared p = prepare(cfg);
    if (dry_run) {
        print_config(cfg, p, grid_cache_status(cfg, p), json);
        return kExitSuccess;
    }
    std::cerr << "loaded " << p.structure.particles.size() << " framework atoms (" << p.structure.speciesSymbols.size()
              << " species) from " << cfg.gcmc.framework_cif.string() << " (supercell " << p.supercell[0] << "x"
              << p.supercell[1] << "x" << p.supercell[2] << ")\n";
    if (p.ideal_gas)
        std::cerr << "no Peng-Robinson critical properties given for \"" << cfg.gcmc.adsorbate.name
                  << "\" -- using ideal-gas fugacity (fugacity = pressure)\n";
    else
        std::cerr << "fugacity(" << cfg.gcmc.temperature << " K, " << cfg.gcmc.pressure_bar << " bar) = " << std::fixed
                  << std::setprecision(3) << p.fugacity << " Pa (Peng-Robinson)\n"
                  << std::defaultfloat << std::setprecision(6);
    const double volume = cell_volume(p.structure.lattice);
    const double mass_kg = p.mass_amu * kAmuToKg;

    DeviceRun d(cfg, p);
    // buildFrameworkEnergyGridIfConfigured (cli/main.cc:300-322): load the cache entry, or build
    // the table on the device (FP64, capped at 1e5 K) and write it
    std::optional<bool> grid_hit;
    std::vector<std::uint32_t> grid_species;
    std::vector<double> grid_tables;
    std::int32_t grid_n[3] = {1, 1, 1};
    if (cfg.gcmc.grid_spacing) {
        const GridKey k(cfg, p);
        std::cerr << "guest-host energy grid requested at " << *cfg.gcmc.grid_spacing << " Ang spacing (cache: "
                  << k.directory << ")...\n";
        const auto t0 = std::chrono::steady_clock::now();
        std::int32_t nt = 0, hit = 0;
        ag::check(aleator_grid_cache_load(k.directory.c_str(), &k.key, grid_n, nullptr, &nt, nullptr, 0, &hit));
        if (hit) {
            grid_species.resize(std::max(nt, 1));
            grid_tables.resize((size_t)grid_n[0] * grid_n[1] * grid_n[2] * (size_t)nt);
            ag::check(aleator_grid_cache_load(k.directory.c_str(), &k.key, grid_n, grid_species.data(), &nt,
                                              grid_tables.data(), grid_tables.size(), &hit));
            grid_species.resize(nt);
        } else {
            ag::check(aleator_gpu_grid_point_counts(&k.lat, *cfg.gcmc.grid_spacing, grid_n));
            grid_species = {p.adsorbate_species};
            grid_tables.resize((size_t)grid_n[0] * grid_n[1] * grid_n[2]);
            ag::check(aleator_gpu_grid_build(d.fw, nullptr, grid_species.data(), 1, 0, grid_n, ALEATOR_PRECISION_FP64,
                                             1.0e5, grid_tables.data()));
            std::string path(4096, '\0');
            ag::check(aleator_grid_cache_file_path(k.directory.c_str(), &k.key, path.data(), path.size()));
            ag::check(aleator_grid_cache_serialize(path.c_str(), &k.key, grid_n, grid_species.data(), 1,
                                                   grid_tables.data()));
        }
        grid_hit = hit != 0;
        std::cerr << "energy grid " << (hit ? "loaded from cache" : "built and cached") << " in " << std::fixed
                  << std::setprecision(3)
                  << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " s\n"
                  << std::defaultfloat << std::setprecision(6);
    }

    fs::create_directories(cfg.run.output_directory);
    {
        const fs::path out = cfg.run.output_directory / "resolved_config.json";
        std::ofstream f(out);
        if (!f) throw std::runtime_error("could not open " + out.string() + " for writing");
        f << resolved_json(cfg, p.supercell, grid_hit);
        std::cerr << "wrote resolved configuration (incl. run.rng_seed=" << cfg.run.rng_seed << ") to " << out.string()
                  << "\n";
    }

    const aleator_molecule_site site{0.0, 0.0, 0.0, p.adsorbate_species, cfg.gcmc.adsorbate.charge};
    const aleator_gcmc_chain chain{p.fugacity, cfg.gcmc.temperature, cfg.run.rng_seed, 0};
    aleator_gcmc* g = nullptr;
    ag::check(aleator_gpu_gcmc_create(d.fw, &site, 1, &chain, 1, 1.5, 0.5, max_molecules, &g));
    struct Free {
        aleator_gcmc* g;
        ~Free() { aleator_gpu_gcmc_destroy(g); }
    } free_g{g};
    // a cached grid without the adsorbate's table fails here (the reference's interpolate throws
    // std::out_of_range at the first step)
    if (cfg.gcmc.grid_spacing)
        ag::check(aleator_gpu_gcmc_set_energy_grid_host(g, grid_n, grid_species.data(), (std::int32_t)grid_species.size(),
                                                        grid_tables.data()));
    aleator_gcmc_stats st{};
    auto check_chain = [&] {
        ag::check(aleator_gpu_gcmc_stats(g, &st));
        if (st.error)
            throw std::runtime_error("the device chain stopped: more than " + std::to_string(max_molecules) +
                                     " molecules (raise --max-molecules)");
    };
    std::cerr << "equilibrating (" << cfg.gcmc.equilibration_steps << " steps)...\n";
    if (cfg.gcmc.equilibration_steps) ag::check(aleator_gpu_gcmc_run(g, cfg.gcmc.equilibration_steps, nullptr));
    check_chain();
    ag::check(aleator_gpu_gcmc_reset_stats(g));
    std::cerr << "production (" << cfg.gcmc.production_steps << " steps)...\n";
    // <N> after every production step (cli/main.cc:538-563): the chain sums its count each step
    const std::size_t per_update = std::max<std::size_t>(1, cfg.gcmc.production_steps / 10);
    std::size_t done = 0;
    while (done < cfg.gcmc.production_steps) {
        const std::size_t batch = std::min(per_update, cfg.gcmc.production_steps - done);
        ag::check(aleator_gpu_gcmc_run(g, batch, nullptr));
        done += batch;
        check_chain();
        char line[160];
        std::snprintf(line, sizeof(line), "  [%3.0f%%] step %zu/%zu  <N> so far = %.3f  molecules = %d\n",
                      100.0 * (double)done / (double)cfg.gcmc.production_steps, done, cfg.gcmc.production_steps,
                      st.sum_molecules / (double)done, st.n_molecules);
        std::cerr << line;
    }
    const double mean_n = st.sum_molecules / (double)cfg.gcmc.production_steps;
    const double per_a3 = mean_n / volume;
    const double mmol_g = mass_kg > 0.0 ? (mean_n / kAvogadro) / mass_kg : 0.0;
    if (json) {
        std::cout << "{\"mean_molecule_count\":" << mean_n << ",\"loading_molecules_per_angstrom3\":" << per_a3
                  << ",\"loading_mmol_per_g\":" << mmol_g << ",\"cell_volume_angstrom3\":" << volume << "}\n";
    } else {
        std::cout << "Final result:\n"
                  << "  <N> (mean molecule count over production) = " << mean_n << "\n"
                  << "  loading                                   = " << per_a3 << " molecules/Ang^3\n"
                  << "                                             = " << mmol_g << " mmol/g\n";
    }
    return kExitSuccess;
}

// Widom K_H of the configured adsorbate in the configured framework at the configured
// temperature: Philox4x32Rng(run.rng_seed, stream), trials from call index 0.
int widom_run(const fs::path& path, std::uint64_t trials, std::uint64_t stream, bool json) {
    const Config cfg = load_gcmc_config(path);
    const Prepared p = prepare(cfg);
    DeviceRun d(cfg, p);
    const aleator_molecule_site site{0.0, 0.0, 0.0, p.adsorbate_species, cfg.gcmc.adsorbate.charge};
    double sum = 0.0;
    ag::check(aleator_gpu_widom(d.fw, &site, 1, cfg.run.rng_seed, stream, 0, trials, cfg.gcmc.temperature,
                                ALEATOR_PRECISION_FP64, nullptr, &sum));
    const double kh = sum / ((double)trials * cfg.gcmc.temperature);
    // Henry regime: <N> = K_H * f * V (fugacity in K / A^3)
    const double volume = cell_volume(p.structure.lattice);
    const double n_henry = kh * p.fugacity * kPascalToInternal * volume;
    if (json) {
        std::cout << std::setprecision(17) << "{\"henry_coefficient_per_kelvin\":" << kh << ",\"trials\":" << trials
                  << ",\"rng_seed\":" << cfg.run.rng_seed << ",\"stream\":" << stream
                  << ",\"fugacity_pa\":" << p.fugacity << ",\"cell_volume_angstrom3\":" << volume
                  << ",\"henry_regime_mean_molecule_count\":" << n_henry << "}\n";
    } else {
        std::cout << "Widom insertion Henry coefficient K_H = " << std::scientific << std::setprecision(6) << kh
                  << " (1/K) over " << std::defaultfloat << trials << " insertions\n"
                  << "  Henry-regime <N> at " << cfg.gcmc.pressure_bar << " bar = " << n_henry << "\n";
    }
    return kExitSuccess;
}

int validate(const fs::path& path) {
    switch (detect_kind(path)) {
        case Kind::Gcmc: {
            const Config cfg = load_gcmc_config(path);
            const Prepared p = prepare(cfg);
            std::cout << path.string() << ": valid GCMC config (\"" << cfg.run.name << "\", "
                      << p.structure.particles.size() << " framework atoms, adsorbate=\"" << cfg.gcmc.adsorbate.name
                      << "\")\n";
            return kExitSuccess;
        }
        case Kind::Pore:
            std::cerr << "error: " << path.string() << " is a [pore] config; pore analysis is not part of this build\n";
            return kExitNotInThisBuild;
        default:
            std::cerr << "error: " << path.string() << " must have exactly one of [gcmc] or [pore] at the top level\n";
            return kExitConfigError;
    }
}

void usage() {
    std::cout << "aleator-hip " << kVersion << "\n\n"
              << "Usage:\n"
              << "  aleator-hip --version\n"
              << "  aleator-hip --help\n"
              << "  aleator-hip gcmc run <config.toml> [--dry-run] [--json] [--max-molecules N]\n"
              << "  aleator-hip widom run <config.toml> [--trials N] [--stream S] [--json]\n"
              << "  aleator-hip validate <config.toml>\n\n"
              << "Runs the configured GCMC chain (or Widom insertions) on one MI355X. A real `gcmc run`\n"
              << "writes <run.output_directory>/resolved_config.json. Exit codes: 0 success, 1\n"
              << "configuration/usage error, 2 a valid config for an engine this build does not have.\n";
}

struct Args {
    std::vector<std::string> positional;
    bool dry_run = false, json = false;
    std::uint64_t trials = 500000, stream = 0;
    int max_molecules = 4096;
};

Args parse_args(const std::vector<std::string>& a, size_t from) {
    Args r;
    for (size_t i = from; i < a.size(); ++i) {
        auto number = [&](const char* flag) -> std::uint64_t {
            if (i + 1 >= a.size()) throw std::invalid_argument(std::string(flag) + " needs a value");
            char* end = nullptr;
            const unsigned long long v = std::strtoull(a[++i].c_str(), &end, 10);
            if (*end || a[i].empty() || a[i][0] == '-') throw std::invalid_argument(std::string(flag) + " needs a non-negative integer");
            return v;
        };
        if (a[i] == "--dry-run") r.dry_run = true;
        else if (a[i] == "--json") r.json = true;
        else if (a[i] == "--trials") r.trials = number("--trials");
        else if (a[i] == "--stream") r.stream = number("--stream");
        else if (a[i] == "--max-molecules") r.max_molecules = (int)std::min<std::uint64_t>(number("--max-molecules"), 1u << 24);
        else r.positional.push_back(a[i]);
    }
    return r;
}

}  // namespace

int main(int argc, char** argv) {
    const std::vector<std::string> a(argv + 1, argv + argc);
    if (a.empty() || a[0] == "--help" || a[0] == "-h") {
        usage();
        return a.empty() ? kExitConfigError : kExitSuccess;
    }
    if (a[0] == "--version") {
        std::cout << kVersion << "\n";
        return kExitSuccess;
    }
    try {
        if (a[0] == "gcmc" || a[0] == "widom") {
            const Args r = parse_args(a, 2);
            if (a.size() < 2 || a[1] != "run" || r.positional.size() != 1) {
                std::cerr << "error: usage: aleator-hip " << a[0] << " run <config.toml> [flags]\n";
                return kExitConfigError;
            }
            if (a[0] == "gcmc") return gcmc_run(r.positional[0], r.dry_run, r.json, r.max_molecules);
            if (r.trials == 0) throw std::invalid_argument("--trials must be positive");
            return widom_run(r.positional[0], r.trials, r.stream, r.json);
        }
        if (a[0] == "validate") {
            if (a.size() != 2) {
                std::cerr << "error: usage: aleator-hip validate <config.toml>\n";
                return kExitConfigError;
            }
            return validate(a[1]);
        }
        if (a[0] == "pore" || a[0] == "bench") {
            std::cerr << "error: \"" << a[0] << "\" is not part of this build (it runs the device path only; "
                      << "bench.py measures it)\n";
            return kExitNotInThisBuild;
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return kExitConfigError;
    }
    std::cerr << "error: unrecognized command \"" << a[0] << "\"\n";
    usage();
    return kExitConfigError;
}
