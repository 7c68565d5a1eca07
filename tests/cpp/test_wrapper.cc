// Tests of the C++23 host layer (include/aleator_gpu.hpp), written like the reference's own
// tests: the reference's value types are stood in for by plain aggregates with the same
// members, and every device result is checked against the oracle (oracle/oracle.h).
//
//   test_wrapper cpu   checks that need no device (exception mapping of the C ABI, cache keys)
//   test_wrapper gpu   the device checks: grid build and interpolate, cache round trip,
//                      EwaldIncrementalState sequence, Widom K_H
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "aleator_gpu.hpp"
extern "C" {
#include "oracle.h"
}

namespace {

// ---- stand-ins with the reference's member names -------------------------------------
struct Lattice {
    std::array<std::array<double, 3>, 3> m;
    const std::array<std::array<double, 3>, 3>& matrix() const { return m; }
};
struct ParticleData {
    std::vector<double> x, y, z, charge;
    std::vector<std::uint32_t> species;
};
struct LennardJonesParameters {
    double epsilon, sigma;
};
struct LennardJones {
    std::vector<LennardJonesParameters> params;
    double rc;
    int trunc = 0, mixing = 0;
    const std::vector<LennardJonesParameters>& perSpecies() const { return params; }
    double cutoff() const { return rc; }
    int truncation() const { return trunc; }
    int mixingRule() const { return mixing; }
};
struct Ewald {
    double a, rc;
    int km;
    double alpha() const { return a; }
    double realSpaceCutoff() const { return rc; }
    int kMax() const { return km; }
};
struct ChargedSite {
    double x, y, z, charge;
};
struct MoleculeSite {
    double x, y, z;
    std::uint32_t ljSpecies;
    double charge, mass;
};

int failures = 0;
void expect(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what.c_str());
    if (!ok) ++failures;
}
template <class E>
bool throws(const std::function<void()>& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}
double rel(double a, double b, double floor = 1.0) { return std::abs(a - b) / std::max(std::abs(b), floor); }

// a synthetic framework from the oracle's generator (424-atom c0 types)
struct System {
    Lattice lat;
    ParticleData p;
    LennardJones lj;
};
System make_system(double a, size_t n, bool charged, double rc) {
    System s;
    s.lat.m = {{{a, 0, 0}, {0, a, 0}, {0, 0, a}}};
    std::vector<double> x(n), y(n), z(n), q(n);
    std::vector<uint32_t> sp(n);
    const double m[3][3] = {{a, 0, 0}, {0, a, 0}, {0, 0, a}};
    const double prop[4] = {0.45, 0.30, 0.17, 0.08};
    orc_synth_framework(m, n, 1, 0, 1.0, prop, 4, charged ? 0.5 : 0.0, 2, x.data(), y.data(), z.data(), q.data(),
                        sp.data());
    s.p = {x, y, z, charged ? q : std::vector<double>{}, sp};
    s.lj.params = {{52.8, 3.43}, {30.2, 3.12}, {22.1, 2.57}, {62.4, 2.46}, {148.0, 3.73}, {27.0, 2.80}, {79.0, 3.05}};
    s.lj.rc = rc;
    return s;
}
orc_lattice olat(const Lattice& l) {
    orc_lattice o;
    double m[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = l.m[i][j];
    orc_lattice_init(&o, m);
    return o;
}
orc_particles opart(System& s) {
    static std::vector<double> zero;
    zero.assign(s.p.x.size(), 0.0);
    return {s.p.x.size(), s.p.x.data(), s.p.y.data(), s.p.z.data(), s.p.charge.empty() ? zero.data() : s.p.charge.data(),
            s.p.species.data()};
}
orc_lj olj(const System& s, std::vector<double>& e, std::vector<double>& g) {
    e.clear();
    g.clear();
    for (auto& p : s.lj.params) {
        e.push_back(p.epsilon);
        g.push_back(p.sigma);
    }
    return {e.data(), g.data(), (uint32_t)e.size(), s.lj.rc, s.lj.trunc, s.lj.mixing};
}

std::filesystem::path g_cif_dir;  // tests/golden/cif, found from the program's own path

void cif_checks() {
    namespace ag = aleator::gpu;
    const ag::StructureData s = ag::readCif(g_cif_dir / "IRMOF-1.cif");
    expect(s.particles.size() == 424 && s.occupancy.size() == 424 && s.sourceLabel.size() == 424,
           "readCif: IRMOF-1 expands to 424 atoms");
    expect((s.speciesSymbols == std::vector<std::string>{"Zn", "O", "C", "H"}), "readCif: species by first appearance");
    expect(s.lattice.matrix()[0][0] == 25.832 && s.lattice.matrix()[1][1] == 25.832, "readCif: cubic cell");
    double mass = 0.0;
    for (double m : s.particles.mass) mass += m;
    expect(std::abs(mass - 6158.936) < 1e-9, "readCif: framework mass");
    expect(throws<ag::CifParseError>([] { (void)ag::readCif(g_cif_dir / "no_such_file.cif"); }),
           "readCif: a missing file raises CifParseError");
    expect(throws<ag::CifParseError>([] { (void)ag::parseCif("data_X\n_cell_length_a 1\n"); }),
           "parseCif: missing cell tags raise CifParseError");
    const ag::StructureData t = ag::parseCif(
        "data_T\n_cell_length_a 10\n_cell_length_b 10\n_cell_length_c 10\n_cell_angle_alpha 90\n"
        "_cell_angle_beta 90\n_cell_angle_gamma 90\nloop_\n_symmetry_equiv_pos_as_xyz\n'x,y,z'\n'-x,-y,-z'\n"
        "loop_\n_atom_site_label\n_atom_site_fract_x\n_atom_site_fract_y\n_atom_site_fract_z\n_atom_site_charge\n"
        "Na1 0.1 0.2 0.3 0.7\nCl1 0.4 0.5 0.6 -0.7\n");
    expect(t.particles.size() == 4 && t.particles.charge[1] == 0.7 && t.particles.charge[3] == -0.7 &&
               t.particles.x[1] == 9.0 && t.sourceLabel[2] == "Cl1",
           "parseCif: charges and labels follow every symmetry image");
}

void cpu_checks() {
    namespace ag = aleator::gpu;
    cif_checks();
    // exception mapping: each status code becomes the reference's type
    expect(throws<std::invalid_argument>([] { ag::check(ALEATOR_INVALID_ARGUMENT); }), "status 1 -> invalid_argument");
    expect(throws<std::out_of_range>([] { ag::check(ALEATOR_OUT_OF_RANGE); }), "status 2 -> out_of_range");
    expect(throws<std::runtime_error>([] { ag::check(ALEATOR_RUNTIME_ERROR); }), "status 3 -> runtime_error");
    expect(throws<ag::NotImplemented>([] { ag::check(ALEATOR_NOT_IMPLEMENTED); }), "status 4 -> NotImplemented");
    // a C-ABI error surfaces with its message
    System s = make_system(12.0, 40, false, 5.0);
    const ag::FrameworkEnergyGridCacheKey k1(s.lj, s.p, s.lat, {4, 4, 1}, 0.5);
    const ag::FrameworkEnergyGridCacheKey k2(s.lj, s.p, s.lat, {1, 4}, 0.5);
    expect(k1.hash() == k2.hash(), "cache key: guest ids sorted and deduplicated");
    const ag::FrameworkEnergyGridCacheKey k3(s.lj, s.p, s.lat, {1, 4}, 0.5000001);
    expect(k1.hash() != k3.hash(), "cache key: spacing is part of the key");
    const auto path = ag::FrameworkEnergyGrid::cacheFilePath("/tmp/c", k1);
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)k1.hash());
    expect(path.filename().string() == std::string("framework_energy_grid_v1_") + hex + ".bin", "cache file name");
}

void gpu_checks() {
    namespace ag = aleator::gpu;
    ag::Device dev(0);
    // ---- FrameworkEnergyGrid: device build equals the oracle restatement
    {
        System s = make_system(14.0, 80, false, 6.5);
        s.lj.trunc = ORC_SHIFTED;
        ag::FrameworkEnergyGrid grid(dev, s.lj, s.p, s.lat, {4, 4}, 0.5);
        const auto n = grid.pointCounts();
        orc_lattice L = olat(s.lat);
        int32_t counts[3] = {n[0], n[1], n[2]};
        std::vector<double> e, g;
        const orc_lj lj = olj(s, e, g);
        orc_particles p = opart(s);
        const uint32_t guest = 4;
        std::vector<double> ref((size_t)n[0] * n[1] * n[2]);
        orc_grid_build(&lj, &p, &L, &guest, 1, counts, 1e5, ref.data());
        double worst = 0.0;
        for (size_t i = 0; i < ref.size(); ++i) worst = std::max(worst, rel(grid.table(4)[i], ref[i]));
        expect(worst <= 1e-10, "FrameworkEnergyGrid FP64 equals the oracle (max rel " + std::to_string(worst) + ")");
        std::vector<std::array<double, 3>> pos;
        for (int i = 0; i < 200; ++i) pos.push_back({-3.0 + 0.11 * i, 0.07 * i, 20.0 - 0.13 * i});
        const auto got = grid.interpolate(4, pos);
        bool same = true;
        for (size_t i = 0; i < pos.size(); ++i)
            same &= got[i] == orc_grid_interpolate(&L, counts, grid.table(4).data(), pos[i].data());
        expect(same, "interpolate equals the reference arithmetic bit for bit");
        expect(throws<std::out_of_range>([&] { (void)grid.interpolate(5, pos[0]); }),
               "interpolate: unknown guest -> out_of_range");
        expect(throws<std::invalid_argument>([&] { ag::FrameworkEnergyGrid(dev, s.lj, s.p, s.lat, {4}, 0.0); }),
               "spacing <= 0 -> invalid_argument");
        // cache: miss then hit, same tables; a foreign file under the name -> runtime_error
        const auto dir = std::filesystem::temp_directory_path() / "aleator_cpp_cache_test";
        std::filesystem::remove_all(dir);
        const ag::FrameworkEnergyGridCacheRequest req{dir, {4}, 0.5};
        bool hit = true;
        auto a = ag::loadOrBuildFrameworkEnergyGrid(dev, s.lj, s.p, s.lat, req, &hit);
        expect(!hit, "loadOrBuild: first call builds");
        auto b = ag::loadOrBuildFrameworkEnergyGrid(dev, s.lj, s.p, s.lat, req, &hit);
        expect(hit && b.table(4) == a.table(4) && b.pointCounts() == a.pointCounts(), "loadOrBuild: second call loads");
        const ag::FrameworkEnergyGridCacheKey key(s.lj, s.p, s.lat, {4}, 0.5);
        const auto file = ag::FrameworkEnergyGrid::cacheFilePath(dir, key);
        {
            std::fstream f(file, std::ios::in | std::ios::out | std::ios::binary);
            f.seekp(4 + 8);  // first byte of the first atom's x
            f.put((char)0x5a);
        }
        expect(throws<std::runtime_error>([&] { (void)ag::FrameworkEnergyGrid::loadFromCache(dev, dir, key); }),
               "cache: stored key differs -> runtime_error");
        std::filesystem::remove_all(dir);
    }
    // ---- EwaldIncrementalState: propose / commit / committedEnergy against the oracle
    {
        System s = make_system(16.0, 60, true, 7.5);
        const Ewald ew{0.35, 7.5, 5};
        ag::EwaldIncrementalState st(dev, ew, s.p, s.lat);
        orc_lattice L = olat(s.lat);
        orc_particles p = opart(s);
        orc_ewald_state ost;
        orc_ewald_state_init(&ost, 0.35, 5, &p, &L);
        expect(rel(st.committedEnergy(), orc_ewald_state_committed(&ost)) <= 1e-10, "committedEnergy equals the oracle");
        const std::vector<ChargedSite> co2 = {{3.0, 4.0, 4.84, -0.35}, {3.0, 4.0, 6.0, 0.70}, {3.0, 4.0, 7.16, -0.35}};
        std::vector<orc_site> oco2;
        for (auto& c : co2) oco2.push_back({c.x, c.y, c.z, c.charge});
        const double d = st.proposeInsertion(co2);
        const double od = orc_ewald_state_delta(&ost, nullptr, 0, oco2.data(), 3, nullptr, nullptr, nullptr, nullptr, nullptr);
        expect(rel(d, od) <= 1e-10, "proposeInsertion equals the oracle");
        st.commitInsertion(co2);
        orc_ewald_state_commit(&ost, nullptr, 0, oco2.data(), 3);
        std::vector<ChargedSite> moved = co2;
        for (auto& c : moved) c.x += 1.3;
        std::vector<orc_site> omoved;
        for (auto& c : moved) omoved.push_back({c.x, c.y, c.z, c.charge});
        const double dm = st.proposeMove(co2, moved);
        const double odm = orc_ewald_state_delta(&ost, oco2.data(), 3, omoved.data(), 3, nullptr, nullptr, nullptr, nullptr,
                                                 nullptr);
        expect(rel(dm, odm) <= 1e-10, "proposeMove after a commit equals the oracle");
        const auto raw = st.rawStateForTesting();
        std::vector<double> oraw(2 * ost.terms.n + 2);
        orc_ewald_state_raw(&ost, oraw.data());
        double worst = 0.0;
        for (size_t i = 0; i < raw.size() && i < oraw.size(); ++i) worst = std::max(worst, std::abs(raw[i] - oraw[i]));
        expect(raw.size() == oraw.size() && worst <= 1e-9, "rawStateForTesting equals the oracle's layout and values");
        orc_ewald_state_free(&ost);
    }
    // ---- GcmcChains with a FrameworkEnergyGrid against the oracle's grid mode
    {
        System s = make_system(14.0, 40, false, 6.5);
        s.lj.trunc = ORC_SHIFTED;
        ag::Framework fw(dev, s.lj, s.p, s.lat);
        ag::FrameworkEnergyGrid grid(dev, s.lj, s.p, s.lat, {4}, 0.5);
        const std::vector<MoleculeSite> ch4 = {{0, 0, 0, 4, 0.0, 16.04}};
        const std::vector<ag::ChainConfig> cfg = {{1e8, 300.0, 5, 0}, {3e7, 300.0, 6, 2}};
        ag::GcmcChains chains(fw, ch4, cfg, 1.5, 0.5, 256, &grid);
        chains.run(2000);
        const auto st = chains.statistics();
        orc_lattice L = olat(s.lat);
        std::vector<double> e, g;
        const orc_lj lj = olj(s, e, g);
        orc_particles p = opart(s);
        const orc_mol_site site = {0, 0, 0, 4, 0.0};
        const auto n = grid.pointCounts();
        const int32_t counts[3] = {n[0], n[1], n[2]};
        const uint32_t species = 4;
        bool same = true;
        for (size_t c = 0; c < cfg.size(); ++c) {
            orc_gcmc og;
            orc_gcmc_init(&og, &lj, &p, &L, &site, 1, cfg[c].fugacityPascal, cfg[c].temperatureKelvin, 1.5, 0.5, 1.0,
                          1.0, 1, 0, cfg[c].seed, cfg[c].stream);
            orc_gcmc_set_grid(&og, counts, &species, 1, grid.table(4).data());
            std::vector<double> trace(2000);
            orc_gcmc_run(&og, 2000, trace.data());
            double sum = 0.0;
            for (double v : trace) sum += v;
            const auto pos = chains.sitePositions((int)c);
            same &= st[c].n_molecules == (int)og.n_mol && pos.size() == og.n_mol && st[c].sum_molecules == sum &&
                    og.n_mol > 0;
            for (size_t m = 0; m < og.n_mol && same; ++m) {
                const size_t i = og.mol[m];
                same &= std::abs(pos[m][0] - og.x[i]) <= 1e-9 && std::abs(pos[m][1] - og.y[i]) <= 1e-9 &&
                        std::abs(pos[m][2] - og.z[i]) <= 1e-9;
            }
            for (int k = 0; k < 4; ++k) same &= st[c].attempted[k] == og.attempted[k] && st[c].accepted[k] == og.accepted[k];
            orc_gcmc_free(&og);
        }
        expect(same, "GcmcChains with a FrameworkEnergyGrid follows the oracle's grid-mode trajectory");
        // a chain saved and loaded into a new chain set takes the same steps; uncharged chains never resync
        const auto rec = chains.saveChain(1);
        ag::GcmcChains copy(fw, ch4, {cfg[1]}, 1.5, 0.5, 256, &grid);
        copy.loadChain(0, rec);
        chains.run(500);
        copy.run(500);
        const auto a1 = chains.sitePositions(1), b1 = copy.sitePositions(0);
        bool resumed = a1.size() == b1.size() && !a1.empty();
        for (size_t m = 0; m < a1.size() && resumed; ++m) resumed &= a1[m] == b1[m];
        expect(resumed, "GcmcChains::saveChain / loadChain: the loaded chain continues step for step");
        const auto drift = chains.maxResyncDrift();
        expect(drift.size() == cfg.size() && drift[0] == 0.0 && drift[1] == 0.0,
               "GcmcChains::maxResyncDrift: uncharged chains report no drift");
        expect(throws<std::invalid_argument>([&] { copy.loadChain(0, std::vector<unsigned char>(rec.begin(), rec.end() - 8)); }),
               "GcmcChains::loadChain: a record of the wrong size -> invalid_argument");
        expect(fw.nonfiniteCount() == 0, "Framework::nonfiniteCount: a clean framework's grid writes no NaN / Inf");
        ag::FrameworkEnergyGrid other(dev, s.lj, s.p, s.lat, {5}, 0.5);
        expect(throws<std::out_of_range>([&] { ag::GcmcChains(fw, ch4, cfg, 1.5, 0.5, 256, &other); }),
               "GcmcChains: grid without the site's species -> out_of_range");
    }
    // ---- Widom K_H against the oracle
    {
        System s = make_system(25.83, 424, false, 12.8);
        ag::Framework fw(dev, s.lj, s.p, s.lat);
        const std::vector<MoleculeSite> ch4 = {{0, 0, 0, 4, 0.0, 16.04}};
        const double kh = ag::widomInsertionHenryCoefficient(fw, ch4, 20000, 298.0, 1, 1);
        orc_lattice L = olat(s.lat);
        std::vector<double> e, g;
        const orc_lj lj = olj(s, e, g);
        orc_particles p = opart(s);
        const orc_mol_site site = {0, 0, 0, 4, 0.0};
        const double okh = orc_widom(&lj, &p, &L, &site, 1, nullptr, 0.0, 1, 1, 0, 20000, 298.0, nullptr);
        expect(rel(kh, okh, 0.0) <= 1e-10, "widomInsertionHenryCoefficient equals the oracle");
    }
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    g_cif_dir = std::filesystem::absolute(argv[0]).parent_path() / ".." / ".." / "golden" / "cif";
    try {
        cpu_checks();
        if (mode == "gpu") gpu_checks();
    } catch (const std::exception& e) {
        std::printf("FAIL uncaught exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
    return failures == 0 ? 0 : 1;
}
