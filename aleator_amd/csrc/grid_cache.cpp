// grid_cache.cpp -- the energy-grid cache file, format v1, read and written byte for byte as
// the reference does (engines/monte_carlo/framework_energy_grid.cc:31-191 key and helpers,
// :286-313 serialize, :315-322 cacheFilePath, :357-408 loadFromCache), so that grids built
// on the device load in the reference's engine and reference-built files load here.
//
// Layout, little-endian: u32 version | key | 3 x i32 point counts | u64 n_tables |
// n_tables x (u32 species, u64 len, f64[len]).  Key: u64 n + f64[n] for each of atom x, y, z;
// u64 n + u32[n] species; 9 x f64 lattice rows; u64 m + m x (f64 epsilon, f64 sigma); f64
// cutoff; u8 truncation; u64 k + u32[k] sorted unique guest species; f64 spacing; f64 cap.
// The file name is framework_energy_grid_v1_<FNV-1a 64 of the key bytes, 16 hex>.bin.
// The reference's writer emits tables in its hash map's order; this one emits them in
// ascending species order, and the reader accepts any order.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "aleator_gpu.h"
#include "internal.hpp"

namespace aleator_gpu {
namespace {

constexpr uint32_t kCacheVersion = 1;
constexpr double kEnergyCap = 1.0e5;  // FrameworkEnergyGrid::kEnergyCapKelvin

template <typename T>
void put(std::string& b, const T& v) {
    b.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

std::vector<uint32_t> sorted_guests(const aleator_grid_cache_key& k) {
    std::vector<uint32_t> g(k.guest_species, k.guest_species + k.n_guest);
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
    return g;
}

void check_key(const aleator_grid_cache_key* k) {
    require(k && k->lattice && k->atoms && k->lj, ALEATOR_INVALID_ARGUMENT, "null argument");
    require(k->n_guest >= 0 && (k->n_guest == 0 || k->guest_species), ALEATOR_INVALID_ARGUMENT,
            "grid cache: guest species");
    const aleator_particles& p = *k->atoms;
    require(p.n == 0 || (p.x && p.y && p.z && p.species), ALEATOR_INVALID_ARGUMENT, "grid cache: atom arrays");
    require(k->lj->n_species == 0 || (k->lj->epsilon && k->lj->sigma), ALEATOR_INVALID_ARGUMENT,
            "grid cache: LJ arrays");
}

// writeKey (framework_energy_grid.cc:79-101) into a byte string
std::string key_bytes(const aleator_grid_cache_key& k) {
    std::string b;
    const aleator_particles& p = *k.atoms;
    for (const double* v : {p.x, p.y, p.z}) {
        put(b, (uint64_t)p.n);
        if (p.n) b.append(reinterpret_cast<const char*>(v), p.n * sizeof(double));
    }
    put(b, (uint64_t)p.n);
    if (p.n) b.append(reinterpret_cast<const char*>(p.species), p.n * sizeof(uint32_t));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) put(b, k.lattice->m[i][j]);
    put(b, (uint64_t)k.lj->n_species);
    for (uint32_t s = 0; s < k.lj->n_species; ++s) {
        put(b, k.lj->epsilon[s]);
        put(b, k.lj->sigma[s]);
    }
    put(b, k.lj->cutoff);
    put(b, (uint8_t)k.lj->truncation);
    const auto g = sorted_guests(k);
    put(b, (uint64_t)g.size());
    if (!g.empty()) b.append(reinterpret_cast<const char*>(g.data()), g.size() * sizeof(uint32_t));
    put(b, k.spacing);
    put(b, kEnergyCap);
    return b;
}

// fnv1a64 (framework_energy_grid.cc:140-150)
uint64_t fnv1a64(const std::string& s) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (unsigned char c : s) {
        h ^= (uint64_t)c;
        h *= 0x100000001b3ULL;
    }
    return h;
}

std::string file_name(const aleator_grid_cache_key& k) {
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)fnv1a64(key_bytes(k)));
    return "framework_energy_grid_v" + std::to_string(kCacheVersion) + "_" + hex + ".bin";
}

void copy_out(const std::string& s, char* out, uint64_t capacity) {
    require(out != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
    require(capacity > s.size(), ALEATOR_INVALID_ARGUMENT, "grid cache: output buffer too small");
    std::memcpy(out, s.c_str(), s.size() + 1);
}

}  // namespace
}  // namespace aleator_gpu

using namespace aleator_gpu;

extern "C" {

aleator_status aleator_grid_cache_key_hash(const aleator_grid_cache_key* key, uint64_t* hash) {
    return guarded([&] {
        check_key(key);
        require(hash != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        *hash = fnv1a64(key_bytes(*key));
    });
}

aleator_status aleator_grid_cache_file_path(const char* directory, const aleator_grid_cache_key* key, char* out,
                                            uint64_t capacity) {
    return guarded([&] {
        check_key(key);
        require(directory != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        copy_out((std::filesystem::path(directory) / file_name(*key)).string(), out, capacity);
    });
}

aleator_status aleator_grid_cache_serialize(const char* path, const aleator_grid_cache_key* key,
                                            const int32_t n_points[3], const uint32_t* table_species,
                                            int32_t n_tables, const double* tables) {
    return guarded([&] {
        check_key(key);
        require(path && n_points && (n_tables == 0 || (table_species && tables)), ALEATOR_INVALID_ARGUMENT,
                "null argument");
        require(n_tables >= 0, ALEATOR_INVALID_ARGUMENT, "grid cache: negative table count");
        uint64_t total = 1;
        for (int i = 0; i < 3; ++i) {
            require(n_points[i] >= 1, ALEATOR_INVALID_ARGUMENT, "grid cache: point counts must be >= 1");
            total *= (uint64_t)n_points[i];
        }
        std::vector<int> order(n_tables);
        for (int t = 0; t < n_tables; ++t) order[t] = t;
        std::sort(order.begin(), order.end(), [&](int x, int y) { return table_species[x] < table_species[y]; });
        for (int t = 1; t < n_tables; ++t)
            require(table_species[order[t]] != table_species[order[t - 1]], ALEATOR_INVALID_ARGUMENT,
                    "grid cache: duplicate table species");
        const std::filesystem::path p(path);
        if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
        std::ofstream out(p, std::ios::binary | std::ios::trunc);
        require((bool)out, ALEATOR_RUNTIME_ERROR, ("FrameworkEnergyGrid::serialize: could not open " + p.string() +
                                                   " for writing").c_str());
        std::string head;
        put(head, kCacheVersion);
        head += key_bytes(*key);
        for (int i = 0; i < 3; ++i) put(head, n_points[i]);
        put(head, (uint64_t)n_tables);
        out.write(head.data(), (std::streamsize)head.size());
        for (int t : order) {
            std::string th;
            put(th, table_species[t]);
            put(th, total);
            out.write(th.data(), (std::streamsize)th.size());
            out.write(reinterpret_cast<const char*>(tables + (size_t)t * total), (std::streamsize)(total * 8));
        }
        require((bool)out, ALEATOR_RUNTIME_ERROR,
                ("FrameworkEnergyGrid::serialize: write failed for " + p.string()).c_str());
    });
}

aleator_status aleator_grid_cache_load(const char* directory, const aleator_grid_cache_key* key,
                                       int32_t n_points[3], uint32_t* table_species, int32_t* n_tables,
                                       double* tables, uint64_t capacity, int32_t* hit) {
    return guarded([&] {
        check_key(key);
        require(directory && n_points && n_tables && hit, ALEATOR_INVALID_ARGUMENT, "null argument");
        const std::filesystem::path path = std::filesystem::path(directory) / file_name(*key);
        *hit = 0;
        *n_tables = 0;
        if (!std::filesystem::exists(path)) return;
        std::ifstream in(path, std::ios::binary);
        require((bool)in, ALEATOR_RUNTIME_ERROR,
                ("FrameworkEnergyGrid::loadFromCache: could not open " + path.string()).c_str());
        std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        size_t pos = 0;
        auto take = [&](void* dst, size_t n) {
            require(pos + n <= data.size(), ALEATOR_RUNTIME_ERROR,
                    ("FrameworkEnergyGrid cache: " + path.string() + " is truncated or corrupt").c_str());
            std::memcpy(dst, data.data() + pos, n);
            pos += n;
        };
        uint32_t version = 0;
        take(&version, 4);
        require(version == kCacheVersion, ALEATOR_RUNTIME_ERROR,
                ("FrameworkEnergyGrid cache: " + path.string() + " has format version " + std::to_string(version) +
                 ", this build expects " + std::to_string(kCacheVersion) + " -- refusing to reinterpret it")
                    .c_str());
        const std::string kb = key_bytes(*key);
        require(pos + kb.size() <= data.size(), ALEATOR_RUNTIME_ERROR,
                ("FrameworkEnergyGrid cache: " + path.string() + " is truncated or corrupt (key read failed)").c_str());
        // FrameworkEnergyGridCacheKey::operator== compares every field exactly, which on the
        // serialized form is byte equality (the same fields, widths and order)
        require(std::memcmp(data.data() + pos, kb.data(), kb.size()) == 0, ALEATOR_RUNTIME_ERROR,
                ("FrameworkEnergyGrid cache: " + path.string() +
                 " exists and has a matching content-hash filename, but its stored key does not exactly match the "
                 "current inputs -- refusing to silently reuse a possibly-stale or colliding cache entry")
                    .c_str());
        pos += kb.size();
        int32_t counts[3];
        take(counts, sizeof(counts));
        uint64_t ntab = 0;
        take(&ntab, 8);
        uint64_t total = 1;
        for (int i = 0; i < 3; ++i) {
            require(counts[i] >= 1, ALEATOR_RUNTIME_ERROR, "FrameworkEnergyGrid cache: corrupt point counts");
            total *= (uint64_t)counts[i];
        }
        struct Tab {
            uint32_t species;
            size_t offset;
        };
        std::vector<Tab> tabs;
        for (uint64_t t = 0; t < ntab; ++t) {
            Tab tb{};
            uint64_t len = 0;
            take(&tb.species, 4);
            take(&len, 8);
            require(len == total, ALEATOR_RUNTIME_ERROR, "FrameworkEnergyGrid cache: table length mismatch");
            require(pos + len * 8 <= data.size(), ALEATOR_RUNTIME_ERROR,
                    ("FrameworkEnergyGrid cache: " + path.string() + " is truncated or corrupt (grid data read failed)")
                        .c_str());
            tb.offset = pos;
            pos += len * 8;
            tabs.push_back(tb);
        }
        std::sort(tabs.begin(), tabs.end(), [](const Tab& x, const Tab& y) { return x.species < y.species; });
        for (int i = 0; i < 3; ++i) n_points[i] = counts[i];
        *n_tables = (int32_t)tabs.size();
        *hit = 1;
        if (!tables) return;  // size query
        require(table_species != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(capacity >= tabs.size() * total, ALEATOR_INVALID_ARGUMENT, "grid cache: output buffer too small");
        for (size_t t = 0; t < tabs.size(); ++t) {
            table_species[t] = tabs[t].species;
            std::memcpy(tables + t * total, data.data() + tabs[t].offset, total * 8);
        }
    });
}

}  // extern "C"
