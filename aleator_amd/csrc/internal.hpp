// internal.hpp -- shared host/device declarations of the gfx950 hot-path library.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "aleator_gpu.h"

namespace aleator_gpu {

// ---------------------------------------------------------------- errors
// Exceptions inside the library are mapped to aleator_status at the C boundary.
struct Error : std::runtime_error {
    aleator_status status;
    Error(aleator_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(aleator_status s, const std::string& msg) { throw Error(s, msg); }

// Message of the last failure on the calling thread (aleator_gpu_last_error).
std::string& last_error_slot();

// Runs `f`, mapping exceptions onto the status codes of the C boundary.
template <typename F>
aleator_status guarded(F&& f) {
    try {
        f();
        last_error_slot().clear();
        return ALEATOR_OK;
    } catch (const Error& e) {
        last_error_slot() = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        last_error_slot() = "host allocation failed";
        return ALEATOR_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        last_error_slot() = e.what();
        return ALEATOR_RUNTIME_ERROR;
    }
}

inline void require(bool ok, aleator_status s, const char* msg) {
    if (!ok) fail(s, msg);
}

#define ALEATOR_HIP_CHECK(expr)                                                                 \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            ::aleator_gpu::fail(ALEATOR_DEVICE_ERROR,                                           \
                                std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

// ---------------------------------------------------------------- limits
constexpr int kMaxTypes = 32;       // distinct framework LJ species
constexpr int kMaxGuests = 4;       // LJ tables per grid launch
constexpr int kMaxSites = 8;        // sites of a Widom / proposal molecule
constexpr int kMaxK = 16;           // kmax bound for the per-axis phase tables
constexpr int kSegLen = 8;          // k-set run length (reciprocal Widom kernel)
constexpr double kCoulomb = 167100.9468982874;  // ewald.hpp:62
constexpr double kPi = 3.14159265358979323846;
constexpr int kErfPolyMax = 28;     // highest degree of the erf kernel polynomial (pair.hpp)

// Coefficients (lowest first) of erf(alpha r)/r as a polynomial in x = r^2 xs - 1 over
// r in [0, rmax] (erf_poly.cpp).  The degree is the smallest of 18, 24 and 28, at least min_deg,
// whose dropped Chebyshev coefficients of int_0^1 exp(-alpha^2 r^2 t^2) dt sum below tol (28
// when none does); it is returned, and coef receives degree + 1 values.
int erf_kernel_poly(double alpha, double rmax, double tol, int min_deg, double* coef, double* xs);
// The same kernel over r in [rmin, rmax] at a given degree, in x = r^2 xs + xo; true when the
// dropped Chebyshev coefficients sum below tol.
bool erf_kernel_poly_range(double alpha, double rmin, double rmax, double tol, int deg, double* coef, double* xs,
                           double* xo);

// ---------------------------------------------------------------- host lattice
// Row-vector lattice (core/geometry/lattice.hpp:21-30) with the reciprocal-vector
// fractional map of lattice.cc:54-60 precomputed: f = c . inv (inv columns are
// (a_j x a_k)/V).
struct HostLattice {
    double m[3][3];
    double inv[3][3];
    double volume;   // signed
    double lperp[3]; // perpendicular widths (lattice.cc:257-284)
};
HostLattice make_lattice(const double m[3][3]);

// Reference k-set (ewald.cc:137-161) in the reference's order.
struct KSet {
    std::vector<int32_t> n;     // [nk][3]
    std::vector<double> k;      // [nk][3]
    std::vector<double> weight; // [nk]
    size_t size() const { return weight.size(); }
};
KSet make_kset(const HostLattice& lat, double alpha, int kmax);

// ---------------------------------------------------------------- objects
// A context and a framework are reference counted.  The caller's destroy call drops the
// caller's reference, and every object created on a context or a framework holds one more, so
// what a child object uses stays valid until the child itself is destroyed, in whatever order
// the destroy calls come (a garbage collector finalises unreachable objects in no set order).
struct Context {
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    // reusable device scratch
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    void* ensure_scratch(size_t bytes);
    // grid tile shape of the per-call grid entry points (set by a grid plan's autotuning)
    int tile_shape = 0;
    int tile_chunk = 512;  // atoms per list-build pass chosen with tile_shape (if allowed, else 128)
    // optional HIP-event timing of each call's dominant kernel, on `stream`
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    void time_begin();
    void time_end();
};

struct Framework {
    std::atomic<int> refs{1};
    Context* ctx = nullptr;
    HostLattice lat{};
    uint64_t n = 0;
    std::vector<double> x, y, z, q; // cartesian input
    std::vector<double> fx, fy, fz; // fractional, wrapped into [0,1)
    std::vector<uint32_t> species;
    std::vector<int32_t> type;          // compact type index per atom
    std::vector<uint32_t> type_species; // species id of each compact type
    // LJ
    std::vector<double> eps, sigma;
    double rc = 0.0;
    int trunc = 0, mixing = 0;
    // device atoms
    double *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr, *d_q = nullptr;
    int32_t* d_type = nullptr;
    // running count of non-finite grid table values written by this framework's builds
    unsigned long long* d_nonfinite = nullptr;
    // Ewald
    bool charged = false;
    double alpha = 0.0, rc_real = 0.0;
    int kmax = 0;
    KSet kset;
    int nk_half = 0;                 // first nk/2 reference terms (lexicographically n < 0)
    int32_t* d_kn = nullptr;         // [nk_half][4] (n0,n1,n2,pad)
    double* d_kw = nullptr;          // [nk_half]
    // the half set as runs of consecutive n2 at fixed (n0, n1), at most kSegLen terms each:
    // (first term index, (n0+K) | (n1+K) << 8 | (n2 of the first term + K) << 16 | length << 24)
    int2* d_kseg = nullptr;
    int n_kseg = 0;
    double *d_sre = nullptr, *d_sim = nullptr;  // framework S(k), half set
    std::vector<double> h_sre, h_sim;           // host copy of the same
    double self_energy = 0.0;
    // add the analytic LJ tail of one insertion to grid tables and Widom energies (off: the
    // reference's outputs)
    bool lj_tail = false;
    // Widom's LJ constants on the device, kept while the inserted site species stay the same
    double* d_widom_tab = nullptr;
    std::vector<uint32_t> widom_tab_species;
};

void retain(Context* c);
void release(Context* c);  // frees the context with its last reference
void retain(Framework* f);
void release(Framework* f);  // frees the framework with its last reference, then drops its context

struct EwaldState {
    Framework* fw = nullptr;
    double *d_sre = nullptr, *d_sim = nullptr; // committed S(k), half set
    double self_energy = 0.0, exclusion_energy = 0.0;
    double* d_out = nullptr;                   // propose result scratch
};

// ---------------------------------------------------------------- device helpers
struct LjPairTable {
    // per (slot, type): sigma^2, 4 eps, V(rc), V'(rc)
    double sig2[kMaxGuests > kMaxSites ? kMaxGuests : kMaxSites][kMaxTypes];
    double eps4[kMaxGuests > kMaxSites ? kMaxGuests : kMaxSites][kMaxTypes];
    double vrc[kMaxGuests > kMaxSites ? kMaxGuests : kMaxSites][kMaxTypes];
    double dvrc[kMaxGuests > kMaxSites ? kMaxGuests : kMaxSites][kMaxTypes];
};

// LennardJones::tailEnergyCorrection (lennard_jones.cc:218-257) of the framework plus one
// molecule of these site species minus that of the framework alone
double lj_tail_delta(const Framework& fw, const uint32_t* site_species, int n_sites);

// mixed parameters and cutoff constants of one (guest species, framework type) pair
// (lennard_jones.cc:15-30, 49-68)
void lj_pair_constants(const Framework& fw, uint32_t guest, uint32_t fw_species, double& sig2,
                       double& eps4, double& vrc, double& dvrc);

// ---------------------------------------------------------------- launchers
struct GridRequest {
    const double* grid_m;  // 3x3 row-major or nullptr
    std::vector<uint32_t> guests;
    bool coulomb = false;
    int32_t n[3];
    int precision = 0;
    double cap = 1e5;
    double* out_device = nullptr;
};
// no_alloc: use the library's fixed device scratch array instead of the context's growable buffer
// (no device allocation; requests larger than kNoAllocScratchBytes fail with invalid argument)
void launch_grid(Framework& fw, const GridRequest& req, bool no_alloc = false);
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;

struct WidomRequest {
    std::vector<aleator_molecule_site> sites;
    uint64_t seed, stream, first_call, n_trials;
    double temperature;
    int precision;
    double* du_device;  // may be null
    double* sum_device;
};
void launch_widom(Framework& fw, const WidomRequest& req);

// structure factor of arbitrary fractional positions over the framework half k-set
void structure_factor(Framework& fw, uint64_t n, const double* d_fx, const double* d_fy,
                      const double* d_fz, const double* d_q, double* d_sre, double* d_sim);

// sum_k w (|S+dS|^2 - |S|^2) over the half set for `n_prop` proposals, each removing
// n_rem and adding n_add sites (fractional coordinates + charge, device buffer
// [n_prop][n_rem+n_add][4]); result (times 2 for the mirror half) in d_out[n_prop].
// When `commit` is set (n_prop == 1), dS is added to S.
void reciprocal_delta(Framework& fw, double* d_sre, double* d_sim, uint64_t n_prop, int n_rem,
                      int n_add, const double* d_sites, double* d_out, bool commit);

} // namespace aleator_gpu
