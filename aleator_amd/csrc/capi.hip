// capi.hip -- the extern "C" boundary of include/aleator_gpu.h.
//
// Host-side responsibilities: argument validation with the reference's error
// semantics, the reference k-set and LJ mixing tables (FP64, the reference's own
// arithmetic), fractional coordinates, device residency of the framework and of
// each S(k) state, and the self/exclusion terms of the Ewald state, which are
// O(sites) host arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "aleator_gpu.h"
#include "internal.hpp"

// Host arithmetic below keeps the reference's operation order: no fused multiply-adds.
#pragma clang fp contract(off)

struct aleator_ctx : aleator_gpu::Context {};
struct aleator_framework : aleator_gpu::Framework {};
struct aleator_ewald_state : aleator_gpu::EwaldState {};

namespace aleator_gpu {

std::string& last_error_slot() {
    static thread_local std::string slot;
    return slot;
}

namespace {

struct Vec {
    double v[3];
};
Vec cross(const double* a, const double* b) {
    return {{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}
double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// rawPotential / rawDerivative (lennard_jones.cc:15-30)
double raw_potential(double r, double eps, double sigma) {
    const double sr = sigma / r;
    const double sr6 = sr * sr * sr * sr * sr * sr;
    const double sr12 = sr6 * sr6;
    return 4.0 * eps * (sr12 - sr6);
}
double raw_derivative(double r, double eps, double sigma) {
    const double sr = sigma / r;
    const double sr6 = sr * sr * sr * sr * sr * sr;
    const double sr12 = sr6 * sr6;
    return (24.0 * eps / r) * (sr6 - 2.0 * sr12);
}

template <typename T>
T* device_upload(const T* host, size_t n, hipStream_t s) {
    T* d = nullptr;
    ALEATOR_HIP_CHECK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) ALEATOR_HIP_CHECK(hipMemcpyAsync(d, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

void frac_of(const HostLattice& lat, const double c[3], double f[3]) {
    for (int j = 0; j < 3; ++j) f[j] = c[0] * lat.inv[0][j] + c[1] * lat.inv[1][j] + c[2] * lat.inv[2][j];
}

// minimum image by fractional rounding: exact for separations below L_perp/2
double min_image_distance(const HostLattice& lat, const double a[3], const double b[3]) {
    double fa[3], fb[3], df[3], d[3];
    frac_of(lat, a, fa);
    frac_of(lat, b, fb);
    for (int j = 0; j < 3; ++j) df[j] = fb[j] - fa[j] - std::rint(fb[j] - fa[j]);
    for (int j = 0; j < 3; ++j) d[j] = df[0] * lat.m[0][j] + df[1] * lat.m[1][j] + df[2] * lat.m[2][j];
    return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// erfOverR (ewald_incremental_state.cc:24-30)
double erf_over_r(double r, double alpha) {
    if (r < 1e-8) return 2.0 * alpha / std::sqrt(kPi);
    return std::erf(alpha * r) / r;
}

// pairExclusionSum (ewald_incremental_state.cc:32-43)
double pair_exclusion_sum(const HostLattice& lat, const aleator_charged_site* s, int n, double alpha) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double a[3] = {s[i].x, s[i].y, s[i].z}, b[3] = {s[j].x, s[j].y, s[j].z};
            sum += s[i].charge * s[j].charge * erf_over_r(min_image_distance(lat, a, b), alpha);
        }
    return sum;
}

double sum_q2(const aleator_charged_site* s, int n) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += s[i].charge * s[i].charge;
    return sum;
}

void check_lattice(const aleator_lattice* l) {
    require(l != nullptr, ALEATOR_INVALID_ARGUMENT, "lattice must not be null");
    const HostLattice h = make_lattice(l->m);
    require(std::abs(h.volume) >= 1e-10, ALEATOR_INVALID_ARGUMENT, "degenerate lattice (volume ~ 0)");
}

// fractional coordinates + charges of sites, [n][4]
std::vector<double> site4_of(const HostLattice& lat, const aleator_charged_site* s, int n) {
    std::vector<double> out(4 * (size_t)std::max(n, 0));
    for (int i = 0; i < n; ++i) {
        const double c[3] = {s[i].x, s[i].y, s[i].z};
        frac_of(lat, c, &out[4 * i]);
        out[4 * i + 3] = s[i].charge;
    }
    return out;
}

// full reference-order S(k) from the device half set
void full_structure_factor(const Framework& fw, const double* d_sre, const double* d_sim,
                           std::vector<double>& re, std::vector<double>& im) {
    std::vector<double> hr(fw.nk_half), hi(fw.nk_half);
    ALEATOR_HIP_CHECK(hipMemcpyAsync(hr.data(), d_sre, fw.nk_half * sizeof(double), hipMemcpyDeviceToHost,
                                     fw.ctx->stream));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(hi.data(), d_sim, fw.nk_half * sizeof(double), hipMemcpyDeviceToHost,
                                     fw.ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
    const size_t nk = fw.kset.size();
    re.resize(nk);
    im.resize(nk);
    for (size_t t = 0; t < nk; ++t) {
        const bool low = (int)t < fw.nk_half;
        const size_t h = low ? t : nk - 1 - t;
        re[t] = hr[h];
        im[t] = low ? hi[h] : -hi[h];
    }
}

double reciprocal_energy(const Framework& fw, const std::vector<double>& re, const std::vector<double>& im) {
    double sum = 0.0;
    for (size_t t = 0; t < fw.kset.size(); ++t) sum += fw.kset.weight[t] * (re[t] * re[t] + im[t] * im[t]);
    return sum * kCoulomb * (2.0 * kPi / std::abs(fw.lat.volume));
}

struct DeviceParticles {
    double *fx = nullptr, *fy = nullptr, *fz = nullptr, *q = nullptr;
    uint64_t n = 0;
    std::vector<double> hq;
    ~DeviceParticles() {
        (void)hipFree(fx);
        (void)hipFree(fy);
        (void)hipFree(fz);
        (void)hipFree(q);
    }
};

void upload_particles(const Framework& fw, const aleator_particles* p, DeviceParticles& d) {
    require(p != nullptr && p->x && p->y && p->z, ALEATOR_INVALID_ARGUMENT, "particles must not be null");
    std::vector<double> fx(p->n), fy(p->n), fz(p->n);
    d.hq.assign(p->n, 0.0);
    for (uint64_t i = 0; i < p->n; ++i) {
        const double c[3] = {p->x[i], p->y[i], p->z[i]};
        double f[3];
        frac_of(fw.lat, c, f);
        fx[i] = f[0];
        fy[i] = f[1];
        fz[i] = f[2];
        if (p->charge) d.hq[i] = p->charge[i];
    }
    d.n = p->n;
    d.fx = device_upload(fx.data(), fx.size(), fw.ctx->stream);
    d.fy = device_upload(fy.data(), fy.size(), fw.ctx->stream);
    d.fz = device_upload(fz.data(), fz.size(), fw.ctx->stream);
    d.q = device_upload(d.hq.data(), d.hq.size(), fw.ctx->stream);
    ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
}

// sum over molecules of pairExclusionSum, with or without the per-molecule -kC factor
double molecules_exclusion(const Framework& fw, const aleator_particles* p, const uint64_t* off,
                           const uint64_t* sites, uint64_t nm, bool per_molecule_factor) {
    double total = 0.0;
    std::vector<aleator_charged_site> buf;
    for (uint64_t m = 0; m < nm; ++m) {
        buf.clear();
        for (uint64_t s = off[m]; s < off[m + 1]; ++s) {
            const uint64_t i = sites[s];
            require(i < p->n, ALEATOR_OUT_OF_RANGE, "molecule site index out of range");
            buf.push_back({p->x[i], p->y[i], p->z[i], p->charge ? p->charge[i] : 0.0});
        }
        const double e = pair_exclusion_sum(fw.lat, buf.data(), (int)buf.size(), fw.alpha);
        total += per_molecule_factor ? -kCoulomb * e : e;
    }
    return total;
}

}  // namespace

// ---------------------------------------------------------------- shared host math
HostLattice make_lattice(const double m[3][3]) {
    HostLattice h{};
    std::memcpy(h.m, m, sizeof(h.m));
    const double* a0 = m[0];
    const double* a1 = m[1];
    const double* a2 = m[2];
    // Lattice::volume (lattice.hpp:35-41)
    const double cx = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double cy = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double cz = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    h.volume = m[0][0] * cx + m[0][1] * cy + m[0][2] * cz;
    const Vec c12 = cross(a1, a2), c20 = cross(a2, a0), c01 = cross(a0, a1);
    for (int r = 0; r < 3; ++r) {
        h.inv[r][0] = c12.v[r] / h.volume;
        h.inv[r][1] = c20.v[r] / h.volume;
        h.inv[r][2] = c01.v[r] / h.volume;
    }
    const Vec* faces[3] = {&c12, &c20, &c01};
    for (int a = 0; a < 3; ++a) {
        const double area = std::sqrt(dot(faces[a]->v, faces[a]->v));
        h.lperp[a] = area > 0 ? std::abs(h.volume) / area : 0.0;
    }
    return h;
}

// buildReciprocalTerms (ewald.cc:137-161), reciprocalVectors (ewald.cc:33-45)
KSet make_kset(const HostLattice& lat, double alpha, int kmax) {
    const double two_pi_over_v = 2.0 * kPi / lat.volume;
    const Vec c12 = cross(lat.m[1], lat.m[2]), c20 = cross(lat.m[2], lat.m[0]), c01 = cross(lat.m[0], lat.m[1]);
    double b[3][3];
    for (int j = 0; j < 3; ++j) {
        b[0][j] = c12.v[j] * two_pi_over_v;
        b[1][j] = c20.v[j] * two_pi_over_v;
        b[2][j] = c01.v[j] * two_pi_over_v;
    }
    const double four_alpha_sq = 4.0 * alpha * alpha;
    const int bound = kmax * kmax + 2;
    KSet ks;
    for (int n0 = -kmax; n0 <= kmax; ++n0)
        for (int n1 = -kmax; n1 <= kmax; ++n1)
            for (int n2 = -kmax; n2 <= kmax; ++n2) {
                if (n0 == 0 && n1 == 0 && n2 == 0) continue;
                if (n0 * n0 + n1 * n1 + n2 * n2 >= bound) continue;
                double k[3];
                for (int j = 0; j < 3; ++j) k[j] = (b[0][j] * n0 + b[1][j] * n1) + b[2][j] * n2;
                const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                ks.n.insert(ks.n.end(), {n0, n1, n2});
                ks.k.insert(ks.k.end(), {k[0], k[1], k[2]});
                ks.weight.push_back(std::exp(-k2 / four_alpha_sq) / k2);
            }
    return ks;
}

void lj_pair_constants(const Framework& fw, uint32_t guest, uint32_t fws, double& sig2, double& eps4,
                       double& vrc, double& dvrc) {
    double eps, sig;
    // mixedParameters (lennard_jones.cc:49-68)
    if (guest == fws) {
        eps = fw.eps[guest];
        sig = fw.sigma[guest];
    } else {
        eps = std::sqrt(fw.eps[guest] * fw.eps[fws]);
        sig = fw.mixing == ALEATOR_MIX_LORENTZ_BERTHELOT ? 0.5 * (fw.sigma[guest] + fw.sigma[fws])
                                                         : std::sqrt(fw.sigma[guest] * fw.sigma[fws]);
    }
    sig2 = sig * sig;
    eps4 = 4.0 * eps;
    vrc = raw_potential(fw.rc, eps, sig);
    dvrc = raw_derivative(fw.rc, eps, sig);
}

void Context::time_begin() {
    if (!timing) return;
    hipEvent_t a, b;
    ALEATOR_HIP_CHECK(hipEventCreate(&a));
    ALEATOR_HIP_CHECK(hipEventCreate(&b));
    events.emplace_back(a, b);
    ALEATOR_HIP_CHECK(hipEventRecord(a, stream));
}

void Context::time_end() {
    if (!timing || events.empty()) return;
    ALEATOR_HIP_CHECK(hipEventRecord(events.back().second, stream));
}

void* Context::ensure_scratch(size_t bytes) {
    if (bytes <= scratch_bytes && scratch) return scratch;
    ALEATOR_HIP_CHECK(hipStreamSynchronize(stream));
    if (scratch) ALEATOR_HIP_CHECK(hipFree(scratch));
    scratch = nullptr;
    scratch_bytes = 0;
    const size_t sz = std::max<size_t>(bytes, 1 << 20);
    ALEATOR_HIP_CHECK(hipMalloc(&scratch, sz));
    scratch_bytes = sz;
    return scratch;
}

void retain(Context* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

void release(Context* c) {
    if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->scratch) (void)hipFree(c->scratch);
    (void)hipStreamDestroy(c->stream);
    delete static_cast<aleator_ctx*>(c);
}

void retain(Framework* f) { f->refs.fetch_add(1, std::memory_order_relaxed); }

void release(Framework* f) {
    if (f->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    Context* c = f->ctx;
    (void)hipStreamSynchronize(c->stream);
    for (void* p : {(void*)f->d_fx, (void*)f->d_fy, (void*)f->d_fz, (void*)f->d_q, (void*)f->d_type,
                    (void*)f->d_kn, (void*)f->d_kw, (void*)f->d_kseg, (void*)f->d_sre, (void*)f->d_sim,
                    (void*)f->d_widom_tab, (void*)f->d_nonfinite})
        if (p) (void)hipFree(p);
    delete static_cast<aleator_framework*>(f);
    release(c);
}

}  // namespace aleator_gpu

using namespace aleator_gpu;

extern "C" {

const char* aleator_gpu_last_error(void) { return last_error_slot().c_str(); }

aleator_status aleator_gpu_init(int32_t device, aleator_ctx** out) {
    return guarded([&] {
        require(out != nullptr, ALEATOR_INVALID_ARGUMENT, "out must not be null");
        int count = 0;
        ALEATOR_HIP_CHECK(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, ALEATOR_INVALID_ARGUMENT, "device index out of range");
        ALEATOR_HIP_CHECK(hipSetDevice(device));
        auto ctx = std::make_unique<aleator_ctx>();
        ctx->device = device;
        ALEATOR_HIP_CHECK(hipStreamCreate(&ctx->stream));
        *out = ctx.release();
    });
}

aleator_status aleator_gpu_timing(aleator_ctx* ctx, int32_t enable, double* total_ms, uint64_t* launches) {
    return guarded([&] {
        require(ctx != nullptr, ALEATOR_INVALID_ARGUMENT, "null context");
        ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        double total = 0.0;
        for (auto& e : ctx->events) {
            float ms = 0.0f;
            ALEATOR_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
            total += ms;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (total_ms) *total_ms = total;
        if (launches) *launches = ctx->events.size();
        ctx->events.clear();
        ctx->timing = enable != 0;
    });
}

aleator_status aleator_gpu_destroy(aleator_ctx* ctx) {
    return guarded([&] {
        if (ctx) release(ctx);
    });
}

void* aleator_gpu_stream(aleator_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

aleator_status aleator_gpu_framework_create(aleator_ctx* ctx, const aleator_lattice* lattice,
                                            const aleator_particles* atoms, const aleator_lj_params* lj,
                                            const aleator_ewald_params* ewald, aleator_framework** out) {
    return guarded([&] {
        require(ctx && lattice && atoms && lj && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        check_lattice(lattice);
        require(lj->cutoff > 0.0, ALEATOR_INVALID_ARGUMENT, "LennardJones: cutoff must be positive");
        require(lj->n_species >= 1 && lj->epsilon && lj->sigma, ALEATOR_INVALID_ARGUMENT,
                "LennardJones: per-species parameters required");
        require(lj->truncation >= 0 && lj->truncation <= 2, ALEATOR_INVALID_ARGUMENT, "unknown truncation");
        require(lj->mixing >= 0 && lj->mixing <= 1, ALEATOR_INVALID_ARGUMENT, "unknown mixing rule");
        require(atoms->n == 0 || (atoms->x && atoms->y && atoms->z && atoms->species), ALEATOR_INVALID_ARGUMENT,
                "framework arrays must not be null");
        require(atoms->n < (1ull << 31), ALEATOR_INVALID_ARGUMENT, "too many framework atoms");
        auto fw = std::make_unique<aleator_framework>();
        fw->ctx = ctx;
        fw->lat = make_lattice(lattice->m);
        fw->n = atoms->n;
        fw->eps.assign(lj->epsilon, lj->epsilon + lj->n_species);
        fw->sigma.assign(lj->sigma, lj->sigma + lj->n_species);
        fw->rc = lj->cutoff;
        fw->trunc = lj->truncation;
        fw->mixing = lj->mixing;
        // Lattice::validateCutoff (lattice.cc:286-300): the grid and Widom kernels rely on it
        for (int a = 0; a < 3; ++a)
            require(!(fw->rc > 0.5 * fw->lat.lperp[a]), ALEATOR_INVALID_ARGUMENT,
                    "cutoff exceeds half the perpendicular width of the lattice");
        const uint64_t n = atoms->n;
        fw->x.assign(atoms->x, atoms->x + n);
        fw->y.assign(atoms->y, atoms->y + n);
        fw->z.assign(atoms->z, atoms->z + n);
        fw->q.assign(n, 0.0);
        if (atoms->charge) fw->q.assign(atoms->charge, atoms->charge + n);
        fw->species.assign(atoms->species, atoms->species + n);
        std::map<uint32_t, int32_t> types;
        fw->fx.resize(n);
        fw->fy.resize(n);
        fw->fz.resize(n);
        fw->type.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            require(fw->species[i] < lj->n_species, ALEATOR_OUT_OF_RANGE, "framework species out of range");
            auto it = types.find(fw->species[i]);
            if (it == types.end()) {
                require(types.size() < (size_t)kMaxTypes, ALEATOR_NOT_IMPLEMENTED,
                        "more than 32 distinct framework species");
                it = types.emplace(fw->species[i], (int32_t)types.size()).first;
                fw->type_species.push_back(fw->species[i]);
            }
            fw->type[i] = it->second;
            const double c[3] = {fw->x[i], fw->y[i], fw->z[i]};
            double f[3];
            frac_of(fw->lat, c, f);
            fw->fx[i] = f[0] - std::floor(f[0]);
            fw->fy[i] = f[1] - std::floor(f[1]);
            fw->fz[i] = f[2] - std::floor(f[2]);
        }
        hipStream_t s = ctx->stream;
        fw->d_fx = device_upload(fw->fx.data(), n, s);
        fw->d_fy = device_upload(fw->fy.data(), n, s);
        fw->d_fz = device_upload(fw->fz.data(), n, s);
        fw->d_q = device_upload(fw->q.data(), n, s);
        fw->d_type = device_upload(fw->type.data(), n, s);
        ALEATOR_HIP_CHECK(hipMalloc(&fw->d_nonfinite, sizeof(unsigned long long)));
        ALEATOR_HIP_CHECK(hipMemsetAsync(fw->d_nonfinite, 0, sizeof(unsigned long long), s));
        if (ewald) {
            require(ewald->alpha > 0.0, ALEATOR_INVALID_ARGUMENT, "Ewald: alpha must be positive");
            require(ewald->real_space_cutoff > 0.0, ALEATOR_INVALID_ARGUMENT,
                    "Ewald: realSpaceCutoff must be positive");
            require(ewald->kmax >= 1, ALEATOR_INVALID_ARGUMENT, "Ewald: kMax must be at least 1");
            require(ewald->kmax <= kMaxK, ALEATOR_NOT_IMPLEMENTED, "Ewald: kMax above 16");
            for (int a = 0; a < 3; ++a)
                require(!(ewald->real_space_cutoff > 0.5 * fw->lat.lperp[a]), ALEATOR_INVALID_ARGUMENT,
                        "real-space cutoff exceeds half the perpendicular width of the lattice");
            double total = 0.0, q2 = 0.0;
            for (double v : fw->q) total += v;
            require(!(std::abs(total) > 1e-8), ALEATOR_INVALID_ARGUMENT,
                    "Ewald: system is not charge-neutral");
            for (double v : fw->q) q2 += v * v;
            fw->charged = true;
            fw->alpha = ewald->alpha;
            fw->rc_real = ewald->real_space_cutoff;
            fw->kmax = ewald->kmax;
            fw->kset = make_kset(fw->lat, fw->alpha, fw->kmax);
            const size_t nk = fw->kset.size();
            fw->nk_half = (int)(nk / 2);
            std::vector<int32_t> kn4(4 * (size_t)fw->nk_half);
            for (int t = 0; t < fw->nk_half; ++t) {
                const size_t mt = nk - 1 - t;
                for (int j = 0; j < 3; ++j) {
                    require(fw->kset.n[3 * t + j] == -fw->kset.n[3 * mt + j], ALEATOR_RUNTIME_ERROR,
                            "k-set is not centrally symmetric");
                    kn4[4 * t + j] = fw->kset.n[3 * t + j];
                }
                kn4[4 * t + 3] = 0;
            }
            fw->d_kn = device_upload(kn4.data(), kn4.size(), s);
            std::vector<int2> seg;
            for (int t = 0; t < fw->nk_half; ++t) {
                const int* n3 = &kn4[4 * t];
                const int len = seg.empty() ? 0 : (seg.back().y >> 24);
                const int* p3 = t > 0 ? &kn4[4 * (t - 1)] : nullptr;
                if (seg.empty() || len == kSegLen || n3[0] != p3[0] || n3[1] != p3[1] || n3[2] != p3[2] + 1) {
                    seg.push_back(make_int2(t, (n3[0] + fw->kmax) | (n3[1] + fw->kmax) << 8 | (n3[2] + fw->kmax) << 16 |
                                                   1 << 24));
                } else {
                    seg.back().y += 1 << 24;
                }
            }
            fw->n_kseg = (int)seg.size();
            fw->d_kseg = device_upload(seg.data(), seg.size(), s);
            fw->d_kw = device_upload(fw->kset.weight.data(), (size_t)fw->nk_half, s);
            ALEATOR_HIP_CHECK(hipMalloc(&fw->d_sre, std::max(1, fw->nk_half) * sizeof(double)));
            ALEATOR_HIP_CHECK(hipMalloc(&fw->d_sim, std::max(1, fw->nk_half) * sizeof(double)));
            structure_factor(*fw, n, fw->d_fx, fw->d_fy, fw->d_fz, fw->d_q, fw->d_sre, fw->d_sim);
            fw->h_sre.resize(fw->nk_half);
            fw->h_sim.resize(fw->nk_half);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(fw->h_sre.data(), fw->d_sre, fw->nk_half * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
            ALEATOR_HIP_CHECK(hipMemcpyAsync(fw->h_sim.data(), fw->d_sim, fw->nk_half * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
            fw->self_energy = -kCoulomb * (fw->alpha / std::sqrt(kPi)) * q2;
        }
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
        retain(ctx);
        *out = fw.release();
    });
}

aleator_status aleator_gpu_framework_destroy(aleator_framework* fw) {
    return guarded([&] {
        if (fw) release(fw);
    });
}

aleator_status aleator_gpu_framework_num_k(const aleator_framework* fw, uint64_t* nk) {
    return guarded([&] {
        require(fw && nk, ALEATOR_INVALID_ARGUMENT, "null argument");
        *nk = fw->kset.size();
    });
}

aleator_status aleator_gpu_framework_nonfinite_count(aleator_framework* fw, uint64_t* count) {
    return guarded([&] {
        require(fw && count, ALEATOR_INVALID_ARGUMENT, "null argument");
        unsigned long long v = 0;
        ALEATOR_HIP_CHECK(hipMemcpyAsync(&v, fw->d_nonfinite, sizeof(v), hipMemcpyDeviceToHost, fw->ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(fw->ctx->stream));
        *count = v;
    });
}

aleator_status aleator_gpu_grid_point_counts(const aleator_lattice* grid_lattice, double spacing,
                                             int32_t n_points[3]) {
    return guarded([&] {
        require(grid_lattice && n_points, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(spacing > 0.0, ALEATOR_INVALID_ARGUMENT, "FrameworkEnergyGrid: spacingAngstrom must be positive");
        check_lattice(grid_lattice);
        const HostLattice h = make_lattice(grid_lattice->m);
        for (int a = 0; a < 3; ++a) {
            const int v = (int)std::lround(h.lperp[a] / spacing);
            n_points[a] = v < 1 ? 1 : v;
        }
    });
}

static void grid_common(aleator_framework* fw, const aleator_lattice* grid_lattice, const uint32_t* guests,
                        int32_t n_guest, int32_t coulomb, const int32_t n_points[3], int32_t precision,
                        double cap, double* out_device, bool no_alloc = false) {
    require(fw && n_points && out_device, ALEATOR_INVALID_ARGUMENT, "null argument");
    require(n_guest >= 0 && (n_guest == 0 || guests), ALEATOR_INVALID_ARGUMENT, "guest species list");
    GridRequest req;
    req.grid_m = grid_lattice ? &grid_lattice->m[0][0] : nullptr;
    if (grid_lattice) check_lattice(grid_lattice);
    for (int g = 0; g < n_guest; ++g) {
        require(guests[g] < fw->eps.size(), ALEATOR_OUT_OF_RANGE, "guest species out of range");
        req.guests.push_back(guests[g]);
    }
    for (int a = 0; a < 3; ++a) req.n[a] = n_points[a];
    req.coulomb = coulomb != 0;
    req.precision = precision;
    req.cap = cap;
    req.out_device = out_device;
    launch_grid(*fw, req, no_alloc);
}

aleator_status aleator_gpu_framework_set_lj_tail(aleator_framework* fw, int32_t enable) {
    return guarded([&] {
        require(fw != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        fw->lj_tail = enable != 0;
    });
}

aleator_status aleator_gpu_grid_build_device(aleator_framework* fw, const aleator_lattice* grid_lattice,
                                             const uint32_t* guest_species, int32_t n_guest, int32_t coulomb,
                                             const int32_t n_points[3], int32_t precision, double cap,
                                             double* out_device) {
    return guarded([&] {
        grid_common(fw, grid_lattice, guest_species, n_guest, coulomb, n_points, precision, cap, out_device);
    });
}

aleator_status aleator_gpu_grid_build_device_noalloc(aleator_framework* fw, const aleator_lattice* grid_lattice,
                                                     const uint32_t* guest_species, int32_t n_guest, int32_t coulomb,
                                                     const int32_t n_points[3], int32_t precision, double cap,
                                                     double* out_device) {
    return guarded([&] {
        grid_common(fw, grid_lattice, guest_species, n_guest, coulomb, n_points, precision, cap, out_device, true);
    });
}

aleator_status aleator_gpu_grid_build(aleator_framework* fw, const aleator_lattice* grid_lattice,
                                      const uint32_t* guest_species, int32_t n_guest, int32_t coulomb,
                                      const int32_t n_points[3], int32_t precision, double cap,
                                      double* out_host) {
    return guarded([&] {
        require(fw && n_points && out_host, ALEATOR_INVALID_ARGUMENT, "null argument");
        const size_t nodes = (size_t)n_points[0] * n_points[1] * n_points[2];
        const size_t tables = (size_t)std::max(n_guest, 0) + (coulomb ? 1 : 0);
        double* d = nullptr;
        ALEATOR_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, nodes * tables) * sizeof(double)));
        try {
            grid_common(fw, grid_lattice, guest_species, n_guest, coulomb, n_points, precision, cap, d);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(out_host, d, nodes * tables * sizeof(double),
                                             hipMemcpyDeviceToHost, fw->ctx->stream));
            ALEATOR_HIP_CHECK(hipStreamSynchronize(fw->ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(fw->ctx->stream);
            (void)hipFree(d);
            throw;
        }
        ALEATOR_HIP_CHECK(hipFree(d));
    });
}

static void widom_common(aleator_framework* fw, const aleator_molecule_site* sites, int32_t n_sites,
                         uint64_t seed, uint64_t stream, uint64_t first, uint64_t n_trials, double temperature,
                         int32_t precision, double* du_device, double* sum_device) {
    require(fw && sites && sum_device, ALEATOR_INVALID_ARGUMENT, "null argument");
    require(n_sites >= 1, ALEATOR_INVALID_ARGUMENT, "widom: a species needs at least one site");
    WidomRequest req;
    req.sites.assign(sites, sites + n_sites);
    req.seed = seed;
    req.stream = stream;
    req.first_call = first;
    req.n_trials = n_trials;
    req.temperature = temperature;
    req.precision = precision;
    req.du_device = du_device;
    req.sum_device = sum_device;
    launch_widom(*fw, req);
}

aleator_status aleator_gpu_widom_device(aleator_framework* fw, const aleator_molecule_site* sites,
                                        int32_t n_sites, uint64_t seed, uint64_t stream, uint64_t first_call_index,
                                        uint64_t n_trials, double temperature, int32_t precision,
                                        double* du_device, double* sum_device) {
    return guarded([&] {
        widom_common(fw, sites, n_sites, seed, stream, first_call_index, n_trials, temperature, precision,
                     du_device, sum_device);
    });
}

aleator_status aleator_gpu_widom(aleator_framework* fw, const aleator_molecule_site* sites, int32_t n_sites,
                                 uint64_t seed, uint64_t stream, uint64_t first_call_index, uint64_t n_trials,
                                 double temperature, int32_t precision, double* du_host, double* sum_boltzmann) {
    return guarded([&] {
        require(fw && sum_boltzmann, ALEATOR_INVALID_ARGUMENT, "null argument");
        double* d = nullptr;
        ALEATOR_HIP_CHECK(hipMalloc(&d, (std::max<uint64_t>(n_trials, 1) + 1) * sizeof(double)));
        try {
            widom_common(fw, sites, n_sites, seed, stream, first_call_index, n_trials, temperature, precision,
                         d + 1, d);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(sum_boltzmann, d, sizeof(double), hipMemcpyDeviceToHost,
                                             fw->ctx->stream));
            if (du_host)
                ALEATOR_HIP_CHECK(hipMemcpyAsync(du_host, d + 1, n_trials * sizeof(double), hipMemcpyDeviceToHost,
                                                 fw->ctx->stream));
            ALEATOR_HIP_CHECK(hipStreamSynchronize(fw->ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(fw->ctx->stream);
            (void)hipFree(d);
            throw;
        }
        ALEATOR_HIP_CHECK(hipFree(d));
    });
}

// ---------------------------------------------------------------- Ewald state
aleator_status aleator_gpu_ewald_state_create(aleator_framework* fw, aleator_ewald_state** out) {
    return guarded([&] {
        require(fw && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(fw->charged, ALEATOR_INVALID_ARGUMENT, "framework was created without Ewald parameters");
        auto st = std::make_unique<aleator_ewald_state>();
        st->fw = fw;
        const size_t bytes = std::max(1, fw->nk_half) * sizeof(double);
        ALEATOR_HIP_CHECK(hipMalloc(&st->d_sre, bytes));
        ALEATOR_HIP_CHECK(hipMalloc(&st->d_sim, bytes));
        ALEATOR_HIP_CHECK(hipMalloc(&st->d_out, sizeof(double)));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(st->d_sre, fw->d_sre, bytes, hipMemcpyDeviceToDevice, fw->ctx->stream));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(st->d_sim, fw->d_sim, bytes, hipMemcpyDeviceToDevice, fw->ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(fw->ctx->stream));
        st->self_energy = fw->self_energy;
        st->exclusion_energy = 0.0;  // the framework carries no exclusion (ewald_incremental_state.cc:82)
        retain(fw);
        *out = st.release();
    });
}

aleator_status aleator_gpu_ewald_state_destroy(aleator_ewald_state* st) {
    return guarded([&] {
        if (!st) return;
        (void)hipStreamSynchronize(st->fw->ctx->stream);
        (void)hipFree(st->d_sre);
        (void)hipFree(st->d_sim);
        (void)hipFree(st->d_out);
        Framework* fw = st->fw;
        delete st;
        release(fw);
    });
}

// computeDelta (ewald_incremental_state.cc:95-136) for one proposal
static double state_delta(aleator_ewald_state* st, const aleator_charged_site* removed, int32_t nr,
                          const aleator_charged_site* added, int32_t na, bool commit, double* self_out,
                          double* excl_out) {
    Framework& fw = *st->fw;
    require(nr >= 0 && na >= 0 && (nr == 0 || removed) && (na == 0 || added), ALEATOR_INVALID_ARGUMENT,
            "site lists");
    require(nr <= kMaxSites && na <= kMaxSites, ALEATOR_INVALID_ARGUMENT, "at most 8 sites per side");
    std::vector<double> s4 = site4_of(fw.lat, removed, nr);
    const std::vector<double> a4 = site4_of(fw.lat, added, na);
    s4.insert(s4.end(), a4.begin(), a4.end());
    double* d_sites = static_cast<double*>(fw.ctx->ensure_scratch(std::max<size_t>(s4.size(), 4) * sizeof(double)));
    if (!s4.empty())
        ALEATOR_HIP_CHECK(hipMemcpyAsync(d_sites, s4.data(), s4.size() * sizeof(double), hipMemcpyHostToDevice,
                                         fw.ctx->stream));
    reciprocal_delta(fw, st->d_sre, st->d_sim, 1, nr, na, d_sites, st->d_out, commit);
    double recip = 0.0;
    ALEATOR_HIP_CHECK(hipMemcpyAsync(&recip, st->d_out, sizeof(double), hipMemcpyDeviceToHost, fw.ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
    const double self_pref = -kCoulomb * (fw.alpha / std::sqrt(kPi));
    const double self = self_pref * (sum_q2(added, na) - sum_q2(removed, nr));
    const double excl = -kCoulomb * (pair_exclusion_sum(fw.lat, added, na, fw.alpha) -
                                     pair_exclusion_sum(fw.lat, removed, nr, fw.alpha));
    *self_out = self;
    *excl_out = excl;
    return recip + self + excl;
}

aleator_status aleator_gpu_ewald_state_propose(aleator_ewald_state* st, const aleator_charged_site* removed,
                                               int32_t n_removed, const aleator_charged_site* added,
                                               int32_t n_added, double* delta_energy) {
    return guarded([&] {
        require(st && delta_energy, ALEATOR_INVALID_ARGUMENT, "null argument");
        double self, excl;
        *delta_energy = state_delta(st, removed, n_removed, added, n_added, false, &self, &excl);
    });
}

aleator_status aleator_gpu_ewald_state_commit(aleator_ewald_state* st, const aleator_charged_site* removed,
                                              int32_t n_removed, const aleator_charged_site* added,
                                              int32_t n_added) {
    return guarded([&] {
        require(st, ALEATOR_INVALID_ARGUMENT, "null argument");
        double self, excl;
        state_delta(st, removed, n_removed, added, n_added, true, &self, &excl);
        st->self_energy += self;
        st->exclusion_energy += excl;
    });
}

aleator_status aleator_gpu_ewald_state_propose_batch(aleator_ewald_state* st, uint64_t n_prop,
                                                     const aleator_charged_site* removed, int32_t nr,
                                                     const aleator_charged_site* added, int32_t na,
                                                     double* delta_energy) {
    return guarded([&] {
        require(st && delta_energy, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(nr >= 0 && na >= 0 && nr <= kMaxSites && na <= kMaxSites, ALEATOR_INVALID_ARGUMENT, "site counts");
        require((nr == 0 || removed) && (na == 0 || added), ALEATOR_INVALID_ARGUMENT, "site lists");
        if (n_prop == 0) return;
        Framework& fw = *st->fw;
        const int ns = nr + na;
        std::vector<double> s4((size_t)n_prop * ns * 4);
        for (uint64_t p = 0; p < n_prop; ++p) {
            const auto r4 = site4_of(fw.lat, removed + p * nr, nr);
            const auto a4 = site4_of(fw.lat, added + p * na, na);
            std::copy(r4.begin(), r4.end(), s4.begin() + p * ns * 4);
            std::copy(a4.begin(), a4.end(), s4.begin() + p * ns * 4 + r4.size());
        }
        double* d = nullptr;
        ALEATOR_HIP_CHECK(hipMalloc(&d, (s4.size() + n_prop) * sizeof(double)));
        try {
            ALEATOR_HIP_CHECK(hipMemcpyAsync(d, s4.data(), s4.size() * sizeof(double), hipMemcpyHostToDevice,
                                             fw.ctx->stream));
            reciprocal_delta(fw, st->d_sre, st->d_sim, n_prop, nr, na, d, d + s4.size(), false);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(delta_energy, d + s4.size(), n_prop * sizeof(double),
                                             hipMemcpyDeviceToHost, fw.ctx->stream));
            ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(fw.ctx->stream);
            (void)hipFree(d);
            throw;
        }
        ALEATOR_HIP_CHECK(hipFree(d));
        const double self_pref = -kCoulomb * (fw.alpha / std::sqrt(kPi));
        for (uint64_t p = 0; p < n_prop; ++p) {
            const aleator_charged_site* a = added + p * na;
            const aleator_charged_site* r = removed + p * nr;
            const double self = self_pref * (sum_q2(a, na) - sum_q2(r, nr));
            const double excl = -kCoulomb * (pair_exclusion_sum(fw.lat, a, na, fw.alpha) -
                                             pair_exclusion_sum(fw.lat, r, nr, fw.alpha));
            delta_energy[p] = delta_energy[p] + self + excl;
        }
    });
}

aleator_status aleator_gpu_ewald_state_committed_energy(aleator_ewald_state* st, double* energy) {
    return guarded([&] {
        require(st && energy, ALEATOR_INVALID_ARGUMENT, "null argument");
        std::vector<double> re, im;
        full_structure_factor(*st->fw, st->d_sre, st->d_sim, re, im);
        *energy = reciprocal_energy(*st->fw, re, im) + st->self_energy + st->exclusion_energy;
    });
}

aleator_status aleator_gpu_ewald_state_resync(aleator_ewald_state* st, const aleator_particles* particles,
                                              const uint64_t* off, const uint64_t* sites, uint64_t nm) {
    return guarded([&] {
        require(st, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(nm == 0 || (off && sites), ALEATOR_INVALID_ARGUMENT, "molecule lists");
        Framework& fw = *st->fw;
        DeviceParticles d;
        upload_particles(fw, particles, d);
        structure_factor(fw, d.n, d.fx, d.fy, d.fz, d.q, st->d_sre, st->d_sim);
        ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
        double q2 = 0.0;
        for (double v : d.hq) q2 += v * v;
        st->self_energy = -kCoulomb * (fw.alpha / std::sqrt(kPi)) * q2;
        st->exclusion_energy = -kCoulomb * molecules_exclusion(fw, particles, off, sites, nm, false);
    });
}

aleator_status aleator_gpu_ewald_state_recompute(aleator_ewald_state* st, const aleator_particles* particles,
                                                 const uint64_t* off, const uint64_t* sites, uint64_t nm,
                                                 double* energy) {
    return guarded([&] {
        require(st && energy, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(nm == 0 || (off && sites), ALEATOR_INVALID_ARGUMENT, "molecule lists");
        Framework& fw = *st->fw;
        DeviceParticles d;
        upload_particles(fw, particles, d);
        double *sre = nullptr, *sim = nullptr;
        const size_t bytes = std::max(1, fw.nk_half) * sizeof(double);
        ALEATOR_HIP_CHECK(hipMalloc(&sre, bytes));
        ALEATOR_HIP_CHECK(hipMalloc(&sim, bytes));
        std::vector<double> re, im;
        try {
            structure_factor(fw, d.n, d.fx, d.fy, d.fz, d.q, sre, sim);
            full_structure_factor(fw, sre, sim, re, im);
        } catch (...) {
            (void)hipStreamSynchronize(fw.ctx->stream);
            (void)hipFree(sre);
            (void)hipFree(sim);
            throw;
        }
        (void)hipFree(sre);
        (void)hipFree(sim);
        double q2 = 0.0;
        for (double v : d.hq) q2 += v * v;
        const double self = -kCoulomb * (fw.alpha / std::sqrt(kPi)) * q2;
        const double excl = molecules_exclusion(fw, particles, off, sites, nm, true);
        *energy = reciprocal_energy(fw, re, im) + self + excl;
    });
}

aleator_status aleator_gpu_ewald_state_raw(aleator_ewald_state* st, double* out, uint64_t capacity,
                                           uint64_t* written) {
    return guarded([&] {
        require(st && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        const size_t nk = st->fw->kset.size();
        require(capacity >= 2 * nk + 2, ALEATOR_INVALID_ARGUMENT, "raw: capacity below 2*Nk+2");
        std::vector<double> re, im;
        full_structure_factor(*st->fw, st->d_sre, st->d_sim, re, im);
        std::copy(re.begin(), re.end(), out);
        std::copy(im.begin(), im.end(), out + nk);
        out[2 * nk] = st->self_energy;
        out[2 * nk + 1] = st->exclusion_energy;
        if (written) *written = 2 * nk + 2;
    });
}

}  // extern "C"
