// grid.hip -- framework energy grids on gfx950: host side of the build and the reciprocal
// pre-contractions.  The tile kernel is in grid_tile.hpp.
//
// Reference: FrameworkEnergyGrid constructor body
// (src/engines/monte_carlo/framework_energy_grid.cc:221-248).
//
// Coulomb table: real space as Ewald::realSpaceParticleEnergy of a unit charge
// (ewald.cc:223-249), reciprocal space as the linear term of computeDelta
// (ewald_incremental_state.cc:117-121).  Every node sits at fractional coordinates
// i/(n*rep) of the k-set's lattice, so exp(i k.r) factorises per axis and the sum over the
// (2K+1)^3 index cube is three small separable contractions in FP64: two pre-kernels here, the
// last one fused into the tile kernel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "grid.hpp"
#include "grid_tile.hpp"
#include "internal.hpp"

namespace aleator_gpu {

namespace {


// T1[n0][n1][iz] = sum_n2 C[n0][n1][n2] * E2[n2][iz]
__global__ void recip_stage_a(const double2* c, const double2* e2, int nk, int nz, double2* t1) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nk * nk * nz) return;
    const int iz = idx % nz;
    const long long n01 = idx / nz;
    double re = 0.0, im = 0.0;
    for (int n2 = 0; n2 < nk; ++n2) {
        const double2 a = c[n01 * nk + n2];
        const double2 b = e2[(long long)n2 * nz + iz];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    t1[idx] = make_double2(re, im);
}

// T2[n0][iz][iy] = sum_n1 T1[n0][n1][iz] * E1[n1][iy]
__global__ void recip_stage_b(const double2* t1, const double2* e1, int nk, int ny, int nz, double2* t2) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nk * nz * ny) return;
    const int iy = idx % ny;
    const int iz = (idx / ny) % nz;
    const int n0 = idx / ((long long)ny * nz);
    double re = 0.0, im = 0.0;
    for (int n1 = 0; n1 < nk; ++n1) {
        const double2 a = t1[((long long)n0 * nk + n1) * nz + iz];
        const double2 b = e1[(long long)n1 * ny + iy];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    t2[idx] = make_double2(re, im);
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

int tile_blocks(const int32_t n[3], int shape) {
    const TileShape& t = kShapes[shape];
    const long long b = (long long)((n[0] + t.tx - 1) / t.tx) * ((n[1] + t.ty - 1) / t.ty) *
                        ((n[2] + t.tz * t.npt - 1) / (t.tz * t.npt));
    if (b > 0x3fffffff) fail(ALEATOR_INVALID_ARGUMENT, "grid: too many nodes");
    return (int)b;
}

// ---------------------------------------------------------------- job: one framework's build
void GridJob::prepare(Framework& fw, const GridRequest& req) {
    ng = (int)req.guests.size();
    coulomb = req.coulomb;
    precision = req.precision;
    if (ng > kMaxGuests) fail(ALEATOR_INVALID_ARGUMENT, "grid: at most 4 guest species per build");
    if (ng == 0 && !req.coulomb) fail(ALEATOR_INVALID_ARGUMENT, "grid: nothing to build");
    if (req.coulomb && !fw.charged)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: a Coulomb table needs a framework created with Ewald parameters");
    if (req.precision != ALEATOR_PRECISION_FP64 && req.precision != ALEATOR_PRECISION_MIXED)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown precision");
    for (int i = 0; i < 3; ++i)
        if (req.n[i] < 1) fail(ALEATOR_INVALID_ARGUMENT, "grid: point counts must be >= 1");
    GridArgs& a = args;
    a = GridArgs{};
    for (int i = 0; i < 3; ++i) {
        a.n[i] = req.n[i];
        for (int j = 0; j < 3; ++j) a.mi[3 * i + j] = fw.lat.m[i][j];
    }
    // grid lattice: the framework lattice or a primitive cell whose rows are the framework rows
    // divided by integers
    double gm[9];
    if (req.grid_m) {
        std::memcpy(gm, req.grid_m, sizeof(gm));
        for (int i = 0; i < 3; ++i) {
            const double rep = std::round(norm3(&a.mi[3 * i]) / norm3(&gm[3 * i]));
            if (rep < 1.0) fail(ALEATOR_INVALID_ARGUMENT, "grid: grid lattice larger than the framework lattice");
            for (int j = 0; j < 3; ++j)
                if (std::abs(gm[3 * i + j] * rep - a.mi[3 * i + j]) > 1e-9 * norm3(&a.mi[3 * i]))
                    fail(ALEATOR_INVALID_ARGUMENT,
                         "grid: grid lattice rows must be the framework lattice rows divided by integers");
            a.rep[i] = rep;
        }
    } else {
        std::memcpy(gm, a.mi, sizeof(gm));
        a.rep[0] = a.rep[1] = a.rep[2] = 1.0;
    }
    std::memcpy(a.mg, gm, sizeof(gm));
    // the tile radius is checked against L_perp for the largest tile of every shape
    rmax = std::max(ng > 0 ? fw.rc : 0.0, req.coulomb ? fw.rc_real : 0.0);
    for (int i = 0; i < 3; ++i) lperp[i] = fw.lat.lperp[i];
    a.fx = fw.d_fx;
    a.fy = fw.d_fy;
    a.fz = fw.d_fz;
    a.q = fw.d_q;
    a.type = fw.d_type;
    a.nonfinite = fw.d_nonfinite;
    a.natoms = (int)fw.n;
    a.rc_lj = fw.rc;
    a.rc2_lj = fw.rc * fw.rc;
    a.rc2_real = fw.rc_real * fw.rc_real;
    a.alpha = fw.alpha;
    a.kc = kCoulomb;
    a.cap = req.cap;
    a.trunc = fw.trunc;
    a.kmax = req.coulomb ? fw.kmax : 0;
    a.out = req.out_device;
    a.stride = (long long)req.n[0] * req.n[1] * req.n[2];

    // the analytic tail of one guest site per table, when the framework asks for it
    for (int g = 0; g < ng; ++g) a.tail[g] = fw.lj_tail ? lj_tail_delta(fw, &req.guests[g], 1) : 0.0;
    // LJ constants per (guest slot, framework type)
    ljtab.assign(4 * kMaxGuests * kMaxTypes, 0.0);
    for (int g = 0; g < ng; ++g) {
        if (req.guests[g] >= fw.eps.size()) fail(ALEATOR_OUT_OF_RANGE, "guest species out of range");
        for (size_t t = 0; t < fw.type_species.size(); ++t) {
            double s2, e4, v, dv;
            lj_pair_constants(fw, req.guests[g], fw.type_species[t], s2, e4, v, dv);
            ljtab[(0 * kMaxGuests + g) * kMaxTypes + t] = s2;
            ljtab[(1 * kMaxGuests + g) * kMaxTypes + t] = e4;
            ljtab[(2 * kMaxGuests + g) * kMaxTypes + t] = v;
            ljtab[(3 * kMaxGuests + g) * kMaxTypes + t] = dv;
        }
    }
    bytes = align256(ljtab.size() * sizeof(double));
    nk = 2 * fw.kmax + 1;
    if (req.coulomb) {
        cube.assign((size_t)nk * nk * nk, make_double2(0.0, 0.0));
        // C(n) = w(n) conj(S(n)) over the full k-set (S(-n) = conj S(n))
        const size_t nkt = fw.kset.size();
        for (size_t t = 0; t < nkt; ++t) {
            const bool low = (int)t < fw.nk_half;
            const size_t h = low ? t : nkt - 1 - t;
            const double re = fw.h_sre[h], im = low ? fw.h_sim[h] : -fw.h_sim[h];
            const int* nv = &fw.kset.n[3 * t];
            const size_t ci = ((size_t)(nv[0] + fw.kmax) * nk + (nv[1] + fw.kmax)) * nk + (nv[2] + fw.kmax);
            cube[ci] = make_double2(fw.kset.weight[t] * re, -fw.kset.weight[t] * im);
        }
        auto phase_table = [&](int axis, std::vector<double2>& e) {
            const long long period = (long long)req.n[axis] * (long long)a.rep[axis];
            e.resize((size_t)nk * req.n[axis]);
            for (int m = 0; m < nk; ++m)
                for (int i = 0; i < req.n[axis]; ++i) {
                    long long p = ((long long)(m - fw.kmax) * i) % period;
                    if (p < 0) p += period;
                    const double ang = 2.0 * kPi * (double)p / (double)period;
                    e[(size_t)m * req.n[axis] + i] = make_double2(std::cos(ang), std::sin(ang));
                }
        };
        phase_table(0, e0);
        phase_table(1, e1);
        phase_table(2, e2);
        off_c = bytes;
        bytes += align256(cube.size() * sizeof(double2));
        off_e0 = bytes;
        bytes += align256(e0.size() * sizeof(double2));
        off_e1 = bytes;
        bytes += align256(e1.size() * sizeof(double2));
        off_e2 = bytes;
        bytes += align256(e2.size() * sizeof(double2));
        off_t1 = bytes;
        bytes += align256((size_t)nk * nk * req.n[2] * sizeof(double2));
        off_t2 = bytes;
        bytes += align256((size_t)nk * req.n[2] * req.n[1] * sizeof(double2));
        a.recip_pref = kCoulomb * (4.0 * kPi / std::abs(fw.lat.volume));
        fit_poly(0);
    }
}

void GridJob::fit_poly(int min_deg) {
    if (!coulomb) return;
    deg = erf_kernel_poly(args.alpha, std::sqrt(args.rc2_real), poly_tolerance(), min_deg, args.poly, &args.poly_xs);
}

double GridJob::tile_radius(int shape) const {
    const TileShape& t = kShapes[shape];
    const int ext[3] = {t.tx, t.ty, t.tz * t.npt};
    double radius = 0.0;
    for (int c = 0; c < 8; ++c) {
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = (((c >> i) & 1) ? 0.5 : -0.5) * (ext[i] - 1) / args.n[i];
        double v[3];
        for (int j = 0; j < 3; ++j) v[j] = d[0] * args.mg[j] + d[1] * args.mg[3 + j] + d[2] * args.mg[6 + j];
        radius = std::max(radius, norm3(v));
    }
    return radius;
}

bool GridJob::shape_fits(int shape) const {
    const double sr = rmax + tile_radius(shape) + 1e-6 * (1.0 + rmax);
    for (int i = 0; i < 3; ++i)
        if (sr >= lperp[i]) return false;
    return true;
}

void GridJob::set_shape(int s, int chunk) {
    // the 8-image scheme needs every image within the search radius to have fractional offsets
    // inside (-1, 1): guaranteed when the search radius is below L_perp
    if (s < 0 || s >= kNumTileShapes) fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown tile shape");
    if (!shape_fits(s))
        fail(ALEATOR_INVALID_ARGUMENT, "grid: cutoff plus tile radius exceeds the perpendicular width");
    const double sr = rmax + tile_radius(s) + 1e-6 * (1.0 + rmax);
    args.search2 = sr * sr;
    // images this close to the tile centre are inside both cutoffs for every node of the tile
    double rmin = 1e300;
    if (ng > 0) rmin = std::min(rmin, args.rc_lj);
    if (coulomb) rmin = std::min(rmin, std::sqrt(args.rc2_real));
    const double inner = rmin - tile_radius(s) - 1e-9 * (1.0 + rmin);
    args.inner2 = inner > 0.0 ? inner * inner : -1.0;
    // shell entries lie beyond `inner` from the tile centre, so every node sees them at r >= r_far:
    // their erf polynomial is fitted over [r_far, rc] at a lower degree, when that meets the tolerance
    args.far_ok = 0;
    const double r_far = inner - tile_radius(s);
    if (coulomb && deg > kFarDrop && r_far > 0.0)
        args.far_ok = erf_kernel_poly_range(args.alpha, r_far, std::sqrt(args.rc2_real), poly_tolerance(),
                                            deg - kFarDrop, args.poly_far, &args.poly_far_xs, &args.poly_far_xo)
                          ? 1
                          : 0;
    // below half the smallest perpendicular width at most one image of an atom, the nearest,
    // is within the search radius (the fractional rounding image is then the nearest one)
    args.single_image = (2.0 * sr < std::min(lperp[0], std::min(lperp[1], lperp[2]))) ? 1 : 0;
    // more than 128 atoms per pass only with one image per atom (a pass then adds <= 1024 entries)
    if (!(chunk == kChunkSmall || (args.single_image && (chunk == 256 || chunk == 512 || chunk == 1024))))
        fail(ALEATOR_INVALID_ARGUMENT,
             "grid: atoms per list pass must be 128, or 256, 512 or 1024 with one image per atom");
    args.chunk = chunk;
    blocks = tile_blocks(args.n, s);
    shape = s;
}

bool GridJob::chunk_fits(int s, int chunk) const {
    if (chunk == kChunkSmall) return true;
    if (chunk != 256 && chunk != 512 && chunk != 1024) return false;
    const double sr = rmax + tile_radius(s) + 1e-6 * (1.0 + rmax);
    return 2.0 * sr < std::min(lperp[0], std::min(lperp[1], lperp[2]));
}

void GridJob::bind(char* base) {
    dev_base = base;
    args.ljtab = reinterpret_cast<const double*>(base);
    if (coulomb) {
        args.t2 = reinterpret_cast<const double2*>(base + off_t2);
        args.e0 = reinterpret_cast<const double2*>(base + off_e0);
    }
}

void GridJob::upload(hipStream_t s) const {
    ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base, ljtab.data(), ljtab.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if (coulomb) {
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_c, cube.data(), cube.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e0, e0.data(), e0.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e1, e1.data(), e1.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e2, e2.data(), e2.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
    }
}

void GridJob::launch_pre(hipStream_t s) const {
    if (!coulomb) return;
    auto* d_c = reinterpret_cast<const double2*>(dev_base + off_c);
    auto* d_e1 = reinterpret_cast<const double2*>(dev_base + off_e1);
    auto* d_e2 = reinterpret_cast<const double2*>(dev_base + off_e2);
    auto* d_t1 = reinterpret_cast<double2*>(dev_base + off_t1);
    auto* d_t2 = reinterpret_cast<double2*>(dev_base + off_t2);
    const long long na = (long long)nk * nk * args.n[2];
    hipLaunchKernelGGL(recip_stage_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, d_c, d_e2, nk, args.n[2],
                       d_t1);
    const long long nb = (long long)nk * args.n[2] * args.n[1];
    hipLaunchKernelGGL(recip_stage_b, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, d_t1, d_e1, nk, args.n[1],
                       args.n[2], d_t2);
}

void launch_tiles(int precision, int shape, int ng, int deg, int ljk, const GridArgs* d_descs, const int* d_prefix,
                  int n_desc, int blocks, hipStream_t s) {
    using Fn = void (*)(int, int, int, const GridArgs*, const int*, int, int, hipStream_t);
    static const Fn kFp64[3] = {launch_tiles_fp64_trunc, launch_tiles_fp64_shift, launch_tiles_fp64_lfs};
    static const Fn kMixed[3] = {launch_tiles_mixed_trunc, launch_tiles_mixed_shift, launch_tiles_mixed_lfs};
    if (ljk < 0 || ljk > 2) fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown LJ truncation");
    (precision == ALEATOR_PRECISION_FP64 ? kFp64 : kMixed)[ljk](shape, ng, deg, d_descs, d_prefix, n_desc, blocks, s);
}

// ---------------------------------------------------------------- allocation-free builds
// The build's device scratch (LJ constants, reciprocal contraction tables, descriptor) in a fixed
// array of this library, for callers that cannot allocate.  Each GPU holds its own copy for the
// life of the process; its address is looked up once per device, with that device current.  All
// calls on one GPU share it, so calls that overlap on one GPU race.
namespace {
__device__ __align__(256) char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mutex;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

void* no_alloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kMaxDevices) fail(ALEATOR_INVALID_ARGUMENT, "grid: device index out of range");
    if (!g_scratch_ptr[device]) {
        int current = 0;
        ALEATOR_HIP_CHECK(hipGetDevice(&current));
        if (current != device) ALEATOR_HIP_CHECK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        if (e != hipSuccess || p == nullptr)
            fail(ALEATOR_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

// single build through the context scratch buffer, or the fixed array (the per-call entry points)
void launch_grid(Framework& fw, const GridRequest& req, bool no_alloc) {
    Context& ctx = *fw.ctx;
    GridJob job;
    job.prepare(fw, req);
    const int shape = job.shape_fits(ctx.tile_shape) ? ctx.tile_shape : 0;
    job.set_shape(shape, job.chunk_fits(shape, ctx.tile_chunk) ? ctx.tile_chunk : 128);
    const size_t off_desc = job.bytes, off_prefix = align256(off_desc + sizeof(GridArgs));
    char* base = nullptr;
    if (no_alloc) {
        if (off_prefix + 256 > kNoAllocScratchBytes)
            fail(ALEATOR_INVALID_ARGUMENT, "grid: this build needs " + std::to_string(off_prefix + 256) +
                                               " bytes of scratch, above the allocation-free limit of " +
                                               std::to_string(kNoAllocScratchBytes) + " bytes");
        base = static_cast<char*>(no_alloc_scratch(ctx.device));
    } else {
        base = static_cast<char*>(ctx.ensure_scratch(off_prefix + 256));
    }
    job.bind(base);
    job.upload(ctx.stream);
    const int prefix[2] = {0, job.blocks};
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_desc, &job.args, sizeof(GridArgs), hipMemcpyHostToDevice, ctx.stream));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_prefix, prefix, sizeof(prefix), hipMemcpyHostToDevice, ctx.stream));
    job.launch_pre(ctx.stream);
    ctx.time_begin();
    launch_tiles(job.precision, job.shape_used(), job.ng, job.deg, lj_kind(job.args.trunc),
                 reinterpret_cast<const GridArgs*>(base + off_desc),
                 reinterpret_cast<const int*>(base + off_prefix), 1, job.blocks, ctx.stream);
    ctx.time_end();
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
