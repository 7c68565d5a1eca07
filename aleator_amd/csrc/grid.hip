// grid.hip -- framework energy grid on gfx950.
//
// Reference: FrameworkEnergyGrid constructor body
// (src/engines/monte_carlo/framework_energy_grid.cc:221-248): for every node,
// E = min(sum over framework atoms of pairEnergy(|minimum image|, mix), cap).
//
// MI355X design.  One workgroup owns a tile of 8x8x4 nodes (one wave per z-slice
// of 8x8).  The workgroup walks the framework atoms in chunks of 128, and for each
// atom tests the 8 periodic images whose fractional offset from the tile centre
// lies in [-1,1)^3; every image within (cutoff + tile radius) of the tile centre is
// appended to an LDS list (structure of arrays: dx, dy, dz relative to the tile
// centre, charge, type) through a deterministic block scan.  Because the cutoff is
// at most half the smallest perpendicular width, at most one image of an atom lies
// within the cutoff of any node, so summing every listed image inside the cutoff is
// the reference's minimum-image sum.  Each thread then streams the LDS list (a
// broadcast read: every lane reads the same entry) against its own node.  The work
// is VALU-bound: the list is reused by 256 nodes, the output is 8 B per node.
//
// Coulomb table: real space as Ewald::realSpaceParticleEnergy of a unit charge
// (ewald.cc:223-249), reciprocal space as the linear term of computeDelta
// (ewald_incremental_state.cc:117-121).  Since every node sits at fractional
// coordinates i/(n*rep) of the k-set's lattice, exp(i k.r) factorises per axis and
// the reciprocal sum over the (2K+1)^3 index cube is evaluated as three small
// separable contractions in FP64 (two pre-kernels here, the last one fused into the
// tile kernel).
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace aleator_gpu {

namespace {

constexpr int kTX = 8, kTY = 8, kTZ = 4;
constexpr int kThreads = kTX * kTY * kTZ;
constexpr int kCap = 1024;   // LDS list capacity (entries)
constexpr int kChunk = 128;  // atoms staged per pass (8 images each <= kCap)

struct GridArgs {
    int n[3];
    int tiles[3];
    double mg[9];   // grid lattice rows
    double mi[9];   // image (framework) lattice rows
    double rep[3];  // image fractional = grid fractional / rep
    const double* fx;
    const double* fy;
    const double* fz;
    const double* q;
    const int32_t* type;
    int natoms;
    int ntypes;
    double search2;
    double rc_lj, rc2_lj;
    double rc2_real, alpha, kc;
    double cap;
    int trunc;
    const double* ljtab;  // [4][kMaxGuests][kMaxTypes]: sig2, eps4, V(rc), V'(rc)
    const double2* t2;    // [2K+1][nz][ny]
    const double2* e0;    // [2K+1][nx]
    int kmax;
    double recip_pref;
    double* out;
    long long stride;
};

template <typename R>
struct Entries {
    R x[kCap], y[kCap], z[kCap], q[kCap];
    int t[kCap];
};

template <typename R>
__device__ inline R rcp(R v);
template <>
__device__ inline double rcp<double>(double v) {
    return 1.0 / v;
}
template <>
__device__ inline float rcp<float>(float v) {
    return __builtin_amdgcn_rcpf(v);
}

template <typename R>
__device__ inline R erfc_over_r(R alpha, R r);
template <>
__device__ inline double erfc_over_r<double>(double alpha, double r) {
    return erfc(alpha * r) / r;
}
template <>
__device__ inline float erfc_over_r<float>(float alpha, float r) {
    return erfcf(alpha * r) * __builtin_amdgcn_rcpf(r);
}

template <typename R, int NG, bool COUL>
__global__ __launch_bounds__(kThreads) void grid_tile_kernel(const GridArgs a) {
    __shared__ Entries<R> ent;
    __shared__ double tab[4][kMaxGuests > 0 ? kMaxGuests : 1][kMaxTypes];
    __shared__ int scan_tmp[kThreads / 64 + 1];

    const int tid = threadIdx.x;
    for (int i = tid; i < 4 * kMaxGuests * kMaxTypes; i += kThreads)
        (&tab[0][0][0])[i] = a.ljtab[i];

    const int tile = blockIdx.x;
    const int bx = tile % a.tiles[0];
    const int by = (tile / a.tiles[0]) % a.tiles[1];
    const int bz = tile / (a.tiles[0] * a.tiles[1]);
    const int tx = tid % kTX, ty = (tid / kTX) % kTY, tz = tid / (kTX * kTY);
    const int ix = bx * kTX + tx, iy = by * kTY + ty, iz = bz * kTZ + tz;
    const bool valid = ix < a.n[0] && iy < a.n[1] && iz < a.n[2];

    // tile centre in grid and image fractional coordinates
    const double cg0 = (bx * kTX + 0.5 * (kTX - 1)) / a.n[0];
    const double cg1 = (by * kTY + 0.5 * (kTY - 1)) / a.n[1];
    const double cg2 = (bz * kTZ + 0.5 * (kTZ - 1)) / a.n[2];
    const double ci0 = cg0 / a.rep[0], ci1 = cg1 / a.rep[1], ci2 = cg2 / a.rep[2];
    // node offset from the tile centre (Cartesian)
    const double d0 = (tx - 0.5 * (kTX - 1)) / a.n[0];
    const double d1 = (ty - 0.5 * (kTY - 1)) / a.n[1];
    const double d2 = (tz - 0.5 * (kTZ - 1)) / a.n[2];
    const R ox = (R)(d0 * a.mg[0] + d1 * a.mg[3] + d2 * a.mg[6]);
    const R oy = (R)(d0 * a.mg[1] + d1 * a.mg[4] + d2 * a.mg[7]);
    const R oz = (R)(d0 * a.mg[2] + d1 * a.mg[5] + d2 * a.mg[8]);

    double acc[NG > 0 ? NG : 1];
#pragma unroll
    for (int g = 0; g < (NG > 0 ? NG : 1); ++g) acc[g] = 0.0;
    double accq = 0.0;

    const R rc2 = (R)a.rc2_lj;
    const R rcr2 = (R)a.rc2_real;
    const R alpha = (R)a.alpha;

    auto compute = [&](int count) {
        __syncthreads();
        if (valid) {
            for (int e = 0; e < count; ++e) {
                const R dx = ent.x[e] - ox;
                const R dy = ent.y[e] - oy;
                const R dz = ent.z[e] - oz;
                const R r2 = dx * dx + dy * dy + dz * dz;
                if (NG > 0 && r2 <= rc2) {
                    const int t = ent.t[e];
                    const R ir2 = rcp<R>(r2);
                    R r = 0;
                    if (a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) r = sqrt(r2);
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        const R s2 = (R)tab[0][g][t] * ir2;
                        const R s6 = s2 * s2 * s2;
                        R v = (R)tab[1][g][t] * (s6 * s6 - s6);
                        if (a.trunc == ALEATOR_LJ_SHIFTED) {
                            v -= (R)tab[2][g][t];
                        } else if (a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) {
                            v = v - (R)tab[2][g][t] - (r - (R)a.rc_lj) * (R)tab[3][g][t];
                        }
                        acc[g] += (double)v;
                    }
                }
                if (COUL && r2 <= rcr2) {
                    const R r = sqrt(r2);
                    accq += (double)(ent.q[e] * erfc_over_r<R>(alpha, r));
                }
            }
        }
        __syncthreads();
    };

    int pending = 0;
    for (int base = 0; base < a.natoms; base += kChunk) {
        const int at = base + tid;
        int cnt = 0;
        unsigned mask = 0;
        double h0 = 0, h1 = 0, h2 = 0;
        if (tid < kChunk && at < a.natoms) {
            h0 = a.fx[at] - ci0;
            h1 = a.fy[at] - ci1;
            h2 = a.fz[at] - ci2;
            h0 -= floor(h0);
            h1 -= floor(h1);
            h2 -= floor(h2);
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                const double cx = g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6];
                const double cy = g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7];
                const double cz = g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8];
                if (cx * cx + cy * cy + cz * cz <= a.search2) {
                    mask |= 1u << s;
                    ++cnt;
                }
            }
        }
        int total;
        const int off = block_exclusive_scan(cnt, scan_tmp, &total);
        if (pending + total > kCap) {
            compute(pending);
            pending = 0;
        }
        if (mask) {
            int w = pending + off;
            const double qa = a.q[at];
            const int ta = a.type[at];
            for (int s = 0; s < 8; ++s) {
                if (!(mask & (1u << s))) continue;
                const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                ent.x[w] = (R)(g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6]);
                ent.y[w] = (R)(g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7]);
                ent.z[w] = (R)(g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8]);
                ent.q[w] = (R)qa;
                ent.t[w] = ta;
                ++w;
            }
        }
        pending += total;
    }
    compute(pending);

    if (!valid) return;
    const long long flat = ((long long)iz * a.n[1] + iy) * a.n[0] + ix;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const double e = acc[g];
        a.out[g * a.stride + flat] = (a.cap < e) ? a.cap : e;  // std::min(e, cap) incl. NaN
    }
    if (COUL) {
        double rec = 0.0;
        const int nk = 2 * a.kmax + 1;
        for (int m = 0; m < nk; ++m) {
            const double2 t = a.t2[((long long)m * a.n[2] + iz) * a.n[1] + iy];
            const double2 e = a.e0[(long long)m * a.n[0] + ix];
            rec += t.x * e.x - t.y * e.y;
        }
        a.out[NG * a.stride + flat] = a.kc * accq + a.recip_pref * rec;
    }
}

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
__global__ void recip_stage_b(const double2* t1, const double2* e1, int nk, int ny, int nz,
                              double2* t2) {
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

template <typename R, int NG, bool COUL>
void launch_tiles(const GridArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((grid_tile_kernel<R, NG, COUL>), dim3(blocks), dim3(kThreads), 0, s, a);
}

template <typename R>
void dispatch(const GridArgs& a, int ng, bool coul, int blocks, hipStream_t s) {
    if (coul) {
        switch (ng) {
            case 0: return launch_tiles<R, 0, true>(a, blocks, s);
            case 1: return launch_tiles<R, 1, true>(a, blocks, s);
            case 2: return launch_tiles<R, 2, true>(a, blocks, s);
            case 3: return launch_tiles<R, 3, true>(a, blocks, s);
            case 4: return launch_tiles<R, 4, true>(a, blocks, s);
        }
    } else {
        switch (ng) {
            case 1: return launch_tiles<R, 1, false>(a, blocks, s);
            case 2: return launch_tiles<R, 2, false>(a, blocks, s);
            case 3: return launch_tiles<R, 3, false>(a, blocks, s);
            case 4: return launch_tiles<R, 4, false>(a, blocks, s);
        }
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unsupported guest count");
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

void launch_grid(Framework& fw, const GridRequest& req) {
    Context& ctx = *fw.ctx;
    const int ng = (int)req.guests.size();
    if (ng > kMaxGuests) fail(ALEATOR_INVALID_ARGUMENT, "grid: at most 4 guest species per build");
    if (ng == 0 && !req.coulomb) fail(ALEATOR_INVALID_ARGUMENT, "grid: nothing to build");
    if (req.coulomb && !fw.charged)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: a Coulomb table needs a framework created with Ewald parameters");
    for (int a = 0; a < 3; ++a)
        if (req.n[a] < 1) fail(ALEATOR_INVALID_ARGUMENT, "grid: point counts must be >= 1");

    GridArgs a{};
    for (int i = 0; i < 3; ++i) {
        a.n[i] = req.n[i];
        for (int j = 0; j < 3; ++j) a.mi[3 * i + j] = fw.lat.m[i][j];
    }
    // grid lattice: the framework lattice or a primitive cell whose rows are the
    // framework rows divided by integers
    double gm[9];
    if (req.grid_m) {
        std::memcpy(gm, req.grid_m, sizeof(gm));
        for (int i = 0; i < 3; ++i) {
            const double ratio = norm3(&a.mi[3 * i]) / norm3(&gm[3 * i]);
            const double rep = std::round(ratio);
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
    for (int i = 0; i < 3; ++i) a.tiles[i] = 0;
    a.tiles[0] = (req.n[0] + kTX - 1) / kTX;
    a.tiles[1] = (req.n[1] + kTY - 1) / kTY;
    a.tiles[2] = (req.n[2] + kTZ - 1) / kTZ;
    const long long blocks = (long long)a.tiles[0] * a.tiles[1] * a.tiles[2];
    if (blocks > 0x7fffffff) fail(ALEATOR_INVALID_ARGUMENT, "grid: too many nodes");

    // tile radius: farthest node of a tile from its centre
    double radius = 0.0;
    for (int c = 0; c < 8; ++c) {
        const double d0 = ((c & 1) ? 0.5 : -0.5) * (kTX - 1) / req.n[0];
        const double d1 = ((c & 2) ? 0.5 : -0.5) * (kTY - 1) / req.n[1];
        const double d2 = ((c & 4) ? 0.5 : -0.5) * (kTZ - 1) / req.n[2];
        double v[3];
        for (int j = 0; j < 3; ++j) v[j] = d0 * gm[j] + d1 * gm[3 + j] + d2 * gm[6 + j];
        radius = std::max(radius, norm3(v));
    }
    const double rmax = std::max(ng > 0 ? fw.rc : 0.0, req.coulomb ? fw.rc_real : 0.0);
    const double sr = rmax + radius + 1e-6 * (1.0 + rmax);
    a.search2 = sr * sr;
    // the 8-image scheme needs every image within the search radius to have fractional
    // offsets inside (-1, 1): guaranteed when the search radius is below L_perp
    for (int i = 0; i < 3; ++i)
        if (sr >= fw.lat.lperp[i])
            fail(ALEATOR_INVALID_ARGUMENT, "grid: cutoff plus tile radius exceeds the perpendicular width");

    a.fx = fw.d_fx;
    a.fy = fw.d_fy;
    a.fz = fw.d_fz;
    a.q = fw.d_q;
    a.type = fw.d_type;
    a.natoms = (int)fw.n;
    a.ntypes = (int)fw.type_species.size();
    a.rc_lj = fw.rc;
    a.rc2_lj = fw.rc * fw.rc;
    a.rc2_real = fw.rc_real * fw.rc_real;
    a.alpha = fw.alpha;
    a.kc = kCoulomb;
    a.cap = req.cap;
    a.trunc = fw.trunc;
    a.kmax = fw.kmax;
    a.out = req.out_device;
    a.stride = (long long)req.n[0] * req.n[1] * req.n[2];

    // LJ constants per (guest slot, framework type)
    std::vector<double> tab(4 * kMaxGuests * kMaxTypes, 0.0);
    for (int g = 0; g < ng; ++g)
        for (size_t t = 0; t < fw.type_species.size(); ++t) {
            double s2, e4, v, dv;
            lj_pair_constants(fw, req.guests[g], fw.type_species[t], s2, e4, v, dv);
            tab[(0 * kMaxGuests + g) * kMaxTypes + t] = s2;
            tab[(1 * kMaxGuests + g) * kMaxTypes + t] = e4;
            tab[(2 * kMaxGuests + g) * kMaxTypes + t] = v;
            tab[(3 * kMaxGuests + g) * kMaxTypes + t] = dv;
        }

    // reciprocal tables
    const int nk = 2 * fw.kmax + 1;
    size_t off_tab = 0, off_c = 0, off_e0 = 0, off_e1 = 0, off_e2 = 0, off_t1 = 0, off_t2 = 0, total = 0;
    auto bump = [&](size_t& o, size_t bytes) {
        o = total;
        total += (bytes + 255) & ~size_t(255);
    };
    bump(off_tab, tab.size() * sizeof(double));
    std::vector<double2> cube, e0, e1, e2;
    if (req.coulomb) {
        cube.assign((size_t)nk * nk * nk, make_double2(0.0, 0.0));
        // C(n) = w(n) conj(S(n)) over the full k-set (S(-n) = conj S(n))
        const size_t nkt = fw.kset.size();
        const std::vector<double>& sre = fw.h_sre;
        const std::vector<double>& sim = fw.h_sim;
        for (size_t t = 0; t < nkt; ++t) {
            const bool low = (int)t < fw.nk_half;
            const size_t h = low ? t : nkt - 1 - t;
            const double re = sre[h], im = low ? sim[h] : -sim[h];
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
        bump(off_c, cube.size() * sizeof(double2));
        bump(off_e0, e0.size() * sizeof(double2));
        bump(off_e1, e1.size() * sizeof(double2));
        bump(off_e2, e2.size() * sizeof(double2));
        bump(off_t1, (size_t)nk * nk * req.n[2] * sizeof(double2));
        bump(off_t2, (size_t)nk * req.n[2] * req.n[1] * sizeof(double2));
        a.recip_pref = kCoulomb * (4.0 * kPi / std::abs(fw.lat.volume));
    }
    char* base = static_cast<char*>(ctx.ensure_scratch(total));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_tab, tab.data(), tab.size() * sizeof(double),
                                     hipMemcpyHostToDevice, ctx.stream));
    a.ljtab = reinterpret_cast<const double*>(base + off_tab);
    if (req.coulomb) {
        ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_c, cube.data(), cube.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, ctx.stream));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_e0, e0.data(), e0.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, ctx.stream));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_e1, e1.data(), e1.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, ctx.stream));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_e2, e2.data(), e2.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, ctx.stream));
        auto* d_c = reinterpret_cast<const double2*>(base + off_c);
        auto* d_e1 = reinterpret_cast<const double2*>(base + off_e1);
        auto* d_e2 = reinterpret_cast<const double2*>(base + off_e2);
        auto* d_t1 = reinterpret_cast<double2*>(base + off_t1);
        auto* d_t2 = reinterpret_cast<double2*>(base + off_t2);
        const long long na = (long long)nk * nk * req.n[2];
        hipLaunchKernelGGL(recip_stage_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, ctx.stream,
                           d_c, d_e2, nk, req.n[2], d_t1);
        const long long nb = (long long)nk * req.n[2] * req.n[1];
        hipLaunchKernelGGL(recip_stage_b, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx.stream,
                           d_t1, d_e1, nk, req.n[1], req.n[2], d_t2);
        a.t2 = d_t2;
        a.e0 = reinterpret_cast<const double2*>(base + off_e0);
    }
    if (req.precision != ALEATOR_PRECISION_FP64 && req.precision != ALEATOR_PRECISION_MIXED)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown precision");
    ctx.time_begin();
    if (req.precision == ALEATOR_PRECISION_FP64)
        dispatch<double>(a, ng, req.coulomb, (int)blocks, ctx.stream);
    else
        dispatch<float>(a, ng, req.coulomb, (int)blocks, ctx.stream);
    ctx.time_end();
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
