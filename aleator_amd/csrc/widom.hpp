// widom.hpp -- batched Widom insertion: shared arguments and the direct-sum kernels (the host side,
// the trial sort and the reciprocal kernel are in widom.hip; widom_fp64.hip and widom_mixed.hip
// instantiate the kernels of each precision).
//
// Reference: MonteCarloEngine::widomInsertionHenryCoefficient
// (src/engines/monte_carlo/monte_carlo_engine.cc:850-917).  Trial t draws, from
// Philox4x32Rng(seed, stream) at call index first + t*d, three uniforms for the reference point
// a0*u0 + a1*u1 + a2*u2 (:882-884) and, for a multi-site molecule, three more for a Shoemake
// quaternion (:56-63, :886-889) that rotates the sites (:68-72).  dU = sum over sites of the LJ
// energy against every framework atom (LennardJones::computeParticleEnergy,
// lennard_jones.cc:158-182) plus, when the framework is charged, the real-space Ewald energy
// (ewald.cc:223-249) and EwaldIncrementalState::proposeInsertion (widom.hip).
//
// Pair arithmetic is FP64 (pair.hpp): 1/r^2 or 1/r refined by NR Newton steps (1: MIXED,
// 2: FP64), LJ as A/r^12 - (B/r^6 + shift [+ slope r]), erfc(alpha r)/r as 1/r minus the erf
// kernel polynomial of degree DEG (0: uncharged).
#pragma once

#include <cstdint>
#include <type_traits>

#include "device.hpp"
#include "internal.hpp"
#include "pair.hpp"

namespace aleator_gpu {
namespace widom_detail {

constexpr int kWThreads = 256;
constexpr int kWCap = 1024;   // LDS list capacity of the cube kernel (entries)
constexpr int kWChunk = 128;  // atoms tested per pass with 8 images each (<= kWCap)

struct WidomArgs {
    uint64_t seed, stream, first, n_trials;
    int draws;  // 3 or 6 per trial
    int nsites;
    double local[kMaxSites][3];
    double charge[kMaxSites];
    double m[9];    // lattice rows
    double inv[9];  // Cartesian -> fractional
    const double* fx;
    const double* fy;
    const double* fz;
    const double* q;
    const int32_t* type;
    int natoms;
    double rc2_lj, rc_lj, rc2_real, alpha, kc;
    int trunc;
    int charged;
    const double* ljtab;  // [4][kMaxSites][kMaxTypes]
    double temperature;
    double self_delta;  // -kC alpha/sqrt(pi) sum q^2 of the molecule
    // charged reciprocal part
    const int4* kn;
    const double* kw;
    const double* sre;
    const double* sim;
    int nk_half, kmax;
    double recip_scale;
    double* du;           // [n_trials] (LJ + real in the direct-sum kernel when charged)
    double* block_sums;   // per workgroup sum of exp(-dU/T)
    int split;            // direct-sum kernel: threads sharing one trial's atoms (1, 2, 4, 8 or 16)
    double poly[kErfPolyMax + 1];  // erf(alpha r)/r in x = r^2 poly_xs - 1 (pair.hpp)
    double poly_xs;
    double tail_delta;  // analytic LJ tail of one insertion (0 unless the framework asks for it)
    double box[3], inv_box[3];  // orthorhombic cells: edge lengths and their inverses
};

// Reference point and rotated sites of trial t, Cartesian, in the reference's
// arithmetic order (no contraction), and their fractional coordinates.
__device__ inline void trial_sites(const WidomArgs& a, uint64_t t, double cart[kMaxSites][3]) {
#pragma clang fp contract(off)
    const uint64_t call = a.first + t * (uint64_t)a.draws;
    const double u0 = philox_uniform(a.seed, a.stream, call);
    const double u1 = philox_uniform(a.seed, a.stream, call + 1);
    const double u2 = philox_uniform(a.seed, a.stream, call + 2);
    double ref[3];
    for (int j = 0; j < 3; ++j) ref[j] = a.m[j] * u0 + a.m[3 + j] * u1 + a.m[6 + j] * u2;
    if (a.nsites == 1) {
        for (int j = 0; j < 3; ++j) cart[0][j] = ref[j] + a.local[0][j];
        return;
    }
    const double v0 = philox_uniform(a.seed, a.stream, call + 3);
    const double v1 = philox_uniform(a.seed, a.stream, call + 4);
    const double v2 = philox_uniform(a.seed, a.stream, call + 5);
    const double r1 = sqrt(1.0 - v0), r2 = sqrt(v0);
    const double th1 = 2.0 * kPi * v1, th2 = 2.0 * kPi * v2;
    const double qw = r2 * cos(th2), qx = r1 * sin(th1), qy = r1 * cos(th1), qz = r2 * sin(th2);
    for (int s = 0; s < a.nsites; ++s) {
        const double vx = a.local[s][0], vy = a.local[s][1], vz = a.local[s][2];
        // t = 2 (qv x v); v' = v + w t + qv x t
        const double tx = (qy * vz - qz * vy) * 2.0;
        const double ty = (qz * vx - qx * vz) * 2.0;
        const double tz = (qx * vy - qy * vx) * 2.0;
        const double rx = (vx + tx * qw) + (qy * tz - qz * ty);
        const double ry = (vy + ty * qw) + (qz * tx - qx * tz);
        const double rz = (vz + tz * qw) + (qx * ty - qy * tx);
        cart[s][0] = ref[0] + rx;
        cart[s][1] = ref[1] + ry;
        cart[s][2] = ref[2] + rz;
    }
}

__device__ inline void to_frac(const WidomArgs& a, const double c[3], double f[3]) {
    for (int j = 0; j < 3; ++j) f[j] = c[0] * a.inv[j] + c[1] * a.inv[3 + j] + c[2] * a.inv[6 + j];
}

struct CellArgs {
    int nc[3];         // cubes per Cartesian axis
    double lo[3];      // bounding-box corner
    double edge;       // cube edge
    double rlist2;     // R^2
    int single_image;  // 2R < min L_perp: only the nearest image of an atom can be within R
    const uint32_t* start;  // [n_cells + 1]
    const uint32_t* order;  // trial indices grouped by cell
};

__device__ inline uint32_t trial_cell(const WidomArgs& a, const CellArgs& c, uint64_t t) {
#pragma clang fp contract(off)
    const uint64_t call = a.first + t * (uint64_t)a.draws;
    const double u0 = philox_uniform(a.seed, a.stream, call);
    const double u1 = philox_uniform(a.seed, a.stream, call + 1);
    const double u2 = philox_uniform(a.seed, a.stream, call + 2);
    int i[3];
    for (int j = 0; j < 3; ++j) {
        const double x = a.m[j] * u0 + a.m[3 + j] * u1 + a.m[6 + j] * u2;  // the trial's reference point
        i[j] = max(0, min((int)floor((x - c.lo[j]) / c.edge), c.nc[j] - 1));
    }
    return (uint32_t)((i[2] * c.nc[1] + i[1]) * c.nc[0] + i[0]);
}

// LJ constants per (site, framework type) in LDS: A = 4 eps sigma^12, B = 4 eps sigma^6,
// shift = V(rc) - rc V'(rc) (0 when truncated, V(rc) when shifted), slope = V'(rc)
// (linear-force-shifted only); and the erf polynomial, read back into vector registers.
template <int NS, int DEG>
__device__ inline void stage_constants(const WidomArgs& a, double4 (&tab)[NS][kMaxTypes], double* s_poly) {
    for (int i = threadIdx.x; i < NS * kMaxTypes; i += blockDim.x) {
        const int s = i / kMaxTypes, ty = i % kMaxTypes;
        const double s2 = a.ljtab[(0 * kMaxSites + s) * kMaxTypes + ty];
        const double e4 = a.ljtab[(1 * kMaxSites + s) * kMaxTypes + ty];
        const double s6 = s2 * s2 * s2;
        const double vrc = a.trunc != ALEATOR_LJ_TRUNCATED ? a.ljtab[(2 * kMaxSites + s) * kMaxTypes + ty] : 0.0;
        const double slope =
            a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? a.ljtab[(3 * kMaxSites + s) * kMaxTypes + ty] : 0.0;
        tab[s][ty] = make_double4(e4 * s6 * s6, e4 * s6, vrc - a.rc_lj * slope, slope);
    }
    if ((int)threadIdx.x <= DEG) s_poly[threadIdx.x] = DEG > 0 ? a.poly[threadIdx.x] : 0.0;
}

// One (site, atom image) pair at squared distance r2: LJ into lj, q_site q_atom erfc(alpha r)/r
// into real.  ONE: the LJ and real-space cutoffs are equal (or the molecule is uncharged).
template <int NR, int DEG, bool LFS, bool ONE>
__device__ inline void widom_pair(double r2, const double4& cg, double qq, double rc2, double rcr2,
                                  const double (&pc)[DEG + 1], double pxs, double& lj, double& real) {
    constexpr bool CH = DEG > 0;
    auto lj_term = [&](double ir2, double r) {
        const double ir6 = ir2 * ir2 * ir2;
        double w = fma(cg.y, ir6, cg.z);
        if constexpr (LFS) w = fma(r, cg.w, w);
        lj += fma(cg.x, ir6 * ir6, -w);
    };
    if constexpr (ONE) {
        if (r2 <= rc2) {
            if constexpr (!CH && !LFS) {
                lj_term(rcp_refined<NR>(r2), 0.0);
            } else {
                const double ir = rsq_refined<NR>(r2);
                lj_term(ir * ir, r2 * ir);
                if constexpr (CH) real = fma(qq, ir - erf_over_r_poly<DEG>(r2, pc, pxs), real);
            }
        }
    } else {
        const double ir = rsq_refined<NR>(r2);
        if (r2 <= rc2) lj_term(ir * ir, r2 * ir);
        if constexpr (CH)
            if (r2 <= rcr2) real = fma(qq, ir - erf_over_r_poly<DEG>(r2, pc, pxs), real);
    }
}

// Direct sum over every framework atom (small batches, or no usable cube grid).  The 256 threads
// of a workgroup take 256/split trials; the split threads of a trial share its atoms (atom j of
// a staged chunk to thread j mod split) and their partial sums meet in LDS in a fixed order, so
// a batch of 1e5 trials still fills the chip.  Atoms are staged through LDS in chunks of 256 as
// fractional coordinates; each pair folds the fractional difference to the nearest image (exact
// for pairs within a cutoff <= L_perp/2).
// ORTHO: orthorhombic cell; atoms and sites stay Cartesian and the nearest image is taken per
// axis, d - L rint(d / L) (exact below L_perp/2, as the fractional fold).
template <bool EXACT, int NS, int DEG, bool LFS, bool ORTHO>
__global__ __launch_bounds__(kWThreads) void widom_split_kernel(const WidomArgs a) {
    constexpr int NR = EXACT ? 2 : 1;
    constexpr bool CH = DEG > 0;
    __shared__ double sfx[kWThreads], sfy[kWThreads], sfz[kWThreads], sq[kWThreads];
    __shared__ int st[kWThreads];
    __shared__ double4 tab[NS][kMaxTypes];
    __shared__ double s_poly[DEG + 1];
    __shared__ double part_lj[kWThreads], part_re[kWThreads];
    __shared__ double red[kWThreads / 64];
    stage_constants<NS, DEG>(a, tab, s_poly);
    __syncthreads();
    double pc[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) pc[k] = s_poly[k];
    const double pxs = a.poly_xs;
    const int split = a.split, tpb = kWThreads / split;
    const int j = threadIdx.x % tpb, part = threadIdx.x / tpb;
    const uint64_t t = (uint64_t)blockIdx.x * tpb + j;
    const bool valid = t < a.n_trials;
    // the trial's sites (fractional, or Cartesian with ORTHO): drawn by the first of its split
    // threads and shared through LDS
    __shared__ double s_fs[kWThreads / 2][NS][3];
    double fs[NS][3];
    double qs[NS];
    auto draw = [&](double (&out)[NS][3]) {
        double cart[kMaxSites][3];
        trial_sites(a, t, cart);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if constexpr (ORTHO) {
                out[s][0] = cart[s][0];
                out[s][1] = cart[s][1];
                out[s][2] = cart[s][2];
            } else {
                to_frac(a, cart[s], out[s]);
            }
        }
    };
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        fs[s][0] = fs[s][1] = fs[s][2] = 0.0;
        qs[s] = a.charge[s];
    }
    if (split == 1) {
        if (valid) draw(fs);
    } else {
        if (valid && part == 0) draw(s_fs[j]);
        __syncthreads();
        if (valid)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                for (int k = 0; k < 3; ++k) fs[s][k] = s_fs[j][s][k];
    }
    double lj = 0.0, real = 0.0;
    const double rc2 = a.rc2_lj, rcr2 = a.rc2_real;
    const bool one = !CH || rc2 == rcr2;
    for (int base = 0; base < a.natoms; base += kWThreads) {
        __syncthreads();
        const int at = base + threadIdx.x;
        if (at < a.natoms) {
            if constexpr (ORTHO) {
                sfx[threadIdx.x] = a.fx[at] * a.box[0];
                sfy[threadIdx.x] = a.fy[at] * a.box[1];
                sfz[threadIdx.x] = a.fz[at] * a.box[2];
            } else {
                sfx[threadIdx.x] = a.fx[at];
                sfy[threadIdx.x] = a.fy[at];
                sfz[threadIdx.x] = a.fz[at];
            }
            sq[threadIdx.x] = a.q[at];
            st[threadIdx.x] = a.type[at];
        }
        __syncthreads();
        const int cnt = min(kWThreads, a.natoms - base);
        if (!valid) continue;
        for (int e = part; e < cnt; e += split) {
            const double ax = sfx[e], ay = sfy[e], az = sfz[e], qa = sq[e];
            const int ty = st[e];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double dx, dy, dz;
                if constexpr (ORTHO) {
                    dx = ax - fs[s][0];
                    dy = ay - fs[s][1];
                    dz = az - fs[s][2];
                    dx = fma(-a.box[0], rint(dx * a.inv_box[0]), dx);
                    dy = fma(-a.box[1], rint(dy * a.inv_box[1]), dy);
                    dz = fma(-a.box[2], rint(dz * a.inv_box[2]), dz);
                } else {
                    double d0 = ax - fs[s][0], d1 = ay - fs[s][1], d2 = az - fs[s][2];
                    d0 -= rint(d0);
                    d1 -= rint(d1);
                    d2 -= rint(d2);
                    dx = d0 * a.m[0] + d1 * a.m[3] + d2 * a.m[6];
                    dy = d0 * a.m[1] + d1 * a.m[4] + d2 * a.m[7];
                    dz = d0 * a.m[2] + d1 * a.m[5] + d2 * a.m[8];
                }
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                if (one)
                    widom_pair<NR, DEG, LFS, true>(r2, tab[s][ty], qs[s] * qa, rc2, rcr2, pc, pxs, lj, real);
                else
                    widom_pair<NR, DEG, LFS, false>(r2, tab[s][ty], qs[s] * qa, rc2, rcr2, pc, pxs, lj, real);
            }
        }
    }
    part_lj[threadIdx.x] = lj;
    part_re[threadIdx.x] = real;
    __syncthreads();
    double boltz = 0.0;
    if (part == 0 && valid) {
        double sl = 0.0, sr = 0.0;
        for (int p = 0; p < split; ++p) {
            sl += part_lj[j + p * tpb];
            sr += part_re[j + p * tpb];
        }
        const double du = (CH ? sl + a.kc * sr : sl) + a.tail_delta;
        a.du[t] = du;
        if (!CH) boltz = exp(-du / a.temperature);
    }
    if constexpr (!CH) {
        const double s = block_sum(boltz, red);
        if (threadIdx.x == 0) a.block_sums[blockIdx.x] = s;
    }
}

// Cube-sorted direct sum (large batches).  Trials are grouped by the cube of their reference
// point in a Cartesian grid of cubes over the lattice's bounding box (about one pass of a
// workgroup's trials per occupied cube).  One workgroup per cube builds, as the grid tile kernel
// does, an LDS list of the framework images within R = max cutoff + cube half-diagonal +
// molecule radius of the cube centre (atom order, block scan), and each thread sums TPT trials'
// sites against it.  Every image within the cutoff of a site lies in the list, and at most one
// image of an atom does (cutoff <= L_perp/2), so the sum is the reference's minimum-image sum,
// taken in atom order per trial.
template <bool EXACT, int NS, int DEG, bool LFS, int TPT>
__global__ __launch_bounds__(kWThreads) void widom_cell_kernel(const WidomArgs a, const CellArgs c) {
    constexpr int NR = EXACT ? 2 : 1;
    constexpr bool CH = DEG > 0;
    __shared__ double2 e_xy[kWCap];
    __shared__ double2 e_zq[kWCap];
    __shared__ int e_t[kWCap];
    __shared__ double4 tab[NS][kMaxTypes];
    __shared__ double s_poly[DEG + 1];
    __shared__ int scan_tmp[kWThreads / 64 + 1];
    const int tid = threadIdx.x;
    const int cell = blockIdx.x;
    const uint32_t lo = c.start[cell], hi = c.start[cell + 1];
    if (lo == hi) return;  // uniform over the workgroup

    stage_constants<NS, DEG>(a, tab, s_poly);
    __syncthreads();
    double pc[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) pc[k] = s_poly[k];
    const double pxs = a.poly_xs;

    // cube centre, Cartesian and fractional
    const int i0 = cell % c.nc[0], i1 = (cell / c.nc[0]) % c.nc[1], i2 = cell / (c.nc[0] * c.nc[1]);
    const double cc[3] = {c.lo[0] + (i0 + 0.5) * c.edge, c.lo[1] + (i1 + 0.5) * c.edge, c.lo[2] + (i2 + 0.5) * c.edge};
    double cf[3];
    to_frac(a, cc, cf);

    double qs[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) qs[s] = a.charge[s];
    const double rc2 = a.rc2_lj, rcr2 = a.rc2_real;
    const bool one = !CH || rc2 == rcr2;
    for (uint32_t pass = lo; pass < hi; pass += kWThreads * TPT) {
        uint64_t tr[TPT];
        bool ok[TPT];
        double ox[TPT][NS], oy[TPT][NS], oz[TPT][NS];
        double lj[TPT], real[TPT];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const uint32_t idx = pass + k * kWThreads + tid;
            ok[k] = idx < hi;
            tr[k] = ok[k] ? c.order[idx] : 0;
            lj[k] = 0.0;
            real[k] = 0.0;
            double cart[kMaxSites][3];
            if (ok[k]) trial_sites(a, tr[k], cart);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                ox[k][s] = ok[k] ? cart[s][0] - cc[0] : 1e6;
                oy[k][s] = ok[k] ? cart[s][1] - cc[1] : 1e6;
                oz[k][s] = ok[k] ? cart[s][2] - cc[2] : 1e6;
            }
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < TPT; ++k) any |= ok[k];

        auto compute = [&](int count) {
            __syncthreads();
            if (any) {
                auto pairs = [&](auto one_tag) {
                    constexpr bool ONE = decltype(one_tag)::value;
                    for (int e = 0; e < count; ++e) {
                        const double2 xy = e_xy[e], zq = e_zq[e];
                        const int ty = e_t[e];
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const double4 cg = tab[s][ty];
                            const double qq = qs[s] * zq.y;
#pragma unroll
                            for (int k = 0; k < TPT; ++k) {
                                const double dx = xy.x - ox[k][s], dy = xy.y - oy[k][s], dz = zq.x - oz[k][s];
                                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                                widom_pair<NR, DEG, LFS, ONE>(r2, cg, qq, rc2, rcr2, pc, pxs, lj[k], real[k]);
                            }
                        }
                    }
                };
                if (one) pairs(std::true_type{});
                else pairs(std::false_type{});
            }
            __syncthreads();
        };

        // atoms per pass: 128 (8 images each <= kWCap), or 1024 (4 per thread, contiguous) when only
        // the nearest image of an atom can be within R
        const int chunk = c.single_image ? 4 * kWThreads : kWChunk;
        const int apt = c.single_image ? 4 : 1;
        int pending = 0;
        for (int base = 0; base < a.natoms; base += chunk) {
            int cnt = 0;
            unsigned mask[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int at = base + tid * apt + j;
                if (j >= apt || tid * apt >= chunk || at >= a.natoms) continue;
                double h0 = a.fx[at] - cf[0], h1 = a.fy[at] - cf[1], h2 = a.fz[at] - cf[2];
                h0 -= floor(h0);
                h1 -= floor(h1);
                h2 -= floor(h2);
                const int near = (h0 >= 0.5 ? 1 : 0) | (h1 >= 0.5 ? 2 : 0) | (h2 >= 0.5 ? 4 : 0);
                for (int sh = 0; sh < 8; ++sh) {
                    if (c.single_image && sh != near) continue;
                    const double g0 = h0 - (sh & 1), g1 = h1 - ((sh >> 1) & 1), g2 = h2 - (sh >> 2);
                    const double x = g0 * a.m[0] + g1 * a.m[3] + g2 * a.m[6];
                    const double y = g0 * a.m[1] + g1 * a.m[4] + g2 * a.m[7];
                    const double z = g0 * a.m[2] + g1 * a.m[5] + g2 * a.m[8];
                    if (x * x + y * y + z * z <= c.rlist2) {
                        mask[j] |= 1u << sh;
                        ++cnt;
                    }
                }
            }
            int total;
            const int off = block_exclusive_scan(cnt, scan_tmp, &total);
            if (pending + total > kWCap) {
                compute(pending);
                pending = 0;
            }
            int w = pending + off;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!mask[j]) continue;
                const int at = base + tid * apt + j;
                const double qa = a.q[at];
                const int ta = a.type[at];
                double h0 = a.fx[at] - cf[0], h1 = a.fy[at] - cf[1], h2 = a.fz[at] - cf[2];
                h0 -= floor(h0);
                h1 -= floor(h1);
                h2 -= floor(h2);
                for (int sh = 0; sh < 8; ++sh) {
                    if (!(mask[j] & (1u << sh))) continue;
                    const double g0 = h0 - (sh & 1), g1 = h1 - ((sh >> 1) & 1), g2 = h2 - (sh >> 2);
                    e_xy[w] = make_double2(g0 * a.m[0] + g1 * a.m[3] + g2 * a.m[6],
                                           g0 * a.m[1] + g1 * a.m[4] + g2 * a.m[7]);
                    e_zq[w] = make_double2(g0 * a.m[2] + g1 * a.m[5] + g2 * a.m[8], qa);
                    e_t[w] = ta;
                    ++w;
                }
            }
            pending += total;
        }
        compute(pending);
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            if (ok[k]) a.du[tr[k]] = (CH ? lj[k] + a.kc * real[k] : lj[k]) + a.tail_delta;
    }
}

// Trials per thread of the cube kernel: 1, or (for large batches) 4 for one site and 2 otherwise.
inline int max_trials_per_thread(int ns) { return ns == 1 ? 4 : 2; }

template <bool EXACT, int DEG, bool LFS>
void launch_cells_deg(const WidomArgs& a, const CellArgs& c, unsigned n_cells, bool wide, hipStream_t s) {
    switch (a.nsites * 2 + (wide ? 1 : 0)) {
#define ALEATOR_WCELL(N, T)                                                                                     \
    case 2 * N + (T > 1 ? 1 : 0):                                                                               \
        hipLaunchKernelGGL((widom_cell_kernel<EXACT, N, DEG, LFS, T>), dim3(n_cells), dim3(kWThreads), 0, s, a, c); \
        return;
        ALEATOR_WCELL(1, 1) ALEATOR_WCELL(1, 4) ALEATOR_WCELL(2, 1) ALEATOR_WCELL(2, 2) ALEATOR_WCELL(3, 1)
        ALEATOR_WCELL(3, 2) ALEATOR_WCELL(4, 1) ALEATOR_WCELL(4, 2) ALEATOR_WCELL(5, 1) ALEATOR_WCELL(5, 2)
#undef ALEATOR_WCELL
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

template <bool EXACT, int DEG, bool LFS>
void launch_split_deg(const WidomArgs& a, unsigned blocks, bool ortho, hipStream_t s) {
    switch (a.nsites) {
#define ALEATOR_WSPLIT(N)                                                                                         \
    case N:                                                                                                       \
        if (ortho)                                                                                                \
            hipLaunchKernelGGL((widom_split_kernel<EXACT, N, DEG, LFS, true>), dim3(blocks), dim3(kWThreads), 0, s, \
                               a);                                                                                \
        else                                                                                                      \
            hipLaunchKernelGGL((widom_split_kernel<EXACT, N, DEG, LFS, false>), dim3(blocks), dim3(kWThreads), 0, \
                               s, a);                                                                             \
        return;
        ALEATOR_WSPLIT(1) ALEATOR_WSPLIT(2) ALEATOR_WSPLIT(3) ALEATOR_WSPLIT(4) ALEATOR_WSPLIT(5)
#undef ALEATOR_WSPLIT
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

// deg: erf polynomial degree (0 when the molecule sees no charges)
template <bool EXACT, bool LFS>
void launch_direct(const WidomArgs& a, const CellArgs* c, unsigned n_cells, bool wide, unsigned blocks, int deg,
                   hipStream_t s) {
    auto go = [&](auto deg_tag) {
        constexpr int D = decltype(deg_tag)::value;
        if (c) launch_cells_deg<EXACT, D, LFS>(a, *c, n_cells, wide, s);
        else launch_split_deg<EXACT, D, LFS>(a, blocks, a.box[0] > 0.0, s);
    };
    switch (deg) {
        case 0: return go(std::integral_constant<int, 0>{});
        case 18: return go(std::integral_constant<int, 18>{});
        case 24: return go(std::integral_constant<int, 24>{});
        case kErfPolyMax: return go(std::integral_constant<int, kErfPolyMax>{});
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: unsupported erf polynomial degree");
}

}  // namespace widom_detail

// The direct-sum kernels of each precision (widom_fp64.hip, widom_mixed.hip): the cube kernel
// when c is given, else the split kernel over `blocks` workgroups.
void widom_direct_fp64(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                       bool wide, unsigned blocks, int deg, bool lfs, hipStream_t s);
void widom_direct_mixed(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells,
                        bool wide, unsigned blocks, int deg, bool lfs, hipStream_t s);

}  // namespace aleator_gpu
