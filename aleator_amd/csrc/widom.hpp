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

#include <algorithm>
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
    int geo;  // split kernel: 0 triclinic, 1 orthorhombic, 2 cubic (kGeo*)
};

// Cell geometry of the split kernel's minimum image: triclinic cells fold fractional differences
// and map them through the lattice; orthorhombic cells take d - L rint(d / L) per Cartesian axis;
// cubic cells fold fractional differences and scale |f|^2 by L^2 (two instructions fewer).
constexpr int kGeoTriclinic = 0, kGeoOrtho = 1, kGeoCubic = 2;

// Reference point and rotated sites of trial t, Cartesian, in the reference's
// arithmetic order (no contraction), and their fractional coordinates.
template <int NS>
__device__ inline void trial_sites(const WidomArgs& a, uint64_t t, double (&cart)[NS][3]) {
#pragma clang fp contract(off)
    const uint64_t call = a.first + t * (uint64_t)a.draws;
    const double u0 = philox_uniform(a.seed, a.stream, call);
    const double u1 = philox_uniform(a.seed, a.stream, call + 1);
    const double u2 = philox_uniform(a.seed, a.stream, call + 2);
    double ref[3];
    for (int j = 0; j < 3; ++j) ref[j] = a.m[j] * u0 + a.m[3 + j] * u1 + a.m[6 + j] * u2;
    if constexpr (NS == 1) {
        for (int j = 0; j < 3; ++j) cart[0][j] = ref[j] + a.local[0][j];
        return;
    }
    const double v0 = philox_uniform(a.seed, a.stream, call + 3);
    const double v1 = philox_uniform(a.seed, a.stream, call + 4);
    const double v2 = philox_uniform(a.seed, a.stream, call + 5);
    const double r1 = sqrt(1.0 - v0), r2 = sqrt(v0);
    const double th1 = 2.0 * kPi * v1, th2 = 2.0 * kPi * v2;
    const double qw = r2 * cos(th2), qx = r1 * sin(th1), qy = r1 * cos(th1), qz = r2 * sin(th2);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
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
// into real.  ONE: the LJ and real-space cutoffs are equal (or the molecule is uncharged).  LJK:
// LJ truncation kind (0 truncated, 1 shifted, 2 linear-force-shifted); V = r^-6 (A r^-6 - B)
// [- shift [- slope r]] from cg = (A, B, shift, slope).
template <int NR, int DEG, int LJK, bool ONE>
__device__ inline void widom_pair(double r2, const double4& cg, double qq, double rc2, double rcr2,
                                  const double (&pc)[DEG + 1], double pxs, double& lj, double& real) {
    constexpr bool CH = DEG > 0;
    constexpr bool LFS = LJK == 2;
    auto lj_term = [&](double ir2, double r) {
        const double ir6 = ir2 * ir2 * ir2;
        lj = fma(ir6, fma(cg.x, ir6, -cg.y), lj);
        if constexpr (LJK >= 1) lj -= cg.z;
        if constexpr (LFS) lj = fma(-r, cg.w, lj);
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

// Trials per lane of the split kernel: two for one-site molecules (two independent dependency
// chains per atom read), one otherwise (the sites already give that).
template <int NS>
constexpr int split_tpt() { return NS == 1 ? 2 : 1; }

// Direct sum over every framework atom (small and mid-size batches, or no usable cube grid).  A
// workgroup of 64 * split * G threads (G = max(1, 4 / split)) takes 64 G TPT trials: wave w holds
// trial group w / split and part w % split, lane l the TPT trials l, l + 64, ... of its group, so
// the loop over atoms is uniform across a wave (part p takes atoms p, p + split, ... of each staged
// pass, read from LDS by every lane at once).  The parts of a trial meet in LDS in a fixed order.
// Atoms are staged through LDS 512 at a time, with each site's LJ constants for the atom's type
// beside them (no per-pair type lookup).  Each pair takes the nearest image as a.geo says (exact
// for pairs within a cutoff <= L_perp/2).  Slot k of a trial group is drawn by part k mod split
// and shared through LDS.  Uncharged trials in a cubic cell work in fractional units: r_f^2 =
// |f|^2 is compared with rc^2 / L^2 and the LJ constants carry the powers of L (A / L^12, B / L^6,
// the slope times L).  Uncharged batches leave one sum of exp(-dU/T) per workgroup.
template <bool EXACT, int NS, int DEG, int LJK>
__global__ __launch_bounds__(1024) void widom_split_kernel(const WidomArgs a) {
    constexpr int NR = EXACT ? 2 : 1;
    constexpr bool CH = DEG > 0;
    constexpr int TPT = split_tpt<NS>();
    constexpr int kStage = 512;  // atoms staged per pass (one pass for up to 512 framework atoms)
    // one LDS area holds, in turn, the staged atoms and the parts' partial sums
    constexpr size_t kStaged = sizeof(double) * kStage * (4 + 2 * NS + (LJK > 0 ? 2 * NS : 0));
    constexpr size_t kParts = sizeof(double) * 1024 * TPT * (CH ? 2 : 1);
    constexpr size_t kArea = kStaged > kParts ? kStaged : kParts;
    __shared__ __align__(16) unsigned char s_area[kArea];
    double* sfx = reinterpret_cast<double*>(s_area);
    double* sfy = sfx + kStage;
    double* sfz = sfy + kStage;
    double* sq = sfz + kStage;
    auto* sab = reinterpret_cast<double2(*)[kStage]>(sq + kStage);  // [NS]: A, B of (site, atom)
    auto* scd = sab + NS;                                             // [NS]: shift, slope (LJK > 0)
    double* part_lj = reinterpret_cast<double*>(s_area);              // [TPT][1024]
    double* part_re = part_lj + 1024 * TPT;
    __shared__ double4 tab[NS][kMaxTypes];
    __shared__ double s_poly[DEG + 1];
    __shared__ double red[16];
    const int tid = threadIdx.x, lane = tid & 63;
    const int geo = a.geo;
    const bool frac_units = !CH && geo == kGeoCubic;
    // the first pass's atoms, loaded before anything waits on memory
    auto load_atom = [&](int at, double& x, double& y, double& z, double& q, int& ty) {
        const double sc0 = geo == kGeoOrtho ? a.box[0] : 1.0;
        const double sc1 = geo == kGeoOrtho ? a.box[1] : 1.0;
        const double sc2 = geo == kGeoOrtho ? a.box[2] : 1.0;
        x = a.fx[at] * sc0;
        y = a.fy[at] * sc1;
        z = a.fz[at] * sc2;
        q = a.q[at];
        ty = a.type[at];
    };
    double ax0 = 0.0, ay0 = 0.0, az0 = 0.0, aq0 = 0.0;
    int aty0 = 0;
    const bool stager = tid < kStage;
    if (stager && tid < a.natoms) load_atom(tid, ax0, ay0, az0, aq0, aty0);
    stage_constants<NS, DEG>(a, tab, s_poly);
    const int split = a.split;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int part = wave % split, group = wave / split;
    const int groups = blockDim.x / (64 * split);
    // slot k of a lane: trial (group * TPT + k) * 64 + lane of this workgroup
    auto local = [&](int k) { return (group * TPT + k) * 64 + lane; };
    auto trial = [&](int k) { return (uint64_t)blockIdx.x * (groups * 64 * TPT) + local(k); };
    // trial sites (fractional, or Cartesian for orthorhombic cells): with one part a lane draws its
    // own; with more, slot k of a group is drawn once, by the group's part k mod split, and shared
    // through LDS (at most two groups of 64 x TPT slots per workgroup then)
    __shared__ double s_fs[128 * TPT][NS][3];
    double fs[TPT][NS][3];
    auto draw = [&](int k, double (&out)[NS][3]) {
        double cart[NS][3];
        trial_sites<NS>(a, trial(k) < a.n_trials ? trial(k) : 0, cart);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (geo == kGeoOrtho) {
                out[s][0] = cart[s][0];
                out[s][1] = cart[s][1];
                out[s][2] = cart[s][2];
            } else {
                to_frac(a, cart[s], out[s]);
            }
        }
    };
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (split == 1) draw(k, fs[k]);
        else if (k % split == part) draw(k, s_fs[local(k)]);
    }
    __syncthreads();  // tab, s_fs
    if (split > 1)
#pragma unroll
        for (int k = 0; k < TPT; ++k)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                for (int c = 0; c < 3; ++c) fs[k][s][c] = s_fs[local(k)][s][c];
    double pc[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) pc[k] = s_poly[k];
    const double pxs = a.poly_xs;
    double qs[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) qs[s] = a.charge[s];
    double lj[TPT], real[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) lj[k] = real[k] = 0.0;
    const double l2 = a.box[0] * a.box[0];  // cubic cells
    const double il2 = 1.0 / l2;
    const double rc2 = frac_units ? a.rc2_lj * il2 : a.rc2_lj, rcr2 = a.rc2_real;
    const bool one = !CH || a.rc2_lj == rcr2;
    // LJ constants in the units of r^2: A L^-12, B L^-6, slope L in fractional units of a cubic cell
    const double sa = frac_units ? il2 * il2 * il2 * il2 * il2 * il2 : 1.0;
    const double sb = frac_units ? il2 * il2 * il2 : 1.0;
    const double sl = frac_units ? a.box[0] : 1.0;
    // the pairs of staged atoms part, part + split, ... < cnt with this lane's trials, minimum
    // image by geometry G
    auto pairs = [&](auto geo_tag, auto one_tag, int cnt) {
        constexpr int G = decltype(geo_tag)::value;
        constexpr bool ONE = decltype(one_tag)::value;
        for (int e = part; e < cnt; e += split) {
            const double ax = sfx[e], ay = sfy[e], az = sfz[e], qa = sq[e];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double2 ab = sab[s][e];
                double2 cd = make_double2(0.0, 0.0);
                if constexpr (LJK > 0) cd = scd[s][e];
                const double4 cg = make_double4(ab.x, ab.y, cd.x, cd.y);
                const double qq = qs[s] * qa;
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    double r2;
                    double d0 = ax - fs[k][s][0], d1 = ay - fs[k][s][1], d2 = az - fs[k][s][2];
                    if constexpr (G == kGeoOrtho) {
                        d0 = fma(-a.box[0], rint(d0 * a.inv_box[0]), d0);
                        d1 = fma(-a.box[1], rint(d1 * a.inv_box[1]), d1);
                        d2 = fma(-a.box[2], rint(d2 * a.inv_box[2]), d2);
                        r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
                    } else {
                        d0 -= rint(d0);
                        d1 -= rint(d1);
                        d2 -= rint(d2);
                        if constexpr (G == kGeoCubic) {
                            r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
                            if constexpr (CH) r2 *= l2;
                        } else {
                            const double dx = d0 * a.m[0] + d1 * a.m[3] + d2 * a.m[6];
                            const double dy = d0 * a.m[1] + d1 * a.m[4] + d2 * a.m[7];
                            const double dz = d0 * a.m[2] + d1 * a.m[5] + d2 * a.m[8];
                            r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                        }
                    }
                    widom_pair<NR, DEG, LJK, ONE>(r2, cg, qq, rc2, rcr2, pc, pxs, lj[k], real[k]);
                }
            }
        }
    };
    for (int base = 0; base < a.natoms; base += kStage) {
        const int at = base + tid;
        if (base > 0) {
            __syncthreads();  // the previous pass's atoms are no longer read
            if (stager && at < a.natoms) load_atom(at, ax0, ay0, az0, aq0, aty0);
        }
        if (stager && at < a.natoms) {
            sfx[tid] = ax0;
            sfy[tid] = ay0;
            sfz[tid] = az0;
            sq[tid] = aq0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double4 c = tab[s][aty0];
                sab[s][tid] = make_double2(c.x * sa, c.y * sb);
                if constexpr (LJK > 0) scd[s][tid] = make_double2(c.z, c.w * sl);
            }
        }
        __syncthreads();
        const int cnt = min(kStage, a.natoms - base);
        if (geo == kGeoCubic) {
            if (one) pairs(std::integral_constant<int, kGeoCubic>{}, std::true_type{}, cnt);
            else pairs(std::integral_constant<int, kGeoCubic>{}, std::false_type{}, cnt);
        } else if (geo == kGeoOrtho) {
            if (one) pairs(std::integral_constant<int, kGeoOrtho>{}, std::true_type{}, cnt);
            else pairs(std::integral_constant<int, kGeoOrtho>{}, std::false_type{}, cnt);
        } else {
            if (one) pairs(std::integral_constant<int, kGeoTriclinic>{}, std::true_type{}, cnt);
            else pairs(std::integral_constant<int, kGeoTriclinic>{}, std::false_type{}, cnt);
        }
    }
    __syncthreads();  // the staged atoms are no longer read
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        part_lj[k * 1024 + tid] = lj[k];
        if constexpr (CH) part_re[k * 1024 + tid] = real[k];
    }
    __syncthreads();
    double boltz = 0.0;
    if (part == 0)
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            if (trial(k) >= a.n_trials) continue;
            double sl_ = 0.0, sr = 0.0;
            for (int p = 0; p < split; ++p) {
                const int i = k * 1024 + (group * split + p) * 64 + lane;
                sl_ += part_lj[i];
                if constexpr (CH) sr += part_re[i];
            }
            const double du = (CH ? sl_ + a.kc * sr : sl_) + a.tail_delta;
            a.du[trial(k)] = du;
            if (!CH) boltz += exp(-du / a.temperature);
        }
    if constexpr (!CH) {
        const double bs = block_sum(boltz, red);
        if (tid == 0) a.block_sums[blockIdx.x] = bs;
    }
}

// Cube-sorted direct sum (large batches).  Trials are grouped by the cube of their reference
// point in a Cartesian grid of cubes over the lattice's bounding box (about one pass of a
// workgroup's trials per occupied cube).  One workgroup per cube builds, as the grid tile kernel
// does, an LDS list of the framework images within R = max cutoff + cube half-diagonal +
// molecule radius of the cube centre (atom order, block scan), and each thread sums TPT trials'
// sites against it (with one trial per thread, a cube of at most 128 or 64 trials splits each
// trial's list between 2 or 4 threads).  Every image within the cutoff of a site lies in the list,
// and at most one image of an atom does (cutoff <= L_perp/2), so the sum is the reference's
// minimum-image sum, taken in atom order per trial (per part, with the parts added in order).
template <bool EXACT, int NS, int DEG, int LJK, int TPT>
__global__ __launch_bounds__(kWThreads) void widom_cell_kernel(const WidomArgs a, const CellArgs c) {
    constexpr int NR = EXACT ? 2 : 1;
    constexpr bool CH = DEG > 0;
    __shared__ double2 e_xy[kWCap];
    __shared__ double2 e_zq[kWCap];
    __shared__ int e_t[kWCap];
    __shared__ double4 tab[NS][kMaxTypes];
    __shared__ double s_poly[DEG + 1];
    __shared__ int scan_tmp[kWThreads / 64 + 1];
    __shared__ double part_lj[TPT == 1 ? kWThreads : 1], part_re[TPT == 1 && CH ? kWThreads : 1];
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
    // a sparsely filled cube (one trial per thread) shares each trial's list among sp threads:
    // slot tid % slots, part tid / slots takes entries part, part + sp, ... (a part spans whole
    // waves), and the parts meet in LDS in a fixed order
    const uint32_t n_cube = hi - lo;
    const int sp = TPT == 1 ? (n_cube <= 64 ? 4 : (n_cube <= 128 ? 2 : 1)) : 1;
    const int slots = kWThreads / sp;
    const int slot = tid % slots, part = tid / slots;
    for (uint32_t pass = lo; pass < hi; pass += slots * TPT) {
        uint64_t tr[TPT];
        bool ok[TPT];
        double ox[TPT][NS], oy[TPT][NS], oz[TPT][NS];
        double lj[TPT], real[TPT];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const uint32_t idx = pass + k * slots + slot;
            ok[k] = idx < hi;
            tr[k] = ok[k] ? c.order[idx] : 0;
            lj[k] = 0.0;
            real[k] = 0.0;
            double cart[NS][3];
            if (ok[k]) trial_sites<NS>(a, tr[k], cart);
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
                    for (int e = part; e < count; e += sp) {
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
                                widom_pair<NR, DEG, LJK, ONE>(r2, cg, qq, rc2, rcr2, pc, pxs, lj[k], real[k]);
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
        if (sp > 1) {  // uniform over the workgroup
            part_lj[tid] = lj[0];
            if constexpr (CH) part_re[tid] = real[0];
            __syncthreads();
            if (part == 0)
                for (int p = 1; p < sp; ++p) {
                    lj[0] += part_lj[p * slots + slot];
                    if constexpr (CH) real[0] += part_re[p * slots + slot];
                }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            if (ok[k] && part == 0) a.du[tr[k]] = (CH ? lj[k] + a.kc * real[k] : lj[k]) + a.tail_delta;
    }
}

// Trials per thread of the cube kernel: 1, or (for large batches) 4 for one site.  Multi-site
// molecules keep one (smaller cubes: at c1's 1e6 CO2 trials, 256 trials per cube ran 11 % faster
// than 512 at two per thread, the list radius shrinking from about 16.1 to 15.5 A).
inline int max_trials_per_thread(int ns) { return ns == 1 ? 4 : 1; }

template <bool EXACT, int DEG, int LJK>
void launch_cells_deg(const WidomArgs& a, const CellArgs& c, unsigned n_cells, bool wide, hipStream_t s) {
    switch (a.nsites * 2 + (wide ? 1 : 0)) {
#define ALEATOR_WCELL(N, T)                                                                                     \
    case 2 * N + (T > 1 ? 1 : 0):                                                                               \
        hipLaunchKernelGGL((widom_cell_kernel<EXACT, N, DEG, LJK, T>), dim3(n_cells), dim3(kWThreads), 0, s, a, c); \
        return;
        ALEATOR_WCELL(1, 1) ALEATOR_WCELL(1, 4) ALEATOR_WCELL(2, 1) ALEATOR_WCELL(3, 1) ALEATOR_WCELL(4, 1)
        ALEATOR_WCELL(5, 1)
#undef ALEATOR_WCELL
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

// threads of a split-kernel workgroup: 64 lanes per group, split waves per group, at least 4 waves
inline int split_block_threads(int split) { return 64 * split * std::max(1, 4 / split); }
// trials per lane of the split kernel (split_tpt) for n sites
inline int split_trials_per_lane(int ns) { return ns == 1 ? 2 : 1; }

template <bool EXACT, int DEG, int LJK>
void launch_split_deg(const WidomArgs& a, unsigned blocks, hipStream_t s) {
    const unsigned threads = (unsigned)split_block_threads(a.split);
    switch (a.nsites) {
#define ALEATOR_WSPLIT(N)                                                                                         \
    case N:                                                                                                       \
        hipLaunchKernelGGL((widom_split_kernel<EXACT, N, DEG, LJK>), dim3(blocks), dim3(threads), 0, s, a);      \
        return;
        ALEATOR_WSPLIT(1) ALEATOR_WSPLIT(2) ALEATOR_WSPLIT(3) ALEATOR_WSPLIT(4) ALEATOR_WSPLIT(5)
#undef ALEATOR_WSPLIT
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

// deg: erf polynomial degree (0 when the molecule sees no charges)
template <bool EXACT, int LJK>
void launch_direct(const WidomArgs& a, const CellArgs* c, unsigned n_cells, bool wide, unsigned blocks, int deg,
                   hipStream_t s) {
    auto go = [&](auto deg_tag) {
        constexpr int D = decltype(deg_tag)::value;
        if (c) launch_cells_deg<EXACT, D, LJK>(a, *c, n_cells, wide, s);
        else launch_split_deg<EXACT, D, LJK>(a, blocks, s);
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

// The direct-sum kernels of each precision and LJ truncation kind ljk (widom_{fp64,mixed}_lj{0,1,2}
// .hip): the cube kernel when c is given, else the split kernel over `blocks` workgroups.
#define ALEATOR_WIDOM_DIRECT(NAME)                                                                       \
    void NAME(const widom_detail::WidomArgs& a, const widom_detail::CellArgs* c, unsigned n_cells, bool wide, \
              unsigned blocks, int deg, hipStream_t s);
ALEATOR_WIDOM_DIRECT(widom_direct_fp64_lj0)
ALEATOR_WIDOM_DIRECT(widom_direct_fp64_lj1)
ALEATOR_WIDOM_DIRECT(widom_direct_fp64_lj2)
ALEATOR_WIDOM_DIRECT(widom_direct_mixed_lj0)
ALEATOR_WIDOM_DIRECT(widom_direct_mixed_lj1)
ALEATOR_WIDOM_DIRECT(widom_direct_mixed_lj2)
#undef ALEATOR_WIDOM_DIRECT

}  // namespace aleator_gpu
