// widom.hip -- batched Widom test-particle insertion on gfx950.
//
// Reference: MonteCarloEngine::widomInsertionHenryCoefficient
// (src/engines/monte_carlo/monte_carlo_engine.cc:850-917).  Trial t draws, from
// Philox4x32Rng(seed, stream) at call index first + t*d, three uniforms for the
// reference point a0*u0 + a1*u1 + a2*u2 (:882-884) and, for a multi-site molecule,
// three more for a Shoemake quaternion (:56-63, :886-889) that rotates the sites
// (:68-72).  dU = sum over sites of the LJ energy against every framework atom
// (LennardJones::computeParticleEnergy, lennard_jones.cc:158-182) plus, when the
// framework is charged, the real-space Ewald energy (ewald.cc:223-249) and
// EwaldIncrementalState::proposeInsertion (reciprocal + self + exclusion,
// ewald_incremental_state.cc:95-136).  K_H = <exp(-dU/T)>/T.
//
// Kernel 1 (one thread per trial): the framework is staged through LDS in chunks of
// 256 atoms as fractional coordinates; every lane reads the same atom (broadcast),
// folds the fractional difference to the nearest image (exact for pairs within a
// cutoff <= L_perp/2) and accumulates LJ and erfc terms for all sites.
// Kernel 2 (charged only; one wave per trial): per-axis phase factors of the sites in
// LDS, then the half k-set sum of recip.hpp.  Both write per-workgroup partial sums of
// exp(-dU/T), reduced in a fixed order by kernel 3.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "recip.hpp"

namespace aleator_gpu {

namespace {

constexpr int kWThreads = 256;

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
    double* du;           // [n_trials] (LJ + real in kernel 1 when charged)
    double* block_sums;   // per workgroup sum of exp(-dU/T)
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

template <typename R, int NS, bool CHARGED>
__global__ __launch_bounds__(kWThreads) void widom_pair_kernel(const WidomArgs a) {
    __shared__ double sfx[kWThreads], sfy[kWThreads], sfz[kWThreads], sq[kWThreads];
    __shared__ int st[kWThreads];
    __shared__ double tab[4][NS][kMaxTypes];
    __shared__ double red[kWThreads / 64];
    for (int i = threadIdx.x; i < 4 * NS * kMaxTypes; i += kWThreads) {
        const int k = i / (NS * kMaxTypes), rem = i % (NS * kMaxTypes);
        (&tab[0][0][0])[i] = a.ljtab[(k * kMaxSites + rem / kMaxTypes) * kMaxTypes + rem % kMaxTypes];
    }
    const uint64_t t = (uint64_t)blockIdx.x * kWThreads + threadIdx.x;
    const bool valid = t < a.n_trials;
    double cart[kMaxSites][3];
    double fs[NS][3];
    if (valid) {
        trial_sites(a, t, cart);
#pragma unroll
        for (int s = 0; s < NS; ++s) to_frac(a, cart[s], fs[s]);
    }
    double lj = 0.0, real = 0.0;
    const R rc2 = (R)a.rc2_lj, rcr2 = (R)a.rc2_real, alpha = (R)a.alpha;
    for (int base = 0; base < a.natoms; base += kWThreads) {
        __syncthreads();
        const int at = base + threadIdx.x;
        if (at < a.natoms) {
            sfx[threadIdx.x] = a.fx[at];
            sfy[threadIdx.x] = a.fy[at];
            sfz[threadIdx.x] = a.fz[at];
            sq[threadIdx.x] = a.q[at];
            st[threadIdx.x] = a.type[at];
        }
        __syncthreads();
        const int cnt = min(kWThreads, a.natoms - base);
        if (!valid) continue;
        for (int j = 0; j < cnt; ++j) {
            const double ax = sfx[j], ay = sfy[j], az = sfz[j];
            const int ty = st[j];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double d0 = ax - fs[s][0], d1 = ay - fs[s][1], d2 = az - fs[s][2];
                d0 -= rint(d0);
                d1 -= rint(d1);
                d2 -= rint(d2);
                const R dx = (R)(d0 * a.m[0] + d1 * a.m[3] + d2 * a.m[6]);
                const R dy = (R)(d0 * a.m[1] + d1 * a.m[4] + d2 * a.m[7]);
                const R dz = (R)(d0 * a.m[2] + d1 * a.m[5] + d2 * a.m[8]);
                const R r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= rc2) {
                    const R ir2 = (R)1 / r2;
                    const R s2 = (R)tab[0][s][ty] * ir2;
                    const R s6 = s2 * s2 * s2;
                    R v = (R)tab[1][s][ty] * (s6 * s6 - s6);
                    if (a.trunc == ALEATOR_LJ_SHIFTED) {
                        v -= (R)tab[2][s][ty];
                    } else if (a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) {
                        v = v - (R)tab[2][s][ty] - (sqrt(r2) - (R)a.rc_lj) * (R)tab[3][s][ty];
                    }
                    lj += (double)v;
                }
                if (CHARGED && r2 <= rcr2) {
                    const R r = sqrt(r2);
                    real += (double)((R)(a.charge[s] * sq[j]) * erfc_over_r<R>(alpha, r));
                }
            }
        }
    }
    double du = lj;
    if (CHARGED) du += a.kc * real;
    double boltz = 0.0;
    if (valid) {
        a.du[t] = du;
        if (!CHARGED) boltz = exp(-du / a.temperature);
    }
    if (!CHARGED) {
        const double s = block_sum(boltz, red);
        if (threadIdx.x == 0) a.block_sums[blockIdx.x] = s;
    }
}

// Reciprocal + self + exclusion, completing dU, for RT trials per workgroup (RT in {4, 8, 16}):
// the RG = 256 / RT consecutive threads of a trial split the runs of the half k-set between them
// (run r to thread r mod RG, in order) and add their partial sums by a fixed shuffle tree.  Per
// run and site, exp(i k.r) = E0[n0] E1[n1] E2[n2] is formed once from the trial's per-axis
// factors in LDS and then stepped along n2 by one complex multiply per term.  The k data
// (weight, S(k)) are read once per run by each trial's thread and shared through L1 by the
// workgroup's trials.
template <int NS>
__global__ __launch_bounds__(256) void widom_recip_runs_kernel(const WidomArgs a, const int2* __restrict__ kseg,
                                                               int n_seg, int rt) {
    extern __shared__ cplx wr_phase[];  // [rt][NS][3][2K + 1]
    __shared__ double s_f[16][NS][3];
    __shared__ double s_excl[16];
    __shared__ double red[16];
    const int rg = 256 / rt;
    const int j = threadIdx.x / rg, g = threadIdx.x % rg;
    const int K = a.kmax, nkk = 2 * K + 1;
    const uint64_t t = (uint64_t)blockIdx.x * rt + j;
    const bool valid = t < a.n_trials;
    if (g == 0 && valid) {
        double cart[kMaxSites][3];
        trial_sites(a, t, cart);
        for (int s = 0; s < NS; ++s) to_frac(a, cart[s], s_f[j][s]);
        // intramolecular exclusion: -kC sum_{i<j} q_i q_j erf(alpha r)/r
        // (ewald_incremental_state.cc:32-43); rigid-molecule distances are far below L_perp/2
        double sum = 0.0;
        for (int i = 0; i < NS; ++i)
            for (int k = i + 1; k < NS; ++k) {
                const double dx = cart[k][0] - cart[i][0], dy = cart[k][1] - cart[i][1], dz = cart[k][2] - cart[i][2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double eor = r < 1e-8 ? 2.0 * a.alpha / sqrt(kPi) : erf(a.alpha * r) / r;
                sum += a.charge[i] * a.charge[k] * eor;
            }
        s_excl[j] = -a.kc * sum;
    }
    __syncthreads();
    // per-axis factors exp(2 pi i m f), m = -K..K: one thread per (site, axis) of the trial
    if (valid && g < 3 * NS)
        phase_row(s_f[j][g / 3][g % 3], K, wr_phase + ((size_t)(j * NS + g / 3) * 3 + g % 3) * nkk);
    __syncthreads();
    double acc = 0.0;
    if (valid) {
        double q[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) q[s] = a.charge[s];
        acc = runs_recip_sum<NS>(wr_phase + (size_t)j * NS * 3 * nkk, q, K, kseg, n_seg, g, rg, a.kw, a.sre, a.sim);
    }
    for (int off = rg / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (g == 0) {
        double boltz = 0.0;
        if (valid) {
            const double du = a.du[t] + a.recip_scale * acc + a.self_delta + s_excl[j];
            a.du[t] = du;
            boltz = exp(-du / a.temperature);
        }
        red[j] = boltz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int i = 0; i < rt; ++i) sum += red[i];
        a.block_sums[blockIdx.x] = sum;
    }
}

// ---------------------------------------------------------------- cell-sorted direct sum
// Trials are grouped by the cell of their reference point in a Cartesian grid of cubes over
// the lattice's bounding box (about one pass of a workgroup's trials per occupied cube).  One workgroup per cell builds,
// as the grid tile kernel does, an LDS list of the framework images within
// R = max cutoff + cell half-diagonal + molecule radius of the cell centre (atom order, block
// scan), and each thread sums TPT trials' sites against it.  Every image within the cutoff of
// a site lies in the list, and at most one image of an atom does (cutoff <= L_perp/2), so the
// sum is the reference's minimum-image sum, taken in atom order per trial.
constexpr int kWCap = 1024;   // LDS list capacity (entries)
constexpr int kWChunk = 128;  // atoms tested per pass with 8 images each (<= kWCap)

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

__global__ void widom_key_kernel(const WidomArgs a, const CellArgs c, uint32_t* key, uint32_t* count) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_trials) return;
    const uint32_t k = trial_cell(a, c, t);
    key[t] = k;
    atomicAdd(&count[k], 1u);
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// start[k] = sum of count[<k]; cursor = start; start[n] = total.  One workgroup of 1024.
__global__ __launch_bounds__(1024) void scan_counts_kernel(const uint32_t* count, uint32_t n, uint32_t* start,
                                                           uint32_t* cursor) {
    __shared__ int tmp[17];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t lo = min(threadIdx.x * per, n), hi = min(lo + per, n);
    int s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += (int)count[i];
    int total;
    int run = block_exclusive_scan(s, tmp, &total);
    for (uint32_t i = lo; i < hi; ++i) {
        start[i] = (uint32_t)run;
        cursor[i] = (uint32_t)run;
        run += (int)count[i];
    }
    if (threadIdx.x == 0) start[n] = (uint32_t)total;
}

__global__ void scatter_kernel(const uint32_t* key, uint64_t n, uint32_t* cursor, uint32_t* order) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    order[atomicAdd(&cursor[key[t]], 1u)] = (uint32_t)t;
}

// Counting sort of the trials by cube with no global atomics (cube counts up to kSortMaxCells):
// block b takes trials [b per, (b + 1) per) and histograms their cubes in LDS; a column scan
// over blocks gives each (block, cube) its offset among the cube's trials; the block then places
// its trials, ranked inside (block, cube) by LDS atomics.
constexpr int kSortThreads = 1024;
constexpr uint32_t kSortMaxCells = 12288;  // 48 KB of LDS counters

__global__ __launch_bounds__(kSortThreads) void cube_hist_kernel(const WidomArgs a, const CellArgs c, uint64_t per,
                                                                 uint32_t n_cells, uint32_t* hist) {
    extern __shared__ uint32_t hcount[];
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hcount[i] = 0u;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(lo + per, a.n_trials);
    for (uint64_t t = lo + threadIdx.x; t < hi; t += kSortThreads) atomicAdd(&hcount[trial_cell(a, c, t)], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hist[(size_t)blockIdx.x * n_cells + i] = hcount[i];
}

// hist[b][cube] <- trials of the cube in blocks before b; count[cube] <- its total
__global__ void cube_colscan_kernel(uint32_t* hist, uint32_t n_blocks, uint32_t n_cells, uint32_t* count) {
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n_cells) return;
    uint32_t run = 0;
    for (uint32_t b = 0; b < n_blocks; ++b) {
        const size_t i = (size_t)b * n_cells + cell;
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
    count[cell] = run;
}

__global__ __launch_bounds__(kSortThreads) void cube_place_kernel(const WidomArgs a, const CellArgs c, uint64_t per,
                                                                  uint32_t n_cells, const uint32_t* hist,
                                                                  const uint32_t* start, uint32_t* order) {
    extern __shared__ uint32_t hrank[];
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hrank[i] = 0u;
    __syncthreads();
    const uint32_t* base = hist + (size_t)blockIdx.x * n_cells;
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(lo + per, a.n_trials);
    for (uint64_t t = lo + threadIdx.x; t < hi; t += kSortThreads) {
        const uint32_t k = trial_cell(a, c, t);
        order[start[k] + base[k] + atomicAdd(&hrank[k], 1u)] = (uint32_t)t;
    }
}

template <typename R>
struct WEntries {
    R x[kWCap], y[kWCap], z[kWCap], q[kWCap];
    int t[kWCap];
    __device__ void put(int w, double x_, double y_, double z_, double q_, int t_, bool) {
        x[w] = (R)x_;
        y[w] = (R)y_;
        z[w] = (R)z_;
        q[w] = (R)q_;
        t[w] = t_;
    }
};
// FP32: (x, y, z, q) packed in one float4, so an entry is two LDS reads; with_charge false: the
// fourth word carries the type bits instead of the charge, and an entry is one LDS read
template <>
struct WEntries<float> {
    float4 p[kWCap];
    int t[kWCap];
    __device__ void put(int w, double x_, double y_, double z_, double q_, int t_, bool with_charge) {
        p[w] = make_float4((float)x_, (float)y_, (float)z_, with_charge ? (float)q_ : __int_as_float(t_));
        t[w] = t_;
    }
};

// LFS: linear-force-shifted LJ (FP32 path: the slope term is compiled in only then).
template <typename R, int NS, bool CHARGED, bool LFS, int TPT>
__global__ __launch_bounds__(kWThreads) void widom_cell_kernel(const WidomArgs a, const CellArgs c) {
    constexpr bool kF32 = std::is_same_v<R, float>;
    __shared__ WEntries<R> ent;
    __shared__ double tab[4][NS][kMaxTypes];
    __shared__ float4 tabf[NS][kMaxTypes];  // A = 4 eps s^12, B = 4 eps s^6, V(rc) - rc slope, slope
    __shared__ float4 erfc_tab[(kF32 && CHARGED) ? kWThreads : 1];
    __shared__ int scan_tmp[kWThreads / 64 + 1];
    const int tid = threadIdx.x;
    const int cell = blockIdx.x;
    const uint32_t lo = c.start[cell], hi = c.start[cell + 1];
    if (lo == hi) return;  // uniform over the workgroup

    if constexpr (kF32) {
        for (int i = tid; i < NS * kMaxTypes; i += kWThreads) {
            const int s = i / kMaxTypes, ty = i % kMaxTypes;
            const double s2 = a.ljtab[(0 * kMaxSites + s) * kMaxTypes + ty];
            const double e4 = a.ljtab[(1 * kMaxSites + s) * kMaxTypes + ty];
            const double s6 = s2 * s2 * s2;
            const double vrc = a.trunc != ALEATOR_LJ_TRUNCATED ? a.ljtab[(2 * kMaxSites + s) * kMaxTypes + ty] : 0.0;
            const double slope =
                a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? a.ljtab[(3 * kMaxSites + s) * kMaxTypes + ty] : 0.0;
            tabf[s][ty] =
                make_float4((float)(e4 * s6 * s6), (float)(e4 * s6), (float)(vrc - a.rc_lj * slope), (float)slope);
        }
    } else {
        for (int i = tid; i < 4 * NS * kMaxTypes; i += kWThreads) {
            const int k = i / (NS * kMaxTypes), rem = i % (NS * kMaxTypes);
            (&tab[0][0][0])[i] = a.ljtab[(k * kMaxSites + rem / kMaxTypes) * kMaxTypes + rem % kMaxTypes];
        }
    }
    // erfc(alpha r) pieces for the FP32 path, as in the grid tile kernel
    float erfc_scale = 0.0f;
    if constexpr (kF32 && CHARGED) {
        const double h = a.alpha * sqrt(a.rc2_real) / (kWThreads - 1);
        const double z0 = tid * h, z1 = z0 + h;
        const double f0 = erfc(z0), f1 = erfc(z1);
        const double d0 = -1.1283791670955126 * exp(-z0 * z0) * h, d1 = -1.1283791670955126 * exp(-z1 * z1) * h;
        erfc_tab[tid] = make_float4((float)f0, (float)d0, (float)(3.0 * (f1 - f0) - (2.0 * d0 + d1)),
                                    (float)(2.0 * (f0 - f1) + (d0 + d1)));
        erfc_scale = (float)(a.alpha / h);
    }

    // cube centre, Cartesian and fractional
    const int i0 = cell % c.nc[0], i1 = (cell / c.nc[0]) % c.nc[1], i2 = cell / (c.nc[0] * c.nc[1]);
    const double cc[3] = {c.lo[0] + (i0 + 0.5) * c.edge, c.lo[1] + (i1 + 0.5) * c.edge, c.lo[2] + (i2 + 0.5) * c.edge};
    double cf[3];
    to_frac(a, cc, cf);

    const R rc2 = (R)a.rc2_lj, rcr2 = (R)a.rc2_real, rc = (R)a.rc_lj, alpha = (R)a.alpha;
    for (uint32_t pass = lo; pass < hi; pass += kWThreads * TPT) {
        uint64_t tr[TPT];
        bool ok[TPT];
        R ox[TPT][NS], oy[TPT][NS], oz[TPT][NS];
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
                ox[k][s] = ok[k] ? (R)(cart[s][0] - cc[0]) : (R)1e6;
                oy[k][s] = ok[k] ? (R)(cart[s][1] - cc[1]) : (R)1e6;
                oz[k][s] = ok[k] ? (R)(cart[s][2] - cc[2]) : (R)1e6;
            }
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < TPT; ++k) any |= ok[k];

        auto compute = [&](int count) {
            __syncthreads();
            if constexpr (kF32) {
                // one branch covers both terms when the LJ and real-space cutoffs agree (or the
                // molecule is uncharged); otherwise each term has its own
                auto pairs = [&](auto same_cutoffs) {
                    constexpr bool kOne = decltype(same_cutoffs)::value;
                    for (int e = 0; e < count; ++e) {
                        const float4 P = ent.p[e];
                        const int ty = CHARGED ? ent.t[e] : __float_as_int(P.w);
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const float4 cg = tabf[s][ty];
                            const float qq = (float)a.charge[s] * P.w;
#pragma unroll
                            for (int k = 0; k < TPT; ++k) {
                                const float dx = P.x - ox[k][s], dy = P.y - oy[k][s], dz = P.z - oz[k][s];
                                const float r2 = dx * dx + dy * dy + dz * dz;
                                // LJ from 1/r^2 (and r, linear-force-shifted only)
                                auto lj_term = [&](float ir2, float r) {
                                    const float ir6 = ir2 * ir2 * ir2;
                                    float w = fmaf(cg.y, ir6, cg.z);
                                    if constexpr (LFS) w = fmaf(r, cg.w, w);
                                    lj[k] += (double)fmaf(cg.x, ir6 * ir6, -w);
                                };
                                auto real_term = [&](float ir) {
                                    const float u = (r2 * ir) * erfc_scale;
                                    const float f = __builtin_amdgcn_fractf(u);
                                    const float4 cf4 = erfc_tab[(int)u];
                                    const float ez4 = fmaf(fmaf(fmaf(cf4.w, f, cf4.z), f, cf4.y), f, cf4.x);
                                    real[k] += (double)(qq * (ez4 * ir));
                                };
                                if constexpr (kOne) {
                                    if (r2 <= rc2) {
                                        if constexpr (!CHARGED && !LFS) {
                                            lj_term(__builtin_amdgcn_rcpf(r2), 0.0f);  // LJ alone needs no 1/r
                                        } else {
                                            const float ir = __builtin_amdgcn_rsqf(r2);
                                            lj_term(ir * ir, r2 * ir);
                                            if constexpr (CHARGED) real_term(ir);
                                        }
                                    }
                                } else {
                                    const float ir = __builtin_amdgcn_rsqf(r2);
                                    if (r2 <= rc2) lj_term(ir * ir, r2 * ir);
                                    if (r2 <= rcr2) real_term(ir);
                                }
                            }
                        }
                    }
                };
                if (any) {
                    if (!CHARGED || rc2 == rcr2)
                        pairs(std::true_type{});
                    else
                        pairs(std::false_type{});
                }
            } else if (any) {
                for (int e = 0; e < count; ++e) {
                    const R ex = ent.x[e], ey = ent.y[e], ez = ent.z[e];
                    const int ty = ent.t[e];
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
#pragma unroll
                        for (int k = 0; k < TPT; ++k) {
                            const double dx = ex - ox[k][s], dy = ey - oy[k][s], dz = ez - oz[k][s];
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            if (r2 <= rc2) {
                                const double ir2 = 1.0 / r2;
                                const double s2 = tab[0][s][ty] * ir2;
                                const double s6 = s2 * s2 * s2;
                                double v = tab[1][s][ty] * (s6 * s6 - s6);
                                if (a.trunc == ALEATOR_LJ_SHIFTED) {
                                    v -= tab[2][s][ty];
                                } else if (a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) {
                                    v = v - tab[2][s][ty] - (sqrt(r2) - rc) * tab[3][s][ty];
                                }
                                lj[k] += v;
                            }
                            if (CHARGED && r2 <= rcr2) {
                                const double r = sqrt(r2);
                                real[k] += (a.charge[s] * ent.q[e]) * erfc_over_r<double>(alpha, r);
                            }
                        }
                    }
                }
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
                    ent.put(w, g0 * a.m[0] + g1 * a.m[3] + g2 * a.m[6], g0 * a.m[1] + g1 * a.m[4] + g2 * a.m[7],
                            g0 * a.m[2] + g1 * a.m[5] + g2 * a.m[8], qa, ta, CHARGED);
                    ++w;
                }
            }
            pending += total;
        }
        compute(pending);
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            if (ok[k]) a.du[tr[k]] = CHARGED ? lj[k] + a.kc * real[k] : lj[k];
    }
}

// Per workgroup sum of exp(-dU/T) over 256 consecutive trials (uncharged: dU is complete).
__global__ __launch_bounds__(kWThreads) void widom_boltz_kernel(const WidomArgs a) {
    __shared__ double red[kWThreads / 64];
    const uint64_t t = (uint64_t)blockIdx.x * kWThreads + threadIdx.x;
    const double b = t < a.n_trials ? exp(-a.du[t] / a.temperature) : 0.0;
    const double s = block_sum(b, red);
    if (threadIdx.x == 0) a.block_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void sum_kernel(const double* parts, uint64_t n, double* out) {
    __shared__ double red[16];
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += 1024) s += parts[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) *out = s;
}

// Trials per thread of the cube kernel: 1, or (for large batches) 4 for one site and 2 otherwise.
inline int max_trials_per_thread(int ns) { return ns == 1 ? 4 : 2; }

template <typename R, bool CH, bool LFS>
void launch_cells(const WidomArgs& a, const CellArgs& c, unsigned n_cells, bool wide, hipStream_t s) {
    switch (a.nsites * 2 + (wide ? 1 : 0)) {
#define CASE(N, T)                                                                                                \
    case 2 * N + (T > 1 ? 1 : 0):                                                                                 \
        hipLaunchKernelGGL((widom_cell_kernel<R, N, CH, LFS, T>), dim3(n_cells), dim3(kWThreads), 0, s, a, c); \
        return;
        CASE(1, 1) CASE(1, 4) CASE(2, 1) CASE(2, 2) CASE(3, 1) CASE(3, 2) CASE(4, 1) CASE(4, 2) CASE(5, 1) CASE(5, 2)
#undef CASE
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

template <typename R, bool CH>
void launch_pair(const WidomArgs& a, unsigned blocks, hipStream_t s) {
    switch (a.nsites) {
#define CASE(N) \
    case N: hipLaunchKernelGGL((widom_pair_kernel<R, N, CH>), dim3(blocks), dim3(kWThreads), 0, s, a); return;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

void launch_recip(const WidomArgs& a, const Framework& fw, unsigned blocks, int rt, hipStream_t s) {
    const size_t lds = (size_t)rt * a.nsites * 3 * (2 * a.kmax + 1) * sizeof(cplx);
    switch (a.nsites) {
#define CASE(N)                                                                                                 \
    case N:                                                                                                     \
        hipLaunchKernelGGL((widom_recip_runs_kernel<N>), dim3(blocks), dim3(256), lds, s, a, fw.d_kseg, fw.n_kseg, \
                           rt);                                                                                 \
        return;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
    }
    fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
}

}  // namespace

void launch_widom(Framework& fw, const WidomRequest& req) {
    Context& ctx = *fw.ctx;
    const int ns = (int)req.sites.size();
    if (ns < 1 || ns > 5) fail(ALEATOR_INVALID_ARGUMENT, "widom: 1 to 5 sites supported");
    if (!(req.temperature > 0.0)) fail(ALEATOR_INVALID_ARGUMENT, "widom: temperature must be positive");
    if (req.n_trials == 0) fail(ALEATOR_INVALID_ARGUMENT, "widom: n_trials must be positive");
    WidomArgs a{};
    a.seed = req.seed;
    a.stream = req.stream;
    a.first = req.first_call;
    a.n_trials = req.n_trials;
    a.draws = ns > 1 ? 6 : 3;
    a.nsites = ns;
    bool any_charge = false;
    double q2 = 0.0;
    for (int s = 0; s < ns; ++s) {
        a.local[s][0] = req.sites[s].x;
        a.local[s][1] = req.sites[s].y;
        a.local[s][2] = req.sites[s].z;
        a.charge[s] = req.sites[s].charge;
        any_charge |= req.sites[s].charge != 0.0;
        q2 += req.sites[s].charge * req.sites[s].charge;
        if (req.sites[s].lj_species >= fw.eps.size())
            fail(ALEATOR_OUT_OF_RANGE, "widom: site LJ species out of range");
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            a.m[3 * i + j] = fw.lat.m[i][j];
            a.inv[3 * i + j] = fw.lat.inv[i][j];
        }
    a.fx = fw.d_fx;
    a.fy = fw.d_fy;
    a.fz = fw.d_fz;
    a.q = fw.d_q;
    a.type = fw.d_type;
    a.natoms = (int)fw.n;
    a.rc2_lj = fw.rc * fw.rc;
    a.rc_lj = fw.rc;
    a.rc2_real = fw.rc_real * fw.rc_real;
    a.alpha = fw.alpha;
    a.kc = kCoulomb;
    a.trunc = fw.trunc;
    const bool charged = fw.charged && any_charge;
    a.charged = charged;
    a.temperature = req.temperature;
    a.self_delta = -kCoulomb * (fw.alpha / std::sqrt(kPi)) * q2;
    a.kn = reinterpret_cast<const int4*>(fw.d_kn);
    a.kw = fw.d_kw;
    a.sre = fw.d_sre;
    a.sim = fw.d_sim;
    a.nk_half = fw.nk_half;
    a.kmax = fw.kmax;
    a.recip_scale = 2.0 * kCoulomb * (2.0 * kPi / std::abs(fw.lat.volume));

    std::vector<double> tab(4 * kMaxSites * kMaxTypes, 0.0);
    for (int s = 0; s < ns; ++s)
        for (size_t t = 0; t < fw.type_species.size(); ++t) {
            double s2, e4, v, dv;
            lj_pair_constants(fw, req.sites[s].lj_species, fw.type_species[t], s2, e4, v, dv);
            tab[(0 * kMaxSites + s) * kMaxTypes + t] = s2;
            tab[(1 * kMaxSites + s) * kMaxTypes + t] = e4;
            tab[(2 * kMaxSites + s) * kMaxTypes + t] = v;
            tab[(3 * kMaxSites + s) * kMaxTypes + t] = dv;
        }
    // Cube grid for the cell-sorted direct sum: about n/1024 trials per occupied cube (so that
    // the cubes keep every CU busy), at least 64 and at most one pass of a workgroup (256
    // threads x trials per thread); list radius R = max cutoff + cube half-diagonal + molecule
    // radius.  It needs R < min L_perp (the 8-image list then holds every image within
    // R); the all-atom kernel runs instead when that fails, or when the sphere of radius R holds
    // more than 70 % of the cell's volume (the list would save too little to pay for the sort).
    CellArgs cg{};
    unsigned n_cells = 0;
    bool wide_cells = false;
    {
        double rmol = 0.0;
        for (int s = 0; s < ns; ++s)
            rmol = std::max(rmol, std::sqrt(a.local[s][0] * a.local[s][0] + a.local[s][1] * a.local[s][1] +
                                            a.local[s][2] * a.local[s][2]));
        const double rmax = std::max(fw.rc, charged ? fw.rc_real : 0.0);
        const double lmin = std::min({fw.lat.lperp[0], fw.lat.lperp[1], fw.lat.lperp[2]});
        const double vol = std::abs(fw.lat.volume);
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int corner = 0; corner < 8; ++corner)
            for (int j = 0; j < 3; ++j) {
                double x = 0.0;
                for (int i = 0; i < 3; ++i)
                    if (corner & (1 << i)) x += fw.lat.m[i][j];
                lo[j] = std::min(lo[j], x);
                hi[j] = std::max(hi[j], x);
            }
        const double per_cell =
            std::clamp((double)req.n_trials / 1024.0, 64.0, (double)kWThreads * max_trials_per_thread(ns));
        wide_cells = per_cell > kWThreads;
        const double edge = std::cbrt(vol / std::max(1.0, (double)req.n_trials / per_cell));
        const double rl = rmax + 0.5 * std::sqrt(3.0) * edge + rmol;
        long long count = 1;
        for (int j = 0; j < 3; ++j) {
            cg.nc[j] = std::max(1, (int)std::ceil((hi[j] - lo[j]) / edge));
            cg.lo[j] = lo[j];
            count *= cg.nc[j];
        }
        cg.edge = edge;
        if (req.n_trials <= 0x7fffffffull && count <= (1 << 22) && rl < lmin &&
            4.0 / 3.0 * kPi * rl * rl * rl <= 0.7 * vol) {
            cg.rlist2 = rl * rl;
            cg.single_image = 2.0 * rl < lmin ? 1 : 0;
            n_cells = (unsigned)count;
        }
    }
    const unsigned pair_blocks = (unsigned)((req.n_trials + kWThreads - 1) / kWThreads);
    const int rec_rt = charged ? runs_molecules_per_block(ns, fw.kmax) : 1;
    if (rec_rt == 0) fail(ALEATOR_NOT_IMPLEMENTED, "widom: phase tables exceed the LDS budget");
    const unsigned rec_blocks = (unsigned)((req.n_trials + rec_rt - 1) / rec_rt);
    const size_t tab_bytes = (tab.size() * sizeof(double) + 255) & ~size_t(255);
    const size_t parts = charged ? rec_blocks : pair_blocks;
    size_t du_bytes = req.du_device ? 0 : ((req.n_trials * sizeof(double) + 255) & ~size_t(255));
    // LDS-histogram sort: about 4096 trials per block, at most 256 blocks
    const bool lds_sort = n_cells > 0 && n_cells <= kSortMaxCells;
    const uint32_t sort_blocks = lds_sort ? (uint32_t)std::min<uint64_t>(256, (req.n_trials + 4095) / 4096) : 0;
    const size_t sort_bytes =
        n_cells ? ((2 * req.n_trials + 3 * (size_t)n_cells + 1 + (size_t)sort_blocks * n_cells) * sizeof(uint32_t) +
                   255) & ~size_t(255)
                : 0;
    char* base = static_cast<char*>(ctx.ensure_scratch(tab_bytes + parts * sizeof(double) + du_bytes + sort_bytes));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                                     ctx.stream));
    a.ljtab = reinterpret_cast<const double*>(base);
    a.block_sums = reinterpret_cast<double*>(base + tab_bytes);
    a.du = req.du_device ? req.du_device
                         : reinterpret_cast<double*>(base + tab_bytes + parts * sizeof(double));
    const bool fp64 = req.precision == ALEATOR_PRECISION_FP64;
    if (!fp64 && req.precision != ALEATOR_PRECISION_MIXED) fail(ALEATOR_INVALID_ARGUMENT, "widom: unknown precision");
    ctx.time_begin();
    if (n_cells) {
        // group the trials by cell: counts, exclusive scan, scatter (order within a cell is
        // arbitrary and does not matter: every trial is summed on its own, into du[t])
        uint32_t* key = reinterpret_cast<uint32_t*>(base + tab_bytes + parts * sizeof(double) + du_bytes);
        uint32_t* order = key + req.n_trials;
        uint32_t* count = order + req.n_trials;
        uint32_t* start = count + n_cells;
        uint32_t* cursor = start + n_cells + 1;
        cg.start = start;
        cg.order = order;
        const unsigned tb = (unsigned)((req.n_trials + 255) / 256);
        if (lds_sort) {
            uint32_t* hist = cursor + n_cells;
            const uint64_t per = (req.n_trials + sort_blocks - 1) / sort_blocks;
            const size_t lds = (size_t)n_cells * sizeof(uint32_t);
            hipLaunchKernelGGL(cube_hist_kernel, dim3(sort_blocks), dim3(kSortThreads), lds, ctx.stream, a, cg, per,
                               n_cells, hist);
            hipLaunchKernelGGL(cube_colscan_kernel, dim3((n_cells + 255) / 256), dim3(256), 0, ctx.stream, hist,
                               sort_blocks, n_cells, count);
            hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, ctx.stream, count, n_cells, start, cursor);
            hipLaunchKernelGGL(cube_place_kernel, dim3(sort_blocks), dim3(kSortThreads), lds, ctx.stream, a, cg, per,
                               n_cells, hist, start, order);
        } else {
            hipLaunchKernelGGL(fill_u32_kernel, dim3((n_cells + 255) / 256), dim3(256), 0, ctx.stream, count,
                               (uint64_t)n_cells);
            hipLaunchKernelGGL(widom_key_kernel, dim3(tb), dim3(256), 0, ctx.stream, a, cg, key, count);
            hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, ctx.stream, count, n_cells, start, cursor);
            hipLaunchKernelGGL(scatter_kernel, dim3(tb), dim3(256), 0, ctx.stream, key, req.n_trials, cursor, order);
        }
        const bool lfs = a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED;
        if (charged) {
            if (fp64) launch_cells<double, true, false>(a, cg, n_cells, wide_cells, ctx.stream);
            else if (lfs) launch_cells<float, true, true>(a, cg, n_cells, wide_cells, ctx.stream);
            else launch_cells<float, true, false>(a, cg, n_cells, wide_cells, ctx.stream);
            launch_recip(a, fw, rec_blocks, rec_rt, ctx.stream);
        } else {
            if (fp64) launch_cells<double, false, false>(a, cg, n_cells, wide_cells, ctx.stream);
            else if (lfs) launch_cells<float, false, true>(a, cg, n_cells, wide_cells, ctx.stream);
            else launch_cells<float, false, false>(a, cg, n_cells, wide_cells, ctx.stream);
            hipLaunchKernelGGL(widom_boltz_kernel, dim3(pair_blocks), dim3(kWThreads), 0, ctx.stream, a);
        }
    } else if (charged) {
        if (fp64) launch_pair<double, true>(a, pair_blocks, ctx.stream);
        else launch_pair<float, true>(a, pair_blocks, ctx.stream);
        launch_recip(a, fw, rec_blocks, rec_rt, ctx.stream);
    } else {
        if (fp64) launch_pair<double, false>(a, pair_blocks, ctx.stream);
        else launch_pair<float, false>(a, pair_blocks, ctx.stream);
    }
    ctx.time_end();
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, ctx.stream, a.block_sums, (uint64_t)parts,
                       req.sum_device);
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
