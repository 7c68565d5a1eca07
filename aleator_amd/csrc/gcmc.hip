// gcmc.hip -- many independent grand-canonical Monte Carlo chains in one persistent launch.
//
// Reference: MonteCarloEngine::run and its single-species moves
// (src/engines/monte_carlo/monte_carlo_engine.cc:405-439 run, :441-484 translation,
// :486-554 rotation, :556-619 insertion, :621-671 deletion, :380-403 Ewald resync every 100
// accepted moves, :814-848 molecule removal), acceptance ratios (gcmc_acceptance.hpp:28-101)
// and the S(k) update it drives (ewald_incremental_state.cc:95-172, 214-247).
//
// MI355X design.  One workgroup of 256 threads owns one chain for the whole launch.  Its
// structure factor S(k) (half set, see recip.hpp) lives in LDS; molecule positions live in
// global memory in the reference's molecule order (erase on delete keeps that order, which is
// what the molecule-index draws select by).  Every random draw of a step is a pure function
// of the chain's Philox call index, so every thread evaluates the same draws and the same
// trial positions with no broadcast.  The O(N) energy scans (framework atoms, other guests)
// and the k-sum are split over the workgroup's threads and reduced through LDS; one thread
// applies the acceptance rule.  Draw order per move matches the reference exactly, so a chain
// follows the reference trajectory for the same (seed, stream) until two energies that differ
// only by rounding straddle an acceptance threshold.
#include <cmath>
#include <vector>

#include "device.hpp"
#include "gcmc.hpp"
#include "internal.hpp"
#include "pair.hpp"
#include "recip.hpp"

namespace aleator_gpu {

namespace {

constexpr int kGThreads = 512;  // 8 waves per chain: two per SIMD, 256 VGPRs each
constexpr int kMaxKHalf = 2048;   // S(k) half set kept in LDS
constexpr int kResyncChunk = 16;  // guest sites per resync pass

constexpr int kDraws = 16;        // uniforms drawn ahead per step (a step takes at most 10)

// The step's uniforms (Philox calls base .. base + kDraws - 1, drawn in parallel into LDS),
// consumed in order by thread 0.
struct Draws {
    const double* u;
    int used;
    __device__ double operator()() { return u[used++]; }
};

// Smallest b with u < cdf[b] over a strictly increasing cdf with cdf[n - 1] = 1 > u: the same
// bin as a binary search, found by the 64 lanes of one wave probing 64 evenly spaced entries per
// round (three dependent loads for 2^18 bins instead of eighteen).  Every lane returns it.
__device__ int wave_cdf_search(const double* cdf, int n, double u) {
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = n - 1;  // invariant: cdf[b] <= u for b < lo, u < cdf[hi]
    while (lo < hi) {
        const long long span = hi - lo + 1;
        const int p = lo + (int)(((long long)(lane + 1) * span) / 64) - 1;
        const bool above = p >= lo && u < cdf[p];
        const unsigned long long mask = __ballot(above);
        const int f = __ffsll(mask) - 1;  // lane 63 probes hi, so the mask is never empty
        const int pf = lo + (int)(((long long)(f + 1) * span) / 64) - 1;
        const int pprev = f == 0 ? lo - 1 : lo + (int)(((long long)f * span) / 64) - 1;
        lo = pprev + 1;
        hi = pf;
    }
    return lo;
}

__device__ inline void frac_of(const GcmcArgs& a, const double c[3], double f[3]) {
    for (int j = 0; j < 3; ++j) f[j] = c[0] * a.inv[j] + c[1] * a.inv[3 + j] + c[2] * a.inv[6 + j];
}

// minimum-image distance^2 between fractional points (exact below L_perp/2)
__device__ inline double dist2(const GcmcArgs& a, double d0, double d1, double d2) {
    d0 -= rint(d0);
    d1 -= rint(d1);
    d2 -= rint(d2);
    const double x = d0 * a.m[0] + d1 * a.m[3] + d2 * a.m[6];
    const double y = d0 * a.m[1] + d1 * a.m[4] + d2 * a.m[7];
    const double z = d0 * a.m[2] + d1 * a.m[5] + d2 * a.m[8];
    return x * x + y * y + z * z;
}

// Guest cell of a fractional position: gc[j] cells along lattice axis j, position wrapped
// into the cell first.
__device__ inline void guest_cell(const GcmcArgs& a, const double f[3], int c[3]) {
    for (int j = 0; j < 3; ++j) {
        const double t = (f[j] - floor(f[j])) * (double)a.gc[j];
        const int k = t >= 0.0 ? (int)t : 0;
        c[j] = k < a.gc[j] - 1 ? k : a.gc[j] - 1;
    }
}

__device__ inline double lj_term(const GcmcArgs& a, const double* tab4, int stride, int idx, double r2) {
    // tab4: [4][stride] = sigma^2, 4 eps, V(rc), V'(rc) (lennard_jones.cc:70-86)
    const double s2 = tab4[idx] / r2;
    const double s6 = s2 * s2 * s2;
    double v = tab4[stride + idx] * (s6 * s6 - s6);
    if (a.trunc == ALEATOR_LJ_SHIFTED)
        v -= tab4[2 * stride + idx];
    else if (a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED)
        v = v - tab4[2 * stride + idx] - (sqrt(r2) - a.rc_lj) * tab4[3 * stride + idx];
    return v;
}

// erfc(alpha r)/r from r^2 as 1/r - erf(alpha r)/r: 1/r from the hardware reciprocal square root
// and two Newton steps, erf(alpha r)/r by Horner's rule over the polynomial coefficients in LDS.
// (The library erfc, inlined into the pair loops, needed more registers than the kernel has.)
template <int DEG>
__device__ inline double erf_horner(const double* poly, double x) {
    double c[DEG + 1];
#pragma unroll
    for (int k = 0; k <= DEG; ++k) c[k] = poly[k];  // independent LDS reads, issued together
    double p = c[DEG];
#pragma unroll
    for (int k = DEG - 1; k >= 0; --k) p = fma(p, x, c[k]);
    return p;
}
__device__ inline double erfc_over_r(const GcmcArgs& a, const double* poly, double r2) {
    const double x = fma(r2, a.erf_xs, -1.0);
    const double p = a.erf_deg == 18 ? erf_horner<18>(poly, x)
                   : a.erf_deg == 24 ? erf_horner<24>(poly, x)
                                     : erf_horner<kErfPolyMax>(poly, x);
    return rsq_refined<2>(r2) - p;
}

// Reduce four doubles over the workgroup of NW waves, in a fixed order; every thread receives
// the totals.
template <int NW>
__device__ inline void block_sum4(double v[4], double* lds /* [NW][4] */) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = 0; i < 4; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) lds[wave * 4 + i] = s;
    }
    __syncthreads();
    for (int i = 0; i < 4; ++i) {
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += lds[4 * w + i];
        v[i] = t;
    }
    __syncthreads();
}

// Framework bin of a fractional position (wrapped into the cell).
__device__ inline int framework_bin(const GcmcArgs& a, const double f[3]) {
    int b[3];
    for (int j = 0; j < 3; ++j) {
        const double t = (f[j] - floor(f[j])) * (double)a.fb[j];
        const int k = t >= 0.0 ? (int)t : 0;
        b[j] = k < a.fb[j] - 1 ? k : a.fb[j] - 1;
    }
    return (b[0] * a.fb[1] + b[1]) * a.fb[2] + b[2];
}

__device__ inline void bin_centre(const GcmcArgs& a, int bin, double c[3]) {
    const int b2 = bin % a.fb[2], b1 = (bin / a.fb[2]) % a.fb[1], b0 = bin / (a.fb[1] * a.fb[2]);
    c[0] = (b0 + 0.5) / a.fb[0];
    c[1] = (b1 + 0.5) / a.fb[1];
    c[2] = (b2 + 0.5) / a.fb[2];
}

// Whether framework atom j lies within sqrt(reach2) of fractional point c (minimum image, no
// contraction, so the counting and filling passes agree).
__device__ inline bool atom_within(const GcmcArgs& a, int j, const double c[3], double reach2) {
#pragma clang fp contract(off)
    double d[3] = {a.fx[j] - c[0], a.fy[j] - c[1], a.fz[j] - c[2]};
    for (int i = 0; i < 3; ++i) d[i] = d[i] - rint(d[i]);
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double x = d[0] * a.m[k] + d[1] * a.m[3 + k] + d[2] * a.m[6 + k];
        r2 = r2 + x * x;
    }
    return r2 <= reach2;
}

// One block per bin: how many framework atoms lie within `reach` of its centre.
__global__ __launch_bounds__(256) void bin_count_kernel(const GcmcArgs a, double reach2, int32_t* counts) {
    __shared__ int warp_tot[5];
    double c[3];
    bin_centre(a, blockIdx.x, c);
    int n = 0;
    for (int j = threadIdx.x; j < a.natoms; j += blockDim.x)
        n += atom_within(a, j, c, reach2) ? 1 : 0;
    int tot = 0;
    (void)block_exclusive_scan(n, warp_tot, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// One block per bin: its atoms, in ascending index order, as (fx, fy, fz, q) and type at start[bin].
__global__ __launch_bounds__(256) void bin_fill_kernel(const GcmcArgs a, double reach2, const int32_t* start,
                                                       double4* pos, int32_t* type) {
    __shared__ int warp_tot[5];
    double c[3];
    bin_centre(a, blockIdx.x, c);
    int base = start[blockIdx.x];
    const int limit = start[blockIdx.x + 1];
    for (int j0 = 0; j0 < a.natoms; j0 += blockDim.x) {
        const int j = j0 + threadIdx.x;
        const int in = (j < a.natoms && atom_within(a, j, c, reach2)) ? 1 : 0;
        int tot = 0;
        const int off = block_exclusive_scan(in, warp_tot, &tot);
        if (in && base + off < limit) {
            pos[base + off] = double4{a.fx[j], a.fy[j], a.fz[j], a.q[j]};
            type[base + off] = a.type[j];
        }
        base += tot;
        __syncthreads();
    }
}

// Bin of the insertion density holding a molecule's anchor (the centroid of its sites):
// fractional coordinates f = c . inv, folded into one primitive image, then binned.
__device__ int anchor_bin(const GcmcArgs& a, const double c[kMaxSites][3]) {
#pragma clang fp contract(off)
    const int ns = a.ns;
    double cen[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < ns; ++s)
        for (int j = 0; j < 3; ++j) cen[j] = cen[j] + c[s][j];
    const double inv_n = 1.0 / (double)ns;
    int b[3];
    for (int j = 0; j < 3; ++j) {
        const double f = (cen[0] * inv_n) * a.inv[j] + (cen[1] * inv_n) * a.inv[3 + j] + (cen[2] * inv_n) * a.inv[6 + j];
        double t = f * (double)a.prop_reps[j];
        t = t - floor(t);
        const double kf = t * (double)a.prop_nb[j];
        const int k = kf >= 0.0 ? (int)kf : 0;  // also maps a NaN coordinate into the table
        b[j] = k < a.prop_nb[j] - 1 ? k : a.prop_nb[j] - 1;
    }
    return (b[0] * a.prop_nb[1] + b[1]) * a.prop_nb[2] + b[2];
}

// Trial positions of one move, in the reference's arithmetic order (no contraction) and draw
// order: translation (:454-461), rotation about the centroid (:512-532, Rodrigues :40-44),
// insertion (:568-590, Shoemake quaternion :56-72).  Deletion draws nothing here.  With an
// insertion density, the insertion anchor comes from it (draws: bin by inverse CDF, found ahead
// as prop_bin, image, position in the bin) and `bias` receives V p(anchor) for insertion and
// deletion.
__device__ void trial_positions(const GcmcArgs& a, Draws& uni, int move, int prop_bin,
                                const double old_c[kMaxSites][3], double new_c[kMaxSites][3], double& bias) {
#pragma clang fp contract(off)
    const int ns = a.ns;
    bias = 1.0;
    if (move == 3 && a.prop_cdf) bias = a.prop_fac[anchor_bin(a, old_c)];
    if (move == 0) {
        double d[3];
        for (int j = 0; j < 3; ++j) d[j] = (2.0 * uni() - 1.0) * a.max_translation;
        for (int s = 0; s < ns; ++s)
            for (int j = 0; j < 3; ++j) new_c[s][j] = old_c[s][j] + d[j];
    } else if (move == 1) {
        // attemptRotation (:512-532), Rodrigues about the centroid (:40-44)
        double cen[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < ns; ++s)
            for (int j = 0; j < 3; ++j) cen[j] = cen[j] + old_c[s][j];
        const double inv_n = 1.0 / (double)ns;
        for (int j = 0; j < 3; ++j) cen[j] = cen[j] * inv_n;
        const double cos_t = 2.0 * uni() - 1.0;
        const double om = 1.0 - cos_t * cos_t;
        const double sin_t = sqrt((0.0 < om) ? om : 0.0);
        const double phi = 2.0 * kPi * uni();
        const double ax[3] = {sin_t * cos(phi), sin_t * sin(phi), cos_t};
        const double angle = (2.0 * uni() - 1.0) * a.max_rotation;
        const double c = cos(angle), sn = sin(angle);
        for (int s = 0; s < ns; ++s) {
            const double v[3] = {old_c[s][0] - cen[0], old_c[s][1] - cen[1], old_c[s][2] - cen[2]};
            const double cr[3] = {ax[1] * v[2] - ax[2] * v[1], ax[2] * v[0] - ax[0] * v[2],
                                  ax[0] * v[1] - ax[1] * v[0]};
            const double dv = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
            for (int j = 0; j < 3; ++j)
                new_c[s][j] = cen[j] + ((v[j] * c + cr[j] * sn) + ax[j] * (dv * (1.0 - c)));
        }
    } else if (move == 2) {
        // attemptInsertion (:568-590)
        double ref[3];
        if (a.prop_cdf) {
            (void)uni();  // the bin draw: prop_bin is the smallest bin with this uniform < cdf[bin]
            const int lo = prop_bin;
            bias = a.prop_fac[lo];
            const int bn[3] = {lo / (a.prop_nb[1] * a.prop_nb[2]), (lo / a.prop_nb[2]) % a.prop_nb[1], lo % a.prop_nb[2]};
            const int n_img = a.prop_reps[0] * a.prop_reps[1] * a.prop_reps[2];
            int img = (int)(uni() * (double)n_img);
            img = img < n_img - 1 ? img : n_img - 1;
            const int im[3] = {img / (a.prop_reps[1] * a.prop_reps[2]), (img / a.prop_reps[2]) % a.prop_reps[1],
                               img % a.prop_reps[2]};
            double f[3];
            for (int j = 0; j < 3; ++j) {
                const double u = uni();
                f[j] = ((double)im[j] + ((double)bn[j] + u) / (double)a.prop_nb[j]) / (double)a.prop_reps[j];
            }
            for (int j = 0; j < 3; ++j) ref[j] = a.m[j] * f[0] + a.m[3 + j] * f[1] + a.m[6 + j] * f[2];
        } else {
            const double u0 = uni(), u1 = uni(), u2 = uni();
            for (int j = 0; j < 3; ++j) ref[j] = a.m[j] * u0 + a.m[3 + j] * u1 + a.m[6 + j] * u2;
        }
        if (ns == 1) {
            for (int j = 0; j < 3; ++j) new_c[0][j] = ref[j] + a.local[0][j];
        } else {
            const double v0 = uni(), v1 = uni(), v2 = uni();
            const double r1 = sqrt(1.0 - v0), r2 = sqrt(v0);
            const double th1 = 2.0 * kPi * v1, th2 = 2.0 * kPi * v2;
            const double qw = r2 * cos(th2), qx = r1 * sin(th1), qy = r1 * cos(th1), qz = r2 * sin(th2);
            for (int s = 0; s < ns; ++s) {
                const double vx = a.local[s][0], vy = a.local[s][1], vz = a.local[s][2];
                const double tx = (qy * vz - qz * vy) * 2.0;
                const double ty = (qz * vx - qx * vz) * 2.0;
                const double tz = (qx * vy - qy * vx) * 2.0;
                new_c[s][0] = ref[0] + ((vx + tx * qw) + (qy * tz - qz * ty));
                new_c[s][1] = ref[1] + ((vy + ty * qw) + (qz * tx - qx * tz));
                new_c[s][2] = ref[2] + ((vz + tz * qw) + (qx * ty - qy * tx));
            }
        }
    }
}

// One workgroup of NT threads per chain.  Thread 0 takes every random draw of the chain, builds
// the move's trial sites and publishes them through LDS (fractional coordinates and charges of
// the removed rows, then the added rows); all threads then split the energy scans and the k-sum;
// thread 0 applies the acceptance rule and publishes the decision.
// CHARGED: the framework carries Ewald charges (CHARGED); the uncharged instances drop the S(k)
// arrays and code.
// The arguments are read through a pointer to a device copy: as a by-value kernel argument the
// compiler kept its many loop-invariant fields in scalar registers and spilled them.
template <int NT, bool CHARGED, bool CELLS>
__global__ __launch_bounds__(NT) void gcmc_chain_kernel(const GcmcArgs* __restrict__ args) {
    const GcmcArgs& a = *args;
    constexpr int NW = NT / 64;
    __shared__ double s_re[kMaxKHalf], s_im[kMaxKHalf];
    // k vectors of the half set, (nx + kmax) | (ny + kmax) << 8 | (nz + kmax) << 16, and weights
    __shared__ int s_kn[CHARGED ? kMaxKHalf : 1];
    __shared__ double s_kw[CHARGED ? kMaxKHalf : 1];
    __shared__ cplx phase[3 * (2 * kMaxK + 1) * 2 * kMaxSites];
    __shared__ cplx rphase[3 * (2 * kMaxK + 1) * kResyncChunk];
    __shared__ double site4[2 * kMaxSites * 4];  // removed rows, then added rows: f0 f1 f2 q
    __shared__ double oldc[kMaxSites][3];        // cartesian positions of the removed sites (thread 0)
    __shared__ double newc[kMaxSites][3];        // cartesian positions of the added sites
    __shared__ double red[4 * NW];
    __shared__ int ctl[3];                        // move (-1: nothing to attempt), molecule, accept
    __shared__ double s_u[kDraws];                // this step's uniforms
    __shared__ uint64_t s_call;                   // Philox call index of the next step's first draw
    __shared__ int s_rbeg[2 * kMaxSites], s_rlen[2 * kMaxSites];  // framework list of each site row
    __shared__ double s_lj[4 * kMaxSites * kMaxTypes];             // site vs framework type LJ constants
    __shared__ double s_erf[CHARGED ? kErfPolyMax + 1 : 1];        // real-space kernel polynomial
    __shared__ int s_cand[CELLS ? kMaxGuestCand : 1];  // guest cell filter: candidate sites, ascending
    __shared__ int s_rowcell[2 * kMaxSites][3];  // guest cell of each site row
    __shared__ int s_wtot[NW];
    // statistics and pending-move state, read and written by thread 0 only; kept in LDS so the
    // other threads hold no registers for them
    __shared__ struct {
        double self_e, excl_e, sum_n;
        double bias;                  // V p(anchor) of the insertion density (1: uniform insertion)
        double pend_self, pend_excl;  // self / exclusion deltas of the pending move
        uint64_t att[4], acc[4];
        uint64_t step_call;           // Philox call index of the current step's first draw
        double max_drift;             // resync consistency check (ChainState::max_drift)
        uint64_t resyncs;
    } t0;

    const int chain = blockIdx.x;
    const int tid = threadIdx.x;
    const ChainParams P = a.params[chain];
    ChainState* st = a.state + chain;
    double* pos = a.mol_pos + (size_t)chain * a.max_mol * a.ns * 3;
    const int ns = a.ns;
    const int nkk = 2 * a.kmax + 1;

    if (CHARGED)
        for (int t = tid; t < a.nk_half; t += NT) {
            s_re[t] = a.sre[(size_t)chain * a.nk_half + t];
            s_im[t] = a.sim[(size_t)chain * a.nk_half + t];
            const int4 n = a.kn[t];
            s_kn[t] = (n.x + a.kmax) | (n.y + a.kmax) << 8 | (n.z + a.kmax) << 16;
            s_kw[t] = a.kw[t];
        }
    for (int t = tid; t < 4 * kMaxSites * kMaxTypes; t += NT) s_lj[t] = a.ljtab[t];
    if (CHARGED && tid <= a.erf_deg) s_erf[tid] = a.erf_poly[tid];
    __syncthreads();

    // tracked by every thread
    int n_mol = st->n_mol;
    int since_resync = st->moves_since_resync;
    int error = st->error;
    if (tid == 0) {
        s_call = st->call;
        t0.self_e = st->self_energy;
        t0.excl_e = st->exclusion_energy;
        t0.sum_n = st->sum_n;
        t0.bias = 1.0;
        t0.pend_self = 0.0;
        t0.pend_excl = 0.0;
        for (int i = 0; i < 4; ++i) {
            t0.att[i] = st->attempted[i];
            t0.acc[i] = st->accepted[i];
        }
        t0.max_drift = st->max_drift;
        t0.resyncs = st->resyncs;
    }
    const double volume = fabs(a.volume);

    __syncthreads();
    // a chain whose accepted insertion would exceed max_mol stops before that step takes effect:
    // `overflow` marks it, the step's counts are taken back and `step` is the number of steps run
    bool overflow = false;
    uint64_t step = 0;
    for (; step < a.n_steps && !error; ++step) {
        // ---------------- this step's uniforms, drawn ahead in parallel
        const uint64_t call = s_call;
        if (tid < kDraws) s_u[tid] = philox_uniform(P.seed, P.stream, call + tid);
        __syncthreads();
        // the insertion bin of a biased insertion (the draw after the move pick), by wave 0
        int prop_bin = 0;
        if (a.prop_cdf && tid < 64 && s_u[0] >= 0.5 && s_u[0] < 0.75)
            prop_bin = wave_cdf_search(a.prop_cdf, a.prop_nbins, s_u[1]);
        Draws uni{s_u, 0};
        // ---------------- trial sites (thread 0)
        if (tid == 0) {
            t0.step_call = call;
            const double pick = uni();
            const int move = pick < 0.25 ? 0 : (pick < 0.5 ? 1 : (pick < 0.75 ? 2 : 3));
            const bool skip = (move != 2 && n_mol == 0) || (move == 1 && ns < 2);
            int mi = -1;
            if (!skip) {
                if (move != 2) {
                    mi = (int)(uni() * (double)n_mol);
                    for (int s = 0; s < ns; ++s)
                        for (int j = 0; j < 3; ++j) oldc[s][j] = pos[((size_t)mi * ns + s) * 3 + j];
                }
                trial_positions(a, uni, move, prop_bin, oldc, newc, t0.bias);
                int row = 0;
                if (move != 2)
                    for (int s = 0; s < ns; ++s, ++row) {
                        frac_of(a, oldc[s], &site4[4 * row]);
                        site4[4 * row + 3] = a.charge[s];
                    }
                if (move != 3)
                    for (int s = 0; s < ns; ++s, ++row) {
                        frac_of(a, newc[s], &site4[4 * row]);
                        site4[4 * row + 3] = a.charge[s];
                    }
            }
            ctl[0] = skip ? -1 : move;
            ctl[1] = mi;
            if (skip) s_call = call + (uint64_t)uni.used;
        }
        __syncthreads();
        const int move = ctl[0], mi = ctl[1];
        if (move < 0) {
            if (tid == 0) {
                t0.sum_n += (double)n_mol;
                if (a.n_trace) a.n_trace[(size_t)chain * a.n_steps + step] = n_mol;
            }
            __syncthreads();  // ctl, s_u and s_call are rewritten next step
            continue;
        }
        const bool has_old = move != 2, has_new = move != 3;
        const int n_rem = has_old ? ns : 0, n_add = has_new ? ns : 0, nsite = n_rem + n_add;

        // ---------------- pair energies: LJ and real-space Ewald, old and new
        double e4[4] = {0.0, 0.0, 0.0, 0.0};  // lj_old, lj_new, real_old, real_new
        // framework: every (site row, atom of that row's bin list) pair in one loop, so that the
        // rows' loads are in flight together (all atoms per row without lists)
        if (tid < nsite) {
            int beg = 0, len = a.natoms;
            if (a.grid && !CHARGED) {
                len = 0;  // the grid holds the whole framework part
            } else if (a.bin_start) {
                const double f[3] = {site4[4 * tid], site4[4 * tid + 1], site4[4 * tid + 2]};
                const int b = framework_bin(a, f);
                beg = a.bin_start[b];
                len = a.bin_start[b + 1] - beg;
            }
            s_rbeg[tid] = beg;
            s_rlen[tid] = len;
            if (CELLS) guest_cell(a, &site4[4 * tid], s_rowcell[tid]);
        }
        __syncthreads();
        {
            int row = 0, row_start = 0, row_len = s_rlen[0];
            for (int i = tid;; i += NT) {
                while (row < nsite && i >= row_start + row_len) {
                    row_start += row_len;
                    if (++row < nsite) row_len = s_rlen[row];
                }
                if (row >= nsite) break;
                const int k = s_rbeg[row] + (i - row_start);
                double4 at;
                int ty;
                if (a.bin_start) {
                    at = a.bin_pos[k];
                    ty = a.bin_type[k];
                } else {
                    at = double4{a.fx[k], a.fy[k], a.fz[k], a.q[k]};
                    ty = a.type[k];
                }
                const bool is_old = row < n_rem;
                const int s = is_old ? row : row - n_rem;
                const double r2 = dist2(a, at.x - site4[4 * row], at.y - site4[4 * row + 1], at.z - site4[4 * row + 2]);
                if (!a.grid && r2 <= a.rc2_lj)
                    e4[is_old ? 0 : 1] += lj_term(a, s_lj, kMaxSites * kMaxTypes, s * kMaxTypes + ty, r2);
                if (CHARGED && r2 <= a.rc2_real) e4[is_old ? 2 : 3] += a.charge[s] * at.w * erfc_over_r(a, s_erf, r2);
            }
        }
        // other guests: every (site row, candidate guest site) pair in one flattened loop over
        // the workgroup, so that a few dozen molecules still occupy every thread.  Without guest
        // cells every site of the other molecules is a candidate.  With them (cells at least the
        // guest cutoff wide, perpendicular to each lattice plane) a site is a candidate when its
        // cell is a periodic neighbour of some row's cell, which holds for every site within the
        // cutoff of a row; the candidates are compacted in ascending site order, the order of
        // the reference's sorted cell-list candidates (monte_carlo_engine.cc:358-366).
        {
            const int gsites = n_mol * ns;
            const bool cells = CELLS && gsites >= a.gc_min_sites;
            int ncand = gsites;
            if (CELLS && cells) {
                const int wave = tid >> 6, lane = tid & 63;
                int out = 0;
                for (int base = 0; base < gsites; base += NT) {
                    const int gs = base + tid;
                    bool take = false;
                    if (gs < gsites && gs / ns != mi) {
                        const double* gp = pos + (size_t)gs * 3;
                        const double gc[3] = {gp[0], gp[1], gp[2]};
                        double gf[3];
                        frac_of(a, gc, gf);
                        int c[3];
                        guest_cell(a, gf, c);
                        for (int r = 0; r < nsite && !take; ++r) {
                            bool near = true;
                            for (int j = 0; j < 3; ++j) {
                                int d = c[j] - s_rowcell[r][j];
                                d = d < 0 ? d + a.gc[j] : d;
                                near = near && (d <= 1 || d == a.gc[j] - 1);
                            }
                            take = near;
                        }
                    }
                    const unsigned long long mask = __ballot(take);
                    if (lane == 0) s_wtot[wave] = __popcll(mask);
                    __syncthreads();
                    int off = out;
                    for (int w = 0; w < wave; ++w) off += s_wtot[w];
                    if (take) s_cand[off + __popcll(mask & ((1ull << lane) - 1ull))] = gs;
                    for (int w = 0; w < NW; ++w) out += s_wtot[w];
                    __syncthreads();
                }
                ncand = out;
            }
            const int total = nsite * ncand;
            for (int i = tid; i < total; i += NT) {
                const int row = i / ncand, k = i - row * ncand;
                const int gs = CELLS && cells ? s_cand[k] : k;
                const int g = gs / ns, u = gs - g * ns;
                if (g == mi) continue;
                const double* gp = pos + (size_t)gs * 3;
                const double gc[3] = {gp[0], gp[1], gp[2]};
                double gf[3];
                frac_of(a, gc, gf);
                const bool is_old = row < n_rem;
                const int s = is_old ? row : row - n_rem;
                const int idx = s * kMaxSites + u;
                const double r2 =
                    dist2(a, gf[0] - site4[4 * row], gf[1] - site4[4 * row + 1], gf[2] - site4[4 * row + 2]);
                if (r2 <= a.rc2_lj) e4[is_old ? 0 : 1] += lj_term(a, a.ljgg, kMaxSites * kMaxSites, idx, r2);
                if (CHARGED && r2 <= a.rc2_real)
                    e4[is_old ? 2 : 3] += a.charge[s] * a.charge[u] * erfc_over_r(a, s_erf, r2);
            }
        }
        // framework LJ from the energy grid: one site row per thread
        if (a.grid && tid < nsite) {
            const bool is_old = tid < n_rem;
            const int s = is_old ? tid : tid - n_rem;
            const double* c = is_old ? oldc[s] : newc[s];
            e4[is_old ? 0 : 1] += grid_interpolate(a.grid_geo, a.grid + (size_t)a.grid_table[s] * a.grid_cells, c[0],
                                                   c[1], c[2]);
        }
        block_sum4<NW>(e4, red);
        // ---------------- reciprocal space: removed = old sites, added = new sites
        double recip = 0.0;
        if (CHARGED) {
            for (int i = tid; i < 3 * nkk * nsite; i += NT) {
                const int s = i % nsite, mm = (i / nsite) % nkk, axis = i / (nsite * nkk);
                double sv, cv;
                sincospi(2.0 * (double)(mm - a.kmax) * site4[4 * s + axis], &sv, &cv);
                phase[i] = {cv, sv};
            }
            __syncthreads();
            double accr = 0.0;
            for (int t = tid; t < a.nk_half; t += NT) {
                const int kp = s_kn[t];
                double dre = 0.0, dim = 0.0;
                for (int s = 0; s < nsite; ++s) {
                    const cplx e = cmul(cmul(phase[(kp & 0xff) * nsite + s], phase[(nkk + ((kp >> 8) & 0xff)) * nsite + s]),
                                        phase[(2 * nkk + (kp >> 16)) * nsite + s]);
                    const double qs = s < n_rem ? -site4[4 * s + 3] : site4[4 * s + 3];
                    dre += qs * e.re;
                    dim += qs * e.im;
                }
                accr += s_kw[t] * (2.0 * (s_re[t] * dre + s_im[t] * dim) + (dre * dre + dim * dim));
            }
            double r4[4] = {accr, 0.0, 0.0, 0.0};
            block_sum4<NW>(r4, red);
            recip = a.recip_scale * r4[0];
        }
        // ---------------- acceptance (thread 0), gcmc_acceptance.hpp:28-101
        // The intramolecular exclusion terms come first, one (site pair, new or old) per lane of
        // wave 0 in the reference's pair order; thread 0 adds them up in that order.
        double ex_new = 0.0, ex_old = 0.0;
        if (CHARGED && tid < 64) {
            const int npair = ns * (ns - 1) / 2;
            double term = 0.0;
            if (tid < 2 * npair) {
                const bool of_new = tid < npair;
                int p = of_new ? tid : tid - npair, i = 0;
                while (p >= ns - 1 - i) {
                    p -= ns - 1 - i;
                    ++i;
                }
                const int j2 = i + 1 + p;
                if (of_new ? has_new : has_old) {
                    const double* fi = &site4[4 * ((of_new ? n_rem : 0) + i)];
                    const double* fj = &site4[4 * ((of_new ? n_rem : 0) + j2)];
                    const double r = sqrt(dist2(a, fj[0] - fi[0], fj[1] - fi[1], fj[2] - fi[2]));
                    term = (a.charge[i] * a.charge[j2]) *
                           (r < 1e-8 ? 2.0 * a.alpha / sqrt(kPi) : erf(a.alpha * r) / r);
                }
            }
            for (int p = 0; p < npair; ++p) {
                ex_new += __shfl(term, p);
                ex_old += __shfl(term, npair + p);
            }
        }
        if (tid == 0) {
            double self_d = 0.0, excl_d = 0.0;
            if (CHARGED) {
                double q2 = 0.0;
                for (int s = 0; s < ns; ++s) q2 += a.charge[s] * a.charge[s];
                self_d = a.self_pref * ((has_new ? q2 : 0.0) - (has_old ? q2 : 0.0));
                excl_d = -a.kc * (ex_new - ex_old);
            }
            const double recip_total = recip + self_d + excl_d;
            const double T = P.temperature;
            double du, ratio;
            if (move <= 1) {
                du = (e4[1] - e4[0]) + a.kc * (e4[3] - e4[2]) + recip_total;
                ratio = exp(-du / T);
            } else if (move == 2) {
                du = e4[1] + a.kc * e4[3] + recip_total;
                ratio = (P.fugacity_internal * volume / T) / (double)(n_mol + 1) * exp(-du / T);
                if (a.prop_cdf) ratio = ratio / t0.bias;
            } else {
                du = -(e4[0] + a.kc * e4[2]) + recip_total;
                ratio = (double)n_mol / (P.fugacity_internal * volume / T) * exp(-du / T);
                if (a.prop_cdf) ratio = ratio * t0.bias;
            }
            const bool accepted = uni() < (ratio < 1.0 ? ratio : 1.0);
            s_call = call + (uint64_t)uni.used;
            t0.att[move] += 1;
            if (accepted) t0.acc[move] += 1;
            t0.pend_self = self_d;
            t0.pend_excl = excl_d;
            ctl[2] = accepted ? 1 : 0;
        }
        __syncthreads();
        if (ctl[2]) {
            if (move == 2 && n_mol >= a.max_mol) {
                error = 1;
                overflow = true;
                break;
            }
            if (tid == 0 && CHARGED) {
                t0.self_e += t0.pend_self;
                t0.excl_e += t0.pend_excl;
            }
            // commit S(k) (the same delta, recomputed, as commit* does, :161-172)
            if (CHARGED) {
                for (int t = tid; t < a.nk_half; t += NT) {
                    const int kp = s_kn[t];
                    double dre = 0.0, dim = 0.0;
                    for (int s = 0; s < nsite; ++s) {
                        const cplx e = cmul(cmul(phase[(kp & 0xff) * nsite + s], phase[(nkk + ((kp >> 8) & 0xff)) * nsite + s]),
                                            phase[(2 * nkk + (kp >> 16)) * nsite + s]);
                        const double qs = s < n_rem ? -site4[4 * s + 3] : site4[4 * s + 3];
                        dre += qs * e.re;
                        dim += qs * e.im;
                    }
                    s_re[t] += dre;
                    s_im[t] += dim;
                }
            }
            // positions
            if (move <= 1) {
                if (tid < ns * 3) pos[(size_t)mi * ns * 3 + tid] = newc[tid / 3][tid % 3];
            } else if (move == 2) {
                if (tid < ns * 3) pos[(size_t)n_mol * ns * 3 + tid] = newc[tid / 3][tid % 3];
                n_mol += 1;
            } else {
                // erase molecule mi, keeping the order of the others (:821)
                const size_t first = (size_t)mi * ns * 3, count = (size_t)(n_mol - 1 - mi) * ns * 3;
                for (size_t base = 0; base < count; base += NT) {
                    double v = 0.0;
                    if (base + tid < count) v = pos[first + ns * 3 + base + tid];
                    __syncthreads();
                    if (base + tid < count) pos[first + base + tid] = v;
                    __syncthreads();
                }
                n_mol -= 1;
            }
            __threadfence_block();
            __syncthreads();
            // maybeResyncEwald (:380-403)
            if (CHARGED && ++since_resync >= 100) {
                since_resync = 0;
                // the committed reciprocal energy before the recompute, for the consistency check
                double e_old = 0.0;
                for (int t = tid; t < a.nk_half; t += NT) {
                    e_old += s_kw[t] * (s_re[t] * s_re[t] + s_im[t] * s_im[t]);
                    s_re[t] = a.fw_sre[t];
                    s_im[t] = a.fw_sim[t];
                }
                const int total = n_mol * ns;
                for (int base = 0; base < total; base += kResyncChunk) {
                    const int cnt = min(kResyncChunk, total - base);
                    __syncthreads();
                    for (int i = tid; i < 3 * nkk * kResyncChunk; i += NT) {
                        const int s = i % kResyncChunk, mm = (i / kResyncChunk) % nkk, axis = i / (kResyncChunk * nkk);
                        cplx v = {0.0, 0.0};
                        if (s < cnt) {
                            const double* gp = pos + (size_t)(base + s) * 3;
                            const double gc[3] = {gp[0], gp[1], gp[2]};
                            double gf[3];
                            frac_of(a, gc, gf);
                            double sv, cv;
                            sincospi(2.0 * (double)(mm - a.kmax) * gf[axis], &sv, &cv);
                            v = {cv, sv};
                        }
                        rphase[i] = v;
                    }
                    __syncthreads();
                    for (int t = tid; t < a.nk_half; t += NT) {
                        const int kp = s_kn[t];
                        double dre = 0.0, dim = 0.0;
                        for (int s = 0; s < cnt; ++s) {
                            const cplx e = cmul(cmul(rphase[(kp & 0xff) * kResyncChunk + s],
                                                     rphase[(nkk + ((kp >> 8) & 0xff)) * kResyncChunk + s]),
                                                rphase[(2 * nkk + (kp >> 16)) * kResyncChunk + s]);
                            const double qs = a.charge[(base + s) % ns];
                            dre += qs * e.re;
                            dim += qs * e.im;
                        }
                        s_re[t] += dre;
                        s_im[t] += dim;
                    }
                }
                // each thread re-reads only the terms it updated itself, so no barrier is needed
                double e_new = 0.0;
                for (int t = tid; t < a.nk_half; t += NT) e_new += s_kw[t] * (s_re[t] * s_re[t] + s_im[t] * s_im[t]);
                double r4[4] = {e_old, e_new, 0.0, 0.0};
                block_sum4<NW>(r4, red);
                // self and exclusion from scratch (ewald_incremental_state.cc:230-246)
                if (tid == 0) {
                    const double committed = a.recip_scale * r4[0] + t0.self_e + t0.excl_e;
                    double q2 = 0.0;
                    for (int s = 0; s < ns; ++s) q2 += a.charge[s] * a.charge[s];
                    t0.self_e = a.fw_self + a.self_pref * (q2 * (double)n_mol);
                    double ex = 0.0;
                    for (int g = 0; g < n_mol; ++g)
                        for (int i = 0; i < ns; ++i)
                            for (int j2 = i + 1; j2 < ns; ++j2) {
                                double fi[3], fj[3];
                                const double* pi = pos + ((size_t)g * ns + i) * 3;
                                const double* pj = pos + ((size_t)g * ns + j2) * 3;
                                const double ci[3] = {pi[0], pi[1], pi[2]}, cj[3] = {pj[0], pj[1], pj[2]};
                                frac_of(a, ci, fi);
                                frac_of(a, cj, fj);
                                const double r = sqrt(dist2(a, fj[0] - fi[0], fj[1] - fi[1], fj[2] - fi[2]));
                                ex += a.charge[i] * a.charge[j2] *
                                      (r < 1e-8 ? 2.0 * a.alpha / sqrt(kPi) : erf(a.alpha * r) / r);
                            }
                    t0.excl_e = -a.kc * ex;
                    // the test of test_ewald_incremental_state.cc:137-222, at every resync: relative
                    // drift of the incrementally kept energy from the from-scratch one; a NaN sticks
                    const double fresh = a.recip_scale * r4[1] + t0.self_e + t0.excl_e;
                    const double drift = fabs(fresh - committed) / fmax(1.0, fabs(fresh));
                    if (!isnan(t0.max_drift) && !(drift <= t0.max_drift)) t0.max_drift = drift;
                    t0.resyncs += 1;
                }
                __syncthreads();
            }
        }
        if (tid == 0) {
            t0.sum_n += (double)n_mol;
            if (a.n_trace) a.n_trace[(size_t)chain * a.n_steps + step] = n_mol;
        }
    }
    __syncthreads();
    if (CHARGED)
        for (int t = tid; t < a.nk_half; t += NT) {
            a.sre[(size_t)chain * a.nk_half + t] = s_re[t];
            a.sim[(size_t)chain * a.nk_half + t] = s_im[t];
        }
    if (tid == 0) {
        if (overflow) {
            // the insertion that did not fit: not counted, and its draws are taken back
            t0.att[2] -= 1;
            t0.acc[2] -= 1;
            s_call = t0.step_call;
        }
        st->call = s_call;
        st->n_mol = n_mol;
        st->moves_since_resync = since_resync;
        st->self_energy = t0.self_e;
        st->exclusion_energy = t0.excl_e;
        st->sum_n = t0.sum_n;
        st->steps += step;
        for (int i = 0; i < 4; ++i) {
            st->attempted[i] = t0.att[i];
            st->accepted[i] = t0.acc[i];
        }
        st->error = error;
        st->max_drift = t0.max_drift;
        st->resyncs = t0.resyncs;
    }
}

}  // namespace

void build_framework_bins(const GcmcArgs& a, double reach, int32_t** start, double4** pos, int32_t** type,
                          hipStream_t s) {
    const int nbins = a.fb[0] * a.fb[1] * a.fb[2];
    int32_t* counts = nullptr;
    ALEATOR_HIP_CHECK(hipMalloc(&counts, nbins * sizeof(int32_t)));
    std::vector<int32_t> h(nbins), st(nbins + 1, 0);
    try {
        hipLaunchKernelGGL(bin_count_kernel, dim3(nbins), dim3(256), 0, s, a, reach * reach, counts);
        ALEATOR_HIP_CHECK(hipGetLastError());
        ALEATOR_HIP_CHECK(hipMemcpyAsync(h.data(), counts, nbins * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipFree(counts);
        throw;
    }
    (void)hipFree(counts);
    for (int b = 0; b < nbins; ++b) st[b + 1] = st[b] + h[b];
    ALEATOR_HIP_CHECK(hipMalloc(start, (nbins + 1) * sizeof(int32_t)));
    ALEATOR_HIP_CHECK(hipMalloc(pos, std::max<size_t>(st[nbins], 1) * sizeof(double4)));
    ALEATOR_HIP_CHECK(hipMalloc(type, std::max<size_t>(st[nbins], 1) * sizeof(int32_t)));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(*start, st.data(), (nbins + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bin_fill_kernel, dim3(nbins), dim3(256), 0, s, a, reach * reach, *start, *pos, *type);
    ALEATOR_HIP_CHECK(hipGetLastError());
    ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
}

void launch_gcmc(const GcmcArgs& a, const GcmcArgs* d_args, int n_chains, hipStream_t s) {
    if (a.nk_half > kMaxKHalf) fail(ALEATOR_NOT_IMPLEMENTED, "gcmc: more than 4096 reciprocal terms");
    // guest cells compile in only where used, so the default full scan keeps its registers
    const bool cells = a.gc[0] > 0;
#define LAUNCH(NTH, CH)                                                                                     \
    do {                                                                                                    \
        if (cells)                                                                                          \
            hipLaunchKernelGGL((gcmc_chain_kernel<NTH, CH, true>), dim3(n_chains), dim3(NTH), 0, s, d_args);  \
        else                                                                                                \
            hipLaunchKernelGGL((gcmc_chain_kernel<NTH, CH, false>), dim3(n_chains), dim3(NTH), 0, s, d_args); \
    } while (0)
    if (a.charged)
        LAUNCH(kGThreads, true);
    else if (a.grid)  // per-step work of a few guests: one wave per chain, barriers near free
        LAUNCH(64, false);
    else
        LAUNCH(kGThreads, false);
#undef LAUNCH
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
