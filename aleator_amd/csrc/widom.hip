// widom.hip -- batched Widom test-particle insertion on gfx950: host side, trial sort and the
// reciprocal kernel (the direct-sum kernels are in widom.hpp).
//
// Reference: MonteCarloEngine::widomInsertionHenryCoefficient
// (src/engines/monte_carlo/monte_carlo_engine.cc:850-917), K_H = <exp(-dU/T)>/T.
//
// Kernel 1, the direct sum of LJ and real-space Coulomb terms per trial: the split kernel (the
// threads of a workgroup share each trial's atoms) or, for large batches, the cube kernel (trials
// sorted by cube, each cube's framework images listed in LDS).  Kernel 2 (charged only): the
// reciprocal, self and exclusion terms of EwaldIncrementalState::proposeInsertion
// (ewald_incremental_state.cc:95-136) per trial, over the half k-set runs of recip.hpp.  Both
// write per-workgroup partial sums of exp(-dU/T), reduced in a fixed order by kernel 3.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "recip.hpp"
#include "widom.hpp"

namespace aleator_gpu {

namespace {

using namespace widom_detail;

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
        double cart[NS][3];
        trial_sites<NS>(a, t, cart);
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
// block b takes trials [b per, (b + 1) per), keeps each trial's cube in key[] and histograms the
// cubes in LDS; a column scan over blocks gives each (block, cube) its offset among the cube's
// trials; the block then places its trials, ranked inside (block, cube) by LDS atomics.
constexpr int kSortThreads = 1024;
constexpr uint64_t kSplitWaves = 6144;
constexpr uint32_t kSortMaxCells = 12288;  // 48 KB of LDS counters
constexpr int kColScanCells = 16;          // cubes per column-scan workgroup

__global__ __launch_bounds__(kSortThreads) void cube_hist_kernel(const WidomArgs a, const CellArgs c, uint64_t per,
                                                                 uint32_t n_cells, uint32_t* hist, uint32_t* key) {
    extern __shared__ uint32_t hcount[];
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hcount[i] = 0u;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(lo + per, a.n_trials);
    for (uint64_t t = lo + threadIdx.x; t < hi; t += kSortThreads) {
        const uint32_t k = trial_cell(a, c, t);
        key[t] = k;
        atomicAdd(&hcount[k], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hist[(size_t)blockIdx.x * n_cells + i] = hcount[i];
}

// hist[b][cube] <- trials of the cube in blocks before b; count[cube] <- its total.  One
// workgroup of 256 threads per kColScanCells cubes (n_blocks <= 256): thread b holds block b's
// counts of those cubes, and each cube's column is scanned across the four waves.
__global__ __launch_bounds__(256) void cube_colscan_kernel(uint32_t* hist, uint32_t n_blocks, uint32_t n_cells,
                                                           uint32_t* count) {
    __shared__ int wtot[4][kColScanCells];
    const uint32_t b = threadIdx.x, c0 = blockIdx.x * kColScanCells;
    const int lane = b & 63, w = b >> 6;
    int v[kColScanCells], incl[kColScanCells];
#pragma unroll
    for (int j = 0; j < kColScanCells; ++j) {
        v[j] = (b < n_blocks && c0 + j < n_cells) ? (int)hist[(size_t)b * n_cells + c0 + j] : 0;
        incl[j] = wave_inclusive_scan(v[j]);
        if (lane == 63) wtot[w][j] = incl[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kColScanCells; ++j) {
        if (c0 + j >= n_cells) continue;
        int off = 0;
        for (int q = 0; q < w; ++q) off += wtot[q][j];
        if (b < n_blocks) hist[(size_t)b * n_cells + c0 + j] = (uint32_t)(off + incl[j] - v[j]);
        if (b == 0) count[c0 + j] = (uint32_t)(wtot[0][j] + wtot[1][j] + wtot[2][j] + wtot[3][j]);
    }
}

__global__ __launch_bounds__(kSortThreads) void cube_place_kernel(const WidomArgs a, uint64_t per, uint32_t n_cells,
                                                                  const uint32_t* hist, const uint32_t* start,
                                                                  const uint32_t* key, uint32_t* order) {
    extern __shared__ uint32_t hrank[];
    for (uint32_t i = threadIdx.x; i < n_cells; i += kSortThreads) hrank[i] = 0u;
    __syncthreads();
    const uint32_t* base = hist + (size_t)blockIdx.x * n_cells;
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(lo + per, a.n_trials);
    for (uint64_t t = lo + threadIdx.x; t < hi; t += kSortThreads) {
        const uint32_t k = key[t];
        order[start[k] + base[k] + atomicAdd(&hrank[k], 1u)] = (uint32_t)t;
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
    // orthorhombic cells (rows along the axes) take the Cartesian nearest image in the direct kernel
    if (fw.lat.m[0][1] == 0.0 && fw.lat.m[0][2] == 0.0 && fw.lat.m[1][0] == 0.0 && fw.lat.m[1][2] == 0.0 &&
        fw.lat.m[2][0] == 0.0 && fw.lat.m[2][1] == 0.0 && fw.lat.m[0][0] > 0.0 && fw.lat.m[1][1] > 0.0 &&
        fw.lat.m[2][2] > 0.0)
        for (int i = 0; i < 3; ++i) {
            a.box[i] = fw.lat.m[i][i];
            a.inv_box[i] = 1.0 / fw.lat.m[i][i];
        }
    a.geo = a.box[0] <= 0.0 ? kGeoTriclinic
          : (a.box[0] == a.box[1] && a.box[1] == a.box[2] ? kGeoCubic : kGeoOrtho);
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
    if (fw.lj_tail) {
        std::vector<uint32_t> site_species;
        for (const auto& s : req.sites) site_species.push_back(s.lj_species);
        a.tail_delta = lj_tail_delta(fw, site_species.data(), ns);
    }

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
    // split kernel: the fewest threads per trial (1 to 16) that give at least kSplitWaves waves
    // (6 per SIMD), so that small batches still fill the chip; 64 trials per wave group
    const uint64_t tpl = (uint64_t)split_trials_per_lane(ns);
    int split = 1;
    while (split < 16 && req.n_trials * (uint64_t)split < 64ull * tpl * kSplitWaves) split *= 2;
    a.split = split;
    const uint64_t split_trials_per_block = (uint64_t)split_block_threads(split) / split * tpl;
    const unsigned split_blocks = (unsigned)((req.n_trials + split_trials_per_block - 1) / split_trials_per_block);
    const int rec_rt = charged ? runs_molecules_per_block(ns, fw.kmax) : 1;
    if (rec_rt == 0) fail(ALEATOR_NOT_IMPLEMENTED, "widom: phase tables exceed the LDS budget");
    const unsigned rec_blocks = (unsigned)((req.n_trials + rec_rt - 1) / rec_rt);
    const size_t tab_bytes = (tab.size() * sizeof(double) + 255) & ~size_t(255);
    const size_t parts = charged ? rec_blocks : (n_cells ? pair_blocks : split_blocks);
    size_t du_bytes = req.du_device ? 0 : ((req.n_trials * sizeof(double) + 255) & ~size_t(255));
    // LDS-histogram sort: about 4096 trials per block, at most 256 blocks
    const bool lds_sort = n_cells > 0 && n_cells <= kSortMaxCells;
    const uint32_t sort_blocks = lds_sort ? (uint32_t)std::min<uint64_t>(256, (req.n_trials + 4095) / 4096) : 0;
    const size_t sort_bytes =
        n_cells ? ((2 * req.n_trials + 3 * (size_t)n_cells + 1 + (size_t)sort_blocks * n_cells) * sizeof(uint32_t) +
                   255) & ~size_t(255)
                : 0;
    char* base = static_cast<char*>(ctx.ensure_scratch(tab_bytes + parts * sizeof(double) + du_bytes + sort_bytes));
    // the LJ constants depend only on the framework and the site species: uploaded once per species list
    std::vector<uint32_t> species_key;
    for (const auto& site : req.sites) species_key.push_back(site.lj_species);
    if (!fw.d_widom_tab) ALEATOR_HIP_CHECK(hipMalloc(&fw.d_widom_tab, tab.size() * sizeof(double)));
    if (species_key != fw.widom_tab_species) {
        ALEATOR_HIP_CHECK(hipMemcpyAsync(fw.d_widom_tab, tab.data(), tab.size() * sizeof(double),
                                         hipMemcpyHostToDevice, ctx.stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx.stream));
        fw.widom_tab_species = species_key;
    }
    a.ljtab = fw.d_widom_tab;
    a.block_sums = reinterpret_cast<double*>(base + tab_bytes);
    a.du = req.du_device ? req.du_device
                         : reinterpret_cast<double*>(base + tab_bytes + parts * sizeof(double));
    const bool fp64 = req.precision == ALEATOR_PRECISION_FP64;
    if (!fp64 && req.precision != ALEATOR_PRECISION_MIXED) fail(ALEATOR_INVALID_ARGUMENT, "widom: unknown precision");
    // erf kernel polynomial of the real-space term (as in the grid build: FP64 to the last bits of
    // a term, MIXED to 1e-12 of it)
    const int deg = charged ? erf_kernel_poly(fw.alpha, fw.rc_real, fp64 ? 1e-17 : 1e-12, 0, a.poly, &a.poly_xs) : 0;
    const int ljk = a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? 2 : (a.trunc == ALEATOR_LJ_SHIFTED ? 1 : 0);
    auto direct = [&](const CellArgs* c, unsigned blocks) {
        using Direct = void (*)(const WidomArgs&, const CellArgs*, unsigned, bool, unsigned, int, hipStream_t);
        static constexpr Direct kDirect[2][3] = {
            {widom_direct_mixed_lj0, widom_direct_mixed_lj1, widom_direct_mixed_lj2},
            {widom_direct_fp64_lj0, widom_direct_fp64_lj1, widom_direct_fp64_lj2}};
        kDirect[fp64 ? 1 : 0][ljk](a, c, n_cells, wide_cells, blocks, deg, ctx.stream);
    };
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
                               n_cells, hist, key);
            hipLaunchKernelGGL(cube_colscan_kernel, dim3((n_cells + kColScanCells - 1) / kColScanCells), dim3(256), 0,
                               ctx.stream, hist, sort_blocks, n_cells, count);
            hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, ctx.stream, count, n_cells, start, cursor);
            hipLaunchKernelGGL(cube_place_kernel, dim3(sort_blocks), dim3(kSortThreads), lds, ctx.stream, a, per,
                               n_cells, hist, start, key, order);
        } else {
            hipLaunchKernelGGL(fill_u32_kernel, dim3((n_cells + 255) / 256), dim3(256), 0, ctx.stream, count,
                               (uint64_t)n_cells);
            hipLaunchKernelGGL(widom_key_kernel, dim3(tb), dim3(256), 0, ctx.stream, a, cg, key, count);
            hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, ctx.stream, count, n_cells, start, cursor);
            hipLaunchKernelGGL(scatter_kernel, dim3(tb), dim3(256), 0, ctx.stream, key, req.n_trials, cursor, order);
        }
        direct(&cg, 0);
        if (charged) {
            launch_recip(a, fw, rec_blocks, rec_rt, ctx.stream);
        } else {
            hipLaunchKernelGGL(widom_boltz_kernel, dim3(pair_blocks), dim3(kWThreads), 0, ctx.stream, a);
        }
    } else {
        direct(nullptr, split_blocks);
        if (charged) launch_recip(a, fw, rec_blocks, rec_rt, ctx.stream);
    }
    ctx.time_end();
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, ctx.stream, a.block_sums, (uint64_t)parts,
                       req.sum_device);
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
