// gcmc_api.hip -- C ABI of the batched GCMC chains (include/aleator_gpu.h).
//
// Host side of MonteCarloEngine's GCMC constructor checks (monte_carlo_engine.cc:174-297):
// positive temperature and fugacity, a neutral species when the framework is charged; the
// fugacity conversion Pa -> K/A^3 (monte_carlo_engine.hpp:184); per-chain device state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "aleator_gpu.h"
#include "gcmc.hpp"
#include "internal.hpp"

struct aleator_gcmc {
    aleator_gpu::Framework* fw = nullptr;
    int n_chains = 0;
    uint32_t site_species[aleator_gpu::kMaxSites] = {};  // LJ species of each site
    aleator_gpu::GcmcArgs args{};
    aleator_gpu::ChainParams* d_params = nullptr;
    aleator_gpu::GcmcArgs* d_args = nullptr;  // device copy of the launch arguments
    aleator_gpu::ChainState* d_state = nullptr;
    double* d_pos = nullptr;
    double* d_sre = nullptr;
    double* d_sim = nullptr;
    double* d_ljtab = nullptr;
    double* d_ljgg = nullptr;
    double* d_prop = nullptr;  // [2][n_bins]: cdf, then n_bins * p
    int32_t* d_bin_start = nullptr;
    double4* d_bin_pos = nullptr;
    int32_t* d_bin_type = nullptr;
    double* d_grid = nullptr;  // energy-grid tables owned by the chains (a copy of the caller's)
    // run without a trace replays one captured launch (recaptured when the launch arguments change)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    aleator_gpu::GcmcArgs graph_args{};
};

namespace aleator_gpu {
namespace {
std::vector<ChainState> download_states(aleator_gcmc* g) {
    std::vector<ChainState> st(g->n_chains);
    ALEATOR_HIP_CHECK(hipMemcpyAsync(st.data(), g->d_state, st.size() * sizeof(ChainState), hipMemcpyDeviceToHost,
                                     g->fw->ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(g->fw->ctx->stream));
    return st;
}

// Bin probabilities p = (1 - mix) w / sum(w) + mix / n, their running sum (the CDF, last entry
// set to 1) and the acceptance factors n (cdf[b] - cdf[b-1]): the probability the inverse-CDF
// draw actually gives bin b, so the correction matches the sampler exactly.
void proposal_tables(const double* w, size_t n, double mix, std::vector<double>& cdf, std::vector<double>& fac) {
#pragma clang fp contract(off)
    double total = 0.0;
    for (size_t b = 0; b < n; ++b) {
        require(std::isfinite(w[b]) && w[b] >= 0.0, ALEATOR_INVALID_ARGUMENT,
                "gcmc proposal: weights must be finite and non-negative");
        total += w[b];
    }
    require(total > 0.0 && std::isfinite(total), ALEATOR_INVALID_ARGUMENT, "gcmc proposal: weights sum to zero");
    cdf.resize(n);
    fac.resize(n);
    double run = 0.0;
    for (size_t b = 0; b < n; ++b) {
        run += (1.0 - mix) * (w[b] / total) + mix / (double)n;
        cdf[b] = run;
    }
    cdf[n - 1] = 1.0;
    double prev = 0.0;
    for (size_t b = 0; b < n; ++b) {
        require(cdf[b] > prev, ALEATOR_INVALID_ARGUMENT,
                "gcmc proposal: a bin has zero probability (raise uniform_mix or its weight)");
        fac[b] = (double)n * (cdf[b] - prev);
        prev = cdf[b];
    }
}

// Installs energy-grid tables (device or host memory, by `kind`) as the chains' own copy.
void set_energy_grid(aleator_gcmc* g, const int32_t n_points[3], const uint32_t* species, int32_t n_tables,
                     const double* tables, hipMemcpyKind kind) {
    require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = g->fw->ctx->stream;
    if (!tables) {
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
        if (g->d_grid) ALEATOR_HIP_CHECK(hipFree(g->d_grid));
        g->d_grid = nullptr;
        g->args.grid = nullptr;
        return;
    }
    require(n_points && species, ALEATOR_INVALID_ARGUMENT, "null argument");
    require(n_tables >= 1 && n_tables <= 256, ALEATOR_INVALID_ARGUMENT, "gcmc energy grid: 1 to 256 tables");
    size_t cells = 1;
    for (int i = 0; i < 3; ++i) {
        require(n_points[i] >= 1, ALEATOR_INVALID_ARGUMENT, "gcmc energy grid: point counts must be >= 1");
        cells *= (size_t)n_points[i];
    }
    require(cells <= (size_t)1 << 31, ALEATOR_INVALID_ARGUMENT, "gcmc energy grid: at most 2^31 nodes per table");
    const Framework& fw = *g->fw;
    GcmcArgs& a = g->args;
    // FrameworkEnergyGrid::interpolate throws std::out_of_range for a species without a table
    // (framework_energy_grid.cc:254-257); the check is made once here for every site
    int table[kMaxSites];
    for (int k = 0; k < a.ns; ++k) {
        table[k] = -1;
        for (int t = 0; t < n_tables && table[k] < 0; ++t)
            if (species[t] == g->site_species[k]) table[k] = t;
        require(table[k] >= 0, ALEATOR_OUT_OF_RANGE, "FrameworkEnergyGrid: no grid table for a site's species");
    }
    double* d = nullptr;
    ALEATOR_HIP_CHECK(hipMalloc(&d, (size_t)n_tables * cells * sizeof(double)));
    try {
        ALEATOR_HIP_CHECK(hipMemcpyAsync(d, tables, (size_t)n_tables * cells * sizeof(double), kind, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipFree(d);
        throw;
    }
    if (g->d_grid) ALEATOR_HIP_CHECK(hipFree(g->d_grid));
    g->d_grid = d;
    a.grid = d;
    a.grid_cells = cells;
    GridGeometry geo{};
    grid_cofactors(fw.lat.m, geo);
    geo.volume = fw.lat.volume;
    for (int i = 0; i < 3; ++i) geo.n[i] = n_points[i];
    a.grid_geo = geo;
    for (int k = 0; k < a.ns; ++k) a.grid_table[k] = table[k];
}
}  // namespace
}  // namespace aleator_gpu

using namespace aleator_gpu;

namespace {

// Guest cells for the guest-guest candidate filter.  mode 0: none; 1: only when every axis has
// at least 4 cells (with 3 or fewer each cell's neighbours are all cells), and in a step only
// from kGuestCellMinSites guest sites on (below, the filter pass and its barriers cost more than
// the pairs it saves); 2: whenever the guest sites fit the filter's capacity, in every step.  A cell is at least the guest cutoff (LJ, and real-space
// Coulomb when charged) wide perpendicular to its lattice plane: L_perp_j / cells_j >= rc.
void set_guest_cells(aleator_gpu::GcmcArgs& a, int mode) {
    using namespace aleator_gpu;
    a.gc[0] = a.gc[1] = a.gc[2] = 0;
    a.gc_min_sites = mode == 1 ? kGuestCellMinSites : 0;
    if (mode == 0) return;
    const bool fits = (int64_t)a.max_mol * a.ns <= kMaxGuestCand;
    require(fits || mode == 1, ALEATOR_OUT_OF_RANGE, "gcmc: guest cells hold at most 4096 guest sites");
    if (!fits) return;
    const double rc = std::sqrt(a.charged ? std::max(a.rc2_lj, a.rc2_real) : a.rc2_lj);
    int n[3];
    for (int j = 0; j < 3; ++j) {
        // |b_j|, b_j = column j of the inverse lattice, is 1 / L_perp_j
        const double b = std::sqrt(a.inv[j] * a.inv[j] + a.inv[3 + j] * a.inv[3 + j] + a.inv[6 + j] * a.inv[6 + j]);
        n[j] = std::max(1, (int)std::floor(1.0 / (b * rc)));
    }
    if (mode == 1 && std::min({n[0], n[1], n[2]}) < 4) return;
    for (int j = 0; j < 3; ++j) a.gc[j] = n[j];
}

}  // namespace

extern "C" {

aleator_status aleator_gpu_gcmc_create(aleator_framework* fw_in, const aleator_molecule_site* sites, int32_t n_sites,
                                       const aleator_gcmc_chain* chains, int32_t n_chains, double max_translation,
                                       double max_rotation, int32_t max_molecules, aleator_gcmc** out) {
    return guarded([&] {
        auto* fw = reinterpret_cast<Framework*>(fw_in);
        require(fw && sites && chains && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(n_sites >= 1, ALEATOR_INVALID_ARGUMENT, "MonteCarloEngine: every GCMC species must have at least one site");
        require(n_sites <= kMaxSites, ALEATOR_NOT_IMPLEMENTED, "gcmc: at most 8 sites per molecule");
        require(n_chains >= 1, ALEATOR_INVALID_ARGUMENT, "gcmc: at least one chain");
        require(max_molecules >= 1, ALEATOR_INVALID_ARGUMENT, "gcmc: max_molecules must be positive");
        double qsum = 0.0;
        for (int s = 0; s < n_sites; ++s) {
            require(sites[s].lj_species < fw->eps.size(), ALEATOR_OUT_OF_RANGE, "site LJ species out of range");
            qsum += sites[s].charge;
        }
        if (fw->charged)
            require(!(std::abs(qsum) > 1e-8), ALEATOR_INVALID_ARGUMENT,
                    "MonteCarloEngine: GCMC species is not individually charge-neutral");
        std::vector<ChainParams> params(n_chains);
        for (int c = 0; c < n_chains; ++c) {
            require(chains[c].temperature > 0.0, ALEATOR_INVALID_ARGUMENT,
                    "MonteCarloEngine: temperatureKelvin must be positive");
            const double f = chains[c].fugacity_pa * (1.0 / (1.380649e-23 * 1.0e30));
            require(f > 0.0, ALEATOR_INVALID_ARGUMENT, "MonteCarloEngine: every fugacityPascal must be positive");
            params[c] = {f, chains[c].temperature, chains[c].seed, chains[c].stream};
        }
        auto g = std::make_unique<aleator_gcmc>();
        g->fw = fw;
        g->n_chains = n_chains;
        for (int k = 0; k < n_sites; ++k) g->site_species[k] = sites[k].lj_species;
        GcmcArgs& a = g->args;
        hipStream_t s = fw->ctx->stream;
        a.fx = fw->d_fx;
        a.fy = fw->d_fy;
        a.fz = fw->d_fz;
        a.q = fw->d_q;
        a.type = fw->d_type;
        a.natoms = (int)fw->n;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                a.m[3 * i + j] = fw->lat.m[i][j];
                a.inv[3 * i + j] = fw->lat.inv[i][j];
            }
        a.volume = fw->lat.volume;
        a.rc_lj = fw->rc;
        a.rc2_lj = fw->rc * fw->rc;
        a.trunc = fw->trunc;
        a.charged = fw->charged ? 1 : 0;
        a.alpha = fw->alpha;
        a.rc2_real = fw->rc_real * fw->rc_real;
        if (fw->charged)
            a.erf_deg = erf_kernel_poly(fw->alpha, fw->rc_real, 1e-17, 0, a.erf_poly, &a.erf_xs);
        a.kc = kCoulomb;
        a.recip_scale = 2.0 * kCoulomb * (2.0 * kPi / std::abs(fw->lat.volume));
        a.self_pref = -kCoulomb * (fw->alpha / std::sqrt(kPi));
        a.fw_self = fw->self_energy;
        a.kn = reinterpret_cast<const int4*>(fw->d_kn);
        a.kw = fw->d_kw;
        a.fw_sre = fw->d_sre;
        a.fw_sim = fw->d_sim;
        a.nk_half = fw->nk_half;
        a.kmax = fw->kmax;
        a.ns = n_sites;
        for (int k = 0; k < n_sites; ++k) {
            a.local[k][0] = sites[k].x;
            a.local[k][1] = sites[k].y;
            a.local[k][2] = sites[k].z;
            a.charge[k] = sites[k].charge;
        }
        a.max_translation = max_translation;
        a.max_rotation = max_rotation;
        a.max_mol = max_molecules;
        set_guest_cells(a, 0);
        // LJ constants: site vs framework type, and site vs site
        std::vector<double> tab(4 * kMaxSites * kMaxTypes, 0.0), gg(4 * kMaxSites * kMaxSites, 0.0);
        for (int k = 0; k < n_sites; ++k) {
            for (size_t t = 0; t < fw->type_species.size(); ++t) {
                double v[4];
                lj_pair_constants(*fw, sites[k].lj_species, fw->type_species[t], v[0], v[1], v[2], v[3]);
                for (int c = 0; c < 4; ++c) tab[(c * kMaxSites + k) * kMaxTypes + t] = v[c];
            }
            for (int u = 0; u < n_sites; ++u) {
                double v[4];
                lj_pair_constants(*fw, sites[k].lj_species, sites[u].lj_species, v[0], v[1], v[2], v[3]);
                for (int c = 0; c < 4; ++c) gg[(c * kMaxSites + k) * kMaxSites + u] = v[c];
            }
        }
        auto upload = [&](auto** dst, const auto* src, size_t n) {
            using T = std::remove_pointer_t<std::remove_pointer_t<decltype(dst)>>;
            ALEATOR_HIP_CHECK(hipMalloc(dst, std::max<size_t>(n, 1) * sizeof(T)));
            if (n) ALEATOR_HIP_CHECK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
        };
        try {
            upload(&g->d_ljtab, tab.data(), tab.size());
            upload(&g->d_ljgg, gg.data(), gg.size());
            upload(&g->d_params, params.data(), params.size());
            ALEATOR_HIP_CHECK(hipMalloc(&g->d_args, sizeof(GcmcArgs)));
            std::vector<ChainState> st(n_chains);
            std::memset(st.data(), 0, st.size() * sizeof(ChainState));
            for (auto& c : st) c.self_energy = fw->charged ? fw->self_energy : 0.0;
            upload(&g->d_state, st.data(), st.size());
            ALEATOR_HIP_CHECK(hipMalloc(&g->d_pos, (size_t)n_chains * max_molecules * n_sites * 3 * sizeof(double)));
            if (fw->charged) {
                const size_t nk = (size_t)fw->nk_half;
                ALEATOR_HIP_CHECK(hipMalloc(&g->d_sre, n_chains * nk * sizeof(double)));
                ALEATOR_HIP_CHECK(hipMalloc(&g->d_sim, n_chains * nk * sizeof(double)));
                for (int c = 0; c < n_chains; ++c) {
                    ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_sre + c * nk, fw->d_sre, nk * sizeof(double),
                                                     hipMemcpyDeviceToDevice, s));
                    ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_sim + c * nk, fw->d_sim, nk * sizeof(double),
                                                     hipMemcpyDeviceToDevice, s));
                }
            }
            ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
            // framework bin lists: bins about 2 A wide; every atom within rc of a point in a bin is
            // within rc + h of its centre (h: half the longest bin diagonal); used only while that
            // reach stays below half the smallest perpendicular width, where the fractional
            // minimum image is exact
            const double rcmax = std::max(fw->rc, fw->charged ? fw->rc_real : 0.0);
            double wmin = fw->lat.lperp[0];
            for (int i = 0; i < 3; ++i) {
                a.fb[i] = std::clamp((int)(fw->lat.lperp[i] / 2.0), 1, 64);
                wmin = std::min(wmin, fw->lat.lperp[i]);
            }
            double hmax = 0.0;
            const int sg[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
            for (auto& d : sg) {
                double v[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) v[j] += d[i] * fw->lat.m[i][j] / a.fb[i];
                hmax = std::max(hmax, 0.5 * std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
            }
            const double reach = rcmax + hmax;
            if (reach < 0.5 * wmin) {
                build_framework_bins(a, reach, &g->d_bin_start, &g->d_bin_pos, &g->d_bin_type, s);
            }
        } catch (...) {
            for (void* p : {(void*)g->d_ljtab, (void*)g->d_ljgg, (void*)g->d_params, (void*)g->d_args,
                            (void*)g->d_state, (void*)g->d_pos, (void*)g->d_sre, (void*)g->d_sim,
                            (void*)g->d_bin_start, (void*)g->d_bin_pos, (void*)g->d_bin_type})
                if (p) (void)hipFree(p);
            throw;
        }
        a.bin_start = g->d_bin_start;
        a.bin_pos = g->d_bin_pos;
        a.bin_type = g->d_bin_type;
        a.ljtab = g->d_ljtab;
        a.ljgg = g->d_ljgg;
        a.params = g->d_params;
        a.state = g->d_state;
        a.mol_pos = g->d_pos;
        a.sre = g->d_sre;
        a.sim = g->d_sim;
        retain(fw);
        *out = g.release();
    });
}

aleator_status aleator_gpu_gcmc_run(aleator_gcmc* g, uint64_t n_steps, int32_t* trace_device) {
    return guarded([&] {
        require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        if (n_steps == 0) return;
        GcmcArgs a = g->args;
        a.n_steps = n_steps;
        a.n_trace = trace_device;
        Context& ctx = *g->fw->ctx;
        // rows a chain does not reach (it stopped on max_molecules) read -1
        if (trace_device)
            ALEATOR_HIP_CHECK(hipMemsetAsync(trace_device, 0xff, (size_t)g->n_chains * n_steps * sizeof(int32_t),
                                             ctx.stream));
        // the kernel reads its arguments from the device copy, refreshed before every launch (a
        // pageable copy: `a` is read before the call returns), never from inside a captured graph
        ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_args, &a, sizeof(a), hipMemcpyHostToDevice, ctx.stream));
        if (g->use_graph && !trace_device) {
            if (!g->graph_exec || std::memcmp(&g->graph_args, &a, sizeof(a)) != 0) {
                if (g->graph_exec) ALEATOR_HIP_CHECK(hipGraphExecDestroy(g->graph_exec));
                if (g->graph) ALEATOR_HIP_CHECK(hipGraphDestroy(g->graph));
                g->graph_exec = nullptr;
                g->graph = nullptr;
                ALEATOR_HIP_CHECK(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
                try {
                    launch_gcmc(a, g->d_args, g->n_chains, ctx.stream);
                } catch (...) {
                    hipGraph_t dead = nullptr;
                    (void)hipStreamEndCapture(ctx.stream, &dead);
                    if (dead) (void)hipGraphDestroy(dead);
                    throw;
                }
                ALEATOR_HIP_CHECK(hipStreamEndCapture(ctx.stream, &g->graph));
                ALEATOR_HIP_CHECK(hipGraphInstantiate(&g->graph_exec, g->graph, nullptr, nullptr, 0));
                g->graph_args = a;
            }
            ctx.time_begin();
            ALEATOR_HIP_CHECK(hipGraphLaunch(g->graph_exec, ctx.stream));
            ctx.time_end();
            return;
        }
        ctx.time_begin();
        launch_gcmc(a, g->d_args, g->n_chains, ctx.stream);
        ctx.time_end();
    });
}

aleator_status aleator_gpu_gcmc_use_graph(aleator_gcmc* g, int32_t enable) {
    return guarded([&] {
        require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        g->use_graph = enable != 0;
    });
}

aleator_status aleator_gpu_gcmc_set_proposal(aleator_gcmc* g, const double* w, const int32_t nb[3],
                                             const int32_t reps[3], double uniform_mix) {
    return guarded([&] {
        require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        hipStream_t s = g->fw->ctx->stream;
        if (!w) {
            ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
            if (g->d_prop) ALEATOR_HIP_CHECK(hipFree(g->d_prop));
            g->d_prop = nullptr;
            g->args.prop_cdf = g->args.prop_fac = nullptr;
            g->args.prop_nbins = 0;
            return;
        }
        require(nb && reps, ALEATOR_INVALID_ARGUMENT, "null argument");
        size_t n = 1;
        for (int i = 0; i < 3; ++i) {
            require(nb[i] >= 1 && reps[i] >= 1, ALEATOR_INVALID_ARGUMENT, "gcmc proposal: bins and images must be >= 1");
            n *= (size_t)nb[i];
        }
        require(n <= (1u << 26), ALEATOR_INVALID_ARGUMENT, "gcmc proposal: at most 2^26 bins");
        require((size_t)reps[0] * reps[1] * reps[2] <= (1u << 20), ALEATOR_INVALID_ARGUMENT,
                "gcmc proposal: at most 2^20 images");
        require(uniform_mix >= 0.0 && uniform_mix <= 1.0, ALEATOR_INVALID_ARGUMENT,
                "gcmc proposal: uniform_mix must be in [0, 1]");
        const GcmcArgs& a = g->args;
        for (int j = 0; j < 3; ++j) {
            double c = 0.0;
            for (int k = 0; k < a.ns; ++k) c += a.local[k][j];
            require(std::abs(c / a.ns) <= 1e-9, ALEATOR_INVALID_ARGUMENT,
                    "gcmc proposal: the species' site offsets must be centred (their mean is the anchor)");
        }
        std::vector<double> cdf, fac;
        proposal_tables(w, n, uniform_mix, cdf, fac);
        double* d = nullptr;
        ALEATOR_HIP_CHECK(hipMalloc(&d, 2 * n * sizeof(double)));
        try {
            ALEATOR_HIP_CHECK(hipMemcpyAsync(d, cdf.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
            ALEATOR_HIP_CHECK(hipMemcpyAsync(d + n, fac.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
            ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        if (g->d_prop) ALEATOR_HIP_CHECK(hipFree(g->d_prop));
        g->d_prop = d;
        GcmcArgs& m = g->args;
        m.prop_cdf = d;
        m.prop_fac = d + n;
        m.prop_nbins = (int)n;
        for (int i = 0; i < 3; ++i) {
            m.prop_nb[i] = nb[i];
            m.prop_reps[i] = reps[i];
        }
    });
}

aleator_status aleator_gpu_gcmc_set_energy_grid(aleator_gcmc* g, const int32_t n_points[3], const uint32_t* species,
                                               int32_t n_tables, const double* tables_device) {
    return guarded([&] { set_energy_grid(g, n_points, species, n_tables, tables_device, hipMemcpyDeviceToDevice); });
}

aleator_status aleator_gpu_gcmc_set_energy_grid_host(aleator_gcmc* g, const int32_t n_points[3],
                                                    const uint32_t* species, int32_t n_tables, const double* tables) {
    return guarded([&] { set_energy_grid(g, n_points, species, n_tables, tables, hipMemcpyHostToDevice); });
}

aleator_status aleator_gpu_gcmc_stats(aleator_gcmc* g, aleator_gcmc_stats* out) {
    return guarded([&] {
        require(g && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto st = download_states(g);
        for (int c = 0; c < g->n_chains; ++c) {
            out[c].n_molecules = st[c].n_mol;
            out[c].error = st[c].error;
            out[c].steps = st[c].steps;
            out[c].sum_molecules = st[c].sum_n;
            for (int i = 0; i < 4; ++i) {
                out[c].attempted[i] = st[c].attempted[i];
                out[c].accepted[i] = st[c].accepted[i];
            }
        }
    });
}

aleator_status aleator_gpu_gcmc_reset_stats(aleator_gcmc* g) {
    return guarded([&] {
        require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        auto st = download_states(g);
        for (auto& c : st) {
            c.sum_n = 0.0;
            c.steps = 0;
            std::memset(c.attempted, 0, sizeof(c.attempted));
            std::memset(c.accepted, 0, sizeof(c.accepted));
            c.max_drift = 0.0;
            c.resyncs = 0;
        }
        ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_state, st.data(), st.size() * sizeof(ChainState), hipMemcpyHostToDevice,
                                         g->fw->ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(g->fw->ctx->stream));
    });
}

aleator_status aleator_gpu_gcmc_positions(aleator_gcmc* g, int32_t chain, double* xyz, uint64_t capacity,
                                          int32_t* n_molecules) {
    return guarded([&] {
        require(g && n_molecules, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(chain >= 0 && chain < g->n_chains, ALEATOR_OUT_OF_RANGE, "chain index out of range");
        const auto st = download_states(g);
        const int n = st[chain].n_mol;
        *n_molecules = n;
        const size_t need = (size_t)n * g->args.ns * 3;
        if (!xyz) return;
        require(capacity >= need, ALEATOR_INVALID_ARGUMENT, "positions: capacity too small");
        ALEATOR_HIP_CHECK(hipMemcpyAsync(xyz, g->d_pos + (size_t)chain * g->args.max_mol * g->args.ns * 3,
                                         need * sizeof(double), hipMemcpyDeviceToHost, g->fw->ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(g->fw->ctx->stream));
    });
}

aleator_status aleator_gpu_gcmc_electrostatic_energy(aleator_gcmc* g, int32_t chain, double* energy) {
    return guarded([&] {
        require(g && energy, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(chain >= 0 && chain < g->n_chains, ALEATOR_OUT_OF_RANGE, "chain index out of range");
        Framework& fw = *g->fw;
        if (!fw.charged) {
            *energy = 0.0;
            return;
        }
        const auto st = download_states(g);
        const size_t nkh = (size_t)fw.nk_half, nk = fw.kset.size();
        std::vector<double> hr(nkh), hi(nkh);
        ALEATOR_HIP_CHECK(hipMemcpyAsync(hr.data(), g->d_sre + chain * nkh, nkh * sizeof(double),
                                         hipMemcpyDeviceToHost, fw.ctx->stream));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(hi.data(), g->d_sim + chain * nkh, nkh * sizeof(double),
                                         hipMemcpyDeviceToHost, fw.ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(fw.ctx->stream));
        // committedEnergy (ewald_incremental_state.cc:85-93) over the full reference-order set
        double sum = 0.0;
        for (size_t t = 0; t < nk; ++t) {
            const size_t h = t < nkh ? t : nk - 1 - t;
            sum += fw.kset.weight[t] * (hr[h] * hr[h] + hi[h] * hi[h]);
        }
        const double recip = sum * kCoulomb * (2.0 * kPi / std::abs(fw.lat.volume));
        *energy = recip + st[chain].self_energy + st[chain].exclusion_energy;
    });
}

aleator_status aleator_gpu_gcmc_set_guest_cells(aleator_gcmc* g, int32_t mode) {
    return guarded([&] {
        require(g != nullptr, ALEATOR_INVALID_ARGUMENT, "null chains");
        require(mode >= 0 && mode <= 2, ALEATOR_INVALID_ARGUMENT, "gcmc: guest cell mode is 0, 1 or 2");
        set_guest_cells(g->args, mode);
    });
}

aleator_status aleator_gpu_gcmc_guest_cells(const aleator_gcmc* g, int32_t cells[3]) {
    return guarded([&] {
        require(g != nullptr && cells != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        for (int j = 0; j < 3; ++j) cells[j] = g->args.gc[j];
    });
}

aleator_status aleator_gpu_gcmc_resync_check(aleator_gcmc* g, double* max_drift, uint64_t* resyncs) {
    return guarded([&] {
        require(g && max_drift && resyncs, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto st = download_states(g);
        for (int c = 0; c < g->n_chains; ++c) {
            max_drift[c] = st[c].max_drift;
            resyncs[c] = st[c].resyncs;
        }
    });
}

}  // extern "C"

namespace aleator_gpu {
namespace {
// Chain checkpoint record (layout documented with aleator_gpu_gcmc_save_chain).
constexpr uint32_t kCheckpointMagic = 0x4b434741u;  // "AGCK"
constexpr uint32_t kCheckpointVersion = 1;
constexpr size_t kCheckpointHeader = 144;

size_t checkpoint_bytes(const aleator_gcmc* g, int n_mol) {
    const size_t nkh = g->fw->charged ? (size_t)g->fw->nk_half : 0;
    return kCheckpointHeader + 16 * nkh + 24 * (size_t)n_mol * g->args.ns;
}

template <class T>
void put(unsigned char* buf, size_t off, T v) {
    std::memcpy(buf + off, &v, sizeof(T));
}
template <class T>
T get(const unsigned char* buf, size_t off) {
    T v;
    std::memcpy(&v, buf + off, sizeof(T));
    return v;
}
}  // namespace
}  // namespace aleator_gpu

extern "C" {

aleator_status aleator_gpu_gcmc_checkpoint_size(aleator_gcmc* g, int32_t chain, uint64_t* bytes) {
    return guarded([&] {
        require(g && bytes, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(chain >= 0 && chain < g->n_chains, ALEATOR_OUT_OF_RANGE, "chain index out of range");
        *bytes = checkpoint_bytes(g, download_states(g)[chain].n_mol);
    });
}

aleator_status aleator_gpu_gcmc_save_chain(aleator_gcmc* g, int32_t chain, void* buffer, uint64_t capacity,
                                           uint64_t* written) {
    return guarded([&] {
        require(g && buffer && written, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(chain >= 0 && chain < g->n_chains, ALEATOR_OUT_OF_RANGE, "chain index out of range");
        const ChainState st = download_states(g)[chain];
        const size_t need = checkpoint_bytes(g, st.n_mol);
        require(capacity >= need, ALEATOR_INVALID_ARGUMENT, "gcmc checkpoint: buffer too small");
        auto* b = static_cast<unsigned char*>(buffer);
        const int nkh = g->fw->charged ? g->fw->nk_half : 0;
        const int ns = g->args.ns;
        put<uint32_t>(b, 0, kCheckpointMagic);
        put<uint32_t>(b, 4, kCheckpointVersion);
        put<int32_t>(b, 8, ns);
        put<int32_t>(b, 12, st.n_mol);
        put<int32_t>(b, 16, nkh);
        put<int32_t>(b, 20, st.moves_since_resync);
        put<uint64_t>(b, 24, st.call);
        put<double>(b, 32, st.self_energy);
        put<double>(b, 40, st.exclusion_energy);
        put<double>(b, 48, st.sum_n);
        put<uint64_t>(b, 56, st.steps);
        for (int i = 0; i < 4; ++i) {
            put<uint64_t>(b, 64 + 8 * i, st.attempted[i]);
            put<uint64_t>(b, 96 + 8 * i, st.accepted[i]);
        }
        put<double>(b, 128, st.max_drift);
        put<uint64_t>(b, 136, st.resyncs);
        hipStream_t s = g->fw->ctx->stream;
        size_t off = kCheckpointHeader;
        if (nkh) {
            ALEATOR_HIP_CHECK(hipMemcpyAsync(b + off, g->d_sre + (size_t)chain * nkh, nkh * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
            off += nkh * sizeof(double);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(b + off, g->d_sim + (size_t)chain * nkh, nkh * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
            off += nkh * sizeof(double);
        }
        if (st.n_mol)
            ALEATOR_HIP_CHECK(hipMemcpyAsync(b + off, g->d_pos + (size_t)chain * g->args.max_mol * ns * 3,
                                             (size_t)st.n_mol * ns * 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
        *written = need;
    });
}

aleator_status aleator_gpu_gcmc_load_chain(aleator_gcmc* g, int32_t chain, const void* buffer, uint64_t size) {
    return guarded([&] {
        require(g && buffer, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(chain >= 0 && chain < g->n_chains, ALEATOR_OUT_OF_RANGE, "chain index out of range");
        const auto* b = static_cast<const unsigned char*>(buffer);
        require(size >= kCheckpointHeader && get<uint32_t>(b, 0) == kCheckpointMagic, ALEATOR_RUNTIME_ERROR,
                "gcmc checkpoint: not a chain checkpoint");
        require(get<uint32_t>(b, 4) == kCheckpointVersion, ALEATOR_RUNTIME_ERROR,
                "gcmc checkpoint: unsupported version");
        const int nkh = g->fw->charged ? g->fw->nk_half : 0;
        const int ns = g->args.ns;
        const int n_mol = get<int32_t>(b, 12);
        require(get<int32_t>(b, 8) == ns && get<int32_t>(b, 16) == nkh, ALEATOR_INVALID_ARGUMENT,
                "gcmc checkpoint: sites or k-set differ from these chains");
        require(n_mol >= 0 && n_mol <= g->args.max_mol, ALEATOR_INVALID_ARGUMENT,
                "gcmc checkpoint: more molecules than max_molecules");
        require(size == checkpoint_bytes(g, n_mol), ALEATOR_INVALID_ARGUMENT, "gcmc checkpoint: wrong size");
        auto all = download_states(g);
        ChainState& st = all[chain];
        st.n_mol = n_mol;
        st.moves_since_resync = get<int32_t>(b, 20);
        st.call = get<uint64_t>(b, 24);
        st.self_energy = get<double>(b, 32);
        st.exclusion_energy = get<double>(b, 40);
        st.sum_n = get<double>(b, 48);
        st.steps = get<uint64_t>(b, 56);
        for (int i = 0; i < 4; ++i) {
            st.attempted[i] = get<uint64_t>(b, 64 + 8 * i);
            st.accepted[i] = get<uint64_t>(b, 96 + 8 * i);
        }
        st.max_drift = get<double>(b, 128);
        st.resyncs = get<uint64_t>(b, 136);
        st.error = 0;
        hipStream_t s = g->fw->ctx->stream;
        size_t off = kCheckpointHeader;
        if (nkh) {
            ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_sre + (size_t)chain * nkh, b + off, nkh * sizeof(double),
                                             hipMemcpyHostToDevice, s));
            off += nkh * sizeof(double);
            ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_sim + (size_t)chain * nkh, b + off, nkh * sizeof(double),
                                             hipMemcpyHostToDevice, s));
            off += nkh * sizeof(double);
        }
        if (n_mol)
            ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_pos + (size_t)chain * g->args.max_mol * ns * 3, b + off,
                                             (size_t)n_mol * ns * 3 * sizeof(double), hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(g->d_state + chain, &st, sizeof(ChainState), hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
    });
}

aleator_status aleator_gpu_gcmc_destroy(aleator_gcmc* g) {
    return guarded([&] {
        if (!g) return;
        (void)hipStreamSynchronize(g->fw->ctx->stream);
        if (g->graph_exec) (void)hipGraphExecDestroy(g->graph_exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        for (void* p : {(void*)g->d_ljtab, (void*)g->d_ljgg, (void*)g->d_params, (void*)g->d_args, (void*)g->d_state,
                        (void*)g->d_pos, (void*)g->d_sre, (void*)g->d_sim, (void*)g->d_prop, (void*)g->d_bin_start,
                        (void*)g->d_bin_pos, (void*)g->d_bin_type, (void*)g->d_grid})
            if (p) (void)hipFree(p);
        Framework* fw = g->fw;
        delete g;
        release(fw);
    });
}

}  // extern "C"
