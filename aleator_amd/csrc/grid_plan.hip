// grid_plan.hip -- reusable, batched, graph-captured grid builds.
//
// A plan lays out every framework's constants in device memory once (LJ tables, reciprocal
// phase tables, descriptors, the prefix table that maps a workgroup to its framework), so
// executing it only launches kernels: the Coulomb pre-contractions of each charged framework
// and ONE tile-kernel launch over the tiles of all frameworks.  With use_graph the launch
// sequence is captured once into a hipGraph and replayed; the capture holds kernel nodes only
// (no host copies, no memsets).  With tile < 0 the plan times every tile shape that fits all
// frameworks with HIP events on the context stream and keeps the fastest (the autotuner).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "aleator_gpu.h"
#include "grid.hpp"
#include "internal.hpp"

struct aleator_grid_plan {
    aleator_gpu::Context* ctx = nullptr;
    std::vector<aleator_gpu::Framework*> fws;  // one reference held on each framework
    std::vector<aleator_gpu::GridJob> jobs;
    int shape = 0;
    int chunk = 128;  // atoms per list-build pass
    int total_blocks = 0;
    char* d_base = nullptr;
    aleator_gpu::GridArgs* d_descs = nullptr;
    int* d_prefix = nullptr;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<float> tuning_ms;
};

namespace aleator_gpu {
namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

void plan_bind_shape(aleator_grid_plan* p, int shape, int chunk) {
    std::vector<GridArgs> descs;
    std::vector<int> prefix(1, 0);
    for (auto& j : p->jobs) {
        // frameworks that may list several images of an atom keep 128 atoms per pass
        j.set_shape(shape, j.chunk_fits(shape, chunk) ? chunk : kChunkSmall);
        descs.push_back(j.args);
        prefix.push_back(prefix.back() + j.blocks);
    }
    p->shape = shape;
    p->chunk = chunk;
    p->total_blocks = prefix.back();
    Context& ctx = *p->ctx;
    ALEATOR_HIP_CHECK(hipMemcpyAsync(p->d_descs, descs.data(), descs.size() * sizeof(GridArgs), hipMemcpyHostToDevice,
                                     ctx.stream));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(p->d_prefix, prefix.data(), prefix.size() * sizeof(int), hipMemcpyHostToDevice,
                                     ctx.stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx.stream));
}

void plan_launch(aleator_grid_plan* p) {
    hipStream_t s = p->ctx->stream;
    for (const auto& j : p->jobs) j.launch_pre(s);
    const GridJob& j0 = p->jobs.front();
    int ljk = 0;
    for (const auto& j : p->jobs) ljk = std::max(ljk, lj_kind(j.args.trunc));
    launch_tiles(j0.precision, p->shape, j0.ng, j0.deg, ljk, p->d_descs, p->d_prefix, (int)p->jobs.size(),
                 p->total_blocks, s);
}

void plan_free(aleator_grid_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_base) (void)hipFree(p->d_base);
}

}  // namespace
}  // namespace aleator_gpu

using namespace aleator_gpu;

extern "C" {

aleator_status aleator_gpu_grid_plan_create(aleator_ctx* ctx_in, int32_t n_frameworks, aleator_framework* const* fws,
                                            const aleator_lattice* grid_lattices, const uint32_t* guest_species,
                                            int32_t n_guest, int32_t coulomb, const int32_t* n_points,
                                            int32_t precision, double cap, double* const* out_device, int32_t tile,
                                            int32_t use_graph, aleator_grid_plan** out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Context*>(ctx_in);
        require(ctx && fws && n_points && out_device && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(n_frameworks >= 1, ALEATOR_INVALID_ARGUMENT, "grid plan: at least one framework");
        require(n_guest >= 0 && (n_guest == 0 || guest_species), ALEATOR_INVALID_ARGUMENT, "guest species list");
        require(tile < kNumTileShapes, ALEATOR_INVALID_ARGUMENT, "grid plan: unknown tile shape");
        auto p = std::make_unique<aleator_grid_plan>();
        p->ctx = ctx;
        p->jobs.resize(n_frameworks);
        size_t total = 0;
        std::vector<size_t> offs(n_frameworks);
        for (int f = 0; f < n_frameworks; ++f) {
            auto* fw = reinterpret_cast<Framework*>(fws[f]);
            require(fw != nullptr && out_device[f] != nullptr, ALEATOR_INVALID_ARGUMENT, "null framework or output");
            require(fw->ctx == ctx, ALEATOR_INVALID_ARGUMENT, "grid plan: framework of another context");
            GridRequest req;
            req.grid_m = grid_lattices ? &grid_lattices[f].m[0][0] : nullptr;
            req.guests.assign(guest_species, guest_species + n_guest);
            for (int a = 0; a < 3; ++a) req.n[a] = n_points[3 * f + a];
            req.coulomb = coulomb != 0;
            req.precision = precision;
            req.cap = cap;
            req.out_device = out_device[f];
            p->jobs[f].prepare(*fw, req);
            offs[f] = total;
            total += align256(p->jobs[f].bytes);
        }
        // one erf polynomial degree for the whole launch: the largest any framework needs
        int deg = 0;
        for (const auto& j : p->jobs) deg = std::max(deg, j.deg);
        for (auto& j : p->jobs)
            if (j.deg != deg) j.fit_poly(deg);
        const size_t off_desc = total;
        total += align256((size_t)n_frameworks * sizeof(GridArgs));
        const size_t off_prefix = total;
        total += align256((size_t)(n_frameworks + 1) * sizeof(int));
        ALEATOR_HIP_CHECK(hipMalloc(&p->d_base, total));
        try {
            for (int f = 0; f < n_frameworks; ++f) {
                p->jobs[f].bind(p->d_base + offs[f]);
                p->jobs[f].upload(ctx->stream);
            }
            p->d_descs = reinterpret_cast<GridArgs*>(p->d_base + off_desc);
            p->d_prefix = reinterpret_cast<int*>(p->d_base + off_prefix);
            // candidate shapes: those whose tile radius fits every framework
            std::vector<int> shapes;
            for (int s = 0; s < kNumTileShapes; ++s) {
                bool ok = true;
                for (const auto& j : p->jobs) ok = ok && j.shape_fits(s);
                if (ok) shapes.push_back(s);
            }
            require(!shapes.empty(), ALEATOR_INVALID_ARGUMENT,
                    "grid: cutoff plus tile radius exceeds the perpendicular width");
            // atoms per list-build pass with a shape: 128, and 256, 512 and 1024 when some framework
            // keeps one image per atom at that shape (the frameworks that do not keep 128)
            auto chunks_for = [&](int s) {
                std::vector<int> c{kChunkSmall};
                for (int k : {256, 512, 1024}) {
                    bool any = false;
                    for (const auto& j : p->jobs) any = any || j.chunk_fits(s, k);
                    if (any) c.push_back(k);
                }
                return c;
            };
            int chosen = shapes.front(), chosen_chunk = chunks_for(shapes.front()).back();
            if (tile >= 0) {
                require(std::find(shapes.begin(), shapes.end(), tile) != shapes.end(), ALEATOR_INVALID_ARGUMENT,
                        "grid plan: tile shape too large for a framework");
                chosen = tile;
                chosen_chunk = chunks_for(tile).back();
            } else {
                // autotune over (tile shape, atoms per pass): one warm launch, then the median of
                // three timed launches each; tuning_ms keeps each shape's best
                hipEvent_t e0, e1;
                ALEATOR_HIP_CHECK(hipEventCreate(&e0));
                ALEATOR_HIP_CHECK(hipEventCreate(&e1));
                float best = 1e30f;
                p->tuning_ms.assign(kNumTileShapes, -1.0f);
                for (int s : shapes) {
                    for (int c : chunks_for(s)) {
                        plan_bind_shape(p.get(), s, c);
                        plan_launch(p.get());
                        float t[3];
                        for (float& v : t) {
                            ALEATOR_HIP_CHECK(hipEventRecord(e0, ctx->stream));
                            plan_launch(p.get());
                            ALEATOR_HIP_CHECK(hipEventRecord(e1, ctx->stream));
                            ALEATOR_HIP_CHECK(hipEventSynchronize(e1));
                            ALEATOR_HIP_CHECK(hipEventElapsedTime(&v, e0, e1));
                        }
                        std::sort(t, t + 3);
                        if (p->tuning_ms[s] < 0.0f || t[1] < p->tuning_ms[s]) p->tuning_ms[s] = t[1];
                        if (t[1] < best) {
                            best = t[1];
                            chosen = s;
                            chosen_chunk = c;
                        }
                    }
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                ctx->tile_shape = chosen;
                ctx->tile_chunk = chosen_chunk;
            }
            plan_bind_shape(p.get(), chosen, chosen_chunk);
            if (use_graph) {
                p->use_graph = true;
                ALEATOR_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                try {
                    plan_launch(p.get());
                } catch (...) {
                    hipGraph_t g;
                    (void)hipStreamEndCapture(ctx->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                ALEATOR_HIP_CHECK(hipStreamEndCapture(ctx->stream, &p->graph));
                ALEATOR_HIP_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
            }
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            plan_free(p.get());
            throw;
        }
        for (int f = 0; f < n_frameworks; ++f) {
            auto* fw = reinterpret_cast<Framework*>(fws[f]);
            retain(fw);
            p->fws.push_back(fw);
        }
        *out = p.release();
    });
}

aleator_status aleator_gpu_grid_plan_execute(aleator_grid_plan* p) {
    return guarded([&] {
        require(p != nullptr, ALEATOR_INVALID_ARGUMENT, "null plan");
        Context& ctx = *p->ctx;
        ctx.time_begin();
        if (p->use_graph)
            ALEATOR_HIP_CHECK(hipGraphLaunch(p->exec, ctx.stream));
        else
            plan_launch(p);
        ctx.time_end();
        ALEATOR_HIP_CHECK(hipGetLastError());
    });
}

aleator_status aleator_gpu_grid_plan_tile(const aleator_grid_plan* p, int32_t* tile, float* tuning_ms) {
    return guarded([&] {
        require(p && tile, ALEATOR_INVALID_ARGUMENT, "null argument");
        *tile = p->shape;
        if (tuning_ms)
            for (int s = 0; s < kNumTileShapes; ++s)
                tuning_ms[s] = p->tuning_ms.empty() ? -1.0f : p->tuning_ms[s];
    });
}

aleator_status aleator_gpu_grid_plan_atoms_per_pass(const aleator_grid_plan* p, int32_t* atoms) {
    return guarded([&] {
        require(p && atoms, ALEATOR_INVALID_ARGUMENT, "null argument");
        *atoms = p->chunk;
    });
}

aleator_status aleator_gpu_grid_plan_erf_degree(const aleator_grid_plan* p, int32_t* degree, int32_t* shell_degree) {
    return guarded([&] {
        require(p && degree && shell_degree, ALEATOR_INVALID_ARGUMENT, "null argument");
        const int deg = p->jobs.empty() ? 0 : p->jobs[0].deg;
        bool split = false;
        for (const auto& j : p->jobs) split = split || j.args.far_ok;
        *degree = deg;
        *shell_degree = split ? deg - kFarDrop : deg;
    });
}

aleator_status aleator_gpu_grid_plan_destroy(aleator_grid_plan* p) {
    return guarded([&] {
        if (!p) return;
        (void)hipStreamSynchronize(p->ctx->stream);
        plan_free(p);
        const std::vector<Framework*> fws = std::move(p->fws);
        delete p;
        for (Framework* fw : fws) release(fw);
    });
}

}  // extern "C"
