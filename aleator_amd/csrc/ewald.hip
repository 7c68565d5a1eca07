// ewald.hip -- EwaldIncrementalState S(k) kernels on gfx950.
//
// Reference: src/forcefield/electrostatics/ewald_incremental_state.cc
//   ctor / resync (:55-83, :214-247): S(k) = sum_i q_i exp(i k.r_i), O(N * Nk);
//   computeDelta  (:95-136): per k, dS from the moved sites and w (|S+dS|^2 - |S|^2).
//
// structure_factor: a workgroup owns 256 k-vectors of the half set and a chunk of 32
// atoms whose per-axis phase factors it stages in LDS; partial sums per chunk are
// reduced in a fixed order by a second kernel (deterministic, no atomics).
// reciprocal_delta: one wave per proposal (see recip.hpp), four proposals per
// workgroup; each proposal's result is written once, by lane 0.
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "recip.hpp"

namespace aleator_gpu {

namespace {

constexpr int kSfK = 256;     // k-vectors per workgroup
constexpr int kSfAtoms = 32;  // atoms per chunk

__global__ __launch_bounds__(kSfK) void sf_partial_kernel(const double* fx, const double* fy,
                                                          const double* fz, const double* q,
                                                          uint64_t n, const int4* kn, int nk_half,
                                                          int K, double2* partial) {
    extern __shared__ cplx sf_phase[];  // [3][2K+1][kSfAtoms]
    const int nkk = 2 * K + 1;
    const uint64_t a0 = (uint64_t)blockIdx.y * kSfAtoms;
    const int na = (int)((n - a0) < (uint64_t)kSfAtoms ? (n - a0) : (uint64_t)kSfAtoms);
    __shared__ double qs[kSfAtoms];
    for (int i = threadIdx.x; i < 3 * nkk * kSfAtoms; i += blockDim.x) {
        const int s = i % kSfAtoms;
        const int m = (i / kSfAtoms) % nkk;
        const int axis = i / (kSfAtoms * nkk);
        cplx v = {0.0, 0.0};
        if (s < na) {
            const double* f = axis == 0 ? fx : (axis == 1 ? fy : fz);
            double sv, cv;
            sincospi(2.0 * (double)(m - K) * f[a0 + s], &sv, &cv);
            v = {cv, sv};
        }
        sf_phase[i] = v;
    }
    if (threadIdx.x < kSfAtoms) qs[threadIdx.x] = threadIdx.x < na ? q[a0 + threadIdx.x] : 0.0;
    __syncthreads();
    const int t = blockIdx.x * kSfK + threadIdx.x;
    if (t >= nk_half) return;
    const int4 nv = kn[t];
    double re = 0.0, im = 0.0;
    for (int s = 0; s < na; ++s) {
        const cplx e0 = sf_phase[(0 * nkk + nv.x + K) * kSfAtoms + s];
        const cplx e1 = sf_phase[(1 * nkk + nv.y + K) * kSfAtoms + s];
        const cplx e2 = sf_phase[(2 * nkk + nv.z + K) * kSfAtoms + s];
        const cplx e = cmul(cmul(e0, e1), e2);
        re += qs[s] * e.re;
        im += qs[s] * e.im;
    }
    partial[(uint64_t)blockIdx.y * nk_half + t] = make_double2(re, im);
}

__global__ void sf_reduce_kernel(const double2* partial, int chunks, int nk_half, double* sre,
                                 double* sim) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk_half) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double2 p = partial[(uint64_t)c * nk_half + t];
        re += p.x;
        im += p.y;
    }
    sre[t] = re;
    sim[t] = im;
}

// sites: [n_prop][ns][4] = (f0, f1, f2, q); ns = n_rem + n_add
__global__ __launch_bounds__(256) void recip_delta_kernel(const double* sites, uint64_t n_prop,
                                                          int n_rem, int n_add, const int4* kn,
                                                          const double* kw, double* sre,
                                                          double* sim, int nk_half, int K,
                                                          double scale, double* out, int commit) {
    extern __shared__ cplx rd_phase[];
    const int ns = n_rem + n_add;
    const int nkk = 2 * K + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cplx* E = rd_phase + (size_t)wave * 3 * nkk * ns;
    const uint64_t p = (uint64_t)blockIdx.x * 4 + wave;
    const bool active = p < n_prop;
    const double* site4 = sites + (active ? p : 0) * ns * 4;
    if (active) stage_phases(lane, ns, K, site4, E);
    __syncthreads();
    if (!active) return;
    const double acc =
        wave_recip_delta(lane, ns, n_rem, K, E, site4, kn, kw, sre, sim, nk_half, commit != 0);
    if (lane == 0) out[p] = scale * acc;
}

// Batched proposals, RT per workgroup (RT = runs_molecules_per_block): the 256 / RT consecutive
// threads of a proposal build its phase rows (one per site and axis) and split the k-set runs,
// adding their partial sums by a fixed shuffle tree.  No commit: S(k) is only read.
template <int NS>
__global__ __launch_bounds__(256) void recip_delta_runs_kernel(const double* sites, uint64_t n_prop, int n_rem,
                                                               const int2* kseg, int n_seg, const double* kw,
                                                               const double* sre, const double* sim, int K,
                                                               double scale, double* out, int rt) {
    extern __shared__ cplx rr_phase[];  // [rt][NS][3][2K + 1]
    const int rg = 256 / rt;
    const int j = threadIdx.x / rg, g = threadIdx.x % rg;
    const int nkk = 2 * K + 1;
    const uint64_t p = (uint64_t)blockIdx.x * rt + j;
    const bool valid = p < n_prop;
    const double* site4 = sites + (valid ? p : 0) * NS * 4;
    cplx* E = rr_phase + (size_t)j * NS * 3 * nkk;
    if (valid && g < 3 * NS) phase_row(site4[4 * (g / 3) + g % 3], K, E + (size_t)g * nkk);
    __syncthreads();
    double acc = 0.0;
    if (valid) {
        double q[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) q[s] = s < n_rem ? -site4[4 * s + 3] : site4[4 * s + 3];
        acc = runs_recip_sum<NS>(E, q, K, kseg, n_seg, g, rg, kw, sre, sim);
    }
    for (int off = rg / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (valid && g == 0) out[p] = scale * acc;
}

template <int NS>
void launch_runs(Framework& fw, const double* sre, const double* sim, uint64_t n_prop, int n_rem,
                 const double* d_sites, double* d_out, int rt, double scale) {
    const size_t lds = (size_t)rt * NS * 3 * (2 * fw.kmax + 1) * sizeof(cplx);
    const unsigned blocks = (unsigned)((n_prop + rt - 1) / rt);
    hipLaunchKernelGGL((recip_delta_runs_kernel<NS>), dim3(blocks), dim3(256), lds, fw.ctx->stream, d_sites, n_prop,
                       n_rem, fw.d_kseg, fw.n_kseg, fw.d_kw, sre, sim, fw.kmax, scale, d_out, rt);
}

}  // namespace

void structure_factor(Framework& fw, uint64_t n, const double* d_fx, const double* d_fy,
                      const double* d_fz, const double* d_q, double* d_sre, double* d_sim) {
    Context& ctx = *fw.ctx;
    const int K = fw.kmax;
    const int chunks = (int)((n + kSfAtoms - 1) / kSfAtoms);
    if (chunks == 0) {
        ALEATOR_HIP_CHECK(hipMemsetAsync(d_sre, 0, fw.nk_half * sizeof(double), ctx.stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(d_sim, 0, fw.nk_half * sizeof(double), ctx.stream));
        return;
    }
    auto* partial = static_cast<double2*>(ctx.ensure_scratch((size_t)chunks * fw.nk_half * sizeof(double2)));
    const size_t lds = (size_t)3 * (2 * K + 1) * kSfAtoms * sizeof(cplx);
    dim3 grid((fw.nk_half + kSfK - 1) / kSfK, chunks);
    hipLaunchKernelGGL(sf_partial_kernel, grid, dim3(kSfK), lds, ctx.stream, d_fx, d_fy, d_fz, d_q, n,
                       reinterpret_cast<const int4*>(fw.d_kn), fw.nk_half, K, partial);
    hipLaunchKernelGGL(sf_reduce_kernel, dim3((fw.nk_half + 255) / 256), dim3(256), 0, ctx.stream,
                       partial, chunks, fw.nk_half, d_sre, d_sim);
    ALEATOR_HIP_CHECK(hipGetLastError());
}

void reciprocal_delta(Framework& fw, double* d_sre, double* d_sim, uint64_t n_prop, int n_rem,
                      int n_add, const double* d_sites, double* d_out, bool commit) {
    Context& ctx = *fw.ctx;
    const int ns = n_rem + n_add;
    if (ns > 2 * kMaxSites) fail(ALEATOR_INVALID_ARGUMENT, "ewald: too many sites in one proposal");
    if (commit && n_prop != 1) fail(ALEATOR_INVALID_ARGUMENT, "ewald: commit takes one proposal");
    const int K = fw.kmax;
    const size_t lds = (size_t)4 * 3 * (2 * K + 1) * (ns > 0 ? ns : 1) * sizeof(cplx);
    const double scale = 2.0 * kCoulomb * (2.0 * kPi / std::abs(fw.lat.volume));
    const unsigned blocks = (unsigned)((n_prop + 3) / 4);
    // many proposals, no commit: the run kernel (several proposals per workgroup)
    const int rt = runs_molecules_per_block(ns, K);
    ctx.time_begin();
    if (!commit && ns >= 1 && ns <= 6 && rt > 0 && n_prop >= 64) {
        switch (ns) {
            case 1: launch_runs<1>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
            case 2: launch_runs<2>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
            case 3: launch_runs<3>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
            case 4: launch_runs<4>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
            case 5: launch_runs<5>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
            default: launch_runs<6>(fw, d_sre, d_sim, n_prop, n_rem, d_sites, d_out, rt, scale); break;
        }
    } else {
        hipLaunchKernelGGL(recip_delta_kernel, dim3(blocks), dim3(256), lds, ctx.stream, d_sites, n_prop, n_rem,
                           n_add, reinterpret_cast<const int4*>(fw.d_kn), fw.d_kw, d_sre, d_sim, fw.nk_half, K,
                           scale, d_out, commit ? 1 : 0);
    }
    ALEATOR_HIP_CHECK(hipGetLastError());
    ctx.time_end();
}

}  // namespace aleator_gpu
