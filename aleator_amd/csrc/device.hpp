// device.hpp -- device-side building blocks shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aleator_gpu {

// ---------------------------------------------------------------- Philox4x32-10
// The bijection of counter_based_rng.cc:13-54 and the draw layout of
// Philox4x32Rng::nextU64/nextUniform (:64-79): key = (seed lo, hi), counter =
// (callIndex lo, hi, stream lo, hi), u64 = out0 << 32 | out1, uniform =
// (u64 >> 11) * 2^-53.  Integer-only, so bit-exact with the host.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
}

__host__ __device__ inline double philox_uniform(uint64_t seed, uint64_t stream, uint64_t call) {
    uint32_t c[4] = {(uint32_t)call, (uint32_t)(call >> 32), (uint32_t)stream,
                     (uint32_t)(stream >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t u = ((uint64_t)c[0] << 32) | (uint64_t)c[1];
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan for blockDim.x <= 1024 (multiple of 64).  `warp_tot`
// is LDS scratch of blockDim.x/64 + 1 ints; returns the exclusive prefix and
// writes the block total to *total.  Deterministic (no atomics).
__device__ inline int block_exclusive_scan(int v, int* warp_tot, int* total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int incl = wave_inclusive_scan(v);
    if (lane == 63) warp_tot[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < nwaves; ++w) {
            const int t = warp_tot[w];
            warp_tot[w] = run;
            run += t;
        }
        warp_tot[nwaves] = run;
    }
    __syncthreads();
    const int excl = warp_tot[wave] + incl - v;
    *total = warp_tot[nwaves];
    return excl;
}

// Block sum of doubles in a fixed order (deterministic); `lds` holds blockDim.x/64.
__device__ inline double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
    __syncthreads();
    return s;  // valid in thread 0
}

// ---------------------------------------------------------------- erfc(alpha r) / r
template <typename R>
__device__ inline R erfc_over_r(R alpha, R r);
template <>
__device__ inline double erfc_over_r<double>(double alpha, double r) {
    return erfc(alpha * r) / r;
}
// FP32 erfc(z)/r, z = alpha r, from r and 1/r, without branches: erfc(z) = t exp(w),
// w = P(t) - z^2, t = 1/(1 + z/2), with the degree-9 Chebyshev fit P of Press et al., Numerical
// Recipes in C (2nd ed.) section 6.2 ("erfcc", fractional error below 1.2e-7 for every z >= 0).
// exp(w) = 2^(w log2 e) goes through the hardware exp2; the roundings of z^2 (recovered exactly
// by an FMA), of P - z^2 (recovered exactly by a two-sum) and of the product by log2 e (an FMA
// split of log2 e into two floats) are carried as the first-order factor 1 + lo ln 2.  FP32
// evaluation against the FP64 erfc: relative error below 5e-7 on [0, 4].
__device__ inline float erfc_over_r_f32(float alpha, float r, float ir) {
#pragma clang fp contract(off)  // the two-sum below needs every product and sum rounded as written
    const float z = alpha * r;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, z, 1.0f));
    float p = 0.17087277f;
    p = fmaf(t, p, -0.82215223f);
    p = fmaf(t, p, 1.48851587f);
    p = fmaf(t, p, -1.13520398f);
    p = fmaf(t, p, 0.27886807f);
    p = fmaf(t, p, -0.18628806f);
    p = fmaf(t, p, 0.09678418f);
    p = fmaf(t, p, 0.37409196f);
    p = fmaf(t, p, 1.00002368f);
    p = fmaf(t, p, -1.26551223f);
    const float h = z * z;
    const float l = fmaf(z, z, -h);  // z^2 = h + l exactly
    const float s = p - h;           // two-sum: p - h = s + e exactly
    const float bb = s - p;
    const float e = (p - (s - bb)) + (-h - bb);
    const float lo = e - l;  // w = s + lo
    constexpr float kLog2e = 1.44269502162933349609375f;  // log2(e) rounded to float
    constexpr float kLog2eLo = 1.925963e-08f;             // log2(e) - kLog2e
    constexpr float kLn2 = 0.693147182464599609375f;
    const float y = s * kLog2e;
    const float ylo = fmaf(s, kLog2e, -y) + fmaf(s, kLog2eLo, lo * kLog2e);
    return t * (__builtin_amdgcn_exp2f(y) * fmaf(ylo, kLn2, 1.0f)) * ir;
}
template <>
__device__ inline float erfc_over_r<float>(float alpha, float r) {
    return erfc_over_r_f32(alpha, r, __builtin_amdgcn_rcpf(r));
}

// ---------------------------------------------------------------- complex helpers
struct cplx {
    double re, im;
};
__device__ inline cplx cmul(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

} // namespace aleator_gpu
