// mlp.hip -- the learned insertion proposal: a Fourier-feature MLP trained and evaluated on the
// GPU with hand-written BF16 MFMA kernels (FP32 accumulation).  No reference counterpart.
//
// Network: 64 Fourier features of the fractional position -> 4 hidden layers of width 256
// with SiLU -> 1 output (214,273 parameters).  Features: sin and cos of 2*pi*(k . f) for 32
// integer frequency vectors k, so the network is periodic in the cell: 2^j e_a (j = 0..5,
// a = 0..2; 18 vectors), 2^j (e_a + e_b) (j = 0..3, three axis pairs; 12), 2^j (1,1,1)
// (j = 0..1; 2).
//
// A training iteration is two kernels:
//  1. mlp_rows_kernel, one workgroup of 8 waves per block of 32 rows: draws the rows, computes
//     their features, runs the 4 hidden layers forward, the head and the loss, and propagates
//     dL/dz back to the first hidden layer, with the block's activations in LDS between layers.
//     Wave w owns output columns [32 w, 32 w + 32) of every layer (2 x 2 tiles of
//     v_mfma_f32_16x16x32_bf16) and keeps their pre-activations in registers from the forward
//     to the backward pass.  Weights are the B operands, read from global memory in MFMA
//     fragment order (one 16-byte load per lane and fragment).  The block's features, hidden
//     activations and dL/dz go to global memory block-transposed (the block's 32 rows of a
//     column contiguous), with per-block sums for the bias and output-layer gradients.
//  2. mlp_grad_adam_kernel: one workgroup per 32 x 32 tile of dW_l = H_{l-1}^T dZ_l over the
//     whole batch (a 32-row block is one MFMA K step whose operands are 16-byte loads of the
//     block-transposed arrays; the 8 waves take contiguous eighths of the blocks and their
//     partial tiles are summed in LDS), followed in the same workgroup by Adam on the tile's
//     FP32 master weights and the refresh of its bf16 fragment copies; further workgroups sum
//     the per-block bias and output-layer gradients and update those parameters.
// Every sum runs in a fixed order, so results are reproducible bit for bit.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (64 lanes): A lane l holds A[l % 16][8 (l / 16) ..
// + 7], B lane l holds B[8 (l / 16) .. + 7][l % 16], D lane l holds D[4 (l / 16) .. + 3][l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace aleator_gpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// bf16 of two floats, round to nearest even (v_cvt_pk_bf16_f32), packed low | high << 16
__device__ inline uint32_t f2bf2(float lo, float hi) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{lo, hi}), bf16x2));
}

// sigmoid(z) = 1 / (1 + exp(-z)) from the hardware exponential and reciprocal (about 1 ulp each);
// silu(z) = z sigmoid(z) and silu'(z) = sigmoid(z) (1 + z (1 - sigmoid(z)))
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

constexpr int kRB = kMlpBlockRows;
constexpr int kRowThreads = 512;             // 8 waves
constexpr int kActStride = kMlpHidden + 8;   // LDS row stride (bf16) of an activation block

enum Mode : int { kPredict = 0, kGrad = 1, kTrain = 2 };

// Position of W_l[i][j] (i < fan_in, j < 256) in the forward copy, the B operand of H W_l:
// fragment (kt = i / 32, nt = j / 16) of 64 lanes x 8 values, lane (i % 32) / 8 * 16 + j % 16,
// value i % 8; fragments nt-major.
__host__ __device__ inline int frag_fwd(int fan_in, int i, int j) {
    const int kt = i >> 5, nt = j >> 4, lane = ((i & 31) >> 3) * 16 + (j & 15);
    return ((nt * (fan_in >> 5) + kt) * 64 + lane) * 8 + (i & 7);
}
// Position of W_l[i][j] in the backward copy, the B operand of dZ W_l^T (K index j, N index i).
__host__ __device__ inline int frag_bwd(int i, int j) {
    const int kt = j >> 5, nt = i >> 4, lane = ((j & 31) >> 3) * 16 + (i & 15);
    return ((nt * 8 + kt) * 64 + lane) * 8 + (j & 7);
}
__host__ __device__ inline int wf_offset(int l) {
    return l == 0 ? 0 : kMlpIn * kMlpHidden + (l - 1) * kMlpHidden * kMlpHidden;
}
__host__ __device__ inline int wb_offset(int l) { return (l - 1) * kMlpHidden * kMlpHidden; }  // l = 1..3
// block-transposed layer inputs of nb blocks: X (64 columns), then H1, H2, H3 (256 each)
__host__ __device__ inline size_t ht_offset(int l, int nb) {
    return l == 0 ? 0 : (size_t)nb * kMlpIn * kRB + (size_t)(l - 1) * nb * kMlpHidden * kRB;
}
constexpr size_t kHtPerBlock = (size_t)(kMlpIn + 3 * kMlpHidden) * kRB;
constexpr size_t kDztPerBlock = (size_t)4 * kMlpHidden * kRB;

__device__ inline bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

struct RowArgs {
    const double* frac;
    const double* target;
    uint64_t n_samples;     // kTrain: rows are drawn from frac[0 .. n_samples)
    uint64_t seed, stream;  // kTrain: row r of iteration c is floor(u n) with u = Philox(seed, stream) at c rows + r
    int rows;
    int mode;
    const int8_t* kvec;
    const uint16_t* wf;
    const uint16_t* wb;
    const float* w;  // fp32 master parameters (biases, w5, b5)
    uint64_t* counters;
    float* y_out;     // kPredict
    uint16_t* ht;     // block-transposed X, H1..H3
    uint16_t* dzt;    // block-transposed dZ1..dZ4
    float* blockpart; // [block][kMlpBlockPart]
};

// acc[i][j] = act[16 i .. + 15][0 .. 32 KT) x W[0 .. 32 KT)[columns of n tile 2 wave + j], the
// W columns from fragments frag (KT K steps per n tile).  All B fragments are loaded first.
template <int KT>
__device__ inline void block_gemm(const uint16_t (*act)[kActStride], const uint16_t* frag, int wave, int lane,
                                  f32x4 (&acc)[2][2]) {
    bf16x8 b[KT][2];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int j = 0; j < 2; ++j) b[kt][j] = ld8(frag + (((2 * wave + j) * KT + kt) * 64 + lane) * 8);
    // every load in flight before the first MFMA (the scheduler would otherwise interleave them
    // and wait on one round trip per K step)
    __builtin_amdgcn_sched_barrier(0);
    const int ar = lane & 15, ak = 8 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
        bf16x8 a[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = ld8(&act[16 * i + ar][32 * kt + ak]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[kt][j], acc[i][j], 0, 0, 0);
    }
}


// 4 consecutive rows of one column (two packed bf16 pairs) into the LDS block and, when dst is
// set, into a block-transposed array with nontemporal stores (read by the next kernel, not by
// this workgroup)
__device__ inline void put4(uint16_t (*out)[kActStride], int row, int col, uint32_t p01, uint32_t p23, uint16_t* dst) {
    out[row][col] = (uint16_t)p01;
    out[row + 1][col] = (uint16_t)(p01 >> 16);
    out[row + 2][col] = (uint16_t)p23;
    out[row + 3][col] = (uint16_t)(p23 >> 16);
    if (dst) {
        __builtin_nontemporal_store(p01, reinterpret_cast<uint32_t*>(dst));
        __builtin_nontemporal_store(p23, reinterpret_cast<uint32_t*>(dst) + 1);
    }
}

// Forward epilogue: Z = acc + bias, H = silu(Z) in bf16 to the LDS block `out` and, when ht is
// set, to the block-transposed array; when training, gz keeps silu'(Z) for the backward pass.
template <bool kKeep>
__device__ inline void forward_out(const f32x4 (&acc)[2][2], f32x4 (&gz)[2][2], const float* bias,
                                   uint16_t (*out)[kActStride], uint16_t* ht, int wave, int lane) {
    const int rq = 4 * (lane >> 4);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = 32 * wave + 16 * j + (lane & 15);
        const float b = bias[col];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            float h[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = acc[i][j][r] + b;
                const float sg = sigmoid(z);
                h[r] = z * sg;
                if (kKeep) gz[i][j][r] = sg * (1.0f + z * (1.0f - sg));
            }
            put4(out, 16 * i + rq, col, f2bf2(h[0], h[1]), f2bf2(h[2], h[3]), ht ? ht + col * kRB + 16 * i + rq : nullptr);
        }
    }
}

// Backward epilogue: dZ = acc * silu'(Z) in bf16 to the LDS block `out` and the block-transposed
// array dzt.
__device__ inline void backward_out(const f32x4 (&acc)[2][2], const f32x4 (&gz)[2][2], uint16_t (*out)[kActStride],
                                    uint16_t* dzt, int wave, int lane) {
    const int rq = 4 * (lane >> 4);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int col = 32 * wave + 16 * j + (lane & 15);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const f32x4 v = acc[i][j] * gz[i][j];
            put4(out, 16 * i + rq, col, f2bf2(v[0], v[1]), f2bf2(v[2], v[3]), dzt + col * kRB + 16 * i + rq);
        }
    }
}

// column sums of a 32-row LDS block of dL/dz (bias gradients of the block): four interleaved
// partial sums, combined in a fixed order
__device__ inline void column_sums(const uint16_t (*blk)[kActStride], float* out) {
    const int c = threadIdx.x;
    if (c >= kMlpHidden) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < kRB; ++r) s[r & 3] += bf2f(blk[r][c]);
    out[c] = (s[0] + s[1]) + (s[2] + s[3]);
}

// One block of 32 rows: features, forward, head; when training also the loss, dL/dz of every
// hidden layer and the block's bias and output-layer gradient sums.
template <bool kTrainPath>
__global__ __launch_bounds__(kRowThreads) void mlp_rows_kernel(const RowArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t act[2][kRB][kActStride];
    __shared__ double s_pos[kRB][3];
    __shared__ float s_target[kRB];
    __shared__ float s_dy[kRB];
    __shared__ float s_err2[kRB];
    __shared__ float s_bias[4][kMlpHidden];
    __shared__ float s_w5[kMlpHidden + 1];  // w5, then b5
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int blk = blockIdx.x, row0 = blk * kRB, nb = gridDim.x;

    // Warm this XCD's L2 with every weight fragment the layers will read, so that each layer's
    // B operands hit in L2 instead of going to memory one layer at a time.  Workgroups are
    // dispatched to the 8 XCDs round robin; the blocks of one XCD split its 128-byte lines.
    uint32_t sink = 0;
    {
        constexpr int kFwdLines = kMlpFwdFrag * 2 / 128;
        constexpr int kLines = kFwdLines + (kTrainPath ? kMlpBwdFrag * 2 / 128 : 0);
        const int xcd = blk & 7, rank = blk >> 3, per_xcd = (nb - xcd + 7) >> 3;
        const int per = (kLines + per_xcd - 1) / per_xcd;
        const int end = min(kLines, (rank + 1) * per);
        for (int line = rank * per + tid; line < end; line += kRowThreads) {
            const uint16_t* p = line < kFwdLines ? a.wf + line * 64 : a.wb + (line - kFwdLines) * 64;
            sink ^= *reinterpret_cast<const uint32_t*>(p);
        }
    }
    if (tid < kMlpHidden) {
#pragma unroll
        for (int l = 0; l < 4; ++l) s_bias[l][tid] = a.w[mlp_offset(2 * l + 1) + tid];
        s_w5[tid] = a.w[mlp_offset(8) + tid];
    } else if (tid == kMlpHidden) {
        s_w5[kMlpHidden] = a.w[mlp_offset(9)];
    }

    if (tid < kRB) {
        const int g = row0 + tid;
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        float t = 0.f;
        if (g < a.rows) {
            uint64_t p = (uint64_t)g;
            if (a.mode == kTrain) {
                const double u = philox_uniform(a.seed, a.stream, a.counters[0] * (uint64_t)a.rows + (uint64_t)g);
                const uint64_t v = (uint64_t)(u * (double)a.n_samples);
                p = v < a.n_samples ? v : a.n_samples - 1;
            }
            f0 = a.frac[3 * p];
            f1 = a.frac[3 * p + 1];
            f2 = a.frac[3 * p + 2];
            if (kTrainPath) t = (float)a.target[p];
        }
        s_pos[tid][0] = f0;
        s_pos[tid][1] = f1;
        s_pos[tid][2] = f2;
        s_target[tid] = t;
    }
    // one more Adam step (read by the update kernel; nothing in this launch reads it)
    if (kTrainPath && a.mode == kTrain && blk == 0 && tid == 0) a.counters[1] += 1;
    __syncthreads();
    // features; rows past the end read zero
    for (int e = tid; e < kRB * kMlpFreqs; e += kRowThreads) {
        const int r = e / kMlpFreqs, c = e % kMlpFreqs;
        uint16_t hs = 0, hc = 0;
        if (row0 + r < a.rows) {
            const double ph = a.kvec[3 * c] * s_pos[r][0] + a.kvec[3 * c + 1] * s_pos[r][1] + a.kvec[3 * c + 2] * s_pos[r][2];
            double s, co;
            sincospi(2.0 * ph, &s, &co);
            hs = f2bf((float)s);
            hc = f2bf((float)co);
        }
        act[0][r][2 * c] = hs;
        act[0][r][2 * c + 1] = hc;
    }
    __syncthreads();
    if (kTrainPath && tid < kMlpIn * 4) {  // X, block-transposed: 8 rows of one column per thread
        const int col = tid >> 2, q = (tid & 3) * 8;
        uint16_t v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = act[0][q + t][col];
        uint4 u;
        u.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
        u.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
        u.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
        u.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
        *reinterpret_cast<uint4*>(a.ht + ((size_t)blk * kMlpIn + col) * kRB + q) = u;
    }
    auto ht_block = [&](int l) -> uint16_t* {
        return kTrainPath ? a.ht + ht_offset(l, nb) + (size_t)blk * kMlpHidden * kRB : nullptr;
    };
    f32x4 acc[2][2];
    f32x4 z[4][2][2];  // silu'(Z) of each hidden layer (training)
    block_gemm<kMlpIn / 32>(act[0], a.wf + wf_offset(0), wave, lane, acc);
    forward_out<kTrainPath>(acc, z[0], s_bias[0], act[1], ht_block(1), wave, lane);
    __syncthreads();
    block_gemm<kMlpHidden / 32>(act[1], a.wf + wf_offset(1), wave, lane, acc);
    forward_out<kTrainPath>(acc, z[1], s_bias[1], act[0], ht_block(2), wave, lane);
    __syncthreads();
    block_gemm<kMlpHidden / 32>(act[0], a.wf + wf_offset(2), wave, lane, acc);
    forward_out<kTrainPath>(acc, z[2], s_bias[2], act[1], ht_block(3), wave, lane);
    __syncthreads();
    block_gemm<kMlpHidden / 32>(act[1], a.wf + wf_offset(3), wave, lane, acc);
    forward_out<kTrainPath>(acc, z[3], s_bias[3], act[0], nullptr, wave, lane);
    __syncthreads();

    // head: y = H4 . w5 + b5, one wave per 4 rows (lane-strided partial sums, then a butterfly)
    const float* w5 = s_w5;
    for (int q = 0; q < kRB / 8; ++q) {
        const int r = wave * (kRB / 8) + q;
        float s = 0.f;
        for (int k = lane; k < kMlpHidden; k += 64) s += bf2f(act[0][r][k]) * w5[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        const float y = s + s_w5[kMlpHidden];
        if (lane == 0) {
            const bool valid = row0 + r < a.rows;
            if (!kTrainPath) {
                if (valid) a.y_out[row0 + r] = y;
            } else {
                const float d = valid ? y - s_target[r] : 0.f;
                s_dy[r] = 2.0f * d / (float)a.rows;
                s_err2[r] = d * d;
            }
        }
    }
    if constexpr (!kTrainPath) {
        asm volatile("" ::"v"(sink));  // consumes the warm-up loads
    } else {
        __syncthreads();
        float* part = a.blockpart + (size_t)blk * kMlpBlockPart;
        // output-layer gradient sums of the block: dw5, db5, squared error
        if (tid < kMlpHidden) {
            float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < kRB; ++r) s[r & 3] += bf2f(act[0][r][tid]) * s_dy[r];
            part[4 * kMlpHidden + tid] = (s[0] + s[1]) + (s[2] + s[3]);
        } else if (tid == kMlpHidden) {
            float s = 0.f;
            for (int r = 0; r < kRB; ++r) s += s_dy[r];
            part[5 * kMlpHidden] = s;
        } else if (tid == kMlpHidden + 1) {
            float s = 0.f;
            for (int r = 0; r < kRB; ++r) s += s_err2[r];
            part[5 * kMlpHidden + 1] = s;
        }
        // dZ4 = dy w5 silu'(Z4)
        {
            uint16_t* dzt4 = a.dzt + ((size_t)3 * nb + blk) * kMlpHidden * kRB;
            const int rq = 4 * (lane >> 4);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = 32 * wave + 16 * j + (lane & 15);
                const float wc = w5[col];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = s_dy[16 * i + rq + r] * wc * z[3][i][j][r];
                    put4(act[1], 16 * i + rq, col, f2bf2(v[0], v[1]), f2bf2(v[2], v[3]), dzt4 + col * kRB + 16 * i + rq);
                }
            }
        }
        __syncthreads();
        // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1}), l = 3, 2, 1 (0-based), with the bias sums of dZ_l
        column_sums(act[1], part + 3 * kMlpHidden);
        block_gemm<kMlpHidden / 32>(act[1], a.wb + wb_offset(3), wave, lane, acc);
        backward_out(acc, z[2], act[0], a.dzt + ((size_t)2 * nb + blk) * kMlpHidden * kRB, wave, lane);
        __syncthreads();
        column_sums(act[0], part + 2 * kMlpHidden);
        block_gemm<kMlpHidden / 32>(act[0], a.wb + wb_offset(2), wave, lane, acc);
        backward_out(acc, z[1], act[1], a.dzt + ((size_t)1 * nb + blk) * kMlpHidden * kRB, wave, lane);
        __syncthreads();
        column_sums(act[1], part + 1 * kMlpHidden);
        block_gemm<kMlpHidden / 32>(act[1], a.wb + wb_offset(1), wave, lane, acc);
        backward_out(acc, z[0], act[0], a.dzt + (size_t)blk * kMlpHidden * kRB, wave, lane);
        __syncthreads();
        column_sums(act[0], part);
        asm volatile("" ::"v"(sink));  // consumes the warm-up loads
    }
}

struct GradArgs {
    const uint16_t* ht;       // block-transposed layer inputs (X, H1..H3)
    const uint16_t* dzt;      // block-transposed dL/dz of the 4 hidden layers
    const float* blockpart;   // [block][kMlpBlockPart]
    int nb;                   // row blocks
    int rows;
    float* w;
    float* grad;
    float* m1;
    float* m2;
    uint16_t* wf;
    uint16_t* wb;
    float lr;
    int adam;                 // 0: gradient and loss only
    uint64_t* counters;
    float* loss_hist;
};

constexpr int kGradThreads = 512;                          // 8 waves
constexpr int kGradWaves = kGradThreads / 64;
constexpr int kGradTiles = (kMlpIn / 32) * 8 + 3 * 8 * 8;  // 32 x 32 tiles: 16 of W1, then 64 each of W2..W4
constexpr int kGradChunk = 8;                              // row blocks whose operands are loaded at once
constexpr int kColsumCols = 16;                            // columns per column-sum workgroup
constexpr int kColsumGroups = kGradThreads / kColsumCols;  // block groups per column
constexpr int kColsumBlocks = (kMlpBlockPart + kColsumCols - 1) / kColsumCols;

// region of parameter i: 0..9 = W1 b1 W2 b2 W3 b3 W4 b4 w5 b5
__device__ inline int param_region(int i) {
    int g = 0;
#pragma unroll
    for (int k = 1; k < 10; ++k) g += i >= mlp_offset(k);
    return g;
}

// bf16 fragment copies of weight i (region g, even and below 8)
__device__ inline void store_bf16_copies(uint16_t* wf, uint16_t* wb, int g, int i, uint16_t v) {
    const int l = g >> 1, e = i - mlp_offset(g);
    const int ii = e / kMlpHidden, jj = e % kMlpHidden;
    wf[wf_offset(l) + frag_fwd(l == 0 ? kMlpIn : kMlpHidden, ii, jj)] = v;
    if (l > 0) wb[wb_offset(l) + frag_bwd(ii, jj)] = v;
}

// Adam on parameter i with gradient g (step t, bias corrections 1 - b^t); returns the new value
__device__ inline float adam_update(const GradArgs& a, int i, float g, float t) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    const float m = b1 * a.m1[i] + (1.f - b1) * g;
    const float v = b2 * a.m2[i] + (1.f - b2) * g * g;
    a.m1[i] = m;
    a.m2[i] = v;
    const float nw = a.w[i] - a.lr * (m / c1) / (sqrtf(v / c2) + eps);
    a.w[i] = nw;
    return nw;
}

// Gradients and, with adam, the Adam step, in one launch.  Workgroups [0, kGradTiles): one 32 x 32
// tile of dW_l = H_{l-1}^T dZ_l over every row: wave w sums an eighth of the row blocks (one MFMA
// K step per 32-row block, operands 16-byte loads of the block-transposed arrays, kGradChunk
// blocks' loads in flight at once), the eight parts are added in wave order, and the tile's
// weights are updated in place, with both bf16 fragment copies written as whole 16-byte
// entries.  Workgroups after them: the per-block records (bias and output-layer gradients,
// squared error) summed over the blocks, 16 columns per workgroup (thread (g, c) takes blocks
// g, g + 32, ... of column c), and the update of those parameters; the workgroup holding the
// loss also advances the iteration counter when training (nothing else in this launch reads it).
// Every sum runs in a fixed order.
__global__ __launch_bounds__(kGradThreads) void mlp_grad_adam_kernel(const GradArgs a) {
    __shared__ float s_acc[kGradWaves][32][33];
    __shared__ float s_w[32][33];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const float t_adam = (float)a.counters[1];  // advanced by the row kernel of this iteration
    if ((int)blockIdx.x >= kGradTiles) {
        float(*s_part)[kColsumCols] = reinterpret_cast<float(*)[kColsumCols]>(&s_acc[0][0][0]);
        const int grp = tid / kColsumCols, c = tid % kColsumCols;
        const int col = kColsumCols * (blockIdx.x - kGradTiles) + c;
        float s = 0.f;
        if (col < kMlpBlockPart) {
#pragma unroll 4
            for (int b = grp; b < a.nb; b += kColsumGroups) s += a.blockpart[(size_t)b * kMlpBlockPart + col];
        }
        s_part[grp][c] = s;
        __syncthreads();
        if (grp != 0 || col >= kMlpBlockPart) return;
        float g = 0.f;
        for (int k = 0; k < kColsumGroups; ++k) g += s_part[k][c];
        if (col == 5 * kMlpHidden + 1) {
            const uint64_t it = a.counters[0];
            a.loss_hist[it % kLossChunk] = g / (float)a.rows;
            if (a.adam) a.counters[0] = it + 1;
            return;
        }
        const int i = col < 4 * kMlpHidden ? mlp_offset(2 * (col / kMlpHidden) + 1) + col % kMlpHidden
                                           : mlp_offset(8) + (col - 4 * kMlpHidden);  // w5, then b5
        a.grad[i] = g;
        if (a.adam) adam_update(a, i, g, t_adam);
        return;
    }
    // tile -> layer l, 32-row tile mt of fan_in, 32-column tile nt of the 256 outputs.  Workgroup
    // b runs on XCD x = b mod 8 (round-robin dispatch); its 26 tiles are layer 0's column tile
    // nt = x (both row tiles), then per hidden layer the 2 x 4 block of row tiles 2 (x / 2) + {0, 1}
    // and column tiles 4 (x mod 2) + {0..3}, so an XCD reads a quarter of a layer's inputs and
    // half of its dL/dz through its L2 instead of nearly all of both
    int l, mt, nt;
    {
        const int x = blockIdx.x % 8, slot = blockIdx.x / 8;
        if (slot < 2) {
            l = 0;
            mt = slot;
            nt = x;
        } else {
            const int r = slot - 2;
            l = 1 + r / 8;
            mt = 2 * (x / 2) + (r % 8) / 4;
            nt = 4 * (x % 2) + r % 4;
        }
    }
    const int cols = l == 0 ? kMlpIn : kMlpHidden;
    const uint16_t* A = a.ht + ht_offset(l, a.nb) + (size_t)(32 * mt) * kRB;
    const uint16_t* B = a.dzt + (size_t)l * a.nb * kMlpHidden * kRB + (size_t)(32 * nt) * kRB;
    const int k8 = 8 * (lane >> 4), li = lane & 15;
    const int per = (a.nb + kGradWaves - 1) / kGradWaves, b0 = wave * per, b1 = min(a.nb, b0 + per);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int base = b0; base < b1; base += kGradChunk) {
        bf16x8 af[kGradChunk][2], bfr[kGradChunk][2];
#pragma unroll
        for (int u = 0; u < kGradChunk; ++u) {
            const int blk = min(base + u, b1 - 1);  // past the end: a repeated block, zeroed below
#pragma unroll
            for (int i = 0; i < 2; ++i) af[u][i] = ld8(A + ((size_t)blk * cols + 16 * i + li) * kRB + k8);
#pragma unroll
            for (int j = 0; j < 2; ++j) bfr[u][j] = ld8(B + ((size_t)blk * kMlpHidden + 16 * j + li) * kRB + k8);
        }
        __builtin_amdgcn_sched_barrier(0);  // the chunk's loads all in flight before the first MFMA
#pragma unroll
        for (int u = 0; u < kGradChunk; ++u) {
            if (base + u >= b1) {  // a zero dL/dz operand adds nothing
#pragma unroll
                for (int j = 0; j < 2; ++j) bfr[u][j] = bf16x8{};
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[u][i], bfr[u][j], acc[i][j], 0, 0, 0);
        }
    }
    {
        const int rq = 4 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) s_acc[wave][16 * i + rq + q][16 * j + li] = acc[i][j][q];
    }
    __syncthreads();
    // the tile's 1024 parameters, 2 consecutive columns per thread
    {
        const int r = tid >> 4, c0 = (tid & 15) * 2;
        const int base_i = mlp_offset(2 * l) + (32 * mt + r) * kMlpHidden + 32 * nt + c0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = c0 + q;
            float g = 0.f;
#pragma unroll
            for (int v = 0; v < kGradWaves; ++v) g += s_acc[v][r][c];
            a.grad[base_i + q] = g;
            if (a.adam) s_w[r][c] = adam_update(a, base_i + q, g, t_adam);
        }
    }
    if (!a.adam) return;
    __syncthreads();
    // bf16 copies as whole 16-byte fragment entries: threads 0-127 the forward copy (8 consecutive
    // input rows of one output column), threads 128-255 the backward copy (8 consecutive output
    // columns of one input row; W2..W4 only)
    const int f = (tid >> 6) & 1;
    uint32_t u[4];
    size_t dst;
    if (tid < 128) {
        const int rr = 8 * (lane >> 4), cc = 16 * f + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) u[e] = f2bf2(s_w[rr + 2 * e][cc], s_w[rr + 2 * e + 1][cc]);
        dst = (size_t)wf_offset(l) + ((size_t)((2 * nt + f) * (cols / 32) + mt) * 64 + lane) * 8;
        *reinterpret_cast<uint4*>(a.wf + dst) = make_uint4(u[0], u[1], u[2], u[3]);
    } else if (l > 0) {
        const int rr = 16 * f + (lane & 15), cc = 8 * (lane >> 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) u[e] = f2bf2(s_w[rr][cc + 2 * e], s_w[rr][cc + 2 * e + 1]);
        dst = (size_t)wb_offset(l) + ((size_t)((2 * mt + f) * 8 + nt) * 64 + lane) * 8;
        *reinterpret_cast<uint4*>(a.wb + dst) = make_uint4(u[0], u[1], u[2], u[3]);
    }
}

// bf16 fragment copies of every weight of W1..W4
__global__ void weights_to_bf16_kernel(const float* w, uint16_t* wf, uint16_t* wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mlp_offset(8)) return;
    const int g = param_region(i);
    if ((g & 1) == 0) store_bf16_copies(wf, wb, g, i, f2bf(w[i]));
}

}  // namespace

// the 32 integer frequency vectors of the features
std::vector<int8_t> mlp_frequencies() {
    std::vector<int8_t> k;
    for (int j = 0; j < 6; ++j)
        for (int a = 0; a < 3; ++a) {
            int8_t v[3] = {0, 0, 0};
            v[a] = (int8_t)(1 << j);
            k.insert(k.end(), v, v + 3);
        }
    const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int j = 0; j < 4; ++j)
        for (auto& p : pairs) {
            int8_t v[3] = {0, 0, 0};
            v[p[0]] = v[p[1]] = (int8_t)(1 << j);
            k.insert(k.end(), v, v + 3);
        }
    for (int j = 0; j < 2; ++j) {
        const int8_t v[3] = {(int8_t)(1 << j), (int8_t)(1 << j), (int8_t)(1 << j)};
        k.insert(k.end(), v, v + 3);
    }
    return k;
}

// parameter offsets (floats) in the packed vector
const int kMlpOffsets[11] = {mlp_offset(0), mlp_offset(1), mlp_offset(2), mlp_offset(3), mlp_offset(4), mlp_offset(5),
                             mlp_offset(6), mlp_offset(7), mlp_offset(8), mlp_offset(9), mlp_offset(10)};

namespace {
int blocks_of(int rows) { return (rows + kRB - 1) / kRB; }

}  // namespace

Mlp::Mlp(Context* c, int max_rows) : ctx(c), cap(max_rows) {
    const auto kv = mlp_frequencies();
    auto alloc = [&](auto** p, size_t n) {
        ALEATOR_HIP_CHECK(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(**p)));
        buffers.push_back(*p);
    };
    const size_t nb = (size_t)blocks_of(cap);
    try {
        alloc(&w, kMlpParams);
        alloc(&grad, kMlpParams);
        alloc(&m1, kMlpParams);
        alloc(&m2, kMlpParams);
        alloc(&wf, kMlpFwdFrag);
        alloc(&wb, kMlpBwdFrag);
        alloc(&kvec, kv.size());
        alloc(&ht, nb * kHtPerBlock);
        alloc(&dzt, nb * kDztPerBlock);
        alloc(&blockpart, nb * kMlpBlockPart);
        alloc(&loss_hist, kLossChunk);
        alloc(&counters, 2);
        ALEATOR_HIP_CHECK(hipMemcpyAsync(kvec, kv.data(), kv.size(), hipMemcpyHostToDevice, ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(w, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(grad, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(wf, 0, kMlpFwdFrag * sizeof(uint16_t), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(wb, 0, kMlpBwdFrag * sizeof(uint16_t), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(m1, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(m2, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(counters, 0, 2 * sizeof(uint64_t), ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        for (void* p : buffers) (void)hipFree(p);
        throw;
    }
    retain(ctx);
}

Mlp::~Mlp() {
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph();
    for (void* p : buffers) (void)hipFree(p);
    release(ctx);
}

void Mlp::set_params(const float* host) {
    ALEATOR_HIP_CHECK(hipMemcpyAsync(w, host, kMlpParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(weights_to_bf16_kernel, dim3((kMlpOffsets[8] + 255) / 256), dim3(256), 0, ctx->stream, w, wf,
                       wb);
    ALEATOR_HIP_CHECK(hipMemsetAsync(m1, 0, kMlpParams * sizeof(float), ctx->stream));
    ALEATOR_HIP_CHECK(hipMemsetAsync(m2, 0, kMlpParams * sizeof(float), ctx->stream));
    ALEATOR_HIP_CHECK(hipMemsetAsync(counters, 0, 2 * sizeof(uint64_t), ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void Mlp::get(float* host, const float* src) const {
    ALEATOR_HIP_CHECK(hipMemcpyAsync(host, src, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

namespace {
// row kernel, weight gradients and the update (with or without Adam) of one pass over `rows`
void training_pass(Mlp& net, const double* frac, const double* target, uint64_t n_samples, int rows, int mode,
                   float lr, uint64_t seed, uint64_t stream) {
    hipStream_t s = net.ctx->stream;
    const int nb = blocks_of(rows);
    RowArgs ra{};
    ra.frac = frac;
    ra.target = target;
    ra.n_samples = n_samples;
    ra.seed = seed;
    ra.stream = stream;
    ra.rows = rows;
    ra.mode = mode;
    ra.kvec = net.kvec;
    ra.wf = net.wf;
    ra.wb = net.wb;
    ra.w = net.w;
    ra.counters = net.counters;
    ra.ht = net.ht;
    ra.dzt = net.dzt;
    ra.blockpart = net.blockpart;
    hipLaunchKernelGGL(mlp_rows_kernel<true>, dim3(nb), dim3(kRowThreads), 0, s, ra);
    GradArgs ga{};
    ga.ht = net.ht;
    ga.dzt = net.dzt;
    ga.blockpart = net.blockpart;
    ga.nb = nb;
    ga.rows = rows;
    ga.w = net.w;
    ga.grad = net.grad;
    ga.m1 = net.m1;
    ga.m2 = net.m2;
    ga.wf = net.wf;
    ga.wb = net.wb;
    ga.lr = lr;
    ga.adam = mode == kTrain ? 1 : 0;
    ga.counters = net.counters;
    ga.loss_hist = net.loss_hist;
    hipLaunchKernelGGL(mlp_grad_adam_kernel, dim3(kGradTiles + kColsumBlocks), dim3(kGradThreads), 0, s, ga);
}
}  // namespace

void Mlp::loss_grad(const double* frac, const double* target, int rows) {
    training_pass(*this, frac, target, (uint64_t)rows, rows, kGrad, 0.0f, 0, 0);
}

void Mlp::iteration(const double* frac, const double* target, uint64_t n_samples, int batch, float lr,
                    uint64_t seed, uint64_t stream) {
    training_pass(*this, frac, target, n_samples, batch, kTrain, lr, seed, stream);
}

void Mlp::drop_graph() {
    for (auto& g : graph_exec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void Mlp::train(const double* frac, const double* target, uint64_t n_samples, int batch, int n_iters, float lr,
                uint64_t seed, uint64_t stream, float* loss_trace) {
    hipStream_t s = ctx->stream;
    ALEATOR_HIP_CHECK(hipMemsetAsync(counters, 0, sizeof(uint64_t), s));  // iteration 0 of this call
    const GraphKey key{frac, target, n_samples, batch, lr, seed, stream};
    if (use_graph && n_iters > 1 && !(graph_exec[0] && key == graph_key)) {
        drop_graph();
        for (int g = 0; g < 2; ++g) {
            hipGraph_t graph = nullptr;
            ALEATOR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < (g ? kGraphIters : 1); ++i) iteration(frac, target, n_samples, batch, lr, seed, stream);
            const hipError_t e = hipStreamEndCapture(s, &graph);
            ALEATOR_HIP_CHECK(e);
            const hipError_t ei = hipGraphInstantiate(&graph_exec[g], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            ALEATOR_HIP_CHECK(ei);
        }
        graph_key = key;
    }
    const bool replay = use_graph && n_iters > 1;
    for (int it0 = 0; it0 < n_iters; it0 += kLossChunk) {
        const int cnt = std::min(kLossChunk, n_iters - it0);
        for (int i = 0; i < cnt;) {
            if (replay && cnt - i >= kGraphIters) {
                ALEATOR_HIP_CHECK(hipGraphLaunch(graph_exec[1], s));
                i += kGraphIters;
            } else {
                if (replay)
                    ALEATOR_HIP_CHECK(hipGraphLaunch(graph_exec[0], s));
                else
                    iteration(frac, target, n_samples, batch, lr, seed, stream);
                ++i;
            }
        }
        ALEATOR_HIP_CHECK(hipGetLastError());
        if (loss_trace)
            ALEATOR_HIP_CHECK(hipMemcpyAsync(loss_trace + it0, loss_hist, cnt * sizeof(float), hipMemcpyDeviceToHost, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
    }
}

void Mlp::predict(const double* frac, int rows, float* out) {
    RowArgs ra{};
    ra.frac = frac;
    ra.rows = rows;
    ra.mode = kPredict;
    ra.kvec = kvec;
    ra.wf = wf;
    ra.wb = wb;
    ra.w = w;
    ra.counters = counters;
    ra.y_out = out;
    hipLaunchKernelGGL(mlp_rows_kernel<false>, dim3(blocks_of(rows)), dim3(kRowThreads), 0, ctx->stream, ra);
}

}  // namespace aleator_gpu

struct aleator_mlp {
    aleator_gpu::Mlp net;
    aleator_mlp(aleator_gpu::Context* c, int rows) : net(c, rows) {}
};

using namespace aleator_gpu;

namespace {
void check_launch() { ALEATOR_HIP_CHECK(hipGetLastError()); }
}  // namespace

extern "C" {

int64_t aleator_gpu_mlp_num_params(void) { return kMlpParams; }

aleator_status aleator_gpu_mlp_frequencies(int8_t* out) {
    return guarded([&] {
        require(out != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto kv = mlp_frequencies();
        std::memcpy(out, kv.data(), kv.size());
    });
}

aleator_status aleator_gpu_mlp_create(aleator_ctx* ctx_in, int32_t max_rows, aleator_mlp** out) {
    return guarded([&] {
        auto* ctx = reinterpret_cast<Context*>(ctx_in);
        require(ctx && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(max_rows >= 1 && max_rows <= (1 << 22), ALEATOR_INVALID_ARGUMENT, "mlp: max_rows must be in [1, 2^22]");
        *out = new aleator_mlp(ctx, max_rows);
    });
}

aleator_status aleator_gpu_mlp_destroy(aleator_mlp* m) {
    return guarded([&] { delete m; });
}

aleator_status aleator_gpu_mlp_init(aleator_mlp* m, uint64_t seed) {
    return guarded([&] {
        require(m != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        std::vector<float> p(kMlpParams, 0.0f);
        for (int l = 0; l < 5; ++l) {
            const int fan_in = l == 0 ? kMlpIn : kMlpHidden;
            const int fan_out = l == 4 ? 1 : kMlpHidden;
            const double a = l == 4 ? std::sqrt(3.0 / fan_in) : std::sqrt(6.0 / fan_in);
            for (int i = 0; i < fan_in * fan_out; ++i) {
                const int o = kMlpOffsets[2 * l] + i;
                p[o] = (float)((2.0 * philox_uniform(seed, 0, (uint64_t)o) - 1.0) * a);
            }
        }
        m->net.set_params(p.data());
    });
}

aleator_status aleator_gpu_mlp_set_params(aleator_mlp* m, const float* params) {
    return guarded([&] {
        require(m && params, ALEATOR_INVALID_ARGUMENT, "null argument");
        m->net.set_params(params);
    });
}

aleator_status aleator_gpu_mlp_get_params(aleator_mlp* m, float* params) {
    return guarded([&] {
        require(m && params, ALEATOR_INVALID_ARGUMENT, "null argument");
        m->net.get(params, m->net.w);
    });
}

aleator_status aleator_gpu_mlp_get_grads(aleator_mlp* m, float* grads) {
    return guarded([&] {
        require(m && grads, ALEATOR_INVALID_ARGUMENT, "null argument");
        m->net.get(grads, m->net.grad);
    });
}

aleator_status aleator_gpu_mlp_predict(aleator_mlp* m, const double* frac, int64_t n, float* out) {
    return guarded([&] {
        require(m && frac && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(n >= 0, ALEATOR_INVALID_ARGUMENT, "mlp: negative row count");
        Mlp& net = m->net;
        for (int64_t r0 = 0; r0 < n; r0 += net.cap) {
            const int rows = (int)std::min<int64_t>(net.cap, n - r0);
            net.predict(frac + 3 * r0, rows, out + r0);
            check_launch();
        }
    });
}

aleator_status aleator_gpu_mlp_loss_grad(aleator_mlp* m, const double* frac, const double* target, int32_t n,
                                         float* loss) {
    return guarded([&] {
        require(m && frac && target, ALEATOR_INVALID_ARGUMENT, "null argument");
        Mlp& net = m->net;
        require(n >= 1 && n <= net.cap, ALEATOR_INVALID_ARGUMENT, "mlp: rows must be in [1, max_rows]");
        ALEATOR_HIP_CHECK(hipMemsetAsync(net.counters, 0, sizeof(uint64_t), net.ctx->stream));  // loss slot 0
        net.loss_grad(frac, target, n);
        check_launch();
        float l = 0.0f;
        ALEATOR_HIP_CHECK(hipMemcpyAsync(&l, net.loss_hist, sizeof(float), hipMemcpyDeviceToHost, net.ctx->stream));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(net.ctx->stream));
        if (loss) *loss = l;
    });
}

aleator_status aleator_gpu_mlp_train(aleator_mlp* m, const double* frac, const double* target, int64_t n_samples,
                                     int32_t batch, int32_t n_iters, float lr, uint64_t seed, uint64_t stream,
                                     float* loss_trace) {
    return guarded([&] {
        require(m && frac && target, ALEATOR_INVALID_ARGUMENT, "null argument");
        Mlp& net = m->net;
        require(n_samples >= 1 && n_samples <= (int64_t)UINT32_MAX, ALEATOR_INVALID_ARGUMENT,
                "mlp: n_samples must be in [1, 2^32)");
        require(batch >= 1 && batch <= net.cap, ALEATOR_INVALID_ARGUMENT, "mlp: batch must be in [1, max_rows]");
        require(n_iters >= 0, ALEATOR_INVALID_ARGUMENT, "mlp: negative iteration count");
        require(lr > 0.0f && std::isfinite(lr), ALEATOR_INVALID_ARGUMENT, "mlp: learning rate must be positive");
        net.train(frac, target, (uint64_t)n_samples, batch, n_iters, lr, seed, stream, loss_trace);
    });
}

aleator_status aleator_gpu_mlp_use_graph(aleator_mlp* m, int32_t enable) {
    return guarded([&] {
        require(m != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        m->net.use_graph = enable != 0;
        if (!m->net.use_graph) m->net.drop_graph();
    });
}

}  // extern "C"
