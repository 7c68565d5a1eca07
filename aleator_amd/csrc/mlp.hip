// mlp.hip -- the learned insertion proposal: a Fourier-feature MLP trained and evaluated on the
// GPU with hand-written BF16 MFMA GEMMs (FP32 accumulation).  No reference counterpart.
//
// Network: 64 Fourier features of the fractional position -> 4 hidden layers of width 256
// with SiLU -> 1 output (214,273 parameters).  Features: sin and cos of 2*pi*(k . f) for 32
// integer frequency vectors k, so the network is periodic in the cell: 2^j e_a (j = 0..5,
// a = 0..2; 18 vectors), 2^j (e_a + e_b) (j = 0..3, three axis pairs; 12), 2^j (1,1,1)
// (j = 0..1; 2).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with A, B in BF16 (arbitrary strides, staged through
// LDS) and FP32 accumulation on v_mfma_f32_16x16x16_bf16: a 256-thread workgroup computes a
// 64 x 64 block of C, each wave a 32 x 32 quarter as 2 x 2 MFMA tiles, K in steps of 16.
// Operand layout of the 16x16x16 instruction (64 lanes): A lane l holds A[l%16][4(l/16)..+3],
// B lane l holds B[4(l/16)..+3][l%16], D lane l holds D[4(l/16)..+3][l%16].
// Epilogues fuse bias + SiLU (forward), the SiLU derivative (backward) and FP32 stores.
// Training is Adam on FP32 master weights; BF16 copies feed the MFMA.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace aleator_gpu {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int kBM = 64, kBN = 64, kBK = 16, kKC = 256;  // tile, MFMA K step, staged K chunk

enum Epilogue : int {
    kStoreF32 = 0,   // C (fp32)
    kBiasSilu = 1,   // Z = C + bias (fp32), H = silu(Z) (bf16)
    kSiluGrad = 2,   // dZ = C * silu'(Zprev) (bf16)
};

struct GemmArgs {
    const uint16_t* a;
    long long sam, sak;
    const uint16_t* b;
    long long sbk, sbn;
    int m, n, k;
    int epilogue;
    float* c;          // fp32 out [m][n] (kStoreF32, kBiasSilu: Z); split-K: partials [split][m][n]
    const float* bias; // [n]
    uint16_t* h;       // bf16 out [m][n] (kBiasSilu: H, kSiluGrad: dZ)
    const float* zref; // [m][n] (kSiluGrad)
    int ones_row;      // row of A that reads as 1 (bias gradients as one more GEMM row), or -1
    int kchunk;        // K per split (a multiple of kBK); gridDim.z splits
};

// Load one element of A / B (bounds and the ones row handled).
__device__ inline uint16_t a_at(const GemmArgs& g, int m, int k, int kend) {
    if (k >= kend) return 0;
    if (m == g.ones_row) return 0x3f80;  // 1.0
    return m < g.m ? g.a[m * g.sam + k * g.sak] : (uint16_t)0;
}
__device__ inline uint16_t b_at(const GemmArgs& g, int k, int n, int kend) {
    return (k < kend && n < g.n) ? g.b[k * g.sbk + n * g.sbn] : (uint16_t)0;
}

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// Stage a kBM x kc chunk of A as As[m][k] and a kc x kBN chunk of B as Bs[n][k] (k contiguous in
// both, the MFMA operand order), with 16-byte global loads along whichever of the two indices
// is contiguous in memory; groups that cross a bound are loaded element by element.
__device__ inline void stage(const GemmArgs& g, int m0, int n0, int k0, int kc, int kend, uint16_t (*As)[kKC + 8],
                             uint16_t (*Bs)[kKC + 8]) {
    const int tid = threadIdx.x;
    // A
    if (g.sak == 1) {  // rows contiguous in k
        for (int i = tid; i < kBM * (kc / 8); i += 256) {
            const int r = i / (kc / 8), c8 = (i % (kc / 8)) * 8;
            const int gm = m0 + r, gk = k0 + c8;
            u16x8 v;
            if (gm < g.m && gm != g.ones_row && gk + 8 <= kend) {
                v = *reinterpret_cast<const u16x8*>(g.a + gm * g.sam + gk);
            } else {
                for (int t = 0; t < 8; ++t) v[t] = a_at(g, gm, gk + t, kend);
            }
            *reinterpret_cast<u16x8*>(&As[r][c8]) = v;
        }
    } else if (g.sam != 1) {  // neither index contiguous: element by element
        for (int i = tid; i < kBM * kc; i += 256) As[i / kc][i % kc] = a_at(g, m0 + i / kc, k0 + i % kc, kend);
    } else {  // columns contiguous in m
        for (int i = tid; i < (kBM / 8) * kc; i += 256) {
            const int k = i / (kBM / 8), r8 = (i % (kBM / 8)) * 8;
            const int gm = m0 + r8, gk = k0 + k;
            u16x8 v;
            if (gm + 8 <= g.m && gk < kend && (g.ones_row < gm || g.ones_row >= gm + 8)) {
                v = *reinterpret_cast<const u16x8*>(g.a + gm + gk * g.sak);
            } else {
                for (int t = 0; t < 8; ++t) v[t] = a_at(g, gm + t, gk, kend);
            }
            for (int t = 0; t < 8; ++t) As[r8 + t][k] = v[t];
        }
    }
    // B
    if (g.sbk == 1) {  // columns of B contiguous in k
        for (int i = tid; i < kBN * (kc / 8); i += 256) {
            const int n = i / (kc / 8), c8 = (i % (kc / 8)) * 8;
            const int gn = n0 + n, gk = k0 + c8;
            u16x8 v;
            if (gn < g.n && gk + 8 <= kend) {
                v = *reinterpret_cast<const u16x8*>(g.b + gk + gn * g.sbn);
            } else {
                for (int t = 0; t < 8; ++t) v[t] = b_at(g, gk + t, gn, kend);
            }
            *reinterpret_cast<u16x8*>(&Bs[n][c8]) = v;
        }
    } else if (g.sbn != 1) {
        for (int i = tid; i < kBN * kc; i += 256) Bs[i / kc][i % kc] = b_at(g, k0 + i % kc, n0 + i / kc, kend);
    } else {  // rows of B contiguous in n
        for (int i = tid; i < (kBN / 8) * kc; i += 256) {
            const int k = i / (kBN / 8), n8 = (i % (kBN / 8)) * 8;
            const int gn = n0 + n8, gk = k0 + k;
            u16x8 v;
            if (gn + 8 <= g.n && gk < kend) {
                v = *reinterpret_cast<const u16x8*>(g.b + gk * g.sbk + gn);
            } else {
                for (int t = 0; t < 8; ++t) v[t] = b_at(g, gk, gn + t, kend);
            }
            for (int t = 0; t < 8; ++t) Bs[n8 + t][k] = v[t];
        }
    }
}

// One workgroup: a kBM x kBN block of C over the K range of its split, staged kKC at a time.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kBM][kKC + 8];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN][kKC + 8];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.k, kbeg + g.kchunk);
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4), li = lane & 15;
    for (int k0 = kbeg; k0 < kend; k0 += kKC) {
        const int kc = min(kKC, (kend - k0 + 15) / 16 * 16);  // whole MFMA steps; the tail is zero
        stage(g, m0, n0, k0, kc, kend, As, Bs);
        __syncthreads();
        for (int kk = 0; kk < kc; kk += kBK) {
            bf16x4 af[2], bfr[2];
            for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x4*>(&As[wm + 16 * i + li][kk + kq]);
            for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x4*>(&Bs[wn + 16 * j + li][kk + kq]);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    const int mrows = g.ones_row >= 0 ? g.ones_row + 1 : g.m;
    const int rq = 4 * (lane >> 4), cl = lane & 15;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * i + rq + r, gn = n0 + wn + 16 * j + cl;
                if (gm >= mrows || gn >= g.n) continue;
                const float v = acc[i][j][r];
                const long long o = (long long)gm * g.n + gn;
                if (g.epilogue == kStoreF32) {
                    g.c[(long long)blockIdx.z * mrows * g.n + o] = v;
                } else if (g.epilogue == kBiasSilu) {
                    const float z = v + g.bias[gn];
                    g.c[o] = z;
                    g.h[o] = f2bf(silu(z));
                } else {
                    g.h[o] = f2bf(v * silu_grad(g.zref[o]));
                }
            }
}

// out[o] = sum over splits of partial[split][o], in split order
__global__ void split_reduce_kernel(const float* partial, int splits, long long n, float* out) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += partial[z * n + o];
    out[o] = s;
}

// Fourier features of fractional positions (optionally gathered through idx), bf16 [rows][64]
__global__ void features_kernel(const double* frac, const uint32_t* idx, int rows, const int8_t* kv,
                                uint16_t* x) {
    const int r = blockIdx.x * blockDim.y + threadIdx.y;
    const int c = threadIdx.x;  // 0..31 frequency
    if (r >= rows) return;
    const long long p = idx ? (long long)idx[r] : r;
    const double f0 = frac[3 * p], f1 = frac[3 * p + 1], f2 = frac[3 * p + 2];
    const double ph = kv[3 * c] * f0 + kv[3 * c + 1] * f1 + kv[3 * c + 2] * f2;
    double s, co;
    sincospi(2.0 * ph, &s, &co);
    x[(long long)r * kMlpIn + 2 * c] = f2bf((float)s);
    x[(long long)r * kMlpIn + 2 * c + 1] = f2bf((float)co);
}

// Output layer and loss: y = H4 . w5 + b5; dy = 2 (y - t) / B; dZ4 = dy w5 silu'(Z4).
// One wave per row.
__global__ void head_kernel(const uint16_t* h4, const float* z4, const float* w5, const float* b5, const double* target,
                            const uint32_t* idx, int rows, float* y_out, float* dy_out, uint16_t* dz4, float* loss_rows,
                            int backward) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    if (r >= rows) return;
    float s = 0.f;
    for (int k = lane; k < kMlpHidden; k += 64) s += bf2f(h4[(long long)r * kMlpHidden + k]) * w5[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const float y = s + b5[0];
    if (lane == 0 && y_out) y_out[r] = y;
    if (!backward) return;
    const long long p = idx ? (long long)idx[r] : r;
    const float d = y - (float)target[p];
    const float dy = 2.0f * d / (float)rows;
    if (lane == 0) {
        dy_out[r] = dy;
        loss_rows[r] = d * d;
    }
    for (int k = lane; k < kMlpHidden; k += 64) {
        const long long o = (long long)r * kMlpHidden + k;
        dz4[o] = f2bf(dy * w5[k] * silu_grad(z4[o]));
    }
}

// Output-layer gradients over one chunk of kHeadChunk rows (one block per chunk): partial[chunk][k] =
// sum_r H4[r][k] dy[r] for k < 256, then sum_r dy[r] and sum_r (y - t)^2.
constexpr int kHeadChunk = 32;
__global__ void head_grad_partial_kernel(const uint16_t* h4, const float* dy, const float* loss_rows, int rows,
                                         float* partial) {
    const int k = threadIdx.x, r0 = blockIdx.x * kHeadChunk;
    const int r1 = min(rows, r0 + kHeadChunk);
    float s = 0.f;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += bf2f(h4[(long long)r * kMlpHidden + k]) * dy[r];
    float* out = partial + (long long)blockIdx.x * (kMlpHidden + 2);
    out[k] = s;
    if (k == 0) {
        float sd = 0.f, sl = 0.f;
        for (int r = r0; r < r1; ++r) {
            sd += dy[r];
            sl += loss_rows[r];
        }
        out[kMlpHidden] = sd;
        out[kMlpHidden + 1] = sl;
    }
}

// dW5, db5 and the mean loss from the chunk partials
// The mean loss goes to loss_hist[iteration mod kLossChunk] (the iteration counter counters[0]).
// One wave per output k (258 blocks of 64): lanes take chunks l, l + 64, ..., then a fixed-order
// wave reduction.
__global__ void head_grad_reduce_kernel(const float* partial, int chunks, int rows, float* dw5, float* db5,
                                        float* loss_hist, const uint64_t* counters) {
    const int k = blockIdx.x, lane = threadIdx.x;  // k: 0 .. 257
    float s = 0.f;
    for (int c = lane; c < chunks; c += 64) s += partial[(long long)c * (kMlpHidden + 2) + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane != 0) return;
    if (k < kMlpHidden)
        dw5[k] = s;
    else if (k == kMlpHidden)
        db5[0] = s;
    else
        loss_hist[counters[0] % kLossChunk] = s / (float)rows;
}

// Adam on fp32 master weights, step t = counters[1] + 1 (bias corrections 1 - b^t); refreshes
// the bf16 copies of the GEMM weights
__global__ void adam_kernel(float* w, const float* grad, float* m1, float* m2, uint16_t* wbf, int n, int n_bf,
                            float lr, float b1, float b2, float eps, const uint64_t* counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = (float)(counters[1] + 1);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float v = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float upd = lr * (a / c1) / (sqrtf(v / c2) + eps);
    const float nw = w[i] - upd;
    w[i] = nw;
    if (i < n_bf) wbf[i] = f2bf(nw);
}

// end of a training iteration: next iteration, one more Adam step
__global__ void step_kernel(uint64_t* counters) {
    counters[0] += 1;
    counters[1] += 1;
}

__global__ void to_bf16_kernel(const float* w, uint16_t* wbf, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wbf[i] = f2bf(w[i]);
}

// batch indices: idx[r] = floor(u * n) from Philox(seed, stream) at call counters[0] * rows + r
__global__ void sample_kernel(uint64_t seed, uint64_t stream, const uint64_t* counters, int rows, uint64_t n,
                              uint32_t* idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double u = philox_uniform(seed, stream, counters[0] * (uint64_t)rows + r);
    uint64_t v = (uint64_t)(u * (double)n);
    idx[r] = (uint32_t)(v < n ? v : n - 1);
}

// Launches C = A B split into at most `splits` K ranges (multiples of kBK); returns the count used.
int gemm(GemmArgs g, hipStream_t s, int splits = 1) {
    const int mrows = g.ones_row >= 0 ? g.ones_row + 1 : g.m;
    g.kchunk = ((g.k + splits - 1) / splits + kBK - 1) / kBK * kBK;
    splits = (g.k + g.kchunk - 1) / g.kchunk;
    dim3 grid((g.n + kBN - 1) / kBN, (mrows + kBM - 1) / kBM, splits);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, g);
    return splits;
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
const int kMlpOffsets[11] = {
    0,                                      // W1 [64][256]
    kMlpIn * kMlpHidden,                    // b1
    kMlpIn * kMlpHidden + kMlpHidden,       // W2 [256][256]
    kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden,  // b2
    kMlpIn * kMlpHidden + 2 * kMlpHidden + kMlpHidden * kMlpHidden,                            // W3
    kMlpIn * kMlpHidden + 2 * kMlpHidden + 2 * kMlpHidden * kMlpHidden,                        // b3
    kMlpIn * kMlpHidden + 3 * kMlpHidden + 2 * kMlpHidden * kMlpHidden,                        // W4
    kMlpIn * kMlpHidden + 3 * kMlpHidden + 3 * kMlpHidden * kMlpHidden,                        // b4
    kMlpIn * kMlpHidden + 4 * kMlpHidden + 3 * kMlpHidden * kMlpHidden,                        // W5 [256]
    kMlpIn * kMlpHidden + 5 * kMlpHidden + 3 * kMlpHidden * kMlpHidden,                        // b5
    kMlpParams};

Mlp::Mlp(Context* c, int max_rows) : ctx(c), cap(max_rows) {
    const auto kv = mlp_frequencies();
    auto alloc = [&](auto** p, size_t n) {
        ALEATOR_HIP_CHECK(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(**p)));
        buffers.push_back(*p);
    };
    try {
        alloc(&w, kMlpParams);
        alloc(&grad, kMlpParams);
        alloc(&m1, kMlpParams);
        alloc(&m2, kMlpParams);
        alloc(&wbf, kMlpOffsets[8]);
        alloc(&kvec, kv.size());
        alloc(&x, (size_t)cap * kMlpIn);
        for (int l = 0; l < 4; ++l) {
            alloc(&z[l], (size_t)cap * kMlpHidden);
            alloc(&h[l], (size_t)cap * kMlpHidden);
            alloc(&dz[l], (size_t)cap * kMlpHidden);
        }
        alloc(&y, cap);
        alloc(&dy, cap);
        alloc(&loss_rows, cap);
        alloc(&idx, cap);
        alloc(&loss_hist, kLossChunk);
        alloc(&counters, 2);
        alloc(&partial, (size_t)kMaxSplits * (kMlpHidden + 1) * kMlpHidden);
        alloc(&head_partial, (size_t)((cap + kHeadChunk - 1) / kHeadChunk) * (kMlpHidden + 2));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(kvec, kv.data(), kv.size(), hipMemcpyHostToDevice, ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(w, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(grad, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(wbf, 0, kMlpOffsets[8] * sizeof(uint16_t), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(m1, 0, kMlpParams * sizeof(float), ctx->stream));
        ALEATOR_HIP_CHECK(hipMemsetAsync(m2, 0, kMlpParams * sizeof(float), ctx->stream));
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
    hipLaunchKernelGGL(to_bf16_kernel, dim3((kMlpOffsets[8] + 255) / 256), dim3(256), 0, ctx->stream, w, wbf,
                       kMlpOffsets[8]);
    ALEATOR_HIP_CHECK(hipMemsetAsync(m1, 0, kMlpParams * sizeof(float), ctx->stream));
    ALEATOR_HIP_CHECK(hipMemsetAsync(m2, 0, kMlpParams * sizeof(float), ctx->stream));
    ALEATOR_HIP_CHECK(hipMemsetAsync(counters, 0, 2 * sizeof(uint64_t), ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void Mlp::get(float* host, const float* src) const {
    ALEATOR_HIP_CHECK(hipMemcpyAsync(host, src, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    ALEATOR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void Mlp::forward(const double* frac, const uint32_t* rows_idx, int rows) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(features_kernel, dim3((rows + 7) / 8), dim3(32, 8), 0, s, frac, rows_idx, rows, kvec, x);
    const uint16_t* in = x;
    int kin = kMlpIn;
    for (int l = 0; l < 4; ++l) {
        GemmArgs g{};
        g.a = in;
        g.sam = kin;
        g.sak = 1;
        g.b = wbf + kMlpOffsets[2 * l];
        g.sbk = kMlpHidden;
        g.sbn = 1;
        g.m = rows;
        g.n = kMlpHidden;
        g.k = kin;
        g.epilogue = kBiasSilu;
        g.c = z[l];
        g.bias = w + kMlpOffsets[2 * l + 1];
        g.h = h[l];
        g.ones_row = -1;
        gemm(g, s);
        in = h[l];
        kin = kMlpHidden;
    }
}

void Mlp::backward(const double* target, const uint32_t* rows_idx, int rows) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(head_kernel, dim3((rows + 3) / 4), dim3(256), 0, s, h[3], z[3], w + kMlpOffsets[8],
                       w + kMlpOffsets[9], target, rows_idx, rows, y, dy, dz[3], loss_rows, 1);
    const int chunks = (rows + kHeadChunk - 1) / kHeadChunk;
    hipLaunchKernelGGL(head_grad_partial_kernel, dim3(chunks), dim3(kMlpHidden), 0, s, h[3], dy, loss_rows, rows,
                       head_partial);
    hipLaunchKernelGGL(head_grad_reduce_kernel, dim3(kMlpHidden + 2), dim3(64), 0, s, head_partial, chunks, rows,
                       grad + kMlpOffsets[8], grad + kMlpOffsets[9], loss_hist, counters);
    const int splits = std::max(1, std::min(kMaxSplits, rows / 256));
    for (int l = 3; l >= 0; --l) {
        // [dW_l; db_l] = [H_{l-1}^T; 1^T] dZ_l   ([kin + 1 x rows] x [rows x 256]), split over rows;
        // W_l and b_l are adjacent in the parameter vector, so the result lands on both
        const uint16_t* hin = l == 0 ? x : h[l - 1];
        const int kin = l == 0 ? kMlpIn : kMlpHidden;
        GemmArgs gw{};
        gw.a = hin;
        gw.sam = 1;
        gw.sak = kin;
        gw.b = dz[l];
        gw.sbk = kMlpHidden;
        gw.sbn = 1;
        gw.m = kin;
        gw.n = kMlpHidden;
        gw.k = rows;
        gw.epilogue = kStoreF32;
        gw.ones_row = kin;
        const long long n_out = (long long)(kin + 1) * kMlpHidden;
        if (splits == 1) {
            gw.c = grad + kMlpOffsets[2 * l];
            gemm(gw, s);
        } else {
            gw.c = partial;
            const int used = gemm(gw, s, splits);
            hipLaunchKernelGGL(split_reduce_kernel, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, s, partial,
                               used, n_out, grad + kMlpOffsets[2 * l]);
        }
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1})   ([rows x 256] x [256 x 256])
        GemmArgs gd{};
        gd.a = dz[l];
        gd.sam = kMlpHidden;
        gd.sak = 1;
        gd.b = wbf + kMlpOffsets[2 * l];
        gd.sbk = 1;
        gd.sbn = kMlpHidden;
        gd.m = rows;
        gd.n = kMlpHidden;
        gd.k = kMlpHidden;
        gd.epilogue = kSiluGrad;
        gd.h = dz[l - 1];
        gd.zref = z[l - 1];
        gd.ones_row = -1;
        gemm(gd, s);
    }
}

void Mlp::iteration(const double* frac, const double* target, uint64_t n_samples, int batch, float lr,
                    uint64_t seed, uint64_t stream) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(sample_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, seed, stream, counters, batch,
                       n_samples, idx);
    forward(frac, idx, batch);
    backward(target, idx, batch);
    hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, w, grad, m1, m2, wbf, kMlpParams,
                       kMlpOffsets[8], lr, 0.9f, 0.999f, 1e-8f, counters);
    hipLaunchKernelGGL(step_kernel, dim3(1), dim3(1), 0, s, counters);
}

void Mlp::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
}

void Mlp::train(const double* frac, const double* target, uint64_t n_samples, int batch, int n_iters, float lr,
                uint64_t seed, uint64_t stream, float* loss_trace) {
    hipStream_t s = ctx->stream;
    ALEATOR_HIP_CHECK(hipMemsetAsync(counters, 0, sizeof(uint64_t), s));  // iteration 0 of this call
    const GraphKey key{frac, target, n_samples, batch, lr, seed, stream};
    if (use_graph && n_iters > 1 && !(graph_exec && key == graph_key)) {
        drop_graph();
        hipGraph_t graph = nullptr;
        ALEATOR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iteration(frac, target, n_samples, batch, lr, seed, stream);
        const hipError_t e = hipStreamEndCapture(s, &graph);
        ALEATOR_HIP_CHECK(e);
        const hipError_t ei = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        ALEATOR_HIP_CHECK(ei);
        graph_key = key;
    }
    const bool replay = use_graph && n_iters > 1;
    for (int it0 = 0; it0 < n_iters; it0 += kLossChunk) {
        const int cnt = std::min(kLossChunk, n_iters - it0);
        for (int i = 0; i < cnt; ++i) {
            if (replay)
                ALEATOR_HIP_CHECK(hipGraphLaunch(graph_exec, s));
            else
                iteration(frac, target, n_samples, batch, lr, seed, stream);
        }
        ALEATOR_HIP_CHECK(hipGetLastError());
        if (loss_trace)
            ALEATOR_HIP_CHECK(hipMemcpyAsync(loss_trace + it0, loss_hist, cnt * sizeof(float), hipMemcpyDeviceToHost, s));
        ALEATOR_HIP_CHECK(hipStreamSynchronize(s));
    }
}

void Mlp::predict(const double* frac, int rows, float* out) {
    forward(frac, nullptr, rows);
    hipLaunchKernelGGL(head_kernel, dim3((rows + 3) / 4), dim3(256), 0, ctx->stream, h[3], z[3], w + kMlpOffsets[8],
                       w + kMlpOffsets[9], nullptr, nullptr, rows, out, nullptr, nullptr, nullptr, 0);
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
        net.forward(frac, nullptr, n);
        net.backward(target, nullptr, n);
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
