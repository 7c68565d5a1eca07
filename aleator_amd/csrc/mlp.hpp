// mlp.hpp -- the Fourier-feature MLP used as a learned insertion density (see mlp.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace aleator_gpu {

constexpr int kMlpIn = 64;       // 32 frequency vectors x (sin, cos)
constexpr int kMlpFreqs = 32;
constexpr int kMlpHidden = 256;  // 4 hidden layers
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + 3 * (kMlpHidden * kMlpHidden + kMlpHidden) + kMlpHidden + 1;
constexpr int kMlpBlockRows = 32;  // rows per workgroup of the row kernel (two 16-row MFMA tiles)
constexpr int kLossChunk = 1 << 16;  // iterations whose losses are kept on the device at once
// per row block: bias gradients of the 4 hidden layers, then dw5 (256), db5 and the summed
// squared error
constexpr int kMlpBlockPart = 4 * kMlpHidden + kMlpHidden + 2;
// bf16 weights in MFMA fragment order: forward copies of W1..W4, backward copies of W2..W4
constexpr int kMlpFwdFrag = kMlpIn * kMlpHidden + 3 * kMlpHidden * kMlpHidden;
constexpr int kMlpBwdFrag = 3 * kMlpHidden * kMlpHidden;

// offset (floats) of part k = 0..9 of the packed parameter vector, W1 b1 W2 b2 W3 b3 W4 b4 w5 b5,
// and the total for k = 10; W_l is stored [fan_in][fan_out] row-major
__host__ __device__ constexpr int mlp_offset(int k) {
    constexpr int first = kMlpIn * kMlpHidden + kMlpHidden;            // W1, b1
    constexpr int layer = kMlpHidden * kMlpHidden + kMlpHidden;        // W_l, b_l (l = 2..4)
    return k == 0 ? 0
         : k == 1 ? kMlpIn * kMlpHidden
         : k < 8  ? first + ((k - 2) >> 1) * layer + ((k - 2) & 1) * kMlpHidden * kMlpHidden
         : first + 3 * layer + (k - 8) * kMlpHidden - (k == 10 ? kMlpHidden - 1 : 0);
}
static_assert(mlp_offset(10) == kMlpParams && mlp_offset(9) == kMlpParams - 1 && mlp_offset(7) == mlp_offset(6) + 65536);
extern const int kMlpOffsets[11];  // mlp_offset(0 .. 10)

std::vector<int8_t> mlp_frequencies();

struct Mlp {
    Context* ctx;
    int cap;  // rows per pass
    std::vector<void*> buffers;
    float* w = nullptr;      // fp32 master parameters
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    uint16_t* wf = nullptr;  // bf16 W1..W4 as B operands of H W (fragment order)
    uint16_t* wb = nullptr;  // bf16 W2..W4 as B operands of dZ W^T (fragment order)
    int8_t* kvec = nullptr;
    // activations of the last pass, per 32-row block and column, the block's rows contiguous:
    // X (64 columns), H1, H2, H3 (256 each)
    uint16_t* ht = nullptr;
    uint16_t* dzt = nullptr;     // dL/dz of the 4 hidden layers, same layout (256 columns each)
    float* blockpart = nullptr;  // [block][kMlpBlockPart]
    float* loss_hist = nullptr;  // per-iteration loss of the current train call chunk
    uint64_t* counters = nullptr;   // device: [0] iteration of the current train call, [1] Adam steps taken
    // a training iteration captured as a hipGraph (kernels only), replayed for every iteration
    // of train calls with the same arguments
    struct GraphKey {
        const double* frac;
        const double* target;
        uint64_t n;
        int batch;
        float lr;
        uint64_t seed, stream;
        bool operator==(const GraphKey&) const = default;
    };
    bool use_graph = true;
    // captured iterations: graph_exec[0] holds one, graph_exec[1] kGraphIters in a row (fewer graph
    // launches per step)
    static constexpr int kGraphIters = 8;
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    GraphKey graph_key{};

    Mlp(Context* c, int max_rows);
    ~Mlp();
    void set_params(const float* host);
    void get(float* host, const float* src) const;
    // loss and gradient of rows 0..rows-1 (mean squared error against target); the mean loss goes
    // to loss_hist[counters[0] mod kLossChunk]
    void loss_grad(const double* frac, const double* target, int rows);
    // one training iteration: sample a minibatch, forward, backward, Adam, advance the counters
    void iteration(const double* frac, const double* target, uint64_t n_samples, int batch, float lr, uint64_t seed,
                   uint64_t stream);
    void train(const double* frac, const double* target, uint64_t n_samples, int batch, int n_iters, float lr,
               uint64_t seed, uint64_t stream, float* loss_trace);
    void drop_graph();
    void predict(const double* frac, int rows, float* out);
};

}  // namespace aleator_gpu
