// mlp.hpp -- the Fourier-feature MLP used as a learned insertion density (see mlp.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace aleator_gpu {

constexpr int kMlpIn = 64;       // 32 frequency vectors x (sin, cos)
constexpr int kMlpFreqs = 32;
constexpr int kMlpHidden = 256;  // 4 hidden layers
constexpr int kMaxSplits = 32;      // row splits of the weight-gradient GEMMs
constexpr int kLossChunk = 1 << 16;  // iterations whose losses are kept on the device at once
constexpr int kMlpParams =kMlpIn * kMlpHidden + kMlpHidden + 3 * (kMlpHidden * kMlpHidden + kMlpHidden) + kMlpHidden + 1;

// offsets (floats) of W1 b1 W2 b2 W3 b3 W4 b4 w5 b5 in the packed parameter vector, then the total;
// W_l is stored [fan_in][fan_out] row-major
extern const int kMlpOffsets[11];

std::vector<int8_t> mlp_frequencies();

struct Mlp {
    Context* ctx;
    int cap;  // rows per pass
    std::vector<void*> buffers;
    float* w = nullptr;      // fp32 master parameters
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    uint16_t* wbf = nullptr; // bf16 copies of W1..W4 (and b1..b4, unused)
    int8_t* kvec = nullptr;
    uint16_t* x = nullptr;   // features [cap][64]
    float* z[4] = {};        // pre-activations [cap][256]
    uint16_t* h[4] = {};     // silu(z) in bf16
    uint16_t* dz[4] = {};    // dL/dz in bf16
    float* y = nullptr;
    float* dy = nullptr;
    float* loss_rows = nullptr;
    uint32_t* idx = nullptr;
    float* loss_hist = nullptr;  // per-iteration loss of the current train call chunk
    float* partial = nullptr;    // split-K partial weight gradients [kMaxSplits][257][256]
    float* head_partial = nullptr;  // output-layer partial sums [row chunk][258]
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
    hipGraphExec_t graph_exec = nullptr;
    GraphKey graph_key{};

    Mlp(Context* c, int max_rows);
    ~Mlp();
    void set_params(const float* host);
    void get(float* host, const float* src) const;
    // forward of rows (gathered through rows_idx when not null) up to the last hidden layer
    void forward(const double* frac, const uint32_t* rows_idx, int rows);
    // head, loss and gradients of the last forward; the mean loss goes to
    // loss_hist[counters[0] mod kLossChunk]
    void backward(const double* target, const uint32_t* rows_idx, int rows);
    // one training iteration: sample a minibatch, forward, backward, Adam, advance the counters
    void iteration(const double* frac, const double* target, uint64_t n_samples, int batch, float lr, uint64_t seed,
                   uint64_t stream);
    void train(const double* frac, const double* target, uint64_t n_samples, int batch, int n_iters, float lr,
               uint64_t seed, uint64_t stream, float* loss_trace);
    void drop_graph();
    void predict(const double* frac, int rows, float* out);
};

}  // namespace aleator_gpu
