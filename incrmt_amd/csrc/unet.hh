// A small U-Net on the GPU that predicts a projection's correction from its
// starting residual: BF16 weights and activations, 3x3 convolutions as
// implicit GEMMs on the MFMA units (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), trained by its own backward kernels and Adam.
//
// No reference counterpart: it feeds the projection solvers' starting
// iterate, the hook the reference calls mgs_base::initial_guess
// (mgs_common.cc:33-66).
#ifndef INCRMT_UNET_HH
#define INCRMT_UNET_HH

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace incrmt {

typedef uint16_t bf16_t;

// FP32 <-> BF16 (round to nearest even; NaN stays NaN)
__host__ __device__ inline bf16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Four levels of the node grid: level l+1 keeps every other point of level l
// ((n + 1) / 2 points). Channels 16/32/48/64 at levels 0..3.
constexpr int kUnetLevels = 4;
constexpr int kUnetCh[kUnetLevels] = {16, 32, 48, 64};
// input channels: starting residual, level set, source, the previous
// solve's correction (zero where a caller has none)
constexpr int kUnetIn = 4;

// Convolution layers in forward order.
enum UnetLayer { kE0a = 0, kE0b, kD1, kE1, kD2, kE2, kD3, kB3, kU2, kU1, kU0, kOut, kUnetLayers };

struct ConvSpec {
    int cin1, cin2;   // input channels (second input: the skip connection of a decoder layer)
    int cout;         // output channels (the output layer has 1)
    int stride;       // 1, or 2 for the down-sampling layers
    int lin, lout;    // input and output levels
    bool relu;
    // derived
    int groups;       // 16-channel groups of the input (cin1 + cin2 rounded up to 16)
    int kp;           // K = 144 groups: index (group * 9 + tap) * 16 + channel % 16
    int cout_p;       // cout rounded up to a multiple of 16
    size_t w_off, b_off;  // offsets into the master parameter array (floats)
};

class UNet {
  public:
    ~UNet();
    // A network for an nx x ny node grid, weights from the seeded He-normal
    // initialisation, on stream s.
    int init(int nx, int ny, unsigned seed, hipStream_t s);
    int nx() const { return nx_[0]; }
    int ny() const { return ny_[0]; }
    // Packed parameter vector (per layer: weights [cout][tap][cin] then the
    // bias [cout]), host order used by get/set and by the gradients.
    size_t nparams() const { return npacked_; }
    int get_params(float *host) const;
    int set_params(const float *host);
    // Forward pass on the device input (nx*ny*kUnetIn bf16, NHWC); the
    // prediction goes to out (nx*ny floats, device).
    int forward(const bf16_t *d_in, float *d_out);
    // Loss 0.5 * mean((prediction - target)^2) of the last forward pass
    // against a device target, and the gradients of every parameter into
    // the master-gradient array (overwritten). *loss is filled after a
    // stream synchronisation when loss is not null.
    int backward(const float *d_target, double *loss);
    // The same for a given d loss / d prediction (device, nx*ny floats).
    int backward_grad(const float *d_gout);
    // Packed gradients of the last backward pass to the host.
    int get_grads(float *host) const;
    // One Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr on
    // the gradients of the last backward pass.
    int adam(float lr);
    // loss of the last backward pass (waits for the stream)
    int read_loss(double *loss);
    // Host-array conveniences for tests: input nx*ny*3 floats (channels
    // 0..2 of each node), output nx*ny floats.
    int forward_host(const float *in3, float *out);
    int backward_host(const float *in3, const float *target, double *loss, float *grads);
    bf16_t *input_buffer() { return in_; }  // device input the caller may fill
    float *output_buffer() { return out_; }
    hipStream_t stream() const { return s_; }

  private:
    int nx_[kUnetLevels] = {0}, ny_[kUnetLevels] = {0};
    hipStream_t s_ = nullptr;
    ConvSpec L_[kUnetLayers];
    size_t nmaster_ = 0, npacked_ = 0;
    float *w32_ = nullptr, *g32_ = nullptr, *m_ = nullptr, *v_ = nullptr;  // master params, grads, Adam moments
    bf16_t *w16_ = nullptr, *wt16_ = nullptr;                              // forward and transposed (backward-data) weights
    size_t wt_off_[kUnetLayers] = {0};
    long adam_t_ = 0;
    // activations (bf16, NHWC) and their gradients (fp32)
    bf16_t *in_ = nullptr;
    bf16_t *act_[kUnetLayers] = {nullptr};   // outputs of each layer (the output layer: none)
    bf16_t *up_[3] = {nullptr};               // up-sampled inputs of kU2, kU1, kU0
    float *out_ = nullptr, *gout_ = nullptr;  // prediction, d loss / d prediction
    float *gact_[kUnetLayers] = {nullptr};    // d loss / d layer output
    float *gup_[3] = {nullptr};
    bf16_t *dpre_ = nullptr;                  // masked output gradient of the layer in hand (bf16)
    const bf16_t *last_in_ = nullptr;         // input of the last forward pass
    double *d_loss_ = nullptr;
    double *h_loss_ = nullptr;
    bool w_dirty_ = true;                     // bf16 copies out of date
    std::vector<void *> allocs_;
    float *h_pack_ = nullptr;                 // pinned staging for get/set
    int alloc(void **p, size_t bytes);
    int refresh_weights();
    int conv_forward(int l, const bf16_t *in1, const bf16_t *in2, bf16_t *out_bf, float *out_f);
    int conv_backward(int l, const bf16_t *in1, const bf16_t *in2, const bf16_t *act_out, const float *gout,
                      float *gin1, float *gin2, bool accum);
};

}  // namespace incrmt

#endif
