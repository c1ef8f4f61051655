// fia_unet.hpp - the learned pressure warm start: a 4-level 3-D U-Net (16/32/48/64 channels,
// 3x3x3 convolutions, BF16 weights and activations on the matrix cores, FP32 accumulation) that
// maps the right-hand side of the pressure system and the obstacle mask to a pressure guess.
// It has no counterpart in the reference, whose solver is amgcl (src/calc_pressure_3d.cpp:141-166).
//
// Layout: activations channels-last ([x][y][z][c], z fastest, like the MAC grids), BF16;
// gradients FP32.  Parameters live as one FP32 array; each 3x3x3 convolution is W[tap][co][ci]
// (tap = (dx+1)*9 + (dy+1)*3 + (dz+1)) followed by its bias[co]; the 1x1 output layer is w[16],
// b.  Grid sizes must be multiples of 8 (three 2x2x2 poolings).
//
//   e0a 2(16)->16  e0b 16->16 = s0 | pool | e1a 16->32  e1b 32->32 = s1 | pool | e2a 32->48
//   e2b 48->48 = s2 | pool | ba 48->64  bb 64->64 | up, cat s2 | d2a 112->48  d2b 48->48 | up, cat
//   s1 | d1a 80->32  d1b 32->32 | up, cat s0 | d0a 48->16  d0b 16->16 | out 16->1 (1x1, linear)
//
// The input has 16 channels, of which only the first two are used (normalised right-hand side,
// obstacle mask); the rest are zero so that every convolution reads whole 16-channel chunks.
#pragma once
#include "fia_internal.hpp"

namespace fia {

constexpr int kUNetConvs = 14;

struct UNetConv {
    int cin, cout;    // input channels (upsampled part first, then the skip), output channels
    int level;        // resolution level (0 = full)
    int cin_real;     // channels that carry data (e0a: 2)
};
extern const UNetConv kUNetTable[kUNetConvs];

// offsets into the FP32 parameter array
size_t unet_weight_offset(int conv);
size_t unet_bias_offset(int conv);
size_t unet_out_offset();       // w[16], then b
size_t unet_param_count();

// Arguments of one 3x3x3 convolution launch (stride 1, zero padding).
struct ConvArgs {
    int nx, ny, nz;                // resolution of input and output
    const uint16_t *srcA;          // first cA channels (half resolution when upA)
    int cA, upA;
    const uint16_t *srcB;          // next cB channels (same resolution), may be null with cB = 0
    int cB;
    const uint16_t *w;             // BF16 [27][cout][cA + cB]
    const float *bias;             // [cout] (forward only)
    int cout;
    void *out;                     // BF16 [vox][cout] (forward) or FP32 [vox][cout] (data gradient)
    int relu;                      // forward: ReLU on the output
};

// mode 0: BF16 output with bias (+ ReLU); 1: FP32 output; 2: FP32 output added to what is there
hipError_t launch_conv3(const ConvArgs &a, int mode, hipStream_t s);
// per-cell flags of the fluid cells for the training loss: bits 0-5 = fluid neighbour at -x, +x,
// -y, +y, -z, +z (0 on obstacle cells); solid may be null
hipError_t launch_fluid_flags(const uint8_t *solid, int nx, int ny, int nz, uint8_t *flags, hipStream_t s);
// weight gradient of a convolution: part (FP32, chunks * 27 * cout * cin) receives per-chunk sums,
// then reduced in chunk order and added into grad_w [27][cout][cin]
size_t conv3_wgrad_chunks(size_t vox);
hipError_t launch_conv3_wgrad(const ConvArgs &in, const uint16_t *dpre, float *part, float *grad_w,
                              hipStream_t s);

// The network: FP32 master parameters, their BF16 forward and transposed (data-gradient) copies,
// gradients, Adam moments, and the activations of one grid size.
class UNet {
public:
    UNet() = default;
    ~UNet();
    UNet(const UNet &) = delete;
    UNet &operator=(const UNet &) = delete;

    hipError_t alloc_params();
    // deterministic He-normal initialisation from seed (host side, then uploaded)
    hipError_t init(unsigned long long seed, hipStream_t s);
    hipError_t set_params(const float *host, hipStream_t s);
    hipError_t get_params(float *host, hipStream_t s);
    // BF16 copies of the current parameters (after init / set / an optimiser step)
    hipError_t pack(hipStream_t s);
    hipError_t ensure_grid(int nx, int ny, int nz);

    // rhs (FP64, nx*ny*nz), solid (may be null): builds the input (rhs / scale, mask) and runs the
    // network; y (FP32) receives the output in units of scale; *scale the normalisation
    // sqrt(mean rhs^2 over fluid cells), computed on the device and read back.
    hipError_t forward(const double *rhs, const uint8_t *solid, float *y, double *scale, hipStream_t s);
    // Loss of the last forward pass against b = rhs / scale on the fluid cells (L: their Neumann
    // Laplacian, no pinned cell; flags: per-cell neighbour bits 0-5) and its gradient dy:
    //   mode 0, residual norm: |L y - b|^2 / |b|^2
    //   mode 1, energy:        (y.L y / 2 - b.y) / |b|^2, the A-norm error of y up to a constant
    // *residual receives |L y - b|^2 / |b|^2 in both modes; loss, residual, dy may be null.
    hipError_t loss(int mode, const uint8_t *flags, const double *rhs, const float *y, float *dy, double *loss,
                    double *residual, hipStream_t s);
    // initial guess of the pressure solve from the last forward pass: x = scale (y - y[pin]) on the
    // unknowns of the pressure system (pcg_flags bit 7), 0 elsewhere
    hipError_t guess(const float *y, long pin, const uint8_t *pcg_flags, double *x, hipStream_t s);
    // backward pass of the last forward pass from dL/dy; gradients are added into grad()
    hipError_t backward(const float *dy, hipStream_t s);
    hipError_t zero_grad(hipStream_t s);
    // Adam on the master parameters with the summed gradients scaled by grad_scale, then pack()
    hipError_t adam_step(float lr, float beta1, float beta2, float eps, float grad_scale, hipStream_t s);

    float *grad() { return grad_.p; }
    float *params() { return param_.p; }
    int nx() const { return nx_; }
    size_t device_bytes() const;

private:
    ConvArgs conv_in(int i) const;       // the input description of conv i (forward gather)
    int nx_ = 0, ny_ = 0, nz_ = 0;
    size_t vox_[4] = {0, 0, 0, 0};
    long adam_t_ = 0;
    DevBuf<float> param_, grad_, m_, v_;
    DevBuf<uint16_t> wf_, wt_;            // BF16 forward [27][co][ci] and flipped-transposed [27][ci][co]
    DevBuf<uint16_t> in0_;                // input [N0][16]
    DevBuf<uint16_t> act_[kUNetConvs];    // conv outputs
    DevBuf<uint16_t> pool_[3];            // pooled s0, s1, s2
    DevBuf<float> gact_[kUNetConvs];      // gradients of the conv outputs
    DevBuf<float> gpool_[3];
    DevBuf<float> gcat_;                  // data gradient of a decoder input (upsampled + skip)
    DevBuf<uint16_t> dpre_;               // BF16 gradient of a pre-activation
    DevBuf<float> wpart_;                 // weight-gradient partial sums
    DevBuf<double> red_;                  // reduction partials
    double *host_red_ = nullptr;          // pinned read-back of reductions
};

}  // namespace fia
