// Learned subgrid closure: a 10 -> 64 -> 64 -> 64 -> 6 GELU network evaluated and
// trained with bfloat16 MFMA tiles (FP32 accumulation) on gfx950.
//
// Flat FP32 parameter vector (the caller's view): for each layer W (out x in,
// row-major) then b (out). On the device the weights also live as bfloat16
// copies laid out for the MFMA operands (rows padded to 32-element K chunks),
// plus transposed copies for back-propagation.
#pragma once

#include "tc_internal.h"

namespace tc {

constexpr int kMlpIn = 10, kMlpHid = 64, kMlpOut = 6;
// offsets in the flat parameter vector
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpHid * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHid, kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid, kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid, kOffB4 = kOffW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 9414
constexpr int kMlpRec = kMlpParams + 2;        // per-block partial record: params, loss, pad

// bfloat16 weight block (unsigned short), element offsets:
constexpr int kW1p = 0;                     // [64][32]   W1, in padded to 32
constexpr int kW2p = kW1p + 64 * 32;        // [64][64]
constexpr int kW3p = kW2p + 64 * 64;        // [64][64]
constexpr int kW4p = kW3p + 64 * 64;        // [16][64]   W4, out padded to 16
constexpr int kT2 = kW4p + 16 * 64;         // [64][64]   W2^T
constexpr int kT3 = kT2 + 64 * 64;          // [64][64]   W3^T
constexpr int kT4 = kT3 + 64 * 64;          // [64][32]   W4^T, out padded to 32
constexpr int kWBlock = kT4 + 64 * 32;      // 21504 elements
// FP32 bias block: b1[64], b2[64], b3[64], b4[16]
constexpr int kBBlock = 64 * 3 + 16;

struct MlpDev {
    const unsigned short* w = nullptr;  // kWBlock bfloat16 values
    const float* b = nullptr;           // kBBlock
};

// Closure inputs at a node are the 9 velocity-gradient components (d u_i / d x_j,
// i-major) times nu / u_tau^2 and the wall distance times inv_len * kYPlusScale;
// outputs are the deviatoric stress XX, YY, ZZ, XY, XZ, YZ over u_tau^2.
constexpr double kYPlusScale = 0.01;

// Host: bfloat16 / FP32 device blocks from the flat vector.
void mlp_pack_device(const float* params, std::vector<unsigned short>& w, std::vector<float>& b);

// y[n][6] = net(x[n][10]) (FP32 in and out).
hipError_t launch_mlp_forward(hipStream_t s, MlpDev w, const float* x, long long n, float* y);
// Loss (mean square over n x 6) gradient: per-block partial records (nblocks x kMlpRec floats).
// Blocks of the gradient kernel for n samples: one per 64-sample chunk, at most
// kMlpTrainMaxBlocks (each writes one partial record).
constexpr int kMlpTrainMaxBlocks = 512;
int mlp_train_blocks(long long n);
hipError_t launch_mlp_partials(hipStream_t s, MlpDev w, const float* x, const float* t, long long n, float* partial);
struct AdamArgs {
    int apply = 0;
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1, bc2 = 1;  // 1 - b1^step, 1 - b2^step
};
// Sums the partial records (fixed order) into grad (FP32) and loss (FP64 sum of
// squares); with a.apply, runs Adam on params/m/v and refreshes the device blocks.
// tmp: kMlpReduceSplit * kMlpRec doubles.
constexpr int kMlpReduceSplit = 16;
hipError_t launch_mlp_reduce(hipStream_t s, const float* partial, int nrec, double* tmp, float* grad, double* loss,
                             const AdamArgs& a, float* params, float* m, float* v, unsigned short* wdev, float* bdev);

// Closure stress at every node of a collocated field (walls: zero), scaled by
// u_tau^2; post: what is kept of the network's output (tc_params::sgs_model - 1):
// all of it, the stress without backscatter, or its dissipative part along the
// deviatoric strain rate.
constexpr int kSgsPostNone = 0, kSgsPostClip = 1, kSgsPostProject = 2;
template <typename T>
hipError_t launch_sgs_tau(hipStream_t s, const Mesh<T>& m, MlpDev w, const T* vel, T* tau, int post);

// Samples at the coarse interior nodes from the 9 filtered fields q (u, v, w, uu,
// vv, ww, uv, uw, vw; each [nz][nyc][nxc]): x [S][10], y [S][6], S = nxc nyc (nz-2).
template <typename T>
hipError_t launch_sgs_samples(hipStream_t s, const Mesh<T>& m, const T* q, int nxc, int nyc, T hxc, T hyc, float* x,
                              float* y);

}  // namespace tc
