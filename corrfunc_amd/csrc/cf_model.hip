// cf_model.hip -- the work-list cost model: a small MLP, its inference and its
// training, as hand-written HIP kernels (gfx950, FP32 arithmetic, FP64 sums
// across workgroups).
//
// Network: kNumFeatures standardised inputs -> kHidden ReLU -> kHidden ReLU ->
// one linear output, which is the standardised log2 cost of a cell pair:
//   x~   = (x - mean) * inv_sd                         (norm[0..F), norm[F..2F))
//   h1   = relu(W1 x~ + b1),  h2 = relu(W2 h1 + b2),  o = W3 . h2 + b3
//   cost = y_mean + y_sd * o                           (norm[2F], norm[2F+1])
// Loss of a minibatch of B samples: L = 1/(2B) sum (o - (y - y_mean)/y_sd)^2.
// Parameter vector layout (row-major, output index first):
//   W1[kHidden][F] | b1[kHidden] | W2[kHidden][kHidden] | b2[kHidden] | W3[kHidden] | b3
#include "cf_device.h"

namespace cfamd {

constexpr int kHidden = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kNumFeatures;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden;
constexpr int kNumParams = kOffB3 + 1;
constexpr int kNumNorm = 2 * kNumFeatures + 2;
constexpr int kGradBlock = 64;  // samples per workgroup of the gradient kernel

__device__ __forceinline__ void mlp_hidden(const float* __restrict__ W, const float* xs, float* h1, float* z2)
{
#pragma unroll
    for (int a = 0; a < kHidden; a++) {
        float z = W[kOffB1 + a];
#pragma unroll
        for (int f = 0; f < kNumFeatures; f++) z = fmaf(W[kOffW1 + a * kNumFeatures + f], xs[f], z);
        h1[a] = z > 0.f ? z : 0.f;
    }
#pragma unroll
    for (int a = 0; a < kHidden; a++) {
        float z = W[kOffB2 + a];
#pragma unroll
        for (int b = 0; b < kHidden; b++) z = fmaf(W[kOffW2 + a * kHidden + b], h1[b], z);
        z2[a] = z;
    }
}

// Predicted log2 cost of n samples (X row-major [n][F]).
// n_dev (optional): a device count that bounds n (graph-captured batches).
__global__ __launch_bounds__(256) void k_mlp_predict(int n, const float* __restrict__ X,
                                                     const float* __restrict__ norm,
                                                     const float* __restrict__ params, float* __restrict__ out,
                                                     const int* __restrict__ n_dev = nullptr)
{
    if (n_dev) n = min(n, *n_dev);
    __shared__ float W[kNumParams];
    __shared__ float nm[kNumNorm];
    for (int p = threadIdx.x; p < kNumParams; p += blockDim.x) W[p] = params[p];
    for (int p = threadIdx.x; p < kNumNorm; p += blockDim.x) nm[p] = norm[p];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xs[kNumFeatures], h1[kHidden], z2[kHidden];
#pragma unroll
    for (int f = 0; f < kNumFeatures; f++) xs[f] = (X[(int64_t)s * kNumFeatures + f] - nm[f]) * nm[kNumFeatures + f];
    mlp_hidden(W, xs, h1, z2);
    float o = W[kOffB3];
#pragma unroll
    for (int a = 0; a < kHidden; a++) o = fmaf(W[kOffW3 + a], z2[a] > 0.f ? z2[a] : 0.f, o);
    out[s] = nm[2 * kNumFeatures] + nm[2 * kNumFeatures + 1] * o;
}

// Per-workgroup gradient of the minibatch loss: samples idx[0..nb) (idx ==
// nullptr: 0..nb), kGradBlock per workgroup; partial[block][kNumParams] and
// partial_loss[block] are summed by k_mlp_reduce. inv_batch = 1/B.
__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(int nb, const int* __restrict__ idx,
                                                         const float* __restrict__ X, const float* __restrict__ Y,
                                                         const float* __restrict__ norm,
                                                         const float* __restrict__ params, float inv_batch,
                                                         float* __restrict__ partial,
                                                         float* __restrict__ partial_loss)
{
    constexpr int S = kGradBlock, HP = kHidden + 1;  // +1: no LDS bank conflicts on column reads
    __shared__ float W[kNumParams];
    __shared__ float nm[kNumNorm];
    __shared__ float xs_s[S][kNumFeatures];
    __shared__ float h1_s[S][HP], h2_s[S][HP], d1_s[S][HP], d2_s[S][HP];
    __shared__ float do_s[S];
    for (int p = threadIdx.x; p < kNumParams; p += blockDim.x) W[p] = params[p];
    for (int p = threadIdx.x; p < kNumNorm; p += blockDim.x) nm[p] = norm[p];
    __syncthreads();
    const int t = threadIdx.x;
    const int s = blockIdx.x * S + t;
    const bool valid = s < nb;
    const int64_t row = valid ? (idx ? idx[s] : s) : 0;
    float xs[kNumFeatures], h1[kHidden], z2[kHidden];
#pragma unroll
    for (int f = 0; f < kNumFeatures; f++) {
        xs[f] = valid ? (X[row * kNumFeatures + f] - nm[f]) * nm[kNumFeatures + f] : 0.f;
        xs_s[t][f] = xs[f];
    }
    mlp_hidden(W, xs, h1, z2);
    float o = W[kOffB3];
#pragma unroll
    for (int a = 0; a < kHidden; a++) o = fmaf(W[kOffW3 + a], z2[a] > 0.f ? z2[a] : 0.f, o);
    const float yt = valid ? (Y[row] - nm[2 * kNumFeatures]) / nm[2 * kNumFeatures + 1] : 0.f;
    const float err = valid ? o - yt : 0.f;
    const float dout = err * inv_batch;
    float loss = 0.5f * err * err * inv_batch;
    float d2[kHidden];
#pragma unroll
    for (int a = 0; a < kHidden; a++) {
        d2[a] = z2[a] > 0.f ? dout * W[kOffW3 + a] : 0.f;
        h1_s[t][a] = h1[a];
        h2_s[t][a] = z2[a] > 0.f ? z2[a] : 0.f;
        d2_s[t][a] = d2[a];
    }
#pragma unroll
    for (int b = 0; b < kHidden; b++) {
        float g = 0.f;
#pragma unroll
        for (int a = 0; a < kHidden; a++) g = fmaf(W[kOffW2 + a * kHidden + b], d2[a], g);
        d1_s[t][b] = h1[b] > 0.f ? g : 0.f;
    }
    do_s[t] = dout;
    __syncthreads();
    for (int p = t; p < kNumParams; p += S) {
        float g = 0.f;
        if (p < kOffB1) {
            const int a = p / kNumFeatures, f = p % kNumFeatures;
            for (int k = 0; k < S; k++) g = fmaf(d1_s[k][a], xs_s[k][f], g);
        } else if (p < kOffW2) {
            const int a = p - kOffB1;
            for (int k = 0; k < S; k++) g += d1_s[k][a];
        } else if (p < kOffB2) {
            const int a = (p - kOffW2) / kHidden, b = (p - kOffW2) % kHidden;
            for (int k = 0; k < S; k++) g = fmaf(d2_s[k][a], h1_s[k][b], g);
        } else if (p < kOffW3) {
            const int a = p - kOffB2;
            for (int k = 0; k < S; k++) g += d2_s[k][a];
        } else if (p < kOffB3) {
            const int b = p - kOffW3;
            for (int k = 0; k < S; k++) g = fmaf(do_s[k], h2_s[k][b], g);
        } else {
            for (int k = 0; k < S; k++) g += do_s[k];
        }
        partial[(int64_t)blockIdx.x * kNumParams + p] = g;
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) loss += __shfl_xor(loss, o2, 64);
    if (t == 0) partial_loss[blockIdx.x] = loss;
}

// grad[p] = sum over workgroups of partial[.][p] (fixed order, FP64), and the loss.
__global__ void k_mlp_reduce(int nblocks, const float* __restrict__ partial, const float* __restrict__ partial_loss,
                             float* __restrict__ grad, float* __restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNumParams) {
        double g = 0.0;
        for (int b = 0; b < nblocks; b++) g += (double)partial[(int64_t)b * kNumParams + p];
        grad[p] = (float)g;
    } else if (p == kNumParams) {
        double l = 0.0;
        for (int b = 0; b < nblocks; b++) l += (double)partial_loss[b];
        *loss = (float)l;
    }
}

// One Adam step (Kingma & Ba 2015, bias-corrected moments). The constants come
// from the host, computed in FP64 and rounded once: 1 - beta1, 1 - beta2 and
// the bias corrections 1 / (1 - beta^step) differ from the same quantities
// formed in FP32 from rounded betas (1 - 0.999f is 1.3e-5 off 0.001).
struct AdamConsts {
    float lr, b1, omb1, b2, omb2, inv_bc1, inv_bc2, eps;
};

static inline AdamConsts adam_consts(int step, float lr, double beta1 = 0.9, double beta2 = 0.999, double eps = 1e-8)
{
    return AdamConsts{lr, (float)beta1, (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2),
                      (float)(1.0 / (1.0 - pow(beta1, step))), (float)(1.0 / (1.0 - pow(beta2, step))), (float)eps};
}

__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, AdamConsts c)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    const float g = grad[p];
    const float mp = c.b1 * m[p] + c.omb1 * g;
    const float vp = c.b2 * v[p] + c.omb2 * (g * g);
    m[p] = mp;
    v[p] = vp;
    const float mh = mp * c.inv_bc1;
    const float vh = vp * c.inv_bc2;
    params[p] -= c.lr * mh / (sqrtf(vh) + c.eps);
}

}  // namespace cfamd
