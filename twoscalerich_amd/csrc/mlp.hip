// mlp.hip - the learned step controller: a 16 -> 64 -> 64 -> 2 MLP with GELU, BF16
// weights and FP32 accumulation, trained and evaluated by the kernels below.
//
// Layout of the FP32 master parameters (row-major, kParams floats):
//   W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
// The network computes with bf16(W) (round to nearest even of the FP32 master), FP32
// biases and FP32 activations; Adam updates the FP32 masters (straight-through).
//
// Mapping on CDNA4: one wavefront per sample, lane j = neuron j of the 64-wide layers.
// Weights (converted to FP32 once per workgroup) and the per-wavefront activations live
// in LDS.  The backward pass keeps each lane's row of dW2 / dW1 in registers across the
// wavefront's samples, writes one partial gradient per wavefront, and k_mlp_reduce sums
// the partials in wavefront order (deterministic).
#include "device.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace tsr {

constexpr int kIn = 16, kH = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kParams = oB3 + kOut;
constexpr int kMlpWaves = 4;   // wavefronts per workgroup

__device__ __forceinline__ float bf16_round(float x)
{
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;   // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// Load bf16(master) weights and FP32 biases into LDS (kParams floats).
__device__ __forceinline__ void load_params(const float* __restrict__ P, float* sp)
{
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
        const bool is_w = (i < oB1) || (i >= oW2 && i < oB2) || (i >= oW3 && i < oB3);
        sp[i] = is_w ? bf16_round(P[i]) : P[i];
    }
    __syncthreads();
}

// Forward for one sample on one wavefront.  x: 16 floats in LDS; returns z1, z2 of this
// lane's neuron and the two outputs (valid in every lane).
__device__ __forceinline__ void forward_one(const float* sp, const float* x, float* a1s, float& z1, float& z2,
                                            float& a2, float y[2])
{
    const int j = threadIdx.x & 63;
    float acc = sp[oB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) acc += sp[oW1 + j * kIn + k] * x[k];
    z1 = acc;
    a1s[j] = gelu(z1);
    __builtin_amdgcn_wave_barrier();
    acc = sp[oB2 + j];
#pragma unroll 8
    for (int k = 0; k < kH; ++k) acc += sp[oW2 + j * kH + k] * a1s[k];
    z2 = acc;
    a2 = gelu(z2);
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
        float p = sp[oW3 + o * kH + j] * a2;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
        y[o] = p + sp[oB3 + o];
    }
}

// out[n][2] = MLP(x[n][16]) for n samples (inputs already standardised).
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_forward(const float* P, const float* X, int n, float* out)
{
    __shared__ float sp[kParams];
    __shared__ float xs[kMlpWaves][kIn];
    __shared__ float a1s[kMlpWaves][kH];
    load_params(P, sp);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int smp = blockIdx.x * kMlpWaves + w; smp < n; smp += gridDim.x * kMlpWaves) {
        if (lane < kIn) xs[w][lane] = X[(size_t)smp * kIn + lane];
        __builtin_amdgcn_wave_barrier();
        float z1, z2, a2, y[2];
        forward_one(sp, xs[w], a1s[w], z1, z2, a2, y);
        if (lane == 0) { out[(size_t)smp * 2] = y[0]; out[(size_t)smp * 2 + 1] = y[1]; }
        __builtin_amdgcn_wave_barrier();
    }
}

// Gradient of L = sum_n sum_o wo[o] (y_no - t_no)^2 * scale over samples [0, n).
// part[wave][kParams] receives each wavefront's partial gradient; loss_part[wave] its loss.
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_grad(const float* P, const float* X, const float* Tg, int n,
                                                             float w0, float w1, float scale, float* part,
                                                             float* loss_part)
{
    __shared__ float sp[kParams];
    __shared__ float xs[kMlpWaves][kIn];
    __shared__ float a1s[kMlpWaves][kH];
    __shared__ float dz2s[kMlpWaves][kH];
    load_params(P, sp);
    const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
    const int gw = blockIdx.x * kMlpWaves + w, nw = gridDim.x * kMlpWaves;
    float gW2[kH], gW1[kIn], gW3[kOut] = {0.f, 0.f}, gb1 = 0.f, gb2 = 0.f, gb3[kOut] = {0.f, 0.f}, loss = 0.f;
#pragma unroll
    for (int k = 0; k < kH; ++k) gW2[k] = 0.f;
#pragma unroll
    for (int k = 0; k < kIn; ++k) gW1[k] = 0.f;
    for (int smp = gw; smp < n; smp += nw) {
        if (j < kIn) xs[w][j] = X[(size_t)smp * kIn + j];
        __builtin_amdgcn_wave_barrier();
        float z1, z2, a2, y[2];
        forward_one(sp, xs[w], a1s[w], z1, z2, a2, y);
        const float e0 = y[0] - Tg[(size_t)smp * 2], e1 = y[1] - Tg[(size_t)smp * 2 + 1];
        loss += scale * (w0 * e0 * e0 + w1 * e1 * e1);
        const float dy0 = 2.f * scale * w0 * e0, dy1 = 2.f * scale * w1 * e1;
        gW3[0] += dy0 * a2;
        gW3[1] += dy1 * a2;
        gb3[0] += dy0;
        gb3[1] += dy1;
        const float dz2 = (sp[oW3 + j] * dy0 + sp[oW3 + kH + j] * dy1) * gelu_d(z2);
        gb2 += dz2;
#pragma unroll
        for (int k = 0; k < kH; ++k) gW2[k] += dz2 * a1s[w][k];
        dz2s[w][j] = dz2;
        __builtin_amdgcn_wave_barrier();
        float da1 = 0.f;
#pragma unroll 8
        for (int r = 0; r < kH; ++r) da1 += sp[oW2 + r * kH + j] * dz2s[w][r];
        const float dz1 = da1 * gelu_d(z1);
        gb1 += dz1;
#pragma unroll
        for (int k = 0; k < kIn; ++k) gW1[k] += dz1 * xs[w][k];
        __builtin_amdgcn_wave_barrier();
    }
    float* g = part + (size_t)gw * kParams;
#pragma unroll
    for (int k = 0; k < kIn; ++k) g[oW1 + j * kIn + k] = gW1[k];
    g[oB1 + j] = gb1;
#pragma unroll
    for (int k = 0; k < kH; ++k) g[oW2 + j * kH + k] = gW2[k];
    g[oB2 + j] = gb2;
    g[oW3 + j] = gW3[0];
    g[oW3 + kH + j] = gW3[1];
    // output biases and the loss: lane sums are identical across lanes
    if (j == 0) { g[oB3] = gb3[0]; g[oB3 + 1] = gb3[1]; loss_part[gw] = loss; }
}

// grad[i] = sum over wavefronts (in order) of part[w][i]; loss = sum of loss_part.
__global__ void k_mlp_reduce(const float* part, const float* loss_part, int nw, float* grad, float* loss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) {
        float s = 0.f;
        for (int w = 0; w < nw; ++w) s += part[(size_t)w * kParams + i];
        grad[i] = s;
    }
    if (i == 0) {
        float s = 0.f;
        for (int w = 0; w < nw; ++w) s += loss_part[w];
        *loss = s;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 masters; step t starts at 1.
__global__ void k_mlp_adam(float* P, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                           int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t));
    const float vh = vi / (1.f - powf(b2, (float)t));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- host
struct Mlp {
    float* P = nullptr;        // FP32 masters
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* part = nullptr;
    float* loss_part = nullptr;
    float* loss = nullptr;
    float* X = nullptr;        // staging: features / targets / outputs
    float* Tg = nullptr;
    float* Y = nullptr;
    int cap = 0;               // samples the staging buffers hold
    int nw = 0;                // wavefronts of k_mlp_grad
    int adam_t = 0;
    float mean[kIn], stdv[kIn];
    bool ready = false;
};

static int ensure_cap(Mlp* M, int n)
{
    if (n <= M->cap) return 0;
    hipFree(M->X); hipFree(M->Tg); hipFree(M->Y);
    const int c = std::max(n, 1024);
    if (hipMalloc(&M->X, sizeof(float) * (size_t)c * kIn) != hipSuccess ||
        hipMalloc(&M->Tg, sizeof(float) * (size_t)c * 2) != hipSuccess ||
        hipMalloc(&M->Y, sizeof(float) * (size_t)c * 2) != hipSuccess)
        return 1;
    M->cap = c;
    return 0;
}

Mlp* mlp_create()
{
    Mlp* M = new Mlp();
    M->nw = 64 * kMlpWaves;
    bool ok = hipMalloc(&M->P, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&M->m, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&M->v, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&M->grad, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&M->part, sizeof(float) * (size_t)M->nw * kParams) == hipSuccess &&
              hipMalloc(&M->loss_part, sizeof(float) * M->nw) == hipSuccess &&
              hipMalloc(&M->loss, sizeof(float)) == hipSuccess;
    if (!ok) { delete M; return nullptr; }
    hipMemset(M->m, 0, sizeof(float) * kParams);
    hipMemset(M->v, 0, sizeof(float) * kParams);
    for (int k = 0; k < kIn; ++k) { M->mean[k] = 0.f; M->stdv[k] = 1.f; }
    return M;
}

void mlp_free(Mlp* M)
{
    if (!M) return;
    for (float* p : {M->P, M->m, M->v, M->grad, M->part, M->loss_part, M->loss, M->X, M->Tg, M->Y}) hipFree(p);
    delete M;
}

int mlp_set(Mlp* M, const float* params, const float* mean, const float* stdv)
{
    if (hipMemcpy(M->P, params, sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (mean) std::memcpy(M->mean, mean, sizeof(M->mean));
    if (stdv) std::memcpy(M->stdv, stdv, sizeof(M->stdv));
    hipMemset(M->m, 0, sizeof(float) * kParams);
    hipMemset(M->v, 0, sizeof(float) * kParams);
    M->adam_t = 0;
    M->ready = true;
    return 0;
}

bool mlp_ready(const Mlp* M) { return M && M->ready; }

// Raw features in, raw outputs out (standardisation of inputs with the stored mean/std).
// Standardised macro-state features X[s][kStat0 + k] = ((float)stats[s][k] - mean) / stdv,
// the arithmetic mlp_infer does on the host (FP64 -> FP32 cast, FP32 subtract and
// correctly rounded divide), so both give the network the same input bit for bit.
constexpr int kStat0 = 10, kNStat = 6;
struct StatNorm { float mean[kNStat], stdv[kNStat]; };
__global__ void k_feat_stats(float* X, const double* stats, int n, StatNorm sn)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
#pragma unroll
    for (int k = 0; k < kNStat; ++k)
        X[(size_t)s * kIn + kStat0 + k] = ((float)stats[(size_t)s * kNStat + k] - sn.mean[k]) / sn.stdv[k];
}

// Inference without a host round trip of its own: features 0..9 from the host (feats,
// standardised here), 10..15 from the macro-state summary already on the device
// (stats, n x 6 FP64), output left in mlp_output(M) ([n][2]) for the caller's next
// readback.
int mlp_infer_device(Mlp* M, const float* feats, int n, const double* stats, hipStream_t st)
{
    if (n <= 0) return 0;
    if (ensure_cap(M, n)) return 2;
    std::vector<float> xs((size_t)n * kIn, 0.0f);
    for (int s = 0; s < n; ++s)
        for (int k = 0; k < kStat0; ++k) xs[(size_t)s * kIn + k] = (feats[(size_t)s * kIn + k] - M->mean[k]) / M->stdv[k];
    if (hipMemcpyAsync(M->X, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    StatNorm sn;
    for (int k = 0; k < kNStat; ++k) { sn.mean[k] = M->mean[kStat0 + k]; sn.stdv[k] = M->stdv[kStat0 + k]; }
    hipLaunchKernelGGL(k_feat_stats, dim3((n + 63) / 64), dim3(64), 0, st, M->X, stats, n, sn);
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::min(64, (n + kMlpWaves - 1) / kMlpWaves)), dim3(64 * kMlpWaves), 0, st,
                       M->P, M->X, n, M->Y);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

const float* mlp_output(const Mlp* M) { return M->Y; }

int mlp_infer(Mlp* M, const float* feats, int n, float* out, hipStream_t st)
{
    if (n <= 0) return 0;
    if (ensure_cap(M, n)) return 2;
    std::vector<float> xs((size_t)n * kIn);
    for (int s = 0; s < n; ++s)
        for (int k = 0; k < kIn; ++k) xs[(size_t)s * kIn + k] = (feats[(size_t)s * kIn + k] - M->mean[k]) / M->stdv[k];
    if (hipMemcpyAsync(M->X, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::min(64, (n + kMlpWaves - 1) / kMlpWaves)), dim3(64 * kMlpWaves), 0, st,
                       M->P, M->X, n, M->Y);
    if (hipMemcpyAsync(out, M->Y, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

}  // namespace tsr

using namespace tsr;

extern "C" {

int tsr_mlp_param_count(void) { return kParams; }

void* tsr_mlp_create(void) { return mlp_create(); }
void tsr_mlp_free(void* h) { mlp_free(static_cast<Mlp*>(h)); }

int tsr_mlp_set_params(void* h, const float* params)
{
    Mlp* M = static_cast<Mlp*>(h);
    return hipMemcpy(M->P, params, sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess ? 0 : 2;
}

int tsr_mlp_get_params(void* h, float* params)
{
    Mlp* M = static_cast<Mlp*>(h);
    return hipMemcpy(params, M->P, sizeof(float) * kParams, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
}

// out[n][2] for already standardised inputs x[n][16].
int tsr_mlp_forward(void* h, const float* x, int n, float* out)
{
    Mlp* M = static_cast<Mlp*>(h);
    if (ensure_cap(M, n)) return 2;
    if (hipMemcpy(M->X, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::max(1, std::min(256, (n + kMlpWaves - 1) / kMlpWaves))),
                       dim3(64 * kMlpWaves), 0, nullptr, M->P, M->X, n, M->Y);
    return hipMemcpy(out, M->Y, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
}

// grad[kParams] and loss of L = scale * sum_n (w0 e0^2 + w1 e1^2) over the given samples.
int tsr_mlp_grad(void* h, const float* x, const float* t, int n, float w0, float w1, float scale, float* grad,
                 float* loss)
{
    Mlp* M = static_cast<Mlp*>(h);
    if (ensure_cap(M, n)) return 2;
    if (hipMemcpy(M->X, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (hipMemcpy(M->Tg, t, sizeof(float) * (size_t)n * 2, hipMemcpyHostToDevice) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_mlp_grad, dim3(M->nw / kMlpWaves), dim3(64 * kMlpWaves), 0, nullptr, M->P, M->X, M->Tg, n, w0,
                       w1, scale, M->part, M->loss_part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 255) / 256), dim3(256), 0, nullptr, M->part, M->loss_part, M->nw,
                       M->grad, M->loss);
    if (grad && hipMemcpy(grad, M->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (loss && hipMemcpy(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return 2;
    return 0;
}

// One Adam step with the gradient of the last tsr_mlp_grad call.
int tsr_mlp_adam(void* h, float lr)
{
    Mlp* M = static_cast<Mlp*>(h);
    M->adam_t += 1;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, nullptr, M->P, M->grad, M->m, M->v, lr,
                       0.9f, 0.999f, 1e-8f, M->adam_t);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// Mini-batch training on the device: x[n][16] (standardised), t[n][2]; epochs passes
// over a seeded permutation; loss_hist[epoch] = mean loss of that epoch.
int tsr_mlp_train(void* h, const float* x, const float* t, int n, int epochs, int batch, float lr, float w0, float w1,
                  unsigned long long seed, float* loss_hist)
{
    Mlp* M = static_cast<Mlp*>(h);
    if (n <= 0 || batch <= 0) return 1;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::vector<float> xb((size_t)batch * kIn), tb((size_t)batch * 2);
    unsigned long long s = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(rnd() % (unsigned long long)(i + 1))]);
        double acc = 0.0;
        for (int b0 = 0; b0 < n; b0 += batch) {
            const int nb = std::min(batch, n - b0);
            for (int r = 0; r < nb; ++r) {
                std::memcpy(&xb[(size_t)r * kIn], &x[(size_t)perm[b0 + r] * kIn], sizeof(float) * kIn);
                tb[(size_t)r * 2] = t[(size_t)perm[b0 + r] * 2];
                tb[(size_t)r * 2 + 1] = t[(size_t)perm[b0 + r] * 2 + 1];
            }
            float loss = 0.f;
            if (tsr_mlp_grad(h, xb.data(), tb.data(), nb, w0, w1, 1.0f / nb, nullptr, &loss)) return 2;
            if (tsr_mlp_adam(h, lr)) return 2;
            acc += (double)loss * nb;
        }
        if (loss_hist) loss_hist[ep] = (float)(acc / n);
    }
    return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

}  // extern "C"
