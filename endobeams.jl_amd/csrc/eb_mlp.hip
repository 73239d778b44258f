// eb_mlp.hip — the learned Newton warm start's network: a fully connected tanh MLP in FP32 with
// hand-written inference, back-propagation and Adam kernels (no library GEMM, no autograd).
//
// Layout: parameters are one FP32 vector; layer l holds W_l (out×in, row-major) then b_l (out).
// Activations of a minibatch are row-major [rows][width] per layer.
//
// Kernels (gfx950, wave64):
//   forward  — one workgroup per 32 rows on the FP32 matrix cores: the input tile sits in LDS and
//              wavefront w computes outputs 32w .. 32w+31 of the 32 rows with
//              v_mfma_f32_32x32x2f32, two input features per instruction;
//   backward — one workgroup per 32 rows: ∂L/∂W and ∂L/∂b of the tile into per-tile partials
//              (summed afterwards in tile order, so training is deterministic), and ∂L/∂Z of the
//              layer below with the same wave-uniform weight scheme;
//   adam     — one lane per parameter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/endobeams_learned.h"
#include "eb_mlp.h"

namespace ebm {

constexpr int FWD_TILE = 32;
constexpr int BWD_TILE = 32;
constexpr int MAXW = 128;

typedef float v16f __attribute__((ext_vector_type(16)));

// Z = Hin Wᵀ + b for 32 rows, wavefront w taking output columns 32w .. 32w+31 (out ≤ 128: four
// wavefronts). v_mfma_f32_32x32x2f32 multiplies a 32×2 by a 2×32 block and accumulates 32×32:
// lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32], and holds column l % 32 of the result,
// rows 8⌊r/4⌋ + 4⌊l/32⌋ + r mod 4 in its register r.
__global__ void __launch_bounds__(256) mlp_forward_layer(int64_t rows, int in, int out, const float* __restrict__ W,
                                                         const float* __restrict__ b, const float* __restrict__ Hin,
                                                         float* __restrict__ Hout, int act) {
    __shared__ float X[FWD_TILE][MAXW + 1];
    const int64_t r0 = (int64_t)blockIdx.x * FWD_TILE;
    const int nr = (int)std::min<int64_t>(FWD_TILE, rows - r0);
    for (int i = threadIdx.x; i < FWD_TILE * in; i += blockDim.x) {
        const int r = i / in, k = i % in;
        X[r][k] = r < nr ? Hin[(r0 + r) * in + k] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (32 * wave >= out) return;   // wave-uniform; no barrier follows
    const int j = 32 * wave + (lane & 31);
    const int kh = lane >> 5;
    v16f acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = 0; k0 < in; k0 += 2) {
        const int k = k0 + kh;
        const float a = k < in ? X[lane & 31][k] : 0.0f;
        const float w = (k < in && j < out) ? W[(int64_t)j * in + k] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w, acc, 0, 0, 0);
    }
    if (j >= out) return;
    const float bj = b[j];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r / 4) + 4 * kh + (r % 4);
        if (row < nr) {
            const float z = acc[r] + bj;
            Hout[(r0 + row) * out + j] = act ? tanhf(z) : z;
        }
    }
}

// MSE loss over rows×out: dZ = 2(Y - T)/(rows·out); per-workgroup loss partials
__global__ void __launch_bounds__(256) mlp_loss(int64_t rows, int out, const float* __restrict__ Y,
                                                const float* __restrict__ T, float* __restrict__ dZ, float* partial) {
    __shared__ float sh[256];
    const int64_t n = rows * out;
    const float scale = 2.0f / (float)n;
    float s = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dZ[i] = scale * d;
        s += d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// one layer of back-propagation for a 32-row tile: partial[tile] = (∂W, ∂b) of the tile, and
// for a hidden input layer dZin = (dZout W) ⊙ (1 - Hin²)
__global__ void __launch_bounds__(256) mlp_backward_layer(int64_t rows, int in, int out, const float* __restrict__ W,
                                                          const float* __restrict__ Hin, const float* __restrict__ dZout,
                                                          float* __restrict__ partial, float* __restrict__ dZin,
                                                          int hidden_in) {
    __shared__ float H[BWD_TILE][MAXW + 1];
    __shared__ float D[BWD_TILE][MAXW + 1];
    const int64_t r0 = (int64_t)blockIdx.x * BWD_TILE;
    const int nr = (int)std::min<int64_t>(BWD_TILE, rows - r0);
    for (int i = threadIdx.x; i < BWD_TILE * in; i += blockDim.x) {
        const int r = i / in, k = i % in;
        H[r][k] = r < nr ? Hin[(r0 + r) * in + k] : 0.0f;
    }
    for (int i = threadIdx.x; i < BWD_TILE * out; i += blockDim.x) {
        const int r = i / out, j = i % out;
        D[r][j] = r < nr ? dZout[(r0 + r) * out + j] : 0.0f;
    }
    __syncthreads();
    const int np = out * in + out;
    float* P = partial + (int64_t)blockIdx.x * np;
    for (int p = threadIdx.x; p < out * in; p += blockDim.x) {
        const int j = p / in, k = p % in;
        float s = 0;
        for (int r = 0; r < BWD_TILE; ++r) s = fmaf(D[r][j], H[r][k], s);
        P[p] = s;
    }
    for (int j = threadIdx.x; j < out; j += blockDim.x) {
        float s = 0;
        for (int r = 0; r < BWD_TILE; ++r) s += D[r][j];
        P[out * in + j] = s;
    }
    if (!hidden_in) return;
    // dH[r][k] = Σ_j D[r][j] W[j][k]: a wavefront per k, lanes over 2 x 32 rows (half-waves per k)
    const int lane = threadIdx.x & 63;
    const int r = lane & (BWD_TILE - 1);
    const int half = lane >> 5;
    const int wave = threadIdx.x >> 6;
    for (int k2 = 2 * wave; k2 < in; k2 += 2 * (blockDim.x >> 6)) {
        const int k = k2 + half;
        if (k >= in) continue;
        float s = 0;
        for (int j = 0; j < out; ++j) s = fmaf(D[r][j], W[(int64_t)j * in + k], s);
        const float h = H[r][k];
        if (r < nr) dZin[(r0 + r) * in + k] = s * (1.0f - h * h);
    }
}

// grad[p] = Σ over tiles of partial[tile][p], in tile order
__global__ void __launch_bounds__(256) mlp_reduce(int ntiles, int np, const float* __restrict__ partial,
                                                  float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float s = 0;
    for (int t = 0; t < ntiles; ++t) s += partial[(int64_t)t * np + p];
    grad[p] = s;
}

__global__ void __launch_bounds__(256) mlp_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                                                float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                float b2, float c1, float c2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(256) mlp_gather_rows(int64_t rows, int width, const int32_t* __restrict__ idx,
                                                       const float* __restrict__ src, float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * width) return;
    const int64_t r = i / width;
    dst[i] = src[(int64_t)idx[r] * width + i % width];
}

}  // namespace ebm

using namespace ebm;

struct eb_mlp {
    std::string err;
    hipStream_t stream = nullptr;
    std::vector<int> width;           // n_layers + 1
    std::vector<int64_t> off;         // parameter offset of each layer's W (b follows)
    int64_t n_params = 0;
    std::vector<void*> allocs;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    // minibatch workspace
    int64_t cap = 0;                   // rows
    std::vector<float*> act;           // per layer output (n_layers), plus the input copy
    float *x = nullptr, *t = nullptr, *dz0 = nullptr, *dz1 = nullptr, *partial = nullptr, *loss_part = nullptr;
    int32_t* idx = nullptr;
    int64_t partial_cap = 0;
    // dataset
    float *dX = nullptr, *dT = nullptr;
    int64_t n_data = 0, data_cap = 0;
    int adam_t = 0;
};

namespace {

#define MLP_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            mlp->err = std::string(#call) + ": " + hipGetErrorString(_e);                              \
            return EB_ERR_DEVICE;                                                                      \
        }                                                                                              \
    } while (0)

template <typename T> int mlp_alloc(eb_mlp* mlp, T*& p, int64_t count) {
    void* q = nullptr;
    MLP_HIP(hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    MLP_HIP(hipMemsetAsync(q, 0, sizeof(T) * (size_t)std::max<int64_t>(count, 1), mlp->stream));
    mlp->allocs.push_back(q);
    p = (T*)q;
    return EB_OK;
}

int layers(const eb_mlp* mlp) { return (int)mlp->width.size() - 1; }

// minibatch workspace for `rows` rows
int ensure_cap(eb_mlp* mlp, int64_t rows) {
    if (rows <= mlp->cap) return EB_OK;
    const int L = layers(mlp);
    int maxw = 0;
    for (int wd : mlp->width) maxw = std::max(maxw, wd);
    mlp->act.assign(L, nullptr);
    for (int l = 0; l < L; ++l)
        if (mlp_alloc(mlp, mlp->act[l], rows * mlp->width[l + 1])) return EB_ERR_DEVICE;
    if (mlp_alloc(mlp, mlp->x, rows * mlp->width[0]) || mlp_alloc(mlp, mlp->t, rows * mlp->width[L]) ||
        mlp_alloc(mlp, mlp->dz0, rows * maxw) || mlp_alloc(mlp, mlp->dz1, rows * maxw) ||
        mlp_alloc(mlp, mlp->idx, rows) || mlp_alloc(mlp, mlp->loss_part, 1024))
        return EB_ERR_DEVICE;
    int64_t np_max = 0;
    for (int l = 0; l < L; ++l) np_max = std::max<int64_t>(np_max, (int64_t)mlp->width[l] * mlp->width[l + 1] + mlp->width[l + 1]);
    const int64_t tiles = (rows + BWD_TILE - 1) / BWD_TILE;
    if (mlp_alloc(mlp, mlp->partial, tiles * np_max)) return EB_ERR_DEVICE;
    // the clears above run on the network's stream; inference may use another stream
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    mlp->cap = rows;
    return EB_OK;
}

int forward_dev(eb_mlp* mlp, const float* X, int64_t rows, float* const* act, hipStream_t stream) {
    const int L = layers(mlp);
    const float* in = X;
    const unsigned grid = (unsigned)((rows + FWD_TILE - 1) / FWD_TILE);
    for (int l = 0; l < L; ++l) {
        const float* W = mlp->w + mlp->off[l];
        const float* b = W + (int64_t)mlp->width[l] * mlp->width[l + 1];
        hipLaunchKernelGGL(mlp_forward_layer, dim3(grid), dim3(256), 0, stream, rows, mlp->width[l],
                           mlp->width[l + 1], W, b, in, act[l], l + 1 < L ? 1 : 0);
        in = act[l];
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return EB_ERR_DEVICE;
    return EB_OK;
}

// forward + loss + backward on the rows in mlp->x / mlp->t; gradient into mlp->g
int loss_grad_dev(eb_mlp* mlp, int64_t rows, float* loss_out) {
    const int L = layers(mlp);
    if (forward_dev(mlp, mlp->x, rows, mlp->act.data(), mlp->stream)) {
        mlp->err = "forward pass launch failed";
        return EB_ERR_DEVICE;
    }
    const int outw = mlp->width[L];
    const int lgrid = (int)std::min<int64_t>(1024, (rows * outw + 255) / 256);
    hipLaunchKernelGGL(mlp_loss, dim3(lgrid), dim3(256), 0, mlp->stream, rows, outw, mlp->act[L - 1], mlp->t,
                       mlp->dz0, mlp->loss_part);
    const unsigned tiles = (unsigned)((rows + BWD_TILE - 1) / BWD_TILE);
    float* dzo = mlp->dz0;
    float* dzi = mlp->dz1;
    for (int l = L - 1; l >= 0; --l) {
        const int in = mlp->width[l], out = mlp->width[l + 1];
        const float* W = mlp->w + mlp->off[l];
        const float* Hin = l > 0 ? mlp->act[l - 1] : mlp->x;
        const int np = out * in + out;
        hipLaunchKernelGGL(mlp_backward_layer, dim3(tiles), dim3(256), 0, mlp->stream, rows, in, out, W, Hin, dzo,
                           mlp->partial, dzi, l > 0 ? 1 : 0);
        hipLaunchKernelGGL(mlp_reduce, dim3((np + 255) / 256), dim3(256), 0, mlp->stream, (int)tiles, np, mlp->partial,
                           mlp->g + mlp->off[l]);
        std::swap(dzo, dzi);
    }
    MLP_HIP(hipGetLastError());
    if (loss_out) {
        std::vector<float> lp(lgrid);
        MLP_HIP(hipMemcpyAsync(lp.data(), mlp->loss_part, sizeof(float) * lgrid, hipMemcpyDeviceToHost, mlp->stream));
        MLP_HIP(hipStreamSynchronize(mlp->stream));
        double s = 0;
        for (float q : lp) s += q;
        *loss_out = (float)(s / (double)(rows * outw));
    }
    return EB_OK;
}

}  // namespace

namespace ebm {

int mlp_infer(eb_mlp* mlp, const float* X_dev, int64_t rows, hipStream_t stream, float* const* act,
              const float** Y_dev) {
    const int st = forward_dev(mlp, X_dev, rows, act, stream);
    *Y_dev = act[layers(mlp) - 1];
    return st;
}

int mlp_layers(const eb_mlp* mlp) { return layers(mlp); }
int mlp_layer_width(const eb_mlp* mlp, int l) { return mlp->width[l + 1]; }

int mlp_width(const eb_mlp* mlp, int which) { return which == 0 ? mlp->width.front() : mlp->width.back(); }

}  // namespace ebm

extern "C" {

int eb_mlp_create(const int32_t* widths, int32_t n_widths, uint64_t seed, eb_mlp** out) {
    if (!widths || !out || n_widths < 2) return EB_ERR_ARG;
    for (int i = 0; i < n_widths; ++i)
        if (widths[i] < 1 || widths[i] > MAXW) return EB_ERR_ARG;
    eb_mlp* mlp = new (std::nothrow) eb_mlp();
    if (!mlp) return EB_ERR_ARG;
    auto bail = [&](int c) { eb_mlp_destroy(mlp); return c; };
    if (hipStreamCreateWithFlags(&mlp->stream, hipStreamNonBlocking) != hipSuccess) return bail(EB_ERR_DEVICE);
    mlp->width.assign(widths, widths + n_widths);
    int64_t n = 0;
    for (int l = 0; l + 1 < n_widths; ++l) {
        mlp->off.push_back(n);
        n += (int64_t)widths[l] * widths[l + 1] + widths[l + 1];
    }
    mlp->n_params = n;
    if (mlp_alloc(mlp, mlp->w, n) || mlp_alloc(mlp, mlp->g, n) || mlp_alloc(mlp, mlp->m, n) || mlp_alloc(mlp, mlp->v, n))
        return bail(EB_ERR_DEVICE);
    // Glorot-uniform weights, zero biases
    std::vector<float> h((size_t)n, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l + 1 < n_widths; ++l) {
        const float a = std::sqrt(6.0f / (float)(widths[l] + widths[l + 1]));
        std::uniform_real_distribution<float> U(-a, a);
        for (int64_t k = 0; k < (int64_t)widths[l] * widths[l + 1]; ++k) h[mlp->off[l] + k] = U(rng);
    }
    if (hipMemcpyAsync(mlp->w, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, mlp->stream) != hipSuccess ||
        hipStreamSynchronize(mlp->stream) != hipSuccess)
        return bail(EB_ERR_DEVICE);
    *out = mlp;
    return EB_OK;
}

int eb_mlp_destroy(eb_mlp* mlp) {
    if (!mlp) return EB_OK;
    if (mlp->stream) hipStreamSynchronize(mlp->stream);
    for (void* p : mlp->allocs) hipFree(p);
    if (mlp->stream) hipStreamDestroy(mlp->stream);
    delete mlp;
    return EB_OK;
}

const char* eb_mlp_last_error(const eb_mlp* mlp) { return mlp ? mlp->err.c_str() : "null network"; }

int eb_mlp_num_params(const eb_mlp* mlp, int64_t* n) {
    if (!mlp || !n) return EB_ERR_ARG;
    *n = mlp->n_params;
    return EB_OK;
}

int eb_mlp_get_params(eb_mlp* mlp, float* out, int64_t n) {
    if (!mlp || !out || n != mlp->n_params) return EB_ERR_ARG;
    MLP_HIP(hipMemcpyAsync(out, mlp->w, sizeof(float) * n, hipMemcpyDeviceToHost, mlp->stream));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    return EB_OK;
}

int eb_mlp_set_params(eb_mlp* mlp, const float* in, int64_t n) {
    if (!mlp || !in || n != mlp->n_params) return EB_ERR_ARG;
    MLP_HIP(hipMemcpyAsync(mlp->w, in, sizeof(float) * n, hipMemcpyHostToDevice, mlp->stream));
    MLP_HIP(hipMemsetAsync(mlp->m, 0, sizeof(float) * n, mlp->stream));
    MLP_HIP(hipMemsetAsync(mlp->v, 0, sizeof(float) * n, mlp->stream));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    mlp->adam_t = 0;
    return EB_OK;
}

int eb_mlp_forward(eb_mlp* mlp, const float* X, int64_t rows, float* Y) {
    if (!mlp || !X || !Y || rows < 1) return EB_ERR_ARG;
    if (ensure_cap(mlp, rows)) return EB_ERR_DEVICE;
    MLP_HIP(hipMemcpyAsync(mlp->x, X, sizeof(float) * rows * mlp->width[0], hipMemcpyHostToDevice, mlp->stream));
    if (forward_dev(mlp, mlp->x, rows, mlp->act.data(), mlp->stream)) {
        mlp->err = "forward pass launch failed";
        return EB_ERR_DEVICE;
    }
    MLP_HIP(hipMemcpyAsync(Y, mlp->act[layers(mlp) - 1], sizeof(float) * rows * mlp->width.back(),
                           hipMemcpyDeviceToHost, mlp->stream));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    return EB_OK;
}

int eb_mlp_loss_gradient(eb_mlp* mlp, const float* X, const float* T, int64_t rows, float* loss, float* grad) {
    if (!mlp || !X || !T || !loss || !grad || rows < 1) return EB_ERR_ARG;
    if (ensure_cap(mlp, rows)) return EB_ERR_DEVICE;
    MLP_HIP(hipMemcpyAsync(mlp->x, X, sizeof(float) * rows * mlp->width[0], hipMemcpyHostToDevice, mlp->stream));
    MLP_HIP(hipMemcpyAsync(mlp->t, T, sizeof(float) * rows * mlp->width.back(), hipMemcpyHostToDevice, mlp->stream));
    if (loss_grad_dev(mlp, rows, loss)) return EB_ERR_DEVICE;
    MLP_HIP(hipMemcpyAsync(grad, mlp->g, sizeof(float) * mlp->n_params, hipMemcpyDeviceToHost, mlp->stream));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    return EB_OK;
}

int eb_mlp_set_dataset(eb_mlp* mlp, const float* X, const float* T, int64_t rows) {
    if (!mlp || !X || !T || rows < 1) return EB_ERR_ARG;
    if (rows > mlp->data_cap) {
        if (mlp_alloc(mlp, mlp->dX, rows * mlp->width[0]) || mlp_alloc(mlp, mlp->dT, rows * mlp->width.back()))
            return EB_ERR_DEVICE;
        mlp->data_cap = rows;
    }
    MLP_HIP(hipMemcpyAsync(mlp->dX, X, sizeof(float) * rows * mlp->width[0], hipMemcpyHostToDevice, mlp->stream));
    MLP_HIP(hipMemcpyAsync(mlp->dT, T, sizeof(float) * rows * mlp->width.back(), hipMemcpyHostToDevice, mlp->stream));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    mlp->n_data = rows;
    return EB_OK;
}

int eb_mlp_train(eb_mlp* mlp, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss_per_epoch) {
    if (!mlp || epochs < 1 || batch < 1 || mlp->n_data < 1) return EB_ERR_ARG;
    const int64_t B = std::min<int64_t>(batch, mlp->n_data);
    if (ensure_cap(mlp, B)) return EB_ERR_DEVICE;
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm((size_t)mlp->n_data);
    for (int64_t i = 0; i < mlp->n_data; ++i) perm[i] = (int32_t)i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int in = mlp->width[0], outw = mlp->width.back();
    for (int ep = 0; ep < epochs; ++ep) {
        MLP_HIP(hipStreamSynchronize(mlp->stream));   // the index copies of the last epoch are done
        std::shuffle(perm.begin(), perm.end(), rng);
        double loss_sum = 0;
        int64_t nb = 0;
        for (int64_t s = 0; s + B <= mlp->n_data; s += B, ++nb) {
            MLP_HIP(hipMemcpyAsync(mlp->idx, perm.data() + s, sizeof(int32_t) * B, hipMemcpyHostToDevice, mlp->stream));
            hipLaunchKernelGGL(mlp_gather_rows, dim3((unsigned)((B * in + 255) / 256)), dim3(256), 0, mlp->stream, B, in,
                               mlp->idx, mlp->dX, mlp->x);
            hipLaunchKernelGGL(mlp_gather_rows, dim3((unsigned)((B * outw + 255) / 256)), dim3(256), 0, mlp->stream, B,
                               outw, mlp->idx, mlp->dT, mlp->t);
            float loss = 0;
            if (loss_grad_dev(mlp, B, (nb % 16 == 0) ? &loss : nullptr)) return EB_ERR_DEVICE;
            if (nb % 16 == 0) loss_sum += loss;
            mlp->adam_t += 1;
            const float c1 = 1.0f - std::pow(b1, (float)mlp->adam_t), c2 = 1.0f - std::pow(b2, (float)mlp->adam_t);
            hipLaunchKernelGGL(mlp_adam, dim3((unsigned)((mlp->n_params + 255) / 256)), dim3(256), 0, mlp->stream,
                               mlp->n_params, mlp->w, mlp->g, mlp->m, mlp->v, lr, b1, b2, c1, c2, eps);
            MLP_HIP(hipGetLastError());
        }
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(loss_sum / std::max<int64_t>(1, (nb + 15) / 16));
    }
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    return EB_OK;
}

}  // extern "C"
