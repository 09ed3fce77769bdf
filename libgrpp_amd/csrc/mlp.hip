// Radial-grid level predictor: a 12 -> 128 -> 128 -> 128 -> 3 multilayer perceptron with SiLU
// activations, trained (forward, backward, Adam) and evaluated by the kernels of this file.
//
// Training is FP32; inference reads BF16 copies of the weights (FP32 accumulation).  Features are
// standardised inside the first layer with a per-feature mean and scale held by the model.
//
// Layout: parameters are one FP32 vector [W1 (12x128) b1 | W2 (128x128) b2 | W3 (128x128) b3 |
// W4 (128x3) b4], W stored input-major (W[i][o]).  A workgroup of 256 threads processes a tile of
// MLP_TILE samples with the activations of the tile in LDS; the backward pass writes per-tile
// gradient partials that a second kernel sums in tile order, so a gradient is deterministic.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "engine.h"

constexpr int NIN = 12, NH = 128, NOUT = 3;
constexpr int MLP_TILE = 32;   // samples per workgroup
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH * NOUT;
constexpr int NPARAM = OFF_B4 + NOUT;

#define MLPCHK(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            eng_set_error("HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__);  \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// one dense layer on a tile: out[s][o] = b[o] + sum_i in[s][i] W[i][o] (LDS in, LDS out)
template <class WT>
__device__ __forceinline__ void dense(const float *in, int nin, const WT *W, const WT *b, float *out, int nout,
                                      int ns)
{
    for (int e = threadIdx.x; e < ns * nout; e += blockDim.x) {
        const int s = e / nout, o = e % nout;
        float acc = (float) b[o];
        for (int i = 0; i < nin; i++) acc += in[s * nin + i] * (float) W[i * nout + o];
        out[e] = acc;
    }
}

// forward pass of the samples idx[t0 .. t0+ns) (rows of X); Z holds the pre-activations of the
// three hidden layers per sample (training), Y the outputs
template <class WT>
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ X, const int *__restrict__ idx, int n,
                                                     const WT *__restrict__ P, const float *__restrict__ norm,
                                                     float *__restrict__ Z, float *__restrict__ Y)
{
    __shared__ float a0[MLP_TILE * NIN], a1[MLP_TILE * NH], a2[MLP_TILE * NH];
    const int t0 = blockIdx.x * MLP_TILE, ns = min(MLP_TILE, n - t0);
    for (int e = threadIdx.x; e < ns * NIN; e += blockDim.x) {
        const int s = e / NIN, f = e % NIN;
        const int row = idx ? idx[t0 + s] : t0 + s;
        a0[e] = (X[(size_t) row * NIN + f] - norm[f]) * norm[NIN + f];
    }
    __syncthreads();
    dense(a0, NIN, P + OFF_W1, P + OFF_B1, a1, NH, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) {
        if (Z) Z[((size_t) (t0 + e / NH) * 3 + 0) * NH + e % NH] = a1[e];
        a1[e] = silu(a1[e]);
    }
    __syncthreads();
    dense(a1, NH, P + OFF_W2, P + OFF_B2, a2, NH, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) {
        if (Z) Z[((size_t) (t0 + e / NH) * 3 + 1) * NH + e % NH] = a2[e];
        a2[e] = silu(a2[e]);
    }
    __syncthreads();
    dense(a2, NH, P + OFF_W3, P + OFF_B3, a1, NH, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) {
        if (Z) Z[((size_t) (t0 + e / NH) * 3 + 2) * NH + e % NH] = a1[e];
        a1[e] = silu(a1[e]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NOUT; e += blockDim.x) {
        const int s = e / NOUT, o = e % NOUT;
        float acc = (float) P[OFF_B4 + o];
        for (int i = 0; i < NH; i++) acc += a1[s * NH + i] * (float) P[OFF_W4 + i * NOUT + o];
        Y[(size_t) (t0 + s) * NOUT + o] = acc;
    }
}

// backward pass of the mean squared error L = (1/(n NOUT)) sum (Y - T)^2 over the n samples of
// the batch; writes this tile's gradient partial G[blockIdx.x][NPARAM]
__global__ __launch_bounds__(256) void k_mlp_backward(const float *__restrict__ X, const int *__restrict__ idx, int n,
                                                      const float *__restrict__ P, const float *__restrict__ norm,
                                                      const float *__restrict__ Z, const float *__restrict__ Y,
                                                      const float *__restrict__ T, float *__restrict__ G)
{
    __shared__ float in0[MLP_TILE * NIN];
    __shared__ float h[MLP_TILE * NH];    // activation feeding the current layer
    __shared__ float d[MLP_TILE * NH];    // gradient w.r.t. the current layer's output
    __shared__ float dn[MLP_TILE * NH];   // gradient w.r.t. the current layer's input
    __shared__ float dy[MLP_TILE * NOUT];
    const int t0 = blockIdx.x * MLP_TILE, ns = min(MLP_TILE, n - t0);
    float *g = G + (size_t) blockIdx.x * NPARAM;
    const float inv = 2.0f / (float) (n * NOUT);
    for (int e = threadIdx.x; e < ns * NOUT; e += blockDim.x) {
        const int s = e / NOUT, o = e % NOUT;
        const int row = idx ? idx[t0 + s] : t0 + s;
        dy[e] = inv * (Y[(size_t) (t0 + s) * NOUT + o] - T[(size_t) row * NOUT + o]);
    }
    for (int e = threadIdx.x; e < ns * NH; e += blockDim.x)
        h[e] = silu(Z[((size_t) (t0 + e / NH) * 3 + 2) * NH + e % NH]);
    __syncthreads();
    // layer 4
    for (int e = threadIdx.x; e < NH * NOUT + NOUT; e += blockDim.x) {
        float acc = 0.0f;
        if (e < NH * NOUT) {
            const int i = e / NOUT, o = e % NOUT;
            for (int s = 0; s < ns; s++) acc += h[s * NH + i] * dy[s * NOUT + o];
            g[OFF_W4 + e] = acc;
        } else {
            const int o = e - NH * NOUT;
            for (int s = 0; s < ns; s++) acc += dy[s * NOUT + o];
            g[OFF_B4 + o] = acc;
        }
    }
    for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) {
        const int s = e / NH, i = e % NH;
        float acc = 0.0f;
        for (int o = 0; o < NOUT; o++) acc += dy[s * NOUT + o] * P[OFF_W4 + i * NOUT + o];
        d[e] = acc * silu_grad(Z[((size_t) (t0 + s) * 3 + 2) * NH + i]);
    }
    __syncthreads();
    // layers 3 and 2 (hidden -> hidden)
    for (int layer = 2; layer >= 1; layer--) {
        const int offW = layer == 2 ? OFF_W3 : OFF_W2, offB = layer == 2 ? OFF_B3 : OFF_B2;
        for (int e = threadIdx.x; e < ns * NH; e += blockDim.x)
            h[e] = silu(Z[((size_t) (t0 + e / NH) * 3 + (layer - 1)) * NH + e % NH]);
        __syncthreads();
        for (int e = threadIdx.x; e < NH * NH + NH; e += blockDim.x) {
            float acc = 0.0f;
            if (e < NH * NH) {
                const int i = e / NH, o = e % NH;
                for (int s = 0; s < ns; s++) acc += h[s * NH + i] * d[s * NH + o];
                g[offW + e] = acc;
            } else {
                const int o = e - NH * NH;
                for (int s = 0; s < ns; s++) acc += d[s * NH + o];
                g[offB + o] = acc;
            }
        }
        for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) {
            const int s = e / NH, i = e % NH;
            float acc = 0.0f;
            for (int o = 0; o < NH; o++) acc += d[s * NH + o] * P[offW + i * NH + o];
            dn[e] = acc * silu_grad(Z[((size_t) (t0 + s) * 3 + (layer - 1)) * NH + i]);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ns * NH; e += blockDim.x) d[e] = dn[e];
        __syncthreads();
    }
    // layer 1 (input -> hidden)
    for (int e = threadIdx.x; e < ns * NIN; e += blockDim.x) {
        const int s = e / NIN, f = e % NIN;
        const int row = idx ? idx[t0 + s] : t0 + s;
        in0[e] = (X[(size_t) row * NIN + f] - norm[f]) * norm[NIN + f];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NIN * NH + NH; e += blockDim.x) {
        float acc = 0.0f;
        if (e < NIN * NH) {
            const int i = e / NH, o = e % NH;
            for (int s = 0; s < ns; s++) acc += in0[s * NIN + i] * d[s * NH + o];
            g[OFF_W1 + e] = acc;
        } else {
            const int o = e - NIN * NH;
            for (int s = 0; s < ns; s++) acc += d[s * NH + o];
            g[OFF_B1 + o] = acc;
        }
    }
}

// G_total[p] = sum over tiles in tile order of G[tile][p]
__global__ void k_mlp_reduce(const float *__restrict__ G, int ntile, float *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float acc = 0.0f;
    for (int t = 0; t < ntile; t++) acc += G[(size_t) t * NPARAM + p];
    out[p] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ void k_to_bf16(const float *__restrict__ P, __hip_bfloat16 *__restrict__ Pb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) Pb[p] = __float2bfloat16(P[p]);
}

// squared error per sample (for the loss of a batch)
__global__ void k_mlp_sqerr(const float *__restrict__ Y, const float *__restrict__ T, const int *__restrict__ idx,
                            int n, float *__restrict__ err)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int row = idx ? idx[s] : s;
    float acc = 0.0f;
    for (int o = 0; o < NOUT; o++) {
        const float dlt = Y[(size_t) s * NOUT + o] - T[(size_t) row * NOUT + o];
        acc += dlt * dlt;
    }
    err[s] = acc;
}

// ------------------------------------------------------------------ host side
struct libgrpp_mlp {
    float *P = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *norm = nullptr;
    __hip_bfloat16 *Pb = nullptr;
    long long step = 0;
    std::vector<float> norm_h = std::vector<float>(2 * NIN, 0.0f);
};

extern "C" int libgrpp_mlp_nparams(void) { return NPARAM; }

extern "C" int libgrpp_mlp_create(unsigned seed, libgrpp_mlp **out)
{
    *out = nullptr;
    if (!eng_device_ok()) return 1;
    libgrpp_mlp *M = new libgrpp_mlp();
    // He-style uniform initialisation, deterministic in the seed; biases zero
    std::vector<float> p(NPARAM, 0.0f);
    std::mt19937 rng(seed);
    auto fill = [&](int off, int fan_in, int count) {
        const float a = std::sqrt(6.0f / (float) fan_in);
        std::uniform_real_distribution<float> U(-a, a);
        for (int k = 0; k < count; k++) p[off + k] = U(rng);
    };
    fill(OFF_W1, NIN, NIN * NH);
    fill(OFF_W2, NH, NH * NH);
    fill(OFF_W3, NH, NH * NH);
    fill(OFF_W4, NH, NH * NOUT);
    for (int f = 0; f < NIN; f++) M->norm_h[NIN + f] = 1.0f;
    MLPCHK(hipMalloc(&M->P, NPARAM * sizeof(float)));
    MLPCHK(hipMalloc(&M->m, NPARAM * sizeof(float)));
    MLPCHK(hipMalloc(&M->v, NPARAM * sizeof(float)));
    MLPCHK(hipMalloc(&M->g, NPARAM * sizeof(float)));
    MLPCHK(hipMalloc(&M->norm, 2 * NIN * sizeof(float)));
    MLPCHK(hipMalloc(&M->Pb, NPARAM * sizeof(__hip_bfloat16)));
    MLPCHK(hipMemcpy(M->P, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
    MLPCHK(hipMemset(M->m, 0, NPARAM * sizeof(float)));
    MLPCHK(hipMemset(M->v, 0, NPARAM * sizeof(float)));
    MLPCHK(hipMemcpy(M->norm, M->norm_h.data(), 2 * NIN * sizeof(float), hipMemcpyHostToDevice));
    *out = M;
    return 0;
}

extern "C" void libgrpp_mlp_destroy(libgrpp_mlp *M)
{
    if (!M) return;
    (void) hipFree(M->P);
    (void) hipFree(M->m);
    (void) hipFree(M->v);
    (void) hipFree(M->g);
    (void) hipFree(M->norm);
    (void) hipFree(M->Pb);
    delete M;
}

// mean[12], scale[12]: inputs enter the network as (x - mean) * scale
extern "C" int libgrpp_mlp_set_normalisation(libgrpp_mlp *M, const float *mean, const float *scale)
{
    for (int f = 0; f < NIN; f++) {
        M->norm_h[f] = mean[f];
        M->norm_h[NIN + f] = scale[f];
    }
    MLPCHK(hipMemcpy(M->norm, M->norm_h.data(), 2 * NIN * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int libgrpp_mlp_params(libgrpp_mlp *M, float *get, const float *set)
{
    if (set) MLPCHK(hipMemcpy(M->P, set, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    if (get) MLPCHK(hipMemcpy(get, M->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

struct DevBuf {
    std::vector<void *> p;
    ~DevBuf()
    {
        for (void *x : p) (void) hipFree(x);
    }
    template <class T>
    T *get(size_t n)
    {
        void *x = nullptr;
        if (hipMalloc(&x, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) return nullptr;
        p.push_back(x);
        return (T *) x;
    }
};

// gradient of the batch loss over all n samples (one batch) -- used to check the kernels
extern "C" int libgrpp_mlp_gradient(libgrpp_mlp *M, int n, const float *X, const float *T, float *grad, float *loss)
{
    if (n <= 0) {
        eng_set_error("mlp_gradient: empty batch");
        return 2;
    }
    DevBuf b;
    float *dX = b.get<float>((size_t) n * NIN), *dT = b.get<float>((size_t) n * NOUT);
    float *dZ = b.get<float>((size_t) n * 3 * NH), *dY = b.get<float>((size_t) n * NOUT);
    const int ntile = (n + MLP_TILE - 1) / MLP_TILE;
    float *dG = b.get<float>((size_t) ntile * NPARAM), *dE = b.get<float>(n);
    if (!dX || !dT || !dZ || !dY || !dG || !dE) {
        eng_set_error("mlp_gradient: device allocation failed");
        return 1;
    }
    MLPCHK(hipMemcpy(dX, X, (size_t) n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MLPCHK(hipMemcpy(dT, T, (size_t) n * NOUT * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_forward<float>, dim3(ntile), dim3(256), 0, 0, dX, nullptr, n, M->P, M->norm, dZ, dY);
    hipLaunchKernelGGL(k_mlp_backward, dim3(ntile), dim3(256), 0, 0, dX, nullptr, n, M->P, M->norm, dZ, dY, dT, dG);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, dG, ntile, M->g);
    hipLaunchKernelGGL(k_mlp_sqerr, dim3((n + 255) / 256), dim3(256), 0, 0, dY, dT, nullptr, n, dE);
    MLPCHK(hipGetLastError());
    MLPCHK(hipMemcpy(grad, M->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> e(n);
    MLPCHK(hipMemcpy(e.data(), dE, n * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (float x : e) s += x;
    if (loss) *loss = (float) (s / ((double) n * NOUT));
    return 0;
}

// minibatch Adam over the n samples for the given epochs; the sample order of each epoch is a
// deterministic shuffle of the seed; *loss receives the mean loss of the last epoch
extern "C" int libgrpp_mlp_train(libgrpp_mlp *M, int n, const float *X, const float *T, int epochs, int batch,
                                 float lr, unsigned seed, float *loss)
{
    if (n <= 0 || batch <= 0 || epochs <= 0) {
        eng_set_error("mlp_train: empty data, batch or epoch count");
        return 2;
    }
    batch = std::min(batch, n);
    DevBuf b;
    float *dX = b.get<float>((size_t) n * NIN), *dT = b.get<float>((size_t) n * NOUT);
    int *dI = b.get<int>(n);
    float *dZ = b.get<float>((size_t) batch * 3 * NH), *dY = b.get<float>((size_t) batch * NOUT);
    const int ntile_max = (batch + MLP_TILE - 1) / MLP_TILE;
    float *dG = b.get<float>((size_t) ntile_max * NPARAM), *dE = b.get<float>(n);
    if (!dX || !dT || !dI || !dZ || !dY || !dG || !dE) {
        eng_set_error("mlp_train: device allocation failed");
        return 1;
    }
    MLPCHK(hipMemcpy(dX, X, (size_t) n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MLPCHK(hipMemcpy(dT, T, (size_t) n * NOUT * sizeof(float), hipMemcpyHostToDevice));
    std::vector<int> order(n);
    std::mt19937 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = 0; i < n; i++) order[i] = i;
        std::shuffle(order.begin(), order.end(), rng);
        MLPCHK(hipMemcpy(dI, order.data(), n * sizeof(int), hipMemcpyHostToDevice));
        double sum = 0.0;
        for (int s0 = 0; s0 < n; s0 += batch) {
            const int nb = std::min(batch, n - s0), ntile = (nb + MLP_TILE - 1) / MLP_TILE;
            hipLaunchKernelGGL(k_mlp_forward<float>, dim3(ntile), dim3(256), 0, 0, dX, dI + s0, nb, M->P, M->norm, dZ,
                               dY);
            hipLaunchKernelGGL(k_mlp_backward, dim3(ntile), dim3(256), 0, 0, dX, dI + s0, nb, M->P, M->norm, dZ, dY, dT,
                               dG);
            hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, dG, ntile, M->g);
            M->step++;
            const float c1 = 1.0f - std::pow(b1, (float) M->step), c2 = 1.0f - std::pow(b2, (float) M->step);
            hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, M->P, M->g, M->m, M->v, lr, b1,
                               b2, eps, c1, c2);
            if (ep == epochs - 1)
                hipLaunchKernelGGL(k_mlp_sqerr, dim3((nb + 255) / 256), dim3(256), 0, 0, dY, dT, dI + s0, nb, dE + s0);
        }
        MLPCHK(hipGetLastError());
        if (ep == epochs - 1) {
            std::vector<float> e(n);
            MLPCHK(hipMemcpy(e.data(), dE, n * sizeof(float), hipMemcpyDeviceToHost));
            for (float x : e) sum += x;
            last = sum / ((double) n * NOUT);
        }
    }
    MLPCHK(hipDeviceSynchronize());
    if (loss) *loss = (float) last;
    return 0;
}

// batched inference; bf16 != 0 evaluates with BF16 copies of the weights (FP32 accumulation)
extern "C" int libgrpp_mlp_predict(libgrpp_mlp *M, int n, const float *X, float *Y, int bf16)
{
    if (n <= 0) return 0;
    DevBuf b;
    float *dX = b.get<float>((size_t) n * NIN), *dY = b.get<float>((size_t) n * NOUT);
    if (!dX || !dY) {
        eng_set_error("mlp_predict: device allocation failed");
        return 1;
    }
    MLPCHK(hipMemcpy(dX, X, (size_t) n * NIN * sizeof(float), hipMemcpyHostToDevice));
    const int ntile = (n + MLP_TILE - 1) / MLP_TILE;
    if (bf16) {
        hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, M->P, M->Pb);
        hipLaunchKernelGGL(k_mlp_forward<__hip_bfloat16>, dim3(ntile), dim3(256), 0, 0, dX, nullptr, n, M->Pb, M->norm,
                           nullptr, dY);
    } else {
        hipLaunchKernelGGL(k_mlp_forward<float>, dim3(ntile), dim3(256), 0, 0, dX, nullptr, n, M->P, M->norm, nullptr,
                           dY);
    }
    MLPCHK(hipGetLastError());
    MLPCHK(hipMemcpy(Y, dY, (size_t) n * NOUT * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}
