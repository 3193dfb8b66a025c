// The learned schedule dispatcher: an FP32 multilayer perceptron
// 8 -> 64 -> 64 -> K (ReLU hidden layers) that maps a product's features to
// the predicted relative log-time of each of K candidate schedules; the
// dispatcher takes the argmin. Training (mean-squared error, Adam) and
// inference run in the HIP kernels of this file.
//
// Parameter vector layout (floats): mu[8] sd[8] (input normalization,
// x' = (x - mu) / sd, not trained) | W1[64][8] b1[64] | W2[64][64] b2[64] |
// W3[K][64] b3[K].
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_internal.h"
#include "ff_mlp.h"
#include "ff_runtime.h"

namespace {

constexpr int NI = FF_MLP_IN, NH = FF_MLP_HIDDEN;

struct Layout {
    int K;
    size_t mu, sd, W1, b1, W2, b2, W3, b3, total;
};

__host__ __device__ inline Layout layout(int K) {
    Layout l;
    l.K = K;
    l.mu = 0;
    l.sd = NI;
    l.W1 = 2 * NI;
    l.b1 = l.W1 + NH * NI;
    l.W2 = l.b1 + NH;
    l.b2 = l.W2 + NH * NH;
    l.W3 = l.b2 + NH;
    l.b3 = l.W3 + (size_t)K * NH;
    l.total = l.b3 + K;
    return l;
}

std::vector<float> g_params;  // host copy of the current parameters (empty: none)
int g_K = 0;
float *g_dparams = nullptr;   // device copy for inference (uploaded when stale)
bool g_dstale = true;

}  // namespace

// ------------------------------------------------------------------ kernels

// Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o]); X normalized on the fly
// when mu != nullptr.
__global__ void k_dense_fwd(const float *X, const float *W, const float *bias, float *Y, int B, int I, int O,
                            int relu, const float *mu, const float *sd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * O) return;
    const int b = t / O, o = t - b * O;
    const float *x = X + (size_t)b * I;
    const float *w = W + (size_t)o * I;
    float acc = bias[o];
    for (int i = 0; i < I; i++) {
        const float xi = mu ? (x[i] - mu[i]) / sd[i] : x[i];
        acc = fmaf(xi, w[i], acc);
    }
    Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// G[b][k] = 2 (Y[b][k] - T[b][k]) / (B K); loss[b] = sum_k (Y - T)^2 / K.
__global__ void k_mse_grad(const float *Y, const float *T, float *G, float *loss, int B, int K) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float s = 0;
    for (int k = 0; k < K; k++) {
        const float d = Y[(size_t)b * K + k] - T[(size_t)b * K + k];
        G[(size_t)b * K + k] = 2.0f * d / (float)(B * K);
        s = fmaf(d, d, s);
    }
    loss[b] = s / (float)K;
}

// GX[b][i] = (sum_o G[b][o] W[o][i]) * (H[b][i] > 0).
__global__ void k_dense_bwd_x(const float *G, const float *W, const float *H, float *GX, int B, int I, int O) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * I) return;
    const int b = t / I, i = t - b * I;
    float acc = 0;
    for (int o = 0; o < O; o++) acc = fmaf(G[(size_t)b * O + o], W[(size_t)o * I + i], acc);
    GX[t] = H[t] > 0.0f ? acc : 0.0f;
}

// dW[o][i] = sum_b G[b][o] X'[b][i] (i < I), dbias[o] = sum_b G[b][o]
// (thread i == I); X normalized on the fly when mu != nullptr.
__global__ void k_dense_grad(const float *G, const float *X, float *dW, float *db, int B, int I, int O,
                             const float *mu, const float *sd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= O * (I + 1)) return;
    const int o = t / (I + 1), i = t - o * (I + 1);
    float acc = 0;
    if (i < I) {
        for (int b = 0; b < B; b++) {
            const float x = X[(size_t)b * I + i];
            acc = fmaf(G[(size_t)b * O + o], mu ? (x - mu[i]) / sd[i] : x, acc);
        }
        dW[(size_t)o * I + i] = acc;
    } else {
        for (int b = 0; b < B; b++) acc += G[(size_t)b * O + o];
        db[o] = acc;
    }
}

// Adam (Kingma & Ba, 2015) on n parameters, bias-corrected step size lr_t.
__global__ void k_adam(float *P, const float *dP, float *m, float *v, int n, float lr_t, float b1, float b2,
                       float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dP[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// Rows idx[r] of src (width w) into dst row r.
__global__ void k_gather(const float *src, const int *idx, float *dst, int rows, int w) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * w) return;
    const int r = t / w, c = t - r * w;
    dst[t] = src[(size_t)idx[r] * w + c];
}

// Fused inference, one sample per thread, parameters staged in LDS: writes
// the K outputs (if Y) and the argmin over candidates allowed by mask bit k
// (if choice).
__global__ __launch_bounds__(256) void k_mlp_infer(const float *X, const float *P, int n, int K,
                                                   const uint32_t *mask, float *Y, int *choice, float margin,
                                                   int dflt) {
    extern __shared__ float sp[];
    const Layout l = layout(K);
    for (size_t i = threadIdx.x; i < l.total; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[NI], h1[NH];
    for (int i = 0; i < NI; i++) x[i] = (X[(size_t)s * NI + i] - sp[l.mu + i]) / sp[l.sd + i];
    for (int o = 0; o < NH; o++) {
        float acc = sp[l.b1 + o];
        for (int i = 0; i < NI; i++) acc = fmaf(x[i], sp[l.W1 + o * NI + i], acc);
        h1[o] = fmaxf(acc, 0.0f);
    }
    float best = INFINITY;
    int arg = -1;
    const uint32_t mk = mask ? mask[s] : 0xffffffffu;
    // layer 2 one unit at a time, folded straight into the K outputs
    float y[FF_MLP_MAXK];
    for (int k = 0; k < K; k++) y[k] = sp[l.b3 + k];
    for (int o = 0; o < NH; o++) {
        float acc = sp[l.b2 + o];
        for (int i = 0; i < NH; i++) acc = fmaf(h1[i], sp[l.W2 + o * NH + i], acc);
        const float h2 = fmaxf(acc, 0.0f);
        for (int k = 0; k < K; k++) y[k] = fmaf(h2, sp[l.W3 + k * NH + o], y[k]);
    }
    float ydef = INFINITY;
    for (int k = 0; k < K; k++) {
        if (Y) Y[(size_t)s * K + k] = y[k];
        if (((mk >> k) & 1) && y[k] < best) best = y[k], arg = k;
        if (k == dflt) ydef = y[k];
    }
    if (margin > 0 && dflt >= 0 && ((mk >> dflt) & 1) && ydef - best <= margin) arg = dflt;
    if (choice) choice[s] = arg;
}

// ------------------------------------------------------------------ host

static unsigned blocks_for(size_t n, unsigned tpb = 256) { return (unsigned)((n + tpb - 1) / tpb); }

// Device workspace of one training or gradient run.
struct MlpDev {
    Layout l;
    int B;
    float *P, *dP, *m, *v;
    float *X, *T, *H1, *H2, *Y, *G3, *G2, *G1, *loss;
    std::vector<void *> owned;

    float *alloc(size_t n) {
        void *p = ff_dalloc(n * sizeof(float));
        owned.push_back(p);
        return (float *)p;
    }
    MlpDev(int K, int B_) : l(layout(K)), B(B_) {
        P = alloc(l.total), dP = alloc(l.total), m = alloc(l.total), v = alloc(l.total);
        X = alloc((size_t)B * NI), T = alloc((size_t)B * K);
        H1 = alloc((size_t)B * NH), H2 = alloc((size_t)B * NH), Y = alloc((size_t)B * K);
        G3 = alloc((size_t)B * K), G2 = alloc((size_t)B * NH), G1 = alloc((size_t)B * NH), loss = alloc(B);
    }
    ~MlpDev() {
        for (void *p : owned) ff_dfree(p);
    }

    // forward + backward of the first nb rows of X / T; gradients into dP
    void step(int nb) {
        const int K = l.K;
        const float *mu = P + l.mu, *sd = P + l.sd;
        hipLaunchKernelGGL(k_dense_fwd, dim3(blocks_for((size_t)nb * NH)), dim3(256), 0, g_st, X, P + l.W1,
                           P + l.b1, H1, nb, NI, NH, 1, mu, sd);
        hipLaunchKernelGGL(k_dense_fwd, dim3(blocks_for((size_t)nb * NH)), dim3(256), 0, g_st, H1, P + l.W2,
                           P + l.b2, H2, nb, NH, NH, 1, nullptr, nullptr);
        hipLaunchKernelGGL(k_dense_fwd, dim3(blocks_for((size_t)nb * K)), dim3(256), 0, g_st, H2, P + l.W3,
                           P + l.b3, Y, nb, NH, K, 0, nullptr, nullptr);
        hipLaunchKernelGGL(k_mse_grad, dim3(blocks_for(nb)), dim3(256), 0, g_st, Y, T, G3, loss, nb, K);
        hipLaunchKernelGGL(k_dense_grad, dim3(blocks_for((size_t)K * (NH + 1))), dim3(256), 0, g_st, G3, H2,
                           dP + l.W3, dP + l.b3, nb, NH, K, nullptr, nullptr);
        hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks_for((size_t)nb * NH)), dim3(256), 0, g_st, G3, P + l.W3,
                           H2, G2, nb, NH, K);
        hipLaunchKernelGGL(k_dense_grad, dim3(blocks_for((size_t)NH * (NH + 1))), dim3(256), 0, g_st, G2, H1,
                           dP + l.W2, dP + l.b2, nb, NH, NH, nullptr, nullptr);
        hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks_for((size_t)nb * NH)), dim3(256), 0, g_st, G2, P + l.W2,
                           H1, G1, nb, NH, NH);
        hipLaunchKernelGGL(k_dense_grad, dim3(blocks_for((size_t)NH * (NI + 1))), dim3(256), 0, g_st, G1, X,
                           dP + l.W1, dP + l.b1, nb, NI, NH, mu, sd);
        FF_CHECK(hipGetLastError());
    }
};

static int check_k(int K) {
    if (K < 1 || K > FF_MLP_MAXK) {
        ff_set_err("candidate count %d outside 1..%d", K, FF_MLP_MAXK);
        return FF_ERR_ARG;
    }
    return FF_OK;
}

extern "C" size_t ff_mlp_param_count(int K) { return K >= 1 && K <= FF_MLP_MAXK ? layout(K).total : 0; }

extern "C" int ff_mlp_set(int K, const float *params, size_t count) {
    if (check_k(K) != FF_OK) return FF_ERR_ARG;
    if (!params || count != layout(K).total) {
        ff_set_err("expected %zu parameters for K = %d", layout(K).total, K);
        return FF_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g_params.assign(params, params + count);
    g_K = K;
    g_dstale = true;
    g_dispatch_gen++;
    return FF_OK;
}

extern "C" int ff_mlp_get(int *K, float *params, size_t count) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (K) *K = g_K;
    if (!g_K) return FF_ERR_ARG;
    if (params) {
        if (count < g_params.size()) return FF_ERR_ARG;
        memcpy(params, g_params.data(), g_params.size() * sizeof(float));
    }
    return FF_OK;
}

extern "C" int ff_mlp_grad(const float *x, const float *t, size_t n, float *loss, float *grad) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_K || !x || !t || n == 0 || n > (1u << 20)) return FF_ERR_ARG;
    ff_ensure_init();
    MlpDev D(g_K, (int)n);
    const Layout &l = D.l;
    ff_copy(D.P, g_params.data(), l.total * 4, hipMemcpyHostToDevice);
    ff_copy(D.X, x, n * NI * 4, hipMemcpyHostToDevice);
    ff_copy(D.T, t, n * g_K * 4, hipMemcpyHostToDevice);
    FF_CHECK(hipMemsetAsync(D.dP, 0, l.total * 4, g_st));
    D.step((int)n);
    std::vector<float> lo(n);
    FF_CHECK(hipStreamSynchronize(g_st));
    if (grad) ff_copy(grad, D.dP, l.total * 4, hipMemcpyDeviceToHost);
    ff_copy(lo.data(), D.loss, n * 4, hipMemcpyDeviceToHost);
    if (loss) {
        double s = 0;
        for (float v : lo) s += v;
        *loss = (float)(s / n);
    }
    return FF_OK;
}

extern "C" int ff_mlp_train(const float *x, const float *t, size_t n, int epochs, int batch, float lr,
                            uint64_t seed, float *loss_out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_K || !x || !t || n == 0 || epochs < 1 || batch < 1 || batch > 65536) return FF_ERR_ARG;
    ff_ensure_init();
    const int K = g_K;
    const int B = (int)(batch < (int)n ? batch : (int)n);
    MlpDev D(K, B);
    const Layout &l = D.l;
    float *Xall = D.alloc(n * NI), *Tall = D.alloc(n * K);
    int *idx = (int *)D.alloc(B);
    ff_copy(Xall, x, n * NI * 4, hipMemcpyHostToDevice);
    ff_copy(Tall, t, n * K * 4, hipMemcpyHostToDevice);
    ff_copy(D.P, g_params.data(), l.total * 4, hipMemcpyHostToDevice);
    FF_CHECK(hipMemsetAsync(D.m, 0, l.total * 4, g_st));
    FF_CHECK(hipMemsetAsync(D.v, 0, l.total * 4, g_st));
    FF_CHECK(hipMemsetAsync(D.dP, 0, l.total * 4, g_st));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    for (size_t i = 0; i < n; i++) perm[i] = (int)i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const size_t ntrain = l.total - l.W1;  // mu, sd are not trained
    long it = 0;
    std::vector<float> lo(B);
    double last_loss = 0;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        double sum = 0;
        size_t seen = 0;
        for (size_t s = 0; s + B <= n; s += B) {
            FF_CHECK(hipMemcpyAsync(idx, perm.data() + s, B * 4, hipMemcpyHostToDevice, g_st));
            hipLaunchKernelGGL(k_gather, dim3(blocks_for((size_t)B * NI)), dim3(256), 0, g_st, Xall, idx, D.X, B,
                               NI);
            hipLaunchKernelGGL(k_gather, dim3(blocks_for((size_t)B * K)), dim3(256), 0, g_st, Tall, idx, D.T, B,
                               K);
            D.step(B);
            it++;
            const float lr_t = lr * sqrtf(1.0f - powf(b2, (float)it)) / (1.0f - powf(b1, (float)it));
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(ntrain)), dim3(256), 0, g_st, D.P + l.W1, D.dP + l.W1,
                               D.m + l.W1, D.v + l.W1, (int)ntrain, lr_t, b1, b2, eps);
            FF_CHECK(hipGetLastError());
            if (e == epochs - 1) {
                FF_CHECK(hipMemcpyAsync(lo.data(), D.loss, B * 4, hipMemcpyDeviceToHost, g_st));
                FF_CHECK(hipStreamSynchronize(g_st));
                for (float v : lo) sum += v;
                seen += B;
            }
        }
        if (e == epochs - 1 && seen) last_loss = sum / seen;
    }
    FF_CHECK(hipStreamSynchronize(g_st));
    ff_copy(g_params.data(), D.P, l.total * 4, hipMemcpyDeviceToHost);
    g_dstale = true;
    g_dispatch_gen++;
    if (loss_out) *loss_out = (float)last_loss;
    return FF_OK;
}

// Inference on device for n feature rows; y (n x K) and/or choice (argmin
// over mask bits; mask NULL allows all; with margin > 0 the candidate dflt
// is kept unless another is predicted faster by more than margin).
int ff_mlp_infer_dev(const float *dx, size_t n, const uint32_t *dmask, float *dy, int *dchoice, float margin,
                     int dflt) {
    if (!g_K) return FF_ERR_ARG;
    const Layout l = layout(g_K);
    if (!g_dparams) g_dparams = (float *)ff_dalloc(layout(FF_MLP_MAXK).total * 4);
    if (g_dstale) {
        FF_CHECK(hipStreamSynchronize(g_st));
        ff_copy(g_dparams, g_params.data(), l.total * 4, hipMemcpyHostToDevice);
        g_dstale = false;
    }
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks_for(n)), dim3(256), l.total * 4, g_st, dx, (const float *)g_dparams,
                       (int)n, g_K, dmask, dy, dchoice, margin, dflt);
    FF_CHECK(hipGetLastError());
    return FF_OK;
}

int ff_mlp_ready(void) { return g_K; }

extern "C" int ff_mlp_predict(const float *x, size_t n, const uint32_t *mask, float *y, int *choice) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_K || !x || n == 0) return FF_ERR_ARG;
    ff_ensure_init();
    const int K = g_K;
    char *buf = (char *)ff_dalloc(n * (NI + K + 2) * 4 + 1024);
    float *dx = (float *)buf, *dy = dx + n * NI;
    int *dc = (int *)(dy + n * K);
    uint32_t *dm = (uint32_t *)(dc + n);
    FF_CHECK(hipMemcpyAsync(dx, x, n * NI * 4, hipMemcpyHostToDevice, g_st));
    if (mask) FF_CHECK(hipMemcpyAsync(dm, mask, n * 4, hipMemcpyHostToDevice, g_st));
    int rc = ff_mlp_infer_dev(dx, n, mask ? dm : nullptr, dy, dc, 0.0f, -1);
    if (rc == FF_OK) {
        FF_CHECK(hipStreamSynchronize(g_st));
        if (y) ff_copy(y, dy, n * K * 4, hipMemcpyDeviceToHost);
        if (choice) ff_copy(choice, dc, n * 4, hipMemcpyDeviceToHost);
    }
    ff_dfree(buf);
    return rc;
}
