// FP32 MLP of the MSM planner: 3 inputs -> 32 -> 32 (ReLU) -> 2 outputs,
// trained on the autotuner's sweeps by the kernels below (forward, backward,
// deterministic gradient reduction, Adam), one training step captured in a
// hipGraph and replayed per epoch.
//
// Parameter layout (float, MLP_PARAMS entries):
//   W1[32][3] b1[32] W2[32][32] b2[32] W3[2][32] b3[2]
// Loss: L = (1/n) sum_i sum_k 0.5 (y_ik - t_ik)^2.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/gpu_msm_ext.h"

namespace {

constexpr int IN = 3, H = 32, OUT = 2;
constexpr int O_W1 = 0, O_B1 = O_W1 + H * IN, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H, O_W3 = O_B2 + H,
              O_B3 = O_W3 + OUT * H, NP = O_B3 + OUT;
static_assert(NP == GM_MLP_PARAMS, "parameter count");

__device__ __forceinline__ void forward(const float* __restrict__ P, const float* x, float* h1, float* h2, float* y) {
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        float a = P[O_B1 + j];
#pragma unroll
        for (int i = 0; i < IN; ++i) a = fmaf(P[O_W1 + j * IN + i], x[i], a);
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < H; ++j) {
        float a = P[O_B2 + j];
        for (int i = 0; i < H; ++i) a = fmaf(P[O_W2 + j * H + i], h1[i], a);
        h2[j] = a > 0.f ? a : 0.f;
    }
#pragma unroll
    for (int k = 0; k < OUT; ++k) {
        float a = P[O_B3 + k];
        for (int i = 0; i < H; ++i) a = fmaf(P[O_W3 + k * H + i], h2[i], a);
        y[k] = a;
    }
}

__global__ void __launch_bounds__(64) k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X, int n,
                                                    float* __restrict__ Y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[IN], h1[H], h2[H], y[OUT];
    for (int i = 0; i < IN; ++i) x[i] = X[s * IN + i];
    forward(P, x, h1, h2, y);
    for (int k = 0; k < OUT; ++k) Y[s * OUT + k] = y[k];
}

// One lane per sample: forward pass, then the gradient of that sample's loss
// term (scaled by 1/n) for every parameter, into G[s][NP].
__global__ void __launch_bounds__(64) k_mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ X,
                                                    const float* __restrict__ T, int n, float* __restrict__ G,
                                                    float* __restrict__ loss) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[IN], h1[H], h2[H], y[OUT], dy[OUT], dh2[H], dh1[H];
    for (int i = 0; i < IN; ++i) x[i] = X[s * IN + i];
    forward(P, x, h1, h2, y);
    const float inv_n = 1.0f / (float)n;
    float l = 0.f;
    for (int k = 0; k < OUT; ++k) {
        const float d = y[k] - T[s * OUT + k];
        l += 0.5f * d * d;
        dy[k] = d * inv_n;
    }
    loss[s] = l * inv_n;
    float* g = G + (size_t)s * NP;
    // layer 3
    for (int k = 0; k < OUT; ++k) {
        g[O_B3 + k] = dy[k];
        for (int i = 0; i < H; ++i) g[O_W3 + k * H + i] = dy[k] * h2[i];
    }
    for (int i = 0; i < H; ++i) {
        float a = 0.f;
        for (int k = 0; k < OUT; ++k) a = fmaf(P[O_W3 + k * H + i], dy[k], a);
        dh2[i] = h2[i] > 0.f ? a : 0.f;
    }
    // layer 2
    for (int j = 0; j < H; ++j) {
        g[O_B2 + j] = dh2[j];
        for (int i = 0; i < H; ++i) g[O_W2 + j * H + i] = dh2[j] * h1[i];
    }
    for (int i = 0; i < H; ++i) {
        float a = 0.f;
        for (int j = 0; j < H; ++j) a = fmaf(P[O_W2 + j * H + i], dh2[j], a);
        dh1[i] = h1[i] > 0.f ? a : 0.f;
    }
    // layer 1
    for (int j = 0; j < H; ++j) {
        g[O_B1 + j] = dh1[j];
        for (int i = 0; i < IN; ++i) g[O_W1 + j * IN + i] = dh1[j] * x[i];
    }
}

// Deterministic sum over samples, one lane per parameter (and the loss).
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ G, const float* __restrict__ loss, int n,
                                                    float* __restrict__ grad, float* __restrict__ loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) {
        float a = 0.f;
        for (int s = 0; s < n; ++s) a += G[(size_t)s * NP + p];
        grad[p] = a;
    } else if (p == NP) {
        float a = 0.f;
        for (int s = 0; s < n; ++s) a += loss[s];
        *loss_out = a;
    }
}

// Adam; the step counter lives on the device so a captured step can be
// replayed.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ P, const float* __restrict__ grad,
                                              float* __restrict__ m, float* __restrict__ v,
                                              const int* __restrict__ step, float lr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int t = *step + 1;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)t));
    const float vh = vv / (1.f - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_step_inc(int* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

struct MlpDev {
    float *P = nullptr, *X = nullptr, *T = nullptr, *G = nullptr, *loss = nullptr, *grad = nullptr, *m = nullptr,
          *v = nullptr, *loss_sum = nullptr;
    int* step = nullptr;
    void release() {
        hipFree(P);
        hipFree(X);
        hipFree(T);
        hipFree(G);
        hipFree(loss);
        hipFree(grad);
        hipFree(m);
        hipFree(v);
        hipFree(loss_sum);
        hipFree(step);
    }
};

int alloc(MlpDev& d, int n) {
    if (hipMalloc(&d.P, NP * 4) != hipSuccess || hipMalloc(&d.X, (size_t)n * IN * 4) != hipSuccess ||
        hipMalloc(&d.T, (size_t)n * OUT * 4) != hipSuccess || hipMalloc(&d.G, (size_t)n * NP * 4) != hipSuccess ||
        hipMalloc(&d.loss, (size_t)n * 4) != hipSuccess || hipMalloc(&d.grad, NP * 4) != hipSuccess ||
        hipMalloc(&d.m, NP * 4) != hipSuccess || hipMalloc(&d.v, NP * 4) != hipSuccess ||
        hipMalloc(&d.loss_sum, 4) != hipSuccess || hipMalloc(&d.step, 4) != hipSuccess) {
        d.release();
        return -3;
    }
    return 0;
}

hipError_t enqueue_grad(hipStream_t s, MlpDev& d, int n) {
    k_mlp_fwd_bwd<<<(n + 63) / 64, 64, 0, s>>>(d.P, d.X, d.T, n, d.G, d.loss);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_mlp_reduce<<<(NP + 1 + 255) / 256, 256, 0, s>>>(d.G, d.loss, n, d.grad, d.loss_sum);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int gm_mlp_forward(const float* params, const float* x, int n, float* y, uint32_t gpu_index) {
    if (!params || !x || !y || n <= 0) return -1;
    if (hipSetDevice((int)gpu_index) != hipSuccess) return -2;
    float *P = nullptr, *X = nullptr, *Y = nullptr;
    if (hipMalloc(&P, NP * 4) != hipSuccess || hipMalloc(&X, (size_t)n * IN * 4) != hipSuccess ||
        hipMalloc(&Y, (size_t)n * OUT * 4) != hipSuccess) {
        hipFree(P);
        hipFree(X);
        hipFree(Y);
        return -3;
    }
    hipError_t e = hipMemcpy(P, params, NP * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(X, x, (size_t)n * IN * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_mlp_forward<<<(n + 63) / 64, 64>>>(P, X, n, Y);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, Y, (size_t)n * OUT * 4, hipMemcpyDeviceToHost);
    hipFree(P);
    hipFree(X);
    hipFree(Y);
    return e == hipSuccess ? 0 : -4;
}

int gm_mlp_grad(const float* params, const float* x, const float* t, int n, float* grad, float* loss,
                uint32_t gpu_index) {
    if (!params || !x || !t || !grad || !loss || n <= 0) return -1;
    if (hipSetDevice((int)gpu_index) != hipSuccess) return -2;
    MlpDev d;
    if (alloc(d, n)) return -3;
    hipError_t e = hipMemcpy(d.P, params, NP * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.X, x, (size_t)n * IN * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.T, t, (size_t)n * OUT * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = enqueue_grad(nullptr, d, n);
    if (e == hipSuccess) e = hipMemcpy(grad, d.grad, NP * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss, d.loss_sum, 4, hipMemcpyDeviceToHost);
    d.release();
    return e == hipSuccess ? 0 : -4;
}

int gm_mlp_train(float* params, const float* x, const float* t, int n, int epochs, float lr, float* loss_curve,
                 uint32_t gpu_index) {
    if (!params || !x || !t || n <= 0 || epochs <= 0) return -1;
    if (hipSetDevice((int)gpu_index) != hipSuccess) return -2;
    MlpDev d;
    if (alloc(d, n)) return -3;
    hipStream_t s = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    std::vector<float> curve;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpy(d.P, params, NP * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.X, x, (size_t)n * IN * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.T, t, (size_t)n * OUT * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d.m, 0, NP * 4);
    if (e == hipSuccess) e = hipMemset(d.v, 0, NP * 4);
    if (e == hipSuccess) e = hipMemset(d.step, 0, 4);
    // one training step: gradients, Adam, step counter
    if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        hipError_t e2 = enqueue_grad(s, d, n);
        if (e2 == hipSuccess) {
            k_adam<<<(NP + 255) / 256, 256, 0, s>>>(d.P, d.grad, d.m, d.v, d.step, lr);
            e2 = hipGetLastError();
        }
        if (e2 == hipSuccess) {
            k_step_inc<<<1, 64, 0, s>>>(d.step);
            e2 = hipGetLastError();
        }
        e = hipStreamEndCapture(s, &g);
        if (e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    const int every = loss_curve ? 1 : 0;
    for (int ep = 0; e == hipSuccess && ep < epochs; ++ep) {
        e = hipGraphLaunch(ge, s);
        if (e == hipSuccess && every) {
            float l;
            e = hipMemcpyAsync(&l, d.loss_sum, 4, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            loss_curve[ep] = l;  // loss before this step's update
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(params, d.P, NP * 4, hipMemcpyDeviceToHost);
    if (ge) hipGraphExecDestroy(ge);
    if (g) hipGraphDestroy(g);
    if (s) hipStreamDestroy(s);
    d.release();
    return e == hipSuccess ? 0 : -4;
}

}  // extern "C"
