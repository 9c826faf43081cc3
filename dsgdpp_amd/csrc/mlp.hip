// Step-size controller network: a 9 -> 64 -> 64 -> 1 perceptron with GELU
// hidden units, FP32, trained and evaluated on the GPU.
//
// Parameters sit in one device array (MLP_PARAMS floats):
//   W1 [64 x 9] (row j = hidden unit j), b1 [64], W2 [64 x 64], b2 [64], W3 [64], b3 [1]
// The training loss is the mean squared error over a batch.  Gradients are
// formed per sample and then summed over samples in sample order by one thread
// per parameter, so a gradient is reproducible bit for bit; Adam runs on the
// device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dsgdpp.h"
#include "device.h"

namespace dsgdpp {

constexpr int MLP_IN = 9, MLP_H = 64;
constexpr int P_W1 = 0, P_B1 = P_W1 + MLP_H * MLP_IN, P_W2 = P_B1 + MLP_H, P_B2 = P_W2 + MLP_H * MLP_H,
              P_W3 = P_B2 + MLP_H, P_B3 = P_W3 + MLP_H, MLP_PARAMS = P_B3 + 1;
constexpr int MLP_CHUNK = 2048;   // samples per gradient pass (per-sample gradient buffer)

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
// d/dz of z * Phi(z) = Phi(z) + z * phi(z)
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// One workgroup (64 lanes = 64 hidden units) per sample.
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                    float* __restrict__ out, float* __restrict__ Z1,
                                                    float* __restrict__ Z2) {
    __shared__ float h1[MLP_H];
    const int n = blockIdx.x, j = threadIdx.x;
    const float* x = X + (int64_t)n * MLP_IN;
    float z1 = P[P_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) z1 = fmaf(P[P_W1 + j * MLP_IN + i], x[i], z1);
    h1[j] = gelu(z1);
    __syncthreads();
    float z2 = P[P_B2 + j];
    for (int k = 0; k < MLP_H; ++k) z2 = fmaf(P[P_W2 + j * MLP_H + k], h1[k], z2);
    const float y = wave_sum(P[P_W3 + j] * gelu(z2));
    if (Z1) {
        Z1[(int64_t)n * MLP_H + j] = z1;
        Z2[(int64_t)n * MLP_H + j] = z2;
    }
    if (j == 0) out[n] = y + P[P_B3];
}

// Per-sample gradient of scale * (out - y)^2 / 2 * 2 = scale * (out - y) * d out.
__global__ __launch_bounds__(64) void k_mlp_backward(const float* __restrict__ P, const float* __restrict__ X,
                                                     const float* __restrict__ Y, const float* __restrict__ out,
                                                     const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                     float* __restrict__ G, float scale) {
    __shared__ float h1[MLP_H], dz2s[MLP_H];
    const int n = blockIdx.x, j = threadIdx.x;
    const float* x = X + (int64_t)n * MLP_IN;
    float* g = G + (int64_t)n * MLP_PARAMS;
    const float z1 = Z1[(int64_t)n * MLP_H + j], z2 = Z2[(int64_t)n * MLP_H + j];
    h1[j] = gelu(z1);
    const float dout = scale * (out[n] - Y[n]);
    g[P_W3 + j] = dout * gelu(z2);
    if (j == 0) g[P_B3] = dout;
    const float dz2 = P[P_W3 + j] * dout * gelu_grad(z2);
    dz2s[j] = dz2;
    g[P_B2 + j] = dz2;
    __syncthreads();
    for (int k = 0; k < MLP_H; ++k) g[P_W2 + j * MLP_H + k] = dz2 * h1[k];
    float dh1 = 0.f;   // lane j = hidden unit j of layer 1
    for (int i = 0; i < MLP_H; ++i) dh1 = fmaf(P[P_W2 + i * MLP_H + j], dz2s[i], dh1);
    const float dz1 = dh1 * gelu_grad(z1);
    g[P_B1 + j] = dz1;
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) g[P_W1 + j * MLP_IN + i] = dz1 * x[i];
}

// grad[p] (+)= sum over the chunk's samples, in sample order
__global__ void k_mlp_sum(const float* __restrict__ G, int n, float* __restrict__ grad, int accumulate) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    float s = accumulate ? grad[p] : 0.f;
    for (int k = 0; k < n; ++k) s += G[(int64_t)k * MLP_PARAMS + p];
    grad[p] = s;
}

// sum of (out - y)^2 over the chunk, in order (one lane)
__global__ void k_mlp_sse(const float* __restrict__ out, const float* __restrict__ Y, int n, double* acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = *acc;
    for (int k = 0; k < n; ++k) {
        const double e = (double)out[k] - (double)Y[k];
        s += e * e;
    }
    *acc = s;
}

__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    const float g = grad[p];
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    P[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace dsgdpp

using namespace dsgdpp;

struct dsgdpp_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* P = nullptr;      // parameters
    float* grad = nullptr;   // gradient
    float* m1 = nullptr;     // Adam moments
    float* m2 = nullptr;
    float* X = nullptr;      // samples (capacity cap)
    float* Y = nullptr;
    float* out = nullptr;
    float* Z1 = nullptr;
    float* Z2 = nullptr;
    float* G = nullptr;      // per-sample gradients of one chunk
    double* sse = nullptr;
    int64_t cap = 0;
    int64_t adam_t = 0;
    std::string err;
};

namespace {

thread_local std::string g_mlp_create_error;

#define MLPC(call)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
void mfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void reserve(dsgdpp_mlp* m, int64_t n) {
    if (n <= m->cap) return;
    mfree(m->X); mfree(m->Y); mfree(m->out); mfree(m->Z1); mfree(m->Z2);
    const int64_t c = std::max<int64_t>(n, 256);
    MLPC(hipMalloc(&m->X, c * MLP_IN * sizeof(float)));
    MLPC(hipMalloc(&m->Y, c * sizeof(float)));
    MLPC(hipMalloc(&m->out, c * sizeof(float)));
    MLPC(hipMalloc(&m->Z1, c * MLP_H * sizeof(float)));
    MLPC(hipMalloc(&m->Z2, c * MLP_H * sizeof(float)));
    m->cap = c;
}

void upload(dsgdpp_mlp* m, const float* X, const float* Y, int64_t n) {
    reserve(m, n);
    MLPC(hipMemcpyAsync(m->X, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    if (Y) MLPC(hipMemcpyAsync(m->Y, Y, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
}

// gradient of the mean squared error over the uploaded batch into m->grad,
// and the mean squared error (before the step) into *loss
void gradient(dsgdpp_mlp* m, int64_t n, double* loss) {
    MLPC(hipMemsetAsync(m->sse, 0, sizeof(double), m->stream));
    const float scale = 2.f / (float)n;
    for (int64_t c0 = 0; c0 < n; c0 += MLP_CHUNK) {
        const int cn = (int)std::min<int64_t>(MLP_CHUNK, n - c0);
        k_mlp_forward<<<cn, 64, 0, m->stream>>>(m->P, m->X + c0 * MLP_IN, m->out + c0, m->Z1 + c0 * MLP_H,
                                                 m->Z2 + c0 * MLP_H);
        k_mlp_backward<<<cn, 64, 0, m->stream>>>(m->P, m->X + c0 * MLP_IN, m->Y + c0, m->out + c0,
                                                  m->Z1 + c0 * MLP_H, m->Z2 + c0 * MLP_H, m->G, scale);
        k_mlp_sum<<<(MLP_PARAMS + 255) / 256, 256, 0, m->stream>>>(m->G, cn, m->grad, c0 > 0 ? 1 : 0);
        k_mlp_sse<<<1, 64, 0, m->stream>>>(m->out + c0, m->Y + c0, cn, m->sse);
    }
    MLPC(hipGetLastError());
    if (loss) {
        double s = 0;
        MLPC(hipMemcpyAsync(&s, m->sse, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        MLPC(hipStreamSynchronize(m->stream));
        *loss = s / (double)n;
    }
}

int fail(dsgdpp_mlp* m, int code, const std::string& what) {
    if (m) m->err = what;
    return code;
}

}  // namespace

extern "C" {

int dsgdpp_mlp_num_params(void) { return MLP_PARAMS; }

int dsgdpp_mlp_create(int device, dsgdpp_mlp** out) {
    if (!out) return DSGDPP_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        g_mlp_create_error = "no HIP device";
        return DSGDPP_E_DEVICE;
    }
    if (device < 0 || device >= count) {
        g_mlp_create_error = "device index out of range";
        return DSGDPP_E_ARG;
    }
    dsgdpp_mlp* m = new dsgdpp_mlp();
    try {
        m->device = device;
        MLPC(hipSetDevice(device));
        MLPC(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        MLPC(hipMalloc(&m->P, MLP_PARAMS * sizeof(float)));
        MLPC(hipMalloc(&m->grad, MLP_PARAMS * sizeof(float)));
        MLPC(hipMalloc(&m->m1, MLP_PARAMS * sizeof(float)));
        MLPC(hipMalloc(&m->m2, MLP_PARAMS * sizeof(float)));
        MLPC(hipMalloc(&m->G, (size_t)MLP_CHUNK * MLP_PARAMS * sizeof(float)));
        MLPC(hipMalloc(&m->sse, sizeof(double)));
        MLPC(hipMemset(m->P, 0, MLP_PARAMS * sizeof(float)));
        MLPC(hipMemset(m->m1, 0, MLP_PARAMS * sizeof(float)));
        MLPC(hipMemset(m->m2, 0, MLP_PARAMS * sizeof(float)));
    } catch (const std::exception& e) {
        g_mlp_create_error = e.what();
        dsgdpp_mlp_destroy(m);
        return DSGDPP_E_DEVICE;
    }
    *out = m;
    return DSGDPP_OK;
}

const char* dsgdpp_mlp_create_error(void) { return g_mlp_create_error.c_str(); }
const char* dsgdpp_mlp_last_error(dsgdpp_mlp* m) { return m ? m->err.c_str() : ""; }

void dsgdpp_mlp_destroy(dsgdpp_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    mfree(m->P); mfree(m->grad); mfree(m->m1); mfree(m->m2); mfree(m->G); mfree(m->sse);
    mfree(m->X); mfree(m->Y); mfree(m->out); mfree(m->Z1); mfree(m->Z2);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int dsgdpp_mlp_set_params(dsgdpp_mlp* m, const float* params) {
    if (!m || !params) return DSGDPP_E_ARG;
    try {
        MLPC(hipSetDevice(m->device));
        MLPC(hipMemcpy(m->P, params, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        MLPC(hipMemset(m->m1, 0, MLP_PARAMS * sizeof(float)));
        MLPC(hipMemset(m->m2, 0, MLP_PARAMS * sizeof(float)));
        m->adam_t = 0;
    } catch (const std::exception& e) {
        return fail(m, DSGDPP_E_DEVICE, e.what());
    }
    return DSGDPP_OK;
}

int dsgdpp_mlp_get_params(dsgdpp_mlp* m, float* params) {
    if (!m || !params) return DSGDPP_E_ARG;
    try {
        MLPC(hipSetDevice(m->device));
        MLPC(hipStreamSynchronize(m->stream));
        MLPC(hipMemcpy(params, m->P, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    } catch (const std::exception& e) {
        return fail(m, DSGDPP_E_DEVICE, e.what());
    }
    return DSGDPP_OK;
}

int dsgdpp_mlp_forward(dsgdpp_mlp* m, const float* X, int64_t n, float* out) {
    if (!m || (n > 0 && (!X || !out)) || n < 0) return DSGDPP_E_ARG;
    if (n == 0) return DSGDPP_OK;
    try {
        MLPC(hipSetDevice(m->device));
        upload(m, X, nullptr, n);
        for (int64_t c0 = 0; c0 < n; c0 += 65535) {
            const int cn = (int)std::min<int64_t>(65535, n - c0);
            k_mlp_forward<<<cn, 64, 0, m->stream>>>(m->P, m->X + c0 * MLP_IN, m->out + c0, nullptr, nullptr);
        }
        MLPC(hipGetLastError());
        MLPC(hipMemcpyAsync(out, m->out, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MLPC(hipStreamSynchronize(m->stream));
    } catch (const std::exception& e) {
        return fail(m, DSGDPP_E_DEVICE, e.what());
    }
    return DSGDPP_OK;
}

int dsgdpp_mlp_gradient(dsgdpp_mlp* m, const float* X, const float* Y, int64_t n, float* grad, double* loss) {
    if (!m || !X || !Y || !grad || n <= 0) return DSGDPP_E_ARG;
    try {
        MLPC(hipSetDevice(m->device));
        upload(m, X, Y, n);
        gradient(m, n, loss);
        MLPC(hipMemcpyAsync(grad, m->grad, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MLPC(hipStreamSynchronize(m->stream));
    } catch (const std::exception& e) {
        return fail(m, DSGDPP_E_DEVICE, e.what());
    }
    return DSGDPP_OK;
}

int dsgdpp_mlp_train(dsgdpp_mlp* m, const float* X, const float* Y, int64_t n, int64_t steps, float lr,
                     double* losses) {
    if (!m || !X || !Y || n <= 0 || steps < 0 || !(lr > 0)) return DSGDPP_E_ARG;
    try {
        MLPC(hipSetDevice(m->device));
        upload(m, X, Y, n);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int64_t s = 0; s < steps; ++s) {
            gradient(m, n, losses ? losses + s : nullptr);
            ++m->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
            k_mlp_adam<<<(MLP_PARAMS + 255) / 256, 256, 0, m->stream>>>(m->P, m->grad, m->m1, m->m2, lr, b1, b2,
                                                                          eps, c1, c2);
            MLPC(hipGetLastError());
        }
        MLPC(hipStreamSynchronize(m->stream));
    } catch (const std::exception& e) {
        return fail(m, DSGDPP_E_DEVICE, e.what());
    }
    return DSGDPP_OK;
}

}  // extern "C"
