// qpc_mlp.hip -- the learned Krylov controller's network (include/qpc_mlp.h):
// forward, backward, loss and Adam kernels for a 7 -> 64 -> 64 -> 2 GELU
// perceptron with FP32 parameters.  One 64-lane wave per sample carries the
// forward and the per-sample backward (a lane per hidden unit, activations in
// LDS); the weight gradients are per-parameter sums over the batch in a fixed
// order, so training is reproducible run to run.
#include "../../include/qpc_gpu.h"
#include "../../include/qpc_mlp.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

namespace {

constexpr int NI = QPC_MLP_IN, NH = QPC_MLP_HID, NO = QPC_MLP_OUT, NP = QPC_MLP_NPARAM;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NO * NH;
static_assert(OB3 + NO == NP, "parameter layout");

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// forward of batch position s = blockIdx.x on sample idx[s] (or s)
__global__ __launch_bounds__(NH) void k_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                           const int* __restrict__ idx, float* __restrict__ Z1,
                                           float* __restrict__ H1, float* __restrict__ Z2, float* __restrict__ H2,
                                           float* __restrict__ Y) {
    __shared__ float sx[NI], sh1[NH], sh2[NH];
    const int s = blockIdx.x, j = threadIdx.x;
    const int b = idx ? idx[s] : s;
    if (j < NI) sx[j] = X[(int64_t)b * NI + j];
    __syncthreads();
    float z = P[OB1 + j];
#pragma unroll
    for (int i = 0; i < NI; ++i) z += P[OW1 + j * NI + i] * sx[i];
    float h = gelu(z);
    Z1[(int64_t)s * NH + j] = z;
    H1[(int64_t)s * NH + j] = h;
    sh1[j] = h;
    __syncthreads();
    z = P[OB2 + j];
    for (int k = 0; k < NH; ++k) z += P[OW2 + j * NH + k] * sh1[k];
    h = gelu(z);
    Z2[(int64_t)s * NH + j] = z;
    H2[(int64_t)s * NH + j] = h;
    sh2[j] = h;
    __syncthreads();
    if (j < NO) {
        float y = P[OB3 + j];
        for (int k = 0; k < NH; ++k) y += P[OW3 + j * NH + k] * sh2[k];
        Y[(int64_t)s * NO + j] = y;
    }
}

// per-sample deltas: DY = dL/dy = (y - t)/n, D2 = dL/dz2, D1 = dL/dz1
__global__ __launch_bounds__(NH) void k_bwd(const float* __restrict__ P, const float* __restrict__ T,
                                           const int* __restrict__ idx, int n, const float* __restrict__ Z1,
                                           const float* __restrict__ Z2, const float* __restrict__ Y,
                                           float* __restrict__ DY, float* __restrict__ D1, float* __restrict__ D2) {
    __shared__ float sdy[NO], sd2[NH];
    const int s = blockIdx.x, j = threadIdx.x;
    const int b = idx ? idx[s] : s;
    if (j < NO) {
        const float d = (Y[(int64_t)s * NO + j] - T[(int64_t)b * NO + j]) / (float)n;
        sdy[j] = d;
        DY[(int64_t)s * NO + j] = d;
    }
    __syncthreads();
    float g = 0.0f;
#pragma unroll
    for (int o = 0; o < NO; ++o) g += P[OW3 + o * NH + j] * sdy[o];
    const float d2 = g * gelu_d(Z2[(int64_t)s * NH + j]);
    D2[(int64_t)s * NH + j] = d2;
    sd2[j] = d2;
    __syncthreads();
    g = 0.0f;
    for (int k = 0; k < NH; ++k) g += P[OW2 + k * NH + j] * sd2[k];
    D1[(int64_t)s * NH + j] = g * gelu_d(Z1[(int64_t)s * NH + j]);
}

// gradient of one parameter per lane: a sum over the batch in order s = 0..n-1
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ X, const int* __restrict__ idx, int n,
                                               const float* __restrict__ H1, const float* __restrict__ H2,
                                               const float* __restrict__ D1, const float* __restrict__ D2,
                                               const float* __restrict__ DY, float* __restrict__ G) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    float acc = 0.0f;
    if (p < OB1) {
        const int j = (p - OW1) / NI, i = (p - OW1) % NI;
        for (int s = 0; s < n; ++s) acc += D1[(int64_t)s * NH + j] * X[(int64_t)(idx ? idx[s] : s) * NI + i];
    } else if (p < OW2) {
        const int j = p - OB1;
        for (int s = 0; s < n; ++s) acc += D1[(int64_t)s * NH + j];
    } else if (p < OB2) {
        const int j = (p - OW2) / NH, k = (p - OW2) % NH;
        for (int s = 0; s < n; ++s) acc += D2[(int64_t)s * NH + j] * H1[(int64_t)s * NH + k];
    } else if (p < OW3) {
        const int j = p - OB2;
        for (int s = 0; s < n; ++s) acc += D2[(int64_t)s * NH + j];
    } else if (p < OB3) {
        const int o = (p - OW3) / NH, k = (p - OW3) % NH;
        for (int s = 0; s < n; ++s) acc += DY[(int64_t)s * NO + o] * H2[(int64_t)s * NH + k];
    } else {
        const int o = p - OB3;
        for (int s = 0; s < n; ++s) acc += DY[(int64_t)s * NO + o];
    }
    G[p] = acc;
}

// L = 1/(2n) sum (y - t)^2, one work-group, fixed reduction order
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ Y, const float* __restrict__ T,
                                              const int* __restrict__ idx, int n, float* __restrict__ out) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int s = threadIdx.x; s < n; s += 256) {
        const int b = idx ? idx[s] : s;
        for (int o = 0; o < NO; ++o) {
            const float d = Y[(int64_t)s * NO + o] - T[(int64_t)b * NO + o];
            acc += d * d;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (2.0f * (float)n);
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                              float* __restrict__ M, float* __restrict__ V, float lr, float c1,
                                              float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    const float g = G[p];
    const float m = 0.9f * M[p] + 0.1f * g;
    const float v = 0.999f * V[p] + 0.001f * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct qpc_mlp {
    hipStream_t stream = nullptr;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *loss = nullptr;
    // dataset and per-batch activations (grown on demand)
    float *X = nullptr, *T = nullptr, *Y = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr,
          *D1 = nullptr, *D2 = nullptr, *DY = nullptr;
    int* idx = nullptr;
    int64_t cap_data = 0, cap_batch = 0;
    std::string err;
};

namespace {

#define MLP_HIP(net, x)                                                     \
    do {                                                                    \
        hipError_t _e = (x);                                                \
        if (_e != hipSuccess) {                                             \
            (net)->err = std::string(#x) + ": " + hipGetErrorString(_e);    \
            return QPC_E_HIP;                                               \
        }                                                                   \
    } while (0)

int reserve(qpc_mlp* m, int64_t ndata, int64_t nbatch) {
    if (ndata > m->cap_data) {
        void* old[] = {m->X, m->T, m->idx};
        for (void* p : old)
            if (p) MLP_HIP(m, hipFree(p));
        m->X = m->T = nullptr;
        m->idx = nullptr;
        m->cap_data = 0;
        MLP_HIP(m, hipMalloc((void**)&m->X, sizeof(float) * NI * ndata));
        MLP_HIP(m, hipMalloc((void**)&m->T, sizeof(float) * NO * ndata));
        MLP_HIP(m, hipMalloc((void**)&m->idx, sizeof(int) * ndata));
        m->cap_data = ndata;
    }
    if (nbatch > m->cap_batch) {
        float** bufs[] = {&m->Y, &m->Z1, &m->H1, &m->Z2, &m->H2, &m->D1, &m->D2, &m->DY};
        for (float** b : bufs)
            if (*b) MLP_HIP(m, hipFree(*b));
        for (float** b : bufs) *b = nullptr;
        m->cap_batch = 0;
        MLP_HIP(m, hipMalloc((void**)&m->Y, sizeof(float) * NO * nbatch));
        MLP_HIP(m, hipMalloc((void**)&m->DY, sizeof(float) * NO * nbatch));
        float** hid[] = {&m->Z1, &m->H1, &m->Z2, &m->H2, &m->D1, &m->D2};
        for (float** b : hid) MLP_HIP(m, hipMalloc((void**)b, sizeof(float) * NH * nbatch));
        m->cap_batch = nbatch;
    }
    return QPC_OK;
}

int check_launch(qpc_mlp* m, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        m->err = std::string(what) + ": " + hipGetErrorString(e);
        return QPC_E_HIP;
    }
    return QPC_OK;
}

// forward (+ optionally backward and weight gradients) of n batch positions
void launch_pass(qpc_mlp* m, const int* idx, int n, bool grad) {
    hipLaunchKernelGGL(k_fwd, dim3(n), dim3(NH), 0, m->stream, m->P, m->X, idx, m->Z1, m->H1, m->Z2, m->H2, m->Y);
    if (grad) {
        hipLaunchKernelGGL(k_bwd, dim3(n), dim3(NH), 0, m->stream, m->P, m->T, idx, n, m->Z1, m->Z2, m->Y, m->DY,
                           m->D1, m->D2);
        hipLaunchKernelGGL(k_wgrad, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->X, idx, n, m->H1, m->H2,
                           m->D1, m->D2, m->DY, m->G);
    }
}

}  // namespace

extern "C" {

int qpc_mlp_create(uint64_t seed, qpc_mlp** out) {
    if (!out) return QPC_E_ARG;
    *out = nullptr;
    qpc_mlp* m = new qpc_mlp();
    auto fail = [&](int rc) {
        qpc_mlp_destroy(m);
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail(QPC_E_HIP);
    float** pv[] = {&m->P, &m->G, &m->M, &m->V};
    for (float** p : pv)
        if (hipMalloc((void**)p, sizeof(float) * NP) != hipSuccess) return fail(QPC_E_HIP);
    if (hipMalloc((void**)&m->loss, sizeof(float)) != hipSuccess) return fail(QPC_E_HIP);
    // Glorot uniform: U(-a, a), a = sqrt(6 / (fan_in + fan_out)); zero biases
    std::vector<float> h(NP, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int fo, int fi) {
        const double a = std::sqrt(6.0 / (fi + fo));
        for (int k = 0; k < fo * fi; ++k) h[off + k] = (float)((2.0 * uniform01(s) - 1.0) * a);
    };
    fill(OW1, NH, NI);
    fill(OW2, NH, NH);
    fill(OW3, NO, NH);
    if (qpc_mlp_set_params(m, h.data()) != QPC_OK) return fail(QPC_E_HIP);
    *out = m;
    return QPC_OK;
}

int qpc_mlp_destroy(qpc_mlp* m) {
    if (!m) return QPC_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* ptrs[] = {m->P, m->G, m->M, m->V, m->loss, m->X, m->T, m->Y, m->Z1, m->H1, m->Z2, m->H2, m->D1, m->D2,
                    m->DY, m->idx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return QPC_OK;
}

const char* qpc_mlp_last_error(const qpc_mlp* m) { return m ? m->err.c_str() : "null network"; }

int qpc_mlp_get_params(qpc_mlp* m, float* p) {
    if (!m || !p) return QPC_E_ARG;
    MLP_HIP(m, hipMemcpyAsync(p, m->P, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return QPC_OK;
}

int qpc_mlp_set_params(qpc_mlp* m, const float* p) {
    if (!m || !p) return QPC_E_ARG;
    MLP_HIP(m, hipMemcpyAsync(m->P, p, sizeof(float) * NP, hipMemcpyHostToDevice, m->stream));
    // a new parameter set restarts the optimiser state (fill kernels, no memset nodes)
    std::vector<float> z(NP, 0.0f);
    MLP_HIP(m, hipMemcpyAsync(m->M, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(m, hipMemcpyAsync(m->V, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return QPC_OK;
}

int qpc_mlp_forward(qpc_mlp* m, const float* X, int64_t n, float* Y) {
    if (!m || !X || !Y || n < 1 || n > (1 << 24)) return QPC_E_ARG;
    int rc = reserve(m, n, n);
    if (rc) return rc;
    MLP_HIP(m, hipMemcpyAsync(m->X, X, sizeof(float) * NI * n, hipMemcpyHostToDevice, m->stream));
    launch_pass(m, nullptr, (int)n, false);
    if ((rc = check_launch(m, "mlp forward"))) return rc;
    MLP_HIP(m, hipMemcpyAsync(Y, m->Y, sizeof(float) * NO * n, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return QPC_OK;
}

int qpc_mlp_loss_grad(qpc_mlp* m, const float* X, const float* T, int64_t n, float* loss, float* grad) {
    if (!m || !X || !T || !loss || !grad || n < 1 || n > (1 << 24)) return QPC_E_ARG;
    int rc = reserve(m, n, n);
    if (rc) return rc;
    MLP_HIP(m, hipMemcpyAsync(m->X, X, sizeof(float) * NI * n, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(m, hipMemcpyAsync(m->T, T, sizeof(float) * NO * n, hipMemcpyHostToDevice, m->stream));
    launch_pass(m, nullptr, (int)n, true);
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, m->stream, m->Y, m->T, nullptr, (int)n, m->loss);
    if ((rc = check_launch(m, "mlp loss/grad"))) return rc;
    MLP_HIP(m, hipMemcpyAsync(grad, m->G, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return QPC_OK;
}

int qpc_mlp_train(qpc_mlp* m, const float* X, const float* T, int64_t n, int64_t epochs, int64_t batch, float lr,
                  uint64_t seed, float* loss_out) {
    if (!m || !X || !T || n < 1 || n > (1 << 24) || epochs < 0 || batch < 1 || !(lr > 0.0f)) return QPC_E_ARG;
    if (batch > n) batch = n;
    int rc = reserve(m, n, n);
    if (rc) return rc;
    MLP_HIP(m, hipMemcpyAsync(m->X, X, sizeof(float) * NI * n, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(m, hipMemcpyAsync(m->T, T, sizeof(float) * NO * n, hipMemcpyHostToDevice, m->stream));
    std::vector<int> perm(n);
    uint64_t s = seed;
    int64_t step = 0;
    for (int64_t e = 0; e < epochs; ++e) {
        for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix64(s) % (uint64_t)(i + 1)]);
        // the permutation is read by the epoch's kernels: wait before it is overwritten
        MLP_HIP(m, hipStreamSynchronize(m->stream));
        MLP_HIP(m, hipMemcpyAsync(m->idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->stream));
        for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
            ++step;
            launch_pass(m, m->idx + b0, (int)batch, true);
            const float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
            hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->P, m->G, m->M, m->V, lr,
                               c1, c2);
        }
        if ((rc = check_launch(m, "mlp train"))) return rc;
    }
    // loss over the whole set
    launch_pass(m, nullptr, (int)n, false);
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, m->stream, m->Y, m->T, nullptr, (int)n, m->loss);
    if ((rc = check_launch(m, "mlp loss"))) return rc;
    float l = 0.0f;
    MLP_HIP(m, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    if (loss_out) *loss_out = l;
    return QPC_OK;
}

}  // extern "C"
