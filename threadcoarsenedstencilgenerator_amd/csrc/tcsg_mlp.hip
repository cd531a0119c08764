// Cost-model network of the learned tuner: a 3-layer MLP
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
// in FP32, trained with Adam on a mean-squared error.  Every kernel here is
// deterministic (fixed summation order, no atomics), so a step reproduces
// exactly; the CPU restatement in oracle/mlp.py checks it.
//
// Parameters live in one device array in the order W1[h][in], b1[h],
// W2[h][h], b2[h], W3[1][h], b3[1]; gradients and the two Adam moments use
// the same layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "tcsg.h"

void tcsg_note_hip_error(int e);  // tcsg_runtime.hip

namespace {

#define MLP_HIP(call)                                     \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) {                           \
            tcsg_note_hip_error((int)e_);                 \
            return TCSG_ERR_HIP;                          \
        }                                                 \
    } while (0)

// Y[b][o] = act(sum_i X[b][i] * W[o][i] + bias[o]); one thread per output.
__global__ void linear_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ bias,
                           float *__restrict__ Y, uint32_t B, uint32_t I, uint32_t O, int relu)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * O) return;
    const uint32_t b = t / O, o = t % O;
    const float *x = X + (size_t)b * I, *w = W + (size_t)o * I;
    float acc = bias[o];
    for (uint32_t i = 0; i < I; ++i) acc = fmaf(x[i], w[i], acc);
    Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dZ = dY * (Y > 0) for a ReLU layer (in place on dY).
__global__ void relu_bwd(float *__restrict__ dY, const float *__restrict__ Y, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !(Y[t] > 0.0f)) dY[t] = 0.0f;
}

// dW[o][i] = sum_b dZ[b][o] * X[b][i]; db[o] = sum_b dZ[b][o].  One thread
// per (o, i) (and per o for db, as i == I).
__global__ void linear_bwd_w(const float *__restrict__ dZ, const float *__restrict__ X, float *__restrict__ dW,
                             float *__restrict__ db, uint32_t B, uint32_t I, uint32_t O)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= O * (I + 1)) return;
    const uint32_t o = t / (I + 1), i = t % (I + 1);
    float acc = 0.0f;
    if (i < I) {
        for (uint32_t b = 0; b < B; ++b) acc = fmaf(dZ[(size_t)b * O + o], X[(size_t)b * I + i], acc);
        dW[(size_t)o * I + i] = acc;
    } else {
        for (uint32_t b = 0; b < B; ++b) acc += dZ[(size_t)b * O + o];
        db[o] = acc;
    }
}

// dX[b][i] = sum_o dZ[b][o] * W[o][i]; one thread per (b, i).
__global__ void linear_bwd_x(const float *__restrict__ dZ, const float *__restrict__ W, float *__restrict__ dX,
                             uint32_t B, uint32_t I, uint32_t O)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * I) return;
    const uint32_t b = t / I, i = t % I;
    float acc = 0.0f;
    for (uint32_t o = 0; o < O; ++o) acc = fmaf(dZ[(size_t)b * O + o], W[(size_t)o * I + i], acc);
    dX[t] = acc;
}

// MSE: dpred[b] = 2 (pred[b] - y[b]) / B; one workgroup sums the loss in a
// fixed order (thread-strided partial sums, then a tree in LDS).
__global__ void mse(const float *__restrict__ pred, const float *__restrict__ y, float *__restrict__ dpred,
                    float *__restrict__ loss, uint32_t B)
{
    __shared__ float part[256];
    float s = 0.0f;
    for (uint32_t b = threadIdx.x; b < B; b += blockDim.x) {
        const float d = pred[b] - y[b];
        dpred[b] = 2.0f * d / (float)B;
        s = fmaf(d, d, s);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = part[0] / (float)B;
}

// Adam (bias-corrected), elementwise over all parameters.
__global__ void adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                     uint32_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

} // namespace

struct tcsg_mlp {
    uint32_t in = 0, h = 0, max_batch = 0;
    size_t np = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *out = nullptr, *d2 = nullptr, *d1 = nullptr, *dout = nullptr, *loss = nullptr;
    uint64_t t = 0;  // Adam step count
    size_t off_w1, off_b1, off_w2, off_b2, off_w3, off_b3;
};

static uint64_t mlp_sm64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" {

size_t tcsg_mlp_num_params(uint32_t in_dim, uint32_t hidden)
{
    return (size_t)hidden * in_dim + hidden + (size_t)hidden * hidden + hidden + hidden + 1;
}

int tcsg_mlp_free(tcsg_mlp *n)
{
    if (!n) return TCSG_OK;
    for (float *q : {n->p, n->g, n->m, n->v, n->h1, n->h2, n->out, n->d2, n->d1, n->dout, n->loss})
        if (q) (void)hipFree(q);
    delete n;
    return TCSG_OK;
}

int tcsg_mlp_create(tcsg_mlp **out, uint32_t in_dim, uint32_t hidden, uint32_t max_batch, uint64_t seed)
{
    if (!out || !in_dim || !hidden || !max_batch) return TCSG_ERR_ARG;
    *out = nullptr;
    auto *n = new tcsg_mlp();
    n->in = in_dim;
    n->h = hidden;
    n->max_batch = max_batch;
    n->np = tcsg_mlp_num_params(in_dim, hidden);
    n->off_w1 = 0;
    n->off_b1 = n->off_w1 + (size_t)hidden * in_dim;
    n->off_w2 = n->off_b1 + hidden;
    n->off_b2 = n->off_w2 + (size_t)hidden * hidden;
    n->off_w3 = n->off_b2 + hidden;
    n->off_b3 = n->off_w3 + hidden;
    // weights U[-1/sqrt(fan_in), 1/sqrt(fan_in)) from SplitMix64, zero biases
    // (the He-uniform bound sqrt(6/fan_in) starts the network far rougher and
    // ranks held-out block shapes measurably worse after training)
    std::vector<float> init(n->np, 0.0f);
    uint64_t s = seed;
    auto fill = [&](size_t off, size_t cnt, uint32_t fan_in) {
        const float lim = 1.0f / std::sqrt((float)fan_in);
        for (size_t i = 0; i < cnt; ++i)
            init[off + i] = (float)(((double)(mlp_sm64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim);
    };
    fill(n->off_w1, (size_t)hidden * in_dim, in_dim);
    fill(n->off_w2, (size_t)hidden * hidden, hidden);
    fill(n->off_w3, hidden, hidden);
    hipError_t e = hipSuccess;
    for (float **q : {&n->p, &n->g, &n->m, &n->v})
        if (e == hipSuccess) e = hipMalloc(q, n->np * sizeof(float));
    const size_t act = (size_t)max_batch * hidden;
    for (float **q : {&n->h1, &n->h2, &n->d2, &n->d1})
        if (e == hipSuccess) e = hipMalloc(q, act * sizeof(float));
    for (float **q : {&n->out, &n->dout})
        if (e == hipSuccess) e = hipMalloc(q, (size_t)max_batch * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&n->loss, sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(n->p, init.data(), n->np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->m, 0, n->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->v, 0, n->np * sizeof(float));
    if (e != hipSuccess) {
        tcsg_note_hip_error((int)e);
        tcsg_mlp_free(n);
        return TCSG_ERR_HIP;
    }
    *out = n;
    return TCSG_OK;
}

int tcsg_mlp_get_params(const tcsg_mlp *n, float *host, size_t count)
{
    if (!n || !host || count != n->np) return TCSG_ERR_ARG;
    MLP_HIP(hipMemcpy(host, n->p, n->np * sizeof(float), hipMemcpyDeviceToHost));
    return TCSG_OK;
}

int tcsg_mlp_set_params(tcsg_mlp *n, const float *host, size_t count)
{
    if (!n || !host || count != n->np) return TCSG_ERR_ARG;
    MLP_HIP(hipMemcpy(n->p, host, n->np * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemset(n->m, 0, n->np * sizeof(float)));
    MLP_HIP(hipMemset(n->v, 0, n->np * sizeof(float)));
    n->t = 0;
    return TCSG_OK;
}

static int mlp_forward(tcsg_mlp *n, const float *x, uint32_t B, hipStream_t s)
{
    const uint32_t I = n->in, H = n->h;
    hipLaunchKernelGGL(linear_fwd, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, x, n->p + n->off_w1,
                       n->p + n->off_b1, n->h1, B, I, H, 1);
    hipLaunchKernelGGL(linear_fwd, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, n->h1, n->p + n->off_w2,
                       n->p + n->off_b2, n->h2, B, H, H, 1);
    hipLaunchKernelGGL(linear_fwd, dim3(blocks_for(B)), dim3(256), 0, s, n->h2, n->p + n->off_w3, n->p + n->off_b3,
                       n->out, B, H, 1u, 0);
    MLP_HIP(hipGetLastError());
    return TCSG_OK;
}

// forward + MSE + backward into n->g
static int mlp_grad(tcsg_mlp *n, const float *x, const float *y, uint32_t B, hipStream_t s)
{
    int rc = mlp_forward(n, x, B, s);
    if (rc) return rc;
    const uint32_t I = n->in, H = n->h;
    hipLaunchKernelGGL(mse, dim3(1), dim3(256), 0, s, n->out, y, n->dout, n->loss, B);
    // layer 3 (linear, 1 output)
    hipLaunchKernelGGL(linear_bwd_w, dim3(blocks_for((size_t)1 * (H + 1))), dim3(256), 0, s, n->dout, n->h2,
                       n->g + n->off_w3, n->g + n->off_b3, B, H, 1u);
    hipLaunchKernelGGL(linear_bwd_x, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, n->dout, n->p + n->off_w3,
                       n->d2, B, H, 1u);
    // layer 2
    hipLaunchKernelGGL(relu_bwd, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, n->d2, n->h2, B * H);
    hipLaunchKernelGGL(linear_bwd_w, dim3(blocks_for((size_t)H * (H + 1))), dim3(256), 0, s, n->d2, n->h1,
                       n->g + n->off_w2, n->g + n->off_b2, B, H, H);
    hipLaunchKernelGGL(linear_bwd_x, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, n->d2, n->p + n->off_w2,
                       n->d1, B, H, H);
    // layer 1
    hipLaunchKernelGGL(relu_bwd, dim3(blocks_for((size_t)B * H)), dim3(256), 0, s, n->d1, n->h1, B * H);
    hipLaunchKernelGGL(linear_bwd_w, dim3(blocks_for((size_t)H * (I + 1))), dim3(256), 0, s, n->d1, x,
                       n->g + n->off_w1, n->g + n->off_b1, B, I, H);
    MLP_HIP(hipGetLastError());
    return TCSG_OK;
}

int tcsg_mlp_forward(tcsg_mlp *n, const float *d_x, uint32_t batch, float *d_out, void *stream)
{
    if (!n || !d_x || !d_out || !batch || batch > n->max_batch) return TCSG_ERR_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const int rc = mlp_forward(n, d_x, batch, s);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(d_out, n->out, batch * sizeof(float), hipMemcpyDeviceToDevice, s));
    return TCSG_OK;
}

int tcsg_mlp_gradients(tcsg_mlp *n, const float *d_x, const float *d_y, uint32_t batch, float *host_grads,
                       float *host_loss)
{
    if (!n || !d_x || !d_y || !host_grads || !batch || batch > n->max_batch) return TCSG_ERR_ARG;
    const int rc = mlp_grad(n, d_x, d_y, batch, nullptr);
    if (rc) return rc;
    MLP_HIP(hipMemcpy(host_grads, n->g, n->np * sizeof(float), hipMemcpyDeviceToHost));
    if (host_loss) MLP_HIP(hipMemcpy(host_loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
    return TCSG_OK;
}

int tcsg_mlp_train_step(tcsg_mlp *n, const float *d_x, const float *d_y, uint32_t batch, float lr, float *d_loss,
                        void *stream)
{
    if (!n || !d_x || !d_y || !batch || batch > n->max_batch || !(lr > 0.0f)) return TCSG_ERR_ARG;
    const hipStream_t s = (hipStream_t)stream;
    int rc = mlp_grad(n, d_x, d_y, batch, s);
    if (rc) return rc;
    n->t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)n->t);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)n->t);
    hipLaunchKernelGGL(adam, dim3(blocks_for(n->np)), dim3(256), 0, s, n->p, n->g, n->m, n->v, (uint32_t)n->np, lr,
                       b1, b2, 1e-8f, c1, c2);
    MLP_HIP(hipGetLastError());
    if (d_loss) MLP_HIP(hipMemcpyAsync(d_loss, n->loss, sizeof(float), hipMemcpyDeviceToDevice, s));
    return TCSG_OK;
}

} // extern "C"
