// The cost model's network on the GPU: a small fully connected MLP (tanh on hidden layers,
// linear output), its forward pass, the masked mean-squared-error loss with the gradient of
// every parameter, and Adam.  No counterpart in the reference.
//
// Parameters are one flat FP32 array, layer by layer: W_l (dims[l+1] x dims[l], row-major)
// then b_l (dims[l+1]).  Activations are batch-major (batch x dims[l]).  Every reduction is a
// fixed-order loop or a fixed LDS tree, so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "xc_internal.hpp"
#include "xc_mlp.hpp"

namespace xc {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void layer_forward(const float* __restrict__ X, uint32_t din,
                                                          const float* __restrict__ W, const float* __restrict__ b,
                                                          uint32_t dout, float* __restrict__ A, uint32_t batch,
                                                          int tanh_act) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx >= (uint64_t)batch * dout) return;
    const uint32_t bi = (uint32_t)(idx / dout), o = (uint32_t)(idx % dout);
    const float* x = X + (uint64_t)bi * din;
    const float* w = W + (uint64_t)o * din;
    float z = b[o];
    for (uint32_t i = 0; i < din; ++i) z = fmaf(x[i], w[i], z);
    A[idx] = tanh_act ? tanhf(z) : z;
}

// loss = sum mask*(y-t)^2 / sum mask;  dY = 2*mask*(y-t) / sum mask.  One block.
// If hist is given, the loss is also written to hist[*step] (step < hist_cap).
__global__ __launch_bounds__(kThreads) void loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                        const float* __restrict__ M, uint32_t count,
                                                        float* __restrict__ loss, float* __restrict__ dY,
                                                        float* __restrict__ hist, const uint32_t* __restrict__ step,
                                                        uint32_t hist_cap) {
    __shared__ float red[kThreads];
    float ms = 0.f;
    for (uint32_t i = threadIdx.x; i < count; i += kThreads) ms += M[i];
    red[threadIdx.x] = ms;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const float msum = red[0];
    __syncthreads();
    const float inv = msum > 0.f ? 1.f / msum : 0.f;
    float acc = 0.f;
    for (uint32_t i = threadIdx.x; i < count; i += kThreads) {
        const float d = Y[i] - T[i];
        const float md = M[i] * d;
        acc = fmaf(md, d, acc);
        dY[i] = 2.f * md * inv;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float l = red[0] * inv;
        *loss = l;
        if (hist && step && *step < hist_cap) hist[*step] = l;
    }
}

// dZ = dA * (1 - A^2) for tanh layers (in place), dA for the linear output layer.
__global__ __launch_bounds__(kThreads) void tanh_backward(const float* __restrict__ A, float* __restrict__ dA,
                                                          uint64_t count) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx >= count) return;
    const float a = A[idx];
    dA[idx] *= (1.f - a * a);
}

// dW[o][i] = sum_b dZ[b][o] X[b][i]  (threads 0 .. dout*din-1),
// db[o]    = sum_b dZ[b][o]          (threads dout*din .. dout*din+dout-1).
__global__ __launch_bounds__(kThreads) void weight_grad(const float* __restrict__ dZ, const float* __restrict__ X,
                                                        uint32_t din, uint32_t dout, uint32_t batch,
                                                        float* __restrict__ gW, float* __restrict__ gb) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint64_t nw = (uint64_t)dout * din;
    if (idx < nw) {
        const uint32_t o = (uint32_t)(idx / din), i = (uint32_t)(idx % din);
        float s = 0.f;
        for (uint32_t bi = 0; bi < batch; ++bi) s = fmaf(dZ[(uint64_t)bi * dout + o], X[(uint64_t)bi * din + i], s);
        gW[idx] = s;
    } else if (idx < nw + dout) {
        const uint32_t o = (uint32_t)(idx - nw);
        float s = 0.f;
        for (uint32_t bi = 0; bi < batch; ++bi) s += dZ[(uint64_t)bi * dout + o];
        gb[o] = s;
    }
}

// dX[b][i] = sum_o dZ[b][o] W[o][i].
__global__ __launch_bounds__(kThreads) void input_grad(const float* __restrict__ dZ, const float* __restrict__ W,
                                                       uint32_t din, uint32_t dout, uint32_t batch,
                                                       float* __restrict__ dX) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx >= (uint64_t)batch * din) return;
    const uint32_t bi = (uint32_t)(idx / din), i = (uint32_t)(idx % din);
    float s = 0.f;
    for (uint32_t o = 0; o < dout; ++o) s = fmaf(dZ[(uint64_t)bi * dout + o], W[(uint64_t)o * din + i], s);
    dX[idx] = s;
}

// Adam with decoupled weight decay.  The step number is `step_val`, or *step_dev + 1 when a
// device counter is given (graph replays advance it with tick_kernel).
__global__ __launch_bounds__(kThreads) void adam_kernel(uint64_t count, float* __restrict__ p,
                                                        const float* __restrict__ g, float* __restrict__ m,
                                                        float* __restrict__ v, const uint32_t* __restrict__ step_dev,
                                                        uint32_t step_val, float lr, float b1, float b2, float eps,
                                                        float wd) {
    const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx >= count) return;
    const float t = step_dev ? (float)(*step_dev + 1) : (float)step_val;
    const float gi = g[idx];
    const float mi = b1 * m[idx] + (1.f - b1) * gi;
    const float vi = b2 * v[idx] + (1.f - b2) * gi * gi;
    m[idx] = mi;
    v[idx] = vi;
    const float mh = mi / (1.f - powf(b1, t));
    const float vh = vi / (1.f - powf(b2, t));
    p[idx] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[idx]);
}

__global__ void tick_kernel(uint32_t* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, (n + kThreads - 1) / kThreads)); }

struct layout {
    std::vector<uint64_t> w_off, b_off;  // parameter offsets per layer
    std::vector<uint64_t> a_off;         // activation offsets in `work` per layer output
    uint64_t params = 0, acts = 0, grad_bufs = 0;
    uint32_t max_dim = 0;
};

layout make_layout(const xc_mlp_shape& sh, uint32_t batch) {
    layout L;
    for (uint32_t l = 0; l < sh.layers; ++l) {
        L.w_off.push_back(L.params);
        L.params += (uint64_t)sh.dims[l + 1] * sh.dims[l];
        L.b_off.push_back(L.params);
        L.params += sh.dims[l + 1];
        L.a_off.push_back(L.acts);
        L.acts += (uint64_t)batch * sh.dims[l + 1];
    }
    for (uint32_t l = 0; l <= sh.layers; ++l) L.max_dim = std::max(L.max_dim, sh.dims[l]);
    L.grad_bufs = 2ull * batch * L.max_dim;
    return L;
}

// Forward pass of all layers into work (activations); returns the output activations.
const float* forward_all(const xc_mlp_shape& sh, const layout& L, const float* P, const float* X, uint32_t batch,
                         float* work, hipStream_t s) {
    const float* in = X;
    for (uint32_t l = 0; l < sh.layers; ++l) {
        const uint32_t din = sh.dims[l], dout = sh.dims[l + 1];
        float* out = work + L.a_off[l];
        hipLaunchKernelGGL(layer_forward, grid_for((uint64_t)batch * dout), dim3(kThreads), 0, s, in, din,
                           P + L.w_off[l], P + L.b_off[l], dout, out, batch, (int)(l + 1 < sh.layers));
        in = out;
    }
    return in;
}

// Backward pass from dY (batch x outputs, in the first gradient buffer) into grad.
void backward_all(const xc_mlp_shape& sh, const layout& L, const float* P, const float* X, uint32_t batch,
                  float* work, float* grad, hipStream_t s) {
    float* gbuf[2] = {work + L.acts, work + L.acts + (uint64_t)batch * L.max_dim};
    int cur = 0;  // gbuf[cur] holds dA of layer l's output
    for (int l = (int)sh.layers - 1; l >= 0; --l) {
        const uint32_t din = sh.dims[l], dout = sh.dims[l + 1];
        const float* A = work + L.a_off[l];
        const float* Xin = l > 0 ? work + L.a_off[l - 1] : X;
        float* dZ = gbuf[cur];
        if (l + 1 < (int)sh.layers)
            hipLaunchKernelGGL(tanh_backward, grid_for((uint64_t)batch * dout), dim3(kThreads), 0, s, A, dZ,
                               (uint64_t)batch * dout);
        hipLaunchKernelGGL(weight_grad, grid_for((uint64_t)dout * din + dout), dim3(kThreads), 0, s, dZ, Xin, din,
                           dout, batch, grad + L.w_off[l], grad + L.b_off[l]);
        if (l > 0) {
            hipLaunchKernelGGL(input_grad, grid_for((uint64_t)batch * din), dim3(kThreads), 0, s, dZ,
                               P + L.w_off[l], din, dout, batch, gbuf[cur ^ 1]);
            cur ^= 1;
        }
    }
}

// SplitMix64: deterministic initial weights from a seed.
uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct dev_buf {
    void* p = nullptr;
    explicit dev_buf(size_t bytes) {
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
            (void)hipGetLastError();
            fail(XC_ERR_NO_MEMORY, "MLP buffer allocation failed");
        }
    }
    ~dev_buf() {
        if (p) (void)hipFree(p);
    }
    float* f() const { return static_cast<float*>(p); }
};

}  // namespace

void mlp_check_shape(const xc_mlp_shape* sh) {
    if (!sh) fail(XC_ERR_INVALID_ARGUMENT, "null MLP shape");
    if (sh->layers < 1 || sh->layers > 7) fail(XC_ERR_INVALID_ARGUMENT, "MLP needs 1..7 layers");
    for (uint32_t l = 0; l <= sh->layers; ++l)
        if (sh->dims[l] < 1 || sh->dims[l] > 4096) fail(XC_ERR_INVALID_ARGUMENT, "MLP layer width must be 1..4096");
}

uint64_t mlp_num_params(const xc_mlp_shape& sh) { return make_layout(sh, 1).params; }

uint64_t mlp_work_size(const xc_mlp_shape& sh, uint32_t batch) {
    const layout L = make_layout(sh, batch);
    return L.acts + L.grad_bufs + 1;  // + the loss scalar
}

void mlp_init(const xc_mlp_shape& sh, uint64_t seed, float* params) {
    const layout L = make_layout(sh, 1);
    uint64_t st = seed;
    for (uint32_t l = 0; l < sh.layers; ++l) {
        const double lim = std::sqrt(6.0 / (sh.dims[l] + sh.dims[l + 1]));  // Glorot uniform
        const uint64_t nw = (uint64_t)sh.dims[l + 1] * sh.dims[l];
        for (uint64_t i = 0; i < nw; ++i) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            params[L.w_off[l] + i] = (float)((2.0 * u - 1.0) * lim);
        }
        for (uint32_t i = 0; i < sh.dims[l + 1]; ++i) params[L.b_off[l] + i] = 0.f;
    }
}

void mlp_forward(const xc_mlp_shape& sh, const float* P, const float* X, uint32_t batch, float* Y, float* work,
                 hipStream_t s) {
    const layout L = make_layout(sh, batch);
    const float* out = forward_all(sh, L, P, X, batch, work, s);
    XC_HIP(hipMemcpyAsync(Y, out, sizeof(float) * batch * sh.dims[sh.layers], hipMemcpyDeviceToDevice, s));
    XC_HIP(hipGetLastError());
}

void mlp_loss_grad(const xc_mlp_shape& sh, const float* P, const float* X, const float* T, const float* M,
                   uint32_t batch, float* loss, float* grad, float* work, hipStream_t s) {
    const layout L = make_layout(sh, batch);
    const float* Y = forward_all(sh, L, P, X, batch, work, s);
    float* dY = work + L.acts;
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(kThreads), 0, s, Y, T, M, batch * sh.dims[sh.layers], loss, dY,
                       (float*)nullptr, (const uint32_t*)nullptr, 0u);
    backward_all(sh, L, P, X, batch, work, grad, s);
    XC_HIP(hipGetLastError());
}

void mlp_adam(uint64_t count, float* p, const float* g, float* m, float* v, uint32_t step, float lr, float b1,
              float b2, float eps, float wd, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, grid_for(count), dim3(kThreads), 0, s, count, p, g, m, v,
                       (const uint32_t*)nullptr, step, lr, b1, b2, eps, wd);
    XC_HIP(hipGetLastError());
}

void mlp_train(const xc_mlp_shape& sh, const float* x_host, const float* t_host, const float* m_host,
               uint32_t batch, const xc_train_options& o, float* params_host, float* loss_host) {
    if (batch == 0) fail(XC_ERR_INVALID_ARGUMENT, "empty training set");
    if (o.epochs == 0 || o.epochs > 1000000) fail(XC_ERR_INVALID_ARGUMENT, "epochs must be 1..1000000");
    if (!(o.lr > 0.f)) fail(XC_ERR_INVALID_ARGUMENT, "learning rate must be positive");
    const layout L = make_layout(sh, batch);
    const uint32_t nin = sh.dims[0], nout = sh.dims[sh.layers];
    dev_buf X(sizeof(float) * batch * nin), T(sizeof(float) * batch * nout), M(sizeof(float) * batch * nout);
    dev_buf P(sizeof(float) * L.params), G(sizeof(float) * L.params), Mo(sizeof(float) * L.params),
        Vo(sizeof(float) * L.params);
    dev_buf W(sizeof(float) * (L.acts + L.grad_bufs + 1)), H(sizeof(float) * o.epochs), step(sizeof(uint32_t));
    XC_HIP(hipMemcpy(X.p, x_host, sizeof(float) * batch * nin, hipMemcpyHostToDevice));
    XC_HIP(hipMemcpy(T.p, t_host, sizeof(float) * batch * nout, hipMemcpyHostToDevice));
    XC_HIP(hipMemcpy(M.p, m_host, sizeof(float) * batch * nout, hipMemcpyHostToDevice));
    XC_HIP(hipMemcpy(P.p, params_host, sizeof(float) * L.params, hipMemcpyHostToDevice));
    std::vector<float> zeros(L.params, 0.f);
    XC_HIP(hipMemcpy(Mo.p, zeros.data(), sizeof(float) * L.params, hipMemcpyHostToDevice));
    XC_HIP(hipMemcpy(Vo.p, zeros.data(), sizeof(float) * L.params, hipMemcpyHostToDevice));
    const uint32_t zero = 0;
    XC_HIP(hipMemcpy(step.p, &zero, sizeof(zero), hipMemcpyHostToDevice));

    hipStream_t s = nullptr;
    XC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct stream_guard {
        hipStream_t s;
        ~stream_guard() { (void)hipStreamDestroy(s); }
    } sg{s};
    float* work = W.f();
    float* loss = work + L.acts + L.grad_bufs;
    uint32_t* stp = static_cast<uint32_t*>(step.p);
    auto epoch = [&] {
        const float* Y = forward_all(sh, L, P.f(), X.f(), batch, work, s);
        hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(kThreads), 0, s, Y, T.f(), M.f(), batch * nout, loss,
                           work + L.acts, H.f(), (const uint32_t*)stp, o.epochs);
        backward_all(sh, L, P.f(), X.f(), batch, work, G.f(), s);
        hipLaunchKernelGGL(adam_kernel, grid_for(L.params), dim3(kThreads), 0, s, L.params, P.f(),
                           (const float*)G.f(), Mo.f(), Vo.f(), (const uint32_t*)stp, 0u, o.lr, 0.9f, 0.999f,
                           1e-8f, o.weight_decay);
        hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(64), 0, s, stp);
    };
    if (o.use_graph) {
        // One epoch is captured once and replayed; the step counter lives in device memory.
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        XC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        epoch();
        const hipError_t ce = hipStreamEndCapture(s, &g);
        XC_HIP(ce);
        struct graph_guard {
            hipGraph_t g;
            hipGraphExec_t* ge;
            ~graph_guard() {
                if (*ge) (void)hipGraphExecDestroy(*ge);
                if (g) (void)hipGraphDestroy(g);
            }
        } gg{g, &ge};
        XC_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        for (uint32_t e = 0; e < o.epochs; ++e) XC_HIP(hipGraphLaunch(ge, s));
        XC_HIP(hipStreamSynchronize(s));
    } else {
        for (uint32_t e = 0; e < o.epochs; ++e) epoch();
        XC_HIP(hipGetLastError());
        XC_HIP(hipStreamSynchronize(s));
    }
    XC_HIP(hipMemcpy(params_host, P.p, sizeof(float) * L.params, hipMemcpyDeviceToHost));
    if (loss_host) XC_HIP(hipMemcpy(loss_host, H.p, sizeof(float) * o.epochs, hipMemcpyDeviceToHost));
}

}  // namespace xc
