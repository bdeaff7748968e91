// On-device error checks, reduced in FP64:
//  * validate_result statistics (src/validate.hpp:134-188): per element d = 0 when both
//    |result| and |valid| are at most 1e-5, else |result - valid| / max(|result|, |valid|);
//    max (first index on ties), mean, population standard deviation, and the two values at
//    the max;
//  * the normalised error e = |x - ref| / den of an output against an FP64 reference, den =
//    corr(|L|, |R|) (0 where den == 0 and x == ref): max, mean and the index of the max.
// Two passes: blocks reduce grid-strided ranges into per-block partials (a workspace), one
// block reduces those.  The host reads five doubles back, so the calls synchronise.
#include <algorithm>
#include <cmath>
#include <functional>

#include "xc_internal.hpp"
#include "xc_methods.hpp"

namespace xc {
namespace {

constexpr int kCheckThreads = 256;
constexpr uint32_t kCheckBlocks = 1024;

struct partial {
    double mx;
    uint64_t imax;
    double sum, sum2;
};

__device__ __forceinline__ void merge(partial& a, const partial& b) {
    // max with the lowest index on ties; NaN never wins (as the CPU loop's d > max)
    if (b.mx > a.mx || (b.mx == a.mx && b.imax < a.imax)) {
        a.mx = b.mx;
        a.imax = b.imax;
    }
    a.sum += b.sum;
    a.sum2 += b.sum2;
}

__device__ partial block_reduce(partial p) {
    __shared__ partial sh[kCheckThreads];
    sh[threadIdx.x] = p;
    __syncthreads();
    for (int w = kCheckThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            partial q = sh[threadIdx.x];
            merge(q, sh[threadIdx.x + w]);
            sh[threadIdx.x] = q;
        }
        __syncthreads();
    }
    return sh[0];
}

template <typename T>
__device__ __forceinline__ double validate_d(T r, T v) {
    const double a = (double)r, b = (double)v;
    const double fa = fabs(a), fb = fabs(b);
    if (fa <= 1e-5 && fb <= 1e-5) return 0.0;
    return fabs(a - b) / (fa > fb ? fa : fb);
}

template <typename T>
__global__ __launch_bounds__(kCheckThreads) void validate_partial_kernel(const T* __restrict__ result,
                                                                         const T* __restrict__ valid, uint64_t count,
                                                                         partial* __restrict__ out) {
    partial p{-INFINITY, ~0ull, 0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double d = validate_d(result[i], valid[i]);
        if (d > p.mx) {  // indices grow along a thread's range: the first maximum stays
            p.mx = d;
            p.imax = i;
        }
        p.sum += d;
        p.sum2 += d * d;
    }
    const partial r = block_reduce(p);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

template <typename TX>
__global__ __launch_bounds__(kCheckThreads) void norm_error_partial_kernel(const TX* __restrict__ x,
                                                                           const double* __restrict__ ref,
                                                                           const double* __restrict__ den,
                                                                           uint64_t count, partial* __restrict__ out) {
    partial p{0.0, ~0ull, 0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double d = fabs((double)x[i] - ref[i]);
        const double e = den[i] > 0 ? d / den[i] : (d == 0 ? 0.0 : INFINITY);
        if (e > p.mx) {  // the first maximum above 0 (all 0: index 0, as the CPU loop)
            p.mx = e;
            p.imax = i;
        }
        p.sum += e;
    }
    const partial r = block_reduce(p);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

__global__ __launch_bounds__(kCheckThreads) void finish_kernel(const partial* __restrict__ parts, uint32_t n,
                                                               partial* __restrict__ out, int empty_max) {
    partial p{empty_max ? -INFINITY : 0.0, ~0ull, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) merge(p, parts[i]);
    const partial r = block_reduce(p);
    if (threadIdx.x == 0) *out = r;
}

uint32_t check_blocks(uint64_t count) {
    const uint64_t b = (count + kCheckThreads - 1) / kCheckThreads;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, kCheckBlocks));
}

partial reduce_on_device(uint64_t count, hipStream_t s, const std::function<void(uint32_t, partial*)>& first,
                         bool empty_max) {
    const uint32_t blocks = check_blocks(count);
    partial* ws = static_cast<partial*>(workspace(WS_CHECK, sizeof(partial) * (blocks + 1), s));
    first(blocks, ws + 1);
    XC_HIP(hipGetLastError());
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kCheckThreads), 0, s, ws + 1, blocks, ws, (int)empty_max);
    XC_HIP(hipGetLastError());
    partial r{};
    XC_HIP(hipMemcpyAsync(&r, ws, sizeof(partial), hipMemcpyDeviceToHost, s));
    XC_HIP(hipStreamSynchronize(s));
    return r;
}

template <typename T>
void validate_result_device(const T* result, const T* valid, uint64_t count, double* out, hipStream_t s) {
    if (count == 0) {
        for (int i = 0; i < 5; ++i) out[i] = NAN;
        return;
    }
    const partial r = reduce_on_device(
        count, s,
        [&](uint32_t blocks, partial* parts) {
            hipLaunchKernelGGL(validate_partial_kernel<T>, dim3(blocks), dim3(kCheckThreads), 0, s, result, valid,
                               count, parts);
        },
        true);
    const double mean = r.sum / (double)count;
    const double var = r.sum2 / (double)count - mean * mean;
    out[0] = r.mx;
    out[1] = mean;
    out[2] = std::sqrt(var > 0 ? var : 0.0);
    uint64_t im = r.imax == ~0ull ? 0 : r.imax;  // every d NaN: the CPU loop reports element 0
    T rv, vv;
    XC_HIP(hipMemcpyAsync(&rv, result + im, sizeof(T), hipMemcpyDeviceToHost, s));
    XC_HIP(hipMemcpyAsync(&vv, valid + im, sizeof(T), hipMemcpyDeviceToHost, s));
    XC_HIP(hipStreamSynchronize(s));
    if (r.imax == ~0ull) out[0] = -INFINITY;
    out[3] = (double)rv;
    out[4] = (double)vv;
}

template <typename TX>
void norm_error_device(const TX* x, const double* ref, const double* den, uint64_t count, double* out,
                       hipStream_t s) {
    if (count == 0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const partial r = reduce_on_device(
        count, s,
        [&](uint32_t blocks, partial* parts) {
            hipLaunchKernelGGL(norm_error_partial_kernel<TX>, dim3(blocks), dim3(kCheckThreads), 0, s, x, ref, den,
                               count, parts);
        },
        false);
    out[0] = r.mx;
    out[1] = r.sum / (double)count;
    out[2] = (double)(r.imax == ~0ull ? 0 : r.imax);
}

}  // namespace

void validate_result_f32(const float* r, const float* v, uint64_t n, double* out, hipStream_t s) {
    validate_result_device(r, v, n, out, s);
}
void validate_result_f64(const double* r, const double* v, uint64_t n, double* out, hipStream_t s) {
    validate_result_device(r, v, n, out, s);
}
void norm_error_f32(const float* x, const double* ref, const double* den, uint64_t n, double* out, hipStream_t s) {
    norm_error_device(x, ref, den, n, out, s);
}
void norm_error_f64(const double* x, const double* ref, const double* den, uint64_t n, double* out, hipStream_t s) {
    norm_error_device(x, ref, den, n, out, s);
}

}  // namespace xc
