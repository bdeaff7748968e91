// FFT method: fused zero-pad scatter -> rocFFT (via hipFFT) R2C -> Hadamard product with
// the conjugate -> C2R -> on-device normalisation (quadrant swap, 1/(rows*cols), crop).
//
// Reference: fft_reduced_transfer_* (src/one_to_many.hpp:679-745, src/one_to_one.hpp:1021-1180,
// src/n_to_mn.hpp:900-1000) and fft_better_n_to_m (src/n_to_m.hpp:700-736) with the kernels of
// src/scatter.cu:15-102 and src/hadamard.cu:18-258; normalisation of src/fft_helpers.hpp:193-261
// (there on the host after the copy back, here a GPU kernel writing the final layout).
#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "xc_methods.hpp"

namespace xc {
namespace {

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };

// Writes every element of the dst matrices: src value inside the placed window, zero
// elsewhere (so the reference's separate cudaMemset of the padded buffer disappears).
template <typename Ti, typename To>
__global__ void scatter_kernel(const Ti* __restrict__ src, To* __restrict__ dst, uint32_t src_rows,
                               uint32_t src_cols, uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row,
                               uint32_t dst_col, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t area = (uint64_t)dst_rows * dst_cols;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t b = i / area;
        const uint32_t off = (uint32_t)(i - b * area);
        const uint32_t y = off / dst_cols, x = off - y * dst_cols;
        const uint32_t sy = y - dst_row, sx = x - dst_col;  // wraps to large values when outside
        To v = To(0);
        if (sy < src_rows && sx < src_cols) v = (To)src[(b * src_rows + sy) * src_cols + sx];
        dst[i] = v;
    }
}

// def[b][i] = ref[b % nref][i] * conj(def[b][i])   (src/hadamard.cu:18-24,33-58)
template <typename C>
__global__ void hadamard_kernel(const C* __restrict__ ref, C* __restrict__ def, uint64_t area, uint32_t nref,
                                uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t b = i / area;
        const uint64_t e = i - b * area;
        const C a = ref[(b % nref) * area + e];
        const C d = def[i];
        def[i] = C{a.x * d.x + a.y * d.y, a.y * d.x - a.x * d.y};
    }
}

// out[l*m + r][i] = left[l][i] * conj(right[r][i])   (src/hadamard.cu:116-149)
template <typename C>
__global__ void hadamard_n_to_m_kernel(const C* __restrict__ left, const C* __restrict__ right, C* __restrict__ out,
                                       uint64_t area, uint32_t m, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t k = i / area;
        const uint64_t e = i - k * area;
        const uint64_t l = k / m, r = k - l * m;
        const C a = left[l * area + e];
        const C d = right[r * area + e];
        out[i] = C{a.x * d.x + a.y * d.y, a.y * d.x - a.x * d.y};
    }
}

// norm[b][ny][nx] = res[b][sy][sx] / (rows*cols) with the quadrant swap of
// normalize_fft_results: sy = ny - nr/2 if ny >= nr/2 else ny + rows/2 + 1 (same for x).
template <typename Ti, typename To>
__global__ void normalize_kernel(const Ti* __restrict__ res, To* __restrict__ norm, uint32_t rows, uint32_t cols,
                                 uint64_t total) {
    const uint32_t nr = rows - 1, nc = cols - 1;
    const uint32_t ey = rows / 2, ex = cols / 2;
    const uint64_t area_n = (uint64_t)nr * nc;
    const Ti scale = Ti(1) / ((Ti)rows * (Ti)cols);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t b = i / area_n;
        const uint32_t off = (uint32_t)(i - b * area_n);
        const uint32_t ny = off / nc, nx = off - ny * nc;
        const uint32_t sy = ny >= nr / 2 ? ny - nr / 2 : ny + ey + 1;
        const uint32_t sx = nx >= nc / 2 ? nx - nc / 2 : nx + ex + 1;
        norm[i] = (To)(res[(b * rows + sy) * cols + sx] * scale);
    }
}

inline uint32_t grid_for(uint64_t total) {
    uint64_t b = (total + 255) / 256;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), 16384);
}

void fft_check(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) fail(XC_ERR_FFT, std::string(what) + " failed with hipfftResult " + std::to_string((int)r));
}

std::mutex g_fft_mutex;

// Cached plans keyed by (device, stream, padded rows, padded cols, batch, type): a plan is
// made on one device, and its internal work buffer must not be shared by work that can run
// at the same time on two streams.
hipfftHandle get_plan(hipStream_t s, uint32_t prow, uint32_t pcol, uint32_t batch, hipfftType type) {
    static std::map<std::tuple<int, hipStream_t, uint32_t, uint32_t, uint32_t, int>, hipfftHandle> plans;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    auto key = std::make_tuple(dev, s, prow, pcol, batch, (int)type);
    auto it = plans.find(key);
    if (it != plans.end()) return it->second;
    hipfftHandle h;
    int dims[2] = {(int)prow, (int)pcol};
    fft_check(hipfftPlanMany(&h, 2, dims, nullptr, 1, 0, nullptr, 1, 0, type, (int)batch), "hipfftPlanMany");
    plans.emplace(key, h);
    return h;
}

template <typename Tt> struct fft_types;
template <> struct fft_types<float> {
    static constexpr hipfftType r2c = HIPFFT_R2C, c2r = HIPFFT_C2R;
    static void fwd(hipfftHandle p, float* in, float2* out) {
        fft_check(hipfftExecR2C(p, in, reinterpret_cast<hipfftComplex*>(out)), "hipfftExecR2C");
    }
    static void inv(hipfftHandle p, float2* in, float* out) {
        fft_check(hipfftExecC2R(p, reinterpret_cast<hipfftComplex*>(in), out), "hipfftExecC2R");
    }
};
template <> struct fft_types<double> {
    static constexpr hipfftType r2c = HIPFFT_D2Z, c2r = HIPFFT_Z2D;
    static void fwd(hipfftHandle p, double* in, double2* out) {
        fft_check(hipfftExecD2Z(p, in, reinterpret_cast<hipfftDoubleComplex*>(out)), "hipfftExecD2Z");
    }
    static void inv(hipfftHandle p, double2* in, double* out) {
        fft_check(hipfftExecZ2D(p, reinterpret_cast<hipfftDoubleComplex*>(in), out), "hipfftExecZ2D");
    }
};

// Ti: data type of inputs/outputs; Tt: transform precision.
template <typename Ti, typename Tt>
void run_fft(const family_args& a, const Ti* L, const Ti* R, Ti* O, bool normalize, hipStream_t s) {
    using C = typename cplx<Tt>::type;
    std::lock_guard<std::mutex> lock(g_fft_mutex);
    const uint32_t H = a.rows, W = a.cols;
    const uint32_t PR = 2 * H, PC = 2 * W;
    const uint64_t parea = (uint64_t)PR * PC;
    const uint64_t sarea = (uint64_t)PR * (PC / 2 + 1);
    const uint32_t nl = a.family == XC_ONE_TO_ONE || a.family == XC_ONE_TO_MANY ? 1 : a.n;
    const uint32_t nr = a.m;
    const uint32_t k = a.pairs;
    const bool all_pairs = a.family == XC_N_TO_M;

    Tt* pad = (Tt*)workspace(WS_FFT_PAD, (nl + nr) * parea * sizeof(Tt), s);
    C* spec = (C*)workspace(WS_FFT_SPEC, (nl + nr) * sarea * sizeof(C), s);
    C* ospec = all_pairs ? (C*)workspace(WS_FFT_OUT, (uint64_t)k * sarea * sizeof(C), s) : spec + nl * sarea;

    uint64_t tot = (uint64_t)nl * parea;
    {
        timed t0("fft_scatter", s);
        hipLaunchKernelGGL((scatter_kernel<Ti, Tt>), dim3(grid_for(tot)), dim3(256), 0, s, L, pad, H, W, PR, PC, 0u,
                           0u, tot);
        tot = (uint64_t)nr * parea;
        hipLaunchKernelGGL((scatter_kernel<Ti, Tt>), dim3(grid_for(tot)), dim3(256), 0, s, R, pad + nl * parea, H, W,
                           PR, PC, 0u, 0u, tot);
    }
    XC_HIP(hipGetLastError());

    hipfftHandle fwd = get_plan(s, PR, PC, nl + nr, fft_types<Tt>::r2c);
    fft_check(hipfftSetStream(fwd, s), "hipfftSetStream");
    {
        timed t1("rocfft_forward", s);
        fft_types<Tt>::fwd(fwd, pad, spec);
    }

    if (all_pairs) {
        tot = (uint64_t)k * sarea;
        timed t2("fft_hadamard", s);
        hipLaunchKernelGGL(hadamard_n_to_m_kernel<C>, dim3(grid_for(tot)), dim3(256), 0, s, spec, spec + nl * sarea,
                           ospec, sarea, nr, tot);
    } else {
        tot = (uint64_t)nr * sarea;
        timed t2("fft_hadamard", s);
        hipLaunchKernelGGL(hadamard_kernel<C>, dim3(grid_for(tot)), dim3(256), 0, s, spec, ospec, sarea, nl, tot);
    }
    XC_HIP(hipGetLastError());
    record_launch("fft_hadamard", dim3(grid_for(tot)), dim3(256), 0, H, W, k, 1);

    hipfftHandle inv = get_plan(s, PR, PC, k, fft_types<Tt>::c2r);
    fft_check(hipfftSetStream(inv, s), "hipfftSetStream");
    const bool direct_out = !normalize && std::is_same<Ti, Tt>::value;
    Tt* res = direct_out ? (Tt*)O : (Tt*)workspace(WS_FFT_RES, (uint64_t)k * parea * sizeof(Tt), s);
    {
        timed t3("rocfft_inverse", s);
        fft_types<Tt>::inv(inv, ospec, res);
    }

    if (normalize) {
        tot = (uint64_t)k * (PR - 1) * (PC - 1);
        timed t4("fft_normalize", s);
        hipLaunchKernelGGL((normalize_kernel<Tt, Ti>), dim3(grid_for(tot)), dim3(256), 0, s, res, O, PR, PC, tot);
    } else if (!direct_out) {
        tot = (uint64_t)k * parea;  // plain precision conversion of the unnormalised output
        hipLaunchKernelGGL((scatter_kernel<Tt, Ti>), dim3(grid_for(tot)), dim3(256), 0, s, res, O, PR, PC, PR, PC,
                           0u, 0u, tot);
    }
    XC_HIP(hipGetLastError());
}

}  // namespace

void run_fft_f32(const family_args& a, const float* L, const float* R, float* O, bool normalize, bool fp64_transform,
                 hipStream_t s) {
    if (fp64_transform)
        run_fft<float, double>(a, L, R, O, normalize, s);
    else
        run_fft<float, float>(a, L, R, O, normalize, s);
}

void run_fft_f64(const family_args& a, const double* L, const double* R, double* O, bool normalize, hipStream_t s) {
    run_fft<double, double>(a, L, R, O, normalize, s);
}

void check_scatter(const void* src, const void* dst, uint32_t src_rows, uint32_t src_cols, uint32_t dst_rows,
                   uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col) {
    if (!src || !dst) fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
    if ((uint64_t)dst_row + src_rows > dst_rows || (uint64_t)dst_col + src_cols > dst_cols)
        fail(XC_ERR_INVALID_ARGUMENT, "scattered matrix does not fit into the destination");
}

void launch_scatter_f32(const float* src, float* dst, uint32_t src_rows, uint32_t src_cols, uint32_t num,
                        uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col, hipStream_t s) {
    const uint64_t tot = (uint64_t)num * dst_rows * dst_cols;
    if (!tot) return;
    hipLaunchKernelGGL((scatter_kernel<float, float>), dim3(grid_for(tot)), dim3(256), 0, s, src, dst, src_rows,
                       src_cols, dst_rows, dst_cols, dst_row, dst_col, tot);
    XC_HIP(hipGetLastError());
}

void launch_scatter_f64(const double* src, double* dst, uint32_t src_rows, uint32_t src_cols, uint32_t num,
                        uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col, hipStream_t s) {
    const uint64_t tot = (uint64_t)num * dst_rows * dst_cols;
    if (!tot) return;
    hipLaunchKernelGGL((scatter_kernel<double, double>), dim3(grid_for(tot)), dim3(256), 0, s, src, dst, src_rows,
                       src_cols, dst_rows, dst_cols, dst_row, dst_col, tot);
    XC_HIP(hipGetLastError());
}

void launch_hadamard_f32(const float* ref, float* def, uint64_t num_complex, uint32_t batch, hipStream_t s) {
    const uint64_t tot = num_complex * batch;
    if (!tot) return;
    hipLaunchKernelGGL(hadamard_kernel<float2>, dim3(grid_for(tot)), dim3(256), 0, s, (const float2*)ref,
                       (float2*)def, num_complex, 1u, tot);
    XC_HIP(hipGetLastError());
}

void launch_hadamard_f64(const double* ref, double* def, uint64_t num_complex, uint32_t batch, hipStream_t s) {
    const uint64_t tot = num_complex * batch;
    if (!tot) return;
    hipLaunchKernelGGL(hadamard_kernel<double2>, dim3(grid_for(tot)), dim3(256), 0, s, (const double2*)ref,
                       (double2*)def, num_complex, 1u, tot);
    XC_HIP(hipGetLastError());
}

void launch_hadamard_n_to_m_f32(const float* l, const float* r, float* o, uint64_t num_complex, uint32_t n,
                                uint32_t m, hipStream_t s) {
    const uint64_t tot = num_complex * n * m;
    if (!tot) return;
    hipLaunchKernelGGL(hadamard_n_to_m_kernel<float2>, dim3(grid_for(tot)), dim3(256), 0, s, (const float2*)l,
                       (const float2*)r, (float2*)o, num_complex, m, tot);
    XC_HIP(hipGetLastError());
}

void launch_hadamard_n_to_m_f64(const double* l, const double* r, double* o, uint64_t num_complex, uint32_t n,
                                uint32_t m, hipStream_t s) {
    const uint64_t tot = num_complex * n * m;
    if (!tot) return;
    hipLaunchKernelGGL(hadamard_n_to_m_kernel<double2>, dim3(grid_for(tot)), dim3(256), 0, s, (const double2*)l,
                       (const double2*)r, (double2*)o, num_complex, m, tot);
    XC_HIP(hipGetLastError());
}

void launch_normalize_f32(const float* res, float* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                          hipStream_t s) {
    const uint64_t tot = (uint64_t)batch * (rows - 1) * (cols - 1);
    if (!tot) return;
    hipLaunchKernelGGL((normalize_kernel<float, float>), dim3(grid_for(tot)), dim3(256), 0, s, res, norm, rows, cols,
                       tot);
    XC_HIP(hipGetLastError());
}

void launch_normalize_f64(const double* res, double* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                          hipStream_t s) {
    const uint64_t tot = (uint64_t)batch * (rows - 1) * (cols - 1);
    if (!tot) return;
    hipLaunchKernelGGL((normalize_kernel<double, double>), dim3(grid_for(tot)), dim3(256), 0, s, res, norm, rows,
                       cols, tot);
    XC_HIP(hipGetLastError());
}

}  // namespace xc
