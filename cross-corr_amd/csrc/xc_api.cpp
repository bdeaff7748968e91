// C ABI of libxcorr.so (include/xcorr.h): argument validation, method dispatch and the
// conversion of library errors into xc_status + a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xc_internal.hpp"
#include "xc_methods.hpp"
#include "xc_mlp.hpp"

namespace {

thread_local std::string g_last_error;
thread_local xc_launch_desc g_last_launch{};

template <typename F>
xc_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return XC_OK;
    } catch (const xc::error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return XC_ERR_NO_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return XC_ERR_INVALID_ARGUMENT;
    } catch (...) {
        g_last_error = "unknown error";
        return XC_ERR_INVALID_ARGUMENT;
    }
}

void check_family(int family, uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, const void* l,
                  const void* r, const void* o) {
    if (family < XC_ONE_TO_ONE || family > XC_N_TO_M) xc::fail(XC_ERR_INVALID_ARGUMENT, "unknown family");
    if (rows == 0 || cols == 0) xc::fail(XC_ERR_INVALID_ARGUMENT, "matrix rows and cols must be positive");
    if (rows > 32768 || cols > 32768) xc::fail(XC_ERR_INVALID_ARGUMENT, "matrix dimension above 32768");
    if (xc::num_results(family, n, m) < 0) {
        static const char* msg[] = {
            "one_to_one needs exactly one left and one right matrix",
            "one_to_many needs exactly one left matrix and at least one right matrix",
            "n_to_mn needs the number of right matrices to be a multiple of the number of left matrices",
            "n_to_m needs at least one left and one right matrix"};
        xc::fail(XC_ERR_INVALID_ARGUMENT, msg[family]);
    }
    if (!l || !r || !o) xc::fail(XC_ERR_INVALID_ARGUMENT, "null matrix pointer");
}

}  // namespace

namespace xc {

void record_launch(const char* kernel, dim3 grid, dim3 block, uint32_t lds, uint32_t rows, uint32_t cols,
                   uint32_t pairs, uint32_t split) {
    xc_launch_desc& d = g_last_launch;
    std::memset(&d, 0, sizeof(d));
    std::strncpy(d.kernel, kernel, sizeof(d.kernel) - 1);
    d.grid[0] = grid.x;
    d.grid[1] = grid.y;
    d.grid[2] = grid.z;
    d.block[0] = block.x;
    d.block[1] = block.y;
    d.block[2] = block.z;
    d.lds_bytes = lds;
    d.rows = rows;
    d.cols = cols;
    d.pairs = pairs;
    d.split = split;
}

}  // namespace xc

extern "C" {

int64_t xc_num_results(xc_family family, uint32_t n, uint32_t m) { return xc::num_results(family, n, m); }

xc_status xc_cross_corr_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                            uint32_t cols, uint32_t n, uint32_t m, const xc_options* opt, xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc_options o{};
        if (opt) o = *opt;
        if (o.precision > XC_PREC_FP16X3)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(o.precision));
        if (o.border > 512) xc::fail(XC_ERR_INVALID_ARGUMENT, "border threshold above 512");
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        xc::run_family_f32(a, left, right, out, o, (hipStream_t)stream);
    });
}

xc_status xc_cross_corr_noalloc_f32(xc_family family, const float* left, const float* right, float* out,
                                    uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        xc::run_fftk_noalloc_f32(a, left, right, out, (hipStream_t)stream);
    });
}

xc_status xc_cross_corr_f64(xc_family family, const double* left, const double* right, double* out,
                            uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, const xc_options* opt,
                            xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc_options o{};
        if (opt) o = *opt;
        if (o.precision != XC_PREC_AUTO && o.precision != XC_PREC_FP32 && o.precision != XC_PREC_FFT64)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "double data supports the direct (FP64) and FFT64 methods only");
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        if (o.precision == XC_PREC_AUTO) o = xc::auto_options_f64(a, left, right, out, (hipStream_t)stream);
        if (o.precision == XC_PREC_FFT64 && (o.tile == 1 || o.tile == 2))
            xc::run_fftk_f64(a, left, right, out, o.split, o.tile == 2, (hipStream_t)stream);
        else if (o.precision == XC_PREC_FFT64)
            xc::run_fft_f64(a, left, right, out, true, (hipStream_t)stream);
        else
            xc::launch_direct_f64(a, left, right, out, o, (hipStream_t)stream);
    });
}

xc_status xc_batch_f32(const xc_pair* pairs, uint32_t num_pairs, const float* left, const float* right, float* out,
                       const xc_options* opt, xc_stream_t stream) {
    return guarded([&] {
        if (num_pairs == 0) return;
        if (!pairs || !left || !right || !out) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        for (uint32_t i = 0; i < num_pairs; ++i) {
            if (pairs[i].rows == 0 || pairs[i].cols == 0)
                xc::fail(XC_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + " has an empty matrix");
            if (pairs[i].rows > 32768 || pairs[i].cols > 32768)
                xc::fail(XC_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + " dimension above 32768");
        }
        xc_options o{};
        if (opt) o = *opt;
        xc::run_batch_f32(pairs, num_pairs, left, right, out, o, (hipStream_t)stream);
    });
}

xc_status xc_batch_plan_create_f32(const xc_pair* pairs, uint32_t num_pairs, const xc_options* opt,
                                   xc_batch_plan** plan) {
    return guarded([&] {
        if (!plan) xc::fail(XC_ERR_INVALID_ARGUMENT, "null plan output pointer");
        *plan = nullptr;
        xc_options o{};
        if (opt) o = *opt;
        *plan = xc::batch_plan_create(pairs, num_pairs, o);
    });
}

xc_status xc_batch_plan_exec_f32(xc_batch_plan* plan, const float* left, const float* right, float* out, int use_graph,
                                 xc_stream_t stream) {
    return guarded([&] { xc::batch_plan_exec(plan, left, right, out, (hipStream_t)stream, use_graph != 0); });
}

xc_status xc_batch_plan_destroy(xc_batch_plan* plan) {
    return guarded([&] { xc::batch_plan_destroy(plan); });
}

xc_status xc_batch_autotune_f32(const xc_pair* pairs, uint32_t num_pairs, const float* left, const float* right,
                                float* out, xc_options* candidates, float* ms, uint32_t capacity, uint32_t* count,
                                xc_options* best, xc_stream_t stream) {
    return guarded([&] {
        if (!pairs || !left || !right || !out) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        std::vector<xc_options> c;
        std::vector<float> t;
        const uint32_t b = xc::batch_autotune_f32(pairs, num_pairs, left, right, out, (hipStream_t)stream, c, t);
        for (uint32_t i = 0; i < capacity && i < c.size(); ++i) {
            if (candidates) candidates[i] = c[i];
            if (ms) ms[i] = t[i];
        }
        if (count) *count = (uint32_t)c.size();
        if (best) *best = c[b];
    });
}

xc_status xc_scatter_f32(const float* src, float* dst, uint32_t src_rows, uint32_t src_cols, uint32_t num_matrices,
                         uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col,
                         xc_stream_t stream) {
    return guarded([&] {
        xc::check_scatter(src, dst, src_rows, src_cols, dst_rows, dst_cols, dst_row, dst_col);
        xc::launch_scatter_f32(src, dst, src_rows, src_cols, num_matrices, dst_rows, dst_cols, dst_row, dst_col,
                               (hipStream_t)stream);
    });
}

xc_status xc_scatter_f64(const double* src, double* dst, uint32_t src_rows, uint32_t src_cols,
                         uint32_t num_matrices, uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row,
                         uint32_t dst_col, xc_stream_t stream) {
    return guarded([&] {
        xc::check_scatter(src, dst, src_rows, src_cols, dst_rows, dst_cols, dst_row, dst_col);
        xc::launch_scatter_f64(src, dst, src_rows, src_cols, num_matrices, dst_rows, dst_cols, dst_row, dst_col,
                               (hipStream_t)stream);
    });
}

xc_status xc_hadamard_f32(const float* ref, float* deformed, uint64_t num_complex, uint32_t batch,
                          xc_stream_t stream) {
    return guarded([&] {
        if (!ref || !deformed) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::launch_hadamard_f32(ref, deformed, num_complex, batch, (hipStream_t)stream);
    });
}

xc_status xc_hadamard_f64(const double* ref, double* deformed, uint64_t num_complex, uint32_t batch,
                          xc_stream_t stream) {
    return guarded([&] {
        if (!ref || !deformed) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::launch_hadamard_f64(ref, deformed, num_complex, batch, (hipStream_t)stream);
    });
}

xc_status xc_hadamard_n_to_m_f32(const float* left, const float* right, float* out, uint64_t num_complex,
                                 uint32_t n, uint32_t m, xc_stream_t stream) {
    return guarded([&] {
        if (!left || !right || !out) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::launch_hadamard_n_to_m_f32(left, right, out, num_complex, n, m, (hipStream_t)stream);
    });
}

xc_status xc_hadamard_n_to_m_f64(const double* left, const double* right, double* out, uint64_t num_complex,
                                 uint32_t n, uint32_t m, xc_stream_t stream) {
    return guarded([&] {
        if (!left || !right || !out) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::launch_hadamard_n_to_m_f64(left, right, out, num_complex, n, m, (hipStream_t)stream);
    });
}

xc_status xc_normalize_fft_f32(const float* res, float* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                               xc_stream_t stream) {
    return guarded([&] {
        if (!res || !norm || rows < 2 || cols < 2) xc::fail(XC_ERR_INVALID_ARGUMENT, "invalid normalize arguments");
        xc::launch_normalize_f32(res, norm, rows, cols, batch, (hipStream_t)stream);
    });
}

xc_status xc_normalize_fft_f64(const double* res, double* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                               xc_stream_t stream) {
    return guarded([&] {
        if (!res || !norm || rows < 2 || cols < 2) xc::fail(XC_ERR_INVALID_ARGUMENT, "invalid normalize arguments");
        xc::launch_normalize_f64(res, norm, rows, cols, batch, (hipStream_t)stream);
    });
}

xc_status xc_fft_cross_corr_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                                uint32_t cols, uint32_t n, uint32_t m, int normalize, int fp64_transform,
                                xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        xc::run_fft_f32(a, left, right, out, normalize != 0, fp64_transform != 0, (hipStream_t)stream);
    });
}

xc_status xc_fft_cross_corr_f64(xc_family family, const double* left, const double* right, double* out,
                                uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int normalize,
                                xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        xc::run_fft_f64(a, left, right, out, normalize != 0, (hipStream_t)stream);
    });
}

// run_ccn_shuffle (src/naive_shuffle.cu:452-484): one_to_one, one shift per thread;
// warps_per_thread_block in [1, 32] as the reference checks it (:461-463), two 32-lane warps
// per 64-lane wave of the block.
xc_status xc_run_ccn_shuffle_f32(const float* left, const float* right, float* out, uint32_t cols, uint32_t rows,
                                 uint32_t search_cols, uint32_t search_rows, uint32_t warps_per_thread_block,
                                 xc_stream_t stream) {
    return guarded([&] {
        if (warps_per_thread_block == 0 || warps_per_thread_block > 32)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "Too many warps per thread block: " +
                                                  std::to_string(warps_per_thread_block) + " (max 32)");
        if (search_cols != 2 * cols - 1 || search_rows != 2 * rows - 1)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "search size must be the full (2*rows-1) x (2*cols-1)");
        check_family(XC_ONE_TO_ONE, rows, cols, 1, 1, left, right, out);
        xc::family_args a{XC_ONE_TO_ONE, rows, cols, 1, 1, 1};
        xc_options o{};
        o.shifts_per_thread = o.right_mats_per_thread = o.left_mats_per_thread = o.rows_per_step = 1;
        o.waves_per_block = (warps_per_thread_block + 1) / 2;
        xc::launch_direct_f32(a, left, right, out, o, (hipStream_t)stream);
    });
}

// run_ccn_shuffle_one_to_many_multirow_both_multimat_right
// (src/naive_shuffle_one_to_many_multirow_both_multimat_right.cu:776-833): the reference's
// argument checks with the limits of its release build (:789-818, CMakeLists.txt:41-43);
// the values select the direct kernel's tile: shifts_per_thread_right_matrix consecutive dy,
// right_matrices_per_thread matrices and left_rows_per_iteration right rows per step.
xc_status xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
    const float* left, const float* right, float* out, uint32_t cols, uint32_t rows, uint32_t search_cols,
    uint32_t search_rows, uint32_t num_right_matrices, uint32_t warps_per_thread_block,
    uint32_t shifts_per_thread_right_matrix, uint32_t right_matrices_per_thread, uint32_t left_rows_per_iteration,
    xc_stream_t stream) {
    return guarded([&] {
        if (warps_per_thread_block == 0 || warps_per_thread_block > 32)
            xc::fail(XC_ERR_INVALID_ARGUMENT,
                     "Cuda block too large: " + std::to_string(warps_per_thread_block) + " (max 32)");
        if (shifts_per_thread_right_matrix == 0 || shifts_per_thread_right_matrix > 4)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "Too many shifts per thread in each right matrix: " +
                                                  std::to_string(shifts_per_thread_right_matrix) + " (max 4)");
        if (right_matrices_per_thread == 0 || right_matrices_per_thread > 4)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "Invalid number of right matrices per thread: " +
                                                  std::to_string(right_matrices_per_thread) + " [1-4]");
        if (left_rows_per_iteration == 0 || left_rows_per_iteration > 4)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "Too many left rows per iteration: " +
                                                  std::to_string(left_rows_per_iteration) + " (max 4)");
        if (search_cols != 2 * cols - 1 || search_rows != 2 * rows - 1)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "search size must be the full (2*rows-1) x (2*cols-1)");
        check_family(XC_ONE_TO_MANY, rows, cols, 1, num_right_matrices, left, right, out);
        xc::family_args a{XC_ONE_TO_MANY, rows, cols, 1, num_right_matrices, num_right_matrices};
        xc_options o{};
        o.shifts_per_thread = shifts_per_thread_right_matrix;
        o.right_mats_per_thread = right_matrices_per_thread;
        o.left_mats_per_thread = 1;
        o.rows_per_step = left_rows_per_iteration;
        o.waves_per_block = (warps_per_thread_block + 1) / 2;
        xc::launch_direct_f32(a, left, right, out, o, (hipStream_t)stream);
    });
}

// ---- method selection and the cost model's MLP
uint32_t xc_num_candidates(void) { return xc::num_candidates(); }

xc_status xc_candidate(uint32_t i, xc_options* opt, char* name, size_t name_len) {
    return guarded([&] {
        if (!opt) xc::fail(XC_ERR_INVALID_ARGUMENT, "null options pointer");
        *opt = xc::candidate(i);
        if (name && name_len) {
            const char* p = opt->precision == XC_PREC_FP32 ? "fp32" : opt->precision == XC_PREC_BF16X3 ? "bf16x3" : "fft64";
            std::snprintf(name, name_len, "%s/%u", p, opt->tile);
        }
    });
}

xc_status xc_workload_features(const xc_workload* w, float* features, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        if (!w || !features) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        const uint32_t n = xc::workload_features(*w, features, capacity);
        if (count) *count = n;
    });
}

xc_status xc_autotune_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                          uint32_t cols, uint32_t n, uint32_t m, uint32_t reps, float* ms, uint32_t* best,
                          xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        const uint32_t b = xc::autotune_f32(a, left, right, out, reps, ms, (hipStream_t)stream);
        if (best) *best = b;
    });
}

xc_status xc_time_candidate_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                                uint32_t cols, uint32_t n, uint32_t m, uint32_t index, uint32_t reps, float* ms,
                                xc_stream_t stream) {
    return guarded([&] {
        check_family(family, rows, cols, n, m, left, right, out);
        if (index >= xc::num_candidates()) xc::fail(XC_ERR_INVALID_ARGUMENT, "candidate index out of range");
        if (!ms) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::family_args a{family, rows, cols, n, m, (uint32_t)xc::num_results(family, n, m)};
        *ms = xc::time_candidate_f32(a, left, right, out, index, reps, (hipStream_t)stream);
    });
}

xc_status xc_autotune_clear(void) {
    return guarded([&] { xc::autotune_clear(); });
}

xc_status xc_cost_model_load(const char* path) {
    return guarded([&] { xc::cost_model_load(path); });
}

xc_status xc_cost_model_predict(const xc_workload* w, float* ms) {
    return guarded([&] {
        if (!w || !ms) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::cost_model_predict(*w, ms);
    });
}

xc_status xc_select(const xc_workload* w, xc_options* chosen, int* source) {
    return guarded([&] {
        if (!w || !chosen) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc_options c{};
        const int src = xc::select_method(*w, c);
        *chosen = c;
        if (source) *source = src;
    });
}

xc_status xc_cost_model_fit(const char* timing_csv, const char* model_path, const xc_train_options* opt,
                            uint32_t holdout_mod) {
    return guarded([&] {
        xc_train_options o{};
        o.epochs = 3000;
        o.lr = 3e-3f;
        o.seed = 1;
        o.use_graph = 1;
        if (opt) o = *opt;
        xc::cost_model_fit(timing_csv, model_path, o, holdout_mod);
    });
}

uint64_t xc_mlp_num_params(const xc_mlp_shape* shape) {
    try {
        xc::mlp_check_shape(shape);
        return xc::mlp_num_params(*shape);
    } catch (...) {
        return 0;
    }
}

uint64_t xc_mlp_work_size(const xc_mlp_shape* shape, uint32_t batch) {
    try {
        xc::mlp_check_shape(shape);
        return xc::mlp_work_size(*shape, batch);
    } catch (...) {
        return 0;
    }
}

xc_status xc_mlp_init(const xc_mlp_shape* shape, uint64_t seed, float* params_host) {
    return guarded([&] {
        xc::mlp_check_shape(shape);
        if (!params_host) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::mlp_init(*shape, seed, params_host);
    });
}

xc_status xc_mlp_forward_f32(const xc_mlp_shape* shape, const float* params, const float* x, uint32_t batch,
                             float* y, float* work, xc_stream_t stream) {
    return guarded([&] {
        xc::mlp_check_shape(shape);
        if (!params || !x || !y || !work) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        if (batch == 0) return;
        xc::mlp_forward(*shape, params, x, batch, y, work, (hipStream_t)stream);
    });
}

xc_status xc_mlp_loss_grad_f32(const xc_mlp_shape* shape, const float* params, const float* x, const float* target,
                               const float* mask, uint32_t batch, float* loss, float* grad, float* work,
                               xc_stream_t stream) {
    return guarded([&] {
        xc::mlp_check_shape(shape);
        if (!params || !x || !target || !mask || !loss || !grad || !work)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        if (batch == 0) xc::fail(XC_ERR_INVALID_ARGUMENT, "empty batch");
        xc::mlp_loss_grad(*shape, params, x, target, mask, batch, loss, grad, work, (hipStream_t)stream);
    });
}

xc_status xc_mlp_adam_f32(uint64_t count, float* params, const float* grad, float* m, float* v, uint32_t step,
                          float lr, float beta1, float beta2, float eps, float weight_decay, xc_stream_t stream) {
    return guarded([&] {
        if (!params || !grad || !m || !v) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        if (step == 0) xc::fail(XC_ERR_INVALID_ARGUMENT, "Adam steps are numbered from 1");
        if (count == 0) return;
        xc::mlp_adam(count, params, grad, m, v, step, lr, beta1, beta2, eps, weight_decay, (hipStream_t)stream);
    });
}

xc_status xc_mlp_train_f32(const xc_mlp_shape* shape, const float* x_host, const float* target_host,
                           const float* mask_host, uint32_t batch, const xc_train_options* opt, float* params_host,
                           float* loss_host) {
    return guarded([&] {
        xc::mlp_check_shape(shape);
        if (!x_host || !target_host || !mask_host || !params_host || !opt)
            xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::mlp_train(*shape, x_host, target_host, mask_host, batch, *opt, params_host, loss_host);
    });
}

xc_status xc_kernel_timing(int enable) {
    return guarded([&] { xc::timing_enable(enable != 0); });
}

xc_status xc_kernel_times(char* buf, size_t len, size_t* needed) {
    return guarded([&] {
        const std::string r = xc::timing_report();
        if (needed) *needed = r.size() + 1;
        if (buf && len) {
            std::strncpy(buf, r.c_str(), len - 1);
            buf[len - 1] = 0;
        }
    });
}

const char* xc_last_error(void) { return g_last_error.c_str(); }

const xc_launch_desc* xc_last_launch(void) { return &g_last_launch; }

const char* xc_version(void) { return "xcorr-mi355x 0.1 (gfx950)"; }

xc_status xc_device_info(char* buf, size_t len) {
    return guarded([&] {
        int dev = 0;
        XC_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        XC_HIP(hipGetDeviceProperties(&p, dev));
        std::string s = std::string(p.name) + " " + p.gcnArchName + " CUs=" + std::to_string(p.multiProcessorCount) +
                        " clock_khz=" + std::to_string(p.clockRate) + " mem_gib=" +
                        std::to_string(p.totalGlobalMem >> 30);
        if (buf && len) {
            std::strncpy(buf, s.c_str(), len - 1);
            buf[len - 1] = 0;
        }
    });
}

xc_status xc_validate_result_f32(const float* result, const float* valid, uint64_t count, double* out,
                                 xc_stream_t stream) {
    return guarded([&] {
        if (!out || (count && (!result || !valid))) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::validate_result_f32(result, valid, count, out, (hipStream_t)stream);
    });
}

xc_status xc_validate_result_f64(const double* result, const double* valid, uint64_t count, double* out,
                                 xc_stream_t stream) {
    return guarded([&] {
        if (!out || (count && (!result || !valid))) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::validate_result_f64(result, valid, count, out, (hipStream_t)stream);
    });
}

xc_status xc_norm_error_f32(const float* x, const double* ref, const double* den, uint64_t count, double* out,
                            xc_stream_t stream) {
    return guarded([&] {
        if (!out || (count && (!x || !ref || !den))) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::norm_error_f32(x, ref, den, count, out, (hipStream_t)stream);
    });
}

xc_status xc_norm_error_f64(const double* x, const double* ref, const double* den, uint64_t count, double* out,
                            xc_stream_t stream) {
    return guarded([&] {
        if (!out || (count && (!x || !ref || !den))) xc::fail(XC_ERR_INVALID_ARGUMENT, "null pointer");
        xc::norm_error_f64(x, ref, den, count, out, (hipStream_t)stream);
    });
}

}  // extern "C"
