// C++ re-creation of the launcher call shapes of the reference's src/kernels.cuh on top of
// the C ABI (xcorr.h): function templates in namespace cross that throw std::runtime_error
// when the library reports an error, so code written against kernels.cuh (src/main.cpp,
// sample_app/src/main.cpp) keeps its structure.  Pointers are device pointers, sizes are
// {x = cols, y = rows} as the reference's dsize2_t (src/types.cuh:154).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "xcorr.h"

namespace cross {

using dsize_t = unsigned int;

struct dsize2_t {
    dsize_t x, y;
};

namespace detail {
inline void check(xc_status s) {
    if (s != XC_OK) throw std::runtime_error(std::string("xcorr: ") + xc_last_error());
}
inline xc_options tune(uint32_t split = 0) {
    xc_options o{};
    o.precision = XC_PREC_FP32;
    o.split = split;
    return o;
}
template <typename T>
inline xc_status family(xc_family f, const T* l, const T* r, T* o, dsize2_t size, dsize_t n, dsize_t m,
                        const xc_options& opt, xc_stream_t s) {
    if constexpr (sizeof(T) == 4)
        return xc_cross_corr_f32(f, (const float*)l, (const float*)r, (float*)o, size.y, size.x, n, m, &opt, s);
    else
        return xc_cross_corr_f64(f, (const double*)l, (const double*)r, (double*)o, size.y, size.x, n, m, &opt, s);
}
inline void check_search(dsize2_t matrix_size, dsize2_t search_size) {
    if (search_size.x != 2 * matrix_size.x - 1 || search_size.y != 2 * matrix_size.y - 1)
        throw std::runtime_error("search size must be the full (2*rows-1) x (2*cols-1) result");
}
inline void check_range(const char* what, dsize_t v, dsize_t max) {
    if (v == 0 || v > max) throw std::runtime_error(std::string("Invalid ") + what + ": " + std::to_string(v));
}
}  // namespace detail

// run_ccn_shuffle, src/kernels.cuh:65-73
template <typename T, typename RES>
void run_ccn_shuffle(const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size,
                     dsize_t warps_per_thread_block, xc_stream_t stream = nullptr) {
    static_assert(sizeof(T) == sizeof(RES), "result type must match the input type");
    detail::check_search(matrix_size, search_size);
    detail::check_range("warps per thread block", warps_per_thread_block, 32);
    detail::check(detail::family<T>(XC_ONE_TO_ONE, left, right, (T*)out, matrix_size, 1, 1, detail::tune(), stream));
}

// run_ccn_shuffle_work_distribution, src/kernels.cuh:75-85 (DIST is accepted for source
// compatibility; the split is a plain partition of the overlap rows)
template <typename DIST, typename T, typename RES>
void run_ccn_shuffle_work_distribution(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                       dsize2_t search_size, dsize_t warps_per_thread_block,
                                       dsize_t max_rows_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_range("warps per thread block", warps_per_thread_block, 32);
    detail::check_range("rows per thread", max_rows_per_thread, 1u << 30);
    const auto opt = detail::tune((matrix_size.y + max_rows_per_thread - 1) / max_rows_per_thread);
    detail::check(detail::family<T>(XC_ONE_TO_ONE, left, right, (T*)out, matrix_size, 1, 1, opt, stream));
}

// run_ccn_shuffle_multimat_right, src/kernels.cuh:87-97
template <typename T, typename RES>
void run_ccn_shuffle_multimat_right(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                    dsize2_t search_size, dsize_t num_right_matrices, dsize_t warps_per_thread_block,
                                    dsize_t right_matrices_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_range("warps per thread block", warps_per_thread_block, 32);
    detail::check_range("right matrices per thread", right_matrices_per_thread, 1u << 30);
    detail::check(detail::family<T>(XC_ONE_TO_MANY, left, right, (T*)out, matrix_size, 1, num_right_matrices,
                                    detail::tune(), stream));
}

// run_ccn_shuffle_one_to_many_multirow_both_multimat_right, src/kernels.cuh:223-236
template <typename T, typename RES>
void run_ccn_shuffle_one_to_many_multirow_both_multimat_right(
    const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size, dsize_t num_right_matrices,
    dsize_t warps_per_thread_block, dsize_t shifts_per_thread_right_matrix, dsize_t right_matrices_per_thread,
    dsize_t left_rows_per_iteration, xc_stream_t stream = nullptr) {
    if constexpr (sizeof(T) == 4) {
        detail::check(xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
            (const float*)left, (const float*)right, (float*)out, matrix_size.x, matrix_size.y, search_size.x,
            search_size.y, num_right_matrices, warps_per_thread_block, shifts_per_thread_right_matrix,
            right_matrices_per_thread, left_rows_per_iteration, stream));
    } else {
        detail::check_search(matrix_size, search_size);
        detail::check_range("warps per thread block", warps_per_thread_block, 32);
        detail::check(detail::family<T>(XC_ONE_TO_MANY, left, right, (T*)out, matrix_size, 1, num_right_matrices,
                                        detail::tune(), stream));
    }
}

// run_ccn_n_to_m_shuffle_multirow_both_multimat_both, src/kernels.cuh:238-254
template <typename T, typename RES>
void run_ccn_n_to_m_shuffle_multirow_both_multimat_both(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                                        dsize2_t search_size, dsize_t num_left_matrices,
                                                        dsize_t num_right_matrices, dsize_t warps_per_thread_block,
                                                        dsize_t, dsize_t, dsize_t, dsize_t,
                                                        xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_range("warps per thread block", warps_per_thread_block, 32);
    detail::check(detail::family<T>(XC_N_TO_M, left, right, (T*)out, matrix_size, num_left_matrices,
                                    num_right_matrices, detail::tune(), stream));
}

// run_scatter, src/kernels.cuh:53-63 (zero padding written by the kernel itself)
template <typename T>
void run_scatter(const T* src, T* dst, dsize2_t src_matrix_size, dsize_t src_num_matrices, dsize2_t dst_matrix_size,
                 dsize2_t dst_pos, dsize_t = 256, dsize_t = 10, xc_stream_t stream = nullptr) {
    if constexpr (sizeof(T) == 4)
        detail::check(xc_scatter_f32((const float*)src, (float*)dst, src_matrix_size.y, src_matrix_size.x,
                                     src_num_matrices, dst_matrix_size.y, dst_matrix_size.x, dst_pos.y, dst_pos.x,
                                     stream));
    else
        detail::check(xc_scatter_f64((const double*)src, (double*)dst, src_matrix_size.y, src_matrix_size.x,
                                     src_num_matrices, dst_matrix_size.y, dst_matrix_size.x, dst_pos.y, dst_pos.x,
                                     stream));
}

}  // namespace cross
