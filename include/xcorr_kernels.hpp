// C++ re-creation of the launcher interface of the reference's src/kernels.cuh on top of the
// C ABI (xcorr.h): the same function templates in namespace cross, with the same argument
// lists, so code written against kernels.cuh (src/main.cpp, src/one_to_many.hpp,
// sample_app/src/main.cpp) compiles with this header instead.  Pointers are device
// pointers; sizes are {x = cols, y = rows} as the reference's dsize2_t (src/types.cuh:154).
//
// Each launcher checks its tuning arguments against the limits of the reference's release
// build (CMakeLists.txt:16-45, the checks in each src/naive_*.cu run_* function) and throws
// std::runtime_error for a rejected value, as the reference does; a zero value, which the
// reference's dispatch turns into a silent no-op, is rejected too.  Accepted values select the
// library's direct-kernel tile (xc_options: shifts, right and left matrices per thread, rows
// per step, waves per block, work distribution), rounded down to the compiled tile catalog.
// Library failures (launch errors included) are thrown as std::runtime_error; the reference
// leaves launch errors to the next synchronising call.  The optional trailing stream argument
// defaults to the null stream, as the reference's cudaStream_t arguments do.
//
// FFT helpers take complex arrays: T is any 8-byte (float2, hipfftComplex) or 16-byte
// (double2, hipfftDoubleComplex) interleaved complex type.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "xcorr.h"

namespace cross {

using dsize_t = unsigned int;

struct dsize2_t {
    dsize_t x, y;
    dsize_t area() const { return x * y; }
};

// Work-distribution policies of src/row_distribution.cuh:34-284, as tags.
struct triangle_distribution {};   // chunks of at most max_rows_per_thread overlap rows
struct rectangle_distribution {};  // ceil(rows / max_rows_per_thread) equal shares per output row
struct no_distribution {};         // one worker per output row

namespace detail {

inline void check(xc_status s) {
    if (s != XC_OK) throw std::runtime_error(std::string("xcorr: ") + xc_last_error());
}

// at most `max` and not zero, else "<what>: <v> (max <max>)"
inline void limit(const char* what, dsize_t v, dsize_t max) {
    if (v == 0 || v > max)
        throw std::runtime_error(std::string(what) + ": " + std::to_string(v) + " (max " + std::to_string(max) + ")");
}

inline void check_search(dsize2_t matrix_size, dsize2_t search_size) {
    if (search_size.x != 2 * matrix_size.x - 1 || search_size.y != 2 * matrix_size.y - 1)
        throw std::runtime_error("search size must be the full (2*rows-1) x (2*cols-1) result");
}

inline void check_warps(dsize_t warps_per_thread_block) {
    limit("Too many warps per thread block", warps_per_thread_block, 32);
}

// Direct-kernel options: a 32-lane "warp" of the reference is half a 64-lane wave.
inline xc_options direct(dsize_t warps_per_thread_block, dsize_t shifts, dsize_t right_mats, dsize_t left_mats,
                         dsize_t rows_per_step) {
    xc_options o{};
    o.precision = XC_PREC_FP32;
    o.waves_per_block = (warps_per_thread_block + 1) / 2;
    o.shifts_per_thread = shifts;
    o.right_mats_per_thread = right_mats;
    o.left_mats_per_thread = left_mats;
    o.rows_per_step = rows_per_step;
    return o;
}

// Work distribution over the overlap rows of each output row (matrix_rows = H).
template <typename DIST>
inline void distribute(xc_options& o, dsize_t max_rows_per_thread, dsize_t matrix_rows) {
    limit("Invalid number of rows per thread", max_rows_per_thread, 1u << 30);
    if constexpr (std::is_same_v<DIST, triangle_distribution>) {
        o.rows_per_block = max_rows_per_thread;
    } else if constexpr (std::is_same_v<DIST, rectangle_distribution>) {
        o.split = (matrix_rows + max_rows_per_thread - 1) / max_rows_per_thread;
    } else {
        static_assert(std::is_same_v<DIST, no_distribution>, "unknown work distribution");
    }
}

template <typename T, typename RES>
inline void family_call(xc_family f, const T* l, const T* r, RES* o, dsize2_t size, dsize_t n, dsize_t m,
                        const xc_options& opt, xc_stream_t s) {
    static_assert(std::is_same_v<T, RES>, "result type must match the input type");
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double matrices");
    if constexpr (std::is_same_v<T, float>)
        check(xc_cross_corr_f32(f, l, r, o, size.y, size.x, n, m, &opt, s));
    else
        check(xc_cross_corr_f64(f, l, r, o, size.y, size.x, n, m, &opt, s));
}

template <typename C>
constexpr bool is_single_complex() {
    static_assert(sizeof(C) == 8 || sizeof(C) == 16, "interleaved complex of float or double");
    return sizeof(C) == 8;
}

}  // namespace detail

// ---- FFT helpers ----------------------------------------------------------------------

// run_hadamard_original, src/kernels.cuh:7-15: deformed[b][i] = ref[i] * conj(deformed[b][i])
// for the subregion_size.area() * subregions_per_pic values of each of batch_size pictures.
template <typename T>
void run_hadamard_original(const T* ref, T* deformed, dsize2_t subregion_size, dsize_t subregions_per_pic,
                           dsize_t batch_size, dsize_t num_threads, xc_stream_t stream = nullptr) {
    detail::limit("Invalid number of threads per block", num_threads, 1024);
    const uint64_t count = (uint64_t)subregion_size.area() * subregions_per_pic;
    if constexpr (detail::is_single_complex<T>())
        detail::check(xc_hadamard_f32((const float*)ref, (float*)deformed, count, batch_size, stream));
    else
        detail::check(xc_hadamard_f64((const double*)ref, (double*)deformed, count, batch_size, stream));
}

namespace detail {
template <typename T>
void hadamard_n_to_m(const T* left, const T* right, T* out, dsize2_t matrix_size, dsize_t left_num,
                     dsize_t right_num, dsize_t threads_per_block, dsize_t min_items_per_thread, xc_stream_t stream) {
    limit("Invalid number of threads per block", threads_per_block, 1024);
    limit("Invalid number of items per thread", min_items_per_thread, 1u << 30);
    const uint64_t count = matrix_size.area();
    if constexpr (is_single_complex<T>())
        check(xc_hadamard_n_to_m_f32((const float*)left, (const float*)right, (float*)out, count, left_num, right_num,
                                     stream));
    else
        check(xc_hadamard_n_to_m_f64((const double*)left, (const double*)right, (double*)out, count, left_num,
                                     right_num, stream));
}
}  // namespace detail

// run_hadamard_n_to_m_over_right, src/kernels.cuh:17-27, and run_hadamard_n_to_m_over_output,
// src/kernels.cuh:29-39: out[l * right_num + r][i] = left[l][i] * conj(right[r][i]).  The two
// reference kernels differ only in how threads are assigned; here both run the one coalesced
// kernel behind xc_hadamard_n_to_m_*.
template <typename T>
void run_hadamard_n_to_m_over_right(const T* left, const T* right, T* out, dsize2_t matrix_size, dsize_t left_num,
                                    dsize_t right_num, dsize_t threads_per_block, dsize_t min_items_per_thread,
                                    xc_stream_t stream = nullptr) {
    detail::hadamard_n_to_m(left, right, out, matrix_size, left_num, right_num, threads_per_block,
                            min_items_per_thread, stream);
}

template <typename T>
void run_hadamard_n_to_m_over_output(const T* left, const T* right, T* out, dsize2_t matrix_size, dsize_t left_num,
                                     dsize_t right_num, dsize_t threads_per_block, dsize_t min_items_per_thread,
                                     xc_stream_t stream = nullptr) {
    detail::hadamard_n_to_m(left, right, out, matrix_size, left_num, right_num, threads_per_block,
                            min_items_per_thread, stream);
}

// run_scatter, src/kernels.cuh:53-63: src matrices into the top-left (at dst_pos) of dst
// matrices; the zero padding is written by the kernel itself (no prior memset needed).
template <typename T>
void run_scatter(const T* src, T* dst, dsize2_t src_matrix_size, dsize_t src_num_matrices, dsize2_t dst_matrix_size,
                 dsize2_t dst_pos, dsize_t threads_per_block = 256, dsize_t items_per_threads = 10,
                 xc_stream_t stream = nullptr) {
    detail::limit("Invalid number of threads per block", threads_per_block, 1024);
    detail::limit("Invalid number of items per thread", items_per_threads, 1u << 30);
    if constexpr (std::is_same_v<T, float>)
        detail::check(xc_scatter_f32(src, dst, src_matrix_size.y, src_matrix_size.x, src_num_matrices,
                                     dst_matrix_size.y, dst_matrix_size.x, dst_pos.y, dst_pos.x, stream));
    else
        detail::check(xc_scatter_f64(src, dst, src_matrix_size.y, src_matrix_size.x, src_num_matrices,
                                     dst_matrix_size.y, dst_matrix_size.x, dst_pos.y, dst_pos.x, stream));
}

// ---- definition-based launchers --------------------------------------------------------

// run_cross_corr_naive_original, src/kernels.cuh:41-51: ref holds subregions_per_pic matrices,
// deformed batch_size pictures of subregions_per_pic matrices each; result b * spp + s is
// corr(ref[s], deformed[b * spp + s]) -- the n_to_mn ordering with n = spp.  One shift per
// thread, as the reference's kernel.
template <typename T, typename RES>
void run_cross_corr_naive_original(const T* ref, const T* deformed, RES* out, dsize2_t subregion_size,
                                   dsize2_t search_size, dsize_t subregions_per_pic, dsize_t batch_size,
                                   xc_stream_t stream = nullptr) {
    detail::check_search(subregion_size, search_size);
    detail::limit("Invalid number of subregions per picture", subregions_per_pic, 1u << 30);
    detail::limit("Invalid batch size", batch_size, 1u << 30);
    detail::family_call(XC_N_TO_MN, ref, deformed, out, subregion_size, subregions_per_pic,
                        subregions_per_pic * batch_size, detail::direct(16, 1, 1, 1, 1), stream);
}

// run_ccn_shuffle, src/kernels.cuh:65-73 (src/naive_shuffle.cu:452-484): one shift per thread.
template <typename T, typename RES>
void run_ccn_shuffle(const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size,
                     dsize_t warps_per_thread_block, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1,
                        detail::direct(warps_per_thread_block, 1, 1, 1, 1), stream);
}

// run_ccn_shuffle_work_distribution, src/kernels.cuh:75-85 (src/naive_shuffle.cu:487-527).
template <typename DIST, typename T, typename RES>
void run_ccn_shuffle_work_distribution(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                       dsize2_t search_size, dsize_t warps_per_thread_block,
                                       dsize_t max_rows_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    auto o = detail::direct(warps_per_thread_block, 1, 1, 1, 1);
    detail::distribute<DIST>(o, max_rows_per_thread, matrix_size.y);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1, o, stream);
}

// run_ccn_shuffle_multimat_right, src/kernels.cuh:87-97 (src/naive_shuffle_multimat_right.cu:705-743).
template <typename T, typename RES>
void run_ccn_shuffle_multimat_right(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                    dsize2_t search_size, dsize_t num_right_matrices, dsize_t warps_per_thread_block,
                                    dsize_t right_matrices_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Invalid number of right matrices per thread", right_matrices_per_thread, 8);
    detail::family_call(XC_ONE_TO_MANY, left, right, out, matrix_size, 1, num_right_matrices,
                        detail::direct(warps_per_thread_block, 1, right_matrices_per_thread, 1, 1), stream);
}

// run_ccn_warp_per_shift, src/kernels.cuh:99-107, and its simple-indexing variant (:109-117):
// the reference reduces each shift over a warp; here they run the direct kernel with one
// shift per thread and shifts_per_thread_block / 2 waves.
template <typename T, typename RES>
void run_ccn_warp_per_shift(const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size,
                            dsize_t shifts_per_thread_block, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::limit("Too many shifts per thread block", shifts_per_thread_block, 32);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1,
                        detail::direct(shifts_per_thread_block, 1, 1, 1, 1), stream);
}

template <typename T, typename RES>
void run_ccn_warp_per_shift_simple_indexing(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                            dsize2_t search_size, dsize_t shifts_per_thread_block,
                                            xc_stream_t stream = nullptr) {
    run_ccn_warp_per_shift(left, right, out, matrix_size, search_size, shifts_per_thread_block, stream);
}

// run_ccn_warp_per_shift_work_distribution, src/kernels.cuh:119-129.
template <typename DIST, typename T, typename RES>
void run_ccn_warp_per_shift_work_distribution(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                              dsize2_t search_size, dsize_t shifts_per_thread_block,
                                              dsize_t max_rows_per_warp, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::limit("Too many shifts per thread block", shifts_per_thread_block, 32);
    auto o = detail::direct(shifts_per_thread_block, 1, 1, 1, 1);
    detail::distribute<DIST>(o, max_rows_per_warp, matrix_size.y);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1, o, stream);
}

// run_ccn_shuffle_multimat_right_work_distribution, src/kernels.cuh:131-143
// (src/naive_shuffle_multimat_right.cu:745-780).
template <typename DIST, typename T, typename RES>
void run_ccn_shuffle_multimat_right_work_distribution(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                                      dsize2_t search_size, dsize_t num_right_matrices,
                                                      dsize_t warps_per_thread_block,
                                                      dsize_t right_matrices_per_thread, dsize_t max_rows_per_thread,
                                                      xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Too many right matrices per thread", right_matrices_per_thread, 8);
    auto o = detail::direct(warps_per_thread_block, 1, right_matrices_per_thread, 1, 1);
    detail::distribute<DIST>(o, max_rows_per_thread, matrix_size.y);
    detail::family_call(XC_ONE_TO_MANY, left, right, out, matrix_size, 1, num_right_matrices, o, stream);
}

// run_ccn_warp_per_shift_shared_mem, src/kernels.cuh:145-160: one_to_many; the reference
// stages right matrices in shared memory per block, which this library's kernel always does.
// shared_mem_row_size, shared_mem_rows, strided_load and column_group_per_block have no
// counterpart and are not read.
template <typename T, typename RES>
void run_ccn_warp_per_shift_shared_mem(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                       dsize2_t search_size, dsize_t num_right_matrices, dsize_t shifts_per_cuda_block,
                                       dsize_t /*shared_mem_row_size*/, dsize_t /*shared_mem_rows*/,
                                       dsize_t right_matrices_per_block, bool /*strided_load*/,
                                       bool /*column_group_per_block*/, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::limit("Too many shifts per block", shifts_per_cuda_block, 32);
    detail::limit("Invalid number of right matrices per block", right_matrices_per_block, 8);
    detail::family_call(XC_ONE_TO_MANY, left, right, out, matrix_size, 1, num_right_matrices,
                        detail::direct(shifts_per_cuda_block, 1, right_matrices_per_block, 1, 1), stream);
}

// run_ccn_shuffle_n_to_m_multimat_both_work_distribution, src/kernels.cuh:162-175
// (src/naive_shuffle_n_to_m_multimat_both.cu:631-679).
template <typename DIST, typename T, typename RES>
void run_ccn_shuffle_n_to_m_multimat_both_work_distribution(const T* left, const T* right, RES* out,
                                                            dsize2_t matrix_size, dsize2_t search_size,
                                                            dsize_t num_left_matrices, dsize_t num_right_matrices,
                                                            dsize_t warps_per_thread_block,
                                                            dsize_t left_matrices_per_thread,
                                                            dsize_t right_matrices_per_thread,
                                                            dsize_t max_rows_per_thread,
                                                            xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Too many left matrices per thread", left_matrices_per_thread, 4);
    detail::limit("Too many right matrices per thread", right_matrices_per_thread, 4);
    auto o = detail::direct(warps_per_thread_block, 1, right_matrices_per_thread, left_matrices_per_thread, 1);
    detail::distribute<DIST>(o, max_rows_per_thread, matrix_size.y);
    detail::family_call(XC_N_TO_M, left, right, out, matrix_size, num_left_matrices, num_right_matrices, o, stream);
}

// run_ccn_shuffle_multirow_right, src/kernels.cuh:177-186 (src/naive_shuffle_multirow_right.cu:520-552):
// right_rows_per_thread consecutive shifts per thread.
template <typename T, typename RES>
void run_ccn_shuffle_multirow_right(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                    dsize2_t search_size, dsize_t warps_per_thread_block,
                                    dsize_t right_rows_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Too many right rows per thread", right_rows_per_thread, 8);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1,
                        detail::direct(warps_per_thread_block, right_rows_per_thread, 1, 1, 1), stream);
}

// run_ccn_shuffle_multirow_both, src/kernels.cuh:188-197 (src/naive_shuffle_multirow_both.cu:634-677):
// max_shifts_per_thread shifts and max_left_rows rows per step.
template <typename T, typename RES>
void run_ccn_shuffle_multirow_both(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                   dsize2_t search_size, dsize_t warps_per_thread_block,
                                   dsize_t max_shifts_per_thread, dsize_t max_left_rows,
                                   xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Too many shifts per thread", max_shifts_per_thread, 8);
    detail::limit("Too many left rows per iteration", max_left_rows, 4);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1,
                        detail::direct(warps_per_thread_block, max_shifts_per_thread, 1, 1, max_left_rows), stream);
}

// run_ccn_block_per_shift, src/kernels.cuh:199-207: one shift per thread, cuda_block_size / 64 waves.
template <typename T, typename RES>
void run_ccn_block_per_shift(const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size,
                             dsize_t cuda_block_size, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::limit("CUDA block too large", cuda_block_size, 1024);
    detail::family_call(XC_ONE_TO_ONE, left, right, out, matrix_size, 1, 1,
                        detail::direct((cuda_block_size + 31) / 32, 1, 1, 1, 1), stream);
}

// run_ccn_shuffle_multirow_right_multimat_right, src/kernels.cuh:209-221
// (src/naive_shuffle_multirow_right_multimat_right.cu:644-691).
template <typename T, typename RES>
void run_ccn_shuffle_multirow_right_multimat_right(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                                   dsize2_t search_size, dsize_t num_right_matrices,
                                                   dsize_t warps_per_thread_block, dsize_t right_rows_per_thread,
                                                   dsize_t right_matrices_per_thread, xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::check_warps(warps_per_thread_block);
    detail::limit("Invalid number of right rows per thread", right_rows_per_thread, 4);
    detail::limit("Invalid number of right matrices per thread", right_matrices_per_thread, 4);
    detail::family_call(XC_ONE_TO_MANY, left, right, out, matrix_size, 1, num_right_matrices,
                        detail::direct(warps_per_thread_block, right_rows_per_thread, right_matrices_per_thread, 1, 1),
                        stream);
}

// run_ccn_shuffle_one_to_many_multirow_both_multimat_right, src/kernels.cuh:223-236
// (src/naive_shuffle_one_to_many_multirow_both_multimat_right.cu:776-833).
template <typename T, typename RES>
void run_ccn_shuffle_one_to_many_multirow_both_multimat_right(
    const T* left, const T* right, RES* out, dsize2_t matrix_size, dsize2_t search_size, dsize_t num_right_matrices,
    dsize_t warps_per_thread_block, dsize_t shifts_per_thread_right_matrix, dsize_t right_matrices_per_thread,
    dsize_t left_rows_per_iteration, xc_stream_t stream = nullptr) {
    if constexpr (std::is_same_v<T, float>) {
        static_assert(std::is_same_v<RES, float>, "result type must match the input type");
        detail::check(xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
            left, right, out, matrix_size.x, matrix_size.y, search_size.x, search_size.y, num_right_matrices,
            warps_per_thread_block, shifts_per_thread_right_matrix, right_matrices_per_thread,
            left_rows_per_iteration, stream));
    } else {
        detail::check_search(matrix_size, search_size);
        detail::limit("Cuda block too large", warps_per_thread_block, 32);
        detail::limit("Too many shifts per thread in each right matrix", shifts_per_thread_right_matrix, 4);
        detail::limit("Invalid number of right matrices per thread", right_matrices_per_thread, 4);
        detail::limit("Too many left rows per iteration", left_rows_per_iteration, 4);
        detail::family_call(XC_ONE_TO_MANY, left, right, out, matrix_size, 1, num_right_matrices,
                            detail::direct(warps_per_thread_block, shifts_per_thread_right_matrix,
                                           right_matrices_per_thread, 1, left_rows_per_iteration),
                            stream);
    }
}

// run_ccn_n_to_m_shuffle_multirow_both_multimat_both, src/kernels.cuh:238-254
// (src/naive_shuffle_n_to_m_multirow_both_multimat_both.cu:953-1021).
template <typename T, typename RES>
void run_ccn_n_to_m_shuffle_multirow_both_multimat_both(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                                        dsize2_t search_size, dsize_t num_left_matrices,
                                                        dsize_t num_right_matrices, dsize_t warps_per_thread_block,
                                                        dsize_t shifts_per_thread_right_matrix,
                                                        dsize_t left_matrices_per_thread,
                                                        dsize_t right_matrices_per_thread,
                                                        dsize_t left_rows_per_iteration,
                                                        xc_stream_t stream = nullptr) {
    detail::check_search(matrix_size, search_size);
    detail::limit("Cuda block too large", warps_per_thread_block, 32);
    detail::limit("Too many shifts per thread in each right matrix", shifts_per_thread_right_matrix, 4);
    detail::limit("Too many left matrices per thread", left_matrices_per_thread, 4);
    detail::limit("Too many right matrices per thread", right_matrices_per_thread, 4);
    detail::limit("Too many left rows per iteration", left_rows_per_iteration, 4);
    detail::family_call(XC_N_TO_M, left, right, out, matrix_size, num_left_matrices, num_right_matrices,
                        detail::direct(warps_per_thread_block, shifts_per_thread_right_matrix,
                                       right_matrices_per_thread, left_matrices_per_thread, left_rows_per_iteration),
                        stream);
}

// The reference keeps copies of two kernels that differ only in how lane 0 selects a value
// (src/naive_shuffle_multirow_both_local_mem.cu, src/naive_shuffle_n_to_m_multimat_both_local_mem.cu,
// declared at src/kernels.cuh:256-281); here they are the same launchers.
namespace local_mem {

template <typename DIST, typename T, typename RES>
void run_ccn_shuffle_n_to_m_multimat_both_work_distribution(const T* left, const T* right, RES* out,
                                                            dsize2_t matrix_size, dsize2_t search_size,
                                                            dsize_t num_left_matrices, dsize_t num_right_matrices,
                                                            dsize_t warps_per_thread_block,
                                                            dsize_t left_matrices_per_thread,
                                                            dsize_t right_matrices_per_thread,
                                                            dsize_t max_rows_per_thread,
                                                            xc_stream_t stream = nullptr) {
    cross::run_ccn_shuffle_n_to_m_multimat_both_work_distribution<DIST>(
        left, right, out, matrix_size, search_size, num_left_matrices, num_right_matrices, warps_per_thread_block,
        left_matrices_per_thread, right_matrices_per_thread, max_rows_per_thread, stream);
}

template <typename T, typename RES>
void run_ccn_shuffle_multirow_both(const T* left, const T* right, RES* out, dsize2_t matrix_size,
                                   dsize2_t search_size, dsize_t warps_per_thread_block,
                                   dsize_t max_shifts_per_thread, dsize_t max_left_rows,
                                   xc_stream_t stream = nullptr) {
    cross::run_ccn_shuffle_multirow_both(left, right, out, matrix_size, search_size, warps_per_thread_block,
                                         max_shifts_per_thread, max_left_rows, stream);
}

}  // namespace local_mem

}  // namespace cross
