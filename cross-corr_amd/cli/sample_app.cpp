// sample_app -- minimal embedding of the library (the reference's sample_app/src/main.cpp:14-76):
// load one left matrix and an array of right matrices from CSV, run the one_to_many
// multirow_both_multimat_right launcher with (8, 2, 2, 2), print the results as CSV on stdout.
#include <hip/hip_runtime.h>

#include <iostream>
#include <stdexcept>
#include <string>

#include "../../include/xcorr_kernels.hpp"
#include "matrix_io.hpp"

namespace {
void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "Usage: " << argv[0] << " <left_input_matrix_path> <right_input_matrices_path>" << std::endl;
        return 1;
    }
    try {
        auto left = xcli::load_csv<float>(argv[1]);
        auto right = xcli::load_csv<float>(argv[2]);
        if (left.rows != right.rows || left.cols != right.cols) {
            std::cerr << "Currently the implementation can only compute cross-correlation for left and right matrices "
                         "of the same size."
                      << std::endl;
            return 2;
        }
        xcli::matrix_array<float> results(2 * left.rows - 1, 2 * left.cols - 1, right.count);
        float *d_left, *d_right, *d_results;
        hip_ok(hipMalloc(&d_left, left.size() * sizeof(float)), "hipMalloc");
        hip_ok(hipMalloc(&d_right, right.size() * sizeof(float)), "hipMalloc");
        hip_ok(hipMalloc(&d_results, results.size() * sizeof(float)), "hipMalloc");
        hip_ok(hipMemcpy(d_left, left.data.data(), left.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hip_ok(hipMemcpy(d_right, right.data.data(), right.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");

        const cross::dsize_t warps_per_thread_block = 8;
        const cross::dsize_t shifts_per_thread_right_matrix = 2;
        const cross::dsize_t right_matrices_per_thread = 2;
        const cross::dsize_t left_rows_per_iteration = 2;
        cross::run_ccn_shuffle_one_to_many_multirow_both_multimat_right(
            d_left, d_right, d_results, cross::dsize2_t{left.cols, left.rows},
            cross::dsize2_t{results.cols, results.rows}, right.count, warps_per_thread_block,
            shifts_per_thread_right_matrix, right_matrices_per_thread, left_rows_per_iteration);

        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_ok(hipGetLastError(), "hipGetLastError");
        hip_ok(hipMemcpy(results.data.data(), d_results, results.size() * sizeof(float), hipMemcpyDeviceToHost),
               "hipMemcpy");
        (void)hipFree(d_results);
        (void)hipFree(d_right);
        (void)hipFree(d_left);
        xcli::store_csv(std::cout, results);
    } catch (std::exception& e) {
        std::cerr << "Exception occured: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
