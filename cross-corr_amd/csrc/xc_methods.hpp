// Method-level entry points behind the C ABI (one per algorithm family of the library).
#pragma once

#include "xc_internal.hpp"

namespace xc {

// Grow-only device workspace, one per purpose; released at process exit.
void* workspace(int slot, size_t bytes, hipStream_t s);
enum : int { WS_FFT_PAD = 0, WS_FFT_SPEC = 1, WS_FFT_OUT = 2, WS_FFT_RES = 3, WS_SPLIT = 4, WS_BATCH = 5 };

// Dispatch of xc_cross_corr_f32 on opt.precision (xc_precision).
void run_family_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                    hipStream_t s);

// Split-precision MFMA implicit GEMM (xc_mfma.hip); `kind` is an xc_precision value.
void launch_mfma_f32(const family_args& a, const float* L, const float* R, float* O, int kind,
                     uint32_t border, uint32_t tile, hipStream_t s);
int mfma_catalog_size();

// Uneven batch (xc_batch.hip).
void run_batch_f32(const xc_pair* pairs, uint32_t num_pairs, const float* L, const float* R, float* O,
                   const xc_options& opt, hipStream_t s);

// FFT method (xc_fft.hip).
void run_fft_f32(const family_args& a, const float* L, const float* R, float* O, bool normalize,
                 bool fp64_transform, hipStream_t s);
void run_fft_f64(const family_args& a, const double* L, const double* R, double* O, bool normalize,
                 hipStream_t s);

void check_scatter(const void* src, const void* dst, uint32_t src_rows, uint32_t src_cols, uint32_t dst_rows,
                   uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col);
void launch_scatter_f32(const float* src, float* dst, uint32_t src_rows, uint32_t src_cols, uint32_t num,
                        uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col, hipStream_t s);
void launch_scatter_f64(const double* src, double* dst, uint32_t src_rows, uint32_t src_cols, uint32_t num,
                        uint32_t dst_rows, uint32_t dst_cols, uint32_t dst_row, uint32_t dst_col, hipStream_t s);
void launch_hadamard_f32(const float* ref, float* def, uint64_t num_complex, uint32_t batch, hipStream_t s);
void launch_hadamard_f64(const double* ref, double* def, uint64_t num_complex, uint32_t batch, hipStream_t s);
void launch_hadamard_n_to_m_f32(const float* l, const float* r, float* o, uint64_t num_complex, uint32_t n,
                                uint32_t m, hipStream_t s);
void launch_hadamard_n_to_m_f64(const double* l, const double* r, double* o, uint64_t num_complex, uint32_t n,
                                uint32_t m, hipStream_t s);
void launch_normalize_f32(const float* res, float* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                          hipStream_t s);
void launch_normalize_f64(const double* res, double* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                          hipStream_t s);

}  // namespace xc
