// Method-level entry points behind the C ABI (one per algorithm family of the library).
#pragma once

#include <vector>

#include "xc_internal.hpp"

namespace xc {

// Grow-only device workspace, one per purpose; released at process exit.
void* workspace(int slot, size_t bytes, hipStream_t s);

// On-device error checks in FP64 (xc_check.hip); they synchronise `s` to return the values.
void validate_result_f32(const float* r, const float* v, uint64_t n, double* out5, hipStream_t s);
void validate_result_f64(const double* r, const double* v, uint64_t n, double* out5, hipStream_t s);
void norm_error_f32(const float* x, const double* ref, const double* den, uint64_t n, double* out3, hipStream_t s);
void norm_error_f64(const double* x, const double* ref, const double* den, uint64_t n, double* out3, hipStream_t s);
enum : int {
    WS_FFT_PAD = 0, WS_FFT_SPEC = 1, WS_FFT_OUT = 2, WS_FFT_RES = 3, WS_SPLIT = 4, WS_BATCH = 5,
    WS_FK_SR = 6, WS_FK_SL = 7, WS_FK_LHAT = 8, WS_FK_Z = 9, WS_FRAME = 10, WS_CHECK = 11, WS_DIRECT = 12
};

// Dispatch of xc_cross_corr_f32 on opt.precision (xc_precision).
void run_family_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                    hipStream_t s);

// Split-precision MFMA implicit GEMM (xc_mfma.hip); `kind` is an xc_precision value.
void launch_mfma_f32(const family_args& a, const float* L, const float* R, float* O, int kind,
                     uint32_t border, uint32_t tile, hipStream_t s);
int mfma_catalog_size();          // MFMA tiles among the cost model's candidates
uint32_t mfma_candidate_tile(int i);  // tile code of MFMA candidate i

// Uneven batch: work list of the persistent kernel (xc_mfma.hip) and its host plan (xc_batch.cpp).
struct batch_pair_desc {
    uint64_t loff, roff, ooff;  // element offsets of the pair's matrices
    uint32_t rows, cols;
    uint32_t table_off, nh;     // frame table of the pair's size in the tables buffer
};
struct batch_item {
    uint32_t pair;
    uint16_t kind;  // 0: MFMA block tile (tx, ty = b); 1: 256 frame outputs starting at b
    uint16_t tx;
    uint32_t b;
};
uint32_t batch_tiles_x(uint32_t cols);
uint32_t batch_tiles_y(uint32_t rows);
uint32_t batch_frame_chunk();
uint32_t batch_default_border(int kind);
uint32_t frame_table_host(uint32_t rows, uint32_t cols, uint32_t thresh, std::vector<uint32_t>& start);
const void* batch_kernel_ptr(int kind);
uint32_t batch_lds_bytes(int kind);
uint32_t batch_blocks_per_cu();

void run_batch_f32(const xc_pair* pairs, uint32_t num_pairs, const float* L, const float* R, float* O,
                   const xc_options& opt, hipStream_t s);
xc_batch_plan* batch_plan_create(const xc_pair* pairs, uint32_t num_pairs, const xc_options& opt);
std::vector<xc_options> batch_candidates(const xc_pair* pairs, uint32_t num_pairs);
uint32_t batch_autotune_f32(const xc_pair* pairs, uint32_t num_pairs, const float* L, const float* R, float* O,
                            hipStream_t s, std::vector<xc_options>& cands, std::vector<float>& ms);
void batch_plan_destroy(xc_batch_plan* plan);  // runs the plan's destructor
void batch_plan_exec(xc_batch_plan* plan, const float* L, const float* R, float* O, hipStream_t s, bool use_graph);

// FFT method (xc_fft.hip).
void run_fft_f32(const family_args& a, const float* L, const float* R, float* O, bool normalize,
                 bool fp64_transform, hipStream_t s);
void run_fft_f64(const family_args& a, const double* L, const double* R, double* O, bool normalize,
                 hipStream_t s);
// One launch sequence of the own-FFT method: nl left and nr right matrices of H x W and
// `pairs` results.  The device tables are null for a regular family call; an uneven batch
// passes, per size class, the element offsets of its matrices (nl left ones, then nr right
// ones), a (left index, right index) word pair per result and each result's offset.
struct fftk_job {
    uint32_t H = 0, W = 0, nl = 0, nr = 0, pairs = 0;
    int family = 0;
    uint32_t n = 0, m = 0;
    const uint64_t* moff = nullptr;
    const uint32_t* ptab = nullptr;
    const uint64_t* ooff = nullptr;
    // Uneven batches: (rows, cols) of every matrix (lefts, then rights), each at most H x W
    // and with the padded lengths of H x W; null when every matrix is H x W.
    const uint32_t* mdims = nullptr;
    // Largest number of pairs per launch of the column and inverse-row kernels (0: as
    // many as fit in 2 GiB of Z).
    uint32_t max_chunk = 0;
    // Row transforms of the power of two >= 2W-1 even where 3 * 2^n is shorter.
    bool pow2_rows = false;
};
struct fftk_ws_sizes {
    size_t sr, sl, lh, z;
    uint32_t chunk;  // pairs per launch of the cols / rows_inv kernels
};
struct fftk_ws {
    void *sr, *sl, *lh, *z;
    uint32_t zchunk;
};
fftk_ws_sizes fftk_sizes(const fftk_job& j);
// Transform lengths of the own-FFT method: columns the power of two >= 2 rows - 1 (at least
// 16); rows that, or 3 * 2^n (n = 4..9) when that is shorter.  Matrices of one job share them.
uint32_t fftk_col_len(uint32_t rows);
uint32_t fftk_row_len(uint32_t cols);
void fftk_prepare();  // device tables the kernels read (call before stream capture)
void fftk_launch_f32(const fftk_job& j, const float* L, const float* R, float* O, const fftk_ws& ws, hipStream_t s);

// FFT method on the library's own FP64 FFT kernels (xc_fftk.hip): rows, cols <= 2048.
// The own FFT without any allocation: scratch memory is a fixed array of the code object
// (rows, cols <= 2048; one pair's spectra must fit it).
void run_fftk_noalloc_f32(const family_args& a, const float* L, const float* R, float* O, hipStream_t s);
// max_chunk: pairs per Z chunk as fftk_job::max_chunk (0 = default); pow2_rows as
// fftk_job::pow2_rows.
void run_fftk_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t max_chunk, bool pow2_rows,
                  hipStream_t s);
void run_fftk_f64(const family_args& a, const double* L, const double* R, double* O, uint32_t max_chunk,
                  bool pow2_rows, hipStream_t s);

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
