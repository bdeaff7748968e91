/*
 * xcorr.h -- C ABI of the MI355X (gfx950) batched 2D full cross-correlation library
 * (libxcorr.so, built from cross-corr_amd/csrc).
 *
 * Contract (src/validate.hpp:94-132 of the reference): for a left ("ref") matrix L
 * and a right ("target") matrix R, both rows x cols, row-major,
 *
 *   out[(dy+rows-1)*(2*cols-1) + (dx+cols-1)] = sum_{r,c} L[r][c] * R[r-dy][c-dx]
 *
 * over the overlap, dy in [-(rows-1), rows-1], dx in [-(cols-1), cols-1].
 * Matrices of one call are stacked back to back; results are (2*rows-1) x (2*cols-1).
 *
 * Conventions shared by every entry point:
 *  - all matrix pointers are device pointers (hipMalloc'ed, or any memory the
 *    GPU can read); the caller owns them; results are fully overwritten, the
 *    caller never pre-zeroes anything;
 *  - work is enqueued on `stream` (NULL = the null stream) and is asynchronous;
 *    the library's scratch memory (and its rocFFT plans) are kept per (device,
 *    stream), so calls on different streams, or from threads using their own
 *    streams, may run at the same time; a stream's scratch grows on demand and
 *    is kept until the process ends;
 *  - no entry point throws: each returns an xc_status, and xc_last_error() gives
 *    the message of the calling thread's last failure;
 *  - arguments are validated on the host before any device work is enqueued.
 */
#ifndef XCORR_H
#define XCORR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* xc_stream_t; /* the same type as hipStream_t */

typedef enum xc_status {
    XC_OK = 0,
    XC_ERR_INVALID_ARGUMENT = 1, /* bad size, count, tuning value or family */
    XC_ERR_HIP = 2,              /* a HIP runtime call failed */
    XC_ERR_FFT = 3,              /* a hipFFT/rocFFT call failed */
    XC_ERR_UNSUPPORTED = 4,      /* valid request the chosen method cannot serve */
    XC_ERR_NO_MEMORY = 5         /* workspace allocation failed */
} xc_status;

/* Families, with the ordering of result k (src/validate.hpp:216-298):
 *  XC_ONE_TO_ONE  n = m = 1            k = 0         : corr(L0, R0)
 *  XC_ONE_TO_MANY n = 1                k = t         : corr(L0, Rt)
 *  XC_N_TO_MN     m % n == 0           k = t         : corr(L[t % n], Rt)   (file order)
 *  XC_N_TO_M      any n, m             k = l*m + t   : corr(Ll, Rt)                      */
typedef enum xc_family {
    XC_ONE_TO_ONE = 0,
    XC_ONE_TO_MANY = 1,
    XC_N_TO_MN = 2,
    XC_N_TO_M = 3
} xc_family;

/* Arithmetic used for float inputs.  Every method accumulates in FP32 with a
 * two-level (per row-pair, then total) sum.
 *  XC_PREC_FP32     exact FP32 products on the VALU (wave64 direct kernel)
 *  XC_PREC_BF16X3   MFMA implicit GEMM, each operand split into bf16 hi + lo,
 *                   hi*hi + hi*lo + lo*hi; outputs whose overlap is smaller than
 *                   the border threshold are recomputed exactly in FP32
 *  XC_PREC_BF16     MFMA, single bf16 product (fast, ~1e-3 relative error)
 *  XC_PREC_FP16     MFMA, single fp16 product (inputs must fit fp16 range)
 *  XC_PREC_FP16X3   MFMA, each operand split into fp16 hi + lo (hi*hi + hi*lo + lo*hi in
 *                   FP32, border outputs exact in FP32 as for bf16x3, default threshold
 *                   16 where bf16x3's is 512); 22 significant bits
 *                   per operand, inputs must fit fp16 range (|x| < 65504; the lo part of
 *                   |x| < 2^-3 loses bits to fp16 subnormals)
 *  XC_PREC_FFT64    FFT method computed in FP64, results in FP32; tile 0: rocFFT D2Z/Z2D,
 *                   tile 1: this library's fused FP64 FFT kernels (rows, cols <= 2048)
 *  XC_PREC_FFT32    FFT method in FP32 (rocFFT R2C/C2R); large error at low-overlap outputs
 *  XC_PREC_AUTO     the library picks (cost model / autotuned table)           */
typedef enum xc_precision {
    XC_PREC_AUTO = 0,
    XC_PREC_FP32 = 1,
    XC_PREC_BF16X3 = 2,
    XC_PREC_BF16 = 3,
    XC_PREC_FP16 = 4,
    XC_PREC_FFT64 = 5,
    XC_PREC_FFT32 = 6,
    XC_PREC_FP16X3 = 7
} xc_precision;

/* Optional tuning; zero-initialise and set what you need (0 = library default). */
typedef struct xc_options {
    uint32_t precision;   /* xc_precision */
    uint32_t tile;        /* kernel-specific tile code from the tile catalog, 0 = default;
                             FP32 (the direct method, float or double data): 1 = register-tile
                             kernel, 2 = lanes-over-pairs kernel (a left and 64 of its rights
                             per wave; not for one_to_one) */
    uint32_t split;       /* FP32: split of the reduction (overlap rows) across blocks;
                             FFT64 tile 1: at most this many pairs per launch of the column
                             and inverse-row kernels (less scratch memory); 0 = auto */
    uint32_t border;      /* BF16X3/FP16X3: overlap-area threshold (1..512) for exact FP32
                             recompute; 0 = default (512 for bf16x3, 16 for fp16x3) */
    /* Register tile and schedule of the direct kernel (XC_PREC_FP32, and double data's
     * direct method); 0 = the library's default for the family and size.  Each value is
     * rounded down to a power of two and the tile to the largest compiled one not above the
     * request in any dimension (xc_last_launch names the tile used).  They correspond to the
     * reference's tuning arguments:                                                       */
    uint32_t shifts_per_thread;     /* consecutive dy per thread, 1..16 (shifts_per_thread,
                                       right_rows_per_thread, shifts_per_thread_right_matrix) */
    uint32_t right_mats_per_thread; /* right matrices per thread, 1..8 (right_matrices_per_thread) */
    uint32_t left_mats_per_thread;  /* left matrices per thread, 1..4, n_to_m (left_matrices_per_thread) */
    uint32_t rows_per_step;         /* right rows per LDS step, 1..4 (left_rows_per_iteration) */
    uint32_t waves_per_block;       /* 64-lane waves per block, 1..8 (warps_per_thread_block / 2) */
    uint32_t rows_per_block;        /* triangle work distribution: at most this many overlap rows
                                       per block, blocks add atomically (max_rows_per_thread);
                                       exclusive with split, which is the rectangle distribution */
    uint32_t reserved[2];
} xc_options;

/* ---- family entry points (the drop-in for the L5 wrappers' run_impl) ----------------
 * Replace cpu_cross_corr_* (src/validate.hpp:216-298) and the run_* launchers of
 * src/kernels.cuh:65-283 that serve each family.  n = number of left matrices,
 * m = number of right matrices, out holds xc_num_results(family, n, m) results.   */
xc_status xc_cross_corr_f32(xc_family family, const float* left, const float* right, float* out,
                            uint32_t rows, uint32_t cols, uint32_t n, uint32_t m,
                            const xc_options* opt, xc_stream_t stream);
xc_status xc_cross_corr_f64(xc_family family, const double* left, const double* right, double* out,
                            uint32_t rows, uint32_t cols, uint32_t n, uint32_t m,
                            const xc_options* opt, xc_stream_t stream);

/* Allocation-free variant of xc_cross_corr_f32 with opt->precision = XC_PREC_FFT64 and
 * opt->tile = 1 (this library's own FP64 FFT kernels), with the same results.  It has no
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc, hipFree or allocating library routine: its scratch memory is a fixed 256 MiB
 * array in the library's code object, one per device the library runs on, held until the
 * process ends.  rows, cols <= 2048 and the spectra of one pair must fit the array (up to
 * about 1450 x 1450; 96 MiB at 1024 x 1024); a call with more pairs than fit at once runs in
 * parts, and a size whose single pair does not fit fails with XC_ERR_UNSUPPORTED.  Calls on
 * one device share the array, so calls that overlap on one device (other streams or threads)
 * race.  The first call on a device looks up the array's address and copies the FFT twiddle
 * tables in (synchronous copies). */
xc_status xc_cross_corr_noalloc_f32(xc_family family, const float* left, const float* right, float* out,
                                    uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, xc_stream_t stream);

/* Number of results of a family, or -1 if the counts are invalid for it
 * (one_to_one_validate_input_size etc., src/one_to_one.hpp:20-22,
 * src/one_to_many.hpp:15-17, src/n_to_mn.hpp:23-25, src/n_to_m.hpp:19-21). */
int64_t xc_num_results(xc_family family, uint32_t n, uint32_t m);

/* ---- uneven batch: pairs of different sizes in one persistent launch ----------------
 * No counterpart in the reference (its n_to_mn wrapper launches one kernel per left
 * matrix on round-robin streams, src/n_to_mn.hpp:553-571).  Pair i correlates
 * left + left_off[i] with right + right_off[i] (both rows[i] x cols[i]) into
 * out + out_off[i]; offsets are in elements and the descriptor arrays are HOST arrays. */
typedef struct xc_pair {
    uint64_t left_off, right_off, out_off;
    uint32_t rows, cols;
} xc_pair;

/* opt->precision: XC_PREC_FP32 (one direct launch per pair), an MFMA method or
 * XC_PREC_FFT64 tile 1 (a plan as below), or XC_PREC_AUTO: the schedule that
 * xc_batch_autotune_f32 measured fastest for this pair list, measured on this call's buffers
 * the first time (XC_AUTO=model: the own-FFT plan when every size fits it, else bf16x3). */
xc_status xc_batch_f32(const xc_pair* pairs, uint32_t num_pairs, const float* left, const float* right,
                       float* out, const xc_options* opt, xc_stream_t stream);

/* Reusable batch plan, captured into a HIP graph on first use and replayed.
 *  - bf16x3 / bf16 / fp16: the work list (MFMA block tiles of every pair, largest first, plus
 *    border-frame chunks for bf16x3) is built and uploaded once; each execution is one launch
 *    of the persistent kernel (kernel arguments are updated in the instantiated graph when
 *    the matrix pointers change).
 *  - XC_PREC_FFT64 tile 1: the own FFT kernels per class of equal transform lengths (pairs
 *    of different sizes with the same padded lengths share one job; a matrix shared
 *    by several pairs is transformed once), classes dealt largest first to opt->split
 *    parallel graph branches (1..8, 0 = 8), each with its own workspaces; re-captured when
 *    the matrix pointers change.  rows, cols <= 2048.
 * xc_batch_f32 keeps a few such plans cached by (pair list, device, stream).  A plan must not
 * be executed concurrently with itself (its work counter and workspaces are shared); destroy
 * synchronises the plan's device. */
typedef struct xc_batch_plan xc_batch_plan;
xc_status xc_batch_plan_create_f32(const xc_pair* pairs, uint32_t num_pairs, const xc_options* opt,
                                   xc_batch_plan** plan);
xc_status xc_batch_plan_exec_f32(xc_batch_plan* plan, const float* left, const float* right, float* out,
                                 int use_graph, xc_stream_t stream);
xc_status xc_batch_plan_destroy(xc_batch_plan* plan);

/* Batch autotuner: builds and times each candidate schedule of this pair list on the
 * caller's buffers -- the own-FFT plan over 1, 2, 4 and 8 parallel branches (when rows and
 * cols <= 2048) and the persistent bf16x3 MFMA plan (left out when the FFT plan fits and the
 * definition's work exceeds 1e13 MACs) -- as graph replays after an untimed
 * capture, remembers the fastest for XC_PREC_AUTO batches of the same pair list and returns
 * it in *best.  candidates and ms (either may be null) receive the first `capacity`
 * candidates and their mean milliseconds per execution; *count (may be null) the number of
 * candidates.  `out` is overwritten; the stream is synchronised. */
xc_status xc_batch_autotune_f32(const xc_pair* pairs, uint32_t num_pairs, const float* left, const float* right,
                                float* out, xc_options* candidates, float* ms, uint32_t capacity, uint32_t* count,
                                xc_options* best, xc_stream_t stream);

/* ---- FFT-method building blocks (src/kernels.cuh:7-63) ------------------------------ */
/* run_scatter (src/kernels.cuh:53-63, src/scatter.cu:15-102): copy src matrices into
 * the top-left (at dst_row, dst_col) of dst matrices and write zeros everywhere else
 * in dst (the reference needs a prior cudaMemset of dst; this one does not). */
xc_status xc_scatter_f32(const float* src, float* dst, uint32_t src_rows, uint32_t src_cols,
                         uint32_t num_matrices, uint32_t dst_rows, uint32_t dst_cols,
                         uint32_t dst_row, uint32_t dst_col, xc_stream_t stream);
xc_status xc_scatter_f64(const double* src, double* dst, uint32_t src_rows, uint32_t src_cols,
                         uint32_t num_matrices, uint32_t dst_rows, uint32_t dst_cols,
                         uint32_t dst_row, uint32_t dst_col, xc_stream_t stream);
/* run_hadamard_original (src/kernels.cuh:7-15, src/hadamard.cu:33-58,176-195):
 * deformed[b][i] = ref[i] * conj(deformed[b][i]) for b < batch, i < num_complex.
 * Complex values are interleaved (re, im) pairs. */
xc_status xc_hadamard_f32(const float* ref, float* deformed, uint64_t num_complex, uint32_t batch,
                          xc_stream_t stream);
xc_status xc_hadamard_f64(const double* ref, double* deformed, uint64_t num_complex, uint32_t batch,
                          xc_stream_t stream);
/* run_hadamard_n_to_m_over_output (src/kernels.cuh:29-39, src/hadamard.cu:116-149):
 * out[l*m + r][i] = left[l][i] * conj(right[r][i]). */
xc_status xc_hadamard_n_to_m_f32(const float* left, const float* right, float* out, uint64_t num_complex,
                                 uint32_t n, uint32_t m, xc_stream_t stream);
xc_status xc_hadamard_n_to_m_f64(const double* left, const double* right, double* out, uint64_t num_complex,
                                 uint32_t n, uint32_t m, xc_stream_t stream);
/* normalize_fft_results (src/fft_helpers.hpp:193-261) on the GPU: res is `batch`
 * unnormalised circular correlations of rows x cols (the padded size); norm gets
 * batch results of (rows-1) x (cols-1). */
xc_status xc_normalize_fft_f32(const float* res, float* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                               xc_stream_t stream);
xc_status xc_normalize_fft_f64(const double* res, double* norm, uint32_t rows, uint32_t cols, uint32_t batch,
                               xc_stream_t stream);
/* Whole FFT method for a family (fft_reduced_transfer_* / fft_better_n_to_m,
 * src/one_to_many.hpp:679-745, src/n_to_m.hpp:700-736).  If `normalize` is 0 the
 * results are the unnormalised (2*rows) x (2*cols) circular correlations the
 * reference stores without -n; otherwise (2*rows-1) x (2*cols-1) as every other
 * method.  Plans are cached per (device, stream, shape, batch, precision). */
xc_status xc_fft_cross_corr_f32(xc_family family, const float* left, const float* right, float* out,
                                uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int normalize,
                                int fp64_transform, xc_stream_t stream);
xc_status xc_fft_cross_corr_f64(xc_family family, const double* left, const double* right, double* out,
                                uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int normalize,
                                xc_stream_t stream);

/* ---- on-device error checks (FP64 reductions) -----------------------------------------
 * xc_validate_result_*: the statistics of the reference's validate_result
 * (src/validate.hpp:134-188) over count elements of device arrays: per element d = 0 when
 * |result| and |valid| are both at most 1e-5, else |result - valid| / max(|result|, |valid|);
 * out[0..4] = max d (first index on ties), mean, population standard deviation, and the
 * result and valid values at the max.  xc_norm_error_*: the normalised error this library
 * reports for a result x against an FP64 reference ref, e = |x - ref| / den with den =
 * corr(|L|, |R|) (e = 0 where den = 0 and x = ref, infinite where den = 0 otherwise);
 * out[0..2] = max e, mean e, index of the max (no counterpart in the reference, whose
 * checks run on the host).  Both read their results back to the host: they synchronise
 * `stream`. */
xc_status xc_validate_result_f32(const float* result, const float* valid, uint64_t count, double* out,
                                 xc_stream_t stream);
xc_status xc_validate_result_f64(const double* result, const double* valid, uint64_t count, double* out,
                                 xc_stream_t stream);
xc_status xc_norm_error_f32(const float* x, const double* ref, const double* den, uint64_t count, double* out,
                            xc_stream_t stream);
xc_status xc_norm_error_f64(const double* x, const double* ref, const double* den, uint64_t count, double* out,
                            xc_stream_t stream);

/* ---- reference run_* launchers with their tuning arguments --------------------------
 * Same argument meaning as src/kernels.cuh:65-73 and :223-236; tuning arguments are
 * validated with the limits of the reference's release build (warps_per_thread_block <= 32,
 * src/naive_shuffle.cu:461-463; shifts, right matrices and left rows per iteration <= 4,
 * src/naive_shuffle_one_to_many_multirow_both_multimat_right.cu:789-818) and select the
 * direct kernel's tile (see xc_options); matrix_size/search_size are {cols, rows}.  The other
 * run_* launchers of kernels.cuh are header templates over xc_cross_corr_* in
 * xcorr_kernels.hpp. */
xc_status xc_run_ccn_shuffle_f32(const float* left, const float* right, float* out,
                                 uint32_t cols, uint32_t rows, uint32_t search_cols, uint32_t search_rows,
                                 uint32_t warps_per_thread_block, xc_stream_t stream);
xc_status xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
    const float* left, const float* right, float* out, uint32_t cols, uint32_t rows,
    uint32_t search_cols, uint32_t search_rows, uint32_t num_right_matrices,
    uint32_t warps_per_thread_block, uint32_t shifts_per_thread_right_matrix,
    uint32_t right_matrices_per_thread, uint32_t left_rows_per_iteration, xc_stream_t stream);

/* ---- method selection: exhaustive autotuner and learned cost model ------------------
 * No counterpart in the reference, whose users pick an algorithm by name
 * (src/main.cpp:368-409, src/run_args.cpp:87-150).  XC_PREC_AUTO resolves a call to one
 * CANDIDATE, a concrete (precision, tile) whose results meet the FP32 accuracy bar
 * (normalised error <= 2e-6): the FP32 direct kernel, each bf16x3 MFMA tile of the tile
 * catalog, and the FFT method in FP64.  In this order AUTO uses
 *   1. the process's autotune table (filled by xc_autotune_f32 and by step 3),
 *   2. the cost model, if one is loaded and the matrix size lies in its training range:
 *      a small MLP that predicts every candidate's time from the workload's features,
 *   3. otherwise an exhaustive autotune on the call's own buffers (every candidate is run
 *      and timed; this synchronises the stream once per candidate).
 * The model is read on first use from $XC_COST_MODEL, else from tuning/cost_model.txt next
 * to libxcorr.so.  $XC_AUTO=autotune skips step 2; $XC_AUTO=model never autotunes (it falls
 * back to bf16x3 with the default tile when the model cannot answer). */
typedef struct xc_workload {
    uint32_t family, rows, cols, n, m;
} xc_workload;

typedef enum xc_select_source {
    XC_SELECT_MODEL = 0,    /* the cost model's prediction */
    XC_SELECT_TUNED = 1,    /* the autotune table */
    XC_SELECT_NONE = 2      /* neither can answer */
} xc_select_source;

uint32_t xc_num_candidates(void);
/* Options of candidate i (and its name "precision/tile", e.g. "bf16x3/4", if name != NULL). */
xc_status xc_candidate(uint32_t i, xc_options* opt, char* name, size_t name_len);
/* The cost model's input features of a workload (count written to *count). */
xc_status xc_workload_features(const xc_workload* w, float* features, uint32_t capacity, uint32_t* count);
/* Exhaustive autotune: runs every candidate on the given buffers (one untimed warm-up, then
 * `reps` runs timed with HIP events on `stream`), writes the mean ms of each candidate
 * (INFINITY for one that failed or was not run) to ms[xc_num_candidates()], the fastest to
 * *best, and records the winner in the autotune table.  Above 1e13 MACs the definition-based
 * candidates are run only if no FFT candidate could run.  `out` holds the last run's results. */
xc_status xc_autotune_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                          uint32_t cols, uint32_t n, uint32_t m, uint32_t reps, float* ms, uint32_t* best,
                          xc_stream_t stream);
/* Time one candidate the way xc_autotune_f32 does (INFINITY if it cannot run the workload). */
xc_status xc_time_candidate_f32(xc_family family, const float* left, const float* right, float* out, uint32_t rows,
                                uint32_t cols, uint32_t n, uint32_t m, uint32_t index, uint32_t reps, float* ms,
                                xc_stream_t stream);
xc_status xc_autotune_clear(void);
/* Load a cost-model file (NULL: the default path).  XC_ERR_UNSUPPORTED if it is missing
 * or was trained for another candidate list. */
xc_status xc_cost_model_load(const char* path);
/* Predicted ms of every candidate (host computation; no GPU needed). */
xc_status xc_cost_model_predict(const xc_workload* w, float* ms);
/* What AUTO picks without autotuning: writes the options and returns the source in *source. */
xc_status xc_select(const xc_workload* w, xc_options* chosen, int* source);
/* Fit the cost model to a timing table written by the tuning collector (CSV: family, rows,
 * cols, n, m, then one column of ms per candidate) on the GPU and write the model file.
 * Rows i with holdout_mod > 0 and i % holdout_mod == 0 are left out of training.  Empty
 * or "nan" cells are unmeasured candidates; AUTO considers a candidate only up to twice the
 * largest work (multiply-adds) it has a measured time for. */
typedef struct xc_train_options {
    uint32_t epochs;
    float lr;
    float weight_decay;
    uint32_t seed;
    uint32_t hidden;    /* width of the two hidden layers (cost_model_fit only) */
    int use_graph;      /* replay each epoch from a HIP graph */
    uint32_t reserved[2];
} xc_train_options;
xc_status xc_cost_model_fit(const char* timing_csv, const char* model_path, const xc_train_options* opt,
                            uint32_t holdout_mod);

/* The cost model's network on the GPU: an MLP with `layers` fully connected layers,
 * dims[0] inputs ... dims[layers] outputs, tanh on hidden layers.  Parameters are one flat
 * FP32 array, layer by layer W (dims[l+1] x dims[l], row-major) then b (dims[l+1]).
 * Pointers other than the *_host ones are device pointers; x, y, target and mask are
 * batch-major; `work` holds xc_mlp_work_size floats. */
typedef struct xc_mlp_shape {
    uint32_t layers;
    uint32_t dims[8];
} xc_mlp_shape;
uint64_t xc_mlp_num_params(const xc_mlp_shape* shape);
uint64_t xc_mlp_work_size(const xc_mlp_shape* shape, uint32_t batch);
/* Glorot-uniform weights and zero biases from a seed (host array). */
xc_status xc_mlp_init(const xc_mlp_shape* shape, uint64_t seed, float* params_host);
xc_status xc_mlp_forward_f32(const xc_mlp_shape* shape, const float* params, const float* x, uint32_t batch,
                             float* y, float* work, xc_stream_t stream);
/* loss = sum mask*(y - target)^2 / sum mask, and d loss / d params into grad. */
xc_status xc_mlp_loss_grad_f32(const xc_mlp_shape* shape, const float* params, const float* x, const float* target,
                               const float* mask, uint32_t batch, float* loss, float* grad, float* work,
                               xc_stream_t stream);
/* One Adam step (decoupled weight decay), step numbered from 1. */
xc_status xc_mlp_adam_f32(uint64_t count, float* params, const float* grad, float* m, float* v, uint32_t step,
                          float lr, float beta1, float beta2, float eps, float weight_decay, xc_stream_t stream);
/* Full-batch training from host arrays; params_host is the start point and receives the
 * result; loss_host (may be NULL) receives the loss of every epoch. */
xc_status xc_mlp_train_f32(const xc_mlp_shape* shape, const float* x_host, const float* target_host,
                           const float* mask_host, uint32_t batch, const xc_train_options* opt, float* params_host,
                           float* loss_host);

/* ---- diagnostics (src/kernel_args.hpp:9-33) ----------------------------------------- */
typedef struct xc_launch_desc {
    char kernel[96];
    uint32_t grid[3];
    uint32_t block[3];
    uint32_t lds_bytes;
    uint32_t rows, cols, pairs, split;
} xc_launch_desc;

/* Kernel timing for measurement: while enabled, every kernel (or rocFFT execution) the
 * library enqueues is bracketed by two HIP events recorded on its own stream.
 * xc_kernel_times synchronises on them and writes one line per kernel name,
 * "name launches total_ms", then forgets them; *needed gets the full text size.
 * Not for launches under stream capture. */
xc_status xc_kernel_timing(int enable);
xc_status xc_kernel_times(char* buf, size_t len, size_t* needed);

const char* xc_last_error(void);             /* thread-local; "" if none */
const xc_launch_desc* xc_last_launch(void);  /* thread-local; last launch this thread enqueued */
const char* xc_version(void);
/* Describe the device the library runs on (name, CUs, clock); returns XC_ERR_HIP without a GPU. */
xc_status xc_device_info(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* XCORR_H */
