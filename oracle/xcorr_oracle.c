/*
 * xcorr_oracle.c -- CPU restatement of the reference's definition of batched
 * 2D full cross-correlation.  TEST INFRASTRUCTURE ONLY: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * this file's shared object (as the checker, or as the timed CPU baseline).
 * The product library never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference checkout).  How each function is pinned
 * (golden vectors, closed forms, invariants, brute force) is listed in the
 * table in tests/test_oracle.py; no function here is "parity unpinned".
 *
 * Layout: matrices are row-major, rows x cols, stacked back to back
 * (src/matrix.hpp:410-619).  The result of one pair is (2*rows-1) x
 * (2*cols-1); element [(dy+rows-1)*(2*cols-1) + (dx+cols-1)] holds
 *     sum_{r,c} ref[r][c] * target[r-dy][c-dx]
 * over the overlap (src/validate.hpp:94-132).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define XCO_EXPORT __attribute__((visibility("default")))

/* hadamard_and_sum + naive_cpu_cross_corr, src/validate.hpp:94-132.
 * Same loop nest: shifts y outer / x inner, overlap rows outer / columns
 * inner, one accumulator of type T, target*ref multiply then add. */
#define XCO_PAIR_BODY(T)                                                       \
    const int R = (int)rows, C = (int)cols;                                    \
    const int SX = 2 * C - 1;                                                  \
    for (int y = -(R - 1); y <= R - 1; ++y) {                                  \
        int y0 = y > 0 ? y : 0, y1 = (R + y) < R ? (R + y) : R;                \
        for (int x = -(C - 1); x <= C - 1; ++x) {                              \
            int x0 = x > 0 ? x : 0, x1 = (C + x) < C ? (C + x) : C;            \
            T sum = (T)0;                                                      \
            for (int yr = y0; yr < y1; ++yr) {                                 \
                const T* rr = ref + (size_t)yr * C;                            \
                const T* tr = target + (size_t)(yr - y) * C;                   \
                for (int xr = x0; xr < x1; ++xr) {                             \
                    T prod = tr[xr - x] * rr[xr];                              \
                    sum += prod;                                               \
                }                                                              \
            }                                                                  \
            res[(size_t)(y + R - 1) * SX + (x + C - 1)] = sum;                 \
        }                                                                      \
    }

XCO_EXPORT void xco_pair_f64(const double* ref, const double* target, double* res,
                             uint32_t rows, uint32_t cols) {
    XCO_PAIR_BODY(double)
}

XCO_EXPORT void xco_pair_f32(const float* ref, const float* target, float* res,
                             uint32_t rows, uint32_t cols) {
    XCO_PAIR_BODY(float)
}

/* corr(|ref|, |target|) in double: the per-output denominator sum |a||b| of
 * the normalised error metric.  Same loop nest as xco_pair_f64. */
XCO_EXPORT void xco_pair_abs_f64(const double* ref, const double* target, double* res,
                                 uint32_t rows, uint32_t cols) {
    const int R = (int)rows, C = (int)cols;
    const int SX = 2 * C - 1;
    for (int y = -(R - 1); y <= R - 1; ++y) {
        int y0 = y > 0 ? y : 0, y1 = (R + y) < R ? (R + y) : R;
        for (int x = -(C - 1); x <= C - 1; ++x) {
            int x0 = x > 0 ? x : 0, x1 = (C + x) < C ? (C + x) : C;
            double sum = 0.0;
            for (int yr = y0; yr < y1; ++yr)
                for (int xr = x0; xr < x1; ++xr)
                    sum += fabs(target[(size_t)(yr - y) * C + xr - x]) * fabs(ref[(size_t)yr * C + xr]);
            res[(size_t)(y + R - 1) * SX + (x + C - 1)] = sum;
        }
    }
}

/* Family orderings.
 *   family 0 one_to_one   src/validate.hpp:216-229   k = 0            : (0, 0)
 *   family 1 one_to_many  src/validate.hpp:231-247   k = t            : (0, t)
 *   family 2 n_to_mn      src/validate.hpp:249-270   k = t*n + r      : (r, t*n + r)
 *   family 3 n_to_m       src/validate.hpp:279-291   k = r*m + t      : (r, t)
 * n and m are the numbers of ref (left) and target (right) matrices.
 * Returns the number of result matrices, or -1 for counts the reference
 * rejects (check_num_result_matrices / the divisibility check).  The
 * n_to_mn result count is m (the reference's allocating overload uses n,
 * src/validate.hpp:274, which then fails its own check at :252). */
XCO_EXPORT int64_t xco_num_results(int family, uint32_t n, uint32_t m) {
    switch (family) {
    case 0: return (n == 1 && m == 1) ? 1 : -1;
    case 1: return (n == 1 && m >= 1) ? (int64_t)m : -1;
    case 2: return (n >= 1 && m >= 1 && m % n == 0) ? (int64_t)m : -1;
    case 3: return (n >= 1 && m >= 1) ? (int64_t)n * m : -1;
    default: return -1;
    }
}

/* k-th result's (left, right) indices for a family; see table above. */
XCO_EXPORT int xco_pair_of_result(int family, uint32_t n, uint32_t m, uint64_t k,
                                  uint32_t* left, uint32_t* right) {
    if (xco_num_results(family, n, m) < 0 || k >= (uint64_t)xco_num_results(family, n, m)) return -1;
    switch (family) {
    case 0: *left = 0; *right = 0; break;
    case 1: *left = 0; *right = (uint32_t)k; break;
    case 2: *left = (uint32_t)(k % n); *right = (uint32_t)k; break;
    case 3: *left = (uint32_t)(k / m); *right = (uint32_t)(k % m); break;
    }
    return 0;
}

typedef struct {
    int family, dbl, absval;
    const void* ref;
    const void* target;
    void* res;
    uint32_t rows, cols, n, m;
    uint64_t k0, k1;
} xco_job;

static void* xco_run_job(void* p) {
    xco_job* j = (xco_job*)p;
    size_t in_sz = (size_t)j->rows * j->cols;
    size_t out_sz = (size_t)(2 * j->rows - 1) * (2 * j->cols - 1);
    for (uint64_t k = j->k0; k < j->k1; ++k) {
        uint32_t l, r;
        xco_pair_of_result(j->family, j->n, j->m, k, &l, &r);
        if (j->dbl) {
            const double* a = (const double*)j->ref + l * in_sz;
            const double* b = (const double*)j->target + r * in_sz;
            double* o = (double*)j->res + k * out_sz;
            if (j->absval) xco_pair_abs_f64(a, b, o, j->rows, j->cols);
            else xco_pair_f64(a, b, o, j->rows, j->cols);
        } else {
            xco_pair_f32((const float*)j->ref + l * in_sz, (const float*)j->target + r * in_sz,
                         (float*)j->res + k * out_sz, j->rows, j->cols);
        }
    }
    return NULL;
}

/* cpu_cross_corr_{one_to_one,one_to_many,n_to_mn,n_to_m}, src/validate.hpp:216-298,
 * over results [k0, k1) split across `threads` POSIX threads (threads <= 1 is the
 * reference's single-threaded loop). */
static int xco_family(int family, int dbl, int absval, const void* ref, const void* target, void* res,
                      uint32_t rows, uint32_t cols, uint32_t n, uint32_t m,
                      uint64_t k0, uint64_t k1, int threads) {
    int64_t total = xco_num_results(family, n, m);
    if (total < 0 || rows == 0 || cols == 0) return -1;
    if (k1 > (uint64_t)total) k1 = (uint64_t)total;
    if (k0 >= k1) return 0;
    if (threads <= 1) {
        xco_job j = {family, dbl, absval, ref, target, res, rows, cols, n, m, k0, k1};
        xco_run_job(&j);
        return 0;
    }
    pthread_t tid[256];
    xco_job jobs[256];
    if (threads > 256) threads = 256;
    uint64_t cnt = k1 - k0;
    if ((uint64_t)threads > cnt) threads = (int)cnt;
    for (int t = 0; t < threads; ++t) {
        uint64_t a = k0 + cnt * t / threads, b = k0 + cnt * (t + 1) / threads;
        xco_job j = {family, dbl, absval, ref, target, res, rows, cols, n, m, a, b};
        jobs[t] = j;
        pthread_create(&tid[t], NULL, xco_run_job, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    return 0;
}

XCO_EXPORT int xco_family_f64(int family, const double* ref, const double* target, double* res,
                              uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int threads) {
    return xco_family(family, 1, 0, ref, target, res, rows, cols, n, m, 0, UINT64_MAX, threads);
}

XCO_EXPORT int xco_family_abs_f64(int family, const double* ref, const double* target, double* res,
                                  uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int threads) {
    return xco_family(family, 1, 1, ref, target, res, rows, cols, n, m, 0, UINT64_MAX, threads);
}

/* Results [k0, k1) only: `res` points at result k0.  Used for bounded CPU
 * baseline samples of large workloads. */
XCO_EXPORT int xco_family_range_f32(int family, const float* ref, const float* target, float* res,
                                    uint32_t rows, uint32_t cols, uint32_t n, uint32_t m,
                                    uint64_t k0, uint64_t k1, int threads) {
    size_t out_sz = (size_t)(2 * rows - 1) * (2 * cols - 1);
    return xco_family(family, 0, 0, ref, target, (float*)res - k0 * out_sz, rows, cols, n, m, k0, k1, threads);
}

XCO_EXPORT int xco_family_f32(int family, const float* ref, const float* target, float* res,
                              uint32_t rows, uint32_t cols, uint32_t n, uint32_t m, int threads) {
    return xco_family(family, 0, 0, ref, target, res, rows, cols, n, m, 0, UINT64_MAX, threads);
}

/* validate_result, src/validate.hpp:134-188: per element
 *   d = 0                              if |a| <= 1e-5 and |b| <= 1e-5
 *   d = |a - b| / max(|a|, |b|)        otherwise
 * (a = result, b = valid), then max, mean and the population standard
 * deviation (Boost lazy variance = E[d^2] - E[d]^2), plus the (result, valid)
 * values at the first element whose difference equals the max.
 * out[0..4] = max, mean, stddev, result value at max, valid value at max. */
XCO_EXPORT void xco_validate_result(const double* result, const double* valid, uint64_t count,
                                    double* out) {
    const double eps = 1e-5;
    double mx = -INFINITY, s = 0.0, s2 = 0.0;
    uint64_t imax = 0;
    for (uint64_t i = 0; i < count; ++i) {
        double a = result[i], b = valid[i], d;
        if (fabs(a) <= eps && fabs(b) <= eps) d = 0.0;
        else d = fabs(a - b) / (fabs(a) > fabs(b) ? fabs(a) : fabs(b));
        if (d > mx) { mx = d; imax = i; }
        s += d;
        s2 += d * d;
    }
    double mean = count ? s / (double)count : NAN;
    double var = count ? s2 / (double)count - mean * mean : NAN;
    out[0] = count ? mx : NAN;
    out[1] = mean;
    out[2] = sqrt(var > 0 ? var : 0.0);
    out[3] = count ? result[imax] : NAN;
    out[4] = count ? valid[imax] : NAN;
}

/* Normalised error: e_k = |x_k - ref_k| / den_k with den = corr(|L|,|R|);
 * outputs with den == 0 count as 0 when x == ref.  out = max, mean, argmax. */
XCO_EXPORT void xco_norm_error(const double* x, const double* ref, const double* den, uint64_t count,
                               double* out) {
    double mx = 0.0, s = 0.0;
    uint64_t im = 0;
    for (uint64_t i = 0; i < count; ++i) {
        double d = fabs(x[i] - ref[i]);
        double e = den[i] > 0 ? d / den[i] : (d == 0 ? 0.0 : INFINITY);
        if (e > mx) { mx = e; im = i; }
        s += e;
    }
    out[0] = mx;
    out[1] = count ? s / (double)count : 0.0;
    out[2] = (double)im;
}

/* normalize_fft_results, src/fft_helpers.hpp:193-261: res is the
 * unnormalised circular correlation of size rows x cols (the zero-padded
 * 2H x 2W transform size); norm is (rows-1) x (cols-1).  Quadrants around
 * the empty row rows/2 and column cols/2 are swapped and divided by rows*cols. */
static void xco_copy_quadrant(const double* src, uint32_t scols, double area, double* tgt, uint32_t tcols,
                              uint32_t sx, uint32_t sy, uint32_t w, uint32_t h, uint32_t tx, uint32_t ty) {
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            tgt[(size_t)(ty + y) * tcols + tx + x] = src[(size_t)(sy + y) * scols + sx + x] / area;
}

XCO_EXPORT void xco_normalize_fft_f64(const double* res, uint32_t rows, uint32_t cols, double* norm) {
    uint32_t ex = cols / 2, ey = rows / 2;
    uint32_t nc = cols - 1, nr = rows - 1;
    double area = (double)rows * (double)cols;
    xco_copy_quadrant(res, cols, area, norm, nc, 0, 0, ex, ey, nc / 2, nr / 2);
    xco_copy_quadrant(res, cols, area, norm, nc, ex + 1, 0, cols - ex - 1, ey, 0, nr / 2);
    xco_copy_quadrant(res, cols, area, norm, nc, 0, ey + 1, ex, rows - ey - 1, nc / 2, 0);
    xco_copy_quadrant(res, cols, area, norm, nc, ex + 1, ey + 1, cols - ex - 1, rows - ey - 1, 0, 0);
}

/* get_interleaved_index, src/matrix.hpp:614-618. */
XCO_EXPORT uint32_t xco_interleaved_index(uint32_t idx, uint32_t group_size, uint32_t num_groups) {
    return (idx % group_size) * num_groups + idx / group_size;
}

/* get_next_iteration_count, src/stopwatch.hpp:57-71 (times in ns). */
XCO_EXPORT uint64_t xco_next_iteration_count(double min_time_ns, double total_ns, uint64_t iterations) {
    if (total_ns >= min_time_ns) return 0;
    double ratio = min_time_ns / total_ns;
    double f = ratio * 1.5;
    if (f > 100.0) f = 100.0;
    if (f < 1.5) f = 1.5;
    return (uint64_t)ceil((double)iterations * f);
}
