// `cross` -- command-line drop-in for the reference's cross-correlation program
// (src/main.cpp:465-683, src/run_args.cpp:17-150), running every GPU algorithm through the
// C ABI of libxcorr.so (include/xcorr.h).
//
//   cross [global options] list
//   cross [global options] run [run options] <alg> <ref_path> <target_path>
//   cross [global options] validate [validate options] <template_data_path> <validate_data_path>...
//   cross [global options] input <alg_type> <rows> <columns> <left_matrices> <right_matrices>
//
// Exit codes as the reference: 0 success, 1 command-line / algorithm-argument errors,
// 2 other exceptions, 3 failure of a computation step.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <optional>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xcorr.h"
#include "json_args.hpp"
#include "matrix_io.hpp"

using namespace xcli;

namespace {

struct usage_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct argument_error : std::runtime_error {  // src/argument_error.hpp:6-15
    using std::runtime_error::runtime_error;
};

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void xc_ok(xc_status s) {
    if (s != XC_OK) throw std::runtime_error(std::string("xcorr: ") + xc_last_error());
}
void fft_ok(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) throw std::runtime_error(std::string(what) + " failed with hipfftResult " + std::to_string(r));
}

std::string last_launch_string() {
    const xc_launch_desc* d = xc_last_launch();
    if (!d || !d->kernel[0]) return "(none)";
    return std::string(d->kernel) + " grid=[" + std::to_string(d->grid[0]) + "," + std::to_string(d->grid[1]) + "," +
           std::to_string(d->grid[2]) + "] block=[" + std::to_string(d->block[0]) + "," + std::to_string(d->block[1]) +
           "," + std::to_string(d->block[2]) + "] lds=" + std::to_string(d->lds_bytes) + " size=" +
           std::to_string(d->rows) + "x" + std::to_string(d->cols) + " pairs=" + std::to_string(d->pairs) +
           " split=" + std::to_string(d->split);
}

// ---------------------------------------------------------------- measurement
// measurement_result + get_next_iteration_count, src/stopwatch.hpp:13-71.
struct measurement {
    double total_ns = 0;
    uint64_t iterations = 0;
    double iteration_ns() const { return iterations ? total_ns / (double)iterations : 0.0; }
};

uint64_t next_iteration_count(double min_ns, const measurement& m) {
    if (m.total_ns >= min_ns) return 0;
    const double ratio = min_ns / m.total_ns;
    return (uint64_t)std::ceil((double)m.iterations * std::max(std::min(ratio * 1.5, 100.0), 1.5));
}

double now_ns() {
    return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::high_resolution_clock::now().time_since_epoch())
        .count();
}

// Repeats `body` until the measured time reaches min_ns (CUDA_ADAPTIVE_MEASURE /
// CPU_ADAPTIVE_MEASURE, src/stopwatch.hpp:125-200).  gpu=true times with HIP events.
measurement adaptive(double min_ns, bool gpu, const std::function<void()>& body) {
    uint64_t iters = 1;
    measurement m;
    hipEvent_t a = nullptr, b = nullptr;
    if (gpu) {
        hip_ok(hipEventCreate(&a), "hipEventCreate");
        hip_ok(hipEventCreate(&b), "hipEventCreate");
    }
    while (iters) {
        if (gpu) {
            hip_ok(hipEventRecord(a, nullptr), "hipEventRecord");
            for (uint64_t i = 0; i < iters; ++i) body();
            hip_ok(hipEventRecord(b, nullptr), "hipEventRecord");
            hip_ok(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0;
            hip_ok(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            m = {ms * 1e6, iters};
        } else {
            const double t0 = now_ns();
            for (uint64_t i = 0; i < iters; ++i) body();
            m = {now_ns() - t0, iters};
        }
        iters = next_iteration_count(min_ns, m);
    }
    if (gpu) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    return m;
}

measurement once(const std::function<void()>& body) {
    const double t0 = now_ns();
    body();
    return {now_ns() - t0, 1};
}

// ---------------------------------------------------------------- validation
// validate_result, src/validate.hpp:134-188.
struct validation_stats {
    double max = 0, mean = 0, stddev = 0, max_result = 0, max_valid = 0;
};

template <typename A, typename B>
validation_stats validate_result(const A& result, const B& valid) {
    if (result.rows != valid.rows || result.cols != valid.cols || result.count != valid.count)
        throw std::runtime_error("Cannot validate matrices of different sizes");
    const size_t n = result.size();
    validation_stats st;
    double s = 0, s2 = 0, mx = -INFINITY;
    size_t im = 0;
    for (size_t i = 0; i < n; ++i) {
        const double a = result.data[i], b = valid.data[i];
        double d = 0;
        if (!(std::fabs(a) <= 1e-5 && std::fabs(b) <= 1e-5)) d = std::fabs(a - b) / std::max(std::fabs(a), std::fabs(b));
        if (d > mx) {
            mx = d;
            im = i;
        }
        s += d;
        s2 += d * d;
    }
    st.max = n ? mx : NAN;
    st.mean = n ? s / n : NAN;
    const double var = n ? s2 / n - st.mean * st.mean : NAN;
    st.stddev = std::sqrt(var > 0 ? var : 0);
    st.max_result = n ? (double)result.data[im] : NAN;
    st.max_valid = n ? (double)valid.data[im] : NAN;
    return st;
}

// simple_logger::result_stats, src/simple_logger.hpp:30-41 + operator<<, src/validate.hpp:76-92.
void print_stats(const validation_stats& st, bool progress, bool header) {
    if (progress) {
        std::cout << "Difference from valid values:\n";
        std::cout << "Max: " << st.max << " (" << std::fixed << st.max_result << " instead of " << st.max_valid
                  << ")\n";
        std::cout.unsetf(std::ios_base::fixed);
        std::cout << std::scientific << "Mean: " << st.mean << "\nStddev: " << st.stddev << "\n";
        std::cout.unsetf(std::ios_base::scientific);
    } else {
        if (header) std::cout << "Max,Mean,Stddev\n";
        std::cout << std::scientific << st.max << "," << st.mean << "," << st.stddev << "\n";
        std::cout.unsetf(std::ios_base::scientific);
    }
}

// k-th result -> (left, right) of a family (src/validate.hpp:216-298).
void pair_of(int family, uint32_t n, uint32_t m, uint64_t k, uint32_t& l, uint32_t& r) {
    switch (family) {
    case XC_ONE_TO_MANY: l = 0; r = (uint32_t)k; break;
    case XC_N_TO_MN: l = (uint32_t)(k % n); r = (uint32_t)k; break;
    case XC_N_TO_M: l = (uint32_t)(k / m); r = (uint32_t)(k % m); break;
    default: l = r = 0;
    }
}

// Host computation of a family: naive_cpu_cross_corr (src/validate.hpp:94-132) per result,
// accumulating in A.  The cpu_* algorithms run it as the reference does: A = T on one thread
// (src/validate.hpp:236-239), so their timings compare with the reference's.  `run -v`
// without a file validates against the same T-typed computation (src/cross_corr.hpp:90-97),
// spread over the host's threads (`threads` = 0) since that is not timed.
template <typename A, typename T>
void cpu_family(int family, const matrix_array<T>& left, const matrix_array<T>& right, matrix_array<T>& out,
                unsigned threads) {
    const int H = (int)left.rows, W = (int)left.cols;
    const int SX = 2 * W - 1;
    const uint64_t k = out.count;
    auto work = [&](uint64_t k0, uint64_t k1) {
        for (uint64_t kk = k0; kk < k1; ++kk) {
            uint32_t l, r;
            pair_of(family, left.count, right.count, kk, l, r);
            const T* ref = left.matrix(l);
            const T* tgt = right.matrix(r);
            T* res = out.matrix((uint32_t)kk);
            for (int y = -(H - 1); y <= H - 1; ++y)
                for (int x = -(W - 1); x <= W - 1; ++x) {
                    A sum = 0;
                    for (int yr = std::max(y, 0); yr < std::min(H + y, H); ++yr)
                        for (int xr = std::max(x, 0); xr < std::min(W + x, W); ++xr)
                            sum += (A)tgt[(size_t)(yr - y) * W + xr - x] * (A)ref[(size_t)yr * W + xr];
                    res[(size_t)(y + H - 1) * SX + x + W - 1] = (T)sum;
                }
        }
    };
    const unsigned avail = threads ? threads : std::thread::hardware_concurrency();
    const unsigned nt = std::max(1u, std::min<unsigned>(avail, (unsigned)k));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, k * t / nt, k * (t + 1) / nt);
    for (auto& t : th) t.join();
}

int64_t num_results(int family, uint32_t n, uint32_t m) { return xc_num_results((xc_family)family, n, m); }

// ---------------------------------------------------------------- algorithms
enum class method { cpu, direct, mfma, fft, autosel };
enum class arg_kind { integer, boolean, text };

struct arg_spec {
    const char* key;
    arg_kind kind;
    long def_int;
    const char* def_text;
};

struct alg_spec {
    std::string name;
    int family;
    method meth;
    int precision;      // xc_precision for direct/mfma
    bool fft_orig;      // host-side zero padding (fft_orig_*) instead of the device scatter
    std::vector<arg_spec> args;
};

const char* kFamilyNames[] = {"one_to_one", "one_to_many", "n_to_mn", "n_to_m"};

std::vector<alg_spec> make_algorithms() {
    using A = arg_spec;
    const A wptb{"warps_per_thread_block", arg_kind::integer, 8, nullptr};
    const A rpt{"rows_per_thread", arg_kind::integer, 10, nullptr};
    const A dist_r{"distribution_type", arg_kind::text, 0, "rectangle"};
    const A dist_t{"distribution_type", arg_kind::text, 0, "triangle"};
    const A streams{"num_cuda_streams", arg_kind::integer, 8, nullptr};
    std::vector<alg_spec> v;
    auto add = [&](std::string name, int fam, method m, int prec, std::vector<arg_spec> args, bool orig = false) {
        v.push_back({std::move(name), fam, m, prec, orig, std::move(args)});
    };
    const int D = XC_PREC_FP32;
    // src/main.cpp:368-409, same names and argument keys/defaults
    for (int f = 0; f < 4; ++f) add(std::string("cpu_") + kFamilyNames[f], f, method::cpu, D, {});
    for (int f = 0; f < 3; ++f) add(std::string("nai_orig_") + kFamilyNames[f], f, method::direct, D, {});
    add("nai_orig_n_to_m", 3, method::direct, D, {streams});
    add("nai_shuffle_one_to_one", 0, method::direct, D, {wptb});
    add("nai_shuffle_work_distribution_one_to_one", 0, method::direct, D, {wptb, rpt, dist_r});
    add("nai_shuffle_multimat_right_one_to_one", 0, method::direct, D, {wptb});
    add("nai_shuffle_multimat_right_one_to_many", 1, method::direct, D,
        {wptb, {"right_matrices_per_thread", arg_kind::integer, 2, nullptr}});
    add("nai_shuffle_multimat_right_work_distribution_one_to_one", 0, method::direct, D, {wptb, rpt, dist_r});
    add("nai_shuffle_multimat_right_work_distribution_one_to_many", 1, method::direct, D,
        {wptb, {"right_matrices_per_thread", arg_kind::integer, 2, nullptr}, rpt, dist_r});
    add("nai_shuffle_multimat_right_work_distribution_n_to_mn", 2, method::direct, D,
        {wptb, {"right_matrices_per_thread", arg_kind::integer, 2, nullptr}, rpt, dist_r, streams});
    add("nai_shuffle_multimat_right_work_distribution_n_to_m", 3, method::direct, D,
        {wptb, {"right_matrices_per_thread", arg_kind::integer, 2, nullptr}, rpt, dist_r, streams});
    const std::vector<A> both_wd{wptb,
                                 {"left_matrices_per_thread", arg_kind::integer, 2, nullptr},
                                 {"right_matrices_per_thread", arg_kind::integer, 2, nullptr},
                                 rpt,
                                 dist_t};
    add("nai_shuffle_multimat_both_work_distribution_local_mem_n_to_m", 3, method::direct, D, both_wd);
    add("nai_shuffle_multimat_both_work_distribution_n_to_m", 3, method::direct, D, both_wd);
    add("nai_shuffle_multirow_both_multimat_both_n_to_m", 3, method::direct, D,
        {wptb,
         {"shifts_per_thread_right_matrix", arg_kind::integer, 2, nullptr},
         {"left_matrices_per_thread", arg_kind::integer, 2, nullptr},
         {"right_matrices_per_thread", arg_kind::integer, 2, nullptr},
         {"left_rows_per_iteration", arg_kind::integer, 2, nullptr}});
    add("nai_shuffle_multirow_right_one_to_one", 0, method::direct, D,
        {wptb, {"right_rows_per_thread", arg_kind::integer, 4, nullptr}});
    const std::vector<A> mrb{wptb,
                             {"shifts_per_thread", arg_kind::integer, 4, nullptr},
                             {"left_rows_per_iteration", arg_kind::integer, 3, nullptr}};
    add("nai_shuffle_multirow_both_one_to_one", 0, method::direct, D, mrb);
    add("nai_shuffle_multirow_both_local_mem_one_to_one", 0, method::direct, D, mrb);
    add("nai_shuffle_multirow_right_multimat_right_one_to_one", 0, method::direct, D,
        {wptb, {"right_rows_per_thread", arg_kind::integer, 4, nullptr}});
    add("nai_shuffle_multirow_right_multimat_right_one_to_many", 1, method::direct, D,
        {wptb, {"right_rows_per_thread", arg_kind::integer, 4, nullptr},
         {"right_matrices_per_thread", arg_kind::integer, 4, nullptr}});
    add("nai_shuffle_multirow_right_multimat_right_n_to_mn", 2, method::direct, D,
        {wptb, {"right_rows_per_thread", arg_kind::integer, 4, nullptr},
         {"right_matrices_per_thread", arg_kind::integer, 4, nullptr}, streams});
    const std::vector<A> mbr{wptb,
                             {"shifts_per_thread_right_matrix", arg_kind::integer, 4, nullptr},
                             {"right_matrices_per_thread", arg_kind::integer, 4, nullptr},
                             {"left_rows_per_iteration", arg_kind::integer, 4, nullptr}};
    add("nai_shuffle_multirow_both_multimat_right_one_to_many", 1, method::direct, D, mbr);
    std::vector<A> mbr_s = mbr;
    mbr_s.push_back(streams);
    add("nai_shuffle_multirow_both_multimat_right_n_to_mn", 2, method::direct, D, mbr_s);
    const A spb8{"shifts_per_thread_block", arg_kind::integer, 8, nullptr};
    add("nai_warp_per_shift_one_to_one", 0, method::direct, D, {spb8});
    add("nai_warp_per_shift_simple_indexing_one_to_one", 0, method::direct, D, {spb8});
    add("nai_warp_per_shift_work_distribution_one_to_one", 0, method::direct, D,
        {spb8, {"rows_per_warp", arg_kind::integer, 3, nullptr}, dist_r});
    for (int f = 0; f < 4; ++f) {
        std::vector<A> a{{"shifts_per_thread_block", arg_kind::integer, f == 0 ? 32 : 16, nullptr},
                         {"shared_mem_row_size", arg_kind::integer, 32, nullptr},
                         {"shared_mem_rows", arg_kind::integer, 0, nullptr},
                         {"strided_load", arg_kind::boolean, 1, nullptr},
                         {"column_group_per_block", arg_kind::boolean, 0, nullptr}};
        if (f > 0) a.push_back({"right_matrices_per_block", arg_kind::integer, 8, nullptr});
        if (f > 1) a.push_back(streams);
        add(std::string("nai_warp_per_shift_shared_mem_") + kFamilyNames[f], f, method::direct, D, a);
    }
    add("nai_block_per_shift_one_to_one", 0, method::direct, D, {{"block_size", arg_kind::integer, 128, nullptr}});
    const A htpb{"hadamard_threads_per_block", arg_kind::integer, 256, nullptr};
    const std::vector<A> red{{"scatter_threads_per_block", arg_kind::integer, 256, nullptr},
                             {"scatter_items_per_thread", arg_kind::integer, 10, nullptr},
                             htpb};
    for (int f = 0; f < 3; ++f) {
        add(std::string("fft_orig_") + kFamilyNames[f], f, method::fft, D, {htpb}, true);
        add(std::string("fft_reduced_transfer_") + kFamilyNames[f], f, method::fft, D, red);
    }
    add("fft_better_n_to_m", 3, method::fft, D, {htpb, {"hadamard_items_per_thread", arg_kind::integer, 10, nullptr}});
    // Methods of this library without a counterpart in the reference.
    const A tile{"tile", arg_kind::integer, 0, nullptr};
    const A border{"border", arg_kind::integer, 0, nullptr};
    for (int f = 0; f < 4; ++f) {
        const std::string fam = kFamilyNames[f];
        add("mfma_bf16x3_" + fam, f, method::mfma, XC_PREC_BF16X3, {tile, border});
        add("mfma_fp16x3_" + fam, f, method::mfma, XC_PREC_FP16X3, {tile, border});
        add("mfma_bf16_" + fam, f, method::mfma, XC_PREC_BF16, {tile});
        add("mfma_fp16_" + fam, f, method::mfma, XC_PREC_FP16, {tile});
        add("auto_" + fam, f, method::autosel, XC_PREC_AUTO, {});
    }
    return v;
}

const std::vector<alg_spec>& algorithms() {
    static const std::vector<alg_spec> a = make_algorithms();
    return a;
}

const alg_spec* find_alg(const std::string& name) {
    for (auto& a : algorithms())
        if (a.name == name) return &a;
    return nullptr;
}

std::vector<std::string> sorted_alg_names() {
    std::vector<std::string> v;
    for (auto& a : algorithms()) v.push_back(a.name);
    std::sort(v.begin(), v.end());
    return v;
}

// Values of an algorithm's JSON arguments (with defaults), validated as the reference does:
// tuning counts must be positive and warps_per_thread_block at most 32
// (src/naive_shuffle.cu:461-463); shared_mem_rows (0 = shifts_per_thread_block) must not be
// below shifts_per_thread_block (src/one_to_one.hpp:573-586).
struct resolved_args {
    std::vector<std::pair<std::string, std::string>> props;  // echoed into the timing CSV
    std::map<std::string, long> ints;
    std::map<std::string, std::string> texts;
};

resolved_args resolve_args(const alg_spec& alg, const json_args& js) {
    resolved_args r;
    for (const auto& a : alg.args) {
        if (a.kind == arg_kind::text) {
            const std::string v = js.get_string(a.key, a.def_text);
            if (std::string(a.key) == "distribution_type" && v != "none" && v != "rectangle" && v != "triangle")
                throw std::runtime_error("Unknown distribution " + v);
            r.texts[a.key] = v;
            r.props.emplace_back(a.key, v);
        } else if (a.kind == arg_kind::boolean) {
            const bool v = js.get_bool(a.key, a.def_int != 0);
            r.ints[a.key] = v;
            r.props.emplace_back(a.key, v ? "1" : "0");
        } else {
            long v = js.get_int(a.key, a.def_int);
            r.ints[a.key] = v;
        }
    }
    if (r.ints.count("shared_mem_rows")) {
        long& rows = r.ints["shared_mem_rows"];
        const long spb = r.ints["shifts_per_thread_block"];
        if (rows == 0) rows = spb;
        if (rows < spb)
            throw argument_error("Invalid number of shared memory rows [" + std::to_string(rows) +
                                 "], must be greater than shifts per block [" + std::to_string(spb) + "]");
    }
    for (const auto& a : alg.args)
        if (a.kind == arg_kind::integer) r.props.emplace_back(a.key, std::to_string(r.ints[a.key]));
    // keep the reference's column order: props in declaration order
    std::vector<std::pair<std::string, std::string>> ordered;
    for (const auto& a : alg.args)
        for (auto& p : r.props)
            if (p.first == a.key) {
                ordered.push_back(p);
                break;
            }
    r.props = ordered;
    return r;
}

// Largest accepted value of a tuning key: the limits of the reference's release build
// (CMakeLists.txt:16-45, checked in each src/naive_*.cu run_* function).  The first entry
// whose algorithm-name fragment occurs in the algorithm's name applies.
struct key_limit {
    const char* key;
    const char* alg;  // "" = any algorithm
    long max;
    const char* message;
};
const key_limit kLimits[] = {
    {"warps_per_thread_block", "multirow_both_multimat", 32, "Cuda block too large"},
    {"warps_per_thread_block", "", 32, "Too many warps per thread block"},
    {"shifts_per_thread_block", "", 32, "Too many shifts per thread block"},
    {"block_size", "", 1024, "CUDA block too large"},
    {"right_matrices_per_block", "", 8, "Invalid number of right matrices per block"},
    {"right_matrices_per_thread", "multimat_both", 4, "Too many right matrices per thread"},
    {"right_matrices_per_thread", "multirow", 4, "Invalid number of right matrices per thread"},
    {"right_matrices_per_thread", "", 8, "Invalid number of right matrices per thread"},
    {"left_matrices_per_thread", "", 4, "Too many left matrices per thread"},
    {"shifts_per_thread_right_matrix", "", 4, "Too many shifts per thread in each right matrix"},
    {"shifts_per_thread", "", 8, "Too many shifts per thread"},
    {"left_rows_per_iteration", "", 4, "Too many left rows per iteration"},
    {"right_rows_per_thread", "multimat_right", 4, "Invalid number of right rows per thread"},
    {"right_rows_per_thread", "", 8, "Too many right rows per thread"},
};

// Launch-time checks of the tuning values (errors of the run step, exit code 3).
void check_tuning(const std::string& alg, const resolved_args& r) {
    for (const auto& kv : r.ints) {
        const std::string& k = kv.first;
        if (k == "strided_load" || k == "column_group_per_block" || k == "tile" || k == "border") continue;
        if (kv.second <= 0) throw std::runtime_error("Invalid value " + std::to_string(kv.second) + " for " + k);
        for (const auto& l : kLimits) {
            if (k != l.key || alg.find(l.alg) == std::string::npos) continue;
            if (kv.second > l.max)
                throw std::runtime_error(std::string(l.message) + ": " + std::to_string(kv.second) + " (max " +
                                         std::to_string(l.max) + ")");
            break;
        }
    }
}

// The direct kernel's tile and schedule for a definition-based algorithm: each tuning key of
// the reference's kernel selects the corresponding dimension of the tile (xc_options in
// xcorr.h); dimensions the algorithm has no key for are 1, as in the reference's kernel
// (one shift, one matrix, one row per step).  A 32-lane warp is half a 64-lane wave;
// nai_orig_* keeps the reference's 16 x 16 thread block (4 waves).
xc_options direct_options(const std::string& alg, const resolved_args& r) {
    auto get = [&](const char* k) -> long { return r.ints.count(k) ? r.ints.at(k) : 0; };
    xc_options o{};
    o.precision = XC_PREC_FP32;
    o.shifts_per_thread = o.right_mats_per_thread = o.left_mats_per_thread = o.rows_per_step = 1;
    o.waves_per_block = 4;
    if (long v = get("warps_per_thread_block")) o.waves_per_block = (uint32_t)((v + 1) / 2);
    if (long v = get("shifts_per_thread_block")) o.waves_per_block = (uint32_t)((v + 1) / 2);
    if (long v = get("block_size")) o.waves_per_block = (uint32_t)((v + 63) / 64);
    for (const char* k : {"shifts_per_thread", "shifts_per_thread_right_matrix", "right_rows_per_thread"})
        if (long v = get(k)) o.shifts_per_thread = (uint32_t)v;
    for (const char* k : {"right_matrices_per_thread", "right_matrices_per_block"})
        if (long v = get(k)) o.right_mats_per_thread = (uint32_t)v;
    if (long v = get("left_matrices_per_thread")) o.left_mats_per_thread = (uint32_t)v;
    if (long v = get("left_rows_per_iteration")) o.rows_per_step = (uint32_t)v;
    // work distribution: at most rows_per_thread (rows_per_warp) overlap rows per worker
    long rpt = get("rows_per_thread");
    if (!rpt) rpt = get("rows_per_warp");
    const std::string dist = r.texts.count("distribution_type") ? r.texts.at("distribution_type") : "none";
    if (rpt > 0 && dist == "triangle") o.rows_per_block = (uint32_t)rpt;
    (void)alg;
    return o;
}

// ---------------------------------------------------------------- run options
enum class bench_type { unknown, none, compute, common_steps, algorithm };

bench_type parse_bench(std::string s) {  // src/run_args.cpp:17-32 (case-insensitive)
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (s == "compute") return bench_type::compute;
    if (s == "commonsteps") return bench_type::common_steps;
    if (s == "algorithm") return bench_type::algorithm;
    if (s == "none") return bench_type::none;
    return bench_type::unknown;
}

struct run_options {
    std::string data_type = "single";
    std::string out;
    std::string times = "measurements.csv";
    bool validate = false;
    std::string validate_path;
    bool normalize = false, append = false, no_progress = false;
    bench_type bench = bench_type::none;
    size_t outer_loops = 1;
    double min_time = 1.0;
    std::string args_path;
    std::string alg, ref_path, target_path;
};

std::string loop_out_path(const std::string& out, size_t loop) {  // src/run_args.cpp:142-150
    std::filesystem::path p(out);
    return (p.parent_path() / (p.stem().string() + "_" + std::to_string(loop) + p.extension().string())).string();
}

// ---------------------------------------------------------------- algorithm lifecycle
// load -> prepare -> transfer -> run -> finalize -> free -> store (src/cross_corr.hpp:21-189).
template <typename T>
class algorithm_run {
public:
    algorithm_run(const alg_spec& spec, const resolved_args& args, double min_ns)
        : spec_(spec), args_(args), min_ns_(min_ns) {}
    ~algorithm_run() { release(); }

    void load(const std::string& ref_path, const std::string& target_path) {
        left_ = load_csv<T>(ref_path);
        right_ = load_csv<T>(target_path);
        if (left_.rows != right_.rows || left_.cols != right_.cols)
            throw std::runtime_error("Invalid input matrix sizes, expected ref and target to be the same size: ref = " +
                                     std::to_string(left_.cols) + "x" + std::to_string(left_.rows) +
                                     " target = " + std::to_string(right_.cols) + "x" + std::to_string(right_.rows));
        if (left_.rows == 0 || left_.cols == 0) throw std::runtime_error("Empty input matrices");
        const int64_t k = num_results(spec_.family, left_.count, right_.count);
        if (k < 0) {
            if (spec_.family == XC_N_TO_MN)
                throw std::runtime_error("Invalid ref and target data counts, " + std::to_string(right_.count) +
                                         " is not divisible by " + std::to_string(left_.count));
            throw std::runtime_error("Invalid number of input matrices for " + std::string(kFamilyNames[spec_.family]));
        }
        k_ = (uint32_t)k;
        if (is_fft() && spec_.family == XC_N_TO_MN) {
            // The FFT products need each left matrix's rights contiguous: position g of the
            // grouped order holds file matrix group_perm_[g] (the reference regroups on load,
            // src/matrix.hpp:461-503, and re-interleaves on store, :522-537).
            const uint32_t n = left_.count, m = right_.count, per = m / n;
            group_perm_.resize(m);
            for (uint32_t t = 0; t < m; ++t) group_perm_[(t % n) * per + t / n] = t;
        }
    }

    bool is_fft() const { return spec_.meth == method::fft; }

    void prepare() {
        const uint32_t H = left_.rows, W = left_.cols;
        out_ = is_fft() ? matrix_array<T>(2 * H, 2 * W, k_) : matrix_array<T>(2 * H - 1, 2 * W - 1, k_);
        if (spec_.meth == method::cpu) return;
        const double t0 = now_ns();
        hip_ok(hipMalloc(&d_left_, left_.size() * sizeof(T)), "hipMalloc");
        hip_ok(hipMalloc(&d_right_, right_.size() * sizeof(T)), "hipMalloc");
        hip_ok(hipMalloc(&d_out_, out_.size() * sizeof(T)), "hipMalloc");
        if (is_fft()) prepare_fft();
        alloc_ns_ = now_ns() - t0;
    }

    void transfer() {
        if (spec_.meth == method::cpu) return;
        const double t0 = now_ns();
        hip_ok(hipMemcpy(d_left_, left_.data.data(), left_.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        if (group_perm_.empty()) {
            hip_ok(hipMemcpy(d_right_, right_.data.data(), right_.size() * sizeof(T), hipMemcpyHostToDevice),
                   "hipMemcpy");
        } else {
            std::vector<T> grouped(right_.size());
            for (uint32_t g = 0; g < right_.count; ++g)
                std::copy(right_.matrix(group_perm_[g]), right_.matrix(group_perm_[g]) + right_.area(),
                          grouped.data() + right_.area() * g);
            hip_ok(hipMemcpy(d_right_, grouped.data(), grouped.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        }
        to_device_ns_ = now_ns() - t0;
        if (is_fft()) {
            // fft_orig pads on the host in the reference (load with relative_zero_padding<2>);
            // here both variants zero-pad on the device, with the padding fused into the scatter.
            auto scatter = [&] {
                xc_ok(xc_scatter<T>(d_left_, d_pad_, H(), W(), nl_, 2 * H(), 2 * W()));
                xc_ok(xc_scatter<T>(d_right_, d_pad_ + (size_t)nl_ * 4 * H() * W(), H(), W(), right_.count, 2 * H(),
                                    2 * W()));
            };
            scatter_ = adaptive(measure_alg() ? min_ns_ : 0, true, scatter);
        }
    }

    void run() {
        switch (spec_.meth) {
        case method::cpu: cpu_family<T, T>(spec_.family, left_, right_, out_, 1); return;
        case method::fft: run_fft(); return;
        default: break;
        }
        xc_options o{};
        if (spec_.meth == method::direct) {
            o = direct_options(spec_.name, args_);
            // rectangle distribution: the same number of workers, ceil(rows / rows_per_thread),
            // for every output row
            long rpt = args_.ints.count("rows_per_thread") ? args_.ints.at("rows_per_thread") : 0;
            if (args_.ints.count("rows_per_warp")) rpt = args_.ints.at("rows_per_warp");
            const bool rect = args_.texts.count("distribution_type") && args_.texts.at("distribution_type") == "rectangle";
            if (rpt > 0 && rect) o.split = (uint32_t)((left_.rows + rpt - 1) / rpt);
        }
        o.precision = spec_.meth == method::autosel ? XC_PREC_AUTO : spec_.precision;
        if (args_.ints.count("tile")) o.tile = (uint32_t)args_.ints.at("tile");
        if (args_.ints.count("border")) o.border = (uint32_t)args_.ints.at("border");
        auto body = [&] { xc_ok(call(o)); };
        kernel_ = adaptive(measure_alg() ? min_ns_ : 0, true, body);
    }

    void finalize() {
        if (spec_.meth == method::cpu) return;
        hip_ok(hipMemcpy(out_.data.data(), d_out_, out_.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        if (is_fft() && spec_.family == XC_N_TO_MN) reinterleave();
    }

    void free() { release(); }

    void store(const std::string& path, bool normalize) const {
        if (is_fft() && normalize) store_csv(path, normalize_fft(out_));
        else store_csv(path, out_);
    }

    validation_stats validate(const std::string& valid_path) const {
        const matrix_array<T> res = is_fft() ? normalize_fft(out_) : out_;
        if (!valid_path.empty()) return validate_result(res, load_csv<T>(valid_path));
        matrix_array<T> valid(2 * H() - 1, 2 * W() - 1, k_);
        cpu_family<T, T>(spec_.family, left_, right_, valid, 0);
        return validate_result(res, valid);
    }

    std::vector<std::string> labels() const {
        if (spec_.meth == method::fft) {
            if (spec_.fft_orig) return {"Forward FFT", "Hadamard", "Inverse FFT", "Allocation", "Plan"};
            return {"Forward FFT", "Hadamard", "Inverse FFT", "ToDevice", "Scatter", "Allocation", "Plan"};
        }
        if (spec_.meth == method::cpu) return {};
        return {"Kernel"};
    }

    std::vector<measurement> measurements() const {
        if (spec_.meth == method::fft) {
            measurement alloc{alloc_ns_, 1}, plan{plan_ns_, 1}, todev{to_device_ns_, 1};
            if (spec_.fft_orig) return {fwd_, hadamard_, inv_, alloc, plan};
            return {fwd_, hadamard_, inv_, todev, scatter_, alloc, plan};
        }
        if (spec_.meth == method::cpu) return {};
        return {kernel_};
    }

    void set_measure_alg(bool b) { measure_alg_ = b; }

private:
    const alg_spec& spec_;
    resolved_args args_;
    double min_ns_;
    bool measure_alg_ = false;
    matrix_array<T> left_, right_, out_;
    uint32_t k_ = 0, nl_ = 1;
    T* d_left_ = nullptr;
    T* d_right_ = nullptr;
    T* d_out_ = nullptr;
    T* d_pad_ = nullptr;
    T* d_spec_ = nullptr;   // interleaved complex spectra of the (padded) inputs
    T* d_ospec_ = nullptr;  // n_to_m: spectra of all pairs
    hipfftHandle fwd_plan_ = 0, inv_plan_ = 0;
    measurement kernel_, fwd_, hadamard_, inv_, scatter_;
    double alloc_ns_ = 0, plan_ns_ = 0, to_device_ns_ = 0;
    std::vector<uint32_t> group_perm_;  // n_to_mn FFT: grouped position -> file index of each right matrix

    uint32_t H() const { return left_.rows; }
    uint32_t W() const { return left_.cols; }
    bool measure_alg() const { return measure_alg_; }

    xc_status call(const xc_options& o) {
        if constexpr (sizeof(T) == 4)
            return xc_cross_corr_f32((xc_family)spec_.family, d_left_, d_right_, d_out_, H(), W(), left_.count,
                                     right_.count, &o, nullptr);
        else
            return xc_cross_corr_f64((xc_family)spec_.family, d_left_, d_right_, d_out_, H(), W(), left_.count,
                                     right_.count, &o, nullptr);
    }

    template <typename U>
    static xc_status xc_scatter(const U* src, U* dst, uint32_t r, uint32_t c, uint32_t n, uint32_t dr, uint32_t dc) {
        if constexpr (sizeof(U) == 4) return xc_scatter_f32(src, dst, r, c, n, dr, dc, 0, 0, nullptr);
        else return xc_scatter_f64(src, dst, r, c, n, dr, dc, 0, 0, nullptr);
    }

    void prepare_fft() {
        const uint32_t n = left_.count, m = right_.count;
        nl_ = (spec_.family == XC_ONE_TO_ONE || spec_.family == XC_ONE_TO_MANY) ? 1 : n;
        const size_t parea = (size_t)4 * H() * W();
        const size_t sarea = (size_t)2 * H() * (W() + 1);  // complex values per spectrum
        hip_ok(hipMalloc(&d_pad_, (nl_ + m) * parea * sizeof(T)), "hipMalloc");
        hip_ok(hipMalloc(&d_spec_, (nl_ + m) * sarea * 2 * sizeof(T)), "hipMalloc");
        if (spec_.family == XC_N_TO_M) hip_ok(hipMalloc(&d_ospec_, (size_t)k_ * sarea * 2 * sizeof(T)), "hipMalloc");
        const double t0 = now_ns();
        int dims[2] = {(int)(2 * H()), (int)(2 * W())};
        const hipfftType r2c = sizeof(T) == 4 ? HIPFFT_R2C : HIPFFT_D2Z;
        const hipfftType c2r = sizeof(T) == 4 ? HIPFFT_C2R : HIPFFT_Z2D;
        fft_ok(hipfftPlanMany(&fwd_plan_, 2, dims, nullptr, 1, 0, nullptr, 1, 0, r2c, (int)(nl_ + m)), "hipfftPlanMany");
        fft_ok(hipfftPlanMany(&inv_plan_, 2, dims, nullptr, 1, 0, nullptr, 1, 0, c2r, (int)k_), "hipfftPlanMany");
        plan_ns_ = now_ns() - t0;
    }

    void run_fft() {
        const uint32_t m = right_.count;
        const size_t sarea = (size_t)2 * H() * (W() + 1);
        const double mn = measure_alg() ? min_ns_ : 0;
        fwd_ = adaptive(mn, false, [&] {
            if constexpr (sizeof(T) == 4)
                fft_ok(hipfftExecR2C(fwd_plan_, (float*)d_pad_, (hipfftComplex*)d_spec_), "hipfftExecR2C");
            else
                fft_ok(hipfftExecD2Z(fwd_plan_, (double*)d_pad_, (hipfftDoubleComplex*)d_spec_), "hipfftExecD2Z");
            hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        });
        // The in-place product overwrites the right spectra, so the timed repetitions redo the
        // forward transform first and time the product on its own with events.
        T* rspec = d_spec_ + nl_ * sarea * 2;
        auto product = [&] {
            if (spec_.family == XC_N_TO_M) {
                if constexpr (sizeof(T) == 4)
                    xc_ok(xc_hadamard_n_to_m_f32(d_spec_, rspec, d_ospec_, sarea, nl_, m, nullptr));
                else
                    xc_ok(xc_hadamard_n_to_m_f64(d_spec_, rspec, d_ospec_, sarea, nl_, m, nullptr));
            } else {
                const uint32_t per = m / nl_;
                for (uint32_t l = 0; l < nl_; ++l) {
                    if constexpr (sizeof(T) == 4)
                        xc_ok(xc_hadamard_f32(d_spec_ + l * sarea * 2, rspec + (size_t)l * per * sarea * 2, sarea, per,
                                              nullptr));
                    else
                        xc_ok(xc_hadamard_f64(d_spec_ + l * sarea * 2, rspec + (size_t)l * per * sarea * 2, sarea, per,
                                              nullptr));
                }
            }
        };
        if (measure_alg() && spec_.family != XC_N_TO_M) {
            hadamard_ = adaptive(mn, true, product);
            // restore the spectra the repetitions conjugated repeatedly
            if constexpr (sizeof(T) == 4)
                fft_ok(hipfftExecR2C(fwd_plan_, (float*)d_pad_, (hipfftComplex*)d_spec_), "hipfftExecR2C");
            else
                fft_ok(hipfftExecD2Z(fwd_plan_, (double*)d_pad_, (hipfftDoubleComplex*)d_spec_), "hipfftExecD2Z");
            product();
        } else {
            hadamard_ = adaptive(mn, true, product);
        }
        T* ispec = spec_.family == XC_N_TO_M ? d_ospec_ : rspec;
        // the inverse transform may overwrite its input: repetitions after the first time
        // garbage, and the result is recomputed once below
        inv_ = adaptive(mn, false, [&] {
            if constexpr (sizeof(T) == 4)
                fft_ok(hipfftExecC2R(inv_plan_, (hipfftComplex*)ispec, (float*)d_out_), "hipfftExecC2R");
            else
                fft_ok(hipfftExecZ2D(inv_plan_, (hipfftDoubleComplex*)ispec, (double*)d_out_), "hipfftExecZ2D");
            hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        });
        if (inv_.iterations > 1) {  // recompute the result once from fresh spectra
            if constexpr (sizeof(T) == 4)
                fft_ok(hipfftExecR2C(fwd_plan_, (float*)d_pad_, (hipfftComplex*)d_spec_), "hipfftExecR2C");
            else
                fft_ok(hipfftExecD2Z(fwd_plan_, (double*)d_pad_, (hipfftDoubleComplex*)d_spec_), "hipfftExecD2Z");
            product();
            if constexpr (sizeof(T) == 4)
                fft_ok(hipfftExecC2R(inv_plan_, (hipfftComplex*)ispec, (float*)d_out_), "hipfftExecC2R");
            else
                fft_ok(hipfftExecZ2D(inv_plan_, (hipfftDoubleComplex*)ispec, (double*)d_out_), "hipfftExecZ2D");
        }
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }

    void reinterleave() {  // results of regrouped rights back to file order
        matrix_array<T> file(out_.rows, out_.cols, out_.count);
        for (uint32_t g = 0; g < out_.count; ++g)
            std::copy(out_.matrix(g), out_.matrix(g) + out_.area(), file.matrix(group_perm_[g]));
        out_ = std::move(file);
    }

    void release() {
        if (fwd_plan_) (void)hipfftDestroy(fwd_plan_);
        if (inv_plan_) (void)hipfftDestroy(inv_plan_);
        fwd_plan_ = inv_plan_ = 0;
        for (T** p : {&d_left_, &d_right_, &d_out_, &d_pad_, &d_spec_, &d_ospec_})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
    }
};

// output_measurements, src/main.cpp:81-121: header "<label>,<label>_iterations,...,<arg keys>".
void output_measurements(const std::string& path, const std::vector<std::string>& labels,
                         const std::vector<measurement>& ms,
                         const std::vector<std::pair<std::string, std::string>>& props, bool append) {
    std::ofstream f;
    if (append) {
        f.open(path, std::ios_base::app);
    } else {
        f.open(path);
        std::string sep;
        for (auto& l : labels) {
            f << sep << l << "," << l << "_iterations";
            sep = ",";
        }
        for (auto& p : props) {
            f << sep << p.first;
            sep = ",";
        }
        f << "\n";
    }
    std::string sep;
    for (auto& m : ms) {
        f << sep << (long long)std::llround(m.iteration_ns()) << "," << m.iterations;
        sep = ",";
    }
    for (auto& p : props) {
        f << sep << p.second;
        sep = ",";
    }
    f << "\n";
}

struct logger {  // simple_logger, src/simple_logger.hpp:14-47
    bool progress;
    void log(const std::string& m) const {
        if (progress) std::cout << m << std::endl;
    }
};

template <typename T>
int run_measurement(const run_options& ro, const alg_spec& spec) {
    const bool compute = ro.bench == bench_type::compute;
    const bool common = ro.bench == bench_type::common_steps;
    logger log{!ro.no_progress && !compute};
    json_args js;
    if (!ro.args_path.empty()) js = json_args::load(ro.args_path);
    const resolved_args args = resolve_args(spec, js);
    const double min_ns = ro.min_time * 1e9;
    algorithm_run<T> alg(spec, args, min_ns);
    alg.set_measure_alg(ro.bench == bench_type::algorithm);

    measurement m_load;
    try {
        log.log("Loading inputs");
        m_load = once([&] { alg.load(ro.ref_path, ro.target_path); });
    } catch (std::exception& e) {
        std::cerr << "Exception occured durign data loading step: " << e.what() << std::endl;
        return 3;
    }

    for (size_t loop = 0; loop < ro.outer_loops; ++loop) {
        std::vector<measurement> steps(7);
        if (loop == 0) steps[0] = m_load;
        auto step = [&](const char* msg, const char* what, int idx, auto&& fn) -> int {
            try {
                log.log(msg);
                const double t0 = now_ns();
                fn();
                if (idx >= 0) steps[idx] = {now_ns() - t0, 1};
            } catch (std::exception& e) {
                std::cerr << "Exception occured durign " << what << " step: " << e.what()
                          << "\nLast kernel launch arguments:\n" << last_launch_string() << std::endl;
                return 3;
            }
            return 0;
        };
        auto computation = [&]() -> int {
            // Tuning values are checked before any device work, so a bad value fails the
            // same way with or without a GPU.
            if (int rc = step("Allocating", "allocation", 1, [&] {
                    check_tuning(spec.name, args);
                    alg.prepare();
                }))
                return rc;
            if (int rc = step("Transfering data", "data transfer", 2, [&] { alg.transfer(); })) return rc;
            if (int rc = step("Running test alg", "computation", -1, [&] {
                    if (common) {
                        steps[3] = adaptive(min_ns, false, [&] {
                            alg.run();
                            if (spec.meth != method::cpu) hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
                        });
                    } else {
                        alg.run();
                    }
                }))
                return rc;
            if (int rc = step("Copying output data to host", "finalization", 4, [&] { alg.finalize(); })) return rc;
            if (int rc = step("Free resources", "free", 5, [&] { alg.free(); })) return rc;
            return 0;
        };
        measurement m_compute;
        int rc = 0;
        if (compute) {
            uint64_t iters = 1;
            while (iters) {
                const double t0 = now_ns();
                for (uint64_t i = 0; i < iters && rc == 0; ++i) rc = computation();
                m_compute = {now_ns() - t0, iters};
                if (rc) break;
                iters = next_iteration_count(min_ns, m_compute);
            }
        } else {
            rc = computation();
        }
        if (rc) return rc;
        try {
            if (spec.meth != method::cpu) hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
            if (spec.meth != method::cpu) hip_ok(hipGetLastError(), "hipGetLastError");
        } catch (std::exception& e) {
            std::cerr << "Exception somewhere in HIP code: " << e.what() << "\nLast kernel launch arguments:\n"
                      << last_launch_string() << std::endl;
            return 3;
        }

        const std::string out = ro.out.empty() ? "" : (ro.outer_loops > 1 ? loop_out_path(ro.out, loop) : ro.out);
        if (!out.empty()) {
            log.log(alg.is_fft() ? "Normalizing and storing results" : "Storing results");
            const double t0 = now_ns();
            alg.store(out, ro.normalize);
            steps[6] = {now_ns() - t0, 1};
        }
        if (ro.bench != bench_type::none) {
            log.log("Collect measurements");
            if (compute)
                output_measurements(ro.times, {"Computation"}, {m_compute}, args.props, ro.append || loop > 0);
            else if (common)
                output_measurements(ro.times, {"Load", "Prepare", "Transfer", "Run", "Finalize", "Free", "Store"},
                                    steps, args.props, ro.append || loop > 0);
            else
                output_measurements(ro.times, alg.labels(), alg.measurements(), args.props, ro.append || loop > 0);
        }
        if (!ro.validate) {
            log.log("No validation");
        } else {
            if (!ro.validate_path.empty()) log.log("Validating results against " + ro.validate_path);
            else log.log("Computing valid results and validating");
            print_stats(alg.validate(ro.validate_path), !ro.no_progress, !ro.append && loop == 0);
        }
    }
    return 0;
}

void print_help(std::ostream& out, const std::string& name) {
    out << "Usage: " << name << " [global options] command [command options]\n"
        << "Commands: \n"
        << "\t" << name << " [global options] list\n"
        << "\t" << name << " [global options] run [run options] <alg> <ref_path> <target_path>\n"
        << "\t" << name << " [global options] validate [validate options] <validate_data_path> <template_data_path>\n"
        << "\t" << name
        << " [global options] input [input options] <alg_type> <rows> <columns> <left_matrices> <right_matrices>\n"
        << "Global options:\n"
        << "  -h [ --help ]                 display this help and exit\n"
        << "Validate options:\n"
        << "  -n [ --normalize ]            Normalize the data to be validated as they are denormalized fft output\n"
        << "  -c [ --csv ]                  CSV output\n"
        << "  -p [ --print_header ]         Print header\n"
        << "Run options:\n"
        << "  -d [ --data_type ] arg (=single)          Data type to use for computation\n"
        << "  -o [ --out ] arg                          Path of the output file to be created\n"
        << "  -t [ --times ] arg (=measurements.csv)    File to store the measured times in\n"
        << "  -v [ --validate ] [=arg(=)]               If validation of the results should be done and optionally path "
           "to a file containing the valid results\n"
        << "  -n [ --normalize ]                        If algorithm is fft, normalize the results\n"
        << "  -a [ --append ]                           Append time measurements without the header\n"
        << "  -p [ --no_progress ]                      Do not print human readable progress\n"
        << "  -b [ --benchmark_type ] arg (=None)       Which part should be measured: Compute, CommonSteps, Algorithm\n"
        << "  -l [ --outer_loops ] arg (=1)             How many measurements of the computation loop should be done\n"
        << "  -m [ --min_time ] arg (=1)                The minimum time to consider measurement statistically relevant, "
           "in seconds\n"
        << "  --args_path arg                           Path to the JSON file containing arguments for the algorithm\n";
}

// Option parser: long (--name, --name=value), short (-x, -xvalue) and positionals.
struct parsed {
    std::map<std::string, std::string> opts;
    std::vector<std::string> pos;
};

parsed parse_args(const std::vector<std::string>& a,
                  const std::map<std::string, std::pair<std::string, int>>& spec) {  // name -> (long, 0 flag / 1 value / 2 implicit)
    parsed p;
    std::map<std::string, std::string> short_to_long;
    for (auto& kv : spec)
        if (kv.first.size() == 1) short_to_long[kv.first] = kv.second.first;
    auto kind_of = [&](const std::string& lng) {
        for (auto& kv : spec)
            if (kv.second.first == lng) return kv.second.second;
        return -1;
    };
    for (size_t i = 0; i < a.size(); ++i) {
        const std::string& s = a[i];
        if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
            std::string name = s.substr(2), value;
            const size_t eq = name.find('=');
            bool has = eq != std::string::npos;
            if (has) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
            }
            const int k = kind_of(name);
            if (k < 0) throw usage_error("unrecognised option '--" + name + "'");
            if (k == 1 && !has) {
                if (i + 1 >= a.size()) throw usage_error("the required argument for option '--" + name + "' is missing");
                value = a[++i];
            } else if (k == 2 && !has && i + 1 < a.size() && a[i + 1][0] != '-' && !find_alg(a[i + 1])) {
                value = a[++i];
            }
            p.opts[name] = value;
        } else if (s.size() >= 2 && s[0] == '-' && !std::isdigit((unsigned char)s[1])) {
            const std::string sh = s.substr(1, 1);
            auto it = short_to_long.find(sh);
            if (it == short_to_long.end()) throw usage_error("unrecognised option '" + s + "'");
            const std::string name = it->second;
            const int k = kind_of(name);
            std::string value;
            if (k == 0 && s.size() > 2) throw usage_error("option '-" + sh + "' takes no value");
            if (k == 1) {
                if (s.size() > 2) value = s.substr(2);
                else if (i + 1 < a.size()) value = a[++i];
                else throw usage_error("the required argument for option '--" + name + "' is missing");
            } else if (k == 2) {
                if (s.size() > 2) value = s.substr(2);
                else if (i + 1 < a.size() && a[i + 1][0] != '-' && !find_alg(a[i + 1])) value = a[++i];
            }
            p.opts[name] = value;
        } else {
            p.pos.push_back(s);
        }
    }
    return p;
}

int cmd_run(const std::vector<std::string>& a, const std::string& prog) {
    const parsed p = parse_args(a, {{"d", {"data_type", 1}},
                                    {"data_type", {"data_type", 1}},
                                    {"o", {"out", 1}},
                                    {"out", {"out", 1}},
                                    {"t", {"times", 1}},
                                    {"times", {"times", 1}},
                                    {"v", {"validate", 2}},
                                    {"validate", {"validate", 2}},
                                    {"n", {"normalize", 0}},
                                    {"normalize", {"normalize", 0}},
                                    {"a", {"append", 0}},
                                    {"append", {"append", 0}},
                                    {"p", {"no_progress", 0}},
                                    {"no_progress", {"no_progress", 0}},
                                    {"b", {"benchmark_type", 1}},
                                    {"benchmark_type", {"benchmark_type", 1}},
                                    {"l", {"outer_loops", 1}},
                                    {"outer_loops", {"outer_loops", 1}},
                                    {"m", {"min_time", 1}},
                                    {"min_time", {"min_time", 1}},
                                    {"args_path", {"args_path", 1}}});
    if (p.pos.size() < 3) throw usage_error("the option '--target_path' is required but missing");
    if (p.pos.size() > 3) throw usage_error("too many positional options have been specified on the command line");
    run_options ro;
    ro.alg = p.pos[0];
    ro.ref_path = p.pos[1];
    ro.target_path = p.pos[2];
    auto get = [&](const char* k) -> const std::string* {
        auto it = p.opts.find(k);
        return it == p.opts.end() ? nullptr : &it->second;
    };
    if (auto v = get("data_type")) ro.data_type = *v;
    if (auto v = get("out")) ro.out = *v;
    if (auto v = get("times")) ro.times = *v;
    if (auto v = get("validate")) {
        ro.validate = true;
        ro.validate_path = *v;
    }
    ro.normalize = get("normalize") != nullptr;
    ro.append = get("append") != nullptr;
    ro.no_progress = get("no_progress") != nullptr;
    if (auto v = get("benchmark_type")) ro.bench = parse_bench(*v);
    try {
        if (auto v = get("outer_loops")) ro.outer_loops = std::stoul(*v);
        if (auto v = get("min_time")) ro.min_time = std::stod(*v);
    } catch (std::exception&) {
        throw usage_error("the argument for option '--outer_loops' or '--min_time' is invalid");
    }
    if (auto v = get("args_path")) ro.args_path = *v;

    const alg_spec* spec = find_alg(ro.alg);
    if (!spec) {  // src/run_args.cpp:114-117
        std::cerr << "Unknown algorithm \"" << ro.alg << "\", expected one of [";
        std::string sep;
        for (auto& n : sorted_alg_names()) {
            std::cerr << sep << n;
            sep = ", ";
        }
        std::cerr << "]" << std::endl;
        print_help(std::cerr, prog);
        return 1;
    }
    if (ro.bench == bench_type::unknown) {
        std::cerr << "Unknown benchmark type" << std::endl;
        print_help(std::cerr, prog);
        return 1;
    }
    if (ro.data_type == "single") return run_measurement<float>(ro, *spec);
    if (ro.data_type == "double") return run_measurement<double>(ro, *spec);
    std::cerr << "Unknown data type " << ro.data_type << "\n";
    print_help(std::cerr, prog);
    return 1;
}

int cmd_validate(const std::vector<std::string>& a) {  // src/main.cpp:56-79, always double
    const parsed p = parse_args(a, {{"n", {"normalize", 0}},
                                    {"normalize", {"normalize", 0}},
                                    {"c", {"csv", 0}},
                                    {"csv", {"csv", 0}},
                                    {"p", {"print_header", 0}},
                                    {"print_header", {"print_header", 0}}});
    if (p.pos.size() < 2) throw usage_error("the option '--validate_data_path' is required but missing");
    const bool normalize = p.opts.count("normalize"), csv = p.opts.count("csv"), header = p.opts.count("print_header");
    const matrix_array<double> valid = load_csv<double>(p.pos[0]);
    for (size_t i = 1; i < p.pos.size(); ++i) {
        if (!csv) std::cout << "Validating " << p.pos[i] << std::endl;
        matrix_array<double> target = load_csv<double>(p.pos[i]);
        if (normalize) target = normalize_fft(target);
        print_stats(validate_result(target, valid), !csv, i == 1 && header);
    }
    return 0;
}

int cmd_input(const std::vector<std::string>& a, const std::string& prog) {  // src/main.cpp:330-359
    const parsed p = parse_args(a, {});
    if (p.pos.size() < 5) throw usage_error("the option '--right_matrices' is required but missing");
    const std::string& type = p.pos[0];
    int fam = -1;
    for (int f = 0; f < 4; ++f)
        if (type == kFamilyNames[f]) fam = f;
    if (fam < 0) {
        std::cerr << "Unknown algorithm type \"" << type << "\", expected one of [n_to_m, n_to_mn, one_to_many, one_to_one]"
                  << std::endl;
        print_help(std::cerr, prog);
        return 1;
    }
    long v[4];
    for (int i = 0; i < 4; ++i) {
        try {
            v[i] = std::stol(p.pos[1 + i]);
        } catch (std::exception&) {
            throw usage_error("the argument ('" + p.pos[1 + i] + "') is invalid");
        }
    }
    const bool ok = v[0] > 0 && v[1] > 0 && num_results(fam, (uint32_t)std::max(0L, v[2]), (uint32_t)std::max(0L, v[3])) > 0 &&
                    v[2] > 0 && v[3] > 0;
    std::cout << (ok ? "Valid" : "Invalid") << "\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string prog = argv[0];
    try {
        std::vector<std::string> args(argv + 1, argv + argc);
        if (args.empty() || args[0] == "-h" || args[0] == "--help") {
            print_help(std::cout, prog);
            return 0;
        }
        const std::string cmd = args[0];
        const std::vector<std::string> rest(args.begin() + 1, args.end());
        for (auto& r : rest)
            if (r == "-h" || r == "--help") {
                print_help(std::cout, prog);
                return 0;
            }
        if (cmd == "run") return cmd_run(rest, prog);
        if (cmd == "validate") return cmd_validate(rest);
        if (cmd == "list") {
            for (auto& n : sorted_alg_names()) std::cout << n << "\n";
            return 0;
        }
        if (cmd == "input") return cmd_input(rest, prog);
        std::cerr << "Unknown command " << cmd << "\n";
        print_help(std::cerr, prog);
        return 1;
    } catch (usage_error& e) {
        std::cerr << "Invalid commandline options: " << e.what() << std::endl;
        return 1;
    } catch (argument_error& e) {
        std::cerr << "Invalid algorithm argument value: " << e.what() << std::endl;
        return 1;
    } catch (std::exception& e) {
        std::cerr << "Exception occured: " << e.what() << std::endl;
        return 2;
    }
}
