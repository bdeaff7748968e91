// Method selection for XC_PREC_AUTO: the candidate list, the cost model's features, the
// model file, host inference of the MLP, the exhaustive autotuner and its table.  No
// counterpart in the reference (its user names the algorithm, src/main.cpp:368-409).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "xc_methods.hpp"
#include "xc_mlp.hpp"

namespace xc {

namespace {

constexpr uint32_t kFeatures = 16;

// ------------------------------------------------------------------ candidates
std::string candidate_name(uint32_t i) {
    const xc_options o = candidate(i);
    const char* p = o.precision == XC_PREC_FP32 ? "fp32" : o.precision == XC_PREC_BF16X3 ? "bf16x3" : "fft64";
    return std::string(p) + "/" + std::to_string(o.tile);
}

// ------------------------------------------------------------------ model
struct model {
    bool loaded = false;
    std::string path;
    std::vector<std::string> cands;
    xc_mlp_shape shape{};
    std::vector<float> params;
    std::vector<double> f_mean, f_std, f_min, f_max, y_mean, y_std;
    std::vector<double> c_max_macs;  // per candidate: largest log2 MACs with a measured time
};

std::mutex g_mu;
model g_model;
bool g_model_tried = false;
std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, int>, uint32_t> g_tuned;
// The model's answer per workload (a prediction costs a few microseconds of host time).
std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> g_predicted;

std::string default_model_path() {
    if (const char* e = std::getenv("XC_COST_MODEL")) return e;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&default_model_path), &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
        return dir + "/tuning/cost_model.txt";
    }
    return "tuning/cost_model.txt";
}

template <typename T>
void read_list(std::istream& in, const char* key, std::vector<T>& out) {
    std::string k;
    size_t n = 0;
    if (!(in >> k >> n) || k != key) fail(XC_ERR_INVALID_ARGUMENT, std::string("cost model file: expected ") + key);
    out.resize(n);
    for (size_t i = 0; i < n; ++i)
        if (!(in >> out[i])) fail(XC_ERR_INVALID_ARGUMENT, std::string("cost model file: short list ") + key);
}

model parse_model(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail(XC_ERR_UNSUPPORTED, "cost model file not found: " + path);
    model m;
    m.path = path;
    std::string magic;
    int version = 0;
    if (!(in >> magic >> version) || magic != "xcorr-cost-model" || version != 2)
        fail(XC_ERR_INVALID_ARGUMENT, "not a cost model file (version 2): " + path);
    read_list(in, "candidates", m.cands);
    std::vector<uint32_t> dims;
    read_list(in, "dims", dims);
    if (dims.size() < 2 || dims.size() > 8) fail(XC_ERR_INVALID_ARGUMENT, "cost model file: bad dims");
    m.shape.layers = (uint32_t)dims.size() - 1;
    for (size_t i = 0; i < dims.size(); ++i) m.shape.dims[i] = dims[i];
    mlp_check_shape(&m.shape);
    read_list(in, "feat_mean", m.f_mean);
    read_list(in, "feat_std", m.f_std);
    read_list(in, "feat_min", m.f_min);
    read_list(in, "feat_max", m.f_max);
    read_list(in, "out_mean", m.y_mean);
    read_list(in, "out_std", m.y_std);
    read_list(in, "cand_max_log2_macs", m.c_max_macs);
    read_list(in, "params", m.params);
    const uint32_t nin = m.shape.dims[0], nout = m.shape.dims[m.shape.layers];
    if (nin != kFeatures || m.f_mean.size() != nin || m.f_std.size() != nin || m.f_min.size() != nin ||
        m.f_max.size() != nin || m.cands.size() != nout || m.y_mean.size() != nout || m.y_std.size() != nout ||
        m.c_max_macs.size() != nout ||
        m.params.size() != mlp_num_params(m.shape))
        fail(XC_ERR_INVALID_ARGUMENT, "cost model file: inconsistent sizes in " + path);
    m.loaded = true;
    return m;
}

void check_candidates(const model& m) {
    const uint32_t n = num_candidates();
    bool ok = m.cands.size() == n;
    for (uint32_t i = 0; ok && i < n; ++i) ok = m.cands[i] == candidate_name(i);
    if (!ok) fail(XC_ERR_UNSUPPORTED, "cost model " + m.path + " was trained for another candidate list");
}

// Host inference (double precision) of the model's MLP; returns predicted ms per candidate.
void predict(const model& m, const float* feat, float* ms) {
    // Features are clamped to the training range first: the network is not trusted outside it.
    std::vector<double> a(feat, feat + kFeatures), b;
    for (uint32_t i = 0; i < kFeatures; ++i)
        a[i] = (std::min(std::max(a[i], m.f_min[i]), m.f_max[i]) - m.f_mean[i]) / m.f_std[i];
    size_t off = 0;
    for (uint32_t l = 0; l < m.shape.layers; ++l) {
        const uint32_t din = m.shape.dims[l], dout = m.shape.dims[l + 1];
        const float* W = m.params.data() + off;
        const float* bias = W + (size_t)din * dout;
        b.assign(dout, 0.0);
        for (uint32_t o = 0; o < dout; ++o) {
            double z = bias[o];
            for (uint32_t i = 0; i < din; ++i) z += (double)W[(size_t)o * din + i] * a[i];
            b[o] = l + 1 < m.shape.layers ? std::tanh(z) : z;
        }
        off += (size_t)din * dout + dout;
        a.swap(b);
    }
    for (size_t c = 0; c < a.size(); ++c) ms[c] = (float)std::exp2(a[c] * m.y_std[c] + m.y_mean[c]);
}

bool in_domain(const model& m, const float* f) {
    // Matrix size features (log2 rows, log2 cols) must lie in the training range; batch
    // counts may extrapolate (time is close to linear in the number of pairs).
    for (int i : {0, 1})
        if (f[i] < m.f_min[i] - 0.25 || f[i] > m.f_max[i] + 0.25) return false;
    return true;
}

void ensure_model_locked() {
    if (g_model_tried) return;
    g_model_tried = true;
    try {
        model m = parse_model(default_model_path());
        check_candidates(m);
        g_model = std::move(m);
    } catch (const error&) {
        g_model = model{};  // AUTO falls back to autotuning
    }
}

std::string auto_mode() {
    const char* e = std::getenv("XC_AUTO");
    return e ? std::string(e) : std::string();
}

// ------------------------------------------------------------------ timing table (CSV)
struct table {
    std::vector<std::string> cands;
    std::vector<xc_workload> w;
    std::vector<std::vector<double>> ms;
};

table read_table(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail(XC_ERR_INVALID_ARGUMENT, "cannot open timing table " + path);
    table t;
    std::string line;
    if (!std::getline(in, line)) fail(XC_ERR_INVALID_ARGUMENT, "empty timing table " + path);
    std::vector<std::string> head;
    {
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) head.push_back(cell);
    }
    static const char* fixed[] = {"family", "rows", "cols", "n", "m"};
    if (head.size() < 6) fail(XC_ERR_INVALID_ARGUMENT, "timing table header too short");
    for (int i = 0; i < 5; ++i)
        if (head[i] != fixed[i]) fail(XC_ERR_INVALID_ARGUMENT, "timing table header must start family,rows,cols,n,m");
    t.cands.assign(head.begin() + 5, head.end());
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string cell;
        std::vector<std::string> cells;
        while (std::getline(ss, cell, ',')) cells.push_back(cell);
        if (cells.size() != head.size()) fail(XC_ERR_INVALID_ARGUMENT, "timing table row with wrong column count");
        xc_workload w{};
        w.family = (uint32_t)std::stoul(cells[0]);
        w.rows = (uint32_t)std::stoul(cells[1]);
        w.cols = (uint32_t)std::stoul(cells[2]);
        w.n = (uint32_t)std::stoul(cells[3]);
        w.m = (uint32_t)std::stoul(cells[4]);
        std::vector<double> v;
        for (size_t i = 5; i < cells.size(); ++i) v.push_back(std::strtod(cells[i].c_str(), nullptr));
        t.w.push_back(w);
        t.ms.push_back(v);
    }
    if (t.w.empty()) fail(XC_ERR_INVALID_ARGUMENT, "timing table has no rows");
    return t;
}

void write_list(std::ostream& out, const char* key, const std::vector<double>& v) {
    out << key << ' ' << v.size();
    char buf[40];
    for (double x : v) {
        std::snprintf(buf, sizeof(buf), " %.9g", x);
        out << buf;
    }
    out << '\n';
}

}  // namespace

uint32_t num_candidates() { return 4 + (uint32_t)mfma_catalog_size(); }

xc_options candidate(uint32_t i) {
    xc_options o{};
    const uint32_t nt = (uint32_t)mfma_catalog_size();
    if (i == 0) {
        o.precision = XC_PREC_FP32;
    } else if (i <= nt) {
        o.precision = XC_PREC_BF16X3;
        o.tile = mfma_candidate_tile((int)i - 1);
    } else if (i >= nt + 1 && i <= nt + 3) {
        o.precision = XC_PREC_FFT64;
        o.tile = i - nt - 1;  // 0: rocFFT, 1: own FFT kernels, 2: own FFT with power-of-two rows
    } else {
        fail(XC_ERR_INVALID_ARGUMENT, "candidate index out of range");
    }
    return o;
}

// Features: log2 of rows, cols, n, m, results and MACs; log2 of the FFT points transformed
// at the reference's padding (2H x 2W); log2 of the own FFT's padded area (fftk_col_len x
// fftk_row_len) over the needed (2H-1)(2W-1); log2 of the largest prime factor of 2H and of
// 2W (how well the reference padding factors); the family one-hot; log2 of the own-FFT
// method's butterfly work (row transforms of every input, two column transforms and one row
// transform per result, each n log2 n at its lengths); log2 of the power-of-two row length
// over the own FFT's row length (0, or log2(4/3) where rows of 3 * 2^n are shorter).
uint32_t workload_features(const xc_workload& w, float* f, uint32_t cap) {
    const int64_t k = num_results((int)w.family, w.n, w.m);
    if (w.family > XC_N_TO_M || k < 0 || w.rows == 0 || w.cols == 0)
        fail(XC_ERR_INVALID_ARGUMENT, "invalid workload for features");
    if (cap < kFeatures) fail(XC_ERR_INVALID_ARGUMENT, "feature buffer too small");
    const double lr = std::log2((double)w.rows), lc = std::log2((double)w.cols);
    const double lk = std::log2((double)k);
    double transforms = 0;
    switch (w.family) {
    case XC_ONE_TO_ONE: transforms = 3; break;
    case XC_ONE_TO_MANY: transforms = 1.0 + 2.0 * w.m; break;
    case XC_N_TO_MN: transforms = (double)w.n + 2.0 * w.m; break;
    default: transforms = (double)w.n + w.m + (double)k; break;
    }
    auto lpf = [](uint32_t x) {
        uint32_t best = 1;
        for (uint32_t p = 2; p * p <= x; ++p)
            while (x % p == 0) {
                best = p;
                x /= p;
            }
        return x > 1 ? std::max(best, x) : best;
    };
    const double pr = fftk_col_len(w.rows), pc = fftk_row_len(w.cols), pc2 = fftk_col_len(w.cols);
    f[0] = (float)lr;
    f[1] = (float)lc;
    f[2] = (float)std::log2((double)w.n);
    f[3] = (float)std::log2((double)w.m);
    f[4] = (float)lk;
    f[5] = (float)(lk + 2 * lr + 2 * lc);
    f[6] = (float)(std::log2(transforms) + 2 + lr + lc);
    f[7] = (float)std::log2(pr * pc / ((2.0 * w.rows - 1) * (2.0 * w.cols - 1)));
    f[8] = (float)std::log2((double)lpf(2 * w.rows));
    f[9] = (float)std::log2((double)lpf(2 * w.cols));
    for (uint32_t i = 0; i < 4; ++i) f[10 + i] = w.family == i ? 1.f : 0.f;
    const double nin = w.family == XC_ONE_TO_ONE ? 2.0 : (w.family == XC_ONE_TO_MANY ? 1.0 + w.m : (double)w.n + w.m);
    const double own = nin * (double)w.rows * pc * std::log2(pc) / 2 + (double)k * (pc / 2 + 1) * pr * std::log2(pr) * 2 +
                       (double)k * pr * pc * std::log2(pc) / 2;
    f[14] = (float)std::log2(own);
    f[15] = (float)std::log2(pc2 / pc);
    return kFeatures;
}

void cost_model_load(const char* path) {
    model m = parse_model(path ? std::string(path) : default_model_path());
    check_candidates(m);
    std::lock_guard<std::mutex> lock(g_mu);
    g_model = std::move(m);
    g_model_tried = true;
    g_predicted.clear();
}

void cost_model_predict(const xc_workload& w, float* ms) {
    float f[kFeatures];
    workload_features(w, f, kFeatures);
    std::lock_guard<std::mutex> lock(g_mu);
    ensure_model_locked();
    if (!g_model.loaded) fail(XC_ERR_UNSUPPORTED, "no cost model loaded");
    predict(g_model, f, ms);
}

int select_method(const xc_workload& w, xc_options& chosen) {
    float f[kFeatures];
    workload_features(w, f, kFeatures);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        dev = -1;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_tuned.find(std::make_tuple(w.family, w.rows, w.cols, w.n, w.m, dev));
    if (it != g_tuned.end()) {
        chosen = candidate(it->second);
        return XC_SELECT_TUNED;
    }
    if (auto_mode() == "autotune") return XC_SELECT_NONE;
    ensure_model_locked();
    if (!g_model.loaded || !in_domain(g_model, f)) return XC_SELECT_NONE;
    const auto wkey = std::make_tuple(w.family, w.rows, w.cols, w.n, w.m);
    auto pit = g_predicted.find(wkey);
    if (pit != g_predicted.end()) {
        chosen = candidate(pit->second);
        return XC_SELECT_MODEL;
    }
    std::vector<float> ms(g_model.cands.size());
    predict(g_model, f, ms.data());
    // A candidate is considered only up to twice the largest work it was timed on.
    uint32_t best = UINT32_MAX;
    for (uint32_t c = 0; c < ms.size(); ++c)
        if (f[5] <= g_model.c_max_macs[c] + 1.0 && (best == UINT32_MAX || ms[c] < ms[best])) best = c;
    if (best == UINT32_MAX) return XC_SELECT_NONE;
    if (g_predicted.size() > 4096) g_predicted.clear();
    g_predicted[wkey] = best;
    chosen = candidate(best);
    return XC_SELECT_MODEL;
}

float time_candidate_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t index,
                         uint32_t reps, hipStream_t s) {
    const xc_options o = candidate(index);
    reps = std::max<uint32_t>(reps, 1);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XC_HIP(hipEventCreate(&e0));
    XC_HIP(hipEventCreate(&e1));
    struct ev_guard {
        hipEvent_t a, b;
        ~ev_guard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    float t = INFINITY;
    try {
        // Warm-up (plans, kernel attributes, workspace), timed too: a slow candidate is not
        // repeated.
        XC_HIP(hipEventRecord(e0, s));
        run_family_f32(a, L, R, O, o, s);
        XC_HIP(hipEventRecord(e1, s));
        XC_HIP(hipEventSynchronize(e1));
        float warm = 0.f;
        XC_HIP(hipEventElapsedTime(&warm, e0, e1));
        const uint32_t r = warm > 250.f ? 1 : reps;
        XC_HIP(hipEventRecord(e0, s));
        for (uint32_t k = 0; k < r; ++k) run_family_f32(a, L, R, O, o, s);
        XC_HIP(hipEventRecord(e1, s));
        XC_HIP(hipEventSynchronize(e1));
        XC_HIP(hipEventElapsedTime(&t, e0, e1));
        t /= (float)r;
    } catch (const error& e) {
        if (e.status == XC_ERR_HIP) throw;  // a device failure is not a slow candidate
        t = INFINITY;
    }
    return t;
}

uint32_t autotune_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t reps, float* ms,
                      hipStream_t s) {
    const uint32_t n = num_candidates();
    uint32_t best = 0;
    float best_ms = INFINITY;
    // The FFT candidates first; beyond 1e13 MACs the definition-based ones (seconds to
    // minutes per run, 10-100x slower than the FFT there) are only timed when no FFT
    // candidate could run.
    const double macs = (double)a.pairs * a.rows * a.rows * a.cols * a.cols;
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < n; ++i)
        if (candidate(i).precision == XC_PREC_FFT64) order.push_back(i);
    const size_t n_fft = order.size();
    for (uint32_t i = 0; i < n; ++i)
        if (candidate(i).precision != XC_PREC_FFT64) order.push_back(i);
    for (size_t oi = 0; oi < order.size(); ++oi) {
        const uint32_t i = order[oi];
        float t = INFINITY;
        if (oi < n_fft || macs <= 1e13 || !std::isfinite(best_ms)) t = time_candidate_f32(a, L, R, O, i, reps, s);
        if (ms) ms[i] = t;
        if (t < best_ms) {
            best_ms = t;
            best = i;
        }
    }
    if (!std::isfinite(best_ms)) fail(XC_ERR_UNSUPPORTED, "no candidate method could run this workload");
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    g_tuned[std::make_tuple((uint32_t)a.family, a.rows, a.cols, a.n, a.m, dev)] = best;
    return best;
}

void autotune_clear() {
    std::lock_guard<std::mutex> lock(g_mu);
    g_tuned.clear();
}

xc_options auto_options_f32(const family_args& a, const float* L, const float* R, float* O, hipStream_t s) {
    const xc_workload w{(uint32_t)a.family, a.rows, a.cols, a.n, a.m};
    xc_options c{};
    if (select_method(w, c) != XC_SELECT_NONE) return c;
    if (auto_mode() == "model") {
        c = xc_options{};
        c.precision = XC_PREC_BF16X3;  // default tile
        return c;
    }
    return candidate(autotune_f32(a, L, R, O, 2, nullptr, s));
}

// Double data: the FP64 direct kernel (default tile) or the own FFT (FP64 throughout), timed
// once per workload and device on the call's buffers; the own FFT alone where the direct
// kernel would take seconds (above 1e12 multiply-adds).
xc_options auto_options_f64(const family_args& a, const double* L, const double* R, double* O, hipStream_t s) {
    static std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, int>, int> tuned;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    const auto key = std::make_tuple((uint32_t)a.family, a.rows, a.cols, a.n, a.m, dev);
    xc_options direct{}, fft{};
    direct.precision = XC_PREC_FP32;  // for double data: the FP64 direct kernel
    fft.precision = XC_PREC_FFT64;
    fft.tile = 1;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = tuned.find(key);
        if (it != tuned.end()) return it->second ? fft : direct;
    }
    const bool fft_ok = a.rows <= 2048 && a.cols <= 2048;
    const double macs = (double)a.pairs * a.rows * a.rows * a.cols * a.cols;
    int pick = fft_ok ? 1 : 0;
    if (fft_ok && macs <= 1e12) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        XC_HIP(hipEventCreate(&e0));
        XC_HIP(hipEventCreate(&e1));
        struct ev_guard {
            hipEvent_t a, b;
            ~ev_guard() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } eg{e0, e1};
        auto time = [&](bool use_fft) {
            float best = INFINITY;
            for (int rep = 0; rep < 2; ++rep) {  // the first run pays workspace and attribute set-up
                XC_HIP(hipEventRecord(e0, s));
                if (use_fft)
                    run_fftk_f64(a, L, R, O, 0, false, s);
                else
                    launch_direct_f64(a, L, R, O, direct, s);
                XC_HIP(hipEventRecord(e1, s));
                XC_HIP(hipEventSynchronize(e1));
                float t = 0.f;
                XC_HIP(hipEventElapsedTime(&t, e0, e1));
                best = std::min(best, t);
            }
            return best;
        };
        pick = time(true) <= time(false) ? 1 : 0;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    tuned[key] = pick;
    return pick ? fft : direct;
}

void cost_model_fit(const char* db_path, const char* model_path, const xc_train_options& o, uint32_t holdout_mod) {
    if (!db_path || !model_path) fail(XC_ERR_INVALID_ARGUMENT, "null path");
    const table t = read_table(db_path);
    const uint32_t nc = num_candidates();
    if (t.cands.size() != nc) fail(XC_ERR_UNSUPPORTED, "timing table has another candidate list");
    for (uint32_t i = 0; i < nc; ++i)
        if (t.cands[i] != candidate_name(i)) fail(XC_ERR_UNSUPPORTED, "timing table column " + t.cands[i] +
                                                                          " is not candidate " + candidate_name(i));
    std::vector<uint32_t> rows;
    for (uint32_t i = 0; i < t.w.size(); ++i)
        if (holdout_mod == 0 || i % holdout_mod != 0) rows.push_back(i);
    if (rows.size() < 4) fail(XC_ERR_INVALID_ARGUMENT, "too few training rows");
    const uint32_t B = (uint32_t)rows.size();

    model m;
    m.cands = t.cands;
    const uint32_t hidden = o.hidden ? o.hidden : 32;
    m.shape.layers = 3;
    m.shape.dims[0] = kFeatures;
    m.shape.dims[1] = hidden;
    m.shape.dims[2] = hidden;
    m.shape.dims[3] = nc;
    mlp_check_shape(&m.shape);

    std::vector<float> X((size_t)B * kFeatures), T((size_t)B * nc), M((size_t)B * nc);
    for (uint32_t r = 0; r < B; ++r) workload_features(t.w[rows[r]], &X[(size_t)r * kFeatures], kFeatures);
    m.f_mean.assign(kFeatures, 0.0);
    m.f_std.assign(kFeatures, 0.0);
    m.f_min.assign(kFeatures, INFINITY);
    m.f_max.assign(kFeatures, -INFINITY);
    for (uint32_t i = 0; i < kFeatures; ++i) {
        double s = 0, s2 = 0;
        for (uint32_t r = 0; r < B; ++r) {
            const double v = X[(size_t)r * kFeatures + i];
            s += v;
            s2 += v * v;
            m.f_min[i] = std::min(m.f_min[i], v);
            m.f_max[i] = std::max(m.f_max[i], v);
        }
        m.f_mean[i] = s / B;
        const double var = s2 / B - m.f_mean[i] * m.f_mean[i];
        m.f_std[i] = var > 1e-12 ? std::sqrt(var) : 1.0;
    }
    for (uint32_t r = 0; r < B; ++r)
        for (uint32_t i = 0; i < kFeatures; ++i) {
            float& v = X[(size_t)r * kFeatures + i];
            v = (float)((v - m.f_mean[i]) / m.f_std[i]);
        }
    m.y_mean.assign(nc, 0.0);
    m.y_std.assign(nc, 1.0);
    for (uint32_t c = 0; c < nc; ++c) {
        double s = 0, s2 = 0;
        uint32_t cnt = 0;
        for (uint32_t r = 0; r < B; ++r) {
            const double v = t.ms[rows[r]][c];
            if (std::isfinite(v) && v > 0) {
                const double y = std::log2(v);
                s += y;
                s2 += y * y;
                ++cnt;
            }
        }
        if (cnt) {
            m.y_mean[c] = s / cnt;
            const double var = s2 / cnt - m.y_mean[c] * m.y_mean[c];
            m.y_std[c] = var > 1e-12 ? std::sqrt(var) : 1.0;
        }
        for (uint32_t r = 0; r < B; ++r) {
            const double v = t.ms[rows[r]][c];
            const bool ok = std::isfinite(v) && v > 0;
            T[(size_t)r * nc + c] = ok ? (float)((std::log2(v) - m.y_mean[c]) / m.y_std[c]) : 0.f;
            M[(size_t)r * nc + c] = ok ? 1.f : 0.f;
        }
    }
    m.c_max_macs.assign(nc, -INFINITY);
    for (uint32_t r = 0; r < B; ++r) {
        float f[kFeatures];
        workload_features(t.w[rows[r]], f, kFeatures);
        for (uint32_t c = 0; c < nc; ++c)
            if (M[(size_t)r * nc + c] > 0.f) m.c_max_macs[c] = std::max(m.c_max_macs[c], (double)f[5]);
    }
    m.params.resize(mlp_num_params(m.shape));
    mlp_init(m.shape, o.seed, m.params.data());
    std::vector<float> loss(o.epochs);
    mlp_train(m.shape, X.data(), T.data(), M.data(), B, o, m.params.data(), loss.data());

    std::ofstream out(model_path);
    if (!out) fail(XC_ERR_INVALID_ARGUMENT, std::string("cannot write ") + model_path);
    out << "xcorr-cost-model 2\n";
    out << "candidates " << nc;
    for (const auto& c : m.cands) out << ' ' << c;
    out << '\n';
    out << "dims " << (m.shape.layers + 1);
    for (uint32_t l = 0; l <= m.shape.layers; ++l) out << ' ' << m.shape.dims[l];
    out << '\n';
    write_list(out, "feat_mean", m.f_mean);
    write_list(out, "feat_std", m.f_std);
    write_list(out, "feat_min", m.f_min);
    write_list(out, "feat_max", m.f_max);
    write_list(out, "out_mean", m.y_mean);
    write_list(out, "out_std", m.y_std);
    write_list(out, "cand_max_log2_macs", m.c_max_macs);
    write_list(out, "params", std::vector<double>(m.params.begin(), m.params.end()));
    char buf[160];
    std::snprintf(buf, sizeof(buf), "trained_rows %u holdout_mod %u epochs %u lr %.6g seed %u final_loss %.6g\n", B,
                  holdout_mod, o.epochs, (double)o.lr, o.seed, (double)loss.back());
    out << buf;
    if (!out) fail(XC_ERR_INVALID_ARGUMENT, std::string("write failed: ") + model_path);
}

}  // namespace xc
