// Uneven batch: pairs of different sizes in one launch of a captured HIP graph, replayed.  The
// graph holds either one persistent MFMA kernel over all pairs (bf16x3 / bf16 / fp16) or, for
// FFT64 tile 1, the own-FFT kernels of every size class on parallel branches.  No counterpart in the reference, whose n_to_mn wrapper launches one
// kernel per left matrix on round-robin streams (src/n_to_mn.hpp:553-571).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "xc_methods.hpp"

struct xc_batch_plan {
    int device = 0;
    int kind = XC_PREC_BF16X3;
    uint32_t thresh = 0;
    uint32_t n_items = 0;
    uint32_t grid = 0, lds = 0;
    xc::batch_item* d_items = nullptr;
    xc::batch_pair_desc* d_pairs = nullptr;
    uint32_t* d_tables = nullptr;
    uint32_t* d_counters = nullptr;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t node = nullptr;
    const float* L = nullptr;
    const float* R = nullptr;
    float* O = nullptr;
    std::vector<xc_pair> pairs;

    // FFT plans: one own-FFT job per size class, classes dealt to parallel graph branches
    // (each with its own stream and workspace).
    bool fft = false;
    std::vector<xc::fftk_job> jobs;
    std::vector<uint32_t> job_branch;
    std::vector<hipStream_t> branch_streams;
    std::vector<xc::fftk_ws> branch_ws;
    std::vector<void*> allocations;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;

    ~xc_batch_plan() {
        // work launched from this plan may still be queued: let its device drain first
        int cur = 0;
        const bool other = hipGetDevice(&cur) == hipSuccess && cur != device && hipSetDevice(device) == hipSuccess;
        (void)hipDeviceSynchronize();
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (hipStream_t b : branch_streams) (void)hipStreamDestroy(b);
        if (fork) (void)hipEventDestroy(fork);
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        for (void* p : {(void*)d_items, (void*)d_pairs, (void*)d_tables, (void*)d_counters})
            if (p) (void)hipFree(p);
        for (void* p : allocations) (void)hipFree(p);
        if (other) (void)hipSetDevice(cur);
    }
};

namespace xc {

namespace {

void check_pairs(const xc_pair* pairs, uint32_t n) {
    if (n && !pairs) fail(XC_ERR_INVALID_ARGUMENT, "null pair array");
    for (uint32_t i = 0; i < n; ++i) {
        if (pairs[i].rows == 0 || pairs[i].cols == 0)
            fail(XC_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + " has an empty matrix");
        if (pairs[i].rows > 32768 || pairs[i].cols > 32768)
            fail(XC_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + " dimension above 32768");
    }
}

// Work estimate of one block tile: left-row steps times columns met (times 64 dy rows).
double tile_cost(uint32_t H, uint32_t W, uint32_t tx, uint32_t ty) {
    const int dyb = (int)ty * 64 - ((int)H - 1), dxb = (int)tx * 256 - ((int)W - 1);
    const int dy_hi = std::min(dyb + 63, (int)H - 1), dx_hi = std::min(dxb + 255, (int)W - 1);
    const int rsteps = std::min((int)H, (int)H + dy_hi) - std::max(0, dyb);
    const int cols = std::min((int)W, (int)W - dxb) - std::max(0, -dx_hi);
    return (double)std::max(rsteps, 0) * std::max(cols, 0) * 64.0;
}

}  // namespace

namespace {

void* plan_alloc(xc_batch_plan* plan, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
        (void)hipGetLastError();
        fail(XC_ERR_NO_MEMORY, "batch plan allocation failed");
    }
    plan->allocations.push_back(p);
    return p;
}

// Own-FFT plan: pairs grouped into classes of equal transform lengths (Pr x Pc, fftk_col_len and
// fftk_row_len), one job per class whose matrices keep their own sizes (the
// job's per-matrix size table); per class the distinct left and right matrices (by offset and
// size) are transformed once and every pair of the class is one result of the job.  Classes
// are dealt to `branches` (1..8) parallel graph branches.
xc_batch_plan* fft_plan_create(const xc_pair* pairs, uint32_t num_pairs, uint32_t branches) {
    auto plan = std::make_unique<xc_batch_plan>();
    plan->fft = true;
    plan->kind = XC_PREC_FFT64;
    plan->n_items = num_pairs;
    plan->pairs.assign(pairs, pairs + num_pairs);
    XC_HIP(hipGetDevice(&plan->device));
    fftk_prepare();
    using mkey = std::tuple<uint64_t, uint32_t, uint32_t>;  // offset, rows, cols
    struct cls {
        std::vector<uint64_t> loffs, roffs, ooffs;
        std::vector<uint32_t> ldims, rdims, ptab;
        std::map<mkey, uint32_t> lidx, ridx;
        uint32_t H = 0, W = 0;
        double work = 0.0;
    };
    std::map<std::pair<uint32_t, uint32_t>, cls> classes;
    for (uint32_t i = 0; i < num_pairs; ++i) {
        const xc_pair& p = pairs[i];
        if (2 * p.rows - 1 > 4096 || 2 * p.cols - 1 > 4096)
            fail(XC_ERR_UNSUPPORTED, "pair " + std::to_string(i) + ": the FFT batch plan takes rows, cols <= 2048");
        cls& c = classes[{fftk_col_len(p.rows), fftk_row_len(p.cols)}];
        auto li = c.lidx.emplace(mkey{p.left_off, p.rows, p.cols}, (uint32_t)c.loffs.size());
        if (li.second) {
            c.loffs.push_back(p.left_off);
            c.ldims.insert(c.ldims.end(), {p.rows, p.cols});
            c.work += (double)p.rows * p.cols;
        }
        auto ri = c.ridx.emplace(mkey{p.right_off, p.rows, p.cols}, (uint32_t)c.roffs.size());
        if (ri.second) {
            c.roffs.push_back(p.right_off);
            c.rdims.insert(c.rdims.end(), {p.rows, p.cols});
            c.work += (double)p.rows * p.cols;
        }
        c.ptab.push_back(li.first->second);
        c.ptab.push_back(ri.first->second);
        c.ooffs.push_back(p.out_off);
        c.H = std::max(c.H, p.rows);
        c.W = std::max(c.W, p.cols);
        c.work += 2.0 * p.rows * p.cols;
    }
    // One device buffer for every class's tables.
    size_t words64 = 0, words32 = 0;
    for (auto& kv : classes) {
        words64 += kv.second.loffs.size() + kv.second.roffs.size() + kv.second.ooffs.size();
        words32 += kv.second.ptab.size() + kv.second.ldims.size() + kv.second.rdims.size();
    }
    std::vector<uint64_t> t64;
    std::vector<uint32_t> t32;
    t64.reserve(words64);
    t32.reserve(words32);
    struct span {
        size_t moff, ooff, ptab, dims;
    };
    std::vector<span> spans;
    std::vector<double> cost;
    for (auto& kv : classes) {
        cls& c = kv.second;
        span sp{t64.size(), 0, t32.size(), 0};
        t64.insert(t64.end(), c.loffs.begin(), c.loffs.end());
        t64.insert(t64.end(), c.roffs.begin(), c.roffs.end());
        sp.ooff = t64.size();
        t64.insert(t64.end(), c.ooffs.begin(), c.ooffs.end());
        t32.insert(t32.end(), c.ptab.begin(), c.ptab.end());
        sp.dims = t32.size();  // lefts then rights, (rows, cols) each
        t32.insert(t32.end(), c.ldims.begin(), c.ldims.end());
        t32.insert(t32.end(), c.rdims.begin(), c.rdims.end());
        spans.push_back(sp);
        fftk_job j{};
        j.H = c.H;
        j.W = c.W;
        j.nl = (uint32_t)c.loffs.size();
        j.nr = (uint32_t)c.roffs.size();
        j.pairs = (uint32_t)c.ooffs.size();
        plan->jobs.push_back(j);
        // work estimate: FFT points transformed (padded lengths are proportional)
        cost.push_back(c.work);
    }
    uint64_t* d64 = static_cast<uint64_t*>(plan_alloc(plan.get(), t64.size() * sizeof(uint64_t)));
    uint32_t* d32 = static_cast<uint32_t*>(plan_alloc(plan.get(), t32.size() * sizeof(uint32_t)));
    XC_HIP(hipMemcpy(d64, t64.data(), t64.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    XC_HIP(hipMemcpy(d32, t32.data(), t32.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (size_t c = 0; c < plan->jobs.size(); ++c) {
        plan->jobs[c].moff = d64 + spans[c].moff;
        plan->jobs[c].ooff = d64 + spans[c].ooff;
        plan->jobs[c].ptab = d32 + spans[c].ptab;
        plan->jobs[c].mdims = d32 + spans[c].dims;
    }
    // Classes, largest first, to the least loaded branch.
    const uint32_t nb = (uint32_t)std::max<size_t>(1, std::min<size_t>(branches, plan->jobs.size()));
    std::vector<size_t> order(plan->jobs.size());
    for (size_t c = 0; c < order.size(); ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<double> load(nb, 0.0);
    plan->job_branch.assign(plan->jobs.size(), 0);
    std::vector<fftk_ws_sizes> need(nb, fftk_ws_sizes{16, 16, 16, 16, 1});
    for (size_t c : order) {
        const uint32_t b = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        plan->job_branch[c] = b;
        load[b] += cost[c];
        const fftk_ws_sizes z = fftk_sizes(plan->jobs[c]);
        need[b].sr = std::max(need[b].sr, z.sr);
        need[b].sl = std::max(need[b].sl, z.sl);
        need[b].lh = std::max(need[b].lh, z.lh);
        need[b].z = std::max(need[b].z, z.z);
    }
    for (uint32_t b = 0; b < nb; ++b) {
        fftk_ws w{};
        w.sr = plan_alloc(plan.get(), need[b].sr);
        w.sl = plan_alloc(plan.get(), need[b].sl);
        w.lh = plan_alloc(plan.get(), need[b].lh);
        w.z = plan_alloc(plan.get(), need[b].z);
        w.zchunk = 0;  // per job below
        plan->branch_ws.push_back(w);
        hipStream_t st = nullptr;
        XC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        plan->branch_streams.push_back(st);
        hipEvent_t e = nullptr;
        XC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        plan->joins.push_back(e);
    }
    XC_HIP(hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming));
    XC_HIP(hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking));
    return plan.release();
}

// Enqueues every class of an FFT plan: on `s` alone, or forked over the branch streams.
void fft_plan_enqueue(xc_batch_plan* plan, const float* L, const float* R, float* O, hipStream_t s, bool branches) {
    if (branches) XC_HIP(hipEventRecord(plan->fork, s));
    std::vector<bool> used(plan->branch_streams.size(), false);
    for (size_t c = 0; c < plan->jobs.size(); ++c) {
        const uint32_t b = plan->job_branch[c];
        hipStream_t st = branches ? plan->branch_streams[b] : s;
        if (branches && !used[b]) {
            XC_HIP(hipStreamWaitEvent(st, plan->fork, 0));
            used[b] = true;
        }
        fftk_ws w = plan->branch_ws[b];
        w.zchunk = fftk_sizes(plan->jobs[c]).chunk;
        fftk_launch_f32(plan->jobs[c], L, R, O, w, st);
    }
    if (branches)
        for (size_t b = 0; b < used.size(); ++b)
            if (used[b]) {
                XC_HIP(hipEventRecord(plan->joins[b], plan->branch_streams[b]));
                XC_HIP(hipStreamWaitEvent(s, plan->joins[b], 0));
            }
}

}  // namespace

xc_batch_plan* batch_plan_create(const xc_pair* pairs, uint32_t num_pairs, const xc_options& opt) {
    check_pairs(pairs, num_pairs);
    if (opt.precision == XC_PREC_FFT64) {
        if (opt.tile != 1) fail(XC_ERR_INVALID_ARGUMENT, "FFT batch plans run the own FFT kernels (tile 1)");
        if (opt.split > 8) fail(XC_ERR_INVALID_ARGUMENT, "FFT batch plans take 1..8 branches (split; 0 = 8)");
        return fft_plan_create(pairs, num_pairs, opt.split ? opt.split : 8);
    }
    const int kind = opt.precision == XC_PREC_AUTO ? XC_PREC_BF16X3 : (int)opt.precision;
    const bool x3 = kind == XC_PREC_BF16X3 || kind == XC_PREC_FP16X3;
    if (!x3 && kind != XC_PREC_BF16 && kind != XC_PREC_FP16)
        fail(XC_ERR_INVALID_ARGUMENT, "batch plans run the MFMA methods (bf16x3, fp16x3, bf16, fp16) or FFT64 tile 1");
    if (opt.border > 512) fail(XC_ERR_INVALID_ARGUMENT, "border threshold above 512");
    const uint32_t thresh = x3 ? (opt.border ? opt.border : batch_default_border(kind)) : 0;

    std::vector<batch_pair_desc> descs(num_pairs);
    std::vector<uint32_t> tables;
    std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> table_of;  // size -> (offset, nh)
    struct costed {
        double cost;
        batch_item item;
    };
    std::vector<costed> items;
    for (uint32_t i = 0; i < num_pairs; ++i) {
        const xc_pair& p = pairs[i];
        batch_pair_desc& d = descs[i];
        d.loff = p.left_off;
        d.roff = p.right_off;
        d.ooff = p.out_off;
        d.rows = p.rows;
        d.cols = p.cols;
        d.table_off = 0;
        d.nh = 0;
        const uint32_t tx_n = batch_tiles_x(p.cols), ty_n = batch_tiles_y(p.rows);
        for (uint32_t ty = 0; ty < ty_n; ++ty)
            for (uint32_t tx = 0; tx < tx_n; ++tx)
                items.push_back({tile_cost(p.rows, p.cols, tx, ty), batch_item{i, 0, (uint16_t)tx, ty}});
        if (thresh > 1) {
            auto key = std::make_pair(p.rows, p.cols);
            auto it = table_of.find(key);
            if (it == table_of.end()) {
                std::vector<uint32_t> start;
                const uint32_t nh = frame_table_host(p.rows, p.cols, thresh, start);
                it = table_of.emplace(key, std::make_pair((uint32_t)tables.size(), nh)).first;
                tables.insert(tables.end(), start.begin(), start.end());
            }
            d.table_off = it->second.first;
            d.nh = it->second.second;
            if (d.nh) {
                const uint32_t total = tables[d.table_off + d.nh] * 4;
                for (uint32_t f0 = 0; f0 < total; f0 += batch_frame_chunk())
                    items.push_back({(double)batch_frame_chunk() * thresh / 4, batch_item{i, 1, 0, f0}});
            }
        }
    }
    // Largest work first: the dynamic queue then balances the tail with small items.
    std::stable_sort(items.begin(), items.end(), [](const costed& a, const costed& b) { return a.cost > b.cost; });

    auto plan = std::make_unique<xc_batch_plan>();
    plan->kind = kind;
    plan->thresh = thresh;
    plan->n_items = (uint32_t)items.size();
    plan->pairs.assign(pairs, pairs + num_pairs);
    XC_HIP(hipGetDevice(&plan->device));
    std::vector<batch_item> flat(items.size());
    for (size_t i = 0; i < items.size(); ++i) flat[i] = items[i].item;
    if (tables.empty()) tables.push_back(0);
    auto upload = [](void** dst, const void* src, size_t bytes) {
        if (hipMalloc(dst, std::max<size_t>(bytes, 16)) != hipSuccess) {
            (void)hipGetLastError();
            fail(XC_ERR_NO_MEMORY, "batch plan allocation failed");
        }
        if (bytes) XC_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    };
    upload((void**)&plan->d_items, flat.data(), flat.size() * sizeof(batch_item));
    upload((void**)&plan->d_pairs, descs.data(), descs.size() * sizeof(batch_pair_desc));
    upload((void**)&plan->d_tables, tables.data(), tables.size() * sizeof(uint32_t));
    const uint32_t zero[2] = {0, 0};
    upload((void**)&plan->d_counters, zero, sizeof(zero));
    int cus = 0;
    XC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, plan->device));
    plan->grid = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)cus * batch_blocks_per_cu(), plan->n_items));
    plan->lds = batch_lds_bytes(kind);
    allow_dynamic_lds(batch_kernel_ptr(kind), plan->lds);
    XC_HIP(hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking));
    return plan.release();
}

void batch_plan_exec(xc_batch_plan* plan, const float* L, const float* R, float* O, hipStream_t s, bool use_graph) {
    if (!plan) fail(XC_ERR_INVALID_ARGUMENT, "null batch plan");
    if (plan->n_items == 0) return;
    if (!L || !R || !O) fail(XC_ERR_INVALID_ARGUMENT, "null matrix pointer");
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    if (dev != plan->device) fail(XC_ERR_INVALID_ARGUMENT, "batch plan used on another device than it was made on");
    if (plan->fft) {
        if (!use_graph) {
            fft_plan_enqueue(plan, L, R, O, s, false);
            return;
        }
        if (plan->exec && (plan->L != L || plan->R != R || plan->O != O)) {
            // the graph bakes the matrix pointers into every node: capture it again
            XC_HIP(hipGraphExecDestroy(plan->exec));
            XC_HIP(hipGraphDestroy(plan->graph));
            plan->exec = nullptr;
            plan->graph = nullptr;
        }
        if (!plan->exec) {
            XC_HIP(hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal));
            try {
                fft_plan_enqueue(plan, L, R, O, plan->capture_stream, true);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(plan->capture_stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipGraph_t g = nullptr;
            XC_HIP(hipStreamEndCapture(plan->capture_stream, &g));
            plan->graph = g;
            XC_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
            plan->L = L;
            plan->R = R;
            plan->O = O;
        }
        record_launch("fft batch graph", dim3((uint32_t)plan->jobs.size()), dim3(256), 0, 0, 0,
                      (uint32_t)plan->pairs.size(), (uint32_t)plan->branch_streams.size());
        timed tg("batch_graph", s);
        XC_HIP(hipGraphLaunch(plan->exec, s));
        return;
    }
    uint32_t n_items = plan->n_items, thresh = plan->thresh;
    void* args[] = {&plan->d_items, &n_items, &plan->d_pairs, &plan->d_tables, &L, &R, &O, &plan->d_counters, &thresh};
    record_launch("batch_kernel (persistent)", dim3(plan->grid), dim3(256), plan->lds, 0, 0,
                  (uint32_t)plan->pairs.size(), 1);
    if (!use_graph) {
        timed tb("batch_kernel", s);
        XC_HIP(hipLaunchKernel(batch_kernel_ptr(plan->kind), dim3(plan->grid), dim3(256), args, plan->lds, s));
        return;
    }
    if (!plan->exec) {
        XC_HIP(hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = hipLaunchKernel(batch_kernel_ptr(plan->kind), dim3(plan->grid), dim3(256), args,
                                              plan->lds, plan->capture_stream);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(plan->capture_stream, &g);
        XC_HIP(le);
        XC_HIP(ce);
        plan->graph = g;
        size_t count = 1;
        XC_HIP(hipGraphGetNodes(plan->graph, &plan->node, &count));
        if (count != 1) fail(XC_ERR_HIP, "unexpected batch graph layout");
        XC_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
        plan->L = L;
        plan->R = R;
        plan->O = O;
    } else if (plan->L != L || plan->R != R || plan->O != O) {
        hipKernelNodeParams p{};
        XC_HIP(hipGraphKernelNodeGetParams(plan->node, &p));
        p.kernelParams = args;
        XC_HIP(hipGraphExecKernelNodeSetParams(plan->exec, plan->node, &p));
        plan->L = L;
        plan->R = R;
        plan->O = O;
    }
    timed tg("batch_graph", s);
    XC_HIP(hipGraphLaunch(plan->exec, s));
}

void batch_plan_destroy(xc_batch_plan* plan) { delete plan; }

namespace {

std::string pairs_key(const xc_pair* pairs, uint32_t num_pairs) {
    std::string key(reinterpret_cast<const char*>(pairs), sizeof(xc_pair) * num_pairs);
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    key.append(reinterpret_cast<const char*>(&dev), sizeof(dev));
    return key;
}

std::mutex g_tuned_mu;
std::map<std::string, xc_options> g_tuned_batches;  // pair list + device -> fastest schedule

}  // namespace

// Schedules the batch autotuner times: the own-FFT plan over 1, 2, 4 and 8 parallel branches
// (when every size fits it, rows and cols <= 2048) and the persistent bf16x3 MFMA plan.  The
// MFMA plan is left out when the FFT plan fits and the definition's work exceeds 1e13 MACs
// (one execution would take about a second; the own FFT wins such batches by 10-100x).
std::vector<xc_options> batch_candidates(const xc_pair* pairs, uint32_t num_pairs) {
    bool fits = true;
    double macs = 0.0;
    for (uint32_t i = 0; i < num_pairs; ++i) {
        fits = fits && pairs[i].rows <= 2048 && pairs[i].cols <= 2048;
        macs += (double)pairs[i].rows * pairs[i].rows * pairs[i].cols * pairs[i].cols;
    }
    std::vector<xc_options> c;
    if (fits)
        for (uint32_t b : {1u, 2u, 4u, 8u}) {
            xc_options o{};
            o.precision = XC_PREC_FFT64;
            o.tile = 1;
            o.split = b;
            c.push_back(o);
        }
    if (!fits || macs <= 1e13) {
        xc_options m{};
        m.precision = XC_PREC_BF16X3;
        c.push_back(m);
    }
    return c;
}

// Times every candidate schedule on the caller's buffers (plan build and graph capture
// untimed, then `reps` graph replays) and remembers the fastest for this pair list; `out` is
// overwritten.  Returns the index of the fastest candidate.
uint32_t batch_autotune_f32(const xc_pair* pairs, uint32_t num_pairs, const float* L, const float* R, float* O,
                            hipStream_t s, std::vector<xc_options>& cands, std::vector<float>& ms) {
    check_pairs(pairs, num_pairs);
    if (num_pairs == 0) fail(XC_ERR_INVALID_ARGUMENT, "empty batch");
    cands = batch_candidates(pairs, num_pairs);
    ms.assign(cands.size(), INFINITY);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XC_HIP(hipEventCreate(&e0));
    XC_HIP(hipEventCreate(&e1));
    struct ev_guard {
        hipEvent_t a, b;
        ~ev_guard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};
    const uint32_t reps = 3;
    for (size_t c = 0; c < cands.size(); ++c) {
        std::unique_ptr<xc_batch_plan, void (*)(xc_batch_plan*)> plan(batch_plan_create(pairs, num_pairs, cands[c]),
                                                                      batch_plan_destroy);
        batch_plan_exec(plan.get(), L, R, O, s, true);  // capture + warm-up
        XC_HIP(hipEventRecord(e0, s));
        for (uint32_t r = 0; r < reps; ++r) batch_plan_exec(plan.get(), L, R, O, s, true);
        XC_HIP(hipEventRecord(e1, s));
        XC_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        XC_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[c] = t / reps;
    }
    const uint32_t best = (uint32_t)(std::min_element(ms.begin(), ms.end()) - ms.begin());
    std::lock_guard<std::mutex> lock(g_tuned_mu);
    if (g_tuned_batches.size() > 64) g_tuned_batches.clear();
    g_tuned_batches[pairs_key(pairs, num_pairs)] = cands[best];
    return best;
}

// One-shot batch: plans are cached by their pair list and options (a few per process).
// AUTO takes the schedule the batch autotuner measured fastest for this pair list (running
// it on the call's buffers the first time), or with XC_AUTO=model the own-FFT plan when
// every size fits it, else bf16x3.
void run_batch_f32(const xc_pair* pairs, uint32_t num_pairs, const float* L, const float* R, float* O,
                   const xc_options& opt, hipStream_t s) {
    if (opt.precision == XC_PREC_FP32) {
        // Exact FP32 products: one direct-kernel launch per pair.
        for (uint32_t i = 0; i < num_pairs; ++i) {
            family_args a{XC_ONE_TO_ONE, pairs[i].rows, pairs[i].cols, 1, 1, 1};
            launch_direct_f32(a, L + pairs[i].left_off, R + pairs[i].right_off, O + pairs[i].out_off, opt, s);
        }
        return;
    }
    xc_options o = opt;
    if (o.precision == XC_PREC_AUTO) {
        bool have = false;
        {
            std::lock_guard<std::mutex> lock(g_tuned_mu);
            auto it = g_tuned_batches.find(pairs_key(pairs, num_pairs));
            if (it != g_tuned_batches.end()) {
                o = it->second;
                have = true;
            }
        }
        if (!have) {
            const char* mode = std::getenv("XC_AUTO");
            if (mode && std::string(mode) == "model") {
                // the first candidate is the own-FFT plan when every size fits it, else bf16x3
                o = batch_candidates(pairs, num_pairs).front();
                if (o.precision == XC_PREC_FFT64) o.split = 8;
            } else {
                std::vector<xc_options> cands;
                std::vector<float> ms;
                o = cands[batch_autotune_f32(pairs, num_pairs, L, R, O, s, cands, ms)];
            }
        }
    }
    static std::mutex mu;
    static std::list<std::pair<std::string, std::unique_ptr<xc_batch_plan>>> cache;
    // per stream too: a plan's workspaces and work counter serve one execution at a time
    std::string key(reinterpret_cast<const char*>(&o), sizeof(o));
    key.append(pairs_key(pairs, num_pairs));
    key.append(reinterpret_cast<const char*>(&s), sizeof(s));
    std::lock_guard<std::mutex> lock(mu);
    xc_batch_plan* plan = nullptr;
    for (auto it = cache.begin(); it != cache.end(); ++it)
        if (it->first == key) {
            cache.splice(cache.begin(), cache, it);
            plan = cache.front().second.get();
            break;
        }
    if (!plan) {
        cache.emplace_front(key, std::unique_ptr<xc_batch_plan>(batch_plan_create(pairs, num_pairs, o)));
        plan = cache.front().second.get();
        if (cache.size() > 8) cache.pop_back();
    }
    batch_plan_exec(plan, L, R, O, s, true);
}

}  // namespace xc
