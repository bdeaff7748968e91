// Method dispatch and device workspace management.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <utility>
#include <vector>

#include "xc_methods.hpp"
#include "xc_mlp.hpp"

namespace xc {

namespace {

struct ws_slot {
    void* ptr = nullptr;
    size_t bytes = 0;
};

std::mutex g_ws_mutex;
// Workspaces per (device, stream): work on one stream runs in order, so one set per stream
// is never used by two calls at once; calls on different streams (or host threads with
// their own streams) get their own sets.  Kept until the process ends.
std::map<std::pair<int, hipStream_t>, std::vector<ws_slot>> g_ws;
// Buffers replaced by a larger one.  They are never freed: another host thread may share
// the stream and hold the old pointer without having enqueued its kernels yet, so no
// synchronisation point here covers its work.  Each replacement at least doubles the
// slot, so what is retired stays below the slot's final size.
std::vector<void*> g_ws_retired;

}  // namespace

void* workspace(int slot, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    auto& v = g_ws[std::make_pair(dev, s)];
    if ((int)v.size() <= slot) v.resize(slot + 1);
    ws_slot& w = v[slot];
    if (w.bytes < bytes) {
        size_t size = w.ptr ? std::max(bytes, 2 * w.bytes) : bytes;
        void* p = nullptr;
        if (hipMalloc(&p, size) != hipSuccess) {
            (void)hipGetLastError();
            size = bytes;
            if (hipMalloc(&p, size) != hipSuccess) {
                (void)hipGetLastError();
                fail(XC_ERR_NO_MEMORY, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
            }
        }
        if (w.ptr) g_ws_retired.push_back(w.ptr);
        w.ptr = p;
        w.bytes = size;
    }
    return w.ptr;
}

void allow_dynamic_lds(const void* kernel, uint32_t bytes) {
    static std::mutex mu;
    static std::set<std::tuple<const void*, int, uint32_t>> done;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (done.count(std::make_tuple(kernel, dev, bytes))) return;
    XC_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done.emplace(kernel, dev, bytes);
}

void run_family_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                    hipStream_t s) {
    switch (opt.precision) {
    case XC_PREC_AUTO:
        run_family_f32(a, L, R, O, auto_options_f32(a, L, R, O, s), s);
        return;
    case XC_PREC_FP32:
        launch_direct_f32(a, L, R, O, opt, s);
        return;
    case XC_PREC_BF16X3:
    case XC_PREC_FP16X3:
    case XC_PREC_BF16:
    case XC_PREC_FP16:
        launch_mfma_f32(a, L, R, O, (int)opt.precision, opt.border, opt.tile, s);
        return;
    case XC_PREC_FFT64:
        if (opt.tile == 1 || opt.tile == 2)
            run_fftk_f32(a, L, R, O, opt.split, opt.tile == 2, s);
        else if (opt.tile == 0)
            run_fft_f32(a, L, R, O, true, true, s);
        else
            fail(XC_ERR_INVALID_ARGUMENT,
                 "FFT64 tile must be 0 (rocFFT), 1 (own FFT kernels) or 2 (own FFT, power-of-two rows)");
        return;
    case XC_PREC_FFT32:
        run_fft_f32(a, L, R, O, true, false, s);
        return;
    default:
        fail(XC_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(opt.precision));
    }
}

}  // namespace xc
