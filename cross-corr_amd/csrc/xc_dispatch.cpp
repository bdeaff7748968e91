// Method dispatch and device workspace management.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
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

}  // namespace

void* workspace(int slot, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    auto& v = g_ws[std::make_pair(dev, s)];
    if ((int)v.size() <= slot) v.resize(slot + 1);
    ws_slot& w = v[slot];
    if (w.bytes < bytes) {
        if (w.ptr) {
            // The old buffer may still be in use by queued work on this stream (and only
            // this stream uses it).
            XC_HIP(hipStreamSynchronize(s));
            XC_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        if (hipMalloc(&w.ptr, bytes) != hipSuccess) {
            (void)hipGetLastError();
            fail(XC_ERR_NO_MEMORY, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
        }
        w.bytes = bytes;
    }
    return w.ptr;
}

void run_family_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                    hipStream_t s) {
    switch (opt.precision) {
    case XC_PREC_AUTO:
        run_family_f32(a, L, R, O, auto_options_f32(a, L, R, O, s), s);
        return;
    case XC_PREC_FP32:
        launch_direct_f32(a, L, R, O, opt.split, s);
        return;
    case XC_PREC_BF16X3:
    case XC_PREC_FP16X3:
    case XC_PREC_BF16:
    case XC_PREC_FP16:
        launch_mfma_f32(a, L, R, O, (int)opt.precision, opt.border, opt.tile, s);
        return;
    case XC_PREC_FFT64:
        if (opt.tile == 1)
            run_fftk_f32(a, L, R, O, s);
        else if (opt.tile == 0)
            run_fft_f32(a, L, R, O, true, true, s);
        else
            fail(XC_ERR_INVALID_ARGUMENT, "FFT64 tile must be 0 (rocFFT) or 1 (own FFT kernels)");
        return;
    case XC_PREC_FFT32:
        run_fft_f32(a, L, R, O, true, false, s);
        return;
    default:
        fail(XC_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(opt.precision));
    }
}

}  // namespace xc
