// Kernel timing for measurement runs: HIP events on each launch's own stream, collected per
// kernel name -- recorded just before and after the launch (`timed` scopes), or attached to
// the kernel's own dispatch (timing_launch_events with hipExtLaunchKernel: the kernel's start
// and end, without the event packets' own processing time).  Off by default; not meant for
// launches under stream capture.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "xc_internal.hpp"

namespace xc {
namespace {

struct record {
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<record> g_done;
thread_local record t_open;

}  // namespace

bool timing_enabled() { return g_on.load(std::memory_order_relaxed); }

void timing_mark(const char* name, hipStream_t s, int phase) {
    hipEvent_t e = nullptr;
    XC_HIP(hipEventCreate(&e));
    XC_HIP(hipEventRecord(e, s));
    if (phase == 0) {
        t_open = record{name, e, nullptr};
        return;
    }
    t_open.b = e;
    std::lock_guard<std::mutex> lock(g_mu);
    g_done.push_back(t_open);
    t_open = record{};
}

void timing_launch_events(const char* name, hipEvent_t* start, hipEvent_t* stop) {
    record r{name, nullptr, nullptr};
    XC_HIP(hipEventCreate(&r.a));
    if (hipEventCreate(&r.b) != hipSuccess) {
        (void)hipEventDestroy(r.a);
        fail(XC_ERR_HIP, "event creation failed");
    }
    *start = r.a;
    *stop = r.b;
    std::lock_guard<std::mutex> lock(g_mu);
    g_done.push_back(r);
}

void timing_enable(bool on) { g_on.store(on); }

std::string timing_report() {
    std::vector<record> recs;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        recs.swap(g_done);
    }
    std::map<std::string, std::pair<uint64_t, double>> acc;
    for (record& r : recs) {
        float ms = 0.f;
        if (r.b) {
            XC_HIP(hipEventSynchronize(r.b));
            XC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        }
        auto& e = acc[r.name];
        e.first += 1;
        e.second += ms;
        (void)hipEventDestroy(r.a);
        if (r.b) (void)hipEventDestroy(r.b);
    }
    std::string out;
    char line[256];
    for (const auto& kv : acc) {
        std::snprintf(line, sizeof(line), "%s %llu %.6f\n", kv.first.c_str(), (unsigned long long)kv.second.first,
                      kv.second.second);
        out += line;
    }
    return out;
}

}  // namespace xc
