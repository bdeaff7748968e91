// Shared definitions of the HIP library: error plumbing, launch descriptors,
// family -> pair mapping, XCD-aware block remapping and kernel entry declarations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/xcorr.h"

namespace xc {

// Error thrown inside the library and converted to an xc_status at the C ABI.
struct error : std::runtime_error {
    xc_status status;
    error(xc_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(xc_status s, const std::string& msg) { throw error(s, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(XC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define XC_HIP(call) ::xc::hip_check((call), #call)

// Records the launch in the calling thread's descriptor (src/kernel_args.cpp:7-55 keeps
// the same facts for its failure messages).
void record_launch(const char* kernel, dim3 grid, dim3 block, uint32_t lds, uint32_t rows, uint32_t cols,
                   uint32_t pairs, uint32_t split);

// Kernel timing (xc_kernel_timing / xc_kernel_times): while enabled, a `timed` scope around
// a launch brackets it with two HIP events recorded on the launch's stream.
bool timing_enabled();
void timing_enable(bool on);
std::string timing_report();  // "name count total_ms" lines; clears the records
void timing_mark(const char* name, hipStream_t s, int phase);
// Two new events, registered under `name`, for the caller to attach to one launch
// (hipExtLaunchKernel's start and stop events).
void timing_launch_events(const char* name, hipEvent_t* start, hipEvent_t* stop);
struct timed {
    const char* name;
    hipStream_t s;
    bool on;
    timed(const char* n, hipStream_t st, bool allow = true) : name(n), s(st), on(allow && timing_enabled()) {
        if (on) timing_mark(name, s, 0);
    }
    ~timed() {
        if (on) timing_mark(name, s, 1);
    }
    timed(timed&& o) noexcept : name(o.name), s(o.s), on(o.on) { o.on = false; }
    timed(const timed&) = delete;
    timed& operator=(const timed&) = delete;
};

// Raises the dynamic-LDS limit of `kernel` to `bytes` on the current device, once per
// (kernel, device, size); thread-safe.
void allow_dynamic_lds(const void* kernel, uint32_t bytes);

// k-th result of a family -> (left index, right index); see the table in xcorr.h.
__host__ __device__ inline void pair_of(int family, uint32_t n, uint32_t m, uint32_t k, uint32_t& l,
                                        uint32_t& r) {
    switch (family) {
    case XC_ONE_TO_MANY: l = 0; r = k; break;
    case XC_N_TO_MN: l = k % n; r = k; break;
    case XC_N_TO_M: l = k / m; r = k % m; break;
    default: l = 0; r = 0; break;
    }
}

inline int64_t num_results(int family, uint32_t n, uint32_t m) {
    switch (family) {
    case XC_ONE_TO_ONE: return (n == 1 && m == 1) ? 1 : -1;
    case XC_ONE_TO_MANY: return (n == 1 && m >= 1) ? (int64_t)m : -1;
    case XC_N_TO_MN: return (n >= 1 && m >= 1 && m % n == 0) ? (int64_t)m : -1;
    case XC_N_TO_M: return (n >= 1 && m >= 1) ? (int64_t)n * m : -1;
    default: return -1;
    }
}

// The MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  Remap so that each XCD receives a contiguous range of logical
// block ids (neighbouring tiles of one pair share that XCD's L2).
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t XCD = 8;
    uint32_t per = nblocks / XCD;
    uint32_t full = per * XCD;
    if (bid >= full) return bid;  // tail blocks keep their id
    return (bid % XCD) * per + bid / XCD;
}

// The 16-byte quad at base + lane_bytes + uniform_bytes, where base and uniform_bytes are
// the same for every lane of the wave: a raw buffer load (the base in a scalar resource, the
// lane's offset in a VGPR, the uniform one in an SGPR), so the lanes need no address
// arithmetic.  The caller keeps both offsets below 2^31.
__device__ inline float4 buffer_quad(const float* base, uint32_t lane_bytes, uint32_t uniform_bytes) {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, 0xffffffff, 0x00020000);
    const i32x4 raw = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)lane_bytes, (int)uniform_bytes, 0);
    return __builtin_bit_cast(float4, raw);
}

// Block order for kernels whose blocks form a grid of units x bands (a unit: the work that
// shares one column range and one group of pairs; a band: one range of dy), dispatched
// round-robin over the 8 XCDs.  by_unit: all bands of a unit on one XCD, units dealt to the
// XCDs in snake order; else a band (all its units) per XCD, bands dealt in turn.  The grid
// has blocks_for_order() blocks; a block maps to (unit, band rank) or to nothing (false).
__host__ __device__ inline uint64_t blocks_for_order(uint32_t units, uint32_t bands, bool by_unit) {
    return by_unit ? ((uint64_t)units + 7) / 8 * 8 * bands : (uint64_t)units * ((bands + 7) / 8 * 8);
}
// The results are the same for every lane (bid = blockIdx.x); readfirstlane says so to the
// compiler, whose uniformity analysis otherwise loses them here and turns the callers' scalar
// loads into vector loads (measured: 8.9 -> 11.2 ms for the FP32 pairs kernel at 64 x 256^2).
__device__ inline bool block_order(uint32_t bid, uint32_t units, uint32_t bands, bool by_unit, uint32_t& unit,
                                   uint32_t& band) {
    const uint32_t xcd = bid % 8, pos = bid / 8;
    if (by_unit) {
        const uint32_t round = pos / bands;
        unit = round * 8 + ((round & 1) ? 7 - xcd : xcd);
        band = pos % bands;
    } else {
        band = pos / units * 8 + xcd;
        unit = pos % units;
    }
    unit = __builtin_amdgcn_readfirstlane(unit);
    band = __builtin_amdgcn_readfirstlane(band);
    return by_unit ? unit < units : band < bands;
}

// Rank k of count in nearest-the-centre-first order: c, c+1, c-1, c+2, ... (c = (count-1)/2).
__host__ __device__ inline int center_first(int k, int count) {
    const int c = (count - 1) / 2;
    return (k & 1) ? c + (k + 1) / 2 : c - k / 2;
}

// Description of one launch over a regular family.
struct family_args {
    int family;
    uint32_t rows, cols, n, m;
    uint32_t pairs;  // number of results
};

// Kernel launchers (defined in the .hip files).
// Direct kernel: opt's split, rows_per_block and register-tile fields (xcorr.h) pick the
// schedule; its precision and tile fields are not read.
void launch_direct_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                       hipStream_t s);
void launch_direct_f64(const family_args& a, const double* L, const double* R, double* O, const xc_options& opt,
                       hipStream_t s);
void launch_zero_f32(float* p, uint64_t count, hipStream_t s);
void launch_zero_f64(double* p, uint64_t count, hipStream_t s);

}  // namespace xc
