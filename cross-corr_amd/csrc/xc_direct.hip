// Wave64 definition-based kernel ("direct"), FP32 and FP64.
//
// Replaces the warp-shuffle family of the reference (src/naive_shuffle.cu:86-267 and its
// multirow / multimat / work-distribution variants, src/naive_shuffle_*.cu): the same
// per-shift sums of src/validate.hpp:94-132, organised for CDNA4 instead of 32-lane
// shuffles:
//   * a wave is 64 consecutive dx of TY consecutive dy; every lane of the wave walks the
//     same (right row j, left column c) sequence, so the left value L[j+dy][c] is
//     wave-uniform and is read with scalar loads into SGPRs (the FMA takes it as its
//     SGPR operand) -- the reference broadcasts it with a shuffle per FMA;
//   * the right row j is staged once per block into LDS with zero padding on both
//     sides, so every lane reads its sliding window R[j][c-dx..] without bounds checks
//     (the reference's bound_checked_loads.cuh:8-53 become padding);
//   * each thread keeps TY outputs; products of one row pair are summed in a fresh
//     accumulator that is added to the total after the row with a compensated (Kahan)
//     add -- a single running FP32 sum, or a plain two-level one, exceeds 1e-6 of
//     sum|a||b| at 256x256 for positive data;
//   * the reduction over right rows can be split over `split` blocks that add their
//     partial sums atomically into a zeroed output ("work distribution",
//     src/naive_shuffle.cu:306-413, src/row_distribution.cuh:34-284, here a plain
//     rectangle split of the overlap rows).
#include <algorithm>

#include "xc_internal.hpp"

namespace xc {
namespace {

constexpr int kWaves = 4;            // waves per block
constexpr int kBX = 64 * kWaves;     // dx covered by one block
constexpr int kPrefetch = 8;         // staged elements kept in registers per thread

template <typename T>
__device__ inline void atomic_add(T* p, T v) {
    atomicAdd(p, v);
}

// Sums, for TY left rows r0..r0+TY-1 against the staged right row `cur`, the products of
// columns [c_lo, c_hi) into acc[y].  Left values are wave-uniform (scalar loads).  A row
// outside [0, H) reads a clamped row and its sum is dropped afterwards.
template <typename T, int TY, int CW>
__device__ __forceinline__ void row_pair(T (&acc)[TY], const T* __restrict__ Lp, int r0, int H, int W,
                                         const T* cur, int P, int c_lo, int c_hi, int c_full) {
    const T* lrow[TY];
#pragma unroll
    for (int y = 0; y < TY; ++y) {
        lrow[y] = Lp + (ptrdiff_t)min(max(r0 + y, 0), H - 1) * W;
        acc[y] = T(0);
    }
    int c0 = c_lo;
    for (; c0 < c_hi && c0 < c_full; c0 += CW, P += CW) {
        T rr[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) rr[t] = cur[P + t];
        T lv[TY][CW];
#pragma unroll
        for (int y = 0; y < TY; ++y)
#pragma unroll
            for (int t = 0; t < CW; ++t) lv[y][t] = lrow[y][c0 + t];
#pragma unroll
        for (int y = 0; y < TY; ++y)
#pragma unroll
            for (int t = 0; t < CW; ++t) acc[y] = fma(lv[y][t], rr[t], acc[y]);
    }
    if (c0 < c_hi) {  // last partial chunk of the row: columns >= W are absent
        T rr[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) rr[t] = cur[P + t];
#pragma unroll
        for (int y = 0; y < TY; ++y)
#pragma unroll
            for (int t = 0; t < CW; ++t) {
                const T l = (c0 + t < W) ? lrow[y][min(c0 + t, W - 1)] : T(0);
                acc[y] = fma(l, rr[t], acc[y]);
            }
    }
#pragma unroll
    for (int y = 0; y < TY; ++y)
        if ((unsigned)(r0 + y) >= (unsigned)H) acc[y] = T(0);
}

// One block: dx in [dxb, dxb + kBX), dy in [dy0, dy0 + TY), right rows [j_begin, j_end).
template <typename T, int TY, int CW>
__global__ __launch_bounds__(256) void direct_kernel(const T* __restrict__ L, const T* __restrict__ R,
                                                     T* __restrict__ O, int family, uint32_t rows,
                                                     uint32_t cols, uint32_t n, uint32_t m, uint32_t split,
                                                     uint32_t tiles_x, uint32_t tiles_y, int len,
                                                     int atomic) {
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    T* buf = reinterpret_cast<T*>(smem_raw);  // 2 x len

    const int H = (int)rows, W = (int)cols;
    const int SX = 2 * W - 1;
    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = bid % tiles_x;
    bid /= tiles_x;
    const uint32_t ty = bid % tiles_y;
    bid /= tiles_y;
    const uint32_t ks = bid % split;
    const uint32_t k = bid / split;
    uint32_t li, ri;
    pair_of(family, n, m, k, li, ri);
    const T* __restrict__ Lp = L + (size_t)li * H * W;
    const T* __restrict__ Rp = R + (size_t)ri * H * W;
    T* __restrict__ Op = O + (size_t)k * SX * (2 * H - 1);

    const int dy0 = (int)ty * TY - (H - 1);
    const int dxb = (int)tx * kBX - (W - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
    const int lane = threadIdx.x & 63;
    const int dxw = dxb + wave * 64;
    const int dx = dxw + lane;

    // Right rows that meet at least one dy of the tile: j in [max(0,-dy), min(H,H-dy)).
    const int dy_last = min(dy0 + TY - 1, H - 1);
    int j_lo = max(0, -dy_last);
    int j_hi = min(H, H - dy0);
    {
        const int cnt = j_hi - j_lo;
        const int a = j_lo + (int)((int64_t)cnt * ks / split);
        const int b = j_lo + (int)((int64_t)cnt * (ks + 1) / split);
        j_lo = a;
        j_hi = b;
    }

    // Left columns this wave needs: c in [max(0, dx), min(W, W + dx)) over its lanes.
    const int dxw_hi = min(dxw + 63, W - 1);
    const int c_lo = max(0, dxw) & ~(CW - 1);
    const int c_hi = min(W, W + dxw_hi);
    const int c_full = W & ~(CW - 1);  // chunks below this need no column check

    // LDS row: buf[P] = R[j][qb + P] (0 outside [0, W)), qb <= min over the block of c - dx.
    const int qb = -((dxb + kBX - 1 + 3) & ~3);

    T tot[TY], comp[TY];
#pragma unroll
    for (int y = 0; y < TY; ++y) tot[y] = comp[y] = T(0);

    if (j_lo < j_hi) {
        auto fetch = [&](int j, int e) -> T {
            const int q = qb + e;
            return (q >= 0 && q < W) ? Rp[(size_t)j * W + q] : T(0);
        };
        for (int e = threadIdx.x; e < len; e += 256) buf[(j_lo & 1) * len + e] = fetch(j_lo, e);
        __syncthreads();

        for (int j = j_lo; j < j_hi; ++j) {
            T* cur = buf + (j & 1) * len;
            T* nxt = buf + ((j + 1) & 1) * len;
            const bool more = j + 1 < j_hi;
            T pre[kPrefetch];
#pragma unroll
            for (int i = 0; i < kPrefetch; ++i) {
                const int e = threadIdx.x + i * 256;
                pre[i] = (more && e < len) ? fetch(j + 1, e) : T(0);
            }

            // Left rows r = j + dy0 + y; rows outside [0, H) contribute nothing.  At the two
            // ends of the j range (at most TY-1 rows each) a missing row reads a clamped
            // row and its partial sum is dropped.
            const int r0 = j + dy0;
            T acc[TY];
            row_pair<T, TY, CW>(acc, Lp, r0, H, W, cur, c_lo - dx - qb, c_lo, c_hi, c_full);
            // Compensated (Kahan) add of the row sums: the error of summing the H row sums
            // in FP32 dominates a plain two-level sum.
#pragma unroll
            for (int y = 0; y < TY; ++y) {
                const T yv = acc[y] - comp[y];
                const T t = tot[y] + yv;
                comp[y] = (t - tot[y]) - yv;
                tot[y] = t;
            }

            if (more) {
#pragma unroll
                for (int i = 0; i < kPrefetch; ++i) {
                    const int e = threadIdx.x + i * 256;
                    if (e < len) nxt[e] = pre[i];
                }
                for (int e = threadIdx.x + kPrefetch * 256; e < len; e += 256) nxt[e] = fetch(j + 1, e);
            }
            __syncthreads();
        }
    }

    if (dx <= W - 1) {
#pragma unroll
        for (int y = 0; y < TY; ++y) {
            const int dy = dy0 + y;
            if (dy <= H - 1) {
                T* o = Op + (size_t)(dy + H - 1) * SX + (dx + W - 1);
                if (atomic) {
                    if (tot[y] != T(0)) atomic_add(o, tot[y]);
                } else {
                    *o = tot[y];
                }
            }
        }
    }
}

template <typename T>
__global__ void zero_kernel(T* __restrict__ p, uint64_t count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = T(0);
}

template <typename T, int TY, int CW>
void launch_direct(const family_args& a, const T* L, const T* R, T* O, uint32_t split, hipStream_t s) {
    const uint32_t SX = 2 * a.cols - 1, SY = 2 * a.rows - 1;
    const uint32_t tiles_x = (SX + kBX - 1) / kBX;
    const uint32_t tiles_y = (SY + TY - 1) / TY;
    if (split == 0) {
        // Enough blocks to fill 256 CUs several times over.
        const uint64_t base = (uint64_t)tiles_x * tiles_y * a.pairs;
        split = 1;
        while (base * split < 2048 && split * 2 <= a.rows / 4 + 1) split *= 2;
    }
    if (split > a.rows) split = a.rows;
    const uint64_t nblocks = (uint64_t)tiles_x * tiles_y * split * a.pairs;
    if (nblocks == 0 || nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    const int len = (int)(((a.cols + kBX + 8) + 3) & ~3u);
    const uint32_t lds = 2u * len * sizeof(T);
    if (lds > 64 * 1024) fail(XC_ERR_UNSUPPORTED, "matrix too wide for the direct kernel's LDS row buffer");
    const int atomic = split > 1 ? 1 : 0;
    if (atomic) {
        const uint64_t count = (uint64_t)a.pairs * SX * SY;
        uint32_t zb = (uint32_t)std::min<uint64_t>((count + 255) / 256, 8192);
        timed tz("zero_kernel", s);
        hipLaunchKernelGGL(zero_kernel<T>, dim3(zb), dim3(256), 0, s, O, count);
    }
    dim3 grid((uint32_t)nblocks), block(256);
    record_launch(sizeof(T) == 4 ? "direct_kernel<f32>" : "direct_kernel<f64>", grid, block, lds, a.rows, a.cols,
                  a.pairs, split);
    timed td("direct_kernel", s);
    hipLaunchKernelGGL((direct_kernel<T, TY, CW>), grid, block, lds, s, L, R, O, a.family, a.rows, a.cols, a.n,
                       a.m, split, tiles_x, tiles_y, len, atomic);
    XC_HIP(hipGetLastError());
}

}  // namespace

void launch_direct_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t split,
                       hipStream_t s) {
    launch_direct<float, 8, 4>(a, L, R, O, split, s);
}

void launch_direct_f64(const family_args& a, const double* L, const double* R, double* O, uint32_t split,
                       hipStream_t s) {
    launch_direct<double, 4, 4>(a, L, R, O, split, s);
}

void launch_zero_f32(float* p, uint64_t count, hipStream_t s) {
    uint32_t zb = (uint32_t)std::min<uint64_t>((count + 255) / 256, 8192);
    if (zb) hipLaunchKernelGGL(zero_kernel<float>, dim3(zb), dim3(256), 0, s, p, count);
}

void launch_zero_f64(double* p, uint64_t count, hipStream_t s) {
    uint32_t zb = (uint32_t)std::min<uint64_t>((count + 255) / 256, 8192);
    if (zb) hipLaunchKernelGGL(zero_kernel<double>, dim3(zb), dim3(256), 0, s, p, count);
}

}  // namespace xc
