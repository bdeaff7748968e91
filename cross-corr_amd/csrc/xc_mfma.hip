// Split-precision MFMA implicit GEMM for float inputs (bf16x3, bf16, fp16).
//
// The definition (src/validate.hpp:94-132) is recast per left row r as a GEMM:
//
//     out[dy][dx] += sum_c'  R[r-dy][c'] * L[r][c'+dx]
//                   =  A_r[dy][c'] * B_r[c'][dx]
//
// with A_r the right-matrix rows r-dy (one per output row dy) and B_r the Hankel matrix of
// left row r (B_r[c'][dx] = L[r][c'+dx]).  Neither operand is materialised in HBM:
//   * the right rows live in an LDS ring of 64*WY+1 rows; step r adds row r-dy_min and
//     drops the oldest, so every right row is loaded once per block and is reused by 64*WY
//     consecutive steps (one per dy);
//   * left row r is staged into LDS four times, shifted by 0..3 elements, so that any
//     8-element Hankel window starts on an 8-byte boundary of one copy;
//   * v_mfma_f32_32x32x16_{bf16,f16}: D rows are dy, D columns are dx, so each lane's
//     results are consecutive dx (coalesced stores);
//   * bf16x3: each float x is split into hi = bf16(x), lo = bf16(x - hi) when staged;
//     hi*hi + hi*lo + lo*hi are accumulated in FP32 (the lo*lo term and the residual below
//     bf16(lo) are dropped: about 1e-5 relative per product);
//   * two-level FP32 accumulation: the products of one (left row, column block) step go
//     to a fresh accumulator that is added to the running total with a compensated
//     (Kahan) add after the step.
// Outputs whose overlap area is below the border threshold are recomputed exactly in FP32
// by frame_kernel: there, few products are summed and the split error does not average out.
#include <algorithm>

#include "xc_methods.hpp"

namespace xc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kKC = 128;              // columns c' per pass over the left rows
constexpr uint32_t kDefaultBorder = 512;  // overlap area below which bf16x3 outputs are recomputed

template <int KIND>
struct split_traits;
template <>
struct split_traits<XC_PREC_BF16X3> {
    static constexpr int planes = 2;
    __device__ static void split(float x, uint16_t& hi, uint16_t& lo) {
        __bf16 h = (__bf16)x;
        __bf16 l = (__bf16)(x - (float)h);
        hi = __builtin_bit_cast(uint16_t, h);
        lo = __builtin_bit_cast(uint16_t, l);
    }
};
template <>
struct split_traits<XC_PREC_BF16> {
    static constexpr int planes = 1;
    __device__ static void split(float x, uint16_t& hi, uint16_t& lo) {
        hi = __builtin_bit_cast(uint16_t, (__bf16)x);
        lo = 0;
    }
};
template <>
struct split_traits<XC_PREC_FP16> {
    static constexpr int planes = 1;
    __device__ static void split(float x, uint16_t& hi, uint16_t& lo) {
        hi = __builtin_bit_cast(uint16_t, (_Float16)x);
        lo = 0;
    }
};

template <int KIND>
__device__ __forceinline__ f32x16 mma(const uint4& a, const uint4& b, f32x16 c) {
    if constexpr (KIND == XC_PREC_FP16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
}

// LDS geometry (bytes) of one block.
template <int WX, int WY, int PLANES>
struct geom {
    static constexpr int BX = 64 * WX;                    // dx per block
    static constexpr int BY = 64 * WY;                    // dy per block
    static constexpr int NSLOT = BY + 1;                  // ring slots (right rows)
    static constexpr int RS = kKC * 2 + 16;               // ring row stride: an odd multiple of 16 B
    static constexpr int LENA = kKC + BX + 64;            // staged left-row elements per copy
    static constexpr int CS = ((LENA * 2 + 255) / 256) * 256 + 64;  // copy stride, 64 mod 256
    static constexpr int RING = NSLOT * PLANES * RS;
    static constexpr int LBUF = PLANES * 4 * CS;          // one left-row buffer (4 shifted copies)
    static constexpr int BYTES = RING + 2 * LBUF;
};

__device__ __forceinline__ int ring_slot(int row, int nslot) {
    int s = row % nslot;
    return s < 0 ? s + nslot : s;
}

template <int KIND, int WX, int WY>
__global__ __launch_bounds__(256, 1) void mfma_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                      float* __restrict__ O, int family, uint32_t rows,
                                                      uint32_t cols, uint32_t n, uint32_t m, uint32_t tiles_x,
                                                      uint32_t tiles_y) {
    using S = split_traits<KIND>;
    constexpr int P = S::planes;
    using G = geom<WX, WY, P>;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    char* ring = smem;
    char* lbuf = smem + G::RING;

    const int H = (int)rows, W = (int)cols;
    const int SX = 2 * W - 1;
    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = bid % tiles_x;
    bid /= tiles_x;
    const uint32_t ty = bid % tiles_y;
    const uint32_t k = bid / tiles_y;
    uint32_t li, ri;
    pair_of(family, n, m, k, li, ri);
    const float* __restrict__ Lp = L + (size_t)li * H * W;
    const float* __restrict__ Rp = R + (size_t)ri * H * W;
    float* __restrict__ Op = O + (size_t)k * SX * (2 * H - 1);

    const int dxb = (int)tx * G::BX - (W - 1);
    const int dyb = (int)ty * G::BY - (H - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int li32 = lane & 31, g = lane >> 5;
    const int wx = wave % WX, wy = wave / WX;
    const int dxw = dxb + 64 * wx;
    const int dyw = dyb + 64 * wy;

    // Left rows meeting the block's dy range: r in [max(0,dy), min(H, H+dy)).
    const int dy_hi_b = min(dyb + G::BY - 1, H - 1);
    const int r_lo = max(0, dyb);
    const int r_hi = min(H, H + dy_hi_b);
    // Columns c' meeting the block's / the wave's dx range: c' in [max(0,-dx), min(W, W-dx)).
    const int dx_hi_b = min(dxb + G::BX - 1, W - 1);
    const int cb_lo = max(0, -dx_hi_b), cb_hi = min(W, W - dxb);
    const int dx_hi_w = min(dxw + 63, W - 1);
    const int cw_lo = max(0, -dx_hi_w), cw_hi = min(W, W - dxw);

    f32x16 tot[2][2], comp[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 16; ++e) tot[u][s][e] = comp[u][s][e] = 0.f;

    // Stage right row `row` (columns [c0, c0+kKC)) into its ring slot.
    auto stage_right = [&](int row, int c0, int e, uint16_t* hi, uint16_t* lo) {
        const int c = c0 + e;
        float v = 0.f;
        if ((unsigned)row < (unsigned)H && c < W) v = Rp[(size_t)row * W + c];
        S::split(v, *hi, *lo);
    };
    auto ring_put = [&](int row, int e, uint16_t hi, uint16_t lo) {
        char* base = ring + ring_slot(row, G::NSLOT) * P * G::RS + e * 2;
        *reinterpret_cast<uint16_t*>(base) = hi;
        if (P == 2) *reinterpret_cast<uint16_t*>(base + G::RS) = lo;
    };
    // Left row r, element e of the staged window starting at column cb: written to the four
    // shifted copies (copy t holds column cb + p + t at position p).
    auto lbuf_put = [&](int buf, int e, uint16_t hi, uint16_t lo) {
        char* base = lbuf + buf * G::LBUF;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int p = e - t;
            if (p >= 0 && p < G::LENA) {
                *reinterpret_cast<uint16_t*>(base + t * G::CS + p * 2) = hi;
                if (P == 2) *reinterpret_cast<uint16_t*>(base + 4 * G::CS + t * G::CS + p * 2) = lo;
            }
        }
    };
    auto left_val = [&](int r, int c) -> float {
        return ((unsigned)r < (unsigned)H && (unsigned)c < (unsigned)W) ? Lp[(size_t)r * W + c] : 0.f;
    };

    constexpr int LE = G::LENA + 3;  // elements loaded per left row
    constexpr int LPT = (LE + 255) / 256;
    constexpr int RPT = kKC / 256 > 0 ? kKC / 256 : 1;

    for (int c0 = (cb_lo / kKC) * kKC; c0 < cb_hi; c0 += kKC) {
        const int cbase = c0 + dxb;  // left column at staged position 0
        // Prefill: right rows r_lo - dy for dy in the block (the ring), left row r_lo.
        for (int idx = threadIdx.x; idx < G::BY * kKC; idx += 256) {
            const int ro = idx / kKC, e = idx % kKC;
            const int row = r_lo - dyb - ro;
            uint16_t hi, lo;
            stage_right(row, c0, e, &hi, &lo);
            ring_put(row, e, hi, lo);
        }
        for (int e = threadIdx.x; e < LE; e += 256) {
            uint16_t hi, lo;
            S::split(left_val(r_lo, cbase + e), hi, lo);
            lbuf_put(0, e, hi, lo);
        }
        __syncthreads();

        // Wave's K chunks of this pass.
        const int q_lo = max(cw_lo, c0), q_hi = min(cw_hi, c0 + kKC);
        const int qk_lo = ((q_lo - c0) & ~15), qk_hi = q_hi - c0;

        for (int r = r_lo; r < r_hi; ++r) {
            const int cur = (r - r_lo) & 1;
            const bool more = r + 1 < r_hi;
            // Prefetch next step's data into registers.
            float pl[LPT];
            float pr[RPT];
#pragma unroll
            for (int i = 0; i < LPT; ++i) {
                const int e = threadIdx.x + i * 256;
                pl[i] = (more && e < LE) ? left_val(r + 1, cbase + e) : 0.f;
            }
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int e = threadIdx.x + i * 256;
                const int row = r + 1 - dyb;
                pr[i] = (more && e < kKC && (unsigned)row < (unsigned)H && c0 + e < W) ? Rp[(size_t)row * W + c0 + e]
                                                                                        : 0.f;
            }

            f32x16 acc[2][2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int e = 0; e < 16; ++e) acc[u][s][e] = 0.f;

            if (q_lo < q_hi) {
                const char* lb = lbuf + cur * G::LBUF;
                // Ring rows of this lane for the two dy sub-tiles: row = r - (dyw + 32u + li32).
                const char* arow[2];
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    arow[u] = ring + ring_slot(r - (dyw + 32 * u + li32), G::NSLOT) * P * G::RS + g * 16;
                for (int qq = qk_lo; qq < qk_hi; qq += 16) {
                    uint4 a[2][P], b[2][P];
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int pl_ = 0; pl_ < P; ++pl_)
                            a[u][pl_] = *reinterpret_cast<const uint4*>(arow[u] + pl_ * G::RS + qq * 2);
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int p = qq + 8 * g + (dxw - dxb) + 32 * s + li32;
                        const int t = p & 3;
#pragma unroll
                        for (int pl_ = 0; pl_ < P; ++pl_) {
                            const uint2* q =
                                reinterpret_cast<const uint2*>(lb + pl_ * 4 * G::CS + t * G::CS + (p - t) * 2);
                            const uint2 v0 = q[0], v1 = q[1];
                            b[s][pl_] = make_uint4(v0.x, v0.y, v1.x, v1.y);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int s = 0; s < 2; ++s) acc[u][s] = mma<KIND>(a[u][0], b[s][0], acc[u][s]);
                    if constexpr (P == 2) {
#pragma unroll
                        for (int u = 0; u < 2; ++u)
#pragma unroll
                            for (int s = 0; s < 2; ++s) {
                                acc[u][s] = mma<KIND>(a[u][0], b[s][1], acc[u][s]);
                                acc[u][s] = mma<KIND>(a[u][1], b[s][0], acc[u][s]);
                            }
                    }
                }
            }
            // Compensated (Kahan) add of this step's sums into the totals.
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const f32x16 yv = acc[u][s] - comp[u][s];
                    const f32x16 t = tot[u][s] + yv;
                    comp[u][s] = (t - tot[u][s]) - yv;
                    tot[u][s] = t;
                }

            if (more) {
#pragma unroll
                for (int i = 0; i < LPT; ++i) {
                    const int e = threadIdx.x + i * 256;
                    if (e < LE) {
                        uint16_t hi, lo;
                        S::split(pl[i], hi, lo);
                        lbuf_put(cur ^ 1, e, hi, lo);
                    }
                }
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int e = threadIdx.x + i * 256;
                    if (e < kKC) {
                        uint16_t hi, lo;
                        S::split(pr[i], hi, lo);
                        ring_put(r + 1 - dyb, e, hi, lo);
                    }
                }
            }
            __syncthreads();
        }
    }

    // D layout of v_mfma_f32_32x32x16: lane -> column (dx), register a -> row
    // 8*(a/4) + 4*(lane/32) + a%4 (dy).
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int dx = dxw + 32 * s + li32;
        if (dx > W - 1) continue;
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < 16; ++a) {
                const int dy = dyw + 32 * u + 8 * (a >> 2) + 4 * g + (a & 3);
                if (dy <= H - 1) Op[(size_t)(dy + H - 1) * SX + (dx + W - 1)] = tot[u][s][a];
            }
    }
}

// Exact FP32 recomputation of the outputs whose overlap area (H-|dy|)*(W-|dx|) is below
// `thresh`.  One block per (pair, dy); threads walk that row's frame outputs.
__global__ __launch_bounds__(64) void frame_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                   float* __restrict__ O, int family, uint32_t rows, uint32_t cols,
                                                   uint32_t n, uint32_t m, uint32_t thresh) {
    const int H = (int)rows, W = (int)cols;
    const int SY = 2 * H - 1, SX = 2 * W - 1;
    const uint32_t k = blockIdx.x / SY;
    const int dy = (int)(blockIdx.x % SY) - (H - 1);
    uint32_t li, ri;
    pair_of(family, n, m, k, li, ri);
    const float* __restrict__ Lp = L + (size_t)li * H * W;
    const float* __restrict__ Rp = R + (size_t)ri * H * W;
    float* __restrict__ Op = O + (size_t)k * SX * SY;
    const int h = H - abs(dy);
    const int wmax = min(W, (int)((thresh + h - 1) / h) - 1);  // overlap widths w with h*w < thresh
    if (wmax <= 0) return;
    const int count = wmax >= W ? SX : 2 * wmax;
    const int r0 = max(0, dy), r1 = min(H, H + dy);
    for (int f = threadIdx.x; f < count; f += 64) {
        int dx;
        if (wmax >= W) dx = f - (W - 1);
        else dx = f < wmax ? -(W - 1) + f : (W - 1) - (f - wmax);
        const int c_lo = max(0, dx), c_hi = min(W, W + dx);
        float tot = 0.f;
        for (int r = r0; r < r1; ++r) {
            const float* lr = Lp + (size_t)r * W;
            const float* rr = Rp + (size_t)(r - dy) * W - dx;
            float acc = 0.f;
            for (int c = c_lo; c < c_hi; ++c) acc = fmaf(lr[c], rr[c], acc);
            tot += acc;
        }
        Op[(size_t)(dy + H - 1) * SX + (dx + W - 1)] = tot;
    }
}

template <int KIND, int WX, int WY>
void launch(const family_args& a, const float* L, const float* R, float* O, hipStream_t s) {
    using G = geom<WX, WY, split_traits<KIND>::planes>;
    const uint32_t SX = 2 * a.cols - 1, SY = 2 * a.rows - 1;
    const uint32_t tiles_x = (SX + G::BX - 1) / G::BX;
    const uint32_t tiles_y = (SY + G::BY - 1) / G::BY;
    const uint64_t nblocks = (uint64_t)tiles_x * tiles_y * a.pairs;
    if (nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    static bool attr_set = false;
    if (!attr_set) {
        XC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&mfma_kernel<KIND, WX, WY>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, G::BYTES));
        attr_set = true;
    }
    dim3 grid((uint32_t)nblocks), block(256);
    record_launch(KIND == XC_PREC_BF16X3 ? "mfma_kernel<bf16x3>" : (KIND == XC_PREC_BF16 ? "mfma_kernel<bf16>"
                                                                                          : "mfma_kernel<fp16>"),
                  grid, block, G::BYTES, a.rows, a.cols, a.pairs, 1);
    hipLaunchKernelGGL((mfma_kernel<KIND, WX, WY>), grid, block, G::BYTES, s, L, R, O, a.family, a.rows, a.cols, a.n,
                       a.m, tiles_x, tiles_y);
    XC_HIP(hipGetLastError());
}

}  // namespace

void launch_frame_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t thresh, hipStream_t s) {
    const uint64_t nblocks = (uint64_t)a.pairs * (2 * a.rows - 1);
    if (nblocks == 0 || thresh == 0) return;
    if (nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    hipLaunchKernelGGL(frame_kernel, dim3((uint32_t)nblocks), dim3(64), 0, s, L, R, O, a.family, a.rows, a.cols,
                       a.n, a.m, thresh);
    XC_HIP(hipGetLastError());
}

void launch_mfma_f32(const family_args& a, const float* L, const float* R, float* O, int kind, uint32_t border,
                     hipStream_t s) {
    switch (kind) {
    case XC_PREC_BF16X3:
        launch<XC_PREC_BF16X3, 4, 1>(a, L, R, O, s);
        launch_frame_f32(a, L, R, O, border ? border : kDefaultBorder, s);
        break;
    case XC_PREC_BF16:
        launch<XC_PREC_BF16, 4, 1>(a, L, R, O, s);
        break;
    case XC_PREC_FP16:
        launch<XC_PREC_FP16, 4, 1>(a, L, R, O, s);
        break;
    default:
        fail(XC_ERR_INVALID_ARGUMENT, "not an MFMA precision");
    }
}

}  // namespace xc
