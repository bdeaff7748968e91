// Wave64 definition-based kernel ("direct"), FP32 and FP64, with a runtime tile catalog:
// the device code, included by xc_direct.hip (float kernels, host planning) and
// xc_direct_f64.hip (double kernels).
//
// Replaces the warp-shuffle family of the reference (src/naive_shuffle.cu:86-267 and its
// multirow / multimat / work-distribution variants, src/naive_shuffle_*.cu): the same
// per-shift sums of src/validate.hpp:94-132, organised for CDNA4 instead of 32-lane
// shuffles.
//
//   * A wave covers 64 consecutive dx (one per lane).  Every lane walks the same (right
//     row j, left column c) sequence, so a left value L[j+dy][c] is wave-uniform: it is
//     read with scalar loads into SGPRs and enters the FMA as its scalar operand (the
//     reference broadcasts it with a shuffle per FMA).
//   * Right rows are staged into LDS with zero padding on both sides, so each lane reads
//     its window R[j][c-dx..] without bounds checks (the reference's
//     bound_checked_loads.cuh:8-53 become padding).  All waves of a block share them.
//   * Register tile of a thread: TY consecutive dy ("shifts per thread",
//     src/naive_shuffle_multirow_right.cu:90-171) x RM right matrices ("right matrices per
//     thread", src/naive_shuffle_multimat_right.cu:116-204) x LM left matrices (n_to_m,
//     src/naive_shuffle_n_to_m_multimat_both.cu:292-431).  Per column a thread does
//     TY*RM*LM FMAs from (TY+JB-1)*LM scalar values and JB*RM LDS values: RM and LM reuse
//     the staged values across matrices, TY and JB (right rows per step, the reference's
//     "left rows per iteration", src/naive_shuffle_multirow_both.cu:90-213) across rows.
//   * Products of one step (JB row pairs) are summed in FP32 in a fresh accumulator that
//     is then added to an FP64 total (float data): a single running FP32 sum exceeds 1e-6
//     of sum|a||b| at 256x256 for positive data.
//   * Work distribution (src/naive_shuffle.cu:306-413, src/row_distribution.cuh:34-284):
//     the right rows of each dy tile can be split over several blocks that add their
//     partial sums atomically into a zeroed output, either in `split` equal shares of the
//     tile's overlap ("rectangle": every tile gets the same number of blocks) or in chunks
//     of at most `rows_per_block` rows ("triangle": tiles with more overlap get more
//     blocks, so every block has about the same work).
#pragma once

#include <type_traits>

#include "xc_internal.hpp"
#include "xc_methods.hpp"

namespace xc {

struct direct_tile {
    int ty, rm, lm, jb;
};
// The compiled kernel of a catalog tile (nullptr if the tile is not in the catalog).
const void* direct_kernel_f32(const direct_tile& t);  // xc_direct.hip
const void* direct_kernel_f64(const direct_tile& t);  // xc_direct_f64.hip

namespace {  // device code: one copy per translation unit


constexpr int kMaxCols = 32768;
// Stands in for left rows outside [0, H): zero-initialised when the code object loads.
__device__ __attribute__((aligned(256))) double g_zero_row[kMaxCols];

struct dk_args {
    int family;
    int H, W;
    uint32_t n, m, mpl;          // mpl: n_to_mn right matrices per left (m / n)
    uint32_t tiles_x, tiles_y;   // output tiles of 64*nwx dx by TY*nwy dy
    uint32_t chunks;             // blocks per tile along the right rows
    uint32_t rows_per_chunk;     // triangle distribution: rows per block; 0 = equal shares
    uint32_t groups_l, groups_r; // groups of LM left and RM right matrices
    uint32_t nwx, nwy;           // waves of a block along dx and along dy
    int len;                     // LDS elements per staged row
    int atomic;                  // 1: add into a zeroed output
};

template <typename T>
struct total_of {
    using type = double;  // FP32 step sums are added into FP64 totals; FP64 sums stay FP64
};

// Columns per chunk of the inner loop: the left values of a chunk fit in 32 scalar
// registers.
template <typename T, int TY, int LM, int JB>
constexpr int cols_per_step() {
    constexpr int per = (TY + JB - 1) * LM * (int)(sizeof(T) / 4);
    return per * 4 <= 32 ? 4 : (per * 2 <= 32 ? 2 : 1);
}

// One block: dx in [dxb, dxb + 64*nwx), dy in [dyb, dyb + TY*nwy), the right rows of its
// chunk, one group of LM left x RM right matrices.
template <typename T, int TY, int RM, int LM, int JB>
__global__ __launch_bounds__(512) void direct_kernel(const T* __restrict__ L, const T* __restrict__ R,
                                                     T* __restrict__ O, dk_args a) {
    constexpr int CW = cols_per_step<T, TY, LM, JB>();
    constexpr int NU = TY + JB - 1;  // left rows a step reads per left matrix
    constexpr int NPF = 8;           // staged elements per thread prefetched into registers
    constexpr int FLUSH = (64 / (CW * JB)) > 0 ? 64 / (CW * JB) : 1;  // chunks per FP32 partial sum
    using Tot = typename total_of<T>::type;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    T* buf = reinterpret_cast<T*>(smem_raw);  // [2][JB][RM][len]

    const int H = a.H, W = a.W;
    const int SX = 2 * W - 1, SY = 2 * H - 1;
    const int nthreads = (int)blockDim.x;
    const int rows_staged = JB * RM;
    const int stage_elems = rows_staged * a.len;

    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = bid % a.tiles_x;
    bid /= a.tiles_x;
    const uint32_t ty = bid % a.tiles_y;
    bid /= a.tiles_y;
    const uint32_t ks = bid % a.chunks;
    bid /= a.chunks;
    const uint32_t gr = bid % a.groups_r;
    const uint32_t gl = bid / a.groups_r;

    // Members of the group: left li[l], right ri[i], whether they exist, result index.
    uint32_t li[LM], ri[RM];
    bool lok[LM], rok[RM];
    if (a.family == XC_N_TO_M) {
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            const uint32_t v = gl * LM + l;
            lok[l] = v < a.n;
            li[l] = lok[l] ? v : a.n - 1;
        }
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            const uint32_t v = gr * RM + i;
            rok[i] = v < a.m;
            ri[i] = rok[i] ? v : a.m - 1;
        }
    } else if (a.family == XC_N_TO_MN) {
        const uint32_t l0 = gr % a.n, qg = gr / a.n;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            lok[l] = l == 0;
            li[l] = l0;
        }
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            const uint32_t q = qg * RM + i;
            rok[i] = q < a.mpl;
            ri[i] = l0 + a.n * (rok[i] ? q : a.mpl - 1);
        }
    } else {  // one_to_one, one_to_many
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            lok[l] = l == 0;
            li[l] = 0;
        }
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            const uint32_t v = gr * RM + i;
            rok[i] = v < a.m;
            ri[i] = rok[i] ? v : a.m - 1;
        }
    }

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int wx = wave % (int)a.nwx, wy = wave / (int)a.nwx;
    const int bxw = 64 * (int)a.nwx;
    const int dxb = (int)tx * bxw - (W - 1);
    const int dyb = (int)ty * (TY * (int)a.nwy) - (H - 1);
    const int dxw = dxb + 64 * wx;
    const int dx = dxw + lane;
    const int dy0 = dyb + TY * wy;

    // Right rows of the block: j in [max(0, -dy_last), min(H, H - dyb)), then its chunk.
    int jb_lo = max(0, -min(dyb + TY * (int)a.nwy - 1, H - 1));
    int jb_hi = min(H, H - dyb);
    if (a.rows_per_chunk) {
        const int s = jb_lo + (int)(ks * a.rows_per_chunk);
        jb_hi = min(jb_hi, s + (int)a.rows_per_chunk);
        jb_lo = s;
    } else {
        const int cnt = jb_hi - jb_lo;
        const int s = jb_lo + (int)((int64_t)cnt * ks / a.chunks);
        jb_hi = jb_lo + (int)((int64_t)cnt * (ks + 1) / a.chunks);
        jb_lo = s;
    }
    if (jb_lo >= jb_hi) return;  // the same for every thread of the block

    // Rows of this wave (its dy may meet fewer of them) and its left columns:
    // c in [max(0, dx), min(W, W + dx)) over its lanes.
    const bool wave_out = dxw <= W - 1 && dy0 <= H - 1;
    const int jw_lo = max(jb_lo, max(0, -min(dy0 + TY - 1, H - 1)));
    const int jw_hi = min(jb_hi, min(H, H - dy0));
    const int dxw_hi = min(dxw + 63, W - 1);
    const int c_lo = max(0, dxw) & ~(CW - 1);
    const int c_hi = min(W, W + dxw_hi);
    const int c_full = W & ~(CW - 1);  // chunks below this need no column check

    // LDS row element e holds column qb + e of its right row (0 outside [0, W)).
    const int qb = -((dxb + bxw - 1 + 3) & ~3);

    Tot tot[TY][RM][LM];
#pragma unroll
    for (int y = 0; y < TY; ++y)
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
            for (int l = 0; l < LM; ++l) tot[y][i][l] = Tot(0);

    // Right matrix of group member i: rbase + rstride * min(rg0 + i, rcap).
    uint32_t rbase = 0, rstride = 1, rg0 = gr * RM, rcap = a.m - 1;
    if (a.family == XC_N_TO_MN) {
        rbase = gr % a.n;
        rstride = a.n;
        rg0 = (gr / a.n) * RM;
        rcap = a.mpl - 1;
    }
    // Staging: thread slot s of a step holds element e = tid + s * nthreads of the JB x RM
    // staged rows; its source (row b of right member i, column q) is the same every step,
    // so the first NPF slots keep it packed as q | b << 20 | i << 24 (~0u: column outside
    // [0, W) or no element).
    constexpr uint32_t kNone = ~0u;
    uint32_t spack[NPF];
#pragma unroll
    for (int s = 0; s < NPF; ++s) {
        const int e = (int)threadIdx.x + s * nthreads;
        const int row = e / a.len, q = qb + (e - row * a.len);
        const int b = row / RM, i = row - b * RM;
        spack[s] = (e < stage_elems && q >= 0 && q < W) ? (uint32_t)q | (uint32_t)b << 20 | (uint32_t)i << 24 : kNone;
    }
    auto source = [&](int i, int j, int q) -> T {
        const uint32_t mat = rbase + rstride * min(rg0 + (uint32_t)i, rcap);
        return R[((size_t)mat * H + j) * W + q];
    };
    auto slot = [&](int j0, int s) -> T {
        const uint32_t v = spack[s];
        const int q = (int)(v & 0xFFFFF), b = (int)((v >> 20) & 0xF), i = (int)(v >> 24);
        return (v != kNone && j0 + b < jb_hi) ? source(i, j0 + b, q) : T(0);
    };
    auto fetch = [&](int j0, int e) -> T {  // elements beyond the NPF slots
        const int row = e / a.len, q = qb + (e - row * a.len);
        const int b = row / RM, i = row - b * RM;
        return (j0 + b < jb_hi && q >= 0 && q < W) ? source(i, j0 + b, q) : T(0);
    };

#pragma unroll
    for (int s = 0; s < NPF; ++s) {
        const int e = (int)threadIdx.x + s * nthreads;
        if (e < stage_elems) buf[e] = slot(jb_lo, s);
    }
    for (int e = (int)threadIdx.x + NPF * nthreads; e < stage_elems; e += nthreads) buf[e] = fetch(jb_lo, e);
    __syncthreads();
    const T* zero = reinterpret_cast<const T*>(g_zero_row);

    int p = 0;
    for (int j0 = jb_lo; j0 < jb_hi; j0 += JB, p ^= 1) {
        const T* cur = buf + p * stage_elems;
        T* nxt = buf + (p ^ 1) * stage_elems;
        const int j1 = j0 + JB;
        const bool more = j1 < jb_hi;
        T pre[NPF];
#pragma unroll
        for (int s = 0; s < NPF; ++s) pre[s] = more ? slot(j1, s) : T(0);

        if (wave_out && j0 < jw_hi && j1 > jw_lo) {
            // Left rows r = j0 + dy0 + u, u < NU; rows outside [0, H) read zeros.
            const T* lrow[LM][NU];
#pragma unroll
            for (int l = 0; l < LM; ++l)
#pragma unroll
                for (int u = 0; u < NU; ++u) {
                    const int r = j0 + dy0 + u;
                    lrow[l][u] = (lok[l] && r >= 0 && r < H) ? L + ((size_t)li[l] * H + r) * W : zero;
                }
            T acc[TY][RM][LM];
#pragma unroll
            for (int y = 0; y < TY; ++y)
#pragma unroll
                for (int i = 0; i < RM; ++i)
#pragma unroll
                    for (int l = 0; l < LM; ++l) acc[y][i][l] = T(0);

            // Values of one chunk of CW columns: right (LDS, per lane) and left (scalar).
            struct chunk {
                T rr[JB][RM][CW];
                T lv[LM][NU][CW];
            };
            const int P0 = c_lo - dx - qb;
            auto load = [&](chunk& v, int c0, auto guarded) {
                const int P = P0 + (c0 - c_lo);
#pragma unroll
                for (int b = 0; b < JB; ++b)
#pragma unroll
                    for (int i = 0; i < RM; ++i)
#pragma unroll
                        for (int t = 0; t < CW; ++t) v.rr[b][i][t] = cur[(b * RM + i) * a.len + P + t];
#pragma unroll
                for (int l = 0; l < LM; ++l)
#pragma unroll
                    for (int u = 0; u < NU; ++u)
#pragma unroll
                        for (int t = 0; t < CW; ++t) {
                            if constexpr (decltype(guarded)::value)
                                v.lv[l][u][t] = (c0 + t < W) ? lrow[l][u][min(c0 + t, W - 1)] : T(0);
                            else
                                v.lv[l][u][t] = lrow[l][u][c0 + t];
                        }
            };
            auto madd = [&](const chunk& v) {
#pragma unroll
                for (int t = 0; t < CW; ++t)
#pragma unroll
                    for (int b = 0; b < JB; ++b)
#pragma unroll
                        for (int y = 0; y < TY; ++y)
#pragma unroll
                            for (int i = 0; i < RM; ++i)
#pragma unroll
                                for (int l = 0; l < LM; ++l)
                                    acc[y][i][l] = fma(v.lv[l][y + b][t], v.rr[b][i][t], acc[y][i][l]);
            };
            // Partial sums are moved into the totals every FLUSH chunks, so that one FP32
            // partial sum holds at most about 64 products.
            auto flush = [&] {
#pragma unroll
                for (int y = 0; y < TY; ++y)
#pragma unroll
                    for (int i = 0; i < RM; ++i)
#pragma unroll
                        for (int l = 0; l < LM; ++l) {
                            tot[y][i][l] += (Tot)acc[y][i][l];
                            acc[y][i][l] = T(0);
                        }
            };
            // Chunks that start below min(c_hi, c_full) need no column check.
            const int c_unguarded = min(c_hi, c_full);
            const int nch = c_unguarded > c_lo ? (c_unguarded - c_lo + CW - 1) / CW : 0;
            chunk A;
            for (int k0 = 0; k0 < nch; k0 += FLUSH) {
                const int k1 = min(nch, k0 + FLUSH);
                for (int k = k0; k < k1; ++k) {
                    load(A, c_lo + k * CW, std::false_type{});
                    madd(A);
                }
                if (k1 < nch) flush();
            }
            const int c_tail = c_lo + nch * CW;
            if (c_tail < c_hi) {  // last partial chunk: columns >= W are absent
                load(A, c_tail, std::true_type{});
                madd(A);
            }
            flush();
        }

        if (more) {
#pragma unroll
            for (int s = 0; s < NPF; ++s) {
                const int e = (int)threadIdx.x + s * nthreads;
                if (e < stage_elems) nxt[e] = pre[s];
            }
            for (int e = (int)threadIdx.x + NPF * nthreads; e < stage_elems; e += nthreads) nxt[e] = fetch(j1, e);
        }
        __syncthreads();
    }

    if (!wave_out || dx > W - 1) return;
#pragma unroll
    for (int l = 0; l < LM; ++l)
#pragma unroll
        for (int i = 0; i < RM; ++i) {
            if (!lok[l] || !rok[i]) continue;
            const uint32_t k = a.family == XC_N_TO_M ? li[l] * a.m + ri[i] : ri[i];
            T* __restrict__ Op = O + (size_t)k * SX * SY;
#pragma unroll
            for (int y = 0; y < TY; ++y) {
                const int dy = dy0 + y;
                if (dy > H - 1) continue;
                T* o = Op + (size_t)(dy + H - 1) * SX + (dx + W - 1);
                const T v = (T)tot[y][i][l];
                if (a.atomic) {
                    if (v != T(0)) atomicAdd(o, v);
                } else {
                    *o = v;
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

}  // namespace

// ---- tile catalog ------------------------------------------------------------------------
// (TY, RM, LM, JB) register tiles compiled for each type: TY in 1..8 (and 16 for one or two
// right matrices), RM in 1..8, LM in 1..4, JB in 1..4, with TY*RM*LM <= 32 accumulators,
// JB*RM*LM <= 16 staged rows read per column and (TY+JB-1)*LM <= 16 left rows per step.  A
// request is rounded down to a catalog entry (round_to_catalog).
#define XC_DIRECT_TILES(X) \
    X(1, 1, 1, 1) X(1, 1, 1, 2) X(1, 1, 1, 4) X(1, 1, 2, 1) X(1, 1, 2, 2) X(1, 1, 2, 4) X(1, 1, 4, 1) \
    X(1, 1, 4, 2) X(1, 1, 4, 4) X(1, 2, 1, 1) X(1, 2, 1, 2) X(1, 2, 1, 4) X(1, 2, 2, 1) X(1, 2, 2, 2) \
    X(1, 2, 2, 4) X(1, 2, 4, 1) X(1, 2, 4, 2) X(1, 4, 1, 1) X(1, 4, 1, 2) X(1, 4, 1, 4) X(1, 4, 2, 1) \
    X(1, 4, 2, 2) X(1, 4, 4, 1) X(1, 8, 1, 1) X(1, 8, 1, 2) X(1, 8, 2, 1) X(2, 1, 1, 1) X(2, 1, 1, 2) \
    X(2, 1, 1, 4) X(2, 1, 2, 1) X(2, 1, 2, 2) X(2, 1, 2, 4) X(2, 1, 4, 1) X(2, 1, 4, 2) X(2, 2, 1, 1) \
    X(2, 2, 1, 2) X(2, 2, 1, 4) X(2, 2, 2, 1) X(2, 2, 2, 2) X(2, 2, 2, 4) X(2, 2, 4, 1) X(2, 2, 4, 2) \
    X(2, 4, 1, 1) X(2, 4, 1, 2) X(2, 4, 1, 4) X(2, 4, 2, 1) X(2, 4, 2, 2) X(2, 4, 4, 1) X(2, 8, 1, 1) \
    X(2, 8, 1, 2) X(2, 8, 2, 1) X(4, 1, 1, 1) X(4, 1, 1, 2) X(4, 1, 1, 4) X(4, 1, 2, 1) X(4, 1, 2, 2) \
    X(4, 1, 2, 4) X(4, 1, 4, 1) X(4, 2, 1, 1) X(4, 2, 1, 2) X(4, 2, 1, 4) X(4, 2, 2, 1) X(4, 2, 2, 2) \
    X(4, 2, 2, 4) X(4, 2, 4, 1) X(4, 4, 1, 1) X(4, 4, 1, 2) X(4, 4, 1, 4) X(4, 4, 2, 1) X(4, 4, 2, 2) \
    X(4, 8, 1, 1) X(4, 8, 1, 2) X(8, 1, 1, 1) X(8, 1, 1, 2) X(8, 1, 1, 4) X(8, 1, 2, 1) X(8, 2, 1, 1) \
    X(8, 2, 1, 2) X(8, 2, 1, 4) X(8, 2, 2, 1) X(8, 4, 1, 1) X(8, 4, 1, 2) X(8, 4, 1, 4) X(16, 1, 1, 1) \
    X(16, 2, 1, 1)

}  // namespace xc
