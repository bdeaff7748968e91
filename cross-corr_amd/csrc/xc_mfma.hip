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
#include <type_traits>

#include "xc_methods.hpp"

namespace xc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMaxBorder = 512;       // largest accepted border threshold
// Overlap area below which split-precision outputs are recomputed exactly.  Measured maximum
// normalised error of the tiles alone per overlap-area octave (u01, u11 and normal data, 32^2
// to 256^2): bf16x3 falls from 2e-5 at area 1 to 2.1e-6 at 256..511 and 1.6e-6 at 512..1023;
// fp16x3 (22 significant bits per operand) is at most 1.1e-6 from area 4 on, 2e-5 below.
constexpr uint32_t kDefaultBorder = 512;        // bf16x3
constexpr uint32_t kDefaultBorderFp16x3 = 16;   // fp16x3

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
struct split_traits<XC_PREC_FP16X3> {
    static constexpr int planes = 2;
    __device__ static void split(float x, uint16_t& hi, uint16_t& lo) {
        _Float16 h = (_Float16)x;
        _Float16 l = (_Float16)(x - (float)h);
        hi = __builtin_bit_cast(uint16_t, h);
        lo = __builtin_bit_cast(uint16_t, l);
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
    if constexpr (KIND == XC_PREC_FP16 || KIND == XC_PREC_FP16X3)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
}

// LDS geometry (bytes) of one block; KC columns c' per pass over the left rows, JR left rows
// per step (between two block barriers).
template <int WX, int WY, int PLANES, int KC, int JR = 1>
struct geom {
    static constexpr int BX = 64 * WX;                    // dx per block
    static constexpr int BY = 64 * WY;                    // dy per block
    // ring slots (right rows): the BY + JR - 1 rows a step reads and the JR the next step adds
    static constexpr int NSLOT = BY + 2 * JR - 1;
    static constexpr int RS = KC * 2 + 16;               // ring row stride: an odd multiple of 16 B
    static constexpr int LENA = KC + BX + 64;            // staged left-row elements per copy
    // Copy stride = 32 mod 128: the 16 lanes of one ds_read_b64 cycle (4 alignments x 4
    // consecutive 8-byte words) then fall on 16 distinct 8-byte bank pairs of the 128-byte
    // LDS row (measured: 4.3x fewer SQ_LDS_BANK_CONFLICT cycles than 64 mod 256).
    static constexpr int CS = ((LENA * 2 + 127) / 128) * 128 + 32;
    // Slot (one ring row, all planes) stride: an odd multiple of 16 B, so that the 32 lanes
    // reading 32 consecutive slots hit distinct LDS banks.
    static constexpr int SS = ((PLANES * RS / 16) % 2) ? PLANES * RS : PLANES * RS + 16;
    static constexpr int RING = NSLOT * SS;
    static constexpr int LBUF = PLANES * 4 * CS;          // one left-row buffer (4 shifted copies)
    static constexpr int BYTES = RING + 2 * JR * LBUF;
};

__device__ __forceinline__ int ring_slot(int row, int nslot) {
    int s = row % nslot;
    return s < 0 ? s + nslot : s;
}

// MFMAs of one left-row step over chunks [q0, q1) of 16 columns.  MASK bit 2u+s selects the
// (dy sub-tile u, dx sub-tile s) accumulators this range feeds; fragments of the next chunk
// are loaded while the current chunk's MFMAs run (the one-past-the-end load stays inside the
// LDS ring / row buffers).
template <int KIND, int P, int MASK, bool PIPE>
__device__ __forceinline__ void chunk_range(f32x16 (&acc)[2][2], const char* const (&arow)[2],
                                            const char* const (&bcol)[2], int q0, int q1, int rs, int cs) {
    auto load = [&](int qq, uint4 (&a)[2][P], uint4 (&b)[2][P]) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int pl = 0; pl < P; ++pl)
                if (MASK & (3 << (2 * u))) a[u][pl] = *reinterpret_cast<const uint4*>(arow[u] + pl * rs + qq * 2);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int pl = 0; pl < P; ++pl)
                if (MASK & (5 << s)) {
                    const uint2* q = reinterpret_cast<const uint2*>(bcol[s] + pl * 4 * cs + qq * 2);
                    const uint2 v0 = q[0], v1 = q[1];
                    b[s][pl] = make_uint4(v0.x, v0.y, v1.x, v1.y);
                }
    };
    uint4 a0[2][P], b0[2][P];
    if (PIPE) load(q0, a0, b0);
    for (int qq = q0; qq < q1; qq += 16) {
        uint4 a1[2][P], b1[2][P];
        if (PIPE) load(qq + 16, a1, b1);
        else load(qq, a0, b0);
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (MASK & (1 << (2 * u + s))) acc[u][s] = mma<KIND>(a0[u][0], b0[s][0], acc[u][s]);
        if constexpr (P == 2) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    if (MASK & (1 << (2 * u + s))) acc[u][s] = mma<KIND>(a0[u][0], b0[s][1], acc[u][s]);
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    if (MASK & (1 << (2 * u + s))) acc[u][s] = mma<KIND>(a0[u][1], b0[s][0], acc[u][s]);
        }
        if (PIPE) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int pl = 0; pl < P; ++pl) {
                    a0[i][pl] = a1[i][pl];
                    b0[i][pl] = b1[i][pl];
                }
        }
    }
}

// Same with the mask known only at run time (uniform branches around each MFMA group).
template <int KIND, int P>
__device__ __forceinline__ void chunk_range_rt(int mask, f32x16 (&acc)[2][2], const char* const (&arow)[2],
                                               const char* const (&bcol)[2], int q0, int q1, int rs, int cs) {
    for (int qq = q0; qq < q1; qq += 16) {
        uint4 a0[2][P], b0[2][P];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int pl = 0; pl < P; ++pl) a0[u][pl] = *reinterpret_cast<const uint4*>(arow[u] + pl * rs + qq * 2);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int pl = 0; pl < P; ++pl) {
                const uint2* q = reinterpret_cast<const uint2*>(bcol[s] + pl * 4 * cs + qq * 2);
                const uint2 v0 = q[0], v1 = q[1];
                b0[s][pl] = make_uint4(v0.x, v0.y, v1.x, v1.y);
            }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (mask & (1 << (2 * u + s))) {
                    acc[u][s] = mma<KIND>(a0[u][0], b0[s][0], acc[u][s]);
                    if constexpr (P == 2) {
                        acc[u][s] = mma<KIND>(a0[u][0], b0[s][1], acc[u][s]);
                        acc[u][s] = mma<KIND>(a0[u][1], b0[s][0], acc[u][s]);
                    }
                }
    }
}

template <int KIND, int P, bool PIPE>
__device__ __forceinline__ void chunk_range_dispatch(int mask, f32x16 (&acc)[2][2], const char* const (&arow)[2],
                                                     const char* const (&bcol)[2], int q0, int q1, int rs, int cs) {
    if constexpr (!PIPE) {
        if (mask == 15) chunk_range<KIND, P, 15, false>(acc, arow, bcol, q0, q1, rs, cs);
        else chunk_range_rt<KIND, P>(mask, acc, arow, bcol, q0, q1, rs, cs);
        return;
    }
    switch (mask) {
#define XC_CASE(M) \
    case M: chunk_range<KIND, P, M, PIPE>(acc, arow, bcol, q0, q1, rs, cs); break;
        XC_CASE(1) XC_CASE(2) XC_CASE(3) XC_CASE(4) XC_CASE(5) XC_CASE(6) XC_CASE(7) XC_CASE(8)
        XC_CASE(9) XC_CASE(10) XC_CASE(11) XC_CASE(12) XC_CASE(13) XC_CASE(14) XC_CASE(15)
#undef XC_CASE
    default: break;
    }
}

// KC: columns per pass; FLUSH: left-row steps summed in `acc` before it is added to the
// total; KAHAN: compensated add of those sums (else plain); OCC: blocks per CU the register
// budget is sized for.
// One block tile (dx tile tx, dy tile ty) of the pair (Lp, Rp) -> Op, both H x W.  Outputs
// whose overlap area is below skip_area are left to the border-frame code (0: write all).
template <int KIND, int WX, int WY, int KC, int FLUSH, bool KAHAN, int OCC, int JR = 1>
__device__ __forceinline__ void mfma_tile(const float* __restrict__ Lp, const float* __restrict__ Rp,
                                          float* __restrict__ Op, const int H, const int W, const uint32_t tx,
                                          const uint32_t ty, const uint32_t skip_area, char* smem) {
    using S = split_traits<KIND>;
    constexpr int P = S::planes;
    using G = geom<WX, WY, P, KC, JR>;
    static_assert(JR == 1 || JR == 2, "one or two left rows per step");
    static_assert(FLUSH % JR == 0, "flushes fall on step boundaries");
    static_assert(JR * (KC / 4) <= 256, "one thread per staged right-row quad");
    constexpr bool PIPE = OCC == 1;  // two waves per SIMD hide LDS latency without it
    char* ring = smem;
    char* lbuf = smem + G::RING;
    const int SX = 2 * W - 1;

    const int dxb = (int)tx * G::BX - (W - 1);
    const int dyb = (int)ty * G::BY - (H - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int li32 = lane & 31, g = lane >> 5;
    const int wx = wave % WX, wy = wave / WX;
    // The wave's two 32-wide dx sub-tiles are 32*WX apart, so that every wave of the block
    // gets one sub-tile from each half of the block's dx range (balanced work per step).
    const int dxs[2] = {dxb + 32 * wx, dxb + 32 * wx + 32 * WX};
    const int dyw = dyb + 64 * wy;

    // Left rows meeting the block's dy range: r in [max(0,dy), min(H, H+dy)).
    const int dy_hi_b = min(dyb + G::BY - 1, H - 1);
    const int r_lo = max(0, dyb);
    const int r_hi = min(H, H + dy_hi_b);
    // Columns c' meeting the block's / the wave's dx range: c' in [max(0,-dx), min(W, W-dx)).
    const int dx_hi_b = min(dxb + G::BX - 1, W - 1);
    const int cb_lo = max(0, -dx_hi_b), cb_hi = min(W, W - dxb);
    int cw_lo = W, cw_hi = 0;
#pragma unroll
    for (int sidx = 0; sidx < 2; ++sidx) {
        const int lo = dxs[sidx], hi = min(lo + 31, W - 1);
        if (lo <= hi) {
            cw_lo = min(cw_lo, max(0, -hi));
            cw_hi = max(cw_hi, min(W, W - lo));
        }
    }

    f32x16 tot[2][2], comp[2][2], acc[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 16; ++e) tot[u][s][e] = comp[u][s][e] = acc[u][s][e] = 0.f;

    // Staging works on quads of 4 consecutive elements, written as one 8-byte LDS store per
    // plane.  Left row: item (copy t, quad j) holds columns cbase + 4j + t .. +3 at positions
    // 4j .. 4j+3 of copy t.  Right row: quad j holds columns c0 + 4j .. +3 of its ring slot.
    static_assert(G::LENA % 4 == 0 && KC % 4 == 0, "quad staging");
    constexpr int LQ = G::LENA / 4;       // quads per left copy
    constexpr int RQ = KC / 4;           // quads per right row
    static_assert(JR * LQ <= 256, "one thread per staged (left row, quad)");
    // Left rows: thread (row i, quad j) loads the 8 values at staged positions 4j .. 4j+7,
    // splits them once and writes quad j of all four shifted copies (as in mfma_rows_tile).
    struct lvals {
        float v[8];
    };
    auto load_left8 = [&](int r, int cbase, int j) -> lvals {
        lvals x;
        const int c = cbase + 4 * j;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            x.v[k] = ((unsigned)r < (unsigned)H && (unsigned)(c + k) < (unsigned)W) ? Lp[(size_t)r * W + c + k] : 0.f;
        return x;
    };
    auto load_right = [&](int row, int c0, int j) -> float4 {
        const int c = c0 + 4 * j;
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            v[k] = ((unsigned)row < (unsigned)H && c + k < W) ? Rp[(size_t)row * W + c + k] : 0.f;
        return make_float4(v[0], v[1], v[2], v[3]);
    };
    auto split4 = [&](float4 v, uint2& hi, uint2& lo) {
        uint16_t h[4], l[4];
        S::split(v.x, h[0], l[0]);
        S::split(v.y, h[1], l[1]);
        S::split(v.z, h[2], l[2]);
        S::split(v.w, h[3], l[3]);
        hi = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
        lo = make_uint2(l[0] | ((uint32_t)l[1] << 16), l[2] | ((uint32_t)l[3] << 16));
    };
    auto put_left8 = [&](int buf, int j, const lvals& x) {
        uint2 h0, l0, h1, l1;
        split4(make_float4(x.v[0], x.v[1], x.v[2], x.v[3]), h0, l0);
        split4(make_float4(x.v[4], x.v[5], x.v[6], x.v[7]), h1, l1);
        auto shifted = [](uint2 a, uint2 b, int t) -> uint2 {  // elements t .. t+3 of (a, b)
            const uint32_t w0 = a.x, w1 = a.y, w2 = b.x, w3 = b.y;
            if (t == 0) return make_uint2(w0, w1);
            if (t == 1) return make_uint2((w0 >> 16) | (w1 << 16), (w1 >> 16) | (w2 << 16));
            if (t == 2) return make_uint2(w1, w2);
            return make_uint2((w1 >> 16) | (w2 << 16), (w2 >> 16) | (w3 << 16));
        };
        char* base = lbuf + buf * G::LBUF + j * 8;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            *reinterpret_cast<uint2*>(base + t * G::CS) = shifted(h0, h1, t);
            if (P == 2) *reinterpret_cast<uint2*>(base + (4 + t) * G::CS) = shifted(l0, l1, t);
        }
    };
    auto put_right = [&](int row, int j, float4 v) {
        uint2 hi, lo;
        split4(v, hi, lo);
        char* base = ring + ring_slot(row, G::NSLOT) * G::SS + j * 8;
        *reinterpret_cast<uint2*>(base) = hi;
        if (P == 2) *reinterpret_cast<uint2*>(base + G::RS) = lo;
    };

    for (int c0 = (cb_lo / KC) * KC; c0 < cb_hi; c0 += KC) {
        const int cbase = c0 + dxb;  // left column at staged position 0
        // Prefill: right rows r - dy for r in [r_lo, r_lo + JR) and dy in the block (the ring),
        // left rows r_lo .. r_lo + JR - 1 into buffers 0 .. JR - 1.
        for (int idx = threadIdx.x; idx < (G::BY + JR - 1) * RQ; idx += 256) {
            const int ro = idx / RQ, j = idx - (idx / RQ) * RQ;
            const int row = r_lo + JR - 1 - dyb - ro;
            put_right(row, j, load_right(row, c0, j));
        }
        if (threadIdx.x < JR * LQ) {
            const int i = threadIdx.x / LQ, j = threadIdx.x - (threadIdx.x / LQ) * LQ;
            put_left8(i, j, load_left8(r_lo + i, cbase, j));
        }
        __syncthreads();

        // Wave's K chunks of this pass, and those of each dx sub-tile (offsets from c0,
        // rounded out to chunks of 16; empty when the sub-tile is outside the output).
        const int q_lo = max(cw_lo, c0), q_hi = min(cw_hi, c0 + KC);
        int ks_lo[2], ks_hi[2];
#pragma unroll
        for (int sidx = 0; sidx < 2; ++sidx) {
            const int lo = dxs[sidx], hi = min(lo + 31, W - 1);
            const int a0 = max(max(0, -hi), c0) - c0, a1 = min(min(W, W - lo), c0 + KC) - c0;
            if (lo <= hi && a0 < a1) {
                ks_lo[sidx] = a0 & ~15;
                ks_hi[sidx] = (a1 + 15) & ~15;
            } else {
                ks_lo[sidx] = ks_hi[sidx] = 0;
            }
        }

        for (int r0 = r_lo; r0 < r_hi; r0 += JR) {
            const int cur = ((r0 - r_lo) / JR) & 1;  // buffer set of this step's left rows
            const bool more = r0 + JR < r_hi;
            // Prefetch the next step's rows into registers.
            const int li_row = threadIdx.x / LQ, lj = threadIdx.x - (threadIdx.x / LQ) * LQ;
            const bool lstage = threadIdx.x < JR * LQ;
            lvals pl{};
            if (more && lstage) pl = load_left8(r0 + JR + li_row, cbase, lj);
            float4 pr = make_float4(0.f, 0.f, 0.f, 0.f);
            if (more && threadIdx.x < JR * RQ)
                pr = load_right(r0 + JR + (int)threadIdx.x / RQ - dyb, c0, (int)threadIdx.x % RQ);

#pragma unroll
            for (int ri = 0; ri < JR; ++ri) {
            const int r = r0 + ri;
            if (r >= r_hi) break;
            // Sub-tiles (32 dy x 32 dx) this row touches: left row r meets dy sub-tile u,
            // chunk [c0+qq, c0+qq+16) meets the columns of dx sub-tile s.
            bool act_u[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int lo = dyw + 32 * u, hi = min(lo + 31, H - 1);
                act_u[u] = lo <= hi && r >= max(0, lo) && r < min(H, H + hi);
            }
            if (q_lo < q_hi && (act_u[0] || act_u[1])) {
                const char* lb = lbuf + (cur * JR + ri) * G::LBUF;
                // Ring rows of this lane for the two dy sub-tiles: row = r - (dyw + 32u + li32).
                const char* arow[2];
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    arow[u] = ring + ring_slot(r - (dyw + 32 * u + li32), G::NSLOT) * G::SS + g * 16;
                const char* bcol[2];
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int p = 8 * g + (dxs[s] - dxb) + li32;  // + qq (a multiple of 16)
                    const int t = p & 3;
                    bcol[s] = lb + t * G::CS + (p - t) * 2;
                }
                // Split the chunk range where the set of active dx sub-tiles changes.  Sub-tile 1
                // lies at larger dx than sub-tile 0, so its column range starts and ends no later.
                const int umask = (act_u[0] ? 3 : 0) | (act_u[1] ? 12 : 0);
                const int k0lo = ks_lo[0], k0hi = ks_hi[0], k1lo = ks_lo[1], k1hi = ks_hi[1];
                const bool e0 = k0lo >= k0hi;  // sub-tile 0 has no columns in this pass
                for (int seg = 0; seg < 3; ++seg) {
                    int q0, q1, mask;
                    if (seg == 0) {  // sub-tile 1 only
                        q0 = k1lo;
                        q1 = e0 ? k1hi : min(k1hi, k0lo);
                        mask = umask & 10;
                    } else if (seg == 1) {  // both
                        q0 = e0 ? 0 : max(k0lo, k1lo);
                        q1 = e0 ? 0 : min(k0hi, k1hi);
                        mask = umask;
                    } else {  // sub-tile 0 only
                        q0 = e0 ? 0 : max(k0lo, k1hi);
                        q1 = e0 ? 0 : k0hi;
                        mask = umask & 5;
                    }
                    if (q0 < q1 && mask)
                        chunk_range_dispatch<KIND, P, PIPE>(mask, acc, arow, bcol, q0, q1, G::RS, G::CS);
                }
            }
            }  // rows of the step

            // Every FLUSH steps (and at the end of a pass) the step sums are added to the
            // totals with a compensated (Kahan) add.
            if (((r0 - r_lo + JR) % FLUSH) == 0 || !more) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        if constexpr (KAHAN) {
                            const f32x16 yv = acc[u][s] - comp[u][s];
                            const f32x16 t = tot[u][s] + yv;
                            comp[u][s] = (t - tot[u][s]) - yv;
                            tot[u][s] = t;
                        } else {
                            tot[u][s] += acc[u][s];
                        }
#pragma unroll
                        for (int e = 0; e < 16; ++e) acc[u][s][e] = 0.f;
                    }
            }

            if (more) {
                if (lstage) put_left8((cur ^ 1) * JR + li_row, lj, pl);
                if (threadIdx.x < JR * RQ)
                    put_right(r0 + JR + (int)threadIdx.x / RQ - dyb, (int)threadIdx.x % RQ, pr);
            }
            __syncthreads();
        }
    }

    // D layout of v_mfma_f32_32x32x16: lane -> column (dx), register a -> row
    // 8*(a/4) + 4*(lane/32) + a%4 (dy).
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int dx = dxs[s] + li32;
        if (dx > W - 1) continue;
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < 16; ++a) {
                const int dy = dyw + 32 * u + 8 * (a >> 2) + 4 * g + (a & 3);
                if (dy <= H - 1 &&
                    (uint32_t)(H - abs(dy)) * (uint32_t)(W - abs(dx)) >= skip_area)
                    Op[(size_t)(dy + H - 1) * SX + (dx + W - 1)] = tot[u][s][a];
            }
    }
}

template <int KIND, int WX, int WY, int KC, int FLUSH, bool KAHAN, int OCC, int JR>
__global__ __launch_bounds__(256, OCC) void mfma_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                      float* __restrict__ O, int family, uint32_t rows,
                                                      uint32_t cols, uint32_t n, uint32_t m, uint32_t tiles_x,
                                                      uint32_t tiles_y, uint32_t skip_area) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int H = (int)rows, W = (int)cols;
    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = bid % tiles_x;
    bid /= tiles_x;
    const uint32_t ty = bid % tiles_y;
    const uint32_t k = bid / tiles_y;
    uint32_t li, ri;
    pair_of(family, n, m, k, li, ri);
    mfma_tile<KIND, WX, WY, KC, FLUSH, KAHAN, OCC, JR>(L + (size_t)li * H * W, R + (size_t)ri * H * W,
                                                   O + (size_t)k * (2 * W - 1) * (2 * H - 1), H, W, tx, ty,
                                                   skip_area, smem);
}

// Row-split block tile: the block's 4 waves compute the same 64 dy x 32*NS dx outputs (2 x NS
// sub-tiles of 32 x 32 each), each wave over its own left rows (wave w takes rows r_lo + w,
// r_lo + w + 4, ...).  Every wave then has the same column ranges and nearly the same work per
// step (in mfma_tile the waves own different dx and their column ranges differ, so the faster
// ones wait at each barrier).  The waves' totals are added through LDS at the end, each of the
// first 2*NS waves finishing one sub-tile.  Staging as mfma_tile with four left rows per step.
template <int NS, int PLANES, int KC>
struct rgeom {
    static constexpr int NW = 4;                          // waves = left rows per step
    static constexpr int BX = 32 * NS, BY = 64;
    static constexpr int NSLOT = BY + 2 * NW - 1;
    static constexpr int RS = KC * 2 + 16;
    static constexpr int LENA = KC + BX + 64;
    static constexpr int CS = ((LENA * 2 + 127) / 128) * 128 + 32;
    static constexpr int SS = ((PLANES * RS / 16) % 2) ? PLANES * RS : PLANES * RS + 16;
    static constexpr int RING = NSLOT * SS;
    static constexpr int LBUF = PLANES * 4 * CS;
    static constexpr int STAGE = RING + 2 * NW * LBUF;
    static constexpr int RED = NW * 2 * NS * 16 * 64 * 4;  // the waves' totals, for the final sum
    static constexpr int BYTES = STAGE > RED ? STAGE : RED;
};

// MFMAs of one left-row step of a one-sub-tile-wide (NS = 1) row-split wave over chunks
// [q0, q1): UM bit u selects dy sub-tile u.
template <int KIND, int P, int UM>
__device__ __forceinline__ void chunk_range1(f32x16 (&acc)[2][1], const char* const (&arow)[2], const char* bcol,
                                             int q0, int q1, int rs, int cs) {
    for (int qq = q0; qq < q1; qq += 16) {
        uint4 a[2][P], b[P];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int pl = 0; pl < P; ++pl)
                if (UM & (1 << u)) a[u][pl] = *reinterpret_cast<const uint4*>(arow[u] + pl * rs + qq * 2);
#pragma unroll
        for (int pl = 0; pl < P; ++pl) {
            const uint2* q = reinterpret_cast<const uint2*>(bcol + pl * 4 * cs + qq * 2);
            const uint2 v0 = q[0], v1 = q[1];
            b[pl] = make_uint4(v0.x, v0.y, v1.x, v1.y);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (UM & (1 << u)) acc[u][0] = mma<KIND>(a[u][0], b[0], acc[u][0]);
        if constexpr (P == 2) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if (UM & (1 << u)) acc[u][0] = mma<KIND>(a[u][0], b[1], acc[u][0]);
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if (UM & (1 << u)) acc[u][0] = mma<KIND>(a[u][1], b[0], acc[u][0]);
        }
    }
}

template <int KIND, int NS, int KC, int FLUSH, int OCC, bool PIPE_ROWS = (OCC == 1)>
__device__ __forceinline__ void mfma_rows_tile(const float* __restrict__ Lp, const float* __restrict__ Rp,
                                               float* __restrict__ Op, const int H, const int W, const uint32_t tx,
                                               const uint32_t ty, const uint32_t skip_area, char* smem) {
    using S = split_traits<KIND>;
    constexpr int P = S::planes;
    using G = rgeom<NS, P, KC>;
    constexpr int NW = G::NW;
    constexpr bool PIPE = PIPE_ROWS;
    char* ring = smem;
    char* lbuf = smem + G::RING;
    const int SX = 2 * W - 1;
    const int dxb = (int)tx * G::BX - (W - 1);
    const int dyb = (int)ty * G::BY - (H - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int li32 = lane & 31, g = lane >> 5;
    const int dxs[2] = {dxb, dxb + 32};  // sub-tile 1 only where NS == 2
    const int dyw = dyb;

    const int dy_hi_b = min(dyb + G::BY - 1, H - 1);
    const int r_lo = max(0, dyb);
    const int r_hi = min(H, H + dy_hi_b);
    const int dx_hi_b = min(dxb + G::BX - 1, W - 1);
    const int cb_lo = max(0, -dx_hi_b), cb_hi = min(W, W - dxb);

    f32x16 tot[2][NS], acc[2][NS];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int e = 0; e < 16; ++e) tot[u][s][e] = acc[u][s][e] = 0.f;

    static_assert(G::LENA % 4 == 0 && KC % 4 == 0, "quad staging");
    constexpr int LQ = G::LENA / 4;  // quads per left copy
    constexpr int RQ = KC / 4;
    static_assert(NW * RQ <= 256, "one thread per staged right-row quad");
    static_assert(NW * LQ <= 256, "one thread per staged (left row, quad)");
    // Left rows: thread (row i, quad j) loads the 8 values at staged positions 4j .. 4j+7
    // (columns cbase + 4j ..), splits them once and writes quad j of all four shifted copies
    // (copy t holds positions 4j + t .. 4j + t + 3), the odd shifts by funnel shifts of the
    // packed pairs.
    struct lvals {
        float v[8];
    };
    auto load_left8 = [&](int r, int cbase, int j) -> lvals {
        lvals x;
        const int c = cbase + 4 * j;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            x.v[k] = ((unsigned)r < (unsigned)H && (unsigned)(c + k) < (unsigned)W) ? Lp[(size_t)r * W + c + k] : 0.f;
        return x;
    };
    auto load_right = [&](int row, int c0, int j) -> float4 {
        const int c = c0 + 4 * j;
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            v[k] = ((unsigned)row < (unsigned)H && c + k < W) ? Rp[(size_t)row * W + c + k] : 0.f;
        return make_float4(v[0], v[1], v[2], v[3]);
    };
    auto split4 = [&](float4 v, uint2& hi, uint2& lo) {
        uint16_t h[4], l[4];
        S::split(v.x, h[0], l[0]);
        S::split(v.y, h[1], l[1]);
        S::split(v.z, h[2], l[2]);
        S::split(v.w, h[3], l[3]);
        hi = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
        lo = make_uint2(l[0] | ((uint32_t)l[1] << 16), l[2] | ((uint32_t)l[3] << 16));
    };
    auto put_left8 = [&](int buf, int j, const lvals& x) {
        uint2 h0, l0, h1, l1;
        split4(make_float4(x.v[0], x.v[1], x.v[2], x.v[3]), h0, l0);
        split4(make_float4(x.v[4], x.v[5], x.v[6], x.v[7]), h1, l1);
        auto shifted = [](uint2 a, uint2 b, int t) -> uint2 {  // elements t .. t+3 of (a, b)
            const uint32_t w0 = a.x, w1 = a.y, w2 = b.x, w3 = b.y;
            if (t == 0) return make_uint2(w0, w1);
            if (t == 1) return make_uint2((w0 >> 16) | (w1 << 16), (w1 >> 16) | (w2 << 16));
            if (t == 2) return make_uint2(w1, w2);
            return make_uint2((w1 >> 16) | (w2 << 16), (w2 >> 16) | (w3 << 16));
        };
        char* base = lbuf + buf * G::LBUF + j * 8;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            *reinterpret_cast<uint2*>(base + t * G::CS) = shifted(h0, h1, t);
            if (P == 2) *reinterpret_cast<uint2*>(base + (4 + t) * G::CS) = shifted(l0, l1, t);
        }
    };
    auto put_right = [&](int row, int j, float4 v) {
        uint2 hi, lo;
        split4(v, hi, lo);
        char* base = ring + ring_slot(row, G::NSLOT) * G::SS + j * 8;
        *reinterpret_cast<uint2*>(base) = hi;
        if (P == 2) *reinterpret_cast<uint2*>(base + G::RS) = lo;
    };

    // The two dx sub-tiles' column ranges (all waves alike), as in mfma_tile.
    int cw_lo = W, cw_hi = 0;
#pragma unroll
    for (int sidx = 0; sidx < NS; ++sidx) {
        const int lo = dxs[sidx], hi = min(lo + 31, W - 1);
        if (lo <= hi) {
            cw_lo = min(cw_lo, max(0, -hi));
            cw_hi = max(cw_hi, min(W, W - lo));
        }
    }
    for (int c0 = (cb_lo / KC) * KC; c0 < cb_hi; c0 += KC) {
        const int cbase = c0 + dxb;
        // Prefill: right rows r - dy for r in [r_lo, r_lo + NW), dy in the block; left rows
        // r_lo .. r_lo + NW - 1 into buffers 0 .. NW - 1.
        for (int idx = threadIdx.x; idx < (G::BY + NW - 1) * RQ; idx += 256) {
            const int ro = idx / RQ, j = idx - (idx / RQ) * RQ;
            const int row = r_lo + NW - 1 - dyb - ro;
            put_right(row, j, load_right(row, c0, j));
        }
        if (threadIdx.x < NW * LQ) {
            const int i = threadIdx.x / LQ, j = threadIdx.x - (threadIdx.x / LQ) * LQ;
            put_left8(i, j, load_left8(r_lo + i, cbase, j));
        }
        __syncthreads();

        const int q_lo = max(cw_lo, c0), q_hi = min(cw_hi, c0 + KC);
        int ks_lo[2] = {0, 0}, ks_hi[2] = {0, 0};
#pragma unroll
        for (int sidx = 0; sidx < NS; ++sidx) {
            const int lo = dxs[sidx], hi = min(lo + 31, W - 1);
            const int a0 = max(max(0, -hi), c0) - c0, a1 = min(min(W, W - lo), c0 + KC) - c0;
            if (lo <= hi && a0 < a1) {
                ks_lo[sidx] = a0 & ~15;
                ks_hi[sidx] = (a1 + 15) & ~15;
            } else {
                ks_lo[sidx] = ks_hi[sidx] = 0;
            }
        }

        int step = 0;
        for (int r0 = r_lo; r0 < r_hi; r0 += NW, ++step) {
            const int cur = step & 1;
            const bool more = r0 + NW < r_hi;
            const int li_row = threadIdx.x / LQ, lj = threadIdx.x - (threadIdx.x / LQ) * LQ;
            const bool lstage = threadIdx.x < NW * LQ;
            lvals pl{};
            if (more && lstage) pl = load_left8(r0 + NW + li_row, cbase, lj);
            float4 pr = make_float4(0.f, 0.f, 0.f, 0.f);
            if (more && threadIdx.x < NW * RQ)
                pr = load_right(r0 + NW + (int)threadIdx.x / RQ - dyb, c0, (int)threadIdx.x % RQ);

            const int r = r0 + wave;  // this wave's left row
            if (r < r_hi) {
                bool act_u[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int lo = dyw + 32 * u, hi = min(lo + 31, H - 1);
                    act_u[u] = lo <= hi && r >= max(0, lo) && r < min(H, H + hi);
                }
                if (q_lo < q_hi && (act_u[0] || act_u[1])) {
                    const char* lb = lbuf + (cur * NW + wave) * G::LBUF;
                    const char* arow[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        arow[u] = ring + ring_slot(r - (dyw + 32 * u + li32), G::NSLOT) * G::SS + g * 16;
                    const char* bcol[2];
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int p = 8 * g + (dxs[s] - dxb) + li32;
                        const int t = p & 3;
                        bcol[s] = lb + t * G::CS + (p - t) * 2;
                    }
                    if constexpr (NS == 1) {
                        const int q0 = ks_lo[0], q1 = ks_hi[0];
                        if (q0 < q1) {
                            if (act_u[0] && act_u[1]) chunk_range1<KIND, P, 3>(acc, arow, bcol[0], q0, q1, G::RS, G::CS);
                            else if (act_u[0]) chunk_range1<KIND, P, 1>(acc, arow, bcol[0], q0, q1, G::RS, G::CS);
                            else chunk_range1<KIND, P, 2>(acc, arow, bcol[0], q0, q1, G::RS, G::CS);
                        }
                    } else {
                    const int umask = (act_u[0] ? 3 : 0) | (act_u[1] ? 12 : 0);
                    const int k0lo = ks_lo[0], k0hi = ks_hi[0], k1lo = ks_lo[1], k1hi = ks_hi[1];
                    const bool e0 = k0lo >= k0hi;
                    for (int seg = 0; seg < 3; ++seg) {
                        int q0, q1, mask;
                        if (seg == 0) {
                            q0 = k1lo;
                            q1 = e0 ? k1hi : min(k1hi, k0lo);
                            mask = umask & 10;
                        } else if (seg == 1) {
                            q0 = e0 ? 0 : max(k0lo, k1lo);
                            q1 = e0 ? 0 : min(k0hi, k1hi);
                            mask = umask;
                        } else {
                            q0 = e0 ? 0 : max(k0lo, k1hi);
                            q1 = e0 ? 0 : k0hi;
                            mask = umask & 5;
                        }
                        if (q0 < q1 && mask)
                            chunk_range_dispatch<KIND, P, PIPE>(mask, acc, arow, bcol, q0, q1, G::RS, G::CS);
                    }
                    }
                }
            }
            if (((step + 1) % FLUSH) == 0 || !more) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        tot[u][s] += acc[u][s];
#pragma unroll
                        for (int e = 0; e < 16; ++e) acc[u][s][e] = 0.f;
                    }
            }
            if (more) {
                if (lstage) put_left8((cur ^ 1) * NW + li_row, lj, pl);
                if (threadIdx.x < NW * RQ)
                    put_right(r0 + NW + (int)threadIdx.x / RQ - dyb, (int)threadIdx.x % RQ, pr);
            }
            __syncthreads();
        }
    }

    // Sum of the waves' totals: each wave writes its 2*NS sub-tiles, then wave w < 2*NS
    // finishes sub-tile w.
    constexpr int NT = 2 * NS;
    float* red = reinterpret_cast<float*>(smem);  // [wave][sub-tile][element][lane]
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int e = 0; e < 16; ++e) red[((wave * NT + NS * u + s) * 16 + e) * 64 + lane] = tot[u][s][e];
    __syncthreads();
    if (wave >= NT) return;
    const int u = wave / NS, s = wave % NS;
    const int dx = dxs[s] + li32;
    if (dx > W - 1) return;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < NW; ++w) v += red[((w * NT + wave) * 16 + a) * 64 + lane];
        // D layout of v_mfma_f32_32x32x16: lane -> column (dx), register a -> row
        // 8*(a/4) + 4*(lane/32) + a%4 (dy).
        const int dy = dyw + 32 * u + 8 * (a >> 2) + 4 * g + (a & 3);
        if (dy <= H - 1 && (uint32_t)(H - abs(dy)) * (uint32_t)(W - abs(dx)) >= skip_area)
            Op[(size_t)(dy + H - 1) * SX + (dx + W - 1)] = v;
    }
}

template <int KIND, int NS, int KC, int FLUSH, int OCC, bool PIPE_ROWS>
__global__ __launch_bounds__(256, OCC) void mfma_rows_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                           float* __restrict__ O, int family, uint32_t rows,
                                                           uint32_t cols, uint32_t n, uint32_t m, uint32_t tiles_x,
                                                           uint32_t tiles_y, uint32_t skip_area) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int H = (int)rows, W = (int)cols;
    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tx = bid % tiles_x;
    bid /= tiles_x;
    const uint32_t ty = bid % tiles_y;
    const uint32_t k = bid / tiles_y;
    uint32_t li, ri;
    pair_of(family, n, m, k, li, ri);
    mfma_rows_tile<KIND, NS, KC, FLUSH, OCC, PIPE_ROWS>(L + (size_t)li * H * W, R + (size_t)ri * H * W,
                                         O + (size_t)k * (2 * W - 1) * (2 * H - 1), H, W, tx, ty, skip_area, smem);
}

// Value of one output (dy, dx) with overlap h x w, computed by one thread: products of FP32
// values are exact in FP64 and their FP64 sum carries a relative error below h*w*2^-53, so the
// result is the exact sum rounded once to FP32.
__device__ __forceinline__ float frame_value(const float* __restrict__ Lp, const float* __restrict__ Rp, int W,
                                             int dy, int dx, int h, int w) {
    const int r0 = max(0, dy), c0 = max(0, dx);
    const float* __restrict__ lp = Lp + (size_t)r0 * W + c0;
    const float* __restrict__ rp = Rp + (size_t)(r0 - dy) * W + (c0 - dx);
    double acc = 0.0;
    for (int r = 0; r < h; ++r, lp += W, rp += W)
        for (int c = 0; c < w; ++c) acc = fma((double)lp[c], (double)rp[c], acc);
    return (float)acc;
}

// (h, w) of frame index hw from a table with start[0..nh] (largest h with start[h-1] <= hw).
__device__ __forceinline__ void frame_hw(const uint32_t* start, uint32_t nh, uint32_t hw, int& h, int& w) {
    int lo = 0, hi = (int)nh - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= hw) lo = mid;
        else hi = mid - 1;
    }
    h = lo + 1;
    w = (int)(hw - start[lo]) + 1;
}

// Exact recomputation of the outputs whose overlap area h*w (h = H-|dy|, w = W-|dx|) is
// below the threshold T.  Every output reads as
//   out = sum_{i<h} sum_{c<w} L[rl+i][cl+c] * R[rr+i][cr+c],
// (rl, rr) = (H-h, 0) where dy >= 0, (0, H-h) where dy < 0; (cl, cr) likewise from dx and w.
// The frame is cut into tasks of up to kFrameGroup outputs that share one loop: a fixed h
// and consecutive widths w = whi-k.  A (the matrix whose column offset moves with w: L
// where dx >= 0, else R) is read from column s0 = W-whi on; output k takes
// A[ra+i][s0+k+c] * B[rb+i][c] over c < ceil(whi/G)*G, A being zero past its last column,
// so each output's sum ends at its own width.  Heights up to Kh (the largest h with h*h < T)
// run this way on row-major copies of the matrices; taller outputs (w < T/Kh) run the same
// code on column-major copies, where they are short and wide.  No task runs more than about
// T/G + sqrt(T) steps.
// A step (row i, group v of G columns) loads 4 quads of A (from column s0 + G*v, which the
// table keeps a multiple of 4 by aligning each group's top width) and 2 quads of B, and does
// G*G FP64 FMAs, with the next step's loads in flight.
// The copies hold quads of 4 consecutive elements of a row, interleaved over the matrices
// (quad q of row r of matrix i at ((r * Q + q) * count + i) * 4), zero past the row end, so
// the lanes, which take consecutive pairs, read 16 consecutive bytes each from one contiguous
// range, and nothing needs a bounds test.  The lanes of a wave run one task for LP
// consecutive pairs (LP = 64: the task is uniform over the wave and its addressing scalar;
// fewer pairs: 64/LP tasks per wave).
// Products of FP32 values are exact in FP64, so each output is its exact sum rounded once.
constexpr int kFrameGroup = 8;
constexpr int kFrameEntries = 48;  // at most 2 * 22 entries for T <= 512

// Padded row length of the copies: the A window of the last group ends at most 14 columns
// past the last column; a multiple of 4.
__host__ __device__ constexpr uint32_t frame_padded(uint32_t len) { return (len + 15 + 3) / 4 * 4; }

// Task table: entry e owns tasks [start[e], start[e+1]) of each quadrant; fixed[e] is the
// fixed height (bit 31 set: on the column-major copies, the fixed width), the entry's widths
// (heights) are vlo[e] .. vmax[e], and its first task's group ends at top[e] (the smallest
// value >= vmax congruent to the row length mod 4); each next task's group ends G lower.
struct frame_plan {
    uint32_t ne;
    uint32_t start[kFrameEntries + 1];
    uint32_t fixed[kFrameEntries];
    uint32_t vlo[kFrameEntries];
    uint32_t vmax[kFrameEntries];
    uint32_t top[kFrameEntries];
};

// The two quad copies of count matrices of rows x cols: row-major Xn (rows x Wp, Wp =
// frame_padded(cols)) and column-major Xc (cols x Hp), zero past each row's end.  A block
// stages 16 matrices x 16 x 16 elements in LDS and writes 16-byte quads, the 16 matrices of
// a quad position contiguous.
__global__ __launch_bounds__(256) void frame_copy_kernel(const float* __restrict__ X, float* __restrict__ Xn,
                                                         float* __restrict__ Xc, uint32_t count, uint32_t rows,
                                                         uint32_t cols, uint32_t tiles_c) {
    __shared__ float tile[16][16][17];
    const uint32_t Wp = frame_padded(cols), Hp = frame_padded(rows);
    const uint32_t r0 = (blockIdx.x / tiles_c) * 16, c0 = (blockIdx.x % tiles_c) * 16;
    const uint32_t i0 = blockIdx.y * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t e = threadIdx.x + 256 * k;
        const uint32_t mt = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
        const uint32_t i = i0 + mt, r = r0 + rr, c = c0 + cc;
        tile[mt][rr][cc] = (i < count && r < rows && c < cols) ? X[((size_t)i * rows + r) * cols + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t e = threadIdx.x + 256 * k;
        const uint32_t mt = e & 15, qd = (e >> 4) & 3, rr = e >> 6;  // matrix, quad, row
        const uint32_t i = i0 + mt;
        if (i >= count) continue;
        const uint32_t r = r0 + rr, c = c0 + 4 * qd;  // row-major quad (r, c .. c+3)
        if (r < rows && c < Wp) {
            const float4 v = make_float4(tile[mt][rr][4 * qd], tile[mt][rr][4 * qd + 1], tile[mt][rr][4 * qd + 2],
                                         tile[mt][rr][4 * qd + 3]);
            *reinterpret_cast<float4*>(Xn + (((size_t)r * (Wp / 4) + c / 4) * count + i) * 4) = v;
        }
        const uint32_t cr = c0 + rr, rq = r0 + 4 * qd;  // column-major quad (column cr, rows rq .. rq+3)
        if (cr < cols && rq < Hp) {
            const float4 v = make_float4(tile[mt][4 * qd][rr], tile[mt][4 * qd + 1][rr], tile[mt][4 * qd + 2][rr],
                                         tile[mt][4 * qd + 3][rr]);
            *reinterpret_cast<float4*>(Xc + (((size_t)cr * (Hp / 4) + rq / 4) * count + i) * 4) = v;
        }
    }
}

template <bool WAVE>
__device__ __forceinline__ uint32_t frame_uniform(uint32_t v) {
    if constexpr (WAVE) return __builtin_amdgcn_readfirstlane(v);
    else return v;
}

// The quad at base + lane_bytes + uniform_bytes: with a wave-uniform base (WAVE) a raw
// buffer load (buffer_quad), else a plain per-lane load.  The host keeps the offsets below
// 2^31.
template <bool WAVE>
__device__ __forceinline__ void frame_load4(const float* __restrict__ base, uint32_t lane_bytes,
                                            uint32_t uniform_bytes, float* v) {
    if constexpr (WAVE) {
        const float4 q = buffer_quad(base, lane_bytes, uniform_bytes);
        v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
    } else {
        const float4 q =
            *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + lane_bytes + uniform_bytes);
        v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
    }
}

// The same quad through the scalar cache, for an address that is uniform over the wave.
__device__ __forceinline__ void frame_sload4(const float* __restrict__ p, float* v) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
}

// MODE 0: a task per LP < 64 lanes; 1: a task per wave; 2: a task per wave whose 64 pairs
// share one left matrix, whose quads are then scalar loads.
template <int MODE>
__global__ __launch_bounds__(256) void frame_kernel(const float* __restrict__ Ln, const float* __restrict__ Rn,
                                                    const float* __restrict__ Lc, const float* __restrict__ Rc,
                                                    float* __restrict__ O, int family, uint32_t rows,
                                                    uint32_t cols, uint32_t n, uint32_t m, uint32_t nl_mats,
                                                    uint32_t nr_mats, uint32_t pairs, uint32_t lp_shift,
                                                    frame_plan tab) {
    constexpr int G = kFrameGroup;
    constexpr bool WAVE = MODE > 0;
    const uint32_t per_quad = tab.start[tab.ne];
    const uint32_t slot = frame_uniform<WAVE>((blockIdx.x * blockDim.x + threadIdx.x) >> lp_shift);
    if (slot >= 4 * per_quad) return;
    const uint32_t q = slot / per_quad, t = slot - q * per_quad;
    int lo = 0, hi = (int)tab.ne - 1;  // entry: the largest with start[.] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab.start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t e = frame_uniform<WAVE>((uint32_t)lo);
    const bool tr = tab.fixed[e] >> 31;  // on the column-major copies
    const int h = (int)(tab.fixed[e] & 0x7fffffffu);
    const int whi = (int)(tab.top[e] - (t - tab.start[e]) * G);
    const int vlo = (int)tab.vlo[e], vmax = (int)tab.vmax[e];
    // the task's own frame: rows Hd x columns Wd, and the signs of its row and column shifts
    const int Hd = tr ? (int)cols : (int)rows, Wd = tr ? (int)rows : (int)cols;
    const bool rneg = tr ? (q & 2) : (q & 1), cneg = tr ? (q & 1) : (q & 2);
    if (rneg && h == Hd) return;  // shift 0 belongs to the non-negative quadrants
    const bool a_left = !cneg;
    const int rl = rneg ? 0 : Hd - h, rr = rneg ? Hd - h : 0;
    const int ra = a_left ? rl : rr, rb = a_left ? rr : rl;
    const int s0 = Wd - whi;  // a multiple of 4
    const int Qd = (int)frame_padded((uint32_t)Wd) / 4;  // quads per row of the copy
    const float* __restrict__ Ab = a_left ? (tr ? Lc : Ln) : (tr ? Rc : Rn);
    const float* __restrict__ Bb = a_left ? (tr ? Rc : Rn) : (tr ? Lc : Ln);
    const int64_t sa = a_left ? nl_mats : nr_mats, sb = a_left ? nr_mats : nl_mats;  // matrices per quad position
    const int ng = (whi + G - 1) / G;
    const int nsteps = h * ng;
    const int SY = 2 * (int)rows - 1, SX = 2 * (int)cols - 1;
    const uint32_t lane_mask = (1u << lp_shift) - 1;
    const uint32_t nchunks = ((pairs - 1) >> lp_shift) + 1;
    for (uint32_t chunk = blockIdx.y; chunk < nchunks; chunk += gridDim.y) {
        const uint32_t pk = (chunk << lp_shift) + (threadIdx.x & lane_mask);
        const bool valid = pk < pairs;  // lanes past the last pair run the last one and store nothing
        uint32_t li, ri;
        pair_of(family, n, m, valid ? pk : pairs - 1, li, ri);
        if constexpr (MODE == 2) li = __builtin_amdgcn_readfirstlane(li);
        const uint32_t la = 16 * (a_left ? li : ri), lb = 16 * (a_left ? ri : li);  // lane byte offsets
        double acc[G];
#pragma unroll
        for (int k = 0; k < G; ++k) acc[k] = 0.0;
        auto sum = [&](const float (&va)[2 * G], const float (&vb)[G]) {
            double da[2 * G - 1];  // A value 2G-1 is never used
#pragma unroll
            for (int j = 0; j < 2 * G - 1; ++j) da[j] = (double)va[j];
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const double bv = (double)vb[j];
#pragma unroll
                for (int k = 0; k < G; ++k) acc[k] = fma(da[j + k], bv, acc[k]);
            }
        };
        // SA, SB: A, B is the shared left matrix, read by scalar loads
        auto run = [&](auto sa_c, auto sb_c) {
            constexpr bool SA = decltype(sa_c)::value, SB = decltype(sb_c)::value;
            // quads of step (row, grp): A row ra+row from column s0 + G*grp, B row rb+row from G*grp
            auto fetch = [&](int row, int grp, float (&va)[2 * G], float (&vb)[G]) {
                const float* __restrict__ ap = Ab + ((int64_t)(ra + row) * Qd + (s0 + G * grp) / 4) * sa * 4;
                const float* __restrict__ bp = Bb + ((int64_t)(rb + row) * Qd + G * grp / 4) * sb * 4;
#pragma unroll
                for (int j = 0; j < G / 2; ++j) {
                    if constexpr (SA) frame_sload4(ap + (j * sa + la / 16) * 4, va + 4 * j);
                    else frame_load4<WAVE>(ap, la, (uint32_t)(16 * j * sa), va + 4 * j);
                }
#pragma unroll
                for (int j = 0; j < G / 4; ++j) {
                    if constexpr (SB) frame_sload4(bp + (j * sb + lb / 16) * 4, vb + 4 * j);
                    else frame_load4<WAVE>(bp, lb, (uint32_t)(16 * j * sb), vb + 4 * j);
                }
            };
            // two register sets in turn: each step's successor is loaded while it is summed
            float xa[2 * G], xb[G], ya[2 * G], yb[G];
            int i = 0, v = 0;
            auto next = [&]() {
                if (++v == ng) {
                    v = 0;
                    ++i;
                }
            };
            fetch(0, 0, xa, xb);
            for (int st = 0; st < nsteps; st += 2) {
                next();
                if (st + 1 < nsteps) fetch(i, v, ya, yb);
                sum(xa, xb);
                if (st + 1 == nsteps) break;
                next();
                if (st + 2 < nsteps) fetch(i, v, xa, xb);
                sum(ya, yb);
            }
        };
        if constexpr (MODE == 2) {
            if (a_left) run(std::true_type{}, std::false_type{});
            else run(std::false_type{}, std::true_type{});
        } else {
            run(std::false_type{}, std::false_type{});
        }
        if (valid) {
            float* __restrict__ op = O + (size_t)pk * SX * SY;
            const int dr = rneg ? -(Hd - h) : Hd - h;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const int x = whi - k;  // the output's width in the task's frame
                if (x < vlo || x > vmax || (cneg && x == Wd)) continue;
                const int dc = cneg ? -(Wd - x) : Wd - x;
                const int dy = tr ? dc : dr, dx = tr ? dr : dc;
                op[(size_t)(dy + (int)rows - 1) * SX + (dx + (int)cols - 1)] = (float)acc[k];
            }
        }
    }
}

template <int KIND, int WX, int WY, int KC, int FLUSH, bool KAHAN, int OCC, int JR = 1>
void launch(const family_args& a, const float* L, const float* R, float* O, uint32_t skip_area, hipStream_t s) {
    using G = geom<WX, WY, split_traits<KIND>::planes, KC, JR>;
    const uint32_t SX = 2 * a.cols - 1, SY = 2 * a.rows - 1;
    const uint32_t tiles_x = (SX + G::BX - 1) / G::BX;
    const uint32_t tiles_y = (SY + G::BY - 1) / G::BY;
    const uint64_t nblocks = (uint64_t)tiles_x * tiles_y * a.pairs;
    if (nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    allow_dynamic_lds(reinterpret_cast<const void*>(&mfma_kernel<KIND, WX, WY, KC, FLUSH, KAHAN, OCC, JR>), G::BYTES);
    dim3 grid((uint32_t)nblocks), block(256);
    record_launch(KIND == XC_PREC_BF16X3   ? "mfma_kernel<bf16x3>"
                  : KIND == XC_PREC_FP16X3 ? "mfma_kernel<fp16x3>"
                  : KIND == XC_PREC_BF16   ? "mfma_kernel<bf16>"
                                           : "mfma_kernel<fp16>",
                  grid, block, G::BYTES, a.rows, a.cols, a.pairs, 1);
    timed tm("mfma_kernel", s);
    hipLaunchKernelGGL((mfma_kernel<KIND, WX, WY, KC, FLUSH, KAHAN, OCC, JR>), grid, block, G::BYTES, s, L, R, O, a.family, a.rows, a.cols, a.n,
                       a.m, tiles_x, tiles_y, skip_area);
    XC_HIP(hipGetLastError());
}

template <int KIND, int NS, int KC, int FLUSH, int OCC, bool PIPE_ROWS = (OCC == 1)>
void launch_rows(const family_args& a, const float* L, const float* R, float* O, uint32_t skip_area, hipStream_t s) {
    using G = rgeom<NS, split_traits<KIND>::planes, KC>;
    const uint32_t SX = 2 * a.cols - 1, SY = 2 * a.rows - 1;
    const uint32_t tiles_x = (SX + G::BX - 1) / G::BX;
    const uint32_t tiles_y = (SY + G::BY - 1) / G::BY;
    const uint64_t nblocks = (uint64_t)tiles_x * tiles_y * a.pairs;
    if (nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    allow_dynamic_lds(reinterpret_cast<const void*>(&mfma_rows_kernel<KIND, NS, KC, FLUSH, OCC, PIPE_ROWS>), G::BYTES);
    dim3 grid((uint32_t)nblocks), block(256);
    record_launch(KIND == XC_PREC_BF16X3   ? "mfma_rows_kernel<bf16x3>"
                  : KIND == XC_PREC_FP16X3 ? "mfma_rows_kernel<fp16x3>"
                  : KIND == XC_PREC_BF16   ? "mfma_rows_kernel<bf16>"
                                           : "mfma_rows_kernel<fp16>",
                  grid, block, G::BYTES, a.rows, a.cols, a.pairs, 1);
    timed tm("mfma_kernel", s);
    hipLaunchKernelGGL((mfma_rows_kernel<KIND, NS, KC, FLUSH, OCC, PIPE_ROWS>), grid, block, G::BYTES, s, L, R, O, a.family, a.rows,
                       a.cols, a.n, a.m, tiles_x, tiles_y, skip_area);
    XC_HIP(hipGetLastError());
}

}  // namespace

void launch_frame_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t thresh, hipStream_t s) {
    if (a.pairs == 0 || thresh <= 1) return;
    if (thresh > (uint32_t)kMaxBorder) fail(XC_ERR_INVALID_ARGUMENT, "border threshold above " + std::to_string(kMaxBorder));
    // Task table: heights 1 .. Kh (the largest h with h * h < T, at most rows) on the
    // row-major copies, then the widths that have frame outputs of height above Kh on the
    // column-major copies.  Each entry's groups of widths (heights) end at top, top - G, ...
    frame_plan tab{};
    uint32_t total = 0, ne = 0;
    auto add = [&](bool column_major, uint32_t fixed, uint32_t vlo, uint32_t vmax, uint32_t row_len) {
        if (ne == (uint32_t)kFrameEntries) fail(XC_ERR_INVALID_ARGUMENT, "border frame: task table overflow");
        const uint32_t top = vmax + (row_len - vmax) % 4;  // <= row_len, row_len - top a multiple of 4
        tab.start[ne] = total;
        tab.fixed[ne] = fixed | (column_major ? 0x80000000u : 0u);
        tab.vlo[ne] = vlo;
        tab.vmax[ne] = vmax;
        tab.top[ne] = top;
        ++ne;
        total += (top - vlo + kFrameGroup) / kFrameGroup;
    };
    uint32_t kh = 0;
    while ((kh + 1) * (kh + 1) < thresh) ++kh;
    kh = std::min(kh, a.rows);
    for (uint32_t h = 1; h <= kh; ++h) add(false, h, 1, std::min<uint32_t>(a.cols, (thresh + h - 1) / h - 1), a.cols);
    for (uint32_t w = 1; w <= a.cols; ++w) {
        const uint32_t hmax = std::min<uint32_t>(a.rows, (thresh + w - 1) / w - 1);
        if (hmax <= kh) break;
        add(true, w, kh + 1, hmax, a.rows);
    }
    if (ne == 0) return;
    tab.ne = ne;
    tab.start[ne] = total;
    // Row-major and column-major quad copies of the left and right matrices.
    const uint32_t nl = a.family == XC_ONE_TO_ONE || a.family == XC_ONE_TO_MANY ? 1 : a.n;
    const uint32_t nr = a.family == XC_ONE_TO_ONE ? 1 : a.m;
    const uint32_t Wp = frame_padded(a.cols), Hp = frame_padded(a.rows);
    const uint64_t per_mat = (uint64_t)a.rows * Wp + (uint64_t)a.cols * Hp;
    float* next = static_cast<float*>(workspace(WS_FRAME, sizeof(float) * per_mat * (nl + nr), s));
    auto copy = [&](const float* X, uint32_t count, const float*& Xn, const float*& Xc) {
        float* xn = next;
        float* xc = next + (uint64_t)a.rows * Wp * count;
        next += per_mat * count;
        const uint64_t tiles_c = (Wp + 15) / 16, tiles = (uint64_t)((Hp + 15) / 16) * tiles_c;
        const uint64_t gy = (count + 15) / 16;
        if (tiles > 0x7fffffffULL || gy > 65535) fail(XC_ERR_INVALID_ARGUMENT, "border frame: matrices too large to copy");
        hipLaunchKernelGGL(frame_copy_kernel, dim3((uint32_t)tiles, (uint32_t)gy), dim3(256), 0, s, X, xn, xc, count,
                           a.rows, a.cols, (uint32_t)tiles_c);
        XC_HIP(hipGetLastError());
        Xn = xn, Xc = xc;
    };
    const float *Ln, *Lc, *Rn, *Rc;
    copy(L, nl, Ln, Lc);
    copy(R, nr, Rn, Rc);
    uint32_t lp_shift = 0;  // lanes per task: the power of two >= pairs, at most 64
    while (lp_shift < 6 && (1u << lp_shift) < a.pairs) ++lp_shift;
    const uint64_t lanes = (uint64_t)4 * total << lp_shift;
    const uint64_t chunks = ((uint64_t)a.pairs + (1u << lp_shift) - 1) >> lp_shift;
    if ((lanes + 255) / 256 > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "border frame too large for one launch");
    dim3 grid((uint32_t)((lanes + 255) / 256), (uint32_t)std::min<uint64_t>(chunks, 65535));
    // One task per wave with scalar addressing when 64 pairs share each task and the buffer
    // loads' offsets (lane: 16 bytes per matrix; uniform: up to 3 quad positions) fit 31 bits.
    const uint64_t most = std::max(nl, nr);
    const bool wave_tasks = lp_shift == 6 && 16 * 4 * most < (1ull << 31);
    // The 64 pairs of a wave share their left matrix (pairs k = 64c .. 64c+63).
    const bool shared_left = a.family == XC_ONE_TO_MANY || (a.family == XC_N_TO_M && a.m % 64 == 0) ||
                             (a.family == XC_N_TO_MN && a.n == 1);
    timed tf("frame_kernel", s);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, Ln, Rn, Lc, Rc, O, a.family, a.rows, a.cols, a.n, a.m, nl, nr,
                           a.pairs, lp_shift, tab);
    };
    if (wave_tasks && shared_left) go(frame_kernel<2>);
    else if (wave_tasks) go(frame_kernel<1>);
    else go(frame_kernel<0>);
    XC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- uneven batch
// Persistent kernel over a work list: MFMA block tiles of pairs of any size, plus (bf16x3)
// chunks of 256 border-frame outputs.  Tiles skip the frame outputs, so the items write
// disjoint outputs and need no ordering.  Blocks take items from an atomic counter; the last
// block to finish resets the counters, so the launch can be replayed (graph) as is.
namespace {

constexpr int kBWX = 4, kBWY = 1, kBKC = 128, kBFLUSH = 8, kBOCC = 2;
constexpr bool kBKAHAN = false;
constexpr uint32_t kFrameChunkOutputs = 256;

template <int KIND>
__global__ __launch_bounds__(256, kBOCC) void batch_kernel(const batch_item* __restrict__ items, uint32_t n_items,
                                                          const batch_pair_desc* __restrict__ pairs,
                                                          const uint32_t* __restrict__ tables,
                                                          const float* __restrict__ L, const float* __restrict__ R,
                                                          float* __restrict__ O, uint32_t* counters,
                                                          uint32_t thresh) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t s_item;
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(&counters[0], 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const batch_item bi = items[it];
        const batch_pair_desc pd = pairs[bi.pair];
        const int H = (int)pd.rows, W = (int)pd.cols;
        const float* Lp = L + pd.loff;
        const float* Rp = R + pd.roff;
        float* Op = O + pd.ooff;
        if (bi.kind == 0) {
            mfma_tile<KIND, kBWX, kBWY, kBKC, kBFLUSH, kBKAHAN, kBOCC>(Lp, Rp, Op, H, W, bi.tx, bi.b,
                                                                      split_traits<KIND>::planes == 2 ? thresh : 0,
                                                                      smem);
        } else {
            const uint32_t* start = tables + pd.table_off;
            const uint32_t total = start[pd.nh] * 4;
            static_assert(kFrameChunkOutputs == 256, "one frame output per thread of the block");
            const uint32_t o = bi.b + threadIdx.x;
            if (o < total) {  // quadrant-major, as frame_kernel
                const uint32_t q = o / start[pd.nh], hw = o - q * start[pd.nh];
                int h, w;
                frame_hw(start, pd.nh, hw, h, w);
                const int dy = (q & 1) ? -(H - h) : (H - h);
                const int dx = (q & 2) ? -(W - w) : (W - w);
                if (!(((q & 1) && dy == 0) || ((q & 2) && dx == 0)))
                    Op[(size_t)(dy + H - 1) * (2 * W - 1) + (dx + W - 1)] = frame_value(Lp, Rp, W, dy, dx, h, w);
            }
        }
    }
    // self-reset for the next launch: the last block out sets both counters back to 0
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&counters[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&counters[0], 0u);
            atomicExch(&counters[1], 0u);
        }
    }
}

}  // namespace

uint32_t batch_tiles_x(uint32_t cols) { return (2 * cols - 1 + 64 * kBWX - 1) / (64 * kBWX); }
uint32_t batch_tiles_y(uint32_t rows) { return (2 * rows - 1 + 64 * kBWY - 1) / (64 * kBWY); }
uint32_t batch_frame_chunk() { return kFrameChunkOutputs; }
uint32_t batch_default_border(int kind) { return kind == XC_PREC_FP16X3 ? kDefaultBorderFp16x3 : kDefaultBorder; }

uint32_t frame_table_host(uint32_t rows, uint32_t cols, uint32_t thresh, std::vector<uint32_t>& start) {
    start.clear();
    uint32_t total = 0;
    for (uint32_t h = 1; h <= rows && h < thresh; ++h) {
        const uint32_t cnt = std::min<uint32_t>(cols, (thresh + h - 1) / h - 1);
        if (cnt == 0) break;
        start.push_back(total);
        total += cnt;
    }
    const uint32_t nh = (uint32_t)start.size();
    start.push_back(total);
    return nh;
}

const void* batch_kernel_ptr(int kind) {
    switch (kind) {
    case XC_PREC_BF16X3: return reinterpret_cast<const void*>(&batch_kernel<XC_PREC_BF16X3>);
    case XC_PREC_BF16: return reinterpret_cast<const void*>(&batch_kernel<XC_PREC_BF16>);
    case XC_PREC_FP16: return reinterpret_cast<const void*>(&batch_kernel<XC_PREC_FP16>);
    case XC_PREC_FP16X3: return reinterpret_cast<const void*>(&batch_kernel<XC_PREC_FP16X3>);
    default: fail(XC_ERR_INVALID_ARGUMENT, "the batch kernel supports bf16x3, fp16x3, bf16 and fp16");
    }
}

uint32_t batch_lds_bytes(int kind) {
    return kind == XC_PREC_BF16X3 || kind == XC_PREC_FP16X3 ? geom<kBWX, kBWY, 2, kBKC>::BYTES
                                                             : geom<kBWX, kBWY, 1, kBKC>::BYTES;
}

uint32_t batch_blocks_per_cu() { return kBOCC; }

// Tile catalog of the split-precision kernel (the runtime replacement of the reference's
// compile-time template limits, CMakeLists.txt:16-45): xc_options.tile selects an entry, 0 the
// default for the size.  Codes 1-8 are the dx-split block tiles (mfma_tile: WX x WY waves own
// different outputs; KC columns per pass, FLUSH, Kahan, blocks per CU, left rows per step),
// 9-12 the row-split tiles (mfma_rows_tile: the waves share the outputs and split the rows).
// Measured on one MI355X (bf16x3 without the frame): 64 x 256^2 one_to_many, tile 9 1.97 ms
// (FLUSH 16 the same within run-to-run noise, at 1.5x the error), 11 2.17, 10 2.68, 12 2.76,
// 4 2.75, 6 4.0; 16 x 64 of 64^2
// n_to_m, tile 6 0.27 ms, 9 0.33, 10 0.37, 4 0.38, 12 0.43.  The default is 6 for matrices up
// to 64 columns (one block covers a pair's outputs) and 9 above.
namespace {
constexpr uint32_t kMfmaCandidates[] = {9, 6, 10, 11, 4, 12};  // the cost model's choices
}
int mfma_catalog_size() { return (int)(sizeof(kMfmaCandidates) / sizeof(kMfmaCandidates[0])); }
uint32_t mfma_candidate_tile(int i) { return kMfmaCandidates[i]; }

namespace {

template <int KIND>
void launch_tile(const family_args& a, const float* L, const float* R, float* O, uint32_t tile, uint32_t skip_area,
                 hipStream_t s) {
    if (tile == 0) tile = a.cols <= 64 ? 6 : 9;
    switch (tile) {
    case 1: launch<KIND, 4, 1, 128, 4, false, 2>(a, L, R, O, skip_area, s); break;
    case 2: launch<KIND, 4, 1, 256, 2, true, 1>(a, L, R, O, skip_area, s); break;
    case 3: launch<KIND, 4, 1, 128, 2, true, 1>(a, L, R, O, skip_area, s); break;
    case 4: launch<KIND, 4, 1, 128, 8, false, 2>(a, L, R, O, skip_area, s); break;
    case 5: launch<KIND, 4, 1, 64, 8, false, 2>(a, L, R, O, skip_area, s); break;
    case 6: launch<KIND, 2, 2, 64, 4, false, 2>(a, L, R, O, skip_area, s); break;
    case 7: launch<KIND, 4, 1, 128, 8, false, 2, 2>(a, L, R, O, skip_area, s); break;
    case 8: launch<KIND, 4, 1, 128, 8, false, 1, 2>(a, L, R, O, skip_area, s); break;
    case 9: launch_rows<KIND, 2, 128, 8, 2>(a, L, R, O, skip_area, s); break;
    case 10: launch_rows<KIND, 2, 64, 8, 2>(a, L, R, O, skip_area, s); break;
    case 11: launch_rows<KIND, 2, 128, 4, 2>(a, L, R, O, skip_area, s); break;
    case 12: launch_rows<KIND, 1, 64, 8, 3>(a, L, R, O, skip_area, s); break;
    default: fail(XC_ERR_INVALID_ARGUMENT, "tile code " + std::to_string(tile) + " is not in the MFMA tile catalog");
    }
}

}  // namespace

void launch_mfma_f32(const family_args& a, const float* L, const float* R, float* O, int kind, uint32_t border,
                     uint32_t tile, hipStream_t s) {
    if (tile > 12)
        fail(XC_ERR_INVALID_ARGUMENT, "tile code " + std::to_string(tile) + " is not in the MFMA tile catalog");
    switch (kind) {
    case XC_PREC_BF16X3: {
        const uint32_t thresh = border ? border : kDefaultBorder;
        launch_tile<XC_PREC_BF16X3>(a, L, R, O, tile, thresh, s);
        launch_frame_f32(a, L, R, O, thresh, s);
        break;
    }
    case XC_PREC_FP16X3: {
        const uint32_t thresh = border ? border : kDefaultBorderFp16x3;
        launch_tile<XC_PREC_FP16X3>(a, L, R, O, tile, thresh, s);
        launch_frame_f32(a, L, R, O, thresh, s);
        break;
    }
    case XC_PREC_BF16:
        launch_tile<XC_PREC_BF16>(a, L, R, O, tile, 0, s);
        break;
    case XC_PREC_FP16:
        launch_tile<XC_PREC_FP16>(a, L, R, O, tile, 0, s);
        break;
    default:
        fail(XC_ERR_INVALID_ARGUMENT, "not an MFMA precision");
    }
}

}  // namespace xc
