// Host side of the wave64 direct kernel (see xc_direct_kernel.hpp): the tile catalog, the
// plan of a launch (tile, block shape, work distribution) and the launchers; the float
// instances of the kernel.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "xc_direct_kernel.hpp"

namespace xc {

const void* direct_kernel_f32(const direct_tile& t) {
#define XC_PICK(A, B, C, D) \
    if (t.ty == A && t.rm == B && t.lm == C && t.jb == D) return reinterpret_cast<const void*>(&direct_kernel<float, A, B, C, D>);
    XC_DIRECT_TILES(XC_PICK)
#undef XC_PICK
    return nullptr;
}

namespace {

constexpr direct_tile kCatalog[] = {
#define XC_ENTRY(A, B, C, D) {A, B, C, D},
    XC_DIRECT_TILES(XC_ENTRY)
#undef XC_ENTRY
};

// The largest catalog tile that is nowhere above the request: most FMAs per staged value
// (TY*RM*LM*JB), then more right matrices, left matrices, shifts.
direct_tile round_to_catalog(const direct_tile& req) {
    direct_tile best{1, 1, 1, 1};
    long best_score = -1;
    for (const auto& c : kCatalog) {
        if (c.ty > req.ty || c.rm > req.rm || c.lm > req.lm || c.jb > req.jb) continue;
        const long score = (long)c.ty * c.rm * c.lm * c.jb * 4096 + c.rm * 256 + c.lm * 16 + c.ty;
        if (score > best_score) {
            best_score = score;
            best = c;
        }
    }
    return best;
}

int pow2_floor(uint32_t v, int cap) {
    int p = 1;
    while (p * 2 <= (int)v && p * 2 <= cap) p *= 2;
    return p;
}

template <typename T>
const void* kernel_of(const direct_tile& t) {
    return sizeof(T) == 4 ? direct_kernel_f32(t) : direct_kernel_f64(t);
}

struct direct_plan {
    direct_tile t;
    uint32_t nwx, nwy, tiles_x, tiles_y, chunks, rows_per_chunk, groups_l, groups_r;
    int len;
    uint32_t lds;
};

template <typename T>
direct_plan plan_direct(const family_args& a, const xc_options& opt) {
    const bool n_to_m = a.family == XC_N_TO_M;
    uint32_t per_left = a.m;  // right matrices that share one left
    if (a.family == XC_N_TO_MN) per_left = a.m / a.n;
    if (a.family == XC_ONE_TO_ONE) per_left = 1;

    // Requested tile, or the default of the family.
    direct_tile t{};
    t.ty = opt.shifts_per_thread ? pow2_floor(opt.shifts_per_thread, 16) : 0;
    t.rm = opt.right_mats_per_thread ? pow2_floor(opt.right_mats_per_thread, 8) : 0;
    t.lm = opt.left_mats_per_thread ? pow2_floor(opt.left_mats_per_thread, 4) : 0;
    t.jb = opt.rows_per_step ? pow2_floor(opt.rows_per_step, 4) : 0;
    // Defaults measured on MI355X (one_to_one 128^2, one_to_many 64 x 256^2, n_to_m 16 x 64 of
    // 64^2): one right matrix: 4 shifts x 4 rows per step; several right matrices per left:
    // 4 shifts x 2 matrices x 2 rows; n_to_m: 2 shifts x 4 right x 2 left matrices x 2 rows.
    const bool many_left = n_to_m && a.n >= 2;
    if (!t.rm) t.rm = many_left ? 4 : (per_left >= 2 ? 2 : 1);
    if (!t.lm) t.lm = many_left ? 2 : 1;
    if (!t.ty) t.ty = many_left ? 2 : 4;
    if (!t.jb) t.jb = t.rm * t.lm == 1 ? 4 : 2;
    // Matrices that do not exist are not tiled.
    t.rm = std::min<int>(t.rm, pow2_floor(per_left, 8));
    t.lm = n_to_m ? std::min<int>(t.lm, pow2_floor(a.n, 4)) : 1;
    t = round_to_catalog(t);
    direct_plan d{};
    d.t = t;
    const uint32_t SX = 2 * a.cols - 1, SY = 2 * a.rows - 1;
    // 8 waves where one block row spans 256 dx or more and there are enough groups, else 4.
    uint32_t waves = opt.waves_per_block ? (uint32_t)pow2_floor(opt.waves_per_block, 8)
                                         : (SX >= 256 && a.pairs >= 8 ? 8 : 4);
    const uint32_t need_x = (SX + 63) / 64;
    d.nwx = 1;
    while (d.nwx * 2 <= waves && d.nwx < need_x) d.nwx *= 2;
    d.nwy = waves / d.nwx;
    d.tiles_x = (SX + 64 * d.nwx - 1) / (64 * d.nwx);
    d.tiles_y = (SY + t.ty * d.nwy - 1) / (t.ty * d.nwy);
    if (n_to_m) {
        d.groups_l = (a.n + t.lm - 1) / t.lm;
        d.groups_r = (a.m + t.rm - 1) / t.rm;
    } else if (a.family == XC_N_TO_MN) {
        d.groups_l = 1;
        d.groups_r = a.n * ((per_left + t.rm - 1) / t.rm);
    } else {
        d.groups_l = 1;
        d.groups_r = (a.m + t.rm - 1) / t.rm;
    }
    // Work distribution over the right rows of each tile.
    if (opt.split && opt.rows_per_block)
        fail(XC_ERR_INVALID_ARGUMENT, "give either split (rectangle) or rows_per_block (triangle), not both");
    d.chunks = 1;
    d.rows_per_chunk = 0;
    const uint64_t base = (uint64_t)d.tiles_x * d.tiles_y * d.groups_r * d.groups_l;
    if (opt.rows_per_block) {
        d.rows_per_chunk = std::max<uint32_t>(opt.rows_per_block, (uint32_t)t.jb);
        d.chunks = (a.rows + d.rows_per_chunk - 1) / d.rows_per_chunk;
    } else if (opt.split) {
        d.chunks = std::min(opt.split, a.rows);
    } else {
        // Enough blocks to fill 256 CUs several times over: triangle chunks.
        uint32_t c = 1;
        while (base * c < 2048 && c * 2 <= a.rows / 8 + 1) c *= 2;
        if (c > 1) {
            d.rows_per_chunk = (a.rows + c - 1) / c;
            d.rows_per_chunk = std::max<uint32_t>(d.rows_per_chunk, (uint32_t)t.jb);
            d.chunks = (a.rows + d.rows_per_chunk - 1) / d.rows_per_chunk;
        }
    }
    const uint32_t cw = 4;
    d.len = (int)(((a.cols + 64 * d.nwx + cw + 8) + 3) & ~3u);
    d.lds = 2u * (uint32_t)(t.jb * t.rm) * (uint32_t)d.len * (uint32_t)sizeof(T);
    return d;
}

// ---------------------------------------------------------------- lanes over pairs
// Definition-based FP32 kernel for batches with many rights per left (one_to_many, n_to_m,
// n_to_mn).  A wave takes one left and 64 of its rights (the pairs (left, right), one per
// lane; a left's last wave may be partial) and a tile of TY consecutive dy x 4 consecutive
// dx, so every lane
// runs the same loops: no lane spends FMAs on another lane's zero padding (the register-tile
// kernel's lanes take 64 consecutive dx and share the union of their column ranges), and
// the left values are wave-uniform scalar loads.  Per right row j and chunk of 4 columns
// c0 .. c0+3 a lane loads one quad of its right row and does TY*4*4 FMAs
//   out(dy0+y, dx0+x) += L[j+dy0+y][c0+t] * R[j][c0+t-dx0-x],   y < TY, x, t < 4,
// the right values R[j][c0-dx0-3 .. c0-dx0+4] being a window of two quads that slides by one
// quad per chunk.  The inputs come from padded copies: the rights as quads interleaved over
// slots (quad position p of row j of slot i at ((j * Q4 + p) * count + i) * 4, column q at
// position q + PL, PL = 8 + (-W mod 4), so every window starts on a quad and reads zeros
// outside [0, W)), the slots ordered so that a wave's 64 rights are consecutive (n_to_mn:
// left l's rights l, l+n, l+2n, ... at slots l*MP .. l*MP+MP-1, MP their count rounded up to
// 64; zero past the last); the lefts row-major with TY-1 zero rows above and below and rows
// padded with zeros to a multiple of 4.  Products are summed in FP32 partials of 64 that are added into
// FP64 totals.  Tiles run nearest-the-centre first (the largest overlaps start first).
struct dp_args {
    int family, H, W;
    uint32_t n, m, pairs, groups;     // groups: waves of 64 rights of one left, all lefts
    uint32_t mpl, wpl;                // rights per left, waves per left
    uint32_t tiles_y, tiles_x, nbx;   // nbx: blocks of 4 x-tiles (one per wave)
    uint32_t r_count;                 // matrices in the right copy
    int Q4, PL, Hl, Wl;               // right copy: quads per row, column offset; left copy: rows, row length
    int by_unit;                      // block order (see the kernel)
};

// Four consecutive elements of a row: 16 bytes (float) or 32 bytes (double).
template <typename T>
struct quad_of;
template <>
struct quad_of<float> {
    using type = float4;
    // the lane's quad of a right-copy quad position (base wave-uniform)
    __device__ static float4 load(const float* base, uint32_t lane_bytes) { return buffer_quad(base, lane_bytes, 0); }
};
template <>
struct quad_of<double> {
    using type = double4;
    __device__ static double4 load(const double* base, uint32_t lane_bytes) {
        const float* b = reinterpret_cast<const float*>(base);
        const double2 lo = __builtin_bit_cast(double2, buffer_quad(b, lane_bytes, 0));
        const double2 hi = __builtin_bit_cast(double2, buffer_quad(b, lane_bytes, 16));
        return make_double4(lo.x, lo.y, hi.x, hi.y);
    }
};

template <typename T, int TY>
__global__ __launch_bounds__(256) void direct_pairs_kernel(const T* __restrict__ Lc, const T* __restrict__ Rq,
                                                           T* __restrict__ O, dp_args a) {
    using Q = typename quad_of<T>::type;
    constexpr int TX = 4, FLUSHC = 16;  // FLUSHC chunks of 4 columns per partial sum
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // Blocks go round-robin to the 8 XCDs; tiles run nearest the centre (largest overlap)
    // first.  Right copies larger than an XCD's L2 (by_unit): a unit (one x-block of 4
    // tiles, one group of 64 pairs) keeps all its dy bands on one XCD; those waves have the
    // same column range, so they take about the same time per right row and read each row
    // while it is in that XCD's L2.  Units are dealt to the XCDs in snake order.  Smaller
    // copies: a band of TY dy (all its x-blocks and groups) per XCD, bands dealt in turn
    // (measured faster where the copy fits the L2).
    uint32_t u, band;
    if (!block_order(blockIdx.x, a.groups * a.nbx, a.tiles_y, a.by_unit, u, band)) return;
    const uint32_t g = u % a.groups;
    const int ty = center_first((int)band, (int)a.tiles_y);
    const int tx = center_first((int)(u / a.groups), (int)a.nbx) * 4 + wave;
    if (tx >= (int)a.tiles_x) return;
    const int H = a.H, W = a.W;
    const int dy0 = -(H - 1) + TY * ty, dx0 = -(W - 1) + TX * tx;
    // the wave's left and its lanes' rights (t: index among the left's rights; lanes past the
    // last right read a zero slot and store nothing)
    const uint32_t li = g / a.wpl, t = (g % a.wpl) * 64 + (uint32_t)lane;
    const bool valid = t < a.mpl;
    const uint32_t slot = (a.family == XC_N_TO_MN ? li * a.wpl * 64 : 0) + t;
    const uint32_t pk = a.family == XC_N_TO_M ? li * a.m + t : (a.family == XC_N_TO_MN ? li + a.n * t : t);
    // right rows: those that meet a dy of the tile; columns: the union over the tile's dx,
    // from a multiple of 4
    const int jlo = max(0, -(dy0 + TY - 1)), jhi = min(H, H - dy0);
    const int clo = max(0, dx0) & ~3;
    const int nch = (min(W, W + dx0 + TX - 1) - clo + 3) >> 2;
    const uint32_t lane_bytes = slot * 4u * (uint32_t)sizeof(T);
    const int64_t rq = (int64_t)a.r_count * 4;  // elements between quad positions of one row
    const T* __restrict__ lt = Lc + ((int64_t)li * a.Hl + dy0 + TY - 1) * a.Wl + clo;  // + j * Wl: row j + dy0
    const T* __restrict__ rt = Rq + (int64_t)((clo - dx0 - 3 + a.PL) >> 2) * rq;       // + j * Q4 * rq: row j
    T acc[TY][TX];
    double tot[TY][TX];
#pragma unroll
    for (int y = 0; y < TY; ++y)
#pragma unroll
        for (int x = 0; x < TX; ++x) acc[y][x] = T(0), tot[y][x] = 0.0;
    int cnt = 0;
    auto flush = [&]() {
#pragma unroll
        for (int y = 0; y < TY; ++y)
#pragma unroll
            for (int x = 0; x < TX; ++x) {
                tot[y][x] += (double)acc[y][x];
                acc[y][x] = T(0);
            }
    };
    for (int j = jlo; j < jhi; ++j) {
        const T* __restrict__ lr = lt + (int64_t)j * a.Wl;
        const T* __restrict__ rr = rt + (int64_t)j * a.Q4 * rq;
        Q w0 = quad_of<T>::load(rr, lane_bytes), w1 = quad_of<T>::load(rr + rq, lane_bytes);
        Q lv[TY];
#pragma unroll
        for (int y = 0; y < TY; ++y) lv[y] = *reinterpret_cast<const Q*>(lr + (int64_t)y * a.Wl);
        for (int ch = 0; ch < nch; ++ch) {
            // the next chunk's values in flight during this one (the copies hold a quad past
            // the last window, the left copy a chunk past the last one)
            const Q w2 = quad_of<T>::load(rr + (ch + 2) * rq, lane_bytes);
            Q ln[TY];
#pragma unroll
            for (int y = 0; y < TY; ++y) ln[y] = *reinterpret_cast<const Q*>(lr + (int64_t)y * a.Wl + 4 * ch + 4);
            const T w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
            for (int y = 0; y < TY; ++y) {
                const T l[4] = {lv[y].x, lv[y].y, lv[y].z, lv[y].w};
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int x = 0; x < TX; ++x) acc[y][x] = fma(l[t], w[t + 3 - x], acc[y][x]);
            }
            w0 = w1;
            w1 = w2;
#pragma unroll
            for (int y = 0; y < TY; ++y) lv[y] = ln[y];
            if (++cnt == FLUSHC) {
                flush();
                cnt = 0;
            }
        }
    }
    flush();
    if (!valid) return;
    T* __restrict__ op = O + (size_t)pk * (2 * W - 1) * (2 * H - 1);
#pragma unroll
    for (int y = 0; y < TY; ++y) {
        const int dy = dy0 + y;
        if (dy > H - 1) break;
#pragma unroll
        for (int x = 0; x < TX; ++x)
            if (dx0 + x <= W - 1) op[(size_t)(dy + H - 1) * (2 * W - 1) + (dx0 + x + W - 1)] = (T)tot[y][x];
    }
}

// Left copy: count matrices of H x W into rows of Wl (zero past W) with pad zero rows above
// and below: element (r, c) of matrix i at (i * Hl + r + pad) * Wl + c.
template <typename T>
__global__ __launch_bounds__(256) void pairs_copy_left_kernel(const T* __restrict__ X, T* __restrict__ Xc,
                                                              uint32_t count, int H, int W, int Hl, int Wl, int pad) {
    const uint64_t total = (uint64_t)count * Hl * Wl;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % (uint64_t)Wl);
        const uint64_t rest = e / (uint64_t)Wl;
        const int r = (int)(rest % (uint64_t)Hl) - pad;
        const uint64_t i = rest / (uint64_t)Hl;
        Xc[e] = (r >= 0 && r < H && c < W) ? X[(i * H + r) * W + c] : T(0);
    }
}

// Right copy: quad position p of row j of slot i at ((j * Q4 + p) * count + i) * 4, holding
// columns 4p - PL .. 4p - PL + 3 (zero outside [0, W)) of the slot's matrix: with ln = 0 slot
// i is matrix i (zero for i >= nmat); with ln lefts (n_to_mn), slot l*MP + t is matrix
// l + ln*t (zero for t >= mpl).  Thread e writes quad e.
template <typename T>
__global__ __launch_bounds__(256) void pairs_copy_right_kernel(const T* __restrict__ X, T* __restrict__ Xq,
                                                               uint32_t count, uint32_t nmat, uint32_t ln,
                                                               uint32_t mp, uint32_t mpl, int H, int W, int Q4, int PL) {
    const uint64_t total = (uint64_t)count * H * Q4;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e % count, rest = e / count;
        const int p = (int)(rest % (uint64_t)Q4), j = (int)(rest / (uint64_t)Q4);
        uint64_t mat = i;
        bool ok = i < nmat;
        if (ln) {
            const uint64_t l = i / mp, t = i % mp;
            mat = l + (uint64_t)ln * t;
            ok = t < mpl;
        }
        const T* __restrict__ row = X + ((ok ? mat : 0) * H + j) * W;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = 4 * p + t - PL;
            Xq[e * 4 + t] = (ok && q >= 0 && q < W) ? row[q] : T(0);
        }
    }
}

constexpr int kPairsTY = 4;

// Rights per left of a family (0: one_to_one, which the kernel does not serve).
uint32_t rights_per_left(const family_args& a) {
    switch (a.family) {
    case XC_ONE_TO_MANY: return a.m;
    case XC_N_TO_M: return a.m;
    case XC_N_TO_MN: return a.m / a.n;
    default: return 0;
    }
}

// Whether the lanes-over-pairs kernel serves this batch by default: at least 48 rights per
// left (a wave's lanes at least 3/4 busy below 64), and its copy of the rights (the register-
// tile kernel needs none) at most 16 GiB.
bool pairs_kernel_applies(const family_args& a) {
    const uint32_t mpl = rights_per_left(a);
    if (mpl < 48) return false;
    const uint64_t lefts = a.family == XC_ONE_TO_MANY ? 1 : a.n;
    const uint64_t slots = (a.family == XC_N_TO_MN ? lefts : 1) * ((mpl + 63) / 64 * 64);
    const uint64_t copy = slots * a.rows * (a.cols + 24) * sizeof(float);
    return copy <= (16ull << 30);
}

// Waves of the lanes-over-pairs kernel.  Measured against the register-tile kernel (one_to_many
// and n_to_m, 16^2 to 256^2): 0.45-0.9x at 128-1024 waves, 1.14x at 2048, 1.6-2.2x from 4096.
uint64_t pairs_kernel_waves(const family_args& a) {
    const uint32_t nl = a.family == XC_ONE_TO_MANY ? 1 : a.n;
    return (uint64_t)nl * ((rights_per_left(a) + 63) / 64) * ((2 * a.rows - 1 + kPairsTY - 1) / kPairsTY) *
           ((2 * a.cols - 1 + 3) / 4);
}

template <typename T>
void launch_direct_pairs(const family_args& a, const T* L, const T* R, T* O, hipStream_t s) {
    constexpr int TY = kPairsTY;
    const int H = (int)a.rows, W = (int)a.cols;
    const uint32_t nl = a.family == XC_ONE_TO_MANY ? 1 : a.n;
    dp_args k{};
    k.family = a.family;
    k.H = H;
    k.W = W;
    k.n = nl;
    k.m = a.m;
    k.pairs = a.pairs;
    k.mpl = rights_per_left(a);
    k.wpl = (k.mpl + 63) / 64;
    k.groups = nl * k.wpl;
    // right slots: the rights in order, padded to a multiple of 64; n_to_mn: per left, its
    // rights, padded to a multiple of 64 each
    const uint32_t nr = a.family == XC_N_TO_MN ? nl * k.wpl * 64 : k.wpl * 64;
    k.tiles_y = (uint32_t)(2 * H - 1 + TY - 1) / TY;
    k.tiles_x = (uint32_t)(2 * W - 1 + 3) / 4;
    k.nbx = (k.tiles_x + 3) / 4;
    k.r_count = nr;
    k.PL = 8 + (4 - W % 4) % 4;
    k.Q4 = (k.PL + W + 12 + 3) / 4;  // a quad past the last window
    k.Hl = H + 2 * (TY - 1);
    k.Wl = (W + 3) / 4 * 4 + 4;       // a chunk past the last one (prefetch)
    if ((uint64_t)nr * 4 * sizeof(T) + 16 >= (1ull << 31)) fail(XC_ERR_UNSUPPORTED, "direct kernel: too many right matrices");
    const uint64_t lsize = (uint64_t)nl * k.Hl * k.Wl, rsize = (uint64_t)nr * H * k.Q4 * 4;
    T* Lc = static_cast<T*>(workspace(WS_DIRECT, sizeof(T) * (lsize + rsize), s));
    T* Rq = Lc + lsize;
    auto blocks_for = [](uint64_t n) { return (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384); };
    hipLaunchKernelGGL(pairs_copy_left_kernel<T>, dim3(blocks_for(lsize)), dim3(256), 0, s, L, Lc, nl, H, W, k.Hl,
                       k.Wl, TY - 1);
    XC_HIP(hipGetLastError());
    hipLaunchKernelGGL(pairs_copy_right_kernel<T>, dim3(blocks_for(rsize / 4)), dim3(256), 0, s, R, Rq, nr, a.m,
                       a.family == XC_N_TO_MN ? nl : 0u, k.wpl * 64, k.mpl, H, W, k.Q4, k.PL);
    XC_HIP(hipGetLastError());
    k.by_unit = rsize * sizeof(T) > (4u << 20);  // larger than one XCD's L2
    const uint64_t nblocks = blocks_for_order(k.groups * k.nbx, k.tiles_y, k.by_unit);
    if (nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    dim3 grid((uint32_t)nblocks), block(256);
    record_launch(sizeof(T) == 4 ? "direct_pairs_kernel<f32> ty4 tx4" : "direct_pairs_kernel<f64> ty4 tx4", grid, block,
                  0, a.rows, a.cols, a.pairs, 1);
    timed td("direct_kernel", s);
    hipLaunchKernelGGL((direct_pairs_kernel<T, TY>), grid, block, 0, s, Lc, Rq, O, k);
    XC_HIP(hipGetLastError());
}

template <typename T>
void launch_direct(const family_args& a, const T* L, const T* R, T* O, const xc_options& opt, hipStream_t s) {
    if (a.cols > (uint32_t)kMaxCols) fail(XC_ERR_UNSUPPORTED, "direct kernel: cols above 32768");
    // tile 0: the lanes-over-pairs kernel where it applies, the batch gives 2048 waves or more
    // and no register tile or work distribution is requested; 1: the register-tile kernel;
    // 2: the lanes-over-pairs kernel (float and double data alike)
    if (opt.tile > 2)
        fail(XC_ERR_INVALID_ARGUMENT, "direct-method tile must be 0 (default), 1 (register tiles) or 2 (lanes over pairs)");
    const bool keys = opt.shifts_per_thread || opt.right_mats_per_thread || opt.left_mats_per_thread ||
                      opt.rows_per_step || opt.waves_per_block || opt.split || opt.rows_per_block;
    if (opt.tile == 2 && rights_per_left(a) == 0)
        fail(XC_ERR_UNSUPPORTED, "direct-method tile 2 (lanes over pairs) does not serve one_to_one");
    if (opt.tile == 2 || (opt.tile == 0 && !keys && pairs_kernel_applies(a) && pairs_kernel_waves(a) >= 2048)) {
        launch_direct_pairs<T>(a, L, R, O, s);
        return;
    }
    direct_plan d = plan_direct<T>(a, opt);
    if (d.lds > 160 * 1024) {
        if (opt.rows_per_step || opt.right_mats_per_thread)
            fail(XC_ERR_UNSUPPORTED, "direct kernel: the staged rows of this tile exceed the LDS");
        xc_options o = opt;
        o.rows_per_step = 1;
        o.right_mats_per_thread = 1;
        d = plan_direct<T>(a, o);
        if (d.lds > 160 * 1024) fail(XC_ERR_UNSUPPORTED, "matrix too wide for the direct kernel's LDS row buffer");
    }
    const uint64_t nblocks = (uint64_t)d.tiles_x * d.tiles_y * d.chunks * d.groups_r * d.groups_l;
    if (nblocks == 0 || nblocks > 0x7fffffffULL) fail(XC_ERR_INVALID_ARGUMENT, "problem too large for one launch");
    dk_args k{};
    k.family = a.family;
    k.H = (int)a.rows;
    k.W = (int)a.cols;
    k.n = a.family == XC_ONE_TO_ONE || a.family == XC_ONE_TO_MANY ? 1 : a.n;
    k.m = a.family == XC_ONE_TO_ONE ? 1 : a.m;
    k.mpl = a.family == XC_N_TO_MN ? a.m / a.n : 0;
    k.tiles_x = d.tiles_x;
    k.tiles_y = d.tiles_y;
    k.chunks = d.chunks;
    k.rows_per_chunk = d.rows_per_chunk;
    k.groups_l = d.groups_l;
    k.groups_r = d.groups_r;
    k.nwx = d.nwx;
    k.nwy = d.nwy;
    k.len = d.len;
    k.atomic = d.chunks > 1 ? 1 : 0;
    if (k.atomic) {
        const uint64_t count = (uint64_t)a.pairs * (2 * a.cols - 1) * (2 * a.rows - 1);
        uint32_t zb = (uint32_t)std::min<uint64_t>((count + 255) / 256, 8192);
        timed tz("zero_kernel", s);
        hipLaunchKernelGGL(zero_kernel<T>, dim3(zb), dim3(256), 0, s, O, count);
    }
    const void* fn = kernel_of<T>(d.t);
    if (!fn) fail(XC_ERR_UNSUPPORTED, "direct kernel: tile not in the catalog");
    if (d.lds > 65536) allow_dynamic_lds(fn, d.lds);
    dim3 grid((uint32_t)nblocks), block(64 * d.nwx * d.nwy);
    char name[96];
    std::snprintf(name, sizeof(name), "direct_kernel<%s> ty%d rm%d lm%d jb%d", sizeof(T) == 4 ? "f32" : "f64", d.t.ty,
                  d.t.rm, d.t.lm, d.t.jb);
    record_launch(name, grid, block, d.lds, a.rows, a.cols, a.pairs, d.chunks);
    timed td("direct_kernel", s);
    void* args[] = {(void*)&L, (void*)&R, (void*)&O, (void*)&k};
    XC_HIP(hipLaunchKernel(fn, grid, block, args, d.lds, s));
}

}  // namespace

void launch_direct_f32(const family_args& a, const float* L, const float* R, float* O, const xc_options& opt,
                       hipStream_t s) {
    launch_direct<float>(a, L, R, O, opt, s);
}

void launch_direct_f64(const family_args& a, const double* L, const double* R, double* O, const xc_options& opt,
                       hipStream_t s) {
    launch_direct<double>(a, L, R, O, opt, s);
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
