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

template <typename T>
void launch_direct(const family_args& a, const T* L, const T* R, T* O, const xc_options& opt, hipStream_t s) {
    if (a.cols > (uint32_t)kMaxCols) fail(XC_ERR_UNSUPPORTED, "direct kernel: cols above 32768");
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
