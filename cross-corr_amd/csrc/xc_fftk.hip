// FFT method on this library's own FP64 FFT kernels (no rocFFT): three fused passes over HBM.
//
// The reference's FFT algorithm (src/one_to_many.hpp:679-745, src/n_to_m.hpp:700-736):
// zero-pad both matrices to 2H x 2W, R2C, multiply the left spectrum by the conjugate of
// each right spectrum (hadamard_original, src/hadamard.cu:33-58), C2R, then normalise on the host
// (src/fft_helpers.hpp:193-261: divide by the padded area, swap quadrants, crop).  Here the
// padded sizes are Pr x Pc, the powers of two >= 2H-1 and >= 2W-1 (any size >= 2N-1 makes
// the circular correlation equal the linear one), and the work is three kernels:
//
//   rows_fwd : two real input rows per complex FFT of length Pc (zero padding implicit),
//              split into the two half spectra, written transposed as S[mat][k][r]
//              (k = 0..Pc/2, only the H nonzero rows r);
//   cols     : per (pair, column k): X = FFT of length Pr of S[right][k][0..H-1] (zero
//              padded), then a second FFT of conj(Lhat) * X, conjugated:
//              Pr * IFFT(Lhat * conj(X)) (Lhat = the left's 2D spectrum, made by the same
//              kernel in "left" mode), of which only the 2H-1 circular rows of the result
//              are kept: Z[pair][k][z_row(rho)];
//   rows_inv : per pair, two output rows per complex FFT of length Pc: the Hermitian half
//              rows of Z are extended, inverse transformed, scaled by 1/(Pr*Pc) and written
//              straight into the cropped (2H-1) x (2W-1) result (quadrant swap by index).
//
// All arithmetic is FP64, in radix-8/4/2 Stockham passes.  Transforms of up to 2048 points
// live in registers (a wave owns each up to 512 points, 2 or 4 waves of a block at 1024 and
// 2048; LDS only for the exchanges between passes); 4096-point transforms are done by a block
// in an LDS slab (padded LDS index against bank conflicts).  Twiddles come from a 4096-entry
// table, read once per kernel into registers; every HBM load of a kernel is issued before its
// first use.  Uneven batches run jobs whose matrices share the padded lengths but keep their
// own sizes (a per-matrix size table).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "xc_methods.hpp"

namespace xc {
namespace {

constexpr int kFT = 256;              // threads per block: slab kernels, block-owned column FFTs
constexpr int kLogMaxP = 12;          // longest transform: 4096 (the twiddle table's length)
constexpr int kSlab = 1 << kLogMaxP;  // twiddle table entries

// LDS slot of element j of transform t (length P = 2^LOGP): one pad slot per 8 values, and a
// transform stride of 9P/8 + 1 slots, so that lanes touching the same j of consecutive
// transforms, or j = 8i, hit different banks.  With P a compile-time constant the pad and
// stride arithmetic folds into immediate LDS offsets.
template <int LOGP>
__device__ __forceinline__ int at(int t, int j) {
    constexpr int S = (1 << LOGP) + (1 << LOGP) / 8 + 1;
    return t * S + j + (j >> 3);
}

// Slab (values per block) of a transform length: 2^11 values, or one transform of 4096.
// Lengths of up to 2048 points are transformed in registers (fft_regs), 4096 in a slab.
template <int LOGP>
struct slab_of {
    static constexpr int LOGS = LOGP <= 11 ? 11 : LOGP;
};
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }  // (-i) * a

// In-register forward DFT of R points: X[k] = sum_j u[j] exp(-2 pi i jk / R).
__device__ __forceinline__ void dft2(double2& a, double2& b) {
    const double2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
__device__ __forceinline__ void dft4(double2& u0, double2& u1, double2& u2, double2& u3) {
    const double2 t0 = cadd(u0, u2), t1 = csub(u0, u2), t2 = cadd(u1, u3), t3 = mul_mi(csub(u1, u3));
    u0 = cadd(t0, t2);
    u2 = csub(t0, t2);
    u1 = cadd(t1, t3);
    u3 = csub(t1, t3);
}
template <int R>
__device__ __forceinline__ void dft(double2 (&u)[R]) {
    if constexpr (R == 2) {
        dft2(u[0], u[1]);
    } else if constexpr (R == 4) {
        dft4(u[0], u[1], u[2], u[3]);
    } else {
        static_assert(R == 8, "radix");
        constexpr double c = 0.70710678118654752440;  // sqrt(1/2)
        double2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
        double2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
        dft4(e0, e1, e2, e3);
        dft4(o0, o1, o2, o3);
        // o_k *= exp(-2 pi i k / 8)
        o1 = make_double2(c * (o1.x + o1.y), c * (o1.y - o1.x));
        o2 = mul_mi(o2);
        o3 = make_double2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
        u[0] = cadd(e0, o0);
        u[4] = csub(e0, o0);
        u[1] = cadd(e1, o1);
        u[5] = csub(e1, o1);
        u[2] = cadd(e2, o2);
        u[6] = csub(e2, o2);
        u[3] = cadd(e3, o3);
        u[7] = csub(e3, o3);
    }
}

// Barrier between the LDS phases of a wave-owned transform: every value a lane reads was
// written by its own wave, so ordering the wave's LDS accesses is enough.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// First radix of a length-2^LOGP transform (8, 2 or 4 as LOGP mod 3 is 0, 1, 2; it takes no
// twiddles) and the number of twiddled radix-8 passes after it.
template <int LOGP>
struct radix_plan {
    static constexpr int REM = LOGP % 3;
    static constexpr int LOGR0 = REM == 0 ? 3 : REM;
    static constexpr int NTW = (LOGP - LOGR0) / 3;
};

// Pass roots of the slab transforms of length 2^LOGP (1024..4096): w[j][q] is the root of the
// j-th twiddled radix-8 pass (p = 2^(LOGR0 + 3j)) for this thread's q-th butterfly, b =
// threadIdx + kFT q.  Kernels load them once, before their first transform: a table read
// inside a pass would wait on memory after every barrier.
template <int LOGP>
struct slab_roots {
    static constexpr int NB8 = (1 << slab_of<LOGP>::LOGS) / 8 / kFT;  // radix-8 butterflies per thread
    double2 w[radix_plan<LOGP>::NTW][NB8];
};
template <int LOGP>
__device__ __forceinline__ slab_roots<LOGP> load_slab_roots(const double2* __restrict__ tw) {
    slab_roots<LOGP> sr;
#pragma unroll
    for (int j = 0; j < radix_plan<LOGP>::NTW; ++j) {
        const int logpp = radix_plan<LOGP>::LOGR0 + 3 * j;
#pragma unroll
        for (int q = 0; q < slab_roots<LOGP>::NB8; ++q) {
            const int i = ((int)threadIdx.x + kFT * q) & ((1 << (LOGP - 3)) - 1);
            sr.w[j][q] = tw[(i & ((1 << logpp) - 1)) << (kLogMaxP - logpp - 3)];
        }
    }
    return sr;
}

// One Stockham radix-R pass over T transforms of length P = 2^LOGP in buf (slots `at`),
// T * P <= 2^LOGS; 2^LOGPP is the product of the radices already applied.  Butterfly i of a
// transform reads elements i + r*P/R and writes (i - k)*R + k + r*p, k = i mod p, after
// multiplying input r by exp(-2 pi i r k / (p R)).  Butterflies go to threads block-strided.
template <int R, int LOGR, int LOGP, int LOGPP>
__device__ __forceinline__ void fft_pass(double2* buf, int T, const slab_roots<LOGP>& sr) {
    constexpr int MAXB = (1 << slab_of<LOGP>::LOGS) / R / kFT;
    static_assert(MAXB >= 1, "slab too small for the radix");
    constexpr int LOGPR = LOGP - LOGR;
    const int nb = T << LOGPR;
    double2 u[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = (int)threadIdx.x + kFT * q;
        if (b < nb) {
            const int t = b >> LOGPR, i = b & ((1 << LOGPR) - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) u[q][r] = buf[at<LOGP>(t, i + (r << LOGPR))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = (int)threadIdx.x + kFT * q;
        if (b < nb) {
            const int t = b >> LOGPR, i = b & ((1 << LOGPR) - 1);
            const int k = i & ((1 << LOGPP) - 1);
            if constexpr (LOGPP > 0) {
                static_assert(R == 8, "only the first pass has another radix");
                const double2 w1 = sr.w[(LOGPP - radix_plan<LOGP>::LOGR0) / 3][q];
                double2 w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    u[q][r] = cmul(u[q][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft<R>(u[q]);
            const int j = ((i - k) << LOGR) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[at<LOGP>(t, j + (r << LOGPP))] = u[q][r];
        }
    }
    __syncthreads();
}

template <int LOGP, int LOGPP>
__device__ __forceinline__ void radix8_passes(double2* buf, int T, const slab_roots<LOGP>& sr) {
    if constexpr (LOGPP < LOGP) {
        fft_pass<8, 3, LOGP, LOGPP>(buf, T, sr);
        radix8_passes<LOGP, LOGPP + 3>(buf, T, sr);
    }
}

// Forward FFT of T transforms of length 2^LOGP (1024..4096) in buf; natural order.  The caller
// has synchronised the block before; the block is synchronised on return.
template <int LOGP>
__device__ __forceinline__ void fft_slab(double2* buf, int T, const slab_roots<LOGP>& sr) {
    static_assert(LOGP >= 10, "shorter transforms run in registers");
    constexpr int LOGR0 = radix_plan<LOGP>::LOGR0;
    fft_pass<1 << LOGR0, LOGR0, LOGP, 0>(buf, T, sr);
    radix8_passes<LOGP, LOGR0>(buf, T, sr);
}

// Transforms of up to 2048 points run in registers: P/8 threads hold one transform, 8 values
// each.  Up to 512 points a transform lives in one wave (BLOCK false: the wave holds 512/P of
// them); 1024 and 2048 points span two and four waves of the block.
template <int LOGP>
struct regs_of {
    static constexpr bool ON = LOGP <= 11;
    static constexpr bool BLOCK = LOGP > 9;
};
template <bool BLOCK>
__device__ __forceinline__ void regs_sync() {
    if constexpr (BLOCK)
        __syncthreads();
    else
        wave_sync();
}

// Pass roots of fft_regs<LOGP> at lane position i: w[j] = exp(-2 pi i k / (8p)), k = i mod p,
// for the j-th twiddled pass (p = 2^(LOGR0 + 3j)).  Kernels load them once, before their
// first transform: a table read inside a pass would wait on memory after every LDS exchange.
template <int LOGP>
struct regs_roots {
    static constexpr int REM = LOGP % 3;
    static constexpr int LOGR0 = REM == 0 ? 3 : REM;  // log2 of the first (untwiddled) radix
    static constexpr int N = (LOGP - LOGR0) / 3;     // twiddled radix-8 passes
    double2 w[N];
};
template <int LOGP>
__device__ __forceinline__ regs_roots<LOGP> load_roots(int i, const double2* __restrict__ tw) {
    regs_roots<LOGP> rr;
#pragma unroll
    for (int j = 0; j < regs_roots<LOGP>::N; ++j) {
        const int logpp = regs_roots<LOGP>::LOGR0 + 3 * j;
        rr.w[j] = tw[(i & ((1 << logpp) - 1)) << (kLogMaxP - logpp - 3)];
    }
    return rr;
}

// Radix-8 passes of fft_regs from p = 2^LOGPP on: middle passes through LDS, the last one
// (p = P/8, k = i) leaves its outputs i + (P/8) r in registers.
template <int LOGP, int LOGPP>
__device__ __forceinline__ void fft_regs_passes(double2 (&u)[8], double2* wb, int tl, int i,
                                                const regs_roots<LOGP>& rr) {
    constexpr int Q = 1 << (LOGP - 3);
    if constexpr (Q >= 8) {
        // at(tl, i + Q r) = at(tl, i) + r (Q + Q/8) when 8 divides Q: one address register and
        // constant offsets (the compiler does not split the shift of the sum by itself, and
        // keeps eight addresses live through both transforms otherwise).
        const int a = at<LOGP>(tl, i);
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = wb[a + r * (Q + (Q >> 3))];
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = wb[at<LOGP>(tl, i + Q * r)];
    }
    const int k = i & ((1 << LOGPP) - 1);
    // The root passes through an empty asm so that its powers are formed here, in this pass:
    // otherwise the compiler hoists the power chains of every pass (and keeps them for a second
    // transform with the same roots), which costs tens of VGPRs.
    double2 w1 = rr.w[(LOGPP - regs_roots<LOGP>::LOGR0) / 3];
    asm volatile("" : "+v"(w1.x), "+v"(w1.y));
    double2 w = w1;
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        u[r] = cmul(u[r], w);
        if (r + 1 < 8) w = cmul(w, w1);
    }
    dft<8>(u);
    regs_sync<regs_of<LOGP>::BLOCK>();
    if constexpr (LOGPP + 3 < LOGP) {
#pragma unroll
        for (int r = 0; r < 8; ++r) wb[at<LOGP>(tl, ((i - k) << 3) + k + (r << LOGPP))] = u[r];
        regs_sync<regs_of<LOGP>::BLOCK>();
        fft_regs_passes<LOGP, LOGPP + 3>(u, wb, tl, i, rr);
    }
}

// FFTs entered and left in registers, for P = 2^LOGP of 16..2048 points (regs_of).  Thread
// (tl, i) of the owning wave (up to 512 points) or block (1024, 2048) -- tl = id / Q,
// i = id mod Q, Q = P/8 -- holds elements i + Q s (s = 0..7) of its transform tl on entry, and
// the transform's outputs at the same positions on return.  The first pass (radix R0 = 8, 2
// or 4 as log2 P mod 3 is 0, 1, 2; 8/R0 butterflies per thread, no twiddles) reads registers
// only and the last radix-8 pass writes registers only, so a kernel that moves those positions
// to and from HBM needs LDS only for the exchanges between passes (one up to 64 points, two up
// to 512, three for 1024 and 2048).  wb is the owner's first transform slot; transform tl lives
// at wb + at<LOGP>(tl, .).  rr holds the pass roots of this thread (load_roots<LOGP>(i, table)).
// A block-owned transform synchronises the block: every thread of it must call.  (tl, i) are
// passed in (regs_pos) so that a kernel calling twice can keep the compiler from holding the
// first call's LDS addresses in registers for the second.
template <int LOGP>
__device__ __forceinline__ void fft_regs(double2 (&u)[8], double2* wb, const regs_roots<LOGP>& rr, int tl,
                                         int i) {
    static_assert(LOGP >= 4 && regs_of<LOGP>::ON, "register FFTs take 16..2048 points");
    constexpr int Q = 1 << (LOGP - 3);
    constexpr int LOGR0 = regs_roots<LOGP>::LOGR0;
    constexpr int R0 = 1 << LOGR0;
    // first pass (p = 1): butterfly b = i + Q m takes elements b + (P/R0) r = u[m + (8/R0) r]
#pragma unroll
    for (int m = 0; m < 8 / R0; ++m) {
        double2 v[R0];
#pragma unroll
        for (int r = 0; r < R0; ++r) v[r] = u[m + (8 / R0) * r];
        dft<R0>(v);
#pragma unroll
        for (int r = 0; r < R0; ++r) wb[at<LOGP>(tl, (i + Q * m) * R0 + r)] = v[r];
    }
    regs_sync<regs_of<LOGP>::BLOCK>();
    fft_regs_passes<LOGP, LOGR0>(u, wb, tl, i, rr);
}

// Position of this thread in the register FFTs of its block: transform t, the first transform
// t0 of the owning wave (0 for block-owned transforms), and i within the transform.
template <int LOGP>
struct regs_pos {
    int t0, t, i;
    __device__ __forceinline__ regs_pos() {
        constexpr int Q = 1 << (LOGP - 3);
        if constexpr (regs_of<LOGP>::BLOCK) {
            t0 = 0;
            t = (int)threadIdx.x >> (LOGP - 3);
        } else {
            t0 = (int)(threadIdx.x >> 6) << (9 - LOGP);
            t = t0 + ((int)(threadIdx.x & 63) >> (LOGP - 3));
        }
        i = (int)threadIdx.x & (Q - 1);
    }
};

// Threads per block of the row kernels (INV: rows_inv): 256, or more for register transforms
// of 1024 and 2048 points so that a block holds four transforms (8 rows: a column's piece of
// S or Z is a whole 128-byte line) -- except rows_inv at 2048 points, where two transforms per
// block (two blocks per CU instead of one) measured faster.  LOGS is the block's slab.
constexpr int rows_threads(int logP, bool inv) {
    return logP == 10 ? 512 : (logP == 11 ? (inv ? 512 : 1024) : 256);
}
template <int LOGP, bool INV>
struct rows_block {
    static constexpr int NT = rows_threads(LOGP, INV);
    static constexpr int LOGS = regs_of<LOGP>::ON ? (NT == 256 ? 11 : (NT == 512 ? 12 : 13)) : slab_of<LOGP>::LOGS;
};

// The block's T transformed row pairs (spectra in natural order in LDS) -> the half spectra
// of rows r0 .. r0 + 2T - 1 (those below Hm), written column-major into Sm (column stride H):
// with z = FFT(a + i b), A[k] = (z[k] + conj(z[P-k]))/2 and B[k] = (z[k] - conj(z[P-k]))/(2i).
template <int LOGPC, int NT>
__device__ __forceinline__ void split_rows(const double2* buf, double2* __restrict__ Sm, uint32_t r0, uint32_t Hm,
                                           uint32_t H, int logT) {
    constexpr int Pc = 1 << LOGPC;
    constexpr uint32_t Kc = Pc / 2 + 1;
    const int rows2 = 2 << logT;
    for (int idx = threadIdx.x; idx < (int)Kc * rows2; idx += NT) {
        const int k = idx >> (logT + 1), rr = idx & (rows2 - 1);
        const uint32_t r = r0 + rr;
        if (r >= Hm) continue;
        const int t = rr >> 1;
        const double2 zk = buf[at<LOGPC>(t, k)];
        const double2 zn = cconj(buf[at<LOGPC>(t, (Pc - k) & (Pc - 1))]);
        double2 v;
        if ((rr & 1) == 0) {
            v = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y + zn.y));
        } else {
            const double2 d = mul_mi(csub(zk, zn));
            v = make_double2(0.5 * d.x, 0.5 * d.y);
        }
        Sm[(size_t)k * H + r] = v;
    }
}

// rows_fwd for register transforms (up to 2048 points): a persistent grid of the blocks that
// fit at once, each taking row groups (items: matrix, group of 2T rows) in turn and loading
// the next item's rows into registers while it transforms and stores the current one, so
// that HBM latency is paid once per block rather than once per item.  Items are dealt so
// that each XCD (blocks b with b mod 8 = x) walks one contiguous eighth of the item range:
// the partial lines of S that neighbouring groups write meet in its L2.  Matrices, offsets
// and sizes as in rows_fwd_slab_kernel below.
template <typename Tin, int LOGPC, int NT = rows_block<LOGPC, false>::NT>
__global__ __launch_bounds__(NT) void rows_fwd_regs_kernel(const Tin* __restrict__ XL, const Tin* __restrict__ XR,
                                                           const uint64_t* __restrict__ moff,
                                                           const uint2* __restrict__ mdims, uint32_t nl, uint32_t H,
                                                           uint32_t W, int logT, uint32_t groups, uint32_t items,
                                                           double2* __restrict__ SL, double2* __restrict__ SR,
                                                           const double2* __restrict__ tw) {
    static_assert(regs_of<LOGPC>::ON, "register transforms only");
    constexpr int Q = 1 << (LOGPC - 3);
    extern __shared__ double2 buf[];
    const regs_pos<LOGPC> ps;
    const int t0 = ps.t0, t = ps.t, i = ps.i;
    const regs_roots<LOGPC> rr = load_roots<LOGPC>(i, tw);
    // this block's items: x * per + k + j * bpx, j = 0, 1, ... below min((x + 1) * per, items)
    const uint32_t nx = gridDim.x >= 8 ? 8 : 1;
    const uint32_t x = blockIdx.x % nx, kb = blockIdx.x / nx, bpx = gridDim.x / nx;
    const uint32_t per = (items + nx - 1) / nx;
    const uint32_t end = min((x + 1) * per, items);
    // where item `it` lives: input rows a, b of this thread, sizes, output
    struct where {
        const Tin* xa;
        const Tin* xb;
        uint32_t Hm, Wm, r0, mat;
        bool left, oka, okb;
    };
    auto locate = [&](uint32_t it) {
        where w;
        const uint32_t matg = it / groups, g = it - matg * groups;
        w.left = matg < nl;
        w.mat = w.left ? matg : matg - nl;
        w.r0 = g << (logT + 1);
        w.Hm = H;
        w.Wm = W;
        if (mdims) {
            const uint2 d = mdims[matg];
            w.Hm = d.x;
            w.Wm = d.y;
        }
        const Tin* Xm = (w.left ? XL : XR) + (moff ? moff[matg] : (size_t)w.mat * H * W);
        const uint32_t a = w.r0 + 2 * t, b = a + 1;
        w.oka = a < w.Hm;
        w.okb = b < w.Hm;
        w.xa = Xm + (w.oka ? (size_t)a * w.Wm : 0);
        w.xb = Xm + (w.okb ? (size_t)b * w.Wm : 0);
        return w;
    };
    // Branch-free loads from clamped (in-bounds) addresses; out-of-range positions are zeroed
    // when the values are used.
    auto fetch = [&](const where& w, Tin (&va)[8], Tin (&vb)[8]) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t c = i + Q * r, cc = c < w.Wm ? c : w.Wm - 1;
            va[r] = w.xa[cc];
            vb[r] = w.xb[cc];
        }
    };
    uint32_t it = x * per + kb;
    if (it >= end) return;
    where cur = locate(it);
    Tin va[8], vb[8];
    fetch(cur, va, vb);
    while (true) {
        const uint32_t nxt = it + bpx;
        const bool more = nxt < end;
        double2 u[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const bool in = i + Q * r < (int)cur.Wm;
            u[r] = make_double2(cur.oka && in ? (double)va[r] : 0.0, cur.okb && in ? (double)vb[r] : 0.0);
        }
        where nw{};
        if (more) {  // the next item's rows are in flight during this item's transform and stores
            nw = locate(nxt);
            fetch(nw, va, vb);
        }
        if (cur.r0 < cur.Hm) {  // block-uniform: a row group beyond a smaller matrix is skipped
            fft_regs<LOGPC>(u, buf + at<LOGPC>(t0, 0), rr, t - t0, i);
#pragma unroll
            for (int r = 0; r < 8; ++r) buf[at<LOGPC>(t, i + Q * r)] = u[r];
            __syncthreads();
            split_rows<LOGPC, NT>(buf, (cur.left ? SL : SR) + (size_t)cur.mat * ((1 << LOGPC) / 2 + 1) * H, cur.r0,
                                  cur.Hm, H, logT);
            __syncthreads();  // the next item's transform reuses the slab
        }
        if (!more) break;
        it = nxt;
        cur = nw;
    }
}

// rows_fwd for 4096-point transforms (the LDS slab): block (matrix, group of 2T rows);
// matrices 0..nl-1 are the left ones (into SL), the rest the right ones (into SR).  Matrix i
// starts at element moff[i] of its input when the offset table is given (uneven batches),
// else at i * H * W.  With mdims (uneven batches), matrix i is mdims[i].x x mdims[i].y, at
// most H x W (the padded lengths are those of H x W); S keeps a column stride of H.
template <typename Tin, int LOGPC, int NT = rows_block<LOGPC, false>::NT>
__global__ __launch_bounds__(NT) void rows_fwd_slab_kernel(const Tin* __restrict__ XL, const Tin* __restrict__ XR,
                                                           const uint64_t* __restrict__ moff,
                                                           const uint2* __restrict__ mdims, uint32_t nl, uint32_t H,
                                                           uint32_t W, int logT, uint32_t groups,
                                                           double2* __restrict__ SL, double2* __restrict__ SR,
                                                           const double2* __restrict__ tw) {
    static_assert(!regs_of<LOGPC>::ON && NT == kFT, "slab transforms");
    constexpr int logPc = LOGPC;
    extern __shared__ double2 buf[];
    // XCD-aware order: the row groups of one matrix run on one XCD, so the partial lines of S
    // that neighbouring groups write (2T rows of a column) merge in that XCD's L2.
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t matg = bid / groups, g = bid - matg * groups;
    const bool left = matg < nl;
    const uint32_t mat = left ? matg : matg - nl;
    const int Pc = 1 << logPc, T = 1 << logT;
    const uint32_t r0 = g << (logT + 1);
    uint32_t Hm = H, Wm = W;
    if (mdims) {
        const uint2 d = mdims[matg];
        Hm = d.x;
        Wm = d.y;
    }
    if (r0 >= Hm) return;  // the whole block: a row group beyond a smaller matrix of the job
    const Tin* Xm = (left ? XL : XR) + (moff ? moff[matg] : (size_t)mat * H * W);
    double2* S = left ? SL : SR;
    // T = 2^LOGS / Pc transforms at most; element idx of the slab is column c of transform t.
    // Branch-free clamped loads first, then the LDS writes.
    constexpr int NI = (1 << slab_of<LOGPC>::LOGS) / kFT;
    const slab_roots<LOGPC> sr = load_slab_roots<LOGPC>(tw);
    const int n_el = T << logPc;
    Tin va[NI], vb[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        const uint32_t t = idx >> logPc, c = idx & (Pc - 1);
        const uint32_t a = r0 + 2 * t, b = a + 1;
        const size_t cc = c < Wm ? c : Wm - 1;
        va[q] = Xm[(size_t)(a < Hm ? a : Hm - 1) * Wm + cc];
        vb[q] = Xm[(size_t)(b < Hm ? b : Hm - 1) * Wm + cc];
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        if (idx >= n_el) continue;
        const uint32_t t = idx >> logPc, c = idx & (Pc - 1);
        const uint32_t a = r0 + 2 * t, b = a + 1;
        const bool in = c < Wm;
        buf[at<LOGPC>(t, c)] = make_double2(a < Hm && in ? (double)va[q] : 0.0, b < Hm && in ? (double)vb[q] : 0.0);
    }
    __syncthreads();
    fft_slab<LOGPC>(buf, T, sr);
    split_rows<LOGPC, NT>(buf, S + (size_t)mat * (Pc / 2 + 1) * H, r0, Hm, H, logT);
}

// Row of Z that keeps circular row rho of a pair with H rows: the result's shifts dy = rho
// (rho < H) stay in place, dy = rho - Pr (rho > Pr - H) move down by Pr - 2H to rows
// H + 1 .. 2H - 1, so that both halves start on the same alignment as in the transform (and
// for Pr = 2H nothing moves); -1 for the rows in between and for rho >= Pr.  Row H of Z is
// never written.
__device__ __forceinline__ int z_row(uint32_t rho, uint32_t H, uint32_t Pr) {
    if (rho < H) return (int)rho;
    if (rho > Pr - H && rho < Pr) return (int)(rho - (Pr - 2 * H));
    return -1;
}

// Output row (dy + H - 1) of row z of Z (z_row's inverse): -1 for row H and rows >= 2H.
__device__ __forceinline__ int out_row_of_z(uint32_t z, uint32_t H) {
    if (z < H) return (int)(z + H - 1);
    if (z > H && z < 2 * H) return (int)(z - H - 1);
    return -1;
}

// Threads per block of the column kernel: one wave for wave-owned register transforms (up to
// 512 points; C = 512 / Pr columns per block, so that a pair's last column group wastes at
// most a wave), 256 for the block-owned and slab transforms.
constexpr int cols_threads(int logP) { return logP <= 9 ? 64 : kFT; }

// cols: block (item, group of C columns k).  mode 0: item = left matrix l, writes its full
// column spectra Lhat[l][k][rho] (the FFT along rho of SL[l][k][0..H), zero padded).  With
// mdims (uneven batches) a matrix has mdims[i].x rows of its own (i = l, or nl + r for the
// right of a pair); S keeps a column stride of H.
// mode 1: item = pair q0 + item (left l, right r from the family's order, or from the pair
// table of an uneven batch): X = the same FFT of SR[r][k], then a second FFT of
// conj(Lhat) * X, conjugated: Pr * IFFT_rho(Lhat * conj(X)), whose circular row rho goes to
// Z[item][k][z_row(rho)] (column stride Ys = 2H; the rows between the result's two halves
// are dropped).
// Blocks are ordered column-group major through the XCD remap, so the pairs that read the
// same Lhat columns run on one XCD and find them in its L2.
template <int LOGPR, int MODE>
__global__ __launch_bounds__(cols_threads(LOGPR)) void cols_kernel(const double2* __restrict__ S, double2* __restrict__ Lhat,
                                                   double2* __restrict__ Z, const uint2* __restrict__ ptab,
                                                   const uint2* __restrict__ mdims, uint32_t nl, int family,
                                                   uint32_t n, uint32_t m, uint32_t q0, uint32_t items, uint32_t H,
                                                   uint32_t Kc, uint32_t Ys, int logC,
                                                   const double2* __restrict__ tw) {
    constexpr int logPr = LOGPR;
    extern __shared__ double2 buf[];
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t cg = bid / items, item = bid - cg * items;
    const int Pr = 1 << logPr, C = 1 << logC;
    const uint32_t k0 = cg << logC;
    uint32_t l = item, r = item;
    if constexpr (MODE == 1) {
        if (ptab) {
            l = ptab[q0 + item].x;
            r = ptab[q0 + item].y;
        } else {
            pair_of(family, n, m, q0 + item, l, r);
        }
    }
    const uint32_t Hq = mdims ? mdims[MODE == 1 ? nl + r : l].x : H;  // rows of this item's matrices
    const double2* Sm = S + (size_t)r * Kc * H;
    double2* Ll = Lhat + (size_t)l * Kc * Pr;
    if constexpr (regs_of<LOGPR>::ON) {
        // C = 8 * threads / Pr: the columns go from HBM to HBM through fft_regs; the product with
        // conj(Lhat) happens in registers between the two FFTs.
        constexpr int Q = 1 << (LOGPR - 3);
        const regs_pos<LOGPR> ps;
        const int t0 = ps.t0, i = ps.i;
        const uint32_t k = k0 + ps.t;
        const bool ok = k < Kc;
        double2* wb = buf + at<LOGPR>(t0, 0);
        const regs_roots<LOGPR> rr = load_roots<LOGPR>(i, tw);
        double2 u[8], lh[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t rho = i + Q * r;
            u[r] = make_double2(0.0, 0.0);
            if (ok && rho < Hq) u[r] = Sm[(size_t)k * H + rho];
        }
        double2* lk = Ll + (size_t)k * Pr;
        // The left spectrum is loaded together with the column (its latency overlaps the
        // first transform's).
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            lh[r] = make_double2(0.0, 0.0);
            if (MODE == 1 && ok) lh[r] = lk[i + Q * r];
        }
        fft_regs<LOGPR>(u, wb, rr, ps.t - t0, i);
        if constexpr (MODE == 0) {
            if (ok) {
#pragma unroll
                for (int r = 0; r < 8; ++r) lk[i + Q * r] = u[r];
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = cmul(cconj(lh[r]), u[r]);
        // An opaque copy of the position: the second transform recomputes its LDS addresses
        // instead of keeping the first one's live across the product (fewer VGPRs).
        int tl2 = ps.t - t0, i2 = i;
        asm volatile("" : "+v"(tl2), "+v"(i2));
        fft_regs<LOGPR>(u, wb, rr, tl2, i2);
        if (ok) {
            // rows from the opaque i2, so that the row indices of the loads are not kept live
            // through both transforms for these stores
            // z_row with constant offsets from two bases: rows below Hq in place, rows from
            // Pr - Hq on down by Pr - 2Hq (for Pr = 2Hq one base and every row, row Hq
            // included, which the row kernel reads as zero)
            double2* zlo = Z + (size_t)item * Kc * Ys + (size_t)k * Ys + i2;
            if (Pr == 2 * (int)Hq) {
#pragma unroll
                for (int r = 0; r < 8; ++r) zlo[Q * r] = cconj(u[r]);
            } else {
                double2* zhi = zlo - (Pr - 2 * (int)Hq);
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const uint32_t rho = (uint32_t)(i2 + Q * r);
                    if (rho < Hq)
                        zlo[Q * r] = cconj(u[r]);
                    else if (rho >= (uint32_t)Pr - Hq)
                        zhi[Q * r] = cconj(u[r]);
                }
            }
        }
        return;
    } else {
    // Slab: C = 2^LOGS / Pr columns; element idx of the slab is row rho of column k0 + t.
    // All loads (the column and, for pairs, the left spectrum) are issued before the first
    // transform, from clamped in-bounds addresses.
    constexpr int NI = (1 << slab_of<LOGPR>::LOGS) / kFT;
    const slab_roots<LOGPR> sr = load_slab_roots<LOGPR>(tw);
    double2 xs[NI], lh[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        const uint32_t k = k0 + (idx >> logPr), rho = idx & (Pr - 1);
        const uint32_t kk = k < Kc ? k : Kc - 1;
        xs[q] = Sm[(size_t)kk * H + (rho < Hq ? rho : Hq - 1)];
        lh[q] = MODE == 1 ? Ll[(size_t)kk * Pr + rho] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        const uint32_t k = k0 + (idx >> logPr), rho = idx & (Pr - 1);
        buf[at<LOGPR>(idx >> logPr, rho)] = (k < Kc && rho < Hq) ? xs[q] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_slab<LOGPR>(buf, C, sr);
    if constexpr (MODE == 0) {
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int idx = threadIdx.x + kFT * q;
            const uint32_t k = k0 + (idx >> logPr);
            if (k < Kc) Ll[(size_t)k * Pr + (idx & (Pr - 1))] = buf[at<LOGPR>(idx >> logPr, idx & (Pr - 1))];
        }
        return;
    }
    // conj(Lhat * conj(X)) = conj(Lhat) * X; its forward FFT, conjugated, is Pr * IFFT.
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        double2& v = buf[at<LOGPR>(idx >> logPr, idx & (Pr - 1))];
        v = k0 + (idx >> logPr) < Kc ? cmul(cconj(lh[q]), v) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_slab<LOGPR>(buf, C, sr);
    double2* Zq = Z + (size_t)item * Kc * Ys;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int idx = threadIdx.x + kFT * q;
        const uint32_t k = k0 + (idx >> logPr);
        const int y = z_row((uint32_t)(idx & (Pr - 1)), Hq, (uint32_t)Pr);
        if (k < Kc && y >= 0) Zq[(size_t)k * Ys + y] = cconj(buf[at<LOGPR>(idx >> logPr, idx & (Pr - 1))]);
    }
    }
}

// rows_inv: block (pair item, group of 2T consecutive rows z0 = 2T*g ... of Z, column stride
// Ys), so that each thread's two rows are adjacent in Z and a group's rows are one aligned
// run.  With mdims (uneven batches) the pair's size is that of its right matrix
// nl + ptab[q].y.  Row z goes to output row out_row_of_z(z) (row H of Z, which holds nothing,
// is read as zero); output column ox = dx + W - 1 takes circular column dx mod Pc.  Pair q's
// result starts at ooff[q] when the table is given, else at q times the result size.
template <typename Tout, int LOGPC, int NT = rows_block<LOGPC, true>::NT>
__global__ __launch_bounds__(NT) void rows_inv_kernel(const double2* __restrict__ Z, Tout* __restrict__ O,
                                                       const uint64_t* __restrict__ ooff,
                                                       const uint2* __restrict__ ptab, const uint2* __restrict__ mdims,
                                                       uint32_t nl, uint32_t q0, uint32_t H, uint32_t W, uint32_t Ys,
                                                       int logT, uint32_t groups, double scale,
                                                       const double2* __restrict__ tw) {
    constexpr int logPc = LOGPC;
    extern __shared__ double2 buf[];
    // XCD-aware order: the row groups of one pair run on one XCD, so the lines of Z that
    // neighbouring groups read in pieces (2T rows of a column) are fetched into its L2 once.
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t item = bid / groups, g = bid - item * groups;
    const int Pc = 1 << logPc, T = 1 << logT;
    const int half = Pc / 2;
    const uint32_t Kc = (uint32_t)half + 1;
    if (mdims) {  // uneven batch: this pair's own size (at most H x W)
        const uint2 d = mdims[nl + ptab[q0 + item].y];
        H = d.x;
        W = d.y;
    }
    const uint32_t SX = 2 * W - 1;
    const uint32_t z0 = g << (logT + 1);
    const int rows2 = 2 << logT;
    if (z0 >= 2 * H) return;  // a smaller pair of an uneven batch
    const double2* Zq = Z + (size_t)item * Kc * Ys;
    regs_roots<LOGPC> rr{};
    slab_roots<LOGPC> sr{};
    if constexpr (regs_of<LOGPC>::ON)
        rr = load_roots<LOGPC>((int)threadIdx.x & ((1 << (LOGPC - 3)) - 1), tw);
    else
        sr = load_slab_roots<LOGPC>(tw);
    // One thread per (column kk, transform t) loads the two rows' values A[kk], B[kk] (32
    // adjacent bytes) and writes both slots kk and Pc-kk of conj(Y), Y[k] = A[k] + i B[k]
    // extended by A[Pc-k] = conj(A[k]), B[Pc-k] = conj(B[k]).
    {
        // At most TMAX = 2^LOGS / Pc transforms, so (Kc << logT) <= (Pc/2 + 1) * TMAX items: all
        // loads first, then the writes.
        constexpr int TMAX = (1 << rows_block<LOGPC, true>::LOGS) >> LOGPC;
        constexpr int NQ = (((1 << (LOGPC - 1)) + 1) * TMAX + NT - 1) / NT;
        const int items = (int)Kc << logT;
        double2 A[NQ], B[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int idx = threadIdx.x + NT * q;
            const int kk = idx >> logT, t = idx & (T - 1);
            const uint32_t za = z0 + 2 * t;
            const double2* zc = Zq + (size_t)kk * Ys + za;
            A[q] = make_double2(0.0, 0.0);
            B[q] = make_double2(0.0, 0.0);
            if (idx < items && out_row_of_z(za, H) >= 0) A[q] = zc[0];
            if (idx < items && out_row_of_z(za + 1, H) >= 0) B[q] = zc[1];
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int idx = threadIdx.x + NT * q;
            if (idx >= items) continue;
            const int kk = idx >> logT, t = idx & (T - 1);
            buf[at<LOGPC>(t, kk)] = cadd(cconj(A[q]), mul_mi(cconj(B[q])));
            if (kk > 0 && kk < half) buf[at<LOGPC>(t, Pc - kk)] = cadd(A[q], mul_mi(B[q]));
        }
    }
    __syncthreads();
    Tout* Oq = O + (ooff ? ooff[q0 + item] : (size_t)(q0 + item) * (2 * H - 1) * SX);
    if constexpr (regs_of<LOGPC>::ON) {
        // T = 8 NT / Pc: the transforms (row pairs) go through fft_regs and are written from
        // registers: thread (t, i) holds circular columns c = i + Q r of rows rho0 + 2t, +1.
        constexpr int Q = 1 << (LOGPC - 3);
        const regs_pos<LOGPC> ps;
        const int t0 = ps.t0, t = ps.t, i = ps.i;
        double2 u[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = buf[at<LOGPC>(t, i + Q * r)];
        // a block-owned transform's first pass overwrites slots other waves still read
        if constexpr (regs_of<LOGPC>::BLOCK) __syncthreads();
        fft_regs<LOGPC>(u, buf + at<LOGPC>(t0, 0), rr, t - t0, i);
        const int oa = out_row_of_z(z0 + 2 * t, H), ob = out_row_of_z(z0 + 2 * t + 1, H);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int c = i + Q * r;
            const int ox = c <= (int)W - 1 ? c + (int)W - 1 : (c >= Pc - (int)W + 1 ? c - (Pc - (int)W + 1) : -1);
            if (ox < 0) continue;
            if (oa >= 0) Oq[(size_t)oa * SX + ox] = (Tout)(u[r].x * scale);
            if (ob >= 0) Oq[(size_t)ob * SX + ox] = (Tout)(-u[r].y * scale);
        }
        return;
    } else {
    fft_slab<LOGPC>(buf, T, sr);
    // IFFT(Y) = conj(FFT(conj(Y))) / Pc: row A is the real part, row B minus the imaginary.
    for (int rr = threadIdx.x >> 6; rr < rows2; rr += NT / 64) {
        const int oy = out_row_of_z(z0 + rr, H);
        if (oy < 0) continue;
        Tout* orow = Oq + (size_t)oy * SX;
        const int t = rr >> 1;
        for (int ox = threadIdx.x & 63; ox < (int)SX; ox += 64) {
            const int c = (ox - (int)(W - 1)) & (Pc - 1);
            const double2 f = buf[at<LOGPC>(t, c)];
            orow[ox] = (Tout)(((rr & 1) ? -f.y : f.x) * scale);
        }
    }
    }
}

// Row transforms of P = 3N points (N = 2^LOGN, 16..512), for widths where 3N >= 2W - 1 is
// shorter than the power of two.  Each of a block's T transforms (row pairs) is split into
// the three sub-sequences x[3n + m] (m = 0, 1, 2) of N points, which go through fft_regs<LOGN>
// as 3T wave-owned transforms (sub-transform s = 3t + m); a radix-3 pass through LDS combines
// them (decimation in time): with F_m the sub-spectra and w = exp(-2 pi i / P),
//   X[k + N j] = F_0[k] + W_3^j w^k F_1[k] + W_3^2j w^2k F_2[k],  j = 0, 1, 2.
// A block holds T transforms (3T N / 8 threads, a multiple of 64): 2 at 1536 points (two
// blocks per CU measured faster than one), 4 from 384 to 768, more below (192 threads).
constexpr int rows3_t(int logN) { return logN >= 9 ? 2 : (logN >= 7 ? 4 : 512 >> logN); }
template <int LOGN>
struct rows3 {
    static constexpr int N = 1 << LOGN, P = 3 * N, Q = N / 8;
    static constexpr int T = rows3_t(LOGN);                 // transforms (row pairs) per block
    static constexpr int NT = 3 * T * N / 8;
    static constexpr int LOGT = T == 2 ? 1 : (T == 4 ? 2 : (T == 8 ? 3 : (T == 16 ? 4 : 5)));
    static constexpr int XS = P + P / 8 + 1;                // LDS stride of a whole transform
    static constexpr int SLOTS = 3 * T * (N + N / 8 + 1) > T * XS ? 3 * T * (N + N / 8 + 1) : T * XS;
    static constexpr int NI = (T * N + NT - 1) / NT;        // radix-3 butterflies per thread
    static_assert(NT % 64 == 0 && NT <= 1024 && (1 << LOGT) == T && 8 * NT == T * P, "rows3 geometry");
};
template <int LOGN>
__device__ __forceinline__ int xat(int t, int p) {
    return t * rows3<LOGN>::XS + p + (p >> 3);
}

// tw3[j] = exp(-2 pi i j / 12288): the P-th roots for P = 3 * 2^n <= 12288, one table per
// device, kept for the process.
constexpr int kTw3 = 3 * kSlab;
const double2* twiddles3();

// Radix-3 butterfly of the combine pass: (a, b, c) = (F_0, w^k F_1, w^2k F_2) -> X[k + N j].
__device__ __forceinline__ void dft3(const double2 a, const double2 b, const double2 c, double2& x0, double2& x1,
                                     double2& x2) {
    constexpr double h = 0.86602540378443864676;  // sqrt(3) / 2
    const double2 sm = cadd(b, c), df = csub(b, c);
    const double2 t1 = make_double2(a.x - 0.5 * sm.x, a.y - 0.5 * sm.y);
    const double2 t2 = make_double2(h * df.y, -h * df.x);  // -i sqrt(3)/2 (b - c)
    x0 = cadd(a, sm);
    x1 = cadd(t1, t2);
    x2 = csub(t1, t2);
}

// Twiddles w^k, w^2k of this thread's radix-3 butterflies (idx = threadIdx + NT q, k = idx mod
// N), read at kernel start like the pass roots (a table read between two LDS phases would
// wait on memory).
template <int LOGN>
struct r3_roots {
    double2 w1[rows3<LOGN>::NI], w2[rows3<LOGN>::NI];
};
template <int LOGN>
__device__ __forceinline__ r3_roots<LOGN> load_r3_roots(const double2* __restrict__ tw3) {
    using G = rows3<LOGN>;
    r3_roots<LOGN> w;
#pragma unroll
    for (int q = 0; q < G::NI; ++q) {
        const int k = ((int)threadIdx.x + G::NT * q) & (G::N - 1);
        w.w1[q] = tw3[k * (kTw3 / G::P)];
        w.w2[q] = tw3[2 * k * (kTw3 / G::P)];
    }
    return w;
}

// The radix-3 pass of the block's T transforms: sub-spectra F_m[k] at at<LOGN>(3t + m, k) in
// buf -> X[t][k + N j] in registers (this thread's butterflies idx = threadIdx + NT q,
// t = idx / N, k = idx mod N).
template <int LOGN>
__device__ __forceinline__ void radix3_pass(const double2* buf, const r3_roots<LOGN>& w,
                                            double2 (&X)[rows3<LOGN>::NI][3]) {
    using G = rows3<LOGN>;
#pragma unroll
    for (int q = 0; q < G::NI; ++q) {
        const int idx = (int)threadIdx.x + G::NT * q;
        const int tt = idx >> LOGN, k = idx & (G::N - 1);
        if (idx < G::T * G::N) {
            const double2 f0 = buf[at<LOGN>(3 * tt, k)];
            const double2 f1 = cmul(w.w1[q], buf[at<LOGN>(3 * tt + 1, k)]);
            const double2 f2 = cmul(w.w2[q], buf[at<LOGN>(3 * tt + 2, k)]);
            dft3(f0, f1, f2, X[q][0], X[q][1], X[q][2]);
        }
    }
}

// rows_fwd for P = 3N: a persistent grid as rows_fwd_regs_kernel (items: matrix, group of 2T
// rows; each XCD walks one contiguous eighth of the items; the next item's rows are loaded
// into registers while the current one is transformed and stored), matrices, offsets, sizes
// and output S as rows_fwd_slab_kernel.
template <typename Tin, int LOGN>
__global__ __launch_bounds__(rows3<LOGN>::NT) void rows_fwd3_kernel(
    const Tin* __restrict__ XL, const Tin* __restrict__ XR, const uint64_t* __restrict__ moff,
    const uint2* __restrict__ mdims, uint32_t nl, uint32_t H, uint32_t W, uint32_t groups, uint32_t items,
    double2* __restrict__ SL, double2* __restrict__ SR, const double2* __restrict__ tw,
    const double2* __restrict__ tw3) {
    using G = rows3<LOGN>;
    constexpr int N = G::N, P = G::P, Q = G::Q, T = G::T;
    constexpr uint32_t Kc = P / 2 + 1;
    extern __shared__ double2 buf[];
    const regs_pos<LOGN> ps;
    const int sub = ps.t, tt = sub / 3, m = sub - 3 * tt, i = ps.i;
    const regs_roots<LOGN> rr = load_roots<LOGN>(i, tw);
    const r3_roots<LOGN> w3 = load_r3_roots<LOGN>(tw3);
    const uint32_t nx = gridDim.x >= 8 ? 8 : 1;
    const uint32_t x = blockIdx.x % nx, kb = blockIdx.x / nx, bpx = gridDim.x / nx;
    const uint32_t per = (items + nx - 1) / nx;
    const uint32_t end = min((x + 1) * per, items);
    struct where {
        const Tin* xa;
        const Tin* xb;
        uint32_t Hm, Wm, r0, mat;
        bool left, oka, okb;
    };
    auto locate = [&](uint32_t it) {
        where w;
        const uint32_t matg = it / groups, g = it - matg * groups;
        w.left = matg < nl;
        w.mat = w.left ? matg : matg - nl;
        w.r0 = g * (2 * T);
        w.Hm = H;
        w.Wm = W;
        if (mdims) {
            const uint2 d = mdims[matg];
            w.Hm = d.x;
            w.Wm = d.y;
        }
        const Tin* Xm = (w.left ? XL : XR) + (moff ? moff[matg] : (size_t)w.mat * H * W);
        const uint32_t a = w.r0 + 2 * tt, b = a + 1;
        w.oka = a < w.Hm;
        w.okb = b < w.Hm;
        w.xa = Xm + (w.oka ? (size_t)a * w.Wm : 0);
        w.xb = Xm + (w.okb ? (size_t)b * w.Wm : 0);
        return w;
    };
    // element n = i + Q r of sub-sequence m is column 3n + m; loads from clamped addresses
    auto fetch = [&](const where& w, Tin (&va)[8], Tin (&vb)[8]) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t c = 3u * (uint32_t)(i + Q * r) + (uint32_t)m, cc = c < w.Wm ? c : w.Wm - 1;
            va[r] = w.xa[cc];
            vb[r] = w.xb[cc];
        }
    };
    uint32_t it = x * per + kb;
    if (it >= end) return;
    where cur = locate(it);
    Tin va[8], vb[8];
    fetch(cur, va, vb);
    while (true) {
        const uint32_t nxt = it + bpx;
        const bool more = nxt < end;
        double2 u[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const bool in = 3u * (uint32_t)(i + Q * r) + (uint32_t)m < cur.Wm;
            u[r] = make_double2(cur.oka && in ? (double)va[r] : 0.0, cur.okb && in ? (double)vb[r] : 0.0);
        }
        where nw{};
        if (more) {
            nw = locate(nxt);
            fetch(nw, va, vb);
        }
        if (cur.r0 < cur.Hm) {  // block-uniform: a row group beyond a smaller matrix is skipped
            fft_regs<LOGN>(u, buf + at<LOGN>(ps.t0, 0), rr, sub - ps.t0, i);
#pragma unroll
            for (int r = 0; r < 8; ++r) buf[at<LOGN>(sub, i + Q * r)] = u[r];
            __syncthreads();
            double2 X[G::NI][3];
            radix3_pass<LOGN>(buf, w3, X);
            __syncthreads();  // X overwrites the sub-spectra
#pragma unroll
            for (int q = 0; q < G::NI; ++q) {
                const int idx = (int)threadIdx.x + G::NT * q;
                if (idx < T * N) {
                    const int t2 = idx >> LOGN, k = idx & (N - 1);
#pragma unroll
                    for (int j = 0; j < 3; ++j) buf[xat<LOGN>(t2, k + N * j)] = X[q][j];
                }
            }
            __syncthreads();
            // split into the two rows' half spectra (as split_rows), column-major into S
            double2* Sm = (cur.left ? SL : SR) + (size_t)cur.mat * Kc * H;
            for (int idx = threadIdx.x; idx < (int)Kc * 2 * T; idx += G::NT) {
                const int k = idx >> (G::LOGT + 1), rw = idx & (2 * T - 1);
                const uint32_t r = cur.r0 + rw;
                if (r >= cur.Hm) continue;
                const int t2 = rw >> 1;
                const double2 zk = buf[xat<LOGN>(t2, k)];
                const double2 zn = cconj(buf[xat<LOGN>(t2, k == 0 ? 0 : P - k)]);
                double2 v;
                if ((rw & 1) == 0) {
                    v = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y + zn.y));
                } else {
                    const double2 d = mul_mi(csub(zk, zn));
                    v = make_double2(0.5 * d.x, 0.5 * d.y);
                }
                Sm[(size_t)k * H + r] = v;
            }
            __syncthreads();  // the next item's transform reuses the slab
        }
        if (!more) break;
        it = nxt;
        cur = nw;
    }
}

// rows_inv for P = 3N: block (pair item, group of 2T rows z0 = 2T g ... of Z), otherwise as
// rows_inv_kernel (Z rows by z_row, output rows out_row_of_z, crop and quadrant swap by index).
template <typename Tout, int LOGN>
__global__ __launch_bounds__(rows3<LOGN>::NT) void rows_inv3_kernel(
    const double2* __restrict__ Z, Tout* __restrict__ O, const uint64_t* __restrict__ ooff,
    const uint2* __restrict__ ptab, const uint2* __restrict__ mdims, uint32_t nl, uint32_t q0, uint32_t H, uint32_t W,
    uint32_t Ys, uint32_t groups, double scale, const double2* __restrict__ tw, const double2* __restrict__ tw3) {
    using G = rows3<LOGN>;
    constexpr int N = G::N, P = G::P, Q = G::Q, T = G::T;
    constexpr uint32_t Kc = P / 2 + 1;
    extern __shared__ double2 buf[];
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t item = bid / groups, g = bid - item * groups;
    if (mdims) {
        const uint2 d = mdims[nl + ptab[q0 + item].y];
        H = d.x;
        W = d.y;
    }
    const uint32_t SX = 2 * W - 1;
    const uint32_t z0 = g * (2 * T);
    if (z0 >= 2 * H) return;
    const double2* Zq = Z + (size_t)item * Kc * Ys;
    const regs_pos<LOGN> ps;
    const int sub = ps.t, tt = sub / 3, m = sub - 3 * tt, i = ps.i;
    const regs_roots<LOGN> rr = load_roots<LOGN>(i, tw);
    const r3_roots<LOGN> w3 = load_r3_roots<LOGN>(tw3);
    // Hermitian extension of the two rows of each transform (conjugated, as rows_inv_kernel)
    {
        constexpr int NQ = ((int)Kc * T + G::NT - 1) / G::NT;
        double2 A[NQ], B[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int idx = threadIdx.x + G::NT * q;
            const int kk = idx >> G::LOGT, t = idx & (T - 1);
            const uint32_t za = z0 + 2 * t;
            const double2* zc = Zq + (size_t)(idx < (int)Kc * T ? kk : 0) * Ys + za;
            A[q] = make_double2(0.0, 0.0);
            B[q] = make_double2(0.0, 0.0);
            if (idx < (int)Kc * T && out_row_of_z(za, H) >= 0) A[q] = zc[0];
            if (idx < (int)Kc * T && out_row_of_z(za + 1, H) >= 0) B[q] = zc[1];
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int idx = threadIdx.x + G::NT * q;
            if (idx >= (int)Kc * T) continue;
            const int kk = idx >> G::LOGT, t = idx & (T - 1);
            buf[xat<LOGN>(t, kk)] = cadd(cconj(A[q]), mul_mi(cconj(B[q])));
            if (kk > 0 && kk < P / 2) buf[xat<LOGN>(t, P - kk)] = cadd(A[q], mul_mi(B[q]));
        }
    }
    __syncthreads();
    double2 u[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) u[r] = buf[xat<LOGN>(tt, 3 * (i + Q * r) + m)];
    __syncthreads();  // the sub-transforms' slots overlap the extended rows
    fft_regs<LOGN>(u, buf + at<LOGN>(ps.t0, 0), rr, sub - ps.t0, i);
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[at<LOGN>(sub, i + Q * r)] = u[r];
    __syncthreads();
    double2 X[G::NI][3];
    radix3_pass<LOGN>(buf, w3, X);
    // IFFT(Y) = conj(FFT(conj(Y))) / P: row A is the real part, row B minus the imaginary
    Tout* Oq = O + (ooff ? ooff[q0 + item] : (size_t)(q0 + item) * (2 * H - 1) * SX);
#pragma unroll
    for (int q = 0; q < G::NI; ++q) {
        const int idx = (int)threadIdx.x + G::NT * q;
        if (idx >= T * N) continue;
        const int t2 = idx >> LOGN, k = idx & (N - 1);
        const int oa = out_row_of_z(z0 + 2 * t2, H), ob = out_row_of_z(z0 + 2 * t2 + 1, H);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int c = k + N * j;
            const int ox = c <= (int)W - 1 ? c + (int)W - 1 : (c >= P - (int)W + 1 ? c - (P - (int)W + 1) : -1);
            if (ox < 0) continue;
            if (oa >= 0) Oq[(size_t)oa * SX + ox] = (Tout)(X[q][j].x * scale);
            if (ob >= 0) Oq[(size_t)ob * SX + ox] = (Tout)(-X[q][j].y * scale);
        }
    }
}

int ilog2_ceil(uint32_t x) {
    int e = 0;
    while ((1u << e) < x) ++e;
    return e;
}

// Twiddle tables in device memory of the code object (no allocation): tw[j] =
// exp(-2 pi i j / 4096) and tw3[j] = exp(-2 pi i j / 12288).  The host computes them and
// copies them in once per device.
__device__ double2 g_tw[kSlab];
__device__ double2 g_tw3[kTw3];

std::vector<double2> twiddle_values() {
    std::vector<double2> h(kSlab);
    const double two_pi = 6.283185307179586476925286766559;
    for (int j = 0; j < kSlab; ++j) {
        // exact at the eighths of the circle, where cos/sin of the rounded argument are not
        const int oct = j % (kSlab / 8);
        if (oct == 0) {
            static const double c8[8][2] = {{1, 0},
                                            {0.70710678118654752440, -0.70710678118654752440},
                                            {0, -1},
                                            {-0.70710678118654752440, -0.70710678118654752440},
                                            {-1, 0},
                                            {-0.70710678118654752440, 0.70710678118654752440},
                                            {0, 1},
                                            {0.70710678118654752440, 0.70710678118654752440}};
            h[j] = make_double2(c8[j / (kSlab / 8)][0], c8[j / (kSlab / 8)][1]);
        } else {
            const double a = two_pi * (double)j / (double)kSlab;
            h[j] = make_double2(std::cos(a), -std::sin(a));
        }
    }
    return h;
}

std::vector<double2> twiddle3_values() {
    std::vector<double2> h(kTw3);
    const double two_pi = 6.283185307179586476925286766559;
    for (int j = 0; j < kTw3; ++j) {
        const double a = two_pi * (double)j / (double)kTw3;
        h[j] = make_double2(std::cos(a), -std::sin(a));
    }
    // exact at the twelfths of the circle (the radix-3 and radix-4 points)
    const double s3 = 0.86602540378443864676;
    const double c12[12][2] = {{1, 0},   {s3, -0.5}, {0.5, -s3}, {0, -1},  {-0.5, -s3}, {-s3, -0.5},
                               {-1, 0},  {-s3, 0.5}, {-0.5, s3}, {0, 1},   {0.5, s3},   {s3, 0.5}};
    for (int q = 0; q < 12; ++q) h[q * (kTw3 / 12)] = make_double2(c12[q][0], c12[q][1]);
    return h;
}

struct device_tables {
    bool ready = false;
    double2* tw = nullptr;
    double2* tw3 = nullptr;
};

// The current device's twiddle tables: on its first use the symbols' addresses are looked up
// and the values copied in (synchronously, before any kernel reads them).
const device_tables& tables_of_device() {
    static std::mutex mu;
    static std::vector<device_tables> per_dev;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1);
    device_tables& t = per_dev[dev];
    if (!t.ready) {
        XC_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&t.tw), HIP_SYMBOL(g_tw)));
        XC_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&t.tw3), HIP_SYMBOL(g_tw3)));
        const std::vector<double2> h = twiddle_values(), h3 = twiddle3_values();
        XC_HIP(hipMemcpy(t.tw, h.data(), sizeof(double2) * kSlab, hipMemcpyHostToDevice));
        XC_HIP(hipMemcpy(t.tw3, h3.data(), sizeof(double2) * kTw3, hipMemcpyHostToDevice));
        t.ready = true;
    }
    return t;
}

const double2* twiddles() { return tables_of_device().tw; }
const double2* twiddles3() { return tables_of_device().tw3; }

struct slab_shape {
    int logS;      // log2 of the slab (values per block)
    bool regs;     // register transforms (regs_of)
    int logT;      // log2 of the transforms per block
    uint32_t lds;  // dynamic LDS bytes
};

// The slab of slab_of<LOGP> and the transforms per block: as many as `want` asks for, within
// the slab.  Register transforms (up to 2048 points) always fill it: 8 values per thread, so a
// block of `threads` holds 8 threads / P of them.
slab_shape shape_for(int logP, uint32_t want, int threads = 256) {
    slab_shape v{};
    v.regs = logP <= 11;  // as regs_of<LOGP>::ON
    // as rows_block<LOGP>::LOGS for the row kernels (threads), slab_of<LOGP>::LOGS otherwise
    v.logS = v.regs ? ilog2_ceil((uint32_t)threads) + 3 : std::max(11, logP);
    int lt = v.regs ? v.logS - logP : 0;
    while (lt + logP < v.logS && (1u << lt) < want) ++lt;
    v.logT = lt;
    v.lds = (uint32_t)(((1u << lt) * ((1u << logP) + (1u << logP) / 8 + 1)) * sizeof(double2));
    return v;
}

// Calls f(std::integral_constant<int, LOGP>) for the run-time transform length 2^logP.
template <typename F>
void with_logp(int logP, F&& f) {
    switch (logP) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    case 10: f(std::integral_constant<int, 10>{}); break;
    case 11: f(std::integral_constant<int, 11>{}); break;
    case 12: f(std::integral_constant<int, 12>{}); break;
    default: fail(XC_ERR_UNSUPPORTED, "own-FFT method: transform length out of range");
    }
}

template <typename K>
void allow_lds(K kernel, uint32_t bytes) {
    if (bytes > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(kernel), bytes);
}

struct fftk_geom {
    int logPr, logPc;  // logPc: of a power-of-two row length
    bool f3;           // row length Pc = 3 * 2^logN (rows3 kernels)
    int logN;
    uint32_t Pr, Pc, Kc;
    slab_shape v1, vc, v3;
    uint32_t g1, cg, g3;
    uint32_t Ys;  // column stride of Z: 2H (the 2H-1 circular rows of the result, z_row)
    size_t s_bytes, col_bytes, z_bytes;
};

fftk_geom geom_of(uint32_t H, uint32_t W, bool pow2_rows = false) {
    if (H == 0 || W == 0 || 2 * H - 1 > (uint32_t)kSlab || 2 * W - 1 > (uint32_t)kSlab)
        fail(XC_ERR_UNSUPPORTED, "own-FFT method: rows and cols must be 1..2048");
    fftk_geom g{};
    g.logPr = std::max(4, ilog2_ceil(2 * H - 1));
    g.Pr = 1u << g.logPr;
    g.Pc = pow2_rows ? fftk_col_len(W) : fftk_row_len(W);
    g.f3 = (g.Pc & (g.Pc - 1)) != 0;
    g.logN = g.f3 ? ilog2_ceil(g.Pc / 3) : 0;
    g.logPc = g.f3 ? 0 : ilog2_ceil(g.Pc);
    g.Kc = g.Pc / 2 + 1;
    // Slab shape per kernel (see shape_for): rows_fwd 2T input rows per block, cols C columns,
    // rows_inv 2T output rows.
    g.vc = shape_for(g.logPr, g.Kc, cols_threads(g.logPr));
    if (g.f3) {  // rows3<logN>: T = 8 NT / P transforms per block of NT threads
        const uint32_t n = g.Pc / 3, t = (uint32_t)rows3_t(g.logN);
        const uint32_t slots = std::max(3 * t * (n + n / 8 + 1), t * (g.Pc + g.Pc / 8 + 1));
        g.v1 = slab_shape{0, true, ilog2_ceil(t), slots * (uint32_t)sizeof(double2)};
        g.v3 = g.v1;
    } else {
        g.v1 = shape_for(g.logPc, (H + 1) / 2, rows_threads(g.logPc, false));
        g.v3 = shape_for(g.logPc, H, rows_threads(g.logPc, true));
    }
    g.g1 = (H + (2u << g.v1.logT) - 1) >> (g.v1.logT + 1);
    g.cg = (g.Kc + (1u << g.vc.logT) - 1) >> g.vc.logT;
    g.g3 = (2 * H + (2u << g.v3.logT) - 1) >> (g.v3.logT + 1);  // a block may be taller
    g.Ys = 2 * H;
    g.s_bytes = sizeof(double2) * (size_t)g.Kc * H;
    g.col_bytes = sizeof(double2) * (size_t)g.Kc * g.Pr;
    g.z_bytes = sizeof(double2) * (size_t)g.Kc * g.Ys;
    return g;
}

// Blocks of `kernel` (threads, dynamic LDS bytes) that the device holds at once, for a
// persistent grid over `items` work items: at most `items`, a multiple of the 8 XCDs when
// there are at least 8.  The occupancy is queried once per kernel and size.
uint32_t resident_grid(const void* kernel, int threads, uint32_t lds, uint32_t items) {
    static std::mutex mu;
    static std::vector<std::tuple<const void*, int, uint32_t, int, int>> known;  // kernel, threads, lds, dev, blocks
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const auto& e : known)
            if (std::get<0>(e) == kernel && std::get<1>(e) == threads && std::get<2>(e) == lds && std::get<3>(e) == dev)
                per_cu = std::get<4>(e);
        if (per_cu == 0) {
            XC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
            per_cu = std::max(per_cu, 1);
            known.emplace_back(kernel, threads, lds, dev, per_cu);
        }
    }
    int cus = 0;
    XC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    uint32_t g = std::min<uint32_t>(items, (uint32_t)cus * (uint32_t)per_cu);
    if (g >= 8) g -= g % 8;
    return std::max<uint32_t>(g, 1);
}

// Launch, with the kernel's own start and end recorded under `name` while kernel timing is on.
template <typename K, typename... Args>
void launch_timed(const char* name, bool timing, K kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t s,
                  Args... args) {
    if (timing && timing_enabled()) {
        hipEvent_t a = nullptr, b = nullptr;
        timing_launch_events(name, &a, &b);
        hipExtLaunchKernelGGL(kernel, grid, block, lds, s, a, b, 0, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
    }
}

template <typename Tio>
void launch_job(const fftk_job& j, const Tio* L, const Tio* R, Tio* O, const fftk_ws& ws, hipStream_t s,
                bool timing) {
    const uint32_t H = j.H, W = j.W;
    const fftk_geom g = geom_of(H, W, j.pow2_rows);
    const double2* tw = twiddles();
    double2* SR = static_cast<double2*>(ws.sr);
    double2* SL = static_cast<double2*>(ws.sl);
    double2* LH = static_cast<double2*>(ws.lh);
    double2* Zw = static_cast<double2*>(ws.z);
    const uint32_t chunk = std::max<uint32_t>(1, std::min<uint32_t>(j.pairs, ws.zchunk));
    const uint2* ptab = reinterpret_cast<const uint2*>(j.ptab);
    const uint2* mdims = reinterpret_cast<const uint2*>(j.mdims);
    const double2* tw3 = g.f3 ? twiddles3() : nullptr;
    if (g.f3) {
        with_logp(g.logN, [&](auto LN) {
            constexpr int LOGN = decltype(LN)::value;
            if constexpr (LOGN <= 9) {
                using G3 = rows3<LOGN>;
                if (G3::SLOTS * sizeof(double2) != g.v1.lds || G3::P != (int)g.Pc)
                    fail(XC_ERR_UNSUPPORTED, "own-FFT method: inconsistent 3N-point row geometry");
                auto k = rows_fwd3_kernel<Tio, LOGN>;
                allow_lds(k, g.v1.lds);
                const uint32_t items = (j.nl + j.nr) * g.g1;
                const uint32_t grid = resident_grid(reinterpret_cast<const void*>(k), G3::NT, g.v1.lds, items);
                launch_timed("fftk_rows_fwd", timing, k, dim3(grid), dim3(G3::NT), g.v1.lds, s, L, R, j.moff, mdims,
                             j.nl, H, W, g.g1, items, SL, SR, tw, tw3);
            }
        });
    } else
    with_logp(g.logPc, [&](auto LP) {
        constexpr int LOGPC = decltype(LP)::value;
        const uint32_t items = (j.nl + j.nr) * g.g1;
        if constexpr (regs_of<LOGPC>::ON) {
            auto k = rows_fwd_regs_kernel<Tio, LOGPC>;
            allow_lds(k, g.v1.lds);
            const uint32_t grid = resident_grid(reinterpret_cast<const void*>(k), rows_threads(LOGPC, false),
                                                g.v1.lds, items);
            launch_timed("fftk_rows_fwd", timing, k, dim3(grid), dim3(rows_threads(LOGPC, false)), g.v1.lds, s, L, R,
                         j.moff, mdims, j.nl, H, W, g.v1.logT, g.g1, items, SL, SR, tw);
        } else {
            auto k = rows_fwd_slab_kernel<Tio, LOGPC>;
            allow_lds(k, g.v1.lds);
            launch_timed("fftk_rows_fwd", timing, k, dim3(items), dim3(kFT), g.v1.lds, s, L, R, j.moff, mdims, j.nl, H,
                         W, g.v1.logT, g.g1, SL, SR, tw);
        }
    });
    auto cols = [&](const char* name, const double2* S, double2* Zout, auto MODE, uint32_t items,
                    uint32_t q0) {
        with_logp(g.logPr, [&](auto LP) {
            auto k = cols_kernel<decltype(LP)::value, decltype(MODE)::value>;
            allow_lds(k, g.vc.lds);
            launch_timed(name, timing, k, dim3(items * g.cg), dim3(cols_threads(g.logPr)), g.vc.lds, s, S, LH, Zout,
                         ptab, mdims, j.nl, j.family, j.n, j.m, q0, items, H, g.Kc, g.Ys, g.vc.logT, tw);
        });
    };
    cols("fftk_cols_left", SL, nullptr, std::integral_constant<int, 0>{}, j.nl, 0);
    const double scale = 1.0 / ((double)g.Pr * (double)g.Pc);
    for (uint32_t q0 = 0; q0 < j.pairs; q0 += chunk) {
        const uint32_t nq = std::min(chunk, j.pairs - q0);
        cols("fftk_cols", SR, Zw, std::integral_constant<int, 1>{}, nq, q0);
        if (g.f3) {
            with_logp(g.logN, [&](auto LN) {
                constexpr int LOGN = decltype(LN)::value;
                if constexpr (LOGN <= 9) {
                    using G3 = rows3<LOGN>;
                    auto k = rows_inv3_kernel<Tio, LOGN>;
                    allow_lds(k, g.v3.lds);
                    launch_timed("fftk_rows_inv", timing, k, dim3(nq * g.g3), dim3(G3::NT), g.v3.lds, s,
                                 (const double2*)Zw, O, j.ooff, ptab, mdims, j.nl, q0, H, W, g.Ys, g.g3, scale, tw,
                                 tw3);
                }
            });
            continue;
        }
        with_logp(g.logPc, [&](auto LP) {
            auto k = rows_inv_kernel<Tio, decltype(LP)::value>;
            allow_lds(k, g.v3.lds);
            launch_timed("fftk_rows_inv", timing, k, dim3(nq * g.g3), dim3(rows_threads(g.logPc, true)), g.v3.lds, s,
                         (const double2*)Zw, O, j.ooff, ptab, mdims, j.nl, q0, H, W, g.Ys,
                         g.v3.logT, g.g3, scale, tw);
        });
    }
    XC_HIP(hipGetLastError());
    record_launch("fftk_rows_inv", dim3(g.g3),
                  dim3(g.f3 ? 3u * (uint32_t)rows3_t(g.logN) * (g.Pc / 3) / 8 : (uint32_t)rows_threads(g.logPc, true)),
                  g.v3.lds, H, W, j.pairs, 1);
}

template <typename Tio>
void run_fftk(const family_args& a, const Tio* L, const Tio* R, Tio* O, uint32_t max_chunk, bool pow2_rows,
              hipStream_t s) {
    fftk_job j{};
    j.max_chunk = max_chunk;
    j.pow2_rows = pow2_rows;
    j.H = a.rows;
    j.W = a.cols;
    j.nl = a.family == XC_ONE_TO_ONE || a.family == XC_ONE_TO_MANY ? 1 : a.n;
    j.nr = a.family == XC_ONE_TO_ONE ? 1 : a.m;
    j.pairs = a.pairs;
    j.family = a.family;
    j.n = a.n;
    j.m = a.m;
    const fftk_ws_sizes z = fftk_sizes(j);
    fftk_ws ws{};
    ws.sr = workspace(WS_FK_SR, z.sr, s);
    ws.sl = workspace(WS_FK_SL, z.sl, s);
    ws.lh = workspace(WS_FK_LHAT, z.lh, s);
    ws.z = workspace(WS_FK_Z, z.z, s);
    ws.zchunk = z.chunk;
    launch_job<Tio>(j, L, R, O, ws, s, true);
}

// ---- allocation-free entry point ---------------------------------------------------------
// The row spectra S, the left spectra Lhat and the column-pass results Z of
// xc_cross_corr_noalloc_f32 live in this array of the code object instead of allocated
// workspaces.  Every device that runs the library holds one copy for the process; calls on
// one device share it, so calls that overlap on one device race (each call overwrites what it
// reads before reading it, so nothing needs clearing between calls).
constexpr size_t kScratchBytes = size_t(256) << 20;
__device__ __attribute__((aligned(256))) char scratch[kScratchBytes];

// The current device's address of `scratch`, looked up on the first call on that device.
char* scratch_of_device() {
    static std::mutex mu;
    static std::vector<char*> ptrs;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        char* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(XC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptrs[dev] = p;
    }
    return ptrs[dev];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// One launch sequence of the allocation-free path: a regular family over a slice of the
// call's matrices (element offsets of its first left, right and result).
struct noalloc_part {
    int family;
    uint32_t n, m, pairs;  // lefts, rights, results of the part
    size_t loff, roff, ooff;
    uint32_t chunk;        // pairs per launch of the column / inverse-row kernels
};

// Scratch bytes of a part with nl left and nr right spectra and `chunk` pairs of Z.
size_t part_bytes(const fftk_geom& g, uint32_t nl, uint32_t nr, uint32_t chunk) {
    return align256(g.s_bytes * nr) + align256(g.s_bytes * nl) + align256(g.col_bytes * nl) +
           align256(g.z_bytes * chunk);
}

// Splits a call into parts whose spectra fit the scratch array: the whole call if it fits,
// else groups of right matrices (one_to_many, n_to_mn: whole groups of n), of left matrices
// against every right (n_to_m), or of right matrices per left (n_to_m when the rights alone do
// not fit).  Each part's results are contiguous in the output.
std::vector<noalloc_part> plan_noalloc(const family_args& a, const fftk_geom& g) {
    const size_t B = kScratchBytes, hw = (size_t)a.rows * a.cols;
    const size_t out_elems = (size_t)(2 * a.rows - 1) * (2 * a.cols - 1);
    const size_t one_pair = part_bytes(g, 1, 1, 1);
    if (one_pair > B)
        fail(XC_ERR_UNSUPPORTED, "allocation-free entry point: one pair of " + std::to_string(a.rows) + " x " +
                                     std::to_string(a.cols) + " needs " + std::to_string(one_pair >> 20) +
                                     " MiB of scratch, above its " + std::to_string(kScratchBytes >> 20) + " MiB");
    std::vector<noalloc_part> parts;
    // Right matrices of `rights` consecutive results, with `nl` lefts, chunk = all pairs.
    auto fit_rights = [&](uint32_t nl, uint32_t unit, uint32_t most) -> uint32_t {
        uint32_t g_best = 0;
        for (uint32_t c = unit; c <= most; c += unit) {
            if (part_bytes(g, nl, c, c) > B) break;
            g_best = c;
        }
        return g_best;
    };
    switch (a.family) {
    case XC_ONE_TO_ONE:
        parts.push_back({XC_ONE_TO_ONE, 1, 1, 1, 0, 0, 0, 1});
        break;
    case XC_ONE_TO_MANY: {
        const uint32_t gr = fit_rights(1, 1, a.m);
        for (uint32_t r0 = 0; r0 < a.m; r0 += gr) {
            const uint32_t c = std::min(gr, a.m - r0);
            parts.push_back({XC_ONE_TO_MANY, 1, c, c, 0, r0 * hw, r0 * out_elems, c});
        }
        break;
    }
    case XC_N_TO_MN: {
        const uint32_t gr = fit_rights(a.n, a.n, a.m);
        if (gr == 0)
            fail(XC_ERR_UNSUPPORTED, "allocation-free entry point: the spectra of " + std::to_string(a.n) +
                                         " left matrices do not fit its scratch array");
        for (uint32_t r0 = 0; r0 < a.m; r0 += gr) {
            const uint32_t c = std::min(gr, a.m - r0);  // a multiple of n, as m is
            parts.push_back({XC_N_TO_MN, a.n, c, c, 0, r0 * hw, r0 * out_elems, c});
        }
        break;
    }
    default: {  // n_to_m
        uint32_t gl = 0;
        for (uint32_t c = 1; c <= a.n; ++c) {
            if (part_bytes(g, c, a.m, a.m) > B) break;
            gl = c;
        }
        if (gl > 0) {
            for (uint32_t l0 = 0; l0 < a.n; l0 += gl) {
                const uint32_t c = std::min(gl, a.n - l0);
                // as many pairs per launch as the rest of the array holds (at least m: gl fits them)
                const size_t rest = B - part_bytes(g, c, a.m, 0);
                uint32_t chunk = (uint32_t)std::min<size_t>((size_t)c * a.m, rest / g.z_bytes);
                while (chunk > a.m && part_bytes(g, c, a.m, chunk) > B) --chunk;
                parts.push_back({XC_N_TO_M, c, a.m, c * a.m, l0 * hw, 0, (size_t)l0 * a.m * out_elems, chunk});
            }
        } else {
            const uint32_t gr = fit_rights(1, 1, a.m);
            for (uint32_t l = 0; l < a.n; ++l)
                for (uint32_t r0 = 0; r0 < a.m; r0 += gr) {
                    const uint32_t c = std::min(gr, a.m - r0);
                    parts.push_back({XC_ONE_TO_MANY, 1, c, c, l * hw, r0 * hw, ((size_t)l * a.m + r0) * out_elems, c});
                }
        }
    }
    }
    return parts;
}

}  // namespace

void run_fftk_noalloc_f32(const family_args& a, const float* L, const float* R, float* O, hipStream_t s) {
    char* base = scratch_of_device();
    const fftk_geom g = geom_of(a.rows, a.cols);
    (void)tables_of_device();
    for (const noalloc_part& p : plan_noalloc(a, g)) {
        fftk_job j{};
        j.H = a.rows;
        j.W = a.cols;
        j.family = p.family;
        j.n = p.n;
        j.m = p.m;
        j.nl = p.family == XC_ONE_TO_ONE || p.family == XC_ONE_TO_MANY ? 1 : p.n;
        j.nr = p.m;
        j.pairs = p.pairs;
        fftk_ws ws{};
        char* q = base;
        ws.sr = q;
        q += align256(g.s_bytes * j.nr);
        ws.sl = q;
        q += align256(g.s_bytes * j.nl);
        ws.lh = q;
        q += align256(g.col_bytes * j.nl);
        ws.z = q;
        ws.zchunk = p.chunk;
        if ((size_t)(q - base) + align256(g.z_bytes * p.chunk) > kScratchBytes)
            fail(XC_ERR_INVALID_ARGUMENT, "allocation-free entry point: internal scratch plan out of bounds");
        launch_job<float>(j, L + p.loff, R + p.roff, O + p.ooff, ws, s, false);
    }
}

uint32_t fftk_col_len(uint32_t rows) {
    uint32_t p = 16;
    while (p < 2 * rows - 1) p <<= 1;
    return p;
}

uint32_t fftk_row_len(uint32_t cols) {
    const uint32_t p = fftk_col_len(cols);
    for (uint32_t n = 16; n <= 512; n <<= 1)
        if (3 * n >= 2 * cols - 1) return 3 * n < p ? 3 * n : p;
    return p;
}

fftk_ws_sizes fftk_sizes(const fftk_job& j) {
    const fftk_geom g = geom_of(j.H, j.W, j.pow2_rows);
    fftk_ws_sizes z{};
    // Z is processed in chunks of pairs of at most ~2 GiB.
    z.chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(j.pairs, (size_t(2) << 30) / g.z_bytes));
    if (j.max_chunk) z.chunk = std::min(z.chunk, j.max_chunk);
    z.sr = std::max<size_t>(g.s_bytes * j.nr, 16);
    z.sl = std::max<size_t>(g.s_bytes * j.nl, 16);
    z.lh = std::max<size_t>(g.col_bytes * j.nl, 16);
    z.z = g.z_bytes * z.chunk;
    return z;
}

void fftk_prepare() {
    (void)twiddles();
    (void)twiddles3();
}

void fftk_launch_f32(const fftk_job& j, const float* L, const float* R, float* O, const fftk_ws& ws, hipStream_t s) {
    launch_job<float>(j, L, R, O, ws, s, false);
}

void run_fftk_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t max_chunk, bool pow2_rows,
                  hipStream_t s) {
    run_fftk<float>(a, L, R, O, max_chunk, pow2_rows, s);
}

void run_fftk_f64(const family_args& a, const double* L, const double* R, double* O, uint32_t max_chunk,
                  bool pow2_rows, hipStream_t s) {
    run_fftk<double>(a, L, R, O, max_chunk, pow2_rows, s);
}

}  // namespace xc
