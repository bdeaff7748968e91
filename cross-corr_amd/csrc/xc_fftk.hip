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
//              Z[pair][k][rho] = Pr * IFFT(Lhat * conj(X))  (Lhat = the left's 2D spectrum,
//              made by the same kernel in "left" mode);
//   rows_inv : per pair, two output rows per complex FFT of length Pc: the Hermitian half
//              rows of Z are extended, inverse transformed, scaled by 1/(Pr*Pc) and written
//              straight into the cropped (2H-1) x (2W-1) result (quadrant swap by index).
//
// All arithmetic is FP64.  Each block transforms a slab of up to 4096 complex values in LDS
// with radix-8/4/2 Stockham passes (register butterflies, padded LDS index against bank
// conflicts); twiddles come from a 4096-entry table.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "xc_methods.hpp"

namespace xc {
namespace {

constexpr int kFT = 256;              // threads per block
constexpr int kLogMaxP = 12;          // longest transform: 4096 (the twiddle table's length)
constexpr int kSlab = 1 << kLogMaxP;  // twiddle table entries

__device__ __forceinline__ int pad(int j) { return j + (j >> 3); }
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

// Barrier between the read and write phases of a pass.  In WAVE mode every transform of the
// slab is owned by one wave for all passes, so ordering that wave's own LDS accesses is enough.
template <bool WAVE>
__device__ __forceinline__ void slab_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// One Stockham radix-R pass over T transforms of length P = 2^logP stored back to back in
// buf (padded index), T * P <= 2^LOGS.  logp: log2 of the product of the radices already
// applied.  Butterfly i of a transform reads elements i + r*P/R and writes (i - k)*R + k + r*p,
// k = i mod p, after multiplying input r by exp(-2 pi i r k / (p R)).  Butterflies go to
// threads block-strided, or (WAVE, T a multiple of 4) as one contiguous range of whole
// transforms per wave.
template <int R, int LOGR, int LOGS, bool WAVE>
__device__ __forceinline__ void fft_pass(double2* buf, int logP, int T, int logp, const double2* __restrict__ tw) {
    constexpr int MAXB = (1 << LOGS) / R / kFT;
    static_assert(MAXB >= 1, "slab too small for the radix");
    const int logPR = logP - LOGR;
    const int nb = T << logPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int per_wave = nb >> 2;
    double2 u[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const bool ok = WAVE ? lane + 64 * q < per_wave : (int)threadIdx.x + kFT * q < nb;
        const int b = WAVE ? wave * per_wave + lane + 64 * q : (int)threadIdx.x + kFT * q;
        if (ok) {
            const int base = (b >> logPR) << logP, i = b & ((1 << logPR) - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) u[q][r] = buf[pad(base + i + (r << logPR))];
        }
    }
    slab_sync<WAVE>();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const bool ok = WAVE ? lane + 64 * q < per_wave : (int)threadIdx.x + kFT * q < nb;
        const int b = WAVE ? wave * per_wave + lane + 64 * q : (int)threadIdx.x + kFT * q;
        if (ok) {
            const int base = (b >> logPR) << logP, i = b & ((1 << logPR) - 1);
            const int k = i & ((1 << logp) - 1);
            if (logp > 0) {
                const double2 w1 = tw[k << (kLogMaxP - logp - LOGR)];
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
            for (int r = 0; r < R; ++r) buf[pad(base + j + (r << logp))] = u[q][r];
        }
    }
    slab_sync<WAVE>();
}

// Forward FFT of T transforms of length 2^logP in buf; natural order.  The caller has
// synchronised the block before; the block is synchronised on return.
template <int LOGS, bool WAVE>
__device__ void fft_slab(double2* buf, int logP, int T, const double2* __restrict__ tw) {
    int logp = 0;
    const int rem = logP % 3;
    if (rem == 1) {
        fft_pass<2, 1, LOGS, WAVE>(buf, logP, T, logp, tw);
        logp += 1;
    } else if (rem == 2) {
        fft_pass<4, 2, LOGS, WAVE>(buf, logP, T, logp, tw);
        logp += 2;
    }
    while (logp < logP) {
        fft_pass<8, 3, LOGS, WAVE>(buf, logP, T, logp, tw);
        logp += 3;
    }
    if constexpr (WAVE) __syncthreads();
}

// rows_fwd: block (matrix, group of 2T rows).
template <typename Tin, int LOGS, bool WAVE>
__global__ __launch_bounds__(kFT) void rows_fwd_kernel(const Tin* __restrict__ X, uint32_t H, uint32_t W, int logPc,
                                                       int logT, uint32_t groups, double2* __restrict__ S,
                                                       const double2* __restrict__ tw) {
    extern __shared__ double2 buf[];
    const uint32_t mat = blockIdx.x / groups, g = blockIdx.x - mat * groups;
    const int Pc = 1 << logPc, T = 1 << logT;
    const uint32_t r0 = g << (logT + 1);
    const Tin* Xm = X + (size_t)mat * H * W;
    for (int idx = threadIdx.x; idx < (T << logPc); idx += kFT) {
        const int t = idx >> logPc, c = idx & (Pc - 1);
        const uint32_t a = r0 + 2 * t, b = a + 1;
        const double re = (a < H && (uint32_t)c < W) ? (double)Xm[(size_t)a * W + c] : 0.0;
        const double im = (b < H && (uint32_t)c < W) ? (double)Xm[(size_t)b * W + c] : 0.0;
        buf[pad(idx)] = make_double2(re, im);
    }
    __syncthreads();
    fft_slab<LOGS, WAVE>(buf, logPc, T, tw);
    const uint32_t Kc = (uint32_t)(Pc / 2 + 1);
    const int rows2 = 2 << logT;
    double2* Sm = S + (size_t)mat * Kc * H;
    for (int idx = threadIdx.x; idx < (int)Kc * rows2; idx += kFT) {
        const int k = idx >> (logT + 1), rr = idx & (rows2 - 1);
        const uint32_t r = r0 + rr;
        if (r >= H) continue;
        const int t = rr >> 1;
        const double2 zk = buf[pad((t << logPc) + k)];
        const double2 zn = cconj(buf[pad((t << logPc) + ((Pc - k) & (Pc - 1)))]);
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

// cols: block (item, group of C columns).  mode 0: item = left matrix l, writes its full
// column spectra Lhat[l][k][rho].  mode 1: item = pair q0 + item, writes
// Z[item][k][rho] = Pr * IFFT(Lhat[l] * conj(FFT(S[r])))  along rho.
template <int LOGS, bool WAVE>
__global__ __launch_bounds__(kFT) void cols_kernel(const double2* __restrict__ S, double2* __restrict__ Lhat,
                                                   double2* __restrict__ Z, int mode, int family, uint32_t n,
                                                   uint32_t m, uint32_t q0, uint32_t H, uint32_t Kc, int logPr,
                                                   int logC, uint32_t cgroups, const double2* __restrict__ tw) {
    extern __shared__ double2 buf[];
    const uint32_t item = blockIdx.x / cgroups, cg = blockIdx.x - item * cgroups;
    const int Pr = 1 << logPr, C = 1 << logC;
    const uint32_t k0 = cg << logC;
    uint32_t l = item, r = item;
    if (mode == 1) pair_of(family, n, m, q0 + item, l, r);
    const double2* Sm = S + (size_t)r * Kc * H;
    for (int idx = threadIdx.x; idx < (C << logPr); idx += kFT) {
        const uint32_t k = k0 + (idx >> logPr), rho = idx & (Pr - 1);
        buf[pad(idx)] = (k < Kc && rho < H) ? Sm[(size_t)k * H + rho] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_slab<LOGS, WAVE>(buf, logPr, C, tw);
    double2* Ll = Lhat + (size_t)l * Kc * Pr;
    if (mode == 0) {
        for (int idx = threadIdx.x; idx < (C << logPr); idx += kFT) {
            const uint32_t k = k0 + (idx >> logPr);
            if (k < Kc) Ll[(size_t)k * Pr + (idx & (Pr - 1))] = buf[pad(idx)];
        }
        return;
    }
    // conj(Lhat * conj(X)) = conj(Lhat) * X; its forward FFT, conjugated, is Pr * IFFT.
    for (int idx = threadIdx.x; idx < (C << logPr); idx += kFT) {
        const uint32_t k = k0 + (idx >> logPr);
        const double2 lh = k < Kc ? Ll[(size_t)k * Pr + (idx & (Pr - 1))] : make_double2(0.0, 0.0);
        buf[pad(idx)] = cmul(cconj(lh), buf[pad(idx)]);
    }
    __syncthreads();
    fft_slab<LOGS, WAVE>(buf, logPr, C, tw);
    double2* Zq = Z + (size_t)item * Kc * Pr;
    for (int idx = threadIdx.x; idx < (C << logPr); idx += kFT) {
        const uint32_t k = k0 + (idx >> logPr);
        if (k < Kc) Zq[(size_t)k * Pr + (idx & (Pr - 1))] = cconj(buf[pad(idx)]);
    }
}

// rows_inv: block (pair item, group of 2T output rows oy).  Output row oy = dy + H - 1 reads
// Z row rho = dy mod Pr; output column ox = dx + W - 1 takes circular column dx mod Pc.
template <typename Tout, int LOGS, bool WAVE>
__global__ __launch_bounds__(kFT) void rows_inv_kernel(const double2* __restrict__ Z, Tout* __restrict__ O,
                                                       uint32_t q0, uint32_t H, uint32_t W, int logPc, int logPr,
                                                       int logT, uint32_t groups, double scale,
                                                       const double2* __restrict__ tw) {
    extern __shared__ double2 buf[];
    const uint32_t item = blockIdx.x / groups, g = blockIdx.x - item * groups;
    const int Pc = 1 << logPc, Pr = 1 << logPr, T = 1 << logT;
    const int half = Pc / 2;
    const uint32_t Kc = (uint32_t)half + 1;
    const uint32_t SY = 2 * H - 1, SX = 2 * W - 1;
    const uint32_t oy0 = g << (logT + 1);
    const double2* Zq = Z + (size_t)item * Kc * Pr;
    // Two phases: even rows (A) set every slot, odd rows (B) add theirs.  The buffer holds
    // conj(Y), Y[k] = A[k] + i B[k] extended by A[Pc-k] = conj(A[k]).
    for (int phase = 0; phase < 2; ++phase) {
        for (int idx = threadIdx.x; idx < (int)Kc << logT; idx += kFT) {
            const int kk = idx >> logT, t = idx & (T - 1);
            const uint32_t oy = oy0 + 2 * t + phase;
            double2 v = make_double2(0.0, 0.0);
            if (oy < SY) {
                const uint32_t rho = (oy - (H - 1)) & (uint32_t)(Pr - 1);
                v = Zq[(size_t)kk * Pr + rho];
            }
            const int base = t << logPc;
            if (phase == 0) {
                buf[pad(base + kk)] = cconj(v);
                if (kk > 0 && kk < half) buf[pad(base + Pc - kk)] = v;
            } else {
                const double2 a = mul_mi(cconj(v)), b = mul_mi(v);
                double2& s0 = buf[pad(base + kk)];
                s0 = cadd(s0, a);
                if (kk > 0 && kk < half) {
                    double2& s1 = buf[pad(base + Pc - kk)];
                    s1 = cadd(s1, b);
                }
            }
        }
        __syncthreads();
    }
    fft_slab<LOGS, WAVE>(buf, logPc, T, tw);
    Tout* Oq = O + (size_t)(q0 + item) * SY * SX;
    const int rows2 = 2 << logT;
    for (int idx = threadIdx.x; idx < rows2 * (int)SX; idx += kFT) {
        const int rr = idx / (int)SX, ox = idx - rr * (int)SX;
        const uint32_t oy = oy0 + rr;
        if (oy >= SY) continue;
        const int c = (ox - (int)(W - 1)) & (Pc - 1);
        const double2 f = buf[pad(((rr >> 1) << logPc) + c)];
        const double v = (rr & 1) ? -f.y : f.x;
        Oq[(size_t)oy * SX + ox] = (Tout)(v * scale);
    }
}

int ilog2_ceil(uint32_t x) {
    int e = 0;
    while ((1u << e) < x) ++e;
    return e;
}

// tw[j] = exp(-2 pi i j / 4096), one table per device, kept for the process.
const double2* twiddles() {
    static std::mutex mu;
    static std::vector<double2*> tables;
    int dev = 0;
    XC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)tables.size() <= dev) tables.resize(dev + 1, nullptr);
    if (!tables[dev]) {
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
        double2* d = nullptr;
        if (hipMalloc(&d, sizeof(double2) * kSlab) != hipSuccess) {
            (void)hipGetLastError();
            fail(XC_ERR_NO_MEMORY, "twiddle table allocation failed");
        }
        XC_HIP(hipMemcpy(d, h.data(), sizeof(double2) * kSlab, hipMemcpyHostToDevice));
        tables[dev] = d;
    }
    return tables[dev];
}

struct slab_shape {
    int logS;      // log2 of the slab (values per block)
    bool wave;     // transforms owned by single waves
    int logT;      // log2 of the transforms per block
    uint32_t lds;  // dynamic LDS bytes
};

slab_shape shape_for(int logP, uint32_t want) {
    slab_shape v{};
    v.logS = logP <= 11 ? 11 : 12;
    v.wave = logP <= 9;  // T >= 4 then fits the 2048-value slab
    int lt = v.wave ? 2 : 0;
    while (lt + logP < v.logS && (1u << lt) < want) ++lt;
    v.logT = lt;
    v.lds = (uint32_t)(((1u << (logP + lt)) * 9 / 8) * sizeof(double2));
    return v;
}

template <typename F>
void with_shape(const slab_shape& v, F&& f) {
    if (v.logS == 12)
        f(std::integral_constant<int, 12>{}, std::false_type{});
    else if (v.wave)
        f(std::integral_constant<int, 11>{}, std::true_type{});
    else
        f(std::integral_constant<int, 11>{}, std::false_type{});
}

template <typename K>
void allow_lds(K kernel, uint32_t bytes) {
    if (bytes > 65536)
        XC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   bytes));
}

template <typename Tio>
void run_fftk(const family_args& a, const Tio* L, const Tio* R, Tio* O, hipStream_t s) {
    const uint32_t H = a.rows, W = a.cols;
    if (2 * H - 1 > (uint32_t)kSlab || 2 * W - 1 > (uint32_t)kSlab)
        fail(XC_ERR_UNSUPPORTED, "own-FFT method: rows and cols must be at most 2048");
    const int logPr = std::max(4, ilog2_ceil(2 * H - 1)), logPc = std::max(4, ilog2_ceil(2 * W - 1));
    const uint32_t Pr = 1u << logPr, Pc = 1u << logPc, Kc = Pc / 2 + 1;
    const uint32_t nl = a.family == XC_ONE_TO_ONE || a.family == XC_ONE_TO_MANY ? 1 : a.n;
    const uint32_t nr = a.family == XC_ONE_TO_ONE ? 1 : a.m;
    const uint32_t pairs = a.pairs;
    const double2* tw = twiddles();

    // Slab shape per kernel: transforms per block, and the variant (slab of 2048 values with
    // wave-owned transforms for P <= 512; 2048 block-wide for P = 1024, 2048; 4096 for 4096).
    const slab_shape v1 = shape_for(logPc, (H + 1) / 2);  // rows_fwd: 2T input rows per block
    const slab_shape vc = shape_for(logPr, Kc);           // cols: C columns per block
    const slab_shape v3 = shape_for(logPc, H);            // rows_inv: 2T output rows per block
    const int logT1 = v1.logT, logC = vc.logT, logT3 = v3.logT;
    const uint32_t g1 = (H + (2u << logT1) - 1) >> (logT1 + 1);
    const uint32_t cg = (Kc + (1u << logC) - 1) >> logC;
    const uint32_t g3 = (2 * H - 1 + (2u << logT3) - 1) >> (logT3 + 1);

    const size_t s_bytes = sizeof(double2) * (size_t)Kc * H;
    const size_t col_bytes = sizeof(double2) * (size_t)Kc * Pr;
    double2* SR = static_cast<double2*>(workspace(WS_FK_SR, s_bytes * nr, s));
    double2* SL = static_cast<double2*>(workspace(WS_FK_SL, s_bytes * nl, s));
    double2* LH = static_cast<double2*>(workspace(WS_FK_LHAT, col_bytes * nl, s));
    // Z is processed in chunks of pairs of at most ~2 GiB.
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(pairs, (size_t(2) << 30) / col_bytes));
    double2* Zw = static_cast<double2*>(workspace(WS_FK_Z, col_bytes * chunk, s));

    auto rows_fwd = [&](const Tio* X, uint32_t nmat, double2* S) {
        with_shape(v1, [&](auto LS, auto WV) {
            auto k = rows_fwd_kernel<Tio, decltype(LS)::value, decltype(WV)::value>;
            allow_lds(k, v1.lds);
            hipLaunchKernelGGL(k, dim3(nmat * g1), dim3(kFT), v1.lds, s, X, H, W, logPc, logT1, g1, S, tw);
        });
    };
    auto cols = [&](const double2* S, double2* Zout, int mode, uint32_t items, uint32_t q0) {
        with_shape(vc, [&](auto LS, auto WV) {
            auto k = cols_kernel<decltype(LS)::value, decltype(WV)::value>;
            allow_lds(k, vc.lds);
            hipLaunchKernelGGL(k, dim3(items * cg), dim3(kFT), vc.lds, s, S, LH, Zout, mode, a.family, a.n, a.m, q0,
                               H, Kc, logPr, logC, cg, tw);
        });
    };
    {
        timed t0("fftk_rows_fwd_left", s);
        rows_fwd(L, nl, SL);
    }
    {
        timed t1("fftk_cols_left", s);
        cols(SL, nullptr, 0, nl, 0);
    }
    {
        timed t2("fftk_rows_fwd", s);
        rows_fwd(R, nr, SR);
    }
    const double scale = 1.0 / ((double)Pr * (double)Pc);
    for (uint32_t q0 = 0; q0 < pairs; q0 += chunk) {
        const uint32_t nq = std::min(chunk, pairs - q0);
        {
            timed t3("fftk_cols", s);
            cols(SR, Zw, 1, nq, q0);
        }
        timed t4("fftk_rows_inv", s);
        with_shape(v3, [&](auto LS, auto WV) {
            auto k = rows_inv_kernel<Tio, decltype(LS)::value, decltype(WV)::value>;
            allow_lds(k, v3.lds);
            hipLaunchKernelGGL(k, dim3(nq * g3), dim3(kFT), v3.lds, s, (const double2*)Zw, O, q0, H, W, logPc, logPr,
                               logT3, g3, scale, tw);
        });
    }
    XC_HIP(hipGetLastError());
    record_launch("fftk_rows_inv", dim3(g3), dim3(kFT), v3.lds, H, W, pairs, 1);
}

}  // namespace

void run_fftk_f32(const family_args& a, const float* L, const float* R, float* O, hipStream_t s) {
    run_fftk<float>(a, L, R, O, s);
}

void run_fftk_f64(const family_args& a, const double* L, const double* R, double* O, hipStream_t s) {
    run_fftk<double>(a, L, R, O, s);
}

}  // namespace xc
