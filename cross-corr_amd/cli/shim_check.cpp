// shim_check -- calls every launcher of include/xcorr_kernels.hpp (the C++ re-creation of the
// reference's src/kernels.cuh interface) the way code written against kernels.cuh calls it,
// on seeded inputs, and writes inputs and results as raw little-endian files for a checker:
//
//   shim_check <outdir>
//
// <outdir>/manifest.txt gets one line per case:
//   <case> corr <family> <rows> <cols> <n> <m> <f32|f64>    files <case>.left .right .out
//   <case> scatter <src_rows> <src_cols> <num> <dst_rows> <dst_cols> <dst_row> <dst_col>   .src .out
//   <case> hadamard <count> <batch> <f32|f64>               files <case>.ref .def .out
//   <case> hadamard_n_to_m <count> <n> <m> <f32|f64>        files <case>.left .right .out
//   <case> rejected <message>                               (a tuning value the launcher refused)
// Exit status 0 when every launcher ran (rejections included as expected), 1 otherwise.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xcorr_kernels.hpp"

namespace {

using cross::dsize2_t;
using cross::dsize_t;

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Deterministic inputs in [-1, 1): a 64-bit LCG.
template <typename T>
std::vector<T> seeded(size_t count, uint64_t seed) {
    std::vector<T> v(count);
    uint64_t s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (auto& x : v) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        x = (T)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    }
    return v;
}

template <typename T>
struct dev {
    T* p = nullptr;
    size_t n = 0;
    explicit dev(size_t count) : n(count) { hip_ok(hipMalloc(&p, count * sizeof(T)), "hipMalloc"); }
    explicit dev(const std::vector<T>& h) : dev(h.size()) {
        hip_ok(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    ~dev() { (void)hipFree(p); }
    std::vector<T> host() const {
        std::vector<T> h(n);
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_ok(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
};

struct writer {
    std::string dir;
    std::ofstream manifest;
    explicit writer(const std::string& d) : dir(d), manifest(d + "/manifest.txt") {
        if (!manifest) throw std::runtime_error("cannot write " + d + "/manifest.txt");
    }
    template <typename T>
    void file(const std::string& name, const std::vector<T>& v) {
        std::ofstream f(dir + "/" + name, std::ios::binary);
        f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
        if (!f) throw std::runtime_error("cannot write " + name);
    }
};

template <typename T>
const char* tname() {
    return sizeof(T) == 4 ? "f32" : "f64";
}

int64_t results(const std::string& family, dsize_t n, dsize_t m) {
    return family == "n_to_m" ? (int64_t)n * m : (int64_t)m;
}

// One correlation case: inputs of n left and m right rows x cols matrices, `run` launches.
template <typename T, typename F>
void corr(writer& w, const std::string& name, const std::string& family, dsize_t rows, dsize_t cols, dsize_t n,
          dsize_t m, uint64_t seed, F run) {
    const auto l = seeded<T>((size_t)n * rows * cols, seed);
    const auto r = seeded<T>((size_t)m * rows * cols, seed + 1);
    dev<T> dl(l), dr(r), dout((size_t)results(family, n, m) * (2 * rows - 1) * (2 * cols - 1));
    run(dl.p, dr.p, dout.p, dsize2_t{cols, rows}, dsize2_t{2 * cols - 1, 2 * rows - 1});
    w.file(name + ".left", l);
    w.file(name + ".right", r);
    w.file(name + ".out", dout.host());
    w.manifest << name << " corr " << family << " " << rows << " " << cols << " " << n << " " << m << " "
               << tname<T>() << "\n";
}

// A launch the launcher must refuse with std::runtime_error.
template <typename F>
bool rejected(writer& w, const std::string& name, F run) {
    try {
        run();
    } catch (const std::runtime_error& e) {
        w.manifest << name << " rejected " << e.what() << "\n";
        return true;
    }
    std::cerr << name << ": a refused tuning value was accepted\n";
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "Usage: " << argv[0] << " <outdir>\n";
        return 2;
    }
    using namespace cross;
    try {
        writer w(argv[1]);
        bool ok = true;

        // ---- definition-based launchers, float
        corr<float>(w, "naive_original", "n_to_mn", 17, 12, 2, 6, 11, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_cross_corr_naive_original(l, r, o, ms, ss, 2, 3);
        });
        corr<float>(w, "shuffle", "one_to_one", 17, 23, 1, 1, 13, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle(l, r, o, ms, ss, 8);
        });
        corr<float>(w, "shuffle_wd_triangle", "one_to_one", 31, 20, 1, 1, 15,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_work_distribution<triangle_distribution>(l, r, o, ms, ss, 8, 5);
                    });
        corr<float>(w, "shuffle_wd_rectangle", "one_to_one", 31, 20, 1, 1, 17,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_work_distribution<rectangle_distribution>(l, r, o, ms, ss, 4, 7);
                    });
        corr<float>(w, "shuffle_wd_none", "one_to_one", 31, 20, 1, 1, 19, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle_work_distribution<no_distribution>(l, r, o, ms, ss, 4, 1);
        });
        corr<float>(w, "multimat_right", "one_to_many", 21, 70, 1, 5, 21, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle_multimat_right(l, r, o, ms, ss, 5, 8, 3);
        });
        corr<float>(w, "warp_per_shift", "one_to_one", 14, 9, 1, 1, 23, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_warp_per_shift(l, r, o, ms, ss, 16);
        });
        corr<float>(w, "warp_per_shift_simple", "one_to_one", 14, 9, 1, 1, 25,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_warp_per_shift_simple_indexing(l, r, o, ms, ss, 8);
                    });
        corr<float>(w, "warp_per_shift_wd", "one_to_one", 40, 9, 1, 1, 27, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_warp_per_shift_work_distribution<triangle_distribution>(l, r, o, ms, ss, 8, 6);
        });
        corr<float>(w, "multimat_right_wd", "one_to_many", 26, 18, 1, 4, 29,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_multimat_right_work_distribution<rectangle_distribution>(l, r, o, ms, ss, 4, 8,
                                                                                                2, 5);
                    });
        corr<float>(w, "warp_per_shift_shared_mem", "one_to_many", 16, 16, 1, 3, 31,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_warp_per_shift_shared_mem(l, r, o, ms, ss, 3, 8, 64, 4, 2, true, false);
                    });
        corr<float>(w, "n_to_m_multimat_both_wd", "n_to_m", 19, 25, 3, 5, 33,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_n_to_m_multimat_both_work_distribution<triangle_distribution>(
                            l, r, o, ms, ss, 3, 5, 8, 2, 3, 4);
                    });
        corr<float>(w, "local_mem_n_to_m_multimat_both_wd", "n_to_m", 19, 25, 3, 5, 35,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        local_mem::run_ccn_shuffle_n_to_m_multimat_both_work_distribution<no_distribution>(
                            l, r, o, ms, ss, 3, 5, 8, 4, 4, 4);
                    });
        corr<float>(w, "multirow_right", "one_to_one", 33, 29, 1, 1, 37, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle_multirow_right(l, r, o, ms, ss, 8, 3);
        });
        corr<float>(w, "multirow_both", "one_to_one", 33, 29, 1, 1, 39, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle_multirow_both(l, r, o, ms, ss, 8, 4, 2);
        });
        corr<float>(w, "local_mem_multirow_both", "one_to_one", 33, 29, 1, 1, 41,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        local_mem::run_ccn_shuffle_multirow_both(l, r, o, ms, ss, 4, 8, 4);
                    });
        corr<float>(w, "block_per_shift", "one_to_one", 12, 30, 1, 1, 43, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_block_per_shift(l, r, o, ms, ss, 256);
        });
        corr<float>(w, "multirow_right_multimat_right", "one_to_many", 27, 34, 1, 5, 45,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_multirow_right_multimat_right(l, r, o, ms, ss, 5, 8, 2, 3);
                    });
        corr<float>(w, "one_to_many_multirow_both_multimat_right", "one_to_many", 40, 24, 1, 3, 47,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_shuffle_one_to_many_multirow_both_multimat_right(l, r, o, ms, ss, 3, 8, 2, 2, 2);
                    });
        corr<float>(w, "n_to_m_multirow_both_multimat_both", "n_to_m", 22, 31, 3, 5, 49,
                    [](auto l, auto r, auto o, auto ms, auto ss) {
                        run_ccn_n_to_m_shuffle_multirow_both_multimat_both(l, r, o, ms, ss, 3, 5, 8, 2, 2, 2, 2);
                    });

        // ---- the same launchers with double data
        corr<double>(w, "shuffle_f64", "one_to_one", 17, 23, 1, 1, 51, [](auto l, auto r, auto o, auto ms, auto ss) {
            run_ccn_shuffle(l, r, o, ms, ss, 8);
        });
        corr<double>(w, "multirow_both_f64", "one_to_one", 25, 19, 1, 1, 53,
                     [](auto l, auto r, auto o, auto ms, auto ss) {
                         run_ccn_shuffle_multirow_both(l, r, o, ms, ss, 8, 8, 4);
                     });
        corr<double>(w, "one_to_many_multirow_both_multimat_right_f64", "one_to_many", 40, 24, 1, 6, 55,
                     [](auto l, auto r, auto o, auto ms, auto ss) {
                         run_ccn_shuffle_one_to_many_multirow_both_multimat_right(l, r, o, ms, ss, 6, 8, 4, 4, 4);
                     });
        corr<double>(w, "n_to_m_multirow_both_multimat_both_f64", "n_to_m", 22, 31, 3, 5, 57,
                     [](auto l, auto r, auto o, auto ms, auto ss) {
                         run_ccn_n_to_m_shuffle_multirow_both_multimat_both(l, r, o, ms, ss, 3, 5, 16, 4, 4, 4, 4);
                     });
        corr<double>(w, "naive_original_f64", "n_to_mn", 9, 14, 3, 6, 59,
                     [](auto l, auto r, auto o, auto ms, auto ss) {
                         run_cross_corr_naive_original(l, r, o, ms, ss, 3, 2);
                     });

        // ---- FFT helpers
        {
            const dsize_t sr = 5, sc = 7, num = 2, dr = 10, dc = 14;
            const auto src = seeded<float>((size_t)num * sr * sc, 61);
            dev<float> ds(src), dd((size_t)num * dr * dc);
            hip_ok(hipMemset(dd.p, 0x7f, dd.n * sizeof(float)), "hipMemset");  // the kernel writes the padding
            run_scatter(ds.p, dd.p, dsize2_t{sc, sr}, num, dsize2_t{dc, dr}, dsize2_t{2, 1}, 256, 10);
            w.file("scatter.src", src);
            w.file("scatter.out", dd.host());
            w.manifest << "scatter scatter " << sr << " " << sc << " " << num << " " << dr << " " << dc << " 1 2\n";
        }
        {
            const dsize2_t sub{6, 4};
            const dsize_t spp = 3, batch = 5, count = sub.area() * spp;
            const auto ref = seeded<float>(2 * (size_t)count, 63);
            const auto def = seeded<float>(2 * (size_t)count * batch, 64);
            dev<float2> dref(count), ddef((size_t)count * batch);
            hip_ok(hipMemcpy(dref.p, ref.data(), ref.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
            hip_ok(hipMemcpy(ddef.p, def.data(), def.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
            run_hadamard_original(dref.p, ddef.p, sub, spp, batch, 128);
            const auto out = ddef.host();
            w.file("hadamard.ref", ref);
            w.file("hadamard.def", def);
            w.file("hadamard.out", out);
            w.manifest << "hadamard hadamard " << count << " " << batch << " f32\n";
        }
        for (int over_output = 0; over_output < 2; ++over_output) {
            const dsize2_t size{9, 4};
            const dsize_t n = 3, m = 4, count = size.area();
            const auto l = seeded<double>(2 * (size_t)count * n, 65 + over_output);
            const auto r = seeded<double>(2 * (size_t)count * m, 67 + over_output);
            dev<double2> dl(count * n), dr(count * m), dout((size_t)count * n * m);
            hip_ok(hipMemcpy(dl.p, l.data(), l.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            hip_ok(hipMemcpy(dr.p, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            if (over_output)
                run_hadamard_n_to_m_over_output(dl.p, dr.p, dout.p, size, n, m, 256, 2);
            else
                run_hadamard_n_to_m_over_right(dl.p, dr.p, dout.p, size, n, m, 256, 2);
            const std::string name = over_output ? "hadamard_n_to_m_over_output" : "hadamard_n_to_m_over_right";
            w.file(name + ".left", l);
            w.file(name + ".right", r);
            w.file(name + ".out", dout.host());
            w.manifest << name << " hadamard_n_to_m " << count << " " << n << " " << m << " f64\n";
        }

        // ---- tuning values the reference's release build refuses
        {
            dev<float> a(64), o(15 * 15);
            const dsize2_t ms{8, 8}, ss{15, 15};
            ok &= rejected(w, "reject_warps", [&] { run_ccn_shuffle(a.p, a.p, o.p, ms, ss, 33); });
            ok &= rejected(w, "reject_right_mats", [&] { run_ccn_shuffle_multimat_right(a.p, a.p, o.p, ms, ss, 1, 8, 9); });
            ok &= rejected(w, "reject_shifts", [&] { run_ccn_shuffle_multirow_both(a.p, a.p, o.p, ms, ss, 8, 9, 2); });
            ok &= rejected(w, "reject_left_rows", [&] { run_ccn_shuffle_multirow_both(a.p, a.p, o.p, ms, ss, 8, 2, 5); });
            ok &= rejected(w, "reject_one_to_many_shifts", [&] {
                run_ccn_shuffle_one_to_many_multirow_both_multimat_right(a.p, a.p, o.p, ms, ss, 1, 8, 5, 2, 2);
            });
            ok &= rejected(w, "reject_n_to_m_left_mats", [&] {
                run_ccn_n_to_m_shuffle_multirow_both_multimat_both(a.p, a.p, o.p, ms, ss, 1, 1, 8, 2, 5, 2, 2);
            });
            ok &= rejected(w, "reject_rows_per_thread",
                           [&] { run_ccn_shuffle_work_distribution<triangle_distribution>(a.p, a.p, o.p, ms, ss, 8, 0); });
            ok &= rejected(w, "reject_search_size", [&] { run_ccn_shuffle(a.p, a.p, o.p, ms, dsize2_t{13, 15}, 8); });
            ok &= rejected(w, "reject_block", [&] { run_ccn_block_per_shift(a.p, a.p, o.p, ms, ss, 2048); });
        }
        w.manifest.flush();
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "shim_check: " << e.what() << "\n";
        return 1;
    }
}
