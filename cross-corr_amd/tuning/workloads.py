"""Seeded random workloads for the cost model's timing table.

A workload is (family, rows, cols, n, m).  Sizes are powers of two from 16 to 1024 for one
workload in three, else log-uniform in [16, 1024] (square three times in four); batch counts
are log-uniform per family, and the total work is capped
at 2e12 multiply-adds (and 512 results) so that every candidate method, the FP32 direct
kernel included, can be timed in well under a second.
"""
import math
import random

FAMILIES = ("one_to_one", "one_to_many", "n_to_mn", "n_to_m")
MAX_MACS = 2e12
MAX_RESULTS = 512
LARGE_MAX_MACS = 2e14      # large workloads: only the FFT methods are timed
LARGE_MAX_OUT_BYTES = 4 << 30


def num_results(family, n, m):
    return {"one_to_one": 1, "one_to_many": m, "n_to_mn": m, "n_to_m": n * m}[family]


def _logu(rng, lo, hi):
    return int(round(2 ** rng.uniform(math.log2(lo), math.log2(hi))))


def random_workloads(count, seed, large=False):
    """`large`: sizes 64..1024 and batches up to LARGE_MAX_MACS multiply-adds (and 4 GiB of
    results), for which only the FFT methods are timed."""
    rng = random.Random(seed)
    lo = 64 if large else 16
    max_macs = LARGE_MAX_MACS if large else MAX_MACS
    out = []
    for _ in range(count):
        fam = FAMILIES[rng.randrange(4)]
        if rng.random() < 1 / 3:
            rows = cols = 2 ** rng.randint(6 if large else 4, 10)
        elif rng.random() < 0.75:
            rows = cols = _logu(rng, lo, 1024)
        else:
            rows, cols = _logu(rng, lo, 1024), _logu(rng, lo, 1024)
        if fam == "one_to_one":
            n, m = 1, 1
        elif fam == "one_to_many":
            n, m = 1, _logu(rng, 1, 256 if not large else 512)
        elif fam == "n_to_mn":
            n = _logu(rng, 1, 16)
            m = n * _logu(rng, 1, 8)
        else:
            n, m = _logu(rng, 1, 32), _logu(rng, 1, 64)
        per_pair = float(rows * cols) ** 2
        out_bytes = 4.0 * (2 * rows - 1) * (2 * cols - 1)
        while num_results(fam, n, m) > 1 and (num_results(fam, n, m) * per_pair > max_macs
                                                or num_results(fam, n, m) > MAX_RESULTS
                                                or num_results(fam, n, m) * out_bytes > LARGE_MAX_OUT_BYTES):
            if fam == "n_to_mn":
                if m > n:
                    m = n * max(1, (m // n) // 2)
                else:
                    n = m = max(1, n // 2)
            elif fam == "n_to_m" and n > m:
                n = max(1, n // 2)
            else:
                m = max(1, m // 2)
        out.append((fam, rows, cols, n, m))
    return out
