"""Order-independent tally accumulation (scn_fix_* in include/scn_fpmath.h): scores turned into
128-bit fixed point and added as four 32-bit limbs give the same words in any order, and the value
read back is the exact sum of the scores within the truncation (2^-64 per score) and 2 ulp."""
import os
import subprocess
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stdlib.h>
#include "scn_fpmath.h"
int main(int argc, char** argv) {
    FILE* f = fopen(argv[1], "rb");
    long n = atol(argv[2]);
    double* v = malloc(sizeof(double) * n);
    if (fread(v, sizeof(double), n, f) != (size_t)n) return 1;
    fclose(f);
    for (int pass = 0; pass < 3; ++pass) {
        uint64_t w[4] = {0, 0, 0, 0};
        for (long k = 0; k < n; ++k) {
            long i = pass == 0 ? k : pass == 1 ? n - 1 - k : (k * 7919) % n;   /* 7919 prime, n not a multiple */
            uint64_t l[4];
            scn_fix_limbs(scn_fix_of(v[i]), l);
            for (int j = 0; j < 4; ++j) w[j] += l[j];
        }
        printf("%llu %llu %llu %llu %.17g\n", (unsigned long long)w[0], (unsigned long long)w[1],
               (unsigned long long)w[2], (unsigned long long)w[3], scn_fix_value(w));
    }
    return 0;
}
"""


def _run(tmp_path, vals):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = str(tmp_path / "probe")
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    exe, "-lm"], check=True)
    data = tmp_path / "v.bin"
    np.asarray(vals, dtype=np.float64).tofile(data)
    out = subprocess.run([exe, str(data), str(len(vals))], capture_output=True, text=True, check=True).stdout
    rows = [l.split() for l in out.strip().split("\n")]
    return [tuple(int(x) for x in r[:4]) for r in rows], [float(r[4]) for r in rows]


def test_fixed_point_sums_are_order_independent_and_exact(tmp_path):
    rng = np.random.default_rng(11)
    n = 200_003
    vals = np.exp(rng.uniform(np.log(1e-6), np.log(1e4), n))
    vals[rng.random(n) < 0.05] *= -1.0
    vals[:3] = [0.0, 1.0, 2.0 ** -70]
    words, values = _run(tmp_path, vals)
    assert words[0] == words[1] == words[2]
    assert values[0] == values[1] == values[2]
    exact = sum(Fraction(float(x)) for x in vals)
    err = abs(Fraction(values[0]) - exact)
    bound = Fraction(n, 2 ** 64) + 2 * Fraction(np.spacing(abs(values[0])))
    assert err <= bound, (float(err), float(bound))


def test_fixed_point_handles_cancellation_and_large_counts(tmp_path):
    # equal and opposite scores cancel to an exact zero; many small scores keep their full sum
    vals = [3.25, -3.25, 1e-3, -1e-3] * 1000 + [0.1] * 100_000
    words, values = _run(tmp_path, vals)
    exact = Fraction(0.1) * 100_000
    assert abs(Fraction(values[0]) - exact) <= Fraction(100_000, 2 ** 64) + 2 * Fraction(np.spacing(values[0]))
