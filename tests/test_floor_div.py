"""scn_floor_div and scn_idiv (include/scn_fpmath.h), the division-free paths of the lattice index
(latUniverse_class.f90 findCell floor((r - corner) / pitch)), get_ijk's integer divisions
(latUniverse_class.f90:489-506) and the mesh-tally bin (spaceMap floor((x - min) / step)): they must
give exactly the reference's floor(fl(x / y)) and a / n.  A C probe compiled with the engine's FP
flags checks them against the plain operations on random values, on values within a few ulps of
every integer multiple of the pitch (where the fast path must fall back), on the pitches and steps
the configurations use, on 2**63-sized pitches of 2D lattices, and on every int32 range edge."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stdlib.h>
#include "scn_fpmath.h"

static uint64_t st = 0x9E3779B97F4A7C15ULL;
static uint64_t nx(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static double ur(void) { return (double)(nx() >> 11) * 0x1.0p-53; }

static long bad = 0, checked = 0, fell_back = 0, random_fell_back = 0;
static int in_random = 0;
static void chk(double x, double y) {
    double inv = 1.0 / y;
    double a = scn_floor_div(x, y, inv), b = floor(x / y);
    double t = x * inv, f = floor(t);
    if (!(t - f > 4.0e-15 * fabs(t) && (f + 1.0) - t > 4.0e-15 * fabs(t))) {
        fell_back++;
        random_fell_back += in_random;
    }
    checked++;
    if (a != b || signbit(a) != signbit(b)) {
        if (bad < 5) printf("floor_div mismatch x=%a y=%a got %a want %a\n", x, y, a, b);
        bad++;
    }
}

int main(void) {
    const double pitches[] = {1.26, 21.42, 0.54, 2.0 * 10.71, 10.71 / 3.0, 0.1, 7.0 / 17.0, 1e-3, 123.456,
                              2.0 * 9223372036854775808.0, 1.0, 0.5};
    const int np = sizeof pitches / sizeof pitches[0];
    for (int k = 0; k < np; ++k) {
        double y = pitches[k];
        for (int n = -400; n <= 400; ++n) {
            double c = n * y;
            double x = c;
            chk(x, y);
            for (int s = 0; s < 8; ++s) { x = nextafter(x, 1e300); chk(x, y); }
            x = c;
            for (int s = 0; s < 8; ++s) { x = nextafter(x, -1e300); chk(x, y); }
            chk(c + 0.5 * y, y);
            chk(-c, -y);
        }
    }
    long n_random = 0;
    for (long i = 0; i < 2000000; ++i) {
        double y = exp((ur() - 0.5) * 20.0) * (nx() & 1 ? 1.0 : -1.0);
        double x = exp((ur() - 0.5) * 30.0) * (nx() & 1 ? 1.0 : -1.0);
        in_random = 1;
        chk(x, y);
        chk(ur() * 1e5 * y, y);
        in_random = 0;
        n_random += 2;
        chk(0.0, y);
        chk(-0.0, y);
    }
    long ibad = 0;
    for (int n = 1; n <= 300; ++n)
        for (int a = 0; a < 200000; ++a)
            if (scn_idiv(a, 1.0 / n) != a / n) ibad++;
    for (long i = 0; i < 2000000; ++i) {
        int n = 1 + (int)(nx() % 2147483647ULL);
        if (i & 1) n = 1 + (int)(nx() % 100000);
        int a = (int)(nx() % 2147483648ULL);
        if (scn_idiv(a, 1.0 / n) != a / n) ibad++;
    }
    const int edges[] = {0, 1, 2, 2147483646, 2147483647};
    for (int i = 0; i < 5; ++i)
        for (int j = 1; j < 5; ++j)
            if (scn_idiv(edges[i], 1.0 / edges[j]) != edges[i] / edges[j]) ibad++;
    printf("%ld %ld %ld %ld %ld %ld\n", checked, bad, fell_back, ibad, random_fell_back, n_random);
    return 0;
}
"""


def test_division_free_floor_and_integer_division_are_exact(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = str(tmp_path / "probe")
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip().split("\n")
    checked, bad, fell_back, ibad, random_fell_back, n_random = (int(x) for x in out[-1].split())
    assert bad == 0, out
    assert ibad == 0, out
    assert checked > 8_000_000
    # the values next to integer multiples force the division; random ones almost never do
    assert fell_back > 10_000
    assert random_fell_back < n_random // 1000
