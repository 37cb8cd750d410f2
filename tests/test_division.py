"""Pin of the shared-reciprocal division of the implicit solve: with
y = RN(1/d), q0 = RN(x y), r = x - q0 d (one FMA), RN(q0 + r y) (one FMA)
equals the IEEE quotient RN(x/d) bit for bit (Markstein's theorem).  Checked
in C on random operands over a wide exponent range and on hard significands
(all-ones, powers of two, neighbours of 1), against the hardware division."""
import os
import subprocess
import textwrap

import pytest

SRC = textwrap.dedent(r"""
    #include <math.h>
    #include <stdint.h>
    #include <stdio.h>
    #include <string.h>
    static uint64_t s = 0x9E3779B97F4A7C15ull;
    static uint64_t nx(void) {  /* SplitMix64 */
      uint64_t z = (s += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    }
    static double mk(uint64_t mant, int e, int neg) {
      uint64_t b = ((uint64_t)(e + 1023) << 52) | (mant & 0xFFFFFFFFFFFFFull);
      double v; memcpy(&v, &b, 8); return neg ? -v : v;
    }
    static double dr(double x, double d, double y) {
      double q0 = x * y; double r = fma(-q0, d, x); return fma(r, y, q0);
    }
    int main(void) {
      long bad = 0, n = 0;
      const uint64_t hard[] = {0, 1, 2, 0xFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFEull,
                               0x8000000000000ull, 0x7FFFFFFFFFFFFull, 0x0000000000001ull};
      for (long t = 0; t < 8000000; ++t) {
        uint64_t mx = nx(), md = nx();
        if ((t & 7) == 1) md = hard[nx() & 7];
        if ((t & 7) == 2) mx = hard[nx() & 7];
        if ((t & 7) == 3) { md = hard[nx() & 7]; mx = hard[nx() & 7]; }
        int ex = (int)(nx() % 121) - 60, ed = (int)(nx() % 121) - 60;
        double x = mk(mx, ex, nx() & 1), d = mk(md, ed, nx() & 1);
        double y = 1.0 / d, a = x / d, b = dr(x, d, y);
        ++n;
        if (memcmp(&a, &b, 8) != 0) {
          if (bad < 5) printf("mismatch x=%a d=%a ieee=%a got=%a\n", x, d, a, b);
          ++bad;
        }
      }
      printf("%ld %ld\n", n, bad);
      return 0;
    }
""")


def test_shared_reciprocal_division_is_ieee(tmp_path):
    c = tmp_path / "div.c"
    exe = tmp_path / "div"
    c.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    n, bad = map(int, out.strip().splitlines()[-1].split())
    assert n == 8000000
    assert bad == 0, out
