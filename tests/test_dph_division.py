"""The far-regime kernel forms dphrec's quotients XX / (4 LU^2 - 1) as
q0 = XX c, q0 + (XX - q0 d) c with an FMA-exact residual and c = RN(1/d).
This checks, with the C compiler's IEEE fma on the host, that the sequence
gives the correctly rounded quotient for every divisor and sampled x."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
int main(void) {
    uint64_t s = 88172645463325252ull;
    long bad = 0, n = 0;
    for (int it = 0; it < 200000; ++it) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        double u = (double)(s >> 11) / 9007199254740992.0;
        double x = 70.0 + 5000.0 * u * u;
        double XX = x * x;
        for (int LU = 2; LU <= 99; ++LU) {
            double d = 4.0 * (double)LU * (double)LU - 1.0;
            double c = 1.0 / d;
            double q0 = XX * c;
            double q1 = fma(fma(-q0, d, XX), c, q0);
            ++n;
            if (q1 != XX / d) ++bad;
        }
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_fma_corrected_quotient_is_correctly_rounded(tmp_path):
    src = tmp_path / "q.c"
    src.write_text(SRC)
    exe = tmp_path / "q"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    n, bad = map(int, subprocess.run([str(exe)], check=True, capture_output=True,
                                     text=True).stdout.split())
    assert n == 200000 * 98 and bad == 0
