"""The division by a known reciprocal that the tapered guide's bounce uses
(mcv_core.h div_recip: q0 = a y, two corrections with the exact remainder,
Markstein's theorem) against IEEE division rounded to nearest, on random
operands over the guides' ranges and on hard cases (divisors with all-ones or
nearly-zero significands, quotients near powers of two).  The checker is the
host's own division; the sequence is compiled with gcc (FMA from libm, no
contraction)."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s0 = 0x243F6A8885A308D3ull, s1 = 0x13198A2E03707344ull;
static uint64_t nx(void) { uint64_t a = s0, b = s1, r = a + b; b ^= a;
  s0 = ((a << 55) | (a >> 9)) ^ b ^ (b << 14); s1 = (b << 36) | (b >> 28); return r; }
static double mk(int e, uint64_t m) { uint64_t u = ((uint64_t)(e + 1023) << 52) | (m & ((1ull << 52) - 1));
  double d; memcpy(&d, &u, 8); return d; }
static double div_recip(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = fma(fma(-b, q0, a), y, q0);
  return fma(fma(-b, q1, a), y, q1);
}
int main(int argc, char** argv) {
  long n = atol(argv[1]), bad = 0;
  for (long i = 0; i < n; ++i) {
    uint64_t r = nx(), ma = nx(), mb = nx();
    int ea = (int)(r % 120) - 60, eb = (int)((r >> 8) % 40) - 15;
    switch ((r >> 16) % 6) {
      case 0: mb |= (1ull << 52) - 1; break;            /* all-ones significand */
      case 1: mb &= 0xffull; break;                     /* divisor just above a power of two */
      case 2: ma = mb; break;                           /* quotient near a power of two */
      case 3: mb = (1ull << 52) - 1 - ((r >> 20) & 0xffff); break;
      default: break;
    }
    double a = mk(ea, ma), b = mk(eb, mb);
    if (r & (1ull << 40)) a = -a;
    if (r & (1ull << 41)) b = -b;
    const double q = a / b, y = 1.0 / b, got = div_recip(a, b, y);
    if (memcmp(&q, &got, 8) != 0) { if (bad < 5) printf("a=%a b=%a q=%a got=%a\n", a, b, q, got); ++bad; }
  }
  printf("%ld %ld\n", n, bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_recip_is_correctly_rounded(tmp_path):
    src = tmp_path / "div.c"
    src.write_text(SRC)
    exe = tmp_path / "div"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    n, bad = map(int, r.stdout.split()[-2:])
    assert n == 20000000 and bad == 0


def test_device_helper_matches_this_sequence():
    """The device helper is the sequence checked above."""
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = open(os.path.join(here, "acc_amd", "csrc", "device", "mcv_core.h")).read()
    body = h[h.index("double div_recip(double a, double b, double y) {"):]
    body = body[:body.index("}") + 1]
    assert "const double q0 = a * y;" in body
    assert "const double q1 = fma(fma(-b, q0, a), y, q0);" in body
    assert "return fma(fma(-b, q1, a), y, q1);" in body
