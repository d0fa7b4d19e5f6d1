"""The FP64 exp of the point kernel (exp_t in csrc/sg_kernels.hip, Tang's
table-driven method) within 1 ulp of glibc's exp.

The 2^(j/64) table and the reduction constants are read from the kernel
source; a C program evaluates the same method with them (fma from libm) on
random arguments over the whole double range and counts ulps against exp().
A wrong table entry or constant shows up as errors of many ulps."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "senga-versiona_amd", "csrc", "sg_kernels.hip")

PROGRAM = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static const double TAB[64] = {%(tab)s};
static double exp_t(double x) {
  const double INV = %(inv)s, L1 = %(l1)s, L2 = %(l2)s;
  const double SH = 0x1.8p52;
  double tsh = fma(x, INV, SH);
  double n = tsh - SH;
  uint64_t bits;
  memcpy(&bits, &tsh, 8);
  int ni = (int)(uint32_t)bits;
  double r = fma(-n, L1, x);
  r = fma(-n, L2, r);
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  double p = fma(q, r * r, r);
  double tj = TAB[ni & 63];
  return ldexp(fma(tj, p, tj), ni >> 6);
}
static int64_t key(double a) { int64_t i; memcpy(&i, &a, 8); return i < 0 ? INT64_MIN - i : i; }
int main(void) {
  const double lo[] = {-745.0, -50.0, -1.0, -1e-3, 700.0}, hi[] = {709.7, 50.0, 1.0, 1e-3, 709.7};
  int64_t worst = 0;
  srand48(7);
  for (int k = 0; k < 5; k++)
    for (long i = 0; i < 400000; i++) {
      double x = lo[k] + (hi[k] - lo[k]) * drand48();
      int64_t d = key(exp_t(x)) - key(exp(x));
      if (d < 0) d = -d;
      if (d > worst) worst = d;
    }
  printf("%%lld %%.17g %%.17g %%.17g %%.17g %%.17g %%.17g %%.17g\n", (long long)worst, exp_t(0.0), exp_t(-800.0),
         exp_t(1.0), exp_t(709.8), exp_t(-1.0e7), exp_t(1.0e7), exp_t(-745.2));
  return 0;
}
"""


def _constants():
    s = open(SRC).read()
    tab = re.search(r"kExp2j\[64\] = \{([^}]*)\}", s).group(1)
    m = re.search(r"const double INV = (\S+), L1 = (\S+), L2 = (\S+);", s)
    return tab, m.group(1), m.group(2), m.group(3)


def test_exp_t_within_one_ulp(tmp_path):
    tab, inv, l1, l2 = _constants()
    assert len(tab.split(",")) == 64
    c = tmp_path / "exp_t.c"
    c.write_text(PROGRAM % {"tab": tab, "inv": inv, "l1": l1, "l2": l2})
    exe = tmp_path / "exp_t"
    r = subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", str(c), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    worst, e0, eu, e1 = int(out[0]), float(out[1]), float(out[2]), float(out[3])
    assert worst <= 1, f"exp_t is {worst} ulp from exp"
    assert e0 == 1.0 and eu == 0.0 and abs(e1 - 2.718281828459045) < 1e-15
    # the ends of the domain |x| < 2^31 ln2/64: overflow to inf, underflow to 0
    eo, em, ep, el = (float(v) for v in out[4:8])
    assert eo == float("inf") and ep == float("inf") and em == 0.0 and el == 0.0
