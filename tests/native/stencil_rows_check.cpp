// Host check: every derivative row in the uniform 5-term form used by the
// bounded direction pass equals the row-by-row form (d1/d2 of sg_stencil.h),
// for all 11 row kinds, on arbitrary data and distinct coefficients.
// Built with hipcc for the host only; exits non-zero on a mismatch.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#undef __device__
#define __device__
#undef __constant__
#define __constant__
#define __builtin_amdgcn_sbfe(v, off, w) (((int)((unsigned)(v) << (32 - (off) - (w)))) >> (32 - (w)))
#include "sg_stencil.h"
using namespace sg;

int main() {
  Params P{};
  for (int k = 0; k < 5; k++) { P.ac[k] = 1.1 + k; P.sc[k] = 2.3 + k; P.s1[k] = 3.7 + k; P.s2[k] = 4.1 + k; }
  for (int k = 0; k < 4; k++) { P.c1[k] = 5.3 + k; P.c2[k] = 6.7 + k; P.c5[k] = 7.1 + k; P.s5[k] = 8.9 + k; }
  for (int k = 0; k < 3; k++) { P.c4[k] = 9.2 + k; P.s4[k] = 10.1 + k; }
  for (int k = 0; k < 2; k++) { P.c3[k] = 11.3 + k; P.s3[k] = 12.7 + k; }
  P.ovdel[0] = 0.7; P.ovdl2[0] = 0.49;
  double fv[17];
  for (int q = 0; q < 17; q++) fv[q] = std::sin(0.37 * q + 0.1) * 3.0 + q * 0.01;
  const int modes[11] = {0, 1, 2, 3, 4, 5, 11, 12, 13, 14, 15};
  int bad = 0;
  for (int mode : modes) {
    Sten S;
    S.mode = mode;
    auto at = [&](int kk) { return fv[kk]; };
    const double a1 = d1<true>(P, S, 0, at), a2 = d2<true>(P, S, 0, at);
    const int row = row_of_mode(mode);
    double w[10];
    row_weights(P, row, w);
    auto fq = [&](int q) { return fv[q + 8]; };
    auto wq = [&](int k) { return w[k]; };
    const double b1 = d1_rows(P, c_row_offsets.w[row], 0, wq, fq);
    const double b2 = d2_rows(P, c_row_offsets.w[row], 0, wq, fq);
    const bool ok = a1 == b1 && a2 == b2;
    std::printf("mode %2d row %2d d1 %.17g %.17g d2 %.17g %.17g %s\n", mode, row, a1, b1, a2, b2, ok ? "ok" : "MISMATCH");
    bad += !ok;
  }
  // mixed derivatives: every row kind against dmix<true> on a 2-D field
  for (int k = 0; k < 5; k++) P.mc[k] = 0.3 + 0.7 * k;
  for (int k = 0; k < 4; k++) { P.m1[k] = 1.9 + k; P.m2[k] = 2.9 - k; P.m5[k] = 0.11 * (k + 1); }
  for (int k = 0; k < 3; k++) P.m4[k] = 1.3 - 0.2 * k;
  for (int k = 0; k < 2; k++) { P.m3[k] = 0.77 + k; P.my1[k] = 0.66 - 0.5 * k; }
  P.ovdel[1] = 1.3;
  double gv[17][17];
  for (int a = 0; a < 17; a++)
    for (int b = 0; b < 17; b++) gv[a][b] = std::cos(0.21 * a + 0.1) * std::sin(0.33 * b + 0.2) + 0.01 * a * b;
  for (int mode : modes) {
    Sten S;
    S.mode = mode;
    auto g = [&](int a, int b) { return gv[a][b]; };
    const double a3 = dmix<true>(P, S, 0, 1, g);
    const int row = row_of_mode(mode);
    double w[7];
    mix_row_weights(P, row, w);
    auto gq = [&](int a, int b) { return gv[a + 8][b + 8]; };
    auto wq = [&](int k) { return w[k]; };
    const double b3 = dmix_rows(P, c_mix_row_offsets.w[row], 0, 1, wq, gq);
    const bool ok = a3 == b3;
    std::printf("mixed mode %2d row %2d %.17g %.17g %s\n", mode, row, a3, b3, ok ? "ok" : "MISMATCH");
    bad += !ok;
  }
  return bad != 0;
}
