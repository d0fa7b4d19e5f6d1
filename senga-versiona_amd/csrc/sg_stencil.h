// sg_stencil.h -- derivative operators of the SENGA+ scheme as device functions.
//
// Interior: explicit 10th-order central first and second derivatives and the
// diagonal-only mixed derivative (src/dfbydx.f:68-110, src/d2fdx2.f:68-110,
// src/d2fdxy.f:74-147).  Bounded ends: the 4th one-sided, 4th skewed, 4th, 6th,
// 8th order rows of src/dfbydx.f:116-228, src/d2fdx2.f:116-260 and the x-end
// rows of src/d2fdxy.f:151-391.  Sums are formed left to right and scaled by
// 1/dx (1/dx^2) afterwards, as the reference does.
//
// A point's neighbours along one direction are described by Sten: element
// offsets for q = -8..8 (periodic directions wrap; bounded directions only
// reach |q| <= 8 on the closure rows, which stay inside the grid) and the row
// kind (0 interior, 1..5 left rows m = 0..4, 11..15 right rows L-m = 0..4).
#pragma once

#include "sg_internal.h"

namespace sg {

struct Sten {
  int o[17];
  int mode;
};

template <bool BND>
__device__ __forceinline__ Sten make_sten(int m, int n, int s) {
  Sten S;
#pragma unroll
  for (int q = -8; q <= 8; q++) {
    int mq = m + q;
    if (!BND) {
      mq = mq < 0 ? mq + n : mq;
      mq = mq >= n ? mq - n : mq;
    }
    S.o[q + 8] = (mq - m) * s;
  }
  S.mode = 0;
  if (BND) {
    if (m < 5) S.mode = 1 + m;
    else if (m > n - 6) S.mode = 11 + (n - 1 - m);
  }
  return S;
}

// f(k) returns the value at offset index k (q = k - 8).
#define FQ(q) f((q) + 8)

template <bool BND, class Fn>
__device__ __forceinline__ double d1(const Params &P, const Sten &S, int dir, Fn f) {
  double r;
  if (!BND || S.mode == 0) {
    double fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3), fd = FQ(4) - FQ(-4),
           fe = FQ(5) - FQ(-5);
    r = P.ac[0] * fa + P.ac[1] * fb + P.ac[2] * fc + P.ac[3] * fd + P.ac[4] * fe;
  } else {
    switch (S.mode) {
      case 1: {
        double fa = FQ(1) - FQ(0), fb = FQ(2) - FQ(0), fc = FQ(3) - FQ(0), fd = FQ(4) - FQ(0);
        r = P.c1[0] * fa + P.c1[1] * fb + P.c1[2] * fc + P.c1[3] * fd;
      } break;
      case 2: {
        double fa = FQ(-1) - FQ(0), fb = FQ(1) - FQ(0), fc = FQ(2) - FQ(0), fd = FQ(3) - FQ(0);
        r = P.c2[0] * fa + P.c2[1] * fb + P.c2[2] * fc + P.c2[3] * fd;
      } break;
      case 3:
      case 13: {
        double fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2);
        r = P.c3[0] * fa + P.c3[1] * fb;
      } break;
      case 4:
      case 14: {
        double fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3);
        r = P.c4[0] * fa + P.c4[1] * fb + P.c4[2] * fc;
      } break;
      case 5:
      case 15: {
        double fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3), fd = FQ(4) - FQ(-4);
        r = P.c5[0] * fa + P.c5[1] * fb + P.c5[2] * fc + P.c5[3] * fd;
      } break;
      case 12: {
        double fa = FQ(0) - FQ(1), fb = FQ(0) - FQ(-1), fc = FQ(0) - FQ(-2), fd = FQ(0) - FQ(-3);
        r = P.c2[0] * fa + P.c2[1] * fb + P.c2[2] * fc + P.c2[3] * fd;
      } break;
      default: {  // 11: last point
        double fa = FQ(0) - FQ(-1), fb = FQ(0) - FQ(-2), fc = FQ(0) - FQ(-3), fd = FQ(0) - FQ(-4);
        r = P.c1[0] * fa + P.c1[1] * fb + P.c1[2] * fc + P.c1[3] * fd;
      } break;
    }
  }
  return r * P.ovdel[dir];
}

template <bool BND, class Fn>
__device__ __forceinline__ double d2(const Params &P, const Sten &S, int dir, Fn f) {
  double r;
  if (!BND || S.mode == 0) {
    double c = FQ(0);
    double ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
           cm = c - FQ(-3), dp = FQ(4) - c, dm = c - FQ(-4), ep = FQ(5) - c, em = c - FQ(-5);
    r = P.sc[0] * (ap - am) + P.sc[1] * (bp - bm) + P.sc[2] * (cp - cm) + P.sc[3] * (dp - dm) +
        P.sc[4] * (ep - em);
  } else {
    double c = FQ(0);
    switch (S.mode) {
      case 1: {
        double ap = FQ(1) - c, bp = FQ(2) - c, cp = FQ(3) - c, dp = FQ(4) - c, ep = FQ(5) - c;
        r = P.s1[0] * ap + P.s1[1] * bp + P.s1[2] * cp + P.s1[3] * dp + P.s1[4] * ep;
      } break;
      case 2: {
        double ap = FQ(-1) - c, bp = FQ(1) - c, cp = FQ(2) - c, dp = FQ(3) - c, ep = FQ(4) - c;
        r = P.s2[0] * ap + P.s2[1] * bp + P.s2[2] * cp + P.s2[3] * dp + P.s2[4] * ep;
      } break;
      case 3:
      case 13: {
        double ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2);
        r = P.s3[0] * (ap - am) + P.s3[1] * (bp - bm);
      } break;
      case 4:
      case 14: {
        double ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
               cm = c - FQ(-3);
        r = P.s4[0] * (ap - am) + P.s4[1] * (bp - bm) + P.s4[2] * (cp - cm);
      } break;
      case 5:
      case 15: {
        double ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
               cm = c - FQ(-3), dp = FQ(4) - c, dm = c - FQ(-4);
        r = P.s5[0] * (ap - am) + P.s5[1] * (bp - bm) + P.s5[2] * (cp - cm) + P.s5[3] * (dp - dm);
      } break;
      case 12: {
        double ap = FQ(1) - c, bp = FQ(-1) - c, cp = FQ(-2) - c, dp = FQ(-3) - c, ep = FQ(-4) - c;
        r = P.s2[0] * ap + P.s2[1] * bp + P.s2[2] * cp + P.s2[3] * dp + P.s2[4] * ep;
      } break;
      default: {
        double ap = FQ(-1) - c, bp = FQ(-2) - c, cp = FQ(-3) - c, dp = FQ(-4) - c, ep = FQ(-5) - c;
        r = P.s1[0] * ap + P.s1[1] * bp + P.s1[2] * cp + P.s1[3] * dp + P.s1[4] * ep;
      } break;
    }
  }
  return r * P.ovdl2[dir];
}

#undef FQ

// Mixed derivative in the (d1,d2) plane; d2 periodic, d1 bounded if BND.
// g(k1, k2) returns the value at offsets q1 = k1 - 8 (along d1), q2 = k2 - 8.
#define G(a, b) g((a) + 8, (b) + 8)
#define DIAG(q, b) (G(q, b) - G(q, -(b)) - G(-(q), b) + G(-(q), -(b)))
#define TR(a1, b1)                                                       \
  (P.my1[0] * (G(a1, 1) - G(a1, -1) - G(b1, 1) + G(b1, -1)) +            \
   P.my1[1] * (G(a1, 2) - G(a1, -2) - G(b1, 2) + G(b1, -2)))

template <bool BND, class Gn>
__device__ __forceinline__ double dmix(const Params &P, const Sten &S1, int d1, int d2, Gn g) {
  double r;
  if (!BND || S1.mode == 0) {
    double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4), fe = DIAG(5, 5);
    r = P.mc[0] * fa + P.mc[1] * fb + P.mc[2] * fc + P.mc[3] * fd + P.mc[4] * fe;
  } else {
    switch (S1.mode) {
      case 1: {
        double fa = TR(1, 0), fb = TR(2, 0), fc = TR(3, 0), fd = TR(4, 0);
        r = P.m1[0] * fa + P.m1[1] * fb + P.m1[2] * fc + P.m1[3] * fd;
      } break;
      case 2: {
        double fa = TR(-1, 0), fb = TR(1, 0), fc = TR(2, 0), fd = TR(3, 0);
        r = P.m2[0] * fa + P.m2[1] * fb + P.m2[2] * fc + P.m2[3] * fd;
      } break;
      case 3:
      case 13: {
        double fa = DIAG(1, 1), fb = DIAG(2, 2);
        r = P.m3[0] * fa + P.m3[1] * fb;
      } break;
      case 4:
      case 14: {
        double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3);
        r = P.m4[0] * fa + P.m4[1] * fb + P.m4[2] * fc;
      } break;
      case 5:
      case 15: {
        double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4);
        r = P.m5[0] * fa + P.m5[1] * fb + P.m5[2] * fc + P.m5[3] * fd;
      } break;
      case 12: {
        double fa = TR(0, 1), fb = TR(0, -1), fc = TR(0, -2), fd = TR(0, -3);
        r = P.m2[0] * fa + P.m2[1] * fb + P.m2[2] * fc + P.m2[3] * fd;
      } break;
      default: {
        double fa = TR(0, -1), fb = TR(0, -2), fc = TR(0, -3), fd = TR(0, -4);
        r = P.m1[0] * fa + P.m1[1] * fb + P.m1[2] * fc + P.m1[3] * fd;
      } break;
    }
  }
  return r * P.ovdel[d1] * P.ovdel[d2];
}

#undef TR
#undef DIAG
#undef G

}  // namespace sg
