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

template <bool BND, class R = double, class Fn>
__device__ __forceinline__ R d1(const Params &P, const Sten &S, int dir, Fn f) {
  R r;
  if (!BND || S.mode == 0) {
    R fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3), fd = FQ(4) - FQ(-4),
           fe = FQ(5) - FQ(-5);
    r = (R)P.ac[0] * fa + (R)P.ac[1] * fb + (R)P.ac[2] * fc + (R)P.ac[3] * fd + (R)P.ac[4] * fe;
  } else {
    switch (S.mode) {
      case 1: {
        R fa = FQ(1) - FQ(0), fb = FQ(2) - FQ(0), fc = FQ(3) - FQ(0), fd = FQ(4) - FQ(0);
        r = (R)P.c1[0] * fa + (R)P.c1[1] * fb + (R)P.c1[2] * fc + (R)P.c1[3] * fd;
      } break;
      case 2: {
        R fa = FQ(-1) - FQ(0), fb = FQ(1) - FQ(0), fc = FQ(2) - FQ(0), fd = FQ(3) - FQ(0);
        r = (R)P.c2[0] * fa + (R)P.c2[1] * fb + (R)P.c2[2] * fc + (R)P.c2[3] * fd;
      } break;
      case 3:
      case 13: {
        R fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2);
        r = (R)P.c3[0] * fa + (R)P.c3[1] * fb;
      } break;
      case 4:
      case 14: {
        R fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3);
        r = (R)P.c4[0] * fa + (R)P.c4[1] * fb + (R)P.c4[2] * fc;
      } break;
      case 5:
      case 15: {
        R fa = FQ(1) - FQ(-1), fb = FQ(2) - FQ(-2), fc = FQ(3) - FQ(-3), fd = FQ(4) - FQ(-4);
        r = (R)P.c5[0] * fa + (R)P.c5[1] * fb + (R)P.c5[2] * fc + (R)P.c5[3] * fd;
      } break;
      case 12: {
        R fa = FQ(0) - FQ(1), fb = FQ(0) - FQ(-1), fc = FQ(0) - FQ(-2), fd = FQ(0) - FQ(-3);
        r = (R)P.c2[0] * fa + (R)P.c2[1] * fb + (R)P.c2[2] * fc + (R)P.c2[3] * fd;
      } break;
      default: {  // 11: last point
        R fa = FQ(0) - FQ(-1), fb = FQ(0) - FQ(-2), fc = FQ(0) - FQ(-3), fd = FQ(0) - FQ(-4);
        r = (R)P.c1[0] * fa + (R)P.c1[1] * fb + (R)P.c1[2] * fc + (R)P.c1[3] * fd;
      } break;
    }
  }
  return r * (R)P.ovdel[dir];
}

template <bool BND, class R = double, class Fn>
__device__ __forceinline__ R d2(const Params &P, const Sten &S, int dir, Fn f) {
  R r;
  if (!BND || S.mode == 0) {
    R c = FQ(0);
    R ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
           cm = c - FQ(-3), dp = FQ(4) - c, dm = c - FQ(-4), ep = FQ(5) - c, em = c - FQ(-5);
    r = (R)P.sc[0] * (ap - am) + (R)P.sc[1] * (bp - bm) + (R)P.sc[2] * (cp - cm) + (R)P.sc[3] * (dp - dm) +
        (R)P.sc[4] * (ep - em);
  } else {
    R c = FQ(0);
    switch (S.mode) {
      case 1: {
        R ap = FQ(1) - c, bp = FQ(2) - c, cp = FQ(3) - c, dp = FQ(4) - c, ep = FQ(5) - c;
        r = (R)P.s1[0] * ap + (R)P.s1[1] * bp + (R)P.s1[2] * cp + (R)P.s1[3] * dp + (R)P.s1[4] * ep;
      } break;
      case 2: {
        R ap = FQ(-1) - c, bp = FQ(1) - c, cp = FQ(2) - c, dp = FQ(3) - c, ep = FQ(4) - c;
        r = (R)P.s2[0] * ap + (R)P.s2[1] * bp + (R)P.s2[2] * cp + (R)P.s2[3] * dp + (R)P.s2[4] * ep;
      } break;
      case 3:
      case 13: {
        R ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2);
        r = (R)P.s3[0] * (ap - am) + (R)P.s3[1] * (bp - bm);
      } break;
      case 4:
      case 14: {
        R ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
               cm = c - FQ(-3);
        r = (R)P.s4[0] * (ap - am) + (R)P.s4[1] * (bp - bm) + (R)P.s4[2] * (cp - cm);
      } break;
      case 5:
      case 15: {
        R ap = FQ(1) - c, am = c - FQ(-1), bp = FQ(2) - c, bm = c - FQ(-2), cp = FQ(3) - c,
               cm = c - FQ(-3), dp = FQ(4) - c, dm = c - FQ(-4);
        r = (R)P.s5[0] * (ap - am) + (R)P.s5[1] * (bp - bm) + (R)P.s5[2] * (cp - cm) + (R)P.s5[3] * (dp - dm);
      } break;
      case 12: {
        R ap = FQ(1) - c, bp = FQ(-1) - c, cp = FQ(-2) - c, dp = FQ(-3) - c, ep = FQ(-4) - c;
        r = (R)P.s2[0] * ap + (R)P.s2[1] * bp + (R)P.s2[2] * cp + (R)P.s2[3] * dp + (R)P.s2[4] * ep;
      } break;
      default: {
        R ap = FQ(-1) - c, bp = FQ(-2) - c, cp = FQ(-3) - c, dp = FQ(-4) - c, ep = FQ(-5) - c;
        r = (R)P.s1[0] * ap + (R)P.s1[1] * bp + (R)P.s1[2] * cp + (R)P.s1[3] * dp + (R)P.s1[4] * ep;
      } break;
    }
  }
  return r * (R)P.ovdl2[dir];
}

#undef FQ

// ---------------------------------------------------------------------------
// Every first/second-derivative row above written as one 5-term form,
//   d1 = sum_k W1[k] * (f(P1[k]) - f(M1[k]))
//   d2 = sum_k W2[k] * ((f(P2[k]) - f(0)) - (f(0) - f(M2[k])))
// summed left to right.  Rows with fewer terms carry W = 0 and P = M = 0,
// which add an exact zero; one-sided second-derivative rows use M2 = 0, whose
// (f(0) - f(0)) is an exact zero.  The result is bit-identical to d1/d2, and
// lanes on different rows run the same instructions, so a wave that mixes
// closure rows with interior points does not serialise over row kinds.
// Row index: 0 interior, 1..5 left rows m = 0..4, 6..10 right rows n-1-m = 0..4
// (mode 1..5 -> row 1..5, mode 11..15 -> row 6..10).

__device__ __forceinline__ int row_of_mode(int mode) { return mode == 0 ? 0 : (mode < 10 ? mode : mode - 5); }

// Offsets (P1, M1, P2, M2: 4 x 5 values in -5..5) of every row, packed as
// signed bytes, four to an int: word w holds entries 4w..4w+3 of the 20.
struct RowOffsets {
  int w[SG_NROW][5];
};
constexpr RowOffsets make_row_offsets() {
  RowOffsets r{};
  for (int row = 0; row < SG_NROW; row++) {
    int o[20] = {};
    auto set = [&](int base, int a, int b, int c, int d, int e) {
      o[base] = a; o[base + 1] = b; o[base + 2] = c; o[base + 3] = d; o[base + 4] = e;
    };
    switch (row) {
      case 0: set(0, 1, 2, 3, 4, 5); set(5, -1, -2, -3, -4, -5); set(10, 1, 2, 3, 4, 5); set(15, -1, -2, -3, -4, -5); break;
      case 1: set(0, 1, 2, 3, 4, 0); set(10, 1, 2, 3, 4, 5); break;
      case 2: set(0, -1, 1, 2, 3, 0); set(10, -1, 1, 2, 3, 4); break;
      case 3: case 8: set(0, 1, 2, 0, 0, 0); set(5, -1, -2, 0, 0, 0); set(10, 1, 2, 0, 0, 0); set(15, -1, -2, 0, 0, 0); break;
      case 4: case 9: set(0, 1, 2, 3, 0, 0); set(5, -1, -2, -3, 0, 0); set(10, 1, 2, 3, 0, 0); set(15, -1, -2, -3, 0, 0); break;
      case 5: case 10: set(0, 1, 2, 3, 4, 0); set(5, -1, -2, -3, -4, 0); set(10, 1, 2, 3, 4, 0); set(15, -1, -2, -3, -4, 0); break;
      case 7: set(5, 1, -1, -2, -3, 0); set(10, 1, -1, -2, -3, -4); break;
      default: set(5, -1, -2, -3, -4, 0); set(10, -1, -2, -3, -4, -5); break;   // 6: last point
    }
    for (int w = 0; w < 5; w++)
      r.w[row][w] = (int)((unsigned)(o[4 * w] & 255) | ((unsigned)(o[4 * w + 1] & 255) << 8) |
                          ((unsigned)(o[4 * w + 2] & 255) << 16) | ((unsigned)(o[4 * w + 3] & 255) << 24));
  }
  return r;
}
__constant__ static const RowOffsets c_row_offsets = make_row_offsets();

// Weights (W1 then W2) of a row, written to w[0..9].  Host side: the table
// is built once per handle (a per-lane switch filling it on the device was
// miscompiled for gfx950 at -O3, storing a wrong weight in rows 6 and 8).
inline void row_weights(const Params &P, int row, double *w) {
  for (int k = 0; k < 10; k++) w[k] = 0.0;
  switch (row) {
    case 0: for (int k = 0; k < 5; k++) { w[k] = P.ac[k]; w[5 + k] = P.sc[k]; } break;
    case 1: case 6: for (int k = 0; k < 4; k++) w[k] = P.c1[k]; for (int k = 0; k < 5; k++) w[5 + k] = P.s1[k]; break;
    case 2: case 7: for (int k = 0; k < 4; k++) w[k] = P.c2[k]; for (int k = 0; k < 5; k++) w[5 + k] = P.s2[k]; break;
    case 3: case 8: for (int k = 0; k < 2; k++) { w[k] = P.c3[k]; w[5 + k] = P.s3[k]; } break;
    case 4: case 9: for (int k = 0; k < 3; k++) { w[k] = P.c4[k]; w[5 + k] = P.s4[k]; } break;
    default: for (int k = 0; k < 4; k++) { w[k] = P.c5[k]; w[5 + k] = P.s5[k]; } break;
  }
}

// Offset entry e (0..19) from the packed words.
__device__ __forceinline__ int row_off(const int packed[5], int e) {
  return __builtin_amdgcn_sbfe(packed[e >> 2], 8 * (e & 3), 8);
}

// First/second derivative in the 5-term form; f(q) gives the value at offset
// q, w the row's 10 weights.
template <class R = double, class Fn, class Wn>
__device__ __forceinline__ R d1_rows(const Params &P, const int packed[5], int dir, Wn w, Fn f) {
  R r = (R)w(0) * (f(row_off(packed, 0)) - f(row_off(packed, 5)));
#pragma unroll
  for (int k = 1; k < 5; k++) r = r + (R)w(k) * (f(row_off(packed, k)) - f(row_off(packed, 5 + k)));
  return r * (R)P.ovdel[dir];
}
template <class R = double, class Fn, class Wn>
__device__ __forceinline__ R d2_rows(const Params &P, const int packed[5], int dir, Wn w, Fn f) {
  const R c = f(0);
  R r = (R)w(5) * ((f(row_off(packed, 10)) - c) - (c - f(row_off(packed, 15))));
#pragma unroll
  for (int k = 1; k < 5; k++) r = r + (R)w(5 + k) * ((f(row_off(packed, 10 + k)) - c) - (c - f(row_off(packed, 15 + k))));
  return r * (R)P.ovdl2[dir];
}

// ---------------------------------------------------------------------------
// Mixed-derivative rows in one uniform form.  With
//   Q(p, m, q) = g(p, q) - g(p, -q) - g(m, q) + g(m, -q)
// every row of dmix below is
//   sum_k W_k * (V_0 * Q(P_k, M_k, q0_k) + V_1 * Q(P_k, M_k, q1_k)),   k = 0..4,
// with V = (1, 0) and q0_k = P_k for the diagonal rows (interior and the 4th,
// 6th, 8th-order rows) and V = (my1[0], my1[1]), q = (1, 2) for the one-sided
// rows.  Unused terms have W = 0 and P = M = q = 0, whose Q is an exact zero,
// and V_1 = 0 multiplies an exact-zero Q, so results are bit-identical to dmix.
struct MixRowOffsets {
  int w[SG_NROW][5];   // (P, M, q0, q1) of term k as signed bytes in word k
};
constexpr MixRowOffsets make_mix_row_offsets() {
  MixRowOffsets r{};
  for (int row = 0; row < SG_NROW; row++) {
    int p[5] = {}, m[5] = {}, q0[5] = {}, q1[5] = {};
    auto diag = [&](int n) {
      for (int k = 0; k < n; k++) { p[k] = k + 1; m[k] = -(k + 1); q0[k] = k + 1; }
    };
    auto tr = [&](int a0, int a1, int a2, int a3, int b0, int b1, int b2, int b3) {
      const int a[4] = {a0, a1, a2, a3}, b[4] = {b0, b1, b2, b3};
      for (int k = 0; k < 4; k++) { p[k] = a[k]; m[k] = b[k]; q0[k] = 1; q1[k] = 2; }
    };
    switch (row) {
      case 0: diag(5); break;
      case 1: tr(1, 2, 3, 4, 0, 0, 0, 0); break;
      case 2: tr(-1, 1, 2, 3, 0, 0, 0, 0); break;
      case 3: case 8: diag(2); break;
      case 4: case 9: diag(3); break;
      case 5: case 10: diag(4); break;
      case 7: tr(0, 0, 0, 0, 1, -1, -2, -3); break;
      default: tr(0, 0, 0, 0, -1, -2, -3, -4); break;   // 6: last point
    }
    for (int k = 0; k < 5; k++)
      r.w[row][k] = (int)((unsigned)(p[k] & 255) | ((unsigned)(m[k] & 255) << 8) |
                          ((unsigned)(q0[k] & 255) << 16) | ((unsigned)(q1[k] & 255) << 24));
  }
  return r;
}
__constant__ static const MixRowOffsets c_mix_row_offsets = make_mix_row_offsets();

// Weights W_0..W_4, V_0, V_1 of a mixed-derivative row (host side).
inline void mix_row_weights(const Params &P, int row, double *w) {
  for (int k = 0; k < 7; k++) w[k] = 0.0;
  w[5] = 1.0;
  switch (row) {
    case 0: for (int k = 0; k < 5; k++) w[k] = P.mc[k]; break;
    case 1: case 6: for (int k = 0; k < 4; k++) w[k] = P.m1[k]; w[5] = P.my1[0]; w[6] = P.my1[1]; break;
    case 2: case 7: for (int k = 0; k < 4; k++) w[k] = P.m2[k]; w[5] = P.my1[0]; w[6] = P.my1[1]; break;
    case 3: case 8: for (int k = 0; k < 2; k++) w[k] = P.m3[k]; break;
    case 4: case 9: for (int k = 0; k < 3; k++) w[k] = P.m4[k]; break;
    default: for (int k = 0; k < 4; k++) w[k] = P.m5[k]; break;
  }
}

// Mixed derivative in the uniform row form; g(a, b) gives the value at offset
// a along d1 and b along d2, w(k) the row's 7 weights, packed its offsets.
template <class R = double, class Gn, class Wn>
__device__ __forceinline__ R dmix_rows(const Params &P, const int packed[5], int d1, int d2, Wn w, Gn g) {
  auto Qv = [&](int p, int m, int q) { return g(p, q) - g(p, -q) - g(m, q) + g(m, -q); };
  const R v0 = (R)w(5), v1 = (R)w(6);
  R r = 0.0;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int pk = packed[k];
    const int p = __builtin_amdgcn_sbfe(pk, 0, 8), m = __builtin_amdgcn_sbfe(pk, 8, 8);
    const int q0 = __builtin_amdgcn_sbfe(pk, 16, 8), q1 = __builtin_amdgcn_sbfe(pk, 24, 8);
    const R t = (R)w(k) * (v0 * Qv(p, m, q0) + v1 * Qv(p, m, q1));
    r = k == 0 ? t : r + t;
  }
  return r * (R)P.ovdel[d1] * (R)P.ovdel[d2];
}

// Mixed derivative in the (d1,d2) plane; d2 periodic, d1 bounded if BND.
// g(k1, k2) returns the value at offsets q1 = k1 - 8 (along d1), q2 = k2 - 8.
#define G(a, b) g((a) + 8, (b) + 8)
#define DIAG(q, b) (G(q, b) - G(q, -(b)) - G(-(q), b) + G(-(q), -(b)))
#define TR(a1, b1)                                                       \
  ((R)P.my1[0] * (G(a1, 1) - G(a1, -1) - G(b1, 1) + G(b1, -1)) +            \
   (R)P.my1[1] * (G(a1, 2) - G(a1, -2) - G(b1, 2) + G(b1, -2)))

template <bool BND, class R = double, class Gn>
__device__ __forceinline__ R dmix(const Params &P, const Sten &S1, int d1, int d2, Gn g) {
  R r;
  if (!BND || S1.mode == 0) {
    R fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4), fe = DIAG(5, 5);
    r = (R)P.mc[0] * fa + (R)P.mc[1] * fb + (R)P.mc[2] * fc + (R)P.mc[3] * fd + (R)P.mc[4] * fe;
  } else {
    switch (S1.mode) {
      case 1: {
        R fa = TR(1, 0), fb = TR(2, 0), fc = TR(3, 0), fd = TR(4, 0);
        r = (R)P.m1[0] * fa + (R)P.m1[1] * fb + (R)P.m1[2] * fc + (R)P.m1[3] * fd;
      } break;
      case 2: {
        R fa = TR(-1, 0), fb = TR(1, 0), fc = TR(2, 0), fd = TR(3, 0);
        r = (R)P.m2[0] * fa + (R)P.m2[1] * fb + (R)P.m2[2] * fc + (R)P.m2[3] * fd;
      } break;
      case 3:
      case 13: {
        R fa = DIAG(1, 1), fb = DIAG(2, 2);
        r = (R)P.m3[0] * fa + (R)P.m3[1] * fb;
      } break;
      case 4:
      case 14: {
        R fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3);
        r = (R)P.m4[0] * fa + (R)P.m4[1] * fb + (R)P.m4[2] * fc;
      } break;
      case 5:
      case 15: {
        R fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4);
        r = (R)P.m5[0] * fa + (R)P.m5[1] * fb + (R)P.m5[2] * fc + (R)P.m5[3] * fd;
      } break;
      case 12: {
        R fa = TR(0, 1), fb = TR(0, -1), fc = TR(0, -2), fd = TR(0, -3);
        r = (R)P.m2[0] * fa + (R)P.m2[1] * fb + (R)P.m2[2] * fc + (R)P.m2[3] * fd;
      } break;
      default: {
        R fa = TR(0, -1), fb = TR(0, -2), fc = TR(0, -3), fd = TR(0, -4);
        r = (R)P.m1[0] * fa + (R)P.m1[1] * fb + (R)P.m1[2] * fc + (R)P.m1[3] * fd;
      } break;
    }
  }
  return r * (R)P.ovdel[d1] * (R)P.ovdel[d2];
}

#undef TR
#undef DIAG
#undef G

}  // namespace sg
