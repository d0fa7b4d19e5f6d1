// sg_kernels.hip -- CDNA4 kernels of one SENGA+ Runge-Kutta substage (FP64).
//
// Per substage (src/senga2.f:113-159):
//   k_boundt   BOUNDT inflow values at x-left            (src/boundt.f:177-218)
//   k_prim     TEMPER + transport + CHRATE + enthalpies   (src/temper.f, src/chrate.f,
//                                                          src/rhscal.f:515-573,1292-1470,3854-3891)
//   k_pass<d>  every derivative along direction d and the RHS terms that
//              involve derivatives along d only           (src/rhscal.f, src/rhsvel.f)
//   k_final    mixed derivatives, cross-direction viscous terms, NSCBC face
//              corrections (src/bounds.f) and the RK register update
//              (src/lincom.f, src/fincom.f)
// Periodic directions use wrapped indices: no halo copies are stored.
#include <math.h>

#include "sg_internal.h"
#include "sg_stencil.h"
#include "sg_mech_h2.h"

namespace sg {

// Point kernels and their helpers: a multiply and an add are fused only
// within one expression, so the fusing does not depend on how the compiler
// lays out the surrounding code (the compiled-mechanism and table-driven
// step loops then give identical results).  The stencil kernels below
// switch back to fusing across statements.
#pragma clang fp contract(on)

#define BLK 256

// Forces x to be materialised here.  Without it the compiler sinks stencil
// sums toward their last use and keeps every neighbour value live across the
// following stages (hundreds of registers in the direction passes).  The empty
// volatile asm is ordered against the LDS barriers, so each derivative is
// formed within its own stage.
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(float &x) { asm volatile("" : "+v"(x)); }

// Scheduling boundary between the phases of a long straight-line kernel, so
// that independent phases are not interleaved (which multiplies live values).
#define SG_PHASE() __builtin_amdgcn_sched_barrier(0)

// Element at byte offset bo of a field.  A 32-bit offset from a uniform field
// base lets loads and stores use the scalar-base addressing form, instead of a
// live 64-bit vector address per field (field byte offsets stay below 2^32,
// which handle creation checks).
template <class T>
__device__ __forceinline__ T &fld(T *base, uint32_t bo) {
  return *(T *)((char *)base + bo);
}

// XCD-aware block order.  Work-groups are dispatched round-robin over the
// chip's 8 XCDs, each with its own L2; consecutive block ids therefore land on
// different L2s and neighbouring tiles (which share stencil halos) miss.  The
// remap gives each XCD one contiguous chunk of logical blocks: hardware block b
// runs on XCD b % 8 as that XCD's (b / 8)-th block, and becomes logical block
// start(XCD) + b / 8.  A bijection on [0, n) for any n.
#define SG_NXCD 8
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n / SG_NXCD, r = n % SG_NXCD;
  const int x = b % SG_NXCD, slot = b / SG_NXCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
}

__device__ __forceinline__ void coords(const Params &P, long idx, int &i, int &j, int &k) {
  int nxy = P.nx * P.ny;
  k = (int)(idx / nxy);
  int r = (int)(idx - (long)k * nxy);
  j = r / P.nx;
  i = r - j * P.nx;
}

// Packed temperature-interval indices: one bit per species (the reference's
// ITNDEX packing, src/com_senga2.h:88-98, with number base NTBASE = 2, which
// holds the two intervals every species has here), species s in bit s.
static_assert(SG_NTINT == 2 && SG_MAX_SPEC <= 32, "one index bit per species in a 32-bit word");
__device__ __forceinline__ int itint_of(uint32_t idx, int s) { return (int)((idx >> s) & 1u); }

// Interval of species s at temperature t: the first interval whose upper
// bound is not below t, the last one otherwise.  With at most two intervals
// (sg_create checks 1 <= ntint <= SG_NTINT) that is one comparison; written
// as a per-lane search loop, the uniform bound was fetched by a vector load
// and waited on inside the loop, once per species.
__device__ __forceinline__ int locate(const sg_config *c, int s, double t) {
  return (c->ntint[s] > 1 && t > c->tinthi[s][0]) ? 1 : 0;
}

__device__ void set_error(int *err, int code, int i, int j, int k) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = i + 1;
    err[2] = j + 1;
    err[3] = k + 1;
  }
}

// Member owning flattened block gb of a launch of kind KIND, and the block's
// index lb within that member's own grid.  The map and table are read-only
// for the kernel's lifetime, so these are uniform scalar loads.
// Field pointers read from the member table are generic pointers to the
// compiler; stating that they point to global memory (a no-op conversion on
// this target) lets every access through them use global instead of flat
// instructions, which would otherwise also count against the LDS wait counter
// and stall the LDS-staged stencils.
template <class T>
__device__ __forceinline__ T *as_global(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}
__device__ __forceinline__ Fields global_fields(const Fields &G) {
  Fields F;
  F.tab = as_global(G.tab);
  F.q = as_global(G.q);
  F.run = as_global(G.run);
  F.prim = as_global(G.prim);
  F.rate = as_global(G.rate);
  F.acc = as_global(G.acc);
  F.grad = as_global(G.grad);
  F.itndex = as_global(G.itndex);
  F.face = as_global(G.face);
  F.err = as_global(G.err);
  F.erk = as_global(G.erk);
  F.emax = as_global(G.emax);
  return F;
}

template <int KIND>
__device__ __forceinline__ const Member &member_at(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                                   int gb, int &lb) {
  const Member &M = mem[map[gb]];
  lb = gb - M.blk0[KIND];
  return M;
}

// ---------------------------------------------------------------------------
// TEMPIN: bisection from TRIN in 50 K steps, src/tempin.f:55-512
template <int NS, class R>
__global__ void __launch_bounds__(BLK) k_tempin(Params P, const sg_config *__restrict__ c, Fields F) {
  long idx = (long)blockIdx.x * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  const long fs = P.fstride;
  const double *q = F.q;
  double d = q[idx], ru = q[fs + idx], rv = q[2 * fs + idx], rw = q[3 * fs + idx];
  double ry[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) ry[s] = q[(5 + s) * fs + idx];
  double ukuk = (ru * ru + rv * rv + rw * rw) / d;
  double tuk2me = 0.5 * ukuk - q[4 * fs + idx];
  const double toltmp = 1.0e-4, tininc = 5.0e1, tlimlo = 2.0e2, tlimhi = 3.0e3;
  auto resid = [&](double t) {
    double tc[6];
    tc[0] = tuk2me;
#pragma unroll
    for (int m = 1; m <= 5; m++) tc[m] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      int it = locate(c, s, t);
      tc[0] = tc[0] + ry[s] * c->amascp[s][it][c->ncenth[s][it] - 1];
      tc[1] = tc[1] + ry[s] * c->amasct[s][it][0];
      for (int m = 2; m <= c->ncpoly[s][it]; m++) tc[m] = tc[m] + ry[s] * c->amasct[s][it][m - 1];
    }
    double r = tc[5];
    for (int m = 4; m >= 1; m--) r = tc[m] + r * t;
    return tc[0] + r * t;
  };
  double tlower = tlimlo, tupper = tlimhi, t = c->trin;
  bool fnconv = true, fail = false;
  double res = resid(t);
  if (fabs(res) < toltmp) {
    fnconv = false;
  } else if (res < 0.0) {
    for (;;) {
      tlower = t;
      t = t + tininc;
      res = resid(t);
      if (fabs(res) < toltmp) { fnconv = false; break; }
      if (res < 0.0) {
        if (t < tlimhi) continue;
        fail = true;
        break;
      }
      tupper = t;
      break;
    }
  } else {
    for (;;) {
      tupper = t;
      t = t - tininc;
      res = resid(t);
      if (fabs(res) < toltmp) { fnconv = false; break; }
      if (res > 0.0) {
        if (t > tlimlo) continue;
        fail = true;
        break;
      }
      tlower = t;
      break;
    }
  }
  if (fail) {
    int i, j, k;
    coords(P, idx, i, j, k);
    set_error(F.err, SG_ETEMPIN, i, j, k);
    return;
  }
  if (fnconv) {
    // bisection: the bracket halves each pass, bounded by 64 passes in double
    for (int it = 0; it < 200; it++) {
      t = 0.5 * (tlower + tupper);
      res = resid(t);
      if (fabs(res) < toltmp) break;
      if (res < 0.0) tlower = t;
      else tupper = t;
    }
  }
  uint32_t nidx = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) nidx += (uint32_t)locate(c, s, t) << s;
  F.itndex[idx] = nidx;
  R *pr = (R *)F.prim;
  pr[P_T * fs + idx] = (R)t;
  double s7 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) s7 = s7 + ry[s] * c->rgspec[s];
  pr[P_P * fs + idx] = (R)(t * s7 - P.pofs);
}

// ---------------------------------------------------------------------------
// BOUNDT for NSBCI3 at x-left (src/boundt.f:177-218; BCDTXL, BCUTXL type 1, BCYTXL)
template <int NS>
__global__ void __launch_bounds__(BLK) k_boundt(const Member *__restrict__ mem, const uint32_t *__restrict__ map) {
  int lb;
  const Member &M = member_at<L_FACE>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  if (!P.inflow_xl) return;
  long t = (long)lb * BLK + threadIdx.x;
  if (t >= (long)P.ny * P.nz) return;
  long p = t * P.nx;  // i = 0
  const long fs = P.fstride;
  double strd = P.drin, stru = P.uin;
  F.q[p] = strd;
  F.q[fs + p] = strd * stru;
  F.q[2 * fs + p] = strd * 0.0;
  F.q[3 * fs + p] = strd * 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) F.q[(5 + s) * fs + p] = strd * P.yrin[s];
}

// BOUNTT for NSBCI3 at x-left (src/bountt.f:195-262)
template <int NS>
__global__ void __launch_bounds__(BLK) k_bountt(const Member *__restrict__ mem, const uint32_t *__restrict__ map) {
  int lb;
  const Member &M = member_at<L_FACE>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  if (!P.inflow_xl) return;
  long t = (long)lb * BLK + threadIdx.x;
  if (t >= (long)P.ny * P.nz) return;
  long p = t * P.nx;
  const long fs = P.fstride;
  double strd = P.drin, stru = P.uin;
  double v[4] = {strd, strd * stru, strd * 0.0, strd * 0.0};
#pragma unroll
  for (int m = 0; m < 4; m++) {
    F.q[m * fs + p] = v[m];
    F.run[m * fs + p] = v[m];
  }
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double y = strd * P.yrin[s];
    F.q[(5 + s) * fs + p] = y;
    F.run[(5 + s) * fs + p] = y;
  }
  // the imposed values carry no RK error (src/bountt.f)
  if (F.erk) {
    for (int m = 0; m < 4; m++) F.erk[m * fs + p] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) F.erk[(5 + s) * fs + p] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Point physics: TEMPER (src/temper.f:58-225), conductivity and viscosity
// (src/rhscal.f:515-573, 3854-3891, src/rhsvel.f:926-980), CHRATE
// (src/chrate.f:146-760), primitive variables, diffusivities
// (src/rhscal.f:1292-1346) and species enthalpies (src/rhscal.f:1444-1470).
// Two-interval polynomials are evaluated for both intervals with uniform
// (scalar) coefficients and the result is chosen per point.
// a * b + c for a uniform c, as one VOP3 FMA that reads c from scalar
// registers.  Left to itself the compiler copies c into vector registers and
// uses the two-address FMAC form (two moves and an FMA in FP64, one move and
// an FMA in FP32) at every Horner step.  Same operation, same rounding.
__device__ __forceinline__ double fma_u(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}
__device__ __forceinline__ float fma_u(float a, float b, float c) {
  float r;
  asm("v_fma_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}
// A uniform FP64 value into vector registers with one 64-bit move (the
// compiler otherwise copies the two halves with two moves), for the leading
// coefficient that starts each Horner chain.
__device__ __forceinline__ double vmov_u(double c) {
  double r;
  asm("v_mov_b64 %0, %1" : "=v"(r) : "s"(c));
  return r;
}

// The temperature interval it when every active lane of the wave holds the
// same one, else -1; wave-uniform (from ballots), so the polynomials below
// then run once with scalar coefficients.
__device__ __forceinline__ int wave_interval(int it) {
  const unsigned long long on = __ballot(it != 0), all = __ballot(1);
  return __builtin_amdgcn_readfirstlane(on == 0ull ? 0 : (on == all ? 1 : -1));
}

// Horner over the zero-padded SG_NCPMAX-term polynomial of interval it.
// When the wave disagrees on it, both intervals are evaluated with uniform
// (scalar) coefficients and the result is chosen per point, which is cheaper
// than choosing every coefficient.  Both forms compute the same sums.  (The
// enthalpy and Gibbs polynomials below keep the two-interval form only: with
// the wave-uniform branch there too the FP64 point kernel spilled.)
__device__ __forceinline__ double horner(const double (&a)[SG_NTINT][SG_NCPMAX], int it, double x) {
  static_assert(SG_NTINT == 2, "two temperature intervals");
  const int u = wave_interval(it);
  if (u >= 0) {
    double f = vmov_u(a[u][SG_NCPMAX - 1]);
#pragma unroll
    for (int m = SG_NCPMAX - 2; m >= 0; m--) f = fma_u(f, x, a[u][m]);
    return f;
  }
  double f0 = vmov_u(a[0][SG_NCPMAX - 1]), f1 = vmov_u(a[1][SG_NCPMAX - 1]);
#pragma unroll
  for (int m = SG_NCPMAX - 2; m >= 0; m--) {
    f0 = fma_u(f0, x, a[0][m]);
    f1 = fma_u(f1, x, a[1][m]);
  }
  return it ? f1 : f0;
}

// c(it) + x * poly(it)(x) for a two-interval table: formed for both intervals
// with scalar coefficients and the result chosen per point (a choice of the
// coefficients instead costs register copies and selects for every one).
__device__ __forceinline__ double horner_c(const double (&a)[SG_NTINT][SG_NCPMAX], const double (&c0)[SG_NTINT],
                                           int it, double x) {
  double f0 = vmov_u(a[0][SG_NCPMAX - 1]), f1 = vmov_u(a[1][SG_NCPMAX - 1]);
#pragma unroll
  for (int m = SG_NCPMAX - 2; m >= 0; m--) {
    f0 = fma_u(f0, x, a[0][m]);
    f1 = fma_u(f1, x, a[1][m]);
  }
  f0 = fma_u(f0, x, c0[0]);
  f1 = fma_u(f1, x, c0[1]);
  return it ? f1 : f0;
}
// Gibbs function gbh(it)/T - gbs(it) ln T - poly(it)(T), both intervals formed,
// one chosen (ovt = 1/T).
__device__ __forceinline__ double gibbs_of(const Tables *tb, int s, int it, double t, double ovt, double lt) {
  double f0 = vmov_u(tb->gb[s][0][SG_NCPMAX - 1]), f1 = vmov_u(tb->gb[s][1][SG_NCPMAX - 1]);
#pragma unroll
  for (int m = SG_NCPMAX - 2; m >= 0; m--) {
    f0 = fma_u(f0, t, tb->gb[s][0][m]);
    f1 = fma_u(f1, t, tb->gb[s][1][m]);
  }
  const double g0 = tb->gbh[s][0] * ovt - tb->gbs[s][0] * lt - f0;
  const double g1 = tb->gbh[s][1] * ovt - tb->gbs[s][1] * lt - f1;
  return it ? g1 : g0;
}

// TEMPER's polynomial sums (src/temper.f:92-128 with the coefficients of
// src/chemin.f:906-916): species s of mass density y in interval it adds its
// terms to tc (E(T) - E) and td (dE/dT).  A wave that agrees on it adds them
// with that interval's scalar coefficients; otherwise y * c(interval) is
// formed as y0 * c(0) + y1 * c(1) with exactly one of y0, y1 zero: the zero
// term adds an exact zero, so every sum is the one formed with the chosen
// interval alone, and the coefficients stay uniform.  Terms above the
// interval's order add y * 0, which leaves the sums unchanged.
__device__ __forceinline__ void temper_add(const Tables *tb, int s, int it, double y, double (&tc)[6],
                                           double (&td)[6]) {
  const int u = wave_interval(it);
  if (u >= 0) {
    tc[0] = tc[0] + y * tb->hcp[s][u];
    tc[1] = tc[1] + y * tb->ct[s][u][0];
    td[1] = tc[1];
#pragma unroll
    for (int m = 2; m <= SG_NCPMAX; m++) {
      tc[m] = tc[m] + y * tb->ct[s][u][m - 1];
      td[m] = td[m] + y * tb->cp[s][u][m - 1];
    }
    return;
  }
  const double y0 = it ? 0.0 : y, y1 = it ? y : 0.0;
  tc[0] = tc[0] + y0 * tb->hcp[s][0];
  tc[0] = tc[0] + y1 * tb->hcp[s][1];
  tc[1] = tc[1] + y0 * tb->ct[s][0][0];
  tc[1] = tc[1] + y1 * tb->ct[s][1][0];
  td[1] = tc[1];
#pragma unroll
  for (int m = 2; m <= SG_NCPMAX; m++) {
    tc[m] = tc[m] + y0 * tb->ct[s][0][m - 1];
    tc[m] = tc[m] + y1 * tb->ct[s][1][m - 1];
    td[m] = td[m] + y0 * tb->cp[s][0][m - 1];
    td[m] = td[m] + y1 * tb->cp[s][1][m - 1];
  }
}

// Reciprocals of a[0..N-1] in place with one division (prefix products,
// Montgomery's trick): 3 (N - 1) products instead of N - 1 more divisions.
template <int N>
__device__ __forceinline__ void batch_recip(double (&a)[N]) {
  double pre[N];
  pre[0] = a[0];
#pragma unroll
  for (int k = 1; k < N; k++) pre[k] = pre[k - 1] * a[k];
  double inv = 1.0 / pre[N - 1];
#pragma unroll
  for (int k = N - 1; k > 0; k--) {
    const double ak = a[k];
    a[k] = inv * pre[k - 1];
    inv = inv * ak;
  }
  a[0] = inv;
}

// Transport fits, zero-padded to SG_NTRC terms.
__device__ __forceinline__ double trfit(const double *a, double x) {
  double f = vmov_u(a[SG_NTRC - 1]);
#pragma unroll
  for (int m = SG_NTRC - 2; m >= 0; m--) f = fma_u(f, x, a[m]);
  return f;
}


// exp(x) in FP64 by Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989): x = (64 m + j) ln2/64 + r, |r| <= ln2/128,
// exp(x) = 2^m 2^(j/64) exp(r), with exp(r) - 1 from its degree-5 Taylor
// polynomial (truncation < 4e-17, a third of an ulp).  n ln2/64 is split so
// that n * L1 is exact.
// xt is the 2^(j/64) table (kExp2j) staged in LDS.  Domain: finite x with
// |x| < 2^31 ln2/64 (about 2.3e7), where n fits the 32-bit exponent split.
// There it is within 1 ulp of a correctly rounded exp (measured against glibc
// on 1.2e8 arguments in [-745, 709.7]), gives inf above 709.78 and 0 below
// -745.2; infinite arguments (which would give NaN) are outside the domain,
// and the kernels never pass them: every call's argument is a rate exponent,
// a transport fit or a falloff term of a temperature that TEMPER bounds.
// NaN propagates.  13 VALU instructions and one LDS read, against ~22 for the
// library exp.
static constexpr double kExp2j[64] = {0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0, 0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0, 0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0, 0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0, 0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0, 0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0};
__device__ __forceinline__ double exp_t(double x, const double *xt) {
  const double INV = 0x1.71547652b82fep+6, L1 = 0x1.62e42ff000000p-7, L2 = -0x1.718432a1b0e26p-41;
  // n = nearest integer to x*64/ln2 by the shifter 1.5*2^52: the fused sum
  // rounds to an integer, whose low 32 bits are n (two's complement)
  const double SH = 0x1.8p52;
  const double tsh = __builtin_fma(x, INV, SH);
  const double n = tsh - SH;
  const int ni = (int)(unsigned)__builtin_bit_cast(unsigned long long, tsh);
  double r = __builtin_fma(-n, L1, x);
  r = __builtin_fma(-n, L2, r);
  double q = __builtin_fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = __builtin_fma(q, r, 1.0 / 6.0);
  q = __builtin_fma(q, r, 0.5);
  const double p = __builtin_fma(q, r * r, r);
  const double tj = xt[ni & 63];
  return __builtin_ldexp(__builtin_fma(tj, p, tj), ni >> 6);
}

// ---------------------------------------------------------------------------
// CHRATE step loop (src/chrate.f:146-760), shared by the two sources of
// mechanism constants: RtMech reads the runtime configuration and derived
// tables; CtMech<T> reads compile-time tables (a mechanism compiled into the
// library, e.g. MechH2), for which the loop is unrolled over the steps, every
// list length, species index and rate constant becomes an immediate, and the
// per-point concentrations and Gibbs functions stay in registers.  Both run
// the same expressions in the same order on the same values, so they give
// the same results.
struct RtMech {
  const sg_config *c;
  const Tables *tb;
  static constexpr bool fixed = false;
  __device__ int nst() const { return c->nstep; }
  __device__ int mb(int st) const { return c->mblist[st]; }
  __device__ int mg(int st) const { return c->mglist[st]; }
  __device__ int ml(int st) const { return c->mllist[st]; }
  __device__ int mt(int st) const { return c->mtlist[st]; }
  __device__ int ms(int st) const { return c->mslist[st]; }
  __device__ double rp(int st, int q) const { return c->rparam[st][q]; }
  __device__ double rl(int j, int q) const { return c->rclind[j][q]; }
  __device__ double rt(int j, int q) const { return c->rctroe[j][q]; }
  __device__ double rs(int j, int q) const { return c->rcsrif[j][q]; }
  __device__ int nrs(int st) const { return c->nrslen[st]; }
  __device__ int nrsp(int st, int m) const { return c->nrspec[st][m]; }
  __device__ int nrc(int st) const { return c->nrclen[st]; }
  __device__ int nrcp(int st, int m) const { return c->nrcpec[st][m]; }
  __device__ double crs(int st, int m) const { return c->crspec[st][m]; }
  __device__ float crsf(int st, int m) const { return tb->crspecf[st][m]; }
  __device__ int nss(int st) const { return c->nsslen[st]; }
  __device__ int nssp(int st, int m) const { return c->nsspec[st][m]; }
  __device__ double dmu(int st, int m) const { return c->diffmu[st][m]; }
  __device__ double dmw(int st, int m) const { return c->diffmw[st][m]; }
  __device__ float dmwf(int st, int m) const { return tb->diffmwf[st][m]; }
  __device__ int nps(int st) const { return c->npslen[st]; }
  __device__ int npsp(int st, int m) const { return c->npspec[st][m]; }
  __device__ int npc(int st) const { return c->npclen[st]; }
  __device__ int npcp(int st, int m) const { return c->npcpec[st][m]; }
  __device__ double cps(int st, int m) const { return c->cpspec[st][m]; }
  __device__ float cpsf(int st, int m) const { return tb->cpspecf[st][m]; }
  __device__ double ovw(int s) const { return c->ovwmol[s]; }
  __device__ float ovwf(int s) const { return tb->ovwf[s]; }
  __device__ double eff(int b, int s) const { return c->effy3b[b][s]; }
  __device__ float efff(int b, int s) const { return tb->effyf[b][s]; }
  __device__ double clnten() const { return c->clnten; }
  __device__ bool dense() const { return tb->ratedense != 0; }
  __device__ double nuw(int st, int u) const { return tb->nuw[st][u]; }
  __device__ float nuwf(int st, int u) const { return tb->nuwf[st][u]; }
};

template <class T>
struct CtMech {
  static constexpr bool fixed = true;
  __device__ static constexpr int nst() { return T::nstep; }
  __device__ static constexpr int mb(int st) { return T::mblist[st]; }
  __device__ static constexpr int mg(int st) { return T::mglist[st]; }
  __device__ static constexpr int ml(int st) { return T::mllist[st]; }
  __device__ static constexpr int mt(int st) { return T::mtlist[st]; }
  __device__ static constexpr int ms(int st) { return T::mslist[st]; }
  __device__ static constexpr double rp(int st, int q) { return T::rparam[st][q]; }
  __device__ static constexpr double rl(int j, int q) { return T::rclind[j][q]; }
  __device__ static constexpr double rt(int j, int q) { return T::rctroe[j][q]; }
  __device__ static constexpr double rs(int j, int q) { return T::rcsrif[j][q]; }
  __device__ static constexpr int nrs(int st) { return T::nrslen[st]; }
  __device__ static constexpr int nrsp(int st, int m) { return T::nrspec[st][m]; }
  __device__ static constexpr int nrc(int st) { return T::nrclen[st]; }
  __device__ static constexpr int nrcp(int st, int m) { return T::nrcpec[st][m]; }
  __device__ static constexpr double crs(int st, int m) { return T::crspec[st][m]; }
  __device__ static constexpr float crsf(int st, int m) { return (float)T::crspec[st][m]; }
  __device__ static constexpr int nss(int st) { return T::nsslen[st]; }
  __device__ static constexpr int nssp(int st, int m) { return T::nsspec[st][m]; }
  __device__ static constexpr double dmu(int st, int m) { return T::diffmu[st][m]; }
  __device__ static constexpr double dmw(int st, int m) { return T::diffmw[st][m]; }
  __device__ static constexpr float dmwf(int st, int m) { return (float)T::diffmw[st][m]; }
  __device__ static constexpr int nps(int st) { return T::npslen[st]; }
  __device__ static constexpr int npsp(int st, int m) { return T::npspec[st][m]; }
  __device__ static constexpr int npc(int st) { return T::npclen[st]; }
  __device__ static constexpr int npcp(int st, int m) { return T::npcpec[st][m]; }
  __device__ static constexpr double cps(int st, int m) { return T::cpspec[st][m]; }
  __device__ static constexpr float cpsf(int st, int m) { return (float)T::cpspec[st][m]; }
  __device__ static constexpr double ovw(int s) { return T::ovwmol[s]; }
  __device__ static constexpr float ovwf(int s) { return (float)T::ovwmol[s]; }
  __device__ static constexpr double eff(int b, int s) { return T::effy3b[b][s]; }
  __device__ static constexpr float efff(int b, int s) { return (float)T::effy3b[b][s]; }
  __device__ static constexpr double clnten() { return T::clnten; }
  __device__ static constexpr bool dense() { return T::ratedense != 0; }
  __device__ static constexpr double nuw(int st, int u) { return T::nuw[st][u]; }
  __device__ static constexpr float nuwf(int st, int u) { return (float)T::nuw[st][u]; }
};

// f(st) for every step: unrolled for compile-time tables
template <class M, class Fn>
__device__ __forceinline__ void for_steps(const M &m, Fn f) {
  if constexpr (M::fixed) {
#pragma unroll
    for (int st = 0; st < M::nst(); st++) f(st);
  } else {
    for (int st = 0; st < m.nst(); st++) f(st);
  }
}

// One step in FP64: rate coefficients with falloff, forward and reverse
// products, added to rate[] (src/chrate.f:177-747).  cyv(s) and gbv(s) give
// the point's molar concentration and Gibbs function of species s, ex(x) the
// exponential.
template <int NS, class M, class CY, class GB, class EX>
__device__ __forceinline__ void chem_step64(const M &m, int st, double t, double lt, double ovt,
                                            const double *ry, CY cyv, GB gbv, double *rate, EX ex) {
  const double ysmall = 1.0e-30, ydenom = 1.0e-15;
  bool flthrd = true;
  double s3 = 0.0;
  const int mb = m.mb(st);
  if (mb != 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) s3 = s3 + ry[s] * (m.ovw(s) * m.eff(mb - 1, s));
  }
  double s2 = m.rp(st, 0) + m.rp(st, 1) * lt - m.rp(st, 2) * ovt;
  double s1 = ex(s2);
  const int ml = m.ml(st);
  if (ml != 0) {
    const int j = ml - 1;
    double s4 = m.rl(j, 0) + m.rl(j, 1) * lt - m.rl(j, 2) * ovt;
    s4 = ex(s4);
    double preduc = s3 * s4 / s1;
    double fornow = m.rl(j, 3) * preduc / (1.0 + preduc);  // rl[3] pre-rounded to REAL
    s1 = s1 * fornow;
    s2 = log(s1);
    flthrd = false;
  }
  const int mt = m.mt(st);
  if (mt != 0) {
    const int j = mt - 1;
    double omalph = 1.0 - m.rt(j, 3);
    double s4 = m.rt(j, 0) + m.rt(j, 1) * lt - m.rt(j, 2) * ovt;
    s4 = ex(s4);
    double preduc = s3 * s4 / s1;
    double trats1 = t * m.rt(j, 4), trats2 = -m.rt(j, 5) * ovt, trats3 = t * m.rt(j, 6);
    double ftcent = omalph * ex(trats3) + m.rt(j, 3) * ex(trats1) + ex(trats2);
    ftcent = log10(ftcent);
    double cfactr = m.rt(j, 7) + m.rt(j, 8) * ftcent;
    double enfact = m.rt(j, 9) + m.rt(j, 10) * ftcent;
    double fornow = log10(preduc) + cfactr;
    fornow = fornow / (enfact - m.rt(j, 11) * fornow);
    fornow = 1.0 + fornow * fornow;
    double fbroad = ftcent / fornow;
    fbroad = ex(fbroad * m.clnten());
    fornow = fbroad * preduc / (1.0 + preduc);
    s1 = s1 * fornow;
    s2 = log(s1);
    flthrd = false;
  }
  const int ms = m.ms(st);
  if (ms != 0) {
    const int j = ms - 1;
    double s4 = m.rs(j, 0) + m.rs(j, 1) * lt - m.rs(j, 2) * ovt;
    s4 = ex(s4);
    double preduc = s3 * s4 / s1;
    double fornow = m.rs(j, 4) * ovt;
    double fbroad = m.rs(j, 3) * ex(fornow);
    fornow = t * m.rs(j, 5);
    fbroad = fbroad + ex(fornow);
    fornow = log10(preduc);
    fornow = 1.0 / (1.0 + log10(fornow));
    fbroad = m.rs(j, 6) * ex(fornow * log(fbroad));
    fornow = fbroad * preduc / (1.0 + preduc);
    s1 = s1 * fornow;
    s2 = log(s1);
    flthrd = false;
  }
  for (int q = 0; q < m.nrs(st); q++) {
    double fornow = cyv(m.nrsp(st, q));
    fornow = fornow > 0.0 ? fornow : 0.0;
    s1 = s1 * fornow;
  }
  for (int q = 0; q < m.nrc(st); q++) {
    double fornow = cyv(m.nrcp(st, q));
    fornow = fornow > ysmall ? fornow : ysmall;
    fornow = ex(m.crs(st, q) * log(fornow));
    fornow = fornow / (1.0 + ydenom * fornow);
    s1 = s1 * fornow;
  }
  if (m.mg(st) != 0) {
    for (int q = 0; q < m.nss(st); q++) s2 = s2 + m.dmu(st, q) * gbv(m.nssp(st, q));
    s2 = ex(s2);
    for (int q = 0; q < m.nps(st); q++) {
      double fornow = cyv(m.npsp(st, q));
      fornow = fornow > 0.0 ? fornow : 0.0;
      s2 = s2 * fornow;
    }
    for (int q = 0; q < m.npc(st); q++) {
      double fornow = cyv(m.npcp(st, q));
      fornow = fornow > ysmall ? fornow : ysmall;
      fornow = ex(m.cps(st, q) * log(fornow));
      fornow = fornow / (1.0 + ydenom * fornow);
      s2 = s2 * fornow;
    }
    s1 = s1 - s2;
  }
  if (mb != 0 && flthrd) s1 = s1 * s3;
  if (m.dense()) {
    // every species once per list: absent species add s1*0, an exact zero;
    // with compile-time tables those terms are not formed at all
#pragma unroll
    for (int u = 0; u < NS; u++)
      if (!M::fixed || m.nuw(st, u) != 0.0) rate[u] = rate[u] + s1 * m.nuw(st, u);
  } else {
    for (int q = 0; q < m.nss(st); q++) {
      int sp = m.nssp(st, q);
      double add = s1 * m.dmw(st, q);
#pragma unroll
      for (int u = 0; u < NS; u++)
        if (u == sp) rate[u] = rate[u] + add;
    }
  }
}

template <int NS, bool FIX>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) k_prim(const Member *__restrict__ mem,
    const uint32_t *__restrict__ map, const sg_config *__restrict__ c, const Tables *__restrict__ tb) {
  int lb;
  const Member &M = member_at<L_POINT>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  // the 2^(j/64) table of exp_t, staged before any thread leaves
  __shared__ double xt[64];
  if (threadIdx.x < 64) xt[threadIdx.x] = kExp2j[threadIdx.x];
  __syncthreads();
  const long idx = (long)lb * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  const uint32_t bo = (uint32_t)idx * 8u;
  const long fs = P.fstride;
  const double *q = F.q;
  double *pr = F.prim;
  double d = fld(q, bo), ru = fld(q + fs, bo), rv = fld(q + 2 * fs, bo), rw = fld(q + 3 * fs, bo);
  double re = fld(q + 4 * fs, bo);
  double ry[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) ry[s] = fld(q + (5 + s) * fs, bo);
  // --- TEMPER (src/temper.f:58-160) ---
  const double ovd = 1.0 / d;   // one reciprocal for every division by rho
  double ukuk = (ru * ru + rv * rv + rw * rw) * ovd;
  double tc[6], td[6];
  tc[0] = 0.5 * ukuk - re;
#pragma unroll
  for (int m = 1; m <= 5; m++) { tc[m] = 0.0; td[m] = 0.0; }
  double s7 = 0.0;
  uint32_t idxo = fld(F.itndex, bo / 2u);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    temper_add(tb, s, itint_of(idxo, s), ry[s], tc, td);
    s7 = s7 + ry[s] * c->rgspec[s];
  }
  const double toltmp = (double)1.0e-10f;  // REAL literal, src/temper.f:33-34
  double t = fld(pr + P_T * fs, bo);
  int iter = 1;
  bool ok = true;
  for (;;) {
    double fp = tc[5], dp = td[5];
#pragma unroll
    for (int m = 4; m >= 1; m--) {
      fp = tc[m] + fp * t;
      dp = td[m] + dp * t;
    }
    fp = tc[0] + fp * t;
    double dt = -fp / dp;
    if (fabs(dt) > toltmp) {
      if (iter < 100) {
        t = t + dt;
        iter++;
        continue;
      }
      ok = false;
    }
    break;
  }
  if (!ok) {
    int i, j, k;
    coords(P, idx, i, j, k);
    set_error(F.err, SG_ETEMPER, i, j, k);
  }
  uint32_t nidx = 0;
  double cp = 0.0;
  int itn[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    int it = locate(c, s, t);
    itn[s] = it;
    nidx += (uint32_t)it << s;
    cp = cp + ry[s] * horner(tb->cp[s], it, t);
  }
  cp = cp * ovd;
  fld(F.itndex, bo / 2u) = nidx;
  SG_PHASE();
  double pres = t * s7;
  fld(pr + P_T * fs, bo) = t;
  fld(pr + P_P * fs, bo) = pres;
  // primitives
  fld(pr + P_U * fs, bo) = ru * ovd;
  fld(pr + P_V * fs, bo) = rv * ovd;
  fld(pr + P_W * fs, bo) = rw * ovd;
  fld(pr + P_E * fs, bo) = re * ovd;
  // --- conductivity (src/rhscal.f:515-573) ---
  double lam, transp, wmomix = 0.0;
  if (!P.mixavg) {
    double fornow = P.alamda * exp_t(P.rlamda * log(t), xt);
    lam = fornow * cp;
    transp = fornow;
  } else {
    transp = log(t / P.tdifgb);
    // the divisions X_k / lambda_k share one (batch_recip; a few ulp from the
    // reference's quotients)
    double c1 = 0.0, c2 = 0.0, c3 = 0.0, ictr[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double ctr = exp_t(trfit(c->condco[s], transp), xt);
      ictr[s] = ctr;
      c1 = c1 + (ry[s] * c->ovwmol[s]) * ctr;
    }
    batch_recip<NS>(ictr);
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double fornow = ry[s] * c->ovwmol[s];
      c2 = c2 + fornow * ictr[s];
      c3 = c3 + fornow;
    }
    c3 = 1.0 / c3;
    c1 = c1 * c3;
    c2 = c2 * c3;
    lam = 0.5 * (c1 + 1.0 / c2);
    wmomix = d * c3;
  }
  fld(pr + P_LAM * fs, bo) = lam;
  if (P.mixw || P.mixp) {
    // ln(mixture molar mass), ln(pressure) and the reciprocal molar mass for
    // the molar-mass and pressure diffusion terms (src/rhscal.f:951-985, 2374)
    fld(pr + P_LNW * fs, bo) = log(wmomix);
    fld(pr + P_LNP * fs, bo) = log(pres);
    fld(pr + P_OVW * fs, bo) = 1.0 / wmomix;
  }
  SG_PHASE();
  // --- chemistry (src/chrate.f:146-760) on rho Y ---
  {
    double rate[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) rate[s] = 0.0;
    const double lt = log(t), ovt = 1.0 / t;
    // Gibbs functions of all species at this temperature, each formed as
    // src/chrate.f:612-632 forms it for every step it enters
    double gibbs[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      int it = itn[s];
      gibbs[s] = gibbs_of(tb, s, it, t, ovt, lt);
    }
    if constexpr (FIX) {
      using MT = CtMech<MechH2>;
      double cy[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) cy[s] = ry[s] * MT::ovw(s);
      for_steps(MT{}, [&](int st) {
        chem_step64<NS>(MT{}, st, t, lt, ovt, ry, [&](int s) { return cy[s]; }, [&](int s) { return gibbs[s]; },
                        rate, [&](double x) { return exp_t(x, xt); });
      });
    } else {
      // Molar concentrations and Gibbs functions of this point in LDS, so
      // that a step's species lists (uniform, known only at run time) index
      // them with one LDS read instead of a select over all species.
      __shared__ double lcy[NS][BLK], lgb[NS][BLK];
      const int tl = threadIdx.x;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        lcy[s][tl] = ry[s] * c->ovwmol[s];
        lgb[s][tl] = gibbs[s];
      }
      const RtMech mr{c, tb};
      for_steps(mr, [&](int st) {
        chem_step64<NS>(mr, st, t, lt, ovt, ry, [&](int s) { return lcy[s][tl]; }, [&](int s) { return lgb[s][tl]; },
                        rate, [&](double x) { return exp_t(x, xt); });
      });
    }
#pragma unroll
    for (int s = 0; s < NS; s++) fld(F.rate + s * fs, bo) = rate[s];
  }
  // --- mass fractions and enthalpies (src/rhscal.f:1444-1470) ---
  double y[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) y[s] = ry[s] * ovd;
  double hmix = 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    fld(pr + prim_y(P, s), bo) = y[s];
    int it = itn[s];
    double h = horner_c(tb->ch[s], tb->hch[s], it, t);
    fld(pr + prim_h(P, s), bo) = h;
    hmix = hmix + h * y[s];
  }
  fld(pr + P_HMIX * fs, bo) = hmix;
  SG_PHASE();
  if (!P.mixavg) {
#pragma unroll
    for (int s = 0; s < NS; s++) fld(pr + prim_rd(P, s), bo) = transp * c->olewis[s];
    fld(pr + P_MU * fs, bo) = transp * P.prantl;
  } else {
    // binary diffusion (src/rhscal.f:1311-1346), with Y_j + tol formed again
    // where it is used rather than held through the pair loop
    double c1t = 0.0;
#pragma unroll
    for (int js = 0; js < NS; js++) c1t = c1t + (y[js] + 1.0e-12);
    // The reference's term (Y_j + tol) W_j^-1 / CTRANS, CTRANS = exp(f) p_g / p,
    // is formed as ((Y_j + tol) W_j^-1 (p / p_g)) * exp(-f): no division or
    // pressure product per pair (a few ulp from the reference's operation
    // order), and the species' final divisions share one (batch_recip).
    const double ppd = pres / P.pdifgb;
    if (tb->diffsym) {
      double yw2p[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) yw2p[s] = ((y[s] + 1.0e-12) * c->ovwmol[s]) * ppd;
      // D_ij = D_ji: one fit per unordered pair a < b, its term added to both
      // sums.  Visiting the pairs row by row delivers the terms of every sum
      // in increasing species order, the reference's order.  The reference
      // sums over every j and subtracts the j = k term; the sums here skip
      // it (no fit and exp for the 9 self pairs; the same sum up to the
      // rounding of that addition and subtraction).
      double cs[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) cs[s] = 0.0;
#pragma unroll
      for (int a = 0; a < NS; a++) {
#pragma unroll
        for (int b = a + 1; b < NS; b++) {
          const double ictd = exp_t(-trfit(c->diffco[a][b], transp), xt);
          cs[a] = cs[a] + yw2p[b] * ictd;
          cs[b] = cs[b] + yw2p[a] * ictd;
        }
      }
      double c2[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) c2[s] = cs[s] * wmomix;
      batch_recip<NS>(c2);
#pragma unroll
      for (int s = 0; s < NS; s++) fld(pr + prim_rd(P, s), bo) = d * (c1t - (y[s] + 1.0e-12)) * c2[s];
    } else {
      double yt[NS], yw2[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        yt[s] = y[s] + 1.0e-12;
        yw2[s] = yt[s] * c->ovwmol[s];
      }
#pragma unroll
      for (int s = 0; s < NS; s++) {
        double c2 = 0.0, self = 0.0;
#pragma unroll
        for (int js = 0; js < NS; js++) {
          const double term = yw2[js] * (exp_t(-trfit(c->diffco[s][js], transp), xt) * ppd);
          c2 = c2 + term;
          if (js == s) self = term;
        }
        double c1 = c1t - yt[s];
        c2 = c2 - self;
        c2 = c2 * wmomix;
        fld(pr + prim_rd(P, s), bo) = d * c1 / c2;
      }
    }
    SG_PHASE();
    // Wilke viscosity (src/rhscal.f:3854-3891).  sqrt(mu_s / mu_j) is formed
    // as sqrt(mu_s) / sqrt(mu_j), with sqrt(mu_s) = exp(fit_s / 2) (one exp
    // instead of exp and square root) and mu_s = sqrt(mu_s)^2; the
    // reciprocals 1/sqrt(mu_j) share one division, and so do the divisions
    // mu_s / sum_s (batch_recip).
    double mus[NS], sq[NS], rsq[NS], yw[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double hf = 0.5 * trfit(c->viscco[s], transp);
      sq[s] = exp_t(hf, xt);
      rsq[s] = sq[s];
      mus[s] = sq[s] * sq[s];
      yw[s] = y[s] * c->ovwmol[s];
    }
    batch_recip<NS>(rsq);
    double c1 = 0.0, wc2[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double c2 = 0.0;
#pragma unroll
      for (int js = 0; js < NS; js++) {
        double fornow = js == s ? 1.0 : sq[s] * rsq[js];
        fornow = 1.0 + fornow * c->wilko2[s][js];
        fornow = c->wilko1[s][js] * fornow * fornow;
        c2 = c2 + yw[js] * fornow;
      }
      wc2[s] = c2;
    }
    batch_recip<NS>(wc2);
#pragma unroll
    for (int s = 0; s < NS; s++) c1 = c1 + yw[s] * (mus[s] * wc2[s]);
    fld(pr + P_MU * fs, bo) = c1;
  }
  // face data for NSCBC: cp and mixture gas constant (src/rhscal.f:72-131)
  if (P.inflow_xl || P.outflow_xr) {
    int i, j, k;
    coords(P, idx, i, j, k);
    if (i == 0 || i == P.nx - 1) {
      long nf = (long)P.ny * P.nz;
      double *fc = F.face + (i == 0 ? 0 : (long)nface(P.ns) * nf);
      long fq = (long)k * P.ny + j;
      fc[F_CP * nf + fq] = cp;
      fc[F_RR * nf + fq] = s7 * ovd;
    }
  }
}

// Mixed form, after the point kernel: FP32 holds exp(x) only for x in about
// (-87, 88), and outside the mechanism's range [P.chlo, P.chhi] (worked out
// on the host) a rate coefficient could underflow to zero before a falloff
// division, or overflow before its product with small concentrations, and
// give NaN (or lose a reverse rate) where FP64 gives finite values.  The
// points whose temperature lies outside that range get their rates formed
// again here in FP64 (chem_step64 with the library exp, from the FP64 rho Y_k
// and the stored FP32 temperature) and stored in FP32; a block with no such
// point returns after reading its temperatures.
template <int NS, bool FIX>
__global__ void __launch_bounds__(BLK) k_rate64(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                                const sg_config *__restrict__ c, const Tables *__restrict__ tb) {
  int lb;
  const Member &M = member_at<L_POINT>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  const long idx = (long)lb * BLK + threadIdx.x;
  const long fs = P.fstride;
  const double t = idx < P.npts ? (double)((const float *)F.prim)[P_T * fs + idx] : 1000.0;
  const bool out = idx < P.npts && !(t >= P.chlo && t <= P.chhi);
  if (!__syncthreads_or(out)) return;
  if (!out) return;   // no barrier below
  double ry[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) ry[s] = F.q[(5 + s) * fs + idx];
  const double lt = log(t), ovt = 1.0 / t;
  double gibbs[NS], rate[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    gibbs[s] = gibbs_of(tb, s, locate(c, s, t), t, ovt, lt);
    rate[s] = 0.0;
  }
  auto ex = [](double x) { return exp(x); };
  if constexpr (FIX) {
    using MT = CtMech<MechH2>;
    for_steps(MT{}, [&](int st) {
      chem_step64<NS>(MT{}, st, t, lt, ovt, ry, [&](int s) { return ry[s] * MT::ovw(s); },
                      [&](int s) { return gibbs[s]; }, rate, ex);
    });
  } else {
    // species lists index the point's values at run time: from LDS, one
    // slot per thread (as in k_prim)
    __shared__ double lcy[NS][BLK], lgb[NS][BLK];
    const int tl = threadIdx.x;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      lcy[s][tl] = ry[s] * c->ovwmol[s];
      lgb[s][tl] = gibbs[s];
    }
    const RtMech mr{c, tb};
    for_steps(mr, [&](int st) {
      chem_step64<NS>(mr, st, t, lt, ovt, ry, [&](int s) { return lcy[s][tl]; }, [&](int s) { return lgb[s][tl]; },
                      rate, ex);
    });
  }
  float *rt = (float *)F.rate;
#pragma unroll
  for (int s = 0; s < NS; s++) rt[s * fs + idx] = (float)rate[s];
}

// Soret / Dufour, after the point kernel: ln(T/TDIFGB), the thermal
// diffusion ratio TDRMIX of each light species relative to the other light
// species (src/rhscal.f:1353-1410) and T * TDRMIX, from the stage register
// and the stored temperature (rounded to FP32 in the mixed form).  A kernel
// of its own, so that the point kernels keep their registers.
template <int NS, class R>
__global__ void __launch_bounds__(BLK) k_tdr(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                             const sg_config *__restrict__ c) {
  int lb;
  const Member &M = member_at<L_POINT>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  if (!P.ltmask) return;
  const Fields F = global_fields(M.F);
  const long idx = (long)lb * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  const long fs = P.fstride;
  R *pr = (R *)F.prim;
  const double d = F.q[idx], ovd = 1.0 / d;
  double ry[NS];
  double c3 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    ry[s] = F.q[(5 + s) * fs + idx];
    c3 = c3 + ry[s] * c->ovwmol[s];
  }
  const double wmomix = d * (1.0 / c3);
  const double t = (double)pr[P_T * fs + idx];
  pr[P_LNT * fs + idx] = (R)log(t / P.tdifgb);
  const double combo2 = t / P.tdifgb;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    if (!((P.ltmask >> k) & 1u)) continue;
    double tdr = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) {
      if (j == k || !((P.ltmask >> j) & 1u)) continue;
      tdr = tdr + (ry[j] * ovd) * c->ovwmol[j] * wmomix * trfit(c->tdrcco[k][j], combo2);
    }
    tdr = tdr * (ry[k] * ovd) * wmomix * c->ovwmol[k];
    pr[prim_tdr(P, k) + idx] = (R)tdr;
    pr[prim_tt(P, k) + idx] = (R)(t * tdr);
  }
}

// Optically thin radiation (src/radcal.f:36-111, called at src/rhscal.f:810):
// the energy equation loses 4 sigma K_P (T^4 - T_ref^4), with the Planck mean
// absorption coefficient K_P = sum over the radiating species r of a_r(T) p_r,
// p_r = rho Y_r R_r T (the conserved rho Y_r, as RHSCAL holds it there).  Added
// in FP64 to the energy accumulator the x pass has just written.
template <int NS, class R>
__global__ void __launch_bounds__(BLK) k_radn(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                              const sg_config *__restrict__ c) {
  int lb;
  const Member &M = member_at<L_POINT>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  if (!P.radn) return;
  const Fields F = global_fields(M.F);
  const long idx = (long)lb * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  const long fs = P.fstride;
  const double t = (double)((const R *)F.prim)[P_T * fs + idx];
  double kp = 0.0;
  for (int r = 0; r < c->nsprad; r++) {
    const int nk = c->nkprad[r];
    double pl = c->akprad[r][nk - 1];
    for (int q = nk - 2; q >= 0; q--) pl = pl * t + c->akprad[r][q];
    const int k = c->nsprid[r];
    kp = kp + pl * (F.q[(5 + k) * fs + idx] * c->rgspec[k] * t);
  }
  const double t4 = t * t * t * t, tr = c->trefrn;
  const double trfrth = tr * tr * tr * tr;
  const double foursb = 4.0 * 5.67037321e-08;   // FOUR * STEFBO (src/com_senga2.h:221-222)
  R *A = (R *)F.acc;
  A[4 * fs + idx] += (R)(-(foursb * kp * (t4 - trfrth)));
}

// FP32 quotient and reciprocal from the hardware reciprocal (v_rcp_f32, about
// 1 ulp; the quotient within 2 ulp) instead of the IEEE-exact FP32 division
// sequence (~10 instructions each), for the FP32 transport and chemistry of
// the mixed-precision point kernel.  Operands there are normal numbers.
__device__ __forceinline__ float rf(float y) { return __builtin_amdgcn_rcpf(y); }
__device__ __forceinline__ float qf(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

// One step of the mixed-precision form: the rate exponents and Gibbs sums
// assembled in FP64, everything else in FP32 (as chem_step64 otherwise).
template <int NS, class M, class CY, class GB>
__device__ __forceinline__ void chem_stepmx(const M &m, int st, float tf, double lt, double ovt, const float *ryf,
                                            CY cyv, GB gbv, float *rate) {
  const float ysmall = 1.0e-30f, ydenom = 1.0e-15f;
  bool flthrd = true;
  float s3 = 0.0f;
  const int mb = m.mb(st);
  if (mb != 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) s3 = s3 + ryf[s] * (m.ovwf(s) * m.efff(mb - 1, s));
  }
  double s2d = m.rp(st, 0) + m.rp(st, 1) * lt - m.rp(st, 2) * ovt;
  float s1 = __expf((float)s2d);
  const int ml = m.ml(st);
  if (ml != 0) {
    const int j = ml - 1;
    float s4 = __expf((float)(m.rl(j, 0) + m.rl(j, 1) * lt - m.rl(j, 2) * ovt));
    float preduc = qf(s3 * s4, s1);
    float fornow = qf((float)m.rl(j, 3) * preduc, 1.0f + preduc);
    s1 = s1 * fornow;
    s2d = (double)__logf(s1);
    flthrd = false;
  }
  const int mt = m.mt(st);
  if (mt != 0) {
    const int j = mt - 1;
    float omalph = 1.0f - (float)m.rt(j, 3);
    float s4 = __expf((float)(m.rt(j, 0) + m.rt(j, 1) * lt - m.rt(j, 2) * ovt));
    float preduc = qf(s3 * s4, s1);
    float trats1 = tf * (float)m.rt(j, 4), trats2 = (float)(-m.rt(j, 5) * ovt), trats3 = tf * (float)m.rt(j, 6);
    float ftcent = omalph * __expf(trats3) + (float)m.rt(j, 3) * __expf(trats1) + __expf(trats2);
    ftcent = __log10f(ftcent);
    float cfactr = (float)m.rt(j, 7) + (float)m.rt(j, 8) * ftcent;
    float enfact = (float)m.rt(j, 9) + (float)m.rt(j, 10) * ftcent;
    float fornow = __log10f(preduc) + cfactr;
    fornow = qf(fornow, enfact - (float)m.rt(j, 11) * fornow);
    fornow = 1.0f + fornow * fornow;
    float fbroad = qf(ftcent, fornow);
    fbroad = __expf(fbroad * (float)m.clnten());
    fornow = qf(fbroad * preduc, 1.0f + preduc);
    s1 = s1 * fornow;
    s2d = (double)__logf(s1);
    flthrd = false;
  }
  const int ms = m.ms(st);
  if (ms != 0) {
    const int j = ms - 1;
    float s4 = __expf((float)(m.rs(j, 0) + m.rs(j, 1) * lt - m.rs(j, 2) * ovt));
    float preduc = qf(s3 * s4, s1);
    float fornow = (float)(m.rs(j, 4) * ovt);
    float fbroad = (float)m.rs(j, 3) * __expf(fornow);
    fornow = tf * (float)m.rs(j, 5);
    fbroad = fbroad + __expf(fornow);
    fornow = __log10f(preduc);
    fornow = rf(1.0f + __log10f(fornow));
    fbroad = (float)m.rs(j, 6) * __expf(fornow * __logf(fbroad));
    fornow = qf(fbroad * preduc, 1.0f + preduc);
    s1 = s1 * fornow;
    s2d = (double)__logf(s1);
    flthrd = false;
  }
  for (int q = 0; q < m.nrs(st); q++) {
    float fornow = cyv(m.nrsp(st, q));
    fornow = fornow > 0.0f ? fornow : 0.0f;
    s1 = s1 * fornow;
  }
  for (int q = 0; q < m.nrc(st); q++) {
    float fornow = cyv(m.nrcp(st, q));
    fornow = fornow > ysmall ? fornow : ysmall;
    fornow = __expf(m.crsf(st, q) * __logf(fornow));
    fornow = qf(fornow, 1.0f + ydenom * fornow);
    s1 = s1 * fornow;
  }
  if (m.mg(st) != 0) {
    for (int q = 0; q < m.nss(st); q++) s2d = s2d + m.dmu(st, q) * gbv(m.nssp(st, q));
    float s2 = __expf((float)s2d);
    for (int q = 0; q < m.nps(st); q++) {
      float fornow = cyv(m.npsp(st, q));
      fornow = fornow > 0.0f ? fornow : 0.0f;
      s2 = s2 * fornow;
    }
    for (int q = 0; q < m.npc(st); q++) {
      float fornow = cyv(m.npcp(st, q));
      fornow = fornow > ysmall ? fornow : ysmall;
      fornow = __expf(m.cpsf(st, q) * __logf(fornow));
      fornow = qf(fornow, 1.0f + ydenom * fornow);
      s2 = s2 * fornow;
    }
    s1 = s1 - s2;
  }
  if (mb != 0 && flthrd) s1 = s1 * s3;
  if (m.dense()) {
#pragma unroll
    for (int u = 0; u < NS; u++)
      if (!M::fixed || m.nuwf(st, u) != 0.0f) rate[u] = rate[u] + s1 * m.nuwf(st, u);
  } else {
    for (int q = 0; q < m.nss(st); q++) {
      int sp = m.nssp(st, q);
      float add = s1 * m.dmwf(st, q);
#pragma unroll
      for (int u = 0; u < NS; u++)
        if (u == sp) rate[u] = rate[u] + add;
    }
  }
}

// Mixed-precision point physics: the temperature Newton iteration, the
// primitive variables and the enthalpies are formed in FP64 as in k_prim and
// stored in FP32; transport and chemistry run in FP32 with the hardware
// exp2/log2 based intrinsics, except that the rate exponents (ln k = ln A +
// n ln T - E/RT, Gibbs sums) are assembled in FP64 before the FP32
// exponential, since they are differences of O(10..100) terms.
template <int NS, bool FIX>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) k_prim_mx(const Member *__restrict__ mem,
    const uint32_t *__restrict__ map, const sg_config *__restrict__ c, const Tables *__restrict__ tb) {
  int lb;
  const Member &M = member_at<L_POINT>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  const long idx = (long)lb * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  const uint32_t bo = (uint32_t)idx * 8u, b4 = (uint32_t)idx * 4u;
  const long fs = P.fstride;
  const double *q = F.q;
  float *pr = (float *)F.prim;
  double d = fld(q, bo), ru = fld(q + fs, bo), rv = fld(q + 2 * fs, bo), rw = fld(q + 3 * fs, bo);
  double re = fld(q + 4 * fs, bo);
  double ry[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) ry[s] = fld(q + (5 + s) * fs, bo);
  // --- TEMPER (src/temper.f:58-160) ---
  const double ovd = 1.0 / d;   // one reciprocal for every division by rho
  double ukuk = (ru * ru + rv * rv + rw * rw) * ovd;
  double tc[6], td[6];
  tc[0] = 0.5 * ukuk - re;
#pragma unroll
  for (int m = 1; m <= 5; m++) { tc[m] = 0.0; td[m] = 0.0; }
  double s7 = 0.0;
  uint32_t idxo = fld(F.itndex, bo / 2u);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    temper_add(tb, s, itint_of(idxo, s), ry[s], tc, td);
    s7 = s7 + ry[s] * c->rgspec[s];
  }
  const double toltmp = (double)1.0e-10f;  // REAL literal, src/temper.f:33-34
  double t = (double)fld(pr + P_T * fs, b4);
  int iter = 1;
  bool ok = true;
  for (;;) {
    double fp = tc[5], dp = td[5];
#pragma unroll
    for (int m = 4; m >= 1; m--) {
      fp = tc[m] + fp * t;
      dp = td[m] + dp * t;
    }
    fp = tc[0] + fp * t;
    double dt = -fp / dp;
    if (fabs(dt) > toltmp) {
      if (iter < 100) {
        t = t + dt;
        iter++;
        continue;
      }
      ok = false;
    }
    break;
  }
  if (!ok) {
    int i, j, k;
    coords(P, idx, i, j, k);
    set_error(F.err, SG_ETEMPER, i, j, k);
  }
  uint32_t nidx = 0;
  double cp = 0.0;
  int itn[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    int it = locate(c, s, t);
    itn[s] = it;
    nidx += (uint32_t)it << s;
    cp = cp + ry[s] * horner(tb->cp[s], it, t);
  }
  cp = cp * ovd;
  fld(F.itndex, bo / 2u) = nidx;
  SG_PHASE();
  double pres = t * s7;
  fld(pr + P_T * fs, b4) = (float)(t);
  fld(pr + P_P * fs, b4) = (float)(pres - P.pofs);
  // primitives
  fld(pr + P_U * fs, b4) = (float)(ru * ovd);
  fld(pr + P_V * fs, b4) = (float)(rv * ovd);
  fld(pr + P_W * fs, b4) = (float)(rw * ovd);
  fld(pr + P_E * fs, b4) = (float)(re * ovd);
  // --- conductivity (src/rhscal.f:515-573), FP32 ---
  const float tf = (float)t;
  float lam = 0.0f, transp = 0.0f, wmomix = 0.0f;
  if (!P.mixavg) {
    float fornow = (float)P.alamda * __expf((float)P.rlamda * __logf(tf));
    lam = fornow * (float)cp;
    transp = fornow;
  }
  float ryf[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) ryf[s] = (float)ry[s];
  const float df = (float)d;
  if (P.mixavg) {
    transp = __logf(tf * (float)(1.0 / P.tdifgb));
    float c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      float f = tb->condf[s][SG_NTRC - 1];
#pragma unroll
      for (int m = SG_NTRC - 2; m >= 0; m--) f = fma_u(f, transp, tb->condf[s][m]);
      float ctr = __expf(f);
      float fornow = ryf[s] * tb->ovwf[s];
      c1 = c1 + fornow * ctr;
      c2 = c2 + qf(fornow, ctr);
      c3 = c3 + fornow;
    }
    c3 = rf(c3);
    c1 = c1 * c3;
    c2 = c2 * c3;
    lam = 0.5f * (c1 + rf(c2));
    wmomix = df * c3;
  }
  fld(pr + P_LAM * fs, b4) = lam;
  if (P.mixw || P.mixp) {
    // ln(mixture molar mass), ln(pressure) and the reciprocal molar mass for
    // the molar-mass and pressure diffusion terms (src/rhscal.f:951-985, 2374),
    // formed in FP64
    double sw = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) sw = sw + ry[s] * c->ovwmol[s];
    const double wm = d * (1.0 / sw);
    fld(pr + P_LNW * fs, b4) = (float)log(wm);
    fld(pr + P_LNP * fs, b4) = (float)log(pres);
    fld(pr + P_OVW * fs, b4) = (float)(1.0 / wm);
  }
  SG_PHASE();
  // --- chemistry (src/chrate.f:146-760) on rho Y ---
  {
    float rate[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) rate[s] = 0.0f;
    const double lt = log(t), ovt = 1.0 / t;
    double gibbs[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      int it = itn[s];
      gibbs[s] = gibbs_of(tb, s, it, t, ovt, lt);
    }
    if constexpr (FIX) {
      using MT = CtMech<MechH2>;
      float cy[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) cy[s] = ryf[s] * MT::ovwf(s);
      for_steps(MT{}, [&](int st) {
        chem_stepmx<NS>(MT{}, st, tf, lt, ovt, ryf, [&](int s) { return cy[s]; }, [&](int s) { return gibbs[s]; },
                        rate);
      });
    } else {
      __shared__ float lcy[NS][BLK];
      __shared__ double lgb[NS][BLK];
      const int tl = threadIdx.x;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        lcy[s][tl] = ryf[s] * tb->ovwf[s];
        lgb[s][tl] = gibbs[s];
      }
      const RtMech mr{c, tb};
      for_steps(mr, [&](int st) {
        chem_stepmx<NS>(mr, st, tf, lt, ovt, ryf, [&](int s) { return lcy[s][tl]; }, [&](int s) { return lgb[s][tl]; },
                        rate);
      });
    }
#pragma unroll
    for (int s = 0; s < NS; s++) fld((float *)F.rate + s * fs, b4) = rate[s];
  }
  // --- mass fractions and enthalpies (src/rhscal.f:1444-1470), FP64, stored FP32 ---
  float y[NS];
  double hmix = 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const double yd = ry[s] * ovd;
    y[s] = (float)yd;
    fld(pr + prim_y(P, s), b4) = y[s];
    int it = itn[s];
    double h = horner_c(tb->ch[s], tb->hch[s], it, t);
    fld(pr + prim_h(P, s), b4) = (float)h;
    hmix = hmix + h * yd;
  }
  fld(pr + P_HMIX * fs, b4) = (float)hmix;
  SG_PHASE();
  if (!P.mixavg) {
#pragma unroll
    for (int s = 0; s < NS; s++) fld(pr + prim_rd(P, s), b4) = transp * tb->olewf[s];
    fld(pr + P_MU * fs, b4) = transp * (float)P.prantl;
  } else {
    // binary diffusion (src/rhscal.f:1311-1346), FP32
    float yt[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) yt[s] = y[s] + 1.0e-12f;
    float c1t = 0.0f;
#pragma unroll
    for (int js = 0; js < NS; js++) c1t = c1t + yt[js];
    const float ppd = (float)(pres / P.pdifgb);
    float yw2[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) yw2[s] = yt[s] * tb->ovwf[s];
    auto fit = [&](const float *a) {
      float f = a[SG_NTRC - 1];
#pragma unroll
      for (int m = SG_NTRC - 2; m >= 0; m--) f = fma_u(f, transp, a[m]);
      return f;
    };
    if (tb->diffsym) {
      // as in the FP64 kernel: one fit per unordered pair a < b, the pressure
      // ratio folded into the pair weights, no self pairs
      float cs[NS], yw2p[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        cs[s] = 0.0f;
        yw2p[s] = yw2[s] * ppd;
      }
#pragma unroll
      for (int a = 0; a < NS; a++) {
#pragma unroll
        for (int b = a + 1; b < NS; b++) {
          const float ictd = __expf(-fit(tb->difff[a][b]));
          cs[a] = cs[a] + yw2p[b] * ictd;
          cs[b] = cs[b] + yw2p[a] * ictd;
        }
      }
#pragma unroll
      for (int s = 0; s < NS; s++) {
        float c1 = c1t - yt[s];
        float c2 = cs[s] * wmomix;
        fld(pr + prim_rd(P, s), b4) = qf(df * c1, c2);
      }
    } else {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        float c2 = 0.0f, self = 0.0f;
#pragma unroll
        for (int js = 0; js < NS; js++) {
          const float term = yw2[js] * (__expf(-fit(tb->difff[s][js])) * ppd);
          c2 = c2 + term;
          if (js == s) self = term;
        }
        float c1 = c1t - yt[s];
        c2 = (c2 - self) * wmomix;
        fld(pr + prim_rd(P, s), b4) = qf(df * c1, c2);
      }
    }
    SG_PHASE();
    // Wilke viscosity (src/rhscal.f:3854-3891), FP32
    float mus[NS], sq[NS], rsq[NS], yw[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      mus[s] = __expf(fit(tb->viscf[s]));
      sq[s] = __builtin_amdgcn_sqrtf(mus[s]);
      rsq[s] = rf(sq[s]);
      yw[s] = y[s] * tb->ovwf[s];
    }
    float c1 = 0.0f;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      float c2 = 0.0f;
#pragma unroll
      for (int js = 0; js < NS; js++) {
        float fornow = js == s ? 1.0f : sq[s] * rsq[js];
        fornow = 1.0f + fornow * tb->wilk2f[s][js];
        fornow = tb->wilk1f[s][js] * fornow * fornow;
        c2 = c2 + yw[js] * fornow;
      }
      c1 = c1 + yw[s] * qf(mus[s], c2);
    }
    fld(pr + P_MU * fs, b4) = c1;
  }
  // face data for NSCBC: cp and mixture gas constant (src/rhscal.f:72-131)
  if (P.inflow_xl || P.outflow_xr) {
    int i, j, k;
    coords(P, idx, i, j, k);
    if (i == 0 || i == P.nx - 1) {
      long nf = (long)P.ny * P.nz;
      double *fc = F.face + (i == 0 ? 0 : (long)nface(P.ns) * nf);
      long fq = (long)k * P.ny + j;
      fc[F_CP * nf + fq] = cp;
      fc[F_RR * nf + fq] = s7 * ovd;
    }
  }
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Direction pass: all derivatives along DIR and the RHS terms separable in
// DIR.  The block owns a tile of points; every field (or product of fields)
// is staged once into LDS with a 5-deep halo along DIR, then each thread
// evaluates its stencils from LDS at compile-time offsets.  Fields staged:
// rho u_d (resident for the whole pass), e, T, lambda, u, v, w, p - p_ref,
// mu, h_mix and per species Y, rho D, h; the products with rho u_d that the
// split convective forms difference (e rho u_d, Y rho u_d and the momentum
// fluxes rho u_d u_b) are formed inside the stencils from two tiles.
// Bounded x direction: the x pass runs as two launches.  Edge tiles, PASS_EW
// points wide at each end of a row, hold every closure row (i < 5, i > nx - 6)
// and evaluate all their stencils in the uniform row form; interior tiles
// cover [PASS_EW, nx - PASS_EW), where every stencil is the interior one and
// reaches no further than 5 points into the edge tiles.  Neither kernel then
// carries the other's stencil code (with both in one kernel, and the form
// chosen per wave, the bounded x pass took 0.345 ms at 128^3 in FP64; the
// interior form alone 0.298 ms, the row form alone 0.331 ms).  xp: 0 every
// x tile of a row (periodic x, and the y and z passes); 1 the interior tiles
// of a bounded row; 2 its edge tiles.  Rows shorter than 2 PASS_EW are all edge
// tiles.
constexpr int PASS_EW = 16;
__host__ __device__ inline int pass_xtiles(int nx, int xp, int tx) {
  if (xp == 0) return (nx + tx - 1) / tx;
  if (xp == 1) return nx >= 2 * PASS_EW ? (nx - 2 * PASS_EW + tx - 1) / tx : 0;
  return nx >= 2 * PASS_EW ? 2 : (nx + PASS_EW - 1) / PASS_EW;
}
// first x index of x tile bx
__host__ __device__ inline int pass_x0(int nx, int xp, int tx, int bx) {
  if (xp == 0) return bx * tx;
  if (xp == 1) return PASS_EW + bx * tx;
  return nx >= 2 * PASS_EW ? (bx == 0 ? 0 : nx - PASS_EW) : bx * PASS_EW;
}

template <int DIR, int NW = 0, int XP = 0>
struct PassTile {
  // Wide tiles (NW = 0), x: 32x8 points (halo 42/32 along x); y, z: 32x16
  // points, 512 threads, so that the 10-point halo along the pass direction
  // costs 26/16 re-reads.  Narrow tiles (NW = 1) have rows of 16 points and
  // twice the rows: no idle lanes when nx is an odd multiple of 16, and a
  // cheaper y/z halo (42/32), at half-length coalesced rows.  The autotuner
  // picks one per direction.  Edge tiles of a bounded x row (XP = 2):
  // PASS_EW x 16 points.
  static constexpr int TX = XP == 2 ? PASS_EW : (NW ? 16 : 32);
  static constexpr int TY = XP == 2 ? 16 : (DIR == 0 ? (NW ? 16 : 8) : (DIR == 1 ? (NW ? 32 : 16) : 1));
  static constexpr int TZ = DIR == 2 ? (NW ? 32 : 16) : 1;
  static constexpr int NT = TX * TY * TZ;   // threads per block
  static constexpr int LX = TX + (DIR == 0 ? 10 : 0);
  static constexpr int LY = TY + (DIR == 1 ? 10 : 0);
  static constexpr int LZ = TZ + (DIR == 2 ? 10 : 0);
  static constexpr int LN = LX * LY * LZ;
  static constexpr int LSTR = DIR == 0 ? 1 : (DIR == 1 ? LX : LX * LY);   // LDS stride along DIR
};

// Occupancy: FP64 passes need their 128 registers (4 waves per SIMD; at 5
// the spills made the y and z passes 40% slower).  The FP32 passes are held
// to 80 registers for 6 waves per SIMD (3 blocks per CU), which the
// latency-bound staging pipeline turns into 7-10% shorter y and z passes and
// 2% shorter interior x passes; the bounded-row edge tiles, which would spill
// heavily at 80, keep 4, and so do the passes with the molar-mass and
// pressure diffusion terms (EXT), which need more.
template <int DIR, bool BND, int NS, class R, int NW, int XP = 0, bool EXT = false>
__global__ void __launch_bounds__((PassTile<DIR, NW, XP>::NT))
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && XP != 2 && !EXT ? 6 : 4))) k_pass(const Member *__restrict__ mem,
                                                                                                       const uint32_t *__restrict__ map) {
  static_assert(XP == 0 || (DIR == 0 && (XP == 1) != BND), "interior x tiles take the periodic form, edge tiles the bounded one");
  using T_ = PassTile<DIR, NW, XP>;
  constexpr int TX = T_::TX, TY = T_::TY, LX = T_::LX, LY = T_::LY, LN = T_::LN, LSTR = T_::LSTR;
  constexpr int NTH = T_::NT;
  // LN staged values per field, padded to whole block-sized rows so that
  // every thread loads and stores unconditionally
  constexpr int NIT = (LN + NTH - 1) / NTH;
  // Two buffers of up to STW fields each: a stage stages one field, or up to
  // three fields under one barrier (e, T and lambda; u, v and w; p and mu; a
  // species' Y_k, rho D_k and h_k).  Grouped stages were measured faster in
  // every direction than one field per barrier (FP64 x pass at 128^3
  // periodic: 0.279 -> 0.265 ms), a few FP64 spills included.
  constexpr int STW = 3;
  __shared__ R tile[2][STW][NIT * NTH];
  const int t = threadIdx.x;
  const int lx = t % TX, ly = (t / TX) % TY, lz = t / (TX * TY);
  int lb;
  const Member &M = member_at<XP == 2 ? L_PASSXE : L_PASSX + DIR>(mem, map, xcd_block(blockIdx.x, gridDim.x), lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  const int nbx = pass_xtiles(P.nx, XP, TX), nby = (P.ny + TY - 1) / TY;
  const int x0 = pass_x0(P.nx, XP, TX, lb % nbx), y0 = ((lb / nbx) % nby) * TY, z0 = (lb / (nbx * nby)) * T_::TZ;
  const int i = x0 + lx, j = y0 + ly, k = z0 + lz;
  const bool active = i < (XP == 1 ? P.nx - PASS_EW : P.nx) && j < P.ny && k < P.nz;
  const long idx = ((long)k * P.ny + j) * P.nx + i;
  const int m = DIR == 0 ? i : (DIR == 1 ? j : k);
  const int n = DIR == 0 ? P.nx : (DIR == 1 ? P.ny : P.nz);
  // closure row of this point (bounded direction only)
  Sten S;
  S.mode = 0;
  if (BND) {
    if (m < 5) S.mode = 1 + m;
    else if (m > n - 6 && m < n) S.mode = 11 + (n - 1 - m);
  }
  // Bounded direction: edge tiles (XP = 2) evaluate every stencil in the
  // uniform 5-term row form (weights from an LDS table, offsets packed per
  // lane; interior points there take the interior row, the same sums).
  __shared__ R wtab[BND ? SG_NROW : 1][10];   // in the working precision
  int pk[5] = {0, 0, 0, 0, 0};
  int row = 0;
  if (BND) {
    if (t < SG_NROW * 10) (&wtab[0][0])[t] = (R)(&F.tab->roww[0][0])[t];
    row = row_of_mode(S.mode);
#pragma unroll
    for (int q = 0; q < 5; q++) pk[q] = c_row_offsets.w[row][q];
  }
  // Row offsets for one stencil.  In FP64 each stencil takes fresh copies, so
  // that its neighbour addresses are formed per stencil rather than held live
  // for the whole pass (they made the FP64 x pass spill); the FP32 pass has
  // the registers and is faster keeping them.
  auto pkc = [&](int (&pq)[5]) {
#pragma unroll
    for (int q = 0; q < 5; q++) {
      pq[q] = pk[q];
      if constexpr (sizeof(R) == 8) asm volatile("" : "+v"(pq[q]));
    }
  };
  // LDS position of this point
  const int lbase = (lz + (DIR == 2 ? 5 : 0)) * LX * LY + (ly + (DIR == 1 ? 5 : 0)) * LX + lx + (DIR == 0 ? 5 : 0);
  // staging: global source index of LDS slot l (wrap when periodic, clamp when bounded)
  auto gsrc = [&](int l) -> long {
    int tx = l % LX, ty = (l / LX) % LY, tz = l / (LX * LY);
    int gx = x0 + tx - (DIR == 0 ? 5 : 0);
    int gy = y0 + ty - (DIR == 1 ? 5 : 0);
    int gz = z0 + tz - (DIR == 2 ? 5 : 0);
    int g = DIR == 0 ? gx : (DIR == 1 ? gy : gz);
    if (BND) {
      g = g < 0 ? 0 : (g >= n ? n - 1 : g);
    } else {
      g = g < 0 ? g + n : (g >= n ? g - n : g);
    }
    if (DIR == 0) gx = g; else if (DIR == 1) gy = g; else gz = g;
    gx = gx < P.nx ? gx : P.nx - 1;
    gy = gy < P.ny ? gy : P.ny - 1;
    gz = gz < P.nz ? gz : P.nz - 1;
    return ((long)gz * P.ny + gy) * P.nx + gx;
  };
  int buf = 0;
  // Stage one field (or product) into the current buffer.  One barrier per
  // stage: the two buffers alternate, and the barrier of stage s orders the
  // reads of stage s-1 before the writes of stage s+1 into the same buffer.
  // Software pipeline: prefetch(fn) issues the global loads of the next
  // field into registers; stage() stores them into the current buffer and
  // synchronises.  Each stage's loads are thus in flight while the previous
  // stage's stencils are evaluated.
  R pf[STW][NIT];
  auto prefetch_to = [&](int f, auto fn) {
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int l = t + it * NTH;
      pf[f][it] = (R)fn(gsrc(l < LN ? l : LN - 1));
    }
  };
  auto prefetch = [&](auto fn) { prefetch_to(0, fn); };
  auto stage_n = [&](int nf) {
#pragma unroll
    for (int f = 0; f < STW; f++)
      if (f < nf) {
#pragma unroll
        for (int it = 0; it < NIT; it++) tile[buf][f][t + it * NTH] = pf[f][it];
      }
    __syncthreads();
  };
  auto stage = [&]() { stage_n(1); };
  auto wq = [&](int q) { return wtab[BND ? row : 0][q]; };
  auto D1f = [&](int f) {
    auto at = [&](int kk) { return tile[buf][f][lbase + (kk - 8) * LSTR]; };
    auto fq = [&](int q) { return tile[buf][f][lbase + q * LSTR]; };
    if (BND) { int pq[5]; pkc(pq); return d1_rows<R>(P, pq, DIR, wq, fq); }
    return d1<false, R>(P, S, DIR, at);
  };
  auto D2f = [&](int f) {
    auto at = [&](int kk) { return tile[buf][f][lbase + (kk - 8) * LSTR]; };
    auto fq = [&](int q) { return tile[buf][f][lbase + q * LSTR]; };
    if (BND) { int pq[5]; pkc(pq); return d2_rows<R>(P, pq, DIR, wq, fq); }
    return d2<false, R>(P, S, DIR, at);
  };
  auto ctrf = [&](int f) { return tile[buf][f][lbase]; };
  auto D1 = [&]() { return D1f(0); };
  auto ctr = [&]() { return ctrf(0); };
  auto next = [&]() { buf ^= 1; };
  const long fs = P.fstride;
  const double *Q = F.q;
  const R *pr = (const R *)F.prim;
  const double *mx = Q + fs, *my = Q + 2 * fs, *mz = Q + 3 * fs;
  const double *md = DIR == 0 ? mx : (DIR == 1 ? my : mz);
  const R *uu = pr + P_U * fs, *vv = pr + P_V * fs, *ww = pr + P_W * fs;
  const R *ee = pr + P_E * fs, *TT = pr + P_T * fs, *pp = pr + P_P * fs;
  const R *la = pr + P_LAM * fs, *mu = pr + P_MU * fs, *hm = pr + P_HMIX * fs;
  const R pref = pp[0];   // PRUN(IPREF,JPREF,KPREF), src/rhsvel.f:741
  // boundary-plane flags of this direction (src/bcinit.f)
  const bool atL = BND && m == 0, atR = BND && m == n - 1;
  const int fl = 2 * DIR, fr = 2 * DIR + 1;
  auto zf = [&](const int *flag) { return (atL && flag[fl]) || (atR && flag[fr]); };
  const bool zcon = zf(P.fcon), zdif = zf(P.fdif), zadb = zf(P.fadb), zvsn = zf(P.fvsn),
             zvst = zf(P.fvst);
  const bool zdifadb = zdif || zadb;
  // ADB zeroing of the correction-velocity divergence acts on any face plane
  const bool zadb_any = (i == 0 && P.fadb[0]) || (i == P.nx - 1 && P.fadb[1]) ||
                        (j == 0 && P.fadb[2]) || (j == P.ny - 1 && P.fadb[3]) ||
                        (k == 0 && P.fadb[4]) || (k == P.nz - 1 && P.fadb[5]);
  // x-face point of an NSCBC case: copies of what the NSCBC kernel reads go
  // to the face arrays (in the x pass only the edge tiles hold such points;
  // the test and the address are formed where they are used, so that they
  // hold no registers through the pass)
  constexpr bool FACES = DIR != 0 || XP == 2;
  auto facew = [&]() { return FACES && active && (P.inflow_xl || P.outflow_xr) && (i == 0 || i == P.nx - 1); };
  const long fnf = (long)P.ny * P.nz;
  auto fcw = [&](int slot) -> double & {
    return F.face[(i == 0 ? 0 : (long)nface(P.ns) * fnf) + (long)slot * fnf + (long)k * P.ny + j];
  };
  // rho u_d stays resident in its own LDS tile for the whole pass: products
  // with it (e rho u_d, Y rho u_d and the momentum fluxes rho u_d u_b) are
  // formed inside the stencils from two staged tiles instead of being staged
  // from memory, which is the same multiplication of the same two values.
  __shared__ R tmd[NIT * NTH];
  // EXT: 1/(mixture molar mass), resident for the species' pressure diffusion
  __shared__ R tow[EXT ? NIT * NTH : 1];
  auto stage_md = [&]() {
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      tmd[t + it * NTH] = pf[0][it];
      if constexpr (EXT) tow[t + it * NTH] = pf[1][it];
    }
    __syncthreads();
  };
  auto atm = [&](int kk) { return tmd[lbase + (kk - 8) * LSTR]; };
  auto fqm = [&](int q) { return tmd[lbase + q * LSTR]; };
  auto D1M = [&]() {
    if (BND) { int pq[5]; pkc(pq); return d1_rows<R>(P, pq, DIR, wq, fqm); }
    return d1<false, R>(P, S, DIR, atm);
  };
  // The product stencil reads twice the operands of a plain one; its loads
  // are kept from being hoisted over the preceding stencils so that they do
  // not hold registers alongside them.
  auto D1Pf = [&](int f) {
    auto atp = [&](int kk) {
      const int l = lbase + (kk - 8) * LSTR;
      return tile[buf][f][l] * tmd[l];
    };
    auto fqp = [&](int q) {
      const int l = lbase + q * LSTR;
      return tile[buf][f][l] * tmd[l];
    };
    SG_PHASE();
    R r;
    if (BND) { int pq[5]; pkc(pq); r = d1_rows<R>(P, pq, DIR, wq, fqp); }
    else r = d1<false, R>(P, S, DIR, atp);
    pin(r);
    SG_PHASE();
    return r;
  };
  // EXT: rho D_k Y_k (pw: times 1 - W_k / W) differentiated along this
  // direction, from the staged Y_k (slot 0) and rho D_k (slot 1) tiles and
  // the resident 1/W tile, formed at every stencil point as the reference
  // forms STORE7 (src/rhscal.f:1977, 2374)
  auto D1PY = [&](bool pw, R wk) {
    auto val = [&](int l) {
      R v = tile[buf][1][l] * tile[buf][0][l];
      if (pw) v = v * (R(1) - wk * tow[l]);
      return v;
    };
    auto atp = [&](int kk) { return val(lbase + (kk - 8) * LSTR); };
    auto fqp = [&](int q) { return val(lbase + q * LSTR); };
    SG_PHASE();
    R r;
    if (BND) { int pq[5]; pkc(pq); r = d1_rows<R>(P, pq, DIR, wq, fqp); }
    else r = d1<false, R>(P, S, DIR, atp);
    pin(r);
    SG_PHASE();
    return r;
  };
  // Momentum fluxes rho u_d u_b of the pass direction d: all three are formed
  // inside the stencils from the resident rho u_d tile and the staged u_b
  // tile.  The reference forms the fluxes rho u*v (in y), rho u*w and rho v*w
  // (in z) as rho u_b * u_d (src/rhsvel.f:348-729); rho u_d * u_b is the same
  // quantity, rounded differently in its last bit, and needs no staged
  // product (one barrier and two field loads fewer in y, two and four in z).
  R flux[3];
  R dm, md0, aD, aE, dp, p0, dmu, mu0;
  R du[3], d2u[3], u0[3];
  auto prefetch_species = [&](int sp) {
    const R *Ys = pr + prim_y(P, sp), *Rs = pr + prim_rd(P, sp), *Hs = pr + prim_h(P, sp);
    prefetch_to(0, [Ys](long g) { return Ys[g]; });
    prefetch_to(1, [Rs](long g) { return Rs[g]; });
    prefetch_to(2, [Hs](long g) { return Hs[g]; });
  };
  auto f_pp = [&](long g) { return pp[g] - pref; };
  auto f_mu = [&](long g) { return mu[g]; };
  const R *lw = pr + P_LNW * fs, *lp = pr + P_LNP * fs, *ow = pr + P_OVW * fs, *lt = pr + P_LNT * fs;
  prefetch([&](long g) { return md[g]; });
  if constexpr (EXT) prefetch_to(1, [&](long g) { return ow[g]; });
  stage_md();
  // e, T and lambda under one barrier, u, v and w under the next, then
  // p - p_ref and mu
  prefetch_to(0, [&](long g) { return ee[g]; });
  prefetch_to(1, [&](long g) { return TT[g]; });
  prefetch_to(2, [&](long g) { return la[g]; });
  dm = D1M(); md0 = tmd[lbase];
  pin(dm); pin(md0);
  aD = -dm;
  stage_n(3);
  prefetch_to(0, [&](long g) { return uu[g]; });
  prefetch_to(1, [&](long g) { return vv[g]; });
  prefetch_to(2, [&](long g) { return ww[g]; });
  // ---------------- continuity and energy convection (src/rhscal.f:141-415)
  {
    R dem = D1Pf(0);
    const R de = D1f(0);
    const R e0 = ctrf(0);
    aE = -R(0.5) * (e0 * dm + dem + md0 * de);
    pin(aE);
    if (DIR == 0 && facew()) {
      fcw(F_E) = (double)e0;
      fcw(F_T) = (double)ctrf(1);
      fcw(F_P) = (double)pp[idx];
      fcw(F_RHO) = Q[idx];
    }
  }
  SG_PHASE();
  // ---------------- heat conduction (src/rhscal.f:427-797)
  {
    R dT = D1f(1), d2T = D2f(1);
    pin(dT); pin(d2T);
    SG_PHASE();
    const R dl = D1f(2);
    const R lam0 = ctrf(2);
    aE = aE + (zcon ? R(0) : dl) * dT;
    aE = aE + lam0 * (zcon ? R(0) : d2T);
    pin(aE);
  }
  next();
  // ---------------- velocity derivatives and rho u_d u_b fluxes, all three
  // from the resident rho u_d tile and the staged u_b tiles
  stage_n(3);
  prefetch_to(0, f_pp);
  prefetch_to(1, f_mu);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    du[a] = D1f(a); d2u[a] = D2f(a); u0[a] = ctrf(a);
    pin(du[a]); pin(d2u[a]); pin(u0[a]);
    flux[a] = D1Pf(a);
    pin(flux[a]);
    SG_PHASE();
  }
  if (facew()) {
    if (DIR == 0) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        fcw(F_U + a) = (double)u0[a];
        fcw(F_DUX + a) = (double)du[a];
      }
    } else if (DIR == 1) {
      fcw(F_DUY) = (double)du[0];
      fcw(F_DVY) = (double)du[1];
    } else {
      fcw(F_DUZ) = (double)du[0];
      fcw(F_DWZ) = (double)du[2];
    }
  }
  next();
  // ---------------- p - p_ref and mu under one barrier
  stage_n(2);
  if constexpr (EXT) {
    prefetch_to(0, [&](long g) { return lw[g]; });
    prefetch_to(1, [&](long g) { return lp[g]; });
    prefetch_to(2, [&](long g) { return lt[g]; });
  } else {
    prefetch_species(0);
  }
  dp = D1f(0); p0 = ctrf(0);
  dmu = D1f(1); mu0 = ctrf(1);
  next();
  pin(dmu); pin(mu0);
  // ---------------- EXT: ln W, ln p and ln(T/TDIFGB) under one barrier
  // (src/rhscal.f:945-1004); 1/W went into its resident tile with rho u_d
  R dlw = R(0), d2lw = R(0), dlp = R(0), d2lp = R(0), dlt = R(0), d2lt = R(0);
  if constexpr (EXT) {
    stage_n(3);
    prefetch_species(0);
    dlw = D1f(0); d2lw = D2f(0);
    dlp = D1f(1); d2lp = D2f(1);
    dlt = D1f(2); d2lt = D2f(2);
    pin(dlw); pin(d2lw); pin(dlp); pin(d2lp); pin(dlt); pin(d2lt);
    next();
  }
  // ---------------- momentum convection (src/rhsvel.f:348-729, 1312-2110)
  R aM[3];
  {
    const R r0 = active ? (R)Q[idx] : R(0);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (a == DIR) aM[a] = -R(0.5) * (md0 * du[a] + flux[a] + u0[a] * dm);
      else aM[a] = -R(0.5) * (flux[a] + u0[a] * dm) - R(0.5) * r0 * u0[DIR] * du[a];
      pin(aM[a]);
    }
  }
  // ---------------- pressure gradient and pressure work (src/rhsvel.f:741-917)
  {
    p0 = p0 + pref + (R)P.pofs;   // stored pressure is p - pofs
    aM[DIR] = aM[DIR] - dp;
    aE = aE - u0[DIR] * dp - p0 * du[DIR];
    pin(aM[DIR]); pin(aE);
    if (facew()) fcw(DIR == 0 ? F_DPDX : (DIR == 1 ? F_DPDY : F_DPDZ)) = dp;
  }
  // ---------------- separable viscous terms (src/rhsvel.f:1462-2305)
#pragma unroll
  for (int a = 0; a < 3; a++) {
    R tv;
    if (a == DIR) tv = zvsn ? R(0) : mu0 * (R)P.fthd * d2u[a];
    else tv = zvst ? R(0) : mu0 * d2u[a];
    aM[a] = aM[a] + tv;
    aE = aE + tv * u0[a];
    pin(aM[a]);
  }
  pin(aE);
  if (active) ((R *)F.grad)[(long)DIR * fs + idx] = dmu;   // d mu / d x_DIR for the final kernel
  // ---------------- species (src/rhscal.f:1030-3944)
  R vt = 0.0, ucor = 0.0;
  R aY[NS], dYz[NS], yv[NS];
#pragma unroll
  for (int sp = 0; sp < NS; sp++) {
    R dY, d2Y, yc, dYm, drD, rdc, dh, hc;
    stage_n(3);
    // a light species under Soret or Dufour stages TDRMIX and T * TDRMIX next
    const bool ltsp = EXT && ((P.ltmask >> sp) & 1u);
    auto prefetch_next = [&]() {
      if (sp + 1 < NS) prefetch_species(sp + 1);
      else prefetch([&](long g) { return hm[g]; });
    };
    if (ltsp) {
      const R *Ts = pr + prim_tdr(P, sp), *TTs = pr + prim_tt(P, sp);
      prefetch_to(0, [Ts](long g) { return Ts[g]; });
      prefetch_to(1, [TTs](long g) { return TTs[g]; });
    } else {
      prefetch_next();
    }
    dY = D1f(0); d2Y = D2f(0); yc = ctrf(0);
    pin(dY); pin(d2Y); pin(yc);
    dYm = D1Pf(0);
    pin(dYm);
    SG_PHASE();
    drD = D1f(1); rdc = ctrf(1);
    pin(drD); pin(rdc);
    SG_PHASE();
    dh = D1f(2);
    hc = ctrf(2);
    R ds7w = R(0), ds7p = R(0), fw = R(0);
    if constexpr (EXT) {
      if (P.mixw) ds7w = D1PY(false, R(0));
      if (P.mixp) {
        const R wk = (R)F.tab->wmol[sp];
        fw = R(1) - wk * tow[lbase];
        ds7p = D1PY(true, wk);
      }
    }
    next();
    // EXT, light species: TDRMIX (slot 0) and T * TDRMIX (slot 1) of this
    // species under one more barrier, the Y_k and rho D_k tiles still in the
    // other buffer
    R tdr0 = R(0), ds7s = R(0), dtt = R(0), ttc = R(0);
    if constexpr (EXT) {
      if (ltsp) {
        stage_n(2);
        prefetch_next();
        const int pb = buf ^ 1;
        tdr0 = ctrf(0);
        ttc = ctrf(1);
        if (P.dufour) dtt = D1f(1);
        if (P.soret) {
          // d(rho D_k Y_k TDRMIX)/dx_d as the reference forms STORE7 (src/rhscal.f:2768)
          auto val = [&](int l) { return tile[pb][1][l] * tile[pb][0][l] * tile[buf][0][l]; };
          auto atp = [&](int kk) { return val(lbase + (kk - 8) * LSTR); };
          auto fqp = [&](int q) { return val(lbase + q * LSTR); };
          SG_PHASE();
          if (BND) { int pq[5]; pkc(pq); ds7s = d1_rows<R>(P, pq, DIR, wq, fqp); }
          else ds7s = d1<false, R>(P, S, DIR, atp);
          SG_PHASE();
          // this stencil read the other buffer, which the next stage
          // overwrites: every wave's reads end before any wave's stores
          // (the barrier of this stage ordered only the stage before)
          __syncthreads();
        }
        pin(tdr0); pin(ttc); pin(dtt); pin(ds7s);
        next();
      }
    }
    // Dufour: the species enthalpy carries R_k T TDRMIX through the heat-flux
    // terms up to the Soret terms (src/rhscal.f:1549-1566, 2832-2845)
    const bool dfsp = EXT && ltsp && P.dufour;
    const R rgk = dfsp ? (R)F.tab->rgs[sp] : R(0);
    const R hcu = dfsp ? hc + rgk * ttc : hc;
    const R dhu = dfsp ? dh + rgk * dtt : dh;
    R ys = -R(0.5) * yc * dm - R(0.5) * dYm - R(0.5) * md0 * dY;
    const R g1 = (zdif ? R(0) : drD) * dY;
    const R g2 = rdc * (zdif ? R(0) : d2Y);
    ys = ys + g1 + g2;
    vt = vt + g1 + g2;
    aE = aE + hcu * ((zdifadb ? R(0) : drD) * dY);
    aE = aE + ((zdifadb ? R(0) : dhu) * dY) * rdc;
    aE = aE + (rdc * (zdifadb ? R(0) : d2Y)) * hcu;
    ucor = ucor + rdc * dY;
    if constexpr (EXT) {
      // molar-mass (src/rhscal.f:1970-2364) and pressure (2365-2759)
      // diffusion terms of this direction.  The heat-flux term with the
      // species enthalpy gradient takes rho D_k in its x component, where the
      // reference's STORE4 holds it at that point instead of dh/dx
      // (src/rhscal.f:1887, 2071-2103, 2467-2499).
      const R s7w = rdc * yc;
      auto term = [&](R s7, R ds7, R d1l, R d2l, R ut, R dut) {
        ucor = ucor + s7 * d1l;
        const R f1 = (zdif ? R(0) : ds7) * d1l;
        ys = ys + f1;
        vt = vt + f1;
        aE = aE + ((zdifadb ? R(0) : ds7) * d1l) * ut;
        aE = aE + ((zdifadb ? R(0) : dut) * d1l) * s7;
        const R f3 = (zdif ? R(0) : d2l) * s7;
        ys = ys + f3;
        vt = vt + f3;
        aE = aE + ((zdifadb ? R(0) : d2l) * s7) * ut;
      };
      const R dhx = DIR == 0 ? rdc : dhu;
      if (P.mixw) term(s7w, ds7w, dlw, d2lw, hcu, dhx);
      if (P.mixp) term(s7w * fw, ds7p, dlp, d2lp, hcu, dhx);
      // Soret (src/rhscal.f:2757-3758): rho D_k Y_k TDRMIX with ln(T/TDIFGB);
      // its heat-flux terms take the species enthalpy without the Dufour part
      // and its gradient afresh
      if (ltsp && P.soret) term(s7w * tdr0, ds7s, dlt, d2lt, hc, dh);
      pin(ys); pin(vt); pin(ucor); pin(aE);
    }
    aY[sp] = ys;
    dYz[sp] = zdif ? R(0) : dY;
    yv[sp] = yc;
    pin(aY[sp]); pin(dYz[sp]); pin(yv[sp]); pin(vt); pin(ucor); pin(aE);
    if (DIR == 0 && BND && facew()) {
      fcw(F_NSCAL + sp) = dY;
      fcw(face_h(P.ns, sp)) = (double)hc;
      fcw(face_y(P.ns, sp)) = (double)yc;
      fcw(face_rate(P.ns, sp)) = (double)((const R *)F.rate)[sp * fs + idx];
    }
  }
  {
    stage();
    const R dhm = D1();
    const R hm0 = ctr();
    aE = aE - hm0 * (zadb_any ? R(0) : vt) - (zdifadb ? R(0) : dhm) * ucor;
  }
  if (!active) return;
  // ---------------- store
  // Every value the stores add to (the chemical rates in x, the accumulators
  // in y and z) is loaded before the first store: the stores could alias
  // those loads as far as the compiler knows, so a load after a store would
  // cost a memory round trip per variable.
  R *A = (R *)F.acc;
  R ay[NS];
#pragma unroll
  for (int sp = 0; sp < NS; sp++) ay[sp] = aY[sp] - yv[sp] * vt - dYz[sp] * ucor;
  R old[5 + NS];
  if (DIR == 0) {
#pragma unroll
    for (int sp = 0; sp < NS; sp++) old[5 + sp] = ((const R *)F.rate)[sp * fs + idx];
    A[idx] = aD;
    A[fs + idx] = aM[0];
    A[2 * fs + idx] = aM[1];
    A[3 * fs + idx] = aM[2];
    A[4 * fs + idx] = aE;
  } else {
#pragma unroll
    for (int v = 0; v < 5 + NS; v++) old[v] = A[v * fs + idx];
    A[idx] = old[0] + aD;
    A[fs + idx] = old[1] + aM[0];
    A[2 * fs + idx] = old[2] + aM[1];
    A[3 * fs + idx] = old[3] + aM[2];
    A[4 * fs + idx] = old[4] + aE;
  }
#pragma unroll
  for (int sp = 0; sp < NS; sp++) A[(5 + sp) * fs + idx] = old[5 + sp] + ay[sp];
}

// ---------------------------------------------------------------------------
// NSCBC face corrections (src/bounds.f:115-206 preparation, NSBCI3 at x-left
// 610-682, NSBCO1 at x-right 1296-1458), one thread per face point.  They are
// additive in the right-hand sides and depend only on fields the passes have
// written, so they are formed here for the 2 ny nz face points and the final
// kernel adds them.  Inside the final kernel they made every wave (each
// holds a face lane when a row is one or two waves long) run the FP64 face
// code, and raised its register count.
#define NSCBC_BLK 64   // one wave per block: the face points spread over every CU
template <int NS, class R>
__global__ void __launch_bounds__(NSCBC_BLK) k_nscbc(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                                     const sg_config *__restrict__ c) {
  int lb;
  const Member &M = member_at<L_FACE2>(mem, map, blockIdx.x, lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  const long nf = (long)P.ny * P.nz;
  const long t = (long)lb * NSCBC_BLK + threadIdx.x;
  if (t >= 2 * nf) return;
  const int side = t >= nf ? 1 : 0;
  const long fq = t - side * nf;
  if (side == 0 ? !P.inflow_xl : !P.outflow_xr) return;
  const int k = (int)(fq / P.ny), j = (int)(fq - (long)k * P.ny);
  const int i = side ? P.nx - 1 : 0;
  const long idx = ((long)k * P.ny + j) * P.nx + i;
  // the face point's values, from the copies the passes wrote into the face
  // arrays (contiguous over the face); each held in the working precision
  // exactly as the passes read it
  const double *fc = F.face + (i == 0 ? 0 : (long)nface(P.ns) * nf) + fq;
  auto fv = [&](int slot) { return (R)fc[slot * nf]; };
  const R u = fv(F_U), v = fv(F_V), w = fv(F_W);
  const R dux = fv(F_DUX), dvy = fv(F_DVY), dwz = fv(F_DWZ);
  R rD = R(0), rU = R(0), rV = R(0), rW = R(0), rE = R(0);
  R rY[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) rY[s] = R(0);
  constexpr bool BND = true;
  {
    const double strg = fc[F_CP * nf], strr = fc[F_RR * nf];
    const double strt = fv(F_T);
    const double strd = fc[F_RHO * nf];
    double gam1 = strg - strr;
    const double stre = fv(F_E) - gam1 * strt;
    gam1 = strr / gam1;
    const double ovgm = 1.0 / gam1;
    double fornow = strg * gam1 * strt;
    const double acou = sqrt(fornow), ova2 = 1.0 / fornow;
    double strh[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) strh[s] = fv(face_h(P.ns, s)) - strg * strt * c->rgspec[s] / strr;
    if (i == 0 && P.inflow_xl) {
      // NSBCI3, src/bounds.f:610-682 (imposed values have zero time derivative)
      // drho/dx at the face (src/rhscal.f:3966-3990), one-sided closure row
      const Sten Sx = make_sten<BND>(i, P.nx, 1);
      const double drdx = d1<BND>(P, Sx, 0, [&](int kk) { return F.q[idx + Sx.o[kk]]; });
      double bcl1 = dux, bcl2 = drdx, bcl3 = fv(F_DVX), bcl4 = fv(F_DWX), bcl5 = fc[F_DPDX * nf];
      fornow = strd * acou * bcl1;
      bcl1 = 0.5 * (u - acou) * (bcl5 - fornow);
      bcl2 = u * (bcl2 - bcl5 * ova2);
      bcl3 = u * bcl3;
      bcl4 = u * bcl4;
      bcl5 = 0.5 * (u + acou) * (bcl5 + fornow);
      fornow = bcl1 - strd * acou * 0.0;
      bcl2 = -0.0 - ova2 * (bcl1 + fornow) - bcl2;
      bcl3 = -0.0 - bcl3;
      bcl4 = -0.0 - bcl4;
      bcl5 = fornow - bcl5;
      rE = rE - bcl2 * stre - bcl3 * strd * v - bcl4 * strd * w - bcl5 * (ova2 * stre + u / acou + ovgm);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        double bcly = fv(face_rate(P.ns, s)) / strd - 0.0 - u * fc[(F_NSCAL + s) * nf];
        rE = rE - bcly * strd * strh[s];
      }
    }
    if (i == P.nx - 1 && P.outflow_xr) {
      // NSBCO1, src/bounds.f:1296-1458 (pointwise in/outflow switch and local-Mach beta off)
      const double bet = 0.050;
      const double strp = (double)fv(F_P) + P.pofs;
      const double t51 = -v * fc[F_DPDY * nf] - w * fc[F_DPDZ * nf];
      const double t52 = -strp * (dvy + dwz);
      const double t2b = -v * fv(F_DUY) - w * fv(F_DUZ);
      const double tt1 = t51 + t52 * (gam1 + 1.0) - strd * acou * t2b;
      double sorp = 0.0;
#pragma unroll
      for (int s = 0; s < NS; s++) sorp = sorp + strh[s] * fv(face_rate(P.ns, s));
      sorp = -sorp * gam1;
      double bcl1 = 0.5 * (u - acou) * (fc[F_DPDX * nf] - strd * acou * dux);
      bcl1 = 0.5 * sorp + P.cobcxr * acou * (strp - P.pinfxr) + 0.5 * (1.0 - bet) * tt1 - bcl1;
      rD = rD - bcl1 * ova2;
      rU = rU - bcl1 * ova2 * (u - acou);
      rV = rV - bcl1 * ova2 * v;
      rW = rW - bcl1 * ova2 * w;
      rE = rE - bcl1 * (ova2 * stre - u / acou + ovgm);
#pragma unroll
      for (int s = 0; s < NS; s++) rY[s] = rY[s] - bcl1 * ova2 * fv(face_y(P.ns, s));
    }
  }
  double *cr = F.face + (long)side * nface(P.ns) * nf + (long)face_corr(P.ns) * nf + fq;
  cr[0] = rD;
  cr[nf] = rU;
  cr[2 * nf] = rV;
  cr[3 * nf] = rW;
  cr[4 * nf] = rE;
#pragma unroll
  for (int s = 0; s < NS; s++) cr[(5 + s) * nf] = rY[s];
}

// ---------------------------------------------------------------------------
// Final kernel: cross-direction viscous terms (src/rhsvel.f:992-2185), NSCBC
// face corrections (added; formed by k_nscbc) and the RK register update
// (src/lincom.f:53-169, src/fincom.f:54-168).
//
// A block owns an FX x FY x FZ tile of points, NPT per thread along z.  The
// six mixed derivatives (d2u/dxdy, d2u/dxdz, d2v/dxdy, d2v/dydz, d2w/dxdz,
// d2w/dydz) read their diagonal neighbours from u, v and w staged one at a
// time into an LDS tile with a 5-deep halo in every direction (wrapped where
// periodic, clamped where bounded), instead of 120 neighbour loads per point.
struct FinalTile {
  static constexpr int FX = 16, FY = 8, FZ = 4;
  static constexpr int NPT = FX * FY * FZ / BLK;   // points per thread
  static constexpr int ZS = BLK / (FX * FY);        // z stride between a thread's points
  static constexpr int GX = FX + 10, GY = FY + 10, GZ = FZ + 10, GN = GX * GY * GZ;
};

// The per-point part of the final kernel at (i, j, k), given its six mixed
// derivatives.
template <bool BND, int NS, class R>
__device__ __forceinline__ void final_point(const Params &P, const Fields &F, const sg_config *__restrict__ c,
                                            int i, int j, int k, int mode, double lhs, double rhsc, double erc,
                                            double *__restrict__ out, R uxy, R vxy, R uxz, R wxz, R vyz, R wyz,
                                            R u, R v, R w, R dux, R dvy, R dwz, R sxy, R sxz, R syz) {
  const long idx = ((long)k * P.ny + j) * P.nx + i;
  const long fs = P.fstride;
  constexpr int ES = (int)sizeof(R);      // bytes per element of the FP32/FP64 fields
  const uint32_t bo = (uint32_t)idx * (uint32_t)ES, bo8 = (uint32_t)idx * 8u;
  const R *pr = (const R *)F.prim;
  const R mu = fld(pr + P_MU * fs, bo);   // u, v, w come from the staged tile
  const R *G = (const R *)F.grad;
  auto gr = [&](int q) { return fld(G + q * fs, bo); };
  const R *A = (const R *)F.acc;
  R rD = fld(A, bo), rU = fld(A + fs, bo), rV = fld(A + 2 * fs, bo), rW = fld(A + 3 * fs, bo),
    rE = fld(A + 4 * fs, bo);
  auto fz = [&](const int *fl, int dir) {
    int m = dir == 0 ? i : (dir == 1 ? j : k);
    int n = dir == 0 ? P.nx : (dir == 1 ? P.ny : P.nz);
    return (m == 0 && fl[2 * dir]) || (m == n - 1 && fl[2 * dir + 1]);
  };
  const R fthd = (R)P.fthd, tthd = (R)P.tthd;
  const R dmux = gr(0), dmuy = gr(1), dmuz = gr(2);
  // normal stresses with viscosity gradients
  {
    R sxx = fthd * dux - tthd * (dvy + dwz);
    rE = rE + mu * sxx * dux;
    R t = fz(P.fvsn, 0) ? R(0) : sxx * dmux;
    rU = rU + t;
    rE = rE + u * t;
    R syy = fthd * dvy - tthd * (dux + dwz);
    rE = rE + mu * syy * dvy;
    t = fz(P.fvsn, 1) ? R(0) : syy * dmuy;
    rV = rV + t;
    rE = rE + v * t;
    R szz = fthd * dwz - tthd * (dux + dvy);
    rE = rE + mu * szz * dwz;
    t = fz(P.fvsn, 2) ? R(0) : szz * dmuz;
    rW = rW + t;
    rE = rE + w * t;
  }
  {  // xy pair
    const R S7 = sxy;   // du/dy + dv/dx
    const R ma = uxy, mb = vxy;
    R t = fz(P.fvsn, 0) ? R(0) : tthd * mu * mb;
    rU = rU - t; rE = rE - u * t;
    t = fz(P.fvsn, 1) ? R(0) : tthd * mu * ma;
    rV = rV - t; rE = rE - v * t;
    t = fz(P.fvst, 1) ? R(0) : mu * mb + S7 * dmuy;
    rU = rU + t; rE = rE + u * t;
    t = fz(P.fvst, 0) ? R(0) : mu * ma + S7 * dmux;
    rV = rV + t; rE = rE + v * t;
    rE = rE + mu * S7 * S7;
  }
  {  // xz pair
    const R S7 = sxz;   // du/dz + dw/dx
    const R ma = uxz, mb = wxz;
    R t = fz(P.fvsn, 0) ? R(0) : tthd * mu * mb;
    rU = rU - t; rE = rE - u * t;
    t = fz(P.fvsn, 2) ? R(0) : tthd * mu * ma;
    rW = rW - t; rE = rE - w * t;
    t = fz(P.fvst, 2) ? R(0) : mu * mb + S7 * dmuz;
    rU = rU + t; rE = rE + u * t;
    t = fz(P.fvst, 0) ? R(0) : mu * ma + S7 * dmux;
    rW = rW + t; rE = rE + w * t;
    rE = rE + mu * S7 * S7;
  }
  {  // yz pair
    const R S7 = syz;   // dv/dz + dw/dy
    const R ma = vyz, mb = wyz;
    R t = fz(P.fvsn, 1) ? R(0) : tthd * mu * mb;
    rV = rV - t; rE = rE - v * t;
    t = fz(P.fvsn, 2) ? R(0) : tthd * mu * ma;
    rW = rW - t; rE = rE - w * t;
    t = fz(P.fvst, 2) ? R(0) : mu * mb + S7 * dmuz;
    rV = rV + t; rE = rE + v * t;
    t = fz(P.fvst, 1) ? R(0) : mu * ma + S7 * dmuy;
    rW = rW + t; rE = rE + w * t;
    rE = rE + mu * S7 * S7;
  }
  R rY[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) rY[s] = fld(A + (5 + s) * fs, bo);
  // ---------------- NSCBC faces: corrections formed by k_nscbc
  if (BND && (i == 0 || i == P.nx - 1)) {
    const long nf = (long)P.ny * P.nz, fq = (long)k * P.ny + j;
    const double *cr = F.face + (i == 0 ? 0 : (long)nface(P.ns) * nf) + (long)face_corr(P.ns) * nf + fq;
    rD = rD + (R)cr[0];
    rU = rU + (R)cr[nf];
    rV = rV + (R)cr[2 * nf];
    rW = rW + (R)cr[3 * nf];
    rE = rE + (R)cr[4 * nf];
#pragma unroll
    for (int s = 0; s < NS; s++) rY[s] = rY[s] + (R)cr[(5 + s) * nf];
  }
  // ---------------- register update
  R Rv[5 + NS];
  Rv[0] = rD; Rv[1] = rU; Rv[2] = rV; Rv[3] = rW; Rv[4] = rE;
#pragma unroll
  for (int s = 0; s < NS; s++) Rv[5 + s] = rY[s];
  if (mode == 2) {
#pragma unroll
    for (int vv2 = 0; vv2 < 5 + NS; vv2++) out[vv2 * fs + idx] = Rv[vv2];
    return;   // no register update
  }
  // The solution register of every variable is loaded before the first
  // store: the stores could alias the loads as far as the compiler knows, so
  // a load after a store would cost a memory round trip per variable.
  double old[5 + NS];
#pragma unroll
  for (int vv2 = 0; vv2 < 5 + NS; vv2++) old[vv2] = fld(F.run + vv2 * fs, bo8);
  if (sizeof(R) == 4 && mode == 1 && P.renorm) {
    // mixed form with renormalisation: FINCOM, then the species of the new
    // solution rescaled in FP64 so that they sum to the new density
    double nw[5 + NS];
#pragma unroll
    for (int vv2 = 0; vv2 < 5 + NS; vv2++) {
      const int slot = vv2 < 5 ? vv2 : 5;
      if (i >= P.rklo[slot]) {
        if (F.erk) fld(F.erk + vv2 * fs, bo8) += erc * (double)Rv[vv2];
        nw[vv2] = old[vv2] + lhs * (double)Rv[vv2];
      } else {
        nw[vv2] = old[vv2];
      }
    }
    if (i >= P.rklo[5]) {
      double sy = 0.0;
#pragma unroll
      for (int s = 0; s < NS; s++) sy = sy + nw[5 + s];
      const double f = nw[0] / sy;
#pragma unroll
      for (int s = 0; s < NS; s++) nw[5 + s] = nw[5 + s] * f;
    }
#pragma unroll
    for (int vv2 = 0; vv2 < 5 + NS; vv2++) {
      const int slot = vv2 < 5 ? vv2 : 5;
      if (i >= P.rklo[slot]) fld(F.run + vv2 * fs, bo8) = nw[vv2];
      fld(F.q + vv2 * fs, bo8) = i >= P.rklo[slot] ? nw[vv2] : (double)Rv[vv2];
    }
    return;
  }
#pragma unroll
  for (int vv2 = 0; vv2 < 5 + NS; vv2++) {
    const int slot = vv2 < 5 ? vv2 : 5;
    double *runv = F.run + vv2 * fs, *qv = F.q + vv2 * fs;
    if (i >= P.rklo[slot]) {
      if (F.erk) fld(F.erk + vv2 * fs, bo8) += erc * (double)Rv[vv2];   // *ERR += RKERR * R
      if (mode == 0) {
        fld(runv, bo8) = old[vv2] + lhs * (double)Rv[vv2];
        fld(qv, bo8) = old[vv2] + rhsc * (double)Rv[vv2];
      } else {
        double nw = old[vv2] + lhs * (double)Rv[vv2];
        fld(runv, bo8) = nw;
        fld(qv, bo8) = nw;
      }
    } else {
      fld(qv, bo8) = (double)Rv[vv2];
    }
  }
}

template <bool BND, int NS, class R>
__global__ void __launch_bounds__(BLK) k_final(const Member *__restrict__ mem, const uint32_t *__restrict__ map,
                                               const sg_config *__restrict__ c, int irk, double *__restrict__ out) {
  using FT = FinalTile;
  constexpr int FX = FT::FX, FY = FT::FY, FZ = FT::FZ, NPT = FT::NPT, GX = FT::GX, GY = FT::GY;
  int lb;
  const Member &M = member_at<L_FINAL>(mem, map, xcd_block(blockIdx.x, gridDim.x), lb);
  const Params P = M.P;
  const Fields F = global_fields(M.F);
  // substage irk: LINCOM (mode 0) or FINCOM at the last substage (mode 1);
  // irk = 0 writes the right-hand side to out (mode 2)
  const int mode = irk == 0 ? 2 : (irk < 5 ? 0 : 1);
  const double lhs = irk ? M.rkl[irk - 1] : 0.0, rhsc = irk ? M.rkr[irk - 1] : 0.0,
               erc = irk ? M.rke[irk - 1] : 0.0;
  // bounded x: mixed-derivative row weights (uniform row form); the field tile
  __shared__ R mtab[BND ? SG_NROW * 7 : 1];   // in the working precision
  __shared__ R tile[FT::GN];
  const int t = threadIdx.x;
  if (BND && t < SG_NROW * 7) mtab[t] = (R)(&F.tab->mixw[0][0])[t];   // ordered by the first staging barrier
  // first-derivative row weights W1 of the x closure rows (the passes' table)
  __shared__ R wtab1[BND ? SG_NROW * 5 : 1];
  if (BND && t < SG_NROW * 5) wtab1[t] = (R)F.tab->roww[t / 5][t % 5];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const int nbx = (nx + FX - 1) / FX, nby = (ny + FY - 1) / FY;
  const int x0 = (lb % nbx) * FX, y0 = ((lb / nbx) % nby) * FY, z0 = (lb / (nbx * nby)) * FZ;
  const int lx = t % FX, ly = (t / FX) % FY, lz0 = t / (FX * FY);
  const int i = x0 + lx, j = y0 + ly;
  // x closure row of this thread's points; waves holding any closure row use
  // the uniform row form, others the interior diagonal stencil
  int row = 0;
  if (BND && i < nx) {
    if (i < 5) row = 1 + i;
    else if (i > nx - 6) row = 6 + (nx - 1 - i);
  }
  const bool xedge = BND && __ballot(row != 0) != 0;
  int pk[5] = {0, 0, 0, 0, 0};
  if (xedge) {
#pragma unroll
    for (int q = 0; q < 5; q++) pk[q] = c_mix_row_offsets.w[row][q];
  }
  Sten S0;
  S0.mode = 0;
  const long fs = P.fstride;
  const R *pr = (const R *)F.prim;
  // halo index along one direction: wrapped when periodic, clamped when
  // bounded; a second clamp keeps the halo of points past the end of a
  // partial tile (whose results are not used) inside the field
  auto hidx = [&](int g, int n, bool bounded) {
    if (bounded) return g < 0 ? 0 : (g >= n ? n - 1 : g);
    g = g < 0 ? g + n : (g >= n ? g - n : g);
    return g < n ? g : n - 1;
  };
  // staging: the tile is GX*GY columns of GZ values; thread t fetches
  // columns t and t + BLK, all loads in flight at once, into registers and
  // then stores them to LDS.  The next field's fetch is issued before the
  // current field's derivatives are evaluated.  Column offsets are per
  // thread, plane offsets are uniform.
  constexpr int NCOL = GX * GY, GZ = FT::GZ;
  static_assert(NCOL > BLK && NCOL <= 2 * BLK, "two columns per thread");
  const bool col2 = t + BLK < NCOL;
  auto coloff = [&](int cc) {
    const int tx = cc % GX, ty = cc / GX;
    return (uint32_t)(hidx(y0 - 5 + ty, ny, false) * nx + hidx(x0 - 5 + tx, nx, BND));
  };
  const uint32_t co0 = coloff(t), co1 = coloff(col2 ? t + BLK : t);
  R stg[2][GZ];
  auto fetch = [&](const R *f) {
#pragma unroll
    for (int z = 0; z < GZ; z++) {
      const uint32_t zo = (uint32_t)hidx(z0 - 5 + z, nz, false) * (uint32_t)(nx * ny);
      stg[0][z] = fld(f, (zo + co0) * (uint32_t)sizeof(R));
      if (col2) stg[1][z] = fld(f, (zo + co1) * (uint32_t)sizeof(R));
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int z = 0; z < GZ; z++) {
      tile[z * NCOL + t] = stg[0][z];
      if (col2) tile[z * NCOL + t + BLK] = stg[1][z];
    }
  };
  auto base = [&](int p) { return (lz0 + p * FT::ZS + 5) * GX * GY + (ly + 5) * GX + (lx + 5); };
  auto mix_x = [&](int B, int d2) {   // d2 = 1: (x, y); d2 = 2: (x, z)
    const int s2 = d2 == 1 ? GX : GX * GY;
    auto g = [&](int a, int b) { return tile[B + a + b * s2]; };
    if (xedge) {
      // fresh copies of the row offsets for every call, so that the
      // neighbour addresses are formed per call instead of being kept live
      // for the whole kernel (they would set its occupancy)
      int pq[5];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        pq[q] = pk[q];
        asm volatile("" : "+v"(pq[q]));
      }
      return dmix_rows<R>(P, pq, 0, d2, [&](int q) { return mtab[BND ? row * 7 + q : 0]; }, g);
    }
    return dmix<false, R>(P, S0, 0, d2, [&](int k1, int k2) { return g(k1 - 8, k2 - 8); });
  };
  auto mix_yz = [&](int B) {
    return dmix<false, R>(P, S0, 1, 2, [&](int k1, int k2) { return tile[B + (k1 - 8) * GX + (k2 - 8) * GX * GY]; });
  };
  // first derivatives of the staged field: x with the closure rows where the
  // wave holds them (the passes' rows and weights), y and z interior (the
  // supported cases are periodic there)
  auto d1x = [&](int B) {
    if (xedge) {
      // a first derivative reads entries 0..9 (P1, M1) of the packed row,
      // words 0..2; all three loads are issued before the fences, which
      // would otherwise wait for each load in turn
      int pq[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 3; q++) pq[q] = c_row_offsets.w[row][q];
#pragma unroll
      for (int q = 0; q < 3; q++) asm volatile("" : "+v"(pq[q]));
      return d1_rows<R>(P, pq, 0, [&](int q) { return wtab1[BND ? row * 5 + q : 0]; },
                        [&](int q) { return tile[B + q]; });
    }
    return d1<false, R>(P, S0, 0, [&](int kk) { return tile[B + kk - 8]; });
  };
  auto d1y = [&](int B) { return d1<false, R>(P, S0, 1, [&](int kk) { return tile[B + (kk - 8) * GX]; }); };
  auto d1z = [&](int B) { return d1<false, R>(P, S0, 2, [&](int kk) { return tile[B + (kk - 8) * GX * GY]; }); };
  // the derivatives, and each point's own u, v, w read back from the tile
  // (the values a global load would give) for the per-point part.  The
  // velocity gradients the viscous terms use are formed here from the tile
  // too, instead of being written by the passes and read back: du/dx, dv/dy,
  // dw/dz and the sums du/dy + dv/dx, du/dz + dw/dx, dv/dz + dw/dy, each
  // added in that order
  R uxy[NPT], uxz[NPT], vxy[NPT], vyz[NPT], wxz[NPT], wyz[NPT], uc[NPT], vc[NPT], wc[NPT];
  R dux[NPT], dvy[NPT], dwz[NPT], sxy[NPT], sxz[NPT], syz[NPT];
  fetch(pr + P_U * fs);
  store();
  __syncthreads();
  fetch(pr + P_V * fs);
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    uc[p] = tile[base(p)];
    uxy[p] = mix_x(base(p), 1);
    uxz[p] = mix_x(base(p), 2);
    pin(uxy[p]); pin(uxz[p]);
    dux[p] = d1x(base(p));
    sxy[p] = d1y(base(p));
    sxz[p] = d1z(base(p));
    pin(dux[p]); pin(sxy[p]); pin(sxz[p]);
  }
  __syncthreads();
  store();
  __syncthreads();
  fetch(pr + P_W * fs);
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    vc[p] = tile[base(p)];
    vxy[p] = mix_x(base(p), 1);
    vyz[p] = mix_yz(base(p));
    pin(vxy[p]); pin(vyz[p]);
    sxy[p] = sxy[p] + d1x(base(p));
    dvy[p] = d1y(base(p));
    syz[p] = d1z(base(p));
    pin(sxy[p]); pin(dvy[p]); pin(syz[p]);
  }
  __syncthreads();
  store();
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    wc[p] = tile[base(p)];
    wxz[p] = mix_x(base(p), 2);
    wyz[p] = mix_yz(base(p));
    pin(wxz[p]); pin(wyz[p]);
    sxz[p] = sxz[p] + d1x(base(p));
    syz[p] = syz[p] + d1y(base(p));
    dwz[p] = d1z(base(p));
    pin(sxz[p]); pin(syz[p]); pin(dwz[p]);
  }
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    const int k = z0 + lz0 + p * FT::ZS;
    if (i < nx && j < ny && k < nz)
      final_point<BND, NS, R>(P, F, c, i, j, k, mode, lhs, rhsc, erc, out, uxy[p], vxy[p], uxz[p], wxz[p],
                              vyz[p], wyz[p], uc[p], vc[p], wc[p], dux[p], dvy[p], dwz[p], sxy[p], sxz[p], syz[p]);
  }
}

// ---------------------------------------------------------------------------
// The nine operators on one field (parity of the stencil device functions).
template <bool BX, bool BY, bool BZ>
__global__ void __launch_bounds__(BLK) k_derivs(Params P, const double *__restrict__ f, double *__restrict__ out) {
  long idx = (long)blockIdx.x * BLK + threadIdx.x;
  if (idx >= P.npts) return;
  int i, j, k;
  coords(P, idx, i, j, k);
  const Sten Sx = make_sten<BX>(i, P.nx, 1);
  const Sten Sy = make_sten<BY>(j, P.ny, (int)P.sy);
  const Sten Sz = make_sten<BZ>(k, P.nz, (int)P.sz);
  const long n = P.npts;
  out[idx] = d1<BX>(P, Sx, 0, [&](int q) { return f[idx + Sx.o[q]]; });
  out[n + idx] = d1<BY>(P, Sy, 1, [&](int q) { return f[idx + Sy.o[q]]; });
  out[2 * n + idx] = d1<BZ>(P, Sz, 2, [&](int q) { return f[idx + Sz.o[q]]; });
  out[3 * n + idx] = d2<BX>(P, Sx, 0, [&](int q) { return f[idx + Sx.o[q]]; });
  out[4 * n + idx] = d2<BY>(P, Sy, 1, [&](int q) { return f[idx + Sy.o[q]]; });
  out[5 * n + idx] = d2<BZ>(P, Sz, 2, [&](int q) { return f[idx + Sz.o[q]]; });
  out[6 * n + idx] = dmix<BX>(P, Sx, 0, 1, [&](int a, int b) { return f[idx + Sx.o[a] + Sy.o[b]]; });
  out[7 * n + idx] = dmix<BX>(P, Sx, 0, 2, [&](int a, int b) { return f[idx + Sx.o[a] + Sz.o[b]]; });
  out[8 * n + idx] = dmix<BY>(P, Sy, 1, 2, [&](int a, int b) { return f[idx + Sy.o[a] + Sz.o[b]]; });
}

// ---------------------------------------------------------------------------
// Launchers, instantiated for the species count SG_NS of this translation
// unit (the build compiles one object per supported count, in parallel).

static inline unsigned grid_for(long n) { return (unsigned)((n + BLK - 1) / BLK); }

template <int NS>
void launch_tempin_t(const Params &P, const sg_config *dcfg, const Fields &F, hipStream_t st) {
  if (P.precision) hipLaunchKernelGGL((k_tempin<NS, float>), dim3(grid_for(P.npts)), dim3(BLK), 0, st, P, dcfg, F);
  else hipLaunchKernelGGL((k_tempin<NS, double>), dim3(grid_for(P.npts)), dim3(BLK), 0, st, P, dcfg, F);
}
template <int NS>
void launch_boundt_t(const Launch &L, hipStream_t st) {
  hipLaunchKernelGGL(k_boundt<NS>, dim3(L.nblk[L_FACE]), dim3(BLK), 0, st, L.mem, L.map[L_FACE]);
}
template <int NS>
void launch_bountt_t(const Launch &L, hipStream_t st) {
  hipLaunchKernelGGL(k_bountt<NS>, dim3(L.nblk[L_FACE]), dim3(BLK), 0, st, L.mem, L.map[L_FACE]);
}
template <int NS>
static void launch_prim_k(const Launch &L, hipStream_t st) {
  const dim3 g(L.nblk[L_POINT]), b(BLK);
  if constexpr (NS == MechH2::nspec) {
    if (L.mechfix) {
      if (L.precision) hipLaunchKernelGGL((k_prim_mx<NS, true>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
      else hipLaunchKernelGGL((k_prim<NS, true>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
      return;
    }
  }
  if (L.precision) hipLaunchKernelGGL((k_prim_mx<NS, false>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
  else hipLaunchKernelGGL((k_prim<NS, false>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
}
template <int NS>
void launch_prim_t(const Launch &L, hipStream_t st) {
  launch_prim_k<NS>(L, st);
  if (L.precision) {   // mixed form: FP64 rates where FP32 cannot hold them
    const dim3 g(L.nblk[L_POINT]), b(BLK);
    bool fix = false;
    if constexpr (NS == MechH2::nspec) fix = L.mechfix != 0;
    if (fix) hipLaunchKernelGGL((k_rate64<NS, NS == MechH2::nspec>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
    else hipLaunchKernelGGL((k_rate64<NS, false>), g, b, 0, st, L.mem, L.map[L_POINT], L.c, L.tab);
  }
  if (L.ext) {   // Soret / Dufour quantities of the members that need them
    const dim3 g(L.nblk[L_POINT]), b(BLK);
    if (L.precision) hipLaunchKernelGGL((k_tdr<NS, float>), g, b, 0, st, L.mem, L.map[L_POINT], L.c);
    else hipLaunchKernelGGL((k_tdr<NS, double>), g, b, 0, st, L.mem, L.map[L_POINT], L.c);
  }
}
template <int NS, class R, int NW, bool EXT>
static void launch_pass_w(int dir, const Launch &L, hipStream_t st) {
  const int k = L_PASSX + dir;
  if (dir == 0) {
    const dim3 b(PassTile<0, NW>::NT);
    if (L.bndx) {
      // interior tiles of the bounded rows, then their edge tiles
      if (L.nblk[k] > 0)
        hipLaunchKernelGGL((k_pass<0, false, NS, R, NW, 1, EXT>), dim3(L.nblk[k]), b, 0, st, L.mem, L.map[k]);
      hipLaunchKernelGGL((k_pass<0, true, NS, R, 0, 2, EXT>), dim3(L.nblk[L_PASSXE]), dim3(PassTile<0, 0, 2>::NT), 0,
                         st, L.mem, L.map[L_PASSXE]);
    } else {
      hipLaunchKernelGGL((k_pass<0, false, NS, R, NW, 0, EXT>), dim3(L.nblk[k]), b, 0, st, L.mem, L.map[k]);
    }
  } else if (dir == 1) {
    hipLaunchKernelGGL((k_pass<1, false, NS, R, NW, 0, EXT>), dim3(L.nblk[k]), dim3(PassTile<1, NW>::NT), 0, st,
                       L.mem, L.map[k]);
  } else {
    hipLaunchKernelGGL((k_pass<2, false, NS, R, NW, 0, EXT>), dim3(L.nblk[k]), dim3(PassTile<2, NW>::NT), 0, st,
                       L.mem, L.map[k]);
  }
}
// EXT: the passes with the molar-mass and pressure diffusion terms
template <int NS, class R>
static void launch_pass_r(int dir, const Launch &L, hipStream_t st) {
  if (L.ext) {
    if (L.pass_nw[dir]) launch_pass_w<NS, R, 1, true>(dir, L, st);
    else launch_pass_w<NS, R, 0, true>(dir, L, st);
  } else {
    if (L.pass_nw[dir]) launch_pass_w<NS, R, 1, false>(dir, L, st);
    else launch_pass_w<NS, R, 0, false>(dir, L, st);
  }
}
template <int NS>
void launch_pass_t(int dir, const Launch &L, hipStream_t st) {
  if (L.precision) launch_pass_r<NS, float>(dir, L, st);
  else launch_pass_r<NS, double>(dir, L, st);
}
template <int NS>
void launch_nscbc_t(const Launch &L, hipStream_t st) {
  const dim3 g(L.nblk[L_FACE2]), b(NSCBC_BLK);
  if (L.precision) hipLaunchKernelGGL((k_nscbc<NS, float>), g, b, 0, st, L.mem, L.map[L_FACE2], L.c);
  else hipLaunchKernelGGL((k_nscbc<NS, double>), g, b, 0, st, L.mem, L.map[L_FACE2], L.c);
}
template <int NS>
void launch_radn_t(const Launch &L, hipStream_t st) {
  const dim3 g(L.nblk[L_POINT]), b(BLK);
  if (L.precision) hipLaunchKernelGGL((k_radn<NS, float>), g, b, 0, st, L.mem, L.map[L_POINT], L.c);
  else hipLaunchKernelGGL((k_radn<NS, double>), g, b, 0, st, L.mem, L.map[L_POINT], L.c);
}
template <int NS>
void launch_final_t(const Launch &L, int irk, double *out, hipStream_t st) {
  const dim3 g(L.nblk[L_FINAL]), b(BLK);
  const uint32_t *map = L.map[L_FINAL];
  if (L.precision) {
    if (L.bndx) hipLaunchKernelGGL((k_final<true, NS, float>), g, b, 0, st, L.mem, map, L.c, irk, out);
    else hipLaunchKernelGGL((k_final<false, NS, float>), g, b, 0, st, L.mem, map, L.c, irk, out);
  } else {
    if (L.bndx) hipLaunchKernelGGL((k_final<true, NS, double>), g, b, 0, st, L.mem, map, L.c, irk, out);
    else hipLaunchKernelGGL((k_final<false, NS, double>), g, b, 0, st, L.mem, map, L.c, irk, out);
  }
}

#ifdef SG_NS
template void launch_tempin_t<SG_NS>(const Params &, const sg_config *, const Fields &, hipStream_t);
template void launch_boundt_t<SG_NS>(const Launch &, hipStream_t);
template void launch_bountt_t<SG_NS>(const Launch &, hipStream_t);
template void launch_prim_t<SG_NS>(const Launch &, hipStream_t);
template void launch_pass_t<SG_NS>(int, const Launch &, hipStream_t);
template void launch_nscbc_t<SG_NS>(const Launch &, hipStream_t);
template void launch_radn_t<SG_NS>(const Launch &, hipStream_t);
template void launch_final_t<SG_NS>(const Launch &, int, double *, hipStream_t);
#endif

#ifdef SG_DERIVS
// ADAPTT error norm (src/adaptt.f:60-160): max over variables and points of
// |ERR| / (|RUN| + norm_v) with the norms of src/dtinit.f:138-145.  A NaN
// never exceeds the running maximum, as in the reference's comparisons.
__global__ void __launch_bounds__(BLK) k_errmax(Params P, Fields F) {
  __shared__ double red[BLK / 64];
  const long idx = (long)blockIdx.x * BLK + threadIdx.x;
  double tot = 0.0;
  if (idx < P.npts) {
    const long fs = P.fstride;
    for (int v = 0; v < 5 + P.ns; v++) {
      const double nrm = v == 0 ? 0.0 : (v < 4 ? 1.0e-6 : (v == 4 ? 1.0e-2 : 1.0e-10));
      const double f = fabs(F.erk[v * fs + idx]) / (fabs(F.run[v * fs + idx]) + nrm);
      if (f > tot) tot = f;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_down(tot, o, 64);
    if (u > tot) tot = u;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLK / 64; w++)
      if (red[w] > tot) tot = red[w];
    // non-negative doubles order like their bit patterns
    atomicMax(F.emax, (unsigned long long)__double_as_longlong(tot));
  }
}

__global__ void __launch_bounds__(BLK) k_zero(double *p, long n) {
  const long idx = (long)blockIdx.x * BLK + threadIdx.x;
  if (idx < n) p[idx] = 0.0;
}

void member_blocks(const Params &P, const int nw[3], int nblk[L_NKIND]) {
  nblk[L_FACE] = (int)grid_for((long)P.ny * P.nz);
  nblk[L_FACE2] = (int)((2L * P.ny * P.nz + NSCBC_BLK - 1) / NSCBC_BLK);
  nblk[L_POINT] = (int)grid_for(P.npts);
  auto tiles = [&](int tx, int ty, int tz) {
    return ((P.nx + tx - 1) / tx) * ((P.ny + ty - 1) / ty) * ((P.nz + tz - 1) / tz);
  };
  nblk[L_FINAL] = tiles(FinalTile::FX, FinalTile::FY, FinalTile::FZ);
  // x pass: every tile of a row (periodic x), or the interior tiles of the
  // bounded rows (L_PASSX) and their edge tiles (L_PASSXE)
  const int xp = P.bnd[0] ? 1 : 0;
  const int txw = nw[0] ? PassTile<0, 1>::TX : PassTile<0, 0>::TX, tyw = nw[0] ? PassTile<0, 1>::TY : PassTile<0, 0>::TY;
  nblk[L_PASSX] = pass_xtiles(P.nx, xp, txw) * ((P.ny + tyw - 1) / tyw) * P.nz;
  nblk[L_PASSXE] = P.bnd[0] ? pass_xtiles(P.nx, 2, PASS_EW) * ((P.ny + PassTile<0, 0, 2>::TY - 1) / PassTile<0, 0, 2>::TY) * P.nz : 0;
  nblk[L_PASSY] = nw[1] ? tiles(PassTile<1, 1>::TX, PassTile<1, 1>::TY, PassTile<1, 1>::TZ)
                        : tiles(PassTile<1, 0>::TX, PassTile<1, 0>::TY, PassTile<1, 0>::TZ);
  nblk[L_PASSZ] = nw[2] ? tiles(PassTile<2, 1>::TX, PassTile<2, 1>::TY, PassTile<2, 1>::TZ)
                        : tiles(PassTile<2, 0>::TX, PassTile<2, 0>::TY, PassTile<2, 0>::TZ);
}

void launch_errmax(const Params &P, const Fields &F, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(BLK), 0, st, (double *)F.emax, 1L);
  hipLaunchKernelGGL(k_errmax, dim3(grid_for(P.npts)), dim3(BLK), 0, st, P, F);
}

void launch_zero(double *p, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(BLK), 0, st, p, n);
}

void launch_derivs(const Params &P, const double *f, double *out, hipStream_t st) {
  dim3 g(grid_for(P.npts)), b(BLK);
  if (P.bnd[0]) hipLaunchKernelGGL((k_derivs<true, false, false>), g, b, 0, st, P, f, out);
  else hipLaunchKernelGGL((k_derivs<false, false, false>), g, b, 0, st, P, f, out);
}
#endif

}  // namespace sg
