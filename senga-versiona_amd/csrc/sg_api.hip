// sg_api.hip -- C ABI of the MI355X SENGA+ right-hand-side library.
//
// The handle owns the device registers of one case and one HIP stream.  The
// substage sequence mirrors src/senga2.f:113-159; see include/senga_gpu.h for
// the contract of each entry point.
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <array>
#include <string>
#include <vector>

#include "sg_internal.h"
#include "sg_stencil.h"
#include "sg_mech_h2.h"

using namespace sg;

namespace {

enum KernelId { K_TEMPIN = 0, K_BOUNDT, K_PRIM, K_PASSX, K_PASSY, K_PASSZ, K_FINAL, K_BOUNTT, K_DERIVS, K_NSCBC, K_RADN, K_COUNT };
const char *kKernelNames[K_COUNT] = {"tempin", "boundt", "prim", "pass_x", "pass_y", "pass_z",
                                     "final", "bountt", "derivs", "nscbc", "radn"};

}  // namespace

// Staging of sg_derivatives_noalloc: the input field and its nine
// derivatives, SG_NOALLOC_MAX_POINTS points each (slots rounded to 32
// elements), in one static device array.  Every GPU holds its own copy for
// the life of the process.  Host code never uses the array's name (a
// host-side stand-in): it uses the address hipGetSymbolAddress returns with
// the device current, kept per device by noalloc_scratch().
static constexpr long kNoAllocSlot = ((long)SG_NOALLOC_MAX_POINTS + 31) / 32 * 32;
static __device__ __align__(256) double scratch[10 * kNoAllocSlot];
static constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];

// The array's address on `device` (which must be current), looked up once
// per device; null if the lookup fails.
static double *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch[device] = (double *)ptr;
  }
  return g_scratch[device];
}

struct sg_batch;

struct sg_handle {
  sg_config cfg;       // host copy (species lists 0-based, REAL-rounded falloff constants)
  sg_config *dcfg;     // device copy
  Tables tab;          // derived tables (host copy)
  Params P;
  Fields F;
  hipStream_t stream;
  int device;
  double *scratch;     // this device's no-allocation staging (noalloc_scratch), or null
  void *base;          // single device allocation
  size_t bytes;
  double etime, tstep;
  double rklhs[5], rkrhs[5], rkerr[5], rktim[5];
  int adaptive;          // NSTPSW = 1: ADAPTT after each step
  double errold, errldr; // controller history
  void *ebase;           // error register allocation (adaptive only)
  size_t ebytes;
  int have_state;
  int mechfix;           // chemistry is the compiled mechanism (mech_compiled)
  int err_code;
  int err_ijk[3];
  std::string err_msg;
  // this case as the single member of its own launches
  Member mem;            // host copy of the device member
  Launch L;
  void *lbase;           // device member and block maps
  size_t lbytes;
  sg_batch *batch;       // batch this handle advances in, or null
  // profiling
  int prof_on;
  hipEvent_t ev[2];
  double prof_ms[K_COUNT];
  int64_t prof_n[K_COUNT];
};

// Derived tables (see Tables).  Transport fits are zero-padded above their
// order in the device copy of the configuration for the same reason.
static void build_tables(sg_config &c, Tables &t) {
  memset(&t, 0, sizeof t);
  for (int s = 0; s < c.nspec; s++) {
    for (int it = 0; it < c.ntint[s]; it++) {
      int np = c.ncpoly[s][it];
      for (int m = 0; m < np; m++) {
        t.cp[s][it][m] = c.amascp[s][it][m];
        t.ct[s][it][m] = c.amasct[s][it][m];
        t.ch[s][it][m] = c.amasch[s][it][m];
        t.gb[s][it][m] = c.amolgb[s][it][m];
      }
      t.hcp[s][it] = c.amascp[s][it][c.ncenth[s][it] - 1];
      t.hch[s][it] = c.amasch[s][it][c.ncenth[s][it] - 1];
      t.gbh[s][it] = c.amolgb[s][it][c.ncenth[s][it] - 1];
      t.gbs[s][it] = c.amolgb[s][it][c.ncenpy[s][it] - 1];
    }
    // an unused second interval repeats the first, so a per-point choice
    // between the two is always well defined
    for (int it = c.ntint[s]; it < SG_NTINT; it++) {
      for (int m = 0; m < SG_NCPMAX; m++) {
        t.cp[s][it][m] = t.cp[s][0][m];
        t.ct[s][it][m] = t.ct[s][0][m];
        t.ch[s][it][m] = t.ch[s][0][m];
        t.gb[s][it][m] = t.gb[s][0][m];
      }
      t.hcp[s][it] = t.hcp[s][0];
      t.hch[s][it] = t.hch[s][0];
      t.gbh[s][it] = t.gbh[s][0];
      t.gbs[s][it] = t.gbs[s][0];
    }
  }
  if (c.mixavg) {
    for (int s = 0; s < SG_MAX_SPEC; s++) {
      for (int m = c.ncovis; m < SG_NTRC; m++) c.viscco[s][m] = 0.0;
      for (int m = c.ncocon; m < SG_NTRC; m++) c.condco[s][m] = 0.0;
      for (int j = 0; j < SG_MAX_SPEC; j++)
        for (int m = c.ncodif; m < SG_NTRC; m++) c.diffco[s][j][m] = 0.0;
    }
  }
  t.ratedense = 1;
  for (int st = 0; st < c.nstep; st++) {
    bool seen[SG_MAX_SPEC] = {};
    for (int m = 0; m < c.nsslen[st]; m++) {
      const int k = c.nsspec[st][m];
      if (k < 0 || k >= c.nspec || seen[k]) {
        t.ratedense = 0;
      } else {
        seen[k] = true;
        t.nuw[st][k] = c.diffmw[st][m];
      }
    }
  }
  for (int s = 0; s < SG_MAX_SPEC; s++) {
    t.ovwf[s] = (float)c.ovwmol[s];
    t.wmol[s] = c.wmolar[s];
    t.rgs[s] = c.rgspec[s];
    t.olewf[s] = (float)c.olewis[s];
    for (int m = 0; m < SG_NTRC; m++) {
      t.condf[s][m] = (float)c.condco[s][m];
      t.viscf[s][m] = (float)c.viscco[s][m];
      for (int j = 0; j < SG_MAX_SPEC; j++) t.difff[s][j][m] = (float)c.diffco[s][j][m];
    }
    for (int j = 0; j < SG_MAX_SPEC; j++) {
      t.wilk1f[s][j] = (float)c.wilko1[s][j];
      t.wilk2f[s][j] = (float)c.wilko2[s][j];
    }
  }
  for (int b = 0; b < SG_MAX_BODY; b++)
    for (int s = 0; s < SG_MAX_SPEC; s++) t.effyf[b][s] = (float)c.effy3b[b][s];
  for (int st = 0; st < SG_MAX_STEP; st++) {
    for (int s = 0; s < SG_MAX_SPEC; s++) t.nuwf[st][s] = (float)t.nuw[st][s];
    for (int m = 0; m < SG_MAX_SSL; m++) {
      t.diffmwf[st][m] = (float)c.diffmw[st][m];
      t.crspecf[st][m] = (float)c.crspec[st][m];
      t.cpspecf[st][m] = (float)c.cpspec[st][m];
    }
  }
  t.diffsym = 1;
  for (int i = 0; i < c.nspec; i++)
    for (int j = i + 1; j < c.nspec; j++)
      for (int m = 0; m < SG_NTRC; m++)
        if (c.diffco[i][j][m] != c.diffco[j][i][m]) t.diffsym = 0;
}

static int set_err(sg_handle *h, int code, const char *msg, int i = 0, int j = 0, int k = 0) {
  if (h) {
    h->err_code = code;
    h->err_ijk[0] = i;
    h->err_ijk[1] = j;
    h->err_ijk[2] = k;
    h->err_msg = msg;
  }
  return code;
}

#define HIPCHK(h, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      char b_[256];                                                                  \
      snprintf(b_, sizeof b_, "%s: %s", #call, hipGetErrorString(e_));               \
      return set_err((h), SG_EDEVICE, b_);                                           \
    }                                                                                \
  } while (0)

// Launch a kernel through fn, timing it with events on the handle's stream
// when profiling is on.
template <class Fn>
static int timed(sg_handle *h, int kid, Fn fn) {
  if (h->prof_on) hipEventRecord(h->ev[0], h->stream);
  fn();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char b[256];
    snprintf(b, sizeof b, "launch of %s failed: %s", kKernelNames[kid], hipGetErrorString(e));
    return set_err(h, SG_EDEVICE, b);
  }
  if (h->prof_on) {
    hipEventRecord(h->ev[1], h->stream);
    hipEventSynchronize(h->ev[1]);
    float ms = 0.f;
    hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
    h->prof_ms[kid] += ms;
    h->prof_n[kid] += 1;
  }
  return 0;
}

// Stencil coefficients (src/dfinit.f:61-477), flags and ranges (src/bcinit.f).
// Whether the mixed form's FP32 rates of the mechanism stay finite and keep
// every term FP64 resolves at temperature T (src/chrate.f:177-725; the Gibbs
// functions as the point kernels form them).  FP32 holds exp(x) for x in
// about (-87, 88).  Rate coefficients that underflow to zero are harmless
// where they only multiply (the FP64 term is as negligible), so the bounds
// are: every forward and reverse exponent below 70 (room for the products
// with concentrations); for a falloff step, its forward exponent above -85
// and its low-pressure exponent in (-80, 70) (the reduced pressure divides
// by the one and the corrected coefficient carries the other into the
// reverse rate), their difference within 80; the Troe and SRI temperature
// exponentials within +-70.
static bool fp32_rates_fit(const sg_config &c, const Tables &t, double T) {
  const double lt = log(T), ovt = 1.0 / T;
  double g[SG_MAX_SPEC];
  for (int s = 0; s < c.nspec; s++) {
    int it = 0;
    while (T > c.tinthi[s][it] && it < c.ntint[s] - 1) it++;
    double f = t.gb[s][it][SG_NCPMAX - 1];
    for (int m = SG_NCPMAX - 2; m >= 0; m--) f = f * T + t.gb[s][it][m];
    g[s] = t.gbh[s][it] * ovt - t.gbs[s][it] * lt - f;
  }
  for (int st = 0; st < c.nstep; st++) {
    const double ef = c.rparam[st][0] + c.rparam[st][1] * lt - c.rparam[st][2] * ovt;
    if (!(ef < 70.0)) return false;
    auto fall = [&](const double *r) {
      const double e4 = r[0] + r[1] * lt - r[2] * ovt;
      return ef > -85.0 && e4 > -80.0 && e4 < 70.0 && fabs(e4 - ef) < 80.0;
    };
    if (c.mllist[st] && !fall(c.rclind[c.mllist[st] - 1])) return false;
    if (c.mtlist[st]) {
      const double *r = c.rctroe[c.mtlist[st] - 1];
      if (!fall(r) || !(fabs(T * r[4]) < 70.0) || !(fabs(r[5] * ovt) < 70.0) || !(fabs(T * r[6]) < 70.0))
        return false;
    }
    if (c.mslist[st]) {
      const double *r = c.rcsrif[c.mslist[st] - 1];
      if (!fall(r) || !(fabs(r[4] * ovt) < 70.0) || !(fabs(T * r[5]) < 70.0)) return false;
    }
    if (c.mglist[st]) {
      double e = ef;
      for (int q = 0; q < c.nsslen[st]; q++) e += c.diffmu[st][q] * g[c.nsspec[st][q]];
      if (!(e < 70.0)) return false;
    }
  }
  return true;
}

// The temperature range around 1000 K in which fp32_rates_fit holds, in
// steps of 1 K between 20 K and 10000 K (an empty range when it fails at
// 1000 K: every point then forms its rates in FP64).
static void fp32_rate_range(const sg_config &c, const Tables &t, double *lo, double *hi) {
  if (!fp32_rates_fit(c, t, 1000.0)) {
    *lo = 1.0;
    *hi = 0.0;
    return;
  }
  double T = 1000.0;
  while (T > 20.0 && fp32_rates_fit(c, t, T - 1.0)) T -= 1.0;
  *lo = T;
  T = 1000.0;
  while (T < 10000.0 && fp32_rates_fit(c, t, T + 1.0)) T += 1.0;
  *hi = T;
}

static int setup_params(sg_handle *h) {
  const sg_config &c = h->cfg;
  Params &P = h->P;
  memset(&P, 0, sizeof P);
  P.nx = c.nx; P.ny = c.ny; P.nz = c.nz; P.ns = c.nspec;
  P.sy = c.nx; P.sz = (long)c.nx * c.ny;
  P.npts = (long)c.nx * c.ny * c.nz;
  P.fstride = (P.npts + 31) / 32 * 32;   // 256-byte aligned fields
  const double L[3] = {c.xgdlen, c.ygdlen, c.zgdlen};
  const int N[3] = {c.nx, c.ny, c.nz};
  for (int d = 0; d < 3; d++) {
    double delta = L[d] / (double)(float)(N[d] - 1);   // src/dfinit.f:42-50
    P.ovdel[d] = 1.0 / delta;
    P.ovdl2[d] = P.ovdel[d] * P.ovdel[d];
  }
  const double acoeff = 5.0 / 3.0, bcoeff = -2.0e1 / 2.1e1, ccoeff = 5.0 / 1.4e1,
               dcoeff = -5.0 / 6.3e1, ecoeff = 1.0 / 1.26e2;
  P.ac[0] = acoeff / 2.0; P.ac[1] = bcoeff / 4.0; P.ac[2] = ccoeff / 6.0; P.ac[3] = dcoeff / 8.0;
  P.ac[4] = ecoeff / 1.0e1;
  const double acoef1 = 4.0, bcoef1 = -3.0, ccoef1 = 4.0 / 3.0, dcoef1 = -1.0 / 4.0;
  const double acoef2 = -1.0 / 4.0, bcoef2 = 3.0 / 2.0, ccoef2 = -1.0 / 2.0, dcoef2 = 1.0 / 1.2e1;
  const double acoef3 = 4.0 / 3.0, bcoef3 = -1.0 / 3.0;
  const double acoef4 = 3.0 / 2.0, bcoef4 = -3.0 / 5.0, ccoef4 = 1.0 / 1.0e1;
  const double acoef5 = 8.0 / 5.0, bcoef5 = -4.0 / 5.0, ccoef5 = 8.0 / 3.5e1, dcoef5 = -1.0 / 3.5e1;
  double c1[4] = {acoef1, bcoef1, ccoef1, dcoef1}, c2[4] = {acoef2, bcoef2, ccoef2, dcoef2};
  memcpy(P.c1, c1, sizeof c1); memcpy(P.c2, c2, sizeof c2);
  P.c3[0] = acoef3 / 2.0; P.c3[1] = bcoef3 / 4.0;
  P.c4[0] = acoef4 / 2.0; P.c4[1] = bcoef4 / 4.0; P.c4[2] = ccoef4 / 6.0;
  P.c5[0] = acoef5 / 2.0; P.c5[1] = bcoef5 / 4.0; P.c5[2] = ccoef5 / 6.0; P.c5[3] = dcoef5 / 8.0;
  P.sc[0] = acoeff; P.sc[1] = bcoeff / 4.0; P.sc[2] = ccoeff / 9.0; P.sc[3] = dcoeff / 1.6e1;
  P.sc[4] = ecoeff / 2.5e1;
  double s1[5] = {-7.7e1 / 6.0, 1.07e2 / 6.0, -1.3e1, 6.1e1 / 1.2e1, -5.0 / 6.0};
  double s2[5] = {5.0 / 6.0, -1.0 / 3.0, 7.0 / 6.0, -1.0 / 2.0, 1.0 / 1.2e1};
  memcpy(P.s1, s1, sizeof s1); memcpy(P.s2, s2, sizeof s2);
  P.s3[0] = acoef3; P.s3[1] = bcoef3 / 4.0;
  P.s4[0] = acoef4; P.s4[1] = bcoef4 / 4.0; P.s4[2] = ccoef4 / 9.0;
  P.s5[0] = acoef5; P.s5[1] = bcoef5 / 4.0; P.s5[2] = ccoef5 / 9.0; P.s5[3] = dcoef5 / 1.6e1;
  P.mc[0] = acoeff / 4.0; P.mc[1] = bcoeff / 4.0 / 4.0; P.mc[2] = ccoeff / 4.0 / 9.0;
  P.mc[3] = dcoeff / 4.0 / 1.6e1; P.mc[4] = ecoeff / 4.0 / 2.5e1;
  P.my1[0] = acoef3 / 2.0; P.my1[1] = bcoef3 / 4.0;
  memcpy(P.m1, c1, sizeof c1); memcpy(P.m2, c2, sizeof c2);
  P.m3[0] = acoef3 / 4.0; P.m3[1] = bcoef3 / 4.0 / 4.0;
  P.m4[0] = acoef4 / 4.0; P.m4[1] = bcoef4 / 4.0 / 4.0; P.m4[2] = ccoef4 / 4.0 / 9.0;
  P.m5[0] = acoef5 / 4.0; P.m5[1] = bcoef5 / 4.0 / 4.0; P.m5[2] = ccoef5 / 4.0 / 9.0;
  P.m5[3] = dcoef5 / 4.0 / 1.6e1;
  P.thrd = 1.0 / 3.0; P.tthd = 2.0 * P.thrd; P.fthd = 2.0 * P.tthd;   // src/com_senga2.h:271-272
  // boundaries: y, z periodic; x periodic or NSBCI3 (left) + NSBCO1 (right)
  for (int f = 2; f < 6; f++)
    if (c.bc_type[f] != SG_BC_PERIODIC) return set_err(h, SG_EINVAL, "y and z must be periodic");
  bool perx = c.bc_type[0] == SG_BC_PERIODIC;
  if (perx != (c.bc_type[1] == SG_BC_PERIODIC)) return set_err(h, SG_EINVAL, "x ends must both be periodic or both bounded");
  for (int v = 0; v < 6; v++) P.rklo[v] = 0;
  if (!perx) {
    if (c.bc_type[0] != SG_BC_INFLOW3 || c.bc_type[1] != SG_BC_OUTFLOW1)
      return set_err(h, SG_EINVAL, "bounded x supports NSBCI3 (13) at x-left with NSBCO1 (21) at x-right");
    if (c.bc_ipar[0][0] != 1) return set_err(h, SG_EINVAL, "NSBCI3 supports the constant-velocity option (I1=1) only");
    if (c.bc_ipar[1][0] != 0 || c.bc_ipar[1][1] != 0)
      return set_err(h, SG_EINVAL, "NSBCO1 supports I1=I2=0 only");
    P.bnd[0] = 1;
    P.inflow_xl = 1; P.outflow_xr = 1;
    P.fcnv[0] = 1; P.fvsn[0] = 1;
    P.rklo[0] = P.rklo[1] = P.rklo[2] = P.rklo[3] = P.rklo[5] = 1;
    P.fcnv[1] = 1; P.fvst[1] = 1; P.fcon[1] = 1; P.fadb[1] = 1; P.fdif[1] = 1;
    P.cobcxr = c.bc_rpar[1][1] * (1.0 - c.bc_rpar[1][2]) / (2.0 * c.xgdlen);
    P.pinfxr = c.bc_rpar[1][0];
    P.uin = c.bc_rpar[0][0];
  }
  double rtin = 0.0;   // src/indata.f:1178-1183
  for (int s = 0; s < c.nspec; s++) rtin = rtin + c.rgspec[s] * c.yrin[s];
  rtin = rtin * c.trin;
  P.drin = c.prin / rtin;
  for (int q = 0; q < c.nspec; q++) P.yrin[q] = c.yrin[q];
  P.mixavg = c.mixavg;
  // the molecular-transport options act with mixture-averaged transport only (src/diffin.f:384-386)
  P.mixw = c.mixavg && (c.physics & SG_PHYS_MIXW) ? 1 : 0;
  P.mixp = c.mixavg && (c.physics & SG_PHYS_MIXP) ? 1 : 0;
  P.radn = (c.physics & SG_PHYS_RADN) ? 1 : 0;
  P.soret = c.mixavg && (c.physics & SG_PHYS_SORET) ? 1 : 0;
  P.dufour = c.mixavg && (c.physics & SG_PHYS_DUFOUR) ? 1 : 0;
  // the thermal diffusion flags FLMTDR/FLMSOR/FLMDUF of src/diffin.f:388-407
  P.ltmask = 0;
  if (P.soret || P.dufour)
    for (int s = 0; s < c.nspec; s++)
      if (c.wmolar[s] <= c.wmltdr + 1.0e-3) P.ltmask |= 1u << s;
  P.precision = c.precision ? 1 : 0;   // 2: the mixed kernels with renormalisation
  P.renorm = c.precision == 2 ? 1 : 0;
  P.pofs = c.precision ? c.prin : 0.0;
  P.alamda = c.alamda; P.rlamda = c.rlamda; P.prantl = c.prantl;
  P.tdifgb = c.tdifgb; P.pdifgb = c.pdifgb;
  return 0;
}

// One-member launch set of a handle: its member record and all-zero block
// maps (every block belongs to member 0, blk0 = 0).
// Whether a configuration's chemistry is, bit for bit, the mechanism compiled
// into the library (MechH2, the 9-species H2-air mechanism): every chemistry
// member of sg_config as sg_create holds it, and the dense rate table derived
// from it.  Then the point kernels run the step loop unrolled over its
// constants, which evaluates the same expressions on the same values.
static int mech_compiled(const sg_config &c, const Tables &t) {
  using T = MechH2;
  if (c.nspec != T::nspec || c.nstep != T::nstep || t.ratedense != T::ratedense) return 0;
#define SAME(f)                                                                      \
  static_assert(sizeof(((const sg_config *)nullptr)->f) == sizeof(T::f), #f " size"); \
  if (memcmp(&c.f, &T::f, sizeof T::f) != 0) return 0;
  SAME(ovwmol) SAME(rparam) SAME(nsslen) SAME(nrslen) SAME(npslen) SAME(nrclen) SAME(npclen) SAME(mblist)
  SAME(mglist) SAME(mllist) SAME(mtlist) SAME(mslist) SAME(nsspec) SAME(nrspec) SAME(npspec) SAME(nrcpec)
  SAME(npcpec) SAME(diffmu) SAME(diffmw) SAME(crspec) SAME(cpspec) SAME(effy3b) SAME(rclind) SAME(rctroe)
  SAME(rcsrif) SAME(clnten)
#undef SAME
  static_assert(sizeof t.nuw == sizeof T::nuw, "nuw size");
  return memcmp(&t.nuw, &T::nuw, sizeof T::nuw) == 0;
}

static int init_launch(sg_handle *h, const int nw[3]) {
  if (h->lbase) {
    hipStreamSynchronize(h->stream);
    hipFree(h->lbase);
    h->lbase = nullptr;
  }
  Launch &L = h->L;
  memset(&L, 0, sizeof L);
  for (int d = 0; d < 3; d++) L.pass_nw[d] = nw[d] ? 1 : 0;
  member_blocks(h->P, L.pass_nw, L.nblk);
  size_t bytes = (sizeof(Member) + 255) / 256 * 256;
  for (int k = 0; k < L_NKIND; k++) bytes += ((size_t)L.nblk[k] * sizeof(uint32_t) + 255) / 256 * 256;
  if (hipMalloc(&h->lbase, bytes) != hipSuccess) {
    h->lbase = nullptr;
    return SG_ENOMEM;
  }
  // zeroed on the handle's own stream: the handle's streams do not wait for
  // the null stream, so a plain hipMemset could land after the member record
  // that refresh_member copies in next
  if (hipMemsetAsync(h->lbase, 0, bytes, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return SG_EDEVICE;
  char *b = (char *)h->lbase;
  L.mem = (const Member *)b;
  b += (sizeof(Member) + 255) / 256 * 256;
  for (int k = 0; k < L_NKIND; k++) {
    L.map[k] = (const uint32_t *)b;
    b += ((size_t)L.nblk[k] * sizeof(uint32_t) + 255) / 256 * 256;
  }
  L.nmem = 1;
  L.c = h->dcfg;
  L.tab = h->F.tab;
  L.ns = h->P.ns;
  L.precision = h->P.precision;
  L.bndx = h->P.bnd[0];
  L.ext = h->P.mixw || h->P.mixp || h->P.soret || h->P.dufour;
  L.radn = h->P.radn;
  L.inflow = h->P.inflow_xl;
  L.mechfix = h->mechfix;
  return 0;
}

static void batch_touch(sg_batch *b);

// Copy the handle's parameters, fields and RK coefficients into its device
// member (stream-ordered, so launches already queued see the old record).
static int refresh_member(sg_handle *h) {
  if (!h->lbase) return 0;
  Member &M = h->mem;
  memset(&M, 0, sizeof M);
  M.P = h->P;
  M.F = h->F;
  for (int r = 0; r < 5; r++) {
    M.rkl[r] = h->rklhs[r];
    M.rkr[r] = h->rkrhs[r];
    M.rke[r] = h->rkerr[r];
  }
  HIPCHK(h, hipMemcpyAsync((void *)h->L.mem, &M, sizeof M, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->batch) batch_touch(h->batch);
  return 0;
}

// A batch advances its member handles together: each substage kernel is one
// launch over all members (see Launch), and one RK step (BOUNDT, five
// substages, BOUNTT) is captured once as a HIP graph and replayed.  Members
// share the species count, precision, x boundary kind and mechanism.
struct sg_batch {
  std::vector<sg_handle *> members;
  hipStream_t stream;
  hipEvent_t ev;
  int device;
  void *lbase;          // device member table and block maps
  size_t lbytes;
  Launch L;
  int dirty_layout;     // membership changed: maps and graph are stale
  int dirty_members;    // a member's record changed: table is stale
  hipGraphExec_t exec;  // captured RK step, or null
};

static void batch_touch(sg_batch *b) { b->dirty_members = 1; }

// Mechanism fields the kernels read through the shared configuration (every
// sg_config member a kernel dereferences, except the per-case inflow state,
// which travels in Params).
#define SG_MECH_FIELDS(X)                                                                              \
  X(nspec) X(nstep) X(mixavg) X(amascp) X(amasct) X(clnten) X(condco) X(cpspec) X(crspec) X(diffco)   \
  X(diffmu) X(diffmw) X(effy3b) X(mblist) X(mglist) X(mllist) X(mslist) X(mtlist) X(ncenth) X(ncpoly) \
  X(npclen) X(npcpec) X(npslen) X(npspec) X(nrclen) X(nrcpec) X(nrslen) X(nrspec) X(nsslen) X(nsspec) \
  X(ntint) X(olewis) X(ovwmol) X(rclind) X(rcsrif) X(rctroe) X(rgspec) X(rparam) X(tinthi) X(viscco)  \
  X(wilko1) X(wilko2) X(ncovis) X(ncocon) X(ncodif)

// the thermal diffusion ratio fits and the radiation data, read from the
// shared configuration by the thermal-diffusion and radiation kernels
#define SG_TDR_FIELDS(X) X(ncotdr) X(wmltdr) X(tdrcco)
#define SG_RADN_FIELDS(X) X(nsprad) X(nsprid) X(nkprad) X(trefrn) X(akprad)

static bool same_mechanism(const sg_handle *a, const sg_handle *b) {
#define CMP(f) if (memcmp(&a->cfg.f, &b->cfg.f, sizeof a->cfg.f) != 0) return false;
  SG_MECH_FIELDS(CMP)
  if ((a->cfg.physics | b->cfg.physics) & (SG_PHYS_SORET | SG_PHYS_DUFOUR)) {
    SG_TDR_FIELDS(CMP)
  }
  if ((a->cfg.physics | b->cfg.physics) & SG_PHYS_RADN) {
    SG_RADN_FIELDS(CMP)
  }
#undef CMP
  // thermodynamic tables (the derivative row weights depend on the grid
  // spacing and stay per member)
  const size_t off = offsetof(Tables, cp);
  return memcmp((const char *)&a->tab + off, (const char *)&b->tab + off, sizeof(Tables) - off) == 0;
}

// Rebuild the device member table (and, when membership changed, the block
// maps).  Waits for the queued batch work, which reads the table.
static int batch_prepare(sg_batch *b) {
  if (!b->dirty_layout && !b->dirty_members) return 0;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return SG_EDEVICE;
  const int nm = (int)b->members.size();
  std::vector<Member> hm(nm);
  int tot[L_NKIND] = {};
  std::vector<std::array<int, L_NKIND>> nb(nm);
  for (int m = 0; m < nm; m++) {
    hm[m] = b->members[m]->mem;
    member_blocks(b->members[m]->P, b->L.pass_nw, nb[m].data());
    for (int k = 0; k < L_NKIND; k++) {
      hm[m].blk0[k] = tot[k];
      long t = (long)tot[k] + nb[m][k];
      if (t >= (1L << 31)) return SG_EINVAL;
      tot[k] = (int)t;
    }
  }
  const size_t mbytes = (sizeof(Member) * nm + 255) / 256 * 256;
  size_t bytes = mbytes;
  for (int k = 0; k < L_NKIND; k++) bytes += ((size_t)tot[k] * sizeof(uint32_t) + 255) / 256 * 256;
  if (b->dirty_layout) {
    if (bytes > b->lbytes) {
      if (b->lbase) hipFree(b->lbase);
      b->lbase = nullptr;
      b->lbytes = 0;
      if (hipMalloc(&b->lbase, bytes) != hipSuccess) return SG_ENOMEM;
      b->lbytes = bytes;
    }
    char *p = (char *)b->lbase;
    b->L.mem = (const Member *)p;
    p += mbytes;
    std::vector<uint32_t> map;
    for (int k = 0; k < L_NKIND; k++) {
      map.assign((size_t)tot[k], 0u);
      for (int m = 0; m < nm; m++)
        for (int q = 0; q < nb[m][k]; q++) map[(size_t)hm[m].blk0[k] + q] = (uint32_t)m;
      b->L.map[k] = (const uint32_t *)p;
      b->L.nblk[k] = tot[k];
      if (hipMemcpyAsync(p, map.data(), map.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream) !=
              hipSuccess ||
          hipStreamSynchronize(b->stream) != hipSuccess)
        return SG_EDEVICE;
      p += ((size_t)tot[k] * sizeof(uint32_t) + 255) / 256 * 256;
    }
    const sg_handle *h0 = b->members[0];
    b->L.nmem = nm;
    b->L.c = h0->dcfg;
    b->L.tab = h0->F.tab;
    b->L.ns = h0->P.ns;
    b->L.precision = h0->P.precision;
    b->L.bndx = h0->P.bnd[0];
    b->L.ext = h0->P.mixw || h0->P.mixp || h0->P.soret || h0->P.dufour;
    b->L.radn = 0;
    for (int m = 0; m < nm; m++) b->L.radn |= b->members[m]->P.radn;
    b->L.mechfix = h0->mechfix;   // members share one mechanism
    b->L.inflow = 0;
    for (int m = 0; m < nm; m++) b->L.inflow |= b->members[m]->P.inflow_xl;
    if (b->exec) hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (hipMemcpyAsync((void *)b->L.mem, hm.data(), sizeof(Member) * nm, hipMemcpyHostToDevice, b->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return SG_EDEVICE;
  b->dirty_layout = b->dirty_members = 0;
  return 0;
}

// Capture one RK step of every member: src/senga2.f:113-159 per substage,
// then BOUNTT.
static int batch_capture(sg_batch *b) {
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return SG_EDEVICE;
  for (int r = 1; r <= 5; r++) {
    if (b->L.inflow) launch_boundt(b->L, b->stream);
    launch_prim(b->L, b->stream);
    launch_pass(0, b->L, b->stream);
    if (b->L.radn) launch_radn(b->L, b->stream);
    launch_pass(1, b->L, b->stream);
    launch_pass(2, b->L, b->stream);
    if (b->L.bndx) launch_nscbc(b->L, b->stream);
    launch_final(b->L, r, nullptr, b->stream);
  }
  if (b->L.inflow) launch_bountt(b->L, b->stream);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || !g) return SG_EDEVICE;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->exec = nullptr;
    return SG_EDEVICE;
  }
  return 0;
}

// Order the handle's stream after every batch step queued so far.
static int batch_fence(sg_batch *b, sg_handle *h) {
  HIPCHK(h, hipEventRecord(b->ev, b->stream));
  HIPCHK(h, hipStreamWaitEvent(h->stream, b->ev, 0));
  return 0;
}

// Every list length and index the kernels dereference: a list of step st
// holds at most SG_MAX_SSL species, each a 0-based species index; third-body
// indices are 1-based below SG_MAX_BODY, falloff indices 1-based below
// SG_MAX_FALL (0 = none).
static bool mechanism_in_range(const sg_config *c) {
  const int ns = c->nspec;
  auto list_ok = [&](int len, const int32_t *sp) {
    if (len < 0 || len > SG_MAX_SSL) return false;
    for (int q = 0; q < len; q++)
      if (sp[q] < 0 || sp[q] >= ns) return false;
    return true;
  };
  for (int st = 0; st < c->nstep; st++) {
    if (!list_ok(c->nsslen[st], c->nsspec[st]) || !list_ok(c->nrslen[st], c->nrspec[st]) ||
        !list_ok(c->npslen[st], c->npspec[st]) || !list_ok(c->nrclen[st], c->nrcpec[st]) ||
        !list_ok(c->npclen[st], c->npcpec[st]))
      return false;
    if (c->mblist[st] < 0 || c->mblist[st] > SG_MAX_BODY) return false;
    if (c->mllist[st] < 0 || c->mllist[st] > SG_MAX_FALL || c->mtlist[st] < 0 || c->mtlist[st] > SG_MAX_FALL ||
        c->mslist[st] < 0 || c->mslist[st] > SG_MAX_FALL)
      return false;
  }
  return true;
}

static int validate_config(const sg_config *cfg) {
  if (!species_supported(cfg->nspec)) return SG_EINVAL;   // compiled species counts
  // radiation data: species in range, at most SG_NKPRAD coefficients
  if (cfg->physics & SG_PHYS_RADN) {
    if (cfg->nsprad < 0 || cfg->nsprad > cfg->nspec) return SG_EINVAL;
    for (int r = 0; r < cfg->nsprad; r++)
      if (cfg->nsprid[r] < 0 || cfg->nsprid[r] >= cfg->nspec || cfg->nkprad[r] < 1 || cfg->nkprad[r] > SG_NKPRAD)
        return SG_EINVAL;
  }
  if (cfg->mixavg && (cfg->physics & (SG_PHYS_SORET | SG_PHYS_DUFOUR)) && (cfg->ncotdr < 1 || cfg->ncotdr > SG_NTRC))
    return SG_EINVAL;
  if (cfg->nstep < 0 || cfg->nstep > SG_MAX_STEP) return SG_EINVAL;
  if (!mechanism_in_range(cfg)) return SG_EINVAL;
  if (cfg->nx < 11 || cfg->ny < 11 || cfg->nz < 11) return SG_EINVAL;
  if ((long)cfg->nx * cfg->ny * 8 >= (1L << 31)) return SG_EINVAL;
  if (cfg->precision < 0 || cfg->precision > 2) return SG_EINVAL;
  if (cfg->mixavg && (cfg->ncovis < 1 || cfg->ncocon < 1 || cfg->ncodif < 1 || cfg->ncovis > SG_NTRC ||
                      cfg->ncocon > SG_NTRC || cfg->ncodif > SG_NTRC))
    return SG_EINVAL;
  for (int s = 0; s < cfg->nspec; s++) {
    if (cfg->ntint[s] < 1 || cfg->ntint[s] > SG_NTINT) return SG_EINVAL;
    for (int it = 0; it < cfg->ntint[s]; it++)
      if (cfg->ncpoly[s][it] < 1 || cfg->ncpoly[s][it] > SG_NCPMAX || cfg->ncenth[s][it] < 1 ||
          cfg->ncenth[s][it] > SG_NCOF || cfg->ncenpy[s][it] < 1 || cfg->ncenpy[s][it] > SG_NCOF)
        return SG_EINVAL;
  }
  return 0;
}

// Device bytes of a case: the register/work block, interval indices, NSCBC
// face data, configuration and table copies, error word (sg_create's single
// allocation) plus its member record and block maps.
static size_t case_bytes(const Params &P, size_t *main_block) {
  const int ns = P.ns;
  const long fs = P.fstride;
  // FP64 registers; primitives, rates, accumulators and gradients in the
  // working precision (FP32 in the mixed mode)
  const size_t esz = P.precision ? sizeof(float) : sizeof(double);
  const long nreg = (5 + ns) * 2, nwork = nprim(ns, P.soret || P.dufour) + ns + (5 + ns) + 3;
  size_t bytes = sizeof(double) * (size_t)fs * nreg + esz * (size_t)fs * nwork;
  bytes = ((bytes + 255) / 256) * 256;
  const size_t itn = ((sizeof(uint32_t) * (size_t)fs + 255) / 256) * 256;
  const size_t facesz = sizeof(double) * 2 * (size_t)nface(ns) * P.ny * P.nz;
  const size_t cfgsz = ((sizeof(sg_config) + 255) / 256) * 256;
  const size_t tabsz = ((sizeof(Tables) + 255) / 256) * 256;
  const size_t main = bytes + itn + ((facesz + 255) / 256) * 256 + cfgsz + tabsz + 256;
  if (main_block) *main_block = main;
  // block maps sized for the larger of the wide and narrow pass tiles
  int nblk[L_NKIND], nb1[L_NKIND];
  const int wide[3] = {0, 0, 0}, narrow[3] = {1, 1, 1};
  member_blocks(P, wide, nblk);
  member_blocks(P, narrow, nb1);
  size_t lb = (sizeof(Member) + 255) / 256 * 256;
  for (int k = 0; k < L_NKIND; k++)
    lb += ((size_t)(nblk[k] > nb1[k] ? nblk[k] : nb1[k]) * sizeof(uint32_t) + 255) / 256 * 256;
  return main + lb;
}

extern "C" {

int sg_abi_version(void) { return SG_ABI_VERSION; }
size_t sg_config_size(void) { return sizeof(sg_config); }

int sg_create(const sg_config *cfg, sg_handle **out) {
  if (!cfg || !out) return SG_EINVAL;
  *out = nullptr;
  if (int rc0 = validate_config(cfg)) return rc0;
  sg_handle *h = new sg_handle();
  h->cfg = *cfg;
  build_tables(h->cfg, h->tab);
  // The reference reads these falloff constants through implicitly typed
  // REAL locals (FLCNST, TALPHA, CFCST1/2, ENCST1/2, DTCNST in
  // src/chrate.f:249,319-330); round them once here.
  for (int j = 0; j < SG_MAX_FALL; j++) {
    h->cfg.rclind[j][3] = (double)(float)h->cfg.rclind[j][3];
    const int tr[6] = {3, 7, 8, 9, 10, 11};
    for (int q = 0; q < 6; q++) h->cfg.rctroe[j][tr[q]] = (double)(float)h->cfg.rctroe[j][tr[q]];
  }
  h->mechfix = mech_compiled(h->cfg, h->tab);
  int rc = setup_params(h);
  fp32_rate_range(h->cfg, h->tab, &h->P.chlo, &h->P.chhi);
  if (!rc && h->P.fstride * (long)sizeof(double) >= (1L << 32)) rc = SG_EINVAL;  // 32-bit field offsets
  if (rc) { delete h; return rc; }
  for (int r = 0; r < SG_NROW; r++) {
    row_weights(h->P, r, h->tab.roww[r]);
    mix_row_weights(h->P, r, h->tab.mixw[r]);
  }
  if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return SG_EDEVICE;
  }
  const Params &P = h->P;
  int ns = P.ns;
  long fs = P.fstride;
  // FP64 registers; primitives, rates, accumulators and gradients in the
  // working precision (FP32 in the mixed mode)
  const size_t esz = P.precision ? sizeof(float) : sizeof(double);
  const size_t total = case_bytes(P, &h->bytes);
  h->lbytes = total - h->bytes;
  size_t bytes = sizeof(double) * (size_t)fs * (5 + ns) * 2 +
                 esz * (size_t)fs * (nprim(ns, P.soret || P.dufour) + ns + (5 + ns) + 3);
  bytes = ((bytes + 255) / 256) * 256;
  const size_t itn = ((sizeof(uint32_t) * (size_t)fs + 255) / 256) * 256;
  const size_t facesz = sizeof(double) * 2 * (size_t)nface(ns) * P.ny * P.nz;
  const size_t cfgsz = ((sizeof(sg_config) + 255) / 256) * 256;
  const size_t tabsz = ((sizeof(Tables) + 255) / 256) * 256;
  if (hipMalloc(&h->base, h->bytes) != hipSuccess) {
    hipStreamDestroy(h->stream);
    delete h;
    return SG_ENOMEM;
  }
  char *b = (char *)h->base;
  Fields &F = h->F;
  F.q = (double *)b; b += sizeof(double) * (5 + ns) * fs;
  F.run = (double *)b; b += sizeof(double) * (5 + ns) * fs;
  F.prim = (double *)b; b += esz * nprim(ns, P.soret || P.dufour) * fs;   // element type per precision
  F.rate = (double *)b; b += esz * ns * fs;
  F.acc = (double *)b; b += esz * (5 + ns) * fs;
  F.grad = (double *)b; b += esz * 3 * fs;
  b = (char *)h->base + bytes;
  F.itndex = (uint32_t *)b; b += itn;
  F.face = (double *)b; b += ((facesz + 255) / 256) * 256;
  h->dcfg = (sg_config *)b; b += cfgsz;
  F.tab = (const Tables *)b; b += tabsz;
  F.err = (int *)b;
  // all on the handle's stream, which does not wait for the null stream
  if (hipMemsetAsync(h->base, 0, h->bytes, h->stream) != hipSuccess ||
      hipMemcpyAsync(h->dcfg, &h->cfg, sizeof(sg_config), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync((void *)F.tab, &h->tab, sizeof(Tables), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    hipFree(h->base);
    hipStreamDestroy(h->stream);
    delete h;
    return SG_EDEVICE;
  }
  hipEventCreate(&h->ev[0]);
  hipEventCreate(&h->ev[1]);
  h->scratch = noalloc_scratch(h->device);   // once per device; failure only disables the no-allocation call
  h->etime = 0.0;
  const int wide[3] = {0, 0, 0};
  if (init_launch(h, wide) != 0) {
    sg_destroy(h);
    return SG_ENOMEM;
  }
  sg_set_timestep(h, 0.0);
  *out = h;
  return 0;
}

int sg_destroy(sg_handle *h) {
  if (!h) return SG_EINVAL;
  if (h->batch) sg_batch_remove(h->batch, h);
  hipStreamSynchronize(h->stream);
  if (h->lbase) hipFree(h->lbase);
  hipEventDestroy(h->ev[0]);
  hipEventDestroy(h->ev[1]);
  hipFree(h->base);
  if (h->ebase) hipFree(h->ebase);
  hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

static int check_device_error(sg_handle *h) {
  int e[4];
  HIPCHK(h, hipMemcpyAsync(e, h->F.err, sizeof e, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (e[0]) {
    const char *msg = e[0] == SG_ETEMPER ? "TEMPER: T iteration failed to converge"
                                         : "TEMPIN: temperature bisection failed";
    char b[200];
    snprintf(b, sizeof b, "%s at point %d %d %d", msg, e[1], e[2], e[3]);
    int z[4] = {0, 0, 0, 0};
    hipMemcpy(h->F.err, z, sizeof z, hipMemcpyHostToDevice);
    return set_err(h, e[0], b, e[1], e[2], e[3]);
  }
  return 0;
}

int sg_upload_state(sg_handle *h, const double *run, double etime) {
  if (!h || !run) return set_err(h, SG_EINVAL, "null argument");
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  const Params &P = h->P;
  size_t fb = sizeof(double) * P.npts;
  for (int v = 0; v < 5 + P.ns; v++) {
    HIPCHK(h, hipMemcpyAsync(h->F.run + v * P.fstride, run + v * P.npts, fb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->F.q + v * P.fstride, run + v * P.npts, fb, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipMemsetAsync(h->F.rate, 0, (P.precision ? sizeof(float) : sizeof(double)) * P.fstride * P.ns, h->stream));
  if (h->F.erk) launch_zero(h->F.erk, (long)(5 + P.ns) * P.fstride, h->stream);
  int rc = timed(h, K_TEMPIN, [&] { launch_tempin(P, h->dcfg, h->F, h->stream); });
  if (rc) return rc;
  rc = check_device_error(h);
  if (rc) return rc;
  h->etime = etime;
  h->have_state = 1;
  return 0;
}

int sg_download_state(sg_handle *h, double *run, double *temp, double *pres, double *rate) {
  if (!h) return SG_EINVAL;
  if (!h->have_state) return set_err(h, SG_ESTATE, "no state uploaded");
  if (h->batch) {
    int rc = batch_fence(h->batch, h);   // after every batch step queued so far
    if (rc) return rc;
  }
  const Params &P = h->P;
  size_t fb = sizeof(double) * P.npts;
  if (run)
    for (int v = 0; v < 5 + P.ns; v++)
      HIPCHK(h, hipMemcpyAsync(run + v * P.npts, h->F.run + v * P.fstride, fb, hipMemcpyDeviceToHost, h->stream));
  if (!P.precision) {
    if (temp) HIPCHK(h, hipMemcpyAsync(temp, h->F.prim + P_T * P.fstride, fb, hipMemcpyDeviceToHost, h->stream));
    if (pres) HIPCHK(h, hipMemcpyAsync(pres, h->F.prim + P_P * P.fstride, fb, hipMemcpyDeviceToHost, h->stream));
    if (rate)
      for (int s = 0; s < P.ns; s++)
        HIPCHK(h, hipMemcpyAsync(rate + s * P.npts, h->F.rate + s * P.fstride, fb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return check_device_error(h);
  }
  // mixed: FP32 fields, the stored pressure is p - pofs
  std::vector<float> tmp((size_t)P.npts);
  const float *prf = (const float *)h->F.prim, *rtf = (const float *)h->F.rate;
  auto fetch = [&](const float *src, double *dst, double add) -> int {
    HIPCHK(h, hipMemcpyAsync(tmp.data(), src, sizeof(float) * P.npts, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (long p = 0; p < P.npts; p++) dst[p] = (double)tmp[p] + add;
    return 0;
  };
  int rc = 0;
  if (temp && !rc) rc = fetch(prf + P_T * P.fstride, temp, 0.0);
  if (pres && !rc) rc = fetch(prf + P_P * P.fstride, pres, P.pofs);
  if (rate)
    for (int s = 0; s < P.ns && !rc; s++) rc = fetch(rtf + s * P.fstride, rate + s * P.npts, 0.0);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (rc) return rc;
  return check_device_error(h);
}

// src/dtinit.f:76-94 (RK4(3)5[2R+]C, Kennedy-Carpenter) and 196-223
int sg_set_timestep(sg_handle *h, double dt) {
  if (!h) return SG_EINVAL;
  const double acofrk[5] = {970286171893.0 / 4311952581923.0, 6584761158862.0 / 12103376702013.0,
                            2251764453980.0 / 15575788980749.0, 26877169314380.0 / 34165994151039.0, 0.0};
  const double bcofrk[5] = {1153189308089.0 / 22510343858157.0, 1772645290293.0 / 4653164025191.0,
                            -1672844663538.0 / 4480602732383.0, 2114624349019.0 / 3568978502595.0,
                            5198255086312.0 / 14908931495163.0};
  const double bhatrk[5] = {1016888040809.0 / 7410784769900.0, 11231460423587.0 / 58533540763752.0,
                            -1563879915014.0 / 6823010717585.0, 606302364029.0 / 971179775848.0,
                            1097981568119.0 / 3980877426909.0};
  h->tstep = dt;
  for (int r = 0; r < 5; r++) {
    h->rklhs[r] = bcofrk[r] * dt;
    h->rkrhs[r] = acofrk[r] * dt;
    h->rkerr[r] = (bcofrk[r] - bhatrk[r]) * dt;
  }
  h->rktim[0] = 0.0;
  for (int r = 1; r < 5; r++) {
    double t = acofrk[r - 1];
    for (int q = 0; q < r - 1; q++) t = t + bcofrk[q];
    h->rktim[r] = t * dt;
  }
  return refresh_member(h);
}

// BOUNDT .. BOUNDS (src/senga2.f:118-131): everything up to the register update.
static int eval_rhs(sg_handle *h) {
  const Launch &L = h->L;
  int rc = 0;
  if (L.inflow) rc = timed(h, K_BOUNDT, [&] { launch_boundt(L, h->stream); });
  if (!rc) rc = timed(h, K_PRIM, [&] { launch_prim(L, h->stream); });
  if (!rc) rc = timed(h, K_PASSX, [&] { launch_pass(0, L, h->stream); });
  if (!rc && L.radn) rc = timed(h, K_RADN, [&] { launch_radn(L, h->stream); });
  if (!rc) rc = timed(h, K_PASSY, [&] { launch_pass(1, L, h->stream); });
  if (!rc) rc = timed(h, K_PASSZ, [&] { launch_pass(2, L, h->stream); });
  if (!rc && L.bndx) rc = timed(h, K_NSCBC, [&] { launch_nscbc(L, h->stream); });
  return rc;
}

int sg_substage(sg_handle *h, int irk) {
  if (!h) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  if (!h->have_state) return set_err(h, SG_ESTATE, "no state uploaded");
  if (irk < 1 || irk > 5) return set_err(h, SG_EINVAL, "irk must be 1..5");
  int rc = eval_rhs(h);
  if (rc) return rc;
  return timed(h, K_FINAL, [&] { launch_final(h->L, irk, nullptr, h->stream); });
}

// ADAPTT (src/adaptt.f:60-330) with the constants of src/dtinit.f:101-170:
// global error norm on the device, PID controller on the host, RK
// coefficients rescaled by the step ratio, error register cleared.
static int adaptt(sg_handle *h) {
  const double errtol = 1.0e-4, errlow = 1.0e-30, trmax = 1.01, trmin = 1.0e-2, tsmax = 1.0,
               tsmin = 1.0e-15;
  const double ctmult = 9.0e-1, ctalph = 4.9e-1 / 3.0, ctbeta = 3.4e-1 / 3.0, ctgama = 1.0e-1 / 3.0;
  launch_errmax(h->P, h->F, h->stream);
  unsigned long long bits = 0;
  HIPCHK(h, hipMemcpyAsync(&bits, h->F.emax, sizeof bits, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double errmax;
  memcpy(&errmax, &bits, sizeof errmax);
  if (errmax < errlow) errmax = errlow;
  double tratio = ctmult * exp(ctalph * log(errtol / errmax) + ctbeta * log(h->errold / errtol) +
                               ctgama * log(errtol / h->errldr));
  h->errldr = h->errold;
  h->errold = errmax;
  if (tratio > trmax) tratio = trmax;
  if (tratio < trmin) tratio = trmin;
  const double tstold = h->tstep;
  h->tstep = h->tstep * tratio;
  if (h->tstep > tsmax) h->tstep = tsmax;
  if (h->tstep < tsmin) h->tstep = tsmin;
  tratio = h->tstep / tstold;
  for (int r = 0; r < 5; r++) {
    h->rklhs[r] = h->rklhs[r] * tratio;
    h->rkrhs[r] = h->rkrhs[r] * tratio;
    h->rkerr[r] = h->rkerr[r] * tratio;
    h->rktim[r] = h->rktim[r] * tratio;
  }
  launch_zero(h->F.erk, (long)(5 + h->P.ns) * h->P.fstride, h->stream);
  int rc = refresh_member(h);
  if (rc) return rc;
  return check_device_error(h);
}

int sg_end_step(sg_handle *h) {
  if (!h) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  h->etime = h->etime + h->tstep;
  int rc = 0;
  if (h->L.inflow) rc = timed(h, K_BOUNTT, [&] { launch_bountt(h->L, h->stream); });
  if (!rc && h->adaptive) rc = adaptt(h);
  return rc;
}

int sg_set_adaptive(sg_handle *h, int on, double errold, double errldr) {
  if (!h || (on && !(errold > 0.0 && errldr > 0.0))) return SG_EINVAL;
  if (on && h->batch) return set_err(h, SG_ESTATE, "batched handles use a fixed time step");
  if (on && !h->ebase) {
    const size_t fb = sizeof(double) * (size_t)(5 + h->P.ns) * h->P.fstride;
    h->ebytes = fb + 256;
    if (hipMalloc(&h->ebase, h->ebytes) != hipSuccess) {
      h->ebase = nullptr;
      return set_err(h, SG_ENOMEM, "error register allocation failed");
    }
    HIPCHK(h, hipMemsetAsync(h->ebase, 0, h->ebytes, h->stream));
  }
  h->adaptive = on ? 1 : 0;
  h->F.erk = on ? (double *)h->ebase : nullptr;
  h->F.emax = on ? (unsigned long long *)((char *)h->ebase + (h->ebytes - 256)) : nullptr;
  if (on) {
    h->errold = errold;
    h->errldr = errldr;
    launch_zero(h->F.erk, (long)(5 + h->P.ns) * h->P.fstride, h->stream);
  }
  return refresh_member(h);
}

int sg_get_timestep(sg_handle *h, double *tstep, double *errold, double *errldr) {
  if (!h) return SG_EINVAL;
  if (tstep) *tstep = h->tstep;
  if (errold) *errold = h->errold;
  if (errldr) *errldr = h->errldr;
  return 0;
}

int sg_step(sg_handle *h, int nsteps) {
  if (!h) return SG_EINVAL;
  for (int n = 0; n < nsteps; n++) {
    for (int r = 1; r <= 5; r++) {
      int rc = sg_substage(h, r);
      if (rc) return rc;
    }
    int rc = sg_end_step(h);
    if (rc) return rc;
  }
  return 0;
}

int sg_rhs(sg_handle *h, int irk, double *rhs) {
  if (!h || !rhs) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  if (!h->have_state) return set_err(h, SG_ESTATE, "no state uploaded");
  (void)irk;
  const Params &P = h->P;
  size_t qbytes = sizeof(double) * P.fstride * (5 + P.ns);
  double *save = nullptr, *out = nullptr;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMalloc((void **)&save, qbytes));
  HIPCHK(h, hipMalloc((void **)&out, qbytes));
  HIPCHK(h, hipMemcpyAsync(save, h->F.q, qbytes, hipMemcpyDeviceToDevice, h->stream));
  int rc = eval_rhs(h);
  if (!rc) rc = timed(h, K_FINAL, [&] { launch_final(h->L, 0, out, h->stream); });
  if (!rc) {
    size_t fb = sizeof(double) * P.npts;
    for (int v = 0; v < 5 + P.ns; v++)
      HIPCHK(h, hipMemcpyAsync(rhs + v * P.npts, out + v * P.fstride, fb, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipMemcpyAsync(h->F.q, save, qbytes, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(save));
  HIPCHK(h, hipFree(out));
  if (rc) return rc;
  return check_device_error(h);
}

int sg_derivatives(sg_handle *h, const double *f, double *out) {
  if (!h || !f || !out) return SG_EINVAL;
  const Params &P = h->P;
  size_t fb = sizeof(double) * P.npts;
  double *df = nullptr, *dout = nullptr;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMalloc((void **)&df, fb));
  HIPCHK(h, hipMalloc((void **)&dout, 9 * fb));
  HIPCHK(h, hipMemcpyAsync(df, f, fb, hipMemcpyHostToDevice, h->stream));
  int rc = timed(h, K_DERIVS, [&] { launch_derivs(P, df, dout, h->stream); });
  if (!rc) HIPCHK(h, hipMemcpyAsync(out, dout, 9 * fb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(df));
  HIPCHK(h, hipFree(dout));
  return rc;
}

int sg_derivatives_noalloc(sg_handle *h, const double *f, double *out) {
  if (!h || !f || !out) return SG_EINVAL;
  const Params &P = h->P;
  if (P.npts > SG_NOALLOC_MAX_POINTS)
    return set_err(h, SG_EINVAL, "grid exceeds SG_NOALLOC_MAX_POINTS points of the no-allocation staging");
  if (!h->scratch) return set_err(h, SG_EDEVICE, "no-allocation staging unavailable on this device");
  size_t fb = sizeof(double) * P.npts;
  double *df = h->scratch, *dout = h->scratch + kNoAllocSlot;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpyAsync(df, f, fb, hipMemcpyHostToDevice, h->stream));
  int rc = timed(h, K_DERIVS, [&] { launch_derivs(P, df, dout, h->stream); });
  if (!rc) HIPCHK(h, hipMemcpyAsync(out, dout, 9 * fb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return rc;
}

int sg_sync(sg_handle *h) {
  if (!h) return SG_EINVAL;
  if (h->batch) {
    int rc = batch_fence(h->batch, h);
    if (rc) return rc;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return check_device_error(h);
}

int sg_last_error(sg_handle *h, int *code, int ijk[3], char *msg, size_t n) {
  if (!h) return SG_EINVAL;
  if (code) *code = h->err_code;
  if (ijk) for (int q = 0; q < 3; q++) ijk[q] = h->err_ijk[q];
  if (msg && n) {
    strncpy(msg, h->err_msg.c_str(), n - 1);
    msg[n - 1] = 0;
  }
  return 0;
}

int sg_get_stream(sg_handle *h, void **stream) {
  if (!h || !stream) return SG_EINVAL;
  *stream = (void *)h->stream;
  return 0;
}

int sg_profile(sg_handle *h, int on) {
  if (!h) return SG_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (on) {
    for (int q = 0; q < K_COUNT; q++) { h->prof_ms[q] = 0.0; h->prof_n[q] = 0; }
  }
  h->prof_on = on ? 1 : 0;
  return 0;
}

int sg_profile_read(sg_handle *h, int idx, char *name, size_t n, double *total_ms, int64_t *count) {
  if (!h || idx < 0 || idx >= K_COUNT) return SG_EINVAL;
  if (name && n) {
    strncpy(name, kKernelNames[idx], n - 1);
    name[n - 1] = 0;
  }
  if (total_ms) *total_ms = h->prof_ms[idx];
  if (count) *count = h->prof_n[idx];
  return 0;
}

int sg_footprint(const sg_config *cfg, size_t *bytes) {
  if (!cfg || !bytes) return SG_EINVAL;
  *bytes = 0;
  if (int rc = validate_config(cfg)) return rc;
  sg_handle *h = new sg_handle();
  h->cfg = *cfg;
  int rc = setup_params(h);
  if (!rc && h->P.fstride * (long)sizeof(double) >= (1L << 32)) rc = SG_EINVAL;
  if (!rc) *bytes = case_bytes(h->P, nullptr);
  delete h;
  return rc;
}

int sg_memory(sg_handle *h, size_t *bytes) {
  if (!h || !bytes) return SG_EINVAL;
  *bytes = h->bytes + h->lbytes + (h->ebase ? h->ebytes : 0);
  return 0;
}

int sg_batch_create(sg_batch **out) {
  if (!out) return SG_EINVAL;
  *out = nullptr;
  sg_batch *b = new sg_batch();
  b->lbase = nullptr;
  b->lbytes = 0;
  b->exec = nullptr;
  b->dirty_layout = b->dirty_members = 0;
  memset(&b->L, 0, sizeof b->L);
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return SG_EDEVICE;
  }
  if (hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess) {
    hipStreamDestroy(b->stream);
    delete b;
    return SG_EDEVICE;
  }
  *out = b;
  return 0;
}

int sg_batch_destroy(sg_batch *b) {
  if (!b) return SG_EINVAL;
  hipStreamSynchronize(b->stream);
  for (sg_handle *h : b->members) h->batch = nullptr;
  b->members.clear();
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->lbase) hipFree(b->lbase);
  hipEventDestroy(b->ev);
  hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int sg_batch_add(sg_batch *b, sg_handle *h) {
  if (!b || !h) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle already advances in a batch");
  if (!h->have_state) return set_err(h, SG_ESTATE, "no state uploaded");
  if (h->adaptive) return set_err(h, SG_EINVAL, "batched handles use a fixed time step");
  if (h->device != b->device) return set_err(h, SG_EINVAL, "handle and batch are on different devices");
  if (b->members.size() >= 65535) return set_err(h, SG_EINVAL, "batch is full");
  if (!b->members.empty()) {
    const sg_handle *h0 = b->members[0];
    if (h->P.ns != h0->P.ns || h->P.precision != h0->P.precision || h->P.bnd[0] != h0->P.bnd[0] ||
        (h->P.mixw || h->P.mixp || h->P.soret || h->P.dufour) !=
            (h0->P.mixw || h0->P.mixp || h0->P.soret || h0->P.dufour))
      return set_err(h, SG_EINVAL,
                     "batch members share species count, precision, x boundary kind and the transport terms");
    if (!same_mechanism(h, h0)) return set_err(h, SG_EINVAL, "batch members share one mechanism");
  }
  // the batch stream runs after the handle's queued work (upload, TEMPIN)
  HIPCHK(h, hipEventRecord(b->ev, h->stream));
  HIPCHK(h, hipStreamWaitEvent(b->stream, b->ev, 0));
  b->members.push_back(h);
  h->batch = b;
  b->dirty_layout = 1;
  return 0;
}

int sg_batch_remove(sg_batch *b, sg_handle *h) {
  if (!b || !h || h->batch != b) return SG_EINVAL;
  // the handle's later work runs after the batch steps queued so far
  int rc = batch_fence(b, h);
  for (size_t m = 0; m < b->members.size(); m++)
    if (b->members[m] == h) {
      b->members.erase(b->members.begin() + (long)m);
      break;
    }
  h->batch = nullptr;
  b->dirty_layout = 1;
  return rc;
}

int sg_batch_size(sg_batch *b, int *n) {
  if (!b || !n) return SG_EINVAL;
  *n = (int)b->members.size();
  return 0;
}

int sg_batch_step(sg_batch *b, int nsteps) {
  if (!b || nsteps < 0) return SG_EINVAL;
  if (b->members.empty() || nsteps == 0) return 0;
  int rc = batch_prepare(b);
  if (rc) return rc;
  if (!b->exec) {
    rc = batch_capture(b);
    if (rc) return rc;
  }
  for (int n = 0; n < nsteps; n++)
    if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) return SG_EDEVICE;
  for (sg_handle *h : b->members)
    for (int n = 0; n < nsteps; n++) h->etime = h->etime + h->tstep;
  return 0;
}

int sg_batch_sync(sg_batch *b, int *failed) {
  if (!b) return SG_EINVAL;
  if (failed) *failed = -1;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return SG_EDEVICE;
  for (size_t m = 0; m < b->members.size(); m++) {
    int rc = check_device_error(b->members[m]);
    if (rc) {
      if (failed) *failed = (int)m;
      return rc;
    }
  }
  return 0;
}

int sg_batch_member(sg_batch *b, int idx, sg_handle **h) {
  if (!b || !h || idx < 0 || idx >= (int)b->members.size()) return SG_EINVAL;
  *h = b->members[(size_t)idx];
  return 0;
}

}  // extern "C"

// ---- Tile tuning ------------------------------------------------------------
// A tuning run times launches on the live state.  It changes nothing a later
// step reads except the stored temperature (the Newton initial guess) and the
// interval indices, which prim rewrites; those are saved and put back, so a
// tuned run computes exactly what an untuned one does.
struct TempSnap {
  void *buf = nullptr;
  size_t tb = 0, ib = 0;
};

static int snap_save(sg_handle *h, TempSnap &sn) {
  const Params &P = h->P;
  const size_t esz = P.precision ? sizeof(float) : sizeof(double);
  sn.tb = esz * (size_t)P.fstride;
  sn.ib = sizeof(uint32_t) * (size_t)P.fstride;
  HIPCHK(h, hipMalloc(&sn.buf, sn.tb + sn.ib));
  const char *pt = (const char *)h->F.prim + esz * (size_t)P_T * P.fstride;
  HIPCHK(h, hipMemcpyAsync(sn.buf, pt, sn.tb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((char *)sn.buf + sn.tb, h->F.itndex, sn.ib, hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

static int snap_restore(sg_handle *h, TempSnap &sn, hipStream_t st) {
  const Params &P = h->P;
  const size_t esz = P.precision ? sizeof(float) : sizeof(double);
  char *pt = (char *)h->F.prim + esz * (size_t)P_T * P.fstride;
  HIPCHK(h, hipStreamSynchronize(st));
  HIPCHK(h, hipMemcpyAsync(pt, sn.buf, sn.tb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->F.itndex, (char *)sn.buf + sn.tb, sn.ib, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipFree(sn.buf));
  sn.buf = nullptr;
  return 0;
}

// Mean duration in ms of reps launches made by fn on stream st (after one
// untimed launch).
template <class Fn>
static double time_launches(hipStream_t st, int reps, Fn fn) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  fn();
  hipEventRecord(e0, st);
  for (int r = 0; r < reps; r++) fn();
  hipEventRecord(e1, st);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return (double)ms / reps;
}

extern "C" {

int sg_set_tiles(sg_handle *h, const int narrow[3]) {
  if (!h || !narrow) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  if (init_launch(h, narrow) != 0) return set_err(h, SG_ENOMEM, "launch table allocation failed");
  return refresh_member(h);
}

int sg_mechanism_compiled(const sg_handle *h, int *compiled) {
  if (!h || !compiled) return SG_EINVAL;
  *compiled = h->mechfix;
  return 0;
}

int sg_get_tiles(sg_handle *h, int narrow[3]) {
  if (!h || !narrow) return SG_EINVAL;
  for (int d = 0; d < 3; d++) narrow[d] = h->L.pass_nw[d];
  return 0;
}

int sg_autotune(sg_handle *h, int reps, double *ms) {
  if (!h) return SG_EINVAL;
  if (h->batch) return set_err(h, SG_ESTATE, "handle advances in a batch");
  if (!h->have_state) return set_err(h, SG_ESTATE, "no state uploaded");
  if (reps <= 0) reps = 5;
  TempSnap sn;
  int rc = snap_save(h, sn);
  if (rc) return rc;
  if (h->L.inflow) launch_boundt(h->L, h->stream);
  launch_prim(h->L, h->stream);
  int nw[3] = {h->L.pass_nw[0], h->L.pass_nw[1], h->L.pass_nw[2]};
  double total = 0.0;
  for (int d = 0; d < 3 && !rc; d++) {
    double tb = 1e30;
    int bw = nw[d];
    for (int w = 0; w < 2 && !rc; w++) {
      int cur[3] = {nw[0], nw[1], nw[2]};
      cur[d] = w;
      if (init_launch(h, cur) != 0) rc = set_err(h, SG_ENOMEM, "launch table allocation failed");
      if (!rc) rc = refresh_member(h);
      if (rc) break;
      const double t = time_launches(h->stream, reps, [&] { launch_pass(d, h->L, h->stream); });
      if (t < tb) { tb = t; bw = w; }
    }
    nw[d] = bw;
    total += tb;
  }
  if (!rc && init_launch(h, nw) != 0) rc = set_err(h, SG_ENOMEM, "launch table allocation failed");
  if (!rc) rc = refresh_member(h);
  int rc2 = snap_restore(h, sn, h->stream);
  if (!rc) rc = rc2;
  if (!rc) rc = refresh_member(h);
  if (!rc && ms) *ms = total;
  return rc;
}

int sg_batch_set_tiles(sg_batch *b, const int narrow[3]) {
  if (!b || !narrow) return SG_EINVAL;
  for (int d = 0; d < 3; d++) b->L.pass_nw[d] = narrow[d] ? 1 : 0;
  b->dirty_layout = 1;
  return 0;
}

int sg_batch_get_tiles(sg_batch *b, int narrow[3]) {
  if (!b || !narrow) return SG_EINVAL;
  for (int d = 0; d < 3; d++) narrow[d] = b->L.pass_nw[d];
  return 0;
}

int sg_batch_autotune(sg_batch *b, int reps, double *ms) {
  if (!b) return SG_EINVAL;
  if (b->members.empty()) return 0;
  if (reps <= 0) reps = 5;
  b->dirty_layout = 1;
  int rc = batch_prepare(b);
  if (rc) return rc;
  std::vector<TempSnap> sn(b->members.size());
  for (size_t m = 0; m < b->members.size() && !rc; m++) rc = snap_save(b->members[m], sn[m]);
  for (sg_handle *h : b->members)
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = SG_EDEVICE;
  if (!rc) {
    if (b->L.inflow) launch_boundt(b->L, b->stream);
    launch_prim(b->L, b->stream);
  }
  double total = 0.0;
  for (int d = 0; d < 3 && !rc; d++) {
    double tb = 1e30;
    int bw = b->L.pass_nw[d];
    for (int w = 0; w < 2 && !rc; w++) {
      b->L.pass_nw[d] = w;
      b->dirty_layout = 1;
      rc = batch_prepare(b);
      if (rc) break;
      const double t = time_launches(b->stream, reps, [&] { launch_pass(d, b->L, b->stream); });
      if (t < tb) { tb = t; bw = w; }
    }
    b->L.pass_nw[d] = bw;
    total += tb;
  }
  b->dirty_layout = 1;
  int rc2 = batch_prepare(b);
  if (!rc) rc = rc2;
  for (size_t m = 0; m < b->members.size(); m++)
    if (sn[m].buf) {
      rc2 = snap_restore(b->members[m], sn[m], b->stream);
      if (!rc) rc = rc2;
    }
  if (!rc && ms) *ms = total;
  return rc;
}

}  // extern "C"
