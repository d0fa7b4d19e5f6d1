/*
 * senga_oracle.c -- CPU restatement of the SENGA+ RK-substage right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY (see senga_oracle.h).  Restated from the reference
 * Fortran in the same evaluation order; every routine cites the reference source lines
 * it follows.  Storage mirrors the reference: "big" arrays carry a 5-deep
 * halo (src/com_senga2.h:47-50, indices -5..n+4 here, 0-based), interior
 * arrays do not.  Periodic halos are filled by wrap-around copies exactly as
 * the single-rank PARFER does (src/parfer.f:832-876).
 *
 * Supported configurations: y and z periodic; x periodic, or NSBCI3 inflow
 * at x-left with NSBCO1 outflow at x-right (src/bounds.f:610-682,1296-1458).
 * Constant-Lewis or mixture-averaged transport, the latter with any of the
 * molar-mass diffusion, pressure diffusion, Soret and Dufour terms; optically
 * thin radiation with either.
 */
#include "senga_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALO 5
#define NBOUND 1
#define NPERI 2

/* numbers as src/com_senga2.h:268-274 */
/* Every loop nest over points below computes each point from inputs that
   the nest does not write (no sums across points), so its outer loop runs
   in parallel when built with OpenMP and the results do not depend on the
   thread count. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

static const double ZERO = 0.0, ONE = 1.0, TWO = 2.0, THREE = 3.0, HALF = 0.5;

struct orc {
  sg_config c;
  int nx, ny, nz, ns;
  int bx, by, bz;
  size_t nb, ni;
  int nend[6];            /* xl,xr,yl,yr,zl,zr: NBOUND or NPERI */
  int fcnv[6], fvsn[6], fvst[6], fcon[6], fadb[6], fdif[6];
  int tlo[3], thi[3];     /* TEMPER ranges ISTALT..ISTOLT (0-based) */
  int blo[3], bhi[3];     /* ISTAB..ISTOB restricted to defined points */
  int rklo[6];            /* x start of RK update for D,U,V,W,E,Y */
  int rkhi[6];
  double delta[3], ovdel[3], ovdl2[3];
  /* stencil coefficients, src/dfinit.f:61-477 */
  double ac[5];                    /* ACOFFX..ECOFFX */
  double c1[4], c2[4], c3[2], c4[3], c5[4];
  double sc[5];                    /* ACOFSX..ECOFSX */
  double s1[5], s2[5], s3[2], s4[3], s5[4];
  double mc[5];                    /* ACOFXY..ECOFXY */
  double my1[2];                   /* ACOFY1,BCOFY1 */
  double m1[4], m2[4], m3[2], m4[3], m5[4];
  double thrd, tthd, fthd;
  double drin, cobcxr, pinfxr;
  double rklhs[5], rkrhs[5], rkerr[5], rktim[5];
  double etime, tstep;
  int adaptive;            /* NSTPSW = 1: ADAPTT after each step */
  double errold, errldr;   /* controller history (ERROLD, ERRLDR) */
  /* registers */
  double *run;   /* interior, 5+ns fields */
  double *err;   /* interior, 5+ns fields */
  double *rhs;   /* big, 5+ns fields */
  /* workspace */
  double *trun, *prun, *transp, *store7, *utmp, *vtmp, *wtmp, *wmomix, *difmix;
  uint32_t *itndex;   /* ITNDEX: NINTMX = 2 words per point (see itint_get) */
  double *st1, *st2, *st3, *st4, *st5, *st6, *divm, *ucor, *vcor, *wcor;
  /* first and second derivatives of ln(mixture molar mass) and ln(pressure)
     (WD1X..WD2Z, PD1X..PD2Z), allocated when NFMIXW / NFMIXP are on */
  double *wd1[3], *wd2[3], *pd1[3], *pd2[3];
  int mixw, mixp;
  /* Soret and Dufour: first and second derivatives of ln(T/TDIFGB) (TD1X..
     TD2Z), the thermal diffusion ratio of the current species (TDRMIX, big)
     and the light species they act for (FLMTDR, bit per species) */
  double *td1[3], *td2[3], *tdrmix;
  int soret, dufour;
  int radn;                /* optically thin radiation (FLRADN) */
  unsigned ltmask;
  double *rate, *rrte;
  /* face arrays (ny*nz, and *ns for species) */
  double *fxl, *fxr;
  int have_state;
};

/* face array slots */
enum { F_STRG, F_STRR, F_T1B, F_STRE, F_STRT, F_BCLT, F_STRD, F_BCL2, F_STRU, F_STRV, F_STRW,
       F_T2B, F_T3B, F_T4B, F_BCL1, F_STRP, F_BCL5, F_T51B, F_T52B, F_BCL3, F_BCL4, F_NSCAL };
enum { FS_RATE, FS_STRY, FS_BCLY, FS_T6B, FS_STRH, F_NSPEC };

#define BI(o, i, j, k) ((((size_t)((k) + HALO)) * (o)->by + ((j) + HALO)) * (o)->bx + ((i) + HALO))
#define II(o, i, j, k) ((((size_t)(k)) * (o)->ny + (j)) * (o)->nx + (i))

static double *FS(orc_t *o, double *face, int slot) { return face + (size_t)slot * o->ny * o->nz; }
static double *FSP(orc_t *o, double *face, int slot, int isp) {
  return face + ((size_t)F_NSCAL + (size_t)slot * o->ns + isp) * o->ny * o->nz;
}
static double *R(orc_t *o, int v) { return o->rhs + (size_t)v * o->nb; }
static double *RUN(orc_t *o, int v) { return o->run + (size_t)v * o->ni; }
static double *ERR(orc_t *o, int v) { return o->err + (size_t)v * o->ni; }

/* ------------------------------------------------------------------------ */
/* set-up: src/dfinit.f:42-477 coefficients, src/bcinit.f flags and ranges   */

static void dfinit(orc_t *o) {
  const double acoeff = 5.0 / 3.0, bcoeff = -2.0e1 / 2.1e1, ccoeff = 5.0 / 1.4e1,
               dcoeff = -5.0 / 6.3e1, ecoeff = 1.0 / 1.26e2;
  double L[3] = {o->c.xgdlen, o->c.ygdlen, o->c.zgdlen};
  int N[3] = {o->nx, o->ny, o->nz};
  for (int d = 0; d < 3; d++) {
    /* DELTAX = XGDLEN/REAL(NXGLBL-1): REAL of a small integer is exact */
    o->delta[d] = L[d] / (double)(float)(N[d] - 1);
    o->ovdel[d] = ONE / o->delta[d];
    o->ovdl2[d] = o->ovdel[d] * o->ovdel[d];
  }
  o->ac[0] = acoeff / 2.0; o->ac[1] = bcoeff / 4.0; o->ac[2] = ccoeff / 6.0;
  o->ac[3] = dcoeff / 8.0; o->ac[4] = ecoeff / 1.0e1;
  const double acoef1 = 4.0, bcoef1 = -3.0, ccoef1 = 4.0 / 3.0, dcoef1 = -1.0 / 4.0;
  o->c1[0] = acoef1; o->c1[1] = bcoef1; o->c1[2] = ccoef1; o->c1[3] = dcoef1;
  const double acoef2 = -1.0 / 4.0, bcoef2 = 3.0 / 2.0, ccoef2 = -1.0 / 2.0, dcoef2 = 1.0 / 1.2e1;
  o->c2[0] = acoef2; o->c2[1] = bcoef2; o->c2[2] = ccoef2; o->c2[3] = dcoef2;
  const double acoef3 = 4.0 / 3.0, bcoef3 = -1.0 / 3.0;
  o->c3[0] = acoef3 / 2.0; o->c3[1] = bcoef3 / 4.0;
  const double acoef4 = 3.0 / 2.0, bcoef4 = -3.0 / 5.0, ccoef4 = 1.0 / 1.0e1;
  o->c4[0] = acoef4 / 2.0; o->c4[1] = bcoef4 / 4.0; o->c4[2] = ccoef4 / 6.0;
  const double acoef5 = 8.0 / 5.0, bcoef5 = -4.0 / 5.0, ccoef5 = 8.0 / 3.5e1, dcoef5 = -1.0 / 3.5e1;
  o->c5[0] = acoef5 / 2.0; o->c5[1] = bcoef5 / 4.0; o->c5[2] = ccoef5 / 6.0; o->c5[3] = dcoef5 / 8.0;
  /* second derivatives */
  o->sc[0] = acoeff; o->sc[1] = bcoeff / 4.0; o->sc[2] = ccoeff / 9.0; o->sc[3] = dcoeff / 1.6e1;
  o->sc[4] = ecoeff / 2.5e1;
  o->s1[0] = -7.7e1 / 6.0; o->s1[1] = 1.07e2 / 6.0; o->s1[2] = -1.3e1; o->s1[3] = 6.1e1 / 1.2e1;
  o->s1[4] = -5.0 / 6.0;
  o->s2[0] = 5.0 / 6.0; o->s2[1] = -1.0 / 3.0; o->s2[2] = 7.0 / 6.0; o->s2[3] = -1.0 / 2.0;
  o->s2[4] = 1.0 / 1.2e1;
  o->s3[0] = acoef3; o->s3[1] = bcoef3 / 4.0;
  o->s4[0] = acoef4; o->s4[1] = bcoef4 / 4.0; o->s4[2] = ccoef4 / 9.0;
  o->s5[0] = acoef5; o->s5[1] = bcoef5 / 4.0; o->s5[2] = ccoef5 / 9.0; o->s5[3] = dcoef5 / 1.6e1;
  /* mixed */
  o->mc[0] = acoeff / 4.0; o->mc[1] = bcoeff / 4.0 / 4.0; o->mc[2] = ccoeff / 4.0 / 9.0;
  o->mc[3] = dcoeff / 4.0 / 1.6e1; o->mc[4] = ecoeff / 4.0 / 2.5e1;
  o->my1[0] = acoef3 / 2.0; o->my1[1] = bcoef3 / 4.0;
  o->m1[0] = acoef1; o->m1[1] = bcoef1; o->m1[2] = ccoef1; o->m1[3] = dcoef1;
  o->m2[0] = acoef2; o->m2[1] = bcoef2; o->m2[2] = ccoef2; o->m2[3] = dcoef2;
  o->m3[0] = acoef3 / 4.0; o->m3[1] = bcoef3 / 4.0 / 4.0;
  o->m4[0] = acoef4 / 4.0; o->m4[1] = bcoef4 / 4.0 / 4.0; o->m4[2] = ccoef4 / 4.0 / 9.0;
  o->m5[0] = acoef5 / 4.0; o->m5[1] = bcoef5 / 4.0 / 4.0; o->m5[2] = ccoef5 / 4.0 / 9.0;
  o->m5[3] = dcoef5 / 4.0 / 1.6e1;
  /* src/com_senga2.h:271-272 */
  o->thrd = ONE / THREE; o->tthd = TWO * o->thrd; o->fthd = TWO * o->tthd;
}

static int bcinit(orc_t *o) {
  const sg_config *c = &o->c;
  for (int f = 0; f < 6; f++) {
    o->fcnv[f] = o->fvsn[f] = o->fvst[f] = o->fcon[f] = o->fadb[f] = o->fdif[f] = 0;
  }
  /* y, z periodic only */
  for (int f = 2; f < 6; f++)
    if (c->bc_type[f] != SG_BC_PERIODIC) return SG_EINVAL;
  int perx = (c->bc_type[0] == SG_BC_PERIODIC);
  if (perx != (c->bc_type[1] == SG_BC_PERIODIC)) return SG_EINVAL;
  for (int f = 0; f < 6; f++) o->nend[f] = NPERI;
  int N[3] = {o->nx, o->ny, o->nz};
  for (int d = 0; d < 3; d++) {
    o->tlo[d] = -HALO; o->thi[d] = N[d] - 1 + HALO;
    o->blo[d] = -HALO; o->bhi[d] = N[d] - 1 + HALO;
  }
  for (int v = 0; v < 6; v++) { o->rklo[v] = 0; o->rkhi[v] = o->nx - 1; }
  if (!perx) {
    if (c->bc_type[0] != SG_BC_INFLOW3 || c->bc_type[1] != SG_BC_OUTFLOW1) return SG_EINVAL;
    if (c->bc_ipar[0][0] != 1) return SG_EINVAL;         /* constant inflow velocity only */
    if (c->bc_ipar[1][0] != 0 || c->bc_ipar[1][1] != 0) return SG_EINVAL;
    o->nend[0] = o->nend[1] = NBOUND;
    o->blo[0] = 0; o->bhi[0] = o->nx - 1;
    /* src/bcinit.f: NSBCI3 x-left */
    o->fcnv[0] = 1; o->fvsn[0] = 1;
    o->rklo[0] = o->rklo[1] = o->rklo[2] = o->rklo[3] = o->rklo[5] = 1;
    o->tlo[0] = 0;
    /* NSBCO1 x-right */
    o->fcnv[1] = 1; o->fvst[1] = 1; o->fcon[1] = 1; o->fadb[1] = 1; o->fdif[1] = 1;
    o->thi[0] = o->nx - 1;
    o->cobcxr = c->bc_rpar[1][1] * (ONE - c->bc_rpar[1][2]) / (TWO * c->xgdlen);
    o->pinfxr = c->bc_rpar[1][0];
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* periodic wrap of the stage register, src/parfer.f:832-876 (single rank)  */

static void wrap(orc_t *o, double *f) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  if (o->nend[0] == NPERI)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int h = 1; h <= HALO; h++) {
          f[BI(o, -h, j, k)] = f[BI(o, nx - h, j, k)];
          f[BI(o, nx - 1 + h, j, k)] = f[BI(o, h - 1, j, k)];
        }
  for (int k = 0; k < nz; k++)
    for (int h = 1; h <= HALO; h++)
      for (int i = -HALO; i < nx + HALO; i++) {
        f[BI(o, i, -h, k)] = f[BI(o, i, ny - h, k)];
        f[BI(o, i, ny - 1 + h, k)] = f[BI(o, i, h - 1, k)];
      }
  for (int h = 1; h <= HALO; h++)
    for (int j = -HALO; j < ny + HALO; j++)
      for (int i = -HALO; i < nx + HALO; i++) {
        f[BI(o, i, j, -h)] = f[BI(o, i, j, nz - h)];
        f[BI(o, i, j, nz - 1 + h)] = f[BI(o, i, j, h - 1)];
      }
}

/* ------------------------------------------------------------------------ */
/* derivative operators                                                     */

/* first derivative along dir, src/dfbydx.f:51-242 (y, z twins identical) */
static void dfbyd(orc_t *o, int dir, const double *f, double *out) {
  int N[3] = {o->nx, o->ny, o->nz};
  long st[3] = {1, o->bx, (long)o->bx * o->by};
  long s = st[dir];
  int n = N[dir], L = n - 1;
  int bl = o->nend[2 * dir] == NBOUND, br = o->nend[2 * dir + 1] == NBOUND;
  const double *a = o->ac;
  OMP_FOR
  for (int k = 0; k < o->nz; k++)
    for (int j = 0; j < o->ny; j++)
      for (int i = 0; i < o->nx; i++) {
        int m = dir == 0 ? i : (dir == 1 ? j : k);
        const double *p = f + BI(o, i, j, k);
        double r;
#define F(q) p[(long)(q) * s]
        if (bl && m < 5) {
          if (m == 0) {
            double fa = F(1) - F(0), fb = F(2) - F(0), fc = F(3) - F(0), fd = F(4) - F(0);
            r = o->c1[0] * fa + o->c1[1] * fb + o->c1[2] * fc + o->c1[3] * fd;
          } else if (m == 1) {
            double fa = F(-1) - F(0), fb = F(1) - F(0), fc = F(2) - F(0), fd = F(3) - F(0);
            r = o->c2[0] * fa + o->c2[1] * fb + o->c2[2] * fc + o->c2[3] * fd;
          } else if (m == 2) {
            double fa = F(1) - F(-1), fb = F(2) - F(-2);
            r = o->c3[0] * fa + o->c3[1] * fb;
          } else if (m == 3) {
            double fa = F(1) - F(-1), fb = F(2) - F(-2), fc = F(3) - F(-3);
            r = o->c4[0] * fa + o->c4[1] * fb + o->c4[2] * fc;
          } else {
            double fa = F(1) - F(-1), fb = F(2) - F(-2), fc = F(3) - F(-3), fd = F(4) - F(-4);
            r = o->c5[0] * fa + o->c5[1] * fb + o->c5[2] * fc + o->c5[3] * fd;
          }
        } else if (br && m > L - 5) {
          int q = L - m;
          if (q == 4) {
            double fa = F(1) - F(-1), fb = F(2) - F(-2), fc = F(3) - F(-3), fd = F(4) - F(-4);
            r = o->c5[0] * fa + o->c5[1] * fb + o->c5[2] * fc + o->c5[3] * fd;
          } else if (q == 3) {
            double fa = F(1) - F(-1), fb = F(2) - F(-2), fc = F(3) - F(-3);
            r = o->c4[0] * fa + o->c4[1] * fb + o->c4[2] * fc;
          } else if (q == 2) {
            double fa = F(1) - F(-1), fb = F(2) - F(-2);
            r = o->c3[0] * fa + o->c3[1] * fb;
          } else if (q == 1) {
            double fa = F(0) - F(1), fb = F(0) - F(-1), fc = F(0) - F(-2), fd = F(0) - F(-3);
            r = o->c2[0] * fa + o->c2[1] * fb + o->c2[2] * fc + o->c2[3] * fd;
          } else {
            double fa = F(0) - F(-1), fb = F(0) - F(-2), fc = F(0) - F(-3), fd = F(0) - F(-4);
            r = o->c1[0] * fa + o->c1[1] * fb + o->c1[2] * fc + o->c1[3] * fd;
          }
        } else {
          double fa = F(1) - F(-1), fb = F(2) - F(-2), fc = F(3) - F(-3), fd = F(4) - F(-4),
                 fe = F(5) - F(-5);
          r = a[0] * fa + a[1] * fb + a[2] * fc + a[3] * fd + a[4] * fe;
        }
#undef F
        out[II(o, i, j, k)] = r * o->ovdel[dir];
      }
}

/* second derivative along dir, src/d2fdx2.f:52-274 */
static void d2fd2(orc_t *o, int dir, const double *f, double *out) {
  int N[3] = {o->nx, o->ny, o->nz};
  long st[3] = {1, o->bx, (long)o->bx * o->by};
  long s = st[dir];
  int n = N[dir], L = n - 1;
  int bl = o->nend[2 * dir] == NBOUND, br = o->nend[2 * dir + 1] == NBOUND;
  const double *a = o->sc;
  OMP_FOR
  for (int k = 0; k < o->nz; k++)
    for (int j = 0; j < o->ny; j++)
      for (int i = 0; i < o->nx; i++) {
        int m = dir == 0 ? i : (dir == 1 ? j : k);
        const double *p = f + BI(o, i, j, k);
        double r;
#define F(q) p[(long)(q) * s]
        if (bl && m < 5) {
          if (m == 0) {
            double ap = F(1) - F(0), bp = F(2) - F(0), cp = F(3) - F(0), dp = F(4) - F(0),
                   ep = F(5) - F(0);
            r = o->s1[0] * ap + o->s1[1] * bp + o->s1[2] * cp + o->s1[3] * dp + o->s1[4] * ep;
          } else if (m == 1) {
            double ap = F(-1) - F(0), bp = F(1) - F(0), cp = F(2) - F(0), dp = F(3) - F(0),
                   ep = F(4) - F(0);
            r = o->s2[0] * ap + o->s2[1] * bp + o->s2[2] * cp + o->s2[3] * dp + o->s2[4] * ep;
          } else if (m == 2) {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2);
            r = o->s3[0] * (ap - am) + o->s3[1] * (bp - bm);
          } else if (m == 3) {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2),
                   cp = F(3) - F(0), cm = F(0) - F(-3);
            r = o->s4[0] * (ap - am) + o->s4[1] * (bp - bm) + o->s4[2] * (cp - cm);
          } else {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2),
                   cp = F(3) - F(0), cm = F(0) - F(-3), dp = F(4) - F(0), dm = F(0) - F(-4);
            r = o->s5[0] * (ap - am) + o->s5[1] * (bp - bm) + o->s5[2] * (cp - cm) +
                o->s5[3] * (dp - dm);
          }
        } else if (br && m > L - 5) {
          int q = L - m;
          if (q == 4) {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2),
                   cp = F(3) - F(0), cm = F(0) - F(-3), dp = F(4) - F(0), dm = F(0) - F(-4);
            r = o->s5[0] * (ap - am) + o->s5[1] * (bp - bm) + o->s5[2] * (cp - cm) +
                o->s5[3] * (dp - dm);
          } else if (q == 3) {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2),
                   cp = F(3) - F(0), cm = F(0) - F(-3);
            r = o->s4[0] * (ap - am) + o->s4[1] * (bp - bm) + o->s4[2] * (cp - cm);
          } else if (q == 2) {
            double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2);
            r = o->s3[0] * (ap - am) + o->s3[1] * (bp - bm);
          } else if (q == 1) {
            double ap = F(1) - F(0), bp = F(-1) - F(0), cp = F(-2) - F(0), dp = F(-3) - F(0),
                   ep = F(-4) - F(0);
            r = o->s2[0] * ap + o->s2[1] * bp + o->s2[2] * cp + o->s2[3] * dp + o->s2[4] * ep;
          } else {
            double ap = F(-1) - F(0), bp = F(-2) - F(0), cp = F(-3) - F(0), dp = F(-4) - F(0),
                   ep = F(-5) - F(0);
            r = o->s1[0] * ap + o->s1[1] * bp + o->s1[2] * cp + o->s1[3] * dp + o->s1[4] * ep;
          }
        } else {
          double ap = F(1) - F(0), am = F(0) - F(-1), bp = F(2) - F(0), bm = F(0) - F(-2),
                 cp = F(3) - F(0), cm = F(0) - F(-3), dp = F(4) - F(0), dm = F(0) - F(-4),
                 ep = F(5) - F(0), em = F(0) - F(-5);
          r = a[0] * (ap - am) + a[1] * (bp - bm) + a[2] * (cp - cm) + a[3] * (dp - dm) +
              a[4] * (ep - em);
        }
#undef F
        out[II(o, i, j, k)] = r * o->ovdl2[dir];
      }
}

/* mixed derivative in the (d1,d2) plane, src/d2fdxy.f:51-391 (interior and
 * d1-direction end closures; d2 must be periodic).  d1 < d2. */
static void d2fdmix(orc_t *o, int d1, int d2, const double *f, double *out) {
  int N[3] = {o->nx, o->ny, o->nz};
  long st[3] = {1, o->bx, (long)o->bx * o->by};
  long s1 = st[d1], s2 = st[d2];
  int L = N[d1] - 1;
  int bl = o->nend[2 * d1] == NBOUND, br = o->nend[2 * d1 + 1] == NBOUND;
  const double *a = o->mc;
  OMP_FOR
  for (int k = 0; k < o->nz; k++)
    for (int j = 0; j < o->ny; j++)
      for (int i = 0; i < o->nx; i++) {
        int m = d1 == 0 ? i : (d1 == 1 ? j : k);
        const double *p = f + BI(o, i, j, k);
        double r;
#define F(a1, a2) p[(long)(a1) * s1 + (long)(a2) * s2]
#define DIAG(q, b) (F(q, b) - F(q, -(b)) - F(-(q), b) + F(-(q), -(b)))
/* 4th-order centred transverse difference of f(a1,.) - f(b1,.) */
#define TR(a1, b1)                                                               \
  (o->my1[0] * (F(a1, 1) - F(a1, -1) - F(b1, 1) + F(b1, -1)) +                    \
   o->my1[1] * (F(a1, 2) - F(a1, -2) - F(b1, 2) + F(b1, -2)))
        if (bl && m < 5) {
          if (m == 0) {
            double fa = TR(1, 0), fb = TR(2, 0), fc = TR(3, 0), fd = TR(4, 0);
            r = o->m1[0] * fa + o->m1[1] * fb + o->m1[2] * fc + o->m1[3] * fd;
          } else if (m == 1) {
            double fa = TR(-1, 0), fb = TR(1, 0), fc = TR(2, 0), fd = TR(3, 0);
            r = o->m2[0] * fa + o->m2[1] * fb + o->m2[2] * fc + o->m2[3] * fd;
          } else if (m == 2) {
            double fa = DIAG(1, 1), fb = DIAG(2, 2);
            r = o->m3[0] * fa + o->m3[1] * fb;
          } else if (m == 3) {
            double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3);
            r = o->m4[0] * fa + o->m4[1] * fb + o->m4[2] * fc;
          } else {
            double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4);
            r = o->m5[0] * fa + o->m5[1] * fb + o->m5[2] * fc + o->m5[3] * fd;
          }
        } else if (br && m > L - 5) {
          int q = L - m;
          if (q == 4) {
            double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4);
            r = o->m5[0] * fa + o->m5[1] * fb + o->m5[2] * fc + o->m5[3] * fd;
          } else if (q == 3) {
            double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3);
            r = o->m4[0] * fa + o->m4[1] * fb + o->m4[2] * fc;
          } else if (q == 2) {
            double fa = DIAG(1, 1), fb = DIAG(2, 2);
            r = o->m3[0] * fa + o->m3[1] * fb;
          } else if (q == 1) {
            double fa = TR(0, 1), fb = TR(0, -1), fc = TR(0, -2), fd = TR(0, -3);
            r = o->m2[0] * fa + o->m2[1] * fb + o->m2[2] * fc + o->m2[3] * fd;
          } else {
            double fa = TR(0, -1), fb = TR(0, -2), fc = TR(0, -3), fd = TR(0, -4);
            r = o->m1[0] * fa + o->m1[1] * fb + o->m1[2] * fc + o->m1[3] * fd;
          }
        } else {
          double fa = DIAG(1, 1), fb = DIAG(2, 2), fc = DIAG(3, 3), fd = DIAG(4, 4),
                 fe = DIAG(5, 5);
          r = a[0] * fa + a[1] * fb + a[2] * fc + a[3] * fd + a[4] * fe;
        }
#undef TR
#undef DIAG
#undef F
        out[II(o, i, j, k)] = r * o->ovdel[d1] * o->ovdel[d2];
      }
}

/* zero a boundary plane of an interior array, src/zeroxl.f:47-53 */
static void zero_face(orc_t *o, int face, double *a) {
  if (face == 0 || face == 1) {
    int i = face == 0 ? 0 : o->nx - 1;
    for (int k = 0; k < o->nz; k++)
      for (int j = 0; j < o->ny; j++) a[II(o, i, j, k)] = ZERO;
  } else if (face == 2 || face == 3) {
    int j = face == 2 ? 0 : o->ny - 1;
    for (int k = 0; k < o->nz; k++)
      for (int i = 0; i < o->nx; i++) a[II(o, i, j, k)] = ZERO;
  } else {
    int k = face == 4 ? 0 : o->nz - 1;
    for (int j = 0; j < o->ny; j++)
      for (int i = 0; i < o->nx; i++) a[II(o, i, j, k)] = ZERO;
  }
}

/* apply a flag set: x faces to ax, y faces to ay, z faces to az */
static void zero_flags(orc_t *o, const int *flag, double *ax, double *ay, double *az) {
  if (flag[0]) zero_face(o, 0, ax);
  if (flag[1]) zero_face(o, 1, ax);
  if (flag[2]) zero_face(o, 2, ay);
  if (flag[3]) zero_face(o, 3, ay);
  if (flag[4]) zero_face(o, 4, az);
  if (flag[5]) zero_face(o, 5, az);
}

/* ------------------------------------------------------------------------ */
/* point physics                                                            */

/* ITNDEX packing, src/com_senga2.h:88-98: base NTBASE = 4 digits, NSPIMX = 15
   species per 32-bit integer, NINTMX = 2 integers per point; species isp is
   digit isp % 15 of integer isp / 15. */
#define NSPIMX 15
static inline int itint_get(const uint32_t *w, int isp) {
  return (int)((w[isp / NSPIMX] >> (2 * (isp % NSPIMX))) & 3u);
}
static inline void itint_put(uint32_t *w, int isp, int it) {
  w[isp / NSPIMX] += (uint32_t)it << (2 * (isp % NSPIMX));
}

/* interval search: first interval whose upper limit is >= T */
static inline int locate(const sg_config *c, int isp, double t) {
  int it = 0;
  while (t > c->tinthi[isp][it] && it < c->ntint[isp] - 1) it++;
  return it;
}

/* TEMPER, src/temper.f:58-225, over the ranges ISTALT..ISTOLT */
static int temper(orc_t *o) {
  const sg_config *c = &o->c;
  const double toltmp = (double)1.0e-10f;   /* REAL literal, src/temper.f:33-34 */
  int fail = 0;
  OMP_FOR
  for (int k = o->tlo[2]; k <= o->thi[2]; k++)
    for (int j = o->tlo[1]; j <= o->thi[1]; j++)
      for (int i = o->tlo[0]; i <= o->thi[0]; i++) {
        size_t b = BI(o, i, j, k);
        double d = R(o, 0)[b], ru = R(o, 1)[b], rv = R(o, 2)[b], rw = R(o, 3)[b];
        double ukuk = (ru * ru + rv * rv + rw * rw) / d;
        double tc[6], td[6];
        tc[0] = HALF * ukuk - R(o, 4)[b];
        for (int q = 1; q <= 5; q++) { tc[q] = ZERO; td[q] = ZERO; }
        double s7 = ZERO;
        const uint32_t *idx = o->itndex + 2 * b;
        for (int isp = 0; isp < o->ns; isp++) {
          int it = itint_get(idx, isp);
          double y = R(o, 5 + isp)[b];
          tc[0] = tc[0] + y * c->amascp[isp][it][c->ncenth[isp][it] - 1];
          tc[1] = tc[1] + y * c->amasct[isp][it][0];
          td[1] = tc[1];
          for (int q = 2; q <= c->ncpoly[isp][it]; q++) {
            tc[q] = tc[q] + y * c->amasct[isp][it][q - 1];
            td[q] = td[q] + y * c->amascp[isp][it][q - 1];
          }
          s7 = s7 + y * c->rgspec[isp];
        }
        double t = o->trun[b];
        int iter = 1, bad = 0;
        for (;;) {
          double fp = tc[5], dp = td[5];
          for (int q = 4; q >= 1; q--) {
            fp = tc[q] + fp * t;
            dp = td[q] + dp * t;
          }
          fp = tc[0] + fp * t;
          double dt = -fp / dp;
          if (fabs(dt) > toltmp) {
            if (iter < 100) {
              t = t + dt;
              iter++;
              continue;
            }
            bad = 1;   /* the reference stops here (src/temper.f:156-167) */
          }
          break;
        }
        if (bad) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
          fail = 1;
          continue;
        }
        o->trun[b] = t;
        uint32_t nidx[2] = {0u, 0u};
        double cp = ZERO;
        for (int isp = 0; isp < o->ns; isp++) {
          int it = locate(c, isp, t);
          itint_put(nidx, isp, it);
          double f = c->amascp[isp][it][c->ncpoly[isp][it] - 1];
          for (int q = c->ncpoly[isp][it] - 2; q >= 0; q--) f = f * t + c->amascp[isp][it][q];
          cp = cp + R(o, 5 + isp)[b] * f;
        }
        o->itndex[2 * b] = nidx[0];
        o->itndex[2 * b + 1] = nidx[1];
        o->transp[b] = cp / d;
        o->store7[b] = s7;
        o->prun[b] = t * s7;
      }
  return fail ? SG_ETEMPER : 0;
}

/* TEMPIN, src/tempin.f:55-512: bisection from TRIN in 50 K steps */
static int tempin(orc_t *o) {
  const sg_config *c = &o->c;
  const double toltmp = 1.0e-4, tininc = 5.0e1, tlimlo = 2.0e2, tlimhi = 3.0e3;
  int fail = 0;
  OMP_FOR
  for (int k = o->tlo[2]; k <= o->thi[2]; k++)
    for (int j = o->tlo[1]; j <= o->thi[1]; j++)
      for (int i = o->tlo[0]; i <= o->thi[0]; i++) {
        size_t b = BI(o, i, j, k);
        double d = R(o, 0)[b], ru = R(o, 1)[b], rv = R(o, 2)[b], rw = R(o, 3)[b];
        double ukuk = (ru * ru + rv * rv + rw * rw) / d;
        double tuk2me = HALF * ukuk - R(o, 4)[b];
        double tlower = tlimlo, tupper = tlimhi, t = c->trin;
        int fnconv = 1, bad = 0;
#define RESID(tt, res)                                                          \
  do {                                                                          \
    double tc_[6];                                                              \
    tc_[0] = tuk2me;                                                            \
    for (int q_ = 1; q_ <= 5; q_++) tc_[q_] = ZERO;                             \
    for (int s_ = 0; s_ < o->ns; s_++) {                                        \
      int it_ = locate(c, s_, (tt));                                            \
      double y_ = R(o, 5 + s_)[b];                                              \
      tc_[0] = tc_[0] + y_ * c->amascp[s_][it_][c->ncenth[s_][it_] - 1];        \
      tc_[1] = tc_[1] + y_ * c->amasct[s_][it_][0];                             \
      for (int q_ = 2; q_ <= c->ncpoly[s_][it_]; q_++)                          \
        tc_[q_] = tc_[q_] + y_ * c->amasct[s_][it_][q_ - 1];                    \
    }                                                                           \
    double r_ = tc_[5];                                                         \
    for (int q_ = 4; q_ >= 1; q_--) r_ = tc_[q_] + r_ * (tt);                   \
    (res) = tc_[0] + r_ * (tt);                                                 \
  } while (0)
        double res;
        RESID(t, res);
        if (fabs(res) < toltmp) {
          fnconv = 0;
        } else if (res < ZERO) {
          for (;;) {
            tlower = t;
            t = t + tininc;
            RESID(t, res);
            if (fabs(res) < toltmp) { fnconv = 0; break; }
            if (res < ZERO) {
              if (t < tlimhi) continue;
              bad = 1;   /* the reference stops (src/tempin.f:238,348) */
              break;
            }
            tupper = t;
            break;
          }
        } else {
          for (;;) {
            tupper = t;
            t = t - tininc;
            RESID(t, res);
            if (fabs(res) < toltmp) { fnconv = 0; break; }
            if (res > ZERO) {
              if (t > tlimlo) continue;
              bad = 1;   /* the reference stops (src/tempin.f:238,348) */
              break;
            }
            tlower = t;
            break;
          }
        }
        if (bad) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
          fail = 1;
          continue;
        }
        if (fnconv) {
          for (;;) {
            t = HALF * (tlower + tupper);
            RESID(t, res);
            if (fabs(res) < toltmp) break;
            if (res < ZERO) tlower = t;
            else tupper = t;
          }
        }
#undef RESID
        o->trun[b] = t;
        uint32_t nidx[2] = {0u, 0u};
        double s7 = ZERO, cp = ZERO;
        for (int isp = 0; isp < o->ns; isp++) {
          int it = locate(c, isp, t);
          itint_put(nidx, isp, it);
          double f = c->amascp[isp][it][c->ncpoly[isp][it] - 1];
          for (int q = c->ncpoly[isp][it] - 2; q >= 0; q--) f = f * t + c->amascp[isp][it][q];
          cp = cp + R(o, 5 + isp)[b] * f;
          s7 = s7 + R(o, 5 + isp)[b] * c->rgspec[isp];
        }
        o->itndex[2 * b] = nidx[0];
        o->itndex[2 * b + 1] = nidx[1];
        o->transp[b] = cp / d;
        o->store7[b] = s7;
        o->prun[b] = t * s7;
      }
  return fail ? SG_ETEMPIN : 0;
}

/* CHRATE, src/chrate.f:146-760, interior; yrhs holds rho Y */
static void chrate(orc_t *o) {
  const sg_config *c = &o->c;
  const double ysmall = 1.0e-30, ydenom = 1.0e-15;   /* src/chrate.f:39-40 */
  int ns = o->ns;
  OMP_FOR
  for (int k = 0; k < o->nz; k++)
    for (int j = 0; j < o->ny; j++)
      for (int i = 0; i < o->nx; i++) {
        size_t b = BI(o, i, j, k), p = II(o, i, j, k);
        double t = o->trun[b];
        const uint32_t *idx = o->itndex + 2 * b;
        double rate[SG_MAX_SPEC];
        for (int s = 0; s < ns; s++) rate[s] = ZERO;
        for (int st = 0; st < c->nstep; st++) {
          int flthrd = 1;
          double s3 = ZERO;
          if (c->mblist[st] != 0) {
            int ib = c->mblist[st] - 1;
            for (int s = 0; s < ns; s++) {
              double ovwmas = c->ovwmol[s] * c->effy3b[ib][s];
              s3 = s3 + R(o, 5 + s)[b] * ovwmas;
            }
          }
          double s2 = c->rparam[st][0] + c->rparam[st][1] * log(t) - c->rparam[st][2] / t;
          double s1 = exp(s2);
          if (c->mllist[st] != 0) {
            const double *rl = c->rclind[c->mllist[st] - 1];
            double s4 = rl[0] + rl[1] * log(t) - rl[2] / t;
            s4 = exp(s4);
            double preduc = s3 * s4 / s1;
            /* FLCNST is an implicitly typed REAL in src/chrate.f:249 */
            double flcnst = (double)(float)rl[3];
            double fornow = flcnst * preduc / (ONE + preduc);
            s1 = s1 * fornow;
            s2 = log(s1);
            flthrd = 0;
          }
          if (c->mtlist[st] != 0) {
            const double *rt = c->rctroe[c->mtlist[st] - 1];
            /* TALPHA, CFCST1/2, ENCST1/2, DTCNST are implicitly typed REALs
             * in src/chrate.f:319-330 */
            double talpha = (double)(float)rt[3];
            double cfcst1 = (double)(float)rt[7], cfcst2 = (double)(float)rt[8];
            double encst1 = (double)(float)rt[9], encst2 = (double)(float)rt[10];
            double dtcnst = (double)(float)rt[11];
            double omalph = ONE - talpha;
            double s4 = rt[0] + rt[1] * log(t) - rt[2] / t;
            s4 = exp(s4);
            double preduc = s3 * s4 / s1;
            double trats1 = t * rt[4];
            double trats2 = -rt[5] / t;
            double trats3 = t * rt[6];
            double ftcent = omalph * exp(trats3) + talpha * exp(trats1) + exp(trats2);
            ftcent = log10(ftcent);
            double cfactr = cfcst1 + cfcst2 * ftcent;
            double enfact = encst1 + encst2 * ftcent;
            double fornow = log10(preduc) + cfactr;
            fornow = fornow / (enfact - dtcnst * fornow);
            fornow = ONE + fornow * fornow;
            double fbroad = ftcent / fornow;
            fbroad = exp(fbroad * c->clnten);
            fornow = fbroad * preduc / (ONE + preduc);
            s1 = s1 * fornow;
            s2 = log(s1);
            flthrd = 0;
          }
          if (c->mslist[st] != 0) {
            const double *rs = c->rcsrif[c->mslist[st] - 1];
            double s4 = rs[0] + rs[1] * log(t) - rs[2] / t;
            s4 = exp(s4);
            double preduc = s3 * s4 / s1;
            double fornow = rs[4] / t;
            double fbroad = rs[3] * exp(fornow);
            fornow = t * rs[5];
            fbroad = fbroad + exp(fornow);
            fornow = log10(preduc);
            fornow = ONE / (ONE + log10(fornow));
            fbroad = rs[6] * exp(fornow * log(fbroad));
            fornow = fbroad * exp(rs[7] * log(t));
            fornow = fbroad * preduc / (ONE + preduc);
            s1 = s1 * fornow;
            s2 = log(s1);
            flthrd = 0;
          }
          for (int q = 0; q < c->nrslen[st]; q++) {
            int s = c->nrspec[st][q];
            double fornow = R(o, 5 + s)[b] * c->ovwmol[s];
            if (!(fornow > ZERO)) fornow = ZERO;   /* MAX(x,0) */
            s1 = s1 * fornow;
          }
          for (int q = 0; q < c->nrclen[st]; q++) {
            int s = c->nrcpec[st][q];
            double fornow = R(o, 5 + s)[b] * c->ovwmol[s];
            if (!(fornow > ysmall)) fornow = ysmall;
            fornow = exp(c->crspec[st][q] * log(fornow));
            fornow = fornow / (ONE + ydenom * fornow);
            s1 = s1 * fornow;
          }
          if (c->mglist[st] != 0) {
            for (int q = 0; q < c->nsslen[st]; q++) {
              int s = c->nsspec[st][q];
              int it = itint_get(idx, s);
              double g = c->amolgb[s][it][c->ncpoly[s][it] - 1];
              for (int cc = c->ncpoly[s][it] - 2; cc >= 0; cc--) g = c->amolgb[s][it][cc] + g * t;
              g = c->amolgb[s][it][c->ncenth[s][it] - 1] / t -
                  c->amolgb[s][it][c->ncenpy[s][it] - 1] * log(t) - g;
              s2 = s2 + c->diffmu[st][q] * g;
            }
            s2 = exp(s2);
            for (int q = 0; q < c->npslen[st]; q++) {
              int s = c->npspec[st][q];
              double fornow = R(o, 5 + s)[b] * c->ovwmol[s];
              if (!(fornow > ZERO)) fornow = ZERO;
              s2 = s2 * fornow;
            }
            for (int q = 0; q < c->npclen[st]; q++) {
              int s = c->npcpec[st][q];
              double fornow = R(o, 5 + s)[b] * c->ovwmol[s];
              if (!(fornow > ysmall)) fornow = ysmall;
              fornow = exp(c->cpspec[st][q] * log(fornow));
              fornow = fornow / (ONE + ydenom * fornow);
              s2 = s2 * fornow;
            }
            s1 = s1 - s2;
          }
          if (c->mblist[st] != 0 && flthrd) s1 = s1 * s3;
          for (int q = 0; q < c->nsslen[st]; q++) {
            int s = c->nsspec[st][q];
            rate[s] = rate[s] + s1 * c->diffmw[st][q];
          }
        }
        for (int s = 0; s < ns; s++) o->rate[(size_t)s * o->ni + p] = rate[s];
      }
}

/* ------------------------------------------------------------------------ */
/* RHSCAL, src/rhscal.f:61-4047                                             */

#define FOR_I(o) OMP_FOR \
  for (int k = 0; k < (o)->nz; k++) for (int j = 0; j < (o)->ny; j++) for (int i = 0; i < (o)->nx; i++)
#define FOR_B(o) OMP_FOR                           \
  for (int k = (o)->blo[2]; k <= (o)->bhi[2]; k++) \
    for (int j = (o)->blo[1]; j <= (o)->bhi[1]; j++) for (int i = (o)->blo[0]; i <= (o)->bhi[0]; i++)
#define FOR_T(o) OMP_FOR                           \
  for (int k = (o)->tlo[2]; k <= (o)->thi[2]; k++) \
    for (int j = (o)->tlo[1]; j <= (o)->thi[1]; j++) for (int i = (o)->tlo[0]; i <= (o)->thi[0]; i++)
#define FOR_FACE(o) OMP_FOR for (int k = 0; k < (o)->nz; k++) for (int j = 0; j < (o)->ny; j++)

static int rhscal(orc_t *o) {
  const sg_config *c = &o->c;
  int ns = o->ns, nx = o->nx;
  size_t ni = o->ni;
  double *D = R(o, 0), *U = R(o, 1), *V = R(o, 2), *W = R(o, 3), *E = R(o, 4);
  double *s1 = o->st1, *s2 = o->st2, *s3 = o->st3, *s4 = o->st4, *s5 = o->st5, *s6 = o->st6;
  double *s7 = o->store7;
  int xf[2] = {0, nx - 1};
  double *face[2] = {o->fxl, o->fxr};
  int rc = temper(o);
  if (rc) return rc;
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) {
        size_t b = BI(o, xf[f], j, k), q = (size_t)k * o->ny + j;
        FS(o, face[f], F_STRG)[q] = o->transp[b];
        FS(o, face[f], F_STRR)[q] = s7[b] / D[b];
      }
  /* mass flux divergence, rhscal.f:141-155 */
  dfbyd(o, 0, U, s1);
  dfbyd(o, 1, V, s2);
  dfbyd(o, 2, W, s3);
  FOR_I(o) {
    size_t p = II(o, i, j, k);
    o->divm[p] = s1[p] + s2[p] + s3[p];
  }
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) {
        size_t p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
        FS(o, face[f], F_T1B)[q] = -(s2[p] + s3[p]);
      }
  /* E: convert to e, rhscal.f:241-249 */
  FOR_T(o) {
    size_t b = BI(o, i, j, k);
    E[b] = E[b] / D[b];
  }
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) FS(o, face[f], F_STRE)[(size_t)k * o->ny + j] = E[BI(o, xf[f], j, k)];
  /* convective terms, rhscal.f:323-415 */
  FOR_I(o) {
    size_t p = II(o, i, j, k);
    s4[p] = E[BI(o, i, j, k)] * o->divm[p];
  }
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = E[b] * U[b]; }
  dfbyd(o, 0, s7, s1);
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = E[b] * V[b]; }
  dfbyd(o, 1, s7, s2);
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = E[b] * W[b]; }
  dfbyd(o, 2, s7, s3);
  FOR_I(o) {
    size_t p = II(o, i, j, k);
    s4[p] = s4[p] + s1[p] + s2[p] + s3[p];
  }
  dfbyd(o, 0, E, s1);
  dfbyd(o, 1, E, s2);
  dfbyd(o, 2, E, s3);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    E[b] = -HALF * (s4[p] + s1[p] * U[b] + s2[p] * V[b] + s3[p] * W[b]);
  }
  /* heat flux, rhscal.f:427-797 */
  dfbyd(o, 0, o->trun, s1);
  dfbyd(o, 1, o->trun, s2);
  dfbyd(o, 2, o->trun, s3);
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) {
        size_t q = (size_t)k * o->ny + j;
        FS(o, face[f], F_STRT)[q] = o->trun[BI(o, xf[f], j, k)];
        FS(o, face[f], F_BCLT)[q] = s1[II(o, xf[f], j, k)];
      }
  FOR_B(o) {
    size_t b = BI(o, i, j, k);
    double fornow = c->alamda * exp(c->rlamda * log(o->trun[b]));
    s7[b] = fornow * o->transp[b];
    o->transp[b] = fornow;
  }
  if (c->mixavg) {
    FOR_B(o) {
      size_t b = BI(o, i, j, k);
      double ctrans[SG_MAX_SPEC];
      o->transp[b] = log(o->trun[b] / c->tdifgb);
      for (int s = 0; s < ns; s++) {
        double fornow = c->condco[s][c->ncocon - 1];
        for (int q = c->ncocon - 2; q >= 0; q--) fornow = fornow * o->transp[b] + c->condco[s][q];
        ctrans[s] = exp(fornow);
      }
      double combo1 = ZERO, combo2 = ZERO, combo3 = ZERO;
      for (int s = 0; s < ns; s++) {
        double fornow = R(o, 5 + s)[b] * c->ovwmol[s];
        combo1 = combo1 + fornow * ctrans[s];
        combo2 = combo2 + fornow / ctrans[s];
        combo3 = combo3 + fornow;
      }
      combo3 = ONE / combo3;
      combo1 = combo1 * combo3;
      combo2 = combo2 * combo3;
      s7[b] = HALF * (combo1 + ONE / combo2);
      o->wmomix[b] = D[b] * combo3;
    }
  }
  dfbyd(o, 0, s7, s4);
  dfbyd(o, 1, s7, s5);
  dfbyd(o, 2, s7, s6);
  zero_flags(o, o->fcon, s4, s5, s6);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    E[b] = E[b] + (s4[p] * s1[p] + s5[p] * s2[p] + s6[p] * s3[p]);
  }
  d2fd2(o, 0, o->trun, s1);
  d2fd2(o, 1, o->trun, s2);
  d2fd2(o, 2, o->trun, s3);
  zero_flags(o, o->fcon, s1, s2, s3);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    E[b] = E[b] + (s1[p] + s2[p] + s3[p]) * s7[b];
  }
  /* optically thin radiation, rhscal.f:810 (src/radcal.f:36-111): the
     species are still the conserved rho Y_k here */
  if (o->radn) {
    const double foursb = 4.0 * 5.67037321e-08;
    const double trfrth = c->trefrn * c->trefrn * c->trefrn * c->trefrn;
    FOR_I(o) {
      size_t b = BI(o, i, j, k);
      double store1 = ZERO;
      for (int r = 0; r < c->nsprad; r++) {
        double fornow = o->trun[b];
        int nk = c->nkprad[r];
        double plspec = c->akprad[r][nk - 1];
        for (int q = nk - 2; q >= 0; q--) plspec = plspec * fornow + c->akprad[r][q];
        int ks = c->nsprid[r];
        fornow = R(o, 5 + ks)[b] * c->rgspec[ks] * o->trun[b];
        store1 = store1 + plspec * fornow;
      }
      double fornow = o->trun[b];
      fornow = fornow * fornow * fornow * fornow;
      E[b] = E[b] - foursb * store1 * (fornow - trfrth);
    }
  }
  /* chemistry, rhscal.f:822-832 */
  chrate(o);
  memcpy(o->rrte, o->rate, sizeof(double) * ni * ns);
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f])
      for (int s = 0; s < ns; s++) FOR_FACE(o) {
          FSP(o, face[f], FS_RATE, s)[(size_t)k * o->ny + j] = o->rate[(size_t)s * ni + II(o, xf[f], j, k)];
        }
  FOR_I(o) {
    size_t p = II(o, i, j, k);
    o->ucor[p] = ZERO; o->vcor[p] = ZERO; o->wcor[p] = ZERO;
    o->vtmp[BI(o, i, j, k)] = ZERO;   /* VTMP is a big array; interior zeroed */
  }
  FOR_B(o) o->wtmp[BI(o, i, j, k)] = ZERO;
  /* first and second derivatives of ln(mixture molar mass) and ln(pressure),
     rhscal.f:945-992 */
  for (int term = 0; term < 2; term++) {
    if (!(term ? o->mixp : o->mixw)) continue;
    double **d1 = term ? o->pd1 : o->wd1, **d2 = term ? o->pd2 : o->wd2;
    FOR_B(o) {
      size_t b = BI(o, i, j, k);
      s7[b] = log(term ? o->prun[b] : o->wmomix[b]);
    }
    for (int d = 0; d < 3; d++) dfbyd(o, d, s7, d1[d]);
    for (int d = 0; d < 3; d++) d2fd2(o, d, s7, d2[d]);
  }
  /* TRANSP holds ln(T/TDIFGB), rhscal.f:994-1004 (FLMIXT) */
  if (o->ltmask) {
    for (int d = 0; d < 3; d++) dfbyd(o, d, o->transp, o->td1[d]);
    for (int d = 0; d < 3; d++) d2fd2(o, d, o->transp, o->td2[d]);
  }
  /* Y: convert, rhscal.f:1013-1028 */
  for (int s = 0; s < ns; s++) {
    double *Y = R(o, 5 + s);
    FOR_T(o) {
      size_t b = BI(o, i, j, k);
      Y[b] = Y[b] / D[b];
    }
  }
  double *vt = o->vtmp;   /* divergence of the correction velocity (interior use) */
  for (int s = 0; s < ns; s++) {
    double *Y = R(o, 5 + s);
    double *rt = o->rate + (size_t)s * ni;
    FOR_I(o) {
      size_t p = II(o, i, j, k);
      rt[p] = rt[p] - HALF * Y[BI(o, i, j, k)] * o->divm[p];
    }
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b] * U[b]; }
    dfbyd(o, 0, s7, s1);
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b] * V[b]; }
    dfbyd(o, 1, s7, s2);
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b] * W[b]; }
    dfbyd(o, 2, s7, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k);
      rt[p] = rt[p] - HALF * (s1[p] + s2[p] + s3[p]);
    }
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b]; }
    dfbyd(o, 0, s7, s1);
    dfbyd(o, 1, s7, s2);
    dfbyd(o, 2, s7, s3);
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) {
          size_t b = BI(o, xf[f], j, k), p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
          FSP(o, face[f], FS_STRY, s)[q] = Y[b];
          FSP(o, face[f], FS_BCLY, s)[q] = s1[p];
          FSP(o, face[f], FS_T6B, s)[q] = -s2[p] * V[b] / D[b] - s3[p] * W[b] / D[b];
        }
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      rt[p] = rt[p] - HALF * (s1[p] * U[b] + s2[p] * V[b] + s3[p] * W[b]);
    }
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = o->transp[b] * c->olewis[s]; }
    if (c->mixavg) {
      FOR_B(o) {
        size_t b = BI(o, i, j, k);
        double ctrans[SG_MAX_SPEC];
        for (int js = 0; js < ns; js++) {
          double fornow = c->diffco[s][js][c->ncodif - 1];
          for (int q = c->ncodif - 2; q >= 0; q--) fornow = fornow * o->transp[b] + c->diffco[s][js][q];
          ctrans[js] = exp(fornow) * c->pdifgb / o->prun[b];
        }
        double combo1 = ZERO, combo2 = ZERO;
        for (int js = 0; js < ns; js++) {
          double fornow = R(o, 5 + js)[b] + 1.0e-12;   /* DFCTOL */
          combo1 = combo1 + fornow;
          combo2 = combo2 + fornow * c->ovwmol[js] / ctrans[js];
        }
        double fornow = Y[b] + 1.0e-12;
        combo1 = combo1 - fornow;
        combo2 = combo2 - fornow * c->ovwmol[s] / ctrans[s];
        combo2 = combo2 * o->wmomix[b];
        o->difmix[b] = D[b] * combo1 / combo2;
        s7[b] = o->difmix[b];
      }
    }
    /* thermal diffusion ratio of species s relative to the other light
       species, rhscal.f:1353-1410 */
    const int flmtdr = (o->ltmask >> s) & 1u;
    const int flmsor = flmtdr && o->soret, flmduf = flmtdr && o->dufour;
    if (flmtdr) {
      FOR_B(o) {
        size_t b = BI(o, i, j, k);
        double tdr = ZERO;
        for (int js = 0; js < ns; js++) {
          if (!((o->ltmask >> js) & 1u) || js == s) continue;
          double combo2 = o->trun[b] / c->tdifgb;
          double fornow = c->tdrcco[s][js][c->ncotdr - 1];
          for (int q = c->ncotdr - 2; q >= 0; q--) fornow = fornow * combo2 + c->tdrcco[s][js][q];
          double ctrans = fornow;
          fornow = R(o, 5 + js)[b] * c->ovwmol[js];
          tdr = tdr + fornow * o->wmomix[b] * ctrans;
        }
        o->tdrmix[b] = tdr * Y[b] * o->wmomix[b] * c->ovwmol[s];
      }
    }
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      o->ucor[p] = o->ucor[p] + s7[b] * s1[p];
      o->vcor[p] = o->vcor[p] + s7[b] * s2[p];
      o->wcor[p] = o->wcor[p] + s7[b] * s3[p];
    }
    /* species enthalpy, rhscal.f:1444-1470 */
    FOR_B(o) {
      size_t b = BI(o, i, j, k);
      int it = itint_get(o->itndex + 2 * b, s);
      double t = o->trun[b];
      double fornow = c->amasch[s][it][c->ncpoly[s][it] - 1];
      for (int q = c->ncpoly[s][it] - 2; q >= 0; q--) fornow = fornow * t + c->amasch[s][it][q];
      o->utmp[b] = c->amasch[s][it][c->ncenth[s][it] - 1] + fornow * t;
      o->wtmp[b] = o->wtmp[b] + o->utmp[b] * Y[b];
    }
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) FSP(o, face[f], FS_STRH, s)[(size_t)k * o->ny + j] = o->utmp[BI(o, xf[f], j, k)];
    /* Dufour: the species enthalpy carries R_k T TDRMIX through the heat-flux
       terms below, rhscal.f:1549-1566 */
    if (flmduf) FOR_B(o) {
        size_t b = BI(o, i, j, k);
        o->utmp[b] = o->utmp[b] + c->rgspec[s] * o->trun[b] * o->tdrmix[b];
      }
    /* diffusivity gradient terms, rhscal.f:1575-1632 */
    dfbyd(o, 0, s7, s4);
    dfbyd(o, 1, s7, s5);
    dfbyd(o, 2, s7, s6);
    zero_flags(o, o->fdif, s4, s5, s6);
    FOR_I(o) {
      size_t p = II(o, i, j, k);
      double fornow = s4[p] * s1[p] + s5[p] * s2[p] + s6[p] * s3[p];
      rt[p] = rt[p] + fornow;
      vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
    }
    zero_flags(o, o->fadb, s4, s5, s6);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      double fornow = s4[p] * s1[p] + s5[p] * s2[p] + s6[p] * s3[p];
      E[b] = E[b] + fornow * o->utmp[b];
    }
    /* enthalpy gradient terms, rhscal.f:1646-1684 */
    dfbyd(o, 0, o->utmp, s4);
    dfbyd(o, 1, o->utmp, s5);
    dfbyd(o, 2, o->utmp, s6);
    zero_flags(o, o->fdif, s4, s5, s6);
    zero_flags(o, o->fadb, s4, s5, s6);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      double fornow = s4[p] * s1[p] + s5[p] * s2[p] + s6[p] * s3[p];
      E[b] = E[b] + fornow * s7[b];
    }
    /* second-derivative terms, rhscal.f:1883-1959 */
    FOR_I(o) s4[II(o, i, j, k)] = s7[BI(o, i, j, k)];
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b]; }
    d2fd2(o, 0, s7, s1);
    d2fd2(o, 1, s7, s2);
    d2fd2(o, 2, s7, s3);
    zero_flags(o, o->fdif, s1, s2, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k);
      double fornow = (s1[p] + s2[p] + s3[p]) * s4[p];
      rt[p] = rt[p] + fornow;
      vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
    }
    zero_flags(o, o->fadb, s1, s2, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      double fornow = (s1[p] + s2[p] + s3[p]) * s4[p];
      E[b] = E[b] + fornow * o->utmp[b];
    }
    /* molar-mass diffusion (rhscal.f:1970-2364) and pressure diffusion
       (rhscal.f:2365-2759): the two blocks differ only in the logarithm
       differentiated and in the factor of DIFMIX*Y.  They find STORE4 holding
       DIFMIX (moved there for the second-derivative terms above) and STORE5,
       STORE6 the zeroed y and z species enthalpy gradients, and their
       "species enthalpy gradient" heat-flux term uses all three as found:
       DIFMIX in the x component.  The wall terms (FXLDFW ...) need wall
       boundaries, which this path does not take. */
    for (int term = 0; term < 2; term++) {
      if (!(term ? o->mixp : o->mixw)) continue;
      double **d1 = term ? o->pd1 : o->wd1, **d2 = term ? o->pd2 : o->wd2;
      FOR_B(o) {
        size_t b = BI(o, i, j, k);
        s7[b] = term ? o->difmix[b] * Y[b] * (ONE - c->wmolar[s] / o->wmomix[b]) : o->difmix[b] * Y[b];
      }
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        o->ucor[p] = o->ucor[p] + s7[b] * d1[0][p];
        o->vcor[p] = o->vcor[p] + s7[b] * d1[1][p];
        o->wcor[p] = o->wcor[p] + s7[b] * d1[2][p];
      }
      dfbyd(o, 0, s7, s1);
      dfbyd(o, 1, s7, s2);
      dfbyd(o, 2, s7, s3);
      zero_flags(o, o->fdif, s1, s2, s3);
      FOR_I(o) {
        size_t p = II(o, i, j, k);
        double fornow = s1[p] * d1[0][p] + s2[p] * d1[1][p] + s3[p] * d1[2][p];
        rt[p] = rt[p] + fornow;
        vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
      }
      zero_flags(o, o->fadb, s1, s2, s3);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = s1[p] * d1[0][p] + s2[p] * d1[1][p] + s3[p] * d1[2][p];
        E[b] = E[b] + fornow * o->utmp[b];
      }
      zero_flags(o, o->fdif, s4, s5, s6);
      zero_flags(o, o->fadb, s4, s5, s6);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = s4[p] * d1[0][p] + s5[p] * d1[1][p] + s6[p] * d1[2][p];
        E[b] = E[b] + fornow * s7[b];
      }
      zero_flags(o, o->fdif, d2[0], d2[1], d2[2]);
      FOR_I(o) {
        size_t p = II(o, i, j, k);
        double fornow = (d2[0][p] + d2[1][p] + d2[2][p]) * s7[BI(o, i, j, k)];
        rt[p] = rt[p] + fornow;
        vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
      }
      zero_flags(o, o->fadb, d2[0], d2[1], d2[2]);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = (d2[0][p] + d2[1][p] + d2[2][p]) * s7[b];
        E[b] = E[b] + fornow * o->utmp[b];
      }
    }
    /* Soret (thermal diffusion) terms, rhscal.f:2757-3758; the Dufour
       enthalpy is taken out again before their heat-flux terms (2832-2845).
       The wall terms (FXLDFW, FXLADW ...) need wall boundaries. */
    if (flmsor) {
      FOR_B(o) {
        size_t b = BI(o, i, j, k);
        s7[b] = o->difmix[b] * Y[b] * o->tdrmix[b];
      }
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        o->ucor[p] = o->ucor[p] + s7[b] * o->td1[0][p];
        o->vcor[p] = o->vcor[p] + s7[b] * o->td1[1][p];
        o->wcor[p] = o->wcor[p] + s7[b] * o->td1[2][p];
      }
      dfbyd(o, 0, s7, s1);
      dfbyd(o, 1, s7, s2);
      dfbyd(o, 2, s7, s3);
      zero_flags(o, o->fdif, s1, s2, s3);
      FOR_I(o) {
        size_t p = II(o, i, j, k);
        double fornow = s1[p] * o->td1[0][p] + s2[p] * o->td1[1][p] + s3[p] * o->td1[2][p];
        rt[p] = rt[p] + fornow;
        vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
      }
      if (flmduf) FOR_B(o) {
          size_t b = BI(o, i, j, k);
          o->utmp[b] = o->utmp[b] - c->rgspec[s] * o->trun[b] * o->tdrmix[b];
        }
      zero_flags(o, o->fadb, s1, s2, s3);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = s1[p] * o->td1[0][p] + s2[p] * o->td1[1][p] + s3[p] * o->td1[2][p];
        E[b] = E[b] + fornow * o->utmp[b];
      }
      dfbyd(o, 0, o->utmp, s4);
      dfbyd(o, 1, o->utmp, s5);
      dfbyd(o, 2, o->utmp, s6);
      zero_flags(o, o->fdif, s4, s5, s6);
      zero_flags(o, o->fadb, s4, s5, s6);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = s4[p] * o->td1[0][p] + s5[p] * o->td1[1][p] + s6[p] * o->td1[2][p];
        E[b] = E[b] + fornow * s7[b];
      }
      zero_flags(o, o->fdif, o->td2[0], o->td2[1], o->td2[2]);
      FOR_I(o) {
        size_t p = II(o, i, j, k);
        double fornow = (o->td2[0][p] + o->td2[1][p] + o->td2[2][p]) * s7[BI(o, i, j, k)];
        rt[p] = rt[p] + fornow;
        vt[BI(o, i, j, k)] = vt[BI(o, i, j, k)] + fornow;
      }
      zero_flags(o, o->fadb, o->td2[0], o->td2[1], o->td2[2]);
      FOR_I(o) {
        size_t p = II(o, i, j, k), b = BI(o, i, j, k);
        double fornow = (o->td2[0][p] + o->td2[1][p] + o->td2[2][p]) * s7[b];
        E[b] = E[b] + fornow * o->utmp[b];
      }
    }
  }
  /* correction velocity terms, rhscal.f:3782-3843 */
  dfbyd(o, 0, o->wtmp, s1);
  dfbyd(o, 1, o->wtmp, s2);
  dfbyd(o, 2, o->wtmp, s3);
  zero_flags(o, o->fdif, s1, s2, s3);
  zero_flags(o, o->fadb, s1, s2, s3);
  FOR_I(o) s4[II(o, i, j, k)] = vt[BI(o, i, j, k)];
  zero_flags(o, o->fadb, s4, s4, s4);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    E[b] = E[b] - o->wtmp[b] * s4[p] - s1[p] * o->ucor[p] - s2[p] * o->vcor[p] - s3[p] * o->wcor[p];
  }
  /* mixture-averaged viscosity, rhscal.f:3854-3891 */
  if (c->mixavg) {
    FOR_B(o) {
      size_t b = BI(o, i, j, k);
      double ctrans[SG_MAX_SPEC];
      for (int s = 0; s < ns; s++) {
        double fornow = c->viscco[s][c->ncovis - 1];
        for (int q = c->ncovis - 2; q >= 0; q--) fornow = fornow * o->transp[b] + c->viscco[s][q];
        ctrans[s] = exp(fornow);
      }
      double combo1 = ZERO;
      for (int s = 0; s < ns; s++) {
        double combo2 = ZERO;
        for (int js = 0; js < ns; js++) {
          double fornow = sqrt(ctrans[s] / ctrans[js]);
          fornow = ONE + fornow * c->wilko2[s][js];
          fornow = c->wilko1[s][js] * fornow * fornow;
          combo2 = combo2 + R(o, 5 + js)[b] * c->ovwmol[js] * fornow;
        }
        double fornow = ctrans[s] / combo2;
        combo1 = combo1 + R(o, 5 + s)[b] * c->ovwmol[s] * fornow;
      }
      o->difmix[b] = combo1;
    }
  }
  /* species source terms, rhscal.f:3899-3944 */
  for (int s = 0; s < ns; s++) {
    double *Y = R(o, 5 + s);
    double *rt = o->rate + (size_t)s * ni;
    FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = Y[b]; }
    dfbyd(o, 0, s7, s1);
    dfbyd(o, 1, s7, s2);
    dfbyd(o, 2, s7, s3);
    zero_flags(o, o->fdif, s1, s2, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Y[b] = rt[p] - Y[b] * vt[b] - s1[p] * o->ucor[p] - s2[p] * o->vcor[p] - s3[p] * o->wcor[p];
    }
  }
  /* density and its gradient for BCs, rhscal.f:3966-4047 */
  if (o->fcnv[0] || o->fcnv[1]) {
    dfbyd(o, 0, D, s1);
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) {
          size_t q = (size_t)k * o->ny + j;
          FS(o, face[f], F_STRD)[q] = D[BI(o, xf[f], j, k)];
          FS(o, face[f], F_BCL2)[q] = s1[II(o, xf[f], j, k)];
        }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RHSVEL, src/rhsvel.f:51-2318                                             */

static void rhsvel(orc_t *o) {
  const sg_config *c = &o->c;
  int nx = o->nx;
  double *D = R(o, 0), *U = R(o, 1), *V = R(o, 2), *W = R(o, 3), *E = R(o, 4);
  double *s1 = o->st1, *s2 = o->st2, *s3 = o->st3, *s4 = o->st4, *s5 = o->st5, *s6 = o->st6;
  double *s7 = o->store7, *ut = o->utmp, *vt = o->vtmp, *wt = o->wtmp, *tr = o->transp;
  int xf[2] = {0, nx - 1};
  double *face[2] = {o->fxl, o->fxr};
  const double thrd = o->thrd, tthd = o->tthd, fthd = o->fthd;
  (void)thrd;
  FOR_B(o) {
    size_t b = BI(o, i, j, k);
    ut[b] = U[b] / D[b];
    vt[b] = V[b] / D[b];
    wt[b] = W[b] / D[b];
  }
  int anyx = o->fcnv[0] || o->fcnv[1];
  /* transverse terms for x faces, rhsvel.f:66-338 */
  if (anyx) {
    dfbyd(o, 1, ut, s1);
    dfbyd(o, 2, ut, s2);
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) {
          size_t b = BI(o, xf[f], j, k), p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
          FS(o, face[f], F_STRU)[q] = ut[b];
          FS(o, face[f], F_STRV)[q] = vt[b];
          FS(o, face[f], F_STRW)[q] = wt[b];
          FS(o, face[f], F_T2B)[q] = -vt[b] * s1[p] - wt[b] * s2[p];
        }
    dfbyd(o, 1, vt, s1);
    dfbyd(o, 2, vt, s2);
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) {
          size_t b = BI(o, xf[f], j, k), p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
          FS(o, face[f], F_T3B)[q] = -vt[b] * s1[p] - wt[b] * s2[p];
        }
    dfbyd(o, 1, wt, s1);
    dfbyd(o, 2, wt, s2);
    for (int f = 0; f < 2; f++)
      if (o->fcnv[f]) FOR_FACE(o) {
          size_t b = BI(o, xf[f], j, k), p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
          FS(o, face[f], F_T4B)[q] = -vt[b] * s1[p] - wt[b] * s2[p];
        }
  }
  /* momentum convective terms, rhsvel.f:348-463 */
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = U[b] * ut[b]; }
  dfbyd(o, 0, s7, s4);
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = U[b] * vt[b]; }
  dfbyd(o, 1, s7, s1);
  dfbyd(o, 0, s7, s5);
  FOR_I(o) { size_t p = II(o, i, j, k); s4[p] = s4[p] + s1[p]; }
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = U[b] * wt[b]; }
  dfbyd(o, 2, s7, s1);
  dfbyd(o, 0, s7, s6);
  FOR_I(o) { size_t p = II(o, i, j, k); s4[p] = s4[p] + s1[p]; }
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = V[b] * vt[b]; }
  dfbyd(o, 1, s7, s1);
  FOR_I(o) { size_t p = II(o, i, j, k); s5[p] = s5[p] + s1[p]; }
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = V[b] * wt[b]; }
  dfbyd(o, 2, s7, s1);
  dfbyd(o, 1, s7, s2);
  FOR_I(o) {
    size_t p = II(o, i, j, k);
    s5[p] = s5[p] + s1[p];
    s6[p] = s6[p] + s2[p];
  }
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = W[b] * wt[b]; }
  dfbyd(o, 2, s7, s1);
  FOR_I(o) { size_t p = II(o, i, j, k); s6[p] = s6[p] + s1[p]; }
  dfbyd(o, 0, ut, s1);
  dfbyd(o, 1, vt, s2);
  dfbyd(o, 2, wt, s3);
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) FS(o, face[f], F_BCL1)[(size_t)k * o->ny + j] = s1[II(o, xf[f], j, k)];
  /* rhsvel.f:622-729 */
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    s7[b] = s4[p] + ut[b] * o->divm[p];
    s4[p] = U[b] * s1[p];
    U[b] = -HALF * (s4[p] + s7[b]);
    s7[b] = s5[p] + vt[b] * o->divm[p];
    s5[p] = V[b] * s2[p];
    V[b] = -HALF * (s5[p] + s7[b]);
    s7[b] = s6[p] + wt[b] * o->divm[p];
    s6[p] = W[b] * s3[p];
    W[b] = -HALF * (s6[p] + s7[b]);
  }
  /* pressure gradient, rhsvel.f:741-917 */
  double prefer = o->prun[BI(o, 0, 0, 0)];
  FOR_B(o) { size_t b = BI(o, i, j, k); s7[b] = o->prun[b] - prefer; }
  dfbyd(o, 0, s7, s4);
  dfbyd(o, 1, s7, s5);
  dfbyd(o, 2, s7, s6);
  for (int f = 0; f < 2; f++)
    if (o->fcnv[f]) FOR_FACE(o) {
        size_t b = BI(o, xf[f], j, k), p = II(o, xf[f], j, k), q = (size_t)k * o->ny + j;
        FS(o, face[f], F_STRP)[q] = o->prun[b];
        FS(o, face[f], F_BCL5)[q] = s4[p];
        FS(o, face[f], F_T3B)[q] = FS(o, face[f], F_T3B)[q] - s5[p] / D[b];
        FS(o, face[f], F_T4B)[q] = FS(o, face[f], F_T4B)[q] - s6[p] / D[b];
        FS(o, face[f], F_T51B)[q] = -vt[b] * s5[p] - wt[b] * s6[p];
        FS(o, face[f], F_T52B)[q] = -o->prun[b] * (s2[p] + s3[p]);
      }
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    U[b] = U[b] - s4[p];
    V[b] = V[b] - s5[p];
    W[b] = W[b] - s6[p];
    E[b] = E[b] - ut[b] * s4[p] - vt[b] * s5[p] - wt[b] * s6[p];
  }
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    E[b] = E[b] - o->prun[b] * (s1[p] + s2[p] + s3[p]);
  }
  /* viscosity, rhsvel.f:926-980 */
  FOR_B(o) { size_t b = BI(o, i, j, k); tr[b] = tr[b] * c->prantl; }   /* PRANTL, src/com_senga2.h:115 */
  if (c->mixavg) FOR_B(o) { size_t b = BI(o, i, j, k); tr[b] = o->difmix[b]; }
  /* TAUXX, rhsvel.f:992-1090 */
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    s6[p] = s2[p] + s3[p];
    s6[p] = fthd * s1[p] - tthd * s6[p];
    E[b] = E[b] + tr[b] * s6[p] * s1[p];
  }
  dfbyd(o, 0, tr, s4);
  FOR_I(o) { size_t p = II(o, i, j, k); s6[p] = s6[p] * s4[p]; }
  zero_flags(o, (int[6]){o->fvsn[0], o->fvsn[1], 0, 0, 0, 0}, s6, s6, s6);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    U[b] = U[b] + s6[p];
    E[b] = E[b] + ut[b] * s6[p];
  }
  /* TAUYY */
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    s6[p] = s1[p] + s3[p];
    s6[p] = fthd * s2[p] - tthd * s6[p];
    E[b] = E[b] + tr[b] * s6[p] * s2[p];
  }
  dfbyd(o, 1, tr, s5);
  FOR_I(o) { size_t p = II(o, i, j, k); s6[p] = s6[p] * s5[p]; }
  zero_flags(o, (int[6]){0, 0, o->fvsn[2], o->fvsn[3], 0, 0}, s6, s6, s6);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    V[b] = V[b] + s6[p];
    E[b] = E[b] + vt[b] * s6[p];
  }
  /* TAUZZ */
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    s1[p] = s1[p] + s2[p];
    s1[p] = fthd * s3[p] - tthd * s1[p];
    E[b] = E[b] + tr[b] * s1[p] * s3[p];
  }
  dfbyd(o, 2, tr, s6);
  FOR_I(o) { size_t p = II(o, i, j, k); s1[p] = s1[p] * s6[p]; }
  zero_flags(o, (int[6]){0, 0, 0, 0, o->fvsn[4], o->fvsn[5]}, s1, s1, s1);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    W[b] = W[b] + s1[p];
    E[b] = E[b] + wt[b] * s1[p];
  }
  /* cross terms: (a,b) velocity pairs with plane (d1,d2), rhsvel.f:1298-2185 */
  struct { int d1, d2; double *va, *vb; double *Ra, *Rb; int fa, fb; double *dmua, *dmub;
           int bcl_a; } pairs[3] = {
      {0, 1, ut, vt, U, V, 0, 2, s4, s5, F_BCL3},
      {0, 2, ut, wt, U, W, 0, 4, s4, s6, F_BCL4},
      {1, 2, vt, wt, V, W, 2, 4, s5, s6, -1}};
  for (int pi = 0; pi < 3; pi++) {
    int d1 = pairs[pi].d1, d2 = pairs[pi].d2;
    double *va = pairs[pi].va, *vb = pairs[pi].vb, *Ra = pairs[pi].Ra, *Rb = pairs[pi].Rb;
    double *ra = (pi == 2) ? vt : ut;   /* velocity multiplying Rb's convective term */
    double *rb = (pi == 0) ? vt : wt;   /* velocity multiplying Ra's convective term */
    int fa = pairs[pi].fa, fb = pairs[pi].fb;   /* face index base of directions d1, d2 */
    double *dmu1 = pairs[pi].dmua, *dmu2 = pairs[pi].dmub;
    /* d va / d x_d2 */
    dfbyd(o, d2, va, s1);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Ra[b] = Ra[b] - HALF * D[b] * rb[b] * s1[p];
    }
    /* d vb / d x_d1 */
    dfbyd(o, d1, vb, s2);
    if (pairs[pi].bcl_a >= 0)
      for (int f = 0; f < 2; f++)
        if (o->fcnv[f]) FOR_FACE(o) FS(o, face[f], pairs[pi].bcl_a)[(size_t)k * o->ny + j] = s2[II(o, xf[f], j, k)];
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Rb[b] = Rb[b] - HALF * D[b] * ra[b] * s2[p];
    }
    FOR_I(o) { size_t p = II(o, i, j, k); s7[BI(o, i, j, k)] = s1[p] + s2[p]; }
    d2fdmix(o, d1, d2, va, s1);
    d2fdmix(o, d1, d2, vb, s2);
    FOR_I(o) { size_t p = II(o, i, j, k); s3[p] = tthd * tr[BI(o, i, j, k)] * s2[p]; }
    zero_flags(o, (int[6]){fa == 0 ? o->fvsn[0] : 0, fa == 0 ? o->fvsn[1] : 0,
                           fa == 2 ? o->fvsn[2] : 0, fa == 2 ? o->fvsn[3] : 0, 0, 0}, s3, s3, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Ra[b] = Ra[b] - s3[p];
      E[b] = E[b] - va[b] * s3[p];
    }
    FOR_I(o) { size_t p = II(o, i, j, k); s3[p] = tthd * tr[BI(o, i, j, k)] * s1[p]; }
    zero_flags(o, (int[6]){0, 0, fb == 2 ? o->fvsn[2] : 0, fb == 2 ? o->fvsn[3] : 0,
                           fb == 4 ? o->fvsn[4] : 0, fb == 4 ? o->fvsn[5] : 0}, s3, s3, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Rb[b] = Rb[b] - s3[p];
      E[b] = E[b] - vb[b] * s3[p];
    }
    d2fd2(o, d2, va, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      s3[p] = s3[p] + s2[p];
      s3[p] = tr[b] * s3[p] + s7[b] * dmu2[p];
    }
    zero_flags(o, (int[6]){0, 0, fb == 2 ? o->fvst[2] : 0, fb == 2 ? o->fvst[3] : 0,
                           fb == 4 ? o->fvst[4] : 0, fb == 4 ? o->fvst[5] : 0}, s3, s3, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Ra[b] = Ra[b] + s3[p];
      E[b] = E[b] + va[b] * s3[p];
    }
    d2fd2(o, d1, vb, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      s3[p] = s3[p] + s1[p];
      s3[p] = tr[b] * s3[p] + s7[b] * dmu1[p];
    }
    zero_flags(o, (int[6]){fa == 0 ? o->fvst[0] : 0, fa == 0 ? o->fvst[1] : 0,
                           fa == 2 ? o->fvst[2] : 0, fa == 2 ? o->fvst[3] : 0, 0, 0}, s3, s3, s3);
    FOR_I(o) {
      size_t p = II(o, i, j, k), b = BI(o, i, j, k);
      Rb[b] = Rb[b] + s3[p];
      E[b] = E[b] + vb[b] * s3[p];
    }
    FOR_I(o) {
      size_t b = BI(o, i, j, k);
      E[b] = E[b] + tr[b] * s7[b] * s7[b];
    }
  }
  /* normal second derivatives, rhsvel.f:2189-2305 */
  d2fd2(o, 0, ut, s1);
  d2fd2(o, 1, vt, s2);
  d2fd2(o, 2, wt, s3);
  zero_flags(o, o->fvsn, s1, s2, s3);
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    double fornow = tr[b] * fthd * s1[p];
    U[b] = U[b] + fornow;
    E[b] = E[b] + fornow * ut[b];
  }
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    double fornow = tr[b] * fthd * s2[p];
    V[b] = V[b] + fornow;
    E[b] = E[b] + fornow * vt[b];
  }
  FOR_I(o) {
    size_t p = II(o, i, j, k), b = BI(o, i, j, k);
    double fornow = tr[b] * fthd * s3[p];
    W[b] = W[b] + fornow;
    E[b] = E[b] + fornow * wt[b];
  }
  FOR_I(o) {
    size_t b = BI(o, i, j, k);
    D[b] = -o->divm[II(o, i, j, k)];
  }
}

/* ------------------------------------------------------------------------ */
/* BOUNDT / BOUNDS / BOUNTT for NSBCI3 (x-left) and NSBCO1 (x-right)        */

/* src/boundt.f:177-218 with BCDTXL/BCUTXL(type 1)/BCYTXL */
static void boundt(orc_t *o) {
  if (o->c.bc_type[0] != SG_BC_INFLOW3) return;
  const sg_config *c = &o->c;
  FOR_FACE(o) {
    size_t b = BI(o, 0, j, k), q = (size_t)k * o->ny + j;
    double strd = o->drin, stru = c->bc_rpar[0][0], strv = ZERO, strw = ZERO;
    FS(o, o->fxl, F_STRD)[q] = strd;
    FS(o, o->fxl, F_STRU)[q] = stru;
    FS(o, o->fxl, F_STRV)[q] = strv;
    FS(o, o->fxl, F_STRW)[q] = strw;
    R(o, 0)[b] = strd;
    R(o, 1)[b] = strd * stru;
    R(o, 2)[b] = strd * strv;
    R(o, 3)[b] = strd * strw;
    for (int s = 0; s < o->ns; s++) {
      FSP(o, o->fxl, FS_STRY, s)[q] = c->yrin[s];
      R(o, 5 + s)[b] = strd * c->yrin[s];
    }
  }
}

/* src/bountt.f:195-262 */
static void bountt(orc_t *o) {
  if (o->c.bc_type[0] != SG_BC_INFLOW3) return;
  const sg_config *c = &o->c;
  FOR_FACE(o) {
    size_t b = BI(o, 0, j, k), p = II(o, 0, j, k);
    double strd = o->drin, stru = c->bc_rpar[0][0];
    R(o, 0)[b] = strd;
    R(o, 1)[b] = strd * stru;
    R(o, 2)[b] = strd * ZERO;
    R(o, 3)[b] = strd * ZERO;
    for (int v = 0; v < 4; v++) { RUN(o, v)[p] = R(o, v)[b]; ERR(o, v)[p] = ZERO; }
    for (int s = 0; s < o->ns; s++) {
      R(o, 5 + s)[b] = strd * c->yrin[s];
      RUN(o, 5 + s)[p] = R(o, 5 + s)[b];
      ERR(o, 5 + s)[p] = ZERO;
    }
  }
}

/* src/bounds.f:115-206 (prep) and 610-682 (NSBCI3 x-left) */
static void bounds_xl(orc_t *o) {
  const sg_config *c = &o->c;
  double *fc = o->fxl;
  double bet = 0.050;
  (void)bet;
  FOR_FACE(o) {
    size_t q = (size_t)k * o->ny + j, b = BI(o, 0, j, k);
    double strg = FS(o, fc, F_STRG)[q], strr = FS(o, fc, F_STRR)[q], strt = FS(o, fc, F_STRT)[q];
    double strh[SG_MAX_SPEC];
    for (int s = 0; s < o->ns; s++)
      strh[s] = FSP(o, fc, FS_STRH, s)[q] - strg * strt * c->rgspec[s] / strr;
    double gam1 = strg - strr;
    double stre = FS(o, fc, F_STRE)[q] - gam1 * strt;
    gam1 = strr / gam1;
    double ovgm = ONE / gam1;
    double fornow = strg * gam1 * strt;
    double acou = sqrt(fornow), ova2 = ONE / fornow;
    double strd = FS(o, fc, F_STRD)[q], stru = FS(o, fc, F_STRU)[q], strv = FS(o, fc, F_STRV)[q],
           strw = FS(o, fc, F_STRW)[q];
    double bcl1 = FS(o, fc, F_BCL1)[q], bcl2 = FS(o, fc, F_BCL2)[q], bcl3 = FS(o, fc, F_BCL3)[q],
           bcl4 = FS(o, fc, F_BCL4)[q], bcl5 = FS(o, fc, F_BCL5)[q];
    const double dudt = ZERO, dddt = ZERO, dvdt = ZERO, dwdt = ZERO;
    fornow = strd * acou * bcl1;
    bcl1 = HALF * (stru - acou) * (bcl5 - fornow);
    bcl2 = stru * (bcl2 - bcl5 * ova2);
    bcl3 = stru * bcl3;
    bcl4 = stru * bcl4;
    bcl5 = HALF * (stru + acou) * (bcl5 + fornow);
    fornow = bcl1 - strd * acou * dudt;
    bcl2 = -dddt - ova2 * (bcl1 + fornow) - bcl2;
    bcl3 = -dvdt - bcl3;
    bcl4 = -dwdt - bcl4;
    bcl5 = fornow - bcl5;
    double *E = R(o, 4);
    E[b] = E[b] - bcl2 * stre - bcl3 * strd * strv - bcl4 * strd * strw -
           bcl5 * (ova2 * stre + stru / acou + ovgm);
    for (int s = 0; s < o->ns; s++) {
      double bcly = FSP(o, fc, FS_RATE, s)[q] / strd - ZERO - stru * FSP(o, fc, FS_BCLY, s)[q];
      E[b] = E[b] - bcly * strd * strh[s];
    }
  }
}

/* src/bounds.f:1196-1458 (prep and NSBCO1 x-right, flags 0) */
static void bounds_xr(orc_t *o) {
  const sg_config *c = &o->c;
  double *fc = o->fxr;
  double bet = 0.050;
  FOR_FACE(o) {
    size_t q = (size_t)k * o->ny + j, b = BI(o, o->nx - 1, j, k);
    double strg = FS(o, fc, F_STRG)[q], strr = FS(o, fc, F_STRR)[q], strt = FS(o, fc, F_STRT)[q];
    double strh[SG_MAX_SPEC];
    for (int s = 0; s < o->ns; s++)
      strh[s] = FSP(o, fc, FS_STRH, s)[q] - strg * strt * c->rgspec[s] / strr;
    double gam1 = strg - strr;
    double stre = FS(o, fc, F_STRE)[q] - gam1 * strt;
    gam1 = strr / gam1;
    double ovgm = ONE / gam1;
    double fornow = strg * gam1 * strt;
    double acou = sqrt(fornow), ova2 = ONE / fornow;
    double strd = FS(o, fc, F_STRD)[q], stru = FS(o, fc, F_STRU)[q], strv = FS(o, fc, F_STRV)[q],
           strw = FS(o, fc, F_STRW)[q], strp = FS(o, fc, F_STRP)[q];
    double tt1 = FS(o, fc, F_T51B)[q] + FS(o, fc, F_T52B)[q] * (gam1 + 1.0) -
                 strd * acou * FS(o, fc, F_T2B)[q];
    double sorp = ZERO;
    for (int s = 0; s < o->ns; s++) sorp = sorp + strh[s] * FSP(o, fc, FS_RATE, s)[q];
    sorp = -sorp * gam1;
    double bcl1 = FS(o, fc, F_BCL1)[q], bcl5 = FS(o, fc, F_BCL5)[q];
    bcl1 = HALF * (stru - acou) * (bcl5 - strd * acou * bcl1);
    bcl1 = HALF * sorp + o->cobcxr * acou * (strp - o->pinfxr) + 0.5 * (1.0 - bet) * tt1 - bcl1;
    R(o, 0)[b] = R(o, 0)[b] - bcl1 * ova2;
    R(o, 1)[b] = R(o, 1)[b] - bcl1 * ova2 * (stru - acou);
    R(o, 2)[b] = R(o, 2)[b] - bcl1 * ova2 * strv;
    R(o, 3)[b] = R(o, 3)[b] - bcl1 * ova2 * strw;
    R(o, 4)[b] = R(o, 4)[b] - bcl1 * (ova2 * stre - stru / acou + ovgm);
    for (int s = 0; s < o->ns; s++)
      R(o, 5 + s)[b] = R(o, 5 + s)[b] - bcl1 * ova2 * FSP(o, fc, FS_STRY, s)[q];
  }
}

/* ------------------------------------------------------------------------ */
/* RK register update, src/lincom.f:53-169, src/fincom.f:54-168             */

static void lincom(orc_t *o, int irk) {
  int nv = 5 + o->ns;
  double lhs = o->rklhs[irk - 1], rhs = o->rkrhs[irk - 1], er = o->rkerr[irk - 1];
  for (int v = 0; v < nv; v++) {
    int slot = v < 5 ? v : 5;
    double *r = R(o, v), *u = RUN(o, v), *e = ERR(o, v);
    OMP_FOR
    for (int k = 0; k < o->nz; k++)
      for (int j = 0; j < o->ny; j++)
        for (int i = o->rklo[slot]; i <= o->rkhi[slot]; i++) {
          size_t p = II(o, i, j, k), b = BI(o, i, j, k);
          e[p] = e[p] + er * r[b];
          double fornow = u[p];
          u[p] = fornow + lhs * r[b];
          r[b] = fornow + rhs * r[b];
        }
  }
}

static void fincom(orc_t *o) {
  int nv = 5 + o->ns;
  double lhs = o->rklhs[4], er = o->rkerr[4];
  for (int v = 0; v < nv; v++) {
    int slot = v < 5 ? v : 5;
    double *r = R(o, v), *u = RUN(o, v), *e = ERR(o, v);
    OMP_FOR
    for (int k = 0; k < o->nz; k++)
      for (int j = 0; j < o->ny; j++)
        for (int i = o->rklo[slot]; i <= o->rkhi[slot]; i++) {
          size_t p = II(o, i, j, k), b = BI(o, i, j, k);
          e[p] = e[p] + er * r[b];
          u[p] = u[p] + lhs * r[b];
          r[b] = u[p];
        }
  }
}

/* ------------------------------------------------------------------------ */
/* public API                                                               */

orc_t *orc_create(const sg_config *cfg, int *err) {
  *err = 0;
  if (cfg->nspec < 1 || cfg->nspec > SG_MAX_SPEC || cfg->nstep < 0 || cfg->nstep > SG_MAX_STEP ||
      cfg->nx < 11 || cfg->ny < 11 || cfg->nz < 11) {
    *err = SG_EINVAL;
    return NULL;
  }
  if (cfg->physics & SG_PHYS_RADN) {
    int bad = cfg->nsprad < 0 || cfg->nsprad > cfg->nspec;
    for (int r = 0; !bad && r < cfg->nsprad; r++)
      bad = cfg->nsprid[r] < 0 || cfg->nsprid[r] >= cfg->nspec || cfg->nkprad[r] < 1 || cfg->nkprad[r] > SG_NKPRAD;
    if (bad) {
      *err = SG_EINVAL;
      return NULL;
    }
  }
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->c = *cfg;
  o->nx = cfg->nx; o->ny = cfg->ny; o->nz = cfg->nz; o->ns = cfg->nspec;
  o->bx = o->nx + 2 * HALO; o->by = o->ny + 2 * HALO; o->bz = o->nz + 2 * HALO;
  o->nb = (size_t)o->bx * o->by * o->bz;
  o->ni = (size_t)o->nx * o->ny * o->nz;
  int rc = bcinit(o);
  if (rc) { free(o); *err = rc; return NULL; }
  dfinit(o);
  /* inflow density, src/indata.f:1178-1183 */
  double rtin = ZERO;
  for (int s = 0; s < o->ns; s++) rtin = rtin + cfg->rgspec[s] * cfg->yrin[s];
  rtin = rtin * cfg->trin;
  o->drin = cfg->prin / rtin;
  int nv = 5 + o->ns;
  o->run = (double *)calloc((size_t)nv * o->ni, sizeof(double));
  o->err = (double *)calloc((size_t)nv * o->ni, sizeof(double));
  o->rhs = (double *)calloc((size_t)nv * o->nb, sizeof(double));
  double **big[] = {&o->trun, &o->prun, &o->transp, &o->store7, &o->utmp, &o->vtmp, &o->wtmp,
                    &o->wmomix, &o->difmix};
  for (size_t q = 0; q < sizeof(big) / sizeof(big[0]); q++) *big[q] = (double *)calloc(o->nb, sizeof(double));
  o->itndex = (uint32_t *)calloc(2 * o->nb, sizeof(uint32_t));
  double **in[] = {&o->st1, &o->st2, &o->st3, &o->st4, &o->st5, &o->st6, &o->divm, &o->ucor,
                   &o->vcor, &o->wcor};
  for (size_t q = 0; q < sizeof(in) / sizeof(in[0]); q++) *in[q] = (double *)calloc(o->ni, sizeof(double));
  o->rate = (double *)calloc(o->ni * o->ns, sizeof(double));
  o->rrte = (double *)calloc(o->ni * o->ns, sizeof(double));
  /* molar-mass and pressure diffusion act with mixture-averaged transport
     only (src/diffin.f:384-386) */
  o->mixw = cfg->mixavg && (cfg->physics & SG_PHYS_MIXW);
  o->mixp = cfg->mixavg && (cfg->physics & SG_PHYS_MIXP);
  o->radn = (cfg->physics & SG_PHYS_RADN) != 0;
  o->soret = cfg->mixavg && (cfg->physics & SG_PHYS_SORET);
  o->dufour = cfg->mixavg && (cfg->physics & SG_PHYS_DUFOUR);
  o->ltmask = 0;
  if (o->soret || o->dufour)   /* src/diffin.f:388-407, WMLTOL = 1.0D-3 */
    for (int s = 0; s < o->ns; s++)
      if (cfg->wmolar[s] <= cfg->wmltdr + 1.0e-3) o->ltmask |= 1u << s;
  for (int d = 0; d < 3; d++) {
    if (o->mixw) { o->wd1[d] = (double *)calloc(o->ni, sizeof(double)); o->wd2[d] = (double *)calloc(o->ni, sizeof(double)); }
    if (o->mixp) { o->pd1[d] = (double *)calloc(o->ni, sizeof(double)); o->pd2[d] = (double *)calloc(o->ni, sizeof(double)); }
    if (o->ltmask) { o->td1[d] = (double *)calloc(o->ni, sizeof(double)); o->td2[d] = (double *)calloc(o->ni, sizeof(double)); }
  }
  if (o->ltmask) o->tdrmix = (double *)calloc(o->nb, sizeof(double));
  size_t fsz = ((size_t)F_NSCAL + (size_t)F_NSPEC * o->ns) * o->ny * o->nz;
  o->fxl = (double *)calloc(fsz, sizeof(double));
  o->fxr = (double *)calloc(fsz, sizeof(double));
  /* pre-initialise T as src/indata.f:1167-1174 */
  for (size_t b = 0; b < o->nb; b++) { o->trun[b] = cfg->trin; o->prun[b] = cfg->prin; }
  orc_set_timestep(o, 0.0);
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  void *ptrs[] = {o->run, o->err, o->rhs, o->trun, o->prun, o->transp, o->store7, o->utmp, o->vtmp,
                  o->wtmp, o->wmomix, o->difmix, o->itndex, o->st1, o->st2, o->st3, o->st4, o->st5,
                  o->st6, o->divm, o->ucor, o->vcor, o->wcor, o->rate, o->rrte, o->fxl, o->fxr};
  for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++) free(ptrs[q]);
  for (int d = 0; d < 3; d++) {
    free(o->wd1[d]); free(o->wd2[d]); free(o->pd1[d]); free(o->pd2[d]); free(o->td1[d]); free(o->td2[d]);
  }
  free(o->tdrmix);
  free(o);
}

/* src/dtinit.f:76-94 (RK4(3)5[2R+]C) and 196-223 */
void orc_set_timestep(orc_t *o, double dt) {
  const double acofrk[5] = {970286171893.0 / 4311952581923.0, 6584761158862.0 / 12103376702013.0,
                            2251764453980.0 / 15575788980749.0, 26877169314380.0 / 34165994151039.0,
                            0.0};
  const double bcofrk[5] = {1153189308089.0 / 22510343858157.0, 1772645290293.0 / 4653164025191.0,
                            -1672844663538.0 / 4480602732383.0, 2114624349019.0 / 3568978502595.0,
                            5198255086312.0 / 14908931495163.0};
  const double bhatrk[5] = {1016888040809.0 / 7410784769900.0, 11231460423587.0 / 58533540763752.0,
                            -1563879915014.0 / 6823010717585.0, 606302364029.0 / 971179775848.0,
                            1097981568119.0 / 3980877426909.0};
  o->tstep = dt;
  for (int r = 0; r < 5; r++) {
    o->rklhs[r] = bcofrk[r] * dt;
    o->rkrhs[r] = acofrk[r] * dt;
    o->rkerr[r] = (bcofrk[r] - bhatrk[r]) * dt;
  }
  o->rktim[0] = ZERO;
  for (int r = 1; r < 5; r++) {
    double t = acofrk[r - 1];
    for (int q = 0; q < r - 1; q++) t = t + bcofrk[q];
    o->rktim[r] = t * dt;
  }
}

int orc_set_state(orc_t *o, const double *run, double etime) {
  int nv = 5 + o->ns;
  memcpy(o->run, run, sizeof(double) * nv * o->ni);
  memset(o->err, 0, sizeof(double) * nv * o->ni);
  /* src/indata.f:1496-1571: RHS <- RUN, wrap, TEMPIN */
  for (int v = 0; v < nv; v++) {
    double *r = R(o, v);
    FOR_I(o) r[BI(o, i, j, k)] = RUN(o, v)[II(o, i, j, k)];
    wrap(o, r);
  }
  o->etime = etime;
  int rc = tempin(o);
  if (rc) return rc;
  memset(o->rrte, 0, sizeof(double) * o->ni * o->ns);
  o->have_state = 1;
  return 0;
}

static int eval_rhs(orc_t *o, int irk) {
  (void)irk;
  int nv = 5 + o->ns;
  boundt(o);
  for (int v = 0; v < nv; v++) wrap(o, R(o, v));
  int rc = rhscal(o);
  if (rc) return rc;
  rhsvel(o);
  if (o->fcnv[0]) bounds_xl(o);
  if (o->fcnv[1]) bounds_xr(o);
  return 0;
}

int orc_rhs(orc_t *o, int irk, double *rhs) {
  if (!o->have_state) return SG_ESTATE;
  int nv = 5 + o->ns;
  double *save = (double *)malloc(sizeof(double) * nv * o->nb);
  memcpy(save, o->rhs, sizeof(double) * nv * o->nb);
  int rc = eval_rhs(o, irk);
  if (!rc)
    for (int v = 0; v < nv; v++) FOR_I(o) rhs[(size_t)v * o->ni + II(o, i, j, k)] = R(o, v)[BI(o, i, j, k)];
  memcpy(o->rhs, save, sizeof(double) * nv * o->nb);
  free(save);
  return rc;
}

int orc_substage(orc_t *o, int irk) {
  if (!o->have_state) return SG_ESTATE;
  if (irk < 1 || irk > 5) return SG_EINVAL;
  int rc = eval_rhs(o, irk);
  if (rc) return rc;
  if (irk < 5) lincom(o, irk);
  else fincom(o);
  return 0;
}

/* ADAPTT, src/adaptt.f:60-330, with the tolerances, error norms and PID
 * controller constants of src/dtinit.f:101-170. */
static void adaptt(orc_t *o) {
  const double errtol = 1.0e-4, errlow = 1.0e-30, trmax = 1.01, trmin = 1.0e-2, tsmax = 1.0,
               tsmin = 1.0e-15;
  const double ctmult = 9.0e-1, ctalph = 4.9e-1 / 3.0, ctbeta = 3.4e-1 / 3.0, ctgama = 1.0e-1 / 3.0;
  int nv = 5 + o->ns;
  double errmax = ZERO;
  for (int v = 0; v < nv; v++) {
    /* ERDNRM = 0, ERUNRM = ERVNRM = ERWNRM = 1e-6, ERENRM = 1e-2, ERYNRM = 1e-10 */
    double nrm = v == 0 ? 0.0 : (v < 4 ? 1.0e-6 : (v == 4 ? 1.0e-2 : 1.0e-10));
    double tot = ZERO;
    const double *e = ERR(o, v), *u = RUN(o, v);
    for (size_t p = 0; p < o->ni; p++) {
      double fornow = fabs(e[p]) / (fabs(u[p]) + nrm);
      if (fornow > tot) tot = fornow;
    }
    if (tot > errmax) errmax = tot;
  }
  if (errmax < errlow) errmax = errlow;
  double tratio = ctmult * exp(ctalph * log(errtol / errmax) + ctbeta * log(o->errold / errtol) +
                               ctgama * log(errtol / o->errldr));
  o->errldr = o->errold;
  o->errold = errmax;
  if (tratio > trmax) tratio = trmax;
  if (tratio < trmin) tratio = trmin;
  double tstold = o->tstep;
  o->tstep = o->tstep * tratio;
  if (o->tstep > tsmax) o->tstep = tsmax;
  if (o->tstep < tsmin) o->tstep = tsmin;
  tratio = o->tstep / tstold;
  for (int r = 0; r < 5; r++) {
    o->rklhs[r] = o->rklhs[r] * tratio;
    o->rkrhs[r] = o->rkrhs[r] * tratio;
    o->rkerr[r] = o->rkerr[r] * tratio;
    o->rktim[r] = o->rktim[r] * tratio;
  }
  memset(o->err, 0, sizeof(double) * nv * o->ni);
}

int orc_set_adaptive(orc_t *o, int on, double errold, double errldr) {
  o->adaptive = on;
  o->errold = errold;
  o->errldr = errldr;
  return 0;
}

int orc_get_timestep(orc_t *o, double *tstep, double *errold, double *errldr) {
  if (tstep) *tstep = o->tstep;
  if (errold) *errold = o->errold;
  if (errldr) *errldr = o->errldr;
  return 0;
}

int orc_end_step(orc_t *o) {
  o->etime = o->etime + o->tstep;
  bountt(o);
  if (o->adaptive) adaptt(o);
  return 0;
}

int orc_step(orc_t *o, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    for (int r = 1; r <= 5; r++) {
      int rc = orc_substage(o, r);
      if (rc) return rc;
    }
    orc_end_step(o);
  }
  return 0;
}

int orc_get(orc_t *o, double *run, double *temp, double *pres, double *rate) {
  if (run) memcpy(run, o->run, sizeof(double) * (5 + o->ns) * o->ni);
  if (temp) FOR_I(o) temp[II(o, i, j, k)] = o->trun[BI(o, i, j, k)];
  if (pres) FOR_I(o) pres[II(o, i, j, k)] = o->prun[BI(o, i, j, k)];
  if (rate) memcpy(rate, o->rrte, sizeof(double) * o->ns * o->ni);
  return 0;
}

int orc_get_stage(orc_t *o, double *stage) {
  for (int v = 0; v < 5 + o->ns; v++) FOR_I(o) stage[(size_t)v * o->ni + II(o, i, j, k)] = R(o, v)[BI(o, i, j, k)];
  return 0;
}

int orc_get_itndex(orc_t *o, int *itndex) {
  FOR_I(o) itndex[II(o, i, j, k)] = (int)o->itndex[2 * BI(o, i, j, k)];
  return 0;
}

int orc_derivatives(orc_t *o, const double *f, double *out9) {
  double *big = (double *)calloc(o->nb, sizeof(double));
  FOR_I(o) big[BI(o, i, j, k)] = f[II(o, i, j, k)];
  wrap(o, big);
  dfbyd(o, 0, big, out9 + 0 * o->ni);
  dfbyd(o, 1, big, out9 + 1 * o->ni);
  dfbyd(o, 2, big, out9 + 2 * o->ni);
  d2fd2(o, 0, big, out9 + 3 * o->ni);
  d2fd2(o, 1, big, out9 + 4 * o->ni);
  d2fd2(o, 2, big, out9 + 5 * o->ni);
  d2fdmix(o, 0, 1, big, out9 + 6 * o->ni);
  d2fdmix(o, 0, 2, big, out9 + 7 * o->ni);
  d2fdmix(o, 1, 2, big, out9 + 8 * o->ni);
  free(big);
  return 0;
}

/* Threads of the parallel loops (OpenMP builds): n > 0 sets the count;
   returns the count in use (1 without OpenMP). */
int orc_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
