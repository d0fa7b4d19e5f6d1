// sg_internal.h -- shared definitions of the HIP right-hand-side library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "senga_gpu.h"

namespace sg {

constexpr int HALO = 5;

// Grid, stencil coefficients and boundary flags handed to every kernel by
// value.  Coefficients are the values src/dfinit.f:61-477 computes.
struct Params {
  int nx, ny, nz, ns;
  long sy, sz;                 // element strides of y and z
  long npts;                   // nx*ny*nz
  long fstride;                // elements between consecutive fields
  int bnd[3];                  // 1: direction bounded (closures), 0: periodic
  double ovdel[3], ovdl2[3];
  double ac[5], c1[4], c2[4], c3[2], c4[3], c5[4];       // first derivative
  double sc[5], s1[5], s2[5], s3[2], s4[3], s5[4];       // second derivative
  double mc[5], my1[2], m1[4], m2[4], m3[2], m4[3], m5[4];  // mixed
  double thrd, tthd, fthd;
  // boundary flags per face xl,xr,yl,yr,zl,zr (src/bcinit.f)
  int fcnv[6], fvsn[6], fvst[6], fcon[6], fadb[6], fdif[6];
  int inflow_xl, outflow_xr;   // NSBCI3 at x-left, NSBCO1 at x-right
  int rklo[6];                 // first x index of the RK update for D,U,V,W,E,Y
  double drin, uin, cobcxr, pinfxr;
  double yrin[SG_MAX_SPEC];    // inflow mass fractions (NSBCI3)
  int precision;               // 0: FP64; 1: mixed (FP32 primitives, rates, RHS; FP64 registers, T)
  int renorm;                  // mixed only: FP64 sum(rho Y_k) = rho renormalisation at FINCOM
  double pofs;                 // stored pressure is p - pofs (mixed: the default pressure, else 0)
  int mixavg;
  int mixw, mixp;              // molar-mass / pressure diffusion (mixture-averaged transport only)
  int soret, dufour;           // thermal diffusion and its heat flux, for the species in ltmask
  int radn;                    // optically thin radiation
  uint32_t ltmask;             // light species (molar mass <= WMLTDR + 1e-3) when soret or dufour
  double alamda, rlamda, prantl, tdifgb, pdifgb;
  // mixed form: the temperatures at which every rate exponent of the
  // mechanism fits FP32's exponential (outside, the rates are formed in FP64)
  double chlo, chhi;
};

// Thermodynamic tables derived from sg_config when a handle is created.  The
// NASA polynomials are held as fixed SG_NCPMAX-term arrays, zero-padded above
// each interval's own order (a zero leading coefficient leaves Horner's result
// bit-identical), with the enthalpy and entropy constants pulled out, so that
// kernels evaluate them with fixed trip counts and uniform loads.
#define SG_NCPMAX 5   // highest cp order the temperature solve accepts (src/temper.f NCTMAX)
#define SG_NROW 11    // derivative row kinds: interior, 5 left and 5 right closure rows
struct Tables {
  double roww[SG_NROW][10];                      // 5-term row weights W1, W2 per row kind
  double mixw[SG_NROW][7];                       // mixed-derivative row weights W0..W4, V0, V1
  double cp[SG_MAX_SPEC][SG_NTINT][SG_NCPMAX];   // amascp polynomial part
  double ct[SG_MAX_SPEC][SG_NTINT][SG_NCPMAX];   // amasct
  double ch[SG_MAX_SPEC][SG_NTINT][SG_NCPMAX];   // amasch
  double gb[SG_MAX_SPEC][SG_NTINT][SG_NCPMAX];   // amolgb
  double hcp[SG_MAX_SPEC][SG_NTINT];             // amascp(ncenth)
  double hch[SG_MAX_SPEC][SG_NTINT];             // amasch(ncenth)
  double gbh[SG_MAX_SPEC][SG_NTINT];             // amolgb(ncenth)
  double gbs[SG_MAX_SPEC][SG_NTINT];             // amolgb(ncenpy)
  // dense per-step rate coefficients: nuw[st][k] = diffmw of species k in
  // step st's list (0 if absent); valid when no species repeats in a list
  double nuw[SG_MAX_STEP][SG_MAX_SPEC];
  double wmol[SG_MAX_SPEC];                      // molar masses (WMOLAR), pressure diffusion
  double rgs[SG_MAX_SPEC];                       // species gas constants (RGSPEC), Dufour
  int diffsym;                                   // binary diffusion fits symmetric in (i, j)
  int ratedense;                                 // nuw is usable
  // FP32 copies of the tables the mixed-precision kernels read, so that
  // uniform coefficients arrive as scalar FP32 loads instead of per-lane
  // FP64 -> FP32 conversions
  float ovwf[SG_MAX_SPEC], olewf[SG_MAX_SPEC];
  float condf[SG_MAX_SPEC][SG_NTRC], viscf[SG_MAX_SPEC][SG_NTRC];
  float difff[SG_MAX_SPEC][SG_MAX_SPEC][SG_NTRC];
  float wilk1f[SG_MAX_SPEC][SG_MAX_SPEC], wilk2f[SG_MAX_SPEC][SG_MAX_SPEC];
  float effyf[SG_MAX_BODY][SG_MAX_SPEC];
  float nuwf[SG_MAX_STEP][SG_MAX_SPEC];
  float diffmwf[SG_MAX_STEP][SG_MAX_SSL], crspecf[SG_MAX_STEP][SG_MAX_SSL], cpspecf[SG_MAX_STEP][SG_MAX_SSL];
};

// Device field set: all arrays are [field][z][y][x] with field stride fstride.
struct Fields {
  const Tables *tab;   // derived thermodynamic tables (device)
  double *q;      // stage register (5+ns): rho, rho u, rho v, rho w, rho E, rho Y
  double *run;    // solution register (5+ns)
  double *prim;   // primitives: see PrimSlot, then Y[ns], rhoD[ns], h[ns]
  double *rate;   // chemical source (ns), kg/m^3/s
  double *acc;    // RHS accumulators (5+ns) + VT (divergence of the correction velocity)
  double *grad;   // dmu/dx, dmu/dy, dmu/dz (the velocity gradients the final kernel forms itself)
  uint32_t *itndex;  // packed temperature interval indices (1 bit per species)
  double *face;   // NSCBC face data, see FaceSlot, [face 0/1][slot][ny*nz]
  int *err;       // [0]=error code, [1..3] = failing point (1-based)
  double *erk;    // RK error register (5+ns), adaptive time step only, else null
  unsigned long long *emax;   // global error-norm maximum (bits of a non-negative double)
};

// P_LNW, P_LNP, P_OVW: ln(mixture molar mass), ln(pressure), 1/(mixture molar
// mass), written when molar-mass or pressure diffusion is on; P_LNT:
// ln(T/TDIFGB), written when Soret or Dufour is on
enum PrimSlot { P_U = 0, P_V, P_W, P_E, P_T, P_P, P_LAM, P_MU, P_HMIX, P_LNW, P_LNP, P_OVW, P_LNT, P_NSCAL };
// face slots: scalars then per-species.  Besides the NSCBC data proper (cp,
// mixture gas constant, pressure and density gradients) the passes copy
// into them, at the x-face points, every value the NSCBC kernel reads
// (velocity, T, E, stored p, rho, the velocity gradients it uses), so that
// it reads contiguous face arrays instead of one cache line per value.
enum FaceSlot { F_CP = 0, F_RR, F_DPDX, F_DPDY, F_DPDZ, F_U, F_V, F_W, F_T, F_E, F_P, F_RHO,
                F_DUX, F_DVX, F_DWX, F_DUY, F_DVY, F_DUZ, F_DWZ, F_NSCAL };
// per species: dY/dx at F_NSCAL + s, h at face_h, Y at face_y, the
// chemical rate at face_rate; then the NSCBC corrections of the 5 + ns
// right-hand sides at face_corr + v (written by k_nscbc)

inline __host__ __device__ long prim_y(const Params &P, int s) { return (long)(P_NSCAL + s) * P.fstride; }
inline __host__ __device__ long prim_rd(const Params &P, int s) { return (long)(P_NSCAL + P.ns + s) * P.fstride; }
inline __host__ __device__ long prim_h(const Params &P, int s) { return (long)(P_NSCAL + 2 * P.ns + s) * P.fstride; }
// with Soret or Dufour on (tdr): per species the thermal diffusion ratio
// TDRMIX and T * TDRMIX follow h
inline __host__ __device__ long prim_tdr(const Params &P, int s) { return (long)(P_NSCAL + 3 * P.ns + s) * P.fstride; }
inline __host__ __device__ long prim_tt(const Params &P, int s) { return (long)(P_NSCAL + 4 * P.ns + s) * P.fstride; }
inline __host__ __device__ int nprim(int ns, int tdr) { return P_NSCAL + (tdr ? 5 : 3) * ns; }
inline __host__ __device__ int face_h(int ns, int s) { return F_NSCAL + ns + s; }
inline __host__ __device__ int face_y(int ns, int s) { return F_NSCAL + 2 * ns + s; }
inline __host__ __device__ int face_rate(int ns, int s) { return F_NSCAL + 3 * ns + s; }
inline __host__ __device__ int face_corr(int ns) { return F_NSCAL + 4 * ns; }
inline __host__ __device__ int nface(int ns) { return face_corr(ns) + 5 + ns; }

// Substage launches cover any number of cases ("members") with one grid per
// kernel kind.  The grid of a kind is the concatenation of every member's own
// grid of that kind: flattened block b belongs to member map[kind][b] and is
// that member's block b - blk0[kind].  A single case is a launch of one member.
// L_FACE2: both x faces; L_PASSXE: the edge tiles of bounded x rows (the x
// pass of a bounded x direction is two launches, interior and edge tiles)
enum LaunchKind { L_FACE = 0, L_FACE2, L_POINT, L_PASSX, L_PASSY, L_PASSZ, L_FINAL, L_PASSXE, L_NKIND };

struct Member {
  Params P;
  Fields F;
  double rkl[5], rkr[5], rke[5];   // RKLHS, RKRHS, RKERR of the member's time step (src/dtinit.f:196-223)
  int blk0[L_NKIND];               // first flattened block of the member per kind
};

// Host-side view of one launch set (device pointers).  Every member shares
// the mechanism (c, tab), the species count, the precision and the x
// boundary kind, which select the kernel instantiation.
struct Launch {
  const Member *mem;
  const uint32_t *map[L_NKIND];
  int nblk[L_NKIND];
  int nmem;
  const sg_config *c;
  const Tables *tab;
  int ns, precision, bndx, inflow;
  int ext;          // a molecular-transport option on: the passes with those terms
  int radn;         // some member has radiation on: the radiation kernel after the x pass
  int pass_nw[3];   // pass tile per direction: 0 wide (rows of 32), 1 narrow (rows of 16)
  int mechfix;      // the chemistry is the compiled mechanism (MechH2): unrolled step loop
};

// member block counts of each kind (sg_kernels.hip)
void member_blocks(const Params &P, const int nw[3], int nblk[L_NKIND]);

// launchers (sg_kernels.hip)
void launch_tempin(const Params &P, const sg_config *dcfg, const Fields &F, hipStream_t st);
void launch_boundt(const Launch &L, hipStream_t st);
void launch_prim(const Launch &L, hipStream_t st);
void launch_pass(int dir, const Launch &L, hipStream_t st);
// NSCBC face corrections (bounded x only), before launch_final
void launch_nscbc(const Launch &L, hipStream_t st);
// optically thin radiation, added to the energy accumulator after the x pass
void launch_radn(const Launch &L, hipStream_t st);
// substage irk (1..5): LINCOM for irk < 5, FINCOM for irk = 5, with each
// member's coefficients; mode 2 (irk = 0): write the RHS to out instead
void launch_final(const Launch &L, int irk, double *out, hipStream_t st);
// adaptive step: max over points and variables of |err|/(|run| + norm) into F.emax
void launch_errmax(const Params &P, const Fields &F, hipStream_t st);
void launch_zero(double *p, long n, hipStream_t st);
void launch_bountt(const Launch &L, hipStream_t st);
void launch_derivs(const Params &P, const double *f, double *out, hipStream_t st);
bool species_supported(int ns);

}  // namespace sg
