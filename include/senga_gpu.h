/*
 * senga_gpu.h -- C ABI of the MI355X right-hand-side library for SENGA+.
 *
 * One call per Runge-Kutta substage replaces the reference's substage body
 *   BOUNDT; PARFER; RHSCAL; RHSVEL; BOUNDS; LINCOM   (src/senga2.f:113-135)
 * and its last-substage variant ending in FINCOM     (src/senga2.f:141-159).
 * The per-step synchronisation BOUNTT (src/senga2.f:171) is sg_end_step.
 *
 * The reference has no plugin interface on this path: its units are
 * argument-less Fortran subroutines that share COMMON blocks
 * (src/com_senga2.h).  The ABI below carries that COMMON state explicitly:
 * sg_config holds what INDATA derives from cont.dat/chem.dat/diff.dat
 * (src/indata.f, src/chemin.f:822-983, src/diffin.f:280-405), the handle owns
 * the state registers of src/com_senga2.h:576-597 in device memory.
 *
 * Conventions
 *   - Host arrays are Fortran column-major interior arrays, x fastest, the
 *     variable (or species) index last:  a(nx,ny,nz[,nvar]).  The solution
 *     register is (rho, rho u, rho v, rho w, rho E, rho Y_1..rho Y_N), i.e.
 *     DRUN, URUN, VRUN, WRUN, ERUN, YRUN of src/com_senga2.h:576-583 in that
 *     order, 5+N fields.
 *   - Species indices inside sg_config are 0-based (the reference's lists
 *     are 1-based).
 *   - All functions return 0 on success, a nonzero SG_E* code otherwise; no
 *     call aborts the process.  sg_last_error gives the failing point.
 *   - One host thread per handle; calls are synchronous from the caller's
 *     view unless documented otherwise.
 */
#ifndef SENGA_GPU_H
#define SENGA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_ABI_VERSION 4

#define SG_MAX_SPEC 18   /* the reference's largest shipped mechanism (casefiles/68step: NSPCMX=18) */
#define SG_MAX_STEP 128
#define SG_MAX_SSL 10    /* NSSMAX: species per step list */
#define SG_MAX_BODY 16
#define SG_MAX_FALL 32   /* Lindemann / Troe / SRI steps of each kind */
#define SG_NCOF 7        /* NASA-7 coefficients per interval */
#define SG_NTINT 2       /* temperature intervals per species */
#define SG_NTRC 4        /* coefficients of the transport fits (diff.dat) */
#define SG_NKPRAD 6      /* coefficients of a Planck mean absorption fit (radn.dat, src/com_senga2.h NCFRMX) */

/* error codes */
#define SG_OK 0
#define SG_EINVAL 1       /* bad argument or unsupported configuration */
#define SG_ENOMEM 2       /* device allocation failed */
#define SG_EDEVICE 3      /* HIP runtime error */
#define SG_ETEMPER 4      /* Newton temperature iteration failed (src/temper.f:150-167) */
#define SG_ETEMPIN 5      /* initial bisection failed (src/tempin.f:238,348) */
#define SG_ESTATE 6       /* call out of order (no state uploaded, ...) */

/* boundary type codes of the reference (src/com_senga2.h NSBCCL block) */
#define SG_BC_PERIODIC 1
#define SG_BC_INFLOW3 13  /* NSBCI3: subsonic reflecting inflow, specified density */
#define SG_BC_OUTFLOW1 21 /* NSBCO1: subsonic non-reflecting outflow */

/* sg_config.physics bits: the reference's optional physics switches from
 * cont.dat (src/contin.f:194-209), set when the switch is on -- the four
 * molecular-transport options only together with mixture-averaged transport,
 * as src/diffin.f:66,384-407 applies them.  All are computed (the transport
 * options with mixture-averaged transport; without it they are inactive, as
 * in the reference). */
#define SG_PHYS_MIXW 1    /* NFMIXW: molar-mass diffusion term (src/rhscal.f:951,1970) */
#define SG_PHYS_MIXP 2    /* NFMIXP: pressure diffusion term (src/rhscal.f:973,2365) */
#define SG_PHYS_SORET 4   /* NFMSOR: thermal diffusion of light species (src/rhscal.f:2760) */
#define SG_PHYS_DUFOUR 8  /* NFMDUF: Dufour heat flux (src/rhscal.f:1552,2834,3108) */
#define SG_PHYS_RADN 16   /* NFRADN: optically thin radiation (src/rhscal.f:810, src/radcal.f) */

typedef struct sg_config {
  /* grid: global = local on one device (src/com_senga2.h:12-34) */
  int32_t nx, ny, nz;
  int32_t nspec, nstep;
  int32_t mixavg;             /* cont.dat mix.av switch (src/diffin.f:64) */
  int32_t precision;          /* 0 = FP64 throughout; 1 = mixed: FP32 primitives,
                                 transport, chemistry, derivatives and RHS
                                 assembly; FP64 RK registers, temperature
                                 Newton iteration, enthalpies and rate
                                 exponents; 2 = mixed as 1, and at the end of
                                 every RK step (FINCOM) the species of the
                                 solution register are renormalised in FP64,
                                 rho Y_k *= rho / sum_j rho Y_j.  The reference
                                 has no such renormalisation (src/lincom.f:294-
                                 333 is commented out) and the FP64 form never
                                 applies it. */
  int32_t physics;            /* SG_PHYS_* bits of switches that are on */
  double xgdlen, ygdlen, zgdlen;  /* domain lengths; spacing L/(N-1) (src/dfinit.f:42-50) */

  /* boundary conditions, faces xl, xr, yl, yr, zl, zr (cont.dat lines 51-56) */
  int32_t bc_type[6];
  int32_t bc_ipar[6][4];
  double bc_rpar[6][4];

  /* default / inflow state (cont.dat "Default initial conditions") */
  double prin, trin, urin, vrin, wrin;
  double yrin[SG_MAX_SPEC];

  /* constant-Lewis transport (src/com_senga2.h:112-115, src/chemin.f:983) */
  double alamda;              /* lambda/cp = alamda * T^rlamda */
  double rlamda;
  double prantl;
  double olewis[SG_MAX_SPEC]; /* 1/Le_k */

  /* thermodynamics (src/chemin.f:858-961), [species][interval][coef] */
  double wmolar[SG_MAX_SPEC], ovwmol[SG_MAX_SPEC], rgspec[SG_MAX_SPEC];
  int32_t ntint[SG_MAX_SPEC];
  int32_t ncpoly[SG_MAX_SPEC][SG_NTINT];   /* number of cp coefficients (5) */
  int32_t ncenth[SG_MAX_SPEC][SG_NTINT];   /* 1-based index of enthalpy constant (6) */
  int32_t ncenpy[SG_MAX_SPEC][SG_NTINT];   /* 1-based index of entropy constant (7) */
  double tinthi[SG_MAX_SPEC][SG_NTINT];
  double amascp[SG_MAX_SPEC][SG_NTINT][SG_NCOF];
  double amasct[SG_MAX_SPEC][SG_NTINT][SG_NCOF];
  double amasch[SG_MAX_SPEC][SG_NTINT][SG_NCOF];
  double amolgb[SG_MAX_SPEC][SG_NTINT][SG_NCOF];

  /* kinetics (src/chrate.f:100-147 names), per step */
  double rparam[SG_MAX_STEP][3];           /* ln A, n, E/R0 */
  int32_t nsslen[SG_MAX_STEP];
  int32_t nsspec[SG_MAX_STEP][SG_MAX_SSL];
  double diffmu[SG_MAX_STEP][SG_MAX_SSL];
  double diffmw[SG_MAX_STEP][SG_MAX_SSL];
  int32_t nrslen[SG_MAX_STEP];
  int32_t nrspec[SG_MAX_STEP][SG_MAX_SSL];
  int32_t npslen[SG_MAX_STEP];
  int32_t npspec[SG_MAX_STEP][SG_MAX_SSL];
  int32_t nrclen[SG_MAX_STEP];
  int32_t nrcpec[SG_MAX_STEP][SG_MAX_SSL];
  double crspec[SG_MAX_STEP][SG_MAX_SSL];
  int32_t npclen[SG_MAX_STEP];
  int32_t npcpec[SG_MAX_STEP][SG_MAX_SSL];
  double cpspec[SG_MAX_STEP][SG_MAX_SSL];
  int32_t mblist[SG_MAX_STEP];   /* 1-based third-body index, 0 = none */
  int32_t mglist[SG_MAX_STEP];   /* nonzero: reversible (Gibbs) step */
  int32_t mllist[SG_MAX_STEP];   /* 1-based Lindemann index, 0 = none */
  int32_t mtlist[SG_MAX_STEP];   /* 1-based Troe index, 0 = none */
  int32_t mslist[SG_MAX_STEP];   /* 1-based SRI index, 0 = none */
  double effy3b[SG_MAX_BODY][SG_MAX_SPEC];
  double rclind[SG_MAX_FALL][4];
  double rctroe[SG_MAX_FALL][12];
  double rcsrif[SG_MAX_FALL][8];
  double clnten;                 /* ln(10) as the reference's CLNTEN (Troe) */

  /* mixture-averaged transport (src/diffin.f:280-405) */
  double pdifgb, tdifgb;
  int32_t ncovis, ncocon, ncodif, reserved1;
  double viscco[SG_MAX_SPEC][SG_NTRC];
  double condco[SG_MAX_SPEC][SG_NTRC];
  double diffco[SG_MAX_SPEC][SG_MAX_SPEC][SG_NTRC];  /* [i][j][coef] = DIFFCO(coef,j,i) */
  double wilko1[SG_MAX_SPEC][SG_MAX_SPEC];             /* [i][j] = WILKO1(j,i) */
  double wilko2[SG_MAX_SPEC][SG_MAX_SPEC];

  /* thermal diffusion ratios for the Soret and Dufour terms (diff.dat,
   * src/diffin.f:180-195), with the reference's antisymmetric fill
   * (src/diffin.f:345-368): [i][j][coef] = TDRCCO(coef,j,i), a polynomial in
   * T/TDIFGB.  They act for the light species, molar mass <= wmltdr + 1e-3
   * (cont.dat WMLTDR, src/diffin.f:388-407). */
  int32_t ncotdr, reserved2;
  double wmltdr;
  double tdrcco[SG_MAX_SPEC][SG_MAX_SPEC][SG_NTRC];

  /* optically thin radiation (NFRADN; radn.dat, src/radcin.f:76-126): nsprad
   * radiating species, species nsprid[r] (0-based) with the Planck mean
   * absorption coefficient sum_c akprad[r][c] T^c over nkprad[r] terms;
   * trefrn: the ambient temperature (cont.dat TREFRN) */
  int32_t nsprad, reserved3;
  int32_t nsprid[SG_MAX_SPEC];
  int32_t nkprad[SG_MAX_SPEC];
  double trefrn;
  double akprad[SG_MAX_SPEC][SG_NKPRAD];
} sg_config;

typedef struct sg_handle sg_handle;

/* ABI identification; sizeof(sg_config) lets bindings check their layout. */
int sg_abi_version(void);
size_t sg_config_size(void);

/* Create a handle on the current HIP device: validates cfg, derives the
 * stencil coefficients (src/dfinit.f), the BC flags and loop ranges
 * (src/bcinit.f), allocates the device registers.  Replaces the table set-up
 * part of INDATA/DFINIT/BCINIT.  Returns SG_EINVAL for a species count that is
 * not compiled in (2..SG_MAX_SPEC are), nx/ny/nz < 11, more than 2^29 points
 * per field, cp fits of more than 5 coefficients (src/temper.f NCTMAX),
 * transport fits or thermal diffusion ratios of more than SG_NTRC
 * coefficients, radiation data out of range (species index, more than
 * SG_NKPRAD coefficients), or a
 * step list, species index, third-body or falloff index out of range. */
int sg_create(const sg_config *cfg, sg_handle **out);
int sg_destroy(sg_handle *h);

/* Upload the solution register run(nx,ny,nz,5+N) and the elapsed time; sets
 * the stage register to it and runs the initial bisection for temperature
 * and the interval index (src/indata.f:1496-1571, src/tempin.f:55-512). */
int sg_upload_state(sg_handle *h, const double *run, double etime);

/* Download any of: run(nx,ny,nz,5+N), temp(nx,ny,nz), pres(nx,ny,nz),
 * rate(nx,ny,nz,N) (chemical source, RRTE of src/rhscal.f:824-832).  NULL
 * pointers are skipped.  T, p and the rate are those of the last RHS
 * evaluation (after upload: from the initial bisection; rate zero). */
int sg_download_state(sg_handle *h, double *run, double *temp, double *pres, double *rate);

/* Fixed time step: RK coefficients times dt (src/dtinit.f:196-223). */
int sg_set_timestep(sg_handle *h, double tstep);

/* One RK substage irk = 1..5 (src/senga2.f:113-159): BOUNDT, periodic wrap,
 * RHSCAL, RHSVEL, BOUNDS, then LINCOM (irk < 5) or FINCOM (irk = 5).
 * Asynchronous on the handle's stream; errors from the temperature Newton
 * iteration are reported by the next synchronising call. */
int sg_substage(sg_handle *h, int irk);

/* End of step: ETIME += TSTEP, BOUNTT, and ADAPTT when the adaptive step
 * is on (src/senga2.f:165-184). */
int sg_end_step(sg_handle *h);

/* Adaptive time step, cont.dat step switch 1 (src/dtinit.f:101-185,
 * src/adaptt.f:60-330): with on = 1 every step accumulates the RK error
 * register (LINCOM/FINCOM) and sg_end_step computes the global error norm,
 * applies the PID step controller and rescales the RK coefficients.  errold
 * and errldr are the controller history, the dump's ERROLD and ERRLDR
 * (1e-4 = ERRTOL on a cold start).  Allocates the error register on first use. */
int sg_set_adaptive(sg_handle *h, int on, double errold, double errldr);

/* Current time step and controller history. */
int sg_get_timestep(sg_handle *h, double *tstep, double *errold, double *errldr);

/* nsteps full steps (5 substages + sg_end_step each). */
int sg_step(sg_handle *h, int nsteps);

/* One right-hand-side evaluation at substage irk without the register update
 * (BOUNDT..BOUNDS); rhs(nx,ny,nz,5+N) receives DRHS..YRHS of the interior.
 * Leaves the registers unchanged. */
int sg_rhs(sg_handle *h, int irk, double *rhs);

/* The nine derivative operators (DFBYDX/Y/Z, D2FDX2/Y2/Z2, D2FDXY/XZ/YZ, in
 * that order) of a host field f(nx,ny,nz) with this handle's end conditions;
 * out(nx,ny,nz,9). */
int sg_derivatives(sg_handle *h, const double *f, double *out);

/* No counterpart in the reference: a variant of sg_derivatives for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same operators, same
 * results; the device copies of f and of the nine outputs live in a fixed
 * static device array instead of per-call allocations, so the call makes no
 * allocation.  Limit: grids of at most SG_NOALLOC_MAX_POINTS points
 * (SG_EINVAL above it); the array (10 x SG_NOALLOC_MAX_POINTS doubles,
 * 40 MiB) is held by every GPU the library runs on until the process ends.
 * All calls on one GPU share that array: calls that overlap on one GPU (from
 * several host threads) race. */
#define SG_NOALLOC_MAX_POINTS (1 << 19)
int sg_derivatives_noalloc(sg_handle *h, const double *f, double *out);

/* Wait for all work of the handle; returns the first pending error. */
int sg_sync(sg_handle *h);

/* Last error: code, failing point (1-based i,j,k or 0) and message. */
int sg_last_error(sg_handle *h, int *code, int ijk[3], char *msg, size_t n);

/* The HIP stream the handle launches on (hipStream_t), for callers that
 * record events around sg_substage / sg_step. */
int sg_get_stream(sg_handle *h, void **stream);

/* Per-kernel timing with HIP events on the handle's stream.  on=1 starts
 * (and clears) accumulation, on=0 stops.  sg_profile_read returns kernel
 * number idx (0-based; SG_EINVAL past the end): its name, summed time in ms
 * and launch count. */
int sg_profile(sg_handle *h, int on);
int sg_profile_read(sg_handle *h, int idx, char *name, size_t n, double *total_ms, int64_t *count);

/* Device bytes held by the handle. */
int sg_memory(sg_handle *h, size_t *bytes);

/* *compiled = 1 when the configuration's chemistry is, bit for bit, the
 * mechanism compiled into the library (the 9-species, 21-step H2-air
 * mechanism of the reference's input/chem.dat): the point kernels then run
 * the CHRATE step loop (src/chrate.f:146-760) unrolled over its constants
 * instead of reading the step tables at run time.  Both evaluate the same
 * arithmetic; any other mechanism takes the table-driven loop.  No
 * counterpart in the reference, which reads chem.dat at run time. */
int sg_mechanism_compiled(const sg_handle *h, int *compiled);

/* ---- Tiles and their autotuner (no counterpart in the reference) ---------
 * narrow[d] selects the tile of the direction-d derivative pass: 0 = rows of
 * 32 points (default), 1 = rows of 16 points and twice the rows (no idle
 * lanes when n_x is an odd multiple of 16, a shorter halo along y and z).
 * No choice changes any result.  sg_autotune times each choice on the
 * handle's current state (reps launches each, 5 if reps <= 0), keeps the
 * fastest, and returns in *ms the tuned passes' time per substage; the state
 * it leaves is the one it found, so tuned and untuned runs compute identical
 * results. */
int sg_set_tiles(sg_handle *h, const int narrow[3]);
int sg_get_tiles(sg_handle *h, int narrow[3]);
int sg_autotune(sg_handle *h, int reps, double *ms);

/* Device bytes sg_create would allocate for cfg (without allocating), for
 * admission by memory budget; SG_EINVAL for a configuration sg_create
 * rejects.  No device access. */
int sg_footprint(const sg_config *cfg, size_t *bytes);

/* ---- Batches: many cases advanced together ------------------------------
 * No counterpart in the reference, which runs one case per job; this is the
 * batch variant of the src/senga2.f:96-194 step loop.  A batch advances its
 * member handles in lockstep: every substage kernel is ONE launch covering
 * all members (their grids concatenated), and one full RK step (BOUNDT and
 * the five substages of src/senga2.f:113-159, then BOUNTT) is captured once
 * as a HIP graph and replayed.  Each member keeps its own grid size, state,
 * boundary data and fixed time step, and computes exactly what it computes
 * alone.  Members share the species count, precision, x boundary kind and
 * mechanism (chem.dat / diff.dat data), else sg_batch_add returns
 * SG_EINVAL.  Adaptive stepping is not available inside a batch.
 *
 * While a handle is in a batch, sg_upload_state, sg_substage, sg_end_step,
 * sg_step and sg_rhs return SG_ESTATE; sg_download_state and sg_sync order
 * themselves after the batch steps queued so far; sg_set_timestep applies
 * from the next sg_batch_step; sg_destroy removes the handle first.  One host
 * thread drives a batch and its members. */
typedef struct sg_batch sg_batch;
int sg_batch_create(sg_batch **out);
/* Detaches the members (they stay valid) and frees the batch. */
int sg_batch_destroy(sg_batch *b);
/* Adds a handle with uploaded state; its queued work completes first. */
int sg_batch_add(sg_batch *b, sg_handle *h);
/* Removes a handle; its later calls run after the queued batch steps. */
int sg_batch_remove(sg_batch *b, sg_handle *h);
int sg_batch_size(sg_batch *b, int *n);
/* Member idx (0-based, in order of addition among current members). */
int sg_batch_member(sg_batch *b, int idx, sg_handle **h);
/* Queue nsteps RK steps of every member (asynchronous). */
int sg_batch_step(sg_batch *b, int nsteps);
/* Wait for the queued steps; on a temperature failure returns its code and
 * the failing member's index in *failed (else -1), whose sg_last_error holds
 * the point. */
int sg_batch_sync(sg_batch *b, int *failed);
/* Pass tiles of the batch's launches (one choice for all members), set or
 * tuned on the members' current states as sg_autotune does for a handle. */
int sg_batch_set_tiles(sg_batch *b, const int narrow[3]);
int sg_batch_get_tiles(sg_batch *b, int narrow[3]);
int sg_batch_autotune(sg_batch *b, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* SENGA_GPU_H */
