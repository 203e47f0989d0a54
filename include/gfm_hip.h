/*
 * gfm_hip.h -- C-ABI of the MI355X combustion-space SIMPLER solver.
 *
 * The reference solver (GFM v4, combustion/Code) keeps all state in the
 * Fortran module GBL_VAR and runs the combustion-space iteration through
 * argument-less subroutines:
 *
 *   SUBROUTINE SPHASE                       C/Sphase.f90:10   (called C/A0main.f90:204)
 *   SUBROUTINE GFLOW                        C/Gflow.f90:6     (called C/Sphase.f90:51)
 *   SUBROUTINE ADLBL3(XT,IB,IE,JB,JE,KB,KE) C/Adlbl3.f90:10   (called C/Gslv3v.f90:93,133,163, C/Gslv3g.f90:96)
 *   subroutine calc_residual(xt,resid,...)  C/Adlbl3.f90:239
 *
 * (C/ = combustion/Code/ of the reference package.)
 *
 * This header is the drop-in boundary for those subroutines.  The host
 * (Fortran through ISO_C_BINDING, or any C-callable language) owns every
 * array in the reference layout: 1-based, column-major, i fastest, on the
 * doubled-index lattice (MP,NP,LP) with MP = 2*nx+4 (cell centres at even
 * indices, faces at odd indices).  The library copies them into
 * device-resident compact arrays at gfm_hip_init / gfm_hip_upload and back
 * at gfm_hip_download.  All pointers are plain host pointers; no
 * language-runtime types cross the boundary.
 *
 * Error convention: every entry point returns 0 on success and a non-zero
 * GFM_E_* code on failure; gfm_hip_last_error() returns a message.  The
 * reference has no return codes (fatal paths call stop_run, C/A0main.f90:415);
 * the Fortran wrapper maps a non-zero return to that call.
 */
#ifndef GFM_HIP_H
#define GFM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFM_ABI_VERSION 6

enum {
  GFM_OK = 0,
  GFM_E_ARG = 1,      /* bad argument / inconsistent dimensions            */
  GFM_E_DEVICE = 2,   /* HIP runtime error (allocation, launch, copy)      */
  GFM_E_STATE = 3,    /* call out of order (e.g. sphase before init)       */
  GFM_E_NUMERIC = 4   /* non-finite value detected in a reduction          */
};

/* Line-relaxation plane order used by ADLBL3.
 *   GFM_ORDER_REFERENCE: x-planes visited forward then backward each sweep,
 *     each plane using the latest x-neighbours (C/Adlbl3.f90:39-51).
 *   GFM_ORDER_ZEBRA: each sweep updates even-numbered planes, then odd,
 *     then even, then odd (every plane still visited twice per sweep) with
 *     the reference's per-plane operator ADLZ unchanged.  Planes of one
 *     colour are independent, so they are solved concurrently. */
enum { GFM_ORDER_REFERENCE = 0, GFM_ORDER_ZEBRA = 1 };

/* Arithmetic of the assembly and line sweeps.  The p' defect, the global
 * balances/reductions and the enthalpy->temperature inversion are always
 * FP64. */
enum { GFM_PREC_FP64 = 0, GFM_PREC_MIXED = 1 };

/* p' solve (C/Gslv3v.f90:149-164).
 *   GFM_PCORR_REFERENCE: zero initial guess, NTIMESG(5) sweeps (the reference).
 *   GFM_PCORR_TOLERANCE: start from the initial guess of pcorr_guess, then
 *     relax one sweep at a time until calc_residual <= pcorr_tol or
 *     pcorr_max_sweeps sweeps.
 * Initial guess in GFM_PCORR_TOLERANCE:
 *   GFM_PGUESS_ZERO: FZ = 0.
 *   GFM_PGUESS_LEARNED: the network attached by gfm_hip_set_pguess maps the
 *     Jacobi-scaled mass source BS/AP to a p' field g; the guess is alpha*g
 *     with alpha minimising ||BS - alpha*A g||_2 (FP64).  FZ = 0 is kept
 *     instead unless calc_residual(alpha*g) < a * calc_residual(0), with
 *     a = pcorr_guess_accept (1 when <= 0: the guess must lower the
 *     residual; smaller values demand a margin, which guards a network
 *     applied outside the states it was trained on).
 * With pcorr_block_corr = 1 each sweep of GFM_PCORR_TOLERANCE is preceded by
 * an additive correction over x-planes (block correction: Settari and Aziz
 * 1973; Patankar 1981, sec. 5.4): one value per x-plane of the box, added
 * to the plane's unpinned cells, chosen so that the residuals of each
 * plane's unpinned cells sum to zero (a tridiagonal system over the planes,
 * FP64), for every plane but the last: the last plane is the reference
 * (value 0), because the p' equation, with no coefficient toward a pinned
 * cell, is singular.  Pinned cells (IBCELL >= 1 or |AP| <= 1e-20) and cells
 * outside the box keep their values. */
enum { GFM_PCORR_REFERENCE = 0, GFM_PCORR_TOLERANCE = 1 };
enum { GFM_PGUESS_ZERO = 0, GFM_PGUESS_LEARNED = 1 };

/* Host arrays of one case, reference layout (see top of file).  Arrays of
 * rank 3 have MP*NP*LP entries; UG holds 3 and GF holds 9 consecutive
 * such blocks (GF(:,:,:,1..9) = fuel, O2, h, k, eps, CO2, N2, -, H2O;
 * C/Var_mod.f90:116-125). */
typedef struct gfm_host_arrays {
  double *x, *dx;        /* (MP)  C/Grid3.f90:64  */
  double *r, *dr, *rs;   /* (NP)                  */
  double *z, *dz;        /* (LP)                  */
  int16_t *ibcell;       /* (MP,NP,LP) cell codes after 26-neighbour propagation, C/Grid3.f90:259-275 */
  double *p, *t, *dnst, *theta, *tmu, *gcp, *glam, *gama, *gte, *crate, *hrate;
  double *bs_hprev;      /* enthalpy-source history, C/Source.f90:511-520 */
  double *ug;            /* (MP,NP,LP,3) */
  double *gf;            /* (MP,NP,LP,9) */
  double *rtim, *delrt, *hrt, *delr, *timed, *hrd;   /* (NR0) kinetic tables, C/Setup.f90:858-888 */
} gfm_host_arrays;

/* Scalars of GBL_VAR read or written by the path (C/Var_mod.f90).  All
 * nondimensional as after SETUP (C/Setup.f90:542-731). */
typedef struct gfm_ctrl {
  /* dimensions */
  int32_t mp, np, lp, nr0;
  int32_t ncells;                 /* open cells, C/Grid3.f90:259-267 */
  int32_t i_ref, j_ref, k_ref;    /* pressure reference (inlet) cell */
  /* options */
  int32_t sweep_order;            /* GFM_ORDER_* */
  int32_t precision;              /* GFM_PREC_* */
  int32_t want_pre_residuals;     /* compute calc_residual before each solve (log only) */
  int32_t want_irc;               /* IRC residual block of ADLBL3 (RESG) */
  int32_t i1st_rad_done;          /* 0: adiabatic walls, C/Extr.f90:194 */
  int32_t isothermal_walls;       /* 1: wall heat sink of C/Source.f90:473-509 with fixed wall T */
  int32_t gray_radiation;         /* GFM_RADSRC_*: radiation source of the enthalpy equation */
  int32_t pcorr_block_corr;       /* 1: x-plane additive correction before each sweep of a p' solve
                                     relaxed to a tolerance (GFM_PCORR_TOLERANCE; not in the reference) */
  /* iteration state */
  int32_t itr_gas, itr_gend, maxgi;
  int32_t want_plane_flows;       /* form FLX1D's per-plane flow rates each iteration (gfm_hip_plane_flows) */
  double bgcon;
  double factor, fadj;
  /* reference scales and nondimensional groups */
  double t0, trdg, h_0, r0, area0, vol0;
  double reyg, reyg2, eun, epsr, taur, ug0, dnst0, qr0, q0, gmfr, tk0, epsi0;
  double gra[3];
  double p_ref;
  double t_mx, t_mn, u_mx, u_mn, p_mx, p_mn, f_mx, f_mn;
  double rf[21], rfl[21], rf_hsource;
  int32_t ntimesg[20];
  double gfin0;                 /* inlet mass flow / GMFR, C/Sbc.f90:375, C/Extr.f90:62 */
  /* wall and radiation model (reference extension; used when enabled) */
  double t_wall, h_g, kappa_gray, sigma_sb;
  /* outputs of the last iteration */
  double res_mass, res_mass_max;
  int32_t imx, jmx, kmx, pad2;
  double p_resid_pre, p_resid, pcorr_resid, h_resid_pre, h_resid;
  double u_resid_pre[3], u_resid[3];
  double gf_resid_pre[10], gf_resid[10];
  double resg[40];                /* RESG(20,2) column-major */
  double gfex, gfin, fuel_out, tmean_ex, q_react, q_exhaust, q_wall;
  double tave, tave_vol, tmax_c;
  double t_av_exit, t_x_exit;     /* C/Sphase.f90:74-89 */
  /* p' solve options (GFM_PCORR_*, GFM_PGUESS_*) and outputs */
  int32_t pcorr_mode, pcorr_guess, pcorr_max_sweeps;
  int32_t pcorr_dump;             /* keep (BS/AP, mask, p') of the last solve for gfm_hip_pcorr_pair */
  int32_t pcorr_sweeps;           /* out: sweeps of the last p' solve */
  int32_t pcorr_fallback;         /* out: 1 if the learned guess was replaced by zero */
  double pcorr_tol;
  double pcorr_resid_start;       /* out: calc_residual of the initial guess */
  double pcorr_alpha;             /* out: scale of the learned guess */
  double pcorr_gain;              /* out: fraction of ||BS||^2 the scaled learned guess removes */
  double pcorr_guess_accept;      /* in: acceptance ratio of the learned guess (see above) */
  double q_wall_bottom, q_wall_top; /* out: the parts of q_wall through cells' bottom (k-) and top (k+) faces:
                                       glass/floor and crown */
  /* half-blocked cells (IBCELL -1 and -2, grid-file digits 9 and 8): the
   * &GRD factors BV_F(1:2) of their volume and BA_F(1:3,1:2) of their face
   * areas (column-major, BA_F(ix,n) = ba_f[(ix-1) + 3*(n-1)]), read by
   * GRID3 (C/Grid3.f90:12-13,252-255) and applied by VOL_C0, AREA_C and VOL1
   * (C/Grid3.f90:389-471, C/Gslv3x.f90:269-286).  1 where no such cells. */
  double bv_f[2], ba_f[6];
} gfm_ctrl;

/* Opaque device-side context. */
typedef struct gfm_ctx gfm_ctx;

/* Create a device context for one case on `device` and upload the host
 * state.  Replaces nothing in the reference: it is the allocation that
 * allocate_large_gas_phase_arrays (C/Setup.f90:740) does for the host. */
int gfm_hip_init(const gfm_ctrl *ctrl, const gfm_host_arrays *h, int device, gfm_ctx **out);

/* Copy host state to the device (after a host-side change such as a
 * restart read, C/Initsv.f90:225) / device state back to the host (before
 * any host consumer: sav3f C/A0main.f90:210, sprint :257). */
int gfm_hip_upload(gfm_ctx *ctx, const gfm_ctrl *ctrl, const gfm_host_arrays *h);
int gfm_hip_download(gfm_ctx *ctx, gfm_ctrl *ctrl, gfm_host_arrays *h);

/* The tail of INITSV (C/Initsv.f90:63-84): INTPV, EXTR(1), INTP on the
 * device state. */
int gfm_hip_initsv_tail(gfm_ctx *ctx, gfm_ctrl *ctrl);

/* One GFLOW call (C/Gflow.f90:6-408): one SIMPLER outer iteration.  Does
 * not advance itr_gas; ctrl carries itr_gas in and the iteration's
 * residuals and balances out. */
int gfm_hip_gflow(gfm_ctx *ctx, gfm_ctrl *ctrl);

/* SPHASE (C/Sphase.f90:37-235): iterate GFLOW while itr_gas < itr_gend and
 * itr_gas < maxgi, stopping after `max_iters` iterations (the host's
 * polling cadence, C/Sphase.f90:106-162) or when res_mass < bgcon
 * (C/Sphase.f90:193).  If res_hist is non-null it receives res_mass of
 * every iteration run (max_iters entries).  Returns the iterations run in
 * *n_done. */
int gfm_hip_sphase(gfm_ctx *ctx, gfm_ctrl *ctrl, int max_iters, double *res_hist, int *n_done);

/* Autotuner of the launch shapes (thread mapping: one x row of the box per
 * row of workgroups, or the (i, j) rows of a plane flattened; workgroups of
 * 64, 128 or 256 threads) of the per-position box kernels (assembly,
 * interpolation, properties, copies): `iters` SIMPLER iterations per
 * candidate shape, each kernel timed with events, then each keeps its
 * fastest shape for the rest of the context's life (the momentum, pressure
 * and scalar assemblies always take the row mapping and are tuned over the
 * workgroup size only).  No counterpart in the reference.  The iterations are ordinary iterations of the run (itr_gas
 * advances by 6*iters; ctrl carries the last one's outputs): every shape
 * computes the same values bit for bit.  A table of the choices (kernel,
 * default and best microseconds per launch, mapping, workgroup size) is
 * written to report (may be NULL). */
int gfm_hip_autotune(gfm_ctx *ctx, gfm_ctrl *ctrl, int iters, char *report, int report_len);

/* ADLBL3 on a host-provided system (C/Adlbl3.f90:10-97): AP, AS(…,3,2),
 * BS and XT on the doubled lattice, NTIMESG(nelg) sweeps over the index
 * box [ib..ie]x[jb..je]x[kb..ke] (stride 2, 1-based; 3 <= begin <= end <=
 * M-2 on each axis, else GFM_E_ARG).  XT is updated in place;
 * RESG(nelg,1:2) is returned in resg[0..1] when want_irc is set.  In a
 * GFM_PREC_MIXED context the coefficients are rounded to FP32 on the
 * device and the sweeps run in FP32, as inside gfm_hip_gflow. */
int gfm_hip_adlbl3(gfm_ctx *ctx, const gfm_ctrl *ctrl, int nelg,
                   const double *ap, const double *as, const double *bs, double *xt,
                   int ib, int ie, int jb, int je, int kb, int ke, double *resg);

/* calc_residual (C/Adlbl3.f90:239-271) on a host-provided system (box as
 * for gfm_hip_adlbl3; FP32-rounded coefficients in a mixed context). */
int gfm_hip_calc_residual(gfm_ctx *ctx, const double *ap, const double *as, const double *bs,
                          const double *xt, int ib, int ie, int jb, int je, int kb, int ke,
                          double *resid);

/* calc_residual, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  A
 * variant of gfm_hip_calc_residual that needs no context: the system is
 * already in device memory in the reference layout (column-major doubled
 * lattice (mp,np,lp), FP64; as holds the 6 blocks AS(:,:,:,1..3,1..2);
 * ibcell int16), ncells is NCELLS, the box is as for gfm_hip_adlbl3, and
 * the result (one double) is written to the device pointer resid.  The
 * two kernels are enqueued on `stream` (a hipStream_t; NULL = the default
 * stream) and the call returns without synchronising.  Partial sums (one
 * per (J,K) row of the box) go to a static device array of
 * GFM_NOALLOC_MAX_ROWS doubles (512 KB), one per GPU, held until the
 * process ends; a box with more rows is rejected with GFM_E_ARG.  Calls
 * that overlap on one GPU share that array and race: serialise them (one
 * stream, or events).  FP64 always (a mixed-precision context's
 * gfm_hip_calc_residual rounds the coefficients to FP32; this does not). */
#define GFM_NOALLOC_MAX_ROWS 65536
int gfm_hip_calc_residual_noalloc(int mp, int np, int lp, int ncells, const int16_t *ibcell, const double *ap,
                                  const double *as, const double *bs, const double *xt, int ib, int ie, int jb, int je,
                                  int kb, int ke, double *resid, void *stream);

/* Attach the network of include/gfm_pnet.h for GFM_PGUESS_LEARNED (NULL
 * detaches).  The case's cell grid (nx, ny, nz of the p' box) must be
 * multiples of 8 and fit the network's buffers; the network must live on
 * the context's device and is not owned by the context.  No counterpart in
 * the reference (its guess is zero, C/Gslv3v.f90:157-161). */
struct gfm_pnet;
int gfm_hip_set_pguess(gfm_ctx *ctx, struct gfm_pnet *net);

/* The training pair of the last p' solve run with pcorr_dump set (the pair
 * of row "p' solve": BS after SOURCE(5) and FZ after the solve): for each
 * cell of the p' box, v = ((k-4)/2*ny + (j-4)/2)*nx + (i-4)/2,
 * x[2v] = BS/AP (0 where the row is pinned), x[2v+1] = 1 for open rows else
 * 0, p[v] = p'. */
int gfm_hip_pcorr_pair(gfm_ctx *ctx, float *x, float *p);

/* FLX1D's cross-sectional flow rates (C/Flx1d.f90:16-73; formed after
 * EXTRV(0) of the velocity update, C/Gslv3v.f90:210-211, and read by the
 * printout C/Sprint.f90:219-228) of the last iteration run with
 * want_plane_flows set: gas mass flow (kg/s) through each x-face plane
 * I = 3..MP-1 in +x and -x, flx(MP,2); the same through the y-face planes,
 * fly(NP,2), and z-face planes, flux_z(LP,2); the enthalpy flows (W) in y
 * and z, flyh(NP,2), fluxh_z(LP,2).  Column-major with 1-based plane index
 * as in the reference (entries at even indices are 0).  Any pointer may be
 * NULL.  GFM_E_STATE if no iteration has formed them. */
int gfm_hip_plane_flows(gfm_ctx *ctx, double *flx, double *fly, double *flyh, double *flux_z, double *fluxh_z);

/* ---------------------------------------------------------------------
 * Minor species and soot (MS=1 runs): the sub-species solve that the
 * reference runs on the converged gas field,
 *
 *   call initsvm; call msflow; call sav3fm      C/A0main.f90:219-222, C/Sphase.f90:216-218
 *   SUBROUTINE INITSVM                           C/Initsv.f90:329-380
 *   SUBROUTINE MSFLOW                            C/Msflow.f90:10-169
 *   SUBROUTINE GSLV3M                            C/Gslv3m.f90:8-286 (SPIN :292-305)
 *   SUBROUTINE SOURCEM(MI1,MI3,NJ1,NJ3,LK1,LK3)  C/Sourcem.f90:8-122 (new soot model)
 *   subroutine flx1dm / extrm / intpm            C/Flx1d.f90:114-158, C/Extr.f90:258-305, C/Intp.f90:348-419
 *
 * Seven lumped species GFM(:,:,:,1..7) = CH4, O2, CO, H2O, CO2, N2, NO and
 * the soot mass fraction smf, transported in the gas field the context
 * holds, with the five-reaction kinetics of GSLV3M (rates RRMS(MZ,NZ,LZ,
 * 1..5), MZ = MP/2 ..., of which 1, 2 and 4 are formed) and the soot
 * formation/oxidation source of SOURCEM.  The species equations reuse the
 * gas solver's coefficient arrays: the rates of the first GSLV3M after
 * INITSVM read the coefficients AS that the last GFLOW left (its last
 * scalar equation), as the reference's GBL_VAR AS does; the species
 * coefficients (SSIGMA) are assembled once after INITSVM and kept (AP1).
 * Device memory for these fields is allocated at the first call of any
 * gfm_hip_ms_* entry point. */
#define GFM_NSP0 7
#define GFM_NRM0 5
typedef struct gfm_ms_ctrl {
  /* inputs (C/Setup.f90:161-185, the relaxation file C/Setup.f90:478-493) */
  int32_t maxms;              /* MAXMS: GSLV3M calls per MSFLOW */
  int32_t react;              /* REACT: solve the soot equation */
  int32_t ntimesm;            /* NTIMESM: ADLBL3 sweeps of every species and of soot */
  int32_t itr_ms;             /* in/out: minor-species iteration count (rs record) */
  int32_t igs0;               /* in/out: 0 = assemble the species coefficients at the next GSLV3M */
  int32_t pad0;
  double bmcon;               /* BMCON: every resid_ms below it ends MSFLOW (after 10 iterations) */
  double ssigma;              /* SSIGMA: Schmidt number of the species equations */
  double rfm[8];              /* RFM(1:8): relaxation of species 1..7 and soot (RFMC = 1 - RFM) */
  double aform, aoxid;        /* soot formation / oxidation pre-exponential constants */
  double esf_r, eso_r;        /* Esf/Ru, Eso/Ru (K) */
  double t0_sf, t0_so;        /* soot formation / oxidation cut-off temperatures (K) */
  double rf_soot_form_source, rf_soot_oxid_source;  /* relaxation of the soot sources */
  double soot_cap;            /* largest soot mass fraction kept */
  double dnst_soot;           /* soot density (kg/m3) */
  double vol_tot;             /* VOL_TOT: volume of the open cells (m3), C/Grid3.f90:286-294 */
  /* outputs of the last GSLV3M / MSFLOW */
  int32_t iterations;         /* GSLV3M calls of the last gfm_hip_msflow */
  int32_t converged;          /* 1 if that call met MSFLOW's convergence test */
  int32_t smf_clipped_cnt;    /* cells clipped at soot_cap */
  int32_t pad1;
  double resid_ms_pre[8], resid_ms[8];  /* calc_residual before / after each solve (8 = soot) */
  double bal_ms;              /* volume-weighted |1 - sum GFM| / VOL_TOT */
  double soot_tot;            /* soot mass (kg) */
  double smf_max, smf_max_noclip;
  double rr_f, rr_c;          /* sum of the CH4 and CO rates times their molar masses */
  double gfm_0[7];            /* species flow through the exit planes, GFM_0(:,2) (kg/s) */
  double gfmex;               /* sum of gfm_0 */
  double ave_chg;             /* mean |change of gfm_0| over the last iteration */
  double svf_r;               /* area-mean soot volume fraction of the exit cells */
} gfm_ms_ctrl;

/* Host arrays of the minor-species state, reference layout:
 * gfm (MP,NP,LP,7), smf, bs_soot_prev, sfp_soot_prev (MP,NP,LP),
 * rrms (MP/2,NP/2,LP/2,5). */
typedef struct gfm_ms_arrays {
  double *gfm, *smf, *bs_soot_prev, *sfp_soot_prev, *rrms;
} gfm_ms_arrays;

/* Copy the minor-species state to / from the device (the arrays sav3fm
 * writes and INITSVM reads, C/Sav3f.f90:96-116, C/Initsv.f90:352-372; the
 * counters itr_ms and igs0 travel in ms). */
int gfm_hip_ms_upload(gfm_ctx *ctx, const gfm_ms_ctrl *ms, const gfm_ms_arrays *a);
int gfm_hip_ms_download(gfm_ctx *ctx, gfm_ms_ctrl *ms, gfm_ms_arrays *a);

/* INITSVM (C/Initsv.f90:329-380) on the device.  restart = 0: a new start
 * (GFM = 0 and smf = 0 everywhere, then the cell centres take CH4, O2,
 * H2O, CO2 and N2 from the gas field; the rates, the soot-source history
 * and itr_ms are kept).  restart = 1: the state of the last
 * gfm_hip_ms_upload is the restart record; inlet cells (IBCELL 2) take
 * the five from the gas field and CO = NO = 0, then extrm.  Both end with
 * intpm and igs0 = 0. */
int gfm_hip_initsvm(gfm_ctx *ctx, gfm_ms_ctrl *ms, int restart);

/* One GSLV3M call (C/Gslv3m.f90:8-286): rates, the seven species
 * equations, the balance, and (react) the soot equation.  Does not
 * advance itr_ms. */
int gfm_hip_gslv3m(gfm_ctx *ctx, gfm_ms_ctrl *ms);

/* MSFLOW (C/Msflow.f90:29-162): up to maxms iterations of itr_ms += 1,
 * GSLV3M, flx1dm, the exit soot fraction, the convergence test (itr_ms >
 * 10 and every resid_ms <= bmcon) and extrm + intpm. */
int gfm_hip_msflow(gfm_ctx *ctx, gfm_ms_ctrl *ms);

/* ---------------------------------------------------------------------
 * Radiation transport (ID_RAD > 0): one full radiation pass of RAD_EMIS
 * (C/Rad_emis.f90:130-229) up to the wall exchange.  The walls with an open
 * neighbour are marked by their normal (C/Rad_emis.f90:155-186); RAD0
 * (C/Rad0.f90:12-172) forms the spectral absorption coefficients AKL of
 * every open cell from H2O, CO2 and soot (RAD_GAS/AH2O/ACO2/A1/CHIWHI/
 * BANDA C/Rad_gas.f90:16-396, SOOTA/SOC C/Rad_soot.f90:26-69) and the
 * band fractions of its emission (DEBX/RIF C/Rad_fun.f90:45-140); the
 * volume loop of QRNF (C/Qrn.f90:58-200) traces every open emitter above
 * 300 K through the enclosure (QREF C/Qre.f90:31-505 with QRTX, QRTY,
 * QRTZ): its band powers leave the six faces, cross the domain slab by
 * slab, spread across each slab by the angles the slab's far face
 * subtends, are absorbed in open cells in proportion to AKL and deposited
 * whole in boundary cells.  QRN1 (C/Qrn.f90:245-290) sums QAL over the
 * bands into QA.  The emission-only pass rad_qe_only (C/Rad_emis.f90:
 * 411-466) is RAD0 with ID_RAD = 1: QE of every open cell, QA kept.
 * The gas state, the minor species H2O and CO2 and the soot fraction are
 * the context's (gfm_hip_ms_upload / gfm_hip_initsvm before the first
 * call).  QE and QA stay on the device for the enthalpy source
 * (gray_radiation = GFM_RADSRC_FIELD: C/Source.f90:437-447).
 * ------------------------------------------------------------------- */
#define GFM_NWL_MAX 16
enum {
  GFM_RADSRC_NONE = 0,   /* gfm_ctrl.gray_radiation: no radiation source */
  GFM_RADSRC_GRAY = 1,   /* the gray closure (reference extension) */
  GFM_RADSRC_FIELD = 2   /* QA - QE of the last radiation pass, C/Source.f90:437-447 */
};
typedef struct gfm_rad_ctrl {
  /* inputs: &RADIN and the band list of the sbc file (C/Sbc.f90:316-355) */
  int32_t nwl;                 /* NWL: spectral bands (1..GFM_NWL_MAX) */
  int32_t have_crown;          /* 1: a crown (walls above height_to_ceiling take eps_c) */
  double wl[GFM_NWL_MAX];      /* WL: band wavelengths (m), the listed microns times 1e-6 */
  double ri;                   /* RI: refraction index */
  double eps_w, eps_m, eps_c;  /* wall, melt-surface and crown emissivity (their AKL) */
  double height_to_ceiling;    /* z (m) from which walls take eps_c */
  double dnst_soot;            /* soot density (kg/m3), C/Var_mod.f90:137 */
  /* outputs */
  int32_t l_ab;                /* open cells counted by QRNF */
  int32_t n_emit;              /* emitters traced (open, T > 300 K) */
  int32_t carry_unset;         /* wall splits on an emitter's own column whose column the reference would
                                  carry over from an earlier emitter (not reproduced: 0 expected) */
  int32_t angle_oob;           /* angle reads past the lattice (boundary cells that are open; 0 expected) */
  double tot_qe_h2o_co2;       /* RAD0: gas emission (W) */
  double tot_qe_soot;          /* RAD0: soot emission (W) */
  double qe_sum_from_open;     /* QRNF: emission of the traced emitters (W) */
  double qa_sum_to_melt, qa_sum_to_wall, qa_sum_to_open, qa_sum_to_exit, qa_sum_to_inlet;  /* (W) */
  double qeg0;                 /* QE of the open cells (W); emission-only pass: the same from RAD0 */
  double qeg_g, qeg_w, qeg_s, qeg_io;  /* QRN1: QA of open, wall, melt-surface, inlet and exit cells (W) */
} gfm_rad_ctrl;

/* Host arrays of the radiation fields, reference layout (cell centres):
 * qe, qa (MP/2,NP/2,LP/2) in W; qal, akl (MP/2,NP/2,LP/2,nwl). */
typedef struct gfm_rad_arrays {
  double *qe, *qa, *qal, *akl;
} gfm_rad_arrays;

/* The full pass (wall marks, RAD0 with ID_RAD = 10, QRNF's volume loop,
 * QRN1).  Outputs in rc; QE, QA, QAL and AKL stay on the device. */
int gfm_hip_rad_emis(gfm_ctx *ctx, gfm_rad_ctrl *rc);
/* rad_qe_only: QE of every open cell from RAD0 with ID_RAD = 1 (qeg0). */
int gfm_hip_rad_qe_only(gfm_ctx *ctx, gfm_rad_ctrl *rc);
/* Copy the device's QE, QA, QAL and AKL to the host (any pointer may be
 * null) / QE and QA from the host (a radiation record, rr####c.d). */
int gfm_hip_rad_download(gfm_ctx *ctx, const gfm_rad_ctrl *rc, gfm_rad_arrays *a);
int gfm_hip_rad_upload(gfm_ctx *ctx, const gfm_rad_arrays *a);
size_t gfm_hip_rad_ctrl_size(void);

/* ---------------------------------------------------------------------
 * Wall radiosity exchange (C/Qwall.f90), the enclosure part of a radiation
 * pass that QRNF runs after its volume loop (C/Qrn.f90:218-221).
 *
 * gfm_hip_wall_setup is prepare_for_wall_radiation_calculation
 * (C/Qwall.f90:54-566, called by INITSV on a new start): the cell codes
 * with each exit tunnel turned into wall up to an exit at its interior end
 * (MAPB), one boundary patch per face between a wall, inlet, exit or melt
 * cell and an open cell -- x-normal faces first, then y, then z, each in
 * (K, J, I) loop order -- with its type, position, normal, the centre of
 * its boundary cell, area and emissivity, its exhaust, the normalised view
 * factors VF(n,m) (preset_vf = 1; kept on the device), the area of all
 * patches on its boundary cell and the totals.  iwall (MP/2,NP/2,LP/2, the
 * wall group of each cell, SBC_B C/Sbc.f90:402-456) may be null.
 *
 * gfm_hip_wallq is one radiation pass's wall part: the patch temperatures
 * and radiosities at the first pass (C/Rad_emis.f90:190-212; later passes
 * refresh the melt patches, :253-261), wall_q_amb (C/Qwall.f90:1585-1659),
 * wallq (C/Qwall.f90:990-1369: qc from QA of the last gfm_hip_rad_emis,
 * the outer loop of calc_q_conv_wall, Gauss-Seidel sweeps over the patches
 * in patch order to the radiosity tolerance, wall temperatures and the melt
 * heat flux qrs, to the melt-flux tolerance; the closing sums, the mean
 * wall temperature, exhausts at a fraction of it, the soot calibration),
 * and QRN_PRN's QA_S0/QA_S1 (C/Qrn.f90:399-402).  Then the state the next
 * gas iterations see (C/Rad_emis.f90:337-355, rad_emis_rr :101-113): the
 * wall temperature at every wall patch face and at its wall cell (with its
 * enthalpy), and each wall and melt patch's q_conv as the enthalpy sink of
 * its open cell (C/Source.f90:473-509), which every later SIMPLER
 * iteration applies while gfm_ctrl.i1st_rad_done is 1 (set it after the
 * first call).  The Gauss-Seidel sweep runs in one workgroup; each row's
 * sum over the other patches is split into the part outside a 64-patch
 * block (all waves) and the block's own part (in order), so sweeps follow
 * the reference's order and differ from it only in the rounding of those
 * sums.
 * ------------------------------------------------------------------- */
#define GFM_NEX_MAX 32     /* exhausts */
#define GFM_NWALG_MAX 64   /* wall property groups */
typedef struct gfm_wall_ctrl {
  /* inputs: &RADIN (C/Sbc.f90:316-330), exhausts (:272-284), wall groups (:288-311) */
  int32_t maxri1, maxri2, minri2, preset_vf;
  int32_t have_crown, nex, nwal0, isoot_cal;
  double eps_w, eps_m, eps_c, height_to_ceiling;
  double t_a, h_a, w_d, w_k, h_g;          /* default wall: ambient T (K), outside coefficient, thickness,
                                              conductivity; inside coefficient h_g (W/m2K) */
  double wall_radiosity_tolerance;         /* 1e-10, C/Qwall.f90:1036 */
  double radiation_boundary_tolerance;     /* 1e-10, C/Qwall.f90:1037 */
  int32_t exh_box[GFM_NEX_MAX][6];         /* exh_i1, j1, k1, i2, j2, k2 */
  int32_t exh_orient[GFM_NEX_MAX], exh_direct[GFM_NEX_MAX], exh_type[GFM_NEX_MAX];
  double exh_frac[GFM_NEX_MAX], exh_fixed[GFM_NEX_MAX];
  double w1_d[GFM_NWALG_MAX], w1_k[GFM_NWALG_MAX], w1_ha[GFM_NWALG_MAX], w1_ta[GFM_NWALG_MAX];
  double q_melt_req, rf_aform, rf_afover;  /* soot calibration (isoot_cal = 1) */
  /* carried between passes (in and out) */
  int32_t i1st_rad_done;                   /* 0: initialise b_T and radiosity from the gas state */
  int32_t iterw;                           /* Gauss-Seidel sweeps so far */
  double avg_wall_T;                       /* mean wall temperature (K); 1600 before the first pass */
  double aform;                            /* soot formation constant (adjusted by the soot calibration) */
  double qa_s0, qa_s1;                     /* QRN_PRN: net radiation to the melt (MW) and its change */
  /* outputs of gfm_hip_wall_setup */
  int32_t b_cnt, iwall_count, imelt_count, inlet_count, iexit_count;
  int32_t pad0;
  double wall_area_total, melt_area_total, inlet_area_total, exit_area_total, boundary_area_total;
  /* outputs of gfm_hip_wallq */
  int32_t sweeps, outer;                   /* Gauss-Seidel sweeps and outer iterations of this call */
  double q_incident_total, q_incident_wall, q_incident_inlet, q_incident_exit, q_incident_melt, qc_avg;
  double qconv_total, qconv_avg, q_wall_loss, qamb_avg, q_melt, qrs_avg;
  double q_rad_inlet, q_rad_outlet, qm_flame, qm_bnds, qm_conv;
  double rel_rad_chg_last;                 /* mean relative radiosity change of the last sweep */
} gfm_wall_ctrl;

/* Host arrays of the patches (patch order) and the lattice fields of the
 * wall exchange.  Any pointer may be null on download; upload (a radiation
 * record, rr####c.d) needs radiosity, q_conv and b_t. */
typedef struct gfm_wall_arrays {
  int16_t *b_typ, *b_i, *b_j, *b_k, *b_orient, *b_direct, *bc_i, *bc_j, *bc_k, *b_exh;  /* (b_cnt) */
  double *patch_area, *b_face_area, *emis, *radiosity, *q_amb, *q_conv, *qc, *qeb, *b_t;  /* (b_cnt) */
  double *vf;          /* (b_cnt, b_cnt) column-major: VF(n,m) = vf[(n-1) + b_cnt*(m-1)] */
  double *vf_sum;      /* (b_cnt): the sum of the row before normalising */
  double *qrs;         /* (MP, NP): melt-surface heat flux (W/m2) */
  double *wall_rad_t;  /* (MP, NP, LP): wall temperature (K) at wall cells, 0 elsewhere */
  double *conv_log;    /* (log_cap): mean relative radiosity change of each sweep of the last call */
  int32_t log_cap, pad;
} gfm_wall_arrays;

int gfm_hip_wall_setup(gfm_ctx *ctx, gfm_wall_ctrl *wc, const int16_t *iwall);
int gfm_hip_wallq(gfm_ctx *ctx, gfm_wall_ctrl *wc);
int gfm_hip_wall_download(gfm_ctx *ctx, const gfm_wall_ctrl *wc, gfm_wall_arrays *a);
int gfm_hip_wall_upload(gfm_ctx *ctx, const gfm_wall_ctrl *wc, const gfm_wall_arrays *a);
size_t gfm_hip_wall_ctrl_size(void);

/* Per-kernel timing with HIP events recorded on the library's stream around
 * each launch.  enable=1 clears the records and starts recording, enable=0
 * stops; every call returns up to `cap` records accumulated so far: name
 * (names + r*name_len, NUL-terminated), total milliseconds, algorithmic
 * bytes moved (the line-relaxation kernel counts 10 words per cell row:
 * 8 coefficients, one read and one write of the unknown; other kernels 0)
 * and launches.  *n receives the record count. */
int gfm_hip_kernel_times(gfm_ctx *ctx, int enable, char *names, int name_len,
                         double *ms, double *bytes, int *launches, int cap, int *n);

int gfm_hip_finalize(gfm_ctx *ctx);

const char *gfm_hip_last_error(void);
int gfm_hip_abi_version(void);
/* sizeof(gfm_ctrl) / sizeof(gfm_ms_ctrl) as compiled into the library, for
 * binding checks. */
size_t gfm_hip_ctrl_size(void);
size_t gfm_hip_ms_ctrl_size(void);

#ifdef __cplusplus
}
#endif
#endif /* GFM_HIP_H */
