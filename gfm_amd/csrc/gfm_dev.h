// gfm_dev.h -- device data model of the MI355X SIMPLER solver.
//
// Every field of the reference lives on the doubled-index lattice
// (MP,NP,LP) (centres at even indices, faces/edges/corners at odd ones,
// C/Var_mod.f90).  On the device a lattice array is stored as its 8 parity
// sub-lattices ("classes"), each a dense x-fastest block.  Within a row the
// compact x index c = I>>1 is split by its parity (the zebra colour of the
// x-plane in the line relaxation) so that planes of one colour are
// adjacent in memory:
//
//   xpos(c)      = (c&1)*HX + (c>>1) + 31,
//   offset(I,J,K) = cls*S + ((K>>1)*NYC + (J>>1))*NXP + xpos(I>>1),
//   cls = (I&1) | (J&1)<<1 | (K&1)<<2,  NXP = 2*HX,  S = NXP*NYC*NZC,
//   HX = ceil((ceil(NXC/2) + 31)/32)*32,  NXC = MP/2+1.
//
// HX is a multiple of 32 elements (128 bytes of FP32, 256 of FP64) and the
// offset 31 puts the first interior plane of either colour (c = 2 or 3,
// c>>1 = 1) at element 32, so a tile of 16 consecutive FP64 planes or 32
// FP32 planes of one colour is exactly one 128-byte L2 line of each row.  Kernels keep the reference's lattice
// indices (I,J,K).  A wavefront walking consecutive I of one class touches
// two contiguous segments; a wavefront walking consecutive planes of one
// colour touches one.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/gfm_hip.h"
#include <stdint.h>

struct DevGeom {
  int MP, NP, LP;
  int NXC, NYC, NZC;
  int HX, NXP;  // half-row stride and row length (multiples of 32)
  long long S;  // entries per class
};

__host__ __device__ __forceinline__ long long lat_index(const DevGeom &gg, int i, int j, int k) {
  int cls = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
  int c = i >> 1;
  int xp = (c & 1) * gg.HX + (c >> 1) + 31;
  return (long long)cls * gg.S + ((long long)(k >> 1) * gg.NYC + (j >> 1)) * gg.NXP + xp;
}

// Everything the kernels read: pointers to device arrays and the scalar
// constants of the case.  Held in __constant__ memory (one bound case at a
// time per device).
struct DevCase {
  DevGeom g;
  // 1-D grid (1-based index i -> x[i-1])
  const double *x, *dx, *r, *dr, *rs, *z, *dz;
  const int16_t *ib;
  const int8_t *intp_code;  // effective INTP weight override per lattice position
  double *p, *t, *dnst, *theta, *tmu, *gcp, *glam, *gama, *gte, *crate, *hrate, *bsh;
  double *ug[3];
  double *gf[9];
  double *fz, *bs, *sfp, *ap;
  double *as[6];  // AS(.,d,m) -> as[(d-1)+3*(m-1)]
  // FP32 coefficient storage of the mixed-precision mode
  float *apf, *bsf;
  float *asf[6];
  int mixed;
  const double *rtim, *delrt, *delr, *timed;
  double *scal;  // device scalar block (see SC_* indices)
  long long pstride;  // stride between partial-sum slots
  int nr0, ncells;
  int i_ref, j_ref, k_ref;
  int njy1v[4], njy2v[4], lkz1v[4], lkz2v[4];
  double t0, trdg, h_0, r0, area0, vol0;
  double reyg, reyg2, eun, epsr, taur, ug0, dnst0, qr0, q0, gmfr, tk0, epsi0;
  double gra[3], p_ref;
  double t_mx, t_mn, u_mx, u_mn, p_mx, p_mn, f_mx, f_mn;
  double rf[21], rfl[21], rf_hsource;
  double t_wall, kappa_gray, sigma_sb, h_g;
  int i1st_rad_done, isothermal_walls, gray_radiation;
  double wt0, stoich;
  double bv_f[2], ba_f[6];  // half-blocked cells: volume and face-area factors (&GRD BV_F, BA_F)
  // QA and QE (W) of the last radiation pass on the cell centres, (id-1) + MZ*((jd-1) + NZ*(kd-1));
  // null until one has run (enthalpy source GFM_RADSRC_FIELD)
  const double *rad_qa, *rad_qe;
  // q_conv (W/m2) of each wall and melt patch of the last wall exchange at its face (lattice), 0 elsewhere;
  // null until one has run (the enthalpy sink of C/Source.f90:473-509 while i1st_rad_done > 0)
  const double *wall_qcf;
};

// device scalar block indices
enum {
  SC_RES_MASS = 0, SC_RES_MAX, SC_IMX, SC_JMX, SC_KMX,
  SC_GFEX, SC_FUEL_OUT, SC_TMEAN_EX, SC_TAREA, SC_AREA_EX,
  SC_GFIN0, SC_GFIN, SC_FACTOR, SC_FADJ, SC_QEXH, SC_QREACT, SC_G0REF,
  SC_TAVE_NUM, SC_TAVE_DEN, SC_TVOL_NUM, SC_TVOL_DEN, SC_TMAX,
  SC_QWALL,  // total convective heat to walls and melt surface (W)
  SC_QWALL_LO, SC_QWALL_HI,  // the parts through bottom (k-) and top (k+) faces
  SC_ITR,  // itr_gas of the running iteration (read by kernels so a captured graph is reusable)
  SC_PG_RMS, SC_PG_ALPHA, SC_PG_FALLBACK,  // learned p' guess: input scale, guess scale, fallback flag
  SC_PG_COS2,  // learned p' guess: fraction of ||BS||^2 the scaled guess removes
  SC_RESID0 = 32,  // 64 residual slots follow
  SC_TEXIT_AREA = 96, SC_TEXIT_SUM, SC_TEXIT_MAX,  // exit cells: area, T*area, max T (SPHASE's exit temperatures)
  // minor species: calc_residual before / after each of the 8 solves, the
  // rate sums, the balance, the soot reductions, the exit soot fraction and
  // the 5 x 7 plane flows MSFLOW reads
  SC_MS_RESPRE = 128, SC_MS_RES = 136,
  SC_MS_RRF = 144, SC_MS_RRC, SC_MS_BAL, SC_MS_SMFMAXNC, SC_MS_SMFMAX, SC_MS_CLIP, SC_MS_SOOT, SC_MS_SVF_A, SC_MS_SVF_S,
  SC_MS_FLX = 160,
  SC_COUNT = 256
};

struct Box {  // lattice index box with a uniform step
  int i0, ni, j0, nj, k0, nk, step;
  int map;  // thread mapping of tunable box kernels (0: one x row per row of workgroups, 1: (i, j) flattened)
};

static inline Box make_box(int i0, int i1, int j0, int j1, int k0, int k1, int step) {
  Box b;
  b.i0 = i0; b.j0 = j0; b.k0 = k0; b.step = step;
  b.map = 0;
  b.ni = (i1 >= i0) ? (i1 - i0) / step + 1 : 0;
  b.nj = (j1 >= j0) ? (j1 - j0) / step + 1 : 0;
  b.nk = (k1 >= k0) ? (k1 - k0) / step + 1 : 0;
  return b;
}
