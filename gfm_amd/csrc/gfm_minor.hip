// gfm_minor.hip -- minor species and soot on the device: INITSVM
// (C/Initsv.f90:329-380), MSFLOW (C/Msflow.f90:10-169), GSLV3M
// (C/Gslv3m.f90:8-286, SPIN :292-305), SOURCEM (C/Sourcem.f90:8-122, the new
// soot model), flx1dm (C/Flx1d.f90:114-158), extrm (C/Extr.f90:258-305) and
// intpm (C/Intp.f90:348-419).  Included at the end of gfm_hip.hip: the
// kernels use the case block g and the helpers of gfm_kernels.hip, the host
// part the context, launch macros and ADLBL3 driver of gfm_hip.hip.
//
// Layout: GFM(:,:,:,1..7), smf, the soot-source history and RRMS(.,1..5)
// are lattice arrays in the parity-class layout of gfm_dev.h (RRMS at the
// cell centres); the species coefficients of the first GSLV3M after
// INITSVM (AP1) likewise.  Every species equation is assembled into the
// gas solver's AP/AS/BS/FZ and solved by the same line relaxation.

#define MP g.g.MP
#define NP g.g.NP
#define LP g.g.LP
#define DX(i) g.dx[(i)-1]
#define DR(j) g.dr[(j)-1]
#define DZ(k) g.dz[(k)-1]
#define IBC(i, j, k) ((int)g.ib[lidx(i, j, k)])

// ---- constants of C/Var_mod.f90:223-235 ----
__constant__ double c_wms[7] = {16.0426, 31.9988, 28.0104, 18.0152, 44.0098, 28.0134, 30.0061};
__constant__ double c_ck0m[5] = {1.66e15, 7.94e14, 2.24e14, 7.94e22, 9.33e14};
__constant__ double c_betam[5] = {0.0, 0.0, 0.0, 0.0, -0.5};
__constant__ double c_odr[5][2] = {{1.46, 0.5217}, {1.6904, 1.57}, {1.0, 0.0}, {0.721, 4.0111}, {1.0, 0.5}};
__constant__ double c_erc[5] = {20643.0, 11613.0, 62281.0, 44369.0, 68899.0};
#define SML1 1.0e-25

struct MsDev {
  double *gfm[7];
  double *smf, *bssp, *sfpsp, *ap1;
  double *rrms[5];
  const long long *prev;  // per cell centre: the previous open cell of GSLV3M's loop, packed i | j<<16 | k<<32; -1 none
};

// SPIN, C/Gslv3m.f90:292-305 (coefficients AS as stored: FP32 in a mixed context)
__device__ __forceinline__ double SPIN(const MsDev &m, int i, int j, int k, long long q, int lsp) {
  const double *f = m.gfm[lsp - 1];
  double g11 = CAS(0, q) * f[lidx(i - 2, j, k)], g12 = CAS(3, q) * f[lidx(i + 2, j, k)];
  double g21 = CAS(1, q) * f[lidx(i, j - 2, k)], g22 = CAS(4, q) * f[lidx(i, j + 2, k)];
  double g31 = CAS(2, q) * f[lidx(i, j, k - 2)], g32 = CAS(5, q) * f[lidx(i, j, k + 2)];
  double s = g11 + g12 + g21 + g22 + g31 + g32;
  return s * g.gmfr / c_wms[lsp - 1];
}

// the NO rate (reactions 4 + 5) of an open cell: what GSLV3M's RR2 holds
// when the cell is done, C/Gslv3m.f90:94-114
__device__ double NO_RATE(const MsDev &m, int i, int j, int k) {
  long long q = lidx(i, j, k);
  double tmp = g.t[q] * g.t0;
  double dnst1 = g.dnst[q] * g.dnst0;
  double vol = DX(i) * DR(j) * DZ(k) * g.vol0;
  double y2 = m.gfm[1][q], y3 = m.gfm[2][q], y6 = m.gfm[5][q];
  double gfm2 = y2 / c_wms[1] * dnst1, gfm3 = y3 / c_wms[2] * dnst1, gfm6 = y6 / c_wms[5] * dnst1;
  double rr2 = 0.0, rr3 = 0.0;
  if (gfm2 > SMALL20 && gfm3 > SMALL20 && gfm6 > SMALL20) {
    double rr1 = pow(gfm3, c_odr[3][0]) * pow(gfm2, c_odr[3][1]) * y6 / 0.725;
    rr2 = c_ck0m[3] * rr1 * exp(-c_erc[3] / tmp) * vol;
  }
  if (gfm2 > SMALL20 && gfm6 > SMALL20) {
    double rr1 = pow(gfm6, c_odr[4][0]) * pow(gfm2, c_odr[4][1]);
    rr3 = c_ck0m[4] * rr1 * exp(-c_erc[4] / tmp) * pow(tmp, c_betam[4]);
    rr3 = rr3 * vol;
  }
  return rr2 + rr3;
}

// reaction rates of GSLV3M, C/Gslv3m.f90:22-115; partials: sum RRMS1 W_CH4, sum RRMS2 W_CO
__global__ void k_ms_rates(Box b, MsDev m, double *part) {
  int i, j, k;
  double sf = 0.0, sc = 0.0;
  if (box_pt(b, i, j, k) && IBC(i, j, k) <= 0) {
    long long q = lidx(i, j, k);
    double tmp = g.t[q] * g.t0;
    double dnst1 = g.dnst[q] * g.dnst0;
    double vol = DX(i) * DR(j) * DZ(k) * g.vol0;
    double gfm1 = m.gfm[0][q] / c_wms[0] * dnst1, gfm2 = m.gfm[1][q] / c_wms[1] * dnst1;
    double gfm3 = m.gfm[2][q] / c_wms[2] * dnst1, gfm5 = m.gfm[4][q] / c_wms[4] * dnst1;
    // RR2 as the previous open cell of the reference's loop left it
    double rr2 = 0.0, rr1, rr3;
    const long long pv = m.prev[q];
    double r1 = m.rrms[0][q], r2 = m.rrms[1][q];
    double fxms1 = SPIN(m, i, j, k, q, 1), fxms2 = SPIN(m, i, j, k, q, 2);
    if (gfm1 <= SML1 || gfm2 <= SML1) {
      r1 = 0.0;
      if (pv >= 0) rr2 = NO_RATE(m, (int)(pv & 0xffff), (int)((pv >> 16) & 0xffff), (int)(pv >> 32));
    } else {
      rr1 = pow(gfm1, c_odr[0][0]) * pow(gfm2, c_odr[0][1]);
      rr2 = c_ck0m[0] * rr1 * exp(-c_erc[0] / tmp) * vol;
      if (rr2 > fxms1) rr2 = fxms1;
      if (fxms2 / 1.5 < rr2) rr2 = fxms2 / 1.5;
      r1 = rr2 * 0.3 + r1 * 0.7;
    }
    fxms2 = fxms2 - rr2 * 1.5;
    if (fxms2 > SMALL20 && gfm3 > SMALL20) {
      rr1 = pow(gfm3, c_odr[1][0]) * pow(gfm2, c_odr[1][1]);
      rr2 = c_ck0m[1] * rr1 * exp(-c_erc[1] / tmp) * vol;
      if (fxms2 * 2 < rr2) rr2 = fxms2 * 2;
      double fxms3 = SPIN(m, i, j, k, q, 3);
      rr2 = fmin(rr2, fxms3);
    } else {
      rr2 = 0.0;
    }
    if (gfm5 > SMALL20) {
      rr1 = pow(gfm5, c_odr[2][0]);
      rr3 = c_ck0m[2] * rr1 * exp(-c_erc[2] / tmp) * vol;
      double fxms5 = SPIN(m, i, j, k, q, 5);
      rr3 = fmin(rr3, fxms5);
    } else {
      rr3 = 0.0;
    }
    rr1 = rr2 - rr3;
    if (rr1 < 0) rr1 = 0.0;
    r2 = rr1 * 0.3 + r2 * 0.7;
    m.rrms[0][q] = r1;
    m.rrms[1][q] = r2;
    m.rrms[3][q] = NO_RATE(m, i, j, k);
    sf = r1 * c_wms[0];
    sc = r2 * c_wms[2];
  }
  double v0 = wave_sum(sf), v1 = wave_sum(sc);
  if (threadIdx.x == 0) {
    part[block_lin()] = v0;
    part[g.pstride + block_lin()] = v1;
  }
}

__global__ void k_copy_to_fz(Box b, const double *src) {
  int i, j, k;
  if (!box_ptf(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  g.fz[q] = src[q];
}

// one species equation: SOURCEM(l) (C/Sourcem.f90:11-38), the coefficients
// (GSLV3G0 with SSIGMA after INITSVM, kept in AP1; C/Gslv3m.f90:143-150) and
// the relaxation (C/Gslv3m.f90:152-169)
__global__ void k_ms_assemble(Box b, MsDev m, int l, int igs0, double sgm, double rfm) {
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  if (IBC(i, j, k) >= 1) {
    SBS(q, 0.0);
    return;
  }
  const double r1 = m.rrms[0][q], r2 = m.rrms[1][q], r4 = m.rrms[3][q];
  double g0;
  switch (l) {
    case 1: g0 = -r1; break;
    case 2: g0 = -r1 * 1.5 - r2 * 0.5 - r4; break;
    case 3: g0 = r1 - r2; break;
    case 4: g0 = r1 * 2.0; break;
    case 5: g0 = r2; break;
    case 6: g0 = -r4; break;
    default: g0 = r4 * 2.0; break;
  }
  double bs = g0 * c_wms[l - 1] / g.gmfr;
  double ap = 0.0;
  if (igs0 == 0) {
    for (int iu = 1; iu <= 3; iu++)
      for (int mm = -1; mm <= 1; mm += 2) {
        int m1 = (mm + 3) / 2;
        const int asl = (iu - 1) + 3 * (m1 - 1);
        int i1, j1, k1, i2, j2, k2;
        VINDX(iu, mm, i, j, k, i1, j1, k1, i2, j2, k2);
        int ibc0 = IBC(i2, j2, k2);
        double a;
        if (ibc0 <= 0) a = FLWD(i1, j1, k1, iu, m1, sgm);
        else if (ibc0 == 2) a = FLW(i2, j2, k2, iu, m1);
        else a = 0.0;
        SAS(asl, q, a);
        ap += g.mixed ? (double)(float)a : a;
      }
    m.ap1[q] = ap;
  } else {
    ap = m.ap1[q];
  }
  const double fz = g.fz[q];
  double vol = VOL_C(i, j, k);
  ap = ap - (0.0 * vol);
  ap = ap / rfm;
  bs = bs + (1.0 - rfm) * fz * ap;
  SAP(q, ap);
  SBS(q, bs);
}

// clamps and relaxation of a species, C/Gslv3m.f90:176-191
__global__ void k_ms_update(Box b, MsDev m, int l, double rfm) {
  int i, j, k;
  if (!box_ptf(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  double fz = g.fz[q];
  if (fz > 1.0) fz = 1.0;
  if (fz < 0) fz = 0.0;
  if (l == 3 && fz > 0.1) fz = 0.1;
  if (l == 7 && fz > 1.0e-2) fz = 1.0e-2;
  double v = (1.0 - rfm) * m.gfm[l - 1][q] + rfm * fz;
  if (v < SMALL20) v = 0.0;
  m.gfm[l - 1][q] = v;
}

// bal_MS partials: |1 - sum GFM| VOL_C, C/Gslv3m.f90:196-217
__global__ void k_ms_balance(Box b, MsDev m, double *part) {
  int i, j, k;
  double s = 0.0;
  if (box_pt(b, i, j, k) && IBC(i, j, k) < 1) {
    long long q = lidx(i, j, k);
    double g0 = 0.0;
#pragma unroll
    for (int l = 0; l < 7; l++) g0 += m.gfm[l][q];
    s = fabs(1.0 - g0) * VOL_C(i, j, k);
  }
  double v = wave_sum(s);
  if (threadIdx.x == 0) part[block_lin()] = v;
}

// the soot equation's sources (SOURCEM(11), C/Sourcem.f90:69-119), kept as
// the next history (zero outside the cell box and at blocked cells), and
// its coefficients (GSLV3G0 with IGS = 1: AP = sum AS) and relaxation
// (C/Gslv3m.f90:225-256), over the cell centres [2, MP] x [2, NP] x [2, LP]
__global__ void k_soot_assemble(Box ball, Box b, MsDev m, int react, int itr_ms, double aform, double aoxid,
                                double esf_r, double eso_r, double t0_sf, double t0_so, double rf_sf, double rf_so,
                                double rfm) {
  int i, j, k;
  if (!box_ptf(ball, i, j, k)) return;
  long long q = lidx(i, j, k);
  const bool inbox = i >= b.i0 && i <= b.i0 + 2 * (b.ni - 1) && j >= b.j0 && j <= b.j0 + 2 * (b.nj - 1) && k >= b.k0 &&
                     k <= b.k0 + 2 * (b.nk - 1);
  const int bc = IBC(i, j, k);
  double bs = 0.0, sfp = 0.0;
  if (inbox && bc <= 0 && react) {
    const double snorm = g.r0 / (g.dnst0 * g.ug0);
    const double e0_sf = (t0_sf < SMALL20) ? 0.0 : exp(-esf_r / t0_sf);
    const double e0_so = (t0_so < SMALL20) ? 0.0 : sqrt(t0_so) * exp(-eso_r / t0_so);
    double tcell = g.t[q] * g.t0;
    if (tcell > t0_sf) {
      bs = aform * m.gfm[0][q] * (exp(-esf_r / tcell) - e0_sf) * snorm;
      if (itr_ms > 1) bs = rf_sf * bs + (1.0 - rf_sf) * m.bssp[q];
    }
    if (tcell > t0_so) {
      sfp = -(aoxid * m.gfm[1][q] * (sqrt(tcell) * exp(-eso_r / tcell) - e0_so) * snorm);
      if (itr_ms > 1) sfp = rf_so * sfp + (1.0 - rf_so) * m.sfpsp[q];
    }
  }
  m.bssp[q] = bs;
  m.sfpsp[q] = sfp;
  if (!inbox) return;
  if (bc >= 1) {
    SBS(q, 0.0);
    return;
  }
  double ap = 0.0;
  for (int iu = 1; iu <= 3; iu++)
    for (int mm = 1; mm <= 2; mm++) ap += CAS((iu - 1) + 3 * (mm - 1), q);
  const double fz = g.fz[q];
  double vol = VOL_C(i, j, k);
  bs = bs * vol;
  ap = ap - (sfp * vol);
  ap = ap / rfm;
  bs = bs + (1.0 - rfm) * fz * ap;
  SAP(q, ap);
  SBS(q, bs);
}

// soot clipping and relaxation, C/Gslv3m.f90:263-284; partials: max FZ,
// max kept value, clipped cells, soot mass
__global__ void k_soot_update(Box b, MsDev m, double rfm, double cap, double *part) {
  int i, j, k;
  double nc = 0.0, mx = 0.0, cnt = 0.0, tot = 0.0;
  if (box_pt(b, i, j, k) && IBC(i, j, k) < 1) {
    long long q = lidx(i, j, k);
    double fz = g.fz[q];
    nc = fz;
    double g0;
    if (fz > cap) { g0 = cap; cnt = 1.0; }
    else if (fz > SMALL20) g0 = fz;
    else g0 = 0.0;
    mx = g0;
    double s = rfm * g0 + (1.0 - rfm) * m.smf[q];
    m.smf[q] = s;
    double vol = DX(i) * DR(j) * DZ(k);
    tot = s * g.dnst[q] * g.dnst0 * vol;
  }
  double a = wave_max(nc), bb = wave_max(mx), cc = wave_sum(cnt), dd = wave_sum(tot);
  if (threadIdx.x == 0) {
    const long long bl = block_lin();
    part[bl] = a;
    part[g.pstride + bl] = bb;
    part[2 * g.pstride + bl] = cc;
    part[3 * g.pstride + bl] = dd;
  }
}
__global__ __launch_bounds__(FIN_NT) void k_fin4(const double *part, int nb, unsigned maxmask, int dst0) {
  double r[4];
  fin_reduce<4>(part, g.pstride, nb, maxmask, r);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) g.scal[dst0 + q] = r[q];
}

// flx1dm's flows through the planes MSFLOW reads (C/Flx1d.f90:114-158):
// block p: 0 = x-plane 5 in -x, 1 = x-plane MP-5 in +x, 2 = x-plane MP-3 in
// +x, 3 = y-plane 5 in -y, 4 = y-plane NP-3 in +y; 7 species each
__global__ __launch_bounds__(256) void k_ms_planes(MsDev m) {
  __shared__ double sh[7][4];
  const int p = blockIdx.x;
  const int ix = (p < 3) ? 1 : 2;
  const int plane = (p == 0) ? 5 : (p == 1) ? MP - 5 : (p == 2) ? MP - 3 : (p == 3) ? 5 : NP - 3;
  const bool neg = (p == 0 || p == 3);
  const int na = (ix == 1) ? (NP - 6) / 2 + 1 : (MP - 6) / 2 + 1;  // J = 4..NP-2 or I = 4..MP-2
  const int nk = (LP - 6) / 2 + 1;                                  // K = 4..LP-2
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < na * nk; t += blockDim.x) {
    const int a = 4 + 2 * (t % na), k = 4 + 2 * (t / na);
    const int i = (ix == 1) ? plane : a, j = (ix == 1) ? a : plane;
    if (IBC(i, j, k) == 1) continue;
    const long long q = lidx(i, j, k);
    double area = AREA_C(i, j, k, ix);
    double g0 = g.theta[q] * g.dnst[q] * g.ug[ix - 1][q] * area * g.gmfr;
    if ((g0 >= 0.0) == neg) continue;
    for (int l = 0; l < 7; l++) acc[l] += fabs(g0) * m.gfm[l][q];
  }
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int l = 0; l < 7; l++) {
    double v = wave_sum(acc[l]);
    if (ln == 0) sh[l][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const int l = threadIdx.x;
    g.scal[SC_MS_FLX + 7 * p + l] = (sh[l][0] + sh[l][1]) + (sh[l][2] + sh[l][3]);
  }
}

// MSFLOW's exit soot volume fraction (C/Msflow.f90:64-82) over the exit
// cells (IBCELL 3) of the inlet/exit list; partials: area, fraction x area
__global__ void k_ms_svf(const long long *lst, int n, MsDev m, double dnst_soot, double *part) {
  int i, j, k;
  double s0 = 0.0, s1 = 0.0;
  if (list_pt(lst, n, i, j, k) && IBC(i, j, k) == 3) {
    double g0 = AREA_C(i, j, k, 3);
    if (g0 < SMALL20) g0 = AREA_C(i, j, k, 2);
    if (g0 < SMALL20) g0 = AREA_C(i, j, k, 1);
    long long q = lidx(i, j, k);
    s0 = g0;
    s1 = m.smf[q] * g.dnst[q] * g.dnst0 / dnst_soot * g0;
  }
  double v0 = wave_sum(s0), v1 = wave_sum(s1);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v0;
    part[g.pstride + blockIdx.x] = v1;
  }
}

// extrm, C/Extr.f90:258-305.  Exit cells: extrapolated from the interior,
// floored at 0, also into the y-face toward the interior cell; smf copied.
__global__ void k_extrm_exit(const long long *lst, int n, MsDev m) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  if ((IBC(i, j, k) % 10) != 3) return;
  int i1, j1, k1, i2, j2, k2;
  int ix = EXTIDX(0, i, j, k, i1, j1, k1, i2, j2, k2);
  if (ix == 4) { i1 = i; j1 = j; k1 = min(LP, k + 2); i2 = i1; j2 = j1; k2 = k1; }
  double f1, f2;
  EXTR1F(f1, f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
  const long long q = lidx(i, j, k), q1 = lidx(i1, j1, k1), q2 = lidx(i2, j2, k2), qf = lidx(i, (j + j1) / 2, k);
  for (int l = 0; l < 7; l++) {
    double v = fmax(m.gfm[l][q1] * f1 + m.gfm[l][q2] * f2, 0.0);
    m.gfm[l][q] = v;
    m.gfm[l][qf] = v;
  }
  m.smf[q] = m.smf[q1];
}
// wall cells with a non-wall neighbour: copy it, and into the face between
__global__ void k_extrm_wall(const long long *lst, int n, MsDev m) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  int i1, j1, k1, i2, j2, k2;
  int ix = EXTIDX(1, i, j, k, i1, j1, k1, i2, j2, k2);
  if (ix == 4) return;  // on the deep-wall list
  const long long q = lidx(i, j, k), q1 = lidx(i1, j1, k1);
  const long long qf = lidx(ix == 1 ? (i + i1) / 2 : i, ix == 2 ? (j + j1) / 2 : j, ix == 3 ? (k + k1) / 2 : k);
  for (int l = 0; l < 7; l++) {
    double v = m.gfm[l][q1];
    m.gfm[l][q] = v;
    m.gfm[l][qf] = v;
  }
}
// wall cells without one copy the cell two steps up in k (the indices EXT1
// leaves after an unsuccessful search) as it was before the wall loop
// reached it: read before any wall is written, then written
__global__ void k_extrm_deep(const long long *lst, int n, MsDev m, double *tmp, int write) {
  int i, j, k;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (!list_pt(lst, n, i, j, k)) return;
  const int k1 = min(LP, k + 2);
  if (!write) {
    const long long q1 = lidx(i, j, k1);
    for (int l = 0; l < 7; l++) tmp[(long long)t * 7 + l] = m.gfm[l][q1];
  } else {
    const long long q = lidx(i, j, k), qf = lidx(i, j, (k + k1) / 2);
    for (int l = 0; l < 7; l++) {
      double v = tmp[(long long)t * 7 + l];
      m.gfm[l][q] = v;
      m.gfm[l][qf] = v;
    }
  }
}

// intpm (C/Intp.f90:348-419): INTP's positions and weights (intp_code) on
// the seven species
template <int DIR>
__global__ void k_intpm(Box b, MsDev m) {
  constexpr int dir = DIR;
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  int bc = IBC(i, j, k);
  if (bc == 1 || bc == 3) return;
  long long q = lidx(i, j, k);
  int code = g.intp_code[q];
  double a = (code == 1) ? 1.0 : (code == 2) ? 0.0 : INTP1(i, j, k, dir);
  double bb = 1.0 - a;
  long long q1, q2;
  if (dir == 3) { q1 = lidx(i, j, k + 1); q2 = lidx(i, j, k - 1); }
  else if (dir == 2) { q1 = lidx(i, j + 1, k); q2 = lidx(i, j - 1, k); }
  else { q1 = lidx(i + 1, j, k); q2 = lidx(i - 1, j, k); }
  double y1[7], y2[7];
#pragma unroll
  for (int l = 0; l < 7; l++) { y1[l] = m.gfm[l][q1]; y2[l] = m.gfm[l][q2]; }
#pragma unroll
  for (int l = 0; l < 7; l++) m.gfm[l][q] = a * y1[l] + bb * y2[l];
}

// INITSVM's cell-centre assignments (C/Initsv.f90:336-366): all centres
// (new start) or the inlet centres (restart) take CH4, O2, H2O, CO2, N2
// from the gas field; at a restart CO and NO of inlets are zeroed
__global__ void k_initsvm_centres(Box b, MsDev m, int restart) {
  int i, j, k;
  if (!box_ptf(b, i, j, k)) return;
  if (restart && IBC(i, j, k) != 2) return;
  long long q = lidx(i, j, k);
  m.gfm[0][q] = g.gf[IYF - 1][q];
  m.gfm[1][q] = g.gf[IYO2 - 1][q];
  m.gfm[3][q] = g.gf[IYH2O - 1][q];
  m.gfm[4][q] = g.gf[IYCO2 - 1][q];
  m.gfm[5][q] = g.gf[IYN2 - 1][q];
  if (restart) {
    m.gfm[2][q] = 0.0;
    m.gfm[6][q] = 0.0;
  }
}

#undef MP
#undef NP
#undef LP
#undef DX
#undef DR
#undef DZ
#undef IBC

// ======================================================================
// host side
// ======================================================================
struct MsHost {
  MsDev dv;
  double *tmp = nullptr;      // deep-wall values between the two extrm passes
  long long *deep = nullptr;  // wall cells without a non-wall neighbour (Fortran linear indices)
  int n_deep = 0;
};

// device arrays of the minor species, GSLV3M's previous-open-cell table
// (its loop: I, then J, then K innermost over the cell box, IBCELL >= 1
// skipped) and extrm's deep-wall list, at the first gfm_hip_ms_* call
static int ms_alloc(gfm_ctx *c) {
  if (c->ms) return GFM_OK;
  const gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  const DevGeom &gg = c->dc.g;
  const long long N = c->ncls;
  MsHost *m = new MsHost();
  c->ms = m;
  int rc;
  for (int l = 0; l < 7; l++)
    if ((rc = dalloc(c, &m->dv.gfm[l], N))) return rc;
  for (int r = 0; r < 5; r++)
    if ((rc = dalloc(c, &m->dv.rrms[r], N))) return rc;
  if ((rc = dalloc(c, &m->dv.smf, N)) || (rc = dalloc(c, &m->dv.bssp, N)) || (rc = dalloc(c, &m->dv.sfpsp, N)) ||
      (rc = dalloc(c, &m->dv.ap1, N)))
    return rc;
  const int16_t *ib = c->ibh.data();
  auto F = [&](int i, int j, int kk) { return (((long long)kk - 1) * NPh + (j - 1)) * MPh + (i - 1); };
  std::vector<long long> prev(N, -1);
  long long last = -1;
  for (int i = 4; i <= MPh - 2; i += 2)
    for (int j = 4; j <= NPh - 2; j += 2)
      for (int kk = 4; kk <= LPh - 2; kk += 2) {
        if (ib[F(i, j, kk)] > 0) continue;
        prev[lat_index(gg, i, j, kk)] = last;
        last = (long long)i | ((long long)j << 16) | ((long long)kk << 32);
      }
  long long *pd;
  if ((rc = dalloc(c, &pd, N))) return rc;
  HCHECK(hipMemcpy(pd, prev.data(), N * sizeof(long long), hipMemcpyHostToDevice));
  m->dv.prev = pd;
  std::vector<long long> deep;
  for (int kk = 2; kk <= LPh; kk += 2)
    for (int j = 2; j <= NPh; j += 2)
      for (int i = 2; i <= MPh; i += 2) {
        if (ib[F(i, j, kk)] % 10 != 1) continue;
        bool found = false;
        for (int d = 1; d <= 3 && !found; d++)
          for (int s = -2; s <= 2 && !found; s += 4) {
            int i1 = std::max(2, std::min(MPh, i + (d == 1) * s)), j1 = std::max(2, std::min(NPh, j + (d == 2) * s));
            int k1 = std::max(2, std::min(LPh, kk + (d == 3) * s));
            found = ib[F(i1, j1, k1)] % 10 != 1;
          }
        if (!found) deep.push_back(F(i, j, kk));
      }
  if (deep.size() > 0x7fffffff / 8) { set_err("deep-wall list too long"); return GFM_E_ARG; }
  m->n_deep = (int)deep.size();
  if ((rc = dalloc(c, &m->deep, (long long)deep.size() + 1)) || (rc = dalloc(c, &m->tmp, 7 * ((long long)deep.size() + 1))))
    return rc;
  if (!deep.empty()) HCHECK(hipMemcpy(m->deep, deep.data(), deep.size() * sizeof(long long), hipMemcpyHostToDevice));
  return GFM_OK;
}

static int ms_check(const gfm_ctx *c, const gfm_ms_ctrl *ms) {
  if (!c || !ms) { set_err("null argument"); return GFM_E_ARG; }
  if (ms->maxms < 0 || ms->ntimesm < 1 || ms->ntimesm > 1000) { set_err("bad maxms %d or ntimesm %d", ms->maxms, ms->ntimesm); return GFM_E_ARG; }
  for (int l = 0; l < 8; l++)
    if (!(ms->rfm[l] > 0.0 && ms->rfm[l] <= 1.0)) { set_err("rfm[%d] = %g outside (0, 1]", l, ms->rfm[l]); return GFM_E_ARG; }
  if (!(ms->vol_tot > 0.0) || !(ms->dnst_soot > 0.0) || !(ms->ssigma > 0.0)) { set_err("vol_tot, dnst_soot and ssigma must be positive"); return GFM_E_ARG; }
  return GFM_OK;
}

static void ms_extrm(gfm_ctx *c) {
  MsHost *h = c->ms;
  const unsigned nbio = (unsigned)std::max(1, (c->n_extr_io + 63) / 64);
  hipLaunchKernelGGL(k_extrm_exit, dim3(nbio), dim3(64), 0, c->stream, c->extr_io, c->n_extr_io, h->dv);
  const unsigned nbd = (unsigned)std::max(1, (h->n_deep + 63) / 64), nbw = (unsigned)std::max(1, (c->n_extr_wall + 63) / 64);
  hipLaunchKernelGGL(k_extrm_deep, dim3(nbd), dim3(64), 0, c->stream, h->deep, h->n_deep, h->dv, h->tmp, 0);
  hipLaunchKernelGGL(k_extrm_wall, dim3(nbw), dim3(64), 0, c->stream, c->extr_wall, c->n_extr_wall, h->dv);
  hipLaunchKernelGGL(k_extrm_deep, dim3(nbd), dim3(64), 0, c->stream, h->deep, h->n_deep, h->dv, h->tmp, 1);
}

static void ms_intpm(gfm_ctx *c) {
  const int MPh = c->ctrl.mp, NPh = c->ctrl.np, LPh = c->ctrl.lp;
  const MsDev &m = c->ms->dv;
  Box b = make_box(2, MPh, 2, NPh, 3, LPh - 1, 2); LAUNCH(k_intpm<3>, b, m);
  b = make_box(2, MPh, 3, NPh - 1, 2, LPh, 2); LAUNCH(k_intpm<2>, b, m);
  b = make_box(2, MPh, 3, NPh - 1, 3, LPh - 1, 2); LAUNCH(k_intpm<2>, b, m);
  for (int pj = 0; pj < 2; pj++)
    for (int pk = 0; pk < 2; pk++) {
      b = make_box(3, MPh - 1, 2 + pj, pj ? NPh - 1 : NPh, 2 + pk, pk ? LPh - 1 : LPh, 2);
      LAUNCH(k_intpm<1>, b, m);
    }
}

// GSLV3M's launches (C/Gslv3m.f90:8-286)
static void ms_gslv3m_launch(gfm_ctx *c, gfm_ms_ctrl *ms) {
  const int MPh = c->ctrl.mp, NPh = c->ctrl.np, LPh = c->ctrl.lp;
  const MsDev &m = c->ms->dv;
  DevCase &d = c->dc;
  Box cen = make_box(4, MPh - 2, 4, NPh - 2, 4, LPh - 2, 2);
  Box cen_all = make_box(2, MPh, 2, NPh, 2, LPh, 2);
  const int nb = (int)box_blocks(cen);
  LAUNCH(k_ms_rates, cen, m, c->part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, nb, 2, SC_MS_RRF, 0);
  for (int l = 1; l <= 7; l++) {
    LAUNCHF(k_copy_to_fz, cen_all, m.gfm[l - 1]);
    LAUNCH(k_ms_assemble, cen, m, l, ms->igs0, ms->ssigma, ms->rfm[l - 1]);
    ms->igs0 = 1;
    residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, SC_MS_RESPRE - SC_RESID0 + l - 1, 0);
    adlbl3_sweeps(c, d.fz, ms->ntimesm, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
    residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, SC_MS_RES - SC_RESID0 + l - 1, 0);
    LAUNCHF(k_ms_update, cen, m, l, ms->rfm[l - 1]);
  }
  LAUNCH(k_ms_balance, cen, m, c->part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, nb, 1, SC_MS_BAL, 0);
  if (!ms->react) return;
  LAUNCHF(k_copy_to_fz, cen_all, m.smf);
  LAUNCHF(k_soot_assemble, cen_all, cen, m, ms->react, ms->itr_ms, ms->aform, ms->aoxid, ms->esf_r, ms->eso_r, ms->t0_sf,
          ms->t0_so, ms->rf_soot_form_source, ms->rf_soot_oxid_source, ms->rfm[7]);
  residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, SC_MS_RESPRE - SC_RESID0 + 7, 0);
  adlbl3_sweeps(c, d.fz, ms->ntimesm, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
  residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, SC_MS_RES - SC_RESID0 + 7, 0);
  LAUNCH(k_soot_update, cen, m, ms->rfm[7], ms->soot_cap, c->part);
  hipLaunchKernelGGL(k_fin4, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, nb, 0x3u, SC_MS_SMFMAXNC);
}

// flx1dm's planes and the exit soot fraction
static void ms_exit_launch(gfm_ctx *c, const gfm_ms_ctrl *ms) {
  const MsDev &m = c->ms->dv;
  hipLaunchKernelGGL(k_ms_planes, dim3(5), dim3(256), 0, c->stream, m);
  const unsigned nbio = (unsigned)std::max(1, (c->n_extr_io + 63) / 64);
  hipLaunchKernelGGL(k_ms_svf, dim3(nbio), dim3(64), 0, c->stream, c->extr_io, c->n_extr_io, m, ms->dnst_soot, c->part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)nbio, 2, SC_MS_SVF_A, 0);
}

static int ms_pull(gfm_ctx *c, gfm_ms_ctrl *ms, double *sc) {
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(sc, c->dc.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  for (int l = 0; l < 8; l++) {
    ms->resid_ms_pre[l] = sc[SC_MS_RESPRE + l];
    ms->resid_ms[l] = sc[SC_MS_RES + l];
  }
  ms->rr_f = sc[SC_MS_RRF];
  ms->rr_c = sc[SC_MS_RRC];
  ms->bal_ms = sc[SC_MS_BAL] * c->ctrl.vol0 / ms->vol_tot;
  ms->smf_max_noclip = sc[SC_MS_SMFMAXNC];
  ms->smf_max = sc[SC_MS_SMFMAX];
  ms->smf_clipped_cnt = (int)sc[SC_MS_CLIP];
  ms->soot_tot = sc[SC_MS_SOOT];
  return GFM_OK;
}

static int ms_scatter_rrms(gfm_ctx *c, const double *h, double *d, int r) {
  const gfm_ctrl &k = c->ctrl;
  const int MZ = k.mp / 2, NZ = k.np / 2, LZ = k.lp / 2;
  std::vector<double> t(c->nlat, 0.0);
  const double *src = h + (size_t)MZ * NZ * LZ * r;
  for (int kd = 1; kd <= LZ; kd++)
    for (int jd = 1; jd <= NZ; jd++)
      for (int id = 1; id <= MZ; id++)
        t[((long long)(2 * kd - 1) * k.np + (2 * jd - 1)) * k.mp + (2 * id - 1)] =
            src[((size_t)(kd - 1) * NZ + (jd - 1)) * MZ + (id - 1)];
  int rc = scatter(c, t.data(), d);
  HCHECK(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" {

size_t gfm_hip_ms_ctrl_size(void) { return sizeof(gfm_ms_ctrl); }

int gfm_hip_ms_upload(gfm_ctx *c, const gfm_ms_ctrl *ms, const gfm_ms_arrays *a) {
  int rc = ms_check(c, ms);
  if (rc) return rc;
  if (!a || !a->gfm || !a->smf || !a->bs_soot_prev || !a->sfp_soot_prev || !a->rrms) { set_err("null array"); return GFM_E_ARG; }
  if ((rc = bind(c)) || (rc = ms_alloc(c))) return rc;
  MsDev &m = c->ms->dv;
  for (int l = 0; l < 7; l++)
    if ((rc = scatter(c, a->gfm + (size_t)l * c->nlat, m.gfm[l]))) return rc;
  if ((rc = scatter(c, a->smf, m.smf)) || (rc = scatter(c, a->bs_soot_prev, m.bssp)) ||
      (rc = scatter(c, a->sfp_soot_prev, m.sfpsp)))
    return rc;
  for (int r = 0; r < 5; r++)
    if ((rc = ms_scatter_rrms(c, a->rrms, m.rrms[r], r))) return rc;
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

int gfm_hip_ms_download(gfm_ctx *c, gfm_ms_ctrl *ms, gfm_ms_arrays *a) {
  if (!c || !a || !a->gfm || !a->smf || !a->bs_soot_prev || !a->sfp_soot_prev || !a->rrms) { set_err("null argument"); return GFM_E_ARG; }
  (void)ms;
  int rc;
  if ((rc = bind(c)) || (rc = ms_alloc(c))) return rc;
  MsDev &m = c->ms->dv;
  for (int l = 0; l < 7; l++)
    if ((rc = gather(c, m.gfm[l], a->gfm + (size_t)l * c->nlat))) return rc;
  if ((rc = gather(c, m.smf, a->smf)) || (rc = gather(c, m.bssp, a->bs_soot_prev)) || (rc = gather(c, m.sfpsp, a->sfp_soot_prev)))
    return rc;
  const gfm_ctrl &k = c->ctrl;
  const int MZ = k.mp / 2, NZ = k.np / 2, LZ = k.lp / 2;
  std::vector<double> t(c->nlat);
  for (int r = 0; r < 5; r++) {
    if ((rc = gather(c, m.rrms[r], t.data()))) return rc;
    HCHECK(hipStreamSynchronize(c->stream));
    double *dst = a->rrms + (size_t)MZ * NZ * LZ * r;
    for (int kd = 1; kd <= LZ; kd++)
      for (int jd = 1; jd <= NZ; jd++)
        for (int id = 1; id <= MZ; id++)
          dst[((size_t)(kd - 1) * NZ + (jd - 1)) * MZ + (id - 1)] =
              t[((long long)(2 * kd - 1) * k.np + (2 * jd - 1)) * k.mp + (2 * id - 1)];
  }
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

int gfm_hip_initsvm(gfm_ctx *c, gfm_ms_ctrl *ms, int restart) {
  int rc = ms_check(c, ms);
  if (rc) return rc;
  if ((rc = bind(c)) || (rc = ms_alloc(c))) return rc;
  MsDev &m = c->ms->dv;
  Box cen_all = make_box(2, c->ctrl.mp, 2, c->ctrl.np, 2, c->ctrl.lp, 2);
  if (!restart) {
    for (int l = 0; l < 7; l++) HCHECK(hipMemsetAsync(m.gfm[l], 0, c->ncls * sizeof(double), c->stream));
    HCHECK(hipMemsetAsync(m.smf, 0, c->ncls * sizeof(double), c->stream));
    LAUNCHF(k_initsvm_centres, cen_all, m, 0);
  } else {
    LAUNCHF(k_initsvm_centres, cen_all, m, 1);
    ms_extrm(c);
  }
  ms->igs0 = 0;
  ms_intpm(c);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

int gfm_hip_gslv3m(gfm_ctx *c, gfm_ms_ctrl *ms) {
  int rc = ms_check(c, ms);
  if (rc) return rc;
  if ((rc = bind(c)) || (rc = ms_alloc(c))) return rc;
  ms_gslv3m_launch(c, ms);
  double sc[SC_COUNT];
  return ms_pull(c, ms, sc);
}

int gfm_hip_msflow(gfm_ctx *c, gfm_ms_ctrl *ms) {
  int rc = ms_check(c, ms);
  if (rc) return rc;
  if ((rc = bind(c)) || (rc = ms_alloc(c))) return rc;
  double sc[SC_COUNT];
  double gfm0[7];
  ms_exit_launch(c, ms);
  if ((rc = ms_pull(c, ms, sc))) return rc;
  for (int l = 0; l < 7; l++) gfm0[l] = sc[SC_MS_FLX + l] + sc[SC_MS_FLX + 7 + l];
  ms->igs0 = 0;
  ms->iterations = 0;
  ms->converged = 0;
  for (int it = 1; it <= ms->maxms; it++) {
    ms->itr_ms += 1;
    ms->iterations = it;
    ms_gslv3m_launch(c, ms);
    ms_exit_launch(c, ms);
    if ((rc = ms_pull(c, ms, sc))) return rc;
    double gfmex = 0.0, chg = 0.0;
    for (int l = 0; l < 7; l++) {
      double v = sc[SC_MS_FLX + l] + sc[SC_MS_FLX + 14 + l] + sc[SC_MS_FLX + 21 + l] + sc[SC_MS_FLX + 28 + l];
      ms->gfm_0[l] = v;
      gfmex += v;
      chg += fabs(v - gfm0[l]);
      gfm0[l] = v;
    }
    ms->gfmex = gfmex;
    ms->ave_chg = chg / 7;
    ms->svf_r = sc[SC_MS_SVF_S] / sc[SC_MS_SVF_A];
    int done = 0;
    if (ms->itr_ms > 10) {
      done = 1;
      for (int l = 0; l < 8; l++)
        if (ms->resid_ms[l] > ms->bmcon) done = 0;
    }
    if (done) { ms->converged = 1; break; }
    ms_extrm(c);
    ms_intpm(c);
  }
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

}  // extern "C"
