// gfm_rad.hip -- radiation transport on the device: one full radiation pass
// of RAD_EMIS (C/Rad_emis.f90:130-229) up to QRN1 -- the wall marks
// (:155-186), RAD0 (C/Rad0.f90:12-172) with the band model of RAD_GAS,
// AH2O, ACO2, A1, CHIWHI, BANDA (C/Rad_gas.f90:16-396) and SOOTA, SOC
// (C/Rad_soot.f90:26-69), the band fractions DEBX/RIF (C/Rad_fun.f90:45-140),
// QRNF's emitter loop (C/Qrn.f90:58-200) with QREF (C/Qre.f90:31-505) and
// QRTX, QRTY, QRTZ (C/Qrtx.f90:15-255, C/Qrty.f90:9-224, C/Qrtz.f90:15-234),
// QRN1's sums (C/Qrn.f90:245-290) -- and rad_qe_only (C/Rad_emis.f90:411-466).
// Included after gfm_minor.hip: the kernels use the case block g, the
// minor-species arrays (H2O, CO2, soot) and the host helpers of gfm_hip.hip.
//
// Layout: the radiation fields live on the cell centres in the reference's
// order (id fastest: cc = (id-1) + MZ*((jd-1) + NZ*(kd-1))), band-major for
// the per-band arrays ([l*ncell + cc]).  QRNF visits emitters in the loop
// order I, J, K innermost (lo = ((id-1)*NZ + (jd-1))*LZ + kd-1); the sums
// are reduced in that order from per-(emitter, axis) partials.
//
// QREF on the device: one 64-lane workgroup per (emitter, axis) at a time
// (a fixed set of workgroups each with its own scratch slot loops over the
// items).  Inside an item the slabs are marched in order; per slab the far
// face's angles are formed by the lanes (first failing index by ballot),
// QRT's first spreading pass runs one row per lane, the second one column
// per lane (the emitter's own column last: see the carried column below),
// and the absorption one plane cell per lane.  Rows and columns touch only
// their own plane entries except the wall split of the second pass, which
// adds into the neighbouring column (global atomics); absorption adds into
// QAL with global atomics (shared by all emitters).  So QAL, QA and the sums
// agree with the sequential pass to rounding, not bit for bit.
//
// The carried column: the wall split of the second pass uses the column
// J2 = J1 -+ 2 (QRTX; I2 in QRTY, QRTZ); on the emitter's own column it
// keeps the value that static variable last had -- from the last split in
// sequential order (columns ascending) or the first pass's inner loop.  The
// lanes record which columns split and which rows ran an inner loop, so the
// own column's value is the sequential one whenever it was set within the
// same emitter; a value that would come from an earlier emitter is counted
// (carry_unset) and replaced by the own column.

#define MP g.g.MP
#define NP g.g.NP
#define LP g.g.LP
#define DX(i) g.dx[(i)-1]
#define DR(j) g.dr[(j)-1]
#define DZ(k) g.dz[(k)-1]

#define RAD_NB GFM_NWL_MAX
#define RAD_PI 3.141592653589793
#define RAD_HCK 1.439e-2
#define RAD_T0R 1.0e2
#define RAD_SIG 5.668e-8

struct RadDev {
  int nwl, MZ, NZ, LZ;
  long long ncell;
  double wl[RAD_NB], deb0[RAD_NB];
  double sgn, dnst_soot, eps_w, eps_m, eps_c, htc;
  int have_crown;
  int16_t *ibc;        // centre codes with the wall marks
  unsigned char *flag; // 1: RAD_GAS ran (open, T >= 300 K, SA and OPL > 0); 2: emitter (open, T > 300 K)
  double *fvs, *opl, *sat4;  // soot volume fraction, optical path, SA*SGN*T^4
  int *src;            // per loop-order index: the last loop-order index up to it with soot, -1 none
  double *aklg, *akl, *deb, *qel0, *qal;  // [l*ncell + cc]
  double *qe, *qa;     // [cc]
  double *cs;          // per cell: gas and soot emission terms of RAD0 ([2*cc], [2*cc+1])
  double *part;        // per (loop-order index, axis): absorbed at melt, exit, inlet, wall, open ([(lo*3+ax)*5 + cat])
  double *scr;         // QREF scratch: per slot 3 planes * pmax * nwl
  long long pmax;
  int *cnt;            // [0] carry_unset, [1] angle_oob
  double *res;         // reduced outputs (RAD_RES_*)
};
enum { RAD_RES_QEG = 0, RAD_RES_QES, RAD_RES_QEOPEN, RAD_RES_MELT, RAD_RES_EXIT, RAD_RES_INLET, RAD_RES_WALL, RAD_RES_OPEN,
       RAD_RES_QAG, RAD_RES_QAIO, RAD_RES_QAS, RAD_RES_QAW, RAD_RES_LAB, RAD_RES_NEMIT, RAD_RES_QE1, RAD_RES_N };

__device__ __forceinline__ void rad_cell_ijk(const RadDev &r, long long cc, int &i, int &j, int &k) {
  const int id = (int)(cc % r.MZ);
  const long long t = cc / r.MZ;
  i = 2 * id + 2;
  j = 2 * (int)(t % r.NZ) + 2;
  k = 2 * (int)(t / r.NZ) + 2;
}
__device__ __forceinline__ long long rad_cc(const RadDev &r, int i, int j, int k) {
  return (long long)(i / 2 - 1) + (long long)r.MZ * ((j / 2 - 1) + (long long)r.NZ * (k / 2 - 1));
}
__device__ __forceinline__ long long rad_lo(const RadDev &r, int i, int j, int k) {
  return ((long long)(i / 2 - 1) * r.NZ + (j / 2 - 1)) * r.LZ + (k / 2 - 1);
}

// ---- the band model (C/Rad_fun.f90, C/Rad_gas.f90, C/Rad_soot.f90) ----
__device__ double rad_per(int m, int n) {  // PER: m!/(m-n)!
  double per = 1.0;
  if (m == 0 || n == 0) return per;
  for (int i = m - n + 1; i <= m; i++) per *= i;
  return per;
}
__device__ double rad_rif3(double x) {  // RIF(3, 0, x), C/Rad_fun.f90:45-91
  if (x > 1.0e16) return 0.0;
  const double y[4] = {1.0, 3.0, 6.0, 6.0};
  double rif = 0.0;
  for (int n = 1; n <= 1000; n++) {
    double g0 = 0.0;
    for (int i = 0; i <= 3; i++) {
      double xp = 1.0, ap = 1.0;
      for (int q = 0; q < 3 - i; q++) xp *= x;
      for (int q = 0; q < i + 1; q++) ap *= (double)n;
      g0 += y[i] * xp / ap;
    }
    g0 *= exp(-(double)n * x);
    rif += g0;
    if (fabs(g0) <= SMALL16) break;
  }
  const double p2 = RAD_PI * RAD_PI;
  rif = rif * 1.5e1 / (p2 * p2);
  return rif > 1.0 ? 1.0 : rif;
}
__device__ void rad_debx(const RadDev &r, double tg0, double *deb) {  // DEBX, C/Rad_fun.f90:122-140
  double eb0 = 0.0;
  for (int l = 0; l < r.nwl - 1; l++) {
    const double eb1 = rad_rif3(r.deb0[l] / tg0);
    deb[l] = eb1 - eb0;
    if (deb[l] <= 1.0e-3) deb[l] = 0.0;
    else eb0 = eb1;
  }
  deb[r.nwl - 1] = 1.0 - eb0;
}
struct RadGas {  // the module variables of one cell's RAD_GAS call
  double t_rad, tt0, p_rad, yco2, yh2o, opl, w_r, pe;
};
__device__ void rad_chiwhi(const RadGas &s, double tt, const double v[3], const int ng[3], const int nd[3], double &chi,
                           double &whi) {  // CHIWHI, C/Rad_gas.f90:308-353
  if (nd[0] != 0 || nd[1] != 0 || nd[2] != 0) {
    double ud = 0.0, xd = 0.0, xy = 0.0, yn = 0.0;
    for (int k = 0; k < 3; k++) {
      const double u = RAD_HCK * v[k] / tt;
      int nv = nd[k] < 0 ? -nd[k] : 0;
      const int ndk = nd[k] < 0 ? -nd[k] : nd[k];
      double y1 = 0.0;
      for (;;) {
        const double xgas = exp(-u * nv) / rad_per(ng[k] - 1, ng[k] - 1);
        const double x1 = xgas * rad_per(nv + ng[k] + ndk - 1, ng[k] + ndk - 1);
        const double x2 = xgas * rad_per(nv + ng[k] - 1, ng[k] - 1);
        if (x1 < 1.0e-10) break;
        xy += x1;
        xd += x2;
        y1 += sqrt(x1);
        nv++;
        if (nv >= 100) break;
      }
      yn += y1 * y1;
      ud -= u * nd[k];
    }
    chi = (1.0 - exp(ud)) * xy / xd;
    whi = yn / xy / sqrt(s.tt0);
  } else {
    chi = 1.0;
    whi = 1.0 / sqrt(s.tt0);
  }
}
__device__ double rad_banda(const RadGas &s, double alp, double bta, double a0) {  // BANDA, C/Rad_gas.f90:360-396
  if (a0 <= 0.0) return 0.0;
  if (alp < 0.0 || bta < 0.0 || s.w_r * s.pe < 0.0) return NAN;
  if (s.w_r * s.pe == 0.0) return 0.0;
  const double eta = bta * s.pe, tau = alp * s.w_r / a0;
  double a_s;
  if (eta >= 1.0) a_s = (tau <= 1.0) ? tau : 1.0 + log(tau);
  else if (tau <= eta) a_s = tau;
  else if (tau <= 1.0 / eta) a_s = 2.0 * sqrt(eta * tau) - eta;
  else a_s = log(eta * tau) + 2.0 - eta;
  return a0 * a_s;
}
__device__ double rad_a1(const RadGas &s, const double v[3], const int ng[3], const double b[3], const int nd[3]) {
  double chi0, whi0, chi, whi;  // A1, C/Rad_gas.f90:284-300
  rad_chiwhi(s, RAD_T0R, v, ng, nd, chi0, whi0);
  rad_chiwhi(s, s.t_rad, v, ng, nd, chi, whi);
  const double c3 = b[2] * sqrt(s.t_rad / RAD_T0R);
  const double c1 = b[0] * chi / chi0;
  const double c2 = 2.0 * sqrt(c1 * c3 * b[1] * whi / whi0);
  return rad_banda(s, c1, c2, c3);
}
__device__ void rad_ah2o(RadGas &s, double aah2o[5]) {  // AH2O, C/Rad_gas.f90:149-213
  const int ng[3] = {1, 1, 1};
  const double v[3] = {3.652e5, 1.595e5, 3.756e5};
  double b[3], chi0, whi0, chi, whi;
  int nd[3] = {0, 0, 0};
  for (int n = 0; n < 5; n++) aah2o[n] = 0.0;
  if (s.yh2o <= 0.0 || s.t_rad <= 3.0e2) return;
  s.pe = s.p_rad + s.p_rad * s.yh2o * (0.86e1 / s.tt0 - 0.5);
  s.w_r = s.p_rad * s.yh2o * s.opl / s.t_rad / 8.208e-2 * 18.0152;
  b[0] = 0.52e4 * exp(-0.176e2 / s.tt0); b[1] = 0.14311; b[2] = 0.284e2;
  aah2o[0] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.412e2; b[1] = 0.9427e-1; b[2] = 0.564e2; nd[1] = 1;
  aah2o[1] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.19; b[1] = 0.13219; b[2] = 0.6e2; nd[1] = 2;
  rad_chiwhi(s, RAD_T0R, v, ng, nd, chi0, whi0);
  rad_chiwhi(s, s.t_rad, v, ng, nd, chi, whi);
  const double t1 = b[0] * chi / chi0, t4 = b[1] * whi / whi0;
  b[0] = 0.23e1; nd[0] = 1; nd[1] = 0;
  rad_chiwhi(s, RAD_T0R, v, ng, nd, chi0, whi0);
  rad_chiwhi(s, s.t_rad, v, ng, nd, chi, whi);
  const double t2 = b[0] * chi / chi0, t5 = b[1] * whi / whi0;
  b[0] = 0.224e2; nd[0] = 0; nd[2] = 1;
  rad_chiwhi(s, RAD_T0R, v, ng, nd, chi0, whi0);
  rad_chiwhi(s, s.t_rad, v, ng, nd, chi, whi);
  const double t3 = b[0] * chi / chi0, t6 = b[1] * whi / whi0;
  const double c1 = t1 + t2 + t3, c3 = b[2] * s.tt0;
  const double bta = 2.0 * sqrt(c3) * (sqrt(t1 * t4) + sqrt(t2 * t5) + sqrt(t3 * t6));
  aah2o[2] = rad_banda(s, c1, bta, c3);
  b[0] = 3.0; b[1] = 0.8169e-1; b[2] = 0.431e2; nd[1] = 1;
  aah2o[3] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.25e1; b[1] = 0.11628; b[2] = 0.32e2; nd[0] = 1; nd[1] = 0;
  aah2o[4] = rad_a1(s, v, ng, b, nd);
  for (int n = 0; n < 5; n++) aah2o[n] *= 1.0e2;
}
__device__ void rad_aco2(RadGas &s, double aaco2[6]) {  // ACO2, C/Rad_gas.f90:224-277
  const int ng[3] = {1, 2, 1};
  const double v[3] = {1.351e5, 6.67e4, 2.396e5};
  double b[3];
  int nd[3];
  s.w_r = s.p_rad * s.yco2 * s.opl / s.t_rad / 8.208e-2 * 44.0098;
  if (s.w_r <= 0.0) return;
  s.pe = pow(s.p_rad + s.p_rad * s.yco2 * 0.3, 0.7);
  b[0] = 0.19e2; b[1] = 0.6157e-1; b[2] = 0.127e2; nd[0] = 0; nd[1] = 1; nd[2] = 0;
  aaco2[0] = rad_a1(s, v, ng, b, nd);
  s.pe = pow(s.p_rad + s.p_rad * s.yco2 * 0.3, 0.8);
  b[0] = 0.36; b[1] = 0.4017e-1; b[2] = 0.134e2; nd[0] = -1; nd[1] = 0; nd[2] = 1;
  aaco2[1] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.24; b[1] = 0.11888; b[2] = 0.101e2;
  aaco2[2] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.11e3; b[1] = 0.24723; b[2] = 0.112e2; nd[0] = 0;
  aaco2[3] = rad_a1(s, v, ng, b, nd);
  s.pe = pow(s.p_rad + s.p_rad * s.yco2 * 0.3, 0.65);
  b[0] = 4.0; b[1] = 0.13341; b[2] = 23.5; nd[0] = 1;
  aaco2[4] = rad_a1(s, v, ng, b, nd);
  b[0] = 0.66e-1; b[1] = 0.39305; b[2] = 0.345e2; nd[0] = 2;
  aaco2[5] = rad_a1(s, v, ng, b, nd);
  for (int n = 0; n < 6; n++) aaco2[n] *= 1.0e2;
}
// SOOTA/SOC, C/Rad_soot.f90:26-69: soot absorption of band l at volume fraction fvs
__device__ double rad_akls(const RadDev &r, int l, double fvs) {
  const double socf[3] = {7.2e26, 4.07e27, 4.47e28}, soch[3] = {1.2e15, 5.90e15, 5.60e15},
               socw[3] = {0.0, 1.25e15, 7.25e15};
  const double wl0 = r.wl[l];
  double xn_k = 1.0, x2nk = 0.0;
  const double w = 2.0 * RAD_PI * 2.9984e8 / wl0;
  for (int j = 0; j < 3; j++) {
    const double a = socw[j] * socw[j] - w * w;
    const double b = a * a + w * w * soch[j] * soch[j];
    if (b < 1.0e-20) break;
    const double g0 = socf[j] * 3.185e3;
    xn_k += g0 * a / b;
    x2nk += g0 * soch[j] * w / b;
  }
  const double g0 = 18.0 * RAD_PI * x2nk / ((xn_k + 2.0) * (xn_k + 2.0) + x2nk * x2nk);
  return g0 * fvs / wl0;
}
// RAD_GAS without SOOTA (C/Rad_gas.f90:16-141): AKLG of the cell
__device__ void rad_aklg(const RadDev &r, RadGas &s, double *aklg) {
  double wnh2o[5][3], wnco2[6][3], wng[11][2], e0[RAD_NB];
  for (int n = 0; n < 5; n++) wnh2o[n][0] = wnh2o[n][1] = wnh2o[n][2] = 0.0;
  for (int n = 0; n < 6; n++) wnco2[n][0] = wnco2[n][1] = wnco2[n][2] = 0.0;
  for (int l = 0; l < r.nwl; l++) aklg[l] = 0.0;
  if (s.yh2o > 1.0e-4) {
    double aa[5];
    wnh2o[1][1] = 1.60e5; wnh2o[2][1] = 3.76e5; wnh2o[3][1] = 5.35e5; wnh2o[4][1] = 7.25e5;
    rad_ah2o(s, aa);
    wnh2o[0][1] = aa[0] / 2;
    wnh2o[0][2] = aa[0];
    for (int n = 1; n < 5; n++) {
      const double g0 = aa[n] / 2;
      wnh2o[n][0] = wnh2o[n][1] - g0;
      wnh2o[n][2] = wnh2o[n][1] + g0;
      if (wnh2o[n - 1][2] > wnh2o[n][0]) wnh2o[n - 1][2] = wnh2o[n][0];
    }
  }
  if (s.yco2 > 1.0e-4) {
    double aa[6] = {0, 0, 0, 0, 0, 0};
    wnco2[0][1] = 6.67e4; wnco2[1][1] = 9.60e4; wnco2[2][1] = 1.06e5; wnco2[3][2] = 2.41e5;
    wnco2[4][1] = 3.66e5; wnco2[5][1] = 5.20e5;
    rad_aco2(s, aa);
    for (int n = 0; n < 6; n++) {
      const double g0 = aa[n] / 2;
      if (n == 0 && g0 > wnco2[0][1]) {
        wnco2[0][0] = 0.0;
        wnco2[0][2] = aa[0];
      } else if (n == 3) {
        wnco2[3][1] = wnco2[3][2] - g0;
        wnco2[3][0] = wnco2[3][1] - g0;
      } else {
        wnco2[n][0] = wnco2[n][1] - g0;
        wnco2[n][2] = wnco2[n][1] + g0;
      }
      if (n == 0) continue;
      if (wnco2[n - 1][2] > wnco2[n][0]) wnco2[n - 1][2] = wnco2[n][0];
    }
  }
  for (int n = 0; n < 5; n++) { wng[n][0] = wnh2o[n][0]; wng[n][1] = wnh2o[n][2]; }
  for (int n = 0; n < 6; n++) { wng[5 + n][0] = wnco2[n][0]; wng[5 + n][1] = wnco2[n][2]; }
  for (int n = 0; n < 10; n++) {
    for (int n2 = n + 1; n2 < 11; n2++) {
      if (wng[n][0] < wng[n2][0]) continue;
      double g0 = wng[n][0]; wng[n][0] = wng[n2][0]; wng[n2][0] = g0;
      g0 = wng[n][1]; wng[n][1] = wng[n2][1]; wng[n2][1] = g0;
    }
    if (n == 0) continue;
    if (wng[n - 1][1] > wng[n][0]) wng[n - 1][1] = wng[n][0];
  }
  double wl2 = 10.0 / r.wl[0];
  for (int l = 0; l < r.nwl; l++) e0[l] = 0.0;
  int n0 = 11;
  for (int n = 1; n <= r.nwl; n++) {
    const double wl1 = (n == r.nwl) ? 0.0 : 1.0 / r.wl[n - 1];
    if (wl2 <= wl1 || wng[10][1] <= wl1) {
      wl2 = wl1;
      continue;
    }
    for (int n1 = n0; n1 >= 1; n1--) {
      if (wng[n1 - 1][1] <= wl1) {
        n0 = n1;
        break;
      }
      if (wng[n1 - 1][1] >= wl1 && wng[n1 - 1][1] <= wl2) {
        if (wng[n1 - 1][0] < wl1) {
          e0[n - 1] += wng[n1 - 1][1] - wl1;
          wng[n1 - 1][1] = wl1;
          n0 = n1;
          break;
        } else {
          e0[n - 1] += wng[n1 - 1][1] - wng[n1 - 1][0];
        }
      }
    }
    if (e0[n - 1] <= 0.0 || wl2 <= wl1) {
      aklg[n - 1] = 0.0;
    } else {
      e0[n - 1] = e0[n - 1] / (wl2 - wl1);
      double g0 = 1.0 - e0[n - 1];
      if (g0 <= 1.0e-20) g0 = 1.0e-20;
      aklg[n - 1] = -log(g0) / s.opl;
    }
    wl2 = wl1;
  }
}

// the wall marks of RAD_EMIS, C/Rad_emis.f90:155-186, into the centre codes
__global__ void k_rad_mark(RadDev r) {
  const long long cc = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cc >= r.ncell) return;
  int i, j, k;
  rad_cell_ijk(r, cc, i, j, k);
  int code = g.ib[lidx(i, j, k)];
  if (code == 1) {
    const int km2 = max(k - 2, 2), kp2 = min(k + 2, LP), jm2 = max(j - 2, 2), jp2 = min(j + 2, NP);
    const int im2 = max(i - 2, 2), ip2 = min(i + 2, MP);
    if (g.ib[lidx(i, j, km2)] <= 0) code = 301;
    else if (g.ib[lidx(i, j, kp2)] <= 0) code = 311;
    else if (g.ib[lidx(i, jm2, k)] <= 0) code = 201;
    else if (g.ib[lidx(i, jp2, k)] <= 0) code = 211;
    else if (g.ib[lidx(im2, j, k)] <= 0) code = 101;
    else if (g.ib[lidx(ip2, j, k)] <= 0) code = 111;
  }
  r.ibc[cc] = (int16_t)code;
}

// RAD0 per cell (C/Rad0.f90:112-166) up to the soot absorption, which
// carries over from cell to cell (k_rad_carry): AKLG, DEB, FVS, OPL, the
// flags; the AKL of melt-surface and marked wall cells.  DEB is formed for
// every open cell above 300 K (QRNF's DEBX of each emitter is the same).
__global__ __launch_bounds__(256) void k_rad0_cell(RadDev r, MsDev m) {
  const long long cc = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cc >= r.ncell) return;
  int i, j, k;
  rad_cell_ijk(r, cc, i, j, k);
  const long long q = lidx(i, j, k);
  const int code = r.ibc[cc];
  const double t_rad = g.t[q] * g.t0;
  unsigned char fl = 0;
  for (int l = 0; l < r.nwl; l++) { r.akl[l * r.ncell + cc] = 0.0; r.aklg[l * r.ncell + cc] = 0.0; }
  r.sat4[cc] = 0.0;
  if (code <= 0 && t_rad >= 300.0) {
    double deb[RAD_NB];
    rad_debx(r, t_rad, deb);
    for (int l = 0; l < r.nwl; l++) r.deb[l * r.ncell + cc] = deb[l];
    if (t_rad > 300.0) fl |= 2;
    RadGas s;
    s.t_rad = t_rad;
    s.tt0 = sqrt(t_rad / RAD_T0R);
    s.p_rad = (1.0 + g.p[q]) * 1.01325e5 / 1.01325e5;
    double sum = 0.0;
    for (int n = 0; n < 7; n++) sum += m.gfm[n][q] / c_wms[n];
    const double wmix = 1.0 / sum;
    s.yco2 = m.gfm[4][q] * wmix / 44.0098;
    s.yh2o = m.gfm[3][q] * wmix / 18.0152;
    s.w_r = 0.0;
    s.pe = 0.0;
    const double fvs = m.smf[q] * g.dnst[q] * g.dnst0 / r.dnst_soot;
    r.fvs[cc] = fvs;
    const double vol = VOL_C(i, j, k);
    const double sa = ((DX(i) * DR(j)) + (DR(j) * DZ(k)) + (DX(i) * DZ(k))) * 2.0;
    if (sa > SMALL16) {
      s.opl = 6.0 * vol / sa;
      if (s.opl > 1.0e-10) {
        fl |= 1;
        r.opl[cc] = s.opl;
        double aklg[RAD_NB];
        rad_aklg(r, s, aklg);
        for (int l = 0; l < r.nwl; l++) r.aklg[l * r.ncell + cc] = aklg[l];
        const double t2 = t_rad * t_rad;
        r.sat4[cc] = sa * r.sgn * (t2 * t2);
      }
    }
  } else if (code == 4) {
    for (int l = 0; l < r.nwl; l++) r.akl[l * r.ncell + cc] = r.eps_m;
  } else if (code > 100) {
    const double e = (g.z[k - 1] >= r.htc || (k == LP && r.have_crown == 0)) ? r.eps_c : r.eps_w;
    for (int l = 0; l < r.nwl; l++) r.akl[l * r.ncell + cc] = e;
  }
  r.flag[cc] = fl;
}

// The soot absorption carried over (SOOTA leaves AKLS unchanged in a cell
// with a volume fraction below 1e-14): src[lo] = the last loop-order index
// <= lo of a cell that ran RAD_GAS with soot, by a prefix maximum over the
// loop order.  One workgroup of 1024 threads, a contiguous chunk each.
#define RAD_SCAN_NT 1024
__global__ __launch_bounds__(RAD_SCAN_NT) void k_rad_carry(RadDev r) {
  __shared__ int sh[RAD_SCAN_NT];
  const int t = threadIdx.x;
  const long long n = r.ncell, ch = (n + RAD_SCAN_NT - 1) / RAD_SCAN_NT;
  const long long lo0 = t * ch, lo1 = min(n, lo0 + ch);
  auto cc_of = [&](long long lo) {
    const int kd = (int)(lo % r.LZ);
    const long long u = lo / r.LZ;
    const int jd = (int)(u % r.NZ), id = (int)(u / r.NZ);
    return (long long)id + (long long)r.MZ * (jd + (long long)r.NZ * kd);
  };
  int last = -1;
  for (long long lo = lo0; lo < lo1; lo++) {
    const long long cc = cc_of(lo);
    if ((r.flag[cc] & 1) && r.fvs[cc] >= 1.0e-14) last = (int)lo;
  }
  sh[t] = last;
  __syncthreads();
  for (int o = 1; o < RAD_SCAN_NT; o <<= 1) {
    const int v = t >= o ? sh[t - o] : -1;
    __syncthreads();
    sh[t] = max(sh[t], v);
    __syncthreads();
  }
  int run = t > 0 ? sh[t - 1] : -1;
  for (long long lo = lo0; lo < lo1; lo++) {
    const long long cc = cc_of(lo);
    if ((r.flag[cc] & 1) && r.fvs[cc] >= 1.0e-14) run = (int)lo;
    r.src[lo] = run;
  }
}

// RAD0's ending per cell: AKLS from the carried source, AKL0 = AKLG + AKLS;
// ID_RAD 10: AKL = EMF(AKL0, OPL) and the gas and soot emission terms;
// ID_RAD 1: QE = SA*SGN*T^4 * sum DEB*EMF(AKL0, OPL).
__global__ void k_rad0_fin(RadDev r, int id_rad) {
  const long long cc = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cc >= r.ncell) return;
  r.cs[2 * cc] = r.cs[2 * cc + 1] = 0.0;
  if (id_rad != 10) r.qe[cc] = 0.0;
  if (!(r.flag[cc] & 1)) return;
  int i, j, k;
  rad_cell_ijk(r, cc, i, j, k);
  const long long s = r.src[rad_lo(r, i, j, k)];
  double fvs_s = 0.0;
  if (s >= 0) {
    const int kd = (int)(s % r.LZ);
    const long long u = s / r.LZ;
    fvs_s = r.fvs[(long long)(u / r.NZ) + (long long)r.MZ * ((u % r.NZ) + (long long)r.NZ * kd)];
  }
  const double opl = r.opl[cc], at4 = r.sat4[cc];
  double sg = 0.0, ss = 0.0, s0 = 0.0;
  for (int l = 0; l < r.nwl; l++) {
    const double akls = s >= 0 ? rad_akls(r, l, fvs_s) : 0.0;
    const double aklg = r.aklg[l * r.ncell + cc];
    const double akl0 = aklg + akls;
    const double deb = r.deb[l * r.ncell + cc];
    if (id_rad == 10) {
      r.akl[l * r.ncell + cc] = EMF(akl0, opl);
      sg += deb * EMF(aklg, opl);
      ss += deb * EMF(akls, opl);
    } else {
      s0 += deb * EMF(akl0, opl);
    }
  }
  if (id_rad == 10) {
    r.cs[2 * cc] = at4 * sg;
    r.cs[2 * cc + 1] = at4 * ss;
  } else {
    r.qe[cc] = at4 * s0;
  }
}

// Each emitter's band powers (QREF C/Qre.f90:64-70: QEL0 = DEB*AKL*SGN*T^4)
// and QE, the sum over the bands of the six faces' powers in QREF's order.
__global__ void k_rad_emit(RadDev r) {
  const long long cc = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cc >= r.ncell) return;
  r.qe[cc] = 0.0;
  if (!(r.flag[cc] & 2)) return;
  int i, j, k;
  rad_cell_ijk(r, cc, i, j, k);
  const double tk = g.t[lidx(i, j, k)] * g.t0, t2 = tk * tk;
  const double g0 = r.sgn * (t2 * t2);
  double qe = 0.0;
  for (int l = 0; l < r.nwl; l++) {
    const double q0 = r.deb[l * r.ncell + cc] * r.akl[l * r.ncell + cc] * g0;
    r.qel0[l * r.ncell + cc] = q0;
    const double qx = q0 * DR(j) * DZ(k), qy = q0 * DX(i) * DZ(k), qz = q0 * DX(i) * DR(j);
    const double ql = ((((((0.0 + qx) + qx) + qy) + qy) + qz) + qz);
    qe = (l == 0) ? ql : qe + ql;
  }
  r.qe[cc] = qe;
}

// ---- QREF ----
// atand of h >= 0 by the same reduction and series as the oracle, with the
// operations in the same order and no contraction, so that angle
// comparisons that tie resolve the same way.
__device__ __attribute__((noinline)) double rad_atand(double h) {
#pragma clang fp contract(off)
  double off = 0.0, sgn = 1.0, off2 = 0.0;
  if (h > 1.0) { h = 1.0 / h; off = RAD_PI / 2.0; sgn = -1.0; }
  double t = h;
  if (h > 0.41421356237309503) { t = (h - 1.0) / (h + 1.0); off2 = RAD_PI / 4.0; }
  const double t2 = t * t;
  double s = 1.0 / 41.0;
  for (int n = 19; n >= 0; n--) s = 1.0 / (double)(2 * n + 1) - t2 * s;
  const double a = off + sgn * (off2 + t * s);
  return a * (180.0 / RAD_PI);
}

struct QrefAx {  // one axis of an item: slab axis A, first spreading axis B, second C
  int A, B, C, a0, b0, c0, Ma, Mb, Mc, PLB, PLC, mk1, mk2, split_gt;
  double *angB, *angC;  // LDS, [(n-1) + M*(col-1)]
  unsigned char *iqt1, *iqtm, *iqt2;
  short *code;                // LDS: cell codes of the current slab, [p]
  double *qtx1, *q2m, *qtx2;  // global scratch, [l*pn + p]
  long long pn;
  double ang0B, ang0C;
};
__device__ __forceinline__ double rad_coord(int d, int n) { return d == 1 ? g.x[n - 1] : d == 2 ? g.r[n - 1] : g.z[n - 1]; }
__device__ __forceinline__ void rad_abc_ijk(const QrefAx &x, int a, int b, int c, int &i, int &j, int &k) {
  i = x.A == 1 ? a : x.B == 1 ? b : c;
  j = x.A == 2 ? a : x.B == 2 ? b : c;
  k = x.A == 3 ? a : x.B == 3 ? b : c;
}
__device__ __forceinline__ int rad_pick(int d, int vi, int vj, int vk) { return d == 1 ? vi : d == 2 ? vj : vk; }
__device__ __forceinline__ long long rad_p(const QrefAx &x, int b, int c) { return (long long)(b / 2 - 1) + (long long)x.PLB * (c / 2 - 1); }

// SANG2 on one side of one spreading axis: ANG(D, n, 2) for n = n0 + s,
// n0 + 3s, ... up to the first angle at or above ang0 (exclusive)
__device__ void rad_sang_side(const QrefAx &x, double *ang, int D, int n0, int s, int M, double dist, double ang0) {
  const int lane = threadIdx.x;
  for (int base = 0;; base += 64) {
    const int nn = n0 + s * (1 + 2 * (base + lane));
    const bool valid = s < 0 ? nn >= 2 : nn <= M;
    double a = 0.0;
    if (valid) a = dist < SMALL16 ? 90.0 : rad_atand(fabs(rad_coord(D, nn) - rad_coord(D, n0)) / dist);
    const bool stop = !valid || a >= ang0;
    const unsigned long long mk = __ballot(stop);
    const int first = mk ? __ffsll((long long)mk) - 1 : 64;
    if (lane < first) ang[(nn - 1) + M] = a;
    if (mk) break;
  }
}

template <int NB>
__device__ void rad_qrt(const RadDev &r, const QrefAx &x, int a10, int a1, int &stale, int &stale_set, int *cnt) {
  const int lane = threadIdx.x;
  const int nwl = r.nwl;
  int a2 = a1 + (a1 - a10) / 2;
  if (a2 > x.Ma) a2 = x.Ma;
  const double dist = fabs(rad_coord(x.A, a2) - rad_coord(x.A, x.a0));
  rad_sang_side(x, x.angB, x.B, x.b0, -1, x.Mb, dist, x.ang0B);
  rad_sang_side(x, x.angB, x.B, x.b0, 1, x.Mb, dist, x.ang0B);
  rad_sang_side(x, x.angC, x.C, x.c0, -1, x.Mc, dist, x.ang0C);
  rad_sang_side(x, x.angC, x.C, x.c0, 1, x.Mc, dist, x.ang0C);
  __syncthreads();
  double q11[NB], q12[NB];
  // first pass: one row c1 = 4 .. Mc-2 per lane
  const int nrow = x.Mc / 2 - 2;
  int p1_val = 0, p1_set = 0;
  for (int base = 0; base < nrow; base += 64) {
    const int c1 = 4 + 2 * (base + lane);
    bool act = base + lane < nrow;
    if (act) {
      if (c1 < x.c0) act = x.angC[c1 + x.Mc] < x.ang0C;  // ANG(C, c1+1, 2)
      else if (c1 > x.c0)
        for (int cp = x.c0 + 2; cp <= c1 && act; cp += 2) act = x.angC[(cp - 2) + x.Mc] < x.ang0C;  // ANG(C, cp-1, 2)
    }
    int my_val = 0, my_set = 0;
    if (act) {
      const int cd2 = c1 / 2;
      for (int bpm = -2; bpm <= 2; bpm += 4) {
        double a11 = 0.0, a12 = 0.0;
        const int bend = bpm < 0 ? 2 : x.Mb;
        for (int b1 = x.b0; bpm < 0 ? b1 >= bend - bpm : b1 <= bend - bpm; b1 += bpm) {
          if (a11 >= x.ang0B) break;
          const int b12 = b1 + bpm / 2;
          a12 = x.angB[b12 - 1];
          const double dang = a12 - a11;
          if (dang <= 1.0e-10) { a11 = a12; continue; }
          double a21 = a11;
          const long long p1 = rad_p(x, b1, c1);
          if (x.iqt1[p1]) { a11 = a12; continue; }
#pragma unroll
          for (int l = 0; l < NB; l++)
            if (l < nwl) {
              const double v = x.qtx1[l * x.pn + p1];
              q11[l] = b1 == x.b0 ? v * 0.5 : v;
              if (b1 != x.b0 || bpm > 0) x.qtx1[l * x.pn + p1] = 0.0;
            }
          if (b1 != x.b0 || bpm > 0) x.iqt1[p1] = 1;
#pragma unroll
          for (int l = 0; l < NB; l++) q12[l] = q11[l];
          int b2;
          for (b2 = b1; bpm < 0 ? b2 >= bend : b2 <= bend; b2 += bpm) {
            const long long p2 = rad_p(x, b2, c1);
            if (x.code[p2] < 1) {
              int b2m = b2 + bpm / 2;
              if (b2m < 1 || b2m > x.Mb) { atomicAdd(&cnt[1], 1); b2m = b2m < 1 ? 1 : x.Mb; }
              const double a22 = x.angB[(b2m - 1) + x.Mb];
              if (a22 <= a21) continue;
              if (a22 < a12) {
                const double fij = (a22 - a21) / dang;
#pragma unroll
                for (int l = 0; l < NB; l++)
                  if (l < nwl && !(q11[l] < SMALL16)) {
                    const double q21 = q11[l] * fij;
                    x.q2m[l * x.pn + p2] += q21;
                    q12[l] -= q21;
                  }
                x.iqtm[p2] = 0;
                a21 = a22;
                continue;
              }
            }
#pragma unroll
            for (int l = 0; l < NB; l++)
              if (l < nwl) x.q2m[l * x.pn + p2] += q12[l];
            x.iqtm[p2] = 0;
            break;
          }
          my_val = b2;
          my_set = 1;
          a11 = a12;
        }
      }
    }
    // the last row (in order) that ran an inner loop gives the carried value
    const unsigned long long mk = __ballot(my_set);
    if (mk) {
      const int hi = 63 - __clzll((long long)mk);
      p1_val = __shfl(my_val, hi);
      p1_set = 1;
    }
  }
  if (p1_set) { stale = p1_val; stale_set = 1; }
  __syncthreads();
  // second pass: one column b1 = 2 .. Mb per lane, the emitter's column last
  const int ncol = x.Mb / 2;
  int lt_split = -1, gt_split = -1;  // highest column below / above b0 with a wall split
  for (int round = 0; round < 2; round++) {
    for (int base = 0; base < ncol; base += 64) {
      const int b1 = 2 + 2 * (base + lane);
      bool act = base + lane < ncol && (round == 0 ? b1 != x.b0 : b1 == x.b0);
      if (act) {
        if (b1 < x.b0) act = x.angB[b1 + x.Mb] < x.ang0B;  // ANG(B, b1+1, 2)
        else if (b1 > x.b0)
          for (int bp = x.b0 + 2; bp <= b1 && act; bp += 2) act = x.angB[(bp - 2) + x.Mb] < x.ang0B;
      }
      int my_split = 0;
      const int cv = round == 1 ? (lt_split >= 0 ? lt_split + 2 : (stale_set ? stale : -1)) : 0;
      if (act) {
        for (int cm = -2; cm <= 2; cm += 4) {
          double a11 = 0.0, a12 = 0.0;
          const int cend = cm < 0 ? 2 : x.Mc;
          for (int c1 = x.c0; cm < 0 ? c1 >= cend - cm : c1 <= cend - cm; c1 += cm) {
            if (a11 >= x.ang0C) break;
            const int c12 = c1 + cm / 2;
            a12 = x.angC[c12 - 1];
            const double dang = a12 - a11;
            if (dang <= 1.0e-10) { a11 = a12; continue; }
            double a21 = a11;
            const long long p1 = rad_p(x, b1, c1);
            if (x.iqtm[p1]) { a11 = a12; continue; }
#pragma unroll
            for (int l = 0; l < NB; l++)
              if (l < nwl) {
                const double v = x.q2m[l * x.pn + p1];
                q11[l] = c1 == x.c0 ? v * 0.5 : v;
                if (c1 != x.c0 || cm > 0) x.q2m[l * x.pn + p1] = 0.0;
              }
            if (c1 != x.c0 || cm > 0) x.iqtm[p1] = 1;
            const int ibc1 = x.code[p1];
            if (ibc1 == x.mk1 || ibc1 == x.mk2 || ibc1 == 4) {
#pragma unroll
              for (int l = 0; l < NB; l++)
                if (l < nwl) atomicAdd(&x.qtx2[l * x.pn + p1], q11[l]);
              x.iqt2[p1] = 0;
              a11 = a12;
              continue;
            }
#pragma unroll
            for (int l = 0; l < NB; l++) q12[l] = q11[l];
            for (int c2 = c1; cm < 0 ? c2 >= cend : c2 <= cend; c2 += cm) {
              const int c2m = c2 + cm / 2;
              const long long p2 = rad_p(x, b1, c2);
              const int ibc2 = x.code[p2];
              bool deposit = false;
              if (ibc1 <= 0 && ibc2 >= 1) {
                deposit = true;
              } else if (ibc1 > x.split_gt && ibc2 == 1) {
                int b2;
                if (b1 < x.b0) b2 = b1 + 2;
                else if (b1 > x.b0) b2 = b1 - 2;
                else {
                  b2 = cv;
                  if (b2 < 0) { atomicAdd(&cnt[0], 1); b2 = b1; }
                  // a first-pass loop that ran to its end leaves the column past the plane
                  if (b2 < 2 || b2 > x.Mb) { atomicAdd(&cnt[1], 1); b2 = b2 < 2 ? 2 : x.Mb; }
                }
                my_split = 1;
                const long long pa = rad_p(x, b2, c2), pb = rad_p(x, b1, c2 - cm);
#pragma unroll
                for (int l = 0; l < NB; l++)
                  if (l < nwl && !(q12[l] < SMALL16)) {
                    const double q21 = q12[l] * 0.5;
                    atomicAdd(&x.qtx2[l * x.pn + pa], q21);
                    atomicAdd(&x.qtx2[l * x.pn + pb], q21);
                  }
                x.iqt2[pa] = 0;
                x.iqt2[pb] = 0;
                break;
              } else if (c2m > x.Mc || c2m < 2) {
                deposit = true;
              } else {
                const double a22 = x.angC[(c2m - 1) + x.Mc];
                if ((a22 - a21) <= 1.0e-10) continue;
                if (a22 < a12) {
                  const double fij = (a22 - a21) / dang;
#pragma unroll
                  for (int l = 0; l < NB; l++)
                    if (l < nwl && !(q11[l] < SMALL16)) {
                      const double q21 = q11[l] * fij;
                      atomicAdd(&x.qtx2[l * x.pn + p2], q21);
                      q12[l] -= q21;
                    }
                  x.iqt2[p2] = 0;
                  a21 = a22;
                  continue;
                }
                deposit = true;
              }
              if (deposit) {
#pragma unroll
                for (int l = 0; l < NB; l++)
                  if (l < nwl) atomicAdd(&x.qtx2[l * x.pn + p2], q12[l]);
                x.iqt2[p2] = 0;
                break;
              }
            }
            a11 = a12;
          }
        }
      }
      if (round == 0) {
        const unsigned long long mlt = __ballot(my_split && b1 < x.b0), mgt = __ballot(my_split && b1 > x.b0);
        if (mlt) lt_split = 2 + 2 * (base + 63 - __clzll((long long)mlt));
        if (mgt) gt_split = 2 + 2 * (base + 63 - __clzll((long long)mgt));
      }
    }
    __syncthreads();
  }
  if (gt_split >= 0) { stale = gt_split - 2; stale_set = 1; }
  else if (lt_split >= 0) { stale = lt_split + 2; stale_set = 1; }
}

// QREF for (emitter, axis) items, C/Qre.f90:95-505: the two directions
// along the axis from the emitter, slab by slab.  Dynamic LDS: the angle
// arrays of the three axes (2*(MP+NP+LP) doubles), the cell codes of the
// current slab and the three flag planes.
template <int NB>
__global__ __launch_bounds__(64) void k_qref(RadDev r) {
  extern __shared__ double rsm[];
  const int lane = threadIdx.x;
  short *codes = reinterpret_cast<short *>(rsm + 2 * (MP + NP + LP));
  unsigned char *flg = reinterpret_cast<unsigned char *>(codes + r.pmax);
  const long long nitem = r.ncell * 3;
  double *slot = r.scr + (long long)blockIdx.x * 3 * r.pmax * r.nwl;
  for (long long item = blockIdx.x; item < nitem; item += gridDim.x) {
    const long long lo = item / 3;
    const int ax = (int)(item % 3);
    const int kd = (int)(lo % r.LZ);
    const long long u = lo / r.LZ;
    const int jd = (int)(u % r.NZ), id = (int)(u / r.NZ);
    const long long cc = (long long)id + (long long)r.MZ * (jd + (long long)r.NZ * kd);
    double *pout = r.part + item * 5;
    if (!(r.flag[cc] & 2)) {
      if (lane < 5) pout[lane] = 0.0;
      continue;
    }
    const int i0 = 2 * id + 2, j0 = 2 * jd + 2, k0 = 2 * kd + 2;
    QrefAx x;
    x.A = ax + 1;
    x.B = ax == 0 ? 2 : 1;
    x.C = ax == 2 ? 2 : 3;
    x.a0 = rad_pick(x.A, i0, j0, k0); x.b0 = rad_pick(x.B, i0, j0, k0); x.c0 = rad_pick(x.C, i0, j0, k0);
    x.Ma = rad_pick(x.A, MP, NP, LP); x.Mb = rad_pick(x.B, MP, NP, LP); x.Mc = rad_pick(x.C, MP, NP, LP);
    x.PLB = x.Mb / 2; x.PLC = x.Mc / 2;
    x.pn = (long long)x.PLB * x.PLC;
    x.mk1 = 100 * x.A + 1; x.mk2 = 100 * x.A + 11;
    x.split_gt = ax == 0 ? 100 : ax == 1 ? 99 : 1;
    x.angB = rsm + rad_pick(x.B, 0, 2 * MP, 2 * (MP + NP));
    x.angC = rsm + rad_pick(x.C, 0, 2 * MP, 2 * (MP + NP));
    x.iqt1 = flg; x.iqtm = flg + x.pn; x.iqt2 = flg + 2 * x.pn;
    x.code = codes;
    x.qtx1 = slot; x.q2m = slot + r.pmax * r.nwl; x.qtx2 = slot + 2 * r.pmax * r.nwl;
    for (long long p = lane; p < x.pn * r.nwl; p += 64) { x.qtx1[p] = 0.0; x.q2m[p] = 0.0; x.qtx2[p] = 0.0; }
    for (long long p = lane; p < 3 * x.pn; p += 64) flg[p] = 1;
    const double db = x.B == 1 ? g.dx[x.b0 - 1] : x.B == 2 ? g.dr[x.b0 - 1] : g.dz[x.b0 - 1];
    const double dc = x.C == 1 ? g.dx[x.c0 - 1] : x.C == 2 ? g.dr[x.c0 - 1] : g.dz[x.c0 - 1];
    const double dbc = db * dc;
    const long long pe = rad_p(x, x.b0, x.c0);
    double s_melt = 0.0, s_exit = 0.0, s_inlet = 0.0, s_wall = 0.0, s_open = 0.0;
    int stale = 0, stale_set = 0;
    for (int dir = 0; dir < 2; dir++) {
      const int apm = dir == 0 ? -2 : 2;
      __syncthreads();
      if (lane < r.nwl) x.qtx1[lane * x.pn + pe] = r.qel0[lane * r.ncell + cc] * dbc;
      if (lane == 0) x.iqt1[pe] = 0;
      // the emitter corner's angles (sfac)
      const double bb = fabs(rad_coord(x.A, x.a0 + apm / 2) - rad_coord(x.A, x.a0));
      if (bb < SMALL16) {
        x.ang0B = 90.0;
        x.ang0C = 90.0;
      } else {
        x.ang0B = rad_atand(fabs(rad_coord(x.B, x.b0 + apm / 2) - rad_coord(x.B, x.b0)) / bb);
        x.ang0C = rad_atand(fabs(rad_coord(x.C, x.c0 + apm / 2) - rad_coord(x.C, x.c0)) / bb);
      }
      for (int n = 1 + lane; n <= x.Mb; n += 64) x.angB[n - 1] = x.angB[n - 1 + x.Mb] = n == x.b0 ? 0.0 : x.ang0B;
      for (int n = 1 + lane; n <= x.Mc; n += 64) x.angC[n - 1] = x.angC[n - 1 + x.Mc] = n == x.c0 ? 0.0 : x.ang0C;
      __syncthreads();
      const int ae = dir == 0 ? 2 : x.Ma;
      for (int a = x.a0 + apm; apm < 0 ? a >= ae : a <= ae; a += apm) {
        for (int p = lane; p < x.pn; p += 64) {
          int i, j, k;
          rad_abc_ijk(x, a, 2 * (p % x.PLB) + 2, 2 * (p / x.PLB) + 2, i, j, k);
          codes[p] = r.ibc[rad_cc(r, i, j, k)];
        }
        __syncthreads();
        rad_qrt<NB>(r, x, a - apm, a, stale, stale_set, r.cnt);
        // absorption in slab a
        for (long long p = lane; p < x.pn; p += 64) {
          if (x.iqt2[p]) continue;
          const int b = 2 * (int)(p % x.PLB) + 2, c = 2 * (int)(p / x.PLB) + 2;
          int i, j, k;
          rad_abc_ijk(x, a, b, c, i, j, k);
          const long long ce = rad_cc(r, i, j, k);
          const int code = codes[p];
          for (int l = 0; l < r.nwl; l++) {
            const double v = x.qtx2[l * x.pn + p];
            if (v < SMALL16) continue;
            if (code >= 1 || a == ae) {
              atomicAdd(&r.qal[l * r.ncell + ce], v);
              if (code == 4) s_melt += v;
              else if (code == 3) s_exit += v;
              else if (code == 2) s_inlet += v;
              else s_wall += v;
            } else {
              const double qa0 = v * r.akl[l * r.ncell + ce];
              atomicAdd(&r.qal[l * r.ncell + ce], qa0);
              x.qtx1[l * x.pn + p] = v - qa0;
              x.iqt1[p] = 0;
              s_open += qa0;
            }
            x.qtx2[l * x.pn + p] = 0.0;
            x.iqt2[p] = 1;
          }
        }
        __syncthreads();
        for (int n = 1 + lane; n <= x.Mb; n += 64) x.angB[n - 1] = x.angB[n - 1 + x.Mb];
        for (int n = 1 + lane; n <= x.Mc; n += 64) x.angC[n - 1] = x.angC[n - 1 + x.Mc];
        __syncthreads();
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      s_melt += __shfl_down(s_melt, o);
      s_exit += __shfl_down(s_exit, o);
      s_inlet += __shfl_down(s_inlet, o);
      s_wall += __shfl_down(s_wall, o);
      s_open += __shfl_down(s_open, o);
    }
    if (lane == 0) { pout[0] = s_melt; pout[1] = s_exit; pout[2] = s_inlet; pout[3] = s_wall; pout[4] = s_open; }
    __syncthreads();
  }
}

// QA = sum of QAL over the bands, C/Qrn.f90:262-271
__global__ void k_rad_qa(RadDev r) {
  const long long cc = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cc >= r.ncell) return;
  double s = r.qal[cc];
  for (int l = 1; l < r.nwl; l++) s += r.qal[l * r.ncell + cc];
  r.qa[cc] = s;
}

// The pass's sums in one workgroup, each thread a contiguous range of the
// loop order, then a fixed tree: RAD0's emission, the traced emission, the
// absorption by destination and by cell type, the open-cell and emitter
// counts; id_rad 1: the emission-only QE.
#define RAD_RED_NT 256
__global__ __launch_bounds__(RAD_RED_NT) void k_rad_reduce(RadDev r, int id_rad) {
  __shared__ double sh[RAD_RED_NT];
  const int t = threadIdx.x;
  const long long n = r.ncell, ch = (n + RAD_RED_NT - 1) / RAD_RED_NT;
  double acc[RAD_RES_N];
  for (int q = 0; q < RAD_RES_N; q++) acc[q] = 0.0;
  for (long long lo = t * ch; lo < min(n, (t + 1) * ch); lo++) {
    const int kd = (int)(lo % r.LZ);
    const long long u = lo / r.LZ;
    const long long cc = (u / r.NZ) + (long long)r.MZ * ((u % r.NZ) + (long long)r.NZ * kd);
    const int code = r.ibc[cc];
    if (id_rad != 10) {
      acc[RAD_RES_QE1] += r.qe[cc];
      continue;
    }
    acc[RAD_RES_QEG] += r.cs[2 * cc];
    acc[RAD_RES_QES] += r.cs[2 * cc + 1];
    if (code < 1) acc[RAD_RES_LAB] += 1.0;
    if (r.flag[cc] & 2) {
      acc[RAD_RES_NEMIT] += 1.0;
      acc[RAD_RES_QEOPEN] += r.qe[cc];
      for (int ax = 0; ax < 3; ax++) {
        const double *pp = r.part + (lo * 3 + ax) * 5;
        acc[RAD_RES_MELT] += pp[0];
        acc[RAD_RES_EXIT] += pp[1];
        acc[RAD_RES_INLET] += pp[2];
        acc[RAD_RES_WALL] += pp[3];
        acc[RAD_RES_OPEN] += pp[4];
      }
    }
    const double qa = r.qa[cc];
    if (code <= 0) acc[RAD_RES_QAG] += qa;
    else if (code == 2 || code == 3) acc[RAD_RES_QAIO] += qa;
    else if (code == 4) acc[RAD_RES_QAS] += qa;
    else acc[RAD_RES_QAW] += qa;
  }
  for (int q = 0; q < RAD_RES_N; q++) {
    sh[t] = acc[q];
    __syncthreads();
    for (int o = RAD_RED_NT / 2; o > 0; o >>= 1) {
      if (t < o) sh[t] += sh[t + o];
      __syncthreads();
    }
    if (t == 0) r.res[q] = sh[0];
    __syncthreads();
  }
}

#undef MP
#undef NP
#undef LP
#undef DX
#undef DR
#undef DZ

// ---------------- host ----------------
struct RadHost {
  RadDev dv;
  double qeg_s_last = 0.0;  // QRN1's melt-surface absorption of the last full pass (QRN_PRN)
  int nwl_cap = 0;
  int nslot = 0;
  size_t lds = 0;
};

static int rad_check(gfm_ctx *c, const gfm_rad_ctrl *rc) {
  if (!c || !rc) { set_err("null argument"); return GFM_E_ARG; }
  if (rc->nwl < 1 || rc->nwl > GFM_NWL_MAX) { set_err("nwl %d outside 1..%d", rc->nwl, GFM_NWL_MAX); return GFM_E_ARG; }
  for (int l = 0; l < rc->nwl; l++)
    if (!(rc->wl[l] > 0.0)) { set_err("band wavelength %d is not positive", l + 1); return GFM_E_ARG; }
  if (!(rc->ri > 0.0) || !(rc->dnst_soot > 0.0)) { set_err("ri and dnst_soot must be positive"); return GFM_E_ARG; }
  if (!c->ms) { set_err("radiation needs the minor-species state (gfm_hip_ms_upload or gfm_hip_initsvm first)"); return GFM_E_STATE; }
  return GFM_OK;
}

// device arrays of the radiation fields, at the first call and when the
// band count grows; the QREF scratch slots
static int rad_alloc(gfm_ctx *c, int nwl) {
  const gfm_ctrl &k = c->ctrl;
  if (!c->rad) c->rad = new RadHost();
  RadHost *h = c->rad;
  RadDev &r = h->dv;
  const int MZ = k.mp / 2, NZ = k.np / 2, LZ = k.lp / 2;
  const long long nc = (long long)MZ * NZ * LZ;
  int rc;
  if (h->nwl_cap == 0) {
    r.MZ = MZ; r.NZ = NZ; r.LZ = LZ; r.ncell = nc;
    if ((rc = dalloc(c, &r.ibc, nc)) || (rc = dalloc(c, &r.flag, nc)) || (rc = dalloc(c, &r.fvs, nc)) ||
        (rc = dalloc(c, &r.opl, nc)) || (rc = dalloc(c, &r.sat4, nc)) || (rc = dalloc(c, &r.src, nc)) ||
        (rc = dalloc(c, &r.qe, nc)) || (rc = dalloc(c, &r.qa, nc)) || (rc = dalloc(c, &r.cs, 2 * nc)) ||
        (rc = dalloc(c, &r.part, 15 * nc)) || (rc = dalloc(c, &r.cnt, 2)) || (rc = dalloc(c, &r.res, RAD_RES_N)))
      return rc;
    r.pmax = std::max(std::max((long long)NZ * LZ, (long long)MZ * LZ), (long long)MZ * NZ);
    c->dc.rad_qa = r.qa;
    c->dc.rad_qe = r.qe;
    g_bound = nullptr;  // the case block changed: bind again
  }
  if (nwl > h->nwl_cap) {
    // the band arrays and the scratch grow (old ones stay with the context until it ends)
    if ((rc = dalloc(c, &r.aklg, nc * nwl)) || (rc = dalloc(c, &r.akl, nc * nwl)) || (rc = dalloc(c, &r.deb, nc * nwl)) ||
        (rc = dalloc(c, &r.qel0, nc * nwl)) || (rc = dalloc(c, &r.qal, nc * nwl)))
      return rc;
    // slots: enough workgroups to fill the chip, the scratch held to 1 GB
    const long long per = 3 * r.pmax * nwl;
    long long ns = std::min<long long>(2048, std::max<long long>(1, (1LL << 30) / (per * (long long)sizeof(double))));
    ns = std::min<long long>(ns, 3 * nc);
    if ((rc = dalloc(c, &r.scr, (size_t)(per * ns)))) return rc;
    h->nslot = (int)ns;
    h->nwl_cap = nwl;
  }
  h->lds = (size_t)2 * (k.mp + k.np + k.lp) * sizeof(double) + (size_t)r.pmax * sizeof(short) + (size_t)3 * r.pmax;
  return GFM_OK;
}

static void rad_set_inputs(gfm_ctx *c, const gfm_rad_ctrl *rc) {
  RadDev &r = c->rad->dv;
  r.nwl = rc->nwl;
  for (int l = 0; l < rc->nwl; l++) { r.wl[l] = rc->wl[l]; r.deb0[l] = RAD_HCK / rc->ri / rc->wl[l]; }
  r.sgn = RAD_SIG * rc->ri * rc->ri;
  r.dnst_soot = rc->dnst_soot;
  r.eps_w = rc->eps_w; r.eps_m = rc->eps_m; r.eps_c = rc->eps_c;
  r.htc = rc->height_to_ceiling;
  r.have_crown = rc->have_crown;
}

static int rad_pass(gfm_ctx *c, gfm_rad_ctrl *rc, int id_rad) {
  int e = rad_check(c, rc);
  if (e) return e;
  if ((e = rad_alloc(c, rc->nwl)) || (e = bind(c))) return e;
  rad_set_inputs(c, rc);
  RadHost *h = c->rad;
  RadDev &r = h->dv;
  MsDev &m = c->ms->dv;
  const unsigned nb = (unsigned)((r.ncell + 255) / 256);
  if (c->prof.on) prof_begin(c);
  hipLaunchKernelGGL(k_rad_mark, dim3(nb), dim3(256), 0, c->stream, r);
  hipLaunchKernelGGL(k_rad0_cell, dim3(nb), dim3(256), 0, c->stream, r, m);
  hipLaunchKernelGGL(k_rad_carry, dim3(1), dim3(RAD_SCAN_NT), 0, c->stream, r);
  hipLaunchKernelGGL(k_rad0_fin, dim3(nb), dim3(256), 0, c->stream, r, id_rad);
  if (c->prof.on) prof_end(c, "k_rad0", 0.0);
  if (id_rad == 10) {
    hipLaunchKernelGGL(k_rad_emit, dim3(nb), dim3(256), 0, c->stream, r);
    HCHECK(hipMemsetAsync(r.qal, 0, sizeof(double) * r.ncell * r.nwl, c->stream));
    HCHECK(hipMemsetAsync(r.cnt, 0, 2 * sizeof(int), c->stream));
    if (c->prof.on) prof_begin(c);
    const int nbk = r.nwl <= 1 ? 1 : r.nwl <= 4 ? 4 : r.nwl <= 10 ? 10 : 16;
    const dim3 grid((unsigned)h->nslot), blk(64);
    if (nbk == 1) hipLaunchKernelGGL(k_qref<1>, grid, blk, h->lds, c->stream, r);
    else if (nbk == 4) hipLaunchKernelGGL(k_qref<4>, grid, blk, h->lds, c->stream, r);
    else if (nbk == 10) hipLaunchKernelGGL(k_qref<10>, grid, blk, h->lds, c->stream, r);
    else hipLaunchKernelGGL(k_qref<16>, grid, blk, h->lds, c->stream, r);
    if (c->prof.on) prof_end(c, "k_qref", 0.0);
    hipLaunchKernelGGL(k_rad_qa, dim3(nb), dim3(256), 0, c->stream, r);
  }
  hipLaunchKernelGGL(k_rad_reduce, dim3(1), dim3(RAD_RED_NT), 0, c->stream, r, id_rad);
  HCHECK(hipGetLastError());
  double res[RAD_RES_N];
  int cnt[2] = {0, 0};
  HCHECK(hipMemcpyAsync(res, r.res, sizeof(res), hipMemcpyDeviceToHost, c->stream));
  if (id_rad == 10) HCHECK(hipMemcpyAsync(cnt, r.cnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  if (id_rad == 10) {
    rc->tot_qe_h2o_co2 = res[RAD_RES_QEG];
    rc->tot_qe_soot = res[RAD_RES_QES];
    rc->qe_sum_from_open = res[RAD_RES_QEOPEN];
    rc->qeg0 = res[RAD_RES_QEOPEN];
    rc->qa_sum_to_melt = res[RAD_RES_MELT];
    rc->qa_sum_to_exit = res[RAD_RES_EXIT];
    rc->qa_sum_to_inlet = res[RAD_RES_INLET];
    rc->qa_sum_to_wall = res[RAD_RES_WALL];
    rc->qa_sum_to_open = res[RAD_RES_OPEN];
    rc->qeg_g = res[RAD_RES_QAG];
    rc->qeg_io = res[RAD_RES_QAIO];
    rc->qeg_s = res[RAD_RES_QAS];
    h->qeg_s_last = res[RAD_RES_QAS];
    rc->qeg_w = res[RAD_RES_QAW];
    rc->l_ab = (int32_t)res[RAD_RES_LAB];
    rc->n_emit = (int32_t)res[RAD_RES_NEMIT];
    rc->carry_unset = cnt[0];
    rc->angle_oob = cnt[1];
  } else {
    rc->qeg0 = res[RAD_RES_QE1];
  }
  for (double v : res)
    if (!isfinite(v)) { set_err("non-finite radiation sum"); return GFM_E_NUMERIC; }
  return GFM_OK;
}

extern "C" {

size_t gfm_hip_rad_ctrl_size(void) { return sizeof(gfm_rad_ctrl); }

int gfm_hip_rad_emis(gfm_ctx *c, gfm_rad_ctrl *rc) { return rad_pass(c, rc, 10); }

int gfm_hip_rad_qe_only(gfm_ctx *c, gfm_rad_ctrl *rc) { return rad_pass(c, rc, 1); }

int gfm_hip_rad_download(gfm_ctx *c, const gfm_rad_ctrl *rc, gfm_rad_arrays *a) {
  if (!c || !rc || !a) { set_err("null argument"); return GFM_E_ARG; }
  if (!c->rad || c->rad->nwl_cap == 0) { set_err("no radiation pass has run"); return GFM_E_STATE; }
  RadDev &r = c->rad->dv;
  if (rc->nwl < 1 || rc->nwl > c->rad->nwl_cap) { set_err("nwl %d not held", rc->nwl); return GFM_E_ARG; }
  HCHECK(hipSetDevice(c->device));
  const size_t nc = (size_t)r.ncell;
  if (a->qe) HCHECK(hipMemcpyAsync(a->qe, r.qe, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (a->qa) HCHECK(hipMemcpyAsync(a->qa, r.qa, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (a->qal) HCHECK(hipMemcpyAsync(a->qal, r.qal, nc * rc->nwl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (a->akl) HCHECK(hipMemcpyAsync(a->akl, r.akl, nc * rc->nwl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

int gfm_hip_rad_upload(gfm_ctx *c, const gfm_rad_arrays *a) {
  if (!c || !a || !a->qe || !a->qa) { set_err("null argument"); return GFM_E_ARG; }
  int rc;
  if ((rc = rad_alloc(c, 1))) return rc;
  RadDev &r = c->rad->dv;
  HCHECK(hipSetDevice(c->device));
  HCHECK(hipMemcpyAsync(r.qe, a->qe, (size_t)r.ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HCHECK(hipMemcpyAsync(r.qa, a->qa, (size_t)r.ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

}  // extern "C"
