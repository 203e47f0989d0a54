/*
 * gfm_oracle.c -- CPU restatement of the GFM v4 combustion-space SIMPLER
 * iteration, used ONLY as the test oracle and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product (gfm_amd/, the HIP library)
 * links, loads or calls this file; tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it to check and to time the GPU path.
 *
 * It works on the reference's own data model: every field on the
 * doubled-index lattice (MP,NP,LP), 1-based, column-major, FP64, with the
 * same loop orders as the Fortran so that floating-point results follow it
 * closely.  Each function cites the reference unit it restates
 * (C/ = combustion/Code/ of the reference package).
 *
 * Deliberate restrictions (checked at entry, documented in the design
 * notes): single phase (NDNP=0), Cartesian (CYL=.false.), turbulent,
 * reacting, steady, SIMPLER; cell codes from -2 (half-blocked) up; no
 * radiation transport (only the optional gray source term).  Unused
 * diagnostic sums (FLX/FLY/FLUX_Z plane fluxes of FLX1D, the Info/gresid
 * file writes) are not formed.
 *
 * Parity pinning: see the headers of tests/test_oracle.py and
 * tests/test_oracle_pins.py for which closed forms, invariants and
 * reference-held data pin each function here (oracle_probe runs single
 * stages for the latter).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "../include/gfm_hip.h"

/* ---- constants of GBL_VAR (C/Var_mod.f90:8-200) ---- */
static const double small20 = 1.0e-20, small16 = 1.0e-16, big = 1.0e20;
static const double wf = 16.0426, wo2 = 31.9988, wco2 = 44.0098, wh2o = 18.0152, wn2 = 28.0134;
static const double cmu = 9.0e-2, ct1 = 1.44, ct2 = 1.92;
static const double cpf1 = 2.156e3 / 1.0057e3;
static const double sigma_eq[12] = {0.7, 0.7, 0.7, 1.0, 1.0, 0.7, 0.7, 0.7, 0.7, 0.7, 0.7, 0.7};
static double wt0, stoich;
static const int iyf = 1, iyo2 = 2, ih = 3, ik = 4, ieps = 5, iyco2 = 6, iyn2 = 7, iyh2o = 9;
static const int lstar = 1, lend = 7;
/* NIST Shomate coefficients (C/Setup.f90:103-141), species order
 * fuel, O2, CO2, N2, H2O; [n][0]=low set, [n][1]=high set. */
static const double a_nist[5][2] = {{-0.703029, 85.81217}, {29.659, 29.659}, {24.99735, 58.16639}, {26.092, 26.092}, {30.092, 41.96426}};
static const double b_nist[5][2] = {{108.4773, 11.26467}, {6.137261, 6.137261}, {55.18696, 2.720074}, {8.218801, 8.218801}, {6.832514, 8.622053}};
static const double c_nist[5][2] = {{-42.52157, -2.114146}, {-1.186521, -1.186521}, {-33.69137, -0.492289}, {-1.976141, -1.976141}, {6.793435, -1.49978}};
static const double d_nist[5][2] = {{5.862788, 0.13819}, {0.09578, 0.09578}, {7.948387, 0.038844}, {0.159274, 0.159274}, {-2.53448, 0.098119}};
static const double e_nist[5][2] = {{0.678565, -26.42221}, {-0.219663, -0.219663}, {-0.136638, -6.447293}, {0.044434, 0.044434}, {0.082139, -11.15764}};
static const double f_nist[5][2] = {{-76.84376, -153.5327}, {-9.861391, -9.861391}, {-403.6075, -425.9186}, {-7.98923, -7.98923}, {-250.881, -272.1797}};
static const double h_nist[5][2] = {{-74.8731, -74.8731}, {0.0, 0.0}, {-393.5224, -393.5224}, {0.0, 0.0}, {-241.8264, -241.8264}};
static const double cutT[5] = {1300.0, 6000.0, 1200.0, 6000.0, 1700.0};
static const int gf_chem_index[5] = {1, 2, 6, 7, 9};
static const int nmax_newton = 10;
static const double tolerance_enth = 1.0e-11;

/* ---- module state (pointers into the caller's arrays) ---- */
static int MP, NP, LP, MPM1, MPM2, NPM1, NPM2, LPM1, LPM2, NR0;
static size_t NN;
static int NJY1, NJY2, LKZ1, LKZ2, NJY1V[4], NJY2V[4], LKZ1V[4], LKZ2V[4];
static double *X_, *DX_, *R_, *DR_, *RS_, *Z_, *DZ_;
static int16_t *IBC;
static double *P_, *T_, *DNST_, *THETA_, *TMU_, *GCP_, *GLAM_, *GAMA_, *GTE_, *CRATE_, *HRATE_, *BSH_;
static double *UG_, *GF_;
static double *RTIM_, *DELRT_, *HRT_, *DELR_, *TIMED_, *HRD_;
static double *FZ_, *BS_, *SFP_, *AP_, *AS_;
static gfm_ctrl *C;
static double SGM = 1.0;
static int NELG = 0;
static double T_new_state = 0.0; /* module variable T_new of enth_to_T */
static double *RMASS_ = NULL;     /* per-cell r_mass of the last SOURCE(5) (-1 where not formed), for diagnosis */
static int STOP_STAGE = 0;        /* probes: 1 = GFLOW returns after the momentum assembly, 2 = after the velocity correction */

#define IX3(i, j, k) ((((size_t)(k)-1) * (size_t)NP + ((size_t)(j)-1)) * (size_t)MP + ((size_t)(i)-1))
#define X(i) X_[(i)-1]
#define DX(i) DX_[(i)-1]
#define R(j) R_[(j)-1]
#define DR(j) DR_[(j)-1]
#define RS(j) RS_[(j)-1]
#define Z(k) Z_[(k)-1]
#define DZ(k) DZ_[(k)-1]
#define IBCELL(i, j, k) ((int)IBC[IX3(i, j, k)])
#define P(i, j, k) P_[IX3(i, j, k)]
#define T(i, j, k) T_[IX3(i, j, k)]
#define DNST(i, j, k) DNST_[IX3(i, j, k)]
#define THETA(i, j, k) THETA_[IX3(i, j, k)]
#define TMU(i, j, k) TMU_[IX3(i, j, k)]
#define GCP(i, j, k) GCP_[IX3(i, j, k)]
#define GLAM(i, j, k) GLAM_[IX3(i, j, k)]
#define GAMA(i, j, k) GAMA_[IX3(i, j, k)]
#define GTE(i, j, k) GTE_[IX3(i, j, k)]
#define CRATE(i, j, k) CRATE_[IX3(i, j, k)]
#define HRATE(i, j, k) HRATE_[IX3(i, j, k)]
#define BSHPREV(i, j, k) BSH_[IX3(i, j, k)]
#define UG(i, j, k, l) UG_[IX3(i, j, k) + NN * (size_t)((l)-1)]
#define GF(i, j, k, l) GF_[IX3(i, j, k) + NN * (size_t)((l)-1)]
#define FZ(i, j, k) FZ_[IX3(i, j, k)]
#define BS(i, j, k) BS_[IX3(i, j, k)]
#define SFP(i, j, k) SFP_[IX3(i, j, k)]
#define AP(i, j, k) AP_[IX3(i, j, k)]
#define AS(i, j, k, d, m) AS_[IX3(i, j, k) + NN * (size_t)(((d)-1) + 3 * ((m)-1))]

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imod10(int a) { return a % 10; } /* Fortran MOD keeps the sign of a, as C % */

/* ======================================================================
 * Geometry, C/Grid3.f90:310-471 (no half-blocked cells)
 * ====================================================================== */
/* AREA_C, C/Grid3.f90:427-471; VOL_C0, C/Grid3.f90:410-421.  Half-blocked
 * cells (IBCELL -1, -2) scale by the &GRD factors BA_F(ix,n) and BV_F(n). */
static double AREA_C(int i, int j, int k, int ix) {
  const int b = IBCELL(i, j, k);
  if (b == 1) return 0.0;
  double area;
  if (ix == 1) area = DR(j) * RS(j) * DZ(k);
  else if (ix == 2) area = DX(i) * DZ(k);
  else area = DX(i) * DR(j) * RS(j);
  if (b == 0) return area;
  if (b <= -1) return area * C->ba_f[(ix - 1) + 3 * (-b - 1)];
  int ibc;
  if (ix == 1) {
    if (i <= MPM2 && IBCELL(i + 2, j, k) <= -1) ibc = -IBCELL(i + 2, j, k);
    else if (i >= 4 && IBCELL(i - 2, j, k) <= -1) ibc = -IBCELL(i - 2, j, k);
    else return area;
  } else if (ix == 2) {
    if (j <= NPM2 && IBCELL(i, j + 2, k) <= -1) ibc = -IBCELL(i, j + 2, k);
    else if (j >= 4 && IBCELL(i, j - 2, k) <= -1) ibc = -IBCELL(i, j - 2, k);
    else return area;
  } else {
    if (k <= LPM2 && IBCELL(i, j, k + 2) <= -1) ibc = -IBCELL(i, j, k + 2);
    else if (k >= 4 && IBCELL(i, j, k - 2) <= -1) ibc = -IBCELL(i, j, k - 2);
    else return area;
  }
  return area * C->ba_f[(ix - 1) + 3 * (ibc - 1)];
}
static double VOL_C0(int i, int j, int k) {
  double vol = DX(i) * DR(j) * RS(j) * DZ(k);
  if (IBCELL(i, j, k) <= -1) vol *= C->bv_f[-IBCELL(i, j, k) - 1];
  return vol;
}
static double VOL_C(int i, int j, int k) {
  double vm, vp;
  if (i & 1) {
    vm = VOL_C0(i - 1, j, k); if (IBCELL(i - 2, j, k) <= 0) vm *= 0.5;
    vp = VOL_C0(i + 1, j, k); if (IBCELL(i + 2, j, k) <= 0) vp *= 0.5;
    return vp + vm;
  } else if (j & 1) {
    vm = VOL_C0(i, j - 1, k); if (IBCELL(i, j - 2, k) <= 0) vm *= 0.5;
    vp = VOL_C0(i, j + 1, k); if (IBCELL(i, j + 2, k) <= 0) vp *= 0.5;
    return vp + vm;
  } else if (k & 1) {
    vm = VOL_C0(i, j, k - 1); if (IBCELL(i, j, k - 2) <= 0) vm *= 0.5;
    vp = VOL_C0(i, j, k + 1); if (IBCELL(i, j, k + 2) <= 0) vp *= 0.5;
    return vp + vm;
  }
  return VOL_C0(i, j, k);
}
static double DXAV(int i, int j, int k) {
  if (IBCELL(i - 2, j, k) >= 1) return X(i + 1) - X(i - 2);
  if (IBCELL(i + 2, j, k) >= 1) return X(i + 2) - X(i - 1);
  return DX(i);
}
static double DRAV(int i, int j, int k) {
  if (IBCELL(i, j - 2, k) >= 1) return (RS(j + 1) + RS(j - 2)) * 0.5 / RS(j) * (R(j + 1) - R(j - 2));
  if (IBCELL(i, j + 2, k) >= 1) return (RS(j + 2) + RS(j - 1)) * 0.5 / RS(j) * (R(j + 2) - R(j - 1));
  return DR(j);
}
static double DZAV(int i, int j, int k) {
  if (IBCELL(i, j, k - 2) >= 1) return Z(k + 1) - Z(k - 2);
  if (IBCELL(i, j, k + 2) >= 1) return Z(k + 2) - Z(k - 1);
  return DZ(k);
}

/* ======================================================================
 * Properties, C/Prop.f90
 * ====================================================================== */
/* DA: exponential-scheme Peclet function, C/Prop.f90:97-117 */
double oracle_da(double d, double f) {
  if (d <= 1.0e-25) return 0.0;
  double fa = fabs(f), pe = fa / d;
  if (pe <= 1.0e-2) return d / (1.0 + pe * (0.5 + pe * (1.0 / 6.0 + pe / 24.0)));
  if (pe <= 20.0) return fa / (exp(pe) - 1.0);
  return 0.0;
}
#define DA oracle_da
/* TGAMF: wall-function factor, C/Gflow.f90:417-426 */
double oracle_tgamf(double b) {
  if (b < 11.5) return 1.0;
  double g0 = b / (log(9.0 * b) * 2.5);
  return dmax(g0, 1.0);
}
/* CONC, C/Prop.f90:28-59 */
static void CONC(int i, int j, int k) {
  double g0 = GF(i, j, k, iyf) + GF(i, j, k, iyo2) + GF(i, j, k, iyco2) + GF(i, j, k, iyn2);
  if (g0 > 1.0) {
    g0 = g0 + GF(i, j, k, iyh2o);
    GF(i, j, k, iyf) /= g0; GF(i, j, k, iyo2) /= g0; GF(i, j, k, iyco2) /= g0;
    GF(i, j, k, iyn2) /= g0; GF(i, j, k, iyh2o) /= g0;
  } else {
    GF(i, j, k, iyh2o) = 1.0 - g0;
  }
}
/* dnst_calc, C/Prop.f90:187-198 */
static void DNST_CALC(int i, int j, int k) {
  CONC(i, j, k);
  double wt = 1.0 / (GF(i, j, k, iyf) / wf + GF(i, j, k, iyco2) / wco2 + GF(i, j, k, iyh2o) / wh2o +
                     GF(i, j, k, iyo2) / wo2 + GF(i, j, k, iyn2) / wn2);
  double temp = (1.0 + P(i, j, k)) / T(i, j, k) * wt / wt0;
  DNST(i, j, k) = dmax(temp, small20);
}
/* THER, C/Prop.f90:130-154 (TURB, AMU=0) */
static void THER(int i, int j, int k) {
  GF(i, j, k, ieps) = dmax(GF(i, j, k, ieps), 1.0e-12);
  double g0 = cmu * C->reyg * C->epsr / (2.0 * C->taur) * GF(i, j, k, ik) * GF(i, j, k, ik) / GF(i, j, k, ieps) * DNST(i, j, k);
  TMU(i, j, k) = g0;
  GLAM(i, j, k) = 1.0;
  GCP(i, j, k) = cpf1;
  GAMA(i, j, k) = 2.0 / C->reyg * (TMU(i, j, k) + 1.0);
}
/* BOUND, C/Prop.f90:243-296 */
static void BOUND(double *f, int il, int ipp) {
  if (ipp == 1) {
    if (il == 0) { *f = dmin(*f, C->p_mx); *f = dmax(*f, C->p_mn); }
    else if (il == 4) { *f = dmin(*f, C->t_mx); *f = dmax(*f, C->t_mn); }
    else if (il == 5) { *f = dmax(*f, 1.0e-6); }
    else { *f = dmin(*f, C->u_mx); *f = dmax(*f, C->u_mn); if (fabs(*f) < 1.0e-12) *f = 0.0; }
  } else if (ipp == 2) {
    if (il == ik) *f = dmax(*f, 1.0e-6);
    else if (il == ieps) *f = dmax(*f, 1.0e-12);
    else if (il == ih) return;
    else if (il == 20) *f = dmax(*f, 1.0e-1);
    else if (il == 8) *f = dmax(*f, 1.0e-14);
    else { *f = dmin(*f, 1.0); *f = dmax(*f, 0.0); }
  }
}
/* Mixture Shomate coefficients with the reference's f_mix/h_mix
 * construction (f_mix = d_mix + Y*F, h_mix = e_mix + Y*H, not
 * accumulated; C/Prop.f90:398-418). */
typedef struct { double a, b, c, d, e, f, h; int lohi[5]; } mixco;
static void mix_add(mixco *m, double y, int n, int s, double sgn) {
  m->a += sgn * y * a_nist[n][s]; m->b += sgn * y * b_nist[n][s]; m->c += sgn * y * c_nist[n][s];
  m->d += sgn * y * d_nist[n][s]; m->e += sgn * y * e_nist[n][s];
  m->f = m->d + sgn * y * f_nist[n][s]; m->h = m->e + sgn * y * h_nist[n][s];
}
static void mix_build(mixco *m, const double y[5], double realT) {
  memset(m, 0, sizeof(*m));
  for (int n = 0; n < 5; n++) {
    int s = (realT < cutT[n]) ? 0 : 1;
    mix_add(m, y[n], n, s, 1.0);
    m->lohi[n] = s + 1;
  }
}
/* enth_to_T: Newton on the Shomate mixture enthalpy, C/Prop.f90:369-505.
 * y = mass fractions (fuel,O2,CO2,N2,H2O), h = normalised enthalpy,
 * tn = normalised temperature (initial guess in, result out). */
double oracle_enth_to_t(const double y[5], double h, double tn, double h_0, double t0) {
  double w_mix = 1.0 / (y[0] / wf + y[2] / wco2 + y[4] / wh2o + y[1] / wo2 + y[3] / wn2);
  double h_s = h * h_0;
  double h_molar = h_s * w_mix / 1.0e6;
  double realT = tn * t0;
  mixco m;
  mix_build(&m, y, realT);
  double T_old = realT / 1000.0, T_new = T_new_state;
  for (int n = 1; n <= nmax_newton; n++) {
    double gp = m.a + T_old * (m.b + T_old * (m.c + T_old * m.d)) + m.e / (T_old * T_old);
    if (gp == 0.0) break;
    double g = T_old * (m.a + T_old * ((m.b / 2.0) + T_old * ((m.c / 3.0) + T_old * (m.d / 4.0)))) - (m.e / T_old) + m.f - m.h;
    double changeT = (g - h_molar) / gp;
    T_new = T_old - changeT;
    if (fabs(changeT) < tolerance_enth) break;
    T_old = T_new;
    for (int nn = 0; nn < 5; nn++) {
      if (nn == 1 || nn == 3) continue;
      if (T_new * 1000.0 < cutT[nn] && m.lohi[nn] != 1) {
        mix_add(&m, y[nn], nn, 1, -1.0);
        m.lohi[nn] = 1;
        mix_add(&m, y[nn], nn, 0, 1.0);
      }
      if (T_new * 1000.0 >= cutT[nn] && m.lohi[nn] != 2) {
        mix_add(&m, y[nn], nn, 0, -1.0);
        m.lohi[nn] = 2;
        mix_add(&m, y[nn], nn, 1, 1.0);
      }
    }
  }
  T_new_state = T_new;
  return T_new * 1000.0 / t0;
}
/* T_to_enth, C/Prop.f90:525-565 */
double oracle_t_to_enth(const double y[5], double tn, double h_0, double t0) {
  double w_mix = 1.0 / (y[0] / wf + y[2] / wco2 + y[4] / wh2o + y[1] / wo2 + y[3] / wn2);
  double realT = tn * t0;
  mixco m;
  mix_build(&m, y, realT);
  double t = realT / 1000.0;
  double ef = t * (m.a + t * ((m.b / 2.0) + t * ((m.c / 3.0) + t * (m.d / 4.0)))) - (m.e / t) + m.f - m.h;
  return ef * 1.0e6 / (w_mix * h_0);
}
static void ygather(int i, int j, int k, double y[5]) {
  for (int n = 0; n < 5; n++) y[n] = GF(i, j, k, gf_chem_index[n]);
}
static void ENTH_TO_T(int i, int j, int k) {
  double y[5]; ygather(i, j, k, y);
  T(i, j, k) = oracle_enth_to_t(y, GF(i, j, k, ih), T(i, j, k), C->h_0, C->t0);
}
static void T_TO_ENTH(int i, int j, int k) {
  double y[5]; ygather(i, j, k, y);
  GF(i, j, k, ih) = oracle_t_to_enth(y, T(i, j, k), C->h_0, C->t0);
}
/* ZY1: table interpolation, C/Source.f90:814-831 */
double oracle_zy1(const double *x, const double *y, double x1, int n) {
  if (x1 < x[0]) return y[0] * x1 / x[0];
  for (int i = 2; i <= n; i++)
    if (x1 < x[i - 1]) return y[i - 2] + (y[i - 1] - y[i - 2]) * ((x1 - x[i - 2]) / (x[i - 1] - x[i - 2]));
  return y[n - 1];
}

/* ======================================================================
 * Momentum assembly helpers, C/Gslv3x.f90 and VINDX C/Dflow.f90:189-219
 * ====================================================================== */
static void VINDX(int iu, int m, int i, int j, int k, int *i1, int *j1, int *k1, int *i2, int *j2, int *k2) {
  int nx[3] = {0, 0, 0};
  nx[iu - 1] = m;
  *i1 = i + nx[0]; *i2 = *i1 + nx[0];
  *j1 = j + nx[1]; *j2 = *j1 + nx[1];
  *k1 = k + nx[2]; *k2 = *k1 + nx[2];
}
static double FLW(int i, int j, int k, int ix, int m) {
  double f = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, ix) * AREA_C(i, j, k, ix);
  return (m == 1) ? dmax(0.0, f) : dmax(0.0, -f);
}
static double FLWD(int i, int j, int k, int ix, int m) {
  double area = AREA_C(i, j, k, ix);
  double ddl = (ix == 1) ? DX(i) : (ix == 2) ? DR(j) : DZ(k);
  double f = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, ix) * area;
  double diff = THETA(i, j, k) * GAMA(i, j, k) / SGM * area / ddl;
  return (m == 1) ? DA(diff, f) + dmax(0.0, f) : DA(diff, f) + dmax(0.0, -f);
}
static double FDDL(int i, int j, int k, int iu) {
  int b = IBCELL(i, j, k);
  if (iu == 1) {
    if (b <= 0) return DX(i);
    if (IBCELL(i + 1, j, k) == 0) return X(i + 1) - X(i);
    if (IBCELL(i - 1, j, k) == 0) return X(i) - X(i - 1);
    return big;
  } else if (iu == 2) {
    if (b <= 0) return DR(j) * RS(j);
    if (IBCELL(i, j + 1, k) == 0) return (R(j + 1) - R(j)) * RS(j);
    if (IBCELL(i, j - 1, k) == 0) return (R(j) - R(j - 1)) * RS(j);
    return big;
  } else {
    if (b <= 0) return DZ(k);
    if (IBCELL(i, j, k + 1) == 0) return Z(k + 1) - Z(k);
    if (IBCELL(i, j, k - 1) == 0) return Z(k) - Z(k - 1);
    return big;
  }
}
/* ECELL, C/Gslv3x.f90:148-200 (LP>=6) ; FFL[d][m], d,m 0-based */
static void ECELL(int i, int j, int k, int iu, double ffl[3][2]) {
  if (iu == 1) {
    for (int m = 0; m < 2; m++) {
      int i1 = m ? i + 1 : i - 1, i2 = m ? i + 2 : i - 2;
      ffl[0][m] = (IBCELL(i2, j, k) >= 1) ? DX(i1) : fabs(X(i) - X(i1));
      ffl[1][m] = ffl[0][m] * DZ(k);
      ffl[2][m] = ffl[0][m] * RS(j) * DR(j);
    }
  } else if (iu == 2) {
    for (int m = 0; m < 2; m++) {
      int j1 = m ? j + 1 : j - 1, j2 = m ? j + 2 : j - 2;
      ffl[1][m] = (IBCELL(i, j2, k) >= 1) ? RS(j1) * DR(j1) : RS(j1) * fabs(R(j) - R(j1));
      ffl[0][m] = ffl[1][m] * DZ(k);
      ffl[2][m] = ffl[1][m] * DX(i);
    }
  } else {
    for (int m = 0; m < 2; m++) {
      int k1 = m ? k + 1 : k - 1, k2 = m ? k + 2 : k - 2;
      ffl[2][m] = (IBCELL(i, j, k2) >= 1) ? DZ(k1) : fabs(Z(k) - Z(k1));
      ffl[0][m] = ffl[2][m] * RS(j) * DR(j);
      ffl[1][m] = ffl[2][m] * DX(i);
    }
  }
}
/* VOL1, C/Gslv3x.f90:266-286 */
static double VOL1(int i, int j, int k, int ix, double ffl[3][2]) {
  double area = AREA_C(i, j, k, ix);
  double vol = (ffl[ix - 1][0] + ffl[ix - 1][1]) * area;
  if (ix == 1) {
    if (IBCELL(i - 1, j, k) <= -1) vol = vol - DX(i - 1) * area / 2.0;
    if (IBCELL(i + 1, j, k) <= -1) vol = vol - DX(i + 1) * area / 2.0;
  } else if (ix == 2) {
    if (IBCELL(i, j - 1, k) <= -1) vol = vol - RS(j - 1) * DR(j - 1) * area / 2.0;
    if (IBCELL(i, j + 1, k) <= -1) vol = vol - RS(j + 1) * DR(j + 1) * area / 2.0;
  } else {
    if (IBCELL(i, j, k - 1) <= -1) vol = vol - DZ(k - 1) * area / 2.0;
    if (IBCELL(i, j, k + 1) <= -1) vol = vol - DZ(k + 1) * area / 2.0;
  }
  return vol;
}
/* GSLV3X1, C/Gslv3x.f90:119-139 */
static double GSLV3X1(int i, int j, int k, int iu) {
  double g0 = 0.0;
  for (int m = -1; m <= 1; m += 2) {
    int m1 = (m + 3) / 2, i1, j1, k1, i2, j2, k2;
    VINDX(iu, m, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2);
    int ibc1 = IBCELL(i2, j2, k2);
    if (ibc1 == 3) AS(i, j, k, iu, m1) = 0.0;
    else if (ibc1 == 2) AS(i, j, k, iu, m1) = FLW(i2, j2, k2, iu, m1);
    else AS(i, j, k, iu, m1) = FLWD(i1, j1, k1, iu, m1);
    AP(i, j, k) += AS(i, j, k, iu, m1);
    g0 += AS(i, j, k, iu, m1) * FZ(i2, j2, k2);
  }
  return g0;
}
/* SRC1: viscous and buoyancy source of momentum, C/Source.f90:664-809 (NDNP=0) */
static void SRC1(int i, int j, int k, int iu1) {
  if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = 0.0; SFP(i, j, k) = 0.0; return; }
  double vis[3][2] = {{0}}, dxn[3], dx1[3] = {0}, dx0[3] = {0};
  dxn[0] = DX(i); dxn[1] = DR(j) * RS(j); dxn[2] = DZ(k);
  for (int m = -1; m <= 1; m += 2) {
    int i1, j1, k1, i2, j2, k2;
    VINDX(iu1, m, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2);
    if (iu1 == 1) { dx0[0] = DX(i); dx1[0] = DX(i1); }
    else if (iu1 == 2) { dx0[1] = DR(j) * RS(j); dx1[1] = DR(j1) * RS(j1); }
    else { dx0[2] = DZ(k); dx1[2] = DZ(k1); }
    double thgm;
    if (IBCELL(i2, j2, k2) == 2) {
      thgm = THETA(i2, j2, k2) * GAMA(i2, j2, k2);
      if (iu1 == 1) dxn[0] += fabs(X(i2) - X(i1));
      if (iu1 == 2) dxn[1] += fabs(R(j2) - R(j1)) * RS(j);
      if (iu1 == 3) dxn[2] += fabs(Z(k2) - Z(k1));
    } else {
      thgm = THETA(i1, j1, k1) * GAMA(i1, j1, k1);
    }
    if (m == -1) vis[iu1 - 1][0] = thgm * (FZ(i, j, k) - FZ(i2, j2, k2)) / dx1[iu1 - 1];
    else vis[iu1 - 1][1] = thgm * (FZ(i2, j2, k2) - FZ(i, j, k)) / dx1[iu1 - 1];
  }
  double vms = (vis[iu1 - 1][1] - vis[iu1 - 1][0]) / dxn[iu1 - 1];
  for (int iu = 1; iu <= 3; iu++) {
    if (iu == iu1) continue;
    int i1 = i, j1 = j, k1 = k;
    for (int m = -1; m <= 1; m += 2) {
      if (iu == 1) i1 = i + m; else if (iu == 2) j1 = j + m; else k1 = k + m;
      double du0;
      if (iu1 == 1) du0 = FZ(i + 1, j1, k1) - FZ(i - 1, j1, k1);
      else if (iu1 == 2) du0 = FZ(i1, j + 1, k1) - FZ(i1, j - 1, k1);
      else du0 = FZ(i1, j1, k + 1) - FZ(i1, j1, k - 1);
      int m1 = (m + 3) / 2;
      vis[iu - 1][m1 - 1] = THETA(i1, j1, k1) * GAMA(i1, j1, k1) * du0 / dx0[iu1 - 1];
      if (iu == 2) vis[iu - 1][m1 - 1] *= RS(j1);
    }
    vms += (vis[iu - 1][1] - vis[iu - 1][0]) / dxn[iu - 1];
  }
  BS(i, j, k) = vms + 0.0 + 0.0;
  BS(i, j, k) = BS(i, j, k) + THETA(i, j, k) * DNST(i, j, k) * C->gra[iu1 - 1];
  SFP(i, j, k) = -0.0;
}

/* ======================================================================
 * Gray radiation source and wall convective flux (reference extension
 * used when ctrl.gray_radiation / ctrl.isothermal_walls are set)
 * ====================================================================== */
/* EMF: band emissivity of a path, C/Rad_fun.f90:97-116 */
double oracle_emf(double xkl, double xl0) {
  const double x1 = 1.0e-3;
  const double g1 = (1.0 - exp(-x1)) / x1;
  double x0 = xkl * xl0;
  return (x0 <= x1) ? x0 * g1 : 1.0 - exp(-x0);
}
/* Net gray radiation into an open cell (W): emission of the gray gas
 * QE = SA*sigma*T^4*EMF(kappa,OPL) with OPL = 6V/SA (C/Rad0.f90:112-157,
 * one band, DEB=1) and absorption from an isothermal black enclosure at
 * T_wall with the same emissivity (Kirchhoff), so QA-QE =
 * SA*sigma*EMF*(Tw^4 - T^4).  Cells below 300 K do not emit
 * (TG_MN, C/Rad0.f90:42). */
/* QA and QE (cell centres, W) of the enthalpy source GFM_RADSRC_FIELD; set by oracle_rad_field */
static const double *RADF_QA = NULL, *RADF_QE = NULL;
/* q_conv of each wall and melt patch at its face (lattice, 0 elsewhere); set by oracle_wall_field */
static const double *WQCF_ = NULL;
void oracle_wall_field(const double *qcf) { WQCF_ = qcf; }
void oracle_rad_field(const double *qa, const double *qe) { RADF_QA = qa; RADF_QE = qe; }
static double gray_qnet(int i, int j, int k) {
  double tk = T(i, j, k) * C->t0;
  double vol = VOL_C(i, j, k);
  double sa = ((DX(i) * DR(j)) + (DR(j) * DZ(k)) + (DX(i) * DZ(k))) * 2.0;
  if (sa <= small16) return 0.0;
  double opl = 6.0 * vol / sa;
  if (opl <= 1.0e-10) return 0.0;
  double e = oracle_emf(C->kappa_gray, opl);
  double tw = C->t_wall * C->t0;
  double qe = (tk >= 300.0) ? sa * C->sigma_sb * tk * tk * tk * tk * e : 0.0;
  double qa = sa * C->sigma_sb * tw * tw * tw * tw * e;
  return qa - qe;
}
/* Convective heat flux (W/m2) from the gas cell (i,j,k) to a wall or melt
 * surface cell at temperature twk (K) across the face at distance ds
 * (calc_q_conv_wall, C/Qwall.f90:1419-1467). */
static double q_conv_face(int i, int j, int k, double ds, double twk) {
  const double kappa_cmu = 0.4 * pow(9.0e-2, 0.25), e_cmu = 9.0 * pow(9.0e-2, 0.25);
  const double pi = 3.141592653589793;
  const double c_htw = 0.4 * ((pi / 4) / sin(pi / 4)) * sqrt(26.0) * (0.7 - 1.0) / pow(0.7, 0.25);
  double cpmix = 1.0057e3 * (cpf1 * GF(i, j, k, iyf) + (2.0e3 / 1.0057e3) * GF(i, j, k, iyco2) +
                             (2.0e3 / 1.0057e3) * GF(i, j, k, iyh2o) + (0.9136e3 / 1.0057e3) * GF(i, j, k, iyo2) +
                             (0.526e3 / 1.0057e3) * GF(i, j, k, iyn2));
  double tgas = T(i, j, k) * C->t0;
  double gmu_t = ((2.0e-15 * tgas + -1.1e-11) * tgas + 4.5e-8) * tgas + 7.0e-6;
  double h_t;
  if (C->tk0 * GF(i, j, k, ik) < 0.1) h_t = 0.0;
  else {
    double sq = sqrt(C->tk0 * GF(i, j, k, ik));
    double dtmp = log(e_cmu * ds * C->dnst0 * DNST(i, j, k) * sq / gmu_t) + (c_htw / sq);
    h_t = (dtmp < 0.01) ? 0.0 : dmin(50.0, cpmix * kappa_cmu * C->dnst0 * DNST(i, j, k) * sq / dtmp);
  }
  return dmax(C->h_g, h_t) * (tgas - twk);
}
/* Wall sink of one open cell: sum over its wall/melt neighbours of
 * q_conv/(ds*r0*qr0) (C/Source.f90:473-509); *qw accumulates q_conv*area. */
static double wall_sink(int i, int j, int k, double *qw) {
  double s = 0.0;
  for (int d = 0; d < 6; d++) {
    int di = (d == 0) ? -2 : (d == 1) ? 2 : 0, dj = (d == 2) ? -2 : (d == 3) ? 2 : 0, dk = (d == 4) ? -2 : (d == 5) ? 2 : 0;
    int b = IBCELL(i + di, j + dj, k + dk);
    if (b != 1 && b != 4) continue;
    double dsw, dsc, area;
    if (di) { dsw = fabs(X(i + di / 2) - X(i)); dsc = DX(i); area = DR(j) * DZ(k); }
    else if (dj) { dsw = fabs(R(j + dj / 2) - R(j)); dsc = DR(j); area = DX(i) * DZ(k); }
    else { dsw = fabs(Z(k + dk / 2) - Z(k)); dsc = DZ(k); area = DX(i) * DR(j); }
    double q = q_conv_face(i, j, k, dsw, T(i + di, j + dj, k + dk) * C->t0);
    *qw += q * area;
    if (d == 4) C->q_wall_bottom += q * area;
    if (d == 5) C->q_wall_top += q * area;
    s += q / (dsc * C->r0 * C->qr0);
  }
  return s;
}

/* ======================================================================
 * Source terms, C/Source.f90:6-658
 * ====================================================================== */
static void SOURCE(int mi1, int mi3, int nj1, int nj3, int lk1, int lk3) {
  int nelg = NELG;
  if (nelg >= 1 && nelg <= 3) {
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) SRC1(i, j, k, nelg);
  } else if (nelg == 4 || nelg == 5) {
    if (nelg == 5) {
      C->res_mass = 0.0; C->res_mass_max = 0.0; C->imx = 2; C->jmx = 2; C->kmx = 2;
      if (RMASS_) for (size_t q = 0; q < NN; q++) RMASS_[q] = -1.0;
    }
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          double ax = AREA_C(i, j, k, 1);
          double frx1 = (THETA(i - 1, j, k) * DNST(i - 1, j, k) * UG(i - 1, j, k, 1)) * ax;
          double frx2 = (THETA(i + 1, j, k) * DNST(i + 1, j, k) * UG(i + 1, j, k, 1)) * ax;
          double ay = AREA_C(i, j, k, 2);
          double fry1 = (THETA(i, j - 1, k) * DNST(i, j - 1, k) * UG(i, j - 1, k, 2)) * ay;
          double fry2 = (THETA(i, j + 1, k) * DNST(i, j + 1, k) * UG(i, j + 1, k, 2)) * ay;
          double az = AREA_C(i, j, k, 3);
          double frz1 = (THETA(i, j, k - 1) * DNST(i, j, k - 1) * UG(i, j, k - 1, 3)) * az;
          double frz2 = (THETA(i, j, k + 1) * DNST(i, j, k + 1) * UG(i, j, k + 1, 3)) * az;
          double rx = frx1 - frx2, ry = fry1 - fry2, rz = frz1 - frz2, evms = 0.0;
          BS(i, j, k) = (rx + ry + rz + evms) / C->eun;
          SFP(i, j, k) = 0.0;
          if (nelg == 5) {
            double r_mass;
            if (IBCELL(i - 1, j, k) == 3 || IBCELL(i + 1, j, k) == 3 || IBCELL(i, j - 1, k) == 3 ||
                IBCELL(i, j + 1, k) == 3 || IBCELL(i, j, k - 1) == 3 || IBCELL(i, j, k + 1) == 3)
              continue;
            double fmax = dmax(dmax(dmax(fabs(frx1), fabs(frx2)), dmax(fabs(fry1), fabs(fry2))),
                               dmax(dmax(fabs(frz1), fabs(frz2)), evms));
            r_mass = (fmax == 0.0) ? 0.0 : fabs(BS(i, j, k)) * C->eun / fmax;
            if (RMASS_) RMASS_[IX3(i, j, k)] = r_mass;
            if (C->res_mass_max < r_mass) { C->imx = i; C->jmx = j; C->kmx = k; C->res_mass_max = r_mass; }
            C->res_mass += r_mass;
          }
        }
    if (nelg == 5) C->res_mass /= C->ncells;
  } else if (nelg == 6) {
    /* fuel: integral reaction submodel, C/Source.f90:104-242 */
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); SFP(i, j, k) = 0.0; CRATE(i, j, k) = 0.0; HRATE(i, j, k) = 0.0; continue; }
          double sc = 0.0;
          double d11 = fabs(UG(i + 1, j, k, 1) / DX(i)), d12 = fabs(UG(i - 1, j, k, 1) / DX(i));
          double d21 = fabs(UG(i, j + 1, k, 2) / DR(j)), d22 = fabs(UG(i, j - 1, k, 2) / DR(j));
          double d31 = fabs(UG(i, j, k + 1, 3) / DZ(k)), d32 = fabs(UG(i, j, k - 1, 3) / DZ(k));
          double dtime = dmax(dmax(dmax(d11, d12), dmax(d21, d22)), dmax(dmax(d31, d32), 1.0e-14));
          dtime = 1.0 / dtime;
          double drtime = dtime * C->r0 / C->ug0;
          double ayf = dmin(GF(i, j, k, iyf), GF(i, j, k, iyo2) / stoich);
          if (ayf < 1.0e-5) { CRATE(i, j, k) = 0.0; HRATE(i, j, k) = 0.0; SFP(i, j, k) = 0.0; BS(i, j, k) = 0.0; continue; }
          double zp1 = GF(i, j, k, iyco2) / wco2, delr1;
          if (zp1 <= small20) delr1 = 0.0;
          else { double zf = GF(i, j, k, iyf) / wf; delr1 = zp1 / (zf + zp1); }
          double time1 = oracle_zy1(DELR_, TIMED_, delr1, NR0);
          double time2 = time1 + drtime;
          double delr2 = oracle_zy1(RTIM_, DELRT_, time2, NR0);
          double ayf0 = ayf / (1.0 - delr1) * DNST(i, j, k) * THETA(i, j, k);
          double sr = ayf0 * (delr2 - delr1) / dtime;
          if (sr <= 0.0) { CRATE(i, j, k) = 0.0; HRATE(i, j, k) = 0.0; SFP(i, j, k) = 0.0; BS(i, j, k) = sc; continue; }
          CRATE(i, j, k) = sr;
          SFP(i, j, k) = -CRATE(i, j, k) / (GF(i, j, k, iyf) + small20);
          BS(i, j, k) = sc;
          HRATE(i, j, k) = CRATE(i, j, k) * C->q0;
        }
  } else if (nelg == 7 || nelg == 13) {
    double g0 = (nelg == 7) ? wco2 / wf : wh2o * 2.0 / wf;
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          SFP(i, j, k) = 0.0;
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); continue; }
          BS(i, j, k) = CRATE(i, j, k) * g0;
        }
  } else if (nelg == 8 || nelg == 9) {
    if (nelg == 8) {
      C->q_react = 0.0;
      C->q_wall = 0.0;
      C->q_wall_bottom = 0.0;
      C->q_wall_top = 0.0;
      for (int i = mi1; i <= mi3; i += 2)
        for (int j = nj1; j <= nj3; j += 2)
          for (int k = lk1; k <= lk3; k += 2) {
            if (IBCELL(i, j, k) >= 1) continue;
            double dudx, dvdx, dwdx, dudr, dvdr, dwdr, dudz, dvdz, dwdz;
            if (IBCELL(i + 1, j, k) == 3 || IBCELL(i - 1, j, k) == 3) { dudx = dvdx = dwdx = 0.0; }
            else {
              dudx = (UG(i + 1, j, k, 1) - UG(i - 1, j, k, 1)) / DX(i);
              dvdx = (UG(i + 1, j, k, 2) - UG(i - 1, j, k, 2)) / DX(i);
              dwdx = (UG(i + 1, j, k, 3) - UG(i - 1, j, k, 3)) / DX(i);
            }
            if (nj3 < 6 || IBCELL(i, j + 1, k) == 3 || IBCELL(i, j - 1, k) == 3) { dudr = dvdr = dwdr = 0.0; }
            else {
              dudr = (UG(i, j + 1, k, 1) - UG(i, j - 1, k, 1)) / DR(j);
              dvdr = (UG(i, j + 1, k, 2) - UG(i, j - 1, k, 2)) / DR(j);
              dwdr = (UG(i, j + 1, k, 3) - UG(i, j - 1, k, 3)) / DR(j);
            }
            if (lk3 < 6 || IBCELL(i, j, k + 1) == 3 || IBCELL(i, j, k - 1) == 3) { dudz = dvdz = dwdz = 0.0; }
            else {
              dudz = (UG(i, j, k + 1, 1) - UG(i, j, k - 1, 1)) / DZ(k);
              dvdz = (UG(i, j, k + 1, 2) - UG(i, j, k - 1, 2)) / DZ(k);
              dwdz = (UG(i, j, k + 1, 3) - UG(i, j, k - 1, 3)) / DZ(k);
            }
            double vr = 0.0;
            GTE(i, j, k) = 2.0 * (dudx * dudx + dvdr * dvdr + vr * vr + dwdz * dwdz) + (dudr + dvdx) * (dudr + dvdx) +
                           (dwdr + dvdz) * (dwdr + dvdz) + (dudz + dwdx) * (dudz + dwdx) -
                           (dudx + dvdr + dwdz) * (dudx + dvdr + dwdz) / 1.5;
          }
      for (int i = mi1; i <= mi3; i += 2)
        for (int j = nj1; j <= nj3; j += 2)
          for (int k = lk1; k <= lk3; k += 2) {
            if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); SFP(i, j, k) = 0.0; continue; }
            double cpmix = cpf1 * GF(i, j, k, iyf) + (2.0e3 / 1.0057e3) * GF(i, j, k, iyco2) +
                           (2.0e3 / 1.0057e3) * GF(i, j, k, iyh2o) + (0.9136e3 / 1.0057e3) * GF(i, j, k, iyo2) +
                           (0.526e3 / 1.0057e3) * GF(i, j, k, iyn2);
            BS(i, j, k) = 0.0 + HRATE(i, j, k);
            SFP(i, j, k) = -0.0 / cpmix;
            double vol = VOL_C(i, j, k) * C->vol0;
            C->q_react += HRATE(i, j, k) * C->qr0 * vol;
            if (C->gray_radiation == GFM_RADSRC_GRAY) BS(i, j, k) += gray_qnet(i, j, k) / (vol * C->qr0);
            else if (C->gray_radiation == GFM_RADSRC_FIELD && RADF_QA) {
              /* q_rad = QA - QE of the last radiation pass, C/Source.f90:437-447 */
              const size_t cc = (((size_t)k / 2 - 1) * (size_t)(NP / 2) + ((size_t)j / 2 - 1)) * (size_t)(MP / 2) + ((size_t)i / 2 - 1);
              BS(i, j, k) += (RADF_QA[cc] - RADF_QE[cc]) / (vol * C->qr0);
            }
            if (C->isothermal_walls) BS(i, j, k) -= wall_sink(i, j, k, &C->q_wall);
            else if (C->i1st_rad_done > 0 && WQCF_) {
              /* the patch q_conv of the last wall exchange, C/Source.f90:473-509: the patches of one
               * open cell in patch order are its -x, +x, -y, +y, -z, +z faces; 0 where none */
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i - 1, j, k)] / (DX(i) * C->r0 * C->qr0);
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i + 1, j, k)] / (DX(i) * C->r0 * C->qr0);
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i, j - 1, k)] / (DR(j) * C->r0 * C->qr0);
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i, j + 1, k)] / (DR(j) * C->r0 * C->qr0);
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i, j, k - 1)] / (DZ(k) * C->r0 * C->qr0);
              BS(i, j, k) = BS(i, j, k) - WQCF_[IX3(i, j, k + 1)] / (DZ(k) * C->r0 * C->qr0);
            }
          }
      if (C->itr_gas > 1) {
        for (int i = mi1; i <= mi3; i += 2)
          for (int j = nj1; j <= nj3; j += 2)
            for (int k = lk1; k <= lk3; k += 2) {
              if (IBCELL(i, j, k) >= 1) continue;
              BS(i, j, k) = C->rf_hsource * BS(i, j, k) + (1.0 - C->rf_hsource) * BSHPREV(i, j, k);
            }
      }
      /* bs_hprev=bs: only cell-centre entries are ever read back */
      for (int i = 2; i <= MP; i += 2)
        for (int j = 2; j <= NP; j += 2)
          for (int k = 2; k <= LP; k += 2) BSHPREV(i, j, k) = BS(i, j, k);
      return;
    }
    /* k: C/Source.f90:524-572 (GTE from the enthalpy call, LSTAR<=3) */
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); SFP(i, j, k) = 0.0; continue; }
          BS(i, j, k) = 2.0 / (C->reyg * C->epsr) * THETA(i, j, k) * TMU(i, j, k) * GTE(i, j, k) - 0.0;
          SFP(i, j, k) = -C->taur * THETA(i, j, k) * DNST(i, j, k) * GF(i, j, k, ieps) / (GF(i, j, k, ik) + small20);
        }
  } else if (nelg == 10) {
    double cbl0 = 2.5 * sqrt(C->epsr) / C->taur * pow(cmu, 0.75);
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          SFP(i, j, k) = 0.0;
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); continue; }
          double cbl = cbl0 * pow(GF(i, j, k, ik), 1.5);
          BS(i, j, k) = 0.0;
          if (MP > 8 && IBCELL(i + 1, j, k) == 1) { BS(i, j, k) += big * cbl / (X(i + 1) - X(i)); SFP(i, j, k) = -big; }
          if (MP > 8 && IBCELL(i - 1, j, k) == 1) { BS(i, j, k) += big * cbl / (X(i) - X(i - 1)); SFP(i, j, k) = -big; }
          if (NP > 8 && IBCELL(i, j + 1, k) == 1) { BS(i, j, k) += big * cbl / (R(j + 1) - R(j)); SFP(i, j, k) = -big; }
          if (NP > 8 && IBCELL(i, j - 1, k) == 1) { BS(i, j, k) += big * cbl / (R(j) - R(j - 1)); SFP(i, j, k) = -big; }
          if (LP > 8 && IBCELL(i, j, k + 1) == 1) { BS(i, j, k) += big * cbl / (Z(k + 1) - Z(k)); SFP(i, j, k) = -big; }
          if (LP > 8 && IBCELL(i, j, k - 1) == 1) { BS(i, j, k) += big * cbl / (Z(k) - Z(k - 1)); SFP(i, j, k) = -big; }
          if (SFP(i, j, k) <= -big) continue;
          double bt = THETA(i, j, k);
          BS(i, j, k) = ct1 * cmu * bt / C->taur * DNST(i, j, k) * GTE(i, j, k) * GF(i, j, k, ik);
          SFP(i, j, k) = -ct2 * C->taur * bt * DNST(i, j, k) * GF(i, j, k, ieps) / (GF(i, j, k, ik) + small20);
        }
  } else if (nelg == 11) {
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); SFP(i, j, k) = 0.0; continue; }
          BS(i, j, k) = 0.0; SFP(i, j, k) = 0.0;
        }
  } else if (nelg == 12) {
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          SFP(i, j, k) = 0.0;
          if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = FZ(i, j, k); continue; }
          BS(i, j, k) = -CRATE(i, j, k) * stoich;
        }
  }
}

/* ======================================================================
 * Line relaxation, C/Adlbl3.f90
 * ====================================================================== */
static double *XT2_, *XT3_; /* (NP,LP) plane buffers */
#define XT2(j, k) XT2_[((size_t)(k)-1) * NP + ((j)-1)]
#define XT3(j, k) XT3_[((size_t)(k)-1) * NP + ((j)-1)]
#define XTA(i, j, k) xt[IX3(i, j, k)]

/* ADLZY: TDMA along y at plane i, line k, C/Adlbl3.f90:147-181 */
static void ADLZY(double *xt, int i, int jb, int je, int k, int ipl, int iml, int jbml) {
  static double *pt = NULL, *qt = NULL;
  static int cap = 0;
  if (cap < NP + 2) { free(pt); free(qt); cap = NP + 2; pt = malloc(sizeof(double) * cap); qt = malloc(sizeof(double) * cap); }
  const int ld = 2;
  int kpl = k + ld, kml = k - ld;
  pt[jbml] = 0.0; qt[jbml] = XTA(i, jbml, k);
  for (int j = jb; j <= je; j += ld) {
    int jml = j - ld;
    double alp = AP(i, j, k) - pt[jml] * AS(i, j, k, 2, 1);
    double bsy = BS(i, j, k);
    if (fabs(alp) <= small20 || IBCELL(i, j, k) >= 1) { pt[j] = 0.0; qt[j] = XTA(i, j, k); }
    else {
      pt[j] = AS(i, j, k, 2, 2) / alp;
      bsy = bsy + AS(i, j, k, 1, 2) * XTA(ipl, j, k) + AS(i, j, k, 1, 1) * XTA(iml, j, k);
      if (LP >= 6) bsy = bsy + AS(i, j, k, 3, 2) * XTA(i, j, kpl) + AS(i, j, k, 3, 1) * XTA(i, j, kml);
      qt[j] = (bsy + AS(i, j, k, 2, 1) * qt[jml]) / alp;
    }
  }
  for (int j = je; j >= jb; j -= ld) XT2(j, k) = XT2(j + ld, k) * pt[j] + qt[j];
}
/* ADLYZ: TDMA along z at plane i, line j, C/Adlbl3.f90:187-221 */
static void ADLYZ(double *xt, int i, int j, int kb, int ke, int ipl, int iml) {
  static double *pt = NULL, *qt = NULL;
  static int cap = 0;
  if (cap < LP + 2) { free(pt); free(qt); cap = LP + 2; pt = malloc(sizeof(double) * cap); qt = malloc(sizeof(double) * cap); }
  const int ld = 2;
  int jpl = j + ld, jml = j - ld, kbml = kb - ld;
  pt[kbml] = 0.0; qt[kbml] = XTA(i, j, kbml);
  for (int k = kb; k <= ke; k += ld) {
    int kml = k - ld;
    double alp = AP(i, j, k) - pt[kml] * AS(i, j, k, 3, 1);
    double bsy = BS(i, j, k);
    if (fabs(alp) <= small20 || IBCELL(i, j, k) >= 1) { pt[k] = 0.0; qt[k] = XTA(i, j, k); }
    else {
      pt[k] = AS(i, j, k, 3, 2) / alp;
      bsy = bsy + AS(i, j, k, 1, 2) * XTA(ipl, j, k) + AS(i, j, k, 1, 1) * XTA(iml, j, k);
      if (NP >= 6) bsy = bsy + AS(i, j, k, 2, 2) * XTA(i, jpl, k) + AS(i, j, k, 2, 1) * XTA(i, jml, k);
      qt[k] = (bsy + AS(i, j, k, 3, 1) * qt[kml]) / alp;
    }
  }
  for (int k = ke; k >= kb; k -= ld) XT3(j, k) = XT3(j, k + ld) * pt[k] + qt[k];
}
/* ADLZ: one plane visit, C/Adlbl3.f90:103-141 */
static void ADLZ(double *xt, int i, int jb, int je, int kb, int ke) {
  const int ld = 2;
  int kepl = ke + ld, jbml = jb - ld, jepl = je + ld, ipl = i + ld, iml = i - ld;
  for (int k = kb; k <= kepl; k += ld)
    for (int j = jb; j <= jepl; j += ld) { XT2(j, k) = XTA(i, j, k); XT3(j, k) = XTA(i, j, k); }
  for (int l = 1; l <= 2; l++) {
    for (int k = kb; k <= ke; k += ld) ADLZY(xt, i, jb, je, k, ipl, iml, jbml);
    for (int j = jb; j <= je; j += ld) ADLYZ(xt, i, j, kb, ke, ipl, iml);
    for (int k = kb; k <= ke; k += ld)
      for (int j = jb; j <= je; j += ld) XTA(i, j, k) = (XT2(j, k) + XT3(j, k)) / 2.0;
  }
}
/* ADLBL3, C/Adlbl3.f90:10-97.  order selects the x-plane visiting order. */
static void ADLBL3_SWEEPS(double *xt, int nim, int ib, int ie, int jb, int je, int kb, int ke) {
  const int ld = 2;
  for (int ns = 1; ns <= nim; ns++) {
    if (C->sweep_order == GFM_ORDER_REFERENCE) {
      for (int i = ib; i <= ie; i += ld) ADLZ(xt, i, jb, je, kb, ke);
      for (int i = ie; i >= ib; i -= ld) ADLZ(xt, i, jb, je, kb, ke);
    } else {
      for (int half = 0; half < 4; half++) {
        int colour = half & 1;
        for (int i = ib + colour * ld; i <= ie; i += 2 * ld) ADLZ(xt, i, jb, je, kb, ke);
      }
    }
  }
}
/* ADLBL3 of a gas-phase equation: NTIMESG(NELG) sweeps, then the IRC
 * residual block (RESG). */
static void ADLBL3(double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  const int ld = 2;
  ADLBL3_SWEEPS(xt, C->ntimesg[NELG - 1], ib, ie, jb, je, kb, ke);
  if (C->want_irc) {
    double res0 = 0.0, resx = 0.0;
    for (int i = ib; i <= ie; i += ld)
      for (int j = jb; j <= je; j += ld)
        for (int k = kb; k <= ke; k += ld) {
          if (AP(i, j, k) == 1.0 || IBCELL(i, j, k) > 0) continue;
          double resc = AS(i, j, k, 1, 2) * XTA(i + ld, j, k) + AS(i, j, k, 1, 1) * XTA(i - ld, j, k) +
                        AS(i, j, k, 2, 2) * XTA(i, j + ld, k) + AS(i, j, k, 2, 1) * XTA(i, j - ld, k) +
                        AS(i, j, k, 3, 2) * XTA(i, j, k + ld) + AS(i, j, k, 3, 1) * XTA(i, j, k - ld) +
                        BS(i, j, k) - AP(i, j, k) * XTA(i, j, k);
          resc = fabs(resc);
          resx = dmax(resx, resc);
          res0 += resc;
        }
    res0 /= C->ncells;
    C->resg[NELG - 1] = res0;
    C->resg[20 + NELG - 1] = resx;
  }
}
/* calc_residual, C/Adlbl3.f90:239-271 */
/* FLX1D, C/Flx1d.f90:16-73: the gas mass flow through each x-face plane
 * I = 3..MP-1 (over J = 2..NP, K = 2..LP), y-face plane J = 3..NP-1 (over
 * I = 4..MP-2, K = LKZ1..LKZ2) and z-face plane K = 3..LP-1 (over
 * I = 4..MP-2, J = NJY1..NJY2), split into the + and - direction, and for
 * y and z the enthalpy flow; faces of wall (1) and melt-surface (4) cells
 * are skipped.  Arrays (M,2) column-major, 1-based plane index; the
 * reference reads them only in its printout (C/Sprint.f90:219-228). */
static double *FLX_, *FLY_, *FLYH_, *FLZ_, *FLZH_;
static int flx_formed;
static void FLX1D(void) {
  FLX_ = realloc(FLX_, sizeof(double) * 2 * MP);
  FLY_ = realloc(FLY_, sizeof(double) * 2 * NP);
  FLYH_ = realloc(FLYH_, sizeof(double) * 2 * NP);
  FLZ_ = realloc(FLZ_, sizeof(double) * 2 * LP);
  FLZH_ = realloc(FLZH_, sizeof(double) * 2 * LP);
  memset(FLX_, 0, sizeof(double) * 2 * MP);
  memset(FLY_, 0, sizeof(double) * 2 * NP);
  memset(FLYH_, 0, sizeof(double) * 2 * NP);
  memset(FLZ_, 0, sizeof(double) * 2 * LP);
  memset(FLZH_, 0, sizeof(double) * 2 * LP);
  const int ih = 3;
  for (int i = 3; i <= MPM1; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        const int ib = IBCELL(i, j, k);
        if (ib == 1 || ib >= 4) continue;
        const double g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, 1) * AREA_C(i, j, k, 1) * C->gmfr;
        if (g0 >= 0.0) FLX_[i - 1] += g0; else FLX_[i - 1 + MP] -= g0;
      }
  for (int j = 3; j <= NPM1; j += 2)
    for (int i = 4; i <= MPM2; i += 2)
      for (int k = LKZ1; k <= LKZ2; k += 2) {
        const int ib = IBCELL(i, j, k);
        if (ib == 1 || ib >= 4) continue;
        const double g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, 2) * AREA_C(i, j, k, 2) * C->gmfr;
        const double h0 = g0 * GF(i, j, k, ih) * C->h_0;
        if (g0 >= 0.0) { FLY_[j - 1] += g0; FLYH_[j - 1] += h0; }
        else { FLY_[j - 1 + NP] -= g0; FLYH_[j - 1 + NP] -= h0; }
      }
  for (int k = 3; k <= LPM1; k += 2)
    for (int i = 4; i <= MPM2; i += 2)
      for (int j = NJY1; j <= NJY2; j += 2) {
        const int ib = IBCELL(i, j, k);
        if (ib == 1 || ib >= 4) continue;
        const double g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, 3) * AREA_C(i, j, k, 3) * C->gmfr;
        const double h0 = g0 * GF(i, j, k, ih) * C->h_0;
        if (g0 >= 0.0) { FLZ_[k - 1] += g0; FLZH_[k - 1] += h0; }
        else { FLZ_[k - 1 + LP] -= g0; FLZH_[k - 1 + LP] -= h0; }
      }
  flx_formed = 1;
}

static double CALC_RESIDUAL(const double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  double resid = 0.0;
  for (int k = kb; k <= ke; k += 2)
    for (int j = jb; j <= je; j += 2)
      for (int i = ib; i <= ie; i += 2) {
        if (IBCELL(i, j, k) > 0) continue;
        double t1 = AS(i, j, k, 1, 1) * XTA(i - 2, j, k), t2 = AS(i, j, k, 1, 2) * XTA(i + 2, j, k);
        double t3 = AS(i, j, k, 2, 1) * XTA(i, j - 2, k), t4 = AS(i, j, k, 2, 2) * XTA(i, j + 2, k);
        double t5 = AS(i, j, k, 3, 1) * XTA(i, j, k - 2), t6 = AS(i, j, k, 3, 2) * XTA(i, j, k + 2);
        double t7 = AP(i, j, k) * XTA(i, j, k), t8 = BS(i, j, k);
        double eq_max = dmax(dmax(dmax(fabs(t1), fabs(t2)), dmax(fabs(t3), fabs(t4))),
                             dmax(dmax(fabs(t5), fabs(t6)), dmax(fabs(t7), fabs(t8))));
        resid += fabs(t1 + t2 + t3 + t4 + t5 + t6 + t8 - t7) / eq_max;
      }
  resid /= C->ncells;
  return dmax(resid, 1.0e-16);
}

/* Additive correction over x-planes (block correction; Settari and Aziz
 * 1973, Patankar 1981 sec. 5.4; not in the reference, used by the p' solve
 * relaxed to a tolerance when pcorr_block_corr is set).  Plane q of the box
 * (i = ib + 2q) gets one value d_q, added to its unpinned cells, such that
 * the residuals r = sum a_nb x_nb + b - a_p x of each plane's unpinned
 * cells sum to zero afterwards, for every plane but the last:
 *   A_q d_q - W_q d_{q-1} - E_q d_{q+1} = R_q,
 *   A_q = sum (a_p - a_nb of the in-plane neighbours that are corrected),
 *   W_q, E_q = sum of a_x- (a_x+) whose neighbour is corrected, R_q = sum r,
 * sums over the plane's unpinned cells.  A cell is corrected when it lies
 * in the box and is not pinned (IBCELL >= 1 or |AP| <= 1e-20, the rows
 * ADLBL3 keeps).  Planes without unpinned cells get d = 0.  The last plane
 * is the reference: its d is 0 and its equation is not solved, because
 * the p' equation has no coefficient toward a pinned cell and is singular
 * (its solution is fixed up to a constant; with coefficients rounded to
 * FP32 it is only nearly so, and solving the full system would add a
 * large constant).  The other planes' system is solved by the Thomas
 * algorithm; a pivot below 1e-12 of its diagonal (a plane closing a block
 * of planes decoupled from the rest) gets d = 0 in the same way.  Sums run
 * over k, then j, then i ascending. */
static int BC_FREE(int i, int j, int k, int ib, int ie, int jb, int je, int kb, int ke) {
  if (i < ib || i > ie || j < jb || j > je || k < kb || k > ke) return 0;
  return IBCELL(i, j, k) <= 0 && fabs(AP(i, j, k)) > small20;
}
static void BLOCK_CORR_X(double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  const int np = (ie - ib) / 2 + 1;
  double *A = calloc((size_t)np, sizeof(double)), *W = calloc((size_t)np, sizeof(double));
  double *E = calloc((size_t)np, sizeof(double)), *R = calloc((size_t)np, sizeof(double));
  double *cp = calloc((size_t)np, sizeof(double)), *dp = calloc((size_t)np, sizeof(double));
  int *nfree = calloc((size_t)np, sizeof(int));
  for (int k = kb; k <= ke; k += 2)
    for (int j = jb; j <= je; j += 2)
      for (int q = 0; q < np; q++) {
        const int i = ib + 2 * q;
        if (!BC_FREE(i, j, k, ib, ie, jb, je, kb, ke)) continue;
        double t1 = AS(i, j, k, 1, 1) * XTA(i - 2, j, k), t2 = AS(i, j, k, 1, 2) * XTA(i + 2, j, k);
        double t3 = AS(i, j, k, 2, 1) * XTA(i, j - 2, k), t4 = AS(i, j, k, 2, 2) * XTA(i, j + 2, k);
        double t5 = AS(i, j, k, 3, 1) * XTA(i, j, k - 2), t6 = AS(i, j, k, 3, 2) * XTA(i, j, k + 2);
        double t7 = AP(i, j, k) * XTA(i, j, k), t8 = BS(i, j, k);
        R[q] += t1 + t2 + t3 + t4 + t5 + t6 + t8 - t7;
        double a = AP(i, j, k);
        if (BC_FREE(i, j - 2, k, ib, ie, jb, je, kb, ke)) a -= AS(i, j, k, 2, 1);
        if (BC_FREE(i, j + 2, k, ib, ie, jb, je, kb, ke)) a -= AS(i, j, k, 2, 2);
        if (BC_FREE(i, j, k - 2, ib, ie, jb, je, kb, ke)) a -= AS(i, j, k, 3, 1);
        if (BC_FREE(i, j, k + 2, ib, ie, jb, je, kb, ke)) a -= AS(i, j, k, 3, 2);
        A[q] += a;
        if (BC_FREE(i - 2, j, k, ib, ie, jb, je, kb, ke)) W[q] += AS(i, j, k, 1, 1);
        if (BC_FREE(i + 2, j, k, ib, ie, jb, je, kb, ke)) E[q] += AS(i, j, k, 1, 2);
        nfree[q]++;
      }
  for (int q = 0; q < np; q++) {
    if (nfree[q] == 0) { A[q] = 1.0; W[q] = E[q] = R[q] = 0.0; }
    const double m = A[q] + (q > 0 ? W[q] * cp[q - 1] : 0.0);
    if (q == np - 1 || !(fabs(m) > 1.0e-12 * fabs(A[q]))) { cp[q] = 0.0; dp[q] = 0.0; continue; }
    cp[q] = -E[q] / m;
    dp[q] = (R[q] + (q > 0 ? W[q] * dp[q - 1] : 0.0)) / m;
  }
  {
    for (int q = np - 2; q >= 0; q--) dp[q] -= cp[q] * dp[q + 1];
    for (int k = kb; k <= ke; k += 2)
      for (int j = jb; j <= je; j += 2)
        for (int q = 0; q < np; q++) {
          const int i = ib + 2 * q;
          if (BC_FREE(i, j, k, ib, ie, jb, je, kb, ke)) XTA(i, j, k) += dp[q];
        }
  }
  free(A); free(W); free(E); free(R); free(cp); free(dp); free(nfree);
}

/* ======================================================================
 * GSLVX1: momentum coefficients and SIMPLER pseudo-velocity,
 * C/Gslv3x.f90:9-113
 * ====================================================================== */
static void GSLVX1(int iu1) {
  int nx[3] = {0, 0, 0};
  nx[iu1 - 1] = 1;
  int mi1 = 4 + nx[0], mi3 = MPM2 - nx[0];
  int nj1 = NJY1V[iu1], nj3 = NJY2V[iu1], lk1 = LKZ1V[iu1], lk3 = LKZ2V[iu1];
  SOURCE(mi1, mi3, nj1, nj3, lk1, lk3);
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) continue;
        AP(i, j, k) = 0.0;
        double g0 = GSLV3X1(i, j, k, iu1);
        double ffl[3][2];
        ECELL(i, j, k, iu1, ffl);
        for (int iu = 1; iu <= 3; iu++) {
          if (iu == iu1) continue;
          for (int m = -1; m <= 1; m += 2) {
            int m1 = (m + 3) / 2, i1, j1, k1, i2, j2, k2;
            VINDX(iu, m, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2);
            if (IBCELL(i1, j1, k1) == 3) { AS(i, j, k, iu, m1) = 0.0; continue; }
            double f = 0.0, diff = 0.0;
            int ii1 = 0, jj1 = 0, kk1 = 0;
            for (int n = -1; n <= 1; n += 2) {
              int n1 = (n + 3) / 2;
              ii1 = i1 + nx[0] * n; jj1 = j1 + nx[1] * n; kk1 = k1 + nx[2] * n;
              int ibc1 = IBCELL(ii1, jj1, kk1);
              if (ibc1 == 3) continue;
              double area = ffl[iu - 1][n1 - 1];
              f += THETA(ii1, jj1, kk1) * DNST(ii1, jj1, kk1) * FZ(ii1, jj1, kk1) * area;
              if (ibc1 == 2) continue;
              double ddl = FDDL(ii1, jj1, kk1, iu);
              diff += THETA(ii1, jj1, kk1) * GAMA(ii1, jj1, kk1) / SGM * area / ddl;
            }
            double fld = (m1 == 1) ? DA(diff, f) + dmax(0.0, f) : DA(diff, f) + dmax(0.0, -f);
            if (IBCELL(ii1, jj1, kk1) == 2 && IBCELL(i2, j2, k2) == 1) {
              double bs0 = fld * (UG(ii1, j2, k2, iu1) - UG(i, j, k, iu1));
              BS(i, j, k) += bs0;
              AS(i, j, k, iu, m1) = 0.0;
              continue;
            }
            AS(i, j, k, iu, m1) = fld;
            AP(i, j, k) += fld;
            g0 += fld * FZ(i2, j2, k2);
          }
        }
        double vol = VOL1(i, j, k, iu1, ffl);
        BS(i, j, k) *= vol;
        AP(i, j, k) += (0.0 - 0.0 - SFP(i, j, k)) * vol;
        if (fabs(AP(i, j, k)) > small20) UG(i, j, k, iu1) = (g0 + BS(i, j, k)) / AP(i, j, k);
      }
}

/* ======================================================================
 * Boundary extrapolation, C/Extr.f90
 * ====================================================================== */
static void EXTR1(double *f1, double *f2, int i, int j, int k, int i1, int j1, int k1, int i2, int j2, int k2, int ix) {
  *f2 = 0.0;
  int ibc0 = imod10(IBCELL(i, j, k));
  if (ix == 3) {
    double dzk = Z(k2) - Z(k1);
    if (dzk != 0.0) { int k0 = (ibc0 == 2) ? (k + k1) / 2 : k; *f2 = (Z(k0) - Z(k1)) / dzk; }
  } else if (ix == 2) {
    double drj = R(j2) - R(j1);
    if (drj != 0.0) { int j0 = (ibc0 == 2) ? (j + j1) / 2 : j; *f2 = (R(j0) - R(j1)) / drj; }
  } else if (ix == 1) {
    double dxi = X(i2) - X(i1);
    if (dxi != 0.0) { int i0 = (ibc0 == 2) ? (i + i1) / 2 : i; *f2 = (X(i0) - X(i1)) / dxi; }
  }
  *f1 = 1.0 - *f2;
}
/* EXT0 / EXT1: extrapolation stencil of a boundary cell, C/Extr.f90:575-661.
 * mode 0 looks for an open neighbour, mode 1 for a non-wall neighbour.
 * Unset outputs keep the caller's previous values, as the Fortran does. */
static void EXTIDX(int mode, int i, int j, int k, int *i1, int *j1, int *k1, int *i2, int *j2, int *k2, int *ix) {
  for (int d = 1; d <= 3; d++) {
    int nx[3] = {0, 0, 0};
    nx[d - 1] = 1;
    for (int m = -2; m <= 2; m += 4) {
      *i1 = imax(2, imin(MP, i + m * nx[0]));
      *j1 = imax(2, imin(NP, j + m * nx[1]));
      *k1 = imax(2, imin(LP, k + m * nx[2]));
      int ok = mode == 0 ? (IBCELL(*i1, *j1, *k1) <= 0) : (imod10(IBCELL(*i1, *j1, *k1)) != 1);
      if (ok) {
        *i2 = imax(2, imin(MP, *i1 + m * nx[0]));
        *j2 = imax(2, imin(NP, *j1 + m * nx[1]));
        *k2 = imax(2, imin(LP, *k1 + m * nx[2]));
        if (imod10(IBCELL(*i2, *j2, *k2)) == 1) { *i2 = *i1; *j2 = *j1; *k2 = *k1; }
        *ix = d;
        return;
      }
    }
  }
  *ix = 4;
}
/* SBC4: copy a cell's state to a neighbour position, C/Sbc.f90:506-522 */
static void SBC4(int i1, int j1, int k1, int i, int j, int k) {
  P(i, j, k) = P(i1, j1, k1); T(i, j, k) = T(i1, j1, k1);
  for (int l = 1; l <= 9; l++) GF(i, j, k, l) = GF(i1, j1, k1, l);
  DNST(i, j, k) = DNST(i1, j1, k1); THETA(i, j, k) = THETA(i1, j1, k1); TMU(i, j, k) = TMU(i1, j1, k1);
  GLAM(i, j, k) = GLAM(i1, j1, k1); GCP(i, j, k) = GCP(i1, j1, k1); GAMA(i, j, k) = GAMA(i1, j1, k1);
}
/* EXTR(1), C/Extr.f90:31-250 */
static int ex_i1 = 2, ex_j1 = 2, ex_k1 = 2, ex_i2 = 2, ex_j2 = 2, ex_k2 = 2, ex_ix = 1;
static void EXTR_GAS(void) {
  C->gfin0 = 0.0;
  C->q_exhaust = 0.0;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        int ibc0 = imod10(IBCELL(i, j, k));
        if (ibc0 != 2 && ibc0 != 3) continue;
        EXTIDX(0, i, j, k, &ex_i1, &ex_j1, &ex_k1, &ex_i2, &ex_j2, &ex_k2, &ex_ix);
        int i1 = ex_i1, j1 = ex_j1, k1 = ex_k1, i2 = ex_i2, j2 = ex_j2, k2 = ex_k2, ix = ex_ix;
        double f1, f2;
        EXTR1(&f1, &f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
        GF(i, j, k, ieps) = GF(i1, j1, k1, ieps);
        P(i, j, k) = P(i1, j1, k1) * f1 + P(i2, j2, k2) * f2;
        if (ibc0 == 2) {
          double g0 = TMU(i, j, k);
          THER(i, j, k);
          TMU(i, j, k) = g0;
          GAMA(i, j, k) = GAMA(i1, j1, k1);
          double area = AREA_C(i, j, k, ix);
          g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, ix) * area;
          C->gfin0 += fabs(g0);
          continue;
        }
        GCP(i, j, k) = GCP(i1, j1, k1) * f1 + GCP(i2, j2, k2) * f2;
        GLAM(i, j, k) = GLAM(i1, j1, k1) * f1 + GLAM(i2, j2, k2) * f2;
        TMU(i, j, k) = TMU(i1, j1, k1);
        GAMA(i, j, k) = GAMA(i1, j1, k1);
        for (int l = lstar; l <= lend; l++) {
          if (l == ieps) continue;
          double g0 = GF(i1, j1, k1, l) * f1 + GF(i2, j2, k2, l) * f2;
          if (l == ik) GF(i, j, k, ik) = dmax(g0, 1.0e-6);
          else if (l == ih) GF(i, j, k, l) = g0;
          else GF(i, j, k, l) = dmax(g0, 0.0);
        }
        ENTH_TO_T(i, j, k);
        DNST_CALC(i, j, k);
        double area = AREA_C(i, j, k, ix) * C->area0;
        double g0 = fabs(UG(i, j, k, ix)) * C->ug0 * area;
        g0 = g0 * THETA(i, j, k) * DNST(i, j, k) * C->dnst0;
        C->q_exhaust += GF(i, j, k, ih) * C->h_0 * g0;
        if (ibc0 == 3) SBC4(i, j, k, (i + i1) / 2, (j + j1) / 2, (k + k1) / 2);
      }
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (imod10(IBCELL(i, j, k)) != 1) continue;
        EXTIDX(1, i, j, k, &ex_i1, &ex_j1, &ex_k1, &ex_i2, &ex_j2, &ex_k2, &ex_ix);
        int i1 = ex_i1, j1 = ex_j1, k1 = ex_k1, i2 = ex_i2, j2 = ex_j2, k2 = ex_k2, ix = ex_ix;
        double f1, f2;
        EXTR1(&f1, &f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
        P(i, j, k) = P(i1, j1, k1) * f1 + P(i2, j2, k2) * f2;
        GCP(i, j, k) = GCP(i1, j1, k1) * f1 + GCP(i2, j2, k2) * f2;
        GLAM(i, j, k) = GLAM(i1, j1, k1) * f1 + GLAM(i2, j2, k2) * f2;
        TMU(i, j, k) = TMU(i2, j2, k2);
        GAMA(i, j, k) = GAMA(i2, j2, k2);
        for (int l = lstar; l <= lend; l++)
          if (l != ih) GF(i, j, k, l) = GF(i1, j1, k1, l);
        if (C->i1st_rad_done == 0 && !C->isothermal_walls) {
          GF(i, j, k, ih) = GF(i1, j1, k1, ih);
          T(i, j, k) = T(i1, j1, k1);
        }
        DNST_CALC(i, j, k);
      }
}
/* EXTRV: exit-face velocities and exit mass flow, C/Extr.f90:358-567 */
static void EXTRV(int mex) {
  double tarea_mean = 0.0, area_ex = 0.0;
  if (mex == 0) { C->gfex = 0.0; C->fuel_out = 0.0; C->tmean_ex = 0.0; }
  /* x-direction */
  for (int i = 3; i <= MPM1; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (imod10(IBCELL(i, j, k)) != 3) continue;
        int done = 0;
        for (int m = -1; m <= 1 && !done; m += 2) {
          if (IBCELL(i + m, j, k) >= 1) continue;
          int i0 = i - m, i1 = i + 2 * m;
          if (UG(i1, j, k, 1) * m > 0.0) UG(i1, j, k, 1) = 0.0;
          double area = AREA_C(i, j, k, 1);
          area_ex += area;
          double g0 = THETA(i1, j, k) * DNST(i1, j, k) * UG(i1, j, k, 1) * area;
          if (mex == 0) {
            C->gfex += fabs(g0);
            C->fuel_out += GF(i, j, k, iyf) * fabs(g0) * C->dnst0 * C->ug0 * C->area0;
            C->tmean_ex += T(i, j, k) * fabs(g0);
            tarea_mean += T(i, j, k) * area;
            done = 1; break;
          }
          UG(i1, j, k, 1) *= C->factor;
          g0 *= C->factor;
          UG(i, j, k, 1) = g0 / (THETA(i, j, k) * DNST(i, j, k) * area);
          BOUND(&UG(i, j, k, 1), 1, 1);
          UG(i0, j, k, 1) = UG(i, j, k, 1);
          done = 1;
        }
        if (done) continue;
        if (j > 2 && IBCELL(i, j - 2, k) <= 0) UG(i, j, k, 1) = UG(i, j - 2, k, 1);
        else if (j < NP && IBCELL(i, j + 2, k) <= 0) UG(i, j, k, 1) = UG(i, j + 2, k, 1);
        else if (k > 2 && IBCELL(i, j, k - 2) <= 0) UG(i, j, k, 1) = UG(i, j, k - 2, 1);
        else if (k < LP && IBCELL(i, j, k + 2) <= 0) UG(i, j, k, 1) = UG(i, j, k + 2, 1);
      }
  /* y-direction (loop order J,K,I as in the reference) */
  for (int j = 3; j <= NPM1; j += 2)
    for (int k = 2; k <= LP; k += 2)
      for (int i = 2; i <= MP; i += 2) {
        if (imod10(IBCELL(i, j, k)) != 3) continue;
        int done = 0;
        for (int m = -1; m <= 1 && !done; m += 2) {
          if (IBCELL(i, j + m, k) >= 1) continue;
          int j0 = j - m, j1 = j + 2 * m;
          if (UG(i, j1, k, 2) * m > 0.0) UG(i, j1, k, 2) = 0.0;
          double area = AREA_C(i, j, k, 2);
          area_ex += area;
          double g0 = THETA(i, j1, k) * DNST(i, j1, k) * UG(i, j1, k, 2) * area;
          if (mex == 0) {
            C->gfex += fabs(g0);
            C->fuel_out += GF(i, j, k, iyf) * fabs(g0) * C->dnst0 * C->ug0 * C->area0;
            C->tmean_ex += T(i, j, k) * fabs(g0);
            tarea_mean += T(i, j, k) * area;
            done = 1; break;
          }
          UG(i, j1, k, 2) *= C->factor;
          g0 *= C->factor;
          UG(i, j, k, 2) = g0 / (THETA(i, j, k) * DNST(i, j, k) * area);
          BOUND(&UG(i, j, k, 2), 2, 1);
          UG(i, j0, k, 2) = UG(i, j, k, 2);
          done = 1;
        }
        if (done) continue;
        if (k > 2 && IBCELL(i, j, k - 2) <= 0) UG(i, j, k, 2) = UG(i, j, k - 2, 2);
        else if (k < LP && IBCELL(i, j, k + 2) <= 0) UG(i, j, k, 2) = UG(i, j, k + 2, 2);
        else if (i > 2 && IBCELL(i - 2, j, k) <= 0) UG(i, j, k, 2) = UG(i - 2, j, k, 2);
        else if (i < MP && IBCELL(i + 2, j, k) <= 0) UG(i, j, k, 2) = UG(i + 2, j, k, 2);
      }
  /* z-direction (loop order K,I,J); the reference skips the exit-face
   * velocity update for z (GOTO 160, C/Extr.f90:529) */
  for (int k = 3; k <= LPM1; k += 2)
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2) {
        if (imod10(IBCELL(i, j, k)) != 3) continue;
        int done = 0;
        for (int m = -1; m <= 1 && !done; m += 2) {
          if (IBCELL(i, j, k + m) >= 1) continue;
          int k1 = k + 2 * m;
          if (UG(i, j, k1, 3) * m > 0.0) UG(i, j, k1, 3) = 0.0;
          double area = AREA_C(i, j, k, 3);
          area_ex += area;
          double g0 = THETA(i, j, k1) * DNST(i, j, k1) * UG(i, j, k1, 3) * area;
          if (mex == 0) {
            C->gfex += fabs(g0);
            C->fuel_out += GF(i, j, k, iyf) * fabs(g0) * C->dnst0 * C->ug0 * C->area0;
            C->tmean_ex += T(i, j, k) * fabs(g0);
            tarea_mean += T(i, j, k) * area;
            done = 1; break;
          }
          UG(i, j, k1, 3) *= C->factor;
          done = 1;
        }
        if (done) continue;
        if (i > 2 && IBCELL(i - 2, j, k) <= 0) UG(i, j, k, 3) = UG(i - 2, j, k, 3);
        else if (i < MP && IBCELL(i + 2, j, k) <= 0) UG(i, j, k, 3) = UG(i + 2, j, k, 3);
        else if (j > 2 && IBCELL(i, j - 2, k) <= 0) UG(i, j, k, 3) = UG(i, j - 2, k, 3);
        else if (j < NP && IBCELL(i, j + 2, k) <= 0) UG(i, j, k, 3) = UG(i, j + 2, k, 3);
      }
  if (mex == 0) {
    C->gfex *= C->gmfr;
    if (C->gfex < small16) C->tmean_ex = tarea_mean * C->t0 / area_ex;
    else C->tmean_ex = C->tmean_ex * C->t0 * C->gmfr / C->gfex;
  }
}

/* ======================================================================
 * Interpolation to non-centre lattice positions, C/Intp.f90
 * ====================================================================== */
static double INTP1(int i, int j, int k, int ix) {
  if (ix == 3) return (DZ(k) <= 0.0 || k > LP) ? 0.0 : (Z(k) - Z(k - 1)) / DZ(k);
  if (ix == 2) return (DR(j) <= 0.0 || j > NP) ? 0.0 : (R(j) - R(j - 1)) / DR(j);
  return (DX(i) <= 0.0 || i > MP) ? 0.0 : (X(i) - X(i - 1)) / DX(i);
}
/* INTP0, C/Intp.f90:145-175.  f1 is the caller's weight and is modified
 * in place at inlet faces, exactly as the Fortran dummy argument. */
static void INTP0(double *f1, int i, int j, int k, int i1, int j1, int k1, int i2, int j2, int k2) {
  if (IBCELL(i, j, k) == 3) return;
  if (IBCELL(i, j, k) == 2) {
    if (IBCELL(i1, j1, k1) == 2 && IBCELL(i2, j2, k2) == 0) *f1 = 1.0;
    else if (IBCELL(i1, j1, k1) == 0 && IBCELL(i2, j2, k2) == 2) *f1 = 0.0;
  }
  double a = *f1, b = 1.0 - a;
  P(i, j, k) = a * P(i1, j1, k1) + b * P(i2, j2, k2);
  T(i, j, k) = a * T(i1, j1, k1) + b * T(i2, j2, k2);
  for (int l = lstar; l <= lend; l++) GF(i, j, k, l) = a * GF(i1, j1, k1, l) + b * GF(i2, j2, k2, l);
  CONC(i, j, k);
  TMU(i, j, k) = a * TMU(i1, j1, k1) + b * TMU(i2, j2, k2);
  GCP(i, j, k) = a * GCP(i1, j1, k1) + b * GCP(i2, j2, k2);
  GLAM(i, j, k) = a * GLAM(i1, j1, k1) + b * GLAM(i2, j2, k2);
  GAMA(i, j, k) = a * GAMA(i1, j1, k1) + b * GAMA(i2, j2, k2);
  THETA(i, j, k) = a * THETA(i1, j1, k1) + b * THETA(i2, j2, k2);
  DNST_CALC(i, j, k);
}
/* INTP, C/Intp.f90:32-72 */
static void INTP(void) {
  for (int i = 2; i <= MP; i += 2) {
    for (int k = 3; k <= LPM1; k += 2) {
      double f1 = INTP1(i, 0, k, 3);
      for (int j = 2; j <= NP; j += 2) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0(&f1, i, j, k, i, j, k + 1, i, j, k - 1);
      }
    }
    for (int j = 3; j <= NPM1; j += 2) {
      double f1 = INTP1(i, j, 0, 2);
      for (int k = 2; k <= LP; k++) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0(&f1, i, j, k, i, j + 1, k, i, j - 1, k);
      }
    }
  }
  for (int i = 3; i <= MPM1; i += 2) {
    double f1 = INTP1(i, 0, 0, 1);
    for (int k = 2; k <= LP; k++)
      for (int j = 2; j <= NP; j++) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0(&f1, i, j, k, i + 1, j, k, i - 1, j, k);
      }
  }
}
/* INTPV1, C/Intp.f90:230-272 */
static void INTPV1(int m, int n, int l, int iu, int ix) {
  int i, j, k;
  if (iu == 1) { i = m; j = n; k = l; }
  else if (iu == 2) { j = m; k = n; i = l; }
  else { k = m; i = n; j = l; }
  int b = IBCELL(i, j, k);
  if (b >= 1 && b != 4) return;
  double f1 = INTP1(i, j, k, ix), f2 = 1.0 - f1;
  if (ix == 1) {
    UG(i, j, k, iu) = UG(i + 1, j, k, iu) * f1 + UG(i - 1, j, k, iu) * f2;
  } else if (ix == 2) {
    if (IBCELL(i, j + 1, k) == 2 && b == 0) { f1 = 0.0; f2 = 1.0; }
    else if (IBCELL(i, j - 1, k) == 2 && b == 0) { f1 = 1.0; f2 = 0.0; }
    UG(i, j, k, iu) = UG(i, j + 1, k, iu) * f1 + UG(i, j - 1, k, iu) * f2;
  } else {
    if (IBCELL(i, j, k + 1) == 2 && b == 0) { f1 = 0.0; f2 = 1.0; }
    else if (IBCELL(i, j, k - 1) == 2 && b == 0) { f1 = 1.0; f2 = 0.0; }
    UG(i, j, k, iu) = UG(i, j, k + 1, iu) * f1 + UG(i, j, k - 1, iu) * f2;
  }
  BOUND(&UG(i, j, k, iu), iu, 1);
}
/* INTPV, C/Intp.f90:84-139 */
static void INTPV(void) {
  for (int iu = 1; iu <= 3; iu++) {
    int m3, n3, l3;
    if (iu == 1) { m3 = MP; n3 = NP; l3 = LP; }
    else if (iu == 2) { m3 = NP; n3 = LP; l3 = MP; }
    else { m3 = LP; n3 = MP; l3 = NP; }
    int ixa = (iu == 1) ? 2 : (iu == 2) ? 3 : 1;
    int ixb = (iu == 1) ? 3 : (iu == 2) ? 1 : 2;
    for (int m = 3; m <= m3 - 1; m += 2) {
      if (n3 > 3)
        for (int l = 2; l <= l3; l += 2)
          for (int n = 3; n <= n3 - 1; n += 2) INTPV1(m, n, l, iu, ixa);
      if (l3 <= 3) continue;
      for (int l = 3; l <= l3 - 1; l += 2)
        for (int n = 2; n <= n3; n++) INTPV1(m, n, l, iu, ixb);
    }
    for (int m = 4; m <= m3; m += 2)
      for (int n = 2; n <= n3; n++)
        for (int l = 2; l <= l3; l++) INTPV1(m, n, l, iu, iu);
  }
}

/* ======================================================================
 * GSLV3V, C/Gslv3v.f90:7-236
 * ====================================================================== */
static void GSLV3V(void) {
  SGM = 1.0;
  for (size_t q = 0; q < NN; q++) AP_[q] = 1.0;
  memset(AS_, 0, sizeof(double) * NN * 6);
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    int nx[3] = {0, 0, 0};
    nx[iu1 - 1] = 1;
    for (int i = 2 + nx[0]; i <= MP - nx[0]; i += 2)
      for (int j = 2 + nx[1]; j <= NP - nx[1]; j += 2)
        for (int k = 2 + nx[2]; k <= LP - nx[2]; k += 2) FZ(i, j, k) = UG(i, j, k, iu1);
  }
  for (int iu1 = 1; iu1 <= 3; iu1++) { NELG = iu1; GSLVX1(iu1); }
  if (STOP_STAGE == 1) return;

  /* pressure equation */
  NELG = 4;
  int mi1 = 4, mi3 = MPM2, nj1 = NJY1, nj3 = NJY2, lk1 = LKZ1, lk3 = LKZ2;
  SOURCE(mi1, mi3, nj1, nj3, lk1, lk3);
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) { BS(i, j, k) = P(i, j, k); continue; }
        AP(i, j, k) = -SFP(i, j, k);
        for (int m = 1; m <= 2; m++) {
          int i1 = (m == 1) ? i - 1 : i + 1;
          if (IBCELL(i1, j, k) <= 0 && fabs(AP(i1, j, k)) > small20) {
            double area = AREA_C(i, j, k, 1);
            AS(i, j, k, 1, m) = THETA(i1, j, k) * THETA(i1, j, k) * DNST(i1, j, k) * area * area / AP(i1, j, k) * DXAV(i1, j, k) / DX(i1);
          }
          AP(i, j, k) += AS(i, j, k, 1, m);
        }
        for (int m = 1; m <= 2; m++) {
          int j1 = (m == 1) ? j - 1 : j + 1;
          if (IBCELL(i, j1, k) <= 0 && fabs(AP(i, j1, k)) > small20) {
            double area = AREA_C(i, j, k, 2);
            AS(i, j, k, 2, m) = THETA(i, j1, k) * THETA(i, j1, k) * DNST(i, j1, k) * area * area / AP(i, j1, k) * DRAV(i, j1, k) / DR(j1);
          }
          AP(i, j, k) += AS(i, j, k, 2, m);
        }
        for (int m = 1; m <= 2; m++) {
          int k1 = (m == 1) ? k - 1 : k + 1;
          if (IBCELL(i, j, k1) <= 0 && fabs(AP(i, j, k1)) > small20) {
            double area = AREA_C(i, j, k, 3);
            AS(i, j, k, 3, m) = THETA(i, j, k1) * THETA(i, j, k1) * DNST(i, j, k1) * area * area / AP(i, j, k1) * DZAV(i, j, k1) / DZ(k1);
          }
          AP(i, j, k) += AS(i, j, k, 3, m);
        }
      }
  if (C->want_pre_residuals) C->p_resid_pre = CALC_RESIDUAL(P_, mi1, mi3, nj1, nj3, lk1, lk3);
  ADLBL3(P_, mi1, mi3, nj1, nj3, lk1, lk3);
  C->p_resid = CALC_RESIDUAL(P_, mi1, mi3, nj1, nj3, lk1, lk3);
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2)
        if (IBCELL(i, j, k) < 1) BOUND(&P(i, j, k), 0, 1);

  /* starred velocities */
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    NELG = iu1;
    int nx[3] = {0, 0, 0};
    nx[iu1 - 1] = 1;
    mi1 = 4 + nx[0]; mi3 = MPM2 - nx[0];
    nj1 = NJY1V[iu1]; nj3 = NJY2V[iu1]; lk1 = LKZ1V[iu1]; lk3 = LKZ2V[iu1];
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          AP(i, j, k) = AP(i, j, k) / C->rf[NELG - 1];
          double vol = VOL_C(i, j, k), dpdx;
          if (iu1 == 1) dpdx = (P(i - 1, j, k) - P(i + 1, j, k)) / DX(i);
          else if (iu1 == 2) dpdx = (P(i, j - 1, k) - P(i, j + 1, k)) / DR(j) / RS(j);
          else dpdx = (P(i, j, k - 1) - P(i, j, k + 1)) / DZ(k);
          BS(i, j, k) = BS(i, j, k) + C->eun * THETA(i, j, k) * vol * dpdx + (1.0 - C->rf[NELG - 1]) * AP(i, j, k) * FZ(i, j, k);
        }
    if (C->want_pre_residuals) C->u_resid_pre[iu1 - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    ADLBL3(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    C->u_resid[iu1 - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          BOUND(&FZ(i, j, k), iu1, 1);
          UG(i, j, k, iu1) = FZ(i, j, k);
        }
  }

  /* pressure correction */
  NELG = 5;
  mi1 = 4; mi3 = MPM2; nj1 = NJY1; nj3 = NJY2; lk1 = LKZ1; lk3 = LKZ2;
  SOURCE(mi1, mi3, nj1, nj3, lk1, lk3);
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) FZ(i, j, k) = 0.0;
  if (C->pcorr_mode == GFM_PCORR_TOLERANCE) {
    /* zero initial guess relaxed one sweep at a time to pcorr_tol (the
       GPU's GFM_PCORR_TOLERANCE with GFM_PGUESS_ZERO; not in the reference),
       each sweep preceded by the x-plane block correction when
       pcorr_block_corr is set */
    const int save = C->ntimesg[NELG - 1];
    C->ntimesg[NELG - 1] = 1;
    double r = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    C->pcorr_resid_start = r;
    int ns = 0;
    while (ns < C->pcorr_max_sweeps && !(r <= C->pcorr_tol)) {
      if (C->pcorr_block_corr) BLOCK_CORR_X(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
      ADLBL3(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
      ns++;
      r = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    }
    C->ntimesg[NELG - 1] = save;
    C->pcorr_sweeps = ns;
    C->pcorr_resid = r;
  } else {
    ADLBL3(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    C->pcorr_resid = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
  }

  /* velocity correction */
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    int nx[3] = {0, 0, 0};
    nx[iu1 - 1] = 1;
    mi1 = 4 + nx[0]; mi3 = MPM2 - nx[0];
    nj1 = NJY1V[iu1]; nj3 = NJY2V[iu1]; lk1 = LKZ1V[iu1]; lk3 = LKZ2V[iu1];
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1 || fabs(AP(i, j, k)) < small20) continue;
          double vol = VOL_C(i, j, k), dpdx;
          if (iu1 == 1) dpdx = (FZ(i - 1, j, k) - FZ(i + 1, j, k)) / DX(i);
          else if (iu1 == 2) dpdx = (FZ(i, j - 1, k) - FZ(i, j + 1, k)) / DR(j) / RS(j);
          else dpdx = (FZ(i, j, k - 1) - FZ(i, j, k + 1)) / DZ(k);
          double g0 = C->eun * THETA(i, j, k) * dpdx * vol / (AP(i, j, k) * C->rf[iu1 - 1]);
          UG(i, j, k, iu1) += g0;
          BOUND(&UG(i, j, k, iu1), iu1, 1);
        }
  }
  if (STOP_STAGE == 2) return;
  EXTRV(0);
  if (C->want_plane_flows) FLX1D();
  C->gfin = C->gfin0 * C->gmfr; /* FLX1D, C/Flx1d.f90:75-77 */
  if (C->itr_gas > 10) {
    if (C->gfex < small20) { C->factor = 1.0; C->fadj = 0.0; }
    else {
      C->factor = C->gfin / C->gfex;
      if (C->factor > C->f_mx) { C->factor = C->f_mx; C->fadj = 0.0; }
      else if (C->factor < C->f_mn) { C->factor = C->f_mn; C->fadj = 0.0; }
    }
  }
  EXTRV(1);
  INTPV();
}

/* ======================================================================
 * GSLV3G / GSLV3G0, C/Gslv3g.f90
 * ====================================================================== */
static void GSLV3G0(int mi1, int mi3, int nj1, int nj3, int lk1, int lk3, int igs) {
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) continue;
        AP(i, j, k) = 0.0;
        for (int iu = 1; iu <= 3; iu++)
          for (int m = -1; m <= 1; m += 2) {
            int m1 = (m + 3) / 2;
            if (igs != 1) {
              int i1, j1, k1, i2, j2, k2;
              VINDX(iu, m, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2);
              int ibc0 = IBCELL(i2, j2, k2);
              if (ibc0 <= 0) AS(i, j, k, iu, m1) = FLWD(i1, j1, k1, iu, m1);
              else if (ibc0 == 2) AS(i, j, k, iu, m1) = FLW(i2, j2, k2, iu, m1);
              else AS(i, j, k, iu, m1) = 0.0;
            }
            AP(i, j, k) += AS(i, j, k, iu, m1);
          }
      }
}
static void GSLV3G(void) {
  int mi1 = 4, mi3 = MPM2, nj1 = NJY1, nj3 = NJY2, lk1 = LKZ1, lk3 = LKZ2;
  for (int l = lstar; l <= lend; l++) {
    if (l == iyf) NELG = 6;
    else if (l == iyco2) NELG = 7;
    else if (l == ih) NELG = 8;
    else if (l == ik) NELG = 9;
    else if (l == ieps) NELG = 10;
    else if (l == iyn2) NELG = 11;
    else if (l == iyo2) NELG = 12;
    else continue;
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2)
        for (int k = 2; k <= LP; k += 2) FZ(i, j, k) = GF(i, j, k, l);
    SOURCE(mi1, mi3, nj1, nj3, lk1, lk3);
    if (l == lstar || SGM != sigma_eq[l - 1]) { SGM = sigma_eq[l - 1]; GSLV3G0(mi1, mi3, nj1, nj3, lk1, lk3, 0); }
    else GSLV3G0(mi1, mi3, nj1, nj3, lk1, lk3, 1);
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          if (fabs(AP(i, j, k)) <= small20) { AP(i, j, k) = 1.0; BS(i, j, k) = FZ(i, j, k); continue; }
          double vol = VOL_C(i, j, k);
          BS(i, j, k) *= vol;
          AP(i, j, k) = AP(i, j, k) - (SFP(i, j, k) * vol);
          AP(i, j, k) = AP(i, j, k) / C->rfl[l - 1];
          BS(i, j, k) = BS(i, j, k) + (1.0 - C->rfl[l - 1]) * FZ(i, j, k) * AP(i, j, k);
        }
    if (C->want_pre_residuals) {
      C->gf_resid_pre[l - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
      if (NELG == 8) C->h_resid_pre = C->gf_resid_pre[l - 1];
    }
    ADLBL3(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    C->gf_resid[l - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          BOUND(&FZ(i, j, k), l, 2);
          GF(i, j, k, l) = FZ(i, j, k);
        }
  }
  C->h_resid = C->gf_resid[ih - 1];
}

/* ======================================================================
 * GFLOW, C/Gflow.f90:6-408
 * ====================================================================== */
static void T_AVE_CALC(void) {
  double tmax = 0.0, tave = 0.0, tvol = 0.0, rvt = 0.0, vt = 0.0;
  for (int k = 4; k <= LP - 2; k += 2)
    for (int j = 4; j <= NP - 2; j += 2)
      for (int i = 4; i <= MP - 2; i += 2) {
        if (IBCELL(i, j, k) != 0) continue;
        double vol = VOL_C(i, j, k);
        tave += T(i, j, k) * DNST(i, j, k) * vol;
        rvt += DNST(i, j, k) * vol;
        tvol += T(i, j, k) * vol;
        vt += vol;
        tmax = dmax(tmax, T(i, j, k));
      }
  C->tave = tave * C->t0 / rvt;
  C->tave_vol = tvol * C->t0 / vt;
  C->tmax_c = tmax * C->t0;
}
static void GFLOW(void) {
  /* theta next to walls */
  for (int i = 2; i <= MP; i++)
    for (int j = 2; j <= NP; j++)
      for (int k = 2; k <= LP; k++) {
        if (IBCELL(i, j, k) >= 1) continue;
        double g1, g2;
        if (IBCELL(i - 1, j, k) == 1) { g1 = (X(i) - X(i - 1)) * (X(i) - X(i - 1)); g2 = (X(i + 2) - X(i - 1)) * (X(i + 2) - X(i - 1)); THETA(i - 1, j, k) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i + 2, j, k) * (g1 / (g1 - g2)); }
        if (IBCELL(i + 1, j, k) == 1) { g1 = (X(i) - X(i + 1)) * (X(i) - X(i + 1)); g2 = (X(i - 2) - X(i + 1)) * (X(i - 2) - X(i + 1)); THETA(i + 1, j, k) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i - 2, j, k) * (g1 / (g1 - g2)); }
        if (IBCELL(i, j - 1, k) == 1) { g1 = (R(j) - R(j - 1)) * (R(j) - R(j - 1)); g2 = (R(j + 2) - R(j - 1)) * (R(j + 2) - R(j - 1)); THETA(i, j - 1, k) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i, j + 2, k) * (g1 / (g1 - g2)); }
        if (IBCELL(i, j + 1, k) == 1) { g1 = (R(j) - R(j + 1)) * (R(j) - R(j + 1)); g2 = (R(j - 2) - R(j + 1)) * (R(j - 2) - R(j + 1)); THETA(i, j + 1, k) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i, j - 2, k) * (g1 / (g1 - g2)); }
        if (IBCELL(i, j, k - 1) == 1) { g1 = (Z(k) - Z(k - 1)) * (Z(k) - Z(k - 1)); g2 = (Z(k + 2) - Z(k - 1)) * (Z(k + 2) - Z(k - 1)); THETA(i, j, k - 1) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i, j, k + 2) * (g1 / (g1 - g2)); }
        if (IBCELL(i, j, k + 1) == 1) { g1 = (Z(k) - Z(k + 1)) * (Z(k) - Z(k + 1)); g2 = (Z(k - 2) - Z(k + 1)) * (Z(k - 2) - Z(k + 1)); THETA(i, j, k + 1) = THETA(i, j, k) * (g2 / (g2 - g1)) + THETA(i, j, k - 2) * (g1 / (g1 - g2)); }
      }
  for (size_t q = 0; q < NN; q++)
    if (IBC[q] == 1) GAMA_[q] = 0.0;
  double rt20 = C->reyg2 * sqrt(C->epsr) * pow(cmu, 0.25);
  for (int i = 2; i <= MP; i++)
    for (int j = 2; j <= NP; j++)
      for (int k = 2; k <= LP; k++) {
        if (IBCELL(i, j, k) >= 1) continue;
        double c = 1.0;
        double blvm = rt20 * DNST(i, j, k) * sqrt(GF(i, j, k, ik)) / c;
        double vgam = c / C->reyg2;
        if (IBCELL(i - 1, j, k) == 1) { double g = vgam * oracle_tgamf(blvm * (X(i) - X(i - 1))); GAMA(i - 1, j, k) = dmax(g, GAMA(i - 1, j, k)); TMU(i - 1, j, k) = C->reyg2 * GAMA(i - 1, j, k); TMU(i - 2, j, k) = TMU(i - 1, j, k); }
        if (IBCELL(i + 1, j, k) == 1) { double g = vgam * oracle_tgamf(blvm * (X(i + 1) - X(i))); GAMA(i + 1, j, k) = dmax(g, GAMA(i + 1, j, k)); TMU(i + 1, j, k) = C->reyg2 * GAMA(i + 1, j, k); TMU(i + 2, j, k) = TMU(i + 1, j, k); }
        if (IBCELL(i, j - 1, k) == 1) { double g = vgam * oracle_tgamf(blvm * (R(j) - R(j - 1))); GAMA(i, j - 1, k) = dmax(g, GAMA(i, j - 1, k)); TMU(i, j - 1, k) = C->reyg2 * GAMA(i, j - 1, k); TMU(i, j - 2, k) = TMU(i, j - 1, k); }
        if (IBCELL(i, j + 1, k) == 1) { double g = vgam * oracle_tgamf(blvm * (R(j + 1) - R(j))); GAMA(i, j + 1, k) = dmax(g, GAMA(i, j + 1, k)); TMU(i, j + 1, k) = C->reyg2 * GAMA(i, j + 1, k); TMU(i, j + 2, k) = TMU(i, j + 1, k); }
        if (IBCELL(i, j, k - 1) == 1) { double g = vgam * oracle_tgamf(blvm * (Z(k) - Z(k - 1))); GAMA(i, j, k - 1) = dmax(g, GAMA(i, j, k - 1)); TMU(i, j, k - 1) = C->reyg2 * GAMA(i, j, k - 1); TMU(i, j, k - 2) = TMU(i, j, k - 1); }
        if (IBCELL(i, j, k + 1) == 1) { double g = vgam * oracle_tgamf(blvm * (Z(k + 1) - Z(k))); GAMA(i, j, k + 1) = dmax(g, GAMA(i, j, k + 1)); TMU(i, j, k + 1) = C->reyg2 * GAMA(i, j, k + 1); TMU(i, j, k + 2) = TMU(i, j, k + 1); }
      }

  GSLV3V();
  if (STOP_STAGE) return;

  /* pressure referencing */
  int i = C->i_ref, j = C->j_ref, k = C->k_ref;
  int i1 = i, j1 = j, k1 = k, i2 = i, j2 = j, k2 = k, nx = 0;
  if (i < MPM2 && IBCELL(i + 2, j, k) == 0) { nx = 1; i1 = i + 2; i2 = (IBCELL(i + 4, j, k) == 0) ? i + 4 : i1; }
  else if (i > 2 && IBCELL(i - 2, j, k) == 0) { nx = 1; i1 = i - 2; i2 = (IBCELL(i - 4, j, k) == 0) ? i - 4 : i1; }
  else if (j < NPM2 && IBCELL(i, j + 2, k) == 0) { nx = 2; j1 = j + 2; j2 = (j < NPM2 && IBCELL(i, j + 4, k) <= 0) ? j + 4 : j1; }
  else if (j > 2 && IBCELL(i, j - 2, k) <= 0) { nx = 2; j1 = j - 2; j2 = (j > 4 && IBCELL(i, j - 4, k) <= 0) ? j - 4 : j1; }
  else if (k < LPM2 && IBCELL(i, j, k + 2) <= 0) { nx = 3; k1 = k + 2; k2 = (k < LPM2 && IBCELL(i, j, k + 4) <= 0) ? k + 4 : k1; }
  else if (k > 2 && IBCELL(i, j, k - 2) <= 0) { nx = 3; k1 = k - 2; k2 = (k > 4 && IBCELL(i, j, k - 4) <= 0) ? k - 4 : k1; }
  double f1, f2;
  EXTR1(&f1, &f2, i, j, k, i1, j1, k1, i2, j2, k2, nx);
  P(i, j, k) = P(i1, j1, k1) * f1 + P(i2, j2, k2) * f2;
  double g0 = P(i1, j1, k1) - C->p_ref;

  GSLV3G();

  for (int ii = 2; ii <= MP; ii += 2)
    for (int jj = 2; jj <= NP; jj += 2)
      for (int kk = 2; kk <= LP; kk += 2) {
        if (IBCELL(ii, jj, kk) >= 1) continue;
        CONC(ii, jj, kk);
        ENTH_TO_T(ii, jj, kk);
        if (T(ii, jj, kk) > C->t_mx) { T(ii, jj, kk) = C->t_mx; T_TO_ENTH(ii, jj, kk); }
        else if (T(ii, jj, kk) < C->t_mn) { T(ii, jj, kk) = C->t_mn; T_TO_ENTH(ii, jj, kk); }
        P(ii, jj, kk) -= g0;
        BOUND(&P(ii, jj, kk), 0, 1);
        DNST_CALC(ii, jj, kk);
        THER(ii, jj, kk);
      }
  T_AVE_CALC();
  EXTR_GAS();
  INTP();
}

/* exit temperature reductions of SPHASE, C/Sphase.f90:74-89 */
static void EXIT_T(void) {
  double g1 = 0.0, tav = 0.0, tx = 0.0;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (IBCELL(i, j, k) != 3) continue;
        double g0 = AREA_C(i, j, k, 3);
        if (g0 < small20) g0 = AREA_C(i, j, k, 2);
        if (g0 < small20) g0 = AREA_C(i, j, k, 1);
        g1 += g0;
        tav += T(i, j, k) * g0;
        if (T(i, j, k) > tx) tx = T(i, j, k);
      }
  C->t_av_exit = (g1 > 0.0) ? tav * C->t0 / g1 : 0.0;
  C->t_x_exit = tx * C->t0;
}

/* ======================================================================
 * Binding and entry points
 * ====================================================================== */
static int bind(gfm_ctrl *c, const gfm_host_arrays *h) {
  C = c;
  MP = c->mp; NP = c->np; LP = c->lp; NR0 = c->nr0;
  if (MP < 8 || NP < 8 || LP < 8 || (MP & 1) || (NP & 1) || (LP & 1)) return GFM_E_ARG;
  MPM1 = MP - 1; MPM2 = MP - 2; NPM1 = NP - 1; NPM2 = NP - 2; LPM1 = LP - 1; LPM2 = LP - 2;
  NN = (size_t)MP * NP * LP;
  NJY1 = 4; NJY2 = NPM2; NJY1V[1] = 4; NJY2V[1] = NPM2; NJY1V[2] = 5; NJY2V[2] = NP - 3; NJY1V[3] = 4; NJY2V[3] = NPM2;
  LKZ1 = 4; LKZ2 = LPM2; LKZ1V[1] = 4; LKZ2V[1] = LPM2; LKZ1V[2] = 4; LKZ2V[2] = LPM2; LKZ1V[3] = 5; LKZ2V[3] = LP - 3;
  X_ = h->x; DX_ = h->dx; R_ = h->r; DR_ = h->dr; RS_ = h->rs; Z_ = h->z; DZ_ = h->dz;
  IBC = h->ibcell;
  P_ = h->p; T_ = h->t; DNST_ = h->dnst; THETA_ = h->theta; TMU_ = h->tmu; GCP_ = h->gcp; GLAM_ = h->glam;
  GAMA_ = h->gama; GTE_ = h->gte; CRATE_ = h->crate; HRATE_ = h->hrate; BSH_ = h->bs_hprev;
  UG_ = h->ug; GF_ = h->gf;
  RTIM_ = h->rtim; DELRT_ = h->delrt; HRT_ = h->hrt; DELR_ = h->delr; TIMED_ = h->timed; HRD_ = h->hrd;
  wt0 = 1.0 / (0.5 / wf + 0.0 / wco2 + 0.0 / wh2o + 0.49 / wo2 + 0.01 / wn2);
  stoich = 2.0 * wo2 / wf;
  for (size_t q = 0; q < NN; q++)
    if (IBC[q] < -2) return GFM_E_ARG;
  static size_t cap = 0;
  if (cap != NN) {
    free(FZ_); free(BS_); free(SFP_); free(AP_); free(AS_); free(XT2_); free(XT3_);
    if (RMASS_) {
      free(RMASS_);
      RMASS_ = calloc(NN, sizeof(double));
      if (!RMASS_) return GFM_E_DEVICE;
      for (size_t q = 0; q < NN; q++) RMASS_[q] = -1.0;
    }
    FZ_ = calloc(NN, sizeof(double)); BS_ = calloc(NN, sizeof(double)); SFP_ = calloc(NN, sizeof(double));
    AP_ = calloc(NN, sizeof(double)); AS_ = calloc(NN * 6, sizeof(double));
    XT2_ = calloc((size_t)(NP + 4) * (LP + 4), sizeof(double)); XT3_ = calloc((size_t)(NP + 4) * (LP + 4), sizeof(double));
    if (!FZ_ || !BS_ || !SFP_ || !AP_ || !AS_ || !XT2_ || !XT3_) return GFM_E_DEVICE;
    cap = NN;
  }
  return GFM_OK;
}

/* tail of INITSV, C/Initsv.f90:63-84 */
int oracle_initsv_tail(gfm_ctrl *c, const gfm_host_arrays *h) {
  int rc = bind(c, h);
  if (rc) return rc;
  INTPV();
  EXTR_GAS();
  INTP();
  return GFM_OK;
}

/* n_iters GFLOW calls with the SPHASE loop control (C/Sphase.f90:37-235).
 * res_hist (may be NULL) receives res_mass per iteration.  Returns the
 * number of iterations run (negative on error). */
int oracle_sphase(gfm_ctrl *c, const gfm_host_arrays *h, int n_iters, double *res_hist) {
  int rc = bind(c, h);
  if (rc) return -rc;
  int n = 0;
  while (n < n_iters && c->itr_gas < c->itr_gend && c->itr_gas < c->maxgi) {
    c->itr_gas += 1;
    GFLOW();
    EXIT_T();
    if (res_hist) res_hist[n] = c->res_mass;
    n++;
    if (c->res_mass < c->bgcon) { c->itr_gend = c->itr_gas; break; }
  }
  return n;
}

/* ADLBL3 on caller arrays (ap, as(…,3,2), bs, xt on the lattice). */
int oracle_adlbl3(gfm_ctrl *c, const gfm_host_arrays *h, int nelg, const double *ap, const double *as,
                  const double *bs, double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  int rc = bind(c, h);
  if (rc) return rc;
  memcpy(AP_, ap, sizeof(double) * NN);
  memcpy(AS_, as, sizeof(double) * NN * 6);
  memcpy(BS_, bs, sizeof(double) * NN);
  NELG = nelg;
  ADLBL3(xt, ib, ie, jb, je, kb, ke);
  return GFM_OK;
}
/* BLOCK_CORR_X on caller arrays (x-plane additive correction of xt). */
int oracle_block_corr_x(gfm_ctrl *c, const gfm_host_arrays *h, const double *ap, const double *as, const double *bs,
                        double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  int rc = bind(c, h);
  if (rc) return rc;
  memcpy(AP_, ap, sizeof(double) * NN);
  memcpy(AS_, as, sizeof(double) * NN * 6);
  memcpy(BS_, bs, sizeof(double) * NN);
  BLOCK_CORR_X(xt, ib, ie, jb, je, kb, ke);
  return GFM_OK;
}
/* FLX1D on the caller's state; and the flows the last FLX1D formed (during
 * an iteration with want_plane_flows set, or this call). */
int oracle_flx1d(gfm_ctrl *c, const gfm_host_arrays *h) {
  int rc = bind(c, h);
  if (rc) return rc;
  FLX1D();
  return GFM_OK;
}
int oracle_plane_flows(double *flx, double *fly, double *flyh, double *flz, double *flzh) {
  if (!flx_formed) return GFM_E_STATE;
  if (flx) memcpy(flx, FLX_, sizeof(double) * 2 * MP);
  if (fly) memcpy(fly, FLY_, sizeof(double) * 2 * NP);
  if (flyh) memcpy(flyh, FLYH_, sizeof(double) * 2 * NP);
  if (flz) memcpy(flz, FLZ_, sizeof(double) * 2 * LP);
  if (flzh) memcpy(flzh, FLZH_, sizeof(double) * 2 * LP);
  return GFM_OK;
}
double oracle_calc_residual(gfm_ctrl *c, const gfm_host_arrays *h, const double *ap, const double *as,
                            const double *bs, const double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  if (bind(c, h)) return -1.0;
  memcpy(AP_, ap, sizeof(double) * NN);
  memcpy(AS_, as, sizeof(double) * NN * 6);
  memcpy(BS_, bs, sizeof(double) * NN);
  return CALC_RESIDUAL(xt, ib, ie, jb, je, kb, ke);
}
/* Expose the last assembled system of the named stage for the staged
 * parity tests: copies AP/AS/BS (and FZ) out. */
void oracle_scratch(double *ap, double *as, double *bs, double *fz) {
  if (ap) memcpy(ap, AP_, sizeof(double) * NN);
  if (as) memcpy(as, AS_, sizeof(double) * NN * 6);
  if (bs) memcpy(bs, BS_, sizeof(double) * NN);
  if (fz) memcpy(fz, FZ_, sizeof(double) * NN);
}
/* Per-cell r_mass of the last SOURCE(5) (C/Source.f90:76-95) on the
 * lattice, -1 at positions where none is formed (blocked, exit-adjacent,
 * non-centre).  Recording starts with the first call (enable = 1). */
int oracle_rmass(int enable, double *out) {
  if (enable && !RMASS_) {
    RMASS_ = calloc(NN, sizeof(double));
    if (!RMASS_) return GFM_E_DEVICE;
    for (size_t q = 0; q < NN; q++) RMASS_[q] = -1.0;
  }
  if (out && RMASS_) memcpy(out, RMASS_, sizeof(double) * NN);
  return GFM_OK;
}
/* Single stages of the iteration on the caller's state, for the tests that
 * pin the oracle against closed forms and invariants:
 *   1 CONC, 2 DNST_CALC, 3 THER at lattice position (i,j,k);
 *   4 q_conv_face(i,j,k, ds = io[0], T_wall = io[1] K) -> io[2];
 *   5 GFLOW up to and including the momentum assembly (GSLVX1 for u, v, w:
 *     AP/AS/BS of the face lattices, UG = pseudo-velocities);
 *   6 GFLOW up to and including the velocity correction (before EXTRV):
 *     UG corrected, FZ = p', BS = SOURCE(5)'s mass source, res_mass set;
 *   7 SOURCE of equation NELG = i over the cell box (BS, SFP; CRATE/HRATE
 *     for the fuel);
 *   8 INTP; 9 INTPV; 10 AREA_C(i,j,k, ix = io[0]) -> io[2]; 11 VOL_C(i,j,k) -> io[2].
 * AP/AS/BS/FZ/SFP are read back with oracle_scratch / oracle_sfp. */
int oracle_probe(gfm_ctrl *c, const gfm_host_arrays *h, int what, int i, int j, int k, double *io) {
  int rc = bind(c, h);
  if (rc) return rc;
  switch (what) {
    case 1: CONC(i, j, k); break;
    case 2: DNST_CALC(i, j, k); break;
    case 3: THER(i, j, k); break;
    case 4: io[2] = q_conv_face(i, j, k, io[0], io[1]); break;
    case 5:
    case 6:
      STOP_STAGE = what - 4;
      GFLOW();
      STOP_STAGE = 0;
      break;
    case 7:
      NELG = i;
      SOURCE(4, MPM2, NJY1, NJY2, LKZ1, LKZ2);
      break;
    case 8: INTP(); break;
    case 9: INTPV(); break;
    case 10: io[2] = AREA_C(i, j, k, (int)io[0]); break;
    case 11: io[2] = VOL_C(i, j, k); break;
    default: return GFM_E_ARG;
  }
  return GFM_OK;
}
void oracle_sfp(double *sfp) { memcpy(sfp, SFP_, sizeof(double) * NN); }
int oracle_abi_version(void) { return GFM_ABI_VERSION; }
size_t oracle_ctrl_size(void) { return sizeof(gfm_ctrl); }

/* SOURCE(5) alone on the caller's state (C/Source.f90:41-101): returns
 * res_mass, writes res_mass_max and its location into the ctrl block. */
double oracle_mass_residual(gfm_ctrl *c, const gfm_host_arrays *h) {
  if (bind(c, h)) return -1.0;
  NELG = 5;
  SOURCE(4, MPM2, NJY1, NJY2, LKZ1, LKZ2);
  return c->res_mass;
}

/* ======================================================================
 * Minor species and soot (MS = 1): INITSVM C/Initsv.f90:329-380, MSFLOW
 * C/Msflow.f90:10-169, GSLV3M C/Gslv3m.f90:8-286 with SPIN :292-305,
 * SOURCEM C/Sourcem.f90:8-122 (the new soot model, ioldsoot = 0),
 * flx1dm C/Flx1d.f90:114-158, extrm C/Extr.f90:258-305, intpm/intp0m
 * C/Intp.f90:348-419.  Constants of C/Var_mod.f90:215-235.
 * ====================================================================== */
static const double wms[7] = {16.0426, 31.9988, 28.0104, 18.0152, 44.0098, 28.0134, 30.0061};
static const double ck0_m[5] = {1.66e15, 7.94e14, 2.24e14, 7.94e22, 9.33e14};
static const double beta_m[5] = {0.0, 0.0, 0.0, 0.0, -0.5};
static const double odr[5][2] = {{1.46, 0.5217}, {1.6904, 1.57}, {1.0, 0.0}, {0.721, 4.0111}, {1.0, 0.5}};
static const double erc[5] = {20643.0, 11613.0, 62281.0, 44369.0, 68899.0};
static const double sml1 = 1.0e-25;

static gfm_ms_ctrl *MSC;
static double *GFM_, *SMF_, *BSSP_, *SFPSP_, *RRMS_, *AP1_;
static int MZ, NZ, LZ;
static double FLXM_[3][7], FLYM_[2][7]; /* FLXM(5,:,2), FLXM(MP-5,:,1), FLXM(MP-3,:,1); FLYM(5,:,2), FLYM(NP-3,:,1) */
#define GFM(i, j, k, l) GFM_[IX3(i, j, k) + NN * (size_t)((l)-1)]
#define SMF(i, j, k) SMF_[IX3(i, j, k)]
#define BSSP(i, j, k) BSSP_[IX3(i, j, k)]
#define SFPSP(i, j, k) SFPSP_[IX3(i, j, k)]
#define AP1(i, j, k) AP1_[IX3(i, j, k)]
#define RRMS(id, jd, kd, n) \
  RRMS_[(((size_t)(kd)-1) * (size_t)NZ + ((jd)-1)) * (size_t)MZ + ((id)-1) + (size_t)MZ * NZ * LZ * (size_t)((n)-1)]

static int ms_bind(gfm_ms_ctrl *ms, const gfm_ms_arrays *a) {
  if (!ms || !a || !a->gfm || !a->smf || !a->bs_soot_prev || !a->sfp_soot_prev || !a->rrms) return GFM_E_ARG;
  MSC = ms;
  GFM_ = a->gfm; SMF_ = a->smf; BSSP_ = a->bs_soot_prev; SFPSP_ = a->sfp_soot_prev; RRMS_ = a->rrms;
  MZ = MP / 2; NZ = NP / 2; LZ = LP / 2;
  static size_t cap = 0;
  if (cap != NN) {
    free(AP1_);
    AP1_ = calloc(NN, sizeof(double));
    if (!AP1_) return GFM_E_DEVICE;
    cap = NN;
  }
  return GFM_OK;
}

/* SPIN, C/Gslv3m.f90:292-305: the inflow of species lsp from the six
 * neighbours through the coefficients AS, in kmol/s */
static double SPIN(int i, int j, int k, int lsp) {
  double g11 = AS(i, j, k, 1, 1) * GFM(i - 2, j, k, lsp), g12 = AS(i, j, k, 1, 2) * GFM(i + 2, j, k, lsp);
  double g21 = AS(i, j, k, 2, 1) * GFM(i, j - 2, k, lsp), g22 = AS(i, j, k, 2, 2) * GFM(i, j + 2, k, lsp);
  double g31 = AS(i, j, k, 3, 1) * GFM(i, j, k - 2, lsp), g32 = AS(i, j, k, 3, 2) * GFM(i, j, k + 2, lsp);
  double f = g11 + g12 + g21 + g22 + g31 + g32;
  return f * C->gmfr / wms[lsp - 1];
}

/* The reaction rates of GSLV3M, C/Gslv3m.f90:22-115 (IRR0 = 5).  RR2 is a
 * local of GSLV3M that carries over from cell to cell: a cell whose CH4 or
 * O2 concentration is at most 1e-25 skips reaction 1 and then reduces the
 * O2 inflow of reaction 2 by 1.5 times the RR2 the previous open cell left
 * (its NO rate, reactions 4 + 5).  The reference leaves RR2 undefined at
 * the first cell of a call; here it starts at 0. */
static void MS_RATES(void) {
  const int mi1 = 4, mi3 = MPM2, nj1 = NJY1, nj3 = NJY2, lk1 = LKZ1, lk3 = LKZ2;
  double rr1, rr2 = 0.0, rr3, rr_f = 0.0, rr_c = 0.0;
  for (int i = mi1; i <= mi3; i += 2) {
    int id2 = i / 2;
    for (int j = nj1; j <= nj3; j += 2) {
      int jd2 = j / 2;
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) > 0) continue;
        int kd2 = k / 2;
        double tmp = T(i, j, k) * C->t0;
        double dnst1 = DNST(i, j, k) * C->dnst0;
        double vol = DX(i) * DR(j) * DZ(k) * C->vol0;
        double gfm1 = GFM(i, j, k, 1) / wms[0] * dnst1, gfm2 = GFM(i, j, k, 2) / wms[1] * dnst1;
        double gfm3 = GFM(i, j, k, 3) / wms[2] * dnst1, gfm5 = GFM(i, j, k, 5) / wms[4] * dnst1;
        double gfm6 = GFM(i, j, k, 6) / wms[5] * dnst1;
        /* reaction 1: CH4 + 1.5 O2 -> CO + 2 H2O */
        double fxms1 = SPIN(i, j, k, 1), fxms2 = SPIN(i, j, k, 2), fxms3, fxms5;
        if (gfm1 <= sml1 || gfm2 <= sml1) {
          RRMS(id2, jd2, kd2, 1) = 0.0;
        } else {
          rr1 = pow(gfm1, odr[0][0]) * pow(gfm2, odr[0][1]);
          rr2 = ck0_m[0] * rr1 * exp(-erc[0] / tmp) * vol;
          if (rr2 > fxms1) rr2 = fxms1;
          if (fxms2 / 1.5 < rr2) rr2 = fxms2 / 1.5;
          RRMS(id2, jd2, kd2, 1) = rr2 * 0.3 + RRMS(id2, jd2, kd2, 1) * 0.7;
        }
        /* reaction 2: CO + 0.5 O2 -> CO2 */
        fxms2 = fxms2 - rr2 * 1.5;
        if (fxms2 > small20 && gfm3 > small20) {
          rr1 = pow(gfm3, odr[1][0]) * pow(gfm2, odr[1][1]);
          rr2 = ck0_m[1] * rr1 * exp(-erc[1] / tmp) * vol;
          if (fxms2 * 2 < rr2) rr2 = fxms2 * 2;
          fxms3 = SPIN(i, j, k, 3);
          rr2 = dmin(rr2, fxms3);
        } else {
          rr2 = 0.0;
        }
        /* reaction 3: CO2 -> CO + 0.5 O2 (the net CO oxidation rate is RR2 - RR3) */
        if (gfm5 > small20) {
          rr1 = pow(gfm5, odr[2][0]);
          rr3 = ck0_m[2] * rr1 * exp(-erc[2] / tmp) * vol;
          fxms5 = SPIN(i, j, k, 5);
          rr3 = dmin(rr3, fxms5);
        } else {
          rr3 = 0.0;
        }
        rr1 = rr2 - rr3;
        if (rr1 < 0) rr1 = 0.0;
        RRMS(id2, jd2, kd2, 2) = rr1 * 0.3 + RRMS(id2, jd2, kd2, 2) * 0.7;
        rr_f += RRMS(id2, jd2, kd2, 1) * wms[0];
        rr_c += RRMS(id2, jd2, kd2, 2) * wms[2];
        /* reactions 4 and 5: N2 + O2 -> 2 NO */
        if (gfm2 > small20 && gfm3 > small20 && gfm6 > small20) {
          rr1 = pow(gfm3, odr[3][0]) * pow(gfm2, odr[3][1]) * GFM(i, j, k, 6) / 0.725;
          rr2 = ck0_m[3] * rr1 * exp(-erc[3] / tmp) * vol;
        } else {
          rr2 = 0.0;
        }
        if (gfm2 > small20 && gfm6 > small20) {
          rr1 = pow(gfm6, odr[4][0]) * pow(gfm2, odr[4][1]);
          rr3 = ck0_m[4] * rr1 * exp(-erc[4] / tmp) * pow(tmp, beta_m[4]);
          rr3 = rr3 * vol;
        } else {
          rr3 = 0.0;
        }
        rr2 = rr2 + rr3;
        RRMS(id2, jd2, kd2, 4) = rr2;
      }
    }
  }
  MSC->rr_f = rr_f;
  MSC->rr_c = rr_c;
}

/* SOURCEM, C/Sourcem.f90:8-122.  nelm 1..7: BS = the species' net
 * production (kg/s, volume included in the rates) / GMFR, SFP = 0.
 * nelm 11: the soot formation (BS, per volume) and oxidation (SFP)
 * sources of the new model, relaxed against the previous call's after the
 * first minor-species iteration; both arrays are cleared everywhere first
 * and saved whole as the next call's history. */
static void SOURCEM(int nelm, int mi1, int mi3, int nj1, int nj3, int lk1, int lk3) {
  if (nelm != 11) {
    for (int i = mi1; i <= mi3; i += 2) {
      int id2 = i / 2;
      for (int j = nj1; j <= nj3; j += 2) {
        int jd2 = j / 2;
        for (int k = lk1; k <= lk3; k += 2) {
          SFP(i, j, k) = 0.0;
          BS(i, j, k) = 0.0;
          if (IBCELL(i, j, k) > 0) continue;
          int kd2 = k / 2;
          double g0;
          switch (nelm) {
            case 1: g0 = -RRMS(id2, jd2, kd2, 1); break;
            case 2: g0 = -RRMS(id2, jd2, kd2, 1) * 1.5 - RRMS(id2, jd2, kd2, 2) * 0.5 - RRMS(id2, jd2, kd2, 4); break;
            case 3: g0 = RRMS(id2, jd2, kd2, 1) - RRMS(id2, jd2, kd2, 2); break;
            case 4: g0 = RRMS(id2, jd2, kd2, 1) * 2.0; break;
            case 5: g0 = RRMS(id2, jd2, kd2, 2); break;
            case 6: g0 = -RRMS(id2, jd2, kd2, 4); break;
            default: g0 = RRMS(id2, jd2, kd2, 4) * 2.0; break;
          }
          BS(i, j, k) = g0 * wms[nelm - 1] / C->gmfr;
        }
      }
    }
    return;
  }
  const double snorm = C->r0 / (C->dnst0 * C->ug0);
  const double e0_sf = (MSC->t0_sf < small20) ? 0.0 : exp(-MSC->esf_r / MSC->t0_sf);
  const double e0_so = (MSC->t0_so < small20) ? 0.0 : sqrt(MSC->t0_so) * exp(-MSC->eso_r / MSC->t0_so);
  memset(BS_, 0, sizeof(double) * NN);
  memset(SFP_, 0, sizeof(double) * NN);
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        double tcell = T(i, j, k) * C->t0;
        if (IBCELL(i, j, k) > 0 || !MSC->react) continue;
        if (tcell > MSC->t0_sf) {
          BS(i, j, k) = MSC->aform * GFM(i, j, k, 1) * (exp(-MSC->esf_r / tcell) - e0_sf) * snorm;
          if (MSC->itr_ms > 1)
            BS(i, j, k) = MSC->rf_soot_form_source * BS(i, j, k) + (1.0 - MSC->rf_soot_form_source) * BSSP(i, j, k);
        }
        if (tcell > MSC->t0_so) {
          SFP(i, j, k) = -(MSC->aoxid * GFM(i, j, k, 2) * (sqrt(tcell) * exp(-MSC->eso_r / tcell) - e0_so) * snorm);
          if (MSC->itr_ms > 1)
            SFP(i, j, k) = MSC->rf_soot_oxid_source * SFP(i, j, k) + (1.0 - MSC->rf_soot_oxid_source) * SFPSP(i, j, k);
        }
      }
  memcpy(BSSP_, BS_, sizeof(double) * NN);
  memcpy(SFPSP_, SFP_, sizeof(double) * NN);
}

/* GSLV3M, C/Gslv3m.f90:8-286 */
static void GSLV3M(void) {
  const int mi1 = 4, mi3 = MPM2, nj1 = NJY1, nj3 = NJY2, lk1 = LKZ1, lk3 = LKZ2;
  MS_RATES();
  for (int l = 1; l <= 7; l++) {
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2)
        for (int k = 2; k <= LP; k += 2) FZ(i, j, k) = GFM(i, j, k, l);
    SOURCEM(l, mi1, mi3, nj1, nj3, lk1, lk3);
    SGM = MSC->ssigma;
    if (MSC->igs0 == 0) {
      GSLV3G0(mi1, mi3, nj1, nj3, lk1, lk3, 0);
      memcpy(AP1_, AP_, sizeof(double) * NN);
      MSC->igs0 = 1;
    } else {
      memcpy(AP_, AP1_, sizeof(double) * NN);
    }
    const double rfm = MSC->rfm[l - 1], rfmc = 1.0 - MSC->rfm[l - 1];
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          double vol = VOL_C(i, j, k);
          AP(i, j, k) = AP(i, j, k) - (SFP(i, j, k) * vol);
          AP(i, j, k) = AP(i, j, k) / rfm;
          BS(i, j, k) = BS(i, j, k) + rfmc * FZ(i, j, k) * AP(i, j, k);
        }
    MSC->resid_ms_pre[l - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    ADLBL3_SWEEPS(FZ_, MSC->ntimesm, mi1, mi3, nj1, nj3, lk1, lk3);
    MSC->resid_ms[l - 1] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
    for (int i = mi1; i <= mi3; i += 2)
      for (int j = nj1; j <= nj3; j += 2)
        for (int k = lk1; k <= lk3; k += 2) {
          if (IBCELL(i, j, k) >= 1) continue;
          if (FZ(i, j, k) > 1.0) FZ(i, j, k) = 1.0;
          if (FZ(i, j, k) < 0) FZ(i, j, k) = 0.0;
          if (l == 3 && FZ(i, j, k) > 0.1) FZ(i, j, k) = 0.1;
          if (l == 7 && FZ(i, j, k) > 1.0e-2) FZ(i, j, k) = 1.0e-2;
          GFM(i, j, k, l) = rfmc * GFM(i, j, k, l) + rfm * FZ(i, j, k);
          if (GFM(i, j, k, l) < small20) GFM(i, j, k, l) = 0.0;
        }
  }
  double bal = 0.0;
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) continue;
        double g0 = 0.0;
        for (int l = 1; l <= 7; l++) g0 += GFM(i, j, k, l);
        bal += fabs(1.0 - g0) * VOL_C(i, j, k);
      }
  MSC->bal_ms = bal * C->vol0 / MSC->vol_tot;

  if (!MSC->react) return;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) FZ(i, j, k) = SMF(i, j, k);
  SOURCEM(11, mi1, mi3, nj1, nj3, lk1, lk3);
  SGM = MSC->ssigma;
  GSLV3G0(mi1, mi3, nj1, nj3, lk1, lk3, 1);
  const double rfm = MSC->rfm[7], rfmc = 1.0 - MSC->rfm[7];
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) continue;
        double vol = VOL_C(i, j, k);
        BS(i, j, k) = BS(i, j, k) * vol;
        AP(i, j, k) = AP(i, j, k) - (SFP(i, j, k) * vol);
        AP(i, j, k) = AP(i, j, k) / rfm;
        BS(i, j, k) = BS(i, j, k) + rfmc * FZ(i, j, k) * AP(i, j, k);
      }
  MSC->resid_ms_pre[7] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
  ADLBL3_SWEEPS(FZ_, MSC->ntimesm, mi1, mi3, nj1, nj3, lk1, lk3);
  MSC->resid_ms[7] = CALC_RESIDUAL(FZ_, mi1, mi3, nj1, nj3, lk1, lk3);
  double soot_tot = 0.0, smf_max = 0.0, smf_max_noclip = 0.0;
  int clipped = 0;
  for (int i = mi1; i <= mi3; i += 2)
    for (int j = nj1; j <= nj3; j += 2)
      for (int k = lk1; k <= lk3; k += 2) {
        if (IBCELL(i, j, k) >= 1) continue;
        smf_max_noclip = dmax(smf_max_noclip, FZ(i, j, k));
        double g0;
        if (FZ(i, j, k) > MSC->soot_cap) { g0 = MSC->soot_cap; clipped++; }
        else if (FZ(i, j, k) > small20) g0 = FZ(i, j, k);
        else g0 = 0.0;
        smf_max = dmax(smf_max, g0);
        SMF(i, j, k) = rfm * g0 + rfmc * SMF(i, j, k);
        double vol = DX(i) * DR(j) * DZ(k);
        soot_tot = soot_tot + SMF(i, j, k) * DNST(i, j, k) * C->dnst0 * vol;
      }
  MSC->soot_tot = soot_tot;
  MSC->smf_max = smf_max;
  MSC->smf_max_noclip = smf_max_noclip;
  MSC->smf_clipped_cnt = clipped;
}

/* flx1dm, C/Flx1d.f90:114-158: the species flow (kg/s) through x-face
 * planes (+x in FLXM(:,:,1), -x in (:,:,2)) and y-face planes; kept for the
 * planes MSFLOW reads (x: 5, MP-5, MP-3; y: 5, NP-3) */
static void FLX1DM(void) {
  memset(FLXM_, 0, sizeof FLXM_);
  memset(FLYM_, 0, sizeof FLYM_);
  const int xi[3] = {5, MP - 5, MP - 3}, xd[3] = {2, 1, 1}, yj[2] = {5, NP - 3}, yd[2] = {2, 1};
  for (int i = 3; i <= MPM1; i += 2)
    for (int j = 4; j <= NPM2; j += 2)
      for (int k = LKZ1; k <= LKZ2; k += 2) {
        if (IBCELL(i, j, k) == 1) continue;
        double area = AREA_C(i, j, k, 1);
        double g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, 1) * area * C->gmfr;
        int id = (g0 >= 0.0) ? 1 : 2;
        for (int q = 0; q < 3; q++)
          if (i == xi[q] && id == xd[q])
            for (int l = 1; l <= 7; l++) FLXM_[q][l - 1] += fabs(g0) * GFM(i, j, k, l);
      }
  for (int j = 3; j <= NPM1; j += 2)
    for (int i = 4; i <= MPM2; i += 2)
      for (int k = LKZ1; k <= LKZ2; k += 2) {
        if (IBCELL(i, j, k) == 1) continue;
        double area = AREA_C(i, j, k, 2);
        double g0 = THETA(i, j, k) * DNST(i, j, k) * UG(i, j, k, 2) * area * C->gmfr;
        int id = (g0 >= 0.0) ? 1 : 2;
        for (int q = 0; q < 2; q++)
          if (j == yj[q] && id == yd[q])
            for (int l = 1; l <= 7; l++) FLYM_[q][l - 1] += fabs(g0) * GFM(i, j, k, l);
      }
}

/* extrm, C/Extr.f90:258-305: exit cells extrapolate the species from the
 * interior (floored at 0; the y-face toward the interior cell takes the
 * same value) and copy smf; wall cells copy their non-wall neighbour and
 * the face between.  Wall cells without a non-wall neighbour copy the
 * cell two steps up in k as it was before this loop reached it. */
static void EXTRM(void) {
  int i1 = 2, j1 = 2, k1 = 2, i2 = 2, j2 = 2, k2 = 2, ix = 4;
  double f1, f2;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (imod10(IBCELL(i, j, k)) != 3) continue;
        EXTIDX(0, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2, &ix);
        EXTR1(&f1, &f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
        for (int l = 1; l <= 7; l++) {
          double g0 = GFM(i1, j1, k1, l) * f1 + GFM(i2, j2, k2, l) * f2;
          GFM(i, j, k, l) = dmax(g0, 0.0);
          int jj = (j + j1) / 2;
          GFM(i, jj, k, l) = GFM(i, j, k, l);
        }
        SMF(i, j, k) = SMF(i1, j1, k1);
      }
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (imod10(IBCELL(i, j, k)) != 1) continue;
        EXTIDX(1, i, j, k, &i1, &j1, &k1, &i2, &j2, &k2, &ix);
        if (ix == 4) { i1 = i; j1 = j; k1 = imin(LP, k + 2); } /* EXT1's indices after an unsuccessful search */
        EXTR1(&f1, &f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
        for (int l = 1; l <= 7; l++) {
          GFM(i, j, k, l) = GFM(i1, j1, k1, l);
          if (ix == 1) GFM((i + i1) / 2, j, k, l) = GFM(i, j, k, l);
          else if (ix == 2) GFM(i, (j + j1) / 2, k, l) = GFM(i, j, k, l);
          else GFM(i, j, (k + k1) / 2, l) = GFM(i, j, k, l);
        }
      }
}

/* intp0m, C/Intp.f90:396-419 */
static void INTP0M(double *f1, int i, int j, int k, int i1, int j1, int k1, int i2, int j2, int k2) {
  if (IBCELL(i, j, k) == 3) return;
  if (IBCELL(i, j, k) == 2) {
    if (IBCELL(i1, j1, k1) == 2 && IBCELL(i2, j2, k2) == 0) *f1 = 1.0;
    else if (IBCELL(i1, j1, k1) == 0 && IBCELL(i2, j2, k2) == 2) *f1 = 0.0;
  }
  double a = *f1, b = 1.0 - a;
  for (int l = 1; l <= 7; l++) GFM(i, j, k, l) = a * GFM(i1, j1, k1, l) + b * GFM(i2, j2, k2, l);
}
/* intpm, C/Intp.f90:348-390 (the loop nest of INTP) */
static void INTPM(void) {
  for (int i = 2; i <= MP; i += 2) {
    for (int k = 3; k <= LPM1; k += 2) {
      double f1 = INTP1(i, 0, k, 3);
      for (int j = 2; j <= NP; j += 2) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0M(&f1, i, j, k, i, j, k + 1, i, j, k - 1);
      }
    }
    for (int j = 3; j <= NPM1; j += 2) {
      double f1 = INTP1(i, j, 0, 2);
      for (int k = 2; k <= LP; k++) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0M(&f1, i, j, k, i, j + 1, k, i, j - 1, k);
      }
    }
  }
  for (int i = 3; i <= MPM1; i += 2) {
    double f1 = INTP1(i, 0, 0, 1);
    for (int k = 2; k <= LP; k++)
      for (int j = 2; j <= NP; j++) {
        if (IBCELL(i, j, k) == 1) continue;
        INTP0M(&f1, i, j, k, i + 1, j, k, i - 1, j, k);
      }
  }
}

/* INITSVM, C/Initsv.f90:329-380.  restart = 0: new start; 1: the arrays
 * hold the restart record (cell centres of GFM and smf, the soot-source
 * history, RRMS; itr_ms in the control block). */
static void INITSVM(int restart) {
  if (!restart) {
    memset(SMF_, 0, sizeof(double) * NN);
    memset(GFM_, 0, sizeof(double) * NN * 7);
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2)
        for (int k = 2; k <= LP; k += 2) {
          GFM(i, j, k, 1) = GF(i, j, k, iyf);
          GFM(i, j, k, 2) = GF(i, j, k, iyo2);
          GFM(i, j, k, 4) = GF(i, j, k, iyh2o);
          GFM(i, j, k, 5) = GF(i, j, k, iyco2);
          GFM(i, j, k, 6) = GF(i, j, k, iyn2);
        }
  } else {
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2)
        for (int k = 2; k <= LP; k += 2) {
          if (IBCELL(i, j, k) != 2) continue;
          GFM(i, j, k, 1) = GF(i, j, k, iyf);
          GFM(i, j, k, 2) = GF(i, j, k, iyo2);
          GFM(i, j, k, 4) = GF(i, j, k, iyh2o);
          GFM(i, j, k, 5) = GF(i, j, k, iyco2);
          GFM(i, j, k, 6) = GF(i, j, k, iyn2);
          GFM(i, j, k, 3) = 0.0;
          GFM(i, j, k, 7) = 0.0;
        }
    EXTRM();
  }
  MSC->igs0 = 0;
  INTPM();
}

/* MSFLOW, C/Msflow.f90:29-162 (without the screen, file and GUI output) */
static void MSFLOW(void) {
  double gfm0[7];
  FLX1DM();
  for (int l = 0; l < 7; l++) gfm0[l] = FLXM_[0][l] + FLXM_[1][l];
  MSC->igs0 = 0;
  MSC->iterations = 0;
  MSC->converged = 0;
  for (int it = 1; it <= MSC->maxms; it++) {
    MSC->itr_ms += 1;
    MSC->iterations = it;
    GSLV3M();
    FLX1DM();
    double gfmex = 0.0, chg = 0.0;
    for (int l = 0; l < 7; l++) {
      double g = FLXM_[0][l] + FLXM_[2][l] + FLYM_[0][l] + FLYM_[1][l];
      MSC->gfm_0[l] = g;
      gfmex += g;
      chg += fabs(g - gfm0[l]);
      gfm0[l] = g;
    }
    MSC->gfmex = gfmex;
    MSC->ave_chg = chg / 7;
    double svf = 0.0, g1 = 0.0;
    for (int i = 2; i <= MP; i += 2)
      for (int j = 2; j <= NP; j += 2)
        for (int k = 2; k <= LP; k += 2) {
          if (IBCELL(i, j, k) != 3) continue;
          double g0 = AREA_C(i, j, k, 3);
          if (g0 < small20) g0 = AREA_C(i, j, k, 2);
          if (g0 < small20) g0 = AREA_C(i, j, k, 1);
          g1 += g0;
          svf += SMF(i, j, k) * DNST(i, j, k) * C->dnst0 / MSC->dnst_soot * g0;
        }
    MSC->svf_r = svf / g1;
    int done = 0;
    if (MSC->itr_ms > 10) {
      done = 1;
      for (int l = 0; l < 8; l++)
        if (MSC->resid_ms[l] > MSC->bmcon) done = 0;
    }
    if (done) { MSC->converged = 1; break; }
    EXTRM();
    INTPM();
  }
}

int oracle_initsvm(gfm_ctrl *c, const gfm_host_arrays *h, gfm_ms_ctrl *ms, const gfm_ms_arrays *a, int restart) {
  int rc = bind(c, h);
  if (!rc) rc = ms_bind(ms, a);
  if (rc) return rc;
  INITSVM(restart);
  return GFM_OK;
}
int oracle_gslv3m(gfm_ctrl *c, const gfm_host_arrays *h, gfm_ms_ctrl *ms, const gfm_ms_arrays *a) {
  int rc = bind(c, h);
  if (!rc) rc = ms_bind(ms, a);
  if (rc) return rc;
  GSLV3M();
  return GFM_OK;
}
int oracle_msflow(gfm_ctrl *c, const gfm_host_arrays *h, gfm_ms_ctrl *ms, const gfm_ms_arrays *a) {
  int rc = bind(c, h);
  if (!rc) rc = ms_bind(ms, a);
  if (rc) return rc;
  MSFLOW();
  return GFM_OK;
}
/* single stages for the pins: 1 the rates (RRMS and rr_f/rr_c), 2 SOURCEM
 * of species nelm (BS/SFP via oracle_scratch/oracle_sfp), 3 extrm, 4 intpm,
 * 5 flx1dm -> out[0..34] = FLXM(5,:,2), FLXM(MP-5,:,1), FLXM(MP-3,:,1),
 * FLYM(5,:,2), FLYM(NP-3,:,1) */
int oracle_ms_probe(gfm_ctrl *c, const gfm_host_arrays *h, gfm_ms_ctrl *ms, const gfm_ms_arrays *a, int what, int nelm,
                    double *out) {
  int rc = bind(c, h);
  if (!rc) rc = ms_bind(ms, a);
  if (rc) return rc;
  switch (what) {
    case 1: MS_RATES(); break;
    case 2: SOURCEM(nelm, 4, MPM2, NJY1, NJY2, LKZ1, LKZ2); break;
    case 3: EXTRM(); break;
    case 4: INTPM(); break;
    case 5:
      FLX1DM();
      for (int l = 0; l < 7; l++) {
        out[l] = FLXM_[0][l]; out[7 + l] = FLXM_[1][l]; out[14 + l] = FLXM_[2][l];
        out[21 + l] = FLYM_[0][l]; out[28 + l] = FLYM_[1][l];
      }
      break;
    default: return GFM_E_ARG;
  }
  return GFM_OK;
}
size_t oracle_ms_ctrl_size(void) { return sizeof(gfm_ms_ctrl); }
/* Zero the oracle's coefficient arrays AP, AS, BS (the state of a device
 * context before its first assembly), for comparisons with a fresh one. */
void oracle_reset_coefficients(void) {
  if (AP_) memset(AP_, 0, sizeof(double) * NN);
  if (AS_) memset(AS_, 0, sizeof(double) * NN * 6);
  if (BS_) memset(BS_, 0, sizeof(double) * NN);
}

/* ======================================================================
 * Radiation transport of one full radiation pass, RAD_EMIS C/Rad_emis.f90:
 * 130-229 up to QRNF's volume loop: the wall marks (C/Rad_emis.f90:155-
 * 186), RAD0 (C/Rad0.f90:12-172) with RAD_GAS, AH2O, ACO2, A1, CHIWHI,
 * BANDA (C/Rad_gas.f90:16-396), SOOTA, SOC (C/Rad_soot.f90:26-69), PER,
 * RIF, EMF, DEBX (C/Rad_fun.f90:14-140); QRNF's emitter loop (C/Qrn.f90:
 * 58-200) with QREF (C/Qre.f90:31-505) and QRTX, QRTY, QRTZ
 * (C/Qrtx.f90:15-255, C/Qrty.f90:9-224, C/Qrtz.f90:15-234); QRN1's sums
 * (C/Qrn.f90:245-290).  rad_qe_only (C/Rad_emis.f90:411-466) is RAD0 with
 * ID_RAD = 1.  Constants of C/Var_mod.f90:133-137, :18.
 *
 * Reference state that carries over and is kept here:
 *  - AKLS (soot absorption) is zeroed once per RAD0 call and SOOTA leaves
 *    it unchanged in a cell with a soot volume fraction below 1e-14, so
 *    such a cell takes the AKLS of the last cell before it (loop order I,
 *    J, K innermost) that had soot;
 *  - QTX1, Q2M, QTX2 and their flags are zeroed once per axis of an
 *    emitter, not between its two directions;
 *  - the wall split of the second spreading pass of QRTX (QRTY, QRTZ)
 *    writes next to the column J2 = J1 -+ 2 (I2 = I1 -+ 2) toward the
 *    emitter; on the emitter's own column J2 (I2) is not assigned and keeps
 *    the value it last had: from an earlier split, or as the first pass's
 *    inner loop variable.  The reference's local variables are static (one
 *    set per subroutine), so that value is carried from call to call; here
 *    too, and RAD_STALE_PREV counts the uses whose value came from an
 *    earlier emitter.
 * ====================================================================== */
static const double hck = 1.439e-2, t0_r = 1.0e2, p0_r = 1.01325e5, pg0_r = 1.01325e5, sig_sb = 5.668e-8;
static const double pi_r = 3.141592653589793;
static gfm_rad_ctrl *RC;
static int NWL, ID_RAD;
static int16_t *IBR_;                   /* cell codes with the wall marks (a copy; the caller's are unchanged) */
static double *AKL_, *QEL_, *QAL_, *QE_, *QA_;
static double WL[GFM_NWL_MAX], DEB[GFM_NWL_MAX], DEB0[GFM_NWL_MAX], AKLG[GFM_NWL_MAX], AKLS[GFM_NWL_MAX],
    AKL0[GFM_NWL_MAX], E0R[GFM_NWL_MAX], QEL0[GFM_NWL_MAX], QEL1[GFM_NWL_MAX];
static double SGN, T_RAD, TT0, P_RAD, YCO2, YH2O, FVS, OPL, W_R, PE;
static double WNH2O[5][3], WNCO2[6][3], WNG[11][2], AAH2O[5], AACO2[6];
static double ANG0[3];
static double *ANGA_[3];                /* ANG1(MP,2), ANG2(NP,2), ANG3(LP,2): [axis][idx-1 + M*(col-1)] */
static double *QTX1_, *Q2M_, *QTX2_;    /* plane arrays of the current axis, (b/2, c/2, L) */
static unsigned char *IQT1_, *IQTM_, *IQT2_;
static int PLB, PLC;                    /* plane extents of the current axis (cells) */
static double qa_sum_bnd, qa_sum_vol, qa_to_open_c, qa_to_melt_c, qa_to_inlet_c, qa_to_exit_c, qa_to_wall_c;
static int RAD_STALE[3];                /* part-1 inner loop variable of QRTX, QRTY, QRTZ (static in the reference) */
static int RAD_STALE_SET[3];            /* set during the current call */
long long RAD_STALE_PREV = 0, RAD_STALE_USES = 0, RAD_OOB = 0;

#define IBR(i, j, k) ((int)IBR_[IX3(i, j, k)])
#define CIX(id, jd, kd) ((((size_t)(kd)-1) * (size_t)NZ + ((size_t)(jd)-1)) * (size_t)MZ + ((size_t)(id)-1))
#define CNN ((size_t)MZ * NZ * LZ)
#define AKL(id, jd, kd, l) AKL_[CIX(id, jd, kd) + CNN * (size_t)((l)-1)]
#define QEL(id, jd, kd, l) QEL_[CIX(id, jd, kd) + CNN * (size_t)((l)-1)]
#define QAL(id, jd, kd, l) QAL_[CIX(id, jd, kd) + CNN * (size_t)((l)-1)]
#define QE(id, jd, kd) QE_[CIX(id, jd, kd)]
#define QA(id, jd, kd) QA_[CIX(id, jd, kd)]
#define PIX(bd, cd) (((size_t)(cd)-1) * (size_t)PLB + ((size_t)(bd)-1))
#define QTX1(bd, cd, l) QTX1_[PIX(bd, cd) + (size_t)PLB * PLC * ((l)-1)]
#define Q2M(bd, cd, l) Q2M_[PIX(bd, cd) + (size_t)PLB * PLC * ((l)-1)]
#define QTX2(bd, cd, l) QTX2_[PIX(bd, cd) + (size_t)PLB * PLC * ((l)-1)]
#define IQT1(bd, cd) IQT1_[PIX(bd, cd)]
#define IQTM(bd, cd) IQTM_[PIX(bd, cd)]
#define IQT2(bd, cd) IQT2_[PIX(bd, cd)]

static int axis_len(int d) { return d == 1 ? MP : d == 2 ? NP : LP; }
static double axis_x(int d, int n) { return d == 1 ? X(n) : d == 2 ? R(n) : Z(n); }
static double *ANGP(int d, int n, int col) { return &ANGA_[d - 1][(size_t)(n - 1) + (size_t)axis_len(d) * (col - 1)]; }
#define ANG(d, n, col) (*ANGP(d, n, col))
/* atand (the degree arc tangent the reference calls) of h >= 0 by reduction
 * to |t| <= tan(pi/8) (1/h above 1, (h-1)/(h+1) above tan(pi/8)) and the
 * series t * sum_{n=0..20} (-t^2)^n/(2n+1) in Horner form: the device code
 * evaluates the same operations in the same order without contraction, so
 * the many angle comparisons of QRT that tie in exact arithmetic resolve
 * the same way on both. */
static double rad_atan(double h) {
  double off = 0.0, sgn = 1.0, off2 = 0.0;
  if (h > 1.0) { h = 1.0 / h; off = pi_r / 2.0; sgn = -1.0; }
  double t = h;
  if (h > 0.41421356237309503) { t = (h - 1.0) / (h + 1.0); off2 = pi_r / 4.0; }
  const double t2 = t * t;
  double s = 1.0 / 41.0;
  for (int n = 19; n >= 0; n--) s = 1.0 / (double)(2 * n + 1) - t2 * s;
  return off + sgn * (off2 + t * s);
}
static double atand_(double h) { return rad_atan(h) * (180.0 / pi_r); }
static double ipow3(double x) { return x * x * x; }
static double ipow4(double x) { double x2 = x * x; return x2 * x2; }

/* PER, C/Rad_fun.f90:14-40: m!/(m-n)! */
static double PER(int m, int n) {
  double per = 1.0;
  if (m < n) return NAN;
  if (m == 0 || n == 0) return per;
  for (int i = m - n + 1; i <= m; i++) per *= i;
  return per;
}
/* RIF, C/Rad_fun.f90:45-91: (15/pi^4) * sum_n e^{-(a+n)x} sum_i m!/(m-i)! x^{m-i}/(a+n)^{i+1},
 * the blackbody fraction above x = hc/(k lambda T) for m = 3, a = 0 */
static double RIF(int m, double a, double x) {
  double rif = 0.0, y[8];
  if (m < 0 || a < 0.0 || m > 7) return NAN;
  if (x > 1.0e16) return 0.0;
  if (x == 0.0) {
    int n = (int)(pow(0.5e6, 1.0 / (m + 1)) - a);
    double g0 = 0.0;
    for (int i = 1; i <= n; i++) g0 += 1.0 / pow(a + i, m + 1);
    rif = PER(m, m) * (g0 + 0.5 * (1.0 / pow(a + n + 1, m) + 1.0 / pow(a + n, m)) / m);
  } else {
    for (int i = 0; i <= m; i++) y[i] = PER(m, i);
    for (int n = 1; n <= 1000; n++) {
      double g0 = 0.0;
      for (int i = 0; i <= m; i++) {
        double xp = 1.0, ap = 1.0;
        for (int q = 0; q < m - i; q++) xp *= x;
        for (int q = 0; q < i + 1; q++) ap *= (a + n);
        g0 += y[i] * xp / ap;
      }
      g0 *= exp(-(a + n) * x);
      rif += g0;
      if (fabs(g0) <= small16) break;
    }
  }
  rif = rif * 1.5e1 / ipow4(pi_r);
  if (rif > 1.0) rif = 1.0;
  return rif;
}
/* DEBX, C/Rad_fun.f90:122-140: the fraction of the blackbody emission at
 * tg0 in each band (bands below 1e-3 join the next) */
static void DEBX(double tg0) {
  double eb0 = 0.0;
  for (int l = 1; l <= NWL - 1; l++) {
    double x1 = DEB0[l - 1] / tg0;
    double eb1 = RIF(3, 0.0, x1);
    DEB[l - 1] = eb1 - eb0;
    if (DEB[l - 1] <= 1.0e-3) DEB[l - 1] = 0.0;
    else eb0 = eb1;
  }
  DEB[NWL - 1] = 1.0 - eb0;
}
/* CHIWHI, C/Rad_gas.f90:308-353 (Weiner and Edwards 1968, eqs A16-A20) */
static void CHIWHI(double tt, const double v[3], const int ng[3], const int nd[3], double *chi, double *whi) {
  if (nd[0] != 0 || nd[1] != 0 || nd[2] != 0) {
    double ud = 0.0, xd = 0.0, xy = 0.0, yn = 0.0;
    for (int k = 0; k < 3; k++) {
      double u = hck * v[k] / tt;
      int nv = 0;
      if (nd[k] < 0) nv = abs(nd[k]);
      int ndk = abs(nd[k]);
      double y1 = 0.0;
      for (;;) {
        double xgas = exp(-u * nv) / PER(ng[k] - 1, ng[k] - 1);
        double x1 = xgas * PER(nv + ng[k] + ndk - 1, ng[k] + ndk - 1);
        double x2 = xgas * PER(nv + ng[k] - 1, ng[k] - 1);
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
    *chi = (1.0 - exp(ud)) * xy / xd;
    *whi = yn / xy / sqrt(TT0);
  } else {
    *chi = 1.0;
    *whi = 1.0 / sqrt(TT0);
  }
}
/* BANDA, C/Rad_gas.f90:360-396 (Edwards and Menard 1964); a negative
 * argument stops the reference's run: here it gives NAN */
static double BANDA(double alp, double bta, double a0) {
  if (a0 <= 0.0) return 0.0;
  if (alp < 0.0 || bta < 0.0 || a0 < 0.0 || W_R * PE < 0.0) return NAN;
  if (W_R * PE == 0.0) return 0.0;
  double eta = bta * PE, tau = alp * W_R / a0, a_s;
  if (eta >= 1.0) {
    a_s = (tau <= 1.0) ? tau : 1.0 + log(tau);
  } else {
    if (tau <= eta) a_s = tau;
    else if (tau <= 1.0 / eta) a_s = 2.0 * sqrt(eta * tau) - eta;
    else a_s = log(eta * tau) + 2.0 - eta;
  }
  return a0 * a_s;
}
/* A1, C/Rad_gas.f90:284-300 */
static double A1R(const double v[3], const int ng[3], const double b[3], const int nd[3]) {
  double chi0, whi0, chi, whi;
  CHIWHI(t0_r, v, ng, nd, &chi0, &whi0);
  CHIWHI(T_RAD, v, ng, nd, &chi, &whi);
  double c3 = b[2] * sqrt(T_RAD / t0_r);
  double c1 = b[0] * chi / chi0;
  double c2 = 2.0 * sqrt(c1 * c3 * b[1] * whi / whi0);
  return BANDA(c1, c2, c3);
}
/* AH2O, C/Rad_gas.f90:149-213: band absorptances of H2O (1/m) */
static void AH2O(void) {
  static const int ng[3] = {1, 1, 1};
  static const double v[3] = {3.652e5, 1.595e5, 3.756e5};
  double aa[5], b[3], chi0, whi0, chi, whi;
  int nd[3];
  for (int n = 0; n < 5; n++) AAH2O[n] = 0.0;
  if (YH2O <= 0.0 || T_RAD <= 3.0e2) return;
  PE = P_RAD + P_RAD * YH2O * (0.86e1 / TT0 - 0.5);
  W_R = P_RAD * YH2O * OPL / T_RAD / 8.208e-2 * wh2o;
  b[0] = 0.52e4 * exp(-0.176e2 / TT0); b[1] = 0.14311; b[2] = 0.284e2;
  nd[0] = nd[1] = nd[2] = 0;
  aa[0] = A1R(v, ng, b, nd);
  b[0] = 0.412e2; b[1] = 0.9427e-1; b[2] = 0.564e2;
  nd[1] = 1;
  aa[1] = A1R(v, ng, b, nd);
  b[0] = 0.19; b[1] = 0.13219; b[2] = 0.6e2;
  nd[1] = 2;
  CHIWHI(t0_r, v, ng, nd, &chi0, &whi0);
  CHIWHI(T_RAD, v, ng, nd, &chi, &whi);
  double t1 = b[0] * chi / chi0, t4 = b[1] * whi / whi0;
  b[0] = 0.23e1;
  nd[0] = 1; nd[1] = 0;
  CHIWHI(t0_r, v, ng, nd, &chi0, &whi0);
  CHIWHI(T_RAD, v, ng, nd, &chi, &whi);
  double t2 = b[0] * chi / chi0, t5 = b[1] * whi / whi0;
  b[0] = 0.224e2;
  nd[0] = 0; nd[2] = 1;
  CHIWHI(t0_r, v, ng, nd, &chi0, &whi0);
  CHIWHI(T_RAD, v, ng, nd, &chi, &whi);
  double t3 = b[0] * chi / chi0, t6 = b[1] * whi / whi0;
  double c1 = t1 + t2 + t3, c3 = b[2] * TT0;
  double bta = 2.0 * sqrt(c3) * (sqrt(t1 * t4) + sqrt(t2 * t5) + sqrt(t3 * t6));
  aa[2] = BANDA(c1, bta, c3);
  b[0] = 3.0; b[1] = 0.8169e-1; b[2] = 0.431e2;
  nd[1] = 1;
  aa[3] = A1R(v, ng, b, nd);
  b[0] = 0.25e1; b[1] = 0.11628; b[2] = 0.32e2;
  nd[0] = 1; nd[1] = 0;
  aa[4] = A1R(v, ng, b, nd);
  for (int n = 0; n < 5; n++) AAH2O[n] = aa[n] * 1.0e2;
}
/* ACO2, C/Rad_gas.f90:224-277: band absorptances of CO2 (1/m); with no
 * CO2 path AACO2 keeps its last values, as in the reference */
static void ACO2(void) {
  static const int ng[3] = {1, 2, 1};
  static const double v[3] = {1.351e5, 6.67e4, 2.396e5};
  double aa[6] = {0, 0, 0, 0, 0, 0}, b[3];
  int nd[3];
  W_R = P_RAD * YCO2 * OPL / T_RAD / 8.208e-2 * wco2;
  if (W_R <= 0.0) return;
  PE = pow(P_RAD + P_RAD * YCO2 * 0.3, 0.7);
  b[0] = 0.19e2; b[1] = 0.6157e-1; b[2] = 0.127e2;
  nd[0] = 0; nd[1] = 1; nd[2] = 0;
  aa[0] = A1R(v, ng, b, nd);
  PE = pow(P_RAD + P_RAD * YCO2 * 0.3, 0.8);
  b[0] = 0.36; b[1] = 0.4017e-1; b[2] = 0.134e2;
  nd[0] = -1; nd[1] = 0; nd[2] = 1;
  aa[1] = A1R(v, ng, b, nd);
  b[0] = 0.24; b[1] = 0.11888; b[2] = 0.101e2;
  aa[2] = A1R(v, ng, b, nd);
  b[0] = 0.11e3; b[1] = 0.24723; b[2] = 0.112e2;
  nd[0] = 0;
  aa[3] = A1R(v, ng, b, nd);
  PE = pow(P_RAD + P_RAD * YCO2 * 0.3, 0.65);
  b[0] = 4.0; b[1] = 0.13341; b[2] = 23.5;
  nd[0] = 1;
  aa[4] = A1R(v, ng, b, nd);
  b[0] = 0.66e-1; b[1] = 0.39305; b[2] = 0.345e2;
  nd[0] = 2;
  aa[5] = A1R(v, ng, b, nd);
  for (int n = 0; n < 6; n++) AACO2[n] = aa[n] * 1.0e2;
}
/* SOC, C/Rad_soot.f90:49-69: optical constants of soot (Lee and Tien 1981) */
static void SOC(double wl0, double *xn_k, double *x2nk) {
  static const double socf[3] = {7.2e26, 4.07e27, 4.47e28}, soch[3] = {1.2e15, 5.90e15, 5.60e15},
                      socw[3] = {0.0, 1.25e15, 7.25e15};
  const double c_0 = 2.9984e8, e2me0 = 3.185e3;
  *xn_k = 1.0;
  *x2nk = 0.0;
  double w = 2.0 * pi_r * c_0 / wl0;
  for (int j = 0; j < 3; j++) {
    double a = socw[j] * socw[j] - w * w;
    double b = a * a + w * w * soch[j] * soch[j];
    if (b < 1.0e-20) return;
    double g0 = socf[j] * e2me0;
    *xn_k += g0 * a / b;
    *x2nk += g0 * soch[j] * w / b;
  }
}
/* SOOTA, C/Rad_soot.f90:26-37: unchanged below a volume fraction of 1e-14 */
static void SOOTA(void) {
  if (FVS < 1.0e-14) return;
  for (int n = 1; n <= NWL; n++) {
    double wl0 = WL[n - 1], xn_k, x2nk;
    SOC(wl0, &xn_k, &x2nk);
    double g0 = 18.0 * pi_r * x2nk / ((xn_k + 2.0) * (xn_k + 2.0) + x2nk * x2nk);
    AKLS[n - 1] = g0 * FVS / wl0;
  }
}
/* RAD_GAS, C/Rad_gas.f90:16-141: AKL0 = AKLG + AKLS */
static void RAD_GAS(void) {
  memset(WNH2O, 0, sizeof(WNH2O));
  memset(WNCO2, 0, sizeof(WNCO2));
  memset(WNG, 0, sizeof(WNG));
  for (int n = 0; n < NWL; n++) AKLG[n] = 0.0;
  if (YH2O > 1.0e-4) {
    WNH2O[0][0] = 0.0;
    WNH2O[1][1] = 1.60e5; WNH2O[2][1] = 3.76e5; WNH2O[3][1] = 5.35e5; WNH2O[4][1] = 7.25e5;
    AH2O();
    WNH2O[0][1] = AAH2O[0] / 2;
    WNH2O[0][2] = AAH2O[0];
    for (int n = 2; n <= 5; n++) {
      double g0 = AAH2O[n - 1] / 2;
      WNH2O[n - 1][0] = WNH2O[n - 1][1] - g0;
      WNH2O[n - 1][2] = WNH2O[n - 1][1] + g0;
      if (WNH2O[n - 2][2] > WNH2O[n - 1][0]) WNH2O[n - 2][2] = WNH2O[n - 1][0];
    }
  }
  if (YCO2 > 1.0e-4) {
    WNCO2[0][1] = 6.67e4; WNCO2[1][1] = 9.60e4; WNCO2[2][1] = 1.06e5; WNCO2[3][2] = 2.41e5;
    WNCO2[4][1] = 3.66e5; WNCO2[5][1] = 5.20e5;
    ACO2();
    for (int n = 1; n <= 6; n++) {
      double g0 = AACO2[n - 1] / 2;
      if (n == 1 && g0 > WNCO2[0][1]) {
        WNCO2[0][0] = 0.0;
        WNCO2[0][2] = AACO2[0];
      } else if (n == 4) {
        WNCO2[3][1] = WNCO2[3][2] - g0;
        WNCO2[3][0] = WNCO2[3][1] - g0;
      } else {
        WNCO2[n - 1][0] = WNCO2[n - 1][1] - g0;
        WNCO2[n - 1][2] = WNCO2[n - 1][1] + g0;
      }
      if (n <= 1) continue;
      if (WNCO2[n - 2][2] > WNCO2[n - 1][0]) WNCO2[n - 2][2] = WNCO2[n - 1][0];
    }
  }
  SOOTA();
  for (int n = 0; n < 5; n++) { WNG[n][0] = WNH2O[n][0]; WNG[n][1] = WNH2O[n][2]; }
  for (int n = 0; n < 6; n++) { WNG[5 + n][0] = WNCO2[n][0]; WNG[5 + n][1] = WNCO2[n][2]; }
  for (int n = 1; n <= 10; n++) {
    for (int n2 = n + 1; n2 <= 11; n2++) {
      if (WNG[n - 1][0] < WNG[n2 - 1][0]) continue;
      double g0 = WNG[n - 1][0]; WNG[n - 1][0] = WNG[n2 - 1][0]; WNG[n2 - 1][0] = g0;
      g0 = WNG[n - 1][1]; WNG[n - 1][1] = WNG[n2 - 1][1]; WNG[n2 - 1][1] = g0;
    }
    if (n <= 1) continue;
    if (WNG[n - 2][1] > WNG[n - 1][0]) WNG[n - 2][1] = WNG[n - 1][0];
  }
  double wl2 = 10.0 / WL[0];
  for (int n = 0; n < NWL; n++) E0R[n] = 0.0;
  int n0 = 11;
  for (int n = 1; n <= NWL; n++) {
    double wl1 = (n == NWL) ? 0.0 : 1.0 / WL[n - 1];
    if (wl2 <= wl1 || WNG[10][1] <= wl1) {
      wl2 = wl1;
      continue;
    }
    for (int n1 = n0; n1 >= 1; n1--) {
      if (WNG[n1 - 1][1] <= wl1) {
        n0 = n1;
        break;
      }
      if (WNG[n1 - 1][1] >= wl1 && WNG[n1 - 1][1] <= wl2) {
        if (WNG[n1 - 1][0] < wl1) {
          E0R[n - 1] += WNG[n1 - 1][1] - wl1;
          WNG[n1 - 1][1] = wl1;
          n0 = n1;
          break;
        } else {
          E0R[n - 1] += WNG[n1 - 1][1] - WNG[n1 - 1][0];
        }
      }
    }
    if (E0R[n - 1] <= 0.0 || wl2 <= wl1) {
      AKLG[n - 1] = 0.0;
    } else {
      E0R[n - 1] = E0R[n - 1] / (wl2 - wl1);
      double g0 = 1.0 - E0R[n - 1];
      if (g0 <= 1.0e-20) g0 = 1.0e-20;
      AKLG[n - 1] = -log(g0) / OPL;
    }
    wl2 = wl1;
  }
  for (int n = 0; n < NWL; n++) AKL0[n] = AKLG[n] + AKLS[n];
}
/* RAD0, C/Rad0.f90:12-172 (T in K).  id_rad 10: AKL of every cell and the
 * emission sums; 1: QE of every open cell into QE_ (rad_qe_only). */
static void RAD0(void) {
  SGN = sig_sb * RC->ri * RC->ri;
  for (int l = 0; l < NWL; l++) DEB0[l] = hck / RC->ri / WL[l];
  RC->tot_qe_h2o_co2 = 0.0;
  RC->tot_qe_soot = 0.0;
  memset(AKL_, 0, sizeof(double) * CNN * NWL);
  for (int l = 0; l < NWL; l++) AKLG[l] = AKLS[l] = AKL0[l] = E0R[l] = DEB[l] = 0.0;
  double qeg0 = 0.0;
  for (int i = 2; i <= MP; i += 2) {
    int id2 = i / 2;
    for (int j = 2; j <= NP; j += 2) {
      int jd2 = j / 2;
      for (int k = 2; k <= LP; k += 2) {
        int kd2 = k / 2;
        T_RAD = T(i, j, k) * C->t0;
        if (IBR(i, j, k) <= 0 && T_RAD >= 300.0) {
          TT0 = sqrt(T_RAD / t0_r);
          P_RAD = (1.0 + P(i, j, k)) * pg0_r / p0_r;
          double sum = 0.0;
          for (int n = 1; n <= 7; n++) sum += GFM(i, j, k, n) / wms[n - 1];
          double wmix_ms = 1.0 / sum;
          YCO2 = GFM(i, j, k, 5) * wmix_ms / wco2;
          YH2O = GFM(i, j, k, 4) * wmix_ms / wh2o;
          FVS = SMF(i, j, k) * DNST(i, j, k) * C->dnst0 / RC->dnst_soot;
          for (int l = 1; l <= NWL; l++) AKL(id2, jd2, kd2, l) = 0.0;
          for (int l = 0; l < NWL; l++) AKL0[l] = 0.0;
          double vol = VOL_C(i, j, k);
          double sa = ((DX(i) * DR(j)) + (DR(j) * DZ(k)) + (DX(i) * DZ(k))) * 2.0;
          if (sa <= small16) continue;
          OPL = 6.0 * vol / sa;
          if (OPL <= 1.0e-10) continue;
          RAD_GAS();
          DEBX(T_RAD);
          double area_sgn_t4 = sa * SGN * ipow4(T_RAD);
          if (ID_RAD == 10) {
            double sg = 0.0, ss = 0.0;
            for (int l = 1; l <= NWL; l++) {
              AKL(id2, jd2, kd2, l) = oracle_emf(AKL0[l - 1], OPL);
              sg += DEB[l - 1] * oracle_emf(AKLG[l - 1], OPL);
              ss += DEB[l - 1] * oracle_emf(AKLS[l - 1], OPL);
            }
            RC->tot_qe_h2o_co2 += area_sgn_t4 * sg;
            RC->tot_qe_soot += area_sgn_t4 * ss;
          } else {
            double s = 0.0;
            for (int l = 1; l <= NWL; l++) s += DEB[l - 1] * oracle_emf(AKL0[l - 1], OPL);
            QE(id2, jd2, kd2) = area_sgn_t4 * s;
            qeg0 += QE(id2, jd2, kd2);
          }
        } else if (IBR(i, j, k) == 4) {
          for (int l = 1; l <= NWL; l++) AKL(id2, jd2, kd2, l) = RC->eps_m;
        } else if (IBR(i, j, k) > 100) {
          double e = (Z(k) >= RC->height_to_ceiling || (k == LP && RC->have_crown == 0)) ? RC->eps_c : RC->eps_w;
          for (int l = 1; l <= NWL; l++) AKL(id2, jd2, kd2, l) = e;
        }
      }
    }
  }
  if (ID_RAD != 10) RC->qeg0 = qeg0;
}

/* Cell (i,j,k) of slab position a on axis A and plane position (b, c) on
 * axes B, C. */
static void rad_cell(int A, int B, int C_, int a, int b, int c, int *i, int *j, int *k) {
  int v[4];
  v[A] = a; v[B] = b; v[C_] = c;
  *i = v[1]; *j = v[2]; *k = v[3];
}

/* One slab step of QRTX (A = x, B = y, C = z), QRTY (y, x, z) or QRTZ (z, x,
 * y): the energy QTX1 that left slab a10 toward slab a1 spreads across the
 * plane by the angles of slab a1's far face, first along B row by row
 * (rows of C from 4 to the last but one) into Q2M, then along C column by
 * column into QTX2.  mk1/mk2: the wall marks of faces normal to A (101/111,
 * 201/211, 301/311); split_gt: the wall split of the second pass needs a
 * cell code above it (100 QRTX, 99 QRTY, 1 QRTZ). */
static void QRT(int A, int B, int C_, int a0, int b0, int c0, int a10, int a1, int mk1, int mk2, int split_gt, int s) {
  const int Ma = axis_len(A), Mb = axis_len(B), Mc = axis_len(C_);
  int a2 = a1 + (a1 - a10) / 2;
  if (a2 > Ma) a2 = Ma;
  double q11[GFM_NWL_MAX], q12[GFM_NWL_MAX];
  /* the angles the far face subtends at the emitter (SANG2 expanded) */
  for (int bm = -1; bm <= 1; bm += 2) {
    int bend = bm < 0 ? 2 : Mb;
    for (int bb = b0 + bm; bm < 0 ? bb >= bend : bb <= bend; bb += 2 * bm) {
      double dist = fabs(axis_x(A, a2) - axis_x(A, a0)), ang;
      if (dist < small16) ang = 90.0;
      else ang = atand_(fabs(axis_x(B, bb) - axis_x(B, b0)) / dist);
      if (ang >= ANG0[B - 1]) break;
      ANG(B, bb, 2) = ang;
    }
  }
  for (int cm = -1; cm <= 1; cm += 2) {
    int cend = cm < 0 ? 2 : Mc;
    for (int cc = c0 + cm; cm < 0 ? cc >= cend : cc <= cend; cc += 2 * cm) {
      double dist = fabs(axis_x(A, a2) - axis_x(A, a0)), ang;
      if (dist < small16) ang = 90.0;
      else ang = atand_(fabs(axis_x(C_, cc) - axis_x(C_, c0)) / dist);
      if (ang >= ANG0[C_ - 1]) break;
      ANG(C_, cc, 2) = ang;
    }
  }
  /* first pass: along B, rows c1 = 4 .. Mc-2 */
  for (int c1 = 4; c1 <= Mc - 2; c1 += 2) {
    if (c1 < c0 && ANG(C_, c1 + 1, 2) >= ANG0[C_ - 1]) continue;
    if (c1 > c0 && ANG(C_, c1 - 1, 2) >= ANG0[C_ - 1]) break;
    int cd2 = c1 / 2;
    for (int bpm = -2; bpm <= 2; bpm += 4) {
      double a11 = 0.0, a12 = 0.0;
      int bend = bpm < 0 ? 2 : Mb;
      for (int b1 = b0; bpm < 0 ? b1 >= bend - bpm : b1 <= bend - bpm; b1 += bpm) {
        if (a11 >= ANG0[B - 1]) break;
        int b12 = b1 + bpm / 2;
        a12 = ANG(B, b12, 1);
        double dang = a12 - a11;
        if (dang <= 1.0e-10) { a11 = a12; continue; }
        double a21 = a11;
        int b1d2 = b1 / 2;
        if (IQT1(b1d2, cd2)) { a11 = a12; continue; }
        if (b1 == b0) {
          for (int l = 1; l <= NWL; l++) q11[l - 1] = QTX1(b1d2, cd2, l) * 0.5;
          if (bpm > 0) for (int l = 1; l <= NWL; l++) QTX1(b1d2, cd2, l) = 0.0;
        } else {
          for (int l = 1; l <= NWL; l++) { q11[l - 1] = QTX1(b1d2, cd2, l); QTX1(b1d2, cd2, l) = 0.0; }
        }
        if (b1 != b0 || bpm > 0) IQT1(b1d2, cd2) = 1;
        for (int l = 0; l < NWL; l++) q12[l] = q11[l];
        int b2;
        for (b2 = b1; bpm < 0 ? b2 >= bend : b2 <= bend; b2 += bpm) {
          int b2d2 = b2 / 2, ii, jj, kk;
          rad_cell(A, B, C_, a1, b2, c1, &ii, &jj, &kk);
          if (IBR(ii, jj, kk) < 1) {
            int b2m = b2 + bpm / 2;
            if (b2m < 1 || b2m > Mb) { RAD_OOB++; b2m = b2m < 1 ? 1 : Mb; }
            double a22 = ANG(B, b2m, 2);
            if (a22 <= a21) continue;
            if (a22 < a12) {
              double fij = (a22 - a21) / dang;
              for (int l = 1; l <= NWL; l++) {
                if (q11[l - 1] < small16) continue;
                double q21 = q11[l - 1] * fij;
                Q2M(b2d2, cd2, l) += q21;
                q12[l - 1] -= q21;
              }
              IQTM(b2d2, cd2) = 0;
              a21 = a22;
              continue;
            }
          }
          for (int l = 1; l <= NWL; l++) Q2M(b2d2, cd2, l) += q12[l - 1];
          IQTM(b2d2, cd2) = 0;
          break;
        }
        RAD_STALE[s] = b2;
        RAD_STALE_SET[s] = 1;
        a11 = a12;
      }
    }
  }
  /* second pass: along C, columns b1 = 2 .. Mb */
  for (int b1 = 2; b1 <= Mb; b1 += 2) {
    if (b1 < b0 && ANG(B, b1 + 1, 2) >= ANG0[B - 1]) continue;
    if (b1 > b0 && ANG(B, b1 - 1, 2) >= ANG0[B - 1]) break;
    int bd2 = b1 / 2;
    for (int cm = -2; cm <= 2; cm += 4) {
      double a11 = 0.0, a12 = 0.0;
      int cend = cm < 0 ? 2 : Mc;
      for (int c1 = c0; cm < 0 ? c1 >= cend - cm : c1 <= cend - cm; c1 += cm) {
        if (a11 >= ANG0[C_ - 1]) break;
        int c12 = c1 + cm / 2;
        a12 = ANG(C_, c12, 1);
        double dang = a12 - a11;
        if (dang <= 1.0e-10) { a11 = a12; continue; }
        double a21 = a11;
        int cd2 = c1 / 2;
        if (IQTM(bd2, cd2)) { a11 = a12; continue; }
        if (c1 == c0) {
          for (int l = 1; l <= NWL; l++) q11[l - 1] = Q2M(bd2, cd2, l) * 0.5;
          if (cm > 0) for (int l = 1; l <= NWL; l++) Q2M(bd2, cd2, l) = 0.0;
        } else {
          for (int l = 1; l <= NWL; l++) { q11[l - 1] = Q2M(bd2, cd2, l); Q2M(bd2, cd2, l) = 0.0; }
        }
        if (c1 != c0 || cm > 0) IQTM(bd2, cd2) = 1;
        int ii, jj, kk;
        rad_cell(A, B, C_, a1, b1, c1, &ii, &jj, &kk);
        const int ibc1 = IBR(ii, jj, kk);
        if (ibc1 == mk1 || ibc1 == mk2 || ibc1 == 4) {
          for (int l = 1; l <= NWL; l++) QTX2(bd2, cd2, l) += q11[l - 1];
          IQT2(bd2, cd2) = 0;
          a11 = a12;
          continue;
        }
        for (int l = 0; l < NWL; l++) q12[l] = q11[l];
        for (int c2 = c1; cm < 0 ? c2 >= cend : c2 <= cend; c2 += cm) {
          int c2m = c2 + cm / 2, c2d2 = c2 / 2;
          rad_cell(A, B, C_, a1, b1, c2, &ii, &jj, &kk);
          const int ibc2 = IBR(ii, jj, kk);
          int deposit = 0;
          if (ibc1 <= 0 && ibc2 >= 1) {
            deposit = 1;
          } else if (ibc1 > split_gt && ibc2 == 1) {
            int b2;
            if (b1 < b0) b2 = b1 + 2;
            else if (b1 > b0) b2 = b1 - 2;
            else {
              b2 = RAD_STALE[s];
              RAD_STALE_USES++;
              if (!RAD_STALE_SET[s]) RAD_STALE_PREV++;
            }
            RAD_STALE[s] = b2;  /* the split assigns the reference's variable too */
            if (b2 < 2 || b2 > Mb) { RAD_OOB++; b2 = b2 < 2 ? 2 : Mb; }  /* a completed first-pass loop leaves it past the end */
            int b2d2 = b2 / 2, c2d2m = c2d2 - cm / 2;
            for (int l = 1; l <= NWL; l++) {
              if (q12[l - 1] < small16) continue;
              double q21 = q12[l - 1] * 0.5;
              QTX2(b2d2, c2d2, l) += q21;
              QTX2(bd2, c2d2m, l) += q21;
            }
            IQT2(b2d2, c2d2) = 0;
            IQT2(bd2, c2d2m) = 0;
            break;
          } else if (c2m > Mc || c2m < 2) {
            deposit = 1;
          } else {
            double a22 = ANG(C_, c2m, 2);
            if ((a22 - a21) <= 1.0e-10) continue;
            if (a22 < a12) {
              double fij = (a22 - a21) / dang;
              for (int l = 1; l <= NWL; l++) {
                if (q11[l - 1] < small16) continue;
                double q21 = q11[l - 1] * fij;
                QTX2(bd2, c2d2, l) += q21;
                q12[l - 1] -= q21;
              }
              IQT2(bd2, c2d2) = 0;
              a21 = a22;
              continue;
            }
            deposit = 1;
          }
          if (deposit) {
            for (int l = 1; l <= NWL; l++) QTX2(bd2, c2d2, l) += q12[l - 1];
            IQT2(bd2, c2d2) = 0;
            break;
          }
        }
        a11 = a12;
      }
    }
  }
}

/* One axis of QREF (C/Qre.f90:95-505): both directions along A from the
 * emitter, slab by slab (the QRT step, then absorption in slab a: open
 * cells take QTX2*AKL and pass the rest on in QTX1, boundary cells and the
 * last slab take all). */
static void QREF_AXIS(int A, int B, int C_, int i0, int j0, int k0, int mk1, int mk2, int split_gt, int s) {
  int v0[4] = {0, i0, j0, k0};
  const int a0 = v0[A], b0 = v0[B], c0 = v0[C_];
  const int Ma = axis_len(A), Mb = axis_len(B), Mc = axis_len(C_);
  const int i0d2 = i0 / 2, j0d2 = j0 / 2, k0d2 = k0 / 2;
  PLB = Mb / 2; PLC = Mc / 2;
  const size_t pn = (size_t)PLB * PLC;
  memset(QTX1_, 0, sizeof(double) * pn * NWL);
  memset(Q2M_, 0, sizeof(double) * pn * NWL);
  memset(QTX2_, 0, sizeof(double) * pn * NWL);
  memset(IQT1_, 1, pn); memset(IQTM_, 1, pn); memset(IQT2_, 1, pn);
  const double (*dd[4]) = {NULL, DX_, DR_, DZ_};
  for (int l = 1; l <= NWL; l++) QEL1[l - 1] = QEL0[l - 1] * dd[B][b0 - 1] * dd[C_][c0 - 1];
  for (int apm = -2; apm <= 2; apm += 4) {
    for (int l = 1; l <= NWL; l++) QTX1(b0 / 2, c0 / 2, l) = QEL1[l - 1];
    IQT1(b0 / 2, c0 / 2) = 0;
    for (int l = 1; l <= NWL; l++) QEL(i0d2, j0d2, k0d2, l) += QEL1[l - 1];
    /* sfac of the emitter's own corner */
    int v1[4] = {0, i0 + apm / 2, j0 + apm / 2, k0 + apm / 2};
    ANG0[A - 1] = 0.0;
    double bb = fabs(axis_x(A, v1[A]) - axis_x(A, a0));
    if (bb < small16) {
      ANG0[B - 1] = 90.0;
      ANG0[C_ - 1] = 90.0;
    } else {
      ANG0[B - 1] = atand_(fabs(axis_x(B, v1[B]) - axis_x(B, b0)) / bb);
      ANG0[C_ - 1] = atand_(fabs(axis_x(C_, v1[C_]) - axis_x(C_, c0)) / bb);
    }
    for (int n = 1; n <= Mb; n++) ANG(B, n, 1) = ANG(B, n, 2) = ANG0[B - 1];
    ANG(B, b0, 1) = ANG(B, b0, 2) = 0.0;
    for (int n = 1; n <= Mc; n++) ANG(C_, n, 1) = ANG(C_, n, 2) = ANG0[C_ - 1];
    ANG(C_, c0, 1) = ANG(C_, c0, 2) = 0.0;
    const int aend = apm < 0 ? 2 : Ma;
    for (int a = a0 + apm; apm < 0 ? a >= aend : a <= aend; a += apm) {
      QRT(A, B, C_, a0, b0, c0, a - apm, a, mk1, mk2, split_gt, s);
      for (int b = 2; b <= Mb; b += 2) {
        int bd2 = b / 2;
        for (int c = 2; c <= Mc; c += 2) {
          int cd2 = c / 2;
          if (IQT2(bd2, cd2)) continue;
          int i, j, k;
          rad_cell(A, B, C_, a, b, c, &i, &j, &k);
          for (int l = 1; l <= NWL; l++) {
            if (QTX2(bd2, cd2, l) < small16) continue;
            const int ib = IBR(i, j, k);
            if (ib >= 1 || a == aend) {
              QAL(i / 2, j / 2, k / 2, l) += QTX2(bd2, cd2, l);
              qa_sum_bnd += QTX2(bd2, cd2, l);
              if (ib == 4) qa_to_melt_c += QTX2(bd2, cd2, l);
              else if (ib == 3) qa_to_exit_c += QTX2(bd2, cd2, l);
              else if (ib == 2) qa_to_inlet_c += QTX2(bd2, cd2, l);
              else qa_to_wall_c += QTX2(bd2, cd2, l);
            } else {
              double qa0 = QTX2(bd2, cd2, l) * AKL(i / 2, j / 2, k / 2, l);
              QAL(i / 2, j / 2, k / 2, l) += qa0;
              QTX1(bd2, cd2, l) = QTX2(bd2, cd2, l) - qa0;
              IQT1(bd2, cd2) = 0;
              qa_sum_vol += qa0;
              qa_to_open_c += qa0;
            }
            QTX2(bd2, cd2, l) = 0.0;
            IQT2(bd2, cd2) = 1;
          }
        }
      }
      for (int n = 1; n <= Mb; n++) ANG(B, n, 1) = ANG(B, n, 2);
      for (int n = 1; n <= Mc; n++) ANG(C_, n, 1) = ANG(C_, n, 2);
    }
  }
}

/* QREF, C/Qre.f90:31-505, for an open emitter (i0,j0,k0) with DEB set */
static void QREF(int i0, int j0, int k0) {
  const int i0d2 = i0 / 2, j0d2 = j0 / 2, k0d2 = k0 / 2;
  qa_sum_bnd = qa_sum_vol = 0.0;
  qa_to_open_c = qa_to_melt_c = qa_to_inlet_c = qa_to_exit_c = qa_to_wall_c = 0.0;
  for (int l = 1; l <= NWL; l++) QEL(i0d2, j0d2, k0d2, l) = 0.0;
  if (IBR(i0, j0, k0) >= 1) return;
  double tk = T(i0, j0, k0) * C->t0;
  double g0 = SGN * ipow4(tk);
  for (int l = 1; l <= NWL; l++) QEL0[l - 1] = DEB[l - 1] * AKL(i0d2, j0d2, k0d2, l) * g0;
  for (int s = 0; s < 3; s++) RAD_STALE_SET[s] = 0;
  QREF_AXIS(1, 2, 3, i0, j0, k0, 101, 111, 100, 0);
  QREF_AXIS(2, 1, 3, i0, j0, k0, 201, 211, 99, 1);
  QREF_AXIS(3, 1, 2, i0, j0, k0, 301, 311, 1, 2);
  double qe = QEL(i0d2, j0d2, k0d2, 1);
  for (int l = 2; l <= NWL; l++) qe += QEL(i0d2, j0d2, k0d2, l);
  QE(i0d2, j0d2, k0d2) = qe;
  RC->qe_sum_from_open += qe;
}

/* the wall marks of RAD_EMIS, C/Rad_emis.f90:155-186, on the copy IBR_ */
static void rad_mark_walls(void) {
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        if (IBR(i, j, k) != 1) continue;
        int km2 = imax(k - 2, 2), kp2 = imin(k + 2, LP), jm2 = imax(j - 2, 2), jp2 = imin(j + 2, NP);
        int im2 = imax(i - 2, 2), ip2 = imin(i + 2, MP);
        int16_t *c = &IBR_[IX3(i, j, k)];
        if (IBR(i, j, km2) <= 0) *c = 301;
        else if (IBR(i, j, kp2) <= 0) *c = 311;
        else if (IBR(i, jm2, k) <= 0) *c = 201;
        else if (IBR(i, jp2, k) <= 0) *c = 211;
        else if (IBR(im2, j, k) <= 0) *c = 101;
        else if (IBR(ip2, j, k) <= 0) *c = 111;
      }
}

static int rad_bind(gfm_ms_ctrl *ms, const gfm_ms_arrays *a, gfm_rad_ctrl *rc, const gfm_rad_arrays *ra) {
  static gfm_ms_ctrl msd;
  static double *dummy = NULL;
  static size_t cap = 0, capn = 0;
  if (!rc || !ra || !ra->qe || !ra->qa || !ra->qal || !ra->akl || !a || !a->gfm || !a->smf) return GFM_E_ARG;
  if (rc->nwl < 1 || rc->nwl > GFM_NWL_MAX || rc->ri <= 0.0 || rc->dnst_soot <= 0.0) return GFM_E_ARG;
  for (int l = 0; l < rc->nwl; l++)
    if (!(rc->wl[l] > 0.0)) return GFM_E_ARG;
  if (!ms) ms = &msd;
  gfm_ms_arrays aa = *a;
  if (!aa.bs_soot_prev || !aa.sfp_soot_prev || !aa.rrms) {
    if (!dummy || cap != NN) { free(dummy); dummy = calloc(NN, sizeof(double)); cap = NN; }
    if (!dummy) return GFM_E_DEVICE;
    if (!aa.bs_soot_prev) aa.bs_soot_prev = dummy;
    if (!aa.sfp_soot_prev) aa.sfp_soot_prev = dummy;
    if (!aa.rrms) aa.rrms = dummy;
  }
  int r = ms_bind(ms, &aa);
  if (r) return r;
  RC = rc;
  NWL = rc->nwl;
  for (int l = 0; l < NWL; l++) WL[l] = rc->wl[l];
  AKL_ = ra->akl; QAL_ = ra->qal; QE_ = ra->qe; QA_ = ra->qa;
  if (capn != NN) {
    free(IBR_); free(QEL_); free(ANGA_[0]); free(ANGA_[1]); free(ANGA_[2]);
    free(QTX1_); free(Q2M_); free(QTX2_); free(IQT1_); free(IQTM_); free(IQT2_);
    size_t pmax = (size_t)(imax(imax(MP * NP, MP * LP), NP * LP) / 4 + 1);
    IBR_ = malloc(sizeof(int16_t) * NN);
    QEL_ = calloc(CNN * GFM_NWL_MAX, sizeof(double));
    ANGA_[0] = calloc(2 * (size_t)MP, sizeof(double)); ANGA_[1] = calloc(2 * (size_t)NP, sizeof(double));
    ANGA_[2] = calloc(2 * (size_t)LP, sizeof(double));
    QTX1_ = calloc(pmax * GFM_NWL_MAX, sizeof(double)); Q2M_ = calloc(pmax * GFM_NWL_MAX, sizeof(double));
    QTX2_ = calloc(pmax * GFM_NWL_MAX, sizeof(double));
    IQT1_ = malloc(pmax); IQTM_ = malloc(pmax); IQT2_ = malloc(pmax);
    if (!IBR_ || !QEL_ || !ANGA_[0] || !ANGA_[1] || !ANGA_[2] || !QTX1_ || !Q2M_ || !QTX2_ || !IQT1_ || !IQTM_ || !IQT2_)
      return GFM_E_DEVICE;
    capn = NN;
  }
  memcpy(IBR_, IBC, sizeof(int16_t) * NN);
  return GFM_OK;
}

/* RAD_EMIS up to QRN1: wall marks, RAD0 (ID_RAD 10), QRNF's volume loop. */
int oracle_rad_emis(gfm_ctrl *c, const gfm_host_arrays *h, const gfm_ms_arrays *a, gfm_rad_ctrl *rc,
                    const gfm_rad_arrays *ra) {
  int r = bind(c, h);
  if (!r) r = rad_bind(NULL, a, rc, ra);
  if (r) return r;
  rad_mark_walls();
  ID_RAD = 10;
  RAD0();
  memset(QE_, 0, sizeof(double) * CNN);
  memset(QEL_, 0, sizeof(double) * CNN * NWL);
  memset(QAL_, 0, sizeof(double) * CNN * NWL);
  memset(QA_, 0, sizeof(double) * CNN);
  for (int d = 0; d < 3; d++) memset(ANGA_[d], 0, sizeof(double) * 2 * axis_len(d + 1));
  rc->qe_sum_from_open = 0.0;
  rc->qa_sum_to_melt = rc->qa_sum_to_exit = rc->qa_sum_to_inlet = rc->qa_sum_to_wall = rc->qa_sum_to_open = 0.0;
  double qeg0 = 0.0;
  int l_ab = 0, n_emit = 0;
  const long long prev0 = RAD_STALE_PREV, oob0 = RAD_OOB;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        const int ibc0 = IBR(i, j, k);
        if (ibc0 < 1) l_ab++;
        if (ibc0 >= 1) continue;
        if (T(i, j, k) * C->t0 <= 300.0) continue;
        DEBX(T(i, j, k) * C->t0);
        QREF(i, j, k);
        n_emit++;
        rc->qa_sum_to_melt += qa_to_melt_c;
        rc->qa_sum_to_exit += qa_to_exit_c;
        rc->qa_sum_to_inlet += qa_to_inlet_c;
        rc->qa_sum_to_wall += qa_to_wall_c;
        rc->qa_sum_to_open += qa_to_open_c;
        if (ibc0 <= 0) qeg0 += QE(i / 2, j / 2, k / 2);
      }
  rc->l_ab = l_ab;
  rc->n_emit = n_emit;
  rc->carry_unset = (int32_t)(RAD_STALE_PREV - prev0);
  rc->angle_oob = (int32_t)(RAD_OOB - oob0);
  rc->qeg0 = qeg0;
  /* QRN1, C/Qrn.f90:245-290 */
  double qg = 0.0, qio = 0.0, qs = 0.0, qw = 0.0;
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int k = 2; k <= LP; k += 2) {
        const int ibc0 = IBR(i, j, k);
        double g0 = 0.0;
        for (int l = 1; l <= NWL; l++) g0 += QAL(i / 2, j / 2, k / 2, l);
        QA(i / 2, j / 2, k / 2) = g0;
        if (ibc0 <= 0) qg += g0;
        else if (ibc0 == 2 || ibc0 == 3) qio += g0;
        else if (ibc0 == 4) qs += g0;
        else qw += g0;
      }
  rc->qeg_g = qg; rc->qeg_io = qio; rc->qeg_s = qs; rc->qeg_w = qw;
  return GFM_OK;
}

/* rad_qe_only, C/Rad_emis.f90:411-466: QE of every open cell (RAD0, ID_RAD 1);
 * QE = 0 elsewhere (qe = qe2). */
int oracle_rad_qe_only(gfm_ctrl *c, const gfm_host_arrays *h, const gfm_ms_arrays *a, gfm_rad_ctrl *rc,
                       const gfm_rad_arrays *ra) {
  int r = bind(c, h);
  if (!r) r = rad_bind(NULL, a, rc, ra);
  if (r) return r;
  ID_RAD = 1;
  memset(QE_, 0, sizeof(double) * CNN);
  RAD0();
  return GFM_OK;
}
/* single functions for the pins: 1 DEBX(x[0]) -> out[0..nwl-1]; 2 RIF(3, 0, x[0]); 3 the band model of one
 * cell state x = (T K, P atm - 1, wmix-free YH2O, YCO2, FVS, OPL) -> out = AKLG, AKLS (nwl each) */
int oracle_rad_probe(gfm_rad_ctrl *rc, int what, const double *x, double *out) {
  if (!rc || rc->nwl < 1 || rc->nwl > GFM_NWL_MAX) return GFM_E_ARG;
  RC = rc;
  NWL = rc->nwl;
  for (int l = 0; l < NWL; l++) WL[l] = rc->wl[l];
  SGN = sig_sb * rc->ri * rc->ri;
  for (int l = 0; l < NWL; l++) DEB0[l] = hck / rc->ri / WL[l];
  switch (what) {
    case 1: DEBX(x[0]); for (int l = 0; l < NWL; l++) out[l] = DEB[l]; break;
    case 2: out[0] = RIF(3, 0.0, x[0]); break;
    case 3:
      for (int l = 0; l < NWL; l++) AKLS[l] = 0.0;
      T_RAD = x[0]; TT0 = sqrt(T_RAD / t0_r); P_RAD = 1.0 + x[1];
      YH2O = x[2]; YCO2 = x[3]; FVS = x[4]; OPL = x[5];
      RAD_GAS();
      for (int l = 0; l < NWL; l++) { out[l] = AKLG[l]; out[NWL + l] = AKLS[l]; }
      break;
    default: return GFM_E_ARG;
  }
  return GFM_OK;
}
size_t oracle_rad_ctrl_size(void) { return sizeof(gfm_rad_ctrl); }
/* uses of the carried wall-split column (total, whose value came from an earlier call), out-of-range angle reads */
void oracle_rad_stats(long long *out) { out[0] = RAD_STALE_USES; out[1] = RAD_STALE_PREV; out[2] = RAD_OOB; }

/* ======================================================================
 * Wall radiosity exchange, C/Qwall.f90: prepare_for_wall_radiation_
 * calculation (:54-566, without the debug prints and the wall-grid file of
 * iwalvis), wall_q_amb (:1585-1659), wallq (:990-1369) with
 * calc_q_conv_wall (:1370-1478) and get_radiation_from_boundaries
 * (:1489-1583, preset view factors); the first-pass patch temperatures of
 * RAD_EMIS (C/Rad_emis.f90:190-212, later passes :253-261), QRN_PRN's
 * QA_S0/QA_S1 (C/Qrn.f90:399-402) and the state the next gas iterations
 * see (C/Rad_emis.f90:337-355 and rad_emis_rr :101-113): T at every wall
 * patch face, and at each wall cell the last wall patch's b_T with its
 * enthalpy.  QRN2 only writes the GUI file (its sums of the removed wall
 * absorption are zero) and is not restated; nor are the GUI poll and the
 * log writes.
 *
 * Reference state kept: the corner variables of the exit search
 * (kend, jend, iend, inew, ...) are not reset between exits, so an exit
 * whose extent or interior end is not found reuses the last one found.
 * ====================================================================== */
static const double sig_w = 5.668e-8, t_mnk_w = 2.0e2, pi_w = 3.141592653589793;
static int16_t *MAPB_ = NULL;
#define MAPB(i, j, k) MAPB_[IX3(i, j, k)]
/* mapb out of the lattice reads as "not open" (the reference never reads there for exits away from edges) */
static int mapb_at(int i, int j, int k) {
  if (i < 1 || i > MP || j < 1 || j > NP || k < 1 || k > LP) return 1;
  return MAPB(i, j, k);
}

/* a write inside the lattice (an exit whose interior end was never found leaves a stale or zero index) */
static void mapb_set(int i, int j, int k, int v) {
  if (i >= 1 && i <= MP && j >= 1 && j <= NP && k >= 1 && k <= LP) MAPB(i, j, k) = (int16_t)v;
}

/* Steps 1-2: MAPB = IBCELL with each exit tunnel walled up to an exit at its interior end, C/Qwall.f90:74-296 */
static void wall_mapb(void) {
  for (size_t q = 0; q < NN; q++) MAPB_[q] = IBC[q];
  int kstart = 0, kend = 0, jstart = 0, jend = 0, istart = 0, iend = 0, inew = 0, jnew = 0, knew = 0;
  /* back wall, i = 2 */
  for (int j = 2; j <= NP; j += 2)
    for (int k = 2; k <= LP; k += 2) {
      if (MAPB(2, j, k) != 3) continue;
      kstart = k;
      for (int kk = k + 2; kk <= LP; kk += 2) if (MAPB(2, j, kk) != 3) { kend = kk - 2; break; }
      jstart = j;
      for (int jj = j + 2; jj <= NP; jj += 2) if (MAPB(2, jj, k) != 3) { jend = jj - 2; break; }
      for (int ii = 4; ii <= MP; ii += 2)
        if (mapb_at(ii, jstart - 2, kstart - 2) == 0 || mapb_at(ii, jend + 2, kend + 2) == 0 ||
            mapb_at(ii, jstart - 2, kend + 2) == 0 || mapb_at(ii, jend + 2, kstart - 2) == 0) { inew = ii - 2; break; }
      for (int ii = inew; ii <= inew + 1; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 3);
      for (int ii = 2; ii <= inew - 1; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 1);
    }
  /* front wall, i = MP */
  for (int j = 2; j <= NP; j += 2)
    for (int k = 2; k <= LP; k += 2) {
      if (MAPB(MP, j, k) != 3) continue;
      kstart = k;
      for (int kk = k + 2; kk <= LP; kk += 2) if (MAPB(MP, j, kk) != 3) { kend = kk - 2; break; }
      jstart = j;
      for (int jj = j + 2; jj <= NP; jj += 2) if (MAPB(MP, jj, k) != 3) { jend = jj - 2; break; }
      for (int ii = MP - 2; ii >= 2; ii -= 2)
        if (mapb_at(ii, jstart - 2, kstart - 2) == 0 || mapb_at(ii, jend + 2, kend + 2) == 0 ||
            mapb_at(ii, jstart - 2, kend + 2) == 0 || mapb_at(ii, jend + 2, kstart - 2) == 0) { inew = ii + 2; break; }
      for (int ii = inew - 1; ii <= inew; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 3);
      for (int ii = inew + 1; ii <= MP; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 1);
    }
  /* left wall, j = 2 */
  for (int i = 2; i <= MP; i += 2)
    for (int k = 2; k <= LP; k += 2) {
      if (MAPB(i, 2, k) != 3) continue;
      kstart = k;
      for (int kk = k + 2; kk <= LP; kk += 2) if (MAPB(i, 2, kk) != 3) { kend = kk - 2; break; }
      istart = i;
      for (int ii = i + 2; ii <= MP; ii += 2) if (MAPB(ii, 2, k) != 3) { iend = ii - 2; break; }
      for (int jj = 4; jj <= NP; jj += 2)
        if (mapb_at(istart - 2, jj, kstart - 2) == 0 || mapb_at(iend + 2, jj, kend + 2) == 0 ||
            mapb_at(istart - 2, jj, kend + 2) == 0 || mapb_at(iend + 2, jj, kstart - 2) == 0) { jnew = jj - 2; break; }
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = jnew; jj <= jnew + 1; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 3);
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = 2; jj <= jnew - 1; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 1);
    }
  /* right wall, j = NP */
  for (int i = 2; i <= MP; i += 2)
    for (int k = 2; k <= LP; k += 2) {
      if (MAPB(i, NP, k) != 3) continue;
      kstart = k;
      for (int kk = k + 2; kk <= LP; kk += 2) if (MAPB(i, NP, kk) != 3) { kend = kk - 2; break; }
      istart = i;
      for (int ii = i + 2; ii <= MP; ii += 2) if (MAPB(ii, NP, k) != 3) { iend = ii - 2; break; }
      for (int jj = NP - 2; jj >= 2; jj -= 2)
        if (mapb_at(istart - 2, jj, kstart - 2) == 0 || mapb_at(iend + 2, jj, kend + 2) == 0 ||
            mapb_at(istart - 2, jj, kend + 2) == 0 || mapb_at(iend + 2, jj, kstart - 2) == 0) { jnew = jj + 2; break; }
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = jnew - 1; jj <= jnew; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 3);
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = jnew + 1; jj <= NP; jj++)
          for (int kk = kstart; kk <= kend; kk++) mapb_set(ii, jj, kk, 1);
    }
  /* top wall, k = LP */
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2) {
      if (MAPB(i, j, LP) != 3) continue;
      istart = i;
      for (int ii = i + 2; ii <= MP; ii += 2) if (MAPB(ii, j, LP) != 3) { iend = ii - 2; break; }
      jstart = j;
      for (int jj = j + 2; jj <= NP; jj += 2) if (MAPB(i, jj, LP) != 3) { jend = jj - 2; break; }
      for (int kk = LP - 2; kk >= 2; kk -= 2)
        if (mapb_at(istart - 2, jstart - 2, kk) == 0 || mapb_at(iend + 2, jend + 2, kk) == 0 ||
            mapb_at(istart - 2, jend + 2, kk) == 0 || mapb_at(iend + 2, jstart - 2, kk) == 0) { knew = kk + 2; break; }
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = knew - 1; kk <= knew; kk++) mapb_set(ii, jj, kk, 3);
      for (int ii = istart; ii <= iend; ii++)
        for (int jj = jstart; jj <= jend; jj++)
          for (int kk = knew + 1; kk <= LP; kk++) mapb_set(ii, jj, kk, 1);
    }
}

/* emissivity of a patch, C/Qwall.f90:406-416, 443-453, 480-492 (z-normal: the crown test at LP-1) */
static double wall_emis(const gfm_wall_ctrl *wc, int code, int k, int ktop) {
  if (code == 1) {
    if (Z(k) >= wc->height_to_ceiling || (k == ktop && wc->have_crown == 0)) return wc->eps_c;
    return wc->eps_w;
  }
  if (code == 4 && ktop == LP - 1) return wc->eps_m;
  return 1.0;
}

/* Steps 3-8 of prepare_for_wall_radiation_calculation.  With wa->b_typ null
 * only counts (wc->b_cnt). */
int oracle_wall_setup(gfm_ctrl *c, const gfm_host_arrays *h, gfm_wall_ctrl *wc, gfm_wall_arrays *wa,
                      const int16_t *iwall) {
  (void)iwall;
  int r = bind(c, h);
  if (r) return r;
  if (!wc || !wa) return GFM_E_ARG;
  free(MAPB_);
  MAPB_ = malloc(sizeof(int16_t) * NN);
  if (!MAPB_) return GFM_E_DEVICE;
  wall_mapb();
  int n = 0;
  const int fill = wa->b_typ != NULL;
  if (fill && wc->b_cnt <= 0) return GFM_E_ARG;
  /* x-normal faces */
  for (int k = 2; k <= LP; k += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int i = 3; i <= MP - 1; i += 2) {
        if (MAPB(i, j, k) < 1) continue;
        if (MAPB(i - 1, j, k) > 0 && MAPB(i + 1, j, k) > 0) continue;
        if (fill) {
          wa->b_typ[n] = MAPB(i, j, k);
          wa->emis[n] = wall_emis(wc, MAPB(i, j, k), k, LP);
          wa->b_orient[n] = 1;
          if (MAPB(i + 1, j, k) == 0) { wa->b_direct[n] = 0; wa->bc_i[n] = i - 1; }
          else { wa->b_direct[n] = 1; wa->bc_i[n] = i + 1; }
          wa->bc_j[n] = j; wa->bc_k[n] = k;
          wa->b_i[n] = i; wa->b_j[n] = j; wa->b_k[n] = k;
          wa->patch_area[n] = DR(j) * DZ(k);
        }
        n++;
      }
  /* y-normal faces */
  for (int k = 2; k <= LP; k += 2)
    for (int j = 3; j <= NP - 1; j += 2)
      for (int i = 2; i <= MP; i += 2) {
        if (MAPB(i, j, k) < 1) continue;
        if (MAPB(i, j - 1, k) > 0 && MAPB(i, j + 1, k) > 0) continue;
        if (fill) {
          wa->b_typ[n] = MAPB(i, j, k);
          wa->emis[n] = wall_emis(wc, MAPB(i, j, k), k, LP);
          wa->b_orient[n] = 2;
          if (MAPB(i, j + 1, k) == 0) { wa->b_direct[n] = 0; wa->bc_j[n] = j - 1; }
          else { wa->b_direct[n] = 1; wa->bc_j[n] = j + 1; }
          wa->bc_i[n] = i; wa->bc_k[n] = k;
          wa->b_i[n] = i; wa->b_j[n] = j; wa->b_k[n] = k;
          wa->patch_area[n] = DX(i) * DZ(k);
        }
        n++;
      }
  /* z-normal faces */
  for (int k = 3; k <= LP - 1; k += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int i = 2; i <= MP; i += 2) {
        if (MAPB(i, j, k) < 1) continue;
        if (MAPB(i, j, k - 1) > 0 && MAPB(i, j, k + 1) > 0) continue;
        if (fill) {
          wa->b_typ[n] = MAPB(i, j, k);
          wa->emis[n] = wall_emis(wc, MAPB(i, j, k), k, LP - 1);
          wa->b_orient[n] = 3;
          if (MAPB(i, j, k + 1) == 0) { wa->b_direct[n] = 0; wa->bc_k[n] = k - 1; }
          else { wa->b_direct[n] = 1; wa->bc_k[n] = k + 1; }
          wa->bc_i[n] = i; wa->bc_j[n] = j;
          wa->b_i[n] = i; wa->b_j[n] = j; wa->b_k[n] = k;
          wa->patch_area[n] = DX(i) * DR(j);
        }
        n++;
      }
  if (!fill) { wc->b_cnt = n; return GFM_OK; }
  if (n != wc->b_cnt) return GFM_E_ARG;
  const int nb = n;
  /* Step 6: exhausts */
  for (int p = 0; p < nb; p++) {
    wa->b_exh[p] = 0;
    if (wa->b_typ[p] != 3) continue;
    for (int m = 0; m < wc->nex; m++) {
      if (wa->b_orient[p] != wc->exh_orient[m] || wa->b_direct[p] != wc->exh_direct[m]) continue;
      const int *e = wc->exh_box[m];
      int hit = 0;
      if (wa->b_orient[p] == 1) hit = wa->bc_j[p] >= e[1] && wa->bc_j[p] <= e[4] && wa->bc_k[p] >= e[2] && wa->bc_k[p] <= e[5];
      else if (wa->b_orient[p] == 2) hit = wa->bc_i[p] >= e[0] && wa->bc_i[p] <= e[3] && wa->bc_k[p] >= e[2] && wa->bc_k[p] <= e[5];
      else hit = wa->bc_i[p] >= e[0] && wa->bc_i[p] <= e[3] && wa->bc_j[p] >= e[1] && wa->bc_j[p] <= e[4];
      if (hit) { wa->b_exh[p] = (int16_t)(m + 1); break; }
    }
  }
  /* Step 7: view factors */
  if (wc->preset_vf == 1) {
    if (!wa->vf || !wa->vf_sum) return GFM_E_ARG;
    for (int p = 0; p < nb; p++) {
      double s = 0.0;
      for (int m = 0; m < nb; m++) {
        double v = 0.0;
        if (m != p) {
          const double ddx = X(wa->b_i[p]) - X(wa->b_i[m]), ddy = R(wa->b_j[p]) - R(wa->b_j[m]),
                       ddz = Z(wa->b_k[p]) - Z(wa->b_k[m]);
          const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
          const double den = pi_w * (d2 * d2);
          const double ex = X(wa->b_i[m]) - X(wa->b_i[p]), ey = R(wa->b_j[m]) - R(wa->b_j[p]), ez = Z(wa->b_k[m]) - Z(wa->b_k[p]);
          const int om = wa->b_orient[m], dp = wa->b_direct[p];
          int vis = 1;
          double num = 0.0, e0 = 0.0;
          if (wa->b_orient[p] == 1) {
            if ((dp == 0 && wa->b_i[m] <= wa->b_i[p]) || (dp == 1 && wa->b_i[m] >= wa->b_i[p])) vis = 0;
            e0 = ex; num = (om == 1) ? ex * ex : fabs(ex * (om == 2 ? ey : ez));
          } else if (wa->b_orient[p] == 2) {
            if ((dp == 0 && wa->b_j[m] <= wa->b_j[p]) || (dp == 1 && wa->b_j[m] >= wa->b_j[p])) vis = 0;
            e0 = ey; num = (om == 2) ? ey * ey : fabs(ey * (om == 1 ? ex : ez));
          } else {
            if ((dp == 0 && wa->b_k[m] <= wa->b_k[p]) || (dp == 1 && wa->b_k[m] >= wa->b_k[p])) vis = 0;
            e0 = ez; num = (om == 3) ? ez * ez : fabs(ez * (om == 1 ? ex : ey));
          }
          (void)e0;
          if (vis) v = num * wa->patch_area[m] / den;
        }
        wa->vf[(size_t)p + (size_t)nb * m] = v;
        s += v;
      }
      wa->vf_sum[p] = s;
    }
    for (int p = 0; p < nb; p++) {
      if (wa->vf_sum[p] == 0.0) continue;
      for (int m = 0; m < nb; m++) wa->vf[(size_t)p + (size_t)nb * m] /= wa->vf_sum[p];
    }
  }
  /* Step 8: patches on the same boundary cell (sameCell1..4) and the totals */
  int *sc = calloc((size_t)4 * nb, sizeof(int));  /* 1-based patch numbers, 0 = none */
  if (!sc) return GFM_E_DEVICE;
#define SC(q, s) sc[(size_t)(s) * nb + (q)]
  for (int p = 0; p < nb; p++)
    for (int m = p + 1; m < nb; m++) {
      if (!(wa->bc_i[p] == wa->bc_i[m] && wa->bc_j[p] == wa->bc_j[m] && wa->bc_k[p] == wa->bc_k[m])) continue;
      const int P1 = p + 1, M1 = m + 1;
      if (SC(p, 0) == 0) { SC(p, 0) = M1; SC(m, 0) = P1; }
      else if (SC(p, 1) == 0) {
        if (SC(p, 0) == M1) continue;
        SC(p, 1) = M1; SC(m, 0) = P1;
        SC(m, 1) = SC(p, 0); SC(SC(p, 0) - 1, 1) = M1;
      } else if (SC(p, 2) == 0) {
        if (SC(p, 0) == M1 || SC(p, 1) == M1) continue;
        SC(p, 2) = M1; SC(m, 0) = P1;
        SC(m, 1) = SC(p, 0); SC(SC(p, 1) - 1, 2) = M1;
        SC(m, 2) = SC(p, 1);
      } else if (SC(p, 3) == 0) {
        if (SC(p, 0) == M1 || SC(p, 1) == M1 || SC(p, 2) == M1) continue;
        SC(p, 3) = M1; SC(m, 0) = P1;
        SC(m, 1) = SC(p, 0); SC(SC(p, 1) - 1, 3) = M1;
        SC(m, 2) = SC(p, 1); SC(SC(p, 2) - 1, 3) = M1;
        SC(m, 3) = SC(p, 2);
      }
    }
  wc->iwall_count = wc->imelt_count = wc->inlet_count = wc->iexit_count = 0;
  wc->wall_area_total = wc->melt_area_total = wc->inlet_area_total = wc->exit_area_total = 0.0;
  for (int p = 0; p < nb; p++) {
    double a = wa->patch_area[p];
    for (int s = 0; s < 4 && SC(p, s); s++) a = a + wa->patch_area[SC(p, s) - 1];
    wa->b_face_area[p] = a;
    const double pa = wa->patch_area[p];
    if (wa->b_typ[p] == 1) { wc->iwall_count++; wc->wall_area_total += pa; }
    else if (wa->b_typ[p] == 4) { wc->imelt_count++; wc->melt_area_total += pa; }
    else if (wa->b_typ[p] == 2) { wc->inlet_count++; wc->inlet_area_total += pa; }
    else { wc->iexit_count++; wc->exit_area_total += pa; }
  }
#undef SC
  free(sc);
  wc->boundary_area_total = wc->wall_area_total + wc->melt_area_total + wc->inlet_area_total + wc->exit_area_total;
  return GFM_OK;
}

/* get_radiation_from_boundaries (preset view factors), C/Qwall.f90:1501-1507 */
static double wall_rad_from(const gfm_wall_arrays *wa, int nb, int p) {
  double s = 0.0;
  for (int m = 0; m < nb; m++) s = s + wa->radiosity[m] * wa->vf[(size_t)p + (size_t)nb * m];
  return s;
}

/* the open cell next to a patch and the face-to-centre distance, C/Qwall.f90:1401-1433 */
static void wall_open_cell(const gfm_wall_arrays *wa, int p, int *ic, int *jc, int *kc, double *ds) {
  int i = wa->b_i[p], j = wa->b_j[p], k = wa->b_k[p];
  const int dn = wa->b_direct[p] == 0;
  if (wa->b_orient[p] == 1) { if (dn) { *ds = X(i + 1) - X(i); i++; } else { *ds = X(i) - X(i - 1); i--; } }
  else if (wa->b_orient[p] == 2) { if (dn) { *ds = R(j + 1) - R(j); j++; } else { *ds = R(j) - R(j - 1); j--; } }
  else { if (dn) { *ds = Z(k + 1) - Z(k); k++; } else { *ds = Z(k) - Z(k - 1); k--; } }
  *ic = i; *jc = j; *kc = k;
}

/* calc_q_conv_wall, C/Qwall.f90:1370-1478 */
static void wall_q_conv(const gfm_wall_ctrl *wc, gfm_wall_arrays *wa, int nb, double *qconv_total) {
  const double kappa_cmu = 0.4 * pow(9.0e-2, 0.25), e_cmu = 9.0 * pow(9.0e-2, 0.25);
  const double c_htw = 0.4 * ((pi_w / 4) / sin(pi_w / 4)) * sqrt(26.0) * (0.7 - 1.0) / pow(0.7, 0.25);
  double tot = 0.0;
  for (int p = 0; p < nb; p++) {
    if (wa->b_typ[p] == 2 || wa->b_typ[p] == 3) { wa->q_conv[p] = 0.0; continue; }
    int i, j, k;
    double ds;
    wall_open_cell(wa, p, &i, &j, &k, &ds);
    const double cpmix = 1.0057e3 * (cpf1 * GF(i, j, k, iyf) + (2.0e3 / 1.0057e3) * GF(i, j, k, iyco2) +
                                     (2.0e3 / 1.0057e3) * GF(i, j, k, iyh2o) + (0.9136e3 / 1.0057e3) * GF(i, j, k, iyo2) +
                                     (0.526e3 / 1.0057e3) * GF(i, j, k, iyn2));
    const double tgas = T(i, j, k) * C->t0;
    const double gmu_t = ((2.0e-15 * tgas + -1.1e-11) * tgas + 4.5e-8) * tgas + 7.0e-6;
    double h_t;
    if (C->tk0 * GF(i, j, k, ik) < 0.1) h_t = 0.0;
    else {
      const double sq = sqrt(C->tk0 * GF(i, j, k, ik));
      const double dtmp = log(e_cmu * ds * C->dnst0 * DNST(i, j, k) * sq / gmu_t) + (c_htw / sq);
      h_t = (dtmp < 0.01) ? 0.0 : dmin(50.0, cpmix * kappa_cmu * C->dnst0 * DNST(i, j, k) * sq / dtmp);
    }
    wa->q_conv[p] = dmax(wc->h_g, h_t) * (tgas - wa->b_t[p]);
    tot = tot + wa->q_conv[p] * wa->patch_area[p];
  }
  *qconv_total = tot;
}

/* wall_q_amb, C/Qwall.f90:1585-1659 */
static void wall_q_amb(gfm_wall_ctrl *wc, gfm_wall_arrays *wa, int nb, const int16_t *iwall) {
  double loss = 0.0;
  const int MZ = MP / 2, NZ = NP / 2;
  for (int p = 0; p < nb; p++) {
    if (wa->b_typ[p] != 1) { wa->q_amb[p] = 0.0; continue; }
    int nwall = 0;
    if (iwall) nwall = iwall[((size_t)(wa->bc_k[p] / 2 - 1) * NZ + (size_t)(wa->bc_j[p] / 2 - 1)) * MZ + (size_t)(wa->bc_i[p] / 2 - 1)];
    double w_k0, h_a0, w_d0, t_a0;
    if (nwall > 0 && nwall <= wc->nwal0) {
      w_k0 = wc->w1_k[nwall - 1]; h_a0 = wc->w1_ha[nwall - 1]; w_d0 = wc->w1_d[nwall - 1]; t_a0 = wc->w1_ta[nwall - 1];
    } else { w_k0 = wc->w_k; h_a0 = wc->h_a; w_d0 = wc->w_d; t_a0 = wc->t_a; }
    const double u_a0 = (w_k0 < 1.0e-20 || h_a0 < 1.0e-20) ? 0.0 : 1.0 / (w_d0 / w_k0 + 1.0 / h_a0);
    wa->q_amb[p] = u_a0 * (wa->b_t[p] - t_a0);
    loss = loss + wa->q_amb[p] * wa->patch_area[p];
  }
  wc->q_wall_loss = loss;
  wc->qamb_avg = loss / wc->wall_area_total;
}

/* One pass's wall part: patch temperatures, wall_q_amb, wallq, QA_S0/QA_S1, the state carried to the gas. */
int oracle_wallq(gfm_ctrl *c, const gfm_host_arrays *h, gfm_wall_ctrl *wc, gfm_wall_arrays *wa, const double *qa,
                 double qeg_s, const int16_t *iwall) {
  int r = bind(c, h);
  if (r) return r;
  if (!wc || !wa || !qa || wc->b_cnt <= 0 || !wa->vf || wc->preset_vf != 1) return GFM_E_ARG;
  const int nb = wc->b_cnt;
  const int MZ = MP / 2, NZ = NP / 2;
#define QAC(i, j, k) qa[((size_t)((k) / 2 - 1) * NZ + (size_t)((j) / 2 - 1)) * MZ + (size_t)((i) / 2 - 1)]
  const double t0 = C->t0;
  /* patch temperatures at the start of the pass, C/Rad_emis.f90:190-212 and 253-261 (T in K) */
  for (int p = 0; p < nb; p++) {
    const int t = wa->b_typ[p];
    if (wc->i1st_rad_done == 0) {
      if (t == 1 || t == 2) wa->b_t[p] = T(wa->bc_i[p], wa->bc_j[p], wa->bc_k[p]) * t0;
      else if (t == 3) {
        const int e = wa->b_exh[p] - 1;
        if (e < 0) return GFM_E_ARG;
        wa->b_t[p] = (wc->exh_type[e] == 0) ? wc->exh_frac[e] * wc->avg_wall_T : wc->exh_fixed[e];
      } else wa->b_t[p] = T(wa->b_i[p], wa->b_j[p], wa->b_k[p]) * t0;
      wa->radiosity[p] = sig_w * pow(wa->b_t[p], 4);
    } else if (t == 4) {
      wa->b_t[p] = T(wa->b_i[p], wa->b_j[p], wa->b_k[p]) * t0;
      wa->radiosity[p] = sig_w * pow(wa->b_t[p], 4);
    }
  }
  wall_q_amb(wc, wa, nb, iwall);
  /* wallq: qc, qeb and the incident sums, C/Qwall.f90:997-1026 */
  wc->q_incident_total = wc->q_incident_wall = wc->q_incident_inlet = wc->q_incident_exit = wc->q_incident_melt = 0.0;
  for (int p = 0; p < nb; p++) {
    wa->qc[p] = QAC(wa->bc_i[p], wa->bc_j[p], wa->bc_k[p]) / wa->b_face_area[p];
    const double inc = wa->qc[p] * wa->patch_area[p];
    wc->q_incident_total += inc;
    switch (wa->b_typ[p]) {
      case 1: wc->q_incident_wall += inc; break;
      case 2: wc->q_incident_inlet += inc; wa->qeb[p] = sig_w * pow(wa->b_t[p], 4); break;
      case 3: wc->q_incident_exit += inc; wa->qeb[p] = sig_w * pow(wa->b_t[p], 4); break;
      case 4: wc->q_incident_melt += inc; wa->qeb[p] = wa->emis[p] * sig_w * pow(wa->b_t[p], 4); break;
    }
  }
  wc->qc_avg = wc->q_incident_total / wc->boundary_area_total;
  const double radi_min = sig_w * pow(t_mnk_w, 4);
  const int nlog = (wa->conv_log && wa->log_cap > 0) ? wa->log_cap : 0;
  double prev_q_melt_chg = 1.0e6, qconv_total = 0.0;
  int nn = 1, nnextra = wc->maxri1, sweeps = 0, outer = 0;
  (void)prev_q_melt_chg;
  while (nn <= wc->maxri1 && nn <= nnextra) {
    int minradi = wc->minri2;
    wall_q_conv(wc, wa, nb, &qconv_total);
    int n = 1, nextra = wc->maxri2;
    while (n <= wc->maxri2 && n <= nextra) {
      double mean = 0.0;
      wc->iterw++;
      for (int p = 0; p < nb; p++) {
        double rfb = 0.0;
        if (wa->b_typ[p] == 1 || wa->emis[p] < 1.0) rfb = wall_rad_from(wa, nb, p);
        const double old = wa->radiosity[p];
        if (wa->b_typ[p] == 1) wa->radiosity[p] = dmax(wa->qc[p] + wa->q_conv[p] - wa->q_amb[p] + rfb, wa->emis[p] * radi_min);
        else wa->radiosity[p] = wa->qeb[p] + (1.0 - wa->emis[p]) * (rfb + wa->qc[p]);
        if (!(old < wc->wall_radiosity_tolerance)) mean = mean + fabs((old - wa->radiosity[p]) / old);
      }
      mean = mean / nb;
      if (sweeps < nlog) wa->conv_log[sweeps] = mean;
      sweeps++;
      wc->rel_rad_chg_last = mean;
      if (mean < wc->wall_radiosity_tolerance && n > minradi)
        if (nextra == wc->maxri2) nextra = n + 3;
      n++;
    }
    /* wall temperatures and the melt-surface flux, C/Qwall.f90:1155-1191 */
    const double prev_q_melt = wc->q_melt;
    double q_melt = 0.0;
    for (int p = 0; p < nb; p++) {
      if (wa->b_typ[p] == 1) {
        const double e = wa->emis[p];
        wa->b_t[p] = pow(dmax(0.0, (wa->radiosity[p] - (wa->q_amb[p] - wa->q_conv[p]) * (1.0 - e) / e) / sig_w), 0.25);
        wa->b_t[p] = dmax(t_mnk_w, wa->b_t[p]);
      } else if (wa->b_typ[p] == 4) {
        const double rfb = wall_rad_from(wa, nb, p);
        const double q = wa->qc[p] + wa->q_conv[p] + rfb - wa->radiosity[p];
        if (wa->qrs) wa->qrs[(size_t)(wa->b_j[p] - 1) * MP + (size_t)(wa->b_i[p] - 1)] = q;
        q_melt = q_melt + q * wa->patch_area[p];
      }
    }
    wc->q_melt = q_melt;
    wall_q_amb(wc, wa, nb, iwall);
    wc->qrs_avg = q_melt / wc->melt_area_total;
    const double q_melt_chg = fabs((prev_q_melt - q_melt) / (q_melt + small16));
    if (q_melt_chg < wc->radiation_boundary_tolerance)
      if (nnextra == wc->maxri1) nnextra = nn + 1;
    prev_q_melt_chg = q_melt_chg;
    outer++;
    nn++;
  }
  wc->qconv_total = qconv_total;
  /* closing sums, C/Qwall.f90:1209-1247 */
  double txa = 0.0;
  wc->q_rad_inlet = wc->q_rad_outlet = wc->qm_flame = wc->qm_bnds = wc->qm_conv = 0.0;
  for (int p = 0; p < nb; p++) {
    const int t = wa->b_typ[p];
    if (t == 1) {
      T(wa->b_i[p], wa->b_j[p], wa->b_k[p]) = wa->b_t[p] / t0;
      if (wa->wall_rad_t) wa->wall_rad_t[IX3(wa->bc_i[p], wa->bc_j[p], wa->bc_k[p])] = wa->b_t[p];
      txa = txa + wa->b_t[p] * wa->patch_area[p];
    } else if (t == 2 || t == 3) {
      const double rfb = wall_rad_from(wa, nb, p);
      const double q = (wa->qc[p] + rfb - wa->qeb[p]) * wa->patch_area[p];
      if (t == 2) wc->q_rad_inlet += q; else wc->q_rad_outlet += q;
    } else if (t == 4) {
      const double rfb = wall_rad_from(wa, nb, p);
      const double e = wa->emis[p];
      const double f_melt = 1.0 - wa->qeb[p] / (e * (wa->qc[p] + rfb) + wa->q_conv[p]);
      wc->qm_flame += f_melt * e * wa->qc[p] * wa->patch_area[p];
      wc->qm_bnds += f_melt * e * rfb * wa->patch_area[p];
      wc->qm_conv += f_melt * wa->q_conv[p] * wa->patch_area[p];
    }
  }
  wc->avg_wall_T = txa / wc->wall_area_total;
  int frac = 0;
  for (int m = 0; m < wc->nex; m++) if (wc->exh_type[m] == 0) frac = 1;
  if (frac)
    for (int p = 0; p < nb; p++) {
      if (wa->b_typ[p] != 3) continue;
      const int e = wa->b_exh[p] - 1;
      wa->b_t[p] = (wc->exh_type[e] == 0) ? wc->exh_frac[e] * wc->avg_wall_T : wc->exh_fixed[e];
    }
  wc->qconv_avg = qconv_total / wc->boundary_area_total;
  /* soot calibration, C/Qwall.f90:1297-1320 */
  if (wc->isoot_cal == 1) {
    if (wc->q_melt < 0.0) wc->q_melt = 0.0;
    const double q_ratio = wc->q_melt / wc->q_melt_req;
    const double a_adj = dmax(dmin(q_ratio, 2.0), 0.5);
    if (a_adj > 1.0) wc->aform = wc->rf_afover * wc->aform / a_adj + (1.0 - wc->rf_afover) * wc->aform;
    else wc->aform = wc->rf_aform * wc->aform / a_adj + (1.0 - wc->rf_aform) * wc->aform;
  }
  /* QRN_PRN, C/Qrn.f90:399-402 (QEW_S and QES0 are zero) */
  const double g0 = qeg_s * 1.0e-6;
  wc->qa_s1 = g0 - wc->qa_s0;
  wc->qa_s0 = g0;
  wc->sweeps = sweeps;
  wc->outer = outer;
  /* the state of the next gas iterations: each wall cell takes its last wall patch's temperature, with its
   * enthalpy (rad_emis_rr, C/Rad_emis.f90:101-113) */
  for (int p = 0; p < nb; p++) {
    if (wa->b_typ[p] != 1) continue;
    T(wa->bc_i[p], wa->bc_j[p], wa->bc_k[p]) = wa->b_t[p] / t0;
  }
  for (int p = 0; p < nb; p++) {
    if (wa->b_typ[p] != 1) continue;
    T_TO_ENTH(wa->bc_i[p], wa->bc_j[p], wa->bc_k[p]);
  }
#undef QAC
  return GFM_OK;
}
size_t oracle_wall_ctrl_size(void) { return sizeof(gfm_wall_ctrl); }
/* INTP on the bound state (test support: face values of a state read from restart records) */
int oracle_intp(gfm_ctrl *c, const gfm_host_arrays *h) {
  int r = bind(c, h);
  if (r) return r;
  INTP();
  return GFM_OK;
}
