// gfm_kernels.hip -- CDNA4 kernels of the combustion-space SIMPLER
// iteration.  Each kernel restates one loop nest of the reference
// (C/ = combustion/Code/ of the GFM v4 package) over the parity-class
// layout of gfm_dev.h; per-point arithmetic follows the reference order of
// operations so that FP64 results track it to rounding.
#include "gfm_dev.h"
#include <string.h>

__constant__ DevCase g;
__constant__ double c_nist[7][5][2];  // a,b,c,d,e,f,h x species x {low,high}
__constant__ double c_cutT[5];

static const double h_nist_tab[7][5][2] = {
    {{-0.703029, 85.81217}, {29.659, 29.659}, {24.99735, 58.16639}, {26.092, 26.092}, {30.092, 41.96426}},
    {{108.4773, 11.26467}, {6.137261, 6.137261}, {55.18696, 2.720074}, {8.218801, 8.218801}, {6.832514, 8.622053}},
    {{-42.52157, -2.114146}, {-1.186521, -1.186521}, {-33.69137, -0.492289}, {-1.976141, -1.976141}, {6.793435, -1.49978}},
    {{5.862788, 0.13819}, {0.09578, 0.09578}, {7.948387, 0.038844}, {0.159274, 0.159274}, {-2.53448, 0.098119}},
    {{0.678565, -26.42221}, {-0.219663, -0.219663}, {-0.136638, -6.447293}, {0.044434, 0.044434}, {0.082139, -11.15764}},
    {{-76.84376, -153.5327}, {-9.861391, -9.861391}, {-403.6075, -425.9186}, {-7.98923, -7.98923}, {-250.881, -272.1797}},
    {{-74.8731, -74.8731}, {0.0, 0.0}, {-393.5224, -393.5224}, {0.0, 0.0}, {-241.8264, -241.8264}}};
static const double h_cutT[5] = {1300.0, 6000.0, 1200.0, 6000.0, 1700.0};

#define SMALL20 1.0e-20
#define SMALL16 1.0e-16
#define BIGV 1.0e20
#define WF 16.0426
#define WO2 31.9988
#define WCO2 44.0098
#define WH2O 18.0152
#define WN2 28.0134
#define CMU 9.0e-2
#define CT1 1.44
#define CT2 1.92
#define CPF1 (2.156e3 / 1.0057e3)

__device__ __forceinline__ long long lidx(int i, int j, int k) { return lat_index(g.g, i, j, k); }

// Coefficients AP, AS(.,d,m), BS of the equation being assembled/solved.
// Plain mode stores them in FP64, mixed mode in FP32 (the assembly and the
// line sweeps then move half the bytes); every reader converts to FP64.
__device__ __forceinline__ double CAP(long long q) { return g.mixed ? (double)g.apf[q] : g.ap[q]; }
__device__ __forceinline__ double CBS(long long q) { return g.mixed ? (double)g.bsf[q] : g.bs[q]; }
__device__ __forceinline__ double CAS(int s, long long q) { return g.mixed ? (double)g.asf[s][q] : g.as[s][q]; }
__device__ __forceinline__ void SAP(long long q, double v) { if (g.mixed) g.apf[q] = (float)v; else g.ap[q] = v; }
__device__ __forceinline__ void SBS(long long q, double v) { if (g.mixed) g.bsf[q] = (float)v; else g.bs[q] = v; }
__device__ __forceinline__ void SAS(int s, long long q, double v) { if (g.mixed) g.asf[s][q] = (float)v; else g.as[s][q] = v; }
#define MP g.g.MP
#define NP g.g.NP
#define LP g.g.LP
#define X(i) g.x[(i)-1]
#define DX(i) g.dx[(i)-1]
#define R(j) g.r[(j)-1]
#define DR(j) g.dr[(j)-1]
#define RS(j) g.rs[(j)-1]
#define Z(k) g.z[(k)-1]
#define DZ(k) g.dz[(k)-1]
#define IBC(i, j, k) ((int)g.ib[lidx(i, j, k)])
#define P_(i, j, k) g.p[lidx(i, j, k)]
#define T_(i, j, k) g.t[lidx(i, j, k)]
#define DN(i, j, k) g.dnst[lidx(i, j, k)]
#define TH(i, j, k) g.theta[lidx(i, j, k)]
#define TMU(i, j, k) g.tmu[lidx(i, j, k)]
#define GCP(i, j, k) g.gcp[lidx(i, j, k)]
#define GLAM(i, j, k) g.glam[lidx(i, j, k)]
#define GAMA(i, j, k) g.gama[lidx(i, j, k)]
#define GTE(i, j, k) g.gte[lidx(i, j, k)]
#define CRATE(i, j, k) g.crate[lidx(i, j, k)]
#define HRATE(i, j, k) g.hrate[lidx(i, j, k)]
#define UG(i, j, k, l) g.ug[(l)-1][lidx(i, j, k)]
#define GF(i, j, k, l) g.gf[(l)-1][lidx(i, j, k)]
#define FZ(i, j, k) g.fz[lidx(i, j, k)]
#define BS(i, j, k) g.bs[lidx(i, j, k)]
#define SFP(i, j, k) g.sfp[lidx(i, j, k)]
#define AP(i, j, k) CAP(lidx(i, j, k))
#define AS(i, j, k, d, m) g.as[((d)-1) + 3 * ((m)-1)][lidx(i, j, k)]

enum { IYF = 1, IYO2 = 2, IH = 3, IK = 4, IEPS = 5, IYCO2 = 6, IYN2 = 7, IYH2O = 9 };

// Position of the thread in a box launched on box_grid: one x row of the
// box per row of workgroups, blockIdx.y and blockIdx.z being j and k
// (uniform over the wavefront: the stencil kernels form their indices in
// scalar registers).
__device__ __forceinline__ bool box_pt(const Box &b, int &i, int &j, int &k) {
  int tx = blockIdx.x * blockDim.x + threadIdx.x;
  if (tx >= b.ni) return false;
  i = b.i0 + b.step * tx;
  j = b.j0 + b.step * (int)blockIdx.y;
  k = b.k0 + b.step * (int)blockIdx.z;
  return true;
}
// The same on box_grid_f, the (i, j) rows of a k-plane flattened along x,
// so that only the last wavefront of a plane has idle lanes: for the light
// kernels over boxes of 193-195 positions per row, whose last wavefront per
// row would otherwise hold 1-3 live lanes; blockIdx.y is k.
__device__ __forceinline__ bool box_ptf(const Box &b, int &i, int &j, int &k) {
  const int tx = blockIdx.x * blockDim.x + threadIdx.x;
  if (tx >= b.ni * b.nj) return false;
  const int jj = tx / b.ni;
  i = b.i0 + b.step * (tx - jj * b.ni);
  j = b.j0 + b.step * jj;
  k = b.k0 + b.step * (int)blockIdx.y;
  return true;
}

// The mapping of a box kernel whose launch shape an autotuner may choose:
// b.map 0 = box_pt on box_grid, 1 = box_ptf on box_grid_f, any block size.
// Each position is handled by one thread either way, so the results do not
// depend on the choice.
__device__ __forceinline__ bool box_at(const Box &b, int &i, int &j, int &k) {
  return b.map ? box_ptf(b, i, j, k) : box_pt(b, i, j, k);
}

// ---------------- wave-level reductions (blockDim == 64) ----------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ long long block_lin() {
  return blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

// ---------------- one-block finalisation of per-block partials ----------------
// A FIN_NT-thread block reduces nb partials of each of NQ slots (slot q at
// part + q*stride) to one value: a sum, or a maximum for the slots whose
// bit is set in maxmask (those values are >= 0, so 0 is the identity).
// Thread t takes partials t + FIN_NT*m into four accumulators used round
// robin (four loads in flight per slot), then the wavefronts and the
// block combine in a tree: a fixed order on every call.  The results are
// valid in thread 0.
#define FIN_NT 1024
template <int NQ>
__device__ void fin_reduce(const double *part, long long stride, int nb, unsigned maxmask, double (&r)[NQ]) {
  __shared__ double sh[NQ][FIN_NT / 64];
  double a[NQ][4];
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int u = 0; u < 4; u++) a[q][u] = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * FIN_NT < nb; b += 4 * FIN_NT) {
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const double v = part[q * stride + b + u * FIN_NT];
        a[q][u] = ((maxmask >> q) & 1) ? fmax(a[q][u], v) : a[q][u] + v;
      }
  }
  for (; b < nb; b += FIN_NT) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const double v = part[q * stride + b];
      a[q][0] = ((maxmask >> q) & 1) ? fmax(a[q][0], v) : a[q][0] + v;
    }
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const bool mx = (maxmask >> q) & 1;
    double v = mx ? fmax(fmax(a[q][0], a[q][1]), fmax(a[q][2], a[q][3])) : (a[q][0] + a[q][1]) + (a[q][2] + a[q][3]);
    v = mx ? wave_max(v) : wave_sum(v);
    if (l == 0) sh[q][w] = v;
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const bool mx = (maxmask >> q) & 1;
      const double v = l < FIN_NT / 64 ? sh[q][l] : 0.0;
      r[q] = mx ? wave_max(v) : wave_sum(v);
    }
  }
  __syncthreads();  // sh is free for the next call
}

// ======================================================================
// geometry, C/Grid3.f90:310-471 (no half-blocked cells)
// ======================================================================
// Face area and volume, C/Grid3.f90:389-471: a half-blocked cell (code -1
// or -2) scales its own face areas by BA_F(ix, n) and its volume by BV_F(n);
// a cell of positive code other than a wall takes BA_F of a half-blocked
// neighbour two lattice steps away along ix.
__device__ __forceinline__ double AREA_C(int i, int j, int k, int ix) {
  const int b = IBC(i, j, k);
  if (b == 1) return 0.0;
  double a;
  if (ix == 1) a = DR(j) * RS(j) * DZ(k);
  else if (ix == 2) a = DX(i) * DZ(k);
  else a = DX(i) * DR(j) * RS(j);
  if (b == 0) return a;
  if (b <= -1) return a * g.ba_f[(ix - 1) + 3 * (-b - 1)];
  int nb;
  if (ix == 1) {
    if (i <= MP - 2 && IBC(i + 2, j, k) <= -1) nb = -IBC(i + 2, j, k);
    else if (i >= 4 && IBC(i - 2, j, k) <= -1) nb = -IBC(i - 2, j, k);
    else return a;
  } else if (ix == 2) {
    if (j <= NP - 2 && IBC(i, j + 2, k) <= -1) nb = -IBC(i, j + 2, k);
    else if (j >= 4 && IBC(i, j - 2, k) <= -1) nb = -IBC(i, j - 2, k);
    else return a;
  } else {
    if (k <= LP - 2 && IBC(i, j, k + 2) <= -1) nb = -IBC(i, j, k + 2);
    else if (k >= 4 && IBC(i, j, k - 2) <= -1) nb = -IBC(i, j, k - 2);
    else return a;
  }
  return a * g.ba_f[(ix - 1) + 3 * (nb - 1)];
}
__device__ __forceinline__ double VOL_C0(int i, int j, int k) {
  const double v = DX(i) * DR(j) * RS(j) * DZ(k);
  const int b = IBC(i, j, k);
  return b <= -1 ? v * g.bv_f[-b - 1] : v;
}
__device__ double VOL_C(int i, int j, int k) {
  double vm, vp;
  if (i & 1) {
    vm = VOL_C0(i - 1, j, k); if (IBC(i - 2, j, k) <= 0) vm *= 0.5;
    vp = VOL_C0(i + 1, j, k); if (IBC(i + 2, j, k) <= 0) vp *= 0.5;
    return vp + vm;
  } else if (j & 1) {
    vm = VOL_C0(i, j - 1, k); if (IBC(i, j - 2, k) <= 0) vm *= 0.5;
    vp = VOL_C0(i, j + 1, k); if (IBC(i, j + 2, k) <= 0) vp *= 0.5;
    return vp + vm;
  } else if (k & 1) {
    vm = VOL_C0(i, j, k - 1); if (IBC(i, j, k - 2) <= 0) vm *= 0.5;
    vp = VOL_C0(i, j, k + 1); if (IBC(i, j, k + 2) <= 0) vp *= 0.5;
    return vp + vm;
  }
  return VOL_C0(i, j, k);
}
__device__ double DXAV(int i, int j, int k) {
  if (IBC(i - 2, j, k) >= 1) return X(i + 1) - X(i - 2);
  if (IBC(i + 2, j, k) >= 1) return X(i + 2) - X(i - 1);
  return DX(i);
}
__device__ double DRAV(int i, int j, int k) {
  if (IBC(i, j - 2, k) >= 1) return (RS(j + 1) + RS(j - 2)) * 0.5 / RS(j) * (R(j + 1) - R(j - 2));
  if (IBC(i, j + 2, k) >= 1) return (RS(j + 2) + RS(j - 1)) * 0.5 / RS(j) * (R(j + 2) - R(j - 1));
  return DR(j);
}
__device__ double DZAV(int i, int j, int k) {
  if (IBC(i, j, k - 2) >= 1) return Z(k + 1) - Z(k - 2);
  if (IBC(i, j, k + 2) >= 1) return Z(k + 2) - Z(k - 1);
  return DZ(k);
}

// ======================================================================
// properties, C/Prop.f90
// ======================================================================
__device__ __forceinline__ double DA(double d, double f) {  // C/Prop.f90:97-117
  if (d <= 1.0e-25) return 0.0;
  double fa = fabs(f), pe = fa / d;
  if (pe <= 1.0e-2) return d / (1.0 + pe * (0.5 + pe * (1.0 / 6.0 + pe / 24.0)));
  if (pe <= 20.0) return fa / (exp(pe) - 1.0);
  return 0.0;
}
__device__ __forceinline__ double TGAMF(double b) {  // C/Gflow.f90:417-426
  if (b < 11.5) return 1.0;
  return fmax(b / (log(9.0 * b) * 2.5), 1.0);
}
__device__ void CONC(long long q) {  // C/Prop.f90:28-59
  double yf = g.gf[0][q], yo = g.gf[1][q], yc = g.gf[5][q], yn = g.gf[6][q];
  double g0 = yf + yo + yc + yn;
  if (g0 > 1.0) {
    g0 = g0 + g.gf[8][q];
    g.gf[0][q] = yf / g0; g.gf[1][q] = yo / g0; g.gf[5][q] = yc / g0; g.gf[6][q] = yn / g0;
    g.gf[8][q] = g.gf[8][q] / g0;
  } else {
    g.gf[8][q] = 1.0 - g0;
  }
}
__device__ void DNST_CALC(long long q) {  // C/Prop.f90:187-198
  CONC(q);
  double wt = 1.0 / (g.gf[0][q] / WF + g.gf[5][q] / WCO2 + g.gf[8][q] / WH2O + g.gf[1][q] / WO2 + g.gf[6][q] / WN2);
  double temp = (1.0 + g.p[q]) / g.t[q] * wt / g.wt0;
  g.dnst[q] = fmax(temp, SMALL20);
}
__device__ void THER(long long q) {  // C/Prop.f90:130-154
  double e = fmax(g.gf[4][q], 1.0e-12);
  g.gf[4][q] = e;
  double kk = g.gf[3][q];
  double t = CMU * g.reyg * g.epsr / (2.0 * g.taur) * kk * kk / e * g.dnst[q];
  g.tmu[q] = t;
  g.glam[q] = 1.0;
  g.gcp[q] = CPF1;
  g.gama[q] = 2.0 / g.reyg * (t + 1.0);
}
__device__ __forceinline__ double BOUND_U(double f) {  // C/Prop.f90:257-260
  f = fmin(f, g.u_mx); f = fmax(f, g.u_mn);
  if (fabs(f) < 1.0e-12) f = 0.0;
  return f;
}
__device__ __forceinline__ double BOUND_P(double f) { return fmax(fmin(f, g.p_mx), g.p_mn); }
__device__ __forceinline__ double BOUND_GF(double f, int l) {  // C/Prop.f90:264-279
  if (l == IK) return fmax(f, 1.0e-6);
  if (l == IEPS) return fmax(f, 1.0e-12);
  if (l == IH) return f;
  return fmax(fmin(f, 1.0), 0.0);
}

struct Mix { double a, b, c, d, e, f, h; int lohi[5]; };
__device__ __forceinline__ void mix_add(Mix &m, double y, int n, int s, double sg) {
  m.a += sg * y * c_nist[0][n][s]; m.b += sg * y * c_nist[1][n][s]; m.c += sg * y * c_nist[2][n][s];
  m.d += sg * y * c_nist[3][n][s]; m.e += sg * y * c_nist[4][n][s];
  m.f = m.d + sg * y * c_nist[5][n][s]; m.h = m.e + sg * y * c_nist[6][n][s];
}
__device__ void mix_build(Mix &m, const double y[5], double realT) {
  m.a = m.b = m.c = m.d = m.e = m.f = m.h = 0.0;
  for (int n = 0; n < 5; n++) {
    int s = (realT < c_cutT[n]) ? 0 : 1;
    mix_add(m, y[n], n, s, 1.0);
    m.lohi[n] = s + 1;
  }
}
__device__ __forceinline__ void ygather(long long q, double y[5]) {
  y[0] = g.gf[0][q]; y[1] = g.gf[1][q]; y[2] = g.gf[5][q]; y[3] = g.gf[6][q]; y[4] = g.gf[8][q];
}
// enth_to_T, C/Prop.f90:369-505 (FP64 Newton with the reference's
// coefficient construction and lo/hi switching)
__device__ void ENTH_TO_T(long long q) {
  double y[5];
  ygather(q, y);
  double w_mix = 1.0 / (y[0] / WF + y[2] / WCO2 + y[4] / WH2O + y[1] / WO2 + y[3] / WN2);
  double h_molar = g.gf[2][q] * g.h_0 * w_mix / 1.0e6;
  double realT = g.t[q] * g.t0;
  Mix m;
  mix_build(m, y, realT);
  double T_old = realT / 1000.0, T_new = T_old;
  for (int n = 1; n <= 10; n++) {
    double gp = m.a + T_old * (m.b + T_old * (m.c + T_old * m.d)) + m.e / (T_old * T_old);
    if (gp == 0.0) break;
    double gg = T_old * (m.a + T_old * ((m.b / 2.0) + T_old * ((m.c / 3.0) + T_old * (m.d / 4.0)))) - (m.e / T_old) + m.f - m.h;
    double ch = (gg - h_molar) / gp;
    T_new = T_old - ch;
    if (fabs(ch) < 1.0e-11) break;
    T_old = T_new;
    for (int nn = 0; nn < 5; nn++) {
      if (nn == 1 || nn == 3) continue;
      if (T_new * 1000.0 < c_cutT[nn] && m.lohi[nn] != 1) { mix_add(m, y[nn], nn, 1, -1.0); m.lohi[nn] = 1; mix_add(m, y[nn], nn, 0, 1.0); }
      if (T_new * 1000.0 >= c_cutT[nn] && m.lohi[nn] != 2) { mix_add(m, y[nn], nn, 0, -1.0); m.lohi[nn] = 2; mix_add(m, y[nn], nn, 1, 1.0); }
    }
  }
  g.t[q] = T_new * 1000.0 / g.t0;
}
__device__ void T_TO_ENTH(long long q) {  // C/Prop.f90:525-565
  double y[5];
  ygather(q, y);
  double w_mix = 1.0 / (y[0] / WF + y[2] / WCO2 + y[4] / WH2O + y[1] / WO2 + y[3] / WN2);
  double realT = g.t[q] * g.t0;
  Mix m;
  mix_build(m, y, realT);
  double t = realT / 1000.0;
  double ef = t * (m.a + t * ((m.b / 2.0) + t * ((m.c / 3.0) + t * (m.d / 4.0)))) - (m.e / t) + m.f - m.h;
  g.gf[2][q] = ef * 1.0e6 / (w_mix * g.h_0);
}
__device__ double ZY1(const double *x, const double *y, double x1, int n) {  // C/Source.f90:814-831
  if (x1 < x[0]) return y[0] * x1 / x[0];
  for (int i = 1; i < n; i++)
    if (x1 < x[i]) return y[i - 1] + (y[i] - y[i - 1]) * ((x1 - x[i - 1]) / (x[i] - x[i - 1]));
  return y[n - 1];
}

// ======================================================================
// GFLOW wall treatment, C/Gflow.f90:16-130
// ======================================================================
// The three stages run over lists of lattice positions built once from
// IBCELL on the host (Fortran linear index (i-1) + MP*((j-1) + NP*(k-1))):
// the open positions with a wall neighbour (theta, wall functions) and the
// wall positions (GAMA = 0).  Every other position of the reference's
// loops leaves the fields unchanged.
__device__ __forceinline__ bool list_pt(const long long *lst, int n, int &i, int &j, int &k) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return false;
  const long long f = lst[t];
  i = (int)(f % MP) + 1;
  j = (int)((f / MP) % NP) + 1;
  k = (int)(f / ((long long)MP * NP)) + 1;
  return true;
}
__global__ void k_theta_wall(const long long *lst, int n) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  double g1, g2;
  if (IBC(i - 1, j, k) == 1) { g1 = (X(i) - X(i - 1)) * (X(i) - X(i - 1)); g2 = (X(i + 2) - X(i - 1)) * (X(i + 2) - X(i - 1)); TH(i - 1, j, k) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i + 2, j, k) * (g1 / (g1 - g2)); }
  if (IBC(i + 1, j, k) == 1) { g1 = (X(i) - X(i + 1)) * (X(i) - X(i + 1)); g2 = (X(i - 2) - X(i + 1)) * (X(i - 2) - X(i + 1)); TH(i + 1, j, k) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i - 2, j, k) * (g1 / (g1 - g2)); }
  if (IBC(i, j - 1, k) == 1) { g1 = (R(j) - R(j - 1)) * (R(j) - R(j - 1)); g2 = (R(j + 2) - R(j - 1)) * (R(j + 2) - R(j - 1)); TH(i, j - 1, k) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i, j + 2, k) * (g1 / (g1 - g2)); }
  if (IBC(i, j + 1, k) == 1) { g1 = (R(j) - R(j + 1)) * (R(j) - R(j + 1)); g2 = (R(j - 2) - R(j + 1)) * (R(j - 2) - R(j + 1)); TH(i, j + 1, k) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i, j - 2, k) * (g1 / (g1 - g2)); }
  if (IBC(i, j, k - 1) == 1) { g1 = (Z(k) - Z(k - 1)) * (Z(k) - Z(k - 1)); g2 = (Z(k + 2) - Z(k - 1)) * (Z(k + 2) - Z(k - 1)); TH(i, j, k - 1) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i, j, k + 2) * (g1 / (g1 - g2)); }
  if (IBC(i, j, k + 1) == 1) { g1 = (Z(k) - Z(k + 1)) * (Z(k) - Z(k + 1)); g2 = (Z(k - 2) - Z(k + 1)) * (Z(k - 2) - Z(k + 1)); TH(i, j, k + 1) = TH(i, j, k) * (g2 / (g2 - g1)) + TH(i, j, k - 2) * (g1 / (g1 - g2)); }
}
__global__ void k_gama_wall_zero(const long long *lst, int n) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  if (IBC(i, j, k) == 1) GAMA(i, j, k) = 0.0;
}
// Wall functions.  Each wall lattice position next to open fluid has one
// open neighbour in a rectangular furnace, so the updates do not collide.
__global__ void k_wall_function(const long long *lst, int n, double rt20) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  double blvm = rt20 * DN(i, j, k) * sqrt(GF(i, j, k, IK));
  double vgam = 1.0 / g.reyg2;
#define WF_UPD(cond, ia, ja, ka, ib_, jb_, kb_, dist)                    \
  if (cond) {                                                             \
    double gg = vgam * TGAMF(blvm * (dist));                              \
    GAMA(ia, ja, ka) = fmax(gg, GAMA(ia, ja, ka));                        \
    TMU(ia, ja, ka) = g.reyg2 * GAMA(ia, ja, ka);                         \
    TMU(ib_, jb_, kb_) = TMU(ia, ja, ka);                                 \
  }
  WF_UPD(IBC(i - 1, j, k) == 1, i - 1, j, k, i - 2, j, k, X(i) - X(i - 1))
  WF_UPD(IBC(i + 1, j, k) == 1, i + 1, j, k, i + 2, j, k, X(i + 1) - X(i))
  WF_UPD(IBC(i, j - 1, k) == 1, i, j - 1, k, i, j - 2, k, R(j) - R(j - 1))
  WF_UPD(IBC(i, j + 1, k) == 1, i, j + 1, k, i, j + 2, k, R(j + 1) - R(j))
  WF_UPD(IBC(i, j, k - 1) == 1, i, j, k - 1, i, j, k - 2, Z(k) - Z(k - 1))
  WF_UPD(IBC(i, j, k + 1) == 1, i, j, k + 1, i, j, k + 2, Z(k + 1) - Z(k))
#undef WF_UPD
}

// ======================================================================
// momentum, C/Gslv3x.f90 + SRC1 C/Source.f90:664-809
// ======================================================================
__device__ __forceinline__ void VINDX(int iu, int m, int i, int j, int k, int &i1, int &j1, int &k1, int &i2, int &j2, int &k2) {
  int n0 = (iu == 1) ? m : 0, n1 = (iu == 2) ? m : 0, n2 = (iu == 3) ? m : 0;
  i1 = i + n0; i2 = i1 + n0; j1 = j + n1; j2 = j1 + n1; k1 = k + n2; k2 = k1 + n2;
}
__device__ double FLW(int i, int j, int k, int ix, int m) {
  double f = TH(i, j, k) * DN(i, j, k) * UG(i, j, k, ix) * AREA_C(i, j, k, ix);
  return (m == 1) ? fmax(0.0, f) : fmax(0.0, -f);
}
__device__ double FLWD(int i, int j, int k, int ix, int m, double sgm) {
  double area = AREA_C(i, j, k, ix);
  double ddl = (ix == 1) ? DX(i) : (ix == 2) ? DR(j) : DZ(k);
  double f = TH(i, j, k) * DN(i, j, k) * UG(i, j, k, ix) * area;
  double diff = TH(i, j, k) * GAMA(i, j, k) / sgm * area / ddl;
  return (m == 1) ? DA(diff, f) + fmax(0.0, f) : DA(diff, f) + fmax(0.0, -f);
}
__device__ double FDDL(int i, int j, int k, int iu) {
  int b = IBC(i, j, k);
  if (iu == 1) {
    if (b <= 0) return DX(i);
    if (IBC(i + 1, j, k) == 0) return X(i + 1) - X(i);
    if (IBC(i - 1, j, k) == 0) return X(i) - X(i - 1);
    return BIGV;
  } else if (iu == 2) {
    if (b <= 0) return DR(j) * RS(j);
    if (IBC(i, j + 1, k) == 0) return (R(j + 1) - R(j)) * RS(j);
    if (IBC(i, j - 1, k) == 0) return (R(j) - R(j - 1)) * RS(j);
    return BIGV;
  }
  if (b <= 0) return DZ(k);
  if (IBC(i, j, k + 1) == 0) return Z(k + 1) - Z(k);
  if (IBC(i, j, k - 1) == 0) return Z(k) - Z(k - 1);
  return BIGV;
}
__device__ void ECELL(int i, int j, int k, int iu, double ffl[3][2]) {  // C/Gslv3x.f90:148-200
  if (iu == 1) {
    for (int m = 0; m < 2; m++) {
      int i1 = m ? i + 1 : i - 1, i2 = m ? i + 2 : i - 2;
      ffl[0][m] = (IBC(i2, j, k) >= 1) ? DX(i1) : fabs(X(i) - X(i1));
      ffl[1][m] = ffl[0][m] * DZ(k);
      ffl[2][m] = ffl[0][m] * RS(j) * DR(j);
    }
  } else if (iu == 2) {
    for (int m = 0; m < 2; m++) {
      int j1 = m ? j + 1 : j - 1, j2 = m ? j + 2 : j - 2;
      ffl[1][m] = (IBC(i, j2, k) >= 1) ? RS(j1) * DR(j1) : RS(j1) * fabs(R(j) - R(j1));
      ffl[0][m] = ffl[1][m] * DZ(k);
      ffl[2][m] = ffl[1][m] * DX(i);
    }
  } else {
    for (int m = 0; m < 2; m++) {
      int k1 = m ? k + 1 : k - 1, k2 = m ? k + 2 : k - 2;
      ffl[2][m] = (IBC(i, j, k2) >= 1) ? DZ(k1) : fabs(Z(k) - Z(k1));
      ffl[0][m] = ffl[2][m] * RS(j) * DR(j);
      ffl[1][m] = ffl[2][m] * DX(i);
    }
  }
}

// FZ = UG on a face class, C/Gslv3v.f90:26-34
__global__ void k_copy_ug_fz(Box b, int iu1) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  g.fz[q] = g.ug[iu1 - 1][q];
}

// SOURCE(IU1) -> SRC1 fused with GSLVX1 for one face (C/Gslv3x.f90:26-111).
// The velocity component is a template parameter, so the per-direction
// arrays below (vis, dxn, dx0, dx1, ffl) are indexed by constants and stay
// in registers.  It takes its position with box_pt only (one x row per row
// of workgroups, launched with the row mapping): j and k stay uniform over
// the wavefront, so the stencil's index and grid terms sit in scalar
// registers and the kernel keeps 4 waves per SIMD (3 for w) without
// scratch; under the mapping chosen at run time they took 136-186 vector
// registers (2-3 waves per SIMD).
template <int IU1>
__global__ __launch_bounds__(256) void k_momentum_assemble(Box b) {
  constexpr int iu1 = IU1;
  // HOIST: the neighbour values a branch on a cell code chooses between
  // are all loaded before the code is known, so no load waits on it; the
  // same values and operations, more registers (y, z faces: 125 and 157
  // VGPRs, 4 and 3 waves per SIMD; the z-face kernel 198 -> 176 us per
  // launch on 192x96x64).  The x-face instance keeps the branches: there
  // the extra registers cost it a wave per SIMD (181 -> 193 us).
  constexpr bool HOIST = IU1 != 1;
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  if (IBC(i, j, k) >= 1) { SBS(q, 0.0); return; }
  // ---- SRC1 ----
  double vis[3][2] = {{0, 0}, {0, 0}, {0, 0}}, dxn[3], dx1[3] = {0, 0, 0}, dx0[3] = {0, 0, 0};
  dxn[0] = DX(i); dxn[1] = DR(j) * RS(j); dxn[2] = DZ(k);
  const double fzc = FZ(i, j, k);
  for (int m = -1; m <= 1; m += 2) {
    int i1, j1, k1, i2, j2, k2;
    VINDX(iu1, m, i, j, k, i1, j1, k1, i2, j2, k2);
    if (iu1 == 1) { dx0[0] = DX(i); dx1[0] = DX(i1); }
    else if (iu1 == 2) { dx0[1] = DR(j) * RS(j); dx1[1] = DR(j1) * RS(j1); }
    else { dx0[2] = DZ(k); dx1[2] = DZ(k1); }
    const bool in2 = IBC(i2, j2, k2) == 2;
    double thgm;
    if (HOIST) {
      const double tg2 = TH(i2, j2, k2) * GAMA(i2, j2, k2), tg1 = TH(i1, j1, k1) * GAMA(i1, j1, k1);
      thgm = in2 ? tg2 : tg1;
    } else {
      thgm = in2 ? TH(i2, j2, k2) * GAMA(i2, j2, k2) : TH(i1, j1, k1) * GAMA(i1, j1, k1);
    }
    if (in2) {
      if (iu1 == 1) dxn[0] += fabs(X(i2) - X(i1));
      if (iu1 == 2) dxn[1] += fabs(R(j2) - R(j1)) * RS(j);
      if (iu1 == 3) dxn[2] += fabs(Z(k2) - Z(k1));
    }
    if (m == -1) vis[iu1 - 1][0] = thgm * (fzc - FZ(i2, j2, k2)) / dx1[iu1 - 1];
    else vis[iu1 - 1][1] = thgm * (FZ(i2, j2, k2) - fzc) / dx1[iu1 - 1];
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
      double v = TH(i1, j1, k1) * GAMA(i1, j1, k1) * du0 / dx0[iu1 - 1];
      if (iu == 2) v *= RS(j1);
      vis[iu - 1][m1 - 1] = v;
    }
    vms += (vis[iu - 1][1] - vis[iu - 1][0]) / dxn[iu - 1];
  }
  double bs = vms;
  bs = bs + TH(i, j, k) * DN(i, j, k) * g.gra[iu1 - 1];
  const double sfp = -0.0;
  // ---- GSLVX1 ----
  double ap = 0.0, g0 = 0.0;
  for (int m = -1; m <= 1; m += 2) {  // GSLV3X1
    int m1 = (m + 3) / 2, i1, j1, k1, i2, j2, k2;
    VINDX(iu1, m, i, j, k, i1, j1, k1, i2, j2, k2);
    const int ibc1 = IBC(i2, j2, k2);
    double a;
    if (HOIST) {
      const double aw = FLW(i2, j2, k2, iu1, m1), ad = FLWD(i1, j1, k1, iu1, m1, 1.0);
      a = ibc1 == 3 ? 0.0 : ibc1 == 2 ? aw : ad;
    } else {
      a = ibc1 == 3 ? 0.0 : ibc1 == 2 ? FLW(i2, j2, k2, iu1, m1) : FLWD(i1, j1, k1, iu1, m1, 1.0);
    }
    SAS((iu1 - 1) + 3 * (m1 - 1), q, a);
    ap += a;
    g0 += a * FZ(i2, j2, k2);
  }
  double ffl[3][2];
  ECELL(i, j, k, iu1, ffl);
  const int nx0 = (iu1 == 1), nx1 = (iu1 == 2), nx2 = (iu1 == 3);
  for (int iu = 1; iu <= 3; iu++) {
    if (iu == iu1) continue;
    for (int m = -1; m <= 1; m += 2) {
      int m1 = (m + 3) / 2, i1, j1, k1, i2, j2, k2;
      VINDX(iu, m, i, j, k, i1, j1, k1, i2, j2, k2);
      const int asl = (iu - 1) + 3 * (m1 - 1);
      if (IBC(i1, j1, k1) == 3) { SAS(asl, q, 0.0); continue; }
      double f = 0.0, diff = 0.0;
      int ii1 = 0, jj1 = 0, kk1 = 0;
      for (int n = -1; n <= 1; n += 2) {
        int n1 = (n + 3) / 2;
        ii1 = i1 + nx0 * n; jj1 = j1 + nx1 * n; kk1 = k1 + nx2 * n;
        const int ibc1 = IBC(ii1, jj1, kk1);
        const double area = ffl[iu - 1][n1 - 1];
        if (HOIST) {
          const long long qq = lidx(ii1, jj1, kk1);
          const double th = g.theta[qq], dn = g.dnst[qq], fzv = g.fz[qq], ga = g.gama[qq];
          const double fnew = f + th * dn * fzv * area;
          f = ibc1 == 3 ? f : fnew;
          if (ibc1 >= 2 && ibc1 <= 3) continue;
          double ddl = FDDL(ii1, jj1, kk1, iu);
          diff += th * ga / 1.0 * area / ddl;
        } else {
          if (ibc1 == 3) continue;
          f += TH(ii1, jj1, kk1) * DN(ii1, jj1, kk1) * FZ(ii1, jj1, kk1) * area;
          if (ibc1 == 2) continue;
          double ddl = FDDL(ii1, jj1, kk1, iu);
          diff += TH(ii1, jj1, kk1) * GAMA(ii1, jj1, kk1) / 1.0 * area / ddl;
        }
      }
      double fld = (m1 == 1) ? DA(diff, f) + fmax(0.0, f) : DA(diff, f) + fmax(0.0, -f);
      if (IBC(ii1, jj1, kk1) == 2 && IBC(i2, j2, k2) == 1) {
        bs += fld * (UG(ii1, j2, k2, iu1) - UG(i, j, k, iu1));
        SAS(asl, q, 0.0);
        continue;
      }
      SAS(asl, q, fld);
      ap += fld;
      g0 += fld * FZ(i2, j2, k2);
    }
  }
  // VOL1 (C/Gslv3x.f90:266-286): less half of a half-blocked neighbour cell
  const double area1 = AREA_C(i, j, k, iu1);
  double vol = (ffl[iu1 - 1][0] + ffl[iu1 - 1][1]) * area1;
  if (iu1 == 1) {
    if (IBC(i - 1, j, k) <= -1) vol -= DX(i - 1) * area1 / 2.0;
    if (IBC(i + 1, j, k) <= -1) vol -= DX(i + 1) * area1 / 2.0;
  } else if (iu1 == 2) {
    if (IBC(i, j - 1, k) <= -1) vol -= RS(j - 1) * DR(j - 1) * area1 / 2.0;
    if (IBC(i, j + 1, k) <= -1) vol -= RS(j + 1) * DR(j + 1) * area1 / 2.0;
  } else {
    if (IBC(i, j, k - 1) <= -1) vol -= DZ(k - 1) * area1 / 2.0;
    if (IBC(i, j, k + 1) <= -1) vol -= DZ(k + 1) * area1 / 2.0;
  }
  bs *= vol;
  ap += (0.0 - 0.0 - sfp) * vol;
  SBS(q, bs);
  SAP(q, ap);
  if (fabs(ap) > SMALL20) UG(i, j, k, iu1) = (g0 + bs) / ap;
}

// SOURCE(4) + pressure-equation coefficients, C/Gslv3v.f90:43-89 and
// C/Source.f90:35-74
__global__ void k_pressure_assemble(Box b) {
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  if (IBC(i, j, k) >= 1) { SBS(q, g.p[q]); return; }
  double ax = AREA_C(i, j, k, 1), ay = AREA_C(i, j, k, 2), az = AREA_C(i, j, k, 3);
  double frx1 = (TH(i - 1, j, k) * DN(i - 1, j, k) * UG(i - 1, j, k, 1)) * ax;
  double frx2 = (TH(i + 1, j, k) * DN(i + 1, j, k) * UG(i + 1, j, k, 1)) * ax;
  double fry1 = (TH(i, j - 1, k) * DN(i, j - 1, k) * UG(i, j - 1, k, 2)) * ay;
  double fry2 = (TH(i, j + 1, k) * DN(i, j + 1, k) * UG(i, j + 1, k, 2)) * ay;
  double frz1 = (TH(i, j, k - 1) * DN(i, j, k - 1) * UG(i, j, k - 1, 3)) * az;
  double frz2 = (TH(i, j, k + 1) * DN(i, j, k + 1) * UG(i, j, k + 1, 3)) * az;
  SBS(q, ((frx1 - frx2) + (fry1 - fry2) + (frz1 - frz2) + 0.0) / g.eun);
  double ap = -0.0;
  for (int d = 1; d <= 3; d++) {
    double area = (d == 1) ? ax : (d == 2) ? ay : az;
    for (int m = 1; m <= 2; m++) {
      int i1 = i, j1 = j, k1 = k;
      int o = (m == 1) ? -1 : 1;
      if (d == 1) i1 += o; else if (d == 2) j1 += o; else k1 += o;
      double a = 0.0;
      // theta and rho loaded with a_p, before the cell code and a_p decide (no load waits on them)
      const long long q1 = lidx(i1, j1, k1);
      const double apf = AP(i1, j1, k1), th = g.theta[q1], dn = g.dnst[q1];
      if (IBC(i1, j1, k1) <= 0 && fabs(apf) > SMALL20) {
        double av = (d == 1) ? DXAV(i1, j1, k1) / DX(i1) : (d == 2) ? DRAV(i1, j1, k1) / DR(j1) : DZAV(i1, j1, k1) / DZ(k1);
        a = th * th * dn * area * area / apf * av;
      }
      SAS((d - 1) + 3 * (m - 1), q, a);
      ap += a;
    }
  }
  SAP(q, ap);
}
__global__ void k_bound_p(Box b) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  g.p[q] = BOUND_P(g.p[q]);
}
// starred-momentum right-hand side, C/Gslv3v.f90:116-130
__global__ void k_starred(Box b, int iu1) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  double ap = CAP(q) / g.rf[iu1 - 1];
  SAP(q, ap);
  double vol = VOL_C(i, j, k), dpdx;
  if (iu1 == 1) dpdx = (P_(i - 1, j, k) - P_(i + 1, j, k)) / DX(i);
  else if (iu1 == 2) dpdx = (P_(i, j - 1, k) - P_(i, j + 1, k)) / DR(j) / RS(j);
  else dpdx = (P_(i, j, k - 1) - P_(i, j, k + 1)) / DZ(k);
  SBS(q, CBS(q) + g.eun * g.theta[q] * vol * dpdx + (1.0 - g.rf[iu1 - 1]) * ap * g.fz[q]);
}
__global__ void k_bound_store_ug(Box b, int iu1) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  double f = BOUND_U(g.fz[q]);
  g.fz[q] = f;
  g.ug[iu1 - 1][q] = f;
}

// SOURCE(5): mass source of the starred velocities and res_mass
// (C/Source.f90:41-101).  Per-block partial sum / max / argmax.
__global__ void k_pcorr_source(Box b, double *part) {
  int i, j, k;
  double rm = 0.0, rmax = 0.0;
  long long key = 0x7fffffffffffffffLL;
  if (box_pt(b, i, j, k) && IBC(i, j, k) < 1) {
    long long q = lidx(i, j, k);
    double ax = AREA_C(i, j, k, 1), ay = AREA_C(i, j, k, 2), az = AREA_C(i, j, k, 3);
    double frx1 = (TH(i - 1, j, k) * DN(i - 1, j, k) * UG(i - 1, j, k, 1)) * ax;
    double frx2 = (TH(i + 1, j, k) * DN(i + 1, j, k) * UG(i + 1, j, k, 1)) * ax;
    double fry1 = (TH(i, j - 1, k) * DN(i, j - 1, k) * UG(i, j - 1, k, 2)) * ay;
    double fry2 = (TH(i, j + 1, k) * DN(i, j + 1, k) * UG(i, j + 1, k, 2)) * ay;
    double frz1 = (TH(i, j, k - 1) * DN(i, j, k - 1) * UG(i, j, k - 1, 3)) * az;
    double frz2 = (TH(i, j, k + 1) * DN(i, j, k + 1) * UG(i, j, k + 1, 3)) * az;
    double bs = ((frx1 - frx2) + (fry1 - fry2) + (frz1 - frz2) + 0.0) / g.eun;
    SBS(q, bs);
    bool exitadj = IBC(i - 1, j, k) == 3 || IBC(i + 1, j, k) == 3 || IBC(i, j - 1, k) == 3 ||
                   IBC(i, j + 1, k) == 3 || IBC(i, j, k - 1) == 3 || IBC(i, j, k + 1) == 3;
    if (!exitadj) {
      double fmx = fmax(fmax(fmax(fabs(frx1), fabs(frx2)), fmax(fabs(fry1), fabs(fry2))), fmax(fmax(fabs(frz1), fabs(frz2)), 0.0));
      double r = (fmx == 0.0) ? 0.0 : fabs(bs) * g.eun / fmx;
      rm = r;
      rmax = r;
      key = ((long long)i * NP + j) * LP + k;
    }
  }
  double s = wave_sum(rm);
  // argmax: largest r, smallest reference loop-order key on ties
  for (int o = 32; o > 0; o >>= 1) {
    double r2 = __shfl_down(rmax, o, 64);
    long long k2 = __shfl_down(key, o, 64);
    if (r2 > rmax || (r2 == rmax && k2 < key)) { rmax = r2; key = k2; }
  }
  if (threadIdx.x == 0) {
    long long bl = block_lin();
    long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    part[bl] = s;
    part[nb + bl] = rmax;
    ((long long *)part)[2 * nb + bl] = key;
  }
}
__global__ __launch_bounds__(FIN_NT) void k_pcorr_finalize(const double *part, long long nb) {
  __shared__ double sm[FIN_NT];
  __shared__ long long sk[FIN_NT];
  double rs[1];
  fin_reduce<1>(part, 0, (int)nb, 0u, rs);
  double m = 0.0;
  long long key = 0x7fffffffffffffffLL;
  for (long long b = threadIdx.x; b < nb; b += FIN_NT) {
    double r = part[nb + b];
    long long k2 = ((const long long *)part)[2 * nb + b];
    if (r > m || (r == m && k2 < key)) { m = r; key = k2; }
  }
  sm[threadIdx.x] = m; sk[threadIdx.x] = key;
  __syncthreads();
  for (int o = FIN_NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      double r2 = sm[threadIdx.x + o];
      long long k2 = sk[threadIdx.x + o];
      if (r2 > sm[threadIdx.x] || (r2 == sm[threadIdx.x] && k2 < sk[threadIdx.x])) { sm[threadIdx.x] = r2; sk[threadIdx.x] = k2; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    g.scal[SC_RES_MASS] = rs[0] / g.ncells;
    g.scal[SC_RES_MAX] = sm[0];
    long long kk = sk[0];
    if (sm[0] > 0.0) {
      g.scal[SC_KMX] = (double)(kk % LP);
      g.scal[SC_JMX] = (double)((kk / LP) % NP);
      g.scal[SC_IMX] = (double)(kk / ((long long)LP * NP));
    } else {
      g.scal[SC_IMX] = g.scal[SC_JMX] = g.scal[SC_KMX] = 2.0;
    }
  }
}
__global__ void k_zero_fz(Box b) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  FZ(i, j, k) = 0.0;
}
// velocity correction, C/Gslv3v.f90:179-204
__global__ void k_vel_correct(Box b, int iu1) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  const double apq = CAP(q);
  if (IBC(i, j, k) >= 1 || fabs(apq) < SMALL20) return;
  double vol = VOL_C(i, j, k), dpdx;
  if (iu1 == 1) dpdx = (FZ(i - 1, j, k) - FZ(i + 1, j, k)) / DX(i);
  else if (iu1 == 2) dpdx = (FZ(i, j - 1, k) - FZ(i, j + 1, k)) / DR(j) / RS(j);
  else dpdx = (FZ(i, j, k - 1) - FZ(i, j, k + 1)) / DZ(k);
  double g0 = g.eun * g.theta[q] * dpdx * vol / (apq * g.rf[iu1 - 1]);
  g.ug[iu1 - 1][q] = BOUND_U(g.ug[iu1 - 1][q] + g0);
}

// ======================================================================
// EXTRV, C/Extr.f90:358-567: exit faces of direction iu.  Partials:
// gfex, fuel_out, Tmean_ex, Tarea, area_ex (5 per block).
// ======================================================================
__global__ void k_extrv(Box b, int iu, int mex, double *part) {
  int i, j, k;
  double s[5] = {0, 0, 0, 0, 0};
  if (box_ptf(b, i, j, k) && (IBC(i, j, k) % 10) == 3) {
    int done = 0;
    for (int m = -1; m <= 1 && !done; m += 2) {
      int ia = i, ja = j, ka = k;
      if (iu == 1) ia += m; else if (iu == 2) ja += m; else ka += m;
      if (IBC(ia, ja, ka) >= 1) continue;
      int i1 = i, j1 = j, k1 = k, i0 = i, j0 = j, k0 = k;
      if (iu == 1) { i1 = i + 2 * m; i0 = i - m; } else if (iu == 2) { j1 = j + 2 * m; j0 = j - m; } else { k1 = k + 2 * m; k0 = k - m; }
      double *u1 = &UG(i1, j1, k1, iu);
      if (*u1 * m > 0.0) *u1 = 0.0;
      double area = AREA_C(i, j, k, iu);
      s[4] += area;
      double g0 = TH(i1, j1, k1) * DN(i1, j1, k1) * (*u1) * area;
      if (mex == 0) {
        s[0] += fabs(g0);
        s[1] += GF(i, j, k, IYF) * fabs(g0) * g.dnst0 * g.ug0 * g.area0;
        s[2] += T_(i, j, k) * fabs(g0);
        s[3] += T_(i, j, k) * area;
        done = 1;
        break;
      }
      double fac = g.scal[SC_FACTOR];
      *u1 = *u1 * fac;
      g0 = g0 * fac;
      if (iu != 3) {
        double u = BOUND_U(g0 / (TH(i, j, k) * DN(i, j, k) * area));
        UG(i, j, k, iu) = u;
        UG(i0, j0, k0, iu) = u;
      }
      done = 1;
    }
    if (!done) {
      // copy from an open neighbour face of the same class, in the
      // reference's order for this direction
      int c[4][3];
      bool inr[4];
      if (iu == 1) {
        c[0][0] = i; c[0][1] = j - 2; c[0][2] = k; inr[0] = j > 2;
        c[1][0] = i; c[1][1] = j + 2; c[1][2] = k; inr[1] = j < NP;
        c[2][0] = i; c[2][1] = j; c[2][2] = k - 2; inr[2] = k > 2;
        c[3][0] = i; c[3][1] = j; c[3][2] = k + 2; inr[3] = k < LP;
      } else if (iu == 2) {
        c[0][0] = i; c[0][1] = j; c[0][2] = k - 2; inr[0] = k > 2;
        c[1][0] = i; c[1][1] = j; c[1][2] = k + 2; inr[1] = k < LP;
        c[2][0] = i - 2; c[2][1] = j; c[2][2] = k; inr[2] = i > 2;
        c[3][0] = i + 2; c[3][1] = j; c[3][2] = k; inr[3] = i < MP;
      } else {
        c[0][0] = i - 2; c[0][1] = j; c[0][2] = k; inr[0] = i > 2;
        c[1][0] = i + 2; c[1][1] = j; c[1][2] = k; inr[1] = i < MP;
        c[2][0] = i; c[2][1] = j - 2; c[2][2] = k; inr[2] = j > 2;
        c[3][0] = i; c[3][1] = j + 2; c[3][2] = k; inr[3] = j < NP;
      }
      for (int n = 0; n < 4; n++) {
        if (inr[n] && IBC(c[n][0], c[n][1], c[n][2]) <= 0) { UG(i, j, k, iu) = UG(c[n][0], c[n][1], c[n][2], iu); break; }
      }
    }
  }
  if (mex != 0) return;
  for (int q = 0; q < 5; q++) {
    double v = wave_sum(s[q]);
    if (threadIdx.x == 0) part[(long long)q * g.pstride + block_lin()] = v;
  }
}

// generic finalize: sum nb partials of each of nq slots into scal[dst0+q]
__global__ __launch_bounds__(FIN_NT) void k_sum_partials(const double *part, int nb, int nq, int dst0, int accumulate) {
  for (int q = 0; q < nq; q++) {
    double r[1];
    fin_reduce<1>(part + (long long)q * g.pstride, 0, nb, 0u, r);
    if (threadIdx.x == 0) g.scal[dst0 + q] = accumulate ? g.scal[dst0 + q] + r[0] : r[0];
  }
}
// FLX1D's GFIN and the outlet flow factor, C/Flx1d.f90:75-77, C/Gslv3v.f90:213-231
__global__ void k_factor() {
  const int itr_gas = (int)g.scal[SC_ITR];
  double *s = g.scal;
  double gfex = s[SC_GFEX] * g.gmfr;
  s[SC_GFEX] = gfex;
  if (gfex < SMALL16) s[SC_TMEAN_EX] = s[SC_TAREA] * g.t0 / s[SC_AREA_EX];
  else s[SC_TMEAN_EX] = s[SC_TMEAN_EX] * g.t0 * g.gmfr / gfex;
  s[SC_GFIN] = s[SC_GFIN0] * g.gmfr;
  if (itr_gas > 10) {
    if (gfex < SMALL20) { s[SC_FACTOR] = 1.0; s[SC_FADJ] = 0.0; }
    else {
      double f = s[SC_GFIN] / gfex;
      if (f > g.f_mx) { f = g.f_mx; s[SC_FADJ] = 0.0; }
      else if (f < g.f_mn) { f = g.f_mn; s[SC_FADJ] = 0.0; }
      s[SC_FACTOR] = f;
    }
  }
}

// ======================================================================
// INTPV, C/Intp.f90:84-139,230-272: velocity component iu interpolated in
// direction ix onto the box positions.
// ======================================================================
__device__ __forceinline__ double INTP1(int i, int j, int k, int ix) {
  if (ix == 3) return (DZ(k) <= 0.0 || k > LP) ? 0.0 : (Z(k) - Z(k - 1)) / DZ(k);
  if (ix == 2) return (DR(j) <= 0.0 || j > NP) ? 0.0 : (R(j) - R(j - 1)) / DR(j);
  return (DX(i) <= 0.0 || i > MP) ? 0.0 : (X(i) - X(i - 1)) / DX(i);
}
// INTPV at one position: velocity component iu interpolated along ix.
__device__ __forceinline__ void intpv_at(int i, int j, int k, int iu, int ix) {
  int bc = IBC(i, j, k);
  if (bc >= 1 && bc != 4) return;
  double f1 = INTP1(i, j, k, ix), f2 = 1.0 - f1;
  double v;
  if (ix == 1) {
    v = UG(i + 1, j, k, iu) * f1 + UG(i - 1, j, k, iu) * f2;
  } else if (ix == 2) {
    if (IBC(i, j + 1, k) == 2 && bc == 0) { f1 = 0.0; f2 = 1.0; }
    else if (IBC(i, j - 1, k) == 2 && bc == 0) { f1 = 1.0; f2 = 0.0; }
    v = UG(i, j + 1, k, iu) * f1 + UG(i, j - 1, k, iu) * f2;
  } else {
    if (IBC(i, j, k + 1) == 2 && bc == 0) { f1 = 0.0; f2 = 1.0; }
    else if (IBC(i, j, k - 1) == 2 && bc == 0) { f1 = 1.0; f2 = 0.0; }
    v = UG(i, j, k + 1, iu) * f1 + UG(i, j, k - 1, iu) * f2;
  }
  UG(i, j, k, iu) = BOUND_U(v);
}
// One stage of INTPV in one launch: up to INTPV_MAXE boxes, each with its
// component and direction; workgroups wg0[e] .. wg0[e+1]-1 take box e, its
// positions flattened (i fastest).  The boxes of a stage write disjoint
// positions and read none that the stage writes.
#define INTPV_MAXE 12
#define INTPV_NT 256
struct IntpvSet {
  int n;
  int iu[INTPV_MAXE], ix[INTPV_MAXE];
  int wg0[INTPV_MAXE + 1];
  Box b[INTPV_MAXE];
};
__global__ __launch_bounds__(INTPV_NT) void k_intpv(const IntpvSet s) {
  const int bid = blockIdx.x;
  int e = 0;
  while (e + 1 < s.n && bid >= s.wg0[e + 1]) e++;
  const Box &b = s.b[e];
  const long long t = (long long)(bid - s.wg0[e]) * INTPV_NT + threadIdx.x;
  if (t >= (long long)b.ni * b.nj * b.nk) return;
  const long long r = t / b.ni;
  const int ii = (int)(t - r * b.ni), jj = (int)(r % b.nj), kk = (int)(r / b.nj);
  intpv_at(b.i0 + b.step * ii, b.j0 + b.step * jj, b.k0 + b.step * kk, s.iu[e], s.ix[e]);
}

// pressure referencing at the reference inlet cell, C/Gflow.f90:182-242
__global__ void k_set_itr(int itr) { g.scal[SC_ITR] = (double)itr; }

__global__ void k_pref() {
  int i = g.i_ref, j = g.j_ref, k = g.k_ref;
  int i1 = i, j1 = j, k1 = k, i2 = i, j2 = j, k2 = k, nx = 0;
  int MPM2 = MP - 2, NPM2 = NP - 2, LPM2 = LP - 2;
  if (i < MPM2 && IBC(i + 2, j, k) == 0) { nx = 1; i1 = i + 2; i2 = (IBC(i + 4, j, k) == 0) ? i + 4 : i1; }
  else if (i > 2 && IBC(i - 2, j, k) == 0) { nx = 1; i1 = i - 2; i2 = (IBC(i - 4, j, k) == 0) ? i - 4 : i1; }
  else if (j < NPM2 && IBC(i, j + 2, k) == 0) { nx = 2; j1 = j + 2; j2 = (j < NPM2 && IBC(i, j + 4, k) <= 0) ? j + 4 : j1; }
  else if (j > 2 && IBC(i, j - 2, k) <= 0) { nx = 2; j1 = j - 2; j2 = (j > 4 && IBC(i, j - 4, k) <= 0) ? j - 4 : j1; }
  else if (k < LPM2 && IBC(i, j, k + 2) <= 0) { nx = 3; k1 = k + 2; k2 = (k < LPM2 && IBC(i, j, k + 4) <= 0) ? k + 4 : k1; }
  else if (k > 2 && IBC(i, j, k - 2) <= 0) { nx = 3; k1 = k - 2; k2 = (k > 4 && IBC(i, j, k - 4) <= 0) ? k - 4 : k1; }
  double f2 = 0.0;
  int ibc0 = IBC(i, j, k) % 10;
  if (nx == 3) { double d = Z(k2) - Z(k1); if (d != 0.0) { int k0 = (ibc0 == 2) ? (k + k1) / 2 : k; f2 = (Z(k0) - Z(k1)) / d; } }
  else if (nx == 2) { double d = R(j2) - R(j1); if (d != 0.0) { int j0 = (ibc0 == 2) ? (j + j1) / 2 : j; f2 = (R(j0) - R(j1)) / d; } }
  else if (nx == 1) { double d = X(i2) - X(i1); if (d != 0.0) { int i0 = (ibc0 == 2) ? (i + i1) / 2 : i; f2 = (X(i0) - X(i1)) / d; } }
  double f1 = 1.0 - f2;
  P_(i, j, k) = P_(i1, j1, k1) * f1 + P_(i2, j2, k2) * f2;
  g.scal[SC_G0REF] = P_(i1, j1, k1) - g.p_ref;
}

// ======================================================================
// gray radiation source and wall convective flux (reference extension,
// enabled by gray_radiation / isothermal_walls)
// ======================================================================
__device__ __forceinline__ double EMF(double xkl, double xl0) {  // C/Rad_fun.f90:97-116
  const double x1 = 1.0e-3;
  const double g1 = (1.0 - exp(-x1)) / x1;
  double x0 = xkl * xl0;
  return (x0 <= x1) ? x0 * g1 : 1.0 - exp(-x0);
}
// net gray radiation into an open cell (W): QA - QE with QE from
// C/Rad0.f90:112-157 (one band) and QA from an isothermal black enclosure
__device__ double GRAY_QNET(int i, int j, int k) {
  double tk = T_(i, j, k) * g.t0;
  double vol = VOL_C(i, j, k);
  double sa = ((DX(i) * DR(j)) + (DR(j) * DZ(k)) + (DX(i) * DZ(k))) * 2.0;
  if (sa <= SMALL16) return 0.0;
  double opl = 6.0 * vol / sa;
  if (opl <= 1.0e-10) return 0.0;
  double e = EMF(g.kappa_gray, opl);
  double tw = g.t_wall * g.t0;
  double qe = (tk >= 300.0) ? sa * g.sigma_sb * tk * tk * tk * tk * e : 0.0;
  double qa = sa * g.sigma_sb * tw * tw * tw * tw * e;
  return qa - qe;
}
// convective flux (W/m2) to a wall or melt cell at twk (K), C/Qwall.f90:1419-1467
__device__ double Q_CONV_FACE(long long q, double ds, double twk) {
  const double kappa_cmu = 0.4 * pow(9.0e-2, 0.25), e_cmu = 9.0 * pow(9.0e-2, 0.25);
  const double pi = 3.141592653589793;
  const double c_htw = 0.4 * ((pi / 4) / sin(pi / 4)) * sqrt(26.0) * (0.7 - 1.0) / pow(0.7, 0.25);
  double cpmix = 1.0057e3 * (CPF1 * g.gf[0][q] + (2.0e3 / 1.0057e3) * g.gf[5][q] + (2.0e3 / 1.0057e3) * g.gf[8][q] +
                             (0.9136e3 / 1.0057e3) * g.gf[1][q] + (0.526e3 / 1.0057e3) * g.gf[6][q]);
  double tgas = g.t[q] * g.t0;
  double gmu_t = ((2.0e-15 * tgas + -1.1e-11) * tgas + 4.5e-8) * tgas + 7.0e-6;
  double h_t;
  if (g.tk0 * g.gf[3][q] < 0.1) h_t = 0.0;
  else {
    double sq = sqrt(g.tk0 * g.gf[3][q]);
    double dtmp = log(e_cmu * ds * g.dnst0 * g.dnst[q] * sq / gmu_t) + (c_htw / sq);
    h_t = (dtmp < 0.01) ? 0.0 : fmin(50.0, cpmix * kappa_cmu * g.dnst0 * g.dnst[q] * sq / dtmp);
  }
  return fmax(g.h_g, h_t) * (tgas - twk);
}
// wall sink of one open cell, C/Source.f90:473-509; qw gets sum q_conv*area,
// qlo / qhi the part through the cell's bottom (k-) / top (k+) face
__device__ double WALL_SINK(int i, int j, int k, double &qw, double &qlo, double &qhi) {
  double s = 0.0;
  long long q = lidx(i, j, k);
  for (int d = 0; d < 6; d++) {
    int di = (d == 0) ? -2 : (d == 1) ? 2 : 0, dj = (d == 2) ? -2 : (d == 3) ? 2 : 0, dk = (d == 4) ? -2 : (d == 5) ? 2 : 0;
    int b = IBC(i + di, j + dj, k + dk);
    if (b != 1 && b != 4) continue;
    double dsw, dsc, area;
    if (di) { dsw = fabs(X(i + di / 2) - X(i)); dsc = DX(i); area = DR(j) * DZ(k); }
    else if (dj) { dsw = fabs(R(j + dj / 2) - R(j)); dsc = DR(j); area = DX(i) * DZ(k); }
    else { dsw = fabs(Z(k + dk / 2) - Z(k)); dsc = DZ(k); area = DX(i) * DR(j); }
    double qc = Q_CONV_FACE(q, dsw, T_(i + di, j + dj, k + dk) * g.t0);
    qw += qc * area;
    if (d == 4) qlo += qc * area;
    if (d == 5) qhi += qc * area;
    s += qc / (dsc * g.r0 * g.qr0);
  }
  return s;
}

// ======================================================================
// scalar transport, C/Gslv3g.f90 + SOURCE(6..12) C/Source.f90:104-654
// ======================================================================
__global__ void k_copy_gf_fz(Box b, int l) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  g.fz[q] = g.gf[l - 1][q];
}
// source + flux coefficients (GSLV3G0) + volume and relaxation terms of one
// scalar equation at one cell.  igs=0 recomputes AS.  part[] collects
// q_react for the enthalpy equation.
// The equation NELG is a template parameter: each instance keeps only its
// own source branch (fewer registers, no branches on it).
template <int NELG>
__global__ void k_scalar_assemble(Box b, int l, int igs, double sgm, double *part) {
  constexpr int nelg = NELG;
  const int itr_gas = (int)g.scal[SC_ITR];
  int i, j, k;
  double qr = 0.0, qw = 0.0, qlo = 0.0, qhi = 0.0;
  if (box_pt(b, i, j, k)) {
    long long q = lidx(i, j, k);
    const double fz = g.fz[q];
    if (IBC(i, j, k) >= 1) {
      SBS(q, fz);
      if (nelg == 6) { g.crate[q] = 0.0; g.hrate[q] = 0.0; }
    } else {
      double bs = 0.0, sfp = 0.0;
      if (nelg == 6) {  // fuel, integral reaction model C/Source.f90:127-242
        double d11 = fabs(UG(i + 1, j, k, 1) / DX(i)), d12 = fabs(UG(i - 1, j, k, 1) / DX(i));
        double d21 = fabs(UG(i, j + 1, k, 2) / DR(j)), d22 = fabs(UG(i, j - 1, k, 2) / DR(j));
        double d31 = fabs(UG(i, j, k + 1, 3) / DZ(k)), d32 = fabs(UG(i, j, k - 1, 3) / DZ(k));
        double dtime = fmax(fmax(fmax(d11, d12), fmax(d21, d22)), fmax(fmax(d31, d32), 1.0e-14));
        dtime = 1.0 / dtime;
        double drtime = dtime * g.r0 / g.ug0;
        double yf = g.gf[0][q];
        double ayf = fmin(yf, g.gf[1][q] / g.stoich);
        double cr = 0.0, hr = 0.0;
        if (ayf >= 1.0e-5) {
          double zp1 = g.gf[5][q] / WCO2, delr1;
          if (zp1 <= SMALL20) delr1 = 0.0;
          else { double zf = yf / WF; delr1 = zp1 / (zf + zp1); }
          double time1 = ZY1(g.delr, g.timed, delr1, g.nr0);
          double delr2 = ZY1(g.rtim, g.delrt, time1 + drtime, g.nr0);
          double ayf0 = ayf / (1.0 - delr1) * g.dnst[q] * g.theta[q];
          double sr = ayf0 * (delr2 - delr1) / dtime;
          if (sr > 0.0) { cr = sr; sfp = -cr / (yf + SMALL20); hr = cr * g.q0; }
        }
        g.crate[q] = cr;
        g.hrate[q] = hr;
        bs = 0.0;
      } else if (nelg == 7) {
        bs = g.crate[q] * (WCO2 / WF);
      } else if (nelg == 12) {
        bs = -g.crate[q] * g.stoich;
      } else if (nelg == 11) {
        bs = 0.0;
      } else if (nelg == 8) {  // enthalpy: GTE, heat release, source history C/Source.f90:278-520
        double dudx, dvdx, dwdx, dudr, dvdr, dwdr, dudz, dvdz, dwdz;
        if (IBC(i + 1, j, k) == 3 || IBC(i - 1, j, k) == 3) { dudx = dvdx = dwdx = 0.0; }
        else {
          dudx = (UG(i + 1, j, k, 1) - UG(i - 1, j, k, 1)) / DX(i);
          dvdx = (UG(i + 1, j, k, 2) - UG(i - 1, j, k, 2)) / DX(i);
          dwdx = (UG(i + 1, j, k, 3) - UG(i - 1, j, k, 3)) / DX(i);
        }
        if (IBC(i, j + 1, k) == 3 || IBC(i, j - 1, k) == 3) { dudr = dvdr = dwdr = 0.0; }
        else {
          dudr = (UG(i, j + 1, k, 1) - UG(i, j - 1, k, 1)) / DR(j);
          dvdr = (UG(i, j + 1, k, 2) - UG(i, j - 1, k, 2)) / DR(j);
          dwdr = (UG(i, j + 1, k, 3) - UG(i, j - 1, k, 3)) / DR(j);
        }
        if (IBC(i, j, k + 1) == 3 || IBC(i, j, k - 1) == 3) { dudz = dvdz = dwdz = 0.0; }
        else {
          dudz = (UG(i, j, k + 1, 1) - UG(i, j, k - 1, 1)) / DZ(k);
          dvdz = (UG(i, j, k + 1, 2) - UG(i, j, k - 1, 2)) / DZ(k);
          dwdz = (UG(i, j, k + 1, 3) - UG(i, j, k - 1, 3)) / DZ(k);
        }
        g.gte[q] = 2.0 * (dudx * dudx + dvdr * dvdr + 0.0 * 0.0 + dwdz * dwdz) + (dudr + dvdx) * (dudr + dvdx) +
                   (dwdr + dvdz) * (dwdr + dvdz) + (dudz + dwdx) * (dudz + dwdx) - (dudx + dvdr + dwdz) * (dudx + dvdr + dwdz) / 1.5;
        double hr = g.hrate[q];
        bs = 0.0 + hr;
        sfp = -0.0;
        double vol0 = VOL_C(i, j, k) * g.vol0;
        qr = hr * g.qr0 * vol0;
        if (g.gray_radiation == GFM_RADSRC_GRAY) {
          bs += GRAY_QNET(i, j, k) / (vol0 * g.qr0);
        } else if (g.gray_radiation == GFM_RADSRC_FIELD) {  // q_rad = QA - QE, C/Source.f90:437-447
          const long long cc = (long long)(i / 2 - 1) + (long long)(MP / 2) * ((j / 2 - 1) + (long long)(NP / 2) * (k / 2 - 1));
          bs += (g.rad_qa[cc] - g.rad_qe[cc]) / (vol0 * g.qr0);
        }
        if (g.isothermal_walls) bs -= WALL_SINK(i, j, k, qw, qlo, qhi);
        else if (g.i1st_rad_done > 0 && g.wall_qcf) {
          // the patches of the last wall exchange, in patch order -x, +x, -y, +y, -z, +z (0 where none)
          bs = bs - g.wall_qcf[lidx(i - 1, j, k)] / (DX(i) * g.r0 * g.qr0);
          bs = bs - g.wall_qcf[lidx(i + 1, j, k)] / (DX(i) * g.r0 * g.qr0);
          bs = bs - g.wall_qcf[lidx(i, j - 1, k)] / (DR(j) * g.r0 * g.qr0);
          bs = bs - g.wall_qcf[lidx(i, j + 1, k)] / (DR(j) * g.r0 * g.qr0);
          bs = bs - g.wall_qcf[lidx(i, j, k - 1)] / (DZ(k) * g.r0 * g.qr0);
          bs = bs - g.wall_qcf[lidx(i, j, k + 1)] / (DZ(k) * g.r0 * g.qr0);
        }
        if (itr_gas > 1) bs = g.rf_hsource * bs + (1.0 - g.rf_hsource) * g.bsh[q];
        g.bsh[q] = bs;
      } else if (nelg == 9) {  // k, C/Source.f90:529-572
        bs = 2.0 / (g.reyg * g.epsr) * g.theta[q] * g.tmu[q] * g.gte[q] - 0.0;
        sfp = -g.taur * g.theta[q] * g.dnst[q] * g.gf[4][q] / (g.gf[3][q] + SMALL20);
      } else if (nelg == 10) {  // eps, C/Source.f90:579-618
        double cbl0 = 2.5 * sqrt(g.epsr) / g.taur * pow(CMU, 0.75);
        double kk = g.gf[3][q];
        double cbl = cbl0 * pow(kk, 1.5);
        bs = 0.0;
        if (MP > 8 && IBC(i + 1, j, k) == 1) { bs += BIGV * cbl / (X(i + 1) - X(i)); sfp = -BIGV; }
        if (MP > 8 && IBC(i - 1, j, k) == 1) { bs += BIGV * cbl / (X(i) - X(i - 1)); sfp = -BIGV; }
        if (NP > 8 && IBC(i, j + 1, k) == 1) { bs += BIGV * cbl / (R(j + 1) - R(j)); sfp = -BIGV; }
        if (NP > 8 && IBC(i, j - 1, k) == 1) { bs += BIGV * cbl / (R(j) - R(j - 1)); sfp = -BIGV; }
        if (LP > 8 && IBC(i, j, k + 1) == 1) { bs += BIGV * cbl / (Z(k + 1) - Z(k)); sfp = -BIGV; }
        if (LP > 8 && IBC(i, j, k - 1) == 1) { bs += BIGV * cbl / (Z(k) - Z(k - 1)); sfp = -BIGV; }
        if (!(sfp <= -BIGV)) {
          double bt = g.theta[q];
          bs = CT1 * CMU * bt / g.taur * g.dnst[q] * g.gte[q] * kk;
          sfp = -CT2 * g.taur * bt * g.dnst[q] * g.gf[4][q] / (kk + SMALL20);
        }
      }
      // GSLV3G0 coefficients, C/Gslv3g.f90:123-152
      double ap = 0.0;
      for (int iu = 1; iu <= 3; iu++)
        for (int m = -1; m <= 1; m += 2) {
          int m1 = (m + 3) / 2;
          const int asl = (iu - 1) + 3 * (m1 - 1);
          double av;
          if (igs != 1) {
            int i1, j1, k1, i2, j2, k2;
            VINDX(iu, m, i, j, k, i1, j1, k1, i2, j2, k2);
            int ibc0 = IBC(i2, j2, k2);
            double a;
            if (ibc0 <= 0) a = FLWD(i1, j1, k1, iu, m1, sgm);
            else if (ibc0 == 2) a = FLW(i2, j2, k2, iu, m1);
            else a = 0.0;
            SAS(asl, q, a);
            av = g.mixed ? (double)(float)a : a;
          } else {
            av = CAS(asl, q);
          }
          ap += av;
        }
      // volume and under-relaxation, C/Gslv3g.f90:64-87
      if (fabs(ap) <= SMALL20) {
        ap = 1.0;
        bs = fz;
      } else {
        double vol = VOL_C(i, j, k);
        bs *= vol;
        ap = ap - (sfp * vol);
        ap = ap / g.rfl[l - 1];
        bs = bs + (1.0 - g.rfl[l - 1]) * fz * ap;
      }
      SAP(q, ap);
      SBS(q, bs);
    }
  }
  if (nelg == 8) {
    double v = wave_sum(qr), w = wave_sum(qw), lo = wave_sum(qlo), hi = wave_sum(qhi);
    if (threadIdx.x == 0) {
      part[block_lin()] = v;
      part[g.pstride + block_lin()] = w;
      part[2 * g.pstride + block_lin()] = lo;
      part[3 * g.pstride + block_lin()] = hi;
    }
  }
}
__global__ void k_bound_store_gf(Box b, int l) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  double f = BOUND_GF(g.fz[q], l);
  g.fz[q] = f;
  g.gf[l - 1][q] = f;
}

// property update of open cells, C/Gflow.f90:302-382
__global__ __launch_bounds__(64) void k_props(Box b) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  if (IBC(i, j, k) >= 1) return;
  long long q = lidx(i, j, k);
  CONC(q);
  ENTH_TO_T(q);
  double t = g.t[q];
  if (t > g.t_mx) { g.t[q] = g.t_mx; T_TO_ENTH(q); }
  else if (t < g.t_mn) { g.t[q] = g.t_mn; T_TO_ENTH(q); }
  g.p[q] = BOUND_P(g.p[q] - g.scal[SC_G0REF]);
  DNST_CALC(q);
  THER(q);
}
// T_ave_calc, C/Utility.f90:1485-1513 (partials: T*rho*V, rho*V, T*V, V, max T)
__global__ void k_tave(Box b, double *part) {
  int i, j, k;
  double s[5] = {0, 0, 0, 0, 0};
  if (box_pt(b, i, j, k) && IBC(i, j, k) == 0) {
    long long q = lidx(i, j, k);
    double vol = VOL_C(i, j, k), t = g.t[q], d = g.dnst[q];
    s[0] = t * d * vol; s[1] = d * vol; s[2] = t * vol; s[3] = vol; s[4] = t;
  }
  for (int n = 0; n < 4; n++) {
    double v = wave_sum(s[n]);
    if (threadIdx.x == 0) part[(long long)n * g.pstride + block_lin()] = v;
  }
  double m = wave_max(s[4]);
  if (threadIdx.x == 0) part[4LL * g.pstride + block_lin()] = m;
}
__global__ __launch_bounds__(FIN_NT) void k_tave_finalize(const double *part, int nb) {
  double r[5];
  fin_reduce<5>(part, g.pstride, nb, 1u << 4, r);
  if (threadIdx.x == 0)
    for (int n = 0; n < 5; n++) g.scal[SC_TAVE_NUM + n] = r[n];
}

// ======================================================================
// EXTR(1), C/Extr.f90:31-250
// ======================================================================
// extrapolation stencil (EXT0 mode 0 / EXT1 mode 1), C/Extr.f90:575-661.
// A cell with no qualifying neighbour (interior corner of the wall
// shell) extrapolates from itself.
__device__ int EXTIDX(int mode, int i, int j, int k, int &i1, int &j1, int &k1, int &i2, int &j2, int &k2) {
  for (int d = 1; d <= 3; d++) {
    int n0 = (d == 1), n1 = (d == 2), n2 = (d == 3);
    for (int m = -2; m <= 2; m += 4) {
      i1 = max(2, min(MP, i + m * n0)); j1 = max(2, min(NP, j + m * n1)); k1 = max(2, min(LP, k + m * n2));
      int b1 = IBC(i1, j1, k1);
      bool ok = (mode == 0) ? (b1 <= 0) : ((b1 % 10) != 1);
      if (ok) {
        i2 = max(2, min(MP, i1 + m * n0)); j2 = max(2, min(NP, j1 + m * n1)); k2 = max(2, min(LP, k1 + m * n2));
        if ((IBC(i2, j2, k2) % 10) == 1) { i2 = i1; j2 = j1; k2 = k1; }
        return d;
      }
    }
  }
  i1 = i; j1 = j; k1 = k; i2 = i; j2 = j; k2 = k;
  return 4;
}
__device__ __forceinline__ void EXTR1F(double &f1, double &f2, int i, int j, int k, int i1, int j1, int k1, int i2, int j2, int k2, int ix) {
  f2 = 0.0;
  int ibc0 = IBC(i, j, k) % 10;
  if (ix == 3) { double d = Z(k2) - Z(k1); if (d != 0.0) { int k0 = (ibc0 == 2) ? (k + k1) / 2 : k; f2 = (Z(k0) - Z(k1)) / d; } }
  else if (ix == 2) { double d = R(j2) - R(j1); if (d != 0.0) { int j0 = (ibc0 == 2) ? (j + j1) / 2 : j; f2 = (R(j0) - R(j1)) / d; } }
  else if (ix == 1) { double d = X(i2) - X(i1); if (d != 0.0) { int i0 = (ibc0 == 2) ? (i + i1) / 2 : i; f2 = (X(i0) - X(i1)) / d; } }
  f1 = 1.0 - f2;
}
// inlets and exits (partials: GFIN0, q_exhaust), over the list of the
// inlet and exit cells of the cell box [2, MP]x[2, NP]x[2, LP] (step 2);
// every other cell of that box leaves the fields unchanged.
__global__ __launch_bounds__(64) void k_extr_inout(const long long *lst, int n, double *part) {
  int i, j, k;
  double s0 = 0.0, s1 = 0.0;
  if (list_pt(lst, n, i, j, k)) {
    int ibc0 = IBC(i, j, k) % 10;
    if (ibc0 == 2 || ibc0 == 3) {
      int i1, j1, k1, i2, j2, k2;
      int ix = EXTIDX(0, i, j, k, i1, j1, k1, i2, j2, k2);
      double f1, f2;
      EXTR1F(f1, f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
      long long q = lidx(i, j, k), q1 = lidx(i1, j1, k1), q2 = lidx(i2, j2, k2);
      g.gf[4][q] = g.gf[4][q1];
      g.p[q] = g.p[q1] * f1 + g.p[q2] * f2;
      if (ibc0 == 2) {
        double t0 = g.tmu[q];
        THER(q);
        g.tmu[q] = t0;
        g.gama[q] = g.gama[q1];
        double area = (ix <= 3) ? AREA_C(i, j, k, ix) : 0.0;
        s0 = fabs(g.theta[q] * g.dnst[q] * ((ix <= 3) ? g.ug[ix - 1][q] : 0.0) * area);
      } else {
        g.gcp[q] = g.gcp[q1] * f1 + g.gcp[q2] * f2;
        g.glam[q] = g.glam[q1] * f1 + g.glam[q2] * f2;
        g.tmu[q] = g.tmu[q1];
        g.gama[q] = g.gama[q1];
        for (int l = 1; l <= 7; l++) {
          if (l == IEPS) continue;
          double v = g.gf[l - 1][q1] * f1 + g.gf[l - 1][q2] * f2;
          if (l == IK) v = fmax(v, 1.0e-6);
          else if (l != IH) v = fmax(v, 0.0);
          g.gf[l - 1][q] = v;
        }
        ENTH_TO_T(q);
        DNST_CALC(q);
        double area = AREA_C(i, j, k, ix) * g.area0;
        double g0 = fabs(g.ug[ix - 1][q]) * g.ug0 * area;
        g0 = g0 * g.theta[q] * g.dnst[q] * g.dnst0;
        s1 = g.gf[2][q] * g.h_0 * g0;
        // SBC4 onto the exit face between the cell and its neighbour
        long long d = lidx((i + i1) / 2, (j + j1) / 2, (k + k1) / 2);
        g.p[d] = g.p[q]; g.t[d] = g.t[q];
        for (int l = 0; l < 9; l++) g.gf[l][d] = g.gf[l][q];
        g.dnst[d] = g.dnst[q]; g.theta[d] = g.theta[q]; g.tmu[d] = g.tmu[q];
        g.glam[d] = g.glam[q]; g.gcp[d] = g.gcp[q]; g.gama[d] = g.gama[q];
      }
    }
  }
  double v0 = wave_sum(s0), v1 = wave_sum(s1);
  if (threadIdx.x == 0) { part[blockIdx.x] = v0; part[g.pstride + blockIdx.x] = v1; }
}
// walls, over the list of the wall cells of the same box
__global__ void k_extr_wall(const long long *lst, int n) {
  int i, j, k;
  if (!list_pt(lst, n, i, j, k)) return;
  if ((IBC(i, j, k) % 10) != 1) return;
  int i1, j1, k1, i2, j2, k2;
  int ix = EXTIDX(1, i, j, k, i1, j1, k1, i2, j2, k2);
  double f1, f2;
  EXTR1F(f1, f2, i, j, k, i1, j1, k1, i2, j2, k2, ix);
  long long q = lidx(i, j, k), q1 = lidx(i1, j1, k1), q2 = lidx(i2, j2, k2);
  g.p[q] = g.p[q1] * f1 + g.p[q2] * f2;
  g.gcp[q] = g.gcp[q1] * f1 + g.gcp[q2] * f2;
  g.glam[q] = g.glam[q1] * f1 + g.glam[q2] * f2;
  g.tmu[q] = g.tmu[q2];
  g.gama[q] = g.gama[q2];
  for (int l = 1; l <= 7; l++)
    if (l != IH) g.gf[l - 1][q] = g.gf[l - 1][q1];
  if (g.i1st_rad_done == 0 && !g.isothermal_walls) {
    g.gf[2][q] = g.gf[2][q1];
    g.t[q] = g.t[q1];
  }
  DNST_CALC(q);
}

// ======================================================================
// INTP, C/Intp.f90:32-72.  The weight of each position is the reference's
// loop-carried F1 (modified in place at inlet faces by INTP0) and is
// precomputed per position in intp_code: 0 -> INTP1 weight, 1 -> 1.0,
// 2 -> 0.0.  dir: 3 = from (k+1,k-1), 2 = (j+1,j-1), 1 = (i+1,i-1).
// ======================================================================
// The 28 neighbour values are loaded before anything is stored (the field
// arrays may alias as far as the compiler knows, so stores between them
// would serialise the loads); CONC and DNST_CALC then run on registers with
// the same operations as on memory.
__device__ __forceinline__ void conc_r(double y[9]) {  // CONC on gf values in registers, C/Prop.f90:28-59
  const double yf = y[0], yo = y[1], yc = y[5], yn = y[6];
  double g0 = yf + yo + yc + yn;
  if (g0 > 1.0) {
    g0 = g0 + y[8];
    y[0] = yf / g0; y[1] = yo / g0; y[5] = yc / g0; y[6] = yn / g0;
    y[8] = y[8] / g0;
  } else {
    y[8] = 1.0 - g0;
  }
}
// The direction DIR of the two source points is a template parameter.
// Stored are only the fields a later kernel reads at these positions: the
// density, theta and Gamma everywhere (assembly fluxes and diffusion), and
// P, T and the mass fractions on the faces and the yz-edges (sources of the
// later INTP stages; faces are read elsewhere too) but not on the xy- and
// xz-edges and corners of the x-stage, where no kernel reads them.  mu_t,
// c_p and lambda are read at cell centres only, so INTP forms none of them
// (the reference interpolates all of them everywhere).
template <int DIR>
__global__ void k_intp(Box b) {
  constexpr int dir = DIR;
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
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
  const double p1 = g.p[q1], p2 = g.p[q2], t1 = g.t[q1], t2 = g.t[q2];
  double y1[7], y2[7];
#pragma unroll
  for (int l = 0; l < 7; l++) { y1[l] = g.gf[l][q1]; y2[l] = g.gf[l][q2]; }
  const double ga1 = g.gama[q1], ga2 = g.gama[q2];
  const double th1 = g.theta[q1], th2 = g.theta[q2];
  double y[9];
  y[7] = 0.0;
  y[8] = g.gf[8][q];
  const double p = a * p1 + bb * p2;
  const double t = a * t1 + bb * t2;
#pragma unroll
  for (int l = 0; l < 7; l++) y[l] = a * y1[l] + bb * y2[l];
  conc_r(y);
  const double ga = a * ga1 + bb * ga2, th = a * th1 + bb * th2;
  // DNST_CALC, C/Prop.f90:187-198
  conc_r(y);
  const double wt = 1.0 / (y[0] / WF + y[5] / WCO2 + y[8] / WH2O + y[1] / WO2 + y[6] / WN2);
  const double temp = (1.0 + p) / t * wt / g.wt0;
  if (DIR != 1 || ((j & 1) == 0 && (k & 1) == 0)) {
    g.p[q] = p;
    g.t[q] = t;
#pragma unroll
    for (int l = 0; l < 7; l++) g.gf[l][q] = y[l];
    g.gf[8][q] = y[8];
  }
  g.gama[q] = ga;
  g.theta[q] = th;
  g.dnst[q] = fmax(temp, SMALL20);
}

// exit temperature of SPHASE, C/Sphase.f90:74-89 (partials: area, T*area, max T)
__global__ void k_exit_t(Box b, double *part) {
  int i, j, k;
  double s0 = 0.0, s1 = 0.0, mx = 0.0;
  if (box_pt(b, i, j, k) && IBC(i, j, k) == 3) {
    double g0 = AREA_C(i, j, k, 3);
    if (g0 < SMALL20) g0 = AREA_C(i, j, k, 2);
    if (g0 < SMALL20) g0 = AREA_C(i, j, k, 1);
    double t = T_(i, j, k);
    s0 = g0; s1 = t * g0; mx = t;
  }
  double v0 = wave_sum(s0), v1 = wave_sum(s1), v2 = wave_max(mx);
  if (threadIdx.x == 0) { part[block_lin()] = v0; part[g.pstride + block_lin()] = v1; part[2 * g.pstride + block_lin()] = v2; }
}
__global__ __launch_bounds__(FIN_NT) void k_exit_t_finalize(const double *part, int nb) {
  double r[3];
  fin_reduce<3>(part, g.pstride, nb, 1u << 2, r);
  if (threadIdx.x == 0)
    for (int n = 0; n < 3; n++) g.scal[SC_TEXIT_AREA + n] = r[n];
}

// ======================================================================
// residual of an assembled system, calc_residual C/Adlbl3.f90:239-271
// (partials of the normalised sum) and the IRC block C/Adlbl3.f90:60-94
// (partials of sum and max of |r|).
// ======================================================================
__global__ void k_residual(Box b, const double *xt, double *part, int irc) {
  int i, j, k;
  double s = 0.0, s2 = 0.0, mx = 0.0;
  if (box_pt(b, i, j, k) && IBC(i, j, k) <= 0) {
    long long q = lidx(i, j, k);
    double t1 = CAS(0, q) * xt[lidx(i - 2, j, k)], t2 = CAS(3, q) * xt[lidx(i + 2, j, k)];
    double t3 = CAS(1, q) * xt[lidx(i, j - 2, k)], t4 = CAS(4, q) * xt[lidx(i, j + 2, k)];
    double t5 = CAS(2, q) * xt[lidx(i, j, k - 2)], t6 = CAS(5, q) * xt[lidx(i, j, k + 2)];
    const double apq = CAP(q);
    double t7 = apq * xt[q], t8 = CBS(q);
    if (!irc) {
      double eq_max = fmax(fmax(fmax(fabs(t1), fabs(t2)), fmax(fabs(t3), fabs(t4))), fmax(fmax(fabs(t5), fabs(t6)), fmax(fabs(t7), fabs(t8))));
      s = fabs(t1 + t2 + t3 + t4 + t5 + t6 + t8 - t7) / eq_max;
    } else if (apq != 1.0) {
      double r = fabs(t2 + t1 + t4 + t3 + t6 + t5 + t8 - t7);
      s2 = r; mx = r;
    }
  }
  if (!irc) {
    double v = wave_sum(s);
    if (threadIdx.x == 0) part[block_lin()] = v;
  } else {
    double v = wave_sum(s2), m = wave_max(mx);
    if (threadIdx.x == 0) { part[block_lin()] = v; part[g.pstride + block_lin()] = m; }
  }
}
__global__ __launch_bounds__(FIN_NT) void k_residual_finalize(const double *part, int nb, int slot, int irc) {
  if (!irc) {
    double r[1];
    fin_reduce<1>(part, 0, nb, 0u, r);
    if (threadIdx.x == 0) g.scal[slot] = fmax(r[0] / g.ncells, 1.0e-16);
  } else {
    double r[2];
    fin_reduce<2>(part, g.pstride, nb, 2u, r);
    if (threadIdx.x == 0) { g.scal[slot] = r[0] / g.ncells; g.scal[slot + 1] = r[1]; }
  }
}

// ======================================================================
// Additive correction over x-planes (block correction; Settari and Aziz
// 1973, Patankar 1981 sec. 5.4) before each sweep of a p' solve relaxed to
// a tolerance.  No counterpart in the reference.  Plane q of the box
// (i = i0 + 2q) gets one value d_q, added to its free cells (in the box,
// IBCELL <= 0, |AP| > 1e-20), with
//   A_q d_q - W_q d_{q-1} - E_q d_{q+1} = R_q,
//   A_q = sum (a_p - a_nb of the free in-plane neighbours),
//   W_q, E_q = sum of a_x- (a_x+) toward a free neighbour,
//   R_q = sum of the residuals a_nb x_nb + b - a_p x,
// sums over the plane's free cells, so that afterwards the residuals of
// each plane's free cells sum to zero, for every plane but the last, the
// reference plane (d = 0, its equation not solved): the p' equation has no
// coefficient toward a pinned cell and is singular, fixed up to a constant
// (with coefficients rounded to FP32 only nearly so, and solving the full
// system would add a large constant).  Four launches, all in a fixed
// order: per-plane partial sums over chunks of BC_ROWS (j,k) rows, the
// chunk sums, the Thomas solve, the update.
// ======================================================================
#define BC_ROWS 16  // (j,k) rows per workgroup of k_bc_part: one per wavefront
#define BC_NV 5     // A, W, E, R, free-cell count
__device__ __forceinline__ bool bc_free(const Box &b, int i, int j, int k) {
  const int ti = (i - b.i0) >> 1, tj = (j - b.j0) >> 1, tk = (k - b.k0) >> 1;
  if (i < b.i0 || ti >= b.ni || j < b.j0 || tj >= b.nj || k < b.k0 || tk >= b.nk) return false;
  const long long q = lidx(i, j, k);
  return g.ib[q] <= 0 && fabs(CAP(q)) > SMALL20;
}
// grid (ceil(ni/64), ceil(nj*nk/BC_ROWS)), 64*BC_ROWS threads: lane = plane,
// wavefront w = row of the chunk, one cell per thread; the rows' sums are
// added in row order.  part[(c*BC_NV + v)*ni + q]
__global__ __launch_bounds__(64 * BC_ROWS) void k_bc_part(Box b, const double *xt, double *part) {
  __shared__ double red[BC_ROWS][BC_NV][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const int r = blockIdx.y * BC_ROWS + w;
  double acc[BC_NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (q < b.ni && r < b.nj * b.nk) {
    const int i = b.i0 + 2 * q, j = b.j0 + 2 * (r % b.nj), k = b.k0 + 2 * (r / b.nj);
    if (bc_free(b, i, j, k)) {
      const long long c = lidx(i, j, k);
      const double a0 = CAS(0, c), a3 = CAS(3, c), a1 = CAS(1, c), a4 = CAS(4, c), a2 = CAS(2, c), a5 = CAS(5, c);
      const double t1 = a0 * xt[lidx(i - 2, j, k)], t2 = a3 * xt[lidx(i + 2, j, k)];
      const double t3 = a1 * xt[lidx(i, j - 2, k)], t4 = a4 * xt[lidx(i, j + 2, k)];
      const double t5 = a2 * xt[lidx(i, j, k - 2)], t6 = a5 * xt[lidx(i, j, k + 2)];
      const double apc = CAP(c), t7 = apc * xt[c], t8 = CBS(c);
      acc[3] = t1 + t2 + t3 + t4 + t5 + t6 + t8 - t7;
      double a = apc;
      if (bc_free(b, i, j - 2, k)) a -= a1;
      if (bc_free(b, i, j + 2, k)) a -= a4;
      if (bc_free(b, i, j, k - 2)) a -= a2;
      if (bc_free(b, i, j, k + 2)) a -= a5;
      acc[0] = a;
      if (bc_free(b, i - 2, j, k)) acc[1] = a0;
      if (bc_free(b, i + 2, j, k)) acc[2] = a3;
      acc[4] = 1.0;
    }
  }
#pragma unroll
  for (int v = 0; v < BC_NV; v++) red[w][v][lane] = acc[v];
  __syncthreads();
  if (w == 0 && q < b.ni)
#pragma unroll
    for (int v = 0; v < BC_NV; v++) {
      double t = red[0][v][lane];
      for (int rr = 1; rr < BC_ROWS; rr++) t += red[rr][v][lane];
      part[((long long)blockIdx.y * BC_NV + v) * b.ni + q] = t;
    }
}
// grid (ceil(ni/64), BC_NV), 1024 threads: plane q = 64*bx + (t mod 64),
// quantity v = by, slice s = t / 64 of the chunks (c = s, s+16, ... with two
// accumulators), the 16 slices added in order: sums[v*ni + q].
__global__ __launch_bounds__(1024) void k_bc_sum(int ni, int nchunk, const double *part, double *sums) {
  __shared__ double red[16][64];
  const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6, v = blockIdx.y;
  const int q = blockIdx.x * 64 + lane;
  double s0 = 0.0, s1 = 0.0;
  if (q < ni) {
    int c = sl;
    for (; c + 16 < nchunk; c += 32) {
      s0 += part[((long long)c * BC_NV + v) * ni + q];
      s1 += part[((long long)(c + 16) * BC_NV + v) * ni + q];
    }
    if (c < nchunk) s0 += part[((long long)c * BC_NV + v) * ni + q];
  }
  red[sl][lane] = s0 + s1;
  __syncthreads();
  if (sl == 0 && q < ni) {
    double t = red[0][lane];
    for (int r = 1; r < 16; r++) t += red[r][lane];
    sums[(long long)v * ni + q] = t;
  }
}
// one workgroup of 256 threads, dynamic LDS BC_NV*ni doubles: the sums
// (A, W, E, R, count) staged in LDS, the Thomas solve over the planes by
// thread 0 there, d[ni] written by all.  A plane without free cells, the
// last plane, and a plane whose pivot falls below 1e-12 of its diagonal
// (closing a block of planes decoupled from the rest) get 0.
__global__ __launch_bounds__(256) void k_bc_solve(int ni, const double *sums, double *d) {
  extern __shared__ double bcs[];
  for (int t = threadIdx.x; t < BC_NV * ni; t += 256) bcs[t] = sums[t];
  __syncthreads();
  double *A = bcs, *W = bcs + ni, *E = bcs + 2 * ni, *R = bcs + 3 * ni, *nf = bcs + 4 * ni;
  double *cp = nf;  // the counts are read before cp overwrites them
  if (threadIdx.x == 0) {
    // c' and d' of the previous plane stay in registers (the chain of the
    // elimination), the plane's sums are independent loads
    double cpm = 0.0, dpm = 0.0;
    for (int q = 0; q < ni; q++) {
      double a = A[q], w = W[q], e = E[q], r = R[q];
      if (nf[q] == 0.0) { a = 1.0; w = e = r = 0.0; }
      const double m = a + (q > 0 ? w * cpm : 0.0);
      double cq = 0.0, dq = 0.0;
      if (q < ni - 1 && fabs(m) > 1.0e-12 * fabs(a)) {
        cq = -e / m;
        dq = (r + (q > 0 ? w * dpm : 0.0)) / m;
      }
      cp[q] = cq;
      R[q] = dq;
      cpm = cq;
      dpm = dq;
    }
    double dn = R[ni - 1];
    for (int q = ni - 2; q >= 0; q--) {
      dn = R[q] - cp[q] * dn;
      R[q] = dn;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ni; q += 256) d[q] = R[q];
}
// ======================================================================
// FLX1D, C/Flx1d.f90:16-73: per face plane, the gas mass flow through it
// in the + and - direction (and for y and z the enthalpy flow), from
// theta*rho*u*AREA_C*GMFR of each face whose cell code is not 1 or >= 4.
// Partials part[term*nitem + item] with item = plane*nv + value, summed in
// term order by k_flow_sum.
// ======================================================================
__device__ __forceinline__ double flow_g0(int i, int j, int k, int d) {
  const int ib = IBC(i, j, k);
  if (ib == 1 || ib >= 4) return 0.0;
  const long long q = lidx(i, j, k);
  return g.theta[q] * g.dnst[q] * g.ug[d - 1][q] * AREA_C(i, j, k, d) * g.gmfr;
}
#define FLOW_ROWS 64  // (j,k) rows per workgroup of k_flow_x, 4 per wavefront
// x-face planes I = b.i0 + 2q (lane q), rows (J, K) of the box; 1,024
// threads, the 16 wavefronts' sums added in order; item = 2q + (-/+).
__global__ __launch_bounds__(1024) void k_flow_x(Box b, double *part) {
  __shared__ double red[16][2][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const int nrow = b.nj * b.nk, r0 = blockIdx.y * FLOW_ROWS;
  double gp = 0.0, gm = 0.0;
  if (q < b.ni) {
    const int i = b.i0 + 2 * q;
    for (int r = r0 + w; r < r0 + FLOW_ROWS && r < nrow; r += 16) {
      const double g0 = flow_g0(i, b.j0 + 2 * (r % b.nj), b.k0 + 2 * (r / b.nj), 1);
      if (g0 >= 0.0) gp += g0; else gm -= g0;
    }
  }
  red[w][0][lane] = gp;
  red[w][1][lane] = gm;
  __syncthreads();
  if (w == 0 && q < b.ni)
#pragma unroll
    for (int v = 0; v < 2; v++) {
      double t = red[0][v][lane];
      for (int r = 1; r < 16; r++) t += red[r][v][lane];
      part[(long long)blockIdx.y * 2 * b.ni + 2 * q + v] = t;
    }
}
// y-face (DIR 2) or z-face (DIR 3) planes: one wavefront per (x chunk, j,
// k) of the box, its 64 faces along x summed; the plane is blockIdx.y (y)
// or blockIdx.z (z), the term (row, x chunk) the other; item = 4*plane +
// (mass +, mass -, enthalpy +, enthalpy -).
template <int DIR>
__global__ void k_flow_yz(Box b, double *part) {
  int i, j, k;
  double g0 = 0.0;
  if (box_pt(b, i, j, k)) g0 = flow_g0(i, j, k, DIR);
  const double h0 = g0 != 0.0 ? g0 * g.gf[IH - 1][lidx(i, j, k)] * g.h_0 : 0.0;
  const double v0 = wave_sum(g0 >= 0.0 ? g0 : 0.0), v1 = wave_sum(g0 < 0.0 ? -g0 : 0.0);
  const double v2 = wave_sum(g0 >= 0.0 ? h0 : 0.0), v3 = wave_sum(g0 < 0.0 ? -h0 : 0.0);
  if (threadIdx.x != 0) return;
  const int plane = DIR == 2 ? blockIdx.y : blockIdx.z, row = DIR == 2 ? blockIdx.z : blockIdx.y;
  const int nplane = DIR == 2 ? b.nj : b.nk;
  double *o = part + ((long long)row * gridDim.x + blockIdx.x) * 4 * nplane + 4 * plane;
  o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
}
// dst[item] = sum over terms of part[term*nitem + item], in term order per
// slice of 16; grid ceil(nitem/64), 1,024 threads.
__global__ __launch_bounds__(1024) void k_flow_sum(int nterm, int nitem, const double *part, double *dst) {
  __shared__ double red[16][64];
  const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int it = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (it < nitem)
    for (int t = sl; t < nterm; t += 16) s += part[(long long)t * nitem + it];
  red[sl][lane] = s;
  __syncthreads();
  if (sl == 0 && it < nitem) {
    double t = red[0][lane];
    for (int r = 1; r < 16; r++) t += red[r][lane];
    dst[it] = t;
  }
}

__global__ void k_bc_apply(Box b, double *xt, const double *d) {
  int i, j, k;
  if (!box_pt(b, i, j, k) || !bc_free(b, i, j, k)) return;
  xt[lidx(i, j, k)] += d[(i - b.i0) >> 1];
}

// ======================================================================
// Learned initial guess of the p' solve.  No counterpart in the reference,
// whose guess is FZ = 0 (C/Gslv3v.f90:157-161).  Cell v of the p' box
// [4, MP-2]x[4, NP-2]x[4, LP-2] (step 2) is v = (kk*ny + jj)*nx + ii with
// ii = (i-4)/2, jj = (j-4)/2, kk = (k-4)/2; a row is open when IBCELL <= 0
// and |AP| > 1e-20 (the rows the line solver does not pin).
// ======================================================================
__device__ __forceinline__ long long pg_cell(int i, int j, int k) {
  const int nx = (MP - 6) / 2 + 1, ny = (NP - 6) / 2 + 1;
  return ((long long)((k - 4) >> 1) * ny + ((j - 4) >> 1)) * nx + ((i - 4) >> 1);
}
__device__ __forceinline__ bool pg_open(int i, int j, int k, long long q) { return IBC(i, j, k) <= 0 && fabs(CAP(q)) > SMALL20; }
// network input x[2v] = BS/AP, x[2v+1] = open; partial sums of (BS/AP)^2
__global__ void k_pg_gather(Box b, float *x, double *part) {
  int i, j, k;
  double s2 = 0.0;
  if (box_pt(b, i, j, k)) {
    const long long q = lidx(i, j, k), v = pg_cell(i, j, k);
    const bool op = pg_open(i, j, k, q);
    const double s = op ? CBS(q) / CAP(q) : 0.0;
    x[2 * v] = (float)s;
    x[2 * v + 1] = op ? 1.f : 0.f;
    s2 = s * s;
  }
  const double w = wave_sum(s2);
  if (threadIdx.x == 0) part[block_lin()] = w;
}
// scale of the input: root mean square of BS/AP over open rows
__global__ __launch_bounds__(FIN_NT) void k_pg_rms(const double *part, int nb) {
  double r[1];
  fin_reduce<1>(part, 0, nb, 0u, r);
  if (threadIdx.x == 0) {
    const double rr = sqrt(r[0] / fmax((double)g.ncells, 1.0));
    g.scal[SC_PG_RMS] = rr > 0.0 ? rr : 1.0;
    g.scal[SC_PG_ALPHA] = 0.0;
    g.scal[SC_PG_FALLBACK] = 0.0;
    g.scal[SC_PG_COS2] = 0.0;
  }
}
__global__ void k_pg_norm(float *x, long long V) {
  const long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v < V) x[2 * v] = (float)((double)x[2 * v] / g.scal[SC_PG_RMS]);
}
// FZ = network output on open rows (FZ is zero elsewhere)
__global__ void k_pg_scatter(Box b, const float *y) {
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  const long long q = lidx(i, j, k);
  if (pg_open(i, j, k, q)) g.fz[q] = (double)y[pg_cell(i, j, k)];
}
// partial sums of BS.(A g), (A g).(A g) and BS.BS, (A g) = AP g - sum AS g_nb, FP64
__global__ void k_pg_ag(Box b, double *part) {
  int i, j, k;
  double bag = 0.0, agag = 0.0, bb = 0.0;
  if (box_pt(b, i, j, k)) {
    const long long q = lidx(i, j, k);
    if (pg_open(i, j, k, q)) {
      const double ag = CAP(q) * FZ(i, j, k) - (CAS(0, q) * FZ(i - 2, j, k) + CAS(3, q) * FZ(i + 2, j, k) +
                                                CAS(1, q) * FZ(i, j - 2, k) + CAS(4, q) * FZ(i, j + 2, k) +
                                                CAS(2, q) * FZ(i, j, k - 2) + CAS(5, q) * FZ(i, j, k + 2));
      bag = CBS(q) * ag;
      agag = ag * ag;
      bb = CBS(q) * CBS(q);
    }
  }
  const double w1 = wave_sum(bag), w2 = wave_sum(agag), w3 = wave_sum(bb);
  if (threadIdx.x == 0) {
    part[block_lin()] = w1;
    part[g.pstride + block_lin()] = w2;
    part[2 * g.pstride + block_lin()] = w3;
  }
}
// alpha = BS.(Ag) / (Ag).(Ag), which leaves ||BS||^2 (1 - cos^2) of the
// residual, cos^2 = (BS.Ag)^2 / (||BS||^2 ||Ag||^2) (kept for diagnostics).
__global__ __launch_bounds__(FIN_NT) void k_pg_alpha(const double *part, int nb) {
  double r[3];
  fin_reduce<3>(part, g.pstride, nb, 0u, r);
  if (threadIdx.x == 0) {
    const double bag = r[0], agag = r[1], bb = r[2];
    const double cos2 = (agag > 0.0 && bb > 0.0) ? bag * bag / (agag * bb) : 0.0;
    g.scal[SC_PG_COS2] = cos2;
    g.scal[SC_PG_ALPHA] = agag > 0.0 ? bag / agag : 0.0;
  }
}
__global__ void k_pg_scale(Box b) {
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  const long long q = lidx(i, j, k);
  g.fz[q] *= g.scal[SC_PG_ALPHA];
}
// keep the guess only if its calc_residual is below `accept` times that of
// FZ = 0 (accept < 1 demands a margin: a prediction that leaves the
// residual nearly as large as zero does can carry wrong long-wave content,
// which the few sweeps of a loose tolerance do not remove)
__global__ void k_pg_fallback(Box b, int slot_guess, int slot_zero, double accept) {
  int i, j, k;
  const bool worse = !(g.scal[SC_RESID0 + slot_guess] < accept * g.scal[SC_RESID0 + slot_zero]);
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) g.scal[SC_PG_FALLBACK] = worse ? 1.0 : 0.0;
  if (!worse || !box_pt(b, i, j, k)) return;
  g.fz[lidx(i, j, k)] = 0.0;
}
// training pair of the last solve: p' on every cell of the p' box
__global__ void k_pg_dump_p(Box b, float *p) {
  int i, j, k;
  if (!box_pt(b, i, j, k)) return;
  p[pg_cell(i, j, k)] = (float)FZ(i, j, k);
}

// ======================================================================
// ADLBL3 line relaxation, C/Adlbl3.f90:10-221.  Each y-line (along J) or
// z-line (along K) of an x-plane is a tridiagonal system solved by
// parallel cyclic reduction.  Rows that the reference's TDMA pins (blocked
// cells, zero pivot) become identity rows carrying the current value; the
// neighbour values along x and along the other in-plane direction come
// from the current XT, as in ADLZY/ADLYZ.
// ======================================================================
#define PCR_MAXN 256

// XT = (XT2 + XT3)/2 on the planes of one sweep step, C/Adlbl3.f90:134-138.
// Consecutive threads take consecutive planes of one colour (contiguous).
template <typename T>
__global__ void k_line_avg(T *xt, const T *x2, const T *x3, int pi0, int pdi, int npl, int jb, int nj, int kb, int nk) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long tot = (long long)npl * nj * nk;
  if (t >= tot) return;
  int pl = (int)(t % npl);
  long long r = t / npl;
  int jj = (int)(r % nj), kk = (int)(r / nj);
  int i = pi0 + pdi * pl, j = jb + 2 * jj, k = kb + 2 * kk;
  long long q = lidx(i, j, k);
  xt[q] = (x2[q] + x3[q]) / (T)2;
}

// The same for two adjacent planes per thread (zebra order: planes of one
// colour are adjacent in memory); the first plane's offset is even, so
// the pair is one aligned 2-element vector.
template <typename T>
__global__ void k_line_avg2(T *xt, const T *x2, const T *x3, int pi0, int pdi, int npl, int jb, int nj, int kb, int nk) {
  typedef T T2 __attribute__((ext_vector_type(2)));
  const int nph = (npl + 1) / 2;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long tot = (long long)nph * nj * nk;
  if (t >= tot) return;
  int ph = (int)(t % nph);
  long long r = t / nph;
  int jj = (int)(r % nj), kk = (int)(r / nj);
  int pl = 2 * ph;
  int i = pi0 + pdi * pl, j = jb + 2 * jj, k = kb + 2 * kk;
  long long q = lidx(i, j, k);
  if (pl + 1 < npl) {
    const T2 a = *reinterpret_cast<const T2 *>(x2 + q), b = *reinterpret_cast<const T2 *>(x3 + q);
    *reinterpret_cast<T2 *>(xt + q) = (a + b) / (T)2;
  } else {
    xt[q] = (x2[q] + x3[q]) / (T)2;
  }
}

// Copies xt on the x-planes i = ib-2 and ie+2 (box rows j, k) into the four
// arrays of the two pairs, so that a pair's average there is xt's value.
template <typename T>
__global__ void k_pair_xbound(const T *xt, T *x2, T *x3, T *x4, T *x5, int ib, int ie, int jb, int nj, int kb, int nk) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 2LL * nj * nk) return;
  const int side = (int)(t & 1);
  const long long r = t >> 1;
  const int j = jb + 2 * (int)(r % nj), k = kb + 2 * (int)(r / nj);
  const long long q = lidx(side ? ie + 2 : ib - 2, j, k);
  const T v = xt[q];
  x2[q] = v; x3[q] = v; x4[q] = v; x5[q] = v;
}

// FP32 working copy of an unknown over a box (mixed precision) and back.
// The copy back skips the rows the line kernel pins (cell code >= 1 or
// |a_p| <= 1e-20, the same test on the same FP32 a_p), which keep their FP64
// values.
__global__ void k_box_narrow(Box b, const double *src, float *dst) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  dst[q] = (float)src[q];
}
__global__ void k_box_widen(Box b, const float *src, double *dst) {
  int i, j, k;
  if (!box_at(b, i, j, k)) return;
  long long q = lidx(i, j, k);
  if (g.ib[q] >= 1 || fabs((double)g.apf[q]) <= SMALL20) return;
  dst[q] = (double)src[q];
}

// Line relaxation, tiled.  A 512-thread workgroup takes LW lines that sit
// at the same (k) -- y-lines -- or the same (j) -- z-lines -- on LW
// consecutive planes of one colour, which are adjacent in memory (16
// planes: one 128-byte line of each FP64 row).  The workgroup loads the
// rows of the tile cooperatively, every load of a row issued before any is
// used, forms the tridiagonal rows of ADLZY/ADLYZ (C/Adlbl3.f90:147-221)
// and stores them in LDS normalised to a unit diagonal, [line][row] with an
// odd row pitch.  The lines are then solved -- with M > 0 by the partition
// solve, four lines per wavefront, 16 lanes of M contiguous rows per line
// (line_partition_solve); with M = 0 by parallel cyclic reduction, two lines
// per wavefront, 32 lanes and up to NU rows per lane each, with
// wavefront-scope synchronisation only -- and the tile is written back
// cooperatively.  A workgroup does y-lines into xy or z-lines into xz.  T
// is the arithmetic and storage type of the sweep (FP64, or FP32 in mixed
// precision, where xt is the FP32 working copy).
template <typename V>
__device__ __forceinline__ V ld_off(const V *base, unsigned byte_off) {
  return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ const T *coef_ap_base() {
  if constexpr (sizeof(T) == 4) return g.apf; else return g.ap;
}
template <typename T>
__device__ __forceinline__ const T *coef_bs_base() {
  if constexpr (sizeof(T) == 4) return g.bsf; else return g.bs;
}
template <typename T>
__device__ __forceinline__ const T *coef_as_base(int s) {
  if constexpr (sizeof(T) == 4) return g.asf[s]; else return g.as[s];
}
// 1/x: the hardware reciprocal, refined by two Newton-Raphson steps in
// FP64 (error below 2 ulp); the hardware FP32 reciprocal (1 ulp) as is.
template <typename T> __device__ __forceinline__ T line_rcp(T x);
template <> __device__ __forceinline__ double line_rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
template <> __device__ __forceinline__ float line_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Orders the LDS accesses of the lanes of one wavefront around the point
// (all lanes must reach it together).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T> struct LinePair { T a, c; };

// One PCR step at a stride s = 32*DU >= 32: the rows e-s and e+s of row
// e = l32 + 32u are this lane's rows u-DU and u+DU, so the step needs no
// exchange between lanes.
template <typename T, int NU, int DU>
__device__ __forceinline__ void pcr_lane_step(T (&ra)[NU], T (&rc)[NU], T (&rd)[NU], int l32, int n, bool act) {
  T na[NU], nc[NU], nd[NU];
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int e = l32 + 32 * u;
    na[u] = ra[u]; nc[u] = rc[u]; nd[u] = rd[u];
    if (!act || e >= n) continue;
    T am = 0, cm = 0, dm = 0, apl = 0, cp = 0, dp = 0;
    if (u - DU >= 0) { am = ra[u - DU]; cm = rc[u - DU]; dm = rd[u - DU]; }
    if (u + DU < NU && e + 32 * DU < n) { apl = ra[u + DU]; cp = rc[u + DU]; dp = rd[u + DU]; }
    const T r = line_rcp<T>((T)1 - ra[u] * cm - rc[u] * apl);
    na[u] = -ra[u] * am * r; nc[u] = -rc[u] * cp * r; nd[u] = (rd[u] - ra[u] * dm - rc[u] * dp) * r;
  }
#pragma unroll
  for (int u = 0; u < NU; u++) { ra[u] = na[u]; rc[u] = nc[u]; rd[u] = nd[u]; }
}

// Value of lane l-K (K > 0) or l+|K| (K < 0) of the same 16-lane DPP row,
// zero where that lane lies outside the row.  Every lane of the wavefront
// must execute it.
template <int K, typename T>
__device__ __forceinline__ T row_shift(T v) {
  constexpr int ctrl = K > 0 ? 0x110 + K : 0x100 - K;  // row_shr:K / row_shl:-K
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xf, 0xf, false));
  } else {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(T, (long long)(unsigned)lo | ((long long)hi << 32));
  }
}

// Row r of a unit-diagonal tridiagonal system, a x_{r-S} + x_r + c x_{r+S} = d.
template <typename T> struct TriRow { T a, c, d; };

// One PCR elimination: row r with its neighbours at -S and +S
// (zero rows where the neighbour does not exist), renormalised.
template <typename T>
__device__ __forceinline__ TriRow<T> pcr_combine(const TriRow<T> &r, const TriRow<T> &m, const TriRow<T> &p) {
  const T q = line_rcp<T>((T)1 - r.a * m.c - r.c * p.a);
  return TriRow<T>{-r.a * m.a * q, -r.c * p.c * q, (r.d - r.a * m.d - r.c * p.d) * q};
}

template <int K, typename T>
__device__ __forceinline__ TriRow<T> row_shift3(const TriRow<T> &v) {
  return TriRow<T>{row_shift<K>(v.a), row_shift<K>(v.c), row_shift<K>(v.d)};
}

// Reduced-system PCR step at lane stride K >= 1 (reduced-row stride 2K):
// each lane's first and last rows combine with the same rows of lanes L-K
// and L+K.
template <int K, typename T>
__device__ __forceinline__ void reduced_step(TriRow<T> &rs, TriRow<T> &re) {
  {
    const TriRow<T> sm = row_shift3<K>(rs), sp = row_shift3<-K>(rs);
    rs = pcr_combine(rs, sm, sp);
  }
  const TriRow<T> em = row_shift3<K>(re), ep = row_shift3<-K>(re);
  re = pcr_combine(re, em, ep);
}

// Solve of one line of up to 16*M rows (M >= 2) by 16 lanes of one DPP row,
// lane L owning the contiguous rows s = L*M .. e = s+M-1 (unit-diagonal
// rows a x_{i-1} + x_i + c x_{i+1} = d, read from the lane's LDS rows acl,
// dl, of which the first nrow exist; the others are zero and decouple,
// since the last row of a line has c = 0).
//  1. Forward elimination inside the lane: rows s+1..e as
//     x_i + al_i x_s + ga_i x_{i+1} = de_i; row e is the lane's last
//     reduced row, x_e + al_e x_s + ga_e x_{e+1} = de_e.
//  2. x_{s+1} = p - q x_s - t x_e by back substitution of those rows;
//     with row s it gives the first reduced row, x_s + A x_{s-1} + C x_e = D.
//  3. The 32 reduced rows (s_0, e_0, s_1, e_1, ...) form a tridiagonal
//     system, solved by parallel cyclic reduction across the lanes with
//     DPP row shifts (reduced strides 1, 2, 4, 8, 16).
//  4. Rows s+1..e-1 from step 1's relations, from e-1 down.
// The solution is written to dl (existing rows).  Every lane of the
// wavefront must call it.
template <typename T, int M>
__device__ __forceinline__ void line_partition_solve(LinePair<T> *acl, T *dl, int nrow) {
  T al[M], ga[M], de[M];
  T a0 = 0, c0 = 0, d0 = 0;
  if (nrow > 0) { const LinePair<T> r = acl[0]; a0 = r.a; c0 = r.c; d0 = dl[0]; }
  al[1] = 0; ga[1] = 0; de[1] = 0;
  if (nrow > 1) { const LinePair<T> r = acl[1]; al[1] = r.a; ga[1] = r.c; de[1] = dl[1]; }
#pragma unroll
  for (int u = 2; u < M; u++) {
    T a = 0, c = 0, d = 0;
    if (u < nrow) { const LinePair<T> r = acl[u]; a = r.a; c = r.c; d = dl[u]; }
    const T r = line_rcp<T>((T)1 - a * ga[u - 1]);
    al[u] = -a * al[u - 1] * r;
    ga[u] = c * r;
    de[u] = (d - a * de[u - 1]) * r;
  }
  T p = 0, q = 0, t = -1;  // M == 2: x_{s+1} = x_e
  if constexpr (M > 2) {
    p = de[M - 2]; q = al[M - 2]; t = ga[M - 2];
#pragma unroll
    for (int u = M - 3; u >= 1; u--) {
      p = de[u] - ga[u] * p;
      q = al[u] - ga[u] * q;
      t = -ga[u] * t;
    }
  }
  const T r0 = line_rcp<T>((T)1 - c0 * q);
  TriRow<T> rs{a0 * r0, -c0 * t * r0, (d0 - c0 * p) * r0};
  TriRow<T> re{al[M - 1], ga[M - 1], de[M - 1]};
  {  // reduced stride 1: s_L couples e_{L-1} and e_L; e_L couples s_L and s_{L+1}
    const TriRow<T> em = row_shift3<1>(re), sp = row_shift3<-1>(rs);
    const TriRow<T> ns = pcr_combine(rs, em, re);
    const TriRow<T> ne = pcr_combine(re, rs, sp);
    rs = ns; re = ne;
  }
  reduced_step<1>(rs, re);
  reduced_step<2>(rs, re);
  reduced_step<4>(rs, re);
  reduced_step<8>(rs, re);
  const T xs = rs.d, xe = re.d;
  if (nrow > 0) dl[0] = xs;
  if (nrow > M - 1) dl[M - 1] = xe;
  T xn = xe;
#pragma unroll
  for (int u = M - 2; u >= 1; u--) {
    xn = de[u] - al[u] * xs - ga[u] * xn;
    if (u < nrow) dl[u] = xn;
  }
}

#define LINES_NT 512
// Waves per SIMD the line kernel is compiled for (register budget), and
// rows of the load phase whose loads are issued together per thread.
#ifndef LINES_MIN_WAVES
#define LINES_MIN_WAVES 6
#endif
#ifndef LINES_LOAD_UNROLL
#define LINES_LOAD_UNROLL 2
#endif
// Operands of one line-kernel launch.  Inside the box the current values of
// the unknown are either in xt or, when a pair is given, the average
// (a + b)/2 of the two line solutions of the previous pass of that colour
// (the value the separate average would have stored in xt); outside the
// box they are in xt, and a pair also holds copies of xt on the two
// x-planes just outside the box (k_pair_xbound).
template <typename T>
struct LineIO {
  const T *xt;
  const T *own2, *own3;  // own-colour planes as a pair, or null (xt)
  const T *oth2, *oth3;  // other-colour planes as a pair, or null (xt)
  T *xy, *xz;            // y-line and z-line solutions of the own-colour planes
  // G = bs + AS(x-) x(i-2) + AS(x+) x(i+2) of the own-colour cells (x0 on
  // pinned rows): stored by the y-line workgroups when gw is given (first
  // pass of a colour visit), read instead of its terms when gr is given
  // (second pass: the other colour's planes have not changed)
  T *gw;
  const T *gr;
  int pi0, pdi, npl;     // own-colour planes i = pi0 + pdi*p, p < npl
  int ib, ie, jb, je, kb, ke;  // the box
};
template <int LW, typename T, int NU, int M>
__global__ __launch_bounds__(LINES_NT, LINES_MIN_WAVES) void k_lines(const LineIO<T> io) {
  extern __shared__ double smd[];
  const T *__restrict__ xt = io.xt;
  const int pi0 = io.pi0, pdi = io.pdi, npl = io.npl, jb = io.jb, je = io.je, kb = io.kb, ke = io.ke;
  const int ny = (je - jb) / 2 + 1, nz = (ke - kb) / 2 + 1;
  const int ntile = (npl + LW - 1) / LW;
  // Workgroups go to the 8 XCDs round robin (workgroup b to XCD b mod 8).
  // XCD x takes the y-lines at o in [x*nz/8, (x+1)*nz/8) and the z-lines at
  // o in [x*ny/8, (x+1)*ny/8), all tiles: workgroups that read the same
  // neighbour rows (adjacent o, adjacent tiles) share one L2, and the XCDs
  // get equal numbers of rows to within one line per tile.  The grid holds
  // 8*ntile*(ceil(nz/8) + ceil(ny/8)) workgroups; the surplus ones return.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int oy0 = xcd * nz / 8, oy1 = (xcd + 1) * nz / 8, oz0 = xcd * ny / 8, oz1 = (xcd + 1) * ny / 8;
  const int cy = ntile * (oy1 - oy0), cz = ntile * (oz1 - oz0);
  int dir, n, o, tile, nlin;
  T *xo;
  if (slot < cy) { dir = 2; n = ny; nlin = nz; tile = slot % ntile; o = oy0 + slot / ntile; xo = io.xy; }
  else if (slot < cy + cz) { dir = 3; n = nz; nlin = ny; tile = (slot - cy) % ntile; o = oz0 + (slot - cy) / ntile; xo = io.xz; }
  else return;
  const bool own_p = io.own2 != nullptr, oth_p = io.oth2 != nullptr;
  const int npad = n | 1;
  LinePair<T> *AC = reinterpret_cast<LinePair<T> *>(smd);
  T *D = reinterpret_cast<T *>(AC + LW * npad);
  const int t = threadIdx.x;
  constexpr int NRG = LINES_NT / LW;
  const int w = t % LW, rg = t / LW;
  const int pl = tile * LW + w;
  const bool valid = pl < npl;
  const int i = pi0 + pdi * (valid ? pl : 0);
  // AS slots along the line (minus, plus) and across it, AS(.,d,m) -> (d-1)+3*(m-1)
  const int s_lm = (dir == 2) ? 1 : 2, s_lp = s_lm + 3;
  const int s_cm = (dir == 2) ? 2 : 1, s_cp = s_cm + 3;
  const int dj = (dir == 2) ? 2 : 0, dk = 2 - dj;  // lattice step along the line
  // arrays of the current values (uniform over the workgroup)
  const bool pcp = own_p && o + 1 < nlin, pcm = own_p && o > 0;
  const T *oth2 = oth_p ? io.oth2 : xt, *oth3 = oth_p ? io.oth3 : xt;
  const T *cp2 = pcp ? io.own2 : xt, *cp3 = pcp ? io.own3 : xt;
  const T *cm2 = pcm ? io.own2 : xt, *cm3 = pcm ? io.own3 : xt;
  const bool g_read = io.gr != nullptr, g_write = io.gw != nullptr && dir == 2;
#pragma unroll LINES_LOAD_UNROLL
  for (int e = rg; e < n; e += NRG) {
    if (!valid) break;
    const int j = (dir == 2) ? jb + 2 * e : jb + 2 * o;
    const int k = (dir == 2) ? kb + 2 * o : kb + 2 * e;
    // 32-bit byte offsets from uniform base pointers (every lattice array
    // of the case is below 4 GB, checked at init)
    const unsigned q = (unsigned)lidx(i, j, k), bq = q * (unsigned)sizeof(T);
    const int code = ld_off(g.ib, 2u * q);
    const T ap = ld_off(coef_ap_base<T>(), bq);
    const T ac_p = ld_off(coef_as_base<T>(s_cp), bq), ac_m = ld_off(coef_as_base<T>(s_cm), bq);
    const T al_m = ld_off(coef_as_base<T>(s_lm), bq), al_p = ld_off(coef_as_base<T>(s_lp), bq);
    // current values: the cross-direction neighbours on the own plane (past
    // the last line of the box they lie outside it, in xt), and for G the
    // x neighbours on the other colour's planes (a pair also holds the
    // planes just outside the box in x, copies of xt there, so these need
    // no test).  Each value is read from two arrays, a pair or xt twice, so
    // that every load is issued unconditionally.
    const unsigned bcp = (unsigned)lidx(i, j + 2 - dj, k + 2 - dk) * (unsigned)sizeof(T);
    const unsigned bcm = (unsigned)lidx(i, j - 2 + dj, k - 2 + dk) * (unsigned)sizeof(T);
    const T xcpa = ld_off(cp2, bcp), xcpb = ld_off(cp3, bcp);
    const T xcma = ld_off(cm2, bcm), xcmb = ld_off(cm3, bcm);
    const T xlm = (e == 0) ? xt[lidx(i, j - dj, k - dk)] : (T)0;
    const T xlp = (e == n - 1) ? xt[lidx(i, j + dj, k + dk)] : (T)0;
    const bool pinned = code >= 1 || fabs((double)ap) <= SMALL20;
    T gx;
    if (g_read) {
      gx = ld_off(io.gr, bq);
    } else {
      const unsigned bxp = (unsigned)lidx(i + 2, j, k) * (unsigned)sizeof(T);
      const unsigned bxm = (unsigned)lidx(i - 2, j, k) * (unsigned)sizeof(T);
      const T bs = ld_off(coef_bs_base<T>(), bq);
      const T ax_p = ld_off(coef_as_base<T>(3), bq), ax_m = ld_off(coef_as_base<T>(0), bq);
      const T xxpa = ld_off(oth2, bxp), xxpb = ld_off(oth3, bxp);
      const T xxma = ld_off(oth2, bxm), xxmb = ld_off(oth3, bxm);
      const T xxp = oth_p ? (xxpa + xxpb) / (T)2 : xxpa, xxm = oth_p ? (xxma + xxmb) / (T)2 : xxma;
      // a pinned row keeps its value, which stays in xt for the whole call
      gx = pinned ? ld_off(xt, bq) : bs + ax_p * xxp + ax_m * xxm;
      if (g_write) *reinterpret_cast<T *>(reinterpret_cast<char *>(io.gw) + bq) = gx;
    }
    const T xcp = pcp ? (xcpa + xcpb) / (T)2 : xcpa, xcm = pcm ? (xcma + xcmb) / (T)2 : xcma;
    T a, c, d;
    if (pinned) {
      a = 0; c = 0; d = gx;
    } else {
      T bsy = gx + ac_p * xcp + ac_m * xcm;
      T am = al_m, apn = al_p;
      if (e == 0) { bsy += am * xlm; am = 0; }
      if (e == n - 1) { bsy += apn * xlp; apn = 0; }
      const T r = line_rcp<T>(ap);
      a = -am * r; c = -apn * r; d = bsy * r;
    }
    AC[w * npad + e] = LinePair<T>{a, c};
    D[w * npad + e] = d;
  }
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  if constexpr (M > 0) {
    // Lines of up to 16*M rows: 16 lanes (one DPP row) per line, four lines
    // per wavefront, each lane M contiguous rows (line_partition_solve).
    // Lanes of absent lines solve zero rows, so that every lane of a
    // wavefront takes part in the row shifts.
    const int sub = lane >> 4, l16 = lane & 15;
    for (int ln0 = 4 * wave; ln0 < LW; ln0 += 4 * (LINES_NT / 64)) {
      const int ln = ln0 + sub;
      const bool act = ln < LW && tile * LW + ln < npl;
      LinePair<T> *ac_ = AC + (act ? ln : 0) * npad;
      T *d_ = D + (act ? ln : 0) * npad;
      const int nrow = act ? min(max(n - l16 * M, 0), M) : 0;
      line_partition_solve<T, M>(ac_ + l16 * M, d_ + l16 * M, nrow);
    }
  } else {
  // PCR on unit-diagonal rows: row e (a, 1, c | d) combined with rows e-s
  // and e+s gives (-a*a_m, 1 - a*c_m - c*a_p, -c*c_p | d - a*d_m - c*d_p),
  // renormalised.  After ceil(log2 n) steps every row is (0, 1, 0 | x).
  const int half = lane >> 5, l32 = lane & 31;
  for (int pr = wave; pr < LW / 2; pr += LINES_NT / 64) {
    const int ln = 2 * pr + half;
    const bool act = tile * LW + ln < npl;  // differs between the halves: every wave_sync stays outside it
    LinePair<T> *ac_ = AC + ln * npad;
    T *d_ = D + ln * npad;
    T ra[NU], rc[NU], rd[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int e = l32 + 32 * u;
      ra[u] = 0; rc[u] = 0; rd[u] = 0;
      if (act && e < n) { const LinePair<T> p = ac_[e]; ra[u] = p.a; rc[u] = p.c; rd[u] = d_[e]; }
    }
    int s = 1;
    for (; s < n && s < 32; s <<= 1) {
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = l32 + 32 * u;
        if (!act || e >= n) continue;
        LinePair<T> pm{0, 0}, pp{0, 0};
        T dm = 0, dp = 0;
        if (e - s >= 0) { pm = ac_[e - s]; dm = d_[e - s]; }
        if (e + s < n) { pp = ac_[e + s]; dp = d_[e + s]; }
        const T r = line_rcp<T>((T)1 - ra[u] * pm.c - rc[u] * pp.a);
        const T na = -ra[u] * pm.a * r, nc = -rc[u] * pp.c * r, nd = (rd[u] - ra[u] * dm - rc[u] * dp) * r;
        ra[u] = na; rc[u] = nc; rd[u] = nd;
      }
      wave_sync();
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = l32 + 32 * u;
        if (act && e < n) { ac_[e] = LinePair<T>{ra[u], rc[u]}; d_[e] = rd[u]; }
      }
      wave_sync();
    }
    if (s < n) {  // strides 32, 64, 128: within a lane
      if (NU > 1) pcr_lane_step<T, NU, 1>(ra, rc, rd, l32, n, act);
      if (NU > 2 && 64 < n) pcr_lane_step<T, NU, 2>(ra, rc, rd, l32, n, act);
      if (NU > 4 && 128 < n) pcr_lane_step<T, NU, 4>(ra, rc, rd, l32, n, act);
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = l32 + 32 * u;
        if (act && e < n) d_[e] = rd[u];
      }
    }
  }
  }
  __syncthreads();
  for (int e = rg; e < n; e += NRG) {
    if (!valid) break;
    const int j = (dir == 2) ? jb + 2 * e : jb + 2 * o;
    const int k = (dir == 2) ? kb + 2 * o : kb + 2 * e;
    xo[lidx(i, j, k)] = D[w * npad + e];
  }
}

// ======================================================================
// layout conversion between the reference's Fortran-ordered lattice and
// the class layout
// ======================================================================
__global__ void k_scatter(const double *src, double *dst, long long n) {
  long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f >= n) return;
  int i = (int)(f % MP) + 1;
  int j = (int)((f / MP) % NP) + 1;
  int k = (int)(f / ((long long)MP * NP)) + 1;
  dst[lidx(i, j, k)] = src[f];
}
// FP32 copy of a coefficient array (mixed-precision contexts).
__global__ void k_narrow(const double *src, float *dst, long long n) {
  long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f < n) dst[f] = (float)src[f];
}
__global__ void k_gather(const double *src, double *dst, long long n) {
  long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f >= n) return;
  int i = (int)(f % MP) + 1;
  int j = (int)((f / MP) % NP) + 1;
  int k = (int)(f / ((long long)MP * NP)) + 1;
  dst[f] = src[lidx(i, j, k)];
}
__global__ void k_scatter_i16(const int16_t *src, int16_t *dst, long long n) {
  long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f >= n) return;
  int i = (int)(f % MP) + 1;
  int j = (int)((f / MP) % NP) + 1;
  int k = (int)(f / ((long long)MP * NP)) + 1;
  dst[lidx(i, j, k)] = src[f];
}

#undef MP
#undef NP
#undef LP
#undef X
#undef DX
#undef R
#undef DR
#undef RS
#undef Z
#undef DZ
#undef IBC
#undef P_
#undef T_
#undef DN
#undef TH
#undef TMU
#undef GCP
#undef GLAM
#undef GAMA
#undef GTE
#undef CRATE
#undef HRATE
#undef UG
#undef GF
#undef FZ
#undef BS
#undef SFP
#undef AP
#undef AS

// ---------------- host-side binding of the constants ----------------
hipError_t dev_bind_case(const DevCase &c) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g), &c, sizeof(DevCase));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_nist), h_nist_tab, sizeof(h_nist_tab));
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_cutT), h_cutT, sizeof(h_cutT));
}
