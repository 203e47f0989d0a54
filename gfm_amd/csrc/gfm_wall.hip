// Wall radiosity exchange (C/Qwall.f90) on the device: the patch list of
// prepare_for_wall_radiation_calculation built on the host from the cell
// codes (C/Qwall.f90:74-566), the view factors, and wallq's loops
// (C/Qwall.f90:990-1369) with calc_q_conv_wall and wall_q_amb.
//
// Layout: per patch arrays in patch order; the normalised view factors
// row-major, VF(n,m) at vf[(n-1)*nb + (m-1)], so that the sums of a
// Gauss-Seidel row read one contiguous row.
//
// One Gauss-Seidel sweep updates the patches in patch order.  The patches
// are taken in blocks of WB = 64, one launch of k_wall_block each: its
// workgroups sum, for the block's rows, VF(i,j)*J(j) over every column
// outside the block (the columns before it hold this sweep's values, those
// after it the last sweep's) in column chunks spread over the chip; the
// last workgroup to finish then, in one wavefront (one row per lane), adds
// the chunk sums in a fixed order and the block's own columns in patch
// order -- a row's block-internal sum starts from the last sweep's values
// of the rows after it and takes each earlier row's new value as that row
// is solved -- and appends each row's relative change to the sweep's
// running sum in patch order.  The sums over the other patches are
// therefore added in another order than the reference's, which changes
// the rounding only; the update order is the reference's.
#pragma once

#define WB 64          // rows per Gauss-Seidel block (one wavefront)
#define WCH 256        // columns per chunk of k_wall_block
#define WROWS 16       // rows per workgroup of k_wall_block
#define WALL_NT 1024   // threads of the one-workgroup kernels
static constexpr double W_SIG = 5.668e-8, W_TMNK = 2.0e2, W_PI = 3.141592653589793;

struct WallDev {
  int nb, nblk, nchunk;
  int mp, np, lp;
  int16_t *typ, *bi, *bj, *bk, *orient, *direct, *bci, *bcj, *bck, *exh;
  int8_t *last;        // 1: the last wall patch of its wall cell (the one whose b_T the cell keeps)
  long long *qopen;    // lattice index of the open cell next to the patch
  long long *qbc;      // lattice index of the boundary cell
  long long *qface;    // lattice index of the patch face
  long long *cbc;      // cell-centre index of the boundary cell (radiation fields)
  double *area, *face_area, *emis, *jrad, *q_amb, *q_conv, *qc, *qeb, *bt, *ua, *ta, *ds, *rfb;
  double *vf, *vf_sum;
  double *vd;          // the diagonal blocks of VF, block b at vd[b*WB*WB]: row a, column c at [a*WB + c]
  double *part;        // (nchunk, WB) partial sums of the current block
  unsigned *cnt;       // workgroups of k_wall_block done (the last one solves the block and resets it)
  double *qrs;         // (MP, NP), reference layout
  double *wrt;         // wall_rad_T, reference layout (MP, NP, LP)
  double *qcf;         // q_conv at patch faces, device lattice
  double *res;         // scalar results (WR_*)
  double *log;         // mean relative change per sweep
  int log_cap;
  // inputs
  int first;           // i1st_rad_done == 0
  int maxri1, maxri2, minri2, nex, isoot_cal, frac_exh;
  double tol_rad, tol_bnd, h_g, wat;
  double exh_frac[GFM_NEX_MAX], exh_fixed[GFM_NEX_MAX];
  int exh_type[GFM_NEX_MAX];
};
enum {
  WR_MEAN = 0, WR_QMELT, WR_QMELT_PREV, WR_QCONV, WR_QLOSS, WR_AVGT, WR_QINC, WR_QINC_W, WR_QINC_I, WR_QINC_E,
  WR_QINC_M, WR_QRIN, WR_QROUT, WR_QMF, WR_QMB, WR_QMC, WR_AVGT_PREV, WR_BAD, WR_N
};

// ---- view factors, C/Qwall.f90:565-661 ----
__device__ double wall_vf_entry(const WallDev &w, int p, int m) {
#pragma clang fp contract(off)
  if (m == p) return 0.0;
  const int ip = w.bi[p], jp = w.bj[p], kp = w.bk[p], im = w.bi[m], jm = w.bj[m], km = w.bk[m];
  const double ddx = g.x[ip - 1] - g.x[im - 1], ddy = g.r[jp - 1] - g.r[jm - 1], ddz = g.z[kp - 1] - g.z[km - 1];
  const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
  const double den = W_PI * (d2 * d2);
  const double ex = g.x[im - 1] - g.x[ip - 1], ey = g.r[jm - 1] - g.r[jp - 1], ez = g.z[km - 1] - g.z[kp - 1];
  const int om = w.orient[m], dp = w.direct[p], op = w.orient[p];
  double num;
  if (op == 1) {
    if ((dp == 0 && im <= ip) || (dp == 1 && im >= ip)) return 0.0;
    num = (om == 1) ? ex * ex : fabs(ex * (om == 2 ? ey : ez));
  } else if (op == 2) {
    if ((dp == 0 && jm <= jp) || (dp == 1 && jm >= jp)) return 0.0;
    num = (om == 2) ? ey * ey : fabs(ey * (om == 1 ? ex : ez));
  } else {
    if ((dp == 0 && km <= kp) || (dp == 1 && km >= kp)) return 0.0;
    num = (om == 3) ? ez * ez : fabs(ez * (om == 1 ? ex : ey));
  }
  return num * w.area[m] / den;
}
// the sum of each row in column order (one thread per row)
__global__ void k_wall_vf_sum(WallDev w) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.nb) return;
  double s = 0.0;
  for (int m = 0; m < w.nb; m++) s = s + wall_vf_entry(w, p, m);
  w.vf_sum[p] = s;
}
__global__ void k_wall_vf(WallDev w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
  if (m >= w.nb) return;
  const double v = wall_vf_entry(w, p, m), s = w.vf_sum[p];
  w.vf[(size_t)p * w.nb + m] = (s == 0.0) ? v : v / s;
}

// ---- helpers of the one-workgroup kernels ----
// sequential sum in patch order of f(p) over the patches with sel(p), by thread 0 (the reference's order)
// (the terms are formed by all threads into the kernel's LDS buffer sh, 4096 at a time, and added there by thread 0)
#define WSEQ 4096
struct WallShared { double terms[WSEQ]; double sum; };
template <typename F>
__device__ double wall_seq_sum(const WallDev &w, WallShared &sh, F f) {
  double s = 0.0;
  for (int p0 = 0; p0 < w.nb; p0 += WSEQ) {
    const int n = min(WSEQ, w.nb - p0);
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) sh.terms[q] = f(p0 + q);
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < n; q++) s = s + sh.terms[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) sh.sum = s;
  __syncthreads();
  return sh.sum;
}

// wall_q_amb, C/Qwall.f90:1585-1659: q_amb and the wall loss
__device__ void wall_q_amb_dev(const WallDev &w, WallShared &sh) {
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x)
    w.q_amb[p] = (w.typ[p] == 1) ? w.ua[p] * (w.bt[p] - w.ta[p]) : 0.0;
  const double loss = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == 1) ? w.q_amb[p] * w.area[p] : 0.0; });
  if (threadIdx.x == 0) w.res[WR_QLOSS] = loss;
}

// the pass's start: patch temperatures (C/Rad_emis.f90:190-212, 253-261), wall_q_amb, qc, qeb and the
// incident sums (C/Qwall.f90:997-1026)
__global__ void __launch_bounds__(WALL_NT) k_wall_begin(WallDev w, const double *qa) {
  __shared__ WallShared sh;
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x) {
    const int t = w.typ[p];
    if (w.first) {
      if (t == 1 || t == 2) w.bt[p] = g.t[w.qbc[p]] * g.t0;
      else if (t == 3) {
        const int e = w.exh[p] - 1;
        w.bt[p] = (w.exh_type[e] == 0) ? w.exh_frac[e] * w.res[WR_AVGT_PREV] : w.exh_fixed[e];
      } else w.bt[p] = g.t[w.qface[p]] * g.t0;
      w.jrad[p] = W_SIG * pow(w.bt[p], 4.0);
    } else if (t == 4) {
      w.bt[p] = g.t[w.qface[p]] * g.t0;
      w.jrad[p] = W_SIG * pow(w.bt[p], 4.0);
    }
  }
  __syncthreads();
  wall_q_amb_dev(w, sh);
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x) {
    w.qc[p] = qa[w.cbc[p]] / w.face_area[p];
    const int t = w.typ[p];
    if (t == 2 || t == 3) w.qeb[p] = W_SIG * pow(w.bt[p], 4.0);
    else if (t == 4) w.qeb[p] = w.emis[p] * W_SIG * pow(w.bt[p], 4.0);
  }
  __syncthreads();
  const double tot = wall_seq_sum(w, sh, [&](int p) { return w.qc[p] * w.area[p]; });
  double s[4];
  for (int t = 1; t <= 4; t++)
    s[t - 1] = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == t) ? w.qc[p] * w.area[p] : 0.0; });
  if (threadIdx.x == 0) {
    w.res[WR_QINC] = tot;
    w.res[WR_QINC_W] = s[0]; w.res[WR_QINC_I] = s[1]; w.res[WR_QINC_E] = s[2]; w.res[WR_QINC_M] = s[3];
  }
}

// calc_q_conv_wall, C/Qwall.f90:1370-1478
__global__ void __launch_bounds__(WALL_NT) k_wall_qconv(WallDev w) {
  __shared__ WallShared sh;
  const double kappa_cmu = 0.4 * pow(9.0e-2, 0.25), e_cmu = 9.0 * pow(9.0e-2, 0.25);
  const double c_htw = 0.4 * ((W_PI / 4) / sin(W_PI / 4)) * sqrt(26.0) * (0.7 - 1.0) / pow(0.7, 0.25);
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x) {
    if (w.typ[p] == 2 || w.typ[p] == 3) { w.q_conv[p] = 0.0; continue; }
    const long long q = w.qopen[p];
    const double cpmix = 1.0057e3 * (CPF1 * g.gf[0][q] + (2.0e3 / 1.0057e3) * g.gf[5][q] + (2.0e3 / 1.0057e3) * g.gf[8][q] +
                                     (0.9136e3 / 1.0057e3) * g.gf[1][q] + (0.526e3 / 1.0057e3) * g.gf[6][q]);
    const double tgas = g.t[q] * g.t0;
    const double gmu_t = ((2.0e-15 * tgas + -1.1e-11) * tgas + 4.5e-8) * tgas + 7.0e-6;
    double h_t;
    if (g.tk0 * g.gf[3][q] < 0.1) h_t = 0.0;
    else {
      const double sq = sqrt(g.tk0 * g.gf[3][q]);
      const double dtmp = log(e_cmu * w.ds[p] * g.dnst0 * g.dnst[q] * sq / gmu_t) + (c_htw / sq);
      h_t = (dtmp < 0.01) ? 0.0 : fmin(50.0, cpmix * kappa_cmu * g.dnst0 * g.dnst[q] * sq / dtmp);
    }
    w.q_conv[p] = fmax(w.h_g, h_t) * (tgas - w.bt[p]);
  }
  const double tot = wall_seq_sum(w, sh, [&](int p) { return w.q_conv[p] * w.area[p]; });
  if (threadIdx.x == 0) { w.res[WR_QCONV] = tot; w.res[WR_MEAN] = 0.0; }
}

// does row p take the radiation of the other patches (C/Qwall.f90:1067-1069)
__device__ __forceinline__ bool wall_sees(const WallDev &w, int p) { return w.typ[p] == 1 || w.emis[p] < 1.0; }

// lane i's value of a double, for a uniform i (two v_readlane)
__device__ __forceinline__ double lane_value(double v, int i) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, i), hi = __builtin_amdgcn_readlane((int)(b >> 32), i);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
// The block's rows in patch order by the first wavefront of the workgroup (row r0 + lane): the chunk sums in
// order, the block's own columns from vd in registers, forward substitution with lane i's value broadcast by
// v_readlane.
__device__ void wall_block_solve(const WallDev &w, int r0, double radi_min, double (*vb)[WB + 1]) {
  const int l = threadIdx.x, row = r0 + l, nr = min(WB, w.nb - r0);
  int ty = 0;
  double old = 0.0, qcv = 0.0, qcn = 0.0, qam = 0.0, qeb = 0.0, em = 1.0;
  if (l < nr) {
    ty = w.typ[row]; old = w.jrad[row]; qcv = w.qc[row]; qcn = w.q_conv[row]; qam = w.q_amb[row];
    qeb = w.qeb[row]; em = w.emis[row];
  }
  double acc = 0.0;
  if (l < nr) {
    for (int c0 = 0; c0 < w.nchunk; c0 += 8) {
      double pv[8];
#pragma unroll
      for (int c = 0; c < 8; c++) pv[c] = (c0 + c < w.nchunk) ? w.part[(size_t)(c0 + c) * WB + l] : 0.0;
#pragma unroll
      for (int c = 0; c < 8; c++)
        if (c0 + c < w.nchunk) acc += pv[c];
    }
  }
  double rv[WB];
#pragma unroll
  for (int b = 0; b < WB; b++) rv[b] = vb[l][b];
#pragma unroll
  for (int b = 0; b < WB; b++) {
    const double jo = lane_value(old, b);
    if (b > l && b < nr) acc += rv[b] * jo;
  }
  const bool sees = ty == 1 || em < 1.0;
  double jn = old;
#pragma unroll
  for (int i = 0; i < WB; i++) {
    if (i < nr) {
      const double rfb = sees ? acc : 0.0;
      const double cand = (ty == 1) ? fmax(qcv + qcn - qam + rfb, em * radi_min) : qeb + (1.0 - em) * (rfb + qcv);
      const double ji = lane_value(cand, i);
      if (l == i) jn = ji;
      if (l > i) acc += rv[i] * ji;
    }
  }
  if (l < nr) w.jrad[row] = jn;
  // the sweep's running sum of relative changes, in patch order
  const double rel = (l < nr && !(old < w.tol_rad)) ? fabs((old - jn) / old) : 0.0;
  double m = w.res[WR_MEAN];
  for (int i = 0; i < nr; i++) m = m + lane_value(rel, i);
  if (l == 0) w.res[WR_MEAN] = m;
}

// One block of a sweep: every workgroup sums its rows over its column chunk; the last
// workgroup to finish (counted on w.cnt) solves the block.  No workgroup waits for another.
__global__ void __launch_bounds__(256) k_wall_block(WallDev w, int r0, double radi_min) {
  __shared__ double red[WROWS][256 / 64];
  __shared__ double vb[WB][WB + 1];
  __shared__ int last;
  const int c0 = blockIdx.x * WCH, rg = blockIdx.y * WROWS;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r1 = min(r0 + WB, w.nb);
  const int m = c0 + threadIdx.x;
  const bool col = m < w.nb && (m < r0 || m >= r1);
  const double jm = col ? w.jrad[m] : 0.0;
  double v[WROWS];
#pragma unroll
  for (int rr = 0; rr < WROWS; rr++) {
    const int row = r0 + rg + rr;
    v[rr] = (col && row < r1) ? w.vf[(size_t)row * w.nb + m] : 0.0;
  }
#pragma unroll
  for (int rr = 0; rr < WROWS; rr++) {
    double s = v[rr] * jm;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) red[rr][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < WROWS && rg + threadIdx.x < WB)
    w.part[(size_t)blockIdx.x * WB + rg + threadIdx.x] =
        ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned total = gridDim.x * gridDim.y;
    last = (atomicAdd(w.cnt, 1u) == total - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // the block's own columns, contiguous in vd
  const double *blk = w.vd + (size_t)(r0 / WB) * WB * WB;
  for (int q = threadIdx.x; q < WB * WB; q += blockDim.x) vb[q / WB][q % WB] = blk[q];
  __syncthreads();
  if (threadIdx.x < 64) wall_block_solve(w, r0, radi_min, vb);
  if (threadIdx.x == 0) atomicExch(w.cnt, 0u);
}

// the diagonal blocks of VF, contiguous (zero outside the patch range)
__global__ void k_wall_vd(WallDev w) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)w.nblk * WB * WB) return;
  const int b = (int)(t / (WB * WB)), a = (int)((t / WB) % WB), c = (int)(t % WB);
  const int r = b * WB + a, q = b * WB + c;
  w.vd[t] = (r < w.nb && q < w.nb) ? w.vf[(size_t)r * w.nb + q] : 0.0;
}

// rad_from_bnds of the rows of the given types (one wavefront per row)
__global__ void __launch_bounds__(256) k_wall_rows(WallDev w, int types) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= w.nb) return;
  if (!((types >> w.typ[row]) & 1)) return;
  const double *vr = w.vf + (size_t)row * w.nb;
  double s = 0.0;
  for (int m = lane; m < w.nb; m += 64) s += vr[m] * w.jrad[m];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) w.rfb[row] = s;
}

// after the sweeps of one outer iteration: wall temperatures, the melt flux, q_melt, wall_q_amb
// (C/Qwall.f90:1155-1191); rfb holds the melt rows' sums
__global__ void __launch_bounds__(WALL_NT) k_wall_outer(WallDev w) {
  __shared__ WallShared sh;
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x) {
    if (w.typ[p] == 1) {
      const double e = w.emis[p];
      const double b = pow(fmax(0.0, (w.jrad[p] - (w.q_amb[p] - w.q_conv[p]) * (1.0 - e) / e) / W_SIG), 0.25);
      w.bt[p] = fmax(W_TMNK, b);
    } else if (w.typ[p] == 4) {
      const double q = w.qc[p] + w.q_conv[p] + w.rfb[p] - w.jrad[p];
      w.qrs[(size_t)(w.bj[p] - 1) * w.mp + (w.bi[p] - 1)] = q;
    }
  }
  __syncthreads();
  const double qm = wall_seq_sum(w, sh, [&](int p) {
    return (w.typ[p] == 4) ? w.qrs[(size_t)(w.bj[p] - 1) * w.mp + (w.bi[p] - 1)] * w.area[p] : 0.0;
  });
  if (threadIdx.x == 0) { w.res[WR_QMELT_PREV] = w.res[WR_QMELT]; w.res[WR_QMELT] = qm; }
  __syncthreads();
  wall_q_amb_dev(w, sh);
}

// the closing loop (C/Qwall.f90:1209-1247), the exit reset, and the state of the next gas iterations: T at wall
// patch faces and wall cells with their enthalpy (C/Rad_emis.f90:337-355, 101-113), q_conv at the faces
__global__ void __launch_bounds__(WALL_NT) k_wall_close(WallDev w) {
  __shared__ WallShared sh;
  const double txa = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == 1) ? w.bt[p] * w.area[p] : 0.0; });
  const double qin = wall_seq_sum(w, sh, [&](int p) {
    return (w.typ[p] == 2) ? (w.qc[p] + w.rfb[p] - w.qeb[p]) * w.area[p] : 0.0; });
  const double qout = wall_seq_sum(w, sh, [&](int p) {
    return (w.typ[p] == 3) ? (w.qc[p] + w.rfb[p] - w.qeb[p]) * w.area[p] : 0.0; });
  auto fm = [&](int p) { const double e = w.emis[p];
                         return 1.0 - w.qeb[p] / (e * (w.qc[p] + w.rfb[p]) + w.q_conv[p]); };
  const double qmf = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == 4) ? fm(p) * w.emis[p] * w.qc[p] * w.area[p] : 0.0; });
  const double qmb = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == 4) ? fm(p) * w.emis[p] * w.rfb[p] * w.area[p] : 0.0; });
  const double qmc = wall_seq_sum(w, sh, [&](int p) { return (w.typ[p] == 4) ? fm(p) * w.q_conv[p] * w.area[p] : 0.0; });
  if (threadIdx.x == 0) {
    w.res[WR_QRIN] = qin; w.res[WR_QROUT] = qout; w.res[WR_QMF] = qmf; w.res[WR_QMB] = qmb; w.res[WR_QMC] = qmc;
  }
  const double avg = txa / w.wat;
  __syncthreads();
  if (threadIdx.x == 0) w.res[WR_AVGT] = avg;
  for (int p = threadIdx.x; p < w.nb; p += blockDim.x) {
    const int t = w.typ[p];
    if (t == 3 && w.frac_exh) {
      const int e = w.exh[p] - 1;
      w.bt[p] = (w.exh_type[e] == 0) ? w.exh_frac[e] * avg : w.exh_fixed[e];
    }
    if (t == 1) {
      g.t[w.qface[p]] = w.bt[p] / g.t0;
      if (w.last[p]) {
        const size_t c = (size_t)(w.bci[p] - 1) + (size_t)w.mp * ((w.bcj[p] - 1) + (size_t)w.np * (w.bck[p] - 1));
        w.wrt[c] = w.bt[p];
        g.t[w.qbc[p]] = w.bt[p] / g.t0;
        T_TO_ENTH(w.qbc[p]);
      }
    }
    w.qcf[w.qface[p]] = (t == 1 || t == 4) ? w.q_conv[p] : 0.0;
  }
}

// q_conv at the wall and melt patch faces (a restored record)
__global__ void k_wall_qcf(WallDev w) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.nb) return;
  w.qcf[w.qface[p]] = (w.typ[p] == 1 || w.typ[p] == 4) ? w.q_conv[p] : 0.0;
}

struct WallHost {
  WallDev dv;
  std::vector<double> log;  // mean relative change of each sweep of the last call
  gfm_wall_ctrl setup;     // inputs and totals of the setup
  std::vector<int16_t> typ;
  size_t cap_nb = 0;
  bool ready = false;
};

// the patch list of prepare_for_wall_radiation_calculation on the host (steps 1-6 and 8)
struct PatchList {
  std::vector<int16_t> typ, bi, bj, bk, orient, direct, bci, bcj, bck, exh;
  std::vector<double> area, emis, face_area;
};
static int wall_build_list(gfm_ctx *c, gfm_wall_ctrl *wc, const int16_t *iwall, PatchList &L, std::vector<double> &ua,
                           std::vector<double> &ta) {
  const gfm_ctrl &k = c->ctrl;
  const int MP = k.mp, NP = k.np, LP = k.lp;
  std::vector<int16_t> mb(c->ibh);
  auto IX = [&](int i, int j, int kk) { return (size_t)(i - 1) + (size_t)MP * ((size_t)(j - 1) + (size_t)NP * (kk - 1)); };
  auto M = [&](int i, int j, int kk) -> int16_t & { return mb[IX(i, j, kk)]; };
  auto at = [&](int i, int j, int kk) -> int {
    if (i < 1 || i > MP || j < 1 || j > NP || kk < 1 || kk > LP) return 1;
    return M(i, j, kk);
  };
  auto fillbox = [&](int i0, int i1, int j0, int j1, int k0, int k1, int v) {
    // inside the lattice only (an exit whose interior end was never found leaves a stale or zero index)
    i0 = std::max(i0, 1); j0 = std::max(j0, 1); k0 = std::max(k0, 1);
    i1 = std::min(i1, MP); j1 = std::min(j1, NP); k1 = std::min(k1, LP);
    for (int ii = i0; ii <= i1; ii++)
      for (int jj = j0; jj <= j1; jj++)
        for (int kk = k0; kk <= k1; kk++) M(ii, jj, kk) = (int16_t)v;
  };
  // exit tunnels (C/Qwall.f90:79-296); the corner variables carry from exit to exit, as the reference's do
  int ks = 0, ke = 0, js = 0, je = 0, is = 0, ie = 0, inew = 0, jnew = 0, knew = 0;
  for (int j = 2; j <= NP; j += 2)
    for (int kk = 2; kk <= LP; kk += 2) {
      if (M(2, j, kk) != 3) continue;
      ks = kk; for (int q = kk + 2; q <= LP; q += 2) if (M(2, j, q) != 3) { ke = q - 2; break; }
      js = j; for (int q = j + 2; q <= NP; q += 2) if (M(2, q, kk) != 3) { je = q - 2; break; }
      for (int ii = 4; ii <= MP; ii += 2)
        if (at(ii, js - 2, ks - 2) == 0 || at(ii, je + 2, ke + 2) == 0 || at(ii, js - 2, ke + 2) == 0 || at(ii, je + 2, ks - 2) == 0) { inew = ii - 2; break; }
      fillbox(inew, inew + 1, js, je, ks, ke, 3);
      fillbox(2, inew - 1, js, je, ks, ke, 1);
    }
  for (int j = 2; j <= NP; j += 2)
    for (int kk = 2; kk <= LP; kk += 2) {
      if (M(MP, j, kk) != 3) continue;
      ks = kk; for (int q = kk + 2; q <= LP; q += 2) if (M(MP, j, q) != 3) { ke = q - 2; break; }
      js = j; for (int q = j + 2; q <= NP; q += 2) if (M(MP, q, kk) != 3) { je = q - 2; break; }
      for (int ii = MP - 2; ii >= 2; ii -= 2)
        if (at(ii, js - 2, ks - 2) == 0 || at(ii, je + 2, ke + 2) == 0 || at(ii, js - 2, ke + 2) == 0 || at(ii, je + 2, ks - 2) == 0) { inew = ii + 2; break; }
      fillbox(inew - 1, inew, js, je, ks, ke, 3);
      fillbox(inew + 1, MP, js, je, ks, ke, 1);
    }
  for (int i = 2; i <= MP; i += 2)
    for (int kk = 2; kk <= LP; kk += 2) {
      if (M(i, 2, kk) != 3) continue;
      ks = kk; for (int q = kk + 2; q <= LP; q += 2) if (M(i, 2, q) != 3) { ke = q - 2; break; }
      is = i; for (int q = i + 2; q <= MP; q += 2) if (M(q, 2, kk) != 3) { ie = q - 2; break; }
      for (int jj = 4; jj <= NP; jj += 2)
        if (at(is - 2, jj, ks - 2) == 0 || at(ie + 2, jj, ke + 2) == 0 || at(is - 2, jj, ke + 2) == 0 || at(ie + 2, jj, ks - 2) == 0) { jnew = jj - 2; break; }
      fillbox(is, ie, jnew, jnew + 1, ks, ke, 3);
      fillbox(is, ie, 2, jnew - 1, ks, ke, 1);
    }
  for (int i = 2; i <= MP; i += 2)
    for (int kk = 2; kk <= LP; kk += 2) {
      if (M(i, NP, kk) != 3) continue;
      ks = kk; for (int q = kk + 2; q <= LP; q += 2) if (M(i, NP, q) != 3) { ke = q - 2; break; }
      is = i; for (int q = i + 2; q <= MP; q += 2) if (M(q, NP, kk) != 3) { ie = q - 2; break; }
      for (int jj = NP - 2; jj >= 2; jj -= 2)
        if (at(is - 2, jj, ks - 2) == 0 || at(ie + 2, jj, ke + 2) == 0 || at(is - 2, jj, ke + 2) == 0 || at(ie + 2, jj, ks - 2) == 0) { jnew = jj + 2; break; }
      fillbox(is, ie, jnew - 1, jnew, ks, ke, 3);
      fillbox(is, ie, jnew + 1, NP, ks, ke, 1);
    }
  for (int i = 2; i <= MP; i += 2)
    for (int j = 2; j <= NP; j += 2) {
      if (M(i, j, LP) != 3) continue;
      is = i; for (int q = i + 2; q <= MP; q += 2) if (M(q, j, LP) != 3) { ie = q - 2; break; }
      js = j; for (int q = j + 2; q <= NP; q += 2) if (M(i, q, LP) != 3) { je = q - 2; break; }
      for (int kk = LP - 2; kk >= 2; kk -= 2)
        if (at(is - 2, js - 2, kk) == 0 || at(ie + 2, je + 2, kk) == 0 || at(is - 2, je + 2, kk) == 0 || at(ie + 2, js - 2, kk) == 0) { knew = kk + 2; break; }
      fillbox(is, ie, js, je, knew - 1, knew, 3);
      fillbox(is, ie, js, je, knew + 1, LP, 1);
    }
  // patches (C/Qwall.f90:393-502)
  auto emis_of = [&](int code, int kk, int ktop) -> double {
    if (code == 1) return (c->hz[kk - 1] >= wc->height_to_ceiling || (kk == ktop && wc->have_crown == 0)) ? wc->eps_c : wc->eps_w;
    if (code == 4 && ktop == LP - 1) return wc->eps_m;
    return 1.0;
  };
  auto add = [&](int code, int i, int j, int kk, int o, int dir, int ci, int cj, int ck, double a, double e) {
    L.typ.push_back((int16_t)code); L.bi.push_back((int16_t)i); L.bj.push_back((int16_t)j); L.bk.push_back((int16_t)kk);
    L.orient.push_back((int16_t)o); L.direct.push_back((int16_t)dir);
    L.bci.push_back((int16_t)ci); L.bcj.push_back((int16_t)cj); L.bck.push_back((int16_t)ck);
    L.area.push_back(a); L.emis.push_back(e);
  };
  const std::vector<double> &DXh = c->hdx, &DRh = c->hdr, &DZh = c->hdz;
  for (int kk = 2; kk <= LP; kk += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int i = 3; i <= MP - 1; i += 2) {
        if (M(i, j, kk) < 1 || (M(i - 1, j, kk) > 0 && M(i + 1, j, kk) > 0)) continue;
        const int d = (M(i + 1, j, kk) == 0) ? 0 : 1;
        add(M(i, j, kk), i, j, kk, 1, d, d == 0 ? i - 1 : i + 1, j, kk, DRh[j - 1] * DZh[kk - 1], emis_of(M(i, j, kk), kk, LP));
      }
  for (int kk = 2; kk <= LP; kk += 2)
    for (int j = 3; j <= NP - 1; j += 2)
      for (int i = 2; i <= MP; i += 2) {
        if (M(i, j, kk) < 1 || (M(i, j - 1, kk) > 0 && M(i, j + 1, kk) > 0)) continue;
        const int d = (M(i, j + 1, kk) == 0) ? 0 : 1;
        add(M(i, j, kk), i, j, kk, 2, d, i, d == 0 ? j - 1 : j + 1, kk, DXh[i - 1] * DZh[kk - 1], emis_of(M(i, j, kk), kk, LP));
      }
  for (int kk = 3; kk <= LP - 1; kk += 2)
    for (int j = 2; j <= NP; j += 2)
      for (int i = 2; i <= MP; i += 2) {
        if (M(i, j, kk) < 1 || (M(i, j, kk - 1) > 0 && M(i, j, kk + 1) > 0)) continue;
        const int d = (M(i, j, kk + 1) == 0) ? 0 : 1;
        add(M(i, j, kk), i, j, kk, 3, d, i, j, d == 0 ? kk - 1 : kk + 1, DXh[i - 1] * DRh[j - 1], emis_of(M(i, j, kk), kk, LP - 1));
      }
  const int nb = (int)L.typ.size();
  if (nb == 0) { set_err("no boundary patch"); return GFM_E_ARG; }
  // exhausts (C/Qwall.f90:513-536)
  L.exh.assign(nb, 0);
  for (int p = 0; p < nb; p++) {
    if (L.typ[p] != 3) continue;
    for (int m = 0; m < wc->nex; m++) {
      if (L.orient[p] != wc->exh_orient[m] || L.direct[p] != wc->exh_direct[m]) continue;
      const int32_t *e = wc->exh_box[m];
      bool hit;
      if (L.orient[p] == 1) hit = L.bcj[p] >= e[1] && L.bcj[p] <= e[4] && L.bck[p] >= e[2] && L.bck[p] <= e[5];
      else if (L.orient[p] == 2) hit = L.bci[p] >= e[0] && L.bci[p] <= e[3] && L.bck[p] >= e[2] && L.bck[p] <= e[5];
      else hit = L.bci[p] >= e[0] && L.bci[p] <= e[3] && L.bcj[p] >= e[1] && L.bcj[p] <= e[4];
      if (hit) { L.exh[p] = (int16_t)(m + 1); break; }
    }
    if (L.exh[p] == 0) { set_err("exit patch %d belongs to no exhaust", p + 1); return GFM_E_ARG; }
  }
  // patches on one boundary cell (sameCell1..4, C/Qwall.f90:685-739): the reference's pairing, run over
  // each cell's own patches in patch order
  std::map<long long, std::vector<int>> cell;
  for (int p = 0; p < nb; p++) cell[((long long)L.bci[p] * 65536 + L.bcj[p]) * 65536 + L.bck[p]].push_back(p);
  std::vector<int> sc((size_t)4 * nb, 0);
  auto SC = [&](int q, int s) -> int & { return sc[(size_t)s * nb + q]; };
  for (auto &kv : cell) {
    const std::vector<int> &v = kv.second;
    for (size_t a = 0; a < v.size(); a++)
      for (size_t b = a + 1; b < v.size(); b++) {
        const int p = v[a], m = v[b], P1 = p + 1, M1 = m + 1;
        if (SC(p, 0) == 0) { SC(p, 0) = M1; SC(m, 0) = P1; }
        else if (SC(p, 1) == 0) {
          if (SC(p, 0) == M1) continue;
          SC(p, 1) = M1; SC(m, 0) = P1; SC(m, 1) = SC(p, 0); SC(SC(p, 0) - 1, 1) = M1;
        } else if (SC(p, 2) == 0) {
          if (SC(p, 0) == M1 || SC(p, 1) == M1) continue;
          SC(p, 2) = M1; SC(m, 0) = P1; SC(m, 1) = SC(p, 0); SC(SC(p, 1) - 1, 2) = M1; SC(m, 2) = SC(p, 1);
        } else if (SC(p, 3) == 0) {
          if (SC(p, 0) == M1 || SC(p, 1) == M1 || SC(p, 2) == M1) continue;
          SC(p, 3) = M1; SC(m, 0) = P1; SC(m, 1) = SC(p, 0); SC(SC(p, 1) - 1, 3) = M1;
          SC(m, 2) = SC(p, 1); SC(SC(p, 2) - 1, 3) = M1; SC(m, 3) = SC(p, 2);
        }
      }
  }
  L.face_area.resize(nb);
  wc->iwall_count = wc->imelt_count = wc->inlet_count = wc->iexit_count = 0;
  wc->wall_area_total = wc->melt_area_total = wc->inlet_area_total = wc->exit_area_total = 0.0;
  for (int p = 0; p < nb; p++) {
    double a = L.area[p];
    for (int s = 0; s < 4 && SC(p, s); s++) a = a + L.area[SC(p, s) - 1];
    L.face_area[p] = a;
    if (L.typ[p] == 1) { wc->iwall_count++; wc->wall_area_total += L.area[p]; }
    else if (L.typ[p] == 4) { wc->imelt_count++; wc->melt_area_total += L.area[p]; }
    else if (L.typ[p] == 2) { wc->inlet_count++; wc->inlet_area_total += L.area[p]; }
    else { wc->iexit_count++; wc->exit_area_total += L.area[p]; }
  }
  wc->boundary_area_total = wc->wall_area_total + wc->melt_area_total + wc->inlet_area_total + wc->exit_area_total;
  // wall_q_amb's parameters per patch (C/Qwall.f90:1612-1645)
  ua.assign(nb, 0.0); ta.assign(nb, 0.0);
  const int MZ = MP / 2, NZ = NP / 2;
  for (int p = 0; p < nb; p++) {
    if (L.typ[p] != 1) continue;
    int nw = 0;
    if (iwall) nw = iwall[((size_t)(L.bck[p] / 2 - 1) * NZ + (size_t)(L.bcj[p] / 2 - 1)) * MZ + (size_t)(L.bci[p] / 2 - 1)];
    double wk, ha, wd, tamb;
    if (nw > 0 && nw <= wc->nwal0) { wk = wc->w1_k[nw - 1]; ha = wc->w1_ha[nw - 1]; wd = wc->w1_d[nw - 1]; tamb = wc->w1_ta[nw - 1]; }
    else { wk = wc->w_k; ha = wc->h_a; wd = wc->w_d; tamb = wc->t_a; }
    ua[p] = (wk < 1.0e-20 || ha < 1.0e-20) ? 0.0 : 1.0 / (wd / wk + 1.0 / ha);
    ta[p] = tamb;
  }
  return GFM_OK;
}

static void wall_set_inputs(WallDev &d, const gfm_wall_ctrl *wc) {
  d.first = wc->i1st_rad_done == 0;
  d.maxri1 = wc->maxri1; d.maxri2 = wc->maxri2; d.minri2 = wc->minri2;
  d.nex = wc->nex; d.isoot_cal = wc->isoot_cal;
  d.tol_rad = wc->wall_radiosity_tolerance; d.tol_bnd = wc->radiation_boundary_tolerance;
  d.h_g = wc->h_g;
  d.frac_exh = 0;
  for (int m = 0; m < wc->nex; m++) {
    d.exh_frac[m] = wc->exh_frac[m]; d.exh_fixed[m] = wc->exh_fixed[m]; d.exh_type[m] = wc->exh_type[m];
    if (wc->exh_type[m] == 0) d.frac_exh = 1;
  }
}

extern "C" {

size_t gfm_hip_wall_ctrl_size(void) { return sizeof(gfm_wall_ctrl); }

int gfm_hip_wall_setup(gfm_ctx *c, gfm_wall_ctrl *wc, const int16_t *iwall) {
  if (!c || !wc) { set_err("null argument"); return GFM_E_ARG; }
  if (wc->nex < 0 || wc->nex > GFM_NEX_MAX || wc->nwal0 < 0 || wc->nwal0 > GFM_NWALG_MAX) {
    set_err("nex %d / nwal0 %d outside 0..%d / 0..%d", wc->nex, wc->nwal0, GFM_NEX_MAX, GFM_NWALG_MAX);
    return GFM_E_ARG;
  }
  int rc;
  if ((rc = bind(c))) return rc;
  PatchList L;
  std::vector<double> ua, ta;
  if ((rc = wall_build_list(c, wc, iwall, L, ua, ta))) return rc;
  const int nb = (int)L.typ.size();
  if ((double)nb * nb * sizeof(double) > 64.0 * (1 << 30)) { set_err("%d patches: view factors above 64 GiB", nb); return GFM_E_ARG; }
  wc->b_cnt = nb;
  if (!c->wall) c->wall = new WallHost();
  WallHost *h = c->wall;
  WallDev &d = h->dv;
  const gfm_ctrl &k = c->ctrl;
  d.mp = k.mp; d.np = k.np; d.lp = k.lp;
  if (h->cap_nb < (size_t)nb) {
    // the arrays of a larger patch list (earlier ones stay with the context until it ends)
    if ((rc = dalloc(c, &d.typ, nb)) || (rc = dalloc(c, &d.bi, nb)) || (rc = dalloc(c, &d.bj, nb)) || (rc = dalloc(c, &d.bk, nb)) ||
        (rc = dalloc(c, &d.orient, nb)) || (rc = dalloc(c, &d.direct, nb)) || (rc = dalloc(c, &d.bci, nb)) ||
        (rc = dalloc(c, &d.bcj, nb)) || (rc = dalloc(c, &d.bck, nb)) || (rc = dalloc(c, &d.exh, nb)) ||
        (rc = dalloc(c, &d.last, nb)) || (rc = dalloc(c, &d.qopen, nb)) || (rc = dalloc(c, &d.qbc, nb)) ||
        (rc = dalloc(c, &d.qface, nb)) || (rc = dalloc(c, &d.cbc, nb)) || (rc = dalloc(c, &d.area, nb)) ||
        (rc = dalloc(c, &d.face_area, nb)) || (rc = dalloc(c, &d.emis, nb)) || (rc = dalloc(c, &d.jrad, nb)) ||
        (rc = dalloc(c, &d.q_amb, nb)) || (rc = dalloc(c, &d.q_conv, nb)) || (rc = dalloc(c, &d.qc, nb)) ||
        (rc = dalloc(c, &d.qeb, nb)) || (rc = dalloc(c, &d.bt, nb)) || (rc = dalloc(c, &d.ua, nb)) ||
        (rc = dalloc(c, &d.ta, nb)) || (rc = dalloc(c, &d.ds, nb)) || (rc = dalloc(c, &d.rfb, nb)) ||
        (rc = dalloc(c, &d.vf_sum, nb)) || (rc = dalloc(c, &d.vf, (size_t)nb * nb)) ||
        (rc = dalloc(c, &d.vd, (size_t)((nb + WB - 1) / WB) * WB * WB)) ||
        (rc = dalloc(c, &d.part, (size_t)((nb + WCH - 1) / WCH) * WB)))
      return rc;
    h->cap_nb = nb;
  }
  if (!d.res) {
    if ((rc = dalloc(c, &d.res, WR_N)) || (rc = dalloc(c, &d.qrs, (size_t)k.mp * k.np)) ||
        (rc = dalloc(c, &d.wrt, (size_t)k.mp * k.np * k.lp)) || (rc = dalloc(c, &d.qcf, (size_t)c->ncls)) ||
        (rc = dalloc(c, &d.log, 1)) || (rc = dalloc(c, &d.cnt, 1)))
      return rc;
    d.log_cap = 0;
  }
  d.nb = nb;
  d.nblk = (nb + WB - 1) / WB;
  d.nchunk = (nb + WCH - 1) / WCH;
  // per patch: the open cell next to it and the distance to its centre (C/Qwall.f90:1401-1433), the lattice
  // positions of the boundary cell and the face, the boundary cell's centre index, the last wall patch of a cell
  std::vector<long long> qo(nb), qb(nb), qf(nb), cb(nb);
  std::vector<double> ds(nb);
  std::vector<int8_t> last(nb, 0);
  std::map<long long, int> lastp;
  const int MZ = k.mp / 2, NZ = k.np / 2;
  for (int p = 0; p < nb; p++) {
    int i = L.bi[p], j = L.bj[p], kk = L.bk[p];
    const bool dn = L.direct[p] == 0;
    if (L.orient[p] == 1) { if (dn) { ds[p] = c->hx[i] - c->hx[i - 1]; i++; } else { ds[p] = c->hx[i - 1] - c->hx[i - 2]; i--; } }
    else if (L.orient[p] == 2) { if (dn) { ds[p] = c->hr[j] - c->hr[j - 1]; j++; } else { ds[p] = c->hr[j - 1] - c->hr[j - 2]; j--; } }
    else { if (dn) { ds[p] = c->hz[kk] - c->hz[kk - 1]; kk++; } else { ds[p] = c->hz[kk - 1] - c->hz[kk - 2]; kk--; } }
    qo[p] = lat_index(c->dc.g, i, j, kk);
    qb[p] = lat_index(c->dc.g, L.bci[p], L.bcj[p], L.bck[p]);
    qf[p] = lat_index(c->dc.g, L.bi[p], L.bj[p], L.bk[p]);
    cb[p] = (long long)(L.bci[p] / 2 - 1) + (long long)MZ * ((L.bcj[p] / 2 - 1) + (long long)NZ * (L.bck[p] / 2 - 1));
    if (L.typ[p] == 1) lastp[qb[p]] = p;
  }
  for (auto &kv : lastp) last[kv.second] = 1;
  auto up = [&](auto *dst, const auto &v) { return hipMemcpy(dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice); };
  HCHECK(up(d.typ, L.typ)); HCHECK(up(d.bi, L.bi)); HCHECK(up(d.bj, L.bj)); HCHECK(up(d.bk, L.bk));
  HCHECK(up(d.orient, L.orient)); HCHECK(up(d.direct, L.direct)); HCHECK(up(d.bci, L.bci)); HCHECK(up(d.bcj, L.bcj));
  HCHECK(up(d.bck, L.bck)); HCHECK(up(d.exh, L.exh)); HCHECK(up(d.last, last)); HCHECK(up(d.qopen, qo));
  HCHECK(up(d.qbc, qb)); HCHECK(up(d.qface, qf)); HCHECK(up(d.cbc, cb)); HCHECK(up(d.area, L.area));
  HCHECK(up(d.face_area, L.face_area)); HCHECK(up(d.emis, L.emis)); HCHECK(up(d.ua, ua)); HCHECK(up(d.ta, ta));
  HCHECK(up(d.ds, ds));
  HCHECK(hipMemset(d.jrad, 0, nb * sizeof(double))); HCHECK(hipMemset(d.bt, 0, nb * sizeof(double)));
  HCHECK(hipMemset(d.q_conv, 0, nb * sizeof(double))); HCHECK(hipMemset(d.q_amb, 0, nb * sizeof(double)));
  HCHECK(hipMemset(d.qrs, 0, (size_t)k.mp * k.np * sizeof(double)));
  HCHECK(hipMemset(d.wrt, 0, (size_t)k.mp * k.np * k.lp * sizeof(double)));
  HCHECK(hipMemset(d.qcf, 0, (size_t)c->ncls * sizeof(double)));
  hipLaunchKernelGGL(k_wall_vf_sum, dim3((nb + 255) / 256), dim3(256), 0, c->stream, d);
  hipLaunchKernelGGL(k_wall_vf, dim3((nb + 255) / 256, nb), dim3(256), 0, c->stream, d);
  hipLaunchKernelGGL(k_wall_vd, dim3((unsigned)(((long long)d.nblk * WB * WB + 255) / 256)), dim3(256), 0, c->stream, d);
  HCHECK(hipMemsetAsync(d.cnt, 0, sizeof(unsigned), c->stream));
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(c->stream));
  h->typ = L.typ;
  h->setup = *wc;
  h->ready = true;
  return GFM_OK;
}

int gfm_hip_wallq(gfm_ctx *c, gfm_wall_ctrl *wc) {
  if (!c || !wc) { set_err("null argument"); return GFM_E_ARG; }
  if (!c->wall || !c->wall->ready) { set_err("no patch list (gfm_hip_wall_setup first)"); return GFM_E_STATE; }
  if (!c->rad || c->rad->nwl_cap == 0) { set_err("no radiation pass has run"); return GFM_E_STATE; }
  WallHost *h = c->wall;
  WallDev &d = h->dv;
  if (wc->b_cnt != d.nb) { set_err("b_cnt %d is not the set-up %d", wc->b_cnt, d.nb); return GFM_E_ARG; }
  if (wc->maxri1 < 1 || wc->maxri2 < 1 || wc->nex < 0 || wc->nex > GFM_NEX_MAX) { set_err("bad loop limits"); return GFM_E_ARG; }
  if (wc->i1st_rad_done == 0)
    for (int p = 0; p < d.nb; p++)
      if (h->typ[p] == 3 && wc->nex == 0) { set_err("exit patches but no exhausts"); return GFM_E_ARG; }
  int rc;
  if ((rc = bind(c))) return rc;
  wall_set_inputs(d, wc);
  d.wat = h->setup.wall_area_total;
  double res[WR_N] = {0};
  res[WR_QMELT] = wc->q_melt;
  res[WR_AVGT_PREV] = wc->avg_wall_T;
  HCHECK(hipMemcpyAsync(d.res, res, sizeof(res), hipMemcpyHostToDevice, c->stream));
  const double radi_min = W_SIG * pow(W_TMNK, 4.0);
  const double zero = 0.0;
  if (c->prof.on) prof_begin(c);
  hipLaunchKernelGGL(k_wall_begin, dim3(1), dim3(WALL_NT), 0, c->stream, d, (const double *)c->rad->dv.qa);
  std::vector<double> log;
  int nn = 1, nnextra = wc->maxri1, sweeps = 0, outer = 0;
  double last_mean = 0.0, q_melt = wc->q_melt;
  while (nn <= wc->maxri1 && nn <= nnextra) {
    const int minradi = wc->minri2;
    hipLaunchKernelGGL(k_wall_qconv, dim3(1), dim3(WALL_NT), 0, c->stream, d);
    int n = 1, nextra = wc->maxri2;
    while (n <= wc->maxri2 && n <= nextra) {
      for (int b = 0; b < d.nblk; b++)
        hipLaunchKernelGGL(k_wall_block, dim3(d.nchunk, WB / WROWS), dim3(256), 0, c->stream, d, b * WB, radi_min);
      HCHECK(hipGetLastError());
      double m;
      HCHECK(hipMemcpyAsync(&m, d.res + WR_MEAN, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HCHECK(hipMemcpyAsync(d.res + WR_MEAN, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
      HCHECK(hipStreamSynchronize(c->stream));
      m = m / d.nb;
      if (!std::isfinite(m)) { set_err("non-finite radiosity change"); return GFM_E_NUMERIC; }
      log.push_back(m);
      wc->iterw++;
      sweeps++;
      last_mean = m;
      if (m < wc->wall_radiosity_tolerance && n > minradi)
        if (nextra == wc->maxri2) nextra = n + 3;
      n++;
    }
    hipLaunchKernelGGL(k_wall_rows, dim3((d.nb + 3) / 4), dim3(256), 0, c->stream, d, 1 << 4);
    hipLaunchKernelGGL(k_wall_outer, dim3(1), dim3(WALL_NT), 0, c->stream, d);
    double qm[2];
    HCHECK(hipMemcpyAsync(qm, d.res + WR_QMELT, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    const double prev = qm[1];
    q_melt = qm[0];
    const double chg = fabs((prev - q_melt) / (q_melt + 1.0e-16));
    if (chg < wc->radiation_boundary_tolerance)
      if (nnextra == wc->maxri1) nnextra = nn + 1;
    outer++;
    nn++;
  }
  hipLaunchKernelGGL(k_wall_rows, dim3((d.nb + 3) / 4), dim3(256), 0, c->stream, d, (1 << 2) | (1 << 3) | (1 << 4));
  hipLaunchKernelGGL(k_wall_close, dim3(1), dim3(WALL_NT), 0, c->stream, d);
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(res, d.res, sizeof(res), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  if (c->prof.on) prof_end(c, "wallq", 0.0);
  for (double v : res)
    if (!std::isfinite(v)) { set_err("non-finite wall exchange sum"); return GFM_E_NUMERIC; }
  const gfm_wall_ctrl &s = h->setup;
  wc->q_incident_total = res[WR_QINC]; wc->q_incident_wall = res[WR_QINC_W]; wc->q_incident_inlet = res[WR_QINC_I];
  wc->q_incident_exit = res[WR_QINC_E]; wc->q_incident_melt = res[WR_QINC_M];
  wc->qc_avg = res[WR_QINC] / s.boundary_area_total;
  wc->qconv_total = res[WR_QCONV]; wc->qconv_avg = res[WR_QCONV] / s.boundary_area_total;
  wc->q_wall_loss = res[WR_QLOSS]; wc->qamb_avg = res[WR_QLOSS] / s.wall_area_total;
  wc->q_melt = q_melt; wc->qrs_avg = q_melt / s.melt_area_total;
  wc->q_rad_inlet = res[WR_QRIN]; wc->q_rad_outlet = res[WR_QROUT];
  wc->qm_flame = res[WR_QMF]; wc->qm_bnds = res[WR_QMB]; wc->qm_conv = res[WR_QMC];
  wc->avg_wall_T = res[WR_AVGT];
  wc->sweeps = sweeps; wc->outer = outer; wc->rel_rad_chg_last = last_mean;
  // soot calibration, C/Qwall.f90:1297-1320
  if (wc->isoot_cal == 1) {
    if (wc->q_melt < 0.0) wc->q_melt = 0.0;
    const double q_ratio = wc->q_melt / wc->q_melt_req;
    const double a_adj = fmax(fmin(q_ratio, 2.0), 0.5);
    if (a_adj > 1.0) wc->aform = wc->rf_afover * wc->aform / a_adj + (1.0 - wc->rf_afover) * wc->aform;
    else wc->aform = wc->rf_aform * wc->aform / a_adj + (1.0 - wc->rf_aform) * wc->aform;
  }
  // QRN_PRN, C/Qrn.f90:399-402
  const double g0 = c->rad->qeg_s_last * 1.0e-6;
  wc->qa_s1 = g0 - wc->qa_s0;
  wc->qa_s0 = g0;
  h->log = log;
  // the next SIMPLER iterations take the patch q_conv as the enthalpy sink
  if (c->dc.wall_qcf != d.qcf) { c->dc.wall_qcf = d.qcf; g_bound = nullptr; }
  return GFM_OK;
}

int gfm_hip_wall_download(gfm_ctx *c, const gfm_wall_ctrl *wc, gfm_wall_arrays *a) {
  if (!c || !wc || !a) { set_err("null argument"); return GFM_E_ARG; }
  if (!c->wall || !c->wall->ready) { set_err("no patch list"); return GFM_E_STATE; }
  WallHost *h = c->wall;
  WallDev &d = h->dv;
  if (wc->b_cnt != d.nb) { set_err("b_cnt %d is not the set-up %d", wc->b_cnt, d.nb); return GFM_E_ARG; }
  HCHECK(hipSetDevice(c->device));
  const size_t nb = d.nb;
  auto dn = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
    return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
  };
  HCHECK(dn(a->b_typ, d.typ, nb * 2)); HCHECK(dn(a->b_i, d.bi, nb * 2)); HCHECK(dn(a->b_j, d.bj, nb * 2));
  HCHECK(dn(a->b_k, d.bk, nb * 2)); HCHECK(dn(a->b_orient, d.orient, nb * 2)); HCHECK(dn(a->b_direct, d.direct, nb * 2));
  HCHECK(dn(a->bc_i, d.bci, nb * 2)); HCHECK(dn(a->bc_j, d.bcj, nb * 2)); HCHECK(dn(a->bc_k, d.bck, nb * 2));
  HCHECK(dn(a->b_exh, d.exh, nb * 2));
  HCHECK(dn(a->patch_area, d.area, nb * 8)); HCHECK(dn(a->b_face_area, d.face_area, nb * 8));
  HCHECK(dn(a->emis, d.emis, nb * 8)); HCHECK(dn(a->radiosity, d.jrad, nb * 8)); HCHECK(dn(a->q_amb, d.q_amb, nb * 8));
  HCHECK(dn(a->q_conv, d.q_conv, nb * 8)); HCHECK(dn(a->qc, d.qc, nb * 8)); HCHECK(dn(a->qeb, d.qeb, nb * 8));
  HCHECK(dn(a->b_t, d.bt, nb * 8)); HCHECK(dn(a->vf_sum, d.vf_sum, nb * 8));
  HCHECK(dn(a->qrs, d.qrs, (size_t)d.mp * d.np * 8));
  HCHECK(dn(a->wall_rad_t, d.wrt, (size_t)d.mp * d.np * d.lp * 8));
  if (a->vf) {
    // row-major on the device, column-major (Fortran) on the host
    std::vector<double> t(nb * nb);
    HCHECK(hipMemcpy(t.data(), d.vf, nb * nb * 8, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < nb; p++)
      for (size_t m = 0; m < nb; m++) a->vf[p + nb * m] = t[p * nb + m];
  }
  if (a->conv_log && a->log_cap > 0) {
    const size_t n = std::min((size_t)a->log_cap, h->log.size());
    for (size_t q = 0; q < n; q++) a->conv_log[q] = h->log[q];
  }
  return GFM_OK;
}

int gfm_hip_wall_upload(gfm_ctx *c, const gfm_wall_ctrl *wc, const gfm_wall_arrays *a) {
  if (!c || !wc || !a || !a->radiosity || !a->q_conv || !a->b_t) { set_err("null argument"); return GFM_E_ARG; }
  if (!c->wall || !c->wall->ready) { set_err("no patch list (gfm_hip_wall_setup first)"); return GFM_E_STATE; }
  WallDev &d = c->wall->dv;
  if (wc->b_cnt != d.nb) { set_err("b_cnt %d is not the set-up %d", wc->b_cnt, d.nb); return GFM_E_ARG; }
  HCHECK(hipSetDevice(c->device));
  const size_t nb = d.nb;
  HCHECK(hipMemcpy(d.jrad, a->radiosity, nb * 8, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(d.q_conv, a->q_conv, nb * 8, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(d.bt, a->b_t, nb * 8, hipMemcpyHostToDevice));
  if (a->q_amb) HCHECK(hipMemcpy(d.q_amb, a->q_amb, nb * 8, hipMemcpyHostToDevice));
  if (a->qc) HCHECK(hipMemcpy(d.qc, a->qc, nb * 8, hipMemcpyHostToDevice));
  if (a->qeb) HCHECK(hipMemcpy(d.qeb, a->qeb, nb * 8, hipMemcpyHostToDevice));
  if (a->wall_rad_t) HCHECK(hipMemcpy(d.wrt, a->wall_rad_t, (size_t)d.mp * d.np * d.lp * 8, hipMemcpyHostToDevice));
  // the enthalpy sink of the restored q_conv
  std::vector<double> zero;
  (void)zero;
  int rc;
  if ((rc = bind(c))) return rc;
  hipLaunchKernelGGL(k_wall_qcf, dim3((d.nb + 255) / 256), dim3(256), 0, c->stream, d);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(c->stream));
  if (c->dc.wall_qcf != d.qcf) { c->dc.wall_qcf = d.qcf; g_bound = nullptr; }
  return GFM_OK;
}

}  // extern "C"
