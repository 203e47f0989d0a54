// gfm_pnet.hip -- the U-Net of the learned p' initial guess (include/gfm_pnet.h):
// BF16 matrix-core convolution kernels (forward, input gradient, weight
// gradient), pooling/upsampling, loss and Adam, and the host driver.
// Included into gfm_hip.hip (one translation unit).
#include "../../include/gfm_pnet.h"

#include <algorithm>
#include <random>

// ---------------------------------------------------------------------
// BF16 storage: uint16 bit patterns, round to nearest even from FP32.
// ---------------------------------------------------------------------
typedef uint16_t pbf;
typedef __bf16 pbf8 __attribute__((ext_vector_type(8)));
typedef float pf4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float pbf_to_f(pbf h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ __forceinline__ pbf pbf_from_f(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (pbf)(u >> 16);
}

// Voxel grid of one level: batch, z, y, x (x fastest); v = ((b*D+z)*H+y)*W+x.
struct PGrid {
  int B, D, H, W;
};
__host__ __device__ __forceinline__ long long pgrid_cells(const PGrid &q) { return (long long)q.B * q.D * q.H * q.W; }

// Raw buffer over an activation tensor for the convolution's gathers: the
// range check (offsets from 2^31 up read as zero) stands for the padding.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pconv_rsrc(const pbf *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<pbf *>(base), (short)0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ pbf8 pconv_load(__amdgpu_buffer_rsrc_t rs, unsigned byte_off) {
  typedef unsigned pu4 __attribute__((ext_vector_type(4)));
  return __builtin_bit_cast(pbf8, (pu4)__builtin_amdgcn_raw_buffer_load_b128(rs, (int)byte_off, 0, 0));
}
// A fragment moved across the 16 lanes of each DPP row: CTRL 0x111
// (row_shr:1) gives lane l the value of lane l-1, 0x101 (row_shl:1) that of
// lane l+1, a lane whose source lies outside the row keeping `old`;
// 0x121 / 0x12F (row_ror:1 / row_ror:15) rotate: lane l gets lane (l-1)
// mod 16 / (l+1) mod 16.  Every lane of the wavefront must execute it.
template <int CTRL>
__device__ __forceinline__ pbf8 pconv_dpp(pbf8 v, pbf8 old) {
  typedef int pi4 __attribute__((ext_vector_type(4)));
  const pi4 a = __builtin_bit_cast(pi4, v), o = __builtin_bit_cast(pi4, old);
  pi4 r;
#pragma unroll
  for (int e = 0; e < 4; e++) r[e] = __builtin_amdgcn_update_dpp(o[e], a[e], CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(pbf8, r);
}

// ---------------------------------------------------------------------
// 3x3x3 (or 1-tap) convolution as an implicit GEMM on the matrix cores:
// out[v][co] = act(bias[co] + sum_{tap,ci} in(v,tap)[ci] * w[tap][co][ci]).
// The input is the channel concatenation [xa (ca channels), xb (cb)], xa
// optionally read at half resolution (nearest upsampling by 2).  Output
// channels [0, split) go to ya (row length split), the rest to yb.
// A wavefront computes MT*16 voxels x 16*NCT channels with
// v_mfma_f32_16x16x32_bf16: A = 16 voxels x 32 input channels (each lane 8
// consecutive channels of one voxel, one 16-byte load), B = 32 input
// channels x 16 output channels from w (each B fragment feeds MT MFMAs),
// D rows = voxels, columns = output channels.  256 threads = 4 wavefronts:
// with KS = 1 each takes its own MT*16 voxels (64*MT per workgroup); with
// KS = 4, for the coarse grids, the four share one tile of MT*16 voxels and
// each runs a quarter of the (tap, chunk) steps, their sums added through
// LDS in a fixed order.
// ---------------------------------------------------------------------
template <int NCT, int MT, int KS>
__global__ __launch_bounds__(256, NCT <= 2 ? 4 : 2) void k_pconv(const pbf *__restrict__ xa, int ca, int up_a,
                                                                  const pbf *__restrict__ xb, int cb,
                                                                  const pbf *__restrict__ w, const float *__restrict__ bias,
                                                                  pbf *__restrict__ ya, int split, pbf *__restrict__ yb,
                                                                  int relu, PGrid pg, int ntap) {
  static_assert(KS == 1 || KS == 4, "steps split over 1 or 4 wavefronts");
  constexpr int CO = NCT * 16, VG = 4 / KS;  // VG voxel tiles per workgroup
  const int cin = ca + cb;
  const unsigned V = (unsigned)pgrid_cells(pg);  // pconv() keeps V and every byte offset below 2^31
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r16 = l & 15, kg = l >> 4;
  const int kw = wave % KS;  // which part of the steps this wavefront runs
  const unsigned vb = blockIdx.x * (VG * 16u * MT) + (wave / KS) * (16u * MT);
  const int Dh = pg.D >> 1, Hh = pg.H >> 1, Wh = pg.W >> 1;
  // Per voxel tile, computed once: the lane's byte offsets of its own voxel
  // in xa and xb (its 8-channel group included), the inverted in-grid mask
  // of the 27 taps (bit t, t = 9(dz+1) + 3(dy+1) + dx+1) and, for a
  // half-resolution xa, the parities of z, y, x.  Inside the tap loop a
  // neighbour then costs a uniform base shift plus, for the lane, an OR of
  // its mask bit into bit 31: a buffer load past the descriptor's range
  // returns zero, which is the zero padding of the convolution.
  unsigned ofa[MT], ofb[MT], ntm[MT];
  bool pz[MT], py[MT], px[MT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    const unsigned v = vb + mt * 16 + r16;
    const bool in = v < V;
    unsigned r = in ? v : 0;
    const int x = (int)(r % (unsigned)pg.W); r /= (unsigned)pg.W;
    const int y = (int)(r % (unsigned)pg.H); r /= (unsigned)pg.H;
    const int z = (int)(r % (unsigned)pg.D);
    const int bt = (int)(r / (unsigned)pg.D);
    const unsigned vo = in ? v : 0;
    ofb[mt] = (vo * (unsigned)cb + 8u * kg) * 2u;
    const unsigned va = up_a ? (((unsigned)(bt * Dh + (z >> 1)) * Hh + (y >> 1)) * Wh + (x >> 1)) : vo;
    ofa[mt] = (va * (unsigned)ca + 8u * kg) * 2u;
    pz[mt] = z & 1; py[mt] = y & 1; px[mt] = x & 1;
    const unsigned mx = (x > 0) | 2u | ((unsigned)(x < pg.W - 1) << 2);
    const unsigned my = (y > 0) | 2u | ((unsigned)(y < pg.H - 1) << 2);
    const unsigned mz = (z > 0) | 2u | ((unsigned)(z < pg.D - 1) << 2);
    const unsigned ey = ((my & 1u) ? 7u : 0u) | ((my & 2u) ? 070u : 0u) | ((my & 4u) ? 0700u : 0u);
    const unsigned ez = ((mz & 1u) ? 0x1FFu : 0u) | ((mz & 2u) ? 0x1FFu << 9 : 0u) | ((mz & 4u) ? 0x1FFu << 18 : 0u);
    const unsigned tm = (ntap == 1) ? 1u : (mx * 01111111111u) & (ey * 0x40201u) & ez;
    ntm[mt] = in ? ~tm : ~0u;
  }
  pf4 acc[MT][NCT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) acc[mt][ct] = pf4{0.f, 0.f, 0.f, 0.f};
  // byte strides of a half-resolution xa, and the lane's offset in w
  const int sza = Hh * Wh * ca * 2, sya = Wh * ca * 2, sxa = ca * 2;
  const unsigned wl = (unsigned)(r16 * cin + 8 * kg) * 2u;
  // One step per (tap, 32-channel chunk).  The fragments of the next step
  // are loaded before the MFMAs of the current one, so a wavefront keeps
  // its gathers in flight while its matrix cores work.
  const int nch = cin / 32, nstep = ntap * nch;
  auto fetch = [&](int st, pbf8 (&A)[MT], pbf8 (&Bf)[NCT]) {
    const int t = st / nch, ch = (st - t * nch) * 32;
    const int dz = (ntap == 1) ? 0 : t / 9 - 1, dy = (ntap == 1) ? 0 : (t / 3) % 3 - 1, dx = (ntap == 1) ? 0 : t % 3 - 1;
    const long long toff = ((long long)dz * pg.H + dy) * pg.W + dx;  // neighbour shift in voxels
    if (ch < ca && up_a) {
      const __amdgpu_buffer_rsrc_t rs = pconv_rsrc(xa + ch);
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        // floor((z+dz)/2) - floor(z/2) is -1, 0 or 1 by dz and the parity of z
        const int oz = dz == 0 ? 0 : dz < 0 ? (pz[mt] ? 0 : -sza) : (pz[mt] ? sza : 0);
        const int oy = dy == 0 ? 0 : dy < 0 ? (py[mt] ? 0 : -sya) : (py[mt] ? sya : 0);
        const int ox = dx == 0 ? 0 : dx < 0 ? (px[mt] ? 0 : -sxa) : (px[mt] ? sxa : 0);
        A[mt] = pconv_load(rs, (ofa[mt] + (unsigned)(oz + oy + ox)) | ((ntm[mt] >> t) << 31));
      }
    } else {
      const bool fa = ch < ca;
      const __amdgpu_buffer_rsrc_t rs = fa ? pconv_rsrc(xa + toff * ca + ch) : pconv_rsrc(xb + toff * cb + (ch - ca));
#pragma unroll
      for (int mt = 0; mt < MT; mt++) A[mt] = pconv_load(rs, (fa ? ofa[mt] : ofb[mt]) | ((ntm[mt] >> t) << 31));
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
      Bf[ct] = *reinterpret_cast<const pbf8 *>(reinterpret_cast<const char *>(w + ((long long)t * CO + ct * 16) * cin + ch) + wl);
  };
  auto mma = [&](const pbf8 (&A)[MT], const pbf8 (&Bf)[NCT]) {
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
#pragma unroll
      for (int mt = 0; mt < MT; mt++) acc[mt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt], Bf[ct], acc[mt][ct], 0, 0, 0);
  };
  // When the wavefront's 16*MT voxels lie in one x-row (W a multiple of
  // 16*MT: the tiles start at multiples of 16*MT), the three x-taps of a
  // (dz, dy) row share their voxels: one step per (row tap, chunk) loads the
  // centre fragments (dx = 0) and one edge fragment, whose lane 0 holds
  // voxel x-1 of the first tile and lane 15 voxel x+1 of the last (the other
  // lanes read past the range: no traffic), and forms dx = -1 and +1 by
  // moving the centre fragments one lane within each 16-lane row, the lane
  // at a tile's edge taking the neighbouring tile's edge lane.  A voxel's
  // x-1 and x+1 then come from the lanes that loaded them with the in-grid
  // test of their own dx = 0 tap, which is the test of this voxel's dx = -1
  // or +1 tap; the first and last voxels of a row are edge-fragment lanes,
  // tested by their own tap bits.
  const bool xrow = ntap == 27 && pg.W % (16 * MT) == 0;
  if (xrow) {
    auto fetch_row = [&](int st, pbf8 (&C)[MT], pbf8 &E) {
      const int rt = st / nch, ch = (st - rt * nch) * 32;
      const int dz = rt / 3 - 1, dy = rt % 3 - 1, t0 = 3 * rt;  // taps t0, t0+1, t0+2: dx = -1, 0, +1
      const long long toff = ((long long)dz * pg.H + dy) * pg.W;
      const bool el = r16 == 0, er = r16 == 15;
      if (ch < ca && up_a) {
        const __amdgpu_buffer_rsrc_t rs = pconv_rsrc(xa + ch);
        // the row's voxels share z and y, so their parities
        const int oz = dz == 0 ? 0 : dz < 0 ? (pz[0] ? 0 : -sza) : (pz[0] ? sza : 0);
        const int oy = dy == 0 ? 0 : dy < 0 ? (py[0] ? 0 : -sya) : (py[0] ? sya : 0);
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
          C[mt] = pconv_load(rs, (ofa[mt] + (unsigned)(oz + oy)) | ((ntm[mt] >> (t0 + 1)) << 31));
        const int ox = el ? (px[0] ? 0 : -sxa) : (px[MT - 1] ? sxa : 0);
        const unsigned eo = (el ? ofa[0] : ofa[MT - 1]) + (unsigned)(oz + oy + ox);
        const unsigned eb = el ? (ntm[0] >> t0) : (ntm[MT - 1] >> (t0 + 2));
        E = pconv_load(rs, (el || er) ? (eo | (eb << 31)) : 0x80000000u);
      } else {
        const bool fa = ch < ca;
        const __amdgpu_buffer_rsrc_t rs = fa ? pconv_rsrc(xa + toff * ca + ch) : pconv_rsrc(xb + toff * cb + (ch - ca));
        const unsigned cs = (unsigned)(fa ? ca : cb) * 2u;  // bytes per voxel
#pragma unroll
        for (int mt = 0; mt < MT; mt++) C[mt] = pconv_load(rs, (fa ? ofa[mt] : ofb[mt]) | ((ntm[mt] >> (t0 + 1)) << 31));
        // voxel x-1 of the first voxel of a row is out of the grid, so its
        // tap bit is set whenever the offset below would wrap
        const unsigned eo = el ? (fa ? ofa[0] : ofb[0]) - cs : (fa ? ofa[MT - 1] : ofb[MT - 1]) + cs;
        const unsigned eb = el ? (ntm[0] >> t0) : (ntm[MT - 1] >> (t0 + 2));
        E = pconv_load(rs, (el || er) ? (eo | (eb << 31)) : 0x80000000u);
      }
    };
    // the weights of a step's three taps (the same for every wavefront,
    // mostly L1 hits) are loaded tap by tap when the step runs
    auto mma_row = [&](int st, const pbf8 (&C)[MT], const pbf8 &E) {
      const int rt = st / nch, ch = (st - rt * nch) * 32, t0 = 3 * rt;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        pbf8 Bw[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
          Bw[ct] = *reinterpret_cast<const pbf8 *>(
              reinterpret_cast<const char *>(w + ((long long)(t0 + d) * CO + ct * 16) * cin + ch) + wl);
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
          pbf8 A = C[mt];
          if (d == 0) {
            const pbf8 lo = mt == 0 ? E : pconv_dpp<0x121>(C[mt > 0 ? mt - 1 : 0], C[mt > 0 ? mt - 1 : 0]);
            A = pconv_dpp<0x111>(C[mt], lo);
          } else if (d == 2) {
            const pbf8 hi = mt == MT - 1 ? E : pconv_dpp<0x12F>(C[mt < MT - 1 ? mt + 1 : mt], C[mt < MT - 1 ? mt + 1 : mt]);
            A = pconv_dpp<0x101>(C[mt], hi);
          }
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) acc[mt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bw[ct], acc[mt][ct], 0, 0, 0);
        }
      }
    };
    const int nrs = 9 * nch, rs1 = (kw + 1) * nrs / KS;
    int s = kw * nrs / KS;
    pbf8 C0[MT], E0, C1[MT], E1;
    fetch_row(s, C0, E0);
    for (; s + 1 < rs1; s += 2) {
      fetch_row(s + 1, C1, E1);
      mma_row(s, C0, E0);
      if (s + 2 < rs1) fetch_row(s + 2, C0, E0);
      mma_row(s + 1, C1, E1);
    }
    if (s < rs1) mma_row(s, C0, E0);
  } else {
  // two fragment sets used in turn (no register copies between steps)
  pbf8 A0[MT], B0[NCT], A1[MT], B1[NCT];
  const int st1 = (kw + 1) * nstep / KS;
  int st = kw * nstep / KS;
  fetch(st, A0, B0);
  for (; st + 1 < st1; st += 2) {
    fetch(st + 1, A1, B1);
    mma(A0, B0);
    if (st + 2 < st1) fetch(st + 2, A0, B0);
    mma(A1, B1);
  }
  if (st < st1) mma(A0, B0);
  }
  if constexpr (KS > 1) {
    __shared__ pf4 red[VG][KS - 1][MT * NCT][64];
    if (kw > 0) {
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) red[wave / KS][kw - 1][mt * NCT + ct][l] = acc[mt][ct];
    }
    __syncthreads();
    if (kw > 0) return;
#pragma unroll
    for (int j = 0; j < KS - 1; j++)
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[mt][ct] += red[wave / KS][j][mt * NCT + ct][l];
  }
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int co = ct * 16 + r16;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const unsigned v = vb + mt * 16 + 4 * kg + e;
        if (v >= V) continue;
        float val = acc[mt][ct][e] + bv;
        if (relu) val = fmaxf(val, 0.f);
        if (co < split) ya[(size_t)v * split + co] = pbf_from_f(val);
        else yb[(size_t)v * (CO - split) + (co - split)] = pbf_from_f(val);
      }
    }
}

// ---------------------------------------------------------------------
// Weight gradient: dw[tap][co][ci] = sum_v dy[v][co] * in(v,tap)[ci] as a
// GEMM over voxels.  A workgroup takes one tap, a 32x32 (co, ci) block and
// a chunk of voxels; 32 voxels at a time are staged through LDS transposed
// ([channel][voxel]) so each wavefront reads its 16x32 fragments with
// 16-byte LDS loads and issues one MFMA per stage.  Partial sums per chunk
// go to part[chunk][tap][co][ci] and are reduced in order afterwards.
// ---------------------------------------------------------------------
#define PW_PITCH 40  // LDS row pitch in bf16 (32 voxels + 8)
__global__ __launch_bounds__(256) void k_pconv_wgrad(const pbf *__restrict__ dy, int co_all, const pbf *__restrict__ xa,
                                                     int ca, int up_a, const pbf *__restrict__ xb, int cb, PGrid pg,
                                                     int ntap, long long vchunk, float *__restrict__ part) {
  __shared__ __attribute__((aligned(16))) pbf sdy[32 * PW_PITCH];
  __shared__ __attribute__((aligned(16))) pbf sx[32 * PW_PITCH];
  const int cin = ca + cb;
  const int nci = cin / 32;
  const int co0 = (blockIdx.z / nci) * 32, ci0 = (blockIdx.z % nci) * 32;
  const int t = blockIdx.y;
  const long long V = pgrid_cells(pg);
  const long long v_beg = (long long)blockIdx.x * vchunk;
  long long v_end = v_beg + vchunk;
  if (v_end > V) v_end = V;
  const int dz = (ntap == 1) ? 0 : t / 9 - 1, dyy = (ntap == 1) ? 0 : (t / 3) % 3 - 1, dx = (ntap == 1) ? 0 : t % 3 - 1;
  const int Dh = pg.D >> 1, Hh = pg.H >> 1, Wh = pg.W >> 1;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r16 = l & 15, kg = l >> 4;
  const int wm = wave >> 1, wn = wave & 1;
  // staging: thread -> voxel sv (0..31), channel quad sq (0..7): 4 channels
  const int sv = threadIdx.x >> 3, sq = threadIdx.x & 7;
  const bool fa = ci0 < ca;
  pf4 acc = pf4{0.f, 0.f, 0.f, 0.f};
  for (long long v0 = v_beg; v0 < v_end; v0 += 32) {
    const long long v = v0 + sv;
    uint2 gdy = make_uint2(0, 0), gx = make_uint2(0, 0);
    if (v < v_end) {
      gdy = *reinterpret_cast<const uint2 *>(dy + v * co_all + co0 + 4 * sq);
      long long r = v;
      const int x = (int)(r % pg.W) + dx; r /= pg.W;
      const int y = (int)(r % pg.H) + dyy; r /= pg.H;
      const int z = (int)(r % pg.D) + dz;
      const long long bt = r / pg.D;
      if (z >= 0 && z < pg.D && y >= 0 && y < pg.H && x >= 0 && x < pg.W) {
        const long long n = fa ? (up_a ? ((bt * Dh + (z >> 1)) * Hh + (y >> 1)) * Wh + (x >> 1) : ((bt * pg.D + z) * pg.H + y) * pg.W + x)
                               : ((bt * pg.D + z) * pg.H + y) * pg.W + x;
        const pbf *src = fa ? xa + n * ca + ci0 : xb + n * cb + (ci0 - ca);
        gx = *reinterpret_cast<const uint2 *>(src + 4 * sq);
      }
    }
    __syncthreads();  // previous stage's fragments are read
    const pbf *pdy = reinterpret_cast<const pbf *>(&gdy), *px = reinterpret_cast<const pbf *>(&gx);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      sdy[(4 * sq + e) * PW_PITCH + sv] = pdy[e];
      sx[(4 * sq + e) * PW_PITCH + sv] = px[e];
    }
    __syncthreads();
    const pbf8 A = *reinterpret_cast<const pbf8 *>(&sdy[(wm * 16 + r16) * PW_PITCH + 8 * kg]);
    const pbf8 Bf = *reinterpret_cast<const pbf8 *>(&sx[(wn * 16 + r16) * PW_PITCH + 8 * kg]);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bf, acc, 0, 0, 0);
  }
  float *out = part + (((long long)blockIdx.x * ntap + t) * co_all) * cin;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int co = co0 + wm * 16 + 4 * kg + e, ci = ci0 + wn * 16 + r16;
    out[(long long)co * cin + ci] = acc[e];
  }
}
// dst[i] = sum over chunks of part[chunk][i], chunk order fixed.
__global__ void k_psum_chunks(const float *__restrict__ part, long long n, int nchunk, float *__restrict__ dst) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += part[(long long)c * n + i];
  dst[i] = s;
}

// bias gradient: db[c] = sum_v dy[v][c]; one workgroup per channel.
__global__ void k_pbias_grad(const pbf *__restrict__ dy, int C, long long V, float *__restrict__ db) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (long long v = threadIdx.x; v < V; v += blockDim.x) s += pbf_to_f(dy[v * C + c]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

// ---------------------------------------------------------------------
// elementwise and resampling kernels (C channels per voxel)
// ---------------------------------------------------------------------
// coarse[v'][c] = mean of the 8 fine children
__global__ void k_ppool(const pbf *__restrict__ fine, pbf *__restrict__ coarse, int C, PGrid cg) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = pgrid_cells(cg) * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long r = i / C;
  const int x = (int)(r % cg.W); r /= cg.W;
  const int y = (int)(r % cg.H); r /= cg.H;
  const int z = (int)(r % cg.D);
  const long long b = r / cg.D;
  const int D = 2 * cg.D, H = 2 * cg.H, W = 2 * cg.W;
  float s = 0.f;
  for (int q = 0; q < 8; q++) {
    const long long f = ((b * D + 2 * z + (q >> 2)) * H + 2 * y + ((q >> 1) & 1)) * W + 2 * x + (q & 1);
    s += pbf_to_f(fine[f * C + c]);
  }
  coarse[i] = pbf_from_f(s * 0.125f);
}
// fine gradient of the average pool, added to `acc` (acc = round(acc + g/8))
__global__ void k_ppool_bwd_add(const pbf *__restrict__ gcoarse, pbf *__restrict__ acc, int C, PGrid fg) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = pgrid_cells(fg) * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long r = i / C;
  const int x = (int)(r % fg.W); r /= fg.W;
  const int y = (int)(r % fg.H); r /= fg.H;
  const int z = (int)(r % fg.D);
  const long long b = r / fg.D;
  const long long cv = ((b * (fg.D >> 1) + (z >> 1)) * (fg.H >> 1) + (y >> 1)) * (fg.W >> 1) + (x >> 1);
  acc[i] = pbf_from_f(pbf_to_f(acc[i]) + pbf_to_f(gcoarse[cv * C + c]) * 0.125f);
}
// coarse gradient of nearest upsampling: sum of the 8 fine children
__global__ void k_punup(const pbf *__restrict__ gfine, pbf *__restrict__ gcoarse, int C, PGrid cg) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = pgrid_cells(cg) * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long r = i / C;
  const int x = (int)(r % cg.W); r /= cg.W;
  const int y = (int)(r % cg.H); r /= cg.H;
  const int z = (int)(r % cg.D);
  const long long b = r / cg.D;
  const int D = 2 * cg.D, H = 2 * cg.H, W = 2 * cg.W;
  float s = 0.f;
  for (int q = 0; q < 8; q++) {
    const long long f = ((b * D + 2 * z + (q >> 2)) * H + 2 * y + ((q >> 1) & 1)) * W + 2 * x + (q & 1);
    s += pbf_to_f(gfine[f * C + c]);
  }
  gcoarse[i] = pbf_from_f(s);
}
// g = g * (y > 0)
__global__ void k_prelu_bwd(pbf *__restrict__ gr, const pbf *__restrict__ y, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n && !(pbf_to_f(y[i]) > 0.f)) gr[i] = 0;
}
// network input: [v][32] bf16 from x[v][2] (float), channels 2..31 zero;
// one thread writes one voxel's 64 bytes
__global__ void k_pinput(const float *__restrict__ x, pbf *__restrict__ x0, long long V) {
  long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= V) return;
  const float2 in = reinterpret_cast<const float2 *>(x)[v];
  uint4 *o = reinterpret_cast<uint4 *>(x0 + v * 32);
  o[0] = make_uint4((uint32_t)pbf_from_f(in.x) | ((uint32_t)pbf_from_f(in.y) << 16), 0u, 0u, 0u);
  o[1] = make_uint4(0u, 0u, 0u, 0u);
  o[2] = make_uint4(0u, 0u, 0u, 0u);
  o[3] = make_uint4(0u, 0u, 0u, 0u);
}
// output layer y[v] = w . d0[v] + b (FP32); C is a multiple of 8, read 16
// bytes at a time
__global__ void k_pout(const pbf *__restrict__ d0, int C, const float *__restrict__ wb, float *__restrict__ y, long long V) {
  long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint4 *row = reinterpret_cast<const uint4 *>(d0 + v * C);
  float s = 0.f;
  for (int q = 0; q < C / 8; q++) {
    const uint4 u = row[q];
    const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int h = 0; h < 4; h++) {
      s += wb[8 * q + 2 * h] * pbf_to_f((pbf)(wd[h] & 0xFFFFu));
      s += wb[8 * q + 2 * h + 1] * pbf_to_f((pbf)(wd[h] >> 16));
    }
  }
  y[v] = s + wb[C];
}
// loss terms: dy = 2 m (y - t) / n_open, per-block partial sums of m (y-t)^2
__global__ void k_ploss(const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ x, long long V,
                        float inv_n, float *__restrict__ dy, double *__restrict__ part) {
  long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double s = 0.0;
  if (v < V) {
    const float m = x[2 * v + 1];
    const float r = m > 0.f ? y[v] - t[v] : 0.f;
    dy[v] = 2.f * r * inv_n;
    s = (double)r * r;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// gradient of the output layer: g_d0[v][c] = dy[v] w[c] (bf16) and the
// per-block partials of dw[c] = sum dy d0[v][c], db = sum dy
__global__ void k_pout_bwd(const float *__restrict__ dy, const pbf *__restrict__ d0, int C, const float *__restrict__ wb,
                           pbf *__restrict__ gd0, float *__restrict__ part, long long V) {
  long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  __shared__ float red[256];
  for (int c = 0; c <= C; c++) {
    float s = 0.f;
    if (v < V) {
      if (c < C) {
        s = dy[v] * pbf_to_f(d0[v * C + c]);
        gd0[v * C + c] = pbf_from_f(dy[v] * wb[c]);
      } else {
        s = dy[v];
      }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)blockIdx.x * (C + 1) + c] = red[0];
    __syncthreads();
  }
}
// sum of per-block partials: dst[i] = sum_b part[b*n + i]
__global__ void k_psum_blocks(const float *__restrict__ part, int nb, int n, float *__restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int b = 0; b < nb; b++) s += part[(long long)b * n + i];
  dst[i] = s;
}
__global__ void k_padam(float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m, float *__restrict__ v,
                        const uint8_t *__restrict__ frozen, long long n, float lr, float b1, float b2, float eps, float c1,
                        float c2) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n || frozen[i]) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// bf16 weights of one conv: wf[t][co][ci] (forward) and wt[26-t][ci][co]
// (input gradient: the transposed, tap-mirrored kernel)
__global__ void k_pcast(const float *__restrict__ p, int ntap, int co, int ci, pbf *__restrict__ wf, pbf *__restrict__ wt) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = (long long)ntap * co * ci;
  if (i >= n) return;
  const int c_in = (int)(i % ci);
  const int c_out = (int)((i / ci) % co);
  const int t = (int)(i / ((long long)ci * co));
  const pbf h = pbf_from_f(p[i]);
  wf[i] = h;
  if (wt) wt[((long long)(ntap - 1 - t) * ci + c_in) * co + c_out] = h;
}

// ---------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------
struct PConv {
  int ca, cb, co, up_a;  // input channels (xa, xb), output channels, xa upsampled
  int ntap;
  size_t off_w, off_b;   // offsets in the flat parameter vector
  pbf *wf = nullptr, *wt = nullptr;
};
struct PLevel {
  PGrid pg;
  long long V;
};

struct gfm_pnet {
  int device = 0;
  int ch[4];
  int max_batch = 0;
  long long max_cells = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  PConv conv[10];
  size_t n_params = 0, n_eff = 0, off_out = 0;
  float *p = nullptr, *gr = nullptr, *m = nullptr, *v = nullptr;
  uint8_t *frozen = nullptr;
  int adam_t = 0;
  // activations (bf16, [v][c]) and gradients
  pbf *x0, *a0, *s0, *p1, *a1, *s1, *p2, *a2, *s2, *p3, *s3, *d2, *d1, *d0;
  pbf *gs0, *gs1, *gs2;    // gradients of the skip outputs s0..s2
  pbf *ga, *gb, *gt;       // gradient scratch, V0 x 128 each
  float *xin, *tin, *y, *dy, *fpart;
  double *lpart;
  float *wpart;
  long long wpart_n = 0;
  // training set on the device
  float *set_x = nullptr, *set_t = nullptr;
  long long set_n = 0, set_cells = 0;
};

template <typename T>
static int palloc(gfm_pnet *n, T **p, size_t cnt) {
  void *q = nullptr;
  HCHECK(hipMalloc(&q, cnt * sizeof(T) + 16));
  HCHECK(hipMemset(q, 0, cnt * sizeof(T) + 16));
  n->allocs.push_back(q);
  *p = (T *)q;
  return GFM_OK;
}
static unsigned pblocks(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// channel widths of the ten convolutions
static void pnet_layout(gfm_pnet *n) {
  const int *c = n->ch;
  const int spec[10][4] = {{32, 0, c[0], 0},    {c[0], 0, c[0], 0}, {c[0], 0, c[1], 0},    {c[1], 0, c[1], 0},
                           {c[1], 0, c[2], 0},  {c[2], 0, c[2], 0}, {c[2], 0, c[3], 0},    {c[3], c[2], c[2], 1},
                           {c[2], c[1], c[1], 1}, {c[1], c[0], c[0], 1}};
  size_t off = 0, eff = 0;
  for (int l = 0; l < 10; l++) {
    PConv &q = n->conv[l];
    q.ca = spec[l][0]; q.cb = spec[l][1]; q.co = spec[l][2]; q.up_a = spec[l][3]; q.ntap = 27;
    q.off_w = off;
    off += (size_t)27 * q.co * (q.ca + q.cb);
    q.off_b = off;
    off += q.co;
    eff += (size_t)27 * q.co * (l == 0 ? 2 : q.ca + q.cb) + q.co;
  }
  n->off_out = off;
  off += c[0] + 1;
  eff += c[0] + 1;
  n->n_params = off;
  n->n_eff = eff;
}

static int pnet_cast(gfm_pnet *n) {
  for (int l = 0; l < 10; l++) {
    PConv &q = n->conv[l];
    const long long cnt = (long long)q.ntap * q.co * (q.ca + q.cb);
    hipLaunchKernelGGL(k_pcast, dim3(pblocks(cnt)), dim3(256), 0, n->stream, n->p + q.off_w, q.ntap, q.co, q.ca + q.cb, q.wf,
                       l == 0 ? nullptr : q.wt);
  }
  HCHECK(hipGetLastError());
  return GFM_OK;
}

// The voxel tile by grid size, so that a launch has about 1,024
// workgroups or more: MT = 4 (accumulators near 64 registers) on the fine
// grids, MT = 2 or 1 on the middle ones, and on the coarsest MT = 1 with
// the steps split over the four wavefronts of a workgroup.  The
// environment variable GFM_PCONV_TILE = 4, 2, 1 or 1x4 fixes the choice
// for every grid (4 means 2 for 96 and 128 output channels).  The choices
// differ only in FP32 summation order: with KS = 4 four partial sums over
// the steps are added at the end.
static int pconv_tile_env() {
  const char *e = getenv("GFM_PCONV_TILE");
  if (!e || !*e || !strcmp(e, "auto")) return 0;
  if (!strcmp(e, "4")) return 4;
  if (!strcmp(e, "2")) return 2;
  if (!strcmp(e, "1")) return 1;
  if (!strcmp(e, "1x4")) return 14;
  return -1;
}
template <int NCT, int MT, int KS>
static void pconv_launch1(hipStream_t s, const pbf *xa, int ca, int up_a, const pbf *xb, int cb, const pbf *w,
                          const float *bias, pbf *ya, int split, pbf *yb, int relu, PGrid pg, int ntap) {
  const long long V = pgrid_cells(pg);
  hipLaunchKernelGGL((k_pconv<NCT, MT, KS>), dim3(pblocks(V, 16 * MT * (4 / KS))), dim3(256), 0, s, xa, ca, up_a, xb, cb,
                     w, bias, ya, split, yb, relu, pg, ntap);
}
template <int NCT, int MT>
static void pconv_launch(hipStream_t s, const pbf *xa, int ca, int up_a, const pbf *xb, int cb, const pbf *w, const float *bias,
                         pbf *ya, int split, pbf *yb, int relu, PGrid pg, int ntap) {
  const long long V = pgrid_cells(pg);
  const int tile = pconv_tile_env();
  if (tile == 4 || (tile == 0 && V >= 1024LL * 64 * MT))
    pconv_launch1<NCT, MT, 1>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap);
  else if (tile == 2 || (tile == 0 && V >= 1024LL * 64 * 2))
    pconv_launch1<NCT, 2, 1>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap);
  else if (tile == 1 || (tile == 0 && V >= 1024LL * 64))
    pconv_launch1<NCT, 1, 1>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap);
  else
    pconv_launch1<NCT, 1, 4>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap);
}
static int pconv(hipStream_t s, const pbf *xa, int ca, int up_a, const pbf *xb, int cb, const pbf *w, const float *bias,
                 pbf *ya, int split, pbf *yb, int co, int relu, PGrid pg, int ntap) {
  // the kernel's gathers use 32-bit byte offsets below 2^31
  const long long V = pgrid_cells(pg);
  if (V * 2 * std::max(std::max(ca, cb), 1) >= (1LL << 31)) {
    set_err("pnet: %lld voxels x %d channels exceed the convolution's 2 GiB tensor limit", V, std::max(ca, cb));
    return GFM_E_ARG;
  }
  if (pconv_tile_env() < 0) {
    set_err("pnet: GFM_PCONV_TILE must be auto, 4, 2, 1 or 1x4");
    return GFM_E_ARG;
  }
  switch (co) {
    case 32: pconv_launch<2, 4>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap); break;
    case 64: pconv_launch<4, 4>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap); break;
    case 96: pconv_launch<6, 2>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap); break;
    case 128: pconv_launch<8, 2>(s, xa, ca, up_a, xb, cb, w, bias, ya, split, yb, relu, pg, ntap); break;
    default: set_err("pnet: unsupported channel count %d", co); return GFM_E_ARG;
  }
  return GFM_OK;
}

static void plevels(const gfm_pnet *, int batch, int d, int h, int w, PLevel L[4]) {
  for (int q = 0; q < 4; q++) {
    L[q].pg = PGrid{batch, d >> q, h >> q, w >> q};
    L[q].V = pgrid_cells(L[q].pg);
  }
}

// forward pass from n->xin (x[v][2], device) to n->y (device) on stream s
static int pnet_forward_dev(gfm_pnet *n, hipStream_t s, int batch, int d, int h, int w) {
  PLevel L[4];
  plevels(n, batch, d, h, w, L);
  const int *c = n->ch;
  PConv *k = n->conv;
  const float *P = n->p;
  int rc;
  hipLaunchKernelGGL(k_pinput, dim3(pblocks(L[0].V)), dim3(256), 0, s, n->xin, n->x0, L[0].V);
#define PCONV(l, xa, up, xb, ya, pg)                                                                                \
  if ((rc = pconv(s, xa, k[l].ca, up, xb, k[l].cb, k[l].wf, P + k[l].off_b, ya, k[l].co, nullptr, k[l].co, 1, pg, 27))) \
    return rc;
  PCONV(0, n->x0, 0, nullptr, n->a0, L[0].pg)
  PCONV(1, n->a0, 0, nullptr, n->s0, L[0].pg)
  hipLaunchKernelGGL(k_ppool, dim3(pblocks(L[1].V * c[0])), dim3(256), 0, s, n->s0, n->p1, c[0], L[1].pg);
  PCONV(2, n->p1, 0, nullptr, n->a1, L[1].pg)
  PCONV(3, n->a1, 0, nullptr, n->s1, L[1].pg)
  hipLaunchKernelGGL(k_ppool, dim3(pblocks(L[2].V * c[1])), dim3(256), 0, s, n->s1, n->p2, c[1], L[2].pg);
  PCONV(4, n->p2, 0, nullptr, n->a2, L[2].pg)
  PCONV(5, n->a2, 0, nullptr, n->s2, L[2].pg)
  hipLaunchKernelGGL(k_ppool, dim3(pblocks(L[3].V * c[2])), dim3(256), 0, s, n->s2, n->p3, c[2], L[3].pg);
  PCONV(6, n->p3, 0, nullptr, n->s3, L[3].pg)
  PCONV(7, n->s3, 1, n->s2, n->d2, L[2].pg)
  PCONV(8, n->d2, 1, n->s1, n->d1, L[1].pg)
  PCONV(9, n->d1, 1, n->s0, n->d0, L[0].pg)
#undef PCONV
  hipLaunchKernelGGL(k_pout, dim3(pblocks(L[0].V)), dim3(256), 0, s, n->d0, c[0], P + n->off_out, n->y, L[0].V);
  HCHECK(hipGetLastError());
  return GFM_OK;
}

// weight and bias gradients of conv l from its output gradient gy (after
// the ReLU mask) and its inputs
static int pnet_wgrad(gfm_pnet *n, hipStream_t s, int l, const pbf *gy, const pbf *xa, const pbf *xb, PGrid pg) {
  PConv &q = n->conv[l];
  const long long V = pgrid_cells(pg);
  const int cin = q.ca + q.cb;
  long long vchunk = 4096;
  int nchunk = (int)((V + vchunk - 1) / vchunk);
  const long long per = 27LL * q.co * cin;
  while ((long long)nchunk * per > n->wpart_n) {
    vchunk *= 2;
    nchunk = (int)((V + vchunk - 1) / vchunk);
  }
  hipLaunchKernelGGL(k_pconv_wgrad, dim3(nchunk, 27, (q.co / 32) * (cin / 32)), dim3(256), 0, s, gy, q.co, xa, q.ca, q.up_a,
                     xb, q.cb, pg, 27, vchunk, n->wpart);
  hipLaunchKernelGGL(k_psum_chunks, dim3(pblocks(per)), dim3(256), 0, s, n->wpart, per, nchunk, n->gr + q.off_w);
  hipLaunchKernelGGL(k_pbias_grad, dim3(q.co), dim3(256), 0, s, gy, q.co, V, n->gr + q.off_b);
  HCHECK(hipGetLastError());
  return GFM_OK;
}

// loss and gradients for the batch in n->xin / n->tin (inv_n = 1 / open
// cells of the batch); the loss partials are left in n->lpart.
static int pnet_loss_grad_dev(gfm_pnet *n, hipStream_t s, int batch, int d, int h, int w, float inv_n) {
  int rc;
  if ((rc = pnet_forward_dev(n, s, batch, d, h, w))) return rc;
  PLevel L[4];
  plevels(n, batch, d, h, w, L);
  const int *c = n->ch;
  PConv *k = n->conv;
  const long long V0 = L[0].V;
  const unsigned nb0 = pblocks(V0);
  hipLaunchKernelGGL(k_ploss, dim3(nb0), dim3(256), 0, s, n->y, n->tin, n->xin, V0, inv_n, n->dy, n->lpart);
  hipLaunchKernelGGL(k_pout_bwd, dim3(nb0), dim3(256), 0, s, n->dy, n->d0, c[0], n->p + n->off_out, n->ga, n->fpart, V0);
  hipLaunchKernelGGL(k_psum_blocks, dim3(1), dim3(256), 0, s, n->fpart, (int)nb0, c[0] + 1, n->gr + n->off_out);
  auto mask = [&](pbf *gr, const pbf *y, long long cnt) {
    hipLaunchKernelGGL(k_prelu_bwd, dim3(pblocks(cnt)), dim3(256), 0, s, gr, y, cnt);
  };
  // input gradient of conv l: out channels ca+cb split at ca
  auto bwd = [&](int l, const pbf *gy, pbf *ya, pbf *yb, PGrid pg) {
    return pconv(s, gy, k[l].co, 0, nullptr, 0, k[l].wt, nullptr, ya, k[l].ca, yb, k[l].ca + k[l].cb, 0, pg, 27);
  };
  auto unup = [&](const pbf *fine, pbf *coarse, int C, const PLevel &cl) {
    hipLaunchKernelGGL(k_punup, dim3(pblocks(cl.V * C)), dim3(256), 0, s, fine, coarse, C, cl.pg);
  };
  auto poolb = [&](const pbf *gcoarse, pbf *acc, int C, const PLevel &fl) {
    hipLaunchKernelGGL(k_ppool_bwd_add, dim3(pblocks(fl.V * C)), dim3(256), 0, s, gcoarse, acc, C, fl.pg);
  };
  // level 0 decoder
  mask(n->ga, n->d0, V0 * c[0]);
  if ((rc = pnet_wgrad(n, s, 9, n->ga, n->d1, n->s0, L[0].pg))) return rc;
  if ((rc = bwd(9, n->ga, n->gt, n->gs0, L[0].pg))) return rc;
  unup(n->gt, n->gb, c[1], L[1]);
  mask(n->gb, n->d1, L[1].V * c[1]);
  if ((rc = pnet_wgrad(n, s, 8, n->gb, n->d2, n->s1, L[1].pg))) return rc;
  if ((rc = bwd(8, n->gb, n->gt, n->gs1, L[1].pg))) return rc;
  unup(n->gt, n->ga, c[2], L[2]);
  mask(n->ga, n->d2, L[2].V * c[2]);
  if ((rc = pnet_wgrad(n, s, 7, n->ga, n->s3, n->s2, L[2].pg))) return rc;
  if ((rc = bwd(7, n->ga, n->gt, n->gs2, L[2].pg))) return rc;
  unup(n->gt, n->gb, c[3], L[3]);
  mask(n->gb, n->s3, L[3].V * c[3]);
  // bottleneck and encoder
  if ((rc = pnet_wgrad(n, s, 6, n->gb, n->p3, nullptr, L[3].pg))) return rc;
  if ((rc = bwd(6, n->gb, n->gt, nullptr, L[3].pg))) return rc;
  poolb(n->gt, n->gs2, c[2], L[2]);
  mask(n->gs2, n->s2, L[2].V * c[2]);
  if ((rc = pnet_wgrad(n, s, 5, n->gs2, n->a2, nullptr, L[2].pg))) return rc;
  if ((rc = bwd(5, n->gs2, n->ga, nullptr, L[2].pg))) return rc;
  mask(n->ga, n->a2, L[2].V * c[2]);
  if ((rc = pnet_wgrad(n, s, 4, n->ga, n->p2, nullptr, L[2].pg))) return rc;
  if ((rc = bwd(4, n->ga, n->gt, nullptr, L[2].pg))) return rc;
  poolb(n->gt, n->gs1, c[1], L[1]);
  mask(n->gs1, n->s1, L[1].V * c[1]);
  if ((rc = pnet_wgrad(n, s, 3, n->gs1, n->a1, nullptr, L[1].pg))) return rc;
  if ((rc = bwd(3, n->gs1, n->ga, nullptr, L[1].pg))) return rc;
  mask(n->ga, n->a1, L[1].V * c[1]);
  if ((rc = pnet_wgrad(n, s, 2, n->ga, n->p1, nullptr, L[1].pg))) return rc;
  if ((rc = bwd(2, n->ga, n->gt, nullptr, L[1].pg))) return rc;
  poolb(n->gt, n->gs0, c[0], L[0]);
  mask(n->gs0, n->s0, V0 * c[0]);
  if ((rc = pnet_wgrad(n, s, 1, n->gs0, n->a0, nullptr, L[0].pg))) return rc;
  if ((rc = bwd(1, n->gs0, n->ga, nullptr, L[0].pg))) return rc;
  mask(n->ga, n->a0, V0 * c[0]);
  if ((rc = pnet_wgrad(n, s, 0, n->ga, n->x0, nullptr, L[0].pg))) return rc;
  HCHECK(hipGetLastError());
  return GFM_OK;
}

static int pnet_check_grid(const gfm_pnet *n, int batch, int d, int h, int w) {
  if (batch < 1 || batch > n->max_batch || d < 8 || h < 8 || w < 8 || (d | h | w) & 7 ||
      (long long)d * h * w > n->max_cells) {
    set_err("pnet: grid %dx%dx%d x %d outside the network's buffers (multiples of 8, <= %lld cells, batch <= %d)", d, h, w,
            batch, n->max_cells, n->max_batch);
    return GFM_E_ARG;
  }
  return GFM_OK;
}
static double pnet_open(const float *x, long long V) {
  double m = 0.0;
  for (long long v = 0; v < V; v++) m += x[2 * v + 1] > 0.f ? 1.0 : 0.0;
  return m;
}

int gfm_pnet_create(int device, const int ch[4], int max_batch, long long max_cells, uint64_t seed, gfm_pnet **out) {
  if (!ch || !out || max_batch < 1 || max_cells < 512) { set_err("pnet: bad argument"); return GFM_E_ARG; }
  for (int q = 0; q < 4; q++)
    if (ch[q] < 32 || ch[q] > 128 || ch[q] % 32) { set_err("pnet: channel widths must be 32, 64, 96 or 128"); return GFM_E_ARG; }
  if (ch[3] + ch[2] > 128 || ch[2] + ch[1] > 128 || ch[1] + ch[0] > 128) {
    set_err("pnet: a decoder input (up + skip channels) exceeds 128");
    return GFM_E_ARG;
  }
  *out = nullptr;
  gfm_pnet *n = new gfm_pnet();
  n->device = device;
  for (int q = 0; q < 4; q++) n->ch[q] = ch[q];
  n->max_batch = max_batch;
  n->max_cells = max_cells;
  auto fail = [&](int r) { gfm_pnet_destroy(n); return r; };
  {
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&n->stream);
    if (e != hipSuccess) { set_err("pnet: %s", hipGetErrorString(e)); return fail(GFM_E_DEVICE); }
  }
  pnet_layout(n);
  int rc;
  const long long V0 = (long long)max_batch * max_cells, V1 = V0 / 8 + 1, V2 = V0 / 64 + 1, V3 = V0 / 512 + 1;
  const int *c = ch;
#define PA(f, cnt) if ((rc = palloc(n, &n->f, (size_t)(cnt)))) return fail(rc);
  PA(p, n->n_params) PA(gr, n->n_params) PA(m, n->n_params) PA(v, n->n_params) PA(frozen, n->n_params)
  PA(x0, V0 * 32) PA(a0, V0 * c[0]) PA(s0, V0 * c[0]) PA(p1, V1 * c[0]) PA(a1, V1 * c[1]) PA(s1, V1 * c[1])
  PA(p2, V2 * c[1]) PA(a2, V2 * c[2]) PA(s2, V2 * c[2]) PA(p3, V3 * c[2]) PA(s3, V3 * c[3])
  PA(d2, V2 * c[2]) PA(d1, V1 * c[1]) PA(d0, V0 * c[0])
  PA(gs0, V0 * c[0]) PA(gs1, V1 * c[1]) PA(gs2, V2 * c[2]) PA(ga, V0 * 128) PA(gb, V0 * 128) PA(gt, V0 * 128)
  PA(xin, V0 * 2) PA(tin, V0) PA(y, V0) PA(dy, V0)
  PA(fpart, (V0 / 256 + 2) * (c[0] + 1)) PA(lpart, V0 / 256 + 2)
#undef PA
  n->wpart_n = 16LL << 20;
  if ((rc = palloc(n, &n->wpart, (size_t)n->wpart_n))) return fail(rc);
  for (int l = 0; l < 10; l++) {
    PConv &q = n->conv[l];
    const size_t cnt = (size_t)27 * q.co * (q.ca + q.cb);
    if ((rc = palloc(n, &q.wf, cnt))) return fail(rc);
    if ((rc = palloc(n, &q.wt, cnt))) return fail(rc);
  }
  // He-normal weights (fan-in = 27 x real input channels), zero biases;
  // layer 0's padding input channels are zero and frozen
  std::vector<float> hp(n->n_params, 0.f);
  std::vector<uint8_t> hf(n->n_params, 0);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int l = 0; l < 10; l++) {
    PConv &q = n->conv[l];
    const int cin = q.ca + q.cb, creal = l == 0 ? 2 : cin;
    const double sd = sqrt(2.0 / (27.0 * creal));
    for (int t = 0; t < 27; t++)
      for (int o = 0; o < q.co; o++)
        for (int i = 0; i < cin; i++) {
          const size_t idx = q.off_w + ((size_t)t * q.co + o) * cin + i;
          if (i < creal) hp[idx] = (float)(sd * nd(rng));
          else hf[idx] = 1;
        }
  }
  for (int q = 0; q < c[0]; q++) hp[n->off_out + q] = (float)(sqrt(1.0 / c[0]) * nd(rng));
  HCHECK(hipMemcpy(n->p, hp.data(), hp.size() * sizeof(float), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(n->frozen, hf.data(), hf.size(), hipMemcpyHostToDevice));
  if ((rc = pnet_cast(n))) return fail(rc);
  HCHECK(hipStreamSynchronize(n->stream));
  *out = n;
  return GFM_OK;
}

int gfm_pnet_destroy(gfm_pnet *n) {
  if (!n) return GFM_OK;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *q : n->allocs) (void)hipFree(q);
  if (n->set_x) (void)hipFree(n->set_x);
  if (n->set_t) (void)hipFree(n->set_t);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return GFM_OK;
}

int gfm_pnet_param_count(const gfm_pnet *n, size_t *n_total, size_t *n_effective) {
  if (!n) { set_err("pnet: null network"); return GFM_E_ARG; }
  if (n_total) *n_total = n->n_params;
  if (n_effective) *n_effective = n->n_eff;
  return GFM_OK;
}
int gfm_pnet_get_params(gfm_pnet *n, float *out, size_t cnt) {
  if (!n || !out || cnt != n->n_params) { set_err("pnet: get_params needs %zu floats", n ? n->n_params : (size_t)0); return GFM_E_ARG; }
  HCHECK(hipSetDevice(n->device));
  HCHECK(hipMemcpyAsync(out, n->p, cnt * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  HCHECK(hipStreamSynchronize(n->stream));
  return GFM_OK;
}
int gfm_pnet_set_params(gfm_pnet *n, const float *in, size_t cnt) {
  if (!n || !in || cnt != n->n_params) { set_err("pnet: set_params needs %zu floats", n ? n->n_params : (size_t)0); return GFM_E_ARG; }
  HCHECK(hipSetDevice(n->device));
  HCHECK(hipMemcpyAsync(n->p, in, cnt * sizeof(float), hipMemcpyHostToDevice, n->stream));
  int rc = pnet_cast(n);
  if (rc) return rc;
  HCHECK(hipStreamSynchronize(n->stream));
  return GFM_OK;
}
int gfm_pnet_forward(gfm_pnet *n, int batch, int d, int h, int w, const float *x, float *y) {
  if (!n || !x || !y) { set_err("pnet: null argument"); return GFM_E_ARG; }
  int rc = pnet_check_grid(n, batch, d, h, w);
  if (rc) return rc;
  HCHECK(hipSetDevice(n->device));
  const long long V = (long long)batch * d * h * w;
  HCHECK(hipMemcpyAsync(n->xin, x, V * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  if ((rc = pnet_forward_dev(n, n->stream, batch, d, h, w))) return rc;
  HCHECK(hipMemcpyAsync(y, n->y, V * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  HCHECK(hipStreamSynchronize(n->stream));
  return GFM_OK;
}
static int pnet_read_loss(gfm_pnet *n, long long V0, double *loss) {
  const unsigned nb = pblocks(V0);
  std::vector<double> part(nb);
  HCHECK(hipMemcpyAsync(part.data(), n->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, n->stream));
  HCHECK(hipStreamSynchronize(n->stream));
  double s = 0.0;
  for (double q : part) s += q;
  *loss = s;
  return GFM_OK;
}
int gfm_pnet_loss_grad(gfm_pnet *n, int batch, int d, int h, int w, const float *x, const float *t, double *loss,
                       float *grads) {
  if (!n || !x || !t) { set_err("pnet: null argument"); return GFM_E_ARG; }
  int rc = pnet_check_grid(n, batch, d, h, w);
  if (rc) return rc;
  HCHECK(hipSetDevice(n->device));
  const long long V = (long long)batch * d * h * w;
  const double nopen = pnet_open(x, V);
  if (nopen <= 0.0) { set_err("pnet: batch has no open cells"); return GFM_E_ARG; }
  HCHECK(hipMemcpyAsync(n->xin, x, V * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  HCHECK(hipMemcpyAsync(n->tin, t, V * sizeof(float), hipMemcpyHostToDevice, n->stream));
  if ((rc = pnet_loss_grad_dev(n, n->stream, batch, d, h, w, (float)(1.0 / nopen)))) return rc;
  double l = 0.0;
  if ((rc = pnet_read_loss(n, V, &l))) return rc;
  if (loss) *loss = l / nopen;
  if (grads) {
    HCHECK(hipMemcpyAsync(grads, n->gr, n->n_params * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HCHECK(hipStreamSynchronize(n->stream));
  }
  return GFM_OK;
}
static int pnet_adam(gfm_pnet *n, float lr, float b1, float b2, float eps) {
  n->adam_t++;
  const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
  hipLaunchKernelGGL(k_padam, dim3(pblocks(n->n_params)), dim3(256), 0, n->stream, n->p, n->gr, n->m, n->v, n->frozen,
                     (long long)n->n_params, lr, b1, b2, eps, c1, c2);
  return pnet_cast(n);
}
int gfm_pnet_adam_step(gfm_pnet *n, float lr, float beta1, float beta2, float eps) {
  if (!n) { set_err("pnet: null network"); return GFM_E_ARG; }
  HCHECK(hipSetDevice(n->device));
  int rc = pnet_adam(n, lr, beta1, beta2, eps);
  if (rc) return rc;
  HCHECK(hipStreamSynchronize(n->stream));
  return GFM_OK;
}
int gfm_pnet_train(gfm_pnet *n, int nsamples, int d, int h, int w, const float *x, const float *t, int epochs, int batch,
                   float lr, uint64_t seed, double *epoch_loss) {
  if (!n || !x || !t || nsamples < 1 || epochs < 1) { set_err("pnet: bad argument"); return GFM_E_ARG; }
  int rc = pnet_check_grid(n, batch, d, h, w);
  if (rc) return rc;
  HCHECK(hipSetDevice(n->device));
  const long long cells = (long long)d * h * w;
  if (n->set_n < nsamples || n->set_cells != cells) {
    if (n->set_x) HCHECK(hipFree(n->set_x));
    if (n->set_t) HCHECK(hipFree(n->set_t));
    n->set_x = n->set_t = nullptr;
    HCHECK(hipMalloc(&n->set_x, (size_t)nsamples * cells * 2 * sizeof(float)));
    HCHECK(hipMalloc(&n->set_t, (size_t)nsamples * cells * sizeof(float)));
    n->set_n = nsamples;
    n->set_cells = cells;
  }
  HCHECK(hipMemcpyAsync(n->set_x, x, (size_t)nsamples * cells * 2 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  HCHECK(hipMemcpyAsync(n->set_t, t, (size_t)nsamples * cells * sizeof(float), hipMemcpyHostToDevice, n->stream));
  std::vector<double> nopen(nsamples);
  for (int q = 0; q < nsamples; q++) nopen[q] = pnet_open(x + (size_t)q * cells * 2, cells);
  std::vector<int> order(nsamples);
  for (int q = 0; q < nsamples; q++) order[q] = q;
  std::mt19937_64 rng(seed);
  for (int e = 0; e < epochs; e++) {
    std::shuffle(order.begin(), order.end(), rng);
    double lsum = 0.0;
    int nbatch = 0;
    for (int b0 = 0; b0 + batch <= nsamples; b0 += batch) {
      double no = 0.0;
      for (int q = 0; q < batch; q++) {
        const int smp = order[b0 + q];
        no += nopen[smp];
        HCHECK(hipMemcpyAsync(n->xin + (size_t)q * cells * 2, n->set_x + (size_t)smp * cells * 2, cells * 2 * sizeof(float),
                              hipMemcpyDeviceToDevice, n->stream));
        HCHECK(hipMemcpyAsync(n->tin + (size_t)q * cells, n->set_t + (size_t)smp * cells, cells * sizeof(float),
                              hipMemcpyDeviceToDevice, n->stream));
      }
      if (no <= 0.0) continue;
      if ((rc = pnet_loss_grad_dev(n, n->stream, batch, d, h, w, (float)(1.0 / no)))) return rc;
      if ((rc = pnet_adam(n, lr, 0.9f, 0.999f, 1e-8f))) return rc;
      double l = 0.0;
      if ((rc = pnet_read_loss(n, (long long)batch * cells, &l))) return rc;
      lsum += l / no;
      nbatch++;
    }
    if (epoch_loss) epoch_loss[e] = nbatch ? lsum / nbatch : 0.0;
  }
  HCHECK(hipStreamSynchronize(n->stream));
  return GFM_OK;
}
