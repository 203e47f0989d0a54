// gfm_hip.hip -- host side of the MI355X SIMPLER solver and its C-ABI
// (include/gfm_hip.h).  One translation unit: the kernels are included so
// that the __constant__ case block is shared without relocatable device
// code.
#include "gfm_kernels.hip"
#include <hip/hip_ext.h>
#include "../../include/gfm_hip.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <map>
#include <mutex>
#include <string>
#include <vector>

static thread_local std::string g_err;
static void set_err(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}
#define HCHECK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return GFM_E_DEVICE;                                                         \
    }                                                                              \
  } while (0)

#include "gfm_pnet.hip"

// residual slots in the device scalar block
enum { RS_P_PRE = 0, RS_P = 1, RS_U_PRE = 2, RS_U = 5, RS_PC = 8, RS_PC0 = 9, RS_GF_PRE = 10, RS_GF = 20, RS_IRC = 30,
       RS_PCG = 60 };

struct ProfRec {
  std::string name;
  double ms = 0.0, bytes = 0.0;
  long launches = 0;
};
struct Prof {
  bool on = false;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<int, double>> pending;  // record index, bytes
  int used = 0;
};

struct MsHost;
struct RadHost;
struct Shape {  // launch shape of a tunable box kernel: thread mapping (Box::map) and workgroup size
  int map, bs;
};
struct gfm_ctx {
  int device = 0;
  // launch shapes of the tunable box kernels chosen by gfm_hip_autotune
  // (kernel name -> shape), the candidate under trial (-1: none) and each
  // kernel's largest workgroup
  std::map<std::string, Shape> tuned;
  std::map<std::string, int> tune_maxbs, tune_map0;
  int tune_cand = -1;
  std::vector<int16_t> ibh;  // cell codes, reference layout (lists built after init)
  std::vector<double> hx, hdx, hr, hdr, hz, hdz;  // the 1-D grid (1-based index i at [i-1])
  struct WallHost *wall = nullptr;  // wall exchange, at gfm_hip_wall_setup
  MsHost *ms = nullptr;       // minor-species state, at the first gfm_hip_ms_* call
  RadHost *rad = nullptr;     // radiation fields, at the first gfm_hip_rad_* call
  Prof prof;
  gfm_ctrl ctrl;
  DevCase dc;
  std::vector<void *> allocs;
  double *xt2 = nullptr, *xt3 = nullptr, *xt4 = nullptr, *xt5 = nullptr, *part = nullptr, *stage = nullptr;
  double *xg = nullptr;  // G of the line kernel between the two passes of a colour visit
  double *bc_part = nullptr, *bc_sums = nullptr, *bc_d = nullptr;  // x-plane block correction of the p' box:
                                                                   // chunk sums, plane sums, plane values
  double *flow_part = nullptr, *flow_x = nullptr, *flow_y = nullptr, *flow_z = nullptr;  // FLX1D partials and sums
  int flows_formed = 0;
  float *xtf = nullptr;  // FP32 working copy of the unknown of ADLBL3 (mixed precision)
  int16_t *stage16 = nullptr;
  // lattice positions of the GFLOW wall treatment: open with a wall
  // neighbour, and wall (Fortran linear indices)
  long long *wall_open = nullptr, *wall_pos = nullptr;
  int n_wall_open = 0, n_wall_pos = 0;
  // cells of EXTR: inlets and exits, walls (Fortran linear indices)
  long long *extr_io = nullptr, *extr_wall = nullptr;
  int n_extr_io = 0, n_extr_wall = 0;
  long long nlat = 0, ncls = 0;
  hipStream_t stream = nullptr;
  double sgm = 1.0;
  // one SIMPLER iteration captured as a graph, replayed every iteration
  bool use_graph = true;
  bool line_pcr = false;  // solve every line by PCR across 32 lanes (GFM_LINE_SOLVE=pcr)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_key = -1;
  // p' solved to a tolerance: the iteration runs as graph A (up to the
  // initial guess), the relaxation loop, graph B (the rest)
  hipGraph_t graph_b = nullptr;
  hipGraphExec_t graph_exec_b = nullptr;
  gfm_pnet *pnet = nullptr;     // learned initial guess (not owned)
  double *pin = nullptr;        // pinned host word for residual checks
  float *dump_x = nullptr, *dump_p = nullptr;  // training pair of the last p' solve
  int pair_ready = 0;
};
static gfm_ctx *g_bound = nullptr;

static int prof_rec(gfm_ctx *c, const char *name) {
  for (size_t q = 0; q < c->prof.recs.size(); q++)
    if (c->prof.recs[q].name == name) return (int)q;
  ProfRec r;
  r.name = name;
  c->prof.recs.push_back(r);
  return (int)c->prof.recs.size() - 1;
}
static void prof_flush(gfm_ctx *c) {
  Prof &p = c->prof;
  if (p.used == 0) return;
  (void)hipStreamSynchronize(c->stream);
  for (int u = 0; u < p.used; u++) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, p.ev[2 * u], p.ev[2 * u + 1]);
    ProfRec &r = p.recs[p.pending[u].first];
    r.ms += ms;
    r.bytes += p.pending[u].second;
    r.launches += 1;
  }
  p.used = 0;
  p.pending.clear();
}
static void prof_begin(gfm_ctx *c) {
  Prof &p = c->prof;
  if ((size_t)(2 * p.used + 2) > p.ev.size()) prof_flush(c);
  if (p.ev.empty()) {
    p.ev.resize(8192);
    for (auto &e : p.ev) (void)hipEventCreate(&e);
  }
  (void)hipEventRecord(p.ev[2 * p.used], c->stream);
}
static void prof_end(gfm_ctx *c, const char *name, double bytes) {
  Prof &p = c->prof;
  (void)hipEventRecord(p.ev[2 * p.used + 1], c->stream);
  p.pending.push_back({prof_rec(c, name), bytes});
  p.used++;
}
// The event pair of the next record, for a launch that records the
// kernel's own start and end (hipExtLaunchKernelGGL): the span excludes
// the launch latency that stream-recorded events around a launch include.
static void prof_ext_events(gfm_ctx *c, hipEvent_t *e0, hipEvent_t *e1) {
  Prof &p = c->prof;
  if ((size_t)(2 * p.used + 2) > p.ev.size()) prof_flush(c);
  if (p.ev.empty()) {
    p.ev.resize(8192);
    for (auto &e : p.ev) (void)hipEventCreate(&e);
  }
  *e0 = p.ev[2 * p.used];
  *e1 = p.ev[2 * p.used + 1];
}
static void prof_ext_done(gfm_ctx *c, const char *name, double bytes) {
  Prof &p = c->prof;
  p.pending.push_back({prof_rec(c, name), bytes});
  p.used++;
}

static int bind(gfm_ctx *c) {
  if (g_bound == c) return GFM_OK;
  HCHECK(hipSetDevice(c->device));
  HCHECK(dev_bind_case(c->dc));
  g_bound = c;
  return GFM_OK;
}

template <typename T>
static int dalloc(gfm_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  HCHECK(hipMalloc(&q, n * sizeof(T)));
  HCHECK(hipMemset(q, 0, n * sizeof(T)));
  c->allocs.push_back(q);
  *p = (T *)q;
  return GFM_OK;
}

// box kernels (box_pt): one x row per row of workgroups
static dim3 box_grid(const Box &b) { return dim3((b.ni + 63) / 64, b.nj, b.nk); }
static long long box_blocks(const Box &b) { return (long long)((b.ni + 63) / 64) * b.nj * b.nk; }
// flattened box kernels (box_ptf): the (i, j) rows of each k-plane along x
static dim3 box_grid_f(const Box &b) { return dim3((unsigned)(((long long)b.ni * b.nj + 63) / 64), b.nk); }
static long long box_blocks_f(const Box &b) { return (((long long)b.ni * b.nj + 63) / 64) * b.nk; }
#define LAUNCH(kern, b, ...)                                                              \
  do {                                                                                    \
    if ((b).ni > 0 && (b).nj > 0 && (b).nk > 0) {                                         \
      if (c->prof.on) prof_begin(c);                                                      \
      hipLaunchKernelGGL(kern, box_grid(b), dim3(64), 0, c->stream, b, ##__VA_ARGS__);   \
      if (c->prof.on) prof_end(c, #kern, 0.0);                                            \
    }                                                                                     \
  } while (0)
// Box kernels whose launch shape (thread mapping and workgroup size) an
// autotuner may choose (gfm_hip_autotune): MAP0 is the default mapping
// (0: box_pt rows, 1: box_ptf flattened) at 64 threads.  Every shape hands
// each position to one thread, so results do not depend on the shape.
// A kernel that reads its position with box_pt itself is launched with
// map0 = ROWS_ONLY: the tuner then chooses only its workgroup size.
static const Shape TUNE_CANDS[6] = {{0, 64}, {0, 128}, {0, 256}, {1, 64}, {1, 128}, {1, 256}};
static const int ROWS_ONLY = 2;
static dim3 box_grid_bs(const Box &b, int bs) {
  return b.map ? dim3((unsigned)(((long long)b.ni * b.nj + bs - 1) / bs), b.nk) : dim3((b.ni + bs - 1) / bs, b.nj, b.nk);
}
template <typename K>
static Shape tune_shape(gfm_ctx *c, const char *name, K kern, int map0) {
  const Shape def = {map0 == 1 ? 1 : 0, 64};
  if (c->tune_cand >= 0) {
    c->tune_map0[name] = map0;
    auto it = c->tune_maxbs.find(name);
    int mx = 64;
    if (it == c->tune_maxbs.end()) {
      hipFuncAttributes a;
      if (hipFuncGetAttributes(&a, reinterpret_cast<const void *>(kern)) == hipSuccess) mx = a.maxThreadsPerBlock;
      c->tune_maxbs[name] = mx;
    } else {
      mx = it->second;
    }
    const Shape s = TUNE_CANDS[c->tune_cand];
    return (s.bs <= mx && !(map0 == ROWS_ONLY && s.map)) ? s : def;
  }
  auto it = c->tuned.find(name);
  return it == c->tuned.end() ? def : it->second;
}
#define LAUNCHT(kern, map0, b, ...)                                                           \
  do {                                                                                       \
    if ((b).ni > 0 && (b).nj > 0 && (b).nk > 0) {                                            \
      Box b_ = (b);                                                                          \
      const Shape sh_ = tune_shape(c, #kern, kern, map0);                                    \
      b_.map = sh_.map;                                                                      \
      if (c->prof.on) prof_begin(c);                                                         \
      hipLaunchKernelGGL(kern, box_grid_bs(b_, sh_.bs), dim3(sh_.bs), 0, c->stream, b_, ##__VA_ARGS__); \
      if (c->prof.on) prof_end(c, #kern, 0.0);                                               \
    }                                                                                        \
  } while (0)
// the same for the flattened box kernels (box_ptf)
#define LAUNCHF(kern, b, ...)                                                              \
  do {                                                                                    \
    if ((b).ni > 0 && (b).nj > 0 && (b).nk > 0) {                                         \
      if (c->prof.on) prof_begin(c);                                                      \
      hipLaunchKernelGGL(kern, box_grid_f(b), dim3(64), 0, c->stream, b, ##__VA_ARGS__);   \
      if (c->prof.on) prof_end(c, #kern, 0.0);                                            \
    }                                                                                     \
  } while (0)

static void fill_devcase_scalars(gfm_ctx *c, const gfm_ctrl *k) {
  DevCase &d = c->dc;
  d.nr0 = k->nr0; d.ncells = k->ncells;
  d.i_ref = k->i_ref; d.j_ref = k->j_ref; d.k_ref = k->k_ref;
  int NP = k->np, LP = k->lp;
  int njy1v[4] = {0, 4, 5, 4}, njy2v[4] = {0, NP - 2, NP - 3, NP - 2};
  int lkz1v[4] = {0, 4, 4, 5}, lkz2v[4] = {0, LP - 2, LP - 2, LP - 3};
  memcpy(d.njy1v, njy1v, sizeof njy1v); memcpy(d.njy2v, njy2v, sizeof njy2v);
  memcpy(d.lkz1v, lkz1v, sizeof lkz1v); memcpy(d.lkz2v, lkz2v, sizeof lkz2v);
  d.t0 = k->t0; d.trdg = k->trdg; d.h_0 = k->h_0; d.r0 = k->r0; d.area0 = k->area0; d.vol0 = k->vol0;
  d.reyg = k->reyg; d.reyg2 = k->reyg2; d.eun = k->eun; d.epsr = k->epsr; d.taur = k->taur; d.ug0 = k->ug0;
  d.dnst0 = k->dnst0; d.qr0 = k->qr0; d.q0 = k->q0; d.gmfr = k->gmfr; d.tk0 = k->tk0; d.epsi0 = k->epsi0;
  for (int q = 0; q < 3; q++) d.gra[q] = k->gra[q];
  d.p_ref = k->p_ref;
  d.t_mx = k->t_mx; d.t_mn = k->t_mn; d.u_mx = k->u_mx; d.u_mn = k->u_mn; d.p_mx = k->p_mx; d.p_mn = k->p_mn;
  d.f_mx = k->f_mx; d.f_mn = k->f_mn;
  for (int q = 0; q < 21; q++) { d.rf[q] = k->rf[q]; d.rfl[q] = k->rfl[q]; }
  d.rf_hsource = k->rf_hsource;
  d.t_wall = k->t_wall;
  d.kappa_gray = k->kappa_gray;
  d.sigma_sb = k->sigma_sb;
  d.h_g = k->h_g;
  d.i1st_rad_done = k->i1st_rad_done;
  d.isothermal_walls = k->isothermal_walls;
  d.gray_radiation = k->gray_radiation;
  for (int q = 0; q < 2; q++) d.bv_f[q] = k->bv_f[q];
  for (int q = 0; q < 6; q++) d.ba_f[q] = k->ba_f[q];
  d.wt0 = 1.0 / (0.5 / WF + 0.0 / WCO2 + 0.0 / WH2O + 0.49 / WO2 + 0.01 / WN2);
  d.stoich = 2.0 * WO2 / WF;
}

// Positions of the GFLOW wall treatment (C/Gflow.f90:16-130), whose loops
// run over i, j, k = 2..MP, 2..NP, 2..LP: the open positions (IBCELL <= 0)
// with a wall (IBCELL == 1) among their six lattice neighbours, where
// theta and the wall functions act, and the wall positions, where GAMA is
// zeroed.  A neighbour beyond MP, NP or LP counts as open, as the device
// holds 0 there.
static void wall_lists(const gfm_ctrl *k, const int16_t *ibf, std::vector<long long> &wo, std::vector<long long> &wp) {
  const int MPh = k->mp, NPh = k->np, LPh = k->lp;
  auto lin = [&](int i, int j, int kk) -> long long { return (((long long)kk - 1) * NPh + (j - 1)) * MPh + (i - 1); };
  auto ib = [&](int i, int j, int kk) -> int {
    if (i > MPh || j > NPh || kk > LPh) return 0;
    return ibf[lin(i, j, kk)];
  };
  wo.clear();
  wp.clear();
  for (int kk = 2; kk <= LPh; kk++)
    for (int j = 2; j <= NPh; j++)
      for (int i = 2; i <= MPh; i++) {
        const int b = ib(i, j, kk);
        if (b == 1) { wp.push_back(lin(i, j, kk)); continue; }
        if (b >= 1) continue;
        if (ib(i - 1, j, kk) == 1 || ib(i + 1, j, kk) == 1 || ib(i, j - 1, kk) == 1 || ib(i, j + 1, kk) == 1 ||
            ib(i, j, kk - 1) == 1 || ib(i, j, kk + 1) == 1)
          wo.push_back(lin(i, j, kk));
      }
}

// Cells of EXTR(1) (C/Extr.f90:31-250) in the cell box [2, MP]x[2, NP]x
// [2, LP] (step 2): inlets and exits (IBCELL mod 10 = 2 or 3) and walls
// (IBCELL mod 10 = 1).
static void extr_lists(const gfm_ctrl *k, const int16_t *ibf, std::vector<long long> &io, std::vector<long long> &wl) {
  const int MPh = k->mp, NPh = k->np, LPh = k->lp;
  io.clear();
  wl.clear();
  for (int kk = 2; kk <= LPh; kk += 2)
    for (int j = 2; j <= NPh; j += 2)
      for (int i = 2; i <= MPh; i += 2) {
        const long long f = (((long long)kk - 1) * NPh + (j - 1)) * MPh + (i - 1);
        const int b = ibf[f] % 10;
        if (b == 2 || b == 3) io.push_back(f);
        else if (b == 1) wl.push_back(f);
      }
}

// sticky INTP weights: replay the reference's INTP loop nest on the host
// (C/Intp.f90:38-70 with INTP0's in-place F1 update, :149-157) and record
// for every position whether F1 is the loop's INTP1 weight (0), forced to
// one (1) or forced to zero (2).
static void intp_codes(const gfm_ctrl *k, const int16_t *ibf, std::vector<int8_t> &code, const DevGeom &gg) {
  const int MPh = k->mp, NPh = k->np, LPh = k->lp;
  auto ib = [&](int i, int j, int kk) -> int { return ibf[(((size_t)kk - 1) * NPh + (j - 1)) * MPh + (i - 1)]; };
  auto li = [&](int i, int j, int kk) -> long long { return lat_index(gg, i, j, kk); };
  code.assign((size_t)gg.S * 8, 0);
  auto visit = [&](int &st, int i, int j, int kk, int i1, int j1, int k1, int i2, int j2, int k2) {
    if (ib(i, j, kk) == 3) return;
    if (ib(i, j, kk) == 2) {
      if (ib(i1, j1, k1) == 2 && ib(i2, j2, k2) == 0) st = 1;
      else if (ib(i1, j1, k1) == 0 && ib(i2, j2, k2) == 2) st = 2;
    }
    code[li(i, j, kk)] = (int8_t)st;
  };
  for (int i = 2; i <= MPh; i += 2) {
    for (int kk = 3; kk <= LPh - 1; kk += 2) {
      int st = 0;
      for (int j = 2; j <= NPh; j += 2) {
        if (ib(i, j, kk) == 1) continue;
        visit(st, i, j, kk, i, j, kk + 1, i, j, kk - 1);
      }
    }
    for (int j = 3; j <= NPh - 1; j += 2) {
      int st = 0;
      for (int kk = 2; kk <= LPh; kk++) {
        if (ib(i, j, kk) == 1) continue;
        visit(st, i, j, kk, i, j + 1, kk, i, j - 1, kk);
      }
    }
  }
  for (int i = 3; i <= MPh - 1; i += 2) {
    int st = 0;
    for (int kk = 2; kk <= LPh; kk++)
      for (int j = 2; j <= NPh; j++) {
        if (ib(i, j, kk) == 1) continue;
        visit(st, i, j, kk, i + 1, j, kk, i - 1, j, kk);
      }
  }
}

// ---------------------------------------------------------------------
// host <-> device state
// ---------------------------------------------------------------------
static int scatter(gfm_ctx *c, const double *h, double *d) {
  HCHECK(hipMemcpyAsync(c->stage, h, c->nlat * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((c->nlat + 255) / 256)), dim3(256), 0, c->stream, c->stage, d, c->nlat);
  return GFM_OK;
}
// Coefficient array of a staged system: FP64 copy, and its FP32 copy in a
// mixed-precision context (the line kernels and residuals read that one).
static int scatter_coef(gfm_ctx *c, const double *h, double *d, float *d32) {
  int rc = scatter(c, h, d);
  if (rc || !c->dc.mixed) return rc;
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((c->ncls + 255) / 256)), dim3(256), 0, c->stream, d, d32, c->ncls);
  return GFM_OK;
}
static int gather(gfm_ctx *c, const double *d, double *h) {
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((c->nlat + 255) / 256)), dim3(256), 0, c->stream, d, c->stage, c->nlat);
  HCHECK(hipMemcpyAsync(h, c->stage, c->nlat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return GFM_OK;
}

#define FIELD_LIST(X_) X_(p) X_(t) X_(dnst) X_(theta) X_(tmu) X_(gcp) X_(glam) X_(gama) X_(gte) X_(crate) X_(hrate)

static int upload_state(gfm_ctx *c, const gfm_ctrl *k, const gfm_host_arrays *h) {
  int rc = bind(c);
  if (rc) return rc;
  DevCase &d = c->dc;
#define UP(f) if ((rc = scatter(c, h->f, d.f))) return rc;
  FIELD_LIST(UP)
#undef UP
  if ((rc = scatter(c, h->bs_hprev, d.bsh))) return rc;
  for (int q = 0; q < 3; q++)
    if ((rc = scatter(c, h->ug + (size_t)q * c->nlat, d.ug[q]))) return rc;
  for (int q = 0; q < 9; q++)
    if ((rc = scatter(c, h->gf + (size_t)q * c->nlat, d.gf[q]))) return rc;
  double sc[SC_COUNT] = {0};
  sc[SC_FACTOR] = k->factor;
  sc[SC_FADJ] = k->fadj;
  sc[SC_GFIN0] = k->gfin0;
  HCHECK(hipMemcpyAsync(d.scal, sc, sizeof sc, hipMemcpyHostToDevice, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

static void pull_scalars(gfm_ctx *c, gfm_ctrl *k, const double *sc) {
  k->res_mass = sc[SC_RES_MASS];
  k->res_mass_max = sc[SC_RES_MAX];
  k->imx = (int)sc[SC_IMX]; k->jmx = (int)sc[SC_JMX]; k->kmx = (int)sc[SC_KMX];
  k->gfex = sc[SC_GFEX]; k->fuel_out = sc[SC_FUEL_OUT]; k->tmean_ex = sc[SC_TMEAN_EX];
  k->gfin0 = sc[SC_GFIN0]; k->gfin = sc[SC_GFIN]; k->factor = sc[SC_FACTOR]; k->fadj = sc[SC_FADJ];
  k->q_exhaust = sc[SC_QEXH]; k->q_react = sc[SC_QREACT]; k->q_wall = sc[SC_QWALL];
  k->q_wall_bottom = sc[SC_QWALL_LO]; k->q_wall_top = sc[SC_QWALL_HI];
  if (sc[SC_TAVE_DEN] > 0.0) {
    k->tave = sc[SC_TAVE_NUM] * k->t0 / sc[SC_TAVE_DEN];
    k->tave_vol = sc[SC_TVOL_NUM] * k->t0 / sc[SC_TVOL_DEN];
    k->tmax_c = sc[SC_TMAX] * k->t0;
  }
  k->t_av_exit = sc[SC_TEXIT_AREA] > 0.0 ? sc[SC_TEXIT_SUM] * k->t0 / sc[SC_TEXIT_AREA] : 0.0;
  k->t_x_exit = sc[SC_TEXIT_MAX] * k->t0;
  const double *r = sc + SC_RESID0;
  if (k->want_pre_residuals) {
    k->p_resid_pre = r[RS_P_PRE];
    for (int q = 0; q < 3; q++) k->u_resid_pre[q] = r[RS_U_PRE + q];
    for (int q = 0; q < 7; q++) k->gf_resid_pre[q] = r[RS_GF_PRE + q];
    k->h_resid_pre = r[RS_GF_PRE + 2];
  }
  k->p_resid = r[RS_P];
  for (int q = 0; q < 3; q++) k->u_resid[q] = r[RS_U + q];
  k->pcorr_resid = r[RS_PC];
  k->pcorr_alpha = sc[SC_PG_ALPHA];
  k->pcorr_gain = sc[SC_PG_COS2];
  k->pcorr_fallback = (int)sc[SC_PG_FALLBACK];
  for (int q = 0; q < 7; q++) k->gf_resid[q] = r[RS_GF + q];
  k->h_resid = r[RS_GF + 2];
  if (k->want_irc)
    for (int q = 0; q < 12; q++) { k->resg[q] = r[RS_IRC + 2 * q]; k->resg[20 + q] = r[RS_IRC + 2 * q + 1]; }
}

// ---------------------------------------------------------------------
// ADLBL3 (C/Adlbl3.f90:10-97) and calc_residual
// ---------------------------------------------------------------------
// Line-kernel instance: tile width (one 128-byte line of each row: 16 FP64
// or 32 FP32 planes; 2 for fewer planes than a tile) and the solve: lines of
// up to 160 rows by 16 lanes with M contiguous rows each (M in 2, 4, 6, 8,
// 10), longer lines by PCR with NU rows per lane across 32 lanes (lines of
// up to 32*NU rows).
static int lines_lw(int npl, size_t word) {
  const int full = word == 4 ? 32 : 16;
  return npl >= full ? full : 2;
}
static int lines_nu(int nmax) { return nmax <= 96 ? (nmax <= 64 ? 2 : 3) : nmax <= 128 ? 4 : nmax <= 160 ? 5 : nmax <= 192 ? 6 : 8; }
static int lines_m(int nmax, bool pcr) {
  if (pcr || nmax > 160) return 0;
  return nmax <= 32 ? 2 : nmax <= 64 ? 4 : nmax <= 96 ? 6 : nmax <= 128 ? 8 : 10;
}
static size_t lines_shm(int lw, int nmax, size_t word) { return 3 * (size_t)lw * (nmax | 1) * word; }
#define LINES_INSTANCES_LW(X_, LW_)                                                                       \
  X_(LW_, 1, 2) X_(LW_, 1, 4) X_(LW_, 1, 6) X_(LW_, 1, 8) X_(LW_, 1, 10) X_(LW_, 2, 0) X_(LW_, 3, 0) X_(LW_, 4, 0) \
  X_(LW_, 5, 0) X_(LW_, 6, 0) X_(LW_, 8, 0)
#define LINES_INSTANCES(X_) LINES_INSTANCES_LW(X_, 32) LINES_INSTANCES_LW(X_, 16) LINES_INSTANCES_LW(X_, 2)
template <typename T>
static void launch_lines(gfm_ctx *c, int lw, int nu, int m, unsigned nblk, size_t shm, hipEvent_t e0, hipEvent_t e1,
                         const LineIO<T> &io) {
#define LINES_CASE(LW_, NU_, M_)                                                                                   \
  if (lw == LW_ && (M_ ? m == M_ : m == 0 && nu == NU_)) {                                                         \
    hipExtLaunchKernelGGL((k_lines<LW_, T, NU_, M_>), dim3(nblk), dim3(LINES_NT), shm, c->stream, e0, e1, 0, io); \
    return;                                                                                                       \
  }
  LINES_INSTANCES(LINES_CASE)
#undef LINES_CASE
}
template <typename T>
static int set_lines_attr(size_t shm) {
#define LINES_ATTR(LW_, NU_, M_)                                                                            \
  HCHECK(hipFuncSetAttribute((const void *)k_lines<LW_, T, NU_, M_>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                             (int)shm));
  LINES_INSTANCES(LINES_ATTR)
#undef LINES_ATTR
  return GFM_OK;
}

// One line-kernel pass: y-line and z-line solves on the planes
// i = pi0 + pdi*p (p < npl) of the box, from the current values io
// describes, into io.xy and io.xz.
template <typename T>
static void line_pass(gfm_ctx *c, const LineIO<T> &io) {
  const int nyl = (io.ke - io.kb) / 2 + 1, nzl = (io.je - io.jb) / 2 + 1;  // y-lines per plane, z-lines per plane
  const int ny = nzl, nz = nyl;                                            // rows of a y-line, of a z-line
  const double cells = (double)io.npl * nyl * nzl;
  // algorithmic bytes per row and direction: 8 coefficients, the current
  // values of the x neighbours (one word from xt, or the two words of a
  // pair whose average the kernel forms) and the written solution, plus one
  // word per cell where G is stored; where G is read, 5 coefficients, G and
  // the written solution
  const double line_bytes = io.gr ? cells * 2.0 * 7.0 * sizeof(T)
                                  : cells * (2.0 * (8.0 + (io.oth2 ? 2.0 : 1.0) + 1.0) + (io.gw ? 1.0 : 0.0)) * sizeof(T);
  const int lw = lines_lw(io.npl, sizeof(T));
  const int ntile = (io.npl + lw - 1) / lw;
  const int nmax = ny > nz ? ny : nz;
  const size_t shm = lines_shm(lw, nmax, sizeof(T));
  const unsigned nblk = (unsigned)(8 * ntile * ((nz + 7) / 8 + (ny + 7) / 8));  // XCD-balanced order (k_lines)
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->prof.on) prof_ext_events(c, &e0, &e1);
  launch_lines<T>(c, lw, lines_nu(nmax), lines_m(nmax, c->line_pcr), nblk, shm, e0, e1, io);
  if (c->prof.on) prof_ext_done(c, "k_lines", line_bytes);
}

// XT = (X2 + X3)/2 on the planes i = pi0 + pdi*p (p < npl) of the box,
// C/Adlbl3.f90:134-138.
template <typename T>
static void line_avg(gfm_ctx *c, T *xt, const T *x2, const T *x3, int pi0, int pdi, int npl, int jb, int je, int kb, int ke) {
  const int nyl = (ke - kb) / 2 + 1, nzl = (je - jb) / 2 + 1;
  const double cells = (double)npl * nyl * nzl;
  // adjacent same-colour planes form aligned pairs when the first plane's
  // x position (lattice_index formula) is even
  const int c0 = pi0 >> 1;
  const bool vec = pdi == 4 && npl > 1 && ((((c0 & 1) * c->dc.g.HX + (c0 >> 1) + 31) & 1) == 0);
  if (c->prof.on) prof_begin(c);
  if (vec) {
    const long long tot2 = (long long)((npl + 1) / 2) * nzl * nyl;
    hipLaunchKernelGGL(k_line_avg2<T>, dim3((unsigned)((tot2 + 255) / 256)), dim3(256), 0, c->stream, xt, x2, x3, pi0, pdi, npl,
                       jb, nzl, kb, nyl);
  } else {
    const long long tot = (long long)npl * nzl * nyl;
    hipLaunchKernelGGL(k_line_avg<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, xt, x2, x3, pi0, pdi, npl,
                       jb, nzl, kb, nyl);
  }
  if (c->prof.on) prof_end(c, "k_line_avg", cells * 3.0 * sizeof(T));
}

// nim sweeps of ADLBL3 on xt over the box.  Reference order: plane by
// plane, forward then backward, each plane the two passes of ADLZ, a y/z
// line solve from the same XT followed by the average.  Zebra order: the
// same on all planes of one colour at once, colours 0, 1, 0, 1 per sweep;
// the averages are not stored between passes: each pass reads the current
// values of a colour's planes as the average of the pair its last pass
// wrote (x2/x3 or x4/x5), and xt receives them once, after the last pass.
template <typename T>
static void sweeps(gfm_ctx *c, T *xt, T *x2, T *x3, T *x4, T *x5, T *xg, int nim, int ib, int ie, int jb, int je, int kb,
                   int ke) {
  const int npl = (ie - ib) / 2 + 1;
  if (c->ctrl.sweep_order == GFM_ORDER_REFERENCE) {
    for (int s = 0; s < nim; s++)
      for (int dirn = 0; dirn < 2; dirn++)
        for (int q = 0; q < npl; q++) {
          const int i = dirn == 0 ? ib + 2 * q : ie - 2 * q;
          for (int l = 0; l < 2; l++) {
            const LineIO<T> io{xt, nullptr, nullptr, nullptr, nullptr, x2, x3, nullptr, nullptr, i, 0, 1, ib, ie, jb, je, kb, ke};
            line_pass<T>(c, io);
            line_avg<T>(c, xt, x2, x3, i, 0, 1, jb, je, kb, ke);
          }
        }
    return;
  }
  T *pr[2][2] = {{x2, x3}, {x4, x5}};
  int cur[2] = {-1, -1};  // pair holding each colour's current values (-1: xt)
  {
    const int nj = (je - jb) / 2 + 1, nk = (ke - kb) / 2 + 1;
    const long long tot = 2LL * nj * nk;
    hipLaunchKernelGGL(k_pair_xbound<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, xt, x2, x3, x4, x5, ib,
                       ie, jb, nj, kb, nk);
  }
  const int ncol[2] = {(npl + 1) / 2, npl / 2};
  for (int s = 0; s < nim; s++)
    for (int half = 0; half < 4; half++) {
      const int col = half & 1;
      if (ncol[col] == 0) continue;
      for (int l = 0; l < 2; l++) {
        const int own = cur[col], oth = cur[1 - col];
        const int tgt = own == 0 ? 1 : 0;  // never the pair this pass reads its own planes from
        const LineIO<T> io{xt, own >= 0 ? pr[own][0] : nullptr, own >= 0 ? pr[own][1] : nullptr,
                           oth >= 0 ? pr[oth][0] : nullptr, oth >= 0 ? pr[oth][1] : nullptr, pr[tgt][0], pr[tgt][1],
                           l == 0 ? xg : nullptr, l == 1 ? xg : nullptr,
                           ib + 2 * col, 4, ncol[col], ib, ie, jb, je, kb, ke};
        line_pass<T>(c, io);
        cur[col] = tgt;
      }
    }
  for (int col = 0; col < 2; col++)
    if (cur[col] >= 0) line_avg<T>(c, xt, pr[cur[col]][0], pr[cur[col]][1], ib + 2 * col, 4, ncol[col], jb, je, kb, ke);
}
static void residual(gfm_ctx *c, const double *xt, int ib, int ie, int jb, int je, int kb, int ke, int slot, int irc) {
  Box b = make_box(ib, ie, jb, je, kb, ke, 2);
  LAUNCH(k_residual, b, xt, c->part, irc);
  hipLaunchKernelGGL(k_residual_finalize, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)box_blocks(b), SC_RESID0 + slot, irc);
}
// nim sweeps of ADLBL3 on xt over the box (FP32 working copy in mixed precision)
static void adlbl3_sweeps(gfm_ctx *c, double *xt, int nim, int ib, int ie, int jb, int je, int kb, int ke) {
  if (!c->dc.mixed) {
    sweeps<double>(c, xt, c->xt2, c->xt3, c->xt4, c->xt5, c->xg, nim, ib, ie, jb, je, kb, ke);
  } else {
    // FP32 sweeps on a working copy of the box and the neighbours it reads
    // (one lattice step of the class outside the box); rows the sweeps pin
    // keep their FP64 values.
    Box ext = make_box(ib - 2, ie + 2, jb - 2, je + 2, kb - 2, ke + 2, 2);
    Box box = make_box(ib, ie, jb, je, kb, ke, 2);
    LAUNCHT(k_box_narrow, 1, ext, xt, c->xtf);
    sweeps<float>(c, c->xtf, reinterpret_cast<float *>(c->xt2), reinterpret_cast<float *>(c->xt3),
                  reinterpret_cast<float *>(c->xt4), reinterpret_cast<float *>(c->xt5), reinterpret_cast<float *>(c->xg), nim,
                  ib, ie, jb, je, kb, ke);
    LAUNCHT(k_box_widen, 1, box, c->xtf, xt);
  }
}
static void adlbl3(gfm_ctx *c, double *xt, int nelg, int ib, int ie, int jb, int je, int kb, int ke) {
  adlbl3_sweeps(c, xt, c->ctrl.ntimesg[nelg - 1], ib, ie, jb, je, kb, ke);
  if (c->ctrl.want_irc) residual(c, xt, ib, ie, jb, je, kb, ke, RS_IRC + 2 * (nelg - 1), 1);
}

// Initial guess of a p' solve relaxed to a tolerance: calc_residual of
// FZ = 0, then (learned guess) the network's prediction from BS/AP,
// scaled by the alpha minimising ||BS - alpha*A g||, kept only if its
// calc_residual is below that of zero.  All on the device (capturable).
static int pcorr_guess(gfm_ctx *c, const Box &cen) {
  DevCase &d = c->dc;
  const gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_PC0, 0);
  if (k.pcorr_guess != GFM_PGUESS_LEARNED || !c->pnet) return GFM_OK;
  gfm_pnet *n = c->pnet;
  const int nx = cen.ni, ny = cen.nj, nz = cen.nk;
  const long long V = (long long)nx * ny * nz;
  const long long nb = box_blocks(cen);
  LAUNCH(k_pg_gather, cen, n->xin, c->part);
  hipLaunchKernelGGL(k_pg_rms, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)nb);
  hipLaunchKernelGGL(k_pg_norm, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, c->stream, n->xin, V);
  if (c->prof.on) prof_begin(c);
  int rc = pnet_forward_dev(n, c->stream, 1, nz, ny, nx);
  if (rc) return rc;
  if (c->prof.on) prof_end(c, "pnet_forward", 0.0);
  LAUNCH(k_pg_scatter, cen, n->y);
  LAUNCH(k_pg_ag, cen, c->part);
  hipLaunchKernelGGL(k_pg_alpha, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)nb);
  LAUNCH(k_pg_scale, cen);
  residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_PCG, 0);
  const double accept = k.pcorr_guess_accept > 0.0 ? k.pcorr_guess_accept : 1.0;
  LAUNCH(k_pg_fallback, cen, RS_PCG, RS_PC0, accept);
  return GFM_OK;
}

// x-plane block correction of xt on the box (p' solve relaxed to a
// tolerance): partial sums, chunk sums, plane solve, update.
static int block_corr_x(gfm_ctx *c, double *xt, int ib, int ie, int jb, int je, int kb, int ke) {
  const Box b = make_box(ib, ie, jb, je, kb, ke, 2);
  const int nchunk = (b.nj * b.nk + BC_ROWS - 1) / BC_ROWS;
  if (nchunk > 65535 || (size_t)BC_NV * b.ni * sizeof(double) > 65536) {
    set_err("block correction: box of %d x-planes and %d x %d rows exceeds the kernels' grid or LDS", b.ni, b.nj, b.nk);
    return GFM_E_ARG;
  }
  hipLaunchKernelGGL(k_bc_part, dim3((b.ni + 63) / 64, nchunk), dim3(64 * BC_ROWS), 0, c->stream, b, xt, c->bc_part);
  hipLaunchKernelGGL(k_bc_sum, dim3((b.ni + 63) / 64, BC_NV), dim3(1024), 0, c->stream, b.ni, nchunk, c->bc_part,
                     c->bc_sums);
  hipLaunchKernelGGL(k_bc_solve, dim3(1), dim3(256), BC_NV * b.ni * sizeof(double), c->stream, b.ni, c->bc_sums,
                     c->bc_d);
  LAUNCH(k_bc_apply, b, xt, c->bc_d);
  return GFM_OK;
}

// Relax p' one sweep at a time until calc_residual <= pcorr_tol (host
// loop: one residual read-back per sweep), each sweep preceded by the
// x-plane block correction when pcorr_block_corr is set.
static int pcorr_relax(gfm_ctx *c) {
  DevCase &d = c->dc;
  gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  auto resid = [&](double *r) {
    residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_PC, 0);
    HCHECK(hipMemcpyAsync(c->pin, d.scal + SC_RESID0 + RS_PC, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    *r = *c->pin;
    return GFM_OK;
  };
  double r;
  int rc = resid(&r);
  if (rc) return rc;
  k.pcorr_resid_start = r;
  int s = 0;
  while (s < k.pcorr_max_sweeps && !(r <= k.pcorr_tol)) {
    if (k.pcorr_block_corr && (rc = block_corr_x(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2))) return rc;
    adlbl3_sweeps(c, d.fz, 1, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
    s++;
    if ((rc = resid(&r))) return rc;
  }
  k.pcorr_sweeps = s;
  if (k.want_irc) residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_IRC + 2 * (5 - 1), 1);
  return GFM_OK;
}

// FLX1D's per-plane flow rates (C/Flx1d.f90:16-73) into flow_x (2 per
// x-face plane), flow_y and flow_z (4 per plane: mass +/-, enthalpy +/-).
static void plane_flows(gfm_ctx *c) {
  const gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  const Box bx = make_box(3, MPh - 1, 2, NPh, 2, LPh, 2);
  const int nch = (bx.nj * bx.nk + FLOW_ROWS - 1) / FLOW_ROWS;
  hipLaunchKernelGGL(k_flow_x, dim3((bx.ni + 63) / 64, nch), dim3(1024), 0, c->stream, bx, c->flow_part);
  hipLaunchKernelGGL(k_flow_sum, dim3((2 * bx.ni + 63) / 64), dim3(1024), 0, c->stream, nch, 2 * bx.ni, c->flow_part,
                     c->flow_x);
  const Box by = make_box(4, MPh - 2, 3, NPh - 1, 4, LPh - 2, 2);
  LAUNCH(k_flow_yz<2>, by, c->flow_part);
  hipLaunchKernelGGL(k_flow_sum, dim3((4 * by.nj + 63) / 64), dim3(1024), 0, c->stream, by.nk * ((by.ni + 63) / 64), 4 * by.nj,
                     c->flow_part, c->flow_y);
  const Box bz = make_box(4, MPh - 2, 4, NPh - 2, 3, LPh - 1, 2);
  LAUNCH(k_flow_yz<3>, bz, c->flow_part);
  hipLaunchKernelGGL(k_flow_sum, dim3((4 * bz.nk + 63) / 64), dim3(1024), 0, c->stream, bz.nj * ((bz.ni + 63) / 64), 4 * bz.nk,
                     c->flow_part, c->flow_z);
}

static void extr(gfm_ctx *c);

// ---------------------------------------------------------------------
// GFLOW, C/Gflow.f90:6-408
// ---------------------------------------------------------------------
// stage 0: the whole iteration; 1: up to the initial guess of a p' solve
// relaxed to a tolerance; 2: from the end of that solve on.
// INTPV (C/Intp.f90:9-30 calls INTP1 over the velocity classes) in three
// launches.  Per component: (1) the two edge classes interpolated from
// the face values, (2) the corner class from the first edge class, (3)
// the four classes between faces along the component's own axis from all
// of these; the three components' arrays are separate, so each stage
// takes all three.
static void intpv_stage(gfm_ctx *c, const Box *bx, const int *iu, const int *ix, int n) {
  IntpvSet s;
  memset(&s, 0, sizeof s);
  int wg = 0;
  for (int q = 0; q < n; q++) {
    const Box &b = bx[q];
    if (b.ni <= 0 || b.nj <= 0 || b.nk <= 0) continue;
    s.b[s.n] = b; s.iu[s.n] = iu[q]; s.ix[s.n] = ix[q]; s.wg0[s.n] = wg;
    wg += (int)(((long long)b.ni * b.nj * b.nk + INTPV_NT - 1) / INTPV_NT);
    s.n++;
  }
  s.wg0[s.n] = wg;
  if (wg == 0) return;
  if (c->prof.on) prof_begin(c);
  hipLaunchKernelGGL(k_intpv, dim3((unsigned)wg), dim3(INTPV_NT), 0, c->stream, s);
  if (c->prof.on) prof_end(c, "k_intpv", 0.0);
}
static void intpv(gfm_ctx *c) {
  const gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  {  // stage 1: edges from faces
    const Box bx[6] = {make_box(3, MPh - 1, 3, NPh - 1, 2, LPh, 2), make_box(3, MPh - 1, 2, NPh, 3, LPh - 1, 2),
                       make_box(2, MPh, 3, NPh - 1, 3, LPh - 1, 2), make_box(3, MPh - 1, 3, NPh - 1, 2, LPh, 2),
                       make_box(3, MPh - 1, 2, NPh, 3, LPh - 1, 2), make_box(2, MPh, 3, NPh - 1, 3, LPh - 1, 2)};
    const int iu[6] = {1, 1, 2, 2, 3, 3}, ix[6] = {2, 3, 3, 1, 1, 2};
    intpv_stage(c, bx, iu, ix, 6);
  }
  {  // stage 2: corners
    const Box cb = make_box(3, MPh - 1, 3, NPh - 1, 3, LPh - 1, 2);
    const Box bx[3] = {cb, cb, cb};
    const int iu[3] = {1, 2, 3}, ix[3] = {3, 1, 2};
    intpv_stage(c, bx, iu, ix, 3);
  }
  {  // stage 3: along the component's own axis
    Box bx[12];
    int iu[12], ix[12], n = 0;
    for (int pj = 0; pj < 2; pj++)
      for (int pk = 0; pk < 2; pk++) {
        bx[n] = make_box(4, MPh, 2 + pj, pj ? NPh - 1 : NPh, 2 + pk, pk ? LPh - 1 : LPh, 2); iu[n] = 1; ix[n] = 1; n++;
      }
    for (int pi = 0; pi < 2; pi++)
      for (int pk = 0; pk < 2; pk++) {
        bx[n] = make_box(2 + pi, pi ? MPh - 1 : MPh, 4, NPh, 2 + pk, pk ? LPh - 1 : LPh, 2); iu[n] = 2; ix[n] = 2; n++;
      }
    for (int pi = 0; pi < 2; pi++)
      for (int pj = 0; pj < 2; pj++) {
        bx[n] = make_box(2 + pi, pi ? MPh - 1 : MPh, 2 + pj, pj ? NPh - 1 : NPh, 4, LPh, 2); iu[n] = 3; ix[n] = 3; n++;
      }
    intpv_stage(c, bx, iu, ix, n);
  }
}

static int gflow(gfm_ctx *c, int stage) {
  const gfm_ctrl &k = c->ctrl;
  DevCase &d = c->dc;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  Box cen = make_box(4, MPh - 2, 4, NPh - 2, 4, LPh - 2, 2);
  Box cen_all = make_box(2, MPh, 2, NPh, 2, LPh, 2);
  Box mom[4];
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    int nx0 = iu1 == 1;
    mom[iu1] = make_box(4 + nx0, MPh - 2 - nx0, d.njy1v[iu1], d.njy2v[iu1], d.lkz1v[iu1], d.lkz2v[iu1], 2);
  }

  if (stage != 2) {
  // C/Gflow.f90:16-130 over the positions where it changes a field
  const double rt20 = k.reyg2 * sqrt(k.epsr) * pow(CMU, 0.25);
  if (c->n_wall_open > 0) {
    const dim3 gw((unsigned)((c->n_wall_open + 255) / 256));
    hipLaunchKernelGGL(k_theta_wall, gw, dim3(256), 0, c->stream, c->wall_open, c->n_wall_open);
  }
  if (c->n_wall_pos > 0)
    hipLaunchKernelGGL(k_gama_wall_zero, dim3((unsigned)((c->n_wall_pos + 255) / 256)), dim3(256), 0, c->stream, c->wall_pos,
                       c->n_wall_pos);
  if (c->n_wall_open > 0) {
    const dim3 gw((unsigned)((c->n_wall_open + 255) / 256));
    hipLaunchKernelGGL(k_wall_function, gw, dim3(256), 0, c->stream, c->wall_open, c->n_wall_open, rt20);
  }

  // ---- GSLV3V ----
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    int nx0 = iu1 == 1, nx1 = iu1 == 2, nx2 = iu1 == 3;
    Box fb = make_box(2 + nx0, MPh - nx0, 2 + nx1, NPh - nx1, 2 + nx2, LPh - nx2, 2);
    LAUNCHT(k_copy_ug_fz, 1, fb, iu1);
  }
  LAUNCHT(k_momentum_assemble<1>, ROWS_ONLY, mom[1]);
  LAUNCHT(k_momentum_assemble<2>, ROWS_ONLY, mom[2]);
  LAUNCHT(k_momentum_assemble<3>, ROWS_ONLY, mom[3]);
  LAUNCHT(k_pressure_assemble, ROWS_ONLY, cen);
  if (k.want_pre_residuals) residual(c, d.p, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_P_PRE, 0);
  adlbl3(c, d.p, 4, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
  residual(c, d.p, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_P, 0);
  LAUNCHT(k_bound_p, 1, cen);
  for (int iu1 = 1; iu1 <= 3; iu1++) {
    const Box &b = mom[iu1];
    int ie = b.i0 + 2 * (b.ni - 1), je = b.j0 + 2 * (b.nj - 1), ke = b.k0 + 2 * (b.nk - 1);
    LAUNCHT(k_starred, 0, b, iu1);
    if (k.want_pre_residuals) residual(c, d.fz, b.i0, ie, b.j0, je, b.k0, ke, RS_U_PRE + iu1 - 1, 0);
    adlbl3(c, d.fz, iu1, b.i0, ie, b.j0, je, b.k0, ke);
    residual(c, d.fz, b.i0, ie, b.j0, je, b.k0, ke, RS_U + iu1 - 1, 0);
    LAUNCHT(k_bound_store_ug, 1, b, iu1);
  }
  // p'
    LAUNCH(k_pcorr_source, cen, c->part);
    hipLaunchKernelGGL(k_pcorr_finalize, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, box_blocks(cen));
    LAUNCHT(k_zero_fz, 1, cen_all);
    if (k.pcorr_mode == GFM_PCORR_REFERENCE) {
      adlbl3(c, d.fz, 5, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
      residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_PC, 0);
    } else {
      int rc = pcorr_guess(c, cen);
      if (rc) return rc;
    }
  }
  if (stage == 1) {
    HCHECK(hipGetLastError());
    return GFM_OK;
  }
  if (k.pcorr_dump) {
    LAUNCH(k_pg_gather, cen, c->dump_x, c->part);
    LAUNCH(k_pg_dump_p, cen, c->dump_p);
  }
  for (int iu1 = 1; iu1 <= 3; iu1++) LAUNCHT(k_vel_correct, 0, mom[iu1], iu1);
  // EXTRV(0), FLX1D, factor, EXTRV(1)
  Box ex[3] = {make_box(3, MPh - 1, 2, NPh, 2, LPh, 2), make_box(2, MPh, 3, NPh - 1, 2, LPh, 2), make_box(2, MPh, 2, NPh, 3, LPh - 1, 2)};
  {
    // three directions accumulate into one set of partials: run them in
    // order and sum each into the scalar block
    for (int q = 0; q < 3; q++) {
      LAUNCHF(k_extrv, ex[q], q + 1, 0, c->part);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)box_blocks_f(ex[q]), 5, SC_GFEX, q > 0);
    }
  }
  if (k.want_plane_flows) plane_flows(c);
  hipLaunchKernelGGL(k_factor, dim3(1), dim3(1), 0, c->stream);
  for (int q = 0; q < 3; q++) LAUNCHF(k_extrv, ex[q], q + 1, 1, c->part);
  intpv(c);
  // pressure reference
  hipLaunchKernelGGL(k_pref, dim3(1), dim3(1), 0, c->stream);

  // ---- GSLV3G ----
  const int Ls[7] = {1, 2, 3, 4, 5, 6, 7};
  const int nelgs[7] = {6, 12, 8, 9, 10, 7, 11};
  const double sigma[7] = {0.7, 0.7, 0.7, 1.0, 1.0, 0.7, 0.7};
  for (int e = 0; e < 7; e++) {
    int l = Ls[e], nelg = nelgs[e];
    LAUNCHT(k_copy_gf_fz, 1, cen_all, l);
    int igs = 1;
    if (e == 0 || c->sgm != sigma[e]) { c->sgm = sigma[e]; igs = 0; }
    switch (nelg) {
      case 6: LAUNCHT(k_scalar_assemble<6>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
      case 7: LAUNCHT(k_scalar_assemble<7>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
      case 8: LAUNCH(k_scalar_assemble<8>, cen, l, igs, c->sgm, c->part); break;
      case 9: LAUNCHT(k_scalar_assemble<9>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
      case 10: LAUNCHT(k_scalar_assemble<10>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
      case 11: LAUNCHT(k_scalar_assemble<11>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
      default: LAUNCHT(k_scalar_assemble<12>, ROWS_ONLY, cen, l, igs, c->sgm, c->part); break;
    }
    if (nelg == 8) {
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)box_blocks(cen), 1, SC_QREACT, 0);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part + c->dc.pstride, (int)box_blocks(cen), 1, SC_QWALL, 0);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part + 2 * c->dc.pstride, (int)box_blocks(cen), 1, SC_QWALL_LO, 0);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part + 3 * c->dc.pstride, (int)box_blocks(cen), 1, SC_QWALL_HI, 0);
    }
    if (k.want_pre_residuals) residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_GF_PRE + l - 1, 0);
    adlbl3(c, d.fz, nelg, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2);
    residual(c, d.fz, 4, MPh - 2, 4, NPh - 2, 4, LPh - 2, RS_GF + l - 1, 0);
    LAUNCHT(k_bound_store_gf, 1, cen, l);
  }
  c->sgm = 1.0;  // GSLV3V resets SGM at the next iteration
  // properties, T_ave, EXTR(1), INTP
  LAUNCHT(k_props, 1, cen_all);
  LAUNCH(k_tave, cen, c->part);
  hipLaunchKernelGGL(k_tave_finalize, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)box_blocks(cen));
  extr(c);

  {
    Box b = make_box(2, MPh, 2, NPh, 3, LPh - 1, 2); LAUNCHT(k_intp<3>, 0, b);
    b = make_box(2, MPh, 3, NPh - 1, 2, LPh, 2); LAUNCHT(k_intp<2>, 0, b);
    b = make_box(2, MPh, 3, NPh - 1, 3, LPh - 1, 2); LAUNCHT(k_intp<2>, 0, b);
    for (int pj = 0; pj < 2; pj++)
      for (int pk = 0; pk < 2; pk++) {
        b = make_box(3, MPh - 1, 2 + pj, pj ? NPh - 1 : NPh, 2 + pk, pk ? LPh - 1 : LPh, 2);
        LAUNCHT(k_intp<1>, 0, b);
      }
  }
  // SPHASE's exit temperatures after the GFLOW call, C/Sphase.f90:74-89
  LAUNCH(k_exit_t, cen_all, c->part);
  hipLaunchKernelGGL(k_exit_t_finalize, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)box_blocks(cen_all));
  HCHECK(hipGetLastError());
  return GFM_OK;
}

// EXTR(1) over the inlet/exit and wall cell lists, C/Extr.f90:31-250
static void extr(gfm_ctx *c) {
  const unsigned nbio = (unsigned)((c->n_extr_io + 63) / 64 > 0 ? (c->n_extr_io + 63) / 64 : 1);
  hipLaunchKernelGGL(k_extr_inout, dim3(nbio), dim3(64), 0, c->stream, c->extr_io, c->n_extr_io, c->part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part, (int)nbio, 1, SC_GFIN0, 0);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(FIN_NT), 0, c->stream, c->part + c->dc.pstride, (int)nbio, 1, SC_QEXH, 0);
  if (c->n_extr_wall > 0)
    hipLaunchKernelGGL(k_extr_wall, dim3((unsigned)((c->n_extr_wall + 255) / 256)), dim3(256), 0, c->stream, c->extr_wall,
                       c->n_extr_wall);
}

static int initsv_tail(gfm_ctx *c) {
  const gfm_ctrl &k = c->ctrl;
  const int MPh = k.mp, NPh = k.np, LPh = k.lp;
  intpv(c);
  Box b;
  extr(c);
  b = make_box(2, MPh, 2, NPh, 3, LPh - 1, 2); LAUNCHT(k_intp<3>, 0, b);
  b = make_box(2, MPh, 3, NPh - 1, 2, LPh, 2); LAUNCHT(k_intp<2>, 0, b);
  b = make_box(2, MPh, 3, NPh - 1, 3, LPh - 1, 2); LAUNCHT(k_intp<2>, 0, b);
  for (int pj = 0; pj < 2; pj++)
    for (int pk = 0; pk < 2; pk++) { b = make_box(3, MPh - 1, 2 + pj, pj ? NPh - 1 : NPh, 2 + pk, pk ? LPh - 1 : LPh, 2); LAUNCHT(k_intp<1>, 0, b); }
  HCHECK(hipGetLastError());
  return GFM_OK;
}

// ---------------------------------------------------------------------
// calc_residual allocating nothing (gfm_hip_calc_residual_noalloc): the
// system in device memory in the reference layout, partial sums in a
// static device array.
// ---------------------------------------------------------------------
#define NOALLOC_NT 256
#define NOALLOC_MAX_DEVICES 64
// Per-row partial sums; host code never passes this name to a kernel (it
// is a host-side stand-in there), only its per-device address from
// hipGetSymbolAddress.
static __device__ __align__(256) double scratch[GFM_NOALLOC_MAX_ROWS];
static double *g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// One workgroup per (J, K) row of the box: the normalised residual terms
// of calc_residual (C/Adlbl3.f90:239-271) of the row's cells, each lane
// summing its cells in i order, the lanes' sums added in a fixed order.
__global__ __launch_bounds__(NOALLOC_NT) void k_resid_rows(int mp, int np, long long nlat, int ni, int nj, int ib, int jb,
                                                           int kb, const int16_t *ibc, const double *ap, const double *as,
                                                           const double *bs, const double *xt, double *part) {
  const int r = blockIdx.x;
  const int j = jb + 2 * (r % nj), k = kb + 2 * (r / nj);
  const long long base = ((long long)(k - 1) * np + (j - 1)) * mp - 1;  // linear index of (i, j, k) is base + i
  const long long sj = 2LL * mp, sk = 2LL * mp * np;
  double s = 0.0;
  for (int t = threadIdx.x; t < ni; t += NOALLOC_NT) {
    const long long q = base + ib + 2 * t;
    if (ibc[q] > 0) continue;
    const double t1 = as[q] * xt[q - 2], t2 = as[3 * nlat + q] * xt[q + 2];
    const double t3 = as[nlat + q] * xt[q - sj], t4 = as[4 * nlat + q] * xt[q + sj];
    const double t5 = as[2 * nlat + q] * xt[q - sk], t6 = as[5 * nlat + q] * xt[q + sk];
    const double t7 = ap[q] * xt[q], t8 = bs[q];
    const double eq_max =
        fmax(fmax(fmax(fabs(t1), fabs(t2)), fmax(fabs(t3), fabs(t4))), fmax(fmax(fabs(t5), fabs(t6)), fmax(fabs(t7), fabs(t8))));
    s += fabs(t1 + t2 + t3 + t4 + t5 + t6 + t8 - t7) / eq_max;
  }
  __shared__ double sh[NOALLOC_NT / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < NOALLOC_NT / 64; w++) v += sh[w];
    part[r] = v;
  }
}
// resid = max(sum of the row sums / ncells, 1e-16), summed in a fixed order
__global__ __launch_bounds__(1024) void k_resid_rows_finalize(const double *part, int nrows, int ncells, double *resid) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (int r = threadIdx.x; r < nrows; r += 1024) v += part[r];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) resid[0] = fmax(sh[0] / ncells, 1.0e-16);
}

// ---------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------
#include "gfm_minor.hip"
#include "gfm_rad.hip"
#include "gfm_wall.hip"

extern "C" {

const char *gfm_hip_last_error(void) { return g_err.c_str(); }
int gfm_hip_abi_version(void) { return GFM_ABI_VERSION; }
size_t gfm_hip_ctrl_size(void) { return sizeof(gfm_ctrl); }

int gfm_hip_init(const gfm_ctrl *k, const gfm_host_arrays *h, int device, gfm_ctx **out) {
  if (!k || !h || !out) { set_err("null argument"); return GFM_E_ARG; }
  *out = nullptr;
  if (k->mp < 8 || k->np < 8 || k->lp < 8 || (k->mp & 1) || (k->np & 1) || (k->lp & 1)) {
    set_err("bad lattice dimensions %d %d %d", k->mp, k->np, k->lp);
    return GFM_E_ARG;
  }
  if ((k->np - 6) / 2 + 1 > 254 || (k->lp - 6) / 2 + 1 > 254) {
    set_err("line length above %d cells is not supported", 254);
    return GFM_E_ARG;
  }
  if (k->sweep_order != GFM_ORDER_REFERENCE && k->sweep_order != GFM_ORDER_ZEBRA) { set_err("bad sweep_order"); return GFM_E_ARG; }
  if (k->precision != GFM_PREC_FP64 && k->precision != GFM_PREC_MIXED) { set_err("bad precision mode %d", k->precision); return GFM_E_ARG; }
  const long long nlat = (long long)k->mp * k->np * k->lp;
  for (long long q = 0; q < nlat; q++)
    if (h->ibcell[q] < -2) { set_err("cell code %d below -2 (half-blocked cells are -1 and -2)", (int)h->ibcell[q]); return GFM_E_ARG; }
  int ndev = 0;
  HCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { set_err("device %d not present (%d devices)", device, ndev); return GFM_E_DEVICE; }
  HCHECK(hipSetDevice(device));
  gfm_ctx *c = new gfm_ctx();
  c->device = device;
  c->ctrl = *k;
  {
    const char *ev = getenv("GFM_HIP_GRAPH");
    c->use_graph = !(ev && ev[0] == '0');
    const char *ls = getenv("GFM_LINE_SOLVE");
    c->line_pcr = ls && strcmp(ls, "pcr") == 0;
  }
  c->nlat = nlat;
  DevGeom &gg = c->dc.g;
  gg.MP = k->mp; gg.NP = k->np; gg.LP = k->lp;
  gg.NXC = k->mp / 2 + 1; gg.NYC = k->np / 2 + 1; gg.NZC = k->lp / 2 + 1;
  gg.HX = ((gg.NXC + 1) / 2 + 31 + 31) / 32 * 32;
  gg.NXP = 2 * gg.HX;
  gg.S = (long long)gg.NXP * gg.NYC * gg.NZC;
  c->ncls = gg.S * 8;
  int rc = 0;
  auto fail = [&](int r) { gfm_hip_finalize(c); return r; };
  if (c->ncls * (long long)sizeof(double) >= (1LL << 32)) {  // the line kernel addresses arrays by 32-bit byte offsets
    set_err("lattice of %lld entries per array exceeds 4 GB of FP64", c->ncls);
    return fail(GFM_E_ARG);
  }
  HCHECK(hipStreamCreate(&c->stream));
  DevCase &d = c->dc;
  const long long N = c->ncls;
#define AL(f) if ((rc = dalloc(c, &d.f, N))) return fail(rc);
  FIELD_LIST(AL)
  AL(bsh) AL(fz) AL(bs) AL(sfp) AL(ap)
#undef AL
  for (int q = 0; q < 3; q++) if ((rc = dalloc(c, &d.ug[q], N))) return fail(rc);
  for (int q = 0; q < 9; q++) if ((rc = dalloc(c, &d.gf[q], N))) return fail(rc);
  for (int q = 0; q < 6; q++) if ((rc = dalloc(c, &d.as[q], N))) return fail(rc);
  d.mixed = (k->precision == GFM_PREC_MIXED) ? 1 : 0;
  if (d.mixed) {
    if ((rc = dalloc(c, &d.apf, N))) return fail(rc);
    if ((rc = dalloc(c, &d.bsf, N))) return fail(rc);
    for (int q = 0; q < 6; q++) if ((rc = dalloc(c, &d.asf[q], N))) return fail(rc);
  }
  if (d.mixed) {
    float *xf;
    if ((rc = dalloc(c, &xf, N))) return fail(rc);
    c->xtf = xf;
  }
  HCHECK(hipHostMalloc((void **)&c->pin, 8 * sizeof(double), hipHostMallocDefault));
  {
    const long long cells = (long long)((k->mp - 6) / 2 + 1) * ((k->np - 6) / 2 + 1) * ((k->lp - 6) / 2 + 1);
    if ((rc = dalloc(c, &c->dump_x, 2 * cells))) return fail(rc);
    if ((rc = dalloc(c, &c->dump_p, cells))) return fail(rc);
    const long long ni = (k->mp - 6) / 2 + 1, nrow = (long long)((k->np - 6) / 2 + 1) * ((k->lp - 6) / 2 + 1);
    if ((rc = dalloc(c, &c->bc_part, ((nrow + BC_ROWS - 1) / BC_ROWS) * BC_NV * ni))) return fail(rc);
    if ((rc = dalloc(c, &c->bc_sums, BC_NV * ni))) return fail(rc);
    if ((rc = dalloc(c, &c->bc_d, ni))) return fail(rc);
    // FLX1D: x planes over all (J, K) rows; y and z planes over interior rows
    const long long px = (k->mp - 2) / 2, py = (k->np - 2) / 2, pz = (k->lp - 2) / 2;  // face planes 3..M-1
    const long long cx = (k->np / 2) * (long long)(k->lp / 2), ix = (k->mp - 6) / 2 + 1;
    const long long nbx = (ix + 63) / 64;
    const long long fx = ((cx + FLOW_ROWS - 1) / FLOW_ROWS) * 2 * px;
    const long long fy = ((k->lp - 6) / 2 + 1) * nbx * 4 * py, fz = ((k->np - 6) / 2 + 1) * nbx * 4 * pz;
    if ((rc = dalloc(c, &c->flow_part, std::max(fx, std::max(fy, fz))))) return fail(rc);
    if ((rc = dalloc(c, &c->flow_x, 2 * px))) return fail(rc);
    if ((rc = dalloc(c, &c->flow_y, 4 * py))) return fail(rc);
    if ((rc = dalloc(c, &c->flow_z, 4 * pz))) return fail(rc);
  }
  if ((rc = dalloc(c, &c->xt2, N))) return fail(rc);
  if ((rc = dalloc(c, &c->xt3, N))) return fail(rc);
  if ((rc = dalloc(c, &c->xt4, N))) return fail(rc);
  if ((rc = dalloc(c, &c->xt5, N))) return fail(rc);
  if ((rc = dalloc(c, &c->xg, N))) return fail(rc);
  if ((rc = dalloc(c, &c->stage, N))) return fail(rc);
  if ((rc = dalloc(c, &c->stage16, N))) return fail(rc);
  // partial sums: 5 slots of the largest reduction grid
  long long maxb = (long long)((gg.NXP + 63) / 64) * gg.NYC * gg.NZC + 64;
  d.pstride = maxb;
  if ((rc = dalloc(c, &c->part, 5 * maxb))) return fail(rc);
  double *scal;
  if ((rc = dalloc(c, &scal, SC_COUNT))) return fail(rc);
  d.scal = scal;
  int16_t *ibd;
  if ((rc = dalloc(c, &ibd, N))) return fail(rc);
  d.ib = ibd;
  int8_t *icode;
  if ((rc = dalloc(c, &icode, N))) return fail(rc);
  d.intp_code = icode;
  double *g1[7];
  const double *h1[7] = {h->x, h->dx, h->r, h->dr, h->rs, h->z, h->dz};
  int n1[7] = {k->mp, k->mp, k->np, k->np, k->np, k->lp, k->lp};
  for (int q = 0; q < 7; q++) {
    if ((rc = dalloc(c, &g1[q], n1[q]))) return fail(rc);
    HCHECK(hipMemcpy(g1[q], h1[q], n1[q] * sizeof(double), hipMemcpyHostToDevice));
  }
  d.x = g1[0]; d.dx = g1[1]; d.r = g1[2]; d.dr = g1[3]; d.rs = g1[4]; d.z = g1[5]; d.dz = g1[6];
  c->hx.assign(h->x, h->x + k->mp); c->hdx.assign(h->dx, h->dx + k->mp);
  c->hr.assign(h->r, h->r + k->np); c->hdr.assign(h->dr, h->dr + k->np);
  c->hz.assign(h->z, h->z + k->lp); c->hdz.assign(h->dz, h->dz + k->lp);
  double *kt[4];
  const double *hk[4] = {h->rtim, h->delrt, h->delr, h->timed};
  for (int q = 0; q < 4; q++) {
    if ((rc = dalloc(c, &kt[q], k->nr0))) return fail(rc);
    HCHECK(hipMemcpy(kt[q], hk[q], k->nr0 * sizeof(double), hipMemcpyHostToDevice));
  }
  d.rtim = kt[0]; d.delrt = kt[1]; d.delr = kt[2]; d.timed = kt[3];
  fill_devcase_scalars(c, k);
  if ((rc = bind(c))) return fail(rc);
  {
    // LDS of k_lines beyond the default 64 KB for long lines
    int nmax = (k->np > k->lp ? k->np : k->lp) / 2;
    size_t shm = lines_shm(d.mixed ? 32 : 16, nmax, d.mixed ? sizeof(float) : sizeof(double));
    if (shm > 64 * 1024) {
      if ((rc = d.mixed ? set_lines_attr<float>(shm) : set_lines_attr<double>(shm))) return fail(rc);
    }
  }
  // cell codes and INTP weight codes
  c->ibh.assign(h->ibcell, h->ibcell + nlat);
  HCHECK(hipMemcpy(c->stage16, h->ibcell, nlat * sizeof(int16_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_scatter_i16, dim3((unsigned)((nlat + 255) / 256)), dim3(256), 0, c->stream, c->stage16, ibd, nlat);
  std::vector<int8_t> code;
  intp_codes(k, h->ibcell, code, gg);
  HCHECK(hipMemcpy(icode, code.data(), N, hipMemcpyHostToDevice));
  {
    std::vector<long long> ls[4];
    wall_lists(k, h->ibcell, ls[0], ls[1]);
    extr_lists(k, h->ibcell, ls[2], ls[3]);
    long long **dst[4] = {&c->wall_open, &c->wall_pos, &c->extr_io, &c->extr_wall};
    int *cnt[4] = {&c->n_wall_open, &c->n_wall_pos, &c->n_extr_io, &c->n_extr_wall};
    for (int q = 0; q < 4; q++) {
      if (ls[q].size() > 0x7fffffff) { set_err("position list too long"); return fail(GFM_E_ARG); }
      *cnt[q] = (int)ls[q].size();
      if ((rc = dalloc(c, dst[q], (long long)ls[q].size() + 1))) return fail(rc);
      if (!ls[q].empty()) HCHECK(hipMemcpy(*dst[q], ls[q].data(), ls[q].size() * sizeof(long long), hipMemcpyHostToDevice));
    }
  }
  if ((rc = upload_state(c, k, h))) return fail(rc);
  *out = c;
  return GFM_OK;
}

int gfm_hip_upload(gfm_ctx *c, const gfm_ctrl *k, const gfm_host_arrays *h) {
  if (!c || !k || !h) { set_err("null argument"); return GFM_E_ARG; }
  c->ctrl = *k;
  fill_devcase_scalars(c, k);
  g_bound = nullptr;
  return upload_state(c, k, h);
}

int gfm_hip_download(gfm_ctx *c, gfm_ctrl *k, gfm_host_arrays *h) {
  if (!c || !h) { set_err("null argument"); return GFM_E_ARG; }
  int rc = bind(c);
  if (rc) return rc;
  DevCase &d = c->dc;
#define DN_(f) if ((rc = gather(c, d.f, h->f))) return rc;
  FIELD_LIST(DN_)
#undef DN_
  if ((rc = gather(c, d.bsh, h->bs_hprev))) return rc;
  for (int q = 0; q < 3; q++) if ((rc = gather(c, d.ug[q], h->ug + (size_t)q * c->nlat))) return rc;
  for (int q = 0; q < 9; q++) if ((rc = gather(c, d.gf[q], h->gf + (size_t)q * c->nlat))) return rc;
  HCHECK(hipStreamSynchronize(c->stream));
  if (k) *k = c->ctrl;
  return GFM_OK;
}

int gfm_hip_initsv_tail(gfm_ctx *c, gfm_ctrl *k) {
  if (!c) { set_err("null context"); return GFM_E_ARG; }
  int rc = bind(c);
  if (rc) return rc;
  if ((rc = initsv_tail(c))) return rc;
  double sc[SC_COUNT];
  HCHECK(hipMemcpyAsync(sc, c->dc.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  c->ctrl.gfin0 = sc[SC_GFIN0];
  c->ctrl.q_exhaust = sc[SC_QEXH];
  if (k) { k->gfin0 = c->ctrl.gfin0; k->q_exhaust = c->ctrl.q_exhaust; }
  return GFM_OK;
}

static void drop_graph(gfm_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->graph_exec_b) (void)hipGraphExecDestroy(c->graph_exec_b);
  if (c->graph_b) (void)hipGraphDestroy(c->graph_b);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_exec_b = nullptr;
  c->graph_b = nullptr;
  c->graph_key = -1;
}

// Launch one GFLOW.  The launch sequence depends only on the options in
// the key, so it is captured once into a graph and replayed; itr_gas
// reaches the kernels through the device scalar block.
static int capture(gfm_ctx *c, int stage, hipGraph_t *gr, hipGraphExec_t *ex) {
  HCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = gflow(c, stage);
  hipGraph_t g0 = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g0);
  if (rc) { if (g0) (void)hipGraphDestroy(g0); return rc; }
  if (e != hipSuccess) { set_err("graph capture failed: %s", hipGetErrorString(e)); return GFM_E_DEVICE; }
  *gr = g0;
  HCHECK(hipGraphInstantiateWithFlags(ex, g0, 0));
  return GFM_OK;
}
static int run_gflow(gfm_ctx *c) {
  hipLaunchKernelGGL(k_set_itr, dim3(1), dim3(1), 0, c->stream, c->ctrl.itr_gas);
  const gfm_ctrl &k = c->ctrl;
  const bool split = k.pcorr_mode == GFM_PCORR_TOLERANCE;
  int rc;
  if (!c->use_graph || c->prof.on) {
    if (!split) return gflow(c, 0);
    if ((rc = gflow(c, 1))) return rc;
    if ((rc = pcorr_relax(c))) return rc;
    return gflow(c, 2);
  }
  const int key = k.sweep_order | (k.want_pre_residuals << 2) | (k.want_irc << 3) | ((int)split << 4) |
                  ((k.pcorr_guess == GFM_PGUESS_LEARNED && c->pnet) << 5) | ((k.pcorr_dump != 0) << 6) |
                  ((k.want_plane_flows != 0) << 7);
  if (!c->graph_exec || key != c->graph_key) {
    drop_graph(c);
    if ((rc = capture(c, split ? 1 : 0, &c->graph, &c->graph_exec))) return rc;
    if (split && (rc = capture(c, 2, &c->graph_b, &c->graph_exec_b))) return rc;
    c->graph_key = key;
  }
  HCHECK(hipGraphLaunch(c->graph_exec, c->stream));
  if (!split) return GFM_OK;
  if ((rc = pcorr_relax(c))) return rc;
  HCHECK(hipGraphLaunch(c->graph_exec_b, c->stream));
  return GFM_OK;
}

int gfm_hip_gflow(gfm_ctx *c, gfm_ctrl *k) {
  if (!c || !k) { set_err("null argument"); return GFM_E_ARG; }
  int rc = bind(c);
  if (rc) return rc;
  c->ctrl.itr_gas = k->itr_gas;
  c->ctrl.sweep_order = k->sweep_order;
  c->ctrl.want_pre_residuals = k->want_pre_residuals;
  c->ctrl.want_irc = k->want_irc;
  c->ctrl.want_plane_flows = k->want_plane_flows;
  if (k->gray_radiation != c->ctrl.gray_radiation || k->i1st_rad_done != c->ctrl.i1st_rad_done) {
    // the radiation source and the wall treatment of EXTR change after the first radiation pass
    if (k->gray_radiation < GFM_RADSRC_NONE || k->gray_radiation > GFM_RADSRC_FIELD) { set_err("bad gray_radiation"); return GFM_E_ARG; }
    if (k->gray_radiation == GFM_RADSRC_FIELD && !c->dc.rad_qa) {
      set_err("gray_radiation = GFM_RADSRC_FIELD needs a radiation pass (gfm_hip_rad_emis) or gfm_hip_rad_upload first");
      return GFM_E_STATE;
    }
    c->ctrl.gray_radiation = k->gray_radiation;
    c->ctrl.i1st_rad_done = k->i1st_rad_done;
    c->dc.gray_radiation = k->gray_radiation;
    c->dc.i1st_rad_done = k->i1st_rad_done;
    g_bound = nullptr;
    if ((rc = bind(c))) return rc;
  }
  if (k->pcorr_mode != GFM_PCORR_REFERENCE && k->pcorr_mode != GFM_PCORR_TOLERANCE) { set_err("bad pcorr_mode"); return GFM_E_ARG; }
  if (k->pcorr_mode == GFM_PCORR_TOLERANCE && (k->pcorr_max_sweeps < 0 || !(k->pcorr_tol >= 0.0))) {
    set_err("pcorr_max_sweeps must be >= 0 and pcorr_tol >= 0");
    return GFM_E_ARG;
  }
  c->ctrl.pcorr_mode = k->pcorr_mode;
  c->ctrl.pcorr_guess = k->pcorr_guess;
  c->ctrl.pcorr_max_sweeps = k->pcorr_max_sweeps;
  c->ctrl.pcorr_tol = k->pcorr_tol;
  c->ctrl.pcorr_block_corr = k->pcorr_block_corr;
  c->ctrl.pcorr_dump = k->pcorr_dump;
  if (c->ctrl.pcorr_guess_accept != k->pcorr_guess_accept) drop_graph(c);  // captured as a kernel argument
  c->ctrl.pcorr_guess_accept = k->pcorr_guess_accept;
  if ((rc = run_gflow(c))) return rc;
  double sc[SC_COUNT];
  HCHECK(hipMemcpyAsync(sc, c->dc.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  pull_scalars(c, &c->ctrl, sc);
  pull_scalars(c, k, sc);
  if (c->ctrl.pcorr_mode == GFM_PCORR_REFERENCE) {
    c->ctrl.pcorr_sweeps = c->ctrl.ntimesg[4];
    c->ctrl.pcorr_resid_start = 0.0;
  }
  if (c->ctrl.pcorr_mode == GFM_PCORR_REFERENCE || c->ctrl.pcorr_guess != GFM_PGUESS_LEARNED || !c->pnet) {
    c->ctrl.pcorr_alpha = 0.0;
    c->ctrl.pcorr_gain = 0.0;
    c->ctrl.pcorr_fallback = 0;
  }
  k->pcorr_sweeps = c->ctrl.pcorr_sweeps;
  k->pcorr_resid_start = c->ctrl.pcorr_resid_start;
  k->pcorr_alpha = c->ctrl.pcorr_alpha;
  k->pcorr_gain = c->ctrl.pcorr_gain;
  k->pcorr_fallback = c->ctrl.pcorr_fallback;
  if (c->ctrl.pcorr_dump) c->pair_ready = 1;
  if (c->ctrl.want_plane_flows) c->flows_formed = 1;
  if (c->prof.on) prof_flush(c);
  return GFM_OK;
}

// Autotuner of the box kernels' launch shapes: iters GFLOW iterations per
// candidate shape (all tunable kernels on the same candidate, each kernel
// timed with events around its launches), then each kernel keeps its
// fastest candidate.  The iterations are ordinary iterations of the run
// (itr_gas advances by 6 * iters): every shape computes the same values.
int gfm_hip_autotune(gfm_ctx *c, gfm_ctrl *k, int iters, char *report, int report_len) {
  if (!c || !k || iters < 1) { set_err("null argument or iters < 1"); return GFM_E_ARG; }
  int rc = bind(c);
  if (rc) return rc;
  const int NC = (int)(sizeof TUNE_CANDS / sizeof TUNE_CANDS[0]);
  std::map<std::string, std::vector<double>> us;
  const bool prof_was = c->prof.on;
  if (prof_was) prof_flush(c);
  std::vector<ProfRec> saved = c->prof.recs;
  c->tuned.clear();
  drop_graph(c);
  for (int cand = 0; cand < NC && rc == GFM_OK; cand++) {
    c->tune_cand = cand;
    c->prof.recs.clear();
    c->prof.on = true;
    for (int it = 0; it < iters; it++) {
      k->itr_gas += 1;
      if ((rc = gfm_hip_gflow(c, k))) break;
      if (!isfinite(k->res_mass)) { set_err("non-finite res_mass at iteration %d", k->itr_gas); rc = GFM_E_NUMERIC; break; }
    }
    prof_flush(c);
    for (const ProfRec &r : c->prof.recs)
      if (c->tune_maxbs.count(r.name) && r.launches > 0) {
        auto &v = us[r.name];
        v.resize(NC, -1.0);
        v[cand] = 1e3 * r.ms / (double)r.launches;
      }
  }
  c->tune_cand = -1;
  c->prof.recs = saved;
  c->prof.on = prof_was;
  c->ctrl.itr_gas = k->itr_gas;
  if (rc) { c->tuned.clear(); drop_graph(c); return rc; }
  std::string rep = "kernel                          default(us)  best(us)  map  workgroup\n";
  for (auto &kv : us) {
    const int mx = c->tune_maxbs[kv.first], map0 = c->tune_map0[kv.first];
    const int def = map0 == 1 ? 3 : 0;
    int best = def;
    for (int q = 0; q < NC; q++)
      if (TUNE_CANDS[q].bs <= mx && !(map0 == ROWS_ONLY && TUNE_CANDS[q].map) && kv.second[q] > 0.0 &&
          kv.second[q] < kv.second[best])
        best = q;
    c->tuned[kv.first] = TUNE_CANDS[best];
    char line[160];
    snprintf(line, sizeof line, "%-30s %12.2f %9.2f %4d %10d\n", kv.first.c_str(), kv.second[def], kv.second[best],
             TUNE_CANDS[best].map, TUNE_CANDS[best].bs);
    rep += line;
  }
  if (report && report_len > 0) snprintf(report, (size_t)report_len, "%s", rep.c_str());
  drop_graph(c);  // the next iteration captures the chosen shapes
  return GFM_OK;
}

int gfm_hip_sphase(gfm_ctx *c, gfm_ctrl *k, int max_iters, double *res_hist, int *n_done) {
  if (!c || !k) { set_err("null argument"); return GFM_E_ARG; }
  int n = 0, rc;
  while (n < max_iters && k->itr_gas < k->itr_gend && k->itr_gas < k->maxgi) {
    k->itr_gas += 1;
    if ((rc = gfm_hip_gflow(c, k))) { if (n_done) *n_done = n; return rc; }
    if (!isfinite(k->res_mass)) { set_err("non-finite res_mass at iteration %d", k->itr_gas); if (n_done) *n_done = n + 1; return GFM_E_NUMERIC; }
    if (res_hist) res_hist[n] = k->res_mass;
    n++;
    if (k->res_mass < k->bgcon) { k->itr_gend = k->itr_gas; break; }
  }
  c->ctrl.itr_gas = k->itr_gas;
  c->ctrl.itr_gend = k->itr_gend;
  if (n_done) *n_done = n;
  return GFM_OK;
}

// A box of ADLBL3/calc_residual reads XT two lattice steps outside itself
// (C/Adlbl3.f90:147-221), so it must lie in [3, M-2] along each axis with
// both ends on the same parity.
static int check_box(const gfm_ctx *c, int ib, int ie, int jb, int je, int kb, int ke) {
  const DevGeom &gg = c->dc.g;
  auto ok = [](int b, int e, int m) { return b >= 3 && b <= e && e <= m - 2 && ((e - b) & 1) == 0; };
  if (!ok(ib, ie, gg.MP) || !ok(jb, je, gg.NP) || !ok(kb, ke, gg.LP)) {
    set_err("box must satisfy 3 <= begin <= end <= M-2 with end-begin even on each axis");
    return GFM_E_ARG;
  }
  return GFM_OK;
}

int gfm_hip_adlbl3(gfm_ctx *c, const gfm_ctrl *k, int nelg, const double *ap, const double *as, const double *bs, double *xt,
                   int ib, int ie, int jb, int je, int kb, int ke, double *resg) {
  if (!c || !k || !ap || !as || !bs || !xt) { set_err("null argument"); return GFM_E_ARG; }
  if (nelg < 1 || nelg > 20) { set_err("bad nelg"); return GFM_E_ARG; }
  if (int rb = check_box(c, ib, ie, jb, je, kb, ke)) return rb;
  int rc = bind(c);
  if (rc) return rc;
  DevCase &d = c->dc;
  c->ctrl.sweep_order = k->sweep_order;
  c->ctrl.want_irc = k->want_irc;
  for (int q = 0; q < 20; q++) c->ctrl.ntimesg[q] = k->ntimesg[q];
  if ((rc = scatter_coef(c, ap, d.ap, d.apf))) return rc;
  for (int q = 0; q < 6; q++) if ((rc = scatter_coef(c, as + (size_t)q * c->nlat, d.as[q], d.asf[q]))) return rc;
  if ((rc = scatter_coef(c, bs, d.bs, d.bsf))) return rc;
  if ((rc = scatter(c, xt, d.fz))) return rc;
  adlbl3(c, d.fz, nelg, ib, ie, jb, je, kb, ke);
  if ((rc = gather(c, d.fz, xt))) return rc;
  double sc[SC_COUNT];
  HCHECK(hipMemcpyAsync(sc, d.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  if (resg && k->want_irc) { resg[0] = sc[SC_RESID0 + RS_IRC + 2 * (nelg - 1)]; resg[1] = sc[SC_RESID0 + RS_IRC + 2 * (nelg - 1) + 1]; }
  return GFM_OK;
}

int gfm_hip_calc_residual(gfm_ctx *c, const double *ap, const double *as, const double *bs, const double *xt, int ib, int ie, int jb,
                          int je, int kb, int ke, double *resid) {
  if (!c || !ap || !as || !bs || !xt || !resid) { set_err("null argument"); return GFM_E_ARG; }
  if (int rb = check_box(c, ib, ie, jb, je, kb, ke)) return rb;
  int rc = bind(c);
  if (rc) return rc;
  DevCase &d = c->dc;
  if ((rc = scatter_coef(c, ap, d.ap, d.apf))) return rc;
  for (int q = 0; q < 6; q++) if ((rc = scatter_coef(c, as + (size_t)q * c->nlat, d.as[q], d.asf[q]))) return rc;
  if ((rc = scatter_coef(c, bs, d.bs, d.bsf))) return rc;
  if ((rc = scatter(c, xt, c->xt2))) return rc;
  residual(c, c->xt2, ib, ie, jb, je, kb, ke, 63, 0);
  double sc[SC_COUNT];
  HCHECK(hipMemcpyAsync(sc, d.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  *resid = sc[SC_RESID0 + 63];
  return GFM_OK;
}

int gfm_hip_calc_residual_noalloc(int mp, int np, int lp, int ncells, const int16_t *ibcell, const double *ap,
                                  const double *as, const double *bs, const double *xt, int ib, int ie, int jb, int je,
                                  int kb, int ke, double *resid, void *stream) {
  // every check before the first HIP call
  if (mp < 8 || np < 8 || lp < 8 || (mp & 1) || (np & 1) || (lp & 1)) {
    set_err("bad lattice dimensions %d %d %d", mp, np, lp);
    return GFM_E_ARG;
  }
  auto ok = [](int b, int e, int m) { return b >= 3 && b <= e && e <= m - 2 && ((e - b) & 1) == 0; };
  if (!ok(ib, ie, mp) || !ok(jb, je, np) || !ok(kb, ke, lp)) {
    set_err("box must satisfy 3 <= begin <= end <= M-2 with end-begin even on each axis");
    return GFM_E_ARG;
  }
  const int ni = (ie - ib) / 2 + 1, nj = (je - jb) / 2 + 1, nk = (ke - kb) / 2 + 1;
  const long long nrows = (long long)nj * nk;
  if (nrows > GFM_NOALLOC_MAX_ROWS) {
    set_err("box of %lld (J,K) rows exceeds the %d partial sums of the static device array", nrows, GFM_NOALLOC_MAX_ROWS);
    return GFM_E_ARG;
  }
  if (ncells <= 0) { set_err("ncells must be positive"); return GFM_E_ARG; }
  if (!ibcell || !ap || !as || !bs || !xt || !resid) { set_err("null argument"); return GFM_E_ARG; }
  int dev = -1;
  HCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { set_err("device %d beyond the %d supported", dev, NOALLOC_MAX_DEVICES); return GFM_E_DEVICE; }
  double *part = nullptr;
  {
    // the array's address on this device, looked up once with the device current
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
      void *p = nullptr;
      HCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_dev[dev] = (double *)p;
    }
    part = g_scratch_dev[dev];
  }
  hipStream_t st = (hipStream_t)stream;
  const long long nlat = (long long)mp * np * lp;
  hipLaunchKernelGGL(k_resid_rows, dim3((unsigned)nrows), dim3(NOALLOC_NT), 0, st, mp, np, nlat, ni, nj, ib, jb, kb, ibcell, ap,
                     as, bs, xt, part);
  hipLaunchKernelGGL(k_resid_rows_finalize, dim3(1), dim3(1024), 0, st, (const double *)part, (int)nrows, ncells, resid);
  HCHECK(hipGetLastError());
  return GFM_OK;
}

int gfm_hip_set_pguess(gfm_ctx *c, gfm_pnet *net) {
  if (!c) { set_err("null context"); return GFM_E_ARG; }
  if (net) {
    const int nx = (c->ctrl.mp - 6) / 2 + 1, ny = (c->ctrl.np - 6) / 2 + 1, nz = (c->ctrl.lp - 6) / 2 + 1;
    int rc = pnet_check_grid(net, 1, nz, ny, nx);
    if (rc) return rc;
    if (net->device != c->device) { set_err("the network lives on device %d, the case on %d", net->device, c->device); return GFM_E_ARG; }
  }
  c->pnet = net;
  drop_graph(c);
  return GFM_OK;
}

int gfm_hip_pcorr_pair(gfm_ctx *c, float *x, float *p) {
  if (!c || !x || !p) { set_err("null argument"); return GFM_E_ARG; }
  if (!c->pair_ready) { set_err("no p' solve has run with pcorr_dump set"); return GFM_E_STATE; }
  int rc = bind(c);
  if (rc) return rc;
  const long long cells = (long long)((c->ctrl.mp - 6) / 2 + 1) * ((c->ctrl.np - 6) / 2 + 1) * ((c->ctrl.lp - 6) / 2 + 1);
  HCHECK(hipMemcpyAsync(x, c->dump_x, 2 * cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipMemcpyAsync(p, c->dump_p, cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  return GFM_OK;
}

int gfm_hip_plane_flows(gfm_ctx *c, double *flx, double *fly, double *flyh, double *flux_z, double *fluxh_z) {
  if (!c) { set_err("null context"); return GFM_E_ARG; }
  if (!c->flows_formed) { set_err("no iteration has run with want_plane_flows set"); return GFM_E_STATE; }
  int rc = bind(c);
  if (rc) return rc;
  const int mp = c->ctrl.mp, np = c->ctrl.np, lp = c->ctrl.lp;
  const int px = (mp - 2) / 2, py = (np - 2) / 2, pz = (lp - 2) / 2;
  std::vector<double> fx(2 * px), fy(4 * py), fz(4 * pz);
  HCHECK(hipMemcpyAsync(fx.data(), c->flow_x, fx.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipMemcpyAsync(fy.data(), c->flow_y, fy.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipMemcpyAsync(fz.data(), c->flow_z, fz.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCHECK(hipStreamSynchronize(c->stream));
  // (M,2) column-major, plane I = 3 + 2q at row I-1
  auto put = [](double *dst, int m, const std::vector<double> &src, int np_, int nv, int off) {
    if (!dst) return;
    std::fill(dst, dst + 2 * (size_t)m, 0.0);
    for (int q = 0; q < np_; q++)
      for (int v = 0; v < 2; v++) dst[(2 + 2 * q) + (size_t)m * v] = src[(size_t)nv * q + off + v];
  };
  put(flx, mp, fx, px, 2, 0);
  put(fly, np, fy, py, 4, 0);
  put(flyh, np, fy, py, 4, 2);
  put(flux_z, lp, fz, pz, 4, 0);
  put(fluxh_z, lp, fz, pz, 4, 2);
  return GFM_OK;
}

int gfm_hip_kernel_times(gfm_ctx *c, int enable, char *names, int name_len, double *ms, double *bytes, int *launches, int cap,
                         int *n) {
  if (!c) { set_err("null context"); return GFM_E_ARG; }
  prof_flush(c);
  if (enable == 1) {
    c->prof.recs.clear();
    c->prof.on = true;
  } else if (enable == 0) {
    c->prof.on = false;
  }
  int cnt = 0;
  for (const ProfRec &r : c->prof.recs) {
    if (cnt >= cap) break;
    if (names && name_len > 0) {
      strncpy(names + (size_t)cnt * name_len, r.name.c_str(), name_len - 1);
      names[(size_t)cnt * name_len + name_len - 1] = 0;
    }
    if (ms) ms[cnt] = r.ms;
    if (bytes) bytes[cnt] = r.bytes;
    if (launches) launches[cnt] = (int)r.launches;
    cnt++;
  }
  if (n) *n = cnt;
  return GFM_OK;
}

int gfm_hip_finalize(gfm_ctx *c) {
  if (!c) return GFM_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) (void)hipFree(p);
  for (auto &e : c->prof.ev) (void)hipEventDestroy(e);
  drop_graph(c);
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (g_bound == c) g_bound = nullptr;
  delete c->ms;
  delete c->rad;
  delete c->wall;
  delete c;
  return GFM_OK;
}

}  // extern "C"
