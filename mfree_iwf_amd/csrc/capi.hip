// C ABI of the MI355X SPH hot path: context, bodies, stepping, parity
// helpers and scene generators.  Kernels live in grid.hip, ul.hip,
// ul32.hip and tl.hip.
#include "../../include/mfree_iwf_amd.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.h"

using namespace mfree;

namespace mfree {


hipError_t alloc8(mfree_ctx* c, Fields8& f, size_t n) {
  for (int k = 0; k < 8; k++) {
    hipError_t e = dalloc(c, &f.f[k], n);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void scene_consts(const mfree_params& p, uint32_t offset, uint32_t n, SceneConst& s) {
  // src/simulation_data.cpp:32-47
  double G = p.E / (2. * (1. + p.nu));
  double K = 2.0 * G * (1 + p.nu) / (3 * (1 - 2 * p.nu));
  double c0 = sqrt(K / p.rho0);
  memset(&s, 0, sizeof(s));
  s.rho0 = p.rho0;
  s.E = p.E;
  s.nu = p.nu;
  s.K = K;
  s.G2 = 2 * G;
  s.c0sq = c0 * c0;
  s.wdeltap = p.wdeltap;
  s.corr_exp = p.corr_exp;
  s.mghn_eps = p.mghn_eps;
  s.alpha = p.alpha;
  s.beta = p.beta;
  s.eta = p.eta;
  s.xsph_eps = p.xsph_eps;
  s.dt = p.dt;
  s.half_dt = 0.5 * p.dt;  // src/rk4.h:82,86
  s.dt6 = p.dt * 1. / 6.;  // src/rk4.h:94
  s.time = p.time;
  s.v_bc = 0.;
  s.offset = offset;
  s.n = n;
  s.cell_offset = offset;
  s.pow4 = (p.corr_exp == 4.0) ? 1u : 0u;
}

// Status of a set of error bits: a non-finite derivative is reported first
// (it makes later positions leave the grid as well).
static mfree_status bits_status(uint32_t bits, const char** msg) {
  if (bits & EB_NONFINITE) return *msg = "non-finite derivative", MFREE_ERR_NONFINITE;
  if (bits & EB_NBH) return *msg = "a particle has >= max_nbh neighbours", MFREE_ERR_NBH_OVERFLOW;
  if (bits & EB_NONEIGH) return *msg = "a particle has no neighbour", MFREE_ERR_NO_NEIGHBOR;
  if (bits & EB_CELLS) return *msg = "cell grid: num_cell >= n or a particle outside the grid", MFREE_ERR_CELLS;
  if (bits & EB_SKIN)
    return *msg = "a particle moved more than half the neighbour-list skin within one step, or faster than h/2 "
                  "per step",
           MFREE_ERR_SKIN;
  return *msg = "unknown device error bits", MFREE_ERR_HIP;
}

mfree_status check_err_bits(mfree_ctx* c) {
  // word 0: the context's bits; words 1..S: each body's (UL kernels)
  const size_t nw = 1 + (size_t)c->S;
  std::vector<uint32_t> w(nw, 0u);
  HIPCHK(c, hipMemcpyAsync(w.data(), c->errb, nw * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->opt.isolate_bodies) {
    // failed bodies stay frozen at their last good state and keep their
    // (sticky) words; the step itself succeeds unless bits were raised
    // outside any body
    uint32_t in_bodies = 0;
    for (uint32_t s = 0; s < c->S; s++) {
      c->body_bits[s] |= w[1 + s];
      in_bodies |= w[1 + s];
    }
    if (!w[0]) return MFREE_OK;
    const uint32_t zero = 0;
    HIPCHK(c, hipMemcpy(c->errb, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    const uint32_t rest = w[0] & ~in_bodies;
    if (!rest) return MFREE_OK;
    const char* msg = nullptr;
    const mfree_status st = bits_status(rest, &msg);
    return fail(c, st, msg);
  }
  if (!w[0]) return MFREE_OK;
  const std::vector<uint32_t> zero(nw, 0u);
  HIPCHK(c, hipMemcpy(c->errb, zero.data(), nw * sizeof(uint32_t), hipMemcpyHostToDevice));
  const char* msg = nullptr;
  const mfree_status st = bits_status(w[0], &msg);
  // name the first body that raised the reported condition
  for (uint32_t s = 0; s < c->S; s++) {
    const char* m2 = nullptr;
    if (w[1 + s] && bits_status(w[1 + s], &m2) == st)
      return fail(c, st, "body " + std::to_string(s) + ": " + msg);
  }
  return fail(c, st, msg);
}

mfree_status autotune_tiles(mfree_ctx* c);

mfree_status commit(mfree_ctx* c) {
  if (c->committed) return MFREE_OK;
  c->S = (uint32_t)c->prm.size();
  c->N = c->off.back();
  if (c->S == 0 || c->N == 0) return fail(c, MFREE_ERR_ARG, "no bodies");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->opt.device)
    return fail(c, MFREE_ERR_HIP, "no HIP device available (the library has no CPU path)");
  HIPCHK(c, hipSetDevice(c->opt.device));
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, c->opt.device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(c, MFREE_ERR_HIP, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t N = c->N;
  const bool tl = tl_method(c->opt.method);
  // the mixed pair passes pack a sorted position into the low 28 bits of a
  // candidate word (the upper four hold the cell offsets)
  if (!tl && c->opt.precision == MFREE_PRECISION_MIXED && N > (size_t)kListMask + 1)
    return fail(c, MFREE_ERR_UNSUPPORTED, "mixed precision UL contexts hold at most 2^28 particles");
  for (int b = 0; b < 2; b++) {
    if (b == 1 && tl) break;  // TL storage is never reordered
    HIPCHK(c, alloc8(c, c->y[b], N));
    HIPCHK(c, dalloc(c, &c->m[b], N));
    HIPCHK(c, dalloc(c, &c->h[b], N));
    HIPCHK(c, dalloc(c, &c->scene[b], N));
    HIPCHK(c, dalloc(c, &c->gid[b], N));
  }
  HIPCHK(c, alloc8(c, c->K, N));
  HIPCHK(c, alloc8(c, c->acc, N));
  // error words: [0] the context's, [1 + s] body s's (UL kernels raise both)
  HIPCHK(c, dalloc(c, &c->errb, 1 + (size_t)c->S));
  HIPCHK(c, hipMemset(c->errb, 0, (1 + (size_t)c->S) * sizeof(uint32_t)));
  c->berr = c->errb + 1;
  c->body_bits.assign(c->S, 0u);
  if (c->opt.isolate_bodies) HIPCHK(c, dalloc(c, &c->perm_keep, N));
  // bodies
  c->h_sc.resize(c->S);
  std::vector<Chunk> ch;
  std::vector<uint32_t> cfirst(c->S + 1, 0);
  for (uint32_t s = 0; s < c->S; s++) {
    scene_consts(c->prm[s], c->off[s], c->off[s + 1] - c->off[s], c->h_sc[s]);
    if (tl) c->h_sc[s].v_bc = c->v_bc[s];
    cfirst[s] = (uint32_t)ch.size();
    for (uint32_t b = c->off[s]; b < c->off[s + 1]; b += kChunk)
      ch.push_back(Chunk{s, b, std::min(b + kChunk, c->off[s + 1]), 0});
  }
  cfirst[c->S] = (uint32_t)ch.size();
  c->nchunks = (uint32_t)ch.size();
  HIPCHK(c, dalloc(c, &c->sc, c->S));
  HIPCHK(c, dalloc(c, &c->geom, c->S));
  HIPCHK(c, dalloc(c, &c->chunks, c->nchunks));
  HIPCHK(c, dalloc(c, &c->chunk_first, c->S + 1));
  HIPCHK(c, dalloc(c, &c->part, 5 * (size_t)c->nchunks));
  HIPCHK(c, hipMemcpy(c->sc, c->h_sc.data(), c->S * sizeof(SceneConst), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->chunks, ch.data(), ch.size() * sizeof(Chunk), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->chunk_first, cfirst.data(), cfirst.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  // binning / records
  HIPCHK(c, dalloc(c, &c->key, N));
  HIPCHK(c, dalloc(c, &c->rank, N));
  HIPCHK(c, dalloc(c, &c->count, N));
  HIPCHK(c, hipMemset(c->count, 0, (size_t)N * sizeof(uint32_t)));  // the binning keeps it zero between calls
  HIPCHK(c, dalloc(c, &c->cell_start, N + 1));
  HIPCHK(c, dalloc(c, &c->perm, N));
  HIPCHK(c, dalloc(c, &c->perm_tmp, 3 * (size_t)N));
  HIPCHK(c, dalloc(c, &c->tsum, (N + kScanTile - 1) / kScanTile + 1));
  HIPCHK(c, dalloc(c, &c->meta, N));
  HIPCHK(c, dalloc(c, &c->tile_off, c->S + 1));
  if (!tl) {
    if (c->opt.precision == MFREE_PRECISION_MIXED) {
      HIPCHK(c, dalloc(c, &c->pos32, N));
      HIPCHK(c, dalloc(c, &c->rec32, N));
      c->use_list = c->opt.pipeline == MFREE_PIPELINE_FUSED && c->opt.neighbor_list >= 0;
      if (c->use_list) {
        c->nl_cap = kListCap;
        c->list_lanes = c->opt.list_lanes ? c->opt.list_lanes : 1;
        c->list_staged = c->opt.list_staging >= 0;
        HIPCHK(c, dalloc(c, &c->nl_col, (size_t)kListCap * N));
        if (c->list_staged && c->list_lanes == 1) HIPCHK(c, dalloc(c, &c->nl16, (size_t)kListCap * N));
        HIPCHK(c, dalloc(c, &c->nl_deg, N));
        HIPCHK(c, dalloc(c, &c->pos32b, N));
        HIPCHK(c, dalloc(c, &c->rec32b, N));
        HIPCHK(c, dalloc(c, &c->vpart, c->nchunks));
      }
    } else {
      HIPCHK(c, dalloc(c, &c->pos64, N));
      HIPCHK(c, dalloc(c, &c->rec64, N));
    }
  }
  // initial state, storage = idx order
  double* src[8] = {c->hx.data(), c->hy.data(), c->hvx.data(), c->hvy.data(),
                    c->hrho.data(), c->hSxx.data(), c->hSxy.data(), c->hSyy.data()};
  for (int k = 0; k < 8; k++)
    HIPCHK(c, hipMemcpy(c->y[0].f[k], src[k], N * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->m[0], c->hm.data(), N * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->h[0], c->hh.data(), N * sizeof(double), hipMemcpyHostToDevice));
  std::vector<uint32_t> sc(N), gid(N);
  for (uint32_t s = 0; s < c->S; s++)
    for (uint32_t i = c->off[s]; i < c->off[s + 1]; i++) {
      sc[i] = s;
      gid[i] = i;
    }
  HIPCHK(c, hipMemcpy(c->scene[0], sc.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->gid[0], gid.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice));
  c->cur = 0;
  {
    mfree_status st = contact_alloc(c);
    if (st != MFREE_OK) return st;
  }
  c->tile_grid = (uint32_t)std::max(8, prop.multiProcessorCount) * 2;
  c->tile_grid = (c->tile_grid + kXcds - 1) / kXcds * kXcds;
  if (tl) {
    mfree_status st = tl_alloc(c->tl, tl_variant(c->opt.method), (uint32_t)N, c->hX.data(), c->hY.data(),
                               c->bc_left, c->bc_right, c->off);
    if (st != MFREE_OK) return fail(c, st, "TL allocation failed");
  }
  if (c->opt.pipeline == MFREE_PIPELINE_REFERENCE) {
    mfree_status st = ref_alloc(c);
    if (st != MFREE_OK) return st;
  }
  c->committed = true;
  if (!tl && c->opt.precision == MFREE_PRECISION_MIXED) {
    c->cell_scale = c->opt.cell_scale >= 1.0 ? c->opt.cell_scale : 1.0;
    if (c->opt.tile_cells > 0) {
      c->tile_cells = (uint32_t)c->opt.tile_cells;
    } else if (c->opt.tile_cells == 0 && !c->use_list) {
      // RK4 steps of a list context do not use the per-stage pass: its shape
      // (derivative evaluations, learned base steps) stays the untiled one
      mfree_status st = autotune_tiles(c);
      if (st != MFREE_OK) return st;
    }
  }
  return MFREE_OK;
}

BinArgs bin_args(mfree_ctx* c, const double* x, const double* y, int b) {
  BinArgs a;
  a.n = c->N;
  a.nscene = c->S;
  a.nchunks = c->nchunks;
  a.ncell_cap = c->N;
  a.deterministic = c->opt.deterministic;
  // src/body.h:135,138: side 2 h_max; the mixed pair pass may use larger
  // cells (the 3x3 block still holds every pair r < 2h, with more candidates)
  a.kernel_width = 2.0 * c->cell_scale;
  a.cell_min = 0.0;
  a.chunks = c->chunks;
  a.chunk_first = c->chunk_first;
  a.sc = c->sc;
  a.geom = c->geom;
  a.x = x;
  a.y = y;
  a.h = c->h[b];
  a.scene = c->scene[b];
  a.gid = c->gid[b];
  a.part = c->part;
  a.key = c->key;
  a.rank = c->rank;
  a.count = c->count;
  a.tsum = c->tsum;
  a.cell_start = c->cell_start;
  a.perm = c->perm;
  a.perm_tmp = c->perm_tmp;
  a.err = c->errb;
  a.berr = tl_method(c->opt.method) ? nullptr : c->berr;
  a.tile_cells = c->tile_cells;
  a.tile_off = c->tile_cells ? c->tile_off : nullptr;
  return a;
}

UlArgs ul_args(mfree_ctx* c, int b) {
  UlArgs u;
  memset(&u, 0, sizeof(u));
  u.n = c->N;
  u.mixed = c->opt.precision == MFREE_PRECISION_MIXED;
  u.method = c->opt.method;
  u.sc = c->sc;
  u.geom = c->geom;
  u.perm = c->perm;
  u.key = c->key;
  u.cell_start = c->cell_start;
  u.m = c->m[b];
  u.h = c->h[b];
  u.scene = c->scene[b];
  u.gid = c->gid[b];
  u.meta = c->meta;
  u.pos64 = c->pos64;
  u.rec64 = c->rec64;
  u.pos32 = c->pos32;
  u.rec32 = c->rec32;
  u.y = c->y[b];
  u.K = c->K;
  u.acc = c->acc;
  u.max_nbh = u.mixed ? 0 : c->opt.max_nbh;
  u.err = c->errb;
  u.berr = c->berr;
  u.nscene = c->S;
  u.tile_cells = c->tile_cells;
  u.tile_off = c->tile_cells ? c->tile_off : nullptr;
  u.tile_grid = c->tile_grid;
  u.xcd_chunk = c->opt.xcd_chunk ? c->opt.xcd_chunk : kXcdChunkDefault;
  u.fcx = c->fc_active ? c->fcx : nullptr;
  u.fcy = c->fc_active ? c->fcy : nullptr;
  return u;
}


TlShared tl_shared(mfree_ctx* c) {
  TlShared t;
  t.st = c->stream;
  t.precision = c->opt.precision;
  t.S = c->S;
  t.N = c->N;
  t.sc = c->sc;
  t.geom = c->geom;
  t.y = c->y[0];
  t.K = c->K;
  t.acc = c->acc;
  t.m = c->m[0];
  t.h = c->h[0];
  t.scene = c->scene[0];
  t.gid = c->gid[0];
  t.chunks = c->chunks;
  t.chunk_first = c->chunk_first;
  t.nchunks = c->nchunks;
  t.part = c->part;
  t.key = c->key;
  t.rank = c->rank;
  t.count = c->count;
  t.tsum = c->tsum;
  t.cell_start = c->cell_start;
  t.perm = c->perm;
  t.err = c->errb;
  t.h_search = c->opt.tl_h_search;
  t.fcx = c->fc_active ? c->fcx : nullptr;
  t.fcy = c->fc_active ? c->fcy : nullptr;
  return t;
}

hipEvent_t take_event(mfree_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

void pair_launch(mfree_ctx* c, const UlArgs& u, std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs) {
  hipEvent_t a = nullptr, b = nullptr;
  if (evs) {
    a = take_event(c);
    b = take_event(c);
    if (a) hipEventRecord(a, c->stream);
  }
  launch_ul_pairs(c->stream, u);
  if (evs && a && b) {
    hipEventRecord(b, c->stream);
    evs->push_back({a, b});
  }
}

// Project autotuner for the mixed UL pair pass: times one complete stage
// (binning, prepass, pair pass in derivative-evaluation mode) of the actual
// bodies for each candidate cell size (side 2 h_max times 1 or 1.25) and
// tile size, and keeps the fastest.  Results are cached per
// (particle-count bucket, body count) for the process.
mfree_status autotune_tiles(mfree_ctx* c) {
  struct Choice {
    uint64_t key;
    uint32_t tile_cells;
    double cell_scale;
  };
  // shared by the process's contexts, which may commit on different threads
  static std::vector<Choice> cache;
  static std::mutex cache_mu;
  const bool forced = c->opt.cell_scale >= 1.0;  // the cell size is given: tiles only, not cached
  const uint64_t key = ((uint64_t)(64 - __builtin_clzll((uint64_t)c->N | 1)) << 32) | c->S;
  if (!forced) {
    std::lock_guard<std::mutex> lk(cache_mu);
    for (auto& e : cache)
      if (e.key == key) {
        c->tile_cells = e.tile_cells;
        c->cell_scale = e.cell_scale;
        return MFREE_OK;
      }
  }
  // trial buffers, released on every return path
  struct Scratch {
    double* d8 = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (d8) (void)hipFree(d8);
    }
  } sr;
  HIPCHK(c, hipMalloc(&sr.d8, 8 * (size_t)c->N * sizeof(double)));
  HIPCHK(c, hipEventCreate(&sr.e0));
  HIPCHK(c, hipEventCreate(&sr.e1));
  double* const d8 = sr.d8;
  const hipEvent_t e0 = sr.e0, e1 = sr.e1;
  const uint32_t cand[] = {0, 8, 12, 16, 24, 32};
  const double scales[] = {1.0, 1.25};
  float best = 1e30f;
  uint32_t best_tc = 0;
  double best_scale = c->cell_scale;
  const int b = c->cur;
  for (double sc : scales) {
    if (forced) sc = c->opt.cell_scale;
    for (uint32_t tc : cand) {
      c->tile_cells = tc;
      c->cell_scale = sc;
      float tmin = 1e30f;
      for (int rep = 0; rep < 4; rep++) {
        hipEventRecord(e0, c->stream);
        launch_binning(c->stream, bin_args(c, c->y[b].f[0], c->y[b].f[1], b));
        UlArgs u = ul_args(c, b);
        u.src = c->y[b];
        launch_ul_prepass(c->stream, u);
        u.stage = 0;
        u.d8 = d8;
        launch_ul_pairs(c->stream, u);
        hipEventRecord(e1, c->stream);
        HIPCHK(c, hipEventSynchronize(e1));
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) tmin = std::min(tmin, ms);  // first run warms up
      }
      if (tmin < best) {
        best = tmin;
        best_tc = tc;
        best_scale = sc;
      }
    }
    if (forced) break;
  }
  c->tile_cells = best_tc;
  c->cell_scale = best_scale;
  if (!forced) {
    std::lock_guard<std::mutex> lk(cache_mu);
    cache.push_back({key, best_tc, best_scale});
  }
  // the timing passes are no steps: no body has failed
  const std::vector<uint32_t> zero(1 + (size_t)c->S, 0u);
  HIPCHK(c, hipMemcpy(c->errb, zero.data(), zero.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return MFREE_OK;
}

// End of a UL step from storage a to b: in a context that isolates its
// bodies, a body whose error word is set keeps the state it started the step
// with (and stays there: the word is sticky); then the time of every body
// that did not fail advances by its dt.
static void end_of_step(mfree_ctx* c, int a, int b, bool advance_time) {
  hipStream_t st = c->stream;
  const bool iso = c->opt.isolate_bodies != 0;
  if (iso)
    hipLaunchKernelGGL(k_restore_failed, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, c->N,
                       c->scene[b], c->berr, c->perm_keep, c->y[a], c->y[b]);
  if (advance_time)
    hipLaunchKernelGGL(k_advance_time, dim3((c->S + 63) / 64), dim3(64), 0, st, c->S, c->sc,
                       iso ? c->berr : nullptr);
  c->cur = b;
}

// Enqueues one RK step of every UL body (src/rk4.h:71-108 with the UL
// compute_time_derivatives of src/body.h:125-148 per stage).
void enqueue_ul_rk(mfree_ctx* c, const RkPlan& plan, std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs) {
  hipStream_t st = c->stream;
  const int a = c->cur, b = 1 - a;
  const SceneConst* sc = plan.sc ? plan.sc : c->sc;
  // stage 1: k1 = y_n; bin y_n, and move storage into the new cell order
  launch_binning(st, bin_args(c, c->y[a].f[0], c->y[a].f[1], a));
  if (plan.after_bin1) plan.after_bin1(a);
  UlArgs u = ul_args(c, a);
  u.sc = sc;
  u.src = c->y[a];
  u.reorder = 1;
  u.dst = c->y[b];
  u.m_dst = c->m[b];
  u.h_dst = c->h[b];
  u.scene_dst = c->scene[b];
  u.gid_dst = c->gid[b];
  u.perm_keep = c->perm_keep;
  launch_ul_prepass(st, u);
  UlArgs p = ul_args(c, b);
  p.sc = sc;
  p.src = c->y[b];
  p.stage = 1;
  p.home_identity = 1;
  p.k1s = plan.k1s;
  if (plan.stage1) plan.stage1(p);
  pair_launch(c, p, evs);
  static const int rk4[] = {2, 3, 4}, rk3[] = {5, 4};
  const int* stages = plan.order == 3 ? rk3 : rk4;
  const int nst = plan.order == 3 ? 2 : 3;
  for (int k = 0; k < nst; k++) {
    launch_binning(st, bin_args(c, c->K.f[0], c->K.f[1], b));
    UlArgs q = ul_args(c, b);
    q.sc = sc;
    q.src = c->K;
    launch_ul_prepass(st, q);
    q.stage = stages[k];
    q.home_identity = 0;
    q.rate_out = plan.rate_out && stages[k] == 4;
    if (stages[k] == 5 || stages[k] == 2) q.k2s = plan.k2s;
    if (stages[k] == 3) q.k3s = plan.k3s;
    pair_launch(c, q, evs);
  }
  end_of_step(c, a, b, plan.advance_time);
}

// One RK4 step of every UL body on the neighbour-list path (mixed
// precision): one binning with the skin, the stage-1 prepass (storage moves
// into the new cell order, records of stage 1), stage 1 (which also writes
// the step's list), then stages 2-4 from the list; each stage writes the
// records of the next one into the other record buffer.
void enqueue_ul_list_step(mfree_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs) {
  hipStream_t st = c->stream;
  const int a = c->cur, b = 1 - a;
  BinArgs ba = bin_args(c, c->y[a].f[0], c->y[a].f[1], a);
  ba.vx = c->y[a].f[2];
  ba.vy = c->y[a].f[3];
  ba.vpart = c->vpart;
  ba.skin_min = kSkinMin;
  ba.skin_vfac = kSkinVfac;
  ba.tile_cells = 0;
  ba.tile_off = nullptr;
  const bool iso = c->opt.isolate_bodies != 0;
  if (!iso) ba.advance = c->sc;
  launch_binning(st, ba);
  UlArgs u = ul_args(c, a);
  u.src = c->y[a];
  u.reorder = 1;
  u.dst = c->y[b];
  u.m_dst = c->m[b];
  u.h_dst = c->h[b];
  u.scene_dst = c->scene[b];
  u.gid_dst = c->gid[b];
  u.perm_keep = c->perm_keep;
  launch_ul_prepass(st, u);
  UlArgs p = ul_args(c, b);
  p.home_identity = 1;
  p.nl_col = c->nl_col;
  p.nl16 = c->nl16;
  p.nl_deg = c->nl_deg;
  p.nl_cap = c->nl_cap;
  float2* pos[2] = {c->pos32, c->pos32b};
  PairRec32* rec[2] = {c->rec32, c->rec32b};
  for (int k = 1; k <= 4; k++) {
    p.stage = k;
    p.src = k == 1 ? c->y[b] : c->K;
    p.pos32 = pos[(k - 1) & 1];
    p.rec32 = rec[(k - 1) & 1];
    p.pos32_out = k < 4 ? pos[k & 1] : nullptr;
    p.rec32_out = k < 4 ? rec[k & 1] : nullptr;
    // the profile times the list-stage kernel (stages 2-4), the dominant one
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (evs && k > 1) {
      e0 = take_event(c);
      e1 = take_event(c);
      if (e0) hipEventRecord(e0, st);
    }
    if (k == 1)
      launch_ul_list_first(st, p, c->list_staged && c->list_lanes == 1);  // as the stages that read the list
    else
      launch_ul_list(st, p, c->list_lanes, c->list_staged);
    if (evs && e0 && e1) {
      hipEventRecord(e1, st);
      evs->push_back({e0, e1});
    }
  }
  end_of_step(c, a, b, iso);
}

void enqueue_ul_step(mfree_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs) {
  if (c->use_list)
    enqueue_ul_list_step(c, evs);
  else
    enqueue_ul_rk(c, RkPlan(), evs);
}

mfree_status run_steps(mfree_ctx* c, uint32_t nsteps) {
  const bool tl = tl_method(c->opt.method);
  if (tl) {
    mfree_status st = tl_ensure_precomputed(c->tl, tl_shared(c));
    if (st != MFREE_OK) return fail(c, st, "TL precompute failed: " + c->tl.msg);
  }
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->prof) {
    c->ev_used = 0;
    e0 = take_event(c);
    e1 = take_event(c);
    HIPCHK(c, hipEventRecord(e0, c->stream));
  }
  for (uint32_t k = 0; k < nsteps; k++) {
    if (c->ref) {
      ref_enqueue_step(c);
      continue;
    }
    if (tl) {
      tl_enqueue_step(c->tl, tl_shared(c), c->prof ? &evs : nullptr, [c]() { return take_event(c); });
      hipLaunchKernelGGL(k_advance_time, dim3((c->S + 63) / 64), dim3(64), 0, c->stream, c->S, c->sc, nullptr);
      continue;
    }
    if (c->opt.use_graph && !c->prof) {
      int a = c->cur;
      if (!c->gexec[a]) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_ul_step(c, nullptr);
        HIPCHK(c, hipStreamEndCapture(c->stream, &g));
        HIPCHK(c, hipGraphInstantiate(&c->gexec[a], g, nullptr, nullptr, 0));
        HIPCHK(c, hipGraphDestroy(g));
        c->cur = a;  // capture did not execute anything
      }
      HIPCHK(c, hipGraphLaunch(c->gexec[a], c->stream));
      c->cur = 1 - a;
    } else {
      enqueue_ul_step(c, c->prof ? &evs : nullptr);
    }
  }
  HIPCHK(c, hipGetLastError());
  if (c->prof) {
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    c->step_ms += ms;
    c->steps += nsteps;
    for (auto& pr : evs) {
      float t = 0.f;
      HIPCHK(c, hipEventElapsedTime(&t, pr.first, pr.second));
      c->pair_ms += t;
      c->pair_n++;
    }
  }
  if (!c->opt.isolate_bodies) {
    for (uint32_t s = 0; s < c->S; s++)
      for (uint32_t k = 0; k < nsteps; k++) c->time[s] = c->time[s] + c->prm[s].dt;
    return check_err_bits(c);
  }
  // isolated bodies: a failed body's time stopped at its last good step, on
  // the device (the same additions of dt as above for the others)
  mfree_status st = check_err_bits(c);
  if (st != MFREE_OK) return st;
  std::vector<SceneConst> hsc(c->S);
  HIPCHK(c, hipMemcpy(hsc.data(), c->sc, c->S * sizeof(SceneConst), hipMemcpyDeviceToHost));
  for (uint32_t s = 0; s < c->S; s++) c->time[s] = hsc[s].time;
  return MFREE_OK;
}

mfree_status body_status(const mfree_ctx* c, uint32_t body) {
  if (!c->committed || body >= c->body_bits.size() || !c->body_bits[body]) return MFREE_OK;
  const char* msg = nullptr;
  return bits_status(c->body_bits[body], &msg);
}

bool valid_particles(const mfree_particles* p, bool tl) {
  if (!p) return false;
  if (!p->x || !p->y || !p->m || !p->rho || !p->h || !p->vx || !p->vy || !p->Sxx || !p->Sxy || !p->Syy)
    return false;
  if (tl && (!p->X || !p->Y)) return false;
  return true;
}

// Storage order -> idx order of one body: dst[gid - o] = src[s] over the
// body's storage range [o, o + n).
__global__ __launch_bounds__(256) void k_to_idx_order(uint32_t n, uint32_t o, const uint32_t* __restrict__ gid,
                                                      const double* __restrict__ src, double* __restrict__ dst) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[gid[o + s]] = src[o + s];
}

void launch_to_idx_order(hipStream_t st, uint32_t n, const uint32_t* gid, const double* src, double* dst) {
  if (n) hipLaunchKernelGGL(k_to_idx_order, dim3((n + 255) / 256), dim3(256), 0, st, n, 0u, gid, src, dst);
}

}  // namespace mfree


extern "C" {

int32_t mfree_abi_version(void) { return MFREE_ABI_VERSION; }

const char* mfree_status_string(mfree_status s) {
  switch (s) {
    case MFREE_OK: return "ok";
    case MFREE_ERR_ARG: return "bad argument";
    case MFREE_ERR_HIP: return "HIP error or no gfx950 device";
    case MFREE_ERR_NBH_OVERFLOW: return "neighbour overflow (MAX_NBH)";
    case MFREE_ERR_NO_NEIGHBOR: return "particle without neighbour";
    case MFREE_ERR_NONFINITE: return "non-finite value";
    case MFREE_ERR_CELLS: return "cell grid overflow (num_cell >= n)";
    case MFREE_ERR_PIPELINE: return "unrecognised action pipeline";
    case MFREE_ERR_CAPACITY: return "buffer too small";
    case MFREE_ERR_UNSUPPORTED: return "unsupported";
    case MFREE_ERR_SCRATCH: return "input exceeds the fixed scratch of an allocation-free entry point";
    case MFREE_ERR_SKIN: return "neighbour-list skin exceeded within a step";
    default: return "unknown status";
  }
}

void mfree_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]) {
  // src/kernel.cpp:20-55
  double h1 = 1. / h;
  double xij = xi - xj;
  double yij = yi - yj;
  double rij = sqrt(xij * xij + yij * yij);
  double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
  double q = rij * h1;
  out[0] = out[1] = out[2] = 0.;
  if (q >= 2.) return;
  if (q >= 1.) {
    out[0] = fac * (0.25 * (2 - q) * (2 - q) * (2 - q));
    if (rij > 1e-12) {
      double der = -0.75 * (2 - q) * (2 - q) * h1 / rij;
      out[1] = xij * der * fac;
      out[2] = yij * der * fac;
    }
    return;
  }
  out[0] = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
  if (rij > 1e-12) {
    double der = -3.0 * q * (1 - 0.75 * q) * h1 / rij;
    out[1] = xij * der * fac;
    out[2] = yij * der * fac;
  }
}

int32_t mfree_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void mfree_default_options(mfree_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->method = MFREE_METHOD_UL_MONAGHAN;
  o->precision = MFREE_PRECISION_FP64;
  o->device = 0;
  o->max_nbh = 200;
  o->use_graph = 1;
  o->deterministic = 1;
  o->tl_h_search = 0.;
  o->tile_cells = 0;
  o->pipeline = MFREE_PIPELINE_FUSED;
  o->xcd_chunk = 0;
  o->cell_scale = 0.;
  o->neighbor_list = 0;
  o->list_lanes = 0;
  o->list_staging = 0;
  o->isolate_bodies = 0;
}

mfree_status mfree_create(const mfree_options* opt, mfree_ctx** out) {
  if (!out) return MFREE_ERR_ARG;
  *out = nullptr;
  mfree_ctx* c = new mfree_ctx();
  if (opt)
    c->opt = *opt;
  else
    mfree_default_options(&c->opt);
  if (c->opt.method != MFREE_METHOD_UL_MONAGHAN && c->opt.method != MFREE_METHOD_UL_GODUNOV &&
      !tl_method(c->opt.method)) {
    delete c;
    return MFREE_ERR_ARG;
  }
  // the co-rotated and potential TL lists exist in FP64
  if ((c->opt.method == MFREE_METHOD_TL_COROTATIONAL || c->opt.method == MFREE_METHOD_TL_POTENTIAL) &&
      c->opt.precision != MFREE_PRECISION_FP64) {
    delete c;
    return MFREE_ERR_UNSUPPORTED;
  }
  // the Godunov pair sums exist in FP64 on the fused pipeline
  if (c->opt.method == MFREE_METHOD_UL_GODUNOV &&
      (c->opt.precision != MFREE_PRECISION_FP64 || c->opt.pipeline != MFREE_PIPELINE_FUSED)) {
    delete c;
    return MFREE_ERR_UNSUPPORTED;
  }
  if (c->opt.precision != MFREE_PRECISION_FP64 && c->opt.precision != MFREE_PRECISION_MIXED) {
    delete c;
    return MFREE_ERR_ARG;
  }
  if (c->opt.pipeline != MFREE_PIPELINE_FUSED && c->opt.pipeline != MFREE_PIPELINE_REFERENCE) {
    delete c;
    return MFREE_ERR_ARG;
  }
  if (c->opt.neighbor_list < -1 || c->opt.neighbor_list > 1 || c->opt.list_staging < -1 || c->opt.list_staging > 1 ||
      !(c->opt.list_lanes == 0 || c->opt.list_lanes == 1 || c->opt.list_lanes == 2 || c->opt.list_lanes == 4)) {
    delete c;
    return MFREE_ERR_ARG;
  }
  if (c->opt.neighbor_list == 1 &&
      (c->opt.method != MFREE_METHOD_UL_MONAGHAN || c->opt.precision != MFREE_PRECISION_MIXED ||
       c->opt.pipeline != MFREE_PIPELINE_FUSED)) {
    delete c;
    return MFREE_ERR_UNSUPPORTED;
  }
  if (c->opt.pipeline == MFREE_PIPELINE_REFERENCE &&
      (c->opt.method != MFREE_METHOD_UL_MONAGHAN || c->opt.precision != MFREE_PRECISION_FP64 ||
       c->opt.max_nbh > 200)) {
    delete c;
    return MFREE_ERR_UNSUPPORTED;
  }
  if (c->opt.isolate_bodies != 0 && c->opt.isolate_bodies != 1) {
    delete c;
    return MFREE_ERR_ARG;
  }
  if (c->opt.isolate_bodies && (tl_method(c->opt.method) || c->opt.pipeline != MFREE_PIPELINE_FUSED)) {
    delete c;
    return MFREE_ERR_UNSUPPORTED;
  }
  *out = c;
  return MFREE_OK;
}

void mfree_destroy(mfree_ctx* c) {
  if (!c) return;
  if (c->committed) hipSetDevice(c->opt.device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; k++)
    if (c->gexec[k]) hipGraphExecDestroy(c->gexec[k]);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  vtk_free(c);
  tl_free(c->tl);
  ref_free(c);
  for (void* p : c->allocs) hipFree(p);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* mfree_last_error(const mfree_ctx* c) { return c ? c->err.c_str() : "null context"; }

mfree_status mfree_add_body(mfree_ctx* c, const mfree_params* p, const mfree_particles* host, uint32_t n,
                            uint32_t* body_id) {
  if (!c || !p) return MFREE_ERR_ARG;
  if (c->committed) return fail(c, MFREE_ERR_ARG, "bodies must be added before the first compute call");
  const bool tl = tl_method(c->opt.method);
  if (!valid_particles(host, tl)) return fail(c, MFREE_ERR_ARG, "missing particle arrays");
  if (n < 2) return fail(c, MFREE_ERR_ARG, "a body needs at least 2 particles");
  // the mixed pair kernel packs sorted positions into 28 bits
  if ((uint64_t)c->off.back() + n >= (1ull << 28)) return fail(c, MFREE_ERR_ARG, "too many particles (max 2^28)");
  if (!(p->dt > 0) || !(p->rho0 > 0)) return fail(c, MFREE_ERR_ARG, "dt and rho0 must be positive");
  bool fc = false;
  for (uint32_t i = 0; i < n; i++) {
    fc = fc || (host->fcx && host->fcx[i] != 0.) || (host->fcy && host->fcy[i] != 0.);
    if (!(host->h[i] > 0) || !(host->rho[i] > 0) || !isfinite(host->x[i]) || !isfinite(host->y[i]))
      return fail(c, MFREE_ERR_ARG, "h and rho must be positive, positions finite");
  }
  // contact forces given with the body (src/particle.h:57): kept for commit
  if (fc && c->opt.pipeline == MFREE_PIPELINE_REFERENCE)
    return fail(c, MFREE_ERR_UNSUPPORTED, "contact forces on the reference-structured pipeline");
  const size_t before = c->off.back();
  if (fc && c->hfcx.empty()) {
    c->hfcx.assign(before, 0.);
    c->hfcy.assign(before, 0.);
  }
  if (!c->hfcx.empty())
    for (uint32_t i = 0; i < n; i++) {
      c->hfcx.push_back(host->fcx ? host->fcx[i] : 0.);
      c->hfcy.push_back(host->fcy ? host->fcy[i] : 0.);
    }
  auto app = [n](std::vector<double>& v, const double* s) { v.insert(v.end(), s, s + n); };
  app(c->hx, host->x);
  app(c->hy, host->y);
  app(c->hX, tl ? host->X : host->x);
  app(c->hY, tl ? host->Y : host->y);
  app(c->hm, host->m);
  app(c->hrho, host->rho);
  app(c->hh, host->h);
  app(c->hvx, host->vx);
  app(c->hvy, host->vy);
  app(c->hSxx, host->Sxx);
  app(c->hSxy, host->Sxy);
  app(c->hSyy, host->Syy);
  if (body_id) *body_id = (uint32_t)c->prm.size();
  c->prm.push_back(*p);
  c->time.push_back(p->time);
  c->off.push_back(c->off.back() + n);
  c->bc_left.emplace_back();
  c->bc_right.emplace_back();
  c->v_bc.push_back(0.);
  return MFREE_OK;
}

mfree_status mfree_set_strip_bc(mfree_ctx* c, uint32_t body, const uint32_t* idx_left, uint32_t nl,
                                const uint32_t* idx_right, uint32_t nr, double v_bc) {
  if (!c || body >= c->prm.size()) return MFREE_ERR_ARG;
  if (!tl_method(c->opt.method)) return fail(c, MFREE_ERR_UNSUPPORTED, "strip BCs are TL only");
  if (c->committed) return fail(c, MFREE_ERR_ARG, "BCs must be set before the first compute call");
  uint32_t n = c->off[body + 1] - c->off[body];
  for (uint32_t i = 0; i < nl; i++)
    if (idx_left[i] >= n) return fail(c, MFREE_ERR_ARG, "BC index out of range");
  for (uint32_t i = 0; i < nr; i++)
    if (idx_right[i] >= n) return fail(c, MFREE_ERR_ARG, "BC index out of range");
  c->bc_left[body].assign(idx_left, idx_left + nl);
  c->bc_right[body].assign(idx_right, idx_right + nr);
  c->v_bc[body] = v_bc;
  return MFREE_OK;
}

mfree_status mfree_tl_precompute(mfree_ctx* c) {
  if (!c) return MFREE_ERR_ARG;
  if (!tl_method(c->opt.method)) return fail(c, MFREE_ERR_UNSUPPORTED, "not a TL context");
  mfree_status st = commit(c);
  if (st) return st;
  st = tl_ensure_precomputed(c->tl, tl_shared(c));
  if (st != MFREE_OK) return fail(c, st, "TL precompute failed: " + c->tl.msg);
  return check_err_bits(c);
}

mfree_status mfree_set_state(mfree_ctx* c, uint32_t body, const mfree_particles* host) {
  if (!c || body >= c->prm.size()) return MFREE_ERR_ARG;
  const bool tl = tl_method(c->opt.method);
  if (!valid_particles(host, false)) return fail(c, MFREE_ERR_ARG, "missing particle arrays");
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t o = c->off[body], n = c->off[body + 1] - o;
  c->state_version++;
  if (c->opt.isolate_bodies && c->body_bits[body]) {
    // a new state thaws a failed body
    const uint32_t zero = 0;
    HIPCHK(c, hipMemcpy(c->berr + body, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    c->body_bits[body] = 0;
  }
  // the body's storage range is [o, o+n) in any order; rewrite it in idx order
  const double* src[8] = {host->x, host->y, host->vx, host->vy, host->rho, host->Sxx, host->Sxy, host->Syy};
  const int b = c->cur;
  for (int k = 0; k < 8; k++)
    HIPCHK(c, hipMemcpy(c->y[b].f[k] + o, src[k], n * sizeof(double), hipMemcpyHostToDevice));
  if (host->fcx || host->fcy) {
    bool nz = false;
    for (uint32_t i = 0; i < n; i++) nz = nz || (host->fcx && host->fcx[i] != 0.) || (host->fcy && host->fcy[i] != 0.);
    if (nz && (st = contact_activate(c))) return st;
    if (host->fcx) HIPCHK(c, hipMemcpy(c->fcx + o, host->fcx, n * sizeof(double), hipMemcpyHostToDevice));
    if (host->fcy) HIPCHK(c, hipMemcpy(c->fcy + o, host->fcy, n * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!tl) {
    HIPCHK(c, hipMemcpy(c->m[b] + o, host->m, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->h[b] + o, host->h, n * sizeof(double), hipMemcpyHostToDevice));
    std::vector<uint32_t> gid(n);
    for (uint32_t i = 0; i < n; i++) gid[i] = o + i;
    HIPCHK(c, hipMemcpy(c->gid[b] + o, gid.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  return MFREE_OK;
}

mfree_status mfree_get_state(mfree_ctx* c, uint32_t body, mfree_particles* host) {
  if (!c || body >= c->prm.size() || !host) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  const uint32_t o = c->off[body], n = c->off[body + 1] - o;
  const int b = c->cur;
  // back to idx order on the device (into the stage buffers K and acc, free
  // between steps), then one copy per field
  double* dst[10] = {host->x, host->y, host->vx, host->vy, host->rho, host->Sxx, host->Sxy, host->Syy,
                     host->m, host->h};
  const double* src[10];
  double* tmp[10];
  for (int k = 0; k < 8; k++) {
    src[k] = c->y[b].f[k];
    tmp[k] = c->K.f[k];
  }
  src[8] = c->m[b];
  src[9] = c->h[b];
  tmp[8] = c->acc.f[0];
  tmp[9] = c->acc.f[1];
  for (int k = 0; k < 10; k++)
    if (dst[k] && n)
      hipLaunchKernelGGL(k_to_idx_order, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, o, c->gid[b], src[k],
                         tmp[k]);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 10; k++)
    if (dst[k]) HIPCHK(c, hipMemcpy(dst[k], tmp[k] + o, n * sizeof(double), hipMemcpyDeviceToHost));
  // contact forces are kept by global idx
  if (host->fcx) HIPCHK(c, hipMemcpy(host->fcx, c->fcx + o, n * sizeof(double), hipMemcpyDeviceToHost));
  if (host->fcy) HIPCHK(c, hipMemcpy(host->fcy, c->fcy + o, n * sizeof(double), hipMemcpyDeviceToHost));
  if (host->X)
    for (uint32_t i = 0; i < n; i++) host->X[i] = c->hX[o + i];
  if (host->Y)
    for (uint32_t i = 0; i < n; i++) host->Y[i] = c->hY[o + i];
  return MFREE_OK;
}

mfree_status mfree_get_time(mfree_ctx* c, uint32_t body, double* t) {
  if (!c || body >= c->prm.size() || !t) return MFREE_ERR_ARG;
  *t = c->time[body];
  return MFREE_OK;
}

mfree_status mfree_body_status(mfree_ctx* c, uint32_t body, mfree_status* status) {
  if (!c || body >= c->prm.size() || !status) return MFREE_ERR_ARG;
  *status = body_status(c, body);
  return MFREE_OK;
}

mfree_status mfree_step(mfree_ctx* c, uint32_t nsteps) {
  if (!c) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  return run_steps(c, nsteps);
}

mfree_status mfree_eval_derivatives(mfree_ctx* c, double* d8, double* aux) {
  if (!c || !d8) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  const size_t N = c->N;
  double *dd = nullptr, *da = nullptr;
  HIPCHK(c, hipMalloc(&dd, 8 * N * sizeof(double)));
  if (aux && hipMalloc(&da, 16 * N * sizeof(double)) != hipSuccess) {
    hipFree(dd);
    return fail(c, MFREE_ERR_HIP, "hipMalloc failed");
  }
  if (tl_method(c->opt.method)) {
    st = tl_ensure_precomputed(c->tl, tl_shared(c));
    if (st == MFREE_OK)
      tl_enqueue_eval(c->tl, tl_shared(c), dd, da);
  } else if (c->ref) {
    ref_enqueue_eval(c, dd);
  } else {
    const int b = c->cur;
    launch_binning(c->stream, bin_args(c, c->y[b].f[0], c->y[b].f[1], b));
    UlArgs u = ul_args(c, b);
    u.src = c->y[b];
    launch_ul_prepass(c->stream, u);
    u.stage = 0;
    u.d8 = dd;
    u.aux = da;
    launch_ul_pairs(c->stream, u);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && st == MFREE_OK) e = hipMemcpy(d8, dd, 8 * N * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && st == MFREE_OK && aux) e = hipMemcpy(aux, da, 16 * N * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(dd);
  if (da) hipFree(da);
  if (st != MFREE_OK) return fail(c, st, "TL precompute failed: " + c->tl.msg);
  if (e != hipSuccess) return fail(c, MFREE_ERR_HIP, hipGetErrorString(e));
  return check_err_bits(c);
}

mfree_status mfree_get_pairs(mfree_ctx* c, uint32_t body, uint32_t* row_ptr, uint32_t* col, uint64_t cap,
                             uint64_t* npairs) {
  if (!c || body >= c->prm.size() || !row_ptr) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  const uint32_t o = c->off[body], n = c->off[body + 1] - o;
  if (tl_method(c->opt.method)) {
    st = tl_ensure_precomputed(c->tl, tl_shared(c));
    if (st != MFREE_OK) return fail(c, st, "TL precompute failed: " + c->tl.msg);
    uint64_t tot = 0;
    st = tl_get_pairs(c->tl, tl_shared(c), body, o, n, row_ptr, col, cap, &tot);
    if (npairs) *npairs = tot;
    if (st == MFREE_ERR_CAPACITY) return fail(c, st, "col buffer too small");
    return st;
  }
  const int b = c->cur;
  BinArgs ba = bin_args(c, c->y[b].f[0], c->y[b].f[1], b);
  ba.kernel_width = 2.0;  // the reference's cells: the lists are its Verlet lists
  ba.tile_cells = 0;
  ba.tile_off = nullptr;
  launch_binning(c->stream, ba);
  uint32_t *dcnt = nullptr, *drow = nullptr, *dcol = nullptr;
  HIPCHK(c, hipMalloc(&dcnt, (size_t)n * sizeof(uint32_t)));
  PairListArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.n = n;
  pa.first = o;
  pa.sc = c->sc;
  pa.geom = c->geom;
  pa.scene = body;
  pa.cell_start = c->cell_start;
  pa.perm = c->perm;
  pa.key = c->key;
  pa.gid = c->gid[b];
  pa.x = c->y[b].f[0];
  pa.y = c->y[b].f[1];
  pa.h = c->h[b];
  pa.row_count = dcnt;
  pa.fill = 0;
  launch_pair_list(c->stream, pa);
  std::vector<uint32_t> cnt(n);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(cnt.data(), dcnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    hipFree(dcnt);
    return fail(c, MFREE_ERR_HIP, hipGetErrorString(e));
  }
  uint64_t tot = 0;
  row_ptr[0] = 0;
  for (uint32_t i = 0; i < n; i++) {
    tot += cnt[i];
    row_ptr[i + 1] = (uint32_t)tot;
  }
  if (npairs) *npairs = tot;
  st = check_err_bits(c);
  if (st != MFREE_OK) {
    hipFree(dcnt);
    return st;
  }
  if (tot > cap || !col) {
    hipFree(dcnt);
    return fail(c, MFREE_ERR_CAPACITY, "col buffer too small");
  }
  e = hipMalloc(&drow, ((size_t)n + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&dcol, std::max<uint64_t>(tot, 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(drow, row_ptr, ((size_t)n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    pa.row_ptr = drow;
    pa.col = dcol;
    pa.fill = 1;
    launch_pair_list(c->stream, pa);
    e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipMemcpy(col, dcol, tot * sizeof(uint32_t), hipMemcpyDeviceToHost);
  hipFree(dcnt);
  if (drow) hipFree(drow);
  if (dcol) hipFree(dcol);
  if (e != hipSuccess) return fail(c, MFREE_ERR_HIP, hipGetErrorString(e));
  return MFREE_OK;
}

mfree_status mfree_profile_enable(mfree_ctx* c, int32_t on) {
  if (!c) return MFREE_ERR_ARG;
  c->prof = on != 0;
  c->pair_ms = c->step_ms = 0.;
  c->pair_n = c->steps = 0;
  return MFREE_OK;
}

mfree_status mfree_profile_read(mfree_ctx* c, double* pair_ms_total, uint64_t* pair_launches, double* step_ms_total,
                                uint64_t* steps) {
  if (!c) return MFREE_ERR_ARG;
  if (pair_ms_total) *pair_ms_total = c->pair_ms;
  if (pair_launches) *pair_launches = c->pair_n;
  if (step_ms_total) *step_ms_total = c->step_ms;
  if (steps) *steps = c->steps;
  return MFREE_OK;
}

mfree_status mfree_calibrate_copy(int32_t device, uint64_t bytes, uint32_t reps, double* gbps) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MFREE_ERR_HIP;
  if (bytes < 8 || reps == 0) return MFREE_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MFREE_ERR_HIP;
  const uint64_t n = bytes / 8;
  double *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  mfree_status rc = MFREE_ERR_HIP;
  if (hipMalloc(&a, n * 8) == hipSuccess && hipMalloc(&b, n * 8) == hipSuccess &&
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess && hipMemsetAsync(a, 0, n * 8, st) == hipSuccess) {
    launch_calib_copy(st, a, b, n);  // warm
    hipEventRecord(e0, st);
    for (uint32_t r = 0; r < reps; r++) launch_calib_copy(st, a, b, n);
    hipEventRecord(e1, st);
    float ms = 0.f;
    if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0) {
      if (gbps) *gbps = 2.0 * (double)(n * 8) * reps / (ms * 1e-3) / 1e9;
      rc = MFREE_OK;
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

int32_t mfree_tile_cells(const mfree_ctx* c) { return c ? (int32_t)c->tile_cells : 0; }

double mfree_cell_scale(const mfree_ctx* c) { return c ? c->cell_scale : 0.0; }

uint64_t mfree_num_particles(const mfree_ctx* c) { return c ? c->off.back() : 0; }
uint32_t mfree_num_bodies(const mfree_ctx* c) { return c ? (uint32_t)c->prm.size() : 0; }

// ---------------------------------------------------------------------------
// scenes

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double unit_open(uint64_t& s) { return ((double)(splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

uint32_t mfree_scene_rings_count(const mfree_rings_spec* sp) {
  if (!sp || sp->nbox < 2 || !(sp->ro > 0) || sp->ri < 0 || sp->ri >= sp->ro) return 0;
  const double dx = 2 * sp->ro / (sp->nbox - 1);
  uint64_t n = 0;
  for (uint32_t i = 0; i < sp->nbox; i++)
    for (uint32_t j = 0; j < sp->nbox; j++) {
      double px = -sp->ro + i * dx, py = -sp->ro + j * dx;
      double dist = sqrt(px * px + py * py);
      if (dist < sp->ro && dist >= sp->ri) n += 2;
    }
  return n < (1ull << 31) ? (uint32_t)n : 0;
}

mfree_status mfree_scene_rings(const mfree_rings_spec* sp, mfree_particles* out, uint32_t cap, uint32_t* n_out) {
  uint32_t n = mfree_scene_rings_count(sp);
  if (n == 0 || !out) return MFREE_ERR_ARG;
  if (n > cap) return MFREE_ERR_CAPACITY;
  if (!out->x || !out->y || !out->m || !out->rho || !out->h || !out->vx || !out->vy) return MFREE_ERR_ARG;
  // src/benchmarks/rings.cpp:40-66
  const double dx = 2 * sp->ro / (sp->nbox - 1);
  std::vector<double> x(n), y(n);
  uint32_t k = 0;
  for (uint32_t i = 0; i < sp->nbox; i++)
    for (uint32_t j = 0; j < sp->nbox; j++) {
      double px = -sp->ro + i * dx, py = -sp->ro + j * dx;
      double dist = sqrt(px * px + py * py);
      if (dist < sp->ro && dist >= sp->ri) {
        x[k] = px - sp->centre;
        y[k] = py;
        k++;
        x[k] = px + sp->centre;
        y[k] = py;
        k++;
      }
    }
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  uint64_t state = sp->seed;
  if (sp->shuffle)
    for (uint32_t i = n - 1; i > 0; i--) {
      uint32_t j = (uint32_t)(splitmix64(state) % ((uint64_t)i + 1));
      std::swap(order[i], order[j]);
    }
  for (uint32_t i = 0; i < n; i++) {
    uint32_t s = order[i];
    out->x[i] = x[s];
    out->y[i] = y[s];
    if (out->X) out->X[i] = x[s];
    if (out->Y) out->Y[i] = y[s];
    out->rho[i] = sp->rho0;
    out->h[i] = sp->hdx * dx;
    out->m[i] = dx * dx * sp->rho0;
    out->vx[i] = (out->x[i] < 0.) ? sp->v0 : -sp->v0;
    out->vy[i] = 0.;
    if (out->Sxx) out->Sxx[i] = 0.;
    if (out->Sxy) out->Sxy[i] = 0.;
    if (out->Syy) out->Syy[i] = 0.;
    if (out->fcx) out->fcx[i] = 0.;
    if (out->fcy) out->fcy[i] = 0.;
  }
  if (sp->jitter > 0)
    for (uint32_t i = 0; i < n; i++) {
      double u1 = unit_open(state), u2 = unit_open(state);
      double r = sqrt(-2.0 * log(u1));
      out->vx[i] += sp->jitter * r * cos(2.0 * M_PI * u2);
      out->vy[i] += sp->jitter * r * sin(2.0 * M_PI * u2);
    }
  if (n_out) *n_out = n;
  return MFREE_OK;
}

uint32_t mfree_scene_plate_count(const mfree_plate_spec* sp, uint32_t* n_left, uint32_t* n_right) {
  if (!sp || !(sp->dx > 0) || !(sp->L > 0) || !(sp->W > 0)) return 0;
  const uint32_t nx = (uint32_t)llround(sp->L / sp->dx) + 1, ny = (uint32_t)llround(sp->W / sp->dx) + 1;
  const double cx = 0.5 * sp->L, cy = 0.5 * sp->W;
  uint64_t n = 0, nl = 0, nr = 0;
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++) {
      double px = i * sp->dx, py = j * sp->dx;
      double dist = sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy));
      if (dist < sp->r_hole) continue;
      n++;
      if (i < sp->strip) nl++;
      if (i + sp->strip >= nx) nr++;
    }
  if (n_left) *n_left = (uint32_t)nl;
  if (n_right) *n_right = (uint32_t)nr;
  return n < (1ull << 31) ? (uint32_t)n : 0;
}

mfree_status mfree_scene_plate(const mfree_plate_spec* sp, mfree_particles* out, uint32_t cap, uint32_t* idx_left,
                               uint32_t* idx_right, uint32_t* n_out) {
  uint32_t nl = 0, nr = 0;
  uint32_t n = mfree_scene_plate_count(sp, &nl, &nr);
  if (n == 0 || !out) return MFREE_ERR_ARG;
  if (n > cap) return MFREE_ERR_CAPACITY;
  if (!out->x || !out->y || !out->m || !out->rho || !out->h || !out->vx || !out->vy) return MFREE_ERR_ARG;
  const uint32_t nx = (uint32_t)llround(sp->L / sp->dx) + 1, ny = (uint32_t)llround(sp->W / sp->dx) + 1;
  const double cx = 0.5 * sp->L, cy = 0.5 * sp->W;
  uint32_t k = 0, a = 0, b = 0;
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++) {
      double px = i * sp->dx, py = j * sp->dx;
      double dist = sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy));
      if (dist < sp->r_hole) continue;
      out->x[k] = px;
      out->y[k] = py;
      if (out->X) out->X[k] = px;
      if (out->Y) out->Y[k] = py;
      out->rho[k] = sp->rho0;
      out->h[k] = sp->hdx * sp->dx;
      out->m[k] = sp->dx * sp->dx * sp->rho0;
      out->vx[k] = 0.;
      out->vy[k] = 0.;
      if (out->Sxx) out->Sxx[k] = 0.;
      if (out->Sxy) out->Sxy[k] = 0.;
      if (out->Syy) out->Syy[k] = 0.;
      if (out->fcx) out->fcx[k] = 0.;
      if (out->fcy) out->fcy[k] = 0.;
      if (i < sp->strip && idx_left) idx_left[a] = k;
      if (i < sp->strip) a++;
      if (i + sp->strip >= nx && idx_right) idx_right[b] = k;
      if (i + sp->strip >= nx) b++;
      k++;
    }
  if (n_out) *n_out = n;
  return MFREE_OK;
}

}  // extern "C"
