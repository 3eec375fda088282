// Mixed-precision Gray-Monaghan pair pass on gfx950.
//
// Same neighbour sums as k_ul_pairs64 (src/derivatives.h:32-66,378-443,
// src/correctors.h:83-169), evaluated in FP32 from the 64-byte records the
// prepass writes:
//  - positions are cell-local (x minus the particle's own cell corner,
//    subtracted in FP64), so x_ij keeps ~1e-11 m resolution instead of the
//    ~1e-8 m of absolute FP32 coordinates;
//  - every FP32 pair term is accumulated into an FP64 sum;
//  - the per-particle epilogue (Jaumann rate, continuity, momentum,
//    advection) and the RK4 state update stay FP64.
//
// Two kernels:
//  k_ul_tiles32  a workgroup takes a tile of `tile_cells` consecutive cells
//                of one cell column, stages every candidate of the three
//                neighbouring columns (positions and records) in LDS once,
//                and evaluates its particles from LDS (scan -> compacted
//                list -> pair terms).  Tiles of all bodies are enumerated on
//                the device; a fixed grid walks them in XCD-contiguous
//                ranges.  A tile whose candidates exceed the LDS capacity is
//                evaluated from global memory instead.
//  k_ul_pairs32  the same per particle straight from global memory, with the
//                candidates inside the cutoff compacted in LDS first.
// The autotuner times both (and several tile sizes) on the actual bodies.
// The neighbour-list path (one list per RK4 step) has its own kernels below.
// Compiled with -ffp-contract=on: a multiply and an add of one expression
// may fuse, never across statements, so every kernel variant of a pass (LDS
// staged or not) computes the same bits.
#include "kernels.h"

namespace mfree {

namespace {

struct Acc {
  double s = 0.;
};

// Each FP32 pair term is added to an FP64 sum: exact conversion, one
// rounding at 2^-53 per term.  On gfx950 an FP64 add issues at the FP32
// rate, so this costs two instructions per term where a compensated FP32
// (Kahan) sum costs four, and it is the more accurate of the two.
__device__ inline void acc_add(Acc& a, float x) { a.s += (double)x; }

__device__ inline double acc_val(const Acc& a) { return a.s; }

struct Sums {
  Acc Sxx_x, Sxy_y, Sxy_x, Syy_y;
  Acc vx_x, vx_y, vy_x, vy_y;
  Acc ax, ay, xs, ys;
};

// What a neighbour contributes (sigma pre-multiplied by 1/rho^2).
struct Nb {
  float vx, vy, rho, qw, m, h, c, sxx, sxy, syy, Rxx, Rxy, Ryy;
};

// Constants of the particle whose sums are accumulated.
struct Me {
  float vx, vy, rho, m, h, c, sxx, sxy, syy, Rxx, Rxy, Ryy;
  float h1, fac, cut2, inv_wdp, corr_exp, alpha, beta, eta2, xeps_mi;
  bool use_as, pow4;
};

__device__ inline Me make_me(const SceneConst* __restrict__ c, const Nb& r) {
  Me me;
  me.vx = r.vx;
  me.vy = r.vy;
  me.rho = r.rho;
  me.m = r.m;
  me.h = r.h;
  me.c = r.c;
  me.sxx = r.sxx;
  me.sxy = r.sxy;
  me.syy = r.syy;
  me.Rxx = r.Rxx;
  me.Rxy = r.Rxy;
  me.Ryy = r.Ryy;
  me.h1 = 1.f / r.h;
  me.fac = (float)(10 * (M_1_PI) / 7.0) * me.h1 * me.h1;
  me.cut2 = 4.0001f * (r.h * r.h);
  me.inv_wdp = (float)(1.0 / c->wdeltap);
  me.corr_exp = (float)c->corr_exp;
  me.use_as = c->wdeltap > 0;
  me.pow4 = c->pow4 != 0;
  me.alpha = (float)c->alpha;
  me.beta = (float)c->beta;
  me.eta2 = (float)(c->eta * c->eta);
  me.xeps_mi = (float)c->xsph_eps * r.m;
  return me;
}

__device__ inline Nb nb_from_rec(const PairRec32& r) {
  Nb n;
  n.vx = r.vx;
  n.vy = r.vy;
  n.rho = r.rho;
  n.qw = r.qw;
  n.m = r.m;
  n.h = r.h;
  n.c = r.c;
  n.sxx = r.sxx_r;
  n.sxy = r.sxy_r;
  n.syy = r.syy_r;
  n.Rxx = r.Rxx;
  n.Rxy = r.Rxy;
  n.Ryy = r.Ryy;
  return n;
}

// All four pair sums of one neighbour at separation (xij, yij), r2 = |x_ij|^2.
// Square roots and reciprocals are the hardware's single instructions
// (v_sqrt_f32, v_rcp_f32: about 1 ulp) rather than the correctly rounded
// sequences; each term then carries about one FP32 rounding more.
__device__ inline void pair_terms(const Me& me, const Nb& nj, float xij, float yij, float r2, Sums& S) {
  const float rij = __builtin_amdgcn_sqrtf(r2);
  const float q = rij * me.h1;
  if (q >= 2.f) return;
  float w, der;
  if (q >= 1.f) {
    const float t = 2.f - q;
    w = me.fac * (0.25f * t * t * t);
    der = -0.75f * t * t * me.h1;
  } else {
    w = me.fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
    der = -3.0f * q * (1.f - 0.75f * q) * me.h1;
  }
  const float gsc = rij > 1e-12f ? der * me.fac * __builtin_amdgcn_rcpf(rij) : 0.f;
  const float wx = xij * gsc, wy = yij * gsc;
  float Rxx = 0.f, Rxy = 0.f, Ryy = 0.f;
  if (me.use_as) {
    float fab = w * me.inv_wdp;
    fab = me.pow4 ? (fab * fab) * (fab * fab) : powf(fab, me.corr_exp);
    Rxx = fab * (me.Rxx + nj.Rxx);
    Rxy = fab * (me.Rxy + nj.Rxy);
    Ryy = fab * (me.Ryy + nj.Ryy);
  }
  const float mj = nj.m;
  acc_add(S.Sxx_x, mj * (me.sxx + nj.sxx + Rxx) * wx);
  acc_add(S.Sxy_y, mj * (me.sxy + nj.sxy + Rxy) * wy);
  acc_add(S.Sxy_x, mj * (me.sxy + nj.sxy + Rxy) * wx);
  acc_add(S.Syy_y, mj * (me.syy + nj.syy + Ryy) * wy);
  const float dvx = nj.vx - me.vx, dvy = nj.vy - me.vy;
  acc_add(S.vx_x, dvx * wx * nj.qw);
  acc_add(S.vx_y, dvx * wy * nj.qw);
  acc_add(S.vy_x, dvy * wx * nj.qw);
  acc_add(S.vy_y, dvy * wy * nj.qw);
  const float vp = -(dvx * xij + dvy * yij);
  const float irhoij = __builtin_amdgcn_rcpf(0.5f * (me.rho + nj.rho));
  if (!(vp >= 0.f)) {
    const float cij = 0.5f * (me.c + nj.c);
    const float hij = 0.5f * (me.h + nj.h);
    const float muij = (hij * vp) * __builtin_amdgcn_rcpf(r2 + me.eta2 * hij * hij);
    const float piij = (-me.alpha * cij * muij + me.beta * muij * muij) * irhoij;
    acc_add(S.ax, -mj * piij * wx);
    acc_add(S.ay, -mj * piij * wy);
  }
  const float fx = -me.xeps_mi * w * irhoij;
  acc_add(S.xs, -fx * dvx);
  acc_add(S.ys, -fx * dvy);
}

// pair_terms as one basic block, for the LDS-staged kernels: the kernel's
// two pieces, the artificial stress and the viscosity are evaluated and
// selected with pair_terms' own predicates, and a pair at q >= 2 (possible
// only inside the cutoff's 1e-4 margin) adds zeros.  Adding +-0 to an FP64
// sum that starts at +0 leaves it unchanged, so the sums are pair_terms'
// bit for bit (for finite states), while the scheduler can overlap one
// pair's square root, reciprocals and loads with the rest of its work.
// AS: 0 no artificial stress (wdeltap <= 0), 1 exponent 4, 2 any exponent.
template <int AS>
__device__ __attribute__((always_inline)) inline void pair_terms_flat(const Me& me, const Nb& nj, float xij, float yij, float r2, Sums& S) {
  const float rij = __builtin_amdgcn_sqrtf(r2);
  const float q = rij * me.h1;
  const bool inside = !(q >= 2.f);
  const float t = 2.f - q;
  const float w_o = me.fac * (0.25f * t * t * t);
  const float der_o = -0.75f * t * t * me.h1;
  const float w_i = me.fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
  const float der_i = -3.0f * q * (1.f - 0.75f * q) * me.h1;
  const float w = inside ? (q >= 1.f ? w_o : w_i) : 0.f;
  const float der = inside ? (q >= 1.f ? der_o : der_i) : 0.f;
  const float gsc = rij > 1e-12f ? der * me.fac * __builtin_amdgcn_rcpf(rij) : 0.f;
  const float wx = xij * gsc, wy = yij * gsc;
  float Rxx = 0.f, Rxy = 0.f, Ryy = 0.f;
  if (AS > 0) {
    float fab = w * me.inv_wdp;
    fab = AS == 1 ? (fab * fab) * (fab * fab) : powf(fab, me.corr_exp);
    Rxx = fab * (me.Rxx + nj.Rxx);
    Rxy = fab * (me.Rxy + nj.Rxy);
    Ryy = fab * (me.Ryy + nj.Ryy);
  }
  const float mj = nj.m;
  acc_add(S.Sxx_x, mj * (me.sxx + nj.sxx + Rxx) * wx);
  acc_add(S.Sxy_y, mj * (me.sxy + nj.sxy + Rxy) * wy);
  acc_add(S.Sxy_x, mj * (me.sxy + nj.sxy + Rxy) * wx);
  acc_add(S.Syy_y, mj * (me.syy + nj.syy + Ryy) * wy);
  const float dvx = nj.vx - me.vx, dvy = nj.vy - me.vy;
  acc_add(S.vx_x, dvx * wx * nj.qw);
  acc_add(S.vx_y, dvx * wy * nj.qw);
  acc_add(S.vy_x, dvy * wx * nj.qw);
  acc_add(S.vy_y, dvy * wy * nj.qw);
  const float vp = -(dvx * xij + dvy * yij);
  const float irhoij = __builtin_amdgcn_rcpf(0.5f * (me.rho + nj.rho));
  const float cij = 0.5f * (me.c + nj.c);
  const float hij = 0.5f * (me.h + nj.h);
  const float muij = (hij * vp) * __builtin_amdgcn_rcpf(r2 + me.eta2 * hij * hij);
  const float pi_c = (-me.alpha * cij * muij + me.beta * muij * muij) * irhoij;
  const float piij = !(vp >= 0.f) ? pi_c : 0.f;
  acc_add(S.ax, -mj * piij * wx);
  acc_add(S.ay, -mj * piij * wy);
  const float fx = -me.xeps_mi * w * irhoij;
  acc_add(S.xs, -fx * dvx);
  acc_add(S.ys, -fx * dvy);
}

// The artificial-stress variant of a particle's scene.
__device__ inline int as_mode(const Me& me) { return !me.use_as ? 0 : me.pow4 ? 1 : 2; }

__device__ inline void jaumann64(double G2, double vx_x, double vx_y, double vy_x, double vy_y, double Sxx,
                                 double Sxy, double Syy, double& t00, double& t01, double& t11) {
  const double e01 = 0.5 * (vx_y + vy_x);
  const double w10 = 0.5 * (vy_x - vx_y);
  const double w01 = 0.5 * (vx_y - vy_x);
  const double s3 = 1. / 3. * (vx_x + vy_y);
  t00 = (G2 * (vx_x - s3) - Sxy * w10) + w01 * Sxy;
  t01 = (G2 * e01 - Syy * w10) + w10 * Sxx;
  t11 = (G2 * (vy_y - s3) - Sxy * w01) + w10 * Sxy;
}

// FP64 epilogue and RK4 stage update of sorted particle i.
__device__ inline void finish(const UlArgs& a, const SceneConst* __restrict__ c, uint32_t i, const Sums& S) {
  const uint32_t home = a.home_identity ? i : a.perm[i];
  const double vxi = a.src.f[2][home], vyi = a.src.f[3][home], rhoi = a.src.f[4][home];
  const double Sxx = a.src.f[5][home], Sxy = a.src.f[6][home], Syy = a.src.f[7][home];
  const double dSxx_x = acc_val(S.Sxx_x) * rhoi, dSxy_y = acc_val(S.Sxy_y) * rhoi;
  const double dSxy_x = acc_val(S.Sxy_x) * rhoi, dSyy_y = acc_val(S.Syy_y) * rhoi;
  const double gvx_x = acc_val(S.vx_x), gvx_y = acc_val(S.vx_y), gvy_x = acc_val(S.vy_x), gvy_y = acc_val(S.vy_y);
  double t00, t01, t11;
  jaumann64(c->G2, gvx_x, gvx_y, gvy_x, gvy_y, Sxx, Sxy, Syy, t00, t01, t11);
  double d[8];
  d[0] = acc_val(S.xs) + vxi;
  d[1] = acc_val(S.ys) + vyi;
  d[2] = acc_val(S.ax) + 1. / rhoi * (dSxx_x + dSxy_y);
  d[3] = acc_val(S.ay) + 1. / rhoi * (dSxy_x + dSyy_y);
  if (a.fcx) {  // contact forces (src/cont_mech.h:44-45)
    const uint32_t g = a.gid[home];
    d[2] = acc_val(S.ax) + (1. / rhoi * (dSxx_x + dSxy_y) + a.fcx[g] / a.m[home]);
    d[3] = acc_val(S.ay) + (1. / rhoi * (dSxy_x + dSyy_y) + a.fcy[g] / a.m[home]);
  }
  d[4] = -rhoi * (gvx_x + gvy_y);
  d[5] = t00;
  d[6] = t01;
  d[7] = t11;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(d[k]);
  if (!ok) raise_err(a.err, a.berr, (uint32_t)(c - a.sc), EB_NONFINITE);
  if (a.stage == 0) {
    const uint32_t gid = a.gid[home];
    double* o = a.d8 + 8 * (size_t)gid;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = d[k];
    if (a.aux) {
      const PairRec32 r = a.rec32[i];
      double* x = a.aux + 16 * (size_t)gid;
      x[0] = r.qw;
      x[1] = c->c0sq * (rhoi - c->rho0);
      x[2] = r.Rxx;
      x[3] = r.Rxy;
      x[4] = r.Ryy;
      x[5] = dSxx_x;
      x[6] = dSxy_y;
      x[7] = dSxy_x;
      x[8] = dSyy_y;
      x[9] = gvx_x;
      x[10] = gvx_y;
      x[11] = gvy_x;
      x[12] = gvy_y;
      x[13] = 0.;
      x[14] = 0.;
      x[15] = 0.;
    }
    return;
  }
  rk_stage_update(a, c, home, d);
}

// Sorted particle i straight from global memory, no buffering (the
// tile kernel's path for tiles whose candidates do not fit in LDS).
__device__ inline void particle_global(const UlArgs& a, uint32_t i) {
  const uint2 mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  // a body whose grid fell back to one cell (g.err) fails the step: scan no
  // cells rather than all n particles, and let the epilogue write finite values
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = g.err ? lo_i : min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const float cellf = (float)g.cell;
  const float2 pi = a.pos32[i];
  const Me me = make_me(c, nb_from_rec(a.rec32[i]));
  Sums S;
  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny;
    const float offx = (float)((int)gi - (int)ni) * cellf;
    uint32_t jb = a.cell_start[rb + lo_j];
    for (uint32_t nj = lo_j; nj < hi_j; nj++) {
      const uint32_t ja = jb;
      jb = a.cell_start[rb + nj + 1];
      const float offy = (float)((int)gj - (int)nj) * cellf;
      for (uint32_t j = ja; j < jb; j++) {
        if (j == i) continue;
        const float2 pj = a.pos32[j];
        const float xij = (pi.x - pj.x) + offx;
        const float yij = (pi.y - pj.y) + offy;
        const float r2 = xij * xij + yij * yij;
        if (r2 > me.cut2) continue;
        pair_terms(me, nb_from_rec(a.rec32[j]), xij, yij, r2, S);
      }
    }
  }
  finish(a, c, i, S);
}

}  // namespace

constexpr int kNbCap = 32;  // compacted neighbours buffered per thread
constexpr int kScanBatch = 4;  // candidate positions loaded together per lane
constexpr uint32_t kMaxGraphDeg = 64;                     // learned-step graph: edge slots per node
constexpr float kSplineNormF = 0.45472840883398667f;      // 10/(7 pi): 2D cubic spline factor times h^2

// Per particle from global memory: scan the 3x3 block, buffer the candidates
// inside the cutoff in LDS (sorted position + cell offsets packed in 32
// bits), then evaluate the pair terms on the compact list with full SIMD
// occupancy.
template <bool GRAPH>
__global__ __launch_bounds__(kBlock) void k_ul_pairs32(UlArgs a) {
  __shared__ uint32_t nb_lds[kNbCap * kBlock];
  uint32_t* const my = nb_lds + threadIdx.x;
  const uint32_t i = xcd_tile_chunked(a.xcd_chunk) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint2 mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  // a body whose grid fell back to one cell (g.err) fails the step: scan no
  // cells rather than all n particles, and let the epilogue write finite values
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = g.err ? lo_i : min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const float cellf = (float)g.cell;
  const float2 pi = a.pos32[i];
  const Me me = make_me(c, nb_from_rec(a.rec32[i]));
  Sums S;
  int nb = 0;
  // learned-step graph (D1 of a learned step): the compacted candidates are
  // a superset of its edges r < 2h (uniform h: h_ij = h); classified in FP32
  // from the cell-relative positions (about 1e-6 relative on r^2), with the
  // FP64 predicate of the graph builder within 4e-6 of the cutoff, so the
  // edge set and order are those of k_graph
  constexpr bool graph = GRAPH;
  uint32_t gd = 0;
  float gws = kSplineNormF;
  const float gcut = 4.f * me.h * me.h;
  // the buffered neighbours in order; the next one's position and record
  // are loaded while the current one is evaluated (the last load repeats the
  // last entry, in bounds)
  auto flush = [&]() {
    uint32_t en = nb > 0 ? my[0] : 0u;
    float2 pn = a.pos32[en & 0x0FFFFFFFu];
    PairRec32 rn = a.rec32[en & 0x0FFFFFFFu];
    for (int k = 0; k < nb; k++) {
      const uint32_t e = en;
      const uint32_t j = e & 0x0FFFFFFFu;
      const float2 pj = pn;
      const PairRec32 rj = rn;
      en = my[min(k + 1, nb - 1) * kBlock];
      pn = a.pos32[en & 0x0FFFFFFFu];
      rn = a.rec32[en & 0x0FFFFFFFu];
      const float xij = (pi.x - pj.x) + (float)((int)((e >> 28) & 3u) - 1) * cellf;
      const float yij = (pi.y - pj.y) + (float)((int)(e >> 30) - 1) * cellf;
      const float r2 = xij * xij + yij * yij;
      if (graph && !(r2 > gcut * (1.f + 4e-6f))) {
        bool edge = true;
        if (r2 >= gcut * (1.f - 4e-6f)) {
          const double4 di = a.gdpos[i], dj = a.gdpos[j];
          const double dx = di.x - dj.x, dy = di.y - dj.y, hij = 0.5 * (di.z + dj.z);
          edge = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)) < __dmul_rn(__dmul_rn(4.0, hij), hij);
        }
        if (edge) {
          const float q = __builtin_amdgcn_sqrtf(r2) * me.h1;
          const float t = fmaxf(2.f - q, 0.f);
          gws += q < 1.f ? kSplineNormF * (1.f - 1.5f * q * q * (1.f - 0.5f * q)) : kSplineNormF * 0.25f * t * t * t;
          if (gd < kMaxGraphDeg) a.gcol[(size_t)gd * a.gstride + i] = j;
          gd++;
        }
      }
      pair_terms(me, nb_from_rec(rj), xij, yij, r2, S);
    }
    nb = 0;
  };
  // Each column of the 3x3 block is one range of consecutive candidates (its
  // 1-3 cells are consecutive in the cell order): its cell boundaries are
  // loaded together and the candidates in batches of kScanBatch positions,
  // so that several gathers are in flight per lane (the kernel is bound by
  // their latency).  Buffer flushes are collective: before each column the
  // lanes flush together if any lane's buffer is likely to overflow in it
  // (half its candidates, about a third of a column lying inside the
  // cutoff, plus one batch).  A flush taken by each lane on its own inside
  // the scan would run once per lane, with the other lanes idle, wherever
  // dense regions fill the buffers at different times; the flush before a
  // batch stays as the guard for the rare lane that still fills its buffer.
  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny + lo_j;
    const uint32_t nc = hi_j - lo_j;  // 1..3 cells
    const uint32_t b0 = a.cell_start[rb], b3 = a.cell_start[rb + nc];
    const uint32_t b1 = nc > 1 ? a.cell_start[rb + 1] : 0xFFFFFFFFu;
    const uint32_t b2 = nc > 2 ? a.cell_start[rb + 2] : 0xFFFFFFFFu;
    if (__any(nb + ((b3 - b0 + 1) >> 1) + kScanBatch > (uint32_t)kNbCap)) flush();
    const float offx = (float)((int)gi - (int)ni) * cellf;
    const uint32_t tx = (gi - ni + 1) << 28;
    for (uint32_t j0 = b0; j0 < b3; j0 += kScanBatch) {
      if (__any(nb > kNbCap - kScanBatch)) flush();  // guard: room for the batch, taken together
      float2 pj[kScanBatch];
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) pj[e] = a.pos32[min(j0 + e, b3 - 1)];
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) {
        const uint32_t j = j0 + e;
        if (j < b3 && j != i) {
          const uint32_t nj = lo_j + (j >= b1 ? 1u : 0u) + (j >= b2 ? 1u : 0u);
          const float xij = (pi.x - pj[e].x) + offx;
          const float yij = (pi.y - pj[e].y) + (float)((int)gj - (int)nj) * cellf;
          if (xij * xij + yij * yij <= me.cut2) my[nb++ * kBlock] = j | tx | ((gj - nj + 1) << 30);
        }
      }
    }
  }
  flush();
  if (graph) {
    if (gd > kMaxGraphDeg) {
      raise_err(a.err, a.berr, mt.y, EB_NBH);
      gd = kMaxGraphDeg;
    }
    a.gdeg[i] = gd;
    a.gwsum[i] = gws;
  }
  finish(a, c, i, S);
}

// ---------------------------------------------------------------------------
// LDS-staged tiles

constexpr int kTileCap = 768;  // staged candidates per tile (10-bit LDS index)
constexpr int kTileNb = 32;    // compacted neighbours buffered per thread

__global__ __launch_bounds__(kBlock) void k_ul_tiles32(UlArgs a) {
  __shared__ float s_px[kTileCap], s_py[kTileCap];
  __shared__ float s_vx[kTileCap], s_vy[kTileCap], s_rho[kTileCap], s_qw[kTileCap], s_m[kTileCap];
  __shared__ float s_h[kTileCap], s_c[kTileCap], s_sxx[kTileCap], s_sxy[kTileCap], s_syy[kTileCap];
  __shared__ float s_Rxx[kTileCap], s_Rxy[kTileCap], s_Ryy[kTileCap];
  __shared__ uint16_t s_nb[kTileNb * kBlock];
  __shared__ uint32_t s_rs[3], s_off[4];

  const uint32_t TC = a.tile_cells;
  const uint32_t total = a.tile_off[a.nscene];
  const uint32_t per = gridDim.x / kXcds;
  const uint32_t xcd = blockIdx.x % kXcds, k0 = blockIdx.x / kXcds;
  const uint32_t range = (total + kXcds - 1) / kXcds;
  const uint32_t t_begin = xcd * range, t_end = min(total, t_begin + range);
  uint16_t* const my = s_nb + threadIdx.x;

  for (uint32_t t = t_begin + k0; t < t_end; t += per) {
    // body of tile t: last s with tile_off[s] <= t
    uint32_t lo = 0, hi = a.nscene;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.tile_off[mid] <= t)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t s = lo;
    const SceneConst* __restrict__ c = a.sc + s;
    const SceneGeom g = a.geom[s];
    const uint32_t lt = t - a.tile_off[s];
    const uint32_t ix = lt / g.tpc;
    const uint32_t iy0 = (lt - ix * g.tpc) * TC, iy1 = min(iy0 + TC, g.ny);
    const uint32_t iy_lo = iy0 > 0 ? iy0 - 1 : 0, iy_hi = min(iy1 + 1, g.ny);
    const uint32_t rb1 = c->cell_offset + ix * g.ny;
    const uint32_t t0 = a.cell_start[rb1 + iy0], t1 = a.cell_start[rb1 + iy1];
    if (t0 == t1) continue;  // empty tile (the same for every thread)

    if (threadIdx.x < 3) {
      const int row = (int)ix - 1 + (int)threadIdx.x;
      uint32_t st = 0, len = 0;
      if (row >= 0 && row < (int)g.nx) {
        const uint32_t rb = c->cell_offset + (uint32_t)row * g.ny;
        st = a.cell_start[rb + iy_lo];
        len = a.cell_start[rb + iy_hi] - st;
      }
      s_rs[threadIdx.x] = st;
      s_off[threadIdx.x + 1] = len;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_off[0] = 0;
      s_off[2] += s_off[1];
      s_off[3] += s_off[2];
    }
    __syncthreads();
    const uint32_t ncand = s_off[3];
    const bool staged = ncand <= (uint32_t)kTileCap;
    if (staged) {
      for (uint32_t q = threadIdx.x; q < ncand; q += blockDim.x) {
        const uint32_t r = q < s_off[1] ? 0 : (q < s_off[2] ? 1 : 2);
        const uint32_t j = s_rs[r] + (q - s_off[r]);
        const float2 p = a.pos32[j];
        const Nb n = nb_from_rec(a.rec32[j]);
        s_px[q] = p.x;
        s_py[q] = p.y;
        s_vx[q] = n.vx;
        s_vy[q] = n.vy;
        s_rho[q] = n.rho;
        s_qw[q] = n.qw;
        s_m[q] = n.m;
        s_h[q] = n.h;
        s_c[q] = n.c;
        s_sxx[q] = n.sxx;
        s_sxy[q] = n.sxy;
        s_syy[q] = n.syy;
        s_Rxx[q] = n.Rxx;
        s_Rxy[q] = n.Rxy;
        s_Ryy[q] = n.Ryy;
      }
    }
    __syncthreads();
    const float cellf = (float)g.cell;
    for (uint32_t i = t0 + threadIdx.x; i < t1; i += blockDim.x) {
      if (!staged) {
        particle_global(a, i);
        continue;
      }
      const uint32_t li = s_off[1] + (i - s_rs[1]);
      Nb ni;
      ni.vx = s_vx[li];
      ni.vy = s_vy[li];
      ni.rho = s_rho[li];
      ni.qw = s_qw[li];
      ni.m = s_m[li];
      ni.h = s_h[li];
      ni.c = s_c[li];
      ni.sxx = s_sxx[li];
      ni.sxy = s_sxy[li];
      ni.syy = s_syy[li];
      ni.Rxx = s_Rxx[li];
      ni.Rxy = s_Rxy[li];
      ni.Ryy = s_Ryy[li];
      const Me me = make_me(c, ni);
      const float pix = s_px[li], piy = s_py[li];
      const uint32_t gj = (a.meta[i].x - c->cell_offset) - ix * g.ny;
      const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
      Sums S;
      int nb = 0;
      auto flush = [&]() {
        for (int k = 0; k < nb; k++) {
          const uint32_t e = my[k * kBlock];
          const uint32_t lj = e & 1023u;
          const float xij = (pix - s_px[lj]) + (float)((int)((e >> 10) & 3u) - 1) * cellf;
          const float yij = (piy - s_py[lj]) + (float)((int)((e >> 12) & 3u) - 1) * cellf;
          Nb nj;
          nj.vx = s_vx[lj];
          nj.vy = s_vy[lj];
          nj.rho = s_rho[lj];
          nj.qw = s_qw[lj];
          nj.m = s_m[lj];
          nj.h = s_h[lj];
          nj.c = s_c[lj];
          nj.sxx = s_sxx[lj];
          nj.sxy = s_sxy[lj];
          nj.syy = s_syy[lj];
          nj.Rxx = s_Rxx[lj];
          nj.Rxy = s_Rxy[lj];
          nj.Ryy = s_Ryy[lj];
          pair_terms(me, nj, xij, yij, xij * xij + yij * yij, S);
        }
        nb = 0;
      };
      for (uint32_t r = 0; r < 3; r++) {
        const int row = (int)ix - 1 + (int)r;
        if (row < 0 || row >= (int)g.nx) continue;
        const uint32_t rb = c->cell_offset + (uint32_t)row * g.ny;
        const float offx = (float)(1 - (int)r) * cellf;
        const uint32_t base = s_off[r] - s_rs[r];
        uint32_t jb = a.cell_start[rb + lo_j];
        // collective flush before a row that may overflow a lane's buffer
        // (as in k_ul_pairs32)
        if (__any(nb + ((a.cell_start[rb + hi_j] - jb + 1) >> 1) > (uint32_t)kTileNb)) flush();
        for (uint32_t nj = lo_j; nj < hi_j; nj++) {
          const uint32_t ja = jb;
          jb = a.cell_start[rb + nj + 1];
          const uint32_t dj = gj + 1 - nj;  // gj - nj + 1 in {0, 1, 2}
          const float offy = (float)((int)dj - 1) * cellf;
          const uint32_t tag = ((2u - r) << 10) | (dj << 12);
          for (uint32_t j = ja; j < jb; j++) {
            if (j == i) continue;
            const uint32_t lj = base + j;
            const float xij = (pix - s_px[lj]) + offx;
            const float yij = (piy - s_py[lj]) + offy;
            if (xij * xij + yij * yij > me.cut2) continue;
            my[nb * kBlock] = (uint16_t)(lj | tag);
            if (++nb == kTileNb) flush();
          }
        }
      }
      flush();
      finish(a, c, i, S);
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// ---------------------------------------------------------------------------
// Neighbour-list path: one list per RK4 step
//
// The step's single binning uses cells of side 2 h_max + skin, with the skin
// at least four times the distance the fastest particle of the body covers
// in one time step.  Stage 1 (k_ul_list_first32) scans the 3x3 block once
// and keeps every candidate within 2 h_i + skin; between stage 1
// and any later stage of the step no pair's separation changes by more than
// the skin, so the list holds every pair with r < 2 h_i at every stage, and
// each stage evaluates exactly the pairs q < 2 of its own positions (the set
// the reference re-bins for at every stage).  The epilogue of each stage
// checks that the particle stays within half the skin of its stage-1
// position (error bit EB_SKIN otherwise).  Sums run in list order.


// Epilogue of a list stage: FP64 derivatives of sorted particle i (storage
// index i: the stage-1 prepass moved storage into sorted order), the RK4
// stage update, and the record of the next stage.  The L lanes of the
// particle hold the same sums; lane q stores the fields f with f % L == q.

// What the epilogue of sorted particle i reads from HBM: the stage's state,
// the step's start state (stage 4: the fields of lane q only), the weighted
// rates so far (stages 2-4, the fields of lane q), mass and smoothing length.
// Each is read and written once per stage: streamed past the caches so that
// they keep the positions and records the neighbour loads re-read.
struct EpiIn {
  double vx, vy, rho, Sxx, Sxy, Syy;
  double y[8], acc[8];
  double m, h;
};

template <int L>
__device__ inline EpiIn epi_load(const UlArgs& a, uint32_t i, uint32_t q) {
  EpiIn e;
  e.vx = ld_nt(&a.src.f[2][i]);
  e.vy = ld_nt(&a.src.f[3][i]);
  e.rho = ld_nt(&a.src.f[4][i]);
  e.Sxx = ld_nt(&a.src.f[5][i]);
  e.Sxy = ld_nt(&a.src.f[6][i]);
  e.Syy = ld_nt(&a.src.f[7][i]);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool mine = k % L == (int)q;
    e.y[k] = a.stage != 4 || mine ? ld_nt(&a.y.f[k][i]) : 0.;
    e.acc[k] = a.stage != 1 && mine ? ld_nt(&a.acc.f[k][i]) : 0.;
  }
  e.m = a.m[i];
  e.h = a.h[i];
  return e;
}

// PRE: the inputs were loaded ahead into e (epi_load); otherwise they are
// loaded here, next to their use, which keeps fewer registers live.
template <int L, bool PRE>
__device__ inline void finish_list(const UlArgs& a, const SceneConst* __restrict__ c, const SceneGeom& g,
                                   uint32_t key, uint32_t i, uint32_t q, const double* v, const EpiIn& e) {
  const double vxi = PRE ? e.vx : ld_nt(&a.src.f[2][i]), vyi = PRE ? e.vy : ld_nt(&a.src.f[3][i]);
  const double rhoi = PRE ? e.rho : ld_nt(&a.src.f[4][i]);
  const double Sxx = PRE ? e.Sxx : ld_nt(&a.src.f[5][i]), Sxy = PRE ? e.Sxy : ld_nt(&a.src.f[6][i]);
  const double Syy = PRE ? e.Syy : ld_nt(&a.src.f[7][i]);
  auto y_k = [&](int k) { return PRE ? e.y[k] : ld_nt(&a.y.f[k][i]); };
  auto acc_k = [&](int k) { return PRE ? e.acc[k] : ld_nt(&a.acc.f[k][i]); };
  const double dSxx_x = v[0] * rhoi, dSxy_y = v[1] * rhoi, dSxy_x = v[2] * rhoi, dSyy_y = v[3] * rhoi;
  double t00, t01, t11;
  jaumann64(c->G2, v[4], v[5], v[6], v[7], Sxx, Sxy, Syy, t00, t01, t11);
  double d[8];
  d[0] = v[10] + vxi;
  d[1] = v[11] + vyi;
  d[2] = v[8] + 1. / rhoi * (dSxx_x + dSxy_y);
  d[3] = v[9] + 1. / rhoi * (dSxy_x + dSyy_y);
  if (a.fcx) {  // contact forces (src/cont_mech.h:44-45)
    const uint32_t g = a.gid[i];
    d[2] = v[8] + (1. / rhoi * (dSxx_x + dSxy_y) + a.fcx[g] / a.m[i]);
    d[3] = v[9] + (1. / rhoi * (dSxy_x + dSyy_y) + a.fcy[g] / a.m[i]);
  }
  d[4] = -rhoi * (v[4] + v[7]);
  d[5] = t00;
  d[6] = t01;
  d[7] = t11;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(d[k]);
  if (!ok && q == 0) raise_err(a.err, a.berr, (uint32_t)(c - a.sc), EB_NONFINITE);
  if (a.stage == 4) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k % L == (int)q) st_nt(&a.y.f[k][i], y_k(k) + c->dt6 * (acc_k(k) + d[k]));
    return;
  }
  const double dts = a.stage == 3 ? c->dt : c->half_dt;
  double kn[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    kn[k] = y_k(k) + d[k] * dts;
    if (k % L == (int)q) {
      st_nt(&a.acc.f[k][i], a.stage == 1 ? d[k] : acc_k(k) + 2 * d[k]);
      if (k >= 2) st_nt(&a.K.f[k][i], kn[k]);  // the positions of a stage live in its records
    }
  }
  const double ux = d[0] * dts, uy = d[1] * dts, half = 0.5 * g.skin;
  if (q == 0 && ux * ux + uy * uy > half * half) raise_err(a.err, a.berr, (uint32_t)(c - a.sc), EB_SKIN);
  const uint32_t local = key - c->cell_offset;
  const uint32_t ix = local / g.ny, iy = local - ix * g.ny;
  float2 pos;
  PairRec32 r;
  make_rec32(c, g.bbmin_x + (double)ix * g.cell, g.bbmin_y + (double)iy * g.cell, kn[0], kn[1], kn[2], kn[3], kn[4],
             kn[5], kn[6], kn[7], PRE ? e.m : a.m[i], PRE ? e.h : a.h[i], pos, r);
  if (q == 0) a.pos32_out[i] = pos;
  const float4* rv = reinterpret_cast<const float4*>(&r);
  float4* out = reinterpret_cast<float4*>(a.rec32_out + i);
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (k % L == (int)q) out[k] = rv[k];
}

// One RK4 stage (a.stage 1-4) from the step's list, L lanes per particle:
// lane q takes the entries q, q + L, ...  In chunks of kListBatch entries per
// lane, the entries inside the cutoff at this stage's positions are first
// compacted in LDS (list word and position only), then evaluated with the
// next one's position and record loaded ahead, so the expensive pair terms
// run with every lane busy (about 70 % of the entries are inside the
// cutoff).  The FP64 partial sums of the L lanes are combined with
// cross-lane shuffles.
constexpr int kListBatch = 32;
// Entry k of particle i in the list: groups of 4 entries, group-major, 16 B
// per particle and group (a wave's 64 particles read a group as 1 KB).
__device__ inline size_t list_at(uint32_t n, uint32_t i, uint32_t k) {
  return ((size_t)(k >> 2) * n + i) * 4 + (k & 3u);
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// Entry k of particle i in the band-slot list: groups of 8 words, group-major.
__device__ inline size_t list16_at(uint32_t n, uint32_t i, uint32_t k) {
  return ((size_t)(k >> 3) * n + i) * 8 + (k & 7u);
}

// Sorted particle i, lane q of L, every neighbour gathered from global
// memory; my = this thread's compaction buffer (kListBatch words, stride kBlock);
// the positions of B entries are loaded together.
template <int L, int B = kScanBatch>
__device__ inline void list_stage_global(const UlArgs& a, uint32_t i, uint32_t q, uint32_t* my) {
  const uint2 mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const float cellf = (float)g.cell;
  const float2 pi = a.pos32[i];
  const Me me = make_me(c, nb_from_rec(a.rec32[i]));
  const uint32_t deg = a.nl_deg[i];
  Sums S;
  auto sep = [&](uint32_t e, float2 pj, float& xij, float& yij) {
    xij = (pi.x - pj.x) + (float)((int)((e >> 28) & 3u) - 1) * cellf;
    yij = (pi.y - pj.y) + (float)((int)(e >> 30) - 1) * cellf;
  };
  for (uint32_t base = q; base < deg; base += kListBatch * L) {
    const uint32_t end = min(deg, base + kListBatch * L);
    int nb = 0;
    for (uint32_t k0 = base; k0 < end; k0 += B * L) {
      uint32_t e[B];
      float2 pj[B];
#pragma unroll
      for (int t = 0; t < B; t++) {
        const uint32_t k = k0 + t * L;
        e[t] = ld_nt(&a.nl_col[list_at(a.n, i, k < end ? k : base)]);  // each list entry is read once
      }
#pragma unroll
      for (int t = 0; t < B; t++) pj[t] = a.pos32[e[t] & kListMask];
#pragma unroll
      for (int t = 0; t < B; t++) {
        float xij, yij;
        sep(e[t], pj[t], xij, yij);
        if (k0 + t * L < end && xij * xij + yij * yij <= me.cut2) my[nb++ * kBlock] = e[t];
      }
    }
    // the compacted entries in order, the next one's position and record
    // loaded while the current one is evaluated (the last load repeats the
    // last entry; with no entry it loads the particle itself)
    uint32_t en = nb > 0 ? my[0] : i;
    float2 pn = a.pos32[en & kListMask];
    PairRec32 rn = a.rec32[en & kListMask];
    for (int t = 0; t < nb; t++) {
      const uint32_t e = en;
      const float2 pj = pn;
      const PairRec32 rj = rn;
      en = my[min(t + 1, nb - 1) * kBlock];
      pn = a.pos32[en & kListMask];
      rn = a.rec32[en & kListMask];
      float xij, yij;
      sep(e, pj, xij, yij);
      pair_terms(me, nb_from_rec(rj), xij, yij, xij * xij + yij * yij, S);
    }
  }
  double v[12] = {acc_val(S.Sxx_x), acc_val(S.Sxy_y), acc_val(S.Sxy_x), acc_val(S.Syy_y),
                  acc_val(S.vx_x),  acc_val(S.vx_y),  acc_val(S.vy_x),  acc_val(S.vy_y),
                  acc_val(S.ax),    acc_val(S.ay),    acc_val(S.xs),    acc_val(S.ys)};
#pragma unroll
  for (int o = 1; o < L; o <<= 1)
#pragma unroll
    for (int t = 0; t < 12; t++) v[t] += __shfl_xor(v[t], o, 64);
  finish_list<L, false>(a, c, g, mt.x, i, q, v, EpiIn{});
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_ul_list32(UlArgs a) {
  static_assert(L == 1 || L == 2 || L == 4, "lanes per particle");
  __shared__ uint32_t nb_lds[kListBatch * kBlock];
  const uint32_t i = xcd_tile_chunked(a.xcd_chunk) * (kBlock / L) + threadIdx.x / L;
  if (i >= a.n) return;  // the L lanes of a particle leave together
  list_stage_global<L>(a, i, threadIdx.x % L, nb_lds + threadIdx.x);
}

// The same stage with the neighbours staged in LDS.  A workgroup takes 256
// consecutive sorted particles, whose cell keys span [k0, k1] of one body.
// Every list entry of tag tx (neighbour cell column = own column + 1 - tx)
// lies in the key range [k0 + (1 - tx) ny - 1, k1 + (1 - tx) ny + 1]: three
// contiguous runs of sorted storage (bands), at most kBandCap candidates in
// all for the bodies this is built for (about 3.3 per particle).  The bands'
// positions and records are copied into LDS once with coalesced loads, and
// each particle's list is evaluated from LDS: the scattered 64-byte record
// gathers of the global path become LDS reads.  The pair sums run over the
// same entries in the same order with the same values, so the result is the
// global path's bit for bit.  A workgroup whose particles span two bodies or
// whose bands exceed kBandCap takes the global path.
constexpr int kBandCap = 1024;  // staged candidates per workgroup (10-bit LDS slot)

struct BandLds {
  float px[kBandCap], py[kBandCap];
  float vx[kBandCap], vy[kBandCap], rho[kBandCap], qw[kBandCap], m[kBandCap], h[kBandCap], c[kBandCap];
  float sxx[kBandCap], sxy[kBandCap], syy[kBandCap], Rxx[kBandCap], Rxy[kBandCap], Ryy[kBandCap];
};
static_assert(sizeof(BandLds) >= sizeof(uint32_t) * kListBatch * kBlock, "global path buffer aliases the bands");

__device__ inline Nb nb_from_lds(const BandLds& s, uint32_t k) {
  Nb n;
  n.vx = s.vx[k];
  n.vy = s.vy[k];
  n.rho = s.rho[k];
  n.qw = s.qw[k];
  n.m = s.m[k];
  n.h = s.h[k];
  n.c = s.c[k];
  n.sxx = s.sxx[k];
  n.sxy = s.sxy[k];
  n.syy = s.syy[k];
  n.Rxx = s.Rxx[k];
  n.Rxy = s.Rxy[k];
  n.Ryy = s.Ryy[k];
  return n;
}

// Pair terms of the nb compacted entries my[0], my[kBlock], ... (LDS slot |
// tx << 10 | ty << 12), in order.  The next entry's fields and the word after
// it are read while the current one is evaluated (reads past the end repeat
// the last entry, and with no entry stay inside the buffer, unused).
template <int AS>
__device__ __attribute__((always_inline)) inline void eval_lds_as(const BandLds& s, const uint16_t* my, int nb, const Me& me, float pix, float piy,
                                   float cellf, Sums& S) {
  const int last = max(nb - 1, 0);
  uint32_t wn = my[0], wnn = my[min(1, last) * kBlock];
  float pxn = s.px[wn & (kBandCap - 1)], pyn = s.py[wn & (kBandCap - 1)];
  Nb nn = nb_from_lds(s, wn & (kBandCap - 1));
  for (int t = 0; t < nb; t++) {
    const uint32_t w = wn;
    const float pxj = pxn, pyj = pyn;
    const Nb nj = nn;
    wn = wnn;
    wnn = my[min(t + 2, last) * kBlock];
    const uint32_t kn = wn & (kBandCap - 1);
    pxn = s.px[kn];
    pyn = s.py[kn];
    nn = nb_from_lds(s, kn);
    const float xij = (pix - pxj) + (float)((int)((w >> 10) & 3u) - 1) * cellf;
    const float yij = (piy - pyj) + (float)((int)(w >> 12) - 1) * cellf;
    pair_terms_flat<AS>(me, nj, xij, yij, xij * xij + yij * yij, S);
  }
}

__device__ __attribute__((always_inline)) inline void eval_lds(const BandLds& s, const uint16_t* my, int nb, const Me& me, float pix, float piy,
                                float cellf, Sums& S) {
  const int as = as_mode(me);
  if (as == 1)
    eval_lds_as<1>(s, my, nb, me, pix, piy, cellf, S);
  else if (as == 0)
    eval_lds_as<0>(s, my, nb, me, pix, piy, cellf, S);
  else
    eval_lds_as<2>(s, my, nb, me, pix, piy, cellf, S);
}

// One thread of the workgroup of sorted particles [i0, i0 + kBlock): the
// three bands (slot = j + base[tx]; band tx ends at slot end[tx]); returns
// the number of staged candidates, or -1 for the global path.
__device__ inline int32_t band_setup(const UlArgs& a, uint32_t i0, uint32_t* base, uint32_t* end) {
  const uint32_t i1 = min(a.n, i0 + kBlock) - 1;
  const uint2 m0 = a.meta[i0], m1 = a.meta[i1];
  if (m0.y != m1.y) return -1;
  const SceneConst* __restrict__ c = a.sc + m0.y;
  const SceneGeom g = a.geom[m0.y];
  const int64_t kb = c->cell_offset, ke = kb + (int64_t)g.num_cell - 1, ny = g.ny;
  uint32_t t = 0;
  for (int b = 0; b < 3; b++) {
    const int64_t ka = max(kb, (int64_t)m0.x + (1 - b) * ny - 1);
    const int64_t kz = min(ke, (int64_t)m1.x + (1 - b) * ny + 1);
    const uint32_t lo = ka <= kz ? a.cell_start[ka] : 0u;
    const uint32_t len = ka <= kz ? a.cell_start[kz + 1] - lo : 0u;
    base[b] = t - lo;
    t += len;
    end[b] = t;
  }
  return t <= (uint32_t)kBandCap ? (int32_t)t : -1;
}

// The whole workgroup: positions and records of the tot staged candidates.
__device__ inline void band_fill(const UlArgs& a, BandLds& s, int32_t tot, const uint32_t* base,
                                 const uint32_t* end) {
  for (int32_t k = threadIdx.x; k < tot; k += kBlock) {
    const uint32_t b = (uint32_t)k < end[0] ? 0 : ((uint32_t)k < end[1] ? 1 : 2);
    const uint32_t j = (uint32_t)k - base[b];
    const float2 p = a.pos32[j];
    const PairRec32 r = a.rec32[j];
    s.px[k] = p.x;
    s.py[k] = p.y;
    s.vx[k] = r.vx;
    s.vy[k] = r.vy;
    s.rho[k] = r.rho;
    s.qw[k] = r.qw;
    s.m[k] = r.m;
    s.h[k] = r.h;
    s.c[k] = r.c;
    s.sxx[k] = r.sxx_r;
    s.sxy[k] = r.sxy_r;
    s.syy[k] = r.syy_r;
    s.Rxx[k] = r.Rxx;
    s.Rxy[k] = r.Rxy;
    s.Ryy[k] = r.Ryy;
  }
}

__global__ __launch_bounds__(kBlock, 2) void k_ul_list_lds32(UlArgs a) {
  __shared__ BandLds s;
  __shared__ uint16_t s_nb[kListBatch * kBlock];
  __shared__ uint32_t s_base[3], s_end[3];
  __shared__ int32_t s_tot;
  const uint32_t i0 = xcd_tile_chunked(a.xcd_chunk) * kBlock;
  if (i0 >= a.n) return;  // the whole workgroup
  if (threadIdx.x == 0) s_tot = band_setup(a, i0, s_base, s_end);
  __syncthreads();
  const int32_t tot = s_tot;
  const uint32_t i = i0 + threadIdx.x;
  // issued first, to arrive while the bands are staged: the epilogue's state
  // and the first kListBatch band-slot words of the list (each read once, 8
  // per load, two per register)
  EpiIn ein;
  uint32_t deg = 0, ew[kListBatch / 2];
  const u32x4* __restrict__ col8 = reinterpret_cast<const u32x4*>(a.nl16) + i;
  if (tot >= 0 && i < a.n) {
    ein = epi_load<1>(a, i, 0);
    deg = a.nl_deg[i];
#pragma unroll
    for (int g = 0; g < kListBatch / 8; g++) {
      const u32x4 w = (uint32_t)(8 * g) < deg ? ld_nt(&col8[(size_t)g * a.n]) : u32x4{0u, 0u, 0u, 0u};
      ew[4 * g] = w.x; ew[4 * g + 1] = w.y; ew[4 * g + 2] = w.z; ew[4 * g + 3] = w.w;
    }
  }
  band_fill(a, s, tot, s_base, s_end);
  __syncthreads();  // the last barrier: threads past the end leave below
  if (i >= a.n) return;
  if (tot < 0) {
    // batches of 8 here: the same bits, and the staged path above and below
    // compiles to a faster schedule (0.177 -> 0.169 ms per stage) than with 4
    list_stage_global<1, 8>(a, i, 0, reinterpret_cast<uint32_t*>(&s) + threadIdx.x);
    return;
  }
  const uint32_t base1 = s_base[1];
  uint16_t* const my = s_nb + threadIdx.x;
  const uint2 mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const float cellf = (float)g.cell;
  const uint32_t li = i + base1;  // the particle's own cell lies in band 1
  const float pix = s.px[li], piy = s.py[li];
  const Me me = make_me(c, nb_from_lds(s, li));
  Sums S;
  for (uint32_t base = 0; base < deg; base += kListBatch) {
    if (base > 0) {  // a list longer than one batch: its next words
#pragma unroll
      for (int g = 0; g < kListBatch / 8; g++) {
        const u32x4 w = base + 8 * g < deg ? ld_nt(&col8[(size_t)((base >> 3) + g) * a.n]) : u32x4{0u, 0u, 0u, 0u};
        ew[4 * g] = w.x; ew[4 * g + 1] = w.y; ew[4 * g + 2] = w.z; ew[4 * g + 3] = w.w;
      }
    }
    int nb = 0;
#pragma unroll
    for (int t = 0; t < kListBatch; t++) {
      // slot | tx << 10 | ty << 12; words past the list's end give an in-range
      // slot through the mask and are not kept
      const uint32_t e = (ew[t >> 1] >> (16 * (t & 1))) & 0xFFFFu, sl = e & (kBandCap - 1);
      const float xij = (pix - s.px[sl]) + (float)((int)((e >> 10) & 3u) - 1) * cellf;
      const float yij = (piy - s.py[sl]) + (float)((int)(e >> 12) - 1) * cellf;
      if (base + t < deg && xij * xij + yij * yij <= me.cut2) my[nb++ * kBlock] = (uint16_t)e;
    }
    eval_lds(s, my, nb, me, pix, piy, cellf, S);
  }
  const double v[12] = {acc_val(S.Sxx_x), acc_val(S.Sxy_y), acc_val(S.Sxy_x), acc_val(S.Syy_y),
                        acc_val(S.vx_x),  acc_val(S.vx_y),  acc_val(S.vy_x),  acc_val(S.vy_y),
                        acc_val(S.ax),    acc_val(S.ay),    acc_val(S.xs),    acc_val(S.ys)};
  finish_list<1, true>(a, c, g, mt.x, i, 0, v, ein);
}

// Stage 1 of a list step in one pass: the 3x3 block of the step's cells is
// scanned once; every candidate within 2 h_i + skin goes to the step's list
// (32-bit words, groups of 4, group-major), those inside the cutoff also to the LDS buffer, which is
// evaluated as in k_ul_pairs32 (collective flushes before a column that may
// overflow a lane's buffer).  The epilogue is the list stages' one.
// Sorted particle i from global memory; my = this thread's buffer (kNbCap
// words, stride kBlock).
__device__ inline void list_first_global(const UlArgs& a, uint32_t i, uint32_t* my) {
  const uint2 mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  // a body whose grid fell back to one cell (g.err) fails the step: scan no
  // cells rather than all n particles, and let the epilogue write finite values
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = g.err ? lo_i : min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const float cellf = (float)g.cell;
  const float2 pi = a.pos32[i];
  const Me me = make_me(c, nb_from_rec(a.rec32[i]));
  const float rl = (float)(2.0 * a.h[i] + g.skin);
  const float rl2 = rl * rl * 1.0001f;
  Sums S;
  int nb = 0;
  uint32_t deg = 0;
  auto flush = [&]() {
    uint32_t en = nb > 0 ? my[0] : 0u;
    float2 pn = a.pos32[en & kListMask];
    PairRec32 rn = a.rec32[en & kListMask];
    for (int k = 0; k < nb; k++) {
      const uint32_t e = en;
      const float2 pj = pn;
      const PairRec32 rj = rn;
      en = my[min(k + 1, nb - 1) * kBlock];
      pn = a.pos32[en & kListMask];
      rn = a.rec32[en & kListMask];
      const float xij = (pi.x - pj.x) + (float)((int)((e >> 28) & 3u) - 1) * cellf;
      const float yij = (pi.y - pj.y) + (float)((int)(e >> 30) - 1) * cellf;
      pair_terms(me, nb_from_rec(rj), xij, yij, xij * xij + yij * yij, S);
    }
    nb = 0;
  };
  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny + lo_j;
    const uint32_t nc = hi_j - lo_j;  // 1..3 consecutive cells
    const uint32_t b0 = a.cell_start[rb], b3 = a.cell_start[rb + nc];
    const uint32_t b1 = nc > 1 ? a.cell_start[rb + 1] : 0xFFFFFFFFu;
    const uint32_t b2 = nc > 2 ? a.cell_start[rb + 2] : 0xFFFFFFFFu;
    // about a fifth of a column's candidates lie inside the cutoff
    if (__any(nb + ((b3 - b0 + 2) >> 2) + kScanBatch > (uint32_t)kNbCap)) flush();
    const float offx = (float)((int)gi - (int)ni) * cellf;
    const uint32_t tx = (gi - ni + 1) << 28;
    for (uint32_t j0 = b0; j0 < b3; j0 += kScanBatch) {
      if (__any(nb > kNbCap - kScanBatch)) flush();  // guard: room for the batch, taken together
      float2 pj[kScanBatch];
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) pj[e] = a.pos32[min(j0 + e, b3 - 1)];
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) {
        const uint32_t j = j0 + e;
        if (j < b3 && j != i) {
          const uint32_t nj = lo_j + (j >= b1 ? 1u : 0u) + (j >= b2 ? 1u : 0u);
          const float xij = (pi.x - pj[e].x) + offx;
          const float yij = (pi.y - pj[e].y) + (float)((int)gj - (int)nj) * cellf;
          const float r2 = xij * xij + yij * yij;
          if (r2 <= rl2) {
            const uint32_t word = j | tx | ((gj - nj + 1) << 30);
            if (deg < a.nl_cap) a.nl_col[list_at(a.n, i, deg)] = word;
            deg++;
            if (r2 <= me.cut2) my[nb++ * kBlock] = word;
          }
        }
      }
    }
  }
  flush();
  if (deg > a.nl_cap) {
    raise_err(a.err, a.berr, mt.y, EB_NBH);
    deg = a.nl_cap;
  }
  a.nl_deg[i] = deg;
  const double v[12] = {acc_val(S.Sxx_x), acc_val(S.Sxy_y), acc_val(S.Sxy_x), acc_val(S.Syy_y),
                        acc_val(S.vx_x),  acc_val(S.vx_y),  acc_val(S.vy_x),  acc_val(S.vy_y),
                        acc_val(S.ax),    acc_val(S.ay),    acc_val(S.xs),    acc_val(S.ys)};
  finish_list<1, false>(a, c, g, mt.x, i, 0, v, EpiIn{});
}

__global__ __launch_bounds__(kBlock) void k_ul_list_first32(UlArgs a) {
  __shared__ uint32_t nb_lds[kNbCap * kBlock];
  const uint32_t i = xcd_tile_chunked(a.xcd_chunk) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  list_first_global(a, i, nb_lds + threadIdx.x);
}

// Stage 1 with the candidates staged in LDS: the workgroup's bands (as for
// the list stages) hold every particle of the 3x3 cell blocks its particles
// scan, so the scan reads positions from LDS and the pairs inside the cutoff
// are evaluated from LDS.  Candidates are visited, listed and summed in the
// global path's order: the same list and the same bits.  The list is written
// as band-slot words (nl16), which the staged list stages read; a workgroup
// whose bands do not fit takes the global path and its 32-bit words, in every
// stage of the step alike (the bands depend on the step's binning only).
static_assert(sizeof(BandLds) >= sizeof(uint32_t) * kNbCap * kBlock, "global path buffer aliases the bands");

__global__ __launch_bounds__(kBlock, 2) void k_ul_list_first_lds32(UlArgs a) {
  __shared__ BandLds s;
  __shared__ uint16_t s_nb[kNbCap * kBlock];
  __shared__ uint32_t s_base[3], s_end[3];
  __shared__ int32_t s_tot;
  const uint32_t i0 = xcd_tile_chunked(a.xcd_chunk) * kBlock;
  if (i0 >= a.n) return;  // the whole workgroup
  if (threadIdx.x == 0) s_tot = band_setup(a, i0, s_base, s_end);
  __syncthreads();
  const int32_t tot = s_tot;
  const uint32_t i = i0 + threadIdx.x;
  const bool live = tot >= 0 && i < a.n;
  // issued first, to arrive while the bands are staged: the epilogue's state
  // and the cell starts of the (up to) three rows of cells the scan visits
  EpiIn ein;
  uint2 mt = make_uint2(0u, 0u);
  if (live) mt = a.meta[i];
  const SceneConst* __restrict__ c = a.sc + mt.y;
  const SceneGeom g = a.geom[mt.y];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = g.err ? lo_i : min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const uint32_t nc = hi_j - lo_j;  // 1..3 consecutive cells per row
  uint32_t cb[3][4];                // per row: starts of its cells (absent: ~0) and its end
  if (live) {
    ein = epi_load<1>(a, i, 0);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const uint32_t ni = lo_i + r;
      const uint32_t rb = c->cell_offset + ni * g.ny + lo_j;
      const bool row = ni < hi_i;
      cb[r][0] = row ? a.cell_start[rb] : 0u;
      cb[r][1] = row && nc > 1 ? a.cell_start[rb + 1] : 0xFFFFFFFFu;
      cb[r][2] = row && nc > 2 ? a.cell_start[rb + 2] : 0xFFFFFFFFu;
      cb[r][3] = row ? a.cell_start[rb + nc] : 0u;
    }
  }
  band_fill(a, s, tot, s_base, s_end);
  __syncthreads();  // the last barrier: threads past the end leave below
  if (i >= a.n) return;
  if (tot < 0) {
    list_first_global(a, i, reinterpret_cast<uint32_t*>(&s) + threadIdx.x);
    return;
  }
  uint16_t* const my = s_nb + threadIdx.x;
  const float cellf = (float)g.cell;
  const uint32_t li = i + s_base[1];  // the particle's own cell lies in band 1
  const float pix = s.px[li], piy = s.py[li];
  const Me me = make_me(c, nb_from_lds(s, li));
  const float rl = (float)(2.0 * ein.h + g.skin);
  const float rl2 = rl * rl * 1.0001f;
  Sums S;
  int nb = 0;
  uint32_t deg = 0;
  auto flush = [&]() {
    eval_lds(s, my, nb, me, pix, piy, cellf, S);
    nb = 0;
  };
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const uint32_t ni = lo_i + r;  // rows past hi_i have b0 = b3 = 0: nothing to scan
    const uint32_t b0 = cb[r][0], b1 = cb[r][1], b2 = cb[r][2], b3 = cb[r][3];
    if (__any(nb + ((b3 - b0 + 2) >> 2) + kScanBatch > (uint32_t)kNbCap)) flush();
    const float offx = (float)((int)gi - (int)ni) * cellf;
    const uint32_t tx = gi - ni + 1;
    const uint32_t base = tx == 0 ? s_base[0] : tx == 1 ? s_base[1] : s_base[2];
    for (uint32_t j0 = b0; j0 < b3; j0 += kScanBatch) {
      // guard: room for the batch; the lanes still scanning flush together,
      // not each on its own with the others idle
      if (__any(nb > kNbCap - kScanBatch)) flush();
      uint32_t sl[kScanBatch];
      float pxj[kScanBatch], pyj[kScanBatch];
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) {
        sl[e] = (min(j0 + e, b3 - 1) + base) & (kBandCap - 1);
        pxj[e] = s.px[sl[e]];
        pyj[e] = s.py[sl[e]];
      }
      // predicated rather than branched: the separation of every slot of the
      // batch is computed, the compaction buffer is written at nb for every
      // slot (nb <= kNbCap - 1 here) and nb, deg advance by the predicates
#pragma unroll
      for (int e = 0; e < kScanBatch; e++) {
        const uint32_t j = j0 + e;
        const uint32_t nj = lo_j + (j >= b1 ? 1u : 0u) + (j >= b2 ? 1u : 0u);
        const float xij = (pix - pxj[e]) + offx;
        const float yij = (piy - pyj[e]) + (float)((int)gj - (int)nj) * cellf;
        const float r2 = xij * xij + yij * yij;
        const bool listed = j < b3 && j != i && r2 <= rl2;
        const uint32_t ty = gj - nj + 1;
        const uint16_t w = (uint16_t)(sl[e] | (tx << 10) | (ty << 12));
        // entries past the capacity overwrite the last slot: the step fails (EB_NBH)
        if (listed) a.nl16[list16_at(a.n, i, min(deg, a.nl_cap - 1))] = w;
        deg += listed ? 1u : 0u;
        my[nb * kBlock] = w;
        nb += (listed && r2 <= me.cut2) ? 1 : 0;
      }
    }
  }
  flush();
  if (deg > a.nl_cap) {
    raise_err(a.err, a.berr, mt.y, EB_NBH);
    deg = a.nl_cap;
  }
  a.nl_deg[i] = deg;
  const double v[12] = {acc_val(S.Sxx_x), acc_val(S.Sxy_y), acc_val(S.Sxy_x), acc_val(S.Syy_y),
                        acc_val(S.vx_x),  acc_val(S.vx_y),  acc_val(S.vy_x),  acc_val(S.vy_y),
                        acc_val(S.ax),    acc_val(S.ay),    acc_val(S.xs),    acc_val(S.ys)};
  finish_list<1, true>(a, c, g, mt.x, i, 0, v, ein);
}

void launch_ul_list_first(hipStream_t st, const UlArgs& a, bool staged) {
  const dim3 grid(xcd_grid((a.n + kBlock - 1) / kBlock));
  if (staged)
    hipLaunchKernelGGL(k_ul_list_first_lds32, grid, dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL(k_ul_list_first32, grid, dim3(kBlock), 0, st, a);
}

void launch_ul_list(hipStream_t st, const UlArgs& a, int lanes, bool staged) {
  const uint32_t per = kBlock / (uint32_t)lanes;
  const dim3 grid(xcd_grid((a.n + per - 1) / per));
  if (lanes == 1 && staged)
    hipLaunchKernelGGL(k_ul_list_lds32, grid, dim3(kBlock), 0, st, a);
  else if (lanes == 4)
    hipLaunchKernelGGL(k_ul_list32<4>, grid, dim3(kBlock), 0, st, a);
  else if (lanes == 2)
    hipLaunchKernelGGL(k_ul_list32<2>, grid, dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL(k_ul_list32<1>, grid, dim3(kBlock), 0, st, a);
}

void launch_ul_pairs32(hipStream_t st, const UlArgs& a) {
  const dim3 grid(xcd_grid((a.n + kBlock - 1) / kBlock));
  if (a.tile_cells && a.tile_off && !a.gcol) {
    hipLaunchKernelGGL(k_ul_tiles32, dim3(a.tile_grid), dim3(kBlock), 0, st, a);
  } else if (a.gcol) {
    hipLaunchKernelGGL(k_ul_pairs32<true>, grid, dim3(kBlock), 0, st, a);
  } else {
    hipLaunchKernelGGL(k_ul_pairs32<false>, grid, dim3(kBlock), 0, st, a);
  }
}

}  // namespace mfree
