// Launch interfaces of the gfx950 kernels (internal to the library).
#pragma once

#include "internal.h"

namespace mfree {

// The 8 integrated fields of the reference's RK4 (src/rk4.h:52-69):
// x, y, vx, vy, rho, Sxx, Sxy, Syy.
struct Fields8 {
  double* f[8];
};

struct BinArgs {
  uint32_t n, nscene, nchunks, ncell_cap;
  int deterministic;
  double kernel_width;          // cell side = max(h_max * kernel_width, cell_min)
  double cell_min;
  const Chunk* chunks;
  const uint32_t* chunk_first;  // nscene+1
  const SceneConst* sc;
  SceneGeom* geom;
  const double *x, *y, *h;      // storage order
  const uint32_t* scene;        // storage order
  const uint32_t* gid;          // storage order
  double* part;                 // 5 per chunk
  uint32_t *key, *rank, *count, *tsum, *cell_start, *perm;
  uint32_t* perm_tmp = nullptr;  // deterministic binning: 3n words (scatter order, ids and keys at the scatter
                                 // slots), rank-sorted into perm (null: in-place sort)
  uint32_t* err;
  uint32_t* berr = nullptr;     // per-body error words (null: context word only)
  uint32_t tile_cells;          // > 0: also enumerate pair tiles of this many cells
  uint32_t* tile_off;           // nscene+1: first tile of each body, [nscene] = total
  // neighbour-list binning (once per step): cells of side 2 h_max + skin with
  // skin = max(skin_min h_max, skin_vfac dt max|v|) per body; vx, vy null:
  // cells of side kernel_width h_max (per-stage binning)
  const double* vx = nullptr;
  const double* vy = nullptr;
  double* vpart = nullptr;      // max |v|^2 per chunk
  double skin_min = 0.0, skin_vfac = 0.0;
  // non-null: the body's time advances by its dt (the step that bins here
  // ends by then; nothing of the step reads the time)
  SceneConst* advance = nullptr;
};

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint32_t n);
void launch_binning(hipStream_t st, const BinArgs& a);
// Exclusive scan of n values: out[i] = sum_{k<i} in[k], out[n] = total;
// tsum holds (n + kScanTile - 1) / kScanTile entries.
void launch_excl_scan(hipStream_t st, const uint32_t* in, uint32_t n, uint32_t* tsum, uint32_t* out);

// UL (Gray-Monaghan) stage.
struct UlArgs {
  uint32_t n;
  int mixed;                    // 0: FP64 pair records, 1: FP32 records
  int method;                   // MFREE_METHOD_UL_MONAGHAN or MFREE_METHOD_UL_GODUNOV (FP64 only)
  // contact forces by global id (src/particle.h:57, set between steps by the
  // penalty contact); null: zero
  const double *fcx, *fcy;
  const SceneConst* sc;
  const SceneGeom* geom;
  const uint32_t* perm;         // sorted position -> storage index
  const uint32_t* key;          // storage order (cell key)
  const uint32_t* cell_start;
  // stage state in storage order (the y_n at stage 1, K otherwise)
  Fields8 src;
  const double *m, *h;
  const uint32_t *scene, *gid;
  // reorder (stage 1): copy y_n and constants into sorted storage
  int reorder;
  Fields8 dst;
  double *m_dst, *h_dst;
  uint32_t *scene_dst, *gid_dst;
  // sorted per-stage records
  uint2* meta;                  // {cell key, body}
  double2* pos64;
  PairRec64* rec64;
  float2* pos32;
  PairRec32* rec32;
  // outputs
  int stage;                    // 1..4: RK4 stage; 0: derivative evaluation only;
                                // 5: second stage of Kutta's third-order method
                                // (stages 1, 5, 4)
  int rate_out;                 // last stage: also acc <- mean rate (y_new = y + dt acc)
  Fields8 k1s, k2s, k3s;        // stage rates k1 (stage 1), k2 (stage 2 or 5), k3 (stage 3), storage
                                // order; fields with a null pointer are not kept
  int home_identity;            // storage index == sorted position
  Fields8 y;                    // y_n (storage order), updated at stage 4
  Fields8 K;                    // next stage state (storage order)
  Fields8 acc;                  // RK accumulator (storage order)
  double* d8;                   // stage 0: derivatives by global id
  double* aux;                  // stage 0: diagnostics by global id (may be null)
  uint32_t max_nbh;
  uint32_t* err;
  uint32_t* berr;               // per-body error words (null: context word only)
  // isolated bodies: stage 1 keeps its permutation (sorted position -> storage
  // index of y_n) so a failed body can be restored at the end of the step
  uint32_t* perm_keep;
  // learned-step graph emitted by the mixed per-particle pass (stage 1 of
  // the learned step's base step, bodies of uniform h): edges r < 2h over
  // sorted positions
  uint32_t* gcol;               // slot-major edge list (null: no graph)
  uint32_t* gdeg;
  float* gwsum;
  const double4* gdpos;         // FP64 (x, y, h) by sorted position (predicate near the cutoff)
  uint32_t gstride;
  // LDS-tiled pair pass (mixed precision)
  uint32_t nscene;
  uint32_t tile_cells;          // 0: per-particle kernel without tiles
  const uint32_t* tile_off;
  uint32_t tile_grid;           // workgroups of the tile kernel (multiple of 8)
  uint32_t xcd_chunk;           // per-particle pair kernels: tiles per XCD chunk (xcd_tile_chunked)
  // neighbour-list path (mixed-precision RK4 steps): one binning and one list
  // per step, cells of side 2 h_max + skin; stages read the list
  uint32_t* nl_col;             // groups of 4 entries, group-major: entry k of particle i at
                                // nl_col[((k / 4) * n + i) * 4 + k % 4],
                                // sorted position | (cell dx + 1) << 28 | (cell dy + 1) << 30
  uint32_t* nl_deg;             // entries per particle
  uint32_t nl_cap;              // slots per particle
  uint16_t* nl16;               // workgroups whose candidates fit the LDS bands (staged stages):
                                // the list as band-slot words slot | (dx + 1) << 10 | (dy + 1) << 12,
                                // groups of 8, group-major: entry k at nl16[((k / 8) * n + i) * 8 + k % 8]
  float2* pos32_out;            // records of the next stage (stages 1-3)
  PairRec32* rec32_out;
};

// Monaghan's artificial stress (src/correctors.h:37-81) without
// transcendentals: with d = (sxx - syy)/2 and r = sqrt(d^2 + sxy^2),
// theta = atan2(2 sxy, sxx - syy)/2 has cos 2theta = d/r and sin 2theta = sxy/r
// (cos 2theta = 1 for r = 0, as atan2(0, 0) = 0), so cos^2 = (1 + cos 2theta)/2,
// sin^2 = (1 - cos 2theta)/2, cos sin = sin 2theta/2, and the rotated normal
// stresses are the eigenvalues (sxx + syy)/2 +- r.  Used where the pair terms
// read R in FP32 (mixed precision).
__device__ inline void art_stress_alg(double sxx, double sxy, double syy, double rhoi, double eps, double& Rxx,
                                      double& Rxy, double& Ryy) {
  const double d = 0.5 * (sxx - syy);
  const double r = sqrt(d * d + sxy * sxy);
  double c2t = 1.0, s2t = 0.0;
  if (r > 0.0) {
    const double ir = 1.0 / r;
    c2t = d * ir;
    s2t = sxy * ir;
  }
  const double ct2 = 0.5 * (1.0 + c2t), sn2 = 0.5 * (1.0 - c2t), cs = 0.5 * s2t;
  const double mean = 0.5 * (sxx + syy);
  const double rot_sxx = mean + r, rot_syy = mean - r;
  const double k = -eps / (rhoi * rhoi);
  const double rot_rxx = rot_sxx > 0 ? k * rot_sxx : 0.0;
  const double rot_ryy = rot_syy > 0 ? k * rot_syy : 0.0;
  Rxx = ct2 * rot_rxx + sn2 * rot_ryy;
  Rxy = cs * (rot_rxx - rot_ryy);
  Ryy = sn2 * rot_rxx + ct2 * rot_ryy;
}

// The mixed-precision pair record of a particle from its FP64 stage state:
// position relative to the corner of its cell (subtracted in FP64), and the
// per-particle actions that precede the pair sums -- quadrature weight
// (src/riemann_integration.h:28-33), EOS (src/material.h:37-64), artificial
// stress -- in FP32.
__device__ inline void make_rec32(const SceneConst* __restrict__ c, double cx, double cy, double x, double y,
                                  double vx, double vy, double rho, double Sxx, double Sxy, double Syy, double m,
                                  double h, float2& pos, PairRec32& r) {
  const double qw = m / rho;
  const double pr = c->c0sq * (rho - c->rho0);
  const double sxx = Sxx - pr, syy = Syy - pr, sxy = Sxy;
  double Rxx, Rxy, Ryy;
  art_stress_alg(sxx, sxy, syy, rho, c->mghn_eps, Rxx, Rxy, Ryy);
  const double rho21 = 1. / (rho * rho);
  pos = make_float2((float)(x - cx), (float)(y - cy));
  r.vx = (float)vx;
  r.vy = (float)vy;
  r.rho = (float)rho;
  r.pad2 = 0.f;
  r.qw = (float)qw;
  r.m = (float)m;
  r.h = (float)h;
  r.c = (float)sqrt(c->K / rho);
  r.sxx_r = (float)(sxx * rho21);
  r.sxy_r = (float)(sxy * rho21);
  r.syy_r = (float)(syy * rho21);
  r.pad0 = 0.f;
  r.Rxx = (float)Rxx;
  r.Rxy = (float)Rxy;
  r.Ryy = (float)Ryy;
  r.pad1 = 0.f;
}

// RK stage update of storage index home from its derivatives d.  RK4
// (src/rk4.h:71-108) runs stages 1-4; Kutta's third-order method runs
// stages 1, 5, 4: k2 at y + dt/2 k1, k3 at y - dt k1 + 2 dt k2,
// y_new = y + dt/6 (k1 + 4 k2 + k3).  acc holds the weighted sum of the
// rates so far.
// Streaming accesses: state read and written once per stage bypasses the
// caches, which then keep the data that neighbour loads re-read.
template <class T>
__device__ inline T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ inline void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }

__device__ inline void rk_stage_update(const UlArgs& a, const SceneConst* __restrict__ c, uint32_t home,
                                       const double* d) {
  if (a.stage == 1) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = ld_nt(&a.y.f[k][home]);
      st_nt(&a.acc.f[k][home], d[k]);
      st_nt(&a.K.f[k][home], yk + d[k] * c->half_dt);
      if (a.k1s.f[k]) a.k1s.f[k][home] = d[k];
    }
  } else if (a.stage == 2 || a.stage == 3) {
    const double dts = a.stage == 2 ? c->half_dt : c->dt;
    const Fields8& ks = a.stage == 2 ? a.k2s : a.k3s;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = ld_nt(&a.y.f[k][home]);
      st_nt(&a.acc.f[k][home], ld_nt(&a.acc.f[k][home]) + 2 * d[k]);
      st_nt(&a.K.f[k][home], yk + d[k] * dts);
      if (ks.f[k]) ks.f[k][home] = d[k];
    }
  } else if (a.stage == 5) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = a.y.f[k][home], k1 = a.acc.f[k][home];
      a.acc.f[k][home] = k1 + 4 * d[k];
      a.K.f[k][home] = yk + c->dt * (2 * d[k] - k1);
      if (a.k2s.f[k]) a.k2s.f[k][home] = d[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = ld_nt(&a.y.f[k][home]), sum = ld_nt(&a.acc.f[k][home]) + d[k];
      st_nt(&a.y.f[k][home], yk + c->dt6 * sum);
      if (a.rate_out) a.acc.f[k][home] = sum * (1. / 6.);
    }
  }
}

void launch_ul_prepass(hipStream_t st, const UlArgs& a);
void launch_ul_pairs(hipStream_t st, const UlArgs& a);
// Neighbour-list path (after the stage-1 binning and prepass): stage 1,
// which also builds the step's list, and the stages 2-4 from the list with
// `lanes` lanes per particle (stage 1, and the stages with one lane:
// neighbours staged in LDS per workgroup when `staged`).
void launch_ul_list_first(hipStream_t st, const UlArgs& a, bool staged);
void launch_ul_list(hipStream_t st, const UlArgs& a, int lanes, bool staged);

// Verlet-list emission in the reference's order (parity helper).
struct PairListArgs {
  uint32_t n;                   // particles of the body
  uint32_t first;               // first sorted position of the body
  const SceneConst* sc;
  const SceneGeom* geom;
  uint32_t scene;
  const uint32_t* cell_start;
  const uint32_t* perm;
  const uint32_t* key;          // storage order
  const uint32_t* gid;          // storage order
  const double *x, *y, *h;      // storage order
  uint32_t* row_count;          // by local idx
  const uint32_t* row_ptr;      // by local idx (fill pass)
  uint32_t* col;                // local idx of neighbours
  int fill;
};
void launch_pair_list(hipStream_t st, const PairListArgs& a);

__global__ void k_advance_time(uint32_t nscene, SceneConst* sc, const uint32_t* berr);
__global__ void k_restore_failed(uint32_t n, const uint32_t* scene_new, const uint32_t* berr, const uint32_t* perm1,
                                 Fields8 y_old, Fields8 y_new);
__global__ void k_advance_time_n(uint32_t nscene, SceneConst* sc, uint32_t nsteps);

void launch_calib_copy(hipStream_t st, const double* a, double* b, uint64_t n);

// glm-compatible 2x2 algebra (column-major m[col][row]) shared by the TL
// kernels: dmat2x2(a, b, c, d) has columns (a, b) and (c, d).
struct M2 {
  double m[2][2];
};
__device__ inline M2 m2(double a, double b, double c, double d) {
  M2 r;
  r.m[0][0] = a;
  r.m[0][1] = b;
  r.m[1][0] = c;
  r.m[1][1] = d;
  return r;
}
// (A*B)[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1]
__device__ inline M2 m2_mul(const M2& a, const M2& b) {
  M2 r;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int w = 0; w < 2; ++w) r.m[c][w] = a.m[0][w] * b.m[c][0] + a.m[1][w] * b.m[c][1];
  return r;
}
__device__ inline M2 m2_scale(double s, const M2& a) {
  return m2(s * a.m[0][0], s * a.m[0][1], s * a.m[1][0], s * a.m[1][1]);
}
__device__ inline double m2_det(const M2& a) { return a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]; }
// adjugate times the reciprocal of the determinant, taken once
__device__ inline M2 m2_inv(const M2& a) {
  const double ood = 1.0 / (a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]);
  return m2(a.m[1][1] * ood, -a.m[0][1] * ood, -a.m[1][0] * ood, a.m[0][0] * ood);
}

}  // namespace mfree
