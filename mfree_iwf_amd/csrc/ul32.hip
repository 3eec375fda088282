// Mixed-precision Gray-Monaghan pair pass on gfx950.
//
// Same neighbour sums as k_ul_pairs64 (src/derivatives.h:32-66,378-443,
// src/correctors.h:83-169), evaluated in FP32 from the 64-byte records the
// prepass writes:
//  - positions are cell-local (x minus the particle's own cell corner,
//    subtracted in FP64), so x_ij keeps ~1e-11 m resolution instead of the
//    ~1e-8 m of absolute FP32 coordinates;
//  - every pair sum is accumulated with Kahan compensation into a
//    (sum, compensation) pair;
//  - the per-particle epilogue (Jaumann rate, continuity, momentum,
//    advection) and the RK4 state update stay FP64.
// Compiled with FMA contraction enabled (FP32 math only in this file).
#include "kernels.h"

namespace mfree {

namespace {

struct Acc {
  float s, c;
};

// Kahan compensated summation: c carries the negated low-order part lost
// by each addition (Kahan 1965).
__device__ inline void acc_add(Acc& a, float x) {
  const float y = x - a.c;
  const float t = a.s + y;
  a.c = (t - a.s) - y;
  a.s = t;
}

__device__ inline double acc_val(const Acc& a) { return (double)a.s - (double)a.c; }

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

}  // namespace

constexpr int kNbCap = 32;


__global__ __launch_bounds__(kBlock) void k_ul_pairs32(UlArgs a) {
  __shared__ uint32_t nb_lds[kNbCap * kBlock];
  uint32_t* const my = nb_lds + threadIdx.x;
  const uint32_t i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint2 mt = a.meta[i];
  const uint32_t s = mt.y;
  const SceneConst* __restrict__ c = a.sc + s;
  const SceneGeom g = a.geom[s];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const float cellf = (float)g.cell;

  const float2 pi = a.pos32[i];
  const PairRec32 ri = a.rec32[i];
  const float hi = ri.h;
  const float h1 = 1.f / hi;
  const float fac = (float)(10 * (M_1_PI) / 7.0) * h1 * h1;
  const float cut2 = 4.0001f * (hi * hi);
  const float inv_wdp = (float)(1.0 / c->wdeltap);
  const float corr_exp = (float)c->corr_exp;
  const bool use_as = c->wdeltap > 0;
  const bool pow4 = c->pow4 != 0;
  const float alpha = (float)c->alpha, beta = (float)c->beta, eta2 = (float)(c->eta * c->eta);
  const float xeps_mi = (float)c->xsph_eps * ri.m;
  const float sxxi_r = ri.sxx * ri.rho21, sxyi_r = ri.sxy * ri.rho21, syyi_r = ri.syy * ri.rho21;

  Acc Sxx_x{0.f, 0.f}, Sxy_y{0.f, 0.f}, Sxy_x{0.f, 0.f}, Syy_y{0.f, 0.f};
  Acc vx_x{0.f, 0.f}, vx_y{0.f, 0.f}, vy_x{0.f, 0.f}, vy_y{0.f, 0.f};
  Acc ax{0.f, 0.f}, ay{0.f, 0.f}, xs{0.f, 0.f}, ys{0.f, 0.f};
  int nb = 0;

  auto pair = [&](uint32_t e) {
    const uint32_t j = e & 0x0FFFFFFFu;
    const float offx = (float)((int)((e >> 28) & 3u) - 1) * cellf;
    const float offy = (float)((int)(e >> 30) - 1) * cellf;
    const float2 pj = a.pos32[j];
    const float xij = (pi.x - pj.x) + offx;
    const float yij = (pi.y - pj.y) + offy;
    const float r2 = xij * xij + yij * yij;
    const float rij = sqrtf(r2);
    const float q = rij * h1;
    if (q >= 2.f) return;
    float w, der;
    if (q >= 1.f) {
      const float t = 2.f - q;
      w = fac * (0.25f * t * t * t);
      der = -0.75f * t * t * h1;
    } else {
      w = fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
      der = -3.0f * q * (1.f - 0.75f * q) * h1;
    }
    const float gsc = rij > 1e-12f ? der * fac / rij : 0.f;
    const float wx = xij * gsc, wy = yij * gsc;
    const PairRec32 rj = a.rec32[j];
    float Rxx = 0.f, Rxy = 0.f, Ryy = 0.f;
    if (use_as) {
      float fab = w * inv_wdp;
      fab = pow4 ? (fab * fab) * (fab * fab) : powf(fab, corr_exp);
      Rxx = fab * (ri.Rxx + rj.Rxx);
      Rxy = fab * (ri.Rxy + rj.Rxy);
      Ryy = fab * (ri.Ryy + rj.Ryy);
    }
    const float mj = rj.m;
    acc_add(Sxx_x, mj * (sxxi_r + rj.sxx * rj.rho21 + Rxx) * wx);
    acc_add(Sxy_y, mj * (sxyi_r + rj.sxy * rj.rho21 + Rxy) * wy);
    acc_add(Sxy_x, mj * (sxyi_r + rj.sxy * rj.rho21 + Rxy) * wx);
    acc_add(Syy_y, mj * (syyi_r + rj.syy * rj.rho21 + Ryy) * wy);
    const float dvx = rj.vx - ri.vx, dvy = rj.vy - ri.vy;
    acc_add(vx_x, dvx * wx * rj.qw);
    acc_add(vx_y, dvx * wy * rj.qw);
    acc_add(vy_x, dvy * wx * rj.qw);
    acc_add(vy_y, dvy * wy * rj.qw);
    const float vp = -(dvx * xij + dvy * yij);
    if (!(vp >= 0.f)) {
      const float cij = 0.5f * (ri.c + rj.c);
      const float hij = 0.5f * (hi + rj.h);
      const float rhoij = 0.5f * (ri.rho + rj.rho);
      const float muij = (hij * vp) / (r2 + eta2 * hij * hij);
      const float piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
      acc_add(ax, -mj * piij * wx);
      acc_add(ay, -mj * piij * wy);
    }
    const float rhoij = 0.5f * (ri.rho + rj.rho);
    const float fx = -xeps_mi * w / rhoij;
    acc_add(xs, -fx * dvx);
    acc_add(ys, -fx * dvy);
  };
  auto flush = [&]() {
    for (int k = 0; k < nb; k++) pair(my[k * kBlock]);
    nb = 0;
  };

  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny;
    const float offx = (float)((int)gi - (int)ni) * cellf;
    uint32_t jb = a.cell_start[rb + lo_j];
    for (uint32_t nj = lo_j; nj < hi_j; nj++) {
      const uint32_t ja = jb;
      jb = a.cell_start[rb + nj + 1];
      const float offy = (float)((int)gj - (int)nj) * cellf;
      const uint32_t tag = ((gi - ni + 1) << 28) | ((gj - nj + 1) << 30);
      for (uint32_t j = ja; j < jb; j++) {
        const float2 pj = a.pos32[j];
        const float xij = (pi.x - pj.x) + offx;
        const float yij = (pi.y - pj.y) + offy;
        const float r2 = xij * xij + yij * yij;
        if (j == i || r2 > cut2) continue;
        my[nb * kBlock] = j | tag;
        if (++nb == kNbCap) flush();
      }
    }
  }
  flush();

  const uint32_t home = a.home_identity ? i : a.perm[i];
  const double vxi = a.src.f[2][home], vyi = a.src.f[3][home], rhoi = a.src.f[4][home];
  const double Sxx = a.src.f[5][home], Sxy = a.src.f[6][home], Syy = a.src.f[7][home];
  const double dSxx_x = acc_val(Sxx_x) * rhoi, dSxy_y = acc_val(Sxy_y) * rhoi;
  const double dSxy_x = acc_val(Sxy_x) * rhoi, dSyy_y = acc_val(Syy_y) * rhoi;
  const double gvx_x = acc_val(vx_x), gvx_y = acc_val(vx_y), gvy_x = acc_val(vy_x), gvy_y = acc_val(vy_y);
  double t00, t01, t11;
  jaumann64(c->G2, gvx_x, gvx_y, gvy_x, gvy_y, Sxx, Sxy, Syy, t00, t01, t11);
  double d[8];
  d[0] = acc_val(xs) + vxi;
  d[1] = acc_val(ys) + vyi;
  d[2] = acc_val(ax) + 1. / rhoi * (dSxx_x + dSxy_y);
  d[3] = acc_val(ay) + 1. / rhoi * (dSxy_x + dSyy_y);
  d[4] = -rhoi * (gvx_x + gvy_y);
  d[5] = t00;
  d[6] = t01;
  d[7] = t11;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(d[k]);
  if (!ok) atomicOr(a.err, EB_NONFINITE);
  if (a.stage == 0) {
    const uint32_t gid = a.gid[home];
    double* o = a.d8 + 8 * (size_t)gid;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = d[k];
    if (a.aux) {
      double* x = a.aux + 16 * (size_t)gid;
      x[0] = ri.qw;
      x[1] = c->c0sq * (rhoi - c->rho0);
      x[2] = ri.Rxx;
      x[3] = ri.Rxy;
      x[4] = ri.Ryy;
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
  if (a.stage == 1) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = a.y.f[k][home];
      a.acc.f[k][home] = d[k];
      a.K.f[k][home] = yk + d[k] * c->half_dt;
    }
  } else if (a.stage == 2 || a.stage == 3) {
    const double dts = a.stage == 2 ? c->half_dt : c->dt;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = a.y.f[k][home];
      a.acc.f[k][home] = a.acc.f[k][home] + 2 * d[k];
      a.K.f[k][home] = yk + d[k] * dts;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double yk = a.y.f[k][home];
      a.y.f[k][home] = yk + c->dt6 * (a.acc.f[k][home] + d[k]);
    }
  }
}

void launch_ul_pairs32(hipStream_t st, const UlArgs& a) {
  hipLaunchKernelGGL(k_ul_pairs32, dim3(xcd_grid((a.n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a);
}

}  // namespace mfree
