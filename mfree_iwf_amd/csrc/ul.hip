// Gray-Monaghan updated-Lagrangian SPH stage on gfx950.
//
// One derivative evaluation of the reference (src/body.h:125-148 with the
// action list of src/benchmarks/rings.cpp:86-96) is three steps here:
//
//  k_ul_prepass   per particle, in cell-sorted order: the per-particle
//                 actions that precede the pair sums -- quadrature weight
//                 (src/riemann_integration.h:28-33), EOS (src/material.h:37-64),
//                 Monaghan artificial stress (src/correctors.h:37-81) -- and a
//                 gather of everything a neighbour contributes into one
//                 128-byte record per particle (coalesced reads in the pair
//                 pass).  At the first stage of a step it also moves the
//                 particles' storage into the new sorted order, so storage
//                 stays spatially coherent.
//  k_ul_pairs     per particle: ONE pass over the 3x3 cell block evaluating
//                 all four pair sums of the reference at once --
//                 derive_stress_monaghan (src/derivatives.h:378-443),
//                 derive_velocity (src/derivatives.h:32-66), artificial
//                 viscosity (src/correctors.h:83-137) and XSPH
//                 (src/correctors.h:139-169) -- then the per-particle
//                 epilogue (Jaumann rate src/material.h:66-100, continuity
//                 src/cont_mech.h:108-126, momentum :32-49, advection
//                 :134-142) and the RK4 stage update (src/rk4.h:52-108).
//
// Each accumulator is summed over neighbours in the reference's order
// (cells ni-major, nj-minor, particles within a cell in idx order), and the
// FP64 path is compiled with -ffp-contract=off, so its arithmetic follows
// the reference's expression order.  Pairs with q >= 2 (or j == i) add an
// exact +-0 to every sum in the reference (W and grad W vanish, and the self
// pair has v_ij = 0, grad W = 0), so they are skipped without changing a bit.
#include "kernels.h"

namespace mfree {

namespace {

// glm-ordered Jaumann stress rate (src/material.h:71-78): returns
// S_t[0][0], S_t[0][1], S_t[1][1] of 2G(epsdot - 1/3 tr I) - S*omega + omega*S
__device__ inline void jaumann(double G2, double vx_x, double vx_y, double vy_x, double vy_y, double Sxx,
                               double Sxy, double Syy, double& t00, double& t01, double& t11) {
  const double e01 = 0.5 * (vx_y + vy_x);
  const double w10 = 0.5 * (vy_x - vx_y);  // omega[0][1]
  const double w01 = 0.5 * (vx_y - vy_x);  // omega[1][0]
  const double tr = vx_x + vy_y;
  const double s3 = 1. / 3. * tr;
  const double D00 = vx_x - s3 * 1.;
  const double D01 = e01 - s3 * 0.;
  const double D11 = vy_y - s3 * 1.;
  const double A00 = G2 * D00, A01 = G2 * D01, A11 = G2 * D11;
  const double SW00 = Sxx * 0. + Sxy * w10;
  const double SW01 = Sxy * 0. + Syy * w10;
  const double SW11 = Sxy * w01 + Syy * 0.;
  const double WS00 = 0. * Sxx + w01 * Sxy;
  const double WS01 = w10 * Sxx + 0. * Sxy;
  const double WS11 = w10 * Sxy + 0. * Syy;
  t00 = (A00 - SW00) + WS00;
  t01 = (A01 - SW01) + WS01;
  t11 = (A11 - SW11) + WS11;
}

// Monaghan artificial stress of one particle (src/correctors.h:37-81).
__device__ inline void art_stress(double sxx, double sxy, double syy, double rhoi, double eps, double& Rxx,
                                  double& Rxy, double& Ryy) {
  const double numer = 2. * sxy;
  const double denom = sxx - syy + 0.;
  const double theta = 0.5 * atan2(numer, denom);
  const double ct = cos(theta);
  const double sn = sin(theta);
  const double ct2 = ct * ct;
  const double sn2 = sn * sn;
  const double rot_sxx = ct2 * sxx + 2.0 * ct * sn * sxy + sn2 * syy;
  const double rot_syy = sn2 * sxx - 2.0 * ct * sn * sxy + ct2 * syy;
  double rot_rxx = 0., rot_ryy = 0.;
  if (rot_sxx > 0) rot_rxx = -eps * rot_sxx / (rhoi * rhoi);
  if (rot_syy > 0) rot_ryy = -eps * rot_syy / (rhoi * rhoi);
  Rxx = ct2 * rot_rxx + sn2 * rot_ryy;
  Rxy = ct * sn * (rot_rxx - rot_ryy);
  Ryy = sn2 * rot_rxx + ct2 * rot_ryy;
}

__device__ inline bool finite8(const double* d) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(d[k]);
  return ok;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_ul_prepass(UlArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t p = a.perm[i];
  const uint32_t s = a.scene[p];
  const SceneConst* __restrict__ c = a.sc + s;
  double v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = ld_nt(&a.src.f[k][p]);
  const double m = ld_nt(&a.m[p]), h = ld_nt(&a.h[p]);
  const double x = v[0], y = v[1], vx = v[2], vy = v[3], rho = v[4], Sxx = v[5], Sxy = v[6], Syy = v[7];
  const double qw = m / rho;
  const double pr = c->c0sq * (rho - c->rho0);
  const double sxx = Sxx - pr, syy = Syy - pr, sxy = Sxy;
  const uint32_t key = a.key[p];
  a.meta[i] = make_uint2(key, s);
  if (!a.mixed) {
    // a Godunov body's record carries p in rho21 and S_yy without the pressure
    const bool gd = a.method == MFREE_METHOD_UL_GODUNOV;
    double Rxx = 0., Rxy = 0., Ryy = 0.;
    if (!gd) art_stress(sxx, sxy, syy, rho, c->mghn_eps, Rxx, Rxy, Ryy);
    const double rho21 = gd ? pr : 1. / (rho * rho);
    const double cs = sqrt(c->K / rho);
    a.pos64[i] = make_double2(x, y);
    PairRec64 r;
    r.vx = vx;
    r.vy = vy;
    r.rho = rho;
    r.rho21 = rho21;
    r.qw = qw;
    r.m = m;
    r.h = h;
    r.c = cs;
    r.sxx = sxx;
    r.sxy = sxy;
    r.syy = gd ? Syy : syy;
    r.Rxx = Rxx;
    r.Rxy = Rxy;
    r.Ryy = Ryy;
    r.pad0 = 0.;
    r.pad1 = 0.;
    a.rec64[i] = r;
  } else {
    const SceneGeom g = a.geom[s];
    const uint32_t local = key - c->cell_offset;
    const uint32_t ix = local / g.ny, iy = local - ix * g.ny;
    const double cx = g.bbmin_x + (double)ix * g.cell;
    const double cy = g.bbmin_y + (double)iy * g.cell;
    float2 pos;
    PairRec32 r;
    make_rec32(c, cx, cy, x, y, vx, vy, rho, Sxx, Sxy, Syy, m, h, pos, r);
    a.pos32[i] = pos;
    a.rec32[i] = r;
  }
  if (a.reorder) {
    if (a.perm_keep) st_nt(&a.perm_keep[i], p);
#pragma unroll
    for (int k = 0; k < 8; k++) st_nt(&a.dst.f[k][i], v[k]);
    st_nt(&a.m_dst[i], m);
    st_nt(&a.h_dst[i], h);
    st_nt(&a.scene_dst[i], s);
    st_nt(&a.gid_dst[i], ld_nt(&a.gid[p]));
  }
}

// Epilogue + RK4 stage update shared by both precisions.  d = derivatives.
__device__ inline void ul_finish(const UlArgs& a, const SceneConst* __restrict__ c, uint32_t i, uint32_t home,
                                 double* d, const double* auxv) {
  if (!finite8(d)) raise_err(a.err, a.berr, (uint32_t)(c - a.sc), EB_NONFINITE);
  if (a.stage == 0) {
    const uint32_t g = a.gid[home];
    double* o = a.d8 + 8 * (size_t)g;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = d[k];
    if (a.aux) {
      double* x = a.aux + 16 * (size_t)g;
#pragma unroll
      for (int k = 0; k < 16; k++) x[k] = auxv[k];
    }
    return;
  }
  rk_stage_update(a, c, home, d);
}

// Neighbour list entries buffered per thread in LDS (strided by the block
// size so that lanes hit distinct banks).
constexpr int kNbCap = 32;
// Candidate positions loaded together per lane in the FP64 scan (8: 1.5 %
// faster than 4 and 3.5 % faster than 2 at 1M particles).
constexpr int kScanBatch64 = 8;

// M = MFREE_METHOD_UL_MONAGHAN: the rings_monaghan action list;
// M = MFREE_METHOD_UL_GODUNOV: the rings_godunov list (src/benchmarks/rings.cpp:151-158),
// derive_velocity, godunov_continuity and godunov_momentum (src/godunov.h:91-189)
// over the same candidate scan.
template <int M>
__global__ __launch_bounds__(kBlock) void k_ul_pairs64(UlArgs a) {
  constexpr bool kGd = M == MFREE_METHOD_UL_GODUNOV;
  __shared__ uint32_t nb_lds[kNbCap * kBlock];
  uint32_t* const my = nb_lds + threadIdx.x;
  const uint32_t i = xcd_tile_chunked(a.xcd_chunk) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint2 mt = a.meta[i];
  const uint32_t s = mt.y;
  const SceneConst* __restrict__ c = a.sc + s;
  const SceneGeom g = a.geom[s];
  const uint32_t local = mt.x - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  // a body whose grid fell back to one cell (g.err) fails the step: scan no
  // cells rather than all n particles, and let the epilogue write finite values
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = g.err ? lo_i : min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);

  const double2 pi = a.pos64[i];
  const PairRec64 ri = a.rec64[i];
  const double hi = ri.h;
  const double h1 = 1. / hi;
  const double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
  const double radius2 = hi * hi * 2 * 2;
  // r2 > cut2 implies q = sqrt(r2)/h >= 2 under any rounding of the
  // reference's evaluation, so the exact test below decides the rest.
  const double cut2 = 4.0000004 * (hi * hi);
  const double wdeltap = c->wdeltap, corr_exp = c->corr_exp;
  const bool use_as = wdeltap > 0;
  const bool pow4 = c->pow4 != 0;
  const double alpha = c->alpha, beta = c->beta, eta = c->eta, xeps = c->xsph_eps;
  const double sxxi_r = ri.sxx * ri.rho21, sxyi_r = ri.sxy * ri.rho21, syyi_r = ri.syy * ri.rho21;

  double Sxx_x = 0., Sxy_y = 0., Sxy_x = 0., Syy_y = 0.;
  double vx_x = 0., vx_y = 0., vy_x = 0., vy_y = 0.;
  double ax = 0., ay = 0., xs = 0., ys = 0.;
  double rho_t = 0.;  // Godunov continuity
  uint32_t cnt = 0;
  int nb = 0;

  // pair terms of one buffered neighbour, in the reference's operation order
  auto pair = [&](uint32_t j, const double2 pj, const PairRec64& rj) {
    const double xij = pi.x - pj.x;
    const double yij = pi.y - pj.y;
    const double r2 = xij * xij + yij * yij;
    const double rij = sqrt(r2);
    const double q = rij * h1;
    if (q >= 2.) return;
    // cubic spline, src/kernel.cpp:20-55
    double w, wx = 0., wy = 0.;
    if (q >= 1.) {
      w = fac * (0.25 * (2 - q) * (2 - q) * (2 - q));
      if (rij > 1e-12) {
        const double der = -0.75 * (2 - q) * (2 - q) * h1 / rij;
        wx = xij * der * fac;
        wy = yij * der * fac;
      }
    } else {
      w = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
      if (rij > 1e-12) {
        const double der = -3.0 * q * (1 - 0.75 * q) * h1 / rij;
        wx = xij * der * fac;
        wy = yij * der * fac;
      }
    }
    if constexpr (kGd) {
      // derive_velocity
      vx_x += (rj.vx - ri.vx) * wx * rj.qw;
      vx_y += (rj.vx - ri.vx) * wy * rj.qw;
      vy_x += (rj.vy - ri.vy) * wx * rj.qw;
      vy_y += (rj.vy - ri.vy) * wy * rj.qw;
      // godunov_continuity with the acoustic Riemann solver (src/godunov.h:41-56,
      // 155-189); p is in rho21, c = sqrt(K/rho)
      const double nx = xij / rij;
      const double ny = yij / rij;
      const double vir = ri.vx * nx + ri.vy * ny;
      const double vjr = rj.vx * nx + rj.vy * ny;
      const double frac = 1. / (ri.rho * ri.c + rj.rho * rj.c);
      const double vstar = frac * (vir * ri.rho * ri.c + vjr * rj.rho * rj.c - ri.rho21 + rj.rho21);
      rho_t += 2 * ri.rho * (vir - vstar) * (wx * nx + wy * ny) * rj.qw;
      // godunov_momentum (src/godunov.h:58-153): R = normalize(x_j - x_i),
      // S = normalize(R rotated by -90 degrees), glm::normalize(v) =
      // v * (1/sqrt(dot(v, v))); the transformation l = glm::mat2x2(R, S) is
      // single precision (its entries are R, S rounded to float); sigma has
      // S_yy without the pressure; SRijstar[1] reads sigma_j twice -- all as
      // the reference computes them.
      // R = x_j - x_i = -x_ij (IEEE subtraction is antisymmetric), so
      // |R|^2 = |S|^2 = r2 and the reference's three square roots of it and
      // its two 1/sqrt are rij and one reciprocal: the same values
      const double Rx = -xij;
      const double Ry = -yij;
      const double Sx = Ry, Sy = -Rx;
      const double iR = 1. / rij;
      const double iS = iR;
      const double Rnx = Rx * iR, Rny = Ry * iR;
      const double Snx = Sx * iS, Sny = Sy * iS;
      const double l00 = (double)(float)Rnx, l01 = (double)(float)Rny;
      const double l10 = (double)(float)Snx, l11 = (double)(float)Sny;
      // sigma * R, then l * (.): (M v)[r] = M[0][r] v.x + M[1][r] v.y
      const double siR0 = ri.sxx * Rnx + ri.sxy * Rny, siR1 = ri.sxy * Rnx + ri.syy * Rny;
      const double sjR0 = rj.sxx * Rnx + rj.sxy * Rny, sjR1 = rj.sxy * Rnx + rj.syy * Rny;
      const double si0 = l00 * siR0 + l10 * siR1;
      const double sj0 = l00 * sjR0 + l10 * sjR1, sj1 = l01 * sjR0 + l11 * sjR1;
      const double vi0 = l00 * ri.vx + l10 * ri.vy, vi1 = l01 * ri.vx + l11 * ri.vy;
      const double vj0 = l00 * rj.vx + l10 * rj.vy, vj1 = l01 * rj.vx + l11 * rj.vy;
      const double rfrac = 1. / (ri.rho * ri.c + rj.rho * rj.c);
      const double SR0 =
          rfrac * (sj0 * ri.rho * ri.c + si0 * rj.rho * rj.c + ri.rho * ri.c * rj.rho * rj.c * (vj0 - vi0));
      const double SR1 =
          rfrac * (sj1 * ri.rho * ri.c + sj1 * rj.rho * rj.c + ri.rho * ri.c * rj.rho * rj.c * (vj1 - vi1));
      // glm::transpose(l) * SR: [r] = l[r][0] SR0 + l[r][1] SR1
      const double SRij0 = l00 * SR0 + l01 * SR1;
      const double SRij1 = l10 * SR0 + l11 * SR1;
      const double mag_r = rij;
      if (fabs(Rx) > 0.) ax += 2. / ri.rho * wx * mag_r / Rx * SRij0 * rj.qw;
      if (fabs(Ry) > 0.) ay += 2. / ri.rho * wy * mag_r / Ry * SRij1 * rj.qw;
      return;
    }
    // derive_stress_monaghan
    double Rxx = 0., Rxy = 0., Ryy = 0.;
    if (use_as) {
      double fab = w / wdeltap;
      fab = pow4 ? (fab * fab) * (fab * fab) : pow(fab, corr_exp);
      Rxx = fab * (ri.Rxx + rj.Rxx);
      Rxy = fab * (ri.Rxy + rj.Rxy);
      Ryy = fab * (ri.Ryy + rj.Ryy);
    }
    const double mj = rj.m;
    Sxx_x += mj * (sxxi_r + rj.sxx * rj.rho21 + Rxx) * wx;
    Sxy_y += mj * (sxyi_r + rj.sxy * rj.rho21 + Rxy) * wy;
    Sxy_x += mj * (sxyi_r + rj.sxy * rj.rho21 + Rxy) * wx;
    Syy_y += mj * (syyi_r + rj.syy * rj.rho21 + Ryy) * wy;
    // derive_velocity
    vx_x += (rj.vx - ri.vx) * wx * rj.qw;
    vx_y += (rj.vx - ri.vx) * wy * rj.qw;
    vy_x += (rj.vy - ri.vy) * wx * rj.qw;
    vy_y += (rj.vy - ri.vy) * wy * rj.qw;
    // correctors_mghn_artificial_viscosity
    const double vijx = ri.vx - rj.vx;
    const double vijy = ri.vy - rj.vy;
    const double vp = vijx * xij + vijy * yij;
    if (!(vp >= 0.)) {
      const double cij = 0.5 * (ri.c + rj.c);
      const double hij = 0.5 * (hi + rj.h);
      const double rhoij = 0.5 * (ri.rho + rj.rho);
      const double muij = (hij * vp) / (r2 + eta * eta * hij * hij);
      const double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
      ax += -mj * piij * wx;
      ay += -mj * piij * wy;
    }
    // correctors_xsph (mass of particle i, src/correctors.h:163)
    const double rhoij = 0.5 * (ri.rho + rj.rho);
    const double fx = -xeps * w * ri.m / rhoij;
    xs += fx * vijx;
    ys += fx * vijy;
  };
  // the buffered neighbours in order; the next one's position and record
  // are loaded while the current one is evaluated (the last load repeats the
  // last entry, in bounds)
  auto flush = [&]() {
    uint32_t jn = nb > 0 ? my[0] : i;
    double2 pn = a.pos64[jn];
    PairRec64 rn = a.rec64[jn];
    for (int k = 0; k < nb; k++) {
      const uint32_t j = jn;
      const double2 pj = pn;
      const PairRec64 rj = rn;
      jn = my[min(k + 1, nb - 1) * kBlock];
      pn = a.pos64[jn];
      rn = a.rec64[jn];
      pair(j, pj, rj);
    }
    nb = 0;
  };

  // scan: the 3x3 block ni-major, nj-minor, particles of a cell in idx
  // order.  Flushes are collective: before each column the lanes flush
  // together if any lane's buffer is likely to overflow in it (half the
  // column's candidates; a flush taken by each lane on its own inside the
  // scan runs once per lane, the others idle, in dense regions); the
  // in-scan flush guards the rare lane that still fills its buffer.
  // Candidates in batches of kScanBatch64 positions per lane (several
  // gathers in flight; the kernel is bound by their latency).
  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny;
    const uint32_t ja = a.cell_start[rb + lo_j], jb = a.cell_start[rb + hi_j];
    if (__any(nb + ((jb - ja + 1) >> 1) + kScanBatch64 > (uint32_t)kNbCap)) flush();
    for (uint32_t j0 = ja; j0 < jb; j0 += kScanBatch64) {
      if (nb > kNbCap - kScanBatch64) flush();  // guard: room for the batch
      double2 pj[kScanBatch64];
#pragma unroll
      for (int e = 0; e < kScanBatch64; e++) pj[e] = a.pos64[min(j0 + e, jb - 1)];
#pragma unroll
      for (int e = 0; e < kScanBatch64; e++) {
        const uint32_t j = j0 + e;
        if (j >= jb) continue;
        const double xij = pi.x - pj[e].x;
        const double yij = pi.y - pj[e].y;
        const double r2 = xij * xij + yij * yij;
        cnt += (r2 <= radius2 + 1e-6) ? 1u : 0u;
        if (j == i || r2 > cut2) continue;
        my[nb++ * kBlock] = j;
      }
    }
  }
  flush();
  if (a.max_nbh && cnt >= a.max_nbh) raise_err(a.err, a.berr, s, EB_NBH);
  if (cnt == 0 && !g.err) raise_err(a.err, a.berr, s, EB_NONEIGH);

  const uint32_t home = a.home_identity ? i : a.perm[i];
  const double rhoi = ri.rho;
  // contact forces (zero unless the penalty contact set them)
  const double fcx = a.fcx ? a.fcx[a.gid[home]] : 0., fcy = a.fcy ? a.fcy[a.gid[home]] : 0.;
  Sxx_x = Sxx_x * rhoi;
  Sxy_y = Sxy_y * rhoi;
  Sxy_x = Sxy_x * rhoi;
  Syy_y = Syy_y * rhoi;
  const double Sxx = a.src.f[5][home], Sxy = a.src.f[6][home], Syy = a.src.f[7][home];
  double t00, t01, t11;
  jaumann(c->G2, vx_x, vx_y, vy_x, vy_y, Sxx, Sxy, Syy, t00, t01, t11);
  double d[8];
  d[0] = xs + ri.vx;
  d[1] = ys + ri.vy;
  if constexpr (kGd) {
    // godunov_momentum adds f_c/m after its pair sum; godunov_continuity
    // assigns rho_t (src/godunov.h:148-152,187)
    d[2] = ax + fcx / ri.m;
    d[3] = ay + fcy / ri.m;
    d[4] = rho_t;
  } else {
    d[2] = ax + (1. / rhoi * (Sxx_x + Sxy_y) + fcx / ri.m);
    d[3] = ay + (1. / rhoi * (Sxy_x + Syy_y) + fcy / ri.m);
    d[4] = 0. - rhoi * (vx_x + vy_y);
  }
  d[5] = 0. + t00;
  d[6] = 0. + t01;
  d[7] = 0. + t11;
  double auxv[16];
  if (a.stage == 0 && a.aux) {
    auxv[0] = ri.qw;
    auxv[1] = c->c0sq * (rhoi - c->rho0);  // zero sums below for Godunov bodies (no R, no stress divergence)
    auxv[2] = ri.Rxx;
    auxv[3] = ri.Rxy;
    auxv[4] = ri.Ryy;
    auxv[5] = Sxx_x;
    auxv[6] = Sxy_y;
    auxv[7] = Sxy_x;
    auxv[8] = Syy_y;
    auxv[9] = vx_x;
    auxv[10] = vx_y;
    auxv[11] = vy_x;
    auxv[12] = vy_y;
    auxv[13] = (double)cnt;
    auxv[14] = 0.;
    auxv[15] = 0.;
  }
  ul_finish(a, c, i, home, d, auxv);
}

// Time advance of every body (src/rk4.h:107, simulation_time::increment_time);
// with per-body error words, a failed (frozen) body keeps its time.
__global__ void k_advance_time(uint32_t nscene, SceneConst* sc, const uint32_t* berr) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nscene && !(berr && berr[s])) sc[s].time = sc[s].time + sc[s].dt;
}

// A failed body of a context that isolates its bodies keeps the state it
// started the step with: the step's storage (sorted at stage 1) takes y_n
// back from the previous storage through the stage-1 permutation.
__global__ __launch_bounds__(kBlock) void k_restore_failed(uint32_t n, const uint32_t* __restrict__ scene_new,
                                                           const uint32_t* __restrict__ berr,
                                                           const uint32_t* __restrict__ perm1, Fields8 y_old,
                                                           Fields8 y_new) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !berr[scene_new[i]]) return;
  const uint32_t p = perm1[i];
#pragma unroll
  for (int k = 0; k < 8; k++) y_new.f[k][i] = y_old.f[k][p];
}

// nsteps additions of dt, one after the other (the same rounding as nsteps
// single advances)
__global__ void k_advance_time_n(uint32_t nscene, SceneConst* sc, uint32_t nsteps) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nscene) return;
  double t = sc[s].time;
  for (uint32_t k = 0; k < nsteps; k++) t = t + sc[s].dt;
  sc[s].time = t;
}

// The reference's Verlet list of one body (src/grid.h:117-177): count pass
// (fill = 0) and fill pass (fill = 1), in idx space.
__global__ __launch_bounds__(kBlock) void k_pair_list(PairListArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const uint32_t i = a.first + t;  // sorted position
  const uint32_t p = a.perm[i];
  const SceneConst* __restrict__ c = a.sc + a.scene;
  const SceneGeom g = a.geom[a.scene];
  const uint32_t local = a.key[p] - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  const int low_i = (int)gi - 1 < 0 ? 0 : (int)gi - 1;
  const int low_j = (int)gj - 1 < 0 ? 0 : (int)gj - 1;
  const int high_i = gi + 2 > g.nx ? g.nx : gi + 2;
  const int high_j = gj + 2 > g.ny ? g.ny : gj + 2;
  const double hi = a.h[p], xi = a.x[p], yi = a.y[p];
  const double radius2 = hi * hi * 2 * 2;
  const uint32_t row = a.gid[p] - c->offset;
  uint32_t o = a.fill ? a.row_ptr[row] : 0, cnt = 0;
  for (int ni = low_i; ni < high_i; ni++) {
    for (int nj = low_j; nj < high_j; nj++) {
      const uint32_t cell = c->cell_offset + (uint32_t)ni * g.ny + (uint32_t)nj;
      for (uint32_t j = a.cell_start[cell]; j < a.cell_start[cell + 1]; j++) {
        const uint32_t q = a.perm[j];
        const double xij = xi - a.x[q];
        const double yij = yi - a.y[q];
        const double r2 = xij * xij + yij * yij;
        if (r2 <= radius2 + 1e-6) {
          if (a.fill) a.col[o + cnt] = a.gid[q] - c->offset;
          cnt++;
        }
      }
    }
  }
  if (!a.fill) a.row_count[row] = cnt;
}

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_ul_prepass(hipStream_t st, const UlArgs& a) {
  hipLaunchKernelGGL(k_ul_prepass, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a);
}

void launch_ul_pairs32(hipStream_t st, const UlArgs& a);

void launch_ul_pairs(hipStream_t st, const UlArgs& a) {
  if (a.mixed)
    launch_ul_pairs32(st, a);
  else if (a.method == MFREE_METHOD_UL_GODUNOV)
    hipLaunchKernelGGL(k_ul_pairs64<MFREE_METHOD_UL_GODUNOV>, dim3(xcd_grid(nblk(a.n, kBlock))), dim3(kBlock), 0,
                       st, a);
  else
    hipLaunchKernelGGL(k_ul_pairs64<MFREE_METHOD_UL_MONAGHAN>, dim3(xcd_grid(nblk(a.n, kBlock))), dim3(kBlock), 0,
                       st, a);
}

void launch_pair_list(hipStream_t st, const PairListArgs& a) {
  hipLaunchKernelGGL(k_pair_list, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a);
}

}  // namespace mfree
