// Total-Lagrangian SPH with Randles-Libersky kernel correction on gfx950.
//
// Set-up (once, like src/benchmarks/tensile.cpp:271-272):
//   neighbour lists of the reference configuration through the cell list
//   (the reference's O(N^2) find_neighbors, src/benchmarks/neighbor_search.h:116-147,
//   gives the same set: predicate r^2 <= 4h^2 + 1e-6, j in idx order); the
//   lists used for computing keep the pairs with q < 2, since the others
//   carry an exactly zero corrected kernel and add +-0 to every sum;
//   k_tl_rl: Shepard sum C, moment matrix B, B^-1 (glm inverse) and the
//   corrected kernel gradients, all FP64 (src/precomp_shape_functions.h:231-290).
// Per RK stage (src/body.h:110-123, actions src/benchmarks/tensile.cpp:276-286):
//   k_tl_pass1: H = grad u (src/derivatives.h:101-149), Fdot (:483-519),
//     L = Fdot F^-1 (src/cont_mech.h:188-203), continuity, EOS, Jaumann rate,
//     P = J F^-1 sigma (src/cont_mech.h:154-169);
//   k_tl_pass2: div P (src/derivatives.h:331-376), artificial viscosity
//     (src/correctors.h:83-137), momentum (src/cont_mech.h:70-87), advection,
//     RK4 stage update; strip boundary conditions (src/benchmarks/tensile.cpp:20-42)
//     act on the stage copies and after the combine, at the step-start time.
// The same two passes carry the other TL action lists (FP64):
//   co-rotated (tensile_corot, src/benchmarks/tensile.cpp:137-211): set-up
//     precomp_sph (plain kernel); pass 1 corotational_anisotropy_tensor
//     (src/corotational.h:32-85), derive_displacement_corotational
//     (src/derivatives.h:202-258), Cauchy-Green stress (src/material.h:102-130),
//     P = J F^-1 sigma with p = 0 (no action of the list sets p); pass 2
//     derive_stress_corotational (src/derivatives.h:521-588), then viscosity;
//   potential (tensile_potential, src/benchmarks/tensile.cpp:291-361): set-up
//     precomp_randles_libersky_symmetric (the reverse pair's kernel, found
//     by index); pass 1 derive_displacement, Green-St Venant stress
//     (src/material.h:132-161), Fe = -2 V F S; pass 2
//     potential_comp_elastic_potential (src/potential.h:33-85), then viscosity.
//   Neither list integrates rho or S (rho_t = S_t = 0); the stage's S is
//   recomputed from H.
// Storage of TL bodies stays in idx order (frozen lists index it directly).
#include <math.h>
#include <string.h>

#include <algorithm>

#include "tl.h"

namespace mfree {

namespace {

struct TlArgs {
  uint32_t n;
  const SceneConst* sc;
  const uint32_t* scene;
  const double *X, *Y;
  const uint8_t* bc;
  const uint32_t* rowp;
  const uint32_t* col;
  const double* w64;
  const float* w32;
  const double* qw;
  const double *m, *h;
  Fields8 src;
  void* rec;  // TlRec64 / TlRec32 per particle, written by pass 1
  double *part8, *diag;
  double2* rot;         // co-rotated: (cos t, sin t) per particle, written by pass 1
  const uint32_t* rev;  // potential: reverse pair of each pair (~0u: none)
  const double *fcx, *fcy;  // contact forces (idx order), or null
  int stage;
  Fields8 y, K, acc;
  double* d8;
  double* aux;
  uint32_t* err;
};

// Position and velocity of a stage particle, with the strip boundary
// conditions applied at time t (src/benchmarks/tensile.cpp:20-42).
__device__ inline void tl_xv(const TlArgs& a, uint32_t j, double t, double& x, double& y, double& vx, double& vy) {
  const uint8_t b = a.bc[j];
  if (b == 0) {
    x = a.src.f[0][j];
    y = a.src.f[1][j];
    vx = a.src.f[2][j];
    vy = a.src.f[3][j];
    return;
  }
  const double vb = a.sc[a.scene[j]].v_bc;
  if (b == 1) {
    x = a.X[j] - vb * t;
    vx = -vb;
  } else {
    x = a.X[j] + vb * t;
    vx = vb;
  }
  y = a.Y[j];
  vy = 0.;
}

struct Sum64 {
  double s = 0.;
  __device__ void add(double x) { s += x; }
  __device__ double val() const { return s; }
};

struct Kahan32 {
  float s = 0.f, c = 0.f;
  __device__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ double val() const { return (double)s - (double)c; }
};

// What pass 2 gathers of a neighbour, written by pass 1 into one record per
// particle: the stage position and velocity with the strip BCs applied, the
// nominal stress P and the viscosity's per-particle factors.  The mixed
// record keeps everything but x and v in FP32 -- pass 2 converts those to
// FP32 at use anyway, so the stored values are the ones it would compute --
// and fits one 64-byte line.
struct __attribute__((aligned(16))) TlRec64 {
  double x, y, vx, vy, Pxx, Pxy, Pyx, Pyy, m, c, h, rho;
};
struct __attribute__((aligned(16))) TlRec32 {
  double x, y, vx, vy;
  float Pxx, Pxy, Pyx, Pyy, m, c, h, rho;
};
static_assert(sizeof(TlRec64) == 96 && sizeof(TlRec32) == 64, "TL record layout");

template <bool MIXED>
struct TlTypes {
  using T = double;
  using A = Sum64;
  using R = TlRec64;
};
template <>
struct TlTypes<true> {
  using T = float;
  using A = Kahan32;
  using R = TlRec32;
};

template <bool MIXED>
__device__ inline void load_w(const TlArgs& a, uint32_t q, typename TlTypes<MIXED>::T* w) {
  if (MIXED) {
    w[0] = a.w32[3 * (size_t)q];
    w[1] = a.w32[3 * (size_t)q + 1];
    w[2] = a.w32[3 * (size_t)q + 2];
  } else {
    w[0] = a.w64[3 * (size_t)q];
    w[1] = a.w64[3 * (size_t)q + 1];
    w[2] = a.w64[3 * (size_t)q + 2];
  }
}

// Pass 1 of a co-rotated body (FP64): rotation, co-rotated displacement
// gradient, Cauchy-Green stress, nominal stress; writes the pass-2 record
// (P), the rotation and the diagnostics.
__device__ inline void tl_pass1_corot(const TlArgs& a, const SceneConst* __restrict__ c, uint32_t i, double t,
                                      double xi, double yi, double vxi, double vyi) {
  const double Xi = a.X[i], Yi = a.Y[i];
  // corotational_anisotropy_tensor
  double Axx = 0., Axy = 0., Ayx = 0., Ayy = 0.;
#pragma unroll 4
  for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
    const uint32_t j = a.col[q];
    const double w0 = a.w64[3 * (size_t)q];
    double xj, yj, vxj, vyj;
    tl_xv(a, j, t, xj, yj, vxj, vyj);
    const double mj = a.m[j];
    const double xji = xj - xi, yji = yj - yi;
    const double Xji = a.X[j] - Xi, Yji = a.Y[j] - Yi;
    Axx += xji * Xji * mj * w0;
    Axy += xji * Yji * mj * w0;
    Ayx += yji * Xji * mj * w0;
    Ayy += yji * Yji * mj * w0;
  }
  const double theta = atan2(Ayx - Axy, Axx + Ayy);
  const double cs = cos(theta), sn = sin(theta);
  const double Rxx = cs, Rxy = -sn, Ryx = sn, Ryy = cs;
  // derive_displacement_corotational with R^-1 = R^T
  const double Rixx_inv = Rxx, Rixy_inv = Ryx, Riyx_inv = Rxy, Riyy_inv = Ryy;
  double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
#pragma unroll 4
  for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
    const uint32_t j = a.col[q];
    const double w1 = a.w64[3 * (size_t)q + 1], w2 = a.w64[3 * (size_t)q + 2];
    double xj, yj, vxj, vyj;
    tl_xv(a, j, t, xj, yj, vxj, vyj);
    const double qw = a.qw[j];
    const double xji = xj - xi, yji = yj - yi;
    const double Xji = a.X[j] - Xi, Yji = a.Y[j] - Yi;
    const double uxji = Rixx_inv * xji + Rixy_inv * yji - Xji;
    const double uyji = Riyx_inv * xji + Riyy_inv * yji - Yji;
    Hxx += uxji * w1 * qw;
    Hxy += uxji * w2 * qw;
    Hyx += uyji * w1 * qw;
    Hyy += uyji * w2 * qw;
  }
  // material_compute_stress_cauchy_green: E = 0.5 (F + F^T) - I
  const double E = c->E, nu = c->nu;
  const double D = E / (1 - nu * nu);
  const M2 F = m2(Hxx + 1., Hyx, Hxy, Hyy + 1.);
  const double e00 = 0.5 * (F.m[0][0] + F.m[0][0]) - 1.;
  const double e01 = 0.5 * (F.m[0][1] + F.m[1][0]) - 0.;
  const double e11 = 0.5 * (F.m[1][1] + F.m[1][1]) - 1.;
  const double Sxx = D * (e00 + nu * e11);
  const double Syy = D * (e00 * nu + e11);
  const double Sxy = D * (1 - nu) * e01;
  // contmech_cauchy_to_nominal, p = 0
  const double p = 0.;
  const M2 F2 = m2(Hxx + 1, Hyx, Hxy, Hyy + 1);
  const double J = m2_det(F2);
  const M2 sigma = m2(Sxx - p, Sxy, Sxy, Syy - p);
  const M2 P = m2_mul(m2_scale(J, m2_inv(F2)), sigma);
  const double rho = a.src.f[4][i];
  TlRec64 r;
  r.x = xi;
  r.y = yi;
  r.vx = vxi;
  r.vy = vyi;
  r.Pxx = P.m[0][0];
  r.Pxy = P.m[1][0];
  r.Pyx = P.m[0][1];
  r.Pyy = P.m[1][1];
  r.m = a.m[i];
  r.c = sqrt(c->K / rho);
  r.h = a.h[i];
  r.rho = rho;
  ((TlRec64*)a.rec)[i] = r;
  a.rot[i] = make_double2(cs, sn);
  if (a.stage == 0 && a.aux) {
    double* dg = a.diag + 16 * (size_t)i;
    const double v[16] = {Hxx, Hxy, Hyx, Hyy, Rxx, Rxy, Ryx, Ryy, r.Pxx, r.Pxy, r.Pyx, r.Pyy, Sxx, Sxy, Syy, 0.};
#pragma unroll
    for (int k = 0; k < 16; k++) dg[k] = v[k];
  }
}

// Pass 1 of a potential-based body (FP64): displacement gradient with the
// Randles-Libersky kernel, Green-St Venant stress, Fe = -2 V F S (stored in
// the record's P slots, same index mapping).
__device__ inline void tl_pass1_potential(const TlArgs& a, const SceneConst* __restrict__ c, uint32_t i, double t,
                                          double xi, double yi, double vxi, double vyi) {
  const double uxi = xi - a.X[i], uyi = yi - a.Y[i];
  double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
#pragma unroll 8
  for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
    const uint32_t j = a.col[q];
    const double w1 = a.w64[3 * (size_t)q + 1], w2 = a.w64[3 * (size_t)q + 2];
    double xj, yj, vxj, vyj;
    tl_xv(a, j, t, xj, yj, vxj, vyj);
    const double uxj = xj - a.X[j], uyj = yj - a.Y[j];
    const double qw = a.qw[j];
    Hxx += (uxj - uxi) * w1 * qw;
    Hxy += (uxj - uxi) * w2 * qw;
    Hyx += (uyj - uyi) * w1 * qw;
    Hyy += (uyj - uyi) * w2 * qw;
  }
  // material_compute_stress_green_stvenant: C = F^T F, E = 0.5 (C - I)
  const double E = c->E, nu = c->nu;
  const double D = E / (1 - nu * nu);
  const M2 F = m2(Hxx + 1., Hyx, Hxy, Hyy + 1.);
  const M2 FT = m2(F.m[0][0], F.m[1][0], F.m[0][1], F.m[1][1]);
  const M2 C = m2_mul(FT, F);
  const double e00 = 0.5 * (C.m[0][0] - 1.), e10 = 0.5 * (C.m[1][0] - 0.), e11 = 0.5 * (C.m[1][1] - 1.);
  const double Sxx = D * (e00 + nu * e11);
  const double Syy = D * (e00 * nu + e11);
  const double Sxy = D * (1 - nu) * e10;
  // Fe = -2*voli*Ji*Si, evaluated ((-2 voli) J) S
  const double voli = a.qw[i];
  const M2 Fe = m2_mul(m2_scale(-2 * voli, m2(Hxx + 1., Hyx, Hxy, Hyy + 1.)), m2(Sxx, Sxy, Sxy, Syy));
  const double rho = a.src.f[4][i];
  TlRec64 r;
  r.x = xi;
  r.y = yi;
  r.vx = vxi;
  r.vy = vyi;
  r.Pxx = Fe.m[0][0];
  r.Pxy = Fe.m[1][0];
  r.Pyx = Fe.m[0][1];
  r.Pyy = Fe.m[1][1];
  r.m = a.m[i];
  r.c = sqrt(c->K / rho);
  r.h = a.h[i];
  r.rho = rho;
  ((TlRec64*)a.rec)[i] = r;
  if (a.stage == 0 && a.aux) {
    double* dg = a.diag + 16 * (size_t)i;
    const double v[16] = {Hxx, Hxy, Hyx, Hyy, r.Pxx, r.Pxy, r.Pyx, r.Pyy, 0., 0., 0., 0., Sxx, Sxy, Syy, 0.};
#pragma unroll
    for (int k = 0; k < 16; k++) dg[k] = v[k];
  }
}

template <bool MIXED, int V>
__global__ __launch_bounds__(kBlock) void k_tl_pass1(TlArgs a) {
  using T = typename TlTypes<MIXED>::T;
  using A = typename TlTypes<MIXED>::A;
  const uint32_t i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const SceneConst* __restrict__ c = a.sc + a.scene[i];
  const double t = c->time;
  double xi, yi, vxi, vyi;
  tl_xv(a, i, t, xi, yi, vxi, vyi);
  if constexpr (V != kTlRl) {
    // no continuity, EOS or stress rate in these lists: rho_t = S_t = 0
    double* pp = a.part8 + 4 * (size_t)i;
    pp[0] = pp[1] = pp[2] = pp[3] = 0.;
    if constexpr (V == kTlCorot)
      tl_pass1_corot(a, c, i, t, xi, yi, vxi, vyi);
    else
      tl_pass1_potential(a, c, i, t, xi, yi, vxi, vyi);
    return;
  }
  const double uxi = xi - a.X[i], uyi = yi - a.Y[i];
  A Hxx, Hxy, Hyx, Hyy, Fxx, Fxy, Fyx, Fyy;
  // unrolled: the next neighbours' loads are issued with the current one's
  // (the sums keep their order); factors measured on the tension plate
  // (12-13 neighbours per row)
#pragma unroll 8
  for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
    const uint32_t j = a.col[q];
    T w[3];
    load_w<MIXED>(a, q, w);
    double xj, yj, vxj, vyj;
    tl_xv(a, j, t, xj, yj, vxj, vyj);
    const double uxj = xj - a.X[j], uyj = yj - a.Y[j];
    const T qw = (T)a.qw[j];
    // derive_displacement and derive_velocity_tl, reference expression order
    Hxx.add((T)(uxj - uxi) * w[1] * qw);
    Hxy.add((T)(uxj - uxi) * w[2] * qw);
    Hyx.add((T)(uyj - uyi) * w[1] * qw);
    Hyy.add((T)(uyj - uyi) * w[2] * qw);
    Fxx.add((T)(vxj - vxi) * w[1] * qw);
    Fxy.add((T)(vxj - vxi) * w[2] * qw);
    Fyx.add((T)(vyj - vyi) * w[1] * qw);
    Fyy.add((T)(vyj - vyi) * w[2] * qw);
  }
  const double hxx = Hxx.val(), hxy = Hxy.val(), hyx = Hyx.val(), hyy = Hyy.val();
  const double fxx = Fxx.val(), fxy = Fxy.val(), fyx = Fyx.val(), fyy = Fyy.val();
  // contmech_velocity_gradient_tl
  const M2 F = m2(hxx + 1., hyx, hxy, hyy + 1.);
  const M2 Fdot = m2(fxx, fyx, fxy, fyy);
  const M2 L = m2_mul(Fdot, m2_inv(F));
  const double vx_x = L.m[0][0], vx_y = L.m[1][0], vy_x = L.m[0][1], vy_y = L.m[1][1];
  const double rho = a.src.f[4][i];
  const double Sxx = a.src.f[5][i], Sxy = a.src.f[6][i], Syy = a.src.f[7][i];
  // contmech_continuity
  const double rho_t = 0. - rho * (vx_x + vy_y);
  // material_eos
  const double p = c->c0sq * (rho - c->rho0);
  // material_stress_rate_jaumann
  const M2 epsdot = m2(vx_x, 0.5 * (vx_y + vy_x), 0.5 * (vx_y + vy_x), vy_y);
  const M2 omega = m2(0., 0.5 * (vy_x - vx_y), 0.5 * (vx_y - vy_x), 0.);
  const M2 S = m2(Sxx, Sxy, Sxy, Syy);
  const double tr = epsdot.m[0][0] + epsdot.m[1][1];
  const double s3 = 1. / 3. * tr;
  const M2 D = m2(epsdot.m[0][0] - s3 * 1., epsdot.m[0][1] - s3 * 0., epsdot.m[1][0] - s3 * 0.,
                  epsdot.m[1][1] - s3 * 1.);
  const M2 A2 = m2_scale(c->G2, D);
  const M2 SW = m2_mul(S, omega), WS = m2_mul(omega, S);
  const double t00 = (A2.m[0][0] - SW.m[0][0]) + WS.m[0][0];
  const double t01 = (A2.m[0][1] - SW.m[0][1]) + WS.m[0][1];
  const double t11 = (A2.m[1][1] - SW.m[1][1]) + WS.m[1][1];
  // contmech_cauchy_to_nominal
  const M2 F2 = m2(hxx + 1, hyx, hxy, hyy + 1);
  const double J = m2_det(F2);
  const M2 sigma = m2(Sxx - p, Sxy, Sxy, Syy - p);
  const M2 P = m2_mul(m2_scale(J, m2_inv(F2)), sigma);
  typename TlTypes<MIXED>::R r;
  r.x = xi;
  r.y = yi;
  r.vx = vxi;
  r.vy = vyi;
  r.Pxx = (T)P.m[0][0];
  r.Pxy = (T)P.m[1][0];
  r.Pyx = (T)P.m[0][1];
  r.Pyy = (T)P.m[1][1];
  r.m = (T)a.m[i];
  r.c = (T)sqrt(c->K / rho);
  r.h = (T)a.h[i];
  r.rho = (T)rho;
  ((typename TlTypes<MIXED>::R*)a.rec)[i] = r;
  double* pp = a.part8 + 4 * (size_t)i;
  pp[0] = rho_t;
  pp[1] = 0. + t00;
  pp[2] = 0. + t01;
  pp[3] = 0. + t11;
  if (a.stage == 0 && a.aux) {
    double* dg = a.diag + 16 * (size_t)i;
    dg[8] = P.m[0][0];
    dg[9] = P.m[1][0];
    dg[10] = P.m[0][1];
    dg[11] = P.m[1][1];
    dg[0] = hxx;
    dg[1] = hxy;
    dg[2] = hyx;
    dg[3] = hyy;
    dg[4] = fxx;
    dg[5] = fxy;
    dg[6] = fyx;
    dg[7] = fyy;
  }
}

// The variants' pair force sums of particle i (FP64), in the reference's
// order: the force action assigns the sum, adds f_c/m (zero), then the
// viscosity adds its pair terms.
template <int V>
__device__ inline void tl_pass2_forces(const TlArgs& a, const SceneConst* __restrict__ c, uint32_t i,
                                       const TlRec64& ri, double& ax, double& ay) {
  const TlRec64* __restrict__ rec = (const TlRec64*)a.rec;
  const double rho0 = c->rho0;
  double vx_t = 0., vy_t = 0.;
  if constexpr (V == kTlCorot) {
    // derive_stress_corotational
    const double2 roti = a.rot[i];
    const double Rxxi = roti.x, Rxyi = -roti.y, Ryxi = roti.y, Ryyi = roti.x;
    const double rhoi21 = 1. / (rho0 * rho0);
#pragma unroll 4
    for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
      const uint32_t j = a.col[q];
      const double w1 = a.w64[3 * (size_t)q + 1], w2 = a.w64[3 * (size_t)q + 2];
      const TlRec64 rj = rec[j];
      const double2 rotj = a.rot[j];
      const double Rxxj = rotj.x, Rxyj = -rotj.y, Ryxj = rotj.y, Ryyj = rotj.x;
      const double mj = rj.m;
      const double rhoj21 = 1. / (rho0 * rho0);
      const double aijx = mj * ((ri.Pxx * rhoi21 + rj.Pxx * rhoj21) * w1 + (ri.Pxy * rhoi21 + rj.Pxy * rhoj21) * w2);
      const double aijy = mj * ((ri.Pyx * rhoi21 + rj.Pyx * rhoj21) * w1 + (ri.Pyy * rhoi21 + rj.Pyy * rhoj21) * w2);
      const double ajix = -aijx, ajiy = -aijy;
      const double aijxprime = aijx * Rxxi + aijy * Rxyi;
      const double aijyprime = aijx * Ryxi + aijy * Ryyi;
      const double ajixprime = ajix * Rxxj + ajiy * Rxyj;
      const double ajiyprime = ajix * Ryxj + ajiy * Ryyj;
      vx_t += 0.5 * (aijxprime - ajixprime);
      vy_t += 0.5 * (aijyprime - ajiyprime);
    }
  } else {
    // potential_comp_elastic_potential: Fe in the P slots, (M v)[r] = M[0][r] v.x + M[1][r] v.y
    const double mi = ri.m;
#pragma unroll 4
    for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
      const uint32_t j = a.col[q];
      const double wij1 = a.w64[3 * (size_t)q + 1], wij2 = a.w64[3 * (size_t)q + 2];
      const uint32_t qr = a.rev[q];
      const double wji1 = qr != ~0u ? a.w64[3 * (size_t)qr + 1] : 0.;
      const double wji2 = qr != ~0u ? a.w64[3 * (size_t)qr + 2] : 0.;
      const TlRec64 rj = rec[j];
      const double fij0 = ri.Pxx * wij1 + ri.Pxy * wij2;
      const double fij1 = ri.Pyx * wij1 + ri.Pyy * wij2;
      const double fji0 = rj.Pxx * wji1 + rj.Pxy * wji2;
      const double fji1 = rj.Pyx * wji1 + rj.Pyy * wji2;
      vx_t += 0.5 * (fji0 - fij0) / mi;
      vy_t += 0.5 * (fji1 - fij1) / mi;
    }
  }
  vx_t += (a.fcx ? a.fcx[i] : 0.) / ri.m;
  vy_t += (a.fcy ? a.fcy[i] : 0.) / ri.m;
  // correctors_mghn_artificial_viscosity
  const double alpha = c->alpha, beta = c->beta, eta = c->eta;
#pragma unroll 4
  for (uint32_t q = a.rowp[i]; q < a.rowp[i + 1]; q++) {
    const uint32_t j = a.col[q];
    const double w1 = a.w64[3 * (size_t)q + 1], w2 = a.w64[3 * (size_t)q + 2];
    const TlRec64 rj = rec[j];
    const double vijx = ri.vx - rj.vx, vijy = ri.vy - rj.vy;
    const double xij = ri.x - rj.x, yij = ri.y - rj.y;
    const double vp = vijx * xij + vijy * yij;
    if (vp >= 0.) continue;
    const double cij = 0.5 * (ri.c + rj.c);
    const double hij = 0.5 * (ri.h + rj.h);
    const double rhoij = 0.5 * (ri.rho + rj.rho);
    const double r2 = xij * xij + yij * yij;
    const double muij = (hij * vp) / (r2 + eta * eta * hij * hij);
    const double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
    vx_t += -rj.m * piij * w1;
    vy_t += -rj.m * piij * w2;
  }
  ax = vx_t;
  ay = vy_t;
}

template <bool MIXED, int V>
__global__ __launch_bounds__(kBlock) void k_tl_pass2(TlArgs a) {
  using T = typename TlTypes<MIXED>::T;
  using A = typename TlTypes<MIXED>::A;
  const uint32_t i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  using R = typename TlTypes<MIXED>::R;
  const SceneConst* __restrict__ c = a.sc + a.scene[i];
  const double t = c->time;
  const double rho0 = c->rho0;
  const R* __restrict__ rec = (const R*)a.rec;
  const R ri = rec[i];
  const T Pxxi = ri.Pxx, Pxyi = ri.Pxy, Pyxi = ri.Pyx, Pyyi = ri.Pyy;
  const T r21 = (T)(1. / (rho0 * rho0));
  const T rho0T = (T)rho0;
  const double xi = ri.x, yi = ri.y, vxi = ri.vx, vyi = ri.vy;
  const T hi = ri.h, rhoi = ri.rho, ci = ri.c;
  const T alpha = (T)c->alpha, beta = (T)c->beta, eta = (T)c->eta;
  A Pxx_x, Pyx_y, Pxy_x, Pyy_y, ax, ay;
  double vx_t = 0., vy_t = 0.;
  if constexpr (V != kTlRl) tl_pass2_forces<V>(a, c, i, ri, vx_t, vy_t);
  constexpr int kUnroll = MIXED ? 16 : 8;
#pragma unroll kUnroll
  for (uint32_t q = a.rowp[i]; q < (V == kTlRl ? a.rowp[i + 1] : 0u); q++) {
    const uint32_t j = a.col[q];
    T w[3];
    load_w<MIXED>(a, q, w);
    const R rj = rec[j];
    const T mj = rj.m;
    // derive_stress_tl
    Pxx_x.add(mj * (Pxxi * r21 + rj.Pxx * r21) * w[1] * rho0T);
    Pyx_y.add(mj * (Pyxi * r21 + rj.Pyx * r21) * w[2] * rho0T);
    Pxy_x.add(mj * (Pxyi * r21 + rj.Pxy * r21) * w[1] * rho0T);
    Pyy_y.add(mj * (Pyyi * r21 + rj.Pyy * r21) * w[2] * rho0T);
    // correctors_mghn_artificial_viscosity
    const T vijx = (T)(vxi - rj.vx), vijy = (T)(vyi - rj.vy);
    const T xij = (T)(xi - rj.x), yij = (T)(yi - rj.y);
    const T vp = vijx * xij + vijy * yij;
    if (!(vp >= (T)0)) {
      const T cij = (T)0.5 * (ci + rj.c);
      const T hij = (T)0.5 * (hi + rj.h);
      const T rhoij = (T)0.5 * (rhoi + rj.rho);
      const T r2 = xij * xij + yij * yij;
      const T muij = (hij * vp) / (r2 + eta * eta * hij * hij);
      const T piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
      ax.add(-mj * piij * w[1]);
      ay.add(-mj * piij * w[2]);
    }
  }
  const double* pp = a.part8 + 4 * (size_t)i;
  const double pxx_x = Pxx_x.val(), pyx_y = Pyx_y.val(), pxy_x = Pxy_x.val(), pyy_y = Pyy_y.val();
  double d[8];
  d[0] = 0. + vxi;  // contmech_advection
  d[1] = 0. + vyi;
  if constexpr (V == kTlRl) {
    d[2] = ax.val() + (1. / rho0 * (pxx_x + pyx_y) + (a.fcx ? a.fcx[i] : 0.) / a.m[i]);  // contmech_momentum_tl
    d[3] = ay.val() + (1. / rho0 * (pxy_x + pyy_y) + (a.fcy ? a.fcy[i] : 0.) / a.m[i]);
  } else {
    d[2] = vx_t;
    d[3] = vy_t;
  }
  d[4] = pp[0];
  d[5] = pp[1];
  d[6] = pp[2];
  d[7] = pp[3];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(d[k]);
  if (!ok) atomicOr(a.err, EB_NONFINITE);
  if (a.stage == 0) {
    double* o = a.d8 + 8 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = d[k];
    if (a.aux) {
      double* x = a.aux + 16 * (size_t)i;
      const double* dg = a.diag + 16 * (size_t)i;
      if constexpr (V == kTlRl) {
#pragma unroll
        for (int k = 0; k < 12; k++) x[k] = dg[k];
        x[12] = pxx_x;
        x[13] = pxy_x;
        x[14] = pyx_y;
        x[15] = pyy_y;
      } else {
#pragma unroll
        for (int k = 0; k < 16; k++) x[k] = dg[k];
      }
    }
    return;
  }
  if (a.stage == 1) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a.acc.f[k][i] = d[k];
      a.K.f[k][i] = a.y.f[k][i] + d[k] * c->half_dt;
    }
  } else if (a.stage == 2 || a.stage == 3) {
    const double dts = a.stage == 2 ? c->half_dt : c->dt;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a.acc.f[k][i] = a.acc.f[k][i] + 2 * d[k];
      a.K.f[k][i] = a.y.f[k][i] + d[k] * dts;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) a.y.f[k][i] = a.y.f[k][i] + c->dt6 * (a.acc.f[k][i] + d[k]);
    // boundary conditions after the combine, at the step-start time (src/rk4.h:105)
    const uint8_t b = a.bc[i];
    if (b) {
      const double vb = c->v_bc;
      a.y.f[0][i] = b == 1 ? a.X[i] - vb * t : a.X[i] + vb * t;
      a.y.f[1][i] = a.Y[i];
      a.y.f[2][i] = b == 1 ? -vb : vb;
      a.y.f[3][i] = 0.;
    }
  }
}

// ---- set-up kernels -------------------------------------------------------

struct SearchArgs {
  uint32_t n;
  const SceneConst* sc;
  const SceneGeom* geom;
  const uint32_t* scene;
  const uint32_t* key;         // storage order
  const uint32_t* cell_start;
  const uint32_t* perm;
  const double *x, *y, *h;
  double h_search;             // > 0: the find_neighbors h; else per-particle h
  int contrib_only;
  uint32_t* count;             // per particle
  const uint32_t* rowp;        // fill pass
  uint32_t* col;
  int fill;
};

// The reference predicate r^2 <= 4 hs^2 + 1e-6 (neighbor_search.h:32-41)
// over the 3x3 block of a cell list whose cell side covers it; with
// contrib_only also q = r/h_i < 2.  Rows are sorted afterwards.
__global__ __launch_bounds__(kBlock) void k_tl_search(SearchArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.scene[i];
  const SceneConst* __restrict__ c = a.sc + s;
  const SceneGeom g = a.geom[s];
  const uint32_t local = a.key[i] - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const double xi = a.x[i], yi = a.y[i], hi = a.h[i];
  const double hs = a.h_search > 0 ? a.h_search : hi;
  const double radius2 = hs * hs * 2 * 2;
  const double h1 = 1. / hi;
  uint32_t cnt = 0;
  const uint32_t o = a.fill ? a.rowp[i] : 0;
  for (uint32_t ni = lo_i; ni < hi_i; ni++) {
    const uint32_t rb = c->cell_offset + ni * g.ny;
    for (uint32_t jj = a.cell_start[rb + lo_j]; jj < a.cell_start[rb + hi_j]; jj++) {
      const uint32_t j = a.perm[jj];
      const double xij = xi - a.x[j];
      const double yij = yi - a.y[j];
      const double r2 = xij * xij + yij * yij;
      if (!(r2 <= radius2 + 1e-6)) continue;
      if (a.contrib_only && !(sqrt(xij * xij + yij * yij) * h1 < 2.)) continue;
      if (a.fill) a.col[o + cnt] = j;
      cnt++;
    }
  }
  if (!a.fill) a.count[i] = cnt;
}

// Rows in idx order (the O(N^2) scan order of find_neighbors).
__global__ __launch_bounds__(kBlock) void k_tl_sort_rows(uint32_t n, const uint32_t* __restrict__ rowp,
                                                         uint32_t* __restrict__ col) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = rowp[i], b = rowp[i + 1];
  for (uint32_t k = a + 1; k < b; k++) {
    const uint32_t v = col[k];
    uint32_t m = k;
    while (m > a && col[m - 1] > v) {
      col[m] = col[m - 1];
      m--;
    }
    col[m] = v;
  }
}

__device__ inline void spline64(double xi, double yi, double xj, double yj, double h, double* out) {
  // src/kernel.cpp:20-55
  const double h1 = 1. / h;
  const double xij = xi - xj;
  const double yij = yi - yj;
  const double rij = sqrt(xij * xij + yij * yij);
  const double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
  const double q = rij * h1;
  out[0] = out[1] = out[2] = 0.;
  if (q >= 2.) return;
  if (q >= 1.) {
    out[0] = fac * (0.25 * (2 - q) * (2 - q) * (2 - q));
    if (rij > 1e-12) {
      const double der = -0.75 * (2 - q) * (2 - q) * h1 / rij;
      out[1] = xij * der * fac;
      out[2] = yij * der * fac;
    }
    return;
  }
  out[0] = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
  if (rij > 1e-12) {
    const double der = -3.0 * q * (1 - 0.75 * q) * h1 / rij;
    out[1] = xij * der * fac;
    out[2] = yij * der * fac;
  }
}

// precomp_randles_libersky (src/precomp_shape_functions.h:231-290), FP64.
__global__ __launch_bounds__(kBlock) void k_tl_rl(uint32_t n, const uint32_t* __restrict__ rowp,
                                                  const uint32_t* __restrict__ col, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ h,
                                                  const double* __restrict__ m, const double* __restrict__ rho,
                                                  double* __restrict__ w64, float* __restrict__ w32,
                                                  double* __restrict__ qw) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i], yi = y[i], hi = h[i];
  qw[i] = m[i] / rho[i];
  double C = 0.;
  for (uint32_t q = rowp[i]; q < rowp[i + 1]; q++) {
    const uint32_t j = col[q];
    const double volj = m[j] / rho[j];
    double w[3];
    spline64(xi, yi, x[j], y[j], hi, w);
    C += w[0] * volj;
  }
  M2 B = m2(0., 0., 0., 0.);
  for (uint32_t q = rowp[i]; q < rowp[i + 1]; q++) {
    const uint32_t j = col[q];
    const double xj = x[j], yj = y[j];
    const double volj = m[j] / rho[j];
    double w[3];
    spline64(xi, yi, xj, yj, hi, w);
    B.m[0][0] += (xj - xi) * w[1] * volj * 1. / C;
    B.m[1][0] += (xj - xi) * w[2] * volj * 1. / C;
    B.m[0][1] += (yj - yi) * w[1] * volj * 1. / C;
    B.m[1][1] += (yj - yi) * w[2] * volj * 1. / C;
  }
  const M2 invB = m2_inv(B);
  for (uint32_t q = rowp[i]; q < rowp[i + 1]; q++) {
    const uint32_t j = col[q];
    double w[3];
    spline64(xi, yi, x[j], y[j], hi, w);
    const double ww = w[0] * 1. / C;
    const double wx = (w[1] * invB.m[0][0] + w[2] * invB.m[1][0]) * 1. / C;
    const double wy = (w[1] * invB.m[0][1] + w[2] * invB.m[1][1]) * 1. / C;
    w64[3 * (size_t)q] = ww;
    w64[3 * (size_t)q + 1] = wx;
    w64[3 * (size_t)q + 2] = wy;
    w32[3 * (size_t)q] = (float)ww;
    w32[3 * (size_t)q + 1] = (float)wx;
    w32[3 * (size_t)q + 2] = (float)wy;
  }
}

// precomp_sph (src/precomp_shape_functions.h:210-229) on the frozen lists:
// the plain kernel of each pair at the set-up positions (co-rotated bodies).
__global__ __launch_bounds__(kBlock) void k_tl_sph(uint32_t n, const uint32_t* __restrict__ rowp,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ h,
                                                   const double* __restrict__ m, const double* __restrict__ rho,
                                                   double* __restrict__ w64, float* __restrict__ w32,
                                                   double* __restrict__ qw) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qw[i] = m[i] / rho[i];
  for (uint32_t q = rowp[i]; q < rowp[i + 1]; q++) {
    double w[3];
    spline64(x[i], y[i], x[col[q]], y[col[q]], h[i], w);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      w64[3 * (size_t)q + k] = w[k];
      w32[3 * (size_t)q + k] = (float)w[k];
    }
  }
}

// precomp_randles_libersky_symmetric (src/precomp_shape_functions.h:292-316):
// for pair q = (i, j), the index of pair (j, i) in j's row (rows are sorted
// by idx, so a binary search finds the reference's first match), or ~0u when
// the lists hold no such pair (its kernel is zero: q >= 2 with h_j).
__global__ __launch_bounds__(kBlock) void k_tl_rev(uint32_t n, const uint32_t* __restrict__ rowp,
                                                   const uint32_t* __restrict__ col, uint32_t* __restrict__ rev) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (uint32_t q = rowp[i]; q < rowp[i + 1]; q++) {
    const uint32_t j = col[q];
    uint32_t lo = rowp[j], hi = rowp[j + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (col[mid] < i)
        lo = mid + 1;
      else
        hi = mid;
    }
    rev[q] = (lo < rowp[j + 1] && col[lo] == i) ? lo : ~0u;
  }
}

inline uint32_t nblk(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

template <typename T>
hipError_t talloc(TlState& t, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    t.allocs.push_back(q);
    *p = (T*)q;
  }
  return e;
}

// src: stage state, read by pass 1 (pass 2 reads pass 1's records); dst:
// next stage state written by pass 2.
TlArgs tl_args(TlState& t, const TlShared& s, Fields8 src, Fields8 dst, int stage) {
  TlArgs a;
  memset(&a, 0, sizeof(a));
  a.n = s.N;
  a.sc = s.sc;
  a.scene = s.scene;
  a.X = t.X;
  a.Y = t.Y;
  a.bc = t.bc;
  a.rowp = t.rowp;
  a.col = t.col;
  a.w64 = t.w64;
  a.w32 = t.w32;
  a.qw = t.qw;
  a.m = s.m;
  a.h = s.h;
  a.src = src;
  a.rec = t.rec;
  a.part8 = t.part8;
  a.diag = t.diag;
  a.rot = t.rot;
  a.rev = t.rev;
  a.fcx = s.fcx;
  a.fcy = s.fcy;
  a.stage = stage;
  a.y = s.y;
  a.K = dst;
  a.acc = s.acc;
  a.err = s.err;
  return a;
}

void launch_stage(const TlState& t, const TlShared& s, const TlArgs& a) {
  const dim3 g(xcd_grid(nblk(a.n))), b(kBlock);
  if (t.variant == kTlCorot) {
    hipLaunchKernelGGL((k_tl_pass1<false, kTlCorot>), g, b, 0, s.st, a);
    hipLaunchKernelGGL((k_tl_pass2<false, kTlCorot>), g, b, 0, s.st, a);
  } else if (t.variant == kTlPotential) {
    hipLaunchKernelGGL((k_tl_pass1<false, kTlPotential>), g, b, 0, s.st, a);
    hipLaunchKernelGGL((k_tl_pass2<false, kTlPotential>), g, b, 0, s.st, a);
  } else if (s.precision == MFREE_PRECISION_MIXED) {
    hipLaunchKernelGGL((k_tl_pass1<true, kTlRl>), g, b, 0, s.st, a);
    hipLaunchKernelGGL((k_tl_pass2<true, kTlRl>), g, b, 0, s.st, a);
  } else {
    hipLaunchKernelGGL((k_tl_pass1<false, kTlRl>), g, b, 0, s.st, a);
    hipLaunchKernelGGL((k_tl_pass2<false, kTlRl>), g, b, 0, s.st, a);
  }
}

// Builds CSR neighbour lists of the current positions; returns them on the
// device (caller owns rowp/col) or only the host row pointers.
mfree_status tl_build_lists(const TlShared& s, int contrib_only, double cell_min, uint32_t** rowp_out,
                            uint32_t** col_out, uint64_t* npairs, std::string& msg) {
  BinArgs b;
  memset(&b, 0, sizeof(b));
  b.n = s.N;
  b.nscene = s.S;
  b.nchunks = s.nchunks;
  b.ncell_cap = s.N;
  b.deterministic = 1;
  b.kernel_width = 2.0;
  b.cell_min = cell_min;
  b.chunks = s.chunks;
  b.chunk_first = s.chunk_first;
  b.sc = s.sc;
  b.geom = s.geom;
  b.x = s.y.f[0];
  b.y = s.y.f[1];
  b.h = s.h;
  b.scene = s.scene;
  b.gid = s.gid;
  b.part = s.part;
  b.key = s.key;
  b.rank = s.rank;
  b.count = s.count;
  b.tsum = s.tsum;
  b.cell_start = s.cell_start;
  b.perm = s.perm;
  b.err = s.err;
  launch_binning(s.st, b);
  uint32_t* cnt = nullptr;
  if (hipMalloc(&cnt, (size_t)s.N * sizeof(uint32_t)) != hipSuccess) {
    msg = "hipMalloc failed";
    return MFREE_ERR_HIP;
  }
  SearchArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.n = s.N;
  sa.sc = s.sc;
  sa.geom = s.geom;
  sa.scene = s.scene;
  sa.key = s.key;
  sa.cell_start = s.cell_start;
  sa.perm = s.perm;
  sa.x = s.y.f[0];
  sa.y = s.y.f[1];
  sa.h = s.h;
  sa.h_search = s.h_search;
  sa.contrib_only = contrib_only;
  sa.count = cnt;
  hipLaunchKernelGGL(k_tl_search, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, sa);
  std::vector<uint32_t> hc(s.N);
  hipError_t e = hipStreamSynchronize(s.st);
  if (e == hipSuccess) e = hipMemcpy(hc.data(), cnt, (size_t)s.N * sizeof(uint32_t), hipMemcpyDeviceToHost);
  uint32_t bits = 0;
  if (e == hipSuccess) e = hipMemcpy(&bits, s.err, sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess || bits) {
    hipFree(cnt);
    msg = e != hipSuccess ? hipGetErrorString(e) : "cell grid overflow in the TL neighbour search";
    return e != hipSuccess ? MFREE_ERR_HIP : MFREE_ERR_CELLS;
  }
  std::vector<uint32_t> hr((size_t)s.N + 1);
  uint64_t tot = 0;
  hr[0] = 0;
  for (uint32_t i = 0; i < s.N; i++) {
    if (hc[i] == 0) {
      hipFree(cnt);
      msg = "particle without neighbour";
      return MFREE_ERR_NO_NEIGHBOR;
    }
    tot += hc[i];
    if (tot >= (1ull << 32)) {
      hipFree(cnt);
      msg = "too many neighbour pairs";
      return MFREE_ERR_CAPACITY;
    }
    hr[i + 1] = (uint32_t)tot;
  }
  uint32_t *drow = nullptr, *dcol = nullptr;
  e = hipMalloc(&drow, hr.size() * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&dcol, std::max<uint64_t>(tot, 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(drow, hr.data(), hr.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    sa.fill = 1;
    sa.rowp = drow;
    sa.col = dcol;
    hipLaunchKernelGGL(k_tl_search, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, sa);
    hipLaunchKernelGGL(k_tl_sort_rows, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, s.N, drow, dcol);
    e = hipStreamSynchronize(s.st);
  }
  hipFree(cnt);
  if (e != hipSuccess) {
    if (drow) hipFree(drow);
    if (dcol) hipFree(dcol);
    msg = hipGetErrorString(e);
    return MFREE_ERR_HIP;
  }
  *rowp_out = drow;
  *col_out = dcol;
  *npairs = tot;
  return MFREE_OK;
}

double host_max_h(const TlShared& s) {
  std::vector<double> h(s.N);
  if (hipMemcpy(h.data(), s.h, (size_t)s.N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1.;
  double m = 0.;
  for (double v : h) m = std::max(m, v);
  return m;
}

}  // namespace

mfree_status tl_alloc(TlState& t, int variant, uint32_t N, const double* X, const double* Y,
                      const std::vector<std::vector<uint32_t>>& left, const std::vector<std::vector<uint32_t>>& right,
                      const std::vector<uint32_t>& off) {
  t.N = N;
  t.variant = variant;
  t.hbc.assign(N, 0);
  for (size_t b = 0; b < left.size(); b++) {
    for (uint32_t k : left[b]) t.hbc[off[b] + k] = 1;
    for (uint32_t k : right[b]) t.hbc[off[b] + k] = 2;
  }
  bool bad = talloc(t, &t.X, N) || talloc(t, &t.Y, N) || talloc(t, &t.bc, N) || talloc(t, &t.qw, N) ||
             talloc(t, &t.rec, 12 * (size_t)N) || talloc(t, &t.part8, 4 * (size_t)N) ||
             talloc(t, &t.diag, 16 * (size_t)N) || (variant == kTlCorot && talloc(t, &t.rot, N));
  for (int k = 0; k < 8 && !bad; k++) bad = talloc(t, &t.Kb.f[k], N) != hipSuccess;
  if (bad) {
    t.msg = "hipMalloc failed";
    return MFREE_ERR_HIP;
  }
  if (hipMemcpy(t.X, X, (size_t)N * sizeof(double), hipMemcpyHostToDevice) ||
      hipMemcpy(t.Y, Y, (size_t)N * sizeof(double), hipMemcpyHostToDevice) ||
      hipMemcpy(t.bc, t.hbc.data(), N, hipMemcpyHostToDevice)) {
    t.msg = "hipMemcpy failed";
    return MFREE_ERR_HIP;
  }
  return MFREE_OK;
}

void tl_free(TlState& t) {
  for (void* p : t.allocs) (void)hipFree(p);
  t.allocs.clear();
}

mfree_status tl_ensure_precomputed(TlState& t, const TlShared& s) {
  if (t.ready) return MFREE_OK;
  const double hmax = host_max_h(s);
  if (!(hmax > 0)) {
    t.msg = "bad smoothing lengths";
    return MFREE_ERR_ARG;
  }
  // cell side >= 2 h_max (with a rounding margin): covers every pair with q < 2
  uint32_t *rowp = nullptr, *col = nullptr;
  uint64_t np = 0;
  mfree_status st = tl_build_lists(s, 1, 2.0 * hmax * (1.0 + 1e-9), &rowp, &col, &np, t.msg);
  if (st != MFREE_OK) return st;
  t.allocs.push_back(rowp);
  t.allocs.push_back(col);
  t.rowp = rowp;
  t.col = col;
  t.npairs = np;
  if (talloc(t, &t.w64, 3 * np) || talloc(t, &t.w32, 3 * np) ||
      (t.variant == kTlPotential && talloc(t, &t.rev, np))) {
    t.msg = "hipMalloc failed";
    return MFREE_ERR_HIP;
  }
  if (t.variant == kTlCorot) {
    hipLaunchKernelGGL(k_tl_sph, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, s.N, t.rowp, t.col, s.y.f[0], s.y.f[1],
                       s.h, s.m, s.y.f[4], t.w64, t.w32, t.qw);
  } else {
    hipLaunchKernelGGL(k_tl_rl, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, s.N, t.rowp, t.col, s.y.f[0], s.y.f[1],
                       s.h, s.m, s.y.f[4], t.w64, t.w32, t.qw);
    if (t.variant == kTlPotential)
      hipLaunchKernelGGL(k_tl_rev, dim3(nblk(s.N)), dim3(kBlock), 0, s.st, s.N, t.rowp, t.col, t.rev);
  }
  if (hipStreamSynchronize(s.st) != hipSuccess) {
    t.msg = "RL precompute failed";
    return MFREE_ERR_HIP;
  }
  t.ready = true;
  return MFREE_OK;
}

void tl_enqueue_step(TlState& t, const TlShared& s, EventPairs* evs, const TakeEvent& take) {
  // stage states: k2 -> K, k3 -> Kb, k4 -> K
  for (int stage = 1; stage <= 4; stage++) {
    const Fields8 src = stage == 1 ? s.y : (stage == 3 ? t.Kb : s.K);
    const Fields8 dst = stage == 2 ? t.Kb : s.K;
    TlArgs a = tl_args(t, s, src, dst, stage);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (evs) {
      e0 = take();
      e1 = take();
      if (e0) hipEventRecord(e0, s.st);
    }
    launch_stage(t, s, a);
    if (evs && e0 && e1) {
      hipEventRecord(e1, s.st);
      evs->push_back({e0, e1});
    }
  }
}

void tl_enqueue_eval(TlState& t, const TlShared& s, double* d8, double* aux) {
  TlArgs a = tl_args(t, s, s.y, s.K, 0);
  a.d8 = d8;
  a.aux = aux;
  launch_stage(t, s, a);
}

mfree_status tl_get_pairs(TlState& t, const TlShared& s, uint32_t body, uint32_t o, uint32_t n, uint32_t* row_ptr,
                          uint32_t* col, uint64_t cap, uint64_t* npairs) {
  (void)body;
  // the reference's full predicate set, r^2 <= 4 hs^2 + 1e-6
  const double hmax = host_max_h(s);
  const double hs = s.h_search > 0 ? s.h_search : hmax;
  const double R = sqrt(hs * hs * 4 + 1e-6) * (1.0 + 1e-9);
  uint32_t *drow = nullptr, *dcol = nullptr;
  uint64_t np = 0;
  mfree_status st = tl_build_lists(s, 0, R, &drow, &dcol, &np, t.msg);
  if (st != MFREE_OK) return st;
  std::vector<uint32_t> hr((size_t)s.N + 1);
  hipError_t e = hipMemcpy(hr.data(), drow, hr.size() * sizeof(uint32_t), hipMemcpyDeviceToHost);
  uint64_t tot = 0;
  if (e == hipSuccess) {
    tot = (uint64_t)hr[o + n] - hr[o];
    if (npairs) *npairs = tot;
    for (uint32_t i = 0; i <= n; i++) row_ptr[i] = hr[o + i] - hr[o];
    if (tot > cap || !col) {
      st = MFREE_ERR_CAPACITY;
    } else {
      std::vector<uint32_t> hcol(tot);
      e = hipMemcpy(hcol.data(), dcol + hr[o], tot * sizeof(uint32_t), hipMemcpyDeviceToHost);
      for (uint64_t k = 0; k < tot; k++) col[k] = hcol[k] - o;
    }
  }
  hipFree(drow);
  hipFree(dcol);
  if (e != hipSuccess) return MFREE_ERR_HIP;
  return st;
}

}  // namespace mfree
