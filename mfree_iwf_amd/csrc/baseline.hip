// Reference-structured FP64 pipeline: the "plain FP64 GPU" baseline that the
// fused kernels are measured against.  It keeps the reference's structure
// and moves it onto the GPU without restructuring:
//
//   per derivative evaluation (src/body.h:125-148)
//     reset of every particle's rates
//     grid geometry, hashing and sort (the counting sort of grid.hip)
//     Verlet lists materialised per particle, MAX_NBH = 200 slots
//       (src/grid.h:117-177: r^2 <= 4 h_i^2 + 1e-6 over the 3x3 cells,
//       the particle itself included)
//     kernel values W, dW/dx, dW/dy materialised per list entry
//       (precomp_sph, src/kernel.cpp:20-55)
//     one kernel per action of src/benchmarks/rings.cpp:86-96
//   per step: RK4 as src/rk4.h:71-108 (stage states y + dt/2 d, y + dt d;
//   y += dt/6 (d1 + 2 d2 + 2 d3 + d4))
//
// Particles stay where they are in storage; every kernel walks the sorted
// order and reaches a particle's fields through the sort permutation (the
// reference sorts an array of particle pointers).  Arithmetic follows the
// reference's expression order (compiled with -ffp-contract=off), so the
// results agree with the fused FP64 path to rounding.
#include <math.h>

#include "ctx.h"

using namespace mfree;

namespace mfree {

constexpr uint32_t kRefMaxNbh = 200;  // src/particle.h:31 (MAX_NBH)

struct RefPipe {
  uint32_t* num_nbh = nullptr;  // by sorted position
  uint32_t* nbh = nullptr;      // [sorted][kRefMaxNbh]: sorted positions
  double* w = nullptr;          // [sorted][kRefMaxNbh][3]: W, W_x, W_y
  // per-particle fields of the actions (storage order)
  double *qw = nullptr, *p = nullptr, *Rxx = nullptr, *Rxy = nullptr, *Ryy = nullptr;
  double *Sxx_x = nullptr, *Sxy_y = nullptr, *Sxy_x = nullptr, *Syy_y = nullptr;
  double *vx_x = nullptr, *vx_y = nullptr, *vy_x = nullptr, *vy_y = nullptr;
  Fields8 d[4];                 // rates of the four stages (storage order)
  Fields8 k;                    // stage state (storage order)
  std::vector<void*> allocs;
};

namespace {

inline uint32_t rblk(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Fields the action kernels share (all storage order, reached via perm).
struct RefArgs {
  uint32_t n;
  const uint32_t *perm, *scene, *key, *cell_start;
  const SceneConst* sc;
  const SceneGeom* geom;
  Fields8 st;          // state being differentiated
  const double *m, *h;
  Fields8 d;           // its rates
  uint32_t *num_nbh, *nbh;
  double* w;
  double *qw, *p, *Rxx, *Rxy, *Ryy;
  double *Sxx_x, *Sxy_y, *Sxy_x, *Syy_y, *vx_x, *vx_y, *vy_x, *vy_y;
  uint32_t max_nbh;
  uint32_t* err;
};

// particle::reset (src/particle.cpp): rates and derived fields to zero.
__global__ void k_ref_reset(RefArgs a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
#pragma unroll
  for (int k = 0; k < 8; k++) a.d.f[k][s] = 0.;
  a.Sxx_x[s] = a.Sxy_y[s] = a.Sxy_x[s] = a.Syy_y[s] = 0.;
  a.vx_x[s] = a.vx_y[s] = a.vy_x[s] = a.vy_y[s] = 0.;
}

// grid<T>::construct_verlet_lists: the 3x3 block ni-major, nj-minor.
__global__ void k_ref_lists(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const uint32_t b = a.scene[si];
  const SceneConst* c = a.sc + b;
  const SceneGeom g = a.geom[b];
  const uint32_t local = a.key[si] - c->cell_offset;
  const uint32_t gi = local / g.ny, gj = local - gi * g.ny;
  const uint32_t lo_i = gi > 0 ? gi - 1 : 0, hi_i = min(gi + 2, g.nx);
  const uint32_t lo_j = gj > 0 ? gj - 1 : 0, hi_j = min(gj + 2, g.ny);
  const double xi = a.st.f[0][si], yi = a.st.f[1][si], hi = a.h[si];
  const double radius2 = hi * hi * 2 * 2;
  uint32_t cnt = 0;
  uint32_t* row = a.nbh + (size_t)i * kRefMaxNbh;
  for (uint32_t ni = lo_i; ni < hi_i; ni++)
    for (uint32_t nj = lo_j; nj < hi_j; nj++) {
      const uint32_t cell = c->cell_offset + ni * g.ny + nj;
      for (uint32_t j = a.cell_start[cell]; j < a.cell_start[cell + 1]; j++) {
        const uint32_t sj = a.perm[j];
        const double xij = xi - a.st.f[0][sj], yij = yi - a.st.f[1][sj];
        const double r2 = xij * xij + yij * yij;
        if (r2 <= radius2 + 1e-6) {
          if (cnt < kRefMaxNbh) row[cnt] = j;
          cnt++;
        }
      }
    }
  if (a.max_nbh && cnt >= a.max_nbh) atomicOr(a.err, EB_NBH);
  if (cnt == 0) atomicOr(a.err, EB_NONEIGH);
  a.num_nbh[i] = min(cnt, kRefMaxNbh);
}

// precomp_sph: W and grad W of every list entry (src/kernel.cpp:20-55).
__global__ void k_ref_precomp(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const double xi = a.st.f[0][si], yi = a.st.f[1][si], h = a.h[si];
  const double h1 = 1. / h;
  const double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
  for (uint32_t k = 0; k < a.num_nbh[i]; k++) {
    const uint32_t sj = a.perm[a.nbh[(size_t)i * kRefMaxNbh + k]];
    const double xij = xi - a.st.f[0][sj], yij = yi - a.st.f[1][sj];
    const double rij = sqrt(xij * xij + yij * yij);
    const double q = rij * h1;
    double w = 0., wx = 0., wy = 0.;
    if (q < 2.) {
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
    }
    double* o = a.w + ((size_t)i * kRefMaxNbh + k) * 3;
    o[0] = w;
    o[1] = wx;
    o[2] = wy;
  }
}

// riemann_update_integration_weights, material_eos,
// correctors_mghn_artificial_stress: three per-particle kernels.
__global__ void k_ref_weights(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  a.qw[s] = a.m[s] / a.st.f[4][s];
}

__global__ void k_ref_eos(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  const SceneConst* c = a.sc + a.scene[s];
  a.p[s] = c->c0sq * (a.st.f[4][s] - c->rho0);
}

__global__ void k_ref_art_stress(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  const SceneConst* c = a.sc + a.scene[s];
  const double rho = a.st.f[4][s], p = a.p[s];
  const double sxx = a.st.f[5][s] - p, sxy = a.st.f[6][s], syy = a.st.f[7][s] - p;
  const double theta = 0.5 * atan2(2. * sxy, sxx - syy + 0.);
  const double ct = cos(theta), sn = sin(theta);
  const double ct2 = ct * ct, sn2 = sn * sn;
  const double rot_sxx = ct2 * sxx + 2.0 * ct * sn * sxy + sn2 * syy;
  const double rot_syy = sn2 * sxx - 2.0 * ct * sn * sxy + ct2 * syy;
  double rot_rxx = 0., rot_ryy = 0.;
  if (rot_sxx > 0) rot_rxx = -c->mghn_eps * rot_sxx / (rho * rho);
  if (rot_syy > 0) rot_ryy = -c->mghn_eps * rot_syy / (rho * rho);
  a.Rxx[s] = ct2 * rot_rxx + sn2 * rot_ryy;
  a.Rxy[s] = ct * sn * (rot_rxx - rot_ryy);
  a.Ryy[s] = sn2 * rot_rxx + ct2 * rot_ryy;
}

// derive_stress_monaghan: rho_i * sum m_j (sigma_i/rho_i^2 + sigma_j/rho_j^2 + R f^n) grad W
__global__ void k_ref_derive_stress(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const SceneConst* c = a.sc + a.scene[si];
  const double rhoi = a.st.f[4][si], pi = a.p[si];
  const double rho21i = 1. / (rhoi * rhoi);
  const double sxxi = (a.st.f[5][si] - pi) * rho21i, sxyi = a.st.f[6][si] * rho21i;
  const double syyi = (a.st.f[7][si] - pi) * rho21i;
  const bool use_as = c->wdeltap > 0;
  const bool pow4 = c->pow4 != 0;
  double Sxx_x = 0., Sxy_y = 0., Sxy_x = 0., Syy_y = 0.;
  for (uint32_t k = 0; k < a.num_nbh[i]; k++) {
    const uint32_t sj = a.perm[a.nbh[(size_t)i * kRefMaxNbh + k]];
    const double* wk = a.w + ((size_t)i * kRefMaxNbh + k) * 3;
    const double w = wk[0], wx = wk[1], wy = wk[2];
    double Rxx = 0., Rxy = 0., Ryy = 0.;
    if (use_as) {
      double fab = w / c->wdeltap;
      fab = pow4 ? (fab * fab) * (fab * fab) : pow(fab, c->corr_exp);
      Rxx = fab * (a.Rxx[si] + a.Rxx[sj]);
      Rxy = fab * (a.Rxy[si] + a.Rxy[sj]);
      Ryy = fab * (a.Ryy[si] + a.Ryy[sj]);
    }
    const double rhoj = a.st.f[4][sj], pj = a.p[sj];
    const double rho21j = 1. / (rhoj * rhoj);
    const double mj = a.m[sj];
    Sxx_x += mj * (sxxi + (a.st.f[5][sj] - pj) * rho21j + Rxx) * wx;
    Sxy_y += mj * (sxyi + a.st.f[6][sj] * rho21j + Rxy) * wy;
    Sxy_x += mj * (sxyi + a.st.f[6][sj] * rho21j + Rxy) * wx;
    Syy_y += mj * (syyi + (a.st.f[7][sj] - pj) * rho21j + Ryy) * wy;
  }
  a.Sxx_x[si] = Sxx_x * rhoi;
  a.Sxy_y[si] = Sxy_y * rhoi;
  a.Sxy_x[si] = Sxy_x * rhoi;
  a.Syy_y[si] = Syy_y * rhoi;
}

// derive_velocity: sum (v_j - v_i) grad W m_j/rho_j
__global__ void k_ref_derive_velocity(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const double vxi = a.st.f[2][si], vyi = a.st.f[3][si];
  double vx_x = 0., vx_y = 0., vy_x = 0., vy_y = 0.;
  for (uint32_t k = 0; k < a.num_nbh[i]; k++) {
    const uint32_t sj = a.perm[a.nbh[(size_t)i * kRefMaxNbh + k]];
    const double* wk = a.w + ((size_t)i * kRefMaxNbh + k) * 3;
    const double wx = wk[1], wy = wk[2];
    const double qwj = a.qw[sj];
    vx_x += (a.st.f[2][sj] - vxi) * wx * qwj;
    vx_y += (a.st.f[2][sj] - vxi) * wy * qwj;
    vy_x += (a.st.f[3][sj] - vyi) * wx * qwj;
    vy_y += (a.st.f[3][sj] - vyi) * wy * qwj;
  }
  a.vx_x[si] = vx_x;
  a.vx_y[si] = vx_y;
  a.vy_x[si] = vy_x;
  a.vy_y[si] = vy_y;
}

// correctors_mghn_artificial_viscosity: vx_t, vy_t += -m_j Pi_ij grad W
// for approaching pairs.
__global__ void k_ref_art_visc(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const SceneConst* c = a.sc + a.scene[si];
  const double xi = a.st.f[0][si], yi = a.st.f[1][si], vxi = a.st.f[2][si], vyi = a.st.f[3][si];
  const double rhoi = a.st.f[4][si], hi = a.h[si];
  const double ci = sqrt(c->K / rhoi);
  double ax = 0., ay = 0.;
  for (uint32_t k = 0; k < a.num_nbh[i]; k++) {
    const uint32_t sj = a.perm[a.nbh[(size_t)i * kRefMaxNbh + k]];
    const double* wk = a.w + ((size_t)i * kRefMaxNbh + k) * 3;
    const double wx = wk[1], wy = wk[2];
    const double xij = xi - a.st.f[0][sj], yij = yi - a.st.f[1][sj];
    const double r2 = xij * xij + yij * yij;
    const double vijx = vxi - a.st.f[2][sj], vijy = vyi - a.st.f[3][sj];
    const double vp = vijx * xij + vijy * yij;
    if (vp >= 0.) continue;
    const double rhoj = a.st.f[4][sj];
    const double cij = 0.5 * (ci + sqrt(c->K / rhoj));
    const double hij = 0.5 * (hi + a.h[sj]);
    const double rhoij = 0.5 * (rhoi + rhoj);
    const double muij = (hij * vp) / (r2 + c->eta * c->eta * hij * hij);
    const double piij = (-c->alpha * cij * muij + c->beta * muij * muij) / rhoij;
    ax += -a.m[sj] * piij * wx;
    ay += -a.m[sj] * piij * wy;
  }
  a.d.f[2][si] += ax;
  a.d.f[3][si] += ay;
}

// correctors_xsph: x_t, y_t += -eps W m_i / rho_ij (v_i - v_j)
__global__ void k_ref_xsph(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t si = a.perm[i];
  const SceneConst* c = a.sc + a.scene[si];
  const double vxi = a.st.f[2][si], vyi = a.st.f[3][si], rhoi = a.st.f[4][si], mi = a.m[si];
  double xs = 0., ys = 0.;
  for (uint32_t k = 0; k < a.num_nbh[i]; k++) {
    const uint32_t sj = a.perm[a.nbh[(size_t)i * kRefMaxNbh + k]];
    const double w = a.w[((size_t)i * kRefMaxNbh + k) * 3];
    const double rhoij = 0.5 * (rhoi + a.st.f[4][sj]);
    const double fx = -c->xsph_eps * w * mi / rhoij;
    xs += fx * (vxi - a.st.f[2][sj]);
    ys += fx * (vyi - a.st.f[3][sj]);
  }
  a.d.f[0][si] += xs;
  a.d.f[1][si] += ys;
}

// material_stress_rate_jaumann (glm order, as the fused kernel)
__global__ void k_ref_jaumann(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  const SceneConst* c = a.sc + a.scene[s];
  const double vx_x = a.vx_x[s], vx_y = a.vx_y[s], vy_x = a.vy_x[s], vy_y = a.vy_y[s];
  const double Sxx = a.st.f[5][s], Sxy = a.st.f[6][s], Syy = a.st.f[7][s];
  const double G2 = c->G2;
  const double e01 = 0.5 * (vx_y + vy_x);
  const double w10 = 0.5 * (vy_x - vx_y);
  const double w01 = 0.5 * (vx_y - vy_x);
  const double s3 = 1. / 3. * (vx_x + vy_y);
  const double A00 = G2 * (vx_x - s3 * 1.), A01 = G2 * (e01 - s3 * 0.), A11 = G2 * (vy_y - s3 * 1.);
  const double SW00 = Sxx * 0. + Sxy * w10, SW01 = Sxy * 0. + Syy * w10, SW11 = Sxy * w01 + Syy * 0.;
  const double WS00 = 0. * Sxx + w01 * Sxy, WS01 = w10 * Sxx + 0. * Sxy, WS11 = w10 * Sxy + 0. * Syy;
  a.d.f[5][s] += (A00 - SW00) + WS00;
  a.d.f[6][s] += (A01 - SW01) + WS01;
  a.d.f[7][s] += (A11 - SW11) + WS11;
}

// contmech_continuity, contmech_momentum_ul, contmech_advection
__global__ void k_ref_continuity(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  a.d.f[4][s] += -a.st.f[4][s] * (a.vx_x[s] + a.vy_y[s]);
}

__global__ void k_ref_momentum(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  const double rho = a.st.f[4][s], m = a.m[s];
  a.d.f[2][s] += 1. / rho * (a.Sxx_x[s] + a.Sxy_y[s]) + 0. / m;
  a.d.f[3][s] += 1. / rho * (a.Sxy_x[s] + a.Syy_y[s]) + 0. / m;
}

__global__ void k_ref_advection(RefArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t s = a.perm[i];
  a.d.f[0][s] += a.st.f[2][s];
  a.d.f[1][s] += a.st.f[3][s];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 8; k++) ok = ok && isfinite(a.d.f[k][s]);
  if (!ok) atomicOr(a.err, EB_NONFINITE);
}

// rk_add: out = y + d * dt_stage (per body)
__global__ void k_ref_rk_add(uint32_t n, const uint32_t* __restrict__ scene, const SceneConst* __restrict__ sc,
                             Fields8 y, Fields8 d, Fields8 out, int full) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SceneConst* c = sc + scene[s];
  const double dts = full ? c->dt : c->half_dt;
#pragma unroll
  for (int k = 0; k < 8; k++) out.f[k][s] = y.f[k][s] + d.f[k][s] * dts;
}

// y += dt/6 (d1 + 2 d2 + 2 d3 + d4)
__global__ void k_ref_rk_final(uint32_t n, const uint32_t* __restrict__ scene, const SceneConst* __restrict__ sc,
                               Fields8 y, Fields8 d1, Fields8 d2, Fields8 d3, Fields8 d4) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SceneConst* c = sc + scene[s];
#pragma unroll
  for (int k = 0; k < 8; k++)
    y.f[k][s] = y.f[k][s] + c->dt6 * (((d1.f[k][s] + 2 * d2.f[k][s]) + 2 * d3.f[k][s]) + d4.f[k][s]);
}

__global__ void k_ref_copy8(uint32_t n, Fields8 src, Fields8 dst) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int k = 0; k < 8; k++) dst.f[k][s] = src.f[k][s];
}

// derivatives by global id (parity helper)
__global__ void k_ref_export(uint32_t n, const uint32_t* __restrict__ gid, Fields8 d, double* __restrict__ d8) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int k = 0; k < 8; k++) d8[8 * (size_t)gid[s] + k] = d.f[k][s];
}

RefArgs ref_args(mfree_ctx* c, const Fields8& st, const Fields8& d) {
  RefPipe* r = c->ref;
  RefArgs a;
  a.n = c->N;
  a.perm = c->perm;
  a.scene = c->scene[0];
  a.key = c->key;
  a.cell_start = c->cell_start;
  a.sc = c->sc;
  a.geom = c->geom;
  a.st = st;
  a.m = c->m[0];
  a.h = c->h[0];
  a.d = d;
  a.num_nbh = r->num_nbh;
  a.nbh = r->nbh;
  a.w = r->w;
  a.qw = r->qw;
  a.p = r->p;
  a.Rxx = r->Rxx;
  a.Rxy = r->Rxy;
  a.Ryy = r->Ryy;
  a.Sxx_x = r->Sxx_x;
  a.Sxy_y = r->Sxy_y;
  a.Sxy_x = r->Sxy_x;
  a.Syy_y = r->Syy_y;
  a.vx_x = r->vx_x;
  a.vx_y = r->vx_y;
  a.vy_x = r->vy_x;
  a.vy_y = r->vy_y;
  a.max_nbh = c->opt.max_nbh;
  a.err = c->errb;
  return a;
}

// One derivative evaluation of state st into d (src/body.h:125-148).
void ref_derivs(mfree_ctx* c, const Fields8& st, const Fields8& d) {
  hipStream_t s = c->stream;
  const uint32_t g = rblk(c->N);
  RefArgs a = ref_args(c, st, d);
  hipLaunchKernelGGL(k_ref_reset, dim3(g), dim3(kBlock), 0, s, a);
  BinArgs b = bin_args(c, st.f[0], st.f[1], 0);
  b.tile_cells = 0;
  b.tile_off = nullptr;
  launch_binning(s, b);
  hipLaunchKernelGGL(k_ref_lists, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_precomp, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_weights, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_eos, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_art_stress, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_derive_stress, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_derive_velocity, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_art_visc, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_xsph, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_jaumann, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_continuity, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_momentum, dim3(g), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_ref_advection, dim3(g), dim3(kBlock), 0, s, a);
}

}  // namespace

mfree_status ref_alloc(mfree_ctx* c) {
  if (c->ref) return MFREE_OK;
  RefPipe* r = new RefPipe();
  c->ref = r;
  const size_t N = c->N;
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    r->allocs.push_back(*p);
    return true;
  };
  bool ok = al((void**)&r->num_nbh, N * 4) && al((void**)&r->nbh, N * kRefMaxNbh * 4) &&
            al((void**)&r->w, N * kRefMaxNbh * 24);
  double** f[] = {&r->qw, &r->p, &r->Rxx, &r->Rxy, &r->Ryy, &r->Sxx_x, &r->Sxy_y,
                  &r->Sxy_x, &r->Syy_y, &r->vx_x, &r->vx_y, &r->vy_x, &r->vy_y};
  for (double** q : f) ok = ok && al((void**)q, N * 8);
  for (int s = 0; s < 4; s++)
    for (int k = 0; k < 8; k++) ok = ok && al((void**)&r->d[s].f[k], N * 8);
  for (int k = 0; k < 8; k++) ok = ok && al((void**)&r->k.f[k], N * 8);
  if (!ok) return fail(c, MFREE_ERR_HIP, "hipMalloc failed (reference-structured pipeline)");
  return MFREE_OK;
}

void ref_free(mfree_ctx* c) {
  if (!c->ref) return;
  for (void* p : c->ref->allocs) (void)hipFree(p);
  delete c->ref;
  c->ref = nullptr;
}

// One RK4 step (src/rk4.h:71-108); storage is never reordered.
void ref_enqueue_step(mfree_ctx* c) {
  RefPipe* r = c->ref;
  hipStream_t s = c->stream;
  const uint32_t g = rblk(c->N);
  const Fields8& y = c->y[0];
  hipLaunchKernelGGL(k_ref_copy8, dim3(g), dim3(kBlock), 0, s, c->N, y, r->k);
  ref_derivs(c, r->k, r->d[0]);
  hipLaunchKernelGGL(k_ref_rk_add, dim3(g), dim3(kBlock), 0, s, c->N, c->scene[0], c->sc, y, r->d[0], r->k, 0);
  ref_derivs(c, r->k, r->d[1]);
  hipLaunchKernelGGL(k_ref_rk_add, dim3(g), dim3(kBlock), 0, s, c->N, c->scene[0], c->sc, y, r->d[1], r->k, 0);
  ref_derivs(c, r->k, r->d[2]);
  hipLaunchKernelGGL(k_ref_rk_add, dim3(g), dim3(kBlock), 0, s, c->N, c->scene[0], c->sc, y, r->d[2], r->k, 1);
  ref_derivs(c, r->k, r->d[3]);
  hipLaunchKernelGGL(k_ref_rk_final, dim3(g), dim3(kBlock), 0, s, c->N, c->scene[0], c->sc, y, r->d[0], r->d[1],
                     r->d[2], r->d[3]);
  hipLaunchKernelGGL(k_advance_time, dim3((c->S + 63) / 64), dim3(64), 0, s, c->S, c->sc, nullptr);
}

void ref_enqueue_eval(mfree_ctx* c, double* d8) {
  RefPipe* r = c->ref;
  ref_derivs(c, c->y[0], r->d[0]);
  hipLaunchKernelGGL(k_ref_export, dim3(rblk(c->N)), dim3(kBlock), 0, c->stream, c->N, c->gid[0], r->d[0], d8);
}

}  // namespace mfree
