// CPU oracle for the SPH particle-interaction hot path of mfree_iwf.
//
// TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Only the test
// suite, the smoke check and the benchmark's CPU-baseline leg load this
// library.  Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
//
// Every routine restates the reference's algorithm in the reference's own
// floating-point operation order (x86-64 g++ without -march: no FMA), so a
// restated expression `a*b*c + d` is evaluated ((a*b)*c)+d exactly as the
// reference source is.  Citations are path:line into the reference tree.
//
// The reference's 2x2 matrix algebra (Jaumann rate, F^-1, P = J F^-1 sigma,
// Randles-Libersky B^-1) is done with glm (a header-only third-party library,
// version not pinned by the reference).  glm's published algorithms are
// restated here: column-major storage m[col][row]; product
// (A*B)[c][r] = A[0][r]*B[c][0] + A[1][r]*B[c][1]; determinant
// m00*m11 - m10*m01; inverse = adjugate * (1/det) with the reciprocal taken
// once.  Parity of these against glm itself is unpinned at the bit level
// (glm is absent) and pinned by closed forms in the tests.

#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------------
// glm-compatible 2x2 (column-major, m[c][r]).
struct M2 {
  double m[2][2];
};

// glm::dmat2x2(a, b, c, d): column 0 = (a, b), column 1 = (c, d).
inline M2 mk(double a, double b, double c, double d) {
  M2 r;
  r.m[0][0] = a;
  r.m[0][1] = b;
  r.m[1][0] = c;
  r.m[1][1] = d;
  return r;
}

inline M2 mul(const M2& a, const M2& b) {
  M2 r;
  for (int c = 0; c < 2; ++c)
    for (int row = 0; row < 2; ++row) r.m[c][row] = a.m[0][row] * b.m[c][0] + a.m[1][row] * b.m[c][1];
  return r;
}

inline M2 smul(double s, const M2& a) {
  M2 r;
  for (int c = 0; c < 2; ++c)
    for (int row = 0; row < 2; ++row) r.m[c][row] = s * a.m[c][row];
  return r;
}

inline M2 sub(const M2& a, const M2& b) {
  M2 r;
  for (int c = 0; c < 2; ++c)
    for (int row = 0; row < 2; ++row) r.m[c][row] = a.m[c][row] - b.m[c][row];
  return r;
}

inline M2 add(const M2& a, const M2& b) {
  M2 r;
  for (int c = 0; c < 2; ++c)
    for (int row = 0; row < 2; ++row) r.m[c][row] = a.m[c][row] + b.m[c][row];
  return r;
}

inline double det(const M2& a) { return a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]; }

inline M2 inv(const M2& a) {
  double ood = 1.0 / (a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1]);
  return mk(a.m[1][1] * ood, -a.m[0][1] * ood, -a.m[1][0] * ood, a.m[0][0] * ood);
}

// ------------------------------------------------------------------------

inline void spline(double xi, double yi, double xj, double yj, double h, double* out) {
  // src/kernel.cpp:20-55
  double h1 = 1. / h;
  double xij = xi - xj;
  double yij = yi - yj;
  double rij = sqrt(xij * xij + yij * yij);
  double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
  double q = rij * h1;
  out[0] = 0.;
  out[1] = 0.;
  out[2] = 0.;
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

inline double G_of(double E, double nu) { return E / (2. * (1. + nu)); }  // src/simulation_data.cpp:32-34
inline double K_of(double E, double nu) {                                 // src/simulation_data.cpp:36-39
  double G = G_of(E, nu);
  return 2.0 * G * (1 + nu) / (3 * (1 - 2 * nu));
}

int set_threads(int threads) {
  int t = threads > 0 ? threads : 1;
  return t;
}

// Per-stage workspace of a UL body: sorted permutation, cell table, Verlet
// lists (sorted-position space, like particle::nbh) and precomputed kernels
// (particle::w).
struct UlStage {
  std::vector<uint32_t> hash;   // per idx
  std::vector<uint32_t> perm;   // sorted position -> idx
  std::vector<int> cells;
  std::vector<uint64_t> rowp;   // per sorted position
  std::vector<uint32_t> nbh;    // sorted positions
  std::vector<double> w;        // 3 per stored pair
  // state gathered into sorted order (the reference's sorted pointer array)
  std::vector<double> x, y, vx, vy, rho, m, h, Sxx, Sxy, Syy, fcx, fcy;
  // per-particle action results in sorted order
  std::vector<double> qw, p, Rxx, Rxy, Ryy, Sxx_x, Sxy_y, Sxy_x, Syy_y, vx_x, vx_y, vy_x, vy_y;
  std::vector<double> x_t, y_t, vx_t, vy_t, rho_t, Sxx_t, Sxy_t, Syy_t;
};

void resize_all(UlStage& s, uint32_t n) {
  std::vector<double>* v[] = {&s.x,     &s.y,     &s.vx,    &s.vy,    &s.rho,   &s.m,     &s.h,
                              &s.Sxx,   &s.Sxy,   &s.Syy,   &s.fcx,   &s.fcy,   &s.qw,    &s.p,
                              &s.Rxx,   &s.Rxy,   &s.Ryy,   &s.Sxx_x, &s.Sxy_y, &s.Sxy_x, &s.Syy_y,
                              &s.vx_x,  &s.vx_y,  &s.vy_x,  &s.vy_y,  &s.x_t,   &s.y_t,   &s.vx_t,
                              &s.vy_t,  &s.rho_t, &s.Sxx_t, &s.Sxy_t, &s.Syy_t};
  for (auto* a : v) a->assign(n, 0.);
  s.hash.assign(n, 0);
  s.perm.resize(n);
  s.rowp.assign((size_t)n + 1, 0);
}

int geometry(uint32_t n, const double* x, const double* y, const double* h, double kernel_width,
             orc_geom* g) {
  // src/grid.h:83-115 (update_geometry); the loop runs over the array in
  // idx order, as the array is idx-ordered when it is called.
  double bminx = 1.7976931348623157e308, bmaxx = -1.7976931348623157e308;
  double bminy = 1.7976931348623157e308, bmaxy = -1.7976931348623157e308;
  double h_max = -1.7976931348623157e308;
  for (uint32_t i = 0; i < n; i++) {
    bminx = fmin(x[i], bminx);
    bminy = fmin(y[i], bminy);
    bmaxx = fmax(x[i], bmaxx);
    bmaxy = fmax(y[i], bmaxy);
    h_max = fmax(h[i], h_max);
  }
  bminx -= 1e-6;
  bminy -= 1e-6;
  bmaxx += 1e-6;
  bmaxy += 1e-6;
  double dx = h_max * kernel_width;
  double lx = bmaxx - bminx;
  double ly = bmaxy - bminy;
  unsigned int nx = ceil(lx / dx);
  unsigned int ny = ceil(ly / dx);
  g->bbmin_x = bminx;
  g->bbmin_y = bminy;
  g->bbmax_x = bmaxx;
  g->bbmax_y = bmaxy;
  g->cell = dx;
  g->nx = nx;
  g->ny = ny;
  g->num_cell = nx * ny;
  if (!(g->num_cell < n)) return ORC_ERR_CELLS;
  return ORC_OK;
}

// Builds s.perm (sorted position -> idx), s.cells, and the Verlet lists in
// sorted-position space.  x, y, h are in idx order.
int build_lists(uint32_t n, const double* x, const double* y, const double* h, int order,
                uint32_t max_nbh, UlStage& s, orc_geom* gout, int nt = 1) {
  orc_geom g;
  int st = geometry(n, x, y, h, 2., &g);
  if (gout) *gout = g;
  if (st) return st;
  // src/grid.h:35-41 (assign_hashes)
  for (uint32_t i = 0; i < n; i++) {
    unsigned int ix = (unsigned int)((x[i] - g.bbmin_x) / g.cell);
    unsigned int iy = (unsigned int)((y[i] - g.bbmin_y) / g.cell);
    s.hash[i] = ix * g.ny + iy;
    if (s.hash[i] >= g.num_cell) return ORC_ERR_CELLS;
  }
  // src/body.h:137: std::sort of the idx-ordered array by hash.
  const uint32_t* hs = s.hash.data();
  if (order == ORC_ORDER_REFERENCE) {
    for (uint32_t i = 0; i < n; i++) s.perm[i] = i;
    std::sort(s.perm.begin(), s.perm.end(), [hs](uint32_t a, uint32_t b) { return hs[a] < hs[b]; });
  } else {
    // a stable sort of the idx-ordered array by hash is a counting sort
    std::vector<uint32_t> start((size_t)g.num_cell + 1, 0);
    for (uint32_t i = 0; i < n; i++) start[hs[i] + 1]++;
    for (uint32_t b = 0; b < g.num_cell; b++) start[b + 1] += start[b];
    for (uint32_t i = 0; i < n; i++) s.perm[start[hs[i]]++] = i;
  }

  // src/grid.h:48-81 (get_cells)
  s.cells.assign((size_t)g.num_cell + 1, -1);
  s.cells[hs[s.perm[0]]] = 0;
  for (uint32_t i = 0; i + 1 < n; i++)
    if (hs[s.perm[i]] != hs[s.perm[i + 1]]) s.cells[hs[s.perm[i + 1]]] = (int)(i + 1);
  s.cells[hs[s.perm[n - 1]] + 1] = (int)n;
  int fix = (int)n;
  for (auto it = s.cells.rbegin(); it != s.cells.rend(); ++it) {
    if (*it == -1)
      *it = fix;
    else
      fix = *it;
  }

  // src/grid.h:117-177 (construct_verlet_lists), lists in sorted positions.
  // Rows are emitted cell by cell, i.e. in sorted-position order; the rows
  // are built in two passes (count, then fill at the counted offsets) so
  // that threads can share the work without changing any row.
  // positions and h in sorted order first: the scan then reads neighbours
  // from consecutive memory instead of through the permutation
  std::vector<double> sx(n), sy(n), sh(n);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const uint32_t k = s.perm[i];
    sx[i] = x[k];
    sy[i] = y[k];
    sh[i] = h[k];
  }
  auto scan_row = [&](uint32_t i, uint32_t* out) -> uint64_t {
    const unsigned int b = hs[s.perm[i]];
    const unsigned int gi = b / g.ny;
    const unsigned int gj = b - gi * g.ny;
    const int low_i = (int)gi - 1 < 0 ? 0 : (int)gi - 1;
    const int low_j = (int)gj - 1 < 0 ? 0 : (int)gj - 1;
    const int high_i = gi + 2 > g.nx ? g.nx : gi + 2;
    const int high_j = gj + 2 > g.ny ? g.ny : gj + 2;
    const double hi = sh[i];
    const double xi = sx[i];
    const double yi = sy[i];
    const double radius2 = hi * hi * 2 * 2;
    uint64_t cnt = 0;
    for (int ni = low_i; ni < high_i; ni++) {
      for (int nj = low_j; nj < high_j; nj++) {
        int c = ni * (int)g.ny + nj;
        for (int j = s.cells[c]; j < s.cells[c + 1]; j++) {
          double xij = xi - sx[j];
          double yij = yi - sy[j];
          double r2 = xij * xij + yij * yij;
          if (r2 <= radius2 + 1e-6) {
            if (out) out[cnt] = (uint32_t)j;
            cnt++;
          }
        }
      }
    }
    return cnt;
  };
  int bad = ORC_OK;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) s.rowp[i + 1] = scan_row((uint32_t)i, nullptr);
  s.rowp[0] = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t cnt = s.rowp[i + 1];
    // the first failing row in emission order decides the status, as in the
    // reference's sequential loop
    if (bad == ORC_OK && cnt >= max_nbh) bad = ORC_ERR_NBH_OVERFLOW;
    if (bad == ORC_OK && cnt == 0) bad = ORC_ERR_NO_NEIGHBOR;
    s.rowp[i + 1] += s.rowp[i];
  }
  if (bad) return bad;
  s.nbh.resize(s.rowp[n]);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) scan_row((uint32_t)i, s.nbh.data() + s.rowp[i]);
  return ORC_OK;
}

void godunov_particle(const orc_consts* c, UlStage& s, const uint64_t* R, const uint32_t* NB, const double* W,
                      bool skip0, int64_t i);

// One UL derivative evaluation on a stage state (idx order): the
// Gray-Monaghan action list (method ORC_METHOD_MONAGHAN) or the Godunov one
// (ORC_METHOD_GODUNOV).
int ul_stage(const orc_consts* c, int method, uint32_t n, const orc_soa* st, int order, uint32_t max_nbh,
             UlStage& s, int threads, double* d8, double* aux) {
  int nt = set_threads(threads);
  int rc = build_lists(n, st->x, st->y, st->h, order, max_nbh, s, nullptr, nt);
  if (rc) return rc;
  const uint32_t* P = s.perm.data();
  // gather into sorted order (the reference's actions run on the sorted array)
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t k = P[i];
    s.x[i] = st->x[k];
    s.y[i] = st->y[k];
    s.vx[i] = st->vx[k];
    s.vy[i] = st->vy[k];
    s.rho[i] = st->rho[k];
    s.m[i] = st->m[k];
    s.h[i] = st->h[k];
    s.Sxx[i] = st->Sxx[k];
    s.Sxy[i] = st->Sxy[k];
    s.Syy[i] = st->Syy[k];
    s.fcx[i] = st->fcx ? st->fcx[k] : 0.;
    s.fcy[i] = st->fcy ? st->fcy[k] : 0.;
    // particle::reset, src/particle.cpp:34-43
    s.x_t[i] = s.y_t[i] = s.rho_t[i] = s.vx_t[i] = s.vy_t[i] = 0.;
    s.Sxx_t[i] = s.Sxy_t[i] = s.Syy_t[i] = 0.;
  }
  const uint64_t* R = s.rowp.data();
  const uint32_t* NB = s.nbh.data();
  s.w.resize(3 * (size_t)R[n]);
  double* W = s.w.data();

  // precomp_sph, src/precomp_shape_functions.h:210-229
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      spline(s.x[i], s.y[i], s.x[j], s.y[j], s.h[i], W + 3 * q);
    }
  }

  const double K = K_of(c->E, c->nu);
  const double G = G_of(c->E, c->nu);

#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    // riemann_update_integration_weights, src/riemann_integration.h:28-33
    s.qw[i] = s.m[i] / s.rho[i];
    // material_eos, src/material.h:37-44
    double c0 = sqrt(K / c->rho0);
    s.p[i] = c0 * c0 * (s.rho[i] - c->rho0);
    if (method == ORC_METHOD_GODUNOV) continue;  // src/benchmarks/rings.cpp:151-158: no artificial stress
    // correctors_mghn_artificial_stress, src/correctors.h:37-81
    double rhoi = s.rho[i];
    double sxx = s.Sxx[i] - s.p[i];
    double syy = s.Syy[i] - s.p[i];
    double sxy = s.Sxy[i];
    double numer = 2. * sxy;
    double denom = sxx - syy + 0.;
    double theta = 0.5 * atan2(numer, denom);
    double ct = cos(theta);
    double sn = sin(theta);
    double ct2 = ct * ct;
    double sn2 = sn * sn;
    double rot_sxx = ct2 * sxx + 2.0 * ct * sn * sxy + sn2 * syy;
    double rot_syy = sn2 * sxx - 2.0 * ct * sn * sxy + ct2 * syy;
    double rot_rxx = 0.;
    double rot_ryy = 0.;
    double eps = c->mghn_eps;
    if (rot_sxx > 0) rot_rxx = -eps * rot_sxx / (rhoi * rhoi);
    if (rot_syy > 0) rot_ryy = -eps * rot_syy / (rhoi * rhoi);
    s.Rxx[i] = ct2 * rot_rxx + sn2 * rot_ryy;
    s.Rxy[i] = ct * sn * (rot_rxx - rot_ryy);
    s.Ryy[i] = sn2 * rot_rxx + ct2 * rot_ryy;
  }

  // The stored pairs include every candidate with r^2 <= 4h^2 + 1e-6; those
  // with q >= 2 have W = 0 and grad W = 0, so each of their terms below is
  // an exact +-0 (for finite states).  In stable order (the validation mode)
  // they are skipped: the sums keep every value of the reference's loops (up
  // to the sign of a zero sum).  In the reference's order every stored pair
  // is evaluated, as the reference does (the CPU baseline times this mode).
  const bool skip0 = order == ORC_ORDER_STABLE;
  const double wdeltap = c->wdeltap, corr_exp = c->corr_exp;
  const double alpha = c->alpha, beta = c->beta, eta = c->eta, xeps = c->xsph_eps;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    if (method == ORC_METHOD_GODUNOV) {
      godunov_particle(c, s, R, NB, W, skip0, i);
      continue;
    }
    // derive_stress_monaghan, src/derivatives.h:378-443
    {
      double Sxxi = s.Sxx[i] - s.p[i];
      double Sxyi = s.Sxy[i];
      double Syyi = s.Syy[i] - s.p[i];
      double Rxxi = s.Rxx[i], Rxyi = s.Rxy[i], Ryyi = s.Ryy[i];
      double rhoi = s.rho[i];
      double rhoi21 = 1. / (rhoi * rhoi);
      double Sxx_x = 0., Sxy_y = 0., Sxy_x = 0., Syy_y = 0.;
      for (uint64_t q = R[i]; q < R[i + 1]; q++) {
        uint32_t j = NB[q];
        const double* w = W + 3 * q;
        if (skip0 && w[0] == 0.) continue;  // q >= 2: W = grad W = 0, every term adds an exact +-0
        double Sxxj = s.Sxx[j] - s.p[j];
        double Sxyj = s.Sxy[j];
        double Syyj = s.Syy[j] - s.p[j];
        double mj = s.m[j];
        double rhoj21 = 1. / (s.rho[j] * s.rho[j]);
        double Rxx = 0, Rxy = 0., Ryy = 0.;
        if (wdeltap > 0) {
          double fab = w[0] / wdeltap;
          fab = pow(fab, corr_exp);
          Rxx = fab * (Rxxi + s.Rxx[j]);
          Rxy = fab * (Rxyi + s.Rxy[j]);
          Ryy = fab * (Ryyi + s.Ryy[j]);
        }
        Sxx_x += mj * (Sxxi * rhoi21 + Sxxj * rhoj21 + Rxx) * w[1];
        Sxy_y += mj * (Sxyi * rhoi21 + Sxyj * rhoj21 + Rxy) * w[2];
        Sxy_x += mj * (Sxyi * rhoi21 + Sxyj * rhoj21 + Rxy) * w[1];
        Syy_y += mj * (Syyi * rhoi21 + Syyj * rhoj21 + Ryy) * w[2];
      }
      s.Sxx_x[i] = Sxx_x * rhoi;
      s.Sxy_y[i] = Sxy_y * rhoi;
      s.Sxy_x[i] = Sxy_x * rhoi;
      s.Syy_y[i] = Syy_y * rhoi;
    }
    // derive_velocity, src/derivatives.h:32-66
    {
      double vxi = s.vx[i], vyi = s.vy[i];
      double vx_x = 0., vx_y = 0., vy_x = 0., vy_y = 0.;
      for (uint64_t q = R[i]; q < R[i + 1]; q++) {
        uint32_t j = NB[q];
        const double* w = W + 3 * q;
        if (skip0 && w[0] == 0.) continue;  // q >= 2: W = grad W = 0, every term adds an exact +-0
        double vxj = s.vx[j], vyj = s.vy[j];
        double quad_weight = s.qw[j];
        vx_x += (vxj - vxi) * w[1] * quad_weight;
        vx_y += (vxj - vxi) * w[2] * quad_weight;
        vy_x += (vyj - vyi) * w[1] * quad_weight;
        vy_y += (vyj - vyi) * w[2] * quad_weight;
      }
      s.vx_x[i] = vx_x;
      s.vx_y[i] = vx_y;
      s.vy_x[i] = vy_x;
      s.vy_y[i] = vy_y;
    }
    // correctors_mghn_artificial_viscosity, src/correctors.h:83-137
    {
      double xi = s.x[i], yi = s.y[i], vxi = s.vx[i], vyi = s.vy[i], hi = s.h[i], rhoi = s.rho[i];
      double ci = sqrt(K / rhoi);
      for (uint64_t q = R[i]; q < R[i + 1]; q++) {
        uint32_t j = NB[q];
        const double* w = W + 3 * q;
        if (skip0 && w[0] == 0.) continue;  // q >= 2: W = grad W = 0, every term adds an exact +-0
        double xj = s.x[j], yj = s.y[j], vxj = s.vx[j], vyj = s.vy[j], hj = s.h[j];
        double mj = s.m[j], rhoj = s.rho[j];
        double cj = sqrt(K / rhoj);
        double vijx = vxi - vxj;
        double vijy = vyi - vyj;
        double xij = xi - xj;
        double yij = yi - yj;
        double vijposij = vijx * xij + vijy * yij;
        if (vijposij >= 0.) continue;
        double cij = 0.5 * (ci + cj);
        double hij = 0.5 * (hi + hj);
        double rhoij = 0.5 * (rhoi + rhoj);
        double r2ij = xij * xij + yij * yij;
        double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
        double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
        s.vx_t[i] += -mj * piij * w[1];
        s.vy_t[i] += -mj * piij * w[2];
      }
    }
    // correctors_xsph, src/correctors.h:139-169 (note: m of particle i)
    {
      double vxi = s.vx[i], vyi = s.vy[i], rhoi = s.rho[i];
      for (uint64_t q = R[i]; q < R[i + 1]; q++) {
        uint32_t j = NB[q];
        const double* w = W + 3 * q;
        if (skip0 && w[0] == 0.) continue;  // q >= 2: W = grad W = 0, every term adds an exact +-0
        double vijx = vxi - s.vx[j];
        double vijy = vyi - s.vy[j];
        double rhoij = 0.5 * (rhoi + s.rho[j]);
        double fac = -xeps * w[0] * s.m[i] / rhoij;
        s.x_t[i] += fac * vijx;
        s.y_t[i] += fac * vijy;
      }
    }
    // material_stress_rate_jaumann, src/material.h:66-100
    {
      double vx_x = s.vx_x[i], vx_y = s.vx_y[i], vy_x = s.vy_x[i], vy_y = s.vy_y[i];
      M2 epsdot = mk(vx_x, 0.5 * (vx_y + vy_x), 0.5 * (vx_y + vy_x), vy_y);
      M2 omega = mk(0., 0.5 * (vy_x - vx_y), 0.5 * (vx_y - vy_x), 0.);
      M2 S = mk(s.Sxx[i], s.Sxy[i], s.Sxy[i], s.Syy[i]);
      M2 I = mk(1., 0., 0., 1.);
      double tr = epsdot.m[0][0] + epsdot.m[1][1];
      M2 St = add(sub(smul(2 * G, sub(epsdot, smul(1. / 3. * tr, I))), mul(S, omega)), mul(omega, S));
      s.Sxx_t[i] += St.m[0][0];
      s.Sxy_t[i] += St.m[0][1];
      s.Syy_t[i] += St.m[1][1];
    }
    // contmech_continuity, src/cont_mech.h:108-126
    s.rho_t[i] -= s.rho[i] * (s.vx_x[i] + s.vy_y[i]);
    // contmech_momentum_ul, src/cont_mech.h:32-49
    s.vx_t[i] += 1. / s.rho[i] * (s.Sxx_x[i] + s.Sxy_y[i]) + s.fcx[i] / s.m[i];
    s.vy_t[i] += 1. / s.rho[i] * (s.Sxy_x[i] + s.Syy_y[i]) + s.fcy[i] / s.m[i];
    // contmech_advection, src/cont_mech.h:134-142
    s.x_t[i] += s.vx[i];
    s.y_t[i] += s.vy[i];
  }

  // un-sort (src/body.h:147): scatter results back to idx order
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t k = P[i];
    double* d = d8 + 8 * (size_t)k;
    d[0] = s.x_t[i];
    d[1] = s.y_t[i];
    d[2] = s.vx_t[i];
    d[3] = s.vy_t[i];
    d[4] = s.rho_t[i];
    d[5] = s.Sxx_t[i];
    d[6] = s.Sxy_t[i];
    d[7] = s.Syy_t[i];
    if (aux) {
      double* a = aux + 16 * (size_t)k;
      a[0] = s.qw[i];
      a[1] = s.p[i];
      a[2] = s.Rxx[i];
      a[3] = s.Rxy[i];
      a[4] = s.Ryy[i];
      a[5] = s.Sxx_x[i];
      a[6] = s.Sxy_y[i];
      a[7] = s.Sxy_x[i];
      a[8] = s.Syy_y[i];
      a[9] = s.vx_x[i];
      a[10] = s.vx_y[i];
      a[11] = s.vy_x[i];
      a[12] = s.vy_y[i];
      a[13] = (double)(R[i + 1] - R[i]);
      a[14] = 0.;
      a[15] = 0.;
    }
  }
  return ORC_OK;
}

// The actions of a Godunov (Parshikov) UL body after the quadrature weight and
// the EOS, src/benchmarks/rings.cpp:151-158: derive_velocity, Jaumann rate,
// godunov_continuity, godunov_momentum (src/godunov.h:91-189), advection.
// Each action loops over all particles in the reference; they read only
// state and the per-particle results of earlier actions, so evaluating them
// per particle i gives the same values.
void godunov_particle(const orc_consts* c, UlStage& s, const uint64_t* R, const uint32_t* NB, const double* W,
                      bool skip0, int64_t i) {
  const double K = K_of(c->E, c->nu);
  const double G = G_of(c->E, c->nu);
  const double xi = s.x[i], yi = s.y[i], vxi = s.vx[i], vyi = s.vy[i], rhoi = s.rho[i];
  // derive_velocity, src/derivatives.h:32-66
  {
    double vx_x = 0., vx_y = 0., vy_x = 0., vy_y = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      if (skip0 && w[0] == 0.) continue;  // q >= 2: W = grad W = 0, every term adds an exact +-0
      double quad_weight = s.qw[j];
      vx_x += (s.vx[j] - vxi) * w[1] * quad_weight;
      vx_y += (s.vx[j] - vxi) * w[2] * quad_weight;
      vy_x += (s.vy[j] - vyi) * w[1] * quad_weight;
      vy_y += (s.vy[j] - vyi) * w[2] * quad_weight;
    }
    s.vx_x[i] = vx_x;
    s.vx_y[i] = vx_y;
    s.vy_x[i] = vy_x;
    s.vy_y[i] = vy_y;
  }
  // material_stress_rate_jaumann, src/material.h:66-100
  {
    double vx_x = s.vx_x[i], vx_y = s.vx_y[i], vy_x = s.vy_x[i], vy_y = s.vy_y[i];
    M2 epsdot = mk(vx_x, 0.5 * (vx_y + vy_x), 0.5 * (vx_y + vy_x), vy_y);
    M2 omega = mk(0., 0.5 * (vy_x - vx_y), 0.5 * (vx_y - vy_x), 0.);
    M2 S = mk(s.Sxx[i], s.Sxy[i], s.Sxy[i], s.Syy[i]);
    M2 I = mk(1., 0., 0., 1.);
    double tr = epsdot.m[0][0] + epsdot.m[1][1];
    M2 St = add(sub(smul(2 * G, sub(epsdot, smul(1. / 3. * tr, I))), mul(S, omega)), mul(omega, S));
    s.Sxx_t[i] += St.m[0][0];
    s.Sxy_t[i] += St.m[0][1];
    s.Syy_t[i] += St.m[1][1];
  }
  // godunov_continuity, src/godunov.h:155-189, with riemann_solver_acoustic
  // (src/godunov.h:41-56); the self pair is skipped by idx (:170)
  {
    double rho_t = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      if ((int64_t)j == i) continue;
      const double* w = W + 3 * q;
      if (skip0 && w[0] == 0.) continue;  // q >= 2: grad W = 0, the term is an exact +-0
      double quad_weight = s.qw[j];
      double dx = xi - s.x[j];
      double dy = yi - s.y[j];
      double mag_r = sqrt(dx * dx + dy * dy);
      double nx = dx / mag_r;
      double ny = dy / mag_r;
      double vir = vxi * nx + vyi * ny;
      double vjr = s.vx[j] * nx + s.vy[j] * ny;
      double rhol = rhoi, rhor = s.rho[j], pl = s.p[i], pr = s.p[j];
      double csl = sqrt(K / rhol);
      double csr = sqrt(K / rhor);
      double frac = 1. / (rhol * csl + rhor * csr);
      double vstar = frac * (vir * rhol * csl + vjr * rhor * csr - pl + pr);
      rho_t += 2 * rhoi * (vir - vstar) * (w[1] * nx + w[2] * ny) * quad_weight;
    }
    s.rho_t[i] = rho_t;
  }
  // godunov_momentum, src/godunov.h:91-153, with godunov_get_trafo (:58-89).
  // glm::normalize(v) = v * (1/sqrt(dot(v, v))).  The transformation
  // t.l = glm::mat2x2(R, S) (:69) is a single-precision matrix: its entries
  // are R and S rounded to float, then widened back when assigned to the
  // double matrix.  SRijstar[1] takes sigma_j twice (:127), and sigma has
  // S_yy without the pressure (:72-73); both as the reference does.
  {
    double ci = sqrt(K / rhoi);
    double sxi = s.Sxx[i] - s.p[i], sxyi = s.Sxy[i], syi = s.Syy[i];
    double vx_t = s.vx_t[i], vy_t = s.vy_t[i];
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      if ((int64_t)j == i) continue;
      const double* w = W + 3 * q;
      if (skip0 && w[0] == 0.) continue;  // q >= 2: grad W = 0, both terms are exact +-0
      double Rx = s.x[j] - xi;
      double Ry = s.y[j] - yi;
      double Sx = Ry;
      double Sy = -Rx;
      double iR = 1. / sqrt(Rx * Rx + Ry * Ry);
      double Rnx = Rx * iR, Rny = Ry * iR;
      double iS = 1. / sqrt(Sx * Sx + Sy * Sy);
      double Snx = Sx * iS, Sny = Sy * iS;
      // columns of l: (R) and (S), single precision
      double l00 = (double)(float)Rnx, l01 = (double)(float)Rny;
      double l10 = (double)(float)Snx, l11 = (double)(float)Sny;
      double sxj = s.Sxx[j] - s.p[j], sxyj = s.Sxy[j], syj = s.Syy[j];
      // sigma * R, glm: (M v)[r] = M[0][r] v.x + M[1][r] v.y
      double siR0 = sxi * Rnx + sxyi * Rny, siR1 = sxyi * Rnx + syi * Rny;
      double sjR0 = sxj * Rnx + sxyj * Rny, sjR1 = sxyj * Rnx + syj * Rny;
      double si0 = l00 * siR0 + l10 * siR1, si1 = l01 * siR0 + l11 * siR1;
      (void)si1;  // the reference computes it, but SRijstar[1] reads sigma_j twice
      double sj0 = l00 * sjR0 + l10 * sjR1, sj1 = l01 * sjR0 + l11 * sjR1;
      double vi0 = l00 * vxi + l10 * vyi, vi1 = l01 * vxi + l11 * vyi;
      double vxj = s.vx[j], vyj = s.vy[j];
      double vj0 = l00 * vxj + l10 * vyj, vj1 = l01 * vxj + l11 * vyj;
      double cj = sqrt(K / s.rho[j]);
      double rhoj = s.rho[j];
      double quad_weight = s.qw[j];
      double Riemann_frac = 1. / (rhoi * ci + rhoj * cj);
      double SR0 = Riemann_frac * (sj0 * rhoi * ci + si0 * rhoj * cj + rhoi * ci * rhoj * cj * (vj0 - vi0));
      double SR1 = Riemann_frac * (sj1 * rhoi * ci + sj1 * rhoj * cj + rhoi * ci * rhoj * cj * (vj1 - vi1));
      // glm::transpose(l) * SR: [r] = l[r][0] SR0 + l[r][1] SR1
      double SRij0 = l00 * SR0 + l01 * SR1;
      double SRij1 = l10 * SR0 + l11 * SR1;
      double rx = s.x[j] - xi;
      double ry = s.y[j] - yi;
      double mag_r = sqrt(rx * rx + ry * ry);
      if (fabs(rx) > 0.) vx_t += 2. / rhoi * w[1] * mag_r / rx * SRij0 * quad_weight;
      if (fabs(ry) > 0.) vy_t += 2. / rhoi * w[2] * mag_r / ry * SRij1 * quad_weight;
    }
    s.vx_t[i] = vx_t + s.fcx[i] / s.m[i];
    s.vy_t[i] = vy_t + s.fcy[i] / s.m[i];
  }
  // contmech_advection, src/cont_mech.h:134-142
  s.x_t[i] += s.vx[i];
  s.y_t[i] += s.vy[i];
}

// The 8 integrated fields of a state, src/rk4.h:52-69.
struct State8 {
  std::vector<double> x, y, vx, vy, rho, Sxx, Sxy, Syy;
  void resize(uint32_t n) {
    for (auto* v : {&x, &y, &vx, &vy, &rho, &Sxx, &Sxy, &Syy}) v->assign(n, 0.);
  }
};

// k_out = y + dt_s * k_in derivatives (rk_add, src/rk4.h:52-69)
void rk_add(uint32_t n, const orc_soa* y, const double* d8, double dts, orc_soa* k, int nt) {
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const double* d = d8 + 8 * (size_t)i;
    k->x[i] = y->x[i] + d[0] * dts;
    k->y[i] = y->y[i] + d[1] * dts;
    k->vx[i] = y->vx[i] + d[2] * dts;
    k->vy[i] = y->vy[i] + d[3] * dts;
    k->rho[i] = y->rho[i] + d[4] * dts;
    k->Sxx[i] = y->Sxx[i] + d[5] * dts;
    k->Sxy[i] = y->Sxy[i] + d[6] * dts;
    k->Syy[i] = y->Syy[i] + d[7] * dts;
  }
}

// y += dt*1./6.*(k1 + 2*k2 + 2*k3 + k4) (src/rk4.h:94-103)
void rk_combine(uint32_t n, orc_soa* y, const double* k1, const double* k2, const double* k3,
                const double* k4, double dt, int nt) {
  double* f[8] = {y->x, y->y, y->vx, y->vy, y->rho, y->Sxx, y->Sxy, y->Syy};
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    for (int c = 0; c < 8; c++) {
      size_t o = 8 * (size_t)i + c;
      f[c][i] += dt * 1. / 6. * (k1[o] + 2 * k2[o] + 2 * k3[o] + k4[o]);
    }
  }
}

// A stage copy of a state: own storage for the 8 integrated fields, shared
// pointers for the constant ones.
struct StageCopy {
  State8 v;
  orc_soa soa;
  void init(uint32_t n, const orc_soa* y) {
    v.resize(n);
    soa = *y;
    soa.x = v.x.data();
    soa.y = v.y.data();
    soa.vx = v.vx.data();
    soa.vy = v.vy.data();
    soa.rho = v.rho.data();
    soa.Sxx = v.Sxx.data();
    soa.Sxy = v.Sxy.data();
    soa.Syy = v.Syy.data();
  }
  void copy_from(uint32_t n, const orc_soa* y) {
    memcpy(soa.x, y->x, n * sizeof(double));
    memcpy(soa.y, y->y, n * sizeof(double));
    memcpy(soa.vx, y->vx, n * sizeof(double));
    memcpy(soa.vy, y->vy, n * sizeof(double));
    memcpy(soa.rho, y->rho, n * sizeof(double));
    memcpy(soa.Sxx, y->Sxx, n * sizeof(double));
    memcpy(soa.Sxy, y->Sxy, n * sizeof(double));
    memcpy(soa.Syy, y->Syy, n * sizeof(double));
  }
};

}  // namespace

// ==========================================================================
// Total-Lagrangian body
struct orc_tl {
  uint32_t n;
  int variant = ORC_TL_RL;
  std::vector<uint64_t> rowp;
  std::vector<uint32_t> nbh;  // idx space
  std::vector<double> w;      // corrected (RL, potential) or plain (co-rotated) kernel, 3 per pair
  std::vector<double> wji;    // potential bodies: the kernel of the reverse pair, 3 per pair
  std::vector<double> C, B4;
  std::vector<double> qw;     // quad_weight = m/rho at creation (src/benchmarks/tensile.cpp:262)
  std::vector<uint32_t> bc_left, bc_right;
  double v_bc = 0.;
};

extern "C" {

void orc_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]) {
  spline(xi, yi, xj, yj, h, out);
}

double orc_G(double E, double nu) { return G_of(E, nu); }
double orc_K(double E, double nu) { return K_of(E, nu); }

int orc_grid_geometry(uint32_t n, const double* x, const double* y, const double* h, double kernel_width,
                      orc_geom* g) {
  if (n == 0 || !x || !y || !h || !g) return ORC_ERR_ARG;
  return geometry(n, x, y, h, kernel_width, g);
}

int orc_verlet_pairs(uint32_t n, const double* x, const double* y, const double* h, int order,
                     uint32_t max_nbh, uint32_t* row_ptr, uint32_t* col, uint64_t cap, orc_geom* g) {
  if (n == 0 || !x || !y || !h || !row_ptr) return ORC_ERR_ARG;
  UlStage s;
  s.hash.assign(n, 0);
  s.perm.resize(n);
  s.rowp.assign((size_t)n + 1, 0);
  int rc = build_lists(n, x, y, h, order, max_nbh, s, g);
  if (rc) return rc;
  if (s.nbh.size() > cap || !col) return ORC_ERR_CAPACITY;
  // convert to idx space: row of idx k = row of its sorted position
  std::vector<uint32_t> cnt(n);
  for (uint32_t i = 0; i < n; i++) cnt[s.perm[i]] = (uint32_t)(s.rowp[i + 1] - s.rowp[i]);
  row_ptr[0] = 0;
  for (uint32_t k = 0; k < n; k++) row_ptr[k + 1] = row_ptr[k] + cnt[k];
  for (uint32_t i = 0; i < n; i++) {
    uint32_t k = s.perm[i];
    uint32_t o = row_ptr[k];
    for (uint64_t q = s.rowp[i]; q < s.rowp[i + 1]; q++) col[o++] = s.perm[s.nbh[q]];
  }
  return ORC_OK;
}

int orc_find_neighbors(uint32_t n, const double* x, const double* y, double h, uint32_t max_nbh,
                       uint32_t* row_ptr, uint32_t* col, uint64_t cap) {
  // src/benchmarks/neighbor_search.h:116-147, distance_r2 :32-41
  if (n == 0 || !x || !y || !row_ptr) return ORC_ERR_ARG;
  uint64_t o = 0;
  row_ptr[0] = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t cnt = 0;
    double xi = x[i], yi = y[i];
    for (uint32_t j = 0; j < n; j++) {
      double xij = xi - x[j];
      double yij = yi - y[j];
      double r2 = xij * xij + yij * yij;
      double radius2 = h * h * 2 * 2;
      if (r2 <= radius2 + 1e-6) {
        if (o >= cap || !col) return ORC_ERR_CAPACITY;
        col[o++] = j;
        cnt++;
      }
    }
    if (cnt >= max_nbh) return ORC_ERR_NBH_OVERFLOW;
    if (cnt == 0) return ORC_ERR_NO_NEIGHBOR;
    row_ptr[i + 1] = (uint32_t)o;
  }
  return ORC_OK;
}

int orc_ul_derivatives(const orc_consts* c, uint32_t n, const orc_soa* s, int order, uint32_t max_nbh,
                       double* d8, double* aux) {
  return orc_ul_derivatives_m(c, ORC_METHOD_MONAGHAN, n, s, order, max_nbh, d8, aux);
}

int orc_ul_derivatives_m(const orc_consts* c, int method, uint32_t n, const orc_soa* s, int order,
                         uint32_t max_nbh, double* d8, double* aux) {
  if (!c || !s || !d8 || n < 2 || (method != ORC_METHOD_MONAGHAN && method != ORC_METHOD_GODUNOV))
    return ORC_ERR_ARG;
  UlStage st;
  resize_all(st, n);
  return ul_stage(c, method, n, s, order, max_nbh, st, 1, d8, aux);
}

int orc_ul_step(const orc_consts* c, uint32_t n, orc_soa* s, double dt, uint32_t nsteps, int order,
                uint32_t max_nbh, int threads) {
  return orc_ul_step_m(c, ORC_METHOD_MONAGHAN, n, s, dt, nsteps, order, max_nbh, threads);
}

int orc_ul_step_m(const orc_consts* c, int method, uint32_t n, orc_soa* s, double dt, uint32_t nsteps, int order,
                  uint32_t max_nbh, int threads) {
  // src/rk4.h:71-108 (step_strong) for a particle_monaghan / particle_ul body
  if (!c || !s || n < 2 || (method != ORC_METHOD_MONAGHAN && method != ORC_METHOD_GODUNOV)) return ORC_ERR_ARG;
  int nt = set_threads(threads);
  UlStage st;
  resize_all(st, n);
  StageCopy k;
  k.init(n, s);
  std::vector<double> d1(8 * (size_t)n), d2(8 * (size_t)n), d3(8 * (size_t)n), d4(8 * (size_t)n);
  for (uint32_t step = 0; step < nsteps; step++) {
    k.copy_from(n, s);
    int rc = ul_stage(c, method, n, &k.soa, order, max_nbh, st, nt, d1.data(), nullptr);
    if (rc) return rc;
    rk_add(n, s, d1.data(), 0.5 * dt, &k.soa, nt);
    if ((rc = ul_stage(c, method, n, &k.soa, order, max_nbh, st, nt, d2.data(), nullptr))) return rc;
    rk_add(n, s, d2.data(), 0.5 * dt, &k.soa, nt);
    if ((rc = ul_stage(c, method, n, &k.soa, order, max_nbh, st, nt, d3.data(), nullptr))) return rc;
    rk_add(n, s, d3.data(), dt, &k.soa, nt);
    if ((rc = ul_stage(c, method, n, &k.soa, order, max_nbh, st, nt, d4.data(), nullptr))) return rc;
    rk_combine(n, s, d1.data(), d2.data(), d3.data(), d4.data(), dt, nt);
    for (uint32_t i = 0; i < n; i++)
      if (!isfinite(s->x[i]) || !isfinite(s->y[i]) || !isfinite(s->vx[i]) || !isfinite(s->vy[i]))
        return ORC_ERR_NONFINITE;
  }
  return ORC_OK;
}

// --------------------------------------------------------------------------
// TL

int orc_tl_create(uint32_t n, const double* x, const double* y, const double* m, const double* rho,
                  const double* h, double h_search, int search, int contrib_only, uint32_t max_nbh,
                  orc_tl** out) {
  return orc_tl_create_m(ORC_TL_RL, n, x, y, m, rho, h, h_search, search, contrib_only, max_nbh, out);
}

int orc_tl_create_m(int variant, uint32_t n, const double* x, const double* y, const double* m, const double* rho,
                    const double* h, double h_search, int search, int contrib_only, uint32_t max_nbh,
                    orc_tl** out) {
  if (!out || n == 0 || !x || !y || !m || !rho || !h) return ORC_ERR_ARG;
  if (variant != ORC_TL_RL && variant != ORC_TL_COROTATIONAL && variant != ORC_TL_POTENTIAL) return ORC_ERR_ARG;
  orc_tl* t = new orc_tl();
  t->n = n;
  t->variant = variant;
  t->rowp.assign((size_t)n + 1, 0);
  t->C.assign(n, 0.);
  t->B4.assign(4 * (size_t)n, 0.);
  t->qw.resize(n);
  for (uint32_t i = 0; i < n; i++) t->qw[i] = m[i] / rho[i];
  const double radius2 = h_search * h_search * 2 * 2;
  auto keep = [&](uint32_t i, uint32_t j) -> bool {
    double xij = x[i] - x[j];
    double yij = y[i] - y[j];
    double r2 = xij * xij + yij * yij;
    if (!(r2 <= radius2 + 1e-6)) return false;
    if (contrib_only) {
      double h1 = 1. / h[i];
      double q = sqrt(xij * xij + yij * yij) * h1;
      if (q >= 2.) return false;
    }
    return true;
  };
  if (search == 0) {
    for (uint32_t i = 0; i < n; i++) {
      for (uint32_t j = 0; j < n; j++)
        if (keep(i, j)) t->nbh.push_back(j);
      t->rowp[i + 1] = t->nbh.size();
    }
  } else {
    // cell list over the same predicate; candidates visited then sorted by
    // idx so that the row equals the O(N^2) scan order.
    double hmax = h_search;
    for (uint32_t i = 0; i < n; i++) hmax = fmax(hmax, h[i]);
    double rad = contrib_only ? 2.0 * hmax * (1.0 + 1e-9) : sqrt(radius2 + 1e-6) * (1.0 + 1e-12);
    double bx = 1e300, by = 1e300, ex = -1e300, ey = -1e300;
    for (uint32_t i = 0; i < n; i++) {
      bx = fmin(bx, x[i]);
      by = fmin(by, y[i]);
      ex = fmax(ex, x[i]);
      ey = fmax(ey, y[i]);
    }
    uint32_t nx = (uint32_t)floor((ex - bx) / rad) + 1, ny = (uint32_t)floor((ey - by) / rad) + 1;
    std::vector<uint32_t> cc((size_t)nx * ny + 1, 0), cell(n), ord(n);
    for (uint32_t i = 0; i < n; i++) {
      uint32_t ix = (uint32_t)((x[i] - bx) / rad), iy = (uint32_t)((y[i] - by) / rad);
      if (ix >= nx) ix = nx - 1;
      if (iy >= ny) iy = ny - 1;
      cell[i] = ix * ny + iy;
      cc[cell[i] + 1]++;
    }
    for (size_t c = 0; c < (size_t)nx * ny; c++) cc[c + 1] += cc[c];
    std::vector<uint32_t> fill(cc.begin(), cc.end() - 1);
    for (uint32_t i = 0; i < n; i++) ord[fill[cell[i]]++] = i;
    std::vector<uint32_t> cand;
    for (uint32_t i = 0; i < n; i++) {
      cand.clear();
      int ix = (int)(cell[i] / ny), iy = (int)(cell[i] % ny);
      for (int a = ix - 1; a <= ix + 1; a++) {
        if (a < 0 || a >= (int)nx) continue;
        for (int b = iy - 1; b <= iy + 1; b++) {
          if (b < 0 || b >= (int)ny) continue;
          size_t c = (size_t)a * ny + b;
          for (uint32_t q = cc[c]; q < cc[c + 1]; q++)
            if (keep(i, ord[q])) cand.push_back(ord[q]);
        }
      }
      std::sort(cand.begin(), cand.end());
      t->nbh.insert(t->nbh.end(), cand.begin(), cand.end());
      t->rowp[i + 1] = t->nbh.size();
    }
  }
  for (uint32_t i = 0; i < n; i++) {
    uint64_t cnt = t->rowp[i + 1] - t->rowp[i];
    if (cnt >= max_nbh) {
      delete t;
      return ORC_ERR_NBH_OVERFLOW;
    }
    if (cnt == 0) {
      delete t;
      return ORC_ERR_NO_NEIGHBOR;
    }
  }
  t->w.assign(3 * t->nbh.size(), 0.);
  if (variant == ORC_TL_COROTATIONAL) {
    // precomp_sph, src/precomp_shape_functions.h:210-229 (tensile_corot,
    // src/benchmarks/tensile.cpp:197-198): the plain kernel at the set-up
    // positions
    for (uint32_t i = 0; i < n; i++)
      for (uint64_t q = t->rowp[i]; q < t->rowp[i + 1]; q++) spline(x[i], y[i], x[t->nbh[q]], y[t->nbh[q]], h[i],
                                                                    t->w.data() + 3 * q);
    *out = t;
    return ORC_OK;
  }
  // precomp_randles_libersky, src/precomp_shape_functions.h:231-290
  for (uint32_t i = 0; i < n; i++) {
    double xi = x[i], yi = y[i], hi = h[i];
    double C = 0.;
    for (uint64_t q = t->rowp[i]; q < t->rowp[i + 1]; q++) {
      uint32_t j = t->nbh[q];
      double volj = m[j] / rho[j];
      double w[3];
      spline(xi, yi, x[j], y[j], hi, w);
      C += w[0] * volj;
    }
    M2 B = mk(0., 0., 0., 0.);
    for (uint64_t q = t->rowp[i]; q < t->rowp[i + 1]; q++) {
      uint32_t j = t->nbh[q];
      double xj = x[j], yj = y[j];
      double volj = m[j] / rho[j];
      double w[3];
      spline(xi, yi, xj, yj, hi, w);
      B.m[0][0] += (xj - xi) * w[1] * volj * 1. / C;
      B.m[1][0] += (xj - xi) * w[2] * volj * 1. / C;
      B.m[0][1] += (yj - yi) * w[1] * volj * 1. / C;
      B.m[1][1] += (yj - yi) * w[2] * volj * 1. / C;
    }
    M2 invB = inv(B);
    t->C[i] = C;
    t->B4[4 * (size_t)i + 0] = B.m[0][0];
    t->B4[4 * (size_t)i + 1] = B.m[0][1];
    t->B4[4 * (size_t)i + 2] = B.m[1][0];
    t->B4[4 * (size_t)i + 3] = B.m[1][1];
    for (uint64_t q = t->rowp[i]; q < t->rowp[i + 1]; q++) {
      uint32_t j = t->nbh[q];
      double w[3];
      spline(xi, yi, x[j], y[j], hi, w);
      t->w[3 * q + 0] = w[0] * 1. / C;
      t->w[3 * q + 1] = (w[1] * invB.m[0][0] + w[2] * invB.m[1][0]) * 1. / C;
      t->w[3 * q + 2] = (w[1] * invB.m[0][1] + w[2] * invB.m[1][1]) * 1. / C;
    }
  }
  if (variant == ORC_TL_POTENTIAL) {
    // precomp_randles_libersky_symmetric, src/precomp_shape_functions.h:292-316:
    // wji of pair (i, j) is j's kernel of its first pair back to i.  A
    // reverse pair dropped from the lists (q >= 2 with h_j) has a zero kernel.
    t->wji.assign(3 * t->nbh.size(), 0.);
    for (uint32_t i = 0; i < n; i++)
      for (uint64_t q = t->rowp[i]; q < t->rowp[i + 1]; q++) {
        uint32_t j = t->nbh[q];
        for (uint64_t k = t->rowp[j]; k < t->rowp[j + 1]; k++)
          if (t->nbh[k] == i) {
            for (int c = 0; c < 3; c++) t->wji[3 * q + c] = t->w[3 * k + c];
            break;
          }
      }
  }
  *out = t;
  return ORC_OK;
}

void orc_tl_destroy(orc_tl* t) { delete t; }

uint64_t orc_tl_num_pairs(const orc_tl* t) { return t ? t->nbh.size() : 0; }

void orc_tl_get(const orc_tl* t, uint32_t* row_ptr, uint32_t* col, double* w) {
  for (uint32_t i = 0; i <= t->n; i++) row_ptr[i] = (uint32_t)t->rowp[i];
  if (col) memcpy(col, t->nbh.data(), t->nbh.size() * sizeof(uint32_t));
  if (w) memcpy(w, t->w.data(), t->w.size() * sizeof(double));
}

void orc_tl_get_moments(const orc_tl* t, double* C, double* B4) {
  if (C) memcpy(C, t->C.data(), t->n * sizeof(double));
  if (B4) memcpy(B4, t->B4.data(), 4 * (size_t)t->n * sizeof(double));
}

int orc_tl_set_bc(orc_tl* t, const uint32_t* idx_left, uint32_t nl, const uint32_t* idx_right, uint32_t nr,
                  double v_bc) {
  if (!t) return ORC_ERR_ARG;
  t->bc_left.assign(idx_left, idx_left + nl);
  t->bc_right.assign(idx_right, idx_right + nr);
  t->v_bc = v_bc;
  return ORC_OK;
}

}  // extern "C"

namespace {

// boundary_conditions::carry_out_boundary_conditions (src/boundary_conditions.h:36-46)
// with left/right_bndry_fun_tl (src/benchmarks/tensile.cpp:20-42).
void tl_apply_bc(const orc_tl* t, orc_soa* s, double time, double* d8) {
  for (uint32_t k : t->bc_left) {
    s->x[k] = s->X[k] - t->v_bc * time;
    s->y[k] = s->Y[k];
    s->vx[k] = -t->v_bc;
    s->vy[k] = 0.;
    if (d8) {
      d8[8 * (size_t)k + 2] = 0.;
      d8[8 * (size_t)k + 3] = 0.;
    }
  }
  for (uint32_t k : t->bc_right) {
    s->x[k] = s->X[k] + t->v_bc * time;
    s->y[k] = s->Y[k];
    s->vx[k] = t->v_bc;
    s->vy[k] = 0.;
    if (d8) {
      d8[8 * (size_t)k + 2] = 0.;
      d8[8 * (size_t)k + 3] = 0.;
    }
  }
}

struct TlWork {
  std::vector<double> Hxx, Hxy, Hyx, Hyy, Fxx, Fxy, Fyx, Fyy, vx_x, vx_y, vy_x, vy_y, p, Pxx, Pxy, Pyx, Pyy,
      Pxx_x, Pxy_x, Pyx_y, Pyy_y;
  void resize(uint32_t n) {
    for (auto* v : {&Hxx, &Hxy, &Hyx, &Hyy, &Fxx, &Fxy, &Fyx, &Fyy, &vx_x, &vx_y, &vy_x, &vy_y, &p, &Pxx, &Pxy,
                    &Pyx, &Pyy, &Pxx_x, &Pxy_x, &Pyx_y, &Pyy_y})
      v->assign(n, 0.);
  }
};

int tl_stage_corot(const orc_consts* c, const orc_tl* t, uint32_t n, orc_soa* s, TlWork& wk, int nt, double* d8,
                   double* aux);
int tl_stage_potential(const orc_consts* c, const orc_tl* t, uint32_t n, orc_soa* s, TlWork& wk, int nt,
                       double* d8, double* aux);

int tl_stage(const orc_consts* c, const orc_tl* t, uint32_t n, orc_soa* s, double time, TlWork& wk, int nt,
             double* d8, double* aux) {
  // src/body.h:110-123: reset, BCs, actions (src/benchmarks/tensile.cpp:276-286)
  for (size_t i = 0; i < 8 * (size_t)n; i++) d8[i] = 0.;
  tl_apply_bc(t, s, time, d8);
  if (t->variant == ORC_TL_COROTATIONAL) return tl_stage_corot(c, t, n, s, wk, nt, d8, aux);
  if (t->variant == ORC_TL_POTENTIAL) return tl_stage_potential(c, t, n, s, wk, nt, d8, aux);
  const uint64_t* R = t->rowp.data();
  const uint32_t* NB = t->nbh.data();
  const double* W = t->w.data();
  const double* QW = t->qw.data();
  const double K = K_of(c->E, c->nu), G = G_of(c->E, c->nu), rho0 = c->rho0;
  // pass 1: derive_displacement (src/derivatives.h:101-149),
  //         derive_velocity_tl (src/derivatives.h:483-519),
  //         contmech_velocity_gradient_tl (src/cont_mech.h:188-203),
  //         contmech_continuity, material_eos, material_stress_rate_jaumann,
  //         contmech_cauchy_to_nominal (src/cont_mech.h:154-169)
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double uxi = s->x[i] - s->X[i], uyi = s->y[i] - s->Y[i];
    double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double uxj = s->x[j] - s->X[j], uyj = s->y[j] - s->Y[j];
      double qw = QW[j];
      Hxx += (uxj - uxi) * w[1] * qw;
      Hxy += (uxj - uxi) * w[2] * qw;
      Hyx += (uyj - uyi) * w[1] * qw;
      Hyy += (uyj - uyi) * w[2] * qw;
    }
    double vxi = s->vx[i], vyi = s->vy[i];
    double Fxx = 0., Fxy = 0., Fyx = 0., Fyy = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double qw = QW[j];
      Fxx += (s->vx[j] - vxi) * w[1] * qw;
      Fxy += (s->vx[j] - vxi) * w[2] * qw;
      Fyx += (s->vy[j] - vyi) * w[1] * qw;
      Fyy += (s->vy[j] - vyi) * w[2] * qw;
    }
    wk.Hxx[i] = Hxx;
    wk.Hxy[i] = Hxy;
    wk.Hyx[i] = Hyx;
    wk.Hyy[i] = Hyy;
    wk.Fxx[i] = Fxx;
    wk.Fxy[i] = Fxy;
    wk.Fyx[i] = Fyx;
    wk.Fyy[i] = Fyy;
    M2 F = mk(Hxx + 1., Hyx, Hxy, Hyy + 1.);
    M2 Fdot = mk(Fxx, Fyx, Fxy, Fyy);
    M2 L = mul(Fdot, inv(F));
    double vx_x = L.m[0][0], vx_y = L.m[1][0], vy_x = L.m[0][1], vy_y = L.m[1][1];
    wk.vx_x[i] = vx_x;
    wk.vx_y[i] = vx_y;
    wk.vy_x[i] = vy_x;
    wk.vy_y[i] = vy_y;
    double* d = d8 + 8 * (size_t)i;
    d[4] -= s->rho[i] * (vx_x + vy_y);
    double c0 = sqrt(K / rho0);
    double p = c0 * c0 * (s->rho[i] - rho0);
    wk.p[i] = p;
    M2 epsdot = mk(vx_x, 0.5 * (vx_y + vy_x), 0.5 * (vx_y + vy_x), vy_y);
    M2 omega = mk(0., 0.5 * (vy_x - vx_y), 0.5 * (vx_y - vy_x), 0.);
    M2 S = mk(s->Sxx[i], s->Sxy[i], s->Sxy[i], s->Syy[i]);
    M2 I = mk(1., 0., 0., 1.);
    double tr = epsdot.m[0][0] + epsdot.m[1][1];
    M2 St = add(sub(smul(2 * G, sub(epsdot, smul(1. / 3. * tr, I))), mul(S, omega)), mul(omega, S));
    d[5] += St.m[0][0];
    d[6] += St.m[0][1];
    d[7] += St.m[1][1];
    M2 F2 = mk(Hxx + 1, Hyx, Hxy, Hyy + 1);
    double J = det(F2);
    M2 sigma = mk(s->Sxx[i] - p, s->Sxy[i], s->Sxy[i], s->Syy[i] - p);
    M2 Pm = mul(smul(J, inv(F2)), sigma);
    wk.Pxx[i] = Pm.m[0][0];
    wk.Pxy[i] = Pm.m[1][0];
    wk.Pyx[i] = Pm.m[0][1];
    wk.Pyy[i] = Pm.m[1][1];
  }
  // pass 2: derive_stress_tl (src/derivatives.h:331-376),
  //         correctors_mghn_artificial_viscosity (src/correctors.h:83-137),
  //         contmech_momentum_tl (src/cont_mech.h:70-87), contmech_advection
  const double alpha = c->alpha, beta = c->beta, eta = c->eta;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double Pxxi = wk.Pxx[i], Pxyi = wk.Pxy[i], Pyxi = wk.Pyx[i], Pyyi = wk.Pyy[i];
    double rhoi21 = 1. / (rho0 * rho0);
    double Pxx_x = 0., Pyx_y = 0., Pxy_x = 0., Pyy_y = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double mj = s->m[j];
      double rhoj21 = 1. / (rho0 * rho0);
      Pxx_x += mj * (Pxxi * rhoi21 + wk.Pxx[j] * rhoj21) * w[1] * rho0;
      Pyx_y += mj * (Pyxi * rhoi21 + wk.Pyx[j] * rhoj21) * w[2] * rho0;
      Pxy_x += mj * (Pxyi * rhoi21 + wk.Pxy[j] * rhoj21) * w[1] * rho0;
      Pyy_y += mj * (Pyyi * rhoi21 + wk.Pyy[j] * rhoj21) * w[2] * rho0;
    }
    wk.Pxx_x[i] = Pxx_x;
    wk.Pyx_y[i] = Pyx_y;
    wk.Pxy_x[i] = Pxy_x;
    wk.Pyy_y[i] = Pyy_y;
    double* d = d8 + 8 * (size_t)i;
    double xi = s->x[i], yi = s->y[i], vxi = s->vx[i], vyi = s->vy[i], hi = s->h[i], rhoi = s->rho[i];
    double ci = sqrt(K / rhoi);
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double xj = s->x[j], yj = s->y[j], vxj = s->vx[j], vyj = s->vy[j], hj = s->h[j];
      double mj = s->m[j], rhoj = s->rho[j];
      double cj = sqrt(K / rhoj);
      double vijx = vxi - vxj, vijy = vyi - vyj;
      double xij = xi - xj, yij = yi - yj;
      double vijposij = vijx * xij + vijy * yij;
      if (vijposij >= 0.) continue;
      double cij = 0.5 * (ci + cj);
      double hij = 0.5 * (hi + hj);
      double rhoij = 0.5 * (rhoi + rhoj);
      double r2ij = xij * xij + yij * yij;
      double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
      double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
      d[2] += -mj * piij * w[1];
      d[3] += -mj * piij * w[2];
    }
    double fcx = s->fcx ? s->fcx[i] : 0., fcy = s->fcy ? s->fcy[i] : 0.;
    d[2] += 1. / rho0 * (Pxx_x + Pyx_y) + fcx / s->m[i];
    d[3] += 1. / rho0 * (Pxy_x + Pyy_y) + fcy / s->m[i];
    d[0] += s->vx[i];
    d[1] += s->vy[i];
    if (aux) {
      double* a = aux + 16 * (size_t)i;
      a[0] = wk.Hxx[i];
      a[1] = wk.Hxy[i];
      a[2] = wk.Hyx[i];
      a[3] = wk.Hyy[i];
      a[4] = wk.Fxx[i];
      a[5] = wk.Fxy[i];
      a[6] = wk.Fyx[i];
      a[7] = wk.Fyy[i];
      a[8] = wk.Pxx[i];
      a[9] = wk.Pxy[i];
      a[10] = wk.Pyx[i];
      a[11] = wk.Pyy[i];
      a[12] = Pxx_x;
      a[13] = Pxy_x;
      a[14] = Pyx_y;
      a[15] = Pyy_y;
    }
  }
  return ORC_OK;
}

// correctors_mghn_artificial_viscosity (src/correctors.h:83-137) of particle
// i over its TL list, added to d[2], d[3].
void tl_viscosity(const orc_consts* c, const orc_tl* t, const orc_soa* s, int64_t i, double* d) {
  const uint64_t* R = t->rowp.data();
  const uint32_t* NB = t->nbh.data();
  const double* W = t->w.data();
  const double K = K_of(c->E, c->nu);
  const double alpha = c->alpha, beta = c->beta, eta = c->eta;
  double xi = s->x[i], yi = s->y[i], vxi = s->vx[i], vyi = s->vy[i], hi = s->h[i], rhoi = s->rho[i];
  double ci = sqrt(K / rhoi);
  for (uint64_t q = R[i]; q < R[i + 1]; q++) {
    uint32_t j = NB[q];
    const double* w = W + 3 * q;
    double xj = s->x[j], yj = s->y[j], vxj = s->vx[j], vyj = s->vy[j], hj = s->h[j];
    double mj = s->m[j], rhoj = s->rho[j];
    double cj = sqrt(K / rhoj);
    double vijx = vxi - vxj, vijy = vyi - vyj;
    double xij = xi - xj, yij = yi - yj;
    double vijposij = vijx * xij + vijy * yij;
    if (vijposij >= 0.) continue;
    double cij = 0.5 * (ci + cj);
    double hij = 0.5 * (hi + hj);
    double rhoij = 0.5 * (rhoi + rhoj);
    double r2ij = xij * xij + yij * yij;
    double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
    double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
    d[2] += -mj * piij * w[1];
    d[3] += -mj * piij * w[2];
  }
}

// A co-rotated TL body, tensile_corot (src/benchmarks/tensile.cpp:137-211):
// corotational_anisotropy_tensor (src/corotational.h:32-85),
// derive_displacement_corotational (src/derivatives.h:202-258),
// material_compute_stress_cauchy_green (src/material.h:102-130),
// contmech_cauchy_to_nominal (src/cont_mech.h:154-169; p keeps its initial
// value 0: no action of this list sets it), derive_stress_corotational
// (src/derivatives.h:521-588), correctors_mghn_artificial_viscosity,
// contmech_advection.  No continuity or stress rate: rho_t = S_t = 0, and the
// stage's S is recomputed from H.
// aux: Hxx, Hxy, Hyx, Hyy, Rxx, Rxy, Ryx, Ryy, Pxx, Pxy, Pyx, Pyy, Sxx, Sxy, Syy, 0.
int tl_stage_corot(const orc_consts* c, const orc_tl* t, uint32_t n, orc_soa* s, TlWork& wk, int nt, double* d8,
                   double* aux) {
  const uint64_t* R = t->rowp.data();
  const uint32_t* NB = t->nbh.data();
  const double* W = t->w.data();
  const double* QW = t->qw.data();
  const double E = c->E, nu = c->nu, rho0 = c->rho0;
  // per particle: rotation (stored in Fxx..Fyy), H, S (in vx_x, vx_y, vy_y), P
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double Xi = s->X[i], Yi = s->Y[i], xi = s->x[i], yi = s->y[i];
    double Axx = 0., Axy = 0., Ayx = 0., Ayy = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double mj = s->m[j];
      double xji = s->x[j] - xi, yji = s->y[j] - yi;
      double Xji = s->X[j] - Xi, Yji = s->Y[j] - Yi;
      Axx += xji * Xji * mj * w[0];
      Axy += xji * Yji * mj * w[0];
      Ayx += yji * Xji * mj * w[0];
      Ayy += yji * Yji * mj * w[0];
    }
    double ax = Axx + Ayy;
    double ay = Ayx - Axy;
    double theta = atan2(ay, ax);
    double Rxx = cos(theta), Rxy = -sin(theta), Ryx = sin(theta), Ryy = cos(theta);
    wk.Fxx[i] = Rxx;
    wk.Fxy[i] = Rxy;
    wk.Fyx[i] = Ryx;
    wk.Fyy[i] = Ryy;
    // derive_displacement_corotational: R_i^-1 = R_i^T
    double Rixx_inv = Rxx, Rixy_inv = Ryx, Riyx_inv = Rxy, Riyy_inv = Ryy;
    double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double quad_weight = QW[j];
      double xji = s->x[j] - xi, yji = s->y[j] - yi;
      double Xji = s->X[j] - Xi, Yji = s->Y[j] - Yi;
      double uxji = Rixx_inv * xji + Rixy_inv * yji - Xji;
      double uyji = Riyx_inv * xji + Riyy_inv * yji - Yji;
      Hxx += uxji * w[1] * quad_weight;
      Hxy += uxji * w[2] * quad_weight;
      Hyx += uyji * w[1] * quad_weight;
      Hyy += uyji * w[2] * quad_weight;
    }
    wk.Hxx[i] = Hxx;
    wk.Hxy[i] = Hxy;
    wk.Hyx[i] = Hyx;
    wk.Hyy[i] = Hyy;
    // material_compute_stress_cauchy_green
    double D = E / (1 - nu * nu);
    M2 F = mk(Hxx + 1., Hyx, Hxy, Hyy + 1.);
    M2 FT = mk(F.m[0][0], F.m[1][0], F.m[0][1], F.m[1][1]);
    M2 Ecg = sub(smul(0.5, add(F, FT)), mk(1., 0., 0., 1.));
    double Sxx = D * (Ecg.m[0][0] + nu * Ecg.m[1][1]);
    double Syy = D * (Ecg.m[0][0] * nu + Ecg.m[1][1]);
    double Sxy = D * (1 - nu) * Ecg.m[0][1];
    wk.vx_x[i] = Sxx;
    wk.vx_y[i] = Sxy;
    wk.vy_y[i] = Syy;
    // contmech_cauchy_to_nominal with p = 0
    double p = 0.;
    M2 F2 = mk(Hxx + 1, Hyx, Hxy, Hyy + 1);
    double J = det(F2);
    M2 sigma = mk(Sxx - p, Sxy, Sxy, Syy - p);
    M2 Pm = mul(smul(J, inv(F2)), sigma);
    wk.Pxx[i] = Pm.m[0][0];
    wk.Pxy[i] = Pm.m[1][0];
    wk.Pyx[i] = Pm.m[0][1];
    wk.Pyy[i] = Pm.m[1][1];
  }
  // derive_stress_corotational (assigns the accelerations, then adds
  // f_c/m), viscosity, advection
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double Pxxi = wk.Pxx[i], Pxyi = wk.Pxy[i], Pyxi = wk.Pyx[i], Pyyi = wk.Pyy[i];
    double Rxxi = wk.Fxx[i], Rxyi = wk.Fxy[i], Ryxi = wk.Fyx[i], Ryyi = wk.Fyy[i];
    double rhoi21 = 1. / (rho0 * rho0);
    double vx_t = 0., vy_t = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double Pxxj = wk.Pxx[j], Pxyj = wk.Pxy[j], Pyxj = wk.Pyx[j], Pyyj = wk.Pyy[j];
      double Rxxj = wk.Fxx[j], Rxyj = wk.Fxy[j], Ryxj = wk.Fyx[j], Ryyj = wk.Fyy[j];
      double mj = s->m[j];
      double rhoj21 = 1. / (rho0 * rho0);
      double aijx = mj * ((Pxxi * rhoi21 + Pxxj * rhoj21) * w[1] + (Pxyi * rhoi21 + Pxyj * rhoj21) * w[2]);
      double aijy = mj * ((Pyxi * rhoi21 + Pyxj * rhoj21) * w[1] + (Pyyi * rhoi21 + Pyyj * rhoj21) * w[2]);
      double ajix = -aijx;
      double ajiy = -aijy;
      double aijxprime = aijx * Rxxi + aijy * Rxyi;
      double aijyprime = aijx * Ryxi + aijy * Ryyi;
      double ajixprime = ajix * Rxxj + ajiy * Rxyj;
      double ajiyprime = ajix * Ryxj + ajiy * Ryyj;
      vx_t += 0.5 * (aijxprime - ajixprime);
      vy_t += 0.5 * (aijyprime - ajiyprime);
    }
    double* d = d8 + 8 * (size_t)i;
    d[2] = vx_t;
    d[3] = vy_t;
    double fcx = s->fcx ? s->fcx[i] : 0., fcy = s->fcy ? s->fcy[i] : 0.;
    d[2] += fcx / s->m[i];
    d[3] += fcy / s->m[i];
    tl_viscosity(c, t, s, i, d);
    d[0] += s->vx[i];
    d[1] += s->vy[i];
    if (aux) {
      double* a = aux + 16 * (size_t)i;
      a[0] = wk.Hxx[i];
      a[1] = wk.Hxy[i];
      a[2] = wk.Hyx[i];
      a[3] = wk.Hyy[i];
      a[4] = Rxxi;
      a[5] = Rxyi;
      a[6] = Ryxi;
      a[7] = Ryyi;
      a[8] = Pxxi;
      a[9] = Pxyi;
      a[10] = Pyxi;
      a[11] = Pyyi;
      a[12] = wk.vx_x[i];
      a[13] = wk.vx_y[i];
      a[14] = wk.vy_y[i];
      a[15] = 0.;
    }
  }
  return ORC_OK;
}

// A potential-based TL body, tensile_potential (src/benchmarks/tensile.cpp:291-361):
// derive_displacement (src/derivatives.h:101-149) with the Randles-Libersky
// kernel, material_compute_stress_green_stvenant (src/material.h:132-161),
// potential_comp_elastic_potential (src/potential.h:33-85) with the reverse
// pairs' kernels, correctors_mghn_artificial_viscosity, contmech_advection.
// No continuity or stress rate: rho_t = S_t = 0.
// aux: Hxx, Hxy, Hyx, Hyy, Fe[0][0], Fe[1][0], Fe[0][1], Fe[1][1], 0, 0, 0, 0, Sxx, Sxy, Syy, 0
// (Fe = -2 V J S, J = F).
int tl_stage_potential(const orc_consts* c, const orc_tl* t, uint32_t n, orc_soa* s, TlWork& wk, int nt,
                       double* d8, double* aux) {
  const uint64_t* R = t->rowp.data();
  const uint32_t* NB = t->nbh.data();
  const double* W = t->w.data();
  const double* WJI = t->wji.data();
  const double* QW = t->qw.data();
  const double E = c->E, nu = c->nu;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double uxi = s->x[i] - s->X[i], uyi = s->y[i] - s->Y[i];
    double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* w = W + 3 * q;
      double uxj = s->x[j] - s->X[j], uyj = s->y[j] - s->Y[j];
      double qw = QW[j];
      Hxx += (uxj - uxi) * w[1] * qw;
      Hxy += (uxj - uxi) * w[2] * qw;
      Hyx += (uyj - uyi) * w[1] * qw;
      Hyy += (uyj - uyi) * w[2] * qw;
    }
    wk.Hxx[i] = Hxx;
    wk.Hxy[i] = Hxy;
    wk.Hyx[i] = Hyx;
    wk.Hyy[i] = Hyy;
    // material_compute_stress_green_stvenant: C = F^T F, E = (C - I)/2
    double D = E / (1 - nu * nu);
    M2 F = mk(Hxx + 1., Hyx, Hxy, Hyy + 1.);
    M2 FT = mk(F.m[0][0], F.m[1][0], F.m[0][1], F.m[1][1]);
    M2 Cm = mul(FT, F);
    M2 Esv = smul(0.5, sub(Cm, mk(1., 0., 0., 1.)));
    double Sxx = D * (Esv.m[0][0] + nu * Esv.m[1][1]);
    double Syy = D * (Esv.m[0][0] * nu + Esv.m[1][1]);
    double Sxy = D * (1 - nu) * Esv.m[1][0];
    wk.vx_x[i] = Sxx;
    wk.vx_y[i] = Sxy;
    wk.vy_y[i] = Syy;
    // Fe = -2*voli*Ji*Si, evaluated ((-2*voli)*Ji)*Si
    M2 Ji = mk(Hxx + 1., Hyx, Hxy, Hyy + 1.);
    M2 Si = mk(Sxx, Sxy, Sxy, Syy);
    double voli = QW[i];
    M2 Fe = mul(smul(-2 * voli, Ji), Si);
    wk.Pxx[i] = Fe.m[0][0];
    wk.Pxy[i] = Fe.m[1][0];
    wk.Pyx[i] = Fe.m[0][1];
    wk.Pyy[i] = Fe.m[1][1];
  }
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    M2 Fei = mk(wk.Pxx[i], wk.Pyx[i], wk.Pxy[i], wk.Pyy[i]);
    double mi = s->m[i];
    double vx_t = 0., vy_t = 0.;
    for (uint64_t q = R[i]; q < R[i + 1]; q++) {
      uint32_t j = NB[q];
      const double* wij = W + 3 * q;
      const double* wji = WJI + 3 * q;
      // fij = Fei * grad_wij: (M v)[r] = M[0][r] v.x + M[1][r] v.y
      double fij0 = Fei.m[0][0] * wij[1] + Fei.m[1][0] * wij[2];
      double fij1 = Fei.m[0][1] * wij[1] + Fei.m[1][1] * wij[2];
      M2 Fej = mk(wk.Pxx[j], wk.Pyx[j], wk.Pxy[j], wk.Pyy[j]);
      double fji0 = Fej.m[0][0] * wji[1] + Fej.m[1][0] * wji[2];
      double fji1 = Fej.m[0][1] * wji[1] + Fej.m[1][1] * wji[2];
      vx_t += 0.5 * (fji0 - fij0) / mi;
      vy_t += 0.5 * (fji1 - fij1) / mi;
    }
    double* d = d8 + 8 * (size_t)i;
    d[2] = vx_t;
    d[3] = vy_t;
    double fcx = s->fcx ? s->fcx[i] : 0., fcy = s->fcy ? s->fcy[i] : 0.;
    d[2] += fcx / mi;
    d[3] += fcy / mi;
    tl_viscosity(c, t, s, i, d);
    d[0] += s->vx[i];
    d[1] += s->vy[i];
    if (aux) {
      double* a = aux + 16 * (size_t)i;
      a[0] = wk.Hxx[i];
      a[1] = wk.Hxy[i];
      a[2] = wk.Hyx[i];
      a[3] = wk.Hyy[i];
      a[4] = wk.Pxx[i];
      a[5] = wk.Pxy[i];
      a[6] = wk.Pyx[i];
      a[7] = wk.Pyy[i];
      a[8] = a[9] = a[10] = a[11] = 0.;
      a[12] = wk.vx_x[i];
      a[13] = wk.vx_y[i];
      a[14] = wk.vy_y[i];
      a[15] = 0.;
    }
  }
  return ORC_OK;
}

}  // namespace

extern "C" {

int orc_tl_derivatives(const orc_consts* c, orc_tl* t, uint32_t n, orc_soa* s, double time, double* d8,
                       double* aux) {
  if (!c || !t || !s || !d8 || n != t->n) return ORC_ERR_ARG;
  TlWork wk;
  wk.resize(n);
  return tl_stage(c, t, n, s, time, wk, 1, d8, aux);
}

int orc_tl_step(const orc_consts* c, orc_tl* t, uint32_t n, orc_soa* s, double dt, uint32_t nsteps,
                double* time, int threads) {
  if (!c || !t || !s || !time || n != t->n) return ORC_ERR_ARG;
  int nt = set_threads(threads);
  TlWork wk;
  wk.resize(n);
  StageCopy k;
  k.init(n, s);
  std::vector<double> d1(8 * (size_t)n), d2(8 * (size_t)n), d3(8 * (size_t)n), d4(8 * (size_t)n);
  for (uint32_t step = 0; step < nsteps; step++) {
    double t0 = *time;
    k.copy_from(n, s);
    int rc = tl_stage(c, t, n, &k.soa, t0, wk, nt, d1.data(), nullptr);
    if (rc) return rc;
    rk_add(n, s, d1.data(), 0.5 * dt, &k.soa, nt);
    tl_stage(c, t, n, &k.soa, t0, wk, nt, d2.data(), nullptr);
    rk_add(n, s, d2.data(), 0.5 * dt, &k.soa, nt);
    tl_stage(c, t, n, &k.soa, t0, wk, nt, d3.data(), nullptr);
    rk_add(n, s, d3.data(), dt, &k.soa, nt);
    tl_stage(c, t, n, &k.soa, t0, wk, nt, d4.data(), nullptr);
    rk_combine(n, s, d1.data(), d2.data(), d3.data(), d4.data(), dt, nt);
    tl_apply_bc(t, s, t0, nullptr);
    *time = t0 + dt;
    for (uint32_t i = 0; i < n; i++)
      if (!isfinite(s->x[i]) || !isfinite(s->y[i]) || !isfinite(s->vx[i]) || !isfinite(s->vy[i]))
        return ORC_ERR_NONFINITE;
  }
  return ORC_OK;
}

}  // extern "C"

// ==========================================================================
// Penalty contact (src/contact.h, src/boundary.h, src/convex_hull.h)

extern "C" {

uint32_t orc_convex_hull(const double* x, const double* y, uint32_t n, uint32_t* idx_out) {
  // src/convex_hull.h:30-95: points as (x, y, index), sorted by leftlower
  // (x, then y); ccw(p1, p2, p3) = (p2 - p1) x (p3 - p1)
  struct P {
    double x, y;
    uint32_t i;
  };
  std::vector<P> pts(n);
  for (uint32_t i = 0; i < n; i++) pts[i] = P{x[i], y[i], i};
  std::sort(pts.begin(), pts.end(), [](const P& a, const P& b) {
    if (a.x < b.x) return true;
    if (a.x > b.x) return false;
    if (a.y < b.y) return true;
    if (a.y > b.y) return false;
    return false;
  });
  auto ccw = [](const P& p1, const P& p2, const P& p3) {
    return (p2.x - p1.x) * (p3.y - p1.y) - (p2.y - p1.y) * (p3.x - p1.x);
  };
  std::vector<P> hull(2 * (size_t)n + 1);
  int k = 0;
  for (int i = 0; i < (int)n; ++i) {
    while (k >= 2 && ccw(hull[k - 2], hull[k - 1], pts[i]) < 0) --k;
    hull[k++] = pts[i];
  }
  for (int i = (int)n - 2, t = k + 1; i >= 0; --i) {
    while (k >= t && ccw(hull[k - 2], hull[k - 1], pts[i]) < 0) --k;
    hull[k++] = pts[i];
  }
  uint32_t m = 0;
  for (int i = 0; i < k - 1; i++) idx_out[m++] = hull[i].i;
  return m;
}

void orc_boundary_update(uint32_t nb, const uint32_t* idx, const double* x, const double* y, double* pts) {
  // src/body.h:152-165: positions first, then src/boundary.h:47-73
  for (uint32_t k = 0; k < nb; k++) {
    pts[4 * (size_t)k] = x[idx[k]];
    pts[4 * (size_t)k + 1] = y[idx[k]];
  }
  auto normalize = [](double vx, double vy, double* o) {
    // glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt(d) = 1/sqrt(d)
    double is = 1. / sqrt(vx * vx + vy * vy);
    o[0] = vx * is;
    o[1] = vy * is;
  };
  for (uint32_t i = 0; i < nb; i++) {
    int left = ((int)i) - 1;
    int right = i + 1;
    if (left < 0) left = nb - 1;
    if (right >= (int)nb) right = 0;
    double px = pts[4 * (size_t)i], py = pts[4 * (size_t)i + 1];
    double tlx = pts[4 * (size_t)left] - px, tly = pts[4 * (size_t)left + 1] - py;
    double trx = pts[4 * (size_t)right] - px, try_ = pts[4 * (size_t)right + 1] - py;
    double nl[2], nr[2];
    normalize(-tly, tlx, nl);
    normalize(try_, -trx, nr);
    pts[4 * (size_t)i + 2] = 0.5 * (nl[0] + nr[0]);
    pts[4 * (size_t)i + 3] = 0.5 * (nl[1] + nr[1]);
  }
}

void orc_contact(uint32_t nb, const double* pts, uint32_t ns, const double* x, const double* y, const double* m,
                 double dt, double* fcx, double* fcy) {
  // src/contact.h:36-80
  const double dt2 = dt * dt;
  const double alpha = 0.1;
  for (uint32_t i = 0; i < ns; i++) {
    fcx[i] = 0.;
    fcy[i] = 0.;
  }
  for (uint32_t i = 0; i < ns; i++) {
    double qx = x[i], qy = y[i];
    // kd_tree::find_closest_point (src/kd_tree.cpp:55-106) returns a point of
    // least (qx - px)^2 + (qy - py)^2
    double best = DBL_MAX;
    uint32_t bk = 0;
    for (uint32_t k = 0; k < nb; k++) {
      double dx = qx - pts[4 * (size_t)k];
      double dy = qy - pts[4 * (size_t)k + 1];
      double dist2 = dx * dx + dy * dy;
      if (dist2 < best) {
        best = dist2;
        bk = k;
      }
    }
    const double* c = pts + 4 * (size_t)bk;
    double dx = qx - c[0];
    double dy = qy - c[1];
    double pdotn = dx * c[2] + dy * c[3];
    bool inside = pdotn <= 0.;
    if (!inside) continue;
    double gN_nx = c[2] * pdotn;
    double gN_ny = c[3] * pdotn;
    double ms = m[i];
    double fx = ms * gN_nx / dt2 * alpha;
    double fy = ms * gN_ny / dt2 * alpha;
    fcx[i] = -fx;
    fcy[i] = -fy;
  }
}

}  // extern "C"

