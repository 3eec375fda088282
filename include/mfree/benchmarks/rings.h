// rings_monaghan(nbox): the reference's two-ring Gray-Monaghan driver
// (src/benchmarks/rings.cpp:20-99) with the reference's constants, on the
// GPU body.  rings_monaghan_params() is the same scene with every
// constant exposed (the benchmark configurations use it).
#ifndef MFREE_BENCHMARKS_RINGS_H_
#define MFREE_BENCHMARKS_RINGS_H_

#include <math.h>

#include <vector>

#include "../body.h"
#include "../convex_hull.h"

struct rings_setup {
  double E = 1e7, nu = 0.4, rho0 = 1.;  // rubber-like, src/benchmarks/rings.cpp:22-24
  double ri = 0.03, ro = 0.04;
  double centre = 0.04 + 0.001;         // ring centres at +-centre
  double hdx = 1.7;
  double speed = 180.;
  double dt = 1e-6;
  double alpha = 1., beta = 1., eta = 0.1;
  double art_stress_eps = 0.3, stress_exponent = 4., xsph_eps = 0.5;
};

inline std::vector<body<particle_monaghan>> rings_monaghan_params(unsigned int nbox, const rings_setup& s) {
  const double dx = 2 * s.ro / (nbox - 1);
  // lattice points inside the annulus, each emitted for the left ring and
  // then the right ring (even / odd idx)
  std::vector<particle_monaghan*> ps;
  for (unsigned int i = 0; i < nbox; i++)
    for (unsigned int j = 0; j < nbox; j++) {
      const double px = -s.ro + i * dx, py = -s.ro + j * dx;
      const double r = sqrt(px * px + py * py);
      if (!(r < s.ro && r >= s.ri)) continue;
      for (int side = 0; side < 2; side++) {
        auto* p = new particle_monaghan((unsigned int)ps.size());
        p->x = side == 0 ? px - s.centre : px + s.centre;
        p->y = py;
        ps.push_back(p);
      }
    }
  const unsigned int n = (unsigned int)ps.size();
  particle_monaghan** arr = new particle_monaghan*[n];
  for (unsigned int i = 0; i < n; i++) {
    arr[i] = ps[i];
    arr[i]->rho = s.rho0;
    arr[i]->h = s.hdx * dx;
    arr[i]->m = dx * dx * s.rho0;
    arr[i]->quad_weight = arr[i]->m / arr[i]->rho;
    arr[i]->vx = arr[i]->x < 0. ? s.speed : -s.speed;
  }
  const double wdeltap = cubic_spline(0, 0, dx, 0, s.hdx * dx).w;
  correction_constants cc(constants_monaghan(wdeltap, s.stress_exponent, s.art_stress_eps),
                          constants_artificial_viscosity(s.alpha, s.beta, s.eta), s.xsph_eps);
  simulation_data sd(physical_constants(s.nu, s.E, s.rho0), cc);
  body<particle_monaghan> b(arr, n, sd, s.dt);
  b.add_action(&riemann_update_integration_weights);
  b.add_action(&material_eos);
  b.add_action(&correctors_mghn_artificial_stress);
  b.add_action(&derive_stress_monaghan);
  b.add_action(&derive_velocity);
  b.add_action(&correctors_mghn_artificial_viscosity);
  b.add_action(&correctors_xsph);
  b.add_action(&material_stress_rate_jaumann);
  b.add_action(&contmech_continuity);
  b.add_action(&contmech_momentum_ul);
  b.add_action(&contmech_advection);
  return std::vector<body<particle_monaghan>>({b});
}

inline std::vector<body<particle_monaghan>> rings_monaghan(unsigned int nbox) {
  return rings_monaghan_params(nbox, rings_setup());
}

// rings_monaghan_contact(nbox) (src/benchmarks/rings.cpp:522-620): the same
// rings as two bodies, particles of each numbered in lattice order, each
// with its convex hull as outline; the main loop's penalty contact
// (contact.h) couples them.
inline std::vector<body<particle_monaghan>> rings_monaghan_contact(unsigned int nbox) {
  const rings_setup s;
  const double dx = 2 * s.ro / (nbox - 1);
  std::vector<particle_monaghan*> side[2];
  for (unsigned int i = 0; i < nbox; i++)
    for (unsigned int j = 0; j < nbox; j++) {
      const double px = -s.ro + i * dx, py = -s.ro + j * dx;
      const double r = sqrt(px * px + py * py);
      if (!(r < s.ro && r >= s.ri)) continue;
      for (int k = 0; k < 2; k++) {
        auto* p = new particle_monaghan((unsigned int)side[k].size());
        p->x = k == 0 ? px - s.centre : px + s.centre;
        p->y = py;
        p->rho = s.rho0;
        p->h = s.hdx * dx;
        p->m = dx * dx * s.rho0;
        p->quad_weight = p->m / p->rho;
        p->vx = k == 0 ? s.speed : -s.speed;
        side[k].push_back(p);
      }
    }
  const double wdeltap = cubic_spline(0, 0, dx, 0, s.hdx * dx).w;
  correction_constants cc(constants_monaghan(wdeltap, s.stress_exponent, s.art_stress_eps),
                          constants_artificial_viscosity(s.alpha, s.beta, s.eta), s.xsph_eps);
  simulation_data sd(physical_constants(s.nu, s.E, s.rho0), cc);
  std::vector<body<particle_monaghan>> bodies;
  for (int k = 0; k < 2; k++) {
    const unsigned int n = (unsigned int)side[k].size();
    particle_monaghan** arr = new particle_monaghan*[n];
    for (unsigned int i = 0; i < n; i++) arr[i] = side[k][i];
    body<particle_monaghan> b(arr, n, sd, s.dt);
    b.set_boundary(new boundary<particle_monaghan>(convex_hull(arr, n), arr));
    b.add_action(&riemann_update_integration_weights);
    b.add_action(&material_eos);
    b.add_action(&correctors_mghn_artificial_stress);
    b.add_action(&derive_stress_monaghan);
    b.add_action(&derive_velocity);
    b.add_action(&correctors_mghn_artificial_viscosity);
    b.add_action(&correctors_xsph);
    b.add_action(&material_stress_rate_jaumann);
    b.add_action(&contmech_continuity);
    b.add_action(&contmech_momentum_ul);
    b.add_action(&contmech_advection);
    bodies.push_back(b);
  }
  return bodies;
}

#endif
