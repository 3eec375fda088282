// tensile_tl(nbox): the reference's Total-Lagrangian tension driver
// (src/benchmarks/tensile.cpp:213-289) with its constants: unit square,
// h = dx, one-column strips pulled at +-10 m/s.
#ifndef MFREE_BENCHMARKS_TENSILE_H_
#define MFREE_BENCHMARKS_TENSILE_H_

#include <vector>

#include "../body.h"
#include "neighbor_search.h"

template <typename T>
static void tl_strip_left(T* p) {
  p->x = p->X - 10 * simulation_time::getInstance().get_time();
  p->y = p->Y;
  p->vx = -10.;
  p->vy = 0.;
  p->vx_t = 0.;
  p->vy_t = 0.;
}

template <typename T>
static void tl_strip_right(T* p) {
  p->x = p->X + 10 * simulation_time::getInstance().get_time();
  p->y = p->Y;
  p->vx = 10.;
  p->vy = 0.;
  p->vx_t = 0.;
  p->vy_t = 0.;
}

inline std::vector<body<particle_tl>> tensile_tl(unsigned int nbox) {
  const double E = 1e7, nu = 0.4, rho0 = 1.;
  const double L = 1.;
  const double dx = L / (nbox - 1);
  const double hdx = 1.;
  const double dt = 1e-6;
  const unsigned int n = nbox * nbox;
  particle_tl** ps = new particle_tl*[n];
  std::vector<unsigned int> left, right;
  unsigned int k = 0;
  for (unsigned int i = 0; i < nbox; i++)
    for (unsigned int j = 0; j < nbox; j++, k++) {
      ps[k] = new particle_tl(k);
      ps[k]->x = ps[k]->X = i * dx;
      ps[k]->y = ps[k]->Y = j * dx;
      if (i == 0) left.push_back(k);
      if (i == nbox - 1) right.push_back(k);
    }
  for (unsigned int i = 0; i < n; i++) {
    ps[i]->rho = rho0;
    ps[i]->h = hdx * dx;
    ps[i]->m = dx * dx * rho0;
    ps[i]->quad_weight = ps[i]->m / ps[i]->rho;
  }
  boundary_conditions<particle_tl> bc;
  bc.add_boundary_condition(left, &tl_strip_left<particle_tl>);
  bc.add_boundary_condition(right, &tl_strip_right<particle_tl>);
  simulation_data sd(physical_constants(nu, E, rho0),
                     correction_constants(constants_monaghan(), constants_artificial_viscosity(1., 1., 0.1), 0., false));
  find_neighbors(ps, n, hdx * dx, distance_euclidian);
  precomp_randles_libersky<particle_tl>(ps, n);
  body<particle_tl> b(ps, n, sd, dt, bc);
  b.add_action(&derive_displacement);
  b.add_action(&derive_velocity_tl);
  b.add_action(&contmech_velocity_gradient_tl);
  b.add_action(&contmech_continuity);
  b.add_action(&material_eos);
  b.add_action(&material_stress_rate_jaumann);
  b.add_action(&contmech_cauchy_to_nominal);
  b.add_action(&derive_stress_tl);
  b.add_action(&correctors_mghn_artificial_viscosity);
  b.add_action(&contmech_momentum_tl);
  b.add_action(&contmech_advection);
  return std::vector<body<particle_tl>>({b});
}

#endif
