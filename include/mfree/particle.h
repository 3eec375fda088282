// Reference-compatible particle classes (src/particle.h:33-149): same class
// names, same public fields, so drivers written against the reference
// compile unchanged.  The state lives on the GPU while a body steps; the
// fields are synchronised by body<T>::get_particles().
#ifndef MFREE_PARTICLE_H_
#define MFREE_PARTICLE_H_

#include <string.h>

#include "kernel.h"

#define MAX_NBH 200

class particle {
 public:
  particle() {}
  explicit particle(unsigned int i) : idx(i) {}
  virtual ~particle() {}

  // src/particle.cpp:34-43
  virtual void reset() { x_t = y_t = rho_t = vx_t = vy_t = Sxx_t = Sxy_t = Syy_t = 0.; }

  unsigned int idx = 0, hash = 0;
  double x = 0., y = 0.;
  double X = 0.;
  double Y = 0.;
  double m = 0., rho = 0., h = 0., quad_weight = 0.;
  double p = 0.;
  double vx = 0., vy = 0.;
  double Sxx = 0., Sxy = 0., Syy = 0.;
  double vx_x = 0., vx_y = 0., vy_x = 0., vy_y = 0.;
  double x_t = 0., y_t = 0.;
  double rho_t = 0.;
  double vx_t = 0., vy_t = 0.;
  double Sxx_t = 0., Sxy_t = 0., Syy_t = 0.;
  double fcx = 0.;
  double fcy = 0.;
  unsigned int num_nbh = 0;
  unsigned int nbh[MAX_NBH] = {};
  kernel_result w[MAX_NBH] = {};
};

class particle_tl : public particle {
 public:
  particle_tl() {}
  explicit particle_tl(unsigned int i) : particle(i) {}
  double Hxx = 0., Hxy = 0., Hyx = 0., Hyy = 0.;
  double Fdotxx = 0.;
  double Fdotxy = 0.;
  double Fdotyx = 0.;
  double Fdotyy = 0.;
  double Pxx = 0., Pxy = 0., Pyx = 0., Pyy = 0.;
  double Pxx_x = 0., Pxy_x = 0., Pyx_y = 0., Pyy_y = 0.;
};

class particle_ul : public particle {
 public:
  particle_ul() {}
  explicit particle_ul(unsigned int i) : particle(i) {}
  double Sxx_x = 0., Sxy_y = 0., Sxy_x = 0., Syy_y = 0.;
};

class particle_monaghan : public particle_ul {
 public:
  particle_monaghan() {}
  explicit particle_monaghan(unsigned int i) : particle_ul(i) {}
  double Rxx = 0., Rxy = 0., Ryy = 0.;
};

#endif
