// Reference-compatible particle classes (src/particle.h:33-149): same class
// names, same public fields and the same reset / copy_into, so drivers
// written against the reference compile unchanged.  The state lives on the GPU while a body steps; the
// fields are synchronised by body<T>::get_particles().
#ifndef MFREE_PARTICLE_H_
#define MFREE_PARTICLE_H_

#include <assert.h>
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

  // src/particle.cpp:45-74: the state and the first num_nbh neighbour
  // entries into *q (the reference's RK4 stage copies use it)
  virtual void copy_into(particle* q) const {
    q->idx = idx;
    q->hash = hash;
    q->x = x;
    q->y = y;
    q->X = X;
    q->Y = Y;
    q->m = m;
    q->rho = rho;
    q->h = h;
    q->quad_weight = quad_weight;
    q->p = p;
    q->vx = vx;
    q->vy = vy;
    q->Sxx = Sxx;
    q->Sxy = Sxy;
    q->Syy = Syy;
    q->vx_x = vx_x;
    q->vx_y = vx_y;
    q->vy_x = vy_x;
    q->vy_y = vy_y;
    q->x_t = x_t;
    q->y_t = y_t;
    q->rho_t = rho_t;
    q->vx_t = vx_t;
    q->vy_t = vy_t;
    q->Sxx_t = Sxx_t;
    q->Sxy_t = Sxy_t;
    q->Syy_t = Syy_t;
    q->fcx = fcx;
    q->fcy = fcy;
    q->num_nbh = num_nbh;
    memcpy(q->nbh, nbh, num_nbh * sizeof(nbh[0]));
    memcpy(q->w, w, num_nbh * sizeof(w[0]));
  }

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

  void copy_into(particle* q) const override {
    particle::copy_into(q);
    particle_tl* t = static_cast<particle_tl*>(q);
    t->Hxx = Hxx;
    t->Hxy = Hxy;
    t->Hyx = Hyx;
    t->Hyy = Hyy;
    t->Fdotxx = Fdotxx;
    t->Fdotxy = Fdotxy;
    t->Fdotyx = Fdotyx;
    t->Fdotyy = Fdotyy;
    t->Pxx = Pxx;
    t->Pxy = Pxy;
    t->Pyx = Pyx;
    t->Pyy = Pyy;
    t->Pxx_x = Pxx_x;
    t->Pxy_x = Pxy_x;
    t->Pyx_y = Pyx_y;
    t->Pyy_y = Pyy_y;
  }
};

// Co-rotated TL particle (src/particle.h:81-91): its rotation tensor.
class particle_corotational : public particle_tl {
 public:
  particle_corotational() {}
  explicit particle_corotational(unsigned int i) : particle_tl(i) {}
  double Rxx = 0., Rxy = 0., Ryx = 0., Ryy = 0.;

  void copy_into(particle* q) const override {
    particle_tl::copy_into(q);
    particle_corotational* t = static_cast<particle_corotational*>(q);
    t->Rxx = Rxx;
    t->Rxy = Rxy;
    t->Ryx = Ryx;
    t->Ryy = Ryy;
  }
};

// Potential-based TL particle (src/particle.h:93-103): the kernel values
// seen from the neighbours.
class particle_potential : public particle_tl {
 public:
  particle_potential() {}
  explicit particle_potential(unsigned int i) : particle_tl(i) {}
  kernel_result wji[MAX_NBH] = {};

  void copy_into(particle* q) const override {
    particle_tl::copy_into(q);
    memcpy(static_cast<particle_potential*>(q)->wji, wji, num_nbh * sizeof(wji[0]));
  }
};

// Weak-form TL particle (src/particle.h:105-114).
class particle_tl_weak : public particle_tl {
 public:
  particle_tl_weak() {}
  explicit particle_tl_weak(unsigned int i) : particle_tl(i) {}
};

class particle_ul : public particle {
 public:
  particle_ul() {}
  explicit particle_ul(unsigned int i) : particle(i) {}
  double Sxx_x = 0., Sxy_y = 0., Sxy_x = 0., Syy_y = 0.;

  void copy_into(particle* q) const override {
    particle::copy_into(q);
    particle_ul* t = static_cast<particle_ul*>(q);
    t->Sxx_x = Sxx_x;
    t->Sxy_y = Sxy_y;
    t->Sxy_x = Sxy_x;
    t->Syy_y = Syy_y;
  }
};

// Weak-form UL particle (src/particle.h:128-139): candidate neighbours of
// its stress point.
class particle_ul_weak : public particle_ul {
 public:
  particle_ul_weak() {}
  explicit particle_ul_weak(unsigned int i) : particle_ul(i) {}
  unsigned int num_cand_nbh = 0;
  unsigned int cand_nbh[MAX_NBH] = {};

  void copy_into(particle* q) const override {
    particle_ul::copy_into(q);
    particle_ul_weak* t = static_cast<particle_ul_weak*>(q);
    t->num_cand_nbh = num_cand_nbh;
    memcpy(t->cand_nbh, cand_nbh, num_cand_nbh * sizeof(cand_nbh[0]));
  }
};

class particle_monaghan : public particle_ul {
 public:
  particle_monaghan() {}
  explicit particle_monaghan(unsigned int i) : particle_ul(i) {}
  double Rxx = 0., Rxy = 0., Ryy = 0.;

  void copy_into(particle* q) const override {
    particle_ul::copy_into(q);
    particle_monaghan* t = static_cast<particle_monaghan*>(q);
    t->Rxx = Rxx;
    t->Rxy = Rxy;
    t->Ryy = Ryy;
  }
};

#endif
