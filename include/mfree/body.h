// Reference-compatible body<T> (src/body.h:35-235) on top of the C ABI.
//
// Drivers build bodies exactly as with the reference: new'ed particle
// arrays, simulation_data, dt, boundary conditions, add_action(&...).  At
// the first step() the registered action list is matched, by function
// identity, against the pipelines of the library:
//   Gray-Monaghan UL  (src/benchmarks/rings.cpp:86-96)   -> particle_monaghan
//   Total-Lagrangian  (src/benchmarks/tensile.cpp:276-286) -> particle_tl
//   Godunov UL        (src/benchmarks/rings.cpp:151-158)  -> particle_ul (FP64)
//   co-rotated TL     (src/benchmarks/tensile.cpp:202-208) -> particle_corotational (FP64)
//   potential TL      (src/benchmarks/tensile.cpp:354-358) -> particle_potential (FP64)
// and the body is moved to the GPU (SoA, idx order).  Any other action
// list, particle type or boundary function is rejected with a message
// (MFREE_ERR_PIPELINE): there is no CPU execution of the actions here.
//
// After step(), get_particles() reads the integrated state back into the
// particle objects (x, y, vx, vy, rho, Sxx, Sxy, Syy).  Derived per-stage
// fields (p, quad_weight, gradients, neighbour lists) are not filled.
#ifndef MFREE_BODY_H_
#define MFREE_BODY_H_

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../mfree_iwf_amd.h"
#include "actions.h"
#include "boundary.h"
#include "boundary_conditions.h"
#include "gauss_integration.h"
#include "particle.h"
#include "precomp_shape_functions.h"
#include "simulation_data.h"
#include "simulation_time.h"

namespace mfree_adapter {

// Arithmetic of the pair sums for bodies created afterwards.
inline int32_t& precision() {
  static int32_t p = MFREE_PRECISION_FP64;
  return p;
}

struct Gpu {
  mfree_ctx* ctx = nullptr;
  ~Gpu() {
    if (ctx) mfree_destroy(ctx);
  }
};

}  // namespace mfree_adapter

// src/rk4.h:224-273: the stepper sets the global dt; the integration itself
// runs on the GPU.
template <typename T>
class rk4 {
 public:
  rk4() {}
  rk4(unsigned int num_part, double dt) {
    (void)num_part;
    simulation_time::getInstance().m_dt = dt;
  }
  rk4(unsigned int num_part, unsigned int num_quad, double dt) {
    (void)num_part;
    (void)num_quad;
    simulation_time::getInstance().m_dt = dt;
  }
};

template <typename T>
class body {
  static_assert(std::is_base_of<particle, T>::value, "T is not derived from particle");

 public:
  body() {}
  body(T** particles, unsigned int n, simulation_data data, double dt,
       boundary_conditions<T> bc = boundary_conditions<T>(), T** quad_points = 0, unsigned int nq = 0,
       gauss_int* gauss_integration = 0)
      : m_data(data), m_bc(bc), m_num_part(n), m_particles(particles), m_stepper(n, nq, dt),
        m_weak(quad_points != 0 || nq != 0 || gauss_integration != 0) {}

  void add_action(void (*action)(body<T>& b)) { m_actions.push_back(action); }

  // src/body.h:196-212: the outline used by the penalty contact.  Its
  // points and normals are kept on the GPU (update_boundary); the host
  // object is refreshed from the particles when it is read.
  void set_boundary(boundary<T>* b) {
    m_boundary = b;
    *m_bnd_sent = false;
  }
  boundary<T>* get_boundary() const {
    if (m_boundary && *m_bnd_stale) {
      m_boundary->update(get_particles());
      *m_bnd_stale = false;
    }
    return m_boundary;
  }

  // src/body.h:223-225 -> src/rk4.h:71-108
  void step() {
    if (!m_gpu) setup();
    mfree_status st = mfree_step(m_gpu->ctx, 1);
    if (st != MFREE_OK) mfree_adapter::die(std::string("step: ") + mfree_last_error(m_gpu->ctx));
    simulation_time::getInstance().increment_time();
    *m_stale = true;
  }

  T** get_particles() const {
    sync();
    return m_particles;
  }
  T** get_cur_particles() const { return get_particles(); }
  unsigned int get_num_part() const { return m_num_part; }
  simulation_data get_sim_data() const { return m_data; }
  bool has_boundary() const { return m_boundary != nullptr; }
  // src/body.h:152-165: the outline follows the particles (on the GPU)
  void update_boundary() {
    if (!m_boundary) return;
    if (!m_gpu) setup();
    if (!*m_bnd_sent) {
      const std::vector<unsigned int>& ix = m_boundary->indexes();
      std::vector<uint32_t> idx(ix.begin(), ix.end());
      if (mfree_set_boundary(m_gpu->ctx, 0, idx.data(), (uint32_t)idx.size()) != MFREE_OK)
        mfree_adapter::die(std::string("set_boundary: ") + mfree_last_error(m_gpu->ctx));
      *m_bnd_sent = true;
    }
    if (mfree_update_boundaries(m_gpu->ctx) != MFREE_OK)
      mfree_adapter::die(std::string("update_boundary: ") + mfree_last_error(m_gpu->ctx));
    *m_bnd_stale = true;
  }
  mfree_ctx* context() const { return m_gpu ? m_gpu->ctx : nullptr; }
  // the GPU changed the body (contact forces): host particles are re-read on access
  void mark_stale() { *m_stale = true; }

  // The library method the registered actions, particle type and boundary
  // conditions map to (MFREE_METHOD_*), or -1 when step() would refuse the
  // body (refusal() says why); needs no GPU.
  int32_t method() const { return refusal().empty() ? pipeline() : -1; }
  std::string refusal() const {
    if (m_weak) return "weak-form and FEM bodies (quadrature points) are not part of this library";
    const int32_t method = pipeline();
    if (method < 0) return "the action list is not one of the supported pipelines (MFREE_ERR_PIPELINE)";
    if (!m_bc.list().empty()) {
      if (!tl(method)) return "boundary conditions on a UL body are not supported";
      double speed = -1.;
      for (const auto& e : m_bc.list()) {
        double a = 0.;
        if (!probe_strip(std::get<1>(e), &a)) return "boundary function is not a velocity strip";
        if (speed >= 0. && fabs(a) != speed) return "strips must move with equal and opposite speed";
        speed = fabs(a);
      }
    }
    return std::string();
  }

 private:
  simulation_data m_data;
  std::vector<void (*)(body<T>&)> m_actions;
  boundary_conditions<T> m_bc;
  unsigned int m_num_part = 0;
  T** m_particles = nullptr;
  rk4<T> m_stepper;
  bool m_weak = false;
  boundary<T>* m_boundary = nullptr;
  std::shared_ptr<mfree_adapter::Gpu> m_gpu;
  std::shared_ptr<bool> m_stale = std::make_shared<bool>(false);
  std::shared_ptr<bool> m_bnd_sent = std::make_shared<bool>(false);   // outline indexes given to the library
  std::shared_ptr<bool> m_bnd_stale = std::make_shared<bool>(false);  // host outline behind the GPU's

  static bool tl(int32_t method) {
    return method == MFREE_METHOD_TL_RL || method == MFREE_METHOD_TL_COROTATIONAL ||
           method == MFREE_METHOD_TL_POTENTIAL;
  }

  int32_t pipeline() const {
    typedef void (*A)(body<T>&);
    const std::vector<A> ul = {&riemann_update_integration_weights<T>,
                               &material_eos<T>,
                               &correctors_mghn_artificial_stress<T>,
                               &derive_stress_monaghan<T>,
                               &derive_velocity<T>,
                               &correctors_mghn_artificial_viscosity<T>,
                               &correctors_xsph<T>,
                               &material_stress_rate_jaumann<T>,
                               &contmech_continuity<T>,
                               &contmech_momentum_ul<T>,
                               &contmech_advection<T>};
    const std::vector<A> tl = {&derive_displacement<T>,
                               &derive_velocity_tl<T>,
                               &contmech_velocity_gradient_tl<T>,
                               &contmech_continuity<T>,
                               &material_eos<T>,
                               &material_stress_rate_jaumann<T>,
                               &contmech_cauchy_to_nominal<T>,
                               &derive_stress_tl<T>,
                               &correctors_mghn_artificial_viscosity<T>,
                               &contmech_momentum_tl<T>,
                               &contmech_advection<T>};
    // src/benchmarks/rings.cpp:151-158
    const std::vector<A> gd = {&riemann_update_integration_weights<T>,
                               &material_eos<T>,
                               &derive_velocity<T>,
                               &material_stress_rate_jaumann<T>,
                               &godunov_continuity<T>,
                               &godunov_momentum<T>,
                               &contmech_advection<T>};
    // src/benchmarks/tensile.cpp:202-208
    const std::vector<A> corot = {&corotational_anisotropy_tensor<T>,
                                  &derive_displacement_corotational<T>,
                                  &material_compute_stress_cauchy_green<T>,
                                  &contmech_cauchy_to_nominal<T>,
                                  &derive_stress_corotational<T>,
                                  &correctors_mghn_artificial_viscosity<T>,
                                  &contmech_advection<T>};
    // src/benchmarks/tensile.cpp:354-358
    const std::vector<A> pot = {&derive_displacement<T>,
                                &material_compute_stress_green_stvenant<T>,
                                &potential_comp_elastic_potential<T>,
                                &correctors_mghn_artificial_viscosity<T>,
                                &contmech_advection<T>};
    if (m_actions == ul && std::is_base_of<particle_monaghan, T>::value) return MFREE_METHOD_UL_MONAGHAN;
    if (m_actions == tl && std::is_base_of<particle_tl, T>::value) return MFREE_METHOD_TL_RL;
    if (m_actions == gd && std::is_base_of<particle_ul, T>::value) return MFREE_METHOD_UL_GODUNOV;
    if (m_actions == corot && std::is_base_of<particle_corotational, T>::value) return MFREE_METHOD_TL_COROTATIONAL;
    if (m_actions == pot && std::is_base_of<particle_potential, T>::value) return MFREE_METHOD_TL_POTENTIAL;
    return -1;
  }

  // Strip velocity a of a boundary function of the form x = X + a t, y = Y,
  // vx = a, vy = 0, vx_t = vy_t = 0, found by evaluating it at t = 0 and 1.
  static bool probe_strip(void (*fn)(T*), double* a) {
    simulation_time& st = simulation_time::getInstance();
    const double t_saved = st.m_time;
    T p0, p1;
    for (T* p : {&p0, &p1}) {
      p->X = 0.25;
      p->Y = 0.5;
      p->x = p->y = p->vx = p->vy = p->vx_t = p->vy_t = 7.0;
    }
    st.m_time = 0.;
    fn(&p0);
    st.m_time = 1.;
    fn(&p1);
    st.m_time = t_saved;
    const double v = p0.vx;
    *a = v;
    return p0.x == 0.25 && p1.x == 0.25 + v && p0.y == 0.5 && p1.y == 0.5 && p1.vx == v && p0.vy == 0. &&
           p1.vy == 0. && p0.vx_t == 0. && p0.vy_t == 0.;
  }

  void setup() {
    const std::string why = refusal();
    if (!why.empty()) mfree_adapter::die("body::step: " + why);
    const int32_t method = pipeline();
    for (unsigned int i = 0; i < m_num_part; i++)
      if (m_particles[i]->idx != i) mfree_adapter::die("body: particle i must have idx == i");
    mfree_options opt;
    mfree_default_options(&opt);
    opt.method = method;
    // the Godunov, co-rotated and potential lists run in FP64 only
    opt.precision = (method == MFREE_METHOD_UL_MONAGHAN || method == MFREE_METHOD_TL_RL) ? mfree_adapter::precision()
                                                                                       : MFREE_PRECISION_FP64;
    if (tl(method)) {
      // the set-up the driver called on the particles before the body
      // existed (find_neighbors + the list's kernel precomputation)
      auto it = mfree_adapter::tl_requests().find((const void*)m_particles);
      const bool found = it != mfree_adapter::tl_requests().end();
      const bool ok = found && (method == MFREE_METHOD_TL_RL            ? it->second.rl
                                : method == MFREE_METHOD_TL_COROTATIONAL ? it->second.sph
                                                                         : it->second.rl_sym);
      if (!ok)
        mfree_adapter::die(method == MFREE_METHOD_TL_RL
                               ? "TL body: find_neighbors and precomp_randles_libersky must be called on its particles"
                           : method == MFREE_METHOD_TL_COROTATIONAL
                               ? "co-rotated body: find_neighbors and precomp_sph must be called on its particles"
                               : "potential body: find_neighbors and precomp_randles_libersky_symmetric must be "
                                 "called on its particles");
      opt.tl_h_search = it->second.h;
    }
    auto gpu = std::make_shared<mfree_adapter::Gpu>();
    if (mfree_create(&opt, &gpu->ctx) != MFREE_OK) mfree_adapter::die("mfree_create failed");
    const physical_constants pc = m_data.get_physical_constants();
    const correction_constants cc = m_data.get_correction_constants();
    mfree_params prm;
    prm.E = pc.E();
    prm.nu = pc.nu();
    prm.rho0 = pc.rho0();
    prm.wdeltap = cc.get_monaghan_const().mghn_wdeltap();
    prm.corr_exp = cc.get_monaghan_const().mghn_corr_exp();
    prm.mghn_eps = cc.get_monaghan_const().mghn_eps();
    prm.alpha = cc.get_art_visc_const().artvisc_alpha();
    prm.beta = cc.get_art_visc_const().artvisc_beta();
    prm.eta = cc.get_art_visc_const().artvisc_eta();
    prm.xsph_eps = cc.xsph_eps();
    prm.dt = simulation_time::getInstance().get_dt();
    prm.time = simulation_time::getInstance().get_time();
    const unsigned int n = m_num_part;
    std::vector<double> f[14];
    for (auto& v : f) v.resize(n);
    for (unsigned int i = 0; i < n; i++) {
      const T* p = m_particles[i];
      f[0][i] = p->x;
      f[1][i] = p->y;
      f[2][i] = p->X;
      f[3][i] = p->Y;
      f[4][i] = p->m;
      f[5][i] = p->rho;
      f[6][i] = p->h;
      f[7][i] = p->vx;
      f[8][i] = p->vy;
      f[9][i] = p->Sxx;
      f[10][i] = p->Sxy;
      f[11][i] = p->Syy;
      f[12][i] = p->fcx;
      f[13][i] = p->fcy;
    }
    mfree_particles hp = {f[0].data(), f[1].data(), f[2].data(),  f[3].data(),  f[4].data(),
                          f[5].data(), f[6].data(), f[7].data(),  f[8].data(),  f[9].data(),
                          f[10].data(), f[11].data(), f[12].data(), f[13].data()};
    uint32_t bid = 0;
    if (mfree_add_body(gpu->ctx, &prm, &hp, n, &bid) != MFREE_OK)
      mfree_adapter::die(std::string("add_body: ") + mfree_last_error(gpu->ctx));
    if (!m_bc.list().empty()) {
      if (!tl(method)) mfree_adapter::die("boundary conditions on a UL body are not supported");
      std::vector<uint32_t> left, right;
      double speed = -1.;
      for (const auto& e : m_bc.list()) {
        double a = 0.;
        if (!probe_strip(std::get<1>(e), &a)) mfree_adapter::die("boundary function is not a velocity strip");
        if (speed >= 0. && fabs(a) != speed) mfree_adapter::die("strips must move with equal and opposite speed");
        speed = fabs(a);
        for (unsigned int k : std::get<0>(e)) (a < 0. ? left : right).push_back(k);
      }
      if (mfree_set_strip_bc(gpu->ctx, bid, left.data(), (uint32_t)left.size(), right.data(),
                             (uint32_t)right.size(), speed) != MFREE_OK)
        mfree_adapter::die(std::string("set_strip_bc: ") + mfree_last_error(gpu->ctx));
    }
    if (tl(method) && mfree_tl_precompute(gpu->ctx) != MFREE_OK)
      mfree_adapter::die(std::string("TL precompute: ") + mfree_last_error(gpu->ctx));
    m_gpu = gpu;
  }

  void sync() const {
    if (!m_gpu || !*m_stale) return;
    const unsigned int n = m_num_part;
    std::vector<double> f[10];
    for (auto& v : f) v.resize(n);
    mfree_particles hp = {f[0].data(), f[1].data(), nullptr,     nullptr,     nullptr,     f[4].data(), nullptr,
                          f[2].data(), f[3].data(), f[5].data(), f[6].data(), f[7].data(), f[8].data(), f[9].data()};
    if (mfree_get_state(m_gpu->ctx, 0, &hp) != MFREE_OK)
      mfree_adapter::die(std::string("get_state: ") + mfree_last_error(m_gpu->ctx));
    for (unsigned int i = 0; i < n; i++) {
      T* p = m_particles[i];
      p->x = f[0][i];
      p->y = f[1][i];
      p->vx = f[2][i];
      p->vy = f[3][i];
      p->rho = f[4][i];
      p->Sxx = f[5][i];
      p->Sxy = f[6][i];
      p->Syy = f[7][i];
      p->fcx = f[8][i];
      p->fcy = f[9][i];
    }
    *m_stale = false;
  }
};

#endif
