// The reference's actions (src/correctors.h, src/cont_mech.h, src/material.h,
// src/derivatives.h, src/godunov.h, src/riemann_integration.h,
// src/interpolation.h, src/corotational.h, src/potential.h): every action a
// driver can register with body<T>::add_action, with the reference's names and
// signature template <typename T> void action(body<T>&).
//
// Here an action is an identity, not CPU code: body<T>::step() matches the
// registered list against the pipelines the library runs on the GPU
// (body.h); calling an action directly is an error.
#ifndef MFREE_ACTIONS_H_
#define MFREE_ACTIONS_H_

#include <stdio.h>
#include <stdlib.h>

#include <string>

template <typename T>
class body;

namespace mfree_adapter {

[[noreturn]] inline void die(const std::string& msg) {
  fprintf(stderr, "mfree: %s\n", msg.c_str());
  abort();
}

[[noreturn]] inline void direct_call(const char* name) {
  die(std::string(name) + ": actions run inside body<T>::step() on the GPU and cannot be called directly");
}

}  // namespace mfree_adapter

#define MFREE_ACTION(name)               \
  template <typename T>                  \
  void name(body<T>&) {                  \
    mfree_adapter::direct_call(#name);   \
  }

// src/correctors.h
MFREE_ACTION(correctors_mghn_artificial_stress)
MFREE_ACTION(correctors_mghn_artificial_viscosity)
MFREE_ACTION(correctors_xsph)
MFREE_ACTION(correctors_smooth_vel)
// src/cont_mech.h
MFREE_ACTION(contmech_momentum_ul)
MFREE_ACTION(contmech_momentum_ul_weak)
MFREE_ACTION(contmech_momentum_tl)
MFREE_ACTION(contmech_momentum_tl_weak)
MFREE_ACTION(contmech_continuity)
MFREE_ACTION(contmech_continuity_gp)
MFREE_ACTION(contmech_advection)
MFREE_ACTION(contmech_advection_gp)
MFREE_ACTION(contmech_velocity_gradient_tl)
MFREE_ACTION(contmech_velocity_gradient_tl_gp)
MFREE_ACTION(contmech_cauchy_to_nominal)
MFREE_ACTION(contmech_cauchy_to_nominal_gp)
// src/material.h
MFREE_ACTION(material_eos)
MFREE_ACTION(material_eos_gp)
MFREE_ACTION(material_stress_rate_jaumann)
MFREE_ACTION(material_stress_rate_jaumann_gp)
MFREE_ACTION(material_compute_stress_cauchy_green)
MFREE_ACTION(material_compute_stress_cauchy_green_gp)
MFREE_ACTION(material_compute_stress_green_stvenant)
MFREE_ACTION(material_compute_stress_green_stvenant_gp)
MFREE_ACTION(material_plasticity_gray2000)
// src/derivatives.h
MFREE_ACTION(derive_velocity)
MFREE_ACTION(derive_velocity_weak)
MFREE_ACTION(derive_velocity_tl)
MFREE_ACTION(derive_velocity_tl_weak)
MFREE_ACTION(derive_displacement)
MFREE_ACTION(derive_displacement_weak)
MFREE_ACTION(derive_displacement_corotational)
MFREE_ACTION(derive_stress_monaghan)
MFREE_ACTION(derive_stress_tl)
MFREE_ACTION(derive_stress_weak_tl)
MFREE_ACTION(derive_stress_weak_ul)
MFREE_ACTION(derive_stress_corotational)
// src/godunov.h
MFREE_ACTION(godunov_momentum)
MFREE_ACTION(godunov_continuity)
// src/riemann_integration.h
MFREE_ACTION(riemann_update_integration_weights)
MFREE_ACTION(riemann_update_integration_weights_gp)
// src/interpolation.h
MFREE_ACTION(interpolation_interpolate_velocity)
// src/corotational.h
MFREE_ACTION(corotational_anisotropy_tensor)
// src/potential.h
MFREE_ACTION(potential_comp_elastic_potential)

#undef MFREE_ACTION

#endif
