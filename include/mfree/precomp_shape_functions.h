// src/precomp_shape_functions.h of the reference: the kernel precomputations
// drivers call on a particle array before the body exists.  They are
// recorded here and carried out on the GPU by the body's first step:
// precomp_randles_libersky (src/precomp_shape_functions.h:231-290) for the
// Total-Lagrangian pipeline, precomp_sph (:210-229) for co-rotated bodies and
// precomp_randles_libersky_symmetric (:292-316) for potential bodies; the
// RKPM variants are accepted and left to the weak-form bodies, which step()
// refuses.
#ifndef MFREE_PRECOMP_SHAPE_FUNCTIONS_H_
#define MFREE_PRECOMP_SHAPE_FUNCTIONS_H_

#include <map>

namespace mfree_adapter {

// Set-up requests of TL drivers (find_neighbors + precomp_randles_libersky
// before the body exists, src/benchmarks/tensile.cpp:271-272), keyed by the
// particle array.
struct TlRequest {
  double h = 0.;
  bool rl = false;      // precomp_randles_libersky
  bool sph = false;     // precomp_sph
  bool rl_sym = false;  // precomp_randles_libersky_symmetric
};
inline std::map<const void*, TlRequest>& tl_requests() {
  static std::map<const void*, TlRequest> r;
  return r;
}

}  // namespace mfree_adapter

// src/precomp_shape_functions.h:231-290: recorded, executed on the GPU.
template <typename T>
void precomp_randles_libersky(T** particles, unsigned int np) {
  (void)np;
  mfree_adapter::tl_requests()[(const void*)particles].rl = true;
}

// src/precomp_shape_functions.h:210-229 (co-rotated bodies; UL bodies
// recompute it every stage): recorded, executed on the GPU.
template <typename T>
void precomp_sph(T** particles, unsigned int n) {
  (void)n;
  mfree_adapter::tl_requests()[(const void*)particles].sph = true;
}
// src/precomp_shape_functions.h:292-316 (potential bodies): recorded,
// executed on the GPU.
template <typename T>
void precomp_randles_libersky_symmetric(T** particles, unsigned int np) {
  (void)np;
  mfree_adapter::tl_requests()[(const void*)particles].rl_sym = true;
}
// :175-208 (RKPM, weak forms).
template <typename T>
void precomp_rkpm(T** particles, unsigned int n) {
  (void)particles;
  (void)n;
}
template <typename T>
void precomp_rkpm(T** from, T** to, unsigned int nto) {
  (void)from;
  (void)to;
  (void)nto;
}

#endif
