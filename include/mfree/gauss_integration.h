// src/gauss_integration.h of the reference: Gauss quadrature of the FEM and
// weak Total-Lagrangian methods.  Not part of this library: the class is
// declared so that the drivers compile, and placing quadrature points
// refuses at run time.
#ifndef MFREE_GAUSS_INTEGRATION_H_
#define MFREE_GAUSS_INTEGRATION_H_

#include <vector>

#include "actions.h"
#include "element.h"

class gauss_int {
  std::vector<element> m_elements;
  unsigned int m_order = 0;

 public:
  gauss_int(std::vector<element> elements, unsigned int order) : m_elements(elements), m_order(order) {}

  template <typename T>
  void update_gauss_points(T** particles, T** gauss_points) {
    (void)particles;
    (void)gauss_points;
    mfree_adapter::die("gauss_int: FEM and weak-form quadrature are not part of this library");
  }
};

#endif
