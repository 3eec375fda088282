// src/benchmarks/utils.h of the reference: mesh_ring builds the nodes and
// quadrilateral elements of the FEM and weak-form ring drivers.  Those
// methods are not part of this library; the function is declared so that
// the drivers compile and refuses at run time.
#ifndef MFREE_BENCH_UTILS_H_
#define MFREE_BENCH_UTILS_H_

#include <vector>

#include "../body.h"
#include "../element.h"

template <typename T>
T** mesh_ring(unsigned int nbox, unsigned int& n, std::vector<element>& elements_in, double rin, double rout) {
  (void)nbox;
  (void)elements_in;
  (void)rin;
  (void)rout;
  n = 0;
  mfree_adapter::die("mesh_ring: FEM and weak-form drivers are not part of this library");
}

#endif
