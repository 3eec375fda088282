// src/convex_hull.h of the reference: the outline of a body, by which the
// contact drivers define its boundary (src/benchmarks/rings.cpp:229-231),
// computed by the library's host code (mfree_convex_hull): Andrew's
// monotone chain over the particles sorted by (x, y), counter-clockwise from
// the lowest-leftmost particle, collinear hull particles kept.
#ifndef MFREE_CONVEX_HULL_H_
#define MFREE_CONVEX_HULL_H_

#include <vector>

#include "../mfree_iwf_amd.h"

template <typename T>
std::vector<unsigned int> convex_hull(T** points, unsigned int num_points) {
  std::vector<double> x(num_points), y(num_points);
  for (unsigned int i = 0; i < num_points; i++) {
    x[i] = points[i]->x;
    y[i] = points[i]->y;
  }
  std::vector<uint32_t> idx(2 * (size_t)num_points + 1);
  const uint32_t k = mfree_convex_hull(x.data(), y.data(), num_points, idx.data());
  return std::vector<unsigned int>(idx.begin(), idx.begin() + k);
}

#endif
