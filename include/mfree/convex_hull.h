// src/convex_hull.h of the reference: the outline of a body, by which the
// contact drivers define its boundary (src/benchmarks/rings.cpp:229-231).
// Andrew's monotone chain (A. M. Andrew, Inf. Process. Lett. 9 (1979) 216):
// particles sorted by (x, y), lower and upper hulls built with a
// counter-clockwise turn test; returns the particle indexes of the hull in
// counter-clockwise order from the lowest-leftmost point, collinear points
// dropped.
#ifndef MFREE_CONVEX_HULL_H_
#define MFREE_CONVEX_HULL_H_

#include <algorithm>
#include <vector>

template <typename T>
std::vector<unsigned int> convex_hull(T** points, unsigned int num_points) {
  std::vector<unsigned int> ord(num_points);
  for (unsigned int i = 0; i < num_points; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [points](unsigned int a, unsigned int b) {
    return points[a]->x < points[b]->x || (points[a]->x == points[b]->x && points[a]->y < points[b]->y);
  });
  auto cross = [points](unsigned int o, unsigned int a, unsigned int b) {
    return (points[a]->x - points[o]->x) * (points[b]->y - points[o]->y) -
           (points[a]->y - points[o]->y) * (points[b]->x - points[o]->x);
  };
  if (num_points < 3) return ord;
  std::vector<unsigned int> h(2 * num_points);
  size_t k = 0;
  for (unsigned int i = 0; i < num_points; i++) {  // lower hull
    while (k >= 2 && cross(h[k - 2], h[k - 1], ord[i]) <= 0) k--;
    h[k++] = ord[i];
  }
  for (size_t i = num_points - 1, t = k + 1; i > 0; i--) {  // upper hull
    while (k >= t && cross(h[k - 2], h[k - 1], ord[i - 1]) <= 0) k--;
    h[k++] = ord[i - 1];
  }
  h.resize(k - 1);  // the first point closes the chain
  return h;
}

#endif
