// src/element.h of the reference: four-node quadrilaterals of the FEM and
// weak-form methods (node indexes lower-left, lower-right, top-right,
// top-left) and the element map of a rectangular node lattice.
#ifndef MFREE_ELEMENT_H_
#define MFREE_ELEMENT_H_

#include <vector>

class element {
 public:
  element() {}
  element(unsigned int ll, unsigned int lr, unsigned int tr, unsigned int tl)
      : ll_idx(ll), lr_idx(lr), tr_idx(tr), tl_idx(tl) {}
  unsigned int ll_idx = 0, lr_idx = 0, tr_idx = 0, tl_idx = 0;
};

// Elements of an (ex+1) x (ey+1) node lattice whose node (i, j) has index
// i (ey+1) + j.
inline std::vector<element> element_map_rectangle(unsigned int ex, unsigned int ey) {
  std::vector<element> e;
  e.reserve((size_t)ex * ey);
  for (unsigned int i = 0; i < ex; i++)
    for (unsigned int j = 0; j < ey; j++) {
      const unsigned int ll = i * (ey + 1) + j;
      e.emplace_back(ll, ll + (ey + 1), ll + (ey + 1) + 1, ll + 1);
    }
  return e;
}

#endif
