// src/boundary_conditions.h:27-51: (index list, void(*)(T*)) pairs.  The
// GPU path applies velocity-strip conditions of the form used by the
// tension drivers (src/benchmarks/tensile.cpp:20-42):
//     x = X + a*t, y = Y, vx = a, vy = 0, vx_t = vy_t = 0.
// body<T> identifies a, for every registered function, by evaluating it on
// probe particles at two times; a function of any other form is rejected
// (MFREE_ERR_PIPELINE) instead of being run on the CPU.
#ifndef MFREE_BOUNDARY_CONDITIONS_H_
#define MFREE_BOUNDARY_CONDITIONS_H_

#include <tuple>
#include <vector>

template <typename T>
class boundary_conditions {
 public:
  boundary_conditions() {}
  void add_boundary_condition(const std::vector<unsigned int> boundary_indexes, void (*boundary_fun)(T*)) {
    m_bc.push_back(std::make_tuple(boundary_indexes, boundary_fun));
  }
  const std::vector<std::tuple<std::vector<unsigned int>, void (*)(T*)>>& list() const { return m_bc; }

 private:
  std::vector<std::tuple<std::vector<unsigned int>, void (*)(T*)>> m_bc;
};

#endif
