// Links the reference's own benchmark drivers (src/benchmarks/rings.cpp and
// tensile.cpp, compiled unchanged against include/mfree) and reports, for
// each driver whose set-up this library accepts, the method body<T>::step()
// would run each body with (MFREE_METHOD_*, -1: refused).  The drivers of the
// weak-form and FEM methods refuse during set-up and are not called.
#include <stdio.h>

#include <vector>

#include "rings.h"
#include "tensile.h"

template <typename T>
static void report(const char* name, std::vector<body<T>> bodies) {
  printf("%s", name);
  for (const auto& b : bodies) printf(" %d:%u", (int)b.method(), b.get_num_part());
  printf("\n");
}

int main() {
  report("rings_monaghan", rings_monaghan(41));
  report("rings_godunov", rings_godunov(41));
  report("rings_monaghan_contact", rings_monaghan_contact(41));
  report("rings_godunov_contact", rings_godunov_contact(41));
  report("rings_corot_contact", rings_corot_contact(41));
  report("rings_potential_contact", rings_potential_contact(41));
  report("rings_tl_contact", rings_tl_contact(41));
  report("tensile_monaghan", tensile_monaghan(21));
  report("tensile_godunov", tensile_godunov(21));
  report("tensile_corot", tensile_corot(21));
  report("tensile_potential", tensile_potential(21));
  report("tensile_tl", tensile_tl(21));
  return 0;
}
