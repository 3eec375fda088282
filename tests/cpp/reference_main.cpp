// A driver written against the reference's surface (the shape of
// src/mfree_iwf.cpp:39-108): build bodies with a reference driver, run the
// main loop (each body steps and updates its outline, then the penalty
// contact of every ordered pair of bodies), read the particles back.
// Writes the final state as text:
//   <n> <time>
//   idx x y vx vy rho Sxx Sxy Syy fcx fcy  (one line per particle)
//
//   reference_main rings|rings_contact|tensile <nbox> <steps> [mixed] [vtk]
// vtk: results/out_%06d.vtk after every 10th step, as the reference's main
// loop writes them (src/mfree_iwf.cpp:82,99-103).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mfree/benchmarks/rings.h"
#include "mfree/benchmarks/tensile.h"
#include "mfree/contact.h"
#include "mfree/vtk_writer.h"

template <typename T>
static int run(std::vector<body<T>> bodies, unsigned int steps, bool vtk) {
  const unsigned int freq = 10;
  const size_t nb = bodies.size();
  for (unsigned int s = 0; s < steps; s++) {
    for (auto& b : bodies) {
      b.step();
      b.update_boundary();
    }
    for (size_t i = 0; i < nb; i++)
      for (size_t j = 0; j < nb; j++)
        if (i != j) contact_apply_body_to_boundary(bodies[i], bodies[j]);
    if (vtk && s % freq == 0) vtk_writer_write(bodies, s);
  }
  for (auto& b : bodies) {
    T** p = b.get_particles();
    printf("%u %.17g\n", b.get_num_part(), simulation_time::getInstance().get_time());
    for (unsigned int i = 0; i < b.get_num_part(); i++)
      printf("%u %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p[i]->idx, p[i]->x, p[i]->y,
             p[i]->vx, p[i]->vy, p[i]->rho, p[i]->Sxx, p[i]->Sxy, p[i]->Syy, p[i]->fcx, p[i]->fcy);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s rings|rings_contact|tensile nbox steps [mixed] [vtk]\n", argv[0]);
    return 2;
  }
  const unsigned int nbox = (unsigned int)atoi(argv[2]), steps = (unsigned int)atoi(argv[3]);
  bool vtk = false;
  for (int a = 4; a < argc; a++) {
    if (!strcmp(argv[a], "mixed")) mfree_adapter::precision() = MFREE_PRECISION_MIXED;
    if (!strcmp(argv[a], "vtk")) vtk = true;
  }
  if (!strcmp(argv[1], "rings")) return run(rings_monaghan(nbox), steps, vtk);
  if (!strcmp(argv[1], "rings_contact")) return run(rings_monaghan_contact(nbox), steps, vtk);
  if (!strcmp(argv[1], "tensile")) return run(tensile_tl(nbox), steps, vtk);
  fprintf(stderr, "unknown driver %s\n", argv[1]);
  return 2;
}
