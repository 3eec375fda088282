// Reference-named results writer: vtk_writer_write (src/vtk_writer.h:27-127)
// on top of the C ABI.  Writes results/out_%06d.vtk (relative to the working
// directory, as the reference does) with every body's particles, bodies in
// order, through mfree_vtk_write_host: the reference's legacy VTK text for
// the same state.  The reference's main loop (src/mfree_iwf.cpp:99-103)
// calls it after every 10th step.
#ifndef MFREE_VTK_WRITER_H_
#define MFREE_VTK_WRITER_H_

#include <stdio.h>

#include <string>
#include <vector>

#include "body.h"

template <typename T>
void vtk_writer_write(const std::vector<body<T>>& bodies, unsigned int step) {
  const size_t nb = bodies.size();
  // SoA copies of the fields the file holds: x, y, h, rho, Sxx, Sxy, Syy, vx, vy
  std::vector<std::vector<double>> col(9 * nb);
  std::vector<mfree_particles> views(nb);
  std::vector<uint32_t> counts(nb);
  for (size_t b = 0; b < nb; b++) {
    T** p = bodies[b].get_particles();
    const unsigned int n = bodies[b].get_num_part();
    std::vector<double>* c = &col[9 * b];
    for (int k = 0; k < 9; k++) c[k].resize(n);
    for (unsigned int i = 0; i < n; i++) {
      c[0][i] = p[i]->x;
      c[1][i] = p[i]->y;
      c[2][i] = p[i]->h;
      c[3][i] = p[i]->rho;
      c[4][i] = p[i]->Sxx;
      c[5][i] = p[i]->Sxy;
      c[6][i] = p[i]->Syy;
      c[7][i] = p[i]->vx;
      c[8][i] = p[i]->vy;
    }
    mfree_particles v = {};
    v.x = c[0].data();
    v.y = c[1].data();
    v.h = c[2].data();
    v.rho = c[3].data();
    v.Sxx = c[4].data();
    v.Sxy = c[5].data();
    v.Syy = c[6].data();
    v.vx = c[7].data();
    v.vy = c[8].data();
    views[b] = v;
    counts[b] = n;
  }
  char buf[256];
  snprintf(buf, sizeof buf, "results/out_%06d.vtk", step);
  const mfree_status st = mfree_vtk_write_host(buf, (uint32_t)nb, views.data(), counts.data());
  if (st != MFREE_OK) mfree_adapter::die(std::string("vtk_writer_write: cannot write ") + buf);
}

#endif  // MFREE_VTK_WRITER_H_
