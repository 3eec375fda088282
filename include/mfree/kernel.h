// src/kernel.h:23-29 -- kernel_result and cubic_spline (evaluated by the
// library's host implementation of src/kernel.cpp:20-55).
#ifndef MFREE_KERNEL_H_
#define MFREE_KERNEL_H_

#include "../mfree_iwf_amd.h"

typedef struct kernel_result {
  double w;
  double w_x;
  double w_y;
} kernel_result;

inline kernel_result cubic_spline(double xi, double yi, double xj, double yj, double h) {
  double o[3];
  mfree_cubic_spline(xi, yi, xj, yj, h, o);
  kernel_result r;
  r.w = o[0];
  r.w_x = o[1];
  r.w_y = o[2];
  return r;
}

#endif
