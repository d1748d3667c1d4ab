// Host build of the device geometry (covof_2d-fortran_amd/csrc/geometry.h)
// for CPU tests: the same source compiled for the host with -ffp-contract=off.
#include "geometry.h"

extern "C" double host_hypot_ref(double x, double y) { return covof::hypot_ref(x, y); }
extern "C" double host_cut_box_area(double a, double b, double c, double d, double nx, double ny, double al) {
  return covof::cut_box_area(a, b, c, d, nx, ny, al);
}
extern "C" double host_compute_alpha(double a, double b, double c, double d, double nx, double ny, double cc) {
  return covof::compute_alpha(a, b, c, d, nx, ny, cc);
}
extern "C" double host_full_box_area(double a, double b, double c, double d) {
  return covof::full_box_area(a, b, c, d);
}
