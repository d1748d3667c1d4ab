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
extern "C" double host_compute_alpha_fast(double a, double b, double c, double d, double nx, double ny,
                                          double cc) {
  return covof::compute_alpha_fast(a, b, c, d, nx, ny, cc);
}

// Random ELVIRA-like compute_alpha calls (grid cells of many resolutions and
// domain sizes, unit normals including axis-aligned and nearly axis-aligned
// ones, fractions including the thresholds): the number of calls whose
// compute_alpha_fast result differs from compute_alpha in any bit.
namespace {
struct Rng {
  unsigned long long s;
  double uni() {  // [0, 1)
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (double)(s >> 11) * 0x1p-53;
  }
};
}  // namespace

extern "C" long host_alpha_fast_mismatches(long n, unsigned long long seed) {
  Rng r{seed * 2654435761ull + 88172645463325252ull};
  const double Ls[] = {1.0, 0.6, 3.7, 1.0e-3, 250.0, 1.5};
  const int res[] = {16, 40, 128, 512, 2048, 4096, 8192};
  long bad = 0;
  for (long k = 0; k < n; ++k) {
    const double L = Ls[(int)(r.uni() * 6)];
    const int nr = res[(int)(r.uni() * 7)];
    const double h = L / nr, h2 = (r.uni() < 0.2) ? h * (0.5 + r.uni()) : h;
    const int i = (int)(r.uni() * (nr + 2)), j = (int)(r.uni() * (nr + 2));
    const double xmin = (double)(i - 1) * h, xmax = (double)i * h;
    const double ymin = (double)(j - 1) * h2, ymax = (double)j * h2;
    double th = 2.0 * 3.141592653589793 * r.uni();
    double nx = cos(th), ny = sin(th);
    const double q = r.uni();
    if (q < 0.08) { nx = (r.uni() < 0.5) ? 1.0 : -1.0; ny = 0.0; }
    else if (q < 0.16) { nx = 0.0; ny = (r.uni() < 0.5) ? 1.0 : -1.0; }
    else if (q < 0.30) {  // nearly axis-aligned
      const double e = pow(10.0, -1.0 - 11.0 * r.uni()) * ((r.uni() < 0.5) ? 1.0 : -1.0);
      const double nr2 = sqrt(1.0 + e * e);
      if (r.uni() < 0.5) { nx = 1.0 / nr2; ny = e / nr2; } else { nx = e / nr2; ny = -1.0 / nr2; }
    } else if (q < 0.36) {  // not normalised
      const double sc = 0.1 + 2.9 * r.uni();
      nx *= sc;
      ny *= sc;
    }
    double c = r.uni();
    const double w = r.uni();
    if (w < 0.05) c = 1.0e-7 * (1.0 + r.uni());
    else if (w < 0.10) c = 1.0 - 1.0e-7 * (1.0 + r.uni());
    else if (w < 0.15) c = pow(10.0, -7.0 + 6.0 * r.uni());
    else if (w < 0.20) c = 1.0 - pow(10.0, -7.0 + 6.0 * r.uni());
    else if (w < 0.22) c = 0.5;
    const double a0 = covof::compute_alpha(xmin, xmax, ymin, ymax, nx, ny, c);
    const double a1 = covof::compute_alpha_fast(xmin, xmax, ymin, ymax, nx, ny, c);
    if (!(a0 == a1) && !(a0 != a0 && a1 != a1)) ++bad;
  }
  return bad;
}
