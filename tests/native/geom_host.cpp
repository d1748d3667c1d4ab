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
    if (q < 0.04) {  // diagonal: at c = 0.5 the line runs through two corners
      nx = ((r.uni() < 0.5) ? 1.0 : -1.0) * 0.7071067811865476;
      ny = ((r.uni() < 0.5) ? 1.0 : -1.0) * 0.7071067811865476;
    } else if (q < 0.08) { nx = (r.uni() < 0.5) ? 1.0 : -1.0; ny = 0.0; }
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

// FixedCut::area against cut_area_w on random boxes and lines whose inside
// set is one to three consecutive corners, a third of them with the line
// through a corner exactly (d_k = 0): the number of calls that differ in any
// bit, or whose inside set FixedCut::init rejects.
extern "C" long host_fixed_cut_mismatches(long n, unsigned long long seed) {
  Rng r{seed * 2654435761ull + 1442695040888963407ull};
  long bad = 0;
  for (long k = 0; k < n; ++k) {
    const double L = (r.uni() < 0.5) ? 1.0 : 250.0 * r.uni();
    const double h = L * pow(10.0, -4.0 * r.uni()), h2 = (r.uni() < 0.3) ? h * (0.2 + 2.0 * r.uni()) : h;
    const double xmin = L * r.uni(), ymin = L * r.uni(), xmax = xmin + h, ymax = ymin + h2;
    double th = 2.0 * 3.141592653589793 * r.uni();
    double nx = cos(th), ny = sin(th);
    if (r.uni() < 0.1) { nx = (r.uni() < 0.5) ? 0.0 : nx; ny = (nx == 0.0) ? 1.0 : 0.0; }
    const double w0 = nx * xmin + ny * ymin, w1 = nx * xmax + ny * ymin;
    const double w2 = nx * xmax + ny * ymax, w3 = nx * xmin + ny * ymax;
    const double w[4] = {w0, w1, w2, w3};
    double lo = w0, hi = w0;
    for (int q = 1; q < 4; ++q) { lo = w[q] < lo ? w[q] : lo; hi = w[q] > hi ? w[q] : hi; }
    double al = -(lo + (hi - lo) * r.uni());
    if (r.uni() < 0.33) al = -w[(int)(r.uni() * 4)];
    const bool i0 = w0 + al <= 0.0, i1 = w1 + al <= 0.0, i2 = w2 + al <= 0.0, i3 = w3 + al <= 0.0;
    const int m = (int)i0 + (int)i1 + (int)i2 + (int)i3;
    if (m < 1 || m > 3) continue;
    covof::FixedCut fc;
    if (!fc.init(xmin, xmax, ymin, ymax, w0, w1, w2, w3, i0, i1, i2, i3)) { ++bad; continue; }
    const double a0 = covof::cut_area_w(xmin, xmax, ymin, ymax, w0, w1, w2, w3, al);
    const double a1 = fc.area(al);
    if (!(a0 == a1)) ++bad;
  }
  return bad;
}
