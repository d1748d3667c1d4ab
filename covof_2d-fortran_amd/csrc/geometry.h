// Interface geometry shared by the VOF kernels (device + host compilable).
//
// Every routine evaluates the reference's IEEE operation sequence in the same
// order, and the library is compiled with -ffp-contract=off, so that results
// are bit-identical to the reference Fortran built without FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace covof {

#define CV_HD __host__ __device__ __forceinline__

// x_face / y_face of src/module_mesh.f90:162-176 for one rank.
CV_HD double face_coord(int i, double d) { return (double)(i - 1) * d; }

// Euclidean norm sqrt(x^2 + y^2) as the reference's HYPOT evaluates it.  The
// reference Fortran's HYPOT intrinsic calls the C library's hypot, which is
// not correctly rounded (about 0.501 ulp worst case), so ELVIRA's normals are
// only reproducible with the same correction.  Method: C. F. Borges, "An
// Improved Algorithm for hypot(a,b)", ACM TOMS 47(1), 2021 -- with
// a >= b >= 0, h = sqrt(a*a + b*b) is corrected once by
// h -= (h^2 - a^2 - b^2) / (2h), where the residual is rearranged around
// delta = h - b (when h <= 2b) or delta = h - a so that it is computed without
// cancellation and without fused multiply-adds.  Power-of-two scaling keeps
// extreme magnitudes finite; it is exact, so moderate arguments take the
// unscaled path.
CV_HD double hypot_ref(double x, double y) {
  double a = fabs(x), b = fabs(y);
  if (a < b) { const double t = a; a = b; b = t; }
  if (!(a <= 1.7976931348623157e308) || !(b == b)) {
    if (a == __builtin_inf() || b == __builtin_inf()) return __builtin_inf();
    return x + y;  // NaN
  }
  if (b == 0.0) return a;
  double s = 1.0;
  if (a > 0x1p+510) { a *= 0x1p-600; b *= 0x1p-600; s = 0x1p+600; }
  else if (b < 0x1p-500) { a *= 0x1p+600; b *= 0x1p+600; s = 0x1p-600; }
  double h = sqrt(a * a + b * b), t1, t2;
  if (h <= 2.0 * b) {
    const double d = h - b;
    t1 = a * (2.0 * d - a);
    t2 = (d - 2.0 * (a - b)) * d;
  } else {
    const double d = h - a;
    t1 = 2.0 * d * (a - 2.0 * b);
    t2 = (4.0 * d - b) * b + d * d;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h * s;
}

// cut_box_area (src/module_interface_reconstruction.f90:33-101): area of the
// part n.x + alpha <= 0 of the box, Sutherland-Hodgman over the corners in
// counter-clockwise order from (xmin, ymin), shoelace sum in absolute
// coordinates.  The shoelace terms are accumulated while the clipped vertices
// are emitted (term(k-1,k) as vertex k appears, term(last, first) at the end):
// the same terms added in the same order as the reference's polygon loop,
// without a vertex array.
// emit() is written with selects rather than branches (the term of the first
// vertex is formed and dropped), so the clip compiles to short straight-line
// code.
struct Shoelace {
  double s, fx, fy, px, py;
  int n;
  CV_HD void emit(double x, double y) {
    const double term = px * y - x * py;
    const bool first = n == 0;
    s = first ? 0.0 : s + term;
    fx = first ? x : fx;
    fy = first ? y : fy;
    px = x;
    py = y;
    ++n;
  }
};

// Edge k -> k+1 of the clip with corner distances d1, d2: the intersection is
// kept when the edge crosses the line (inside -> outside or outside ->
// inside), then the end corner when it is inside -- the reference's four
// cases in their emission order.
CV_HD void clip_edge(Shoelace &P, double d1, double d2, double x1, double y1, double x2,
                     double y2) {
  const bool in1 = d1 <= 0.0, in2 = d2 <= 0.0;
  if (in1 != in2) {
    const double t = -d1 / (d2 - d1);
    P.emit(x1 + t * (x2 - x1), y1 + t * (y2 - y1));
  }
  if (in2) P.emit(x2, y2);
}

// Area of the box cut by the line whose corner values without alpha are
// w0..w3 (w_k = nx*x_k + ny*y_k, corners (xmin,ymin),(xmax,ymin),(xmax,ymax),
// (xmin,ymax)).  d_k = w_k + alpha is the reference's (nx*x + ny*y) + alpha.
CV_HD double cut_area_w(double xmin, double xmax, double ymin, double ymax, double w0, double w1,
                        double w2, double w3, double alpha) {
  const double d0 = w0 + alpha, d1 = w1 + alpha, d2 = w2 + alpha, d3 = w3 + alpha;
  Shoelace P;
  P.n = 0;
  P.s = 0.0;
  P.fx = P.fy = P.px = P.py = 0.0;
  clip_edge(P, d0, d1, xmin, ymin, xmax, ymin);
  clip_edge(P, d1, d2, xmax, ymin, xmax, ymax);
  clip_edge(P, d2, d3, xmax, ymax, xmin, ymax);
  clip_edge(P, d3, d0, xmin, ymax, xmin, ymin);
  if (P.n < 3) return 0.0;
  const double s = P.s + (P.px * P.fy - P.fx * P.py);
  return 0.5 * fabs(s);
}

CV_HD double cut_box_area(double xmin, double xmax, double ymin, double ymax, double nx,
                          double ny, double alpha) {
  if ((xmax - xmin) <= 1.0e-14 || (ymax - ymin) <= 1.0e-14) return 0.0;
  return cut_area_w(xmin, xmax, ymin, ymax, nx * xmin + ny * ymin, nx * xmax + ny * ymin,
                    nx * xmax + ny * ymax, nx * xmin + ny * ymax, alpha);
}

// cut_box_area of a full cell's segment (n = (1,0), alpha = -1e9: every
// corner is inside whenever |x| < 1e9), without the clip: the four corners
// are emitted in the clip's order -- (xmax,ymin), (xmax,ymax), (xmin,ymax),
// (xmin,ymin) -- so the shoelace terms, and their order, are the clip's.
CV_HD double full_box_area(double xmin, double xmax, double ymin, double ymax) {
  if ((xmax - xmin) <= 1.0e-14 || (ymax - ymin) <= 1.0e-14) return 0.0;
  double s = 0.0;
  s = s + (xmax * ymax - xmax * ymin);
  s = s + (xmax * ymax - xmin * ymax);
  s = s + (xmin * ymin - xmin * ymax);
  s = s + (xmin * ymin - xmax * ymin);
  return 0.5 * fabs(s);
}

// compute_alpha (src/module_interface_reconstruction.f90:105-147): 45-step
// bisection of the line constant so that the kept area equals c * box area.
// The corner products n.x_k do not depend on alpha and are formed once.
CV_HD double compute_alpha(double xmin, double xmax, double ymin, double ymax, double nx,
                           double ny, double c) {
  const double total = (xmax - xmin) * (ymax - ymin);
  if (c <= 1.0e-7) return 1.0e9;
  if (c >= 1.0 - 1.0e-7) return -1.0e9;
  const double cc = c < 1.0 ? c : 1.0;
  const double target = (0.0 > cc ? 0.0 : cc) * total;
  const double v1 = nx * xmin + ny * ymin;
  const double v2 = nx * xmax + ny * ymin;
  const double v3 = nx * xmax + ny * ymax;
  const double v4 = nx * xmin + ny * ymax;
  const double m34 = v3 < v4 ? v3 : v4, M34 = v3 > v4 ? v3 : v4;
  const double m234 = v2 < m34 ? v2 : m34, M234 = v2 > M34 ? v2 : M34;
  const double vmin = v1 < m234 ? v1 : m234, vmax = v1 > M234 ? v1 : M234;
  double lo = -vmax - 1.0e-12;
  double hi = -vmin + 1.0e-12;
  const bool degenerate = (xmax - xmin) <= 1.0e-14 || (ymax - ymin) <= 1.0e-14;
  for (int it = 0; it < 45; ++it) {
    const double mid = 0.5 * (lo + hi);
    const double a = degenerate ? 0.0 : cut_area_w(xmin, xmax, ymin, ymax, v1, v2, v3, v4, mid);
    if (a < target)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Area of {a X + b Y <= g} inside [0, W] x [0, H] for a, b >= 0 in closed
// form: with the roles of x and y arranged so that a W <= b H, a triangle
// (g <= a W), a trapezoid across the width (g <= b H) or the box less a
// triangle.  The constants depend on the box and the normal only; evaluate()
// is branch-free (all three forms, then selects), so the lanes of a wave that
// sit in different cases do not serialise.  Not bit-exact: used with an
// error margin only.
struct HalfPlaneArea {
  double W, H, aW, bH, k2, ib, WH;
  CV_HD void init(double W_, double H_, double a, double b) {
    if (a * W_ > b * H_) {
      const double t = a;
      a = b;
      b = t;
      const double u = W_;
      W_ = H_;
      H_ = u;
    }
    W = W_;
    H = H_;
    aW = a * W;
    bH = b * H;
    k2 = 0.5 / (a * b);  // inf when a == 0: its form is then never selected
    ib = 1.0 / b;
    WH = W * H;
  }
  CV_HD double evaluate(double g) const {
    const double tri = g * g * k2;
    const double trap = W * (g - 0.5 * aW) * ib;
    const double r = aW + bH - g;
    const double top = WH - r * r * k2;
    double v = g <= bH ? trap : top;
    v = g <= aW ? tri : v;
    v = g >= aW + bH ? WH : v;
    return g <= 0.0 ? 0.0 : v;
  }
};

// compute_alpha with the same result bit for bit and fewer clips.  The
// bisection's step k only needs the sign of a(mid) - target, where a is
// cut_area_w's floating-point area.  That area differs from the exact area A
// of the line's cut by at most Ef: the shoelace terms are cross products of
// vertices in absolute coordinates (|x| <= X, |y| <= Y, |p| <= Rp), each
// product rounded (2u XY per term, u = 2^-53), the terms and partial sums
// bounded by Rp * diag; the corner values n.p + mid carry at most ~4u S
// (S = |nx| X + |ny| Y) and every computed cut vertex lies within ~3 of
// those of the exact line, so the cut polygon differs from the exact one
// within a band of that width along a chord of at most diag.  A is evaluated
// in closed form from the box corner of least n.p (HalfPlaneArea; its
// offset vmin carries ~2u S, moving the area by at most 2u S diag / |n|).
// Steps whose closed-form area is farther than E (four times the sum of
// those bounds) from the target have the same outcome as the clip and skip
// it; the steps near the root -- about the last 20 of 45 at cell sizes of
// 1e-3 -- evaluate cut_area_w exactly as compute_alpha does.
CV_HD double compute_alpha_fast(double xmin, double xmax, double ymin, double ymax, double nx,
                                double ny, double c) {
  const double total = (xmax - xmin) * (ymax - ymin);
  if (c <= 1.0e-7) return 1.0e9;
  if (c >= 1.0 - 1.0e-7) return -1.0e9;
  const double cc = c < 1.0 ? c : 1.0;
  const double target = (0.0 > cc ? 0.0 : cc) * total;
  const double v1 = nx * xmin + ny * ymin;
  const double v2 = nx * xmax + ny * ymin;
  const double v3 = nx * xmax + ny * ymax;
  const double v4 = nx * xmin + ny * ymax;
  const double m34 = v3 < v4 ? v3 : v4, M34 = v3 > v4 ? v3 : v4;
  const double m234 = v2 < m34 ? v2 : m34, M234 = v2 > M34 ? v2 : M34;
  const double vmin = v1 < m234 ? v1 : m234, vmax = v1 > M234 ? v1 : M234;
  double lo = -vmax - 1.0e-12;
  double hi = -vmin + 1.0e-12;
  const bool degenerate = (xmax - xmin) <= 1.0e-14 || (ymax - ymin) <= 1.0e-14;
  // certification bounds
  const double W = xmax - xmin, Hh = ymax - ymin;
  const double a = fabs(nx), b = fabs(ny);
  const double nn = sqrt(a * a + b * b);
  const double X = fabs(xmin) > fabs(xmax) ? fabs(xmin) : fabs(xmax);
  const double Y = fabs(ymin) > fabs(ymax) ? fabs(ymin) : fabs(ymax);
  const double Rp = sqrt(X * X + Y * Y), diag = sqrt(W * W + Hh * Hh), S = a * X + b * Y;
  const double amin = a < b ? a : b;
  const bool cert = !degenerate && nn > 0.5 && nn < 2.0 && X < 1.0e100 && Y < 1.0e100 &&
                    (amin == 0.0 || amin > 1.0e-150);
  const double E = 4.0 * 0x1p-53 * (8.0 * X * Y + 20.0 * Rp * diag + 30.0 * S * diag / nn + 8.0 * W * Hh);
  HalfPlaneArea hp;
  hp.init(W, Hh, a, b);
  for (int it = 0; it < 45; ++it) {
    const double mid = 0.5 * (lo + hi);
    int below = -1;  // 1: a(mid) < target, 0: a(mid) >= target
    if (cert) {
      const double ah = hp.evaluate(-(mid + vmin));
      if (ah < target - E) below = 1;
      else if (ah > target + E) below = 0;
    }
    if (below < 0) {
      const double ar = degenerate ? 0.0 : cut_area_w(xmin, xmax, ymin, ymax, v1, v2, v3, v4, mid);
      below = ar < target ? 1 : 0;
    }
    if (below)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Full / empty short-circuit thresholds of reconstruct_plic / _elvira (:298-310).
CV_HD bool is_mixed(double c) { return !(c <= 1.0e-6) && !(c >= 1.0 - 1.0e-6); }

}  // namespace covof
