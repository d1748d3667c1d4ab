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
// form, and its inverse: with the roles of x and y arranged so that
// a W <= b H, a triangle (g <= a W, area g^2 / 2ab), a trapezoid across the
// width (g <= b H, area W (g - a W / 2) / b) or the box less a triangle.  Not
// bit-exact: used with an error margin only.
struct HalfPlaneArea {
  double W, aW, bH, WH, ab2, bW, T1;
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
    aW = a * W_;
    bH = b * H_;
    WH = W_ * H_;
    ab2 = 2.0 * a * b;
    bW = b / W_;
    T1 = 0.5 * aW * W_ / b;  // area at g = a W (0 when a == 0)
  }
  // g with area(g) = T, for 0 < T < W H
  CV_HD double inverse(double T) const {
    if (T <= T1) return sqrt(ab2 * T);
    if (T <= WH - T1) return T * bW + 0.5 * aW;
    return (aW + bH) - sqrt(ab2 * (WH - T));
  }
};

// cut_area_w for lines whose inside corners (d_k <= 0) are a fixed set: one to
// three consecutive corners, entered across edge ea (outside -> inside) and
// left across edge eb.  The clip then always emits the same vertex sequence --
// a rotation of the cycle X_a, C_f .. C_l, X_b (crossings X, inside corners C
// from the first to the last in counter-clockwise order) that starts where
// edge 0's emission starts -- so its shoelace sum is that of the cycle's terms
//   V1 = t(X_a, C_f), K1, K2 (corner-to-corner, constant), V2 = t(C_l, X_b),
//   V3 = t(X_b, X_a),      t(p, q) = p.x q.y - q.x p.y,
// added in the rotated order from a zero start.  Only the two crossings depend
// on alpha; each is the clip's own t = -d1 / (d2 - d1), x1 + t (x2 - x1),
// y1 + t (y2 - y1).  A missing K is a zero added after a nonzero partial sum
// or to a zero one, which changes at most the sign of a zero, and the area is
// |s| / 2: area() equals cut_area_w bit for bit for every alpha that keeps the
// inside set.  The rotation (cls) is fixed by which of corners 0, 1, 2 are
// inside: corner 0 outside -> V1 first; 0 in, 1 out -> V3 first; 0, 1 in,
// 2 out -> V2 first; 0, 1, 2 in -> K2 first.
struct FixedCut {
  double wa1, wa2, xa, ya, dxa, dya;  // entry edge: corner values, start corner, x2 - x1, y2 - y1
  double wb1, wb2, xb, yb, dxb, dyb;  // exit edge
  double cfx, cfy, clx, cly;          // first and last inside corners
  double K1, K2;
  int cls;

  CV_HD static double pick(int k, double a0, double a1, double a2, double a3) {
    return k == 0 ? a0 : k == 1 ? a1 : k == 2 ? a2 : a3;
  }
  // false when the inside set is not one to three consecutive corners
  CV_HD bool init(double xmin, double xmax, double ymin, double ymax, double w0, double w1, double w2,
                  double w3, bool i0, bool i1, bool i2, bool i3) {
    const int m = (int)i0 + (int)i1 + (int)i2 + (int)i3;
    const int nin = (int)(!i0 && i1) + (int)(!i1 && i2) + (int)(!i2 && i3) + (int)(!i3 && i0);
    if (m < 1 || m > 3 || nin != 1) return false;
    const int ea = (!i0 && i1) ? 0 : (!i1 && i2) ? 1 : (!i2 && i3) ? 2 : 3;
    const int eb = (i0 && !i1) ? 0 : (i1 && !i2) ? 1 : (i2 && !i3) ? 2 : 3;
    const int f = (ea + 1) & 3, f1 = (ea + 2) & 3, f2 = (ea + 3) & 3;
    const int ea1 = f, eb1 = (eb + 1) & 3;
    wa1 = pick(ea, w0, w1, w2, w3);
    wa2 = pick(ea1, w0, w1, w2, w3);
    wb1 = pick(eb, w0, w1, w2, w3);
    wb2 = pick(eb1, w0, w1, w2, w3);
    xa = pick(ea, xmin, xmax, xmax, xmin);
    ya = pick(ea, ymin, ymin, ymax, ymax);
    dxa = pick(ea1, xmin, xmax, xmax, xmin) - xa;
    dya = pick(ea1, ymin, ymin, ymax, ymax) - ya;
    xb = pick(eb, xmin, xmax, xmax, xmin);
    yb = pick(eb, ymin, ymin, ymax, ymax);
    dxb = pick(eb1, xmin, xmax, xmax, xmin) - xb;
    dyb = pick(eb1, ymin, ymin, ymax, ymax) - yb;
    cfx = pick(f, xmin, xmax, xmax, xmin);
    cfy = pick(f, ymin, ymin, ymax, ymax);
    clx = pick(eb, xmin, xmax, xmax, xmin);
    cly = pick(eb, ymin, ymin, ymax, ymax);
    const double c1x = pick(f1, xmin, xmax, xmax, xmin), c1y = pick(f1, ymin, ymin, ymax, ymax);
    const double c2x = pick(f2, xmin, xmax, xmax, xmin), c2y = pick(f2, ymin, ymin, ymax, ymax);
    K1 = m >= 2 ? cfx * c1y - c1x * cfy : 0.0;
    K2 = m >= 3 ? c1x * c2y - c2x * c1y : 0.0;
    cls = !i0 ? 0 : !i1 ? 1 : !i2 ? 2 : 3;
    return true;
  }
  CV_HD double area(double alpha) const {
    const double a1 = wa1 + alpha, a2 = wa2 + alpha;
    const double ta = -a1 / (a2 - a1);
    const double pax = xa + ta * dxa, pay = ya + ta * dya;
    const double b1 = wb1 + alpha, b2 = wb2 + alpha;
    const double tb = -b1 / (b2 - b1);
    const double pbx = xb + tb * dxb, pby = yb + tb * dyb;
    const double V1 = pax * cfy - cfx * pay;
    const double V2 = clx * pby - pbx * cly;
    const double V3 = pbx * pay - pax * pby;
    // the four rotations, then a select (no branches in a wave whose lanes differ)
    const double s0 = (((V1 + K1) + K2) + V2) + V3;
    const double s1 = (((V3 + V1) + K1) + K2) + V2;
    const double s2 = (((V2 + V3) + V1) + K1) + K2;
    const double s3 = (((K2 + V2) + V3) + V1) + K1;
    const double s = cls == 0 ? s0 : cls == 1 ? s1 : cls == 2 ? s2 : s3;
    return 0.5 * fabs(s);
  }
};

// compute_alpha with the same result bit for bit and far fewer operations.
// The bisection's step k only needs the sign of a(mid) - target, where a is
// cut_area_w's floating-point area.  That area differs from the exact area A
// of the line's cut by at most Ef: the shoelace terms are cross products of
// vertices in absolute coordinates (|x| <= X, |y| <= Y, |p| <= Rp), each
// product rounded (2u XY per term, u = 2^-53), the terms and partial sums
// bounded by Rp * diag; the corner values n.p + mid carry at most ~4u S
// (S = |nx| X + |ny| Y) and every computed cut vertex lies within ~3 of
// those of the exact line, so the cut polygon differs from the exact one
// within a band of that width along a chord of at most diag.  A is known in
// closed form from the box corner of least n.p (HalfPlaneArea; its offset
// vmin carries ~2u S, moving the area by at most 2u S diag / |n|); E is four
// times the sum of those bounds.  Inverting the closed form at target +- 2E
// gives a window [m_lo, m_hi] of line constants: below it A > target + E
// (a(mid) > target, the step moves lo), above it A < target - E (the step
// moves hi); the window's own rounding (a few u of S and diag in mid) moves
// A by far less than E.  Only the steps whose mid falls in the window -- about
// the last 20 of 45 at cell sizes of 1e-3 -- evaluate the clip, and where the
// window keeps every corner on one side (the corner values are monotone in
// mid) they evaluate it through FixedCut.
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
  // the window of steps evaluated exactly, within the bracket (every mid lies strictly inside it)
  double m_lo = lo, m_hi = hi;
  if (cert) {
    const double E = 4.0 * 0x1p-53 * (8.0 * X * Y + 20.0 * Rp * diag + 30.0 * S * diag / nn + 8.0 * W * Hh);
    HalfPlaneArea hp;
    hp.init(W, Hh, a, b);
    const double Tp = target + 2.0 * E, Tm = target - 2.0 * E;
    if (Tp < hp.WH) {
      const double m = -hp.inverse(Tp) - vmin;
      m_lo = m > m_lo ? m : m_lo;
    }
    if (Tm > 0.0) {
      const double m = -hp.inverse(Tm) - vmin;
      m_hi = m < m_hi ? m : m_hi;
    }
  }
  FixedCut fc;
  const bool fixed = !degenerate &&
                     ((v1 + m_lo <= 0.0) == (v1 + m_hi <= 0.0)) && ((v2 + m_lo <= 0.0) == (v2 + m_hi <= 0.0)) &&
                     ((v3 + m_lo <= 0.0) == (v3 + m_hi <= 0.0)) && ((v4 + m_lo <= 0.0) == (v4 + m_hi <= 0.0)) &&
                     fc.init(xmin, xmax, ymin, ymax, v1, v2, v3, v4, v1 + m_lo <= 0.0, v2 + m_lo <= 0.0,
                             v3 + m_lo <= 0.0, v4 + m_lo <= 0.0);
  // below: a(mid) < target.  Two loops, so that the general clip's operands
  // are not held through the common one.
  if (fixed) {
    for (int it = 0; it < 45; ++it) {
      const double mid = 0.5 * (lo + hi);
      const bool below = mid < m_lo ? false : mid > m_hi ? true : fc.area(mid) < target;
      if (below)
        hi = mid;
      else
        lo = mid;
    }
  } else {
    for (int it = 0; it < 45; ++it) {
      const double mid = 0.5 * (lo + hi);
      bool below;
      if (mid < m_lo) {
        below = false;
      } else if (mid > m_hi) {
        below = true;
      } else {
        const double ar = degenerate ? 0.0 : cut_area_w(xmin, xmax, ymin, ymax, v1, v2, v3, v4, mid);
        below = ar < target;
      }
      if (below)
        hi = mid;
      else
        lo = mid;
    }
  }
  return 0.5 * (lo + hi);
}

// Full / empty short-circuit thresholds of reconstruct_plic / _elvira (:298-310).
CV_HD bool is_mixed(double c) { return !(c <= 1.0e-6) && !(c >= 1.0 - 1.0e-6); }

}  // namespace covof
