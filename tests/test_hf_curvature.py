"""Height-function curvature of the oracle (a curvature option with no
reference counterpart: 3 x 7 column heights of C, Cummins, Francois & Kothe,
Computers & Structures 83 (2005) 425-434), pinned by what the mathematics
fixes: zero on straight interfaces, 1/R on circles with exact volume
fractions (the sign of the reference's -div(n): + for a liquid drop, - for a
bubble), and invariance under exchanging x and y."""
import numpy as np
import pytest
from scipy import integrate


def _line_fractions(n, slope, y0):
    """Exact liquid fractions below y = y0 + slope * x on the unit square."""
    h = 1.0 / n
    C = np.zeros((n + 2, n + 2))
    for j in range(n):
        for i in range(n):
            xa, xb, ya, yb = i * h, (i + 1) * h, j * h, (j + 1) * h
            f = lambda x: min(max(y0 + slope * x - ya, 0.0), h)  # noqa: E731
            pts = [x for x in ((ya - y0) / slope, (yb - y0) / slope) if xa < x < xb]
            C[j + 1, i + 1] = integrate.quad(f, xa, xb, points=pts or None, epsabs=1e-15, epsrel=1e-13)[0] / h / h
    return C


def _disc_fractions(n, xc, yc, R):
    """Exact fractions of the disc (x - xc)^2 + (y - yc)^2 <= R^2 (adaptive
    quadrature of the clipped chord length, breakpoints at its kinks)."""
    h = 1.0 / n
    C = np.zeros((n + 2, n + 2))
    for j in range(n):
        for i in range(n):
            xa, xb, ya, yb = i * h, (i + 1) * h, j * h, (j + 1) * h
            # cells fully in / out by their corner distances
            d = [np.hypot(x - xc, y - yc) for x in (xa, xb) for y in (ya, yb)]
            dmin = np.hypot(max(xa - xc, 0, xc - xb), max(ya - yc, 0, yc - yb))
            if max(d) <= R:
                C[j + 1, i + 1] = 1.0
                continue
            if dmin >= R:
                continue

            def chord(x):
                s = np.sqrt(max(R * R - (x - xc) ** 2, 0.0))
                return max(0.0, min(yb, yc + s) - max(ya, yc - s))

            pts = [xc - R, xc + R]
            for yy in (ya, yb):
                if abs(yy - yc) < R:
                    w = np.sqrt(R * R - (yy - yc) ** 2)
                    pts += [xc - w, xc + w]
            pts = sorted(p for p in pts if xa < p < xb)
            C[j + 1, i + 1] = integrate.quad(chord, xa, xb, points=pts or None, epsabs=1e-15,
                                             epsrel=1e-13, limit=200)[0] / h / h
    return C


def _halo(C, n):
    C[0, :], C[-1, :] = C[1, :], C[-2, :]
    C[:, 0], C[:, -1] = C[:, 1], C[:, -2]
    return C


def _kappa(orc, C, n, method):
    m = orc.mesh(n, n, bc=(1, 1, 1, 1))
    k, _, _ = orc.curvature_csf(m, C, C, 1.0, 1, method=method)
    return k


def test_straight_line_has_zero_curvature(orc):
    n = 48
    C = _halo(_line_fractions(n, 0.3, 0.41), n)
    k = _kappa(orc, C, n, 1)
    kr = _kappa(orc, C, n, 0)
    band = (C[1:-1, 1:-1] > 1e-6) & (C[1:-1, 1:-1] < 1 - 1e-6)
    band[:, :4] = band[:, -4:] = False  # away from the side walls
    hf = k[1:-1, 1:-1][band]
    assert band.sum() > 40
    # heights of a straight line are linear: h'' = 0 up to rounding
    assert np.abs(hf).max() * (1.0 / n) < 1e-10, np.abs(hf).max()
    print(f"\nline: |kappa| dx <= {np.abs(hf).max() / n:.1e} (reference method {np.abs(kr[1:-1, 1:-1][band]).max() / n:.1e})")


@pytest.mark.parametrize("liquid_inside", [True, False], ids=["drop", "bubble"])
def test_circle_curvature_is_one_over_radius(orc, liquid_inside):
    n, R = 64, 12.0 / 64
    C = _disc_fractions(n, 0.5 + 0.3 / n, 0.5 - 0.2 / n, R)
    if not liquid_inside:
        C[1:-1, 1:-1] = 1.0 - C[1:-1, 1:-1]
    C = _halo(C, n)
    k = _kappa(orc, C, n, 1)
    kr = _kappa(orc, C, n, 0)
    band = (C[1:-1, 1:-1] > 1e-6) & (C[1:-1, 1:-1] < 1 - 1e-6)
    sign = 1.0 if liquid_inside else -1.0
    rel = k[1:-1, 1:-1][band] * R * sign - 1.0
    relr = kr[1:-1, 1:-1][band] * R * sign - 1.0
    print(f"\nR = 12 cells: height function |kR - 1| mean {np.abs(rel).mean():.2e} max {np.abs(rel).max():.2e}; "
          f"reference method mean {np.abs(relr).mean():.2e} max {np.abs(relr).max():.2e}")
    assert np.abs(rel).mean() < 0.01 and np.abs(rel).max() < 0.05
    # the reference's estimate is far from 1/R at this resolution: the option matters
    assert np.abs(relr).mean() > 5 * np.abs(rel).mean()


def test_exchanging_x_and_y_transposes_the_curvature(orc):
    n = 64
    C = _halo(_disc_fractions(n, 0.47, 0.55, 10.0 / 64), n)
    k = _kappa(orc, C, n, 1)
    kt = _kappa(orc, np.ascontiguousarray(C.T), n, 1)
    assert np.allclose(kt, k.T, rtol=0, atol=1e-12 * np.abs(k).max())


def test_exact_disc_fractions_of_the_workloads():
    """workloads.disc_cell_fractions (closed-form quadrant areas): equal to the
    adaptive quadrature of the clipped chord on every cell, and the fractions
    of a disc sum to pi R^2."""
    from covof_amd import workloads as wl
    n, R, xc, yc = 40, 9.3 / 40, 0.52, 0.47
    h = 1.0 / n
    X0, Y0 = np.meshgrid(np.arange(n) * h, np.arange(n) * h)
    Ce = wl.disc_cell_fractions(X0, X0 + h, Y0, Y0 + h, xc, yc, R)
    Cq = _disc_fractions(n, xc, yc, R)[1:-1, 1:-1]
    assert np.abs(Ce - Cq).max() < 1e-12
    assert abs(Ce.sum() * h * h - np.pi * R * R) < 1e-13
