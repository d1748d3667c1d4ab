"""The synthetic workloads' exact cell fractions (CPU only): the rotated
ellipses of the VOF workload against the exact disc-cell areas (a = b), the
ellipse's area pi a b, a fine sub-column quadrature, and the symmetry
(a, b, th) = (b, a, th + pi/2)."""
import math

import numpy as np

from covof_amd import workloads as wl


def _one(n, xc, yc, a, b, th, exact=True, sub=32):
    dx = 1.0 / n
    C = np.zeros((n + 2, n + 2))
    if exact:
        wl._ellipse_fraction_exact(C, n, n, dx, dx, xc, yc, a * dx, b * dx, th)
    else:
        wl._ellipse_fraction(C, n, n, dx, dx, xc, yc, a * dx, b * dx, th, sub=sub)
    return C


def test_circle_equals_disc_fractions():
    n, r = 128, 17.3
    dx = 1.0 / n
    C = _one(n, 0.4731, 0.5119, r, r, 0.7)
    X0, Y0 = np.meshgrid((np.arange(1, n + 1) - 1.0) * dx, (np.arange(1, n + 1) - 1.0) * dx)
    D = wl.disc_cell_fractions(X0, X0 + dx, Y0, Y0 + dx, 0.4731, 0.5119, r * dx)
    D = np.where(D < 1e-12, 0.0, np.where(D > 1.0 - 1e-12, 1.0, D))
    assert np.abs(C[1:-1, 1:-1] - D).max() < 1e-12


def test_ellipse_area_and_symmetry():
    n = 256
    dx = 1.0 / n
    for a, b, th in ((30.0, 12.0, 0.3), (7.3, 3.1, 0.0), (20.0, 45.0, 2.5), (64.0, 9.5, 1.1)):
        C = _one(n, 0.503, 0.47, a, b, th)
        assert abs(C.sum() * dx * dx - math.pi * a * b * dx * dx) <= 1e-13 * math.pi * a * b * dx * dx
        assert C.min() >= 0.0 and C.max() <= 1.0
        Cs = _one(n, 0.503, 0.47, b, a, th + 0.5 * math.pi)
        assert np.abs(C - Cs).max() < 1e-12


def test_ellipse_fractions_vs_fine_quadrature():
    """Midpoint quadrature over 4096 sub-columns per cell (exact in y) agrees to
    its own error (largest at the vertical tangents)."""
    n = 96
    C = _one(n, 0.41, 0.57, 23.0, 11.0, 0.9)
    Q = _one(n, 0.41, 0.57, 23.0, 11.0, 0.9, exact=False, sub=4096)
    assert np.abs(C - Q).max() < 2e-5


def test_vof_case_exact_by_default():
    C, u, v, dt = wl.vof_case(256, seed=3, semi=(5.0, 40.0))
    C2 = wl.ellipses(256, 256, 128, 5.0, 40.0, 3, exact=True)
    wl.fill_halo_periodic(C2, 256, 256)
    assert np.array_equal(C, C2)
