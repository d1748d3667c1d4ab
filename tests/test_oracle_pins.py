"""Pins of the CPU oracle that do not need the reference build: closed forms,
invariants and brute force that a dropped term, a wrong sign or a transposed
index would fail.  (The reference's own golden vectors pin geometry,
reconstruction and advection bit-for-bit in test_oracle_golden.py.)"""
import math

import numpy as np
import pytest
from scipy import integrate

from conftest import ROOT  # noqa: F401  (puts the oracle on sys.path)

PER, WALL, NEU = 3, 4, 2


# ------------------------------------------------------------------ geometry

def _area_by_quadrature(xmin, xmax, ymin, ymax, nx, ny, alpha):
    """Area of {n.x + alpha <= 0} inside the box by integrating the kept
    height over x (kinks passed to quad)."""
    h = ymax - ymin

    def height(x):
        if ny == 0.0:
            return h if nx * x + alpha <= 0.0 else 0.0
        yc = (-alpha - nx * x) / ny
        return float(np.clip(yc - ymin, 0.0, h)) if ny > 0 else float(np.clip(ymax - yc, 0.0, h))

    pts = []
    if nx != 0.0:
        for y in (ymin, ymax):
            x = (-alpha - ny * y) / nx
            if xmin < x < xmax:
                pts.append(x)
    val, _ = integrate.quad(height, xmin, xmax, points=pts or None, epsabs=1e-15, epsrel=1e-13, limit=200)
    return val


def test_cut_box_area_closed_forms(orc):
    A = orc.cut_box_area
    # axis-aligned cut x <= 0.3 of [0.1,0.7]x[0.2,0.45]: (0.3-0.1)*0.25
    assert A(0.1, 0.7, 0.2, 0.45, 1.0, 0.0, -0.3) == pytest.approx(0.2 * 0.25, rel=1e-14)
    # y >= 0.4 kept with n = (0,-1): (0.45-0.4)*0.6
    assert A(0.1, 0.7, 0.2, 0.45, 0.0, -1.0, 0.4) == pytest.approx(0.05 * 0.6, rel=1e-13)
    # diagonal through the centre of the unit box: one half
    s = 1.0 / math.sqrt(2.0)
    assert A(0.0, 1.0, 0.0, 1.0, s, s, -s) == pytest.approx(0.5, rel=1e-14)
    # corner triangle x + y <= t: t^2/2
    t = 0.3
    assert A(0.0, 1.0, 0.0, 1.0, s, s, -t * s) == pytest.approx(t * t / 2, rel=1e-13)
    # full / empty and degenerate boxes
    assert A(0.0, 1.0, 0.0, 2.0, 0.6, 0.8, -1e9) == pytest.approx(2.0, rel=1e-15)
    assert A(0.0, 1.0, 0.0, 2.0, 0.6, 0.8, 1e9) == 0.0
    assert A(0.3, 0.3 + 1e-15, 0.0, 1.0, 1.0, 0.0, -10.0) == 0.0
    assert A(0.0, 1.0, 0.5, 0.5, 1.0, 0.0, -10.0) == 0.0


def test_cut_box_area_vs_quadrature(orc):
    rng = np.random.default_rng(0)
    for _ in range(300):
        x0, y0 = rng.uniform(-2, 2, size=2)
        w, h = rng.uniform(0.01, 1.5, size=2)
        th = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(th), math.sin(th)
        # a line through a random point of the box
        px, py = x0 + rng.uniform(0, w), y0 + rng.uniform(0, h)
        alpha = -(nx * px + ny * py)
        a = orc.cut_box_area(x0, x0 + w, y0, y0 + h, nx, ny, alpha)
        q = _area_by_quadrature(x0, x0 + w, y0, y0 + h, nx, ny, alpha)
        assert abs(a - q) <= 1e-11 * w * h, (a, q)


def test_compute_alpha_round_trip_and_thresholds(orc):
    rng = np.random.default_rng(1)
    for _ in range(300):
        x0, y0 = rng.uniform(-1, 1, size=2)
        w, h = rng.uniform(0.01, 0.5, size=2)
        th = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(th), math.sin(th)
        c = rng.uniform(1e-6, 1 - 1e-6)
        al = orc.compute_alpha(x0, x0 + w, y0, y0 + h, nx, ny, c)
        a = orc.cut_box_area(x0, x0 + w, y0, y0 + h, nx, ny, al)
        # 45 halvings of a bracket of width <= |n|_1 * (w + h) + 2e-12; the area
        # moves by at most the chord length (<= hypot(w, h)) per unit of alpha
        bracket = (abs(nx) * w + abs(ny) * h + 2e-12) * 2.0 ** -45
        assert abs(a - c * w * h) <= bracket * math.hypot(w, h) + 1e-15, (a, c * w * h)
    # monotone in C
    cs = np.linspace(0.01, 0.99, 50)
    al = [orc.compute_alpha(0.0, 1.0, 0.0, 1.0, 0.6, 0.8, c) for c in cs]
    assert all(a1 > a2 for a1, a2 in zip(al, al[1:]))
    assert orc.compute_alpha(0.0, 1.0, 0.0, 1.0, 0.6, 0.8, 1e-7) == 1e9
    assert orc.compute_alpha(0.0, 1.0, 0.0, 1.0, 0.6, 0.8, 1 - 1e-7) == -1e9


def _line_field(orc, n, nx_, ny_, alpha):
    """Exact volume fractions of the half-plane n.x + alpha <= 0 on the unit
    square with n x n cells (halo ring included)."""
    dx = 1.0 / n
    C = np.zeros((n + 2, n + 2))
    for j in range(n + 2):
        for i in range(n + 2):
            x0, y0 = (i - 1) * dx, (j - 1) * dx
            C[j, i] = orc.cut_box_area(x0, x0 + dx, y0, y0 + dx, nx_, ny_, alpha) / (dx * dx)
    return C


def test_elvira_reproduces_straight_lines(orc):
    """ELVIRA's defining property: a straight line is reconstructed exactly in
    every mixed cell whose 3x3 block it crosses from side to side (through
    both vertical or both horizontal block edges), where the column (row)
    sums of the block are exact heights.

    The reference's height-function candidates are (-s_y dH/dx, s_y) and
    (s_x, -s_x dW/dy) (src/module_interface_reconstruction.f90:395-401): the
    sign factor on the slope is right when the liquid lies below (left of) the
    line, n_y > 0 (n_x > 0), and inverts the slope otherwise, so only those
    orientations are exact; the oracle keeps that behaviour (the golden
    vectors pin it bit-for-bit) and this test asserts both sides of it."""
    n = 16
    dx = 1.0 / n
    m = orc.mesh(n, n, bc=(NEU, NEU, NEU, NEU))
    total = 0
    for th, off in ((0.3, 0.45), (1.1, 0.52), (2.0, 0.1), (5.9, -0.55), (0.7, 0.2), (2.8, 0.1), (4.3, 0.0)):
        nx_, ny_ = math.cos(th), math.sin(th)
        alpha = -(nx_ * 0.5 + ny_ * 0.5) + off * 0.3
        C = _line_field(orc, n, nx_, ny_, alpha)
        sx, sy, sa = orc.reconstruct_interface(m, orc.ELVIRA, C)
        checked = 0
        for j in range(2, n):
            for i in range(2, n):
                if not (1e-6 < C[j, i] < 1 - 1e-6):
                    continue
                xlo, xhi, ylo, yhi = (i - 2) * dx, (i + 1) * dx, (j - 2) * dx, (j + 1) * dx
                through_x = ny_ != 0 and all(ylo < (-alpha - nx_ * x) / ny_ < yhi for x in (xlo, xhi))
                through_y = nx_ != 0 and all(xlo < (-alpha - ny_ * y) / nx_ < xhi for y in (ylo, yhi))
                if not (through_x or through_y):
                    continue
                exact = abs(sx[j, i] - nx_) < 1e-9 and abs(sy[j, i] - ny_) < 1e-9
                if (through_x and ny_ > 0) or (through_y and nx_ > 0):
                    assert exact, (i, j, sx[j, i], sy[j, i])
                    assert abs(sa[j, i] - alpha) < 1e-9
                    checked += 1
                elif abs(nx_ * ny_) > 0.05:
                    assert not exact, (th, i, j)
        total += checked
    assert total > 40


def test_youngs_normal_is_gradient_direction_of_linear_field(orc):
    """For a linear C field (all cells mixed) the Youngs stencil is the exact
    gradient, so the PLIC normal is -grad C / |grad C|."""
    n = 12
    m = orc.mesh(n, n, bc=(NEU, NEU, NEU, NEU))
    gx, gy = 0.013, -0.021
    jj, ii = np.meshgrid(np.arange(n + 2), np.arange(n + 2), indexing="ij")
    C = 0.5 + gx * (ii - 6.5) + gy * (jj - 6.5)
    assert C.min() > 0.1 and C.max() < 0.9
    sx, sy, _ = orc.reconstruct_interface(m, orc.PLIC, C)
    g = math.hypot(gx, gy)
    assert np.allclose(sx[1:n + 1, 1:n + 1], -gx / g, atol=1e-12)
    assert np.allclose(sy[1:n + 1, 1:n + 1], -gy / g, atol=1e-12)


@pytest.mark.parametrize("method", ["PLIC", "ELVIRA"])
def test_wy_volume_conservation_vortex(orc, method):
    """Discretely divergence-free (nodal stream function) periodic velocity:
    the total volume telescopes to rounding, |dV|/V < 1e-12."""
    n = 48
    m = orc.mesh(n, n, bc=(PER,) * 4)
    dx = 1.0 / n
    xs = np.arange(n + 3) * dx - dx
    X, Y = np.meshgrid(xs, xs)  # node (i, j) at ((i-1) dx, (j-1) dy)
    psi = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) / (2 * np.pi)
    u = (psi[1:n + 3, 0:n + 2] - psi[0:n + 2, 0:n + 2]) / dx
    v = -(psi[0:n + 2, 1:n + 3] - psi[0:n + 2, 0:n + 2]) / dx
    C = _line_field(orc, n, 0.6, 0.8, -0.55)
    C[0, :], C[n + 1, :] = C[n, :], C[1, :]
    C[:, 0], C[:, n + 1] = C[:, n], C[:, 1]
    dt = 0.25 * dx / max(abs(u).max(), abs(v).max())
    V0 = C[1:n + 1, 1:n + 1].sum()
    sc = 0
    meth = orc.PLIC if method == "PLIC" else orc.ELVIRA
    for _ in range(20):
        C, sc = orc.wy_step(m, meth, C, u, v, dt, sc)
    assert abs(C[1:n + 1, 1:n + 1].sum() - V0) / V0 < 1e-12
    assert C.min() >= 0.0 and C.max() <= 1.0


def test_wy_exact_translation(orc):
    """Uniform u with u dt = dx moves every fraction by exactly one cell."""
    n = 24
    m = orc.mesh(n, n, bc=(PER,) * 4)
    dx = 1.0 / n
    C = _line_field(orc, n, 0.8, -0.6, -0.1)
    C = np.clip(C + 0.0, 0, 1)
    C[0, :], C[n + 1, :] = C[n, :], C[1, :]
    C[:, 0], C[:, n + 1] = C[:, n], C[:, 1]
    u = np.ones((n + 2, n + 2))
    v = np.zeros((n + 2, n + 2))
    C1, _ = orc.wy_step(m, orc.PLIC, C, u, v, dx, 0)
    assert np.allclose(C1[1:n + 1, 2:n + 1], C[1:n + 1, 1:n], atol=1e-12)
    assert np.allclose(C1[1:n + 1, 1], C[1:n + 1, n], atol=1e-12)


# ------------------------------------------------------------------- Poisson

def _dense(orc, m, co):
    n = m.nx * m.ny
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = orc.poisson_matvec(m, co, e.reshape(m.ny, m.nx)).ravel()
    return A


def _assemble(orc, nx, ny, bc, rho_u=None, rho_v=None, seed=0):
    m = orc.mesh(nx, ny, bc=bc)
    shp = (ny + 2, nx + 2)
    rng = np.random.default_rng(seed)
    u, v = rng.normal(size=shp), rng.normal(size=shp)
    ru = np.ones(shp) if rho_u is None else rho_u
    rv = np.ones(shp) if rho_v is None else rho_v
    return m, orc.poisson_assemble(m, u, v, ru, rv, 1.0), (u, v, ru, rv)


@pytest.mark.parametrize("bc", [(NEU,) * 4, (PER,) * 4, (WALL,) * 4])
def test_constant_coefficient_spectrum(orc, bc):
    """rho = 1, dt = 1, dx = dy: the operator's eigenvalues are the closed forms
    of the 5-point Laplacian with ghost p = 0 (NEUMANN velocity BC), periodic
    wrap, or dropped wall faces, plus the 1e-12*dvol shift."""
    nx, ny = 7, 5
    m, co, _ = _assemble(orc, nx, ny, bc)
    lx, ly = 1.0 / nx, 1.0 / ny
    scale = ly / lx  # k_x = dt*dy/dx, k_y = dt*dx/dy
    kx, ky = scale, 1.0 / scale
    if bc[0] == NEU:
        ex = [2 - 2 * math.cos(k * math.pi / (nx + 1)) for k in range(1, nx + 1)]
        ey = [2 - 2 * math.cos(k * math.pi / (ny + 1)) for k in range(1, ny + 1)]
    elif bc[0] == PER:
        ex = [2 - 2 * math.cos(2 * math.pi * k / nx) for k in range(nx)]
        ey = [2 - 2 * math.cos(2 * math.pi * k / ny) for k in range(ny)]
    else:
        ex = [2 - 2 * math.cos(math.pi * k / nx) for k in range(nx)]
        ey = [2 - 2 * math.cos(math.pi * k / ny) for k in range(ny)]
    want = np.sort([kx * a + ky * b + 1e-12 * lx * ly for a in ex for b in ey])
    A = _dense(orc, m, co)
    assert np.array_equal(A, A.T)
    got = np.sort(np.linalg.eigvalsh(A))
    assert np.allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("bc", [(PER, PER, WALL, WALL), (WALL, NEU, PER, PER), (WALL,) * 4])
def test_variable_coefficient_quadratic_form(orc, bc):
    """x.Ax = sum over faces k_f (x_L - x_R)^2 (+ k_f x^2 on NEUMANN edge faces)
    + 1e-12*dvol |x|^2, with k_f = dt*dy/(rho_u dx) on x faces and
    dt*dx/(rho_v dy) on y faces; and A is symmetric positive definite."""
    nx, ny = 6, 5
    rng = np.random.default_rng(3)
    shp = (ny + 2, nx + 2)
    ru, rv = rng.uniform(1, 1000, size=shp), rng.uniform(1, 1000, size=shp)
    # face densities of a periodic direction come from a wrapped halo: the
    # first and last faces are the same face
    if bc[0] == PER:
        ru[:, nx + 1] = ru[:, 1]
    if bc[2] == PER:
        rv[ny + 1, :] = rv[1, :]
    m, co, _ = _assemble(orc, nx, ny, bc, ru, rv)
    A = _dense(orc, m, co)
    assert np.array_equal(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0
    dx, dy = 1.0 / nx, 1.0 / ny
    x = rng.normal(size=(ny, nx))
    q = 1e-12 * dx * dy * (x * x).sum()
    for j in range(1, ny + 1):
        for i in range(1, nx + 2):  # x face i between cells i-1 and i
            k = dy / (ru[j, i] * dx)
            if 1 < i <= nx:
                q += k * (x[j - 1, i - 2] - x[j - 1, i - 1]) ** 2
            elif i == 1 and bc[0] == PER:
                q += k * (x[j - 1, nx - 1] - x[j - 1, 0]) ** 2
            elif i == 1 and bc[0] == NEU:
                q += k * x[j - 1, 0] ** 2
            elif i == nx + 1 and bc[1] == NEU:
                q += k * x[j - 1, nx - 1] ** 2
    for j in range(1, ny + 2):
        for i in range(1, nx + 1):
            k = dx / (rv[j, i] * dy)
            if 1 < j <= ny:
                q += k * (x[j - 2, i - 1] - x[j - 1, i - 1]) ** 2
            elif j == 1 and bc[2] == PER:
                q += k * (x[ny - 1, i - 1] - x[0, i - 1]) ** 2
            elif j == 1 and bc[2] == NEU:
                q += k * x[0, i - 1] ** 2
            elif j == ny + 1 and bc[3] == NEU:
                q += k * x[ny - 1, i - 1] ** 2
    xa = (x.ravel() * (A @ x.ravel())).sum()
    assert xa == pytest.approx(q, rel=1e-12)


def test_sgs_is_explicit_symmetric_gauss_seidel(orc):
    """Walls: z = SGS(r) solves (D - L) D^-1 (D - U) z = r with D, -L, -U the
    diagonal and strict lower / upper parts of A in lexicographic order."""
    nx, ny = 6, 5
    rng = np.random.default_rng(4)
    shp = (ny + 2, nx + 2)
    m, co, _ = _assemble(orc, nx, ny, (WALL,) * 4, rng.uniform(1, 1000, size=shp), rng.uniform(1, 1000, size=shp))
    A = _dense(orc, m, co)
    D = np.diag(np.diag(A))
    Lm, Um = -np.tril(A, -1), -np.triu(A, 1)
    M = (D - Lm) @ np.linalg.inv(D) @ (D - Um)
    r = rng.normal(size=(ny, nx))
    z = orc.poisson_sgs(m, co, r)
    zr = np.linalg.solve(M, r.ravel())
    assert np.allclose(z.ravel(), zr, rtol=0, atol=1e-12 * np.abs(zr).max())
    # and Jacobi is r / a_P
    zj = orc.poisson_precond(m, 1, co, r)
    assert np.array_equal(zj, r / co["aP"])


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_sp_pcg_and_bicgstab2_vs_direct_solve(orc, precond):
    nx, ny = 10, 8
    rng = np.random.default_rng(5)
    shp = (ny + 2, nx + 2)
    ru, rv = rng.uniform(1, 100, size=shp), rng.uniform(1, 100, size=shp)
    rv[ny + 1, :] = rv[1, :]
    # SGS with a periodic wrap is not a symmetric preconditioner (the forward
    # sweep reads the wrapped value as 0, the backward sweep its forward
    # value), so the SGS case runs on walls
    bc = (WALL,) * 4 if precond == 2 else (WALL, WALL, PER, PER)
    m, co, _ = _assemble(orc, nx, ny, bc, ru, rv)
    A = _dense(orc, m, co)
    b = rng.normal(size=(ny, nx))
    b -= b.mean()
    xd = np.linalg.solve(A, b.ravel())
    xd -= xd.mean()
    x, it, rr = orc.sp_pcg(m, precond, co, b, np.zeros((ny, nx)), tol=1e-13, maxit=5000)
    assert rr <= 1e-13 and it > 0
    assert np.allclose(x.ravel() - x.mean(), xd, rtol=0, atol=1e-9 * np.abs(xd).max())
    if precond != 2:
        xb, itb, rb = orc.bicgstab2(m, precond, co, b, np.zeros((ny, nx)), tol=1e-13, maxit=5000)
        assert np.allclose(xb.ravel() - xb.mean(), xd, rtol=0, atol=1e-8 * np.abs(xd).max())


def test_projection_makes_velocity_discretely_divergence_free(orc):
    """After p solves A p = b(u*) and the correction u -= (dt/rho_u) dp/dx,
    v -= (dt/rho_v) dp/dy, the discrete divergence is the solve residual."""
    nx, ny = 12, 10
    rng = np.random.default_rng(6)
    shp = (ny + 2, nx + 2)
    ru, rv = rng.uniform(1, 1000, size=shp), rng.uniform(1, 1000, size=shp)
    m = orc.mesh(nx, ny, bc=(WALL,) * 4)
    u, v = rng.normal(size=shp), rng.normal(size=shp)
    u[:, 1] = u[:, nx + 1] = 0.0  # wall faces carry no flux
    v[1, :] = v[ny + 1, :] = 0.0
    dt = 1e-2
    co = orc.poisson_assemble(m, u, v, ru, rv, dt)
    x, _, _ = orc.sp_pcg(m, 1, co, co["b"], np.zeros((ny, nx)), tol=1e-14, maxit=10000)
    p = orc.sp_finish_pressure(m, x, np.zeros(shp))
    un, vn = orc.sp_correct(m, p, ru, rv, dt, u, v)
    dx, dy = 1.0 / nx, 1.0 / ny

    def div(uu, vv):
        return (uu[1:ny + 1, 2:nx + 2] - uu[1:ny + 1, 1:nx + 1]) * dy + (vv[2:ny + 2, 1:nx + 1] - vv[1:ny + 1, 1:nx + 1]) * dx

    d0, d1 = np.abs(div(u, v)).max(), np.abs(div(un, vn)).max()
    assert d1 <= 1e-9 * d0, (d0, d1)  # a sign or index slip leaves O(d0)


# ------------------------------------------------- smoothing, curvature, CSF

def _circle_profile(n, R, w, lx=1.0):
    """Smooth indicator tanh((R - r) / w) / 2 + 1/2 of a centred disc (halo field)."""
    h = lx / n
    xc = (np.arange(n + 2) - 0.5) * h
    X, Y = np.meshgrid(xc, xc)
    r = np.hypot(X - 0.5 * lx, Y - 0.5 * lx)
    return 0.5 + 0.5 * np.tanh((R - r) / w)


def test_curvature_of_a_disc_and_a_plane(orc):
    """kappa = -div(n) with n = grad G / |grad G|: +1/r on a smooth disc
    (positive for the liquid side; second-order: the error falls 3x as the
    profile widens from 6 to 10 cells); 0 on a planar (linear) field away from
    the walls."""
    n, R = 128, 0.3
    m = orc.mesh(n, n, 1.0, 1.0, (WALL,) * 4)
    h = 1.0 / n
    xc = (np.arange(n + 2) - 0.5) * h
    X, Y = np.meshgrid(xc, xc)
    r = np.hypot(X - 0.5, Y - 0.5)
    ring = np.abs(r - R) < 0.5 * h
    assert ring.sum() > 100
    errs = []
    for w in (0.05, 0.08):
        curv, _, _ = orc.curvature_csf(m, _circle_profile(n, R, w), _circle_profile(n, R, w), 1.0)
        errs.append(np.abs(curv[ring] * r[ring] - 1.0).max())
    assert errs[0] < 0.01 and errs[1] < errs[0] / 2.5, errs
    G = _circle_profile(n, R, 0.02)
    curv, fx, fy = orc.curvature_csf(m, G, G, 1.0)
    # fx = sigma * kappa_face * dC/dx: with sigma = 1 and kappa ~ 1/R it points to the centre
    assert (fx[ring[:, :]] * np.sign(X[ring] - 0.5 + 1e-9) <= 1e-9).all()
    # planar field: constant normal -> zero divergence
    Gp = 0.37 * X + 0.11 * Y
    curv, fx, fy = orc.curvature_csf(m, Gp, Gp, 2.0)
    inner = (slice(2, n), slice(2, n))
    assert np.abs(curv[inner]).max() < 1e-9
    # the force is the face curvature times the face gradient (sigma scales it)
    curv2, fx2, fy2 = orc.curvature_csf(m, G, G, 3.0)
    curv1, fx1, fy1 = orc.curvature_csf(m, G, G, 1.0)
    assert np.array_equal(curv1, curv2)
    assert np.allclose(fx2, 3.0 * fx1, rtol=1e-15, atol=0) and np.allclose(fy2, 3.0 * fy1, rtol=1e-15, atol=0)


def test_curvature_mirror_symmetry_and_switch_off(orc):
    """Mirroring the field in x mirrors the curvature and flips fx (a wrong
    index or sign in one direction breaks it); sigma <= 0 zeroes all three."""
    n = 48
    m = orc.mesh(n, n, 1.0, 1.0, (PER,) * 4)
    rng = np.random.default_rng(4)
    G = _circle_profile(n, 0.2, 0.03) + 0.05 * rng.random((n + 2, n + 2))
    G[:, 0], G[:, n + 1] = G[:, n], G[:, 1]
    G[0, :], G[n + 1, :] = G[n, :], G[1, :]
    curv, fx, fy = orc.curvature_csf(m, G, G, 1.0)
    Gm = G[:, ::-1].copy()  # i -> n + 1 - i
    cm, fxm, fym = orc.curvature_csf(m, Gm, Gm, 1.0)
    assert np.allclose(cm[1:n + 1, 1:n + 1], curv[1:n + 1, 1:n + 1][:, ::-1], rtol=0, atol=1e-9)
    # x-faces: face i of the mirror is face n + 2 - i of the original
    assert np.allclose(fxm[1:n + 1, 1:n + 2], -fx[1:n + 1, 1:n + 2][:, ::-1], rtol=0, atol=1e-7)
    assert np.allclose(fym[1:n + 2, 1:n + 1], fy[1:n + 2, 1:n + 1][:, ::-1], rtol=0, atol=1e-7)
    z = orc.curvature_csf(m, G, G, 0.0, fx0=np.ones_like(G), fy0=np.ones_like(G))
    assert all(np.array_equal(a, np.zeros_like(G)) for a in z)


def test_diffusion_smoothing_constant_and_fourier_mode(orc):
    """Explicit diffusion keeps a constant bitwise and damps a periodic
    Fourier mode by (1 - dtau * 4/h^2 * sin^2(pi k / n)) per step."""
    n, steps = 64, 5
    m = orc.mesh(n, n, 1.0, 1.0, (PER,) * 4)
    Cc = np.full((n + 2, n + 2), 0.625)
    assert np.array_equal(orc.smooth_diffusion(m, Cc, steps), Cc)
    h = 1.0 / n
    k = 3
    xc = (np.arange(n + 2) - 0.5) * h
    mode = np.cos(2 * np.pi * k * xc)[None, :] * np.ones((n + 2, 1))
    out = orc.smooth_diffusion(m, mode, steps)
    dtau = 0.2 * h * h
    g = 1.0 - dtau * 4.0 / h ** 2 * math.sin(math.pi * k / n) ** 2
    assert np.abs(out - g ** steps * mode).max() < 1e-13
    # walls: zero-gradient halos, total mass conserved (the flux through a wall is zero)
    mw = orc.mesh(n, n, 1.0, 1.0, (WALL,) * 4)
    C = _circle_profile(n, 0.2, 0.01)
    out = orc.smooth_diffusion(mw, C, steps)
    assert abs(out[1:n + 1, 1:n + 1].sum() - C[1:n + 1, 1:n + 1].sum()) < 1e-10
    assert np.array_equal(out[0, 1:n + 1], out[1, 1:n + 1]) and np.array_equal(out[1:n + 1, 0], out[1:n + 1, 1])


def test_svof_smoothing_solves_the_helmholtz_filter(orc):
    """25 symmetric Gauss-Seidel sweeps of (I - D Lap) C_s = C: on a periodic
    Fourier mode C_s approaches C / (1 + D * 4/h^2 * sin^2(pi k / n)); a
    constant is kept; the result is clamped to [0, 1]."""
    n, width = 64, 1.0
    m = orc.mesh(n, n, 1.0, 1.0, (PER,) * 4)
    h = 1.0 / n
    k = 2
    xc = (np.arange(n + 2) - 0.5) * h
    mode = 0.5 + 0.25 * np.cos(2 * np.pi * k * xc)[None, :] * np.ones((n + 2, 1))
    out = orc.svof_smooth(m, mode, width)
    lam = 1.0 + width * 4.0 * math.sin(math.pi * k / n) ** 2  # D / h^2 = width
    ref = 0.5 + 0.25 * np.cos(2 * np.pi * k * xc)[None, :] / lam * np.ones((n + 2, 1))
    assert np.abs(out - ref).max() < 1e-6
    Cc = np.full((n + 2, n + 2), 0.4)
    assert np.abs(orc.svof_smooth(m, Cc, width) - 0.4).max() < 1e-15
    mw = orc.mesh(n, n, 1.0, 1.0, (WALL,) * 4)
    C = (_circle_profile(n, 0.25, 0.001) > 0.5).astype(float)
    out = orc.svof_smooth(mw, C, 2.0)
    assert out.min() >= 0.0 and out.max() <= 1.0
    assert 0.0 < out[n // 2, n // 2 + int(0.25 * n)] < 1.0  # the jump is smoothed


# ------------------------------------------------------------ momentum block

def _mom_case(orc, nx, ny, bc, opts, seed=0, mu_scale=1.0, lx=1.0, ly=1.0):
    rng = np.random.default_rng(seed)
    shp = (ny + 2, nx + 2)
    f = {k: rng.standard_normal(shp) for k in ("u", "v", "u_old", "v_old", "u_nm1", "v_nm1", "fx", "fy")}
    f["rho_u"] = 1.0 + rng.random(shp)
    f["rho_v"] = 1.0 + rng.random(shp)
    f["mu"] = mu_scale * (0.01 + 0.02 * rng.random(shp))
    f["mu_node"] = mu_scale * (0.01 + 0.02 * rng.random(shp))
    return orc.Momentum(orc.mesh(nx, ny, lx, ly, bc=bc), f, opts, 1e-2), f


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 1, 1), (3, 3, 3, 3), (1, 2, 4, 1)],
                         ids=["walls", "channel", "periodic", "inflow-outflow"])
def test_momentum_diag_is_the_operator_diagonal(orc, bc):
    """Away from the boundary rows, diag_u / diag_v of the preconditioner
    blocks equal the diagonal entries A e_k of the operator for Stokes and
    for upwind convection with Picard linearisation (the upwind flux puts
    rho max(F,0) on the diagonal); the two functions are written apart."""
    nx, ny = 12, 10
    for opts in (dict(enable_convection=0), dict(enable_convection=1, conv_scheme=1, conv_lin=1)):
        M, _ = _mom_case(orc, nx, ny, bc, opts, seed=1)
        d = M.diag()
        for k in range(M.n):
            if k < M.u_nx * ny:
                i, j = k % M.u_nx + 1, k // M.u_nx + 1
                inner = 3 <= i <= nx - 2 and 3 <= j <= ny - 2
            else:
                kv = k - M.u_nx * ny
                i, j = kv % nx + 1, kv // nx + 1
                inner = 3 <= i <= nx - 2 and 3 <= j <= ny - 2
            if not inner:
                continue
            e = np.zeros(M.n)
            e[k] = 1.0
            akk = M.matvec(e)[k]
            assert abs(akk - d[k]) <= 1e-12 * abs(d[k]), (opts, k, akk, d[k])


def test_momentum_operator_closed_forms(orc):
    """Mass term: with zero viscosity and no convection, A x = (rho/dt) dvol x
    on every interior row; a uniform velocity in a periodic box feels no
    viscous stress (A x = mass x exactly); the operator is linear in x."""
    nx, ny = 10, 8
    M, f = _mom_case(orc, nx, ny, (3, 3, 3, 3), dict(enable_convection=0), seed=2, mu_scale=0.0)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(M.n)
    y = M.matvec(x)
    dvol = (1.0 / nx) * (1.0 / ny)
    mass_u = (f["rho_u"][1:ny + 1, 1:nx + 1] / 1e-2 * dvol).reshape(-1)
    mass_v = (f["rho_v"][1:ny + 1, 1:nx + 1] / 1e-2 * dvol).reshape(-1)
    mass = np.concatenate([mass_u, mass_v])
    assert np.allclose(y, mass * x, rtol=1e-14, atol=0)
    M2, f2 = _mom_case(orc, nx, ny, (3, 3, 3, 3), dict(enable_convection=0), seed=4)
    c = np.full(M2.n, 0.7)
    mass2 = np.concatenate([(f2["rho_u"][1:ny + 1, 1:nx + 1] / 1e-2 * dvol).reshape(-1),
                            (f2["rho_v"][1:ny + 1, 1:nx + 1] / 1e-2 * dvol).reshape(-1)])
    assert np.allclose(M2.matvec(c), mass2 * 0.7, rtol=1e-14, atol=0)
    M3, _ = _mom_case(orc, nx, ny, (1, 2, 4, 1), dict(conv_scheme=3, conv_lin=3), seed=5)
    a, b = rng.standard_normal(M3.n), rng.standard_normal(M3.n)
    lhs = M3.matvec(2.0 * a - 3.0 * b)
    rhs = 2.0 * M3.matvec(a) - 3.0 * M3.matvec(b)
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max()


@pytest.mark.parametrize("solver", [1, 3], ids=["bicgstab2", "cg"])
def test_momentum_solve_vs_dense(orc, solver):
    """solve_linear_system on the momentum operator against a dense solve of
    the matrix assembled column by column from the operator."""
    nx, ny = 9, 7
    M, _ = _mom_case(orc, nx, ny, (1, 1, 1, 1), dict(enable_convection=0), seed=6)
    A = np.column_stack([M.matvec(np.eye(M.n)[k]) for k in range(M.n)])
    b = M.rhs()
    x, it, _ = M.solve(b, np.zeros(M.n), solver, 1, 1e-13, 4000)
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max(), it


def test_threaded_wy_step_is_the_serial_one(orc):
    """The CPU baseline runs the oracle's reconstruction / WY cell loops on
    several threads: their iterations are independent, so the step is the
    serial one bit for bit."""
    from covof_amd import workloads as wl
    n = 96
    C = wl.ellipses(n, n, 10, 3.0, 14.0, seed=5)
    wl.fill_halo_periodic(C, n, n)
    u, v = wl.stream_velocity(n, n, 6, seed=6, periodic=True)
    dt = 0.25 / n / max(abs(u).max(), abs(v).max())
    m = orc.mesh(n, n, bc=(3, 3, 3, 3))
    out = []
    for threads in (1, 4):
        orc.set_threads(threads)
        Cw, sc = C.copy(), 0
        for _ in range(3):
            Cw, sc = orc.wy_step(m, orc.ELVIRA, Cw, u, v, dt, sc)
        out.append(Cw)
    orc.set_threads(1)
    assert np.array_equal(out[0], out[1])


def test_cpu_baseline_solve_converges(orc):
    """The timed CPU baseline solve (OpenMP Jacobi-PCG) reaches the tolerance
    with the iteration count of the serial SP loop within a few."""
    from covof_amd import workloads as wl
    n = 64
    case = wl.poisson_case(n, seed=0, n_circles=4, rmin=4, rmax=12)
    m = orc.mesh(n, n, bc=case["bc"])
    co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
    sec, it, rr = orc.bench_jacobi_pcg_solve(m, co, co["b"], 1e-10, 100000, 2)
    _, it_s, _ = orc.sp_pcg(m, 1, co, co["b"], np.zeros((n, n)), tol=1e-10, maxit=100000)
    assert rr <= 1e-10 and abs(it - it_s) <= max(3, it_s // 50), (it, it_s, rr)
