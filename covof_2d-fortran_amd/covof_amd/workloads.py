"""Seeded synthetic inputs for the hot-path workloads (numpy, host side).

Fields use the reference's halo layout (``src/module_fields.f90:129-169``):
a numpy array of shape ``(ny+2, nx+2)`` indexed ``F[j, i]`` for Fortran
``F(i, j)``, ``i, j = 0 .. n+1``.  Face velocities follow the MAC placement of
``src/module_fields.f90:33-35``: ``u(i, j)`` at ``(x_face(i), y_center(j))``,
``v(i, j)`` at ``(x_center(i), y_face(j))``.

Workloads
---------
* :func:`poisson_case` -- variable-density pressure Poisson problem: seeded
  random circles (sharp 0/1 indicator, one 3x3 box-smoothing pass), density
  ``rho = 1000 C + (1 - C)``, zero-mean Gaussian right-hand side, walls.
* :func:`vof_case` -- seeded random ellipses (area fractions by chord
  integration over sub-columns, overlaps combined by a clamped sum) in a
  periodic unit square, advected by an exactly (discretely) divergence-free
  velocity built from a nodal stream function of seeded Fourier modes.
* :func:`vortex_velocity` -- the reference's reversible shear-vortex velocity
  (``src/module_simulation.f90:2254-2288``).
"""
from __future__ import annotations

import math

import numpy as np

BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, BC_FREESLIP = 1, 2, 3, 4


def x_face(i, dx):
    """``x_face(i) = real(i-1) * dx`` (``src/module_mesh.f90:162-166``)."""
    return (np.asarray(i, dtype=np.float64) - 1.0) * dx


def x_center(i, dx):
    """``x_center(i) = (real(i-1) + 0.5) * dx`` (``src/module_mesh.f90:182-186``)."""
    return ((np.asarray(i, dtype=np.float64) - 1.0) + 0.5) * dx


# ----------------------------------------------------------------------------- VOF

def _ellipse_fraction(C, nx, ny, dx, dy, xc, yc, a, b, th, sub=32):
    """Add the area fraction of one rotated ellipse to the interior of ``C``."""
    ct_, st_ = math.cos(th), math.sin(th)
    ext_x = math.sqrt((a * ct_) ** 2 + (b * st_) ** 2)
    ext_y = math.sqrt((a * st_) ** 2 + (b * ct_) ** 2)
    i0 = max(1, int(math.floor((xc - ext_x) / dx)) + 1 - 1)
    i1 = min(nx, int(math.floor((xc + ext_x) / dx)) + 1 + 1)
    j0 = max(1, int(math.floor((yc - ext_y) / dy)) + 1 - 1)
    j1 = min(ny, int(math.floor((yc + ext_y) / dy)) + 1 + 1)
    if i0 > i1 or j0 > j1:
        return
    ii = np.arange(i0, i1 + 1)
    jj = np.arange(j0, j1 + 1)
    # quadratic in y' for every sub-column x': A y'^2 + B y' + Cq <= 0
    A = (st_ / a) ** 2 + (ct_ / b) ** 2
    xs = (ii[:, None] - 1.0 + (np.arange(sub)[None, :] + 0.5) / sub) * dx - xc  # (ni, sub)
    B = 2.0 * xs * (ct_ * st_ / a ** 2 - st_ * ct_ / b ** 2)
    Cq = xs ** 2 * ((ct_ / a) ** 2 + (st_ / b) ** 2) - 1.0
    disc = B * B - 4.0 * A * Cq
    ok = disc > 0.0
    sq = np.sqrt(np.where(ok, disc, 0.0))
    ylo = np.where(ok, (-B - sq) / (2.0 * A), 0.0) + yc
    yhi = np.where(ok, (-B + sq) / (2.0 * A), 0.0) + yc
    y0 = (jj - 1.0) * dy
    y1 = jj * dy
    lo = np.clip(ylo[None, :, :], y0[:, None, None], y1[:, None, None])
    hi = np.clip(yhi[None, :, :], y0[:, None, None], y1[:, None, None])
    frac = np.where(ok[None, :, :], hi - lo, 0.0).mean(axis=2) / dy  # (nj, ni)
    C[j0:j1 + 1, i0:i1 + 1] += frac


def _disc_polygon_area(px, py):
    """Area of the unit disc inside the polygon with vertices (px[k], py[k])
    (arrays of shape (K, ...), counter-clockwise): the sum over the edges P->Q
    of the signed area of the disc inside the triangle (0, P, Q) -- where the
    segment runs inside the disc, the triangle's area 0.5 P x Q; outside it,
    the circular sector 0.5 angle(P, Q)."""
    def ang(ux, uy, vx, vy):
        return np.arctan2(ux * vy - uy * vx, ux * vx + uy * vy)

    total = np.zeros(px.shape[1:])
    K = px.shape[0]
    for k in range(K):
        Px, Py, Qx, Qy = px[k], py[k], px[(k + 1) % K], py[(k + 1) % K]
        dx, dy = Qx - Px, Qy - Py
        A = dx * dx + dy * dy
        B = 2.0 * (Px * dx + Py * dy)
        Cq = Px * Px + Py * Py - 1.0
        disc = B * B - 4.0 * A * Cq
        cross = disc > 0.0
        sq = np.sqrt(np.where(cross, disc, 0.0))
        A2 = np.where(A > 0.0, 2.0 * A, 1.0)
        t1 = np.where(cross, np.clip((-B - sq) / A2, 0.0, 1.0), 1.0)
        t2 = np.where(cross, np.clip((-B + sq) / A2, 0.0, 1.0), 1.0)
        m1x, m1y = Px + t1 * dx, Py + t1 * dy
        m2x, m2y = Px + t2 * dx, Py + t2 * dy
        total += 0.5 * (ang(Px, Py, m1x, m1y) + (m1x * m2y - m1y * m2x) + ang(m2x, m2y, Qx, Qy))
    return total


def ellipse_cell_fractions(x0, x1, y0, y1, xc, yc, a, b, th):
    """Exact area fraction of each cell [x0, x1] x [y0, y1] inside the ellipse
    of semi-axes a, b rotated by th about (xc, yc): the map
    q = diag(1/a, 1/b) R(th)^T (p - c) takes the ellipse to the unit disc and
    the cell to a parallelogram, whose area inside the disc, times a b, is the
    cell's covered area."""
    ct, st = math.cos(th), math.sin(th)
    xs = np.stack([x0, x1, x1, x0]) - xc
    ys = np.stack([y0, y0, y1, y1]) - yc
    qx = (ct * xs + st * ys) / a
    qy = (-st * xs + ct * ys) / b
    area = _disc_polygon_area(qx, qy) * (a * b)
    f = area / ((x1 - x0) * (y1 - y0))
    return np.where(f < 1e-12, 0.0, np.where(f > 1.0 - 1e-12, 1.0, f))  # round-off of the closed form


def _ellipse_fraction_exact(C, nx, ny, dx, dy, xc, yc, a, b, th):
    """Add the exact area fraction of one rotated ellipse to the interior of ``C``."""
    ct_, st_ = math.cos(th), math.sin(th)
    ext_x = math.sqrt((a * ct_) ** 2 + (b * st_) ** 2)
    ext_y = math.sqrt((a * st_) ** 2 + (b * ct_) ** 2)
    i0 = max(1, int(math.floor((xc - ext_x) / dx)))
    i1 = min(nx, int(math.floor((xc + ext_x) / dx)) + 2)
    j0 = max(1, int(math.floor((yc - ext_y) / dy)))
    j1 = min(ny, int(math.floor((yc + ext_y) / dy)) + 2)
    if i0 > i1 or j0 > j1:
        return
    X0, Y0 = np.meshgrid((np.arange(i0, i1 + 1) - 1.0) * dx, (np.arange(j0, j1 + 1) - 1.0) * dy)
    C[j0:j1 + 1, i0:i1 + 1] += ellipse_cell_fractions(X0, X0 + dx, Y0, Y0 + dy, xc, yc, a, b, th)


def ellipses(nx, ny, n, semi_min_cells, semi_max_cells, seed, lx=1.0, ly=1.0, exact=False):
    """Clamped union of ``n`` seeded random ellipses; interior only (halo zero).
    Each ellipse's cell fractions by 32-point sub-column quadrature, or its
    exact cell areas (``exact``)."""
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    C = np.zeros((ny + 2, nx + 2))
    add = _ellipse_fraction_exact if exact else _ellipse_fraction
    for _ in range(n):
        a = rng.uniform(semi_min_cells, semi_max_cells) * dx
        b = rng.uniform(semi_min_cells, semi_max_cells) * dx
        th = rng.uniform(0.0, math.pi)
        xc = rng.uniform(0.0, lx)
        yc = rng.uniform(0.0, ly)
        add(C, nx, ny, dx, dy, xc, yc, a, b, th)
    np.clip(C, 0.0, 1.0, out=C)
    return C


def fill_halo_periodic(F, nx, ny):
    """Single-rank halo exchange with both directions periodic
    (``src/module_mpi_comm.f90:46-56``)."""
    F[:, 0] = F[:, nx]
    F[:, nx + 1] = F[:, 1]
    F[0, :] = F[ny, :]
    F[ny + 1, :] = F[1, :]
    return F


def fill_halo_zero_gradient(F, nx, ny):
    """Copy the first/last interior row and column outwards (walls)."""
    F[1:ny + 1, 0] = F[1:ny + 1, 1]
    F[1:ny + 1, nx + 1] = F[1:ny + 1, nx]
    F[0, :] = F[1, :]
    F[ny + 1, :] = F[ny, :]
    return F


def stream_velocity(nx, ny, nmodes, seed, lx=1.0, ly=1.0, kmax=4, periodic=True, amplitude=1.0):
    """Exactly divergence-free MAC velocity from a nodal stream function.

    ``psi`` is sampled at nodes ``(x_face(i), y_face(j))``; ``u = dpsi/dy`` and
    ``v = -dpsi/dx`` by one-sided node differences, so the discrete divergence
    ``(u(i+1,j)-u(i,j)) dy + (v(i,j+1)-v(i,j)) dx`` telescopes to rounding.
    Periodic: Fourier modes with integer wave vectors ``0 < |k| <= kmax``, the
    node row/column ``n+1`` is the wrap copy of ``1``.  Walls: ``psi`` is the
    same sum multiplied by ``sin^2(pi x) sin^2(pi y)``, so the wall-normal
    velocity vanishes.
    """
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    ks = []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=2)
        if 0 < k[0] ** 2 + k[1] ** 2 <= kmax ** 2:
            ks.append(k)
    amp = rng.normal(size=nmodes)
    amp = amplitude * amp / np.abs(amp).sum()
    phs = rng.uniform(0.0, 2.0 * math.pi, size=nmodes)
    xn = x_face(np.arange(1, nx + 2), dx)  # nodes i = 1..nx+1
    yn = x_face(np.arange(1, ny + 2), dy)
    X, Y = np.meshgrid(xn, yn)  # (ny+1, nx+1)
    psi = np.zeros_like(X)
    for (kx, ky), a, ph in zip(ks, amp, phs):
        psi += a * np.sin(2.0 * math.pi * (kx * X / lx + ky * Y / ly) + ph) / (2.0 * math.pi)
    if periodic:
        psi[:, nx] = psi[:, 0]
        psi[ny, :] = psi[0, :]
    else:
        psi *= (np.sin(math.pi * X / lx) ** 2) * (np.sin(math.pi * Y / ly) ** 2)
        psi[:, 0] = psi[:, nx] = 0.0
        psi[0, :] = psi[ny, :] = 0.0
    u = np.zeros((ny + 2, nx + 2))
    v = np.zeros((ny + 2, nx + 2))
    # u(i,j), i = 1..nx+1, j = 1..ny ; psi index [j-1, i-1]
    u[1:ny + 1, 1:nx + 2] = (psi[1:ny + 1, 0:nx + 1] - psi[0:ny, 0:nx + 1]) / dy
    # v(i,j), i = 1..nx, j = 1..ny+1
    v[1:ny + 2, 1:nx + 1] = -(psi[0:ny + 1, 1:nx + 1] - psi[0:ny + 1, 0:nx]) / dx
    if periodic:
        u[1:ny + 1, 0] = u[1:ny + 1, nx]
        v[0, 1:nx + 1] = v[ny, 1:nx + 1]
    return u, v


def vof_case(n, seed=0, n_ellipses=128, semi=(20.0, 200.0), nmodes=8, cfl=0.25, exact=True):
    """Periodic ``n x n`` VOF advection case (the ellipses' cell fractions
    exact); returns ``(C, u, v, dt)``."""
    C = ellipses(n, n, n_ellipses, semi[0], semi[1], seed, exact=exact)
    fill_halo_periodic(C, n, n)
    u, v = stream_velocity(n, n, nmodes, seed + 1, periodic=True)
    vmax = max(np.abs(u).max(), np.abs(v).max())
    dt = cfl * (1.0 / n) / vmax
    return C, u, v, dt


def vortex_velocity(nx, ny, t, dt, T=6.0, lx=1.0, ly=1.0, walls=True):
    """Reversible shear vortex of ``src/module_simulation.f90:2254-2288``."""
    dx, dy = lx / nx, ly / ny
    cf = math.cos(math.pi * (t + 0.5 * dt) / T)
    u = np.zeros((ny + 2, nx + 2))
    v = np.zeros((ny + 2, nx + 2))
    u_nx = nx + 1 if walls else nx
    v_ny = ny + 1 if walls else ny
    iu = np.arange(1, u_nx + 1)
    jc = np.arange(1, ny + 1)
    sx = np.sin(math.pi * x_face(iu, dx))
    sy = np.sin(2.0 * math.pi * x_center(jc, dy))
    u[1:ny + 1, 1:u_nx + 1] = -(sx * sx)[None, :] * sy[:, None] * cf
    ic = np.arange(1, nx + 1)
    jv = np.arange(1, v_ny + 1)
    sx2 = np.sin(2.0 * math.pi * x_center(ic, dx))
    sy2 = np.sin(math.pi * x_face(jv, dy))
    v[1:v_ny + 1, 1:nx + 1] = sx2[None, :] * (sy2 * sy2)[:, None] * cf
    return u, v


def disc(nx, ny, xc, yc, r, lx=1.0, ly=1.0, sub=16):
    """Area fraction of a disc by ``sub x sub`` point sampling (as the
    reference's bubble initial condition, ``src/module_interface_vof.f90:128-163``)."""
    dx, dy = lx / nx, ly / ny
    C = np.zeros((ny + 2, nx + 2))
    s = (np.arange(sub) + 0.5) / sub
    xs = ((np.arange(nx)[:, None] + s[None, :]) * dx).reshape(-1)
    ys = ((np.arange(ny)[:, None] + s[None, :]) * dy).reshape(-1)
    inside = ((xs[None, :] - xc) ** 2 + (ys[:, None] - yc) ** 2) <= r * r
    C[1:ny + 1, 1:nx + 1] = inside.reshape(ny, sub, nx, sub).mean(axis=(1, 3))
    return C


# ------------------------------------------------------------------------- Poisson

def circles_indicator(n, n_circles, rmin_cells, rmax_cells, seed, lx=1.0):
    """Sharp 0/1 union of seeded circles, then one 3x3 box-average pass
    (edges replicated); interior ``(n, n)``."""
    rng = np.random.default_rng(seed)
    dx = lx / n
    xc = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc)
    ind = np.zeros((n, n))
    for _ in range(n_circles):
        r = rng.uniform(rmin_cells, rmax_cells) * dx
        cx, cy = rng.uniform(0.0, lx, size=2)
        # work on the bounding box only
        i0 = max(0, int((cx - r) / dx) - 1)
        i1 = min(n, int((cx + r) / dx) + 2)
        j0 = max(0, int((cy - r) / dx) - 1)
        j1 = min(n, int((cy + r) / dx) + 2)
        sub = (X[j0:j1, i0:i1] - cx) ** 2 + (Y[j0:j1, i0:i1] - cy) ** 2 <= r * r
        ind[j0:j1, i0:i1] = np.maximum(ind[j0:j1, i0:i1], sub)
    P = np.pad(ind, 1, mode="edge")
    sm = sum(P[1 + dj:n + 1 + dj, 1 + di:n + 1 + di] for dj in (-1, 0, 1) for di in (-1, 0, 1)) / 9.0
    return sm


def poisson_case(n, seed=0, n_circles=64, rmin=16.0, rmax=128.0, rho_liquid=1000.0,
                 rho_gas=1.0):
    """Variable-density Poisson inputs in the reference's field layout.

    Returns ``dict(C, u, v, rho_u, rho_v, p, rhs, dt, bc)``: ``C`` with a
    zero-gradient halo; ``rho_u``/``rho_v`` are the arithmetic face means of
    ``rho = rho_liquid C + rho_gas (1 - C)`` (``src/module_interface_vof.f90:538-591``);
    ``u`` is zero and ``v`` carries the right-hand side: with ``dx = dy`` the
    Poisson right-hand side ``b = -[(u_E-u_W) dy + (v_N-v_S) dx]`` of
    ``src/module_simulation.f90:2483-2486`` equals ``rhs`` (zero-mean
    N(0,1), seed ``seed``) up to the rounding of one telescoping sum, and the
    exact ``b`` is returned as ``rhs``.
    """
    ind = circles_indicator(n, n_circles, rmin, rmax, seed)
    C = np.zeros((n + 2, n + 2))
    C[1:n + 1, 1:n + 1] = ind
    fill_halo_zero_gradient(C, n, n)
    rho = C * rho_liquid + (1.0 - C) * rho_gas
    rho_u = np.ones((n + 2, n + 2))
    rho_v = np.ones((n + 2, n + 2))
    rho_u[1:n + 1, 1:n + 2] = 0.5 * (rho[1:n + 1, 0:n + 1] + rho[1:n + 1, 1:n + 2])
    rho_v[1:n + 2, 1:n + 1] = 0.5 * (rho[0:n + 1, 1:n + 1] + rho[1:n + 2, 1:n + 1])
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(n, n))
    g -= g.mean()
    dx = 1.0 / n
    # v(i, j+1) - v(i, j) = -g(i, j) / dx  with v(i, 1) = 0  (column cumsum)
    v = np.zeros((n + 2, n + 2))
    v[2:n + 2, 1:n + 1] = np.cumsum(-g / dx, axis=0)
    u = np.zeros((n + 2, n + 2))
    rhs = -((v[2:n + 2, 1:n + 1] - v[1:n + 1, 1:n + 1]) * dx)
    return dict(C=C, u=u, v=v, rho_u=rho_u, rho_v=rho_v, p=np.zeros((n + 2, n + 2)),
                rhs=rhs, dt=1.0, bc=(BC_FREESLIP,) * 4)


def momentum_case(nx, ny, seed=0, n_circles=16, rmin=4.0, rmax=24.0, rho_liquid=1000.0, rho_gas=100.0,
                  mu_liquid=10.0, mu_gas=1.0, amplitude=0.1, lx=1.0, ly=None, bc=(BC_DIRICHLET,) * 4):
    """Inputs of the SP momentum block in the reference's field layout.

    Two-phase properties from seeded circles (arithmetic face densities,
    cell viscosities and arithmetic-mean node viscosities as the shapes of
    ``update_physical_properties``, src/module_interface_vof.f90:479-616),
    a smooth velocity field (a sum of Fourier modes, amplitude ``amplitude``)
    as the current state, ``u_old``/``v_old`` and ``u_nm1``/``v_nm1`` as
    nearby states, and small smooth forces standing in for the CSF terms.
    Returns ``dict`` of halo arrays named as the C-ABI's momentum fields, and
    ``dt`` (CFL 0.25 on the velocity).
    """
    ly = lx * ny / nx if ly is None else ly
    n = max(nx, ny)
    ind = circles_indicator(n, n_circles, rmin, rmax, seed)[:ny, :nx]
    C = np.zeros((ny + 2, nx + 2))
    C[1:ny + 1, 1:nx + 1] = ind
    fill_halo_zero_gradient(C, nx, ny)
    rho = C * rho_liquid + (1.0 - C) * rho_gas
    mu = C * mu_liquid + (1.0 - C) * mu_gas
    f = {}
    f["rho_u"] = rho.copy()
    f["rho_u"][:, 1:] = 0.5 * (rho[:, :-1] + rho[:, 1:])
    f["rho_v"] = rho.copy()
    f["rho_v"][1:, :] = 0.5 * (rho[:-1, :] + rho[1:, :])
    f["mu"] = mu
    f["mu_node"] = mu.copy()
    f["mu_node"][1:, 1:] = 0.25 * (mu[:-1, :-1] + mu[1:, :-1] + mu[:-1, 1:] + mu[1:, 1:])
    rng = np.random.default_rng(seed + 7)
    x = (np.arange(nx + 2) - 0.5) / nx * lx
    y = (np.arange(ny + 2) - 0.5) / ny * ly
    X, Y = np.meshgrid(x, y)

    def smooth(phase):
        out = np.zeros((ny + 2, nx + 2))
        for _ in range(4):
            kx, ky = rng.integers(1, 4, size=2)
            out += np.sin(2 * np.pi * (kx * X / lx + ky * Y / ly) + phase + rng.uniform(0, 2 * np.pi))
        return out / 4.0

    for k in ("u", "v"):
        base = amplitude * smooth(0.0)
        f[k] = base
        f[k + "_old"] = base + 0.01 * amplitude * smooth(1.0)
        f[k + "_nm1"] = base + 0.02 * amplitude * smooth(2.0)
    f["fx_sigma"] = 10.0 * smooth(3.0)
    f["fy_sigma"] = 10.0 * smooth(4.0)
    umax = max(np.abs(f["u"]).max(), np.abs(f["v"]).max(), 1e-30)
    dt = 0.25 * min(lx / nx, ly / ny) / umax
    return f, dt, bc, lx, ly


def _disc_quadrant(x, y, R):
    """Area of {(t, u): t^2 + u^2 <= R^2, t <= x, u <= y} (disc centred at the
    origin), closed form from the primitive of sqrt(R^2 - t^2); arrays."""
    def S(t):
        t = np.clip(t, -R, R)
        return 0.5 * (t * np.sqrt(np.maximum(R * R - t * t, 0.0)) + R * R * np.arcsin(t / R))
    X = np.clip(x, -R, R)
    T = np.sqrt(np.maximum(R * R - y * y, 0.0))
    mid = y * (np.clip(X, -T, T) + T) + S(np.clip(X, -T, T)) - S(-T)  # |t| < T: chord part below y
    sides = 2.0 * (S(np.minimum(X, -T)) - S(-R)) + 2.0 * (S(np.maximum(X, T)) - S(T))  # |t| >= T, y >= 0
    return np.where(y >= R, 2.0 * (S(X) - S(-R)), np.where(y <= -R, 0.0, np.where(y >= 0.0, sides + mid, mid)))


def disc_cell_fractions(x0, x1, y0, y1, cx, cy, R):
    """Exact area fraction of each cell [x0, x1] x [y0, y1] inside the disc of
    radius R centred at (cx, cy), by inclusion-exclusion of quadrant areas."""
    a = (_disc_quadrant(x1 - cx, y1 - cy, R) - _disc_quadrant(x0 - cx, y1 - cy, R)
         - _disc_quadrant(x1 - cx, y0 - cy, R) + _disc_quadrant(x0 - cx, y0 - cy, R))
    return np.clip(a / ((x1 - x0) * (y1 - y0)), 0.0, 1.0)


def bubbles_case(nx, ny, n_bubbles=256, rmin=8.0, rmax=32.0, seed=0, lx=2.0, ly=1.0, sub=4, exact=False):
    """Liquid (C = 1) holding ``n_bubbles`` seeded non-overlapping circular
    bubbles (C = 0), radii uniform in [rmin, rmax] cells, at least two cells
    apart and from the walls; volume fractions by sub x sub sampling per cell,
    or exact disc-cell areas (``exact``); zero-gradient halo.  Returns the halo
    array ``C`` of shape (ny+2, nx+2)."""
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    h = min(dx, dy)
    placed = []
    tries = 0
    while len(placed) < n_bubbles and tries < 200000:
        tries += 1
        r = rng.uniform(rmin, rmax) * h
        cx = rng.uniform(r + 2 * h, lx - r - 2 * h)
        cy = rng.uniform(r + 2 * h, ly - r - 2 * h)
        if all((cx - x) ** 2 + (cy - y) ** 2 > (r + q + 2 * h) ** 2 for x, y, q in placed):
            placed.append((cx, cy, r))
    C = np.ones((ny + 2, nx + 2))
    off = (np.arange(sub) + 0.5) / sub
    for cx, cy, r in placed:
        i0, i1 = max(1, int((cx - r) / dx)), min(nx, int((cx + r) / dx) + 2)
        j0, j1 = max(1, int((cy - r) / dy)), min(ny, int((cy + r) / dy) + 2)
        if exact:
            xi = (np.arange(i0, i1 + 1) - 1) * dx
            yj = (np.arange(j0, j1 + 1) - 1) * dy
            X0, Y0 = np.meshgrid(xi, yj)
            gas = disc_cell_fractions(X0, X0 + dx, Y0, Y0 + dy, cx, cy, r)
            gas = np.where(gas < 1e-12, 0.0, np.where(gas > 1.0 - 1e-12, 1.0, gas))  # round-off of the closed form
        else:
            xs = (np.arange(i0, i1 + 1)[:, None] - 1 + off[None, :]).reshape(-1) * dx
            ys = (np.arange(j0, j1 + 1)[:, None] - 1 + off[None, :]).reshape(-1) * dy
            inside = ((xs[None, :] - cx) ** 2 + (ys[:, None] - cy) ** 2 <= r * r).astype(np.float64)
            gas = inside.reshape(j1 - j0 + 1, sub, i1 - i0 + 1, sub).mean(axis=(1, 3))
        C[j0:j1 + 1, i0:i1 + 1] -= gas
    fill_halo_zero_gradient(C, nx, ny)
    return C, len(placed)
