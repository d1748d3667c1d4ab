"""Pins of the torch restatement (wy_autograd.py) that serves as the
reference gradient of the multi-step kinetic-energy adjoint: its forward
against the oracle (which the compiled reference's fixtures pin), and its
autograd against its own central differences and against closed forms of
the geometry's derivatives."""
import numpy as np
import pytest
import torch

import wy_autograd as W


def _field(nx, ny, periodic, seed):
    from covof_amd import workloads as wl
    C = 1.0 - wl.ellipses(nx, ny, 4, 2.0, 6.0, seed=seed)
    (wl.fill_halo_periodic if periodic else wl.fill_halo_zero_gradient)(C, nx, ny)
    u, v = wl.stream_velocity(nx, ny, 4, seed + 1, periodic=periodic)
    dt = 0.35 * (1.0 / nx) / max(np.abs(u).max(), np.abs(v).max())
    return C, u, v, dt


@pytest.mark.parametrize("method", ["PLIC", "ELVIRA"])
@pytest.mark.parametrize("periodic", [False, True], ids=["walls", "periodic"])
def test_wy_steps_match_oracle(orc, method, periodic):
    """Four Weymouth-Yue steps (both pass orders) within 1e-12 of the oracle."""
    nx, ny = 24, 20
    C, u, v, dt = _field(nx, ny, periodic, seed=5)
    m = orc.mesh(nx, ny, bc=(3,) * 4 if periodic else (1,) * 4)
    g = W.Grid(nx, ny, periodic=periodic)
    Co, sc = C.copy(), 0
    Ct = torch.tensor(C)
    with torch.no_grad():
        for s in range(4):
            Co, sc = orc.wy_step(m, {"PLIC": orc.PLIC, "ELVIRA": orc.ELVIRA}[method], Co, u, v, dt, sc)
            Ct = W.wy_step(Ct, torch.tensor(u), torch.tensor(v), g, dt, s, method, two_phase=False)
            assert np.abs(Ct.numpy() - Co).max() < 1e-12, s


def test_projection_matches_oracle(orc):
    """The dense projection against the oracle's assembly, Jacobi-PCG to
    1e-13, mean removal and correction (walls, density ratio 1000)."""
    nx, ny = 24, 20
    C, u, v, dt = _field(nx, ny, False, seed=7)
    m = orc.mesh(nx, ny, bc=(1,) * 4)
    g = W.Grid(nx, ny)
    Cc = C.copy()
    rho, rho_u, rho_v, _, _ = orc.update_properties(m, Cc, 1000.0, 1.0, 1.0, 1.0)
    us, vs = W.predictor(torch.tensor(u), torch.tensor(v), None, None, g, dt, 0.0, -0.98, True)
    us, vs = us.numpy(), vs.numpy()
    co = orc.poisson_assemble(m, us, vs, rho_u, rho_v, dt)
    x, it, rr = orc.sp_pcg(m, 1, co, co["b"], np.zeros((ny, nx)), tol=1e-13, maxit=20000)
    p = orc.sp_finish_pressure(m, x, np.zeros_like(C))
    uo, vo = orc.sp_correct(m, p, rho_u, rho_v, dt, us, vs)
    with torch.no_grad():
        ut, vt = W.projection(torch.tensor(us), torch.tensor(vs), torch.tensor(Cc), g, dt, 1000.0, 1.0)
    scale = np.abs(us).max()
    assert np.abs(ut.numpy() - uo).max() < 1e-9 * scale
    assert np.abs(vt.numpy() - vo).max() < 1e-9 * scale
    # the projected field is divergence-free
    div = (ut[1:ny + 1, 2:] - ut[1:ny + 1, 1:nx + 1]) * g.dy + (vt[2:, 1:nx + 1] - vt[1:ny + 1, 1:nx + 1]) * g.dx
    assert float(div.abs().max()) < 1e-12 * scale


def test_geometry_derivatives_closed_form():
    """d area / d alpha = -L, d area / d n = -L M (chord length L, midpoint
    M), d alpha / d c = -|box| / L, d alpha / d n = -M, on random boxes and
    lines crossing them."""
    rng = np.random.default_rng(3)
    n = 64
    x0 = torch.tensor(rng.uniform(-1, 1, n))
    y0 = torch.tensor(rng.uniform(-1, 1, n))
    x1 = x0 + torch.tensor(rng.uniform(0.1, 0.5, n))
    y1 = y0 + torch.tensor(rng.uniform(0.1, 0.5, n))
    th = rng.uniform(0, 2 * np.pi, n)
    nx = torch.tensor(np.cos(th), requires_grad=True)
    ny = torch.tensor(np.sin(th), requires_grad=True)
    c = torch.tensor(rng.uniform(0.05, 0.95, n), requires_grad=True)
    al = W.alpha_of(x0, x1, y0, y1, nx, ny, c)
    # the cut area is the requested one
    with torch.no_grad():
        A, L = W.cut_area(x0, x1, y0, y1, nx, ny, al)
        assert torch.allclose(A, c * (x1 - x0) * (y1 - y0), rtol=0, atol=1e-14)
    gc, gx, gy = torch.autograd.grad(al.sum(), (c, nx, ny))
    # midpoint of the chord: the two points where the line crosses the box's edges
    M = []
    for k in range(n):
        a, b = float(nx[k].detach()), float(ny[k].detach())
        pts = []
        X0, X1, Y0, Y1, A0 = float(x0[k]), float(x1[k]), float(y0[k]), float(y1[k]), float(al[k].detach())
        for (px, py, qx, qy) in ((X0, Y0, X1, Y0), (X1, Y0, X1, Y1), (X1, Y1, X0, Y1), (X0, Y1, X0, Y0)):
            d0, d1 = a * px + b * py + A0, a * qx + b * qy + A0
            if (d0 <= 0) != (d1 <= 0):
                t = d0 / (d0 - d1)
                pts.append((px + t * (qx - px), py + t * (qy - py)))
        assert len(pts) == 2
        M.append(((pts[0][0] + pts[1][0]) / 2, (pts[0][1] + pts[1][1]) / 2))
    M = np.array(M)
    Lv = L.numpy()
    box = ((x1 - x0) * (y1 - y0)).numpy()
    np.testing.assert_allclose(gc.numpy(), -box / Lv, rtol=1e-9)
    np.testing.assert_allclose(gx.numpy(), -M[:, 0], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(gy.numpy(), -M[:, 1], rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("method", ["PLIC", "ELVIRA"])
def test_coupled_autograd_vs_own_differences(method):
    """Three steps of the coupled model: autograd against central
    differences (h = 1e-6) along directions on the mixed cells of C0."""
    nx, ny, N = 24, 20, 3
    C, u, v, dt = _field(nx, ny, False, seed=3)
    g = W.Grid(nx, ny)
    Ct = torch.tensor(C, requires_grad=True)
    W.coupled_ke(Ct, torch.tensor(u), torch.tensor(v), g, N, dt, 1000.0, 1.0, method=method).backward()
    I = (slice(1, ny + 1), slice(1, nx + 1))
    ok = (C[I] >= 0.02) & (C[I] <= 0.98)
    rng = np.random.default_rng(9)
    gm = Ct.grad.numpy()
    scale = float(np.sqrt((gm[I][ok] ** 2).sum()))
    h = 1e-6
    for _ in range(3):
        d = np.zeros_like(C)
        d[I] = rng.normal(size=ok.shape) * ok
        with torch.no_grad():
            Jp = W.coupled_ke(torch.tensor(C + h * d), torch.tensor(u), torch.tensor(v), g, N, dt, 1000.0, 1.0,
                              method=method).item()
            Jm = W.coupled_ke(torch.tensor(C - h * d), torch.tensor(u), torch.tensor(v), g, N, dt, 1000.0, 1.0,
                              method=method).item()
        fd = (Jp - Jm) / (2 * h)
        ad = float((gm * d).sum())
        assert abs(fd - ad) < 1e-6 * max(abs(ad), scale), (fd, ad)
