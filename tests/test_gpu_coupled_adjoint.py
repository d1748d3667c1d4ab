"""The checkpointed multi-step adjoint (KineticEnergyAdjoint) and its
Weymouth-Yue step adjoint (covof_vof_advect_adjoint_d), against torch
autograd of the dense float64 restatement in wy_autograd.py (computed on the
CPU) and against central finite differences of the GPU forward.

The transport map is piecewise smooth (donors, cell classes, ELVIRA's
choice, the clamp, c_d): both references take their branches from their own
forward values, and the finite differences move only cells that are mixed
with 0.02 <= C0 <= 0.98, by h = 1e-6, so that no branch changes.  Where
rounding alone decides a branch, the adjoint's rule fixes the derivative
(both models follow it): a clamp that removes at most 1e-12 passes the
derivative, and an ELVIRA winner that ties to rounding with a lower-index
candidate of the same normal (Youngs' normal and the central height-function
one coincide wherever the outer rows of the 3 x 3 block are uniform) is
differentiated along that candidate.  On a smooth field ELVIRA's candidates
nearly tie everywhere with different normals, so the finite differences run
on the bubble fields only.
"""
import numpy as np
import pytest
import torch

import wy_autograd as W
from conftest import to_gpu

pytestmark = pytest.mark.gpu

RHO1, RHO2 = 1000.0, 1.0   # liquid (C = 1), gas
CASES = [("PLIC", "bubbles"), ("ELVIRA", "bubbles"), ("ELVIRA", "smooth")]


def _smooth(nx, ny, seed):
    """0.5 + 0.42 f with f a normalised sum of six Fourier modes (periodic)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    f = np.zeros_like(X)
    for _ in range(6):
        kx, ky = rng.integers(1, 4, size=2)
        f += rng.normal() * np.sin(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
    C = np.zeros((ny + 2, nx + 2))
    C[1:ny + 1, 1:nx + 1] = 0.5 + 0.42 * f / np.abs(f).max()
    return C


def _case(nx, ny, periodic, seed, cfl=0.35, field="bubbles"):
    from covof_amd import workloads as wl
    if field == "bubbles":
        C = 1.0 - wl.ellipses(nx, ny, 4, 2.0, 6.0, seed=seed)
    else:
        C = _smooth(nx, ny, seed)
    (wl.fill_halo_periodic if periodic else wl.fill_halo_zero_gradient)(C, nx, ny)
    u, v = wl.stream_velocity(nx, ny, 4, seed + 1, periodic=periodic)
    dt = cfl * (1.0 / nx) / max(np.abs(u).max(), np.abs(v).max())
    return C, u, v, dt


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("method,field", CASES)
@pytest.mark.parametrize("periodic", [False, True], ids=["walls", "periodic"])
@pytest.mark.parametrize("step_count", [0, 1], ids=["x_first", "y_first"])
def test_wy_step_adjoint_vs_autograd(cv, method, field, periodic, step_count):
    """dJ/dC^n (halo entries included), dJ/du, dJ/dv of J = sum(w C^{n+1})
    for one step, against autograd; the forward values against the GPU's."""
    nx, ny = 24, 20
    C, u, v, dt = _case(nx, ny, periodic, seed=11 + step_count, field=field)
    bc = (3,) * 4 if periodic else (1,) * 4
    ctx = cv.Context(nx, ny, bc=bc)
    w = np.random.default_rng(5).normal(size=C.shape)
    for two_phase in ((True,) if periodic else (True, False)):
        g = W.Grid(nx, ny, periodic=periodic)
        Ct = torch.tensor(C, requires_grad=True)
        ut = torch.tensor(u, requires_grad=True)
        vt = torch.tensor(v, requires_grad=True)
        out = W.wy_step(Ct, ut, vt, g, dt, step_count, method, two_phase=two_phase)
        (torch.tensor(w) * out).sum().backward()
        Cg = to_gpu(C)
        ctx.advect_vof(Cg, to_gpu(u), to_gpu(v), dt, reconstruction_method=method, step_count=step_count,
                       two_phase=two_phase, rho1=RHO1, rho2=RHO2)
        assert _rel(Cg.cpu().numpy(), out.detach().numpy()) < 1e-11
        cb, ub, vb = ctx.vof_advect_adjoint(to_gpu(C), to_gpu(u), to_gpu(v), dt, to_gpu(w), step_count, method,
                                            two_phase=two_phase)
        e = (_rel(cb.cpu().numpy(), Ct.grad.numpy()), _rel(ub.cpu().numpy(), ut.grad.numpy()),
             _rel(vb.cpu().numpy(), vt.grad.numpy()))
        print(f"[{method} {field} {'per' if periodic else 'walls'} sc={step_count} two_phase={two_phase}] "
              f"dC {e[0]:.2e} du {e[1]:.2e} dv {e[2]:.2e}", flush=True)
        assert max(e) < 1e-9, e
        # folding the input halo = the gradient of the step of the halo-copied field
        cbf, _, _ = ctx.vof_advect_adjoint(to_gpu(C), to_gpu(u), to_gpu(v), dt, to_gpu(w), step_count, method,
                                           two_phase=two_phase, fold_input_halo=True)
        Ci = torch.tensor(C[1:ny + 1, 1:nx + 1], requires_grad=True)
        full = g.copy_halo(torch.nn.functional.pad(Ci, (1, 1, 1, 1)))
        outf = W.wy_step(full, torch.tensor(u), torch.tensor(v), g, dt, step_count, method, two_phase=two_phase)
        (torch.tensor(w) * outf).sum().backward()
        cbf = cbf.cpu().numpy()
        assert _rel(cbf[1:ny + 1, 1:nx + 1], Ci.grad.numpy()) < 1e-9
        assert not cbf[g.ring.numpy()].any()


def _coupled(cv, method, field, nx=24, ny=20):
    C, u, v, dt = _case(nx, ny, False, seed=3, cfl=0.3, field=field)
    ctx = cv.Context(nx, ny, bc=(1,) * 4)
    run = cv.KineticEnergyAdjoint(ctx, RHO1, RHO2, dt, gy=-0.98, reconstruction=method, tol=1e-14)
    return ctx, run, C, u, v, dt


def _chained(run, C, g, dt, method, N):
    """torch autograd of each step at the GPU run's stored states, chained
    backwards as the checkpointed adjoint chains them (the same branches)."""
    T = lambda a: a.detach().cpu().clone().requires_grad_()
    CN, uN, vN = T(run.C[N]), T(run.u[N]), T(run.v[N])
    W.kinetic_energy(CN, uN, vN, g, RHO1, RHO2).backward()
    Cb, Ub, Vb = CN.grad, uN.grad, vN.grad
    Au, Av = torch.zeros_like(Ub), torch.zeros_like(Vb)
    for n in range(N - 1, -1, -1):
        Cn, un, vn = T(run.C[n]), T(run.u[n]), T(run.v[n])
        first = n == 0
        u1 = None if first else T(run.u[n - 1])
        v1 = None if first else T(run.v[n - 1])
        Cn1 = W.wy_step(Cn, un, vn, g, dt, n, method, two_phase=True)
        us, vs = W.predictor(un, vn, u1, v1, g, dt, 0.0, -0.98, first)
        u2, v2 = W.projection(us, vs, Cn1, g, dt, RHO1, RHO2)
        ((Cb * Cn1).sum() + (Ub * u2).sum() + (Vb * v2).sum()).backward()
        Cb, Ub, Vb = Cn.grad, un.grad + Au, vn.grad + Av
        if not first:
            Au, Av = u1.grad, v1.grad
    Ci = torch.tensor(C, requires_grad=True)
    (g.copy_halo(Ci) * Cb).sum().backward()
    return Ci.grad, Ub, Vb


@pytest.mark.parametrize("method,field", CASES)
def test_coupled_adjoint_vs_autograd(cv, method, field):
    """dJ/dC0, dJ/du0, dJ/dv0 of the kinetic energy after 16 steps on 24 x 20
    walls against torch autograd of the dense model: of the whole 16-step
    model from C0 (its own trajectory and branches), and of each step at the
    GPU run's stored states chained backwards; the trajectories agree step
    by step."""
    nx, ny, N = 24, 20, 16
    ctx, run, C, u, v, dt = _coupled(cv, method, field, nx, ny)
    C0, u0, v0 = to_gpu(C).requires_grad_(), to_gpu(u).requires_grad_(), to_gpu(v).requires_grad_()
    J = cv.kinetic_energy_after(run, C0, u0, v0, N)
    J.backward()
    g = W.Grid(nx, ny)
    Ct = torch.tensor(C, requires_grad=True)
    ut = torch.tensor(u, requires_grad=True)
    vt = torch.tensor(v, requires_grad=True)
    Jt = W.coupled_ke(Ct, ut, vt, g, N, dt, RHO1, RHO2, gy=-0.98, method=method)
    Jt.backward()
    eJ = abs(J.item() - Jt.item()) / abs(Jt.item())
    e = (_rel(C0.grad.cpu().numpy(), Ct.grad.numpy()), _rel(u0.grad.cpu().numpy(), ut.grad.numpy()),
         _rel(v0.grad.cpu().numpy(), vt.grad.numpy()))
    print(f"[{method} {field} N={N}] J {J.item():.10e} (rel {eJ:.2e}); iters {run.iters}; dC0 {e[0]:.2e} "
          f"du0 {e[1]:.2e} dv0 {e[2]:.2e}", flush=True)
    assert eJ < 1e-10
    assert max(e) < 1e-9, e
    # the trajectories, step by step
    Cc, uu, vv, u1, v1 = g.copy_halo(torch.tensor(C)), torch.tensor(u), torch.tensor(v), None, None
    with torch.no_grad():
        for n in range(N):
            Cc = W.wy_step(Cc, uu, vv, g, dt, n, method, two_phase=True)
            us, vs = W.predictor(uu, vv, u1, v1, g, dt, 0.0, -0.98, n == 0)
            u1, v1 = uu, vv
            uu, vv = W.projection(us, vs, Cc, g, dt, RHO1, RHO2)
            assert _rel(run.C[n + 1].cpu().numpy(), Cc.numpy()) < 1e-11
            assert _rel(run.u[n + 1].cpu().numpy(), uu.numpy()) < 1e-9
    ch = _chained(run, C, g, dt, method, N)
    ec = tuple(_rel(a.cpu().numpy(), b.numpy()) for a, b in zip((C0.grad, u0.grad, v0.grad), ch))
    print(f"[{method} {field}] chained at the stored states: dC0 {ec[0]:.2e} du0 {ec[1]:.2e} dv0 {ec[2]:.2e}",
          flush=True)
    assert max(ec) < 1e-9, ec


@pytest.mark.parametrize("method", ["PLIC", "ELVIRA"])
def test_coupled_adjoint_fd(cv, method):
    """Central differences of the GPU forward (h = 1e-6) along 16 seeded
    directions on the mixed cells of C0 against <dJ/dC0, d>.  The error is
    relative to the larger of |<dJ/dC0, d>| and the expected size of a random
    direction's derivative, |dJ/dC0| over the perturbed cells (d has unit
    normal entries there): the forward's rounding noise (~1e-13 in J after 16
    steps) limits a difference quotient to ~1e-7 absolute, which a direction
    whose derivative nearly cancels would otherwise turn into a large
    relative error."""
    nx, ny, N = 24, 20, 16
    ctx, run, C, u, v, dt = _coupled(cv, method, "bubbles", nx, ny)
    ug, vg = to_gpu(u), to_gpu(v)
    run.forward(to_gpu(C), ug, vg, N)
    dC, _, _ = run.backward()
    dC = dC.cpu().numpy()
    I = (slice(1, ny + 1), slice(1, nx + 1))
    ok = (C[I] >= 0.02) & (C[I] <= 0.98)
    assert ok.sum() >= 8
    rng = np.random.default_rng(2024)
    h = 1e-6
    scale = float(np.sqrt((dC[I][ok] ** 2).sum()))
    worst = 0.0
    for k in range(16):
        d = np.zeros_like(C)
        d[I] = rng.normal(size=ok.shape) * ok
        Jp = run.forward(to_gpu(C + h * d), ug, vg, N)
        Jm = run.forward(to_gpu(C - h * d), ug, vg, N)
        fd = (Jp - Jm) / (2.0 * h)
        ad = float((dC * d).sum())
        err = abs(fd - ad) / max(abs(ad), scale)
        worst = max(worst, err)
        print(f"[{method} dir {k}] fd {fd:.12e} ad {ad:.12e} rel {err:.2e} (to |ad| {abs(fd - ad) / abs(ad):.2e}, "
              f"scale {scale:.3e})", flush=True)
    assert worst < 1e-6
