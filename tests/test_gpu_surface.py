"""The steps advect_vof runs after the transport -- SVOF Helmholtz smoothing,
diffusion smoothing, curvature and CSF forces -- on the GPU, bit-for-bit
against the golden vectors of the compiled reference and against the oracle."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, to_gpu

pytestmark = pytest.mark.gpu

METHOD = {0: "PLIC", 1: "ELVIRA"}


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "surface_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_surface_steps_golden(cv, path):
    """advect_vof (WEYMOUTH_YUE) followed by the smoothings and the
    curvature / CSF call, step by step, against the reference's state."""
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    ctx = cv.Context(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    C = to_gpu(g["C0"])
    u, v = to_gpu(g["u"]), to_gpu(g["v"])
    Cs, Ct = ctx.new_field(1.0), ctx.new_field(1.0)  # the reference's initial values
    out = {k: ctx.new_field(0.0) for k in ("curvature", "fx_sigma", "fy_sigma")}
    svof, diff = int(g["svof"]), int(g["diff"])
    for s in range(int(g["nsteps"])):
        ctx.advect_vof(C, u, v, float(g["dt"]), reconstruction_method=METHOD[int(g["method"])])
        if svof:
            ctx.smooth_svof_helmholtz(C, float(g["width"]), C_s=Cs)
        if diff:
            ctx.smooth_diffusion(C, int(g["dsteps"]), C_tilde=Ct)
        G = Ct if diff else (Cs if svof else None)
        ctx.compute_curvature_and_csf(C, float(g["sigma"]), G=G, enable=bool(int(g["tension"])), out=out)
        got = {"C": C, "C_s": Cs, "C_tilde": Ct, **out}
        for name, t in got.items():
            ref = g[name][s]
            val = t.cpu().numpy()
            assert np.array_equal(val, ref), f"step {s + 1} {name}: max diff {np.abs(val - ref).max()}"


def _field(n_x, n_y, seed, smooth=True):
    from covof_amd import workloads as wl
    C = wl.ellipses(n_x, n_y, 10, 3.0, 0.2 * min(n_x, n_y), seed=seed)
    rng = np.random.default_rng(seed)
    if not smooth:
        C = C + 0.01 * rng.random(C.shape)
    return C


@pytest.mark.parametrize("bc", [(4, 4, 4, 4), (3, 3, 3, 3), (3, 3, 1, 1), (2, 2, 3, 3)],
                         ids=["walls", "periodic", "periodic-x", "periodic-y"])
def test_surface_kernels_vs_oracle(cv, orc, bc):
    """Each kernel alone on a larger grid and every halo combination:
    curvature + CSF (normals from a smoothed field), diffusion (1, 2 and 5
    steps), SVOF smoothing; the force arrays' untouched entries keep their
    prior contents."""
    nx, ny = 150, 98
    m = orc.mesh(nx, ny, 1.3, 0.85, bc)
    ctx = cv.Context(nx, ny, 1.3, 0.85, bc=bc)
    C = _field(nx, ny, 5)
    Cg = to_gpu(C)
    for steps in (1, 2, 5):
        Ct = ctx.smooth_diffusion(Cg, steps).cpu().numpy()
        assert np.array_equal(Ct, orc.smooth_diffusion(m, C, steps)), steps
    Cs = ctx.smooth_svof_helmholtz(Cg, 1.7).cpu().numpy()
    assert np.array_equal(Cs, orc.svof_smooth(m, C, 1.7))
    rng = np.random.default_rng(9)
    f0x, f0y = rng.random(C.shape), rng.random(C.shape)
    for G in (C, Ct, Cs):
        out = {"curvature": ctx.new_field(5.0), "fx_sigma": to_gpu(f0x), "fy_sigma": to_gpu(f0y)}
        ctx.compute_curvature_and_csf(Cg, 0.0728, G=to_gpu(G), out=out)
        curv, fx, fy = orc.curvature_csf(m, G, C, 0.0728, 1, f0x, f0y)
        assert np.array_equal(out["curvature"].cpu().numpy(), curv)
        assert np.array_equal(out["fx_sigma"].cpu().numpy(), fx)
        assert np.array_equal(out["fy_sigma"].cpu().numpy(), fy)
    # switched off: everything zero
    out = ctx.compute_curvature_and_csf(Cg, 0.0728, enable=False,
                                        out={"curvature": ctx.new_field(5.0), "fx_sigma": to_gpu(f0x),
                                             "fy_sigma": to_gpu(f0y)})
    assert all(float(t.abs().max()) == 0.0 for t in out.values())


@pytest.mark.parametrize("bc", [(4, 4, 4, 4), (3, 3, 3, 3), (3, 3, 1, 1), (2, 2, 3, 3)],
                         ids=["walls", "periodic", "periodic-x", "periodic-y"])
def test_height_function_curvature_vs_oracle(cv, orc, bc):
    """The height-function curvature option (no reference counterpart; its
    accuracy is pinned on the CPU: zero on lines, 1/R on circles) and the CSF
    forces from it, bitwise against the oracle for every halo combination,
    with the normals from C and from the diffusion-smoothed field; most
    interface cells take the height-function value."""
    nx, ny = 150, 98
    m = orc.mesh(nx, ny, 1.3, 0.85, bc)
    ctx = cv.Context(nx, ny, 1.3, 0.85, bc=bc)
    C = _field(nx, ny, 5)
    Cg = to_gpu(C)
    Ct = orc.smooth_diffusion(m, C, 3)
    for G in (C, Ct):
        out = ctx.compute_curvature_and_csf(Cg, 0.0728, G=to_gpu(G), method="height_function")
        curv, fx, fy = orc.curvature_csf(m, G, C, 0.0728, 1, method=1)
        for k, ref in (("curvature", curv), ("fx_sigma", fx), ("fy_sigma", fy)):
            got = out[k].cpu().numpy()
            assert np.array_equal(got, ref), f"{k}: max diff {np.abs(got - ref).max()}"
        k0, _, _ = orc.curvature_csf(m, G, C, 0.0728, 1, method=0)
        band = (C[1:-1, 1:-1] > 1e-6) & (C[1:-1, 1:-1] < 1 - 1e-6)
        changed = (curv[1:-1, 1:-1] != k0[1:-1, 1:-1]) & band
        assert changed.sum() > 0.8 * band.sum(), (changed.sum(), band.sum())


def test_surface_rejects_aliasing(cv):
    ctx = cv.Context(16, 16)
    C = ctx.new_field(0.5)
    with pytest.raises(cv.CovofError):
        ctx.smooth_diffusion(C, 3, C_tilde=C)
    with pytest.raises(cv.CovofError):
        ctx.smooth_svof_helmholtz(C, 1.0, C_s=C)
    with pytest.raises(cv.CovofError):
        ctx.compute_curvature_and_csf(C, 1.0, out={"curvature": C, "fx_sigma": ctx.new_field(),
                                                   "fy_sigma": ctx.new_field()})


def test_csf_full_size_properties(cv):
    """2048^2 periodic workload: the curvature of the diffusion-smoothed
    field is finite and below 2/dx, and a face between two equal pure cells
    carries no force."""
    from covof_amd import workloads as wl
    n = 2048
    C, _, _, _ = wl.vof_case(n, seed=0)
    ctx = cv.Context(n, n, bc=(3, 3, 3, 3))
    Cg = to_gpu(C)
    Ct = ctx.smooth_diffusion(Cg, 3)
    out = ctx.compute_curvature_and_csf(Cg, 0.07, G=Ct)
    curv = out["curvature"].cpu().numpy()
    fx = out["fx_sigma"].cpu().numpy()
    assert np.isfinite(curv).all() and np.abs(curv).max() < 2.0 * n
    Ci = C[1:n + 1, 1:n + 1]
    far = (Ci == 0.0) | (Ci == 1.0)
    # a face's force needs a C jump across it: pure regions with pure neighbours carry none
    pure = far[:, 1:] & far[:, :-1] & (Ci[:, 1:] == Ci[:, :-1])
    assert np.abs(fx[1:n + 1, 2:n + 1][pure]).max() == 0.0
