"""Wall-clock times of the compiled reference (the "R" CPU baselines) on this
container's cores, through the harness's timing modes.

Builds (if needed) the reference with its own optimisation flags
(``-O3 -march=native``, CMakeLists.txt:9) into ``oracle/_ref/fast`` and runs
it single-rank (the harness hands whole fields to one rank; with more ranks
the reference's SP Poisson operator ignores the inter-rank halos, so only
np = 1 is the same algorithm):

* ``vof``: ELVIRA + Weymouth-Yue steps of the VOF workload (128 seeded
  ellipses, divergence-free Fourier velocity, CFL 0.25, periodic) at 2048^2
  and 4096^2 (``tadvect``);
* ``sp``: one standard-projection step without convection on the
  variable-density Poisson workload (64 seeded circles, rho 1 vs 1000, the
  Gaussian right-hand side carried by v) at 256^2 and 512^2 (``tsp``): the
  predictor alone, then the step, whose difference is the Poisson block
  (the reference's hard-coded CG + symmetric Gauss-Seidel to 1e-7, at most
  1000 iterations) and the correction;
* ``plain``: the reference program itself on the plain-method case (config
  1: 128 x 256 rising bubble, 200 steps), its output files included.

Runs only where the reference tree exists; nothing here travels or is
measured on the GPU.

    python oracle/time_reference.py [vof|sp|plain|all]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "covof_2d-fortran_amd"))
from covof_amd import workloads as wl  # noqa: E402

OUT = os.path.join(HERE, "_ref", "fast")
HARNESS = os.path.join(OUT, "covof_ref_harness")
FFLAGS = "-O3 -march=native -I/opt/conda/include"


def build():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", HERE, "ref", "R=_ref/fast", f"FFLAGS={FFLAGS}"], check=True,
                       stdout=subprocess.DEVNULL)


def run(mode: str, payload: bytes) -> bytes:
    with tempfile.TemporaryDirectory(dir=OUT) as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            f.write(payload)
        subprocess.run([HARNESS, mode, fi, fo], check=True, cwd=td,
                       env=dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1"))
        with open(fo, "rb") as f:
            return f.read()


def i32(*v):
    return np.asarray(v, dtype="<i4").tobytes()


def f64(*v):
    return np.asarray(v, dtype="<f8").tobytes()


def vof(n: int, nsteps: int):
    C, u, v, dt = wl.vof_case(n, seed=0)
    payload = (i32(n, n, 3, 3, 3, 3, 1, 0, nsteps) + f64(1.0, 1.0, dt, 1.0, 1.0, 1.0, 1.0)
               + C.astype("<f8").tobytes() + u.astype("<f8").tobytes() + v.astype("<f8").tobytes())
    t, vol = np.frombuffer(run("tadvect", payload), dtype="<f8")
    per = t / nsteps
    print(f"| VOF ELVIRA + WY {n}^2 (R, np=1) | {nsteps} steps | {per * 1e3:.1f} ms/step | "
          f"{n * n / per:.3e} cell-updates/s | volume {vol:.12e} |", flush=True)


def sp(n: int):
    c = wl.poisson_case(n, seed=0, n_circles=64, rmin=n / 128, rmax=n / 16)
    payload = (i32(n, n, 4, 4, 4, 4) + f64(1.0, 1.0, 1.0e-3, 1000.0, 1.0, 1.0e-3, 1.0e-5, 0.0, 0.0)
               + c["C"].astype("<f8").tobytes() + c["u"].astype("<f8").tobytes()
               + c["v"].astype("<f8").tobytes())
    out = run("tsp", payload)
    t_pred, t_step = np.frombuffer(out[:16], dtype="<f8")
    iters, iters_u = np.frombuffer(out[16:24], dtype="<i4")
    t_p = t_step - t_pred
    it_p = int(iters) - int(iters_u)
    print(f"| SP Poisson block {n}^2, CG + SGS to 1e-7 (R, np=1) | {it_p} iterations | {t_p:.2f} s "
          f"(predictor {t_pred:.2f} s, {int(iters_u)} iterations) | {t_p / max(it_p, 1) * 1e3:.2f} ms/iteration |",
          flush=True)


def plain():
    """The reference program on the plain-method case (config 1's JSON):
    200 steps of the 128 x 256 rising bubble, its own output included."""
    import json
    case = {"case_name": "plain", "mode": "NAVIER_STOKES", "mesh": {"nx": 128, "ny": 256, "lx": 1.0, "ly": 2.0},
            "physics": {"enable_two_phase": True, "enable_surface_tension": True, "sigma": 24.5, "gy": -0.98,
                        "convective_scheme": "TVD_SUPERBEE", "fluid1": {"rho": 1000.0, "mu": 10.0},
                        "fluid2": {"rho": 100.0, "mu": 1.0}},
            "time": {"time_scheme": "BDF2", "dt": 1e-3, "t_final": 0.2},
            "solver": {"type": "STANDARD_PROJECTION", "linear_solver": "CG", "preconditioner": "JACOBI",
                       "tol": 1e-12, "max_iter": 500},
            "vof": {"advection_method": "WEYMOUTH_YUE", "reconstruction_method": "PLIC"},
            "boundary_conditions": {"left": {"type": "FREE_SLIP"}, "right": {"type": "FREE_SLIP"},
                                    "bottom": {"type": "DIRICHLET"}, "top": {"type": "DIRICHLET"}},
            "initial_condition": {"type": "BUBBLE", "center_x": 0.5, "center_y": 0.5, "radius": 0.25}}
    with tempfile.TemporaryDirectory(dir=OUT) as td:
        path = os.path.join(td, "plain.json")
        with open(path, "w") as f:
            json.dump(case, f)
        t = time.perf_counter()
        subprocess.run([HARNESS, "sim", path, os.path.join(td, "o.bin")], check=True, cwd=td,
                       stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1"))
        t = time.perf_counter() - t
    print(f"| plain-method case 128x256, 200 steps (R, np=1) | {t:.1f} s | {t / 200 * 1e3:.0f} ms/step |", flush=True)


def main():
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    build()
    t0 = time.time()
    if what in ("vof", "all"):
        vof(2048, 4)
        vof(4096, 2)
    if what in ("sp", "all"):
        sp(256)
        sp(512)
    if what in ("plain", "all"):
        plain()
    print(f"(wall {time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
