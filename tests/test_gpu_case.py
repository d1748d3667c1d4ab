"""A reference JSON case run on the GPU (covof_amd.case: the reference's
initialisation and time loop, every step on the GPU, the reference's output
files) against the compiled reference program's own run of the same file."""
import json
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _arrays(text):
    """name -> flat float array of every DataArray of a .vtr file."""
    out = {}
    for m in re.finditer(r'Name="(\w+)"[^>]*format="ascii">\n(.*?)</DataArray>', text, re.S):
        out[m.group(1)] = np.array([float(x) for x in m.group(2).split()])
    return out


@pytest.mark.parametrize("fixture", ["case_bubble_json.npz", "case_bubble_mono_json.npz"])
def test_json_case_matches_reference_program(cv, tmp_path, fixture):
    """The standard-projection case (CG + Jacobi predictor, the reference's
    CG + SSOR Poisson block) and the same case with the fully coupled solver
    (MONOLITHIC, BiCGSTAB(2) + block-triangular Schur-SSOR)."""
    from covof_amd import case
    print(f"\n[json case] {fixture}", flush=True)
    g = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    cfg = json.loads(str(g["json"]))
    out = tmp_path / "out"
    cfg["output"]["output_dir"] = str(out)
    path = tmp_path / "case.json"
    path.write_text(json.dumps(cfg))
    run = case.run_case(case.load_config(str(path)), log=lambda *a: None)
    assert run.steps == 13
    # step 0: initial condition, properties, curvature: the same file
    assert (out / "bubble_p0_000000.vtr").read_text() == str(g["vtr_0"])
    # history: same rows; step, time and dt columns identical, the others
    # within the rounding the solves' summation order leaves
    got = (out / "bubble_history.csv").read_text().splitlines()
    exp = str(g["history"]).splitlines()
    assert len(got) == len(exp) and got[0] == exp[0] and got[1] == exp[1]
    worst = 0.0
    for a, b in zip(got[2:], exp[2:]):
        fa, fb = a.split(","), b.split(",")
        assert fa[:3] == fb[:3], (a, b)
        for k, (x, y) in enumerate(zip(fa[3:], fb[3:])):
            x, y = float(x), float(y)
            if k == 5:  # max_div: a solve-residual quantity (1e-8 here)
                assert abs(x - y) <= 1e-9, (a, b)
            else:
                worst = max(worst, abs(x - y) / max(abs(y), 1e-300))
    print(f"history: worst relative difference {worst:.2e}")
    assert worst <= 2e-6
    # the last step's file: every array within 1e-9 of the reference's scale
    for step in (5, 10, 13):
        A = _arrays((out / f"bubble_p0_{step:06d}.vtr").read_text())
        B = _arrays(str(g[f"vtr_{step}"]))
        assert A.keys() == B.keys()
        for k in B:
            scale = max(np.abs(B[k]).max(), 1e-300)
            tol = 1e-6 if k == "Divergence" else 1e-9
            err = np.abs(A[k] - B[k]).max() / scale
            assert err <= tol, (step, k, err)
