"""The reference's own driver (BD_cuda/main.cpp, readfile.cpp) linked
against libbdhip.so through include/func_cuda.h, run on the reference's
input deck (tests/golden/testdata).

The driver seeds its noise with time(0) (main.cpp:60-62), so only what the
noise cannot change is compared: the step-0 energy (force pass 1 at the
initial state) against the compiled reference's energy of that state, and
the finiteness and shape of every output file.  The LH binding Monte Carlo
is switched off (unbind = Pbind = 0) so that the bound set stays fixed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from decks import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "bd_ref_driver")
DECK = os.path.join(GOLDEN, "testdata")


def run_driver(tmp, steps, hi_set):
    for f in os.listdir(DECK):
        shutil.copy(os.path.join(DECK, f), tmp)
    setup = os.path.join(tmp, "setup.txt")
    out = []
    for line in open(setup):
        key = line.split()[0] if line.split() else ""
        if key == "number_of_steps":
            line = f"number_of_steps {steps}\n"
        elif key == "frequency_of_sampling":
            line = "frequency_of_sampling 1\n"
        elif key in ("unbind", "Pbind"):
            line = f"{key} 0\n"
        out.append(line)
    os.chmod(setup, 0o644)
    open(setup, "w").writelines(out)
    env = {k: v for k, v in os.environ.items() if k != "BD_HI_SET"}
    if hi_set == "main":
        env["BD_HI_SET"] = "main"
    print(f"running the reference driver ({steps} steps, hydrodynamic set {hi_set})", flush=True)
    p = subprocess.run([DRIVER], cwd=tmp, env=env, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "failed" not in p.stderr, p.stderr
    return p


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/bd_ref_driver not built")
@pytest.mark.parametrize("hi_set", ["all", "main"])
def test_reference_driver_runs_on_the_hip_library(tmp_path, hi_set):
    run_driver(str(tmp_path), 3, hi_set)
    E = np.loadtxt(tmp_path / "Energy.txt")
    assert E.shape == (3,) and np.all(np.isfinite(E))
    _, st, g = load_golden("testdata_deck")
    # Energy.txt carries the default 6 significant digits (main.cpp:494)
    assert E[0] == pytest.approx(float(g["force_all_E"]), rel=5e-6)
    # restart files: positions and frames after the last step, finite and
    # within a few nm of the start (three 1 ps steps)
    r_end = np.loadtxt(tmp_path / "restart_r.txt").reshape(-1)
    assert r_end.shape == st.r.shape and np.all(np.isfinite(r_end))
    assert np.max(np.abs(r_end - st.r)) < 1.0
    for f in ("restart_a.txt", "restart_b.txt", "restart_c.txt", "restart_tail.txt", "restart_lh.txt"):
        v = np.loadtxt(tmp_path / f)
        assert np.all(np.isfinite(v)), f
    xyz = open(tmp_path / "out.xyz").read()
    assert xyz.count("\n") > 100
