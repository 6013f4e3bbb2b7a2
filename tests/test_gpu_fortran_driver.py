"""The Fortran time loop (jd_lepic_amd/fortran/lepic_driver.f90) drives the
hot path through iso_c_binding: a short run on a small box must finish,
conserve the particle count up to wall losses and ionization, and print the
reference-style phase timers."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "jd_lepic_amd", "fortran", "lepic_driver")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_runs(tmp_path):
    (tmp_path / "lepic.nml").write_text("&lepic n=64, ppc=20, nsteps=12, ng=6, precision=1 /\n")
    out = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    counts = [l for l in lines if "particles:" in l][0].split()
    ne, ni = int(counts[2]), int(counts[4])
    loaded = 64 * 64 * 20
    assert 0.9 * loaded < ne < 1.1 * loaded and 0.9 * loaded < ni < 1.1 * loaded
    assert any("<t> (ms):" in l for l in lines)
