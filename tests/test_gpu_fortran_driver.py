"""The Fortran time loop (jd_lepic_amd/fortran/lepic_driver.f90) drives the
hot path through iso_c_binding: a short run on a small box must finish,
conserve the particle count up to wall losses and ionization, and print the
reference-style phase timers."""
import os
import subprocess
import sys

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


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_backup_restart_and_moments(tmp_path):
    """navg/nbak/restart: the run averages moments, writes a backup in the
    reference's format at step 7 and a second run restarts from it."""
    (tmp_path / "DATA.BAK").mkdir()
    (tmp_path / "lepic.nml").write_text("&lepic n=64, ppc=10, nsteps=8, ng=6, precision=1, navg=2, nbak=6 /\n")
    out = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "backup written at step        7" in out.stdout
    te = [l for l in out.stdout.splitlines() if "Te(centre)" in l][0].split()
    assert int(te[1]) == 4 and 1.0 < float(te[-1]) < 6.0      # 3 eV electrons, 4 averages
    bak = tmp_path / "DATA.BAK" / "particles0.bak"
    sys.path.insert(0, ROOT)
    import jd_lepic_amd as L
    t, g, pc, counts = L.backup_header(bak, 2)
    assert counts.shape == (1, 2) and 0.9 * 64 * 64 * 10 < counts[0, 0] < 1.1 * 64 * 64 * 10
    (tmp_path / "lepic.nml").write_text("&lepic n=64, ppc=10, nsteps=2, ng=6, precision=1, restart=1 /\n")
    out2 = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stdout + out2.stderr
    counts2 = [l for l in out2.stdout.splitlines() if "particles:" in l][0].split()
    assert abs(int(counts2[2]) - counts[0, 0]) < 0.01 * counts[0, 0]


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("precision,mode", [(0, 0), (1, 1)])
def test_fortran_driver_heating(tmp_path, precision, mode):
    """flag_heat = 1: the driver calls eheating every ns_heat steps before
    the push (check mode on the FP64 state, production mode on the mixed
    one); eheat_type = 1 replaces heated velocities by 3 eV Maxwellian draws
    from 3 eV electrons, so the tally is small next to eheat_type = 0, where
    every kick adds energy on average."""
    tallies = []
    for etype in (0, 1):
        (tmp_path / "lepic.nml").write_text(
            f"&lepic n=64, ppc=10, nsteps=6, ng=6, precision={precision}, mcc_mode={mode}, flag_heat=1, "
            f"ns_heat=2, flag_c=1, eheat_type={etype}, nudt=0.2 /\n")
        out = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        tallies.append(float([l for l in out.stdout.splitlines() if "heating P_loss" in l][0].split()[-1]))
    assert tallies[0] > 0.0 and abs(tallies[1]) < 0.3 * tallies[0]
