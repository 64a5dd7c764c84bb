"""The souffle-gpu command-line driver: the reference's flags, outfiles/ writers and restart."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_lib import REPO, TEMPLATE

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(REPO, "capsid-souffle_amd", "bin", "souffle-gpu")
FLAGS = ["-f", "trimer_MVM", "-S", "20", "-C", "800", "-c", "100", "-E", "2.1", "-t", "0.005", "-K", "298"]


def _run(cwd, extra):
    r = subprocess.run([DRIVER] + FLAGS + extra, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _rows(path):
    rows = []
    for line in open(path):
        tok = line.split()
        if tok and tok[0] != "time":
            rows.append([float(x) for x in tok])
    return rows


def test_driver_outputs_and_resume(tmp_path):
    os.makedirs(tmp_path / "infiles")
    shutil.copy(TEMPLATE, tmp_path / "infiles" / "trimer_MVM")
    _run(tmp_path, ["-T", "100", "-W", "50", "-w", "50", "-M", "50", "-R", "false"])
    out = tmp_path / "outfiles"
    rows = _rows(out / "energy.out")
    # initial row (11 columns) then loop rows at a = 0, 50, 100 (12 columns)
    assert [r[0] for r in rows] == [0, 0, 50, 100]
    assert len(rows[0]) == 11 and all(len(r) == 12 for r in rows[1:])
    assert abs(rows[0][1] - 1.5) < 1e-4        # KE/N = 3T/2 at T = 1
    assert abs(rows[0][8] - 1.0) < 1e-4        # temperature column
    for r in rows:
        assert np.isfinite(r).all()
    for a in (0, 50, 100):
        assert (out / f"restart_{a}.out").exists()
    assert open(out / "ovito.lammpstrj").read().count("ITEM: TIMESTEP") == 3
    # resume from the newest restart file (restart_100.out) and run to a = 150
    _run(tmp_path, ["-T", "150", "-W", "50", "-w", "50", "-M", "1000", "-R", "true"])
    rows2 = _rows(out / "energy.out")
    assert [r[0] for r in rows2[len(rows):]] == [0, 100, 150]
    assert (out / "restart_150.out").exists()


def test_driver_oligomers_and_uncut_yukawa(tmp_path):
    os.makedirs(tmp_path / "infiles")
    shutil.copy(TEMPLATE, tmp_path / "infiles" / "trimer_MVM")
    _run(tmp_path, ["-T", "200", "-W", "100", "-w", "1000", "-M", "100", "-R", "false", "--oligomers", "3.0"])
    out = tmp_path / "outfiles"
    rows = _rows(out / "oligomers.out")
    assert [r[0] for r in rows] == [0, 100, 200]
    for r in rows:
        ms = np.array(r[3:])
        assert len(ms) == 20 and ms.sum() == r[1] and (np.arange(1, 21) * ms).sum() == 20
        assert r[2] == np.nonzero(ms)[0].max() + 1
    # analysis.ms: frame average of (oligomers of size k) x k, as src/analysis.cpp:267-285
    avg = np.array([float(x) for x in open(out / "analysis.ms").read().split()])
    want = np.mean([np.array(r[3:]) * np.arange(1, 21) for r in rows], axis=0)
    assert np.allclose(avg, want, rtol=1e-5) and abs(avg.sum() - 20) < 1e-4
    # -e 0: the Yukawa over every charged pair of the box (far field beyond 2.5 diameters)
    d2 = tmp_path / "uncut"
    os.makedirs(d2 / "infiles")
    shutil.copy(TEMPLATE, d2 / "infiles" / "trimer_MVM")
    r = subprocess.run([DRIVER, "-f", "trimer_MVM", "-S", "20", "-C", "200", "-c", "10", "-E", "2.1", "-t", "0.005",
                        "-e", "0", "-T", "100", "-W", "50", "-R", "false"], cwd=d2, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = _rows(d2 / "outfiles" / "energy.out")
    assert [x[0] for x in rows] == [0, 0, 50, 100] and all(np.isfinite(x).all() for x in rows)


def test_driver_encapsulation(tmp_path):
    """-p: immobile big beads (src/md.cpp:183-185), capenergy.out rows (src/md.cpp:347-352,
    625-630), rb_<a>.out beside each restart file (src/md.cpp:537-545), the big beads in each
    movie frame (src/md.cpp:649-676) and the rb restart (src/initialize.cpp:574-600)."""
    os.makedirs(tmp_path / "infiles")
    shutil.copy(TEMPLATE, tmp_path / "infiles" / "trimer_MVM")
    _run(tmp_path, ["-T", "100", "-W", "50", "-w", "50", "-M", "50", "-R", "false", "-p", "true", "-m", "2"])
    out = tmp_path / "outfiles"
    rows = _rows(out / "capenergy.out")
    assert [r[0] for r in rows] == [0, 0, 50, 100]
    for r in rows:
        assert len(r) == 6 and np.isfinite(r).all()
        assert r[1] == 0 and r[5] == 0              # immobile, zero velocity
        assert r[2] >= 0                            # shifted WCA energy
        assert abs(r[4] - (r[1] + r[2] + r[3])) <= 1e-5 * max(1.0, abs(r[4]))
    big = [np.loadtxt(out / f"rb_{a}.out") for a in (0, 50, 100)]
    for b in big:
        assert b.shape == (2, 6) and np.array_equal(b, big[0])
    text = open(out / "ovito.lammpstrj").read()
    n_sub = 20 * 105
    assert text.count("ITEM: NUMBER OF ATOMS\n%d\n" % (n_sub + 2)) == 3
    last = text.strip().splitlines()[-1].split()
    assert int(last[0]) == n_sub + 2 and int(last[1]) == 3 and float(last[6]) == -78
    # resume: subunits from restart_100.out, big beads from rb_100.out
    _run(tmp_path, ["-T", "150", "-W", "50", "-w", "50", "-M", "1000", "-R", "true", "-p", "true", "-m", "2"])
    rows2 = _rows(out / "capenergy.out")
    assert [r[0] for r in rows2] == [0, 100, 150]
    assert np.array_equal(np.loadtxt(out / "rb_150.out"), big[0])
