"""The command-line driver without a GPU: options, messages and the wall-normal
meshes it builds (no engine is created for these)."""
import math

import numpy as np
import pytest

from tests import cases
from tests.driver_replay import DRIVER, run_driver

# the reference's own options (staggered/io.cpp:107-117)
REFERENCE_OPTIONS = ["-d,--device", "--nz", "-t,--TimeMax", "-a,--TimeStartAveraging", "--TimeStepsMax", "--tau",
                     "--CFLmin", "--CFLmax", "-s,--SProgress", "-f,--FProgress", "-p,--PProgress"]


def mesh(tmp_path, *args):
    r = run_driver(["--print-mesh", 1, *args], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    nz = int(lines[0].split()[1])
    return nz, np.array([float(v) for v in lines[1:] if v.strip()])


def test_driver_help_lists_reference_options(tmp_path):
    r = run_driver(["--help"], tmp_path)
    assert r.returncode == 0
    for opt in REFERENCE_OPTIONS:
        assert opt in r.stdout, opt


def test_driver_rejects_unknown_option_like_reference(tmp_path):
    # CLI11's message, then exit(0) (staggered/main.cpp:192)
    r = run_driver(["--bogus", "1"], tmp_path)
    assert r.returncode == 0
    assert "The following argument was not expected: --bogus" in r.stdout
    r = run_driver(["--tau"], tmp_path)
    assert "required argument missing" in r.stdout
    r = run_driver(["--tau", "abc"], tmp_path)
    assert "Could not convert" in r.stdout
    assert not list(tmp_path.iterdir())  # nothing written


def test_driver_sinh_mesh_matches_reference_formula(tmp_path):
    for nz in (16, 17, 32):
        n, z = mesh(tmp_path, "--nz", nz)
        assert n == nz
        np.testing.assert_allclose(z, cases.sinh_mesh(nz), rtol=0, atol=1e-15)
    n, z = mesh(tmp_path, "--nz", 20, "--stretch", 3.0)
    np.testing.assert_allclose(z, cases.sinh_mesh(20, stretch=3.0), rtol=0, atol=1e-15)


def test_driver_tanh_and_uniform_meshes(tmp_path):
    n, z = mesh(tmp_path, "--mesh", "tanh", "--gamma", 1.8, "--nz", 24)
    assert n == 24
    np.testing.assert_allclose(z, cases.tanh_faces(24, 1.8), rtol=0, atol=1e-14)
    n, z = mesh(tmp_path, "--scheme", "ab3", "--mesh", "tanh", "--gamma", 2.0, "--nz", 33)
    assert n == 33
    np.testing.assert_allclose(z, cases.ab3_mesh_tanh(33, 2.0), rtol=0, atol=1e-14)
    n, z = mesh(tmp_path, "--mesh", "uniform", "--nz", 10)
    np.testing.assert_allclose(z, np.linspace(0, 2, 11), rtol=0, atol=1e-15)


def test_driver_geometric_mesh_is_symmetric_progression(tmp_path):
    # collocated_AB3/main.cpp:205-226 in wall units: first step y1, growth by
    # the ratio while below the cap, a uniform core up to the centre, mirrored
    re_tau = 180.0
    n, z = mesh(tmp_path, "--scheme", "ab3", "--Re_tau", re_tau, "--geo_y1", 1.0, "--geo_ratio", 1.1,
                "--geo_ybulk", 5.0)
    zp = z * re_tau
    assert n == len(z)
    np.testing.assert_allclose(zp + zp[::-1], 2 * re_tau, rtol=0, atol=1e-9)
    steps = np.diff(zp[: n // 2 + 1])
    assert steps[0] == pytest.approx(1.0)
    k = int(math.ceil(math.log(5.0) / math.log(1.1)))  # growing steps below the cap: 1.1^j < 5
    np.testing.assert_allclose(steps[1:k], steps[: k - 1] * 1.1, rtol=1e-12)
    core = steps[k:]
    assert np.all(core < 5.0) and np.ptp(core) < 1e-9
    # the reference's own parameters (y1 = 50: the cap is reached at once)
    n, z = mesh(tmp_path, "--scheme", "ab3", "--Re_tau", re_tau)
    assert z[1] * re_tau == pytest.approx(50.0)
    core = np.diff(z[1: n // 2 + 1] * re_tau)
    assert np.ptp(core) < 1e-9 and core[0] < 2.2


def test_driver_mesh_from_file(tmp_path):
    zz = cases.tanh_faces(12, 1.5)
    (tmp_path / "z.txt").write_text("\n".join("%.17g" % v for v in zz))
    n, z = mesh(tmp_path, "--mesh", "file", "--zfile", "z.txt")
    assert n == 12
    np.testing.assert_array_equal(z, zz)
    r = run_driver(["--print-mesh", 1, "--mesh", "file", "--zfile", "z.txt", "--nz", 10], tmp_path)
    assert "does not match the mesh" in r.stdout


def test_driver_without_device_reports_and_exits(tmp_path, monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = run_driver(["--nx", 8, "--ny", 8, "--nz", 8, "--TimeStepsMax", 1], tmp_path)
    assert r.returncode == 0
    assert r.stdout.strip().splitlines()[-1].startswith("tc_create:")


def test_driver_binary_exists():
    import os
    assert os.access(DRIVER, os.X_OK)


def test_driver_srk_defaults_are_the_reference_mesh(tmp_path):
    """--scheme srk without mesh options: collocated_SRK/main.cpp:312-333 (Re_tau 180,
    first step 1, ratio 1.3, steps capped at 10 wall units), 49 nodes."""
    n, z = mesh(tmp_path, "--scheme", "srk")
    zp = [0.0, 1.0]
    while (zp[-1] - zp[-2]) * 1.3 < 10.0:
        zp.append(zp[-1] + (zp[-1] - zp[-2]) * 1.3)
    m = int((180.0 - zp[-1]) / 10.0) + 1
    hb = (180.0 - zp[-1]) / m
    zp += [zp[-1] + hb * (i + 1) for i in range(m - 1)] + [180.0]
    zp += [360.0 - v for v in zp[-2::-1]]
    assert n == len(zp) == 49
    np.testing.assert_allclose(z * 180.0, zp, rtol=1e-14, atol=1e-12)


def test_driver_srk_rejects_unknown_method(tmp_path):
    r = run_driver(["--scheme", "srk", "--imex", "rk4", "--print-mesh", 1], tmp_path)
    assert r.returncode == 0 and "unknown IMEX method rk4" in r.stdout
