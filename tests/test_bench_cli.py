"""The benchmark's contract on the CPU: one GPU only, the workload table, and the
algorithmic bytes it prices the kernels with (no GPU call)."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import turbchannel_amd as tca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("name", sorted(bench.WORKLOADS))
def test_workload_definitions(name):
    scheme, nx, ny, nz, zz, kw, dt = bench.problem(name, tca.FP64)
    collocated = scheme in (tca.SCHEME_AB3, tca.SCHEME_SRK)
    assert len(zz) == (nz if collocated else nz + 1)  # nodes (collocated) or faces (staggered)
    assert zz[0] == 0.0 and zz[-1] == 2.0 and np.all(np.diff(zz) > 0)
    assert nx % 2 == 0 and ny % 2 == 0 and dt > 0
    assert math.isclose(kw["hx"] * nx, 2 * math.pi) and math.isclose(kw["hy"] * ny, math.pi)
    # bulk units: nu = 2 / Re_b, u_tau = Re_tau * nu, forcing u_tau^2 / H
    assert math.isclose(kw["sour_dpdx"], (kw["visc"] * kw["inv_length_scale"]) ** 2)
    if scheme == tca.SCHEME_SRK:
        assert kw["cfr_ubulk"] == 0.0
    else:
        assert kw["cfr_ubulk"] == 1.0


def test_headline_workload_is_the_fourth_order_dns():
    scheme, nx, ny, nz, zz, kw, dt = bench.problem("vv_192x160x128", tca.FP64)
    assert (scheme, nx, ny, nz, dt) == (tca.SCHEME_VV, 192, 160, 128, 1e-3)
    assert math.isclose(2.0 / kw["visc"], 5600.0)


def test_algorithmic_reals_per_step():
    """Reals per grid point and step of the staggered DNS: 42 with the z-solve's
    two sweeps over the spectrum counted (4), 40 with the spectrum counted once."""
    r = bench.REALS_PER_POINT
    assert r["flux"] == 21 and r["project"] == 7 and r["zsolve"] == 4
    assert r["fft_y"] == r["ifft_y"] == 2
    total = r["flux"] + r["fft_x"] + r["fft_y"] + r["zsolve"] + r["ifft_y"] + r["ifft_x"] + r["project"]
    assert total == 42 and total - 2 == 40
